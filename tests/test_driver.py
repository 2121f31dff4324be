"""The run driver's budget arithmetic and the result record of the
reference's island driver (IslandEA::Execute / PrintResult,
src/island_EA.cc:78-160; header line :24; per-run seed src/main.cc:138;
parameter string DE_CUDA::GetParameters, CUDA_EA_library/DE_CUDA.cc:18-56).
Numbers in the record are written as a C++ stream writes a double by default
(six significant digits, %g): the expected strings below follow that rule."""
import ctypes
import math

import pytest

# the header line the reference writes into a new SPEO_GPU.csv (island_EA.cc:24)
REFERENCE_HEADER = ("function_ID,run_ID,dim,best_fitness,time_period,communication_percentage,"
                    "total_FEs,node_num,GPU_num,island_num,total_pop_size,subisland_size,interval,"
                    "migration_size,migration_rate,buffer_capacity,migration_topology,buffer_manage,"
                    "regroup_option,configure_EA,EA_parameters")

# strategy names of DE_CUDA::GetParameters, IDs 0..8
STRATEGIES = ["current/1/bin", "current/2/bin", "current-to-best/1/bin", "current-to-best/2/bin",
              "rand/1/bin", "rand/2/bin", "best/1/bin", "best/2/bin", "current_to_rand/1/bin"]


def _record(speo, **over):
    fields = dict(function_ID=9, run_ID=3, dim=10, best_fitness=0.12345678, time_period=1.5,
                  communication_percentage=0.0, total_FEs=100000, node_num=1, GPU_num=1,
                  island_num=4, total_pop_size=400, subisland_size=100, interval=50,
                  migration_size=2, migration_rate=0.5, buffer_capacity=32,
                  migration_topology="ring", buffer_manage="none",
                  regroup_option="static_and_ordered", configure_EA="constant",
                  EA_parameters="CR/F=0.9/0.5_rand/1/bin")
    fields.update(over)
    return speo.result_record(**fields)


def test_header_is_the_references(speo):
    assert speo.result_header() == REFERENCE_HEADER
    assert len(REFERENCE_HEADER.split(",")) == 21


def test_record_columns_in_order(speo):
    row = _record(speo)
    assert row == ("9,3,10,0.123457,1.5,0,100000,1,1,4,400,100,50,2,0.5,32,ring,none,"
                   "static_and_ordered,constant,CR/F=0.9/0.5_rand/1/bin")
    assert len(row.split(",")) == 21


@pytest.mark.parametrize("value,text", [(1e-10, "1e-10"), (123456789.0, "1.23457e+08"),
                                        (100000.0, "100000"), (1e6, "1e+06"), (0.1, "0.1"),
                                        (2.5e-5, "2.5e-05"), (11340.7, "11340.7"),
                                        (math.inf, "inf"), (math.nan, "nan")])
def test_record_number_format(speo, value, text):
    assert _record(speo, best_fitness=value).split(",")[3] == text


def test_parameters_string(speo):
    for sid, name in enumerate(STRATEGIES):
        assert speo.de_parameters(0.9, 0.5, sid) == "CR/F=0.9/0.5_" + name
    # the reference's real may be float: 0.9f printed as a double is still "0.9"
    f32 = ctypes.c_float(0.9).value
    assert speo.de_parameters(f32, ctypes.c_float(0.5).value, 4) == "CR/F=0.9/0.5_rand/1/bin"
    assert speo.de_parameters(0.15, 0.8, 4) == "CR/F=0.15/0.8_rand/1/bin"
    with pytest.raises(speo.SpeoError, match="strategy"):
        speo.de_parameters(0.9, 0.5, 9)


def test_record_rejects_short_buffer_and_null_fields(speo):
    from speo_de import _lib
    L = speo.lib()
    rec = _lib.RunRecord()
    rec.migration_topology = rec.buffer_manage = rec.regroup_option = b"x"
    rec.configure_EA = rec.EA_parameters = b"y"
    small = ctypes.create_string_buffer(8)
    assert L.speo_result_record(ctypes.byref(rec), small, 8) == _lib.SPEO_ERR_ARG
    assert "short" in speo.last_error()
    rec.EA_parameters = None
    big = ctypes.create_string_buffer(256)
    assert L.speo_result_record(ctypes.byref(rec), big, 256) == _lib.SPEO_ERR_ARG


def test_run_seed(speo):
    # (node_ID + function_ID) * (run_ID + function_ID)
    assert speo.run_seed(0, 1, 3) == 4
    assert speo.run_seed(2, 5, 10) == 105
    assert speo.run_seed(0, 30, 0) == 900


@pytest.mark.parametrize("max_fes,current,per,gens", [(1000, 100, 100, 9), (1001, 100, 100, 10),
                                                      (100, 100, 100, 0), (50, 100, 100, 0),
                                                      (1000, 40, 40, 24), (1000, 100, 0, 0),
                                                      (10**12, 8192, 8192, 122070312)])
def test_generations_for_fes(speo, max_fes, current, per, gens):
    # the budget loop: while current < max: one generation, current += per
    assert speo.generations_for_fes(max_fes, current, per) == gens
    if 0 < per and gens < 100:
        c, g = current, 0
        while c < max_fes:
            c += per
            g += 1
        assert g == gens
