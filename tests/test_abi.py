"""The C-ABI library loads, exports every symbol include/speo_de.h declares, and
its structs have the reference's layouts (CUDA_EA_library/EA_CUDA.h:21-39,
config.h:48-59 with real = float).  Host-side argument validation is checked
without touching a GPU."""
import ctypes

import pytest


def test_exports_every_declared_symbol(speo):
    from speo_de import _lib
    declared = _lib.header_symbols()
    assert len(declared) >= 30
    L = speo.lib()
    missing = [n for n in declared if not hasattr(L, n)]
    assert missing == []
    # and the ctypes table covers every declaration
    assert sorted(_lib._SIGS) == declared


def test_struct_layouts(speo):
    from speo_de import _lib
    assert ctypes.sizeof(_lib.GPU_Population) == 16
    assert ctypes.sizeof(_lib.GPU_IslandInfo) == 16
    assert _lib.GPU_IslandInfo.permutated_index.offset == 8
    assert ctypes.sizeof(_lib.DEInfo) == 16
    assert [f.offset for f in (_lib.DEInfo.CR, _lib.DEInfo.F, _lib.DEInfo.strategy_ID,
                               _lib.DEInfo.pop_size)] == [0, 4, 8, 12]
    assert ctypes.sizeof(_lib.ProblemInfo) == 36
    assert _lib.ProblemInfo.max_bound.offset == 28 and _lib.ProblemInfo.min_bound.offset == 32
    assert ctypes.sizeof(_lib.SPEO_RandState) == 16


def test_next_pow2_dim(speo):
    f = speo.lib().SPEO_NextPow2Dim
    assert [f(d) for d in (1, 10, 32, 33, 100, 128, 129, 1000)] == [32, 32, 32, 64, 128, 128, 256,
                                                                     1024]


def test_argument_errors_are_returned(speo):
    from speo_de import _lib
    L = speo.lib()
    h = ctypes.c_void_p()
    rc = L.speo_problem_create(1, 1, None, None, None, None, None, None, None, ctypes.byref(h))
    assert rc == _lib.SPEO_ERR_ARG
    assert "dim" in speo.last_error()
    rc = L.speo_engine_create(None, None, ctypes.byref(h))
    assert rc == _lib.SPEO_ERR_ARG
    with pytest.raises(speo.SpeoError):
        speo.generate_problem_data(1, 0)


def test_wide_eval_switch_is_host_side(speo):
    # the kernel switch of the wide 16-bit evaluations: 1 by default, -1
    # reads it, and no call touches a GPU
    prev, launches = speo.wide_eval(-1)
    assert prev == 1 and launches >= 0
    assert speo.wide_eval(0)[0] == 1
    assert speo.wide_eval(-1)[0] == 0
    assert speo.wide_eval(1)[0] == 0
    assert speo.wide_eval(-1) == (1, launches)
