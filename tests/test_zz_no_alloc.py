"""speo_evaluate_noalloc (no device allocation; a fixed per-GPU workspace)
against speo_evaluate, bit for bit, on two calls in a row, the second with a
smaller input, and its size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(speo, p, X, f, precision=0, n=None):
    """The C entry point itself (argument types as the header declares them)."""
    n = X.shape[0] if n is None else n
    return speo.lib().speo_evaluate_noalloc(p.handle, ctypes.c_void_p(X.data_ptr()), X.shape[1],
                                            n, ctypes.c_void_p(f.data_ptr()), precision, None)


def test_noalloc_equals_evaluate_on_two_calls(cuda, speo, oracle):
    import torch
    rng = np.random.default_rng(5)
    # call 1: a three-component composition at D = 1000 (per-chunk and
    # per-component workspace)
    D1, n1 = 1000, 96
    os_ = np.stack([oracle.gen_shift(s, D1) for s in (1, 2, 3)])
    Ms = np.stack([oracle.gen_rotation(s, D1) for s in (1, 2, 3)])
    p1 = speo.Problem([speo.OBJ_ELLIPTIC, speo.OBJ_RASTRIGIN, speo.OBJ_ACKLEY], os_, Ms,
                      sigma=[10.0, 20.0, 30.0], bias=[0.0, 100.0, 200.0])
    X1 = torch.from_numpy(rng.uniform(-100, 100, (n1, D1)).astype(np.float32)).to(cuda)
    for prec in (0, 3):
        ref = torch.zeros(n1, dtype=torch.float32, device=cuda)
        got = torch.full((n1,), -1.0, dtype=torch.float32, device=cuda)
        speo.evaluate(p1, X1, ref, precision=prec)
        assert _noalloc(speo, p1, X1, got, prec) == 0, speo.last_error()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
    # call 2: a smaller input (one component, D = 200, fewer rows); the
    # workspace keeps the previous call's contents, which must not matter
    D2, n2 = 200, 23
    o, M = oracle.gen_problem(9, D2)
    p2 = speo.Problem([speo.OBJ_DISCUS], o[None, :], M[None, :, :])
    X2 = torch.from_numpy(rng.uniform(-100, 100, (n2, D2)).astype(np.float32)).to(cuda)
    ref = torch.zeros(n2, dtype=torch.float32, device=cuda)
    got = torch.full((n2,), -1.0, dtype=torch.float32, device=cuda)
    speo.evaluate(p2, X2, ref)
    assert _noalloc(speo, p2, X2, got) == 0, speo.last_error()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
    np.testing.assert_allclose(got.cpu().numpy(),
                               oracle.evaluate(speo.OBJ_DISCUS, X2.cpu().numpy(), o, M), rtol=1e-5)
    p1.close()
    p2.close()


def test_noalloc_rejects_inputs_beyond_its_workspace(cuda, speo, oracle):
    import torch
    D = 1000
    o, M = oracle.gen_problem(4, D)
    p = speo.Problem([speo.OBJ_RASTRIGIN], o[None, :], M[None, :, :])
    X = torch.zeros((1, D), dtype=torch.float32, device=cuda)
    f = torch.zeros(1, dtype=torch.float32, device=cuda)
    # 16 floats per row at D = 1000 (8 chunks x 2 sums): 70000 rows need more
    # than the 2^20 the workspace holds; the call is refused before anything
    # is launched or read (X holds one row only)
    rc = _noalloc(speo, p, X, f, n=70000)
    assert rc == -1  # SPEO_ERR_ARG
    assert "workspace" in speo.last_error()
    # at the limit it runs: 65536 rows x 16 floats = 2^20
    Xl = torch.from_numpy(np.random.default_rng(2).uniform(-100, 100, (65536, D))
                          .astype(np.float32)).to(cuda)
    ref = torch.zeros(65536, dtype=torch.float32, device=cuda)
    got = torch.zeros(65536, dtype=torch.float32, device=cuda)
    speo.evaluate(p, Xl, ref)
    assert _noalloc(speo, p, Xl, got) == 0, speo.last_error()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
    p.close()
