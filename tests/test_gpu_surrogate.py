"""GPU parity of the surrogate MLP kernels against the CPU oracle
(oracle/surrogate.py, pinned by finite differences): forward and gradients
with the same BF16 operand rounding, Adam, Spearman; the engine's screening
path; and what the trained network gains over plain DE at an equal number of
true evaluations."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 256


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


class Net:
    def __init__(self, speo, D, batch, seed=3):
        self.L = speo.lib()
        self.h = ctypes.c_void_p()
        speo.check(self.L.speo_mlp_create(D, batch, seed, ctypes.byref(self.h)))
        self.D, self.Kp = D, (D + 31) // 32 * 32
        self.n = int(self.L.speo_mlp_num_params(self.h))
        assert self.n == H * self.Kp + H + H * H + H + H + 4

    def flat(self):
        out = np.zeros(self.n, np.float32)
        self.L.speo_mlp_params(self.h, out.ctypes.data_as(ctypes.c_void_p), None)
        return out

    def set_flat(self, f):
        f = np.ascontiguousarray(f, np.float32)
        self.L.speo_mlp_params(self.h, None, f.ctypes.data_as(ctypes.c_void_p))

    def to_dict(self, f):
        D, Kp = self.D, self.Kp
        o = 0
        W1 = f[o:o + H * Kp].reshape(H, Kp)[:, :D].astype(np.float64); o += H * Kp
        b1 = f[o:o + H].astype(np.float64); o += H
        W2 = f[o:o + H * H].reshape(H, H).astype(np.float64); o += H * H
        b2 = f[o:o + H].astype(np.float64); o += H
        w3 = f[o:o + H].astype(np.float64); o += H
        b3 = f[o:o + 1].astype(np.float64)
        return {"W1": W1, "b1": b1, "W2": W2, "b2": b2, "w3": w3, "b3": b3}

    def close(self):
        self.L.speo_mlp_destroy(self.h)


def _random_params(net, rng):
    f = net.flat()
    f = f + rng.normal(0, 0.05, f.size).astype(np.float32)
    f[: H * net.Kp].reshape(H, net.Kp)[:, net.D:] = 0.0
    return f


def test_mlp_forward_matches_oracle(cuda, speo):
    import surrogate as SO
    import torch
    rng = np.random.default_rng(0)
    D, n = 100, 300
    net = Net(speo, D, 512)
    f = _random_params(net, rng)
    net.set_flat(f)
    X = rng.uniform(-100, 100, (n, 104)).astype(np.float32)
    y = torch.zeros(n, dtype=torch.float32, device=cuda)
    Xd = _t(X, cuda)          # keep a reference: a freed temporary may be reused
    speo.check(net.L.speo_mlp_forward(net.h, ctypes.c_void_p(Xd.data_ptr()), 104, n,
                                      ctypes.c_void_p(y.data_ptr()), None))
    torch.cuda.synchronize()
    ref, _ = SO.forward(net.to_dict(f), X[:, :D].astype(np.float64), round_bf16=True)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(y.cpu().numpy(), ref, atol=2e-4 * scale)
    # the fused single-kernel inference path gives the same predictions
    y2 = torch.zeros(n, dtype=torch.float32, device=cuda)
    speo.check(net.L.speo_mlp_infer(net.h, ctypes.c_void_p(Xd.data_ptr()), 104, n,
                                    ctypes.c_void_p(y2.data_ptr()), None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(y2.cpu().numpy(), ref, atol=2e-4 * scale)
    net.close()


@pytest.mark.parametrize("D", [10, 30, 64, 100, 200])
def test_mlp_infer_dims(cuda, speo, D):
    import surrogate as SO
    import torch
    rng = np.random.default_rng(D)
    n = 130
    net = Net(speo, D, 256)
    f = _random_params(net, rng)
    net.set_flat(f)
    X = rng.uniform(-100, 100, (n, D)).astype(np.float32)
    y = torch.zeros(n, dtype=torch.float32, device=cuda)
    Xd = _t(X, cuda)
    speo.check(net.L.speo_mlp_infer(net.h, ctypes.c_void_p(Xd.data_ptr()), D, n,
                                    ctypes.c_void_p(y.data_ptr()), None))
    torch.cuda.synchronize()
    ref, _ = SO.forward(net.to_dict(f), X.astype(np.float64), round_bf16=True)
    np.testing.assert_allclose(y.cpu().numpy(), ref, atol=2e-4 * np.abs(ref).max())
    net.close()


def test_mlp_gradients_match_oracle(cuda, speo):
    import surrogate as SO
    rng = np.random.default_rng(1)
    D, n = 100, 256
    net = Net(speo, D, 512)
    f = _random_params(net, rng)
    net.set_flat(f)
    X = rng.uniform(-100, 100, (n, D)).astype(np.float32)
    t = rng.normal(3.0, 2.0, n).astype(np.float32)
    mu, sd = 3.0, 2.0
    loss = np.zeros(1, np.float32)
    g = np.zeros(net.n, np.float32)
    Xd, td = _t(X, cuda), _t(t, cuda)
    speo.check(net.L.speo_mlp_loss_grad(net.h, ctypes.c_void_p(Xd.data_ptr()), D, n,
                                        ctypes.c_void_p(td.data_ptr()), mu, sd,
                                        loss.ctypes.data_as(ctypes.c_void_p),
                                        g.ctypes.data_as(ctypes.c_void_p)))
    p = net.to_dict(f)
    L, gr = SO.loss_and_grads(p, X.astype(np.float64), (t.astype(np.float64) - mu) / sd,
                              round_bf16=True)
    assert loss[0] == pytest.approx(L, rel=1e-3)
    gd = net.to_dict(g)
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        scale = np.abs(gr[k]).max()
        np.testing.assert_allclose(gd[k], gr[k], atol=3e-3 * scale, err_msg=k)
    net.close()


def test_adam_step_matches_oracle(cuda, speo):
    import surrogate as SO
    rng = np.random.default_rng(2)
    D, n = 30, 128
    net = Net(speo, D, 256)
    f = _random_params(net, rng)
    net.set_flat(f)
    X = rng.uniform(-100, 100, (n, D)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    g = np.zeros(net.n, np.float32)
    Xd, td = _t(X, cuda), _t(t, cuda)
    for step in range(1, 4):
        before = net.flat().astype(np.float64)
        speo.check(net.L.speo_mlp_loss_grad(net.h, ctypes.c_void_p(Xd.data_ptr()), D, n,
                                            ctypes.c_void_p(td.data_ptr()), 0.0, 1.0, None,
                                            g.ctypes.data_as(ctypes.c_void_p)))
        if step == 1:
            m = {"a": np.zeros(net.n)}
            v = {"a": np.zeros(net.n)}
        p2, m, v = SO.adam_step({"a": before}, {"a": g.astype(np.float64)}, m, v, step, lr=1e-3)
        speo.check(net.L.speo_mlp_adam_step(net.h, 1e-3))
        after = net.flat()
        np.testing.assert_allclose(after, p2["a"], rtol=1e-5, atol=1e-7)
    net.close()


def _batch_slots(oracle, seed, rnd, step, batch, n_valid):
    """The archive slots of a training minibatch: word b % 4 of Philox block
    {b / 4, round, step, 33} as u01, times n_valid in FP32, truncated, clamped,
    held-out slots (slot % 8 == 7) moved down by one."""
    b = np.arange(batch)
    w = oracle.philox_np(seed, b >> 2, rnd, step, 33)
    word = np.choose(b & 3, w).astype(np.uint32)
    u = ((word >> 8).astype(np.float64) + 1.0) * 2.0 ** -24
    slot = (u.astype(np.float32) * np.float32(n_valid)).astype(np.int64)
    slot = np.minimum(slot, n_valid - 1)
    return np.where(slot % 8 == 7, slot - 1, slot)


@pytest.mark.parametrize("D", [100, 30])
def test_fused_train_step_equals_gather_grad_adam(cuda, speo, oracle, D):
    # the engine's two-kernel training step (archive gather, gradients, Adam,
    # BF16 refresh) against the slots drawn on the host, the gradient entry
    # point and the Adam entry point: the same parameters, bit for bit, over
    # three steps; the gradients themselves are pinned to the oracle above
    rng = np.random.default_rng(7)
    Kp, n_valid, batch, seed, rnd = (D + 31) // 32 * 32, 3000, 500, 0xABCDEF12345, 5
    AX = np.zeros((4096, Kp), np.float32)
    AX[:, :D] = rng.uniform(-100, 100, (4096, D))
    Ay = rng.normal(4.0, 1.5, 4096).astype(np.float32)
    a, b = Net(speo, D, 512), Net(speo, D, 512)
    f = _random_params(a, rng)
    a.set_flat(f)
    b.set_flat(f)
    AXd, Ayd = _t(AX, cuda), _t(Ay, cuda)
    for step in range(3):
        la = np.zeros(1, np.float32)
        speo.check(a.L.speo_mlp_train_step(a.h, ctypes.c_void_p(AXd.data_ptr()),
                                           ctypes.c_void_p(Ayd.data_ptr()), n_valid, seed, rnd,
                                           step, batch, 4.0, 1.5, 1e-3,
                                           la.ctypes.data_as(ctypes.c_void_p)))
        slots = _batch_slots(oracle, seed, rnd, step, batch, n_valid)
        assert slots.min() >= 0 and slots.max() < n_valid and not np.any(slots % 8 == 7)
        Xb, tb = _t(AX[slots, :D], cuda), _t(Ay[slots], cuda)
        lb = np.zeros(1, np.float32)
        speo.check(b.L.speo_mlp_loss_grad(b.h, ctypes.c_void_p(Xb.data_ptr()), D, batch,
                                          ctypes.c_void_p(tb.data_ptr()), 4.0, 1.5,
                                          lb.ctypes.data_as(ctypes.c_void_p), None))
        speo.check(b.L.speo_mlp_adam_step(b.h, 1e-3))
        assert la[0] == lb[0]
        np.testing.assert_array_equal(a.flat(), b.flat())
    assert not np.array_equal(a.flat(), f)
    a.close()
    b.close()


def test_spearman_kernel(cuda, speo):
    import surrogate as SO
    import torch
    rng = np.random.default_rng(4)
    for n in (2, 17, 1024, 4096):
        a = rng.normal(size=n).astype(np.float32)
        b = (a + rng.normal(scale=1.0, size=n)).astype(np.float32)
        a[::9] = a[0]
        out = torch.zeros(1, dtype=torch.float32, device=cuda)
        ad, bd = _t(a, cuda), _t(b, cuda)   # both alive while the kernel is enqueued
        speo.check(speo.lib().speo_spearman(ctypes.c_void_p(ad.data_ptr()),
                                            ctypes.c_void_p(bd.data_ptr()), n,
                                            ctypes.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        assert out.item() == pytest.approx(SO.spearman(a, b), abs=1e-5)


def _problems(speo, oracle, D, objs):
    out = []
    for k, obj in enumerate(objs):
        o, M = oracle.gen_problem(500 + k, D)
        out.append(speo.Problem([obj], o[None, :], M[None, :, :]))
    return out


def test_inactive_surrogate_equals_plain(cuda, speo, oracle):
    # rho_min above any correlation: candidate 0 (the plain trial) every time,
    # while candidates are still generated, scored and the network trained.
    probs = _problems(speo, oracle, 20, [1, 2])
    kw = dict(seed=9, rows_per_block=32, interval=10, migration_size=2)
    plain = speo.Engine(probs, [64] * 4, [0, 1, 0, 1], mode=speo.MODE_PLAIN, **kw)
    surr = speo.Engine(probs, [64] * 4, [0, 1, 0, 1], mode=speo.MODE_PLAIN,
                       surrogate_candidates=8, surrogate_train_interval=4, surrogate_archive=2048,
                       surrogate_steps=5, surrogate_batch=256, surrogate_rho_min=2.0, **kw)
    plain.run(24)
    surr.run(24)
    Xp, fp = plain.read()
    Xs, fs = surr.read()
    np.testing.assert_array_equal(Xp, Xs)
    np.testing.assert_array_equal(fp, fs)
    rho, act, loss = surr.surrogate_stats()
    assert (act == 0).all()
    assert surr.stats()["surrogate_rounds"] >= 5
    assert surr.stats()["evaluations"] == plain.stats()["evaluations"]


def test_surrogate_gain_over_plain(cuda, speo, oracle):
    """At an equal number of true evaluations, screening 8 candidates with the
    trained network reaches a lower error than plain DE (elliptic, D=30)."""
    probs = _problems(speo, oracle, 30, [1])
    kw = dict(seed=21, rows_per_block=32, interval=20, migration_size=2)
    sizes, pidx = [128] * 4, [0] * 4
    plain = speo.Engine(probs, sizes, pidx, mode=speo.MODE_PLAIN, **kw)
    surr = speo.Engine(probs, sizes, pidx, mode=speo.MODE_PLAIN, surrogate_candidates=8,
                       surrogate_train_interval=10, surrogate_archive=16384, surrogate_steps=60,
                       surrogate_batch=512, surrogate_lr=1e-3, **kw)
    G = 200
    plain.run(G)
    surr.run(G)
    _, _, fp = plain.best()
    _, _, fs = surr.best()
    rho, act, loss = surr.surrogate_stats()
    print(f"\nplain best {fp.min():.4e}  surrogate best {fs.min():.4e}  rho {rho[0]:.3f} "
          f"active {act[0]} loss {loss[0]:.4f}")
    assert plain.stats()["evaluations"] == surr.stats()["evaluations"]
    assert rho[0] > 0.5
    assert fs.min() < fp.min()


def test_surrogate_rounds_under_graphs_equal_eager(cuda, speo, oracle):
    # once the networks screen and the archives are full, each round (the
    # generations up to a training point, a migration among them or not, and
    # the training) is captured once and replayed: the same bits as eager
    # launches, generation counter, archive heads and Adam steps included
    D, C = 30, 4
    probs = []
    for k, obj in enumerate((speo.OBJ_RASTRIGIN, speo.OBJ_ELLIPTIC)):
        o, M = oracle.gen_problem(300 + k, D)
        probs.append(speo.Problem([obj], o[None, :], M[None, :, :]))
    kw = dict(seed=17, interval=10, migration_size=2, mode=speo.MODE_PLAIN, surrogate_candidates=C,
              surrogate_train_interval=5, surrogate_archive=1024, surrogate_steps=5,
              surrogate_batch=256, surrogate_rho_min=-1.0)
    eager = speo.Engine(probs, [64] * 8, [i % 2 for i in range(8)], use_graph=False, **kw)
    graph = speo.Engine(probs, [64] * 8, [i % 2 for i in range(8)], use_graph=True, **kw)
    for chunk in (12, 3, 25):  # rounds start at generations 15, 20, ..., 35
        eager.run(chunk)
        graph.run(chunk)
        Xe, fe = eager.read()
        Xg, fg = graph.read()
        np.testing.assert_array_equal(Xe, Xg)
        np.testing.assert_array_equal(fe, fg)
    se, sg = eager.stats(), graph.stats()
    assert se["surrogate_rounds"] == sg["surrogate_rounds"] == 8
    assert sg["graph_launches"] >= 4 and se["graph_launches"] == 0
    rho_e, act_e, _ = eager.surrogate_stats()
    rho_g, act_g, _ = graph.surrogate_stats()
    np.testing.assert_array_equal(rho_e, rho_g)
    for e in (eager, graph):
        e.close()
    for p in probs:
        p.close()
