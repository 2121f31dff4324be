"""Per-row error bounds of FP32 (and split-precision) objective values against
the FP64 oracle.

A weighted-sphere objective (elliptic, bent cigar, discus: sum_j w_j z_j^2)
computed from a rotation z = M (x - o) in a precision with unit roundoff u has
relative error up to about u kappa, with the row's condition number

    kappa = 2 sum_j w_j |z_j| S_j / sum_j w_j z_j^2,   S_j = sum_k |M_jk| |x_k - o_k|,

because each z_j carries an error of order u S_j.  A row whose heavily
weighted coordinate cancels (z_j small against the terms it sums) has a large
kappa for any implementation; discus, whose value is 1e6 z_0^2 plus the rest,
has such rows often.  The kernels compute a column that outweighs all others
together (discus's z_0; elliptic and bent cigar at small D) in FP64, which
leaves every FP32 row within `base`: FP32 rows are held to `base`.  Rows of
the 16-bit precisions (u > 2^-24) are held to max(base, 4 u kappa); for the
other objectives the bound is `base`.

Schwefel (objective 8) sums g(v_j), v_j = z_j + 420.97, and rows are held to
`base`.  The function jumps (by up to about 400) where |v_j| crosses 1000,
1500, ...: a row with a coordinate within d_j = 8 u (S_j + |v_j|), a bound on
the FP32 error of v_j, of such a point has no well-defined FP32 value and is
left out (schwefel_rows).
"""
import numpy as np

U_FP32 = 2.0 ** -24


def sphere_weights(obj, D):
    j = np.arange(D, dtype=np.float64)
    if obj == 4:
        return np.where(j == 0, 1.0, 1.0e6)
    if obj == 5:
        return np.where(j == 0, 1.0e6, 1.0)
    return 10.0 ** (6.0 * j / (D - 1))


def row_rtol(obj, X, o, M, base=1e-5, u=U_FP32):
    X = np.asarray(X, np.float64)
    if obj not in (1, 4, 5) or u <= U_FP32:
        return np.full(X.shape[0], base)
    a = X - o[None, :]
    z = a @ M.T
    S = np.abs(a) @ np.abs(M).T
    w = sphere_weights(obj, o.size)[None, :]
    num, den = 2.0 * (w * np.abs(z) * S).sum(1), (w * z * z).sum(1)
    kappa = np.divide(num, den, out=np.zeros_like(num), where=den > 0)  # x = o: exact
    return np.maximum(base, 4.0 * u * kappa)


def composition_rtol(objs, X, os_, Ms, base=1e-5, u=U_FP32):
    """Every fit_c = lambda g_c + bias_c >= 0 and f is their convex combination,
    so f's relative error is at most the largest of the components'."""
    return np.max([row_rtol(ob, X, os_[c], Ms[c], base, u) for c, ob in enumerate(objs)], axis=0)


def assert_rows(f, ref, rtol):
    rel = np.abs(np.asarray(f, np.float64) - ref) / np.abs(ref)
    bad = np.flatnonzero(rel > rtol)
    assert bad.size == 0, [(int(i), float(rel[i]), float(rtol[i])) for i in bad[:5]]


def schwefel_rows(X, o, M, rate, base=1e-5, u=U_FP32):
    """(per-row rtol, rows within rounding distance of a jump) for Schwefel."""
    a = (np.asarray(X, np.float64) - o[None, :]) * rate
    v = a @ M.T + 4.209687462275036e+002
    d = 8.0 * u * (np.abs(a) @ np.abs(M).T + np.abs(v))
    av = np.abs(v)
    dist = np.abs(av - 500.0 * np.rint(av / 500.0))
    near = ((av >= 1000.0 - d) & (dist <= d)).any(1)
    return np.full(av.shape[0], base), near
