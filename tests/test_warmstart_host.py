"""Host-side pieces of the warm start (no GPU): the 1-D coarse-space
operators of the correction filter and the filter itself."""
import numpy as np

from igap4_amd import warmstart as ws


def _open_knots(n, graded=False):
    t = np.arange(1, n) / n
    if graded:
        t = t ** 1.3
    return np.concatenate([[0.0] * 3, t, [1.0] * 3])


def test_bspline_rows_partition_and_closed_form():
    """Open quadratic B-splines: partition of unity everywhere, and on a
    uniform interior span the three values are the textbook (1-u)^2/2,
    (1+2u-2u^2)/2, u^2/2."""
    n = 8
    kv = _open_knots(n)
    x = np.linspace(0, 1, 101)
    N = ws._bs2_rows(kv, n + 2, x)
    assert np.abs(N.sum(1) - 1).max() < 1e-14
    u = 0.3
    row = ws._bs2_rows(kv, n + 2, (3 + u) / n)[0]
    assert np.allclose(row[3:6], [(1 - u) ** 2 / 2, (1 + 2 * u - 2 * u * u) / 2, u * u / 2], atol=1e-14)


def test_coarse_ops_refinement_and_interpolation():
    """P expresses every coarse B-spline in the fine basis (knot insertion,
    also on a graded knot vector) and T interpolates: T P = I."""
    for graded in (False, True):
        n = 12
        kf = _open_knots(n, graded)
        P, T = ws.coarse_ops_1d(kf, n)
        kc = np.concatenate([[0.0] * 3, kf[2 + 2 * np.arange(1, n // 2)], [1.0] * 3])
        x = np.linspace(0, 1, 211)
        Nf, Nc = ws._bs2_rows(kf, n + 2, x), ws._bs2_rows(kc, n // 2 + 2, x)
        assert np.abs(Nf @ P - Nc).max() < 1e-13
        assert np.abs(T @ P - np.eye(n // 2 + 2)).max() < 1e-13


class _Mesh:
    def __init__(self, nelem, periodic=(0, 0, 0)):
        self.nelem = nelem
        self.periodic = periodic
        self.nb = tuple(n + 2 for n in nelem)
        self.knots = [_open_knots(n) for n in nelem]


def test_coarse_filter_projection():
    """The filter is a projection (idempotent), keeps fields of the coarse
    space unchanged and cuts a checkerboard (highest-frequency) field to less
    than 40% of its norm (an interpolatory, not an orthogonal, projection:
    the open-knot end layers keep part of it); an odd direction is left
    alone."""
    m = _Mesh((8, 4, 5))
    f = ws.CoarseFilter(m)
    rng = np.random.default_rng(0)
    c = rng.standard_normal(int(np.prod(m.nb)) * 3)
    once = f(c)
    assert np.abs(f(once) - once).max() < 1e-12
    # a field in the coarse space: P applied to coarse values, per direction
    Px = ws.coarse_ops_1d(m.knots[0], 8)[0]
    Py = ws.coarse_ops_1d(m.knots[1], 4)[0]
    cc = rng.standard_normal((Px.shape[1], Py.shape[1], m.nb[2], 3))
    fine = np.einsum("ia,jb,abkc->ijkc", Px, Py, cc).reshape(-1)
    assert np.abs(f(fine) - fine).max() < 1e-12
    i, j, k = np.meshgrid(*[np.arange(n) for n in m.nb], indexing="ij")
    chk = np.repeat(((-1.0) ** (i + j))[..., None], 3, axis=3).reshape(-1)
    assert np.linalg.norm(f(chk)) < 0.4 * np.linalg.norm(chk)


def test_residual_shifted_target_gives_the_residual_gradient():
    """The network trained on the shifted target descends the linearised
    residual of its candidate: the mean square gradient at the shifted target
    (the network oracle, float64) times N / (2 alpha) equals the central
    difference of |rq + K (km y(p))|^2 / 2 in the last layer's parameters, for a
    symmetric K; the target shifted the other way (a planted sign error)
    gives the opposite gradient and fails."""
    from oracle import convnet as cn
    rng = np.random.default_rng(7)
    grid, cin, cout, dil = (4, 3, 3), 9, 3, (1, 2, 1, 1)
    V = int(np.prod(grid))
    n = V * cout
    p = 0.2 * rng.standard_normal(cn.NL * cn.LAYER_P)
    x = rng.standard_normal((V, cin))
    B = rng.standard_normal((n, n))
    K = B @ B.T / n + np.eye(n)
    km = (rng.random(n) > 0.2).astype(np.float64)
    rq = rng.standard_normal(n)
    alpha = 0.37

    def y_of(q):
        return cn.forward(q, x, grid, cin, cout, dil, emulate=False)[0].reshape(-1)

    def loss(q):
        r = rq + K @ (km * y_of(q))
        return 0.5 * float(r @ r)

    y = y_of(p)
    r, t = ws.residual_shifted_target(y, rq, lambda v: K @ v, km, alpha)
    assert np.allclose(r, rq + K @ (km * y), rtol=0, atol=1e-12)
    g = cn.gradients(p, x, t, grid, cin, cout, dil, emulate=False)[1] * n / (2 * alpha)
    # a direction in the last (linear) layer: the loss is then a quadratic
    # in the step, and its central difference is exact up to rounding
    d = np.zeros(p.size)
    d[(cn.NL - 1) * cn.LAYER_P:] = rng.standard_normal(cn.LAYER_P)
    eps = 1e-4
    fd = (loss(p + eps * d) - loss(p - eps * d)) / (2 * eps)
    assert abs(g @ d - fd) <= 1e-6 * abs(fd), (g @ d, fd)
    t_wrong = 2 * y - t
    g_wrong = cn.gradients(p, x, t_wrong, grid, cin, cout, dil, emulate=False)[1] * n / (2 * alpha)
    assert abs(g_wrong @ d - fd) > 0.5 * abs(fd)
