"""Closed forms of the boundary data of Residual_iga for open (or uniform
periodic) quadratic B-spline patches, independent of the code under test.

* Dirichlet (dirichlet_iga, src/assmcore_iga.c:861-934): the affine data
  u_i = sum_d par[3 i + d] xi_d is reproduced exactly by quadratic
  B-splines whose control values are the data at the Greville abscissae
  (k[a+1] + k[a+2]) / 2 of the in-face directions (a periodic in-face
  direction included: its owned functions take the same values, unwrapped);
  the normal coordinate is
  the face's knot value (the far face evaluates at knot_end - 1e-15).  Both
  orders' layers take the face data; faces, orders and dofs are inserted in
  the reference's order (last one wins).
* Neumann (src/assmcore_iga.c:273-405 with the constant table of
  neumann_default_iga): in knot-space measure, order 0 puts
  -t (int N_a)(int N_b) on the face layer, order 1 puts -t (int N_a)(int N_b)
  N'_c(face) on the first two layers, with int_0^1 N_a = (k[a+3] - k[a]) / 3
  (uniform periodic: h) and the end derivatives of open quadratic B-splines
  N'_0(0) = -2/h0, N'_1(0) = 2/h0 (h0 the first knot span).  The reference
  evaluates the far face at the last knot itself, where evalN's half-open
  spans leave only the degenerate last function of each degree
  (src/meshcore_iga.c:348): there every derivative it returns is 0 (the
  compiled reference gives N'_{n-1}(1) = N'_{n-2}(1) = 0, and +-2/h1 at
  1 - 1e-15), so the far face's order-1 load is zero.
"""
import numpy as np

FACE_DIRS = {0: (1, 2, 0), 1: (0, 2, 1), 2: (0, 1, 2)}   # face // 2 -> (in-face d0, d1, normal)


def greville(kv, nfun):
    kv = np.asarray(kv, dtype=np.float64)
    return 0.5 * (kv[1:nfun + 1] + kv[2:nfun + 2])


def integrals(kv, nfun, periodic):
    kv = np.asarray(kv, dtype=np.float64)
    if periodic:
        return np.full(nfun, kv[3] - kv[2])
    return (kv[3:nfun + 3] - kv[:nfun]) / 3.0


def _put(arr, dn, layer, plane):
    """arr[i, j, k] over the owned grid: add plane (shape nb[d0] x nb[d1], the
    in-face directions d0 < d1) at index ``layer`` of the normal direction dn."""
    idx = [slice(None)] * 3
    idx[dn] = layer
    arr[tuple(idx)] += plane


def dirichlet(knots, nelem, periodic, bc_type, par):
    """(mask, g) over the dofs 3 cp + i of the owned grid."""
    nb = [n if p else n + 2 for n, p in zip(nelem, periodic)]
    mask = np.zeros(tuple(nb) + (3,), dtype=np.int32)
    g = np.zeros(tuple(nb) + (3,))
    for f in range(6):
        d0, d1, dn = FACE_DIRS[f // 2]
        if periodic[dn]:
            continue
        even = f % 2 == 0
        z = 0.0 if even else float(knots[dn][-1]) - 1e-15
        gam0, gam1 = greville(knots[d0], nb[d0]), greville(knots[d1], nb[d1])
        for order in (0, 1):
            layer = order if even else nb[dn] - 1 - order
            for i in range(3):
                if bc_type[6 * f + 3 * order + i] != 0:
                    continue
                vals = par[3 * i + d0] * gam0[:, None] + par[3 * i + d1] * gam1[None, :] + par[3 * i + dn] * z
                idx = [slice(None)] * 3
                idx[dn] = layer
                sl = tuple(idx) + (i,)
                mask[sl] = 1
                g[sl] = vals
    return mask.reshape(-1), g.reshape(-1)


def neumann(knots, nelem, periodic, bc_type, t):
    """The Neumann load F (the residual's U-independent part) over the dofs."""
    nb = [n if p else n + 2 for n, p in zip(nelem, periodic)]
    F = np.zeros(tuple(nb) + (3,))
    for f in range(6):
        d0, d1, dn = FACE_DIRS[f // 2]
        if periodic[dn]:
            continue
        even = f % 2 == 0
        kn = np.asarray(knots[dn], dtype=np.float64)
        if even:
            h = kn[3] - kn[2]
            layers, dN = (0, 1), (-2.0 / h, 2.0 / h)
        else:
            layers, dN = (nb[dn] - 1, nb[dn] - 2), (0.0, 0.0)
        outer = np.outer(integrals(knots[d0], nb[d0], periodic[d0]), integrals(knots[d1], nb[d1], periodic[d1]))
        for order in (0, 1):
            for i in range(3):
                if bc_type[6 * f + 3 * order + i] != 1:
                    continue
                tv = t[6 * f + 3 * order + i]
                if tv == 0.0:
                    continue
                if order == 0:
                    _put(F[..., i], dn, layers[0], -tv * outer)
                else:
                    for layer, dv in zip(layers, dN):
                        _put(F[..., i], dn, layer, -tv * dv * outer)
    return F.reshape(-1)


def static_jet_of_gradient(A):
    """The 30-slot jet of a displacement with constant physical gradient A[k][d]
    (value and second derivatives zero)."""
    u = np.zeros(30)
    for k in range(3):
        for d in range(3):
            u[10 * k + 1 + d] = A[k][d]
    return u
