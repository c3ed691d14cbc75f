"""CPU oracle of the warm-start network (test infrastructure only).

Restates, in numpy float64, the network that include/igap_ws.h documents:
4 layers of 3x3x3 convolutions with zero padding and per-layer dilation, 16
channels between layers, ReLU after layers 0..2, a linear last layer, mean-
square loss over the first ``cout`` output channels, and Adam (beta1 0.9,
beta2 0.999, eps 1e-8, bias-corrected).  The network has no counterpart in
the reference (IGAP4 starts every step from the previous solution,
src/interface.c:252-259,319); it is pinned here by finite differences and a
brute-force convolution (tests/test_ws_oracle.py).

``emulate=True`` rounds to bfloat16 (round to nearest even) exactly where the
GPU kernels hold BF16 data: the packed input, the weight copies, the
activations between layers and the back-propagated pre-activation gradients.
Sums are float64 (the GPU accumulates in float32).

Parameter layout (float32 vector, as igap_ws_get_params): per layer
W[27][16 out][16 in] with tap t = 9(dx+1) + 3(dy+1) + (dz+1), then b[16].
Arrays x, y, target are channels-last [V][c] over the grid, v = (i*ny+j)*nz+k.
"""
from __future__ import annotations

import numpy as np

NL = 4
CH = 16
LAYER_W = 27 * CH * CH
LAYER_P = LAYER_W + CH
B1, B2, EPS = 0.9, 0.999, 1e-8


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def taps(dil):
    for t in range(27):
        yield t, dil * (t // 9 - 1), dil * ((t // 3) % 3 - 1), dil * (t % 3 - 1)


def shift(a, dx, dy, dz):
    """s[x,y,z] = a[x+dx, y+dy, z+dz] with zeros outside the grid."""
    nx, ny, nz = a.shape[:3]
    s = np.zeros_like(a)
    xs = slice(max(0, -dx), min(nx, nx - dx))
    ys = slice(max(0, -dy), min(ny, ny - dy))
    zs = slice(max(0, -dz), min(nz, nz - dz))
    xd = slice(max(0, dx), min(nx, nx + dx))
    yd = slice(max(0, dy), min(ny, ny + dy))
    zd = slice(max(0, dz), min(nz, nz + dz))
    s[xs, ys, zs] = a[xd, yd, zd]
    return s


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    Ws, bs = [], []
    for l in range(NL):
        o = l * LAYER_P
        Ws.append(p[o:o + LAYER_W].reshape(27, CH, CH))
        bs.append(p[o + LAYER_W:o + LAYER_P])
    return Ws, bs


def conv(a, W, dil, sgn=1, transpose=False):
    """out[v][co] = sum_t sum_ci a[v + sgn off_t][ci] W[t][co][ci]
    (transpose: W[t][ci][co], i.e. the input-gradient form)."""
    out = np.zeros(a.shape[:3] + (CH,))
    for t, dx, dy, dz in taps(dil):
        Wt = W[t].T if not transpose else W[t]
        out += shift(a, sgn * dx, sgn * dy, sgn * dz) @ Wt
    return out


def forward(params, x, grid, cin, cout, dil, emulate=True):
    """y[V][cout] and the per-layer inputs (for backward)."""
    r = bf16 if emulate else (lambda z: np.asarray(z, dtype=np.float64))
    nx, ny, nz = grid
    Ws, bs = unpack(params)
    Wr = [r(W) for W in Ws]
    a0 = np.zeros((nx, ny, nz, CH))
    a0[..., :cin] = np.asarray(x, dtype=np.float64).reshape(nx, ny, nz, cin)
    acts = [r(a0)]
    for l in range(NL - 1):
        z = conv(acts[l], Wr[l], dil[l]) + bs[l]
        acts.append(r(np.maximum(z, 0.0)))
    z = conv(acts[-1], Wr[-1], dil[-1]) + bs[-1]
    y = z[..., :cout].reshape(-1, cout)
    return y, (acts, Wr)


def gradients(params, x, target, grid, cin, cout, dil, emulate=True):
    """(loss, grad) with grad in the parameter layout."""
    r = bf16 if emulate else (lambda z: np.asarray(z, dtype=np.float64))
    nx, ny, nz = grid
    V = nx * ny * nz
    y, (acts, Wr) = forward(params, x, grid, cin, cout, dil, emulate)
    e = y - np.asarray(target, dtype=np.float64).reshape(-1, cout)
    loss = float(np.sum(e * e) / (V * cout))
    dz = np.zeros((nx, ny, nz, CH))
    dz[..., :cout] = (2.0 * e / (V * cout)).reshape(nx, ny, nz, cout)
    dz = r(dz)
    grad = np.zeros(NL * LAYER_P)
    for l in range(NL - 1, -1, -1):
        gW = np.zeros((27, CH, CH))
        a2 = acts[l].reshape(-1, CH)
        d2 = dz.reshape(-1, CH)
        for t, dx, dy, dzz in taps(dil[l]):
            gW[t] = d2.T @ shift(acts[l], dx, dy, dzz).reshape(-1, CH)   # [co][ci]
        o = l * LAYER_P
        grad[o:o + LAYER_W] = gW.reshape(-1)
        grad[o + LAYER_W:o + LAYER_P] = d2.sum(axis=0)
        if l > 0:
            dA = conv(dz, np.transpose(Wr[l], (0, 2, 1)), dil[l], sgn=-1)
            dz = r(dA * (acts[l] > 0.0))
        del a2
    return loss, grad


def adam(params, grad, m, s, step, lr):
    """One bias-corrected Adam step (step counts from 1); returns (p, m, s)."""
    m = B1 * m + (1 - B1) * grad
    s = B2 * s + (1 - B2) * grad * grad
    c1, c2 = 1 - B1 ** step, 1 - B2 ** step
    p = params - lr * (m / c1) / (np.sqrt(s / c2) + EPS)
    return p, m, s


def conv_bruteforce(a, W, dil):
    """Triple-loop convolution of one layer (the definition, for pinning)."""
    nx, ny, nz, _ = a.shape
    out = np.zeros((nx, ny, nz, CH))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for t, dx, dy, dz in taps(dil):
                    ii, jj, kk = i + dx, j + dy, k + dz
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        out[i, j, k] += W[t] @ a[ii, jj, kk]
    return out
