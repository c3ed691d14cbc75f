"""Pins the warm-start network oracle (oracle/convnet.py) without a GPU:
the vectorised convolution equals the triple-loop definition, the analytic
gradient equals central finite differences of the loss, the BF16 rounding is
round-to-nearest-even, and Adam's first step has its closed form."""
import numpy as np
import pytest

from oracle import convnet as cn

GRID = (4, 3, 3)
CIN, COUT = 3, 2
DIL = (1, 2, 1, 1)


def random_params(rng, scale=0.3):
    p = np.zeros(cn.NL * cn.LAYER_P)
    for l in range(cn.NL):
        o = l * cn.LAYER_P
        W = rng.standard_normal((27, cn.CH, cn.CH)) * scale / np.sqrt(27 * cn.CH)
        if l == 0:
            W[:, :, CIN:] = 0.0
        if l == cn.NL - 1:
            W[:, COUT:, :] = 0.0
        p[o:o + cn.LAYER_W] = W.reshape(-1)
        p[o + cn.LAYER_W:o + cn.LAYER_P] = rng.standard_normal(cn.CH) * 0.1
    return p


def test_conv_equals_definition():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(GRID + (cn.CH,))
    W = rng.standard_normal((27, cn.CH, cn.CH))
    for dil in (1, 2):
        assert np.allclose(cn.conv(a, W, dil), cn.conv_bruteforce(a, W, dil), atol=1e-12)


def test_input_gradient_is_the_adjoint_convolution():
    """<conv(a, W), d> = <a, conv_T(d, W)> for the sgn=-1 transposed form."""
    rng = np.random.default_rng(2)
    a = rng.standard_normal(GRID + (cn.CH,))
    d = rng.standard_normal(GRID + (cn.CH,))
    W = rng.standard_normal((27, cn.CH, cn.CH))
    lhs = np.sum(cn.conv(a, W, 2) * d)
    rhs = np.sum(a * cn.conv(d, np.transpose(W, (0, 2, 1)), 2, sgn=-1))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = random_params(rng)
    V = int(np.prod(GRID))
    x = rng.standard_normal((V, CIN))
    t = rng.standard_normal((V, COUT))
    loss, g = cn.gradients(p, x, t, GRID, CIN, COUT, DIL, emulate=False)
    h = 1e-6
    # entries in every layer, weights and biases, where the gradient is not tiny
    checked = 0
    for l in range(cn.NL):
        o = l * cn.LAYER_P
        cand = [o + i for i in range(cn.LAYER_P) if abs(g[o + i]) > 1e-6]
        for idx in rng.choice(cand, size=min(6, len(cand)), replace=False):
            pp, pm = p.copy(), p.copy()
            pp[idx] += h
            pm[idx] -= h
            fp, _ = cn.gradients(pp, x, t, GRID, CIN, COUT, DIL, emulate=False)
            fm, _ = cn.gradients(pm, x, t, GRID, CIN, COUT, DIL, emulate=False)
            fd = (fp - fm) / (2 * h)
            assert abs(fd - g[idx]) < 1e-6 * max(1.0, abs(g[idx])) + 1e-9, (l, idx, fd, g[idx])
            checked += 1
    assert checked >= 16


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -(one + 3 * ulp / 2), 3.1415926])
    r = cn.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                  # tie to even (mantissa 0)
    assert r[2] == one + 2 * ulp        # tie to even (mantissa 2)
    assert r[3] == 1.0
    assert r[4] == -(one + 2 * ulp)
    assert abs(r[5] - 3.140625) < 1e-12


def test_adam_first_step_closed_form():
    """With zero moments, step 1 moves every parameter by lr * sign(g) (up to eps)."""
    g = np.array([0.5, -2.0, 1e-3])
    p, m, s = cn.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-5)
    assert np.allclose(m, 0.1 * g) and np.allclose(s, 0.001 * g * g)


@pytest.mark.parametrize("emulate", [False, True])
def test_untrained_last_layer_zero_gives_zero_output(emulate):
    rng = np.random.default_rng(4)
    p = random_params(rng)
    o = (cn.NL - 1) * cn.LAYER_P
    p[o:o + cn.LAYER_P] = 0.0
    V = int(np.prod(GRID))
    y, _ = cn.forward(p, rng.standard_normal((V, CIN)), GRID, CIN, COUT, DIL, emulate=emulate)
    assert np.all(y == 0.0)
