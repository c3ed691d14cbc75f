"""Learned warm start on the GPU (include/igap_ws.h): the BF16 matrix-core
layout, forward values and gradients against the CPU oracle with the same
BF16 rounding (oracle/convnet.py), Adam against the oracle's Adam, training
that lowers the loss, and the load-step ablation of the learned term against
the quadratic extrapolation it corrects."""
import numpy as np
import pytest

from oracle import convnet as cn

pytestmark = pytest.mark.gpu

GRID = (10, 6, 5)
CIN, COUT = 9, 3
DIL = (1, 2, 4, 1)


def random_params(rng):
    p = np.zeros(cn.NL * cn.LAYER_P, dtype=np.float32)
    for l in range(cn.NL):
        o = l * cn.LAYER_P
        fan = 27 * (CIN if l == 0 else cn.CH)
        W = rng.uniform(-1, 1, (27, cn.CH, cn.CH)) * np.sqrt(6.0 / fan)
        if l == 0:
            W[:, :, CIN:] = 0.0
        if l == cn.NL - 1:
            W[:, COUT:, :] = 0.0
        p[o:o + cn.LAYER_W] = W.reshape(-1)
        p[o + cn.LAYER_W:o + cn.LAYER_P] = rng.standard_normal(cn.CH) * 0.05
    return p


@pytest.fixture(scope="module")
def ws(igap):
    from igap4_amd import warmstart
    return warmstart


def test_bf16_mfma_layout(ws):
    """Result register r of lane l holds D[4(l>>4)+r][l&15] (f32 16x16 rule)."""
    out = ws.selftest_mfma_bf16()
    A = np.array([[1 + i + 0.25 * k for k in range(16)] for i in range(16)])
    B = np.array([[2 - 0.5 * n + k for n in range(16)] for k in range(16)])
    D = A @ B
    exp = np.array([[D[4 * (l >> 4) + r, l & 15] for l in range(64)] for r in range(4)])
    assert np.array_equal(out, exp.astype(np.float32))


def test_forward_matches_oracle(ws):
    rng = np.random.default_rng(0)
    p = random_params(rng)
    V = int(np.prod(GRID))
    x = rng.standard_normal((V, CIN)).astype(np.float32)
    with ws.WarmStartNet(GRID, CIN, COUT, DIL) as net:
        net.set_params(p)
        y = net.forward(x)
    y_ref, _ = cn.forward(p, x, GRID, CIN, COUT, DIL, emulate=True)
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    print(f"  forward rel err {err:.2e}")
    # the same BF16 roundings on both sides: only the FP32 summation order
    # differs (measured ~1e-7); a dropped tap or a wrong dilation is O(1)
    assert err < 1e-5


def test_gradients_match_oracle(ws):
    rng = np.random.default_rng(1)
    p = random_params(rng)
    V = int(np.prod(GRID))
    x = rng.standard_normal((V, CIN)).astype(np.float32)
    t = rng.standard_normal((V, COUT)).astype(np.float32)
    with ws.WarmStartNet(GRID, CIN, COUT, DIL) as net:
        net.set_params(p)
        loss, g = net.gradients(x, t)
    loss_ref, g_ref = cn.gradients(p, x, t, GRID, CIN, COUT, DIL, emulate=True)
    assert abs(loss - loss_ref) < 1e-3 * loss_ref
    for l in range(cn.NL):
        o = l * cn.LAYER_P
        for name, sl in (("W", slice(o, o + cn.LAYER_W)), ("b", slice(o + cn.LAYER_W, o + cn.LAYER_P))):
            a, b = g[sl], g_ref[sl]
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print(f"  layer {l} d{name} rel err {err:.2e}")
            assert err < 1e-5


def test_adam_step_matches_oracle(ws):
    rng = np.random.default_rng(2)
    p = random_params(rng)
    V = int(np.prod(GRID))
    x = rng.standard_normal((V, CIN)).astype(np.float32)
    t = rng.standard_normal((V, COUT)).astype(np.float32)
    with ws.WarmStartNet(GRID, CIN, COUT, DIL) as net:
        net.set_params(p)
        _, g = net.gradients(x, t)
        net.train_step(x, t, 1e-3)
        p1 = net.get_params()
    z = np.zeros_like(p, dtype=np.float64)
    p_ref, _, _ = cn.adam(p.astype(np.float64), g.astype(np.float64), z, z, 1, 1e-3)
    assert np.max(np.abs(p1 - p_ref)) < 1e-6


def test_training_lowers_the_loss(ws):
    """Fit a smooth linear map of the inputs (what a second-order
    extrapolation is): the loss falls by more than 10x."""
    rng = np.random.default_rng(3)
    V = int(np.prod(GRID))
    xs, ts = [], []
    for _ in range(8):
        x = rng.standard_normal((V, CIN)).astype(np.float32)
        xs.append(x)
        ts.append((0.5 * (x[:, 0:3] - x[:, 3:6])).astype(np.float32))
    with ws.WarmStartNet(GRID, CIN, COUT, DIL, seed=5) as net:
        first = None
        for ep in range(150):
            tot = sum(net.train_step(x, t, 3e-3) for x, t in zip(xs, ts)) / len(xs)
            first = tot if first is None else first
        print(f"  loss {first:.3e} -> {tot:.3e}")
        assert tot < 0.1 * first


@pytest.mark.parametrize("solver", ["ilu0-f32", "mg-mixed"])
def test_warm_start_ablation_small_beam(ws, igap, solver):
    """The warm-start comparison on a 16x4x4 beam (a ramp of 60 steps to a
    large tip displacement): trained on steps 1..30, evaluated on 31..60 from
    the same state with the previous solution (plain FP64), linear and
    quadratic extrapolation and the learned guess (FP32 tangent with ILU(0),
    or FP64 tangent with the FP32 multigrid preconditioner).  Every arm
    reaches the plain FP64 answer to 1e-8.

    Learned against quadratic (the same candidate without the network): the
    learned term is used only if it lowered the residual of the quadratic
    extrapolation on the held-out training steps.  When it is used, the
    learned arm must need no more Krylov iterations than the quadratic arm
    on the evaluated steps (what the network gains; the candidates' residual
    ratios are printed beside); when the held-out test switches it off, the
    learned arm must be the quadratic arm exactly (same iterates, same
    counts)."""
    pc, prec = solver.split("-")
    s = ws.run(nelem=(16, 4, 4), nsteps=60, ntrain=30, tip_total=0.6, epochs=80, lr=2e-3, verbose=False, pc=pc,
               prec=prec, coarse_filter=2)
    for arm in ("plain", "linear", "quadratic", "learned"):
        print(arm, s[arm])
        assert s[arm]["all_converged"]
        assert s[arm]["final_rel_diff_vs_plain"] < 1e-8
    print({k: s[k] for k in ("loss_first", "loss_last", "val_loss_best", "val_loss_zero_correction",
                             "val_residual_ratio_learned_over_quadratic", "correction_used")})
    assert s["loss_last"] < s["loss_first"]
    q, l = s["quadratic"], s["learned"]
    rq, rl = q["median_residual_ratio_candidate_over_linear"], l["median_residual_ratio_candidate_over_linear"]
    print(f"learned vs quadratic on steps 31..60: candidate residual / linear {rl:.3g} vs {rq:.3g}; "
          f"newton {l['newton_its']} vs {q['newton_its']}, ksp {l['ksp_its']} vs {q['ksp_its']}")
    if s["correction_used"]:
        assert l["ksp_its"] <= q["ksp_its"]
    else:
        assert rl == rq and l["newton_its"] == q["newton_its"] and l["ksp_its"] == q["ksp_its"]
        assert l["final_rel_diff_vs_plain"] == q["final_rel_diff_vs_plain"]


def test_residual_training_lowers_the_candidate_residual(ws):
    """Training on the candidate's linearised residual (through
    igap_gpu_tangent_apply) on the 16x4x4 ramp: some epoch scores a lower
    held-out residual ratio learned / quadratic than the mean-square-trained
    network it starts from (a gradient of the wrong sign raises it every
    epoch), and every arm still reaches the plain answer to 1e-8."""
    s = ws.run(nelem=(16, 4, 4), nsteps=60, ntrain=30, tip_total=0.6, epochs=40, lr=2e-3, verbose=False, pc="mg",
               prec="mixed", coarse_filter=0, residual_epochs=20, residual_lr=2e-3, neval=10,
               arms=("plain", "quadratic", "learned"))
    rt = s["residual_training"]
    print({k: rt[k] for k in ("val_ratio_start", "val_ratio_best", "alpha")}, "used:", s["correction_used"])
    assert rt["val_ratio_best"] < rt["val_ratio_start"]
    for arm in ("plain", "quadratic", "learned"):
        print(arm, s[arm])
        assert s[arm]["all_converged"]
        assert s[arm]["final_rel_diff_vs_plain"] < 1e-8
