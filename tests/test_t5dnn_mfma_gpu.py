"""The MFMA training steps of the T5 DNN (lst_t5dnn_train_step_bf16 and the split-operand
lst_t5dnn_train_step_bf16x3):
gradients against float64 autograd with the error stated per precision,
determinism, ragged sizes, a training run, and the chain's settable working
point against the oracle (GPU)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rows(geom_dir):
    import tracklooper_amd as lst
    from tracklooper_amd import synth, t5dnn
    es = lst.LSTESData(geom_dir)
    events = [synth.generate_event("pu", 40 + i, n_pu=200) for i in range(2)]
    x, y = t5dnn.collect_rows(es, events)
    return es, events, x, y


def bf16(a):
    """float32 -> bfloat16 (round to nearest even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def emulated_bf16_step(x, y, p, split=False):
    """The MFMA step's arithmetic in float64, rounding to bfloat16 exactly where the kernel does: every GEMM
    operand as bf16(v) (plain), or as the pair bf16(v), bf16(v - bf16(v)) with the three products
    lo*hi + hi*lo + hi*hi (split).  Pins the kernel's GEMM operand layouts, masks and sums independently of
    the rounding error itself."""
    from tracklooper_amd import t5dnn
    o, P = 0, {}
    for name, k in t5dnn.TENSORS:
        P[name] = np.asarray(p[o:o + k], np.float64)
        o += k
    n = len(y)

    def M(a, b):
        ah, bh = bf16(a), bf16(b)
        r = ah @ bh
        if split:
            r = r + bf16(a - ah) @ bh + ah @ bf16(b - bh)
        return r

    carried = (lambda v: bf16(v) + bf16(v - bf16(v))) if split else bf16  # the value a GEMM operand carries
    W0, W2, W4 = P["wgtT_0"].reshape(38, 32), P["wgtT_2"].reshape(32, 32), P["wgtT_4"]
    H1 = np.maximum(M(x, W0) + P["bias_0"], 0)
    H2 = np.maximum(M(H1, W2) + P["bias_2"], 0)
    z = H2 @ W4 + P["bias_4"][0]
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dl = (1 / (1 + np.exp(-z)) - y) / n
    dH2 = (H2 > 0) * dl[:, None] * W4[None, :]
    dH1 = (H1 > 0) * M(dH2, W2.T)
    g = np.concatenate([carried(dH1).sum(0), M(np.asarray(x, np.float64).T, dH1).ravel(), dH2.sum(0),
                        M(H1.T, dH2).ravel(), [dl.sum()], H2.T @ dl])
    return g, loss


@pytest.mark.parametrize("split,tol", [(False, 2e-3), (True, 1e-4)])
def test_bf16_step_matches_its_float64_emulation(rows, split, tol):
    """The kernel against the same arithmetic in float64 with the same bfloat16 roundings: only the FP32
    accumulation order differs."""
    import torch
    from tracklooper_amd import t5dnn
    _, _, x, y = rows
    rng = np.random.default_rng(2)
    p = t5dnn.load_params() + rng.normal(0, 0.01, t5dnn.N_PARAMS).astype(np.float32)
    for xs, ys in ((x[:8192], y[:8192]), (rng.normal(0, 1, (3000, 38)).astype(np.float32),
                                          (rng.uniform(size=3000) < 0.3).astype(np.float32))):
        g_ref, loss_ref = emulated_bf16_step(xs, ys, p, split)
        P = torch.as_tensor(p, device="cuda")
        G = torch.zeros_like(P)
        L = torch.zeros(1, device="cuda")
        t5dnn.train_step_bf16(torch.as_tensor(xs, device="cuda"), torch.as_tensor(ys, device="cuda"), P, G, L,
                              split)
        g = G.cpu().double().numpy()
        rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
        print(f"split={split} kernel vs emulation: gradient L2 rel {rel:.2e}, loss {float(L):.7f} vs {loss_ref:.7f}")
        assert rel < tol
        assert abs(float(L) - loss_ref) < 1e-5 * max(1.0, loss_ref)


def test_bf16_and_fp32_gradients_against_float64(rows):
    """The error of each precision against float64 autograd, stated: the FP32 kernels agree to a few 1e-7.
    Plain BF16 GEMM operands (8-bit mantissa) on the raw T5 features (z up to 270 cm, radii up to 110 cm)
    move the logits by about 0.1 and give a loss within a few 1e-3 and a gradient within about 10-25 % in L2
    (observed 0.08-0.14 on PU200 rows); the split operands (16-bit mantissa, three MFMAs per product) bring
    both to about 1e-5 (gradient below 1e-3, loss below 5e-4 asserted)."""
    from tracklooper_amd import t5dnn
    _, _, x, y = rows
    assert len(y) > 8192 and 0.02 < y.mean() < 0.98
    rng = np.random.default_rng(1)
    p = t5dnn.load_params() + rng.normal(0, 0.01, t5dnn.N_PARAMS).astype(np.float32)
    err = t5dnn.gradient_errors(x, y, p, n=8192)
    print(err)
    assert err["fp32"]["grad_l2_rel"] < 1e-5 and err["fp32"]["loss_rel"] < 1e-5
    assert err["bf16_mfma"]["grad_l2_rel"] < 0.3 and err["bf16_mfma"]["loss_rel"] < 0.05
    assert err["bf16x3_mfma"]["grad_l2_rel"] < 1e-3 and err["bf16x3_mfma"]["loss_rel"] < 5e-4


@pytest.mark.parametrize("split", [False, True])
def test_bf16_step_deterministic_ragged_and_empty(rows, split):
    import torch
    from tracklooper_amd import t5dnn
    _, _, x, y = rows
    P = torch.as_tensor(t5dnn.load_params(), device="cuda")
    for n in (1, 63, 64, 1000, 8192):
        X = torch.as_tensor(x[:n], device="cuda")
        Y = torch.as_tensor(y[:n], device="cuda")
        g1, g2 = torch.zeros_like(P), torch.zeros_like(P)
        l1, l2 = torch.zeros(1, device="cuda"), torch.zeros(1, device="cuda")
        t5dnn.train_step_bf16(X, Y, P, g1, l1, split)
        t5dnn.train_step_bf16(X, Y, P, g2, l2, split)
        assert torch.equal(g1, g2) and torch.equal(l1, l2), n
        assert bool(torch.isfinite(g1).all()) and float(l1) > 0
    # the gradient of n rows equals the mean of the per-chunk gradients weighted by rows (chunks of 64)
    X = torch.as_tensor(x[:128], device="cuda")
    Y = torch.as_tensor(y[:128], device="cuda")
    ga, gb, g = torch.zeros_like(P), torch.zeros_like(P), torch.zeros_like(P)
    t5dnn.train_step_bf16(X[:64].contiguous(), Y[:64].contiguous(), P, ga, None, split)
    t5dnn.train_step_bf16(X[64:].contiguous(), Y[64:].contiguous(), P, gb, None, split)
    t5dnn.train_step_bf16(X, Y, P, g, None, split)
    torch.testing.assert_close(g, 0.5 * (ga + gb), rtol=1e-5, atol=1e-7)


def test_bf16_training_lowers_the_loss(rows):
    from tracklooper_amd import t5dnn
    _, _, x, y = rows
    p, losses = t5dnn.train_bf16(x, y, t5dnn.load_params(), epochs=5, batch=8192, lr=2e-3)
    assert losses[-1] < losses[0] and np.all(np.isfinite(p))
    s0 = t5dnn.scores(x, t5dnn.load_params())
    s1 = t5dnn.scores(x, p)
    bce = lambda s: float(-np.mean(y * np.log(np.clip(s, 1e-7, 1)) + (1 - y) * np.log(np.clip(1 - s, 1e-7, 1))))
    assert bce(s1) < bce(s0)


def test_working_point_setting_matches_oracle(rows, oracle_geometry):
    import tracklooper_amd as lst
    from oracle import orc
    from tests.parity import compare
    from tracklooper_amd import t5dnn
    es, events, _, _ = rows
    orc.lib().orc_t5dnn_set_working_point.argtypes = [C.c_float]
    assert abs(t5dnn.get_chain_working_point() - 0.6177366) < 1e-7
    try:
        for wp in (0.3, 0.9):
            t5dnn.set_chain_working_point(wp)
            orc.lib().orc_t5dnn_set_working_point(wp)
            prod = lst.Event(es, t5_mode=lst.LST_T5_DNN).run_lst(events[1], "tc")
            ref = orc.Event(oracle_geometry, t5_mode=1).run_lst(events[1], "tc")
            compare(prod, ref, "tc")
    finally:
        t5dnn.set_chain_working_point(float(t5dnn.LSTWP2))
        orc.lib().orc_t5dnn_set_working_point(float(t5dnn.LSTWP2))
