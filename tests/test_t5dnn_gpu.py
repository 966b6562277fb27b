"""T5 DNN training kernels (GPU): logits and gradients of lst_t5dnn_forward /
backward against float64 PyTorch autograd, the recorded chain scores, and a
short training run whose parameters the chain and the oracle then use alike."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rows(geom_dir):
    import tracklooper_amd as lst
    from tracklooper_amd import synth, t5dnn
    es = lst.LSTESData(geom_dir)
    events = [synth.generate_event("pu", 20 + i, n_pu=50) for i in range(2)]
    e = lst.Event(es, t5_mode=lst.LST_T5_DNN).record_t5dnn().run_lst(events[0], "t5")
    rec = e.t5dnn_record()
    x, y = t5dnn.collect(es, events)
    return es, events, rec, x, y


def test_forward_matches_float64_and_chain_scores(rows):
    import torch
    from tracklooper_amd import t5dnn
    _, _, rec, _, _ = rows
    p = t5dnn.load_params()
    X = torch.as_tensor(rec["features"], device="cuda")
    P = torch.as_tensor(p, device="cuda")
    logit = t5dnn.mlp(X, P).cpu().double()
    ref = t5dnn.reference_logits(X.cpu().double(), P.cpu().double())
    # float32 accumulation of input terms up to ~1e3 (z coordinates times weights) against float64:
    # a few 1e-5 of absolute error on logits of order 1
    torch.testing.assert_close(logit, ref, rtol=1e-4, atol=1e-4)
    # the chain's score of every recorded candidate is the sigmoid of the same float32 logit
    np.testing.assert_allclose(torch.sigmoid(logit).numpy(), rec["score"], rtol=2e-6, atol=2e-7)


def test_gradients_match_float64_autograd(rows):
    import torch
    from tracklooper_amd import t5dnn
    _, _, _, x, y = rows
    n = min(4096, len(y))
    rng = np.random.default_rng(0)
    p = t5dnn.load_params() + rng.normal(0, 0.01, t5dnn.N_PARAMS).astype(np.float32)
    X = torch.as_tensor(x[:n], device="cuda")
    Y = torch.as_tensor(y[:n], device="cuda")
    P = torch.as_tensor(p, device="cuda").requires_grad_(True)
    loss = torch.nn.functional.binary_cross_entropy_with_logits(t5dnn.mlp(X, P), Y)
    loss.backward()
    P64 = torch.as_tensor(p, dtype=torch.float64).requires_grad_(True)
    X64, Y64 = X.cpu().double(), Y.cpu().double()
    f = lambda q: torch.nn.functional.binary_cross_entropy_with_logits(t5dnn.reference_logits(X64, q), Y64)
    l64 = f(P64)
    l64.backward()
    g, g64 = P.grad.cpu().double(), P64.grad
    assert abs(float(loss) - float(l64)) < 1e-5 * max(1.0, abs(float(l64)))
    # float32 kernels (same accumulation error as the forward pass) against float64 autograd
    torch.testing.assert_close(g, g64, rtol=1e-3, atol=1e-4 * float(g64.abs().max()))
    # the float64 reference itself: central differences on a few parameters
    for q in (0, 100, 40 * 32 + 5, 2304, 2336):
        e = torch.zeros_like(P64)
        e[q] = 1e-6
        fd = (f(P64.detach() + e) - f(P64.detach() - e)) / 2e-6
        assert abs(float(fd) - float(g64[q])) < 1e-6 + 1e-4 * abs(float(g64[q]))


def test_training_lowers_the_loss_and_the_chain_uses_trained_weights(rows, oracle_geometry):
    from oracle import orc
    import tracklooper_amd as lst
    from tests.parity import compare
    from tracklooper_amd import t5dnn
    es, events, _, x, y = rows
    assert 0.05 < y.mean() < 0.95
    p, losses = t5dnn.train(x, y, t5dnn.load_params(), epochs=4, batch=2048, lr=3e-3)
    assert losses[-1] < losses[0]
    ref_p = t5dnn.load_params()
    try:
        t5dnn.set_chain_weights(p)
        np.testing.assert_array_equal(t5dnn.get_chain_weights().view(np.uint32), p.view(np.uint32))
        orc.lib().orc_t5dnn_set_weights(np.ascontiguousarray(p).ctypes.data_as(C.c_void_p))
        prod = lst.Event(es, t5_mode=lst.LST_T5_DNN).run_lst(events[1], "tc")
        ref = orc.Event(oracle_geometry, t5_mode=1).run_lst(events[1], "tc")
        compare(prod, ref, "tc")
    finally:
        t5dnn.set_chain_weights(ref_p)
        orc.lib().orc_t5dnn_set_weights(np.ascontiguousarray(ref_p).ctypes.data_as(C.c_void_p))
