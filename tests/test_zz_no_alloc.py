"""lst_t5dnn_train_step_bf16_noalloc (no allocation, a fixed per-device array
for the chunk partials) against lst_t5dnn_train_step_bf16 (library-allocated
workspace): bit-identical gradient and loss on two calls in a row, the second
with a smaller input, and the capacity error above its maximum (GPU)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 32768  # LST_T5DNN_NOALLOC_MAX_ROWS


def _lib():
    import tracklooper_amd as lst
    L = lst.lib()
    P = C.c_void_p
    for name in ("lst_t5dnn_train_step_bf16", "lst_t5dnn_train_step_bf16_noalloc"):
        getattr(L, name).argtypes = [P] * 5 + [C.c_uint32, P]
        getattr(L, name).restype = C.c_int
    return L


def _inputs(n, seed):
    from tracklooper_amd import t5dnn
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 38)).astype(np.float32)
    y = (rng.uniform(size=n) < 0.3).astype(np.float32)
    p = t5dnn.load_params() + rng.normal(0, 0.01, t5dnn.N_PARAMS).astype(np.float32)
    return x, y, p


def _step(L, name, X, Y, P, n):
    import torch
    G = torch.full((P.numel(),), float("nan"), device="cuda")
    loss = torch.full((1,), float("nan"), device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    rc = getattr(L, name)(X.data_ptr(), Y.data_ptr(), P.data_ptr(), G.data_ptr(), loss.data_ptr(), n, s)
    torch.cuda.synchronize()
    return rc, G.cpu().numpy(), loss.cpu().numpy()


def test_noalloc_step_matches_allocating_step_twice():
    import torch
    L = _lib()
    # the maximum first, then a smaller ragged size: the second call reads only its own chunks' partials
    for i, n in enumerate((MAX_ROWS, 3001)):
        x, y, p = _inputs(n, 7 + i)
        X, Y, P = (torch.as_tensor(a, device="cuda") for a in (x, y, p))
        rc0, g0, l0 = _step(L, "lst_t5dnn_train_step_bf16", X, Y, P, n)
        rc1, g1, l1 = _step(L, "lst_t5dnn_train_step_bf16_noalloc", X, Y, P, n)
        assert rc0 == 0 and rc1 == 0
        assert np.isfinite(g0).all() and np.isfinite(l0).all()
        assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32)), f"call {i} (n={n}): gradients differ"
        assert np.array_equal(l0.view(np.uint32), l1.view(np.uint32)), f"call {i} (n={n}): losses differ"


def test_noalloc_step_rejects_oversize():
    import torch
    L = _lib()
    # the size check comes before any device work: the small buffers are never read
    X, Y, P = (torch.zeros(k, device="cuda") for k in (38, 1, 4))
    G = torch.zeros(4, device="cuda")
    rc = L.lst_t5dnn_train_step_bf16_noalloc(X.data_ptr(), Y.data_ptr(), P.data_ptr(), G.data_ptr(), None,
                                             MAX_ROWS + 1, None)
    assert rc == 6  # LST_E_CAPACITY
