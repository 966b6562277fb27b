"""T5 DNN: parameters, training rows, and training on the HIP kernels.

The network is the reference's T5 classifier (SDL/NeuralNetwork.h:21-165):
38 inputs, Linear(38,32) ReLU Linear(32,32) ReLU Linear(32,1) sigmoid, used
by the DNN selection mode with the working point LSTWP2 (SDL/Quintuplet.h:
2921-2941).  Parameters are packed as bias_0[32], wgtT_0[38][32],
bias_2[32], wgtT_2[32][32], bias_4[1], wgtT_4[32][1], the order of the
chain's parameter files (tracklooper_amd/data/t5dnn/*.inc).

Training rows are the engine's record of every T5 candidate that reaches the
DNN (``Event.record_t5dnn``); a row is signal when all 10 of its hits belong
to one simulated track.  The forward and backward passes run in the library's
kernels (``lst_t5dnn_forward`` / ``lst_t5dnn_backward``) behind a
``torch.autograd.Function``; the optimiser and the loss are plain PyTorch.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PARAM_DIR = os.path.join(HERE, "data", "t5dnn")
TENSORS = [("bias_0", 32), ("wgtT_0", 38 * 32), ("bias_2", 32), ("wgtT_2", 32 * 32), ("bias_4", 1), ("wgtT_4", 32)]
N_PARAMS = sum(n for _, n in TENSORS)  # LST_T5DNN_NPARAMS
LSTWP2 = np.float32(0.6177366)


def load_params(directory: str = PARAM_DIR) -> np.ndarray:
    """Packed float32 parameters from a directory of parameter files."""
    out = []
    for name, n in TENSORS:
        with open(os.path.join(directory, name + ".inc")) as f:
            vals = [float.fromhex(l.strip().rstrip(",").rstrip("f")) for l in f if l.strip()]
        if len(vals) != n:
            raise ValueError(f"{name}: {len(vals)} values, expected {n}")
        out.extend(vals)
    return np.array(out, np.float32)


def write_params(params: np.ndarray, directory: str) -> None:
    """Write packed parameters as parameter files (C99 hex floats, one per line)."""
    p = np.asarray(params, np.float32)
    if p.size != N_PARAMS:
        raise ValueError(f"{p.size} parameters, expected {N_PARAMS}")
    os.makedirs(directory, exist_ok=True)
    o = 0
    for name, n in TENSORS:
        with open(os.path.join(directory, name + ".inc"), "w") as f:
            for v in p[o:o + n]:
                f.write(float.hex(float(v)) + "f,\n")
        o += n


def labels(hit_idxs: np.ndarray, hit_trk: np.ndarray) -> np.ndarray:
    """1 when all 10 hits of a candidate belong to one simulated track, else 0."""
    t = np.asarray(hit_trk)[np.asarray(hit_idxs, np.int64)]
    return ((t >= 0) & (t == t[:, :1])).all(1).astype(np.float32)


def collect(es, events, capacity: int = 1 << 19):
    """Training rows and labels of the candidates at the DNN call site of each event."""
    import tracklooper_amd as lst
    xs, ys = [], []
    for ev in events:
        e = lst.Event(es, t5_mode=lst.LST_T5_DNN).record_t5dnn(capacity).run_lst(ev, "t5")
        rec = e.t5dnn_record()
        e.close()
        xs.append(rec["features"])
        ys.append(labels(rec["hit_idxs"], ev["truth_hit_trk"]))
    return np.concatenate(xs), np.concatenate(ys)


def _lib():
    import tracklooper_amd as lst
    L = lst.lib()
    P = C.c_void_p
    L.lst_t5dnn_forward.argtypes = [P] * 5 + [C.c_uint32, P]
    L.lst_t5dnn_backward.argtypes = [P] * 8 + [C.c_uint32, P]
    L.lst_t5dnn_set_weights.argtypes = [P]
    L.lst_t5dnn_get_weights.argtypes = [P]
    return lst, L


def set_chain_weights(params: np.ndarray) -> None:
    """Make every event's DNN mode use these parameters (process-wide)."""
    lst, L = _lib()
    p = np.ascontiguousarray(params, np.float32)
    lst._check(L.lst_t5dnn_set_weights(p.ctypes.data_as(C.c_void_p)))


def get_chain_weights() -> np.ndarray:
    lst, L = _lib()
    p = np.zeros(N_PARAMS, np.float32)
    lst._check(L.lst_t5dnn_get_weights(p.ctypes.data_as(C.c_void_p)))
    return p


def _mlp_function():
    import torch

    class MLP(torch.autograd.Function):
        """logits = MLP(x; params) on the library's HIP kernels (float32, on the GPU)."""

        @staticmethod
        def forward(ctx, x, params):
            lst, L = _lib()
            n = x.shape[0]
            h1 = torch.empty((n, 32), device=x.device, dtype=torch.float32)
            h2 = torch.empty_like(h1)
            logit = torch.empty(n, device=x.device, dtype=torch.float32)
            s = torch.cuda.current_stream(x.device).cuda_stream
            lst._check(L.lst_t5dnn_forward(x.data_ptr(), params.data_ptr(), h1.data_ptr(), h2.data_ptr(),
                                           logit.data_ptr(), n, s))
            ctx.save_for_backward(x, params, h1, h2)
            return logit

        @staticmethod
        def backward(ctx, dlogit):
            lst, L = _lib()
            x, params, h1, h2 = ctx.saved_tensors
            dlogit = dlogit.contiguous()
            d1 = torch.empty_like(h1)
            d2 = torch.empty_like(h2)
            grad = torch.empty_like(params)
            s = torch.cuda.current_stream(x.device).cuda_stream
            lst._check(L.lst_t5dnn_backward(x.data_ptr(), params.data_ptr(), h1.data_ptr(), h2.data_ptr(),
                                            dlogit.data_ptr(), d2.data_ptr(), d1.data_ptr(), grad.data_ptr(),
                                            x.shape[0], s))
            return None, grad

    return MLP


def mlp(x, params):
    """Differentiable logits of the T5 DNN (x: [n, 38] float32 CUDA tensor, params: [N_PARAMS])."""
    return _mlp_function().apply(x.contiguous(), params)


def reference_logits(x, params):
    """Plain PyTorch logits (any dtype/device): the check for the kernels."""
    import torch
    o = 0
    parts = {}
    for name, n in TENSORS:
        parts[name] = params[o:o + n]
        o += n
    h = torch.relu(x @ parts["wgtT_0"].reshape(38, 32) + parts["bias_0"])
    h = torch.relu(h @ parts["wgtT_2"].reshape(32, 32) + parts["bias_2"])
    return (h @ parts["wgtT_4"].reshape(32, 1) + parts["bias_4"])[:, 0]


def train(x: np.ndarray, y: np.ndarray, params0: np.ndarray, epochs: int = 10, batch: int = 8192,
          lr: float = 1e-3, seed: int = 0, device: str = "cuda") -> tuple[np.ndarray, list[float]]:
    """Adam on the binary cross-entropy of the logits; returns the parameters and the epoch losses."""
    import torch
    g = torch.Generator().manual_seed(seed)
    X = torch.as_tensor(x, dtype=torch.float32, device=device)
    Y = torch.as_tensor(y, dtype=torch.float32, device=device)
    P = torch.nn.Parameter(torch.as_tensor(params0, dtype=torch.float32, device=device).clone())
    opt = torch.optim.Adam([P], lr=lr)
    losses = []
    for _ in range(epochs):
        perm = torch.randperm(X.shape[0], generator=g).to(device)
        tot = 0.0
        for i in range(0, X.shape[0], batch):
            idx = perm[i:i + batch]
            opt.zero_grad()
            loss = torch.nn.functional.binary_cross_entropy_with_logits(mlp(X[idx], P), Y[idx], reduction="sum")
            (loss / idx.numel()).backward()
            opt.step()
            tot += float(loss)
        losses.append(tot / X.shape[0])
    return P.detach().cpu().numpy().astype(np.float32), losses


def main(argv=None):
    """Collect rows from synthetic PU events, train from the reference parameters, write them."""
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--events", type=int, default=8)
    ap.add_argument("--n-pu", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="t5dnn_trained", help="directory for the parameter files")
    args = ap.parse_args(argv)
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    es = lst.LSTESData(synth.geometry_dir())
    events = [synth.generate_event("pu", 1000 + i, n_pu=args.n_pu) for i in range(args.events)]
    x, y = collect(es, events)
    p, losses = train(x, y, load_params(), epochs=args.epochs, lr=args.lr)
    write_params(p, args.out)
    print(json.dumps({"rows": int(len(y)), "signal_fraction": float(y.mean()), "losses": losses,
                      "out": args.out}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
