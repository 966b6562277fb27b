"""T5 DNN: parameters, training rows, and training on the HIP kernels.

The network is the reference's T5 classifier (SDL/NeuralNetwork.h:21-165):
38 inputs, Linear(38,32) ReLU Linear(32,32) ReLU Linear(32,1) sigmoid, used
by the DNN selection mode with the working point LSTWP2 (SDL/Quintuplet.h:
2921-2941).  Parameters are packed as bias_0[32], wgtT_0[38][32],
bias_2[32], wgtT_2[32][32], bias_4[1], wgtT_4[32][1], the order of the
chain's parameter files (tracklooper_amd/data/t5dnn/*.inc).

Training rows are the engine's record of every T5 candidate that reaches the
DNN (``Event.record_t5dnn``); a row is signal when all 10 of its hits belong
to one simulated track.  Two training paths run on the library's kernels:

* FP32: forward and backward (``lst_t5dnn_forward`` / ``lst_t5dnn_backward``)
  behind a ``torch.autograd.Function``, with PyTorch's loss and Adam;
* BF16 on the matrix cores (``train_bf16``): one fused kernel per step for
  forward, loss and backward (``lst_t5dnn_train_step_bf16``, GEMMs in BF16 with
  FP32 accumulation) and the library's Adam (``lst_t5dnn_adam``); Python only
  draws the minibatches.

``evaluate_selection`` runs the chain with the plain chi-squared cut and with
the network (a given parameter set and working point) on the same events and
compares efficiency, fake rate and duplicate rate (``tracklooper_amd.metrics``).
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
    L.lst_t5dnn_train_step_bf16.argtypes = [P] * 5 + [C.c_uint32, P]
    L.lst_t5dnn_train_step_bf16x3.argtypes = [P] * 5 + [C.c_uint32, P]
    L.lst_t5dnn_adam.argtypes = [P] * 4 + [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, P]
    L.lst_t5dnn_set_working_point.argtypes = [C.c_float]
    L.lst_t5dnn_get_working_point.restype = C.c_float
    return lst, L


def set_chain_working_point(wp: float) -> None:
    """The score a T5 candidate must exceed in the chain's DNN mode (process-wide)."""
    lst, L = _lib()
    lst._check(L.lst_t5dnn_set_working_point(float(wp)))


def get_chain_working_point() -> float:
    return float(_lib()[1].lst_t5dnn_get_working_point())


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


def train_step_bf16(X, Y, P, G, loss=None, split: bool = False):
    """One fused forward / loss / backward on the matrix cores: G <- d(mean BCE)/dP, loss <- mean BCE.
    split: every GEMM operand as two BF16 terms, three MFMAs per product (about FP32 accuracy)."""
    import torch
    lst, L = _lib()
    s = torch.cuda.current_stream(X.device).cuda_stream
    fn = L.lst_t5dnn_train_step_bf16x3 if split else L.lst_t5dnn_train_step_bf16
    lst._check(fn(X.data_ptr(), Y.data_ptr(), P.data_ptr(), G.data_ptr(), None if loss is None else loss.data_ptr(),
                  X.shape[0], s))


def train_bf16(x: np.ndarray, y: np.ndarray, params0: np.ndarray, epochs: int = 10, batch: int = 8192,
               lr: float = 1e-3, seed: int = 0, device: str = "cuda", split: bool = True
               ) -> tuple[np.ndarray, list[float]]:
    """Adam on the mean binary cross-entropy, every step one fused MFMA kernel (split BF16 operands, or plain
    BF16 with split=False) plus the library's Adam; returns the parameters and the mean loss of each epoch."""
    import torch
    lst, L = _lib()
    g = torch.Generator().manual_seed(seed)
    X = torch.as_tensor(x, dtype=torch.float32, device=device)
    Y = torch.as_tensor(y, dtype=torch.float32, device=device)
    P = torch.as_tensor(params0, dtype=torch.float32, device=device).clone()
    G, M, V = torch.zeros_like(P), torch.zeros_like(P), torch.zeros_like(P)
    s = torch.cuda.current_stream(X.device).cuda_stream
    losses, step = [], 0
    n = X.shape[0]
    for _ in range(epochs):
        perm = torch.randperm(n, generator=g).to(device)
        Xp, Yp = X[perm], Y[perm]  # one gather per epoch; the batches are then contiguous slices
        batch_loss = torch.zeros(-(-n // batch), dtype=torch.float32, device=device)
        for b, i in enumerate(range(0, n, batch)):
            xb, yb = Xp[i:i + batch], Yp[i:i + batch]
            train_step_bf16(xb, yb, P, G, batch_loss[b:], split)
            step += 1
            lst._check(L.lst_t5dnn_adam(P.data_ptr(), G.data_ptr(), M.data_ptr(), V.data_ptr(), step, lr, 0.9, 0.999,
                                        1e-8, s))
        rows = torch.full_like(batch_loss, float(batch), dtype=torch.float64)
        rows[-1] = n - batch * (len(rows) - 1)
        losses.append(float((batch_loss.double() * rows).sum()) / n)
    return P.cpu().numpy().astype(np.float32), losses


def gradient_errors(x: np.ndarray, y: np.ndarray, params: np.ndarray, n: int = 8192) -> dict:
    """Relative errors of the FP32, the BF16 MFMA and the split-BF16 MFMA gradients (and losses) of the mean
    BCE against float64 autograd on the first n rows: max |g - g64| / max |g64| and ||g - g64|| / ||g64||."""
    import torch
    n = min(n, len(y))
    X = torch.as_tensor(x[:n], dtype=torch.float32, device="cuda")
    Y = torch.as_tensor(y[:n], dtype=torch.float32, device="cuda")
    P = torch.as_tensor(params, dtype=torch.float32, device="cuda")
    P64 = torch.as_tensor(params, dtype=torch.float64).requires_grad_(True)
    l64 = torch.nn.functional.binary_cross_entropy_with_logits(reference_logits(X.cpu().double(), P64),
                                                               Y.cpu().double())
    l64.backward()
    g64 = P64.grad.numpy()
    P32 = P.clone().requires_grad_(True)
    l32 = torch.nn.functional.binary_cross_entropy_with_logits(mlp(X, P32), Y)
    l32.backward()
    g32 = P32.grad.cpu().double().numpy()
    res = [("fp32", g32, float(l32.detach()))]
    for name, split in (("bf16_mfma", False), ("bf16x3_mfma", True)):
        G = torch.zeros_like(P)
        lb = torch.zeros(1, dtype=torch.float32, device="cuda")
        train_step_bf16(X, Y, P, G, lb, split)
        res.append((name, G.cpu().double().numpy(), float(lb)))
    scale, norm = np.abs(g64).max(), np.linalg.norm(g64)
    out = {"rows": n}
    for name, gr, lo in res:
        out[name] = {"grad_max_rel": float(np.abs(gr - g64).max() / scale),
                     "grad_l2_rel": float(np.linalg.norm(gr - g64) / norm),
                     "loss_rel": abs(lo - float(l64)) / abs(float(l64))}
    return out


def collect_rows(es, events, capacity: int = 1 << 19):
    """Training rows (features, labels) of many events, one event handle reused."""
    import tracklooper_amd as lst
    e = lst.Event(es, t5_mode=lst.LST_T5_DNN).record_t5dnn(capacity)
    xs, ys = [], []
    for ev in events:
        e.run_lst(ev, "t5")
        rec = e.t5dnn_record()
        xs.append(rec["features"])
        ys.append(labels(rec["hit_idxs"], ev["truth_hit_trk"]))
    e.close()
    return np.concatenate(xs), np.concatenate(ys)


def scores(x: np.ndarray, params: np.ndarray, batch: int = 1 << 20) -> np.ndarray:
    """Sigmoid scores of rows with the FP32 forward kernel (the chain's arithmetic)."""
    import torch
    P = torch.as_tensor(params, dtype=torch.float32, device="cuda")
    out = []
    for i in range(0, len(x), batch):
        X = torch.as_tensor(x[i:i + batch], dtype=torch.float32, device="cuda")
        out.append(torch.sigmoid(mlp(X, P)).cpu().numpy())
    return np.concatenate(out) if out else np.zeros(0, np.float32)


def selection_metrics(es, events, t5_mode: int, params: np.ndarray | None = None, wp: float | None = None,
                      slots: int = 64) -> dict:
    """Efficiency, fake and duplicate rates of the chain on events (batched engine) in a T5 selection mode;
    in DNN mode with the given parameters and working point (restored afterwards)."""
    import tracklooper_amd as lst
    from tracklooper_amd import metrics
    old_p, old_wp = get_chain_weights(), get_chain_working_point()
    try:
        if params is not None:
            set_chain_weights(params)
        if wp is not None:
            set_chain_working_point(wp)
        b = lst.Batch(es, slots=min(slots, len(events)), t5_mode=t5_mode)
        outs = b.run(events)
        b.close()
    finally:
        set_chain_weights(old_p)
        set_chain_working_point(old_wp)
    return metrics.summarize([metrics.event_metrics(o, ev) for o, ev in zip(outs, events)])


def evaluate_selection(es, events, params: np.ndarray, wp: float) -> dict:
    """The plain chi-squared cut, the shipped network at LSTWP2, and `params` at `wp`, on the same events."""
    keys = ("efficiency", "fake_rate", "duplicate_rate", "track_candidates", "sim_tracks")
    res = {"plain_chi2": selection_metrics(es, events, 0),
           "shipped_dnn_lstwp2": selection_metrics(es, events, 1, load_params(), float(LSTWP2)),
           "trained_dnn": selection_metrics(es, events, 1, params, wp)}
    return {k: {q: v[q] for q in keys} | {"by_type": v["by_type"]} for k, v in res.items()}


def choose_working_point(es, events, params: np.ndarray, max_eff_loss: float = 0.0025,
                         grid=tuple(np.round(np.arange(0.05, 0.96, 0.05), 2))) -> tuple[float, list]:
    """The highest threshold (fewest fakes) whose efficiency on these events is at most `max_eff_loss` below
    the plain chi-squared cut's; returns it and the scan."""
    plain = selection_metrics(es, events, 0)
    scan, best = [], float(grid[0])
    for wp in grid:
        m = selection_metrics(es, events, 1, params, float(wp))
        scan.append({"wp": float(wp), "efficiency": m["efficiency"], "fake_rate": m["fake_rate"]})
        if m["efficiency"] >= plain["efficiency"] - max_eff_loss:
            best = float(wp)
    return best, [{"plain_efficiency": plain["efficiency"], "plain_fake_rate": plain["fake_rate"]}] + scan


def main(argv=None):
    """Collect rows from synthetic PU200 events, train the T5 DNN (MFMA steps, or FP32 autograd) from
    the shipped parameters, report the gradient errors against float64, write the parameter files.  With
    --val-events and --eval-events: choose the trained network's working point on the validation events, then
    compare the plain chi-squared cut, the shipped network and the trained one on the evaluation events."""
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--events", type=int, default=8)
    ap.add_argument("--min-rows", type=int, default=0, help="add events until this many rows are collected")
    ap.add_argument("--n-pu", type=int, default=200)
    ap.add_argument("--geometry", default="OT800_IT615_pt0.8")
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--precision", choices=("bf16x3", "bf16", "fp32"), default="bf16x3",
                    help="MFMA steps with split BF16 operands, with plain BF16 operands, or the FP32 autograd path")
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default="t5dnn_trained", help="directory for the parameter files")
    ap.add_argument("--val-events", type=int, default=0, help="events for the working-point choice")
    ap.add_argument("--eval-events", type=int, default=0, help="events for the selection comparison")
    args = ap.parse_args(argv)
    import time
    from tracklooper_amd import synth
    pt = 0.6 if "pt0.6" in args.geometry else 0.8
    specs = lambda lo, hi: [("pu", 5000 + i, args.n_pu, 2, args.geometry) for i in range(lo, hi)]
    # every event before the GPU is used (the pool forks); validation and evaluation seeds follow the training ones
    n_val, n_eval = (args.val_events, args.eval_events) if args.eval_events else (0, 0)
    held = synth.generate_many(specs(100000, 100000 + n_val + n_eval), args.workers) if n_eval else []
    val_events, eval_events = held[:n_val], held[n_val:]
    events = synth.generate_many(specs(0, args.events), args.workers)
    print(f"generated {len(events)} training and {len(held)} held-out events", file=sys.stderr, flush=True)
    import tracklooper_amd as lst
    es = lst.LSTESData(synth.geometry_dir(args.geometry), pt_cut=pt)
    x, y = collect_rows(es, events)
    n_ev = len(events)
    while len(y) < args.min_rows:  # more events until enough rows (a pool cannot fork now: serial)
        more = synth.generate_many(specs(n_ev, n_ev + 8), 1)
        n_ev += len(more)
        xm, ym = collect_rows(es, more)
        x, y = np.concatenate([x, xm]), np.concatenate([y, ym])
    print(f"{len(y)} rows from {n_ev} events", file=sys.stderr, flush=True)
    p0 = load_params()
    errors = gradient_errors(x, y, p0)
    t0 = time.perf_counter()
    if args.precision == "fp32":
        p, losses = train(x, y, p0, epochs=args.epochs, batch=args.batch, lr=args.lr)
    else:
        p, losses = train_bf16(x, y, p0, epochs=args.epochs, batch=args.batch, lr=args.lr,
                               split=args.precision == "bf16x3")
    dt = time.perf_counter() - t0
    write_params(p, args.out)
    res = {"events": n_ev, "rows": int(len(y)), "signal_fraction": float(y.mean()),
           "precision": args.precision, "batch": args.batch, "epochs": args.epochs,
           "train_seconds": dt, "steps_per_s": args.epochs * -(-len(y) // args.batch) / dt,
           "losses": losses, "gradient_errors_vs_float64": errors, "out": args.out}
    print(json.dumps(res), flush=True)
    if eval_events:
        wp, scan = choose_working_point(es, val_events or eval_events, p)
        print(f"working point {wp}", file=sys.stderr, flush=True)
        res = {"geometry": args.geometry, "val_events": len(val_events), "eval_events": len(eval_events),
               "working_point": wp, "wp_scan": scan, "selection": evaluate_selection(es, eval_events, p, wp)}
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
