"""Encoding wizard on the GPU: exhaustive trial sizing and a learned selector trained on its choices.

The reference picks an expression per column by encoding it with every expression of the pool and keeping the
smallest (ExpColEncoder<PT>::encode_all + AllExpEncodedCol::GetBest, src/encoder/exp_col_encoder.cpp:19-28,
117-129). Here, per 65536-value int64 row-group:

* exhaustive -- `fls_wizard_sizes` computes the exact encoded size under every candidate in closed form on the GPU
  (what `encode_from_memory` would return), `fls_wizard_best` takes the argmin;
* learned    -- `fls_wizard_features` reads 4 of the 64 vectors and computes 64 statistics, the MLP
  (64-256-256-K over the K candidates, BF16 MFMA, FP32 accumulation) proposes the expression.

The MLP is trained on the GPU (fls_mlp_forward_backward / fls_mlp_adam_step) on labels from the exhaustive sizer.
Either choice is then encoded by the same bit-exact codec: the selector changes sizes, never decoded values.

    python -m fastlanes_amd.wizard [--columns 512] [--log2n 20]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import struct
import time

import numpy as np

import fastlanes_amd as fa
from fastlanes_amd import workloads as wl

RG = 65536
N_FEATURES = 64
_vp, _u64, _u32 = C.c_void_p, C.c_uint64, C.c_uint32
WEIGHTS_MAGIC = b"FLSWIZ01"
FEATURE_SCHEMA = "v3:est16,log16,bw,runs,dict,order,mag,sample4,patch2,dpatch2,bias"


def _lib():
    L = fa.lib()
    if getattr(L, "_wizard_bound", False):
        return L
    L.fls_wizard_n_candidates.restype = _u32
    L.fls_wizard_candidate_expr.argtypes = [_u32]
    L.fls_wizard_candidate_expr.restype = C.c_uint8
    L.fls_wizard_sizes.argtypes = [_vp, _u64, _vp, _vp]
    L.fls_wizard_best.argtypes = [_vp, _u64, _vp, _vp]
    L.fls_wizard_features.argtypes = [_vp, _u64, _vp, _vp]
    L.fls_mlp_create.argtypes = [_u32, _u32, _u64, C.POINTER(_vp)]
    L.fls_mlp_destroy.argtypes = [_vp]
    L.fls_mlp_destroy.restype = None
    L.fls_mlp_n_params.restype = _u64
    L.fls_mlp_get_params.argtypes = [_vp, _vp]
    L.fls_mlp_set_params.argtypes = [_vp, _vp]
    L.fls_mlp_get_grads.argtypes = [_vp, _vp]
    L.fls_mlp_forward_backward.argtypes = [_vp, _vp, _vp, _u32, C.POINTER(C.c_float), _vp]
    L.fls_mlp_adam_step.argtypes = [_vp, C.c_float, _vp]
    L.fls_mlp_predict.argtypes = [_vp, _vp, _u64, _vp, _vp]
    L.fls_mlp_get_logits.argtypes = [_vp, _vp, _u32]
    L._wizard_bound = True
    return L


def candidates() -> list[int]:
    L = _lib()
    return [int(L.fls_wizard_candidate_expr(k)) for k in range(L.fls_wizard_n_candidates())]


def exhaustive_sizes(dev_values: int, n_rg: int, dev_sizes: int, stream: int = 0) -> None:
    """dev_sizes[r * K + k] = exact encoded bytes of row-group r under candidate k."""
    fa.check(_lib().fls_wizard_sizes(dev_values, n_rg, dev_sizes, stream), "wizard_sizes")


def get_best(dev_sizes: int, n_rg: int, dev_choice: int, stream: int = 0) -> None:
    fa.check(_lib().fls_wizard_best(dev_sizes, n_rg, dev_choice, stream), "wizard_best")


def features(dev_values: int, n_rg: int, dev_features: int, stream: int = 0) -> None:
    fa.check(_lib().fls_wizard_features(dev_values, n_rg, dev_features, stream), "wizard_features")


class MLP:
    """64-256-256-K selector (fls_mlp_*)."""

    def __init__(self, n_classes: int, max_batch: int = 1024, seed: int = 0):
        L = _lib()
        h = _vp()
        fa.check(L.fls_mlp_create(n_classes, max_batch, seed, C.byref(h)), "mlp_create")
        self.h, self.n_classes, self.max_batch = h, n_classes, max_batch
        self.n_params = int(L.fls_mlp_n_params())

    def params(self) -> np.ndarray:
        p = np.empty(self.n_params, dtype=np.float32)
        fa.check(_lib().fls_mlp_get_params(self.h, p.ctypes.data), "mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        fa.check(_lib().fls_mlp_set_params(self.h, p.ctypes.data), "mlp_set_params")

    def grads(self) -> np.ndarray:
        g = np.empty(self.n_params, dtype=np.float32)
        fa.check(_lib().fls_mlp_get_grads(self.h, g.ctypes.data), "mlp_get_grads")
        return g

    def logits(self, batch: int) -> np.ndarray:
        z = np.empty((batch, 32), dtype=np.float32)
        fa.check(_lib().fls_mlp_get_logits(self.h, z.ctypes.data, batch), "mlp_get_logits")
        return z

    def forward_backward(self, dev_x: int, dev_y: int, batch: int, stream: int = 0, want_loss=True):
        loss = C.c_float(0)
        fa.check(_lib().fls_mlp_forward_backward(self.h, dev_x, dev_y, batch, C.byref(loss) if want_loss else None,
                                                 stream), "mlp_forward_backward")
        return float(loss.value)

    def adam_step(self, lr: float, stream: int = 0) -> None:
        fa.check(_lib().fls_mlp_adam_step(self.h, lr, stream), "mlp_adam_step")

    def predict(self, dev_x: int, n: int, dev_choice: int, stream: int = 0) -> None:
        fa.check(_lib().fls_mlp_predict(self.h, dev_x, n, dev_choice, stream), "mlp_predict")

    def save(self, path: str) -> None:
        """Weights blob: magic, version, classes, parameter count, feature-schema tag, float32 parameters."""
        tag = FEATURE_SCHEMA.encode()
        with open(path, "wb") as f:
            f.write(WEIGHTS_MAGIC + struct.pack("<IIQI", 1, self.n_classes, self.n_params, len(tag)) + tag)
            f.write(self.params().tobytes())

    def load(self, path: str) -> None:
        with open(path, "rb") as f:
            blob = f.read()
        if blob[:8] != WEIGHTS_MAGIC:
            raise fa.FlsError(f"{path}: not a wizard weights file")
        ver, kc, n, tl = struct.unpack_from("<IIQI", blob, 8)
        tag = blob[28: 28 + tl].decode()
        if ver != 1 or kc != self.n_classes or n != self.n_params or tag != FEATURE_SCHEMA:
            raise fa.FlsError(f"{path}: incompatible weights ({ver}, {kc}, {n}, {tag})")
        self.set_params(np.frombuffer(blob, dtype=np.float32, offset=28 + tl, count=n))

    def __del__(self):
        if getattr(self, "h", None) and fa._lib is not None:
            fa._lib.fls_mlp_destroy(self.h)
            self.h = None


def sample_estimate_choice(feat: np.ndarray) -> np.ndarray:
    """The selector without the network: argmin of the sample-based size estimates (features 0 .. K-1)."""
    return np.argmin(feat[:, :len(candidates())], axis=1).astype(np.int32)


def _events(torch, stream):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def build_dataset(torch, dev, n_columns=512, n=1 << 20, seed=99):
    """Device array of all columns' values (n_columns * n int64) and the column id of every row-group."""
    per = n // RG
    d_vals = torch.empty(n_columns * n, dtype=torch.int64, device=dev)
    fams = []
    for c, (fam, x) in enumerate(wl.wizard_columns(n_columns, n, seed)):
        d_vals[c * n:(c + 1) * n] = torch.from_numpy(x).to(dev, non_blocking=False)
        fams.append(fam)
    col = np.repeat(np.arange(n_columns), per)
    return d_vals, col, fams


def train(torch, mlp: MLP, d_feat, d_lab, idx: np.ndarray, steps: int, batch: int, lr: float, dev, stream, seed=0):
    """Minibatch Adam on the rows `idx` of (features, labels); returns the loss curve (every 50 steps)."""
    g = np.random.default_rng(seed)
    sp = int(stream.cuda_stream)
    d_idx = torch.from_numpy(idx.astype(np.int64)).to(dev)
    xb = torch.empty((batch, N_FEATURES), dtype=torch.float32, device=dev)
    yb = torch.empty(batch, dtype=torch.int32, device=dev)
    curve = []
    for s in range(steps):
        sel = d_idx[torch.from_numpy(g.integers(0, idx.size, batch)).to(dev)]
        torch.index_select(d_feat, 0, sel, out=xb)
        torch.index_select(d_lab, 0, sel, out=yb)
        want = s % 50 == 0 or s == steps - 1
        loss = mlp.forward_backward(xb.data_ptr(), yb.data_ptr(), batch, sp, want_loss=want)
        mlp.adam_step(lr * (0.5 * (1 + np.cos(np.pi * s / steps))) + lr * 0.02, sp)
        if want:
            curve.append((s, round(loss, 4)))
    return curve


def run(dev, stream, n_columns=512, n=1 << 20, steps=3000, batch=512, lr=2e-3, seed=99, reps=5, return_model=False):
    """Config-5 study: exhaustive labels, training, held-out regret and selection times. Returns a dict (and the
    trained MLP with return_model)."""
    import torch

    sp = int(stream.cuda_stream)
    t0 = time.time()
    d_vals, col, fams = build_dataset(torch, dev, n_columns, n, seed)
    n_rg = d_vals.numel() // RG
    K = len(candidates())
    d_sizes = torch.empty((n_rg, K), dtype=torch.int64, device=dev)
    d_best = torch.empty(n_rg, dtype=torch.int32, device=dev)
    d_feat = torch.empty((n_rg, N_FEATURES), dtype=torch.float32, device=dev)
    d_pred = torch.empty(n_rg, dtype=torch.int32, device=dev)
    t_build = time.time() - t0

    def exhaustive():
        exhaustive_sizes(d_vals.data_ptr(), n_rg, d_sizes.data_ptr(), sp)
        get_best(d_sizes.data_ptr(), n_rg, d_best.data_ptr(), sp)

    mlp = MLP(K, max_batch=max(batch, 8192), seed=seed)

    def learned():
        features(d_vals.data_ptr(), n_rg, d_feat.data_ptr(), sp)
        mlp.predict(d_feat.data_ptr(), n_rg, d_pred.data_ptr(), sp)

    exhaustive()
    features(d_vals.data_ptr(), n_rg, d_feat.data_ptr(), sp)
    torch.cuda.synchronize()
    test = (col % 4) == 3  # every fourth column is held out
    tr_idx = np.nonzero(~test)[0]
    t1 = time.time()
    curve = train(torch, mlp, d_feat, d_best, tr_idx, steps, batch, lr, dev, stream)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    learned()
    torch.cuda.synchronize()

    def timed(fn):
        ts = []
        for _ in range(reps + 1):
            a, b = _events(torch, stream)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ts.append(a.elapsed_time(b) * 1e-3)
        return float(np.median(ts[1:]))

    t_ex, t_le = timed(exhaustive), timed(learned)
    sizes = d_sizes.cpu().numpy().astype(np.float64)
    best = d_best.cpu().numpy()
    pred = d_pred.cpu().numpy()
    feat = d_feat.cpu().numpy()
    base = sample_estimate_choice(feat)
    rows = np.arange(n_rg)

    def regret(choice, mask):
        return float(sizes[rows[mask], choice[mask]].sum() / sizes[rows[mask], best[mask]].sum() - 1.0)

    res = {
        "columns": n_columns, "values_per_column": n, "rowgroups": n_rg, "candidates": candidates(),
        "host_data_s": round(t_build, 1), "train_s": round(t_train, 2), "train_steps": steps, "batch": batch,
        "loss_curve": curve[:: max(1, len(curve) // 8)] + [curve[-1]],
        "exhaustive_ms": round(t_ex * 1e3, 3), "learned_ms": round(t_le * 1e3, 3),
        "speedup": round(t_ex / t_le, 2),
        "test_regret": regret(pred, test), "train_regret": regret(pred, ~test),
        "test_accuracy": float((pred[test] == best[test]).mean()),
        "no_network_test_regret": regret(base, test),
        "no_network_test_accuracy": float((base[test] == best[test]).mean()),
        "label_histogram": np.bincount(best, minlength=K).tolist(),
        "per_family_test_regret": {f: regret(pred, test & (np.array(fams)[col] == f)) for f in wl.WIZARD_FAMILIES},
    }
    res["meets_targets"] = bool(res["speedup"] >= 10 and res["test_regret"] <= 0.02)
    return (res, mlp) if return_model else res


def bench(dev):
    import torch

    return run(dev, torch.cuda.current_stream(dev))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--columns", type=int, default=512)
    ap.add_argument("--log2n", type=int, default=20)
    ap.add_argument("--steps", type=int, default=3000)
    a = ap.parse_args()
    import torch

    dev = torch.device("cuda:0")
    print(json.dumps(run(dev, torch.cuda.current_stream(dev), a.columns, 1 << a.log2n, a.steps)), flush=True)


if __name__ == "__main__":
    main()
