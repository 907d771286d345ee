"""Learned aberration initializer: a small CNN trained on the GPU with the
library's own bf16/fp32-accumulate kernels (``csrc/ssb_learn.hip``).

No reference counterpart exists.  The network maps trotter feature maps --
for 8 fixed scan frequencies Q, the window of G(Q, k) over the detector
around the BF disk as |G|/max|G| and the phase of G(k) conj(G(k_centre))
(16 channels, 32 x 32) -- to (C10, C30, rotation).  It is trained on the
weak-phase SSB forward model G_b(Q) = Gamma_b(Q; theta) Phi(Q) + noise in
the engine's own conventions, generated on the GPU per batch.  Its output
only seeds a local batched search on the exact SSB objective, which scores
it; the global batched search stays the fallback.

Architecture: four 3x3 stride-2 convolutions (16->32->64->64->64, ReLU,
32x32 -> 2x2), then 256 -> 256 (ReLU) -> 3; 163k parameters.  Targets are
scaled to O(1): (C10 / 500 A, C30 / 2e5 A, rotation / 5 deg).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from .compute.hip import _abi

CONV = ((16, 32), (32, 64), (64, 64), (64, 64))
FC = ((256, 256), (256, 3))
WIN = 32
NQ = 8
SCALE = np.array([500.0, 2.0e5, math.radians(5.0)])


def _lib():
    lib = _abi.lib()
    vp, i64, ci, cf = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_float
    sig = {
        "qnn_gemm_bf16": [vp, i64, i64, vp, i64, i64, vp, i64, i64, vp, ci, ci, ci, ci, ci, vp],
        "qnn_gemm_bf16_splitk": [vp, i64, i64, vp, i64, i64, vp, ci, ci, ci, ci, vp, vp],
        "qnn_im2col": [vp, ci, ci, ci, ci, vp, vp],
        "qnn_col2im": [vp, ci, ci, ci, ci, vp, vp],
        "qnn_cast_bf16": [vp, i64, vp, vp],
        "qnn_relu_bwd": [vp, vp, i64, vp, vp, vp],
        "qnn_colsum_workspace": [ci, ci],
        "qnn_colsum": [vp, ci, ci, vp, vp, ctypes.c_longlong, vp],
        "qnn_mse": [vp, vp, ci, vp, vp, vp],
        "qnn_adam": [vp, vp, vp, vp, i64, cf, cf, cf, cf, ci, vp, vp],
        "qnn_features_synth": [vp, vp, ci, cf, cf, cf, cf, vp, ctypes.c_uint, vp, vp],
    }
    for name, argtypes in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ci
    lib.qnn_colsum_workspace.restype = i64
    return lib


def _call(name, *args):
    rc = getattr(_lib(), name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed with status {rc}")


def _p(t):
    return ctypes.c_void_p(0) if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(torch, dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


@dataclass
class FeatureConfig:
    """Optics and the 8 fixed scan frequencies (1/A) the features sample."""
    recip: float          # detector reciprocal sampling, 1/A per pixel
    semiangle_rad: float
    ang_rad: float        # angular sampling per detector pixel
    wavelength: float
    q: np.ndarray         # [8, 2] (qx, qy)

    @staticmethod
    def for_scan(n: int, scan_step: float, semiangle_rad: float, ang_rad: float, wavelength: float):
        """Q on the scan's own frequency grid: radii at 0.3, 0.6, 0.9 of the
        aperture radius and four directions, as (row, col) grid indices."""
        dq = 1.0 / (n * scan_step)
        rad = semiangle_rad / wavelength
        picks = []
        for frac, angles in ((0.3, (0.0, 90.0)), (0.6, (0.0, 45.0, 90.0, 135.0)), (0.9, (30.0, 120.0))):
            for a in angles:
                t = math.radians(a)
                picks.append((int(round(frac * rad * math.cos(t) / dq)), int(round(frac * rad * math.sin(t) / dq))))
        idx = np.array([[r % n, c % n] for r, c in picks], dtype=np.int32)
        freq = np.fft.fftfreq(n, scan_step)
        q = np.stack([freq[idx[:, 0]], freq[idx[:, 1]]], axis=1)
        cfg = FeatureConfig(ang_rad / wavelength, semiangle_rad, ang_rad, wavelength, q)
        cfg.index = idx
        return cfg


class LearnedInit:
    """The network, its parameters (fp32 master + bf16 copies) and Adam state."""

    def __init__(self, device=None, seed: int = 0):
        import torch
        self.torch = torch
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        rng = np.random.default_rng(seed)
        self.params = []      # (W [out, in], b [out]) fp32 on the GPU
        for cin, cout in CONV:
            k = 9 * cin
            self.params.append((rng.standard_normal((cout, k)) * math.sqrt(2.0 / k), np.zeros(cout)))
        for fin, fout in FC:
            scale = math.sqrt(2.0 / fin) if fout != 3 else math.sqrt(1.0 / fin)
            self.params.append((rng.standard_normal((fout, fin)) * scale, np.zeros(fout)))
        self.params = [(torch.tensor(w, dtype=torch.float32, device=self.dev),
                        torch.tensor(b, dtype=torch.float32, device=self.dev)) for w, b in self.params]
        self.half = [torch.empty(w.shape, dtype=torch.int16, device=self.dev) for w, _ in self.params]
        self.adam = [tuple(torch.zeros_like(t) for t in (w, w, b, b)) for w, b in self.params]
        self.step = 0
        self._sync_half()

    # ------------------------------------------------------------- plumbing
    def _s(self):
        return _stream(self.torch, self.dev)

    def _cast(self, x, out):
        _call("qnn_cast_bf16", _p(x), int(x.numel()), _p(out), self._s())
        return out

    def _sync_half(self):
        for (w, _), h in zip(self.params, self.half):
            self._cast(w, h)

    def _gemm(self, A, sam, sak, B, sbk, sbn, C, scm, scn, bias, M, N, K, relu=False, acc=False):
        _call("qnn_gemm_bf16", _p(A), sam, sak, _p(B), sbk, sbn, _p(C), scm, scn, _p(bias), M, N, K,
              1 if relu else 0, 1 if acc else 0, self._s())

    def _colsum(self, dz, M, N, out):
        """Bias gradient: column sums of dz[M][N] (qnn_colsum, two fixed-order passes)."""
        need = int(_lib().qnn_colsum_workspace(M, N))
        ws = self.torch.empty(max(need, 1), dtype=self.torch.float32, device=self.dev)
        _call("qnn_colsum", _p(dz), M, N, _p(out), _p(ws), need, self._s())

    def _gemm_wgrad(self, A, sam, sak, B, sbk, sbn, C, M, N, K):
        """Weight gradient C[M][N] with a long reduction K, split over K."""
        splits = max(1, min(64, K // 1024))
        if splits == 1:
            self._gemm(A, sam, sak, B, sbk, sbn, C, N, 1, None, M, N, K)
            return
        ws = self.torch.empty(splits * M * N, dtype=self.torch.float32, device=self.dev)
        _call("qnn_gemm_bf16_splitk", _p(A), sam, sak, _p(B), sbk, sbn, _p(C), M, N, K, splits, _p(ws), self._s())

    def state_dict(self) -> dict:
        out = {}
        for i, (w, b) in enumerate(self.params):
            out[f"w{i}"] = w.cpu().numpy()
            out[f"b{i}"] = b.cpu().numpy()
        return out

    def load_state_dict(self, state: dict) -> None:
        for i, (w, b) in enumerate(self.params):
            w.copy_(self.torch.as_tensor(state[f"w{i}"], dtype=self.torch.float32))
            b.copy_(self.torch.as_tensor(state[f"b{i}"], dtype=self.torch.float32))
        self._sync_half()

    # -------------------------------------------------------------- forward
    def forward(self, X):
        """X: [B, 32, 32, 16] fp32 NHWC on the GPU -> [B, 3] (scaled targets)."""
        torch = self.torch
        Bn = int(X.shape[0])
        cache = []
        h = X.contiguous()
        H = WIN
        for li, (cin, cout) in enumerate(CONV):
            Ho = H // 2
            M, K = Bn * Ho * Ho, 9 * cin
            cols = torch.empty((M, K), dtype=torch.int16, device=self.dev)
            _call("qnn_im2col", _p(h), Bn, H, H, cin, _p(cols), self._s())
            y = torch.empty((M, cout), dtype=torch.float32, device=self.dev)
            w, b = self.params[li]
            self._gemm(cols, K, 1, self.half[li], 1, K, y, cout, 1, b, M, cout, K, relu=True)
            cache.append((h, cols, y, H))
            h, H = y, Ho
        flat = h.reshape(Bn, -1)
        for fi, (fin, fout) in enumerate(FC):
            li = len(CONV) + fi
            xh = self._cast(flat, torch.empty((Bn, fin), dtype=torch.int16, device=self.dev))
            y = torch.empty((Bn, fout), dtype=torch.float32, device=self.dev)
            self._gemm(xh, fin, 1, self.half[li], 1, fin, y, fout, 1, self.params[li][1], Bn, fout, fin,
                       relu=(fi == 0))
            cache.append((flat, xh, y, None))
            flat = y
        self._cache = cache
        return flat

    # ------------------------------------------------------------- backward
    def backward(self, dpred):
        """Gradients of the last forward for d loss / d pred; returns [(dW, db)]."""
        torch = self.torch
        grads = [None] * len(self.params)
        dy = dpred.contiguous()
        Bn = int(dy.shape[0])
        for fi in reversed(range(len(FC))):
            li = len(CONV) + fi
            fin, fout = FC[fi]
            x, xh, y, _ = self._cache[li]
            dz = torch.empty_like(dy)
            dzh = torch.empty(dy.shape, dtype=torch.int16, device=self.dev)
            _call("qnn_relu_bwd", _p(y) if fi == 0 else None, _p(dy), int(dy.numel()), _p(dz), _p(dzh), self._s())
            dw = torch.empty((fout, fin), dtype=torch.float32, device=self.dev)
            self._gemm_wgrad(dzh, 1, fout, xh, fin, 1, dw, fout, fin, Bn)
            db = torch.empty(fout, dtype=torch.float32, device=self.dev)
            self._colsum(dz, Bn, fout, db)
            grads[li] = (dw, db)
            dx = torch.empty((Bn, fin), dtype=torch.float32, device=self.dev)
            self._gemm(dzh, fout, 1, self.half[li], fin, 1, dx, fin, 1, None, Bn, fin, fout)
            dy = dx
        for li in reversed(range(len(CONV))):
            cin, cout = CONV[li]
            x, cols, y, H = self._cache[li]
            Ho = H // 2
            M, K = Bn * Ho * Ho, 9 * cin
            dyv = dy.reshape(M, cout)
            dz = torch.empty((M, cout), dtype=torch.float32, device=self.dev)
            dzh = torch.empty((M, cout), dtype=torch.int16, device=self.dev)
            _call("qnn_relu_bwd", _p(y), _p(dyv), M * cout, _p(dz), _p(dzh), self._s())
            dw = torch.empty((cout, K), dtype=torch.float32, device=self.dev)
            self._gemm_wgrad(dzh, 1, cout, cols, K, 1, dw, cout, K, M)
            db = torch.empty(cout, dtype=torch.float32, device=self.dev)
            self._colsum(dz, M, cout, db)
            grads[li] = (dw, db)
            if li > 0:
                dcols = torch.empty((M, K), dtype=torch.float32, device=self.dev)
                self._gemm(dzh, cout, 1, self.half[li], K, 1, dcols, K, 1, None, M, K, cout)
                dx = torch.empty((Bn, H, H, cin), dtype=torch.float32, device=self.dev)
                _call("qnn_col2im", _p(dcols), Bn, H, H, cin, _p(dx), self._s())
                dy = dx
        return grads

    def loss_and_grad(self, X, target):
        torch = self.torch
        pred = self.forward(X)
        dpred = torch.empty_like(pred)
        loss = torch.empty(1, dtype=torch.float32, device=self.dev)
        _call("qnn_mse", _p(pred), _p(target), int(pred.numel()), _p(dpred), _p(loss), self._s())
        return loss, self.backward(dpred)

    def apply(self, grads, lr: float, b1=0.9, b2=0.999, eps=1e-8) -> None:
        self.step += 1
        for i, ((w, b), (dw, db), (mw, vw, mb, vb)) in enumerate(zip(self.params, grads, self.adam)):
            _call("qnn_adam", _p(w), _p(dw), _p(mw), _p(vw), int(w.numel()), lr, b1, b2, eps, self.step,
                  _p(self.half[i]), self._s())
            _call("qnn_adam", _p(b), _p(db), _p(mb), _p(vb), int(b.numel()), lr, b1, b2, eps, self.step, None,
                  self._s())

    # ---------------------------------------------------------------- data
    def synth_batch(self, cfg: FeatureConfig, rng: np.random.Generator, n: int, ranges=None, seed: int = 0):
        """Features and scaled targets of n forward-model samples drawn from ``ranges``
        (C10 A, C30 A, rotation rad) with log-uniform noise."""
        torch = self.torch
        lo = np.array([r[0] for r in ranges]) if ranges else -SCALE
        hi = np.array([r[1] for r in ranges]) if ranges else SCALE
        theta = lo + (hi - lo) * rng.random((n, 3))
        snr = np.exp(rng.uniform(np.log(0.05), np.log(1.5), n))
        th = torch.tensor(theta, dtype=torch.float32, device=self.dev)
        sn = torch.tensor(snr, dtype=torch.float32, device=self.dev)
        q = np.ascontiguousarray(cfg.q.astype(np.float32).reshape(-1))
        X = torch.empty((n, WIN, WIN, 2 * NQ), dtype=torch.float32, device=self.dev)
        _call("qnn_features_synth", _p(th), _p(sn), n, cfg.recip, cfg.semiangle_rad, cfg.ang_rad, cfg.wavelength,
              q.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint(seed & 0xffffffff), _p(X), self._s())
        target = torch.tensor(theta / SCALE, dtype=torch.float32, device=self.dev)
        return X, target

    def train(self, cfg: FeatureConfig, steps: int = 2000, batch: int = 256, lr: float = 1e-3, seed: int = 0,
              log_every: int = 0):
        """Adam on fresh forward-model batches; returns the loss history."""
        rng = np.random.default_rng(seed)
        hist = []
        for it in range(steps):
            X, t = self.synth_batch(cfg, rng, batch, seed=seed * 1000003 + it)
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * it / steps))
            loss, grads = self.loss_and_grad(X, t)
            self.apply(grads, rate)
            if log_every and (it % log_every == 0 or it == steps - 1):
                hist.append((it, float(loss.item())))
        return hist

    def predict(self, X) -> np.ndarray:
        """(C10 A, C30 A, rotation rad) for each feature map."""
        pred = self.forward(X)
        return pred.cpu().numpy().astype(np.float64) * SCALE


def features_from_engine(engine, cfg: FeatureConfig):
    """Features of a resident Fourier stack (``HipSSBEngine`` after ``prepare_g``)."""
    torch = engine.torch
    out = torch.empty((1, WIN, WIN, 2 * NQ), dtype=torch.float32, device=engine.plan.device)
    r = np.ascontiguousarray(cfg.index[:, 0].astype(np.int32))
    c = np.ascontiguousarray(cfg.index[:, 1].astype(np.int32))
    _abi.call("ssb_learn_features", engine.plan.handle, r.ctypes.data_as(_abi._ip), c.ctypes.data_as(_abi._ip),
              _abi.ptr(out), engine._stream())
    return out
