"""Learned initial guess: an in -> H -> H -> out GELU MLP trained on signals the
forward model itself generates, run on the matrix cores.

Workflow (all device work is done by the HIP kernels of csrc/mlp_lib.hip and
the generated module's pca_synth; torch only holds the buffers):

    net = MLPInit(IVIM, consts, hidden=64)
    data, truth = net.synthesize(10_000_000, seed=..., snr=(20, 100))
    net.fit_standardization(truth, data)
    X, T = net.prepare(data, truth)
    net.train(X, T, epochs=2, batch=8192)
    net.infer(data_t, lb_t, ub_t, params_t, fallback=(model, consts))  # MLP guess out

The network sees the signal normalised by its first measurement (b = 0 for
the diffusion protocols) and predicts standardised, transformed parameters:
a parameter relative to that signal level (S0), log-parameters for the rates
spanning decades, the rest as is. LM then refines from the prediction; voxels
whose prediction is not finite get the default log-linear guess, computed in
the inference kernel itself. Training: mean squared error on the standardised
targets, Adam.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import lm
from .models import ModelSpec

_P = ctypes.c_void_p
_I32, _I64, _F = ctypes.c_int32, ctypes.c_int64, ctypes.c_float

EPI_STORE, EPI_BIAS_GELU, EPI_BIAS, EPI_GELU_GRAD = 0, 1, 2, 3


class InferArgs(ctypes.Structure):
    _fields_ = [("data", _P), ("w1", _P), ("b1", _P), ("w2", _P), ("b2", _P), ("w3", _P),
                ("b3", _P), ("mu", _P), ("sd", _P), ("lb", _P), ("ub", _P), ("params", _P),
                ("N", _I64), ("in_", _I32), ("out", _I32), ("hidden", _I32),
                ("norm_index", _I32), ("scale_param", _I32), ("logshift_param", _I32),
                ("log_mask", _I32), ("reserved", _I32), ("fallback_P", _P),
                ("fallback_K", _I32), ("fallback_kind", _I32)]


def _bind():
    L = _lib.lib()
    if getattr(L, "_mlp_bound", False):
        return L
    L.pca_mlp_gemm.argtypes = [_P, _I64, _I32, _P, _I64, _I32, _P, _I64, _I32, _I32, _I32, _I32,
                               _P, _P, _P, _I32, _P]
    L.pca_mlp_mse_grad.argtypes = [_P, _P, _P, _P, _I64, _F, _P]
    L.pca_mlp_colsum.argtypes = [_P, _I64, _I32, _I32, _P, _P]
    L.pca_mlp_gemm_ordered.argtypes = [_P, _I64, _I32, _P, _I64, _I32, _P, _I64, _I32, _I32, _I32,
                                       _I32, _P, _I64, _P]
    L.pca_mlp_colsum_ordered.argtypes = [_P, _I64, _I32, _I32, _P, _P, _I64, _P]
    L.pca_mlp_adam.argtypes = [_P, _P, _P, _P, _I64, _F, _F, _F, _F, _I32, _F, _P]
    L.pca_mlp_fill.argtypes = [_P, _I64, _F, _P]
    L.pca_mlp_to_bf16.argtypes = [_P, _P, _I64, _P]
    L.pca_mlp_to_f16.argtypes = [_P, _P, _I64, _P]
    L.pca_mlp_prepare.argtypes = [_P, _P, _I64, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P,
                                  _P, _P, _P]
    L.pca_mlp_infer.argtypes = [ctypes.POINTER(InferArgs), _P]
    L.pca_mlp_train_step.argtypes = [_P, _I64, _P, _I32, _P, _I32, _I32, _I32,
                                     ctypes.POINTER(_I32), _I32, _P, _P, _P]
    L.pca_mlp_grad_reduce.argtypes = [_P, _P, _I32, _I32, _P, _P, _P, _P, _P, _F, _F, _F, _F,
                                      _I32, _F, _P]
    for n in ("pca_mlp_gemm", "pca_mlp_mse_grad", "pca_mlp_colsum", "pca_mlp_adam",
              "pca_mlp_fill", "pca_mlp_to_bf16", "pca_mlp_to_f16", "pca_mlp_prepare",
              "pca_mlp_infer", "pca_mlp_train_step", "pca_mlp_grad_reduce",
              "pca_mlp_gemm_ordered", "pca_mlp_colsum_ordered"):
        getattr(L, n).restype = ctypes.c_int
    L._mlp_bound = True
    return L


def _stream(dev=None):
    return torch.cuda.current_stream(dev).cuda_stream


def gemm(A, lda, transA, B, ldb, transB, C, ldc, M, N, K, epi=EPI_STORE, bias=None, Z=None,
         Zin=None, splits=1):
    """C = op(A) op(B) (+ epilogue) on the matrix cores (bf16 operands, fp32
    accumulate); arguments as pca_mlp_gemm. Tensors are fp32 on the device."""
    L = _bind()
    p = lambda t: None if t is None else t.data_ptr()
    _lib.check(L.pca_mlp_gemm(p(A), lda, int(transA), p(B), ldb, int(transB), p(C), ldc, M, N, K,
                              epi, p(bias), p(Z), p(Zin), splits, _stream()), "pca_mlp_gemm")


def _pad_in(n):
    return 16 if n <= 16 else (32 if n <= 32 else 64)


class MLPInit:
    def __init__(self, model: ModelSpec, consts, hidden: int = 128, norm_index: int = 0,
                 scale_param: int | None = None, logshift_param: int | None = None,
                 log_params=(), device="cuda", seed: int = 0):
        self.model = model
        self.consts = np.asarray(consts, np.float64)
        if self.consts.ndim == 1:
            self.consts = self.consts[None, :]
        self.IN = int(self.consts.shape[1])
        self.INP = _pad_in(self.IN)
        self.OUT = model.nparam
        self.H = int(hidden)
        if (self.IN > 64 or self.OUT > 32 or self.H not in (32, 64, 128, 256)
                or (self.H < 128 and (self.IN > 16 or self.OUT > 16))):
            raise ValueError("MLPInit supports <= 64 measurements, <= 32 parameters, hidden "
                             "32|64|128|256 (32, 64: <= 16 measurements and parameters)")
        self.norm_index = norm_index
        self.scale_param = -1 if scale_param is None else int(scale_param)
        self.logshift_param = -1 if logshift_param is None else int(logshift_param)
        self.log_mask = sum(1 << int(j) for j in log_params)
        self.device = torch.device(device)
        H, I, O = self.H, self.INP, self.OUT
        shapes = [("w1", (H, I)), ("b1", (H,)), ("w2", (H, H)), ("b2", (H,)), ("w3", (O, H)),
                  ("b3", (O,))]
        self.offsets = {}
        off = 0
        for name, shp in shapes:
            off = (off + 7) // 8 * 8   # 16-byte aligned blocks: 4-value weight loads
            self.offsets[name] = (off, shp)
            off += int(np.prod(shp))
        self.nparams = off
        rng = np.random.default_rng(seed)
        w = np.zeros(off, np.float32)
        for name, (o, shp) in self.offsets.items():
            if name.startswith("w"):
                fan_in = shp[1] if name != "w1" else self.IN
                bound = math.sqrt(6.0 / (fan_in + shp[0]))
                blk = rng.uniform(-bound, bound, shp).astype(np.float32)
                if name == "w1":
                    blk[:, self.IN:] = 0.0
                w[o:o + blk.size] = blk.ravel()
        kw = dict(dtype=torch.float32, device=self.device)
        self.w = torch.tensor(w, **kw)
        self.g = torch.zeros(off, **kw)
        self.m = torch.zeros(off, **kw)
        self.v = torch.zeros(off, **kw)
        self.wh = torch.zeros(off, dtype=torch.int16, device=self.device)   # fp16 bits
        self.mu = torch.zeros(O, **kw)
        self.sd = torch.ones(O, **kw)
        self.step = 0
        self.history = []
        self._buf = None

    # -- views ------------------------------------------------------------------
    def _view(self, t, name):
        o, shp = self.offsets[name]
        return t[o:o + int(np.prod(shp))].view(*shp)

    # -- data -------------------------------------------------------------------
    def synthesize(self, n: int, seed: int, snr=(20.0, 100.0), lo=None, hi=None, truth=None,
                   chunk: int = 1 << 21):
        """Signals (M, n) and parameters (nparam, n) from the model's own
        generated forward kernel (pca_synth) with Rician noise; parameters
        uniform in [lo, hi] unless `truth` (nparam, n) is given."""
        m = self.model
        prog = lm.get_program(lm.LMSpec(m.expr, m.pars, m.consts, self.IN))
        dev = self.device
        data = torch.empty((self.IN, n), dtype=torch.float32, device=dev)
        tru = torch.empty((self.OUT, n), dtype=torch.float32, device=dev)
        lo_t = torch.tensor(np.asarray(lo if lo is not None else m.lb, np.float32), device=dev)
        hi_t = torch.tensor(np.asarray(hi if hi is not None else m.ub, np.float32), device=dev)
        ucbuf, ucbytes = lm.uconsts_buffer(self.consts.reshape(-1), "float32")
        k = prog.kernel("pca_synth")

        class SynthArgs(ctypes.Structure):
            _fields_ = [("data", _P), ("truth", _P), ("truth_in", _P), ("lo", _P), ("hi", _P),
                        ("snr_lo", ctypes.c_double), ("snr_hi", ctypes.c_double),
                        ("seed", ctypes.c_uint64), ("N", _I64), ("offset", _I64)]

        L = _lib.lib()
        for s0 in range(0, n, chunk):
            cnt = min(chunk, n - s0)
            dsub = torch.empty((self.IN, cnt), dtype=torch.float32, device=dev)
            tsub = torch.empty((self.OUT, cnt), dtype=torch.float32, device=dev)
            tin = None
            if truth is not None:
                tin = truth[:, s0:s0 + cnt].to(torch.float32).contiguous()
            a = SynthArgs(dsub.data_ptr(), tsub.data_ptr(), None if tin is None else tin.data_ptr(),
                          lo_t.data_ptr(), hi_t.data_ptr(), float(snr[0]), float(snr[1]),
                          int(seed) & ((1 << 64) - 1), cnt, s0)
            params = (_P * 2)(ctypes.cast(ctypes.pointer(a), _P), ctypes.cast(ucbuf, _P))
            _lib.check(L.pca_launch(k, (cnt + 255) // 256, 1, 1, 256, 1, 1, 0, _stream(),
                                    params), "pca_synth")
            data[:, s0:s0 + cnt] = dsub
            tru[:, s0:s0 + cnt] = tsub
        return data, tru

    def _transform(self, truth, data):
        """Training-target transform (the inverse of what inference applies)."""
        t = truth.double().clone()
        s = data[self.norm_index].double()
        s = torch.where(s > 0, s, torch.ones_like(s))
        if self.scale_param >= 0:
            t[self.scale_param] /= s
        if self.logshift_param >= 0:
            t[self.logshift_param] -= torch.log(s)
        for j in range(self.OUT):
            if (self.log_mask >> j) & 1:
                t[j] = torch.log(t[j])
        return t

    def fit_standardization(self, truth, data):
        t = self._transform(truth, data)
        self.mu.copy_(t.mean(dim=1).float())
        self.sd.copy_(t.std(dim=1).clamp_min(1e-12).float())

    def prepare(self, data, truth):
        """X (INP, n) normalised signals, T (n, OUT) standardised targets."""
        n = data.shape[1]
        X = torch.empty((self.INP, n), dtype=torch.float32, device=self.device)
        T = torch.empty((n, self.OUT), dtype=torch.float32, device=self.device)
        L = _bind()
        _lib.check(L.pca_mlp_prepare(data.data_ptr(), truth.data_ptr(), n, self.IN, self.INP,
                                     self.OUT, self.norm_index, self.scale_param,
                                     self.logshift_param, self.log_mask, self.mu.data_ptr(),
                                     self.sd.data_ptr(), X.data_ptr(), T.data_ptr(), _stream()),
                   "pca_mlp_prepare")
        return X, T

    # -- training -----------------------------------------------------------------
    def _buffers(self, B):
        if self._buf is None or self._buf["B"] != B:
            kw = dict(dtype=torch.float32, device=self.device)
            H, O = self.H, self.OUT
            self._buf = {"B": B, "Z1": torch.empty((B, H), **kw), "H1": torch.empty((B, H), **kw),
                         "Z2": torch.empty((B, H), **kw), "H2": torch.empty((B, H), **kw),
                         "Y": torch.empty((B, O), **kw), "dY": torch.empty((B, O), **kw),
                         "dZ2": torch.empty((B, H), **kw), "dZ1": torch.empty((B, H), **kw),
                         "loss": torch.zeros(1, **kw)}
            # partial products of the split weight gradients and the chunk
            # column sums, summed in a fixed order (no atomics)
            split = max(1, min(64, B // 256))
            nws = max(split * H * max(H, self.INP), min(1024, (B + 63) // 64) * H)
            self._buf["ws"] = torch.empty(nws, **kw)
        return self._buf

    def forward_backward(self, X, T, off: int, B: int, ldx: int):
        """Gradients of the batch [off, off+B) into self.g (zeroed first);
        returns nothing (the loss accumulates into the loss buffer)."""
        L = _bind()
        b = self._buffers(B)
        H, O, I = self.H, self.OUT, self.INP
        w = {k: self._view(self.w, k) for k in self.offsets}
        g = {k: self._view(self.g, k) for k in self.offsets}
        Xb = X[:, off:]                         # (INP, ldx) column slice, ld = ldx
        Tb = T[off:off + B]
        split = max(1, min(64, B // 256))
        # forward
        gemm(Xb, ldx, 1, w["w1"], I, 1, b["H1"], H, B, H, I, EPI_BIAS_GELU, w["b1"], b["Z1"])
        gemm(b["H1"], H, 0, w["w2"], H, 1, b["H2"], H, B, H, H, EPI_BIAS_GELU, w["b2"], b["Z2"])
        gemm(b["H2"], H, 0, w["w3"], H, 1, b["Y"], O, B, O, H, EPI_BIAS, w["b3"])
        _lib.check(L.pca_mlp_mse_grad(b["Y"].data_ptr(), Tb.data_ptr(), b["dY"].data_ptr(),
                                      b["loss"].data_ptr(), B * O, 1.0 / (B * O), _stream()),
                   "pca_mlp_mse_grad")
        # backward: weight gradients and bias sums without atomics, so a step
        # repeats bit for bit
        _lib.check(L.pca_mlp_fill(self.g.data_ptr(), self.nparams, 0.0, _stream()), "fill")
        ws, nws = b["ws"].data_ptr(), b["ws"].numel()

        def wgrad(A, lda, Bm, ldb, transB, C, M, N):
            _lib.check(L.pca_mlp_gemm_ordered(A.data_ptr(), lda, 1, Bm.data_ptr(), ldb, transB,
                                              C.data_ptr(), N, M, N, B, split, ws, nws,
                                              _stream()), "pca_mlp_gemm_ordered")

        def bsum(X, N, out):
            _lib.check(L.pca_mlp_colsum_ordered(X.data_ptr(), N, B, N, out.data_ptr(), ws, nws,
                                                _stream()), "pca_mlp_colsum_ordered")

        wgrad(b["dY"], O, b["H2"], H, 0, g["w3"], O, H)
        bsum(b["dY"], O, g["b3"])
        gemm(b["dY"], O, 0, w["w3"], H, 0, b["dZ2"], H, B, H, O, EPI_GELU_GRAD, Zin=b["Z2"])
        wgrad(b["dZ2"], H, b["H1"], H, 0, g["w2"], H, H)
        bsum(b["dZ2"], H, g["b2"])
        gemm(b["dZ2"], H, 0, w["w2"], H, 0, b["dZ1"], H, B, H, H, EPI_GELU_GRAD, Zin=b["Z1"])
        wgrad(b["dZ1"], H, Xb, ldx, 1, g["w1"], H, I)
        bsum(b["dZ1"], H, g["b1"])

    # -- fused training step ------------------------------------------------------
    FUSED_ROWS = 64   # batch rows per workgroup of pca_mlp_train_step

    def fused_supported(self, B: int) -> bool:
        """pca_mlp_train_step covers hidden 32 and 64 (16 padded inputs, <= 16
        outputs) and 128 (16, 32 or 64 padded inputs, <= 32 outputs), batches
        of a multiple of 64 rows; wider networks train on the GEMM path."""
        if B <= 0 or B % self.FUSED_ROWS:
            return False
        if self.H in (32, 64):
            return self.INP == 16 and self.OUT <= 16
        return self.H == 128

    def _fused_ws(self, B):
        nwg = B // self.FUSED_ROWS
        ws = getattr(self, "_ws", None)
        if ws is None or ws["nwg"] != nwg:
            kw = dict(dtype=torch.float32, device=self.device)
            # zeros once: entries between the parameter blocks are never written
            ws = {"nwg": nwg, "part": torch.zeros((nwg, self.nparams), **kw),
                  "loss_part": torch.zeros(nwg, **kw), "loss": torch.zeros(1, **kw),
                  "offsets": (_I32 * 6)(*[self.offsets[k][0]
                                          for k in ("w1", "b1", "w2", "b2", "w3", "b3")])}
            self._ws = ws
        return ws

    def fused_step(self, X, T, off: int, B: int, ldx: int, lr=None, b1=0.9, b2=0.999, eps=1e-8,
                   weight_decay=0.0):
        """One training step on the batch [off, off+B) in two kernels:
        pca_mlp_train_step (forward, loss gradient and backward, one
        workgroup per 64 rows, partial gradients to a workspace) and
        pca_mlp_grad_reduce (the sums into self.g, the loss into the fused
        loss buffer, and with lr given one Adam step). With lr None the
        weights are left as they are."""
        if not self.fused_supported(B):
            raise ValueError(f"fused training step: hidden {self.H}, {self.INP} padded inputs, "
                             f"{self.OUT} outputs, batch {B} not covered")
        if off < 0 or off + B > ldx or tuple(X.shape) != (self.INP, ldx) \
                or T.shape[0] < off + B or T.shape[1] != self.OUT:
            raise ValueError("fused training step: X (in_padded, ldx), T (>= off + B, out) "
                             "and off + B <= ldx")
        if not (X.is_contiguous() and T.is_contiguous() and X.dtype == T.dtype == torch.float32):
            raise ValueError("fused training step: contiguous float32 X and T")
        ws = self._fused_ws(B)
        self._fused_launch(ws, X.data_ptr() + 4 * off, T.data_ptr() + 4 * off * self.OUT, B, ldx,
                           lr, b1, b2, eps, weight_decay)

    def _fused_launch(self, ws, xp, tp, B, ldx, lr, b1=0.9, b2=0.999, eps=1e-8,
                      weight_decay=0.0):
        L = _bind()
        st = _stream()
        _lib.check(L.pca_mlp_train_step(xp, ldx, tp, B, self.w.data_ptr(), self.H, self.INP,
                                        self.OUT, ws["offsets"], self.nparams,
                                        ws["part"].data_ptr(), ws["loss_part"].data_ptr(), st),
                   "pca_mlp_train_step")
        if lr is None:
            wp = mp = vp = None
            step = 0
        else:
            self.step += 1
            wp, mp, vp, step = self.w.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self.step
        _lib.check(L.pca_mlp_grad_reduce(ws["part"].data_ptr(), ws["loss_part"].data_ptr(),
                                         ws["nwg"], self.nparams, self.g.data_ptr(),
                                         ws["loss"].data_ptr(), wp, mp, vp, lr or 0.0, b1, b2,
                                         eps, step, weight_decay, st), "pca_mlp_grad_reduce")

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8, weight_decay=0.0):
        L = _bind()
        self.step += 1
        # the padded input columns of w1 carry no gradient and stay zero
        _lib.check(L.pca_mlp_adam(self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(),
                                  self.v.data_ptr(), self.nparams, lr, b1, b2, eps, self.step,
                                  weight_decay, _stream()), "pca_mlp_adam")

    def train(self, X, T, epochs: int = 2, batch: int = 8192, lr: float = 2e-3,
              lr_min: float = 2e-5, log=None, fused=None):
        """Adam with a cosine learning-rate schedule over epochs x (n // batch)
        steps. fused None: the fused step where it covers the network and the
        batch, else the GEMM path; True or False forces one."""
        n = X.shape[1]
        steps_per_epoch = n // batch
        total = max(1, epochs * steps_per_epoch)
        if fused is None:
            fused = self.fused_supported(batch)
        loss_t = self._fused_ws(batch)["loss"] if fused else self._buffers(batch)["loss"]
        for e in range(epochs):
            loss_t.zero_()
            for s in range(steps_per_epoch):
                k = e * steps_per_epoch + s
                lr_k = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / total))
                if fused:
                    if k == 0:   # argument checks once; the loop launches directly
                        self.fused_step(X, T, 0, batch, n, lr=lr_k)
                        ws = self._fused_ws(batch)
                        xp, tp = X.data_ptr(), T.data_ptr()
                    else:
                        self._fused_launch(ws, xp + 4 * s * batch,
                                           tp + 4 * s * batch * self.OUT, batch, n, lr_k)
                else:
                    self.forward_backward(X, T, s * batch, batch, n)
                    self.adam_step(lr_k)
            loss = float(loss_t.item()) / max(1, steps_per_epoch)
            self.history.append(loss)
            if log:
                log(f"mlp epoch {e}: mean loss {loss:.5f}")
        self.finalize()
        return self.history

    def finalize(self):
        """fp16 copies of the weights for inference."""
        L = _bind()
        _lib.check(L.pca_mlp_to_f16(self.w.data_ptr(), self.wh.data_ptr(), self.nparams,
                                    _stream()), "pca_mlp_to_f16")

    # -- inference ------------------------------------------------------------------
    def infer(self, data_t, lb_t, ub_t, params_t, stream=None, fallback=None):
        """params_t (out, N): the clipped MLP guess out. Voxels whose outputs
        are not finite get the log-linear guess of ``fallback`` = (model,
        consts), computed in the same kernel; with fallback None they keep
        what params_t held."""
        L = _bind()
        N = int(data_t.shape[1])
        fP, fK, fkind = None, 0, 0
        if fallback is not None:
            from . import init as _init
            P, fkind = _init.loglinear_spec(fallback[0], fallback[1], data_t.device)
            if P.shape[1] != self.IN:
                raise ValueError(f"fallback guess is for {P.shape[1]} measurements, the "
                                 f"network takes {self.IN}")
            fP, fK = P.data_ptr(), int(P.shape[0])
        wb = {k: self._view(self.wh, k) for k in self.offsets}
        wf = {k: self._view(self.w, k) for k in self.offsets}
        a = InferArgs(data_t.data_ptr(), wb["w1"].data_ptr(), wf["b1"].data_ptr(),
                      wb["w2"].data_ptr(), wf["b2"].data_ptr(), wb["w3"].data_ptr(),
                      wf["b3"].data_ptr(), self.mu.data_ptr(), self.sd.data_ptr(),
                      lb_t.data_ptr(), ub_t.data_ptr(), params_t.data_ptr(), N, self.IN,
                      self.OUT, self.H, self.norm_index, self.scale_param, self.logshift_param,
                      self.log_mask, 0, fP, fK, fkind)
        self._infer_args = a  # kept alive for graph capture
        _lib.check(L.pca_mlp_infer(ctypes.byref(a), _stream() if stream is None else stream),
                   "pca_mlp_infer")

    # -- persistence ------------------------------------------------------------------
    def state_dict(self):
        return {"w": self.w.cpu(), "mu": self.mu.cpu(), "sd": self.sd.cpu(), "H": self.H,
                "IN": self.IN, "OUT": self.OUT, "history": list(self.history)}

    def load_state_dict(self, st):
        self.w.copy_(st["w"].to(self.device))
        self.mu.copy_(st["mu"].to(self.device))
        self.sd.copy_(st["sd"].to(self.device))
        self.finalize()


def ivim_mlp(consts, hidden=64, device="cuda", seed=0) -> MLPInit:
    """The IVIM initialiser: S0 relative to the b=0 signal, log D_1 and log D_2.
    Hidden width 64: on the 16-b-value volume its guess needs the same LM
    iteration cap (6) as width 128 at 58% of the inference time; width 32
    needs a cap of 8."""
    from .models import IVIM
    return MLPInit(IVIM, consts, hidden=hidden, norm_index=0, scale_param=0,
                   log_params=(2, 3), device=device, seed=seed)


def dki_mlp(consts, hidden=256, device="cuda", seed=0) -> MLPInit:
    """The kurtosis-model initialiser (61 -> 256 -> 256 -> 22): signals
    relative to the b=0 measurement, lnS0 relative to its log."""
    from .models import DKI
    return MLPInit(DKI, consts, hidden=hidden, norm_index=0, logshift_param=0, device=device,
                   seed=seed)
