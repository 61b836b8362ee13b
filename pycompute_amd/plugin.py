"""The F_GradF plugin path (pycompute/cuda/lsqnonlin.py:32-38, 814-816, 939-941).

A user supplies the model as device code instead of an expression: two
HipFunction objects whose device functions have the reference signatures

    void <gradf>(const real* params, const real* consts, const real* data,
                 const real* lam, real* f, real* g, real* h, real* hl,
                 int tid, int Nprobs)   // fghhl: accumulate f, g, h, hl
    void <f>(const real* params, const real* consts, const real* data,
             real* f, int tid, int Nprobs)  // accumulate f += sum r^2

(real = float or double) operating on (rows, N) global arrays. They are
linked with code_gen_walking (their get_deps() included) ahead of the plugin
kernel (csrc/plugin_kernel.hip), which keeps the reference's global-memory
workspaces and runs all iterations in one launch. consts are passed through
untouched, in whatever layout the user's code indexes.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from . import lm
from .program import Module, gen_deps_dict, read_csrc


class PluginArgs(ctypes.Structure):
    _fields_ = [("a", _lib.LMArgs), ("g", ctypes.c_void_p), ("h", ctypes.c_void_p),
                ("hl", ctypes.c_void_p), ("ftp", ctypes.c_void_p), ("params_tp", ctypes.c_void_p),
                ("step", ctypes.c_void_p)]


class PluginLM:
    per_voxel = True  # consts are handed to the user's code as given

    def __init__(self, f_gradf, ndata: int, dtype: str, block: int):
        self.fg = f_gradf
        self.ndata = int(ndata)
        self.dtype = dtype
        self.block = int(block)
        t = "f" if dtype == "float32" else "d"
        # reference naming: ..._{n}_{c}_{M}_{t}_{name} (lsqnonlin.py:814-816)
        self.funcid = (f"second_order_levenberg_marquardt_{f_gradf.nparam}_{f_gradf.nconst}_"
                       f"{self.ndata}_{t}_{f_gradf.name}")
        self._program = None

    def source(self) -> str:
        deps = {}
        for fn in (self.fg.gradf, self.fg.f):
            gen_deps_dict(fn, deps)
            deps.setdefault(fn.get_device_funcid(), fn.get_device_code())
        user = "\n".join(deps.values())
        n = self.fg.nparam
        head = [f"// generated: {self.funcid} (plugin)",
                f"#define PCA_NP {n}", f"#define PCA_M {self.ndata}",
                f"#define PCA_BLOCK {self.block}",
                f"#define PCA_GRADF {self.fg.gradf.get_device_funcid()}",
                f"#define PCA_FOBJ {self.fg.f.get_device_funcid()}"]
        if self.dtype == "float64":
            head.append("#define PCA_REAL_DOUBLE 1")
        return "\n".join(["\n".join(head), read_csrc("strict_math.h"), read_csrc("prelude.hip"),
                          read_csrc("lm_solver.hip"), user, read_csrc("plugin_kernel.hip")])

    def program(self, consts_per_voxel: bool = True):
        if self._program is None:
            src = self.source()

            class _Prog:
                pass

            p = _Prog()
            p.source = src
            p.module = Module(src, self.funcid)
            p.kernel = p.module.kernel
            self._program = p
        return self._program

    def run(self, solver, iters, tol, mu, eta, acc, dec):
        prog = self.program()
        N = solver.Nprobs
        n, nh = self.fg.nparam, self.fg.nparam * (self.fg.nparam + 1) // 2
        kw = dict(dtype=solver.dtype, device=solver.data_t.device)
        if getattr(solver, "g_t", None) is None or solver.g_t.shape != (n, N):
            solver.g_t = torch.zeros((n, N), **kw)
            solver.h_t = torch.zeros((nh, N), **kw)
            solver.hl_t = torch.zeros((nh, N), **kw)
            solver.step_t = torch.zeros((n, N), **kw)
        if solver.ftp_t is None:
            solver.ftp_t = torch.zeros((1, N), **kw)
            solver.pars_tp_t = torch.zeros((n, N), **kw)
        a = solver.lm_args(iters, tol, mu, eta, acc, dec)
        a.consts = solver.consts_t.data_ptr()
        a.g_out = a.h_out = a.hl_out = a.step_out = None
        P = PluginArgs(a, solver.g_t.data_ptr(), solver.h_t.data_ptr(), solver.hl_t.data_ptr(),
                       solver.ftp_t.data_ptr(), solver.pars_tp_t.data_ptr(),
                       solver.step_t.data_ptr())
        params = (ctypes.c_void_p * 1)(ctypes.cast(ctypes.pointer(P), ctypes.c_void_p))
        grid = (N + self.block - 1) // self.block
        _lib.check(_lib.lib().pca_launch(prog.kernel("pca_lm_plugin"), grid, 1, 1, self.block, 1,
                                         1, 0, lm.stream_handle(), params), "pca_lm_plugin")
