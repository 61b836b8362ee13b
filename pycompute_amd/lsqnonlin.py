"""Drop-in host API of the reference's LM fitter (pycompute/cuda/lsqnonlin.py),
over torch tensors on a ROCm device.

    solm = SecondOrderLevenbergMarquardt(expr, pars_str, consts_str, ndata=M, dtype=torch.float32)
    solm.setup(pars_t, consts_t, data_t, lower_bound_t, upper_bound_t)
    solm.run(iters, tol, mu, eta, acc, dec)
    solm.pars_t, solm.first_f, solm.last_f, solm.step_type_t, solm.lam_t ...

Both constructor forms of the reference are accepted: the committed one
``(model_expr: Expr | None, f_gradf: F_GradF | None, ndata, dtype, write_to_file)``
(lsqnonlin.py:950) and the notebook one ``(expr, pars_str, consts_str, ndata,
dtype, write_to_file)`` (test/ivim/test_ivim.ipynb:366).

Differences from the reference, by design:
  * run() is ONE launch that performs all ``iters`` iterations in registers
    (the reference launches once per iteration); results are those of the
    reference's per-iteration loop, including first_f/last_f and resumption.
  * Scalars are converted to the kernel's element type (the reference passes a
    Python float to a ``float`` parameter).
  * Voxel-uniform consts are detected and passed by value; per-voxel consts
    use the reference (nconst*M, N) layout with index (k*M+i)*N+v.
  * The workspace attributes g_t/h_t/hl_t/step_t hold the last iteration's
    values when ``keep_workspace`` is on (default for n <= 8); pars_tp_t and
    ftp_t are internal to the kernel and stay None.
"""
from __future__ import annotations

import copy
import ctypes

import torch

from . import _lib
from . import chunked
from . import lm
from .program import HipFunction
from .sym import hipgen


class Expr:
    """Model specification (lsqnonlin.py:21-28). nonlin_terms / nonlin_pars are
    stored and unused, as in the reference."""

    def __init__(self, expr: str, pars_str: list, consts_str: list,
                 nonlin_terms: list | None = None, nonlin_pars: list | None = None):
        self.expr = expr
        self.pars_str = list(pars_str)
        self.consts_str = list(consts_str)
        self.nonlin_terms = nonlin_terms
        self.nonlin_pars = nonlin_pars


class F_GradF:
    """Plugin hook (lsqnonlin.py:32-38): user-supplied model code replacing the
    symbolic one. ``f`` and ``gradf`` are HipFunction objects; see
    pycompute_amd.plugin for the device-function signatures they provide."""

    def __init__(self, f: HipFunction, gradf: HipFunction, nparam: int, nconst: int, name: str):
        self.f = f
        self.gradf = gradf
        self.nparam = nparam
        self.nconst = nconst
        self.name = name


def _bind(args, kwargs, names, defaults):
    """Bind positional/keyword arguments of one of the reference's signatures."""
    if len(args) > len(names):
        raise TypeError(f"too many positional arguments (expected at most {len(names)})")
    out = dict(defaults)
    for n, v in zip(names, args):
        out[n] = v
    for k, v in kwargs.items():
        if k not in names:
            raise TypeError(f"unexpected keyword argument {k!r}")
        if k in names[:len(args)]:
            raise TypeError(f"got multiple values for argument {k!r}")
        out[k] = v
    missing = [n for n in names if n not in out]
    if missing:
        raise TypeError(f"missing arguments: {missing}")
    return out


def _check_device_tensor(t, name, dtype):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on a ROCm device, got {type(t)}")
    if not t.is_cuda:
        raise ValueError(f"{name} must be on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"{name} has dtype {t.dtype}, expected {dtype}")


def _as_rows(t, rows, N, name):
    """View t as contiguous (rows, N); accepts (rows, N) or (1, rows, N)."""
    if t.dim() == 3 and t.shape[0] == 1:
        t = t[0]
    if t.dim() == 1 and rows == 1:
        t = t.unsqueeze(0)
    if tuple(t.shape) != (rows, N):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected ({rows}, {N})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous (rows, N)")
    return t


def split_consts(consts_t, nconst, M, N, uniform=None):
    """Return (uniform_values | None, per_voxel_tensor | None).

    Accepted layouts: (nconst*M, N) or (1, nconst*M, N) per voxel (the
    reference's replicated layout, test_ivim.ipynb:111-112), (nconst*M,),
    (nconst, M) or (nconst*M, 1) uniform, or a stride-0 expand of a uniform
    column. ``uniform=None`` detects replicated columns with one device
    comparison.
    """
    c = consts_t
    if c.dim() == 3 and c.shape[0] == 1:
        c = c[0]
    ncm = nconst * M
    if c.dim() == 1 or (c.dim() == 2 and tuple(c.shape) == (nconst, M)):
        if c.numel() != ncm:
            raise ValueError(f"consts has {c.numel()} values, expected nconst*M = {ncm}")
        return c.reshape(-1).detach().cpu().tolist(), None
    if c.dim() != 2 or c.shape[0] != ncm or c.shape[1] not in (1, N):
        raise ValueError(f"consts has shape {tuple(consts_t.shape)}, expected ({ncm}, {N})")
    if N == 0 and c.shape[1] == 0:             # no voxels: nothing to tell apart
        return None, c.contiguous()
    if c.shape[1] == 1 or (N > 0 and c.stride(1) == 0):
        return c[:, 0].detach().cpu().tolist(), None
    if uniform is None:
        uniform = bool(torch.equal(c, c[:, :1].expand_as(c)))
    if uniform:
        return c[:, 0].detach().cpu().tolist(), None
    if not c.is_contiguous():
        raise ValueError("per-voxel consts must be C-contiguous (nconst*M, N)")
    return None, c


class SecondOrderLevenbergMarquardt(HipFunction):
    """Batched second-order LM with the GMW81 step (lsqnonlin.py:949-1073)."""

    def __init__(self, *args, block: int = lm.DEFAULT_BLOCK, fast_exp: bool = False,
                 keep_workspace: bool | None = None, strict: bool = False,
                 compaction: bool = True, chunk: int = 8, max_chunks: int = 0, unroll: int = 0,
                 waves_per_eu: int = 0, fast_div: bool = False, wide: bool | None = None,
                 tile: int = 136, reassoc: bool = False, sched_group: int = 0,
                 factored: bool = False, vpl: int = 1, mfma: bool = False, exp_rec: int = 0,
                 uc_lds: bool = False, quad: int = 0, quad_waves: int = 0, **kwargs):
        if (args and isinstance(args[0], str)) or "expr" in kwargs:
            # notebook form: (expr, pars_str, consts_str, ndata, dtype, write_to_file)
            a = _bind(args, kwargs, ("expr", "pars_str", "consts_str", "ndata", "dtype",
                                     "write_to_file"), {"dtype": torch.float32,
                                                        "write_to_file": False})
            model_expr, f_gradf = Expr(a["expr"], a["pars_str"], a["consts_str"]), None
        else:
            # committed form: (model_expr, f_gradf, ndata, dtype, write_to_file)
            a = _bind(args, kwargs, ("model_expr", "f_gradf", "ndata", "dtype", "write_to_file"),
                      {"model_expr": None, "f_gradf": None, "dtype": torch.float32,
                       "write_to_file": False})
            model_expr, f_gradf = a["model_expr"], a["f_gradf"]
        ndata, dtype, write_to_file = a["ndata"], a["dtype"], a["write_to_file"]
        if (model_expr is None) == (f_gradf is None):
            raise RuntimeError("exactly one of model_expr and f_gradf must be given")
        if ndata is None or int(ndata) <= 0:
            raise ValueError("ndata must be a positive integer")
        self.expr_true = model_expr is not None
        if not self.expr_true:
            from .plugin import PluginLM
            self._plugin = PluginLM(f_gradf, int(ndata), lm.dtype_name(dtype), block)
        else:
            self._plugin = None
        self.model_expr = copy.deepcopy(model_expr)
        self.f_gradf = f_gradf
        self.nparam = len(model_expr.pars_str) if self.expr_true else f_gradf.nparam
        self.nhes = self.nparam * (self.nparam + 1) // 2
        self.nconst = len(model_expr.consts_str) if self.expr_true else f_gradf.nconst
        self.ndata = int(ndata)
        self.dtype_name = lm.dtype_name(dtype)
        self.dtype = lm.torch_dtype(self.dtype_name)
        self.type_str = "float" if self.dtype_name == "float32" else "double"
        self.write_to_file = write_to_file
        self.expr = model_expr.expr if self.expr_true else None
        self.pars_str = list(model_expr.pars_str) if self.expr_true else None
        self.consts_str = list(model_expr.consts_str) if self.expr_true else None
        self.block = int(block)
        self.fast_exp = bool(fast_exp)
        # strict: reproducible transcendentals and no FMA contraction, so the
        # CPU oracle replays the fit bit for bit
        self.strict = bool(strict)
        # compaction: run the iterations in chunks of `chunk`, re-packing the
        # unconverged voxels between chunks (identical results, fuller waves)
        self.compaction = bool(compaction)
        self.chunk = int(chunk)
        self.max_chunks = int(max_chunks)
        self._ws = None
        self._uc_t = None
        self._chunk_state = None
        self._knobs = dict(unroll=int(unroll), waves_per_eu=int(waves_per_eu),
                           fast_div=bool(fast_div), reassoc=bool(reassoc),
                           sched_group=int(sched_group), factored=bool(factored),
                           vpl=int(vpl), exp_rec=int(exp_rec), uc_lds=bool(uc_lds))
        self.keep_workspace = (self.nparam <= 8) if keep_workspace is None else bool(keep_workspace)
        # wide: the kernels for models with more parameters than one lane's
        # registers hold (n > 8): csrc/lm_glm.hip for generalised linear
        # models with voxel-uniform consts, csrc/lm_wide.hip (Hessian kept in
        # LDS) for any other model. None = use them whenever they apply.
        # mfma: the GLM kernel with the Hessian on the matrix cores (float32);
        # quad: the GLM kernel with 4 lanes per voxel (quad=4 or True; float32,
        # n <= 24), quad_waves its occupancy hint (wavefronts per SIMD)
        self.wide, self.tile, self.mfma = wide, int(tile), bool(mfma)
        self.quad = 4 if quad is True else int(quad)
        self.quad_waves = int(quad_waves)
        self._glm, self._glm_tab, self._wide_prog = None, None, None
        # more consts than the kernel-argument block holds: per-voxel layout
        self._uc_fits = self.nconst * self.ndata <= _lib.PCA_UC_MAX
        if self.expr_true:
            self._spec_uniform = self._spec(False)
            self.funcid = self._spec_uniform.funcid()
            if wide is not False:
                from . import wide as _wide
                self._glm = _wide.glm_program(self._spec_uniform, self.tile, self.mfma,
                                              self.quad, self.quad_waves)
                if self.mfma and self._glm is None:
                    raise ValueError("mfma needs a float32 generalised linear model with n > 8 "
                                     "whose tables fit in the LDS")
                if self.quad and self._glm is None:
                    raise ValueError("quad needs a float32 generalised linear model with "
                                     "9 to 24 parameters")
                if wide and self._glm is None and _wide.wide_program(
                        self._spec(not self._uc_fits), self.tile) is None:
                    raise ValueError("the wide kernels need n > 8 and a packed Hessian that "
                                     "fits in the 160 KB of LDS of a workgroup")
        else:
            self.funcid = self._plugin.funcid
        self.Nprobs = 1
        self.run_func = None
        self.program = None
        self.build()

    def _spec(self, consts_per_voxel: bool) -> lm.LMSpec:
        return lm.LMSpec(self.expr, tuple(self.pars_str), tuple(self.consts_str), self.ndata,
                         self.dtype_name, self.block, consts_per_voxel, self.fast_exp,
                         self.strict, **self._knobs)

    # -- CudaFunction protocol ------------------------------------------------
    def get_device_funcid(self):
        return self.funcid

    def get_kernel_funcid(self):
        return "k_" + self.funcid

    def get_device_code(self):
        return self.code

    def get_kernel_code(self):
        return ""

    def get_deps(self):
        return []

    def build(self, consts_per_voxel: bool | None = None):
        """Generate and compile the module (eager, as lsqnonlin.py:973). By
        default the uniform-consts module, or the per-voxel one when nconst*M
        exceeds the kernel-argument block."""
        if consts_per_voxel is None:
            consts_per_voxel = not self._uc_fits
        if not self.expr_true:
            self.program = self._plugin.program(consts_per_voxel)
            self.code = self.program.source
        else:
            spec = self._spec(consts_per_voxel)
            try:
                self.program = lm.get_program(spec)
            except _lib.PcaError:
                with open("on_compile_fail.hip", "w") as fh:
                    fh.write(lm.build_source(spec, hipgen.res_jac_hes(
                        spec.expr, spec.pars, spec.consts)))
                raise
            self.code = self.program.source
        if self.write_to_file:
            with open(self.funcid + ".hip", "w") as fh:
                fh.write(self.code)
        self.run_func = self.program
        return self.code

    # -- fit --------------------------------------------------------------------
    def setup(self, pars_t, consts_t, data_t, lower_bound_t, upper_bound_t,
              uniform_consts: bool | None = None):
        """Keep references to the caller's tensors (pars_t is updated in place)
        and allocate the state: lam=5, step_type=1 (lsqnonlin.py:1036-1056)."""
        for t, nm in ((pars_t, "pars_t"), (data_t, "data_t"), (lower_bound_t, "lower_bound_t"),
                      (upper_bound_t, "upper_bound_t"), (consts_t, "consts_t")):
            _check_device_tensor(t, nm, self.dtype)
        d = data_t[0] if data_t.dim() == 3 and data_t.shape[0] == 1 else data_t
        if d.dim() != 2 or d.shape[0] != self.ndata:
            raise ValueError(f"data_t has shape {tuple(data_t.shape)}, expected ({self.ndata}, N)")
        N = int(d.shape[1])
        self.Nprobs = N
        self.data_t = _as_rows(d, self.ndata, N, "data_t")
        self.pars_t = _as_rows(pars_t, self.nparam, N, "pars_t")
        self.lower_bound_t = _as_rows(lower_bound_t, self.nparam, N, "lower_bound_t")
        self.upper_bound_t = _as_rows(upper_bound_t, self.nparam, N, "upper_bound_t")
        self.consts_t = consts_t
        if not self.expr_true:
            # plugin code indexes consts itself: passed through as given
            self._uc, self._cpv = None, consts_t.contiguous()
            self.consts_t = self._cpv
        else:
            self._uc, self._cpv = split_consts(consts_t, self.nconst, self.ndata, N,
                                               uniform_consts)
            if self._uc is not None and not self._uc_fits:
                # uniform values beyond the argument block: replicate them
                self._cpv = torch.tensor(self._uc, dtype=self.dtype,
                                         device=data_t.device)[:, None].expand(
                                             -1, N).contiguous()
                self._uc = None
        if (self._cpv is not None) != self.program_is_per_voxel():
            self.build(consts_per_voxel=self._cpv is not None)
        if self._uc is not None:
            # the given consts and the derived rows the generated code reads
            self._ucx = self.program.uconsts(self._uc)
            self._ucbuf, self._ucbytes = lm.uconsts_buffer(self._ucx, self.dtype_name)
        else:
            self._ucx = None
            self._ucbuf, self._ucbytes = None, 0
        self._glm_tab = None
        if self._glm is not None and self._uc is not None:
            self._glm_tab = self._glm.table(self._uc, self.data_t.device)
        self._wide_prog = None
        if (self.expr_true and self._glm_tab is None and self.wide is not False
                and self.nparam > 8):
            from . import wide as _wide
            self._wide_prog = _wide.wide_program(self._spec(self._cpv is not None), self.tile)
        dev = self.data_t.device
        kw = dict(dtype=self.dtype, device=dev)
        self.lam_t = torch.full((1, N), 5.0, **kw)
        self.step_type_t = torch.ones((1, N), dtype=torch.int8, device=dev)
        self.f_t = torch.zeros((1, N), **kw)
        self.first_f = torch.zeros((1, N), **kw)
        self.last_f = torch.zeros((1, N), **kw)
        self.n_iter_t = torch.zeros((1, N), dtype=torch.int32, device=dev)
        if self.keep_workspace:
            self.step_t = torch.zeros((self.nparam, N), **kw)
            self.g_t = torch.zeros((self.nparam, N), **kw)
            self.h_t = torch.zeros((self.nhes, N), **kw)
            self.hl_t = torch.zeros((self.nhes, N), **kw)
        else:
            self.step_t = self.g_t = self.h_t = self.hl_t = None
        self.ftp_t = None
        self._ws = None
        self._uc_t = None
        if self._chunk_state is not None:
            self._chunk_state[1].destroy_graph()
        self._chunk_state = None
        self.pars_tp_t = None

    def program_is_per_voxel(self) -> bool:
        if not self.expr_true:
            return self._plugin.per_voxel
        return self.program is not None and self.program.spec.consts_per_voxel

    def lm_args(self, iters, tol, mu, eta, acc, dec) -> _lib.LMArgs:
        P = lm.ptr
        return _lib.LMArgs(
            data=P(self.data_t), consts=P(self._cpv), lb=P(self.lower_bound_t),
            ub=P(self.upper_bound_t), params=P(self.pars_t), lam=P(self.lam_t),
            step_type=P(self.step_type_t), f=P(self.f_t), first_f=P(self.first_f),
            last_f=P(self.last_f), n_iter=P(self.n_iter_t), g_out=P(self.g_t),
            h_out=P(self.h_t), hl_out=P(self.hl_t), step_out=P(self.step_t),
            N=self.Nprobs, iters=int(iters), reserved=0, tol=float(tol), mu=float(mu),
            eta=float(eta), acc=float(acc), dec=float(dec))

    def run(self, iters: int, tol=1e-8, mu=0.25, eta=0.75, acc=0.2, dec=2.0):
        """Run ``iters`` LM iterations on every voxel (one launch on the current
        stream; asynchronous like the reference). Voxels with step_type 0 are
        converged and skipped."""
        if not hasattr(self, "data_t"):
            raise RuntimeError("setup() must be called before run()")
        if iters <= 0 or self.Nprobs == 0:       # nothing to fit
            return None
        if not self.expr_true:
            return self._plugin.run(self, iters, tol, mu, eta, acc, dec)
        if self._glm_tab is not None:
            self._glm.launch(self.lm_args(iters, tol, mu, eta, acc, dec), self._glm_tab)
            return None
        if self._wide_prog is not None:
            a = self.lm_args(iters, tol, mu, eta, acc, dec)
            a.h_out = a.hl_out = None           # the wide kernels write g and step only
            self._wide_prog.launch(a, self._ucbuf, self._ucbytes)
            return None
        if self._use_chunked():
            r = self._chunked_run(tol, mu, eta, acc, dec)
            r.set_iters(iters)
            r.launch(chunked.chunk_schedule(int(iters), self.chunk, self.max_chunks))
            return None
        a = self.lm_args(iters, tol, mu, eta, acc, dec)
        k = self.program.kernel("pca_lm_fit")
        _lib.check(_lib.lib().pca_lm_launch(k, ctypes.byref(a), self._ucbuf, self._ucbytes,
                                            self.block, lm.stream_handle()), "pca_lm_launch")
        return None

    # -- chunked execution with convergence compaction ------------------------
    def _use_chunked(self) -> bool:
        if not self.compaction:
            return False
        # per-voxel consts ride along in the queues only when they fit in registers
        return self._cpv is None or self.nconst * self.ndata <= lm.CV_REGS

    def _chunked_run(self, tol, mu, eta, acc, dec):
        key = (float(tol), float(mu), float(eta), float(acc), float(dec))
        if self._chunk_state is not None and self._chunk_state[0] == key:
            return self._chunk_state[1]
        dev = self.data_t.device
        if self._ws is None:
            ncm = self.nconst * self.ndata if self._cpv is not None else 0
            self._ws = chunked.Workset(self.Nprobs, self.nparam, self.dtype, dev, M=self.ndata,
                                       ncm_per_voxel=ncm)
            if self._ucx is not None:
                self._uc_t = torch.tensor(self._ucx, dtype=self.dtype, device=dev)
        a = self.lm_args(0, tol, mu, eta, acc, dec)
        r = chunked.ChunkedRun(self.program, [(self.program, a, self._uc_t, self._ws)], dev,
                               self.block)
        self._chunk_state = (key, r)
        return r

    def reset_state(self):
        """Restore the setup() state on the device (lambda 5, step_type 1, f 0,
        n_iter 0) with one kernel on the current stream."""
        launch_reset_state(self.program, self.lm_args(0, 1e-8, 0.25, 0.75, 0.2, 2.0),
                           self.Nprobs, lm.stream_handle())

    def capture_graph(self, iters: int, tol=1e-8, mu=0.25, eta=0.75, acc=0.2, dec=2.0,
                      init=None):
        """Capture a whole fit -- optionally init(stream) (a kernel writing the
        initial guess into pars_t), then the chunked run starting from the
        setup() state (the first chunk takes lambda 5, step_type 1, f 0 and
        n_iter 0 itself, as reset_state() would write them) -- into a hipGraph;
        replay it with replay_graph()."""
        if not self._use_chunked():
            raise RuntimeError("graph capture needs the chunked path (compaction=True)")
        r = self._chunked_run(tol, mu, eta, acc, dec)
        r.set_iters(iters)
        r.capture(chunked.chunk_schedule(int(iters), self.chunk, self.max_chunks), init,
                  fresh=True)
        return r

    def replay_graph(self):
        if self._chunk_state is None or self._chunk_state[1]._graph is None:
            raise RuntimeError("capture_graph() first")
        self._chunk_state[1].replay()

    # -- extras -----------------------------------------------------------------
    def final_f(self):
        """SSR at the current (final, clamped) parameters, as a new (1, N) tensor."""
        out = torch.zeros((1, self.Nprobs), dtype=self.dtype, device=self.data_t.device)
        F.launch(self.program, self.pars_t, self.data_t, self._cpv, self._ucbuf, self._ucbytes,
                 None, out, self.block)
        return out


def launch_reset_state(program, args, N, stream):
    """pca_lm_reset_state over N voxels: lambda 5, step_type 1, f 0, n_iter 0
    (the state setup() allocates, lsqnonlin.py:1048-1056)."""
    grid = max(1, min((int(N) + 255) // 256, 4096))
    params = (ctypes.c_void_p * 1)(ctypes.cast(ctypes.pointer(args), ctypes.c_void_p))
    _lib.check(_lib.lib().pca_launch(program.kernel("pca_lm_reset_state"), grid, 1, 1, 256, 1,
                                     1, 0, stream, params), "pca_lm_reset_state")


class F(HipFunction):
    """Standalone SSR kernel (lsqnonlin.py:120-199): run(pars, consts, data,
    step_type, f) accumulates sum_i r_i^2 onto f for voxels with step_type != 0."""

    def __init__(self, *args, block: int = lm.DEFAULT_BLOCK, **kwargs):
        if (args and isinstance(args[0], str)) or "expr" in kwargs:
            # notebook form: F(expr, pars_str, consts_str, ndata, dtype) (test_ivim.ipynb:574)
            a = _bind(args, kwargs, ("expr", "pars_str", "consts_str", "ndata", "dtype",
                                     "write_to_file"), {"dtype": torch.float32,
                                                        "write_to_file": False})
            model_expr = Expr(a["expr"], a["pars_str"], a["consts_str"])
        else:
            # committed form: F(model_expr, ndata, dtype, write_to_file) (lsqnonlin.py:120)
            a = _bind(args, kwargs, ("model_expr", "ndata", "dtype", "write_to_file"),
                      {"dtype": torch.float32, "write_to_file": False})
            model_expr = copy.deepcopy(a["model_expr"])
        ndata, dtype, write_to_file = a["ndata"], a["dtype"], a["write_to_file"]
        self.model_expr = model_expr
        self.ndata = int(ndata)
        self.dtype_name = lm.dtype_name(dtype)
        self.dtype = lm.torch_dtype(self.dtype_name)
        self.block = block
        self.write_to_file = write_to_file
        self._programs = {}
        self.funcid = lm.LMSpec(model_expr.expr, tuple(model_expr.pars_str),
                                tuple(model_expr.consts_str), self.ndata,
                                self.dtype_name).funcid("f")

    def get_device_funcid(self):
        return self.funcid

    def _program(self, per_voxel):
        p = self._programs.get(per_voxel)
        if p is None:
            e = self.model_expr
            p = lm.get_program(lm.LMSpec(e.expr, tuple(e.pars_str), tuple(e.consts_str),
                                         self.ndata, self.dtype_name, self.block, per_voxel))
            self._programs[per_voxel] = p
            if self.write_to_file:
                with open(self.funcid + ".hip", "w") as fh:
                    fh.write(p.source)
        return p

    @staticmethod
    def launch(program, pars, data, cpv, ucbuf, ucbytes, step_type, f, block):
        N = int(data.shape[-1])
        a = _lib.SSRArgs(params=lm.ptr(pars), data=lm.ptr(data), consts=lm.ptr(cpv),
                         step_type=lm.ptr(step_type), f=lm.ptr(f), N=N)
        k = program.kernel("pca_ssr")
        _lib.check(_lib.lib().pca_ssr_launch(k, ctypes.byref(a), ucbuf, ucbytes, block,
                                             lm.stream_handle()), "pca_ssr_launch")

    def run(self, pars, consts, data, step_type, f):
        e = self.model_expr
        for t, nm in ((pars, "pars"), (data, "data"), (f, "f")):
            _check_device_tensor(t, nm, self.dtype)
        d = data[0] if data.dim() == 3 else data
        N = int(d.shape[-1])
        pars = _as_rows(pars, len(e.pars_str), N, "pars")
        d = _as_rows(d, self.ndata, N, "data")
        if f.numel() != N or not f.is_contiguous():
            raise ValueError("f must be a contiguous array of N values")
        uc, cpv = split_consts(consts, len(e.consts_str), self.ndata, N)
        prog = self._program(cpv is not None)
        ucbuf, ucbytes = lm.uconsts_buffer(uc, self.dtype_name) if uc is not None else (None, 0)
        st = None
        if step_type is not None:
            if step_type.dtype != torch.int8 or step_type.numel() != N:
                raise ValueError("step_type must be an int8 array of N values (or None)")
            st = step_type
        self.launch(prog, pars, d, cpv, ucbuf, ucbytes, st, f, self.block)


def model_eval(expr: Expr, pars, consts, M: int, dtype=torch.float32, want=("m", "J", "d2"),
               block: int = lm.DEFAULT_BLOCK):
    """Model value (M, N), Jacobian (n, M, N) and packed second derivatives
    (nh, M, N) at pars (n, N), computed on the GPU by the generated code."""
    dn = lm.dtype_name(dtype)
    td = lm.torch_dtype(dn)
    _check_device_tensor(pars, "pars", td)
    n, N = pars.shape
    uc, cpv = split_consts(consts, len(expr.consts_str), M, N)
    prog = lm.get_program(lm.LMSpec(expr.expr, tuple(expr.pars_str), tuple(expr.consts_str), M,
                                    dn, block, cpv is not None))
    nh = n * (n + 1) // 2
    kw = dict(dtype=td, device=pars.device)
    m = torch.empty((M, N), **kw) if "m" in want else None
    J = torch.empty((n, M, N), **kw) if "J" in want else None
    d2 = torch.empty((nh, M, N), **kw) if "d2" in want else None
    ucbuf, ucbytes = lm.uconsts_buffer(uc, dn) if uc is not None else (None, 0)
    a = _lib.EvalArgs(params=lm.ptr(pars.contiguous()), consts=lm.ptr(cpv), m=lm.ptr(m),
                      J=lm.ptr(J), d2=lm.ptr(d2), N=N)
    _lib.check(_lib.lib().pca_eval_launch(prog.kernel("pca_eval"), ctypes.byref(a), ucbuf,
                                          ucbytes, block, lm.stream_handle()), "pca_eval_launch")
    return m, J, d2
