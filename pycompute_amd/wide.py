"""Programs of the wide-model kernels for models with more parameters than one
lane's registers hold (n > 8): csrc/lm_glm.hip for generalised linear models
(voxel-uniform consts), csrc/lm_wide.hip for any other generated model (the
Hessian tiles accumulated in registers and kept in LDS)."""
from __future__ import annotations

import ctypes
import threading

import numpy as np
import torch

from . import _lib, lm
from .program import BASE_OPTIONS, Module, read_csrc
from .sym import glm, hipgen

WIDE_MIN_PARAMS = 9
DEFAULT_TILE = 136
LDS_LIMIT = 160 * 1024   # bytes of LDS one gfx950 workgroup may allocate


def tile_rows(n: int, tile_max: int) -> list:
    """First row of each tile (and n at the end): runs of whole rows of the
    packed lower triangle with at most tile_max entries, at least one row
    (the host restatement of tile_row in csrc/lm_glm.hip)."""
    rows, r = [0], 0
    while r < n:
        e = 0
        while True:
            e += r + 1
            r += 1
            if not (r < n and e + r + 1 <= tile_max):
                break
        rows.append(r)
    return rows


def park_entries(n: int, tile_max: int) -> int:
    """Hessian entries parked in LDS (all tiles but the last)."""
    r = tile_rows(n, tile_max)[-2]
    return r * (r + 1) // 2


def lds_bytes(n: int, tile_max: int, dtype: str) -> int:
    """LDS of one 64-lane workgroup: one column of parked entries per lane."""
    return max(park_entries(n, tile_max), 1) * 64 * (8 if dtype == "float64" else 4)


def mfma_lds_bytes(n: int, M: int) -> int:
    """LDS of one 256-thread workgroup of pca_lm_glm_mfma: the entry-product
    table (M rows of 4 x (groups padded to 4, plus 4) floats), the samples of
    every voxel (4 wavefronts x M x 64 floats) and g parked over the solve
    (4 x 64 x (n + 1) floats)."""
    nq = (n * (n + 1) // 2 + 3) // 4
    rs = (nq + 3) // 4 * 4 + 4
    return M * 4 * rs * 4 + 4 * M * 64 * 4 + 4 * 64 * (n + 1) * 4


QUAD_MAX_PARAMS = 24   # pca_lm_glm_quad: rows of the pivoted system per voxel


class GLMProgram:
    """pca_lm_glm (Hessian tiles accumulated on the VALU, parked in LDS), or
    with mfma=True pca_lm_glm_mfma (Hessian accumulated on the matrix cores,
    one voxel per lane; float only), or with quad = 4 (True)
    pca_lm_glm_quad (four lanes per voxel, the rows of the pivoted system
    spread over them; float only, at most 24 parameters; quad_waves is its
    occupancy hint)."""

    def __init__(self, spec: lm.LMSpec, tile_max: int = DEFAULT_TILE, mfma: bool = False,
                 quad: int = 0, quad_waves: int = 0):
        self.spec = spec
        self.mfma = bool(mfma)
        self.quad = quad_lanes(quad)
        # occupancy hint of the quad kernel (wavefronts per SIMD; 0 = none)
        self.quad_waves = int(quad_waves) if self.quad else 0
        if self.mfma and self.quad:
            raise ValueError("mfma and quad are different GLM kernels; pick one")
        self.dag = hipgen.res_jac_hes(spec.expr, spec.pars, spec.consts)
        self.form = glm.glm_form(self.dag)
        if self.form is None:
            raise ValueError("the wide kernel needs a model phi(eta) with eta affine in the "
                             "parameters")
        self.tile_max = int(tile_max)
        n = spec.nparam
        self.lds_bytes = lds_bytes(n, self.tile_max, spec.dtype)
        if self.lds_bytes > LDS_LIMIT:
            raise ValueError(f"tile {self.tile_max} parks {self.lds_bytes} bytes of LDS per "
                             f"workgroup (limit {LDS_LIMIT}); use a larger tile")
        if self.mfma:
            if spec.dtype != "float32":
                raise ValueError("the matrix-core GLM kernel is float32 only")
            if mfma_lds_bytes(n, spec.M) > LDS_LIMIT:
                raise ValueError(f"the matrix-core GLM kernel needs {mfma_lds_bytes(n, spec.M)} "
                                 f"bytes of LDS for M = {spec.M} (limit {LDS_LIMIT})")
        if self.quad:
            if spec.dtype != "float32":
                raise ValueError("the quad GLM kernel is float32 only")
            if n > QUAD_MAX_PARAMS:
                raise ValueError(f"the quad GLM kernel takes at most {QUAD_MAX_PARAMS} "
                                 f"parameters, got {n}")
        self.stride = glm.table_stride(n)
        self.funcid = spec.funcid() + f"_glm{self.tile_max}" + ("_mfma" if self.mfma else "") \
            + (f"_quad{self.quad}w{self.quad_waves}" if self.quad else "")
        head = [f"// generated: {self.funcid}", f"// model: {spec.expr}",
                f"#define PCA_NP {n}", f"#define PCA_M {spec.M}", f"#define PCA_TS {self.stride}",
                f"#define PCA_TILE_MAX {self.tile_max}", f"#define PCA_BLOCK {spec.block}"]
        if spec.dtype == "float64":
            head.append("#define PCA_REAL_DOUBLE 1")
        if spec.fast_exp:
            head.append("#define PCA_FAST_EXP 1")
        if self.form.phi == "identity":
            head.append("#define PCA_GLM_IDENTITY 1")
        if spec.reassoc:
            head.append("#define PCA_FAST_SOLVE 1")
        if self.quad:
            head.append(f"#define PCA_LPV {self.quad}")
            if self.quad_waves:
                head.append(f"#define PCA_QUAD_WAVES {self.quad_waves}")
            kernels = [read_csrc("lm_glm_quad.hip")]
        else:
            kernels = [read_csrc("lm_glm.hip")] + ([read_csrc("lm_glm_mfma.hip")]
                                                   if self.mfma else [])
        self.source = "\n".join(["\n".join(head), read_csrc("strict_math.h"),
                                 read_csrc("prelude.hip"), read_csrc("lm_solver.hip")] + kernels)
        options = list(BASE_OPTIONS)
        if spec.fast_div:
            options = [o for o in options if o != "-fhip-fp32-correctly-rounded-divide-sqrt"]
            options.append("-fno-hip-fp32-correctly-rounded-divide-sqrt")
        self.module = Module(self.source, self.funcid, options)
        fl = hipgen.flops_per_iteration(self.dag, spec.M)
        self.flops = glm_flops(n, spec.M, fl)

    def table(self, uc, device):
        """Coefficient table for voxel-uniform consts (k*M+i ordering); with
        mfma, the pair (table, entry-product table)."""
        c = np.asarray(uc, np.float64).reshape(self.spec.nconst, self.spec.M)
        tab = glm.coefficient_table(self.form, self.dag, c)
        if self.quad:
            return torch.tensor(quad_table(tab, self.spec.nparam, self.quad),
                                dtype=torch.float32, device=device)
        t = torch.tensor(tab, dtype=lm.torch_dtype(self.spec.dtype), device=device)
        if not self.mfma:
            return t
        return t, torch.tensor(entry_products(tab, self.spec.nparam), dtype=torch.float32,
                               device=device)

    def launch(self, args: _lib.LMArgs, tab, stream=None):
        s = lm.stream_handle() if stream is None else stream
        if self.quad:
            L = self.quad
            if (tab.shape != (self.spec.M, L, quad_slots(self.spec.nparam, L))
                    or tab.dtype != torch.float32 or not tab.is_contiguous()):
                raise ValueError("the quad kernel needs the float32 table of "
                                 "GLMProgram.table()")
            t = ctypes.c_void_p(tab.data_ptr())
            params = (ctypes.c_void_p * 2)(ctypes.cast(ctypes.pointer(args), ctypes.c_void_p),
                                           ctypes.cast(ctypes.pointer(t), ctypes.c_void_p))
            vpb = 256 // L                        # voxels of L lanes per workgroup
            grid = (int(args.N) + vpb - 1) // vpb
            _lib.check(_lib.lib().pca_launch(self.module.kernel("pca_lm_glm_quad"), grid, 1, 1,
                                             256, 1, 1, 0, s, params), "pca_lm_glm_quad")
            return
        if self.mfma:
            t, t2 = (ctypes.c_void_p(x.data_ptr()) for x in tab)
            params = (ctypes.c_void_p * 3)(ctypes.cast(ctypes.pointer(args), ctypes.c_void_p),
                                           ctypes.cast(ctypes.pointer(t), ctypes.c_void_p),
                                           ctypes.cast(ctypes.pointer(t2), ctypes.c_void_p))
            grid = (int(args.N) + 255) // 256
            _lib.check(_lib.lib().pca_launch(self.module.kernel("pca_lm_glm_mfma"), grid, 1, 1,
                                             256, 1, 1, 0, s, params), "pca_lm_glm_mfma")
            return
        t = ctypes.c_void_p(tab.data_ptr())
        params = (ctypes.c_void_p * 2)(ctypes.cast(ctypes.pointer(args), ctypes.c_void_p),
                                       ctypes.cast(ctypes.pointer(t), ctypes.c_void_p))
        grid = (int(args.N) + 63) // 64
        _lib.check(_lib.lib().pca_launch(self.module.kernel("pca_lm_glm"), grid, 1, 1, 64, 1, 1,
                                         0, s, params), "pca_lm_glm")


def quad_lanes(quad) -> int:
    """Lanes per voxel of the quad GLM kernel: 0 (off) or 4 (True). The
    kernel also compiles for 2 lanes per voxel (PCA_LPV 2), but with 11 row
    groups per lane the compiler keeps the solve's loops rolled over arrays in
    scratch, and the 22-parameter fit ran 12x slower; the host does not offer
    it."""
    lanes = 4 if quad is True else int(quad)
    if lanes not in (0, 4):
        raise ValueError(f"quad must be 0 or 4 lanes per voxel, got {quad}")
    return lanes


def quad_perm(n: int, a: int) -> int:
    """Parameter of row a of the pivoted system: the reference's diagonal
    pivot is the fixed rotation [n-1, 0, 1, ..., n-2] (lm_solver.hip perm())."""
    return n - 1 if a == 0 else a - 1


def quad_slots(n: int, L: int) -> int:
    """Floats per (measurement, lane) of pca_lm_glm_quad's table: the lane's
    NT = ceil(n / L) rows, a0, padded to a multiple of 4."""
    return (-(-n // L) + 1 + 3) // 4 * 4


def quad_table(tab: np.ndarray, n: int, L: int = 4) -> np.ndarray:
    """(M, L, quad_slots) float64 table of pca_lm_glm_quad with L lanes per
    voxel from the coefficient table: [i, r, t] = A_{perm(L t + r)} of
    measurement i for L t + r < n (else 0), [i, r, NT] = a0 (NT = ceil(n /
    L)), zero padding."""
    if n > QUAD_MAX_PARAMS:
        raise ValueError(f"at most {QUAD_MAX_PARAMS} parameters")
    if L not in (2, 4):
        raise ValueError("2 or 4 lanes per voxel")
    M, nt = tab.shape[0], -(-n // L)
    out = np.zeros((M, L, quad_slots(n, L)))
    for r in range(L):
        for t in range(nt):
            a = L * t + r
            if a < n:
                out[:, r, t] = tab[:, quad_perm(n, a)]
        out[:, r, nt] = tab[:, 2 * n]
    return out


def entry_products(tab: np.ndarray, n: int) -> np.ndarray:
    """(M, 4, RS) float64 table of pca_lm_glm_mfma: t2[i, r, q] = A_ij A_ik
    of packed entry e = 4q + r = j(j+1)/2 + k at measurement i (0 past the
    last entry), rows padded to RS = (groups rounded up to 4) + 4."""
    M = tab.shape[0]
    nh = n * (n + 1) // 2
    nq = (nh + 3) // 4
    rs = (nq + 3) // 4 * 4 + 4
    prod = np.zeros((M, 4 * nq))
    e = 0
    for j in range(n):
        for k in range(j + 1):
            prod[:, e] = tab[:, j] * tab[:, k]
            e += 1
    out = np.zeros((M, 4, rs))
    out[:, :, :nq] = prod.reshape(M, nq, 4).transpose(0, 2, 1)
    return out


def glm_flops(n: int, M: int, generic: dict) -> dict:
    """Algorithmic FLOPs of one iteration of the GLM formulation (the
    parameter-dependent work; the consts-only table is evaluated once):
    per measurement eta (2n), r, w, f (6), g and damping (2n + 3n + 3),
    w A_j (n) and the packed update (2 n(n+1)/2); the trial pass (2n + 3 per
    measurement); the solve and step as in the generic count."""
    nh = n * (n + 1) // 2
    per_meas = 2 * n + 6 + 5 * n + 3 + n + 2 * nh
    trial = 2 * n + 3
    f_it = M * (per_meas + trial) + generic["F_solve"] + generic["F_step"]
    f_conv = M * per_meas + generic["F_solve"] + 4 * n + 2
    return {"F_it": f_it, "F_conv": f_conv, "exp_it": 2 * M, "exp_conv": M,
            "F_solve": generic["F_solve"], "F_step": generic["F_step"]}


class WideProgram:
    """The general wide kernel (pca_lm_wide) for a generated model with n > 8
    parameters: the model struct of the generated module, the packed Hessian
    in LDS (n(n+1)/2 columns of 64 lanes), tiles of at most tile_max entries
    accumulated in registers."""

    def __init__(self, spec: lm.LMSpec, tile_max: int = DEFAULT_TILE):
        self.spec = spec
        self.dag = hipgen.res_jac_hes(spec.expr, spec.pars, spec.consts)
        self.tile_max = int(tile_max)
        n = spec.nparam
        self.lds_bytes = wide_lds_bytes(n, spec.dtype)
        if self.lds_bytes > LDS_LIMIT:
            raise ValueError(f"{n} parameters keep {self.lds_bytes} bytes of Hessian in LDS per "
                             f"workgroup (limit {LDS_LIMIT})")
        self.funcid = spec.funcid() + f"_wide{self.tile_max}" + ("_strict" if spec.strict else "")
        head = [f"// generated: {self.funcid}", f"// model: {spec.expr}",
                f"#define PCA_M {spec.M}", f"#define PCA_UC_MAX {_lib.PCA_UC_MAX}",
                f"#define PCA_TILE_MAX {self.tile_max}"]
        if spec.dtype == "float64":
            head.append("#define PCA_REAL_DOUBLE 1")
        if spec.consts_per_voxel:
            head.append("#define PCA_CONSTS_PER_VOXEL 1")
        if spec.fast_exp:
            head.append("#define PCA_FAST_EXP 1")
        if spec.strict:
            head.append("#define PCA_STRICT 1")
        if spec.reassoc:
            head.append("#define PCA_FAST_SOLVE 1")
        model = hipgen.emit_model(self.dag, "Model", spec.dtype, reassoc=spec.reassoc)
        self.source = "\n".join(["\n".join(head), read_csrc("strict_math.h"),
                                 read_csrc("prelude.hip"), model, read_csrc("lm_solver.hip"),
                                 read_csrc("lm_wide.hip")])
        options = list(BASE_OPTIONS)
        if spec.strict:
            options.append("-ffp-contract=off")
        if spec.fast_div:
            options = [o for o in options if o != "-fhip-fp32-correctly-rounded-divide-sqrt"]
            options.append("-fno-hip-fp32-correctly-rounded-divide-sqrt")
        self.module = Module(self.source, self.funcid, options)
        self.flops = hipgen.flops_per_iteration(self.dag, spec.M)

    def launch(self, args: _lib.LMArgs, ucbuf, ucbytes, stream=None):
        s = lm.stream_handle() if stream is None else stream
        _lib.check(_lib.lib().pca_lm_launch(self.module.kernel("pca_lm_wide"), ctypes.byref(args),
                                            ucbuf, ucbytes, 64, s), "pca_lm_launch (wide)")


def wide_lds_bytes(n: int, dtype: str) -> int:
    """LDS of the general wide kernel: the packed Hessian, one column per lane."""
    return n * (n + 1) // 2 * 64 * (8 if dtype == "float64" else 4)


_programs: dict = {}
_lock = threading.Lock()
_key_locks: dict = {}
_wide_programs: dict = {}


def wide_program(spec: lm.LMSpec, tile_max: int = DEFAULT_TILE) -> WideProgram | None:
    """The general wide-kernel program for ``spec`` (any generated model with
    n > 8 whose Hessian fits in LDS), or None."""
    if spec.nparam < WIDE_MIN_PARAMS or wide_lds_bytes(spec.nparam, spec.dtype) > LDS_LIMIT:
        return None
    if not spec.consts_per_voxel and spec.nconst * spec.M > _lib.PCA_UC_MAX:
        return None
    key = (spec, tile_max)
    with _lock:
        if key in _wide_programs:
            return _wide_programs[key]
        klock = _key_locks.setdefault(("wide",) + key, threading.Lock())
    with klock:
        with _lock:
            if key in _wide_programs:
                return _wide_programs[key]
        prog = WideProgram(spec, tile_max)
        with _lock:
            _wide_programs[key] = prog
        return prog


def glm_program(spec: lm.LMSpec, tile_max: int = DEFAULT_TILE,
                mfma: bool = False, quad: int = 0, quad_waves: int = 0) -> GLMProgram | None:
    """The wide-kernel program for ``spec``, or None when the model is not a
    GLM, is small enough for the register kernel, or parks more LDS than a
    workgroup may hold at this tile (with mfma: is not float32, or its
    tables exceed the LDS; with quad (4 lanes per voxel): is not float32 or
    has more than 24 parameters)."""
    if (spec.nparam < WIDE_MIN_PARAMS or spec.strict
            or lds_bytes(spec.nparam, tile_max, spec.dtype) > LDS_LIMIT):
        return None
    if mfma and (spec.dtype != "float32"
                 or mfma_lds_bytes(spec.nparam, spec.M) > LDS_LIMIT):
        return None
    quad = quad_lanes(quad)
    if quad and (mfma or spec.dtype != "float32" or spec.nparam > QUAD_MAX_PARAMS
                 or spec.M * quad * quad_slots(spec.nparam, quad) * 4 > LDS_LIMIT):
        return None
    quad_waves = int(quad_waves) if quad else 0
    key = (spec, tile_max, bool(mfma), quad, quad_waves)
    with _lock:
        if key in _programs:
            return _programs[key]
        klock = _key_locks.setdefault(key, threading.Lock())
    with klock:                 # different specifications compile concurrently
        with _lock:
            if key in _programs:
                return _programs[key]
        dag = hipgen.res_jac_hes(spec.expr, spec.pars, spec.consts)
        prog = GLMProgram(spec, tile_max, mfma, quad, quad_waves) if glm.glm_form(dag) else None
        with _lock:
            _programs[key] = prog
        return prog
