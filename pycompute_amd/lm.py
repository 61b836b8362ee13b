"""Generated LM programs: one hiprtc module per (model, M, dtype, layout).

A module holds three kernels over the same generated model struct:
``pca_lm_fit`` (csrc/lm_kernel.hip), ``pca_ssr`` and ``pca_eval``
(csrc/eval_kernel.hip). Programs are cached per specification.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import torch

from . import _lib
from .program import BASE_OPTIONS, Module, read_csrc
from .sym import hipgen

DEFAULT_BLOCK = 256


def dtype_name(dtype) -> str:
    """Normalise torch/numpy/str dtypes to 'float32' / 'float64'; anything else
    is an error (cuda_program.py:99-120 accepts only f32/f64)."""
    s = str(dtype).replace("torch.", "").replace("numpy.", "").replace("<class '", "").strip("'>")
    s = {"float": "float32", "double": "float64", "f": "float32", "d": "float64",
         "f4": "float32", "f8": "float64"}.get(s, s)
    if s not in ("float32", "float64"):
        raise RuntimeError(f"only float32 and float64 are supported, got {dtype}")
    return s


def torch_dtype(name: str):
    return torch.float32 if name == "float32" else torch.float64


@dataclass(frozen=True)
class LMSpec:
    expr: str
    pars: tuple
    consts: tuple
    M: int
    dtype: str = "float32"
    block: int = DEFAULT_BLOCK
    consts_per_voxel: bool = False
    fast_exp: bool = False
    strict: bool = False
    # tuning knobs (pycompute_amd/autotune.py): measurement-loop unroll factor
    # (0 = full for n <= 8, else 1), waves per SIMD hint (0 = compiler's
    # choice), faster non-correctly-rounded FP32 divide/sqrt
    unroll: int = 0
    waves_per_eu: int = 0
    fast_div: bool = False

    @property
    def nparam(self):
        return len(self.pars)

    @property
    def nconst(self):
        return len(self.consts)

    @property
    def type_char(self):
        return "f" if self.dtype == "float32" else "d"

    def funcid(self, prefix="second_order_levenberg_marquardt") -> str:
        """Reference naming: {fn}_{n}_{c}_{M}_{f|d}_{md5(expr)[1:13]}
        (lsqnonlin.py:810-818, cuda_program.py:77-86, sym/util.py:61-62)."""
        return (f"{prefix}_{self.nparam}_{self.nconst}_{self.M}_{self.type_char}_"
                f"{hipgen.expr_hash(self.expr, 12)}")


class LMProgram:
    def __init__(self, spec: LMSpec):
        if spec.M <= 0:
            raise ValueError("ndata must be positive")
        if spec.nconst * spec.M > _lib.PCA_UC_MAX and not spec.consts_per_voxel:
            raise ValueError(f"nconst*M = {spec.nconst * spec.M} exceeds {_lib.PCA_UC_MAX}; "
                             "use per-voxel consts")
        self.spec = spec
        self.dag = hipgen.res_jac_hes(spec.expr, spec.pars, spec.consts)
        self.funcid = spec.funcid()
        self.source = build_source(spec, self.dag)
        if spec.strict and (spec.fast_exp or spec.fast_div):
            raise ValueError("the strict build excludes fast_exp and fast_div")
        options = list(BASE_OPTIONS)
        if spec.strict:
            options.append("-ffp-contract=off")
        if spec.fast_div:
            options = [o for o in options if o != "-fhip-fp32-correctly-rounded-divide-sqrt"]
            options.append("-fno-hip-fp32-correctly-rounded-divide-sqrt")
        self.module = Module(self.source, self.funcid + ("_strict" if spec.strict else ""),
                             options)
        self.flops = hipgen.flops_per_iteration(self.dag, spec.M)

    @property
    def nhes(self):
        return self.dag.nhes

    def kernel(self, name: str):
        return self.module.kernel(name)


def build_source(spec: LMSpec, dag) -> str:
    sname = "PcaModel_" + hipgen.expr_hash(spec.expr, 12)
    n = spec.nparam
    # Full unrolling of the measurement loop keeps consts in scalar registers;
    # large models keep the loop rolled to bound code size.
    unroll = spec.unroll if spec.unroll > 0 else (spec.M if n <= 8 else 1)
    head = [f"// generated: {spec.funcid()}",
            f"// model: {spec.expr}",
            f"// pars: {', '.join(spec.pars)}; consts: {', '.join(spec.consts)}; M = {spec.M}",
            f"#define PCA_M {spec.M}",
            f"#define PCA_BLOCK {spec.block}",
            f"#define PCA_UNROLL_M {unroll}",
            f"#define PCA_UC_MAX {_lib.PCA_UC_MAX}"]
    if spec.dtype == "float64":
        head.append("#define PCA_REAL_DOUBLE 1")
    if spec.consts_per_voxel:
        head.append("#define PCA_CONSTS_PER_VOXEL 1")
    if spec.fast_exp:
        head.append("#define PCA_FAST_EXP 1")
    if spec.strict:
        head.append("#define PCA_STRICT 1")
    if spec.waves_per_eu > 0:
        head.append(f"#define PCA_WAVES_PER_EU {spec.waves_per_eu}")
    return "\n".join([
        "\n".join(head),
        read_csrc("strict_math.h"),
        read_csrc("prelude.hip"),
        hipgen.emit_model(dag, sname, spec.dtype),
        f"typedef {sname} Model;",
        read_csrc("lm_solver.hip"),
        read_csrc("lm_kernel.hip"),
        read_csrc("eval_kernel.hip"),
    ])


_programs: dict = {}
_plock = threading.Lock()


def get_program(spec: LMSpec) -> LMProgram:
    with _plock:
        p = _programs.get(spec)
        if p is None:
            p = LMProgram(spec)
            _programs[spec] = p
        return p


# ---------------------------------------------------------------------------
# launch helpers
# ---------------------------------------------------------------------------

def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def uconsts_buffer(consts_flat, dtype: str):
    """ctypes block with the voxel-uniform consts (k*M+i ordering)."""
    ct = ctypes.c_float if dtype == "float32" else ctypes.c_double
    buf = (ct * _lib.PCA_UC_MAX)()
    vals = [float(x) for x in consts_flat]
    if len(vals) > _lib.PCA_UC_MAX:
        raise ValueError("too many uniform consts")
    for i, x in enumerate(vals):
        buf[i] = x
    return buf, ctypes.sizeof(ct) * _lib.PCA_UC_MAX


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream
