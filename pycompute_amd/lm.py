"""Generated LM programs: one hiprtc module per (model, M, dtype, layout).

A module holds three kernels over the same generated model struct:
``pca_lm_fit`` (csrc/lm_kernel.hip), ``pca_ssr`` and ``pca_eval``
(csrc/eval_kernel.hip). Programs are cached per specification.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .program import BASE_OPTIONS, Module, read_csrc
from .sym import hipgen

DEFAULT_BLOCK = 256
# per-voxel consts values a lane keeps in registers in the chunked kernel
# (PCA_CV_REGS in csrc/lm_kernel.hip); larger per-voxel layouts run in the
# single-launch kernel, which reads them from global memory
CV_REGS = 64


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
    # reassociated FMA form of the Hessian update and a reciprocal-based
    # solve (fast builds; results equal the reference's up to rounding)
    reassoc: bool = False
    # register fence every sched_group measurements of the unrolled loops
    # (0 = none): bounds the measurements the compiler keeps in flight, so
    # fewer registers and more waves per SIMD
    sched_group: int = 0
    # factored accumulation (with reassoc): parameter-only factors applied
    # once per iteration instead of once per measurement
    factored: bool = False
    # voxels per lane of the chunked kernel (1 or 2; 64 * vpl voxels per
    # wavefront): a lane iterates vpl voxels in lockstep
    vpl: int = 1
    # exponential recurrence (factored builds with voxel-uniform consts whose
    # exponential factor rows are arithmetic progressions): exp(k_i x)
    # directly every exp_rec measurements, one multiply in between (0 = off)
    exp_rec: int = 0
    # the chunk kernel copies the voxel-uniform consts tables to LDS and reads
    # them there (lower latency than the vector cache)
    uc_lds: bool = False

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
    """One module. ``members`` (LMSpecs with voxel-uniform consts) are further
    models compiled into the same pca_lm_chunk kernel for batched launches."""

    def __init__(self, spec: LMSpec, members: tuple = ()):
        if spec.M <= 0 or any(m.M <= 0 for m in members):
            raise ValueError("ndata must be positive")
        if spec.nconst * spec.M > _lib.PCA_UC_MAX and not spec.consts_per_voxel:
            raise ValueError(f"nconst*M = {spec.nconst * spec.M} exceeds {_lib.PCA_UC_MAX}; "
                             "use per-voxel consts")
        if len(members) > 3:
            raise ValueError("at most 4 models share one module")
        self.spec = spec
        self.members = tuple(members)
        self.dag = hipgen.res_jac_hes(spec.expr, spec.pars, spec.consts)
        self.member_dags = [hipgen.res_jac_hes(m.expr, m.pars, m.consts) for m in members]
        self.funcid = spec.funcid()
        if members:
            ids = "|".join(m.funcid() for m in (spec,) + self.members)
            self.funcid = f"batched_levenberg_marquardt_{len(ids.split('|'))}_" \
                          f"{hipgen.expr_hash(ids, 12)}"
        # per model: the derived consts rows its factored code reads
        self.derived: list = []
        self.source = build_source(spec, self.dag, list(zip(self.members, self.member_dags)),
                                   self.derived)
        if spec.strict and (spec.fast_exp or spec.fast_div or spec.reassoc):
            raise ValueError("the strict build excludes fast_exp, fast_div and reassoc")
        if spec.factored and not spec.reassoc:
            raise ValueError("the factored accumulation needs reassoc")
        if spec.vpl not in (1, 2):
            raise ValueError(f"vpl must be 1 or 2, got {spec.vpl}")
        if spec.vpl > 1 and spec.sched_group > 0:
            raise ValueError("register fences (sched_group) apply to vpl = 1 only")
        if spec.exp_rec < 0:
            raise ValueError(f"exp_rec must be >= 0, got {spec.exp_rec}")
        if spec.exp_rec > 0:
            for s_ in (spec,) + tuple(members):
                if not s_.factored or s_.consts_per_voxel or _unroll(s_) != s_.M:
                    raise ValueError("exp_rec needs the factored build with voxel-uniform "
                                     "consts and a fully unrolled measurement loop")
        options = list(BASE_OPTIONS)
        if spec.strict:
            options.append("-ffp-contract=off")
        if spec.fast_div:
            options = [o for o in options if o != "-fhip-fp32-correctly-rounded-divide-sqrt"]
            options.append("-fno-hip-fp32-correctly-rounded-divide-sqrt")
        self.module = Module(self.source, self.funcid + ("_strict" if spec.strict else ""),
                             options)
        self.flops = hipgen.flops_per_iteration(self.dag, spec.M)
        self.member_flops = [hipgen.flops_per_iteration(d, m.M)
                             for m, d in zip(self.members, self.member_dags)]

    @property
    def nhes(self):
        return self.dag.nhes

    def uconsts(self, consts_flat, model: int = 0) -> list:
        """The voxel-uniform consts table of model ``model`` as its kernels
        read it: the NC*M given values (k*M+i ordering) followed by the rows
        of its derived consts, evaluated in float64."""
        vals = [float(x) for x in consts_flat]
        ent = self.derived[model] if self.derived else []
        if not ent:
            return vals
        s = (self.spec,) + self.members
        dag = (self.dag, *self.member_dags)[model]
        M = s[model].M
        rows = np.asarray(vals, dtype=np.float64).reshape(dag.nconst, M)
        exp2 = self.spec.fast_exp and self.spec.dtype == "float32"
        return vals + [float(x) for x in
                       hipgen.derived_values(ent, dag, rows, exp2).ravel()]

    def kernel(self, name: str):
        return self.module.kernel(name)


def _unroll(spec: LMSpec) -> int:
    # Full unrolling of the measurement loop keeps consts in scalar registers;
    # large models keep the loop rolled to bound code size.
    return spec.unroll if spec.unroll > 0 else (spec.M if spec.nparam <= 8 else 1)


def derive_consts(spec: LMSpec) -> bool:
    """Whether the module reads const-only subexpressions of its factored
    accumulation from derived consts rows (voxel-uniform consts only)."""
    return spec.factored and not spec.consts_per_voxel


def build_source(spec: LMSpec, dag, members=(), derived: list | None = None) -> str:
    """HIP source of one module. ``members``: further (LMSpec, dag) pairs that
    share the module's pca_lm_chunk kernel as Model1, Model2, ... (the
    launch/code-generation knobs of ``spec`` apply to all of them). A list
    as ``derived`` receives, per model, the entries of its derived consts
    (hipgen.DerivedConsts; empty when it has none)."""
    models = [(spec, dag)] + list(members)
    head = [f"// generated: {spec.funcid()}"]
    for i, (s, _) in enumerate(models):
        head += [f"// model {i}: {s.expr}",
                 f"// pars: {', '.join(s.pars)}; consts: {', '.join(s.consts)}; M = {s.M}"]
    head += [f"#define PCA_M {spec.M}",
             f"#define PCA_BLOCK {spec.block}",
             f"#define PCA_UNROLL_M {_unroll(spec)}",
             f"#define PCA_UC_MAX {_lib.PCA_UC_MAX}"]
    if members:
        head.append(f"#define PCA_NMODELS {len(models)}")
        for i, (s, _) in enumerate(models):
            head += [f"#define PCA_M{i} {s.M}", f"#define PCA_UNROLL_M{i} {_unroll(s)}"]
    if spec.dtype == "float64":
        head.append("#define PCA_REAL_DOUBLE 1")
    if spec.consts_per_voxel:
        head.append("#define PCA_CONSTS_PER_VOXEL 1")
        if spec.nconst * spec.M > CV_REGS:
            head.append("#define PCA_NO_CHUNK 1")
    if spec.fast_exp:
        head.append("#define PCA_FAST_EXP 1")
    if spec.strict:
        head.append("#define PCA_STRICT 1")
    if spec.waves_per_eu > 0:
        head.append(f"#define PCA_WAVES_PER_EU {spec.waves_per_eu}")
    if spec.reassoc:
        head.append("#define PCA_FAST_SOLVE 1")
    if spec.factored:
        head.append("#define PCA_FACTORED 1")
    if spec.sched_group > 0:
        head.append(f"#define PCA_SCHED_GROUP {spec.sched_group}")
    if spec.vpl > 1:
        head.append(f"#define PCA_VPL {spec.vpl}")
    if spec.exp_rec > 0:
        head.append(f"#define PCA_EXP_REC {spec.exp_rec}")
    if spec.uc_lds:
        head.append("#define PCA_UC_LDS 1")
    structs, typedefs, entries = {}, [], {}
    maxM: dict = {}
    for s, _ in models:
        sname = "PcaModel_" + hipgen.expr_hash(s.expr, 12)
        maxM[sname] = max(maxM.get(sname, 0), s.M)
    for i, (s, d) in enumerate(models):
        sname = "PcaModel_" + hipgen.expr_hash(s.expr, 12)
        if sname not in structs:
            ent = [] if derive_consts(spec) else None
            txt = hipgen.emit_model(d, sname, spec.dtype, reassoc=spec.reassoc,
                                    factored=spec.factored, derived=ent,
                                    exp_rec=spec.exp_rec > 0)
            if ent and (d.nconst + len(ent)) * maxM[sname] > _lib.PCA_UC_MAX:
                # the derived rows would not fit the consts block: computed
                # per measurement instead
                if spec.exp_rec > 0:
                    raise ValueError("exp_rec: the derived consts rows exceed the "
                                     f"{_lib.PCA_UC_MAX}-value consts block")
                ent = None
                txt = hipgen.emit_model(d, sname, spec.dtype, reassoc=spec.reassoc,
                                        factored=spec.factored)
            structs[sname], entries[sname] = txt, ent or []
        for r, (kind, node) in enumerate(entries[sname]):
            what = "exp factor, times log2(e) in fast-exp float builds" \
                if kind == "expk" else "value"
            head.append(f"// model {i} derived const c[{d.nconst + r}] ({what}): "
                        f"{hipgen.E.to_str(node)}")
        if derived is not None:
            derived.append(entries[sname])
        if members:
            typedefs.append(f"typedef {sname} Model{i};")
        if i == 0:
            typedefs.append(f"typedef {sname} Model;")
    return "\n".join([
        "\n".join(head),
        read_csrc("strict_math.h"),
        read_csrc("prelude.hip"),
        "\n".join(structs.values()),
        "\n".join(typedefs),
        read_csrc("lm_solver.hip"),
        read_csrc("lm_kernel.hip"),
        read_csrc("eval_kernel.hip"),
    ])


_programs: dict = {}
_plock = threading.Lock()
_key_locks: dict = {}


def get_program(spec: LMSpec, members: tuple = ()) -> LMProgram:
    """The program of ``spec`` (and batch ``members``), compiled once per
    process; different specifications compile concurrently."""
    key = (spec, tuple(members))
    with _plock:
        p = _programs.get(key)
        if p is not None:
            return p
        klock = _key_locks.setdefault(key, threading.Lock())
    with klock:
        with _plock:
            p = _programs.get(key)
        if p is None:
            p = LMProgram(spec, tuple(members))
            with _plock:
                _programs[key] = p
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
