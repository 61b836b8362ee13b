"""The C-ABI library loads, exports every entry point include/pycompute_amd.h
declares, compiles with hiprtc, and rejects bad arguments -- all without a
GPU (no kernel is launched here)."""
import ctypes
import os
import re

import pytest

from pycompute_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pycompute_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pca_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("pca_compile", "pca_module_load", "pca_lm_launch", "pca_ssr_launch",
              "pca_eval_launch", "pca_bpolyfit", "pca_loglinear_init", "pca_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.pca_abi_version() == _lib.ABI_VERSION


def test_struct_layouts_match_header():
    # pca_lm_args: 15 pointers, int64 N, 2 x int32, 5 doubles
    assert ctypes.sizeof(_lib.LMArgs) == 15 * 8 + 8 + 8 + 5 * 8
    assert _lib.LMArgs.tol.offset == 136
    assert ctypes.sizeof(_lib.SSRArgs) == 6 * 8
    assert ctypes.sizeof(_lib.EvalArgs) == 6 * 8


def test_hiprtc_compile_ok_and_error():
    code = _lib.compile_source('extern "C" __global__ void k(float* x) { x[threadIdx.x] = 1.0f; }',
                               "probe.hip", ["--offload-arch=gfx950", "-O3"])
    assert code[:4] == b"\x7fELF"
    with pytest.raises(_lib.PcaError) as e:
        _lib.compile_source("this is not C++", "bad.hip", ["--offload-arch=gfx950"])
    assert "hiprtc" in str(e.value)


def test_argument_errors_without_gpu():
    L = _lib.lib()
    a = _lib.LMArgs()
    a.N = 10                                    # per-voxel arrays required for N > 0
    rc = L.pca_lm_launch(None, ctypes.byref(a), None, 0, 256, None)
    assert rc == -1 and b"NULL" in L.pca_last_error()
    L.pca_bpolyfit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                               ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                               ctypes.c_void_p]
    assert L.pca_bpolyfit(1, 0, 1, 4, 9, 1, 10, 0, None) == -1
    assert b"deg" in L.pca_last_error()
    big = (ctypes.c_char * (_lib.PCA_UC_MAX * 8 + 8))()
    a.data = a.lb = a.ub = a.params = a.lam = a.step_type = a.f = 1
    rc = L.pca_lm_launch(None, ctypes.byref(a), big, ctypes.sizeof(big), 256, None)
    assert rc == -4 and b"PCA_UC_MAX" in L.pca_last_error()


def test_mlp_infer_args_layout_and_validation():
    """pca_mlp_infer_args as the header lays it out (the optional fallback
    fields after `reserved`), and argument errors caught before any launch."""
    from pycompute_amd import mlp
    A = mlp.InferArgs
    assert A.N.offset == 12 * 8 and A.reserved.offset == 12 * 8 + 8 + 7 * 4
    assert A.fallback_P.offset == 136 and A.fallback_kind.offset == 148
    assert ctypes.sizeof(A) == 152
    L = mlp._bind()
    a = A(1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1000, 16, 4, 128, 0, 0, -1, 12, 0,
          None, 0, 0)
    a.hidden = 100                                   # not 128 or 256
    assert L.pca_mlp_infer(ctypes.byref(a), None) == -1
    assert b"hidden" in L.pca_last_error()
    a.hidden = 128
    a.fallback_P, a.fallback_K, a.fallback_kind = 1, 2, 1   # T2 guess for a 4-output net
    assert L.pca_mlp_infer(ctypes.byref(a), None) == -1
    assert b"fallback" in L.pca_last_error()
    a.w1 = None
    assert L.pca_mlp_infer(ctypes.byref(a), None) == -1
    assert b"NULL" in L.pca_last_error()


def test_mlp_train_step_validation():
    """The fused training step and the gradient reduction reject what their
    kernels do not cover before any launch."""
    from pycompute_amd import mlp
    L = mlp._bind()
    offs = (ctypes.c_int32 * 6)(0, 1024, 1088, 5184, 5248, 5504)   # hidden 64, 16 in, 4 out
    np_ = 5512

    def step(B=8192, hidden=64, inp=16, out=4, n=np_, ldx=8192):
        return L.pca_mlp_train_step(1, ldx, 1, B, 1, hidden, inp, out, offs, n, 1, 1, None)

    for kw, msg in (({"B": 100}, b"multiple of 64"), ({"hidden": 256}, b"hidden"),
                    ({"inp": 32}, b"hidden"), ({"out": 17}, b"hidden"),
                    ({"n": 5000}, b"outside np"), ({"ldx": 64}, b"bad argument")):
        assert step(**kw) == -1, kw
        assert msg in L.pca_last_error(), (kw, L.pca_last_error())
    assert L.pca_mlp_grad_reduce(1, 1, 0, np_, 1, None, None, None, None, 1e-3, 0.9, 0.999,
                                 1e-8, 1, 0.0, None) == -1
    assert L.pca_mlp_grad_reduce(1, 1, 128, np_, 1, None, 1, 1, 1, 1e-3, 0.9, 0.999, 1e-8, 0,
                                 0.0, None) == -1          # Adam needs step >= 1
    assert b"pca_mlp_grad_reduce" in L.pca_last_error()
    # the ordered GEMM and column sums check their workspace size
    assert L.pca_mlp_gemm_ordered(1, 64, 1, 1, 64, 0, 1, 64, 64, 64, 8192, 32, 1,
                                  32 * 64 * 64 - 1, None) == -1
    assert b"workspace" in L.pca_last_error()
    assert L.pca_mlp_colsum_ordered(1, 64, 8192, 64, 1, 1, 128 * 64 - 1, None) == -1
    assert b"workspace" in L.pca_last_error()


def test_zero_voxel_calls_accept_empty_arrays():
    """With no voxels the per-voxel arrays of an empty tensor have no address;
    the entry points return PCA_OK without launching (no GPU needed)."""
    from pycompute_amd import mlp
    L = _lib.lib()
    P = (ctypes.c_double * 8)()
    L.pca_loglinear_init.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p]
    assert L.pca_loglinear_init(ctypes.addressof(P), 2, 4, None, 1, 2, None, None, None, 0, 0,
                                None) == 0
    assert L.pca_loglinear_init(ctypes.addressof(P), 2, 4, None, 1, 2, None, None, None, 5, 0,
                                None) == -1
    L.pca_ivim_canonicalize.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_void_p]
    assert L.pca_ivim_canonicalize(None, 0, 0, None) == 0
    Lm = mlp._bind()
    a = mlp.InferArgs(None, 1, 1, 1, 1, 1, 1, 1, 1, None, None, None, 0, 16, 4, 128, 0, 0, -1,
                      12, 0, None, 0, 0)
    assert Lm.pca_mlp_infer(ctypes.byref(a), None) == 0
    a.N = 10
    assert Lm.pca_mlp_infer(ctypes.byref(a), None) == -1
