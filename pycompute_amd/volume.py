"""Volume I/O around the fit (test/ivim/test_ivim.ipynb cells 5, 19, 27 and
:441-446).

    y = to_voxels(vol)                    # (X, Y, Z, M) -> (M, N), N = X*Y*Z, C order
    c = replicated_consts(b, N)           # (1, M, N) view of one protocol (no copy)
    ... fit ...
    canonicalize_ivim(pars_t)             # D_1 >= D_2, f -> 1 - f where swapped
    maps = to_volume(pars_t, vol.shape)   # (n, N) -> (X, Y, Z, n)

Layouts follow the notebook: data_flat = transpose(data, (3,0,1,2)).reshape(M,
N) with non-finite samples set to 0 (nan_to_num); results come back with
transpose(reshape(p, (n, X, Y, Z)), (1, 2, 3, 0)). The permutations run on the
GPU (torch copies); canonicalisation is the kernel pca_ivim_canonicalize.

read_nifti / write_nifti: NIfTI-1 single-file images (.nii, .nii.gz) by the
published header layout (348-byte header, data at vox_offset, scl_slope /
scl_inter applied on read), memory-mapped when uncompressed.
"""
from __future__ import annotations

import ctypes
import gzip
import struct

import numpy as np
import torch

from . import _lib

# NIfTI-1 datatype codes
_NIFTI_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64,
                 256: np.int8, 512: np.uint16, 768: np.uint32}
_NIFTI_CODES = {np.dtype(v): k for k, v in _NIFTI_DTYPES.items()}


def to_voxels(vol, device="cuda", dtype=torch.float32):
    """(X, Y, Z, M) signals -> contiguous (M, N) on the GPU, non-finite -> 0."""
    t = torch.as_tensor(np.asarray(vol) if not isinstance(vol, torch.Tensor) else vol)
    if t.dim() != 4:
        raise ValueError(f"expected an (X, Y, Z, M) volume, got shape {tuple(t.shape)}")
    t = t.to(device=device, dtype=dtype)
    M = t.shape[3]
    y = t.permute(3, 0, 1, 2).reshape(M, -1).contiguous()
    return torch.nan_to_num(y, nan=0.0, posinf=0.0, neginf=0.0)


def to_volume(p_t, shape):
    """(n, N) per-voxel values -> (X, Y, Z, n), shape = (X, Y, Z[, ...])."""
    X, Y, Z = (int(s) for s in shape[:3])
    n = p_t.shape[0]
    if p_t.shape[1] != X * Y * Z:
        raise ValueError(f"{p_t.shape[1]} voxels do not fill a {X}x{Y}x{Z} volume")
    return p_t.reshape(n, X, Y, Z).permute(1, 2, 3, 0)


def replicated_consts(consts, N, device="cuda", dtype=torch.float32):
    """The reference's replicated consts layout (1, nconst*M, N) as a stride-0
    view of one protocol (the solver detects it as voxel-uniform)."""
    c = torch.as_tensor(np.asarray(consts, np.float64).reshape(-1), dtype=dtype, device=device)
    return c.reshape(1, -1, 1).expand(1, c.numel(), int(N))


def canonicalize_ivim(pars_t, stream=None):
    """In place on (4, N) IVIM parameters (S0, f, D_1, D_2): where D_1 < D_2,
    swap them and set f = 1 - f (test_ivim.ipynb:441-446)."""
    if not (isinstance(pars_t, torch.Tensor) and pars_t.is_cuda):
        raise ValueError("pars_t must be a torch tensor on the GPU")
    if pars_t.dim() != 2 or pars_t.shape[0] != 4 or not pars_t.is_contiguous():
        raise ValueError("pars_t must be a contiguous (4, N) tensor")
    if pars_t.dtype not in (torch.float32, torch.float64):
        raise RuntimeError("only float32 and float64 are supported")
    L = _lib.lib()
    L.pca_ivim_canonicalize.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                        ctypes.c_void_p]
    s = torch.cuda.current_stream(pars_t.device).cuda_stream if stream is None else stream
    _lib.check(L.pca_ivim_canonicalize(pars_t.data_ptr(), int(pars_t.shape[1]),
                                       int(pars_t.dtype == torch.float64), s),
               "pca_ivim_canonicalize")
    return pars_t


# ---------------------------------------------------------------------------
# NIfTI-1
# ---------------------------------------------------------------------------

def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def read_nifti(path, mmap: bool = True):
    """(array, header) of a single-file NIfTI-1 image. The array is in the
    file's axis order (X, Y, Z, ...), scaled by scl_slope / scl_inter when
    those are set (then float32/float64)."""
    with _open(path, "rb") as fh:
        hdr = fh.read(348)
    if len(hdr) != 348:
        raise ValueError("truncated NIfTI header")
    endian = "<" if struct.unpack("<i", hdr[0:4])[0] == 348 else ">"
    if struct.unpack(endian + "i", hdr[0:4])[0] != 348:
        raise ValueError("not a NIfTI-1 file (sizeof_hdr != 348)")
    magic = hdr[344:348]
    if magic not in (b"n+1\x00",):
        raise ValueError(f"unsupported NIfTI magic {magic!r} (single-file n+1 only)")
    dim = struct.unpack(endian + "8h", hdr[40:56])
    datatype, bitpix = struct.unpack(endian + "hh", hdr[70:74])
    pixdim = struct.unpack(endian + "8f", hdr[76:108])
    vox_offset = struct.unpack(endian + "f", hdr[108:112])[0]
    scl_slope, scl_inter = struct.unpack(endian + "ff", hdr[112:120])
    if datatype not in _NIFTI_DTYPES:
        raise ValueError(f"unsupported NIfTI datatype {datatype}")
    nd = dim[0]
    shape = tuple(int(d) for d in dim[1:1 + nd])
    dt = np.dtype(_NIFTI_DTYPES[datatype]).newbyteorder(endian)
    off = int(vox_offset)
    count = int(np.prod(shape))
    if mmap and not str(path).endswith(".gz"):
        arr = np.memmap(path, dtype=dt, mode="r", offset=off, shape=shape, order="F")
    else:
        with _open(path, "rb") as fh:
            fh.read(off)
            buf = fh.read(count * dt.itemsize)
        arr = np.frombuffer(buf, dtype=dt, count=count).reshape(shape, order="F")
    if scl_slope not in (0.0, 1.0) or scl_inter != 0.0:
        slope = scl_slope if scl_slope != 0.0 else 1.0
        out_t = np.float64 if arr.dtype == np.float64 else np.float32
        arr = arr.astype(out_t) * out_t(slope) + out_t(scl_inter)
    header = {"dim": shape, "datatype": datatype, "bitpix": bitpix,
              "pixdim": tuple(pixdim[1:1 + nd]), "vox_offset": off,
              "scl_slope": scl_slope, "scl_inter": scl_inter}
    return arr, header


def write_nifti(path, arr, pixdim=None):
    """Write an array (X, Y, Z[, T]) as a single-file NIfTI-1 image (little
    endian, voxel data after a 352-byte header + extension flag)."""
    a = np.asarray(arr)
    if a.dtype not in _NIFTI_CODES:
        a = a.astype(np.float32)
    code = _NIFTI_CODES[a.dtype]
    nd = a.ndim
    if not 1 <= nd <= 7:
        raise ValueError("NIfTI-1 holds 1 to 7 dimensions")
    dim = [nd] + list(a.shape) + [1] * (7 - nd)
    pd = [1.0] + list(pixdim if pixdim is not None else [1.0] * nd) + [1.0] * (7 - nd)
    hdr = bytearray(348)
    struct.pack_into("<i", hdr, 0, 348)
    struct.pack_into("<8h", hdr, 40, *dim)
    struct.pack_into("<hh", hdr, 70, code, a.dtype.itemsize * 8)
    struct.pack_into("<8f", hdr, 76, *pd[:8])
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<ff", hdr, 112, 1.0, 0.0)
    struct.pack_into("<B", hdr, 123, 10 | 8)   # xyzt_units: mm, s
    hdr[344:348] = b"n+1\x00"
    with _open(path, "wb") as fh:
        fh.write(bytes(hdr))
        fh.write(b"\x00\x00\x00\x00")             # no extensions
        fh.write(np.asfortranarray(a.astype(a.dtype.newbyteorder("<"))).tobytes(order="F"))
