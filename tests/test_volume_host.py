"""Volume I/O (no GPU): NIfTI-1 round trips and header layout, and the
notebook's (X, Y, Z, M) <-> (M, N) layout restated in numpy."""
import gzip
import struct

import numpy as np
import pytest

from pycompute_amd import volume


@pytest.mark.parametrize("suffix", [".nii", ".nii.gz"])
@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.float64])
def test_nifti_round_trip(tmp_path, suffix, dtype):
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((5, 4, 3, 7)) * 100).astype(dtype)
    p = tmp_path / ("img" + suffix)
    volume.write_nifti(p, a, pixdim=(2.0, 2.0, 3.0, 1.0))
    b, h = volume.read_nifti(p)
    assert b.shape == a.shape and h["dim"] == a.shape
    np.testing.assert_array_equal(np.asarray(b), a)
    assert h["pixdim"] == (2.0, 2.0, 3.0, 1.0) and h["vox_offset"] == 352
    raw = (gzip.open(p, "rb") if suffix.endswith(".gz") else open(p, "rb")).read(348)
    assert struct.unpack("<i", raw[:4])[0] == 348 and raw[344:348] == b"n+1\x00"
    # first voxel of the data block is a[0,0,0,0] (column-major on disk)
    full = (gzip.open(p, "rb") if suffix.endswith(".gz") else open(p, "rb")).read()
    first = np.frombuffer(full[352:352 + a.itemsize], dtype=a.dtype)[0]
    assert first == a[0, 0, 0, 0]


def test_nifti_scaling(tmp_path):
    a = np.arange(24, dtype=np.int16).reshape(2, 3, 4)
    p = tmp_path / "s.nii"
    volume.write_nifti(p, a)
    raw = bytearray(open(p, "rb").read())
    struct.pack_into("<ff", raw, 112, 0.5, 10.0)       # scl_slope, scl_inter
    open(p, "wb").write(bytes(raw))
    b, h = volume.read_nifti(p)
    np.testing.assert_allclose(b, a * 0.5 + 10.0)
    assert b.dtype == np.float32


def test_nifti_rejects_other_formats(tmp_path):
    p = tmp_path / "bad.nii"
    open(p, "wb").write(b"\x00" * 400)
    with pytest.raises(ValueError):
        volume.read_nifti(p)
