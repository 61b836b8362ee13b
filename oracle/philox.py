"""Philox4x32-10 counter-based generator in numpy (test infrastructure).

Algorithm of Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy
as 1, 2, 3" (SC 2011): 10 rounds of
    (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2,
    c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0),  k += (W0, W1)
with M0 = 0xD2511F53, M1 = 0xCD9E8D57, W0 = 0x9E3779B9, W1 = 0xBB67AE85.
Used to replay the device's random-search candidates (csrc/eval_kernel.hip,
pca::random_candidate) on the host.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c, k):
    """c: (4, ...) uint32 counters, k: (2, ...) uint32 keys -> (4, ...) uint32."""
    c = [np.asarray(x, np.uint32).copy() for x in c]
    k0, k1 = (np.asarray(x, np.uint32).copy() for x in k)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c[0].astype(np.uint64)
            p1 = M1 * c[2].astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK).astype(np.uint32)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
            k0 = (k0 + W0).astype(np.uint32)
            k1 = (k1 + W1).astype(np.uint32)
    return np.stack(c)


def random_candidates(seed, n_voxels, nparam, draws, lo, hi, dtype=np.float32):
    """Candidates (len(draws), nparam, n_voxels) of the device's random search:
    candidate k of voxel v uses counter (v, v >> 32, k, 0x7a110000 + j // 4)."""
    lo = np.asarray(lo, dtype)
    hi = np.asarray(hi, dtype)
    v = np.arange(n_voxels, dtype=np.uint64)
    k = (np.uint32(seed & 0xFFFFFFFF), np.uint32((seed >> 32) & 0xFFFFFFFF))
    out = np.empty((len(draws), nparam, n_voxels), dtype)
    for a, d in enumerate(draws):
        for j0 in range(0, nparam, 4):
            c = (v.astype(np.uint32), (v >> np.uint64(32)).astype(np.uint32),
                 np.full(n_voxels, d, np.uint32), np.full(n_voxels, 0x7A110000 + j0 // 4,
                                                          np.uint32))
            r = philox4x32_10(c, (np.full(n_voxels, k[0]), np.full(n_voxels, k[1])))
            for q in range(4):
                j = j0 + q
                if j < nparam:
                    u = ((r[q].astype(np.float64) + 0.5) * 2.3283064365386963e-10).astype(dtype)
                    out[a, j] = lo[j] + (hi[j] - lo[j]) * u
    return out
