"""Seeded synthetic workloads (numpy PCG64, so every machine draws the same
inputs). Rician noise: y = sqrt((S + sigma n1)^2 + (sigma n2)^2), n1, n2 ~ N(0,1),
sigma = S0_true / SNR.

Workloads:
  ivim_c1: IVIM, b = {0,10,20,50,100,200,300,500,700,1000}, SNR 50
  ivim_c2: IVIM, 16 b-values linspace(0, 800), SNR ~ U[20, 100] per voxel
  t2:      S0 exp(-TE/T2), TE = 10..320 ms (32 echoes), T2 ~ U[20, 200] ms
  se:      S0 exp(-(b D)^alpha), 16 b-values linspace(50, 800), alpha ~ U[0.5, 1]
           (b = 0 is left out: d/dalpha (bD)^alpha = (bD)^alpha log(bD) is 0*log 0 = NaN there)
  dki:     61 measurements: b = 0 plus 30 directions at b = 1000 and 2000
IVIM truth: S0 ~ U[0.8,1.2], f ~ U[0.05,0.4], D_2 = D ~ U[0.5e-3,3e-3],
D_1 = Dp ~ U[5e-3,0.1] (the reference's parameter order [S0, f, D_1, D_2]).
"""
from __future__ import annotations

import numpy as np

from .models import dki_design

B_C1 = np.array([0, 10, 20, 50, 100, 200, 300, 500, 700, 1000], dtype=np.float64)
B_C2 = np.linspace(0.0, 800.0, 16)
TE_T2 = np.arange(10.0, 321.0, 10.0)
B_SE = np.linspace(50.0, 800.0, 16)


def _rician(S, sigma, rng):
    n1 = rng.standard_normal(S.shape)
    n2 = rng.standard_normal(S.shape)
    return np.sqrt((S + sigma * n1) ** 2 + (sigma * n2) ** 2)


def ivim_signal(p, b):
    S0, f, D1, D2 = p
    b = b[:, None]
    return S0 * (f * np.exp(-b * D1) + (1 - f) * np.exp(-b * D2))


def ivim_truth(N, rng):
    return np.stack([rng.uniform(0.8, 1.2, N), rng.uniform(0.05, 0.4, N),
                     rng.uniform(5e-3, 0.1, N), rng.uniform(0.5e-3, 3e-3, N)])


def ivim(N, b, seed, snr=50.0):
    """snr: a number or a (lo, hi) range drawn per voxel."""
    rng = np.random.Generator(np.random.PCG64(seed))
    truth = ivim_truth(N, rng)
    S = ivim_signal(truth, b)
    snr_v = rng.uniform(snr[0], snr[1], N) if isinstance(snr, tuple) else np.full(N, float(snr))
    y = _rician(S, truth[0] / snr_v, rng)
    return {"consts": np.asarray(b, np.float64)[None, :], "data": y.astype(np.float32),
            "truth": truth, "snr": snr_v}


def ivim_c1(N=4096, seed=1):
    return ivim(N, B_C1, seed, 50.0)


def ivim_c2(N=1 << 20, seed=2):
    return ivim(N, B_C2, seed, (20.0, 100.0))


def t2(N, seed=3, snr=40.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    truth = np.stack([rng.uniform(0.8, 1.2, N), rng.uniform(20.0, 200.0, N)])
    S = truth[0] * np.exp(-TE_T2[:, None] / truth[1])
    y = _rician(S, truth[0] / snr, rng)
    return {"consts": TE_T2[None, :], "data": y.astype(np.float32), "truth": truth}


def se(N, seed=4, snr=40.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    truth = np.stack([rng.uniform(0.8, 1.2, N), rng.uniform(0.5e-3, 3e-3, N),
                      rng.uniform(0.5, 1.0, N)])
    S = truth[0] * np.exp(-(B_SE[:, None] * truth[1]) ** truth[2])
    y = _rician(S, truth[0] / snr, rng)
    return {"consts": B_SE[None, :], "data": y.astype(np.float32), "truth": truth}


def dki_protocol(seed=5):
    """(4, 61) consts: b = 0 then 30 directions (uniform on the sphere) at
    b = 1000 and b = 2000."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = rng.standard_normal((3, 30))
    g /= np.linalg.norm(g, axis=0, keepdims=True)
    b = np.concatenate([[0.0], np.full(30, 1000.0), np.full(30, 2000.0)])
    gg = np.concatenate([np.zeros((3, 1)), g, g], axis=1)
    return np.concatenate([b[None], gg])


def _random_rotations(N, rng):
    q = rng.standard_normal((N, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)


def dki_truth(N, rng):
    """22 parameters: lnS0, 6 D elements of an SPD tensor with eigenvalues in
    [0.1e-3, 3e-3], 15 K = MD^2 W elements for a kurtosis tensor W that is
    isotropic with K0 ~ U[0, 2] plus a 10% random perturbation."""
    R = _random_rotations(N, rng)
    ev = rng.uniform(0.1e-3, 3e-3, (N, 3))
    D = np.einsum("nij,nj,nkj->nik", R, ev, R)
    md = ev.mean(axis=1)
    k0 = rng.uniform(0.0, 2.0, N)
    W = np.zeros((N, 15))
    W[:, 0:3] = k0[:, None]
    W[:, 9:12] = k0[:, None] / 3.0
    W += 0.1 * k0[:, None] * rng.uniform(-1, 1, (N, 15))
    p = np.empty((22, N))
    p[0] = np.log(rng.uniform(0.8, 1.2, N))
    p[1:7] = np.stack([D[:, 0, 0], D[:, 1, 1], D[:, 2, 2], D[:, 0, 1], D[:, 0, 2], D[:, 1, 2]])
    p[7:] = (md ** 2)[None, :] * W.T
    return p


def dki(N, seed=6, snr=30.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    consts = dki_protocol()
    truth = dki_truth(N, rng)
    A = dki_design(consts)            # (61, 22)
    S = np.exp(A @ truth)             # (61, N)
    y = _rician(S, np.exp(truth[0]) / snr, rng)
    return {"consts": consts, "data": y.astype(np.float32), "truth": truth}


def mixed(n_t2=400_000, n_ivim=400_000, n_se=200_000, seed=7, snr=40.0):
    """Mixed-model workload: mono-exponential T2 (32 echoes), IVIM (16
    b-values 0-800) and stretched exponential (16 b-values), Rician SNR 40,
    each with its own seeded stream."""
    return {"t2": t2(n_t2, seed, snr), "ivim": ivim(n_ivim, B_C2, seed + 1, snr),
            "se": se(n_se, seed + 2, snr)}


def dki_truth_gpu(N, seed, device="cuda"):
    """The kurtosis-model parameter distribution of dki_truth, drawn on the GPU
    with a seeded torch generator (training sets of millions of voxels); float64
    (22, N)."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    kw = dict(generator=gen, device=device, dtype=torch.float64)
    q = torch.randn((N, 4), **kw)
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.T
    R = torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        torch.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        torch.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    ev = 0.1e-3 + (3e-3 - 0.1e-3) * torch.rand((N, 3), **kw)
    D = torch.einsum("nij,nj,nkj->nik", R, ev, R)
    md = ev.mean(dim=1)
    k0 = 2.0 * torch.rand(N, **kw)
    W = torch.zeros((N, 15), device=device, dtype=torch.float64)
    W[:, 0:3] = k0[:, None]
    W[:, 9:12] = k0[:, None] / 3.0
    W += 0.1 * k0[:, None] * (2.0 * torch.rand((N, 15), **kw) - 1.0)
    p = torch.empty((22, N), device=device, dtype=torch.float64)
    p[0] = torch.log(0.8 + 0.4 * torch.rand(N, **kw))
    p[1:7] = torch.stack([D[:, 0, 0], D[:, 1, 1], D[:, 2, 2], D[:, 0, 1], D[:, 0, 2], D[:, 1, 2]])
    p[7:] = (md ** 2)[None, :] * W.T
    return p
