"""Volume layout and IVIM canonicalisation on the GPU
(test/ivim/test_ivim.ipynb cells 5, 19 and :441-446)."""
import numpy as np
import pytest
import torch

from pycompute_amd import fit, synth, volume
from pycompute_amd.lsqnonlin import F
from pycompute_amd.models import IVIM

pytestmark = pytest.mark.gpu


def test_layout_round_trip(gpu):
    rng = np.random.default_rng(1)
    vol = rng.standard_normal((6, 5, 4, 10)).astype(np.float32)
    vol[1, 2, 3, 4] = np.nan
    vol[0, 0, 0, 0] = np.inf
    y = volume.to_voxels(vol)
    want = np.nan_to_num(np.transpose(vol, (3, 0, 1, 2)).reshape(10, -1), nan=0.0, posinf=0.0,
                         neginf=0.0)
    np.testing.assert_array_equal(y.cpu().numpy(), want)
    back = volume.to_volume(y, vol.shape).cpu().numpy()
    np.testing.assert_array_equal(back, np.nan_to_num(vol, nan=0.0, posinf=0.0, neginf=0.0))


def test_canonicalize_matches_notebook_and_keeps_the_model(gpu):
    rng = np.random.default_rng(2)
    N = 10000
    p = np.stack([rng.uniform(0.8, 1.2, N), rng.uniform(0, 1, N), rng.uniform(1e-4, 0.05, N),
                  rng.uniform(1e-4, 0.05, N)]).astype(np.float32)
    t = torch.tensor(p, device="cuda")
    volume.canonicalize_ivim(t)
    torch.cuda.synchronize()
    want = p.T.copy()                              # the notebook works on (voxels, 4)
    gemask = want[:, 2] < want[:, 3]
    tmp = want[gemask, 2].copy()
    want[gemask, 2] = want[gemask, 3].copy()
    want[gemask, 3] = tmp.copy()
    want[gemask, 1] = 1 - want[gemask, 1]
    np.testing.assert_array_equal(t.cpu().numpy(), want.T)
    assert 0.3 < gemask.mean() < 0.7
    # the model is symmetric: the SSR against any data is unchanged (to rounding)
    b = synth.B_C2
    y = torch.rand((16, N), device="cuda")
    f0 = torch.zeros((1, N), device="cuda")
    f1 = torch.zeros((1, N), device="cuda")
    Fo = F(IVIM.expr, list(IVIM.pars), list(IVIM.consts), 16, torch.float32)
    cb = torch.tensor(b, dtype=torch.float32, device="cuda")
    Fo.run(torch.tensor(p, device="cuda"), cb, y, None, f0)
    Fo.run(t, cb, y, None, f1)
    torch.cuda.synchronize()
    np.testing.assert_allclose(f1.cpu().numpy(), f0.cpu().numpy(), rtol=1e-5)


def test_volume_fit_end_to_end(gpu):
    X, Y, Z = 16, 12, 8
    w = synth.ivim(X * Y * Z, synth.B_C2, seed=9, snr=(20.0, 100.0))
    vol = np.transpose(w["data"].reshape(16, X, Y, Z), (1, 2, 3, 0))    # (X, Y, Z, M)
    y = volume.to_voxels(vol)
    np.testing.assert_array_equal(y.cpu().numpy(), w["data"])
    N = y.shape[1]
    consts = volume.replicated_consts(synth.B_C2, N)
    prob = fit.make_problem("ivim", synth.B_C2, w["data"])
    p0 = fit.default_init(prob)
    s = fit.make_solver(prob)
    pars = p0.clone()
    s.setup(pars, consts, y, prob.lb_t, prob.ub_t)       # the reference's (1, M, N) layout
    s.run(100)
    volume.canonicalize_ivim(pars)
    maps = volume.to_volume(pars, vol.shape)
    torch.cuda.synchronize()
    assert tuple(maps.shape) == (X, Y, Z, 4)
    assert bool((maps[..., 2] >= maps[..., 3]).all())
    ref = fit.fit(prob, iters=100, params0=p0.clone())
    volume.canonicalize_ivim(ref.pars_t)
    assert torch.equal(pars, ref.pars_t)
