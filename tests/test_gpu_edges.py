"""Edge sizes on the GPU: fits of 1, 63, 65 and 257 voxels equal the same
voxels of a larger fit bit for bit (voxels are independent, and partial
wavefronts and workgroups are masked), for the register kernel (single launch
and chunked), the wide-model kernel and the fused MLP inference; a fit of no
voxels is a no-op."""
import numpy as np
import pytest
import torch

from pycompute_amd import fit, mlp, synth, volume
from pycompute_amd.models import IVIM

pytestmark = pytest.mark.gpu

SIZES = (1, 63, 65, 257)
STATE = ("pars_t", "f_t", "lam_t", "step_type_t", "n_iter_t")


def _sub(prob, n):
    return fit.make_problem(prob.model, prob.consts, prob.data_t[:, :n].contiguous())


@pytest.mark.parametrize("compaction", [False, True])
def test_register_kernel_ragged_sizes(gpu, compaction):
    w = synth.ivim_c1(1000, seed=31)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    ref = fit.fit(prob, iters=60, params0=p0.clone(), compaction=compaction, chunk=8)
    torch.cuda.synchronize()
    for n in SIZES:
        s = fit.fit(_sub(prob, n), iters=60, params0=p0[:, :n].clone(), compaction=compaction,
                    chunk=8)
        torch.cuda.synchronize()
        for k in STATE:
            assert torch.equal(getattr(s, k), getattr(ref, k)[:, :n]), (n, k)


def test_wide_kernel_ragged_sizes(gpu):
    w = synth.dki(200, seed=32)
    prob = fit.make_problem("dki", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    ref = fit.fit(prob, iters=30, params0=p0.clone(), wide=True)
    torch.cuda.synchronize()
    for n in (1, 65):
        s = fit.fit(_sub(prob, n), iters=30, params0=p0[:, :n].clone(), wide=True)
        torch.cuda.synchronize()
        for k in STATE:
            assert torch.equal(getattr(s, k), getattr(ref, k)[:, :n]), (n, k)


def test_inference_ragged_sizes(gpu):
    w = synth.ivim_c2(1000, seed=33)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    net = mlp.ivim_mlp(synth.B_C2, seed=2)
    net.finalize()
    fb = (prob.model, prob.consts)
    ref = torch.zeros((4, prob.N), device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, ref, fallback=fb)
    for n in SIZES:
        sp = _sub(prob, n)
        out = torch.zeros((4, n), device="cuda")
        net.infer(sp.data_t, sp.lb_t, sp.ub_t, out, fallback=fb)
        torch.cuda.synchronize()
        assert torch.equal(out, ref[:, :n]), n


def test_zero_voxels_is_a_no_op(gpu):
    w = synth.ivim_c1(8, seed=34)
    prob = _sub(fit.make_problem("ivim", w["consts"], w["data"]), 0)
    assert prob.N == 0
    p0 = fit.default_init(prob)
    assert tuple(p0.shape) == (4, 0)
    s = fit.fit(prob, iters=10, params0=p0)
    assert tuple(s.pars_t.shape) == (4, 0) and tuple(s.final_f().shape) == (1, 0)
    net = mlp.ivim_mlp(synth.B_C1, seed=0)
    net.finalize()
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, p0, fallback=(prob.model, prob.consts))
    volume.canonicalize_ivim(p0)
    torch.cuda.synchronize()
