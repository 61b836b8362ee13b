"""pca_lm_chunked_launch_static, the fit that allocates nothing (its queues
and counters live in a file-scope device array), called through the C-ABI:
two calls in a row, the second on fewer voxels than the first, each equal
bit for bit to pca_lm_chunked_launch (caller-allocated queues) with the same
chunk schedule; sizes beyond the array are refused before any launch."""
import ctypes

import pytest
import torch

from pycompute_amd import _lib, chunked, fit, synth

pytestmark = pytest.mark.gpu

STATE = ("pars_t", "step_type_t", "lam_t", "f_t", "first_f", "last_f", "n_iter_t")
KNOBS = dict(fast_exp=True, fast_div=True, reassoc=True, factored=True, chunk=8)


def _static_fit(prob, p0, iters, chunk):
    s = fit.make_solver(prob, compaction=True, keep_workspace=False, **KNOBS)
    s.setup(p0.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    uc = torch.tensor(s._ucx, dtype=torch.float32, device=prob.data_t.device)
    args = (_lib.LMArgs * 1)(s.lm_args(iters, 1e-8, 0.25, 0.75, 0.2, 2.0))
    ucp = (ctypes.c_void_p * 1)(uc.data_ptr())
    npar = (ctypes.c_int32 * 1)(4)
    sched = chunked.chunk_schedule(iters, chunk)
    k = (ctypes.c_int32 * len(sched))(*sched)
    waves = (prob.N + 63) // 64
    grid = max(1, min((waves * 64 + s.block - 1) // s.block, 8192))
    L = _lib.lib()
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.pca_lm_chunked_launch_static(s.program.kernel("pca_lm_chunk"), args, ucp, npar, 1, 0,
                                        k, len(sched), grid, s.block, 0, stream)
    _lib.check(rc, "pca_lm_chunked_launch_static")
    torch.cuda.synchronize()
    return s


def _reference_fit(prob, p0, iters):
    s = fit.make_solver(prob, compaction=True, keep_workspace=False, **KNOBS)
    s.setup(p0.clone(), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    s.run(iters)
    torch.cuda.synchronize()
    return s


@pytest.mark.parametrize("n,seed", [(300_000, 71), (70_001, 72)], ids=["first", "smaller"])
def test_static_equals_allocating_launch(gpu, n, seed):
    w = synth.ivim(n, synth.B_C2, seed=seed, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = fit.default_init(prob)
    ref = _reference_fit(prob, p0, 100)
    got = _static_fit(prob, p0, 100, KNOBS["chunk"])
    for k in STATE:
        assert torch.equal(getattr(ref, k), getattr(got, k)), k


def test_two_calls_in_a_row(gpu):
    """A large call, then a smaller one: the second does not see the first's
    queue contents or counters."""
    outs = []
    for n, seed in ((400_000, 73), (5_000, 74)):
        w = synth.ivim(n, synth.B_C2, seed=seed, snr=(20.0, 100.0))
        prob = fit.make_problem("ivim", w["consts"], w["data"])
        p0 = fit.default_init(prob)
        got = _static_fit(prob, p0, 100, 8)
        ref = _reference_fit(prob, p0, 100)
        outs.append(all(torch.equal(getattr(ref, k), getattr(got, k)) for k in STATE))
    assert outs == [True, True]


def test_oversize_is_refused_before_launch(gpu):
    """A voxel count whose queues exceed the array is refused on the host
    (PCA_ERR_LIMIT) before the size is used for anything."""
    L = _lib.lib()
    cap = L.pca_static_scratch_bytes()
    assert cap == 96 << 20
    w = synth.ivim(64, synth.B_C2, seed=75, snr=50.0)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    s = fit.make_solver(prob, compaction=True, **KNOBS)
    s.setup(fit.default_init(prob), prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    a = s.lm_args(10, 1e-8, 0.25, 0.75, 0.2, 2.0)
    a.N = cap // 16           # two queues of 4-parameter float32 entries need more
    args = (_lib.LMArgs * 1)(a)
    npar = (ctypes.c_int32 * 1)(4)
    k = (ctypes.c_int32 * 1)(10)
    rc = L.pca_lm_chunked_launch_static(s.program.kernel("pca_lm_chunk"), args, None, npar, 1,
                                        0, k, 1, 1, 256, 0, None)
    assert rc == -4 and b"static scratch" in L.pca_last_error()
