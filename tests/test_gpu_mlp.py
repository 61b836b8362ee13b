"""GPU checks of the learned initialiser: the bf16-MFMA GEMM against torch,
training gradients against the CPU oracle (oracle/mlp.py, bf16 rounding at the
same points), the fp16 fused inference kernels against the float64 oracle
forward, the in-kernel fallback, the synthetic signal generator against the
model, and what the trained network gains."""
import numpy as np
import pytest
import torch

import parity
from oracle import mlp as omlp
from pycompute_amd import fit, synth
from pycompute_amd import mlp
from pycompute_amd.models import IVIM, T2, ModelSpec

pytestmark = pytest.mark.gpu


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


@pytest.mark.parametrize("transA,transB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 36, 70), (4, 128, 1000), (513, 4, 128)])
def test_gemm_matches_torch(gpu, transA, transB, M, N, K):
    g = torch.Generator(device="cuda").manual_seed(M * 7 + N + K)
    A = torch.randn((K, M) if transA else (M, K), device="cuda", generator=g)
    B = torch.randn((N, K) if transB else (K, N), device="cuda", generator=g)
    ref = (_bf(A).T if transA else _bf(A)) @ (_bf(B).T if transB else _bf(B))
    for splits in (1, 3):
        C = torch.zeros((M, N), device="cuda")
        mlp.gemm(A, A.shape[1], transA, B, B.shape[1], transB, C, N, M, N, K, splits=splits)
        torch.cuda.synchronize()
        err = (C - ref).abs().max().item() / ref.abs().max().item()
        assert err < 1e-5, (splits, err)


def test_gemm_epilogues(gpu):
    g = torch.Generator(device="cuda").manual_seed(3)
    M, N, K = 130, 48, 40
    A = torch.randn((M, K), device="cuda", generator=g)
    B = torch.randn((K, N), device="cuda", generator=g)
    bias = torch.randn(N, device="cuda", generator=g)
    acc = _bf(A) @ _bf(B)
    C = torch.empty((M, N), device="cuda")
    Z = torch.empty((M, N), device="cuda")
    mlp.gemm(A, K, 0, B, N, 0, C, N, M, N, K, mlp.EPI_BIAS_GELU, bias, Z)
    torch.cuda.synchronize()
    z = (acc + bias).double().cpu().numpy()
    np.testing.assert_allclose(Z.cpu().numpy(), z, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(C.cpu().numpy(), omlp.gelu(z), rtol=1e-4, atol=1e-5)
    C2 = torch.empty((M, N), device="cuda")
    mlp.gemm(A, K, 0, B, N, 0, C2, N, M, N, K, mlp.EPI_GELU_GRAD, Zin=Z)
    torch.cuda.synchronize()
    np.testing.assert_allclose(C2.cpu().numpy(), acc.double().cpu().numpy() * omlp.gelu_grad(z),
                               rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("M,N,ld", [(1, 4, 4), (63, 22, 22), (8192, 4, 4), (8192, 64, 64),
                                    (777, 300, 301), (100, 256, 260)])
def test_colsum_matches_float64(gpu, M, N, ld):
    """The bias-gradient column sums (out += X.sum(0) over the first N of ld
    columns) against float64, with several column tiles, strided rows and
    more row chunks than rows."""
    from pycompute_amd import _lib
    g = torch.Generator(device="cuda").manual_seed(M + N)
    X = torch.randn((M, ld), device="cuda", generator=g)
    out = torch.randn(N, device="cuda", generator=g)
    want = out.double() + X[:, :N].double().sum(0)
    L = mlp._bind()
    _lib.check(L.pca_mlp_colsum(X.data_ptr(), ld, M, N, out.data_ptr(),
                                torch.cuda.current_stream().cuda_stream), "colsum")
    torch.cuda.synchronize()
    err = (out.double() - want).abs().max().item()
    assert err <= 1e-5 * max(1.0, float(np.sqrt(M))), err


@pytest.mark.parametrize("transB,M,N,K,splits", [(0, 4, 64, 8192, 32), (0, 64, 64, 8192, 32),
                                                  (1, 64, 16, 8192, 32), (0, 100, 36, 1000, 3),
                                                  (1, 256, 256, 4096, 16), (0, 17, 5, 40, 1)])
def test_gemm_ordered_matches_torch_and_repeats(gpu, transB, M, N, K, splits):
    """pca_mlp_gemm_ordered (transA, the weight-gradient form): against torch
    on bf16-rounded operands, into a C with a wider row stride, and the same
    bits on a second call (split partials summed in order, no atomics)."""
    from pycompute_amd import _lib
    g = torch.Generator(device="cuda").manual_seed(M + N + K)
    A = torch.randn((K, M), device="cuda", generator=g)
    B = torch.randn((N, K) if transB else (K, N), device="cuda", generator=g)
    ref = _bf(A).T @ (_bf(B).T if transB else _bf(B))
    L = mlp._bind()
    ws = torch.empty(splits * M * N, device="cuda")
    outs = []
    for _ in range(2):
        C = torch.full((M, N + 3), 7.0, device="cuda")
        _lib.check(L.pca_mlp_gemm_ordered(A.data_ptr(), M, 1, B.data_ptr(), B.shape[1], transB,
                                          C.data_ptr(), N + 3, M, N, K, splits, ws.data_ptr(),
                                          ws.numel(), torch.cuda.current_stream().cuda_stream),
                   "gemm_ordered")
        torch.cuda.synchronize()
        assert torch.all(C[:, N:] == 7.0)                 # the pad columns untouched
        outs.append(C[:, :N].clone())
    err = (outs[0] - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-5, err
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("M,N", [(8192, 4), (8192, 256), (100, 22)])
def test_colsum_ordered_matches_float64_and_repeats(gpu, M, N):
    from pycompute_amd import _lib
    g = torch.Generator(device="cuda").manual_seed(M * N)
    X = torch.randn((M, N), device="cuda", generator=g)
    base = torch.randn(N, device="cuda", generator=g)
    want = base.double() + X.double().sum(0)
    L = mlp._bind()
    ws = torch.empty(1024 * N, device="cuda")
    outs = []
    for _ in range(2):
        out = base.clone()
        _lib.check(L.pca_mlp_colsum_ordered(X.data_ptr(), N, M, N, out.data_ptr(), ws.data_ptr(),
                                            ws.numel(), torch.cuda.current_stream().cuda_stream),
                   "colsum_ordered")
        torch.cuda.synchronize()
        outs.append(out)
    assert (outs[0].double() - want).abs().max().item() <= 1e-5 * max(1.0, float(np.sqrt(M)))
    assert torch.equal(outs[0], outs[1])


def test_gemm_path_training_repeats(gpu):
    """Two trainings on the GEMM path (forced; the path of width-256 networks)
    from the same initial weights and data end with the same weight bits."""
    ws = []
    for _ in range(2):
        net = mlp.ivim_mlp(synth.B_C2, seed=11)
        data, truth = net.synthesize(1 << 18, seed=5, snr=(20.0, 100.0),
                                     lo=[0.8, 0.05, 5e-3, 0.5e-3], hi=[1.2, 0.4, 0.1, 3e-3])
        net.fit_standardization(truth, data)
        X, T = net.prepare(data, truth)
        net.train(X, T, epochs=1, batch=8192, fused=False)
        torch.cuda.synchronize()
        ws.append(net.w.clone())
    assert torch.equal(ws[0], ws[1])


def _small_net(seed=0):
    net = mlp.ivim_mlp(synth.B_C2, seed=seed)
    return net


def test_training_gradients_match_oracle(gpu):
    net = _small_net()
    rng = np.random.default_rng(0)
    B = 512
    X = rng.uniform(0.1, 1.2, (net.INP, B)).astype(np.float32)
    X[net.IN:] = 0
    T = rng.standard_normal((B, net.OUT)).astype(np.float32)
    Xt = torch.tensor(X, device="cuda")
    Tt = torch.tensor(T, device="cuda")
    net._buffers(B)["loss"].zero_()
    net.forward_backward(Xt, Tt, 0, B, B)
    torch.cuda.synchronize()
    w = {k: net._view(net.w, k).double().cpu().numpy() for k in net.offsets}
    loss, g, _ = omlp.forward_backward(X.T.astype(np.float64), T.astype(np.float64), w["w1"],
                                       w["b1"], w["w2"], w["b2"], w["w3"], w["b3"])
    assert abs(net._buf["loss"].item() - loss) / loss < 1e-4
    for k in net.offsets:
        got = net._view(net.g, k).double().cpu().numpy()
        rel = np.linalg.norm(got - g[k]) / np.linalg.norm(g[k])
        print(f"grad {k}: relative error {rel:.2e}", flush=True)
        assert rel < 2e-3, (k, rel)


def _fused_case(case):
    if case == "dki128":
        return mlp.dki_mlp(synth.dki(16, seed=1)["consts"], hidden=128, seed=3)
    return mlp.ivim_mlp(synth.B_C2, hidden={"ivim32": 32, "ivim64": 64, "ivim128": 128}[case],
                        seed=4)


@pytest.mark.parametrize("case", ["ivim32", "ivim64", "ivim128", "dki128"])
def test_fused_step_gradients_match_oracle(gpu, case):
    """pca_mlp_train_step + pca_mlp_grad_reduce (no Adam): loss and every
    parameter block's gradient against the float64 oracle with bf16 operands,
    and against the GEMM path; 640 batch rows (10 workgroups), inputs read from
    a column offset of a wider X. dki128: 61 inputs (64 padded), 22 outputs."""
    net = _fused_case(case)
    rng = np.random.default_rng(5)
    B, off, n = 640, 192, 1024
    X = rng.uniform(0.1, 1.2, (net.INP, n)).astype(np.float32)
    X[net.IN:] = 0
    T = rng.standard_normal((n, net.OUT)).astype(np.float32)
    Xt, Tt = torch.tensor(X, device="cuda"), torch.tensor(T, device="cuda")
    print(f"[fused step {case}] H={net.H} in={net.INP} out={net.OUT} B={B}", flush=True)
    net._fused_ws(B)["loss"].zero_()
    w0 = net.w.clone()
    net.fused_step(Xt, Tt, off, B, n)
    torch.cuda.synchronize()
    assert torch.equal(net.w, w0)                     # no Adam without lr
    fused_loss = net._fused_ws(B)["loss"].item()
    fused_g = net.g.clone()
    net._buffers(B)["loss"].zero_()
    net.forward_backward(Xt, Tt, off, B, n)
    torch.cuda.synchronize()
    w = {k: net._view(net.w, k).double().cpu().numpy() for k in net.offsets}
    loss, g, _ = omlp.forward_backward(X[:, off:off + B].T.astype(np.float64),
                                       T[off:off + B].astype(np.float64), w["w1"], w["b1"],
                                       w["w2"], w["b2"], w["w3"], w["b3"])
    assert abs(fused_loss - loss) / loss < 1e-4, (fused_loss, loss)
    for k in net.offsets:
        got = net._view(fused_g, k).double().cpu().numpy()
        gemm = net._view(net.g, k).double().cpu().numpy()
        rel = np.linalg.norm(got - g[k]) / np.linalg.norm(g[k])
        rel_gemm = np.linalg.norm(got - gemm) / np.linalg.norm(g[k])
        print(f"[fused step {case}] grad {k}: vs oracle {rel:.2e}, vs GEMM path {rel_gemm:.2e}",
              flush=True)
        assert rel < 2e-3, (k, rel)
        assert rel_gemm < 2e-3, (k, rel_gemm)
    pad = net._view(fused_g, "w1")[:, net.IN:]
    assert torch.all(pad == 0)                        # padded inputs carry no gradient


def test_fused_step_adam_equals_adam_kernel(gpu):
    """The Adam update inside pca_mlp_grad_reduce equals pca_mlp_adam applied
    to the same reduced gradient, bit for bit, over three steps; the reduction
    is deterministic (the same batch twice gives the same gradient bits)."""
    net = _fused_case("ivim64")
    ref = _fused_case("ivim64")
    rng = np.random.default_rng(6)
    B = 1024
    X = torch.tensor(rng.uniform(0.1, 1.2, (net.INP, B)).astype(np.float32), device="cuda")
    X[net.IN:] = 0
    T = torch.tensor(rng.standard_normal((B, net.OUT)).astype(np.float32), device="cuda")
    for it in range(3):
        lr = 1e-3 * (it + 1)
        ref.fused_step(X, T, 0, B, B)                 # gradient only
        g1 = ref.g.clone()
        ref.fused_step(X, T, 0, B, B)
        assert torch.equal(ref.g, g1)
        ref.adam_step(lr)
        net.fused_step(X, T, 0, B, B, lr=lr)
        torch.cuda.synchronize()
        assert torch.equal(net.g, ref.g)
        for k in ("w", "m", "v"):
            assert torch.equal(getattr(net, k), getattr(ref, k)), (it, k)


def test_fused_training_matches_gemm_training(gpu):
    """Two epochs on 1M forward-model signals: the fused step and the GEMM
    path reach the same loss (within 2%) from the same initial weights."""
    hist = {}
    for fused in (True, False):
        net = mlp.ivim_mlp(synth.B_C2, seed=7)
        data, truth = net.synthesize(1 << 20, seed=99, snr=(20.0, 100.0),
                                     lo=[0.8, 0.05, 5e-3, 0.5e-3], hi=[1.2, 0.4, 0.1, 3e-3])
        net.fit_standardization(truth, data)
        X, T = net.prepare(data, truth)
        hist[fused] = net.train(X, T, epochs=2, batch=8192, fused=fused)
    print(f"loss history fused {hist[True]}, GEMM path {hist[False]}", flush=True)
    assert hist[True][-1] < hist[True][0]
    assert abs(hist[True][-1] - hist[False][-1]) / hist[False][-1] < 0.02


def test_fused_inference_matches_oracle_forward(gpu):
    net = _small_net(seed=1)
    w = synth.ivim_c2(3000, seed=9)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    net.mu.copy_(torch.tensor([0.05, 0.2, -3.5, -6.5]))
    net.sd.copy_(torch.tensor([0.1, 0.1, 0.5, 0.5]))
    net.finalize()
    params = torch.zeros((4, prob.N), device="cuda")
    lb = torch.full_like(params, -1e30)
    ub = torch.full_like(params, 1e30)
    net.infer(prob.data_t, lb, ub, params)
    torch.cuda.synchronize()
    y = w["data"].astype(np.float64)
    s = y[0]
    X = np.zeros((net.INP, prob.N))
    X[:net.IN] = y / s
    wt = {k: net._view(net.w, k).double().cpu().numpy() for k in net.offsets}
    out = omlp.forward(X.T, wt["w1"], wt["b1"], wt["w2"], wt["b2"], wt["w3"], wt["b3"],
                       round_bf16=False)
    mu, sd = net.mu.double().cpu().numpy(), net.sd.double().cpu().numpy()
    p = mu + sd * out
    p[:, 2:] = np.exp(p[:, 2:])
    p[:, 0] *= s
    got = params.double().cpu().numpy().T
    rel = np.abs(got - p) / np.abs(p)
    print("inference rel err p99", np.percentile(rel, 99), "max", rel.max(), flush=True)
    assert np.percentile(rel, 99) < 2e-2 and np.median(rel) < 2e-3


def _plain_net(model, consts, hidden, seed):
    """A network without output transforms (mu 0, sd 1): inference returns
    the raw forward pass, clipped to [-1e30, 1e30]."""
    return mlp.MLPInit(model, consts, hidden=hidden, norm_index=0, seed=seed)


_FAKE6 = ModelSpec("fake6", "a0", ("a0", "a1", "a2", "a3", "a4", "a5"), ("b",), (0,) * 6,
                   (1,) * 6)


@pytest.mark.parametrize("case", ["ivim32", "ivim64", "ivim128", "ivim256", "dki128", "dki256",
                                  "six128"])
def test_inference_variants_match_oracle_forward(gpu, case):
    """Every instantiation family of the fused inference kernel against the
    float64 forward pass of the same weights: LDS-staged weights with the
    output count fixed (IVIM, H=128) or at run time (six outputs), and
    weights from global memory (H=256; 61 inputs)."""
    if case.startswith("dki"):
        w = synth.dki(2000, seed=3)
        model = fit.MODELS["dki"]
    else:
        w = synth.ivim_c2(2000, seed=9)
        model = _FAKE6 if case == "six128" else IVIM
    hidden = int(case[-3:]) if case[-3:].isdigit() else int(case[-2:])
    net = _plain_net(model, w["consts"], hidden, seed=4)
    net.finalize()
    y = np.asarray(w["data"], np.float64)
    N = y.shape[1]
    data_t = torch.tensor(w["data"], dtype=torch.float32, device="cuda")
    out = torch.zeros((net.OUT, N), device="cuda")
    big = torch.full_like(out, 1e30)
    net.infer(data_t, -big, big, out)
    torch.cuda.synchronize()
    X = np.zeros((net.INP, N))
    X[:net.IN] = y / y[0]
    wt = {k: net._view(net.w, k).double().cpu().numpy() for k in net.offsets}
    ref = omlp.forward(X.T, wt["w1"], wt["b1"], wt["w2"], wt["b2"], wt["w3"], wt["b3"],
                       round_bf16=False)
    got = out.double().cpu().numpy().T
    scale = np.abs(ref) + 0.1 * np.sqrt(np.mean(ref ** 2))
    rel = np.abs(got - ref) / scale
    # yardstick: the same forward pass with bf16-rounded operands and the
    # exact GELU (the oracle's model of bf16 MFMA inference)
    bf = omlp.forward(X.T, wt["w1"], wt["b1"], wt["w2"], wt["b2"], wt["w3"], wt["b3"])
    rel_bf = np.abs(bf - ref) / scale
    print(case, "inference err p99", np.percentile(rel, 99), "median", np.median(rel),
          "| bf16 operands p99", np.percentile(rel_bf, 99), "median", np.median(rel_bf),
          flush=True)
    assert np.percentile(rel, 99) <= np.percentile(rel_bf, 99)
    assert np.median(rel) <= np.median(rel_bf) and np.percentile(rel, 99) < 5e-2


def test_inference_fallback_is_the_default_guess(gpu):
    """Voxels whose network outputs are not finite get the log-linear default
    guess from the same kernel, identical to pca_loglinear_init; without a
    fallback they keep what params held."""
    net = _small_net(seed=1)
    w = synth.ivim_c2(3000, seed=9)
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    net.finalize()
    ref = fit.default_init(prob)
    # every output non-finite
    net.sd.fill_(float("inf"))
    p = torch.full((4, prob.N), -7.0, device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, p, fallback=(prob.model, prob.consts))
    q = torch.full((4, prob.N), -7.0, device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, q)
    torch.cuda.synchronize()
    assert torch.equal(p, ref)
    assert bool((q == -7.0).all())
    # only the voxels with a NaN sample: the rest keep the network's guess
    net.sd.fill_(1.0)
    data = prob.data_t.clone()
    bad = torch.arange(0, prob.N, 7, device="cuda")
    data[3, bad] = float("nan")
    r = torch.full((4, prob.N), -7.0, device="cuda")
    net.infer(data, prob.lb_t, prob.ub_t, r, fallback=(prob.model, prob.consts))
    good = torch.full((4, prob.N), -7.0, device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, good)
    dref = fit._init.default_init(prob.model, prob.consts, data, prob.lb_t, prob.ub_t)
    torch.cuda.synchronize()
    mask = torch.zeros(prob.N, dtype=torch.bool, device="cuda")
    mask[bad] = True
    assert torch.equal(r[:, mask], dref[:, mask])
    assert torch.equal(r[:, ~mask], good[:, ~mask])
    with pytest.raises(Exception):
        net.infer(prob.data_t, prob.lb_t, prob.ub_t, r, fallback=(T2, [[10.0] * 16]))


def test_synth_matches_model(gpu):
    from pycompute_amd.lsqnonlin import Expr, model_eval
    net = _small_net()
    truth = torch.tensor(synth.ivim_truth(4096, np.random.default_rng(3)), dtype=torch.float32,
                         device="cuda")
    data, tru = net.synthesize(4096, seed=5, snr=(1e12, 1e12), truth=truth)
    m, _, _ = model_eval(Expr(IVIM.expr, IVIM.pars, IVIM.consts), truth,
                         torch.tensor(synth.B_C2, dtype=torch.float32, device="cuda"), 16,
                         want=("m",))
    torch.cuda.synchronize()
    assert torch.equal(tru, truth)
    np.testing.assert_allclose(data.cpu().numpy(), m.cpu().numpy(), rtol=1e-5)
    # sampled parameters stay inside the box; noise has the requested level
    lo, hi = [0.8, 0.05, 5e-3, 0.5e-3], [1.2, 0.4, 0.1, 3e-3]
    d2, t2 = net.synthesize(200000, seed=6, snr=(50.0, 50.0), lo=lo, hi=hi)
    t2 = t2.cpu().numpy()
    for j in range(4):
        assert t2[j].min() >= lo[j] and t2[j].max() <= hi[j]
    m2, _, _ = model_eval(Expr(IVIM.expr, IVIM.pars, IVIM.consts), torch.tensor(t2, device="cuda"),
                          torch.tensor(synth.B_C2, dtype=torch.float32, device="cuda"), 16,
                          want=("m",))
    # at b = 800 the signal is well above sigma: residual std ~ sigma = S0 / 50
    r = (d2[-1] - m2[-1]) / torch.tensor(t2[0], device="cuda") * 50.0
    assert abs(float(r.std()) - 1.0) < 0.05


def test_trained_network_gains_over_default_init(gpu):
    """Trained on 2M forward-model signals, the MLP guess lets LM reach the
    total final SSR of plain LM (default guess, 100 iterations) with far fewer
    iterations."""
    net = _small_net(seed=2)
    lo, hi = [0.8, 0.05, 5e-3, 0.5e-3], [1.2, 0.4, 0.1, 3e-3]
    data, truth = net.synthesize(1 << 21, seed=1234, snr=(20.0, 100.0), lo=lo, hi=hi)
    net.fit_standardization(truth, data)
    X, T = net.prepare(data, truth)
    hist = net.train(X, T, epochs=2, batch=8192)
    print("loss history", hist, flush=True)
    # f and D_1 are poorly determined by noisy signals, so the standardised
    # loss stays well above 0; what matters is the LM gain below
    assert hist[-1] < hist[0]
    w = synth.ivim(65536, synth.B_C2, seed=77, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    plain = fit.fit(prob, iters=100)
    f_plain = plain.final_f().double().sum().item()
    it_plain = plain.n_iter_t.double().mean().item()
    p0 = fit.default_init(prob)
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, p0)
    res = {}
    for cap in (10, 15, 20):
        s = fit.fit(prob, iters=cap, params0=p0.clone())
        res[cap] = (s.final_f().double().sum().item() / f_plain - 1.0,
                    s.n_iter_t.double().mean().item())
    torch.cuda.synchronize()
    print(f"plain: total SSR {f_plain:.4f}, {it_plain:.1f} iterations/voxel; mlp: {res}", flush=True)
    best = [c for c, (d, it) in res.items() if d <= 1e-4]
    assert best, res
    assert res[min(best)][1] * 2.0 < it_plain


def test_mlp_initialised_graph_replays_equal_eager(gpu):
    """The headline pipeline -- state reset, inference with the fused
    fallback, capped chunked LM -- captured as one hipGraph and replayed
    several times with eager work in between equals the eager sequence on
    every replay."""
    net = _small_net(seed=3)
    w = synth.ivim(16384, synth.B_C2, seed=5, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    net.mu.copy_(torch.tensor([0.05, 0.2, -3.5, -6.5]))
    net.sd.copy_(torch.tensor([0.1, 0.1, 0.5, 0.5]))
    net.finalize()
    fb = (prob.model, prob.consts)
    p0 = torch.empty((4, prob.N), device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, p0, fallback=fb)
    ref = fit.fit(prob, iters=6, params0=p0.clone(), compaction=True, chunk=24)
    torch.cuda.synchronize()
    pars = torch.empty_like(p0)
    s = fit.make_solver(prob, compaction=True, chunk=24)
    s.setup(pars, prob.consts_t, prob.data_t, prob.lb_t, prob.ub_t)
    s.capture_graph(6, init=lambda st: net.infer(prob.data_t, prob.lb_t, prob.ub_t, pars,
                                                 stream=st, fallback=fb))
    for rep in range(3):
        pars.fill_(float("nan"))
        s.lam_t.fill_(123.0)
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        s.replay_graph()
        torch.cuda.synchronize()
        assert bool(torch.isfinite(junk))
        for k in ("pars_t", "step_type_t", "lam_t", "f_t", "n_iter_t"):
            assert torch.equal(getattr(ref, k), getattr(s, k)), (rep, k)


IVIM_BOX = ([0.8, 0.05, 5e-3, 0.5e-3], [1.2, 0.4, 0.1, 3e-3])
FASTEST = dict(fast_exp=True, fast_div=True, reassoc=True, factored=True, waves_per_eu=2,
               chunk=16)


def _trained_ivim_net(samples=10_000_000, seed=0):
    """The headline initialiser: 16 -> 64 -> 64 -> 4, trained on forward-model
    signals of a separate seed (the benchmark's recipe)."""
    net = mlp.ivim_mlp(synth.B_C2, seed=seed)
    data, truth = net.synthesize(samples, seed=20240613, snr=(20.0, 100.0), lo=IVIM_BOX[0],
                                 hi=IVIM_BOX[1])
    net.fit_standardization(truth, data)
    X, T = net.prepare(data, truth)
    del data, truth
    net.train(X, T, epochs=2, batch=8192)
    return net


def test_mlp_guess_lm_matches_oracle(gpu):
    """C4 leg: LM refined from the network's guess (the fastest build, the
    headline cap of 6 iterations) against the oracle run from the same guess,
    on 65,536 voxels, to the parity bars of tests/parity.py."""
    net = _trained_ivim_net(2_000_000)
    w = synth.ivim(65536, synth.B_C2, seed=61, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    p0 = torch.empty((4, prob.N), device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, p0, fallback=(prob.model, prob.consts))
    s = fit.fit(prob, iters=6, params0=p0.clone(), **FASTEST)
    torch.cuda.synchronize()
    y = prob.data_t.cpu().numpy()
    r32, r64 = parity.oracle_pair("ivim", prob.consts, y, prob.lb_t.cpu().numpy(),
                                  prob.ub_t.cpu().numpy(), p0.cpu().numpy(), 6)
    parity.check("C4 MLP guess + 6 LM iterations", "ivim", prob.consts, y,
                 s.pars_t.cpu().numpy(), r32, r64, gpu_first_f=s.first_f.cpu().numpy()[0])


def test_mlp_pipeline_4m_matched_residual(gpu):
    """The C4 pipeline at full size: the network trained on 10M signals, the
    iteration cap chosen on a held-out validation volume, then the
    4,194,304-voxel volume fitted from the MLP guess at that cap reaches a
    total final SSR within 1e-4 of plain LM's (default guess, 100
    iterations), in far fewer iterations; the per-voxel ratio of the two
    residuals is printed."""
    net = _trained_ivim_net()
    fb = lambda pr: (pr.model, pr.consts)     # noqa: E731
    wv = synth.ivim(1 << 18, synth.B_C2, seed=99, snr=(20.0, 100.0))
    pv = fit.make_problem("ivim", wv["consts"], wv["data"])
    ref_v = float(fit.fit(pv, iters=100, **FASTEST).final_f().double().sum())
    gv = torch.empty((4, pv.N), device="cuda")
    net.infer(pv.data_t, pv.lb_t, pv.ub_t, gv, fallback=fb(pv))
    cap = None
    for c in (2, 3, 4, 5, 6, 8, 10, 12, 15, 20, 25, 30):
        if float(fit.fit(pv, iters=c, params0=gv.clone(), **FASTEST).final_f().double().sum()) \
                <= ref_v * (1 + 1e-4):
            cap = c
            break
    assert cap is not None
    del pv, gv
    w = synth.ivim(1 << 22, synth.B_C2, seed=4, snr=(20.0, 100.0))
    prob = fit.make_problem("ivim", w["consts"], w["data"])
    del w
    plain = fit.fit(prob, iters=100, **FASTEST)
    f_plain = plain.final_f().double()[0]
    it_plain = float(plain.n_iter_t.double().mean())
    del plain
    g = torch.empty((4, prob.N), device="cuda")
    net.infer(prob.data_t, prob.lb_t, prob.ub_t, g, fallback=fb(prob))
    s = fit.fit(prob, iters=cap, params0=g, **FASTEST)
    f_mlp = s.final_f().double()[0]
    it_mlp = float(s.n_iter_t.double().mean())
    tot = float(f_mlp.sum()) / float(f_plain.sum()) - 1.0
    ratio = (f_mlp / f_plain).cpu().numpy()
    q = np.percentile(ratio, [0.1, 1, 50, 99, 99.9])
    print(f"4M: cap {cap}, total SSR {tot:+.3e} vs plain, iterations {it_mlp:.2f} vs "
          f"{it_plain:.2f}; per-voxel SSR_mlp/SSR_plain p0.1 p1 p50 p99 p99.9 = {q}",
          flush=True)
    assert tot <= 1e-4
    assert it_mlp * 4 < it_plain
    assert bool(torch.isfinite(s.pars_t).all())
    assert bool((s.pars_t >= prob.lb_t).all()) and bool((s.pars_t <= prob.ub_t).all())
