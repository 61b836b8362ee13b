"""Host-side logic of the drop-in API (no GPU): constructor forms, consts
layouts, dtype handling, init matrices, synthetic workloads."""
import numpy as np
import pytest
import torch

import oracle
from pycompute_amd import init, lm, synth
from pycompute_amd.lsqnonlin import Expr, F, SecondOrderLevenbergMarquardt, split_consts
from pycompute_amd.models import IVIM, MODELS, T2


def test_constructor_forms():
    a = SecondOrderLevenbergMarquardt(IVIM.expr, list(IVIM.pars), list(IVIM.consts), ndata=21,
                                      dtype=np.float32, write_to_file=False)
    b = SecondOrderLevenbergMarquardt(Expr(IVIM.expr, list(IVIM.pars), list(IVIM.consts), None,
                                           None), None, 21, torch.float32)
    assert a.funcid == b.funcid == "second_order_levenberg_marquardt_4_1_21_f_f7def86f0a03"
    assert a.nparam == 4 and a.nhes == 10 and a.nconst == 1 and a.ndata == 21
    assert a.code == b.code
    c = SecondOrderLevenbergMarquardt(model_expr=Expr(T2.expr, T2.pars, T2.consts), f_gradf=None,
                                      ndata=32, dtype="float64")
    assert c.dtype == torch.float64 and c.funcid.endswith("_d_" + c.funcid.split("_")[-1])


def test_constructor_errors():
    with pytest.raises(RuntimeError):
        SecondOrderLevenbergMarquardt(None, None, 10, torch.float32)
    with pytest.raises(RuntimeError):
        SecondOrderLevenbergMarquardt(IVIM.expr, list(IVIM.pars), list(IVIM.consts), 10,
                                      torch.float16)
    with pytest.raises(TypeError):
        SecondOrderLevenbergMarquardt(IVIM.expr, list(IVIM.pars), list(IVIM.consts), 10,
                                      ndata=10)


def test_f_constructor_forms():
    a = F(IVIM.expr, list(IVIM.pars), list(IVIM.consts), 21, torch.float32)
    b = F(Expr(IVIM.expr, IVIM.pars, IVIM.consts), 21, torch.float32)
    assert a.funcid == b.funcid == "f_4_1_21_f_f7def86f0a03"


def test_setup_rejects_host_tensors():
    s = SecondOrderLevenbergMarquardt(IVIM.expr, list(IVIM.pars), list(IVIM.consts), 10)
    z = torch.zeros((4, 8))
    with pytest.raises(ValueError):
        s.setup(z, torch.zeros(10), torch.zeros((10, 8)), z, z)


def test_split_consts_layouts():
    M, N = 5, 7
    b = torch.arange(M, dtype=torch.float32)
    uc, pv = split_consts(b, 1, M, N)
    assert pv is None and uc == b.tolist()
    uc, pv = split_consts(b.reshape(1, M, 1).expand(1, M, N), 1, M, N)
    assert pv is None and uc == b.tolist()
    tiled = b.reshape(M, 1).repeat(1, N)
    uc, pv = split_consts(tiled, 1, M, N)            # replicated columns detected
    assert pv is None and uc == b.tolist()
    uc, pv = split_consts(tiled, 1, M, N, uniform=False)
    assert uc is None and pv is tiled
    varying = tiled + torch.arange(N)[None, :]
    uc, pv = split_consts(varying, 1, M, N)
    assert uc is None and pv is varying
    with pytest.raises(ValueError):
        split_consts(torch.zeros(M + 1), 1, M, N)
    uc, pv = split_consts(torch.zeros((4, M)), 4, M, N)  # (nconst, M)
    assert pv is None and len(uc) == 4 * M


def test_dtype_names():
    assert lm.dtype_name(torch.float32) == "float32"
    assert lm.dtype_name(np.float64) == "float64"
    assert lm.dtype_name(np.dtype("float32")) == "float32"
    assert lm.dtype_name("float") == "float32"
    with pytest.raises(RuntimeError):
        lm.dtype_name(torch.int32)


def test_ivim_split():
    lo, hi = init.ivim_split(np.arange(21) * 50.0)     # 21-point protocol: literal index split
    assert lo.tolist() == list(range(10)) and hi.tolist() == list(range(15, 21))
    lo, hi = init.ivim_split(synth.B_C2)
    assert synth.B_C2[lo].max() <= 140 and synth.B_C2[hi].min() >= 500
    lo, hi = init.ivim_split(synth.B_C1)
    assert lo.tolist() == [0, 1, 2, 3, 4] and hi.tolist() == [7, 8, 9]


@pytest.mark.parametrize("name", ["ivim", "t2", "se", "dki"])
def test_init_matrix_is_least_squares(name):
    """P log y reproduces the least-squares fits of the default initial guess."""
    m = MODELS[name]
    w = {"ivim": lambda: synth.ivim_c1(64, 1), "t2": lambda: synth.t2(64, 3),
         "se": lambda: synth.se(64, 4), "dki": lambda: synth.dki(64, 6)}[name]()
    P = init.init_matrix(m, w["consts"])
    ly = np.log(w["data"].astype(np.float64))
    c = P @ ly
    if name == "ivim":
        lo, hi = init.ivim_split(w["consts"][0])
        b = w["consts"][0]
        np.testing.assert_allclose(c[0:2], oracle.bpolyfit(-b[lo], ly[lo], 1), rtol=1e-9,
                                   atol=1e-12)
        np.testing.assert_allclose(c[2:4], oracle.bpolyfit(-b[hi], ly[hi], 1), rtol=1e-9,
                                   atol=1e-12)
    elif name in ("t2", "se"):
        np.testing.assert_allclose(c, oracle.bpolyfit(-w["consts"][0], ly, 1), rtol=1e-9,
                                   atol=1e-12)
    else:
        from pycompute_amd.models import dki_design
        A = dki_design(w["consts"])
        np.testing.assert_allclose(c, np.linalg.lstsq(A, ly, rcond=None)[0], rtol=1e-6,
                                   atol=1e-9)


def test_dki_noise_free_log_linear_init_is_exact():
    from pycompute_amd.models import dki_design
    w = synth.dki(32, seed=6)
    A = dki_design(w["consts"])
    y = np.exp(A @ w["truth"])
    P = init.init_matrix(MODELS["dki"], w["consts"])
    np.testing.assert_allclose(P @ np.log(y), w["truth"], rtol=1e-6, atol=1e-12)


def test_synthetic_workloads_are_seeded():
    a, b = synth.ivim_c2(1000, seed=2), synth.ivim_c2(1000, seed=2)
    np.testing.assert_array_equal(a["data"], b["data"])
    assert a["data"].shape == (16, 1000) and a["data"].dtype == np.float32
    t = a["truth"]
    assert t[0].min() >= 0.8 and t[0].max() <= 1.2 and t[1].min() >= 0.05 and t[1].max() <= 0.4
    assert t[2].min() >= 5e-3 and t[3].max() <= 3e-3
    d = synth.dki(100, seed=6)
    assert d["data"].shape == (61, 100) and d["consts"].shape == (4, 61)
    g = d["consts"][1:, 1:]
    np.testing.assert_allclose(np.linalg.norm(g, axis=0), 1.0, rtol=1e-12)


def test_packed_matrix_helpers():
    from pycompute_amd.program import compact_to_LD, compact_to_full
    rng = np.random.default_rng(3)
    A = rng.standard_normal((5, 5))
    A = A + A.T
    packed = np.array([A[i, j] for i in range(5) for j in range(i + 1)])
    np.testing.assert_array_equal(compact_to_full(packed), A)
    L, D = compact_to_LD(packed)
    assert np.allclose(np.diag(L), 1.0) and np.allclose(np.triu(L, 1), 0.0)
    np.testing.assert_array_equal(np.diag(D), np.diag(A))
    np.testing.assert_array_equal(np.tril(L, -1), np.tril(A, -1))


def test_code_gen_walking_links_post_order_once():
    from pycompute_amd.program import HipFunction, code_gen_walking

    class Fn(HipFunction):
        def __init__(self, fid, deps=()):
            self.fid, self.deps = fid, list(deps)

        def get_device_funcid(self):
            return self.fid

        def get_device_code(self):
            return f"/*{self.fid}*/"

        def get_deps(self):
            return self.deps

    a = Fn("a")
    b = Fn("b", [a])
    c = Fn("c", [a, b])
    code = code_gen_walking(Fn("top", [b, c]))
    order = [s for s in code.split() if s.startswith("/*")]
    assert order == ["/*a*/", "/*b*/", "/*c*/", "/*top*/"]


def test_autotune_pick_ties_and_agreement():
    from pycompute_amd import autotune
    r = [{"knobs": {"k": 0}, "ms": 4.0, "agree": 1.0},
         {"knobs": {"k": 1}, "ms": 2.010, "agree": 0.99},
         {"knobs": {"k": 2}, "ms": 2.000, "agree": 0.99},
         {"knobs": {"k": 3}, "ms": 1.0, "agree": 0.5}]     # fast but disagrees
    assert autotune.pick(r)["knobs"] == {"k": 1}           # 0.5% slower: a tie, earlier wins
    r[1]["ms"] = 2.05
    assert autotune.pick(r)["knobs"] == {"k": 2}
    # agreement just under 98%, or a one-sided residual tail, disqualifies
    r[2]["agree"] = 0.975
    assert autotune.pick(r)["knobs"] == {"k": 1}
    r[2]["agree"] = 0.99
    r[2].update(worse=0.004, better=0.001)              # 0.4% worse vs 0.1% better
    assert autotune.pick(r)["knobs"] == {"k": 1}
    r[2].update(worse=0.0025, better=0.001)             # within 2x + 0.1%
    assert autotune.pick(r)["knobs"] == {"k": 2}


def test_chunk_schedule():
    from pycompute_amd.chunked import chunk_schedule
    assert chunk_schedule(100, 16) == [16] * 6 + [4]
    assert chunk_schedule(100, 16, 4) == [16, 16, 16, 52]
    assert chunk_schedule(30, 16, 4) == [16, 14]
    assert chunk_schedule(6, 16, 4) == [6]
    assert chunk_schedule(0, 16, 4) == []
    for it in range(1, 120):
        for k in (1, 5, 16):
            for m in (0, 1, 2, 4):
                s = chunk_schedule(it, k, m)
                assert sum(s) == it and all(x >= 1 for x in s)
                assert m == 0 or len(s) <= m


def test_autotune_skips_candidates_that_do_not_apply(monkeypatch):
    """tune() skips a knob set whose solver cannot be bound to the sample
    (ValueError at construction or setup, as the exponential recurrence on a
    protocol that is not an arithmetic progression raises), re-tunes when the
    cached choice no longer applies, and re-raises for the reference
    candidate. CUDA timing and the cache file are replaced by stand-ins."""
    import types
    import torch
    from pycompute_amd import autotune

    class Ev:
        t = 0.0

        def __init__(self, enable_timing=True):
            pass

        def record(self):
            Ev.t += 1.0
            self.at = Ev.t

        def elapsed_time(self, other):
            return other.at - self.at

    monkeypatch.setattr(torch.cuda, "get_device_properties",
                        lambda d: types.SimpleNamespace(gcnArchName="gfx950"))
    monkeypatch.setattr(torch.cuda, "current_device", lambda: 0)
    monkeypatch.setattr(torch.cuda, "Event", Ev)
    monkeypatch.setattr(torch.cuda, "synchronize", lambda: None)
    db = {}
    monkeypatch.setattr(autotune, "_load", lambda: db)
    monkeypatch.setattr(autotune, "_store", lambda d: None)

    class Solver:
        def __init__(self, **kn):
            self.kn = kn

        def run(self, iters):
            Ev.t += 0.5 if self.kn.get("exp_rec") else 2.0   # the recurrence "is faster"

        def final_f(self):
            return torch.ones((1, 8))

    def make(**kn):
        return Solver(**kn)

    def setup(s):
        if s.kn.get("exp_rec") and not uniform[0]:
            raise ValueError("not an arithmetic progression")

    cands = [{"exp_rec": 0, "k": 0}, {"exp_rec": 4, "k": 1}, {"exp_rec": 0, "k": 2}]
    uniform = [True]
    best = autotune.tune(make, setup, "x", 4, "float32", cands=cands, reps=1)
    assert best["exp_rec"] == 4
    uniform[0] = False      # same key, a protocol the cached choice does not fit
    best = autotune.tune(make, setup, "x", 4, "float32", cands=cands, reps=1)
    assert best["exp_rec"] == 0
    with pytest.raises(ValueError):
        autotune.tune(make, setup, "y", 4, "float32", cands=[{"exp_rec": 4}], reps=1)
