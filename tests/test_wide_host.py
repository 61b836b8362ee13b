"""Host side of the wide-model kernel: GLM structure detection and the
coefficient table (no GPU)."""
import numpy as np
import pytest

from pycompute_amd import synth, wide
from pycompute_amd.models import DKI, IVIM, SE, T2, dki_design
from pycompute_amd.sym import glm, hipgen


def test_dki_is_exp_glm_and_table_is_the_design_matrix():
    dag = hipgen.res_jac_hes(DKI.expr, DKI.pars, DKI.consts)
    form = glm.glm_form(dag)
    assert form is not None and form.phi == "exp" and len(form.coef) == 22
    consts = synth.dki_protocol()
    tab = glm.coefficient_table(form, dag, consts)
    n = 22
    assert tab.shape == (61, glm.table_stride(n)) and glm.table_stride(n) == 48
    # the closed-form design matrix of log S (models.dki_design)
    np.testing.assert_allclose(tab[:, :n], dki_design(consts), rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(tab[:, n:2 * n], tab[:, :n] ** 2)
    np.testing.assert_array_equal(tab[:, 2 * n], 0.0)
    np.testing.assert_array_equal(tab[:, 2 * n + 1:], 0.0)


@pytest.mark.parametrize("m", [IVIM, T2, SE])
def test_nonlinear_models_are_not_glm(m):
    assert glm.glm_form(hipgen.res_jac_hes(m.expr, m.pars, m.consts)) is None


def test_identity_glm_with_offset():
    expr = "3*t + a + b*cos(t) + c*sin(2*t)"
    dag = hipgen.res_jac_hes(expr, ["a", "b", "c"], ["t"])
    form = glm.glm_form(dag)
    assert form.phi == "identity"
    t = np.linspace(0, 6, 7)
    tab = glm.coefficient_table(form, dag, t[None, :])
    np.testing.assert_allclose(tab[:, 0], 1.0)
    np.testing.assert_allclose(tab[:, 1], np.cos(t), rtol=1e-12)
    np.testing.assert_allclose(tab[:, 2], np.sin(2 * t), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(tab[:, 6], 3 * t, rtol=1e-12)       # offset a0 = eta(p = 0)


def test_exp_of_nonlinear_exponent_is_rejected():
    dag = hipgen.res_jac_hes("exp(a*b*t)", ["a", "b"], ["t"])
    assert glm.glm_form(dag) is None


@pytest.mark.parametrize("n,tile", [(22, 136), (22, 100), (22, 170), (22, 40), (22, 10),
                                    (11, 136), (9, 1)])
def test_tile_rows_partition_the_packed_triangle(n, tile):
    """Tiles are consecutive runs of whole rows, each as long as fits in
    tile entries (a row longer than tile gets a tile of its own), covering
    rows 0..n-1; the parked entries are those of all tiles but the last."""
    from pycompute_amd import wide
    rows = wide.tile_rows(n, tile)
    assert rows[0] == 0 and rows[-1] == n and all(a < b for a, b in zip(rows, rows[1:]))
    tri = lambda r: r * (r + 1) // 2
    for r0, r1 in zip(rows, rows[1:]):
        size = tri(r1) - tri(r0)
        assert size <= tile or r1 == r0 + 1
        if r1 < n:   # maximal: the next row would not fit
            assert size + r1 + 1 > tile
    assert wide.park_entries(n, tile) == tri(rows[-2])
    if tile >= tri(n):
        assert rows == [0, n] and wide.park_entries(n, tile) == 0


def test_wide_kernel_lds_budget():
    """The GLM kernel's LDS park holds every tile but the last, one column per
    lane; the general wide kernel keeps the whole packed Hessian. Both fit
    the 160 KB a gfx950 workgroup may allocate for the 22-parameter model in
    FP32 and FP64; past the limit the programs decline."""
    from pycompute_amd import wide
    for tile in (40, 100, 136, 170):
        assert wide.lds_bytes(22, tile, "float32") <= wide.LDS_LIMIT
        assert wide.lds_bytes(22, tile, "float64") <= wide.LDS_LIMIT
    assert wide.LDS_LIMIT == 160 * 1024
    assert wide.lds_bytes(22, 136, "float32") == 136 * 64 * 4
    assert wide.wide_lds_bytes(22, "float32") == 253 * 64 * 4
    assert wide.wide_lds_bytes(22, "float64") == 253 * 64 * 8
    # 36 parameters: 666 entries, 170 KB in FP32, over the limit
    assert wide.wide_lds_bytes(36, "float32") > wide.LDS_LIMIT
    assert wide.lds_bytes(60, 40, "float32") > wide.LDS_LIMIT


def test_mfma_entry_product_table_layout():
    """pca_lm_glm_mfma's table: t2[i, r, q] = A_ij A_ik of packed entry
    e = 4q + r = j(j+1)/2 + k, zero past the last entry, rows of
    RS = (groups rounded up to 4) + 4 floats; checked against a direct loop
    over (j, k) on the kurtosis design."""
    from pycompute_amd import wide
    consts = synth.dki_protocol()
    A = dki_design(consts)
    M, n = A.shape
    tab = np.zeros((M, glm.table_stride(n)))
    tab[:, :n] = A
    t2 = wide.entry_products(tab, n)
    nh = n * (n + 1) // 2
    nq = (nh + 3) // 4
    assert t2.shape == (M, 4, (nq + 3) // 4 * 4 + 4) == (61, 4, 68)
    for j in range(n):
        for k in range(j + 1):
            e = j * (j + 1) // 2 + k
            np.testing.assert_array_equal(t2[:, e % 4, e // 4], A[:, j] * A[:, k])
    mask = np.ones(t2.shape, bool)
    for e in range(nh):
        mask[:, e % 4, e // 4] = False
    assert np.all(t2[mask] == 0)
    # LDS of a 256-thread workgroup: the table, the samples and g over the solve
    assert wide.mfma_lds_bytes(22, 61) == 61 * 4 * 68 * 4 + 4 * 61 * 64 * 4 + 4 * 64 * 23 * 4
    assert wide.mfma_lds_bytes(22, 61) <= wide.LDS_LIMIT < wide.mfma_lds_bytes(22, 80)


def test_quad_table_layout():
    """pca_lm_glm_quad's table for L lanes per voxel: [i, r, t] is column
    perm(L t + r) of the coefficient table (the reference's fixed pivot
    rotation), 0 past the last parameter, a0 in slot ceil(n / L) of every
    lane, zero padding to a multiple of 4 floats."""
    rng = np.random.default_rng(3)
    for L in (2, 4):
        for n in (9, 11, 22, 24):
            tab = rng.standard_normal((5, glm.table_stride(n)))
            q = wide.quad_table(tab, n, L)
            nt = -(-n // L)
            assert q.shape == (5, L, wide.quad_slots(n, L)) and q.shape[2] % 4 == 0
            assert q.shape[2] >= nt + 1
            for r in range(L):
                for t in range(q.shape[2]):
                    a = L * t + r
                    if t < nt:
                        want = tab[:, n - 1 if a == 0 else a - 1] if a < n else 0.0
                    else:
                        want = tab[:, 2 * n] if t == nt else 0.0
                    np.testing.assert_array_equal(q[:, r, t], want)
    with pytest.raises(ValueError):
        wide.quad_table(np.zeros((5, glm.table_stride(25))), 25)
    for bad in (2, 3):
        with pytest.raises(ValueError):
            wide.quad_lanes(bad)
    assert wide.quad_lanes(True) == 4 and wide.quad_lanes(False) == 0
