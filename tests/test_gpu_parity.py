"""GPU parity: the HIP path, called through the C-ABI, against the CPU oracle and the
reference's golden values.

Tolerances (FP64 path):
  AO / MO VGL       max |gpu - ref| <= 1e-12 * max|ref| per component (summation order only)
  Jastrow factors   |gpu - ref| <= 1e-12 * max(1, |ref|)
  Jastrow GL        max |gpu - ref| <= 1e-11 * max|ref| per component
Goldens are checked with the reference's own absolute tolerances.
"""
import ctypes

import numpy as np
import pytest

import oracle
import qmckl_amd as q
import systems
from qmckl_amd import system as S

pytestmark = pytest.mark.gpu


def make_ctx(s, x, W, jastrow=True, mo=True):
    ctx = q.context_create()
    S.load(ctx, s, jastrow=jastrow, mo=mo)
    S.set_coords(ctx, s, x, W)
    return ctx


def vgl_close(gpu, ref, rel, what):
    for c in range(ref.shape[1]):
        scale = max(np.abs(ref[:, c]).max(), 1e-300)
        err = np.abs(gpu[:, c] - ref[:, c]).max()
        assert err <= rel * scale, f"{what} component {c}: max abs err {err:.3e}, scale {scale:.3e}"


def golden_check(var, arr, file):
    gs = systems.goldens(var, file)
    assert gs
    for g in gs:
        v = arr[tuple(g["idx"])]
        assert abs(v - g["value"]) < g["tol"], (var, g["line"], g["idx"], v, g["value"])


# ---------------------------------------------------------------- AO
def test_ao_vgl_chbrclf_goldens_and_oracle():
    s, x, W = systems.chbrclf()
    ctx = q.context_create()
    S.load(ctx, s, jastrow=False, mo=False)
    pts = np.ascontiguousarray(x[:, : s.elec_num])
    P = pts.shape[1]
    q.check(ctx, q.lib().qmckl_set_point(ctx, b"T", P, q.ptr(pts), pts.size))
    ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num).reshape(P, 5, s.ao_num)
    golden_check("ao_vgl", ao, "qmckl_ao.org")
    vgl_close(ao, oracle.ao_vgl(s, pts), 1e-12, "ao_vgl")
    val = q.get_array(ctx, "qmckl_get_ao_basis_ao_value", P * s.ao_num).reshape(P, s.ao_num)
    golden_check("ao_value", val, "qmckl_ao.org")
    np.testing.assert_array_equal(val, ao[:, 0, :])
    sh = q.get_array(ctx, "qmckl_get_ao_basis_shell_vgl", P * 5 * s.shell_num).reshape(P, 5, s.shell_num)
    golden_check("shell_vgl", sh, "qmckl_ao.org")
    pr = q.get_array(ctx, "qmckl_get_ao_basis_primitive_vgl", P * 5 * s.prim_num).reshape(P, 5, s.prim_num)
    golden_check("prim_vgl", pr, "qmckl_ao.org")
    vgl_close(pr, oracle.prim_vgl(s, pts), 1e-13, "prim_vgl")
    q.context_destroy(ctx)


@pytest.mark.parametrize("name,W", [("small", 256), ("medium", 8)])
def test_ao_vgl_synthetic(name, W):
    s, x, W = S.workload(name, walk_num=W)
    ctx = make_ctx(s, x, W, jastrow=False, mo=False)
    P = W * s.elec_num
    ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num).reshape(P, 5, s.ao_num)
    vgl_close(ao, oracle.ao_vgl(s, x), 1e-12, f"ao_vgl {name}")
    q.context_destroy(ctx)


# ---------------------------------------------------------------- MO
@pytest.mark.parametrize("name,W", [("chbrclf", 2), ("small", 256), ("medium", 8)])
def test_mo_vgl_fp64(name, W):
    s, x, W = systems.chbrclf() if name == "chbrclf" else S.workload(name, walk_num=W)
    ctx = make_ctx(s, x, W, jastrow=False)
    P = W * s.elec_num
    nmo = s.mo_num_kept
    mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo).reshape(P, 5, nmo)
    ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
    vgl_close(mo, ref, 1e-12, f"mo_vgl {name}")
    # invariant of org/qmckl_mo.org:3747: mo_value equals the value row of mo_vgl
    val = q.get_array(ctx, "qmckl_get_mo_basis_mo_value", P * nmo).reshape(P, nmo)
    np.testing.assert_array_equal(val, mo[:, 0, :])
    q.context_destroy(ctx)


@pytest.mark.parametrize("mode,bound", [(q.MO_FP64, 1e-12), (q.MO_FP32, 2e-5), (q.MO_BF16X3, 2e-6)])
def test_ao_mo_vgl_high_l(mode, bound):
    """s to h shells (the AO polynomial's power loop beyond f) with AO factors off 1 (folded
    into the fused kernel's C_t operands): the fused AO->MO product in each mode, then the
    AO getter, against the oracle."""
    s, x, W = systems.high_l()
    ctx = make_ctx(s, x, W, jastrow=False)
    q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, mode))
    P = W * s.elec_num
    nmo = s.mo_num_kept
    ao_ref = oracle.ao_vgl(s, x)
    mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo).reshape(P, 5, nmo)
    vgl_close(mo, oracle.mo_vgl(oracle.kept_mo_coef(s), ao_ref), bound, f"mo_vgl high-l mode {mode}")
    ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num).reshape(P, 5, s.ao_num)
    vgl_close(ao, ao_ref, 1e-12, "ao_vgl high-l")
    q.context_destroy(ctx)


def test_mo_vgl_after_ao_getter_reuses_resident_ao():
    s, x, W = S.workload("small", walk_num=64)
    ctx = make_ctx(s, x, W, jastrow=False)
    P = W * s.elec_num
    q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num)
    mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * s.mo_num_kept).reshape(P, 5, -1)
    ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
    vgl_close(mo, ref, 1e-12, "mo_vgl")
    q.context_destroy(ctx)


@pytest.mark.parametrize("mode,bound", [(q.MO_FP32, 2e-5), (q.MO_BF16X3, 2e-6)])
@pytest.mark.parametrize("name,W", [("chbrclf", 2), ("small", 256), ("medium", 8)])
def test_mo_vgl_split_precision(mode, bound, name, W):
    """Split-precision products stay within `bound` * max|component| of the FP64 result."""
    s, x, W = systems.chbrclf() if name == "chbrclf" else S.workload(name, walk_num=W)
    ctx = make_ctx(s, x, W, jastrow=False)
    q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, mode))
    P = W * s.elec_num
    nmo = s.mo_num_kept
    mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo).reshape(P, 5, nmo)
    ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
    vgl_close(mo, ref, bound, f"mo_vgl mode {mode} {name}")
    q.context_destroy(ctx)


# ---------------------------------------------------------------- distances
def test_distances_chbrclf():
    s, x, W = systems.chbrclf()
    ctx = make_ctx(s, x, W, jastrow=False, mo=False)
    Ne, Nn = s.elec_num, s.nucl_num
    ee = q.get_array(ctx, "qmckl_get_electron_ee_distance", W * Ne * Ne)
    en = q.get_array(ctx, "qmckl_get_electron_en_distance", W * Ne * Nn).reshape(W, Ne, Nn)
    golden_check("ee_distance", ee, "qmckl_electron.org")
    golden_check("en_distance", en, "qmckl_electron.org")
    np.testing.assert_allclose(ee.reshape(W, Ne, Ne), oracle.ee_distance(x, Ne, W), rtol=1e-14, atol=0)
    q.context_destroy(ctx)


# ---------------------------------------------------------------- Jastrow
def _n2_ctx():
    s, x, W = systems.n2()
    ctx = q.context_create()
    L = q.lib()
    q.check(ctx, L.qmckl_set_nucleus_num(ctx, s.nucl_num))
    nc = np.ascontiguousarray(s.nucl_coord)
    q.check(ctx, L.qmckl_set_nucleus_coord(ctx, b"T", q.ptr(nc), nc.size))
    ch = np.ascontiguousarray(s.nucl_charge)
    q.check(ctx, L.qmckl_set_nucleus_charge(ctx, q.ptr(ch), ch.size))
    q.check(ctx, L.qmckl_set_electron_num(ctx, s.up, s.dn))
    S.set_coords(ctx, s, x, W)
    S.load  # parameters below follow org/qmckl_jastrow_champ.org:2420-2480
    q.check(ctx, L.qmckl_set_jastrow_champ_spin_independent(ctx, 0))
    q.check(ctx, L.qmckl_set_jastrow_champ_aord_num(ctx, s.aord))
    q.check(ctx, L.qmckl_set_jastrow_champ_bord_num(ctx, s.bord))
    q.check(ctx, L.qmckl_set_jastrow_champ_cord_num(ctx, s.cord))
    q.check(ctx, L.qmckl_set_jastrow_champ_type_nucl_num(ctx, s.type_nucl_num))
    keep = [np.ascontiguousarray(a) for a in (s.type_nucl_vector, s.a_vector, s.b_vector, s.c_vector, s.rescale_en)]
    q.check(ctx, L.qmckl_set_jastrow_champ_type_nucl_vector(ctx, q.ptr(keep[0]), keep[0].size))
    q.check(ctx, L.qmckl_set_jastrow_champ_a_vector(ctx, q.ptr(keep[1]), keep[1].size))
    q.check(ctx, L.qmckl_set_jastrow_champ_b_vector(ctx, q.ptr(keep[2]), keep[2].size))
    q.check(ctx, L.qmckl_set_jastrow_champ_c_vector(ctx, q.ptr(keep[3]), keep[3].size))
    q.check(ctx, L.qmckl_set_jastrow_champ_rescale_factor_en(ctx, q.ptr(keep[4]), keep[4].size))
    q.check(ctx, L.qmckl_set_jastrow_champ_rescale_factor_ee(ctx, s.rescale_ee))
    assert L.qmckl_jastrow_champ_provided(ctx)
    return ctx, s, x, W


def test_jastrow_n2_goldens():
    ctx, s, x, W = _n2_ctx()
    Ne, Nn, C = s.elec_num, s.nucl_num, s.cord
    g = lambda fn, n: q.get_array(ctx, fn, n)
    golden_check("factor_ee", g("qmckl_get_jastrow_champ_factor_ee", W), "qmckl_jastrow_champ.org")
    golden_check("factor_en", g("qmckl_get_jastrow_champ_factor_en", W), "qmckl_jastrow_champ.org")
    golden_check("factor_een", g("qmckl_get_jastrow_champ_factor_een", W), "qmckl_jastrow_champ.org")
    golden_check("factor_een_gl", g("qmckl_get_jastrow_champ_factor_een_gl", W * 4 * Ne).reshape(W, 4, Ne),
                 "qmckl_jastrow_champ.org")
    golden_check("asymp_jasb", g("qmckl_get_jastrow_champ_asymp_jasb", 2), "qmckl_jastrow_champ.org")
    golden_check("asymp_jasa", g("qmckl_get_jastrow_champ_asymp_jasa", 1), "qmckl_jastrow_champ.org")
    inter = {
        "ee_distance_rescaled": ("qmckl_get_jastrow_champ_ee_distance_rescaled", (W * Ne * Ne,)),
        "en_distance_rescaled": ("qmckl_get_jastrow_champ_en_distance_rescaled", (W, Nn, Ne)),
        "een_rescaled_e": ("qmckl_get_jastrow_champ_een_rescaled_e", (W, C + 1, Ne, Ne)),
        "een_rescaled_e_gl": ("qmckl_get_jastrow_champ_een_rescaled_e_gl", (W, C + 1, Ne, 4, Ne)),
        "een_rescaled_n": ("qmckl_get_jastrow_champ_een_rescaled_n", (W, C + 1, Nn, Ne)),
        "een_rescaled_n_gl": ("qmckl_get_jastrow_champ_een_rescaled_n_gl", (W, C + 1, Nn, 4, Ne)),
        "tmp_c": ("qmckl_get_jastrow_champ_tmp_c", (W, C, C + 1, Nn, Ne)),
        "dtmp_c": ("qmckl_get_jastrow_champ_dtmp_c", (W, C, C + 1, Nn, 4, Ne)),
    }
    for var, (fn, shape) in inter.items():
        golden_check(var, g(fn, int(np.prod(shape))).reshape(shape), "qmckl_jastrow_champ.org")
    q.context_destroy(ctx)


def _jastrow_gpu(ctx, W, Ne):
    g = lambda fn, n: q.get_array(ctx, fn, n)
    return {
        "value": g("qmckl_get_jastrow_champ_value", W),
        "gl": g("qmckl_get_jastrow_champ_gl", W * 4 * Ne).reshape(W, 4, Ne),
        "factor_ee": g("qmckl_get_jastrow_champ_factor_ee", W),
        "factor_en": g("qmckl_get_jastrow_champ_factor_en", W),
        "factor_een": g("qmckl_get_jastrow_champ_factor_een", W),
        "factor_ee_gl": g("qmckl_get_jastrow_champ_factor_ee_gl", W * 4 * Ne).reshape(W, 4, Ne),
        "factor_en_gl": g("qmckl_get_jastrow_champ_factor_en_gl", W * 4 * Ne).reshape(W, 4, Ne),
        "factor_een_gl": g("qmckl_get_jastrow_champ_factor_een_gl", W * 4 * Ne).reshape(W, 4, Ne),
    }


def _compare_jastrow(got, ref, tag, factor_tol=1e-12):
    for k in ("factor_ee", "factor_en", "factor_een"):
        err = np.abs(got[k] - ref[k]) / np.maximum(1.0, np.abs(ref[k]))
        assert err.max() <= factor_tol, f"{tag} {k}: {err.max():.3e}"
    # value = exp(J): relative error ~ |dJ|; J < -100 is flushed to exactly 0 by both, and
    # exp(J) overflows to +inf in both above J = 709.78
    J = ref["factor_ee"] + ref["factor_en"] + ref["factor_een"]
    flushed = J < -100.0
    np.testing.assert_array_equal(got["value"][flushed], 0.0)
    over = J > 709.79
    assert np.all(np.isposinf(ref["value"][over])) and np.all(np.isposinf(got["value"][over])), tag
    live = ~flushed & ~over & (np.abs(J - 709.78) > 1e-6)
    if live.any():
        rel = np.abs(got["value"][live] - ref["value"][live]) / ref["value"][live]
        assert rel.max() <= 1e-12 * max(1.0, np.abs(J[live]).max()), f"{tag} value {rel.max():.3e}"
    for k in ("factor_ee_gl", "factor_en_gl", "factor_een_gl", "gl"):
        # the total GL carries exp(J): compared on the walkers where it is finite
        rows = ~over if k == "gl" else np.ones(J.shape, bool)
        if not rows.any():
            continue
        for c in range(4):
            scale = np.abs(ref[k][rows, c]).max()
            err = np.abs(got[k][rows, c] - ref[k][rows, c]).max()
            assert err <= 1e-11 * max(scale, 1e-300), f"{tag} {k}[{c}]: {err:.3e} vs {scale:.3e}"


def test_jastrow_n2_vs_oracle():
    ctx, s, x, W = _n2_ctx()
    _compare_jastrow(_jastrow_gpu(ctx, W, s.elec_num), oracle.jastrow(s, x, W), "n2")
    q.context_destroy(ctx)


@pytest.mark.parametrize("name,W", [("small", 256), ("medium", 8), ("small", 13), ("small", 600), ("medium", 27)])
def test_jastrow_synthetic_vs_oracle(name, W):
    """The persistent e-e-n kernel runs one workgroup per CU over a walker-major range of
    (walker, tile) items.  Up to 256 items every workgroup takes one; 600 walkers of one
    tile and 27 walkers of 10 tiles give workgroups two or three items whose ranges start
    and end inside a walker, so the per-walker inputs are reloaded mid-range."""
    s, x, W = S.workload(name, walk_num=W)
    ctx = make_ctx(s, x, W, mo=False)
    got = _jastrow_gpu(ctx, W, s.elec_num)
    _compare_jastrow(got, oracle.jastrow(s, x, W), name)
    grad = q.get_array(ctx, "qmckl_get_jastrow_champ_grad", W * 3 * s.elec_num).reshape(W, 3, s.elec_num)
    np.testing.assert_array_equal(grad, got["gl"][:, :3, :])
    q.context_destroy(ctx)


@pytest.mark.parametrize("nn,up,dn,aord,bord,cord", [
    (20, 19, 18, 5, 5, 2), (20, 19, 18, 3, 4, 3), (17, 24, 16, 5, 5, 4), (33, 40, 39, 2, 6, 6),
    (20, 19, 18, 5, 5, 7), (9, 12, 11, 4, 3, 8), (3, 3, 2, 5, 5, 5), (1, 1, 0, 5, 5, 5),
    (25, 30, 29, 5, 5, 5), (32, 20, 20, 4, 4, 4), (18, 21, 20, 5, 5, 3),
    (40, 120, 120, 5, 5, 5), (20, 200, 200, 5, 5, 6), (37, 180, 170, 5, 5, 5), (50, 200, 190, 4, 4, 4)])
def test_jastrow_orders_and_shapes_vs_oracle(nn, up, dn, aord, bord, cord):
    """Every e-e-n order the kernels are built for (2..8), with electron and nucleus counts
    off the 16-wide tiles and the 32-electron chunks, down to a single electron.  In the
    LDS-resident e-e-n kernel a last nucleus tile of at most 4 nuclei runs on the 4x4x4
    products (20, 17, 33, 3, 1, 18 nuclei), with both units of a pair in one product when it
    holds at most 2 (17, 33, 1, 18), one of 9 or 16 nuclei on the 16x16x4 ones (25, 32).
    The last four shapes do not fit the resident kernel: 240 electrons x 40 nuclei run the
    pass-based kernel (one wave per (k, q) unit) in one pass over j, 400 electrons at cord 6
    the same kernel in two passes; 350 electrons x 37 nuclei (cord 5) and 390 x 50 (cord 4)
    the balanced pass-based kernel in two passes, the last of its pairs of nucleus tiles
    holding one tile (of 5 and 2 nuclei)."""
    s, x, W = S.jastrow_workload(nn, up, dn, aord, bord, cord, walk_num=8 if up + dn < 200 else 2)
    ctx = make_ctx(s, x, W, mo=False)
    _compare_jastrow(_jastrow_gpu(ctx, W, s.elec_num), oracle.jastrow(s, x, W), s.name)
    q.context_destroy(ctx)


VALUE_GETTERS = ["qmckl_get_jastrow_champ_factor_ee", "qmckl_get_jastrow_champ_factor_en",
                 "qmckl_get_jastrow_champ_factor_een", "qmckl_get_jastrow_champ_value"]


@pytest.mark.parametrize("nn,up,dn,aord,bord,cord", [
    (20, 19, 18, 5, 5, 2), (20, 19, 18, 3, 4, 3), (17, 24, 16, 5, 5, 4), (33, 40, 39, 2, 6, 6),
    (25, 30, 29, 5, 5, 5), (3, 3, 2, 5, 5, 5), (18, 21, 20, 5, 5, 3),
    (37, 180, 170, 5, 5, 5), (50, 200, 190, 4, 4, 4), (40, 120, 120, 5, 5, 5), (20, 19, 18, 5, 5, 7)])
def test_jastrow_values_only(nn, up, dn, aord, bord, cord):
    """The value and factor getters with no current GL run the values-only evaluation where
    the LDS-resident e-e-n kernel applies, and the balanced pass-based one (350 electrons at
    cord 5, with an odd number of nucleus tiles, and 390 at cord 4): the value component of
    the products alone, the fold of J alone.  The three factors and e^J equal those of the
    full evaluation bit for bit, through the host and the device getters, and the oracle's
    to 1e-12.  240 electrons (the one-wave-per-unit pass-based kernel) and cord 7 take them
    from the full evaluation."""
    s, x, W = S.jastrow_workload(nn, up, dn, aord, bord, cord, walk_num=8 if up + dn < 200 else 2)
    ctx = make_ctx(s, x, W, mo=False)
    dev = q.DeviceBuffer(8 * W)
    try:
        vals = {n: q.get_array(ctx, n, W) for n in VALUE_GETTERS}  # values first: values-only path
        S.set_coords(ctx, s, x, W)
        q.check(ctx, q.lib().qmckl_amd_get_jastrow_champ_value_device(ctx, dev.ptr, W))
        q.check(ctx, q.lib().qmckl_amd_synchronize(ctx))
        dval = dev.to_array((W,))
        S.set_coords(ctx, s, x, W)
        q.get_array(ctx, "qmckl_get_jastrow_champ_gl", W * 4 * s.elec_num)  # the full evaluation
        full = {n: q.get_array(ctx, n, W) for n in VALUE_GETTERS}
        for n in VALUE_GETTERS:
            assert np.array_equal(vals[n], full[n]), f"{s.name} {n}"
        assert np.array_equal(dval, full["qmckl_get_jastrow_champ_value"])
        ref = oracle.jastrow(s, x, W, want=("factor_ee", "factor_en", "factor_een"))
        for k in ("factor_ee", "factor_en", "factor_een"):
            err = np.abs(vals[f"qmckl_get_jastrow_champ_{k}"] - ref[k]) / np.maximum(1.0, np.abs(ref[k]))
            assert err.max() <= 1e-12, f"{s.name} {k}: {err.max():.3e}"
    finally:
        dev.free()
        q.context_destroy(ctx)


def test_jastrow_value_then_gl_evaluates_once():
    """A caller that reads the value and then the GL at every step gets the values-only
    evaluation on its first step only: once it has asked for the GL after a value request,
    the next value request runs the full evaluation (the GL request then finds it done), and
    a step without a GL request returns to the values-only evaluation one step later.  The
    Jastrow totals' profiling slot (4) counts the full evaluations (the values-only one
    launches its total under the e-e-n slot)."""
    s, x, W = S.workload("medium", walk_num=8)
    ctx = make_ctx(s, x, W, mo=False)
    L = q.lib()
    ms = np.zeros(5)
    cnt = np.zeros(5, np.int64)

    def full_evals(steps, gl):
        q.check(ctx, L.qmckl_amd_get_kernel_times(ctx, q.ptr(ms), q.ptr(cnt), 5))
        for i in range(steps):
            S.set_coords(ctx, s, x + 0.01 * i, W)
            q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
            if gl:
                q.get_array(ctx, "qmckl_get_jastrow_champ_gl", W * 4 * s.elec_num)
        q.check(ctx, L.qmckl_amd_get_kernel_times(ctx, q.ptr(ms), q.ptr(cnt), 5))
        return int(cnt[4]), int(cnt[3])

    try:
        q.check(ctx, L.qmckl_amd_set_profiling(ctx, 1))
        assert full_evals(1, gl=True) == (1, 2)  # values-only, then the full evaluation for the GL
        assert full_evals(3, gl=True) == (3, 3)  # the full evaluation on each value request
        assert full_evals(1, gl=False) == (1, 1)  # still full: the GL was asked after the last value
        assert full_evals(2, gl=False) == (0, 2)  # values-only
        q.check(ctx, L.qmckl_amd_set_profiling(ctx, 0))
    finally:
        q.context_destroy(ctx)


def test_coordinate_update_recomputes():
    """A new set of coordinates invalidates every cached result (date semantics)."""
    s, x, W = S.workload("small", walk_num=16)
    ctx = make_ctx(s, x, W)
    Ne = s.elec_num
    v1 = q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
    m1 = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", W * Ne * 5 * s.mo_num_kept)
    x2 = x + 0.1
    S.set_coords(ctx, s, x2, W)
    v2 = q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
    m2 = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", W * Ne * 5 * s.mo_num_kept)
    assert not np.array_equal(v1, v2) and not np.array_equal(m1, m2)
    np.testing.assert_allclose(v2, oracle.jastrow(s, x2, W)["value"], rtol=1e-12)
    q.context_destroy(ctx)


# ---------------------------------------------------------------- device entry points
def test_large_host_getters_through_pinned_staging():
    """Host getters above 16 MiB go through two pinned 32 MiB staging buffers in chunks: an
    86 MB MO VGL (3 chunks, both buffers) and a 1 MB GL, against the device entry points
    read back directly, bit for bit, twice (the staging buffers are reused)."""
    s, x, W = S.workload("medium", walk_num=192)
    ctx = make_ctx(s, x, W)
    L = q.lib()
    Ne, nmo = s.elec_num, s.mo_num_kept
    n = W * Ne * 5 * nmo
    assert 8 * n > 2 * 32 * 2**20
    mo = q.DeviceBuffer(8 * n)
    gl = q.DeviceBuffer(8 * W * 4 * Ne)
    xd = q.DeviceBuffer.from_array(x)
    for _ in range(2):
        host_mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", n)
        host_gl = q.get_array(ctx, "qmckl_get_jastrow_champ_gl", W * 4 * Ne)
        q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
        q.check(ctx, L.qmckl_amd_get_mo_basis_mo_vgl_device(ctx, mo.ptr, n))
        q.check(ctx, L.qmckl_amd_get_jastrow_champ_gl_device(ctx, gl.ptr, W * 4 * Ne))
        q.synchronize(ctx)
        np.testing.assert_array_equal(mo.to_array(n), host_mo)
        np.testing.assert_array_equal(gl.to_array(W * 4 * Ne), host_gl)
    q.context_destroy(ctx)


def test_device_entry_points_match_host_getters():
    s, x, W = S.workload("small", walk_num=32)
    ctx = make_ctx(s, x, W)
    L = q.lib()
    Ne, nmo = s.elec_num, s.mo_num_kept
    P = W * Ne
    host_mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo)
    host_v = q.get_array(ctx, "qmckl_get_jastrow_champ_value", W)
    host_gl = q.get_array(ctx, "qmckl_get_jastrow_champ_gl", W * 4 * Ne)
    host_ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num)
    xd = q.DeviceBuffer.from_array(x)
    mo = q.DeviceBuffer(8 * P * 5 * nmo)
    ao = q.DeviceBuffer(8 * P * 5 * s.ao_num)
    val = q.DeviceBuffer(8 * W)
    gl = q.DeviceBuffer(8 * W * 4 * Ne)
    q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x.size))
    q.check(ctx, L.qmckl_amd_get_ao_basis_ao_vgl_device(ctx, ao.ptr, P * 5 * s.ao_num))
    q.check(ctx, L.qmckl_amd_get_mo_basis_mo_vgl_device(ctx, mo.ptr, P * 5 * nmo))
    q.check(ctx, L.qmckl_amd_get_jastrow_champ_value_device(ctx, val.ptr, W))
    q.check(ctx, L.qmckl_amd_get_jastrow_champ_gl_device(ctx, gl.ptr, W * 4 * Ne))
    q.synchronize(ctx)
    np.testing.assert_array_equal(ao.to_array(P * 5 * s.ao_num), host_ao)
    np.testing.assert_array_equal(mo.to_array(P * 5 * nmo), host_mo)
    np.testing.assert_array_equal(val.to_array(W), host_v)
    np.testing.assert_array_equal(gl.to_array(W * 4 * Ne), host_gl)
    back = np.empty(3 * P)
    q.check(ctx, L.qmckl_get_electron_coord(ctx, b"T", q.ptr(back), back.size))
    np.testing.assert_array_equal(back.reshape(3, P), x)
    # the 'N' layout is refused for device coordinates (INVALID_ARG_2)
    assert L.qmckl_amd_set_electron_coord_device(ctx, b"N", W, xd.ptr, x.size) == 2
    q.context_destroy(ctx)
