"""GPU parity at the benchmark configurations and on the branches the other parity files do
not reach, through the C-ABI, against the CPU oracle:

* the large configuration (200 nuclei, 1000 electrons, 5000 AOs, 500 MOs, precision 35): MO
  VGL in FP64 and in 3xBF16, Jastrow factors and GL;
* the split-precision error contract of the 3xBF16 AO->MO product: for each VGL component,
  max |MO_split - MO_FP64| <= 1e-7 * max |MO_FP64| for the value and the three gradient
  components, <= 1e-6 * max |MO_FP64| for the Laplacian, on the small, medium and large
  configurations, the molecules of the batched one and the s-to-h-shell system; CHBrClF
  (Br and Cl cores) within 4e-7 on the value; the FP32 MFMA product (FP32 operands) within
  1e-6 on every component;
* the spin-independent e-e term, r_power in the nucleus screening range, the rescaled
  distance GL intermediates, the e-e / e-n potentials, the nuclear repulsion, the
  `_inplace` getters and factor_een_grad.

FP64 tolerances as in test_gpu_parity.py: MO <= 1e-12 of scale per component, Jastrow factors
<= 1e-12 relative, GL <= 1e-11 of scale.
"""
import ctypes

import numpy as np
import pytest

import oracle
import qmckl_amd as q
import systems
from qmckl_amd import system as S
from test_gpu_parity import _compare_jastrow, _jastrow_gpu, make_ctx, vgl_close

pytestmark = pytest.mark.gpu

# the split-precision contract: value and gradient components, Laplacian
SPLIT_BOUND = (1e-7, 1e-7, 1e-7, 1e-7, 1e-6)


def split_errors(mo, ref):
    """max |mo - ref| / max |ref| per VGL component."""
    return [float(np.abs(mo[:, c] - ref[:, c]).max() / np.abs(ref[:, c]).max()) for c in range(5)]


# FP32 MFMA products: FP32 operands (24-bit significands) and FP32 sums per 32-AO chunk
FP32_BOUND = (1e-6,) * 5
# CHBrClF: at the worst MO value the terms sum_k |AO_k C_kj| are 4x max |MO|, and the FP32
# sums inside the matrix instructions round at 2^-24 of them; measured 3.0e-7 on the value
SPLIT_BOUND_CHBRCLF = (4e-7,) + SPLIT_BOUND[1:]


def assert_split_contract(mo, ref, tag, bound=SPLIT_BOUND):
    err = split_errors(mo, ref)
    print(f"{tag}: split-precision MO error per component {['%.2e' % e for e in err]}", flush=True)
    for c, (e, b) in enumerate(zip(err, bound)):
        assert e <= b, f"{tag} component {c}: {e:.3e} > {b:.0e}"


def _mo(ctx, s, W, mode):
    q.check(ctx, q.lib().qmckl_amd_set_mo_precision(ctx, mode))
    P, nmo = W * s.elec_num, s.mo_num_kept
    return q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo).reshape(P, 5, nmo)


# ---------------------------------------------------------------- large configuration
def test_large_config_mo_and_jastrow_vs_oracle():
    """200 nuclei, 1000 electrons, 5000 AOs, 500 occupied MOs at precision 35 (one walker):
    the pass-based e-e-n kernel and the 5000-AO fused products."""
    s, x, W = S.workload("large", walk_num=1)
    print(f"large: Nn={s.nucl_num} Ne={s.elec_num} Nao={s.ao_num} Nmo={s.mo_num_kept}", flush=True)
    ctx = make_ctx(s, x, W)
    ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
    vgl_close(_mo(ctx, s, W, q.MO_FP64), ref, 1e-12, "mo_vgl large fp64")
    assert_split_contract(_mo(ctx, s, W, q.MO_BF16X3), ref, "large bf16x3")
    # factor_en sums Ne Nn = 200 000 pair terms, each with its asymptote subtracted, and
    # factor_ee 500 000: the summation orders of the GPU (fixed trees) and of the oracle
    # (sequential) differ by a few 1e-12 there, so the factors get 1e-11
    _compare_jastrow(_jastrow_gpu(ctx, W, s.elec_num), oracle.jastrow(s, x, W), "large", factor_tol=1e-11)
    q.context_destroy(ctx)


# ---------------------------------------------------------------- split-precision contract
@pytest.mark.parametrize("mode", [q.MO_BF16X3, q.MO_FP32])
@pytest.mark.parametrize("name,W", [("chbrclf", 2), ("small", 256), ("medium", 8), ("high_l", 16)])
def test_split_precision_error_contract(mode, name, W):
    if name == "chbrclf":
        s, x, W = systems.chbrclf()
    elif name == "high_l":
        s, x, W = systems.high_l()
    else:
        s, x, W = S.workload(name, walk_num=W)
    ctx = make_ctx(s, x, W, jastrow=False)
    mo64 = _mo(ctx, s, W, q.MO_FP64)
    ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
    vgl_close(mo64, ref, 1e-12, f"mo_vgl {name} fp64")
    # against the library's own FP64 product (what bench.py reports) and against the oracle
    got = _mo(ctx, s, W, mode)
    bound = FP32_BOUND if mode == q.MO_FP32 else SPLIT_BOUND_CHBRCLF if name == "chbrclf" else SPLIT_BOUND
    assert_split_contract(got, mo64, f"{name} mode {mode} vs fp64", bound)
    assert_split_contract(got, ref, f"{name} mode {mode} vs oracle", bound)
    q.context_destroy(ctx)


# ---------------------------------------------------------------- branches
def test_spin_independent_jastrow_vs_oracle():
    s, x, W = S.jastrow_workload(20, 19, 18, 5, 5, 5, walk_num=8)
    s.spin_independent = 1
    ctx = make_ctx(s, x, W, mo=False)
    got = _jastrow_gpu(ctx, W, s.elec_num)
    ref = oracle.jastrow(s, x, W)
    _compare_jastrow(got, ref, "spin independent")
    jb = q.get_array(ctx, "qmckl_get_jastrow_champ_asymp_jasb", 2)
    np.testing.assert_allclose(jb, oracle.asymp_jasb(s), rtol=1e-14)
    q.context_destroy(ctx)


def test_r_power_in_nucleus_range():
    """r_power = -1 on every shell of one nucleus makes its range 0 (its AOs vanish off the
    nucleus); r_power = 2 on the other widens its range (no change)."""
    s, x, W = S.workload("small", walk_num=32)
    rp = np.zeros(s.shell_num, np.int32)
    rp[s.nucleus_index[0]:s.nucleus_index[0] + s.nucleus_shell_num[0]] = -1
    rp[s.nucleus_index[1]:s.nucleus_index[1] + s.nucleus_shell_num[1]] = 2
    s.r_power = rp
    ctx = make_ctx(s, x, W, jastrow=False)
    P = W * s.elec_num
    ao_ref = oracle.ao_vgl(s, x)
    ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num).reshape(P, 5, s.ao_num)
    vgl_close(ao, ao_ref, 1e-12, "ao_vgl r_power")
    got_rp = q.get_array(ctx, "qmckl_get_ao_basis_r_power", s.shell_num, np.int32)
    np.testing.assert_array_equal(got_rp, rp)
    ctx2 = make_ctx(s, x, W, jastrow=False)
    mo = _mo(ctx2, s, W, q.MO_FP64)
    vgl_close(mo, oracle.mo_vgl(oracle.kept_mo_coef(s), ao_ref), 1e-12, "mo_vgl r_power")
    for c in (ctx, ctx2):
        q.context_destroy(c)


def test_rescaled_distance_gl_getters_vs_oracle():
    s, x, W = S.jastrow_workload(5, 6, 5, 5, 5, 5, walk_num=4)
    ctx = make_ctx(s, x, W, mo=False)
    Ne, Nn = s.elec_num, s.nucl_num
    ref = oracle.jastrow(s, x, W, want=("ee_rescaled", "ee_rescaled_gl", "en_rescaled", "en_rescaled_gl"))
    for key, fn, shape in [
        ("ee_rescaled", "qmckl_get_jastrow_champ_ee_distance_rescaled", (W, Ne, Ne)),
        ("ee_rescaled_gl", "qmckl_get_jastrow_champ_ee_distance_rescaled_gl", (W, Ne, Ne, 4)),
        ("en_rescaled", "qmckl_get_jastrow_champ_en_distance_rescaled", (W, Nn, Ne)),
        ("en_rescaled_gl", "qmckl_get_jastrow_champ_en_distance_rescaled_gl", (W, Nn, Ne, 4)),
    ]:
        got = q.get_array(ctx, fn, int(np.prod(shape))).reshape(shape)
        np.testing.assert_allclose(got, ref[key], rtol=1e-13, atol=1e-14, err_msg=key)
    q.context_destroy(ctx)


def test_potentials_and_nuclear_repulsion():
    """V_ee = sum_{i<j} 1/r_ij, V_en = -sum_{i,a} Z_a / r_ia (org/qmckl_electron.org:1022-1192,
    1477-1658) against numpy, and V_nn on CHBrClF against the reference's value
    318.2309879436158 (org/qmckl_nucleus.org:1098-1099) and the closed form."""
    s, x, W = systems.chbrclf()
    ctx = make_ctx(s, x, W, jastrow=False, mo=False)
    L = q.lib()
    Ne = s.elec_num
    vee = np.empty(W)
    ven = np.empty(W)
    q.check(ctx, L.qmckl_get_electron_ee_potential(ctx, q.ptr(vee)))
    q.check(ctx, L.qmckl_get_electron_en_potential(ctx, q.ptr(ven)))
    xw = x.reshape(3, W, Ne)
    for w in range(W):
        r = np.linalg.norm(xw[:, w, :, None] - xw[:, w, None, :], axis=0)
        iu = np.triu_indices(Ne, 1)
        assert abs(vee[w] - (1.0 / r[iu]).sum()) <= 1e-12 * abs(vee[w])
        R = np.linalg.norm(xw[:, w, :, None] - s.nucl_coord[:, None, :], axis=0)
        assert abs(ven[w] + (s.nucl_charge[None, :] / R).sum()) <= 1e-12 * abs(ven[w])
    rep = ctypes.c_double()
    q.check(ctx, L.qmckl_get_nucleus_repulsion(ctx, ctypes.byref(rep)))
    assert abs(rep.value - 318.2309879436158) < 1e-10
    nn = q.get_array(ctx, "qmckl_get_nucleus_nn_distance", s.nucl_num ** 2).reshape(s.nucl_num, s.nucl_num)
    ref_nn = np.linalg.norm(s.nucl_coord[:, :, None] - s.nucl_coord[:, None, :], axis=0)
    np.testing.assert_allclose(nn, ref_nn, rtol=1e-14, atol=1e-14)
    q.context_destroy(ctx)


def test_inplace_getters_and_een_grad():
    """The reference's Python caller reads MO VGL through the `_inplace` getter
    (python/test/test_api.py:47): every `_inplace` getter returns what its plain getter
    does; factor_een_grad is the gradient rows of factor_een_gl."""
    s, x, W = S.workload("small", walk_num=16)
    ctx = make_ctx(s, x, W)
    P, nmo, nao, Ne = W * s.elec_num, s.mo_num_kept, s.ao_num, s.elec_num
    for plain, inplace, n in [
        ("qmckl_get_mo_basis_mo_vgl", "qmckl_get_mo_basis_mo_vgl_inplace", P * 5 * nmo),
        ("qmckl_get_mo_basis_mo_value", "qmckl_get_mo_basis_mo_value_inplace", P * nmo),
        ("qmckl_get_ao_basis_ao_vgl", "qmckl_get_ao_basis_ao_vgl_inplace", P * 5 * nao),
        ("qmckl_get_ao_basis_ao_value", "qmckl_get_ao_basis_ao_value_inplace", P * nao),
    ]:
        np.testing.assert_array_equal(q.get_array(ctx, inplace, n), q.get_array(ctx, plain, n), err_msg=inplace)
    gl = q.get_array(ctx, "qmckl_get_jastrow_champ_factor_een_gl", W * 4 * Ne).reshape(W, 4, Ne)
    grad = q.get_array(ctx, "qmckl_get_jastrow_champ_factor_een_grad", W * 3 * Ne).reshape(W, 3, Ne)
    np.testing.assert_array_equal(grad, gl[:, :3, :])
    ref = oracle.jastrow(s, x, W, want=("factor_een_gl",))["factor_een_gl"]
    for c in range(3):
        assert np.abs(grad[:, c] - ref[:, c]).max() <= 1e-11 * np.abs(ref[:, c]).max()
    assert q.lib().qmckl_get_jastrow_champ_factor_een_grad(ctx, q.ptr(np.empty(4)), 4) == 3
    q.context_destroy(ctx)


def test_context_copy_after_device_coordinates():
    """A copy taken right after device coordinates were set holds those coordinates (the
    copy reads them back on the context's stream)."""
    s, x, W = S.workload("small", walk_num=64)
    ctx = make_ctx(s, x, W)
    L = q.lib()
    x2 = np.ascontiguousarray(x + 0.25)
    xd = q.DeviceBuffer.from_array(x2)
    q.check(ctx, L.qmckl_amd_set_electron_coord_device(ctx, b"T", W, xd.ptr, x2.size))
    c2 = L.qmckl_context_copy(ctx)
    assert c2 != 0
    back = np.empty(x2.size)
    q.check(c2, L.qmckl_get_electron_coord(c2, b"T", q.ptr(back), back.size))
    np.testing.assert_array_equal(back.reshape(x2.shape), x2)
    v = q.get_array(c2, "qmckl_get_jastrow_champ_value", W)
    np.testing.assert_array_equal(v, q.get_array(ctx, "qmckl_get_jastrow_champ_value", W))
    for c in (ctx, c2):
        q.context_destroy(c)


# ---------------------------------------------------------------- Slater ('S') shells
def test_slater_basis_vs_oracle():
    """A basis of Slater shells (r^n e^(-alpha r), n = r_power 0 and 1; s, p and d shells)
    through the AO and shell VGL getters and the fused AO->MO product in FP64 (1e-12 of scale)
    and 3xBF16 (the split-precision contract), against the oracle; one electron of one walker
    sits exactly on a nucleus (the r = 0 branch).  The primitive VGL and the Hessians are
    Gaussian-only here, as the primitive VGL is in the reference: both answer QMCKL_FAILURE."""
    s, x, W = S.slater_workload(walk_num=8)
    x = x.copy()
    x[:, 0] = s.nucl_coord[:, 0]  # walker 0, electron 0 on nucleus 0
    ctx = make_ctx(s, x, W)
    try:
        P = W * s.elec_num
        ao = q.get_array(ctx, "qmckl_get_ao_basis_ao_vgl", P * 5 * s.ao_num).reshape(P, 5, s.ao_num)
        vgl_close(ao, oracle.ao_vgl(s, x), 1e-12, "slater ao_vgl")
        sh = q.get_array(ctx, "qmckl_get_ao_basis_shell_vgl", P * 5 * s.shell_num).reshape(P, 5, s.shell_num)
        vgl_close(sh, oracle.shell_vgl(s, x), 1e-12, "slater shell_vgl")
        ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
        S.set_coords(ctx, s, x, W)  # a new date: the MO product evaluates the AOs itself
        vgl_close(_mo(ctx, s, W, q.MO_FP64), ref, 1e-12, "slater mo_vgl fp64")
        # the point on nucleus 0 holds the s AOs at their cusp, the largest terms of the
        # product: the FP32 sums of the leading bf16 product inside the matrix instructions
        # reach 1.3e-7 of the largest MO value there (the exception of the batched molecules)
        assert_split_contract(_mo(ctx, s, W, q.MO_BF16X3), ref, "slater bf16x3",
                              bound=(1.5e-7,) * 4 + SPLIT_BOUND[4:])
        out = np.zeros(P * 5 * s.prim_num)
        assert q.lib().qmckl_get_ao_basis_primitive_vgl(ctx, out.ctypes.data, out.size) == q.QMCKL_FAILURE
        out = np.zeros(P * 12 * s.ao_num)
        assert q.lib().qmckl_get_ao_basis_ao_hessian(ctx, out.ctypes.data, out.size) == q.QMCKL_FAILURE
    finally:
        q.context_destroy(ctx)


def test_low_precision_default_product():
    """At a numerical precision of 24 bits, with no MO precision set, the product runs in
    3xBF16 (the reference's optimised build switches to single precision at 26 bits or
    fewer): within the split-precision contract of the FP64 oracle at the same precision."""
    s, x, W = S.workload("small", walk_num=16)
    s.precision = 24
    ctx = make_ctx(s, x, W)
    try:
        P, nmo = W * s.elec_num, s.mo_num_kept
        mo = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * nmo).reshape(P, 5, nmo)
        ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
        assert_split_contract(mo, ref, "precision 24, default product")
        assert np.abs(mo - ref).max() > 0.0  # not the FP64 product
        vgl_close(_mo(ctx, s, W, q.MO_FP64), ref, 1e-12, "precision 24, FP64 set")
    finally:
        q.context_destroy(ctx)
