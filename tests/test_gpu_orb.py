"""GPU: orbital Hessians, forces on the orbitals and the MO cusp correction through the C-ABI,
against oracle/orbitals.py (itself pinned by finite differences and the cusp conditions):

* shell Hessian [P][3][4][shell], AO Hessian [3][P][4][ao], MO Hessian [3][P][3][mo] on the
  small, high-l and CHBrClF systems at 1e-12 of each component's scale;
* forces on the AO values, MO values, gradients and Laplacians (and the _inplace getters);
* qmckl_set_mo_basis_r_cusp: MO VGL, MO values and MO Hessian with the correction at points
  inside and outside the spheres, against the oracle's own fit; the caller's points are kept;
  qmckl_mo_basis_select_mo fits again for the kept MOs; 3xBF16 products with the correction;
  the argument errors.
"""
import numpy as np
import pytest

import oracle
import qmckl_amd as q
import systems
from oracle import orbitals as OO
from qmckl_amd import system as S

pytestmark = pytest.mark.gpu


def _sys(name):
    if name == "small":
        s, x, W = S.workload("small", walk_num=4)
    elif name == "high_l":
        s, x, W = systems.high_l(walk_num=3)
    else:
        s, x, W = systems.chbrclf()
    return s, x, W


def _ctx(s, x, W):
    ctx = q.context_create()
    S.load(ctx, s, jastrow=False)
    S.set_coords(ctx, s, x, W)
    return ctx


def close(got, ref, rel, what, axis=None):
    """|got - ref| <= rel * max|ref| per slice along `axis` (each component its own scale)."""
    if axis is None:
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(got - ref).max()
        assert err <= rel * scale, f"{what}: {err:.3e} of scale {scale:.3e}"
        return
    for i in range(ref.shape[axis]):
        close(np.take(got, i, axis), np.take(ref, i, axis), rel, f"{what}[{i}]")


@pytest.mark.parametrize("name", ["small", "high_l", "chbrclf"])
def test_hessians(name):
    print(f"hessians {name}", flush=True)
    s, x, W = _sys(name)
    ctx = _ctx(s, x, W)
    P = x.shape[1]
    sh = q.get_array(ctx, "qmckl_get_ao_basis_shell_hessian", P * 12 * s.shell_num).reshape(P, 3, 4, -1)
    close(sh, OO.shell_hessian(s, x), 1e-12, "shell hessian", axis=2)
    ah = q.get_array(ctx, "qmckl_get_ao_basis_ao_hessian", 12 * P * s.ao_num).reshape(3, P, 4, -1)
    ref = OO.ao_hessian(s, x)
    close(ah, ref, 1e-12, "ao hessian", axis=2)
    mo = s.mo_num_kept
    mh = q.get_array(ctx, "qmckl_get_mo_basis_mo_hessian", 9 * P * mo).reshape(3, P, 3, mo)
    close(mh, OO.mo_hessian(s, x), 1e-12, "mo hessian", axis=2)
    q.context_destroy(ctx)


@pytest.mark.parametrize("name", ["small", "high_l"])
def test_orbital_forces(name):
    print(f"orbital forces {name}", flush=True)
    s, x, W = _sys(name)
    ctx = _ctx(s, x, W)
    P, Nn, ao, mo = x.shape[1], s.nucl_num, s.ao_num, s.mo_num_kept
    ao_ref = oracle.ao_vgl(s, x)
    H = OO.ao_hessian(s, x)
    fa = q.get_array(ctx, "qmckl_get_forces_ao_value", Nn * P * 3 * ao).reshape(Nn, P, 3, ao)
    close(fa, OO.forces_ao_value(s, x, ao_ref), 1e-12, "forces ao value")
    for fn in ("qmckl_get_forces_mo_value", "qmckl_get_forces_mo_value_inplace"):
        fm = q.get_array(ctx, fn, Nn * P * 3 * mo).reshape(Nn, P, 3, mo)
        close(fm, OO.forces_mo_value(s, x, ao_ref), 1e-12, fn)
    for fn in ("qmckl_get_forces_mo_g", "qmckl_get_forces_mo_g_inplace"):
        fg = q.get_array(ctx, fn, Nn * 9 * P * mo).reshape(Nn, 3, P, 3, mo)
        close(fg, OO.forces_mo_g(s, x, H), 1e-12, fn)
    fl = q.get_array(ctx, "qmckl_get_forces_mo_l", Nn * 3 * P * mo).reshape(Nn, 3, P, mo)
    close(fl, OO.forces_mo_l(s, x, H), 1e-12, "forces mo l")
    q.context_destroy(ctx)


def _cusp_points(s, r_cusp, rng, n_in=6, n_out=6):
    """Points inside the spheres (radius 0.3 - 0.9 r_cusp) and outside, [3][n]."""
    pts = []
    for i in range(n_in):
        a = i % s.nucl_num
        u = rng.normal(size=3)
        pts.append(s.nucl_coord[:, a] + rng.uniform(0.3, 0.9) * r_cusp[a] * u / np.linalg.norm(u))
    for i in range(n_out):
        a = i % s.nucl_num
        u = rng.normal(size=3)
        pts.append(s.nucl_coord[:, a] + rng.uniform(1.2, 3.0) * r_cusp[a] * u / np.linalg.norm(u))
    return np.ascontiguousarray(np.array(pts).T)


@pytest.mark.parametrize("mode", [q.MO_FP64, q.MO_BF16X3])
def test_mo_cusp_correction(mode):
    s, x, W = S.workload("small", walk_num=4)
    s.keep_mo = None
    ctx = _ctx(s, x, W)
    L = q.lib()
    q.check(ctx, L.qmckl_amd_set_mo_precision(ctx, mode))
    r_cusp = np.array([0.25, 0.3])
    q.check(ctx, L.qmckl_set_mo_basis_r_cusp(ctx, q.ptr(r_cusp), r_cusp.size))
    # the caller's points are back
    back = np.empty(x.size)
    q.check(ctx, L.qmckl_get_point(ctx, b"T", q.ptr(back), back.size))
    np.testing.assert_array_equal(back.reshape(x.shape), x)
    cp = OO.cusp_params(s, r_cusp)
    pts = _cusp_points(s, r_cusp, np.random.default_rng(7))
    n = pts.shape[1]
    q.check(ctx, L.qmckl_set_point(ctx, b"T", n, q.ptr(pts), pts.size))
    mo = s.mo_num
    got = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", n * 5 * mo).reshape(n, 5, mo)
    ref = OO.mo_vgl_cusp(s, pts, r_cusp, cp)
    # 3xBF16: the s AOs are taken out in FP64 but the product's split error on their (large,
    # core) contribution stays: ~1e-7 of the s part, up to ~1e-6 of the corrected MO
    tol = 1e-10 if mode == q.MO_FP64 else 4e-6
    close(got, ref, tol, "cusp mo vgl", axis=1)
    val = q.get_array(ctx, "qmckl_get_mo_basis_mo_value", n * mo).reshape(n, mo)
    np.testing.assert_array_equal(val, got[:, 0])
    if mode == q.MO_FP64:
        mh = q.get_array(ctx, "qmckl_get_mo_basis_mo_hessian", 9 * n * mo).reshape(3, n, 3, mo)
        close(mh, OO.mo_hessian(s, pts, r_cusp, cp), 1e-10, "cusp mo hessian", axis=2)
    # the correction changes the MOs only inside the spheres
    plain = oracle.mo_vgl(s.mo_coef, oracle.ao_vgl(s, pts))
    assert np.abs(ref[6:] - plain[6:]).max() <= 1e-14 * np.abs(plain).max()
    assert np.abs(ref[:6] - plain[:6]).max() > 1e-3
    # select_mo fits again for the kept MOs
    keep = np.zeros(mo, np.int32)
    keep[[0, 3, 5, 8]] = 1
    q.check(ctx, L.qmckl_mo_basis_select_mo(ctx, q.ptr(keep), keep.size))
    s.keep_mo = keep
    got2 = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", n * 5 * 4).reshape(n, 5, 4)
    ref2 = OO.mo_vgl_cusp(s, pts, r_cusp, OO.cusp_params(s, r_cusp))
    close(got2, ref2, tol, "cusp mo vgl after select_mo", axis=1)
    q.context_destroy(ctx)


def test_mo_cusp_errors():
    s, x, W = S.workload("small", walk_num=2)
    ctx = _ctx(s, x, W)
    L = q.lib()
    r = np.array([0.25, 0.25])
    assert L.qmckl_set_mo_basis_r_cusp(ctx, None, 2) == 2
    assert L.qmckl_set_mo_basis_r_cusp(ctx, q.ptr(r), 1) == 3
    bad = np.array([3.0 / 7.0, 0.25])  # 2 R Z - 6 = 0 for Z = 7
    assert L.qmckl_set_mo_basis_r_cusp(ctx, q.ptr(bad), 2) == 2
    # no correction was left behind: the MOs are the plain ones
    P = x.shape[1]
    mo = s.mo_num_kept
    got = q.get_array(ctx, "qmckl_get_mo_basis_mo_vgl", P * 5 * mo).reshape(P, 5, mo)
    ref = oracle.mo_vgl(oracle.kept_mo_coef(s), oracle.ao_vgl(s, x))
    close(got, ref, 1e-12, "mo vgl", axis=1)
    q.context_destroy(ctx)
