"""The orbital-Hessian, MO-cusp and orbital-force oracle (oracle/orbitals.py) against what the
mathematics fixes:

* shell and AO Hessians (and the gradient of the Laplacian) against 8th-order central finite
  differences of the C oracle's shell and AO VGL in the electron coordinates, on s..h shells;
  the AO Hessian also by the reference's own protocol (finite differences in the nuclear
  coordinates, org/qmckl_ao.org:9753-9830);
* the MO Hessian against finite differences of the MO gradient, with and without the cusp;
* the forces on the AO values and on the MO values, gradients and Laplacians against finite
  differences of the AO / MO VGL in the nuclear coordinates (a dropped sign or a wrong nucleus
  block fails);
* the cusp fit: the corrected MOs are continuous in value, gradient and Laplacian through the
  sphere r = r_cusp, and satisfy the cusp condition c1 = -Z (c0 + eta) at the nucleus.
"""
import copy

import numpy as np
import pytest

import oracle
import systems
from oracle import orbitals as OO
from qmckl_amd import system as S

FD = np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0.0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])


def _fd(f, x, i, h):
    """d f / d x_i by the 9-point stencil."""
    acc = 0.0
    for m, c in zip(range(-4, 5), FD):
        if c == 0.0:
            continue
        y = x.copy()
        y[i] += m * h
        acc = acc + c * f(y)
    return acc / h


@pytest.fixture(scope="module", params=["small", "high_l"])
def system(request):
    """The 'small' system (two N nuclei, s-d shells, all 30 MOs) and the high-l one (three
    nuclei, s-h shells, AO factors in [0.5, 2]); 12 electron positions each."""
    if request.param == "small":
        s, x, W = S.workload("small", walk_num=2)
        s.keep_mo = None
    else:
        s, x, W = systems.high_l(walk_num=2)
    return s, x[:, :12]


def test_shell_and_ao_hessian_match_finite_differences(system):
    s, x = system
    P = x.shape[1]
    H = OO.ao_hessian(s, x)
    Hs = OO.shell_hessian(s, x)
    h = 1e-3
    for i in range(3):
        d_ao = np.stack([_fd(lambda y: oracle.ao_vgl(s, y.reshape(3, 1))[0], x[:, p].copy(), i, h)
                         for p in range(P)])  # [P][5][ao]
        ref = np.concatenate([d_ao[:, 1:4], d_ao[:, 4:5]], axis=1)  # d_i grad, d_i lap
        assert np.abs(H[i] - ref).max() <= 1e-8 * np.abs(ref).max(), i
        d_sh = np.stack([_fd(lambda y: oracle.shell_vgl(s, y.reshape(3, 1))[0], x[:, p].copy(), i, h)
                         for p in range(P)])
        ref_s = np.concatenate([d_sh[:, 1:4], d_sh[:, 4:5]], axis=1)
        assert np.abs(Hs[:, i] - ref_s).max() <= 1e-8 * np.abs(ref_s).max(), i


def test_ao_hessian_by_the_reference_protocol(system):
    """sum_A -d/dR_A of the AO gradient and Laplacian (org/qmckl_ao.org:9753-9830)."""
    s, x = system
    H = OO.ao_hessian(s, x)
    h = 1e-3
    for k in range(3):
        acc = 0.0
        for a in range(s.nucl_num):
            def f(Rak, a=a):
                t = copy.deepcopy(s)
                t.nucl_coord[k, a] = Rak[0]
                return oracle.ao_vgl(t, x)
            acc = acc - _fd(f, np.array([s.nucl_coord[k, a]]), 0, h)
        ref = acc[:, 1:5]
        assert np.abs(H[k] - ref).max() <= 1e-8 * np.abs(ref).max()


def test_orbital_forces_match_nuclear_finite_differences(system):
    s, x = system
    C = oracle.kept_mo_coef(s)
    fa = OO.forces_ao_value(s, x)
    fm = OO.forces_mo_value(s, x)
    fg = OO.forces_mo_g(s, x)
    fl = OO.forces_mo_l(s, x)
    # small steps: the screening (alpha r^2 > 36.7 for precision 53) drops h-shell values of
    # ~1e-10 (the r^5 factor), a jump a wider stencil straddles
    h = 2.5e-4
    for a in range(s.nucl_num):
        for n in range(3):
            def f(Rk, a=a, n=n):
                t = copy.deepcopy(s)
                t.nucl_coord[n, a] = Rk[0]
                return oracle.ao_vgl(t, x)
            dao = _fd(f, np.array([s.nucl_coord[n, a]]), 0, h)  # [P][5][ao]
            dmo = np.einsum("pck,ik->pci", dao, C)
            assert np.abs(fa[a][:, n] - dao[:, 0]).max() <= 1e-8 * np.abs(dao[:, 0]).max()
            assert np.abs(fm[a][:, n] - dmo[:, 0]).max() <= 1e-8 * np.abs(dmo[:, 0]).max()
            assert np.abs(fg[a, n] - dmo[:, 1:4]).max() <= 1e-8 * np.abs(dmo[:, 1:4]).max()
            assert np.abs(fl[a, n] - dmo[:, 4]).max() <= 1e-8 * np.abs(dmo[:, 4]).max()


def test_mo_hessian_matches_finite_differences(system):
    s, x = system
    Hm = OO.mo_hessian(s, x)
    C = oracle.kept_mo_coef(s)
    h = 1e-3
    for i in range(3):
        d = np.stack([_fd(lambda y: oracle.mo_vgl(C, oracle.ao_vgl(s, y.reshape(3, 1)))[0], x[:, p].copy(), i, h)
                      for p in range(x.shape[1])])
        assert np.abs(Hm[i] - d[:, 1:4]).max() <= 1e-8 * np.abs(d[:, 1:4]).max()


def _cusp_system(system):
    s, _ = system
    if s.name == "high-l":
        pytest.skip("the cusp checks run on the small system")
    r_cusp = np.full(s.nucl_num, 0.25)
    assert np.all(np.abs(2 * r_cusp * s.nucl_charge - 6) > 1e-6)
    return s, r_cusp, OO.cusp_params(s, r_cusp)


def test_cusp_fit_is_continuous_and_satisfies_the_cusp_condition(system):
    s, r_cusp, cp = _cusp_system(system)
    rng = np.random.default_rng(3)
    zero = np.zeros_like(cp)
    eta = OO.mo_vgl_cusp(s, s.nucl_coord, r_cusp, zero)[:, 0, :]
    for a in range(s.nucl_num):
        # cusp condition: d/dr of the spherical average at the nucleus is -Z times the value
        np.testing.assert_allclose(cp[a, 1], -s.nucl_charge[a] * (cp[a, 0] + eta[a]), rtol=1e-10,
                                   atol=1e-10 * np.abs(cp[a, 1]).max())
        for _ in range(3):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            pts = np.stack([s.nucl_coord[:, a] + r_cusp[a] * (1 + e) * u for e in (-1e-9, 1e-9)], axis=1)
            v = OO.mo_vgl_cusp(s, pts, r_cusp, cp)
            scale = np.abs(v).max(axis=(0, 2))
            assert np.all(np.abs(v[0] - v[1]).max(axis=1) <= 1e-6 * scale), (a, np.abs(v[0] - v[1]).max(axis=1), scale)
    # without the fit the MOs jump at the sphere
    u = np.array([0.0, 0.6, 0.8])
    pts = np.stack([s.nucl_coord[:, 0] + r_cusp[0] * (1 + e) * u for e in (-1e-9, 1e-9)], axis=1)
    v0 = OO.mo_vgl_cusp(s, pts, r_cusp, zero)
    assert np.abs(v0[0, 0] - v0[1, 0]).max() > 1e-3 * np.abs(v0[:, 0]).max()


def test_cusp_corrected_gradient_laplacian_and_hessian_are_derivatives(system):
    s, r_cusp, cp = _cusp_system(system)
    rng = np.random.default_rng(4)
    u = rng.normal(size=(3, 3))
    pts = s.nucl_coord[:, [0, 0, 1]] + 0.15 * u / np.linalg.norm(u, axis=0)  # inside the spheres
    v = OO.mo_vgl_cusp(s, pts, r_cusp, cp)
    Hm = OO.mo_hessian(s, pts, r_cusp, cp)
    h = 1e-4
    for p in range(3):
        f = lambda y: OO.mo_vgl_cusp(s, y.reshape(3, 1), r_cusp, cp)[0]
        lap = 0.0
        for i in range(3):
            d = _fd(f, pts[:, p].copy(), i, h)
            assert np.abs(v[p, 1 + i] - d[0]).max() <= 1e-7 * np.abs(d[0]).max()
            assert np.abs(Hm[i, p] - d[1:4]).max() <= 1e-6 * np.abs(d[1:4]).max()
            lap = lap + d[1 + i]
        assert np.abs(v[p, 4] - lap).max() <= 1e-6 * np.abs(lap).max()
