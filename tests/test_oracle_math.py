"""Pins the oracle beyond the reference's golden values, with checks that do not restate the
oracle's own formulas:

* AO values against an independent numpy closed form (contracted Gaussian times the cartesian
  monomial, in the reference's polynomial order), and AO gradients/Laplacians against
  central finite differences of the oracle's own values;
* MO VGL against a dense numpy product;
* distances against numpy norms;
* the e-e-n Jastrow value against a brute-force triple sum over (i, j, nucleus) of the
  defining CHAMP expression, and every Jastrow gradient/Laplacian (e-e, e-n, e-e-n, total)
  against central finite differences of the corresponding value;
* value == exp(J) (0 below J = -100) and gl == e^J (grad J, lap J + |grad J|^2).

Each check is chosen so that a plausible slip in the oracle (a wrong polynomial order or
power, a dropped normalisation factor, a sign or index error in a derivative, a missing
e-e-n term) fails it.
"""

import numpy as np
import pytest

import oracle
from qmckl_amd import system as S


# ------------------------------------------------------------------------ AO closed form
def _ao_closed_form(s, x):
    """AO values [P][ao] from the definition: ao_factor * x^a y^b z^c * shell_factor *
    sum_p coefficient_p prim_factor_p exp(-alpha_p r^2), cartesian order a = l..0,
    b = l-a..0 for each shell (org/qmckl_ao.org, polynomial order of :6270-6300), with the
    primitives where alpha r^2 > -log(2^(1 - precision)) left out, the screening rule of
    org/qmckl_ao.org:118-120 (its prim_factor grows as alpha^((2l+3)/4), so beyond the
    f shells a screened tight primitive is not negligible against its shell)."""
    cutoff = -np.log(2.0 ** (1 - s.precision))
    P = x.shape[1]
    out = np.zeros((P, s.ao_num))
    rp = np.zeros(s.shell_num, np.int64) if s.r_power is None else np.asarray(s.r_power, np.int64)
    k = 0
    for a in range(s.nucl_num):
        R = s.nucl_coord[:, a][:, None]
        d = x - R
        r2 = (d * d).sum(0)
        shells = range(s.nucleus_index[a], s.nucleus_index[a] + s.nucleus_shell_num[a])
        # nucleus range: max over its primitives of (r_power + 1) / alpha
        # (org/qmckl_ao.org:3195-3205); points beyond r^2 > cutoff * range see no AO of it
        rng_a = max([0.0] + [(rp[sh] + 1) / s.exponent[p] for sh in shells
                             for p in range(s.shell_prim_index[sh], s.shell_prim_index[sh] + s.shell_prim_num[sh])])
        far = r2 > cutoff * rng_a
        for sh in shells:
            l = int(s.shell_ang_mom[sh])
            p0, np_ = int(s.shell_prim_index[sh]), int(s.shell_prim_num[sh])
            rad = np.zeros(P)
            for p in range(p0, p0 + np_):
                ar2 = s.exponent[p] * r2
                rad += np.where(ar2 > cutoff, 0.0, s.coefficient[p] * s.prim_factor[p] * np.exp(-ar2))
            rad *= s.shell_factor[sh]
            rad[far] = 0.0
            for pa in range(l, -1, -1):
                for pb in range(l - pa, -1, -1):
                    pc = l - pa - pb
                    out[:, k] = s.ao_factor[k] * d[0] ** pa * d[1] ** pb * d[2] ** pc * rad
                    k += 1
    assert k == s.ao_num
    return out


def _points(s, n, seed):
    rng = np.random.default_rng(seed)
    c = s.nucl_coord[:, rng.integers(0, s.nucl_num, n)]
    return np.ascontiguousarray(c + rng.normal(0.0, 0.8, (3, n)))


def _ao_system(name):
    """A synthetic workload, or the s-to-h-shell system with AO factors off 1."""
    if name == "high_l":
        import systems

        return systems.high_l(walk_num=1)[0]
    return S.workload(name, walk_num=1)[0]


@pytest.mark.parametrize("name", ["small", "medium", "high_l"])
def test_ao_values_match_closed_form(name):
    s = _ao_system(name)
    s.precision = 53
    x = _points(s, 24, 1)
    got = oracle.ao_vgl(s, x)[:, 0, :]
    ref = _ao_closed_form(s, x)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * scale)


def test_ao_values_with_r_power_match_closed_form():
    """r_power enters the nucleus range (r_power + 1) / alpha: r_power = -1 on every shell of
    the first nucleus gives it range 0 (its AOs vanish at every point off the nucleus),
    larger powers on the second nucleus widen its range (no change: the primitive screening
    already stops within the r_power = 0 range)."""
    s = _ao_system("small")
    s.precision = 53
    rp = np.zeros(s.shell_num, np.int32)
    n0 = slice(s.nucleus_index[0], s.nucleus_index[0] + s.nucleus_shell_num[0])
    n1 = slice(s.nucleus_index[1], s.nucleus_index[1] + s.nucleus_shell_num[1])
    rp[n0] = -1
    rp[n1] = 2
    s.r_power = rp
    x = _points(s, 24, 3)
    got = oracle.ao_vgl(s, x)
    ref = _ao_closed_form(s, x)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got[:, 0, :], ref, rtol=1e-12, atol=1e-14 * scale)
    nao0 = sum((l + 1) * (l + 2) // 2 for l in s.shell_ang_mom[n0])
    assert np.all(got[:, :, :nao0] == 0.0) and np.abs(got[:, 0, nao0:]).max() > 0.0
    s.r_power = None
    base = oracle.ao_vgl(s, x)
    np.testing.assert_array_equal(got[:, :, nao0:], base[:, :, nao0:])


@pytest.mark.parametrize("name", ["small", "high_l"])
def test_ao_gradient_and_laplacian_match_finite_differences(name):
    s = _ao_system(name)
    x = _points(s, 12, 2)
    vgl = oracle.ao_vgl(s, x)
    h = 1e-4
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        fp = oracle.ao_vgl(s, x + e)[:, 0, :]
        fm = oracle.ao_vgl(s, x - e)[:, 0, :]
        fd = (fp - fm) / (2 * h)
        scale = np.abs(vgl[:, 1 + d, :]).max()
        assert np.abs(fd - vgl[:, 1 + d, :]).max() <= 1e-6 * scale, f"d/d{'xyz'[d]}"
    h = 2e-3
    lap = -6.0 * vgl[:, 0, :]
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        lap = lap + oracle.ao_vgl(s, x + e)[:, 0, :] + oracle.ao_vgl(s, x - e)[:, 0, :]
    lap /= h * h
    scale = np.abs(vgl[:, 4, :]).max()
    assert np.abs(lap - vgl[:, 4, :]).max() <= 2e-4 * scale


def test_mo_vgl_is_the_dense_product():
    s, x, W = S.workload("small", walk_num=4)
    ao = oracle.ao_vgl(s, x)
    coef = oracle.kept_mo_coef(s)
    got = oracle.mo_vgl(coef, ao)
    ref = np.einsum("pck,jk->pcj", ao, coef)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


def test_distances_are_norms():
    s, x, W = S.workload("small", walk_num=3)
    Ne = s.elec_num
    ee = oracle.ee_distance(x, Ne, W)
    en = oracle.en_distance(x, s.nucl_coord)
    xw = x.reshape(3, W, Ne)
    ref_ee = np.linalg.norm(xw[:, :, :, None] - xw[:, :, None, :], axis=0)  # [W][j][i]
    np.testing.assert_allclose(ee, ref_ee.transpose(0, 2, 1), rtol=1e-14, atol=1e-14)
    ref_en = np.linalg.norm(x[:, :, None] - s.nucl_coord[:, None, :], axis=0)
    np.testing.assert_allclose(en, ref_en, rtol=1e-14)


# ------------------------------------------------------------------------ Jastrow
def _een_brute_force(s, x, W):
    """J_eeN from its defining sum: over nuclei a, ordered pairs i != j and the terms
    (l, k, m) of dimension dim_c, c_{n,type(a)} * g_e(r_ij)^k * g_a(r_j)^m * g_a(r_i)^(m+l),
    with g_e = exp(-kappa_ee r), g_a = exp(-kappa_a R) (org/qmckl_jastrow_champ.org,
    factor_een_naive :14019)."""
    Ne, Nn, C = s.elec_num, s.nucl_num, s.cord
    terms = []
    for p in range(2, C + 1):
        for k in range(p - 1, -1, -1):
            lmax = p - k if k != 0 else p - k - 2
            for l in range(lmax, -1, -1):
                if (p - k - l) % 2 == 0:
                    terms.append((l, k, (p - k - l) // 2))
    xw = x.reshape(3, W, Ne)
    J = np.zeros(W)
    kap = s.rescale_en[s.type_nucl_vector]
    for w in range(W):
        r = np.linalg.norm(xw[:, w, :, None] - xw[:, w, None, :], axis=0)
        G = np.exp(-s.rescale_ee * r)
        off = ~np.eye(Ne, dtype=bool)
        R = np.linalg.norm(xw[:, w, :, None] - s.nucl_coord[:, None, :], axis=0)  # [i][a]
        g = np.exp(-kap[None, :] * R)
        for n, (l, k, m) in enumerate(terms):
            Gk = np.where(off, G ** k, 0.0)
            c = s.c_vector[s.type_nucl_vector, n]  # [a]
            # sum_{i,j} Gk[i,j] g[j,a]^m g[i,a]^(m+l)
            J[w] += np.einsum("ij,ja,ia,a->", Gk, g ** m, g ** (m + l), c)
    return J


def _ee_brute_force(s, x, W):
    """J_ee from its definition (org/qmckl_jastrow_champ.org:2340-2380, 3582-3650): over
    pairs i < j with f = (1 - exp(-kappa r_ij)) / kappa,
    s_ij b0 f / (1 + b1 f) + sum_{k>=2} b_k f^k - (s_ij b0 / kappa / (1 + b1 / kappa)
    + sum_{k>=2} b_k / kappa^k), where s_ij = 1/2 for a same-spin pair unless the Jastrow is
    spin independent, else 1 (the shift makes each pair term vanish as r -> infinity)."""
    Ne, up, kap, b = s.elec_num, s.up, s.rescale_ee, s.b_vector
    xw = x.reshape(3, W, Ne)
    J = np.zeros(W)
    for w in range(W):
        for j in range(Ne):
            for i in range(j):
                f = (1.0 - np.exp(-kap * np.linalg.norm(xw[:, w, i] - xw[:, w, j]))) / kap
                same = (j < up) or (i >= up)
                sij = 1.0 if (s.spin_independent == 1 or not same) else 0.5
                term = sij * b[0] * f / (1.0 + b[1] * f) + sum(b[k] * f ** k for k in range(2, s.bord + 1))
                shift = sij * b[0] / kap / (1.0 + b[1] / kap) + sum(b[k] / kap ** k for k in range(2, s.bord + 1))
                J[w] += term - shift
    return J


@pytest.mark.parametrize("spin_independent", [0, 1])
def test_ee_value_matches_brute_force(spin_independent):
    s, x, W = S.jastrow_workload(3, 4, 3, 5, 5, 5, walk_num=3)
    s.spin_independent = spin_independent
    got = oracle.jastrow(s, x, W, want=("factor_ee",))["factor_ee"]
    np.testing.assert_allclose(got, _ee_brute_force(s, x, W), rtol=1e-12, atol=1e-12)


def test_ee_gl_spin_independent_matches_finite_differences():
    s, x, W = S.jastrow_workload(3, 3, 2, 5, 5, 5, walk_num=2)
    s.spin_independent = 1
    got = oracle.jastrow(s, x, W, want=("factor_ee_gl",))["factor_ee_gl"]
    fd = _fd_gl(s, x, W, "factor_ee")
    for c, tol in ((0, 1e-6), (1, 1e-6), (2, 1e-6), (3, 1e-4)):
        scale = max(np.abs(got[:, c]).max(), 1e-3)
        assert np.abs(fd[:, c] - got[:, c]).max() <= tol * scale, f"component {c}"


@pytest.mark.parametrize("cfg", [(3, 3, 2, 5, 5, 5), (4, 4, 3, 3, 4, 3), (2, 3, 3, 5, 5, 7)])
def test_een_value_matches_brute_force(cfg):
    s, x, W = S.jastrow_workload(*cfg, walk_num=3)
    got = oracle.jastrow(s, x, W, want=("factor_een",))["factor_een"]
    ref = _een_brute_force(s, x, W)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12)


def _fd_gl(s, x, W, key):
    """Central-difference gradient and Laplacian of the per-walker value `key` w.r.t. each
    electron, [W][4][Ne]."""
    Ne = s.elec_num
    f0 = oracle.jastrow(s, x, W, want=(key,))[key]
    out = np.zeros((W, 4, Ne))
    h1, h2 = 1e-5, 1e-3
    xw = x.reshape(3, W, Ne)
    for i in range(Ne):
        lap = -6.0 * f0
        for d in range(3):
            vals = {}
            for h, sgn in ((h1, 1), (h1, -1), (h2, 1), (h2, -1)):
                y = xw.copy()
                y[d, :, i] += sgn * h
                vals[(h, sgn)] = oracle.jastrow(s, np.ascontiguousarray(y.reshape(3, W * Ne)), W, want=(key,))[key]
            out[:, d, i] = (vals[(h1, 1)] - vals[(h1, -1)]) / (2 * h1)
            lap = lap + vals[(h2, 1)] + vals[(h2, -1)]
        out[:, 3, i] = lap / (h2 * h2)
    return out


@pytest.mark.parametrize("key", ["factor_ee", "factor_en", "factor_een"])
def test_jastrow_gl_matches_finite_differences(key):
    s, x, W = S.jastrow_workload(3, 3, 2, 5, 5, 5, walk_num=2)
    got = oracle.jastrow(s, x, W, want=(key + "_gl",))[key + "_gl"]
    fd = _fd_gl(s, x, W, key)
    for c, tol in ((0, 1e-6), (1, 1e-6), (2, 1e-6), (3, 1e-4)):
        scale = max(np.abs(got[:, c]).max(), 1e-3)
        assert np.abs(fd[:, c] - got[:, c]).max() <= tol * scale, f"{key} component {c}"


def test_jastrow_value_and_gl_identities():
    s, x, W = S.jastrow_workload(3, 3, 2, 5, 5, 5, walk_num=4)
    r = oracle.jastrow(s, x, W)
    J = r["factor_ee"] + r["factor_en"] + r["factor_een"]
    np.testing.assert_allclose(r["value"], np.where(J < -100.0, 0.0, np.exp(J)), rtol=1e-14)
    g = r["factor_ee_gl"] + r["factor_en_gl"] + r["factor_een_gl"]
    ref = g.copy()
    ref[:, 3] = g[:, 3] + (g[:, :3] ** 2).sum(1)
    np.testing.assert_allclose(r["gl"], ref * r["value"][:, None, None], rtol=1e-13, atol=1e-300)
    # and the total GL is the finite-difference GL of e^J
    fd = _fd_gl(s, x, W, "value")
    for c, tol in ((0, 1e-6), (1, 1e-6), (2, 1e-6), (3, 1e-4)):
        scale = np.abs(r["gl"][:, c]).max()
        assert np.abs(fd[:, c] - r["gl"][:, c]).max() <= tol * scale, f"gl component {c}"


def test_rescaled_distance_gl_match_finite_differences():
    """ee_rescaled_gl[w][j][i][:] and en_rescaled_gl[w][a][i][:] are the gradient and
    Laplacian, in the coordinates of electron i, of the rescaled distances
    f(r) = (1 - exp(-kappa r)) / kappa of ee_rescaled[w][j][i] and en_rescaled[w][a][i]
    (org/qmckl_jastrow_champ.org:3077-3120, 6601-6700; org/qmckl_distance.org:1345-1509)."""
    s, x, W = S.jastrow_workload(3, 3, 2, 5, 5, 5, walk_num=2)
    Ne = s.elec_num
    keys = ("ee_rescaled", "en_rescaled", "ee_rescaled_gl", "en_rescaled_gl")
    r = oracle.jastrow(s, x, W, want=keys)
    xw = x.reshape(3, W, Ne)
    h1, h2 = 1e-5, 1e-3

    def at(y):
        o = oracle.jastrow(s, np.ascontiguousarray(y.reshape(3, W * Ne)), W, want=keys[:2])
        return o["ee_rescaled"], o["en_rescaled"]

    for i in range(Ne):
        lap_ee = -6.0 * r["ee_rescaled"][:, :, i]
        lap_en = -6.0 * r["en_rescaled"][:, :, i]
        for d in range(3):
            v = {}
            for h, sg in ((h1, 1), (h1, -1), (h2, 1), (h2, -1)):
                y = xw.copy()
                y[d, :, i] += sg * h
                v[(h, sg)] = at(y)
            fd_ee = (v[(h1, 1)][0] - v[(h1, -1)][0]) / (2 * h1)
            fd_en = (v[(h1, 1)][1] - v[(h1, -1)][1]) / (2 * h1)
            off = np.arange(Ne) != i
            np.testing.assert_allclose(fd_ee[:, off, i], r["ee_rescaled_gl"][:, off, i, d], atol=1e-8)
            np.testing.assert_allclose(fd_en[:, :, i], r["en_rescaled_gl"][:, :, i, d], atol=1e-8)
            lap_ee = lap_ee + v[(h2, 1)][0][:, :, i] + v[(h2, -1)][0][:, :, i]
            lap_en = lap_en + v[(h2, 1)][1][:, :, i] + v[(h2, -1)][1][:, :, i]
        off = np.arange(Ne) != i
        np.testing.assert_allclose((lap_ee / h2 ** 2)[:, off], r["ee_rescaled_gl"][:, off, i, 3], atol=2e-5)
        np.testing.assert_allclose(lap_en / h2 ** 2, r["en_rescaled_gl"][:, :, i, 3], atol=2e-5)


# ------------------------------------------------------------------------ Slater shells
def _slater_closed_form(s, x):
    """AO values [P][ao] of a Slater basis from the definition: ao_factor * x^a y^b z^c *
    shell_factor * sum_p coefficient_p prim_factor_p r^n e^(-alpha_p r) (n = r_power of the
    shell), with the screening of org/qmckl_ao.org: the AOs of a nucleus vanish where
    r^2 > cutoff * range (the AO assembly, :8515) or r > cutoff * range (the Slater shells,
    :4970), range = max (r_power + 1) / alpha; a primitive is left out where alpha r > cutoff."""
    cutoff = -np.log(2.0 ** (1 - s.precision))
    P = x.shape[1]
    out = np.zeros((P, s.ao_num))
    k = 0
    for a in range(s.nucl_num):
        d = x - s.nucl_coord[:, a][:, None]
        r = np.sqrt((d * d).sum(0))
        shells = range(s.nucleus_index[a], s.nucleus_index[a] + s.nucleus_shell_num[a])
        rng_a = max((s.r_power[sh] + 1) / s.exponent[p] for sh in shells
                    for p in range(s.shell_prim_index[sh], s.shell_prim_index[sh] + s.shell_prim_num[sh]))
        far = (r * r > cutoff * rng_a) | (r > cutoff * rng_a)
        for sh in shells:
            l, n = int(s.shell_ang_mom[sh]), int(s.r_power[sh])
            rad = np.zeros(P)
            for p in range(s.shell_prim_index[sh], s.shell_prim_index[sh] + s.shell_prim_num[sh]):
                ar = s.exponent[p] * r
                rad += np.where(ar > cutoff, 0.0, s.coefficient[p] * s.prim_factor[p] * r ** n * np.exp(-ar))
            rad *= s.shell_factor[sh]
            rad[far] = 0.0
            for pa in range(l, -1, -1):
                for pb in range(l - pa, -1, -1):
                    pc = l - pa - pb
                    out[:, k] = s.ao_factor[k] * d[0] ** pa * d[1] ** pb * d[2] ** pc * rad
                    k += 1
    return out


@pytest.mark.parametrize("precision", [53, 30])
def test_slater_ao_values_match_closed_form(precision):
    s, x, W = S.slater_workload(walk_num=4)
    s.precision = precision
    pts = np.concatenate([x, _points(s, 24, 5)], axis=1)
    got = oracle.ao_vgl(s, pts)[:, 0, :]
    ref = _slater_closed_form(s, pts)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_slater_ao_gradient_and_laplacian_match_finite_differences():
    s, x, W = S.slater_workload(walk_num=2)
    vgl = oracle.ao_vgl(s, x)
    h = 1e-4
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        fd = (oracle.ao_vgl(s, x + e)[:, 0, :] - oracle.ao_vgl(s, x - e)[:, 0, :]) / (2 * h)
        assert np.abs(fd - vgl[:, 1 + d, :]).max() <= 1e-6 * np.abs(vgl[:, 1 + d, :]).max(), f"d/d{'xyz'[d]}"
    h = 2e-3
    lap = -6.0 * vgl[:, 0, :]
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        lap = lap + oracle.ao_vgl(s, x + e)[:, 0, :] + oracle.ao_vgl(s, x - e)[:, 0, :]
    lap /= h * h
    assert np.abs(lap - vgl[:, 4, :]).max() <= 2e-4 * np.abs(vgl[:, 4, :]).max()


def test_slater_shells_at_the_reference_test_point():
    """The inputs of the reference's own Slater test (org/qmckl_ao.org:4284-4297): X =
    (1.1, 2.2, 3.3), R = (0.1, 1.2, -2.3), five s shells with r^i e^(-0.5 i r), i = 1..5:
    shell values from numpy, gradients and Laplacians by central finite differences of the
    oracle's values."""
    s, _, _ = S.slater_workload(walk_num=1)
    five = dict(nucl_charge=np.array([1.0]), nucl_coord=np.array([[0.1], [1.2], [-2.3]]),
                shell_ang_mom=np.zeros(5, np.int32), shell_prim_num=np.ones(5, np.int64),
                shell_prim_index=np.arange(5, dtype=np.int64), nucleus_index=np.zeros(1, np.int64),
                nucleus_shell_num=np.array([5], np.int64), shell_factor=np.ones(5),
                exponent=0.5 * np.arange(1, 6), coefficient=np.ones(5), prim_factor=np.ones(5),
                ao_factor=np.ones(5), r_power=np.arange(1, 6, dtype=np.int32))
    t = S.System(up=1, dn=0, mo_coef=np.eye(5), type_nucl_vector=np.zeros(1, np.int64), a_vector=s.a_vector[:1],
                 b_vector=s.b_vector, c_vector=s.c_vector[:1], rescale_en=s.rescale_en[:1], basis_type="S",
                 precision=53, **five)
    X = np.array([[1.1], [2.2], [3.3]])
    r = float(np.linalg.norm(X[:, 0] - t.nucl_coord[:, 0]))
    sh = oracle.shell_vgl(t, X)[0]
    np.testing.assert_allclose(sh[0], r ** np.arange(1, 6) * np.exp(-0.5 * np.arange(1, 6) * r), rtol=1e-14)
    h = 1e-5
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        fd = (oracle.shell_vgl(t, X + e)[0, 0] - oracle.shell_vgl(t, X - e)[0, 0]) / (2 * h)
        np.testing.assert_allclose(sh[1 + d], fd, rtol=1e-8)
    h = 1e-3
    lap = -6.0 * sh[0]
    for d in range(3):
        e = np.zeros((3, 1))
        e[d] = h
        lap = lap + oracle.shell_vgl(t, X + e)[0, 0] + oracle.shell_vgl(t, X - e)[0, 0]
    np.testing.assert_allclose(sh[4], lap / h ** 2, rtol=1e-5)


def test_slater_shells_on_a_nucleus():
    """At r = 0 (org/qmckl_ao.org:4985-4995): a shell with r_power 0 has the sum of its
    normalised coefficients as value, every other shell 0; gradients and Laplacians 0."""
    s, _, _ = S.slater_workload(walk_num=1)
    X = np.ascontiguousarray(s.nucl_coord[:, :1])
    sh = oracle.shell_vgl(s, X)[0]
    for q in range(s.nucleus_shell_num[0]):
        p0, npr = s.shell_prim_index[q], s.shell_prim_num[q]
        want = (s.coefficient[p0:p0 + npr] * s.prim_factor[p0:p0 + npr]).sum() if s.r_power[q] == 0 else 0.0
        assert sh[0, q] == pytest.approx(want * s.shell_factor[q], rel=1e-15)
        assert np.all(sh[1:, q] == 0.0)
