"""Pins the CPU oracle on every golden value of the reference's own tests.

The values and tolerances are those of the reference's assert statements
(tests/golden/goldens.json, extracted from org/qmckl_ao.org, org/qmckl_electron.org and
org/qmckl_jastrow_champ.org with the file and line of each assert).
"""
import numpy as np
import pytest

import oracle
import systems


def _check(var, arr, file):
    gs = systems.goldens(var, file)
    assert gs, var
    for g in gs:
        v = arr[tuple(g["idx"])]
        assert abs(v - g["value"]) < g["tol"], (var, g["line"], g["idx"], v, g["value"])
    return len(gs)


def test_ao_goldens_chbrclf():
    s, x, W = systems.chbrclf()
    pts = x[:, : s.elec_num]  # the reference sets the 68 points of walker 0 (org/qmckl_ao.org:9270-9280)
    ao = oracle.ao_vgl(s, pts)
    n = _check("ao_vgl", ao, "qmckl_ao.org")
    n += _check("ao_value", ao[:, 0, :], "qmckl_ao.org")
    n += _check("shell_vgl", oracle.shell_vgl(s, pts), "qmckl_ao.org")
    n += _check("prim_vgl", oracle.prim_vgl(s, pts), "qmckl_ao.org")
    assert n == 46


def test_distance_goldens_chbrclf():
    s, x, W = systems.chbrclf()
    ee = oracle.ee_distance(x, s.elec_num, W)
    en = oracle.en_distance(x, s.nucl_coord).reshape(W, s.elec_num, s.nucl_num)
    _check("ee_distance", ee.ravel(), "qmckl_electron.org")
    _check("en_distance", en, "qmckl_electron.org")


JAS = {
    "ee_distance_rescaled": lambda J: J["ee_rescaled"].ravel(),
    "en_distance_rescaled": lambda J: J["en_rescaled"],
    "een_rescaled_e": lambda J: J["een_e"],
    "een_rescaled_e_gl": lambda J: J["een_e_gl"],
    "een_rescaled_n": lambda J: J["een_n"],
    "een_rescaled_n_gl": lambda J: J["een_n_gl"],
    "tmp_c": lambda J: J["tmp_c"],
    "dtmp_c": lambda J: J["dtmp_c"],
    "factor_ee": lambda J: J["factor_ee"],
    "factor_en": lambda J: J["factor_en"],
    "factor_een": lambda J: J["factor_een"],
    "factor_een_gl": lambda J: J["factor_een_gl"],
}


@pytest.mark.parametrize("var", sorted(JAS))
def test_jastrow_goldens_n2(var):
    s, x, W = systems.n2()
    J = oracle.jastrow(s, x, W, want=oracle.JASTROW_FIELDS)
    _check(var, JAS[var](J), "qmckl_jastrow_champ.org")


def test_asymptotic_goldens_n2():
    s, x, W = systems.n2()
    _check("asymp_jasa", oracle.asymp_jasa(s), "qmckl_jastrow_champ.org")
    _check("asymp_jasb", oracle.asymp_jasb(s), "qmckl_jastrow_champ.org")


def test_lkpm_and_c_full_match_n2_header():
    # include/n2.h holds the reference's (l, k, p, m) table, stored c-index by c-index.
    d = systems.data()
    ref = d["n2/n2_lkpm_of_cindex"].reshape(23, 4).T.astype(np.int64)
    assert oracle.dim_c_vector(5) == 23
    np.testing.assert_array_equal(oracle.lkpm(5), ref)
