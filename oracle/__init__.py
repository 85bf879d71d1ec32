"""CPU oracle -- TEST INFRASTRUCTURE ONLY.

ctypes binding of liboracle.so, a plain-C FP64 restatement of the reference's doc-path
algorithms (qmckl_oracle.c cites the reference file:line of every function).  Only the
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py use it, and
only as the checker; the product never does.

How each function is pinned (tests/test_oracle_*.py):
  ao_vgl, shell_vgl, prim_vgl  golden values of the reference tests on CHBrClF
                               (org/qmckl_ao.org:4704-4712, 5248-5252, 8400-8405, 9355-9384),
                               an independent numpy closed form (Gaussian x monomial), and
                               central finite differences of the value for the gradient
                               and Laplacian.
  mo_vgl                       numpy dense product ao_vgl . C^T (exact linear algebra).
  ee/en distances              golden values (org/qmckl_electron.org:1009-1018, 1458-1473)
                               and numpy norms.
  jastrow                      golden values on N2 for asymp_jasa/b, rescaled distances,
                               een_rescaled_e/_gl/_n/_gl, tmp_c, dtmp_c, factor_ee/en/een,
                               factor_een_gl (org/qmckl_jastrow_champ.org, 964 asserts in
                               total, tests/golden/goldens.json); an independent brute-force
                               numpy evaluation of J over (i, j, alpha) triples from the
                               defining formula; central finite differences of J in the
                               electron coordinates for the gradients and Laplacians; and
                               value == exp(J_ee + J_en + J_eeN).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_L = None
_P = ctypes.c_void_p
_I = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_cutoff.restype = _D
        L.oracle_cutoff.argtypes = [_I32]
        L.oracle_prim_vgl.argtypes = [_I, _I, _P, _P, _P, _I32, _I, _P, _P]
        L.oracle_shell_vgl.argtypes = [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I, _P, _P]
        L.oracle_ao_vgl.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I, _P,
                                    _P]
        L.oracle_mo_vgl.argtypes = [_I, _I, _I, _P, _P, _P]
        L.oracle_ee_distance.argtypes = [_I, _I, _P, _P]
        L.oracle_en_distance.argtypes = [_I, _I, _I, _P, _P, _P]
        L.oracle_dim_c_vector.restype = _I
        L.oracle_dim_c_vector.argtypes = [_I]
        L.oracle_lkpm.argtypes = [_I, _P]
        L.oracle_asymp_jasa.argtypes = [_I, _I, _P, _P, _P]
        L.oracle_asymp_jasb.argtypes = [_I, _P, _D, _I32, _P]
        L.oracle_jastrow.argtypes = [_I, _I, _I, _I, _P, _P, _I, _P, _I, _I, _I, _P, _P, _P, _D, _P, _I32] + [_P] * 18
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def _basis_args(s):
    return [
        _c(s.nucl_coord), _c(s.nucleus_index, np.int64), _c(s.nucleus_shell_num, np.int64),
        _c(s.shell_ang_mom, np.int32), _c(s.shell_prim_num, np.int64), _c(s.shell_prim_index, np.int64),
        _c(s.shell_factor), _c(s.exponent), _c(s.coefficient), _c(s.prim_factor),
        _c(np.zeros(s.shell_num) if getattr(s, "r_power", None) is None else s.r_power, np.int32),
    ]


def _slater(s):
    return 1 if getattr(s, "basis_type", "G") == "S" else 0


def ao_vgl(s, coord_T, precision=None):
    """AO VGL [P][5][ao] of the points coord_T [3][P]."""
    x = _c(coord_T)
    P = x.shape[1]
    out = np.empty((P, 5, s.ao_num))
    args = _basis_args(s)
    af = _c(s.ao_factor)
    prec = s.precision if precision is None else precision
    lib().oracle_ao_vgl(s.nucl_num, s.shell_num, s.prim_num, s.ao_num, *[_p(a) for a in args], _p(af),
                        prec, _slater(s), P, _p(x), _p(out))
    return out


def shell_vgl(s, coord_T, precision=None):
    x = _c(coord_T)
    P = x.shape[1]
    out = np.empty((P, 5, s.shell_num))
    args = _basis_args(s)
    prec = s.precision if precision is None else precision
    lib().oracle_shell_vgl(s.nucl_num, s.shell_num, s.prim_num, *[_p(a) for a in args], prec, _slater(s), P, _p(x),
                           _p(out))
    return out


def prim_vgl(s, coord_T, precision=None):
    x = _c(coord_T)
    P = x.shape[1]
    out = np.empty((P, 5, s.prim_num))
    npi = np.empty(s.nucl_num + 1, np.int64)
    for a in range(s.nucl_num):
        npi[a] = s.shell_prim_index[s.nucleus_index[a]]
    npi[s.nucl_num] = s.prim_num
    prec = s.precision if precision is None else precision
    nc, ex = _c(s.nucl_coord), _c(s.exponent)
    lib().oracle_prim_vgl(s.nucl_num, s.prim_num, _p(npi), _p(nc), _p(ex), prec, P, _p(x), _p(out))
    return out


def mo_vgl(coefficient, ao):
    """MO VGL [P][5][mo] from AO VGL [P][5][ao] and coefficient [mo][ao]."""
    c = _c(coefficient)
    a = _c(ao)
    P, _, nao = a.shape
    out = np.empty((P, 5, c.shape[0]))
    lib().oracle_mo_vgl(nao, c.shape[0], P, _p(c), _p(a), _p(out))
    return out


def kept_mo_coef(s):
    return s.mo_coef if s.keep_mo is None else s.mo_coef[np.asarray(s.keep_mo) != 0]


def ee_distance(coord_T, elec_num, walk_num):
    x = _c(coord_T)
    out = np.empty((walk_num, elec_num, elec_num))
    lib().oracle_ee_distance(elec_num, walk_num, _p(x), _p(out))
    return out


def en_distance(coord_T, nucl_coord):
    x = _c(coord_T)
    n = _c(nucl_coord)
    P = x.shape[1]
    out = np.empty((P, n.shape[1]))
    lib().oracle_en_distance(P, n.shape[1], 1, _p(x), _p(n), _p(out))
    return out


def dim_c_vector(cord):
    return int(lib().oracle_dim_c_vector(cord))


def lkpm(cord):
    d = dim_c_vector(cord)
    out = np.empty((4, max(d, 1)), np.int64)
    lib().oracle_lkpm(cord, _p(out))
    return out[:, :d]


JASTROW_FIELDS = ["ee_rescaled", "ee_rescaled_gl", "en_rescaled", "en_rescaled_gl", "een_e", "een_e_gl", "een_n",
                  "een_n_gl", "tmp_c", "dtmp_c", "factor_ee", "factor_ee_gl", "factor_en", "factor_en_gl",
                  "factor_een", "factor_een_gl", "value", "gl"]


def jastrow(s, coord_T, walk_num, want=("factor_ee", "factor_en", "factor_een", "factor_ee_gl", "factor_en_gl",
                                         "factor_een_gl", "value", "gl")):
    """CHAMP Jastrow of all walkers; returns a dict of the requested outputs."""
    W, Ne, Nn, C = walk_num, s.elec_num, s.nucl_num, s.cord
    shapes = {
        "ee_rescaled": (W, Ne, Ne), "ee_rescaled_gl": (W, Ne, Ne, 4), "en_rescaled": (W, Nn, Ne),
        "en_rescaled_gl": (W, Nn, Ne, 4), "een_e": (W, C + 1, Ne, Ne), "een_e_gl": (W, C + 1, Ne, 4, Ne),
        "een_n": (W, C + 1, Nn, Ne), "een_n_gl": (W, C + 1, Nn, 4, Ne), "tmp_c": (W, C, C + 1, Nn, Ne),
        "dtmp_c": (W, C, C + 1, Nn, 4, Ne), "factor_ee": (W,), "factor_ee_gl": (W, 4, Ne), "factor_en": (W,),
        "factor_en_gl": (W, 4, Ne), "factor_een": (W,), "factor_een_gl": (W, 4, Ne), "value": (W,),
        "gl": (W, 4, Ne),
    }
    out = {k: np.zeros(shapes[k]) for k in want}
    x = _c(coord_T)
    keep = [x, _c(s.nucl_coord), _c(s.type_nucl_vector, np.int64), _c(s.a_vector), _c(s.b_vector),
            _c(s.c_vector if s.cord > 1 else np.zeros(1)), _c(s.rescale_en)]
    lib().oracle_jastrow(W, Ne, s.up, Nn, _p(keep[0]), _p(keep[1]), s.type_nucl_num, _p(keep[2]), s.aord, s.bord,
                         s.cord, _p(keep[3]), _p(keep[4]), _p(keep[5]), float(s.rescale_ee), _p(keep[6]),
                         int(s.spin_independent), *[_p(out.get(k)) for k in JASTROW_FIELDS])
    return out


def asymp_jasa(s):
    out = np.empty(s.type_nucl_num)
    a, k = _c(s.a_vector), _c(s.rescale_en)
    lib().oracle_asymp_jasa(s.aord, s.type_nucl_num, _p(a), _p(k), _p(out))
    return out


def asymp_jasb(s):
    out = np.empty(2)
    b = _c(s.b_vector)
    lib().oracle_asymp_jasb(s.bord, _p(b), float(s.rescale_ee), int(s.spin_independent), _p(out))
    return out


# ----------------------------------------------------------------------------------------
# Parameter derivatives and nuclear forces of the Jastrow (qmckl_oracle_deriv.c)
# ----------------------------------------------------------------------------------------
DERIV_FIELDS = ["asymp_jasb_pderiv", "asymp_jasa_pderiv", "factor_ee_pderiv", "factor_ee_gl_pderiv",
                "factor_en_pderiv", "factor_en_gl_pderiv", "factor_een_pderiv", "factor_een_gl_pderiv",
                "forces_jastrow_en", "forces_jastrow_en_g", "forces_jastrow_en_l", "forces_tmp_c",
                "forces_dtmp_c", "forces_een_n", "forces_jastrow_een", "forces_jastrow_een_g",
                "forces_jastrow_een_l"]


def c_full(s):
    """c_vector_full [dim_c][Nn]: the c-coefficients of each nucleus's type."""
    d = dim_c_vector(s.cord)
    return np.ascontiguousarray(np.asarray(s.c_vector)[np.asarray(s.type_nucl_vector), :d].T)


def jastrow_derivatives(s, coord_T, walk_num):
    """Parameter derivatives and nuclear forces of the Jastrow, in the reference's C layouts
    (see qmckl_oracle_deriv.c); returns a dict keyed by DERIV_FIELDS."""
    L = lib()
    W, Ne, Nn, C, T = walk_num, s.elec_num, s.nucl_num, s.cord, s.type_nucl_num
    I = jastrow(s, coord_T, W, want=("ee_rescaled", "ee_rescaled_gl", "en_rescaled", "en_rescaled_gl", "een_e",
                                      "een_e_gl", "een_n", "een_n_gl", "tmp_c", "dtmp_c"))
    edist = en_distance(coord_T, s.nucl_coord)
    typ = _c(s.type_nucl_vector, np.int64)
    a, b, kap = _c(s.a_vector), _c(s.b_vector), _c(s.rescale_en)
    d = dim_c_vector(C)
    lk = np.ascontiguousarray(lkpm(C)) if C >= 2 else np.zeros((4, 1), np.int64)
    cf = c_full(s) if C >= 2 else np.zeros((1, Nn))
    o = {
        "asymp_jasb_pderiv": np.zeros((2, s.bord + 1)), "asymp_jasa_pderiv": np.zeros((T, s.aord + 1)),
        "factor_ee_pderiv": np.zeros(s.bord + 1), "factor_ee_gl_pderiv": np.zeros((s.bord + 1, Ne, 4)),
        "factor_en_pderiv": np.zeros((T, s.aord + 1)), "factor_en_gl_pderiv": np.zeros((T, s.aord + 1, Ne, 4)),
        "factor_een_pderiv": np.zeros((T, max(d, 1))), "factor_een_gl_pderiv": np.zeros((T, max(d, 1), Ne, 4)),
        "forces_jastrow_en": np.zeros((W, Nn, 3)), "forces_jastrow_en_g": np.zeros((W, Nn, 3, Ne, 3)),
        "forces_jastrow_en_l": np.zeros((W, Nn, 3)), "forces_tmp_c": np.zeros((W, C, C + 1, Nn, 4, Ne)),
        "forces_dtmp_c": np.zeros((W, C, C + 1, 4, Nn, 4, Ne)), "forces_een_n": np.zeros((W, C + 1, 3, Nn, 4, Ne)),
        "forces_jastrow_een": np.zeros((W, Nn, 3)), "forces_jastrow_een_g": np.zeros((W, 3, Nn, 3, Ne)),
        "forces_jastrow_een_l": np.zeros((W, 3, Nn)),
    }
    def P(arr):  # explicit pointer: these functions have no declared argtypes
        return ctypes.c_void_p(arr.ctypes.data)

    L.oracle_asymp_jasb_pderiv(_I(s.bord), P(b), _D(s.rescale_ee), _I32(int(s.spin_independent)),
                               P(o["asymp_jasb_pderiv"]))
    L.oracle_asymp_jasa_pderiv(_I(s.aord), _I(T), P(a), P(kap), P(o["asymp_jasa_pderiv"]))
    L.oracle_factor_ee_pderiv(_I(W), _I(Ne), _I(s.up), _I(s.bord), P(b), P(I["ee_rescaled"]),
                              P(o["asymp_jasb_pderiv"]), _I32(int(s.spin_independent)), P(o["factor_ee_pderiv"]))
    L.oracle_factor_ee_gl_pderiv(_I(W), _I(Ne), _I(s.up), _I(s.bord), P(b), P(I["ee_rescaled"]),
                                 P(I["ee_rescaled_gl"]), _I32(int(s.spin_independent)), P(o["factor_ee_gl_pderiv"]))
    L.oracle_factor_en_pderiv(_I(W), _I(Ne), _I(Nn), _I(T), P(typ), _I(s.aord), P(a), P(I["en_rescaled"]),
                              P(o["asymp_jasa_pderiv"]), P(o["factor_en_pderiv"]))
    L.oracle_factor_en_gl_pderiv(_I(W), _I(Ne), _I(Nn), _I(T), P(typ), _I(s.aord), P(a), P(I["en_rescaled"]),
                                 P(I["en_rescaled_gl"]), P(o["factor_en_gl_pderiv"]))
    L.oracle_forces_jastrow_en(_I(W), _I(Ne), _I(Nn), P(typ), _I(s.aord), P(a), P(I["en_rescaled"]),
                               P(I["en_rescaled_gl"]), P(o["forces_jastrow_en"]))
    L.oracle_forces_jastrow_en_g(_I(W), _I(Ne), _I(Nn), P(typ), _I(s.aord), P(a), P(kap), P(edist),
                                 P(I["en_rescaled"]), P(I["en_rescaled_gl"]), P(o["forces_jastrow_en_g"]))
    L.oracle_forces_jastrow_en_l(_I(W), _I(Ne), _I(Nn), P(typ), _I(s.aord), P(a), P(kap), P(edist),
                                 P(I["en_rescaled"]), P(I["en_rescaled_gl"]), P(o["forces_jastrow_en_l"]))
    if C >= 2:
        L.oracle_factor_een_pderiv(_I(W), _I(Ne), _I(Nn), _I(T), P(typ), _I(C), _I(d), P(lk), P(I["tmp_c"]),
                                   P(I["een_n"]), P(o["factor_een_pderiv"]))
        L.oracle_factor_een_gl_pderiv(_I(W), _I(Ne), _I(Nn), _I(T), P(typ), _I(C), _I(d), P(lk), P(I["tmp_c"]),
                                      P(I["dtmp_c"]), P(I["een_n"]), P(I["een_n_gl"]), P(o["factor_een_gl_pderiv"]))
        L.oracle_forces_tmp_c(_I(W), _I(Ne), _I(Nn), _I(C), P(I["een_e"]), P(I["een_n_gl"]), P(o["forces_tmp_c"]))
        L.oracle_forces_dtmp_c(_I(W), _I(Ne), _I(Nn), _I(C), P(I["een_e_gl"]), P(I["een_n_gl"]),
                               P(o["forces_dtmp_c"]))
        L.oracle_forces_een_n(_I(W), _I(Ne), _I(Nn), P(typ), _I(C), P(kap), P(edist), P(I["een_n"]),
                              P(I["een_n_gl"]), P(o["forces_een_n"]))
        L.oracle_forces_jastrow_een(_I(W), _I(Ne), _I(Nn), _I(C), _I(d), P(cf), P(lk), P(I["een_n"]),
                                    P(I["een_n_gl"]), P(I["tmp_c"]), P(o["forces_tmp_c"]), P(o["forces_jastrow_een"]))
        args = (P(cf), P(lk), P(I["tmp_c"]), P(I["dtmp_c"]), P(o["forces_tmp_c"]), P(o["forces_dtmp_c"]),
                P(o["forces_een_n"]), P(I["een_n"]), P(I["een_n_gl"]))
        L.oracle_forces_jastrow_een_g(_I(W), _I(Ne), _I(Nn), _I(C), _I(d), *args, P(o["forces_jastrow_een_g"]))
        L.oracle_forces_jastrow_een_l(_I(W), _I(Ne), _I(Nn), _I(C), _I(d), *args, P(o["forces_jastrow_een_l"]))
    o["factor_een_pderiv"] = o["factor_een_pderiv"][:, :d]
    o["factor_een_gl_pderiv"] = o["factor_een_gl_pderiv"][:, :d]
    return o
