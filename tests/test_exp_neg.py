"""Accuracy of the e^x used by the radial parts of the AO kernels (exp_neg.hpp).

The header's functions are compiled for the host with hipcc and checked against exact
values (50-digit decimal arithmetic) and against numpy's exp over the argument range the
screening allows, x = -alpha r^2 in [-cutoff, 0] with cutoff = (p - 1) ln 2 <= 36.04.
"""
import ctypes
import os
import shutil
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.fixture(scope="module")
def exp_lib(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("expneg") / "libexpneg.so"
    subprocess.run([HIPCC, "-O2", "-fPIC", "-shared", os.path.join(HERE, "exp_neg_host.cpp"), "-o", str(out)],
                   check=True, capture_output=True)
    return ctypes.CDLL(str(out))


def _eval(lib, x, deg):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib.exp_neg_host(x.ctypes.data_as(ctypes.c_void_p), y.ctypes.data_as(ctypes.c_void_p),
                     ctypes.c_long(x.size), ctypes.c_int(deg))
    return y


def _args(n=400_000, seed=3):
    rng = np.random.default_rng(seed)
    edges = [0.0, -0.0, -1e-300, -1e-17, -0.34657359027997264, -0.3465735902799727, -36.04365338911715,
             -0.6931471805599453, -1.3862943611198906, -20.0, -36.0, -100.0, -708.0]
    return np.concatenate([-rng.uniform(0.0, 36.05, n), -rng.uniform(0.0, 1e-3, n // 10), edges])


def _exact_ulps(x, y):
    getcontext().prec = 50
    e = Decimal(float(x)).exp()
    return float(abs(Decimal(float(y)) - e) / Decimal(float(np.spacing(float(e)))))


def test_exp_neg_within_one_ulp(exp_lib):
    x = _args()
    y = _eval(exp_lib, x, 13)
    ref = np.exp(x)
    ulp = np.abs(y - ref) / np.spacing(ref)
    assert ulp.max() <= 1.0  # numpy's exp is itself within 1 ulp
    rng = np.random.default_rng(5)
    sample = np.concatenate([rng.choice(x.size, 1500, replace=False), np.argsort(ulp)[-200:],
                             np.arange(x.size - 13, x.size)])
    worst = max(_exact_ulps(x[i], y[i]) for i in sample)
    assert worst < 1.0, worst
    assert y[x.size - 13] == 1.0 and y[x.size - 12] == 1.0  # e^0, e^-0


def test_exp_neg_f32_relative_error(exp_lib):
    x = _args(seed=4)
    y = _eval(exp_lib, x, 9)
    ref = np.exp(x)
    rel = np.abs(y - ref) / ref
    assert rel.max() < 1e-11, rel.max()  # truncation bound |r|^10 / 10! = 7.1e-12
    assert rel.max() > 1e-14  # the degree-9 series is the one evaluated: its truncation shows
