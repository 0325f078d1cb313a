"""The FP64 erfc/exp pair of the kinetic fluxes (mage_amd/csrc/gam_erfc.h),
compiled for the host from the same header, against 40-digit mpmath values.

The reference evaluates erfc and exp with its compiler's libm
(stress.F90:474-475, 577); neither is pinned by a reference test, so the
pin is the mathematical value: relative error <= 1e-15 for |z| <= 20, the
2 - erfc(|z|) reflection for z < 0, and the exact end values."""
import ctypes as C
import os
import subprocess

import mpmath as mp
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include "gam_erfc.h"
extern "C" void eval(int n, const double *z, double *y, double *e)
{
    for (int i = 0; i < n; i++) y[i] = gam_erfc_exp(z[i], e[i]);
}
'''


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("erfc")
    src = d / "erfc_host.cpp"
    src.write_text(SRC)
    so = d / "liberfc_host.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-std=c++17",
                    "-I", os.path.join(ROOT, "mage_amd", "csrc"), str(src), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    DP = C.POINTER(C.c_double)
    L.eval.argtypes = [C.c_int, DP, DP, DP]
    return L


def _eval(L, z):
    z = np.ascontiguousarray(z, dtype=np.float64)
    y = np.empty_like(z)
    e = np.empty_like(z)
    DP = C.POINTER(C.c_double)
    L.eval(len(z), z.ctypes.data_as(DP), y.ctypes.data_as(DP), e.ctypes.data_as(DP))
    return y, e


def test_erfc_and_gaussian_against_mpmath(lib):
    mp.mp.dps = 40
    z = np.concatenate([np.linspace(0.0, 1.0, 401), np.linspace(1.0, 6.0, 801),
                        np.linspace(6.0, 20.0, 401), [1e-300, 1e-8, 0.5, 2.0]])
    z = np.concatenate([z, -z])
    y, e = _eval(lib, z)
    ry = np.array([float(mp.erfc(mp.mpf(float(v)))) for v in z])
    re = np.array([float(mp.exp(-mp.mpf(float(v)) ** 2)) for v in z])
    assert np.max(np.abs(y - ry) / ry) <= 1e-15
    assert np.max(np.abs(e - re) / re) <= 1e-15


def test_erfc_end_values(lib):
    y, e = _eval(lib, np.array([0.0, 40.0, -40.0, np.inf, -np.inf]))
    assert y[0] == 1.0 and e[0] == 1.0
    assert y[1] == 0.0 and y[2] == 2.0 and e[1] == 0.0
    assert y[3] == 0.0 and y[4] == 2.0


def test_generator_reproduces_header():
    """gam_erfc.h is exactly what gen_erfc.py writes."""
    out = subprocess.run(["python3", os.path.join(ROOT, "mage_amd", "csrc", "gen_erfc.py")],
                         check=True, stdout=subprocess.PIPE, text=True).stdout
    assert out == open(os.path.join(ROOT, "mage_amd", "csrc", "gam_erfc.h")).read()
