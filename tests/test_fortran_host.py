"""The Fortran host path: ISO_C_BINDING module + gamera.x-style driver
(mage_amd/fortran) over the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from helpers import rel_linf
from mage_amd import gamhip, problems as pb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "mage_amd", "fortran", "gamerax_hip.x")


def test_fortran_driver_is_built_against_libgamhip():
    """The driver exists and its dynamic linker resolves libgamhip.so."""
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", EXE], stdout=subprocess.PIPE, text=True).stdout
    line = [l for l in out.splitlines() if "libgamhip.so" in l]
    assert line and "not found" not in line[0]


@pytest.mark.gpu
def test_fortran_driver_matches_c_abi_and_oracle(tmp_path):
    """gamerax_hip.x (Fortran host) steps Orszag-Tang 10 times; the same
    initial state stepped through the Python binding gives the identical
    state (same kernels), and the oracle agrees to 1e-10."""
    N, nk, steps = 32, 8, 10
    out = tmp_path / "ot.bin"
    r = subprocess.run([EXE, str(N), str(steps), str(out)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "kZCs" in r.stdout
    Ni, Nk = N + 8, nk + 8
    c, f = (Ni, Ni, Nk), (Ni + 1, Ni + 1, Nk + 1)
    raw = np.fromfile(out, dtype=np.float64)
    sizes = [np.prod(c) * 5, np.prod(f) * 3, np.prod(c) * 5, np.prod(f) * 3, np.prod(c) * 3]
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    G0 = parts[0].reshape(c + (5,), order="F")
    F0 = parts[1].reshape(f + (3,), order="F")
    G1 = parts[2].reshape(c + (5,), order="F")
    F1 = parts[3].reshape(f + (3,), order="F")
    B1 = parts[4].reshape(c + (3,), order="F")

    p = pb.box_params("7UP")
    xyz = pb.cartesian_corners(N, N, nk, lo=(0, 0, -0.5 * nk / N), hi=(1, 1, 0.5 * nk / N))
    g = gamhip.GamHip(N, N, nk, p)
    g.grid_from_corners(xyz)
    st = pb.empty_state(N, N, nk)
    st["Gas"][...] = G0
    st["magFlux"][...] = F0
    g.put_state(st)
    g.init_run()
    g.run(steps)
    s = g.get_state()
    assert np.array_equal(s["Gas"], G1) and np.array_equal(s["magFlux"], F1)
    assert np.array_equal(s["Bxyz"], B1)

    grid = g.get_grid()
    grid["xyz"] = xyz
    o = po.Oracle(N, N, nk, p, grid)
    o.set_state(st)
    o.init_run()
    for _ in range(steps):
        o.step()
    assert rel_linf(G1, o.st["Gas"]) <= 1e-10
    assert rel_linf(F1, o.st["magFlux"]) <= 1e-10
    g.close()
