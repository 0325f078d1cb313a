"""Restart files (include/gamres.h): the layout of the reference's
writeH5Res/readH5Restart (gioH5.F90:636-858).

CPU: host-array round trip, the on-disk structure as h5dump reports it
(dataset names, reversed Fortran extents, root attributes), the older
layout without ghosts, error paths.  GPU: a context restarted from a file
continues the run of the context that wrote it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.skipif(not os.path.exists(gamhip.RES_LIBPATH),
                                reason="libgamres.so not built (no HDF5)")

N = (16, 8, 8)


def _state(rng, N, ghosts=True):
    g = 2 * pb.NG if ghosts else 0
    c = (N[0] + g, N[1] + g, N[2] + g)
    f = (c[0] + 1, c[1] + 1, c[2] + 1)
    return {"Gas": np.asfortranarray(rng.standard_normal(c + (5,))),
            "magFlux": np.asfortranarray(rng.standard_normal(f + (3,))),
            "Bxyz": np.asfortranarray(rng.standard_normal(c + (3,)))}


META = dict(nOut=3, nRes=7, ts=120, t=1.25, ot=1.2375, dt=0.0125, dt0=0.0)


def test_round_trip_host_arrays(tmp_path):
    rng = np.random.default_rng(5)
    st, ost = _state(rng, N), _state(rng, N)
    xyz = pb.cartesian_corners(*N)
    p = tmp_path / "run.gam.Res.00007.h5"
    gamhip.res_write(p, N, META, xyz, st, ost)
    meta, s2, o2 = gamhip.res_read(p, N)
    assert meta == META
    for k in ("Gas", "magFlux"):
        assert np.array_equal(s2[k], st[k]) and np.array_equal(o2[k], ost[k])


def test_small_dt0_reads_as_zero(tmp_path):
    rng = np.random.default_rng(6)
    st, ost = _state(rng, N), _state(rng, N)
    p = tmp_path / "r.h5"
    gamhip.res_write(p, N, dict(META, dt0=1e-12), None, st, ost)
    assert gamhip.res_read(p, N)[0]["dt0"] == 0.0       # gioH5.F90:650-654, 814-820
    gamhip.res_write(p, N, dict(META, dt0=2.5e-3), None, st, ost)
    assert gamhip.res_read(p, N)[0]["dt0"] == 2.5e-3


@pytest.mark.skipif(shutil.which("h5dump") is None and not os.path.exists("/opt/conda/bin/h5dump"),
                    reason="h5dump not available")
def test_file_structure_matches_reference_layout(tmp_path):
    """Datasets carry the Fortran extents reversed (the Fortran HDF5 API's
    convention), scalars are root attributes."""
    rng = np.random.default_rng(7)
    st, ost = _state(rng, N), _state(rng, N)
    xyz = pb.cartesian_corners(*N)
    p = tmp_path / "r.h5"
    gamhip.res_write(p, N, META, xyz, st, ost)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([h5dump, "-H", str(p)], check=True, stdout=subprocess.PIPE, text=True).stdout
    Ni, Nj, Nk = (n + 8 for n in N)
    for name, dims in [("Gas", (1, 5, Nk, Nj, Ni)), ("oGas", (1, 5, Nk, Nj, Ni)),
                       ("magFlux", (3, Nk + 1, Nj + 1, Ni + 1)), ("Bxyz", (3, Nk, Nj, Ni)),
                       ("X", (Nk + 1, Nj + 1, Ni + 1))]:
        block = out[out.index('DATASET "%s"' % name):]
        block = block[:block.index("}")]
        assert "H5T_IEEE_F64LE" in block
        assert "( %s )" % ", ".join(str(d) for d in dims) in block, (name, block)
    for a in ("nOut", "nRes", "ts", "t", "ot", "dt", "dt0"):
        assert 'ATTRIBUTE "%s"' % a in out
    assert out.count('ATTRIBUTE "Units"') >= 9


def test_reads_restart_without_ghosts(tmp_path):
    """Older restarts hold the active cells only (gioH5.F90:767-773); they
    fill the active block, ghosts stay as they were."""
    rng = np.random.default_rng(8)
    st, ost = _state(rng, N, ghosts=False), _state(rng, N, ghosts=False)
    p = tmp_path / "old.h5"
    # write with the active extents by calling the writer with nG = 0
    import ctypes as C
    L = gamhip.load_restart()
    d = gamhip.Dims(*N, 0)
    m = gamhip.ResMeta(**META)
    P = gamhip._p
    assert L.gam_res_write(str(p).encode(), C.byref(d), C.byref(m), P(None), P(st["Gas"]),
                           P(st["magFlux"]), P(None), P(ost["Gas"]), P(ost["magFlux"]), P(None)) == 0
    meta, s2, o2 = gamhip.res_read(p, N)
    a = (slice(4, -4),) * 3
    f = (slice(4, -4),) * 3
    assert np.array_equal(s2["Gas"][a], st["Gas"])
    assert np.array_equal(o2["magFlux"][f], ost["magFlux"])
    assert not s2["Gas"][:4].any()


def test_wrong_grid_and_missing_file(tmp_path):
    rng = np.random.default_rng(9)
    st, ost = _state(rng, N), _state(rng, N)
    p = tmp_path / "r.h5"
    gamhip.res_write(p, N, META, None, st, ost)
    with pytest.raises(gamhip.GamError, match="another size"):
        gamhip.res_read(p, (16, 16, 8))
    with pytest.raises(gamhip.GamError, match="unable to open"):
        gamhip.res_read(tmp_path / "nope.h5", N)


@pytest.mark.gpu
def test_restarted_context_continues_the_run(tmp_path):
    """Run A: 10 steps, write a restart.  Run B: a new context on the same
    grid reads it.  The restart carries State, oState, t, ts and dt0
    exactly; Model%dt is recomputed from the restored state (init.F90:238),
    i.e. after the boundary pass, where the step computed it before
    (gamapp.F90:81-86), so the two can differ in the last digits.  A then
    takes the same CalcDT in place, and both run 10 more steps: the runs
    must agree."""
    from helpers import shell_case
    gA, o, grid, st, p = shell_case((24, 24, 32))
    gA.init_run()
    gA.run(10)
    f = tmp_path / "shell.gam.Res.00000.h5"
    gamhip.write_restart(gA, f, grid["xyz"], nOut=0, nRes=0)
    a10 = gA.get_state(), gA.get_state(gamhip.GAM_OSTATE), gA.get_time()
    dtA = gA.calc_dt()
    gA.set_time(a10[2][0], dtA, a10[2][2])
    gA.run(10)
    a20 = gA.get_state()

    gB = gamhip.GamHip(24, 24, 32, p)
    gB.grid_from_corners(grid["xyz"])
    gB.add_dipole_b0(pb.SHELL_M0)
    gB.set_wind(pb.constant_wind())
    nOut, nRes, dt = gamhip.read_restart(gB, f)
    assert (nOut, nRes) == (0, 1)
    b10, bo10 = gB.get_state(), gB.get_state(gamhip.GAM_OSTATE)
    for k in ("Gas", "magFlux", "Bxyz"):
        assert np.array_equal(b10[k], a10[0][k]), k
    for k in ("Gas", "magFlux"):
        assert np.array_equal(bo10[k], a10[1][k]), k
    t, dtB, ts = gB.get_time()
    assert t == a10[2][0] and ts == a10[2][2]
    assert dtB == dt == dtA
    assert abs(dtB - a10[2][1]) <= 1e-5 * a10[2][1]
    gB.run(10)
    b20 = gB.get_state()
    for k in ("Gas", "magFlux", "Bxyz"):
        den = np.max(np.abs(a20[k]))
        assert np.max(np.abs(b20[k] - a20[k])) <= 1e-13 * den, k
    gA.close()
    gB.close()
