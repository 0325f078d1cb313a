"""Shared set-up for parity tests: build a case on the GPU and in the oracle
from the same corners, metrics and initial state."""
import numpy as np

import pyoracle as po
from mage_amd import gamhip, problems as pb


def rel_linf(a, b):
    """max|a-b| / max|b| (relative L-infinity)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den


def box_case(N=(16, 16, 16), rmeth="7UP", seed=pb.SEED0, gpu=True, **over):
    Nip, Njp, Nkp = N
    xyz = pb.cartesian_corners(Nip, Njp, Nkp)
    p = pb.box_params(rmeth)
    p.update(over)
    g = None
    if gpu:
        g = gamhip.GamHip(Nip, Njp, Nkp, p)
        g.grid_from_corners(xyz)
        grid = g.get_grid()
    else:
        grid = po.corners2grid(Nip, Njp, Nkp, xyz)
    grid["xyz"] = xyz
    st = pb.periodic_box_state(Nip, Njp, Nkp, grid["xyzcc"], xyz, seed=seed)
    o = po.Oracle(Nip, Njp, Nkp, p, grid)
    o.set_state(st)
    if g is not None:
        g.put_state(st)
    return g, o, grid, st, p


def shell_case(N=(32, 32, 32), rmeth="7UP", seed=pb.SEED0, wind=None, state="quiet", **over):
    """The warped shell with the dipole, Boris and the inflow BC.  state
    "quiet" is the benchmark IC (v = 0, no residual B); "mhd" has seeded
    flow and a seeded residual field (problems.shell_mhd_state)."""
    Nip, Njp, Nkp = N
    xyz = pb.shell_corners(Nip, Njp, Nkp, seed=seed)
    p = pb.shell_params(rmeth)
    p.update(over)
    wind = wind or pb.constant_wind()
    g = gamhip.GamHip(Nip, Njp, Nkp, p)
    g.grid_from_corners(xyz)
    g.add_dipole_b0(pb.SHELL_M0)
    g.set_wind(wind)
    grid = g.get_grid()
    grid.update(g.get_background())
    grid["xyz"] = xyz
    if state == "mhd":
        st = pb.shell_mhd_state(Nip, Njp, Nkp, grid["xyzcc"], xyz, seed=seed)
    else:
        st = pb.shell_state(Nip, Njp, Nkp, grid["xyzcc"], seed=seed)
    o = po.Oracle(Nip, Njp, Nkp, p, grid, wind=wind)
    o.set_state(st)
    g.put_state(st)
    return g, o, grid, st, p


def compare_state(g, o, keys=("Gas", "magFlux", "Bxyz", "Efld")):
    out = g.get_state()
    return {k: rel_linf(out[k], o.st[k]) for k in keys}
