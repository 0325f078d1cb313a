"""Ensembles: members of uneven size stepped by batched launches in a
captured hipGraph must give what each member gives stepped on its own
(which the parity tests pin to the oracle).

Tolerance: relative L-infinity 1e-12 per array after the steps (the batched
and the single launches run the same kernel bodies; the test reports the
largest difference seen).
"""
import numpy as np
import pytest

from helpers import rel_linf
from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

TOL = 1e-12
KEYS = ("Gas", "magFlux", "Bxyz", "Efld")


def member(spec, rmeth="7UP", **over):
    Nip, Njp, Nkp = spec["N"]
    p = pb.shell_params(rmeth)
    p.update(over)
    g = gamhip.GamHip(Nip, Njp, Nkp, p)
    xyz = pb.shell_corners(Nip, Njp, Nkp, seed=spec["seed"])
    g.grid_from_corners(xyz)
    g.add_dipole_b0(pb.SHELL_M0)
    g.set_wind(pb.constant_wind(vx=spec["vx"], bz=spec["bz"]))
    xyzcc = g.get_grid(fields=("xyzcc",))["xyzcc"]
    g.put_state(pb.shell_state(Nip, Njp, Nkp, xyzcc, seed=spec["seed"]))
    g.init_run()
    return g


SPECS = pb.ensemble_members(n=5, lo=(16, 16, 16), hi=(40, 32, 40))


def _diff(a, b):
    sa, sb = a.get_state(), b.get_state()
    return max(rel_linf(sa[k], sb[k]) for k in KEYS)


@pytest.mark.parametrize("concurrent", ["1", "0"])
def test_ensemble_matches_members_stepped_alone(concurrent, monkeypatch):
    """concurrent: GAM_CONCURRENT at context creation; "1" (the default)
    records the members' flux/EMF branches and the batched graph keeps
    them, "0" records a single chain."""
    monkeypatch.setenv("GAM_CONCURRENT", concurrent)
    print("ensemble sizes", [s["N"] for s in SPECS], "concurrent", concurrent, flush=True)
    batched = [member(s) for s in SPECS]
    alone = [member(s) for s in SPECS]
    ens = gamhip.GamEnsemble(batched)
    assert ens.launches_per_step() > 10
    # an odd step count, then more steps: both buffer roles of the graph
    ens.run(3)
    for g in alone:
        g.run(3)
    worst = max(_diff(a, b) for a, b in zip(batched, alone))
    print("after 3 steps: max rel diff %.3e" % worst, flush=True)
    assert worst <= TOL
    ens.run(4)
    for g in alone:
        g.run(4)
    for a, b in zip(batched, alone):
        ta, tb = a.get_time(), b.get_time()
        assert ta == tb
        assert _diff(a, b) <= TOL
        assert abs(a.divb()) <= 1e-12 * np.sum(np.abs(a.get_state()["magFlux"]))
    # a member stepped twice on its own keeps its buffer role
    batched[0].run(2)
    alone[0].run(2)
    ens.run(1)
    for g in alone:
        g.run(1)
    worst = max(_diff(a, b) for a, b in zip(batched, alone))
    print("after 10 steps: max rel diff %.3e" % worst, flush=True)
    assert worst <= TOL
    ens.close()
    for g in batched + alone:
        g.close()


def test_ensemble_rejects_odd_outside_steps_and_mixed_variants():
    a, b = member(SPECS[0]), member(SPECS[1])
    ens = gamhip.GamEnsemble([a, b])
    a.run(1)
    with pytest.raises(gamhip.GamError) as ei:
        ens.run(1)
    assert ei.value.code == -3
    ens.close()
    c = member(SPECS[2], rmeth="8CENT")
    with pytest.raises(gamhip.GamError) as ei:
        gamhip.GamEnsemble([a, c])
    assert ei.value.code == -1
    with pytest.raises(gamhip.GamError):
        gamhip.GamEnsemble([a, a])
    for g in (a, b, c):
        g.close()


def test_ensemble_of_one_matches_plain_run():
    s = SPECS[3]
    x, y = member(s), member(s)
    ens = gamhip.GamEnsemble([x])
    ms = ens.time_steps(5)
    assert ms > 0.0
    y.run(5)
    assert _diff(x, y) <= TOL
    ens.close()
    x.close()
    y.close()


def test_ensemble_refuses_members_changed_after_creation():
    """A member whose inflow series (or grid, background, plan) changes after
    the ensemble was recorded makes gam_ens_run refuse, instead of replaying
    launches that point at freed or stale device memory."""
    N = (16, 16, 24)
    gs = []
    for q in range(2):
        g = gamhip.GamHip(*N, pb.shell_params("7UP"))
        g.grid_from_corners(pb.shell_corners(*N, seed=pb.SEED0 + q))
        g.add_dipole_b0(pb.SHELL_M0)
        g.set_wind(pb.constant_wind())
        g.put_state(pb.shell_state(*N, g.get_grid(fields=("xyzcc",))["xyzcc"]))
        g.init_run()
        gs.append(g)
    ens = gamhip.GamEnsemble(gs)
    ens.run(2)
    gs[1].set_wind(pb.varying_wind(8))       # a different series length reallocates it
    with pytest.raises(gamhip.GamError) as ei:
        ens.run(2)
    assert ei.value.code == -3    # GAM_E_STATE
    ens.close()
    ens = gamhip.GamEnsemble(gs)              # recorded again: runs
    ens.run(2)
    ens.close()
    for g in gs:
        g.close()
