"""The adjoint's memory paths give the gradients of the full-tape path.

Large grids (the 256x256x384 shell) run the reverse sweep with only part
of an interval's forward intermediates taped (the other steps recompute
their predictor, face velocities and fluxes inside the adjoint step) and
without the second tape pair (the interval recompute then runs before the
interval's adjoint steps, not beside them).  gam_adj_create_budget forces
those paths on a small grid: the MHD shell with the dipole, Boris, 7UP and
a time-varying inflow, 7 steps with checkpoints every 3 (intervals of 3,
3 and 1 steps).  Every path must reproduce the full-tape, overlapped
gradients (State and oState gas and face fluxes, inflow series) to 1e-12
relative; only the series gradient's atomic adds may land in another
order.
"""
import numpy as np
import pytest

from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu


def _case():
    N = (24, 16, 24)
    g = gamhip.GamHip(*N, pb.shell_params("7UP"))
    xyz = pb.shell_corners(*N)
    g.grid_from_corners(xyz)
    g.add_dipole_b0(pb.SHELL_M0)
    wind = pb.varying_wind(24, t0=-60.0, t1=60.0)
    g.set_wind(wind)
    xyzcc = g.get_grid(fields=("xyzcc",))["xyzcc"]
    g.put_state(pb.shell_mhd_state(*N, xyzcc, xyz, p0=0.1))
    g.init_run()
    g.run(2)
    grid = g.get_grid(fields=("xyzcc", "volume"))
    r = np.sqrt(np.sum(grid["xyzcc"] ** 2, axis=-1))
    w = np.asfortranarray(grid["volume"] * (r >= 12.0))
    return g, w, len(wind["tW"])


def _snapshot(g):
    return g.get_state(gamhip.GAM_STATE), g.get_state(gamhip.GAM_OSTATE), g.get_time()


def _restore(g, snap):
    S, O, (t, dt, ts) = snap
    g.put_state(S, gamhip.GAM_STATE)
    g.put_state(O, gamhip.GAM_OSTATE)
    g.set_time(t, dt, ts)


def _grads(g, w, nw, snap, **budget):
    _restore(g, snap)       # every run starts from the same state
    adj = gamhip.GamAdjoint(g, 7, 3, **budget)
    info = adj.tape_info()
    adj.set_weights(w)
    J = adj.forward()
    out = list(adj.backward(fluxes=True)) + list(adj.wind_grad(nw))
    adj.close()
    return J, out, info


def test_adjoint_tape_budgets_match_full_tape():
    print("adjoint tape budgets: 24x16x24 MHD shell, 7 steps, ckpt 3", flush=True)
    g, w, nw = _case()
    snap = _snapshot(g)
    J0, ref, info0 = _grads(g, w, nw, snap)
    assert info0 == (3, True), info0
    names = ("State gas", "oState gas", "State flux", "oState flux", "series Q", "series B")
    for a in ref:
        assert np.max(np.abs(a)) > 0.0
    for budget, want in (({"overlap": False}, (3, False)), ({"max_taped": 2}, (2, False)),
                         ({"max_taped": 0, "overlap": False}, (0, False))):
        print("adjoint tape budget %s" % budget, flush=True)
        J, got, info = _grads(g, w, nw, snap, **budget)
        assert info == want, (budget, info)
        assert J == J0
        for nm, a, b in zip(names, got, ref):
            err = np.max(np.abs(a - b)) / np.max(np.abs(b))
            assert err <= 1e-12, (budget, nm, err)
    g.close()


def test_adjoint_after_the_context_stepped_again():
    """Steps of the context between gam_adj_forward and gam_adj_backward
    rewrite its own forward intermediates; the backward must not take them
    for the recorded last step's (it recomputes them): the gradients equal
    those of a backward right after the forward."""
    print("adjoint after further steps: 24x16x24 MHD shell, 7 steps, ckpt 3, no taped steps", flush=True)
    g, w, nw = _case()
    snap = _snapshot(g)
    J0, ref, _ = _grads(g, w, nw, snap, max_taped=0, overlap=False)
    _restore(g, snap)
    adj = gamhip.GamAdjoint(g, 7, 3, max_taped=0, overlap=False)
    adj.set_weights(w)
    J = adj.forward()
    g.run(3)
    got = list(adj.backward(fluxes=True)) + list(adj.wind_grad(nw))
    adj.close()
    assert J == J0
    for a, b in zip(got, ref):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    g.close()


def _grads_dt(g, w, nw, snap, **budget):
    _restore(g, snap)
    adj = gamhip.GamAdjoint(g, 7, 3, **budget)
    adj.set_weights(w)
    adj.set_dt_mode(3)
    J = adj.forward()
    out = list(adj.backward(fluxes=True)) + list(adj.wind_grad(nw)) + [adj.dt_grad(7)]
    adj.close()
    return J, out


def test_dt_following_adjoint_on_every_tape_path():
    """With every dt following CalcDT (mode 3), the reduced tape paths (the
    interval's last step up to its fluxes only, taking its CalcDT record
    from the forward run; its intermediates from the context's buffers
    when untaped) give the full-tape gradients, dt sensitivities included."""
    print("dt-following adjoint on the tape paths: 24x16x24 MHD shell, 7 steps, ckpt 3", flush=True)
    g, w, nw = _case()
    snap = _snapshot(g)
    J0, ref = _grads_dt(g, w, nw, snap)
    assert np.max(np.abs(ref[-1])) > 0.0
    for budget in ({"overlap": False}, {"max_taped": 2}, {"max_taped": 0, "overlap": False}):
        print("adjoint tape budget %s, dt mode 3" % budget, flush=True)
        J, got = _grads_dt(g, w, nw, snap, **budget)
        assert J == J0
        for a, b in zip(got, ref):
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), budget
    g.close()
