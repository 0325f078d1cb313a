"""dt sensitivities of the discrete adjoint (gam_adj_set_dt_mode) against
central finite differences.

Mode 1: dJ/d(dt_n) of every step with the other steps' dt held, against
replays of the recorded dt sequence with dt_n moved by +-e
(gam_adj_replay_dt).  These derivatives run through the update (dt dH,
dt dM), the constrained transport (dt curl E), the EMF's vDiff cap
CFL edge / dt and the predictor's AB2 weight dt_n / (2 dt_(n-1)).

Mode 2: the gradient of the run whose dt follows the state (every dt after
the first is CalcDT of the state the step before leaves, before its
boundary pass, as in the reference's stepGamera), against central
differences of free-dt replays (gam_adj_replay_dt without a dt sequence)
and of the oracle stepping with its own CalcDT.  The first step's dt is
the caller's input in both.  The mode-0 gradient (every dt held) is
checked against the same free-dt differences, and must miss them: the dt
path carries a real share of the derivative.

Mode 3: mode 2 with the first dt also CalcDT of the start State (after
its boundary pass, as gam_init_run computes it): against replays whose
every dt is CalcDT's, and, through the torch bridge (StepObjective with
dt_follows=True, whose forward runs gam_init_run), against central
differences of the bridge's own forward.

Cases: the MHD shell (seeded flow and residual field, dipole background,
Boris, 7UP, time-varying inflow; J = pressure integral over r >= 12), and
the hydrodynamic and the MHD periodic boxes (J = a smoothly weighted
pressure integral).
Differences are the best of a few steps.  Bars: 1e-6 relative for the
dt derivatives; for state directions 1e-6 relative or, where the
directional derivative cancels to ~1 % of sum_c |g_c d_c| (the smoothly
weighted box objective), 1e-7 of that sum.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import box_case, shell_case
from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

NSTEPS = 4
GAMMA = 5.0 / 3.0


def _smooth(shape, rng, scale, ghost=4):
    Ni, Nj, Nk, nv = shape
    x = np.linspace(0, 1, Ni)[:, None, None]
    y = np.linspace(0, 1, Nj)[None, :, None]
    z = np.linspace(0, 1, Nk)[None, None, :]
    d = np.zeros(shape, order="F")
    for v in range(nv):
        k = rng.integers(1, 4, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        d[..., v] = (np.sin(2 * np.pi * k[0] * x + ph[0]) * np.sin(2 * np.pi * k[1] * y + ph[1])
                     * np.sin(2 * np.pi * k[2] * z + ph[2])) * scale[v] * 1e-2
    if ghost:
        d[:ghost] = d[-ghost:] = 0.0
        d[:, :ghost] = d[:, -ghost:] = 0.0
        d[:, :, :ghost] = d[:, :, -ghost:] = 0.0
    return d


def _shell():
    N = (16, 16, 24)
    wind = pb.varying_wind(24, t0=-60.0, t1=60.0)
    g, o, grid, st, p = shell_case(N, "7UP", wind=wind, state="mhd")
    r = np.sqrt(np.sum(grid["xyzcc"] ** 2, axis=-1))
    w = np.asfortranarray(grid["volume"] * (r >= 12.0))
    for ax in range(3):
        sl = [slice(None)] * 3
        sl[ax] = slice(0, 4)
        w[tuple(sl)] = 0.0
        sl[ax] = slice(-4, None)
        w[tuple(sl)] = 0.0
    return g, o, grid, st, p, w, N, wind


def _vary(g, st, seed):
    """Smooth variation in every gas variable: the box state's uniform
    pressure puts every limiter difference of it exactly on a branch
    switch, where a central difference averages two one-sided slopes."""
    gas = np.array(st["Gas"], order="F")
    scale = np.max(np.abs(gas), axis=(0, 1, 2))
    d = _smooth(gas.shape, np.random.default_rng(seed), scale, ghost=0)
    st = dict(st)
    st["Gas"] = np.asfortranarray(gas + 3.0 * d)
    g.put_state(st)
    return st


def _box():
    N = (16, 16, 16)
    g, o, grid, st, p = box_case(N, "7UP", doMHD=0)
    return g, o, grid, _vary(g, st, 5), p, _box_weights(grid), N, None


def _mhdbox():
    N = (16, 16, 16)
    g, o, grid, st, p = box_case(N, "7UP")
    return g, o, grid, _vary(g, st, 6), p, _box_weights(grid), N, None


def _box_weights(grid):
    vol = grid["volume"]
    Ni, Nj, Nk = vol.shape
    x = np.linspace(0, 1, Ni)[:, None, None]
    y = np.linspace(0, 1, Nj)[None, :, None]
    return np.asfortranarray(vol * (1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)) * np.ones((1, 1, Nk)))


def _case(name):
    return {"shell": _shell, "box": _box, "mhdbox": _mhdbox}[name]()


def _record(g, w, mode):
    """gam_init_run, every cell's B from the face fluxes, then the recorded
    forward and the reverse sweep with dt mode `mode`."""
    g.init_run()
    if g.params.get("doMHD", 1):
        g.bflux2fld(gamhip.GAM_STATE)
        g.bflux2fld(gamhip.GAM_OSTATE)
    adj = gamhip.GamAdjoint(g, NSTEPS, 2)
    adj.set_ostate_bc_time(True)
    adj.set_weights(w)
    adj.set_dt_mode(mode)
    J = adj.forward()
    return adj, J


@pytest.mark.parametrize("case", ["shell", "box", "mhdbox"])
def test_dt_partials_match_dt_finite_differences(case):
    g, o, grid, st, p, w, N, wind = _case(case)
    print("dt partials, %s %s, %d steps" % (case, N, NSTEPS), flush=True)
    adj, J = _record(g, w, 1)
    gas0 = np.asfortranarray(st["Gas"])
    mf0 = None   # the recorded start's face fluxes
    gS, gO = adj.backward()
    mu = adj.dt_grad(NSTEPS)
    dts = adj.step_dt(NSTEPS)
    # the recorded dt sequence replays the recorded run
    Jr = adj.replay_dt(gas0, gas0, mf0, mf0, dts)
    assert Jr == pytest.approx(J, rel=1e-13)
    worst = 0.0
    for n in range(NSTEPS):
        best = None
        for e in (1e-4, 1e-5, 1e-6, 1e-7):
            h = e * dts[n]
            dp, dm = dts.copy(), dts.copy()
            dp[n] += h
            dm[n] -= h
            fd = (adj.replay_dt(gas0, gas0, mf0, mf0, dp) - adj.replay_dt(gas0, gas0, mf0, mf0, dm)) / (2 * h)
            rel = abs(fd - mu[n]) / max(abs(fd), abs(mu[n]), 1e-300)
            print("  %s dt_%d e %.0e: FD %.12e" % (case, n, e, fd), flush=True)
            best = rel if best is None else min(best, rel)
        print("%s dJ/d(dt_%d): adjoint %.12e, best rel vs FD %.2e" % (case, n, mu[n], best), flush=True)
        worst = max(worst, best)
    assert worst <= 1e-6, (case, worst)
    adj.close()
    g.close()


@pytest.mark.parametrize("case", ["shell", "box", "mhdbox"])
def test_dt_following_gradient_matches_free_dt_differences(case):
    g, o, grid, st, p, w, N, wind = _case(case)
    print("dt-following gradient, %s %s, %d steps" % (case, N, NSTEPS), flush=True)
    gas0 = np.asfortranarray(st["Gas"])
    mf0 = None   # the recorded start's face fluxes
    adj, J = _record(g, w, 2)
    gS, gO = adj.backward()
    grad2 = gS + gO
    mu = adj.dt_grad(NSTEPS)
    dts = adj.step_dt(NSTEPS)
    nw = len(wind["tW"]) if wind is not None else 0
    if nw:
        gQ2, gB2 = adj.wind_grad(nw)
    adj.set_dt_mode(0)
    gS0, gO0 = adj.backward()
    grad0 = gS0 + gO0
    if nw:
        gQ0, gB0 = adj.wind_grad(nw)
    # the free-dt replay of the unperturbed start is the recorded run
    assert adj.replay_dt(gas0, gas0, mf0, mf0) == pytest.approx(J, rel=1e-13)
    print("%s: dt_n %s, dJ/d(dt_n) %s" % (case, np.array2string(dts, precision=6),
                                          np.array2string(mu, precision=6)), flush=True)

    # the oracle's run with its own CalcDT after the first step
    def start(oo):
        oo.init_run()
        if p["doMHD"]:
            oo.recompute_cell_b()

    def J_oracle(gas, wnd=wind):
        oo = po.Oracle(*N, p, grid, wind=wnd) if wnd is not None else po.Oracle(*N, p, grid)
        s = dict(st)
        s["Gas"] = gas
        oo.set_state(s)
        start(oo)
        oo.M.dt = dts[0]
        for _ in range(NSTEPS):
            oo.step()
        rho = oo.st["Gas"][..., 0]
        ir = 1.0 / rho
        v2 = sum((oo.st["Gas"][..., 1 + x] * ir) ** 2 for x in range(3))
        P = np.maximum((GAMMA - 1.0) * (oo.st["Gas"][..., 4] - 0.5 * rho * v2), p["pFloor"])
        return float(np.sum(w * P))

    assert J_oracle(gas0) == pytest.approx(J, rel=1e-10)
    rng = np.random.default_rng({"shell": 777, "box": 778, "mhdbox": 779}[case])
    scale = np.max(np.abs(gas0), axis=(0, 1, 2))
    gap = []
    for q in range(3):
        d = _smooth(gas0.shape, rng, scale)
        a2 = float(np.sum(grad2 * d))
        a0 = float(np.sum(grad0 * d))
        gd = float(np.sum(np.abs(grad2 * d)))
        best2 = best0 = bestO = None
        for e in (1e-3, 1e-4, 1e-5, 1e-6):
            fd = (adj.replay_dt(gas0 + e * d, gas0 + e * d, mf0, mf0) -
                  adj.replay_dt(gas0 - e * d, gas0 - e * d, mf0, mf0)) / (2 * e)
            r2 = (abs(fd - a2) / max(abs(fd), abs(a2)), abs(fd - a2) / gd)
            r0 = abs(fd - a0) / max(abs(fd), abs(a0))
            best2 = r2 if best2 is None or r2[0] < best2[0] else best2
            best0 = r0 if best0 is None else min(best0, r0)
            if e >= 1e-5:
                fo = (J_oracle(gas0 + e * d) - J_oracle(gas0 - e * d)) / (2 * e)
                ro = (abs(fo - a2) / max(abs(fo), abs(a2)), abs(fo - a2) / gd)
                bestO = ro if bestO is None or ro[0] < bestO[0] else bestO
        print("%s direction %d: adjoint %.12e (dt held %.12e), best rel vs free-dt FD %.2e (%.1e of sum|g d|; "
              "held-dt gradient %.2e), vs oracle FD %.2e (%.1e), |a2 - a0| / sum|g d| %.2e" % (
                  case, q, a2, a0, best2[0], best2[1], best0, bestO[0], bestO[1], abs(a2 - a0) / gd), flush=True)
        # directional derivatives that cancel to ~1 % of sum |g d| are held
        # to that sum
        assert best2[0] <= 1e-6 or best2[1] <= 1e-7, (case, q, best2)
        assert bestO[0] <= 1e-6 or bestO[1] <= 1e-7, (case, q, bestO)
        gap.append(best0 / max(best2[0], 1e-16))
    # the held-dt gradient is not the gradient of this run
    assert max(gap) >= 10.0, gap
    # the inflow series (density and pressure columns, By and Bz; the
    # velocity moves the lag, a step function)
    for q in range(2 if nw else 0):
        dQ = np.zeros(wind["Q"].shape, order="F")
        dQ[:, 0] = rng.standard_normal(nw) * 1e-2
        dQ[:, 4] = rng.standard_normal(nw) * 1e-5
        dB = np.zeros(wind["B"].shape, order="F")
        dB[:, 1:] = rng.standard_normal((nw, 2)) * 1e-3
        a2 = float(np.sum(gQ2 * dQ) + np.sum(gB2 * dB))
        a0 = float(np.sum(gQ0 * dQ) + np.sum(gB0 * dB))
        best2 = bestO = None
        for e in (1e-2, 1e-3, 1e-4, 1e-5):
            wp, wm = dict(wind), dict(wind)
            wp["Q"], wm["Q"] = wind["Q"] + e * dQ, wind["Q"] - e * dQ
            wp["B"], wm["B"] = wind["B"] + e * dB, wind["B"] - e * dB
            g.set_wind(wp)
            Jp = adj.replay_dt(gas0, gas0, mf0, mf0)
            g.set_wind(wm)
            Jm = adj.replay_dt(gas0, gas0, mf0, mf0)
            g.set_wind(wind)
            fd = (Jp - Jm) / (2 * e)
            r2 = abs(fd - a2) / max(abs(fd), abs(a2))
            best2 = r2 if best2 is None else min(best2, r2)
            if e >= 1e-4:
                fo = (J_oracle(gas0, wp) - J_oracle(gas0, wm)) / (2 * e)
                ro = abs(fo - a2) / max(abs(fo), abs(a2))
                bestO = ro if bestO is None else min(bestO, ro)
        print("%s series direction %d: adjoint %.12e (dt held %.12e), best rel vs free-dt FD %.2e, vs oracle "
              "FD %.2e" % (case, q, a2, a0, best2, bestO), flush=True)
        assert best2 <= 1e-6 and bestO <= 1e-6, (case, q, best2, bestO)
    adj.close()
    g.close()


def _directional(adj, grad, gas0, d, replay, gd_floor=1e-7):
    """Best (relative, of sum |g d|) agreement of <grad, d> with central
    differences of replay over a few steps."""
    a = float(np.sum(grad * d))
    gd = float(np.sum(np.abs(grad * d)))
    best = None
    for e in (1e-3, 1e-4, 1e-5, 1e-6):
        fd = (replay(gas0 + e * d) - replay(gas0 - e * d)) / (2 * e)
        c = (abs(fd - a) / max(abs(fd), abs(a)), abs(fd - a) / gd)
        best = c if best is None or c[0] < best[0] else best
    return a, best


@pytest.mark.parametrize("case", ["shell", "box"])
def test_dt_from_the_start_state(case):
    g, o, grid, st, p, w, N, wind = _case(case)
    print("dt mode 3, %s %s, %d steps" % (case, N, NSTEPS), flush=True)
    gas0 = np.asfortranarray(st["Gas"])
    g.init_run()
    if p["doMHD"]:
        # the replay's start form: every cell's B from the face fluxes; the
        # first dt is CalcDT of that state
        g.bflux2fld(gamhip.GAM_STATE)
        g.bflux2fld(gamhip.GAM_OSTATE)
        t, _, ts = g.get_time()
        g.set_time(t, g.calc_dt(), ts)
    adj = gamhip.GamAdjoint(g, NSTEPS, 2)
    adj.set_ostate_bc_time(True)
    adj.set_weights(w)
    adj.set_dt_mode(3)
    J = adj.forward()
    gS, gO = adj.backward()
    grad3 = gS + gO
    zeros = np.zeros(NSTEPS)
    assert adj.replay_dt(gas0, gas0, None, None, zeros) == pytest.approx(J, rel=1e-13)
    rng = np.random.default_rng({"shell": 787, "box": 788}[case])
    scale = np.max(np.abs(gas0), axis=(0, 1, 2))
    for q in range(3):
        d = _smooth(gas0.shape, rng, scale)
        a, best = _directional(adj, grad3, gas0, d, lambda x: adj.replay_dt(x, x, None, None, zeros))
        print("%s mode 3 direction %d: adjoint %.12e, best rel vs all-CalcDT FD %.2e (%.1e of sum|g d|)"
              % (case, q, a, best[0], best[1]), flush=True)
        assert best[0] <= 1e-6 or best[1] <= 1e-7, (case, q, best)
    adj.close()
    g.close()


def test_torch_bridge_with_dt_following_the_state():
    import torch
    from mage_amd.autograd import StepObjective
    g, o, grid, st, p, w, N, wind = _case("box")
    g.init_run()
    base = np.asfortranarray(st["Gas"])
    obj = StepObjective(g, NSTEPS, 2, w, dt_follows=True)
    held = StepObjective(g, NSTEPS, 2, w)   # made from the same start
    gas0 = torch.from_numpy(np.ascontiguousarray(base)).requires_grad_(True)
    J = obj(gas0)
    J.backward()
    grad = gas0.grad.numpy()
    gas1 = torch.from_numpy(np.ascontiguousarray(base)).requires_grad_(True)
    held(gas1).backward()
    grad0 = gas1.grad.numpy()

    def J_of(x):
        with torch.no_grad():
            return float(obj(torch.from_numpy(np.ascontiguousarray(x))))

    rng = np.random.default_rng(799)
    scale = np.max(np.abs(base), axis=(0, 1, 2))
    for q in range(2):
        d = _smooth(base.shape, rng, scale)
        a, best = _directional(None, grad, base, d, J_of)
        a0, best0 = _directional(None, grad0, base, d, J_of)
        print("bridge dt_follows direction %d: <grad, d> %.12e best rel %.2e (%.1e of sum|g d|); "
              "held-dt gradient %.2e" % (q, a, best[0], best[1], best0[0]), flush=True)
        assert best[0] <= 1e-6 or best[1] <= 1e-7, (q, best)
        assert best0[0] > 10 * best[0]
    obj.close()
    held.close()
    g.close()
