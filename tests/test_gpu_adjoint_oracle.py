"""The device adjoint against finite differences of the CPU reference.

The oracle (the C restatement of the reference's step) is run forward
from perturbed initial states with each step's dt held at the values of
its own unperturbed run (the adjoint does not differentiate the CFL dt),
and the central difference of its objective is compared with the device
adjoint's directional derivative.  The case is the adjoint workload's
physics on a small grid: the warped shell with seeded flow and residual
field, the dipole background, Boris (Ca = 10), 7UP and a time-varying
inflow series; J = sum_c w_c P_c of the final State with w the cell
volume over r >= 12.

The start state is what gam_init_run makes of gas0 (the boundary pass,
oState = State) with every cell's B then formed from the face fluxes
(gam_adj_replay's order), on the device and in the oracle alike; the
device gradient is dJ/dState + dJ/doState with oState's ghosts folded at
State's time (gam_adj_set_ostate_bc_time), and the oracle is perturbed
the same way: gas0 +- e d, then the same start.
Directions: seeded smooth State-gas directions, directions in the inflow
series' density, pressure and field columns (the velocity moves the lag,
a step function), and face-flux fields on the I, J and K faces: smooth
ones, held to 1e-5 of sum |g d| because their directional derivatives
nearly cancel, and the gradient's own direction, held to 1e-5 relative.
Bar for the gas and series directions: 1e-6 relative at the best of the
steps e = 1e-3 .. 1e-6 (of 1 % of each field's scale).
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import shell_case
from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

NSTEPS = 4
GAMMA = 5.0 / 3.0


def _pressure_objective(gas, w, pfloor):
    rho = gas[..., 0]
    ir = 1.0 / rho
    v2 = sum((gas[..., 1 + x] * ir) ** 2 for x in range(3))
    P = np.maximum((GAMMA - 1.0) * (gas[..., 4] - 0.5 * rho * v2), pfloor)
    return float(np.sum(w * P))


def _smooth(shape, rng, scale):
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
    d[:4] = d[-4:] = 0.0
    d[:, :4] = d[:, -4:] = 0.0
    d[:, :, :4] = d[:, :, -4:] = 0.0
    return d


def test_adjoint_matches_finite_differences_of_the_oracle():
    N = (16, 16, 24)
    wind = pb.varying_wind(24, t0=-60.0, t1=60.0)
    print("adjoint vs oracle FD: %s MHD shell, %d steps" % (N, NSTEPS), flush=True)
    g, o, grid, st, p = shell_case(N, "7UP", wind=wind, state="mhd")
    r = np.sqrt(np.sum(grid["xyzcc"] ** 2, axis=-1))
    w = np.asfortranarray(grid["volume"] * (r >= 12.0))
    w[:4] = w[-4:] = 0.0
    w[:, :4] = w[:, -4:] = 0.0
    w[:, :, :4] = w[:, :, -4:] = 0.0

    def start(oo):
        """gam_init_run's state, then every cell's B from the face fluxes
        after the boundary pass: the adjoint's start (gam_adj_replay's
        order).  The reference's set-up forms cell B before the boundary
        pass (init.F90:200-244), so there the ghost cells' B would not
        follow the interior face fluxes."""
        oo.init_run()
        oo.recompute_cell_b()

    # the oracle's own dt sequence from gas0
    start(o)
    dts = []
    for _ in range(NSTEPS):
        dts.append(o.M.dt)
        o.step()
    J0 = _pressure_objective(o.st["Gas"], w, p["pFloor"])

    def J_oracle(gas, wnd, mf=None):
        oo = po.Oracle(*N, p, grid, wind=wnd)
        s = dict(st)
        s["Gas"] = gas
        if mf is not None:
            s["magFlux"] = mf
        oo.set_state(s)
        start(oo)
        for n in range(NSTEPS):
            oo.M.dt = dts[n]
            oo.step()
        return _pressure_objective(oo.st["Gas"], w, p["pFloor"])

    # device adjoint of the same forward (the same start, then the steps)
    g.init_run()
    g.bflux2fld(gamhip.GAM_STATE)
    g.bflux2fld(gamhip.GAM_OSTATE)
    adj = gamhip.GamAdjoint(g, NSTEPS, 2)
    adj.set_ostate_bc_time(True)
    adj.set_weights(w)
    Jd = adj.forward()
    assert Jd == pytest.approx(J0, rel=1e-11)
    gS, gO, gM, gMO = adj.backward(fluxes=True)
    grad = gS + gO
    gradM = gM + gMO
    gQ, gB = adj.wind_grad(len(wind["tW"]))
    adj.close()
    g.close()

    rng = np.random.default_rng(4242)
    gas0 = np.asfortranarray(st["Gas"])
    scale = np.max(np.abs(gas0), axis=(0, 1, 2))
    results = []
    for q in range(3):
        d = _smooth(gas0.shape, rng, scale)
        ad = float(np.sum(grad * d))
        best = None
        for e in (1e-3, 1e-4, 1e-5, 1e-6):
            fd = (J_oracle(gas0 + e * d, wind) - J_oracle(gas0 - e * d, wind)) / (2 * e)
            rel = abs(fd - ad) / max(abs(fd), abs(ad))
            best = rel if best is None else min(best, rel)
        print("gas direction %d: device adjoint %.12e, best rel vs oracle FD %.2e" % (q, ad, best), flush=True)
        results.append(("gas", best))
    for q in range(2):
        dQ = np.zeros(wind["Q"].shape, order="F")
        dQ[:, 0] = rng.standard_normal(len(wind["tW"])) * 1e-2
        dQ[:, 4] = rng.standard_normal(len(wind["tW"])) * 1e-5
        dB = np.zeros(wind["B"].shape, order="F")
        dB[:, 1:] = rng.standard_normal((len(wind["tW"]), 2)) * 1e-3
        ad = float(np.sum(gQ * dQ) + np.sum(gB * dB))
        best = None
        for e in (1e-2, 1e-3, 1e-4, 1e-5):
            wp, wm = dict(wind), dict(wind)
            wp["Q"], wm["Q"] = wind["Q"] + e * dQ, wind["Q"] - e * dQ
            wp["B"], wm["B"] = wind["B"] + e * dB, wind["B"] - e * dB
            fd = (J_oracle(gas0, wp) - J_oracle(gas0, wm)) / (2 * e)
            rel = abs(fd - ad) / max(abs(fd), abs(ad))
            best = rel if best is None else min(best, rel)
        print("series direction %d: device adjoint %.12e, best rel vs oracle FD %.2e" % (q, ad, best), flush=True)
        results.append(("series", best))
    # face-flux directions.  Smooth fields on the faces give directional
    # derivatives that nearly cancel (|<g, d>| ~ 1e-3 of sum |g d| on the
    # shell), so relative agreement there measures the difference quotient's
    # noise, not the gradient: they are held to 1e-6 of sum |g d|.  Along the
    # gradient itself (d = g on the faces of one direction) nothing cancels
    # and the relative bar applies.  The difference quotients along face
    # fluxes bottom out near 1e-6 (the limiters' switch points lie closer to
    # the base point than in the gas directions), so these rows are held to
    # 1e-5, ten times the smallest error measured over the step sizes.
    mf0 = np.asfortranarray(st["magFlux"])
    mscale = np.max(np.abs(mf0))
    for q in range(3):
        for kind in ("smooth", "gradient"):
            d = np.zeros(mf0.shape, order="F")
            if kind == "smooth":
                d[..., q] = _smooth(mf0.shape[:3] + (1,), rng, [mscale])[..., 0]
            else:
                d[4:-4, 4:-4, 4:-4, q] = gradM[4:-4, 4:-4, 4:-4, q]
                d *= 1e-2 * mscale / np.max(np.abs(d))
            ad = float(np.sum(gradM * d))
            gd = float(np.sum(np.abs(gradM * d)))
            best = None
            for e in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
                fd = (J_oracle(gas0, wind, mf0 + e * d) - J_oracle(gas0, wind, mf0 - e * d)) / (2 * e)
                q2 = (abs(fd - ad) / max(abs(fd), abs(ad)), abs(fd - ad) / gd)
                best = q2 if best is None or q2[0] < best[0] else best
            print("face-flux %s direction %s: <g,d> %.6e, sum|g d| %.3e, best rel %.2e, of sum|g d| %.2e"
                  % (kind, "IJK"[q], ad, gd, best[0], best[1]), flush=True)
            results.append(("flux %s %s" % (kind, "IJK"[q]), best[0] if kind == "gradient" else best[1]))
    assert all(b <= (1e-5 if kind.startswith("flux") else 1e-6) for kind, b in results), results
