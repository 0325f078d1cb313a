"""Discrete adjoint of the hydrodynamic step against central finite
differences of the same device step (dt held at the recorded values in
both), on the periodic box (8CENT and 7UP) and on a warped shell with
zero-gradient, inflow and periodic boundaries (MHD with the dipole
background and Boris), including the gradient with respect to the inflow
time series.

Gradients are with respect to the interiors of the starting State and
oState gas (ghost cells follow through the boundary conditions, in the
adjoint and in the replays alike).  For seeded smooth directions d the
adjoint's <g, d> must match (J(Gas + e d) - J(Gas - e d)) / (2 e) at the
best of four steps, e = 1e-2 ... 1e-5 of 1 % of each field's scale
(relative perturbations 1e-4 to 1e-7).  The difference quotient is
squeezed from both sides: at small steps by the round-off of the forward
step (the error grows like 1/e, about 1e-16 J / (e |d|)), at large ones by
the PDM limiter's branch switches that the perturbation crosses (the
adjoint differentiates the branch taken, as intended away from switches).
A wrong adjoint misses at every step; the best one must agree, for
State directions, to 1e-4 relative to the derivative and to 1e-6 of
sum_c |g_c d_c| (the size of the terms the directional derivative sums,
free of their cancellation).  oState directions move J 10-1000x less
(oState enters only through the AB2 slope), so no step escapes both the
round-off floor and the limiter switches as cleanly: measured best
agreement 1e-6 to 2e-3 relative; they are held to 1e-3 relative or 1e-5
of sum_c |g_c d_c|.
The states get smooth variation in every variable first: where a variable
is exactly uniform every limiter difference is exactly 0, a branch switch,
and a central difference there averages two one-sided slopes instead of
measuring the derivative.
"""
import numpy as np
import pytest

from helpers import box_case
from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _directions(shape, n, seed):
    rng = np.random.default_rng(seed)
    Ni, Nj, Nk, _ = shape
    x = np.linspace(0, 1, Ni)[:, None, None]
    y = np.linspace(0, 1, Nj)[None, :, None]
    z = np.linspace(0, 1, Nk)[None, None, :]
    out = []
    for _ in range(n):
        d = np.zeros(shape, order="F")
        for v in range(5):
            k = rng.integers(1, 4, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            d[..., v] = (np.sin(2 * np.pi * k[0] * x + ph[0]) * np.sin(2 * np.pi * k[1] * y + ph[1])
                         * np.sin(2 * np.pi * k[2] * z + ph[2]))
        out.append(d)
    return out


def _generic(g, seed):
    """Add smooth variation to every gas variable of State and oState: with
    an exactly uniform variable (rho = 1 or p = 1 in the case states) every
    limiter difference is exactly 0, i.e. on a branch switch, where a
    central difference averages the two one-sided slopes."""
    S = g.get_state(gamhip.GAM_STATE)
    O = g.get_state(gamhip.GAM_OSTATE)
    t = g.get_time()[0]
    scale = np.max(np.abs(S["Gas"]), axis=(0, 1, 2))
    d = _directions(S["Gas"].shape, 1, seed)[0]
    for v in range(5):
        S["Gas"][..., v] += 0.03 * scale[v] * d[..., v]
        O["Gas"][..., v] += 0.03 * scale[v] * d[..., v]
    g.put_state(S, gamhip.GAM_STATE)
    g.put_state(O, gamhip.GAM_OSTATE)
    return t


def _weights(g):
    vol = g.get_grid(fields=("volume",))["volume"]
    Ni, Nj, Nk = vol.shape
    x = np.linspace(0, 1, Ni)[:, None, None]
    y = np.linspace(0, 1, Nj)[None, :, None]
    return np.asfortranarray(vol * (1.0 + 0.5 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)) * np.ones((1, 1, Nk)))


def _shell_weights(g):
    """The adjoint workload's objective: pressure integral over 3 <= r <= 6."""
    grid = g.get_grid(fields=("xyzcc", "volume"))
    r = np.sqrt(np.sum(grid["xyzcc"] ** 2, axis=-1))
    return np.asfortranarray(grid["volume"] * ((r >= 3.0) & (r <= 6.0)))


def _outer_weights(g, r0=20.0):
    """Pressure integral over r >= r0: the cells the inflow reaches within
    a few steps."""
    grid = g.get_grid(fields=("xyzcc", "volume"))
    r = np.sqrt(np.sum(grid["xyzcc"] ** 2, axis=-1))
    return np.asfortranarray(grid["volume"] * (r >= r0))


def _fd_check(g, nsteps, ckpt, ndir, seed, tag, weights=None, ostate=True):
    s0 = g.get_state(gamhip.GAM_STATE)["Gas"].copy(order="F")
    o0 = g.get_state(gamhip.GAM_OSTATE)["Gas"].copy(order="F")
    adj = gamhip.GamAdjoint(g, nsteps, ckpt)
    adj.set_weights(_weights(g) if weights is None else weights)
    J = adj.forward()
    gS, gO = adj.backward()
    assert adj.replay(s0, o0) == pytest.approx(J, rel=1e-14, abs=0.0)
    scale = np.max(np.abs(s0), axis=(0, 1, 2))
    worst = 0.0
    for q, d in enumerate(_directions(s0.shape, ndir, seed)):
        d = d * scale * 1e-2
        which = q % 2 if ostate else 0   # alternate: State gas, oState gas
        ad = np.sum(gS * d) if which == 0 else np.sum(gO * d)
        gd = np.sum(np.abs(gS * d if which == 0 else gO * d))
        best = None
        for e in (1e-2, 1e-3, 1e-4, 1e-5):
            if which == 0:
                fd = (adj.replay(s0 + e * d, o0) - adj.replay(s0 - e * d, o0)) / (2 * e)
            else:
                fd = (adj.replay(s0, o0 + e * d) - adj.replay(s0, o0 - e * d)) / (2 * e)
            rel = abs(fd - ad) / max(abs(fd), abs(ad))
            print("%s dir %d (%s) e %.0e: fd %.12e adj %.12e rel %.2e of sum|g d| %.2e" % (
                tag, q, "State" if which == 0 else "oState", e, fd, ad, rel, abs(fd - ad) / gd), flush=True)
            if best is None or rel < best[0]:
                best = (rel, abs(fd - ad) / gd)
        if which == 0:
            assert best[0] <= 1e-4 and best[1] <= 1e-6, (tag, q, best)
            worst = max(worst, best[1])
        else:
            assert best[0] <= 1e-3 or best[1] <= 1e-5, (tag, q, best)
    f_ms, b_ms = adj.times()
    print("%s: worst State-direction error %.2e of sum|g d|; forward %.2f ms, reverse %.2f ms for %d steps" % (tag, worst, f_ms, b_ms, nsteps), flush=True)
    adj.close()
    return worst


@pytest.mark.parametrize("rmeth", ["8CENT", "7UP"])
def test_adjoint_box_matches_finite_differences(rmeth):
    g = box_case((16, 16, 16), rmeth, doMHD=0)[0]
    g.init_run()
    _generic(g, 7)
    g.run(2)              # a non-trivial oState (AB2 slope), both states BC-consistent
    assert _fd_check(g, 6, 4, 6, 101, "box " + rmeth) <= TOL
    g.close()


def test_adjoint_shell_zero_gradient_matches_finite_differences():
    N = (24, 16, 24)
    p = pb.shell_params("7UP")
    p.update(doMHD=0, doBoris=0, doBackground=0,
             bc=["zerograd", "zerograd", "zerograd", "zerograd", "periodic", "periodic"])
    g = gamhip.GamHip(*N, p)
    g.grid_from_corners(pb.shell_corners(*N))
    xyzcc = g.get_grid(fields=("xyzcc",))["xyzcc"]
    st = pb.shell_state(*N, xyzcc, p0=0.1)
    rng = np.random.default_rng(5)
    st["Gas"][..., 1] = 0.05 * np.sin(xyzcc[..., 1])   # some momentum
    st["Gas"][..., 4] += 0.5 * st["Gas"][..., 1] ** 2
    g.put_state(st)
    g.init_run()
    _generic(g, 9)
    g.run(2)
    assert _fd_check(g, 5, 2, 6, 202, "shell") <= TOL
    g.close()


def test_adjoint_refuses_unsupported_contexts():
    """Boris or a background without MHD is refused; the inflow-series
    gradient needs an inflow BC."""
    N = (16, 16, 16)
    g = box_case(N, "7UP", doMHD=0)[0]
    g.init_run()
    adj = gamhip.GamAdjoint(g, 2, 1)
    adj.set_weights(_weights(g))
    adj.forward()
    adj.backward()
    with pytest.raises(gamhip.GamError) as ei:
        adj.wind_grad(2)
    assert ei.value.code == -3
    adj.close()
    g.close()
    p = pb.box_params("7UP")
    p.update(doMHD=0, doBoris=1)
    g = gamhip.GamHip(*N, p)
    g.grid_from_corners(pb.cartesian_corners(*N))
    with pytest.raises(gamhip.GamError) as ei:
        gamhip.GamAdjoint(g, 4, 2)
    assert ei.value.code == -1
    g.close()


def _fd_check_flux(g, nsteps, ckpt, ndir, seed, tag, weights=None, ostate=True):
    """Face-flux directions of an MHD run against difference quotients of
    gam_adj_replay.  Two kinds: seeded smooth fields on the faces (State
    magFlux, and with ostate=True alternately oState's) and, per face
    direction, the State gradient itself on that direction's active faces.
    A smooth field's directional derivative can nearly cancel (on the
    shells |<g, d>| is 1e-3 to 1e-2 of sum |g d|), so its error is taken
    relative to sum |g d|; along the gradient nothing cancels and the error
    is relative.  Returns the best error over the steps, per direction."""
    S = g.get_state(gamhip.GAM_STATE)
    O = g.get_state(gamhip.GAM_OSTATE)
    s0, o0 = S["Gas"].copy(order="F"), O["Gas"].copy(order="F")
    m0, mo0 = S["magFlux"].copy(order="F"), O["magFlux"].copy(order="F")
    adj = gamhip.GamAdjoint(g, nsteps, ckpt)
    adj.set_weights(_weights(g) if weights is None else weights)
    adj.forward()
    gS, gO, gM, gMO = adj.backward(fluxes=True)
    scale = np.max(np.abs(m0))
    cases = []
    for q, d in enumerate(_directions(m0.shape[:3] + (5,), ndir, seed)):
        cases.append(("smooth", q % 2 if ostate else 0, np.asfortranarray(d[..., :3]) * scale * 1e-2))
    for c in range(3):
        d = np.zeros(m0.shape, order="F")
        d[4:-4, 4:-4, 4:-4, c] = gM[4:-4, 4:-4, 4:-4, c]
        cases.append(("gradient", 0, d * (1e-2 * scale / np.max(np.abs(d)))))
    bests = []
    for q, (kind, which, d) in enumerate(cases):
        gw = gM if which == 0 else gMO
        ad = float(np.sum(gw * d))
        gd = float(np.sum(np.abs(gw * d)))
        best = None
        for e in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
            if which == 0:
                fd = (adj.replay(s0, o0, m0 + e * d, mo0) - adj.replay(s0, o0, m0 - e * d, mo0)) / (2 * e)
            else:
                fd = (adj.replay(s0, o0, m0, mo0 + e * d) - adj.replay(s0, o0, m0, mo0 - e * d)) / (2 * e)
            err = abs(fd - ad) / (gd if kind == "smooth" else max(abs(fd), abs(ad)))
            print("%s flux %s dir %d (%s) e %.0e: fd %.12e adj %.12e sum|g d| %.3e err %.2e" % (
                tag, kind, q, "State" if which == 0 else "oState", e, fd, ad, gd, err), flush=True)
            best = err if best is None else min(best, err)
        bests.append(best)
    adj.close()
    print("%s flux directions, best error (smooth: of sum|g d|; gradient: relative): %s"
          % (tag, ["%.1e" % b for b in bests]), flush=True)
    return bests


@pytest.mark.parametrize("rmeth", ["8CENT", "7UP"])
def test_adjoint_mhd_box_matches_finite_differences(rmeth):
    """MHD on the periodic box: EMF, constrained transport, magnetic
    kinetic flux with HOGS, Maxwell update, all in the adjoint."""
    g = box_case((16, 16, 16), rmeth)[0]
    g.init_run()
    _generic(g, 11)
    g.run(2)
    assert _fd_check(g, 6, 4, 6, 303, "mhd box " + rmeth) <= TOL
    assert max(_fd_check_flux(g, 6, 4, 4, 404, "mhd box " + rmeth)) <= 1e-5
    g.close()


@pytest.mark.parametrize("rmeth", ["7UP"])
def test_adjoint_mhd_boris_background_matches_finite_differences(rmeth):
    """MHD with Boris correction (Ca = 2, so the Boris factor matters) and a
    dipole background field on a periodic box off the origin: the B0 cross
    terms, the Boris vA, mag-HOGS and the CellBoris update in the adjoint."""
    N = (16, 16, 16)
    p = pb.box_params(rmeth)
    p.update(doBoris=1, Ca=2.0, doBackground=1)
    g = gamhip.GamHip(*N, p)
    xyz = pb.cartesian_corners(*N, lo=(1.0, 1.0, 1.0), hi=(2.0, 2.0, 2.0))
    g.grid_from_corners(xyz)
    g.add_dipole_b0(-1.0)
    g.put_state(pb.periodic_box_state(*N, g.get_grid(fields=("xyzcc",))["xyzcc"], xyz))
    g.init_run()
    _generic(g, 13)
    g.run(2)
    assert _fd_check(g, 6, 4, 6, 505, "boris+B0 " + rmeth) <= TOL
    assert max(_fd_check_flux(g, 6, 4, 4, 606, "boris+B0 " + rmeth)) <= 1e-5
    g.close()


def test_adjoint_mhd_shell_walls_matches_finite_differences():
    """The adjoint workload's physics without its inflow: warped shell with
    the dipole background, Boris (Ca = 10), 7UP, zero-gradient r and theta
    walls (the outer r wall rebuilding ghost fluxes from the last cell's B),
    periodic phi, seeded flow and a seeded residual field."""
    N = (24, 16, 24)
    p = pb.shell_params("7UP")
    p["bc"] = ["zerograd", "zerograd", "zerograd", "zerograd", "periodic", "periodic"]
    g = gamhip.GamHip(*N, p)
    xyz = pb.shell_corners(*N)
    g.grid_from_corners(xyz)
    g.add_dipole_b0(pb.SHELL_M0)
    xyzcc = g.get_grid(fields=("xyzcc",))["xyzcc"]
    g.put_state(pb.shell_mhd_state(*N, xyzcc, xyz))
    g.init_run()
    _generic(g, 15)
    g.run(2)
    w = _shell_weights(g)
    # State directions only: with the masked objective the oState
    # sensitivities sit at the difference quotient's noise floor (the oState
    # path is the predictor's AB2 slope, checked on the boxes above)
    assert _fd_check(g, 5, 3, 4, 707, "shell walls", w, ostate=False) <= TOL
    # face-flux directions with the objective over r >= 12: over 3 <= r <= 6
    # J is ~50 and these directional derivatives ~5e-4, so the difference
    # quotients reach their round-off floor before their truncation error
    # has died (measured errors 1e-5 to 1e-4 at every step 1e-3 .. 1e-7);
    # over r >= 12 the same directions agree to ~1e-6 at e = 1e-3
    assert max(_fd_check_flux(g, 5, 3, 3, 808, "shell walls", _outer_weights(g, 12.0), ostate=False)) <= 1e-5
    g.close()


def _wind_series(n=16, seed=0):
    """A time-varying inflow whose lag across the r = 30 boundary spans
    about ten samples; every column varies, the flow is off the x axis and
    the IMF has a Bx from ByC, BzC."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 7)
    s = 2 * np.pi * np.arange(n) / n
    Q = np.zeros((n, 5), order="F")
    Q[:, 0] = 1.0 + 0.2 * np.sin(s + ph[0])
    Q[:, 1] = -0.4 + 0.05 * np.sin(2 * s + ph[1])
    Q[:, 2] = 0.04 * np.sin(s + ph[2])
    Q[:, 3] = 0.03 * np.cos(s + ph[3])
    Q[:, 4] = 1e-3 * (1.0 + 0.3 * np.sin(s + ph[4]))
    B = np.zeros((n, 3), order="F")
    B[:, 1] = 0.02 * np.sin(s + ph[5])
    B[:, 2] = -0.05 + 0.02 * np.cos(s + ph[6])
    return {"tW": np.linspace(-120.0, 120.0, n), "Q": Q, "B": B, "Bx0": 0.01, "ByC": 0.1, "BzC": -0.05,
            "xyzW": (0.0, 0.0, 0.0)}


def _fd_check_wind(g, wind, nsteps, ckpt, seed, tag, weights, field=True):
    """Directions in the inflow series: all columns; density, pressure and
    field only (these leave the lag, a step function of the velocity, in
    place); velocity only.  Best relative agreement over e = 1e-2 .. 1e-6 of
    1 % of each column's scale."""
    S = g.get_state(gamhip.GAM_STATE)
    O = g.get_state(gamhip.GAM_OSTATE)
    s0, o0 = S["Gas"].copy(order="F"), O["Gas"].copy(order="F")
    adj = gamhip.GamAdjoint(g, nsteps, ckpt)
    adj.set_weights(weights)
    adj.forward()
    adj.backward()
    n = len(wind["tW"])
    gQ, gB = adj.wind_grad(n)
    assert np.all(gB[:, 0] == 0.0)
    assert np.count_nonzero(gQ[:, 0]) >= 3, gQ
    if field:   # the inflow field enters through the ghost face fluxes (MHD)
        assert np.count_nonzero(gB[:, 2]) >= 3, gB
    else:
        assert np.all(gB == 0.0)
    rng = np.random.default_rng(seed)
    qs = np.max(np.abs(wind["Q"]), axis=0)
    qs[2:4] = np.max(np.abs(wind["Q"][:, 1]))
    bs = np.max(np.abs(wind["B"]))
    bests = []
    for kind in ("all", "rho-p-B", "velocity"):
        dQ = np.asfortranarray(rng.standard_normal((n, 5)) * qs * 1e-2)
        dB = np.zeros((n, 3), order="F")
        dB[:, 1:] = rng.standard_normal((n, 2)) * bs * 1e-2
        if kind == "rho-p-B":
            dQ[:, 1:4] = 0.0
        elif kind == "velocity":
            dQ[:, [0, 4]] = 0.0
            dB[:] = 0.0
        ad = np.sum(gQ * dQ) + np.sum(gB * dB)
        best = None
        for e in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
            J = []
            for sg in (1.0, -1.0):
                w = dict(wind)
                w["Q"] = wind["Q"] + sg * e * dQ
                w["B"] = wind["B"] + sg * e * dB
                g.set_wind(w)
                J.append(adj.replay(s0, o0))
            fd = (J[0] - J[1]) / (2 * e)
            rel = abs(fd - ad) / max(abs(fd), abs(ad))
            print("%s series dir %s e %.0e: fd %.12e adj %.12e rel %.2e" % (tag, kind, e, fd, ad, rel), flush=True)
            best = rel if best is None else min(best, rel)
        g.set_wind(wind)
        bests.append(best)
    adj.close()
    print("%s series directions, best relative agreement: %s" % (tag, ["%.1e" % b for b in bests]), flush=True)
    return bests


def test_adjoint_mhd_shell_inflow_matches_finite_differences():
    """The adjoint workload's physics: warped shell, dipole background,
    Boris (Ca = 10), 7UP, a time-varying inflow on the outer r face,
    zero-gradient inner r and theta walls, periodic phi, seeded flow and a
    seeded residual field.  Gas and face-flux directions of the start state
    and directions in the inflow series."""
    N = (24, 16, 24)
    g = gamhip.GamHip(*N, pb.shell_params("7UP"))
    xyz = pb.shell_corners(*N)
    g.grid_from_corners(xyz)
    g.add_dipole_b0(pb.SHELL_M0)
    wind = _wind_series()
    g.set_wind(wind)
    xyzcc = g.get_grid(fields=("xyzcc",))["xyzcc"]
    g.put_state(pb.shell_mhd_state(*N, xyzcc, xyz))
    g.init_run()
    _generic(g, 17)
    g.run(2)
    w = _outer_weights(g)
    assert _fd_check(g, 5, 3, 4, 909, "shell inflow", w, ostate=False) <= TOL
    assert max(_fd_check_flux(g, 5, 3, 2, 910, "shell inflow", w, ostate=False)) <= 1e-5
    assert max(_fd_check_wind(g, wind, 5, 3, 911, "shell inflow", w)) <= 1e-5
    g.close()


def test_adjoint_hydro_shell_inflow_matches_finite_differences():
    """The inflow BC's transpose without magnetic terms (hydrodynamic shell,
    the inflow on the outer r face): gas directions of the start state and
    rho / pressure / velocity directions of the series."""
    N = (24, 16, 24)
    p = pb.shell_params("7UP")
    p.update(doMHD=0, doBoris=0, doBackground=0)
    g = gamhip.GamHip(*N, p)
    g.grid_from_corners(pb.shell_corners(*N))
    wind = _wind_series(seed=3)
    g.set_wind(wind)
    g.put_state(pb.shell_state(*N, g.get_grid(fields=("xyzcc",))["xyzcc"], p0=0.1))
    g.init_run()
    _generic(g, 21)
    g.run(2)
    w = _outer_weights(g)
    assert _fd_check(g, 5, 3, 4, 1001, "hydro inflow", w, ostate=False) <= TOL
    assert max(_fd_check_wind(g, wind, 5, 3, 1002, "hydro inflow", w, field=False)) <= 1e-5
    g.close()


def test_adjoint_backward_repeats():
    """The reverse sweep takes the last interval from the recorded forward
    run's tapes; a second sweep recomputes every interval from its
    checkpoint and must give the same gradients (to the order in which
    atomic adds land)."""
    g = box_case((16, 16, 16), "7UP")[0]
    g.init_run()
    _generic(g, 19)
    g.run(1)
    adj = gamhip.GamAdjoint(g, 7, 3)
    adj.set_weights(_weights(g))
    adj.forward()
    first = adj.backward(fluxes=True)
    second = adj.backward(fluxes=True)
    for a, b in zip(first, second):
        scale = np.max(np.abs(a))
        assert scale > 0.0
        assert np.max(np.abs(a - b)) <= 1e-12 * scale
    adj.close()
    g.close()


def test_torch_autograd_bridge():
    """StepObjective (torch.autograd.Function over the C ABI): the tensor
    gradient is the adjoint of State plus oState at the start (both begin
    as gas0), and it matches central differences of frozen-dt replays."""
    import torch
    from mage_amd.autograd import StepObjective
    g = box_case((16, 16, 16), "7UP", doMHD=0)[0]
    g.init_run()
    _generic(g, 3)
    g.run(1)
    base = g.get_state()["Gas"]
    gas0 = torch.from_numpy(np.ascontiguousarray(base)).requires_grad_(True)
    obj = StepObjective(g, 4, 2, _weights(g))
    J = obj(gas0)
    J.backward()
    grad = gas0.grad.numpy()
    assert np.all(grad[:4] == 0.0) and np.all(grad[-4:] == 0.0)   # ghost entries
    d = _directions(base.shape, 1, 13)[0] * np.max(np.abs(base), axis=(0, 1, 2)) * 1e-2
    ad = np.sum(grad * d)
    best = min(abs((obj.adj.replay(base + e * d, base + e * d) - obj.adj.replay(base - e * d, base - e * d))
                   / (2 * e) - ad) / abs(ad) for e in (1e-2, 1e-3, 1e-4, 1e-5))
    print("torch bridge: J %.12e, <grad, d> %.6e, best relative difference %.2e" % (J.item(), ad, best), flush=True)
    assert best <= 1e-4
    obj.close()
    g.close()


def test_torch_bridge_inflow_shell_and_reuse():
    """StepObjective with the inflow series as an input, on the MHD shell
    with a time-varying inflow: the gradients with respect to gas0 and to
    the series match central differences of the bridge's forward with each
    step's dt held (gam_adj_replay with oState's ghosts at State time, as
    the bridge's gam_init_run fills them; the CFL dt itself is not
    differentiated), and J = obj(a) + obj(b) before one backward gives each
    call its own gradient."""
    import torch
    from mage_amd.autograd import StepObjective
    N = (24, 16, 24)
    g = gamhip.GamHip(*N, pb.shell_params("7UP"))
    xyz = pb.shell_corners(*N)
    g.grid_from_corners(xyz)
    g.add_dipole_b0(pb.SHELL_M0)
    wind = _wind_series(seed=5)
    g.set_wind(wind)
    xyzcc = g.get_grid(fields=("xyzcc",))["xyzcc"]
    g.put_state(pb.shell_mhd_state(*N, xyzcc, xyz, p0=0.1))
    g.init_run()
    _generic(g, 23)
    g.run(2)
    base = g.get_state()["Gas"]
    obj = StepObjective(g, 4, 2, _outer_weights(g), wind=True)
    Q0 = torch.from_numpy(np.ascontiguousarray(wind["Q"]))
    B0 = torch.from_numpy(np.ascontiguousarray(wind["B"]))
    gas0 = torch.from_numpy(np.ascontiguousarray(base)).requires_grad_(True)
    Q = Q0.clone().requires_grad_(True)
    B = B0.clone().requires_grad_(True)
    J = obj(gas0, Q, B)
    J.backward()
    ggas, gQ, gB = gas0.grad.numpy(), Q.grad.numpy(), B.grad.numpy()
    assert np.all(ggas[:4] == 0.0)
    rng = np.random.default_rng(29)
    d = _directions(base.shape, 1, 31)[0] * np.max(np.abs(base), axis=(0, 1, 2)) * 1e-2
    dQ = rng.standard_normal(wind["Q"].shape) * np.max(np.abs(wind["Q"]), axis=0) * 1e-2
    dQ[:, 1:4] = 0.0      # keep the lag (a step function of the velocity) in place
    dB = np.zeros(wind["B"].shape)
    dB[:, 1:] = rng.standard_normal((len(wind["tW"]), 2)) * 1e-4
    td = torch.from_numpy(np.ascontiguousarray(d))
    with torch.no_grad():
        obj(gas0.detach(), Q0, B0)           # record the unperturbed run for the replays

    def J_gas(e):
        x = np.asfortranarray(base + e * d)
        return obj.adj.replay(x, x)

    def J_series(e):
        w = dict(wind)
        w["Q"] = wind["Q"] + e * dQ
        w["B"] = wind["B"] + e * dB
        g.set_wind(w)
        b = np.asfortranarray(base)
        J = obj.adj.replay(b, b)
        g.set_wind(wind)
        return J
    for name, ad, fJ in (("gas", np.sum(ggas * d), J_gas), ("series", np.sum(gQ * dQ) + np.sum(gB * dB), J_series)):
        best = None
        for e in (1e-2, 1e-3, 1e-4, 1e-5):
            fd = (fJ(e) - fJ(-e)) / (2 * e)
            rel = abs(fd - ad) / max(abs(fd), abs(ad))
            best = rel if best is None else min(best, rel)
        print("torch bridge shell inflow %s: <grad, d> %.8e best rel %.2e" % (name, ad, best), flush=True)
        assert best <= 1e-5, (name, best)
    # two calls recorded before one backward: each keeps its own gradient
    a = gas0.detach().clone().requires_grad_(True)
    b = (gas0.detach() + 0.5 * td).requires_grad_(True)
    (obj(a, Q0, B0) + obj(b, Q0, B0)).backward()
    gb = b.grad.clone()
    b2 = b.detach().clone().requires_grad_(True)
    obj(b2, Q0, B0).backward()
    assert torch.allclose(a.grad, gas0.grad, rtol=1e-12, atol=0.0)
    assert torch.allclose(gb, b2.grad, rtol=1e-12, atol=0.0)
    assert not torch.allclose(a.grad, gb, rtol=1e-6, atol=0.0)
    obj.close()
    g.close()
