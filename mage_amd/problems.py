"""Grids, initial conditions and model settings of the benchmark workloads.

Host-side input producers (the reference's genGridXML / prob.F90 role,
init.F90:564-790): corner coordinates, seeded initial states and Model
knobs.  All arrays are numpy arrays in the reference's Fortran layout with
ghost lower bound isg = 1 - nG (nG = 4).

Workloads
  periodic_box  64^3 uniform Cartesian periodic box (Fourier-mode rho, v, A)
  shell         log-r spherical shell r in [2.5, 30], theta in
                [0.03pi, 0.97pi], phi periodic, 5% seeded warp of the theta
                lines, dipole background, inflow on the outer shell
Random numbers come from SplitMix64 (Steele, Lea & Flood 2014) seeded from
31337 (kdefs.F90:28), so every run is reproducible bit for bit.
"""
import math

import numpy as np

NG = 4
SEED0 = 31337
MASK64 = (1 << 64) - 1

# 12-point Gauss-Legendre nodes/weights (quadrature.F90:9-26)
GA = np.array([0.1252334085114689154724, 0.3678314989981801937527, 0.5873179542866174472967,
               0.7699026741943046870369, 0.9041172563704748566785, 0.9815606342467192506906,
               -0.1252334085114689154724, -0.3678314989981801937527, -0.5873179542866174472967,
               -0.7699026741943046870369, -0.9041172563704748566785, -0.9815606342467192506906])
GW = np.array([0.2491470458134027850006, 0.233492536538354808761, 0.203167426723065921749,
               0.160078328543346226335, 0.1069393259953184309603, 0.0471753363865118271946] * 2)


class SplitMix64:
    """SplitMix64 generator; uniform() returns doubles in [0,1) from the top 53 bits."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo, hi):
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def model_params(**over):
    """Model_T knobs with the reference defaults (init.F90:312-484)."""
    p = dict(gamma=5.0 / 3.0, CFL=0.3, Vd0=0.5, Ca=1.0e12, cHogH=0.25, cHogM=0.25, pdmb=0.75,
             dFloor=1.0e-12, pFloor=1.0e-12, tFin=1.0e30, dt0=0.0, limDT0=1.0e-3,
             doMHD=1, doBoris=0, doBackground=0, doLFMLim=1, rmeth=0,
             bc=["periodic"] * 6)
    p.update(over)
    return p


def index_ranges(N):
    """Ghost-inclusive corner indices 1-nG .. N+nG+1."""
    return np.arange(1 - NG, N + NG + 2)


# -------------------------------------------------------------- grids
def cartesian_corners(Nip, Njp, Nkp, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Uniform Cartesian corners, x = min + (i-1)*dx (init.F90:663-709)."""
    dx = [(hi[q] - lo[q]) / n for q, n in enumerate((Nip, Njp, Nkp))]
    I = index_ranges(Nip)
    J = index_ranges(Njp)
    K = index_ranges(Nkp)
    x = lo[0] + (I - 1) * dx[0]
    y = lo[1] + (J - 1) * dx[1]
    z = lo[2] + (K - 1) * dx[2]
    xyz = np.zeros((len(I), len(J), len(K), 3), order="F")
    xyz[..., 0] = x[:, None, None]
    xyz[..., 1] = y[None, :, None]
    xyz[..., 2] = z[None, None, :]
    return xyz


def shell_corners(Nip, Njp, Nkp, r0=2.5, r1=30.0, th0=0.03 * math.pi, th1=0.97 * math.pi,
                  warp=0.05, seed=SEED0):
    """Warped spherical-shell corners.

    r is log-spaced over [r0, r1]; theta is uniform over [th0, th1] in the
    active range and continues exponentially towards the poles in the ghost
    layers so that no ghost cell crosses the axis; phi covers [0, 2pi).  The
    theta lines get a seeded non-orthogonal displacement
        dtheta = warp*(th1-th0)*sin(pi*s)*sin(2pi*u + phi + psi)
    with s, u the normalised theta and log-r indices (s clipped to [0,1]).
    """
    rng = SplitMix64(seed)
    psi = 2.0 * math.pi * rng.uniform()
    I = index_ranges(Nip)
    J = index_ranges(Njp)
    K = index_ranges(Nkp)
    u = (I - 1) / Nip
    r = r0 * (r1 / r0) ** u
    s = (J - 1) / Njp
    span = th1 - th0
    th = th0 + s * span
    lo = s < 0
    hi = s > 1
    th[lo] = th0 * np.exp(s[lo] * span / th0)
    th[hi] = math.pi - (math.pi - th1) * np.exp((1 - s[hi]) * span / (math.pi - th1))
    ph = 2.0 * math.pi * (K - 1) / Nkp
    R = r[:, None, None]
    U = u[:, None, None]
    S = s[None, :, None]
    # the warp vanishes outside the active theta range so ghost lines near the
    # poles stay monotone
    TH = th[None, :, None] + warp * span * np.sin(math.pi * np.clip(S, 0.0, 1.0)) * np.sin(
        2 * math.pi * U + ph[None, None, :] + psi)
    PH = ph[None, None, :]
    xyz = np.zeros((len(I), len(J), len(K), 3), order="F")
    xyz[..., 0] = R * np.sin(TH) * np.cos(PH)
    xyz[..., 1] = R * np.sin(TH) * np.sin(PH)
    xyz[..., 2] = R * np.cos(TH)
    return xyz


# -------------------------------------------------------------- states
def empty_state(Nip, Njp, Nkp):
    c = (Nip + 2 * NG, Njp + 2 * NG, Nkp + 2 * NG)
    f = (c[0] + 1, c[1] + 1, c[2] + 1)
    return {"Gas": np.zeros(c + (5,), order="F"), "magFlux": np.zeros(f + (3,), order="F"),
            "Bxyz": np.zeros(c + (3,), order="F"), "Efld": np.zeros(c + (3,), order="F"),
            "time": 0.0}


def prim2con(gamma, rho, vx, vy, vz, p):
    """CellP2C (gamutils.F90:362-383) on arrays."""
    ke = 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, p / (gamma - 1) + ke], axis=-1)


def fourier_modes(rng, n, kmax=4):
    """n seeded modes: integer wavevector 0<|k|<=kmax, phase, unit direction."""
    modes = []
    while len(modes) < n:
        k = np.array([rng.randint(-kmax, kmax) for _ in range(3)], dtype=float)
        nk = np.linalg.norm(k)
        if nk == 0 or nk > kmax:
            continue
        ph = 2 * math.pi * rng.uniform()
        d = np.array([rng.uniform() - 0.5 for _ in range(3)])
        d = d / np.linalg.norm(d)
        modes.append((k, ph, d))
    return modes


def _modesum(modes, X, Y, Z, scale=1.0):
    out = np.zeros(np.broadcast(X, Y, Z).shape)
    for k, ph, _ in modes:
        out += np.cos(2 * math.pi * (k[0] * X + k[1] * Y + k[2] * Z) / scale + ph)
    return out


def _vecsum(modes, X, Y, Z, amp):
    out = [np.zeros(np.broadcast(X, Y, Z).shape) for _ in range(3)]
    for k, ph, d in modes:
        c = amp * np.cos(2 * math.pi * (k[0] * X + k[1] * Y + k[2] * Z) + ph)
        for q in range(3):
            out[q] += d[q] * c
    return out


def vecpot_fluxes(xyz, Afun):
    """Face fluxes from edge line integrals of A (VectorPot2Flux,
    gridutils.F90:193-237) with the 12-point Gauss edge rule
    (GaussianEdgeIntegral, quadrature.F90:347-369)."""
    ni, nj, nk, _ = xyz.shape          # corner counts = Ni+1, ...
    Ni, Nj, Nk = ni - 1, nj - 1, nk - 1
    lA = np.zeros((Ni, Nj, Nk, 3), order="F")
    e0 = xyz[:-2, :-2, :-2]
    for d in range(3):
        sl = [slice(0, -2)] * 3
        sl[d] = slice(1, -1)
        e1 = xyz[tuple(sl)]
        dR2 = 0.5 * (e0 - e1)
        xb = 0.5 * (e0 + e1)
        acc = np.zeros(e0.shape)
        for q in range(12):
            xp = xb + GA[q] * dR2
            A = Afun(xp[..., 0], xp[..., 1], xp[..., 2])
            for c in range(3):
                acc[..., c] += 0.5 * GW[q] * A[c]
        lA[:-1, :-1, :-1, d] = np.sum(acc * (e1 - e0), axis=-1)
    mf = np.zeros((Ni + 1, Nj + 1, Nk + 1, 3), order="F")
    a = lA
    s = (slice(0, Ni - 1), slice(0, Nj - 1), slice(0, Nk - 1))
    ip = (slice(1, Ni), slice(0, Nj - 1), slice(0, Nk - 1))
    jp = (slice(0, Ni - 1), slice(1, Nj), slice(0, Nk - 1))
    kp = (slice(0, Ni - 1), slice(0, Nj - 1), slice(1, Nk))
    mf[s + (0,)] = a[s + (1,)] - a[kp + (1,)] + a[jp + (2,)] - a[s + (2,)]
    mf[s + (1,)] = -(a[s + (0,)] - a[kp + (0,)] + a[ip + (2,)] - a[s + (2,)])
    mf[s + (2,)] = a[s + (0,)] - a[jp + (0,)] + a[ip + (1,)] - a[s + (1,)]
    return mf


def periodic_box_state(Nip, Njp, Nkp, xyzcc, xyz, gamma=5.0 / 3.0, seed=SEED0):
    """Seeded periodic-box IC on cell centres xyzcc(Ni,Nj,Nk,3):
    rho = 1 + 0.1*sum_8 cos(2pi k.x + phi), p = 1, v = 8 modes of amplitude
    0.1, B = curl A with A = 8 modes of amplitude 0.05 (face fluxes by edge
    integration, so div B vanishes to round-off).  Point values at the
    volume centroids."""
    rng = SplitMix64(seed)
    rmodes = fourier_modes(rng, 8)
    vmodes = fourier_modes(rng, 8)
    amodes = fourier_modes(rng, 8)
    st = empty_state(Nip, Njp, Nkp)
    X, Y, Z = xyzcc[..., 0], xyzcc[..., 1], xyzcc[..., 2]
    rho = 1.0 + 0.1 * _modesum(rmodes, X, Y, Z)
    vx, vy, vz = _vecsum(vmodes, X, Y, Z, 0.1)
    p = np.ones_like(rho)
    st["Gas"][...] = prim2con(gamma, rho, vx, vy, vz, p)
    st["magFlux"][...] = vecpot_fluxes(xyz, lambda x, y, z: _vecsum(amodes, x, y, z, 0.05))
    return st


def shell_state(Nip, Njp, Nkp, xyzcc, gamma=5.0 / 3.0, seed=SEED0, p0=1.0e-3):
    """Shell IC: rho = 1, v = 0, residual B = 0, p = p0*(1 + 0.5*f) with f a
    seeded smooth field in [-1,1] (8 modes on the 60-wide box)."""
    rng = SplitMix64(seed + 1)
    modes = fourier_modes(rng, 8)
    st = empty_state(Nip, Njp, Nkp)
    X, Y, Z = xyzcc[..., 0], xyzcc[..., 1], xyzcc[..., 2]
    f = _modesum(modes, X, Y, Z, scale=60.0) / 8.0
    rho = np.ones_like(X)
    p = p0 * (1.0 + 0.5 * f)
    z = np.zeros_like(X)
    st["Gas"][...] = prim2con(gamma, rho, z, z, z, p)
    return st


def shell_mhd_state(Nip, Njp, Nkp, xyzcc, xyz, gamma=5.0 / 3.0, seed=SEED0, p0=0.05, vamp=0.3,
                    aamp=0.3):
    """Shell IC with every MHD term active: rho = 1 + 0.2*f1, p = p0*(1 +
    0.5*f2), v = vamp * (8 seeded modes)/8 per component, residual B from a
    seeded vector potential (8 modes of amplitude aamp on the 60-wide box)
    through edge line integrals, so div(dB) vanishes to round-off.  The f are
    seeded smooth fields in [-1, 1]; |v| spans ~0.05-0.3 across the shell."""
    rng = SplitMix64(seed + 2)
    rmodes = fourier_modes(rng, 8)
    pmodes = fourier_modes(rng, 8)
    vmodes = [fourier_modes(rng, 8) for _ in range(3)]
    amodes = fourier_modes(rng, 8)
    st = empty_state(Nip, Njp, Nkp)
    X, Y, Z = xyzcc[..., 0], xyzcc[..., 1], xyzcc[..., 2]
    rho = 1.0 + 0.2 * _modesum(rmodes, X, Y, Z, scale=60.0) / 8.0
    p = p0 * (1.0 + 0.5 * _modesum(pmodes, X, Y, Z, scale=60.0) / 8.0)
    v = [vamp * _modesum(vm, X, Y, Z, scale=60.0) / 8.0 for vm in vmodes]
    st["Gas"][...] = prim2con(gamma, rho, v[0], v[1], v[2], p)

    def A(x, y, z):
        out = [np.zeros(np.broadcast(x, y, z).shape) for _ in range(3)]
        for k, ph, d in amodes:
            c = aamp * np.cos(2 * math.pi * (k[0] * x + k[1] * y + k[2] * z) / 60.0 + ph)
            for q in range(3):
                out[q] += d[q] * c
        return out
    st["magFlux"][...] = vecpot_fluxes(xyz, A)
    return st


def constant_wind(rho=1.0, vx=-0.4, p=1.0e-3, bz=-0.05):
    """A two-sample constant inflow series covering every time the runs reach."""
    tW = np.array([-1.0e3, 1.0e3])
    Q = np.zeros((2, 5), order="F")
    Q[:, 0] = rho
    Q[:, 1] = vx
    Q[:, 4] = p
    B = np.zeros((2, 3), order="F")
    B[:, 2] = bz
    return {"tW": tW, "Q": Q, "B": B, "Bx0": 0.0, "ByC": 0.0, "BzC": 0.0, "xyzW": (0.0, 0.0, 0.0)}


def varying_wind(n=200, t0=-100.0, t1=100.0, seed=SEED0):
    """An n-sample inflow series on [t0, t1] whose density, velocity,
    pressure and IMF vary smoothly (a few harmonics each) around a
    southward-IMF solar wind like constant_wind's; the flow is a little off
    the x axis and ByC, BzC give the IMF a Bx."""
    rng = np.random.default_rng(seed)
    s = np.linspace(0.0, 2.0 * np.pi, n)

    def wave(amp):
        ph = rng.uniform(0, 2 * np.pi, 3)
        return amp * (np.sin(s + ph[0]) + 0.5 * np.sin(2 * s + ph[1]) + 0.25 * np.sin(3 * s + ph[2])) / 1.75
    Q = np.zeros((n, 5), order="F")
    Q[:, 0] = 1.0 + wave(0.2)
    Q[:, 1] = -0.4 + wave(0.05)
    Q[:, 2] = wave(0.04)
    Q[:, 3] = wave(0.03)
    Q[:, 4] = 1.0e-3 * (1.0 + wave(0.3))
    B = np.zeros((n, 3), order="F")
    B[:, 1] = wave(0.02)
    B[:, 2] = -0.05 + wave(0.02)
    return {"tW": np.linspace(t0, t1, n), "Q": Q, "B": B, "Bx0": 0.01, "ByC": 0.1, "BzC": -0.05,
            "xyzW": (0.0, 0.0, 0.0)}


# Dipole moment giving |B0| = 1 at r = 2.5 on the equator, oriented like the
# Earth's (equatorial B0 along +z): M0 = -(2.5)^3.
SHELL_M0 = -(2.5 ** 3)


def ensemble_members(n=32, seed=SEED0, lo=(48, 48, 64), hi=(96, 96, 128)):
    """The ensemble workload: n shell runs of uneven size.

    Grid sizes come from SplitMix64(seed) as multiples of 8 in
    [lo, hi] per axis (i, j, k drawn in that order, member by member).
    Member m uses seed + m for its grid warp and initial state, and draws
    its constant inflow from SplitMix64(seed + m): vx in [-0.8, -0.3],
    Bz in [-0.1, 0.1].  Returns dicts with N, seed, vx, bz.
    """
    rng = SplitMix64(seed)
    out = []
    for m in range(n):
        N = tuple(8 * rng.randint(lo[a] // 8, hi[a] // 8) for a in range(3))
        r = SplitMix64(seed + m)
        vx = -0.8 + 0.5 * r.uniform()
        bz = -0.1 + 0.2 * r.uniform()
        out.append({"N": N, "seed": seed + m, "vx": vx, "bz": bz})
    return out


def box_params(rmeth="7UP"):
    return model_params(rmeth={"8CENT": 0, "7UP": 1}[rmeth], bc=["periodic"] * 6)


def shell_params(rmeth="7UP"):
    return model_params(rmeth={"8CENT": 0, "7UP": 1}[rmeth], doBackground=1, doBoris=1, Ca=10.0,
                        bc=["zerograd", "wind", "zerograd", "zerograd", "periodic", "periodic"])


# -------------------------------------------------------------- reference case decks
# The reference's whole-step case tests (tests/cases/*.xml, checked by
# tests/cases/testCases.pf) with the initial states of prob.F90.  Decks with
# one cell along j or k run here with 8 uniform cells along that axis and
# periodic walls: every flux and EMF difference along it is then exactly
# zero, and the dt bound of the collapsed axis stays above the in-plane one.
def cfl_from_pdmb(pdmb, max_cfl=0.3):
    """Default CFL from the PDM parameter (init.F90:348-353)."""
    return min(0.5 / (pdmb + 0.5), max_cfl)


CASES = {
    # name: cells, domain lo/hi, tFin, Model_T overrides
    "adv1d": ((128, 8, 8), (-1.0, -0.5, -0.5), (1.0, 0.5, 0.5), 2.0,
              dict(gamma=1.4, pdmb=4.0, doMHD=0, rmeth=1)),
    "adv1d_8cent": ((128, 8, 8), (-1.0, -0.5, -0.5), (1.0, 0.5, 0.5), 2.0,
                    dict(gamma=1.4, pdmb=4.0, doMHD=0, rmeth=0)),
    "adv2d": ((256, 256, 8), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 1.0, dict(pdmb=4.0)),
    "loop2d": ((256, 256, 8), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 1.0, dict(pdmb=4.0)),
    "alfven": ((128, 256, 8), (0.0, 0.0, -0.5), (2.0, 1.15470053838, 0.5), 1.0, {}),
    "ot2d": ((128, 128, 8), (0.0, 0.0, -0.5), (1.0, 1.0, 0.5), 1.0, {}),
}


def case_params(name):
    N, lo, hi, tFin, over = CASES[name]
    p = model_params(tFin=tFin, bc=["periodic"] * 6)
    p.update(over)
    p["CFL"] = cfl_from_pdmb(p["pdmb"])
    return p


def case_corners(name):
    N, lo, hi, _, _ = CASES[name]
    return cartesian_corners(*N, lo=lo, hi=hi)


def _gauss_box_average(xyz, fun, nq=12):
    """Cell averages of fun(x, y) over axis-aligned cells by the nq x nq
    Gauss rule in the plane (GaussianVolumeIntegral/volume, quadrature.F90,
    for functions that do not depend on z).  fun returns a tuple of arrays."""
    x0, x1 = xyz[:-1, 0, 0, 0], xyz[1:, 0, 0, 0]
    y0, y1 = xyz[0, :-1, 0, 1], xyz[0, 1:, 0, 1]
    acc = None
    for a in range(nq):
        xq = 0.5 * (x0 + x1) + 0.5 * GA[a] * (x1 - x0)
        for b in range(nq):
            yq = 0.5 * (y0 + y1) + 0.5 * GA[b] * (y1 - y0)
            vals = fun(xq[:, None], yq[None, :])
            w = 0.25 * GW[a] * GW[b]
            if acc is None:
                acc = [w * v for v in vals]
            else:
                for q, v in enumerate(vals):
                    acc[q] = acc[q] + w * v
    return acc


def case_state(name, grid, xyz):
    """Initial State of a reference case deck (prob.F90)."""
    p = case_params(name)
    gam = p["gamma"]
    st = empty_state(*[n - 2 * NG for n in grid["xyzcc"].shape[:3]])
    xc, yc = grid["xyzcc"][..., 0], grid["xyzcc"][..., 1]
    zero = np.zeros_like(xc)
    pi = math.pi
    if name.startswith("adv1d"):
        # initADV1D (prob.F90:671-736): four shapes on rho0 = 0.1, vx = 1, P = 1
        r1 = np.exp(-((xc + 0.7) / 0.05) ** 2)
        r2 = np.where(np.abs(xc + 0.3) <= 0.1, 1.0, 0.0)
        r3 = np.where(np.abs(xc - 0.1) <= 0.1, 1.0 - np.abs(10.0 * (xc - 0.1)), 0.0)
        arg = 0.1 ** 2 - (xc - 0.5) ** 2
        r4 = np.where(arg >= 0, 10.0 * np.sqrt(np.maximum(arg, 0.0)), 0.0)
        rho = 0.1 + r1 + r2 + r3 + r4
        st["Gas"][...] = prim2con(gam, rho, zero + 1.0, zero, zero, zero + 1.0)
    elif name in ("adv2d", "loop2d"):
        # initADV2D (prob.F90:739-796) / initLoop (prob.F90:436-500)
        V0, alpha = 1.41421356237, 45.0
        vx = V0 * math.sin(alpha * pi / 180.0)
        vy = V0 * math.cos(alpha * pi / 180.0)
        rho = np.ones_like(xc)
        if name == "adv2d":
            rho = np.where(np.sqrt(xc ** 2 + yc ** 2) < 0.3, 2.0, 1.0)
        st["Gas"][...] = prim2con(gam, rho, zero + vx, zero + vy, zero, zero + 1.0)
        if name == "loop2d":
            def A(x, y, z):
                r = np.sqrt(x ** 2 + y ** 2)
                return (0 * x, 0 * x, np.maximum(0.0, 1.0e-3 * (0.3 - r)))
            st["magFlux"][...] = vecpot_fluxes(xyz, A)
    elif name == "alfven":
        # initALFVEN (prob.F90:1027-1083): GasIC2State + VectorPot2Flux
        alp, Bp = 1.0471975512, 1.0
        ca, sa = math.cos(alp), math.sin(alp)

        def W(x, y):
            ph = 2 * pi * (x * ca + y * sa)
            return (1.0 + 0 * x, -0.1 * np.sin(ph) * sa, 0.1 * np.sin(ph) * ca, 0.1 * np.cos(ph),
                    0.1 + 0 * x)
        D, vx, vy, vz, P = _gauss_box_average(xyz, W)
        mx, my, mz = D * vx, D * vy, D * vz
        g = st["Gas"]
        g[..., 0] = D[:, :, None]
        g[..., 1] = mx[:, :, None]
        g[..., 2] = my[:, :, None]
        g[..., 3] = mz[:, :, None]
        g[..., 4] = (P / (gam - 1.0) + 0.5 * (mx ** 2 + my ** 2 + mz ** 2) / D)[:, :, None]

        def A(x, y, z):
            ph = 2 * pi * (x * ca + y * sa)
            return (Bp * sa * z, -1.0 * Bp * ca * z + 0.1 * np.sin(ph) / (2 * pi * ca),
                    0.1 * np.cos(ph) / (2 * pi))
        st["magFlux"][...] = vecpot_fluxes(xyz, A)
    elif name == "ot2d":
        # initOT2D (prob.F90:74-119)
        rho0, P0, B0 = 25.0 / (36.0 * pi), 5.0 / (12.0 * pi), 1.0 / math.sqrt(4.0 * pi)
        vx, vy = -np.sin(2 * pi * yc), np.sin(2 * pi * xc)
        st["Gas"][...] = prim2con(gam, zero + rho0, vx, vy, zero, zero + P0)
        f = st["magFlux"]
        face = grid["face"]
        f[:-1, :-1, :-1, 0] = (-B0 * np.sin(2 * pi * yc)) * face[:-1, :-1, :-1, 0]
        f[:-1, :-1, :-1, 1] = (B0 * np.sin(4 * pi * xc)) * face[:-1, :-1, :-1, 1]
    else:
        raise KeyError(name)
    return st
