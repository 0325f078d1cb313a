"""Pins of the oracle's step functions (CPU) by what the mathematics fixes.

tests/test_oracle_pins.py pins the cell-level pieces (conversions,
reconstruction weights, PDM, the two kinetic fluxes, metrics, CellDT); this
file pins the functions that combine them inside AdvanceMHD and the
boundary conditions:

  CalcElecField (fields.F90:70-232)     exact EMF of a uniform state on rotated
                                        and sheared grids; the dissipative
                                        speed |v_perp| + vA, its CFL cap and
                                        its Boris limit, from scalings that
                                        hold edge by edge
  ebGeom Teb (init.F90:1075-1137)       closed form on a sheared grid
  gamCellBxyz (gridutils.F90:648-680)   uniform B recovered exactly on planar
                                        faces, at second order on the warped
                                        shell
  Fluxes (stress.F90:204-429)           L = R: area x (Euler flux + Maxwell
                                        stress with B0 terms), no HOGS or
                                        mag-HOGS term; the nuke term
                                        -Ca*dU*A fires exactly when
                                        sqrt((|vL|^2+|vR|^2)/2) >= 1.5 Ca
  CellBoris (mhdgroup.F90:336-441)      parallel force unchanged, perpendicular
                                        force / (1+alpha), Ca -> inf plain sum
  Predictor (mhdgroup.F90:444-615)      primitives and fluxes linear in time
                                        give the exact half state
  AddB0 (background.F90:20-167)         bFlux0 = loop integral of A = M x r/r^3
                                        (Stokes), fcB0/edgB0/B0 against
                                        independent quadratures
  InterpWind/RefreshWind/GetWindAt/     sample choice with discrete walls, the
  wgtWind (apps/wind.F90:114-203,       front normal, the lag, the ramp
  766-868)

Each test names the mistake it is built to catch.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle as po
from mage_amd import problems as pb

L = po.lib()
DP = po.DP
NG = 4
GAM = 5.0 / 3.0


def vp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(DP)


# ------------------------------------------------------------ grids and states
def rotation(seed):
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


SHEAR = np.array([[1.0, 0.31, -0.17], [0.12, 0.93, 0.26], [-0.21, 0.14, 1.08]])


def affine_corners(N, A, x0=(0.3, -0.2, 0.1)):
    """Corners x = x0 + A @ (i, j, k) over the ghost-inclusive index ranges;
    the columns of A are the lattice vectors of one index step."""
    I = pb.index_ranges(N[0]).astype(float)
    J = pb.index_ranges(N[1]).astype(float)
    K = pb.index_ranges(N[2]).astype(float)
    xyz = np.zeros((len(I), len(J), len(K), 3), order="F")
    for q in range(3):
        xyz[..., q] = (x0[q] + A[q, 0] * I[:, None, None] + A[q, 1] * J[None, :, None]
                       + A[q, 2] * K[None, None, :])
    return xyz


def vector_areas(xyz):
    """Vector area of every face of the corner grid, orientation +d, as the
    exact integral of n dA over the bilinear face: half the cross product of
    its diagonals.  Shape (Ni+1, Nj+1, Nk+1, 3 comp, 3 dir); the last
    corner layer of each direction stays zero."""
    ni, nj, nk, _ = xyz.shape
    out = np.zeros((ni, nj, nk, 3, 3), order="F")
    p = xyz
    # I faces: corners (j,k) (j+1,k) (j+1,k+1) (j,k+1); J x K = +I
    d1 = p[:, 1:, 1:] - p[:, :-1, :-1]
    d2 = p[:, :-1, 1:] - p[:, 1:, :-1]
    out[:, :-1, :-1, :, 0] = 0.5 * np.cross(d1, d2)
    # J faces: corners (k,i) (k+1,i) (k+1,i+1) (k,i+1); K x I = +J
    d1 = p[1:, :, 1:] - p[:-1, :, :-1]
    d2 = p[1:, :, :-1] - p[:-1, :, 1:]
    out[:-1, :, :-1, :, 1] = 0.5 * np.cross(d1, d2)
    # K faces: corners (i,j) (i+1,j) (i+1,j+1) (i,j+1); I x J = +K
    d1 = p[1:, 1:, :] - p[:-1, :-1, :]
    d2 = p[:-1, 1:, :] - p[1:, :-1, :]
    out[:-1, :-1, :, :, 2] = 0.5 * np.cross(d1, d2)
    return out


def uniform_state(N, grid, xyz, rho, v, P, B):
    st = pb.empty_state(*N)
    st["Gas"][...] = pb.prim2con(GAM, rho, v[0], v[1], v[2], P)
    Av = vector_areas(xyz)
    for d in range(3):
        st["magFlux"][..., d] = Av[..., :, d] @ np.asarray(B, dtype=float)
    return st


def oracle_on(N, xyz, grid=None, **over):
    grid = grid if grid is not None else po.corners2grid(*N, xyz)
    grid["xyz"] = xyz
    p = pb.box_params("7UP")
    p.update(over)
    return po.Oracle(*N, p, grid), grid, p


def act(N):
    return tuple(slice(NG, NG + n) for n in N)


def edge_box(N, eD):
    """Active edges of direction eD (fields.F90:100-175), numpy slices."""
    hi = [NG + n for n in N]
    sl = [slice(NG, h + 1) for h in hi]
    sl[eD] = slice(NG, hi[eD])
    return tuple(sl)


def uniform_background(N, grid, xyz, B0):
    """Background arrays of a uniform field B0, built without the oracle's
    AddB0: face averages, vector-area fluxes, edge components along the edge
    triad, cell values, and no stress divergence."""
    c = tuple(n + 2 * NG for n in N)
    B0 = np.asarray(B0, dtype=float)
    Av = vector_areas(xyz)
    g = {"fcB0": np.zeros(c + (3, 3), order="F"), "edgB0": np.zeros(c + (2, 3), order="F"),
         "B0": np.zeros(c + (3,), order="F"), "bFlux0": np.zeros(c + (3,), order="F"),
         "dpB0": np.zeros(c + (3,), order="F")}
    for d in range(3):
        g["fcB0"][..., :, d] = B0
        g["bFlux0"][..., d] = (Av[:-1, :-1, :-1, :, d] @ B0)
        Te = grid["Te"][:-1, :-1, :-1, :, d]
        g["edgB0"][..., 0, d] = Te[..., 3:6] @ B0
        g["edgB0"][..., 1, d] = Te[..., 6:9] @ B0
    g["B0"][...] = B0
    return g


# ------------------------------------------------------------ CalcElecField
@pytest.mark.parametrize("kind", ["rotated", "sheared"])
@pytest.mark.parametrize("bg", [False, True])
def test_emf_of_uniform_state_is_minus_v_cross_b(kind, bg):
    """CalcElecField (fields.F90:70-232) on a uniform state: every active
    edge gets E = -(v x (B + B0)) . a_e, a_e the edge vector, with no
    dissipative part (L = R, so Jd = 0) whatever Vd0.  On a rotated and on a
    sheared lattice the face normals, the edge triads Te and the 2x2 map Teb
    all differ from the axes.  Catches a transposed Teb inverse, swapped
    TAN1/TAN2 projections, a sign error in v1*b2 - v2*b1, a missing edge
    length, a dropped or mis-projected edgB0, and a dissipative term that
    does not vanish for L = R."""
    N = (16, 8, 8)
    h = 0.05
    A = h * (rotation(7) if kind == "rotated" else SHEAR)
    xyz = affine_corners(N, A)
    rho, v, P, B = 1.3, np.array([0.31, -0.22, 0.17]), 0.8, np.array([0.41, -0.33, 0.52])
    B0 = np.array([-0.27, 0.45, 0.19]) if bg else np.zeros(3)
    for Vd0 in (0.5, 7.0):
        o, grid, p = oracle_on(N, xyz, Vd0=Vd0, doBackground=int(bg))
        if bg:
            grid.update(uniform_background(N, grid, xyz, B0))
            o, grid, p = oracle_on(N, xyz, grid=grid, Vd0=Vd0, doBackground=1)
        st = uniform_state(N, grid, xyz, rho, v, P, B)
        o.set_state(st)
        o.init_run()
        o.advance(o.M.dt)
        E = o.st["Efld"]
        scale = np.linalg.norm(v) * np.linalg.norm(B + B0) * h
        for eD in range(3):
            ref = -np.dot(np.cross(v, B + B0), A[:, eD])
            got = E[edge_box(N, eD) + (eD,)]
            assert np.max(np.abs(got - ref)) <= 1e-12 * scale, (kind, bg, Vd0, eD)


def _emf_runs(N=(16, 8, 8), v=(0.0, 0.0, 0.0), rho=1.0, bscale=1.0, dt=1e-9, **over):
    """Edge EMFs of a smooth non-uniform field (seeded vector potential) on a
    uniform-density box at rest or in uniform motion."""
    xyz = pb.cartesian_corners(*N, hi=(1.0, 0.5, 0.5))
    o, grid, p = oracle_on(N, xyz, **over)
    rng = pb.SplitMix64(4242)
    modes = pb.fourier_modes(rng, 6, kmax=2)
    st = pb.empty_state(*N)
    st["Gas"][...] = pb.prim2con(GAM, np.full(grid["volume"].shape, rho), v[0], v[1], v[2],
                                 np.full(grid["volume"].shape, 0.6))
    st["magFlux"][...] = bscale * pb.vecpot_fluxes(
        xyz, lambda x, y, z: pb._vecsum(modes, 2 * x, 2 * y, 2 * z, 0.05))
    o.set_state(st)
    o.init_run()
    o.M.dt = dt
    o.advance(dt)
    return o.st["Efld"].copy(), grid


def _dissipative(eD, **kw):
    """E(Vd0=1) - E(Vd0=0) on the active edges of direction eD:
    min(vDiff, CFL*l/dt) * Jd * l."""
    E1, grid = _emf_runs(Vd0=1.0, **kw)
    E0, _ = _emf_runs(Vd0=0.0, **kw)
    s = edge_box((16, 8, 8), eD) + (eD,)
    return E1[s] - E0[s], E0[s], grid


def test_emf_dissipative_speed_and_cfl_cap():
    """vDiff = |v_perp| + vA, capped at CFL*l/Model%dt (fields.F90:204-218),
    checked edge by edge through scalings of the dissipative part
    D = E(Vd0=1) - E(Vd0=0) = min(vDiff, CFL l/dt) Jd l:
      * at rest and uncapped, density x4 halves D (vA = |b|/sqrt(rho_c));
        flux x2 multiplies D by 4 (vA and Jd both linear in B);
      * in uniform motion, D(v) - D(0) = |v_perp| Jd l, with |v_perp| the
        velocity component across the edge, so the ideal part is absent
        from D;
      * with the cap, D(dt) = Jd l min(vA, CFL l/dt) exactly, with Jd l
        and vA taken from the uncapped runs, for a dt at which some edges
        are capped and some are not.
    Catches a vA without the square root of density, a cap in the wrong
    units (CFL*dt/l), a cap applied to vA alone or missing, and a
    dissipative speed that includes the parallel velocity."""
    for eD in range(3):
        D0, _, _ = _dissipative(eD)
        mask = np.abs(D0) > 1e-6 * np.max(np.abs(D0))
        assert mask.sum() > 0.5 * mask.size
        D4, _, _ = _dissipative(eD, rho=4.0)
        assert np.max(np.abs(D4[mask] - 0.5 * D0[mask])) <= 1e-12 * np.max(np.abs(D0))
        D2, _, _ = _dissipative(eD, bscale=2.0)
        assert np.max(np.abs(D2[mask] - 4.0 * D0[mask])) <= 1e-12 * np.max(np.abs(D2))
        v = np.array([0.07, -0.05, 0.04])
        Dv, _, _ = _dissipative(eD, v=v)
        vperp = math.sqrt(np.sum(v ** 2) - v[eD] ** 2)
        Jl = (Dv - D0) / vperp                    # Jd * l
        vA = np.where(mask, D0 / np.where(mask, Jl, 1.0), 0.0)
        assert np.all(vA[mask] > 0)
        # the EMF edge length along eD
        le = [1.0 / 16, 0.5 / 8, 0.5 / 8][eD]
        # a dt that caps about half of the edges
        cfl = 0.3
        dt = cfl * le / float(np.median(vA[mask]))
        Dc, _, _ = _dissipative(eD, dt=dt)
        ref = Jl * np.minimum(vA, cfl * le / dt)
        capped = mask & (vA > cfl * le / dt)
        assert 0.2 * mask.sum() < capped.sum() < 0.8 * mask.sum()
        assert np.max(np.abs(Dc[mask] - ref[mask])) <= 1e-10 * np.max(np.abs(D0))


def test_emf_boris_alfven_speed():
    """With Boris on, vA -> Ca vA / sqrt(Ca^2 + vA^2) in the dissipative
    speed (fields.F90:206-208): at rest, D_Boris / D = Ca / sqrt(Ca^2 + vA^2)
    edge by edge, tending to Ca/vA when vA >> Ca and to 1 when Ca >> vA.
    Catches a Boris factor applied to the wrong speed, Ca^2 in place of Ca,
    or a missing square root."""
    for eD in range(3):
        D0, _, _ = _dissipative(eD)
        mask = np.abs(D0) > 1e-6 * np.max(np.abs(D0))
        v = np.array([0.07, -0.05, 0.04])
        Dv, _, _ = _dissipative(eD, v=v)
        vperp = math.sqrt(np.sum(v ** 2) - v[eD] ** 2)
        vA = np.where(mask, D0 * vperp / np.where(mask, Dv - D0, 1.0), 0.0)
        for Ca in (0.5 * float(np.median(vA[mask])), 1e-3, 1e6):
            Db, _, _ = _dissipative(eD, doBoris=1, Ca=Ca)
            ratio = Db[mask] / D0[mask]
            ref = Ca / np.sqrt(Ca * Ca + vA[mask] ** 2)
            assert np.max(np.abs(ratio - ref)) <= 1e-9, (eD, Ca)


def test_emf_dissipative_term_removes_magnetic_energy():
    """The dissipative EMF Vd0*vDiff*Jd*l (fields.F90:218-220) is Ohm's law
    E = eta J along the edge (Jd = db2 - db1 is the edge current), so with
    the gas at rest -- no ideal EMF -- the constrained-transport update
    lowers the magnetic energy sum |B|^2 V, the more the larger Vd0.
    Catches a dissipative term with the wrong sign (which the scaling pins
    above cannot see) or Jd taken as db1 - db2."""
    N = (16, 8, 8)
    W = []
    for Vd0 in (0.0, 1.0, 2.0):
        xyz = pb.cartesian_corners(*N, hi=(1.0, 0.5, 0.5))
        o, grid, p = oracle_on(N, xyz, Vd0=Vd0)
        rng = pb.SplitMix64(4242)
        modes = pb.fourier_modes(rng, 6, kmax=2)
        st = pb.empty_state(*N)
        one = np.ones(grid["volume"].shape)
        st["Gas"][...] = pb.prim2con(GAM, one, 0 * one, 0 * one, 0 * one, 0.6 * one)
        st["magFlux"][...] = pb.vecpot_fluxes(
            xyz, lambda x, y, z: pb._vecsum(modes, 2 * x, 2 * y, 2 * z, 0.05))
        o.set_state(st)
        o.init_run()
        o.advance(o.M.dt)
        a = act(N)
        W.append(float(np.sum(np.sum(o.st["Bxyz"][a] ** 2, axis=-1) * grid["volume"][a])))
    assert W[1] < W[0] and W[2] < W[1], W


# ------------------------------------------------------------ Teb
def test_teb_closed_form_on_sheared_grid():
    """ebGeom (init.F90:1075-1137) on an affinely sheared lattice: every face
    of direction d has the normal n_d = a_T1 x a_T2 / |.|, and the edge triad
    of direction e is eN = a_e/|a_e|, eT1 = (eAvg x eN)/|.| with eAvg along
    the third lattice vector, eT2 = eN x eT1.  Teb(XNQI:YNQI, e) is the
    normalised (n_T1.eT1, n_T1.eT2) and Teb(XNQJ:YNQJ, e) the same for
    n_T2.  Catches T1/T2 swapped between the two Teb pairs, the wrong face
    direction, or projections onto the wrong triad vectors."""
    N = (16, 8, 8)
    A = 0.05 * SHEAR
    xyz = affine_corners(N, A)
    g = po.corners2grid(*N, xyz)
    a = [A[:, q] for q in range(3)]
    T = {0: (1, 2), 1: (2, 0), 2: (0, 1)}
    n = [np.cross(a[T[d][0]], a[T[d][1]]) for d in range(3)]
    n = [x / np.linalg.norm(x) for x in n]
    eavg = {0: a[2], 1: a[0], 2: a[1]}   # init.F90:1000-1017
    for e in range(3):
        eN = a[e] / np.linalg.norm(a[e])
        t1 = np.cross(eavg[e], eN)
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(eN, t1)
        t2 /= np.linalg.norm(t2)
        s = edge_box(N, e)
        assert np.max(np.abs(g["Te"][s + (slice(3, 6), e)] - t1)) <= 1e-14
        assert np.max(np.abs(g["Te"][s + (slice(6, 9), e)] - t2)) <= 1e-14
        for q, fd in enumerate(T[e]):
            x, y = n[fd] @ t1, n[fd] @ t2
            nrm = math.hypot(x, y)
            ref = np.array([x, y]) / nrm
            sl = tuple(slice(NG, NG + nn + 1) for nn in N) + (slice(2 * q, 2 * q + 2), e)
            got = g["Teb"][sl]
            assert np.max(np.abs(got - ref)) <= 1e-13, (e, q)


# ------------------------------------------------------------ gamCellBxyz
@pytest.mark.parametrize("kind", ["rotated", "sheared"])
def test_cell_b_exact_for_uniform_field_on_planar_faces(kind):
    """gamCellBxyz (gridutils.F90:648-680): B = [sum +-Phi_f x_f - (div Phi)
    x_c]/V.  For a uniform field and planar faces, Phi_f x_f summed over the
    closed surface is exactly B V, so every cell returns B to round-off.
    Catches a wrong sign of a face term, a face paired with the wrong
    centroid, or a missing divergence correction."""
    N = (16, 8, 8)
    A = 0.05 * (rotation(3) if kind == "rotated" else SHEAR)
    xyz = affine_corners(N, A)
    g = po.corners2grid(*N, xyz)
    g["xyz"] = xyz
    B = np.array([0.41, -0.33, 0.52])
    mf = np.zeros(tuple(s for s in xyz.shape[:3]) + (3,), order="F")
    Av = vector_areas(xyz)
    for d in range(3):
        mf[..., d] = Av[..., :, d] @ B
    o = po.Oracle(*N, pb.box_params(), g)
    out = np.zeros(g["volume"].shape + (3,), order="F")
    L.orc_bflux2fld(C.byref(o.D), C.byref(o.M), C.byref(o.G), po.ptr(mf), po.ptr(out))
    interior = tuple(slice(0, n - 1) for n in out.shape[:3])
    assert np.max(np.abs(out[interior] - B)) <= 1e-13


def _shell_cell_b_error(N):
    xyz = pb.shell_corners(*N)
    g = po.corners2grid(*N, xyz)
    B = np.array([0.41, -0.33, 0.52])
    Av = vector_areas(xyz)
    mf = np.zeros(xyz.shape[:3] + (3,), order="F")
    for d in range(3):
        mf[..., d] = Av[..., :, d] @ B
    o = po.Oracle(*N, pb.box_params(), g)
    out = np.zeros(g["volume"].shape + (3,), order="F")
    L.orc_bflux2fld(C.byref(o.D), C.byref(o.M), C.byref(o.G), po.ptr(mf), po.ptr(out))
    return np.max(np.linalg.norm(out[act(N)] - B, axis=-1)) / np.linalg.norm(B)


def test_cell_b_uniform_field_on_warped_shell_converges():
    """On the warped shell the faces are not planar and the area centroids
    stand in for the flux-weighted ones, so a uniform B comes back with an
    error of second order in the cell size: small at 16x16x16 and about
    four times smaller at 32x32x32.  A wrong sign, face or centroid gives an
    O(1) error that does not shrink."""
    e1 = _shell_cell_b_error((16, 16, 16))
    e2 = _shell_cell_b_error((32, 32, 32))
    assert e1 < 2e-2
    assert e2 < e1 / 3.0


# ------------------------------------------------------------ Fluxes
def _euler_flux(rho, v, P, n):
    vn = v @ n
    E = P / (GAM - 1) + 0.5 * rho * (v @ v)
    return np.concatenate([[rho * vn], rho * v * vn + P * n, [(E + P) * vn]])


@pytest.mark.parametrize("boris", [False, True])
@pytest.mark.parametrize("bg", [False, True])
def test_face_fluxes_of_uniform_state(boris, bg):
    """Fluxes (stress.F90:204-429) composes the two kinetic fluxes, the HOGS
    and Boris mag-HOGS terms, the nuke term and the face area.  For a
    uniform state (L = R) every face carries area x (Euler flux) and
    area x (Maxwell stress of B + B0 minus that of B0), and nothing else:
    HOGS (-cHog VaD dW) and mag-HOGS (-cHogM VaD bbD dV/Ca^2) vanish, and the
    nuke, although |v| exceeds 1.5 Ca when Boris is on, adds Ca*dU = 0.
    Checked on a sheared grid with large cHog coefficients.  Catches a
    wrong face area or normal, a non-vanishing dissipative term, Bn not
    divided by the area, or a dropped B0 cross term."""
    N = (16, 8, 8)
    A = 0.05 * SHEAR
    xyz = affine_corners(N, A)
    rho, v, P, B = 1.3, np.array([0.31, -0.22, 0.17]), 0.8, np.array([0.41, -0.33, 0.52])
    B0 = np.array([-0.27, 0.45, 0.19]) if bg else np.zeros(3)
    over = dict(cHogH=3.0, cHogM=2.0, doBoris=int(boris), Ca=0.2, doBackground=int(bg))
    o, grid, p = oracle_on(N, xyz, **over)
    if bg:
        grid.update(uniform_background(N, grid, xyz, B0))
        o, grid, p = oracle_on(N, xyz, grid=grid, **over)
    st = uniform_state(N, grid, xyz, rho, v, P, B)
    o.set_state(st)
    o.init_run()
    o.advance(o.M.dt)
    Ni, Nj, Nk = (n + 2 * NG for n in N)
    gF = o.solver_array(0, (Ni, Nj, Nk, 5, 3))
    mF = o.solver_array(1, (Ni, Nj, Nk, 3, 3))
    Av = vector_areas(xyz)
    for d in range(3):
        a = Av[NG + 2, NG + 2, NG + 2, :, d]
        area = np.linalg.norm(a)
        nrm = a / area
        Bt = B + B0
        refg = area * _euler_flux(rho, v, P, nrm)
        refm = area * ((0.5 * Bt @ Bt * nrm - Bt * (Bt @ nrm))
                       - (0.5 * B0 @ B0 * nrm - B0 * (B0 @ nrm)))
        hi = [NG + n for n in N]
        sl = tuple(slice(NG, hi[q] + (1 if q == d else 0)) for q in range(3))
        got_g = gF[sl + (slice(None), d)]
        got_m = mF[sl + (slice(None), d)]
        assert np.max(np.abs(got_g - refg)) <= 1e-13 * np.max(np.abs(refg)), (d, "gas")
        mscale = area * (Bt @ Bt + B0 @ B0)
        assert np.max(np.abs(got_m - refm)) <= 1e-13 * mscale, (d, "mag")


def _nuke_box(Ca, vmag):
    """Hydrodynamic box with Boris on: uniform velocity of size vmag along
    a fixed direction, uniform pressure, density with a smooth variation
    along i."""
    N = (16, 8, 8)
    xyz = pb.cartesian_corners(*N)
    o, grid, p = oracle_on(N, xyz, doMHD=0, doBoris=1, Ca=Ca)
    vdir = np.array([0.6, -0.48, 0.64])
    v = vmag * vdir / np.linalg.norm(vdir)
    x = grid["xyzcc"][..., 0]
    rho = 1.0 + 0.3 * np.sin(2 * math.pi * x)
    st = pb.empty_state(*N)
    st["Gas"][...] = pb.prim2con(GAM, rho, np.full_like(x, v[0]), np.full_like(x, v[1]),
                                 np.full_like(x, v[2]), np.full_like(x, 0.7))
    o.set_state(st)
    o.init_run()
    o.M.dt = 1e-4
    o.advance(1e-4)
    Ni, Nj, Nk = (n + 2 * NG for n in N)
    gF = o.solver_array(0, (Ni, Nj, Nk, 5, 3))
    U = o.half_state()["Gas"]
    return gF, U, grid


def test_nuke_term_fires_at_threshold_and_subtracts_ca_du():
    """The Boris 'nuclear option' (stress.F90:399-417): where
    Vn = sqrt(|vL|^2 + |vR|^2)/sqrt(2.0) >= 1.5 Ca the gas flux loses
    Ca * (U(i) - U(i-1)) * A, the half-state conserved jump across the face;
    below the threshold nothing changes.  sqrt(2.0) is the single-precision
    value 1.41421353816986..., so with |vL| = |vR| = |v| the nuke fires
    exactly when |v| * 1.0000000168... >= 1.5 Ca.  Ca placed 1e-8 above
    |v|/1.5 (between the single- and double-precision thresholds) must
    fire; 3e-8 above must not.  Catches a missing face area or a wrong sign
    on the nuke term, L/R states in place of cell values, a threshold on
    |v| alone or at a different factor, and the double-precision sqrt(2)."""
    vmag = 0.9
    gF_off, U, grid = _nuke_box(1e6, vmag)
    face = grid["face"]
    for rel, fires in ((1e-8, True), (3e-8, False), (-1e-3, True), (1e-3, False)):
        Ca = vmag / 1.5 * (1.0 + rel)
        gF, U2, _ = _nuke_box(Ca, vmag)
        assert np.array_equal(U, U2)
        for d in range(3):
            hi = [NG + n for n in (16, 8, 8)]
            sl = tuple(slice(NG, hi[q] + (1 if q == d else 0)) for q in range(3))
            slm = tuple(slice(NG - (1 if q == d else 0), hi[q]) for q in range(3))
            dU = U[sl] - U[slm]
            diff = gF[sl + (slice(None), d)] - gF_off[sl + (slice(None), d)]
            scale = np.max(np.abs(gF_off[sl + (slice(None), d)]))
            ref = -Ca * dU * face[sl + (d,)][..., None] if fires else 0.0 * dU
            assert np.max(np.abs(diff - ref)) <= 1e-13 * scale, (rel, d)
            if fires and d == 0:
                assert np.max(np.abs(ref)) > 1e-3 * scale


# ------------------------------------------------------------ CellBoris
def _boris(U2, U0, B2, B0, dH, dM, dt, **over):
    M = po.make_model(pb.model_params(**over))
    u2, pu2 = vp(U2)
    _, pu0 = vp(U0)
    _, pb2 = vp(B2)
    _, pb0 = vp(B0)
    _, pdh = vp(dH)
    _, pdm = vp(dM)
    L.orc_cell_boris(C.byref(M), pu2, pu0, pb2, pb0, pdh, pdm, float(dt))
    return u2


def _boris_case(F_gas, F_mag, dRho=0.0, Ca=0.7):
    rho, v0, P = 1.4, np.array([0.2, -0.1, 0.3]), 0.5
    U0 = pb.prim2con(GAM, rho, *v0, P)
    dt = 0.01
    dH = np.concatenate([[dRho / dt], np.asarray(F_gas) / dt, [0.0]])
    dM = np.asarray(F_mag, dtype=float) / dt
    U2 = U0 + dt * dH                    # the Reynolds update before CellBoris
    B2 = np.array([0.9, 0.4, -0.6])
    B0 = np.array([0.7, 0.5, -0.4])
    out = _boris(U2, U0, B2, B0, dH, dM, dt, Ca=Ca)
    B1 = 0.5 * (B2 + B0)
    D2 = U2[0]
    alpha = (B1 @ B1) / (D2 * Ca * Ca)
    bh = B1 / np.linalg.norm(B1)
    W2 = np.empty(5)
    W2[0] = U2[0]
    W2[1:4] = U2[1:4] / U2[0]
    W2[4] = (GAM - 1) * (U2[4] - 0.5 * U2[0] * (W2[1:4] @ W2[1:4]))
    return out, U0, B1, bh, alpha, v0, rho, W2


def test_cell_boris_force_split():
    """CellBoris (mhdgroup.F90:336-441), alpha = |B_half|^2/(rho Ca^2):
      * a force along B_half passes unchanged (momentum += F);
      * a force across B_half is divided by 1 + alpha;
      * a density change at fixed momentum forces adds
        alpha/(1+alpha) dRho v0_perp;
      * the pressure is that of the Reynolds-updated state.
    Catches the parallel/perpendicular split applied the other way round,
    alpha with the wrong B (new or old instead of their mean) or without
    rho, and pCon on the parallel part."""
    # parallel force
    out, U0, B1, bh, alpha, v0, rho, W2 = _boris_case(0.03 * np.array([0.9, 0.4, -0.6]) / 1.0, np.zeros(3))
    F = 0.03 * np.array([0.9, 0.4, -0.6])
    Fpar = (F @ bh) * bh
    Fperp = F - Fpar
    ref = U0[1:4] + Fpar + Fperp / (1 + alpha)
    assert np.max(np.abs(out[1:4] - ref)) <= 1e-15
    assert abs(W2[4] - ((GAM - 1) * (out[4] - 0.5 * (out[1:4] @ out[1:4]) / out[0]))) <= 1e-14
    # a purely perpendicular force, gas and magnetic parts added
    perp = np.cross(bh, np.array([0.3, -0.2, 0.5]))
    out, U0, B1, bh, alpha, v0, rho, W2 = _boris_case(0.02 * perp, 0.01 * perp)
    ref = U0[1:4] + 0.03 * perp / (1 + alpha)
    assert np.max(np.abs(out[1:4] - ref)) <= 1e-15
    assert alpha > 0.5
    # density change only
    out, U0, B1, bh, alpha, v0, rho, W2 = _boris_case(np.zeros(3), np.zeros(3), dRho=0.05)
    vperp = v0 - (v0 @ bh) * bh
    ref = U0[1:4] + alpha / (1 + alpha) * 0.05 * vperp
    assert np.max(np.abs(out[1:4] - ref)) <= 1e-15


def test_cell_boris_small_alpha_and_null_field():
    """Small alpha (Ca large but alpha > TINY): pCon -> 1 and the update tends
    to dPGas + (dPMag across B_half) -- the magnetic force enters only
    through its perpendicular part ("only using dPMag in perp",
    mhdgroup.F90:372-376), its parallel numerical residue is dropped.  At
    alpha <= TINY (the null-field branch, 377-379) the update is the plain
    sum dPGas + dPMag.  Catches the magnetic parallel part kept in the alpha
    branch, or the null-field branch taken on the wrong side of TINY."""
    F_g = np.array([0.02, -0.01, 0.015])
    F_m = np.array([-0.005, 0.012, 0.004])
    for Ca in (1e3, 1e5):
        out, U0, B1, bh, alpha, v0, rho, W2 = _boris_case(F_g, F_m, dRho=0.01, Ca=Ca)
        assert 1e-12 < alpha < 1e-5
        Fm_perp = F_m - (F_m @ bh) * bh
        ref = U0[1:4] + F_g + Fm_perp
        assert np.max(np.abs(out[1:4] - ref)) <= 1e-15 + 2 * alpha, Ca
        assert np.max(np.abs(out[1:4] - (U0[1:4] + F_g + F_m))) > 1e-4
    out, U0, B1, bh, alpha, v0, rho, W2 = _boris_case(F_g, F_m, dRho=0.01, Ca=1e12)
    assert alpha <= 1e-12
    assert np.max(np.abs(out[1:4] - (U0[1:4] + F_g + F_m))) <= 1e-15


# ------------------------------------------------------------ Predictor
def test_predictor_exact_for_linear_time_dependence():
    """Predictor/PredictCell/ExtrapAB2 (mhdgroup.F90:444-615, 709-717): with
    oState at t - odt and State at t whose primitives and face fluxes are
    linear in time, the half state at t + dt/2 is exact (AB2 extrapolates
    lines exactly).  odt differs from dt.  Catches a slope with the wrong
    sign or time base, dt in place of dt/2, extrapolation of conserved
    instead of primitive variables (kinetic energy is not linear), and a
    flux extrapolation that is missing."""
    N = (16, 8, 8)
    xyz = pb.cartesian_corners(*N)
    o, grid, p = oracle_on(N, xyz)
    x = grid["xyzcc"]
    rng = np.random.default_rng(5)
    W0 = np.stack([1.0 + 0.2 * np.sin(2 * math.pi * x[..., 0]), 0.1 * np.cos(2 * math.pi * x[..., 1]),
                   -0.05 + 0 * x[..., 0], 0.08 * np.sin(2 * math.pi * x[..., 2]),
                   0.7 + 0.1 * np.cos(2 * math.pi * x[..., 0])], axis=-1)
    Wd = np.stack([0.3 + 0 * x[..., 0], -0.2 + 0.1 * x[..., 1], 0.4 + 0 * x[..., 0],
                   0.1 * x[..., 2], -0.5 + 0 * x[..., 0]], axis=-1)
    mf0 = pb.vecpot_fluxes(xyz, lambda a, b, c: (0.05 * np.sin(2 * math.pi * b), 0.04 * np.cos(2 * math.pi * c),
                                                  0.03 * np.sin(2 * math.pi * a)))
    mfd = rng.standard_normal(mf0.shape) * 0.01
    t, odt, dt = 2.0, 0.013, 0.021

    def state(tt):
        W = W0 + (tt - t) * Wd
        st = pb.empty_state(*N)
        st["Gas"][...] = pb.prim2con(GAM, W[..., 0], W[..., 1], W[..., 2], W[..., 3], W[..., 4])
        st["magFlux"][...] = mf0 + (tt - t) * mfd
        st["time"] = tt
        return st
    S, O = state(t), state(t - odt)
    o.set_state(S)
    o.ost.update({k: np.asfortranarray(O[k]) for k in ("Gas", "magFlux", "Bxyz", "Efld")})
    o.O = o._state(o.ost)
    o.O.time = t - odt
    o.S.time = t
    o.M.dt = dt
    o.advance(dt)
    H = o.half_state()
    ref = state(t + 0.5 * dt)
    a = act(N)
    assert np.max(np.abs(H["Gas"][a] - ref["Gas"][a])) <= 1e-13 * np.max(np.abs(ref["Gas"][a]))
    assert np.max(np.abs(H["magFlux"] - ref["magFlux"])) <= 1e-14 * np.max(np.abs(ref["magFlux"]))


# ------------------------------------------------------------ AddB0
GX, GWT = np.polynomial.legendre.leggauss(20)


def _dipole_B(M0, p):
    r = np.linalg.norm(p, axis=-1, keepdims=True)
    m = np.array([0.0, 0.0, M0])
    return 3 * (p @ m)[..., None] * p / r ** 5 - m / r ** 3


def _dipole_A(M0, p):
    r = np.linalg.norm(p, axis=-1, keepdims=True)
    return np.cross(np.array([0.0, 0.0, M0]), p) / r ** 3


def _edge_avg(f, e0, e1):
    """Average of f over the straight segment e0 -> e1 (20-point Gauss)."""
    acc = 0.0
    for x, w in zip(GX, GWT):
        acc = acc + 0.5 * w * f(0.5 * (e0 + e1) + 0.5 * x * (e1 - e0))
    return acc


def _loop_flux(M0, c):
    """Loop integral of A around the corner loop c[0]..c[3] (Stokes)."""
    s = 0.0
    for q in range(4):
        e0, e1 = c[q], c[(q + 1) % 4]
        s = s + np.sum(_edge_avg(lambda p: _dipole_A(M0, p), e0, e1) * (e1 - e0), axis=-1)
    return s


def _face_corners(xyz, d):
    """Corner loops of every face of direction d, ordered counter-clockwise
    about +d; arrays over the cell index box (Ni, Nj, Nk)."""
    p = xyz
    c = lambda di, dj, dk: p[di:di + p.shape[0] - 1, dj:dj + p.shape[1] - 1, dk:dk + p.shape[2] - 1]
    if d == 0:
        return [c(0, 0, 0), c(0, 1, 0), c(0, 1, 1), c(0, 0, 1)]
    if d == 1:
        return [c(0, 0, 0), c(0, 0, 1), c(1, 0, 1), c(1, 0, 0)]
    return [c(0, 0, 0), c(1, 0, 0), c(1, 1, 0), c(0, 1, 0)]


def _bilinear_face_quad(f, c, n=20):
    """Integral of f and of 1 over the bilinear face c (n x n Gauss)."""
    X, Wt = np.polynomial.legendre.leggauss(n)
    I = 0.0
    Ar = 0.0
    for a, wa in zip(X, Wt):
        for b, wb in zip(X, Wt):
            u, v = 0.5 * (a + 1), 0.5 * (b + 1)
            p = (1 - u) * (1 - v) * c[0] + u * (1 - v) * c[1] + u * v * c[2] + (1 - u) * v * c[3]
            xu = (1 - v) * (c[1] - c[0]) + v * (c[2] - c[3])
            xv = (1 - u) * (c[3] - c[0]) + u * (c[2] - c[1])
            jac = np.linalg.norm(np.cross(xu, xv), axis=-1, keepdims=True)
            w = 0.25 * wa * wb * jac
            I = I + w * f(p)
            Ar = Ar + w
    return I, Ar[..., 0]


@pytest.fixture(scope="module")
def dipole_shell():
    N = (16, 16, 16)
    xyz = pb.shell_corners(*N)
    g = po.corners2grid(*N, xyz)
    g["xyz"] = xyz
    po.add_dipole_b0(*N, g, pb.SHELL_M0)
    return N, xyz, g


def test_dipole_face_flux_is_loop_integral_of_vector_potential(dipole_shell):
    """AddB0 bFlux0 (background.F90:20-167, GaussianFaceFlux): the flux of
    the dipole B = curl(M x r / r^3) through each bilinear face equals the
    loop integral of A around its four straight edges (Stokes), computed
    here with 20-point Gauss edges.  Catches a flux with the wrong
    orientation, the wrong corner order, the wrong moment sign or a
    quadrature that is not over the face's own surface."""
    N, xyz, g = dipole_shell
    refs = [_loop_flux(pb.SHELL_M0, _face_corners(xyz, d)) for d in range(3)]
    # the phi faces lie in meridional planes, where the axisymmetric dipole
    # has no normal component: their flux is zero up to round-off, so every
    # direction is held to the scale of the largest face flux
    scale = max(np.max(np.abs(r)) for r in refs)
    for d in range(3):
        got = g["bFlux0"][..., d]
        rel = np.max(np.abs(got - refs[d])) / scale
        assert rel <= 1e-11, (d, rel)
    assert np.max(np.abs(refs[0])) > 0.1 * scale and np.max(np.abs(refs[1])) > 0.1 * scale


def test_dipole_face_and_edge_averages(dipole_shell):
    """fcB0 is the area average of B0 over each face (fInt/Face), edgB0 the
    edge average of B0 projected on the edge triad's TAN1/TAN2, B0 the cell
    average.  Checked against independent 20-point face and edge
    quadratures and a Gauss product rule over each trilinear cell.
    Catches averages taken over the wrong face or edge, projections on the
    wrong triad vectors, or a missing Jacobian."""
    N, xyz, g = dipole_shell
    f = lambda p: _dipole_B(pb.SHELL_M0, p)
    for d in range(3):
        I, Ar = _bilinear_face_quad(f, _face_corners(xyz, d))
        ref = I / g["face"][:-1, :-1, :-1, d][..., None]
        got = g["fcB0"][..., :, d]
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) <= 1e-11, d
        assert np.max(np.abs(Ar - g["face"][:-1, :-1, :-1, d])) / np.max(Ar) <= 1e-12
    p = xyz
    # some edge components vanish by the dipole's symmetry (the phi edges'
    # field lies in the meridional plane): all are held to the largest
    refs = {}
    for e in range(3):
        sl1 = [slice(0, s - 1) for s in p.shape[:3]]
        sl1[e] = slice(1, p.shape[e])
        e0 = p[:-1, :-1, :-1]
        e1 = p[tuple(sl1)]
        avg = _edge_avg(f, e0, e1)
        Te = g["Te"][:-1, :-1, :-1, :, e]
        for q, (lo, hi) in enumerate(((3, 6), (6, 9))):
            refs[e, q] = np.sum(avg * Te[..., lo:hi], axis=-1)[edge_box(N, e)]
    scale = max(np.max(np.abs(r)) for r in refs.values())
    for (e, q), ref in refs.items():
        got = g["edgB0"][..., q, e][edge_box(N, e)]
        assert np.max(np.abs(got - ref)) / scale <= 1e-12, (e, q)
    # cell averages over the trilinear cells of the active range
    X, Wt = np.polynomial.legendre.leggauss(14)
    a = act(N)
    c = [[[p[di:di + p.shape[0] - 1, dj:dj + p.shape[1] - 1, dk:dk + p.shape[2] - 1][a]
           for dk in (0, 1)] for dj in (0, 1)] for di in (0, 1)]
    acc = 0.0
    vol = 0.0
    for u, wu in zip(X, Wt):
        for v, wv in zip(X, Wt):
            for w, ww in zip(X, Wt):
                s, t, r = 0.5 * (u + 1), 0.5 * (v + 1), 0.5 * (w + 1)
                wts = {(0, 0, 0): (1 - s) * (1 - t) * (1 - r), (1, 0, 0): s * (1 - t) * (1 - r),
                       (0, 1, 0): (1 - s) * t * (1 - r), (1, 1, 0): s * t * (1 - r),
                       (0, 0, 1): (1 - s) * (1 - t) * r, (1, 0, 1): s * (1 - t) * r,
                       (0, 1, 1): (1 - s) * t * r, (1, 1, 1): s * t * r}
                pt = sum(wts[k] * c[k[0]][k[1]][k[2]] for k in wts)
                xs = ((1 - t) * (1 - r) * (c[1][0][0] - c[0][0][0]) + t * (1 - r) * (c[1][1][0] - c[0][1][0])
                      + (1 - t) * r * (c[1][0][1] - c[0][0][1]) + t * r * (c[1][1][1] - c[0][1][1]))
                xt = ((1 - s) * (1 - r) * (c[0][1][0] - c[0][0][0]) + s * (1 - r) * (c[1][1][0] - c[1][0][0])
                      + (1 - s) * r * (c[0][1][1] - c[0][0][1]) + s * r * (c[1][1][1] - c[1][0][1]))
                xr = ((1 - s) * (1 - t) * (c[0][0][1] - c[0][0][0]) + s * (1 - t) * (c[1][0][1] - c[1][0][0])
                      + (1 - s) * t * (c[0][1][1] - c[0][1][0]) + s * t * (c[1][1][1] - c[1][1][0]))
                jac = np.abs(np.sum(xs * np.cross(xt, xr), axis=-1))[..., None]
                wq = 0.125 * wu * wv * ww * jac
                acc = acc + wq * f(pt)
                vol = vol + wq
    ref = acc / vol
    got = g["B0"][a]
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) <= 1e-10


# ------------------------------------------------------------ inflow series
def _wind_model(wind):
    return po.make_model(pb.shell_params(), wind)


def _interp(M, t):
    out, po_ = vp(np.zeros(8))
    L.orc_wind_interp(C.byref(M), float(t), po_)
    return out


def test_interp_wind_discrete_walls():
    """InterpWind/InterpWgts (apps/wind.F90:766-868) with discrete walls
    (doWindInterp = F): for tMin < t < tMax the sample is
    i0 = floor((t - tMin)/dtW) (0-based), the last interval maps to the
    second-to-last sample, t >= tMax gives the last sample, t <= tMin the
    first sample with B set to zero; Bx is replaced by Bx0 + ByC By + BzC
    Bz.  dtW is the smallest sample spacing, so on an uneven series the
    index follows that formula and not the containing interval.  Catches
    off-by-one sample choice, interpolation where walls are meant, and a
    dropped Bx expansion or null wind."""
    n = 6
    tW = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0])
    Q = np.zeros((n, 5), order="F")
    Q[:, 0] = 10.0 + np.arange(n)
    Q[:, 1] = -0.4 - 0.01 * np.arange(n)
    Q[:, 2] = 0.02 * np.arange(n)
    Q[:, 3] = -0.03 * np.arange(n)
    Q[:, 4] = 1e-3 * (1 + np.arange(n))
    B = np.zeros((n, 3), order="F")
    B[:, 0] = 99.0
    B[:, 1] = 0.1 * np.arange(n)
    B[:, 2] = -0.2 * np.arange(n)
    wind = {"tW": tW, "Q": Q, "B": B, "Bx0": 0.01, "ByC": 0.3, "BzC": -0.2, "xyzW": (0, 0, 0)}
    M = _wind_model(wind)
    for t, i0 in ((0.5, 0), (1.0, 1), (1.999, 1), (3.2, 3), (4.5, 4), (4.999999, 4), (5.0, 5), (7.0, 5)):
        out = _interp(M, t)
        assert out[0] == Q[i0, 0] and out[1] == Q[i0, 4], (t, i0)
        assert np.array_equal(out[2:5], Q[i0, 1:4])
        assert out[6] == B[i0, 1] and out[7] == B[i0, 2]
        assert out[5] == 0.01 + 0.3 * B[i0, 1] - 0.2 * B[i0, 2]
    out = _interp(M, -0.5)
    assert out[0] == Q[0, 0] and np.all(out[5:8] == 0.0)
    # uneven series: dtW = 0.5 (smallest spacing), index floor((t-0)/0.5)
    tW2 = np.array([0.0, 0.5, 2.0, 4.0, 6.0, 8.0])
    M2 = _wind_model(dict(wind, tW=tW2))
    for t, i0 in ((0.6, 1), (1.2, 2), (1.6, 3), (2.2, 4), (7.0, 4)):
        assert _interp(M2, t)[0] == Q[i0, 0], (t, i0)


def test_wind_front_lag_and_weight():
    """RefreshWind (apps/wind.F90:114-155): the front velocity is
    vFr = -|V0| (n.x) n with n = (1, -ByC, -BzC)/sqrt(1 + ByC^2 + BzC^2),
    the normal of the planes that contain every IMF Bx0=0 + ByC By + BzC Bz
    direction.  GetWindAt (157-175): a point at xyzW + s vFr is served the
    wind of time t - s.  wgtWind (180-203): RampDown of the angle between
    -vHat and the radial direction, 1 up to 90 degrees, exp(1 - 1/(1 -
    ((th-90)/90)^2)) beyond, 0 at 180.  Catches a front that is not
    normal to the IMF plane, a lag with the wrong sign or scale, and a
    ramp on the wrong angle."""
    n = 200
    wind = pb.varying_wind(n)
    wind["xyzW"] = (1.5, -0.5, 0.25)
    M = _wind_model(wind)
    t = 3.7
    vF, pv = vp(np.zeros(3))
    L.orc_wind_front(C.byref(M), t, pv)
    V0 = _interp(M, t)[2:5]
    nn = np.array([1.0, -wind["ByC"], -wind["BzC"]])
    nn /= np.linalg.norm(nn)
    ref = -np.linalg.norm(V0) * nn[0] * nn
    assert np.max(np.abs(vF - ref)) <= 1e-15
    for By, Bz in ((0.3, -0.1), (-0.2, 0.5)):
        assert abs(vF @ np.array([wind["ByC"] * By + wind["BzC"] * Bz, By, Bz])) <= 1e-15
    for s in (-40.0, -3.3, 0.0, 12.5):
        x, px = vp(np.array(wind["xyzW"]) + s * vF)
        out, po_ = vp(np.zeros(8))
        L.orc_wind_at(C.byref(M), pv, px, t, po_)
        ref = _interp(M, t - s)
        tl = t - s
        # a point exactly on a sample wall may round either way
        walls = np.min(np.abs(wind["tW"] - tl))
        if walls > 1e-9:
            assert np.array_equal(out, ref), s
    vhat = V0 / np.linalg.norm(V0)
    perp = np.cross(vhat, np.array([0.0, 0.0, 1.0]))
    perp /= np.linalg.norm(perp)
    for ang in (0.0, 45.0, 90.0, 120.0, 150.0, 180.0):
        th = math.radians(ang)
        nhat = -vhat * math.cos(th) + perp * math.sin(th)
        x, px = vp(20.0 * nhat)
        w = L.orc_wind_wgt(C.byref(M), pv, px, t)
        # the weight uses the wind at the lagged time, whose direction
        # differs slightly; recompute the angle from that wind
        out, po_ = vp(np.zeros(8))
        L.orc_wind_at(C.byref(M), pv, px, t, po_)
        vh = out[2:5] / np.linalg.norm(out[2:5])
        dth = math.degrees(math.acos(np.clip(-vh @ nhat, -1, 1)))
        if dth <= 90.0:
            ref = 1.0
        elif dth >= 180.0:
            ref = 0.0
        else:
            sc = (dth - 90.0) / 90.0
            ref = math.exp(1.0 - 1.0 / (1.0 - sc * sc))
        assert abs(w - ref) <= 1e-12, ang
    assert L.orc_ramp_down(90.0, 90.0, 90.0) == 1.0
    assert L.orc_ramp_down(180.0, 90.0, 90.0) == 0.0
    assert abs(L.orc_ramp_down(135.0, 90.0, 90.0) - math.exp(-1.0 / 3.0)) <= 1e-16
