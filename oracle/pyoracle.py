"""ctypes front end of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product (mage_amd) never does.  Arrays are numpy arrays in
the reference's Fortran layout, as produced by mage_amd.problems.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the same source built for speed (-O3, AVX2/FMA): the CPU baseline only
LIB_FAST = os.path.join(HERE, "liboracle_fast.so")

BC = {"periodic": 0, "zerograd": 1, "sphin": 2, "sphout": 3, "wind": 4}
DP = C.POINTER(C.c_double)


class Dims(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "Nip", "Njp", "Nkp", "nG", "Ni", "Nj", "Nk", "is_", "ie", "js", "je", "ks", "ke",
        "isg", "ieg", "jsg", "jeg", "ksg", "keg", "isDT", "ieDT", "jsDT", "jeDT", "ksDT", "keDT")] + [
        ("nc", C.c_long), ("nf", C.c_long)]


class Model(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "CFL", "Vd0", "Ca", "cHogH", "cHogM", "pdmb", "dFloor", "pFloor",
        "t", "tFin", "dt", "dt0", "limDT0")] + [
        ("ts", C.c_int), ("doMHD", C.c_int), ("doBoris", C.c_int), ("doBackground", C.c_int),
        ("doLFMLim", C.c_int), ("rmeth", C.c_int), ("bc", C.c_int * 6),
        ("nWind", C.c_int), ("tW", DP), ("Qw", DP), ("Bw", DP),
        ("Bx0", C.c_double), ("ByC", C.c_double), ("BzC", C.c_double), ("xyzW", C.c_double * 3)]


GRID_FIELDS = ("xyz", "xyzcc", "xfc", "di", "dj", "dk", "volume", "Tf", "Te", "Teb", "face",
               "edge", "fcB0", "edgB0", "B0", "bFlux0", "dpB0")


class Grid(C.Structure):
    _fields_ = [(n, DP) for n in GRID_FIELDS]


class State(C.Structure):
    _fields_ = [("Gas", DP), ("magFlux", DP), ("Efld", DP), ("Bxyz", DP), ("time", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(fast=False):
    """The strict parity build, or with fast=True the speed build."""
    path = LIB_FAST if fast else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_dims_init.argtypes = [C.POINTER(Dims), C.c_int, C.c_int, C.c_int]
        L.orc_cell_c2p.argtypes = [C.POINTER(Model), DP, DP]
        L.orc_cell_p2c.argtypes = [C.POINTER(Model), DP, DP]
        L.orc_pdm.argtypes = [C.c_double] * 5
        L.orc_pdm.restype = C.c_double
        L.orc_pdm_lr.argtypes = [C.c_double, DP, C.c_double, DP, DP]
        L.orc_recon_lr.argtypes = [C.POINTER(Model), DP, DP, DP, DP, DP]
        L.orc_gas_kin_flux.argtypes = [C.POINTER(Model), C.c_int, DP, DP, DP, DP]
        L.orc_mag_kin_flux.argtypes = [C.POINTER(Model), C.c_int, DP, DP, C.c_double, DP, DP,
                                       C.c_double, DP, DP, DP]
        L.orc_cell_dt.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                                  C.POINTER(State), C.c_int, C.c_int, C.c_int]
        L.orc_cell_dt.restype = C.c_double
        L.orc_corners2grid.argtypes = [C.POINTER(Dims), C.POINTER(Grid)]
        L.orc_add_dipole_b0.argtypes = [C.POINTER(Dims), C.POINTER(Grid), C.c_double]
        L.orc_dipole.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, DP]
        L.orc_solver_new.argtypes = [C.POINTER(Dims)]
        L.orc_solver_new.restype = vp
        L.orc_solver_free.argtypes = [vp]
        L.orc_solver_array.argtypes = [vp, C.c_int]
        L.orc_solver_array.restype = DP
        L.orc_bflux2fld.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid), DP, DP]
        L.orc_calc_dt.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                                  C.POINTER(State), DP]
        L.orc_enforce_bcs.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                                      C.POINTER(State)]
        L.orc_advance_mhd.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                                      C.POINTER(State), C.POINTER(State), vp, C.c_double]
        L.orc_step.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                               C.POINTER(State), C.POINTER(State), vp]
        L.orc_init_run.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                                   C.POINTER(State), C.POINTER(State)]
        L.orc_divb.argtypes = [C.POINTER(Dims), C.POINTER(Model), C.POINTER(Grid),
                               C.POINTER(State), DP]
        L.orc_divb.restype = C.c_double
        L.orc_cell_boris.argtypes = [C.POINTER(Model), DP, DP, DP, DP, DP, DP, C.c_double]
        L.orc_wind_interp.argtypes = [C.POINTER(Model), C.c_double, DP]
        L.orc_wind_front.argtypes = [C.POINTER(Model), C.c_double, DP]
        L.orc_wind_at.argtypes = [C.POINTER(Model), DP, DP, C.c_double, DP]
        L.orc_wind_wgt.argtypes = [C.POINTER(Model), DP, DP, C.c_double]
        L.orc_wind_wgt.restype = C.c_double
        L.orc_ramp_down.argtypes = [C.c_double] * 3
        L.orc_ramp_down.restype = C.c_double
        L.orc_nuke_count.argtypes = [C.c_int]
        L.orc_nuke_count.restype = C.c_long
        _libs[path] = L
    return _libs[path]


def nuke_count(reset=True):
    """Faces on which the oracle's nuke term fired since the last reset."""
    return int(lib().orc_nuke_count(1 if reset else 0))


def ptr(a):
    if a is None:
        return DP()
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need Fortran-contiguous float64"
    return a.ctypes.data_as(DP)


def make_model(p, wind=None):
    """Model struct from a parameter dict (keys as mage_amd.problems.model_params)."""
    m = Model()
    for k in ("gamma", "CFL", "Vd0", "Ca", "cHogH", "cHogM", "pdmb", "dFloor", "pFloor",
              "tFin", "dt0", "limDT0"):
        setattr(m, k, float(p[k]))
    for k in ("doMHD", "doBoris", "doBackground", "doLFMLim", "rmeth"):
        setattr(m, k, int(p[k]))
    for s in range(6):
        m.bc[s] = BC[p["bc"][s]] if isinstance(p["bc"][s], str) else int(p["bc"][s])
    m.t = 0.0
    m.dt = 0.0
    m.ts = 0
    m._keep = []
    if wind is not None:
        tW = np.ascontiguousarray(wind["tW"], dtype=np.float64)
        Q = np.asfortranarray(wind["Q"], dtype=np.float64)
        B = np.asfortranarray(wind["B"], dtype=np.float64)
        m._keep = [tW, Q, B]
        m.nWind = len(tW)
        m.tW, m.Qw, m.Bw = ptr(tW), ptr(Q), ptr(B)
        m.Bx0, m.ByC, m.BzC = wind.get("Bx0", 0.0), wind.get("ByC", 0.0), wind.get("BzC", 0.0)
        for q, v in enumerate(wind.get("xyzW", (0.0, 0.0, 0.0))):
            m.xyzW[q] = v
    return m


class Oracle:
    """One oracle run: dims, model, grid arrays and the State/oState pair."""

    def __init__(self, Nip, Njp, Nkp, params, grid, wind=None, fast=False):
        self.L = lib(fast)
        self.D = Dims()
        self.L.orc_dims_init(C.byref(self.D), Nip, Njp, Nkp)
        self.M = make_model(params, wind)
        self.grid = grid
        self.G = Grid()
        for k in GRID_FIELDS:
            if k in grid and grid[k] is not None:
                setattr(self.G, k, ptr(grid[k]))
        self.solver = self.L.orc_solver_new(C.byref(self.D))
        self.S = None
        self.O = None

    def __del__(self):
        try:
            if self.solver:
                self.L.orc_solver_free(self.solver)
        except Exception:
            pass

    @staticmethod
    def _state(arrs):
        s = State()
        s.Gas, s.magFlux, s.Efld, s.Bxyz = (ptr(arrs[k]) for k in ("Gas", "magFlux", "Efld", "Bxyz"))
        s.time = float(arrs.get("time", 0.0))
        return s

    def set_state(self, st):
        """st: dict of Fortran arrays Gas, magFlux, Bxyz, Efld (+time)."""
        self.st = {k: np.asfortranarray(np.array(st[k], dtype=np.float64)) for k in ("Gas", "magFlux", "Bxyz", "Efld")}
        self.ost = {k: v.copy(order="F") for k, v in self.st.items()}
        self.S = self._state(self.st)
        self.S.time = float(st.get("time", 0.0))
        self.O = self._state(self.ost)
        self.O.time = self.S.time

    def init_run(self):
        return self.L.orc_init_run(C.byref(self.D), C.byref(self.M), C.byref(self.G),
                                   C.byref(self.S), C.byref(self.O))

    def step(self):
        return self.L.orc_step(C.byref(self.D), C.byref(self.M), C.byref(self.G),
                               C.byref(self.S), C.byref(self.O), self.solver)

    def advance(self, dt):
        self.L.orc_advance_mhd(C.byref(self.D), C.byref(self.M), C.byref(self.G),
                               C.byref(self.S), C.byref(self.O), self.solver, float(dt))

    def calc_dt(self):
        dt = C.c_double()
        rc = self.L.orc_calc_dt(C.byref(self.D), C.byref(self.M), C.byref(self.G),
                                C.byref(self.S), C.byref(dt))
        return rc, dt.value

    def enforce_bcs(self):
        self.L.orc_enforce_bcs(C.byref(self.D), C.byref(self.M), C.byref(self.G), C.byref(self.S))

    def recompute_cell_b(self):
        """Cell B of State and oState from their own face fluxes over every
        cell, ghosts included (bFlux2Fld after the boundary pass)."""
        for st in (self.st, self.ost):
            self.L.orc_bflux2fld(C.byref(self.D), C.byref(self.M), C.byref(self.G), ptr(st["magFlux"]),
                                 ptr(st["Bxyz"]))

    def divb(self):
        return self.L.orc_divb(C.byref(self.D), C.byref(self.M), C.byref(self.G), C.byref(self.S), DP())

    def solver_array(self, which, shape):
        p = self.L.orc_solver_array(self.solver, which)
        n = int(np.prod(shape))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(shape, order="F").copy(order="F")

    def half_state(self):
        Ni, Nj, Nk = self.D.Ni, self.D.Nj, self.D.Nk
        return {
            "Gas": self.solver_array(5, (Ni, Nj, Nk, 5)),
            "magFlux": self.solver_array(6, (Ni + 1, Nj + 1, Nk + 1, 3)),
            "Bxyz": self.solver_array(7, (Ni, Nj, Nk, 3)),
        }

    @property
    def time(self):
        return self.M.t, self.M.dt, self.M.ts


def corners2grid(Nip, Njp, Nkp, xyz):
    """Oracle metric pre-compute from corners xyz(Ni+1,Nj+1,Nk+1,3)."""
    L = lib()
    D = Dims()
    L.orc_dims_init(C.byref(D), Nip, Njp, Nkp)
    Ni, Nj, Nk = D.Ni, D.Nj, D.Nk
    f = (Ni + 1, Nj + 1, Nk + 1)
    c = (Ni, Nj, Nk)
    g = {
        "xyz": np.asfortranarray(xyz, dtype=np.float64),
        "xyzcc": np.zeros(c + (3,), order="F"), "xfc": np.zeros(f + (3, 3), order="F"),
        "di": np.zeros(c, order="F"), "dj": np.zeros(c, order="F"), "dk": np.zeros(c, order="F"),
        "volume": np.zeros(c, order="F"), "Tf": np.zeros(f + (9, 3), order="F"),
        "Te": np.zeros(f + (9, 3), order="F"), "Teb": np.zeros(f + (4, 3), order="F"),
        "face": np.zeros(f + (3,), order="F"), "edge": np.zeros(f + (3,), order="F"),
    }
    G = Grid()
    for k, v in g.items():
        setattr(G, k, ptr(v))
    L.orc_corners2grid(C.byref(D), C.byref(G))
    return g


def add_dipole_b0(Nip, Njp, Nkp, grid, M0):
    L = lib()
    D = Dims()
    L.orc_dims_init(C.byref(D), Nip, Njp, Nkp)
    c = (D.Ni, D.Nj, D.Nk)
    for k, sh in (("fcB0", (3, 3)), ("edgB0", (2, 3)), ("B0", (3,)), ("bFlux0", (3,)), ("dpB0", (3,))):
        grid[k] = np.zeros(c + sh, order="F")
    G = Grid()
    for k in GRID_FIELDS:
        if k in grid:
            setattr(G, k, ptr(grid[k]))
    L.orc_add_dipole_b0(C.byref(D), C.byref(G), float(M0))
    return grid
