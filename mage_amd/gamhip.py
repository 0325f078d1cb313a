"""Python binding of libgamhip (include/gamhip.h) over ctypes.

Mirrors the reference's operator interface for the GAMERA advance step
(rgorby/MAGE src/gamera): a ``GamHip`` context holds one Grid/State/Solver
set on the GPU and exposes AdvanceMHD, CalcDT, EnforceBCs, bFlux2Fld and
stepGamera.  Host arrays are numpy arrays in the reference's Fortran
layout: cell arrays (Ni,Nj,Nk[,comp..]) and face arrays
(Ni+1,Nj+1,Nk+1[,comp..]), order='F', with Ni = Nip + 2*nG.

There is no CPU fallback: if the library or a GPU is missing, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libgamhip.so")

GAM_OK = 0
GAM_STATE, GAM_OSTATE, GAM_HFSTATE = 0, 1, 2
BC = {"periodic": 0, "zerograd": 1, "sphin": 2, "sphout": 3, "wind": 4}
RECON = {"8CENT": 0, "7UP": 1}
ERRORS = {-1: "GAM_E_ARG", -2: "GAM_E_HIP", -3: "GAM_E_STATE", -4: "GAM_E_DT_BAD",
          -5: "GAM_E_DT_DISASTER", -6: "GAM_E_NOMEM"}

# Every symbol include/gamhip.h declares
EXPORTS = (
    "gam_last_error", "gam_create", "gam_destroy", "gam_get_dims", "gam_set_grid", "gam_grid_from_corners",
    "gam_get_grid", "gam_set_background", "gam_add_dipole_b0", "gam_get_background",
    "gam_set_wind", "gam_put_state", "gam_get_state", "gam_set_time", "gam_get_time", "gam_set_dt0", "gam_get_dt0",
    "gam_init_run", "gam_advance_mhd", "gam_calc_dt", "gam_enforce_bcs", "gam_bflux2fld",
    "gam_step", "gam_run", "gam_advance_model", "gam_sync", "gam_divb", "gam_time_steps", "gam_set_profiling",
    "gam_kernel_times", "gam_kernel_group_name", "gam_get_solver",
    "gam_autotune", "gam_get_plan", "gam_set_plan", "gam_tune_slot_name", "gam_tune_candidates",
    "gam_ens_create", "gam_ens_run", "gam_ens_time_steps", "gam_ens_launches_per_step", "gam_ens_destroy",
    "gam_divb_noalloc",
)

DP = C.POINTER(C.c_double)
TUNE_SLOTS = 8


class GamError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "GAM_E?"), code, msg))
        self.code = code


class Dims(C.Structure):
    _fields_ = [("Nip", C.c_int), ("Njp", C.c_int), ("Nkp", C.c_int), ("nG", C.c_int)]


class Model(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "CFL", "Vd0", "Ca", "cHogH", "cHogM", "pdmb", "dFloor", "pFloor",
        "tFin", "dt0", "limDT0")] + [(n, C.c_int) for n in (
            "doMHD", "doBoris", "doBackground", "doLFMLim", "rmeth")] + [("bc", C.c_int * 6)]


GRID_FIELDS = ("xyzcc", "xfc", "di", "dj", "dk", "volume", "Tf", "Te", "Teb", "face", "edge")
BG_FIELDS = ("fcB0", "edgB0", "B0", "bFlux0", "dpB0")


class GridArrays(C.Structure):
    _fields_ = [(n, DP) for n in GRID_FIELDS]


class BgArrays(C.Structure):
    _fields_ = [(n, DP) for n in BG_FIELDS]


_lib = None


def load():
    """Load libgamhip.so (built by mage_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise OSError("libgamhip.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIBPATH)
    vp = C.c_void_p
    L.gam_last_error.restype = C.c_char_p
    L.gam_kernel_group_name.restype = C.c_char_p
    L.gam_kernel_group_name.argtypes = [C.c_int]
    L.gam_create.argtypes = [C.POINTER(vp), C.POINTER(Dims), C.POINTER(Model)]
    L.gam_destroy.argtypes = [vp]
    L.gam_set_grid.argtypes = [vp, C.POINTER(GridArrays)]
    L.gam_get_grid.argtypes = [vp, C.POINTER(GridArrays)]
    L.gam_grid_from_corners.argtypes = [vp, DP]
    L.gam_set_background.argtypes = [vp, C.POINTER(BgArrays)]
    L.gam_get_background.argtypes = [vp, C.POINTER(BgArrays)]
    L.gam_add_dipole_b0.argtypes = [vp, C.c_double]
    L.gam_set_wind.argtypes = [vp, C.c_int, DP, DP, DP, C.c_double, C.c_double, C.c_double, DP]
    L.gam_put_state.argtypes = [vp, C.c_int, DP, DP, DP, DP, C.c_double]
    L.gam_get_state.argtypes = [vp, C.c_int, DP, DP, DP, DP, DP]
    L.gam_set_time.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.gam_get_time.argtypes = [vp, DP, DP, C.POINTER(C.c_int)]
    L.gam_get_dims.argtypes = [vp, C.POINTER(Dims)]
    L.gam_set_dt0.argtypes = [vp, C.c_double]
    L.gam_get_dt0.argtypes = [vp, DP]
    L.gam_init_run.argtypes = [vp, DP]
    L.gam_advance_mhd.argtypes = [vp, C.c_double]
    L.gam_calc_dt.argtypes = [vp, DP]
    L.gam_enforce_bcs.argtypes = [vp]
    L.gam_bflux2fld.argtypes = [vp, C.c_int]
    L.gam_step.argtypes = [vp, DP]
    L.gam_run.argtypes = [vp, C.c_int]
    L.gam_advance_model.argtypes = [vp, C.c_double, C.POINTER(C.c_int)]
    L.gam_sync.argtypes = [vp]
    L.gam_divb.argtypes = [vp, DP]
    L.gam_time_steps.argtypes = [vp, C.c_int, DP]
    L.gam_set_profiling.argtypes = [vp, C.c_int]
    L.gam_kernel_times.argtypes = [vp, DP, C.POINTER(C.c_int)]
    L.gam_get_solver.argtypes = [vp, C.c_int, DP]
    L.gam_autotune.argtypes = [vp, C.c_int]
    L.gam_get_plan.argtypes = [vp, C.POINTER(C.c_int)]
    L.gam_set_plan.argtypes = [vp, C.POINTER(C.c_int)]
    L.gam_tune_slot_name.restype = C.c_char_p
    L.gam_tune_slot_name.argtypes = [C.c_int]
    L.gam_tune_candidates.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.gam_ens_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.gam_ens_run.argtypes = [vp, C.c_int]
    L.gam_ens_time_steps.argtypes = [vp, C.c_int, DP]
    L.gam_ens_launches_per_step.argtypes = [vp, C.POINTER(C.c_int)]
    L.gam_ens_destroy.argtypes = [vp]
    _lib = L
    return L


def _p(a):
    if a is None:
        return DP()
    if a.dtype != np.float64 or not a.flags.f_contiguous:
        raise ValueError("arrays must be float64 and Fortran-contiguous")
    return a.ctypes.data_as(DP)


def model_struct(p):
    m = Model()
    for k in ("gamma", "CFL", "Vd0", "Ca", "cHogH", "cHogM", "pdmb", "dFloor", "pFloor",
              "tFin", "dt0", "limDT0"):
        setattr(m, k, float(p[k]))
    for k in ("doMHD", "doBoris", "doBackground", "doLFMLim", "rmeth"):
        setattr(m, k, int(p[k]))
    for s in range(6):
        b = p["bc"][s]
        m.bc[s] = BC[b] if isinstance(b, str) else int(b)
    return m


class GamHip:
    """One device context: Grid, State, oState, StateHf and solver storage."""

    def __init__(self, Nip, Njp, Nkp, params):
        self.L = load()
        self.Nip, self.Njp, self.Nkp = Nip, Njp, Nkp
        self.Ni, self.Nj, self.Nk = Nip + 8, Njp + 8, Nkp + 8
        self.params = dict(params)
        self.ctx = C.c_void_p()
        self.wind = None
        d = Dims(Nip, Njp, Nkp, 4)
        m = model_struct(params)
        self._chk(self.L.gam_create(C.byref(self.ctx), C.byref(d), C.byref(m)))

    def _chk(self, rc):
        if rc != GAM_OK:
            raise GamError(rc, self.L.gam_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "ctx", None):
            self.L.gam_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- shapes
    @property
    def cshape(self):
        return (self.Ni, self.Nj, self.Nk)

    @property
    def fshape(self):
        return (self.Ni + 1, self.Nj + 1, self.Nk + 1)

    # ---- grid
    def grid_from_corners(self, xyz):
        self._chk(self.L.gam_grid_from_corners(self.ctx, _p(np.asfortranarray(xyz, dtype=np.float64))))

    def set_grid(self, g):
        a = GridArrays()
        keep = {}
        for k in GRID_FIELDS:
            keep[k] = np.asfortranarray(g[k], dtype=np.float64)
            setattr(a, k, _p(keep[k]))
        self._chk(self.L.gam_set_grid(self.ctx, C.byref(a)))

    def get_grid(self, fields=GRID_FIELDS):
        c, f = self.cshape, self.fshape
        shapes = {"xyzcc": c + (3,), "xfc": f + (3, 3), "di": c, "dj": c, "dk": c, "volume": c,
                  "Tf": f + (9, 3), "Te": f + (9, 3), "Teb": f + (4, 3), "face": f + (3,), "edge": f + (3,)}
        out = {k: np.zeros(shapes[k], order="F") for k in fields}
        a = GridArrays()
        for k in fields:
            setattr(a, k, _p(out[k]))
        self._chk(self.L.gam_get_grid(self.ctx, C.byref(a)))
        return out

    def add_dipole_b0(self, M0):
        self._chk(self.L.gam_add_dipole_b0(self.ctx, float(M0)))

    def set_background(self, b):
        a = BgArrays()
        keep = {}
        for k in BG_FIELDS:
            keep[k] = np.asfortranarray(b[k], dtype=np.float64)
            setattr(a, k, _p(keep[k]))
        self._chk(self.L.gam_set_background(self.ctx, C.byref(a)))

    def get_background(self):
        c = self.cshape
        shapes = {"fcB0": c + (3, 3), "edgB0": c + (2, 3), "B0": c + (3,), "bFlux0": c + (3,), "dpB0": c + (3,)}
        out = {k: np.zeros(s, order="F") for k, s in shapes.items()}
        a = BgArrays()
        for k in BG_FIELDS:
            setattr(a, k, _p(out[k]))
        self._chk(self.L.gam_get_background(self.ctx, C.byref(a)))
        return out

    def set_wind(self, wind):
        tW = np.asfortranarray(wind["tW"], dtype=np.float64)
        Q = np.asfortranarray(wind["Q"], dtype=np.float64)
        B = np.asfortranarray(wind["B"], dtype=np.float64)
        xyzW = np.asfortranarray(wind.get("xyzW", (0.0, 0.0, 0.0)), dtype=np.float64)
        self._chk(self.L.gam_set_wind(self.ctx, len(tW), _p(tW), _p(Q), _p(B),
                                      float(wind.get("Bx0", 0.0)), float(wind.get("ByC", 0.0)),
                                      float(wind.get("BzC", 0.0)), _p(xyzW)))
        self.wind = dict(wind)

    # ---- state
    def put_state(self, st, which=GAM_STATE):
        arrs = {k: (np.asfortranarray(st[k], dtype=np.float64) if st.get(k) is not None else None)
                for k in ("Gas", "magFlux", "Bxyz", "Efld")}
        self._chk(self.L.gam_put_state(self.ctx, which, _p(arrs["Gas"]), _p(arrs["magFlux"]),
                                       _p(arrs["Bxyz"]), _p(arrs["Efld"]), float(st.get("time", 0.0))))

    def get_state(self, which=GAM_STATE):
        c, f = self.cshape, self.fshape
        out = {"Gas": np.zeros(c + (5,), order="F"), "magFlux": np.zeros(f + (3,), order="F"),
               "Bxyz": np.zeros(c + (3,), order="F"), "Efld": np.zeros(c + (3,), order="F")}
        t = C.c_double()
        self._chk(self.L.gam_get_state(self.ctx, which, _p(out["Gas"]), _p(out["magFlux"]),
                                       _p(out["Bxyz"]), _p(out["Efld"]), C.byref(t)))
        out["time"] = t.value
        return out

    def get_solver(self, which):
        c = self.cshape
        out = np.zeros(c + ((5, 3) if which == 0 else (3, 3)), order="F")
        self._chk(self.L.gam_get_solver(self.ctx, which, _p(out)))
        return out

    def set_time(self, t, dt, ts):
        self._chk(self.L.gam_set_time(self.ctx, float(t), float(dt), int(ts)))

    def set_dt0(self, dt0):
        self._chk(self.L.gam_set_dt0(self.ctx, float(dt0)))

    def get_dt0(self):
        v = C.c_double()
        self._chk(self.L.gam_get_dt0(self.ctx, C.byref(v)))
        return v.value

    def get_time(self):
        t, dt, ts = C.c_double(), C.c_double(), C.c_int()
        self._chk(self.L.gam_get_time(self.ctx, C.byref(t), C.byref(dt), C.byref(ts)))
        return t.value, dt.value, ts.value

    # ---- operators
    def init_run(self):
        dt = C.c_double()
        self._chk(self.L.gam_init_run(self.ctx, C.byref(dt)))
        return dt.value

    def advance_mhd(self, dt):
        self._chk(self.L.gam_advance_mhd(self.ctx, float(dt)))

    def calc_dt(self):
        dt = C.c_double()
        self._chk(self.L.gam_calc_dt(self.ctx, C.byref(dt)))
        return dt.value

    def enforce_bcs(self):
        self._chk(self.L.gam_enforce_bcs(self.ctx))

    def bflux2fld(self, which=GAM_STATE):
        self._chk(self.L.gam_bflux2fld(self.ctx, which))

    def step(self, sync=True):
        if sync:
            dt = C.c_double()
            self._chk(self.L.gam_step(self.ctx, C.byref(dt)))
            return dt.value
        self._chk(self.L.gam_step(self.ctx, DP()))
        return None

    def run(self, nsteps):
        self._chk(self.L.gam_run(self.ctx, int(nsteps)))

    def advance_model(self, dt):
        """gamAdvanceModel: step until Model%t >= t0 + dt (at least once);
        returns the number of steps taken."""
        n = C.c_int()
        self._chk(self.L.gam_advance_model(self.ctx, float(dt), C.byref(n)))
        return n.value

    def sync(self):
        self._chk(self.L.gam_sync(self.ctx))

    def divb(self):
        v = C.c_double()
        self._chk(self.L.gam_divb(self.ctx, C.byref(v)))
        return v.value

    def time_steps(self, nsteps):
        ms = C.c_double()
        self._chk(self.L.gam_time_steps(self.ctx, int(nsteps), C.byref(ms)))
        return ms.value

    def set_profiling(self, on):
        self._chk(self.L.gam_set_profiling(self.ctx, 1 if on else 0))

    # ---- launch plans
    def autotune(self, reps=3):
        """Time every launch shape of every tunable slot on this grid and keep
        the fastest (the run's state is left as it was); returns the plan."""
        self._chk(self.L.gam_autotune(self.ctx, int(reps)))
        return self.get_plan()

    def get_plan(self):
        v = (C.c_int * TUNE_SLOTS)()
        self._chk(self.L.gam_get_plan(self.ctx, v))
        return {self.L.gam_tune_slot_name(q).decode(): v[q] for q in range(TUNE_SLOTS)}

    def set_plan(self, plan):
        names = [self.L.gam_tune_slot_name(q).decode() for q in range(TUNE_SLOTS)]
        v = (C.c_int * TUNE_SLOTS)(*[int(plan[n]) for n in names])
        self._chk(self.L.gam_set_plan(self.ctx, v))

    def kernel_times(self):
        ms = (C.c_double * 8)()
        n = (C.c_int * 8)()
        self._chk(self.L.gam_kernel_times(self.ctx, ms, n))
        return {self.L.gam_kernel_group_name(q).decode(): (ms[q], n[q]) for q in range(8)}


class GamEnsemble:
    """Independent contexts stepped together: one launch per step kernel
    over all members, one captured hipGraph per step.  The members stay
    usable as GamHip objects (state transfers show the ensemble's
    progress); they must outlive the ensemble."""

    def __init__(self, members):
        self.L = load()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.ctx for m in self.members])
        self.ens = C.c_void_p()
        self._chk(self.L.gam_ens_create(C.byref(self.ens), len(self.members), arr))

    def _chk(self, rc):
        if rc != GAM_OK:
            raise GamError(rc, self.L.gam_last_error().decode())
        return rc

    def run(self, nsteps):
        self._chk(self.L.gam_ens_run(self.ens, int(nsteps)))

    def time_steps(self, nsteps):
        ms = C.c_double()
        self._chk(self.L.gam_ens_time_steps(self.ens, int(nsteps), C.byref(ms)))
        return ms.value

    def launches_per_step(self):
        n = C.c_int()
        self._chk(self.L.gam_ens_launches_per_step(self.ens, C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "ens", None):
            self.L.gam_ens_destroy(self.ens)
            self.ens = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- adjoint
# Every symbol include/gamadj.h declares (exported by libgamhip)
ADJ_EXPORTS = ("gam_adj_create", "gam_adj_create_budget", "gam_adj_tape_info", "gam_adj_set_ostate_bc_time",
               "gam_adj_destroy",
               "gam_adj_set_weights", "gam_adj_forward", "gam_adj_backward", "gam_adj_replay",
               "gam_adj_wind_grad", "gam_adj_times", "gam_adj_set_dt_mode", "gam_adj_dt_grad",
               "gam_adj_replay_dt", "gam_adj_step_dt")


def _load_adj(L):
    vp = C.c_void_p
    L.gam_adj_create.argtypes = [C.POINTER(vp), vp, C.c_int, C.c_int]
    L.gam_adj_create_budget.argtypes = [C.POINTER(vp), vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.gam_adj_tape_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.gam_adj_set_ostate_bc_time.argtypes = [vp, C.c_int]
    L.gam_adj_destroy.argtypes = [vp]
    L.gam_adj_set_weights.argtypes = [vp, DP]
    L.gam_adj_forward.argtypes = [vp, DP]
    L.gam_adj_backward.argtypes = [vp, DP, DP, DP, DP]
    L.gam_adj_replay.argtypes = [vp, DP, DP, DP, DP, DP]
    L.gam_adj_wind_grad.argtypes = [vp, DP, DP]
    L.gam_adj_times.argtypes = [vp, DP, DP]
    L.gam_adj_set_dt_mode.argtypes = [vp, C.c_int]
    L.gam_adj_dt_grad.argtypes = [vp, DP]
    L.gam_adj_replay_dt.argtypes = [vp, DP, DP, DP, DP, DP, DP]
    L.gam_adj_step_dt.argtypes = [vp, DP]
    return L


class GamAdjoint:
    """Discrete adjoint of nsteps steps of a context (any step option and
    BC kind) for the objective J = sum_c w_c P_c of the final State, with
    State/oState checkpoints every ckpt_every steps; wind_grad gives the
    gradient with respect to the inflow series."""

    def __init__(self, g, nsteps, ckpt_every, max_taped=None, overlap=True):
        """max_taped / overlap: the tape budget (gam_adj_create_budget);
        by default as many taped steps as HBM allows, with the overlapped
        sweep where a second tape pair fits."""
        self.L = _load_adj(load())
        self.g = g
        self.adj = C.c_void_p()
        if max_taped is None and overlap:
            self._chk(self.L.gam_adj_create(C.byref(self.adj), g.ctx, int(nsteps), int(ckpt_every)))
        else:
            self._chk(self.L.gam_adj_create_budget(C.byref(self.adj), g.ctx, int(nsteps), int(ckpt_every),
                                                   -1 if max_taped is None else int(max_taped),
                                                   1 if overlap else 0))

    def set_ostate_bc_time(self, at_state_time):
        """oState's ghosts at State%time (True, what init_run produces) or
        at oState%time (False, the default); next forward on."""
        self._chk(self.L.gam_adj_set_ostate_bc_time(self.adj, 1 if at_state_time else 0))

    def tape_info(self):
        """(steps of an interval with taped intermediates, overlapped sweep)"""
        kt, ov = C.c_int(), C.c_int()
        self._chk(self.L.gam_adj_tape_info(self.adj, C.byref(kt), C.byref(ov)))
        return kt.value, bool(ov.value)

    def _chk(self, rc):
        if rc != GAM_OK:
            raise GamError(rc, self.L.gam_last_error().decode())
        return rc

    def set_weights(self, w):
        w = np.asfortranarray(w, dtype=np.float64)
        if w.shape != self.g.cshape[:3]:
            raise ValueError("weights must be (Ni, Nj, Nk)")
        self._chk(self.L.gam_adj_set_weights(self.adj, _p(w)))

    def forward(self):
        J = C.c_double()
        self._chk(self.L.gam_adj_forward(self.adj, C.byref(J)))
        return J.value

    def backward(self, fluxes=False):
        """Gradients with respect to the starting State and oState gas, and
        with fluxes=True also their magFlux: (gS, gO) or (gS, gO, gMF, gMFO).
        gMF is the gradient of a run whose start has every cell's B formed
        from the face fluxes after the boundary pass (after init_run, call
        bflux2fld on GAM_STATE and GAM_OSTATE)."""
        shp = self.g.cshape[:3] + (5,)
        gS = np.zeros(shp, order="F")
        gO = np.zeros(shp, order="F")
        if not fluxes:
            self._chk(self.L.gam_adj_backward(self.adj, _p(gS), _p(gO), None, None))
            return gS, gO
        fshp = self.g.fshape + (3,)
        gM = np.zeros(fshp, order="F")
        gMO = np.zeros(fshp, order="F")
        self._chk(self.L.gam_adj_backward(self.adj, _p(gS), _p(gO), _p(gM), _p(gMO)))
        return gS, gO, gM, gMO

    def replay(self, Gas, oGas=None, MF=None, oMF=None):
        J = C.c_double()
        arr = [None if x is None else np.asfortranarray(x, dtype=np.float64) for x in (Gas, oGas, MF, oMF)]
        self._chk(self.L.gam_adj_replay(self.adj, *[_p(x) for x in arr], C.byref(J)))
        return J.value

    def set_dt_mode(self, mode):
        """0: dt of every step held (default); 1: also dJ/d(dt_n) with the
        other steps' dt held; 2: dt after the first step follows CalcDT of
        the state (gradients of that run, dJ/d(dt_n) total); 3: the first
        dt too (CalcDT of the start State, as after init_run).  Modes 2
        and 3 take effect at the next forward."""
        self._chk(self.L.gam_adj_set_dt_mode(self.adj, int(mode)))

    def dt_grad(self, nsteps):
        """dJ/d(dt_n), n = 0..nsteps-1, of the last backward."""
        g = np.zeros(nsteps)
        self._chk(self.L.gam_adj_dt_grad(self.adj, _p(g)))
        return g

    def step_dt(self, nsteps):
        """dt of each step of the recorded run."""
        d = np.zeros(nsteps)
        self._chk(self.L.gam_adj_step_dt(self.adj, _p(d)))
        return d

    def replay_dt(self, Gas, oGas=None, MF=None, oMF=None, dts=None):
        """replay with the dt of step n = dts[n] (0: CalcDT's, for the
        first step of the start after its boundary pass), or (dts None) the
        recorded first dt and CalcDT's after it."""
        J = C.c_double()
        arr = [None if x is None else np.asfortranarray(x, dtype=np.float64) for x in (Gas, oGas, MF, oMF)]
        d = None if dts is None else np.ascontiguousarray(dts, dtype=np.float64)
        self._chk(self.L.gam_adj_replay_dt(self.adj, *[_p(x) for x in arr], _p(d), C.byref(J)))
        return J.value

    def wind_grad(self, n):
        """Gradient with respect to the inflow series of the last backward:
        (gQ (n,5), gB (n,3)) in the layout set_wind takes."""
        gQ = np.zeros((n, 5), order="F")
        gB = np.zeros((n, 3), order="F")
        self._chk(self.L.gam_adj_wind_grad(self.adj, _p(gQ), _p(gB)))
        return gQ, gB

    def times(self):
        f = C.c_double()
        b = C.c_double()
        self._chk(self.L.gam_adj_times(self.adj, C.byref(f), C.byref(b)))
        return f.value, b.value

    def close(self):
        if getattr(self, "adj", None):
            self.L.gam_adj_destroy(self.adj)
            self.adj = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- restart files
RES_LIBPATH = os.path.join(HERE, "libgamres.so")
NG = 4
RES_EXPORTS = ("gam_res_last_error", "gam_res_write", "gam_res_read", "gam_write_restart",
               "gam_read_restart")


class ResMeta(C.Structure):
    """gam_res_meta: Model_T/State_T scalars of a restart."""
    _fields_ = [("nOut", C.c_int), ("nRes", C.c_int), ("ts", C.c_int), ("t", C.c_double),
                ("ot", C.c_double), ("dt", C.c_double), ("dt0", C.c_double)]


_reslib = None


def load_restart():
    """Load libgamres.so (restart files over HDF5); raises if it was not built."""
    global _reslib
    if _reslib is not None:
        return _reslib
    load()
    if not os.path.exists(RES_LIBPATH):
        raise OSError("libgamres.so is not built (HDF5 missing when __graft_entry__.build() ran)")
    L = C.CDLL(RES_LIBPATH)
    vp = C.c_void_p
    L.gam_res_last_error.restype = C.c_char_p
    L.gam_res_write.argtypes = [C.c_char_p, C.POINTER(Dims), C.POINTER(ResMeta)] + [DP] * 7
    L.gam_res_read.argtypes = [C.c_char_p, C.POINTER(Dims), C.POINTER(ResMeta)] + [DP] * 4
    L.gam_write_restart.argtypes = [vp, C.c_char_p, DP, C.c_int, C.c_int]
    L.gam_read_restart.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), DP]
    _reslib = L
    return L


def _res_chk(L, rc):
    if rc != 0:
        raise GamError(rc, L.gam_res_last_error().decode())


def res_write(path, dims, meta, xyz, st, ost):
    """Write a restart from host arrays: st/ost dicts with Gas, magFlux and
    optionally Bxyz; meta a dict of ResMeta fields."""
    L = load_restart()
    d = Dims(*dims, NG)
    m = ResMeta(**meta)
    _res_chk(L, L.gam_res_write(str(path).encode(), C.byref(d), C.byref(m), _p(xyz),
                                _p(st["Gas"]), _p(st["magFlux"]), _p(st.get("Bxyz")),
                                _p(ost["Gas"]), _p(ost["magFlux"]), _p(ost.get("Bxyz"))))


def res_read(path, dims):
    """Read a restart into host arrays: (meta dict, state dict, oState dict)."""
    L = load_restart()
    Nip, Njp, Nkp = dims
    c = (Nip + 2 * NG, Njp + 2 * NG, Nkp + 2 * NG)
    f = (c[0] + 1, c[1] + 1, c[2] + 1)
    st = {"Gas": np.zeros(c + (5,), order="F"), "magFlux": np.zeros(f + (3,), order="F")}
    ost = {"Gas": np.zeros(c + (5,), order="F"), "magFlux": np.zeros(f + (3,), order="F")}
    d = Dims(Nip, Njp, Nkp, NG)
    m = ResMeta()
    _res_chk(L, L.gam_res_read(str(path).encode(), C.byref(d), C.byref(m), _p(st["Gas"]),
                               _p(st["magFlux"]), _p(ost["Gas"]), _p(ost["magFlux"])))
    return {k: getattr(m, k) for k, _ in ResMeta._fields_}, st, ost


def write_restart(g, path, xyz=None, nOut=0, nRes=0):
    """WriteRestart of a GamHip context."""
    L = load_restart()
    _res_chk(L, L.gam_write_restart(g.ctx, str(path).encode(), _p(xyz), int(nOut), int(nRes)))


def read_restart(g, path):
    """Restart a GamHip context from a file: returns (nOut, nRes, dt)."""
    L = load_restart()
    nOut, nRes, dt = C.c_int(), C.c_int(), C.c_double()
    _res_chk(L, L.gam_read_restart(g.ctx, str(path).encode(), C.byref(nOut), C.byref(nRes),
                                   C.byref(dt)))
    return nOut.value, nRes.value, dt.value
