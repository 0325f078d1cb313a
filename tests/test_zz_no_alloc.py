"""gam_divb_noalloc (no device allocation, scratch in a static device array)
against gam_divb (per-block partials in the context's own buffer).

Both return the total |div B| over the active cells (DivB,
gridutils.F90:717-770).  The inputs are face-flux arrays with a real
divergence (random fluxes, so every cell contributes), on a shell with the
dipole's background fluxes and on a smaller box without, called one after
the other: the second, smaller call must not see anything the first left
in the shared scratch.  The two reductions add the same terms in different
orders, so they agree to round-off (<= 1e-12 relative).
"""
import ctypes as C

import numpy as np
import pytest

from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)


def _hip():
    hip = C.CDLL("libamdhip64.so.7")          # the runtime libgamhip is linked against
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _dev(hip, a, keep):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    keep.append(p)
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0          # host to device
    return C.cast(p, DP)


def _case(N, shell, seed):
    if shell:
        g = gamhip.GamHip(*N, pb.shell_params("7UP"))
        g.grid_from_corners(pb.shell_corners(*N))
        g.add_dipole_b0(pb.SHELL_M0)
        g.set_wind(pb.constant_wind())        # the shell's outer BC needs its series
        b0 = g.get_background()["bFlux0"]
    else:
        g = gamhip.GamHip(*N, pb.box_params("7UP"))
        g.grid_from_corners(pb.cartesian_corners(*N))
        b0 = None
    rng = np.random.default_rng(seed)
    st = pb.empty_state(*N)
    st["Gas"][..., 0] = 1.0
    st["Gas"][..., 4] = 1.0
    st["magFlux"][...] = rng.standard_normal(st["magFlux"].shape)
    g.put_state(st)
    return g, st["magFlux"], b0


def test_divb_noalloc_matches_divb_twice():
    L = gamhip.load()
    L.gam_divb_noalloc.argtypes = [C.POINTER(gamhip.Dims), DP, DP, C.c_void_p, DP]
    hip = _hip()
    keep = []
    try:
        out = np.zeros(1)
        dout = _dev(hip, out, keep)
        for q, (N, shell) in enumerate((((32, 24, 40), True), ((16, 16, 16), False))):
            print("divb noalloc call %d: %s %s" % (q, N, "shell with bFlux0" if shell else "box"), flush=True)
            g, mf, b0 = _case(N, shell, 100 + q)
            ref = g.divb()
            dmf = _dev(hip, np.asfortranarray(mf), keep)
            db0 = _dev(hip, np.asfortranarray(b0), keep) if b0 is not None else None
            d = gamhip.Dims(*N, 4)
            assert L.gam_divb_noalloc(C.byref(d), dmf, db0, None, dout) == 0, L.gam_last_error()
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(out.ctypes.data, C.cast(dout, C.c_void_p), 8, 2) == 0   # device to host
            print("divb noalloc call %d: %.15e vs gam_divb %.15e" % (q, out[0], ref), flush=True)
            assert ref > 0.0
            assert abs(out[0] - ref) <= 1e-12 * ref
            g.close()
        # sizes beyond the supported maximum are refused before any launch
        big = gamhip.Dims(2048, 2048, 1024, 4)
        assert L.gam_divb_noalloc(C.byref(big), dmf, None, None, dout) == -1     # GAM_E_ARG
    finally:
        for p in keep:
            hip.hipFree(p)
