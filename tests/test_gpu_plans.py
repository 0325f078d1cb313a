"""Launch plans: every launch shape of every tunable slot gives the results
of the default plan (relative L-infinity 1e-12 after the steps; the shapes
only change which thread computes which point, and the j/k flux forms run
the same per-face arithmetic), and gam_autotune leaves the run where it was.
"""
import pytest

from helpers import rel_linf, shell_case, box_case
from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

TOL = 1e-12
KEYS = ("Gas", "magFlux", "Bxyz", "Efld")


def _diff(a, b):
    sa, sb = a.get_state(), b.get_state()
    return max(rel_linf(sa[k], sb[k]) for k in KEYS)


def _candidates(slot):
    L = gamhip.load()
    import ctypes as C
    v = (C.c_int * 16)()
    n = L.gam_tune_candidates(slot, v)
    return [v[q] for q in range(n)]


@pytest.mark.parametrize("kind", ["shell", "box"])
def test_every_shape_matches_default(kind):
    N = (24, 16, 24) if kind == "shell" else (16, 16, 16)
    mk = (lambda: shell_case(N)[0]) if kind == "shell" else (lambda: box_case(N, "8CENT")[0])
    ref = mk()
    ref.init_run()
    ref.run(3)
    base = ref.get_plan()
    names = list(base)
    worst = 0.0
    for q, name in enumerate(names):
        for v in _candidates(q):
            if v == base[name]:
                continue
            g = mk()
            plan = dict(base)
            plan[name] = v
            g.set_plan(plan)
            assert g.get_plan() == plan
            g.init_run()
            g.run(3)
            d = _diff(g, ref)
            worst = max(worst, d)
            assert d <= TOL, (name, v, d)
            g.close()
    print("%s: all shapes within %.3e of the default plan" % (kind, worst), flush=True)
    ref.close()


def test_autotune_keeps_the_run():
    a = shell_case((40, 32, 40))[0]
    b = shell_case((40, 32, 40))[0]
    a.init_run()
    b.init_run()
    a.run(2)
    b.run(2)
    t0 = a.get_time()
    plan = a.autotune(reps=2)
    print("tuned plan", plan, flush=True)
    assert a.get_time() == t0
    for q, (name, v) in enumerate(plan.items()):
        assert v in _candidates(q), (name, v)
    assert _diff(a, b) == 0.0
    a.run(4)
    b.run(4)
    assert a.get_time() == b.get_time()
    assert _diff(a, b) <= TOL
    a.close()
    b.close()


def test_ensemble_with_mixed_plans():
    specs = pb.ensemble_members(n=3, lo=(16, 16, 16), hi=(32, 24, 32))
    from test_gpu_ensemble import member
    batched = [member(s) for s in specs]
    alone = [member(s) for s in specs]
    p = batched[1].get_plan()
    p["flux_j"] = 1
    p["update"] = 2
    batched[1].set_plan(p)
    ens = gamhip.GamEnsemble(batched)
    plain = gamhip.GamEnsemble(alone[:1])
    assert ens.launches_per_step() > plain.launches_per_step()
    plain.close()
    ens.run(5)
    for g in alone:
        g.run(5)
    for a, b in zip(batched, alone):
        assert _diff(a, b) <= TOL
    ens.close()
    for g in batched + alone:
        g.close()
