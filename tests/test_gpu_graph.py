"""Steps replayed as a graph of two (gam_run, gam_time_steps) against the
same steps enqueued one by one (gam_step): the graph replays the very
launches of a step, so the states must agree bit for bit, across both
State-buffer parities, over several replays of each graph, and after a
launch argument changes (launch plan, inflow series), which drops the
captured graphs."""
import numpy as np
import pytest

from helpers import box_case, shell_case
from mage_amd import gamhip, problems as pb

pytestmark = pytest.mark.gpu

KEYS = ("Gas", "magFlux", "Bxyz", "Efld")


def _same(a, b):
    sa, sb = a.get_state(), b.get_state()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    assert a.get_time() == b.get_time()


def _plain(g, n):
    for _ in range(n):
        g.step()


@pytest.mark.parametrize("kind", ["shell", "box"])
def test_graph_steps_match_plain_steps(kind):
    N = (24, 16, 24) if kind == "shell" else (16, 16, 16)
    mk = (lambda: shell_case(N)[0]) if kind == "shell" else (lambda: box_case(N, "7UP")[0])
    a, b = mk(), mk()
    a.init_run()
    b.init_run()
    a.run(9)            # four replays from buffer 0, one plain step
    a.run(6)            # three replays of the graph from buffer 1
    a.time_steps(4)     # two more from buffer 1
    _plain(b, 19)
    _same(a, b)
    a.close()
    b.close()


def test_graph_dropped_when_launch_arguments_change():
    N = (24, 16, 24)
    a, b = shell_case(N)[0], shell_case(N)[0]
    a.init_run()
    b.init_run()
    a.run(4)
    _plain(b, 4)
    plan = a.get_plan()
    plan["flux_j"] = 1 if plan["flux_j"] != 1 else 4
    a.set_plan(plan)
    b.set_plan(plan)
    a.run(4)
    _plain(b, 4)
    _same(a, b)
    wind = pb.constant_wind(vx=-0.5, bz=-0.08)
    a.set_wind(wind)
    b.set_wind(wind)
    a.run(4)
    _plain(b, 4)
    _same(a, b)
    a.close()
    b.close()
