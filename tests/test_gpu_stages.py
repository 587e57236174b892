"""GPU test of the per-stage timers (mc_stage_*), the counterpart of the
reference's fine timers (vel_/pres_/sca_*_time, FlowSolver.hpp:128-141):
every stage of a SIMPLE iteration is bracketed, the brackets are disjoint on
the one stream, so their sum lies within the host's wall time of the same
iterations, and nothing is timed while the stages are off."""
import time

import pytest

from minicombust_app_amd import mcflow as mc

pytestmark = pytest.mark.gpu


def test_stage_times_cover_the_iteration():
    print("stages m=16", flush=True)
    with mc.Context(16, mode=mc.MODE_SIMPLE, p_rtol=1e-8) as g:
        g.simple_iteration(0)
        g.sync()
        assert all(v["calls"] == 0 for v in g.stage_read().values())
        g.stage_enable(True)
        t0 = time.perf_counter()
        for it in range(1, 4):
            g.simple_iteration(it % 3)
        g.sync()
        wall_ms = 1e3 * (time.perf_counter() - t0)
        st = g.stage_read()
        total = sum(v["ms"] for v in st.values())
        print("stage ms: %s; sum %.3f of wall %.3f" % (
            ", ".join("%s %.3f" % (k, v["ms"]) for k, v in st.items()), total, wall_ms), flush=True)
        for name, v in st.items():
            assert v["calls"] >= 3 and v["ms"] > 0.0, name
        assert total <= wall_ms
        g.stage_enable(False)
        g.stage_reset()
        g.simple_iteration(1)
        g.sync()
        assert all(v["calls"] == 0 and v["ms"] == 0.0 for v in g.stage_read().values())
