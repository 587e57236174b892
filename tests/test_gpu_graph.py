"""GPU test of the pressure PCG iterations replayed as HIP graphs
(mc_options.pcg_graph = 1; off by default) against the same solves launched
kernel by kernel (pcg_graph = 0): the graph holds the same kernels in the same
order, so every field and the mass flux must agree bit for bit.  Each graph is
replayed many times with the SIMPLE iteration's eager work (assembly,
BiCGStab, corrections, host polls) between replays; the runs cross coarse-
level refreshes (every 4th solve), a change of the smoothing sweeps (the
graphs are captured again) and both pressure forms."""
import numpy as np
import pytest

from minicombust_app_amd import mcflow as mc
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]


@pytest.mark.parametrize("psolve", [mc.PSOLVE_FP64, mc.PSOLVE_MIXED])
def test_graph_replay_equals_eager_launches(psolve):
    m = 32
    print("pcg graphs m=%d psolve %d" % (m, psolve), flush=True)
    out = {}
    for graph in (0, 1):
        with mc.Context(m, mode=mc.MODE_SIMPLE, psolve=psolve, p_rtol=1e-8, mom_rtol=1e-5, p_recycle=2,
                        pcg_graph=graph) as g:
            for it in range(6):
                g.simple_iteration(it % 3)
            g.set_options(amg_nu=2)   # other smoothing weights: captured again
            for it in range(6, 9):
                g.simple_iteration(it % 3)
            st = g.stats()
            assert st["p_relres_max"] <= 1e-8 * 1.0001
            out[graph] = ({f: g.phi(f).copy() for f in FIELDS}, g.get(mc.ARR_MFLUX).copy(), st["p_iters"])
    print("PCG iterations: eager %d, graphs %d" % (out[0][2], out[1][2]), flush=True)
    assert out[0][2] == out[1][2]
    for f in FIELDS:
        assert np.array_equal(out[0][0][f], out[1][0][f]), po.FIELD_NAMES[f]
    assert np.array_equal(out[0][1], out[1][1])
