"""GPU tests of the pressure solve's deflation-space start: loop-1 solves
start from the Galerkin projection x0 = W y, (W^T A W) y = W^T b, onto the
last p_recycle loop-1 solutions (and the learned guess when the learned form
uses it), then run the unchanged PCG-AMG to its 1e-8 relative residual.

* the start's residual ratio ||b - A x0|| / ||b - mean b|| that the library
  reports equals the one numpy's projection gives on the same vectors and
  system, and the solve meets its tolerance;
* recycling lowers the PCG iteration count of the solves and leaves the
  fields within what the 1e-8 solves allow of the zero-started run;
* a net whose weights are NaN is dropped from W: with no recycled vectors
  the solve starts from zero and every field equals the learned = 0 run bit
  for bit, with them the recycled start is kept.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from minicombust_app_amd import learned
from minicombust_app_amd import mcflow as mc
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FIELDS = [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]
BENCH = dict(mode=mc.MODE_SIMPLE, p_rtol=1e-8, mom_rtol=1e-5)


def _pressure_system(g):
    rp = g.get(mc.ARR_ROW_PTR)
    col = g.get(mc.ARR_COL)
    A = sp.csr_matrix((g.get(mc.ARR_PA_VALS), col, rp), shape=(g.N, g.N))
    return A, g.get(mc.ARR_PB)


@pytest.mark.parametrize("psolve", [mc.PSOLVE_FP64, mc.PSOLVE_MIXED])
def test_start_is_the_galerkin_projection(psolve):
    """After 4 SIMPLE iterations at modifier 16 (p_recycle = 2) the two
    newest loop-1 solutions are read back; the next loop-1 system is
    assembled stage by stage and solved: the library's start residual ratio
    equals numpy's Galerkin projection onto those two vectors to 1e-6
    (relative), the projection leaves under a tenth of the zero start's
    residual, and the solve meets 1e-8."""
    m = 16
    print("deflation start m=%d psolve %d" % (m, psolve), flush=True)
    with mc.Context(m, psolve=psolve, p_recycle=2, **BENCH) as g:
        for it in range(4):
            g.simple_iteration(it % 3)
        W = np.stack([g.get(mc.ARR_PDEFL), g.get(mc.ARR_PDEFL + 1)], axis=1)
        assert np.all(np.abs(W.mean(axis=0)) <= 1e-12 * np.abs(W).max(axis=0))
        g.get_phi_gradients()
        g.momentum_assemble()
        g.momentum_solve()
        g.pressure_assemble()
        A, b = _pressure_system(g)
        bt = b - b.mean()
        AW = A @ W
        y = np.linalg.solve(W.T @ AW, W.T @ bt)
        ratio_np = np.linalg.norm(bt - AW @ y) / np.linalg.norm(bt)
        s0 = g.stats()
        st = g.pressure_solve()
        s1 = g.stats()
        assert s1["p_defl_starts"] == s0["p_defl_starts"] + 1
        assert s1["p_defl_vectors"] - s0["p_defl_vectors"] == 2
        ratio = s1["p_defl_ratio_sum"] - s0["p_defl_ratio_sum"]
        print("start residual ratio: library %.6e, numpy %.6e" % (ratio, ratio_np), flush=True)
        assert ratio_np < 0.1
        assert abs(ratio - ratio_np) <= 1e-6 * ratio_np + 1e-14
        assert st.converged and st.relres <= 1e-8 * 1.0001
        x = g.phi(po.PP)[:g.N]
        assert np.linalg.norm(bt - A @ x) <= 1.0001e-8 * np.linalg.norm(bt)


@pytest.mark.parametrize("psolve", [mc.PSOLVE_FP64, mc.PSOLVE_MIXED])
def test_recycling_cuts_pressure_iterations(psolve):
    """Modifier 64, the benchmark's options, 9 SIMPLE iterations with and
    without recycling (p_recycle = 2): fewer PCG iterations with it, every
    solve at 1e-8, fields within 1e-7 of each field's maximum of the
    zero-started run."""
    m = 64
    print("recycling m=%d psolve %d" % (m, psolve), flush=True)
    out = {}
    for rec in (0, 2):
        with mc.Context(m, psolve=psolve, p_recycle=rec, **BENCH) as g:
            for it in range(9):
                g.simple_iteration(it % 3)
            st = g.stats()
            assert st["p_relres_max"] <= 1e-8 * 1.0001, st["p_relres_max"]
            out[rec] = (st, {f: g.phi(f)[:g.N].copy() for f in FIELDS})
    it0, it2 = out[0][0]["p_iters"], out[2][0]["p_iters"]
    starts = out[2][0]["p_defl_starts"]
    print("PCG iterations: zero start %d, recycled %d (%d recycled starts, mean ratio %.2e)"
          % (it0, it2, starts, out[2][0]["p_defl_ratio_sum"] / max(starts, 1)), flush=True)
    assert starts >= 7
    assert it2 < it0
    for f in FIELDS:
        a, b = out[2][1][f], out[0][1][f]
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
        assert err <= 1e-7, (po.FIELD_NAMES[f], err)


def test_nan_net_falls_back():
    """Weights set to NaN: every net output is NaN.  learned = 2 without
    recycling gives the learned = 0 fields bit for bit (the guess is dropped,
    the solve starts from zero); with p_recycle = 2 the recycled start is
    still taken and the solves meet 1e-8."""
    m = 16
    print("NaN net m=%d" % m, flush=True)
    w = learned.load_weights()
    bad = np.full_like(w, np.nan)
    runs = {}
    for name, kw, weights in [("plain", dict(learned=0), None), ("nan", dict(learned=2), bad),
                              ("nan_rec", dict(learned=2, p_recycle=2), bad)]:
        with mc.Context(m, psolve=mc.PSOLVE_MIXED, **kw, **BENCH) as g:
            if weights is not None:
                g.mpnn_load(weights)
            for it in range(6):
                g.simple_iteration(it % 3)
            st = g.stats()
            runs[name] = (st, {f: g.phi(f).copy() for f in FIELDS})
    assert runs["nan"][0]["p_guesses"] == 0 and runs["nan"][0]["p_defl_starts"] == 0
    for f in FIELDS:
        assert np.array_equal(runs["nan"][1][f], runs["plain"][1][f]), po.FIELD_NAMES[f]
    st = runs["nan_rec"][0]
    assert st["p_guesses"] == 0 and st["p_defl_starts"] >= 4
    assert st["p_relres_max"] <= 1e-8 * 1.0001
