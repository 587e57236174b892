"""Parity of the HIP SIMPLE step with the CPU oracle.

Stage tests feed both implementations the same state (the oracle's, pushed
into the HIP context) and compare the stage's outputs.  Stages without a
linear solve are compared bit for bit: the kernels perform the reference's
operations in the reference's order without FP contraction.  Solves are
compared to a tolerance set by the solver tolerance (both sides solve to
1e-12 relative residual here).
"""
import numpy as np
import pytest

from minicombust_app_amd import mcflow as mc
from oracle import pyoracle as po
from tests.helpers import (CELL_ARRAYS, FACE_ARRAYS, assert_bits, assert_close, oracle_pressure_assemble,
                           push_state)

pytestmark = pytest.mark.gpu

TIGHT = dict(mom_rtol=1e-12, p_rtol=1e-12)


def make_pair(m, mode=po.MODE_COMPAT, **kw):
    opts = dict(TIGHT)
    opts.update(kw)
    o = po.Oracle(m, mode=mode, mom_rtol=opts["mom_rtol"], p_rtol=opts["p_rtol"], p_maxit=100000)
    # warm_start=0: the oracle starts every BiCGStab from zero, as the
    # reference does (FlowSolver.inl:1348)
    g = mc.Context(m, mode=mode, mom_rtol=opts["mom_rtol"], p_rtol=opts["p_rtol"], p_maxit=5000, warm_start=0)
    return o, g


@pytest.mark.parametrize("m", [1, 2, 3, 4, 8])
def test_mesh_bits(m):
    print("mesh m=%d" % m, flush=True)
    om = po.oracle_mesh(m)
    with mc.Context(m) as g:
        assert np.array_equal(g.get(mc.ARR_FACES).reshape(-1, 2), om["faces"])
        assert np.array_equal(g.get(mc.ARR_CELL_FACES).reshape(-1, 6), om["cell_faces"])
        assert np.array_equal(g.get(mc.ARR_CELL_NEIGHBOURS).reshape(-1, 6), om["cell_neighbours"])
        assert np.array_equal(g.get(mc.ARR_CELLS).reshape(-1, 8), om["cells"])
        assert_bits(g.get(mc.ARR_POINTS).reshape(-1, 3), om["points"], "points")
        assert np.array_equal(g.get(mc.ARR_CELLS_PER_POINT), om["cells_per_point"])
        assert_bits(g.get(mc.ARR_CELL_CENTER).reshape(-1, 3), om["cell_centers"], "cell centres")


@pytest.mark.parametrize("m", [2, 5])
def test_geometry_and_init_bits(m):
    print("geometry m=%d" % m, flush=True)
    o = po.Oracle(m)
    with mc.Context(m) as g:
        for name, aid in [("face_areas", mc.ARR_FACE_AREA), ("face_lambdas", mc.ARR_FACE_LAMBDA),
                          ("face_rlencos", mc.ARR_FACE_RLENCOS)]:
            assert_bits(g.get(aid), o.face(name), name)
        assert_bits(g.get(mc.ARR_FACE_NORMAL).reshape(-1, 3), o.facevec("face_normals"), "normals")
        assert_bits(g.get(mc.ARR_FACE_CENTER).reshape(-1, 3), o.facevec("face_centers"), "face centres")
        assert_bits(g.get(mc.ARR_CELL_VOLUME), o.cell("cell_volumes"), "volumes")
        for f in range(po.NFIELDS):
            assert_bits(g.phi(f), o.phi(f), "initial phi[%s]" % po.FIELD_NAMES[f])
        rp, col = o.csr()
        assert np.array_equal(g.get(mc.ARR_ROW_PTR), rp)
        assert np.array_equal(g.get(mc.ARR_COL), col)
    o.close()


def test_stages_bits_and_solves():
    m = 4
    print("stages m=%d" % m, flush=True)
    o, g = make_pair(m)
    # A4: gradients
    o.get_phi_gradients()
    g.get_phi_gradients()
    for f in [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]:
        assert_bits(g.grad(f), o.grad(f), "grad %s" % po.FIELD_NAMES[f])
    # A6 + A11: set_up_field
    o.set_up_field()
    g.set_up_field()
    assert_bits(g.get(mc.ARR_AU)[:o.N], o.cell("A_U"), "A_U (precomp)")
    assert_bits(g.get(mc.ARR_MFLUX), o.face("mflux"), "mass flux (set_up_field)")
    assert_bits(g.get(mc.ARR_FF0), o.face("ff0"), "ff0 (set_up_field)")
    assert_bits(g.get(mc.ARR_SU)[:o.N], o.cell("S_U"), "S_U (set_up_field)")
    # A7-A9: momentum assembly
    o.momentum_assemble()
    g.momentum_assemble()
    assert_bits(g.get(mc.ARR_AU)[:o.N], o.cell("A_U"), "A_U")
    assert_bits(g.get(mc.ARR_SU)[:o.N], o.cell("S_U"), "S_U")
    assert_bits(g.get(mc.ARR_A_VALS), o.vals("A_vals"), "momentum CSR values")
    assert_bits(g.get(mc.ARR_FF0), o.face("ff0"), "ff0 (UVW)")
    assert_bits(g.get(mc.ARR_FF1), o.face("ff1"), "ff1 (UVW)")
    for f in [po.U, po.V, po.W]:
        assert_bits(g.phi(f)[o.N:], o.phi(f)[o.N:], "boundary slots %s" % po.FIELD_NAMES[f])
    # A10: momentum solves
    o.momentum_solve(0)
    o.momentum_solve(1)
    o.momentum_solve(2)
    assert_bits(g.get(mc.ARR_SV)[:o.N], o.cell("S_V"), "S_V")
    assert_bits(g.get(mc.ARR_SW)[:o.N], o.cell("S_W"), "S_W")
    st = g.momentum_solve()
    assert st.converged
    for f in [po.U, po.V, po.W]:
        assert_close(g.phi(f)[:o.N], o.phi(f)[:o.N], "momentum %s" % po.FIELD_NAMES[f], 1e-11)
    # A11 + A12 from identical state
    push_state(o, g)
    oracle_pressure_assemble(o)
    g.pressure_assemble()
    assert_bits(g.get(mc.ARR_MFLUX), o.face("mflux"), "mass flux")
    assert_bits(g.get(mc.ARR_FF0), o.face("ff0"), "ff0 (pressure)")
    assert_bits(g.get(mc.ARR_SU)[:o.N], o.cell("S_U"), "pressure RHS")
    assert_bits(g.get(mc.ARR_AV)[:o.N], o.cell("A_V"), "pressure diagonal")
    assert_bits(g.get(mc.ARR_PA_VALS), o.vals("PA_vals"), "pressure CSR values")
    assert_bits(g.get(mc.ARR_PB), o.cell("S_U"), "Pressure_b")
    # A13: pressure solve
    o.solve_pressure_matrix()
    st = g.pressure_solve()
    assert st.converged and st.relres <= 1e-12
    pp_o = o.phi(po.PP)[:o.N]
    assert_close(g.phi(po.PP)[:o.N], pp_o, "pressure correction", 1e-9)
    o.close()
    g.close()


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
def test_calculate_pressure_from_same_state(mode):
    m = 4
    print("calculate_pressure m=%d mode=%d" % (m, mode), flush=True)
    o, g = make_pair(m, mode)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    push_state(o, g)
    o.calculate_pressure()
    g.calculate_pressure()
    for f in [po.U, po.V, po.W, po.P]:
        assert_close(g.phi(f), o.phi(f), "after calculate_pressure %s" % po.FIELD_NAMES[f], 1e-9)
    assert_close(g.get(mc.ARR_MFLUX), o.face("mflux"), "mass flux after correction", 1e-9)
    assert_close(g.grad(po.P), o.grad(po.P), "grad P", 1e-7, scale=max(1e-12, np.abs(o.grad(po.P)).max()))
    o.close()
    g.close()


@pytest.mark.parametrize("t", range(7))
def test_scalar_assemble_bits(t):
    m = 4
    print("scalar %d m=%d" % (t, m), flush=True)
    o, g = make_pair(m)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    o.calculate_pressure()
    for tt in range(t):
        o.scalar_solve(tt, [po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP][tt])
    push_state(o, g)
    field = [po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP][t]
    o.scalar_assemble(t, field)
    g.scalar_assemble(t)
    exact = t != po.TERBED   # ED wall values use pow(), libm vs device ulp
    check = assert_bits if exact else (lambda a, b, w: assert_close(a, b, w, 1e-14))
    check(g.get(mc.ARR_SU)[:o.N], o.cell("S_U"), "scalar %d RHS" % t)
    check(g.get(mc.ARR_AV)[:o.N], o.cell("A_V"), "scalar %d diagonal" % t)
    check(g.get(mc.ARR_A_VALS), o.vals("A_vals"), "scalar %d CSR values" % t)
    check(g.phi(field)[o.N:], o.phi(field)[o.N:], "scalar %d slots" % t)
    o.close()
    g.close()


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
@pytest.mark.parametrize("m", [4, 8])
def test_timesteps(m, mode):
    print("timesteps m=%d mode=%d" % (m, mode), flush=True)
    o, g = make_pair(m, mode)
    for ts in range(2):
        o.timestep()
        g.timestep()
        for f in [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]:
            assert_close(g.phi(f), o.phi(f), "ts %d field %s" % (ts, po.FIELD_NAMES[f]), 1e-8)
        assert_close(g.get(mc.ARR_MFLUX), o.face("mflux"), "ts %d mass flux" % ts, 1e-8)
    o.close()
    g.close()


@pytest.mark.parametrize("psolve", [mc.PSOLVE_FP64, mc.PSOLVE_MIXED])
@pytest.mark.parametrize("m", [16, 32])
def test_pressure_solve_forms(m, psolve):
    """Both pressure forms reach 1e-8 on the assembled pressure-correction system:
    residual checked with an independent SpMV, solution against the oracle's
    Jacobi-PCG solve at 1e-12 (mean-zero representatives)."""
    print("pressure solve m=%d psolve=%d" % (m, psolve), flush=True)
    o = po.Oracle(m, mode=po.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=200000)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    oracle_pressure_assemble(o)
    b = o.cell("S_U").copy()
    o.solve_pressure_matrix()
    x_ref = o.phi(po.PP)[:o.N].copy()
    g = mc.Context(m, mode=mc.MODE_SIMPLE, psolve=psolve, p_rtol=1e-8)
    push_state(o, g)
    x, st = g.solve_pressure(b)
    assert st.converged
    bp = b - b.mean()
    res = np.linalg.norm(bp - g.spmv(x, which=1)) / np.linalg.norm(bp)
    assert res <= 1.05e-8, res
    assert abs(x.mean()) <= 1e-12 * max(1.0, np.abs(x).max())
    assert_close(x, x_ref, "pressure correction vs oracle", 1e-6)
    o.close()
    g.close()


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
def test_timesteps_m16(mode):
    test_timesteps(16, mode)


@pytest.mark.parametrize("warm", [0, 1])
def test_scalar_batch_equals_single(warm):
    """The 4-RHS passive-scalar solve reproduces four single solves bit for bit
    (from zero, and warm-started from the current fields)."""
    m = 8
    o = po.Oracle(m, mode=po.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    o.calculate_pressure()
    o.scalar_solve(po.TERBTE, po.TE)
    o.scalar_solve(po.TERBED, po.ED)
    o.scalar_solve(po.TEMP, po.TEM)
    g1 = mc.Context(m, mode=mc.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12, warm_start=warm)
    g2 = mc.Context(m, mode=mc.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12, warm_start=warm)
    push_state(o, g1)
    push_state(o, g2)
    for t in [po.FUEL, po.PROG, po.VARFU, po.VARPR]:
        g1.scalar_solve(t)
    g2.scalar_batch_solve()
    for f in [po.FUL, po.PRO, po.VARF, po.VARP]:
        assert_bits(g2.phi(f), g1.phi(f), "batched scalar %s" % po.FIELD_NAMES[f])
    for t, f in [(po.FUEL, po.FUL), (po.PROG, po.PRO), (po.VARFU, po.VARF), (po.VARPR, po.VARP)]:
        o.scalar_solve(t, f)
        assert_close(g2.phi(f), o.phi(f), "batched scalar vs oracle %s" % po.FIELD_NAMES[f], 1e-9)
    g1.close()
    g2.close()
    o.close()


LIMITED = [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]


def _rough(o, seed, rel=0.1):
    """Seeded noise on every limited field (cells and slots): its local extrema
    make the gradient limiter act in many cells."""
    rng = np.random.default_rng(seed)
    for f in LIMITED:
        ph = o.phi(f)
        ph[:] = ph + rng.standard_normal(ph.size) * (rel * np.abs(ph).max() + 1e-6)


def test_gradient_limiter_bits():
    """limit_phi_gradients (FlowSolver.inl:804-893) bit for bit: the C-ABI
    entry after the gradients, and the limiter fused into the gradient pass of
    a SIMPLE iteration with limit_grad = 1."""
    m = 5
    print("gradient limiter m=%d" % m, flush=True)
    o, g = make_pair(m)
    _rough(o, 11)
    push_state(o, g)
    o.get_phi_gradients()
    g.get_phi_gradients()
    unl = {f: o.grad(f).copy() for f in LIMITED}
    o.limit_phi_gradients()
    g.limit_phi_gradients()
    acted = 0
    for f in LIMITED:
        assert_bits(g.grad(f), o.grad(f), "limited grad %s" % po.FIELD_NAMES[f])
        acted += int(np.sum(np.any(o.grad(f)[:o.N] != unl[f][:o.N], axis=1)))
    assert acted > o.N
    g.close()
    o.close()
    # fused form, from a state one timestep in with 1% multiplicative noise
    # (the iteration's solves need positive TE / ED)
    o = po.Oracle(m)
    o.timestep()
    rng = np.random.default_rng(12)
    for f in LIMITED:
        ph = o.phi(f)
        ph[:] = ph * (1.0 + 1e-2 * rng.standard_normal(ph.size))
    g2 = mc.Context(m, limit_grad=1)
    push_state(o, g2)
    o.get_phi_gradients()
    unl = {f: o.grad(f).copy() for f in LIMITED}
    o.limit_phi_gradients()
    assert sum(int(np.sum(np.any(o.grad(f)[:o.N] != unl[f][:o.N], axis=1))) for f in LIMITED) > o.N
    g2.simple_iteration(1)
    # P and PP are re-differentiated later in the iteration; the other ten
    # keep the gradients of its first pass
    for f in [po.U, po.V, po.W, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]:
        assert_bits(g2.grad(f), o.grad(f), "fused limited grad %s" % po.FIELD_NAMES[f])
    o.close()
    g2.close()


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
def test_timesteps_limited(mode):
    """Whole timesteps with the limiter on both sides (FLOW_SOLVER_LIMIT_GRAD = 1)."""
    m = 8
    print("limited timesteps m=%d mode=%d" % (m, mode), flush=True)
    o = po.Oracle(m, mode=mode, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=100000, limit_grad=1)
    g = mc.Context(m, mode=mode, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=5000, limit_grad=1, warm_start=0)
    for ts in range(2):
        o.timestep()
        g.timestep()
        for f in LIMITED:
            assert_close(g.phi(f), o.phi(f), "limited ts %d field %s" % (ts, po.FIELD_NAMES[f]), 1e-8)
        assert_close(g.get(mc.ARR_MFLUX), o.face("mflux"), "limited ts %d mass flux" % ts, 1e-8)
    o.close()
    g.close()


def test_write_vtk_from_device(tmp_path):
    """output_data: mc_write_vtk downloads the fields and writes the same bytes
    as the host writer on those arrays (whose layout the reference's own
    writer pins in tests/test_io_cpu.py); file name as VisitWriter.hpp:131."""
    from tests.test_io_cpu import write_vtk
    m = 4
    print("vtk m=%d" % m, flush=True)
    with mc.Context(m) as g:
        g.timestep()
        path = g.write_vtk(str(tmp_path / "minicombust"), 3)
        assert path.endswith("minicombust_flow0_timestep3.vtk")
        ours = open(path, "rb").read()
        n = g.N
        fields = [g.phi(f)[:n] for f in [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF,
                                           po.VARP]]
        cc = g.get(mc.ARR_CELL_CENTER).reshape(-1, 3)
    assert ours == write_vtk(str(tmp_path / "host.vtk"), cc, fields)


def test_particle_coupling_bits():
    """Flow side of update_flow_field (FlowSolver.inl:496-700): particle terms
    added in receive order over two calls, and node values interpolated from a
    cell set, bit for bit against the oracle (which takes the set in ascending
    order, the order each node sums its cells in)."""
    import ctypes as C
    m = 4
    print("particle coupling m=%d" % m, flush=True)
    o, g = make_pair(m)
    o.timestep()
    push_state(o, g)
    rng = np.random.default_rng(21)
    up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint64))   # noqa: E731
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))   # noqa: E731
    pt = np.zeros((o.N, 5))
    g.particle_terms_reset()
    for call in range(2):
        n = 3000
        cells = rng.integers(0, o.N, size=n).astype(np.uint64)
        terms = rng.standard_normal((n, 5)) * 10.0 ** rng.integers(-6, 6, size=(n, 5))
        o.L.mco_particle_terms_add(o.sp, n, up(cells), dp(terms), dp(pt))
        g.particle_terms_add(cells, terms)
        assert_bits(g.get(mc.ARR_PARTICLE_TERMS).reshape(-1, 5), pt, "particle terms after call %d" % call)
    g.particle_terms_reset()
    assert not np.any(g.get(mc.ARR_PARTICLE_TERMS))
    # interpolation set: a random cluster of cells and their face neighbours
    M = o.mesh_arrays()
    seed_cells = rng.choice(o.N, size=12, replace=False)
    nb = M["cell_neighbours"][seed_cells].reshape(-1)
    cset = np.unique(np.concatenate([seed_cells, nb[nb < o.N]])).astype(np.uint64)
    nodes = np.unique(M["cells"][cset.astype(np.int64)].reshape(-1)).astype(np.uint64)
    nodes = np.concatenate([nodes, np.array([0, M["points"].shape[0] - 1], dtype=np.uint64)])
    out = np.zeros((nodes.size, 5))
    o.L.mco_interpolate_to_nodes(o.sp, cset.size, up(cset), nodes.size, up(nodes), dp(out))
    shuffled = rng.permutation(np.concatenate([cset, cset[:5]]))
    assert_bits(g.interpolate_to_nodes(shuffled, nodes), out, "node values")
    o.close()
    g.close()


@pytest.mark.parametrize("m", [4, 12])
def test_repair_nonneg_bits(m):
    """The TE / ED non-negativity repair (FlowSolver.inl:2956-2985) bit for bit:
    the reference's in-order sweep against the library's relaxation passes, on
    smooth fields with large negative regions (long chains of lower
    neighbours) plus noise; twice per field, so the second call starts from
    the first one's pass history."""
    print("repair m=%d" % m, flush=True)
    o, g = make_pair(m)
    M = o.mesh_arrays()
    cc = M["cell_centers"]
    rng = np.random.default_rng(30 + m)
    for field in [po.TE, po.ED]:
        for rep in range(2):
            k = rng.uniform(5.0, 20.0, size=(3, 3))
            ph = o.phi(field)
            ph[:o.N] = (np.cos(cc @ k[0]) * np.cos(cc @ k[1] + 1.0) + 0.3 * np.sin(cc @ k[2])
                        + 0.2 * rng.standard_normal(o.N))
            assert np.mean(ph[:o.N] < 0) > 0.2
            push_state(o, g)
            o.repair_nonneg(field)
            g.repair_nonneg(field)
            assert_bits(g.phi(field), o.phi(field), "repaired %s (call %d)" % (po.FIELD_NAMES[field], rep))
            assert np.all(o.phi(field)[:o.N] > 0)
    o.close()
    g.close()


def test_amg_coarse_levels_reused_between_refreshes():
    """amg_refresh = 4 keeps the coarse AMG levels of an earlier pressure matrix
    for three solves in four; the fine level always holds the current matrix,
    so every solve still meets its tolerance on the current system and the
    fields track those of a hierarchy refreshed at every solve."""
    print("amg refresh m=16", flush=True)
    runs = {}
    for r in (1, 4):
        with mc.Context(16, mode=mc.MODE_SIMPLE, p_rtol=1e-10, mom_rtol=1e-12, amg_refresh=r) as g:
            for it in range(9):
                g.simple_iteration(it % 3)
            st = g.stats()
            assert st["p_relres_max"] <= 1.05e-10, (r, st["p_relres_max"])
            runs[r] = (st["p_iters"], {f: g.phi(f).copy() for f in (po.U, po.V, po.W, po.P, po.TE, po.TEM)})
    it1, f1 = runs[1]
    it4, f4 = runs[4]
    assert it4 <= 1.2 * it1 + 2, (it1, it4)
    for f in f1:
        assert_close(f4[f], f1[f], "field %s, refresh 4 vs 1" % po.FIELD_NAMES[f], 1e-7)


@pytest.mark.parametrize("psolve", [mc.PSOLVE_FP64, mc.PSOLVE_MIXED])
def test_amg_tail_is_the_same_cycle(psolve):
    """The single-workgroup pass over the small AMG levels (amg_tail = 1)
    forms every row as the per-pass launches do (amg_tail = 0): the pressure
    solves agree bit for bit, iteration for iteration."""
    print("amg tail psolve=%d" % psolve, flush=True)
    m = 32   # 65,536 cells: levels of 8,192 and 1,024 rows and a dense 128 below the fine one,
             # the tail from the 1,024-row level
    o = po.Oracle(m, mode=po.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    oracle_pressure_assemble(o)
    b = o.cell("S_U").copy()
    out = {}
    for tail in (0, 1):
        with mc.Context(m, mode=mc.MODE_SIMPLE, psolve=psolve, p_rtol=1e-10, amg_tail=tail) as g:
            push_state(o, g)
            x, st = g.solve_pressure(b)
            assert st.converged
            out[tail] = (x.copy(), st.iterations)
    assert out[0][1] == out[1][1]
    assert_bits(out[1][0], out[0][0], "pressure correction, tail vs per-pass launches")
    o.close()


def test_cell_visiting_order_does_not_change_results():
    """The bricks and the order in which the cell-centric kernels visit them
    (cell_tile, cell_ychunk) change only which workgroup forms a cell's row:
    three SIMPLE iterations give the same bits for every field and flux."""
    print("visiting order m=12", flush=True)
    runs = []
    for tile, yc in ((0, 0), (64, 1), (4, 5), (256, 16), (1, 12)):
        with mc.Context(12, mode=mc.MODE_SIMPLE, cell_tile=tile, cell_ychunk=yc) as g:
            for it in range(3):
                g.simple_iteration(it)
            runs.append(({f: g.phi(f).copy() for f in range(po.NFIELDS)}, g.get(mc.ARR_MFLUX).copy()))
    for k, (fields, mflux) in enumerate(runs[1:], 1):
        for f in fields:
            assert_bits(fields[f], runs[0][0][f], "order %d field %s" % (k, po.FIELD_NAMES[f]))
        assert_bits(mflux, runs[0][1], "order %d mass flux" % k)


@pytest.mark.parametrize("m", [13, 16])
def test_scalar_batch_beside_ed_solve_same_bits(m):
    """The batched TEM / FUL / PRO / VARF / VARP assembly runs on a second
    stream beside the ED solve unless the BiCGStab kernel class is being
    timed: the two orders give the same bits for every field, the last
    system's matrix, diagonal and right-hand side, and the face arrays, over
    four SIMPLE iterations (m = 13: bricks cut by the mesh's edges)."""
    print("batch beside ED solve m=%d" % m, flush=True)
    runs = []
    for timed in (False, True):
        with mc.Context(m, mode=mc.MODE_SIMPLE) as g:
            if timed:
                g.profile(True, classes=["bicg_spmv"])
            for it in range(4):
                g.simple_iteration(it)
            arrays = {name: g.get(a).copy() for name, a in CELL_ARRAYS + FACE_ARRAYS}
            arrays["A_vals"] = g.get(mc.ARR_A_VALS).copy()
            runs.append(({f: g.phi(f).copy() for f in range(po.NFIELDS)}, arrays))
    for f in runs[0][0]:
        assert_bits(runs[0][0][f], runs[1][0][f], "beside vs in order: field %s" % po.FIELD_NAMES[f])
    for name in runs[0][1]:
        assert_bits(runs[0][1][name], runs[1][1][name], "beside vs in order: %s" % name)


@pytest.mark.parametrize("guess", ["current", "worse_than_zero"])
def test_warm_start_momentum_solve(guess):
    """warm_start = 1: the momentum BiCGStab starts from the current velocity
    and still meets ||b - A x|| <= rtol ||b|| on every component; a guess whose
    residual is not below ||b|| is dropped, and the solve is then the zero
    start's, bit for bit."""
    print("warm start guess=%s" % guess, flush=True)
    m, rtol = 16, 1e-6
    rng = np.random.default_rng(5)
    out = {}
    for warm in (0, 1):
        with mc.Context(m, mode=mc.MODE_SIMPLE, mom_rtol=rtol, warm_start=0) as g:
            for it in range(3):   # the same state for both starts
                g.simple_iteration(it)
            g.set_options(warm_start=warm)
            g.get_phi_gradients()
            g.momentum_assemble()
            N = g.N
            if guess == "worse_than_zero":
                for f in (po.U, po.V, po.W):
                    v = g.phi(f)
                    v[:N] = 1e6 * rng.standard_normal(N)
                    g.set(mc.ARR_PHI + f, v)
            bs = [g.get(a)[:N].copy() for a in (mc.ARR_SU, mc.ARR_SV, mc.ARR_SW)]
            st = g.momentum_solve()
            assert st.converged
            xs = [g.phi(f)[:N].copy() for f in (po.U, po.V, po.W)]
            for b, x in zip(bs, xs):
                res = np.linalg.norm(b - g.spmv(x, which=0)) / np.linalg.norm(b)
                assert res <= 1.05 * rtol, (warm, res)
            out[warm] = (xs, st.iterations)
    if guess == "worse_than_zero":
        assert out[1][1] == out[0][1]
        for k in range(3):
            assert_bits(out[1][0][k], out[0][0][k], "component %d, dropped guess vs zero start" % k)
    else:
        assert out[1][1] <= out[0][1], (out[1][1], out[0][1])


@pytest.mark.parametrize("eq", [po.TERBTE, po.TERBED])
def test_warm_start_te_solve_meets_criterion(eq):
    """The TE and ED solves' starting residual b - A x0 is formed inside the
    assembly kernel (no separate product), including the ED wall rows whose
    fixed value differs from the current field: the warm-started solve still
    meets ||b - A x|| <= rtol ||b|| on the assembled system, and agrees with
    the zero start to the accuracy both solves promise."""
    field = po.TE if eq == po.TERBTE else po.ED
    print("warm start %s" % po.FIELD_NAMES[field], flush=True)
    m, rtol = 16, 1e-7
    out = {}
    for warm in (0, 1):
        with mc.Context(m, mode=mc.MODE_SIMPLE, mom_rtol=rtol, warm_start=0) as g:
            for it in range(3):
                g.simple_iteration(it)
            g.set_options(warm_start=warm)
            g.get_phi_gradients()
            if eq == po.TERBED:
                g.scalar_solve(po.TERBTE)
            g.scalar_solve(eq)
            N = g.N
            b = g.get(mc.ARR_SU)[:N].copy()
            x = g.phi(field)[:N].copy()
            res = np.linalg.norm(b - g.spmv(x, which=0)) / np.linalg.norm(b)
            assert res <= 1.05 * rtol, (warm, res)
            out[warm] = x
    # the repair after the solve only lifts negative values, none here
    assert np.all(out[1] > 0)
    assert_close(out[1], out[0], "%s warm vs zero start" % po.FIELD_NAMES[field], 1e-5)


def test_fp16_smoothing_copy_keeps_the_tolerance():
    """amg_half = 1: the mixed form's fine-level smoothing reads an FP16 copy of
    D^-1 A.  Only the preconditioner changes: the solve still reaches 1e-8 on
    the assembled system (independent FP64 product) and agrees with the FP32
    smoothing's solution to the accuracy both reach."""
    print("fp16 smoothing m=32", flush=True)
    m = 32
    o = po.Oracle(m, mode=po.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    oracle_pressure_assemble(o)
    b = o.cell("S_U").copy()
    xs = {}
    for half in (0, 1):
        with mc.Context(m, mode=mc.MODE_SIMPLE, psolve=mc.PSOLVE_MIXED, p_rtol=1e-8, amg_half=half) as g:
            push_state(o, g)
            x, st = g.solve_pressure(b)
            assert st.converged
            bp = b - b.mean()
            res = np.linalg.norm(bp - g.spmv(x, which=1)) / np.linalg.norm(bp)
            assert res <= 1.05e-8, (half, res)
            xs[half] = x
    assert_close(xs[1], xs[0], "pressure correction, FP16 vs FP32 smoothing", 1e-6)
    o.close()


FIELDS_ALL = [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
def test_reference_cpu_case_m20_ten_timesteps(mode):
    """The reference's own CPU case: modifier 20 (16,000 cells), default
    boundary conditions and initial fields, 10 timesteps (30 SIMPLE
    iterations).  Every field and the mass flux agree with the oracle to 1e-8
    of the field's maximum after each timestep (both sides solve to 1e-12)."""
    print("m=20 x 10 timesteps mode=%d" % mode, flush=True)
    o, g = make_pair(20, mode)
    for ts in range(10):
        o.timestep()
        g.timestep()
        for f in FIELDS_ALL:
            assert_close(g.phi(f), o.phi(f), "timestep %d field %s" % (ts, po.FIELD_NAMES[f]), 1e-8)
        assert_close(g.get(mc.ARR_MFLUX), o.face("mflux"), "timestep %d mass flux" % ts, 1e-8)
    o.close()
    g.close()


def test_fp64_gpu_case_m64_against_oracle():
    """The FP64 GPU case's mesh, modifier 64 (524,288 cells), checked against
    the oracle's method on the same mesh for one timestep (3 SIMPLE
    iterations, simple mode): fields and mass flux to 1e-8."""
    print("m=64 x 3 iterations", flush=True)
    o, g = make_pair(64, po.MODE_SIMPLE)
    for it in range(3):
        o.simple_iteration(it)
        g.simple_iteration(it)
    for f in FIELDS_ALL:
        assert_close(g.phi(f), o.phi(f), "field %s" % po.FIELD_NAMES[f], 1e-8)
    assert_close(g.get(mc.ARR_MFLUX), o.face("mflux"), "mass flux", 1e-8)
    o.close()
    g.close()


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
def test_host_mesh_context(mode):
    """A context built from host mesh arrays (the reference's Mesh<double>
    layout, mc_ctx_create: consecutive-cell order, faces in reference order)
    against the oracle for one timestep, and against the generated-box context
    bit for bit (same sums in the same order, only the storage differs)."""
    print("host mesh m=6 mode=%d" % mode, flush=True)
    m = 6
    mesh = po.oracle_mesh(m)
    o = po.Oracle(m, mode=mode, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=100000)
    kw = dict(mode=mode, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=5000, warm_start=0)
    with mc.Context(mesh=mesh, **kw) as gh, mc.Context(m, **kw) as gb:
        assert gh.N == gb.N and gh.F == gb.F and gh.nnz == gb.nnz
        assert np.array_equal(gh.get(mc.ARR_FACES), gb.get(mc.ARR_FACES))
        o.timestep()
        gh.timestep()
        gb.timestep()
        for f in FIELDS_ALL:
            assert_close(gh.phi(f), o.phi(f), "host mesh field %s" % po.FIELD_NAMES[f], 1e-8)
            assert_bits(gh.phi(f), gb.phi(f), "host mesh vs box field %s" % po.FIELD_NAMES[f])
        assert_bits(gh.get(mc.ARR_MFLUX), gb.get(mc.ARR_MFLUX), "host mesh vs box mass flux")
    o.close()


def test_cpp_driver_matches_the_library(tmp_path):
    """The C++ driver with the reference's call surface (host FlowSolver:
    constructor, timestep(), output_data(t), print_logger_stats, compute_time)
    runs 2 timesteps at modifier 4 and writes the flow VTK file the ctypes
    binding writes after the same 2 timesteps, byte for byte."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(mc.LIB_PATH), "minicombust_amd")
    assert os.path.exists(exe), "C++ driver not built"
    env = dict(os.environ, MINICOMBUST_CELLS="4", MINICOMBUST_ITERS="2", MINICOMBUST_OUTPUT_ITER="2")
    print("driver m=4", flush=True)
    out = subprocess.run([exe], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Flow Compute Time" in out.stdout
    for title in ("Velocity Field Solve Timings", "Pressure Solve Timings", "Scalar (TEM, FUL, PRO, VARF, VARP)"):
        assert title in out.stdout, title
    drv = (tmp_path / "out" / "flow" / "minicombust_flow0_timestep2.vtk").read_bytes()
    with mc.Context(4) as g:   # the driver's defaults: compat mode, FP64 pressure solve
        g.timestep()
        g.timestep()
        ours = open(g.write_vtk(str(tmp_path / "lib"), 2), "rb").read()
    assert drv == ours


@pytest.mark.parametrize("mode", [po.MODE_COMPAT, po.MODE_SIMPLE])
def test_reference_load_mesh_into_the_dropin(tmp_path, mode):
    """The reference's own load_mesh output (Mesh<double>, built from the
    reference's source into oracle/_ref/ref_flow_driver) handed unchanged to
    the drop-in FlowSolver(&mpi_config, mesh, delta, stdout) under the name
    src/minicombust.cpp uses: after 2 timesteps at modifier 8 the fields and
    mass fluxes equal the generated-box context's bit for bit and the
    oracle's within 1e-8."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "ref_flow_driver")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/ref_flow_driver is built only where the reference sources are present")
    m, steps = 8, 2
    name = "simple" if mode == po.MODE_SIMPLE else "compat"
    out_bin = tmp_path / "fields.bin"
    print("reference load_mesh -> drop-in FlowSolver m=%d %s" % (m, name), flush=True)
    out = subprocess.run([exe, str(m), str(steps), name, str(out_bin)], cwd=str(tmp_path), capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Flow Compute Time" in out.stdout
    raw = out_bin.read_bytes()
    N, F = np.frombuffer(raw[:16], dtype=np.int64)
    vals = np.frombuffer(raw[16:], dtype=np.float64)
    assert vals.size == 11 * N + F
    kw = dict(mode=mode, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=5000, warm_start=0)
    o = po.Oracle(m, mode=mode, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=100000)
    with mc.Context(m, **kw) as g:
        assert g.N == N and g.F == F
        for _ in range(steps):
            g.timestep()
            o.timestep()
        for k, f in enumerate(FIELDS_ALL):
            drv = vals[k * N:(k + 1) * N]
            assert_bits(drv, g.phi(f)[:N], "drop-in vs box field %s" % po.FIELD_NAMES[f])
            assert_close(drv, o.phi(f)[:N], "drop-in vs oracle field %s" % po.FIELD_NAMES[f], 1e-8)
        assert_bits(vals[11 * N:], g.get(mc.ARR_MFLUX), "drop-in vs box mass flux")
    o.close()


def test_error_paths_leave_the_context_usable():
    """Refusals come back as status codes with a message and change nothing:
    a box modifier out of range, options mc_set_options refuses (the previous
    ones stay in force), array calls with a wrong size, an unknown id or a
    read-only array, particle terms read before any were added.  A momentum
    solve stopped by its iteration limit returns MC_ERR_NOCONV; after
    mc_reset_fields the context's next timestep matches a fresh context's."""
    print("error paths m=4", flush=True)
    with pytest.raises(mc.MCError, match="modifier out of range"):
        mc.Context(0)
    kw = dict(mode=po.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12, p_maxit=5000, warm_start=0)
    with mc.Context(4, **kw) as g, mc.Context(4, **kw) as fresh:
        L, h = g.L, g.h
        for bad in [dict(amg_nu=0), dict(warm_start=5), dict(cell_tile=3)]:
            with pytest.raises(mc.MCError, match="rc=%d" % mc.MC_ERR_ARG):
                g.set_options(**bad)
        assert g.opts.amg_nu == fresh.opts.amg_nu and g.opts.warm_start == 0
        buf = np.zeros(g.N + 6)
        assert L.mc_array_get(h, mc.ARR_PHI, buf.ctypes.data, g.N) == mc.MC_ERR_ARG
        assert "expected %d" % (g.N + 6) in L.mc_last_error(h).decode()
        assert L.mc_array_get(h, 9999, buf.ctypes.data, g.N) == mc.MC_ERR_ARG
        assert "unknown array id" in L.mc_last_error(h).decode()
        faces = g.get(mc.ARR_FACES)
        assert L.mc_array_set(h, mc.ARR_FACES, faces.ctypes.data, faces.size) == mc.MC_ERR_ARG
        assert "read-only" in L.mc_last_error(h).decode()
        terms = np.zeros(5 * g.N)
        assert L.mc_array_get(h, mc.ARR_PARTICLE_TERMS, terms.ctypes.data, terms.size) == mc.MC_ERR_STATE
        g.set_options(mom_maxit=1, mom_rtol=1e-30)
        g.get_phi_gradients()
        g.set_up_field()
        g.momentum_assemble()
        with pytest.raises(mc.MCError, match="rc=%d.*did not converge" % mc.MC_ERR_NOCONV):
            g.momentum_solve()
        g.set_options(mom_maxit=fresh.opts.mom_maxit, mom_rtol=fresh.opts.mom_rtol)
        g.reset_fields()
        g.timestep()
        fresh.timestep()
        for f in FIELDS_ALL:
            assert_close(g.phi(f), fresh.phi(f), "after the errors, field %s" % po.FIELD_NAMES[f], 1e-10)


def test_pressure_form_switched_between_iterations():
    """mc_set_options may change the pressure solve's form on a running
    context (the FP32 hierarchy copies exist in both forms): FP64, then
    mixed, then FP64 again over six SIMPLE iterations (m = 16, simple mode,
    1e-10) gives the fields of an all-FP64 run to 1e-8 of each field's
    maximum, every solve meeting its tolerance."""
    print("psolve switch m=16", flush=True)
    kw = dict(mode=po.MODE_SIMPLE, p_rtol=1e-10, mom_rtol=1e-12, warm_start=0)
    with mc.Context(16, psolve=mc.PSOLVE_FP64, **kw) as g, mc.Context(16, psolve=mc.PSOLVE_FP64, **kw) as ref:
        for it in range(6):
            g.set_options(psolve=mc.PSOLVE_MIXED if it in (2, 3) else mc.PSOLVE_FP64)
            g.simple_iteration(it % 3)
            ref.simple_iteration(it % 3)
        assert g.stats()["p_relres_max"] <= 1e-10 * 1.0001
        for f in FIELDS_ALL:
            assert_close(g.phi(f), ref.phi(f), "switched form, field %s" % po.FIELD_NAMES[f], 1e-8)


def test_amg_rebuild_releases_the_old_hierarchy():
    """Changing amg_coarse_max rebuilds the AMG hierarchy at the next solve;
    the old levels are freed first, so the context's AMG memory stays near
    one hierarchy's (aggregates follow the matrix values, so a rebuild at a
    later iteration may differ by a few aggregates; four rebuilds kept would
    be five times it), and every solve still meets its tolerance (m = 16,
    simple mode, mixed form)."""
    print("amg rebuild m=16", flush=True)
    with mc.Context(16, mode=po.MODE_SIMPLE, psolve=mc.PSOLVE_MIXED, p_rtol=1e-10, warm_start=0) as g:
        g.simple_iteration(0)
        amg0 = g.memory()["amg"]
        assert amg0 > 0
        for it, cmax in enumerate([32, 128, 32, 128]):
            g.set_options(amg_coarse_max=cmax)
            g.simple_iteration(1 + it % 2)
            assert g.memory()["amg"] < 1.5 * amg0
        assert g.stats()["p_relres_max"] <= 1e-10 * 1.0001


def test_amg_rebuild_between_assembly_and_solve():
    """mc_pressure_assemble, then mc_set_options with a new amg_coarse_max,
    then mc_pressure_solve: the rebuilt hierarchy's fine level is refreshed
    from the assembled matrix (the assembly's diagonal copies went to the
    released hierarchy), so the solve is the one of a context whose options
    changed before the assembly, bit for bit, and meets its tolerance."""
    print("amg rebuild after assembly m=16", flush=True)
    kw = dict(mode=po.MODE_SIMPLE, psolve=mc.PSOLVE_MIXED, p_rtol=1e-10, mom_rtol=1e-12, warm_start=0)
    out = {}
    for order in ("options_first", "assembly_first"):
        with mc.Context(16, **kw) as g:
            g.simple_iteration(0)
            if order == "options_first":
                g.set_options(amg_coarse_max=32)
            g.pressure_assemble()
            if order == "assembly_first":
                g.set_options(amg_coarse_max=32)
            st = g.pressure_solve()
            assert st.converged and st.relres <= 1e-10 * 1.0001, (order, st.relres)
            out[order] = (g.phi(po.PP).copy(), st.iterations)
    assert out["options_first"][1] == out["assembly_first"][1]
    assert_bits(out["assembly_first"][0], out["options_first"][0], "PP, rebuild after the assembly")


def test_fp16_copy_switched_on_after_the_hierarchy():
    """amg_half switched from 0 to 1 on a context whose hierarchy exists
    rebuilds it with the FP16 smoothing copy: the next pressure solve is the
    one of a context that had amg_half = 1 throughout and rebuilt at the same
    point (amg_coarse_max changed), bit for bit, on the same system."""
    print("amg_half 0 -> 1 m=16", flush=True)
    o = po.Oracle(16, mode=po.MODE_SIMPLE, mom_rtol=1e-12, p_rtol=1e-12)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    oracle_pressure_assemble(o)
    b = o.cell("S_U").copy()
    kw = dict(mode=po.MODE_SIMPLE, psolve=mc.PSOLVE_MIXED, p_rtol=1e-10)
    out = {}
    for case in ("half_later", "half_always"):
        first = dict(amg_half=0) if case == "half_later" else dict(amg_half=1, amg_coarse_max=64)
        with mc.Context(16, **dict(kw, **first)) as g:
            push_state(o, g)
            x, st = g.solve_pressure(b)
            assert st.converged
            g.set_options(amg_half=1, amg_coarse_max=128)
            x, st = g.solve_pressure(b)
            assert st.converged
            out[case] = (x, st.iterations)
    assert out["half_later"][1] == out["half_always"][1]
    assert_bits(out["half_later"][0], out["half_always"][0], "pressure correction, FP16 copy switched on later")
    o.close()
