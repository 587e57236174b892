"""CPU tests: pin the oracle (the C restatement in oracle/) against the
reference's own mesh generator, closed forms of the uniform box, exactness
properties of the least-squares gradient, invariants of the assembled
systems, an independent sparse solver, and the mass conservation the pressure
correction must deliver.  None of these needs a GPU.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MESH_KEYS = ["points", "cells", "cell_neighbours", "faces", "cell_faces", "cells_per_point", "cell_centers",
             "boundary_types"]


def same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


# ---------------------------------------------------------------- A1 mesh --
@pytest.mark.parametrize("m", [1, 2, 3])
def test_oracle_mesh_equals_reference_golden(m):
    """Box mesh of the oracle vs the reference generator's output (fixture)."""
    ref = np.load(os.path.join(GOLDEN, "ref_mesh_m%d.npz" % m))
    ora = po.oracle_mesh(m)
    for k in MESH_KEYS:
        assert same_bits(ora[k], ref[k]), "mesh array %s differs at m=%d" % (k, m)


@pytest.mark.skipif(not os.path.exists(po.REF_MESH_DUMP), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("m", [4, 5, 8, 20, 64])
def test_oracle_mesh_equals_reference_generator(m):
    """Same, at larger sizes, running the reference's compiled generator."""
    with tempfile.TemporaryDirectory() as d:
        ref = po.ref_mesh(m, d)
    ora = po.oracle_mesh(m)
    for k in MESH_KEYS:
        assert same_bits(ora[k], ref[k]), "mesh array %s differs at m=%d" % (k, m)


@pytest.mark.parametrize("m", [2, 5])
def test_mesh_counts_and_face_numbering(m):
    """Cell/face counts of the 2m x m x m box and the closed-form face count
    before each cell (faces numbered by first visiting cell, side order)."""
    M = po.oracle_mesh(m)
    N, F = 2 * m ** 3, 6 * m ** 3 + 5 * m ** 2
    assert M["cells"].shape[0] == N and M["faces"].shape[0] == F
    faces = M["faces"].astype(np.int64)
    assert np.all(faces[:, 0] < faces[:, 1])
    assert np.sum(faces[:, 1] >= N) == 10 * m * m
    nx, nxny = 2 * m, 2 * m * m
    cf = M["cell_faces"].astype(np.int64)
    for c in range(N):
        z = c // nxny
        before = 3 * c + min(c, nxny) + -(-c // nx) + z * nx + min(c % nxny, nx)
        assert cf[c].max() >= before  # first new face of cell c
        assert np.sort(cf[c])[-1] < F
    # every interior face appears in exactly two cells, every boundary face in one
    counts = np.bincount(cf.ravel(), minlength=F)
    assert np.all(counts[faces[:, 1] < N] == 2) and np.all(counts[faces[:, 1] >= N] == 1)


# ------------------------------------------------------------ A2 geometry --
@pytest.mark.parametrize("m", [2, 5])
def test_geometry_closed_forms(m):
    """Uniform box of edge 0.05: areas, normals, volumes, lambdas, rlencos and
    centres have closed forms (FlowSolver.hpp:252-397, Mesh.hpp:54-67)."""
    o = po.Oracle(m)
    N, F = o.N, o.F
    M = o.mesh_arrays()
    faces = M["faces"].astype(np.int64)
    interior = faces[:, 1] < N
    h = 0.05
    np.testing.assert_allclose(o.face("face_areas"), h * h, rtol=1e-13)
    n = o.facevec("face_normals")
    assert np.all(np.count_nonzero(np.abs(n) > 1e-18, axis=1) == 1)  # axis aligned
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), h * h, rtol=1e-13)
    np.testing.assert_allclose(o.cell("cell_volumes"), h ** 3, rtol=1e-12)
    np.testing.assert_allclose(o.cell("cell_densities"), 1.2, rtol=0)
    lam = o.face("face_lambdas")
    np.testing.assert_allclose(lam[interior], 0.5, rtol=1e-12)
    np.testing.assert_allclose(lam[~interior], 1.0, rtol=0)
    rl = o.face("face_rlencos")
    np.testing.assert_allclose(rl[interior], h, rtol=1e-12)
    np.testing.assert_allclose(rl[~interior], 2 * h, rtol=1e-12)
    # cell centres: ((i + 1/2) h, (j + 1/2) h, (k + 1/2) h), x fastest
    nx, ny = 2 * m, m
    c = np.arange(N)
    ijk = np.stack([c % nx, (c // nx) % ny, c // (nx * ny)], axis=1)
    np.testing.assert_allclose(M["cell_centers"], (ijk + 0.5) * h, rtol=1e-12, atol=1e-15)
    # normals point from cell0 to cell1 (interior) and outward (boundary)
    fc = o.facevec("face_centers")
    cc = M["cell_centers"]
    assert np.all(np.einsum("ij,ij->i", n, fc - cc[faces[:, 0]]) > 0)


# ---------------------------------------------------------- A4/A5 gradient --
def _interior_cells(M, N):
    nb = M["cell_neighbours"]
    return np.all(nb < N, axis=1)


@pytest.mark.parametrize("field", [po.U, po.TEM])
def test_gradient_exact_on_linear_fields(field):
    """LSQ reproduces the gradient of a linear field exactly in the interior."""
    o = po.Oracle(5)
    M = o.mesh_arrays()
    cc = M["cell_centers"]
    a = np.array([3.25, -1.5, 0.75])
    phi = o.phi(field)
    phi[:o.N] = cc @ a + 2.0
    o.get_phi_gradients()
    g = o.grad(field)[:o.N]
    inner = _interior_cells(M, o.N)
    assert inner.sum() > 0
    np.testing.assert_allclose(g[inner], np.broadcast_to(a, g[inner].shape), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("pressure", [0, 1])
def test_gradient_matches_lstsq(pressure):
    """Every cell's gradient equals numpy's least-squares solution over its six
    faces: dX = x_nb - x_c (boundary: face centre - x_c), dphi = phi_nb - phi_c
    (boundary: side slot - phi_c, or 0 for the pressure form)."""
    m = 3
    o = po.Oracle(m)
    N = o.N
    M = o.mesh_arrays()
    rng = np.random.default_rng(7)
    field = po.P if pressure else po.U
    phi = o.phi(field)
    phi[:] = rng.standard_normal(N + 6)
    o.get_phi_gradient(field, pressure)
    g = o.grad(field)[:N]
    cc = M["cell_centers"]
    fcen = o.facevec("face_centers")
    faces = M["faces"].astype(np.int64)
    for c in range(N):
        rows, rhs = [], []
        for f in M["cell_faces"][c].astype(np.int64):
            c0, c1 = faces[f]
            if c1 < N:
                other = c1 if c0 == c else c0
                rows.append(cc[other] - cc[c])
                rhs.append(phi[other] - phi[c])
            else:
                rows.append(fcen[f] - cc[c])
                rhs.append(0.0 if pressure else phi[c1] - phi[c])
        sol = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
        np.testing.assert_allclose(g[c], sol, rtol=1e-9, atol=1e-9 * np.abs(sol).max())


# ------------------------------------------------------- A9/A12 assembly --
def _state_after_momentum(m, mode=po.MODE_SIMPLE):
    o = po.Oracle(m, mode=mode, p_solver=1)
    o.get_phi_gradients()
    o.set_up_field()
    o.calculate_UVW()
    return o


def _pressure_assemble(o):
    """The entry of calculate_pressure (FlowSolver.inl:2333-2360): A_phi.V,
    S_phi.U, mass fluxes and face coefficients zeroed, then the Rhie-Chow
    mass flux and the pressure-correction system."""
    o.cell("A_V")[:] = 0.0
    o.cell("S_U")[:] = 0.0
    o.face("mflux")[:] = 0.0
    o.face("ff0")[:] = 0.0
    o.face("ff1")[:] = 0.0
    o.calculate_mass_flux()
    o.setup_pressure_matrix()


def _csr(o, which):
    rp, col = o.csr()
    return sp.csr_matrix((o.vals(which).copy(), col, rp), shape=(o.N, o.N))


def test_pressure_matrix_invariants():
    """Pressure-correction matrix: symmetric, off-diagonals <= 0, zero row
    sums (pure Neumann), and a right-hand side whose sum is the net boundary
    mass flux (FlowSolver.inl:2106-2201)."""
    o = _state_after_momentum(5)
    _pressure_assemble(o)
    A = _csr(o, "PA_vals")
    assert (A != A.T).nnz == 0
    off = A - sp.diags(A.diagonal())
    assert off.data.max() <= 0.0
    rows = np.asarray(A.sum(axis=1)).ravel()
    assert np.abs(rows).max() <= 1e-12 * A.diagonal().max()
    assert np.all(A.diagonal() > 0)
    # RHS = -(net outflow) per cell; its sum is minus the net boundary outflow
    M = o.mesh_arrays()
    faces = M["faces"].astype(np.int64)
    mf = o.face("mflux")
    bnd = faces[:, 1] >= o.N
    np.testing.assert_allclose(o.vec("Pb").sum(), -mf[bnd].sum(), rtol=1e-9, atol=1e-15)


def test_momentum_matrix_properties():
    """Momentum matrix after under-relaxation (URF 0.5): positive diagonal,
    strict row dominance, off-diagonal entries equal to the face coefficients
    ff.cell1 (row cell0) and ff.cell0 (row cell1) (FlowSolver.inl:1186-1297)."""
    o = _state_after_momentum(4)
    A = _csr(o, "A_vals")
    d = A.diagonal()
    off = abs(A - sp.diags(d))
    assert np.all(d > 0)
    assert np.all(d > np.asarray(off.sum(axis=1)).ravel())
    M = o.mesh_arrays()
    faces = M["faces"].astype(np.int64)
    ff0, ff1 = o.face("ff0"), o.face("ff1")
    inter = np.nonzero(faces[:, 1] < o.N)[0]
    Ad = A.todok()
    for f in inter[:200]:
        c0, c1 = faces[f]
        assert Ad[c0, c1] == ff1[f] and Ad[c1, c0] == ff0[f]


# -------------------------------------------------------- A10/A13 solvers --
def _ptr(a, t=np.float64):
    import ctypes as C
    return a.ctypes.data_as(C.POINTER(C.c_double if t == np.float64 else C.c_int64))


def test_bicgstab_against_direct_solve():
    """Oracle BiCGStab (no preconditioner, x0 = 0) on the momentum system:
    residual within rtol, solution close to scipy's direct solve."""
    import ctypes as C
    o = _state_after_momentum(4)
    rp, col = o.csr()
    vals = o.vals("A_vals").copy()
    b = o.cell("S_U").copy()
    x = np.zeros(o.N)
    rel = C.c_double(0.0)
    its = po.lib().mco_bicgstab(o.N, _ptr(rp, np.int64), _ptr(col, np.int64), _ptr(vals), _ptr(b), _ptr(x),
                                1e-10, 1000, C.byref(rel))
    A = sp.csr_matrix((vals, col, rp), shape=(o.N, o.N))
    assert its > 0
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b) * 1.0001
    xd = spla.spsolve(A.tocsc(), b)
    np.testing.assert_allclose(x, xd, rtol=1e-8, atol=1e-12 * np.abs(xd).max())


def test_bicgstab_warm_start():
    """The oracle's warm-started BiCGStab (the library's simple-mode option):
    from a perturbed solution it meets ||b - A x|| <= rtol ||b|| in fewer
    iterations than from zero; from the exact solution it does no iteration
    and returns the guess; from a guess worse than zero it is the zero start,
    bit for bit (the guess is dropped)."""
    import ctypes as C
    o = _state_after_momentum(4)
    rp, col = o.csr()
    vals = o.vals("A_vals").copy()
    b = o.cell("S_U").copy()
    A = sp.csr_matrix((vals, col, rp), shape=(o.N, o.N))
    xd = spla.spsolve(A.tocsc(), b)
    L = po.lib()

    def solve(x0, warm, rtol=1e-8):
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        rel = C.c_double(0.0)
        its = L.mco_bicgstab_x0(o.N, _ptr(rp, np.int64), _ptr(col, np.int64), _ptr(vals), _ptr(b), _ptr(x),
                                rtol, 1000, C.byref(rel), warm)
        return x, its

    rng = np.random.default_rng(4)
    x_zero, it_zero = solve(np.zeros(o.N), 0)
    x_warm, it_warm = solve(xd * (1.0 + 1e-4 * rng.standard_normal(o.N)), 1)
    assert np.linalg.norm(b - A @ x_warm) <= 1e-8 * np.linalg.norm(b) * 1.0001
    assert 0 < it_warm < it_zero, (it_warm, it_zero)
    x_ex, it_ex = solve(xd, 1, rtol=1e-6)
    assert it_ex == 0 and np.array_equal(x_ex, xd)
    x_bad, it_bad = solve(1e6 * rng.standard_normal(o.N), 1)
    assert it_bad == it_zero and same_bits(x_bad, x_zero)


@pytest.mark.parametrize("solver", ["pcg_singular", "pcg_amg"])
def test_pressure_solvers_against_direct_solve(solver):
    """Oracle PCG (Jacobi checker and aggregation AMG) on the singular
    pressure system: mean-zero solution equal to the direct solution of the
    system with the null space removed."""
    import ctypes as C
    o = _state_after_momentum(5)
    _pressure_assemble(o)
    rp, col = o.csr()
    vals = o.vals("PA_vals").copy()
    b = o.vec("Pb").copy()
    bp = b - b.mean()
    x = np.zeros(o.N)
    rel, bn = C.c_double(0.0), C.c_double(0.0)
    fn = getattr(po.lib(), "mco_" + solver)
    fn(o.N, _ptr(rp, np.int64), _ptr(col, np.int64), _ptr(vals), _ptr(b), _ptr(x), 1e-10, 0.0, 5000,
       C.byref(rel), C.byref(bn))
    A = sp.csr_matrix((vals, col, rp), shape=(o.N, o.N))
    ones = np.ones(o.N) / np.sqrt(o.N)
    K = (A + sp.csr_matrix(np.outer(ones, ones) * A.diagonal().mean())).tocsc()
    xd = spla.spsolve(K, bp)
    xd -= xd.mean()
    xm = x - x.mean()
    assert np.linalg.norm(bp - A @ x) <= 1.01e-10 * np.linalg.norm(bp)
    np.testing.assert_allclose(xm, xd, rtol=1e-6, atol=1e-7 * np.abs(xd).max())


# -------------------------------------------------- A11-A15 conservation --
def _divergence(M, N, mf):
    faces = M["faces"].astype(np.int64)
    d = np.zeros(N)
    np.add.at(d, faces[:, 0], mf)
    inter = faces[:, 1] < N
    np.add.at(d, faces[inter, 1], -mf[inter])
    return d


def test_pressure_correction_conserves_mass():
    """Simple mode: after calculate_pressure the face mass fluxes are
    divergence free per cell to the pressure-solve tolerance, starting from
    the Rhie-Chow fluxes' imbalance (a sign or index error in the mass flux,
    the pressure matrix or the corrections breaks this)."""
    m = 6
    before_o = _state_after_momentum(m)
    _pressure_assemble(before_o)
    M = before_o.mesh_arrays()
    before = np.linalg.norm(_divergence(M, before_o.N, before_o.face("mflux").copy()))
    o = _state_after_momentum(m)
    o.calculate_pressure()
    after = np.linalg.norm(_divergence(M, o.N, o.face("mflux").copy()))
    assert before > 1e-6
    assert after <= 1e-7 * before, (before, after)


def test_timestep_fields_finite_and_bounded():
    """Three reference timesteps at m = 4: every field finite, TE/ED >= 0
    after the non-negativity repair (FlowSolver.inl:2960-3076)."""
    o = po.Oracle(4, mode=po.MODE_SIMPLE, p_solver=1)
    for _ in range(3):
        o.timestep()
    for f in [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]:
        assert np.all(np.isfinite(o.phi(f)[:o.N])), po.FIELD_NAMES[f]
    assert o.phi(po.TE)[:o.N].min() >= 0 and o.phi(po.ED)[:o.N].min() >= 0


LIMITED = [po.U, po.V, po.W, po.P, po.TE, po.ED, po.TEM, po.FUL, po.PRO, po.VARF, po.VARP]


def _face_values(o, M, phi):
    """phi across each of the six faces of every cell (boundary: side slot)."""
    N = o.N
    faces = M["faces"].astype(np.int64)
    cf = M["cell_faces"].astype(np.int64)
    c0, c1 = faces[cf, 0], faces[cf, 1]
    cells = np.arange(N)[:, None]
    other = np.where(c0 == cells, c1, c0)      # boundary faces: c1 = N + side
    return phi[other]


def test_limiter_keeps_face_values_within_neighbour_bounds():
    """Venkatakrishnan's limiter (FlowSolver.inl:804-873): every limited
    gradient is alpha times the unlimited one with 0 <= alpha <= 1, and the
    limited reconstruction phi_c + g . (x_f - x_c) stays within the min / max of
    the cell and its neighbours at every face, up to the 6e-6 threshold below
    which a face does not limit (f(r) <= r for r >= 0 gives the bound)."""
    m = 4
    o = po.Oracle(m)
    M = o.mesh_arrays()
    N = o.N
    rng = np.random.default_rng(3)
    for f in LIMITED:
        ph = o.phi(f)
        ph[:] = ph + rng.standard_normal(ph.size) * (0.1 * np.abs(ph).max() + 1e-3)
    o.get_phi_gradients()
    unl = {f: o.grad(f)[:N].copy() for f in LIMITED}
    o.limit_phi_gradients()
    fc = o.facevec("face_centers")
    ds = fc[M["cell_faces"].astype(np.int64)] - M["cell_centers"][:, None, :]   # (N, 6, 3)
    limited = 0
    for f in LIMITED:
        ph = o.phi(f)
        g, g0 = o.grad(f)[:N], unl[f]
        nbv = _face_values(o, M, ph)
        pmax = np.maximum(ph[:N], nbv.max(axis=1))
        pmin = np.minimum(ph[:N], nbv.min(axis=1))
        # g = alpha g0, one alpha per cell
        j = np.argmax(np.abs(g0), axis=1)
        big = np.abs(g0[np.arange(N), j]) > 0
        alpha = np.where(big, g[np.arange(N), j] / np.where(big, g0[np.arange(N), j], 1.0), 1.0)
        assert np.all(alpha >= -1e-15) and np.all(alpha <= 1.0 + 1e-15)
        assert np.allclose(g, alpha[:, None] * g0, rtol=1e-14, atol=1e-300)
        rec = ph[:N, None] + np.einsum("nk,nsk->ns", g, ds)
        tol = 6e-6 + 1e-12 * np.abs(ph).max()
        assert np.all(rec <= pmax[:, None] + tol) and np.all(rec >= pmin[:, None] - tol)
        limited += int(np.sum(alpha < 1.0 - 1e-12))
    assert limited > N   # the rough fields are limited in many cells


def test_limiter_leaves_linear_fields_and_flattens_extrema():
    """On a linear field the interior gradient is kept (r = 2 on every face,
    f(2) = 1); at a cell raised above all its neighbours deltamax = 0, so any
    face with g . ds > 0 gives r = 0 and the gradient is set to zero."""
    o = po.Oracle(5)
    M = o.mesh_arrays()
    N = o.N
    cc = M["cell_centers"]
    a = np.array([3.25, -1.5, 0.75])
    inner = _interior_cells(M, N)
    spike = int(np.nonzero(inner)[0][len(np.nonzero(inner)[0]) // 2])
    ph = o.phi(po.U)
    ph[:N] = cc @ a + 2.0
    o.get_phi_gradients()
    g0 = o.grad(po.U)[:N].copy()
    o.limit_phi_gradients()
    g = o.grad(po.U)[:N]
    np.testing.assert_allclose(g[inner], g0[inner], rtol=1e-12, atol=0)
    ph[spike] += 10.0
    o.get_phi_gradients()
    assert np.abs(o.grad(po.U)[spike]).max() > 1.0
    o.limit_phi_gradients()
    assert np.all(o.grad(po.U)[spike] == 0.0)


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_particle_terms_add_in_receive_order():
    """FlowSolver.inl:675-683: entries add to their cell in receive order,
    which numpy's unbuffered np.add.at reproduces addition for addition."""
    o = po.Oracle(3)
    rng = np.random.default_rng(4)
    n = 500
    cells = rng.integers(0, o.N, size=n).astype(np.uint64)
    terms = rng.standard_normal((n, 5)) * 10.0 ** rng.integers(-6, 6, size=(n, 5))
    pt = np.zeros((o.N, 5))
    o.L.mco_particle_terms_add(o.sp, n, _u64p(cells), _dp(terms), _dp(pt))
    ref = np.zeros((o.N, 5))
    np.add.at(ref, cells.astype(np.int64), terms)
    assert np.array_equal(pt.view(np.uint64), ref.view(np.uint64))


def test_node_interpolation_closed_forms():
    """interpolate_to_nodes (FlowSolver.inl:297-335, 428-493): with exact
    gradients of linear fields every cell contributes phi(x_node) / 8, so a node
    whose cells are all in the set gets phi(x_node) plus the dummy gas state
    (Mesh.hpp:114-116) times its missing cells / 8; a node outside the set
    keeps only that dummy term."""
    o = po.Oracle(3)
    M = o.mesh_arrays()
    N = o.N
    cc, pts = M["cell_centers"], M["points"]
    cpp = M["cells_per_point"].astype(np.float64)
    rng = np.random.default_rng(6)
    coef = {}
    for f in [po.U, po.V, po.W, po.P, po.TEM]:
        a = rng.standard_normal(3)
        b = rng.standard_normal()
        coef[f] = (a, b)
        o.phi(f)[:N] = cc @ a + b
        o.grad(f)[:N] = a
    cells = np.arange(N, dtype=np.uint64)[::-1].copy()       # every cell, reversed order
    nodes = np.arange(pts.shape[0], dtype=np.uint64)
    out = np.zeros((nodes.size, 5))
    o.L.mco_interpolate_to_nodes(o.sp, cells.size, _u64p(cells), nodes.size, _u64p(nodes), _dp(out))
    dummy = np.array([50.0, 0.0, 0.0, 4000.0, 2000.0])
    for q, f in enumerate([po.U, po.V, po.W, po.P, po.TEM]):
        a, b = coef[f]
        expect = dummy[q] * ((8.0 - cpp) / 8.0) + cpp / 8.0 * (pts @ a + b)
        np.testing.assert_allclose(out[:, q], expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    # only the first cell in the set: its 8 nodes get 1/8 of phi, all others the dummy term
    one = np.array([5], dtype=np.uint64)
    out1 = np.zeros((nodes.size, 5))
    o.L.mco_interpolate_to_nodes(o.sp, 1, _u64p(one), nodes.size, _u64p(nodes), _dp(out1))
    mine = set(int(v) for v in M["cells"][5])
    for k in range(nodes.size):
        base = dummy * ((8.0 - cpp[k]) / 8.0)
        if k in mine:
            phi_n = np.array([pts[k] @ coef[f][0] + coef[f][1] for f in [po.U, po.V, po.W, po.P, po.TEM]])
            np.testing.assert_allclose(out1[k], base + phi_n / 8.0, rtol=1e-12, atol=1e-12)
        else:
            assert np.array_equal(out1[k], base)


def _interior_face_pairs(o, A):
    M = o.mesh_arrays()
    faces = M["faces"].astype(np.int64)
    inter = np.nonzero(faces[:, 1] < o.N)[0]
    c0, c1 = faces[inter, 0], faces[inter, 1]
    Ad = A.tocsr()
    a = np.asarray(Ad[c0, c1]).ravel()   # row cell0, column cell1
    b = np.asarray(Ad[c1, c0]).ravel()
    return inter, c0, c1, a, b


@pytest.mark.parametrize("which", ["momentum", "TEM", "FUL", "VARP"])
def test_face_coefficients_are_upwind_convection_plus_diffusion(which):
    """Momentum and scalar matrices (A7, A9, A16): on every interior face the
    antisymmetric part of the coefficient pair is the face's mass flux
    (first-order upwind convection, cell0 -> cell1 positive), the rest is a
    diffusion coefficient that is positive and, divided by the face's
    rlencos, the same on every face (one effective viscosity per equation);
    and every cell without a boundary face satisfies
    A_c * URF + sum of its off-diagonals = rho V / dt (the unsteady term
    alone, before under-relaxation).  A swapped owner / neighbour, a flipped
    flux sign or a lost term breaks one of these."""
    o = _state_after_momentum(4)
    urf = 0.5
    if which != "momentum":
        o.calculate_pressure()
        t, f = {"TEM": (po.TEMP, po.TEM), "FUL": (po.FUEL, po.FUL), "VARP": (po.VARPR, po.VARP)}[which]
        o.scalar_assemble(t, f)
        urf = 0.95
    A = _csr(o, "A_vals")
    mflux = o.face("mflux")
    inter, c0, c1, a, b = _interior_face_pairs(o, A)
    mf = mflux[inter]
    scale = np.maximum(np.abs(a), np.abs(mf))
    assert np.all(np.abs((a - b) - mf) <= 8 * np.finfo(float).eps * scale)
    D = -(a + b) / 2 - np.abs(mf) / 2
    assert np.all(D > 0)
    visc = D / o.face("face_rlencos")[inter]
    assert np.ptp(visc) <= 1e-10 * np.abs(visc).max()
    # cells with six interior faces
    inner = np.all(o.mesh_arrays()["cell_neighbours"] < o.N, axis=1)
    d = A.diagonal()
    offsum = np.asarray((A - sp.diags(d)).sum(axis=1)).ravel()
    unsteady = o.cell("cell_densities") * o.cell("cell_volumes") / o.opts.delta
    lhs = d * urf + offsum
    np.testing.assert_allclose(lhs[inner], unsteady[inner], rtol=1e-11, atol=0)


@pytest.mark.parametrize("which", ["TEM", "FUL", "VARF"])
def test_deferred_correction_vanishes_on_the_orthogonal_box(which):
    """Scalar source assembly (A16): on the box every face normal is parallel
    to the centre-to-centre vector, so each interior face's explicit diffusion
    (Visac grad . n) equals its implicit counterpart (VisFace grad . Xpn) and
    the deferred correction cancels for any gradient.  A cell away from the
    boundary therefore carries only its cell terms, rho V / dt phi and the
    under-relaxation (1 - URF) A_c phi.  Checked on rough random fields, so
    the gradients are arbitrary; a flipped sign or swapped operand in the face
    source terms shows at the scale of the diffusion flux.  (The momentum
    sources also hold the transpose-gradient stress, which does not cancel.)"""
    rng = np.random.default_rng(12)
    t, f = {"TEM": (po.TEMP, po.TEM), "FUL": (po.FUEL, po.FUL), "VARF": (po.VARFU, po.VARF)}[which]
    o = _state_after_momentum(4)
    o.calculate_pressure()
    ph = o.phi(f)
    ph[:] = rng.uniform(0.0, 1.0, size=ph.size) * (1000.0 if which == "TEM" else 1.0)
    o.get_phi_gradients()
    o.scalar_assemble(t, f)
    M = o.mesh_arrays()
    inner = np.all(M["cell_neighbours"] < o.N, axis=1)
    unsteady = o.cell("cell_densities") * o.cell("cell_volumes") / o.opts.delta
    A = o.cell("A_V")
    S = o.cell("S_U")
    phi = ph[:o.N]
    expect = unsteady * phi + (1.0 - 0.95) * A * phi
    np.testing.assert_allclose(S[inner], expect[inner], rtol=1e-12, atol=0)
    assert np.abs(o.grad(f)[:o.N]).max() > 1.0   # the gradients are far from zero
