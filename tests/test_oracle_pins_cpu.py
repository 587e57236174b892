"""CPU pins of the oracle's physics against what the mathematics and the
reference's constants fix: closed forms on the uniform box (edge h = 0.05,
face area A = h^2, volume V = h^3, rho = 1.2, mu = 1.4e-5, dt = 1e-5), with
chosen fields and gradients written straight into the oracle's state, so that
each stage is checked on inputs whose answer is known without the oracle.
A plausible slip in the restatement (a dropped term, a wrong sign, index,
factor or axis) changes at least one of these numbers.

Covered: A3 initial fields and boundary slots, the CSR pattern, A6
precomp_AU, A7 deferred-correction stress sources and wall / inlet / outlet
rows, A8 cell sources and under-relaxation, A11 / A12 Rhie-Chow flux and
pressure coefficient, A14 correction step (compat's cell-0 rule included),
A15 boundary pressure slots, the non-orthogonal flux update, A16 k-epsilon
production, dissipation and wall functions.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

H = 0.05
AREA = H * H
VOL = H ** 3
RHO = 1.2
MU = 1.4e-5
RDT = RHO * VOL / 1e-5            # rho V / dt = 15
WALL, INLET, OUTLET = 1, 2, 3
# outward unit normal of each box side: FRONT(z-), BACK(z+), LEFT(x-), RIGHT(x+), DOWN(y-), UP(y+)
SIDE_NORMAL = np.array([[0, 0, -1], [0, 0, 1], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0]], dtype=np.float64)
TE_IN = 37.5                      # 3/2 (0.1 |U_in|)^2 with |U_in| = 50
CMU75 = 0.09 ** 0.75
ED_IN = CMU75 * TE_IN ** 1.5
EV_IN = MU + RHO * 0.09 ** 0.25 / np.sqrt(TE_IN)   # mu + rho C_mu k^2 / eps at the inlet


class Box:
    """Mesh facts of an oracle state, derived from its arrays."""

    def __init__(self, o):
        M = o.mesh_arrays()
        self.o = o
        self.N = o.N
        self.faces = M["faces"].astype(np.int64)
        self.cf = M["cell_faces"].astype(np.int64)
        self.nb = M["cell_neighbours"]
        self.cc = M["cell_centers"]
        self.bt = M["boundary_types"].astype(np.int64)
        self.bnd = self.faces[:, 1] >= self.N
        self.inter = ~self.bnd
        self.side = np.where(self.bnd, self.faces[:, 1] - self.N, -1)
        nbound = np.sum(self.nb >= np.uint64(self.N), axis=1)
        self.interior_cells = np.nonzero(nbound == 0)[0]
        # the highest face id of each side and its cell: the writer of the side's slot
        self.last_face = {s: int(np.max(np.nonzero(self.side == s)[0])) for s in range(6)}

    def boundary_faces_of(self, c):
        return [f for f in self.cf[c] if self.bnd[f]]


def _fresh(m, **kw):
    return po.Oracle(m, mode=kw.pop("mode", po.MODE_SIMPLE), **kw)


# ------------------------------------------------------------------ A3 --
def test_initial_fields_and_slots_are_the_reference_constants():
    """Mesh.hpp:114-120 (dummy gas state) and FlowSolver.hpp:403-485
    (boundary slots by type; CPU wall temperature 2000): the values the
    reference text gives, computed by hand here."""
    o = _fresh(3, mode=po.MODE_COMPAT)
    b = Box(o)
    N = o.N
    cells = {po.U: 50.0, po.V: 0.0, po.W: 0.0, po.P: 4000.0, po.PP: 0.0, po.TE: 1e-4, po.ED: 0.0, po.TP: 0.0,
             po.TEM: 2000.0, po.FUL: 0.1, po.PRO: 0.001, po.VARF: 0.1, po.VARP: 0.001}
    for f, v in cells.items():
        assert np.all(o.phi(f)[:N] == v), po.FIELD_NAMES[f]
    slots = {
        INLET: {po.U: 50.0, po.V: 0.0, po.W: 0.0, po.P: 4000.0, po.PP: 0.0, po.TE: 37.5, po.ED: ED_IN,
                po.TP: 0.0, po.TEM: 2000.0, po.FUL: 0.1, po.PRO: 0.0, po.VARF: 0.1, po.VARP: 0.0},
        OUTLET: {po.U: 50.0, po.V: 0.0, po.W: 0.0, po.P: 4000.0, po.PP: 0.0, po.TE: 1e-4, po.ED: 0.0,
                 po.TP: 0.0, po.TEM: 2000.0, po.FUL: 0.1, po.PRO: 0.001, po.VARF: 0.1, po.VARP: 0.001},
        WALL: {po.U: 0.0, po.V: 0.0, po.W: 0.0, po.P: 4000.0, po.PP: 0.0, po.TE: 1e-4, po.ED: 0.0,
               po.TP: 0.0, po.TEM: 2000.0, po.FUL: 0.0, po.PRO: 0.0, po.VARF: 0.0, po.VARP: 0.0},
    }
    assert list(b.bt) == [WALL, WALL, INLET, OUTLET, WALL, WALL]
    for side in range(6):
        for f, v in slots[b.bt[side]].items():
            got = o.phi(f)[N + side]
            assert abs(got - v) <= 1e-15 * max(abs(v), 1.0), (side, po.FIELD_NAMES[f], got, v)
    assert abs(o.s.inlet_effective_viscosity - EV_IN) <= 1e-15 * EV_IN
    assert o.s.effective_viscosity == MU
    assert np.all(o.cell("cell_densities") == RHO)
    np.testing.assert_allclose(o.cell("cell_volumes"), VOL, rtol=1e-14)


def test_csr_pattern_is_the_seqaij_pattern():
    """nnz = N + 2 F_int = 14 m^3 - 10 m^2; each row holds the cell and its
    interior neighbours, columns strictly ascending; diag_idx points at the
    diagonal."""
    m = 4
    o = _fresh(m)
    b = Box(o)
    rp, col = o.csr()
    N = o.N
    F_int = int(np.sum(b.inter))
    assert F_int == 6 * m ** 3 - 5 * m ** 2
    assert rp[-1] == N + 2 * F_int == 14 * m ** 3 - 10 * m ** 2
    diag = po._arr(o.s.diag_idx, N, np.int64)
    for c in range(N):
        row = col[rp[c]:rp[c + 1]]
        assert np.all(np.diff(row) > 0), c
        nbs = {int(x) for x in b.nb[c] if x < np.uint64(N)}
        assert set(row.tolist()) == nbs | {c}, c
        assert col[diag[c]] == c


# ------------------------------------------------------------------ A6 --
def test_precomp_AU_closed_form():
    """precomp_AU (FlowSolver.inl:1065-1113): A_U = rho V / dt plus, per
    inlet or outlet face, mu rlencos_b - min(mflux, 0) with rlencos_b =
    A / (h / 2) = 0.1; walls add nothing."""
    o = _fresh(4)
    b = Box(o)
    rng = np.random.default_rng(1)
    mf = o.face("mflux")
    mf[:] = rng.uniform(-1e-3, 1e-3, size=o.F)
    o.precomp_AU()
    want = np.full(o.N, RDT)
    for f in np.nonzero(b.bnd)[0]:
        if b.bt[b.side[f]] in (INLET, OUTLET):
            want[b.faces[f, 0]] += MU * 0.1 - min(mf[f], 0.0)
    np.testing.assert_allclose(o.cell("A_U"), want, rtol=1e-14)


# ------------------------------------------------------------------ A7 --
def _set_velocity(o, b, a, c, d, lin, rng):
    """U = a x^2 + l_U.x, V = c x y + l_V.x, W = d x z + l_W.x with their exact
    gradients written into the gradient arrays."""
    x, y, z = b.cc[:, 0], b.cc[:, 1], b.cc[:, 2]
    X = b.cc
    N = o.N
    U, V, W = o.phi(po.U), o.phi(po.V), o.phi(po.W)
    U[:N] = a * x * x + X @ lin[0]
    V[:N] = c * x * y + X @ lin[1]
    W[:N] = d * x * z + X @ lin[2]
    gU, gV, gW = o.grad(po.U), o.grad(po.V), o.grad(po.W)
    gU[:N] = np.stack([2 * a * x, 0 * x, 0 * x], axis=1) + lin[0]
    gV[:N] = np.stack([c * y, c * x, 0 * x], axis=1) + lin[1]
    gW[:N] = np.stack([d * z, 0 * x, d * x], axis=1) + lin[2]


def test_momentum_stress_sources_closed_form():
    """Deferred-correction stress sources of calculate_flux_UVW
    (FlowSolver.inl:1377-1653, GammaBlend = 0): on the orthogonal box a face
    with normal e_k adds mu A (d u_k / d x_i) to equation i (the transpose
    stress; the normal-gradient part is implicit), with the face gradient the
    mean of its cells'.  For U = a x^2, V = c x y, W = d x z (+ any linear
    field) a cell away from the boundary therefore gets
    S_U = mu V (2a + c + d) and S_V = S_W = 0 (the discrete div of grad u^T)."""
    o = _fresh(6)
    b = Box(o)
    rng = np.random.default_rng(2)
    a, c, d = 300.0, -170.0, 90.0
    lin = rng.uniform(-20, 20, size=(3, 3))
    _set_velocity(o, b, a, c, d, lin, rng)
    o.face("mflux")[:] = 0.0
    for name in ("S_U", "S_V", "S_W", "A_U", "A_V", "A_W"):
        o.cellslot(name)[:] = 0.0
    o.calculate_flux_UVW()
    ic = b.interior_cells
    assert ic.size > 20
    scale = MU * AREA * 2 * abs(a) * 2 * H * 12   # largest single face term
    np.testing.assert_allclose(o.cell("S_U")[ic], MU * VOL * (2 * a + c + d), rtol=0, atol=1e-9 * scale)
    assert np.max(np.abs(o.cell("S_V")[ic])) <= 1e-9 * scale
    assert np.max(np.abs(o.cell("S_W")[ic])) <= 1e-9 * scale
    # zero mass flux: both coefficients of every interior face are -mu rlencos = -mu A / h
    np.testing.assert_allclose(o.face("ff0")[b.inter], -MU * AREA / H, rtol=1e-14)
    np.testing.assert_allclose(o.face("ff1")[b.inter], -MU * AREA / H, rtol=1e-14)


def _boundary_rows_closed_form(o, b, U, V, W):
    """Expected A (any component) and S_U, S_V, S_W of calculate_flux_UVW with
    zero gradients and zero mass flux: per wall face mu A / (h/2) on the
    diagonal and (mu A / (h/2)) (u . n) n on the right (the tangential part
    cancels against the wall shear mu A u_t / (h/2)); per inlet face
    ev_in A / (h/2) and ev_in A / (h/2) * (50, 0, 0); per outlet face
    mu A / (h/2) and mu A / (h/2) * u_cell."""
    N = o.N
    A = np.zeros(N)
    S = np.zeros((N, 3))
    coef = AREA / (H / 2)
    for f in np.nonzero(b.bnd)[0]:
        c0 = b.faces[f, 0]
        t = b.bt[b.side[f]]
        u = np.array([U[c0], V[c0], W[c0]])
        if t == WALL:
            n = SIDE_NORMAL[b.side[f]]
            A[c0] += MU * coef
            S[c0] += MU * coef * (u @ n) * n
        elif t == INLET:
            A[c0] += EV_IN * coef
            S[c0] += EV_IN * coef * np.array([50.0, 0.0, 0.0])
        else:
            A[c0] += MU * coef
            S[c0] += MU * coef * u
    return A, S


def test_momentum_boundary_rows_closed_form():
    """Wall, inlet and outlet faces of calculate_flux_UVW with zero gradients
    and zero mass flux (FlowSolver.inl:1560-1650), and the boundary slots they
    write (the last face of each side wins: inlet 50, outlet its cell's
    velocity, walls 0)."""
    o = _fresh(4)
    b = Box(o)
    rng = np.random.default_rng(3)
    N = o.N
    for f in (po.U, po.V, po.W):
        o.phi(f)[:N] = rng.uniform(-30, 60, size=N)
        o.grad(f)[:] = 0.0
    o.face("mflux")[:] = 0.0
    for name in ("S_U", "S_V", "S_W", "A_U", "A_V", "A_W"):
        o.cellslot(name)[:] = 0.0
    U, V, W = (o.phi(f)[:N].copy() for f in (po.U, po.V, po.W))
    o.calculate_flux_UVW()
    A, S = _boundary_rows_closed_form(o, b, U, V, W)
    for k, name in enumerate(("S_U", "S_V", "S_W")):
        np.testing.assert_allclose(o.cell(name), S[:, k], rtol=1e-12, atol=1e-18)
    for name in ("A_U", "A_V", "A_W"):
        np.testing.assert_allclose(o.cell(name), A, rtol=1e-13)
    for side in range(6):
        c = b.faces[b.last_face[side], 0]
        want = {WALL: (0, 0, 0), INLET: (50, 0, 0), OUTLET: (U[c], V[c], W[c])}[b.bt[side]]
        got = tuple(o.phi(f)[N + side] for f in (po.U, po.V, po.W))
        assert got == tuple(float(x) for x in want), (side, got, want)


# ------------------------------------------------------------------ A8 --
def test_momentum_cell_sources_and_under_relaxation_closed_form():
    """calculate_UVW's cell terms (FlowSolver.inl:1689-1758): buoyancy
    -0.001 rho V (T - 2000) g with g = (0, -9.81, 0), -grad P V, the unsteady
    rho V / dt on both sides; then under-relaxation 0.5 of the U row:
    diag = (A + sum of the interior coefficients mu A / h) / 0.5 and
    b = S + 0.5 diag U."""
    o = _fresh(4)
    b = Box(o)
    rng = np.random.default_rng(4)
    N = o.N
    for f in (po.U, po.V, po.W):
        o.phi(f)[:N] = rng.uniform(-30, 60, size=N)
        o.grad(f)[:] = 0.0
    o.phi(po.TEM)[:N] = rng.uniform(300, 2500, size=N)
    gP = rng.uniform(-1e3, 1e3, size=(N, 3))
    o.grad(po.P)[:N] = gP
    o.face("mflux")[:] = 0.0
    U, V, W = (o.phi(f)[:N].copy() for f in (po.U, po.V, po.W))
    T = o.phi(po.TEM)[:N].copy()
    o.momentum_assemble()
    A, S = _boundary_rows_closed_form(o, b, U, V, W)
    buoy_y = 9.81 * 0.001 * RHO * VOL * (T - 2000.0)
    S_pre = S + np.stack([0 * T, buoy_y, 0 * T], axis=1) - gP * VOL + RDT * np.stack([U, V, W], axis=1)
    np.testing.assert_allclose(o.cell("S_V"), S_pre[:, 1], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(o.cell("S_W"), S_pre[:, 2], rtol=1e-12, atol=1e-15)
    n_int = np.sum(b.nb < np.uint64(N), axis=1)
    diag = (A + RDT + n_int * MU * AREA / H) / 0.5
    rp, col = o.csr()
    Ad = o.vals("A_vals")
    di = po._arr(o.s.diag_idx, N, np.int64)
    np.testing.assert_allclose(Ad[di], diag, rtol=1e-13)
    np.testing.assert_allclose(o.vec("b"), S_pre[:, 0] + 0.5 * diag * U, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------- A11/A12 --
def test_rhie_chow_flux_and_pressure_coefficient_closed_form():
    """calculate_mass_flux (FlowSolver.inl:1880-2104) with uniform velocity
    and P = a x^3 + p.x (exact gradient): the face value of a uniform field
    is the field, central differences of the cubic leave -a h^3 / 2 across an
    x face, so mflux = rho u . S + ApV a h^3 / 2 on x faces and rho u . S on
    y, z faces, with ApV = rho (1/(2 A_U0) + 1/(2 A_U1)) A V / h; both face
    coefficients are -ApV; inlet and outlet carry 1.2 * 50 * (area normal)_x
    (the outlet fixed to the inlet velocity), walls zero."""
    o = _fresh(5)
    b = Box(o)
    rng = np.random.default_rng(5)
    N = o.N
    u = np.array([41.0, -3.0, 7.5])
    for k, f in enumerate((po.U, po.V, po.W)):
        o.phi(f)[:] = u[k]
        o.grad(f)[:] = 0.0
    a, p = 2.0e5, rng.uniform(-100, 100, size=3)
    x = b.cc[:, 0]
    o.phi(po.P)[:N] = a * x ** 3 + b.cc @ p
    o.grad(po.P)[:N] = np.stack([3 * a * x * x, 0 * x, 0 * x], axis=1) + p
    AU = o.cellslot("A_U")
    AU[:N] = rng.uniform(10.0, 40.0, size=N)
    for name in ("mflux", "ff0", "ff1"):
        o.face(name)[:] = 0.0
    o.cellslot("S_U")[:] = 0.0
    o.calculate_mass_flux()
    mf, ff0, ff1 = o.face("mflux"), o.face("ff0"), o.face("ff1")
    fi = np.nonzero(b.inter)[0]
    c0, c1 = b.faces[fi, 0], b.faces[fi, 1]
    axis = np.argmax(np.abs(b.cc[c1] - b.cc[c0]), axis=1)
    apv = RHO * (0.5 / AU[c0] + 0.5 / AU[c1]) * AREA * VOL / H
    want = RHO * u[axis] * AREA + np.where(axis == 0, apv * a * H ** 3 / 2, 0.0)
    np.testing.assert_allclose(mf[fi], want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(ff0[fi], -apv, rtol=1e-13)
    np.testing.assert_allclose(ff1[fi], -apv, rtol=1e-13)
    for f in np.nonzero(b.bnd)[0]:
        t = b.bt[b.side[f]]
        want_b = {WALL: 0.0, INLET: -1.2 * 50 * AREA, OUTLET: 1.2 * 50 * AREA}[t]
        assert abs(mf[f] - want_b) <= 1e-15, (f, mf[f], want_b)


# ------------------------------------------------------------ A14/A15 --
def _correction_state(o, b, rng, g):
    N = o.N
    for f in (po.U, po.V, po.W, po.P):
        o.phi(f)[:N] = rng.uniform(-50, 50, size=N)
    o.cellslot("A_U")[:N] = rng.uniform(10.0, 40.0, size=N)
    o.face("mflux")[:] = rng.uniform(-1e-2, 1e-2, size=o.F)
    ff0 = o.face("ff0")
    ff0[:] = 0.0
    ff0[b.inter] = -rng.uniform(1e-8, 1e-6, size=int(np.sum(b.inter)))
    o.phi(po.PP)[:N] = b.cc @ g + 3.0


def test_pressure_correction_step_closed_form():
    """One correction loop (FlowSolver.inl:2391-2433) in simple mode with a
    linear PP: P += 0.2 PP everywhere, u -= grad PP V / A_U (grad PP exact away
    from the boundary), every interior face flux += ff0 (PP1 - PP0) =
    ff0 g . (x1 - x0), and each PP slot is its side's last-face cell value."""
    o = _fresh(5)
    b = Box(o)
    rng = np.random.default_rng(6)
    N = o.N
    g = np.array([120.0, -45.0, 30.0])
    _correction_state(o, b, rng, g)
    P0 = o.phi(po.P)[:N].copy()
    vel0 = np.stack([o.phi(f)[:N].copy() for f in (po.U, po.V, po.W)], axis=1)
    mf0 = o.face("mflux").copy()
    PP = o.phi(po.PP)[:N].copy()
    AU = o.cell("A_U").copy()
    o.apply_pressure_correction()
    assert np.array_equal(o.phi(po.P)[:N], P0 + 0.2 * PP)
    ic = b.interior_cells
    vel = np.stack([o.phi(f)[:N] for f in (po.U, po.V, po.W)], axis=1)
    np.testing.assert_allclose(vel[ic], vel0[ic] - np.outer(VOL / AU[ic], g), rtol=1e-12, atol=1e-12)
    fi = np.nonzero(b.inter)[0]
    d = b.cc[b.faces[fi, 1]] - b.cc[b.faces[fi, 0]]
    np.testing.assert_allclose(o.face("mflux")[fi], mf0[fi] + o.face("ff0")[fi] * (d @ g), rtol=1e-12, atol=1e-18)
    assert np.array_equal(o.face("mflux")[b.bnd], mf0[b.bnd])
    for side in range(6):
        assert o.phi(po.PP)[N + side] == PP[b.faces[b.last_face[side], 0]]


def test_compat_corrects_cell_zero_once():
    """The reference CPU path's quirk (FlowSolver.inl:2416-2433): in compat
    mode the first loop of the first pressure solve corrects cell 0 only and
    clears first_pressure, after which no cell is corrected; face fluxes are
    corrected every loop."""
    o = _fresh(4, mode=po.MODE_COMPAT)
    b = Box(o)
    rng = np.random.default_rng(7)
    N = o.N
    g = np.array([50.0, 20.0, -10.0])
    _correction_state(o, b, rng, g)
    assert o.s.first_pressure == 1
    before = {f: o.phi(f)[:N].copy() for f in (po.U, po.V, po.W, po.P)}
    mf0 = o.face("mflux").copy()
    o.apply_pressure_correction()
    assert o.s.first_pressure == 0
    for f, v in before.items():
        now = o.phi(f)[:N]
        assert now[0] != v[0], po.FIELD_NAMES[f]
        assert np.array_equal(now[1:], v[1:]), po.FIELD_NAMES[f]
    mf1 = o.face("mflux").copy()
    assert not np.array_equal(mf1[b.inter], mf0[b.inter])
    mid = {f: o.phi(f)[:N].copy() for f in (po.U, po.V, po.W, po.P)}
    o.apply_pressure_correction()
    for f, v in mid.items():
        assert np.array_equal(o.phi(f)[:N], v), po.FIELD_NAMES[f]
    assert not np.array_equal(o.face("mflux")[b.inter], mf1[b.inter])


def test_update_P_slots_closed_form():
    """update_P (FlowSolver.inl:2292-2315): the outlet slot copies its cell,
    every other side extrapolates P + grad P . (h/2) n_out from the cell of
    the side's last face; Update_P_at_boundaries (:2275-2290) copies that
    cell's value on every side."""
    o = _fresh(4)
    b = Box(o)
    rng = np.random.default_rng(8)
    N = o.N
    P = o.phi(po.P)
    P[:N] = rng.uniform(3000, 5000, size=N)
    gP = rng.uniform(-500, 500, size=(N, 3))
    o.grad(po.P)[:N] = gP
    o.update_P(po.P)
    for side in range(6):
        c = b.faces[b.last_face[side], 0]
        want = P[c] if b.bt[side] == OUTLET else P[c] + gP[c] @ (SIDE_NORMAL[side] * H / 2)
        assert abs(P[N + side] - want) <= 1e-12 * abs(want), (side, P[N + side], want)
    o.update_P_at_boundaries(po.P)
    for side in range(6):
        assert P[N + side] == P[b.faces[b.last_face[side], 0]]


def test_non_orthogonal_flux_update_vanishes_on_the_box():
    """update_mass_flux (FlowSolver.inl:1812-1878): on the orthogonal box both
    cells' centres project onto the face normal through the face centre, so
    the correction is zero for any grad PP: fluxes and RHS unchanged."""
    o = _fresh(4)
    rng = np.random.default_rng(9)
    N = o.N
    o.grad(po.PP)[:N] = rng.uniform(-1e4, 1e4, size=(N, 3))
    o.face("ff0")[:] = -rng.uniform(1e-8, 1e-6, size=o.F)
    mf = o.face("mflux")
    mf[:] = rng.uniform(-1, 1, size=o.F)
    before = mf.copy()
    o.cellslot("S_U")[:] = 0.0
    o.update_mass_flux()
    assert np.array_equal(mf, before)
    assert not np.any(o.cell("S_U"))


# ------------------------------------------------------------------ A16 --
def _turbulence_state(o, rng):
    N = o.N
    for f in (po.U, po.V, po.W):
        o.phi(f)[:N] = rng.uniform(-30, 60, size=N)
        o.grad(f)[:N] = rng.uniform(-1e3, 1e3, size=(N, 3))
    o.phi(po.TE)[:N] = rng.uniform(0.1, 40.0, size=N)
    o.phi(po.ED)[:N] = rng.uniform(0.1, 80.0, size=N)
    o.cellslot("S_U")[:] = 0.0
    o.cellslot("A_V")[:] = 0.0


def test_te_production_dissipation_and_wall_function():
    """solveTurbulenceModels for TE (FlowSolver.inl:2746-2838): the
    production uses VisT = ev - ev = 0, so away from walls TP = 0 and only the
    dissipation rho ED / TE V enters the diagonal; in a wall cell the log-law
    wall function gives TP = tau_w u_tau / (0.419 d) with d = h / 2,
    tau_w = mu |u_t| / d, u_tau = sqrt(tau_w / rho) from the tangential
    velocity of the cell's last wall face (S = TP V), and each wall face adds
    rho C_mu^0.75 sqrt(TE) / (0.419 d) V to the diagonal."""
    o = _fresh(4)
    b = Box(o)
    rng = np.random.default_rng(10)
    N = o.N
    _turbulence_state(o, rng)
    U, V, W = (o.phi(f)[:N].copy() for f in (po.U, po.V, po.W))
    TE, ED = o.phi(po.TE)[:N].copy(), o.phi(po.ED)[:N].copy()
    o.solve_turbulence_models(po.TERBTE)
    TP, S, A = o.phi(po.TP)[:N], o.cell("S_U"), o.cell("A_V")
    d = H / 2
    wall_faces = {}
    for f in np.nonzero(b.bnd)[0]:
        if b.bt[b.side[f]] == WALL:
            wall_faces.setdefault(int(b.faces[f, 0]), []).append(int(f))
    diss = RHO * ED / (TE + 1e-18) * VOL
    for c in range(N):
        if c not in wall_faces:
            assert TP[c] == 0.0 and S[c] == 0.0, c
            assert abs(A[c] - diss[c]) <= 1e-14 * diss[c], c
            continue
        last = max(wall_faces[c])
        n = SIDE_NORMAL[b.side[last]]
        up = np.array([U[c], V[c], W[c]])
        ut = up - (up @ n) * n
        tau = MU * np.linalg.norm(ut) / d
        tp = tau * np.sqrt(tau / RHO) / (0.419 * d)
        assert abs(TP[c] - tp) <= 1e-12 * tp, (c, TP[c], tp)
        assert abs(S[c] - tp * VOL) <= 1e-12 * tp * VOL, c
        want_a = diss[c] + len(wall_faces[c]) * RHO * CMU75 * np.sqrt(TE[c]) / (0.419 * d) * VOL
        assert abs(A[c] - want_a) <= 1e-12 * want_a, c


def test_ed_sources_and_wall_row():
    """solveTurbulenceModels for ED (FlowSolver.inl:2838-2883): away from
    walls S = 1.44 ED / TE V TP and diagonal 1.92 ED / TE V rho; a wall cell's
    ED is fixed to C_mu^0.75 TE^1.5 / (0.419 d): its row becomes the identity
    (S = that value, A = 1, its face coefficients zero), so after
    under-relaxation (0.5) the assembled row is 2 ED = 2 value with no
    off-diagonal and the solve returns that value."""
    o = _fresh(4)
    b = Box(o)
    rng = np.random.default_rng(11)
    N = o.N
    _turbulence_state(o, rng)
    o.phi(po.TP)[:N] = rng.uniform(0.0, 5e3, size=N)
    TE, ED, TP = (o.phi(f)[:N].copy() for f in (po.TE, po.ED, po.TP))
    o.solve_turbulence_models(po.TERBED)
    S, A = o.cell("S_U"), o.cell("A_V")
    wall_cells = {int(b.faces[f, 0]) for f in np.nonzero(b.bnd)[0] if b.bt[b.side[f]] == WALL}
    fact = ED / (TE + 1e-18) * VOL
    dis = CMU75 * TE ** 1.5 / (0.419 * H / 2)
    for c in range(N):
        if c in wall_cells:
            assert abs(S[c] - dis[c]) <= 1e-13 * dis[c] and A[c] == 1.0, c
            assert abs(o.phi(po.ED)[c] - dis[c]) <= 1e-13 * dis[c], c
        else:
            assert abs(S[c] - 1.44 * fact[c] * TP[c]) <= 1e-13 * abs(1.44 * fact[c] * TP[c]) + 1e-300, c
            assert abs(A[c] - 1.92 * fact[c] * RHO) <= 1e-13 * 1.92 * fact[c] * RHO, c
    # the assembled system: wall rows are 2 x = 2 dis with no off-diagonal
    o2 = _fresh(4)
    _turbulence_state(o2, np.random.default_rng(11))
    o2.scalar_assemble(po.TERBED, po.ED)
    rp, col = o2.csr()
    vals, rhs = o2.vals("A_vals"), o2.vec("b")
    TE2 = o2.phi(po.TE)[:N]
    for c in sorted(wall_cells):
        row = slice(rp[c], rp[c + 1])
        cols, v = col[row], vals[row]
        assert np.all(v[cols != c] == 0.0) and v[cols == c][0] == 2.0, c
        dis2 = CMU75 * TE2[c] ** 1.5 / (0.419 * H / 2)
        assert abs(rhs[c] - 2 * dis2) <= 1e-13 * dis2, c
