"""ctypes view of the CPU oracle (oracle/mc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by ``bench.py``'s cpu_baseline leg, always as the checker, never as the
product path.  Builds ``oracle/_build/libmc_oracle.so`` on first use if it is
missing (gcc is part of the image).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmc_oracle.so")
REF_MESH_DUMP = os.path.join(HERE, "_ref", "ref_mesh_dump")
REF_VTK_DUMP = os.path.join(HERE, "_ref", "ref_vtk_dump")

U, V, W, P, PP, TE, ED, TP, TEM, FUL, PRO, VARF, VARP = range(13)
NFIELDS = 13
TERBTE, TERBED, TEMP, FUEL, PROG, VARFU, VARPR = range(7)
MODE_COMPAT, MODE_SIMPLE = 0, 1
FIELD_NAMES = ["U", "V", "W", "P", "PP", "TE", "ED", "TP", "TEM", "FUL", "PRO", "VARF", "VARP"]


class Vec(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Mesh(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("n_cells", C.c_int64), ("n_faces", C.c_int64), ("n_points", C.c_int64),
        ("points", C.POINTER(Vec)), ("cells", C.POINTER(C.c_uint64)),
        ("cell_neighbours", C.POINTER(C.c_uint64)), ("faces", C.POINTER(C.c_uint64)),
        ("cell_faces", C.POINTER(C.c_uint64)), ("cells_per_point", C.POINTER(C.c_uint8)),
        ("cell_centers", C.POINTER(Vec)), ("boundary_types", C.c_uint64 * 6),
    ]


class Opts(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("delta", C.c_double), ("mom_rtol", C.c_double), ("mom_maxit", C.c_int),
        ("p_rtol", C.c_double), ("p_maxit", C.c_int), ("solve_scalars", C.c_int), ("p_solver", C.c_int),
        ("limit_grad", C.c_int), ("warm_start", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("mom_iters", C.c_int64), ("scalar_iters", C.c_int64), ("p_iters", C.c_int64),
        ("p_solves", C.c_int64), ("p_loops", C.c_int64), ("last_p_relres", C.c_double),
    ]


_DP = C.POINTER(C.c_double)
_VP = C.POINTER(Vec)
_IP = C.POINTER(C.c_int64)


class State(C.Structure):
    _fields_ = [
        ("mesh", Mesh), ("opts", Opts), ("stats", Stats), ("N", C.c_int64), ("F", C.c_int64),
        ("face_areas", _DP), ("face_lambdas", _DP), ("face_rlencos", _DP),
        ("face_normals", _VP), ("face_centers", _VP), ("cell_volumes", _DP), ("cell_densities", _DP),
        ("phi", _DP * NFIELDS), ("grad", _VP * NFIELDS),
        ("A_U", _DP), ("A_V", _DP), ("A_W", _DP), ("S_U", _DP), ("S_V", _DP), ("S_W", _DP),
        ("ff0", _DP), ("ff1", _DP), ("mflux", _DP),
        ("nnz", C.c_int64), ("row_ptr", _IP), ("col", _IP), ("diag_idx", _IP),
        ("face_idx01", _IP), ("face_idx10", _IP),
        ("A_vals", _DP), ("b", _DP), ("PA_vals", _DP), ("Pb", _DP), ("u", _DP),
        ("first_pressure", C.c_int), ("timestep_count", C.c_int64),
        ("effective_viscosity", C.c_double), ("inlet_effective_viscosity", C.c_double),
        ("pressure_correction_max_last", C.c_double), ("pressure_loops_last", C.c_int),
        ("p_atol", C.c_double), ("p_bnorm_last", C.c_double),
    ]


_lib = None


def build():
    """Compile the oracle library (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref():
    """Compile the reference mesh generator driver (needs /root/reference)."""
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        SP = C.POINTER(State)
        L.mco_create.restype = SP
        L.mco_create.argtypes = [C.c_int64, C.POINTER(Opts)]
        L.mco_default_opts.argtypes = [C.POINTER(Opts)]
        L.mco_destroy.argtypes = [SP]
        for name in ["mco_get_phi_gradients", "mco_limit_phi_gradients", "mco_precomp_AU", "mco_set_up_field", "mco_calculate_flux_UVW",
                     "mco_momentum_assemble", "mco_calculate_UVW", "mco_calculate_mass_flux",
                     "mco_setup_pressure_matrix", "mco_solve_pressure_matrix", "mco_update_mass_flux",
                     "mco_calculate_pressure", "mco_timestep", "mco_apply_pressure_correction"]:
            getattr(L, name).argtypes = [SP]
        L.mco_get_phi_gradient.argtypes = [SP, C.c_int, C.c_int]
        L.mco_repair_nonneg.argtypes = [SP, C.c_int]
        _UP = C.POINTER(C.c_uint64)
        L.mco_particle_terms_add.argtypes = [SP, C.c_int64, _UP, _DP, _DP]
        L.mco_interpolate_to_nodes.argtypes = [SP, C.c_int64, _UP, C.c_int64, _UP, _DP]
        L.mco_momentum_solve.argtypes = [SP, C.c_int]
        L.mco_update_P_at_boundaries.argtypes = [SP, C.c_int]
        L.mco_update_P.argtypes = [SP, C.c_int]
        L.mco_flux_scalar.argtypes = [SP, C.c_int, C.c_int]
        L.mco_solve_turbulence_models.argtypes = [SP, C.c_int]
        L.mco_scalar_assemble.argtypes = [SP, C.c_int, C.c_int]
        L.mco_scalar_solve.argtypes = [SP, C.c_int, C.c_int]
        L.mco_simple_iteration.argtypes = [SP, C.c_int]
        L.mco_mesh_build.argtypes = [C.c_int64, C.POINTER(Mesh)]
        L.mco_mesh_free.argtypes = [C.POINTER(Mesh)]
        L.mco_pcg_singular.restype = C.c_int
        L.mco_pcg_singular.argtypes = [C.c_int64, _IP, _IP, _DP, _DP, _DP, C.c_double, C.c_double, C.c_int, _DP,
                                       _DP]
        L.mco_bicgstab.restype = C.c_int
        L.mco_bicgstab.argtypes = [C.c_int64, _IP, _IP, _DP, _DP, _DP, C.c_double, C.c_int, _DP]
        L.mco_bicgstab_x0.restype = C.c_int
        L.mco_bicgstab_x0.argtypes = [C.c_int64, _IP, _IP, _DP, _DP, _DP, C.c_double, C.c_int, _DP, C.c_int]
        L.mco_pcg_amg.restype = C.c_int
        L.mco_pcg_amg.argtypes = [C.c_int64, _IP, _IP, _DP, _DP, _DP, C.c_double, C.c_double, C.c_int, _DP, _DP]
        L.mco_spmv.argtypes = [C.c_int64, _IP, _IP, _DP, _DP, _DP]
        _lib = L
    return _lib


def _arr(ptr, n, dtype=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


def default_opts(**kw):
    o = Opts()
    lib().mco_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Oracle:
    """Owns one oracle state; numpy views alias the C arrays (no copies)."""

    def __init__(self, m, **opts):
        self.L = lib()
        self.opts = default_opts(**opts)
        self.sp = self.L.mco_create(m, C.byref(self.opts))
        if not self.sp:
            raise RuntimeError("mco_create failed")
        self.s = self.sp.contents
        self.m = m
        self.N = self.s.N
        self.F = self.s.F

    def close(self):
        if self.sp:
            self.L.mco_destroy(self.sp)
            self.sp = None

    def __del__(self):
        self.close()

    # ---- views ----
    def phi(self, f):
        return _arr(self.s.phi[f], self.N + 6)

    def grad(self, f):
        return np.ctypeslib.as_array(C.cast(self.s.grad[f], _DP), shape=((self.N + 6) * 3,)).reshape(self.N + 6, 3)

    def cell(self, name):
        return _arr(getattr(self.s, name), self.N)

    def cellslot(self, name):
        return _arr(getattr(self.s, name), self.N + 6)

    def face(self, name):
        return _arr(getattr(self.s, name), self.F)

    def facevec(self, name):
        return np.ctypeslib.as_array(C.cast(getattr(self.s, name), _DP), shape=(self.F * 3,)).reshape(self.F, 3)

    def csr(self):
        s = self.s
        rp = _arr(s.row_ptr, self.N + 1, np.int64).copy()
        col = _arr(s.col, s.nnz, np.int64).copy()
        return rp, col

    def vals(self, which="A_vals"):
        return _arr(getattr(self.s, which), self.s.nnz)

    def vec(self, name):
        return _arr(getattr(self.s, name), self.N)

    def mesh_arrays(self):
        return mesh_arrays(self.s.mesh)

    # ---- stages ----
    def __getattr__(self, name):
        fn = getattr(lib(), "mco_" + name)
        return lambda *a: fn(self.sp, *a)


def mesh_arrays(M):
    N, F, NP = M.n_cells, M.n_faces, M.n_points
    pts = np.ctypeslib.as_array(C.cast(M.points, _DP), shape=(NP * 3,)).reshape(NP, 3).copy()
    cc = np.ctypeslib.as_array(C.cast(M.cell_centers, _DP), shape=(N * 3,)).reshape(N, 3).copy()
    return {
        "points": pts,
        "cells": _arr(M.cells, 8 * N, np.uint64).reshape(N, 8).copy(),
        "cell_neighbours": _arr(M.cell_neighbours, 6 * N, np.uint64).reshape(N, 6).copy(),
        "faces": _arr(M.faces, 2 * F, np.uint64).reshape(F, 2).copy(),
        "cell_faces": _arr(M.cell_faces, 6 * N, np.uint64).reshape(N, 6).copy(),
        "cells_per_point": np.ctypeslib.as_array(M.cells_per_point, shape=(NP,)).copy(),
        "cell_centers": cc,
        "boundary_types": np.array(list(M.boundary_types), dtype=np.uint64),
    }


def oracle_mesh(m):
    M = Mesh()
    rc = lib().mco_mesh_build(m, C.byref(M))
    if rc:
        raise RuntimeError("mco_mesh_build failed")
    out = mesh_arrays(M)
    lib().mco_mesh_free(C.byref(M))
    return out


def read_ref_mesh(path):
    """Parse the binary written by oracle/_ref/ref_mesh_dump."""
    with open(path, "rb") as f:
        m, N, F, NP = np.frombuffer(f.read(32), dtype=np.int64)
        def take(dtype, count):
            return np.frombuffer(f.read(np.dtype(dtype).itemsize * count), dtype=dtype)
        out = {"m": int(m)}
        out["points"] = take(np.float64, 3 * NP).reshape(NP, 3)
        out["cells"] = take(np.uint64, 8 * N).reshape(N, 8)
        out["cell_neighbours"] = take(np.uint64, 6 * N).reshape(N, 6)
        out["faces"] = take(np.uint64, 2 * F).reshape(F, 2)
        out["cell_faces"] = take(np.uint64, 6 * N).reshape(N, 6)
        out["cells_per_point"] = take(np.uint8, NP)
        out["cell_centers"] = take(np.float64, 3 * N).reshape(N, 3)
        out["boundary_types"] = take(np.uint64, 6)
    return out


def vtk_test_fields(n, seed):
    """Deterministic field values for the VTK writer tests: 11 x n values
    spanning many magnitudes, with signs, zeros and integral values."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((11, n)) * 10.0 ** rng.integers(-9, 9, size=(11, n))
    f[:, ::7] = 0.0
    f[:, 3::11] = np.round(f[:, 3::11])
    f[:, 5::13] = 50.0
    return f


def ref_vtk(m, fields, tmpdir, timestep=7):
    """The reference's own VisitWriter::write_flow output (oracle/_ref) for
    the box mesh m and fields (11 x N: U V W P TE ED TEM FUL PRO VARF VARP)."""
    if not os.path.exists(REF_VTK_DUMP):
        build_ref()
    fields = np.ascontiguousarray(fields, dtype=np.float64)
    inp = os.path.join(tmpdir, "vtk_fields.bin")
    with open(inp, "wb") as f:
        np.array([fields.shape[1]], dtype=np.int64).tofile(f)
        fields.tofile(f)
    prefix = os.path.join(tmpdir, "ref")
    subprocess.run([REF_VTK_DUMP, str(m), inp, prefix, str(timestep)], check=True, timeout=600,
                   stdout=subprocess.DEVNULL)
    with open(prefix + "_flow0_timestep%d.vtk" % timestep, "rb") as f:
        return f.read()


def ref_mesh(m, tmpdir):
    """Run the reference's own mesh generator (oracle/_ref) and parse its output."""
    if not os.path.exists(REF_MESH_DUMP):
        build_ref()
    path = os.path.join(tmpdir, "ref_mesh_%d.bin" % m)
    subprocess.run([REF_MESH_DUMP, str(m), path], check=True, timeout=600,
                   stdout=subprocess.DEVNULL)
    return read_ref_mesh(path)
