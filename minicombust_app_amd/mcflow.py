"""ctypes binding of libmcflow.so (include/mcflow.h).

This is the Python-side FFI of the C-ABI, the same binding a ctypes user of
the library would write.  It loads the in-tree build
``minicombust_app_amd/_lib/libmcflow.so`` (or the file named by the MCFLOW_LIB
environment variable) and raises if that file is missing: there is no
fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MCFLOW_LIB") or os.path.join(HERE, "_lib", "libmcflow.so")

MC_OK, MC_ERR_ARG, MC_ERR_HIP, MC_ERR_NOMEM, MC_ERR_STATE, MC_ERR_NOCONV, MC_ERR_NODEV = range(7)
(U, V, W, P, PP, TE, ED, TP, TEM, FUL, PRO, VARF, VARP) = range(13)
NFIELDS = 13
TERBTE, TERBED, TEMP, FUEL, PROG, VARFU, VARPR = range(7)
MODE_COMPAT, MODE_SIMPLE = 0, 1
PSOLVE_FP64, PSOLVE_MIXED = 0, 1

ARR_PHI, ARR_GRAD = 0, 16
ARR_AU, ARR_AV, ARR_AW, ARR_SU, ARR_SV, ARR_SW = range(32, 38)
ARR_FF0, ARR_FF1, ARR_MFLUX = 40, 41, 42
(ARR_FACE_AREA, ARR_FACE_LAMBDA, ARR_FACE_RLENCOS, ARR_FACE_NORMAL, ARR_FACE_CENTER,
 ARR_CELL_CENTER, ARR_CELL_VOLUME, ARR_CELL_DENSITY) = range(48, 56)
ARR_ROW_PTR, ARR_COL, ARR_A_VALS, ARR_PA_VALS, ARR_PB = range(64, 69)
ARR_FACES, ARR_CELL_FACES, ARR_CELL_NEIGHBOURS, ARR_CELLS, ARR_POINTS, ARR_CELLS_PER_POINT = range(72, 78)
ARR_PARTICLE_TERMS = 80
ARR_PDEFL = 84   # + k: k-th newest recycled loop-1 pressure solution (read-only)

KC_NAMES = ["grad11", "grad1", "mom_asm", "p_asm", "scalar_asm", "pcg_spmv64", "pcg_spmv32", "vcycle",
            "flux_correct", "mpnn_fwd", "jac64", "jac32", "bicg_spmv", "scalar_batch", "ir_resid", "mpnn_bwd"]
# kernel symbol (as a profiler lists it) of each single-kernel class (p_asm
# times two kernels, mc::k_mass_faces then mc::k_mass_rows)
KC_KERNELS = {"grad11": "mc::k_grad<11, false>", "grad1": "mc::k_grad<1, false>", "mom_asm": "mc::k_momentum",
              "scalar_asm": "mc::k_scalar(", "pcg_spmv64": "mc::k_pcg_pq<-7, double>",
              "pcg_spmv32": "mc::k_pcg_pq<7, float>", "flux_correct": "mc::k_flux_correct",
              "scalar_batch": "mc::k_scalar_batch<true>", "ir_resid": "mc::k_ir_resid<-7>"}
KC_COUNT = len(KC_NAMES)
# stages of the SIMPLE iteration (MC_ST_*)
ST_NAMES = ["grad", "vel_flux", "vel_solve", "pres_flux", "pres_solve", "pres_correct", "turb_flux", "turb_solve",
            "sca_flux", "sca_solve"]
ST_COUNT = len(ST_NAMES)
MEM_NAMES = ["mesh", "fields", "matrix", "amg", "work", "mpnn", "dataset"]


class Options(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("psolve", C.c_int), ("learned", C.c_int), ("solve_scalars", C.c_int),
        ("delta", C.c_double), ("mom_rtol", C.c_double), ("mom_maxit", C.c_int),
        ("p_rtol", C.c_double), ("p_maxit", C.c_int),
        ("amg_nu", C.c_int), ("amg_omega", C.c_double), ("amg_cscale", C.c_double), ("amg_coarse_max", C.c_int),
        ("mixed_inner_rtol", C.c_double), ("check_every", C.c_int), ("inlet_scale", C.c_double),
        ("amg_cheb_ratio", C.c_double), ("cell_tile", C.c_int), ("limit_grad", C.c_int), ("amg_refresh", C.c_int),
        ("cell_ychunk", C.c_int), ("warm_start", C.c_int), ("amg_half", C.c_int),
        ("amg_tail", C.c_int), ("p_recycle", C.c_int), ("pcg_graph", C.c_int),
    ]


class SolveStatus(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("relres", C.c_double), ("outer", C.c_int),
                ("guess_used", C.c_int), ("bnorm", C.c_double)]


class Stats(C.Structure):
    _fields_ = [
        ("simple_iterations", C.c_int64), ("mom_iters", C.c_int64), ("scalar_iters", C.c_int64),
        ("p_iters", C.c_int64), ("p_solves", C.c_int64), ("p_loops", C.c_int64), ("p_outer", C.c_int64),
        ("last_p_relres", C.c_double), ("p_solve_ms", C.c_double), ("last_p_max", C.c_double),
        ("last_loops", C.c_int), ("p_relres_max", C.c_double), ("clip_cells", C.c_int64),
        ("clip_passes", C.c_int64), ("host_polls", C.c_int64), ("p_guesses", C.c_int64),
        ("p_guess_ratio_sum", C.c_double), ("p_defl_starts", C.c_int64), ("p_defl_ratio_sum", C.c_double),
        ("p_defl_vectors", C.c_int64), ("amg_setups", C.c_int64), ("amg_setup_ms", C.c_double),
        ("p_loops_skipped", C.c_int64),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64), ("n_faces", C.c_int64), ("n_points", C.c_int64),
        ("points", C.POINTER(C.c_double)), ("cells", C.POINTER(C.c_uint64)),
        ("cell_neighbours", C.POINTER(C.c_uint64)), ("faces", C.POINTER(C.c_uint64)),
        ("cell_faces", C.POINTER(C.c_uint64)), ("boundary_types", C.c_uint64 * 6),
    ]


class MCError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libmcflow.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MCError("libmcflow.so not built: run `python -m minicombust_app_amd.build` (looked for %s)"
                          % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P_ = C.c_void_p
        L.mc_options_default.argtypes = [C.POINTER(Options)]
        L.mc_ctx_create_box.argtypes = [C.c_int64, C.POINTER(Options), C.POINTER(P_)]
        L.mc_ctx_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(Options), C.POINTER(P_)]
        L.mc_ctx_destroy.argtypes = [P_]
        L.mc_last_error.argtypes = [P_]
        L.mc_last_error.restype = C.c_char_p
        L.mc_ctx_sizes.argtypes = [P_] + [C.POINTER(C.c_int64)] * 4
        L.mc_set_options.argtypes = [P_, C.POINTER(Options)]
        for n in ["mc_timestep", "mc_get_phi_gradients", "mc_limit_phi_gradients", "mc_set_up_field", "mc_momentum_assemble",
                  "mc_pressure_assemble", "mc_calculate_pressure", "mc_update_P", "mc_sync", "mc_profile_reset",
                  "mc_stage_reset",
                  "mc_scalar_batch_solve", "mc_reset_fields"]:
            getattr(L, n).argtypes = [P_]
        L.mc_simple_iteration.argtypes = [P_, C.c_int]
        L.mc_write_vtk.argtypes = [P_, C.c_char_p, C.c_int64, C.c_int]
        L.mc_repair_nonneg.argtypes = [P_, C.c_int]
        L.mc_particle_terms_reset.argtypes = [P_]
        L.mc_particle_terms_add.argtypes = [P_, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_int64]
        L.mc_particle_terms_add_noalloc.argtypes = [P_, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_int64]
        L.mc_interpolate_to_nodes.argtypes = [P_, C.POINTER(C.c_uint64), C.c_int64, C.POINTER(C.c_uint64), C.c_int64,
                                              C.POINTER(C.c_double)]
        L.mc_get_phi_gradient.argtypes = [P_, C.c_int, C.c_int]
        L.mc_momentum_solve.argtypes = [P_, C.POINTER(SolveStatus)]
        L.mc_pressure_solve.argtypes = [P_, C.POINTER(SolveStatus)]
        L.mc_scalar_assemble.argtypes = [P_, C.c_int]
        L.mc_scalar_solve.argtypes = [P_, C.c_int]
        L.mc_update_boundary_slots.argtypes = [P_, C.c_int]
        L.mc_spmv.argtypes = [P_, C.c_int, C.c_void_p, C.c_void_p]
        L.mc_solve_pressure_host.argtypes = [P_, C.c_void_p, C.c_void_p, C.POINTER(SolveStatus)]
        L.mc_array_get.argtypes = [P_, C.c_int, C.c_void_p, C.c_int64]
        L.mc_array_set.argtypes = [P_, C.c_int, C.c_void_p, C.c_int64]
        L.mc_get_stats.argtypes = [P_, C.POINTER(Stats)]
        L.mc_state_flags.argtypes = [P_, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.mc_set_state_flags.argtypes = [P_, C.c_int, C.c_int64]
        L.mc_profile_enable.argtypes = [P_, C.c_int]
        L.mc_profile_read.argtypes = [P_, C.c_void_p, C.c_void_p, C.c_void_p]
        L.mc_stage_enable.argtypes = [P_, C.c_int]
        L.mc_stage_read.argtypes = [P_, C.c_void_p, C.c_void_p]
        L.mc_copy_probe.argtypes = [P_, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.mc_ctx_memory.argtypes = [P_, C.c_void_p]
        L.mc_mpnn_nparams.restype = C.c_int64
        L.mc_mpnn_nparams.argtypes = []
        L.mc_mpnn_load.argtypes = [P_, C.c_void_p, C.c_int64]
        L.mc_mpnn_save.argtypes = [P_, C.c_void_p, C.c_int64]
        L.mc_mpnn_collect.argtypes = [P_, C.c_int64]
        L.mc_mpnn_dataset_size.argtypes = [P_, C.POINTER(C.c_int64)]
        L.mc_mpnn_pair_get.argtypes = [P_, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.mc_mpnn_forward_pair.argtypes = [P_, C.c_int64, C.c_void_p]
        L.mc_mpnn_grad.argtypes = [P_, C.c_int64, C.c_void_p, C.POINTER(C.c_double)]
        L.mc_mpnn_train_step.argtypes = [P_, C.c_int64, C.c_double, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def default_options(**kw):
    o = Options()
    lib().mc_options_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def mpnn_nparams():
    """Number of FP32 parameters of the learned-guess net."""
    return int(lib().mc_mpnn_nparams())


class Context:
    """One GPU flow-solver context (the FlowSolver object of the reference)."""

    def __init__(self, m=None, mesh=None, **opts):
        self.L = lib()
        self.opts = default_options(**opts)
        h = C.c_void_p()
        if mesh is None:
            rc = self.L.mc_ctx_create_box(int(m), C.byref(self.opts), C.byref(h))
        else:
            self._mesh_keep = mesh
            d = MeshDesc()
            d.n_cells = mesh["cells"].shape[0]
            d.n_faces = mesh["faces"].shape[0]
            d.n_points = mesh["points"].shape[0]
            for k in ["points", "cells", "cell_neighbours", "faces", "cell_faces"]:
                a = np.ascontiguousarray(mesh[k])
                mesh[k] = a
                t = C.c_double if a.dtype == np.float64 else C.c_uint64
                setattr(d, k, a.ctypes.data_as(C.POINTER(t)))
            for i in range(6):
                d.boundary_types[i] = int(mesh["boundary_types"][i])
            rc = self.L.mc_ctx_create(C.byref(d), C.byref(self.opts), C.byref(h))
        self.h = h
        if rc != MC_OK:
            msg = self.L.mc_last_error(h).decode() if h.value else ""
            if h.value:
                self.L.mc_ctx_destroy(h)
            self.h = None
            raise MCError("context creation failed (rc=%d): %s" % (rc, msg))
        n, f, z, p = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self.L.mc_ctx_sizes(self.h, C.byref(n), C.byref(f), C.byref(z), C.byref(p))
        self.N, self.F, self.nnz, self.NP = n.value, f.value, z.value, p.value

    def close(self):
        if getattr(self, "h", None):
            self.L.mc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != MC_OK:
            raise MCError("%s failed (rc=%d): %s" % (what, rc, self.L.mc_last_error(self.h).decode()))

    # ---- SIMPLE step and stages ----
    def timestep(self):
        self._check(self.L.mc_timestep(self.h), "mc_timestep")

    def simple_iteration(self, inner):
        self._check(self.L.mc_simple_iteration(self.h, inner), "mc_simple_iteration")

    def get_phi_gradients(self):
        self._check(self.L.mc_get_phi_gradients(self.h), "mc_get_phi_gradients")

    def write_vtk(self, prefix, timestep, rank=0):
        """The reference's flow VTK file: <prefix>_flow<rank>_timestep<timestep>.vtk."""
        self._check(self.L.mc_write_vtk(self.h, prefix.encode(), timestep, rank), "mc_write_vtk")
        return "%s_flow%d_timestep%d.vtk" % (prefix, rank, timestep)

    def particle_terms_reset(self):
        self._check(self.L.mc_particle_terms_reset(self.h), "mc_particle_terms_reset")

    def particle_terms_add(self, cells, terms):
        """terms: (n, 5) particle_aos rows (momentum xyz, energy, fuel) for cells (n,)."""
        cells = np.ascontiguousarray(cells, dtype=np.uint64)
        terms = np.ascontiguousarray(terms, dtype=np.float64).reshape(-1, 5)
        assert terms.shape[0] == cells.shape[0]
        self._check(self.L.mc_particle_terms_add(self.h, cells.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                 terms.ctypes.data_as(C.POINTER(C.c_double)), cells.shape[0]),
                    "mc_particle_terms_add")

    def particle_terms_add_noalloc(self, cells, terms):
        """particle_terms_add through the entry point that allocates no device memory."""
        cells = np.ascontiguousarray(cells, dtype=np.uint64)
        terms = np.ascontiguousarray(terms, dtype=np.float64).reshape(-1, 5)
        assert terms.shape[0] == cells.shape[0]
        self._check(self.L.mc_particle_terms_add_noalloc(self.h, cells.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                         terms.ctypes.data_as(C.POINTER(C.c_double)),
                                                         cells.shape[0]),
                    "mc_particle_terms_add_noalloc")

    def interpolate_to_nodes(self, cells, nodes):
        """(len(nodes), 5) flow_aos rows (vel xyz, pressure, temp)."""
        cells = np.ascontiguousarray(cells, dtype=np.uint64)
        nodes = np.ascontiguousarray(nodes, dtype=np.uint64)
        out = np.zeros((nodes.shape[0], 5))
        self._check(self.L.mc_interpolate_to_nodes(self.h, cells.ctypes.data_as(C.POINTER(C.c_uint64)), cells.shape[0],
                                                   nodes.ctypes.data_as(C.POINTER(C.c_uint64)), nodes.shape[0],
                                                   out.ctypes.data_as(C.POINTER(C.c_double))),
                    "mc_interpolate_to_nodes")
        return out

    def repair_nonneg(self, field):
        self._check(self.L.mc_repair_nonneg(self.h, int(field)), "mc_repair_nonneg")

    def limit_phi_gradients(self):
        self._check(self.L.mc_limit_phi_gradients(self.h), "mc_limit_phi_gradients")

    def get_phi_gradient(self, field, pressure):
        self._check(self.L.mc_get_phi_gradient(self.h, field, int(pressure)), "mc_get_phi_gradient")

    def set_up_field(self):
        self._check(self.L.mc_set_up_field(self.h), "mc_set_up_field")

    def momentum_assemble(self):
        self._check(self.L.mc_momentum_assemble(self.h), "mc_momentum_assemble")

    def momentum_solve(self):
        st = SolveStatus()
        self._check(self.L.mc_momentum_solve(self.h, C.byref(st)), "mc_momentum_solve")
        return st

    def pressure_assemble(self):
        self._check(self.L.mc_pressure_assemble(self.h), "mc_pressure_assemble")

    def pressure_solve(self):
        st = SolveStatus()
        self._check(self.L.mc_pressure_solve(self.h, C.byref(st)), "mc_pressure_solve")
        return st

    def calculate_pressure(self):
        self._check(self.L.mc_calculate_pressure(self.h), "mc_calculate_pressure")

    def scalar_assemble(self, t):
        self._check(self.L.mc_scalar_assemble(self.h, t), "mc_scalar_assemble")

    def scalar_solve(self, t):
        self._check(self.L.mc_scalar_solve(self.h, t), "mc_scalar_solve")

    def scalar_batch_solve(self):
        self._check(self.L.mc_scalar_batch_solve(self.h), "mc_scalar_batch_solve")

    def update_boundary_slots(self, field):
        self._check(self.L.mc_update_boundary_slots(self.h, field), "mc_update_boundary_slots")

    def update_P(self):
        self._check(self.L.mc_update_P(self.h), "mc_update_P")

    def sync(self):
        self._check(self.L.mc_sync(self.h), "mc_sync")

    def set_options(self, **kw):
        """Change options; when the library refuses them (MCError) the context
        and this object keep the previous ones."""
        new = Options.from_buffer_copy(self.opts)
        for k, v in kw.items():
            setattr(new, k, v)
        self._check(self.L.mc_set_options(self.h, C.byref(new)), "mc_set_options")
        self.opts = new

    # ---- linear algebra ----
    def spmv(self, x, which=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.N)
        self._check(self.L.mc_spmv(self.h, which, x.ctypes.data, y.ctypes.data), "mc_spmv")
        return y

    def solve_pressure(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty(self.N)
        st = SolveStatus()
        self._check(self.L.mc_solve_pressure_host(self.h, b.ctypes.data, x.ctypes.data, C.byref(st)),
                    "mc_solve_pressure_host")
        return x, st

    # ---- arrays ----
    def _count(self, aid):
        NS = self.N + 6
        if ARR_PHI <= aid < ARR_PHI + NFIELDS:
            return NS, np.float64
        if ARR_GRAD <= aid < ARR_GRAD + NFIELDS:
            return NS * 3, np.float64
        table = {
            ARR_AU: (NS, np.float64), ARR_AV: (NS, np.float64), ARR_AW: (NS, np.float64),
            ARR_SU: (NS, np.float64), ARR_SV: (NS, np.float64), ARR_SW: (NS, np.float64),
            ARR_FF0: (self.F, np.float64), ARR_FF1: (self.F, np.float64), ARR_MFLUX: (self.F, np.float64),
            ARR_FACE_AREA: (self.F, np.float64), ARR_FACE_LAMBDA: (self.F, np.float64),
            ARR_FACE_RLENCOS: (self.F, np.float64), ARR_FACE_NORMAL: (3 * self.F, np.float64),
            ARR_FACE_CENTER: (3 * self.F, np.float64), ARR_CELL_CENTER: (3 * self.N, np.float64),
            ARR_CELL_VOLUME: (self.N, np.float64), ARR_CELL_DENSITY: (self.N, np.float64),
            ARR_ROW_PTR: (self.N + 1, np.int32), ARR_COL: (self.nnz, np.int32),
            ARR_A_VALS: (self.nnz, np.float64), ARR_PA_VALS: (self.nnz, np.float64), ARR_PB: (self.N, np.float64),
            ARR_FACES: (2 * self.F, np.uint64), ARR_CELL_FACES: (6 * self.N, np.uint64),
            ARR_CELL_NEIGHBOURS: (6 * self.N, np.uint64), ARR_CELLS: (8 * self.N, np.uint64),
            ARR_POINTS: (3 * self.NP, np.float64), ARR_CELLS_PER_POINT: (self.NP, np.uint8),
            ARR_PARTICLE_TERMS: (5 * self.N, np.float64),
            ARR_PDEFL: (self.N, np.float64), ARR_PDEFL + 1: (self.N, np.float64),
            ARR_PDEFL + 2: (self.N, np.float64), ARR_PDEFL + 3: (self.N, np.float64),
        }
        return table[aid]

    def get(self, aid):
        n, dt = self._count(aid)
        out = np.empty(n, dtype=dt)
        self._check(self.L.mc_array_get(self.h, aid, out.ctypes.data, n), "mc_array_get(%d)" % aid)
        return out

    def set(self, aid, arr):
        n, dt = self._count(aid)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if a.size != n:
            raise ValueError("array %d: expected %d elements, got %d" % (aid, n, a.size))
        self._check(self.L.mc_array_set(self.h, aid, a.ctypes.data, n), "mc_array_set(%d)" % aid)

    def phi(self, f):
        return self.get(ARR_PHI + f)

    def grad(self, f):
        return self.get(ARR_GRAD + f).reshape(-1, 3)

    def stats(self):
        s = Stats()
        self._check(self.L.mc_get_stats(self.h, C.byref(s)), "mc_get_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def flags(self):
        fp, ts = C.c_int(), C.c_int64()
        self.L.mc_state_flags(self.h, C.byref(fp), C.byref(ts))
        return fp.value, ts.value

    def set_flags(self, first_pressure, timestep_count):
        self._check(self.L.mc_set_state_flags(self.h, int(first_pressure), int(timestep_count)), "flags")

    # ---- profiling ----
    def profile(self, on=True, classes=None):
        """Time kernel classes with HIP events: all classes, or those named."""
        if not on:
            mask = 0
        elif classes is None:
            mask = (1 << KC_COUNT) - 1
        else:
            mask = sum(1 << KC_NAMES.index(k) for k in classes)
        self._check(self.L.mc_profile_enable(self.h, mask), "mc_profile_enable")

    def memory(self):
        """Device bytes held by the context per category."""
        b = np.zeros(len(MEM_NAMES), dtype=np.int64)
        self._check(self.L.mc_ctx_memory(self.h, b.ctypes.data), "mc_ctx_memory")
        return {k: int(v) for k, v in zip(MEM_NAMES, b)}

    def reset_fields(self):
        """Back to the constructor's fields (mesh, geometry, AMG kept)."""
        self._check(self.L.mc_reset_fields(self.h), "mc_reset_fields")

    # ---- learned initial guess ----
    def mpnn_load(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.L.mc_mpnn_load(self.h, w.ctypes.data, w.size), "mc_mpnn_load")

    def mpnn_save(self):
        w = np.empty(mpnn_nparams(), dtype=np.float32)
        self._check(self.L.mc_mpnn_save(self.h, w.ctypes.data, w.size), "mc_mpnn_save")
        return w

    def mpnn_collect(self, max_pairs):
        self._check(self.L.mc_mpnn_collect(self.h, int(max_pairs)), "mc_mpnn_collect")

    def mpnn_dataset_size(self):
        n = C.c_int64(0)
        self._check(self.L.mc_mpnn_dataset_size(self.h, C.byref(n)), "mc_mpnn_dataset_size")
        return n.value

    def mpnn_pair(self, pair):
        f = np.empty((self.N, 4), dtype=np.float32)
        e = np.empty((self.N, 6), dtype=np.float32)
        t = np.empty(self.N, dtype=np.float32)
        self._check(self.L.mc_mpnn_pair_get(self.h, int(pair), f.ctypes.data, e.ctypes.data, t.ctypes.data),
                    "mc_mpnn_pair_get")
        return f, e, t

    def mpnn_forward_pair(self, pair):
        y = np.empty(self.N, dtype=np.float32)
        self._check(self.L.mc_mpnn_forward_pair(self.h, int(pair), y.ctypes.data), "mc_mpnn_forward_pair")
        return y

    def mpnn_grad(self, pair):
        g = np.empty(mpnn_nparams(), dtype=np.float32)
        loss = C.c_double(0.0)
        self._check(self.L.mc_mpnn_grad(self.h, int(pair), g.ctypes.data, C.byref(loss)), "mc_mpnn_grad")
        return g, loss.value

    def mpnn_train_step(self, pair, lr, want_loss=False):
        loss = C.c_double(float("nan"))
        self._check(self.L.mc_mpnn_train_step(self.h, int(pair), float(lr), C.byref(loss) if want_loss else None),
                    "mc_mpnn_train_step")
        return loss.value

    def copy_probe(self, nbytes, reps=10):
        """Average ms of a plain device copy of `nbytes` (read + write)."""
        ms = C.c_double(0.0)
        self._check(self.L.mc_copy_probe(self.h, int(nbytes), int(reps), C.byref(ms)), "mc_copy_probe")
        return ms.value

    def profile_reset(self):
        self._check(self.L.mc_profile_reset(self.h), "mc_profile_reset")

    def stage_enable(self, on=True):
        """Time the SIMPLE iteration's stages (MC_ST_*) with HIP events."""
        self._check(self.L.mc_stage_enable(self.h, 1 if on else 0), "mc_stage_enable")

    def stage_reset(self):
        self._check(self.L.mc_stage_reset(self.h), "mc_stage_reset")

    def stage_read(self):
        ms = np.zeros(ST_COUNT)
        calls = np.zeros(ST_COUNT, dtype=np.int64)
        self._check(self.L.mc_stage_read(self.h, ms.ctypes.data, calls.ctypes.data), "mc_stage_read")
        return {ST_NAMES[i]: {"ms": float(ms[i]), "calls": int(calls[i])} for i in range(ST_COUNT)}

    def profile_read(self):
        ms = np.zeros(KC_COUNT)
        calls = np.zeros(KC_COUNT, dtype=np.int64)
        by = np.zeros(KC_COUNT)
        self._check(self.L.mc_profile_read(self.h, ms.ctypes.data, calls.ctypes.data, by.ctypes.data),
                    "mc_profile_read")
        return {KC_NAMES[i]: {"ms": float(ms[i]), "calls": int(calls[i]), "bytes": float(by[i])}
                for i in range(KC_COUNT)}
