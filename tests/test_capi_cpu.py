"""CPU tests of the C-ABI boundary (include/mcflow.h, libmcflow.so): the
library loads, exports every function the header declares, its structs have
the layout the ctypes binding assumes, and the host-only entry points behave.
No call here touches a GPU."""
import ctypes as C
import os
import re
import subprocess
import sys
import tempfile

import pytest

from minicombust_app_amd import mcflow as mc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mcflow.h")


def _ensure_built():
    if not os.path.exists(mc.LIB_PATH):
        from minicombust_app_amd import build
        build.build_lib()
    return mc.LIB_PATH


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = _declared_functions()
    for must in ["mc_ctx_create", "mc_ctx_create_box", "mc_ctx_destroy", "mc_timestep", "mc_simple_iteration",
                 "mc_get_phi_gradients", "mc_momentum_assemble", "mc_pressure_assemble", "mc_pressure_solve",
                 "mc_calculate_pressure", "mc_scalar_solve", "mc_mpnn_train_step", "mc_last_error"]:
        assert must in names


def test_library_exports_every_declared_function():
    L = C.CDLL(_ensure_built())
    missing = [n for n in _declared_functions() if not hasattr(L, n)]
    assert not missing, "declared in mcflow.h but not exported: %s" % missing


def _c_layout():
    """sizeof/offsetof of the header's structs, from a program built with gcc."""
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "mcflow.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(mc_options) S(mc_solve_status) S(mc_stats) S(mc_mesh_desc)
  O(mc_options, delta) O(mc_options, p_rtol) O(mc_options, amg_cscale) O(mc_options, check_every)
  O(mc_stats, p_solve_ms) O(mc_stats, host_polls) O(mc_stats, amg_setup_ms) O(mc_solve_status, bnorm)
  O(mc_mesh_desc, cell_faces) O(mc_mesh_desc, boundary_types)
  printf("MC_KC_COUNT %d\n", (int)MC_KC_COUNT);
  printf("MC_ST_COUNT %d\n", (int)MC_ST_COUNT);
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c11", "-I" + os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (ln.split() for ln in out.splitlines())}


def test_ctypes_structs_match_the_header():
    lay = _c_layout()
    assert C.sizeof(mc.Options) == lay["mc_options"]
    assert C.sizeof(mc.SolveStatus) == lay["mc_solve_status"]
    assert C.sizeof(mc.Stats) == lay["mc_stats"]
    assert C.sizeof(mc.MeshDesc) == lay["mc_mesh_desc"]
    assert mc.Options.delta.offset == lay["mc_options.delta"]
    assert mc.Options.p_rtol.offset == lay["mc_options.p_rtol"]
    assert mc.Options.amg_cscale.offset == lay["mc_options.amg_cscale"]
    assert mc.Options.check_every.offset == lay["mc_options.check_every"]
    assert mc.Stats.p_solve_ms.offset == lay["mc_stats.p_solve_ms"]
    assert mc.Stats.host_polls.offset == lay["mc_stats.host_polls"]
    assert mc.Stats.amg_setup_ms.offset == lay["mc_stats.amg_setup_ms"]
    assert mc.SolveStatus.bnorm.offset == lay["mc_solve_status.bnorm"]
    assert mc.MeshDesc.cell_faces.offset == lay["mc_mesh_desc.cell_faces"]
    assert mc.MeshDesc.boundary_types.offset == lay["mc_mesh_desc.boundary_types"]
    assert mc.KC_COUNT == lay["MC_KC_COUNT"]
    assert mc.ST_COUNT == lay["MC_ST_COUNT"]


def test_options_default():
    """mc_options_default: the reference's settings (compat mode, dt 1e-5,
    PETSc BiCGStab rtol 1e-5) and the pressure solve to 1e-8."""
    L = C.CDLL(_ensure_built())
    L.mc_options_default.argtypes = [C.POINTER(mc.Options)]
    o = mc.Options()
    assert L.mc_options_default(C.byref(o)) == mc.MC_OK
    assert o.mode == mc.MODE_COMPAT and o.psolve == mc.PSOLVE_FP64
    assert o.delta == 1e-5 and o.mom_rtol == 1e-5 and o.p_rtol == 1e-8
    assert 8 <= o.amg_coarse_max <= 128 and 1 <= o.amg_nu <= 8
    # the solver settings the measurements were taken with: three
    # Chebyshev-weighted sweeps on [0.5, 2], inner FP32 solves to 1e-5,
    # coarse levels refreshed every 4th solve, FP16 smoothing copy
    assert o.amg_nu == 3 and o.amg_cheb_ratio == 4.0 and o.mixed_inner_rtol == 1e-5
    assert o.amg_refresh == 4 and o.amg_half == 1 and o.amg_tail == 1 and o.pcg_graph == 0
    assert L.mc_options_default(None) == mc.MC_ERR_ARG


def test_null_context_is_an_argument_error():
    """Entry points validate their context instead of crashing."""
    L = C.CDLL(_ensure_built())
    for name in ["mc_timestep", "mc_sync", "mc_get_phi_gradients", "mc_pressure_assemble"]:
        fn = getattr(L, name)
        fn.argtypes = [C.c_void_p]
        assert fn(None) == mc.MC_ERR_ARG, name
    L.mc_ctx_destroy.argtypes = [C.c_void_p]
    assert L.mc_ctx_destroy(None) == mc.MC_OK


@pytest.mark.parametrize("field,value", [("mode", 7), ("psolve", -1), ("amg_nu", 0), ("amg_coarse_max", 4096),
                                         ("cell_tile", 48), ("warm_start", 2), ("amg_half", 3), ("p_rtol", 0.0),
                                         ("inlet_scale", -1.0), ("amg_refresh", 0)])
def test_bad_options_are_refused_before_any_device_call(field, value):
    """Context creation checks the options before it asks for a device: the
    call returns MC_ERR_ARG and leaves no context behind (so this runs on a
    machine with no GPU)."""
    L = C.CDLL(_ensure_built())
    L.mc_options_default.argtypes = [C.POINTER(mc.Options)]
    L.mc_ctx_create_box.argtypes = [C.c_int64, C.POINTER(mc.Options), C.POINTER(C.c_void_p)]
    o = mc.Options()
    assert L.mc_options_default(C.byref(o)) == mc.MC_OK
    setattr(o, field, value)
    h = C.c_void_p(12345)
    assert L.mc_ctx_create_box(4, C.byref(o), C.byref(h)) == mc.MC_ERR_ARG
    assert h.value is None


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
