"""The allocation-free particle-terms entry point (mc_particle_terms_add_noalloc)
against mc_particle_terms_add, through the C-ABI: the same terms array bit for
bit after two calls in a row (the second with fewer rows, so the fixed
scratch still holds the first call's larger contents), and a call above the
fixed scratch size refused with a message.  Runs alone in a fresh process and
after the other GPU tests (the scratch array keeps the last call's contents)."""
import ctypes as C

import numpy as np
import pytest

from minicombust_app_amd import mcflow as mc
from tests.helpers import assert_bits

pytestmark = pytest.mark.gpu


def _rows(rng, N, n):
    cells = rng.integers(0, N, size=n).astype(np.uint64)
    terms = rng.standard_normal((n, 5)) * 10.0 ** rng.integers(-6, 6, size=(n, 5))
    return cells, terms


def test_noalloc_particle_terms_match_the_allocating_entry():
    m = 16
    print("no-alloc particle terms m=%d" % m, flush=True)
    rng = np.random.default_rng(77)
    with mc.Context(m) as ga, mc.Context(m) as gn:
        # before any terms array exists the entry point refuses (it allocates nothing)
        cells, terms = _rows(rng, gn.N, 10)
        rc = gn.L.mc_particle_terms_add_noalloc(gn.h, cells.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                terms.ctypes.data_as(C.POINTER(C.c_double)), cells.size)
        assert rc == mc.MC_ERR_STATE
        assert "mc_particle_terms_reset" in gn.L.mc_last_error(gn.h).decode()
        ga.particle_terms_reset()
        gn.particle_terms_reset()
        for call, n in enumerate([200000, 3001]):
            cells, terms = _rows(rng, ga.N, n)
            # many repeats per cell: each cell's rows are added in receive order
            ga.particle_terms_add(cells, terms)
            gn.particle_terms_add_noalloc(cells, terms)
            assert_bits(gn.get(mc.ARR_PARTICLE_TERMS), ga.get(mc.ARR_PARTICLE_TERMS),
                        "particle terms after call %d (n = %d)" % (call, n))
        # above the fixed scratch: refused with the bytes it would need, terms unchanged
        before = gn.get(mc.ARR_PARTICLE_TERMS)
        big = 2_000_000
        cells = np.zeros(big, dtype=np.uint64)
        terms = np.zeros((big, 5))
        rc = gn.L.mc_particle_terms_add_noalloc(gn.h, cells.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                terms.ctypes.data_as(C.POINTER(C.c_double)), big)
        assert rc == mc.MC_ERR_ARG
        assert "above the fixed" in gn.L.mc_last_error(gn.h).decode()
        assert_bits(gn.get(mc.ARR_PARTICLE_TERMS), before, "particle terms after the refused call")
