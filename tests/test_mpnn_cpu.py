"""CPU tests of the learned-guess oracle (oracle/mpnn_ref.py): its backward
pass against central finite differences, and the bfloat16 rounding helper
against exact cases.  These pin the checker the GPU gradient test uses."""
import numpy as np
import pytest

from oracle import mpnn_ref as mr
from oracle import pyoracle as po


def small_graph(m=2, seed=3):
    M = po.oracle_mesh(m)
    N = M["cells"].shape[0]
    nb = M["cell_neighbours"].astype(np.int64)
    nbr = np.where(nb < N, nb, -1)
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((N, mr.F_IN))
    e = np.where(nbr >= 0, -rng.uniform(0.5, 1.5, size=nbr.shape), 0.0)
    t = rng.standard_normal(N)
    t -= t.mean()
    return f, nbr, e, t


def loss_of(flat, f, nbr, e, t):
    y, _ = mr.forward(flat, f, nbr, e)
    return mr.loss_and_dy(y, t)[0]


def test_param_count():
    assert mr.N_PARAMS == 4 * 64 + 64 + 3 * (5 * 64 * 64 + 64 * 64 + 5 * 64) + 64 + 1


def test_backward_matches_finite_differences():
    f, nbr, e, t = small_graph()
    flat = mr.init_params(5)
    # make biases non-zero so their gradients are exercised
    rng = np.random.default_rng(11)
    flat = flat + 0.05 * rng.standard_normal(flat.size)
    y, cache = mr.forward(flat, f, nbr, e)
    L, dy = mr.loss_and_dy(y, t)
    g = mr.backward(flat, cache, dy)
    idx = rng.choice(flat.size, size=120, replace=False)
    # include every small parameter group
    names = [n for n, _ in mr.PARAM_SHAPES]
    starts = np.cumsum([0] + [int(np.prod(s)) for _, s in mr.PARAM_SHAPES])
    idx = np.unique(np.concatenate([idx, starts[:-1]]))
    eps = 1e-6
    for i in idx:
        fp = flat.copy(); fp[i] += eps
        fm = flat.copy(); fm[i] -= eps
        fd = (loss_of(fp, f, nbr, e, t) - loss_of(fm, f, nbr, e, t)) / (2 * eps)
        k = np.searchsorted(starts, i, side="right") - 1
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (names[k], i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    r = mr.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                          # tie -> even
    assert r[2] == np.float32(1.0 + 4 * 2.0 ** -8)  # tie -> even (up)
    assert r[3] == -2.5
    assert np.isfinite(r[4])
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_bf16_forward_close_to_fp64():
    f, nbr, e, t = small_graph()
    flat = mr.init_params(5)
    y64, _ = mr.forward(flat, f, nbr, e)
    y16, _ = mr.forward(flat, f, nbr, e, bf16_ops=True)
    assert np.max(np.abs(y16 - y64)) <= 0.05 * np.max(np.abs(y64)) + 1e-3
