"""GPU tests of the learned pressure-correction guess through the C-ABI:
forward pass and gradient against the CPU oracle (oracle/mpnn_ref.py, whose
own gradient is pinned by finite differences), and what a short training run
on the GPU gains over the plain solve."""
import numpy as np
import pytest

from minicombust_app_amd import learned
from minicombust_app_amd import mcflow as mc
from oracle import mpnn_ref as mr

pytestmark = pytest.mark.gpu

UINT64_MAX = np.uint64(0xFFFFFFFFFFFFFFFF)


def neighbours(ctx):
    nb = ctx.get(mc.ARR_CELL_NEIGHBOURS).reshape(ctx.N, 6)
    return np.where(nb == UINT64_MAX, -1, nb.astype(np.int64)).astype(np.int64)


def context_with_pairs(m, pairs, w):
    ctx = mc.Context(m, mode=mc.MODE_SIMPLE, p_rtol=1e-8)
    ctx.mpnn_load(w)
    ctx.mpnn_collect(pairs)
    it = 0
    while ctx.mpnn_dataset_size() < pairs:
        ctx.simple_iteration(it % 3)
        it += 1
    ctx.mpnn_collect(0)
    return ctx


def test_parameter_count_matches_oracle():
    assert mc.mpnn_nparams() == mr.N_PARAMS
    assert learned.init_weights().size == mr.N_PARAMS


@pytest.mark.parametrize("m", [4, 8])
def test_forward_matches_oracle(m):
    print("mpnn forward m=%d" % m, flush=True)
    w = mr.init_params(5).astype(np.float32)
    ctx = context_with_pairs(m, 2, w)
    f, e, t = ctx.mpnn_pair(0)
    nbr = neighbours(ctx)
    assert np.all((e != 0) == (nbr >= 0))
    y_gpu = ctx.mpnn_forward_pair(0)
    # the inference kernels keep activations in BF16: against the oracle
    # doing the same, the remaining differences are FP32 summation order and
    # the BF16 roundings it flips (one BF16 step is 2^-8 of a value)
    y_act, _ = mr.forward(w.astype(np.float64), f, nbr, e, bf16_ops=True, bf16_act=True)
    scale = np.max(np.abs(y_act))
    err = np.abs(y_gpu - y_act)
    print("m=%d: inference vs oracle (BF16 activations): max %.2e, mean %.2e of max |y| %.3e"
          % (m, err.max() / scale, err.mean() / scale, scale), flush=True)
    assert err.max() <= 1e-2 * scale and err.mean() <= 1e-3 * scale
    # and against the FP32-activation forward the training pass uses
    y_ref, _ = mr.forward(w.astype(np.float64), f, nbr, e, bf16_ops=True)
    assert np.max(np.abs(y_gpu - y_ref)) <= 2e-2 * np.max(np.abs(y_ref)), np.max(np.abs(y_gpu - y_ref))
    ctx.close()


def test_gradient_matches_oracle():
    print("mpnn gradient m=4", flush=True)
    rng = np.random.default_rng(3)
    w = (mr.init_params(5) + 0.02 * rng.standard_normal(mr.N_PARAMS)).astype(np.float32)
    ctx = context_with_pairs(4, 1, w)
    f, e, t = ctx.mpnn_pair(0)
    nbr = neighbours(ctx)
    g_gpu, loss_gpu = ctx.mpnn_grad(0)
    y, cache = mr.forward(w.astype(np.float64), f, nbr, e, bf16_ops=True)
    loss_ref, dy = mr.loss_and_dy(y.astype(np.float64), t.astype(np.float64))
    g_ref = mr.backward(w.astype(np.float64), cache, dy)
    assert abs(loss_gpu - loss_ref) <= 2e-2 * abs(loss_ref)
    starts = np.cumsum([0] + [int(np.prod(s)) for _, s in mr.PARAM_SHAPES])
    for k, (name, _) in enumerate(mr.PARAM_SHAPES):
        a, b = g_gpu[starts[k]:starts[k + 1]].astype(np.float64), g_ref[starts[k]:starts[k + 1]]
        nb_ = np.linalg.norm(b)
        if nb_ < 1e-6 * np.linalg.norm(g_ref):
            continue
        rel = np.linalg.norm(a - b) / nb_
        assert rel <= 6e-2, (name, rel)
    cos = float(g_gpu @ g_ref / (np.linalg.norm(g_gpu) * np.linalg.norm(g_ref)))
    assert cos >= 0.998, cos
    ctx.close()


def test_training_gains_over_plain_solve():
    """Train on pairs from inlet-perturbed m = 16 runs, then solve held-out
    systems (another inlet scale) from the untrained net's guess, the trained
    net's guess and zero: the trained guess must leave a smaller residual
    (||b - A x0|| / ||b - mean b||, mean over the solves that used it) than the
    untrained one, below 1 (better than x0 = 0), meet the same tolerance and
    not need more PCG iterations than the plain solve."""
    print("mpnn training m=16", flush=True)
    w0 = learned.init_weights()
    ctx = mc.Context(16, mode=mc.MODE_SIMPLE, p_rtol=1e-8)
    ctx.mpnn_load(w0)
    learned.generate(ctx, 120, seed=11, iters_per_case=10)
    held0 = [ctx.mpnn_grad(p)[1] for p in range(108, 120)]
    losses, held1 = learned.train(ctx, 600, lr=2e-3, seed=5, holdout=0.1)
    print("held-out loss %.4f -> %.4f" % (np.mean(held0), held1), flush=True)
    assert held1 < 0.8 * np.mean(held0)
    w = ctx.mpnn_save()
    ctx.close()

    def run(weights, use):
        c = mc.Context(16, mode=mc.MODE_SIMPLE, p_rtol=1e-8, inlet_scale=1.13, learned=2 if use else 0)
        c.mpnn_load(weights)
        for it in range(9):
            c.simple_iteration(it % 3)
        st = c.stats()
        c.close()
        return st

    off, untrained, trained = run(w, False), run(w0, True), run(w, True)
    ratio = lambda st: st["p_guess_ratio_sum"] / max(st["p_guesses"], 1)
    print("plain: %d PCG iterations in %d solves | untrained guess: %d its, %d used, residual ratio %.4f | "
          "trained guess: %d its, %d used, residual ratio %.4f"
          % (off["p_iters"], off["p_solves"], untrained["p_iters"], untrained["p_guesses"], ratio(untrained),
             trained["p_iters"], trained["p_guesses"], ratio(trained)), flush=True)
    assert trained["p_guesses"] >= 1
    assert ratio(trained) < 1.0
    assert trained["p_guesses"] == 0 or untrained["p_guesses"] == 0 or ratio(trained) < ratio(untrained)
    assert trained["p_relres_max"] <= 1e-8 * 1.0001
    assert trained["p_iters"] <= off["p_iters"]
