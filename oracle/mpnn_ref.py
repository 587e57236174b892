"""CPU oracle of the learned pressure-correction guess (TEST INFRASTRUCTURE
ONLY: tests and the benchmark's baseline use it as a checker; the library
never calls it).

The net has no counterpart in the reference; its definition lives in the HIP
library (minicombust_app_amd/csrc/mc_mpnn.hip) and this file states the same
function in numpy so the GPU forward and backward can be checked.  Its own
backward is pinned by finite differences (tests/test_mpnn_cpu.py).

Graph: cells are nodes; each cell has up to six neighbours (side order, -1 at
the boundary); messages run both ways over every interior face.

    features f_i (4): (b_i - mean b) / s_b, A_ii / s_d, V_i / mean V, q_i / s_q
        (b: pressure-correction RHS, A_ii: its diagonal, q: net mass outflow,
         s_b = rms(b - mean b), s_q = rms(q), s_d = mean(A_ii));
        edge feature e_ij = A_ij / s_d
    h = relu(f Wenc + benc)
    3 layers:  P = h Wa,  Q = h Wb
               s_i = sum_j relu(P_i + Q_j + e_ij we + b1)
               a_i = s_i W2 + deg_i b2
               u_i = relu([h_i, a_i] Wn1 + bn1)
               h_i = h_i + u_i Wn2 + bn2
    y = h wdec + bdec          (the guess, up to the solve's Galerkin scaling)
    loss = mean((y - mean(y) - t)^2),  t = x* / rms(x*) (x* mean-zero)

GEMM operands are rounded to bfloat16 (round to nearest even) with float32
accumulation when `bf16_ops=True`, as the MFMA kernels compute them; with
`bf16_act=True` the activations h are also stored in bfloat16 after the
encoder and after every layer (the residual adds the stored value), as the
fused inference kernels keep them.
"""
import numpy as np

H = 64
F_IN = 4
LAYERS = 3

# parameter order of the flat weight vector (shapes for H = 64)
PARAM_SHAPES = [("Wenc", (F_IN, H)), ("benc", (H,))]
for _l in range(LAYERS):
    PARAM_SHAPES += [("Wa%d" % _l, (H, H)), ("Wb%d" % _l, (H, H)), ("we%d" % _l, (H,)), ("b1%d" % _l, (H,)),
                     ("W2%d" % _l, (H, H)), ("b2%d" % _l, (H,)), ("Wn1%d" % _l, (2 * H, H)), ("bn1%d" % _l, (H,)),
                     ("Wn2%d" % _l, (H, H)), ("bn2%d" % _l, (H,))]
PARAM_SHAPES += [("wdec", (H,)), ("bdec", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)


def unpack(flat):
    out, k = {}, 0
    for name, shp in PARAM_SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(flat[k:k + n], dtype=np.float64).reshape(shp)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in PARAM_SHAPES])


def init_params(seed=1234):
    """Deterministic initial weights (He-uniform for ReLU layers, small decoder)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in PARAM_SHAPES:
        if len(shp) == 2:
            lim = np.sqrt(6.0 / shp[0])
            d[name] = rng.uniform(-lim, lim, size=shp)
        else:
            d[name] = np.zeros(shp)
    for l in range(LAYERS):
        d["we%d" % l] = rng.uniform(-1.0, 1.0, size=(H,))
        d["Wn2%d" % l] *= 0.1
    d["wdec"] = rng.uniform(-0.1, 0.1, size=(H,))
    return pack(d)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def _mm(x, w, lowp):
    if lowp:
        return (bf16(x).astype(np.float64) @ bf16(w).astype(np.float64)).astype(np.float32)
    return x @ w


def features(b, diag, vol, q):
    bc = b - b.mean()
    s_b = np.sqrt(np.mean(bc * bc))
    s_b = s_b if s_b > 0 else 1.0
    s_q = np.sqrt(np.mean(q * q))
    s_q = s_q if s_q > 0 else 1.0
    s_d = np.mean(diag)
    f = np.stack([bc / s_b, diag / s_d, vol / np.mean(vol), q / s_q], axis=1)
    return f, s_b, s_d


def forward(flat, f, nbr, e, bf16_ops=False, bf16_act=False):
    """f: (N, 4) features; nbr: (N, 6) neighbour ids (-1 none); e: (N, 6) edge
    features.  Returns y (N,) and the cache the backward needs."""
    p = unpack(flat)
    dt = np.float32 if bf16_ops else np.float64
    f = f.astype(dt)
    e = e.astype(dt)
    mask = nbr >= 0
    nb = np.where(mask, nbr, 0)
    deg = mask.sum(axis=1).astype(dt)
    c = {"f": f, "mask": mask, "nb": nb, "deg": deg, "e": e}
    h = np.maximum(_mm(f, p["Wenc"].astype(dt), bf16_ops) + p["benc"].astype(dt), 0)
    if bf16_act:
        h = bf16(h)
    c["h0"] = h
    for l in range(LAYERS):
        P = _mm(h, p["Wa%d" % l].astype(dt), bf16_ops)
        Q = _mm(h, p["Wb%d" % l].astype(dt), bf16_ops)
        z = P[:, None, :] + Q[nb] + e[:, :, None] * p["we%d" % l].astype(dt) + p["b1%d" % l].astype(dt)
        r = np.where(mask[:, :, None], np.maximum(z, 0), 0)
        s = r.sum(axis=1)
        a = _mm(s, p["W2%d" % l].astype(dt), bf16_ops) + deg[:, None] * p["b2%d" % l].astype(dt)
        ha = np.concatenate([h, a], axis=1)
        zu = _mm(ha, p["Wn1%d" % l].astype(dt), bf16_ops) + p["bn1%d" % l].astype(dt)
        u = np.maximum(zu, 0)
        hn = h + _mm(u, p["Wn2%d" % l].astype(dt), bf16_ops) + p["bn2%d" % l].astype(dt)
        c[l] = {"h": h, "P": P, "Q": Q, "z": z, "s": s, "ha": ha, "zu": zu, "u": u}
        h = bf16(hn) if bf16_act else hn
    c["hL"] = h
    y = _mm(h, p["wdec"].astype(dt)[:, None], bf16_ops)[:, 0] + p["bdec"].astype(dt)[0]
    return y, c


def loss_and_dy(y, t):
    d = y - y.mean() - t
    L = float(np.mean(d * d))
    dy = (2.0 / y.size) * d
    dy = dy - dy.mean()
    return L, dy


def backward(flat, c, dy):
    """Gradient of sum(dy * y) with respect to the flat parameters (float64)."""
    p = unpack(flat)
    g = {k: np.zeros_like(v) for k, v in p.items()}
    mask, nb, deg, e = c["mask"], c["nb"], c["deg"], c["e"]
    h = c["hL"].astype(np.float64)
    dy = dy.astype(np.float64)
    g["wdec"] = h.T @ dy
    g["bdec"] = np.array([dy.sum()])
    dh = dy[:, None] * p["wdec"][None, :]
    for l in reversed(range(LAYERS)):
        cl = c[l]
        hin = cl["h"].astype(np.float64)
        # h' = h + u Wn2 + bn2
        g["Wn2%d" % l] = cl["u"].astype(np.float64).T @ dh
        g["bn2%d" % l] = dh.sum(axis=0)
        du = dh @ p["Wn2%d" % l].T
        dzu = du * (cl["zu"] > 0)
        g["Wn1%d" % l] = cl["ha"].astype(np.float64).T @ dzu
        g["bn1%d" % l] = dzu.sum(axis=0)
        dha = dzu @ p["Wn1%d" % l].T
        dh_new = dh + dha[:, :H]
        da = dha[:, H:]
        g["W2%d" % l] = cl["s"].astype(np.float64).T @ da
        g["b2%d" % l] = (deg[:, None] * da).sum(axis=0)
        ds = da @ p["W2%d" % l].T
        G = np.where(mask[:, :, None] & (cl["z"] > 0), ds[:, None, :], 0.0)   # (N, 6, H)
        dP = G.sum(axis=1)
        dQ = np.zeros_like(dP)
        np.add.at(dQ, nb[mask], G[mask])
        g["we%d" % l] = (G * e[:, :, None]).sum(axis=(0, 1))
        g["b1%d" % l] = G.sum(axis=(0, 1))
        g["Wa%d" % l] = hin.T @ dP
        g["Wb%d" % l] = hin.T @ dQ
        dh = dh_new + dP @ p["Wa%d" % l].T + dQ @ p["Wb%d" % l].T
    dz0 = dh * (c["h0"] > 0)
    g["Wenc"] = c["f"].astype(np.float64).T @ dz0
    g["benc"] = dz0.sum(axis=0)
    return pack(g)
