"""NumPy oracle of the learned preconditioner's graph MLP (test oracle only).

No reference counterpart (the reference has no learned component).  This is
the CPU statement the HIP BF16 forward/backward kernels are checked against;
the oracle itself is pinned by finite differences of its loss
(tests/test_learned_oracle.py).

Network (per node i of a mesh graph N(i) = tet-edge neighbours, self
excluded):
    x_i  = poly_features(...)_i (16 values of D, see there)
    h0   = relu(x W0 + b0)                       W0: 16 x 64
    m1_i = mean_{j in N(i)} h0_j
    h1   = relu(h0 Ws1 + m1 Wn1 + b1)            64 x 64 each
    m2   = mean_{j in N(i)} h1_j
    h2   = relu(h1 Ws2 + m2 Wn2 + b2)
    o    = h2 Wo + bo                            Wo: 64 x 3 (padded to 64 x 32 on the GPU)
The node mean of o predicts log10(b/a) - 2.5 of the Chebyshev interval of
the block-Jacobi passes at degrees 8, 12, 16 (predict_log_ratio).
Loss of a batch: mean over samples of sum_i |o_i - t_i|^2 / den.

With ``bf16=True`` the weights and every stored activation (x, h0, m1, h1,
m2, h2) are rounded to bfloat16 (round to nearest even) exactly where the
GPU kernels store them; sums are exact (float64 here, float32 on the GPU).
"""
from __future__ import annotations

import numpy as np

HID = 64
IN = 16
OUT = 3
PARAM_SHAPES = [("W0", (IN, HID)), ("b0", (HID,)), ("Ws1", (HID, HID)), ("Wn1", (HID, HID)), ("b1", (HID,)),
                ("Ws2", (HID, HID)), ("Wn2", (HID, HID)), ("b2", (HID,)), ("Wo", (HID, OUT)), ("bo", (OUT,))]


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    r = ((u + rounding) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isfinite(a), out, a.astype(np.float64))


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in PARAM_SHAPES:
        if name.startswith("b"):
            p[name] = np.zeros(shp)
        else:
            fan_in = shp[0] if name != "W0" else 14
            p[name] = rng.normal(size=shp) * np.sqrt(2.0 / fan_in) * (0.1 if name == "Wo" else 1.0)
    p["W0"][14:] = 0.0
    return p


def graph_from_bsr(rowptr, colidx):
    """Neighbour CSR without the diagonal, from a full-storage BSR pattern."""
    nbr_ptr, nbr = [0], []
    for i in range(len(rowptr) - 1):
        cols = [c for c in colidx[rowptr[i]:rowptr[i + 1]] if c != i]
        nbr.extend(cols)
        nbr_ptr.append(len(nbr))
    return np.asarray(nbr_ptr, dtype=np.int64), np.asarray(nbr, dtype=np.int64)


def poly_features(D, active_nodes):
    """Per-node input of the learned polynomial from the full symmetric FEM
    block D (SciPy sparse, 3x3 node blocks) and the coupled nodes:
    x_i = [D_ii / s_i (9, row-major), log10(s_i / mean_{j in N(i)} s_j) / 2,
    sum_{j in N(i)} |D_ij|_F / |D_ii|_F / 4, |N(i)| / 16, coupled,
    log10(nb) / 6, 0, 0], s_i = tr(D_ii) / 3,
    N(i) = the other block columns stored in block row i.  Returns x (nb, 16)
    and the neighbour CSR (nbr_ptr, nbr)."""
    Db = D.tobsr(blocksize=(3, 3))
    Db.sort_indices()
    nb = Db.shape[0] // 3
    ptr, idx, blk = Db.indptr, Db.indices, Db.data
    x = np.zeros((nb, IN))
    diag = np.zeros((nb, 3, 3))
    for i in range(nb):
        for p in range(ptr[i], ptr[i + 1]):
            if idx[p] == i:
                diag[i] = blk[p]
    s = np.trace(diag, axis1=1, axis2=2) / 3.0
    s = np.where(s > 0, s, 1.0)
    act = np.zeros(nb, dtype=bool)
    act[np.asarray(active_nodes, dtype=np.int64)] = True
    nbr_ptr, nbr = [0], []
    for i in range(nb):
        cols = [int(idx[p]) for p in range(ptr[i], ptr[i + 1]) if idx[p] != i]
        offn = sum(np.linalg.norm(blk[p]) for p in range(ptr[i], ptr[i + 1]) if idx[p] != i)
        nbr.extend(cols)
        nbr_ptr.append(len(nbr))
        deg = len(cols)
        x[i, :9] = (diag[i] / s[i]).ravel()
        ms = np.mean(s[cols]) if deg else 0.0
        x[i, 9] = 0.5 * np.log10(s[i] / ms) if deg and ms > 0 else 0.0
        dn = np.linalg.norm(diag[i])
        x[i, 10] = 0.25 * offn / dn if dn > 0 else 0.0
        x[i, 11] = deg / 16.0
        x[i, 12] = 1.0 if act[i] else 0.0
        x[i, 13] = np.log10(nb) / 6.0
    return x, np.asarray(nbr_ptr, dtype=np.int64), np.asarray(nbr, dtype=np.int64)


def predict_log_ratio(p, x, nbr_ptr, nbr, degree, bf=True):
    """log10(b/a) the network picks for `degree` passes: the node means of its
    3 outputs are log10(b/a) - 2.5 at degrees 8, 12 and 16; other degrees
    interpolate linearly between neighbouring channels (clamped to [8, 16]).
    Returns (log10 ratio, channel means)."""
    o, _ = forward(p, x, nbr_ptr, nbr, bf=bf)
    u = o.reshape(-1, OUT).mean(axis=0)
    t = min(max((degree - 8) / 4.0, 0.0), 2.0)
    c0 = int(np.floor(t))
    c1 = min(c0 + 1, 2)
    return float((1 - (t - c0)) * u[c0] + (t - c0) * u[c1] + 2.5), u


def aggregate(h, nbr_ptr, nbr):
    """mean over neighbours along axis -2 (nodes)."""
    deg = np.diff(nbr_ptr)
    out = np.zeros_like(h)
    for i in range(len(deg)):
        if deg[i]:
            out[..., i, :] = h[..., nbr[nbr_ptr[i]:nbr_ptr[i + 1]], :].sum(axis=-2) / deg[i]
    return out


def aggregate_T(g, nbr_ptr, nbr):
    """adjoint of aggregate: out_j = sum_{i: j in N(i)} g_i / deg_i."""
    deg = np.diff(nbr_ptr)
    out = np.zeros_like(g)
    for i in range(len(deg)):
        if deg[i]:
            out[..., nbr[nbr_ptr[i]:nbr_ptr[i + 1]], :] += g[..., i:i + 1, :] / deg[i]
    return out


def forward(p, x, nbr_ptr, nbr, bf=False):
    R = bf16 if bf else (lambda v: v)
    P = {k: R(v) for k, v in p.items()} if bf else p
    x = R(x)
    a0 = x @ P["W0"] + P["b0"]
    h0 = R(np.maximum(a0, 0))
    m1 = R(aggregate(h0, nbr_ptr, nbr))
    a1 = h0 @ P["Ws1"] + m1 @ P["Wn1"] + P["b1"]
    h1 = R(np.maximum(a1, 0))
    m2 = R(aggregate(h1, nbr_ptr, nbr))
    a2 = h1 @ P["Ws2"] + m2 @ P["Wn2"] + P["b2"]
    h2 = R(np.maximum(a2, 0))
    o = h2 @ P["Wo"] + P["bo"]
    cache = dict(x=x, h0=h0, m1=m1, h1=h1, m2=m2, h2=h2, P=P)
    return o, cache


def loss_and_grad(p, x, t, nbr_ptr, nbr, bf=False):
    """x: (B, nb, 16), t: (B, nb, 3). Returns loss, grads (dict), o."""
    o, c = forward(p, x, nbr_ptr, nbr, bf)
    B = x.shape[0]
    den = np.sum(t ** 2, axis=(1, 2))
    den = np.where(den > 0, den, 1.0)
    diff = o - t
    loss = float(np.mean(np.sum(diff ** 2, axis=(1, 2)) / den))
    do = 2.0 * diff / den[:, None, None] / B
    P = c["P"]
    g = {}
    g["Wo"] = np.einsum("bni,bnj->ij", c["h2"], do)
    g["bo"] = do.sum(axis=(0, 1))
    dh2 = do @ P["Wo"].T
    da2 = dh2 * (c["h2"] > 0)
    g["Ws2"] = np.einsum("bni,bnj->ij", c["h1"], da2)
    g["Wn2"] = np.einsum("bni,bnj->ij", c["m2"], da2)
    g["b2"] = da2.sum(axis=(0, 1))
    dh1 = da2 @ P["Ws2"].T + aggregate_T(da2 @ P["Wn2"].T, nbr_ptr, nbr)
    da1 = dh1 * (c["h1"] > 0)
    g["Ws1"] = np.einsum("bni,bnj->ij", c["h0"], da1)
    g["Wn1"] = np.einsum("bni,bnj->ij", c["m1"], da1)
    g["b1"] = da1.sum(axis=(0, 1))
    dh0 = da1 @ P["Ws1"].T + aggregate_T(da1 @ P["Wn1"].T, nbr_ptr, nbr)
    da0 = dh0 * (c["h0"] > 0)
    g["W0"] = np.einsum("bni,bnj->ij", c["x"], da0)
    g["b0"] = da0.sum(axis=(0, 1))
    return loss, g, o


def adam_step(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] ** 2
        mh = m[k] / (1 - b1 ** step)
        vh = v[k] / (1 - b2 ** step)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p, m, v
