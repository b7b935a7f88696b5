"""Synthetic coupled ABD/FEM Newton systems in the plugin's input format.

The HSC plugin receives the assembled Hessian as an upper-triangle BCOO of
3x3 blocks (libuipc's ``SolvingInfo::A()``, integration.patch:7,17-20 of the
reference) plus a right-hand side.  This module builds such systems for the
benchmark workloads:

* ``cube_system``   -- one FEM block on a structured grid split 5 tets/cube,
  coupled to ABD bodies (the 2,178-DOF correctness system).
* ``torus_system``  -- several tire-like tori (hollow torus, Tet4) with rim
  couplings to ABD hubs (the 78k-tet vehicle-scale system).
* ``batch_systems`` -- many uneven torus systems (the batched workload).

FEM Hessian: per-tet Neo-Hookean tangent at the rest state (identical to the
linear-elastic tangent, C = mu (d_ik d_jl + d_il d_jk) + lambda d_ij d_kl),
projected SPD by eigenvalue clamping (the recipe of the reference's
log_neo_hookean_function.h:106-150), pulled back to 3x3 node blocks.  Barrier
terms kappa*I3 (kappa log-uniform) sit on chosen nodes.

Coupling: each coupled body (hub) has a dense 12 x 3*n_rim block of the
requested rank.  A penalty coupling 0.5*k_s*|x - J q|^2 puts k_s*I3 on the
coupled nodes' diagonal blocks; the coupling block is scaled so that
||E_h||_2^2 = 0.5 * k_s * lambda_min(A_h), which keeps the Schur complement
D - E^T A^-1 E >= D - 0.5 k_s I_rim positive definite (H is SPD).

DOF order: FEM first (3 per node), then ABD (12 per body) -- the order of the
reference's NumPy method (core.py:161-167).  The plugin boundary accepts any
offsets.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class HSCSystem:
    """A coupled system in plugin input format.

    rows, cols: int32 block indices (one stored block per unordered pair,
    rows <= cols); vals: (nnz, 9) float64, each 3x3 block column-major
    (Eigen's default, as libuipc stores it).
    """

    rows: np.ndarray
    cols: np.ndarray
    vals: np.ndarray
    b: np.ndarray
    fo: int
    fn: int
    qo: int
    qm: int
    info: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.fn + self.qm

    @property
    def nnz(self) -> int:
        return int(self.rows.shape[0])

    def dense(self) -> np.ndarray:
        """Full symmetric H (small systems only)."""
        N = self.N
        H = np.zeros((N, N))
        for r, c, v in zip(self.rows, self.cols, self.vals):
            B = v.reshape(3, 3).T  # column-major -> row-major
            H[3 * r:3 * r + 3, 3 * c:3 * c + 3] += B
            if r != c:
                H[3 * c:3 * c + 3, 3 * r:3 * r + 3] += B.T
        return H

    def sparse(self):
        """Full symmetric H as scipy CSR."""
        import scipy.sparse as sp
        B = self.vals.reshape(-1, 3, 3).transpose(0, 2, 1)  # row-major blocks
        ii = (3 * self.rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2)
        jj = (3 * self.cols[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1)
        off = self.rows != self.cols
        data = np.concatenate([B.ravel(), B[off].transpose(0, 2, 1).ravel()])
        ri = np.concatenate([ii.ravel(), jj[off].transpose(0, 2, 1).ravel()])
        ci = np.concatenate([jj.ravel(), ii[off].transpose(0, 2, 1).ravel()])
        return sp.csr_matrix((data, (ri, ci)), shape=(self.N, self.N))


# ---------------------------------------------------------------- meshes

def _hex_tets5(c):
    """Split hex with corner ids c[8] (c[dx+2dy+4dz]) into 5 tets; the two
    mirror patterns alternate with cell parity so neighbouring faces match."""
    even = np.array([[0, 1, 2, 4], [1, 3, 2, 7], [1, 4, 5, 7], [2, 4, 7, 6], [1, 2, 4, 7]])
    odd = np.array([[1, 0, 3, 5], [0, 2, 3, 6], [0, 5, 4, 6], [3, 5, 6, 7], [0, 3, 5, 6]])
    return even, odd


def box_mesh(nx, ny, nz, h):
    """Structured (nx,ny,nz)-cube grid of spacing h, 5 tets per cube."""
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nodes = np.stack([ix, iy, iz], -1).reshape(-1, 3).astype(np.float64) * h
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    corners = np.stack([nid(ci + dx, cj + dy, ck + dz)
                        for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)], -1)
    even, odd = _hex_tets5(None)
    par = (ci + cj + ck) % 2
    tets = np.where(par[:, None, None] == 0, corners[:, even], corners[:, odd]).reshape(-1, 4)
    return nodes, tets


def torus_mesh(R, r_out, r_in, n_theta, n_phi, n_rho, center=(0.0, 0.0, 0.0)):
    """Hollow (tire-like) torus: major radius R, wall from r_in to r_out.

    Structured in (theta, phi, rho), periodic in theta and phi (both counts
    must be even for the alternating 5-tet split), 5 tets per cell.
    Axis of revolution = z.  Returns nodes, tets, outer_surface_ids.
    """
    if n_theta % 2 or n_phi % 2:
        raise ValueError("n_theta and n_phi must be even")
    th = 2 * np.pi * np.arange(n_theta) / n_theta
    ph = 2 * np.pi * np.arange(n_phi) / n_phi
    rh = np.linspace(r_in, r_out, n_rho + 1)
    T, P, Rr = np.meshgrid(th, ph, rh, indexing="ij")
    rad = R + Rr * np.cos(P)
    nodes = np.stack([rad * np.cos(T), rad * np.sin(T), Rr * np.sin(P)], -1).reshape(-1, 3)
    nodes = nodes + np.asarray(center)[None, :]
    nid = lambda i, j, k: ((i % n_theta) * n_phi + (j % n_phi)) * (n_rho + 1) + k  # noqa: E731
    ci, cj, ck = np.meshgrid(np.arange(n_theta), np.arange(n_phi), np.arange(n_rho), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    corners = np.stack([nid(ci + dx, cj + dy, ck + dz)
                        for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)], -1)
    even, odd = _hex_tets5(None)
    par = (ci + cj + ck) % 2
    tets = np.where(par[:, None, None] == 0, corners[:, even], corners[:, odd]).reshape(-1, 4)
    # orient every tet positively
    X = nodes[tets]
    vol = np.einsum("ij,ij->i", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0])
    neg = vol < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    I, J, K = np.meshgrid(np.arange(n_theta), np.arange(n_phi), np.arange(n_rho + 1), indexing="ij")
    outer = np.flatnonzero((K == n_rho).ravel())
    return nodes, tets, outer


# ---------------------------------------------------------------- FEM Hessian

def neo_hookean_rest_tangent(E, nu):
    """9x9 d2Psi/dF2 of Neo-Hookean at F = I, projected SPD (eigenvalue clamp)."""
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    d = np.eye(3)
    C = (mu * np.einsum("ik,jl->ijkl", d, d) + mu * np.einsum("il,jk->ijkl", d, d)
         + lam * np.einsum("ij,kl->ijkl", d, d))
    H = C.reshape(9, 9)
    w, Q = np.linalg.eigh(0.5 * (H + H.T))
    return (Q * np.maximum(w, 0.0)) @ Q.T


def fem_blocks(nodes, tets, E, nu):
    """Assembled node-block stiffness: returns (I, J, B[k,3,3]) with every
    ordered pair (I, J) present once (full symmetric pattern)."""
    X = nodes[tets]                                   # (t,4,3)
    Dm = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], -1)  # columns
    vol = np.abs(np.linalg.det(Dm)) / 6.0
    Dinv = np.linalg.inv(Dm)                          # (t,3,3)
    dN = np.empty((len(tets), 4, 3))
    dN[:, 1:, :] = Dinv                               # row a-1 of Dm^-1 = grad N_a
    dN[:, 0, :] = -Dinv.sum(axis=1)
    HF = neo_hookean_rest_tangent(E, nu).reshape(3, 3, 3, 3)
    K = np.einsum("t,taj,ijkl,tbl->taibk", vol, dN, HF, dN)   # (t,4,3,4,3)
    a_idx, b_idx = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    I = tets[:, a_idx.ravel()].ravel()
    J = tets[:, b_idx.ravel()].ravel()
    Bk = K.transpose(0, 1, 3, 2, 4).reshape(len(tets), 16, 3, 3).reshape(-1, 3, 3)
    nn = len(nodes)
    key = I.astype(np.int64) * nn + J
    uk, inv = np.unique(key, return_inverse=True)
    B = np.zeros((len(uk), 9))
    for c in range(9):
        B[:, c] = np.bincount(inv, weights=Bk.reshape(-1, 9)[:, c], minlength=len(uk))
    return (uk // nn).astype(np.int64), (uk % nn).astype(np.int64), B.reshape(-1, 3, 3)


def _log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def _spd_body(rng):
    M = rng.normal(size=(12, 12))
    return M @ M.T + 12.0 * np.eye(12)


def _coupling_block(rng, rank, n_cols, k_s, A_body, noise_rel=0.0):
    """12 x n_cols block of exact rank `rank` (+ relative noise), scaled so
    that ||E||_2^2 = 0.5 * k_s * lambda_min(A_body)."""
    G = rng.normal(size=(12, rank))
    F = rng.normal(size=(n_cols, rank))
    E0 = G @ F.T
    s = np.sqrt(0.5 * k_s * np.linalg.eigvalsh(A_body)[0]) / np.linalg.norm(E0, 2)
    E = s * E0
    if noise_rel:
        E = E + noise_rel * np.abs(E).max() * rng.normal(size=E.shape)
        E *= np.sqrt(0.5 * k_s * np.linalg.eigvalsh(A_body)[0]) / np.linalg.norm(E, 2)
    return E


# ---------------------------------------------------------------- assembly

def _to_bcoo(nb_fem, fem_I, fem_J, fem_B, bodies, couplings, rng_b, N):
    """Build the upper-triangle BCOO (row <= col) of
    H = [[D, E^T], [E, A]] with FEM node blocks first, ABD block rows after.

    bodies: list of 12x12 SPD; couplings: list of (body, fem_node_ids, E_blk
    (12 x 3*len(ids))).
    """
    keep = fem_I <= fem_J
    rows = [fem_I[keep]]
    cols = [fem_J[keep]]
    vals = [fem_B[keep].transpose(0, 2, 1).reshape(-1, 9)]  # row-major -> col-major
    q0 = nb_fem
    for t, Ab in enumerate(bodies):
        for bi in range(4):
            for bj in range(bi, 4):
                blk = Ab[3 * bi:3 * bi + 3, 3 * bj:3 * bj + 3]
                rows.append(np.array([q0 + 4 * t + bi]))
                cols.append(np.array([q0 + 4 * t + bj]))
                vals.append(blk.T.reshape(1, 9))
    for body, ids, Eb in couplings:
        # H[fem node f, ABD block row q] = (E[q rows, f cols])^T
        for bi in range(4):
            blk = Eb[3 * bi:3 * bi + 3, :].reshape(3, len(ids), 3)  # (q,f,c)
            HT = blk.transpose(1, 2, 0)                             # (f, fem comp, q comp)
            rows.append(np.asarray(ids, dtype=np.int64))
            cols.append(np.full(len(ids), q0 + 4 * body + bi))
            vals.append(HT.transpose(0, 2, 1).reshape(-1, 9))       # col-major
    r = np.concatenate(rows).astype(np.int32)
    c = np.concatenate(cols).astype(np.int32)
    v = np.concatenate(vals).astype(np.float64)
    order = np.lexsort((c, r))
    return r[order], c[order], np.ascontiguousarray(v[order])


def _finish(nb_fem, fem_I, fem_J, fem_B, bodies, couplings, rng, info):
    n = 3 * nb_fem
    m = 12 * len(bodies)
    N = n + m
    r, c, v = _to_bcoo(nb_fem, fem_I, fem_J, fem_B, bodies, couplings, rng, N)
    b = rng.normal(size=N)
    nnzb_full = int(np.sum(fem_I >= 0))
    info.update(nb=nb_fem, n=n, m=m, N=N, nnzb_D=nnzb_full, bcoo_nnz=int(len(r)),
                active_nodes=int(sum(len(ids) for _, ids, _ in couplings)))
    return HSCSystem(r, c, v, b, fo=0, fn=n, qo=n, qm=m, info=info)


def _add_diag(fem_I, fem_J, fem_B, node_ids, kappa):
    diag = np.flatnonzero(fem_I == fem_J)
    pos = np.empty(fem_I.max() + 1, dtype=np.int64)
    pos[fem_I[diag]] = diag
    fem_B[pos[node_ids]] += kappa[:, None, None] * np.eye(3)[None]


def cube_system(seed=0, nx=10, ny=10, nz=5, h=0.02, E=1e8, nu=0.45,
                barrier_frac=0.10, barrier=(1e3, 1e6), n_rim_side=8, rank=6,
                noise_rel=1e-8, n_bodies=1):
    """The 2,178-DOF correctness system: 10x10x5 cubes of 0.02 m, 5 tets per
    cube (2,500 Tet4, 726 nodes), one ABD body coupled through a dense
    12 x (3*64) rank-6 (+1e-8 noise) block on an 8x8 rim patch of the top
    face, barrier kappa*I3 on 10% of the nodes."""
    rng = np.random.default_rng(seed)
    nodes, tets = box_mesh(nx, ny, nz, h)
    nb = len(nodes)
    I, J, B = fem_blocks(nodes, tets, E, nu)
    nbar = int(round(barrier_frac * nb))
    bar_ids = np.sort(rng.choice(nb, nbar, replace=False))
    _add_diag(I, J, B, bar_ids, _log_uniform(rng, *barrier, nbar))
    # rim: 8x8 patch in the middle of the top face (k = nz)
    off = (nx + 1 - n_rim_side) // 2
    ii, jj = np.meshgrid(np.arange(off, off + n_rim_side), np.arange(off, off + n_rim_side), indexing="ij")
    rim = ((ii * (ny + 1) + jj) * (nz + 1) + nz).ravel()
    k_s = E * h
    _add_diag(I, J, B, rim, np.full(len(rim), k_s))
    bodies = [_spd_body(rng) for _ in range(n_bodies)]
    Eb = _coupling_block(rng, rank, 3 * len(rim), k_s, bodies[0], noise_rel)
    info = dict(kind="cube", tets=int(len(tets)), E=E, nu=nu, seed=seed, rank=rank, k_s=k_s)
    return _finish(nb, I, J, B, bodies, [(0, rim, Eb)], rng, info)


def _torus_block(rng, R, r_out, r_in, n_theta, n_phi, n_rho, E, nu, n_rim,
                 barrier_frac, barrier, center):
    nodes, tets, outer = torus_mesh(R, r_out, r_in, n_theta, n_phi, n_rho, center)
    I, J, B = fem_blocks(nodes, tets, E, nu)
    # rim = outer-surface nodes nearest the axis of revolution (inner equator)
    rad = np.hypot(nodes[outer, 0] - center[0], nodes[outer, 1] - center[1])
    rim = np.sort(outer[np.argsort(rad, kind="stable")[:n_rim]])
    rest = np.setdiff1d(outer, rim)
    nbar = max(1, int(round(barrier_frac * len(outer))))
    bar = np.sort(rng.choice(rest, nbar, replace=False))
    h = 2 * np.pi * R / n_theta
    k_s = E * h
    _add_diag(I, J, B, bar, _log_uniform(rng, *barrier, nbar))
    _add_diag(I, J, B, rim, np.full(len(rim), k_s))
    return nodes, tets, I, J, B, rim, k_s


def torus_system(seed=1, n_tori=7, n_bodies=12, youngs=(1e10,) * 4 + (1e8,) * 3, nu=0.45,
                 R=0.45, r_out=0.12, r_in=0.06, n_theta=70, n_phi=16, n_rho=2,
                 n_rim=300, rank=8, barrier_frac=0.05, barrier=(1e3, 1e6)):
    """The vehicle-scale system: 12 ABD bodies, 7 hollow tori (R 0.45 m,
    r 0.12 m, 11,200 Tet4 each = 78,400 tets), E = 1e10 Pa on 4 and 1e8 Pa
    on 3; torus t is coupled to hub body t by a rank-8 block over its 300
    innermost outer-surface nodes; barrier kappa*I3 on 5% of outer-surface
    nodes."""
    rng = np.random.default_rng(seed)
    if len(youngs) != n_tori:
        raise ValueError("one Young's modulus per torus")
    Is, Js, Bs, rims, ks, nt = [], [], [], [], [], 0
    base = 0
    for t in range(n_tori):
        center = (1.2 * t, 0.0, 0.0)
        nodes, tets, I, J, B, rim, k_s = _torus_block(
            rng, R, r_out, r_in, n_theta, n_phi, n_rho, youngs[t], nu, n_rim,
            barrier_frac, barrier, center)
        Is.append(I + base)
        Js.append(J + base)
        Bs.append(B)
        rims.append(rim + base)
        ks.append(k_s)
        nt += len(tets)
        base += len(nodes)
    bodies = [_spd_body(rng) for _ in range(n_bodies)]
    couplings = []
    for t in range(n_tori):
        hub = t % n_bodies
        Eb = _coupling_block(rng, rank, 3 * len(rims[t]), ks[t], bodies[hub])
        couplings.append((hub, rims[t], Eb))
    info = dict(kind="tori", tets=int(nt), tori=n_tori, bodies=n_bodies, seed=seed,
                youngs=list(map(float, youngs)), nu=nu, rank=rank)
    return _finish(base, np.concatenate(Is), np.concatenate(Js), np.concatenate(Bs),
                   bodies, couplings, rng, info)


def replicate_fem(sysm, copies, seed=0):
    """`copies` block-diagonal copies of a generated system's FEM part, its
    ABD part kept once and coupled to the first copy only: a larger system
    with the same per-row structure (bandwidth measurements with the working
    set beyond the 256 MB Infinity Cache).  Needs the generator's layout
    (FEM blocks first, then the ABD block rows)."""
    if sysm.fo != 0 or sysm.qo != sysm.fn:
        raise ValueError("expects FEM blocks first, then ABD")
    nb = sysm.fn // 3
    r, c, v = sysm.rows.astype(np.int64), sysm.cols.astype(np.int64), sysm.vals
    ff = c < nb                      # FEM-FEM (rows <= cols)
    shift = (copies - 1) * nb        # ABD block rows move past the copies
    rows = [r[ff] + k * nb for k in range(copies)] + [r[~ff] + np.where(r[~ff] >= nb, shift, 0)]
    cols = [c[ff] + k * nb for k in range(copies)] + [c[~ff] + shift]
    vals = [v[ff]] * copies + [v[~ff]]
    R = np.concatenate(rows).astype(np.int32)
    C = np.concatenate(cols).astype(np.int32)
    V = np.ascontiguousarray(np.concatenate(vals))
    fn = copies * sysm.fn
    b = np.random.default_rng(seed).normal(size=fn + sysm.qm)
    info = dict(sysm.info, copies=copies, nb=copies * nb, n=fn, N=fn + sysm.qm, bcoo_nnz=int(len(R)))
    return HSCSystem(R, C, V, b, fo=0, fn=fn, qo=fn, qm=sysm.qm, info=info)


def batch_specs(seed=2, count=64, bodies=(1, 16), tori=(1, 8), tets=(2000, 15000), ranks=(4, 16),
                youngs=(1e8, 1e10)):
    """Per-system parameters of the uneven batch, drawn in order from one
    stream: U{1..16} bodies, U{1..8} tori of log-U[2k, 15k] tets (E
    log-uniform), coupling rank U{4..16}, and the seed of the system's own
    stream (mesh barriers, bodies, couplings, right-hand side)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        nbod = int(rng.integers(bodies[0], bodies[1] + 1))
        ntor = int(rng.integers(tori[0], tori[1] + 1))
        rank = int(rng.integers(ranks[0], ranks[1] + 1))
        target = _log_uniform(rng, tets[0], tets[1], ntor)
        ys = _log_uniform(rng, youngs[0], youngs[1], ntor)
        sub_seed = int(rng.integers(1 << 31))
        out.append(dict(index=s, bodies=nbod, tori=ntor, rank=rank, tets=target, youngs=ys, seed=sub_seed))
    return out


def batch_system(spec):
    """One system of the uneven batch from its spec (see batch_specs); the
    coupling rank is capped at 12 per hub block, the most a 12-row block can
    carry."""
    sub = np.random.default_rng(spec["seed"])
    Is, Js, Bs, rims, ks, base, nt = [], [], [], [], [], 0, 0
    for t in range(spec["tori"]):
        n_phi, n_rho = 12, 2
        n_theta = max(4, 2 * int(round(spec["tets"][t] / (5 * n_phi * n_rho) / 2)))
        n_outer = n_theta * n_phi
        nodes, tt, I, J, B, rim, k_s = _torus_block(
            sub, 0.45, 0.12, 0.06, n_theta, n_phi, n_rho, spec["youngs"][t], 0.45,
            min(300, n_outer // 4), 0.05, (1e3, 1e6), (1.2 * t, 0.0, 0.0))
        Is.append(I + base)
        Js.append(J + base)
        Bs.append(B)
        rims.append(rim + base)
        ks.append(k_s)
        nt += len(tt)
        base += len(nodes)
    bods = [_spd_body(sub) for _ in range(spec["bodies"])]
    couplings = []
    for t in range(spec["tori"]):
        hub = t % spec["bodies"]
        Eb = _coupling_block(sub, min(spec["rank"], 12), 3 * len(rims[t]), ks[t], bods[hub])
        couplings.append((hub, rims[t], Eb))
    info = dict(kind="batch", index=spec["index"], tets=int(nt), tori=spec["tori"], bodies=spec["bodies"],
                rank=spec["rank"])
    return _finish(base, np.concatenate(Is), np.concatenate(Js), np.concatenate(Bs), bods, couplings, sub, info)


def map_processes(fn, items, workers):
    """fn over items in `workers` fresh processes (spawned, so no parent
    state such as an initialised GPU runtime is inherited), results in
    order; workers <= 1 maps in this process."""
    items = list(items)
    if workers <= 1 or len(items) <= 1:
        return [fn(x) for x in items]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(workers, len(items)), mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(fn, items))


def batch_systems(seed=2, count=64, bodies=(1, 16), tori=(1, 8), tets=(2000, 15000),
                  ranks=(4, 16), youngs=(1e8, 1e10), progress=None, workers=1):
    """The uneven batch: `count` systems from batch_specs, built in `workers`
    processes (the same systems for any worker count)."""
    specs = batch_specs(seed, count, bodies, tori, tets, ranks, youngs)
    out = map_processes(batch_system, specs, workers)
    if progress:
        for s, sysm in enumerate(out):
            progress(s, sysm)
    return out


def _affine_jacobian(xbar):
    """3 x 12 dx/dq of x = p + A xbar, q = (p, rows of A)."""
    J = np.zeros((3, 12))
    J[:, :3] = np.eye(3)
    for i in range(3):
        J[i, 3 + 3 * i:6 + 3 * i] = xbar
    return J


def contact_scene(seed=4, nx=6, ny=6, nz=3, h=0.02, E=1e8, nu=0.45, n_bodies=3, verts_per_body=40,
                  n_contacts=600, fixed_body=None, n_fixed_vertices=0, fem_stiffness=True):
    """ABD bodies touching an FEM box through point contacts, split as the
    coupling assembly sees it: the contact list (ABD vertex, FEM vertex,
    3x3 contact block H = -K_c, the ABD-row / FEM-column block of the
    penalty 0.5 (x_a - x_f)^T K_c (x_a - x_f)) and the general triplets of
    everything else (ABD blocks m_v sum J^T J + kappa diag(0, I9) + sum
    J^T K_c J, FEM stiffness + m I + sum K_c), ABD block rows first (qo = 0,
    fo = qm).  K_c = k_n n n^T + k_t (I - n n^T), n a random unit normal.
    With fem_stiffness=False the FEM triplets hold only the mass and contact
    diagonal terms (the caller adds the elastic Hessian, e.g. from the
    device generator); the mesh is returned as "nodes" and "tets"."""
    rng = np.random.default_rng(seed)
    nodes, tets = box_mesh(nx, ny, nz, h)
    nf = len(nodes)
    I, J, B = fem_blocks(nodes, tets, E, nu)
    if not fem_stiffness:
        I = J = np.arange(nf, dtype=np.int64)
        B = np.zeros((nf, 3, 3))
    kf = E * h
    diag = np.flatnonzero(I == J)
    pos = np.empty(nf, dtype=np.int64)
    pos[I[diag]] = diag
    B[diag] += 0.05 * kf * np.eye(3)[None]
    v2b = np.repeat(np.arange(n_bodies), verts_per_body).astype(np.int32)
    xbar = rng.uniform(-0.05, 0.05, size=(len(v2b), 3))
    top = np.flatnonzero(np.isclose(nodes[:, 2], nodes[:, 2].max()))
    av = rng.integers(0, len(v2b), n_contacts).astype(np.int32)
    fv = rng.choice(top, n_contacts).astype(np.int32)
    nrm = rng.normal(size=(n_contacts, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    kn = kf * rng.uniform(0.5, 2.0, n_contacts)
    K = (kn[:, None, None] * (0.9 * nrm[:, :, None] * nrm[:, None, :] + 0.1 * np.eye(3)[None]))
    body_fixed = np.zeros(n_bodies, dtype=np.int32)
    if fixed_body is not None:
        body_fixed[fixed_body] = 1
    vertex_fixed = np.zeros(nf, dtype=np.int32)
    vertex_fixed[rng.choice(top, n_fixed_vertices, replace=False)] = 1
    # ABD blocks
    A = np.zeros((n_bodies, 12, 12))
    mv = 0.05 * kf
    for v in range(len(v2b)):
        Jv = _affine_jacobian(xbar[v])
        A[v2b[v]] += mv * Jv.T @ Jv
    A[:, 3:, 3:] += 0.5 * kf * np.eye(9)[None]
    for c in range(n_contacts):
        Jc = _affine_jacobian(xbar[av[c]])
        A[v2b[av[c]]] += Jc.T @ K[c] @ Jc
        B[pos[fv[c]]] += K[c]
    qm = 12 * n_bodies
    rows, cols, vals = [], [], []
    for t in range(n_bodies):
        for bi in range(4):
            for bj in range(bi, 4):
                rows.append(4 * t + bi)
                cols.append(4 * t + bj)
                vals.append(A[t, 3 * bi:3 * bi + 3, 3 * bj:3 * bj + 3].T.ravel())
    up = I <= J
    R = np.concatenate([np.asarray(rows), I[up] + qm // 3]).astype(np.int32)
    C = np.concatenate([np.asarray(cols), J[up] + qm // 3]).astype(np.int32)
    V = np.concatenate([np.asarray(vals), B[up].transpose(0, 2, 1).reshape(-1, 9)])
    return dict(rows=R, cols=C, vals=np.ascontiguousarray(V), abd_vertex=av, fem_vertex=fv, hess=-K,
                v2b=v2b, xbar=xbar, body_fixed=body_fixed, vertex_fixed=vertex_fixed, qo=0, qm=qm, fo=qm,
                fn=3 * nf, abd_block_offset=0, fem_block_offset=qm // 3, N=qm + 3 * nf,
                b=rng.normal(size=qm + 3 * nf), nodes=nodes, tets=tets)


def from_dense(D, A, E, b):
    """Pack a dense FEM-first system (core.py layout: D n x n, A m x m,
    E m x n) into plugin input format; zero blocks are dropped."""
    n, m = len(D), len(A)
    H = np.zeros((n + m, n + m))
    H[:n, :n], H[n:, n:], H[n:, :n], H[:n, n:] = D, A, E, E.T
    nbk = (n + m) // 3
    Hb = H.reshape(nbk, 3, nbk, 3).transpose(0, 2, 1, 3)
    I, J = np.nonzero(np.abs(Hb).sum(axis=(2, 3)) > 0)
    keep = I <= J
    I, J = I[keep], J[keep]
    v = Hb[I, J].transpose(0, 2, 1).reshape(-1, 9)
    return HSCSystem(I.astype(np.int32), J.astype(np.int32), np.ascontiguousarray(v),
                     np.asarray(b, dtype=np.float64).copy(), fo=0, fn=n, qo=n, qm=m,
                     info=dict(kind="dense", nb=n // 3, n=n, m=m, N=n + m))
