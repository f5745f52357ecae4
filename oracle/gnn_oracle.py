"""CPU restatement (numpy, FP64) of the learned block-Jacobi correction of the
reduced camera system and of its covisibility-graph network, for tests only
(test infrastructure: the product path never imports it).

There is no counterpart in the reference (maplab's Ceres ITERATIVE_SCHUR
uses the fixed SCHUR_JACOBI preconditioner, algorithms/map-optimization/
include/map-optimization/solver-options.h:57-68), so this module states the
mathematics directly:

* preconditioner: S_ii = L_i L_i', Sh_ij = L_i^-1 S_ij L_j^-T, lower
  neighbours P(i) = {j < i : S_ij != 0}; T1 = Sh_iP, T2 = T1 (Sh_PP - I),
  T3 = T2 (Sh_PP - I); K_i = theta_i1 T1 + theta_i2 T2; G = (I - K) L^-1;
  M^-1 = G'G. Per row the Kaporin objective t_i = tr[(G S G')_ii] is
  6 - 2 theta'c + theta'Q theta with c = (a, b),
  Q = [[a + b, b + e], [b + e, e + d]], a = <T1,T1>, b = <T1,T2>,
  e = <T2,T2>, d = <T2,T3>.
* network: the forward and backward passes of the 3-layer message-passing
  network (hidden width 64), with the BF16 rounding of the activations at
  the points the device rounds them (bf16=True) or none (bf16=False, the
  exact function whose gradient finite differences check).

* deflated PCG (Saad, Yeung, Erhel, Guyomarc'h, SIAM J. Sci. Comput. 21,
  2000) over the hat-function coarse space of the vertex order, with any SPD
  preconditioner: x0 = W E^-1 W'b, E = W'SW, and each search direction
  p = z - W E^-1 (SW)'z + beta p made S-orthogonal to W.

Pinned by (tests/test_oracle_gnn.py): finite differences of the loss for the
gradient; t_i(theta) against tr[(G S G')_ii] formed densely; theta = 0
giving block-Jacobi; M^-1 symmetric positive definite for random theta; the
deflated PCG reducing to plain PCG without coarse columns, solving S x = b,
and converging in m iterations when W spans all but m eigenvectors.
"""
from __future__ import annotations

import numpy as np

H = 64
NODE_F, EDGE_F, TERMS, LAYERS = 9, 3, 2, 3


# ------------------------------------------------------------ preconditioner

def blocks(S):
    n = S.shape[0] // 6
    Sb = S.reshape(n, 6, n, 6).transpose(0, 2, 1, 3)
    nz = np.abs(Sb).max(axis=(2, 3)) > 0
    return n, Sb, nz


def normalized(S):
    """(n, Sb, nz, Linv, Sh): the blocks and the normalized blocks."""
    n, Sb, nz = blocks(S)
    Linv = np.stack([np.linalg.inv(np.linalg.cholesky(0.5 * (Sb[i, i] + Sb[i, i].T))) for i in range(n)])
    Sh = np.einsum('iab,ijbc,jdc->ijad', Linv, Sb, Linv)
    return n, Sb, nz, Linv, Sh


def row_terms(Sh, nz, i):
    """(P, T1, T2, T3) of row i (T as 6 x 6|P| matrices)."""
    P = [j for j in range(i) if nz[i, j]]
    if not P:
        return P, None, None, None
    SPP = Sh[np.ix_(P, P)].transpose(0, 2, 1, 3).reshape(6 * len(P), 6 * len(P))
    E = SPP - np.eye(6 * len(P))
    T1 = Sh[i, P].transpose(1, 0, 2).reshape(6, -1)
    T2 = T1 @ E
    T3 = T2 @ E
    return P, T1, T2, T3


def traces(S):
    """Per row (a, b, e, d, has) as float64 (n x 5)."""
    n, Sb, nz, Linv, Sh = normalized(S)
    out = np.zeros((n, 5))
    for i in range(n):
        P, T1, T2, T3 = row_terms(Sh, nz, i)
        if T1 is None:
            continue
        out[i] = [np.sum(T1 * T1), np.sum(T1 * T2), np.sum(T2 * T2), np.sum(T2 * T3), 1.0]
    return out


def kmatrix(S, theta):
    """Dense G = (I - K) L^-1 of the coefficients theta (n x 2)."""
    n, Sb, nz, Linv, Sh = normalized(S)
    K = np.zeros_like(S)
    Li = np.zeros_like(S)
    for i in range(n):
        Li[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Linv[i]
        P, T1, T2, _ = row_terms(Sh, nz, i)
        if T1 is None:
            continue
        Ki = theta[i, 0] * T1 + theta[i, 1] * T2
        for k, j in enumerate(P):
            K[6 * i:6 * i + 6, 6 * j:6 * j + 6] = Ki[:, 6 * k:6 * k + 6]
    return (np.eye(S.shape[0]) - K) @ Li


def kaporin_rows(qc, theta):
    a, b, e, d = qc[:, 0], qc[:, 1], qc[:, 2], qc[:, 3]
    t1, t2 = theta[:, 0], theta[:, 1]
    return 6 - 2 * (t1 * a + t2 * b) + t1 * ((a + b) * t1 + (b + e) * t2) + t2 * ((b + e) * t1 + (e + d) * t2)


def pcg_iterations(S, b, Minv, tol=1e-6, maxit=5000):
    x = np.zeros_like(b)
    r = b.copy()
    z = Minv(r)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for k in range(maxit):
        q = S @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= tol * nb:
            return k + 1
        z = Minv(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return -1


def hat_basis(n, K):
    """Hat functions over the order of n vertices with K nodes (6 K columns,
    column 6 k + c for node k and component c): vertex i at position
    s = i (K - 1) / (n - 1) has weights (1 - t, t) on nodes (n0, n0 + 1),
    n0 = min(floor(s), K - 2), t = s - n0."""
    W = np.zeros((6 * n, 6 * K))
    if K <= 0:
        return W
    for i in range(n):
        s = i * (K - 1) / (n - 1)
        n0 = min(int(s), K - 2)
        t = s - n0
        for c in range(6):
            W[6 * i + c, 6 * n0 + c] = 1.0 - t
            W[6 * i + c, 6 * (n0 + 1) + c] = t
    return W


def default_hat_nodes(n):
    """Coarse nodes of the default deflation (min(15, n / 16); none below 2
    or when n < 2 K)."""
    K = min(15, n // 16)
    return 0 if K < 2 or n < 2 * K else K


def deflated_pcg(S, b, Minv, W, tol=1e-6, maxit=5000):
    """Deflated PCG iterations to |r| <= tol |b| (-1: not within maxit) and
    the solution."""
    nb = np.linalg.norm(b)
    if W.shape[1] == 0:
        SW = np.zeros((S.shape[0], 0))
        Einv = np.zeros((0, 0))
    else:
        SW = S @ W
        Einv = np.linalg.inv(W.T @ SW)
    x = W @ (Einv @ (W.T @ b))
    r = b - SW @ (Einv @ (W.T @ b))
    if np.linalg.norm(r) <= tol * nb:
        return 0, x
    z = Minv(r)
    p = z - W @ (Einv @ (SW.T @ z))
    rz = r @ z
    for k in range(maxit):
        q = S @ p
        a = rz / (p @ q)
        x = x + a * p
        r = r - a * q
        if np.linalg.norm(r) <= tol * nb:
            return k + 1, x
        z = Minv(r)
        rz2 = r @ z
        p = z - W @ (Einv @ (SW.T @ z)) + (rz2 / rz) * p
        rz = rz2
    return -1, x


# ------------------------------------------------------------ network

def layout():
    """Offsets of the FP32 parameter vector (same order as the device)."""
    o, L = 0, {}

    def take(name, n):
        nonlocal o
        L[name] = (o, n)
        o += n
    take('w_in', NODE_F * H)
    take('b_in', H)
    for l in range(LAYERS):
        take(f'w_m{l}', H * H)
        take(f'b_m{l}', H)
        take(f'a_g{l}', EDGE_F)
        take(f'c_g{l}', 1)
        take(f'w_u{l}', 2 * H * H)
        take(f'b_u{l}', H)
    take('w_out', H * TERMS)
    take('b_out', TERMS)
    return L, o


LAYOUT, NUM_PARAMS = layout()
SHAPES = {'w_in': (NODE_F, H), 'w_out': (H, TERMS)}
for _l in range(LAYERS):
    SHAPES[f'w_m{_l}'] = (H, H)
    SHAPES[f'w_u{_l}'] = (2 * H, H)


def unpack(params):
    out = {}
    for k, (o, n) in LAYOUT.items():
        v = np.asarray(params[o:o + n], dtype=np.float64)
        out[k] = v.reshape(SHAPES[k]) if k in SHAPES else v
    return out


def pack(grads):
    out = np.zeros(NUM_PARAMS)
    for k, (o, n) in LAYOUT.items():
        out[o:o + n] = np.asarray(grads[k]).reshape(-1)
    return out


def round_bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(rng, theta0=(0.5, -4.0)):
    """He-initialized weights; the output bias starts at theta0."""
    p = np.zeros(NUM_PARAMS)
    P = unpack(p)
    for k, shape in SHAPES.items():
        P[k] = rng.normal(size=shape) * np.sqrt(2.0 / shape[0])
    P['w_out'] *= 0.01
    P['b_out'] = np.asarray(theta0, dtype=np.float64)
    return pack(P).astype(np.float32)


class Graph:
    """CSR by destination row: slots [row_ptr[i], row_ptr[i+1]) of row i with
    source col[e]; slots with col == row are not edges; rev[e] is the slot
    of the reverse edge."""

    def __init__(self, row_ptr, col, rev, node_feat, edge_feat):
        self.row_ptr, self.col, self.rev = np.asarray(row_ptr), np.asarray(col), np.asarray(rev)
        self.x = np.asarray(node_feat, dtype=np.float64)
        self.ef = np.asarray(edge_feat, dtype=np.float64)
        self.n = len(self.row_ptr) - 1
        self.dst = np.repeat(np.arange(self.n), np.diff(self.row_ptr))
        self.edge = self.col != self.dst
        self.deg = np.bincount(self.dst[self.edge], minlength=self.n)


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def forward(params, g: Graph, bf16=True):
    """theta (n x 2) and the cache of the backward pass."""
    rb = round_bf16 if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    P = unpack(params)
    h = rb(np.maximum(g.x @ P['w_in'] + P['b_in'], 0.0))
    cache = {'h0': h, 'layers': []}
    e = np.where(g.edge)[0]
    for l in range(LAYERS):
        wm, wu = rb(P[f'w_m{l}']), rb(P[f'w_u{l}'])
        y = rb(h @ wm + P[f'b_m{l}'])
        gate = _sigmoid(g.ef @ P[f'a_g{l}'] + P[f'c_g{l}'][0])
        acc = np.zeros((g.n, H))
        np.add.at(acc, g.dst[e], gate[e, None] * y[g.col[e]])
        agg = rb(np.where(g.deg[:, None] > 0, acc / np.maximum(g.deg, 1)[:, None], 0.0))
        ha = np.concatenate([h, agg], axis=1)
        hn = rb(np.maximum(ha @ wu + P[f'b_u{l}'], 0.0))
        cache['layers'].append({'h': h, 'y': y, 'gate': gate, 'ha': ha, 'wm': wm, 'wu': wu})
        h = hn
    theta = h @ P['w_out'] + P['b_out']
    cache['h3'] = h
    return theta, cache


def loss_and_grad(params, g: Graph, qc, bf16=True):
    """Loss sum_i has_i log t_i / sum has and its gradient (as the device)."""
    rb = round_bf16 if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    P = unpack(params)
    theta, c = forward(params, g, bf16)
    qc = np.asarray(qc, dtype=np.float64)
    has = qc[:, 4] != 0
    inv = 1.0 / max(has.sum(), 1)
    t = np.maximum(kaporin_rows(qc, theta), 1e-12)
    loss = inv * np.sum(np.log(t[has]))
    a, b, e_, d = qc[:, 0], qc[:, 1], qc[:, 2], qc[:, 3]
    t1, t2 = theta[:, 0], theta[:, 1]
    dth = np.zeros_like(theta)
    dth[:, 0] = inv * (2 * ((a + b) * t1 + (b + e_) * t2) - 2 * a) / t
    dth[:, 1] = inv * (2 * ((b + e_) * t1 + (e_ + d) * t2) - 2 * b) / t
    dth[~has] = 0.0
    G = {}
    h3 = c['h3']
    G['w_out'] = h3.T @ dth
    G['b_out'] = dth.sum(0)
    dz = rb((dth @ P['w_out'].T) * (h3 > 0))
    e = np.where(g.edge)[0]
    for l in reversed(range(LAYERS)):
        L = c['layers'][l]
        G[f'w_u{l}'] = L['ha'].T @ dz
        G[f'b_u{l}'] = dz.sum(0)
        dha = dz @ L['wu'].T
        dagg = dha[:, H:]
        gate = L['gate']
        # dy_j = sum over slots e of row j (src i) of gate(rev e) / deg_i dagg_i
        dyacc = np.zeros((g.n, H))
        er = g.rev[e]
        np.add.at(dyacc, g.dst[e], (gate[er] / np.maximum(g.deg[g.col[e]], 1))[:, None] * dagg[g.col[e]])
        dy = rb(dyacc)
        dg = np.einsum('ef,ef->e', dagg[g.dst[e]], L['y'][g.col[e]]) / np.maximum(g.deg[g.dst[e]], 1)
        ds = dg * gate[e] * (1 - gate[e])
        G[f'a_g{l}'] = ds @ g.ef[e]
        G[f'c_g{l}'] = np.array([ds.sum()])
        G[f'w_m{l}'] = L['h'].T @ dy
        G[f'b_m{l}'] = dy.sum(0)
        dh = dha[:, :H] + dy @ L['wm'].T
        if l > 0:
            dz = rb(dh * (L['h'] > 0))
        else:
            dz0 = dh * (c['h0'] > 0)
            G['w_in'] = g.x.T @ dz0
            G['b_in'] = dz0.sum(0)
    return loss, pack(G), theta
