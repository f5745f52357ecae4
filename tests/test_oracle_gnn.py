"""Pins the CPU restatement of the learned preconditioner and of its network
(oracle/gnn_oracle.py): the Kaporin row objective against the dense matrix
G S G', the SPD and block-Jacobi properties of M^-1 = G'G, the network's
gradient against finite differences, and the BF16 rounding."""
import numpy as np
import pytest

from oracle import gnn_oracle as go


def random_graph(rng, n=14, p=0.35):
    """A random symmetric covisibility pattern (CSR by row, diagonal slots
    included like the reduced camera system) with random features."""
    A = rng.random((n, n)) < p
    A = A | A.T
    np.fill_diagonal(A, True)
    row_ptr, col = [0], []
    for i in range(n):
        js = np.where(A[i])[0]
        col += list(js)
        row_ptr.append(len(col))
    col = np.array(col)
    pos = {(i, j): k for i in range(n) for k, j in zip(range(row_ptr[i], row_ptr[i + 1]), col[row_ptr[i]:row_ptr[i + 1]])}
    rev = np.array([pos[(j, i)] for i in range(n) for j in col[row_ptr[i]:row_ptr[i + 1]]])
    x = rng.normal(size=(n, go.NODE_F))
    ef = rng.normal(size=(len(col), go.EDGE_F))
    qc = np.zeros((n, 5))
    for i in range(n):
        B = rng.normal(size=(6, 6)) * 0.3
        a = 2.0 + rng.random()
        qc[i] = [a, -0.5 * a * rng.random(), 0.3 + rng.random(), 0.2 * rng.random(), float(i > 0)]
    return go.Graph(row_ptr, col, rev, x, ef), qc


def block_spd(rng, n=7, p=0.5):
    """Block-sparse SPD matrix (6x6 blocks) with a random symmetric pattern."""
    A = rng.random((n, n)) < p
    A = A | A.T
    np.fill_diagonal(A, True)
    J = rng.normal(size=(6 * n + 20, 6 * n))
    S = J.T @ J
    mask = np.kron(A, np.ones((6, 6)))
    S = S * mask + 6 * n * np.eye(6 * n)  # diagonally dominant enough to stay SPD
    return S


def test_kaporin_rows_equal_the_diagonal_blocks_of_GSG():
    rng = np.random.default_rng(1)
    S = block_spd(rng)
    qc = go.traces(S)
    for _ in range(3):
        theta = rng.normal(size=(S.shape[0] // 6, 2)) * [0.7, 2.0]
        G = go.kmatrix(S, theta)
        GSG = G @ S @ G.T
        direct = np.array([np.trace(GSG[6 * i:6 * i + 6, 6 * i:6 * i + 6]) for i in range(S.shape[0] // 6)])
        np.testing.assert_allclose(go.kaporin_rows(qc, theta), direct, rtol=1e-10)


def test_zero_theta_is_block_jacobi_and_every_theta_gives_spd():
    rng = np.random.default_rng(2)
    S = block_spd(rng)
    n = S.shape[0] // 6
    G = go.kmatrix(S, np.zeros((n, 2)))
    Minv = G.T @ G
    for i in range(n):
        blk = S[6 * i:6 * i + 6, 6 * i:6 * i + 6]
        np.testing.assert_allclose(Minv[6 * i:6 * i + 6, 6 * i:6 * i + 6], np.linalg.inv(blk), rtol=1e-9, atol=1e-12)
        for j in range(n):
            if j != i:
                assert np.abs(Minv[6 * i:6 * i + 6, 6 * j:6 * j + 6]).max() == 0.0
    for _ in range(3):
        G = go.kmatrix(S, rng.normal(size=(n, 2)) * 3)
        ev = np.linalg.eigvalsh(G.T @ G)
        assert ev.min() > 0


def test_trace_identities():
    # <T1, T3> = <T2, T2> (Sh_PP - I is symmetric), which the row objective uses.
    rng = np.random.default_rng(3)
    S = block_spd(rng, n=8)
    n, Sb, nz, Linv, Sh = go.normalized(S)
    for i in range(1, n):
        P, T1, T2, T3 = go.row_terms(Sh, nz, i)
        if T1 is not None:
            assert abs(np.sum(T1 * T3) - np.sum(T2 * T2)) <= 1e-10 * max(1.0, np.sum(T2 * T2))


def test_hat_basis_interpolates_linear_functions():
    # Partition of unity and exact interpolation of the node index: the
    # weights of vertex i sum to 1 and place it at s = i (K - 1) / (n - 1).
    n, K = 37, 5
    W = go.hat_basis(n, K)
    for c in range(6):
        Wc = W[c::6, c::6]
        np.testing.assert_allclose(Wc.sum(axis=1), 1.0, rtol=0, atol=1e-15)
        np.testing.assert_allclose(Wc @ np.arange(K), np.arange(n) * (K - 1) / (n - 1), rtol=0, atol=1e-13)
        assert np.abs(W[c::6, [k for k in range(6 * K) if k % 6 != c]]).max() == 0.0
    assert go.default_hat_nodes(49) == 3 and go.default_hat_nodes(31) == 0 and go.default_hat_nodes(2000) == 15


def test_deflated_pcg_reduces_to_pcg_and_solves():
    rng = np.random.default_rng(5)
    S = block_spd(rng, n=40)
    b = rng.normal(size=S.shape[0])
    n = S.shape[0] // 6
    G = go.kmatrix(S, rng.normal(size=(n, 2)) * 0.5)
    Minv = lambda r: G.T @ (G @ r)  # noqa: E731
    it0, _ = go.deflated_pcg(S, b, Minv, np.zeros((S.shape[0], 0)), tol=1e-8)
    assert it0 == go.pcg_iterations(S, b, Minv, tol=1e-8)
    it, x = go.deflated_pcg(S, b, Minv, go.hat_basis(n, 4), tol=1e-10)
    assert it > 0
    np.testing.assert_allclose(x, np.linalg.solve(S, b), rtol=1e-7, atol=1e-9 * np.abs(np.linalg.solve(S, b)).max())


def test_deflation_of_all_but_m_eigenvectors_takes_m_iterations():
    # W spanning every eigenvector of S but m leaves an operator with m
    # distinct eigenvalues on the complement: CG ends in m steps. A dropped
    # projection, a wrong sign or a transposed SW breaks this.
    rng = np.random.default_rng(6)
    Q, _ = np.linalg.qr(rng.normal(size=(30, 30)))
    lam = np.linspace(1.0, 50.0, 30)
    S = (Q * lam) @ Q.T
    b = rng.normal(size=30)
    for m in (1, 3):
        W = Q[:, m:] @ rng.normal(size=(30 - m, 30 - m))  # any basis of that span
        it, x = go.deflated_pcg(S, b, lambda r: r.copy(), W, tol=1e-10)
        assert 0 < it <= m
        np.testing.assert_allclose(x, np.linalg.solve(S, b), rtol=1e-8)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    g, qc = random_graph(rng)
    params = go.init_params(rng).astype(np.float64)
    params += rng.normal(size=params.shape) * 0.05   # every layer's parameters matter
    loss, grad, _ = go.loss_and_grad(params, g, qc, bf16=False)
    idx = np.concatenate([rng.choice(np.arange(o, o + n), size=min(n, 6), replace=False)
                          for o, n in go.LAYOUT.values()])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(params)
        e[k] = h
        lp, _, _ = go.loss_and_grad(params + e, g, qc, bf16=False)
        lm, _, _ = go.loss_and_grad(params - e, g, qc, bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (k, fd, grad[k])
    # A random direction through all parameters.
    d = rng.normal(size=params.shape)
    lp, _, _ = go.loss_and_grad(params + h * d, g, qc, bf16=False)
    lm, _, _ = go.loss_and_grad(params - h * d, g, qc, bf16=False)
    assert abs((lp - lm) / (2 * h) - grad @ d) <= 1e-6 * max(1.0, abs(grad @ d))


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38, 1e-40])
    r = go.round_bf16(x)
    np.testing.assert_array_equal(r[:4], [1.0, 1.0, 1.0 + 2.0 ** -7, -2.5])
    assert np.isfinite(r[4]) and r[5] >= 0
    # Every BF16 value is a float32 with the low 16 bits zero.
    v = go.round_bf16(np.random.default_rng(5).normal(size=1000)).astype(np.float32).view(np.uint32)
    assert np.all((v & 0xFFFF) == 0)
