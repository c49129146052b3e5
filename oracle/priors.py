"""FP64 NumPy restatement of the reference priors (core/seis_prior.py).

TEST INFRASTRUCTURE ONLY (parity oracle); the product never imports it.

Generalised from the hard-coded 70x70 grid to (nz, nx).  Every prior exposes
``basis`` (n_params x N), ``lam`` and ``cost_grad(x) -> (cost, grad)`` where
cost already includes lam (Prior.compute_cost_and_gradient, :47-65).
"""
from __future__ import annotations

import numpy as np


class RowTotalVariation:
    """core/seis_prior.py:67-109: each row constant; smoothed |row diffs|."""

    def __init__(self, nz, nx, lam=1e-8, epsilon=0.1):
        self.nz, self.nx, self.lam, self.epsilon = nz, nx, lam, epsilon
        self.N = nz + 1
        B = np.zeros((nz * nx + 1, nz + 1))
        for i in range(nz):
            m = np.zeros((nz, nx))
            m[i, :] = 1.0
            B[:-1, i] = m.ravel()
        B[-1, nz] = 1.0
        self.basis = B

    def cost_grad(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        nz = self.nz
        diff = np.diff(x[:nz])
        cpi = np.sqrt(diff ** 2 + self.epsilon ** 2)
        cost = np.mean(cpi)
        sign = diff / cpi
        g = np.zeros_like(x)
        g[0] = -sign[0]
        g[1:nz - 1] = sign[:-1] - sign[1:]
        g[nz - 1] = sign[-1]
        g = g / (nz - 1)
        return self.lam * cost, self.lam * g


class TotalVariation:
    """core/seis_prior.py:112-177: smoothed 2-D TV on the identity basis."""

    def __init__(self, nz, nx, lam=1e-8, epsilon=0.1):
        self.nz, self.nx, self.lam, self.epsilon = nz, nx, lam, epsilon
        self.N = nz * nx + 1
        self.basis = np.eye(self.N)

    def cost_grad(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        xm = x[:-1].reshape(self.nz, self.nx)
        d1 = np.diff(xm, axis=0)
        d2 = np.diff(xm, axis=1)
        diff = np.concatenate([d1.ravel(), d2.ravel()])
        cost = np.mean(np.sqrt(diff ** 2 + self.epsilon ** 2))
        N = diff.size
        g1 = d1 / np.sqrt(d1 ** 2 + self.epsilon ** 2) / N
        g2 = d2 / np.sqrt(d2 ** 2 + self.epsilon ** 2) / N
        gm = np.zeros_like(xm)
        gm[1:, :] += g1
        gm[:-1, :] -= g1
        gm[:, 1:] += g2
        gm[:, :-1] -= g2
        g = np.concatenate([gm.ravel(), [0.0]])
        return self.lam * cost, self.lam * g


class L2Smoothness:
    """lam * x^T P x with P = D^T D, D the first differences along z and x.

    Not in the reference: the Gauss-Newton workload asks for an L2 smoothness
    prior, and the reference's GN (core/seis_invert.py:63-107) needs an
    explicit precision matrix ``P`` (cost lam x^T P x, gradient 2 lam P x, as
    SquaredExponential at core/seis_prior.py:232-242).
    """

    def __init__(self, nz, nx, lam=1e-8):
        self.nz, self.nx, self.lam = nz, nx, lam
        self.N = nz * nx + 1
        self.basis = np.eye(self.N)
        n = nz * nx
        rows = []
        idx = np.arange(n).reshape(nz, nx)
        D = []
        for a, b in [(idx[1:, :], idx[:-1, :]), (idx[:, 1:], idx[:, :-1])]:
            for i, j in zip(a.ravel(), b.ravel()):
                r = np.zeros(n)
                r[i], r[j] = 1.0, -1.0
                D.append(r)
        self.D = np.array(D)
        self.P = self.D.T @ self.D

    def cost_grad(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        xv = x[:-1]
        Px = self.P @ xv
        cost = float(xv @ Px)
        g = np.concatenate([2 * Px, [0.0]])
        return self.lam * cost, self.lam * g


class SquaredExponential:
    """core/seis_prior.py:178-242 on an (nz, nx) grid centred as :199-207.
    K from the SE kernel, a constant offset, a depth slope and noise; P =
    K^-1 (:210).  transform (:213-219): basis U[:, s > cutoff] sqrt(s) from
    the SVD of K, min_vel appended, and P = I."""

    def __init__(self, nz, nx, lam=1e-8, length_scale=np.log(32.4), noise=0.1, sigma=183.4, sigma_mean=520.0,
                 sigma_slope=31.4, transform=False, svd_cutoff=0.0):
        self.nz, self.nx, self.lam = nz, nx, lam
        y = (np.arange(nz) - nz // 2)[:, None] + np.zeros((1, nx))
        x = (np.arange(nx) - nx // 2)[None, :] + np.zeros((nz, 1))
        x = x.reshape(-1, 1)
        y = y.reshape(-1, 1)
        dist = np.sqrt((x - x.T) ** 2 + (y - y.T) ** 2)
        K = sigma ** 2 * np.exp(-dist ** 2 / (2 * length_scale ** 2)) + sigma_mean ** 2
        K = K + sigma_slope ** 2 * (y @ y.T) + noise ** 2 * np.eye(nz * nx)
        self.K = K
        self.P = np.linalg.inv(K)
        if transform:
            U, sv, _ = np.linalg.svd(K)
            keep = sv > svd_cutoff
            B = np.zeros((nz * nx + 1, int(keep.sum()) + 1))
            B[:-1, :-1] = U[:, keep] @ np.diag(np.sqrt(sv[keep]))
            B[-1, -1] = 1.0
            self.basis = B
            self.P = np.eye(B.shape[1] - 1)
        else:
            self.basis = np.eye(nz * nx + 1)
        self.N = self.basis.shape[1]

    def cost_grad(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        Px = self.P @ x[:-1]
        return self.lam * float(x[:-1] @ Px), self.lam * np.concatenate([2 * Px, [0.0]])


def label_thresholded_components(A, X, connectivity=4):
    """core/seis_numerics.py:81-124: connected components of the grid graph
    whose edges join neighbouring pixels (4- or 8-connectivity) with
    |A[n] - A[p]| <= X.  Labels 0..n_labels-1 are numbered in order of each
    component's smallest flat pixel index, the order scipy's
    connected_components (the reference's call, :117-120) assigns for an
    undirected graph.  Restated with union-find (smallest index as root);
    pinned against scipy itself in tests/test_oracle.py."""
    A = np.asarray(A, dtype=np.float64)
    H, W = A.shape
    N = H * W
    parent = np.arange(N)

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    neigh = [(1, 0), (0, 1)] if connectivity == 4 else [(1, 0), (0, 1), (1, 1), (1, -1)]
    for i in range(H):
        for j in range(W):
            for di, dj in neigh:   # each undirected edge once; |a-b| is symmetric (:105-110)
                ni, nj = i + di, j + dj
                if 0 <= ni < H and 0 <= nj < W and abs(A[ni, nj] - A[i, j]) <= X:
                    a, b = find(i * W + j), find(ni * W + nj)
                    if a != b:
                        parent[max(a, b)] = min(a, b)
    roots = np.array([find(i) for i in range(N)])
    order = {r: k for k, r in enumerate(sorted(set(roots.tolist())))}
    labels = np.array([order[r] for r in roots.tolist()], dtype=np.int64)
    return labels.reshape(H, W), len(order)


class RestrictFlatAreas:
    """core/seis_prior.py:244-281: areas that are flat (4-connected, steps of
    at most diff_threshold) in a velocity guess become single parameters.

    The basis is built as in _adapt (:258-272): one indicator column per
    label, empty columns dropped, min_vel appended (block_diag), every column
    scaled to sum 1.  cost_grad evaluates the underlying prior (identity
    basis, :252-254) at basis @ x and maps its gradient back (:274-281)."""

    def __init__(self, underlying, diff_threshold=1.0):
        self.underlying = underlying
        self.diff_threshold = diff_threshold
        self.lam = underlying.lam
        self.nz, self.nx = underlying.nz, underlying.nx
        self.basis = np.eye(self.nz * self.nx + 1)
        self.N = self.basis.shape[1]

    def adapt(self, vec_guess):
        v = np.asarray(vec_guess, dtype=np.float64).reshape(-1)[:self.nz * self.nx].reshape(self.nz, self.nx)
        labels, count = label_thresholded_components(v, self.diff_threshold, connectivity=4)
        n = self.nz * self.nx
        bf = np.zeros((n, count + 1))
        bf[np.arange(n), labels.ravel()] = 1.0
        bf = bf[:, bf.sum(axis=0) > 0]
        B = np.zeros((n + 1, bf.shape[1] + 1))
        B[:n, :bf.shape[1]] = bf
        B[n, -1] = 1.0
        self.basis = B / B.sum(axis=0)
        self.N = self.basis.shape[1]
        if hasattr(self.underlying, "P"):   # precision on the area coordinates (Gauss-Newton)
            Bv = self.basis[:n, :-1]
            self.P = Bv.T @ self.underlying.P @ Bv
        return labels, count

    def cost_grad(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        c, g = self.underlying.cost_grad(self.basis @ x)
        return c, self.basis.T @ g
