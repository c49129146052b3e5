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
