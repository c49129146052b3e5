"""Priors for the inversion loops (core/seis_prior.py), on device in float64.

Each prior carries ``basis`` (n_params x N, or None for the identity), the
weight ``lam`` and ``cost_grad(x) -> (cost, grad)`` with lam applied
(Prior.compute_cost_and_gradient, core/seis_prior.py:47-65).  Priors usable by
Gauss-Newton also provide ``apply_P(xv)`` (the precision matrix acting on the
velocity part, core/seis_invert.py:78,82).
"""
from __future__ import annotations

import math

import numpy as np

import torch


class Prior:
    basis: torch.Tensor | None = None
    lam: float = 1.0

    def __init__(self, nz: int, nx: int, device="cuda"):
        self.nz, self.nx = nz, nx
        self.device = torch.device(device)

    @property
    def N(self) -> int:
        return self.nz * self.nx + 1 if self.basis is None else self.basis.shape[1]

    def to_basis(self, vec: torch.Tensor) -> torch.Tensor:
        """x = solve(B^T B, B^T vec) (core/seis_invert.py:72,121)."""
        if self.basis is None:
            return vec.clone()
        B = self.basis
        # small dense solve on the host (LAPACK), like the dense GN solve
        return torch.linalg.solve((B.T @ B).cpu(), (B.T @ vec).cpu()).to(vec.device)

    def from_basis(self, x: torch.Tensor) -> torch.Tensor:
        return x if self.basis is None else self.basis @ x

    def is_identity(self) -> bool:
        """True when to_basis / from_basis leave vectors unchanged."""
        return self.basis is None

    def basis_T(self, g: torch.Tensor) -> torch.Tensor:
        return g if self.basis is None else self.basis.T @ g

    def from_basis_rows(self, X: torch.Tensor) -> torch.Tensor:
        """from_basis of every row of X (S x n_basis)."""
        return X if self.basis is None else X @ self.basis.T

    def cost_rows(self, X: torch.Tensor) -> torch.Tensor:
        """Prior cost (lam included) of every row of X, as a tensor; the
        Gauss-Newton line search evaluates all its trial points with it."""
        return torch.stack([torch.as_tensor(self.cost_grad(x)[0], dtype=X.dtype, device=X.device) for x in X])

    # ---- the reference's Prior interface (core/seis_prior.py:17-65), so code
    # written against it (model.prior.λ = ..., prior.basis_vectors,
    # prior.compute_cost_and_gradient(x, True)) runs unchanged
    prepped: bool = False

    @property
    def λ(self) -> float:
        return self.lam

    @λ.setter
    def λ(self, value: float):
        self.lam = value

    @property
    def basis_vectors(self) -> torch.Tensor:
        """Dense basis, (nz*nx + 1) x N float64 (the identity for identity-basis
        priors, built on first use)."""
        if self.basis is not None:
            return self.basis
        if getattr(self, "_eye", None) is None:
            self._eye = torch.eye(self.nz * self.nx + 1, dtype=torch.float64, device=self.device)
        return self._eye

    def prep(self):
        """Prior.prep (:27-35): the basis is built at construction here."""
        self.prepped = True

    def adapt(self, velocity_guess):
        """Prior.adapt (:37-45); priors without an adaptive basis keep theirs."""

    def compute_cost_and_gradient(self, x, compute_gradient: bool = False):
        """Prior.compute_cost_and_gradient (:47-65): x of shape (N, 1);
        returns the 0-d cost (lam applied), with the (N, 1) gradient when
        compute_gradient."""
        x = torch.as_tensor(x)
        if tuple(x.shape) != (self.N, 1):
            raise ValueError(f"x must have shape ({self.N}, 1), got {tuple(x.shape)}")
        cost, grad = self.cost_grad(x.reshape(-1).to(device=self.device, dtype=torch.float64))
        cost = torch.as_tensor(cost, dtype=torch.float64, device=self.device)
        return (cost, grad.reshape(-1, 1)) if compute_gradient else cost


class RowTotalVariation(Prior):
    """Rows constant, smoothed |row differences| (core/seis_prior.py:67-109)."""

    def __init__(self, nz, nx, lam=1e-8, epsilon=0.1, device="cuda"):
        super().__init__(nz, nx, device)
        self.lam, self.epsilon = lam, epsilon
        B = torch.zeros(nz * nx + 1, nz + 1, dtype=torch.float64)
        for i in range(nz):
            B[i * nx:(i + 1) * nx, i] = 1.0
        B[-1, nz] = 1.0
        self.basis = B.to(self.device)

    def cost_grad(self, x):
        nz = self.nz
        diff = torch.diff(x[:nz])
        cpi = torch.sqrt(diff ** 2 + self.epsilon ** 2)
        cost = cpi.mean()
        sign = diff / cpi
        g = torch.zeros_like(x)
        g[0] = -sign[0]
        g[1:nz - 1] = sign[:-1] - sign[1:]
        g[nz - 1] = sign[-1]
        return self.lam * cost, self.lam * g / (nz - 1)


class TotalVariation(Prior):
    """Smoothed 2-D total variation, identity basis (core/seis_prior.py:112-177)."""

    def __init__(self, nz, nx, lam=1e-8, epsilon=0.1, device="cuda"):
        super().__init__(nz, nx, device)
        self.lam, self.epsilon = lam, epsilon

    def cost_grad(self, x):
        xm = x[:-1].reshape(self.nz, self.nx)
        d1 = torch.diff(xm, dim=0)
        d2 = torch.diff(xm, dim=1)
        e2 = self.epsilon ** 2
        n = d1.numel() + d2.numel()
        cost = (torch.sqrt(d1 ** 2 + e2).sum() + torch.sqrt(d2 ** 2 + e2).sum()) / n
        g1 = d1 / torch.sqrt(d1 ** 2 + e2) / n
        g2 = d2 / torch.sqrt(d2 ** 2 + e2) / n
        gm = torch.zeros_like(xm)
        gm[1:, :] += g1
        gm[:-1, :] -= g1
        gm[:, 1:] += g2
        gm[:, :-1] -= g2
        g = torch.cat([gm.reshape(-1), torch.zeros(1, dtype=x.dtype, device=x.device)])
        return self.lam * cost, self.lam * g


class L2Smoothness(Prior):
    """lam * x^T D^T D x over the velocity part, D = first differences in z and x.

    The Gauss-Newton workload's smoothness prior; like SquaredExponential
    (core/seis_prior.py:232-242) its cost is lam x^T P x and its gradient
    2 lam P x, and P = D^T D is applied matrix-free.
    """

    def __init__(self, nz, nx, lam=1e-8, device="cuda"):
        super().__init__(nz, nx, device)
        self.lam = lam

    def apply_P(self, xv):
        xm = xv.reshape(self.nz, self.nx)
        d1 = torch.diff(xm, dim=0)
        d2 = torch.diff(xm, dim=1)
        out = torch.zeros_like(xm)
        out[1:, :] += d1
        out[:-1, :] -= d1
        out[:, 1:] += d2
        out[:, :-1] -= d2
        return out.reshape(-1)

    def cost_grad(self, x):
        xv = x[:-1]
        Px = self.apply_P(xv)
        cost = torch.dot(xv, Px)
        g = torch.cat([2 * Px, torch.zeros(1, dtype=x.dtype, device=x.device)])
        return self.lam * cost, self.lam * g

    def cost_rows(self, X):
        """lam * x^T P x per row, P x applied to all rows at once (x^T P x =
        |D_z x|^2 + |D_x x|^2)."""
        xm = X[:, :-1].reshape(-1, self.nz, self.nx)
        d1 = torch.diff(xm, dim=1)
        d2 = torch.diff(xm, dim=2)
        return self.lam * ((d1 * d1).sum(dim=(1, 2)) + (d2 * d2).sum(dim=(1, 2)))


_SE_CACHE: dict = {}


def _se_factors(nz, nx, length_scale, noise, sigma, sigma_mean, sigma_slope, transform, svd_cutoff, device):
    """K^-1, or the SVD-mode basis, of the SE kernel: factored once per
    parameter set on the host (LAPACK, float64) and shared read-only on each
    device, as the reference prepares its prior once (Prior.prep,
    core/seis_prior.py:29-35)."""
    key = (nz, nx, length_scale, noise, sigma, sigma_mean, sigma_slope, transform, svd_cutoff, str(device))
    if key in _SE_CACHE:
        return _SE_CACHE[key]
    y = (torch.arange(nz, dtype=torch.float64) - nz // 2)[:, None] + torch.zeros(1, nx, dtype=torch.float64)
    x = (torch.arange(nx, dtype=torch.float64) - nx // 2)[None, :] + torch.zeros(nz, 1, dtype=torch.float64)
    x = x.reshape(-1, 1)
    y = y.reshape(-1, 1)
    d2 = (x - x.T) ** 2 + (y - y.T) ** 2
    K = sigma ** 2 * torch.exp(-d2 / (2 * length_scale ** 2)) + sigma_mean ** 2
    K = K + sigma_slope ** 2 * (y @ y.T) + noise ** 2 * torch.eye(nz * nx, dtype=torch.float64)
    if transform:
        s, U = torch.linalg.eigh(K)
        s, U = s.flip(0), U.flip(1)
        keep = s > svd_cutoff
        k = int(keep.sum())
        B = torch.zeros(nz * nx + 1, k + 1, dtype=torch.float64)
        B[:-1, :k] = U[:, keep] * torch.sqrt(s[keep])[None, :]
        B[-1, k] = 1.0
        out = (B.to(device), None)
    else:
        out = (None, torch.linalg.inv(K).to(device))
    _SE_CACHE[key] = out
    return out


class SquaredExponential(Prior):
    """Gaussian-process prior (core/seis_prior.py:178-242).

    K = sigma^2 exp(-d^2 / 2 l^2) + sigma_mean^2 + sigma_slope^2 y y^T + noise^2 I
    over the (nz, nx) grid centred as in the reference (:199-207).  Without
    ``transform`` the basis is the identity and P = K^-1 (:210); with it the
    basis is U sqrt(s) over the modes of K with s > svd_cutoff, plus min_vel,
    and P = I (:213-219).  K is symmetric positive definite, so its SVD is
    its eigendecomposition: the modes come from ``eigh``, in descending
    order; a mode's sign may differ from the SVD's, which leaves every
    velocity and cost unchanged.  The factors are computed once per
    parameter set and device (``_se_factors``) and shared."""

    def __init__(self, nz, nx, lam=1e-8, length_scale=math.log(32.4), noise=0.1, sigma=183.4,
                 sigma_mean=520.0, sigma_slope=31.4, transform=False, svd_cutoff=0.0, device="cuda"):
        super().__init__(nz, nx, device)
        self.lam = lam
        self.transform = transform
        B, P = _se_factors(nz, nx, float(length_scale), float(noise), float(sigma), float(sigma_mean),
                           float(sigma_slope), bool(transform), float(svd_cutoff), self.device)
        if transform:
            self.basis = B
        self.P = P                                 # None: identity on the modes

    def apply_P(self, xv):
        return xv.clone() if self.P is None else self.P @ xv

    def cost_grad(self, x):
        xv = x[:-1]
        Px = self.apply_P(xv)
        cost = torch.dot(xv, Px)
        g = torch.cat([2 * Px, torch.zeros(1, dtype=x.dtype, device=x.device)])
        return self.lam * cost, self.lam * g

    def cost_rows(self, X):
        Xv = X[:, :-1]
        PX = Xv if self.P is None else Xv @ self.P     # P symmetric
        return self.lam * (Xv * PX).sum(dim=1)


def label_flat(grids, threshold: float = 1.0):
    """Flat areas of velocity grids on the GPU (libseis ``seis_label_flat``,
    the label_thresholded_components of core/seis_numerics.py:81-124 with
    4-connectivity): neighbours join when their velocities (compared in
    float64) differ by at most ``threshold``.  ``grids`` is a list of (nz, nx)
    tensors; returns (labels, counts): int64 (nz, nx) label tensors on the
    grids' device, numbered as scipy numbers components, and the component
    counts."""
    import ctypes

    from . import _lib
    if not grids:
        return [], []
    dev = grids[0].device
    if dev.type != "cuda":
        raise RuntimeError("label_flat needs a ROCm GPU (torch.cuda)")
    flat = torch.cat([g.reshape(-1).to(device=dev, dtype=torch.float64) for g in grids]).contiguous()
    n = len(grids)
    nz = (ctypes.c_int32 * n)(*[int(g.shape[0]) for g in grids])
    nx = (ctypes.c_int32 * n)(*[int(g.shape[1]) for g in grids])
    labels = torch.empty(flat.numel(), dtype=torch.int32, device=dev)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    with torch.cuda.device(dev):
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().seis_label_flat(nz, nx, n, ctypes.c_void_p(flat.data_ptr()), float(threshold),
                                              ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(counts.data_ptr()),
                                              stream), "seis_label_flat")
    out, pos = [], 0
    for g in grids:
        k = int(g.shape[0]) * int(g.shape[1])
        out.append(labels[pos:pos + k].to(torch.int64).reshape(g.shape[0], g.shape[1]))
        pos += k
    return out, counts.cpu().tolist()


class RestrictFlatAreas(Prior):
    """Flat areas of a velocity guess become single parameters
    (core/seis_prior.py:244-281).  ``adapt(vec_guess)`` labels the areas on
    the GPU (``label_flat``) and sets the basis of _adapt (:258-272): one
    column per area (indicator / area size, so columns sum to 1) and min_vel;
    the basis is applied through the labels instead of a dense matrix.
    Costs and gradients are the underlying prior's (identity basis, :253)
    at basis @ x, mapped back by basis^T (:274-281).  Before ``adapt`` the
    basis is the identity."""

    def __init__(self, underlying: Prior, diff_threshold: float = 1.0):
        super().__init__(underlying.nz, underlying.nx, underlying.device)
        if underlying.basis is not None:
            raise ValueError("the underlying prior must have the identity basis (core/seis_prior.py:253)")
        self.underlying = underlying
        self.diff_threshold = diff_threshold
        self.lam = underlying.lam
        self.labels = None

    @property
    def N(self) -> int:
        return self.nz * self.nx + 1 if self.labels is None else self.count + 1

    @property
    def underlying_prior(self):
        """The reference's attribute name (core/seis_prior.py:247)."""
        return self.underlying

    def adapt(self, vec_guess):
        """_adapt (:256-272) on a parameter vector, or on the reference's
        Velocity object (``.data`` the (nz, nx) grid, ``.min_vel``)."""
        n = self.nz * self.nx
        if hasattr(vec_guess, "data") and hasattr(vec_guess, "min_vel") and not torch.is_tensor(vec_guess):
            vec_guess = torch.as_tensor(np.asarray(vec_guess.data, dtype=np.float64)).reshape(-1)
        v = torch.as_tensor(vec_guess).reshape(-1)[:n].to(self.device, torch.float64).reshape(self.nz, self.nx)
        (lab,), (count,) = label_flat([v], self.diff_threshold)
        self.labels = lab.reshape(-1)
        self.count = count
        self.sizes = torch.bincount(self.labels, minlength=count).to(torch.float64)
        return lab, count

    def _scatter(self, g):          # sum over each area of the velocity part
        out = torch.zeros(self.count, dtype=g.dtype, device=g.device)
        return out.index_add_(0, self.labels, g)

    def to_basis(self, vec):
        """solve(B^T B, B^T vec): B^T B is diagonal (1/size), so each area's
        coordinate is the sum of its velocities."""
        if self.labels is None:
            return vec.clone()
        n = self.nz * self.nx
        return torch.cat([self._scatter(vec[:n]), vec[n:]])

    def is_identity(self) -> bool:
        return self.labels is None

    @property
    def basis_vectors(self) -> torch.Tensor:
        """The dense basis of _adapt (:258-272) once adapted."""
        return self.dense_basis() if self.labels is not None else Prior.basis_vectors.fget(self)

    def from_basis(self, x):
        if self.labels is None:
            return x
        return torch.cat([x[self.labels] / self.sizes[self.labels], x[-1:]])

    def from_basis_rows(self, X):
        if self.labels is None:
            return X
        return torch.cat([X[:, self.labels] / self.sizes[self.labels], X[:, -1:]], dim=1)

    def basis_T(self, g):
        if self.labels is None:
            return g
        n = self.nz * self.nx
        return torch.cat([self._scatter(g[:n]) / self.sizes, g[n:]])

    def dense_basis(self) -> torch.Tensor:
        """The (nz*nx+1) x N basis matrix of _adapt (for callers that take a
        dense basis)."""
        n = self.nz * self.nx
        if self.labels is None:
            return torch.eye(n + 1, dtype=torch.float64, device=self.device)
        B = torch.zeros(n + 1, self.count + 1, dtype=torch.float64, device=self.device)
        B[torch.arange(n, device=self.device), self.labels] = 1.0 / self.sizes[self.labels]
        B[n, self.count] = 1.0
        return B

    def cost_grad(self, x):
        c, g = self.underlying.cost_grad(self.from_basis(x))
        return c, self.basis_T(g)

    def cost_rows(self, X):
        return self.underlying.cost_rows(self.from_basis_rows(X))

    def apply_P(self, xv):
        """B_v^T P B_v on the area coordinates (Gauss-Newton), for an
        underlying prior with a precision matrix."""
        if self.labels is None:
            return self.underlying.apply_P(xv)
        Pv = self.underlying.apply_P(xv[self.labels] / self.sizes[self.labels])
        return self._scatter(Pv) / self.sizes
