"""Pins the FP64 oracle (restatement of core/seis_forward2.py, seis_numerics.py,
seis_invert.py, seis_prior.py) before it is trusted as the parity reference.

The reference's own tests need competition data that is absent here
(core/seis_test.py:120), so each oracle function is pinned by closed forms,
exact adjoint identities, brute force on tiny grids, physical symmetries and a
library special case (torch.optim.LBFGS).  The reference's own recipes are
restated where they apply (dot test and Taylor test, core/seis_test.py:26-44;
prior and cost FD tests, :48-89)."""
import os

import numpy as np
import pytest
import torch

from oracle import fd2d, invert, priors
from seismic_amd import synth

TINY = dict(nz=4, nx=5, nbc=3, nt=150)     # 10 x 11 padded, 2 shots
SMALL = dict(nz=12, nx=16, nbc=8, nt=200)


def _layered(g, seed=0):
    rng = np.random.default_rng(seed)
    v = np.repeat(rng.uniform(1500, 4500, (g.nz, 1)), g.nx, 1)
    v += rng.uniform(-100, 100, (g.nz, g.nx))
    return fd2d.default_vec(g, v)


# ------------------------------------------------------------ closed forms
def test_ricker_closed_form():
    """core/seis_forward2.py:20-40 with f=15, dt=1e-3: nw=147, nc=74, peak 1 at k=nc."""
    w = fd2d.ricker(15, 1e-3)
    assert len(w) == 147
    assert w[73] == 1.0                      # k = nc = 74 -> beta = 0
    np.testing.assert_allclose(w[:73], w[74:][::-1], rtol=0, atol=1e-15)
    k = 10
    beta = ((74 - k) * 15 * 1e-3 * np.pi) ** 2
    assert w[k - 1] == pytest.approx((1 - 2 * beta) * np.exp(-beta), rel=1e-14)
    s = fd2d.source_wavelet(fd2d.Geometry())
    assert len(s) == 999 and np.all(s[147:] == 0)


def test_damping_profile_and_corners():
    """AbcCoef2D (:46-59): kappa0 = 3 ln(1e7)/(2 (nbc-1) dx), quadratic, corners use x."""
    g = fd2d.Geometry()
    d = fd2d.damping(g)
    k0 = 3 * np.log(1e7) / (2 * 1190.0)
    assert k0 == pytest.approx(0.0203169, rel=1e-5)
    assert d.shape == (310, 310)
    assert d[0, 150] == pytest.approx(k0)            # outer top row, interior column
    assert d[150, 0] == pytest.approx(k0)            # outer left column
    assert d[119, 150] == 0.0 and d[150, 119] == 0.0  # innermost pad cell
    assert np.all(d[120:190, 120:190] == 0)
    # corner cell (z=5, x=3): x profile (j = 116), not the z profile (j = 114)
    assert d[5, 3] == pytest.approx(k0 * (116 / 119) ** 2)
    assert d[300, 305] == pytest.approx(k0 * (115 / 119) ** 2)


def test_reference_geometry_indices():
    """Shots at x=[0,17,34,52,69] on padded row 121; receivers row 121, cols 120..189."""
    g = fd2d.Geometry()
    assert g.src_x == (0, 17, 34, 52, 69)
    assert g.src_cells() == [(121, 120), (121, 137), (121, 154), (121, 172), (121, 189)]
    igz, igx = g.rcv_cells()
    assert np.all(igz == 121) and np.array_equal(igx, np.arange(120, 190))
    assert g.src_cells()[2][0] * 310 + g.src_cells()[2][1] == 121 * 310 + 154   # src_idx (:94)


def test_prep_run_algebra():
    g = fd2d.Geometry(**SMALL)
    vec = _layered(g)
    v, alpha, t1, t2 = fd2d.prep_run(vec, g)
    damp = fd2d.damping(g)
    kappa = vec[-1] * damp * g.dt
    np.testing.assert_allclose(alpha, (v * g.dt / g.dx) ** 2, rtol=1e-15)
    np.testing.assert_allclose(t1, 2 - 5 * alpha - kappa, rtol=1e-14)
    np.testing.assert_allclose(t2, 1 - kappa, rtol=1e-15)
    # edge padding
    vv = vec[:-1].reshape(g.nz, g.nx)
    assert v[0, 0] == vv[0, 0] and v[-1, -1] == vv[-1, -1] and v[g.nbc + 3, 0] == vv[3, 0]


def test_prep_run_diff_is_derivative():
    """prep_run_diff (:346-359) equals the central difference of prep_run."""
    g = fd2d.Geometry(**SMALL)
    vec = _layered(g, 1)
    dvec = np.random.default_rng(2).normal(size=vec.shape)
    v, _, _, _ = fd2d.prep_run(vec, g)
    _, dt1, dt2, da = fd2d.prep_run_diff(dvec, v, g)
    h = 1e-3
    p = fd2d.prep_run(vec + h * dvec, g)
    m = fd2d.prep_run(vec - h * dvec, g)
    for ana, i in [(da, 1), (dt1, 2), (dt2, 3)]:
        np.testing.assert_allclose(ana, (p[i] - m[i]) / (2 * h), rtol=1e-7, atol=1e-12)


def test_unpad_is_adjoint_of_edge_pad():
    """<pad_edge(x), y> == <x, unpad(y)> (core/seis_numerics.py:14-79)."""
    rng = np.random.default_rng(3)
    for nz, nx, nbc in [(7, 9, 4), (2, 3, 5), (70, 70, 120)]:
        x = rng.normal(size=(nz, nx))
        y = rng.normal(size=(nz + 2 * nbc, nx + 2 * nbc))
        a = np.sum(np.pad(x, nbc, mode="edge") * y)
        b = np.sum(x * fd2d.unpad_edge_padded_gradient(y, nbc))
        assert a == pytest.approx(b, rel=1e-12)


def test_first_steps_closed_form():
    """p[2] = s_mod[0] delta_src; p[3] = temp1 p[2] + alpha lapg(p[2]) + s_mod[1] delta_src.
    Pins the injection amplitude (v_src dt)^2 (:206), the output index t <-> p[t+2]
    (:249) and the stencil coefficients c2, c3 (:80-81)."""
    g = fd2d.Geometry(nz=10, nx=10, nbc=6, nt=150, n_shots=1, src_x=(4,), rcv_x=tuple(range(10)))
    vec = _layered(g, 4)
    v, alpha, t1, t2 = fd2d.prep_run(vec, g)
    s = fd2d.source_wavelet(g)
    seis = fd2d.vel_to_seis(vec, g)[0].reshape(g.nt, g.n_rcv)
    zs, xs = g.src_cells()[0]
    amp = (v[zs, xs] * g.dt) ** 2
    p2 = amp * s[0]
    # receivers share the source row: sample t=0 is p[2] on that row
    assert seis[0, 4] == pytest.approx(p2, rel=1e-15)
    assert np.all(seis[0, np.arange(10) != 4] == 0)
    p3_src = t1[zs, xs] * p2 + amp * s[1]
    assert seis[1, 4] == pytest.approx(p3_src, rel=1e-13)
    assert seis[1, 5] == pytest.approx(alpha[zs, xs + 1] * (4 / 3) * p2, rel=1e-13)
    assert seis[1, 6] == pytest.approx(alpha[zs, xs + 2] * (-1 / 12) * p2, rel=1e-13)
    assert seis[1, 7] == 0.0


def test_homogeneous_arrival_time():
    """Direct-wave moveout between two far-field receivers is their offset
    difference over v (pins the (dt/dx)^2 scaling of alpha; the 2-D phase lag
    cancels in the difference)."""
    g = fd2d.Geometry(nz=30, nx=120, nbc=30, nt=650, n_shots=1, src_x=(10,))
    for v in (2000.0, 3000.0):
        vec = fd2d.default_vec(g, np.full((g.nz, g.nx), v))
        seis = fd2d.vel_to_seis(vec, g)[0].reshape(g.nt, g.n_rcv)
        t60, t100 = np.argmax(seis[:, 60]), np.argmax(seis[:, 100])
        expected = 40 * g.dx / v / g.dt
        assert abs((t100 - t60) - expected) <= 2, (v, t100 - t60, expected)


# ------------------------------------------------------------ adjoint identities
def test_dot_product_reference_recipe():
    """core/seis_test.py:26-38: dm = 1e-4 N(0,1) seed 0, y = data + N(0,1) seed 0;
    |<J dm, s-y> - <dm, J^T (s-y)>| / |<J dm, s-y>| < 1e-10 in FP64."""
    g = fd2d.Geometry(**SMALL)
    vec = _layered(g, 5)
    data = fd2d.vel_to_seis(fd2d.default_vec(g, vec[:-1].reshape(g.nz, g.nx) * 1.03), g)[0]
    dm = 1e-4 * np.random.default_rng(seed=0).normal(0, 1, vec.shape)
    y = data + np.random.default_rng(seed=0).normal(0, 1, data.shape)
    s, ds, gr = fd2d.vel_to_seis(vec, g, vec_diff=dm, vec_adjoint=y, adjoint_on_residual=True)
    a = np.sum((s - y) * ds)
    b = np.sum(gr * dm)
    assert abs(a - b) / abs(a) < 1e-10


def test_taylor_reference_recipe():
    """core/seis_test.py:41-44: rms(s(v+dm)-s(v)-J dm)/rms(s(v+dm)-s(v)) < 1e-5."""
    g = fd2d.Geometry(**SMALL)
    vec = _layered(g, 6)
    dm = 1e-4 * np.random.default_rng(seed=0).normal(0, 1, vec.shape)
    s, ds, _ = fd2d.vel_to_seis(vec, g, vec_diff=dm)
    s2 = fd2d.vel_to_seis(vec + dm, g)[0]
    rms = lambda a: np.sqrt(np.mean(a ** 2))
    assert rms(s2 - s - ds) / rms(s2 - s) < 1e-5


def test_brute_force_jacobian_tiny():
    """Column-by-column central-difference Jacobian (every velocity and min_vel)
    against the tangent pass, and J^T e_k against the adjoint pass."""
    g = fd2d.Geometry(**TINY, n_shots=2)
    vec = _layered(g, 7)
    n = len(vec)
    h = 1e-2
    J_fd = np.stack([(fd2d.vel_to_seis(vec + h * e, g)[0] - fd2d.vel_to_seis(vec - h * e, g)[0]) / (2 * h)
                     for e in np.eye(n)], axis=1)
    J_tan = np.stack([fd2d.vel_to_seis(vec, g, vec_diff=e)[1] for e in np.eye(n)], axis=1)
    np.testing.assert_allclose(J_tan, J_fd, rtol=0, atol=1e-6 * np.abs(J_fd).max())
    rng = np.random.default_rng(8)
    w = rng.normal(size=J_fd.shape[0])
    jt = fd2d.vel_to_seis(vec, g, vec_adjoint=w)[2]
    np.testing.assert_allclose(jt, J_tan.T @ w, rtol=1e-10, atol=1e-12 * np.abs(jt).max())


# ------------------------------------------------------------ symmetries
def test_mirror_symmetry():
    """Laterally symmetric v: shot 0 == flip_x(shot 4) (used by core/seis_model.py:155)."""
    g = fd2d.Geometry(**SMALL)
    rng = np.random.default_rng(9)
    half = rng.uniform(1500, 4500, (g.nz, g.nx // 2))
    v = np.concatenate([half, half[:, ::-1]], axis=1)
    assert g.src_x == (0, 4, 8, 11, 15)
    seis = fd2d.vel_to_seis(fd2d.default_vec(g, v), g)[0].reshape(g.n_shots, g.nt, g.n_rcv)
    np.testing.assert_allclose(seis[0], seis[4][:, ::-1], rtol=0, atol=1e-12 * np.abs(seis).max())
    np.testing.assert_allclose(seis[1], seis[3][:, ::-1], rtol=0, atol=1e-12 * np.abs(seis).max())


def test_reciprocity():
    """Source and receivers on one row: seis_i[t, x_j] == seis_j[t, x_i]
    (symmetric spatial operator after scaling by 1/alpha, source scaled by alpha dx^2)."""
    g = fd2d.Geometry(**SMALL)
    vec = _layered(g, 10)
    seis = fd2d.vel_to_seis(vec, g)[0].reshape(g.n_shots, g.nt, g.n_rcv)
    xs = g.src_x
    scale = np.abs(seis).max()
    for i in range(g.n_shots):
        for j in range(g.n_shots):
            np.testing.assert_allclose(seis[i, :, xs[j]], seis[j, :, xs[i]], rtol=0, atol=1e-10 * scale)


# ------------------------------------------------------------ priors
@pytest.mark.parametrize("cls", [priors.RowTotalVariation, priors.TotalVariation, priors.L2Smoothness])
def test_prior_gradient_fd(cls):
    """core/seis_test.py:48-69: <g, d> vs cost(x+d)-cost(x), d = 1e-6 N(0,1), rel < 1e-6."""
    p = cls(12, 16, lam=15.0)
    x = np.random.default_rng(seed=0).normal(0, 1, p.N)
    d = 1e-6 * np.random.default_rng(seed=0).normal(0, 1, p.N)
    c, gr = p.cost_grad(x)
    c2, _ = p.cost_grad(x + d)
    assert abs(gr @ d - (c2 - c)) / abs(c2 - c) < 1e-6


def test_row_tv_basis():
    p = priors.RowTotalVariation(12, 16)
    assert p.basis.shape == (12 * 16 + 1, 13)
    x = np.arange(13.0)
    v = p.basis @ x
    assert np.all(v[:-1].reshape(12, 16) == np.arange(12.0)[:, None]) and v[-1] == 12.0


def test_l2_smoothness_closed_form():
    p = priors.L2Smoothness(5, 6, lam=2.0)
    xm = np.random.default_rng(1).normal(size=(5, 6))
    x = np.concatenate([xm.ravel(), [3.0]])
    ref = 2.0 * (np.sum(np.diff(xm, axis=0) ** 2) + np.sum(np.diff(xm, axis=1) ** 2))
    assert p.cost_grad(x)[0] == pytest.approx(ref, rel=1e-13)


def test_cost_and_gradient_fd():
    """core/seis_test.py:71-89 with the RowTV prior, lam 1e-9, dx = 1e-4 N(0,1): rel < 1e-4."""
    g = fd2d.Geometry(**SMALL)
    p = priors.RowTotalVariation(g.nz, g.nx, lam=1e-9)
    target = fd2d.vel_to_seis(_layered(g, 11), g)[0]
    base = 3000 + 100 * np.random.default_rng(seed=0).normal(0, 1, p.N)
    base[-1] = 1800
    off = 1e-4 * np.random.default_rng(seed=0).normal(0, 1, p.N)
    c2 = invert.cost_and_gradient(base + off, target, p, p.basis, g)[0]
    c, gr, _, _ = invert.cost_and_gradient(base, target, p, p.basis, g, compute_gradient=True)
    assert abs(gr @ off - (c2 - c)) / abs(c2 - c) < 1e-4


# ------------------------------------------------------------ optimisers
def _rosen(x):
    f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
    g = np.zeros_like(x)
    g[:-1] = -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    g[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return f, g


@pytest.mark.parametrize("max_iter,hist", [(30, None), (60, None), (40, 10)])
def test_bfgs_matches_torch_lbfgs(max_iter, hist):
    """bfgs is a port of torch 2.7 LBFGS (core/seis_numerics.py:126-129): same
    iterates as torch.optim.LBFGS(lr=1, line_search_fn='strong_wolfe') in FP64.
    The installed torch adds a g.d stopping test and caps each line search at
    the remaining evaluation budget; bfgs reproduces both with torch_compat=True
    and without it follows the reference wherever they do not bind."""
    x0 = np.array([-1.2, 1.0, -0.5, 0.8, 1.3, -0.7])
    x_ref = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.LBFGS([x_ref], lr=1, max_iter=max_iter, tolerance_grad=1e-7,
                            tolerance_change=1e-9, history_size=hist or 10 ** 6,
                            line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        f, g = _rosen(x_ref.detach().numpy())
        x_ref.grad = torch.tensor(g)
        return torch.tensor(f)

    opt.step(closure)
    x, _ = invert.bfgs(_rosen, x0, max_iter, 1e-7, 1e-9, history_size=hist, torch_compat=True)
    np.testing.assert_allclose(x, x_ref.detach().numpy(), rtol=1e-8, atol=1e-9)
    if max_iter <= 30:  # neither torch-only rule binds here: reference semantics agree
        x25, _ = invert.bfgs(_rosen, x0, max_iter, 1e-7, 1e-9, history_size=hist)
        np.testing.assert_array_equal(x25, x)


def test_cg_reproduces_dense_solve():
    rng = np.random.default_rng(12)
    M = rng.normal(size=(15, 15))
    A = M @ M.T + 15 * np.eye(15)
    b = rng.normal(size=15)
    x = invert.conjugate_gradient(lambda v: A @ v, b, 15)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)


def test_gn_cg_equals_dense_gn_small_basis():
    """With CG steps >= basis size, GN-CG reproduces the reference's dense GN pass
    (core/seis_invert.py:63-107) on a RowTV-style basis (n = nz+1)."""
    g = fd2d.Geometry(nz=6, nx=8, nbc=5, nt=160)
    p = priors.L2Smoothness(g.nz, g.nx, lam=1e-9)
    # restrict to the RowTV basis so the dense J is cheap; P in that basis
    rt = priors.RowTotalVariation(g.nz, g.nx)
    B = rt.basis

    class RowL2:
        basis = B
        lam = 1e-9
        P = (B[:-1, :-1].T @ p.P @ B[:-1, :-1])

        def cost_grad(self, x):
            xv = x[:-1]
            return self.lam * xv @ self.P @ xv, np.concatenate([2 * self.lam * self.P @ xv, [0.0]])

    prior = RowL2()
    true = fd2d.default_vec(g, np.repeat(np.linspace(1800, 3500, g.nz)[:, None], g.nx, 1))
    target = fd2d.vel_to_seis(true, g)[0]
    start = B @ np.concatenate([np.linspace(2000, 3000, g.nz), [1800.0]])
    v_dense, s_dense, _ = invert.gauss_newton_dense(start, target, prior, g, n_scales=11)
    v_cg, s_cg, _ = invert.gauss_newton_cg(start, target, prior, g, n_cg=B.shape[1] + 4, n_scales=11)
    assert s_dense == s_cg
    np.testing.assert_allclose(v_cg, v_dense, rtol=1e-6)


def test_golden_fixtures_match_oracle():
    """The committed full-size fixtures are what the oracle computes now
    (shot 0 forward, and the fixture's misfit from the oracle's 5-shot forward)."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    f = np.load(os.path.join(gold, "ref70_fwd.npz"))
    d = np.load(os.path.join(gold, "ref70_grad.npz"))
    g = fd2d.Geometry()
    vec = synth.to_vec(synth.curve_fault_model(70, 70, seed=7))
    s0 = fd2d.forward_only(vec, g, shots=[0])
    assert np.array_equal(s0.reshape(g.nt, g.n_rcv).astype(np.float32), f["seis"][0])
    full = fd2d.forward_only(vec, g).reshape(-1)
    m = float(np.sum((full - d["target"].astype(np.float64)) ** 2))
    assert abs(m - float(d["misfit"])) <= 1e-12 * float(d["misfit"])
    assert d["grad"].shape == (g.nz * g.nx + 1,) and np.all(np.isfinite(d["grad"]))


def _scipy_labels(A, X, connectivity):
    """The reference's construction (core/seis_numerics.py:98-120): directed
    adjacency for every pixel and neighbour with |diff| <= X, then scipy's
    connected_components on the undirected graph."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    H, W = A.shape
    neigh = [(1, 0), (-1, 0), (0, 1), (0, -1)]
    if connectivity == 8:
        neigh += [(1, 1), (1, -1), (-1, 1), (-1, -1)]
    r, c = [], []
    for i in range(H):
        for j in range(W):
            for di, dj in neigh:
                ni, nj = i + di, j + dj
                if 0 <= ni < H and 0 <= nj < W and abs(A[ni, nj] - A[i, j]) <= X:
                    r.append(i * W + j)
                    c.append(ni * W + nj)
    adj = coo_matrix((np.ones(len(r), dtype=bool), (r, c)), shape=(H * W, H * W)).tocsr()
    n, lab = connected_components(adj, directed=False, return_labels=True)
    return lab.reshape(H, W), n


@pytest.mark.parametrize("shape,X,conn", [((9, 13), 1.0, 4), ((20, 20), 1.0, 4), ((17, 6), 0.0, 4),
                                          ((12, 12), 2.0, 8), ((1, 30), 1.0, 4), ((30, 1), 1.0, 4)])
def test_label_thresholded_components_matches_scipy(shape, X, conn):
    """The union-find restatement gives scipy's labels (numbering included)
    on integer-valued grids, where steps equal to the threshold occur."""
    rng = np.random.default_rng(shape[0] * 31 + shape[1])
    for trial in range(4):
        A = np.round(rng.normal(0, 1.2 + trial, shape))
        lab, n = priors.label_thresholded_components(A, X, conn)
        ref, nref = _scipy_labels(A, X, conn)
        assert n == nref
        assert np.array_equal(lab, ref)


def test_label_thresholded_components_special_cases():
    flat = np.full((5, 7), 3.0)
    lab, n = priors.label_thresholded_components(flat, 0.0)
    assert n == 1 and np.all(lab == 0)
    ramp = np.arange(35, dtype=np.float64).reshape(5, 7) * 2.0
    lab, n = priors.label_thresholded_components(ramp, 1.0)
    assert n == 35 and np.array_equal(lab.ravel(), np.arange(35))
    two = np.zeros((4, 6))
    two[:, 3:] = 10.0
    lab, n = priors.label_thresholded_components(two, 1.0)
    assert n == 2 and np.all(lab[:, :3] == 0) and np.all(lab[:, 3:] == 1)


def test_restrict_flat_areas_basis_and_gradient():
    """Basis of RestrictFlatAreas (core/seis_prior.py:258-272): one column per
    flat area plus min_vel, each summing to 1 and each pixel in exactly one
    column; its gradient (:274-281) matches central differences."""
    rng = np.random.default_rng(3)
    nz, nx = 8, 9
    v = np.round(rng.normal(2000, 1.0, (nz, nx)))
    v[:, :4] += 300.0
    p = priors.RestrictFlatAreas(priors.TotalVariation(nz, nx, lam=1.0), diff_threshold=1.0)
    labels, count = p.adapt(np.concatenate([v.ravel(), [1500.0]]))
    B = p.basis
    assert B.shape == (nz * nx + 1, count + 1)
    np.testing.assert_allclose(B.sum(axis=0), 1.0)
    assert np.all((B[:-1] > 0).sum(axis=1) == 1)
    for k in range(count):
        assert np.all((B[:-1, k] > 0) == (labels.ravel() == k))
    x = rng.normal(0, 100, p.N)
    c, g = p.cost_grad(x)
    h = 1e-3
    for k in rng.choice(p.N, 5, replace=False):
        e = np.zeros(p.N)
        e[k] = h
        fd = (p.cost_grad(x + e)[0] - p.cost_grad(x - e)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k]))


def test_squared_exponential_oracle_closed_forms():
    """SE prior (core/seis_prior.py:196-220): K P = I; with every SVD mode
    kept, B_v B_v^T = K; the product's eigh modes span the same space (B B^T
    equal) and give the same costs."""
    from seismic_amd import priors as ppri
    o = priors.SquaredExponential(6, 7, lam=1.5)
    np.testing.assert_allclose(o.K @ o.P, np.eye(42), atol=1e-6)
    ot = priors.SquaredExponential(6, 7, lam=1.5, transform=True, svd_cutoff=0.0)
    Bv = ot.basis[:-1, :-1]
    np.testing.assert_allclose(Bv @ Bv.T, o.K, rtol=1e-9, atol=1e-6 * np.abs(o.K).max())
    for cut in (1.0, 1e4):
        ot = priors.SquaredExponential(6, 7, lam=1.5, transform=True, svd_cutoff=cut)
        pt = ppri.SquaredExponential(6, 7, lam=1.5, transform=True, svd_cutoff=cut, device="cpu")
        Bp = pt.basis.numpy()
        assert Bp.shape == ot.basis.shape
        np.testing.assert_allclose(Bp @ Bp.T, ot.basis @ ot.basis.T, atol=1e-7 * np.abs(o.K).max())
        x = np.random.default_rng(1).normal(0, 3, ot.N)
        assert float(pt.cost_grad(torch.tensor(x))[0]) == pytest.approx(ot.cost_grad(x)[0], rel=1e-12)


def test_dispatch_kpis_on_constructed_cases():
    """The dispatch KPIs (core/seis_model.py:133, :204-215) on inputs whose
    values follow from the definitions: a mirrored shot pair gives 0 and an
    offset c on the last shot gives exactly |c|; round(v/3) bins of known
    sizes; areas of known sizes separated by steps just above 10 m/s, joined
    by steps of exactly 10."""
    rng = np.random.default_rng(3)
    s = rng.normal(0, 1, (5, 30, 9))
    s[4] = s[0][:, ::-1]
    assert invert.kpi_flat_vel(s.ravel(), 5, 30, 9) == 0.0
    s[4] += 0.25
    assert invert.kpi_flat_vel(s.ravel(), 5, 30, 9) == pytest.approx(0.25, rel=1e-12)
    s[1] += 7.0                                   # only shots 0 and 4 enter
    assert invert.kpi_flat_vel(s.ravel(), 5, 30, 9) == pytest.approx(0.25, rel=1e-12)

    # bins: 3000 and 3001 share bin 1000 (17 cells), 3003 is bin 1001 (12), 3006 bin 1002 (6)
    v = np.array([3000.0] * 10 + [3001.0] * 7 + [3003.0] * 12 + [3006.0] * 6)
    assert invert.kpi_style_b(rng.permutation(v).reshape(5, 7)) == 17

    # three vertical strips of widths 3, 5, 2 on 6 rows, steps 10.5 between them
    g = np.zeros((6, 10))
    g[:, 3:8] = 10.5
    g[:, 8:] = 21.0
    assert invert.kpi_fault_a(g) == 6 * 5 + 6 * 3
    g[:, 8:] = 20.5                               # a step of exactly 10 joins the last two strips
    assert invert.kpi_fault_a(g) == 6 * 7 + 6 * 3
    assert invert.kpi_fault_a(np.full((4, 4), 2000.0)) == 16    # one area: every cell
    assert invert.kpi_fault_a(np.arange(12.0).reshape(3, 4) * 11) == 2    # all singletons


# ------------------------------------------------------------ GN, priors and the LBFGS closure by closed forms
class LinearV2S:
    """s(vec) = J vec with the vel_to_seis call convention (J dvec, J^T dd,
    J^T (s - dd)): a propagator whose Gauss-Newton model is exact."""

    def __init__(self, J):
        self.J = J

    def __call__(self, vec, vec_diff=None, vec_adjoint=None, adjoint_on_residual=False):
        s = self.J @ vec
        ds = None if vec_diff is None else self.J @ vec_diff
        gr = None
        if vec_adjoint is not None:
            gr = self.J.T @ ((s - vec_adjoint) if adjoint_on_residual else vec_adjoint)
        return s, ds, gr


def _row_l2(nz, nx, lam):
    """L2 smoothness restricted to the row basis (rows constant, min_vel)."""
    p = priors.L2Smoothness(nz, nx, lam=lam)
    B = priors.RowTotalVariation(nz, nx).basis

    class RowL2:
        basis = B
        P = B[:-1, :-1].T @ p.P @ B[:-1, :-1]
        D = p.D @ B[:-1, :-1]

        def cost_grad(self, x):
            xv = x[:-1]
            return self.lam * xv @ self.P @ xv, np.concatenate([2 * self.lam * self.P @ xv, [0.0]])

    RowL2.lam = lam
    return RowL2()


def _gn_closed_form(prior, J, t):
    """argmin_x lam |D x_v|^2 + |J B x - t|^2 / N (x_v = velocity part),
    from a least-squares solve of the stacked system: the minimiser one
    Gauss-Newton step reaches on a linear model."""
    N = len(t)
    B = prior.basis
    D = np.hstack([prior.D, np.zeros((prior.D.shape[0], 1))])
    M = np.vstack([np.sqrt(prior.lam) * D, J @ B / np.sqrt(N)])
    rhs = np.concatenate([np.zeros(D.shape[0]), t / np.sqrt(N)])
    return B @ np.linalg.lstsq(M, rhs, rcond=None)[0]


@pytest.mark.parametrize("row_basis", [False, True])
def test_gauss_newton_reaches_closed_form_minimiser(row_basis):
    """On a linear model s = J vec a GN pass (core/seis_invert.py:63-107)
    lands on the minimiser of lam x^T P x + |J B x - t|^2 / N: with 13 line
    search points the grid holds scale 1 (to 2e-16).  A slip in the rhs or A
    scaling (a factor N, lam, 2 or a sign) moves the result."""
    rng = np.random.default_rng(5)
    nz, nx, lam = 4, 5, 0.37
    prior = _row_l2(nz, nx, lam) if row_basis else priors.L2Smoothness(nz, nx, lam=lam)
    n = nz * nx + 1
    J = rng.normal(size=(90, n))
    t = rng.normal(size=90) * 10
    start = prior.basis @ rng.normal(size=prior.basis.shape[1])     # inside the basis' span
    ref = _gn_closed_form(prior, J, t)
    v2s = LinearV2S(J)
    g = fd2d.Geometry(**SMALL)
    vd, sd, _ = invert.gauss_newton_dense(start, t, prior, g, v2s=v2s, n_scales=13)
    vc, sc, _ = invert.gauss_newton_cg(start, t, prior, g, n_cg=prior.basis.shape[1] + 5, v2s=v2s, n_scales=13)
    assert abs(sd - 1.0) < 1e-15 and abs(sc - 1.0) < 1e-15
    np.testing.assert_allclose(vd, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    np.testing.assert_allclose(vc, ref, rtol=1e-7, atol=1e-7 * np.abs(ref).max())


def test_gn_rhs_is_minus_half_cost_gradient():
    """The GN right-hand side (:77-78) is -1/2 the gradient of the cost
    lam x^T P x + mean((s - t)^2) (:19-49), whose gradient is pinned by finite
    differences here (propagator: the oracle on a small grid)."""
    g = fd2d.Geometry(**SMALL)
    p = priors.L2Smoothness(g.nz, g.nx, lam=3e-10)
    target = fd2d.vel_to_seis(_layered(g, 11), g)[0]
    x = _layered(g, 4) + np.concatenate([20 * np.random.default_rng(3).normal(size=g.nz * g.nx), [0.0]])
    c, gr, _, _ = invert.cost_and_gradient(x, target, p, p.basis, g, compute_gradient=True)
    off = 1e-3 * np.random.default_rng(7).normal(0, 1, p.N)
    c2 = invert.cost_and_gradient(x + off, target, p, p.basis, g)[0]
    c1 = invert.cost_and_gradient(x - off, target, p, p.basis, g)[0]
    assert abs(gr @ off - (c2 - c1) / 2) / abs((c2 - c1) / 2) < 1e-6
    rhs = invert.gn_rhs(x, target, p, lambda vec, **kw: fd2d.vel_to_seis(vec, g, **kw))
    np.testing.assert_allclose(rhs, -0.5 * gr, rtol=1e-11, atol=1e-12 * np.abs(gr).max())


def test_tv_priors_closed_forms():
    """Costs of core/seis_prior.py:67-177 on fields whose value is known: a
    constant field costs lam*eps (every smoothed |difference| is eps); a ramp
    of step h along z costs lam*sqrt(h^2 + eps^2) for the row prior (mean over
    the nz-1 row differences), and for the 2-D prior the mean over all
    (nz-1)*nx vertical and nz*(nx-1) horizontal differences."""
    nz, nx, lam, eps, h = 7, 9, 2.5, 0.1, 3.0
    rt, tv = priors.RowTotalVariation(nz, nx, lam=lam, epsilon=eps), priors.TotalVariation(nz, nx, lam=lam, epsilon=eps)
    const = np.full(nz + 1, 1234.0)
    assert rt.cost_grad(const)[0] == pytest.approx(lam * eps, rel=1e-14)
    assert tv.cost_grad(np.full(nz * nx + 1, 1234.0))[0] == pytest.approx(lam * eps, rel=1e-14)
    ramp_rows = np.concatenate([1500.0 + h * np.arange(nz), [1500.0]])
    assert rt.cost_grad(ramp_rows)[0] == pytest.approx(lam * np.sqrt(h * h + eps * eps), rel=1e-14)
    ramp = np.concatenate([np.repeat(1500.0 + h * np.arange(nz), nx), [1500.0]])
    n1, n2 = (nz - 1) * nx, nz * (nx - 1)
    ref = lam * (n1 * np.sqrt(h * h + eps * eps) + n2 * eps) / (n1 + n2)
    assert tv.cost_grad(ramp)[0] == pytest.approx(ref, rel=1e-14)
    # the product's priors are the same functions
    from seismic_amd import priors as ppri
    prt = ppri.RowTotalVariation(nz, nx, lam=lam, epsilon=eps, device="cpu")
    ptv = ppri.TotalVariation(nz, nx, lam=lam, epsilon=eps, device="cpu")
    assert float(prt.cost_grad(torch.tensor(ramp_rows))[0]) == pytest.approx(lam * np.sqrt(h * h + eps * eps), rel=1e-14)
    assert float(ptv.cost_grad(torch.tensor(ramp))[0]) == pytest.approx(ref, rel=1e-14)


def test_lbfgs_closure_scaling_and_error_tolerance(monkeypatch):
    """The closure of seis_to_vel_lbfgs (core/seis_invert.py:125-138): cost and
    gradient of cost_and_gradient scaled by 1e15, and both zeroed when the
    data misfit mean((s - t)^2) is below seis_error_tolerance."""
    rng = np.random.default_rng(8)
    nz, nx = 3, 4
    p = priors.RowTotalVariation(nz, nx, lam=1e-3)
    J = rng.normal(size=(50, nz * nx + 1))
    v2s = LinearV2S(J)
    x = np.concatenate([rng.normal(2000, 100, nz), [1500.0]])
    t_exact = J @ (p.basis @ x)
    seen = {}

    def fake_bfgs(func, x0, *a, **k):
        seen["func"] = func
        return x0, 1

    monkeypatch.setattr(invert, "bfgs", fake_bfgs)
    g = fd2d.Geometry(**SMALL)
    t_off = t_exact + 0.5
    invert.lbfgs_invert(p.basis @ x, t_off, p, g, max_iter=1, v2s=v2s, seis_error_tolerance=0.2)
    f, gr = seen["func"](x)
    cost_prior = p.cost_grad(x)[0]
    assert f == pytest.approx(1e15 * (cost_prior + 0.25), rel=1e-13)       # mean residual^2 = 0.25 > 0.2
    ref_g = 1e15 * (p.cost_grad(x)[1] + (2.0 / 50) * p.basis.T @ (J.T @ np.full(50, -0.5)))   # s - t = -0.5
    np.testing.assert_allclose(gr, ref_g, rtol=1e-12)
    invert.lbfgs_invert(p.basis @ x, t_off, p, g, max_iter=1, v2s=v2s, seis_error_tolerance=0.3, scaling=2.0)
    f, gr = seen["func"](x)
    assert f == 0.0 and not np.any(gr)                                     # 0.25 < 0.3: zeroed
    invert.lbfgs_invert(p.basis @ x, t_exact, p, g, max_iter=1, v2s=v2s, scaling=2.0)
    f, gr = seen["func"](x)
    assert f == pytest.approx(2.0 * cost_prior, rel=1e-13)                 # tolerance 0: never zeroed
