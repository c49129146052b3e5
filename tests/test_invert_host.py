"""Host logic of the inversion loops, on the CPU: the product's machines
(seismic_amd.invert) driven by the FP64 oracle propagator must reproduce the
oracle's restatements of core/seis_numerics.py:307-448 (bfgs) and
core/seis_invert.py:63-152 (GN, LBFGS) step for step."""
import numpy as np
import pytest
import torch

import host_lbfgs
from oracle import fd2d
from oracle import invert as oinv
from oracle import priors as opri
from seismic_amd import invert, priors, synth
from seismic_amd.geometry import Geometry

KW = dict(nz=10, nx=12, nbc=8, nt=160)


class OracleExecutor:
    """Executes machine ops with the FP64 oracle propagator on the CPU."""

    def __init__(self, og, target):
        self.og, self.t = og, target

    def execute(self, pending):
        out = {}
        for s, op in pending.items():
            if op.kind == "fev":
                seis, _, gr = fd2d.vel_to_seis(op.vec.numpy(), self.og, vec_adjoint=self.t,
                                               adjoint_on_residual=True)
                out[s] = (float(np.sum((seis - self.t) ** 2)), torch.tensor(gr))
            elif op.kind == "dir":
                out[s] = host_lbfgs.direction(op.step)
            elif op.kind == "gnp":
                _, jdv, _ = fd2d.vel_to_seis(op.vec.numpy(), self.og, vec_diff=op.dvec.numpy())
                out[s] = torch.tensor(fd2d.vel_to_seis(op.vec.numpy(), self.og, vec_adjoint=jdv)[2])
            else:
                out[s] = [float(np.sum((fd2d.vel_to_seis(v.numpy(), self.og)[0] - self.t) ** 2))
                          for v in op.vecs]
        return out


def _setup(seed=0):
    g, og = Geometry(**KW), fd2d.Geometry(**KW)
    true = synth.make_model("flat", g.nz, g.nx, seed)
    target = fd2d.vel_to_seis(synth.to_vec(true), og)[0]
    start = synth.to_vec(synth.depth_linear_start(g.nz, g.nx, true.min(), true.max()), min_vel=true.min())
    return g, og, target, start


@pytest.mark.parametrize("hist", [None, 3])
def test_lbfgs_machine_equals_oracle(hist):
    g, og, target, start = _setup(1)
    ref, nfev = oinv.lbfgs_invert(start, target, opri.RowTotalVariation(g.nz, g.nx), og, max_iter=6,
                                  tolerance_change=1e-9, history_size=hist)
    r = invert.Request(geom=g, target=torch.tensor(target), prior=priors.RowTotalVariation(g.nz, g.nx, device="cpu"),
                       start=torch.tensor(start))
    m = invert.lbfgs_machine(r, max_iter=6, tolerance_change=1e-9, history_size=hist)
    out = invert.run_machines({0: m}, OracleExecutor(og, target))[0].numpy()
    assert r.meta["stats"]["fev"] == nfev
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-7)


def test_gn_machine_equals_oracle():
    g, og, target, start = _setup(2)
    op = opri.L2Smoothness(g.nz, g.nx, lam=1e-9)
    ref, _, _ = oinv.gauss_newton_cg(start, target, op, og, n_cg=4, n_scales=7)
    r = invert.Request(geom=g, target=torch.tensor(target),
                       prior=priors.L2Smoothness(g.nz, g.nx, lam=1e-9, device="cpu"), start=torch.tensor(start))
    m = invert.gn_machine(r, n_outer=1, n_cg=4, n_scales=7)
    out = invert.run_machines({0: m}, OracleExecutor(og, target))[0].numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-7)
    st = r.meta["stats"]
    assert (st["fev"], st["gnp"], st["fwd"]) == (1, 4, 7) and st["misfit_hist"] == [st["misfit"]]


def test_priors_match_oracle():
    rng = np.random.default_rng(0)
    for P, O in [(priors.RowTotalVariation, opri.RowTotalVariation), (priors.TotalVariation, opri.TotalVariation),
                 (priors.L2Smoothness, opri.L2Smoothness)]:
        p, o = P(7, 9, lam=3.0, device="cpu"), O(7, 9, lam=3.0)
        x = rng.normal(0, 50, p.N)
        c, gr = p.cost_grad(torch.tensor(x))
        c2, g2 = o.cost_grad(x)
        assert float(c) == pytest.approx(c2, rel=1e-12)
        np.testing.assert_allclose(gr.numpy(), g2, rtol=1e-12, atol=1e-15)
    se = priors.SquaredExponential(6, 8, lam=2.0, device="cpu")
    x = torch.tensor(rng.normal(0, 50, 6 * 8 + 1))
    d = 1e-6 * torch.tensor(rng.normal(0, 1, 6 * 8 + 1))
    c, gr = se.cost_grad(x)
    c2, _ = se.cost_grad(x + d)
    assert abs(float((gr * d).sum()) - float(c2 - c)) / abs(float(c2 - c)) < 1e-5


def test_l2_apply_P_matches_dense():
    p = priors.L2Smoothness(5, 6, device="cpu")
    o = opri.L2Smoothness(5, 6)
    x = np.random.default_rng(1).normal(size=30)
    np.testing.assert_allclose(p.apply_P(torch.tensor(x)).numpy(), o.P @ x, rtol=1e-12)


def test_prior_rows_match_single_evaluations():
    """The line search's batched prior costs and basis expansion equal the
    one-vector evaluations row by row."""
    rng = np.random.default_rng(2)
    for p in (priors.RowTotalVariation(7, 9, lam=3.0, device="cpu"), priors.TotalVariation(7, 9, lam=3.0, device="cpu"),
              priors.L2Smoothness(7, 9, lam=3.0, device="cpu"), priors.SquaredExponential(6, 8, lam=2.0, device="cpu")):
        X = torch.tensor(rng.normal(0, 50, (5, p.N)))
        rows = p.cost_rows(X)
        one = torch.stack([torch.as_tensor(p.cost_grad(x)[0], dtype=X.dtype) for x in X])
        np.testing.assert_allclose(rows.numpy(), one.numpy(), rtol=1e-12)
        V = p.from_basis_rows(X)
        for k in range(5):
            np.testing.assert_allclose(V[k].numpy(), p.from_basis(X[k]).numpy(), rtol=1e-12, atol=1e-12)


def _flat_prior_cpu(o):
    """RestrictFlatAreas on the CPU with the oracle's labels (adapt labels on
    the GPU; the basis algebra is host torch)."""
    p = priors.RestrictFlatAreas(priors.L2Smoothness(o.nz, o.nx, lam=o.lam, device="cpu"), o.diff_threshold)
    lab, count = o._labels
    p.labels = torch.tensor(lab.ravel())
    p.count = count
    p.sizes = torch.bincount(p.labels, minlength=count).to(torch.float64)
    return p


def test_gn_machine_restrict_flat_areas_equals_oracle():
    """The GN machine on the RestrictFlatAreas basis (labels, sums and
    scatters instead of the dense matrix) reproduces the oracle's GN-CG on
    the dense basis of _adapt, step for step."""
    g, og = Geometry(**KW), fd2d.Geometry(**KW)
    true = synth.make_model("flat", g.nz, g.nx, 4)
    true[3:6, 4:8] = 3000.0                       # an area that is not a whole row
    target = fd2d.vel_to_seis(synth.to_vec(true), og)[0]
    guess = synth.to_vec(true * 0.97, min_vel=1500.0)
    o = opri.RestrictFlatAreas(opri.L2Smoothness(g.nz, g.nx, lam=1e-9), diff_threshold=1.0)
    o._labels = o.adapt(guess)
    ref, _, _ = oinv.gauss_newton_cg(guess, target, o, og, n_cg=4, n_scales=7)
    r = invert.Request(geom=g, target=torch.tensor(target), prior=_flat_prior_cpu(o), start=torch.tensor(guess))
    m = invert.gn_machine(r, n_outer=1, n_cg=4, n_scales=7)
    out = invert.run_machines({0: m}, OracleExecutor(og, target))[0].numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-7)


def test_chain_machine_equals_oracle_chain():
    """Stages chained as kgs.ChainedModel (LBFGS with 2-D TV, LBFGS with row
    TV, one Gauss-Newton pass on the SE prior's SVD modes with as many CG
    steps as modes) reproduce the oracle's chain, step for step."""
    from seismic_amd import schedule
    g, og, target, start = _setup(5)
    stages = [schedule.Stage(lambda nz, nx, d: priors.TotalVariation(nz, nx, lam=1e-8, device="cpu"), maxiter=4,
                             tolerance_grad=1e-7),
              schedule.Stage(lambda nz, nx, d: priors.RowTotalVariation(nz, nx, lam=1e-8, device="cpu"), maxiter=3,
                             tolerance_change=1e-9),
              schedule.Stage(lambda nz, nx, d: priors.SquaredExponential(nz, nx, lam=1e-12, transform=True,
                                                                         svd_cutoff=1e4, device="cpu"),
                             maxiter=0, n_scales=7)]
    ostages = [dict(prior=lambda: opri.TotalVariation(g.nz, g.nx, lam=1e-8), maxiter=4, tolerance_grad=1e-7),
               dict(prior=lambda: opri.RowTotalVariation(g.nz, g.nx, lam=1e-8), maxiter=3, tolerance_change=1e-9),
               dict(prior=lambda: opri.SquaredExponential(g.nz, g.nx, lam=1e-12, transform=True, svd_cutoff=1e4),
                    maxiter=0, n_scales=7)]
    ref = oinv.chain_invert(start, target, ostages, og)
    r = invert.Request(geom=g, target=torch.tensor(target), prior=None, start=torch.tensor(start))
    m = schedule.chain_machine(r, schedule.Schedule(stages))
    out = invert.run_machines({0: m}, OracleExecutor(og, target))[0].numpy()
    np.testing.assert_allclose(out, ref, rtol=1e-7, atol=1e-6)
    assert len(r.meta["stages"]) == 3


def test_schedule_kpis_match_oracle():
    from seismic_amd import schedule
    rng = np.random.default_rng(8)
    s = rng.normal(0, 1, (5, 40, 12))
    s[4] = s[0][:, ::-1] + 1e-6 * rng.normal(0, 1, (40, 12))
    assert schedule.kpi_flat_vel(torch.tensor(s.ravel()), 5, 40, 12) == pytest.approx(
        oinv.kpi_flat_vel(s.ravel(), 5, 40, 12), rel=1e-12)
    v = np.round(rng.normal(3000, 20, (13, 17)))
    assert schedule.kpi_style_b(torch.tensor(v)) == oinv.kpi_style_b(v)
    se = priors.SquaredExponential(6, 8, lam=1e-12, device="cpu")
    o = opri.SquaredExponential(6, 8, lam=1e-12)
    x = rng.normal(2500, 300, 6 * 8 + 1)
    # K^-1 of the SE kernel is ill-conditioned (sigma_mean^2 offset): agreement to 1e-7
    assert schedule.kpi_style_a(torch.tensor(x), se.P) == pytest.approx(float(x[:-1] @ o.P @ x[:-1]), rel=1e-7)


def test_reference_model_classes_map_to_stages():
    """InversionModel / ChainedModel with the reference's attribute names
    (core/seis_invert.py:52-58, core/kaggle_support.py:384, 512-523) run the
    same chain as the explicit stages; per-model rounding rounds that
    model's output; a prior for another grid is refused."""
    from seismic_amd import schedule
    g, og, target, start = _setup(6)
    tv = priors.TotalVariation(g.nz, g.nx, lam=1e-8, device="cpu")
    rtv = priors.RowTotalVariation(g.nz, g.nx, lam=1e-8, device="cpu")
    chain = schedule.ChainedModel([schedule.InversionModel(tv, maxiter=3, lbfgs_tolerance_grad=1e-7),
                                   schedule.InversionModel(rtv, maxiter=2, lbfgs_tolerance_change=1e-9,
                                                           round_results=True)])
    sts = chain.stages()
    assert [st.maxiter for st in sts] == [3, 2] and [st.round_results for st in sts] == [False, True]
    explicit = [schedule.Stage(lambda nz, nx, d: priors.TotalVariation(nz, nx, lam=1e-8, device="cpu"), maxiter=3),
                schedule.Stage(lambda nz, nx, d: priors.RowTotalVariation(nz, nx, lam=1e-8, device="cpu"), maxiter=2,
                               tolerance_change=1e-9, round_results=True)]
    outs = []
    for stages in (sts, explicit):
        r = invert.Request(geom=g, target=torch.tensor(target), prior=None, start=torch.tensor(start))
        m = schedule.chain_machine(r, schedule.Schedule(stages))
        outs.append(invert.run_machines({0: m}, OracleExecutor(og, target))[0])
    assert torch.equal(outs[0], outs[1])
    assert torch.equal(outs[0], torch.round(outs[0]))
    wrong = schedule.InversionModel(priors.TotalVariation(g.nz + 1, g.nx, device="cpu"), maxiter=1)
    with pytest.raises(ValueError):
        wrong.stages()[0].prior(g.nz, g.nx, "cpu")
    with pytest.raises(ValueError):
        schedule.InversionModel().stages()


class LinearExecutor:
    """Machine ops on a linear model s = J vec (exact Gauss-Newton model)."""

    def __init__(self, J, t):
        self.J, self.t = J, t

    def execute(self, pending):
        out = {}
        J, t = self.J, self.t
        for s, op in pending.items():
            if op.kind == "fev":
                r = J @ op.vec.numpy() - t
                out[s] = (float(r @ r), torch.tensor(J.T @ r))
            elif op.kind == "gnp":
                out[s] = torch.tensor(J.T @ (J @ op.dvec.numpy()))
            elif op.kind == "dir":
                out[s] = host_lbfgs.direction(op.step)
            else:
                out[s] = [float(np.sum((J @ v.numpy() - t) ** 2)) for v in op.vecs]
        return out


@pytest.mark.parametrize("n_cg", [None, 30])
def test_gn_machine_reaches_closed_form_minimiser(n_cg):
    """The product's GN pass (dense solve, or CG with more steps than
    parameters) on a linear model lands on the minimiser of
    lam x^T P x + |J x - t|^2 / N, computed by least squares on the stacked
    system; pins the machine's rhs and system scaling (core/seis_invert.py:77-92)."""
    g = Geometry(nz=4, nx=5, nbc=2, nt=150)
    rng = np.random.default_rng(6)
    n, N, lam = g.n_params, g.n_data, 0.37
    J = rng.normal(size=(N, n))
    t = rng.normal(size=N) * 10
    D = np.hstack([opri.L2Smoothness(g.nz, g.nx).D, np.zeros((opri.L2Smoothness(g.nz, g.nx).D.shape[0], 1))])
    M = np.vstack([np.sqrt(lam) * D, J / np.sqrt(N)])
    ref = np.linalg.lstsq(M, np.concatenate([np.zeros(D.shape[0]), t / np.sqrt(N)]), rcond=None)[0]
    r = invert.Request(geom=g, target=torch.tensor(t), prior=priors.L2Smoothness(g.nz, g.nx, lam=lam, device="cpu"),
                       start=torch.tensor(rng.normal(size=n)))
    out = invert.run_machines({0: invert.gn_machine(r, n_outer=1, n_cg=n_cg, n_scales=13)}, LinearExecutor(J, t))[0]
    np.testing.assert_allclose(out.numpy(), ref, rtol=1e-7, atol=1e-7 * np.abs(ref).max())


@pytest.mark.parametrize("scaling,tol", [(1e15, 0.0), (1.0, 0.0), (1e15, 0.02)])
def test_lbfgs_closure_matches_oracle_on_linear_model(scaling, tol):
    """The product's LBFGS closure (scaling, seis_error_tolerance zeroing,
    core/seis_invert.py:125-138) against the oracle's (pinned by
    test_lbfgs_closure_scaling_and_error_tolerance) on a linear model: same
    iterate, same number of evaluations, also when the misfit tolerance
    stops the run early."""
    from test_oracle import LinearV2S
    g = Geometry(nz=6, nx=5, nbc=2, nt=150)
    og = fd2d.Geometry(nz=6, nx=5, nbc=2, nt=150)
    rng = np.random.default_rng(9)
    J = rng.normal(size=(g.n_data, g.n_params)) / 50
    op = opri.RowTotalVariation(g.nz, g.nx, lam=1e-4)
    true = op.basis @ np.concatenate([np.linspace(1800, 2600, g.nz), [1700.0]])
    t = J @ true
    start = op.basis @ np.concatenate([np.full(g.nz, 2200.0), [1700.0]])
    ref, nfev = oinv.lbfgs_invert(start, t, op, og, max_iter=12, tolerance_change=1e-9, scaling=scaling,
                                  seis_error_tolerance=tol, v2s=LinearV2S(J))
    r = invert.Request(geom=g, target=torch.tensor(t), prior=priors.RowTotalVariation(g.nz, g.nx, lam=1e-4, device="cpu"),
                       start=torch.tensor(start))
    m = invert.lbfgs_machine(r, max_iter=12, tolerance_change=1e-9, scaling=scaling, seis_error_tolerance=tol)
    out = invert.run_machines({0: m}, LinearExecutor(J, t))[0].numpy()
    assert r.meta["stats"]["fev"] == nfev
    if tol > 0:
        assert nfev < 12 and r.meta["stats"]["misfit"] < tol    # stopped by the misfit tolerance
    np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-7)


def test_reference_prior_interface():
    """The reference's Prior surface (core/seis_prior.py:17-65): λ,
    basis_vectors, prep / adapt, compute_cost_and_gradient(x (N, 1),
    compute_gradient), on the product's priors; the same values as their
    own cost_grad and basis."""
    rng = np.random.default_rng(4)
    for p in (priors.RowTotalVariation(7, 9, device="cpu"), priors.TotalVariation(7, 9, device="cpu"),
              priors.L2Smoothness(7, 9, lam=2e-8, device="cpu")):
        p.prep()
        assert p.prepped
        p.λ = 3e-7
        assert p.lam == 3e-7 and p.λ == 3e-7
        B = p.basis_vectors
        assert tuple(B.shape) == (7 * 9 + 1, p.N) and B.dtype == torch.float64
        x = torch.tensor(rng.normal(2000, 50, (p.N, 1)))
        c = p.compute_cost_and_gradient(x)
        c2, g2 = p.compute_cost_and_gradient(x, True)
        c3, g3 = p.cost_grad(x.reshape(-1))
        assert c.shape == () and float(c) == float(c2) == float(c3)
        assert tuple(g2.shape) == (p.N, 1) and torch.equal(g2.reshape(-1), g3)
        with pytest.raises(ValueError):
            p.compute_cost_and_gradient(x.reshape(-1))
        p.adapt(torch.zeros(7 * 9 + 1))          # no adaptive basis: unchanged
        assert torch.equal(p.basis_vectors, B)
