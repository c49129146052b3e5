"""The continuous-batching server on the GPU: served LBFGS requests of two
grid sizes, arriving as a Poisson process on fewer slots than requests
(checkpointed plans), give each request's result bit for bit as it comes out
run alone on the plain path (full-history plan, one model per launch); the
streaming mode gives the same results."""
import numpy as np
import pytest
import torch

from oracle import fd2d
from seismic_amd import invert, priors, serve, synth
from seismic_amd.geometry import Geometry

pytestmark = pytest.mark.gpu


def _requests():
    geoms = [Geometry(nz=20, nx=24, nbc=16, nt=300), Geometry(nz=16, nx=30, nbc=16, nt=300)]
    reqs = []
    for i in range(6):
        g = geoms[i % 2]
        og = fd2d.Geometry(nz=g.nz, nx=g.nx, nbc=g.nbc, nt=g.nt)
        true = synth.curve_fault_model(g.nz, g.nx, 80 + i)
        target = torch.tensor(fd2d.vel_to_seis(synth.to_vec(true), og)[0], dtype=torch.float32, device="cuda")
        start = synth.to_vec(synth.depth_linear_start(g.nz, g.nx, 1500.0, 4000.0), min_vel=1500.0)
        reqs.append(invert.Request(geom=g, target=target, prior=priors.L2Smoothness(g.nz, g.nx, lam=1e-8),
                                   start=torch.tensor(start, dtype=torch.float64, device="cuda"), rid=i))
    return geoms, reqs


def _clone(reqs):
    return [invert.Request(geom=r.geom, target=r.target, prior=r.prior, start=r.start.clone(), rid=r.rid)
            for r in reqs]


def test_served_lbfgs_equals_alone():
    geoms, reqs = _requests()
    kw = dict(max_iter=8, history_size=5, tolerance_grad=1e-7, tolerance_change=0.0)
    alone = invert.invert(_clone(reqs), method="lbfgs", batched=False, **kw)
    srv = serve.InversionServer([geoms[0], geoms[1], geoms[0], geoms[1]], hbm_budget=1 << 40, ckpt_every=50,
                                device="cuda", **kw)
    recs = srv.run(_clone(reqs), serve.poisson_arrivals(6, 300.0, seed=2))
    assert all(not r.failed for r in recs)
    assert max(n for _, n in srv.log) > 1                      # requests shared launches
    for rec, a in zip(recs, alone):
        assert torch.equal(rec.result, a), rec.rid
    srv2 = serve.InversionServer([geoms[0], geoms[1], geoms[0], geoms[1]], hbm_budget=1 << 40, ckpt_every=50,
                                 device="cuda", **kw)
    recs2 = srv2.run_stream(iter(_clone(reqs)))
    for rec in recs2:
        assert torch.equal(rec.result, alone[rec.rid]), rec.rid


def test_server_keeps_full_history_when_it_fits():
    """ckpt_every=None keeps every slot's full forward history when the
    plan fits the budget (one byte less and it checkpoints every 50 steps; a
    budget below the checkpointed plan is refused); the plan is reserved
    against the budget when the server is built; results are those of the
    plain run either way."""
    from seismic_amd._lib import CAP_ADJOINT, CAP_FORWARD, CAP_TANGENT
    from seismic_amd.propagator import plan_bytes
    geoms, reqs = _requests()
    slots = [geoms[0], geoms[1], geoms[0], geoms[1]]
    caps = CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT
    full = plan_bytes(slots, ckpt_every=0, caps=caps)
    kw = dict(max_iter=8, history_size=5, tolerance_grad=1e-7, tolerance_change=0.0)
    small = serve.InversionServer(slots, hbm_budget=full - 1, ckpt_every=None, device="cuda", **kw)
    assert small.ckpt_every == 50 and small.used == plan_bytes(slots, ckpt_every=50, caps=caps)
    small.ex.close()
    with pytest.raises(ValueError):        # not even the checkpointed plan fits
        serve.InversionServer(slots, hbm_budget=plan_bytes(slots, ckpt_every=50, caps=caps) - 1, ckpt_every=None,
                              device="cuda", **kw)
    srv = serve.InversionServer(slots, hbm_budget=1 << 40, ckpt_every=None, device="cuda", **kw)
    assert srv.ckpt_every == 0
    alone = invert.invert(_clone(reqs), method="lbfgs", batched=False, **kw)
    recs = srv.run(_clone(reqs), serve.poisson_arrivals(6, 300.0, seed=4))
    for rec, a in zip(recs, alone):
        assert torch.equal(rec.result, a), rec.rid


def test_served_lbfgs_at_config5_sizes():
    """The serving workload at its sizes: LBFGS requests (30 iterations,
    memory 10, L2 smoothness prior) on 70x70, 100x100 and 140x140 grids
    (nbc 120, nt 1000, 5 shots, a receiver on every surface column),
    curved-fault models, 6 requests arriving as a Poisson process on 3 slots
    (one per size, full forward histories): every served result equals the
    same request run alone on the plain path, bit for bit."""
    from seismic_amd.propagator import Propagator
    geoms = [Geometry(nz=n, nx=n, nt=1000) for n in (70, 100, 140)]
    reqs = []
    for i in range(6):
        g = geoms[i % 3]
        P = Propagator([g], caps=1)
        target = P.forward(torch.tensor(synth.to_vec(synth.curve_fault_model(g.nz, g.nx, 200 + i)))).clone()
        P.close()
        start = synth.to_vec(synth.depth_linear_start(g.nz, g.nx, 1500.0, 4000.0), min_vel=1500.0)
        reqs.append(invert.Request(geom=g, target=target, prior=priors.L2Smoothness(g.nz, g.nx, lam=1e-8),
                                   start=torch.tensor(start, dtype=torch.float64, device="cuda"), rid=i))
    kw = dict(max_iter=30, history_size=10, tolerance_grad=1e-7, tolerance_change=0.0)
    alone = invert.invert(_clone(reqs), method="lbfgs", batched=False, **kw)
    srv = serve.InversionServer(geoms, hbm_budget=1 << 40, ckpt_every=None, device="cuda", **kw)
    assert srv.ckpt_every == 0
    recs = srv.run(_clone(reqs), serve.poisson_arrivals(6, 100.0, seed=3))
    print("served at config-5 sizes: in flight", max(n for _, n in srv.log),
          "latency p50/p99", serve.percentiles([r.latency for r in recs]), flush=True)
    assert all(not r.failed for r in recs)
    assert max(n for _, n in srv.log) == 3                     # all three slots shared launches
    for rec, a in zip(recs, alone):
        assert rec.stats["n_iter"] >= 10, rec.stats
        assert torch.equal(rec.result, a), rec.rid
    srv.ex.close()
