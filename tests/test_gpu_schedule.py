"""Chained schedules and family dispatch (core/seis_model.py) on the GPU:
the flat-patch KPI on the labelling kernel against the oracle, dispatch of
a laterally symmetric and a faulted 70x70 case, and batched chains of
different families equal to the same chains run one at a time.  The chain
logic itself is pinned step for step against the oracle's chain on the CPU
(tests/test_invert_host.py)."""
import numpy as np
import pytest
import torch

from oracle import fd2d
from oracle import invert as oinv
from seismic_amd import invert, schedule, synth
from seismic_amd.geometry import Geometry

pytestmark = pytest.mark.gpu


def test_kpi_fault_a_matches_oracle():
    rng = np.random.default_rng(4)
    grids = [synth.curve_fault_model(70, 70, 3), np.round(rng.normal(3000, 8, (30, 40))),
             np.full((12, 12), 2500.0), synth.make_model("flat", 70, 70, 2)]
    for v in grids:
        assert schedule.kpi_fault_a(torch.tensor(v, device="cuda")) == oinv.kpi_fault_a(v)


def _request(true, start, g, og):
    target = fd2d.vel_to_seis(synth.to_vec(true), og)[0]
    return invert.Request(geom=g, target=torch.tensor(target, dtype=torch.float32, device="cuda"), prior=None,
                          start=torch.tensor(start, dtype=torch.float64, device="cuda"))


def test_dispatch_symmetric_and_faulted():
    """The reference's KPIs pick FlatVel for a laterally invariant model
    (shot 0 mirrors shot 4) and a TV2D family for a faulted one."""
    kw = dict(nt=400)
    g, og = Geometry(**kw), fd2d.Geometry(**kw)
    flat = synth.make_model("flat", 70, 70, 5)
    fault = synth.curve_fault_model(70, 70, 6)
    start = synth.to_vec(synth.depth_linear_start(70, 70, 1500, 4000), min_vel=1500.0)
    P = schedule.style_a_precision(70, 70, "cuda")
    assert schedule.classify(_request(flat, start, g, og), P) == "FlatVel"
    fam = schedule.classify(_request(fault, synth.to_vec(fault), g, og), P)
    print("faulted family", fam, flush=True)
    assert fam in ("TV2D", "TV2Deasy")


def test_batched_chains_equal_plain():
    """TV2D (three stages, the last on RestrictFlatAreas), FlatVel, Style_A
    (SE modes, then a dense Gauss-Newton pass), Style_B (two SE stages, the
    second stopping on the misfit) and TV2Deasy chains in test mode, batched
    together and run one at a time: identical results, lower misfit than the
    start (for Style_A, after its LBFGS stage)."""
    kw = dict(nz=20, nx=24, nbc=16, nt=300)
    g, og = Geometry(**kw), fd2d.Geometry(**kw)
    reqs = []
    fams = ["TV2D", "FlatVel", "Style_A", "TV2D", "Style_B", "TV2Deasy"]
    for i, fam in enumerate(fams):
        # each family on a model of its kind: layered for FlatVel, smooth for the SE
        # families, faulted for the TV ones
        true = {"FlatVel": lambda: synth.make_model("flat", g.nz, g.nx, 20 + i),
                "Style_A": lambda: synth.make_model("grf", g.nz, g.nx, 20 + i),
                "Style_B": lambda: synth.make_model("grf", g.nz, g.nx, 20 + i)}.get(
                    fam, lambda: synth.curve_fault_model(g.nz, g.nx, 20 + i))()
        start = synth.to_vec(np.round(true * 0.96), min_vel=1500.0)
        reqs.append(_request(true, start, g, og))
    clone = lambda: [invert.Request(geom=r.geom, target=r.target, prior=None, start=r.start.clone()) for r in reqs]
    rb = clone()
    vb, _ = schedule.infer(rb, test_mode=True, batched=True, families=fams)
    vp, _ = schedule.infer(clone(), test_mode=True, batched=False, families=fams)
    for a, b, r, q, fam in zip(vb, vp, reqs, rb, fams):
        assert torch.equal(a, b), fam
        m = lambda v: np.mean((fd2d.vel_to_seis(v.cpu().numpy(), og)[0] - r.target.cpu().numpy()) ** 2)
        m0 = m(r.start)
        if fam == "Style_A":
            # The test-mode Gauss-Newton stage keeps only the SE modes above 1e4 and, as
            # core/seis_invert.py:71,104 does, adds its step to the unprojected model; on a
            # 20x24 grid that leaves an unphysical model.  The LBFGS stage on the modes must
            # still lower the misfit, and the GN line-search minimum must be finite.
            s1, s2 = q.meta["stages"]
            print(fam, "misfit", m0, "-> stage 1", s1["misfit"], "GN line search", s2["misfit"], flush=True)
            assert s1["misfit"] < m0 and np.isfinite(s2["misfit"])
            continue
        m1 = m(a)
        print(fam, "misfit", m0, "->", m1, flush=True)
        assert np.isfinite(m1) and m1 < m0


def test_reference_model_classes_on_gpu():
    """An InversionModel set up as the reference's FlatVel model
    (core/seis_model.py:15-27, test-mode length) and wrapped in a rounding
    ChainedModel gives the FlatVel chain of schedule.infer bit for bit."""
    kw = dict(nz=20, nx=24, nbc=16, nt=300)
    g, og = Geometry(**kw), fd2d.Geometry(**kw)
    reqs = []
    for i in range(3):
        true = synth.make_model("flat", g.nz, g.nx, 60 + i)
        reqs.append(_request(true, synth.to_vec(np.round(true * 0.96), min_vel=1500.0), g, og))
    clone = lambda: [invert.Request(geom=r.geom, target=r.target, prior=None, start=r.start.clone()) for r in reqs]
    from seismic_amd import priors
    m = schedule.InversionModel(priors.RowTotalVariation(g.nz, g.nx, lam=1e-8, epsilon=0.1, device="cuda"),
                                maxiter=30, lbfgs_tolerance_change=1e-9, lbfgs_tolerance_grad=1e-7)
    out = schedule.ChainedModel([m], round_results=True).infer(clone())
    ref, _ = schedule.infer(clone(), test_mode=True, families=["FlatVel"] * 3)
    for a, b in zip(out, ref):
        assert torch.equal(a, b)


def test_served_chains_equal_batched():
    """Classified chains served with continuous batching (Poisson arrivals,
    4 slots for 6 requests, checkpointed plans) give the families and the
    results of schedule.infer on the same requests, bit for bit."""
    from seismic_amd import serve
    kw = dict(nz=20, nx=24, nbc=16, nt=300)
    g, og = Geometry(**kw), fd2d.Geometry(**kw)
    reqs = []
    for i in range(6):
        true = synth.make_model("flat", g.nz, g.nx, 70 + i) if i % 2 == 0 else synth.curve_fault_model(g.nz, g.nx, 70 + i)
        r = _request(true, synth.to_vec(np.round(true * 0.96), min_vel=1500.0), g, og)
        r.rid = i
        reqs.append(r)
    clone = lambda: [invert.Request(geom=r.geom, target=r.target, prior=None, start=r.start.clone(), rid=r.rid)
                     for r in reqs]
    ref, fams = schedule.infer(clone(), test_mode=True)
    srv = serve.InversionServer([g] * 4, hbm_budget=1 << 40, ckpt_every=50, history_size=100,
                                machine=schedule.serving_machine(test_mode=True))
    recs = srv.run(clone(), serve.poisson_arrivals(6, 200.0, seed=3))
    print("families", fams, [r.stats.get("family") for r in recs], flush=True)
    assert [r.stats["family"] for r in recs] == fams
    for rec, b in zip(recs, ref):
        assert not rec.failed and torch.equal(rec.result, b)
    assert max(n for _, n in srv.log) > 1          # requests shared launches


def test_flatvel_chain_matches_oracle_chain_at_full_size():
    """The classified chain at the reference size against the oracle's chain:
    a flat-layered 70x70 model (nbc 120, nt 999) is dispatched to FlatVel by
    the mirrored-shot KPI (core/seis_model.py:155-157) and runs the FlatVel
    model in test mode (:15-27: row TV, 30 LBFGS iterations, rounding); the
    oracle's chain is its LBFGS inversion of the same data from the same start
    (tests/golden/ref70_lbfgs.npz), rounded (core/kaggle_support.py:472-474)."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref70_lbfgs.npz"))
    g = Geometry()
    req = invert.Request(geom=g, target=torch.tensor(d["target"], device="cuda"), prior=None,
                         start=torch.tensor(d["start"], dtype=torch.float64, device="cuda"))
    vecs, fams = schedule.infer([req], test_mode=True, batched=True)
    assert fams == ["FlatVel"]
    v = vecs[0].cpu().numpy()
    ref = np.round(d["v"])
    mae = np.abs(v[:-1] - ref[:-1]).mean()
    same = float(np.mean(v[:-1] == ref[:-1]))
    print(f"FlatVel chain 70x70: MAE vs oracle chain {mae:.4f} m/s, identical cells {same:.4f}", flush=True)
    assert mae < 1.0 and same > 0.9 and abs(v[-1] - ref[-1]) <= 1.0
