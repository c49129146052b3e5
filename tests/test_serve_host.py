"""Serving loop host logic on the CPU with a fake executor: admission by slot
and HBM footprint, continuous batching (late arrivals join running batches),
per-request results independent of the mix, NaN fallback to the start model
(core/seis_invert.py:146-149)."""
import time

import numpy as np
import pytest
import torch

import host_lbfgs
from seismic_amd import invert, serve
from seismic_amd.geometry import Geometry


class QuadPrior:
    """Zero prior on the identity basis (enough to drive the LBFGS machine)."""

    basis = None
    lam = 0.0
    device = torch.device("cpu")

    def __init__(self, n):
        self.n = n

    def to_basis(self, v):
        return v.clone()

    def from_basis(self, x):
        return x

    def basis_T(self, g):
        return g

    def cost_grad(self, x):
        return torch.tensor(0.0, dtype=torch.float64), torch.zeros_like(x)


class FakeExecutor:
    """'fev' of a separable quadratic: misfit = sum (x - c)^2 * N, J^T r = N (x - c)."""

    def __init__(self, nan_rids=(), seconds=0.01):
        self.centers = {}
        self.nan_targets = set(nan_rids)
        self.nan = set()
        self.calls = []
        self.seconds = seconds      # stands in for the batched launch time

    def set_target(self, slot, target):
        # a request's "observed data" is its quadratic's centre; the last entry tags NaN requests
        self.centers[slot] = target[:-1]
        self.nan.discard(slot)
        if int(target[-1]) in self.nan_targets:
            self.nan.add(slot)

    def execute(self, pending):
        self.calls.append(sorted(pending))
        time.sleep(self.seconds)
        out = {}
        for s, op in pending.items():
            if op.kind == "dir":
                out[s] = host_lbfgs.direction(op.step)
                continue
            assert op.kind == "fev"
            c = self.centers[s]
            r = op.vec - c
            N = float(c.numel())
            f = float((r * r).sum()) * N
            g = r * N
            if s in self.nan:
                f, g = float("nan"), g * float("nan")
            out[s] = (f, g)
        return out


def _server(slots, budget, fake, footprint=100):
    srv = serve.InversionServer.__new__(serve.InversionServer)
    srv.slots = list(slots)
    srv.ckpt_every = 0
    srv.kw = dict(max_iter=5, history_size=3, tolerance_grad=1e-12, tolerance_change=0.0)
    srv.machine = lambda req: invert.lbfgs_machine(req, **srv.kw)     # the default machine
    srv.budget = budget
    srv.ex = fake
    srv._fp = {g: footprint for g in set(slots)}
    srv.used = 0
    srv.free = {}
    for i, g in enumerate(slots):
        srv.free.setdefault(g, []).append(i)
    srv.log = []
    return srv


def _req(g, rid, n=7, center=None):
    c = torch.ones(n, dtype=torch.float64) if center is None else center
    return invert.Request(geom=g, target=torch.cat([c, torch.tensor([float(rid)], dtype=torch.float64)]),
                          prior=QuadPrior(n), start=torch.full((n,), 3.0, dtype=torch.float64), rid=rid)


def test_admission_respects_slots_and_budget():
    g1, g2 = Geometry(nz=4, nx=4, nbc=2, nt=150), Geometry(nz=5, nx=5, nbc=2, nt=150)
    srv = _server([g1, g1, g2], budget=250, fake=None)
    a, b, c, d = _req(g1, 0), _req(g1, 1), _req(g2, 2), _req(g1, 3)
    assert srv._try_admit(a) is not None
    assert srv._try_admit(b) is not None
    assert srv._try_admit(d) is None          # no free 4x4 slot
    assert srv._try_admit(c) is None          # 5x5 slot free but the budget is used up
    srv._release(0, a)
    assert srv._try_admit(c) is not None


def test_continuous_batching_and_results_match_alone():
    g = Geometry(nz=4, nx=4, nbc=2, nt=150)
    n = 7
    centers = {i: torch.linspace(0, 1, n, dtype=torch.float64) * (i + 1) for i in range(4)}
    fake = FakeExecutor()
    srv = _server([g] * 4, budget=10 ** 9, fake=fake)
    reqs = [_req(g, i, n, centers[i]) for i in range(4)]
    arrivals = [0.0, 0.0, 0.015, 0.025]
    recs = srv.run(reqs, arrivals)
    assert all(not r.failed for r in recs)
    assert any(len(c) >= 3 for c in fake.calls)   # late arrivals merged into running batches
    for i, r in enumerate(recs):
        s1 = _server([g], budget=10 ** 9, fake=FakeExecutor())
        r1 = s1.run([_req(g, i, n, centers[i])], [0.0])[0]
        assert torch.equal(r.result, r1.result)
        assert r.latency >= 0 and r.finished >= r.admitted >= r.arrival - 1e-9


def test_first_fit_admission_does_not_block_other_sizes():
    """A request waiting for a busy 4x4 slot does not hold back a 5x5 request
    behind it in the queue."""
    g1, g2 = Geometry(nz=4, nx=4, nbc=2, nt=150), Geometry(nz=5, nx=5, nbc=2, nt=150)
    srv = _server([g1, g2], budget=10 ** 9, fake=FakeExecutor())
    reqs = [_req(g1, 0), _req(g1, 1), _req(g2, 2, n=7)]
    recs = srv.run(reqs, [0.0, 0.0, 0.0])
    assert all(not r.failed for r in recs)
    assert recs[2].admitted < recs[0].finished      # admitted while request 0 held the 4x4 slot
    assert recs[1].admitted >= recs[0].finished     # request 1 waited for that slot


def test_nan_request_falls_back_to_start():
    g = Geometry(nz=4, nx=4, nbc=2, nt=150)
    fake = FakeExecutor(nan_rids=[1])
    srv = _server([g, g], budget=10 ** 9, fake=fake)
    reqs = [_req(g, 0), _req(g, 1)]
    recs = srv.run(reqs, [0.0, 0.0])
    assert torch.equal(recs[1].result, reqs[1].start)
    assert not torch.isnan(recs[0].result).any()


def test_run_stream_refills_slots_and_matches_alone():
    """run_stream pulls the next request whenever a slot frees: 6 requests on
    2 slots all finish in pull order, at most 2 run at once, every slot and
    byte is released at the end, and each result equals that request run
    alone; a request with no slot of its size, or over the budget of an
    idle server, is refused."""
    g = Geometry(nz=4, nx=4, nbc=2, nt=150)
    n = 7
    centers = {i: torch.linspace(0, 1, n, dtype=torch.float64) * (i + 2) for i in range(6)}
    fake = FakeExecutor()
    srv = _server([g, g], budget=10 ** 9, fake=fake)
    pulled = []

    def source():
        for i in range(6):
            pulled.append(i)
            yield _req(g, i, n, centers[i])
    recs = srv.run_stream(source())
    assert [r.rid for r in recs] == list(range(6)) and pulled == list(range(6))
    assert all(not r.failed for r in recs) and max(len(c) for c in fake.calls) <= 2
    assert srv.used == 0 and sorted(srv.free[g]) == [0, 1]
    for r in recs:
        r1 = _server([g], budget=10 ** 9, fake=FakeExecutor()).run([_req(g, r.rid, n, centers[r.rid])], [0.0])[0]
        assert torch.equal(r.result, r1.result)
    other = Geometry(nz=5, nx=5, nbc=2, nt=150)
    with pytest.raises(ValueError):
        _server([g], budget=10 ** 9, fake=FakeExecutor()).run_stream([_req(other, 0)])
    with pytest.raises(ValueError):
        _server([g], budget=10, fake=FakeExecutor()).run_stream([_req(g, 0)])


def test_run_refuses_requests_it_can_never_admit():
    """run() raises instead of waiting forever for a request with no slot of
    its size, or with a footprint above the whole budget; requests before
    it are still served (they arrive earlier)."""
    g, other = Geometry(nz=4, nx=4, nbc=2, nt=150), Geometry(nz=5, nx=5, nbc=2, nt=150)
    with pytest.raises(ValueError, match="no slot"):
        _server([g], budget=10 ** 9, fake=FakeExecutor()).run([_req(g, 0), _req(other, 1)], [0.0, 0.01])
    with pytest.raises(ValueError, match="budget"):
        _server([g], budget=10, fake=FakeExecutor()).run([_req(g, 0)], [0.0])
    srv = _server([g, g], budget=250, fake=FakeExecutor())     # two slots, room for two footprints
    recs = srv.run([_req(g, i) for i in range(3)], [0.0, 0.0, 0.0])
    assert all(not r.failed for r in recs) and srv.used == 0


def test_poisson_arrivals_and_percentiles():
    a = serve.poisson_arrivals(2000, 100.0, seed=1)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 100.0) < 10.0
    p = serve.percentiles([1.0, 2.0, 3.0, 4.0])
    assert p["p50"] == pytest.approx(2.5) and p["max"] == 4.0


def test_server_reserves_its_plans_against_the_budget():
    """The plans (all slots' wavefields, and the line-search fan-out plan when
    the machines run line searches) are reserved at construction; requests
    are admitted against what is left, by their own state."""
    from seismic_amd._lib import CAP_ADJOINT, CAP_FORWARD, CAP_TANGENT
    from seismic_amd.propagator import plan_bytes
    g = Geometry(nz=8, nx=9, nbc=4, nt=150)
    caps = CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT
    plan = plan_bytes([g, g], ckpt_every=0, caps=caps)
    fp = serve.request_footprint(g, 10)
    srv = serve.InversionServer([g, g], hbm_budget=plan + 2 * fp, ckpt_every=0, device="cpu")
    assert srv.used == plan and srv.reserved == plan
    a, b = _req(g, 0), _req(g, 1)
    assert srv._try_admit(a) is not None and srv._try_admit(b) is not None
    assert srv.used == plan + 2 * fp
    tight = serve.InversionServer([g, g], hbm_budget=plan + fp, ckpt_every=0, device="cpu")
    assert tight._try_admit(a) is not None and tight._try_admit(b) is None     # one request's state fits
    fan = invert.fanout_plan_bytes([g, g])
    gn = serve.InversionServer([g, g], hbm_budget=plan + fan + fp, ckpt_every=0, device="cpu",
                               machine=lambda r: invert.gn_machine(r))
    assert gn.reserved == plan + fan
    with pytest.raises(ValueError):
        serve.InversionServer([g, g], hbm_budget=plan - 1, ckpt_every=0, device="cpu")
