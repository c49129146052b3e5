"""Full-size parity of the kernel configs the workloads actually run.

The batched and served plans take their launch configs from the tuned table
(``autotune.apply``): streamed forward and Born kernels over column strips
and row segments, 128x64 adjoint and record windows at 6 steps per launch,
whole-column strips for the 1600-model line search.  These tests build such
plans at the reference geometry (nbc 120, nt 999) with the golden models
inside them and check those models against the fixtures the FP64 oracle
wrote (tests/golden/make_golden.py): seismograms 1e-5, gradients 1e-4 (relative
L2), the dot-product test through the Born pass 1e-6, and batch invariance
(each model's results bit-identical alone and in the batch)."""
import os
import sys

import numpy as np
import pytest
import torch

from conftest import rel
from seismic_amd import autotune, synth
from seismic_amd._lib import CAP_FORWARD
from seismic_amd.geometry import Geometry
from seismic_amd.propagator import Propagator

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_golden  # noqa: E402

SEIS_TOL, GRAD_TOL, DOT_TOL = 1e-5, 1e-4, 1e-6


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _gold70():
    return np.load(os.path.join(GOLD, "ref70_fwd.npz")), np.load(os.path.join(GOLD, "ref70_grad.npz"))


def _vec70():
    return synth.to_vec(synth.curve_fault_model(70, 70, seed=7))


def _filler(g, i):
    return synth.to_vec(synth.make_model(("flat", "curved", "faulted", "grf")[i % 4], g.nz, g.nx, 100 + i))


def _configs(P):
    return {k: P.get_config(k) for k in ("forward", "record", "tangent", "adjoint", "born") if autotune._has(P, k)}


def test_tuned_batch32_plan_golden():
    """The 32-model Gauss-Newton plan of the benchmark (full history): the
    golden 70x70 model at slot 13 of 32.  Forward (streamed strips),
    gradient (128x96 record and 128x64 adjoint windows, k = 6), Born on the
    saved forward (streamed, 160-row segments) in a dot-product test, the
    fused tangent bit-identical to the Born pass."""
    g = Geometry()
    fwd, gold = _gold70()
    B, at = 32, 13
    vecs = [_filler(g, i) for i in range(B)]
    vecs[at] = _vec70()
    P = Propagator([g] * B)
    applied = autotune.apply(P)
    print("batch32 configs", _configs(P), flush=True)
    assert applied["born"][3] == 2 and applied["forward"][3] == 2 and applied["adjoint"] == (128, 64, 6, 1)
    v = torch.tensor(np.concatenate(vecs))
    sl = slice(int(P.seis_off[at]), int(P.seis_off[at + 1]))
    vl = slice(int(P.vec_off[at]), int(P.vec_off[at + 1]))
    seis = P.forward(v)
    s13 = _np(seis[sl]).reshape(g.n_shots, g.nt, g.n_rcv)[fwd["shots"]]
    r_fwd = rel(s13, fwd["seis"])
    dd = seis.clone()
    dd[sl] = torch.tensor(gold["target"], device="cuda")
    s_g, misfit, grad = P.gradient(v, dd, on_residual=True)
    r_grad = rel(_np(grad[vl]), gold["grad"])
    r_mis = abs(float(misfit[at]) - float(gold["misfit"])) / float(gold["misfit"])
    dm = torch.tensor(np.concatenate([np.concatenate([np.random.default_rng(11 + i).normal(0, 20, g.nz * g.nx), [3.0]])
                                      for i in range(B)]))
    jdm = P.tangent_saved(v, dm)
    res = _np(s_g[sl]) - gold["target"].astype(np.float64)
    a = float(np.dot(res, _np(jdm[sl])))
    b = float(np.dot(_np(grad[vl]), _np(dm[vl])))
    print(f"batch32 slot {at}: forward {r_fwd:.2e}, gradient {r_grad:.2e}, misfit {r_mis:.2e}, "
          f"dot {abs(a - b) / abs(a):.2e}", flush=True)
    assert r_fwd < SEIS_TOL and r_grad < GRAD_TOL and r_mis < GRAD_TOL
    assert abs(a - b) / abs(a) < DOT_TOL
    _, jdm_fused = P.tangent(v, dm)
    assert torch.equal(jdm_fused, jdm)
    # the same model alone on a plan with its own tuned configs: bit-identical
    P1 = Propagator([g])
    autotune.apply(P1)
    s1, m1, g1 = P1.gradient(torch.tensor(vecs[at]), dd[sl], on_residual=True)
    assert torch.equal(s1, s_g[sl]) and torch.equal(g1, grad[vl]) and float(m1[0]) == float(misfit[at])
    P.close()
    P1.close()


def test_tuned_linesearch_plan_golden():
    """The benchmark's line-search fan-out (1600 forward-only 70x70 models,
    streamed whole-column strips of 128, 5 levels): the golden model at slot 777, its
    seismograms and its misfit against the fixtures."""
    g = Geometry()
    fwd, gold = _gold70()
    n, at = 1600, 777
    base = [_filler(g, i) for i in range(8)]
    vecs = [base[i % 8] for i in range(n)]
    vecs[at] = _vec70()
    P = Propagator([g] * n, caps=CAP_FORWARD)
    applied = autotune.apply(P)
    print("linesearch configs", _configs(P), flush=True)
    assert applied["forward"] == (128, 0, 5, 2)
    v = torch.tensor(np.concatenate(vecs))
    seis = P.forward(v)
    sl = slice(int(P.seis_off[at]), int(P.seis_off[at + 1]))
    r_fwd = rel(_np(seis[sl]).reshape(g.n_shots, g.nt, g.n_rcv)[fwd["shots"]], fwd["seis"])
    dd = seis.clone()
    dd[sl] = torch.tensor(gold["target"], device="cuda")
    mis = P.forward_misfit(v, dd)
    r_mis = abs(float(mis[at]) - float(gold["misfit"])) / float(gold["misfit"])
    print(f"linesearch slot {at}: forward {r_fwd:.2e}, misfit {r_mis:.2e}", flush=True)
    assert r_fwd < SEIS_TOL and r_mis < GRAD_TOL
    assert float(mis[at - 1]) == 0.0 and float(mis[at + 1]) == 0.0     # the other models fit their own data
    P.close()


def _size_case(i):
    nz, nx = make_golden.SIZES[i]
    g = Geometry(nz=nz, nx=nx)
    d = np.load(os.path.join(GOLD, f"ref{nz}x{nx}.npz"))
    vec = synth.to_vec(make_golden.size_model(i, nz, nx))
    dd = make_golden.adjoint_probe(g.n_shots, g.nt, g.n_rcv)
    return g, d, vec, dd


@pytest.mark.parametrize("ckpt", [0, 50])
def test_mixed_sizes_plan_golden(ckpt):
    """A plan of 70x70, 100x100, 70(z)x140(x) and 140x140 models (three of
    each size, tuned configs; full history or checkpoints every 50 steps):
    the fixture models' shot-0 seismograms and J^T dd against the oracle's,
    and each fixture model bit-identical alone (its own tuned plan) and in the
    mix."""
    cases = [_size_case(i) for i in range(3)]
    g70 = Geometry()
    geoms, vecs, dds, where = [], [], [], {}
    for rep in range(3):
        for ci, (g, d, vec, dd) in enumerate([(g70, None, _vec70(), None)] + cases):
            if rep == 1 and d is not None:
                where[ci] = len(geoms)
                vecs.append(vec)
                dds.append(dd)
            else:
                vecs.append(_filler(g, len(geoms)))
                dds.append(np.random.default_rng(len(geoms)).standard_normal(g.n_data).astype(np.float32))
            geoms.append(g)
    P = Propagator(geoms, ckpt_every=ckpt)
    autotune.apply(P)
    print(f"mixed plan ({len(geoms)} models, ckpt {ckpt}) configs", _configs(P), flush=True)
    v = torch.tensor(np.concatenate(vecs))
    seis, _, grad = P.gradient(v, torch.tensor(np.concatenate(dds)), on_residual=False)
    for ci, m in where.items():
        g, d, vec, dd = cases[ci - 1]
        s = _np(seis[int(P.seis_off[m]):int(P.seis_off[m + 1])]).reshape(g.n_shots, g.nt, g.n_rcv)[0]
        gr = grad[int(P.vec_off[m]):int(P.vec_off[m + 1])]
        r_s, r_g = rel(s, d["seis0"]), rel(_np(gr), d["grad"])
        print(f"  {g.nz}x{g.nx} at slot {m}: seismogram {r_s:.2e}, gradient {r_g:.2e}", flush=True)
        assert r_s < SEIS_TOL and r_g < GRAD_TOL
        P1 = Propagator([g], ckpt_every=ckpt)
        autotune.apply(P1)
        s1, _, g1 = P1.gradient(torch.tensor(vec), torch.tensor(dd), on_residual=False)
        assert torch.equal(s1, seis[int(P.seis_off[m]):int(P.seis_off[m + 1])]) and torch.equal(g1, gr)
        P1.close()
    P.close()
