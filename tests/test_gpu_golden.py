"""Full-size parity against the committed golden vectors (reference geometry:
70x70, nbc 120, nt 999, 5 shots; fixtures written by tests/golden/make_golden.py
from the FP64 oracle, so these tests need no CPU propagation at run time)."""
import os

import numpy as np
import pytest
import torch

from conftest import rel
from seismic_amd import synth
from seismic_amd.geometry import Geometry
from seismic_amd.propagator import Propagator, vel_to_seis

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEIS_TOL = 1e-5
GRAD_TOL = 1e-4
DOT_TOL = 1e-6


def _load(name):
    return np.load(os.path.join(GOLD, name))


def _vec():
    return synth.to_vec(synth.curve_fault_model(70, 70, seed=7))


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def test_forward_golden():
    g = Geometry()
    d = _load("ref70_fwd.npz")
    out = _np(Propagator([g]).forward(torch.tensor(_vec()))).reshape(g.n_shots, g.nt, g.n_rcv)[d["shots"]]
    r = rel(out, d["seis"])
    print("forward vs golden rel", r, flush=True)
    assert r < SEIS_TOL


@pytest.mark.parametrize("ckpt", [0, 50])
def test_gradient_golden(ckpt):
    g = Geometry()
    d = _load("ref70_grad.npz")
    P = Propagator([g], ckpt_every=ckpt)
    _, misfit, grad = P.gradient(torch.tensor(_vec()), torch.tensor(d["target"]), on_residual=True)
    r = rel(_np(grad), d["grad"])
    m = float(misfit[0])
    print(f"ckpt {ckpt}: gradient rel {r:.3e}  misfit {m:.6e} vs {float(d['misfit']):.6e}", flush=True)
    assert r < GRAD_TOL
    assert abs(m - float(d["misfit"])) / float(d["misfit"]) < GRAD_TOL


def test_dropin_golden():
    """vel_to_seis with the reference's default shapes and FP64 columns."""
    d = _load("ref70_grad.npz")
    f = _load("ref70_fwd.npz")
    vec = torch.tensor(_vec()).reshape(-1, 1)
    s, _, gr = vel_to_seis(vec, None, torch.tensor(d["target"], dtype=torch.float64).reshape(-1, 1),
                           adjoint_on_residual=True)
    g = Geometry()
    assert s.shape == (g.n_data, 1) and gr.shape == (g.n_params, 1)
    assert rel(_np(s).reshape(g.n_shots, g.nt, g.n_rcv)[f["shots"]], f["seis"]) < SEIS_TOL
    assert rel(_np(gr), d["grad"]) < GRAD_TOL


def test_dot_product_full_size():
    """<J dm, r> = <dm, J^T r> at the reference size, r the golden residual."""
    g = Geometry()
    d = _load("ref70_grad.npz")
    vec = _vec()
    rng = np.random.default_rng(11)
    dm = np.concatenate([rng.normal(0, 20, g.nz * g.nx), [3.0]])
    P = Propagator([g])
    s, jdm = P.tangent(torch.tensor(vec), torch.tensor(dm))
    _, _, jt = P.gradient(torch.tensor(vec), torch.tensor(d["target"]), on_residual=True)
    r = _np(s) - d["target"].astype(np.float64)
    a = float(np.dot(r, _np(jdm)))
    b = float(np.dot(_np(jt), dm))
    print("full-size dot", a, b, abs(a - b) / abs(a), flush=True)
    assert abs(a - b) / abs(a) < DOT_TOL
