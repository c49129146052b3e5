"""Write full-size golden vectors from the FP64 oracle (test fixtures).

Reference geometry (core/seis_forward2.py:69-101): 70x70, dx 10 m, nbc 120,
nt 999, dt 1 ms, 15 Hz Ricker, shots at x = 0, 17, 34, 52, 69, 70 receivers.
Model: seeded curved layers + dipping fault (seismic_amd.synth.curve_fault_model,
seed 7).  Target: the forward of the same model with +2 % velocity, rounded
to float32 (the precision observed data is handed over in).

  ref70_fwd.npz   seismograms of shots 0 and 2 (float32, (2, 999, 70))
  ref70_grad.npz  the float32 target (349650), J^T (F(v) - target) for all 5
                  shots (float64, 4901) and the misfit sum((F(v) - target)^2)

The grid sizes of the batched and served workloads (100x100, 70x140 (z x x),
140x140; nbc 120, nt 999, shots at round(linspace(0, nx-1, 5)), receivers on
every surface column), one curved-fault model each (seed 20 + size index):

  ref{nz}x{nx}.npz  seismogram of shot 0 (float32, (999, nx)) and J^T dd for
                    the float32 data vector dd = adjoint_probe(...) (float64)

An LBFGS inversion as the reference's FlatVel model runs it in test mode
(core/seis_model.py:15-27: row total variation, lam 1e-8, 30 iterations,
tolerance_change 1e-9; core/seis_invert.py:109-152): a flat-layered 70x70
model (seed 3), started from its rows smoothed over 13 rows:

  ref70_lbfgs.npz   float32 observed data (the oracle forward of the true
                    model), the start vector, the oracle's inverted vector,
                    its evaluation count and data misfit

Run from the repo root: python tests/golden/make_golden.py [ref70] [sizes]
[lbfgs70]  (~5 minutes each part; lbfgs70 uses 5 processes).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import fd2d  # noqa: E402
from seismic_amd import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [(100, 100), (70, 140), (140, 140)]


def size_model(i, nz, nx):
    """The velocity model of the i-th fixture size (float64 (nz, nx))."""
    return synth.curve_fault_model(nz, nx, seed=20 + i)


def adjoint_probe(n_shots, nt, n_rcv, seed=5):
    """Data-space vector the size fixtures back-propagate: seeded standard
    normal samples in float32 (the GPU tests regenerate it)."""
    return np.random.default_rng(seed).standard_normal(n_shots * nt * n_rcv).astype(np.float32)


def main(parts=("ref70", "sizes", "lbfgs70")):
    if "ref70" in parts:
        write_ref70()
    if "sizes" in parts:
        write_sizes()
    if "lbfgs70" in parts:
        write_lbfgs70()


def lbfgs70_case():
    """(true model (70, 70), start vector) of the LBFGS fixture."""
    true = synth.make_model("flat", 70, 70, 3)
    rows = true[:, 0]
    k = np.ones(13) / 13
    smooth = np.convolve(np.pad(rows, 6, mode="edge"), k, mode="valid")
    return true, synth.to_vec(np.repeat(smooth[:, None], 70, 1), min_vel=float(true.min()))


def _shot(a):
    vec, g, s, kw = a
    kw = dict(kw)
    if kw.get("vec_adjoint") is not None:
        kw["vec_adjoint"] = kw["vec_adjoint"].reshape(g.n_shots, g.nt, g.n_rcv)[s].reshape(-1)
    return fd2d.vel_to_seis(vec, g, shots=[s], **kw)


def write_lbfgs70():
    from concurrent.futures import ProcessPoolExecutor

    from oracle import invert as oinv
    from oracle import priors as opri
    g = fd2d.Geometry()
    true, start = lbfgs70_case()
    with ProcessPoolExecutor(g.n_shots) as pool:
        def v2s(vec, **kw):
            outs = list(pool.map(_shot, [(vec, g, s, kw) for s in range(g.n_shots)]))
            gr = None if outs[0][2] is None else np.sum([o[2] for o in outs], axis=0)
            return np.concatenate([o[0] for o in outs]), None, gr

        target32 = v2s(synth.to_vec(true))[0].astype(np.float32)
        v, nfev = oinv.lbfgs_invert(start, target32.astype(np.float64), opri.RowTotalVariation(70, 70, lam=1e-8), g,
                                    max_iter=30, tolerance_change=1e-9, v2s=v2s)
        misfit = float(np.mean((v2s(v)[0] - target32.astype(np.float64)) ** 2))
    np.savez_compressed(os.path.join(HERE, "ref70_lbfgs.npz"), target=target32, start=start, v=v,
                        nfev=np.array(nfev), misfit=np.array(misfit))
    print("written lbfgs70", nfev, misfit, np.abs(v - synth.to_vec(true))[:-1].mean())


def write_ref70():
    g = fd2d.Geometry()
    v = synth.curve_fault_model(70, 70, seed=7)
    vec = synth.to_vec(v)
    target32 = fd2d.forward_only(synth.to_vec(v * 1.02), g).reshape(-1).astype(np.float32)
    target = target32.astype(np.float64)
    seis, _, grad = fd2d.vel_to_seis(vec, g, vec_adjoint=target, adjoint_on_residual=True)
    s3 = seis.reshape(5, g.nt, g.n_rcv)
    np.savez_compressed(os.path.join(HERE, "ref70_fwd.npz"), seis=s3[[0, 2]].astype(np.float32),
                        shots=np.array([0, 2]))
    np.savez_compressed(os.path.join(HERE, "ref70_grad.npz"), grad=grad, target=target32,
                        misfit=np.array(np.sum((seis - target) ** 2)))
    print("written", np.abs(s3).max(), np.linalg.norm(grad))


def write_sizes():
    for i, (nz, nx) in enumerate(SIZES):
        gs = fd2d.Geometry(nz=nz, nx=nx)
        vs = synth.to_vec(size_model(i, nz, nx))
        dd = adjoint_probe(gs.n_shots, gs.nt, gs.n_rcv).astype(np.float64)
        seis, _, grad = fd2d.vel_to_seis(vs, gs, vec_adjoint=dd, adjoint_on_residual=False)
        s0 = seis.reshape(gs.n_shots, gs.nt, gs.n_rcv)[0]
        np.savez_compressed(os.path.join(HERE, f"ref{nz}x{nx}.npz"), seis0=s0.astype(np.float32), grad=grad)
        print("written", nz, nx, np.abs(s0).max(), np.linalg.norm(grad))


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("ref70", "sizes", "lbfgs70"))
