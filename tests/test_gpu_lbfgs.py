"""LBFGS direction kernel (seis_lbfgs_direction, the reference's
lbfgs_backward / lbfgs_forward kernels with the history update of
core/seis_numerics.py:350-392) against the oracle's restatement
(oracle.invert.lbfgs_history_update / lbfgs_two_loop) over whole sequences
of iterations: bounded and unbounded histories, rejected pairs (y.s <=
1e-10), sizes from 1 to 19601 parameters; batch invariance; the scalars."""
import numpy as np
import pytest
import torch

from oracle import invert as oinv
from seismic_amd import invert


def _sequence(n, cap, steps, seed, reject_at=()):
    """(gradient, step) pairs: gradients of a convex quadratic along a
    descent path, with noise on the steps in ``reject_at``; the pairs y.s of
    the random steps take both signs, so pairs are kept and rejected."""
    rng = np.random.default_rng(seed)
    diag = rng.uniform(0.5, 5.0, n)
    x = rng.normal(0, 1, n)
    out = []
    for k in range(steps):
        g = diag * x + (0.3 * rng.normal(0, 1, n) if k in reject_at else 0.0)
        out.append((g, float(rng.uniform(0.2, 1.5))))
        x = x - 0.1 * g
    return out


def _run(cases):
    """Drive several histories through the kernel, one launch per
    iteration for all of them; returns per case the list of (d, scalars)
    and checks every step against the oracle."""
    hists = [invert.LbfgsHistory(n, cap, "cuda") for n, cap, _ in cases]
    host = [{"dirs": [], "stps": [], "ro": [], "H": 1.0} for _ in cases]
    prev = [None] * len(cases)
    results = [[] for _ in cases]
    steps = max(len(seq) for _, _, seq in cases)
    for k in range(steps):
        live = [i for i, (_, _, seq) in enumerate(cases) if k < len(seq)]
        batch = []
        for i in live:
            g, t = cases[i][2][k]
            gp, dp, tp = prev[i] if prev[i] is not None else (None, None, None)
            batch.append(invert.LbfgsStep(hists[i], torch.tensor(g, device="cuda"),
                                          None if gp is None else torch.tensor(gp, device="cuda"),
                                          None if dp is None else torch.tensor(dp, device="cuda"), tp, k == 0))
        outs = invert.lbfgs_directions(batch)
        for i, (d, sc) in zip(live, outs):
            g, t = cases[i][2][k]
            d = d.cpu().numpy()
            h = host[i]
            if k == 0:
                ref = -g
            else:
                gp, dp, tp = prev[i]
                h["H"] = oinv.lbfgs_history_update(h["dirs"], h["stps"], h["ro"], h["H"], g, gp, dp, tp,
                                                   cases[i][1])
                ref = oinv.lbfgs_two_loop(h["dirs"], h["stps"], h["ro"], h["H"], g)
            err = np.abs(d - ref).max() / np.abs(ref).max()
            assert err < 1e-11, (i, k, err)
            assert abs(sc[0] - g @ ref) <= 1e-11 * np.linalg.norm(g) * np.linalg.norm(ref)
            np.testing.assert_allclose(sc[1:], [np.abs(ref).max(), np.abs(g).sum(), np.abs(g).max()], rtol=1e-11)
            results[i].append((d, sc))
            prev[i] = (g, d, t)
    return results


@pytest.mark.gpu
def test_lbfgs_direction_matches_oracle_sequences():
    cases = [(71, 30, _sequence(71, 30, 14, 0)),                       # unbounded (cap = max_iter)
             (4901, 10, _sequence(4901, 10, 16, 1, reject_at=(5, 9))),  # history size 10, rejected pairs
             (19601, 3, _sequence(19601, 3, 9, 2)),                     # ring wraps several times
             (1, 2, _sequence(1, 2, 6, 3)),
             (1500, 1, _sequence(1500, 1, 7, 4, reject_at=(3,)))]
    print("lbfgs sequences:", [(n, cap, len(s)) for n, cap, s in cases], flush=True)
    _run(cases)


@pytest.mark.gpu
def test_lbfgs_direction_batch_invariance():
    """Each history's directions are bit-identical alone and in a batch of
    20 (two launches of 16 + 4)."""
    cases = [(300 + 97 * i, 4 + i % 3, _sequence(300 + 97 * i, 4 + i % 3, 6, 10 + i)) for i in range(20)]
    batched = _run(cases)
    for i in (0, 7, 19):
        alone = _run([cases[i]])[0]
        for (d1, s1), (d2, s2) in zip(batched[i], alone):
            assert np.array_equal(d1, d2) and s1 == s2


@pytest.mark.gpu
def test_lbfgs_direction_rejects_bad_sizes():
    h = invert.LbfgsHistory(40000, 3, "cuda")
    with pytest.raises(ValueError):
        invert.lbfgs_directions([invert.LbfgsStep(h, torch.zeros(40000, dtype=torch.float64, device="cuda"),
                                                  None, None, 1.0, True)])
