"""On-disk formats on the GPU: the loader's pinned-slab / side-stream path
delivers the files' rows to HBM bit for bit while GPU work runs between
batches, and ``infer_dataset`` over competition-shaped files (classified,
chained inversions, result cache) equals the same chains run from the
tensors directly, then reads everything back from its cache."""
import os

import numpy as np
import pytest
import torch

from oracle import fd2d
from oracle import formats as of
from seismic_amd import datasets as ds
from seismic_amd import invert, schedule, synth
from seismic_amd.geometry import Geometry

pytestmark = pytest.mark.gpu


def _save(path, arr):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, arr)


@pytest.mark.parametrize("batch,depth", [(3, 2), (4, 1), (2, 3)])
def test_loader_rows_reach_hbm(tmp_path, batch, depth):
    rng = np.random.default_rng(batch)
    root = str(tmp_path)
    for k, n in enumerate((4, 3)):
        _save(f"{root}/train_samples/F/seis_{k}.npy", rng.normal(0, 1, (n, 5, 1000, 70)).astype(np.float32))
        _save(f"{root}/train_samples/F/vel_{k}.npy", rng.uniform(1500, 4500, (n, 1, 70, 70)).astype(np.float32))
    samples = ds.load_train(root)
    seen = []
    for b in ds.Loader(samples, batch=batch, depth=depth, device="cuda"):
        assert b.seis.is_cuda and b.vel.is_cuda
        x = torch.randn(2048, 2048, device="cuda")
        (x @ x).sum().item()                         # GPU work between batches
        seen.append((b.samples, b.seis.cpu().numpy(), b.vel.cpu().numpy()))
    flat = [s for ss, _, _ in seen for s in ss]
    assert flat == samples
    seis = np.concatenate([a for _, a, _ in seen])
    vel = np.concatenate([v for _, _, v in seen])
    for j, s in enumerate(samples):
        d = dict(seis_file=s.seis_file, seis_ind=s.seis_ind, vel_file=s.vel_file, vel_ind=s.vel_ind)
        np.testing.assert_array_equal(seis[j].astype(np.float64), of.read_seismogram(d).ravel())
        np.testing.assert_array_equal(vel[j], of.read_velocity(d)[1])


def test_infer_dataset_with_cache(tmp_path):
    """Four layered 70x70 models as test files (nt 300 of the 1000 rows
    used), classified and inverted in batches of 3 (a ragged last batch);
    equal to schedule.infer on the same requests at once and to the served
    pipeline (3 slots, samples streamed in as slots free up); the second call
    reads every result from the cache; submission and score follow (the
    score against the oracle's)."""
    nt = 300
    g, og = Geometry(nt=nt), fd2d.Geometry(nt=nt)
    root = str(tmp_path / "data")
    trues = [synth.make_model("flat", 70, 70, 40 + i) for i in range(4)]
    for i, t in enumerate(trues):
        seis = np.zeros((5, 1000, 70), np.float32)
        seis[:, :nt, :] = fd2d.vel_to_seis(synth.to_vec(t), og)[0].reshape(5, nt, 70)
        _save(f"{root}/test/s{i:02d}.npy", seis)
    samples = ds.load_test(root)
    start = torch.tensor(synth.to_vec(synth.depth_linear_start(70, 70, 1500, 4500), min_vel=1500.0),
                         dtype=torch.float64, device="cuda")
    cache = ds.ResultCache(str(tmp_path / "cache"), "Default")
    out, fams = ds.infer_dataset(samples, start=lambda s, seis: start.clone(), cache=cache, batch=3,
                                 test_mode=True, nt=nt)
    print("families", fams, flush=True)
    assert fams == ["FlatVel"] * 4
    targets = [torch.tensor(np.load(s.seis_file)[:, :nt, :].ravel(), device="cuda") for s in samples]
    reqs = [invert.Request(geom=g, target=t, prior=None, start=start.clone(), rid=i) for i, t in enumerate(targets)]
    ref, rfams = schedule.infer(reqs, test_mode=True, batched=True)
    assert rfams == fams
    for a, b in zip(out, ref):
        assert torch.equal(a, b)
    served, fams_s = ds.infer_dataset(samples, start=lambda s, seis: start.clone(), batch=3, test_mode=True, nt=nt,
                                      served=True)
    assert fams_s == fams
    for a, b in zip(out, served):
        assert torch.equal(a, b)
    again, fams2 = ds.infer_dataset(samples, cache=cache, batch=3, test_mode=True, nt=nt)
    assert fams2 == [None] * 4
    for a, b in zip(out, again):
        assert torch.equal(a.cpu(), b.cpu())
    sub = str(tmp_path / "submission.csv")
    ds.write_submission(samples, [v[:-1].reshape(70, 70) for v in out], sub)
    assert len(open(sub, newline="").read().split("\r\n")) == 1 + 4 * 70 + 1
    score, per, res = ds.score_metric(torch.tensor(np.stack(trues), device="cuda"), torch.stack(out), fams)
    oscore, oper, ores = of.score_metric(trues, [v[:-1].reshape(70, 70).cpu().numpy() for v in out], fams)
    print("score", score, per, flush=True)
    assert res == pytest.approx(ores, rel=1e-12) and score == pytest.approx(oscore, rel=1e-12)
