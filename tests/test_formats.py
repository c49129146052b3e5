"""On-disk formats (core/kaggle_support.py:192-345, 414-476, 529-587) on the
CPU: the oracle's restatement pinned by hand-written expectations (the sample
list of a constructed directory tree, literal CSV lines, a score worked by
hand), then the product against the oracle (sample lists, the rows read from
memory-mapped files, cache names, submission bytes, scores).  The reference's
own data is absent; the trees here are synthetic files of the competition's
shapes."""
import os

import numpy as np
import pytest
import torch

from oracle import formats as of
from seismic_amd import datasets as ds


def _tree(root, rng):
    """train_samples/A: seis_2_1 (3 samples), seis_1_1 (2); train_samples/B:
    data/data1 (2), data/data2 (1) with model/; test: two files."""
    def put(path, arr):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.save(path, arr)
    seis = lambda n: rng.normal(0, 1, (n, 5, 1000, 70)).astype(np.float32)
    vel = lambda n: rng.uniform(1500, 4500, (n, 1, 70, 70)).astype(np.float32)
    a = os.path.join(root, "train_samples", "A")
    put(f"{a}/seis_2_1.npy", seis(3)); put(f"{a}/vel_2_1.npy", vel(3))
    put(f"{a}/seis_1_1.npy", seis(2)); put(f"{a}/vel_1_1.npy", vel(2))
    b = os.path.join(root, "train_samples", "B")
    put(f"{b}/data/data1.npy", seis(2)); put(f"{b}/model/model1.npy", vel(2))
    put(f"{b}/data/data2.npy", seis(1)); put(f"{b}/model/model2.npy", vel(1))
    put(os.path.join(root, "test", "zz9.npy"), seis(1)[0])
    put(os.path.join(root, "test", "aa1.npy"), seis(1)[0])


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    root = str(tmp_path_factory.mktemp("data"))
    _tree(root, np.random.default_rng(0))
    return root


def test_oracle_sample_order(tree):
    got = [(os.path.relpath(s["seis_file"], tree), s["seis_ind"], os.path.relpath(s["vel_file"], tree), s["family"])
           for s in of.list_train(tree)]
    want = [("train_samples/A/seis_1_1.npy", k, "train_samples/A/vel_1_1.npy", "A") for k in range(2)] + \
           [("train_samples/A/seis_2_1.npy", k, "train_samples/A/vel_2_1.npy", "A") for k in range(3)] + \
           [("train_samples/B/data/data1.npy", k, "train_samples/B/model/model1.npy", "B") for k in range(2)] + \
           [("train_samples/B/data/data2.npy", 0, "train_samples/B/model/model2.npy", "B")]
    assert got == want
    val = [(os.path.basename(s["seis_file"]), s["seis_ind"]) for s in of.list_train(tree, validation_only=True)]
    assert val == [("seis_1_1.npy", 0), ("seis_1_1.npy", 1), ("data1.npy", 0), ("data1.npy", 1)]
    assert [os.path.basename(s["seis_file"]) for s in of.list_test(tree)] == ["aa1.npy", "zz9.npy"]
    assert of.cache_name(of.list_train(tree)[3]) == "seis_2_1__A__1"
    assert of.cache_name(of.list_test(tree)[0]) == "aa1__test__None"


def test_oracle_reads_rows(tree):
    s = of.list_train(tree)[3]
    raw = np.load(s["seis_file"])
    assert of.read_seismogram(s).shape == (5, 999, 70)
    np.testing.assert_array_equal(of.read_seismogram(s), raw[1, :, :999, :])
    np.testing.assert_array_equal(of.read_seismogram(s, load_last_row=True), raw[1])
    g, vec = of.read_velocity(s)
    np.testing.assert_array_equal(g, np.load(s["vel_file"])[1, 0])
    assert vec.shape == (4901,) and vec[-1] == g.min()


def test_oracle_submission_lines(tmp_path):
    grid = np.full((70, 70), 3000.0)
    grid[0, 1] = 2234.56          # float32 2234.56005859375
    grid[0, 3] = 9000.0           # clipped to 4500
    grid[1, 1] = 2500.5
    s = dict(seis_file="/x/abc.npy", seis_ind=None, family="test")
    p = str(tmp_path / "o.csv")
    of.write_submission([s], [grid], p)
    lines = open(p, newline="").read().split("\r\n")
    assert lines[0] == "oid_ypos," + ",".join(f"x_{x}" for x in range(1, 70, 2))
    assert lines[1] == "abc_y_0,2234.56005859375,4500.0," + ",".join(["3000.0"] * 33)
    assert lines[2] == "abc_y_1,2500.5," + ",".join(["3000.0"] * 34)
    assert len(lines) == 72 and lines[-1] == ""
    of.write_submission([s], [grid], p, do_round=True)
    lines = open(p, newline="").read().split("\r\n")
    assert lines[1] == "abc_y_0,2235,4500," + ",".join(["3000"] * 33)
    assert lines[2] == "abc_y_1,2500," + ",".join(["3000"] * 34)      # half to even


def test_oracle_score_by_hand():
    t = [np.full((2, 2), 3000.0), np.full((2, 2), 2000.0)]
    g = [np.array([[3000.4, 3000.5], [3001.5, 2990.0]]), np.full((2, 2), 2000.0)]
    score, per, res = of.score_metric(t, g, ["F", "G"])
    assert res == [pytest.approx((0 + 0 + 2 + 10) / 4), 0.0]
    assert per == {"F": pytest.approx(3.0), "G": 0.0} and score == pytest.approx(1.5)


# ------------------------------------------------------------ product against the oracle
def _as_dict(s):
    return dict(seis_file=s.seis_file, seis_ind=s.seis_ind, vel_file=s.vel_file, vel_ind=s.vel_ind,
                family=s.family, is_train=s.is_train)


@pytest.mark.parametrize("val", [False, True])
def test_sample_lists_match_oracle(tree, val):
    assert [_as_dict(s) for s in ds.load_train(tree, validation_only=val)] == of.list_train(tree, val)
    assert [_as_dict(s) for s in ds.load_test(tree)] == of.list_test(tree)
    for s in ds.load_train(tree) + ds.load_test(tree):
        assert s.cache_name() == of.cache_name(_as_dict(s))


def test_batch_reads_match_oracle(tree):
    samples = ds.load_train(tree) + ds.load_test(tree)
    rng = np.random.default_rng(1)
    order = [samples[i] for i in rng.permutation(len(samples))] + samples     # runs and scattered rows
    for nt in (999, 1000, 37):
        out = np.zeros((len(order) + 2, 5 * nt * 70), np.float32)
        ds.read_seismograms(order, out, nt)
        for j, s in enumerate(order):
            ref = of.read_seismogram(_as_dict(s), load_last_row=True)[:, :nt, :]
            np.testing.assert_array_equal(out[j].astype(np.float64), ref.ravel())
        assert not out[len(order):].any()
    train = ds.load_train(tree)
    vel = np.zeros((len(train), 4901))
    ds.read_velocities(train, vel)
    for j, s in enumerate(train):
        np.testing.assert_array_equal(vel[j], of.read_velocity(_as_dict(s))[1])


@pytest.mark.parametrize("batch,depth", [(3, 2), (8, 1), (1, 3)])
def test_loader_host_path(tree, batch, depth):
    """The loader's batching and slab rotation (host copies; the HBM path is
    a GPU test)."""
    train = ds.load_train(tree)
    got = list(ds.Loader(train, batch=batch, depth=depth, device="cpu"))
    assert [len(b.samples) for b in got] == [min(batch, len(train) - i) for i in range(0, len(train), batch)]
    flat = [s for b in got for s in b.samples]
    assert flat == train
    seis = torch.cat([b.seis for b in got]).numpy()
    vel = torch.cat([b.vel for b in got]).numpy()
    for j, s in enumerate(train):
        np.testing.assert_array_equal(seis[j].astype(np.float64), of.read_seismogram(_as_dict(s)).ravel())
        np.testing.assert_array_equal(vel[j], of.read_velocity(_as_dict(s))[1])
    assert all(b.vel is None for b in ds.Loader(ds.load_test(tree), batch=2, device="cpu"))
    with pytest.raises(ValueError):
        ds.Loader(train, nt=1001)
    with pytest.raises(ValueError):
        ds.Loader(ds.load_test(tree), velocities=True)


def test_result_cache(tree, tmp_path):
    train = ds.load_train(tree)
    c = ds.ResultCache(str(tmp_path), "Default")
    assert c.get(train[0]) is None
    v = np.random.default_rng(2).normal(3000, 100, 4901)
    c.put(train[0], torch.tensor(v))
    np.testing.assert_array_equal(c.get(train[0]), v)
    assert sorted(os.listdir(tmp_path / "Default")) == [train[0].cache_name() + ".npy"]
    assert ds.ResultCache(str(tmp_path), "Default", read=False).get(train[0]) is None
    assert c.get(train[1]) is None


@pytest.mark.parametrize("do_round,do_range", [(False, True), (True, True), (False, False), (True, False)])
def test_submission_bytes_match_oracle(tmp_path, do_round, do_range):
    rng = np.random.default_rng(3)
    grids = [rng.uniform(1000, 5000, (70, 70)),                          # non-integers, some out of range
             np.round(rng.uniform(1400, 4600, (70, 70))),                 # integral floats
             np.round(rng.uniform(1400, 4600, (70, 70))) + 0.5,           # ties for the rounding
             np.where(rng.random((70, 70)) < 0.5, -0.0, 1e17) * rng.choice([1, -1], (70, 70))]
    grids[0][3, 5] = np.nan
    grids[0][4, 7] = 1e-7
    samples = [ds.Sample(f"/d/test/f{k}.npy") for k in range(len(grids))]
    a, b = str(tmp_path / "a.csv"), str(tmp_path / "b.csv")
    ds.write_submission(samples, [torch.tensor(g) for g in grids], a, do_round, do_range)
    with np.errstate(invalid="ignore"):
        of.write_submission([_as_dict(s) for s in samples], grids, b, do_round, do_range)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_score_matches_oracle():
    rng = np.random.default_rng(4)
    t = rng.uniform(1500, 4500, (6, 70, 70))
    g = t + rng.normal(0, 20, t.shape)
    g[0] = np.round(g[0]) + 0.5
    fams = ["A", "B", "A", "C", "B", "A"]
    score, per, res = ds.score_metric(torch.tensor(t), torch.tensor(g), fams)
    oscore, oper, ores = of.score_metric(list(t), list(g), fams)
    assert res == pytest.approx(ores, rel=1e-14) and per == pytest.approx(oper, rel=1e-14)
    assert score == pytest.approx(oscore, rel=1e-14)


def test_submission_and_loader_edges(tree, tmp_path):
    """The writer refuses grids that are not 70x70; the loader reports its
    batch count and yields nothing for an empty sample list."""
    with pytest.raises(ValueError):
        ds.write_submission([ds.Sample("/d/test/f.npy")], [np.zeros((70, 71))], str(tmp_path / "x.csv"))
    train = ds.load_train(tree)
    assert len(ds.Loader(train, batch=3, device="cpu")) == (len(train) + 2) // 3
    assert list(ds.Loader([], batch=3, device="cpu")) == []
