"""seis_label_flat_static, the labelling entry point that allocates nothing,
against seis_label_flat (same kernel, grid table staged in a stream-ordered
allocation): identical labels and counts on two calls in a row, the second
with fewer and smaller grids, and a clear error above its grid limit."""
import ctypes

import numpy as np
import pytest
import torch

from seismic_amd import _lib


def _grids(seed, shapes):
    """Integer-valued grids: random cells (many small areas) and layered
    blocks (a few large ones)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (nz, nx) in enumerate(shapes):
        if i % 2 == 0:
            v = rng.integers(0, 3, (nz, nx)).astype(np.float64)
        else:
            v = np.repeat(np.cumsum(rng.integers(0, 3, nz)).astype(np.float64)[:, None], nx, axis=1)
            v[rng.random((nz, nx)) < 0.05] += 7.0
        out.append(v)
    return out


def _call(name, grids, threshold):
    flat = torch.tensor(np.concatenate([g.ravel() for g in grids]), dtype=torch.float64, device="cuda")
    n = len(grids)
    nz = (ctypes.c_int32 * n)(*[g.shape[0] for g in grids])
    nx = (ctypes.c_int32 * n)(*[g.shape[1] for g in grids])
    labels = torch.full((flat.numel(),), -7, dtype=torch.int32, device="cuda")
    counts = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(_lib.lib(), name)(nz, nx, n, ctypes.c_void_p(flat.data_ptr()), float(threshold),
                                  ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(counts.data_ptr()), stream)
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return labels.cpu().numpy(), counts.cpu().numpy()


@pytest.mark.gpu
def test_static_labelling_equals_allocating_entry_point():
    calls = [([(70, 70), (100, 100), (20, 300), (156, 157), (1, 1), (5, 1)], 0, 0.5),
             ([(3, 4), (70, 70)], 1, 1.0)]          # second call: fewer, smaller grids
    for shapes, seed, thr in calls:
        print(f"labelling {len(shapes)} grids {shapes}, threshold {thr}", flush=True)
        grids = _grids(seed, shapes)
        lab_a, cnt_a = _call("seis_label_flat", grids, thr)
        lab_b, cnt_b = _call("seis_label_flat_static", grids, thr)
        assert np.array_equal(cnt_a, cnt_b)
        assert np.array_equal(lab_a, lab_b)
        assert (lab_b >= 0).all() and cnt_b.min() >= 1 and cnt_b.max() > 10
        pos = 0
        for g, c in zip(grids, cnt_b):
            lab = lab_b[pos:pos + g.size]
            pos += g.size
            assert lab.max() == c - 1 and lab[0] == 0


@pytest.mark.gpu
def test_static_labelling_refuses_too_many_grids():
    n = _lib.LABEL_STATIC_MAX_GRIDS + 1
    flat = torch.zeros(n, dtype=torch.float64, device="cuda")
    ones = (ctypes.c_int32 * n)(*([1] * n))
    labels = torch.empty(n, dtype=torch.int32, device="cuda")
    counts = torch.empty(n, dtype=torch.int32, device="cuda")
    rc = _lib.lib().seis_label_flat_static(ones, ones, n, ctypes.c_void_p(flat.data_ptr()), 1.0,
                                           ctypes.c_void_p(labels.data_ptr()), ctypes.c_void_p(counts.data_ptr()),
                                           None)
    assert rc == -1
    assert str(_lib.LABEL_STATIC_MAX_GRIDS) in _lib.lib().seis_last_error().decode()
