"""libseis.so loads and exports exactly what include/seis.h declares; the
host-only entry point (seis_plan_bytes) validates arguments without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from seismic_amd import _lib
from seismic_amd.geometry import Geometry
from seismic_amd.propagator import plan_bytes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "seis.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(seis_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_plan_bytes_reference_geometry():
    g = Geometry()
    full = plan_bytes([g], ckpt_every=0)
    ck = plan_bytes([g], ckpt_every=50)
    P = 310 * 310 * 4
    # full history: 5 shots x (nt+2) levels of q on the padded grid
    assert 5 * 1001 * P <= full < 5 * 1001 * P + 10 * 2 ** 20 + 5 * 12 * P
    assert ck < full / 5
    fwd_only = plan_bytes([g], caps=_lib.CAP_FORWARD)
    assert fwd_only < 5 * 4 * P + 10 * 2 ** 20
    assert plan_bytes([g] * 4, ckpt_every=50) > 3.9 * ck


@pytest.mark.parametrize("bad", [
    dict(nz=1), dict(nbc=1), dict(src_x=(70,)), dict(src_z=70), dict(rcv_x=tuple(range(60, 71))),
])
def test_plan_bytes_rejects_bad_geometry(bad):
    g = Geometry(**bad) if "rcv_x" not in bad and "src_x" not in bad else None
    if g is None:
        kw = dict(bad)
        g = Geometry(**kw)
    with pytest.raises(ValueError):
        plan_bytes([g])


def test_plan_bytes_rejects_grids_beyond_32bit_offsets():
    """A padded grid above INT32_MAX / 4 cells is refused (one wavefield is
    addressed with 32-bit byte offsets); one just below is accepted."""
    kw = dict(nbc=8, nt=200, n_shots=1, src_x=(5,), rcv_x=(1, 2, 3))
    with pytest.raises(ValueError, match="model too large"):
        plan_bytes([Geometry(nz=23200 - 16, nx=23200 - 16, **kw)], caps=_lib.CAP_FORWARD)
    assert plan_bytes([Geometry(nz=23168 - 16, nx=23168 - 16, **kw)], caps=_lib.CAP_FORWARD) > 0


def test_plan_bytes_rejects_bad_options():
    g = Geometry(nz=10, nx=10, nbc=5, nt=200)
    with pytest.raises(ValueError):
        plan_bytes([g], ckpt_every=1)
    with pytest.raises(ValueError):
        plan_bytes([g], caps=_lib.CAP_ADJOINT)
    with pytest.raises(ValueError):
        plan_bytes([g, Geometry(nz=10, nx=10, nbc=5, nt=300)])


def test_last_error_message():
    g = Geometry(nz=10, nx=10, nbc=5, nt=200)
    with pytest.raises(ValueError, match="ckpt_every"):
        plan_bytes([g], ckpt_every=1)


@pytest.mark.parametrize("src_z,rcv_z", [(1, 1), (0, 0), (0, 3), (4, 0), (2, 5)])
def test_geometry_rows_follow_adjust_sr(src_z, rcv_z):
    """The rows handed to the library are the oracle's adjust_sr rows
    (core/seis_forward2.py:60-68: a depth of 0 moves to row 1)."""
    from oracle import fd2d
    from seismic_amd.propagator import model_desc
    kw = dict(nz=12, nx=16, nbc=8, nt=50, src_x=(0, 7, 15), rcv_x=(2, 5, 8, 11, 14), src_z=src_z, rcv_z=rcv_z)
    d = model_desc(Geometry(**kw))
    og = fd2d.Geometry(**kw)
    igz, igx = og.rcv_cells()
    assert all(isz - og.nbc == d.src_z for isz, _ in og.src_cells())
    assert [isx - og.nbc for _, isx in og.src_cells()] == list(d.src_x[:d.n_shots])
    assert set(igz - og.nbc) == {d.rcv_z}
    assert list(igx - og.nbc) == [d.rcv_x0 + r * d.rcv_dx for r in range(d.n_rcv)]


def test_nt_below_wavelet_length_is_rejected():
    """ricker (core/seis_forward2.py:20-30) raises when nt is shorter than the
    wavelet (147 samples at 15 Hz, dt 1 ms); so does the geometry here."""
    assert len(Geometry(nt=147).wavelet()) == 147
    with pytest.raises(ValueError):
        Geometry(nt=146).wavelet()


def test_autotune_apply_falls_back_when_a_config_is_refused():
    """apply() takes the nearest table bucket the plan accepts: a streamed
    config refused for a short padded grid falls through to the next bucket,
    and with none accepted the plan keeps its defaults."""
    from seismic_amd import autotune

    class FakePlan:
        caps = 1
        ckpt_every = 0
        geoms = [Geometry(nz=2, nx=5, nbc=6, nt=150)]   # 14 padded rows

        def __init__(self):
            self.cfg = {}

        def set_config(self, kind, rx, rz, k, strips=None):
            if strips == 2 and 14 < 4 * k:
                raise ValueError("streamed forward needs padded grids of at least 4k rows")
            self.cfg[kind] = (rx, rz, k, strips)

    b = autotune.key(FakePlan(), "forward").split(":")[1]
    table = {f"forward:{b}:full": [128, 0, 4, 2], f"forward:{int(b) + 1}:full": [64, 32, 3, 1]}
    P = FakePlan()
    assert autotune.apply(P, table) == {"forward": (64, 32, 3, 1)}
    P = FakePlan()
    assert autotune.apply(P, {f"forward:{b}:full": [128, 0, 4, 2]}) == {} and P.cfg == {}
