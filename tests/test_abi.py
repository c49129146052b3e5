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
    # full history: 5 shots x (nt+1) steps x (Lf(p), q) pairs of padded fields
    assert full >= 5 * 1000 * 2 * P
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
