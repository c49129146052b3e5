"""Batched propagator plans and the drop-in ``vel_to_seis``.

``Propagator`` wraps one libseis plan: a fixed list of models (each with its
own ``Geometry``), device workspaces and the captured hipGraphs of the time
loops.  Inputs and outputs are torch tensors on the GPU; model m's slice of a
batched tensor is ``[vec_off[m]:vec_off[m+1]]`` (parameters) or
``[seis_off[m]:seis_off[m+1]]`` (data).

``vel_to_seis`` keeps the reference signature (core/seis_forward2.py:146-150)
and return tuple (:331) so that callers written against the reference work
unchanged on one GPU.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import CAP_ADJOINT, CAP_FORWARD, CAP_TANGENT
from .geometry import Geometry


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def model_desc(g: Geometry) -> _lib.ModelDesc:
    if g.n_shots > _lib.MAX_SHOTS:
        raise ValueError(f"at most {_lib.MAX_SHOTS} shots per model")
    d = _lib.ModelDesc()
    d.nz, d.nx, d.dx = g.nz, g.nx, g.dx
    d.n_shots = g.n_shots
    d.src_z = g.src_row
    for i, x in enumerate(g.src_x):
        d.src_x[i] = x
    d.rcv_z = g.rcv_row
    d.rcv_x0 = g.rcv_x[0]
    d.rcv_dx = g.rcv_x[1] - g.rcv_x[0] if g.n_rcv > 1 else 1
    d.n_rcv = g.n_rcv
    return d


def _descs(geoms: Sequence[Geometry]):
    keys = {g.time_key() for g in geoms}
    if len(keys) != 1:
        raise ValueError("all models of a plan must share nt, dt, nbc and freq")
    arr = (_lib.ModelDesc * len(geoms))(*[model_desc(g) for g in geoms])
    g0 = geoms[0]
    wl = np.ascontiguousarray(g0.wavelet(), dtype=np.float64)
    td = _lib.TimeDesc()
    td.nt, td.dt, td.nbc = g0.nt, g0.dt, g0.nbc
    td.wavelet = wl.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return arr, td, wl


def plan_bytes(geoms: Sequence[Geometry], ckpt_every: int = 0,
               caps: int = CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT) -> int:
    """Device bytes a plan would allocate (runs on the host only)."""
    arr, td, _ = _descs(geoms)
    n = _lib.lib().seis_plan_bytes(arr, len(geoms), ctypes.byref(td), ckpt_every, caps)
    if n < 0:
        _lib.check(int(n), "seis_plan_bytes")
    return int(n)


class Propagator:
    """A libseis plan over a fixed batch of models."""

    def __init__(self, geoms: Sequence[Geometry] | Geometry, ckpt_every: int = 0,
                 caps: int = CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT, device="cuda"):
        if isinstance(geoms, Geometry):
            geoms = [geoms]
        self.geoms = list(geoms)
        self.n_models = len(self.geoms)
        self.ckpt_every = ckpt_every
        self.caps = caps
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("Propagator needs a ROCm GPU (torch.cuda)")
        L = _lib.lib()
        arr, td, self._wavelet = _descs(self.geoms)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(L.seis_plan_create(arr, self.n_models, ctypes.byref(td), ckpt_every, caps,
                                          ctypes.byref(handle)), "seis_plan_create")
        self._h = handle
        vo = (ctypes.c_int64 * (self.n_models + 1))()
        so = (ctypes.c_int64 * (self.n_models + 1))()
        _lib.check(L.seis_plan_layout(self._h, vo, so), "seis_plan_layout")
        self.vec_off = np.array(vo[:], dtype=np.int64)
        self.seis_off = np.array(so[:], dtype=np.int64)
        self.n_vec = int(self.vec_off[-1])
        self.n_seis = int(self.seis_off[-1])
        self.device_bytes = int(L.seis_plan_device_bytes(self._h))
        self._gen = 0          # bumped by every call that rewrites wavefield state
        self._saved_gen = -1   # generation of the last forward_save
        # Full-history plans keep every step of a saving forward (gradient,
        # forward_save, gn_product); the *_saved passes reuse it.  _lin holds
        # the parameter vector each model's saved forward ran at.
        full = ckpt_every == 0 or ckpt_every >= self.geoms[0].nt
        self.full_history = bool(full and (caps & CAP_ADJOINT))
        self._lin = torch.full((self.n_vec,), float("nan"), device=self.device) if self.full_history else None
        self._act_idx = None   # parameter indices of the active models (None = all)

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().seis_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- helpers
    def _vec(self, v: torch.Tensor) -> torch.Tensor:
        v = torch.as_tensor(v).reshape(-1)
        if v.numel() != self.n_vec:
            raise ValueError(f"parameter vector has {v.numel()} entries, plan expects {self.n_vec}")
        return v.to(device=self.device, dtype=torch.float32).contiguous()

    def _data(self, d: torch.Tensor) -> torch.Tensor:
        d = torch.as_tensor(d).reshape(-1)
        if d.numel() != self.n_seis:
            raise ValueError(f"data vector has {d.numel()} entries, plan expects {self.n_seis}")
        return d.to(device=self.device, dtype=torch.float32).contiguous()

    KINDS = {"forward": _lib.KIND_FORWARD, "tangent": _lib.KIND_TANGENT, "adjoint": _lib.KIND_ADJOINT,
             "born": _lib.KIND_BORN, "record": _lib.KIND_RECORD}

    def set_config(self, kind: str, rx: int, rz: int, k: int, strips: int | None = None):
        """Launch config of one kernel kind: rx x rz LDS window, k steps per
        launch, row ownership (1 = consecutive rows per wavefront, 0 =
        round-robin; None keeps the current one).  strips = 2 streams
        rx-column strips down rz-row segments (rz = 0: whole columns) with k
        pipelined time levels."""
        if strips is None:
            strips = self.get_config(kind)[3]
        _lib.check(_lib.lib().seis_plan_set_config(self._h, self.KINDS[kind], rx, rz, k, int(strips)),
                   "seis_plan_set_config")

    def set_stream_threshold(self, min_active_tiles: int):
        """Streamed configs run a pass only with at least this many active
        strip segments, else the tiled tail window runs it (same results);
        -1: the device's CU count (default), 0: always stream."""
        _lib.check(_lib.lib().seis_plan_set_stream_threshold(self._h, int(min_active_tiles)),
                   "seis_plan_set_stream_threshold")

    def get_config(self, kind: str):
        """(rx, rz, k, strips)."""
        v = [ctypes.c_int32() for _ in range(4)]
        _lib.check(_lib.lib().seis_plan_get_config(self._h, self.KINDS[kind], *[ctypes.byref(x) for x in v]),
                   "seis_plan_get_config")
        return tuple(x.value for x in v)

    def set_active(self, mask: Sequence[bool]):
        mask = [bool(a) for a in mask]
        m = (ctypes.c_int32 * self.n_models)(*[1 if a else 0 for a in mask])
        _lib.check(_lib.lib().seis_set_active_async(self._h, m, _stream()), "seis_set_active_async")
        if all(mask):
            self._act_idx = None
            return
        # parameter indices of the active models, kept per mask: a serving
        # loop cycles through few masks, and each upload would wait for the stream
        key = tuple(mask)
        cache = self.__dict__.setdefault("_act_cache", {})
        if key not in cache:
            if len(cache) >= 64:
                cache.pop(next(iter(cache)))
            idx = [np.arange(self.vec_off[i], self.vec_off[i + 1]) for i, a in enumerate(mask) if a]
            cache[key] = torch.from_numpy(np.concatenate(idx) if idx else np.zeros(0, np.int64)).to(self.device)
        self._act_idx = cache[key]

    def _note_saved(self, v: torch.Tensor):
        """The active models' saved forward now belongs to v."""
        if self._lin is None:
            return
        if self._act_idx is None:
            self._lin.copy_(v)
        else:
            self._lin.index_copy_(0, self._act_idx, v.index_select(0, self._act_idx))

    def saved_matches(self, vec) -> bool:
        """True when every active model's saved forward ran at vec (one
        device-to-host sync), i.e. the *_saved passes apply."""
        if self._lin is None or not (self.caps & CAP_TANGENT):
            return False
        v = self._vec(vec)
        if self._act_idx is None:
            return bool(torch.equal(v, self._lin))
        return bool(torch.equal(v.index_select(0, self._act_idx), self._lin.index_select(0, self._act_idx)))

    def split_vec(self, v: torch.Tensor):
        return [v[self.vec_off[i]:self.vec_off[i + 1]] for i in range(self.n_models)]

    def split_seis(self, s: torch.Tensor):
        return [s[self.seis_off[i]:self.seis_off[i + 1]] for i in range(self.n_models)]

    # ------------------------------------------------------------- passes
    def forward(self, vec, out: torch.Tensor | None = None) -> torch.Tensor:
        self._gen += 1
        v = self._vec(vec)
        seis = out if out is not None else torch.empty(self.n_seis, device=self.device)
        _lib.check(_lib.lib().seis_forward(self._h, _ptr(v), _ptr(seis), _stream()), "seis_forward")
        return seis

    def forward_misfit(self, vec, dd, seis_out: torch.Tensor | None = None) -> torch.Tensor:
        """misfit[m] = sum((F(vec) - dd)^2) per model, float64 (n_models,)."""
        self._gen += 1
        v, d = self._vec(vec), self._data(dd)
        misfit = torch.empty(self.n_models, device=self.device, dtype=torch.float64)
        _lib.check(_lib.lib().seis_forward_misfit(self._h, _ptr(v), _ptr(d), _ptr(seis_out), _ptr(misfit),
                                                  _stream()), "seis_forward_misfit")
        return misfit

    def tangent(self, vec, dvec):
        self._gen += 1
        v, dv = self._vec(vec), self._vec(dvec)
        seis = torch.empty(self.n_seis, device=self.device)
        dseis = torch.empty(self.n_seis, device=self.device)
        _lib.check(_lib.lib().seis_tangent(self._h, _ptr(v), _ptr(dv), _ptr(seis), _ptr(dseis),
                                           _stream()), "seis_tangent")
        return seis, dseis

    def gradient(self, vec, dd, on_residual: bool = True):
        """(seis, misfit[n_models] float64 or None, grad float64)."""
        self._gen += 1
        v, d = self._vec(vec), self._data(dd)
        seis = torch.empty(self.n_seis, device=self.device)
        grad = torch.empty(self.n_vec, device=self.device, dtype=torch.float64)
        misfit = torch.empty(self.n_models, device=self.device, dtype=torch.float64) if on_residual else None
        _lib.check(_lib.lib().seis_gradient(self._h, _ptr(v), _ptr(d), int(bool(on_residual)), _ptr(seis),
                                            _ptr(misfit), _ptr(grad), _stream()), "seis_gradient")
        self._note_saved(v)
        return seis, misfit, grad

    def forward_save(self, vec) -> tuple:
        """Forward that keeps its checkpoints for a later ``adjoint``.
        Returns (seis, token); the token must be passed to ``adjoint``."""
        self._gen += 1
        v = self._vec(vec)
        seis = torch.empty(self.n_seis, device=self.device)
        _lib.check(_lib.lib().seis_forward_save(self._h, _ptr(v), _ptr(seis), _stream()), "seis_forward_save")
        self._note_saved(v)
        self._saved_gen = self._gen
        return seis, self._gen

    def saved_valid(self, token: int) -> bool:
        return token == self._saved_gen == self._gen

    def adjoint(self, vec, dd, token: int, on_residual: bool = False):
        """Reverse pass from the forward_save identified by ``token``:
        (misfit or None, grad float64).  Raises if another pass ran since."""
        if not self.saved_valid(token):
            raise RuntimeError("the saved forward was overwritten by a later pass on this plan")
        v, d = self._vec(vec), self._data(dd)
        grad = torch.empty(self.n_vec, device=self.device, dtype=torch.float64)
        misfit = torch.empty(self.n_models, device=self.device, dtype=torch.float64) if on_residual else None
        _lib.check(_lib.lib().seis_adjoint(self._h, _ptr(v), _ptr(d), int(bool(on_residual)), _ptr(misfit),
                                           _ptr(grad), _stream()), "seis_adjoint")
        return misfit, grad   # the reverse pass leaves the saved forward intact

    def gn_product(self, vec, dvec):
        """(J dvec, J^T J dvec) at vec."""
        self._gen += 1
        v, dv = self._vec(vec), self._vec(dvec)
        dseis = torch.empty(self.n_seis, device=self.device)
        grad = torch.empty(self.n_vec, device=self.device, dtype=torch.float64)
        _lib.check(_lib.lib().seis_gn_product(self._h, _ptr(v), _ptr(dv), _ptr(dseis), _ptr(grad),
                                              _stream()), "seis_gn_product")
        self._note_saved(v)
        return dseis, grad

    def _check_lin(self, v, check: bool):
        if check and not self.saved_matches(v):
            raise RuntimeError("vec is not the point of the saved forward of every active model "
                               "(run gradient, forward_save or gn_product at it first)")

    def tangent_saved(self, vec, dvec, check: bool = True) -> torch.Tensor:
        """J dvec from the saved forward at vec (no forward recomputation);
        bit-identical to ``tangent(vec, dvec)[1]``.  ``check`` verifies that
        vec is the saved point."""
        v, dv = self._vec(vec), self._vec(dvec)
        self._check_lin(v, check)
        dseis = torch.empty(self.n_seis, device=self.device)
        _lib.check(_lib.lib().seis_tangent_saved(self._h, _ptr(v), _ptr(dv), _ptr(dseis), _stream()),
                   "seis_tangent_saved")
        return dseis

    def gn_product_saved(self, vec, dvec, check: bool = True):
        """(J dvec, J^T J dvec) from the saved forward at vec; bit-identical
        to ``gn_product``."""
        v, dv = self._vec(vec), self._vec(dvec)
        self._check_lin(v, check)
        dseis = torch.empty(self.n_seis, device=self.device)
        grad = torch.empty(self.n_vec, device=self.device, dtype=torch.float64)
        _lib.check(_lib.lib().seis_gn_product_saved(self._h, _ptr(v), _ptr(dv), _ptr(dseis), _ptr(grad),
                                                    _stream()), "seis_gn_product_saved")
        return dseis, grad


# ----------------------------------------------------------------- drop-in
_default: dict = {}


def default_propagator(geometry: Geometry | None = None, ckpt_every: int = 0) -> Propagator:
    """One cached single-model plan per geometry (the reference's module globals)."""
    g = geometry or Geometry()
    key = (g, ckpt_every)
    if key not in _default:
        from . import autotune
        P = Propagator([g], ckpt_every=ckpt_every)
        autotune.apply(P)
        _default[key] = P
    return _default[key]


def vel_to_seis(vec, vec_diff=None, vec_adjoint=None, adjoint_on_residual=False,
                geometry: Geometry | None = None):
    """Drop-in for core/seis_forward2.py:146 vel_to_seis.

    Returns (result, result_diff, result_adjoint) as float64 column tensors:
    result = seismogram (n_data, 1); result_diff = J @ vec_diff; and
    result_adjoint = J^T @ vec_adjoint, or J^T @ (result - vec_adjoint) when
    ``adjoint_on_residual``.  Computation runs in float32 on the GPU.
    """
    g = geometry or Geometry()
    P = default_propagator(g)
    vec = torch.as_tensor(vec)
    if tuple(vec.shape) != (g.n_params, 1):
        raise ValueError(f"vec must have shape ({g.n_params}, 1), got {tuple(vec.shape)}")
    if vec_diff is not None and tuple(torch.as_tensor(vec_diff).shape) != (g.n_params, 1):
        raise ValueError(f"vec_diff must have shape ({g.n_params}, 1)")
    if vec_adjoint is not None and tuple(torch.as_tensor(vec_adjoint).shape) != (g.n_data, 1):
        raise ValueError(f"vec_adjoint must have shape ({g.n_data}, 1)")
    result_diff = None
    result_adjoint = None
    if vec_adjoint is not None:
        seis, _, grad = P.gradient(vec, vec_adjoint, on_residual=adjoint_on_residual)
        result_adjoint = grad.reshape(-1, 1)
        if vec_diff is not None:
            # the gradient's forward was saved at vec: no second forward
            dseis = P.tangent_saved(vec, vec_diff, check=False)
            result_diff = dseis.double().reshape(-1, 1)
    elif vec_diff is not None:
        seis, dseis = P.tangent(vec, vec_diff)
        result_diff = dseis.double().reshape(-1, 1)
    else:
        seis = P.forward(vec)
    torch.cuda.current_stream().synchronize()
    return seis.double().reshape(-1, 1), result_diff, result_adjoint
