"""Seeded synthetic velocity models (no dataset is available offline).

Styles follow the OpenFWI families the reference was built for
(FlatVel / CurveVel / FlatFault / CurveFault / Style): flat layers, curved
(sinusoidal) layers, faulted layers and a Gaussian-random-field texture over a
layered base.  Velocities are clipped to [1500, 4500] m/s.
"""
from __future__ import annotations

import numpy as np

VMIN, VMAX = 1500.0, 4500.0


def _layer_velocities(rng, n):
    v = np.sort(rng.uniform(VMIN, VMAX, n))
    if rng.random() < 0.3:  # occasional low-velocity layer
        k = rng.integers(1, n) if n > 1 else 0
        v[k] = rng.uniform(VMIN, v[k])
    return v


def flat_layers(nz, nx, rng, n_layers=None):
    n = int(n_layers or rng.integers(3, 7))
    depths = np.sort(rng.choice(np.arange(3, nz - 2), n - 1, replace=False))
    vel = _layer_velocities(rng, n)
    v = np.empty((nz, nx))
    z = np.arange(nz)[:, None]
    idx = np.searchsorted(depths, z, side="right").repeat(nx, 1)
    v[:] = vel[idx]
    return v


def _interfaces(nz, nx, rng, n, amp_range):
    depths = np.sort(rng.choice(np.arange(5, nz - 5), n - 1, replace=False)).astype(float)
    x = np.arange(nx)
    surf = []
    for d in depths:
        amp = rng.uniform(*amp_range)
        k = rng.uniform(0.5, 2.0) * 2 * np.pi / nx
        ph = rng.uniform(0, 2 * np.pi)
        surf.append(d + amp * np.sin(k * x + ph))
    return np.array(surf)


def curved_layers(nz, nx, rng, n_layers=None, amp_range=(5, 15), fault=False):
    n = int(n_layers or rng.integers(3, 7))
    surf = _interfaces(nz, nx, rng, n, amp_range)
    vel = _layer_velocities(rng, n)
    z = np.arange(nz)[:, None].astype(float)
    if fault:
        # one dipping fault: cells right of the fault line are shifted by `throw`
        x0 = rng.uniform(0.3, 0.7) * nx
        dip = rng.uniform(-1.0, 1.0)
        throw = rng.uniform(3, 10) * rng.choice([-1, 1])
        xx = np.arange(nx)[None, :]
        right = xx > x0 + dip * (z - nz / 2)
        zz = np.where(right, z - throw, z)
    else:
        zz = np.broadcast_to(z, (nz, nx))
    idx = np.zeros((nz, nx), dtype=int)
    for s in surf:
        idx += (zz >= s[None, :]).astype(int)
    return vel[np.clip(idx, 0, n - 1)]


def faulted_layers(nz, nx, rng, n_layers=None):
    return curved_layers(nz, nx, rng, n_layers, amp_range=(0, 0.5), fault=True)


def grf_texture(nz, nx, rng, corr=None, sigma=None):
    """Gaussian random field (squared-exponential spectrum) over a layered base."""
    base = flat_layers(nz, nx, rng)
    ell = float(corr or rng.uniform(5, 20))
    sig = float(sigma or rng.uniform(100, 400))
    kz = np.fft.fftfreq(nz)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    spec = np.exp(-2 * (np.pi * ell) ** 2 * (kz ** 2 + kx ** 2))
    noise = rng.normal(size=(nz, nx)) + 1j * rng.normal(size=(nz, nx))
    field = np.real(np.fft.ifft2(noise * np.sqrt(spec)))
    field = field / (field.std() + 1e-12) * sig
    return base + field


STYLES = {"flat": flat_layers, "curved": curved_layers, "faulted": faulted_layers,
          "grf": grf_texture}


def make_model(style: str, nz: int, nx: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = STYLES[style](nz, nx, rng)
    return np.clip(v, VMIN, VMAX)


def curve_fault_model(nz, nx, seed):
    """True model of the single-inversion workload: curved layers (sinusoidal
    interfaces, amplitude 5-15 cells) plus one dipping fault (throw 3-10 cells)."""
    rng = np.random.default_rng(seed)
    return np.clip(curved_layers(nz, nx, rng, amp_range=(5, 15), fault=True), VMIN, VMAX)


def depth_linear_start(nz, nx, v_top=1500.0, v_bottom=4000.0):
    """Starting model: velocity linear in depth from v_top to v_bottom."""
    return np.repeat(np.linspace(v_top, v_bottom, nz)[:, None], nx, axis=1)


def to_vec(v: np.ndarray, min_vel: float | None = None) -> np.ndarray:
    """[v.ravel(), min_vel] with min_vel = min(v) by default (Velocity.to_vector)."""
    v = np.asarray(v, dtype=np.float64)
    mv = float(v.min()) if min_vel is None else float(min_vel)
    return np.concatenate([v.ravel(), [mv]])
