"""Acquisition geometry of one model (reference module globals, generalised).

core/seis_forward2.py:69-101 hard-codes a 70x70 grid, dx=10 m, nbc=120 pad
cells, nt=999 steps of dt=1 ms, a 15 Hz Ricker wavelet, 5 shots at columns
[0, 17, 34, 52, 69] of physical row 1 and 70 receivers on every column of
physical row 1.  ``Geometry`` carries the same quantities per model so that
grids of other sizes (100x100, 70x140, 140x140) run through the same kernels.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def ricker(freq: float, dt: float, nt: int) -> np.ndarray:
    """Ricker wavelet zero-padded to nt samples (core/seis_forward2.py:20-45)."""
    nw = int(2.2 / freq / dt)
    nw = 2 * (nw // 2) + 1
    nc = nw // 2 + 1
    k = np.arange(1, nw + 1)
    beta = ((nc - k) * freq * dt * np.pi) ** 2
    w0 = (1.0 - 2.0 * beta) * np.exp(-beta)
    if nt < len(w0):
        raise ValueError("nt is smaller than the wavelet length")
    w = np.zeros(nt)
    w[:len(w0)] = w0
    return w


@dataclass(frozen=True)
class Geometry:
    """Grid, time axis and acquisition of one velocity model.

    Rows are depth (z), columns lateral (x).  ``src_z``/``rcv_z`` are the
    depths of shots and receivers in grid cells (z / dx).  As adjust_sr
    (core/seis_forward2.py:60-68) resolves them, a depth of 0 is moved down to
    physical row 1 (``src_row``/``rcv_row``), so the reference's z = dx and
    z = 0 both inject and record on row 1.
    """

    nz: int = 70
    nx: int = 70
    dx: float = 10.0
    nbc: int = 120
    nt: int = 999
    dt: float = 1e-3
    freq: float = 15.0
    n_shots: int = 5
    src_z: int = 1
    rcv_z: int = 1
    src_x: tuple = field(default=None)
    rcv_x: tuple = field(default=None)

    def __post_init__(self):
        if self.src_x is None:
            xs = np.round(np.linspace(0, self.nx - 1, self.n_shots)).astype(int)
            object.__setattr__(self, "src_x", tuple(int(v) for v in xs))
        if self.rcv_x is None:
            object.__setattr__(self, "rcv_x", tuple(range(self.nx)))
        object.__setattr__(self, "src_x", tuple(int(v) for v in self.src_x))
        object.__setattr__(self, "rcv_x", tuple(int(v) for v in self.rcv_x))
        object.__setattr__(self, "n_shots", len(self.src_x))
        r = np.asarray(self.rcv_x)
        if len(r) > 1 and not (np.all(np.diff(r) == r[1] - r[0]) and r[1] > r[0]):
            raise ValueError("receivers must sit at evenly spaced, increasing columns")

    @property
    def src_row(self) -> int:
        """Physical row of the shots (adjust_sr, core/seis_forward2.py:62,65-66)."""
        return self.src_z + (1 if self.src_z == 0 else 0)

    @property
    def rcv_row(self) -> int:
        """Physical row of the receivers (adjust_sr, core/seis_forward2.py:64,67)."""
        return self.rcv_z + (1 if self.rcv_z == 0 else 0)

    @property
    def NZ(self) -> int:
        return self.nz + 2 * self.nbc

    @property
    def NX(self) -> int:
        return self.nx + 2 * self.nbc

    @property
    def n_rcv(self) -> int:
        return len(self.rcv_x)

    @property
    def n_params(self) -> int:
        """Length of the parameter vector: velocities + min_vel."""
        return self.nz * self.nx + 1

    @property
    def n_data(self) -> int:
        """Length of the data vector: shots x time x receivers."""
        return self.n_shots * self.nt * self.n_rcv

    @property
    def padded_cells(self) -> int:
        return self.NZ * self.NX

    def wavelet(self) -> np.ndarray:
        return ricker(self.freq, self.dt, self.nt)

    def time_key(self):
        """Fields that every model of one plan must share."""
        return (self.nt, float(self.dt), self.nbc, float(self.freq))
