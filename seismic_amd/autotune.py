"""Autotuner for the step kernels' launch configuration.

Each kernel kind (forward, forward recording the history, fused tangent,
adjoint, Born on the saved forward) runs an rx x rz cell window per workgroup, advances k time steps
per launch and deals window rows to wavefronts either as consecutive strips
or round-robin (libseis ``seis_plan_set_config``); every kind can also
march rx-column strips down rz-row segments (strips = 2; the tuner does not
try the streamed adjoint or saving forward).  The best choice depends on how many cells one
time step covers (batch size x model size: a single 70x70 model leaves most
of the GPU idle with large windows) and on whether the forward writes the full
history.  ``tune`` times every candidate on a live plan; ``apply`` looks a
plan's configs up in a table keyed by kind, log2 of the cells per step and
history mode, and sets them.  The default table, ``tuned_configs.json``,
holds results measured on MI355X; results do not depend on the config
(bit-identical wavefields), only speed does.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np
import torch

from . import synth
from ._lib import CAP_ADJOINT, CAP_TANGENT

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_configs.json")

_WINDOWS = {
    "forward": [(64, 32, 3), (128, 32, 4), (64, 64, 4), (128, 64, 4), (128, 64, 6), (128, 64, 8), (256, 32, 3),
                (128, 96, 6), (128, 96, 8), (128, 128, 6), (128, 128, 8)],
    "tangent": [(64, 32, 3), (64, 32, 4), (128, 32, 3), (128, 32, 4), (64, 64, 3), (64, 64, 4)],
    "adjoint": [(64, 32, 3), (64, 32, 4), (128, 32, 3), (128, 32, 4), (64, 64, 4), (64, 64, 5), (128, 64, 4),
                (128, 64, 5), (128, 64, 6), (128, 64, 7)],
    "born": [(64, 32, 3), (128, 32, 3), (64, 64, 3), (64, 64, 4), (128, 64, 4), (128, 64, 5), (128, 64, 6),
             (256, 32, 3)],
}
_WINDOWS["record"] = [(64, 32, 3), (128, 32, 4), (64, 64, 4), (128, 64, 4), (128, 64, 6), (128, 64, 8),
                      (256, 32, 3), (128, 96, 6), (128, 128, 6)]
# (rx, rz, k, strips): every window with both row ownerships
CANDIDATES = {kind: [w + (st,) for w in ws for st in (1, 0)] for kind, ws in _WINDOWS.items()}
# row-streamed forward (strips = 2): rx-column strips, rz-row segments (0:
# whole columns), k levels
CANDIDATES["forward"] += [(rx, rz, k, 2) for rx in (128, 192) for rz in (0, 40, 80, 160) for k in (3, 4, 5)]
CANDIDATES["born"] += [(rx, rz, k, 2) for rx in (64, 128) for rz in (0, 80, 160) for k in (3, 4)]
# row-streamed fused forward + tangent: 128- or 192-column strips
CANDIDATES["tangent"] += [(rx, rz, k, 2) for rx in (128, 192) for rz in (0, 80, 160) for k in (3, 4)]
# the row-streamed adjoint and saving forward (strips = 2) are not default
# candidates: timed alone on a full batch they win (adjoint by ~10 %, saving
# forward by ~20 %), but inside served and chained inversions, where the
# active models change from launch to launch, the tiled kernels keep those
# workloads faster, which a table keyed by the plan's size cannot see
STREAM_ADJOINT = [(rx, rz, k, 2) for rx in (64, 128) for rz in (0, 80, 160) for k in (4, 6)]
STREAM_RECORD = [(rx, rz, k, 2) for rx in (128, 192) for rz in (0, 80, 160) for k in (3, 4)]
# tuning order: the adjoint and Born passes run after the forward that
# records their history
KINDS = ("forward", "record", "tangent", "adjoint", "born")


def _has(P, kind: str) -> bool:
    if kind == "tangent":
        return bool(P.caps & CAP_TANGENT)
    if kind == "adjoint":
        return bool(P.caps & CAP_ADJOINT)
    if kind == "born":
        return bool(P.caps & CAP_TANGENT) and P.full_history
    if kind == "record":
        return bool(P.caps & CAP_ADJOINT)
    return True


def cells_per_step(P) -> int:
    return sum(g.n_shots * g.padded_cells for g in P.geoms)


def key(P, kind: str) -> str:
    mode = "full" if P.ckpt_every == 0 or P.ckpt_every >= P.geoms[0].nt else "ckpt"
    return f"{kind}:{int(math.log2(max(1, cells_per_step(P))))}:{mode}"


def load_table(path: str = TABLE) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def apply(P, table: dict | None = None) -> dict:
    """Set P's configs from the table (nearest bucket of the same kind and
    mode); returns what was applied."""
    table = load_table() if table is None else table
    applied = {}
    for kind in KINDS:
        if not _has(P, kind):
            continue
        k = key(P, kind)
        kk, b, mode = k.split(":")
        cands = [(abs(int(e.split(":")[1]) - int(b)), e) for e in table
                 if e.split(":")[0] == kk and e.split(":")[2] == mode]
        if not cands:
            continue
        # nearest bucket first; a config the plan cannot take (a streamed one
        # on padded grids under 4k rows) falls through to the next
        for _, e in sorted(cands):
            cfg = tuple(table[e])
            try:
                P.set_config(kind, *cfg)
            except ValueError:
                continue
            applied[kind] = cfg
            break
    return applied


def _timed(fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune(P, kinds=KINDS, reps: int = 2, verbose: bool = False) -> dict:
    """Time every candidate config of each kind on plan P (synthetic models of
    the plan's shapes) and leave P on the fastest.  Returns {key: (cfg, s)}."""
    vec = torch.tensor(np.concatenate([synth.to_vec(synth.make_model("curved", g.nz, g.nx, i % 7))
                                       for i, g in enumerate(P.geoms)]), dtype=torch.float32, device="cuda")
    dd = P.forward(vec).clone()
    # the forward kind runs plain forwards and, on checkpointed adjoint plans,
    # the saving forward (checkpoints only); the record kind runs the saving
    # forward of full-history plans and the segment recompute inside the
    # adjoint pass of checkpointed ones
    ckpt_adj = bool(P.caps & CAP_ADJOINT) and not P.full_history
    run = {"forward": (lambda: P.forward_save(vec)) if ckpt_adj else (lambda: P.forward(vec)),
           "record": (lambda: P.gradient(vec, dd)) if ckpt_adj else (lambda: P.forward_save(vec)),
           "tangent": lambda: P.tangent(vec, vec),
           "adjoint": lambda: P.gradient(vec, dd), "born": lambda: P.tangent_saved(vec, vec, check=False)}
    out = {}
    for kind in kinds:
        if not _has(P, kind):
            continue
        if kind == "born":
            P.gradient(vec, dd)   # the saved forward the Born pass reads
        res = {}
        for cfg in CANDIDATES[kind]:
            try:
                P.set_config(kind, *cfg)
            except ValueError:
                continue
            res[cfg] = _timed(run[kind], reps)
            if verbose:
                print(f"[tune] {key(P, kind)} {cfg}: {res[cfg] * 1e3:.2f} ms", flush=True)
        best = min(res, key=res.get)
        P.set_config(kind, *best)
        out[key(P, kind)] = (best, res[best])
    return out


def merge(table: dict, results: dict) -> dict:
    t = dict(table)
    for k, (cfg, _) in results.items():
        t[k] = list(cfg)
    return t


def main():
    """Tune the plans the inversion loops create and write the merged table."""
    import argparse

    from ._lib import CAP_FORWARD
    from .geometry import Geometry
    from .propagator import Propagator
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()
    g70 = Geometry(nt=1000)
    sizes = [Geometry(nz=n, nx=n, nt=1000) for n in (70, 100, 140)]
    work = [
        ("plain", [g70], 0, 7),
        ("batch16", [g70] * 16, 0, 7),
        ("batch32", [g70] * 32, 0, 7),
        ("batch16-ckpt", [g70] * 16, 50, 7),
        ("batch64-ckpt", [g70] * 64, 50, 7),
        ("serve-mix", [sizes[0]] * 12 + [sizes[1]] * 8 + [sizes[2]] * 4, 50, 7),
        ("linesearch-fanout", [g70] * 1600, 0, CAP_FORWARD),
        ("linesearch-plain", [g70], 0, CAP_FORWARD),
    ]
    table = load_table(args.out)
    for name, geoms, ck, caps in work:
        P = Propagator(geoms, ckpt_every=ck, caps=caps)
        res = tune(P, reps=args.reps, verbose=True)
        for k, (cfg, t) in res.items():
            print(f"[tune] {name:18s} {k:22s} best {cfg} {t * 1e3:.2f} ms", flush=True)
        table = merge(table, res)
        P.close()
        del P
        torch.cuda.empty_cache()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
