"""Run each propagator pass once on a batch plan, for counter collection.

Usage (under rocprofv3):
    python -m seismic_amd.kernel_probe [--batch 16] [--kind forward|tangent|born|adjoint|all]

Launches: one warm call (graph capture), then the measured call of each
requested pass, then a calibration copy of exactly 1 GiB (torch copy_) whose
source was pushed out of the 256 MB last-level cache first, so that its
FETCH_SIZE / WRITE_SIZE readings scale the counters to HBM bytes.  Prints
a JSON line with the launch geometry of each pass.
"""
from __future__ import annotations

import argparse
import json
import math

import numpy as np
import torch

from . import autotune, synth
from .geometry import Geometry
from .propagator import Propagator


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--kind", default="all")
    ap.add_argument("--nt", type=int, default=1000)
    args = ap.parse_args()
    g = Geometry(nt=args.nt)
    B = args.batch
    P = Propagator([g] * B, ckpt_every=0)
    autotune.apply(P)
    vec = torch.tensor(np.concatenate([synth.to_vec(synth.curve_fault_model(70, 70, i)) for i in range(B)]),
                       dtype=torch.float32, device="cuda")
    dd = P.forward(vec).clone()
    kinds = ["forward", "tangent", "born", "adjoint"] if args.kind == "all" else [args.kind]
    P.gradient(vec, dd)               # saved forward for the Born pass
    run = {"forward": lambda: P.forward(vec), "tangent": lambda: P.tangent(vec, vec),
           "born": lambda: P.tangent_saved(vec, vec, check=False), "adjoint": lambda: P.gradient(vec, dd)}
    for k in kinds:
        run[k]()                      # capture
    torch.cuda.synchronize()
    for k in kinds:
        run[k]()
    torch.cuda.synchronize()
    n = 256 * 2 ** 20
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    flush = torch.zeros(n, device="cuda")      # evicts x from the last-level cache
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del flush
    info = {"batch": B, "cells_per_step": B * g.n_shots * g.padded_cells, "calibration_copy_bytes": 4 * n}
    for k in ["forward", "tangent", "born", "adjoint"]:
        rx, rz, kk, st = P.get_config(k)
        info[k] = {"window": [rx, rz], "k": kk, "strips": st,
                   "launches_per_pass": math.ceil((g.nt + (1 if k != "adjoint" else 0)) / kk)}
    print(json.dumps(info))


if __name__ == "__main__":
    main()
