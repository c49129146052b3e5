"""Run each step-kernel pass of a Gauss-Newton pass once, for counter collection.

Usage (under rocprofv3):
    python -m seismic_amd.kernel_probe [--batch 16] [--fanout 50] [--kind all|<pass>]

Passes, with the tuned configs the inversion executor uses:
    linesearch_forward  forward of the line-search plan (batch x fanout models)
    saving_forward      forward that stores the q history (batch plan)
    born                tangent on the saved forward (batch plan)
    adjoint             reverse pass with imaging (batch plan)
    tangent             fused forward + tangent (batch plan; outside the GN loop)

Each pass runs twice (graph capture, then the measured call), then a
calibration copy of exactly 1 GiB (torch copy_) whose source was pushed out
of the 256 MB last-level cache first, so that its FETCH_SIZE / WRITE_SIZE
readings scale the counters to HBM bytes.  Prints a JSON line with the launch
geometry of each pass.
"""
from __future__ import annotations

import argparse
import json
import math

import numpy as np
import torch

from . import autotune, synth
from ._lib import CAP_FORWARD
from .geometry import Geometry
from .propagator import Propagator

PASSES = ("linesearch_forward", "saving_forward", "born", "adjoint", "tangent")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--fanout", type=int, default=50)
    ap.add_argument("--kind", default="all")
    ap.add_argument("--nt", type=int, default=1000)
    args = ap.parse_args()
    g = Geometry(nt=args.nt)
    B = args.batch
    P = Propagator([g] * B, ckpt_every=0)
    autotune.apply(P)
    Pf = Propagator([g] * (B * args.fanout), caps=CAP_FORWARD)
    autotune.apply(Pf)
    models = [synth.to_vec(synth.curve_fault_model(70, 70, i)) for i in range(B)]
    vec = torch.tensor(np.concatenate(models), dtype=torch.float32, device="cuda")
    vec_f = torch.tensor(np.concatenate([models[i % B] for i in range(Pf.n_models)]), dtype=torch.float32,
                         device="cuda")
    # reference data from the history-writing forward, so that every dispatch
    # of the plain forward kernel below belongs to the line-search pass
    dd = P.forward_save(vec)[0].clone()
    kinds = list(PASSES) if args.kind == "all" else [args.kind]
    for _ in range(2):                # capture, then the measured call
        for k in kinds:
            if k == "linesearch_forward":
                Pf.forward(vec_f)
            elif k == "saving_forward":
                P.forward_save(vec)
            elif k == "born":
                P.forward_save(vec)
                P.tangent_saved(vec, vec, check=False)
            elif k == "adjoint":
                _, tok = P.forward_save(vec)
                P.adjoint(vec, dd, tok, on_residual=True)
            elif k == "tangent":
                P.tangent(vec, vec)
            else:
                raise ValueError(k)
        torch.cuda.synchronize()
    n = 256 * 2 ** 20
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    flush = torch.zeros(n, device="cuda")      # evicts x from the last-level cache
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    del flush
    info = {"batch": B, "fanout_models": Pf.n_models, "calibration_copy_bytes": 4 * n}
    kind_of = {"linesearch_forward": (Pf, "forward"), "saving_forward": (P, "record"), "born": (P, "born"),
               "adjoint": (P, "adjoint"), "tangent": (P, "tangent")}
    for k in PASSES:
        plan, kind = kind_of[k]
        rx, rz, kk, st = plan.get_config(kind)
        steps = g.nt + (0 if k == "adjoint" else 1)
        info[k] = {"window": [rx, rz], "k": kk, "strips": st, "models": plan.n_models,
                   "launches_per_pass": math.ceil(steps / kk),
                   "cell_updates_per_launch": plan.n_models * g.n_shots * g.padded_cells * kk}
    print(json.dumps(info))


if __name__ == "__main__":
    main()
