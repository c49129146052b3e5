"""Benchmark: batched Gauss-Newton inversions per second vs the plain path.

Workload (one MI355X): synthetic 70x70 models, curved layers with a dipping
fault (seeded), depth-linear start 1500 -> 4000 m/s, L2 smoothness prior,
5 shots, 70 receivers, nt = 1000 steps of 1 ms.  One inversion = 10
Gauss-Newton passes, each with 20 CG steps on J^T J (one Born pass on the
pass's saved forward + one adjoint pass per step) and the reference's
50-point line search.

A *step* is one Gauss-Newton pass over the whole batch.  ``value`` is the
batch's inversions per second (steps/10 inversions per model); the same
inversion run one model at a time (the plain path: same kernels, batch 1,
hipGraph time loops) is timed beside it, and a served LBFGS run with Poisson
arrivals reports p50/p99 latency.  ``roofline`` is the dominant kernel's
algorithmic HBM bytes per launch over its average launch time (HIP events on
the launch stream); ``cpu_baseline`` is the NumPy float64 oracle timed on a
bounded sample on the host.

Usage: python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12                     # MI355X HBM3E bytes/s
# L2 (TCC): 8 XCDs x 16 channels x 128 B/clk x 2.4 GHz (the CDNA3 L2
# organisation, kept by CDNA4; AMD publishes no MI355X L2 figure)
L2_PEAK = 8 * 16 * 128 * 2.4e9
# Algorithmic HBM bytes per cell-update (float32, one padded cell advanced one
# step for one shot): forward reads p[t-1], p[t], alpha, kappa and writes
# p[t+1]; the fused tangent also reads dp[t-1], dp[t], dalpha, dkappa and
# writes dp[t+1]; the Born pass reads dp[t-1], dp[t], alpha, kappa, dalpha,
# dkappa and the saved q[t+1] and writes dp[t+1]; the adjoint reads lam[t+1],
# lam[t+2], alpha, kappa and p[t], p[t+1] and writes lam[t] (the per-unit
# figure of the adjoint + imaging pass; the kernel reads one saved q level per
# step).  The saving forward also writes q[t+1] of every cell.
ALG_BYTES = {"linesearch_forward": 20, "saving_forward": 24, "tangent": 40, "born": 32, "adjoint": 28}
KERNEL = {"linesearch_forward": "k_fwd", "saving_forward": "k_fwd (history)", "tangent": "k_fwd (tangent)",
          "born": "k_born", "adjoint": "k_adj"}
STREAMED = {"linesearch_forward": "k_fws", "saving_forward": "k_fws (history)", "tangent": "k_fws (tangent)",
            "born": "k_bns", "adjoint": "k_ads"}
N_OUTER, N_CG, N_SCALES = 10, 20, 50


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_requests(B, g, seed0, device, lam):
    import torch
    from seismic_amd import invert, priors, synth
    from seismic_amd.propagator import Propagator
    trues = [synth.curve_fault_model(g.nz, g.nx, seed0 + i) for i in range(B)]
    P = Propagator([g] * B, caps=1, device=device)
    data = P.forward(torch.tensor(np.concatenate([synth.to_vec(t) for t in trues])))
    torch.cuda.synchronize()
    start = synth.to_vec(synth.depth_linear_start(g.nz, g.nx, 1500.0, 4000.0), min_vel=1500.0)
    reqs = []
    for i in range(B):
        tgt = data[int(P.seis_off[i]):int(P.seis_off[i + 1])].clone()
        reqs.append(invert.Request(geom=g, target=tgt, prior=priors.L2Smoothness(g.nz, g.nx, lam=lam, device=device),
                                   start=torch.tensor(start, dtype=torch.float64, device=device), rid=i))
    P.close()
    return reqs, trues


def _copy_request(r):
    from seismic_amd import invert
    return invert.Request(geom=r.geom, target=r.target, prior=r.prior, start=r.start.clone(), rid=r.rid)


def run_gn(executor, reqs, n_outer):
    """Run n_outer GN passes of every request through `executor`; returns (seconds, models)."""
    import torch
    from seismic_amd import invert
    for i, r in enumerate(reqs):
        executor.set_target(i, r.target)
        r.meta.clear()
    machines = {i: invert.gn_machine(r, n_outer=n_outer, n_cg=N_CG, n_scales=N_SCALES) for i, r in enumerate(reqs)}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if executor.max_batch == 1:
        res = {}
        for i in machines:
            res.update(invert.run_machines({i: machines[i]}, executor))
    else:
        res = invert.run_machines(machines, executor)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, [res[i] for i in range(len(reqs))]


def kernel_roofline(ex, B, g, reps=3):
    """Average launch time of each step kernel as the GN pass runs it: the
    line-search forward on the executor's fan-out plan (B x 50 models), the
    saving forward, Born and adjoint passes on the batch plan.  HIP events on
    torch's current stream, which is the stream the plans launch on."""
    import torch
    from seismic_amd import synth
    P = ex._grad_plan(list(range(B)))
    P.set_active([True] * B)
    P._mask = tuple([True] * B)
    Pf = max((v for k, v in ex._plans.items() if k[0] == "fB"), key=lambda q: q.n_models)
    models = [synth.to_vec(synth.curve_fault_model(g.nz, g.nx, i)) for i in range(B)]
    vec = torch.tensor(np.concatenate(models), dtype=torch.float32, device="cuda")
    vec_f = torch.tensor(np.concatenate([models[i % B] for i in range(Pf.n_models)]), dtype=torch.float32,
                         device="cuda")
    dd = P.forward(vec).clone()
    s = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    t_ls = timed(lambda: Pf.forward(vec_f))
    t_save = timed(lambda: P.forward_save(vec))
    t_tan = timed(lambda: P.tangent(vec, vec))
    _, tok = P.forward_save(vec)
    t_adj = timed(lambda: P.adjoint(vec, dd, tok, on_residual=True))
    t_born = timed(lambda: P.tangent_saved(vec, vec, check=False))
    out = {}
    for name, plan, kind, t_pass, steps in [("linesearch_forward", Pf, "forward", t_ls, g.nt + 1),
                                            ("saving_forward", P, "record", t_save, g.nt + 1),
                                            ("tangent", P, "tangent", t_tan, g.nt + 1),
                                            ("born", P, "born", t_born, g.nt + 1),
                                            ("adjoint", P, "adjoint", t_adj, g.nt)]:
        rx, rz, k, strips = plan.get_config(kind)
        launches = math.ceil(steps / k)
        cells = plan.n_models * g.n_shots * g.padded_cells
        kname = STREAMED[name] if strips == 2 else KERNEL[name]
        out[name] = {"kernel": kname, "pass_s": t_pass, "launches": launches, "launch_s": t_pass / launches,
                     "cell_updates_per_launch": cells * k, "k": k, "window": [rx, rz], "strips": strips,
                     "models": plan.n_models}
    return out


def per_step_configs(P):
    """Every kernel kind of plan P at one time step per launch (64x32
    windows, round-robin rows): the reference's launch pattern, one kernel
    per step, on this library's kernels."""
    for kind in ("forward", "record", "tangent", "adjoint", "born"):
        try:
            P.set_config(kind, 64, 32, 1, 0)
        except (ValueError, RuntimeError):
            pass                          # a kind this plan does not run


def measured_traffic(name, kd, table="hbm_traffic.json", field="bytes_per_launch"):
    """HBM (or, with table="l2_traffic.json", L2) bytes per launch of this
    kernel and config from the rocprofv3 summaries in profiles/, or None
    when the committed measurement was taken on another config."""
    path = os.path.join(ROOT, "profiles", table)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tab = json.load(f)
    e = tab.get(name)
    if not e or e.get("window") != kd["window"] or e.get("k") != kd["k"] or e.get("strips") != kd["strips"] \
            or e.get("models") != kd["models"]:
        return None
    return e[field] if field in e else None


def hbm_info(traffic, kd):
    """Counter-measured HBM bytes per launch (profiles/, rocprofv3 FETCH_SIZE /
    WRITE_SIZE) over this run's launch time, against the 8 TB/s peak: the
    fraction of HBM bandwidth the kernel really uses, beside the algorithmic
    roofline fraction (temporal blocking moves fewer bytes than the
    per-step algorithmic count)."""
    if traffic is None:
        return None
    achieved = traffic / kd["launch_s"]
    return {"bytes_per_launch": traffic, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK}


def l2_info(name, kd):
    """L2 bytes per launch (rocprofv3 TCC_READ_sum / TCC_WRITE_sum, profiles/)
    over this run's launch time, against L2_PEAK; None on another config."""
    b = measured_traffic(name, kd, "l2_traffic.json", "l2_bytes_per_launch")
    if b is None:
        return None
    hit = measured_traffic(name, kd, "l2_traffic.json", "l2_hit_rate")
    achieved = b / kd["launch_s"]
    return {"traffic": b, "achieved": achieved / 1e9, "peak": L2_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / L2_PEAK, "hit_rate": hit}


def cpu_baseline(g, lbfgs_iters=30):
    """The FP64 NumPy oracle (the reference algorithm restated) on the host
    cores, on bounded samples of the GPU workloads: one forward, one fev
    (forward + adjoint gradient), one Gauss-Newton product (the oracle's
    tangent then adjoint) and one FP32 forward, for as many 70x70 models at
    once as fill the cores at one process per shot; GN inversions/s from those
    measured pass times and the workload's pass counts.  Also one whole
    30-iteration LBFGS inversion (memory 10, L2 smoothness prior, the
    serving workload's), measured."""
    from concurrent.futures import ProcessPoolExecutor
    from oracle import fd2d
    from oracle import invert as oinv
    from oracle import priors as opri
    from seismic_amd import synth
    og = fd2d.Geometry(nz=g.nz, nx=g.nx, nbc=g.nbc, nt=g.nt, dt=g.dt, dx=g.dx, freq=g.freq)
    n = og.n_shots
    cores = len(os.sched_getaffinity(0))
    n_models = max(1, cores // n)
    workers = min(cores, n * n_models)
    vecs = [synth.to_vec(synth.curve_fault_model(g.nz, g.nx, m)) for m in range(n_models)]
    target = fd2d.forward_only(synth.to_vec(synth.depth_linear_start(g.nz, g.nx)), og).reshape(-1)
    dv = np.concatenate([np.random.default_rng(1).normal(0, 20, g.nz * g.nx), [3.0]])
    units = [(vecs[m], og, s) for m in range(n_models) for s in range(n)]

    def timed(pool, fn, args):
        t0 = time.perf_counter()
        list(pool.map(fn, args))
        return time.perf_counter() - t0

    with ProcessPoolExecutor(workers) as ex:
        list(ex.map(_cpu_fwd, units[:workers]))            # start the workers
        t_fwd = timed(ex, _cpu_fwd, units)
        t_fev = timed(ex, _cpu_fev, [u + (target,) for u in units])
        t_gnp = timed(ex, _cpu_gnp, [u + (dv,) for u in units])
        t_f32 = timed(ex, _cpu_fwd32, units)
        # one LBFGS inversion: its fevs run one process per shot
        true = synth.curve_fault_model(g.nz, g.nx, 100)
        data = np.concatenate(list(ex.map(_cpu_fwd, [(synth.to_vec(true), og, s) for s in range(n)]))).reshape(-1)
        start = synth.to_vec(synth.depth_linear_start(g.nz, g.nx), min_vel=1500.0)

        def v2s(vec, vec_adjoint=None, adjoint_on_residual=False):
            outs = list(ex.map(_cpu_shot, [(vec, og, s, vec_adjoint, adjoint_on_residual) for s in range(n)]))
            seis = np.concatenate([o[0] for o in outs])
            return seis, None, (None if vec_adjoint is None else np.sum([o[2] for o in outs], axis=0))

        t0 = time.perf_counter()
        _, nfev = oinv.lbfgs_invert(start, data, opri.L2Smoothness(g.nz, g.nx, lam=1e-8), og, lbfgs_iters,
                                    history_size=10, v2s=v2s)
        t_lbfgs = time.perf_counter() - t0
    t_inv = N_OUTER * (t_fev + N_CG * t_gnp + N_SCALES * t_fwd) / n_models
    cells = n * (g.nt + 1) * og.NZ * og.NX * n_models
    t0 = time.perf_counter()
    from oracle import priors as opri2
    opri2.label_thresholded_components(_flat_grid(0), 1.0)
    t_label = time.perf_counter() - t0
    return {"value": 1.0 / t_inv, "unit": "inversions/s", "cores": workers, "kind": "port",
            "s_per_forward": t_fwd / n_models, "s_per_fev": t_fev / n_models, "s_per_gn_product": t_gnp / n_models,
            "s_per_forward_fp32": t_f32 / n_models, "fevs_per_s": n_models / t_fev,
            "forward_cell_updates_per_s": cells / t_fwd, "forward_fp32_cell_updates_per_s": cells / t_f32,
            "lbfgs_inversion": {"iterations": lbfgs_iters, "history": 10, "fevs": nfev, "seconds": t_lbfgs,
                                "inversions_per_s": 1.0 / t_lbfgs, "processes": n},
            "label_flat_s_per_grid": t_label,
            "sample": f"NumPy oracle, {g.nz}x{g.nx} nt={g.nt}, {n_models} models x {n} shots in {workers} "
                      f"processes on {cores} host cores: per model a forward {t_fwd / n_models:.2f} s, "
                      f"a forward+adjoint gradient {t_fev / n_models:.2f} s, a GN product (tangent then adjoint) "
                      f"{t_gnp / n_models:.2f} s, all measured (float64; float32 forward "
                      f"{t_f32 / n_models:.2f} s); value = measured pass times x {N_OUTER} GN passes x "
                      f"(1 fev + {N_CG} products + {N_SCALES} line-search forwards); one measured "
                      f"{lbfgs_iters}-iteration LBFGS inversion: {t_lbfgs:.1f} s ({nfev} fevs)"}


def _cpu_fwd(a):
    from oracle import fd2d
    vec, og, s = a
    return fd2d.forward_only(vec, og, shots=[s])


def _cpu_fwd32(a):
    from oracle import fd2d
    vec, og, s = a
    return fd2d.forward_only(vec, og, dtype=np.float32, shots=[s])


def _cpu_fev(a):
    from oracle import fd2d
    vec, og, s, target = a
    tgt = target.reshape(og.n_shots, og.nt, og.n_rcv)[s].reshape(-1)
    return fd2d.vel_to_seis(vec, og, vec_adjoint=tgt, adjoint_on_residual=True, shots=[s])[2]


def _cpu_gnp(a):
    """One shot's share of the oracle's Gauss-Newton product J^T J dv."""
    from oracle import fd2d
    vec, og, s, dv = a
    jdv = fd2d.vel_to_seis(vec, og, vec_diff=dv, shots=[s])[1]
    return fd2d.vel_to_seis(vec, og, vec_adjoint=jdv, shots=[s])[2]


def _cpu_shot(a):
    from oracle import fd2d
    vec, og, s, dd, resid = a
    if dd is not None:
        dd = dd.reshape(og.n_shots, og.nt, og.n_rcv)[s].reshape(-1)
    return fd2d.vel_to_seis(vec, og, vec_adjoint=dd, adjoint_on_residual=resid, shots=[s])


def serving_run(n_req, rate, device, seed=100, sweep=(50.0, 200.0, 400.0), sweep_requests=36, ckpt_every=None):
    """Served LBFGS inversions (30 iterations, memory 10, smoothness prior) of
    mixed grid sizes arriving as a Poisson process: `n_req` requests at
    `rate` on as many slots, then a rate sweep with `sweep_requests` requests
    on the same slots (requests queue for a free slot of their size).  The
    server keeps full forward histories when they fit its 200 GiB budget
    (ckpt_every=None), else checkpoints every 50 steps."""
    import torch
    from seismic_amd import invert, priors, serve, synth
    from seismic_amd.geometry import Geometry
    from seismic_amd.propagator import Propagator
    sizes = [(70, 70), (100, 100), (140, 140)]

    def size_mix(n):
        return [0] * (n // 2) + [1] * (n // 3) + [2] * (n - n // 2 - n // 3)

    n_max = max(n_req, sweep_requests)
    mix_all = size_mix(n_max)
    geoms = {sz: Geometry(nz=sz[0], nx=sz[1], nt=1000) for sz in sizes}
    trues = [synth.curve_fault_model(*sizes[si], seed + i) for i, si in enumerate(mix_all)]
    targets = [None] * n_max
    for si, sz in enumerate(sizes):
        ids = [i for i, k in enumerate(mix_all) if k == si]
        if not ids:
            continue
        Pf = Propagator([geoms[sz]] * len(ids), caps=1)
        d = Pf.forward(torch.tensor(np.concatenate([synth.to_vec(trues[i]) for i in ids])))
        for j, i in enumerate(ids):
            targets[i] = d[int(Pf.seis_off[j]):int(Pf.seis_off[j + 1])].clone()
        Pf.close()
    by_size = {si: [i for i, k in enumerate(mix_all) if k == si] for si in range(3)}

    def make_reqs(n):
        out = []
        taken = {si: 0 for si in range(3)}
        for rid, si in enumerate(size_mix(n)):
            i = by_size[si][taken[si]]
            taken[si] += 1
            g = geoms[sizes[si]]
            start = synth.to_vec(synth.depth_linear_start(g.nz, g.nx), min_vel=1500.0)
            out.append(invert.Request(geom=g, target=targets[i], prior=priors.L2Smoothness(g.nz, g.nx, lam=1e-8),
                                      start=torch.tensor(start, device=device), rid=rid))
        return out

    slots = [geoms[sizes[si]] for si in size_mix(n_req)]
    srv = serve.InversionServer(slots, hbm_budget=200 * 2 ** 30, ckpt_every=ckpt_every, max_iter=30, history_size=10)
    srv.kw["max_iter"] = 1                      # capture the server's graphs before timing
    srv.run(make_reqs(n_req), [0.0] * n_req)
    srv.kw["max_iter"] = 30

    def timed_run(n, r, sd):
        srv.log.clear()
        reqs = make_reqs(n)
        arrivals = serve.poisson_arrivals(n, r, sd)
        srv.start_timing()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        recs = srv.run(reqs, arrivals)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        busy = srv.gpu_busy_s()
        return {"requests": n, "rate_req_per_s": r, "inversions_per_s": n / wall, "wall_s": wall,
                "gpu_busy_s": busy, "host_time_share": max(0.0, 1.0 - busy / wall),
                "ckpt_every": srv.ckpt_every,
                "latency_s": serve.percentiles([x.latency for x in recs]),
                "queue_wait_s": serve.percentiles([x.admitted - x.arrival for x in recs]),
                "max_in_flight": max(k for _, k in srv.log) if srv.log else 0,
                "failed": int(sum(x.failed for x in recs)), "records": recs}, reqs

    main_run, reqs = timed_run(n_req, rate, seed)
    served_results = [rec.result for rec in main_run.pop("records")]
    sweep_runs = [timed_run(sweep_requests, r, seed + 7)[0] for r in sweep]
    for sw in sweep_runs:
        sw.pop("records")
    srv.ex.close()
    # plain path, like for like: every request of the same mix inverted
    # alone, one after the other (same kernels, one model per plan)
    ex = invert.Executor([geoms[sz] for sz in sizes], ckpt_every=0, max_batch=1)
    for j, sz in enumerate(sizes):          # capture each size's graphs before timing
        r = make_reqs(n_req)[[si for si in size_mix(n_req)].index(j)]
        ex.set_target(j, r.target)
        invert.run_machines({j: invert.lbfgs_machine(r, max_iter=1, history_size=10)}, ex)
    plain_reqs = make_reqs(n_req)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_results = []
    for r in plain_reqs:
        j = sizes.index((r.geom.nz, r.geom.nx))
        ex.set_target(j, r.target)
        plain_results.append(invert.run_machines({j: invert.lbfgs_machine(r, max_iter=30, history_size=10)},
                                                 ex)[j])
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    ex.close()
    same = all(torch.equal(a, b) for a, b in zip(served_results, plain_results))
    out = dict(main_run)
    out.update({"slots": n_req, "sizes": [f"{a}x{b}" for a, b in sizes],
                "plain_same_mix": {"inversions_per_s": n_req / t_plain, "seconds": t_plain,
                                   "results_equal_served": same},
                "speedup_vs_plain_same_mix": out["inversions_per_s"] / (n_req / t_plain),
                "rate_sweep": sweep_runs})
    return out


def _timed_s(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def mixed_batch_run(device, n_models=512, ckpt=32, seed=300):
    """Many models of uneven size in one plan: 512 models of 70x70, 100x100 and
    70(z)x140(x) (flat, curved, faulted, GRF-textured), 5 shots, n_rcv = nx,
    nt = 1000.  Every launch advances all 2,560 wavefields.  Times one forward
    pass and one fev (misfit + gradient, checkpoints every `ckpt` steps)."""
    import torch
    from seismic_amd import autotune, synth
    from seismic_amd._lib import CAP_ADJOINT, CAP_FORWARD
    from seismic_amd.geometry import Geometry
    from seismic_amd.propagator import Propagator
    sizes = [(70, 70), (100, 100), (70, 140)]
    styles = ["flat", "curved", "faulted", "grf"]
    geoms = {sz: Geometry(nz=sz[0], nx=sz[1], nt=1000) for sz in sizes}
    gl = [geoms[sizes[i % 3]] for i in range(n_models)]
    vec = np.concatenate([synth.to_vec(synth.make_model(styles[(i // 3) % 4], g.nz, g.nx, seed + i))
                          for i, g in enumerate(gl)])
    cells = sum(g.n_shots * g.padded_cells for g in gl)
    v = torch.tensor(vec, dtype=torch.float32, device=device)
    Pf = Propagator(gl, caps=CAP_FORWARD, device=device)
    autotune.apply(Pf)
    dd = Pf.forward(v * 1.01).clone()
    t_fwd = _timed_s(lambda: Pf.forward(v), 2)
    Pf.close()
    Pg = Propagator(gl, ckpt_every=ckpt, caps=CAP_FORWARD | CAP_ADJOINT, device=device)
    autotune.apply(Pg)
    t_fev = _timed_s(lambda: Pg.gradient(v, dd), 1)
    gbytes = Pg.device_bytes
    Pg.close()
    torch.cuda.empty_cache()
    steps = 1001
    return {"models": n_models, "sizes": [f"{a}x{b}" for a, b in sizes], "wavefields": n_models * 5,
            "cells_per_step": cells, "forward_s": t_fwd,
            "forward_gcell_updates_per_s": cells * steps / t_fwd / 1e9,
            "fev_s": t_fev, "fevs_per_s": n_models / t_fev, "fev_ckpt_every": ckpt,
            "fev_plan_gb": gbytes / 1e9}


def autograd_run(device, n_models=64, ckpt=50, reps=3):
    """torch.autograd through the propagator: 64 70x70 models, checkpoints
    every 50 steps, loss = data misfit + small-CNN prior; one step = forward,
    backward (segment recompute + adjoint) and the prior."""
    import torch
    from seismic_amd import autotune, synth
    from seismic_amd.autograd import SmallCNNPrior, propagate
    from seismic_amd.geometry import Geometry
    from seismic_amd.propagator import Propagator
    g = Geometry(nt=1000)
    vel = np.stack([synth.to_vec(synth.curve_fault_model(70, 70, 500 + i)) for i in range(n_models)])
    P = Propagator([g] * n_models, ckpt_every=ckpt, device=device)
    autotune.apply(P)
    obs = P.forward(torch.tensor(vel * 1.01).reshape(-1)).reshape(n_models, -1).clone()
    prior = SmallCNNPrior(g.nz, g.nx).to(device)
    base = torch.tensor(vel, dtype=torch.float32, device=device)

    def step():
        v = base.clone().requires_grad_(True)
        loss = ((propagate(v, P) - obs) ** 2).mean(dim=1).sum() + 0.5 * prior(v).sum()
        loss.backward()

    t = _timed_s(step, reps)
    gbytes = P.device_bytes
    P.close()
    return {"models": n_models, "ckpt_every": ckpt, "step_s": t, "model_gradients_per_s": n_models / t,
            "plan_gb": gbytes / 1e9}


def flat_areas_run(device, n_models=32, reps=5):
    """Flat-area labelling of RestrictFlatAreas (label_thresholded_components,
    core/seis_numerics.py:81-124) for a batch of 70x70 guesses in one launch
    (the host baseline of the same labelling is in cpu_baseline)."""
    import torch
    from seismic_amd import priors
    grids = [torch.tensor(_flat_grid(i), device=device) for i in range(n_models)]
    t = _timed_s(lambda: priors.label_flat(grids, 1.0), reps)
    _, counts = priors.label_flat(grids, 1.0)
    return {"models": n_models, "grid": "70x70", "gpu_s": t, "gpu_grids_per_s": n_models / t,
            "mean_areas": float(np.mean(counts))}


def schedule_run(device, n_each=8):
    """Classified, chained inversions (core/seis_model.py, test-mode stage
    lengths): n_each laterally invariant layered models (FlatVel) and n_each
    faulted ones (TV2D: two TV stages and one on the flat areas), all chains
    batched over one executor."""
    import torch
    from seismic_amd import invert, schedule, synth
    from seismic_amd.geometry import Geometry
    from seismic_amd.propagator import Propagator
    g = Geometry(nt=1000)
    trues = [synth.make_model("flat", 70, 70, 900 + i) for i in range(n_each)] + \
            [synth.curve_fault_model(70, 70, 950 + i) for i in range(n_each)]
    P = Propagator([g] * len(trues), caps=1, device=device)
    obs = P.forward(torch.tensor(np.concatenate([synth.to_vec(t) for t in trues]), device=device)).clone()
    P.close()
    reqs = []
    for i, t in enumerate(trues):
        start = synth.to_vec(np.round(t * 0.97), min_vel=1500.0)
        reqs.append(invert.Request(geom=g, target=obs[i * g.n_data:(i + 1) * g.n_data].clone(), prior=None,
                                   start=torch.tensor(start, dtype=torch.float64, device=device), rid=i))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out, fams = schedule.infer(reqs, test_mode=True, batched=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fevs = [sum(st.get("fev", 0) for st in r.meta.get("stages", [])) for r in reqs]
    mae = [float(np.abs(o.cpu().numpy()[:-1] - t.ravel()).mean()) for o, t in zip(out, trues)]
    return {"requests": len(reqs), "families": {f: fams.count(f) for f in sorted(set(fams))}, "wall_s": dt,
            "inversions_per_s": len(reqs) / dt, "fevs_per_request_mean": float(np.mean(fevs)),
            "mae_to_true_m_per_s": {f: float(np.mean([m for m, ff in zip(mae, fams) if ff == f]))
                                    for f in sorted(set(fams))}, "stage_lengths": "test mode"}


def served_schedule_run(device, n_each=8, slots=12, rate=50.0):
    """The classified chains of schedule_run served with continuous batching:
    2 n_each requests arriving as a Poisson process at `rate` on `slots`
    70x70 slots (later arrivals wait for a free slot; full forward history,
    which fits), each classified on admission and run through its family's
    chain (test-mode stage lengths)."""
    import torch
    from seismic_amd import invert, schedule, serve, synth
    from seismic_amd.geometry import Geometry
    from seismic_amd.propagator import Propagator
    g = Geometry(nt=1000)
    trues = [synth.make_model("flat", 70, 70, 1300 + i) for i in range(n_each)] + \
            [synth.curve_fault_model(70, 70, 1350 + i) for i in range(n_each)]
    order = np.random.default_rng(5).permutation(len(trues))
    trues = [trues[i] for i in order]
    P = Propagator([g] * len(trues), caps=1, device=device)
    obs = P.forward(torch.tensor(np.concatenate([synth.to_vec(t) for t in trues]), device=device)).clone()
    P.close()
    reqs = [invert.Request(geom=g, target=obs[i * g.n_data:(i + 1) * g.n_data].clone(), prior=None,
                           start=torch.tensor(synth.to_vec(np.round(t * 0.97), min_vel=1500.0),
                                              dtype=torch.float64, device=device), rid=i)
            for i, t in enumerate(trues)]
    arrivals = serve.poisson_arrivals(len(reqs), rate, seed=7)
    srv = serve.InversionServer([g] * slots, hbm_budget=200 << 30, ckpt_every=None, history_size=100,
                                machine=schedule.serving_machine(test_mode=True), device=device)
    srv.start_timing()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = srv.run(reqs, arrivals)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    busy = srv.gpu_busy_s()
    fams = [r.stats.get("family") for r in recs]
    lat = [r.latency for r in recs]
    return {"requests": len(reqs), "slots": slots, "rate_per_s": rate, "wall_s": wall,
            "gpu_busy_s": busy, "host_time_share": max(0.0, 1.0 - busy / wall),
            "inversions_per_s": len(reqs) / wall, "latency_s": serve.percentiles(lat),
            "families": {f: fams.count(f) for f in sorted(set(fams))},
            "max_requests_per_launch": max(n for _, n in srv.log), "failed": sum(r.failed for r in recs),
            "stage_lengths": "test mode"}


def formats_run(device, n_files=4, per_file=32, batch=32, reps=3):
    """Competition-shaped files (core/kaggle_support.py:192-345, 554-587):
    n_files x per_file training samples ((N, 5, 1000, 70) float32 seismograms
    from the GPU forward of layered models, (N, 1, 70, 70) velocities)
    written to a temporary directory, then streamed to HBM by the loader
    (page-cached reads: the files were just written) and written out as a
    submission file."""
    import shutil
    import tempfile
    import torch
    from seismic_amd import datasets, synth
    from seismic_amd.geometry import Geometry
    from seismic_amd.propagator import Propagator
    g = Geometry(nt=1000)
    n = n_files * per_file
    trues = np.stack([synth.make_model("flat", 70, 70, 1200 + i) for i in range(per_file)]).astype(np.float32)
    P = Propagator([g] * per_file, caps=1, device=device)
    seis = P.forward(torch.tensor(np.concatenate([synth.to_vec(t) for t in trues]), device=device))
    seis = seis.reshape(per_file, 5, 1000, 70).cpu().numpy()
    P.close()
    root = tempfile.mkdtemp(prefix="seis_formats_")
    try:
        for k in range(n_files):
            d = os.path.join(root, "train_samples", "FlatVel_A")
            os.makedirs(d, exist_ok=True)
            np.save(os.path.join(d, f"seis_{k}.npy"), seis)
            np.save(os.path.join(d, f"vel_{k}.npy"), trues[:, None])
        samples = datasets.load_train(root)
        assert len(samples) == n
        best = float("inf")
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for b in datasets.Loader(samples, batch=batch, nt=999, device=device):
                pass
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        nbytes = n * (5 * 999 * 70 * 4 + 4901 * 8)
        grids = [torch.tensor(trues[i % per_file] + 0.37, device=device) for i in range(n)]
        sub = os.path.join(root, "submission.csv")
        t0 = time.perf_counter()
        datasets.write_submission(samples, grids, sub)
        t_sub = time.perf_counter() - t0
        rows = n * 70
    finally:
        shutil.rmtree(root, ignore_errors=True)
    return {"samples": n, "batch": batch, "loader_s": best, "loader_samples_per_s": n / best,
            "loader_gb_per_s": nbytes / best / 1e9, "submission_rows": rows, "submission_s": t_sub,
            "submission_rows_per_s": rows / t_sub, "reads": "page-cached"}


def _flat_grid(i):
    from seismic_amd import synth
    return np.round(synth.curve_fault_model(70, 70, 700 + i) * 0.98)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="GN passes timed (10 = one full inversion)")
    ap.add_argument("--warmup", type=int, default=1, help="GN passes run untimed first")
    ap.add_argument("--batch", type=int, default=32, help="models per batched launch")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--lam", type=float, default=1e-8)
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=100.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the 512-model and autograd legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); inversions are independent, "
              "so N GPUs would run N copies of this workload", file=sys.stderr)
        sys.exit(2)
    import torch
    from seismic_amd import invert
    from seismic_amd.geometry import Geometry
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = "cuda"
    g = Geometry(nt=1000)
    B = args.batch
    log(f"[bench] building {B} requests")
    reqs, trues = make_requests(B, g, 0, dev, args.lam)

    # ---- batched path
    ex_b = invert.Executor([g] * B, ckpt_every=0)
    log(f"[bench] warmup {args.warmup} GN pass(es), batch {B}")
    run_gn(ex_b, reqs, args.warmup)
    log(f"[bench] timing {args.steps} GN pass(es), batch {B}")
    torch.cuda.synchronize()
    t_b, v_b = run_gn(ex_b, reqs, args.steps)
    value = B * (args.steps / N_OUTER) / t_b
    mis_b = [r.meta["stats"]["misfit"] for r in reqs]
    hist_b0 = list(reqs[0].meta["stats"]["misfit_hist"])
    log(f"[bench] batched: {t_b:.2f} s, {value:.4f} inversions/s")

    # ---- roofline of the step kernels as the GN pass runs them
    kr = kernel_roofline(ex_b, B, g)
    # passes per GN iteration: the 50-point line search (one pass of the
    # fan-out plan), 1 saving forward, 20 Born passes (J^T J products on the
    # saved forward) and 21 adjoints
    per_outer = {"linesearch_forward": 1, "saving_forward": 1, "tangent": 0, "born": N_CG, "adjoint": 1 + N_CG}
    share = {k: per_outer[k] * kr[k]["pass_s"] for k in kr}
    dom = max(share, key=share.get)
    kd = kr[dom]
    achieved = ALG_BYTES[dom] * kd["cell_updates_per_launch"] / kd["launch_s"]
    traffic = measured_traffic(dom, kd)
    log(f"[bench] kernels: " + ", ".join(f"{k} {v['launch_s']*1e6:.1f} us/launch" for k, v in kr.items()))

    # ---- plain path: same kernels, one model at a time (its own copy of
    # request 0, so the batched run's misfit history stays intact)
    ex_p = invert.Executor([g], ckpt_every=0, max_batch=1)
    preq = [_copy_request(reqs[0])]
    run_gn(ex_p, preq, 1)
    log(f"[bench] timing plain path, {args.plain_steps} GN pass(es)")
    t_p, v_p = run_gn(ex_p, preq, args.plain_steps)
    hist_p0 = list(preq[0].meta["stats"]["misfit_hist"])
    plain = 1 * (args.plain_steps / N_OUTER) / t_p
    log(f"[bench] plain: {t_p:.2f} s, {plain:.4f} inversions/s")
    ex_p.close()
    # the same plain inversion shaped like the reference's loop: one time step
    # per launch (k = 1, 64x32 windows, no temporal blocking) inside the
    # captured graphs
    ex_s = invert.Executor([g], ckpt_every=0, max_batch=1, configure=per_step_configs)
    sreq = [_copy_request(reqs[0])]
    run_gn(ex_s, sreq, 1)
    log("[bench] timing plain path with one step per launch, 1 GN pass")
    t_s, _ = run_gn(ex_s, sreq, 1)
    hist_s0 = list(sreq[0].meta["stats"]["misfit_hist"])
    plain_step = (1 / N_OUTER) / t_s
    log(f"[bench] plain, one step per launch: {t_s:.2f} s per GN pass, {plain_step:.4f} inversions/s")
    ex_s.close()

    ex_b.close()
    torch.cuda.empty_cache()
    extras = {}
    if not args.no_extras:
        log("[bench] 512 mixed-size models in one plan")
        extras["mixed_batch_512"] = mixed_batch_run(dev)
        log(f"[bench] mixed 512: forward {extras['mixed_batch_512']['forward_s']:.2f} s, "
            f"fev {extras['mixed_batch_512']['fev_s']:.2f} s")
        log("[bench] autograd, 64 models")
        extras["autograd_64"] = autograd_run(dev)
        log(f"[bench] autograd: {extras['autograd_64']['step_s']:.2f} s per step")
        log("[bench] flat-area labelling, 32 grids")
        extras["flat_areas_32"] = flat_areas_run(dev)
        log(f"[bench] flat areas: {extras['flat_areas_32']['gpu_s'] * 1e3:.3f} ms per batch of 32")
        log("[bench] classified chained inversions, 16 requests")
        extras["schedules_16"] = schedule_run(dev)
        log(f"[bench] schedules: {extras['schedules_16']['inversions_per_s']:.2f} inversions/s "
            f"{extras['schedules_16']['families']}")
        log("[bench] served classified chains, 16 requests on 12 slots")
        extras["served_schedules_16"] = served_schedule_run(dev)
        log(f"[bench] served chains: {extras['served_schedules_16']['inversions_per_s']:.2f} inversions/s, "
            f"p50 {extras['served_schedules_16']['latency_s']['p50']:.2f} s, "
            f"p99 {extras['served_schedules_16']['latency_s']['p99']:.2f} s")
        log("[bench] on-disk formats, 128 samples")
        extras["formats_128"] = formats_run(dev)
        log(f"[bench] formats: loader {extras['formats_128']['loader_gb_per_s']:.2f} GB/s, submission "
            f"{extras['formats_128']['submission_rows_per_s']:.0f} rows/s")

    serving = None
    if not args.no_serve:
        log(f"[bench] serving {args.serve_requests} LBFGS requests at {args.serve_rate}/s")
        serving = serving_run(args.serve_requests, args.serve_rate, dev)
        log(f"[bench] serving: {serving['inversions_per_s']:.3f} inv/s p50 {serving['latency_s']['p50']:.2f} s")
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(g)

    out = {
        "metric": "inversions/sec batched+served vs plain at fixed misfit; p50/p99 latency",
        "value": value,
        "unit": "inversions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_b / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded curved-fault velocity models; data from the GPU forward of the true model)",
        "config": {"workload": "Gauss-Newton inversion, 70x70 dx=10 m, nbc=120, 5 shots, 70 receivers, "
                               "nt=1000 dt=1 ms, 10 GN x 20 CG + 50-point line search, L2 smoothness prior",
                   "batch_models": B, "step": "one GN pass over the batch", "ckpt_every": 0},
        "plain": {"inversions_per_s": plain, "seconds_per_gn_pass": t_p / args.plain_steps},
        "speedup_vs_plain": value / plain,
        "plain_one_step_per_launch": {"inversions_per_s": plain_step, "seconds_per_gn_pass": t_s,
                                      "speedup": value / plain_step,
                                      "configs": "k=1, 64x32 windows, every kernel kind"},
        "misfit": {"batched_mean_final": float(np.mean(mis_b)), "after_gn_passes": args.steps,
                   "model0_batched": hist_b0,
                   "model0_plain": hist_p0,
                   "model0_plain_one_step_per_launch": hist_s0,
                   "equal": hist_b0[:len(hist_p0)] == hist_p0 and hist_b0[:len(hist_s0)] == hist_s0,
                   "passes_compared": len(hist_p0)},
        "serving": serving,
        "extras": extras,
        "roofline": {"bound": "hbm", "kernel": kd["kernel"], "pass": dom,
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "launch_us": kd["launch_s"] * 1e6, "alg_bytes_per_cell_update": ALG_BYTES[dom],
                     "cell_updates_per_launch": kd["cell_updates_per_launch"], "k": kd["k"],
                     "window": kd["window"], "strips": kd["strips"], "models": kd["models"],
                     "hbm_measured": hbm_info(traffic, kd),
                     "l2": l2_info(dom, kd)},
        "kernels": {k: {"kernel": v["kernel"], "launch_us": v["launch_s"] * 1e6, "k": v["k"], "window": v["window"],
                        "strips": v["strips"], "models": v["models"],
                        "gcell_updates_per_s": v["cell_updates_per_launch"] / v["launch_s"] / 1e9,
                        "alg_GBps": ALG_BYTES[k] * v["cell_updates_per_launch"] / v["launch_s"] / 1e9,
                        "share_of_gn_pass_s": share[k],
                        "hbm_measured": hbm_info(measured_traffic(k, v), v), "l2": l2_info(k, v)}
                    for k, v in kr.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
