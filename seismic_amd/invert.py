"""Inversion loops driving the propagator (core/seis_invert.py, core/seis_numerics.py).

Every loop is written as a *machine*: a generator that yields propagator
requests (``Op``) and is resumed with their results.  One machine per
inversion request; an ``Executor`` runs the pending ops of many machines as
batched launches (continuous batching), or one at a time (the plain path).
Control flow and float64 vector algebra stay on the host side in torch; every
wavefield operation goes through libseis.

  lbfgs_machine   bfgs + _strong_wolfe + _cubic_interpolate, core/seis_numerics.py:138-448,
                  with the cost closure of seis_to_vel_lbfgs (core/seis_invert.py:125-138)
  gn_machine      Gauss-Newton passes of seis_to_vel_gn (core/seis_invert.py:63-107) with
                  the dense J^T J solve replaced by conjugate gradients on J^T J products
  bfgs            reference signature, driving lbfgs_machine on one model
  cost_and_gradient  reference signature (core/seis_invert.py:19-49)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Callable, Sequence

import numpy as np
import torch

from . import autotune
from ._lib import CAP_ADJOINT, CAP_FORWARD, CAP_TANGENT
from .geometry import Geometry
from .propagator import Propagator, vel_to_seis


# --------------------------------------------------------------------- ops
@dataclass
class Op:
    """A propagator request from a machine.

    kind 'fev': vec -> (misfit, g) with misfit = sum((F(vec) - target)^2) and
                g = J(vec)^T (F(vec) - target)                     (float64)
    kind 'gnp': (vec, dvec) -> J(vec)^T J(vec) dvec                 (float64)
    kind 'fwd': vecs (list) -> [sum((F(v) - target)^2) for v in vecs]
    kind 'dir': step (LbfgsStep) -> (d, (g.d, max|d|, sum|g|, max|g|)): the
                LBFGS history update and two-loop direction
    """

    kind: str
    vec: Any = None
    dvec: Any = None
    vecs: Any = None
    step: Any = None


@dataclass
class Request:
    """One inversion: geometry, observed data (float32, device) and prior."""

    geom: Geometry
    target: torch.Tensor
    prior: Any
    start: torch.Tensor                  # float64 (n_params,)
    rid: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_data(self) -> int:
        return self.geom.n_data


# --------------------------------------------------------------------- LBFGS
class LbfgsHistory:
    """Device-resident (s, y) history of one LBFGS run: ring buffers of
    cap + 1 rows (libseis ``seis_lbfgs_direction``).  ``cap`` is the history
    size; the reference keeps every pair (cap = max_iter,
    core/seis_numerics.py:336-340).  Buffers are allocated on first use."""

    def __init__(self, n: int, cap: int, device):
        self.n, self.cap, self.device = int(n), max(1, int(cap)), torch.device(device)
        self.S = self.Y = self.rho = self.al = self.ring = self.hdiag = None

    def allocate(self):
        if self.S is None:
            R = self.cap + 1
            f64 = dict(dtype=torch.float64, device=self.device)
            self.S = torch.zeros(R, self.n, **f64)
            self.Y = torch.zeros(R, self.n, **f64)
            self.rho = torch.zeros(R, **f64)
            self.al = torch.zeros(R, **f64)
            self.ring = torch.zeros(2, dtype=torch.int32, device=self.device)
            self.hdiag = torch.ones(1, **f64)

    def nbytes(self) -> int:
        return 8 * ((self.cap + 1) * (2 * self.n + 2) + 1) + 8


@dataclass
class LbfgsStep:
    """One direction request: gradient g, the previous iterate's gradient,
    direction and step (ignored when ``first``)."""

    hist: LbfgsHistory
    g: torch.Tensor
    g_prev: torch.Tensor | None
    d_prev: torch.Tensor | None
    t: float
    first: bool


def lbfgs_directions(steps: Sequence[LbfgsStep]) -> list:
    """LBFGS directions of many requests in one launch per 16
    (``seis_lbfgs_direction``): returns [(d, (g.d, max|d|, sum|g|, max|g|))]
    with one device-to-host read for all the scalars."""
    import ctypes

    from . import _lib
    if not steps:
        return []
    dev = steps[0].g.device
    if dev.type != "cuda":
        raise RuntimeError("lbfgs_directions needs a ROCm GPU (torch.cuda)")
    scal = torch.empty(len(steps), 4, dtype=torch.float64, device=dev)
    reqs = (_lib.LbfgsReq * len(steps))()
    keep = []
    outs = []
    ptr = lambda t: None if t is None else t.data_ptr()
    for i, st in enumerate(steps):
        h = st.hist
        if h.n > _lib.LBFGS_MAX_N:
            raise ValueError(f"LBFGS direction supports at most {_lib.LBFGS_MAX_N} parameters, got {h.n}")
        h.allocate()
        g = st.g.to(torch.float64).contiguous()
        gp = None if st.first else st.g_prev.to(torch.float64).contiguous()
        dp = None if st.first else st.d_prev.to(torch.float64).contiguous()
        d = torch.empty(h.n, dtype=torch.float64, device=dev)
        keep += [g, gp, dp]
        outs.append(d)
        r = reqs[i]
        r.S, r.Y, r.rho, r.al = ptr(h.S), ptr(h.Y), ptr(h.rho), ptr(h.al)
        r.ring, r.hdiag = ptr(h.ring), ptr(h.hdiag)
        r.g, r.g_prev, r.d_prev = ptr(g), ptr(gp), ptr(dp)
        r.t = 0.0 if st.first else float(st.t)
        r.d, r.scal = ptr(d), scal[i].data_ptr()
        r.n, r.cap, r.first = h.n, h.cap, int(bool(st.first))
    with torch.cuda.device(dev):
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(_lib.lib().seis_lbfgs_direction(reqs, len(steps), stream), "seis_lbfgs_direction")
    sc = scal.cpu().tolist()
    return [(d, tuple(v)) for d, v in zip(outs, sc)]


def _cubic_interpolate(x1, f1, g1, x2, f2, g2, bounds=None):
    """core/seis_numerics.py:279-304 (scalars as Python floats)."""
    if bounds is not None:
        xmin_bound, xmax_bound = bounds
    else:
        xmin_bound, xmax_bound = (x1, x2) if x1 <= x2 else (x2, x1)
    d1 = g1 + g2 - 3 * (f1 - f2) / (x1 - x2)
    d2_square = d1 ** 2 - g1 * g2
    if d2_square >= 0:
        d2 = math.sqrt(d2_square)
        if x1 <= x2:
            min_pos = x2 - (x2 - x1) * ((g2 + d2 - d1) / (g2 - g1 + 2 * d2))
        else:
            min_pos = x1 - (x1 - x2) * ((g1 + d2 - d1) / (g1 - g2 + 2 * d2))
        return min(max(min_pos, xmin_bound), xmax_bound)
    return (xmin_bound + xmax_bound) / 2.0


def _strong_wolfe(obj, x, t, d, f, g, gtd, c1=1e-4, c2=0.9, tolerance_change=1e-9, max_ls=25, d_norm=None):
    """core/seis_numerics.py:138-277 as a sub-machine; ``obj(x, t, d)`` is a
    machine returning (f: float, g: tensor).  ``d_norm`` = max|d| when the
    caller has it."""
    if d_norm is None:
        d_norm = float(d.abs().max())
    g = g.clone()
    f_new, g_new = yield from obj(x, t, d)
    ls_func_evals = 1
    gtd_new = float(g_new.dot(d))
    t_prev, f_prev, g_prev, gtd_prev = 0, f, g, gtd
    done = False
    ls_iter = 0
    while ls_iter < max_ls:
        if f_new > (f + c1 * t * gtd) or (ls_iter > 1 and f_new >= f_prev):
            bracket, bracket_f = [t_prev, t], [f_prev, f_new]
            bracket_g, bracket_gtd = [g_prev, g_new.clone()], [gtd_prev, gtd_new]
            break
        if abs(gtd_new) <= -c2 * gtd:
            bracket, bracket_f, bracket_g = [t], [f_new], [g_new]
            done = True
            break
        if gtd_new >= 0:
            bracket, bracket_f = [t_prev, t], [f_prev, f_new]
            bracket_g, bracket_gtd = [g_prev, g_new.clone()], [gtd_prev, gtd_new]
            break
        min_step = t + 0.01 * (t - t_prev)
        max_step = t * 10
        tmp = t
        t = _cubic_interpolate(t_prev, f_prev, gtd_prev, t, f_new, gtd_new, bounds=(min_step, max_step))
        t_prev, f_prev, g_prev, gtd_prev = tmp, f_new, g_new.clone(), gtd_new
        f_new, g_new = yield from obj(x, t, d)
        ls_func_evals += 1
        gtd_new = float(g_new.dot(d))
        ls_iter += 1
    if ls_iter == max_ls:
        bracket, bracket_f, bracket_g = [0, t], [f, f_new], [g, g_new]
    insuf_progress = False
    low_pos, high_pos = (0, 1) if bracket_f[0] <= bracket_f[-1] else (1, 0)
    while not done and ls_iter < max_ls:
        if abs(bracket[1] - bracket[0]) * d_norm < tolerance_change:
            break
        t = _cubic_interpolate(bracket[0], bracket_f[0], bracket_gtd[0],
                               bracket[1], bracket_f[1], bracket_gtd[1])
        eps = 0.1 * (max(bracket) - min(bracket))
        if min(max(bracket) - t, t - min(bracket)) < eps:
            if insuf_progress or t >= max(bracket) or t <= min(bracket):
                if abs(t - max(bracket)) < abs(t - min(bracket)):
                    t = max(bracket) - eps
                else:
                    t = min(bracket) + eps
                insuf_progress = False
            else:
                insuf_progress = True
        else:
            insuf_progress = False
        f_new, g_new = yield from obj(x, t, d)
        ls_func_evals += 1
        gtd_new = float(g_new.dot(d))
        ls_iter += 1
        if f_new > (f + c1 * t * gtd) or f_new >= bracket_f[low_pos]:
            bracket[high_pos] = t
            bracket_f[high_pos] = f_new
            bracket_g[high_pos] = g_new.clone()
            bracket_gtd[high_pos] = gtd_new
            low_pos, high_pos = (0, 1) if bracket_f[0] <= bracket_f[1] else (1, 0)
        else:
            if abs(gtd_new) <= -c2 * gtd:
                done = True
            elif gtd_new * (bracket[high_pos] - bracket[low_pos]) >= 0:
                bracket[high_pos] = bracket[low_pos]
                bracket_f[high_pos] = bracket_f[low_pos]
                bracket_g[high_pos] = bracket_g[low_pos]
                bracket_gtd[high_pos] = bracket_gtd[low_pos]
            bracket[low_pos] = t
            bracket_f[low_pos] = f_new
            bracket_g[low_pos] = g_new.clone()
            bracket_gtd[low_pos] = gtd_new
    t = bracket[low_pos]
    return bracket_f[low_pos], bracket_g[low_pos], t, ls_func_evals


def bfgs_machine(func, x0: torch.Tensor, max_iter: int, tolerance_grad: float,
                 tolerance_change: float, history_size: int | None = None):
    """core/seis_numerics.py:307-448 as a machine.  ``func(x)`` is a machine
    returning (loss: float, grad: tensor).  history_size=None keeps every
    (s, y) pair like the reference; an int keeps the most recent ones.  The
    history update and two-loop recursion (the reference's lbfgs_backward /
    lbfgs_forward kernels) run on the device as a ``dir`` op, batched over
    requests by the executor.  Returns (x, n_evals, n_iter)."""
    lr = 1
    max_eval = max_iter * 5 // 4
    loss, flat_grad = yield from func(x0)
    x = x0.clone()
    current_evals = 1
    if float(flat_grad.abs().max()) <= tolerance_grad:
        return x, current_evals, 0
    hist = LbfgsHistory(x0.numel(), history_size if history_size is not None else max_iter, x0.device)
    n_iter = 0
    prev_flat_grad = None
    d = t = None
    while n_iter < max_iter:
        n_iter += 1
        d, (gtd, d_max, g_sum, _) = yield Op("dir", step=LbfgsStep(hist, flat_grad, prev_flat_grad, d, t,
                                                                   n_iter == 1))
        prev_flat_grad = flat_grad
        prev_loss = loss
        if n_iter == 1:
            t = min(1.0, 1.0 / g_sum) * lr
        else:
            t = lr
        x_init = x.clone()

        def obj(xx, tt, dd):
            return (yield from func(xx + tt * dd))

        loss, flat_grad, t, ls_func_evals = yield from _strong_wolfe(obj, x_init, t, d, loss, flat_grad, gtd,
                                                                     d_norm=d_max)
        x.add_(d, alpha=t)
        opt_cond = float(flat_grad.abs().max()) <= tolerance_grad
        current_evals += ls_func_evals
        if n_iter == max_iter:
            break
        if current_evals >= max_eval:
            break
        if opt_cond:
            break
        if d_max * t <= tolerance_change:      # max|d t| (t > 0)
            break
        if abs(loss - prev_loss) < tolerance_change:
            break
    return x, current_evals, n_iter


def _basis_of(prior):
    return getattr(prior, "basis", None)


def lbfgs_machine(req: Request, max_iter: int = 30, tolerance_grad: float = 1e-7,
                  tolerance_change: float = 0.0, scaling: float = 1e15,
                  seis_error_tolerance: float = 0.0, history_size: int | None = None):
    """InversionModel.seis_to_vel_lbfgs (core/seis_invert.py:109-152) as a machine.
    Returns the inverted parameter vector (float64, n_params)."""
    prior = req.prior
    N = req.n_data
    x_guess = prior.to_basis(req.start)
    stats = req.meta.setdefault("stats", {"fev": 0})

    def func(x):
        vec = prior.from_basis(x)
        misfit, gres = yield Op("fev", vec=vec)
        stats["fev"] += 1
        cost_prior, grad_prior = prior.cost_grad(x)
        cost_res = misfit / N
        cost = float(cost_prior) + cost_res
        grad = grad_prior + (2.0 / N) * prior.basis_T(gres)
        stats["misfit"] = cost_res
        if cost_res < seis_error_tolerance:
            cost, grad = 0.0 * cost, 0.0 * grad
        return cost * scaling, grad * scaling

    x, n_evals, n_iter = yield from bfgs_machine(func, x_guess, max_iter, tolerance_grad,
                                                 tolerance_change, history_size)
    stats["n_iter"] = n_iter
    if torch.isnan(x).any():
        return req.start.clone()       # core/seis_invert.py:146-149
    return prior.from_basis(x)


# --------------------------------------------------------------------- Gauss-Newton
def conjugate_gradient_machine(apply_A, b: torch.Tensor, n_iter: int):
    """CG from 0 on an SPD operator given as a machine ``apply_A(p)``.

    The scalars stay 0-d device tensors, so a batch of machines resumed
    after one shared product queues its updates without a host round trip
    each.  An exactly converged residual (rr = 0) makes every later step
    zero (the guarded a and beta), which leaves x as a break would."""
    x = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rr = r.dot(r)
    if float(rr) == 0.0:
        return x
    zero = torch.zeros((), dtype=b.dtype, device=b.device)
    for _ in range(n_iter):
        Ap = yield from apply_A(p)
        pAp = p.dot(Ap)
        a = torch.where(pAp != 0, rr / pAp, zero)
        x.add_(a * p)
        r.sub_(a * Ap)
        rr_new = r.dot(r)
        p = r + torch.where(rr != 0, rr_new / rr, zero) * p
        rr = rr_new
    return x


def gn_machine(req: Request, n_outer: int = 10, n_cg: int | None = 20, n_scales: int = 50):
    """n_outer passes of seis_to_vel_gn (core/seis_invert.py:63-107).

    Each pass: rhs = -B^T J^T r / N - lam P x  (:77-78); solve
    (lam P (+) 0 + J^T J / N) dx = rhs by n_cg CG steps, with J^T J products
    from one tangent + one adjoint pass each (replacing the dense J of
    :84-92); line search over linspace(-0.1, 1.1, n_scales) of the forward
    cost (:95-102).  n_cg = None solves densely as the reference does
    (:89-92): the matrix is assembled from one product per basis vector and
    solved with torch.linalg.solve, for bases too ill-conditioned for CG (the
    SE prior's SVD modes).  The prior must provide apply_P.
    """
    prior = req.prior
    N = req.n_data
    vec = req.start.clone()
    scales = np.linspace(-0.1, 1.1, n_scales)
    stats = req.meta.setdefault("stats", {"fev": 0, "gnp": 0, "fwd": 0})
    for _ in range(n_outer):
        x_guess = prior.to_basis(vec)
        v_fev = prior.from_basis(x_guess)
        misfit, gres = yield Op("fev", vec=v_fev)
        stats["fev"] += 1
        rhs = -prior.basis_T(gres) / N
        rhs[:-1] -= prior.lam * prior.apply_P(x_guess[:-1])
        # J is taken at vec (:86); when the projected fev point is vec itself
        # the products name that same vector, so the executor reuses the
        # fev's saved forward (identity basis: always; else one check per pass)
        identity = prior.is_identity() if hasattr(prior, "is_identity") else prior.basis is None
        lin = v_fev if (identity or bool(torch.equal(v_fev, vec))) else vec

        def apply_A(p):
            jtj = yield Op("gnp", vec=lin, dvec=prior.from_basis(p))
            stats["gnp"] += 1
            out = prior.basis_T(jtj) / N
            out[:-1] += prior.lam * prior.apply_P(p[:-1])
            return out

        if n_cg is None:
            cols = []
            for i in range(rhs.numel()):
                e = torch.zeros_like(rhs)
                e[i] = 1.0
                cols.append((yield from apply_A(e)))
            dx = torch.linalg.solve(torch.stack(cols, dim=1).cpu(), rhs.cpu()).to(rhs.device)   # host LAPACK
        else:
            dx = yield from conjugate_gradient_machine(apply_A, rhs, n_cg)
        # all trial points at once: row k is x_guess + scales[k] * dx
        trial = x_guess[None, :] + torch.as_tensor(scales, dtype=dx.dtype, device=dx.device)[:, None] * dx[None, :]
        misfits = yield Op("fwd", vecs=list(prior.from_basis_rows(trial)))
        stats["fwd"] += len(scales)
        costs = (prior.cost_rows(trial).cpu().numpy() + np.asarray(misfits, dtype=np.float64) / N).tolist()
        best = float(scales[int(np.argmin(costs))])
        vec = vec + best * prior.from_basis(dx)
        stats["misfit"] = min(costs)
        stats.setdefault("misfit_hist", []).append(min(costs))
    return vec


# --------------------------------------------------------------------- executor
class Executor:
    """Runs the pending ops of a set of machines on fixed slots.

    ``slots`` lists one geometry per slot; request i of a batch occupies slot
    i.  ``max_batch`` > 1 runs each op kind as one launch over every slot that
    has one pending (inactive slots are skipped by the kernels); max_batch = 1
    is the plain path: every op runs alone on a single-model plan.  With
    ``reuse_forward`` (full-history plans), J^T J products at the point of the
    models' last saving forward skip the forward recomputation.
    """

    def __init__(self, slots: Sequence[Geometry], ckpt_every: int = 0, max_batch: int | None = None,
                 fwd_fanout: int = 50, reuse_forward: bool = True, device="cuda",
                 configure: Callable | None = None):
        self.slots = list(slots)
        self.configure = configure    # called on each new plan after the tuned configs are applied
        self.n = len(self.slots)
        self.max_batch = self.n if max_batch is None else max_batch
        self.ckpt_every = ckpt_every
        self.fwd_fanout = fwd_fanout
        self.reuse_forward = reuse_forward
        self.device = torch.device(device)
        self.targets: list = [None] * self.n
        self._plans: dict = {}

    def close(self):
        """Free every plan's device memory (the executor can be used again)."""
        for P in self._plans.values():
            P.close()
        self._plans.clear()

    # plan cache --------------------------------------------------------
    def _plan(self, key, geoms, caps):
        if key not in self._plans:
            P = Propagator(geoms, ckpt_every=self.ckpt_every, caps=caps, device=self.device)
            autotune.apply(P)
            if self.configure is not None:
                self.configure(P)
            self._plans[key] = P
        return self._plans[key]

    def _grad_plan(self, slot_ids):
        if self.max_batch == 1:
            g = self.slots[slot_ids[0]]
            return self._plan(("g1", g), [g], CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT)
        return self._plan(("gB",), self.slots, CAP_FORWARD | CAP_TANGENT | CAP_ADJOINT)

    def set_target(self, slot: int, target: torch.Tensor):
        self.targets[slot] = target.to(self.device, torch.float32).reshape(-1)
        self._dd = None
        self._dd_fan = None

    def _targets_cat(self):
        if getattr(self, "_dd", None) is None:
            self._dd = torch.cat([t if t is not None else torch.zeros(g.n_data, device=self.device)
                                  for t, g in zip(self.targets, self.slots)])
        return self._dd

    # execution ---------------------------------------------------------
    def execute(self, pending: dict) -> dict:
        """pending: slot -> Op.  Returns slot -> result."""
        out = {}
        by_kind: dict = {}
        for s, op in pending.items():
            by_kind.setdefault(op.kind, []).append(s)
        for kind, ids in by_kind.items():
            ids = sorted(ids)
            chunks = [ids[i:i + self.max_batch] for i in range(0, len(ids), self.max_batch)]
            for ch in chunks:
                if kind == "fev":
                    out.update(self._fev(ch, pending))
                elif kind == "gnp":
                    out.update(self._gnp(ch, pending))
                elif kind == "fwd":
                    out.update(self._fwd(ch, pending))
                elif kind == "dir":
                    out.update(zip(ch, lbfgs_directions([pending[i].step for i in ch])))
                else:
                    raise ValueError(kind)
        return out

    def _pack(self, P: Propagator, ids, vecs):
        if list(ids) == list(range(P.n_models)):   # every model: one concatenation
            return torch.cat([v.reshape(-1) for v in vecs]).to(device=self.device, dtype=torch.float32)
        buf = torch.zeros(P.n_vec, dtype=torch.float32, device=self.device)
        for i, v in zip(ids, vecs):
            a, b = int(P.vec_off[i]), int(P.vec_off[i + 1])
            buf[a:b] = v
        return buf

    def _local(self, ids):
        """(plan, local slot ids) for a chunk."""
        P = self._grad_plan(ids)
        if self.max_batch == 1:
            return P, [0]
        live = set(ids)
        mask = tuple(i in live for i in range(self.n))
        if getattr(P, "_mask", None) != mask:
            P.set_active(mask)
            P._mask = mask
        return P, list(ids)

    @staticmethod
    def _note_saving(P, ids, loc, pending):
        """Record, per model of plan P, which slot's op vector its saved
        forward history now belongs to (host bookkeeping: the vector object
        itself, so the check costs no device round trip)."""
        rec = P.__dict__.setdefault("_saved_for", {})
        for i, l in zip(ids, loc):
            rec[l] = (i, pending[i].vec)

    @staticmethod
    def _saved_here(P, ids, loc, pending) -> bool:
        rec = P.__dict__.get("_saved_for", {})
        return all(l in rec and rec[l][0] == i and rec[l][1] is pending[i].vec for i, l in zip(ids, loc))

    def _fev(self, ids, pending):
        P, loc = self._local(ids)
        v = self._pack(P, loc, [pending[i].vec for i in ids])
        dd = self.targets[ids[0]] if self.max_batch == 1 else self._targets_cat()
        _, misfit, grad = P.gradient(v, dd, on_residual=True)
        self._note_saving(P, ids, loc, pending)
        mis = misfit.cpu().tolist()
        return {i: (mis[l], grad[int(P.vec_off[l]):int(P.vec_off[l + 1])].clone()) for i, l in zip(ids, loc)}

    def _gnp(self, ids, pending):
        P, loc = self._local(ids)
        v = self._pack(P, loc, [pending[i].vec for i in ids])
        dv = self._pack(P, loc, [pending[i].dvec for i in ids])
        if self.reuse_forward and P.full_history and self._saved_here(P, ids, loc, pending):
            # every model's last saving forward (its fev or previous product)
            # ran at this very vector: tangent from the saved history
            _, grad = P.gn_product_saved(v, dv, check=False)
        else:
            _, grad = P.gn_product(v, dv)
            self._note_saving(P, ids, loc, pending)
        return {i: grad[int(P.vec_off[l]):int(P.vec_off[l + 1])].clone() for i, l in zip(ids, loc)}

    def _fanout_plan(self):
        """One forward-only plan of ``fwd_fanout`` models per slot (slot i
        owns models i*F .. i*F+F-1): every line search of every slot runs on
        it, the models in use activated per launch, so its size is fixed."""
        F = self.fwd_fanout
        P = self._plan(("fB",), [g for g in self.slots for _ in range(F)], CAP_FORWARD)
        return P

    def _fan_targets(self, P):
        if getattr(self, "_dd_fan", None) is None:
            F = self.fwd_fanout
            self._dd_fan = torch.cat([(t if t is not None else torch.zeros(g.n_data, device=self.device))
                                      for t, g in zip(self.targets, self.slots) for _ in range(F)])
        return self._dd_fan

    def _fwd(self, ids, pending):
        """Line-search forwards: every trial model of every op in one launch
        (batched; rounds of ``fwd_fanout`` trial points per slot) or one
        model per launch (plain)."""
        out = {}
        if self.max_batch == 1:
            for i in ids:
                g = self.slots[i]
                P = self._plan(("f1", g), [g], CAP_FORWARD)
                ms = []
                for v in pending[i].vecs:
                    ms.append(P.forward_misfit(P._vec(v), self.targets[i]))
                out[i] = [float(m[0]) for m in torch.stack(ms).cpu()]
            return out
        F = self.fwd_fanout
        P = self._fanout_plan()
        dd = self._fan_targets(P)
        res = {i: [] for i in ids}
        n_rounds = max((len(pending[i].vecs) + F - 1) // F for i in ids)
        for r in range(n_rounds):
            part = {i: pending[i].vecs[r * F:(r + 1) * F] for i in ids}
            models = [i * F + k for i in ids for k in range(len(part[i]))]
            if not models:
                continue
            live = set(models)
            mask = tuple(m in live for m in range(P.n_models))
            if getattr(P, "_mask", None) != mask:
                P.set_active(mask)
                P._mask = mask
            v = self._pack(P, models, [x for i in ids for x in part[i]])
            mis = P.forward_misfit(v, dd).cpu()
            for i in ids:
                res[i] += mis[[i * F + k for k in range(len(part[i]))]].tolist()
        return res


def fanout_plan_bytes(slots: Sequence[Geometry], fanout: int = 50) -> int:
    """Device bytes of an executor's line-search fan-out plan over ``slots``."""
    from .propagator import plan_bytes
    return plan_bytes([g for g in slots for _ in range(fanout)], ckpt_every=0, caps=CAP_FORWARD)


def run_machines(machines: dict, executor: Executor, on_finish: Callable | None = None) -> dict:
    """Drive machines (slot -> generator) to completion through ``executor``."""
    pending = {}
    results = {}
    for s, m in machines.items():
        try:
            pending[s] = next(m)
        except StopIteration as e:     # nothing to evaluate (e.g. zero iterations)
            results[s] = e.value
            if on_finish:
                on_finish(s, e.value)
    while pending:
        res = executor.execute(pending)
        for s, r in res.items():
            try:
                pending[s] = machines[s].send(r)
            except StopIteration as e:
                results[s] = e.value
                del pending[s]
                if on_finish:
                    on_finish(s, e.value)
    return results


# --------------------------------------------------------------------- reference signatures
def cost_and_gradient(x, target, prior, basis_functions, compute_gradient=False, geometry=None):
    """core/seis_invert.py:19-49 on the GPU propagator (single model).

    x (N,1) float64, target (n_data,1).  Returns (cost, grad, cost_prior,
    cost_residual) or (cost, cost_prior, cost_residual)."""
    x = torch.as_tensor(x, dtype=torch.float64).reshape(-1)
    cost_prior, grad_prior = prior.cost_grad(x.to(prior.device))
    B = basis_functions
    vec = (B @ x.to(prior.device)) if B is not None else x.to(prior.device)
    target = torch.as_tensor(target)
    if compute_gradient:
        s, _, s_adj = vel_to_seis(vec.reshape(-1, 1), vec_adjoint=target.reshape(-1, 1),
                                  adjoint_on_residual=True, geometry=geometry)
    else:
        s = vel_to_seis(vec.reshape(-1, 1), geometry=geometry)[0]
    t = target.to(s.device, torch.float64).reshape(-1, 1)
    cost_residual = torch.mean((s - t) ** 2)
    if compute_gradient:
        g = s_adj.reshape(-1)
        gres = (2.0 / len(s)) * ((B.T @ g) if B is not None else g)
        return cost_prior + cost_residual, (grad_prior + gres).reshape(-1, 1), cost_prior, cost_residual
    return cost_prior + cost_residual, cost_prior, cost_residual


def bfgs(cost_and_gradient_func, x0, max_iter, tolerance_grad, tolerance_change, history_size=None):
    """core/seis_numerics.py:307 signature: cost_and_gradient_func(x) -> (loss, grad)."""
    def func(x):
        f, g = cost_and_gradient_func(x)
        return float(f), g
        yield  # pragma: no cover  (makes this a generator)

    m = bfgs_machine(func, torch.as_tensor(x0), max_iter, tolerance_grad, tolerance_change, history_size)
    try:
        op = next(m)
        while True:
            if op.kind != "dir":
                raise RuntimeError("bfgs machine yielded a propagator op")
            op = m.send(lbfgs_directions([op.step])[0])
    except StopIteration as e:
        return e.value[0]


def invert(requests: Sequence[Request], method: str = "lbfgs", batched: bool = True,
           ckpt_every: int = 0, reuse_forward: bool = True, **kw) -> list:
    """Invert a list of requests; batched=False runs the plain one-at-a-time path."""
    ex = Executor([r.geom for r in requests], ckpt_every=ckpt_every,
                  max_batch=None if batched else 1, reuse_forward=reuse_forward)
    machines = {}
    for i, r in enumerate(requests):
        ex.set_target(i, r.target)
        machines[i] = (lbfgs_machine(r, **kw) if method == "lbfgs" else gn_machine(r, **kw))
    if batched:
        res = run_machines(machines, ex)
        return [res[i] for i in range(len(requests))]
    out = []
    for i in range(len(requests)):
        out.append(run_machines({i: machines[i]}, ex)[i])
    return out
