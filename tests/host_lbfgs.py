"""Host stand-in for the ``dir`` op of the inversion machines in CPU tests:
the oracle's history update and two-loop recursion (oracle.invert,
core/seis_numerics.py:350-392) on per-history Python lists.  Test
infrastructure only; the product runs the op on the GPU."""
import numpy as np
import torch

from oracle import invert as oinv


def direction(step):
    """(d, (g.d, max|d|, sum|g|, max|g|)) for an ``invert.LbfgsStep``."""
    h = step.hist
    g = step.g.detach().cpu().numpy().astype(np.float64)
    if step.first or not hasattr(h, "_host"):
        h._host = {"dirs": [], "stps": [], "ro": [], "H": 1.0}
    st = h._host
    if step.first:
        d = -g
    else:
        st["H"] = oinv.lbfgs_history_update(st["dirs"], st["stps"], st["ro"], st["H"], g,
                                            step.g_prev.detach().cpu().numpy(), step.d_prev.detach().cpu().numpy(),
                                            step.t, h.cap)
        d = oinv.lbfgs_two_loop(st["dirs"], st["stps"], st["ro"], st["H"], g)
    dt = torch.tensor(d, dtype=torch.float64, device=step.g.device)
    return dt, (float(g @ d), float(np.abs(d).max()), float(np.abs(g).sum()), float(np.abs(g).max()))
