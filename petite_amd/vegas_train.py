"""VEGAS map training: the adaptive maps of the accept/reject draws,
regenerated on the GPU.

The reference trains its maps with the vegas package
(src/PETITE/utilities/generate_integrators.py:49-100; iteration counts at
all_processes.py:1085-1096) and ships them as pickled AdaptiveMap grids, some
of which are missing (the 400 GeV Brem/PairProd/MuonBrem maps).  Here
``petite_map_accumulate`` samples every grid energy of a process through the
context's current map and sums (J f)^2 per increment; ``refine_nodes`` then
moves the nodes as vegas's AdaptiveMap.adapt does (published algorithm:
G. P. Lepage, J. Comput. Phys. 27 (1978) 192; J. Comput. Phys. 439 (2021)
110386):

1. average importance per increment, a_i = sum_f_i / n_i (0 when empty);
2. smoothing, t_0 = (7 a_0 + a_1)/8, t_i = (a_{i-1} + 6 a_i + a_{i+1})/8,
   t_{N-1} = (a_{N-2} + 7 a_{N-1})/8, normalised to sum 1;
3. damping, t_i -> ((1 - t_i) / ln(1/t_i))^alpha;
4. new nodes splitting the cumulative damped importance into N equal parts,
   linear inside an old increment; the end points stay.

A map changes only the efficiency of the accept/reject draws (their accepted
density is the integrand wherever max_F bounds J f), and its end points, which
are kept, act as the cut.  Sampling here is uniform in y (vegas adds
stratification on top of the same map adaptation).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib

TRAIN_SEED = 0x7A11AB1E


def refine_nodes(nodes, sum_f, n_f, alpha=0.5):
    """One adaptation of one dimension: nodes (N+1,) float64, sum_f and n_f
    (N,) per increment -> new nodes (N+1,) float64 (unchanged when no
    importance was seen)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    N = len(nodes) - 1
    n = np.asarray(n_f, dtype=np.float64)
    a = np.where(n > 0, np.asarray(sum_f, dtype=np.float64) / np.where(n > 0, n, 1.0), 0.0)
    if N > 1 and alpha > 0:
        t = np.empty(N)
        t[0] = (7.0 * a[0] + a[1]) / 8.0
        t[-1] = (7.0 * a[-1] + a[-2]) / 8.0
        t[1:-1] = (6.0 * a[1:-1] + a[:-2] + a[2:]) / 8.0
    else:
        t = a.copy()
    s = t.sum()
    if not (s > 0.0) or not np.isfinite(s):
        return nodes.copy()
    t = t / s
    if alpha > 0:
        pos = (t > 0.0) & (t < 1.0)
        t = np.where(pos, ((1.0 - np.where(pos, t, 0.5)) / np.log(1.0 / np.where(pos, t, 0.5))) ** alpha,
                     np.where(t >= 1.0, 1.0, 0.0))
    cum = np.concatenate([[0.0], np.cumsum(t)])
    target = cum[-1] * np.arange(1, N) / N
    i = np.searchsorted(cum, target, side="right") - 1
    i = np.clip(i, 0, N - 1)
    frac = (target - cum[i]) / np.where(t[i] > 0, t[i], 1.0)
    new = np.empty(N + 1)
    new[0], new[-1] = nodes[0], nodes[-1]
    new[1:-1] = nodes[i] + frac * (nodes[i + 1] - nodes[i])
    return new


def train_process(ctx, process, n_iter=10, n_per_energy=1 << 16, alpha=0.5, seed=TRAIN_SEED, train=None):
    """Train the context's grid of `process` in place (n_iter adaptations at
    every grid energy, or at the energy indices where the boolean mask
    `train` is set) and return the new float32 nodes (n_E, stride)."""
    from .constants import all_process_code
    code = all_process_code[process]
    g = ctx.desc.grid[code]
    n_E, stride, ndim = int(g.n_E), int(g.stride), int(g.ndim)
    if n_E <= 0:
        raise ValueError(f"train_process: {process} has no grid in this context")
    ninc = [int(g.ninc[d]) for d in range(ndim)]
    E, nodes32, cnt = ctx.tables_grid(process)
    todo = np.ones(n_E, dtype=bool) if train is None else np.asarray(train, dtype=bool)
    if todo.shape != (n_E,):
        raise ValueError("train_process: the mask needs one entry per grid energy")
    nodes = np.asarray(nodes32, dtype=np.float32).reshape(n_E, stride).copy()
    dev = torch.device("cuda")
    sum_f = torch.zeros(n_E * stride, device=dev, dtype=torch.float64)
    n_f = torch.zeros(n_E * stride, device=dev, dtype=torch.int32)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for it in range(int(n_iter)):
        sum_f.zero_()
        n_f.zero_()
        _lib.check(ctx.lib.petite_map_accumulate(ctx.ctx, code, int(n_per_energy), int(seed), it, sum_f.data_ptr(),
                                                 n_f.data_ptr(), stream), "petite_map_accumulate")
        S = sum_f.cpu().numpy().reshape(n_E, stride)
        Nn = n_f.cpu().numpy().reshape(n_E, stride)
        new = nodes.astype(np.float64)
        for ie in np.nonzero(todo)[0]:
            off = 0
            for d in range(ndim):
                N = ninc[d]
                new[ie, off:off + N + 1] = refine_nodes(new[ie, off:off + N + 1], S[ie, off:off + N], Nn[ie, off:off + N],
                                                        alpha)
                off += N + 1
        nodes = new.astype(np.float32)
        _lib.check(ctx.lib.petite_set_grid(ctx.ctx, code, nodes.ctypes.data, nodes.size), "petite_set_grid")
    return nodes
