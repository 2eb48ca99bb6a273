"""VEGAS adaptive-map transform, accept/reject draw and max_F scan (oracle).

Map transform (the vegas AdaptiveMap, called by the reference at
src/PETITE/shower.py:433,453; published VEGAS, Lepage 1978/2021): per
dimension with N increments and nodes g_0..g_N,
  i = floor(y N),  x = g_i + (g_{i+1} - g_i)(y N - i),  J_d = N (g_{i+1} - g_i),
and the candidate weight is J(y) f(x) with J = prod_d J_d.  With one
hypercube vegas reports wgt = J/neval; neval cancels against max_F (which
find_maxes.py:105-110 records in the same units), so here max_F is in units
of J f directly.

Accept/reject follows draw_sample, shower.py:401-465: candidates j = 0, 1, ...
in order; accept the first with max_F * U_j < J_j f(x_j), where f is
evaluated at the true energy E and the map is the one at
LU_Key = argmin|E_i - E| + 1.  Uniforms come from the Philox streams of
``oracle.philox`` (PURPOSE_VEGAS_A, one block per candidate).

Parity of the map: "parity unpinned" at the bit level for vegas itself (not
installed); pinned through the sigma-table integral identity and the
accept/reject law tests.
"""
import numpy as np

from . import philox
from .constants import PROCESS_CODE
from .integrands import INTEGRAND

MAXF_SEED = 0x5EEDF00D
MAXF_SAMPLES = 32768  # ~ n_trials (100) x neval (300) of utilities/find_maxes.py:206-210


def map_points(grid, ie, y):
    """y: (n, d) uniforms -> x (n, d), J (n,)."""
    n, d = y.shape
    x = np.empty_like(y)
    J = np.ones(n)
    for k in range(d):
        g = grid.dim_nodes(ie, k)
        N = len(g) - 1
        yN = y[:, k] * N
        i = np.floor(yN).astype(np.int64)
        i = np.clip(i, 0, N - 1)
        fr = yN - i
        dg = g[i + 1] - g[i]
        x[:, k] = g[i] + dg * fr
        J *= N * dg
    return x, J


def candidate_uniforms(seed, sid, j, ndim):
    """(y (n, ndim), u_accept (n,)) for candidate indices j (array).  One
    Philox block per candidate: y from the high 23 bits of its words; with
    four dimensions the accept uniform is made of the unused low bits (9 of
    word 0, 9 of word 1, 5 of word 2), otherwise it is word 3."""
    w = philox.words(seed, sid, j, philox.PURPOSE_VEGAS_A)
    a = philox.to_uniform(w)
    if ndim <= 3:
        return a[..., :ndim], a[..., 3]
    m9, m5 = np.uint32(0x1FF), np.uint32(0x1F)
    low = (w[..., 0] & m9) | ((w[..., 1] & m9) << np.uint32(9)) | ((w[..., 2] & m5) << np.uint32(18))
    return a[..., :4], philox.to_uniform(low << np.uint32(9))


def scan_maxf(tables, process, ie, ev, n=MAXF_SAMPLES, seed=MAXF_SEED, ndim=None):
    """max over n map candidates of J f at the grid energy (find_maxes.py:105-110)."""
    grid = tables.grids[process] if hasattr(tables, "grids") else tables
    ndim = grid.ndim if ndim is None else ndim
    sid = np.uint64((PROCESS_CODE[process] << 32) | ie)
    j = np.arange(n, dtype=np.uint64)
    y, _ = candidate_uniforms(seed, sid, j, ndim)
    x, J = map_points(grid, ie, y)
    w = J * INTEGRAND[process](x, ev)
    return float(np.max(w)), float(np.mean(w))


def draw(grid, process, E, ev, max_f, seed, sid, max_trials=3_000_000, chunk=256, vectorised=True):
    """Return (x, trials, w) of the first accepted candidate, or (None, trials, 0)."""
    ie = grid.lu_key(E)
    ndim = grid.ndim
    f = INTEGRAND[process]
    j0 = 0
    while j0 < max_trials:
        m = min(chunk if vectorised else 1, max_trials - j0)
        j = np.arange(j0, j0 + m, dtype=np.uint64)
        y, ua = candidate_uniforms(seed, sid, j, ndim)
        x, J = map_points(grid, ie, y.reshape(m, ndim))
        w = J * f(x, ev)
        acc = np.nonzero(max_f * ua < w)[0]
        if len(acc):
            k = int(acc[0])
            return x[k], j0 + k + 1, float(w[k])
        j0 += m
    return None, max_trials, 0.0
