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
installed); pinned through the sigma-table integral identity
(tests/test_oracle_tables.py) and the law of the accepted points
(tests/test_oracle_draw_law.py).
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


# ------------------------------------------------------------ map training --
def train_accumulate(grid, process, ie, ev, seed, iteration, n):
    """(sum_f, n_f) of one grid energy, in the node layout of the grid (one
    unused slot per dimension): samples k = 0..n-1 of iteration `iteration`
    take Philox block iteration*n + k of stream (code << 32 | ie), purpose
    MAP_TRAIN; per dimension the sample's increment gets (J f)^2 and a count
    (the data vegas's AdaptiveMap.adapt consumes; utilities/
    generate_integrators.py:66-74 trains the reference's maps with it)."""
    code = PROCESS_CODE[process]
    sid = np.uint64((code << 32) | ie)
    j = np.arange(n, dtype=np.uint64) + np.uint64(iteration) * np.uint64(n)
    w = philox.words(seed, sid, j, philox.PURPOSE_MAP_TRAIN)
    ndim = grid.ndim
    y = philox.to_uniform(w)[:, :ndim]
    x, J = map_points(grid, ie, y)
    f = J * INTEGRAND[process](x, ev)
    f2 = f * f
    ok = np.isfinite(f2)
    sizes = [len(grid.dim_nodes(ie, d)) for d in range(ndim)]
    sum_f = np.zeros(sum(sizes))
    n_f = np.zeros(sum(sizes), dtype=np.int64)
    off = 0
    for d in range(ndim):
        N = sizes[d] - 1
        i = np.floor(y[:, d] * N).astype(np.int64)
        np.add.at(sum_f, off + i[ok], f2[ok])
        np.add.at(n_f, off + i[ok], 1)
        off += N + 1
    return sum_f, n_f


def refine(nodes, sum_f, n_f, alpha=0.5):
    """One vegas map adaptation of one dimension (AdaptiveMap.adapt; Lepage,
    J. Comput. Phys. 27 (1978) 192 and 439 (2021) 110386), written as its
    sequential loops: per-increment average of (J f)^2, smoothing
    (7,1)/8 at the ends and (1,6,1)/8 inside, normalisation, damping
    ((1 - t)/ln(1/t))^alpha, then the walk that places node k where the
    cumulative damped importance reaches k/N of its total."""
    nodes = [float(v) for v in nodes]
    N = len(nodes) - 1
    avg = [sum_f[i] / n_f[i] if n_f[i] > 0 else 0.0 for i in range(N)]
    if alpha > 0 and N > 1:
        t = [0.0] * N
        t[0] = (7.0 * avg[0] + avg[1]) / 8.0
        t[N - 1] = (7.0 * avg[N - 1] + avg[N - 2]) / 8.0
        for i in range(1, N - 1):
            t[i] = (6.0 * avg[i] + avg[i - 1] + avg[i + 1]) / 8.0
    else:
        t = list(avg)
    total = sum(t)
    if not (total > 0.0) or not np.isfinite(total):
        return np.array(nodes)
    t = [v / total for v in t]
    if alpha > 0:
        t = [((1.0 - v) / np.log(1.0 / v)) ** alpha if 0.0 < v < 1.0 else (1.0 if v >= 1.0 else 0.0) for v in t]
    per = sum(t) / N
    new = [nodes[0]]
    acc = 0.0   # damped importance accumulated up to the start of increment j
    j = 0
    for k in range(1, N):
        goal = k * per
        while j < N - 1 and acc + t[j] <= goal:
            acc += t[j]
            j += 1
        frac = (goal - acc) / t[j] if t[j] > 0 else 0.0
        new.append(nodes[j] + frac * (nodes[j + 1] - nodes[j]))
    new.append(nodes[N])
    return np.array(new)
