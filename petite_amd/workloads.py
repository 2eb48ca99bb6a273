"""Synthetic primary lists of the benchmark workloads (no data needed).

batch_primaries: the batched uneven-shower workload, n primaries drawn
uniformly from {e-, e+, gamma}, energies log-uniform in [0.1, 100] GeV along
+z with a Gaussian angular spread of 1 mrad per projected angle (theta_x,
theta_y ~ N(0, 1 mrad), direction proportional to (theta_x, theta_y, 1)),
on shell, weight 1, shower index and stream id = position.
"""
import numpy as np


def batch_primaries(n=100_000, seed=1234):
    from .batch import PRIMARY_DTYPE
    from .constants import m_electron
    rng = np.random.default_rng(seed)
    pid = rng.choice(np.array([11, -11, 22], dtype=np.int32), size=n)
    E = np.exp(rng.uniform(np.log(0.1), np.log(100.0), n))
    tx = rng.normal(0.0, 1e-3, n)
    ty = rng.normal(0.0, 1e-3, n)
    m = np.where(pid == 22, 0.0, m_electron)
    p = np.sqrt(E * E - m * m)
    nrm = np.sqrt(tx * tx + ty * ty + 1.0)
    prim = np.zeros(n, dtype=PRIMARY_DTYPE)
    prim["E"] = E
    prim["px"], prim["py"], prim["pz"] = p * tx / nrm, p * ty / nrm, p / nrm
    prim["mass"] = m
    prim["weight"] = 1.0
    prim["pid"] = pid
    prim["shower"] = np.arange(n, dtype=np.int32)
    prim["sid"] = np.arange(n, dtype=np.uint64)
    return prim
