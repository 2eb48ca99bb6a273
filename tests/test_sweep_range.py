"""The sweep-range arithmetic of k_sweep_setup (petite_amd/csrc/sweep.h),
compiled for the host with g++: new particles plus resumed ones never exceed
the sweep bound (each files at most one draw request into a region of
`bound` requests per process), the range never passes the store count or the
capacity, and it is as large as those limits allow."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <cstdint>
#include <cstdio>
#include "sweep.h"
int main() {
  long long a, b, c, d, e, f;
  while (std::scanf("%lld %lld %lld %lld %lld %lld", &a, &b, &c, &d, &e, &f) == 6) {
    const SweepRange r = sweep_range(a, b, c, d, e, f);
    std::printf("%lld %lld %lld\n", (long long)r.first, (long long)r.n, (long long)r.nres);
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("sweep")
    (d / "t.cpp").write_text(SRC)
    out = d / "t"
    try:
        subprocess.run(["g++", "-std=c++17", "-D__host__=", "-D__device__=", "-I",
                        os.path.join(ROOT, "petite_amd", "csrc"), str(d / "t.cpp"), "-o", str(out)], check=True)
    except (FileNotFoundError, subprocess.CalledProcessError):
        pytest.skip("no C++ compiler")
    return str(out)


def _run(exe, cases):
    inp = "\n".join(" ".join(str(int(v)) for v in c) for c in cases) + "\n"
    res = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    return np.array([[int(v) for v in line.split()] for line in res if line.strip()], dtype=np.int64)


def test_bound_limited_sweep_with_hand_ons(exe):
    # 2^23 new particles waiting, 2^20 handed on: the sweep takes 2^23 - 2^20 new ones
    b = 1 << 23
    r = _run(exe, [(0, 3 * b, 1 << 30, b, 1 << 20, 1 << 20)])[0]
    assert r.tolist() == [0, b - (1 << 20), 1 << 20]


def test_invariants_on_random_cases(exe):
    rng = np.random.default_rng(3)
    n = 20000
    first = rng.integers(0, 1 << 30, n)
    now = first + rng.integers(-5, 1 << 24, n)
    cap = first + rng.integers(-5, 1 << 25, n)
    bound = rng.integers(64, 1 << 24, n)
    rcap = np.minimum(bound, rng.integers(1, 1 << 21, n))
    handed = rng.integers(0, 1 << 22, n)
    cases = np.stack([first, now, cap, bound, handed, rcap], 1)
    out = _run(exe, cases)
    f, m, nres = out[:, 0], out[:, 1], out[:, 2]
    assert np.array_equal(f, first)
    assert np.all(m >= 0) and np.all(nres >= 0)
    assert np.all(nres == np.minimum(handed, rcap))
    assert np.all(m + nres <= np.maximum(bound, nres))             # never past the region
    assert np.all(first + m <= np.maximum(np.minimum(now, cap), first))  # never past the store or its capacity
    avail = np.maximum(np.minimum(now, cap) - first, 0)
    short = m < avail
    assert np.all(m[short] + nres[short] == bound[short])          # only the bound cuts a sweep short
