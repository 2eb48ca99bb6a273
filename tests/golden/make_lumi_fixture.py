"""Rows of the reference's radiative-return luminosity table, as test data.

    python tests/golden/make_lumi_fixture.py

The reference tabulates L(s, x) = integral over u of the transformed lepton
luminosity (src/PETITE/radiative_return.py:61-79) in lumi_integral_list
(src/PETITE/lumi_integral_data.py: rows [s, x, L], 50 values of s in
[1e-3, 10] GeV^2 by x = m_V^2 / s on a log/linear grid).  This script reads
that file as text (nothing is imported or executed) and writes a sample of
its rows with L > 0 to tests/golden/lumi_integral_rows.npz (arrays s, x, L),
which tests/test_oracle_dark_ann_lumi.py compares against the oracle's DarkAnn
integrand.  Only this script reads the reference tree, and only here.
"""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/PETITE/lumi_integral_data.py"
OUT = os.path.join(HERE, "lumi_integral_rows.npz")
NUM = r"[-+]?\d+\.\d+e[-+]\d+"


def main():
    text = open(SRC).read()
    rows = np.array([[float(a), float(b), float(c)]
                     for a, b, c in re.findall(r"\[\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\]", text)])
    rows = rows[(rows[:, 2] > 0.0) & (rows[:, 1] < 0.995)]
    rng = np.random.default_rng(20240611)
    # s up to 0.2 GeV^2 covers electrons to ~200 GeV on electrons at rest (s = 2 m (E + m))
    pick = rows[rows[:, 0] <= 0.2]
    sel = pick[rng.choice(len(pick), size=48, replace=False)]
    sel = sel[np.lexsort((sel[:, 1], sel[:, 0]))]
    np.savez(OUT, s=sel[:, 0], x=sel[:, 1], L=sel[:, 2], n_rows_in_table=len(rows))
    print(f"{len(sel)} of {len(rows)} rows -> {OUT}")


if __name__ == "__main__":
    main()
