"""Detector acceptance (oracle): transverse_position and detector_cut,
src/PETITE/shower.py:815-865.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Records are arrays rather than Particle objects: p (n, 4) = (E, px, py, pz)
and r (n, 4) = (x, y, z, weight), the production momentum and point, as the
reference reads them with get_p0()/get_r0() and get_ids()["weight"].  The
arithmetic follows the reference's numpy order: T = (z - z0)/pz, x0 + T px,
then the 2-norm over the transverse pair, compared strictly with the inner
and outer radii.  As in the reference, the detector positions are the z
values of the planes (shower.py:852 passes them as transverse_position's z),
the energy window is strict on both sides (shower.py:842-843), records
moving away from a plane still pass if their backward extrapolation lands
in the annulus, and pz = 0 gives no pass (NaN/inf radius).

Pinned by closed forms (tests/test_oracle_detector.py): straight lines from
the origin at known angles, shifted production points, the annulus
boundaries, the energy window, pz = 0 and pz < 0, and the weight sums.
"""
import numpy as np


def transverse_position(r0, p0, z):
    """[x, y] at plane(s) z of one record (shower.py:815-823)."""
    x0, y0, z0 = (float(v) for v in r0[:3])
    _, px, py, pz = (float(v) for v in p0[:4])
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        T = (z - z0) / pz
    return [x0 + T * px, y0 + T * py]


def pass_matrix(p, r, det_z, radius, inner_radius=0.0, energy_cut=None):
    """(in_window (n,), passes (n_det, n)) for record arrays p, r (float64)."""
    p = np.asarray(p, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    det_z = np.atleast_1d(np.asarray(det_z, dtype=np.float64))
    n = len(p)
    if energy_cut is not None:
        E = p[:, 0]
        win = (E < energy_cut[1]) & (E > energy_cut[0])
    else:
        win = np.ones(n, dtype=bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        T = (det_z[None, :] - r[:, 2:3]) / p[:, 3:4]
        xf = r[:, 0:1] + T * p[:, 1:2]
        yf = r[:, 1:2] + T * p[:, 2:3]
        rT = np.sqrt(xf * xf + yf * yf)
        ok = (rT > inner_radius) & (rT < radius)
    ok &= win[:, None]
    return win, ok.T


def detector_cut(p, r, det_z, radius, method="Sample", energy_cut=None, inner_radius=0.0):
    """detector_cut over record arrays (shower.py:825-865).  "Sample" returns
    index arrays into the energy-filtered records per plane (the reference
    returns the Particle objects at those positions)."""
    win, ok = pass_matrix(p, r, det_z, radius, inner_radius, energy_cut)
    det_z = np.atleast_1d(det_z)
    idx = np.nonzero(win)[0]
    if len(idx) == 0:
        return [[] for _ in det_z] if method == "Sample" else [0.0 for _ in det_z]
    sub = ok[:, idx]
    w = np.asarray(r, dtype=np.float64)[idx, 3]
    if method == "Sample":
        return [np.nonzero(sub[d])[0] for d in range(len(det_z))]
    if method == "SampleW":
        return sub
    if method == "Efficiency":
        return [np.sum(w[sub[d]]) / np.sum(w) for d in range(len(det_z))]
    if method == "TotalWeight":
        return [np.sum(w[sub[d]]) for d in range(len(det_z))]
    return None
