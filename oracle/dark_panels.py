"""Panel quadrature of the dark weights and rates (oracle, float64; test
infrastructure only).

The reference builds DarkShower's track-length weight arrays
(construct_*_weight_array, src/PETITE/dark_shower.py:337-399) and its 10-bin
interaction-energy rates (_d_rate_d_E_*, 454-533) with scipy.integrate.quad
over the survival factor of the SM interaction integrals
(src/PETITE/shower.py:298-354); oracle.dark.OracleDark restates that with
quad.  The device computes the same integrals with a fixed panel rule
(petite_dark_quadrature); this module restates that rule in numpy so the
device is checked integral by integral (to 1e-9), while the rule itself is
checked against quad (oracle.dark: 5e-4 on weights, 2e-4 on rates) and the
caches shipped with the reference:

* the integrand n sigma(E) / (dE/dx) * exp(-sum_c [II_c(Ei) - II_c(E)] /
  (dE/dx)), with II_c the running integral of channel c's n sigma, exact for
  its piecewise-linear interpolant (trapezoids over the union of its nodes and
  the grid it is kept on);
* panels: geometric edges Ei - d, d = (scale / 8) 1.35^k, scale = 10 mfp(Ei)
  dE/dx / 10, kept where lo < Ei - d < hi, plus every table node strictly
  inside (lo, hi); 8-point Gauss-Legendre on each panel.

Pinned by: tests/test_dark_tables.py (quad, shipped caches).
"""
import numpy as np

from .constants import cmtom

GL_X, GL_W = np.polynomial.legendre.leggauss(8)

# survival channels per species (shower.py:322-354)
SURVIVAL = {11: ("Brem", "Moller"), -11: ("Brem", "Ann", "Bhabha"), 13: ("MuonBrem", "MuonE"),
            -13: ("MuonBrem", "MuonE")}
FAMILIES = (("brem_elec", "DarkBrem", 11), ("brem_positron", "DarkBrem", -11), ("muon_brem", "DarkMuonBrem", 13),
            ("annihilation", "DarkAnn", -11))


def lin0(xg, yg, x):
    x = np.asarray(x, dtype=np.float64)
    return np.where((x < xg[0]) | (x > xg[-1]), 0.0, np.interp(x, xg, yg))


def cumint(xg, yg, nodes):
    """Integral from nodes[0] to each node of the piecewise-linear (xg, yg),
    zero outside [xg[0], xg[-1]] (interp1d with fill 0, as the reference's
    quad integrates it): trapezoids over the union of the nodes and the xg
    inside them, summed in order; a segment outside the table adds nothing."""
    nodes = np.asarray(nodes, dtype=np.float64)
    pts = np.union1d(nodes, xg[(xg > nodes[0]) & (xg < nodes[-1])])
    f = np.interp(pts, xg, yg)
    mid = 0.5 * (pts[1:] + pts[:-1])
    seg = np.where((mid < xg[0]) | (mid > xg[-1]), 0.0, 0.5 * (f[1:] + f[:-1]) * np.diff(pts))
    cum = np.concatenate([[0.0], np.cumsum(seg)])
    return np.interp(nodes, pts, cum)


def loss_tables(T):
    """{channel: (x, y, grid)} of the SM interaction integrals: each channel's
    n sigma table, integrated on its own grid, except MuonBrem, whose integral
    is the one of MuonE on the MuonBrem grid (shower.py:316)."""
    out = {}
    for P in ("Brem", "PairProd", "Ann", "Comp", "Moller", "Bhabha", "MuonE"):
        xg, yg = T.nsigma_tab[P]
        out[P] = (xg, yg, xg)
    mex, mey = T.nsigma_tab["MuonE"]
    out["MuonBrem"] = (mex, mey, T.nsigma_tab["MuonBrem"][0])
    return out


class Panels:
    """The panel rule on the tables of an oracle.dark.OracleDark."""

    def __init__(self, od):
        self.od = od
        self.T = od.T
        self.dEdx = self.T.dEdx
        self.loss = loss_tables(self.T)
        self.ii = {P: (grid, cumint(x, y, grid)) for P, (x, y, grid) in self.loss.items()}

    def _II(self, P, E):
        xg, yg = self.ii[P]
        return lin0(xg, yg, E)

    def survival(self, pid, E, Ei):
        E = np.asarray(E, dtype=np.float64)
        d = sum(self._II(P, Ei) - self._II(P, E) for P in SURVIVAL[pid])
        with np.errstate(over="ignore"):
            out = np.exp(-d / self.dEdx / cmtom)
        return np.where((d < 0.0) | (E > Ei), 0.0, out)

    def integrand(self, proc, pid, E, Ei):
        k = self.dEdx * cmtom
        n = self.od.NS[proc](np.asarray(E, dtype=np.float64))
        if proc != "DarkAnn":
            n = np.where(n < 1.0e-18, 0.0, n)
        return n / k * self.survival(pid, E, Ei)

    def mfp(self, pid, E):
        return self.T.mfp(pid, E)

    @staticmethod
    def panels(lo, hi, Ei, scale, nodes):
        if hi <= lo:
            return np.array([lo, hi])
        edges = [hi]
        d = scale / 8.0
        while Ei - d > lo:
            if Ei - d < hi:
                edges.append(Ei - d)
            d *= 1.35
        edges.append(lo)
        inner = nodes[(nodes > lo) & (nodes < hi)]
        return np.unique(np.concatenate([np.array(edges), inner]))

    def integrate(self, proc, pid, lo, hi, Ei, nodes):
        if hi <= lo:
            return 0.0
        scale = max(10.0 * self.mfp(pid, Ei) * self.dEdx, 1e-9)
        e = self.panels(lo, hi, Ei, scale / 10.0, nodes)
        a, b = e[:-1], e[1:]
        h = 0.5 * (b - a)
        c = 0.5 * (b + a)
        x = c[:, None] + h[:, None] * GL_X[None, :]
        f = self.integrand(proc, pid, x.ravel(), Ei).reshape(x.shape)
        return float(np.sum(h * (f @ GL_W)))

    def nodes(self, proc, pid):
        pts = [self.ii[P][0] for P in SURVIVAL[pid]]
        pts.append(10.0 ** self.od.NS[proc].lx)
        return np.unique(np.concatenate(pts))

    def energies(self, proc):
        od = self.od
        return {"DarkAnn": od.DAS[:, 0], "DarkMuonBrem": od.DMB[:, 0]}.get(proc, od.DBS[:, 0])

    def weight_array(self, proc, pid):
        Ei_all = self.energies(proc)
        nodes = self.nodes(proc, pid)
        W = np.array([self.integrate(proc, pid, Ei_all[0], Ei, Ei, nodes) for Ei in Ei_all])
        return np.column_stack([Ei_all, W])

    def drate(self, proc, pid, Ei, lo_table):
        lo = max(Ei - 10.0 * self.mfp(pid, Ei) * self.dEdx, lo_table)
        edges = np.linspace(lo, Ei, 11)
        nodes = self.nodes(proc, pid)
        w = np.array([self.integrate(proc, pid, edges[i], edges[i + 1], Ei, nodes) for i in range(10)])
        return np.column_stack([0.5 * (edges[1:] + edges[:-1]), w])

    def tables(self, use_cache=True):
        """(W, D) in DarkTables' layout: W[family] = (n, 2) energies and
        weights; D[family] = (energies, rows (n, 10, 2), counts).  The brem and
        muon-brem families come from the shipped caches where they cover the
        (mass, material) pair (dark_shower.py:402-420), the bound-electron
        annihilation is always computed (450-451)."""
        od = self.od
        z, key, mat = od.z, od.key, self.T.material
        cached = use_cache and f"weights__brem_elec_weights__{key}__{mat}" in z.files
        W, D = {}, {}
        for name, proc, pid in FAMILIES:
            if cached and name != "annihilation":
                W[name] = np.array(z[f"weights__{name}_weights__{key}__{mat}"])
            else:
                W[name] = self.weight_array(proc, pid)
            ck = f"drate__{name}_drate__{key}__{mat}"
            if use_cache and name != "annihilation" and ck in z.files:
                arr = np.array(z[ck])
                D[name] = (np.array(z[f"drateE__{name}_drate__{key}__{mat}"]), arr[:, :10, :],
                           np.array(z[f"dratecnt__{name}_drate__{key}__{mat}"]))
            else:
                Es = self.energies(proc)
                D[name] = (Es, np.stack([self.drate(proc, pid, Ei, Es[0]) for Ei in Es]), np.full(len(Es), 10))
        return W, D
