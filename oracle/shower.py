"""SM shower transport (oracle, float64, scalar Python).

Restates Shower.propagate_particle (src/PETITE/shower.py:509-601),
sample_scattering (467-507), draw_sample (401-465, via oracle.vegas.draw),
the process choice and child bookkeeping of generate_shower (603-708),
Particle.lose_energy / rotation_matrix (src/PETITE/particle.py:143-153,
176-185) and the fast MCS (oracle.moliere).

Order of processing does not change results here: every particle owns a
Philox stream (oracle.philox), children's stream ids derive from the parent
stream, so the reference's FIFO sweep order and any other order give the same
shower.

Other species (mesons, neutrinos) end where they start: the reference
propagates none of them and its meson decay returns no products
(shower.py:640-653, particle.py:391-423).

Counting convention: one particle-step is one charged substep (shower.py
559-581, the one that ends in a hard interaction included), one closing
increment (583-598) or one photon flight (538-553).

Pinned by: exponential free-path law (KS), per-shower energy budget, and the
reference's multiplicity anchor (tests/test_oracle_physics.py).
"""
import math

import numpy as np

from . import philox, vegas
from .constants import PROCESS_PIDS, cmtom, mass_dict, m_electron
from .kinematics import sm_final_state
from .moliere import scatter


PROPAGATED = (22, 11, -11, 13, -13)


class OParticle:
    __slots__ = ("pid", "mass", "p0", "r0", "pf", "rf", "sid", "gen", "process", "parent", "parent_pid",
                 "weight", "k", "ended", "steps", "shower")

    def __init__(self, pid, mass, p0, r0, sid, gen=0, process="Input", parent=-1, parent_pid=22,
                 weight=1.0, k=0, shower=0):
        self.pid, self.mass = pid, mass
        self.p0 = np.array(p0, dtype=np.float64)
        self.r0 = np.array(r0, dtype=np.float64)
        self.pf = self.p0.copy()
        self.rf = self.r0.copy()
        self.sid = np.uint64(sid)
        self.gen, self.process, self.parent, self.parent_pid = gen, process, parent, parent_pid
        self.weight, self.k, self.ended, self.steps, self.shower = weight, k, False, 0, shower


def lose_energy(pf, mass, value):
    """Particle.lose_energy, particle.py:143-153."""
    E0, px, py, pz = pf
    p30 = math.sqrt(px * px + py * py + pz * pz)
    Eu = E0 - value
    if Eu <= mass:
        Eu = mass
    p3f = math.sqrt(Eu * Eu - mass * mass)
    if p3f > 0.0:
        return np.array([Eu, px / p30 * p3f, py / p30 * p3f, pz / p30 * p3f])
    return np.array([mass, 0.0, 0.0, 0.0])


def rotation_matrix(pf):
    """Particle.rotation_matrix, particle.py:176-185."""
    _, px, py, pz = pf
    th = math.acos(pz / math.sqrt(px * px + py * py + pz * pz))
    ph = math.atan2(py, px)
    ct, st, cp, sp = math.cos(th), math.sin(th), math.cos(ph), math.sin(ph)
    return np.array([[ct * cp, -sp, st * cp], [ct * sp, cp, st * sp], [-st, 0.0, ct]])


class Stats:
    def __init__(self):
        self.steps = 0
        self.draws = 0
        self.trials = 0
        self.maxf_violations = 0
        self.no_sample = 0
        self.deposit = {}
        self.final = {}
        self.dropped = {}
        self.atomic = {}

    def add(self, d, shower, v):
        d[shower] = d.get(shower, 0.0) + v


class OracleShower:
    """CPU restatement of Shower for one material/table set."""

    def __init__(self, tables, min_energy, seed=0, maxf=None, fudge=1.0, max_trials=3_000_000, MS=True,
                 rescale_MCS=1.0, vectorised_draw=True):
        self.T = tables
        self.min_energy = float(min_energy)
        self.seed = int(seed)
        self.maxf = {} if maxf is None else maxf  # (process, ie) -> max_F
        self.fudge = fudge
        self.max_trials = max_trials
        self.MS = MS
        self.rescale = rescale_MCS
        self.vectorised_draw = vectorised_draw
        self.stats = Stats()

    # -- tables -------------------------------------------------------------
    def max_f(self, process, ie):
        key = (process, ie)
        if key not in self.maxf:
            g = self.T.grids[process]
            ev = self.T.event_info(process, g.E[ie])
            self.maxf[key] = vegas.scan_maxf(self.T, process, ie, ev)[0]
        return self.maxf[key]

    def pme(self, p):
        return max(self.T.min_calc.get(p.pid, 0.0), self.min_energy, p.mass)

    # -- transport ------------------------------------------------------------
    def propagate(self, p):
        T, seed, st = self.T, self.seed, self.stats
        if p.pid not in PROPAGATED:
            # shower.py:640-653 propagates photons, e+- and mu+- only; a
            # neutrino is ended where it is, and a decaying meson's
            # decay_particle (particle.py:391-423) returns no products
            p.ended = True
            return
        pme = self.pme(p)
        if p.p0[0] < pme:
            p.ended = True
            return
        if p.pid == 22:
            mfp = T.mfp(22, p.p0[0])
            u = philox.block(seed, p.sid, 0, philox.PURPOSE_STEP)
            dist = mfp * math.log(1.0 / (1.0 - u[0]))
            v = p.p0[1:]
            p.rf = p.r0 + v / np.linalg.norm(v) * dist
            p.steps = 1
            st.steps += 1
            p.ended = True
            return
        dEdx = T.dEdx
        pf, rf = p.pf.copy(), p.rf.copy()
        k = 0
        hard = False
        delta_z = 0.0
        while not hard and pf[0] >= pme:
            mfp = T.mfp(p.pid, pf[0])
            w = philox.words(seed, p.sid, k, philox.PURPOSE_STEP)
            u = philox.to_uniform(w)
            delta_z = mfp / (6.0 + 14.0 * u[1])
            if u[0] > math.exp(-delta_z / mfp):
                hard = True
            else:
                E_before = pf[0]
                pf = lose_energy(pf, p.mass, dEdx * delta_z)
                st.add(st.deposit, p.shower, E_before - pf[0])
                pn = math.sqrt(pf[1] ** 2 + pf[2] ** 2 + pf[3] ** 2)
                if pn > 0.0:
                    rf = rf + pf[1:] / pn * delta_z
                if self.MS:
                    pf = scatter(pf, T.rho * (delta_z / cmtom), T.A, T.Z, self.rescale, mass_dict[p.pid], u[2], u[3],
                                 1.0 if int(w[3]) & 1 else -1.0)
            k += 1
        wc = philox.words(seed, p.sid, 0, philox.PURPOSE_CLOSE)
        uc = philox.to_uniform(wc)
        distC = uc[0]
        if pf[0] < pme:
            last = distC * delta_z
        else:
            mfp = T.mfp(p.pid, pf[0])
            last = mfp * math.log(1.0 / (1.0 + (math.exp(-delta_z / mfp) - 1) * distC))
        E_before = pf[0]
        pf = lose_energy(pf, p.mass, dEdx * last)
        st.add(st.deposit, p.shower, E_before - pf[0])
        pn = math.sqrt(pf[1] ** 2 + pf[2] ** 2 + pf[3] ** 2)
        if pn > 0.0:
            rf = rf + pf[1:] / pn * last
        if self.MS:
            pf = scatter(pf, T.rho * (last / cmtom), T.A, T.Z, self.rescale, m_electron, uc[1], uc[2],
                         1.0 if int(wc[2]) & 1 else -1.0)
        p.pf, p.rf = pf, rf
        p.steps = k + 1
        st.steps += k + 1
        p.ended = True

    # -- interaction ----------------------------------------------------------
    def choose_process(self, p):
        """Categorical choice over n*sigma at pf[0], shower.py:665-698."""
        chans = self.T.CHANNELS.get(p.pid)
        if chans is None:
            return None
        E = p.pf[0]
        w = np.array([float(self.T.nsigma(c, E)) for c in chans])
        SC = w.sum()
        if SC == 0.0 or np.isnan(SC):
            return None
        u = philox.block(self.seed, p.sid, 0, philox.PURPOSE_INTERACT)[0]
        c = np.cumsum(w)
        k = int(np.searchsorted(c, u * SC, side="right"))
        return chans[min(k, len(chans) - 1)]

    def sample_scattering(self, p, process):
        """shower.py:467-507; returns list of (pid, p_lab) or None."""
        E0 = p.pf[0]
        if E0 <= self.pme(p):
            return None
        RM = rotation_matrix(p.pf)
        g = self.T.grids[process]
        ie = g.lu_key(E0)
        mf = self.max_f(process, ie) * self.fudge
        ev = self.T.event_info(process, E0)
        x, trials, w = vegas.draw(g, process, E0, ev, mf, self.seed, p.sid, self.max_trials,
                                  vectorised=self.vectorised_draw)
        st = self.stats
        st.draws += 1
        st.trials += trials
        if x is None:
            st.no_sample += 1
            raise RuntimeError(f"No Sample Found {process} {E0} {ie}")
        if w > mf:
            st.maxf_violations += 1
        u_az = philox.block(self.seed, p.sid, 0, philox.PURPOSE_INTERACT)[1]
        a, b = sm_final_state(process, E0, p.mass, x, u_az)
        out = []
        for k, v in enumerate((a, b)):
            pid = PROCESS_PIDS[process][k] or p.pid
            out.append((pid, np.concatenate([[v[0]], RM @ v[1:]])))
        return out

    # -- driver ---------------------------------------------------------------
    def generate_shower(self, pid, p0, r0=(0.0, 0.0, 0.0), sid=0, shower=0, mass=None):
        """Returns the list of all particles (primary first), generate_shower semantics."""
        m = mass_dict[pid] if mass is None else mass
        prim = OParticle(pid, m, p0, r0, sid, shower=shower)
        out = [prim]
        if prim.p0[0] < self.min_energy:
            prim.ended = True
            return out
        st = self.stats
        i = 0
        while i < len(out):
            p = out[i]
            i += 1
            self.propagate(p)
            proc = self.choose_process(p)
            kids = self.sample_scattering(p, proc) if proc is not None else None
            if kids is None:
                st.add(st.final, shower, p.pf[0])
                continue
            if proc in ("Comp", "Ann", "Moller", "Bhabha", "MuonE"):
                st.add(st.atomic, shower, m_electron)
            for k, (cpid, pl) in enumerate(kids):
                if pl[0] > self.min_energy:
                    c = OParticle(cpid, mass_dict[cpid], pl, p.rf, philox.derive_sid(self.seed, p.sid, k, philox.PURPOSE_CHILD),
                                  gen=p.gen + 1, process=proc, parent=i - 1,
                                  parent_pid=p.pid, weight=p.weight, k=k, shower=shower)
                    out.append(c)
                else:
                    st.add(st.dropped, shower, pl[0])
        return out
