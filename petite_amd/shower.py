"""Shower: drop-in for PETITE's ``Shower`` with the transport on the GPU.

Same constructor and methods as src/PETITE/shower.py:98-708 of the reference.
``generate_shower`` returns the list of ``Particle`` objects of one shower;
``generate_showers`` runs many primaries in one batched device run and returns
a device-resident ``ParticleBatch``.  The per-particle work (propagation,
process choice, accept/reject draws, kinematics, children) runs in
libpetite_hip; there is no CPU path.

Random numbers: ``seed`` keys a Philox4x32-10 generator; the n-th call of
generate_shower (n = 0, 1, ...) uses primary stream id n, so a Shower built
with the same seed reproduces the same sequence of showers.
"""
import os
import warnings

import numpy as np
import torch

from . import _lib
from .batch import PRIMARY_DTYPE, ParticleBatch
from .constants import (all_process_code, cmtom, dimensionalities, mass_dict, process_code, process_PIDS,
                        target_information)
from .detector import detector_cut, transverse_position  # noqa: F401  (module functions of shower.py:815-865)
from .device import DeviceContext
from .particle import Particle
from .tables import MaterialTables, table_set_for

_CHARGED = (11, -11, 13, -13)
# cross-section table whose energies a dark map is trained on: the maps of the
# bound-electron processes were made on the free-electron energies
_DARK_SIGMA_TABLE = {"DarkAnn": "DarkAnn_free", "DarkComp": "DarkComp_free"}


def _seed_from(seed):
    if seed is None:
        return int.from_bytes(os.urandom(8), "little")
    return int(seed) & 0xFFFFFFFFFFFFFFFF


class Shower:
    """Representation of a shower (src/PETITE/shower.py:98)."""

    def __init__(self, dict_dir, target_material, min_energy, maxF_fudge_global=1, max_n_integrators=int(1e4),
                 fast_MCS_mode=True, seed=None, rescale_MCS=1):
        if not fast_MCS_mode:
            raise NotImplementedError("only the fast (Lynch-Dahl) MCS mode is implemented; the reference's Bethe "
                                      "mode raises TypeError for thick steps (moliere.py:244)")
        self.set_dict_dir(dict_dir)
        self.set_target_material(target_material)
        self.min_energy = min_energy
        self._seed = _seed_from(seed)
        self._n_calls = 0
        self._maxF_fudge_global = maxF_fudge_global
        self._max_n_integrators = max_n_integrators
        self._neval = 300  # utilities/generate_integrators.py:206-207
        self.set_material_properties()
        self.set_n_targets()
        self._tables = MaterialTables(target_material, table_set_for(dict_dir))
        self.set_cross_sections()
        self.set_MCS_rescale_factor(rescale_MCS)
        self._ctx = {}

    # ----------------------------------------------------------- accessors --
    def set_dict_dir(self, value):
        self._dict_dir = value

    def get_dict_dir(self):
        return self._dict_dir

    def set_target_material(self, value):
        self._target_material = value

    def get_target_material(self):
        return self._target_material

    def set_material_properties(self):
        ti = target_information[self.get_target_material()]
        self._ZTarget, self._ATarget, self._rhoTarget, self._dEdx = ti["Z_T"], ti["A_T"], ti["rho"], ti["dEdx"]

    def get_material_properties(self):
        return self._ZTarget, self._ATarget, self._rhoTarget, self._dEdx

    def set_n_targets(self):
        from .constants import m_proton_grams
        ZT, AT, rhoT, _ = self.get_material_properties()
        self._nTarget = rhoT / m_proton_grams / AT
        self._nElecs = self._nTarget * ZT

    def get_n_targets(self):
        return self._nTarget, self._nElecs

    def set_cross_sections(self):
        t = self._tables
        self._brem_cross_section = t.xsec["Brem"]
        self._pair_production_cross_section = t.xsec["PairProd"]
        self._annihilation_cross_section = t.xsec["Ann"]
        self._compton_cross_section = t.xsec["Comp"]
        self._moller_cross_section = t.xsec["Moller"]
        self._bhabha_cross_section = t.xsec["Bhabha"]
        self._muonbrem_cross_section = t.xsec["MuonBrem"]
        self._muone_cross_section = t.xsec["MuonE"]
        self._minimum_calculable_energy = dict(t.min_calc)
        self._Egamma_min = t.Eg_min
        self._Ee_min = t.Ee_min

    def get_brem_cross_section(self):
        return self._brem_cross_section

    def get_pairprod_cross_section(self):
        return self._pair_production_cross_section

    def get_annihilation_cross_section(self):
        return self._annihilation_cross_section

    def get_compton_cross_section(self):
        return self._compton_cross_section

    def get_moller_cross_section(self):
        return self._moller_cross_section

    def get_bhabha_cross_section(self):
        return self._bhabha_cross_section

    def get_muonbrem_cross_section(self):
        return self._muonbrem_cross_section

    def get_muone_cross_section(self):
        return self._muone_cross_section

    def set_MCS_rescale_factor(self, rescale_MCS):
        self._MCS_rescale_factor = rescale_MCS
        self._ctx = {}

    # n*sigma [1/cm] per process and species (shower.py:273-368)
    def _NSigma(self, process, E):
        return self._tables.nsigma(process, E)

    def _NSigmaBrem(self, E):
        return self._NSigma("Brem", E)

    def _NSigmaPP(self, E):
        return self._NSigma("PairProd", E)

    def _NSigmaAnn(self, E):
        return self._NSigma("Ann", E)

    def _NSigmaComp(self, E):
        return self._NSigma("Comp", E)

    def _NSigmaMoller(self, E):
        return self._NSigma("Moller", E)

    def _NSigmaBhabha(self, E):
        return self._NSigma("Bhabha", E)

    def _NSigmaMuonE(self, E):
        return self._NSigma("MuonE", E)

    def _NSigmaMuonBrem(self, E):
        return self._NSigma("MuonBrem", E)

    def _NSigmaElectron(self, E):
        return self._NSigmaBrem(E) + self._NSigmaMoller(E)

    def _NSigmaPhoton(self, E):
        return self._NSigmaPP(E) + self._NSigmaComp(E)

    def _NSigmaPositron(self, E):
        return self._NSigmaBrem(E) + self._NSigmaBhabha(E) + self._NSigmaAnn(E)

    def _NSigmaMuon(self, E):
        return self._NSigmaMuonBrem(E) + self._NSigmaMuonE(E)

    def get_mfp(self, particle):
        """Mean free path in metres (shower.py:370-389)."""
        if not isinstance(particle, Particle) and isinstance(particle, (list, np.ndarray, tuple)):
            PID, Energy = particle
        else:
            PID, Energy = particle.get_ids()["PID"], particle.get_pf()[0]
        if PID == 22:
            ns = self._NSigmaPhoton(Energy)
        elif PID == 11:
            ns = self._NSigmaElectron(Energy)
        elif PID == -11:
            ns = self._NSigmaPositron(Energy)
        elif abs(PID) == 13:
            ns = self._NSigmaMuon(Energy)
        else:
            raise ValueError(f"get_mfp: PID {PID} not propagated")
        ns = float(ns)
        return 1.0e12 if ns <= 0.0 else cmtom / ns

    def BF_positron_brem(self, Energy):
        b0, b1 = self._NSigmaBrem(Energy), self._NSigmaAnn(Energy)
        return b0 / (b0 + b1)

    def BF_photon_pairprod(self, Energy):
        b0, b1 = self._NSigmaPP(Energy), self._NSigmaComp(Energy)
        return b0 / (b0 + b1)

    # -------------------------------------------------------------- device --
    def train_maps(self, processes=None, n_iter=10, n_per_energy=1 << 16, alpha=0.5, seed=None, energies=None):
        """Retrain the VEGAS maps of the given processes (default: every process
        with a map) on the GPU, starting from the current ones, and rescan
        max_F (petite_amd.vegas_train).  The reference ships its maps trained
        with the vegas package (utilities/generate_integrators.py); a retrained
        map keeps the end points and changes only the draws' efficiency.
        energies="sigma" first puts each map on its process's cross-section
        energies (e.g. up to 400 GeV in the 400 GeV set, whose Brem/PairProd/
        MuonBrem/DarkBrem/DarkMuonBrem maps are missing; DarkAnn and DarkComp
        take their free-electron tables' energies); an array puts every map on
        those energies."""
        from .tables import regrid
        from .vegas_train import TRAIN_SEED
        if energies is not None:
            if processes is None or (isinstance(energies, str) and energies != "sigma"):
                raise ValueError('train_maps: energies needs the processes, and takes "sigma" or an array')
            dark = getattr(self, "_dark", None)
            for P in processes:
                if P in self._tables.grids:
                    grids, E_sig = self._tables.grids, lambda: self._tables.nsigma_tab[P][0]
                elif dark is not None and P in dark.grids:
                    grids, E_sig = dark.grids, lambda: dark.xsec[_DARK_SIGMA_TABLE.get(P, P)][:, 0]
                else:
                    raise ValueError(f"train_maps: {P} has no map")
                regrid(grids, P, E_sig() if isinstance(energies, str) else energies)
            for c in self._ctx.values():
                c.close()
            self._ctx = {}
        ctx = self.device_context()
        if processes is None:
            processes = [P for P in all_process_code if ctx.desc.grid[all_process_code[P]].n_E > 0]
        for P in processes:
            E, _, cnt = ctx.tables_grid(P)
            nodes = ctx.train_map(P, n_iter=n_iter, n_per_energy=n_per_energy, alpha=alpha,
                                  seed=TRAIN_SEED if seed is None else seed)
            grids = self._tables.grids if P in self._tables.grids else self._dark.grids
            grids[P] = (E, nodes.ravel(), cnt)
            for other in self._ctx.values():
                if other is not ctx:
                    other.set_grid(P, nodes)

    def device_context(self, MS=True):
        key = bool(MS)
        if key not in self._ctx:
            self._ctx[key] = DeviceContext(self._tables, self.min_energy, maxF_fudge=self._maxF_fudge_global,
                                           max_trials=int(self._max_n_integrators) * self._neval, MS=MS,
                                           rescale_MCS=self._MCS_rescale_factor)
        return self._ctx[key]

    def draw_samples(self, Einc, process="PairProd", seed=None, sid=None):
        """Batched draw_sample: one accepted point per energy (device tensors x, trials)."""
        E = np.atleast_1d(np.asarray(Einc, dtype=np.float64))
        n = len(E)
        code = all_process_code[process]
        if sid is None:
            sid = np.arange(n, dtype=np.uint64)
        ctx = self.device_context()
        x, trials, w, rc = ctx.draw(np.full(n, code, dtype=np.int32), E, sid, self._seed if seed is None else seed)
        if rc == _lib.PETITE_ERR_NO_SAMPLE:
            raise Exception("No Sample Found", process)
        return x[:, : dimensionalities[process]], trials

    def draw_sample(self, Einc, LU_Key=-1, process="PairProd", VB=False):
        """One accepted point (shower.py:401-465); the energy's own map is used (LU_Key ignored)."""
        sid = self._seed_counter()
        x, trials = self.draw_samples([Einc], process=process, sid=np.array([sid], dtype=np.uint64))
        xs = x[0].double().cpu().numpy()
        if VB:
            return np.concatenate([xs, [int(trials[0])]])
        return xs

    def _seed_counter(self):
        c = self._n_calls
        self._n_calls += 1
        return (1 << 63) | c

    def _particle_min_energy(self, p):
        ids = p.get_ids()
        return max(self._minimum_calculable_energy.get(ids["PID"], 0.0), self.min_energy, ids["mass"] or 0.0)

    def sample_scattering(self, p0, process, VB=False, sid=None):
        """The two particles of a hard interaction of p0 by `process`
        (shower.py:467-507): draw_sample at p0's final energy, the process's
        kinematics rotated onto p0's final direction, children at p0's final
        position with the reference's ids.  None at or below the particle's
        minimum energy.  The draw, the kinematic azimuth and the children's
        stream ids come from Philox stream `sid` (default: the next call
        counter of this Shower), on the GPU (petite_sample_scattering)."""
        E0 = float(p0.get_pf()[0])
        ids = p0.get_ids()
        if E0 <= self._particle_min_energy(p0):
            return None
        if process not in process_code:
            raise Exception("Your process is not in the list")
        code = process_code[process]
        from .device import VERTEX_DTYPE
        v = np.zeros(1, dtype=VERTEX_DTYPE)
        pf = np.asarray(p0.get_pf(), dtype=np.float64)
        rf = np.asarray(p0.get_rf(), dtype=np.float64)
        v["E"], v["px"], v["py"], v["pz"] = pf
        v["x"], v["y"], v["z"] = rf
        v["weight"] = float(ids["weight"])
        v["pid"] = int(ids["PID"])
        v["generation"] = int(ids["generation_number"])
        v["sid"] = self._seed_counter() if sid is None else int(sid)
        ctx = self.device_context()
        store, trials, rc = ctx.sample_scattering(v, [code], self._seed)
        t = int(trials[0])
        if rc == _lib.PETITE_ERR_NO_SAMPLE or t <= 0:
            raise Exception("No Sample Found", process, E0)
        if VB:
            print(f"{process}: sample found after {t} candidates")
        p = store["p0"][:2].double().cpu().numpy()
        pid = store["pid"][:2].cpu().numpy()
        out = []
        for k in range(2):
            d = {"PID": int(pid[k]), "parent_PID": ids["PID"], "ID": 2 * ids["ID"] + k, "parent_ID": ids["ID"],
                 "generation_number": ids["generation_number"] + 1, "generation_process": process,
                 "weight": ids["weight"], "mass": mass_dict[int(pid[k])]}
            out.append(Particle(p[k], rf, d))
        return out

    def propagate_particle(self, Part0, Losses=False, MS=False, sid=None):
        """Transport to the next hard interaction (shower.py:509-601), as
        generate_shower calls it: photons with Losses=False and MS=False (one
        exponential flight), charged leptons with Losses = the material's
        dE/dx [GeV/m] and MS = the multiple-scattering flag (substeps with
        energy loss, and scattering when MS).  Other combinations are not
        supported on the device (NotImplementedError).  Random numbers from
        Philox stream `sid` (default: the next call counter of this Shower)."""
        if Part0.get_ended() is True:
            Part0.set_rf(Part0.get_rf())
            return Part0
        if (np.linalg.norm(np.asarray(Part0.get_p0()) - np.asarray(Part0.get_pf())) > 0.0) \
                or (np.linalg.norm(np.asarray(Part0.get_r0()) - np.asarray(Part0.get_rf())) > 0.0):
            raise ValueError("propagate_particle() should only be called for a particle with pf = p0 and rf = r0 "
                             "and get_ended() == False")
        pid = int(Part0.get_ids()["PID"])
        if Part0.get_p0()[0] < self._particle_min_energy(Part0):
            Part0.set_ended(True)
            return Part0
        if pid == 22:
            if Losses or MS:
                raise NotImplementedError("propagate_particle: photons propagate with Losses=False, MS=False "
                                          "(generate_shower's call) on the device")
        elif pid in _CHARGED:
            if not Losses or abs(float(Losses) - self._tables.dEdx) > 1e-12 * self._tables.dEdx:
                raise NotImplementedError("propagate_particle: charged leptons propagate with Losses = the "
                                          f"material's dE/dx ({self._tables.dEdx} GeV/m) on the device")
        else:
            raise NotImplementedError(f"propagate_particle: PID {pid} is not propagated by generate_shower")
        from .device import _empty_store
        dev = torch.device("cuda")
        store = _empty_store(1, dev)
        rec = self._primary_record(Part0, 0, self._seed_counter() if sid is None else int(sid))
        store["p0"][0] = torch.tensor([rec["E"][0], rec["px"][0], rec["py"][0], rec["pz"][0]])
        store["r0"][0] = torch.tensor([rec["x"][0], rec["y"][0], rec["z"][0], rec["weight"][0]])
        store["pid"][0] = pid
        store["meta"][0] = int(Part0.get_ids()["generation_number"]) & 0xFFFF
        store["sid"][0] = int(np.asarray(rec["sid"][0]).view(np.int64))
        ctx = self.device_context(MS=bool(MS) if pid != 22 else True)
        ctx.propagate(store, 1, self._seed)
        Part0.set_pf(store["pf"][0].double().cpu().numpy())
        Part0.set_rf(store["rf"][0, :3].double().cpu().numpy())
        Part0.set_ended(True)
        return Part0

    # ------------------------------------------------------------- showers --
    @staticmethod
    def _primary_record(p0, shower, sid):
        pid = int(p0.get_ids()["PID"])
        m_ref = mass_dict.get(pid, 0.0)
        E, px, py, pz = [float(v) for v in np.asarray(p0.get_p0(), dtype=np.float64)]
        r = np.asarray(p0.get_r0(), dtype=np.float64)
        m_ids = p0.get_ids()["mass"]
        # the reference rounds an inferred mass to 6 decimals (particle.py:125-131): only a
        # larger mismatch (e.g. a massless charged primary) is reported
        if pid in _CHARGED and (m_ids is None or abs(float(m_ids) - m_ref) > 1e-6):
            warnings.warn(f"primary PID {pid} has mass {m_ids}; transporting it on shell with m = {m_ref} "
                          "(the reference's multiple scattering is NaN for a massless charged lepton)")
            pn = np.sqrt(px * px + py * py + pz * pz)
            pnew = np.sqrt(max(E * E - m_ref * m_ref, 0.0))
            if pn > 0:
                px, py, pz = px / pn * pnew, py / pn * pnew, pz / pn * pnew
        rec = np.zeros(1, dtype=PRIMARY_DTYPE)
        rec["E"], rec["px"], rec["py"], rec["pz"] = E, px, py, pz
        rec["x"], rec["y"], rec["z"] = r[0], r[1], r[2]
        rec["mass"] = m_ref
        rec["weight"] = float(p0.get_ids()["weight"])
        rec["pid"] = pid
        rec["shower"] = shower
        rec["sid"] = sid
        return rec

    def generate_showers(self, primaries, seed=None, GlobalMS=True, capacity=None):
        """Run many primaries in one device run; returns a ParticleBatch.

        primaries: list of Particle, or a structured array of PRIMARY_DTYPE."""
        if isinstance(primaries, np.ndarray) and primaries.dtype == PRIMARY_DTYPE:
            prim = primaries
        else:
            recs = [self._primary_record(p, i, i) for i, p in enumerate(primaries)]
            prim = np.concatenate(recs) if recs else np.zeros(0, dtype=PRIMARY_DTYPE)
        keep = prim["E"] >= self.min_energy
        ctx = self.device_context(MS=GlobalMS)
        store, n, stats, energy, rc = ctx.run_showers(prim, self._seed if seed is None else seed, capacity)
        if rc == _lib.PETITE_ERR_NO_SAMPLE:
            raise Exception("No Sample Found")
        batch = ParticleBatch(store, n, stats, energy, prim)
        batch.below_cut = ~keep
        return batch

    def generate_shower(self, p0, VB=False, GlobalMS=True):
        """All particles of one shower, primary first (shower.py:603-708)."""
        if VB:
            print("Starting shower, initial particle with ID Info")
            print(p0.get_ids())
            print("Initial four-momenta:")
            print(p0.get_p0())
        p0.set_ended(False)
        sid = self._n_calls
        self._n_calls += 1
        if p0.get_p0()[0] < self.min_energy:
            q = Particle(p0.get_p0(), p0.get_r0(), dict(p0.get_ids()))
            q.set_ended(True)
            return [q]
        rec = self._primary_record(p0, 0, sid)
        ctx = self.device_context(MS=GlobalMS)
        store, n, stats, energy, rc = ctx.run_showers(rec, self._seed)
        if rc == _lib.PETITE_ERR_NO_SAMPLE:
            raise Exception("No Sample Found")
        batch = ParticleBatch(store, n, stats, energy, rec)
        out = batch.to_particles()
        ids0 = dict(p0.get_ids())
        out[0].set_ids(ids0)
        return out
