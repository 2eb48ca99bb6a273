"""DarkShower: drop-in for PETITE's ``DarkShower`` with the dark-vector
production on the GPU.

Same constructor and methods as src/PETITE/dark_shower.py:66-849 of the
reference.  ``generate_dark_shower`` takes an SM shower (a list of Particle,
a ParticleBatch, or a primary to shower first) and returns
``(SM_shower, dark_vectors)``.  ``generate_dark_showers`` reweights a whole
device-resident ParticleBatch in one call and returns a ``DarkBatch``.

Random numbers: each (SM particle, dark process) pair draws from its own
Philox stream derived from the particle's stream id and the process code, so
the vectors do not depend on batch composition or launch shapes.  Particles
passed as a list get stream id = their list position.
"""
import numpy as np
import torch

from .batch import ParticleBatch
from . import _lib
from .constants import (TWO_BODY_CODE, PID_DARK_VECTOR, all_process_code, alpha_em, dark_process_codes,
                        dimensionalities_dark, mass_dict,
                        meson_twobody_branchingratios)
from .dark_tables import DarkTables
from .device import DeviceContext
from .particle import Particle
from .shower import Shower

_DARK_NAMES = {8: "DarkBrem", 9: "DarkAnn", 10: "DarkComp", 11: "DarkMuonBrem", TWO_BODY_CODE: "TwoBody_BSMDecay"}
_MESONS = (111, 221, 331)


class DarkBatch:
    """Device-resident dark vectors: one entry per (SM particle, active process)
    with a positive weight; ``valid`` marks the entries that hold a vector
    (produce_bsm_particle returned a Particle)."""

    def __init__(self, out, n, stats, active, bound_electron=True, mV=None):
        self.n = int(n)
        self._out = out
        self.stats = stats
        self.active = list(active)
        self.bound_electron = bound_electron
        self.mV = mV

    def __len__(self):
        return self.n

    def tensor(self, key):
        return self._out[key][: self.n]

    @property
    def p(self):
        return self.tensor("p")

    @property
    def r(self):
        return self.tensor("r")[:, :3]

    @property
    def weight(self):
        return self.tensor("r")[:, 3]

    @property
    def parent(self):
        return self.tensor("parent")

    @property
    def process(self):
        return self.tensor("meta") & 0xFF

    @property
    def valid(self):
        return (self.tensor("meta") & 0x100) != 0

    @property
    def yields(self):
        """Sum of vector weights per process (float64, accumulated on the device)."""
        return dict(zip(("DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem", "TwoBody_BSMDecay"), self.stats["yield"]))

    def order(self):
        """Indices of the valid entries in the reference's order: SM record
        order, then the active-process order."""
        parent = self.parent.cpu().numpy().astype(np.int64)
        proc = self.process.cpu().numpy()
        rank = np.zeros(16, dtype=np.int64)
        for i, c in enumerate(self.active):
            rank[c] = i
        ok = self.valid.cpu().numpy()
        idx = np.nonzero(ok)[0]
        return idx[np.lexsort((rank[proc[idx]], parent[idx]))]

    def to_particles(self, sm_particles):
        """Particle objects of the vectors (dark_shower.py:786-804); sm_particles
        is the SM record as a list of Particle, indexed like the store slots."""
        idx = self.order()
        p = self.p.cpu().numpy().astype(np.float64)
        r = self.tensor("r").cpu().numpy().astype(np.float64)
        parent = self.parent.cpu().numpy()
        proc = self.process.cpu().numpy()
        out = []
        for e in idx:
            par = sm_particles[int(parent[e])].get_ids()
            name = _DARK_NAMES[int(proc[e])]
            if self.bound_electron and name in ("DarkAnn", "DarkComp"):
                name += "_bound"
            k = 1 if name == "TwoBody_BSMDecay" else 0  # dark_shower.py:838 vs 790
            ids = {"PID": PID_DARK_VECTOR, "parent_PID": par["PID"], "ID": 2 * par["ID"] + k,
                   "parent_ID": par["ID"], "generation_number": par["generation_number"] + 1,
                   "generation_process": name, "weight": float(r[e, 3])}
            if self.mV is not None:
                # TwoBody_BSMDecay sets m_V (dark_shower.py:839); produce_bsm_particle sets none and
                # Particle infers it from the float64 vector, rounded to 6 decimals (particle.py:125-131),
                # which the float32 vector here cannot reproduce
                ids["mass"] = float(self.mV) if k else round(float(self.mV), 6)
            out.append(Particle(p[e], r[e, :3], ids))
        return out


class DarkShower(Shower):
    """A class to reprocess an existing EM shower to generate dark photons
    (src/PETITE/dark_shower.py:66)."""

    def __init__(self, dict_dir, target_material, min_energy, mV_in_GeV, mode="exact", maxF_fudge_global=1,
                 max_n_integrators=int(1e4), kinetic_mixing=1.0, Zeff=29.508, bound_electron=True, g_e=None,
                 active_processes=None, fast_MCS_mode=True, rescale_MCS=1, seed=None, below="clamp",
                 masses="shipped"):
        """The reference's constructor (dark_shower.py:69-84).  Extra: below
        ("clamp" or the reference's "wrap") for m_V under the lightest mass,
        and masses: "shipped" picks among the reference's six trained masses,
        "all" also among the masses made by petite_amd.tablegen.dark_masses."""
        super().__init__(dict_dir, target_material, min_energy, maxF_fudge_global=maxF_fudge_global,
                         max_n_integrators=max_n_integrators, fast_MCS_mode=fast_MCS_mode, seed=seed,
                         rescale_MCS=rescale_MCS)
        if mode not in ("exact", "approx"):
            raise Exception("Mode not valid. Chose exact or approx.")
        self.active_processes = list(dark_process_codes if active_processes is None else active_processes)
        self.kinetic_mixing = kinetic_mixing
        self.bound_electron = bound_electron
        self.Zeff = Zeff
        self.g_e = g_e if g_e is not None else kinetic_mixing * np.sqrt(4 * np.pi * alpha_em)
        self._dark = DarkTables(self._tables, mV_in_GeV, mode=mode, Zeff=Zeff, bound_electron=bound_electron,
                                kinetic_mixing=kinetic_mixing, below=below, masses=masses)
        self._dark.eps2 = self.g_e ** 2 / (4 * np.pi * alpha_em)
        self._mV = self._dark.mV
        self._mV_estimator = self._dark.mV_est
        self._mV_list = list(self._dark.mV_list)
        self._resonant_annihilation_energy = self._dark.E_res
        self._compton_threshold_energy = self._dark.E_C
        self._minimum_calculable_dark_energy = self._dark.min_calc_dark

    # ----------------------------------------------------------- accessors --
    def get_mV(self):
        return self._mV

    def get_mV_list(self):
        return self._mV_list

    def get_DarkBremXSec(self):
        return self._dark.xsec["DarkBrem"]

    def get_DarkAnnXSec(self):
        return self._dark.xsec["DarkAnn"]

    def get_DarkCompXSec(self):
        return self._dark.xsec["DarkComp"]

    def get_DarkMuonBremXSec(self):
        return self._dark.xsec["DarkMuonBrem"]

    def _NSigmaDarkBrem(self, E):
        return self._dark.ns["DarkBrem"](E)

    def _NSigmaDarkAnn(self, E):
        return self._dark.ns["DarkAnn"](E)

    def _NSigmaDarkComp(self, E):
        return self._dark.ns["DarkComp"](E)

    def _NSigmaDarkMuonBrem(self, E):
        return self._dark.ns["DarkMuonBrem"](E)

    def _device_active(self):
        return [TWO_BODY_CODE if p == "TwoBody_BSMDecay" else all_process_code[p] for p in self.active_processes
                if p in all_process_code or p == "TwoBody_BSMDecay"]

    def device_context(self, MS=True):
        key = bool(MS)
        if key not in self._ctx:
            ctx = DeviceContext(self._tables, self.min_energy, maxF_fudge=self._maxF_fudge_global,
                                max_trials=int(self._max_n_integrators) * self._neval, MS=MS,
                                rescale_MCS=self._MCS_rescale_factor, dark=self._dark)
            self._train_flat_maps(ctx)
            ctx.dark_setup(self._device_active())
            self._ctx[key] = ctx
        return self._ctx[key]

    def _train_flat_maps(self, ctx):
        """Adapt on the GPU the maps of a generated mass (DarkTables starts
        them flat on the generator's domains), with the reference's iteration
        counts (all_processes.py:1085-1096), once per DarkShower: the trained
        nodes replace the flat ones for every later context.  The float64
        atomics of the accumulation make the nodes reproducible to rounding,
        not bit for bit, between runs."""
        from .dark_tables import MAP_ITERATIONS
        src = self._dark.map_source
        for P in [P for P, how in src.items() if how == "untrained"]:
            nodes = ctx.train_map(P, n_iter=MAP_ITERATIONS[P])
            E, _, cnt = self._dark.grids[P]
            self._dark.grids[P] = (E, nodes.ravel(), cnt)
            src[P] = "trained"

    def dark_yield(self, record, mV, eps2=1.0):
        """eps^2 * Y(m_V) of a shower record, differentiable in m_V and eps^2
        (torch autograd; see petite_amd.dark_yield)."""
        from .dark_yield import DarkYield
        if getattr(self, "_yield_model", None) is None:
            self._yield_model = DarkYield(self._tables, self.min_energy, active_processes=self.active_processes,
                                          Zeff=self.Zeff, ctx=self.device_context())
        return self._yield_model(record, mV, eps2)

    # ------------------------------------------------------------- weights --
    def GetBSMWeights(self, particle, process):
        """Dark weight of one SM particle (dark_shower.py:595-647)."""
        if isinstance(particle, (list, np.ndarray, tuple)):
            PID, energy_initial = particle
        else:
            PID, energy_initial = particle.get_ids()["PID"], particle.get_p0()[0]
        PID = int(PID)
        if PID in _MESONS:
            if process != "TwoBody_BSMDecay":
                return 0.0
            m_meson = particle.get_ids()["mass"] if isinstance(particle, Particle) else mass_dict[PID]
            r = self._mV / m_meson
            if r >= 1.0:
                return 0.0
            return 2 * self.kinetic_mixing ** 2 * (1.0 - r ** 2) ** 3 * meson_twobody_branchingratios[PID]
        if process not in ("DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"):
            return 0.0
        return float(self._dark.bsm_weight(np.array([PID]), np.array([float(energy_initial)]), process)[0])

    # ------------------------------------------------------------ sampling --
    def draw_dark_sample(self, Einc, LU_Key=-1, process="DarkBrem", VB=False, sid=None):
        """One accepted point of a dark process's accept/reject draw at energy
        Einc (dark_shower.py:649-704), or None when no sample is found within
        max_n_integrators * neval candidates.  The map is the one at LU_Key =
        argmin|E_i - Einc| + 1 (LU_Key is ignored, as draw_sample's is).  With
        VB the candidate count is appended.  Random numbers from Philox stream
        `sid` (default: the next call counter of this DarkShower)."""
        if process not in ("DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"):
            raise Exception("Your process is not in the list")
        s = self._seed_counter() if sid is None else int(sid)
        ctx = self.device_context()
        x, trials, w, rc = ctx.draw(np.array([all_process_code[process]], dtype=np.int32),
                                    np.array([float(Einc)]), np.array([s], dtype=np.uint64), self._seed)
        t = int(trials[0])
        if rc == _lib.PETITE_ERR_NO_SAMPLE or t <= 0:
            print("No Sample Found", process, Einc)
            return None
        xs = x[0, : dimensionalities_dark[process]].double().cpu().numpy()
        return np.concatenate([xs, [t]]) if VB else xs

    def draw_pe_sample(self, sid=None):
        """The bound electron's momentum [GeV] from its 1s momentum density,
        effective charge Zeff (dark_shower.py:706-719), on the GPU."""
        from .device import draw_pe
        s = self._seed_counter() if sid is None else int(sid)
        pe = float(draw_pe([s], self._seed, self.Zeff)[0])
        if pe < 0.0:
            raise RuntimeError("draw_pe_sample: no sample in 65536 accept/reject pairs")
        return pe

    def produce_bsm_particle(self, p_original, process, weight=None, VB=False, sid=None):
        """The dark vector an SM particle makes by `process`
        (dark_shower.py:721-804): interaction energy from the drate tables,
        scattering over the distance and energy loss, the threshold shortcuts,
        draw_dark_sample and the process's kinematics, on the GPU
        (petite_dark_reweight of a one-particle record).  Returns a Particle,
        or None where the reference returns None (no sample, a drate row that
        sums to zero).  The vector's weight is weight (default GetBSMWeights)
        times the particle's weight; vectors are produced where GetBSMWeights
        is positive (where generate_dark_shower asks for them), otherwise
        ValueError.  The pair's random numbers come from the Philox stream
        derived from `sid` (default: the next call counter) and the process."""
        if process not in ("DarkBrem", "DarkAnn", "DarkComp", "DarkMuonBrem"):
            raise Exception("Your process is not in the list")
        if process not in self.active_processes:
            raise ValueError(f"produce_bsm_particle: {process} is not among this DarkShower's active processes")
        wg_dev = self.GetBSMWeights(p_original, process)
        if not wg_dev > 0.0:
            raise ValueError(f"produce_bsm_particle: GetBSMWeights is 0 for this particle and {process}; vectors "
                             "are produced where it is positive")
        store = self._store_from_particles([p_original])
        store["sid"][0] = int(np.uint64(self._seed_counter() if sid is None else int(sid)).view(np.int64))
        ctx = self.device_context()
        out, n_e, stats = ctx.dark_reweight(store, 1, self._seed)
        code = all_process_code[process]
        meta = out["meta"][:n_e].cpu().numpy()
        hit = np.nonzero((meta & 0xFF) == code)[0]
        if len(hit) == 0 or not (meta[hit[0]] & 0x100):
            return None
        e = int(hit[0])
        p = out["p"][e].double().cpu().numpy()
        w = float(out["r"][e, 3])
        if weight is not None:
            w = w / wg_dev * float(weight) if wg_dev > 0 else 0.0
        ids = p_original.get_ids()
        name = process + ("_bound" if self.bound_electron and process in ("DarkAnn", "DarkComp") else "")
        d = {"PID": PID_DARK_VECTOR, "parent_PID": ids["PID"], "ID": 2 * ids["ID"], "parent_ID": ids["ID"],
             "generation_number": ids["generation_number"] + 1, "generation_process": name, "weight": w,
             "mass": round(float(self._mV), 6)}
        if VB:
            print(f"{process}: vector from {p_original.get_ids()['PID']} at E0 = {p_original.get_p0()[0]}")
        return Particle(p, np.asarray(p_original.get_rf(), dtype=np.float64), d)

    # ------------------------------------------------------------- records --
    @staticmethod
    def _store_from_particles(particles):
        """Device store of a list of Particle (slot = list position, stream id = position)."""
        n = len(particles)
        dev = torch.device("cuda")
        p0 = np.zeros((n, 4), dtype=np.float32)
        pf = np.zeros((n, 4), dtype=np.float32)
        r0 = np.zeros((n, 4), dtype=np.float32)
        rf = np.zeros((n, 4), dtype=np.float32)
        pid = np.zeros(n, dtype=np.int32)
        for i, q in enumerate(particles):
            ids = q.get_ids()
            p0[i] = q.get_p0()
            pf[i] = q.get_pf()
            r0[i, :3] = q.get_r0()
            r0[i, 3] = ids["weight"]
            rf[i, :3] = q.get_rf()
            pid[i] = ids["PID"]
        t = lambda a: torch.from_numpy(a).to(dev)
        return {"p0": t(p0), "pf": t(pf), "r0": t(r0), "rf": t(rf), "pid": t(pid),
                "parent": torch.full((n,), -1, device=dev, dtype=torch.int32),
                "meta": torch.zeros(n, device=dev, dtype=torch.int32),
                "shower": torch.zeros(n, device=dev, dtype=torch.int32),
                "sid": torch.arange(n, device=dev, dtype=torch.int64)}

    def generate_dark_showers(self, record, seed=None, GlobalMS=True):
        """Dark vectors of a whole SM record (ParticleBatch or list of Particle) as a DarkBatch."""
        if isinstance(record, ParticleBatch):
            store, n = record._store, record.n
        else:
            store, n = self._store_from_particles(record), len(record)
        ctx = self.device_context(MS=GlobalMS)
        out, n_e, stats = ctx.dark_reweight(store, n, self._seed if seed is None else seed)
        return DarkBatch(out, n_e, stats, ctx.dark_active, self.bound_electron, self._mV)

    def generate_dark_shower(self, ExDir=None, SParams=None):
        """(SM shower, dark vectors) of an existing shower or of a new one from
        the primary SParams (dark_shower.py:806-849)."""
        if ExDir is None and SParams is None:
            print("Need an existing SM shower-file directory or SM incident particle to run dark shower")
            return None
        if ExDir is not None and isinstance(ExDir, str):
            if not ExDir.endswith(".npz"):
                raise NotImplementedError("ExDir as a file path: the reference's pickled Particle lists are not "
                                          "loaded; pass a record written by ParticleBatch.save (.npz), the list "
                                          "of Particle or a ParticleBatch")
            ExDir = ParticleBatch.load(ExDir)
        if ExDir is not None and isinstance(ExDir, ParticleBatch):
            dark = self.generate_dark_showers(ExDir)
            sm = ExDir.to_particles()
            return sm, dark.to_particles(sm)
        if ExDir is not None and isinstance(ExDir, list):
            sm = ExDir
        elif isinstance(SParams, Particle):
            sm = self.generate_shower(SParams)
        else:
            raise ValueError("Provided SParams must be a `Particle' class object")
        if len(sm) == 0:
            return sm, []
        dark = self.generate_dark_showers(sm)
        return sm, dark.to_particles(sm)
