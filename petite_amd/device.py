"""Device context: one libpetite_hip context per (material, table set, cuts).

Device memory for inputs and outputs is allocated with PyTorch (ROCm) and
handed to the C ABI as raw pointers; the context holds the table copies.
"""
import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from .constants import all_process_code, mass_dict
from .tables import MaterialTables, build_blobs

MAXF_SEED = 0x5EEDF00D
MAXF_SAMPLES = 32768  # ~ n_trials (100) x neval (300) of utilities/find_maxes.py:206-210


VERTEX_DTYPE = np.dtype([("E", "f8"), ("px", "f8"), ("py", "f8"), ("pz", "f8"), ("x", "f8"), ("y", "f8"),
                         ("z", "f8"), ("weight", "f8"), ("pid", "i4"), ("generation", "i4"), ("sid", "u8")])
STORE_KEYS = ("p0", "r0", "pf", "rf", "pid", "parent", "meta", "shower", "sid")


def _empty_store(capacity, dev):
    return {k: torch.zeros((capacity, 4) if k in ("p0", "r0", "pf", "rf") else (capacity,), device=dev,
                           dtype=torch.float32 if k in ("p0", "r0", "pf", "rf") else
                           (torch.int64 if k == "sid" else torch.int32)) for k in STORE_KEYS}


def _store_struct(store):
    return _lib.Store(*(store[k].data_ptr() for k in STORE_KEYS), int(store["pid"].shape[0]))


def draw_pe(sid, seed, Zeff):
    """petite_draw_pe: bound-electron momenta [GeV] for the given stream ids
    (-1 where no sample came within 65536 pairs)."""
    _require_gpu()
    lib = _lib.load()
    dev = torch.device("cuda")
    s_t = torch.as_tensor(np.asarray(sid, dtype=np.uint64).view(np.int64), device=dev)
    out = torch.empty(s_t.numel(), device=dev, dtype=torch.float64)
    _lib.check(lib.petite_draw_pe(s_t.data_ptr(), s_t.numel(), int(seed), float(Zeff), out.data_ptr(),
                                  _stream_handle()), "petite_draw_pe")
    return out


def _stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("petite_amd needs a HIP device (MI355X); no GPU is visible")


class DeviceContext:
    def __init__(self, tables: MaterialTables, min_energy, maxF_fudge=1.0, max_trials=int(1e4) * 300, MS=True,
                 rescale_MCS=1.0, dark=None, maxf_samples=MAXF_SAMPLES, mT=None):
        """mT: the target mass the integrands see (default A_T, the event_info
        of draw_sample, shower.py:435); the cross-section tables were made
        with the nuclear mass (utilities/find_maxes.py:88), which a context
        that regenerates them passes here."""
        _require_gpu()
        self.lib = _lib.load()
        self.tables = tables
        self.dark = dark
        desc, b64, b32 = build_blobs(tables, dark)
        desc.min_energy = float(min_energy)
        desc.rescale_MCS = float(rescale_MCS)
        desc.maxF_fudge = float(maxF_fudge)
        desc.max_trials = int(max_trials)
        desc.MS = 1 if MS else 0
        if mT is not None:
            desc.mT = float(mT)
        if dark is not None:
            desc.mV = float(dark.mV)
            desc.Zeff = float(dark.Zeff)
        self.desc = desc
        self._b64, self._b32 = b64, b32
        ctx = C.c_void_p()
        _lib.check(self.lib.petite_ctx_create(C.byref(desc), b64.ctypes.data, b64.size, b32.ctypes.data, b32.size,
                                              C.byref(ctx)), "petite_ctx_create")
        self.ctx = ctx
        # launch sizes tuned on this kind of device (petite_amd.autotune), unless
        # PETITE_TUNING=default asks for the built-in ones
        self.tuning_source = "default"
        if os.environ.get("PETITE_TUNING", "") != "default":
            from .autotune import load
            pt = load()
            if pt is not None:
                self.set_tuning(pt)
                self.tuning_source = "table"
        self.maxf = {}
        self.sigma_est = {}
        for P, code in all_process_code.items():
            if desc.grid[code].n_E > 0:
                self.scan_maxf(P, maxf_samples)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.petite_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on=True):
        _lib.check(self.lib.petite_set_timing(self.ctx, 1 if on else 0), "petite_set_timing")

    def set_tuning(self, tuning=None):
        """Launch and queue sizes of run_showers (dict of PetiteTuning fields;
        missing or non-positive fields keep their defaults; None restores all
        defaults).  Returns the values in effect."""
        if tuning is None:
            _lib.check(self.lib.petite_set_tuning(self.ctx, None), "petite_set_tuning")
        else:
            t = _lib.Tuning.from_dict(tuning) if isinstance(tuning, dict) else tuning
            _lib.check(self.lib.petite_set_tuning(self.ctx, C.byref(t)), "petite_set_tuning")
        return self.get_tuning()

    def get_tuning(self):
        t = _lib.Tuning()
        _lib.check(self.lib.petite_get_tuning(self.ctx, C.byref(t)), "petite_get_tuning")
        return t.as_dict()

    def scan_maxf(self, process, n_samples=MAXF_SAMPLES, seed=MAXF_SEED):
        code = all_process_code[process]
        n_E = self.desc.grid[code].n_E
        mx = np.zeros(n_E)
        mean = np.zeros(n_E)
        _lib.check(self.lib.petite_maxf_scan(self.ctx, code, int(n_samples), int(seed), mx.ctypes.data,
                                             mean.ctypes.data, _stream_handle()), "petite_maxf_scan")
        self.maxf[process] = mx
        self.sigma_est[process] = mean
        return mx, mean

    def tables_grid(self, process):
        """(E, nodes float32, node counts) of the process's map in this context."""
        if process in self.tables.grids:
            return self.tables.grids[process]
        if self.dark is not None and process in self.dark.grids:
            return self.dark.grids[process]
        raise KeyError(f"{process}: no map in this context")

    def set_grid(self, process, nodes):
        """Replace the process's map nodes (n_E * stride float32) and rescan max_F."""
        code = all_process_code[process]
        v = np.ascontiguousarray(nodes, dtype=np.float32).ravel()
        _lib.check(self.lib.petite_set_grid(self.ctx, code, v.ctypes.data, v.size), "petite_set_grid")
        self.scan_maxf(process)

    def train_map(self, process, **kw):
        """Train the process's map in this context (vegas_train.train_process)
        and rescan max_F; returns the new nodes (n_E, stride) float32."""
        from .vegas_train import train_process
        nodes = train_process(self, process, **kw)
        self.scan_maxf(process)
        return nodes

    def set_maxf(self, process, values):
        code = all_process_code[process]
        v = np.ascontiguousarray(values, dtype=np.float64)
        _lib.check(self.lib.petite_set_maxf(self.ctx, code, v.ctypes.data, v.size), "petite_set_maxf")
        self.maxf[process] = v.copy()

    def draw(self, proc, E, sid, seed, noalloc=False):
        """Batched draws; proc/E/sid are 1-D host arrays or cuda tensors.
        noalloc=True uses petite_vegas_draw_noalloc (no device allocation by
        the library; at most _lib.NOALLOC_MAX_DRAWS requests)."""
        dev = torch.device("cuda")
        proc_t = torch.as_tensor(np.asarray(proc, dtype=np.int32) if not torch.is_tensor(proc) else proc,
                                 device=dev, dtype=torch.int32).contiguous()
        E_t = torch.as_tensor(np.asarray(E, dtype=np.float64) if not torch.is_tensor(E) else E,
                              device=dev, dtype=torch.float64).contiguous()
        if torch.is_tensor(sid):
            sid_t = sid.to(dev).contiguous()
        else:
            sid_t = torch.as_tensor(np.asarray(sid, dtype=np.uint64).view(np.int64), device=dev)
        n = proc_t.numel()
        if not (E_t.numel() == n and sid_t.numel() == n):
            raise ValueError("draw: proc, E and sid must have the same length")
        x = torch.empty((n, 4), device=dev, dtype=torch.float32)
        trials = torch.empty(n, device=dev, dtype=torch.int32)
        w = torch.empty(n, device=dev, dtype=torch.float32)
        fn = self.lib.petite_vegas_draw_noalloc if noalloc else self.lib.petite_vegas_draw
        rc = fn(self.ctx, proc_t.data_ptr(), E_t.data_ptr(), sid_t.data_ptr(), n, int(seed), x.data_ptr(),
                trials.data_ptr(), w.data_ptr(), _stream_handle())
        if rc not in (_lib.PETITE_OK, _lib.PETITE_ERR_NO_SAMPLE):
            _lib.check(rc, "petite_vegas_draw_noalloc" if noalloc else "petite_vegas_draw")
        return x, trials, w, rc

    def sample_scattering(self, vertices, proc, seed):
        """petite_sample_scattering: vertices is a structured array of
        VERTEX_DTYPE, proc the SM process codes.  Returns (store dict of 2n
        child entries, trials (n,), rc)."""
        v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
        n = len(v)
        dev = torch.device("cuda")
        v_t = torch.from_numpy(v.view(np.uint8).copy()).to(dev)
        p_t = torch.as_tensor(np.asarray(proc, dtype=np.int32), device=dev)
        if p_t.numel() != n:
            raise ValueError("sample_scattering: one process per vertex")
        store = _empty_store(max(2 * n, 1), dev)
        S = _store_struct(store)
        trials = torch.zeros(max(n, 1), device=dev, dtype=torch.int32)
        rc = self.lib.petite_sample_scattering(self.ctx, v_t.data_ptr(), p_t.data_ptr(), n, int(seed), C.byref(S),
                                               trials.data_ptr(), _stream_handle())
        if rc not in (_lib.PETITE_OK, _lib.PETITE_ERR_NO_SAMPLE):
            _lib.check(rc, "petite_sample_scattering")
        return store, trials[:n], rc

    def propagate(self, store, n, seed):
        """petite_propagate over the first n records of a store dict (p0, r0,
        pid, meta, shower, sid filled in); pf and rf are written in place."""
        S = _store_struct(store)
        _lib.check(self.lib.petite_propagate(self.ctx, C.byref(S), int(n), int(seed), _stream_handle()),
                   "petite_propagate")

    def eval_points(self, proc, E, words):
        """J*f and x of single candidates given their Philox words (n x 4 uint32)."""
        dev = torch.device("cuda")
        proc_t = torch.as_tensor(np.asarray(proc, dtype=np.int32), device=dev)
        E_t = torch.as_tensor(np.asarray(E, dtype=np.float64), device=dev)
        w_np = np.ascontiguousarray(words, dtype=np.uint32).view(np.int32)
        words_t = torch.as_tensor(w_np, device=dev)
        n = proc_t.numel()
        x = torch.empty((n, 4), device=dev, dtype=torch.float64)
        w = torch.empty(n, device=dev, dtype=torch.float64)
        _lib.check(self.lib.petite_eval_points(self.ctx, proc_t.data_ptr(), E_t.data_ptr(), words_t.data_ptr(), n,
                                               x.data_ptr(), w.data_ptr(), _stream_handle()), "petite_eval_points")
        return x.cpu().numpy(), w.cpu().numpy()

    def run_showers(self, primaries, seed, capacity=None):
        """primaries: structured numpy array of PRIMARY_DTYPE.  Returns (store dict of tensors, n, stats, energy)."""
        from .batch import PRIMARY_DTYPE
        prim = np.ascontiguousarray(primaries, dtype=PRIMARY_DTYPE)
        n = len(prim)
        dev = torch.device("cuda")
        if n == 0:  # nothing to transport: an empty record (the C entry point rejects n = 0)
            store = {k: torch.empty((0, 4) if k in ("p0", "r0", "pf", "rf") else (0,), device=dev,
                                    dtype=torch.float32 if k in ("p0", "r0", "pf", "rf") else
                                    (torch.int64 if k == "sid" else torch.int32))
                     for k in ("p0", "r0", "pf", "rf", "pid", "parent", "meta", "shower", "sid")}
            return store, 0, _lib.Stats().as_dict(), torch.zeros((0, 4), device=dev, dtype=torch.float64), \
                _lib.PETITE_OK
        prim_t = torch.from_numpy(prim.view(np.uint8).copy()).to(dev)
        n_showers = int(prim["shower"].max()) + 1 if n else 0
        if capacity is None:
            # records per unit of E / cutoff: an electromagnetic shower leaves
            # about 1.5 (100k lead: 1.5, 10 GeV graphite: 0.7), a muon with its
            # knock-on electrons' showers about 4 (400 GeV in lead); a store
            # that fills up costs a second run
            e = np.maximum(prim["E"], 0.0)
            per = np.where(np.abs(prim["pid"]) == 13, 5.0, 2.5)
            capacity = int(np.sum(per * e / max(self.desc.min_energy, 1e-3)) + 64 * n + 1024)
        capacity = int(min(capacity, 2 ** 31 - 1))
        while True:
            store = {
                "p0": torch.empty((capacity, 4), device=dev, dtype=torch.float32),
                "r0": torch.empty((capacity, 4), device=dev, dtype=torch.float32),
                "pf": torch.empty((capacity, 4), device=dev, dtype=torch.float32),
                "rf": torch.empty((capacity, 4), device=dev, dtype=torch.float32),
                "pid": torch.empty(capacity, device=dev, dtype=torch.int32),
                "parent": torch.empty(capacity, device=dev, dtype=torch.int32),
                "meta": torch.empty(capacity, device=dev, dtype=torch.int32),
                "shower": torch.empty(capacity, device=dev, dtype=torch.int32),
                "sid": torch.empty(capacity, device=dev, dtype=torch.int64),
            }
            energy = torch.zeros((max(n_showers, 1), 4), device=dev, dtype=torch.float64)
            S = _lib.Store(*(store[k].data_ptr() for k in ("p0", "r0", "pf", "rf", "pid", "parent", "meta", "shower",
                                                            "sid")), capacity)
            st = _lib.Stats()
            rc = self.lib.petite_run_showers(self.ctx, prim_t.data_ptr(), n, int(seed), C.byref(S), energy.data_ptr(),
                                             C.byref(st), _stream_handle())
            if rc == _lib.PETITE_ERR_CAPACITY and capacity < 2 ** 31 - 1:
                capacity = int(min(max(2 * capacity, int(st.n_particles * 1.2)), 2 ** 31 - 1))
                del store
                continue
            if rc not in (_lib.PETITE_OK, _lib.PETITE_ERR_NO_SAMPLE):
                _lib.check(rc, "petite_run_showers")
            return store, int(st.n_particles), st.as_dict(), energy, rc

    # ------------------------------------------------------------ dark sector --
    def dark_setup(self, active_codes):
        """Upload the dark tables of the context's DarkTables for the given process codes."""
        from .dark_tables import build_dark_blob
        if self.dark is None:
            raise ValueError("dark_setup: the context was built without dark tables")
        desc, blob = build_dark_blob(self.dark, active_codes)
        self._dark_desc, self._dark_blob = desc, blob
        _lib.check(self.lib.petite_dark_setup(self.ctx, C.byref(desc), blob.ctypes.data, blob.size),
                   "petite_dark_setup")
        self.dark_active = list(active_codes)

    def dark_reweight(self, store, n_sm, seed, capacity=None):
        """Dark vectors of the first n_sm records of a particle store (dict of device tensors)."""
        dev = torch.device("cuda")
        n_sm = int(n_sm)
        if capacity is None:
            capacity = max(1, n_sm * max(1, len(self.dark_active)))
        out = {"p": torch.empty((capacity, 4), device=dev, dtype=torch.float32),
               "r": torch.empty((capacity, 4), device=dev, dtype=torch.float32),
               "parent": torch.empty(capacity, device=dev, dtype=torch.int32),
               "meta": torch.empty(capacity, device=dev, dtype=torch.int32)}
        S = _lib.Store(*(store[k].data_ptr() for k in ("p0", "r0", "pf", "rf", "pid", "parent", "meta", "shower",
                                                        "sid")), int(store["pid"].shape[0]))
        O = _lib.DarkOut(out["p"].data_ptr(), out["r"].data_ptr(), out["parent"].data_ptr(), out["meta"].data_ptr(),
                         capacity)
        st = _lib.DarkStats()
        _lib.check(self.lib.petite_dark_reweight(self.ctx, C.byref(S), n_sm, int(seed), C.byref(O), C.byref(st),
                                                 _stream_handle()), "petite_dark_reweight")
        return out, int(st.n_entries), st.as_dict()

