"""ctypes binding of libpetite_hip.so (declarations: include/petite_hip.h).

The library is loaded from this package directory only; there is no CPU
fallback.  ``load()`` raises if the shared object is missing or if no HIP
device is present when a compute entry point is used.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpetite_hip.so")

NPROC = 12
PETITE_OK = 0
PETITE_ERR_ARG = 1
PETITE_ERR_HIP = 2
PETITE_ERR_NO_SAMPLE = 3
PETITE_ERR_CAPACITY = 4
PETITE_ERR_INTERNAL = 5

EXPORTED = ["petite_ctx_create", "petite_ctx_destroy", "petite_last_error", "petite_maxf_scan", "petite_set_maxf",
            "petite_vegas_draw", "petite_eval_points", "petite_run_showers", "petite_set_timing", "petite_dark_setup",
            "petite_dark_reweight", "petite_yield_setup", "petite_dark_yield", "petite_detector_cut",
            "petite_map_accumulate", "petite_set_grid", "petite_set_tuning", "petite_get_tuning", "petite_vegas_draw_noalloc", "petite_sample_scattering",
            "petite_propagate", "petite_draw_pe", "petite_dark_quadrature", "petite_yield_bound"]
NOALLOC_MAX_DRAWS = 16384
MAX_MASSES = 16
DARK_FAMILIES = ["brem_elec", "brem_positron", "muon_brem", "annihilation"]
DARK_DRATE_BINS = 10


class GridDesc(C.Structure):
    _fields_ = [("n_E", C.c_int32), ("ndim", C.c_int32), ("ninc", C.c_int32 * 4), ("E_off", C.c_int64),
                ("node_off", C.c_int64), ("stride", C.c_int64), ("lnE0", C.c_double), ("inv_dlog", C.c_double)]


class SigmaDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad", C.c_int32), ("E_off", C.c_int64), ("v_off", C.c_int64),
                ("lnE0", C.c_double), ("inv_dlog", C.c_double)]


class TableDesc(C.Structure):
    _fields_ = [("Z", C.c_double), ("A", C.c_double), ("rho", C.c_double), ("dEdx", C.c_double),
                ("mT", C.c_double), ("min_energy", C.c_double), ("Eg_min", C.c_double), ("Ee_min", C.c_double),
                ("rescale_MCS", C.c_double), ("maxF_fudge", C.c_double), ("min_calc", C.c_double * 5),
                ("mV", C.c_double), ("Zeff", C.c_double), ("max_trials", C.c_int64), ("MS", C.c_int32),
                ("pad", C.c_int32), ("grid", GridDesc * NPROC), ("sigma", SigmaDesc * NPROC)]


class Primary(C.Structure):
    _fields_ = [("E", C.c_double), ("px", C.c_double), ("py", C.c_double), ("pz", C.c_double),
                ("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("mass", C.c_double),
                ("weight", C.c_double), ("pid", C.c_int32), ("shower", C.c_int32), ("sid", C.c_uint64)]


class Vertex(C.Structure):
    _fields_ = [("E", C.c_double), ("px", C.c_double), ("py", C.c_double), ("pz", C.c_double),
                ("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("weight", C.c_double),
                ("pid", C.c_int32), ("generation", C.c_int32), ("sid", C.c_uint64)]


class Store(C.Structure):
    _fields_ = [("p0", C.c_void_p), ("r0", C.c_void_p), ("pf", C.c_void_p), ("rf", C.c_void_p),
                ("pid", C.c_void_p), ("parent", C.c_void_p), ("meta", C.c_void_p), ("shower", C.c_void_p),
                ("sid", C.c_void_p), ("capacity", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("n_steps", C.c_int64), ("n_draws", C.c_int64),
                ("n_trials", C.c_int64), ("n_maxf_viol", C.c_int64), ("n_no_sample", C.c_int64),
                ("n_iterations", C.c_int64), ("n_dropped", C.c_int64), ("ms_transport", C.c_double),
                ("ms_draw", C.c_double), ("ms_emit", C.c_double), ("launches_transport", C.c_int64),
                ("launches_draw", C.c_int64), ("launches_emit", C.c_int64),
                ("n_draws_proc", C.c_int64 * 12), ("n_trials_proc", C.c_int64 * 12),
                ("n_viol_proc", C.c_int64 * 12), ("ms_tail", C.c_double), ("launches_tail", C.c_int64),
                ("n_steps_tail", C.c_int64)]

    def as_dict(self):
        return {f: (float(getattr(self, f)) if t is C.c_double else
                    [int(v) for v in getattr(self, f)] if f.endswith("_proc") else int(getattr(self, f)))
                for f, t in self._fields_}


class Tuning(C.Structure):
    """PetiteTuning: launch and queue sizes of petite_run_showers (0 = default)."""
    _fields_ = [("batch_lanes", C.c_int32), ("yield_steps", C.c_int32), ("transport_waves", C.c_int32),
                ("emit_waves", C.c_int32), ("draw_waves", C.c_int32 * 8), ("min_group_lanes", C.c_int32),
                ("sweep_group", C.c_int32), ("group_requests", C.c_int32), ("small_sweep", C.c_int32),
                ("sweep_bound", C.c_int64),
                ("resume_cap", C.c_int64), ("small_yield_steps", C.c_int32), ("graph", C.c_int32),
                ("tail_sweep", C.c_int32), ("tail_waves", C.c_int32), ("tail_depth", C.c_int32),
                ("tail_budget", C.c_int32)]

    def as_dict(self):
        return {f: ([int(v) for v in getattr(self, f)] if f == "draw_waves" else int(getattr(self, f)))
                for f, _ in self._fields_ if f != "pad"}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k, v in d.items():
            if k == "draw_waves":
                for i, w in enumerate(v):
                    t.draw_waves[i] = int(w)
            elif k != "pad":
                setattr(t, k, int(v))
        return t


class QTable(C.Structure):
    _fields_ = [("off", C.c_int64), ("n", C.c_int32), ("pad", C.c_int32)]


class DarkQuadJob(C.Structure):
    """PetiteDarkQuadJob: one weight or drate job of petite_dark_quadrature."""
    _fields_ = [("ns", QTable), ("loss", QTable * 3), ("grid", QTable * 3), ("mfp", QTable * 3), ("nodes", QTable),
                ("energies", QTable), ("n_loss", C.c_int32), ("n_mfp", C.c_int32), ("mode", C.c_int32),
                ("zero_small", C.c_int32), ("lo", C.c_double), ("dEdx", C.c_double), ("out_off", C.c_int64)]


class BoundDesc(C.Structure):
    """PetiteBoundDesc: bound-electron inputs of the m_V-continuous yield."""
    _fields_ = [("Zeff", C.c_double), ("ne_conv", C.c_double), ("m_e", C.c_double), ("alpha", C.c_double),
                ("ann_floor", C.c_double), ("comp_floor", C.c_double), ("rel_step", C.c_double), ("dEdx", C.c_double),
                ("top_ann", C.c_double * 16), ("top_comp", C.c_double * 16), ("loss", QTable * 3),
                ("mfp", QTable * 3), ("nodes", QTable)]


class DarkDesc(C.Structure):
    _fields_ = [("w_n", C.c_int32 * 4), ("w_off", C.c_int64 * 4), ("d_n", C.c_int32 * 4), ("d_off", C.c_int64 * 4),
                ("comp_n", C.c_int32), ("n_active", C.c_int32), ("comp_off", C.c_int64), ("active", C.c_int32 * 8),
                ("eps2", C.c_double), ("E_res", C.c_double), ("E_C", C.c_double), ("min_brem", C.c_double),
                ("min_ann", C.c_double), ("min_comp", C.c_double), ("min_muon_brem", C.c_double),
                ("w_lnE0", C.c_double * 4), ("w_inv_dlog", C.c_double * 4), ("comp_l0", C.c_double),
                ("comp_inv_dl", C.c_double), ("meson_w", C.c_double * 3)]


class DarkOut(C.Structure):
    _fields_ = [("p", C.c_void_p), ("r", C.c_void_p), ("parent", C.c_void_p), ("meta", C.c_void_p),
                ("capacity", C.c_int64)]


class DarkStats(C.Structure):
    _fields_ = [("n_entries", C.c_int64), ("n_vectors", C.c_int64), ("n_draws", C.c_int64), ("n_trials", C.c_int64),
                ("n_no_sample", C.c_int64), ("n_zero_rate", C.c_int64), ("n_pe_fail", C.c_int64), ("pad", C.c_int64),
                ("yield_", C.c_double * 5), ("ms_select", C.c_double), ("ms_draw", C.c_double),
                ("ms_emit", C.c_double), ("launches", C.c_int64), ("ms_meson", C.c_double)]

    def as_dict(self):
        d = {f: (float(getattr(self, f)) if t is C.c_double else int(getattr(self, f)))
             for f, t in self._fields_ if f not in ("yield_", "pad")}
        d["yield"] = [float(v) for v in self.yield_]
        return d


class PetiteError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[petite {code}] {msg}")
        self.code = code


_lib = None


def load():
    """Load the in-tree libpetite_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m petite_amd.build`")
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
    lib.petite_ctx_create.argtypes = [C.POINTER(TableDesc), vp, i64, vp, i64, C.POINTER(vp)]
    lib.petite_ctx_create.restype = C.c_int
    lib.petite_ctx_destroy.argtypes = [vp]
    lib.petite_ctx_destroy.restype = None
    lib.petite_last_error.argtypes = []
    lib.petite_last_error.restype = C.c_char_p
    lib.petite_set_timing.argtypes = [vp, i32]
    lib.petite_set_timing.restype = C.c_int
    lib.petite_maxf_scan.argtypes = [vp, i32, i32, u64, vp, vp, vp]
    lib.petite_maxf_scan.restype = C.c_int
    lib.petite_set_maxf.argtypes = [vp, i32, vp, i32]
    lib.petite_set_maxf.restype = C.c_int
    lib.petite_vegas_draw.argtypes = [vp, vp, vp, vp, i64, u64, vp, vp, vp, vp]
    lib.petite_vegas_draw.restype = C.c_int
    lib.petite_vegas_draw_noalloc.argtypes = [vp, vp, vp, vp, i64, u64, vp, vp, vp, vp]
    lib.petite_vegas_draw_noalloc.restype = C.c_int
    lib.petite_eval_points.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp]
    lib.petite_eval_points.restype = C.c_int
    lib.petite_run_showers.argtypes = [vp, vp, i64, u64, C.POINTER(Store), vp, C.POINTER(Stats), vp]
    lib.petite_run_showers.restype = C.c_int
    lib.petite_yield_bound.argtypes = [vp, C.POINTER(BoundDesc), vp, i64]
    lib.petite_yield_bound.restype = C.c_int
    lib.petite_dark_quadrature.argtypes = [C.POINTER(DarkQuadJob), i32, vp, i64, vp, i64]
    lib.petite_dark_quadrature.restype = C.c_int
    lib.petite_dark_setup.argtypes = [vp, C.POINTER(DarkDesc), vp, i64]
    lib.petite_dark_setup.restype = C.c_int
    lib.petite_dark_reweight.argtypes = [vp, C.POINTER(Store), i64, u64, C.POINTER(DarkOut), C.POINTER(DarkStats), vp]
    lib.petite_dark_reweight.restype = C.c_int
    lib.petite_yield_setup.argtypes = [vp, i32, vp, vp, vp, i64]
    lib.petite_yield_setup.restype = C.c_int
    lib.petite_dark_yield.argtypes = [vp, C.POINTER(Store), i64, i32, vp, vp, vp, vp]
    lib.petite_dark_yield.restype = C.c_int
    f64 = C.c_double
    lib.petite_detector_cut.argtypes = [vp, vp, i64, vp, i32, f64, f64, i32, f64, f64, vp, vp, vp]
    lib.petite_detector_cut.restype = C.c_int
    lib.petite_map_accumulate.argtypes = [vp, i32, i64, u64, C.c_uint32, vp, vp, vp]
    lib.petite_map_accumulate.restype = C.c_int
    lib.petite_set_grid.argtypes = [vp, i32, vp, i64]
    lib.petite_set_grid.restype = C.c_int
    lib.petite_sample_scattering.argtypes = [vp, vp, vp, i64, u64, C.POINTER(Store), vp, vp]
    lib.petite_sample_scattering.restype = C.c_int
    lib.petite_propagate.argtypes = [vp, C.POINTER(Store), i64, u64, vp]
    lib.petite_propagate.restype = C.c_int
    lib.petite_draw_pe.argtypes = [vp, i64, u64, f64, vp, vp]
    lib.petite_draw_pe.restype = C.c_int
    lib.petite_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.petite_set_tuning.restype = C.c_int
    lib.petite_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.petite_get_tuning.restype = C.c_int
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != PETITE_OK:
        msg = load().petite_last_error().decode(errors="replace")
        raise PetiteError(rc, f"{what}: {msg}" if what else msg)
