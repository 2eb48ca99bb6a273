"""m_V-continuous dark-vector yield of a shower record, differentiable in m_V
and eps^2.

The reference evaluates the dark weights only at its trained masses
(src/PETITE/dark_shower.py:158-176 picks one).  Here the yield per unit eps^2,
Y(m_V) = sum_i w_i sum_P W_P(E0_i; m_V) with W_P = GetBSMWeights (595-647),
interpolates each particle's weight between the trained masses m_j < m_{j+1}:
W_j^(1-t) W_{j+1}^t (t = ln(m_V/m_j) / ln(m_{j+1}/m_j)) where both are positive,
(1-t) W_j + t W_{j+1} otherwise; at a trained mass it equals the weight sum of
DarkShower.generate_dark_shower's entries.  The bound-electron processes
(DarkAnn, DarkComp) are not interpolated: their weights are a resonance and a
threshold that move with m_V, so the device builds their tables at each
requested m_V (closed-form cross sections, DarkAnn track-length weights;
petite_yield_bound), and their dY/dm_V is the central difference of tables
at m_V (1 +- 1e-5).  The trained masses are the
shipped ones and those of the generated mass pack
(petite_amd.tablegen.dark_masses); a mass outside [lightest, heaviest] raises
PetiteError (petite_dark_yield rejects it: nothing to interpolate).  One float64 pass of the
k_dark_yield kernel returns Y and dY/dm_V for up to 16 masses at a time;
dY/d(eps^2) = Y.  ``DarkYield.__call__`` wraps it in a torch.autograd.Function.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from .constants import all_process_code, dark_process_codes
from .constants import GeVsqcm2, alpha_em, m_electron
from .dark_tables import MFP_ORDER, SURVIVAL, DarkTables, build_dark_blob, quadrature_loss_tables
from .device import DeviceContext, _stream_handle
from .tables import MaterialTables


def _store_struct(store):
    return _lib.Store(*(store[k].data_ptr() for k in ("p0", "r0", "pf", "rf", "pid", "parent", "meta", "shower",
                                                       "sid")), int(store["pid"].shape[0]))


class DarkYield:
    """Yield model of one material and table set over the trained masses."""

    def __init__(self, tables: MaterialTables, min_energy=0.010, masses=None, active_processes=None, Zeff=29.508,
                 ctx=None):
        self.tables = tables
        active = list(dark_process_codes if active_processes is None else active_processes)
        self.active_codes = [all_process_code[p] for p in active if p in all_process_code]
        self.dark = {}
        probe = DarkTables(tables, 1.0, Zeff=Zeff, masses="all")
        self.masses = sorted(float(m) for m in (probe.mV_list if masses is None else masses))
        if len(self.masses) > _lib.MAX_MASSES:
            raise ValueError(f"at most {_lib.MAX_MASSES} trained masses")
        descs = (_lib.DarkDesc * len(self.masses))()
        parts, n = [], 0
        for j, m in enumerate(self.masses):
            dt = probe if m == probe.mV else DarkTables(tables, m, Zeff=Zeff, masses="all")
            if dt.mV != m:
                raise ValueError(f"mass {m} is not a trained mass ({dt.mV_list})")
            dt.eps2 = 1.0
            self.dark[m] = dt
            d, blob = build_dark_blob(dt, self.active_codes)
            for f in range(len(_lib.DARK_FAMILIES)):
                d.w_off[f] += n
                d.d_off[f] += n
            d.comp_off += n
            descs[j] = d
            parts.append(blob)
            n += blob.size
        self._blob = np.concatenate(parts)
        self._descs = descs
        self._bound, self._bound_blob = self._bound_inputs(Zeff)
        self.ctx = ctx if ctx is not None else DeviceContext(tables, min_energy)
        self._setup()

    def _bound_inputs(self, Zeff, rel_step=1e-5):
        """Inputs of petite_yield_bound: the positron survival and mean-free-path
        tables of the material and the grid tops of each trained mass."""
        sm = self.tables
        d = _lib.BoundDesc()
        d.Zeff, d.ne_conv, d.m_e, d.alpha = float(Zeff), float(sm.ne * GeVsqcm2), m_electron, alpha_em
        d.ann_floor, d.comp_floor = float(sm.min_calc[-11]), float(sm.min_calc[22])
        d.rel_step, d.dEdx = rel_step, float(sm.dEdx)
        for j, m in enumerate(self.masses):
            d.top_ann[j] = float(self.dark[m].xsec["DarkBrem"][-1, 0])
            d.top_comp[j] = float(self.dark[m].xsec["DarkComp_free"][-1, 0])
        parts, n = [], 0

        def put(*arrays):
            nonlocal n
            off = n
            for a in arrays:
                parts.append(np.ascontiguousarray(a, dtype=np.float64).ravel())
                n += parts[-1].size
            return off

        loss = quadrature_loss_tables(sm)
        for c, P in enumerate(SURVIVAL[-11]):
            x, y, _ = loss[P]
            d.loss[c].off, d.loss[c].n = put(x, y), len(x)
        for c, P in enumerate(MFP_ORDER[-11]):
            x, y = sm.nsigma_tab[P]
            d.mfp[c].off, d.mfp[c].n = put(x, y), len(x)
        nodes = np.unique(np.concatenate([loss[P][2] for P in SURVIVAL[-11]]))
        d.nodes.off, d.nodes.n = put(nodes), len(nodes)
        return d, np.concatenate(parts)

    def _setup(self):
        """Load this model's tables into the context (a context holds one
        model's tables; models sharing a context reload theirs when used)."""
        m = np.ascontiguousarray(self.masses, dtype=np.float64)
        _lib.check(self.ctx.lib.petite_yield_setup(self.ctx.ctx, len(self.masses), m.ctypes.data,
                                                   C.cast(self._descs, C.c_void_p), self._blob.ctypes.data,
                                                   self._blob.size), "petite_yield_setup")
        if any(c in self.active_codes for c in (all_process_code["DarkAnn"], all_process_code["DarkComp"])):
            _lib.check(self.ctx.lib.petite_yield_bound(self.ctx.ctx, C.byref(self._bound), self._bound_blob.ctypes.data,
                                                       self._bound_blob.size), "petite_yield_bound")
        self.ctx._yield_owner = self

    def evaluate(self, record, mV, grad=True):
        """(Y, dY/dm_V) per unit eps^2 at the masses mV (host float64 arrays);
        grad=False skips the derivative (returns None for it)."""
        store, n = (record._store, record.n) if hasattr(record, "_store") else record
        if getattr(self.ctx, "_yield_owner", None) is not self:
            self._setup()
        mv = np.ascontiguousarray(np.atleast_1d(mV), dtype=np.float64)
        Y = np.zeros_like(mv)
        dY = np.zeros_like(mv) if grad else None
        S = _store_struct(store)
        _lib.check(self.ctx.lib.petite_dark_yield(self.ctx.ctx, C.byref(S), int(n), mv.size, mv.ctypes.data,
                                                  Y.ctypes.data, dY.ctypes.data if grad else None,
                                                  _stream_handle()), "petite_dark_yield")
        return Y, dY

    def __call__(self, record, mV, eps2):
        """eps2 * Y(mV) as a differentiable float64 tensor (mV: 1-D tensor; eps2: scalar or like mV)."""
        mV = torch.as_tensor(mV, dtype=torch.float64)
        eps2 = torch.as_tensor(eps2, dtype=torch.float64, device=mV.device)
        return _YieldFn.apply(mV, eps2, self, record)


class _YieldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mV, eps2, model, record):
        Y, dY = model.evaluate(record, mV.detach().cpu().numpy())
        Yt = torch.as_tensor(Y, dtype=torch.float64, device=mV.device).reshape(mV.shape)
        dYt = torch.as_tensor(dY, dtype=torch.float64, device=mV.device).reshape(mV.shape)
        ctx.save_for_backward(Yt, dYt, eps2)
        return eps2 * Yt

    @staticmethod
    def backward(ctx, g):
        Yt, dYt, eps2 = ctx.saved_tensors
        g_m = g * eps2 * dYt
        g_e = g * Yt
        if eps2.dim() == 0:
            g_e = g_e.sum()
        else:
            g_e = g_e.sum_to_size(eps2.shape)
        return g_m, g_e, None, None
