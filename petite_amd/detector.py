"""Detector acceptance: drop-ins for the reference's ``transverse_position``
and ``detector_cut`` (src/PETITE/shower.py:815-865), computed on the GPU by
``petite_detector_cut``.

``detector_cut`` takes the reference's arguments.  ``particle_list`` may be a
list of Particle (the reference's input; the outputs are the reference's
lists), a ParticleBatch (the SM record on the device) or a DarkBatch (its
valid vectors, in the reference's order).  For the device batches "Sample"
returns, per detector, a LongTensor of positions in the energy-filtered
record instead of a list of Particle objects.

As in the reference, ``detector_positions`` are the z of the detector planes
(shower.py:852 passes them as transverse_position's z), each record is
extrapolated in a straight line from its production point along its
production momentum, and the energy window is strict on both sides.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib

MAX_DETECTORS = 64


def transverse_position(particle, z):
    """Determines the transverse position of a particle at a given z
    (shower.py:815-823)."""
    x0, y0, z0 = particle.get_r0()
    E, px, py, pz = particle.get_p0()
    T = (z - z0) / pz
    xf, yf = x0 + T * px, y0 + T * py
    return [xf, yf]


_GATHER_ROWS = 1 << 26


def _take(t, idx):
    """t.index_select(0, idx) in pieces whose output stays below 2^31 bytes:
    on this ROCm/torch stack a single row gather with a larger output returned
    wrong rows past that point."""
    row = t[0:1].numel() * t.element_size()
    if idx.numel() * row < (1 << 31):
        return t.index_select(0, idx)
    return torch.cat([t.index_select(0, idx[i:i + _GATHER_ROWS]) for i in range(0, idx.numel(), _GATHER_ROWS)])


def _records(particle_list):
    """(p, r, kind, selection) on the device: float32 (n, 4) production
    momentum and (x, y, z, weight)."""
    from .batch import ParticleBatch
    from .dark_shower import DarkBatch
    if isinstance(particle_list, ParticleBatch):
        return particle_list.p0.contiguous(), particle_list.r0.contiguous(), "batch", None
    if isinstance(particle_list, DarkBatch):
        idx = torch.as_tensor(particle_list.order(), device=particle_list.p.device, dtype=torch.long)
        return (_take(particle_list.p, idx).contiguous(), _take(particle_list.tensor("r"), idx).contiguous(),
                "batch", idx)
    parts = list(particle_list)
    n = len(parts)
    p = np.zeros((n, 4), dtype=np.float32)
    r = np.zeros((n, 4), dtype=np.float32)
    for i, q in enumerate(parts):
        p[i] = q.get_p0()
        r[i, :3] = q.get_r0()
        r[i, 3] = q.get_ids()["weight"]
    dev = torch.device("cuda")
    return torch.from_numpy(p).to(dev), torch.from_numpy(r).to(dev), "list", parts


def cut_masks(p, r, detector_positions, detector_radius, energy_cut=None, detector_inner_radius=0.0):
    """Device pass masks and weight sums of float32 (n, 4) record tensors:
    (mask uint64 tensor (n,), bit d = passes plane d; wsum float64 (n_det+1,):
    passing weight per plane, then the weight inside the energy window)."""
    if not torch.cuda.is_available():
        raise RuntimeError("petite_amd needs a HIP device (MI355X); no GPU is visible")
    z = np.ascontiguousarray(np.atleast_1d(np.asarray(detector_positions, dtype=np.float64)))
    if not 1 <= len(z) <= MAX_DETECTORS:
        raise ValueError(f"detector_cut: 1 to {MAX_DETECTORS} detector planes")
    lib = _lib.load()
    n = int(p.shape[0])
    dev = p.device
    mask = torch.empty(n, device=dev, dtype=torch.int64)
    wsum = torch.empty(len(z) + 1, device=dev, dtype=torch.float64)
    has = energy_cut is not None
    emin, emax = (float(energy_cut[0]), float(energy_cut[1])) if has else (0.0, 0.0)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.petite_detector_cut(p.data_ptr() if n else None, r.data_ptr() if n else None, n, z.ctypes.data,
                                       len(z), float(detector_radius), float(detector_inner_radius), int(has), emin,
                                       emax, mask.data_ptr() if n else None, wsum.data_ptr(), stream),
               "petite_detector_cut")
    return mask, wsum


def detector_cut(particle_list, detector_positions, detector_radius, method="Sample", energy_cut=None,
                 detector_inner_radius=0.0):
    """Places an imaginary detector of a certain size at a certain distance
    from the beam origin (shower.py:825-865); methods "Sample", "SampleW",
    "Efficiency", "TotalWeight"."""
    z = np.atleast_1d(np.asarray(detector_positions, dtype=np.float64))
    n_det = len(z)
    p, r, kind, sel = _records(particle_list)
    mask, wsum = cut_masks(p, r, z, detector_radius, energy_cut, detector_inner_radius)
    if energy_cut is not None:
        E = p[:, 0].double()  # the kernel's comparison, in float64
        win = (E < float(energy_cut[1])) & (E > float(energy_cut[0]))
    else:
        win = torch.ones(p.shape[0], dtype=torch.bool, device=p.device)
    keep = torch.nonzero(win).flatten()
    if keep.numel() == 0:
        return [[] for _ in range(n_det)] if method == "Sample" else [0.0 for _ in range(n_det)]
    m = _take(mask, keep)
    bits = torch.stack([((m >> d) & 1).bool() for d in range(n_det)])  # (n_det, n_filtered)
    if method == "Sample":
        if kind == "list":
            kept = keep.cpu().numpy()
            b = bits.cpu().numpy()
            return [[sel[kept[i]] for i in np.nonzero(b[d])[0]] for d in range(n_det)]
        return [torch.nonzero(bits[d]).flatten() for d in range(n_det)]
    if method == "SampleW":
        return bits.cpu().numpy() if kind == "list" else bits
    w = wsum.cpu().numpy()
    if method == "Efficiency":
        with np.errstate(divide="ignore", invalid="ignore"):
            return [np.float64(w[d]) / np.float64(w[n_det]) for d in range(n_det)]
    if method == "TotalWeight":
        return [np.float64(w[d]) for d in range(n_det)]
    return None
