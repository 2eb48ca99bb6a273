"""petite_vegas_draw_noalloc (no allocation; its scratch is one static device
array per GPU) against petite_vegas_draw on the same requests: identical
trial counts, points and J*f, on two calls in a row, the second smaller, so
that state the first call leaves in the shared array cannot leak into the
second.  Called through the C ABI (ctypes)."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from petite_amd import _lib
from petite_amd.constants import process_code
from petite_amd.device import DeviceContext, _stream_handle
from petite_amd.tables import MaterialTables

SEED = 99
PROCS = ["Brem", "PairProd", "Comp", "Ann", "Moller", "Bhabha"]


def _requests(T, n, rng):
    procs = rng.choice(PROCS, size=n)
    E = np.empty(n)
    for i, P in enumerate(procs):
        xg, yg = T.nsigma_tab[P]
        lo = xg[np.argmax(yg > 0)]
        E[i] = np.exp(rng.uniform(np.log(lo * 1.05), np.log(xg[-1] * 0.99)))
    code = np.array([process_code[P] for P in procs], dtype=np.int32)
    sid = rng.integers(0, 2 ** 62, size=n, dtype=np.int64)
    return code, E, sid


def _call(ctx, fn, code, E, sid):
    dev = torch.device("cuda")
    n = len(code)
    p_t = torch.as_tensor(code, device=dev)
    E_t = torch.as_tensor(E, device=dev)
    s_t = torch.as_tensor(sid, device=dev)
    x = torch.full((n, 4), -7.0, device=dev, dtype=torch.float32)
    tr = torch.full((n,), -7, device=dev, dtype=torch.int32)
    w = torch.full((n,), -7.0, device=dev, dtype=torch.float32)
    rc = fn(ctx.ctx, p_t.data_ptr(), E_t.data_ptr(), s_t.data_ptr(), n, SEED, x.data_ptr(), tr.data_ptr(), w.data_ptr(),
            _stream_handle())
    torch.cuda.synchronize()
    return rc, x.cpu().numpy(), tr.cpu().numpy(), w.cpu().numpy()


def test_noalloc_matches_allocating_draws_on_two_calls():
    mt = MaterialTables("graphite")
    ctx = DeviceContext(mt, 0.010)
    lib = _lib.load()
    rng = np.random.default_rng(4)
    for n in (12000, 700):
        code, E, sid = _requests(mt, n, rng)
        rc_a, xa, ta, wa = _call(ctx, lib.petite_vegas_draw, code, E, sid)
        rc_b, xb, tb, wb = _call(ctx, lib.petite_vegas_draw_noalloc, code, E, sid)
        assert rc_a == rc_b == _lib.PETITE_OK, (rc_a, rc_b, lib.petite_last_error())
        assert np.array_equal(ta, tb), n
        assert np.all(tb > 0)
        assert np.array_equal(xa, xb), n
        assert np.array_equal(wa, wb), n


def test_noalloc_rejects_too_many_requests():
    ctx = DeviceContext(MaterialTables("graphite"), 0.010)
    lib = _lib.load()
    n = _lib.NOALLOC_MAX_DRAWS + 1
    dev = torch.device("cuda")
    code = torch.zeros(n, dtype=torch.int32, device=dev)
    E = torch.ones(n, dtype=torch.float64, device=dev)
    sid = torch.arange(n, dtype=torch.int64, device=dev)
    x = torch.empty((n, 4), dtype=torch.float32, device=dev)
    tr = torch.empty(n, dtype=torch.int32, device=dev)
    w = torch.empty(n, dtype=torch.float32, device=dev)
    rc = lib.petite_vegas_draw_noalloc(ctx.ctx, code.data_ptr(), E.data_ptr(), sid.data_ptr(), n, SEED, x.data_ptr(),
                                       tr.data_ptr(), w.data_ptr(), _stream_handle())
    assert rc == _lib.PETITE_ERR_ARG
    assert b"PETITE_NOALLOC_MAX_DRAWS" in lib.petite_last_error()
