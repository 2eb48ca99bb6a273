"""The C ABI: libpetite_hip.so loads without a GPU, exports every entry point
include/petite_hip.h declares, and its ctypes structures match the C layout."""
import ctypes
import os
import re
import subprocess

import pytest

from petite_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "petite_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(petite_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "petite_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(PetiteTableDesc), '
                   'sizeof(PetitePrimary), sizeof(PetiteStore), sizeof(PetiteStats), offsetof(PetiteTableDesc, grid), '
                   'sizeof(PetiteGridDesc), sizeof(PetiteSigmaDesc), sizeof(PetiteDarkDesc), '
                   'offsetof(PetiteDarkDesc, eps2), sizeof(PetiteDarkOut), sizeof(PetiteDarkStats), '
                   'offsetof(PetiteDarkStats, yield)); printf("%zu %zu\\n", sizeof(PetiteTuning), '
                   'offsetof(PetiteTuning, sweep_bound)); printf("%zu %zu %zu %zu %zu\\n", sizeof(PetiteDarkQuadJob), '
                   'offsetof(PetiteDarkQuadJob, lo), sizeof(PetiteBoundDesc), offsetof(PetiteBoundDesc, loss), '
                   'sizeof(PetiteQTable)); return 0;}\n')
    exe = tmp_path / "sz"
    try:
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    except (FileNotFoundError, subprocess.CalledProcessError):
        pytest.skip("no C compiler")
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(_lib.TableDesc), ctypes.sizeof(_lib.Primary), ctypes.sizeof(_lib.Store),
            ctypes.sizeof(_lib.Stats), _lib.TableDesc.grid.offset, ctypes.sizeof(_lib.GridDesc),
            ctypes.sizeof(_lib.SigmaDesc), ctypes.sizeof(_lib.DarkDesc), _lib.DarkDesc.eps2.offset,
            ctypes.sizeof(_lib.DarkOut), ctypes.sizeof(_lib.DarkStats), _lib.DarkStats.yield_.offset,
            ctypes.sizeof(_lib.Tuning), _lib.Tuning.sweep_bound.offset, ctypes.sizeof(_lib.DarkQuadJob),
            _lib.DarkQuadJob.lo.offset, ctypes.sizeof(_lib.BoundDesc), _lib.BoundDesc.loss.offset,
            ctypes.sizeof(_lib.QTable)]
    assert got == want


def test_error_string_without_gpu():
    lib = _lib.load()
    rc = lib.petite_ctx_create(None, None, 0, None, 0, None)
    assert rc == _lib.PETITE_ERR_ARG
    assert b"bad arguments" in lib.petite_last_error()


def test_primary_dtype_matches_struct():
    from petite_amd.batch import PRIMARY_DTYPE
    assert PRIMARY_DTYPE.itemsize == ctypes.sizeof(_lib.Primary)


def test_detector_cut_rejects_bad_arguments_without_gpu():
    lib = _lib.load()
    z = (ctypes.c_double * 1)(1.0)
    assert lib.petite_detector_cut(None, None, 0, z, 0, 1.0, 0.0, 0, 0.0, 0.0, None, None, None) == _lib.PETITE_ERR_ARG
    assert lib.petite_detector_cut(None, None, 5, z, 1, 1.0, 0.0, 0, 0.0, 0.0, None, None, None) == _lib.PETITE_ERR_ARG
    assert lib.petite_detector_cut(None, None, 0, None, 1, 1.0, 0.0, 0, 0.0, 0.0, None, None, None) == _lib.PETITE_ERR_ARG
    assert b"petite_detector_cut" in lib.petite_last_error()


def test_detector_limit_matches_header():
    from petite_amd import detector
    m = re.search(r"#define PETITE_MAX_DETECTORS (\d+)", open(HEADER).read())
    assert int(m.group(1)) == detector.MAX_DETECTORS


def test_tuning_rejects_null_context_without_gpu():
    lib = _lib.load()
    t = _lib.Tuning()
    assert lib.petite_set_tuning(None, ctypes.byref(t)) == _lib.PETITE_ERR_ARG
    assert lib.petite_get_tuning(None, ctypes.byref(t)) == _lib.PETITE_ERR_ARG
    assert b"petite_get_tuning" in lib.petite_last_error()


def test_tuning_dict_round_trip():
    d = {"batch_lanes": 40, "yield_steps": 32, "transport_waves": 16, "emit_waves": 24,
         "draw_waves": [12, 2, 8, 2, 2, 2, 2, 2], "min_group_lanes": 4, "sweep_group": 8, "group_requests": 2, "small_sweep": 4096, "small_yield_steps": 80, "graph": 1,
         "sweep_bound": 1 << 22, "resume_cap": 1 << 19,
         "tail_sweep": 2048, "tail_waves": 4, "tail_depth": 16, "tail_budget": 32}
    assert _lib.Tuning.from_dict(d).as_dict() == d
