"""The C-ABI library loads and exports every symbol include/rh_gpu.h
declares; host-side argument checks work without a GPU."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "rh_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rh_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list(gpu_mod):
    assert declared_symbols() == sorted(gpu_mod.EXPORTS)


def test_library_exports_all_symbols(gpu_mod):
    L = gpu_mod.load_library()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_null_arguments_rejected(gpu_mod):
    L = gpu_mod.load_library()
    assert L.rh_destroy(None) == gpu_mod.RH_ERR_ARG
    assert L.rh_create(None, None, 0, None) == gpu_mod.RH_ERR_ARG
    assert L.rh_step(None, 1) == gpu_mod.RH_ERR_ARG
    assert L.rh_last_error(None) == b"null context"
