"""The C-ABI library loads and exports every symbol include/rh_gpu.h and
include/rh_denoise.h declare; host-side argument checks work without a GPU."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="rh_gpu.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rh_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list(gpu_mod):
    assert declared_symbols() == sorted(gpu_mod.EXPORTS)


def test_library_exports_all_symbols(gpu_mod):
    L = gpu_mod.load_library()
    for header in ("rh_gpu.h", "rh_denoise.h"):
        for name in declared_symbols(header):
            assert hasattr(L, name), name


def test_denoiser_header_matches_binding(gpu_mod):
    from rhybrid_amd import denoise
    L = denoise._lib()
    for name in declared_symbols("rh_denoise.h"):
        assert getattr(L, name).argtypes is not None, name


def test_null_arguments_rejected(gpu_mod):
    L = gpu_mod.load_library()
    assert L.rh_destroy(None) == gpu_mod.RH_ERR_ARG
    assert L.rh_create(None, None, 0, None) == gpu_mod.RH_ERR_ARG
    assert L.rh_step(None, 1) == gpu_mod.RH_ERR_ARG
    assert L.rh_last_error(None) == b"null context"


def test_denoiser_null_arguments_rejected(gpu_mod):
    from rhybrid_amd import denoise
    L = denoise._lib()
    h = C.c_void_p()
    assert L.rh_dn_create(0, C.byref(h)) == denoise.RH_DN_ERR_ARG
    assert L.rh_dn_forward(None, None, None, 1, 16, 4, 4) == denoise.RH_DN_ERR_ARG
    assert L.rh_dn_train_step(None, None, None, 1, 16, 4, 4, 0.0, None) == denoise.RH_DN_ERR_ARG
    assert L.rh_dn_last_error(None) == b"null denoiser"
