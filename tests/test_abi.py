"""The C-ABI library loads without a GPU and exports every entry point that
include/imcom_gpu.h declares (no compute calls here)."""
import ctypes
import os
import re

from imcom_amd import gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "imcom_gpu.h")).read()
    return sorted(set(re.findall(r"\b(imcom_gpu_\w+)\s*\(", txt)))


def test_header_matches_binding_table():
    assert _declared() == sorted(gpu.EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_no_gpu_means_loud_failure():
    # without a visible GPU the context cannot be created: no silent fallback
    import torch
    if torch.cuda.is_available():
        return
    try:
        gpu.ImcomGPU(0)
    except gpu.ImcomError:
        return
    raise AssertionError("ImcomGPU constructed without a GPU")
