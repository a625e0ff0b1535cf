"""The C-ABI library loads and exports every function include/hpm_mi355x.h
declares; host-side argument checks behave without a GPU."""
import ctypes
import os
import re

import pytest

import hpm_mi355x as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hpm_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("hpm_create", "hpm_destroy", "hpm_set_views", "hpm_run_patchmatch", "hpm_get_planes",
                 "hpm_cost_fixed", "hpm_jbu_depth", "hpm_jbu_prior", "hpm_set_prior", "hpm_set_coarse"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(H.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(H.EXPORTS) <= set(declared_functions())


def test_struct_layouts_match_the_reference_camera():
    # struct Camera (main.h:39-47): 9 + 9 + 3 floats, 2 ints, 2 floats = 100 bytes
    assert ctypes.sizeof(H.hpm_camera) == 100
    assert ctypes.sizeof(H.hpm_params) == 11 * 4


def test_default_params_are_the_reference_defaults():
    p = H.default_params()
    # PatchMatchParams, HPM.h:32-55; 3 iterations (HPM.cu:3013)
    assert (p.patch_size, p.radius_increment, p.top_k, p.num_iterations) == (11, 2, 4, 3)
    assert (p.sigma_spatial, p.sigma_color) == (5.0, 3.0)
    assert p.precision == H.PREC_FP32_EXACT


def test_error_codes_without_device():
    L = H.lib()
    # null output pointer is rejected before touching the device
    assert L.hpm_create(None, 0) == -1
    # a device index that does not exist is a device error, never an exit()
    h = ctypes.c_void_p()
    assert L.hpm_create(ctypes.byref(h), 4096) == -3
    assert L.hpm_last_error(None) == b"null context"
    assert L.hpm_jbu_depth(None, 4, 4, None, 2, 2, None) == -1
    assert L.hpm_set_u8_pairs_min_bytes(None, ctypes.c_longlong(0)) == -1
