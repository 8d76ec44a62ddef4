"""The C-ABI library loads on a host without a GPU, exports every entry point
that include/ssim2hip.h declares, and validates arguments before touching the
device (no compute calls here)."""
import ctypes
import re

from conftest import ROOT

HEADER = ROOT / "include" / "ssim2hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(ssim2_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    import ssimulacra2_hip as s2
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(s2.lib, n)]
    assert not missing, missing


def test_status_strings_and_version():
    import ssimulacra2_hip as s2
    assert s2.lib.ssim2_status_string(0) == b"SSIM2_OK"
    assert s2.lib.ssim2_status_string(6) == b"SSIM2_ADMISSION_REFUSED"
    assert s2.lib.ssim2_version() >= 100


def test_argument_validation_without_device():
    import ssimulacra2_hip as s2
    L = s2.lib
    out = ctypes.c_void_p()
    assert L.ssim2_create(7, 100, None, ctypes.byref(out)) == 1          # INVALID_SIZE
    assert b"7x100" in L.ssim2_last_error()
    assert L.ssim2_create(100, 40000, None, ctypes.byref(out)) == 1      # beyond maximum
    assert L.ssim2_create(64, 64, None, None) == 8                       # INVALID_ARGUMENT
    d = ctypes.c_double()
    assert L.ssim2_get_score(None, ctypes.byref(d)) == 8
    assert L.ssim2_destroy(None) == 0
    assert L.ssim2_image_alloc(0, 10, 1, None) == 8


def test_resident_reference_argument_validation_without_device():
    import ssimulacra2_hip as s2
    L = s2.lib
    rid = ctypes.c_uint32()
    img = s2._Image()
    assert L.ssim2_batch_add_reference(None, ctypes.byref(img), None, ctypes.byref(rid)) == 8
    assert L.ssim2_batch_remove_reference(None, 1) == 8
    assert L.ssim2_batch_compute_ref_async(None, None, 1, None) == 8
    assert L.ssim2_server_add_reference(None, ctypes.byref(img), ctypes.byref(rid)) == 8
    assert L.ssim2_server_remove_reference(None, 1) == 8
    assert L.ssim2_server_submit_ref(None, 1, ctypes.byref(img), 0) == 8


def test_bf16_storage_refused_without_device():
    import ssimulacra2_hip as s2
    L = s2.lib
    o = s2._BatchOptions(ctypes.sizeof(s2._BatchOptions), 2, 4, 0, 100_000, 0)
    out = ctypes.c_void_p()
    assert L.ssim2_batch_create(ctypes.byref(o), ctypes.byref(out)) == 8
    assert b"bfloat16" in L.ssim2_last_error()
    assert not out.value
