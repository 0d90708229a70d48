"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/apsep/apsep.h declares, and rejects bad arguments before any
HIP call (status codes instead of the reference's exit-on-error,
src/apsep.cuh:39-47)."""
import ctypes
import os
import re

import pytest

import apsep_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include", "apsep")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", "apsep", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(apsep_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    L = A.lib()
    declared = _declared()
    assert declared, "no declarations parsed"
    assert declared == set(A.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    assert L.apsep_version().decode().count(".") == 2


def test_status_strings():
    L = A.lib()
    for st in range(10):
        s = L.apsep_status_string(st).decode()
        assert s and s != "unknown status"
    assert L.apsep_status_string(99).decode() == "unknown status"


def test_scratch_create_rejects_bad_args():
    L = A.lib()
    h = ctypes.c_void_p()
    assert L.apsep_scratch_create(0, 10, 0, None) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_create(7, 10, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_create(-1, 10, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_create(0, -1, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_create(0, 10, 2, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_create(0, 2**31, 0, ctypes.byref(h)) == A.ERR_TOO_LARGE
    assert not h.value


@pytest.mark.skipif(_gpu(), reason="checks the no-device path")
def test_scratch_create_without_device_reports_device_error():
    h = ctypes.c_void_p()
    assert A.lib().apsep_scratch_create(1, 1000, 0, ctypes.byref(h)) == A.ERR_DEVICE
    assert not h.value
    with pytest.raises(A.ApsepError) as e:
        A.Scratch("spt", 100)
    assert e.value.status == A.ERR_DEVICE


def test_run_and_launch_argument_checks():
    L = A.lib()
    assert L.apsep_run(None, 16, 32, 4, None) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_set_persistent_blocks(None, 8) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_set_persistent(None, 1) == A.ERR_INVALID_ARG
    assert L.apsep_scratch_set_cooperative(None, 1) == A.ERR_INVALID_ARG
    refused = ctypes.c_int64(0)
    assert L.apsep_scratch_cooperative_refusals(None, ctypes.byref(refused)) == A.ERR_INVALID_ARG
    code = ctypes.c_int32(0)
    assert L.apsep_batch_plan_check(None, ctypes.byref(code)) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_tuned_persistent(None, 3, ctypes.byref(code)) == A.ERR_INVALID_ARG
    # n <= 0 is a no-op, as the reference's run*/launch* return early
    assert L.apsep_launch(1, None, None, 0, 0) == A.OK
    assert L.apsep_launch(1, None, None, -5, 0) == A.OK
    assert L.apsep_launch(9, 16, 4096, 4, 0) == A.ERR_INVALID_ARG
    assert L.apsep_launch(1, 16, 4096, 4, 3) == A.ERR_INVALID_ARG
    assert L.apsep_launch(1, None, 4096, 4, 0) == A.ERR_INVALID_ARG
    assert L.apsep_launch(1, 4096, None, 4, 0) == A.ERR_INVALID_ARG
    # alignment: dwordx4 loads/stores need 16-byte aligned pointers
    assert L.apsep_launch(1, 4100, 1 << 20, 4, 0) == A.ERR_ALIGN
    assert L.apsep_launch(1, 4096, (1 << 20) + 8, 4, 0) == A.ERR_ALIGN
    # overlapping input/output
    assert L.apsep_launch(1, 4096, 4096 + 16, 100, 0) == A.ERR_INVALID_ARG
    assert L.apsep_launch(1, 4096, 1 << 40, 2**31, 0) == A.ERR_TOO_LARGE


def test_profile_argument_checks():
    L = A.lib()
    t = A.KernelTimes()
    assert L.apsep_time_kernels(None, 4096, 8192, 16, 0, 1, None,
                                ctypes.byref(t)) == A.ERR_INVALID_ARG


def test_batch_plan_argument_checks():
    L = A.lib()
    h = ctypes.c_void_p()
    assert L.apsep_batch_plan_create(100, 4, 0, None) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_create(-1, 4, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_create(100, -4, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_create(100, 4, 2, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_create(2**31, 4, 0, ctypes.byref(h)) == A.ERR_TOO_LARGE
    assert not h.value
    assert L.apsep_batch_plan_destroy(None) == A.OK
    assert L.apsep_batch_plan_use_graph(None, 1) == A.ERR_INVALID_ARG
    assert L.apsep_run_segmented(None, 4096, 8192, 16384, 1, 10, None) == A.ERR_INVALID_ARG
    assert L.apsep_batch_plan_set_tile_lanes(None, 256) == A.ERR_INVALID_ARG
    lanes = ctypes.c_int32(7)
    assert L.apsep_batch_plan_tile_lanes(None, ctypes.byref(lanes)) == A.ERR_INVALID_ARG
    assert lanes.value == 7


def test_server_argument_checks():
    L = A.lib()
    h = ctypes.c_void_p()
    jid = ctypes.c_int64()
    assert L.apsep_server_create(1 << 20, 1000, 16, 0, None) == A.ERR_INVALID_ARG
    assert L.apsep_server_create(8, 1000, 16, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_server_create(1 << 20, 0, 16, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_server_create(1 << 20, 1000, 0, 0, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_server_create(1 << 20, 1000, 16, 5, ctypes.byref(h)) == A.ERR_INVALID_ARG
    assert L.apsep_server_create(1 << 40, 2**31, 16, 0, ctypes.byref(h)) == A.ERR_TOO_LARGE
    assert not h.value
    assert L.apsep_server_destroy(None) == A.OK
    assert L.apsep_server_submit(None, 4096, 10, ctypes.byref(jid)) == A.ERR_INVALID_ARG
    assert L.apsep_server_poll(None, None) == A.ERR_INVALID_ARG
    assert L.apsep_server_release(None, 0) == A.ERR_INVALID_ARG
    assert L.apsep_server_drain(None) == A.ERR_INVALID_ARG


def test_scratch_bytes():
    b0 = A.scratch_bytes("spt", 0)
    b1 = A.scratch_bytes("spt", 1 << 20)
    b2 = A.scratch_bytes("wstl", 1 << 25)
    assert b0 == 0 and 0 < b1 < b2
    # summaries are a small fraction of the data (chunk mins = 1/64 of it)
    assert b2 < (1 << 25) * 4 // 20
    with pytest.raises(A.ApsepError):
        A.scratch_bytes("spt", -1)


def test_header_shims_compile_on_host():
    """apsep.hpp (reference-named C++ templates) compiles with g++ against
    the C header alone (HIP types stay out of the boundary)."""
    import subprocess
    import tempfile
    src = r'''
#include "apsep/apsep.hpp"
void f(const int* in, int* out, int n) {
    auto s = allocSPTScratch<int, 128, 4>(n);
    runSPT<int, 128, 4>(in, out, n, s);
    freeSPTScratch<int, 128, 4>(s);
    auto w = allocWSTLScratch<int>(n);
    runWSTL<int>(in, out, n, w);
    freeWSTLScratch<int>(w);
    launchSPT<int, 128, 4, true>(in, out, n);
    launchWSTL<int>(in, out, n);
}
'''
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.cpp")
        open(p, "w").write(src)
        r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I",
                            os.path.join(ROOT, "include"), p],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
