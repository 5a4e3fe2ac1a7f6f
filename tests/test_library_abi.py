"""The HIP library loads and exports every entry point declared in include/matfun_hip.h (no GPU needed: only the
symbols are looked up, nothing is called that touches a device)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "matfun_amd", "libmatfun_hip.so")


def declared():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("smf_hip_eval", "smf_hip_eval_matrix", "gmf_hip_eval", "smf_hip_version", "smf_hip_target",
              "smf_hip_eval_pow", "smf_hip_eval_matrix_pow", "gmf_hip_eval_pow"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for n in declared():
        assert hasattr(lib, n), n
    lib.smf_hip_version.restype = ctypes.c_char_p
    lib.smf_hip_target.restype = ctypes.c_char_p
    assert lib.smf_hip_target() == b"gfx950"
    assert lib.smf_hip_version().startswith(b"0.1.3")


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_invalid_arguments_rejected_before_any_device_work():
    """-1 / -2 / n == 0 paths return before touching the device (cuda/tests/test_cuda_api.cu:113-131), so they can be
    checked on a host without a GPU"""
    lib = ctypes.CDLL(LIB)
    args = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 5
    for f in (lib.smf_hip_eval, lib.smf_hip_eval_matrix, lib.gmf_hip_eval):
        f.argtypes, f.restype = args, ctypes.c_int
    dummy = ctypes.c_void_p(0x1000)
    for f in (lib.smf_hip_eval, lib.smf_hip_eval_matrix, lib.gmf_hip_eval):
        assert f(9, 0.0, 10, dummy, dummy, None, None, None) == -1          # unknown kind
        assert f(-1, 0.0, 10, dummy, dummy, None, None, None) == -1
        assert f(2, 0.0, 10, dummy, dummy, None, None, None) == -1          # pow needs alpha > 0
        assert f(2, -1.0, 10, dummy, dummy, None, None, None) == -1
        assert f(2, float("nan"), 10, dummy, dummy, None, None, None) == -1
        assert f(0, 0.0, 10, dummy, dummy, None, dummy, None) == -1          # H without J
        assert f(0, 0.0, -1, dummy, dummy, None, None, None) == -1          # n < 0
        assert f(0, 0.0, 0, dummy, dummy, None, None, None) == 0            # n == 0: nothing to do
    # grids beyond the limit: -2 without a launch (2^31 blocks of 32 / 64 matrices, and 2^40)
    assert lib.smf_hip_eval(0, 0.0, (2 ** 31 - 1) * 32 + 1, dummy, dummy, None, None, None) == -2
    assert lib.smf_hip_eval_matrix(0, 0.0, 2 ** 40, dummy, dummy, None, None, None) == -2
    assert lib.gmf_hip_eval(0, 0.0, (2 ** 31 - 1) * 64 + 1, dummy, dummy, None, None, None) == -2


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_store_probe_rejects_before_device_work():
    lib = ctypes.CDLL(LIB)
    for f in (lib.mf_hip_store_probe, lib.mf_hip_store_probe_f32):
        f.argtypes, f.restype = [ctypes.c_int64, ctypes.c_int] + [ctypes.c_void_p] * 4, ctypes.c_int
        d = ctypes.c_void_p(0x1000)
        assert f(-1, 9, d, None, None, None) == -1                    # n < 0
        assert f(10, 7, d, None, None, None) == -1                    # nf not 6 or 9
        assert f(10, 9, None, None, None, None) == -1                 # F NULL
        assert f(10, 9, d, None, d, None) == -1                       # H without J
        assert f(10, 9, ctypes.c_void_p(0x1008), d, d, None) == -1    # not 16-byte aligned
        assert f(0, 9, None, None, None, None) == 0                   # n == 0: nothing to do
    for f in (lib.mf_hip_store_probe_in, lib.mf_hip_store_probe_in_f32):
        f.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4
        f.restype = ctypes.c_int
        d = ctypes.c_void_p(0x1000)
        assert f(10, None, 9, 9, d, None, None, None) == -1           # no input rows
        assert f(10, d, 7, 9, d, None, None, None) == -1              # nin not 6 or 9
        assert f(10, d, 9, 7, d, None, None, None) == -1              # nf not 6 or 9
        assert f(-1, d, 9, 9, d, None, None, None) == -1              # n < 0
        assert f(10, d, 9, 9, d, None, d, None) == -1                 # H without J
        assert f(0, None, 9, 9, None, None, None, None) == 0          # n == 0: nothing to do


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_per_matrix_exponent_entry_points_reject_before_device_work():
    lib = ctypes.CDLL(LIB)
    args = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5
    fns = (lib.smf_hip_eval_pow, lib.smf_hip_eval_matrix_pow, lib.gmf_hip_eval_pow, lib.smf_hip_eval_pow_mixed,
           lib.smf_hip_eval_matrix_pow_mixed, lib.gmf_hip_eval_pow_mixed)
    dummy = ctypes.c_void_p(0x1000)
    for f in fns:
        f.argtypes, f.restype = args, ctypes.c_int
        assert f(dummy, -1, dummy, dummy, None, None, None) == -1          # n < 0
        assert f(dummy, 10, dummy, dummy, None, dummy, None) == -1         # H without J
        assert f(None, 10, dummy, dummy, None, None, None) == -1           # no exponents
        assert f(None, 0, None, None, None, None, None) == 0               # n == 0: nothing to do
    assert lib.smf_hip_eval_pow(dummy, 2 ** 40, dummy, dummy, None, None, None) == -2
    assert lib.gmf_hip_eval_pow(dummy, (2 ** 31 - 1) * 64 + 1, dummy, dummy, None, None, None) == -2


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_hencky_entry_points_reject_before_device_work():
    lib = ctypes.CDLL(LIB)
    dummy = ctypes.c_void_p(0x1000)
    for f in (lib.mf_hip_hencky, lib.mf_hip_hencky_mixed):
        f.argtypes, f.restype = [ctypes.c_int64] + [ctypes.c_void_p] * 7, ctypes.c_int
        assert f(-1, dummy, dummy, dummy, dummy, dummy, dummy, None) == -1      # n < 0
        assert f(10, dummy, dummy, None, None, None, dummy, None) == -1         # K without H
        assert f(10, dummy, None, None, None, dummy, None, None) == -1          # g without J
        assert f(10, None, dummy, dummy, dummy, None, None, None) == -1         # no L
        assert f(0, None, None, None, None, None, None, None) == 0              # n == 0: nothing to do
        assert f(10, dummy, dummy, dummy, None, None, None, None) == 0          # no output asked for
    for f in (lib.smf_hip_hencky_matrix, lib.smf_hip_hencky, lib.gmf_hip_hencky):
        f.argtypes, f.restype = [ctypes.c_int64] + [ctypes.c_void_p] * 6, ctypes.c_int
        assert f(-1, dummy, dummy, dummy, dummy, dummy, None) == -1             # n < 0
        assert f(10, None, dummy, None, None, None, None) == -1                 # outputs without A
        assert f(0, None, None, None, None, None, None) == 0
        assert f(10, dummy, None, None, None, None, None) == 0                  # no output asked for


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_plan_rejects_invalid_groups_and_2_31_matrices_before_device_work():
    """mf_hip_plan_create validates every group before it builds anything: an invalid group, or 2^31 matrices or more
    in all (the batch kernels may keep input rows as 32-bit values), give -1 and no plan"""
    class Group(ctypes.Structure):
        _fields_ = [("family", ctypes.c_int32), ("kind", ctypes.c_int32), ("order", ctypes.c_int32),
                    ("count", ctypes.c_int32), ("alpha", ctypes.c_double)]
    lib = ctypes.CDLL(LIB)
    f = lib.mf_hip_plan_create
    f.argtypes, f.restype = [ctypes.POINTER(Group), ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int
    plan = ctypes.c_void_p(0x1234)
    big = (Group * 2)(Group(0, 1, 2, 2 ** 30, 0.0), Group(1, 0, 1, 2 ** 30, 0.0))
    assert f(big, 2, ctypes.byref(plan)) == -1 and not plan.value
    bad = (Group * 2)(Group(0, 1, 2, 10, 0.0), Group(2, 0, 1, 10, 0.0))           # family 2
    assert f(bad, 2, ctypes.byref(plan)) == -1 and not plan.value
    neg = (Group * 1)(Group(1, 2, 1, 10, float("nan")))                            # pow with a NaN exponent
    assert f(neg, 1, ctypes.byref(plan)) == -1 and not plan.value


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_sorted_entry_points_reject_before_device_work():
    """gmf_hip_eval_sorted*: gmf_hip_eval's checks, plus -1 for a missing or too small workspace (fewer than
    min(n, 64) matrices), all before anything is launched; the workspace size grows with n"""
    lib = ctypes.CDLL(LIB)
    lib.gmf_hip_sort_workspace.argtypes, lib.gmf_hip_sort_workspace.restype = [ctypes.c_int64], ctypes.c_size_t
    ws = lib.gmf_hip_sort_workspace
    assert ws(-1) == 0 and ws(64) < ws(65) < ws(1 << 20) and ws(1 << 20) < 600 * (1 << 20)
    dummy = ctypes.c_void_p(0x1000)
    big = 1 << 40
    args = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_size_t, ctypes.c_void_p]
    for f in (lib.gmf_hip_eval_sorted, lib.gmf_hip_eval_sorted_mixed):
        f.argtypes, f.restype = args, ctypes.c_int
        assert f(9, 0.0, 10, dummy, dummy, None, None, dummy, big, None) == -1           # unknown kind
        assert f(2, 0.0, 10, dummy, dummy, None, None, dummy, big, None) == -1           # pow needs alpha > 0
        assert f(0, 0.0, 10, dummy, dummy, None, dummy, dummy, big, None) == -1          # H without J
        assert f(0, 0.0, -1, dummy, dummy, None, None, dummy, big, None) == -1           # n < 0
        assert f(0, 0.0, 10, dummy, dummy, None, None, None, big, None) == -1            # no workspace
        assert f(0, 0.0, 10, dummy, dummy, None, None, dummy, ws(0), None) == -1         # holds no matrix
        assert f(0, 0.0, 0, None, None, None, None, None, 0, None) == 0                  # n == 0
        assert f(0, 0.0, (2 ** 31 - 1) * 64 + 1, dummy, dummy, None, None, dummy, big, None) == -2
    pargs = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_size_t, ctypes.c_void_p]
    for f in (lib.gmf_hip_eval_pow_sorted, lib.gmf_hip_eval_pow_sorted_mixed):
        f.argtypes, f.restype = pargs, ctypes.c_int
        assert f(None, 10, dummy, dummy, None, None, dummy, big, None) == -1             # no exponents
        assert f(dummy, 10, dummy, dummy, None, dummy, dummy, big, None) == -1           # H without J
        assert f(dummy, 100, dummy, dummy, None, None, dummy, ws(0), None) == -1         # holds no matrix
        assert f(None, 0, None, None, None, None, None, 0, None) == 0


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built (python -m matfun_amd.build)")
def test_sorted_workspace_of_n_always_suffices():
    """the workspace rule: the fixed tables alone (gmf_hip_sort_workspace(0)) are refused for every n (only rejections
    are called here: dummy pointers must never reach a launch; the GPU tests run the accepted sizes,
    tests/test_gpu_sorted.py)"""
    lib = ctypes.CDLL(LIB)
    lib.gmf_hip_sort_workspace.argtypes, lib.gmf_hip_sort_workspace.restype = [ctypes.c_int64], ctypes.c_size_t
    f = lib.gmf_hip_eval_sorted
    f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_size_t,
                                                                                            ctypes.c_void_p]
    dummy = ctypes.c_void_p(0x1000)
    for n in (1, 2, 63, 64, 65, 1000):
        small = lib.gmf_hip_sort_workspace(0) + 4 * min(n, 64) - 8         # (room for fewer than min(n, 64))
        assert f(0, 0.0, n, dummy, dummy, None, None, dummy, small, None) == -1
