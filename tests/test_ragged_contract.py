"""Ragged-batch C-ABI contract (CPU): entry validation, status codes, plan-state rules,
extents and the exec guards.  Nothing here launches a kernel; data pointers are dummy
addresses that are never dereferenced.
"""
import ctypes
import math

import pytest

from sip_amd import _native as N

OK = N.ACL_SUCCESS
INVALID = N.ACL_ERROR_INVALID_PARAM
INTERNAL = N.ACL_ERROR_INTERNAL_ERROR
NOT_SUPPORTED = N.ACL_ERROR_API_NOT_SUPPORT
C2C, C2R, R2C = N.ASCEND_FFT_C2C, N.ASCEND_FFT_C2R, N.ASCEND_FFT_R2C
FWD, INV = N.ASCEND_FFT_FORWARD, N.ASCEND_FFT_INVERSE


@pytest.fixture(scope="module")
def L():
    return N.lib()


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


class Handle:
    def __init__(self, L):
        self.L = L
        self.h = ctypes.c_void_p()
        assert L.asdFftCreate(ctypes.byref(self.h)) == OK

    def __enter__(self):
        return self.h

    def __exit__(self, *exc):
        self.L.asdFftDestroy(self.h)


def entries(*rows):
    arr = (N.RaggedEntry * len(rows))()
    for i, (t, d, n, c, i0, o0) in enumerate(rows):
        arr[i].type, arr[i].direction, arr[i].n, arr[i].count = t, d, n, c
        arr[i].in_offset, arr[i].out_offset = i0, o0
    return arr


def make(L, h, rows, flags=0):
    arr = entries(*rows)
    return L.asdFftExtMakePlanRagged(h, arr, len(rows), flags)


GOOD = [(C2C, FWD, 1000, 3, 0, 0), (R2C, FWD, 4096, 2, 24000, 24000), (C2R, INV, 75, 4, 40384, 56800)]


def test_valid_plan_and_info(L):
    with Handle(L) as h:
        assert make(L, h, GOOD) == OK
        ie, oe = ctypes.c_int64(), ctypes.c_int64()
        by, fl = ctypes.c_double(), ctypes.c_double()
        assert L.asdFftExtRaggedInfo(h, ctypes.byref(ie), ctypes.byref(oe), ctypes.byref(by), ctypes.byref(fl)) == OK
        # input: c2c 3*8000, r2c 2*4*4096 from 24000, c2r 4*8*38 from 40384
        assert ie.value == max(24000, 24000 + 2 * 4 * 4096, 40384 + 4 * 8 * 38)
        assert oe.value == max(24000, 24000 + 2 * 8 * 2049, 56800 + 4 * 4 * 75)
        want_bytes = 3 * 16 * 1000 + 2 * (4 * 4096 + 8 * 2049) + 4 * (8 * 38 + 4 * 75)
        assert by.value == want_bytes
        want_flops = sum(5 * n * math.log2(n) * c for _, _, n, c, _, _ in GOOD)
        assert abs(fl.value - want_flops) < 1e-6 * want_flops
        need = ctypes.c_size_t()
        assert L.asdFftExtDescribePlan(h, None, 0, ctypes.byref(need)) == OK
        buf = ctypes.create_string_buffer(need.value)
        L.asdFftExtDescribePlan(h, buf, len(buf), None)
        text = buf.value.decode()
        assert "ragged plan: 3 entries, 3 classes" in text
        assert "c2c:f:1000" in text and "r2c:f:4096" in text and "c2r:i:75" in text
        steps = ctypes.c_int32()
        assert L.asdFftExtPlanSteps(h, ctypes.byref(steps), None) == OK
        assert steps.value == 1  # one register budget -> one launch


def test_two_register_budgets_one_launch_or_two(L):
    """Classes of two register budgets share one launch by default; with
    ASDFFT_RAGGED_PER_BUDGET each budget gets a launch of its own."""
    rows = [(C2C, FWD, 4096, 1, 0, 0), (C2C, FWD, 16384, 1, 32768, 32768)]
    with Handle(L) as h:
        assert make(L, h, rows) == OK
        steps = ctypes.c_int32()
        arr = (ctypes.c_double * 2)()
        assert L.asdFftExtPlanSteps(h, ctypes.byref(steps), arr) == OK
        assert steps.value == 1
        assert arr[0] == 16.0 * 4096 + 16.0 * 16384
    with Handle(L) as h:
        assert L.asdFftExtMakePlanRagged(h, entries(*rows), len(rows), N.ASDFFT_RAGGED_PER_BUDGET) == OK
        steps = ctypes.c_int32()
        arr = (ctypes.c_double * 2)()
        assert L.asdFftExtPlanSteps(h, ctypes.byref(steps), arr) == OK
        assert steps.value == 2
        assert sorted(arr) == [16.0 * 4096, 16.0 * 16384]


@pytest.mark.parametrize("row,code", [
    ((0x13, FWD, 64, 1, 0, 0), INVALID),          # unknown type
    ((N.ASCEND_STFT_C2C, FWD, 64, 1, 0, 0), INVALID),
    ((C2C, 0x12, 64, 1, 0, 0), INVALID),          # invalid direction
    ((C2C, FWD, 0, 1, 0, 0), INVALID),            # zero length
    ((C2C, FWD, 64, 0, 0, 0), INVALID),           # zero count
    ((C2C, FWD, 64, -1, 0, 0), INVALID),
    ((C2C, FWD, 64, 1, -8, 0), INVALID),          # negative offset
    ((C2C, FWD, 64, 1, 4, 0), INVALID),           # complex input not 8-byte aligned
    ((C2R, FWD, 75, 1, 0, 2), INVALID),           # real output not 4-byte aligned
    ((R2C, FWD, 64, 1, 4, 0), INVALID),           # even R2C reads float pairs: 8-byte aligned
    ((C2C, FWD, 10007, 1, 0, 0), NOT_SUPPORTED),  # prime factor above 19
    ((C2C, FWD, 32768, 1, 0, 0), NOT_SUPPORTED),  # above one workgroup's LDS
    ((C2C, FWD, 1, 1, 0, 0), NOT_SUPPORTED),      # length 1 has no Stockham pass
    ((R2C, FWD, 65536, 1, 0, 0), NOT_SUPPORTED),
])
def test_invalid_entries(L, row, code):
    with Handle(L) as h:
        assert make(L, h, [row]) == code
        # a rejected plan does not initialise the handle
        assert make(L, h, [(C2C, FWD, 64, 1, 0, 0)]) == OK


def test_odd_r2c_needs_only_float_alignment(L):
    with Handle(L) as h:
        assert make(L, h, [(R2C, FWD, 75, 1, 4, 0)]) == OK


def test_overlapping_outputs_rejected(L):
    with Handle(L) as h:
        assert make(L, h, [(C2C, FWD, 64, 2, 0, 0), (C2C, FWD, 64, 1, 2048, 512)]) == INVALID
    with Handle(L) as h:  # touching ranges are fine
        assert make(L, h, [(C2C, FWD, 64, 2, 0, 0), (C2C, FWD, 64, 1, 2048, 1024)]) == OK


def test_plan_rules(L):
    arr = entries(*GOOD)
    with Handle(L) as h:
        assert L.asdFftExtMakePlanRagged(h, None, 1, 0) == INVALID
        assert L.asdFftExtMakePlanRagged(h, arr, 0, 0) == INVALID
        assert L.asdFftExtMakePlanRagged(h, arr, len(GOOD), 0x8) == INVALID  # unknown flag
        assert L.asdFftExtMakePlanRagged(h, arr, len(GOOD), N.ASDFFT_RAGGED_GRAPH) == OK
        assert L.asdFftExtMakePlanRagged(h, arr, len(GOOD), 0) == INVALID  # already initialised
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, N.ASCEND_FFT_HORIZONTAL) == OK
        assert L.asdFftExtMakePlanRagged(h, arr, len(GOOD), 0) == INVALID
        assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x900000)) == INVALID
        assert L.asdFftExtRaggedInfo(h, None, None, None, None) == INVALID
    assert L.asdFftExtMakePlanRagged(ctypes.c_void_p(0x1234), arr, len(GOOD), 0) == INVALID
    assert L.asdFftExtExecRagged(ctypes.c_void_p(0x1234), None, None) == INVALID


def test_exec_guards(L):
    with Handle(L) as h:
        assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x900000)) == INVALID  # no plan
        assert make(L, h, GOOD) == OK
        # the plain exec entry points refuse a ragged plan
        t = N.AclTensor((3, 1000), N.ACL_COMPLEX64, 0x1000)
        assert L.asdFftExecC2C(h, t.ptr, t.ptr) == INVALID
        t.close()
        assert L.asdFftExtExecRagged(h, None, ctypes.c_void_p(0x900000)) == INTERNAL
        assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x1000), None) == INTERNAL
        # output range starts inside the input range
        assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x10000), ctypes.c_void_p(0x10000 + 4096)) == INVALID
        # not all-C2C: in place is refused
        assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x10000), ctypes.c_void_p(0x10000)) == INVALID
        if not has_gpu():
            assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x900000)) == INTERNAL


def test_in_place_allowed_only_for_matching_c2c(L):
    with Handle(L) as h:
        assert make(L, h, [(C2C, FWD, 64, 2, 0, 0), (C2C, INV, 128, 1, 1024, 1024)]) == OK
        if not has_gpu():
            # passes the overlap guard, then stops at the missing device
            assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x10000), ctypes.c_void_p(0x10000)) == INTERNAL
    with Handle(L) as h:
        assert make(L, h, [(C2C, FWD, 64, 2, 0, 1024), (C2C, INV, 128, 1, 1024, 0)]) == OK
        assert L.asdFftExtExecRagged(h, ctypes.c_void_p(0x10000), ctypes.c_void_p(0x10000)) == INVALID


def test_get_type_reports_first_entry(L):
    with Handle(L) as h:
        assert make(L, h, [(R2C, FWD, 64, 1, 0, 0), (C2C, FWD, 64, 1, 256, 520)]) == OK
        t = ctypes.c_int()
        assert L.asdFftGetType(h, ctypes.byref(t)) == OK
        assert t.value == R2C


# ---- serving loop C ABI (no kernel launches) ------------------------------------------------
def test_serve_create_guards(L):
    h = ctypes.c_void_p()
    assert L.asdFftExtServeCreate(None, None) == INVALID
    bad = N.ServeConfig(0, 9, 0, 1, 0)  # more than 8 slots
    assert L.asdFftExtServeCreate(ctypes.byref(h), ctypes.byref(bad)) == INVALID
    if not has_gpu():
        assert L.asdFftExtServeCreate(ctypes.byref(h), None) == INTERNAL
        assert not h.value


def test_serve_unknown_server(L):
    fake = ctypes.c_void_p(0x1234)
    r = N.ServeRequest()
    t = ctypes.c_uint64()
    assert L.asdFftExtServeSubmit(fake, ctypes.byref(r), ctypes.byref(t)) == INVALID
    assert L.asdFftExtServeWait(fake, 1, 0, None) == INVALID
    assert L.asdFftExtServeStats(fake, ctypes.byref(N.ServeCounters())) == INVALID
    assert L.asdFftExtServeDestroy(fake) == INVALID
    assert L.asdFftExtServeReplay(fake, None, None, 0, None) == INVALID


def test_workload_schedules_fit_their_launch():
    """Every class of the 4000-transform workload fits the workgroup of its launch (threads per
    transform <= block): with wide EMAX-16 workgroups of up to 1024 threads (plus one launch for
    the classes that need a larger register budget), in the default single launch, and in the
    per-budget form."""
    import collections
    import os
    import re

    import bench
    from sip_amd.fft import RaggedPlan
    groups = collections.OrderedDict()
    for k, n in sorted(bench.ragged_workload()):
        groups[(k, n)] = groups.get((k, n), 0) + 1
    entries = [(k, n, c, False) for (k, n), c in groups.items()]
    forms = []
    for wide, per_budget in (("1", False), ("0", False), (None, True)):
        if wide is None:
            os.environ.pop("ASDFFT_RAGGED_WIDE16", None)
        else:
            os.environ["ASDFFT_RAGGED_WIDE16"] = wide
        try:
            plan = RaggedPlan(entries, per_budget=per_budget)
            text = plan.describe()
            plan.close()
        finally:
            os.environ.pop("ASDFFT_RAGGED_WIDE16", None)
        block = None
        launches = 0
        for line in text.splitlines():
            m = re.search(r"launch emax=(\d+) block=(\d+)", line)
            if m:
                block = int(m.group(2))
                launches += 1
                assert block <= (1024 if m.group(1) == "16" else 512), line
                continue
            m = re.search(r" T=(\d+) tpb=(\d+)", line)
            if m:
                assert int(m.group(1)) * int(m.group(2)) <= block, line
        forms.append(launches)
    assert forms[0] <= 2 and forms[1] == 1 and forms[2] == 3, forms
