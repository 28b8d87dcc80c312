"""Planner tests (CPU): radix schedules, launch shapes and workspace sizes for the
benchmark configurations, read back through asdFftExtDescribePlan."""
import ctypes
import re

import pytest

from sip_amd import _native as N
from sip_amd.fft import Plan


def steps(desc):
    return [l.strip() for l in desc.splitlines() if l.strip().startswith("[")]


def field(line, name):
    m = re.search(rf"\b{name}=([\w\.x]+)", line)
    return m.group(1) if m else None


@pytest.mark.parametrize("n,radices", [(4096, "16x16x16"), (1000, "10x10x10"), (2187, "9x9x9x3"),
                                       (256, "16x16"), (16, "16"), (17, "17"), (6, "6"), (2048, "16x16x8")])
def test_radix_schedule(n, radices):
    p = Plan("c2c", (n,), 2048)  # large enough that the small-launch order rule does not apply
    s = steps(p.describe())
    assert len(s) == 1 and field(s[0], "radices") == radices
    p.close()


def test_small_launches_take_the_smallest_radix_first():
    """Row C2C launches of at most 16 MiB of input whose radices are all below 12 run the smallest
    radix first (2187 x 256: 3x9x9x9); larger launches and sequences with radix 12 or 16 keep the
    fewest-pass order."""
    for n, b, want in [(2187, 256, "3x9x9x9"), (2187, 2048, "9x9x9x3"), (6000, 256, "6x10x10x10"),
                       (1000, 256, "10x10x10"), (4096, 512, "16x16x16"), (1536, 256, "16x16x6")]:
        p = Plan("c2c", (n,), b)
        assert field(steps(p.describe())[0], "radices") == want, (n, b)
        p.close()


def test_one_launch_per_single_kernel_length():
    for n in (1000, 2187, 6000, 4096, 16384):
        p = Plan("c2c", (n,), 256)
        assert len(steps(p.describe())) == 1 and p.workspace_size == 0
        p.close()


def test_large_length_is_two_pass_four_step():
    p = Plan("c2c", (65536,), 64)
    s = steps(p.describe())
    assert len(s) == 2
    assert "[in->work]" in s[0] and "[work->out]" in s[1]
    assert p.workspace_size == 65536 * 64 * 8
    assert p.step_bytes() == [65536 * 64 * 16.0] * 2
    p.close()


def test_real_even_uses_half_length_complex():
    for kind in ("r2c", "c2r"):
        p = Plan(kind, (8192,), 1024)
        s = steps(p.describe())
        assert len(s) == 1 and field(s[0], "n") == "4096"
        p.close()


def test_real_odd_uses_full_length():
    p = Plan("r2c", (2187,), 2)
    s = steps(p.describe())
    assert field(s[0], "n") == "2187" and field(s[0], "variant") == "4"
    p.close()


def test_multidim_plans_have_no_transposes(monkeypatch):
    p = Plan("c2c", (2048, 2048), 4)
    s = steps(p.describe())
    assert len(s) == 2 and "[in->out]" in s[0] and "[out->out]" in s[1] and p.workspace_size == 0
    # rows in fours through one LDS plane (variant 24: rows n + q 512 with the columns' first
    # radix-4 stage), then 512-point columns on one LDS plane (variant 15), in place on the output
    assert field(s[0], "variant") == "24"
    assert field(s[1], "variant") == "15" and field(s[1], "n") == "512" and field(s[1], "lds") == "34816"
    p.close()
    monkeypatch.setenv("ASDFFT_QUAD2D", "0")
    p = Plan("c2c", (2048, 2048), 4)
    s = steps(p.describe())
    # paired rows (variant 21: rows n and n + 1024 with the columns' first radix-2 stage), then
    # 1024-point columns on one LDS plane (variant 15), both in place on the output
    assert field(s[0], "variant") == "21"
    assert field(s[1], "variant") == "15" and field(s[1], "n") == "1024" and field(s[1], "lds") == "69632"
    p.close()
    monkeypatch.setenv("ASDFFT_PAIR2D", "0")
    p = Plan("c2c", (2048, 2048), 4)
    s = steps(p.describe())
    # without pairing: the 2048-point columns exchange through one LDS plane (variant 15),
    # 16 columns per workgroup (128-byte runs) in 136 KiB
    assert field(s[1], "variant") == "15" and field(s[1], "tpb") == "16" and field(s[1], "lds") == "139264"
    p.close()
    p = Plan("c2c", (1024, 1024), 4)  # 1024-point columns: 16 per workgroup on one plane, two per CU
    s1 = steps(p.describe())[1]
    assert field(s1, "variant") == "15" and field(s1, "tpb") == "16" and field(s1, "lds") == "69632"
    p.close()
    monkeypatch.delenv("ASDFFT_PAIR2D")
    monkeypatch.delenv("ASDFFT_QUAD2D")
    # 512-point columns: two planes of 16 columns fit only twice in a CU, so the one-plane kernel
    # (four per CU) takes them; 256-point columns fit four times either way and keep two planes
    p = Plan("c2c", (512, 512), 4)
    s1 = steps(p.describe())[1]
    assert field(s1, "variant") == "15" and field(s1, "lds") == "34816", s1
    p.close()
    p = Plan("c2c", (256, 256), 4)
    assert field(steps(p.describe())[1], "variant") == "1"
    p.close()
    p = Plan("c2c", (256, 256, 256), 1)
    assert len(steps(p.describe())) == 3 and p.workspace_size == 0
    p.close()
    p = Plan("c2r", (64, 96, 80), 2)
    s = steps(p.describe())
    assert len(s) == 3 and "[in->work]" in s[0] and "[work->out]" in s[-1]
    assert p.workspace_size == 2 * 64 * 96 * 41 * 8
    p.close()


def test_lds_within_gfx950_limit():
    for kind, dims, b in [("c2c", (16384,), 8), ("c2c", (2048, 2048), 4), ("c2c", (4096, 4096), 1),
                          ("r2c", (32768,), 2), ("c2c", (1 << 20,), 1), ("c2c", (19 * 17 * 13,), 3)]:
        p = Plan(kind, dims, b)
        for line in steps(p.describe()):
            lds = int(field(line, "lds") or 0)
            assert lds <= 160 * 1024, line
            if "stockham" in line:
                # run-time kernels are bounded at 512 threads; a specialised kernel's bound is its
                # own workgroup, at most the hardware's 1024
                assert int(field(line, "block")) <= (1024 if "static=" in line else 512), line
        p.close()


def test_every_smooth_length_up_to_16384_has_a_plan():
    """{2^a 3^b 5^c} lengths whose fewest-pass radix sequence needs more threads than a
    workgroup holds (e.g. 14400, 15625) fall back to another multiset or a larger register
    budget instead of failing."""
    import bench
    from sip_amd.fft import Plan, RaggedPlan
    for n in bench.smooth_lengths(2, 16384):
        for kind in ("c2c", "r2c", "c2r"):
            Plan(kind, (n,), 2).close()
    work = bench.ragged_workload()
    p = RaggedPlan([(k, n, 1, False) for k, n in work])
    try:
        assert "4000 entries" in p.describe()
    finally:
        p.close()


def test_large_real_lengths_plan():
    """Real lengths beyond one workgroup: even N (N/2 > 16384) run the four-step on N/2 with one
    post- or pre-twiddle pass; odd N > 16384 run it on N between two elementwise passes."""
    for kind, n, first, last in [("r2c", 65536, "four-step A", "r2c-post"), ("c2r", 65536, "c2r-pre", "four-step B"),
                                 ("r2c", 131072, "four-step A", "r2c-post"), ("r2c", 59049, "promote", "truncate"),
                                 ("c2r", 59049, "hermitian-extend", "real-part"), ("r2c", 40000, "four-step A", "r2c-post")]:
        p = Plan(kind, (n,), 2)
        s = steps(p.describe())
        m = n // 2 if n % 2 == 0 else n
        assert first in s[0] and last in s[-1], (kind, n, s)
        assert p.workspace_size == 2 * m * 2 * 8, (kind, n, p.workspace_size)
        p.close()


# Column passes whose in-transform element offsets pass 2^31: (n - 1) * stride for the axis-0
# pass of a 3D plan (stride Y * Z) and for VERTICAL 1D (stride = batch).  The kernels take
# 64-bit strides and form every element offset as a 64-bit product, so such plans are accepted
# and described with their full spans (nothing is allocated or launched here).
@pytest.mark.parametrize("make,n,stride", [
    (lambda L, h: L.asdFftMakePlan3D(h, 4096, 1024, 520, N.ASCEND_FFT_C2C, N.ASCEND_FFT_FORWARD, 1), 4096, 1024 * 520),
    (lambda L, h: L.asdFftMakePlan1D(h, 2048, N.ASCEND_FFT_C2C, N.ASCEND_FFT_FORWARD, 1100000,
                                     N.ASCEND_FFT_VERTICAL), 2048, 1100000),
])
def test_column_spans_beyond_int32(make, n, stride):
    import re
    L = N.lib()
    h = ctypes.c_void_p()
    assert L.asdFftCreate(ctypes.byref(h)) == 0
    try:
        assert make(L, h) == 0
        buf = ctypes.create_string_buffer(1 << 14)
        need = ctypes.c_size_t()
        assert L.asdFftExtDescribePlan(h, buf, len(buf), ctypes.byref(need)) == 0
        text = buf.value.decode()
        spans = [(int(m.group(1)), int(m.group(2)), int(m.group(3)), int(m.group(4)))
                 for m in re.finditer(r"n=(\d+) .*stride=(\d+)/(\d+) span=(\d+)/", text)]
        hit = [sp for sp in spans if sp[0] == n]
        assert hit, text
        nn, s_in, s_out, span = hit[0]
        assert s_in == stride and s_out == stride
        assert span == (n - 1) * stride > 2 ** 31
    finally:
        L.asdFftDestroy(h)


def test_lane_butterfly_rows_are_opt_in(monkeypatch):
    """ASDFFT_WAVE=1 puts contiguous C2C rows of n = e * lanes (e in {16, 12, 10, 9}, lanes a
    power of two from 4 to 64) on the lane-butterfly kernel; unset or 0, and for every other
    length or layout, the Stockham kernels keep them."""
    want = {1024: (16, 64), 256: (16, 16), 64: (16, 4), 768: (12, 64), 640: (10, 64), 576: (9, 64), 36: (9, 4)}
    for n, (e, lanes) in want.items():
        p = Plan("c2c", (n,), 512)
        assert "wave n=" not in p.describe()
        p.close()
        monkeypatch.setenv("ASDFFT_WAVE", "1")
        p = Plan("c2c", (n,), 512)
        s = steps(p.describe())
        assert len(s) == 1 and f"wave n={n} count=512 points/lane={e} lanes={lanes}" in s[0], s
        p.close()
        monkeypatch.delenv("ASDFFT_WAVE")
    monkeypatch.setenv("ASDFFT_WAVE", "1")
    for n in (1000, 4096, 2048, 32, 15 * 64):  # no instance: 1000 = 8 x 125, 4096 / 16 = 256 lanes
        p = Plan("c2c", (n,), 64)
        assert "wave n=" not in p.describe(), n
        p.close()
    p = Plan("c2c", (1024,), 64, vertical=True)  # strided transforms stay on the column kernels
    assert "wave n=" not in p.describe()
    p.close()


def test_in_place_steps_read_and_write_the_same_elements():
    """A step whose source and destination are the same buffer must address its input and its
    output identically (stride and span), or one workgroup could overwrite what another has yet
    to read.  The paired 2D plans once read their column inputs at stride Y and wrote at 2Y in
    place, and raced on the GPU."""
    shapes = [("c2c", (2048, 2048), 4), ("c2c", (1024, 1024), 16), ("c2c", (2048, 1024), 2), ("c2c", (1000, 1000), 1),
              ("c2c", (64, 96, 80), 2), ("c2c", (256, 256, 256), 1), ("c2c", (512, 512, 8), 3), ("r2c", (2048, 2048), 1),
              ("c2r", (512, 1024), 2), ("r2c", (64, 64, 64), 2), ("c2c", (65536,), 8), ("c2c", (131072,), 2)]
    checked = 0
    for kind, dims, batch in shapes:
        p = Plan(kind, dims, batch)
        for line in steps(p.describe()):
            m = re.match(r"\[(\w+)->(\w+)\]", line)
            if m is None or m.group(1) != m.group(2):
                continue
            st = re.search(r"stride=(\d+)/(\d+) span=(\d+)/(\d+)", line)
            if st is None:
                continue
            checked += 1
            assert st.group(1) == st.group(2) and st.group(3) == st.group(4), (kind, dims, batch, line)
        p.close()
    assert checked >= 4


def test_2d_rows_in_fours(monkeypatch):
    """2D C2C with both lengths in {1024, 2048}: the row pass takes four rows per workgroup and
    the columns' first radix-4 stage (variant 23), the quarter-length columns run in place on
    the one-plane kernel (variant 15) at stride 4Y.  ASDFFT_QUAD2D=0 falls back to the pairs."""
    for (X, Y), b in (((2048, 2048), 4), ((1024, 1024), 16), ((2048, 1024), 2)):
        p = Plan("c2c", (X, Y), b)
        s = steps(p.describe())
        p.close()
        assert len(s) == 2 and field(s[0], "variant") == "24" and field(s[0], "tpb") == "4", s
        assert field(s[1], "variant") == "15" and field(s[1], "n") == str(X // 4), s
        assert f"stride={4 * Y}/{4 * Y}" in s[1], s
        monkeypatch.setenv("ASDFFT_QUAD2D", "0")
        p = Plan("c2c", (X, Y), b)
        s = steps(p.describe())
        p.close()
        monkeypatch.delenv("ASDFFT_QUAD2D")
        assert field(s[0], "variant") in ("0", "21"), s


def test_streamed_input_on_multi_pass_plans(monkeypatch):
    """The first launch of a plan whose later launch reads back what it wrote (2D, 3D, the
    four-step) loads its input non-temporally where its kernel has a streamed path (row
    passes, the four-step's first pass); single-launch plans, first passes that do not read
    interleaved complex points (odd-length R2C rows), the twiddle-preloading small-launch
    kernel, first column passes (multi-dimensional C2R) and plans whose input and intermediate
    together are under the 256 MiB device cache (65536 x 64, 128^3 x 4, 1000^2 x 16) do not.
    ASDFFT_STREAM_IN=0 turns it off, =1 on wherever the first kernel has the path."""
    marked = [("c2c", (2048, 2048), 4), ("c2c", (256, 256, 256), 1), ("c2c", (65536,), 256),
              ("r2c", (4096, 4096), 2), ("c2c", (1000, 1000), 20)]
    unmarked = [("c2c", (4096,), 256), ("r2c", (8192,), 16), ("r2c", (4095, 4095), 2), ("c2c", (1, 4096), 3),
                ("c2r", (64, 64), 2), ("c2r", (256, 2046), 1), ("c2c", (1000, 1000), 1),
                ("c2c", (65536,), 64), ("c2c", (128, 128, 128), 4), ("c2c", (1000, 1000), 16)]
    for kind, dims, batch in marked + unmarked:
        p = Plan(kind, dims, batch)
        s = steps(p.describe())
        p.close()
        want = (kind, dims, batch) in marked
        assert ("streamed-input" in s[0]) == want, (kind, dims, batch, s)
        assert not any("streamed-input" in line for line in s[1:]), s
    monkeypatch.setenv("ASDFFT_STREAM_IN", "0")
    for kind, dims, batch in marked:
        p = Plan(kind, dims, batch)
        assert "streamed-input" not in p.describe(), (kind, dims, batch)
        p.close()
    monkeypatch.setenv("ASDFFT_STREAM_IN", "1")
    for kind, dims, batch in [("c2c", (65536,), 64), ("c2c", (128, 128, 128), 4)]:
        p = Plan(kind, dims, batch)
        assert "streamed-input" in steps(p.describe())[0], (kind, dims, batch)
        p.close()


def test_4096_rows_with_one_transform_per_cu_run_wide(monkeypatch):
    """4096-point C2C rows with at most one transform per CU run four radix-8 passes on 512
    threads x 8 points (a specialised instance); two or more per CU keep 16x16x16 on 256 threads.
    ASDFFT_WIDE8=0 turns it off."""
    for b, radices, threads in ((256, "8x8x8x8", "512"), (1, "8x8x8x8", "512"), (512, "16x16x16", "256"),
                                (4096, "16x16x16", "256")):
        p = Plan("c2c", (4096,), b)
        s = steps(p.describe())[0]
        p.close()
        assert field(s, "radices") == radices and field(s, "T") == threads, (b, s)
    monkeypatch.setenv("ASDFFT_WIDE8", "0")
    p = Plan("c2c", (4096,), 256)
    s = steps(p.describe())[0]
    p.close()
    assert field(s, "radices") == "16x16x16", s
