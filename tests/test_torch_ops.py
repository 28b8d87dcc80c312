"""torch_sip operators of libasdfft_torch.so (CPU): registration, the argument checks of the
reference's plugin operators (sip_pta/csrc/fft/*.cpp; TORCH_CHECK -> RuntimeError) and the
plan LRU controls.  CPU tensors pass every check and are then refused: the operators have no
host path."""
import pytest

torch = pytest.importorskip("torch")

from sip_amd import _native as N  # noqa: E402


@pytest.fixture(scope="module")
def ops():
    return N.torch_ops()


def test_schemas_registered(ops):
    for name in ("asd_fft_c2c", "asd_fft_c2c_sep", "asd_fft_r2c", "asd_fft_c2r", "asd_fft_shift",
                 "asd_fft_shift_auto", "asd_ifft_shift", "asd_ifft_shift_auto", "asd_fft_plan_cache_size",
                 "asd_fft_plan_cache_set_capacity", "asd_fft_plan_cache_clear"):
        assert hasattr(ops, name), name
    s = str(torch.ops.torch_sip.asd_ifft_shift.default._schema)
    assert "int dist=0" in s and "float scale=1." in s, s
    assert "(Tensor, Tensor)" in str(torch.ops.torch_sip.asd_fft_c2c_sep.default._schema)


def test_c2c_checks(ops):
    with pytest.raises(RuntimeError, match="complex64"):
        ops.asd_fft_c2c(torch.zeros(4, 8), True)
    with pytest.raises(RuntimeError, match="2, 3 or 4 dimensions"):
        ops.asd_fft_c2c(torch.zeros(8, dtype=torch.complex64), True)
    with pytest.raises(RuntimeError, match="2, 3 or 4 dimensions"):
        ops.asd_fft_c2c(torch.zeros(1, 2, 2, 2, 2, dtype=torch.complex64), True)
    with pytest.raises(RuntimeError, match="HIP device"):
        ops.asd_fft_c2c(torch.zeros(4, 8, dtype=torch.complex64), True)


def test_real_checks(ops):
    with pytest.raises(RuntimeError, match="Float32"):
        ops.asd_fft_r2c(torch.zeros(4, 8, dtype=torch.complex64), True)
    with pytest.raises(RuntimeError, match="COMPLEX64"):
        ops.asd_fft_c2r(torch.zeros(4, 8), True)
    with pytest.raises(RuntimeError, match="2, 3 or 4 dimensions"):
        ops.asd_fft_r2c(torch.zeros(8), True)
    with pytest.raises(RuntimeError, match="HIP device"):
        ops.asd_fft_c2r(torch.zeros(4, 5, dtype=torch.complex64), False)


def test_shift_checks(ops):
    c = torch.zeros(2, 4, 6, dtype=torch.complex64)
    with pytest.raises(RuntimeError, match="complex64"):
        ops.asd_fft_shift(torch.zeros(2, 4, 6), 4, 6, 2, 0)
    with pytest.raises(RuntimeError, match="must be positive"):
        ops.asd_fft_shift(c, 0, 6, 2, 0)
    with pytest.raises(RuntimeError, match=r"must be >= row \* col"):
        ops.asd_fft_shift(c.reshape(-1), 4, 6, 2, 10)
    with pytest.raises(RuntimeError, match="needs at least"):
        ops.asd_fft_shift(c.reshape(-1), 4, 6, 3, 0)
    with pytest.raises(RuntimeError, match="must match explicit"):
        ops.asd_fft_shift(c, 6, 4, 2, 0)  # same numel, other shape (issue of the reference: #373)
    with pytest.raises(RuntimeError, match="auto mode expects a 3D"):
        ops.asd_ifft_shift_auto(c.reshape(8, 6))
    with pytest.raises(RuntimeError, match="HIP device"):
        ops.asd_ifft_shift(c, 4, 6, 2, 0, 0.5)


def test_python_wrappers(ops):
    from sip_amd import fft_shift, ifft_shift
    with pytest.raises(ValueError, match="auto-infer requires a 3D"):
        fft_shift(torch.zeros(8, 6, dtype=torch.complex64))
    with pytest.raises(ValueError, match="auto-infer requires a 3D"):
        ifft_shift(torch.zeros(8, 6, dtype=torch.complex64))
    with pytest.raises(RuntimeError, match="HIP device"):
        fft_shift(torch.zeros(2, 4, 6, dtype=torch.complex64))


def test_plan_cache_controls(ops):
    ops.asd_fft_plan_cache_clear()
    assert ops.asd_fft_plan_cache_size() == 0
    ops.asd_fft_plan_cache_set_capacity(3)
    ops.asd_fft_plan_cache_set_capacity(10)
    assert ops.asd_fft_plan_cache_size() == 0
