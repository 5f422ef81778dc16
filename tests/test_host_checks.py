"""CPU tests of the host mirror's argument checks: tensors the device entry points cannot
take (host tensors, a wrong dtype, a too short buffer, a malformed batch layout) raise
ValueError before any handle is created or any library function runs, so a bad call never
hands a host pointer or a short buffer to a kernel."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.fixture(autouse=True)
def no_library_calls(sep, monkeypatch):
    """Any handle creation or library call fails the test loudly."""
    def forbidden(*a, **k):
        raise AssertionError("the library was called before the arguments were checked")
    monkeypatch.setattr(sep, "default_handle", forbidden)
    monkeypatch.setattr(sep, "lib", forbidden)


def _t(*shape, dtype=torch.float64):
    return torch.zeros(*shape, dtype=dtype)


def test_bulk_apply_rejects_host_tensors_and_dtype_mismatch(sep):
    A = _t(8, 6).t()  # column-major, CPU
    g = _t(4, 7, 2)
    with pytest.raises(ValueError, match="GPU tensor"):
        sep.bulk_wave_order_2x2_rmul_(A, g)
    with pytest.raises(ValueError, match="numpy"):
        sep.bulk_wave_order_2x2_rmul_(np.asfortranarray(np.zeros((6, 8))), g)


def test_qr_algorithm_rejects_host_tensors(sep):
    H = sep.SymTridiagonal(_t(6), _t(5))
    with pytest.raises(ValueError, match="GPU tensor"):
        sep.qr_algorithm_(H, None)
    Hn = sep.SymTridiagonal(np.ones(6), np.ones(5))
    with pytest.raises(ValueError, match="numpy"):
        sep.qr_algorithm_(Hn, _t(6, 6))


def test_sort_and_backward_reject_host_tensors(sep):
    with pytest.raises(ValueError, match="GPU tensor"):
        sep.sort_eigen(_t(5))
    with pytest.raises(ValueError, match="GPU tensor"):
        sep.eig_backward(_t(4), _t(4, 4), _t(4), _t(4, 4))
    with pytest.raises(ValueError, match="sort=True needs device tensors"):
        sep.eigen_decomp_(sep.SymTridiagonal(np.ones(4), np.ones(3)), sort=True)


def test_batch_layout_checks(sep):
    with pytest.raises(ValueError, match="one length"):
        sep._batch_layout([4, 6], [0], [0, 16])
    with pytest.raises(ValueError, match="non-negative"):
        sep._batch_layout([4], [-1], [0])
    ns, od, oq, need_d, need_q = sep._batch_layout([4, 6], [0, 4], [0, 16])
    assert (need_d, need_q) == (10, 52)
    with pytest.raises(ValueError, match="GPU tensor"):
        sep.qr_batched_([4, 6], [0, 4], [0, 16], _t(10), _t(10), _t(52))
    with pytest.raises(ValueError, match="GPU tensor"):
        sep.eig_backward_batched([4], [0], [0], _t(4), _t(16), _t(4), _t(16))


def test_require_device_checks_dtype_length_and_stride(sep):
    class FakeCuda:
        """Just enough of a tensor for _require_device (no GPU here)."""
        __module__ = "torch.fake"

        def __init__(self, n, dtype=torch.float64, stride=1, dim=1):
            self.device = torch.device("cuda", 0)
            self.dtype, self._n, self._s, self._d = dtype, n, stride, dim

        def numel(self):
            return self._n

        def stride(self, i):
            return self._s

        def dim(self):
            return self._d

    sep._require_device("x", FakeCuda(8), dtype=torch.float64, min_len=8, vector=True)
    with pytest.raises(ValueError, match="dtype"):
        sep._require_device("x", FakeCuda(8, torch.float32), dtype=torch.float64)
    with pytest.raises(ValueError, match="at least 9"):
        sep._require_device("x", FakeCuda(8), min_len=9)
    with pytest.raises(ValueError, match="stride 1"):
        sep._require_device("x", FakeCuda(8, stride=2), vector=True)
