"""gabe_means_vars_noalloc (the entry point that allocates nothing) against
gabe_means_vars, through the C ABI, on two calls in a row, the second with a
smaller input: bitwise equal for FP64 fields (the same kernel, workgroup
count and host summation), and the FP32 path against a two-pass float64
reduction of the same floats (means within 1e-15 of the base, variances
within 1e-9 relative)."""
import numpy as np
import pytest

from gabe_amd.lattice import GABE_FP32, GABE_FP64, GabeError, means_vars_noalloc
from tests.helpers import configs, lattice_from_oracle, seeded_oracle

pytestmark = pytest.mark.gpu


class DeviceCopy:
    """Caller-owned device memory holding a copy of a host array, allocated
    through the HIP runtime the library itself links (/opt/rocm/lib)."""

    def __init__(self, a):
        import ctypes as C
        self._hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        assert self._hip.hipMalloc(C.byref(self.ptr), C.c_size_t(a.nbytes)) == 0
        assert self._hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        self.nbytes = a.nbytes

    def data_ptr(self):
        return self.ptr.value

    def close(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = None


def device_copy(a):
    return DeviceCopy(a)


@pytest.mark.parametrize("precision", [GABE_FP64, GABE_FP32])
def test_noalloc_matches_means_vars_twice(precision):
    for N in (32, 16):   # the second call on a smaller input
        cfg = configs.lambda_phi4_config(N=N, calc_gws=0)
        o = seeded_oracle(cfg, amp=2e-3)
        lat = lattice_from_oracle(o, cfg, GABE_FP64)
        lat.step(3)
        f, _ = lat.download_state()
        if precision == GABE_FP64:
            ref = lat.meansvars()
            dev = device_copy(f)
            got = means_vars_noalloc(dev.data_ptr(), GABE_FP64, N ** 3, 2, stream=None)
            assert np.array_equal(got, ref), (N, got, ref)
        else:
            base = f.reshape(2, -1).mean(axis=1)
            x = (f.reshape(2, -1) - base[:, None]).astype(np.float32)
            dev = device_copy(x)
            got = means_vars_noalloc(dev.data_ptr(), GABE_FP32, N ** 3, 2, base=base, stream=None)
            xd = x.astype(np.float64)
            for fld in range(2):
                m = xd[fld].mean()
                v = ((xd[fld] - m) ** 2).mean()
                assert abs(got[2 * fld] - (base[fld] + m)) <= 1e-15 * max(abs(base[fld]), 1e-3)
                assert abs(got[2 * fld + 1] - v) <= 1e-9 * v
        dev.close()
        lat.close()


def test_noalloc_rejects_bad_sizes():
    dev = device_copy(np.zeros(8))
    with pytest.raises(GabeError):
        means_vars_noalloc(dev.data_ptr(), GABE_FP64, 0, 1)
    with pytest.raises(GabeError):
        means_vars_noalloc(dev.data_ptr(), GABE_FP64, (1 << 33) + 1, 1)
    with pytest.raises(GabeError):
        means_vars_noalloc(dev.data_ptr(), GABE_FP64, 8, 3)
    dev.close()
