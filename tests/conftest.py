import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "elvibrot-tnumtana_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libevrhip.so)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def evr():
    import evrhip
    evrhip._lib.load()
    return evrhip


def hh_potential(q, lam=0.111803):
    """Henon-Heiles potential of the reference's 6D/21D working tests
    (Working_tests/MPI_tests/6D_Davidson_S0/sub_system_HenonHeiles.f:33-48)."""
    D = q.shape[1]
    V = 0.5 * (q ** 2).sum(1)
    for i in range(D - 1):
        V = V + lam * (q[:, i] ** 2 * q[:, i + 1] - q[:, i + 1] ** 3 / 3.0)
    return V


class DeviceBuffer:
    """Device memory for tests that call the C ABI with device pointers,
    through the HIP runtime the library itself links (no second runtime in
    the process)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes
            import evrhip
            evrhip._lib.load()                       # the library's runtime first
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
            h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        import ctypes
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = self.hip().hipMalloc(ctypes.byref(p), self.nbytes)
        if rc != 0:
            raise RuntimeError("hipMalloc failed: %d" % rc)
        self.ptr = p.value

    @classmethod
    def of(cls, a):
        import numpy as np
        a = np.asfortranarray(a, dtype=np.float64)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        import numpy as np
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.nbytes <= self.nbytes
        rc = self.hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1)     # host to device
        assert rc == 0, rc

    def fill_nan(self):
        rc = self.hip().hipMemset(self.ptr, 0xFF, self.nbytes)               # all-ones bits: NaN
        assert rc == 0, rc
        assert self.hip().hipDeviceSynchronize() == 0    # before work on non-blocking streams

    @classmethod
    def sync(cls, stream):
        """Wait for a (non-blocking) stream of the library: plain hipMemcpy does not."""
        import ctypes
        rc = cls.hip().hipStreamSynchronize(ctypes.c_void_p(stream))
        assert rc == 0, rc

    def download(self, shape):
        import numpy as np
        out = np.zeros(shape, order="F")
        assert out.nbytes <= self.nbytes
        rc = self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2)  # device to host
        assert rc == 0, rc
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
