"""dsph_energy_sums_noalloc against dsph_diagnostics (the entry point whose sums
it restates over caller-owned device arrays, allocating nothing): two calls in a
row on the same GPU, the second with fewer particles, each within 1e-14 of the
context's compensated sums; and a size beyond the fixed scratch is refused
before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from deorbit_amd import dsph, ics

pytestmark = pytest.mark.gpu


class Hip:
    """The HIP runtime libdsph.so is linked against (device buffers for the test)."""

    def __init__(self):
        dsph.load_library()
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0      # host to device
        self.bufs.append(p)
        return p

    def download(self, p, n):
        out = np.empty(n, np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def context_sums(ic):
    s = dsph.make_system(ic, dsph.FP64)
    s.init()
    s.step(1e-4)
    d = s.diagnostics()
    st = {f: s.get(f) for f in ("mass", "vel", "u")}
    s.close()
    return d, st


def test_noalloc_sums_match_dsph_diagnostics_on_two_calls(gpu):
    lib = dsph.load_library()
    hip = Hip()
    try:
        out = hip.upload(np.full(8, np.nan))
        # 36,937 particles (three chunks), then 1,217 (one chunk) through the same scratch
        for ic in (ics.impact_37k(), ics.tiny_impact()):
            d, st = context_sums(ic)
            n = len(st["mass"])
            v = st["vel"]
            ptrs = [hip.upload(st["mass"]), hip.upload(v[:, 0]), hip.upload(v[:, 1]), hip.upload(v[:, 2]),
                    hip.upload(st["u"])]
            assert lib.dsph_energy_sums_noalloc(*ptrs, n, None, out) == 0, lib.dsph_last_error()
            assert hip.rt.hipDeviceSynchronize() == 0
            got = hip.download(out, 8)
            want = [d["mass"], d["e_kin"], d["e_int"], d["e_tot"], *d["mom"]]
            print(f"n={n}: " + " ".join(f"{g:.17g}/{w:.17g}" for g, w in zip(got, want)), flush=True)
            for g, w in zip(got[:4], want[:4]):
                assert abs(g - w) <= 1e-14 * abs(w)
            pscale = float(np.abs(st["mass"][:, None] * v).sum())
            for g, w in zip(got[4:7], want[4:7]):
                assert abs(g - w) <= 1e-14 * pscale
            assert got[7] == 0.0
        # beyond the scratch: refused with DSPH_E_RANGE, nothing launched (out unchanged)
        before = hip.download(out, 8)
        rc = lib.dsph_energy_sums_noalloc(*ptrs, 67_108_865, None, out)
        assert rc == -5 and b"exceeds" in lib.dsph_last_error()
        assert hip.rt.hipDeviceSynchronize() == 0
        assert np.array_equal(hip.download(out, 8), before)
    finally:
        hip.free()
