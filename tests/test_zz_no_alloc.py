"""The allocation-free path sum (chm_fm_chi_noalloc) through the C ABI: on device buffers
it gives what chm_fm_chi gives for the same inputs, on two calls in a row (the second with
fewer configurations, so the shared scratch array holds the first call's values), flags a
configuration with an out-of-range species with NaN rows, runs on a caller's stream, and
refuses a call whose intermediate arrays exceed the scratch array."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers and a stream for the test, from the HIP runtime the library itself
    links (libamdhip64.so.7, already loaded with it)."""

    def __init__(self):
        from chimera_amd import _lib
        _lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                           ("hipStreamDestroy", [C.c_void_p]),
                           ("hipStreamSynchronize", [C.c_void_p]),
                           ("hipDeviceSynchronize", [])):
            getattr(self.h, name).argtypes = args
            getattr(self.h, name).restype = C.c_int

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), nbytes))
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self.ok(self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))        # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        self.ok(self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2))    # device to host
        return out


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tern():
    from chimera_amd import workloads as W
    _say("[build] noalloc t108")
    return W.Problem(W.Workload("na108", "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, False, 2))


def _configs(P, n, seed):
    rng = np.random.default_rng(seed)
    sp = np.stack([rng.permutation(P.cfg0) for _ in range(n)]).astype(np.int32)
    pos = np.repeat(P.cell.ideal[None], n, 0) + rng.normal(0, 0.02, (n,) + P.cell.ideal.shape)
    return np.ascontiguousarray(pos), sp


def _noalloc(P, pos, sp, precision, stream=None):
    from chimera_amd import _lib
    hip = _Hip()
    fm = P.fm
    n = sp.shape[0]
    dpos, dsp = hip.put(pos), hip.put(sp.astype(np.int32))
    dchi = hip.put(np.full((n, fm.nsp, fm.nk), 7.0))
    try:
        _lib.check(_lib.load().chm_fm_chi_noalloc(fm.ctx.h, n, dpos, dsp, dchi, precision, stream))
        if stream:
            hip.ok(hip.h.hipStreamSynchronize(stream))
        fm.ctx.sync()
        return hip.get(dchi, (n, fm.nsp, fm.nk), np.float64)
    finally:
        for p in (dpos, dsp, dchi):
            hip.h.hipFree(p)


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_chm_fm_chi(tern, precision):
    P = tern
    prec = "f64" if precision == 0 else "f32"
    for n, seed in ((3, 1), (1, 2)):                 # second call smaller than the first
        _say(f"[test] noalloc {prec} nconf={n}")
        pos, sp = _configs(P, n, seed)
        want = P.fm.chi(pos, sp, normalise=False, precision=prec)
        got = _noalloc(P, pos, sp, precision)
        assert np.all(np.isfinite(got))
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-12 * scale
        assert np.array_equal(got, want)            # the same kernels in the same order


def test_noalloc_on_caller_stream_and_bad_species(tern):
    P = tern
    pos, sp = _configs(P, 2, 5)
    want = P.fm.chi(pos, sp, normalise=False)
    hip = _Hip()
    st = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(st)))
    try:
        got = _noalloc(P, pos, sp, 0, stream=st.value)
    finally:
        hip.h.hipStreamDestroy(st)
    assert np.array_equal(got, want)
    bad = sp.copy()
    bad[1, 17] = 3                                    # species index out of range (nsp = 3)
    got = _noalloc(P, pos, bad, 0)
    assert np.array_equal(got[0], want[0])
    assert np.all(np.isnan(got[1]))


def test_noalloc_rejects_oversized_call(tern):
    """Sizes beyond the scratch array fail on the host before any launch."""
    from chimera_amd import _lib
    fm = tern.fm
    n = 100000                                        # 100000 x 47 KB of distances alone
    hip = _Hip()
    d = hip.alloc(256)
    try:
        with pytest.raises(ValueError, match="scratch"):
            _lib.check(_lib.load().chm_fm_chi_noalloc(fm.ctx.h, n, d, d, d, 0, None))
    finally:
        hip.h.hipFree(d)
