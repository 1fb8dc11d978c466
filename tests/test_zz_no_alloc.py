"""cp3_atmos_calculate_async_static, the entry point that allocates nothing
(its slab table and job descriptor live in a static device array), against
cp3_atmos_calculate_async on the same object: bitwise the same results, on
two calls in a row with the second on a smaller grid, in each precision and
solve mode, then the refusals (a table over the array's size, inputs not yet
resident).  Called through the C ABI with ctypes; the output buffers are
the test's own hipMalloc allocations."""
import ctypes as C

import numpy as np
import pytest

from cudaprob3_amd import workloads as W

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


class DevBuf:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        # a fill the kernels must overwrite everywhere; hipMemset returns
        # before the fill is done and runs on the null stream, which the
        # library's non-blocking streams do not wait for, so the fill is
        # completed here before any kernel writes the buffer
        assert hip().hipMemset(self.p, 0x7f, C.c_size_t(nbytes)) == 0
        assert hip().hipDeviceSynchronize() == 0

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def make(precision, earth, nC, nE, mode="auto", poly=None, params=W.NOMINAL):
    from cudaprob3_amd import AtmosPropagator
    p = AtmosPropagator(nC, nE, precision=precision)
    p.setSolveMode(mode)
    p.setEnergyList(W.log_grid(nE, 0.5, 50.0))
    p.setCosineList(W.cos_grid(nC))
    t12, t13, t23, d, k, m12, m23 = params.for_nutype(0)
    p.setMNSMatrix(t12, t13, t23, d, k)
    p.setNeutrinoMasses(m12, m23)
    if poly is not None:
        p.setDensity(poly.radii, poly.a, poly.b, poly.c, poly.yps)
    else:
        p.setDensity(earth.radii, earth.rhos, earth.yps)
    p.setProductionHeight(22.0)
    return p


def static_call(p, nutype, buf):
    from cudaprob3_amd import _native as N
    lib = N.load()
    rc = lib.cp3_atmos_calculate_async_static(p._h, int(nutype), buf.p)
    return rc, lib.cp3_last_error().decode()


def compare(p, nutype, precision):
    """The object's own calculation, then the static variant into a fresh
    buffer: the same bits."""
    from cudaprob3_amd import _native as N
    p.calculateProbabilities(nutype)
    own = np.transpose(p.probabilities(), (2, 0, 1))  # [9][nC][nE]
    el = 8 if precision == "fp64" else 4
    buf = DevBuf(9 * p.n_cosines * p.n_energies * el)
    try:
        rc, msg = static_call(p, nutype, buf)
        assert rc == 0, msg
        N.check(N.load().cp3_atmos_wait(p._h))
        got = buf.get(np.float64 if el == 8 else np.float32,
                      (9, p.n_cosines, p.n_energies)).astype(np.float64)
    finally:
        buf.free()
    assert np.array_equal(got, own)
    return got


@pytest.mark.parametrize("precision", ["mixed", "fp64", "fp32"])
@pytest.mark.parametrize("mode", ["table", "shared"])
def test_static_equals_allocating_two_calls(precision, mode):
    e59, e4 = W.earth_59(), W.prem4()
    print("static entry point", precision, mode, flush=True)
    # first call: 59 shells on 400 x 400 (mixed: 60 x 448 x 40 B table)
    big = make(precision, e59, 400, 400, mode)
    first = compare(big, 0, precision)
    # second call in a row, smaller: 4 shells on 37 x 70, both types
    small = make(precision, e4, 37, 70, mode)
    compare(small, 1, precision)
    compare(small, 0, precision)
    # and the first object again, after the array held the smaller table
    again = compare(big, 0, precision)
    assert np.array_equal(again, first)
    big.close()
    small.close()


def test_static_polynomial_and_averaging():
    """The per-trajectory solve of a polynomial model (no table) and the
    averaging kernels (table) through the static variant."""
    p = make("mixed", None, 50, 90, poly=W.prem4_poly())
    compare(p, 0, "mixed")
    p.close()
    p = make("fp64", W.prem4(), 40, 70)
    nb = 3
    p.SetNumberOfProductionHeightBinsForAveraging(nb)
    p.setProductionHeightAveraging(True)
    prob = np.full(2 * 3 * 40 * 70 * nb, 1.0 / nb)
    p.setProductionHeightList(prob, [10.0, 20.0, 30.0, 40.0])
    # a normal calculation uploads the bin shifts; then the two agree
    p.calculateProbabilities(0)
    compare(p, 1, "fp64")
    p.close()


def test_static_refusals():
    from cudaprob3_amd import _native as N
    e59 = W.earth_59()
    # 59 shells x 70000 energies: a 168 MB table, over the 16 MiB array
    p = make("mixed", e59, 2, 70000)
    p.calculateProbabilities(0)
    buf = DevBuf(9 * 2 * 70000 * 4)
    try:
        rc, msg = static_call(p, 0, buf)
        assert rc == N.CP3_ERR_INVALID
        assert "static array" in msg
        # the shared solve mode needs no table: it runs
        p.setSolveMode("shared")
        rc, msg = static_call(p, 0, buf)
        assert rc == 0, msg
        N.check(N.load().cp3_atmos_wait(p._h))
    finally:
        buf.free()
    p.close()
    # inputs changed since the last upload
    p = make("mixed", W.prem4(), 20, 30)
    buf = DevBuf(9 * 20 * 30 * 4)
    try:
        rc, msg = static_call(p, 0, buf)
        assert rc == N.CP3_ERR_STATE and "uploads them first" in msg
        p.calculateProbabilities(0)
        assert static_call(p, 0, buf)[0] == 0
        N.check(N.load().cp3_atmos_wait(p._h))
        p.setEnergyList(W.log_grid(30, 1.0, 10.0))
        assert static_call(p, 0, buf)[0] == N.CP3_ERR_STATE
        # parameter changes need no upload
        p.calculateProbabilities(0)
        p.setMNSMatrix(0.6, 0.15, 0.8, 0.3, 1)
        assert static_call(p, 1, buf)[0] == 0
        N.check(N.load().cp3_atmos_wait(p._h))
        assert N.load().cp3_atmos_calculate_async_static(p._h, 0, None) == N.CP3_ERR_INVALID
        assert N.load().cp3_atmos_calculate_async_static(p._h, 2, buf.p) == N.CP3_ERR_INVALID
    finally:
        buf.free()
    p.close()
