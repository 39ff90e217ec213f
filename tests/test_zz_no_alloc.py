"""The allocation-free g(r) entry point (gs_radial_distribution_noalloc)
against gs_radial_distribution, through the C ABI: bit-identical bin counts
on two calls in a row, the second with a smaller universe and fewer bins,
so anything the first call left in the shared scratch area would show."""
import ctypes

import numpy as np
import pytest

from genesis_amd.inputs import mixture

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


@pytest.fixture(scope="module")
def hip():
    # the HIP runtime libgenesis_gpu.so links: device buffers for the output
    from genesis_amd import _lib
    _lib.load()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def noalloc_counts(hip, u, r_max, nbins):
    from genesis_amd import _lib
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 8 * nbins) == 0
    try:
        # garbage in the output: the entry point must clear it itself
        assert hip.hipMemset(dev, 0x5A, 8 * nbins) == 0
        assert hip.hipDeviceSynchronize() == 0
        rm, dr = ctypes.c_double(), ctypes.c_double()
        _lib.check(u.lib.gs_radial_distribution_noalloc(u.h, r_max, nbins, dev, ctypes.byref(rm), ctypes.byref(dr)))
        assert hip.hipDeviceSynchronize() == 0
        out = np.zeros(nbins, dtype=np.uint64)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, 8 * nbins, 2) == 0
    finally:
        hip.hipFree(dev)
    return out, rm.value, dr.value


def test_noalloc_counts_equal_the_allocating_entry_point_on_two_calls(hip):
    from genesis_amd.universe import GpuUniverse
    cases = [(32768, 1002, 0.25, 200), (3000, 5, 0.2, 37)]
    for n, seed, frac, nbins in cases:
        w = mixture(n=n, seed=seed, steps=1)
        r_max = frac * w.physics.box[0]
        say(f"n={n} r_max={r_max:.3f} bins={nbins}")
        with GpuUniverse(w.physics, w.state) as u:
            u.step(3)
            ref, _, rm_ref, dr_ref = u.radial_distribution(r_max, nbins)
            got, rm, dr = noalloc_counts(hip, u, r_max, nbins)
        assert (rm, dr) == (rm_ref, dr_ref)
        assert ref.sum() > 0
        assert np.array_equal(got, ref)


def test_noalloc_rejects_bins_beyond_its_limit(hip):
    from genesis_amd import _lib
    from genesis_amd.universe import GpuUniverse
    w = mixture(n=512, seed=3, steps=1)
    with GpuUniverse(w.physics, w.state) as u:
        for nb in (0, 65537):
            rc = u.lib.gs_radial_distribution_noalloc(u.h, 4.0, nb, ctypes.c_void_p(256), None, None)
            assert rc == _lib.GS_E_INVALID
            assert b"65536" in u.lib.gs_last_error()
