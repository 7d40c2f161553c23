"""The allocation-free apply_sink (gm_apply_sink_device) against gm_apply_sink
(needs an MI355X).

gm_apply_sink_device takes device buffers and reduces through the library's
static per-GPU scratch; gm_apply_sink stages host buffers through the
solver's own buffers and scratch. Both run the same kernels in the same
order, so their results agree bitwise. Two solvers of different sizes are
called one after the other, the second smaller, so the shared scratch is
reused with fewer modes; the last call enqueues on a caller's stream. Device
buffers and the stream come from the HIP runtime libgm.so itself uses.
"""
import ctypes as C

import numpy as np
import pytest

import gm

pytestmark = pytest.mark.gpu


def hip():
    """The HIP runtime libgm.so is linked with (one runtime in the process),
    for device buffers and a stream."""
    gm.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    assert paths, "libamdhip64 is not loaded"
    h = C.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


H2D, D2H = 1, 2


def solver(l_max, Nr, L):
    s = gm.CrankNicolsonModes(Nr=Nr, L=L, dt=0.5, epsilon=0.0, l_max=l_max, mu=0.2, M_c=0.05, q=0.5, a=0.0,
                              R=20.0, counter=False, absorption=True, save_every=100, ramp_periods=0.0,
                              epsilon_smear=4.0)
    s.set_flags(evolve_orbit=0)
    rng = np.random.default_rng(l_max)
    x = rng.standard_normal(s.size) + 1j * rng.standard_normal(s.size)
    s.set_initial_state_full(x)
    s.initialize()
    s.begin_run(0.5 * L, 2)
    # the absorbers ramp in over one initial period: step once past it
    s.set_scalars(t=2.0 * s.P)
    s.step(1, diag=False)
    assert s.scalars()["sink_active"] == 1
    assert s.sink_strength(1) > 0 and s.sink_strength(2) > 0
    return s


def check(s, seed, stream=None):
    h = hip()
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(s.size) + 1j * rng.standard_normal(s.size)
    ref = s.apply_sink(x)
    assert np.linalg.norm(ref) > 0
    nbytes = x.nbytes
    din, dout = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(din), nbytes) == 0 and h.hipMalloc(C.byref(dout), nbytes) == 0
    try:
        junk = np.full(s.size, np.nan + 1j * np.nan)
        assert h.hipMemcpy(din, x.ctypes.data, nbytes, H2D) == 0
        assert h.hipMemcpy(dout, junk.ctypes.data, nbytes, H2D) == 0
        if stream is None:
            s.apply_sink_device(din.value, dout.value)
        else:
            s.apply_sink_device(din.value, dout.value, stream)
            assert h.hipStreamSynchronize(C.c_void_p(stream)) == 0
        out = np.empty_like(x)
        assert h.hipMemcpy(out.ctypes.data, dout, nbytes, D2H) == 0
    finally:
        h.hipFree(din)
        h.hipFree(dout)
    assert np.array_equal(out, ref), np.max(np.abs(out - ref))


def test_allocation_free_apply_sink_equals_apply_sink():
    big = solver(8, 2049, 256.0)
    check(big, 1)
    small = solver(4, 513, 128.0)
    check(small, 2)
    h = hip()
    st = C.c_void_p()
    assert h.hipStreamCreate(C.byref(st)) == 0
    try:
        check(big, 3, stream=st.value)
    finally:
        h.hipStreamDestroy(st)
    big.close()
    small.close()
