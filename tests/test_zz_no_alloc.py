"""cup2d_find_max_u_noalloc against cup2d_find_max_u, through the C-ABI.

The caller side (output buffer, stream) is set up with the HIP runtime the
library itself links, loaded through ctypes; the entry point under test
allocates nothing.  Two calls in a row on one stream, the second on a
smaller plane, each compared bit for bit with the context's findMaxU (a max
is exact), then a NaN cell (kept by both) and the size checks."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu case] {request.node.name}", flush=True)
    sys.stdout.flush()


def _hip():
    # by soname: the runtime already loaded with the library (torch's copy, when
    # torch was imported first); a second HIP runtime in the process sees no device
    h = ctypes.CDLL("libamdhip64.so.7")
    for n in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate", "hipStreamDestroy",
              "hipStreamSynchronize"):
        getattr(h, n).restype = ctypes.c_int
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return h


def _sim(cup, bpdx, bpdy, seed, uinf):
    g = cup.Simulation(bpdx, bpdy, extent=1.0, nu=1e-3)
    g.reset()
    rng = np.random.default_rng(seed)
    g.set_field(cup.VEL, rng.standard_normal((2, g.NY, g.NX)))
    g.set_scalars(time=0.0, step=0, uinf=uinf)
    return g


def test_find_max_u_noalloc_matches_context_entry_point(cup):
    hip = _hip()
    out = ctypes.c_void_p()
    stream = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out), 8) == 0
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # a 512x256 plane, then a smaller 64x32 one, on one stream
        for bpdx, bpdy, seed, uinf in [(16, 8, 7, (0.3, -0.2)), (2, 1, 8, (-1.5, 0.4))]:
            g = _sim(cup, bpdx, bpdy, seed, uinf)
            ref = g.find_max_u()
            cup.find_max_u_noalloc(g.device_ptr(cup.VEL, 0), g.device_ptr(cup.VEL, 1),
                                   g.NX * g.NY, uinf, out.value, stream.value)
            assert hip.hipStreamSynchronize(stream) == 0
            got = ctypes.c_double()
            assert hip.hipMemcpy(ctypes.byref(got), out, 8, 2) == 0  # device to host
            vel = g.get_field(cup.VEL)
            expect = max(np.abs(vel[0] + uinf[0]).max(), np.abs(vel[1] + uinf[1]).max(),
                         np.abs(vel).max())
            print(f"  {g.NX}x{g.NY}: noalloc {got.value!r} ctx {ref!r}", flush=True)
            assert got.value == ref == expect
            g.close()
        # a NaN cell is kept (the context's findMaxU reports it as an error)
        g = _sim(cup, 4, 2, 9, (0.0, 0.0))
        vel = g.get_field(cup.VEL)
        vel[1, 5, 77] = np.nan
        g.set_field(cup.VEL, vel)
        cup.find_max_u_noalloc(g.device_ptr(cup.VEL, 0), g.device_ptr(cup.VEL, 1),
                               g.NX * g.NY, (0.0, 0.0), out.value, stream.value)
        assert hip.hipStreamSynchronize(stream) == 0
        got = ctypes.c_double()
        assert hip.hipMemcpy(ctypes.byref(got), out, 8, 2) == 0
        assert np.isnan(got.value)
        # sizes beyond the maximum, and empty planes, are refused with a message
        for n in ((1 << 28) + 1, 0):
            with pytest.raises(cup.Cup2dError) as e:
                cup.find_max_u_noalloc(g.device_ptr(cup.VEL, 0), g.device_ptr(cup.VEL, 1),
                                       n, (0.0, 0.0), out.value, stream.value)
            assert e.value.code == cup.ERR_ARG and "noalloc" in str(e.value)
        g.close()
    finally:
        hip.hipStreamDestroy(stream)
        hip.hipFree(out)
