"""imcom_gpu_bisect_kappa_learned_noalloc against imcom_gpu_bisect_kappa_learned,
through the C ABI: the same kernels, so K, U, the evaluation counts and the
bracket hits must be bitwise equal.  Two calls in a row on one GPU share the
fixed scratch array; the second has a smaller tile (fewer inputs and output
pixels) than the first, so stale scratch from the first call would show.
Sizes beyond the fixed scratch are rejected before anything is enqueued."""
import ctypes

import numpy as np
import pytest

from imcom_amd import kappa_net, synth
from imcom_amd.gpu import ImcomError, load_library

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers for the test through the HIP runtime the library is
    linked against (its libamdhip64, already loaded with the library)."""

    def __init__(self):
        load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                                  ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _reference_and_device_inputs(g, d, tile, params):
    g.set_tile(tile)
    g.lookup_A(tile.npoly, tile.npad)
    _, L, _ = g.eigen_A(out_Q=False, out_L=True)
    g.lookup_B(tile.npoly, tile.npad)
    P = g.build_P(out=True)
    K, U, ev, hits = g.bisect_kappa_learned()
    dev = dict(L=d.put(L), P=d.put(P.ravel(order="F")), N=d.put(tile.Ndiag), params=d.put(params))
    return (K, U, ev, hits), dev


def _noalloc(g, d, tile, dev):
    m = tile.m
    f64 = np.full(m, np.nan)
    i32 = np.full(m, -1, dtype=np.int32)
    K, US, ev, hits = d.put(f64), d.put(f64), d.put(i32), d.put(i32)
    g.bisect_kappa_learned_noalloc(
        tile.n, m, dev["L"], dev["P"], tile.n, dev["N"], tile.C_a, dev["params"], K, US, ev,
        hits, zsig=3.0, UnotS=tile.USB, US_max=tile.US_max, k_min=tile.kappa_min,
        k_max=tile.kappa_max, tol=tile.US_tol, maxnits=tile.maxNbis)
    g.synchronize()
    return d.get(K, f64), d.get(US, f64), d.get(ev, i32), d.get(hits, i32)


def test_noalloc_matches_learned_bisection_two_calls(gpu_ctx):
    params = kappa_net.load_weights()
    assert params is not None, "trained bracket weights missing"
    gpu_ctx.mlp_set_params(params)
    first = synth.small()                                      # n = 1024, m = 576
    second = synth.tiny(seed=11, nexp=2, npix_in=10, nout=9)   # n = 200, m = 81
    assert second.n < first.n and second.m < first.m
    d = _Dev()
    try:
        cases = [_reference_and_device_inputs(gpu_ctx, d, t, params) for t in (first, second)]
        for tile, (ref, dev) in zip((first, second), cases):
            print(f"[case] noalloc learned bisection n={tile.n} m={tile.m}", flush=True)
            got = _noalloc(gpu_ctx, d, tile, dev)
            for name, a, b in zip(("K", "U", "evals", "hits"), got, ref):
                assert np.array_equal(a, b), (name, tile.n, tile.m)
            assert np.all(got[1] < tile.US_max)
    finally:
        d.free()


def test_noalloc_rejects_sizes_beyond_scratch(gpu_ctx):
    print("[case] noalloc size limits", flush=True)
    for n, m in ((32769, 16), (64, 16385), (0, 4)):
        with pytest.raises(ImcomError):
            gpu_ctx.bisect_kappa_learned_noalloc(n, m, 1, 1, max(n, 1), 1, 1.0, 1, 1, 1)
