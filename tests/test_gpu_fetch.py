"""Asynchronous stage-output copies (imcom_gpu_fetch_async / fetch_wait):
each fetched array equals the stage's synchronous output, also when the next
stage overwrites the device buffer right after the fetch is issued (that
stage waits for the copy on the device)."""
import numpy as np
import pytest

from imcom_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    from imcom_amd.gpu import ImcomGPU
    with ImcomGPU(0) as ctx:
        yield ctx


def test_fetch_matches_synchronous_outputs(g):
    from imcom_amd.gpu import ImcomError, IMCOM_PREC_FP64
    print("[fetch] small tile: A, Q, L, B, P, T", flush=True)
    t = synth.small()
    g.set_tile(t)
    A_sync = g.lookup_A(t.npoly, t.npad, out=True)
    g.lookup_A(t.npoly, t.npad)
    fA = g.fetch_async("A")
    Q_sync, L_sync, _ = g.eigen_A(out=True)   # overwrites A on the device
    g.fetch_wait("A")
    assert np.array_equal(fA, A_sync)
    with pytest.raises(ImcomError, match="A not built"):
        g.fetch_async("A")                    # the reduction has overwritten A
    fQ, fL = g.fetch_async("Q"), g.fetch_async("L")
    B_sync = g.lookup_B(t.npoly, t.npad, out=True)
    fB = g.fetch_async("B")
    P_sync = g.build_P(out=True)
    fP = g.fetch_async("P")
    g.build_P()                               # rewrites P after the copy
    g.bisect_kappa()
    T_sync = g.build_T(IMCOM_PREC_FP64, out=True)
    fT = g.fetch_async("T")
    g.build_T(IMCOM_PREC_FP64, out=False)
    g.fetch_wait()
    for name, a, b in (("Q", fQ, Q_sync), ("L", fL, L_sync), ("B", fB, B_sync),
                       ("P", fP, P_sync), ("T", fT, T_sync)):
        assert a.shape == b.shape and np.array_equal(a, b), name
    assert fQ.flags.f_contiguous and fT.flags.f_contiguous


def test_fetch_into_caller_buffer_and_set_tile_waits(g):
    print("[fetch] caller buffer, set_tile with a copy in flight", flush=True)
    t = synth.small()
    g.set_tile(t)
    A_sync = g.lookup_A(t.npoly, t.npad, out=True)
    out = np.empty((t.n, t.n), order="F")
    assert g.fetch_async("A", out=out) is out
    g.set_tile(t)                             # waits for the copy before any reallocation
    assert np.array_equal(out, A_sync)
    with pytest.raises(ValueError):
        g.fetch_async("A", out=np.empty((t.n, t.n)))  # C order
