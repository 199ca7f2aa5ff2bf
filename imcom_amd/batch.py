"""Many tiles as several concurrent batched solves.

One batched solve (imcom_gpu_batch_solve) advances all its tiles column by
column in lockstep; its last columns, where only the largest tiles are still
being reduced, are short latency-bound launches.  Several batched solves on
separate contexts (each with its own HIP stream), issued from separate host
threads, fill those gaps with each other's work.  The tiles, sorted by
size, are cut into runs of about equal eigendecomposition work (sum of n^3):
the first solve takes the few largest tiles, the last the many small ones.

Measured on the 64-tile workload (n from 1024 to 5766): one solve 962 ms;
tiles dealt round-robin by size over two solves 905 ms, three 872 ms, four
986 ms; three runs of equal n^3 work (4, 9 and 51 tiles) 829 ms, against
893 ms for three runs of equal tile count; four and six runs of equal work
1004 and 1015 ms, also with eight hardware queues per process
(GPU_MAX_HW_QUEUES=8), so three is the default.  A context opened before
these and still open beside them (it was idle) made the three solves take
1055 ms, probably by sharing a hardware queue with one of them; contexts
already closed did no harm.  Open a StreamedBatch before any context that
stays open beside it.
"""
from __future__ import annotations

import threading

import numpy as np

from . import autotune
from .gpu import ImcomGPU

DEFAULT_STREAMS = 3


def split_tiles(ns, nstreams: int = DEFAULT_STREAMS) -> list[list[int]]:
    """Tile indices, largest first, cut into at most `nstreams` runs of about
    equal sum of n^3 (empty runs dropped)."""
    if nstreams < 1:
        raise ValueError("nstreams must be >= 1")
    if len(ns) == 0:
        raise ValueError("no tiles")
    order = np.argsort(np.asarray(ns), kind="stable")[::-1]
    work = np.cumsum(np.asarray(ns, dtype=np.float64)[order] ** 3)
    cuts = [int(np.searchsorted(work, work[-1] * j / nstreams)) + 1 for j in range(1, nstreams)]
    groups = [[int(i) for i in part] for part in np.split(order, cuts)]
    return [g for g in groups if g]


class StreamedBatch:
    """Tiles split over `nstreams` batched solves, solved concurrently."""

    def __init__(self, tiles, nstreams: int = DEFAULT_STREAMS, device: int = 0, weights=None,
                 tune: bool = True):
        self.groups = split_tiles([t.n for t in tiles], nstreams)
        self.where = {}  # tile index -> (solve, index within it)
        self.ctxs = []
        try:
            for s, gidx in enumerate(self.groups):
                g = ImcomGPU(device)
                self.ctxs.append(g)
                if tune:
                    autotune.apply(g, max(tiles[i].n for i in gidx))
                if weights is not None:
                    g.mlp_set_params(weights)
                g.batch_set([tiles[i] for i in gidx])
                for j, i in enumerate(gidx):
                    self.where[i] = (s, j)
        except Exception:
            self.close()
            raise
        self.ntiles = len(tiles)

    def solve(self, **kw):
        """imcom_gpu_batch_solve on every context at once (keyword arguments
        as ImcomGPU.batch_solve); returns when all have finished.  The
        contexts' workspaces are allocated first, on this thread, so that no
        device allocation runs while another thread captures a solve graph."""
        from .gpu import IMCOM_PREC_FP64
        for g in self.ctxs:
            g.batch_reserve(kw.get("precision_P", IMCOM_PREC_FP64),
                            kw.get("precision_T", IMCOM_PREC_FP64))
        errors = []

        def work(g):
            try:
                g.batch_solve(**kw)
            except Exception as e:  # re-raised on the calling thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(g,)) for g in self.ctxs]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]

    def result(self, i, want_T=False):
        s, j = self.where[i]
        return self.ctxs[s].batch_result(j, want_T=want_T)

    def close(self):
        for g in self.ctxs:
            g.close()
        self.ctxs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
