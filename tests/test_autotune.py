"""The eigensolver autotuner's choice logic (imcom_amd/autotune.py), on the
CPU: size classes, table lookup and default, the argmin rule of tune() (run
against a stand-in context whose stage times are given), and the committed
table's consistency."""
import numpy as np

from imcom_amd import autotune as AT


def test_size_classes():
    assert AT.size_class(1) == 1024
    assert AT.size_class(1024) == 1024
    assert AT.size_class(1025) == 2048
    assert AT.size_class(16384) == 16384
    assert AT.size_class(32768) == 16384      # beyond the last class: the largest


def test_choice_and_default():
    table = {"2048": {"nb": 16, "nbb": 64}, "16384": {"nb": 64, "nbb": 256}}
    assert AT.choice(1500, table) == (16, 64)
    assert AT.choice(20000, table) == (64, 256)
    assert AT.choice(3000, table) == AT.DEFAULT   # class 4096 not tuned


class _FakeCtx:
    """Stands in for ImcomGPU: records set_tuning and reports a reduction
    plus back-transform time that depends on (nb, nbb) through a given cost."""

    def __init__(self, cost):
        self.cost = cost
        self.tuning = []
        self.cur = None
        self.n = 0

    def set_tile(self, t):
        self.n = t.n

    def set_tuning(self, nb, nbb):
        self.tuning.append((nb, nbb))
        self.cur = (nb, nbb)

    def set_A(self, A):
        assert A.shape == (self.n, self.n) and np.allclose(A, A.T)

    def set_profiling(self, level):
        pass

    def eigen_A(self, out=True):
        return None, None, 1.0

    def stage_times(self):
        ms = self.cost(self.n, *self.cur)
        return {"sytrd": (0.75 * ms, 1), "ormtr": (0.25 * ms, 1), "stedc": (123.0, 1)}


def test_tune_picks_the_fastest_and_restores_the_default():
    best = {256: (32, 128), 512: (64, 256)}

    def cost(n, nb, nbb):  # minimum at best[n]; the D&C time (stedc) must not matter
        b = best[n]
        return 10.0 + abs(nb - b[0]) + abs(nbb - b[1]) / 64.0

    g = _FakeCtx(cost)
    table = AT.tune(g, classes=(256, 512), repeats=2)
    for n, b in best.items():
        e = table[str(n)]
        assert (e["nb"], e["nbb"]) == b
        assert set(e["ms"]) == {f"{nb},{nbb}" for nb, nbb in AT.CANDIDATES}
        assert e["ms"][f"{b[0]},{b[1]}"] == min(e["ms"].values())
    assert g.tuning[-1] == AT.DEFAULT


def test_apply_sets_the_tables_choice():
    g = _FakeCtx(lambda n, nb, nbb: 1.0)
    assert AT.apply(g, 16000, {"16384": {"nb": 64, "nbb": 256}}) == (64, 256)
    assert g.tuning == [(64, 256)]


def test_committed_table_is_consistent():
    table = AT.load_table()
    assert table, "imcom_amd/data/tuning.json missing"
    for cls, e in table.items():
        assert int(cls) in AT.CLASSES
        assert (e["nb"], e["nbb"]) in AT.CANDIDATES
        ms = e["ms"]
        assert ms[f"{e['nb']},{e['nbb']}"] == min(ms.values()), cls
