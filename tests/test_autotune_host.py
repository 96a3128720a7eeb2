"""CPU checks of the autotuner's selection logic (mdacp_amd.autotune) against a
mock domain whose kernel times are a known function of the options."""
from mdacp_amd import autotune as A
from mdacp_amd import mdgpu


class MockDomain:
    """Times: rebuild fastest at search (8, 0); force fastest at waves 4, 72 KiB, unroll 16, brick 3."""

    def __init__(self, ensemble=False):
        self.opts = {mdgpu.OPT_SEARCH_WAVES: 16, mdgpu.OPT_SEARCH_STAGE: 1, mdgpu.OPT_WAVES: 8,
                     mdgpu.OPT_LDS_KB: 96, mdgpu.OPT_UNROLL: 8, mdgpu.OPT_BRICK: 4 | 4 << 8 | 4 << 16,
                     mdgpu.OPT_PRUNE: 0, mdgpu.OPT_ROW_SORT: 8, mdgpu.OPT_DISPLACEMENT_CHECK: 1}
        self.clones = []
        self.closed = False
        if ensemble:
            self.R = 3

    def clone(self, **kw):
        c = MockDomain(hasattr(self, "R"))
        c.opts = dict(self.opts)
        self.clones.append(c)
        return c

    def set_option(self, o, v):
        self.opts[o] = v

    def get_option(self, o):
        return self.opts[o]

    def calculate(self, n):
        return 0

    def make_pair_list(self):
        pass

    def kinetic_energy(self):
        return 0.0

    def observe_replicas_raw(self):
        return None

    def enable_timing(self, on):
        pass

    def timing(self):
        o = self.opts
        b = o[mdgpu.OPT_BRICK] & 0xff
        t = 1.0 + 0.1 * abs(b - 3) + 0.05 * (o[mdgpu.OPT_WAVES] != 4) + 0.01 * abs(o[mdgpu.OPT_LDS_KB] - 72) / 8 \
            + 0.02 * (o[mdgpu.OPT_UNROLL] != 16) + 0.03 * (o[mdgpu.OPT_ROW_SORT] != 1)
        return dict(force_ms=t * 10, force_launches=10, total_ms=0.0, rebuilds=0, rebuild_ms=0.0)

    def close(self):
        self.closed = True


def test_picks_the_fastest_shapes(monkeypatch):
    times = {}

    def fake_rebuild(d, reps=3):
        w, s = d.get_option(mdgpu.OPT_SEARCH_WAVES), d.get_option(mdgpu.OPT_SEARCH_STAGE)
        return times.setdefault((w, s), 5.0 + abs(w - 8) * 0.1 + s * 0.3)

    monkeypatch.setattr(A, "_rebuild_ms", fake_rebuild)
    # the cadence without displacement checks is 5% faster per step on this mock
    monkeypatch.setattr(A, "_step_ms", lambda d, steps: 1.0 - 0.05 * (d.get_option(mdgpu.OPT_DISPLACEMENT_CHECK) == 0))
    d = MockDomain()
    res = A.autotune(d, steps=2)
    ch = res["choice"]
    assert ch["row_sort"] == 1 and d.opts[mdgpu.OPT_ROW_SORT] == 1
    assert ch["displacement_check"] == 0 and d.opts[mdgpu.OPT_DISPLACEMENT_CHECK] == 0
    assert set(res["layout"]) == {str(v) for v in A.LAYOUT} and set(res["cadence"]) == {str(v) for v in A.CADENCE}
    assert (ch["search_waves"], ch["search_stage"]) == (8, 0)
    assert (ch["waves"], ch["lds_kb"], ch["unroll"]) == (4, 72, 16)
    assert ch["brick"] == (3, 3, 3)
    # applied to the caller, and the clone was released
    assert d.opts[mdgpu.OPT_SEARCH_WAVES] == 8 and d.opts[mdgpu.OPT_SEARCH_STAGE] == 0
    assert d.opts[mdgpu.OPT_WAVES] == 4 and d.opts[mdgpu.OPT_LDS_KB] == 72 and d.opts[mdgpu.OPT_UNROLL] == 16
    assert d.opts[mdgpu.OPT_BRICK] == 3 | 3 << 8 | 3 << 16
    assert all(c.closed for c in d.clones)
    assert set(res["search"]) == {f"waves={w} stage={s}" for w, s in A.SEARCH}
    assert set(res["force"]) == {f"waves={w} lds_kb={k} unroll={u}" for w, k, u in A.FORCE}


def test_ensembles_keep_their_per_replica_bricks(monkeypatch):
    monkeypatch.setattr(A, "_rebuild_ms", lambda d, reps=3: 1.0)
    monkeypatch.setattr(A, "_step_ms", lambda d, steps: 1.0)
    d = MockDomain(ensemble=True)
    brick0 = d.opts[mdgpu.OPT_BRICK]
    res = A.autotune(d, steps=2)
    assert "brick" not in res["choice"] and d.opts[mdgpu.OPT_BRICK] == brick0


class LuckyDomain(MockDomain):
    """Force times: the current shape (8 wavefronts, 78 KiB, 16 ahead, not on the grid) takes
    1.0; (4, 72, 32) takes 0.9 the first time it is timed over the grid's steps and 1.05 after
    that (a lucky timing); every other shape takes 1.1."""

    def __init__(self, ensemble=False):
        super().__init__(ensemble)
        self.opts[mdgpu.OPT_LDS_KB] = 78
        self.opts[mdgpu.OPT_UNROLL] = 16
        self.seen = {}
        self.steps = 0

    def clone(self, **kw):
        c = LuckyDomain(hasattr(self, "R"))
        c.opts = dict(self.opts)
        self.clones.append(c)
        return c

    def calculate(self, n):
        self.steps = n
        return 0

    def timing(self):
        o = self.opts
        f = (o[mdgpu.OPT_WAVES], o[mdgpu.OPT_LDS_KB], o[mdgpu.OPT_UNROLL])
        if f == (8, 78, 16):
            t = 1.0
        elif f == (4, 72, 32):
            k = self.seen.get(f, 0) + (self.steps > 3)
            self.seen[f] = k
            t = 0.9 if (self.steps > 3 and k == 1) else 1.05
        else:
            t = 1.1
        return dict(force_ms=t * 10, force_launches=10, total_ms=0.0, rebuilds=0, rebuild_ms=0.0)


def test_a_lucky_timing_does_not_displace_the_current_shape(monkeypatch):
    monkeypatch.setattr(A, "_rebuild_ms", lambda d, reps=3: 5.0)
    monkeypatch.setattr(A, "_step_ms", lambda d, steps: 1.0)
    d = LuckyDomain()
    res = A.autotune(d, steps=30)
    ch = res["choice"]
    assert "waves=8 lds_kb=78 unroll=16" in res["force"]  # the current shape is timed on the grid
    assert res["force"]["waves=4 lds_kb=72 unroll=32"] == 0.9  # and loses to the lucky one there
    assert (ch["waves"], ch["lds_kb"], ch["unroll"]) == (8, 78, 16)  # but not when timed again
    assert d.opts[mdgpu.OPT_LDS_KB] == 78 and d.opts[mdgpu.OPT_UNROLL] == 16
