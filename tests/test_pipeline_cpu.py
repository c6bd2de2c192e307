"""The chunk pipeline's host protocol on the CPU (propagate_seeds.ChunkPipeline):
chunks handed round-robin to forked handles and collected in order, a chunk
whose streamlines exceed its slots run again with more, buffers grown and
reused, and the points of every chunk downloaded into a PackedSink.  The
device side is a stand-in with the bounded-chunk semantics of
cuslines_chunk_async: with `cap` slots only the seeds whose slots all lie
below the capacity are tracked (a prefix), and the totals report the
streamline count, the tracked slots and their points."""

import numpy as np
import pytest

from gpustreamlines_amd.array_sequence import PackedSink
from gpustreamlines_amd.propagate_seeds import ChunkPipeline


def _slines_of(seed):
    """Streamlines of one seed: 1-4 of them, 2-9 points, values naming the
    seed, streamline and point."""
    k = int(seed[0]) % 4 + 1
    out = []
    for j in range(k):
        n = 2 + (int(seed[1]) + j) % 8
        p = np.zeros((n, 3), np.float32)
        p[:, 0] = seed[0]
        p[:, 1] = j
        p[:, 2] = np.arange(n)
        out.append(p)
    return out


class FakeBuf:
    def __init__(self):
        self.data = None

    def upload(self, a):
        self.data = np.array(a, copy=True)


class FakeChunk:
    made = []

    def __init__(self, dt, nmax, cap):
        self.dt, self.nmax, self.cap = dt, nmax, cap
        self.sd = FakeBuf()
        self.graph = None
        self.closed = False
        self.enqueued = []
        self.rng_chunks = []
        FakeChunk.made.append(self)

    def enqueue(self, n, rng_chunk, graph=False):
        assert not self.closed and n <= self.nmax
        seeds = self.sd.data[:n]
        self.enqueued.append((n, graph))
        self.rng_chunks.append(rng_chunk)
        if graph:
            self.graph = "g"
        sl = [s for seed in seeds for s in _slines_of(seed)]
        offs = np.concatenate([[0], np.cumsum([len(_slines_of(s)) for s in seeds])])
        tracked = int(offs[offs <= self.cap].max())
        self.res = (len(sl), tracked, sl[:tracked])

    def totals(self):
        nsl, tracked, sl = self.res
        return nsl, tracked, int(sum(len(s) for s in sl))

    def collect(self, n, with_seeds=True, points_out=None):
        nsl, tracked, sl = self.res
        pts = np.concatenate(sl) if sl else np.zeros((0, 3), np.float32)
        if points_out is not None:
            assert points_out.shape == pts.shape
            points_out[:] = pts
            pts = points_out
        return {"totals": self.totals(), "points": pts,
                "sline_len": np.array([len(s) for s in sl], np.int32)}

    def close(self):
        self.closed = True


class FakeTracker:
    def __init__(self):
        self.forks = []

    def fork(self):
        f = FakeTracker()
        self.forks.append(f)
        return f

    def close(self):
        pass


def _pipe(depth, **kw):
    FakeChunk.made = []
    p = ChunkPipeline(FakeTracker(), depth=depth, **kw)
    p._CB = FakeChunk
    return p


def _expected(seeds, chunk):
    pts, lens = [], []
    for c in range(0, len(seeds), chunk):
        for s in seeds[c:c + chunk]:
            for sl in _slines_of(s):
                pts.append(sl)
                lens.append(len(sl))
    return np.concatenate(pts), np.array(lens)


def _seeds(n):
    i = np.arange(n, dtype=np.float32)
    return np.stack([i, (i * 7) % 13, i % 3], 1).astype(np.float32)


@pytest.mark.parametrize("depth,slots", [(1, 8), (3, 8), (4, 0.3)])
def test_pipeline_order_and_reruns(depth, slots):
    seeds = _seeds(47)
    pipe = _pipe(depth)
    out = list(pipe.run(seeds, 10, slots_per_seed=slots))
    pts, lens = _expected(seeds, 10)
    assert len(out) == 5
    np.testing.assert_array_equal(np.concatenate([o[0] for o in out]), pts)
    np.testing.assert_array_equal(np.concatenate([o[1] for o in out]), lens)
    # too few slots: every chunk ran again, without its graph, on larger buffers
    assert (pipe.reruns > 0) == (slots < 1)
    if slots < 1:
        assert any(not g for cb in FakeChunk.made for _, g in cb.enqueued)
    pipe.close()
    assert all(cb.closed for cb in FakeChunk.made)


def test_pipeline_into_a_sink_and_length_filter():
    seeds = _seeds(33)
    pipe = _pipe(2, minlen=4, maxlen=7)
    sink = PackedSink(len(seeds))
    for c, (p, n) in enumerate(pipe.run(seeds, 8, sink=sink)):
        sink.add(p, n, min(8, len(seeds) - 8 * c))
    pipe.close()
    got = sink.result(minlen=4, maxlen=7)
    pts, lens = _expected(seeds, 8)
    keep = (lens >= 4) & (lens <= 7)
    np.testing.assert_array_equal(got._lengths, lens[keep])
    np.testing.assert_array_equal(got._data, pts[np.repeat(keep, lens)])
    # without a sink the pipeline applies its own filter
    pipe = _pipe(2, minlen=4, maxlen=7)
    out = list(pipe.run(seeds, 8))
    pipe.close()
    np.testing.assert_array_equal(np.concatenate([o[1] for o in out]), lens[keep])
    np.testing.assert_array_equal(np.concatenate([o[0] for o in out]), pts[np.repeat(keep, lens)])


def test_pipeline_passes_chunk_relative_streams():
    """Launches of several reference chunks keep the streams chunk-relative:
    every launch (and rerun) gets rng_chunk; without it a launch is one
    chunk (rng_chunk = its seed count)."""
    seeds = _seeds(25)
    pipe = _pipe(2, rng_chunk=4)
    list(pipe.run(seeds, 12, slots_per_seed=0.3))
    pipe.close()
    assert pipe.reruns > 0
    assert {r for cb in FakeChunk.made for r in cb.rng_chunks} == {4}
    pipe = _pipe(2)
    list(pipe.run(seeds, 12))
    pipe.close()
    assert sorted(r for cb in FakeChunk.made for r in cb.rng_chunks) == [1, 12, 12]
