"""Tractogram output on the CPU: chunking, the min/max length filter, the
packed StreamlineArray, and the TRX / StatefulTractogram writers, driven by a
stand-in seed propagator.  trx-python, nibabel and DIPY are not installed
here, so minimal stand-ins of the few calls the writers make are injected;
what is checked is the writers' own arithmetic (affine, offsets, growth)."""

import sys
import types

import numpy as np
import pytest

from gpustreamlines_amd.generic_tracker import GenericTracker


class FakePropagator:
    """Streamline j of seed i: (i % 3) + 1 directions, 5 + 97 * ((i + j) % 7)
    points along x from the seed."""

    def __init__(self, minlen=0, maxlen=np.inf):
        self.ngpus = 1
        self.minlen, self.maxlen = minlen, maxlen
        self.calls = []

    def propagate(self, seeds):
        self.calls.append(len(seeds))
        pts, lens = [], []
        for i, s in enumerate(np.asarray(seeds, np.float32)):
            for j in range(int(s[1]) % 3 + 1):
                n = 5 + 97 * ((i + j) % 7)
                p = np.repeat(s[None], n, 0)
                p[:, 0] += np.arange(n, dtype=np.float32) * 0.25
                pts.append(p)
                lens.append(n)
        lens = np.array(lens, np.int32)
        offs = np.zeros(len(lens) + 1, np.int64)
        np.cumsum(lens, out=offs[1:])
        self.slines = [np.concatenate(pts) if pts else np.zeros((0, 3), np.float32)]
        self.sline_lens = [lens]
        self.sline_offs = [offs]


class Tracker(GenericTracker):
    def __init__(self, chunk_size, **kw):
        self.chunk_size = chunk_size
        self.seed_propagator = FakePropagator(**kw)


def _seeds(n):
    i = np.arange(n, dtype=np.float32)
    return np.stack([i % 11, i % 5, i % 3], axis=1).astype(np.float32)


def _expected(seeds, chunk, minlen=0, maxlen=np.inf):
    fp = FakePropagator()
    out = []
    for k in range(0, len(seeds), chunk):
        fp.propagate(seeds[k:k + chunk])
        o, L = fp.sline_offs[0], fp.sline_lens[0]
        for j, n in enumerate(L):
            if minlen <= n <= maxlen:
                out.append(fp.slines[0][o[j]:o[j] + n])
    return out


@pytest.mark.parametrize("minlen,maxlen", [(0, np.inf), (100, 400)])
def test_generate_streamlines_chunks_and_filter(minlen, maxlen):
    seeds = _seeds(23)
    tr = Tracker(chunk_size=5, minlen=minlen, maxlen=maxlen)
    got = tr.generate_streamlines(seeds)
    want = _expected(seeds, 5, minlen, maxlen)
    assert tr.seed_propagator.calls == [5, 5, 5, 5, 3]
    assert len(got) == len(want)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)


class SinkPropagator(FakePropagator):
    """The same streamlines, its points written into the caller's
    PackedSink when one is set (as the device propagator downloads them)."""

    def __init__(self, **kw):
        super().__init__(**kw)
        self.sink = None

    def propagate(self, seeds):
        super().propagate(seeds)
        if self.sink is not None:
            dest = self.sink.take(len(self.slines[0]), len(seeds))
            dest[:] = self.slines[0]
            self.slines = [dest]


@pytest.mark.parametrize("minlen,maxlen", [(0, np.inf), (100, 400)])
def test_generate_streamlines_into_a_sink(minlen, maxlen, monkeypatch):
    """Points downloaded straight into one buffer give the same output; the
    buffer is sized once from the first chunk's points per seed (every later
    chunk fits in the 15 % headroom here) and never moved again."""
    from gpustreamlines_amd import array_sequence

    seeds = _seeds(40)
    tr = Tracker(chunk_size=8, minlen=minlen, maxlen=maxlen)
    tr.seed_propagator = SinkPropagator(minlen=minlen, maxlen=maxlen)
    sinks = []

    class Spy(array_sequence.PackedSink):
        def __init__(self, n):
            super().__init__(n)
            sinks.append(self)

    from gpustreamlines_amd import generic_tracker

    monkeypatch.setattr(generic_tracker, "PackedSink", Spy)
    got = tr.generate_streamlines(seeds)
    want = _expected(seeds, 8, minlen, maxlen)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    assert tr.seed_propagator.sink is None
    assert sinks[0].grows <= 2


def test_packed_sink_growth_order_and_misuse():
    from gpustreamlines_amd.array_sequence import PackedSink

    rng = np.random.default_rng(1)
    sink = PackedSink(total_seeds=100)
    parts = []
    for k in range(10):
        lens = rng.integers(1, 30, size=rng.integers(0, 12))
        pts = rng.normal(size=(int(lens.sum()), 3)).astype(np.float32)
        parts.append((pts, lens))
        if k % 2:
            sink.take(len(pts), 10)[:] = pts  # a producer that downloads into the sink
            sink.add(sink._open[0], lens, 10)
        else:
            sink.add(pts, lens, 10)  # a producer with its own array: copied in
    out = sink.result()
    np.testing.assert_array_equal(out._data, np.concatenate([p for p, _ in parts]))
    np.testing.assert_array_equal(out._lengths, np.concatenate([n for _, n in parts]))
    np.testing.assert_array_equal(out._offsets[1:], np.cumsum(out._lengths)[:-1])
    kept = sink.result(minlen=10, maxlen=20)
    assert ((kept._lengths >= 10) & (kept._lengths <= 20)).all()
    assert sum(len(s) for s in kept) == kept._lengths.sum() == len(kept._data)
    # a view taken and not added, or added out of order, is refused
    s2 = PackedSink()
    a, b = s2.take(3), s2.take(2)
    with pytest.raises(RuntimeError):
        s2.add(b, [2])
    with pytest.raises(RuntimeError):
        s2.result()
    s2.add(a, [3])
    s2.add(b, [1, 1])
    assert len(s2.result()) == 3
    assert len(PackedSink().result()) == 0


class FakeTrxFile:
    """The calls the TRX writer makes on trx.trx_file_memmap.TrxFile."""

    def __init__(self, reference=None, nb_streamlines=None, nb_vertices=None, init_as=None):
        self.streamlines = types.SimpleNamespace()
        if init_as is None:
            self.streamlines._data = np.zeros((0, 3), np.float32)
            self.streamlines._offsets = np.zeros(0, np.int64)
            self.streamlines._lengths = np.zeros(0, np.uint32)
            self.header = {"NB_STREAMLINES": 0, "NB_VERTICES": 0}
            self.resizes = []
        else:
            t = init_as.streamlines
            self.streamlines._data = np.zeros((nb_vertices, 3), t._data.dtype)
            self.streamlines._offsets = np.zeros(nb_streamlines, t._offsets.dtype)
            self.streamlines._lengths = np.zeros(nb_streamlines, t._lengths.dtype)
            self.header = {"NB_STREAMLINES": nb_streamlines, "NB_VERTICES": nb_vertices}
            self.resizes = []

    def resize(self, nb_streamlines=None, nb_vertices=None):
        s = self.streamlines
        for name, n in (("_data", nb_vertices), ("_offsets", nb_streamlines), ("_lengths", nb_streamlines)):
            old = getattr(s, name)
            new = np.zeros((n,) + old.shape[1:], old.dtype)
            k = min(n, len(old))
            new[:k] = old[:k]
            setattr(s, name, new)
        self.header = {"NB_STREAMLINES": nb_streamlines, "NB_VERTICES": nb_vertices}
        self.resizes.append((nb_streamlines, nb_vertices))


@pytest.fixture
def fake_trx(monkeypatch):
    trx = types.ModuleType("trx")
    mm = types.ModuleType("trx.trx_file_memmap")
    mm.TrxFile = FakeTrxFile
    trx.trx_file_memmap = mm
    monkeypatch.setitem(sys.modules, "trx", trx)
    monkeypatch.setitem(sys.modules, "trx.trx_file_memmap", mm)


@pytest.mark.parametrize("dtype", [np.float32, np.float16])
def test_generate_trx_world_coordinates_and_growth(fake_trx, dtype):
    seeds = _seeds(17)
    tr = Tracker(chunk_size=4)
    tr.set_compression_parameters(pos_dtype=dtype)
    ref = types.SimpleNamespace(affine=np.array([[2, 0, 0, -80], [0, 2, 0, -100], [0, 0, 2, -60], [0, 0, 0, 1.0]]))
    trx = tr.generate_trx(seeds, ref)
    want = _expected(seeds, 4)
    lens = np.array([len(w) for w in want])
    s = trx.streamlines
    assert len(s._lengths) == len(want) and len(s._data) == lens.sum()
    np.testing.assert_array_equal(s._lengths, lens)
    np.testing.assert_array_equal(s._offsets, np.concatenate([[0], np.cumsum(lens)[:-1]]))
    world = np.concatenate(want).astype(np.float64) * 2 + np.array([-80, -100, -60])
    np.testing.assert_array_equal(s._data, world.astype(dtype))
    assert s._data.dtype == dtype
    assert len(trx.resizes) >= 2  # grew past the 100-point guess, then trimmed


def test_generate_sft_uses_packed_arrays(monkeypatch):
    seeds = _seeds(9)
    tr = Tracker(chunk_size=4)
    dipy = types.ModuleType("dipy")
    io = types.ModuleType("dipy.io")
    st = types.ModuleType("dipy.io.stateful_tractogram")
    st.Space = types.SimpleNamespace(VOX="vox")
    st.StatefulTractogram = lambda seq, ref, space: types.SimpleNamespace(seq=seq, ref=ref, space=space)
    nib = types.ModuleType("nibabel")
    nst = types.ModuleType("nibabel.streamlines")
    asq = types.ModuleType("nibabel.streamlines.array_sequence")
    asq.ArraySequence = lambda: types.SimpleNamespace()
    for name, mod in (("dipy", dipy), ("dipy.io", io), ("dipy.io.stateful_tractogram", st), ("nibabel", nib),
                      ("nibabel.streamlines", nst), ("nibabel.streamlines.array_sequence", asq)):
        monkeypatch.setitem(sys.modules, name, mod)
    sft = tr.generate_sft(seeds, "ref")
    want = _expected(seeds, 4)
    assert sft.space == "vox" and sft.ref == "ref"
    np.testing.assert_array_equal(sft.seq._data, np.concatenate(want))
    np.testing.assert_array_equal(sft.seq._lengths, [len(w) for w in want])


class RecordingPropagator(FakePropagator):
    def __init__(self, **kw):
        super().__init__(**kw)
        self.rng_chunk = None
        self.rng_chunks = []

    def propagate(self, seeds):
        self.rng_chunks.append(self.rng_chunk)
        super().propagate(seeds)


def test_chunks_grouped_into_launches():
    """On one GPU, launch_seeds groups whole chunks into one propagate call
    and sets rng_chunk to the chunk size, so each chunk keeps its own random
    streams; launch_seeds 0 keeps one chunk per call."""
    seeds = _seeds(23)
    tr = Tracker(chunk_size=5)
    tr.seed_propagator = RecordingPropagator()
    tr.launch_seeds = 10
    tr.generate_streamlines(seeds)
    assert tr.seed_propagator.calls == [10, 10, 3]
    assert tr.seed_propagator.rng_chunks == [5, 5, 5]
    tr2 = Tracker(chunk_size=5)
    tr2.seed_propagator = RecordingPropagator()
    tr2.launch_seeds = 0
    tr2.generate_streamlines(seeds)
    assert tr2.seed_propagator.calls == [5, 5, 5, 5, 3]
    assert tr2.seed_propagator.rng_chunks == [None] * 5
