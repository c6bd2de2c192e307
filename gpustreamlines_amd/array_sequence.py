"""Streamline containers.

The reference returns nibabel ``ArraySequence`` objects
(cuslines/cuda_python/cu_propagate_seeds.py:261-262).  When nibabel is
installed the same type is returned; otherwise ``StreamlineArray``, a packed
(points, offsets, lengths) sequence with the parts of that interface the
tracking API uses (``len``, indexing, iteration, ``_data``/``_offsets``/
``_lengths``).
"""

from __future__ import annotations

import numpy as np

MEGABYTE = 1024 * 1024


class StreamlineArray:
    def __init__(self, iterable=None, buffer_size=0):
        arrs = [np.asarray(a, dtype=np.float32).reshape(-1, 3) for a in (iterable or [])]
        self._lengths = np.array([len(a) for a in arrs], dtype=np.int64)
        self._offsets = np.zeros(len(arrs), dtype=np.int64)
        if len(arrs):
            self._offsets[1:] = np.cumsum(self._lengths)[:-1]
            self._data = np.concatenate(arrs, axis=0)
        else:
            self._data = np.zeros((0, 3), dtype=np.float32)

    @classmethod
    def from_packed(cls, data, lengths):
        s = cls()
        s._data = np.asarray(data, dtype=np.float32).reshape(-1, 3)
        s._lengths = np.asarray(lengths, dtype=np.int64)
        s._offsets = np.zeros(len(s._lengths), dtype=np.int64)
        if len(s._lengths):
            s._offsets[1:] = np.cumsum(s._lengths)[:-1]
        return s

    def __len__(self):
        return len(self._lengths)

    def __getitem__(self, i):
        o = int(self._offsets[i])
        return self._data[o: o + int(self._lengths[i])]

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]


class PackedSink:
    """One growing host buffer that the chunks of a tracking run download
    their packed points into, in chunk order, so the packed output needs no
    concatenation pass over all points.

    ``take(npts, nseeds)`` appends room for a chunk's points and returns it
    (a view the caller fills at once: a later ``take`` may move the buffer,
    copying what was filled); ``add(pts, lens, nseeds)`` records the chunk's
    lengths, and copies ``pts`` in when it is not the view ``take`` gave.
    With ``total_seeds`` known, the first growth sizes the buffer for the
    whole run from the points per seed seen so far (15 % headroom); pages
    past the points written are never touched, so the headroom costs
    address space, not memory."""

    HEADROOM = 1.15

    def __init__(self, total_seeds: int = 0):
        self.total_seeds = int(total_seeds)
        self._buf = np.empty((0, 3), dtype=np.float32)
        self._n = 0
        self._seeds = 0
        self._open = []
        self._lens = []
        self.grows = 0

    def _grow(self, need):
        cap = max(need, len(self._buf) + len(self._buf) // 2)
        if 0 < self._seeds < self.total_seeds:
            cap = max(need, int(need / self._seeds * self.total_seeds * self.HEADROOM) + 1)
        new = np.empty((cap, 3), dtype=np.float32)
        new[: self._n] = self._buf[: self._n]
        self._buf = new
        self.grows += 1

    def take(self, npts: int, nseeds: int = 0) -> np.ndarray:
        npts = int(npts)
        self._seeds += int(nseeds)
        need = self._n + npts
        if need > len(self._buf):
            self._grow(need)
        view = self._buf[self._n: need]
        self._n = need
        self._open.append(view)
        return view

    def add(self, pts, lens, nseeds: int = 0):
        lens = np.asarray(lens, dtype=np.int64)
        if self._open:
            if pts is not self._open[0]:
                raise RuntimeError("add the points of the chunks in the order they were taken")
            self._open.pop(0)
        else:
            pts = np.asarray(pts, dtype=np.float32).reshape(-1, 3)
            self.take(len(pts), nseeds)[:] = pts
            self._open.pop()
        self._lens.append(lens)

    def result(self, minlen=0, maxlen=np.inf) -> StreamlineArray:
        """The packed streamlines whose point count lies in [minlen, maxlen]."""
        if self._open:
            raise RuntimeError("points taken but never added")
        lens = np.concatenate(self._lens) if self._lens else np.zeros(0, dtype=np.int64)
        if int(lens.sum()) != self._n:
            raise RuntimeError("lengths do not cover the points")
        data = self._buf[: self._n]
        keep = (lens >= minlen) & (lens <= maxlen)
        if not keep.all():
            data, lens = data[np.repeat(keep, lens)], lens[keep]
        return StreamlineArray.from_packed(data, lens)


def make_array_sequence(iterable, buffer_size=0):
    try:
        from nibabel.streamlines.array_sequence import ArraySequence
    except ImportError:
        return StreamlineArray(iterable, buffer_size)
    return ArraySequence(iterable, buffer_size)
