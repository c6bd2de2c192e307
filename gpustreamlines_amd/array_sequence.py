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


def make_array_sequence(iterable, buffer_size=0):
    try:
        from nibabel.streamlines.array_sequence import ArraySequence
    except ImportError:
        return StreamlineArray(iterable, buffer_size)
    return ArraySequence(iterable, buffer_size)
