"""Chunking and tractogram output (cuslines/generic_tracker.py).

``generate_sft`` and ``generate_trx`` keep the reference's signatures and
import DIPY / nibabel / trx-python lazily; ``generate_streamlines`` returns
the packed streamlines without any of them.  All three consume the packed
points and lengths each seed chunk's device compaction returns.
"""

from __future__ import annotations

import logging

import numpy as np

from .array_sequence import PackedSink, StreamlineArray

logger = logging.getLogger("GPUStreamlines")


class GenericTracker:
    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        return False

    def set_compression_parameters(self, pos_dtype=np.float32, linearize=False, tol_error=0.1,
                                   max_segment_length=10):
        """Compression of TRX output (generic_tracker.py:20-52)."""
        self.pos_dtype = pos_dtype
        self.linearize = linearize
        self.tol_error = tol_error
        self.max_segment_length = max_segment_length

    def _ngpus(self):
        return getattr(self, "ngpus", 1)

    def _pos_dtype(self):
        return getattr(self, "pos_dtype", np.float32)

    def _linearize(self):
        return getattr(self, "linearize", False)

    def _tol_error(self):
        return getattr(self, "tol_error", 0.1)

    def _max_segment_length(self):
        return getattr(self, "max_segment_length", 10)

    def _divide_chunks(self, seeds):
        global_chunk_sz = self.chunk_size * self._ngpus()
        nchunks = (seeds.shape[0] + global_chunk_sz - 1) // global_chunk_sz
        return global_chunk_sz, nchunks

    def _raw_chunks(self, seeds, sink=None):
        """Per launch and GPU: (points [n, 3] float32 voxel coordinates,
        lengths int64, seeds) as the device compaction returns them, before
        the min/max length filter.  With ``sink`` (a PackedSink) the points
        are downloaded into it and the caller adds each part to it before
        asking for the next."""
        seeds = np.asarray(seeds)
        gsz, nchunks = self._divide_chunks(seeds)
        sp = self.seed_propagator
        # On one GPU several chunks go into one launch: each chunk keeps its
        # own random streams (rng_chunk), so the output is the same as chunk
        # by chunk, and the persistent kernel's tail of long streamlines is
        # paid once per launch instead of once per chunk.
        per = 1
        if self._ngpus() == 1:
            per = max(1, int(getattr(self, "launch_seeds", 0) or 0) // gsz)
            depth = int(getattr(self, "chunk_pipeline", 1) or 1)
            if depth > 1 and nchunks > per:
                # consecutive launches overlapping on handles that share the
                # device data: one chunk per launch with `depth` in flight,
                # or `per` chunks per launch (chunk-relative streams) with
                # two in flight, so one launch's points are downloaded while
                # the next one tracks
                from .propagate_seeds import ChunkPipeline

                step = per * gsz
                pipe = ChunkPipeline(self.device_trackers[0], depth=depth if per == 1 else 2,
                                     graphs=bool(getattr(self, "chunk_graphs", True)),
                                     rng_chunk=gsz if per > 1 else None)
                try:
                    for c, (pts, lens) in enumerate(pipe.run(seeds, step, sink=sink)):
                        yield pts, lens, len(seeds[c * step:(c + 1) * step])
                finally:
                    pipe.close()
                return
        from .propagate_seeds import seed_shares

        sp.rng_chunk = gsz if per > 1 else None
        direct = sink is not None and hasattr(sp, "sink")
        if direct:
            sp.sink = sink
        try:
            for idx in range(0, nchunks, per):
                part = seeds[idx * gsz: (idx + per) * gsz]
                sp.propagate(part)
                shares = seed_shares(len(part), sp.ngpus)
                for ii in range(sp.ngpus):
                    lens = sp.sline_lens[ii]
                    if lens is not None:
                        yield sp.slines[ii], lens.astype(np.int64), shares[ii][1]
        finally:
            if direct:
                sp.sink = None

    def _packed_chunks(self, seeds):
        """Per seed chunk: (points [n, 3] float32 voxel coordinates, lengths
        int64) of the streamlines kept by the min/max length filter, packed as
        the device compaction returns them."""
        sp = self.seed_propagator
        for pts, lens, _ in self._raw_chunks(seeds):
            keep = (lens >= sp.minlen) & (lens <= sp.maxlen)
            if keep.all():
                if len(lens):
                    yield pts, lens
            elif keep.any():
                # points are packed in streamline order: one mask over all of them
                yield pts[np.repeat(keep, lens)], lens[keep]

    def generate_streamlines(self, seeds) -> StreamlineArray:
        """All kept streamlines of ``seeds`` (voxel coordinates), packed: the
        chunks download into one buffer (PackedSink), filtered once at the
        end."""
        sp = self.seed_propagator
        sink = PackedSink(len(seeds))
        for pts, lens, n in self._raw_chunks(seeds, sink):
            sink.add(pts, lens, n)
        return sink.result(sp.minlen, sp.maxlen)

    def generate_sft(self, seeds, ref_img):
        """StatefulTractogram in voxel space (generic_tracker.py:75-93), built
        from the packed points without a per-streamline copy."""
        from dipy.io.stateful_tractogram import Space, StatefulTractogram
        from nibabel.streamlines.array_sequence import ArraySequence

        packed = self.generate_streamlines(seeds)
        seq = ArraySequence()
        seq._data = packed._data
        seq._offsets = packed._offsets
        seq._lengths = packed._lengths
        return StatefulTractogram(seq, ref_img, Space.VOX)

    def generate_trx(self, seeds, ref_img):
        """TRX tractogram in world (RAS mm) coordinates (generic_tracker.py:95-172).

        Streams each chunk's packed points into a memory-mapped TrxFile: voxel
        to world with the image affine, optional linearisation
        (set_compression_parameters), cast to the position dtype, and the
        file's capacity grown to 1.5x what the chunk needs when it runs out
        (first guess: 2 streamlines of 100 points per seed)."""
        from trx.trx_file_memmap import TrxFile

        dtype = self._pos_dtype()
        affine = np.asarray(ref_img.affine, dtype=np.float64)
        template = TrxFile(reference=ref_img)
        template.streamlines._data = template.streamlines._data.astype(dtype)
        template.streamlines._offsets = template.streamlines._offsets.astype(np.uint64)
        cap_sl = 2 * int(np.asarray(seeds).shape[0])
        cap_pts = 100 * cap_sl
        trx = TrxFile(nb_streamlines=cap_sl, nb_vertices=cap_pts, init_as=template)
        n_sl = n_pts = 0
        for pts, lens in self._packed_chunks(seeds):
            world = pts.astype(np.float64) @ affine[:3, :3].T + affine[:3, 3]
            if self._linearize():
                world, lens = _compress(world, lens, self._tol_error(), self._max_segment_length())
            data = world.astype(dtype)
            need_sl, need_pts = n_sl + len(lens), n_pts + len(data)
            if need_sl > cap_sl or need_pts > cap_pts:
                cap_sl, cap_pts = max(cap_sl, need_sl * 3 // 2), max(cap_pts, need_pts * 3 // 2)
                logger.info("TRX capacity -> %d streamlines, %d points", cap_sl, cap_pts)
                trx.resize(nb_streamlines=cap_sl, nb_vertices=cap_pts)
            offsets = np.zeros(len(lens), dtype=np.uint64)
            np.cumsum(lens[:-1], out=offsets[1:])
            trx.streamlines._data[n_pts:need_pts] = data
            trx.streamlines._offsets[n_sl:need_sl] = offsets + np.uint64(n_pts)
            trx.streamlines._lengths[n_sl:need_sl] = lens
            n_sl, n_pts = need_sl, need_pts
        trx.resize(nb_streamlines=n_sl, nb_vertices=n_pts)
        return trx


def _compress(points, lens, tol_error, max_segment_length):
    """DIPY's streamline linearisation of packed streamlines; returns packed."""
    from dipy.tracking.streamlinespeed import compress_streamlines

    bounds = np.concatenate([[0], np.cumsum(lens)])
    comp = compress_streamlines([points[bounds[k]:bounds[k + 1]] for k in range(len(lens))], tol_error=tol_error,
                                max_segment_length=max_segment_length)
    new_lens = np.array([len(c) for c in comp], dtype=np.int64)
    return (np.concatenate(comp) if len(comp) else np.zeros((0, 3))), new_lens
