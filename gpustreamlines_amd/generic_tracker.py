"""Chunking and tractogram output (cuslines/generic_tracker.py).

``generate_sft`` and ``generate_trx`` keep the reference's signatures and
import DIPY / nibabel / trx-python lazily; ``generate_streamlines`` returns
the packed streamlines without any of them.
"""

from __future__ import annotations

import logging

import numpy as np

from .array_sequence import StreamlineArray, make_array_sequence

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

    def generate_streamlines(self, seeds) -> StreamlineArray:
        """All kept streamlines of ``seeds`` (voxel coordinates), packed."""
        seeds = np.asarray(seeds)
        gsz, nchunks = self._divide_chunks(seeds)
        data, lens = [], []
        for idx in range(nchunks):
            sp = self.seed_propagator
            sp.propagate(seeds[idx * gsz: (idx + 1) * gsz])
            for ii in range(sp.ngpus):
                L = sp.sline_lens[ii]
                keep = (L >= sp.minlen) & (L <= sp.maxlen)
                if keep.all():
                    data.append(sp.slines[ii])
                    lens.append(L)
                else:
                    for jj in np.nonzero(keep)[0]:
                        o = sp.sline_offs[ii][jj]
                        data.append(sp.slines[ii][o: o + L[jj]])
                    lens.append(L[keep])
        if not data:
            return StreamlineArray()
        return StreamlineArray.from_packed(np.concatenate(data), np.concatenate(lens))

    def generate_sft(self, seeds, ref_img):
        """StatefulTractogram in voxel space (generic_tracker.py:75-93)."""
        from dipy.io.stateful_tractogram import Space, StatefulTractogram

        global_chunk_sz, nchunks = self._divide_chunks(seeds)
        buffer_size = 0
        generators = []
        for idx in range(nchunks):
            self.seed_propagator.propagate(seeds[idx * global_chunk_sz: (idx + 1) * global_chunk_sz])
            buffer_size += self.seed_propagator.get_buffer_size()
            generators.append(list(self.seed_propagator.as_generator()))
        seq = make_array_sequence((s for g in generators for s in g), buffer_size)
        return StatefulTractogram(seq, ref_img, Space.VOX)

    def generate_trx(self, seeds, ref_img):
        """Memory-mapped TRX file, grown by doubling (generic_tracker.py:95-172)."""
        from nibabel.streamlines.array_sequence import ArraySequence
        from nibabel.streamlines.tractogram import Tractogram
        from trx.trx_file_memmap import TrxFile

        global_chunk_sz, nchunks = self._divide_chunks(seeds)
        sl_len_guess = 100
        sl_per_seed_guess = 2
        n_sls_guess = sl_per_seed_guess * seeds.shape[0]
        trx_reference = TrxFile(reference=ref_img)
        trx_reference.streamlines._data = trx_reference.streamlines._data.astype(self._pos_dtype())
        trx_reference.streamlines._offsets = trx_reference.streamlines._offsets.astype(np.uint64)
        trx_file = TrxFile(nb_streamlines=n_sls_guess, nb_vertices=n_sls_guess * sl_len_guess,
                           init_as=trx_reference)
        offsets_idx = 0
        sls_data_idx = 0
        for idx in range(int(nchunks)):
            self.seed_propagator.propagate(seeds[idx * global_chunk_sz: (idx + 1) * global_chunk_sz])
            tractogram = Tractogram(self.seed_propagator.as_array_sequence(), affine_to_rasmm=ref_img.affine)
            if len(tractogram) == 0:
                continue
            tractogram.to_world()
            sls = tractogram.streamlines
            if self._linearize():
                from dipy.tracking.streamlinespeed import compress_streamlines

                sls = ArraySequence(compress_streamlines(sls, tol_error=self._tol_error(),
                                                         max_segment_length=self._max_segment_length()))
            sls._data = sls._data.astype(self._pos_dtype())
            new_offsets_idx = offsets_idx + len(sls._offsets)
            new_sls_data_idx = sls_data_idx + len(sls._data)
            if (new_offsets_idx > trx_file.header["NB_STREAMLINES"]
                    or new_sls_data_idx > trx_file.header["NB_VERTICES"]):
                logger.info("TRX resizing...")
                trx_file.resize(nb_streamlines=new_offsets_idx * 2, nb_vertices=new_sls_data_idx * 2)
            trx_file.streamlines._data[sls_data_idx:new_sls_data_idx] = sls._data
            trx_file.streamlines._offsets[offsets_idx:new_offsets_idx] = sls_data_idx + sls._offsets
            trx_file.streamlines._lengths[offsets_idx:new_offsets_idx] = sls._lengths
            offsets_idx = new_offsets_idx
            sls_data_idx = new_sls_data_idx
        trx_file.resize()
        return trx_file
