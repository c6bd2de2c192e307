"""The seed split across GPUs, rehearsed with processes on the CPU (gloo,
world size 2): each rank tracks its contiguous share of a chunk
(``seed_shares``, cu_propagate_seeds.py:45-54) with the CPU oracle standing
in for its GPU, the ranks exchange their streamline counts with one
all-gather, and the offsets from ``streamline_offsets`` place every rank's
streamlines.  With random streams restarting at each GPU's share (the
reference's chunk-relative streams, rng_chunk = the per-GPU share) the merged
output equals one process tracking the whole chunk."""

import os
import socket

import numpy as np


def _free_port():
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


def _rank(rank, world, port, out_dir):
    import torch
    import torch.distributed as dist

    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd.propagate_seeds import seed_shares, streamline_offsets
    from oracle.oracle import Oracle, set_num_threads

    set_num_threads(2)
    dist.init_process_group("gloo", init_method=f"tcp://127.0.0.1:{port}", rank=rank, world_size=world)
    try:
        w = W.build("csa", shape=(24, 26, 22))
        seeds = W.seeds_for(w, 90, placement="mixed", seed=3)
        shares = seed_shares(len(seeds), world)
        lo, n = shares[rank]
        per = shares[0][1]
        o = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot).track(
            seeds[lo: lo + n], rng_chunk=per)
        lens = o["sline_len"].astype(np.int64)
        counts = [torch.zeros(1, dtype=torch.int64) for _ in range(world)]
        dist.all_gather(counts, torch.tensor([len(lens)], dtype=torch.int64))
        offs = streamline_offsets([int(c.item()) for c in counts])
        pts = np.concatenate([o["slines"][i, :m] for i, m in enumerate(o["sline_len"])]) if len(lens) else \
            np.zeros((0, 3), np.float32)
        np.savez(os.path.join(out_dir, f"rank{rank}.npz"), lens=lens, pts=pts, seed=o["sline_seed"] + lo,
                 offs=offs)
    finally:
        dist.destroy_process_group()


def test_seed_shares_cover_the_chunk():
    from gpustreamlines_amd.propagate_seeds import seed_shares

    for n, g in [(0, 1), (1, 2), (7, 2), (8, 4), (25000, 3), (10, 4)]:
        sh = seed_shares(n, g)
        assert len(sh) == g
        assert sum(c for _, c in sh) == n
        nxt = 0
        for lo, c in sh:
            assert c >= 0
            if c:
                assert lo == nxt
                nxt = lo + c
        assert max(c for _, c in sh) == -(-n // g)


def test_two_rank_split_equals_one_process(tmp_path):
    import torch.multiprocessing as mp

    from gpustreamlines_amd import workloads as W
    from gpustreamlines_amd.propagate_seeds import seed_shares
    from oracle.oracle import Oracle

    world = 2
    mp.start_processes(_rank, args=(world, _free_port(), str(tmp_path)), nprocs=world, join=True,
                       start_method="spawn")
    w = W.build("csa", shape=(24, 26, 22))
    seeds = W.seeds_for(w, 90, placement="mixed", seed=3)
    per = seed_shares(len(seeds), world)[0][1]
    ref = Oracle(w.params, w.dataf, w.stop_map, w.vertices, w.edges, boot=w.boot).track(seeds, rng_chunk=per)
    parts = [np.load(tmp_path / f"rank{r}.npz") for r in range(world)]
    offs = parts[0]["offs"]
    assert np.array_equal(offs, parts[1]["offs"])
    assert offs[-1] == len(ref["sline_len"])
    merged_lens = np.concatenate([p["lens"] for p in parts])
    assert np.array_equal(merged_lens, ref["sline_len"])
    for r, p in enumerate(parts):
        assert np.array_equal(ref["sline_seed"][offs[r]:offs[r + 1]], p["seed"])
    ref_pts = np.concatenate([ref["slines"][i, :m] for i, m in enumerate(ref["sline_len"])])
    assert np.array_equal(np.concatenate([p["pts"] for p in parts]), ref_pts)
