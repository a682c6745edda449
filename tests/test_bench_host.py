"""Host logic of bench.py: how a job's sources are cut into timed steps.

A step must be a uniform sample of whole 64-source tiles in the library's tile
order, and the steps together must hold every source exactly once; the credit
per step (sources of the slice) and the chain / non-chain split that prices
its algorithmic bytes follow from the slices.
"""
import types

import numpy as np

import bench


def fake_job(sizes, seed):
    rng = np.random.default_rng(seed)
    comps, orders = [], []
    for n in sizes:
        k = int(rng.integers(0, n // 3 + 1))
        comps.append(types.SimpleNamespace(n=n, chain_nodes=np.sort(rng.choice(n, k, replace=False))))
        orders.append(rng.permutation(n))
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return comps, offsets, orders


def test_slices_partition_sources_into_whole_tiles():
    sizes = [3, 64, 65, 200, 1000, 7]
    comps, offsets, orders = fake_job(sizes, 1)
    for n_slices in (1, 3, 4, 64):
        slices = bench.make_slices(comps, offsets, orders, n_slices)
        assert len(slices) == n_slices
        allsrc = np.concatenate([s[0] for s in slices])
        assert np.array_equal(np.sort(allsrc), np.arange(offsets[-1]))  # each source once
        # tiles in order: tile i of the job (components in turn, 64 sources of
        # the tuned order each) goes to slice i mod n_slices
        tiles = [(ci, orders[ci][t:t + 64]) for ci, c in enumerate(comps) for t in range(0, c.n, 64)]
        for s, (srcs, bfs, chn) in enumerate(slices):
            mine = tiles[s::n_slices]
            want = np.sort(np.concatenate([offsets[ci] + v for ci, v in mine])) if mine else np.zeros(0)
            assert np.array_equal(srcs, want)
            for ci, c in enumerate(comps):
                v = np.concatenate([t for cj, t in mine if cj == ci] or [np.zeros(0, np.int64)])
                x = int(np.isin(v, c.chain_nodes).sum())
                assert chn[ci] == x and bfs[ci] == len(v) - x
