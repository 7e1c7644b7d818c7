"""The benchmark inputs' generator (CPU): SplitMix64 against its published
outputs, tile values keyed by (seed, workload, tensor, tile) and element
ordinal (so they do not depend on where a tile lies in the pool or on the
other tiles), and the value ranges and moments of each kind."""
import numpy as np
import torch

from tiledarray_amd import workloads as W


def test_mix64_matches_splitmix64_reference_outputs():
    # SplitMix64 seeded with 0: the first three outputs of the reference
    # generator (Vigna's splitmix64.c)
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    state = torch.arange(1, 4, dtype=torch.int64) * W._s64(W._GAMMA)
    got = [int(v) & W._M64 for v in W.mix64(state)]
    assert got == expect
    assert [W._mix_int(i * W._GAMMA) for i in (1, 2, 3)] == expect


def test_tile_values_do_not_depend_on_placement():
    # the same three tiles packed in two different orders (and one absent
    # tile in between) hold the same values
    vols = np.array([7, 0, 13, 5], np.int64)
    offs1 = np.array([0, -1, 7, 20], np.int64)
    offs2 = np.array([18, -1, 0, 13], np.int64)
    p1 = torch.zeros(25, dtype=torch.float64)
    p2 = torch.zeros(25, dtype=torch.float64)
    W.keyed_fill(p1, offs1, vols, W.CFG_BLOCKSPARSE, 0, batch=8)
    W.keyed_fill(p2, offs2, vols, W.CFG_BLOCKSPARSE, 0)
    for t in (0, 2, 3):
        a = p1[offs1[t]:offs1[t] + vols[t]]
        b = p2[offs2[t]:offs2[t] + vols[t]]
        assert torch.equal(a, b)
    # another tensor or workload key gives other values
    p3 = torch.zeros(25, dtype=torch.float64)
    W.keyed_fill(p3, offs1, vols, W.CFG_BLOCKSPARSE, 1)
    assert not torch.equal(p1, p3)


def test_element_draw_is_the_keyed_counter():
    # element e of tile t: mix64(key(t) + (e + 1) * gamma), top 53 bits to
    # [-1, 1)
    pool = torch.zeros(6, dtype=torch.float64)
    W.keyed_fill(pool, np.array([0]), np.array([6]), W.CFG_DENSE, 1, seed=12345)
    key = W.tile_key(12345, W.CFG_DENSE, 1, 0)
    for e in range(6):
        z = W._mix_int(key + (e + 1) * W._GAMMA)
        assert pool[e].item() == (z >> 11) * 2.0 ** -52 - 1.0


def test_kinds_ranges_and_moments():
    n = 200000
    offs, vols = np.array([0, 100000]), np.array([100000, 100000])
    u = torch.zeros(n, dtype=torch.float64)
    W.keyed_fill(u, offs, vols, W.CFG_TOT, 0)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean().item()) < 0.01 and abs(u.var().item() - 1 / 3) < 0.01
    g = torch.zeros(n, dtype=torch.float64)
    W.keyed_fill(g, offs, vols, W.CFG_LADDER, 1, kind="normal")
    assert torch.isfinite(g).all()
    assert abs(g.mean().item()) < 0.01 and abs(g.var().item() - 1.0) < 0.02
    i = torch.zeros(n, dtype=torch.float32)
    W.keyed_fill(i, offs, vols, W.CFG_DENSE, 0, kind="int")
    vals = torch.unique(i)
    assert vals.min().item() == -8 and vals.max().item() == 8 and vals.numel() == 17
    assert torch.equal(i, torch.round(i))
