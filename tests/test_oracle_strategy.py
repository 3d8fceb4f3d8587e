"""Pins of the oracle's progressive load strategy.

* costs: the oracle uses exact integers; here they are compared with the
  reference's own double-precision evaluation (dataloader.cu:59-68, pow and
  ceil in float64);
* search: the oracle resolves ties with a closed-form key.  Here it is
  compared with an emulation, written from dataloader.cu:73-133, of what the
  reference's 33 blocks x 1024 threads actually do: each thread's first strict
  maximum over pos = tid + 1024k, the 1024-way halving tree per block and the
  64-slot halving tree over blocks.
"""
import math

import numpy as np
import pytest


def ref_cost_table():
    cost = np.zeros((4, 33), np.uint64)
    for l in range(4):
        p = 1.0 if l == 0 else 1.25
        for b in range(33):
            shift = 32 - b
            err = 0.0 if shift == 0 else float(1 << shift)
            err *= (math.pow(p, 12 - 3 * l - 1) + math.pow(p, 12 - 3 * l - 2) + math.pow(p, 12 - 3 * l - 3))
            cost[l, b] = math.ceil(err)
    return cost


def test_costs_match_double_evaluation(oracle):
    ref = ref_cost_table()
    for l in range(4):
        for b in range(33):
            assert oracle.strategy_cost(l, b) == int(ref[l, b]), (l, b)
    assert oracle.strategy_cost(0, 0) == 3 * 2**32
    assert oracle.strategy_cost(1, 16) == 953125
    assert oracle.strategy_cost(3, 31) == 8  # ceil(2 * 61/16)


def ref_search(cum, target_cost):
    """Emulation of findOptimalStrategy<int,4><<<33,1024>>> (dataloader.cu:30-177)."""
    cum = np.asarray(cum, np.uint64).reshape(4, 32).astype(np.int64)
    value = np.zeros((4, 33), np.int64)
    for i in range(4):
        value[i, 0] = cum[i, 31]
        value[i, 32] = 0
        for j in range(31, 0, -1):
            value[i, j] = cum[i, 31] - cum[i, j - 1]
    cost = ref_cost_table().astype(np.int64)
    pos = np.arange(33 ** 3)
    b1, b2, b3 = pos // 1089, (pos // 33) % 33, pos % 33
    v123 = value[1][b1] + value[2][b2] + value[3][b3]
    c123 = cost[1][b1] + cost[2][b2] + cost[3][b3]
    npad = 36 * 1024
    value_blk = [0] * 65
    maxbid = [0] * 65
    target_blk = [[32] * 4 for _ in range(65)]
    for b0 in range(33):
        v = v123 + value[0][b0]
        c = c123 + cost[0][b0]
        vv = np.zeros(npad, np.int64)
        vv[:pos.size] = np.where(c <= target_cost, v, 0)
        grid = vv.reshape(36, 1024)  # grid[k, tid] = value at pos = tid + 1024 k
        tmax = grid.max(axis=0)
        first = np.argmax(grid == tmax[None, :], axis=0)
        max_value = [int(x) for x in tmax]
        maxtid = list(range(1024))
        stride = 512
        while stride > 0:
            for t in range(stride):
                if max_value[t] < max_value[t + stride]:
                    max_value[t] = max_value[t + stride]
                    maxtid[t] = maxtid[t + stride]
            stride >>= 1
        wt = maxtid[0]
        value_blk[b0] = max_value[0]
        maxbid[b0] = b0
        if max_value[0] > 0:
            p = wt + 1024 * int(first[wt])
            target_blk[b0] = [b0, p // 1089, (p // 33) % 33, p % 33]
        else:
            target_blk[b0] = [32, 32, 32, 32]
    stride = 32
    while stride > 0:
        for t in range(stride):
            if value_blk[t] < value_blk[t + stride]:
                value_blk[t] = value_blk[t + stride]
                maxbid[t] = maxbid[t + stride]
        stride >>= 1
    return target_blk[maxbid[0]], int(value_blk[0])


def _cum_from_sizes(sizes):
    return np.cumsum(np.asarray(sizes, np.uint64).reshape(-1)).astype(np.uint64)


@pytest.mark.parametrize("seed", range(6))
def test_search_matches_reference_reductions(oracle, seed):
    rng = np.random.default_rng(seed)
    if seed % 3 == 0:
        sizes = rng.integers(2, 5000, 128)          # generic
    elif seed % 3 == 1:
        sizes = np.full(128, 2)                      # every plane one empty chunk: massive ties
    else:
        sizes = rng.choice([2, 4, 6], 128)           # many ties
    cum = _cum_from_sizes(sizes)
    for target in [0, 7, 30, 1000, 95399, 953124, 4 * 10**6, int(rng.integers(0, 2**24))]:
        ours, ov = oracle.strategy_search(cum, target)
        ref, rv = ref_search(cum, target)
        assert list(ours) == list(ref), (seed, target, ov, rv)
        assert ov == rv


def test_find_strategy_update_rules(oracle):
    rng = np.random.default_rng(4)
    cum = _cum_from_sizes(rng.integers(100, 4000, 128))
    eb = 1e-6
    b, e, ml, used = oracle.find_strategy(cum, [0] * 4, [0] * 4, eb, 1e-2)
    assert used and list(b) == [0] * 4
    ch, _ = oracle.strategy_search(cum, int(math.floor(1e-2 * 0.954 / eb)) - 1)
    assert list(e) == [max(0, c) for c in ch]
    assert ml == 3  # coarsest level advanced
    # a tighter target never lowers end; begin becomes the previous end
    b2, e2, ml2, _ = oracle.find_strategy(cum, b, e, eb, 1e-4)
    assert list(b2) == list(e)
    assert all(x >= y for x, y in zip(e2, e))
    # below the bound: load everything that is left
    b3, e3, ml3, used3 = oracle.find_strategy(cum, b2, e2, eb, 0.5e-6)
    assert not used3 and list(e3) == [32] * 4 and list(b3) == list(e2)
    expect = 3 if b3[0] < 32 else 2 if b3[1] < 32 else 1 if b3[2] < 32 else 0 if b3[3] < 32 else 4
    assert ml3 == expect
