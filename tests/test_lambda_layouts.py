"""Lambda layouts of the generated start, on the CPU.

Between the passes of a generated-start evaluation lambda sits in the order
that makes the next high pass's tile one contiguous 2^12 block.  These tests
restate the layout (``lay_addr`` / ``lay_fits`` of
``gpu_simulation_amd/csrc/rie_kernels.hpp``) and the engine's pass plan and
check what the kernels rely on: each layout is a permutation of the index
bits below n; under its own layout a high pass's tile is an aligned block
whose offsets appear in order; the register bits of every pass's load mapping
stay contiguous in the layouts it reads and writes, wherever the engine turns
the layouts on (30 qubits among them); and the chain of layouts hands every
amplitude from the pass that writes it to the pass that reads it.  The
kernels themselves are compared with and without layouts on the GPU.
"""

from __future__ import annotations

import numpy as np
import pytest

T = 12


def lay_addr(i, h):
    if h == 0:
        return i
    low = i & 7
    blk = (i >> h) & 511
    mid = (i >> 3) & ((1 << (h - 3)) - 1)
    top = (i >> (h + 9)) << (h + 9)
    return low | (blk << 3) | (mid << 12) | top


def lay_fits(h, n, first, count):
    if h == 0:
        return True
    if h < 3 or h + 9 > n:
        return False
    p0 = lay_addr(1 << first, h)
    return all(lay_addr(1 << (first + b), h) == p0 << b for b in range(1, count))


def high_groups(n):
    """h0 of every high pass as the engine plans them (a = 3, groups of at
    most 9 wires above the 12-bit low tile, h0 = min(g0, n - 9))."""
    rest, width = n - T, T - 3
    groups = (rest + width - 1) // width
    g0, out = T, []
    for gi in range(groups):
        bg = rest // groups + (1 if gi < rest % groups else 0)
        out.append(min(g0, n - width))
        g0 += bg
    return out


def layouts_fit(n, r_low=4, r_high=4):
    """Engine::lam_layouts_fit: the low pass writes the layout of the last
    high pass, high pass j reads its own and writes high pass j-1's."""
    h0s = high_groups(n)
    K = len(h0s)
    if not lay_fits(h0s[K - 1], n, T - r_low, r_low):
        return False
    for j in range(K):
        first = h0s[j] + (T - r_high - 3)
        out = h0s[j - 1] if j > 0 else 0
        if not (lay_fits(h0s[j], n, first, r_high) and lay_fits(out, n, first, r_high)):
            return False
    return True


def tile_base(g, a, h0):
    gap = h0 - a
    return ((g & ((1 << gap) - 1)) << a) | ((g >> gap) << (h0 + T - a))


def tile_index(base, o, a, h0):
    return base | (o & ((1 << a) - 1)) | ((o >> a) << h0)


@pytest.mark.parametrize("n,h", [(13, 4), (16, 7), (21, 12), (30, 12), (30, 21), (31, 22)])
def test_layout_is_a_bit_permutation(n, h):
    # every index bit below n lands on a distinct address bit below n
    images = [lay_addr(1 << b, h) for b in range(n)]
    assert all(v and v & (v - 1) == 0 for v in images)
    assert sorted(images) == [1 << b for b in range(n)]
    # and the map is linear over disjoint bits
    rng = np.random.default_rng(n + h)
    for x in rng.integers(0, 1 << n, 50):
        x = int(x)
        want = 0
        for b in range(n):
            if (x >> b) & 1:
                want |= images[b]
        assert lay_addr(x, h) == want


@pytest.mark.parametrize("n", [13, 16, 21, 25, 30])
def test_own_layout_makes_each_high_tile_one_block(n):
    rng = np.random.default_rng(n)
    for h0 in high_groups(n):
        tiles = 1 << (n - T)
        for g in [0, tiles - 1] + [int(x) for x in rng.integers(0, tiles, 6)]:
            base = tile_base(g, 3, h0)
            start = lay_addr(base, h0)
            assert start % (1 << T) == 0
            offs = [lay_addr(tile_index(base, o, 3, h0), h0) - start for o in range(1 << T)]
            assert offs == list(range(1 << T))


def test_layouts_are_on_at_the_headline_geometry():
    assert high_groups(30) == [12, 21]
    assert layouts_fit(30)
    assert layouts_fit(30, r_low=3, r_high=3)
    # where the last group's register bits straddle the layout below it
    # (31 qubits: bits 27..30 against layout 19's edge at bit 28) the engine
    # keeps the natural order
    assert high_groups(31) == [12, 19, 22]
    assert not layouts_fit(31)
    # the qubit counts with layouts (both register shapes of the high passes)
    assert [n for n in range(13, 35) if layouts_fit(n)] == [
        13, 14, 15, 16, 17, 21, 25, 26, 27, 28, 29, 30, 32, 33, 34]


@pytest.mark.parametrize("n", list(range(13, 35)))
def test_register_bits_stay_contiguous_where_layouts_fit(n):
    # lay_fits is what the launchers check; here its verdict is recomputed by
    # brute force over the register index for every pass the plan makes
    if not layouts_fit(n):
        return
    h0s = high_groups(n)
    K = len(h0s)

    def contiguous(h, bits):
        pos = [lay_addr(1 << b, h).bit_length() - 1 for b in bits]
        return pos == list(range(pos[0], pos[0] + len(bits)))

    assert contiguous(h0s[K - 1], range(T - 4, T))
    for j in range(K):
        first = h0s[j] + (T - 4 - 3)
        assert contiguous(h0s[j], range(first, first + 4))
        assert contiguous(h0s[j - 1] if j > 0 else 0, range(first, first + 4))


def _lay_vec(i, h):
    if h == 0:
        return i
    return ((i & 7) | (((i >> h) & 511) << 3) | (((i >> 3) & ((1 << (h - 3)) - 1)) << 12)
            | ((i >> (h + 9)) << (h + 9)))


@pytest.mark.parametrize("n", [16, 21, 25, 30])
def test_kernel_address_arithmetic_lands_on_the_layout(n):
    # The kernels form an address as lay(base) + lay(thread offset) + k *
    # lay(register step) (lay_tile_addr, store_lam); reader and writer of a
    # boundary agree when that equals lay(index) for every (tile, thread,
    # register).  Low pass: offset t | k << 8 (mapping 0 of the low tile);
    # high pass with first high bit h0: offset (t & 7) | (t >> 3) << h0,
    # register step 2^(h0 + 5).  All tiles for n <= 22, a sample at 30.
    assert layouts_fit(n)
    h0s = high_groups(n)
    K = len(h0s)
    tiles = 1 << (n - T)
    if n <= 21:
        gs = np.arange(tiles, dtype=np.int64)
    else:
        rng = np.random.default_rng(n)
        gs = np.unique(np.concatenate([[0, tiles - 1], rng.integers(0, tiles, 64)])).astype(np.int64)
    t = np.arange(256, dtype=np.int64)
    k = np.arange(16, dtype=np.int64)
    G, Tt, Kk = np.meshgrid(gs, t, k, indexing="ij")

    def check(base, off, step, h):
        index = base + off + Kk * step
        addr = _lay_vec(base, h) + _lay_vec(off, h) + Kk * _lay_vec(np.int64(step), h)
        assert np.array_equal(addr, _lay_vec(index, h))
        if len(gs) == tiles:  # the pass covers every index once
            assert np.array_equal(np.sort(index.ravel()), np.arange(1 << n))
            assert np.array_equal(np.sort(addr.ravel()), np.arange(1 << n))

    # low pass store into the last high pass's layout
    check(G << T, Tt, 1 << 8, h0s[K - 1])
    for j in range(K):
        h0 = h0s[j]
        base = ((G & ((1 << (h0 - 3)) - 1)) << 3) | ((G >> (h0 - 3)) << (h0 + 9))
        off = (Tt & 7) | ((Tt >> 3) << h0)
        check(base, off, 1 << (h0 + 5), h0)                        # its read
        check(base, off, 1 << (h0 + 5), h0s[j - 1] if j > 0 else 0)  # its write
