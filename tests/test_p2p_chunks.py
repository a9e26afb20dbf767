"""CPU replay of the batched point-to-point step's work split (runtime/batched_p2p.h): the
scan's chunk offsets, each wave's equal range of chunks, the 64-ary search for the wave's
first pair, the forward walk over empty pairs, and each chunk's piece of its destination.
Over random pair lists and wave counts, every destination byte of every pair is written
exactly once and no byte outside the pairs is written; every piece's aligned interior fits
one chunk (at most 8 vectors per lane) with heads and tails under 16 bytes."""
import numpy as np
import pytest

SHIFT = 13
CHUNK = 1 << SHIFT
WAVE = 64


def chunks(d, b):
    return 0 if b == 0 else ((d + b - 1) >> SHIFT) - (d >> SHIFT) + 1


def replay(pairs, nw):
    off = [0]
    for d, b in pairs:
        off.append(off[-1] + chunks(d, b))
    total, n = off[-1], len(pairs)
    writes = {}
    for w in range(nw):
        c0, c1 = total * w // nw, total * (w + 1) // nw
        if c0 >= c1:
            continue
        lo, hi = 0, n
        while hi - lo > 1:                       # p2p_copy_kernel's search
            step = (hi - lo + WAVE - 1) // WAVE
            k = max(l for l in range(WAVE) if lo + l * step < hi and off[lo + l * step] <= c0)
            lo += k * step
            hi = min(lo + step, hi)
        i, nxt = lo, off[lo + 1]
        assert off[i] <= c0 < nxt
        for c in range(c0, c1):
            while c >= nxt:                      # skips empty pairs
                i += 1
                nxt = off[i + 1]
            d, b = pairs[i]
            blk = ((d >> SHIFT) + c - off[i]) << SHIFT
            plo, phi = max(blk, d), min(blk + CHUNK, d + b)
            assert plo < phi and phi - plo <= CHUNK
            a, bb = (plo + 15) & ~15, phi & ~15
            if a < bb:
                assert a - plo < 16 and phi - bb < 16 and (bb - a) // 16 <= WAVE * (CHUNK // 16 // WAVE)
            for x in range(plo, phi):
                writes[x] = writes.get(x, 0) + 1
    return writes


@pytest.mark.parametrize("seed", range(6))
def test_every_destination_byte_written_once(seed):
    rng = np.random.default_rng(seed)
    for _ in range(40):
        n = int(rng.integers(1, 50))
        cur = int(rng.integers(0, 100000))
        pairs = []
        for _ in range(n):
            b = int(rng.choice([0, 1, 15, 16, 17, 8191, 8192, 8193, 20000, int(rng.integers(0, 70000))]))
            pairs.append((cur, b))
            cur += b + int(rng.integers(0, 50))
        nw = int(rng.choice([1, 3, 64, 4096]))
        w = replay(pairs, nw)
        expected = {x for d, b in pairs for x in range(d, d + b)}
        assert set(w) == expected
        assert all(v == 1 for v in w.values())
