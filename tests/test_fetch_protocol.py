"""Model check of the traversal's work-fetch protocol (eos_amd/csrc/trace.h,
fetch_items), run on the CPU: a restatement of the wave-uniform bookkeeping
(ranged cursors, guided chunk sizes, hand-out to idle lanes) driven by random
lane completions.  Properties: every item of a launch is issued exactly once
and every wave terminates, for item counts from 0 up, fewer waves than ranges,
every refill threshold and ray batch.  (The GPU tests check every item's
result; this pins the protocol itself, including schedules the GPU rarely
produces.)"""
import math
import random

import pytest

RANGES = 8


def _c_div(a, b):
    """C integer division (truncates toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _run(n, n_waves, thresh, batch, seed, grid_waves=None):
    """n_waves take part (the launch's other waves, up to grid_waves, left at
    once because the items cannot occupy them); range of wave w: its
    workgroup's index mod 8 (four waves per workgroup)."""
    rnd = random.Random(seed)
    cursors = [0] * RANGES
    issued = [0] * n
    waves = [dict(item=[-1] * 64, drained=False, nxt=0, end=0, seen=0,
                  rng=(w // 4 if grid_waves else w) % RANGES, tried=0, done=False)
             for w in range(n_waves)]
    wpr = max(1, (grid_waves or n_waves) // RANGES)
    steps = 0
    while not all(w["done"] for w in waves):
        steps += 1
        assert steps < 200_000, "a wave did not terminate"
        for wi in rnd.sample(range(n_waves), n_waves):
            w = waves[wi]
            if w["done"]:
                continue
            item = w["item"]
            idle = [lane for lane in range(64) if item[lane] < 0]
            nidle = len(idle)
            if nidle >= thresh and (not w["drained"] or w["nxt"] < w["end"]):
                avail, old = w["end"] - w["nxt"], w["nxt"]
                base = take = 0
                if avail < nidle and not w["drained"]:
                    want = nidle - avail
                    for _ in range(RANGES):
                        r = w["rng"]
                        lo, hi = n * r // RANGES, n * (r + 1) // RANGES
                        size = max(min(_c_div(hi - lo - w["seen"], 2 * wpr), batch), want)
                        got = cursors[r]
                        cursors[r] += size
                        if got < hi - lo:
                            w["seen"] = got + size
                            base, take = lo + got, min(size, hi - lo - got)
                            break
                        w["rng"], w["seen"] = (r + 1) % RANGES, 0
                        w["tried"] += 1
                        if w["tried"] == RANGES:
                            w["drained"] = True
                            break
                for rank, lane in enumerate(idle):
                    my = old + rank if rank < avail else (base + rank - avail if rank - avail < take else -1)
                    if my >= 0:
                        assert 0 <= my < n
                        item[lane] = my
                        issued[my] += 1
                if avail >= nidle:
                    w["nxt"] = old + nidle
                else:
                    used = min(nidle - avail, take)
                    w["nxt"], w["end"] = (base + used, base + take) if take > used else (0, 0)
            if all(x < 0 for x in item):
                if w["drained"] and w["nxt"] >= w["end"]:
                    w["done"] = True
                continue
            for lane in range(64):
                if item[lane] >= 0 and rnd.random() < 0.35:
                    item[lane] = -1
    assert all(c == 1 for c in issued)


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 64, 65, 1000, 9001])
@pytest.mark.parametrize("n_waves", [1, 3, 8, 17])
def test_every_item_once_and_every_wave_ends(n, n_waves):
    for thresh in (1, 16, 64):
        for batch in (16, 128):
            _run(n, n_waves, thresh, batch, n * 31 + n_waves * 7 + thresh + batch)


@pytest.mark.parametrize("n", [1, 3, 40, 200, 5000])
def test_launches_with_few_items_use_only_the_waves_they_need(n):
    """A launch with few items on the full persistent grid (5120 waves): only
    the first ceil(n / items per wave) waves take part (ranges by workgroup),
    chunk sizes still computed for the whole grid; every item once."""
    for per_wave in (4, 16, 64):
        need = max(1, -(-n // per_wave))
        _run(n, min(need, 64), 1, 128, n + per_wave, grid_waves=5120)


def _scatter_map(n, b, rounds=1, lanes=1 << 30):
    """Restatement of the scattered claim order (trace.h Scatter): claim ->
    item for a launch of n items in blocks of b."""
    on = rounds > 0 and n > 1 and n <= lanes * rounds
    b = max(b, 1)
    m = n // b
    q = max(int(m * 0.6180339887498949), 1)
    while m > 1 and math.gcd(q, m) != 1:
        q += 1
    out = []
    for i in range(n):
        if not on or i >= m * b:
            out.append(i)
        else:
            k, w = divmod(i, b)
            out.append((k * q) % m * b + w)
    return out


@pytest.mark.parametrize("b", [1, 2, 4, 7, 32])
def test_scattered_claim_order_is_a_bijection(b):
    """Every item is claimed exactly once under the scattered order, for item
    counts that are and are not multiples of the block, including counts whose
    golden-ratio multiplier needs stepping to become coprime."""
    for n in list(range(0, 70)) + [97, 128, 1000, 1024, 4096, 65536 + 3, 266688]:
        got = _scatter_map(n, b)
        assert sorted(got) == list(range(n)), (n, b)
        if n >= 64 and b == 1:
            # neighbouring claims land about n / golden ratio^2 apart or more
            assert min(abs(got[i + 1] - got[i]) for i in range(min(n - 1, 200))) >= n // 4
