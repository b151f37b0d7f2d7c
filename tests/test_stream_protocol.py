"""CPU model of the stream-staged SpMV's ring hand-over (no GPU).

The stream kernel's streaming waves keep finished segment sums in per-wave
LDS rings and hand them to one writer wave per work-group through three
counters per wave (pub: ordinals below it are final; con: ordinals below it
are written and their slots free; done).  This model rebuilds the stream
layout of small matrices the way the analysis does (segments per row, a
stable order by panel, 512-entry steps, 16384-entry slices, seven slices per
work-group), then replays each work-group's streaming waves and writer in a
random interleaving with the kernel's rules:

  streaming wave, per step: publish pub = open; if the step would span a ring
  or more from the oldest unwritten ordinal, wait until the writer has
  written all it will (con at least open's line less its hold-back) and, if
  the step still does not fit, write its segments directly (with the
  held-back tail) and hand the ring back at the step's end;
  writer: write [con, pub) in whole 16-ordinal lines until the wave is done.

It checks that every wait ends, that the slots a step writes never hold
ordinals the writer has not written, and that every segment ordinal is
written to HBM exactly once (by the writer, a direct step or a closed
slice's end).

Every wait of the kernel is bounded (kStreamSpinLimit sleeps): a streaming
wave that reaches the bound records a fault, stops writing and reports done;
a writer that finds nothing to write that many passes in a row records a
fault and leaves.  The model runs the same rules with a spin bound and shows
that a work-group whose writer is gone (the kernel's test setting, bound 0)
ends with the fault recorded and no ordinal written twice.
"""
import random

import numpy as np
import pytest

from asc_sparse_hip import gen as G

TILE, SLICE, WAVES, RING, LINE = 512, 32 * 512, 7, 512, 16
M32 = 0xFFFFFFFF


def s32(v):
    v &= M32
    return v - (1 << 32) if v >= 1 << 31 else v


def stream_layout(offsets, cols_idx, cols, long_min, width):
    """Head flags of the padded stream and the slices {start, end, open
    ordinal, closed}, grouped per work-group."""
    panels = -(-cols // width)
    panels = (panels + 7) // 8 * 8
    pc = (cols - 1) // panels + 1
    lens = np.diff(offsets)
    listed = np.nonzero(lens >= long_min)[0]
    entries = np.concatenate([np.arange(offsets[r], offsets[r + 1]) for r in listed]) \
        if len(listed) else np.zeros(0, np.int64)
    rowi = np.repeat(np.arange(len(listed)), lens[listed])
    q = cols_idx[entries] // pc
    head = np.ones(len(entries), bool)
    head[1:] = (rowi[1:] != rowi[:-1]) | (q[1:] != q[:-1])
    seg_start = np.nonzero(head)[0]
    seg_len = np.diff(np.append(seg_start, len(entries)))
    seg_panel = q[seg_start]
    order = np.argsort(seg_panel, kind="stable")
    lens_s = seg_len[order]
    dest = np.concatenate([[0], np.cumsum(lens_s)[:-1]]) if len(order) else np.zeros(0, np.int64)
    nseg = len(order)
    pseg = np.searchsorted(seg_panel[order], np.arange(panels + 1), "left")
    total = int(lens_s.sum())
    raw_start = [int(dest[pseg[p]]) if pseg[p] < nseg else total for p in range(panels)] + [total]
    heads, wgs, padded = [], [], 0
    for p in range(panels):
        raw, ln = raw_start[p], raw_start[p + 1] - raw_start[p]
        plen = -(-ln // TILE) * TILE
        hp = np.zeros(plen + 1, bool)  # one past the end: no head
        hp[dest[pseg[p]:pseg[p + 1]] - raw] = True
        slices = []
        ns = -(-ln // SLICE)
        for j in range(ns):
            b, e = j * SLICE, min((j + 1) * SLICE, ln)
            lo, hi = pseg[p], pseg[p + 1] - 1  # segment holding raw + b
            while lo < hi:
                mid = (lo + hi + 1) >> 1
                if dest[mid] <= raw + b:
                    lo = mid
                else:
                    hi = mid - 1
            ordinal = (int(lo) - (1 if dest[lo] == raw + b else 0)) & M32
            closed = j + 1 == ns or bool(hp[e])
            slices.append((b, e, ordinal, closed))
        for j in range(0, ns, WAVES):
            wgs.append((hp, slices[j:j + WAVES]))
        padded += plen
    return wgs, nseg


def run_workgroup(hp, slices, rng, written, spin_limit=100_000, writer=True):
    """Replays one work-group; returns True when a wait gave up (the
    kernel's fault word)."""
    lds = {}
    fault = False
    waves = []
    for w, (s0, s1, o, closed) in enumerate(slices):
        lds["pub", w] = lds["con", w] = o
        lds["done", w] = 0
        waves.append({"s0": s0, "s1": s1, "open": o, "t": s0, "closed": closed, "spins": 0})
    for w, st in enumerate(waves):  # after the barrier
        st["con_c"] = (st["open"] + (1 if hp[st["s0"]] else 0)) & M32
        lds["con", w] = st["con_c"]
    finished, live = set(), set(range(len(waves))) | ({"writer"} if writer else set())
    idle = 0

    def write(o):
        written[o] = written.get(o, 0) + 1

    steps = 0
    while live:
        steps += 1
        assert steps < 5_000_000, "no progress"
        a = rng.choice(sorted(live, key=str))
        if a == "writer":
            worked = False
            for w in range(len(waves)):
                if w in finished:
                    continue
                d, p, c = lds["done", w], lds["pub", w], lds["con", w]
                upto = p if d else p & ~(LINE - 1)
                if s32(upto - c) > 0:
                    for k in range(s32(upto - c)):
                        write((c + k) & M32)
                    lds["con", w] = upto
                    worked = True
                if d:
                    finished.add(w)
            if len(finished) == len(waves):
                live.discard("writer")
            elif worked:
                idle = 0
            else:
                idle += 1
                if idle > spin_limit:  # bounded: record the fault and leave
                    fault = True
                    live.discard("writer")
            continue
        st = waves[a]
        t, s1, op = st["t"], st["s1"], st["open"]
        total = int(hp[t:min(t + TILE, s1)].sum())
        end = (op + total) & M32
        lds["pub", a] = op

        def span(c):
            return (end - (c if s32(c - op) < 0 else op)) & M32

        if span(st["con_c"]) >= RING:
            st["con_c"] = lds["con", a]
            if not (span(st["con_c"]) < RING or s32(st["con_c"] - ((op & ~(LINE - 1)) & M32)) >= 0):
                if st["spins"] >= spin_limit:  # bounded: give up, write nothing more
                    fault = True
                    lds["done", a] = 1
                    live.discard(a)
                    continue
                st["spins"] += 1
                continue
        st["spins"] = 0
        if span(st["con_c"]) >= RING:  # direct step
            assert lds["con", a] == st["con_c"] and s32(st["con_c"] - ((op & ~(LINE - 1)) & M32)) >= 0
            for k in range(max(0, s32(op - st["con_c"]))):  # the held-back tail
                write((st["con_c"] + k) & M32)
            for k in range(total):
                if k == 0 and hp[t] and t == st["s0"]:
                    continue  # a slice opening on a head leaves that ordinal to the previous slice
                write((op + k) & M32)
            lds["con", a] = lds["pub", a] = st["con_c"] = end
        else:  # slots [oldest unwritten, end] are distinct in the ring
            assert span(lds["con", a]) < RING
        st["open"], st["t"] = end, t + TILE
        if not t + TILE < s1:
            lds["pub", a] = end
            lds["done", a] = 1
            if st["closed"]:
                write(end)
            live.discard(a)
    return fault


def _check(offsets, idx, cols, long_min, width, seed):
    wgs, nseg = stream_layout(offsets, idx.astype(np.int64), cols, long_min, width)
    rng, written = random.Random(seed), {}
    for hp, slices in wgs:
        assert not run_workgroup(hp, slices, rng, written), "a wait reached its bound"
    bad = [o for o in range(nseg) if written.get(o, 0) != 1]
    assert not bad, f"{len(bad)} of {nseg} ordinals not written exactly once, first {bad[:5]}"
    assert set(written) <= set(range(nseg))


@pytest.mark.parametrize("width,long_min", [(16384, 1), (300, 16), (4, 1)])
def test_ring_handover_uniform(width, long_min):
    # width 4: 1024 panels of 4 columns (short panel streams, many slices that
    # open on a head); width 300: the first slice of every panel opens on a
    # head.
    m = G.uniform16(4096)
    _check(m.offsets, m.indices, m.cols, long_min, width, seed=width)


def _one_per_row(rows=60000, cols=8 * 16384, seed=9):
    # Every entry is its own segment: full steps hold 512 heads (direct steps).
    rng = np.random.default_rng(seed)
    return np.arange(rows + 1, dtype=np.int64), rng.integers(0, cols, rows).astype(np.int64), cols


def test_ring_handover_direct_steps():
    offsets, idx, cols = _one_per_row()
    _check(offsets, idx, cols, 1, 16384, seed=3)


def test_ring_handover_zipf():
    m = G.make(G.ZIPF, 1 << 12, 1 << 12, seed=23, values=G.VALUES_NORMAL)
    _check(m.offsets, m.indices, m.cols, 1, 64, seed=5)


def test_model_catches_the_unbounded_wait():
    # Planted mistake: waiting for con to reach open exactly never ends for a
    # slice that opens on a head (it starts with con one past open) when its
    # first step holds more segments than the ring has slots; with a
    # 256-slot ring a full step of 512 segments does.
    offsets, idx, cols = _one_per_row()
    wgs, _ = stream_layout(offsets, idx, cols, 1, 16384)
    hp, slices = wgs[0]
    s0, s1, o, _ = slices[0]
    total = int(hp[s0:min(s0 + TILE, s1)].sum())
    con = (o + 1) & M32  # opens on a head
    end = (o + total) & M32
    ring = 256
    assert hp[s0] and total > ring
    assert not ((end - con) & M32 < ring or con == o)  # the planted exit test: never true
    line = (o & ~(LINE - 1)) & M32
    assert s32(con - line) >= 0  # the kernel's exit test holds at once


def _five_per_row(rows=40000, cols=16384, seed=12):
    # One panel, rows of 5 entries: about 100 segments per 512-entry step, so
    # each wave's ring fills after a few steps and needs the writer.
    rng = np.random.default_rng(seed)
    idx = np.sort(np.stack([rng.choice(cols, 5, replace=False) for _ in range(rows)]), 1)
    return np.arange(0, 5 * rows + 1, 5, dtype=np.int64), idx.reshape(-1).astype(np.int64), cols


def test_waits_give_up_without_a_writer():
    # The kernel's test setting: no writer wave and every wait bounded at 0.
    # Each streaming wave reaches a full ring, gives up, and the work-group
    # ends with the fault recorded; nothing is written twice.
    offsets, idx, cols = _five_per_row()
    wgs, nseg = stream_layout(offsets, idx, cols, 1, 16384)
    rng, written = random.Random(4), {}
    for hp, slices in wgs:
        assert run_workgroup(hp, slices, rng, written, spin_limit=0, writer=False)
    assert all(v == 1 for v in written.values())
    assert len(written) < nseg


@pytest.mark.parametrize("limit", [0, 1, 3])
def test_tight_bounds_end_with_a_fault_or_a_complete_result(limit):
    # Whatever the interleaving, tight bounds end every wave and the writer:
    # the work-group either completes (each ordinal once) or records a fault
    # (no ordinal twice).
    offsets, idx, cols = _five_per_row(rows=20000)
    wgs, nseg = stream_layout(offsets, idx, cols, 1, 16384)
    rng, written = random.Random(limit), {}
    faults = [run_workgroup(hp, slices, rng, written, spin_limit=limit) for hp, slices in wgs]
    assert all(v == 1 for v in written.values())
    if not any(faults):
        assert len(written) == nseg
