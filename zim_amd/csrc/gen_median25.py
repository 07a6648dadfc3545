"""Writes median25.inc: a branch-free selection network for the median (13th
smallest) of 25 values.

Construction: Batcher's odd-even merge sort on 32 wires, the 7 extra wires
held at +inf and constant-folded away, then every comparator whose outputs
cannot reach wire 12 removed (and comparators of which only the min or only
the max is live reduced to that half). Correctness is checked exhaustively
over all 2^25 zero-one inputs (the 0-1 principle applies to comparator
networks). Run from this directory:
    python gen_median25.py > median25.inc
"""
import sys

import numpy as np

N, W, TARGET = 25, 32, 12


def batcher(n):
    ops = []

    def merge(lo, hi, r):
        step = r * 2
        if step < hi - lo:
            merge(lo, hi, step)
            merge(lo + r, hi, step)
            for i in range(lo + r, hi - r, step):
                ops.append((i, i + r))
        else:
            ops.append((lo, lo + r))

    def sort(lo, hi):
        if hi - lo >= 1:
            mid = lo + (hi - lo) // 2
            sort(lo, mid)
            sort(mid + 1, hi)
            merge(lo, hi, 1)

    sort(0, n - 1)
    return ops


def build():
    # wire contents: ('v', id) for a variable, ('inf',) for +inf
    wires = [("v", i) for i in range(N)] + [("inf",)] * (W - N)
    nxt = N
    prog = []  # (out_id_min, out_id_max, in_a, in_b)
    for a, b in batcher(W):
        wa, wb = wires[a], wires[b]
        if wa[0] == "inf" and wb[0] == "inf":
            continue
        if wb[0] == "inf":
            continue
        if wa[0] == "inf":
            wires[a], wires[b] = wb, wa
            continue
        lo, hi = nxt, nxt + 1
        nxt += 2
        prog.append((lo, hi, wa[1], wb[1]))
        wires[a], wires[b] = ("v", lo), ("v", hi)
    out = wires[TARGET][1]
    live = {out}
    kept = []
    for lo, hi, ia, ib in reversed(prog):
        need_lo, need_hi = lo in live, hi in live
        if need_lo or need_hi:
            kept.append((lo if need_lo else None, hi if need_hi else None, ia, ib))
            live.add(ia)
            live.add(ib)
    kept.reverse()
    return kept, out


def verify(kept, out):
    total = 1 << N
    chunk = 1 << 20
    for start in range(0, total, chunk):
        idx = np.arange(start, start + chunk, dtype=np.int64)
        vals = {i: ((idx >> i) & 1).astype(np.uint8) for i in range(N)}
        for lo, hi, ia, ib in kept:
            a, b = vals[ia], vals[ib]
            if lo is not None:
                vals[lo] = np.minimum(a, b)
            if hi is not None:
                vals[hi] = np.maximum(a, b)
        ones = np.zeros(chunk, np.int64)
        for i in range(N):
            ones += (idx >> i) & 1
        truth = (ones >= N - TARGET).astype(np.uint8)  # 13th smallest is 1 iff >= 13 ones
        if not np.array_equal(vals[out], truth):
            raise SystemExit("median network failed the 0-1 check")


def main():
    kept, out = build()
    verify(kept, out)
    n_ops = sum((lo is not None) + (hi is not None) for lo, hi, _, _ in kept)
    print(f"// median of v0..v24 in {n_ops} min/max operations "
          "(generated by gen_median25.py, checked on all 2^25 0-1 inputs)")
    for lo, hi, ia, ib in kept:
        if lo is not None:
            print(f"const float t{lo} = fminf(t{ia}, t{ib});")
        if hi is not None:
            print(f"const float t{hi} = fmaxf(t{ia}, t{ib});")
    print(f"#define PSGC_MEDIAN25_RESULT t{out}")


if __name__ == "__main__":
    sys.exit(main())
