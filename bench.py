"""Throughput of the reconstruction path on one MI355X.

Workload (full-size single-pair shape, batched): `--pairs` independently
seeded 2448x2048 rectified pairs, three-frequency heterodyne 4-step fringes
(70/64/59 across a 1280-px projector), uint16 12-bit, rendered on the GPU
before timing. One step = decode L + decode R + reconstruct (order match,
5x5 median, Q reprojection, 5x5 Gaussian sigma 1, SOR k=30, MLS order 2 at
1.5 mm) of every pair, frames already resident in HBM.

Prints one JSON line: metric, value (Mpts/s over all pairs, one point =
one left rectified pixel), per-path numbers, the dominant kernel's HBM
roofline and the CPU restatement's throughput on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts/s end-to-end 5MP 3-freq, batched vs per-pair, at ≤1e-3 mm vs FP64"
HBM_PEAK_GBS = 8000.0
ROWS, COLS = 2048, 2448


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=16)
    ap.add_argument("--path", choices=["plain"], default="plain")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-rows", type=int, default=ROWS)
    return ap.parse_args()


def algorithmic_bytes(family: str, npix: int) -> float:
    """Compulsory HBM bytes of one launch of a kernel family over npix pixels
    (inputs read once, outputs written once)."""
    per = {
        "decode": 12 * 2 + 8 + 4 + 8 + 0,   # 12 u16 frames -> phase f64, k i32, mod f64
        "post_smooth": 8 + 8,
        "match": 8 + 8 + 4 + 4 + 8 + 4 + 4,  # l/r phase, l/r k -> d64, d32, col
        "median5": 4 + 4,
        "reproject": 4 + 12 + 1,
        "gaussian": 13 + 13,
        "sor_knn": 13 + 8,
        "sor_stats": 8 + 12 + 13,
        "mls": 13 + 12,
    }[family]
    return float(per) * npix


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs shard across GPUs with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import zim_amd
    from zim_amd import _capi, synth
    from tests import scenes

    dev = torch.device("cuda:0")
    ctx = zim_amd.Context(0)
    dcfg = scenes.c2_decode_config()
    rcfg = scenes.c2_recon_config()

    print(f"[bench] rendering {args.pairs} pairs of {COLS}x{ROWS} on the GPU", file=sys.stderr,
          flush=True)
    t0 = time.time()
    pairs = []
    for i in range(args.pairs):
        p = synth.render_c2(1000 + i, device=dev)
        pairs.append(p)
    torch.cuda.synchronize()
    print(f"[bench] rendered in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    npix = ROWS * COLS
    outs = []
    for _ in pairs:
        outs.append(dict(
            lph=torch.empty((ROWS, COLS), dtype=torch.float64, device=dev),
            rph=torch.empty((ROWS, COLS), dtype=torch.float64, device=dev),
            lk=torch.empty((ROWS, COLS), dtype=torch.int32, device=dev),
            rk=torch.empty((ROWS, COLS), dtype=torch.int32, device=dev),
            mod=torch.empty((ROWS, COLS), dtype=torch.float64, device=dev),
            xyz=torch.empty((ROWS, COLS, 3), dtype=torch.float32, device=dev),
            P=torch.empty((ROWS, COLS), dtype=torch.uint8, device=dev),
        ))
    lib = ctx.lib
    import ctypes as C
    Qa = [np.ascontiguousarray(p.Q, np.float64) for p in pairs]
    Ra = [np.ascontiguousarray(p.R1, np.float64) for p in pairs]

    def run_plain():
        ctx.bind_stream()
        for p, o, q, r in zip(pairs, outs, Qa, Ra):
            for frames, ph, k in ((p.left, o["lph"], o["lk"]), (p.right, o["rph"], o["rk"])):
                arr = (C.c_void_p * len(frames))(*[f.data_ptr() for f in frames])
                _capi.check(lib.psgc_decode(ctx.h, C.byref(dcfg), arr, len(frames),
                                            _capi.SAMPLE_U16, ROWS, COLS, COLS * 2,
                                            _capi.ptr(ph), _capi.ptr(k), _capi.ptr(o["mod"]),
                                            None, _capi.MEM_DEVICE))
            _capi.check(lib.psgc_reconstruct(ctx.h, C.byref(rcfg), _capi.ptr(o["lph"]),
                                             _capi.ptr(o["rph"]), _capi.ptr(o["lk"]),
                                             _capi.ptr(o["rk"]), ROWS, COLS, 70, _capi.ptr(q),
                                             _capi.ptr(r), _capi.ptr(o["xyz"]),
                                             _capi.ptr(o["P"]), None, None, _capi.MEM_DEVICE))

    run = run_plain
    for w in range(args.warmup):
        run()
        torch.cuda.synchronize()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)

    families = ["decode", "post_smooth", "match", "median5", "reproject", "gaussian", "sor_knn",
                "sor_stats", "mls"]
    lib.psgc_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    prof = {}
    for fam in families:
        n = C.c_int64()
        ms = C.c_double()
        _capi.check(lib.psgc_profile_read(fam.encode(), C.byref(n), C.byref(ms)))
        if n.value:
            prof[fam] = (n.value, ms.value)
    lib.psgc_profile_enable(0)
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step", file=sys.stderr,
          flush=True)

    pts = npix * args.pairs
    value = pts / (ms_per_step * 1e-3) / 1e6
    valid = sum(int((o["P"] != 0).sum().item()) for o in outs)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dom_name, (dom_n, dom_ms) = dom
    # launches of the dominant family per step: decode runs per camera, mls per iteration
    avg_ms = dom_ms / dom_n
    alg = algorithmic_bytes(dom_name, npix)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    breakdown = {k: round(v[1] / args.steps, 3) for k, v in prof.items()}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pairs[0], dcfg, rcfg, args.cpu_sample_rows)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 decode/match/triangulation, f32 cloud filters",
        "data": "synthetic (seeded bump height fields rendered on the GPU)",
        "config": {"workload": f"{args.pairs} seeded 2448x2048 pairs, 3-freq 70/64/59 4-step "
                               "u16, SOR k=30, MLS order 2 r=1.5 mm, Gaussian 5x5 s=1",
                   "pairs": args.pairs, "path": args.path, "rows": ROWS, "cols": COLS},
        "valid_mpts_per_s": round(valid / args.steps / (ms_per_step * 1e-3) / 1e6 * args.steps, 3)
        if False else round(valid / (ms_per_step * 1e-3) / 1e6, 3),
        "stage_ms_per_step": breakdown,
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "avg_launch_ms": round(avg_ms, 4),
                     "alg_bytes_per_launch": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(pair, dcfg, rcfg, sample_rows):
    """The CPU restatement of the reference path (oracle/), timed on the
    host's cores over one pair (rows [0, sample_rows))."""
    import numpy as np

    import oracle
    cores = os.cpu_count() or 1
    left = [f[:sample_rows].cpu().numpy() for f in pair.left]
    right = [f[:sample_rows].cpu().numpy() for f in pair.right]
    print(f"[bench] cpu baseline: one pair, {sample_rows} rows, {cores} threads", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    L = oracle.decode(dcfg, left)
    R = oracle.decode(dcfg, right)
    oracle.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, pair.Q, pair.R1)
    dt = time.perf_counter() - t0
    pts = sample_rows * COLS
    return {"value": round(pts / dt / 1e6, 4), "unit": "Mpts/s", "cores": cores, "kind": "port",
            "sample": f"1 pair, {sample_rows}x{COLS} px, decode L+R + reconstruct, {dt:.1f} s"}


if __name__ == "__main__":
    main()
