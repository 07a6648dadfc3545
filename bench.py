"""Throughput of the reconstruction path on one MI355X.

Workload: `--pairs` independently seeded 2448x2048 rectified pairs (the
full-size single-pair shape), three-frequency heterodyne 4-step fringes
(70/64/59 across a 1280-px projector), uint16 12-bit, rendered on the GPU
before timing. One step = decode L + decode R + reconstruct (order match,
5x5 median, Q reprojection, 5x5 Gaussian sigma 1, SOR k=30, MLS order 2 at
1.5 mm in the smallest window that holds that radius for the pair, 2
iterations) of every pair, frames resident in HBM.

value = Mpts/s of the batched pipeline (one hipGraph replay per step; one
point = one left rectified pixel). Beside it, on the same pairs: the plain
per-pair GPU path (one launch per stage per pair, one-thread-per-pixel
kernels, a sync after each pair, as the reference GPU path is structured)
and the tiled kernels run one pair at a time (one graph per pair, a sync
after each), which separates what batching gains from what the kernels
gain. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts/s end-to-end 5MP 3-freq, batched vs per-pair, at ≤1e-3 mm vs FP64"
HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 1 << 30
ROWS, COLS = 2048, 2448


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--plain-pairs", type=int, default=None,
                    help="pairs timed through the per-pair paths (default: all)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-per-pair-tiled", action="store_true")
    ap.add_argument("--cpu-sample-rows", type=int, default=ROWS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the measurement, copy 1 GiB device-to-device once (a kernel "
                         "of known traffic that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--c5", action="store_true",
                    help="instead: the merged-cloud filters (16 views of one 2448x2048 scene "
                         "merged into one unorganized cloud; SOR k=50, MLS order 2 at 1 mm, "
                         "Gaussian), timed per filter")
    ap.add_argument("--c5-views", type=int, default=16)
    ap.add_argument("--c3", action="store_true",
                    help="instead: one batch of 64 uneven pairs (640x480 to 2448x2048; "
                         "3-, 4- or 8-step three-frequency fringes or a Gray-code schedule), "
                         "batched against the same pairs one at a time")
    ap.add_argument("--c4", action="store_true",
                    help="instead: the differentiation workload at full size (1280x1024 plane "
                         "target): adjoint timing, central-difference check, refinement")
    ap.add_argument("--ref-timing", action="store_true",
                    help="instead: the configuration of the reference's own published GPU "
                         "timing (5120x5120 8-bit, 107/100/94 x 8 steps, MLS off), one pair at "
                         "a time, device-resident and host-to-host")
    return ap.parse_args()


# Compulsory HBM bytes per pixel of one launch of each kernel family
# (inputs read once, outputs written once), for the tiled batch pipeline.
ALG_BYTES_PER_PX = {
    "decode": 12 * 2 + 8 + 4,         # 12 u16 frames -> raw phase f64 + order i32
    "match": 2 * (8 + 4) + 4,         # l/r raw phase + order -> f32 disparity
    "median5": 4 + 4,
    "repro_gauss": 4 + 12 + 1,        # disparity -> Gaussian xyz + mask
    "sor_knn": 12 + 1 + 8,            # xyz + mask -> mean distance f64
    "sor_stats": 8,
    "mls": 12 + 1 + 8 + 12,           # xyz, mask, mean distance -> xyz
}


MLS_RADIUS = 1.5  # mm


# VALU lane-instructions per left pixel of one pass over the 64 pairs
# (SQ_INSTS_VALU x 64 / pixels; MLS: the window-class launches of one
# iteration, per kernel variant) on this workload, from the rocprofv3 counter
# run summarized in profiles/c2_64pair_valu_counters.md. With the live pass
# time they give the wave-instruction rate, set against the rate at which
# FP32 FMAs reach MI355X's published vector peak: 157.3 TFLOP/s / (2 flops x
# 64 lanes) = 1229 G wave-instructions/s.
VALU_PER_PX_PROFILED = {"mls/0": 11400, "mls/1": 11152, "decode": 1889, "sor_knn": 1649,
                        "match": 573, "repro_gauss": 504, "median5": 271}
VALU_PEAK_GINST_S = 157.3e3 / 2 / 64
# The published FP32 vector peak counts packed FP32 FMAs (two per lane per
# instruction); packed FP32 is disabled in this build, so single-lane FMAs
# issue at most half that rate.
VALU_UNPACKED_PEAK_GINST_S = VALU_PEAK_GINST_S / 2


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (pairs would "
              "shard across GPUs with no exchange, one process per GPU)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import zim_amd
    from zim_amd import _capi, synth
    from tests import scenes

    if args.c5:
        return run_c5(args)
    if args.c3:
        return run_c3(args)
    if args.c4:
        return run_c4(args)
    if args.ref_timing:
        return run_ref_timing(args)

    dev = torch.device("cuda:0")
    ctx = zim_amd.Context(0)
    lib = ctx.lib
    dcfg = scenes.c2_decode_config()

    log = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    log(f"rendering {args.pairs} pairs of {COLS}x{ROWS} on the GPU")
    t0 = time.time()
    pairs = [synth.render_c2(1000 + i, device=dev) for i in range(args.pairs)]
    torch.cuda.synchronize()
    log(f"rendered in {time.time() - t0:.1f}s")
    npix = ROWS * COLS
    t0 = time.time()
    windows, ring_left = mls_windows(
        [dict(left=p.left, right=p.right, decode=dcfg, pha_f=70, Q=p.Q, R1=p.R1) for p in pairs],
        [scenes.c2_window(p, MLS_RADIUS) for p in pairs], ctx)
    rcfgs = [scenes.c2_recon_config(mls_window_size=w) for w in windows]
    log(f"MLS half windows {min(windows)}-{max(windows)}, mean {sum(windows) / len(windows):.2f} "
        f"(chosen in {time.time() - t0:.1f}s; in-radius neighbours left on an outer ring: {ring_left})")

    # ---------------------------------------------------------- batched path
    t0 = time.time()
    batch = zim_amd.Batch([dict(left=p.left, right=p.right, decode=dcfg, recon=rc, pha_f=70,
                                Q=p.Q, R1=p.R1) for p, rc in zip(pairs, rcfgs)], ctx, graph=True,
                          autotune=True)
    torch.cuda.synchronize()
    grids = {s: batch.grid(s) for s in ("decode", "decode_rows", "match", "match_threads",
                                        "median5", "repro_gauss", "sor_knn", "mls",
                                        "mls_variant")}
    log(f"batch planned, autotuned and captured in {time.time() - t0:.1f}s; grids {grids}")
    for w in range(args.warmup):
        batch.run()
        torch.cuda.synchronize()
    log("warmup done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    value = npix * args.pairs / (ms_per_step * 1e-3) / 1e6
    valid = sum(int((o["P"] != 0).sum().item()) for o in batch.outputs)
    log(f"batched: {ms_per_step:.2f} ms/step, {value:.1f} Mpts/s")

    # per-kernel durations: eager replay of the same launches with HIP events
    lib.psgc_profile_enable(1)
    prof_steps = 2
    for _ in range(prof_steps):
        batch.run_eager()
    torch.cuda.synchronize()
    prof = {}
    for fam in list(ALG_BYTES_PER_PX) + ["gray_fix", "ror"]:
        n = C.c_int64()
        ms = C.c_double()
        _capi.check(lib.psgc_profile_read(fam.encode(), C.byref(n), C.byref(ms)))
        if n.value:
            prof[fam] = (n.value, ms.value)
    lib.psgc_profile_enable(0)
    stage_ms = {k: round(v[1] / prof_steps, 3) for k, v in prof.items()}
    dom_name, (dom_n, dom_ms) = max(prof.items(), key=lambda kv: kv[1][1])
    # one pass of a batched stage covers every pair once: one launch, except
    # the MLS, whose iterations are a pass each, launched per window class
    passes = prof_steps * (rcfgs[0].mls_iterations if dom_name == "mls" else 1)
    avg_ms = dom_ms / passes
    launches_per_pass = dom_n / passes
    alg = ALG_BYTES_PER_PX[dom_name] * npix * args.pairs * (2 if dom_name == "decode" else 1)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    compute = None
    vkey = f"mls/{grids['mls_variant']}" if dom_name == "mls" else dom_name
    if vkey in VALU_PER_PX_PROFILED and args.pairs == 64:
        gi = VALU_PER_PX_PROFILED[vkey] / 64 * npix * args.pairs / (avg_ms * 1e-3) / 1e9
        compute = {"bound": "valu", "wave_insts_per_px": round(VALU_PER_PX_PROFILED[vkey] / 64, 1),
                   "achieved_ginst_s": round(gi, 1),
                   "peak_ginst_s": round(VALU_PEAK_GINST_S, 1),
                   "peak_from": "157.3 TFLOP/s FP32 vector / (2 flops per FMA x 64 lanes)",
                   "frac": round(gi / VALU_PEAK_GINST_S, 3),
                   "insts_from": "rocprofv3 SQ_INSTS_VALU on this workload (profile-derived), "
                                 "profiles/c2_64pair_valu_counters.md"}
    log(f"stage ms/step {stage_ms}")
    # the device's measured FMA issue ceilings (FP32 and FP64 chains, no
    # packed instructions), outside the timed region
    c32, c64 = C.c_double(), C.c_double()
    _capi.check(lib.psgc_probe_valu_rate(ctx.h, C.byref(c32), C.byref(c64)))
    fma_ceiling = {"fp32": round(c32.value, 1), "fp64": round(c64.value, 1)}
    log(f"measured FMA issue ceilings (G wave-instructions/s): {fma_ceiling}")
    if compute is not None:
        compute["peak_unpacked_ginst_s"] = round(VALU_UNPACKED_PEAK_GINST_S, 1)
        compute["peak_unpacked_from"] = ("the published peak counts packed FP32 FMAs (2 per lane "
                                         "per instruction); packed FP32 is disabled in this build")
        compute["fma_ceiling_measured_ginst_s"] = fma_ceiling
        compute["frac_of_measured_fp32_fma_ceiling"] = round(compute["achieved_ginst_s"]
                                                             / fma_ceiling["fp32"], 3)

    n_sub = args.plain_pairs if args.plain_pairs else args.pairs

    def agreement(outs):
        """Mask pixels and max |xyz| difference against the batched outputs."""
        mism, maxerr = 0, 0.0
        for o, bo in zip(outs, batch.outputs):
            mism += int((o["P"] != bo["P"]).sum().item())
            m = bo["P"] != 0
            if m.any():
                maxerr = max(maxerr, float((o["xyz"][m] - bo["xyz"][m]).abs().max().item()))
        return mism, maxerr

    # ------------------------------------- tiled kernels, one pair at a time
    per_pair_tiled = None
    if not args.no_per_pair_tiled:
        singles = [zim_amd.Batch([dict(left=p.left, right=p.right, decode=dcfg, recon=rc,
                                       pha_f=70, Q=p.Q, R1=p.R1)], ctx, graph=True)
                   for p, rc in zip(pairs[:n_sub], rcfgs[:n_sub])]

        def run_singles():
            for b in singles:
                b.run()
                torch.cuda.synchronize()

        run_singles()
        t0 = time.perf_counter()
        run_singles()
        single_ms = (time.perf_counter() - t0) * 1e3
        mism, maxerr = agreement([b.outputs[0] for b in singles])
        per_pair_tiled = {"mpts_per_s": round(npix * n_sub / (single_ms * 1e-3) / 1e6, 3),
                          "ms_per_pair": round(single_ms / n_sub, 3), "pairs": n_sub,
                          "mask_mismatch_px": mism, "xyz_max_diff_mm": maxerr}
        for b in singles:
            b.close()
        del singles
        torch.cuda.empty_cache()
        log(f"tiled kernels, one pair at a time: {per_pair_tiled}")

    # ----------------------------------------------------- plain per-pair path
    plain = None
    if not args.no_plain:
        ctx.set_kernels("plain")
        ctx.bind_stream()
        sub = pairs[:n_sub]
        sub_rc = rcfgs[:n_sub]
        bufs = [dict(lph=torch.empty((ROWS, COLS), dtype=torch.float64, device=dev),
                     rph=torch.empty((ROWS, COLS), dtype=torch.float64, device=dev),
                     lk=torch.empty((ROWS, COLS), dtype=torch.int32, device=dev),
                     rk=torch.empty((ROWS, COLS), dtype=torch.int32, device=dev),
                     mod=torch.empty((ROWS, COLS), dtype=torch.float64, device=dev),
                     xyz=torch.empty((ROWS, COLS, 3), dtype=torch.float32, device=dev),
                     P=torch.empty((ROWS, COLS), dtype=torch.uint8, device=dev)) for _ in sub]
        Qa = [np.ascontiguousarray(p.Q, np.float64) for p in sub]
        Ra = [np.ascontiguousarray(p.R1, np.float64) for p in sub]

        def run_plain():
            for p, o, q, r, rcfg in zip(sub, bufs, Qa, Ra, sub_rc):
                for frames, ph, k in ((p.left, o["lph"], o["lk"]), (p.right, o["rph"], o["rk"])):
                    arr = (C.c_void_p * len(frames))(*[f.data_ptr() for f in frames])
                    _capi.check(lib.psgc_decode(ctx.h, C.byref(dcfg), arr, len(frames),
                                                _capi.SAMPLE_U16, ROWS, COLS, COLS * 2,
                                                _capi.ptr(ph), _capi.ptr(k), _capi.ptr(o["mod"]),
                                                None, _capi.MEM_DEVICE))
                _capi.check(lib.psgc_reconstruct(ctx.h, C.byref(rcfg), _capi.ptr(o["lph"]),
                                                 _capi.ptr(o["rph"]), _capi.ptr(o["lk"]),
                                                 _capi.ptr(o["rk"]), ROWS, COLS, 70,
                                                 _capi.ptr(q), _capi.ptr(r), _capi.ptr(o["xyz"]),
                                                 _capi.ptr(o["P"]), None, None,
                                                 _capi.MEM_DEVICE))
                _capi.check(lib.psgc_synchronize(ctx.h))

        run_plain()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_plain()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1e3
        plain_val = npix * len(sub) / (plain_ms * 1e-3) / 1e6
        # the two paths must agree on the pairs both ran
        mism, maxerr = agreement(bufs)
        plain = {"mpts_per_s": round(plain_val, 3), "ms_per_pair": round(plain_ms / len(sub), 3),
                 "pairs": len(sub), "mask_mismatch_px": mism, "xyz_max_diff_mm": maxerr}
        ctx.set_kernels("tiled")
        log(f"plain per-pair: {plain}")

    if args.copy_probe:
        src = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {COPY_PROBE_BYTES} bytes read and written once")
        del src, dst

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, ref = cpu_baseline(pairs[0], dcfg, rcfgs[0], args.cpu_sample_rows, log)
        if ref is not None:
            # the oracle's output for the whole of pair 0 checks the batched one
            rx, rP = ref
            out = batch.outputs[0]
            gP = out["P"].cpu().numpy()
            gx = out["xyz"].cpu().numpy()
            m = rP != 0
            parity = {"pair": 0, "points": int(m.sum()), "mls_half_window": windows[0],
                      "mask_mismatch_px": int((gP != rP).sum()),
                      "xyz_max_diff_mm": float(np.abs(gx[m] - rx[m]).max()) if m.any() else 0.0,
                      "xyz_p99999_diff_mm": float(np.quantile(np.abs(gx[m] - rx[m]).max(-1),
                                                              0.99999)) if m.any() else 0.0}
            log(f"full-pair parity vs oracle: {parity}")

    # HBM bytes per pass of the dominant kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE on this workload (profiles/c2_64pair_traffic.*)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "c2_64pair_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        if (tj.get("workload") == {"pairs": args.pairs, "rows": ROWS, "cols": COLS,
                                   "mls_windows": "per pair"}
                and vkey in tj.get("kernels", {})):
            traffic = int(tj["kernels"][vkey]["traffic_bytes"])
    # whole pipeline: frames of both cameras read once, xyz + mask written once
    e2e_bytes = (24 * 2 + 13) * npix * args.pairs
    e2e_gbs = e2e_bytes / (ms_per_step * 1e-3) / 1e9

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
        "dtype": "f64",
        "data": "synthetic (seeded bump height fields rendered on the GPU)",
        "config": {"workload": f"{args.pairs} seeded 2448x2048 pairs, 3-freq 70/64/59 4-step "
                               "u16 12-bit; median 5x5, Gaussian 5x5 s=1, SOR k=30 (9x9 window), "
                               "MLS order 2 r=1.5 mm (per pair the smallest half window whose "
                               "outer ring holds no neighbour within r; 2 iterations)",
                   "pairs": args.pairs, "rows": ROWS, "cols": COLS, "path": "batched graph",
                   "mls_half_windows": {str(w): windows.count(w) for w in sorted(set(windows))},
                   "mls_ring_points_left": ring_left},
        "valid_mpts_per_s": round(valid / (ms_per_step * 1e-3) / 1e6, 3),
        "per_pair_plain": plain,
        "per_pair_tiled": per_pair_tiled,
        "speedup_vs_per_pair": round(value / plain["mpts_per_s"], 2) if plain else None,
        "batching_gain": (round(value / per_pair_tiled["mpts_per_s"], 2)
                          if per_pair_tiled else None),
        "stage_ms_per_step": stage_ms,
        "autotuned_grids": grids,
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_from": ("profile-derived: rocprofv3 FETCH_SIZE/WRITE_SIZE of "
                                      "this workload, profiles/c2_64pair_traffic.json")
                     if traffic else None,
                     "ms_per_pass": round(avg_ms, 4), "launches_per_pass": launches_per_pass,
                     "alg_bytes_per_pass": alg,
                     "end_to_end": {"alg_bytes_per_pt": 61, "achieved": round(e2e_gbs, 2),
                                    "frac": round(e2e_gbs / HBM_PEAK_GBS, 5)},
                     "compute": compute},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    }
    print(json.dumps(line), flush=True)
    batch.close()


def run_ref_timing(args):
    """The reference's published GPU figure is decode L+R + setData/remap +
    reconstruct (match, median, reprojection, Gaussian, ROR; MLS off) of one
    5120x5120 8-bit pair with 24 frames per camera in about 1.5 s on its GPU
    (JSLAB_Dual_GPU移植说明.md:90-100, 128-133), 17.5 Mpx/s. Here the same
    configuration (synthetic scene; the remap is the identity for a rectified
    pair and is skipped, as the reference's GPU path leaves it on the host) is
    timed one pair at a time: frames resident in HBM with the tiled and with
    the plain kernels, and host-to-host (frames from host memory, every map
    back to the host between decode and reconstruct, as the reference's
    launchers do)."""
    import numpy as np
    import torch

    import zim_amd
    from zim_amd import synth

    log = lambda m: print(f"[bench-ref] {m}", file=sys.stderr, flush=True)  # noqa: E731
    size = 5120
    dev = torch.device("cuda:0")
    ctx = zim_amd.Context(0)
    p = synth.render_ref_timing(3000, device=dev, size=size)
    torch.cuda.synchronize()
    dcfg = zim_amd.DecodeConfig.default(freq=(107, 100, 94), shift_num=(8, 8, 8),
                                        modulation_threshold=0.3)
    rcfg = zim_amd.ReconConfig.default_gpu(use_mls=0, radius_search=0.7,
                                           radius_min_neighbors=10, radius_window_size=10)
    hostL = [f.cpu().numpy() for f in p.left]
    hostR = [f.cpu().numpy() for f in p.right]

    def device_pair():
        L = zim_amd.decode(dcfg, p.left, ctx, want_mod=False, want_valid=False)
        R = zim_amd.decode(dcfg, p.right, ctx, want_mod=False, want_valid=False)
        return zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 107, p.Q, p.R1, ctx)

    def host_pair():
        L = zim_amd.decode(dcfg, hostL, ctx, want_mod=False, want_valid=False)
        R = zim_amd.decode(dcfg, hostR, ctx, want_mod=False, want_valid=False)
        return zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 107, p.Q, p.R1, ctx)

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        return best * 1e3

    res = {}
    res["device_tiled_ms"] = timed(device_pair)
    out = device_pair()
    torch.cuda.synchronize()
    valid = int((out[1] != 0).sum().item())
    ctx.set_kernels("plain")
    res["device_plain_ms"] = timed(device_pair)
    ctx.set_kernels("tiled")
    res["host_to_host_ms"] = timed(host_pair)
    npx = size * size
    ref_mpx = 26.21 / 1.5
    line = {"metric": "reference timing configuration: Mpx/s per pair (5120x5120 u8, 107/100/94 "
                      "x 8 steps, MLS off)",
            "value": round(npx / (res["device_tiled_ms"] * 1e-3) / 1e6, 3), "unit": "Mpx/s",
            "n_gpus": 1, "valid_points": valid,
            **{k: round(v, 3) for k, v in res.items()},
            "mpx_s": {k.replace("_ms", ""): round(npx / (v * 1e-3) / 1e6, 1) for k, v in res.items()},
            "reference_gpu_mpx_s": round(ref_mpx, 2),
            "vs_reference_gpu": {k.replace("_ms", ""): round(npx / (v * 1e-3) / 1e6 / ref_mpx, 1)
                                 for k, v in res.items()},
            "dtype": "f64", "data": "synthetic (seeded bump height field, 8-bit frames)"}
    log(f"times {res}")
    print(json.dumps(line), flush=True)


def run_c4(args):
    """Differentiation at full size: a 1280x1024 pair of a flat plane target
    at 700 mm rendered with the true calibration (f = 1600 px, B = 120 mm,
    three-frequency 4-step fringes, uint16), decoded and matched on the GPU.
    Under a seeded perturbed calibration (focal +-1 %, baseline +-0.5 %,
    principal points +-3 px, rectification rotation +-0.05 deg) the FP64
    triangulation, the plane-fit loss and its adjoint (autograd Function
    over the hand-written kernels) are timed; the gradient is checked
    against central differences (h = 1e-4 in units of those scales) over
    the conditioned points; then a plane refinement runs from the perturbed
    calibration."""
    import numpy as np
    import torch

    import zim_amd
    from zim_amd import autograd as zag
    from zim_amd import synth
    from tests import scenes

    log = lambda m: print(f"[bench-c4] {m}", file=sys.stderr, flush=True)  # noqa: E731
    dev = torch.device("cuda:0")
    ctx = zim_amd.Context(0)
    p = synth.render_c4(4, device=dev)
    dcfg = scenes.c2_decode_config()
    L = zim_amd.decode(dcfg, p.left, ctx)
    R = zim_amd.decode(dcfg, p.right, ctx)
    rcfg = zim_amd.ReconConfig.default_gpu(use_mls=0, outlier_mode=0)
    _, _, _, mc = zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, p.Q, p.R1, ctx)
    phL, phR, mc = L[0].contiguous(), R[0].contiguous(), mc.contiguous()
    cam = p.cam
    theta0 = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy, cam.baseline, 0, 0, 0],
                          dtype=torch.float64)
    nominal = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy], dtype=torch.float64)
    g = np.random.default_rng(7)
    d = np.zeros(9)
    d[0] = cam.f * 0.01 * g.uniform(-1, 1)
    d[1:5] = 3.0 * g.uniform(-1, 1, 4)
    d[5] = cam.baseline * 0.005 * g.uniform(-1, 1)
    d[6:9] = math.radians(0.05) * g.uniform(-1, 1, 3)
    th_p = theta0 + torch.from_numpy(d)
    scale = zag.default_scale(theta0)

    # the loss sums over a fixed point set (as the integer match is fixed):
    # the points conditioned at th_p and triangulated at every step below
    h = 1e-4
    steps = []
    for k in range(9):
        for sgn in (1.0, -1.0):
            e = torch.zeros(9, dtype=torch.float64)
            e[k] = sgn * h * scale[k]
            steps.append(e)
    with torch.no_grad():
        M = zag.point_mask(zag.triangulate(th_p, phL, phR, mc, nominal, ctx)[1])
        for e in steps:
            M = M & (zag.triangulate(th_p + e, phL, phR, mc, nominal, ctx)[1].reshape(-1) != 0)

    def fwd_bwd():
        th = th_p.clone().requires_grad_(True)
        xyz, valid = zag.triangulate(th, phL, phR, mc, nominal, ctx)
        loss = zag.plane_loss(xyz, M)
        loss.backward()
        torch.cuda.synchronize()
        return th.grad, int(M.sum().item())

    fwd_bwd()
    t = time.perf_counter()
    for _ in range(5):
        grad, nvalid = fwd_bwd()
    step_ms = (time.perf_counter() - t) * 1e3 / 5
    grad = grad * scale
    fd = torch.zeros(9, dtype=torch.float64)
    with torch.no_grad():
        for k in range(9):
            lp = zag.plane_loss(zag.triangulate(th_p + steps[2 * k], phL, phR, mc, nominal, ctx)[0], M)
            lm = zag.plane_loss(zag.triangulate(th_p + steps[2 * k + 1], phL, phR, mc, nominal,
                                                ctx)[0], M)
            fd[k] = (lp - lm) / (2 * h)
    rel = (grad - fd).abs() / grad.abs().clamp_min(1e-3 * grad.abs().max())
    rels = {h: rel}
    log(f"h = {h:g}: FD rel err per parameter {[f'{x:.1e}' for x in rel.tolist()]}")
    log(f"forward+backward {step_ms:.2f} ms over {nvalid} valid points")
    t = time.perf_counter()
    theta, hist = zag.refine_plane(th_p, theta0, phL, phR, L[1].contiguous(), R[1].contiguous(),
                                   nominal, 70, outer=4, inner=4, ctx=ctx)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t
    log(f"refinement: flatness {hist[0][0]:.3e} -> {hist[-1][1]:.3e} in {ref_s:.2f} s")
    npx = cam.rows * cam.cols
    line = {"metric": "calibration adjoint: forward + backward ms per 1280x1024 pair, and its "
                      "agreement with central differences",
            "value": round(step_ms, 3), "unit": "ms", "n_gpus": 1, "higher_is_better": False,
            "mpx_per_s": round(npx / (step_ms * 1e-3) / 1e6, 3), "loss_points": nvalid,
            "fd_max_rel_err_by_h": {f"{h:g}": float(r.max()) for h, r in rels.items()},
            "fd_rel_err_by_h": {f"{h:g}": [float(x) for x in r] for h, r in rels.items()},
            "refine": {"outer": 4, "inner": 4, "seconds": round(ref_s, 3),
                       "flatness_start": float(hist[0][0]), "flatness_end": float(hist[-1][1])},
            "dtype": "f64",
            "data": "synthetic (seeded plane target; calibration perturbed with a seeded draw)"}
    print(json.dumps(line), flush=True)


C3_SIZES = ((480, 640), (1024, 1280), (1200, 1920), (2048, 2448))
C3_STEPS = (3, 4, 8)
C3_SCHEDULES = ("heterodyne", "gray")


def run_c3(args):
    """Uneven batch: `--pairs` pairs (default 64), each with a resolution
    drawn uniformly from C3_SIZES, a step count from {3, 4, 8} and a
    three-frequency (70/64/59 heterodyne) or Gray-code (period 16 px plus
    6-bit Gray code with complements) schedule, seeded; every pair is the
    config-1 scene (plane tilted 20 deg at 600 mm plus 3 spheres, f = 800 px
    per 640 columns, B = 100 mm) in uint16 12-bit. Filters as in the
    full-size workload (Gaussian 5x5, SOR k=30, MLS order 2 at 1.5 mm in the
    smallest window that holds it for the pair). value = Mpts/s of the
    autotuned batch (one graph replay per step, points = left pixels of
    every pair). The autotuner also tunes the row match per image width
    (its own launch, grid and threads per row for each width) against one
    launch for all widths; both are timed here. The same pairs run one at a
    time through the same tiled kernels (one graph per pair: what batching
    alone gains) and through the plain per-pair path beside it; their masks
    and xyz must agree with the batch."""
    import numpy as np
    import torch

    import zim_amd
    from zim_amd import synth
    from tests import scenes

    log = lambda m: print(f"[bench-c3] {m}", file=sys.stderr, flush=True)  # noqa: E731
    dev = torch.device("cuda:0")
    ctx = zim_amd.Context(0)
    rng = np.random.default_rng(3)
    specs, npts, starts = [], 0, []
    t0 = time.time()
    for i in range(args.pairs):
        rows, cols = C3_SIZES[rng.integers(len(C3_SIZES))]
        steps = C3_STEPS[rng.integers(len(C3_STEPS))]
        gray = C3_SCHEDULES[rng.integers(len(C3_SCHEDULES))] == "gray"
        p = synth.render_c3(4000 + i, device=dev, rows=rows, cols=cols, steps=steps, gray=gray)
        if gray:
            dcfg, pha_f = scenes.c1_gray_config(shift_num=(steps, 0, 0)), 50
        else:
            dcfg, pha_f = scenes.c2_decode_config(shift_num=(steps,) * 3), 70
        specs.append(dict(left=p.left, right=p.right, decode=dcfg, pha_f=pha_f, Q=p.Q, R1=p.R1,
                          _key=f"{cols}x{rows} {steps}-step {'gray' if gray else 'heterodyne'}"))
        starts.append(scenes.c2_window(p, MLS_RADIUS))
        npts += rows * cols
    torch.cuda.synchronize()
    mix = {}
    for sp in specs:
        mix[sp["_key"]] = mix.get(sp["_key"], 0) + 1
    log(f"rendered {len(specs)} pairs ({npts / 1e6:.1f} Mpts) in {time.time() - t0:.1f}s: {mix}")
    base = [{k: v for k, v in sp.items() if k != "_key"} for sp in specs]
    t0 = time.time()
    windows, ring_left = mls_windows(base, starts, ctx)
    pairs = [dict(sp, recon=scenes.c2_recon_config(mls_window_size=w))
             for sp, w in zip(base, windows)]
    log(f"MLS half windows {min(windows)}-{max(windows)} (chosen in {time.time() - t0:.1f}s; "
        f"in-radius neighbours left on an outer ring: {ring_left})")

    t0 = time.time()
    batch = zim_amd.Batch(pairs, ctx, graph=True, autotune=True)
    torch.cuda.synchronize()
    grids = {s: batch.grid(s) for s in ("decode", "decode_rows", "match", "match_threads",
                                        "match_groups", "median5", "repro_gauss", "sor_knn",
                                        "mls", "mls_variant")}
    widths = sorted({c for _, c in C3_SIZES})
    per_width = {}
    for w in widths:
        try:
            per_width[str(w)] = {"grid": batch.grid(f"match@{w}"),
                                 "threads": batch.grid(f"match_threads@{w}")}
        except zim_amd.PsgcError:
            pass
    log(f"batch planned, autotuned and captured in {time.time() - t0:.1f}s; grids {grids}; "
        f"row match per width {per_width}")

    def timed(steps):
        for _ in range(args.warmup):
            batch.run()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            batch.run()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / steps

    ms_per_step = timed(args.steps)
    value = npts / (ms_per_step * 1e-3) / 1e6
    valid = sum(int((o["P"] != 0).sum().item()) for o in batch.outputs)
    log(f"batched: {ms_per_step:.2f} ms/step, {value:.1f} Mpts/s, {valid} valid points")
    # the other row-match mode, for the record (same outputs)
    match_modes = None
    if per_width:
        chosen = 1 if grids["match_groups"] else 0
        ref = [(o["P"].clone(), o["xyz"].clone()) for o in batch.outputs]
        batch.set_grid("match_groups", 1 - chosen)
        other = timed(max(1, args.steps // 2))
        same = all(torch.equal(P, o["P"]) and torch.equal(x, o["xyz"])
                   for (P, x), o in zip(ref, batch.outputs))
        batch.set_grid("match_groups", chosen)
        match_modes = {"per_width_ms_per_step": round(other if not chosen else ms_per_step, 3),
                       "single_ms_per_step": round(ms_per_step if not chosen else other, 3),
                       "chosen": "per width" if chosen else "single", "outputs_identical": same}
        log(f"row match modes: {match_modes}")
        del ref

    def agreement(outs):
        mism, maxerr = 0, 0.0
        for (xyz, P), bo in zip(outs, batch.outputs):
            mism += int((P != bo["P"]).sum().item())
            m = bo["P"] != 0
            if m.any():
                maxerr = max(maxerr, float((xyz[m] - bo["xyz"][m]).abs().max().item()))
        return mism, maxerr

    # the same tiled kernels one pair at a time (one captured graph per pair,
    # default launch shapes, a sync after each): what batching alone gains
    singles = [zim_amd.Batch([pr], ctx, graph=True) for pr in pairs]

    def run_singles():
        for b in singles:
            b.run()
            torch.cuda.synchronize()

    run_singles()
    t0 = time.perf_counter()
    run_singles()
    single_ms = (time.perf_counter() - t0) * 1e3
    single_val = npts / (single_ms * 1e-3) / 1e6
    mism, maxerr = agreement([(b.outputs[0]["xyz"], b.outputs[0]["P"]) for b in singles])
    per_pair_tiled = {"mpts_per_s": round(single_val, 3), "ms": round(single_ms, 3),
                      "mask_mismatch_px": mism, "xyz_max_diff_mm": maxerr}
    for b in singles:
        b.close()
    del singles
    torch.cuda.empty_cache()
    log(f"tiled kernels one pair at a time: {per_pair_tiled}")

    # the same pairs one at a time through the plain kernels, a sync per pair
    ctx.set_kernels("plain")

    def run_plain():
        outs = []
        for pr in pairs:
            L = zim_amd.decode(pr["decode"], pr["left"], ctx)
            R = zim_amd.decode(pr["decode"], pr["right"], ctx)
            xyz, P, _, _ = zim_amd.reconstruct(pr["recon"], L[0], R[0], L[1], R[1], pr["pha_f"],
                                               pr["Q"], pr["R1"], ctx)
            torch.cuda.synchronize()
            outs.append((xyz, P))
        return outs

    run_plain()
    t0 = time.perf_counter()
    outs = run_plain()
    plain_ms = (time.perf_counter() - t0) * 1e3
    ctx.set_kernels("tiled")
    plain_val = npts / (plain_ms * 1e-3) / 1e6
    mism, maxerr = agreement(outs)
    log(f"plain per-pair: {plain_ms:.1f} ms, {plain_val:.1f} Mpts/s; mask mismatch {mism}, "
        f"xyz max diff {maxerr}")
    line = {"metric": "Mpts/s of one batch of 64 uneven pairs, batched vs per-pair",
            "value": round(value, 3), "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": True, "dtype": "f64",
            "data": "synthetic (seeded plane-sphere scenes rendered on the GPU)",
            "valid_points": valid,
            "config": {"workload": "uneven batch of config-1 scenes, u16 12-bit",
                       "pairs": len(pairs), "points": npts, "mix": mix, "path": "batched graph",
                       "mls_half_windows": {str(w): windows.count(w)
                                            for w in sorted(set(windows))},
                       "mls_ring_points_left": ring_left},
            "row_match_per_width": per_width, "row_match_modes": match_modes,
            "per_pair_plain": {"mpts_per_s": round(plain_val, 3), "ms": round(plain_ms, 3),
                               "mask_mismatch_px": mism, "xyz_max_diff_mm": maxerr},
            "per_pair_tiled": per_pair_tiled,
            "speedup_vs_per_pair": round(value / plain_val, 2),
            "batching_gain": round(value / single_val, 2), "autotuned_grids": grids}
    print(json.dumps(line), flush=True)
    batch.close()


def run_c5(args):
    """Merged-cloud filters: --c5-views views (default 16) of one seeded
    multi-object scene (a plane plus 12 spheres and boxes resting on it),
    seen from ring poses (elevation 55-75 deg, 650-800 mm away), each a
    2448x2048 pair with 70/64/59 4-step fringes in uint16 12-bit,
    reconstructed on the GPU without filters (FP64 triangulation), moved
    into the world frame with each view's pose and merged into one
    unorganized cloud resident in HBM; then SOR (k = 50, exact k nearest),
    MLS (order 2, r = 1.0 mm, sigma 0.5 mm, one iteration) and a Gaussian
    (sigma 0.5 mm, r = 1.0 mm) over it in FP32, each timed with the grid
    build included. The kNN voxel comes from a timing sweep (the autotuner);
    the radius filters run on cells of the radius and of half of it, and
    the faster of the two counts in the total. Before the filters, one step
    of the per-view calibration refinement (batched triangulation, plane-fit
    loss over each view's pixels of the plane, batched adjoint) over the
    views is timed on its own."""
    import torch

    import zim_amd
    from zim_amd import synth
    from tests import scenes

    log = lambda m: print(f"[bench-c5] {m}", file=sys.stderr, flush=True)  # noqa: E731
    dev = torch.device("cuda:0")
    ctx = zim_amd.Context(0)
    dcfg = scenes.c2_decode_config()
    rcfg = scenes.c2_recon_config(use_mls=0, outlier_mode=0)
    t0 = time.time()
    parts, views, plane_px, plane_z = [], [], [], []
    for v in range(args.c5_views):
        p, Rv, Cv, obj = synth.render_c5_view(7, v, device=dev, n_views=args.c5_views)
        L = zim_amd.decode(dcfg, p.left, ctx)
        R = zim_amd.decode(dcfg, p.right, ctx)
        xyz, P, _, mc = zim_amd.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, p.Q, p.R1, ctx)
        m = P != 0
        w = synth.to_world(xyz[m], Rv, Cv)
        parts.append(w)
        plane_z.append(w[(obj[m] == 0)][:, 2])
        views.append((L[0].contiguous(), R[0].contiguous(), mc.contiguous()))
        plane_px.append((obj == 0).reshape(-1))
        cam = p.cam
        del p, L, R, xyz, P, m, obj
    cloud = torch.cat(parts).contiguous()
    pz = torch.cat(plane_z)
    del parts, plane_z
    torch.cuda.synchronize()
    n = cloud.shape[0]
    plane_med = float(pz.abs().median().item())
    log(f"{args.c5_views} views merged: {n} points in {time.time() - t0:.1f}s; "
        f"plane points at world |z| median {plane_med:.3f} mm")
    del pz

    # one step of the per-view calibration refinement: the batched FP64
    # triangulation of every view under its own perturbed calibration
    # (focal +-1 %, baseline +-0.5 %, principal points +-3 px, rotation
    # +-0.05 deg; seeded), a plane-fit loss per view, and the batched adjoint
    from zim_amd import autograd as zag
    g = torch.Generator().manual_seed(5)
    V = len(views)
    th0 = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy, cam.baseline, 0, 0, 0],
                       dtype=torch.float64)
    scale = torch.tensor([0.01 * cam.f, 3, 3, 3, 3, 0.005 * cam.baseline] +
                         [math.radians(0.05)] * 3, dtype=torch.float64)
    thetas = (th0 + (2 * torch.rand(V, 9, generator=g, dtype=torch.float64) - 1) * scale)
    nominal = torch.tensor([cam.f, cam.cx, cam.cy, cam.cx, cam.cy], dtype=torch.float64)

    def refine_step():
        th = thetas.clone().requires_grad_(True)
        outs = zag.triangulate_views(th, [v[0] for v in views], [v[1] for v in views],
                                     [v[2] for v in views], nominal, ctx)
        # each view's plane pixels (the calibration target)
        loss = zag.plane_losses([(x, zag.point_mask(ok) & pm)
                                 for (x, ok), pm in zip(outs, plane_px)]).sum()
        loss.backward()
        torch.cuda.synchronize()
        return th.grad

    refine_step()
    t = time.perf_counter()
    for _ in range(3):
        grad = refine_step()
    adj_ms = (time.perf_counter() - t) * 1e3 / 3
    del views, plane_px
    log(f"batched adjoint step over {V} views: {adj_ms:.1f} ms "
        f"(|grad| finite: {bool(torch.isfinite(grad).all())})")

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        return best * 1e3

    # the library's voxel autotuner for the k-nearest search
    voxel, tune = zim_amd.cloud_tune_voxel(cloud, None, 50, (0.5, 0.7, 1.0, 1.4, 2.0), 4, ctx)
    log(f"kNN voxel sweep (ms): {tune} -> {voxel}")
    res = {}
    res["sor_ms"] = timed(lambda: zim_amd.cloud_sor(cloud, None, 50, 1.0, voxel, 4, ctx))
    keep, md, st = zim_amd.cloud_sor(cloud, None, 50, 1.0, voxel, 4, ctx)
    # the radius filters on cells of the radius (3x3x3 blocks) and of half
    # of it (5x5x5 blocks); the faster counts in the total
    for vx, tag in ((1.0, "r"), (0.5, "half_r")):
        res[f"mls_ms_voxel_{tag}"] = timed(
            lambda: zim_amd.cloud_mls(cloud, keep, 0.5, 2, 1.0, 1, 1, vx, ctx), reps=1)
        res[f"gauss_ms_voxel_{tag}"] = timed(
            lambda: zim_amd.cloud_gaussian(cloud, keep, 0.5, 1.0, vx, ctx))
        log(f"voxel {vx}: mls {res[f'mls_ms_voxel_{tag}']:.1f} ms, "
            f"gauss {res[f'gauss_ms_voxel_{tag}']:.1f} ms")
    res["mls_ms"] = min(res["mls_ms_voxel_r"], res["mls_ms_voxel_half_r"])
    res["gauss_ms"] = min(res["gauss_ms_voxel_r"], res["gauss_ms_voxel_half_r"])
    torch.cuda.synchronize()
    total = res["sor_ms"] + res["mls_ms"] + res["gauss_ms"]
    line = {"metric": "merged-cloud filters Mpts/s (SOR k=50 + MLS order 2 r=1 mm + Gaussian)",
            "value": round(n / (total * 1e-3) / 1e6, 3), "unit": "Mpts/s", "n_gpus": 1,
            "points": n, "views": args.c5_views, "voxel_knn": voxel, "voxel_sweep_ms": tune,
            "kept_after_sor": int(keep.sum().item()), "sor_stats": [float(x) for x in st.cpu()],
            "refine_step_ms": round(adj_ms, 3),
            "refine_step_mpx_per_s": round(V * 2048 * 2448 / (adj_ms * 1e-3) / 1e6, 3),
            **{k: round(v, 3) for k, v in res.items()}, "dtype": "f32",
            "plane_world_z_median_mm": round(plane_med, 4),
            "data": f"synthetic ({args.c5_views} ring views of one seeded plane + 12 spheres "
                    f"and boxes scene, merged in the world frame)"}
    print(json.dumps(line), flush=True)


def mls_windows(specs, starts, ctx, radius=MLS_RADIUS, largest=14):
    """Per pair, the MLS half window that holds the radius: the pipeline runs
    once without MLS (its output is the MLS input cloud), and from the
    depth-based estimate `starts[i]` upward the first half window whose outer
    ring has no valid neighbour within the radius (FP64 distances, a 1e-6
    relative margin over the filter's FP32 test) is taken. specs: Batch pair
    dicts (their recon configs are replaced by the same chain without MLS).
    Returns the windows and the in-radius neighbours left on the outer ring
    of pairs that reached `largest`."""
    import torch
    import zim_amd
    from tests import scenes
    probe = zim_amd.Batch([dict(sp, recon=scenes.c2_recon_config(use_mls=0)) for sp in specs],
                          ctx, graph=False)
    outs = probe.run()
    out, left = [], 0
    for hw, o in zip(starts, outs):
        ring = scenes.ring_points(o["xyz"], o["P"], hw, radius * (1 + 1e-6))
        while ring and hw < largest:
            hw += 1
            ring = scenes.ring_points(o["xyz"], o["P"], hw, radius * (1 + 1e-6))
        out.append(hw)
        left += ring
    probe.close()
    del outs
    torch.cuda.empty_cache()
    return out, left


def cpu_baseline(pair, dcfg, rcfg, sample_rows, log):
    """The CPU restatement of the reference path (oracle/, OpenMP), timed on
    this host's cores over a row band of one pair. For the whole pair it also
    returns the oracle's (xyz, P), which check the batched output."""
    import oracle
    cores = oracle.num_threads()
    left = [f[:sample_rows].cpu().numpy() for f in pair.left]
    right = [f[:sample_rows].cpu().numpy() for f in pair.right]
    log(f"cpu baseline: one pair, rows [0,{sample_rows}), {cores} threads")
    oracle.decode(dcfg, [f[:8] for f in left])  # load + thread pool warm-up
    runs, dec, rec = [], [], []
    for _ in range(5):   # median of five runs of the whole sample
        t0 = time.perf_counter()
        L = oracle.decode(dcfg, left)
        R = oracle.decode(dcfg, right)
        t1 = time.perf_counter()
        rx, rP, _, _ = oracle.reconstruct(rcfg, L[0], R[0], L[1], R[1], 70, pair.Q, pair.R1)
        t2 = time.perf_counter()
        runs.append(t2 - t0)
        dec.append(t1 - t0)
        rec.append(t2 - t1)
    med = lambda v: sorted(v)[len(v) // 2]  # noqa: E731
    dt = med(runs)
    pts = sample_rows * COLS
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    res = {"value": round(pts / dt / 1e6, 4), "unit": "Mpts/s", "cores": cores, "kind": "port",
           "cpu": model,
           # the reference's stage boundaries (decode L+R, reconstruct; the
           # pairs are rendered rectified, so there is no remap stage)
           "stage_s": {"decode_lr": round(med(dec), 3), "reconstruct": round(med(rec), 3)},
           "sample": f"rows [0,{sample_rows}) of one 2448x2048 pair ({pts / 1e6:.2f} Mpts): "
                     f"decode L+R + reconstruct, median of 5 runs {dt:.2f} s "
                     f"(range {min(runs):.2f}-{max(runs):.2f} s)"}
    return res, ((rx, rP) if sample_rows == ROWS else None)


if __name__ == "__main__":
    main()
