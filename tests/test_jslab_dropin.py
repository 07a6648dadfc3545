"""The C++ drop-in (zim_amd/jslab/psgc_jslab.cpp) against the reference's
launcher declarations (JSLAB_PhaseShiftDecoder_V2_GPU.h:50-60,
JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:131-137, JSLAB_MLS_Filter.h:42-47).

No OpenCV here: tests/jslab_mock/ holds a small cv::Mat and the reference
headers' declarations (Config structs with their defaults, the launcher
prototypes, the MLS launcher's default arguments). The glue and a separate
caller translation unit are compiled on their own and linked with
libpsgc.so, as a drop-in build of the reference would link them."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "jslab_mock")
LIBDIR = os.path.join(ROOT, "zim_amd")


def build_caller(tmp_path):
    flags = ["-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{MOCK}",
             f"-I{os.path.join(ROOT, 'include')}"]
    objs = []
    for src in (os.path.join(ROOT, "zim_amd", "jslab", "psgc_jslab.cpp"),
                os.path.join(MOCK, "caller.cpp")):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        r = subprocess.run(["g++", *flags, "-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        objs.append(obj)
    exe = str(tmp_path / "caller")
    r = subprocess.run(["g++", *objs, "-o", exe, f"-L{LIBDIR}", "-l:libpsgc.so",
                        f"-Wl,-rpath,{LIBDIR}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_dropin_compiles_links_and_reports_errors(tmp_path):
    """The glue defines the launchers the caller binds to (a template or
    inline definition would leave them undefined at link time), and a failing
    library call reaches the caller as an exception with the library's
    message (here: no HIP device, unless one is present)."""
    exe = build_caller(tmp_path)
    nm = subprocess.run(["nm", "-C", "--defined-only", str(tmp_path / "psgc_jslab.cpp.o")],
                        capture_output=True, text=True).stdout
    for sig in ("JSLAB_DecodePhase_GPU(std::vector<cv::Mat, std::allocator<cv::Mat> > const&, "
                "JSLAB_PhaseShiftDecoder::Config const&, cv::Mat&, cv::Mat&, cv::Mat&)",
                "JSLAB_DecodePhase_GPU_Warmup()", "JSLAB_DecodePhase_GPU_Release()",
                "JSLAB_ReconstructPipeline_GPU(cv::Mat const&, cv::Mat const&, cv::Mat const&, "
                "cv::Mat const&, int, double, double, double const*, double const*, "
                "JSLAB_PhaseEpipolarReconstructor_V2_GPU::Config const&, cv::Mat&, cv::Mat&, "
                "cv::Mat&)",
                "applyMLSFilterToPointCloud_GPU(cv::Mat const&, cv::Mat const&, cv::Mat&, int, "
                "float, int, float, int)"):
        assert f" T {sig}" in nm, sig
    r = subprocess.run([exe, "probe"], capture_output=True, text=True, timeout=120)
    if r.returncode == 0:
        assert r.stdout.strip() == "ok"
    else:
        assert r.returncode == 3 and r.stdout.startswith("error: psgc:"), (r.returncode, r.stdout)


def _u8(frames):
    """The 12-bit frames scaled into 8 bit (the reference GPU decoder's
    input depth)."""
    return [np.clip(np.round(f.astype(np.float64) * (255.0 / 2047.0)), 0, 255).astype(np.uint8)
            for f in frames]


@pytest.mark.gpu
def test_dropin_launchers_equal_the_oracle(tmp_path):
    """Decode (8-bit frames, padded rows on the left camera), the
    reconstruction launcher and both MLS calls, through the reference's
    declarations, against the oracle with the reference GPU semantics: the
    launcher leaves the post-smooth to its caller, the pipeline has no MLS,
    and the MLS call without its optional arguments is order 0, radius
    3 sigma, one iteration."""
    exe = build_caller(tmp_path)
    p, left, right = scenes.c2_small(3)
    L8, R8 = _u8(left), _u8(right)
    rows, cols = L8[0].shape
    d = tmp_path
    np.concatenate([f.ravel() for f in L8]).tofile(d / "frames_L.u8")
    np.concatenate([f.ravel() for f in R8]).tofile(d / "frames_R.u8")
    np.concatenate([np.ravel(p.Q), np.ravel(p.R1)]).astype(np.float64).tofile(d / "QR.f64")
    thr, r_search, hw_def, hw_full, sigma, order, radius, iters = 0.3, 1.0, 3, 5, 0.5, 2, 1.5, 2
    (d / "meta.txt").write_text(f"{rows} {cols} 12 70 {thr} {r_search} {hw_def} {hw_full} "
                                f"{sigma} {order} {radius} {iters}\n")
    print(f"[dropin] running the caller on {rows}x{cols}", flush=True)
    r = subprocess.run([exe, "run", str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)

    def load(name, dt, shape):
        return np.fromfile(d / name, dt).reshape(shape)

    dcfg = scenes.c2_decode_config(post_smooth=0)
    refs = []
    for s, frames in (("L", L8), ("R", R8)):
        ref = oracle.decode(dcfg, frames)
        refs.append(ref)
        assert np.array_equal(load(f"k_{s}.i32", np.int32, (rows, cols)), ref[1])
        assert np.abs(load(f"phase_{s}.f64", np.float64, (rows, cols)) - ref[0]).max() <= 1e-5
        assert np.abs(load(f"mod_{s}.f64", np.float64, (rows, cols)) - ref[2]).max() <= 1e-9
    (Lp, Lk), (Rp, Rk) = (refs[0][0], refs[0][1]), (refs[1][0], refs[1][1])
    rc = scenes.ReconConfig.default_gpu(use_mls=0, radius_search=r_search)
    rx, rP, rdisp, _ = oracle.reconstruct(rc, Lp, Rp, Lk, Rk, 70, p.Q, p.R1)
    gP = load("P.u8", np.uint8, (rows, cols))
    gx = load("xyz.f32", np.float32, (rows, cols, 3))
    assert rP.sum() > 0.2 * rows * cols
    assert np.array_equal(gP, rP)
    assert np.abs(gx - rx).max() <= 1e-3
    assert np.abs(load("disp.f32", np.float32, (rows, cols)) - rdisp).max() <= 1e-3
    ma = oracle.mls(rx, rP, hw_def, sigma, 0, 0.0, 1, 1)
    mb = oracle.mls(rx, rP, hw_full, sigma, order, radius, iters, 1)
    assert np.abs(load("mls_default.f32", np.float32, (rows, cols, 3)) - ma).max() <= 1e-3
    assert np.abs(load("mls_full.f32", np.float32, (rows, cols, 3)) - mb).max() <= 1e-3
    assert np.abs(mb - rx).max() > 1e-3  # MLS moved the points
