"""C ABI checks that need no GPU: the library loads, exports every entry
point include/psgc.h declares, reports a missing device cleanly, and its
structs and defaults match the Python mirror (which mirrors the reference
Config structs)."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import torch

from zim_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(psgc_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_internal_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True).stdout
    syms = [l.split()[-1] for l in out.splitlines() if " T " in l]
    assert all(s.startswith("psgc_") for s in syms), [s for s in syms if not s.startswith("psgc_")]


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-device path")
def test_no_device_is_reported():
    lib = _capi.load()
    assert lib.psgc_device_count() == 0
    h = C.c_void_p()
    rc = lib.psgc_create(C.byref(h), 0)
    assert rc == _capi.PSGC_E_NODEVICE
    assert b"device" in lib.psgc_last_error()
    assert lib.psgc_warmup() == _capi.PSGC_E_NODEVICE


def test_null_context_is_invalid():
    lib = _capi.load()
    cfg = _capi.DecodeConfig.default()
    rc = lib.psgc_decode(None, C.byref(cfg), None, 0, 0, 1, 1, 0, None, None, None, None, 0)
    assert rc == _capi.PSGC_E_INVALID
    rc = lib.psgc_reconstruct(None, None, None, None, None, None, 1, 1, 1, None, None, None,
                              None, None, None, 0)
    assert rc == _capi.PSGC_E_INVALID


def test_struct_defaults_match_reference_configs():
    lib = _capi.load()
    d = _capi.DecodeConfig()
    lib.psgc_decode_config_default(C.byref(d))
    ref = _capi.DecodeConfig.default()
    assert bytes(d) == bytes(ref)
    r = _capi.ReconConfig()
    lib.psgc_recon_config_default(C.byref(r))
    assert bytes(r) == bytes(_capi.ReconConfig.default())
    lib.psgc_recon_config_default_gpu(C.byref(r))
    assert bytes(r) == bytes(_capi.ReconConfig.default_gpu())
    # values of JSLAB_PhaseShiftDecoder.h:30-52 and
    # JSLAB_PhaseEpipolarReconstructor_V2.h:21-54 / _V2_GPU.h:30-60
    assert ref.modulation_threshold == 1.0 and ref.preblur_sigma == 3.0
    rc = _capi.ReconConfig.default()
    assert (rc.grad_threshold, rc.radius_min_neighbors, rc.radius_window_size) == (1.0, 10, 10)
    assert abs(rc.radius_search - 0.7) < 1e-7 and rc.mls_iterations == 3
    rg = _capi.ReconConfig.default_gpu()
    assert (rg.radius_min_neighbors, rg.mls_window_size, rg.mls_iterations) == (5, 25, 2)


def test_rectify_ideal_closed_form():
    lib = _capi.load()
    Q = np.zeros(16)
    R1 = np.zeros(9)
    assert lib.psgc_rectify_ideal(2400.0, 2400.0, 1223.5, 1023.5, 1200.0, 150.0,
                                  _capi.ptr(Q), _capi.ptr(R1)) == 0
    assert np.allclose(R1, np.eye(3).ravel())
    # w = (d - (cx1 - cx2)) / B, Z = f / w
    d = 500.0
    w = Q[14] * d + Q[15]
    assert abs(Q[11] / w - 2400.0 * 150.0 / (d - 23.5)) < 1e-9
    assert lib.psgc_rectify_ideal(0.0, 1.0, 0, 0, 0, 1.0, _capi.ptr(Q), _capi.ptr(R1)) \
        == _capi.PSGC_E_INVALID


def test_header_compiles_as_c_and_links(tmp_path):
    """include/psgc.h is valid C99 and C++17 and links against the library."""
    import subprocess
    inc = os.path.join(ROOT, "include")
    libdir = os.path.join(ROOT, "zim_amd")
    c_src = tmp_path / "t.c"
    c_src.write_text('#include "psgc.h"\n#include <stdio.h>\nint main(void){\n'
                     '  psgc_decode_config d; psgc_decode_config_default(&d);\n'
                     '  printf("%d %d\\n", psgc_abi_version(), d.shift_num[0]);\n'
                     '  return psgc_device_count() < 0; }\n')
    exe = tmp_path / "t"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{inc}", str(c_src), "-o",
                        str(exe), f"-L{libdir}", "-l:libpsgc.so", f"-Wl,-rpath,{libdir}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split() == ["1", "8"], out
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "psgc.h"\nint main(){ return psgc_abi_version() == 1 ? 0 : 1; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{inc}", str(cpp), "-o",
                        str(tmp_path / "tpp"), f"-L{libdir}", "-l:libpsgc.so",
                        f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_write_ply_binary_layout(tmp_path):
    """Binary little-endian PLY as main.cpp:132-162 writes it."""
    lib = _capi.load()
    pts = np.array([[1.5, -2.0, 700.25], [0.0, 1.0, 650.0]], np.float32)
    gray = np.array([10, 200], np.uint8)
    path = str(tmp_path / "a.ply")
    assert lib.psgc_write_ply(path.encode(), _capi.ptr(pts), _capi.ptr(gray), 2) == 0
    data = open(path, "rb").read()
    head, body = data.split(b"end_header\n", 1)
    assert b"format binary_little_endian 1.0" in head and b"element vertex 2" in head
    assert b"property uchar red" in head
    rec = np.frombuffer(body, dtype=np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                                              ("r", "u1"), ("g", "u1"), ("b", "u1")]))
    assert rec["x"].tolist() == [1.5, 0.0] and rec["z"].tolist() == [700.25, 650.0]
    assert rec["g"].tolist() == [10, 200]
    path2 = str(tmp_path / "b.ply")
    assert lib.psgc_write_ply(path2.encode(), _capi.ptr(pts), None, 2) == 0
    body2 = open(path2, "rb").read().split(b"end_header\n", 1)[1]
    assert np.array_equal(np.frombuffer(body2, "<f4").reshape(2, 3), pts)


def test_write_raw_f64_layout(tmp_path):
    """Raw row-major doubles plus the shape file, as saveRaw64F writes them
    (JSLAB_PhaseDisparityExporter.cpp:115-141)."""
    lib = _capi.load()
    m = np.arange(12, dtype=np.float64).reshape(3, 4) * 0.5 - 1.25
    raw, info = tmp_path / "phase.raw", tmp_path / "phase.txt"
    assert lib.psgc_write_raw_f64(str(raw).encode(), str(info).encode(), _capi.ptr(m), 3, 4) == 0
    assert np.array_equal(np.fromfile(raw, "<f8").reshape(3, 4), m)
    lines = info.read_text().splitlines()
    assert lines[:3] == ["rows 3", "cols 4", "type CV_64FC1 (double)"]
    raw2 = tmp_path / "empty.raw"
    assert lib.psgc_write_raw_f64(str(raw2).encode(), None, None, 0, 4) == 0
    assert raw2.stat().st_size == 0
    assert lib.psgc_write_raw_f64(str(raw2).encode(), None, None, 2, 2) != 0


def test_struct_layouts_match_ctypes(tmp_path):
    """The ctypes mirrors in zim_amd/_capi.py have the C structs' sizes and
    field offsets (a drifted field would silently corrupt batch descriptors)."""
    import ctypes as C
    import subprocess
    inc = os.path.join(ROOT, "include")
    fields = {"psgc_decode_config": _capi.DecodeConfig, "psgc_recon_config": _capi.ReconConfig,
              "psgc_pair_desc": _capi.PairDesc, "psgc_plane_view": _capi.PlaneView}
    lines = ['#include "psgc.h"', "#include <stdio.h>", "#include <stddef.h>", "int main(void){"]
    for cname, cls in fields.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0; }")
    src = tmp_path / "l.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "l"
    r = subprocess.run(["gcc", "-std=c99", f"-I{inc}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, cls in fields.items():
        assert got[(cname, "sizeof")] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_load_coef_bin(tmp_path):
    """The calibration coefficient file of JSLAB_StereoCalibParams_V2::
    loadMatBinary: int32 type (CV_32FC4 = 29), rows, cols, row-major data."""
    import zim_amd
    g = np.random.default_rng(0)
    a = g.normal(size=(5, 7, 4)).astype(np.float32)
    path = tmp_path / "c_rz.bin"
    path.write_bytes(np.array([29, 5, 7], "<i4").tobytes() + a.tobytes())
    assert np.array_equal(zim_amd.load_coef_bin(path), a)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(np.array([5, 5, 7], "<i4").tobytes() + a.tobytes())   # CV_32FC1
    with pytest.raises(_capi.PsgcError):
        zim_amd.load_coef_bin(bad)
    short = tmp_path / "short.bin"
    short.write_bytes(np.array([29, 5, 7], "<i4").tobytes() + a.tobytes()[:-4])
    with pytest.raises(_capi.PsgcError):
        zim_amd.load_coef_bin(short)
