"""Writes the golden fixtures (inputs and expected outputs) of the
reconstruction path, computed by the CPU restatement in oracle/:

    python tests/golden/make_golden.py

c2_tiny.npz  48x64 three-frequency u16 pair -> decode maps, reconstruction
c1_tiny.npz  60x80 Gray-code u8 pair        -> decode maps
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests import scenes  # noqa: E402
from zim_amd import synth  # noqa: E402


def main():
    p = synth.render_c2(77, rows=48, cols=64, f=2400.0 * 64 / 2448.0)
    left = np.stack([f.numpy() for f in p.left])
    right = np.stack([f.numpy() for f in p.right])
    dc = scenes.c2_decode_config()
    L = oracle.decode(dc, list(left))
    R = oracle.decode(dc, list(right))
    rc = scenes.c2_recon_config(mls_window_size=3)
    xyz, P, disp, mc = oracle.reconstruct(rc, L[0], R[0], L[1], R[1], 70, p.Q, p.R1)
    np.savez_compressed(os.path.join(HERE, "c2_tiny.npz"), left=left, right=right,
                        Q=np.array(p.Q), R1=np.array(p.R1),
                        l_phase=L[0], l_k=L[1], l_mod=L[2], l_valid=L[3],
                        r_phase=R[0], r_k=R[1], xyz=xyz, P=P, disp=disp, match_col=mc)
    p1 = synth.render_c1(78, rows=60, cols=80)
    left1 = np.stack([f.numpy() for f in p1.left])
    g = oracle.decode(scenes.c1_gray_config(), list(left1))
    np.savez_compressed(os.path.join(HERE, "c1_tiny.npz"), left=left1, phase=g[0], k=g[1],
                        mod=g[2], valid=g[3])
    print("wrote c2_tiny.npz, c1_tiny.npz")


if __name__ == "__main__":
    main()
