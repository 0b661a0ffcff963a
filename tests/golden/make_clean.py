"""Writes tests/golden/clean_blobs.npz: the flags of the reference's clean_hex_mesh passes
(tagging_uneven_element, clean_non_manifold_ve, drop_small_pieces; grid_meshing/
grid_hex_meshing.cpp:1983-2124, compiled from the reference sources into oracle/_ref/clean_ref
by oracle/ref/Makefile) on 40 seeded random blobs of 6^3 to 14^3 cell grids: a smoothed random
field thresholded to the inside flags, half of the meshes with their vertices and hexes
renumbered at random. The flags after each pass are stored with the inputs, and the
medial-surface tags of the final flags (:1966-1981): vertex flags and the tagged faces.

Run in the development container:  python tests/golden/make_clean.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import refio  # noqa: E402

BIN = os.path.join(REPO, "oracle", "_ref", "clean_ref")
CORN = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def grid(n):
    v = n + 1
    c = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")[::-1], -1).reshape(-1, 3)
    p = c[:, None, :] + CORN[None]
    return (p[..., 0] + v * (p[..., 1] + v * p[..., 2])).astype(np.int32), v ** 3


def blob(n, rng):
    f = rng.random((n + 2, n + 2, n + 2))
    for _ in range(2):     # box smoothing
        f = (f + np.roll(f, 1, 0) + np.roll(f, -1, 0) + np.roll(f, 1, 1) + np.roll(f, -1, 1) + np.roll(f, 1, 2) +
             np.roll(f, -1, 2)) / 7
    f = f[1:-1, 1:-1, 1:-1]
    return (f > np.quantile(f, rng.uniform(0.35, 0.6))).reshape(-1)   # z-major, x fastest


def run(nV, H, flag):
    with tempfile.TemporaryDirectory() as td:
        pin, pout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(pin, "wb") as f:
            f.write(struct.pack("<ii", nV, H.shape[0]))
            f.write(np.ascontiguousarray(H, dtype="<i4").tobytes())
            f.write(np.ascontiguousarray(flag, dtype="<i4").tobytes())
        subprocess.run([BIN, pin, pout], check=True, stdout=subprocess.DEVNULL)
        r = refio.read_out(pout)
    out = {k: v.astype(bool) for k, v in r.items() if k != "medial_f"}
    out["medial_f"] = r["medial_f"].astype(np.int32).reshape(-1, 4)
    return out


def main():
    rng = np.random.default_rng(5)
    out = {}
    for i in range(40):
        n = int(rng.integers(6, 15))
        H, nV = grid(n)
        flag = blob(n, rng)
        if i % 2:
            vperm = rng.permutation(nV)
            hperm = rng.permutation(H.shape[0])
            H = vperm[H][hperm].astype(np.int32)
            flag = flag[hperm]
        r = run(nV, H, flag)
        out["m%d_nV" % i] = nV
        out["m%d_H" % i] = H
        out["m%d_in" % i] = flag
        for k in ("tag", "manifold", "final", "medial_v", "medial_f"):
            out["m%d_%s" % (i, k)] = r[k]
    out["count"] = 40
    path = os.path.join(HERE, "clean_blobs.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
