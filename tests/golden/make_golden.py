"""Writes the golden fixtures under tests/golden/ from the reference's own
src/polesitter.h, compiled by oracle/Makefile into oracle/_ref/.

    make -C oracle && python tests/golden/make_golden.py

Each fixture is a compressed .npz holding the inputs, the outputs of one
public-API tick (ps_prepare_particles + ps_calc_forces, root = min + range/2,
as tests/accuracy.c:493-503 calls it) and the reference's octree dumped in
breadth-first, child-slot order (= canonical level order):
    in_x, in_y, in_z, in_m                    inputs
    prep_codes, min, max, range               ps_prepare_particles
    x, y, z, m, id, codes, fx, fy, fz         after ps_calc_forces (sorted)
    t_depth, t_leaf, t_first, t_cnt, t_parent, t_geom, t_mp, t_loc
Builds: strict_mt / strict_st = -O2 -ffp-contract=off -fno-fast-math
-march=x86-64-v2 (scalar paths) with / without PS_MULTITHREADING;
shipped_mt = -O3 -ffast-math -march=x86-64-v3 (AVX2 paths).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from polesitter_amd import workloads as W  # noqa: E402

THETA = 2.0
THREADS = 4

CASES = {
    # name: (generator, builds)
    "uniform512": (lambda: W.ref_fill_uniform(512), ("strict_mt", "strict_st", "shipped_mt")),
    "clusters512": (lambda: W.ref_fill_two_clusters(512), ("strict_mt", "strict_st", "shipped_mt")),
    "c1_1024": (lambda: W.uniform_cube(1024, seed=1), ("strict_mt", "strict_st", "shipped_mt")),
    "plummer2048": (lambda: W.plummer(2048, seed=11), ("strict_mt",)),
    "dupes300": (lambda: _dupes(300), ("strict_mt",)),
}


def _dupes(n):
    """Many particles on a few exact positions: deep depth-10 leaves."""
    x, y, z, m = W.uniform_cube(n, seed=3)
    k = np.arange(n) % 7
    x, y, z = x[k].copy(), y[k].copy(), z[k].copy()
    x[::5] += np.float32(1e-4)
    return x, y, z, m


def make(name, gen, build):
    x, y, z, m = gen()
    r = O.ref_tick(x, y, z, m, theta=THETA, thrd_cnt=THREADS, build=build, dump=True)
    codes, lo, hi, rg = O.ref_prepare(x, y, z, build=build)
    t = r["tree"]
    out = dict(in_x=x, in_y=y, in_z=z, in_m=m, prep_codes=codes,
               min=np.float32(lo), max=np.float32(hi), range=np.float32(rg),
               theta=np.float32(THETA),
               x=r["x"], y=r["y"], z=r["z"], m=r["m"], id=r["id"], codes=r["codes"],
               fx=r["fx"], fy=r["fy"], fz=r["fz"],
               t_depth=t["depth"], t_leaf=t["leaf"], t_first=t["first"], t_cnt=t["cnt"],
               t_parent=t["parent"], t_geom=t["geom"], t_mp=t["mp"], t_loc=t["loc"])
    path = os.path.join(HERE, f"{name}_{build}.npz")
    np.savez_compressed(path, **out)
    return path, t["n_nodes"]


def main():
    for b in ("strict_mt", "strict_st", "shipped_mt"):
        if not O.ref_available(b):
            raise SystemExit(f"oracle/_ref/libref_{b}.so missing: run make -C oracle")
    for name, (gen, builds) in CASES.items():
        for b in builds:
            path, nn = make(name, gen, b)
            print(f"{os.path.basename(path)}: {nn} nodes, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
