"""Writes the workload fixtures of the parity tests: for every benchmark
configuration, the SHA-256 of its seeded synthetic image(s), the SHA-256 of
the oracle's node maps, and effective tensors solved by the CPU FP64 oracle
(oracle/chfem_oracle.c) at a tight tolerance.

    python tests/golden/make_configs.py --parts c1,c2,c3,c4,c5
    python tests/golden/make_configs.py --parts headline --dirs 0,3

Each part writes tests/golden/config_<part>.json (merged with an existing
file, so load cases of a long part can be added in separate runs).  The
oracle is the checker here; the GPU tests compare the HIP library against
these files and check the image hashes first, so a fixture only applies to
the image it was made from.

Parts:
  c1        32^3 thresholded Gaussian random field (sigma 2, volume fraction
            0.4, seed 1), conductivity k 1 / 10, 3 load cases, tol 1e-10
  c2        128^3 overlapping spheres (radius 6-12, fraction 0.3, seed 2),
            E 1 / 10, nu 0.3 / 0.2, 6 load cases, tol 1e-9; NBN key hash
  headline  256^3 spheres of the same recipe, seed 3; NBN key hash; the load
            cases named by --dirs, tol 1e-8
  c3        256^3 porous spheres (solid radius 6-12, porosity 0.35, seed 4
            bumped until the pore space percolates): fluid key, DOF id,
            DOF->node and border key hashes, velocity and border node counts
  c4        64 samples (seed 100 + s, edges in [48, 160]; even thermal GRF
            k 1 / 50, odd elastic spheres E 1 / 20): image hashes of all,
            effective tensors of the 4 smallest samples of each kind at
            tol 1e-10
  c5        512^3 spheres, seed 5: image hash
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from chfem_amd import synthetic  # noqa: E402
from chfem_amd.batchbench import make_samples  # noqa: E402
from oracle import oracle as O  # noqa: E402

ELASTIC_PROPS = [1.0, 0.3, 10.0, 0.2]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def path(part):
    return os.path.join(HERE, f"config_{part}.json")


def load(part):
    p = path(part)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return {}


def save(part, d):
    with open(path(part), "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {path(part)}", flush=True)


def solve_dirs(m, dirs, tol, d, log):
    """Solves the listed load cases with the oracle's FP64 Jacobi-PCG and
    stores the C columns (C[:, dir]) and iteration counts in d."""
    cols = d.setdefault("columns", {})
    for k in dirs:
        t0 = time.time()
        x, ok, it, res = m.solve(k, tol=tol, maxit=100000)
        assert ok, (k, it, res)
        C = m.update_c(k, x).reshape(m.cdim, m.cdim)
        cols[str(k)] = {"column": C[:, k].tolist(), "iterations": it, "residual": res,
                        "seconds": time.time() - t0}
        log(f"  load {k}: {it} iterations, residual {res:.2e}, {time.time() - t0:.1f} s")


def part_c1(log):
    img = synthetic.grf((32, 32, 32), sigma=2.0, vf=0.4, seed=1)
    m = O.Model(O.THERMAL, (32, 32, 32), synthetic.elem_map(img, [0, 255]), [1.0, 10.0])
    d = {"workload": "thermal 32^3 GRF sigma 2 vf 0.4 seed 1, k 1 (colour 0) / 10 (colour 255)",
         "image_sha256": sha(img), "shape": list(img.shape), "tol": 1e-10, "keys_sha256": sha(m.keys())}
    solve_dirs(m, range(3), 1e-10, d, log)
    return d


def part_spheres(n, seed, dirs, tol, log, d):
    img = synthetic.spheres((n, n, n), rmin=6, rmax=12, fraction=0.3, seed=seed)
    h = sha(img)
    if d.get("image_sha256") not in (None, h):
        raise RuntimeError("image hash changed: remove the old fixture first")
    m = O.Model(O.ELASTIC, (n, n, n), synthetic.elem_map(img, [0, 255]), ELASTIC_PROPS)
    d.update({"workload": f"elastic {n}^3 spheres r 6-12 fraction 0.3 seed {seed}, "
                          "E/nu 1/0.3 (colour 0) and 10/0.2 (colour 255)",
              "image_sha256": h, "shape": [n, n, n], "tol": tol, "keys_sha256": sha(m.keys()),
              "elem_sha256": sha(m.elem)})
    solve_dirs(m, dirs, tol, d, log)
    return d


def part_c3(log):
    img, seed = synthetic.porous_spheres((256, 256, 256), rmin=6, rmax=12, porosity=0.35, seed=4)
    m = O.Model(O.FLUID, (256, 256, 256), synthetic.elem_map(img, [0, 255]), [0.0])
    fk, ids, d2n, bk = m.fluid_maps()
    return {"workload": "Stokes 256^3 porous spheres r 6-12 porosity <= 0.35, pore colour 0",
            "seed": seed, "image_sha256": sha(img), "shape": [256, 256, 256],
            "porosity": float((img == 0).mean()), "nvel": int(m.nvel), "nborder": int(m.nborder),
            "ndof": int(m.ndof), "fluid_keys_sha256": sha(fk), "dof_id_sha256": sha(ids),
            "dof2node_sha256": sha(d2n), "border_keys_sha256": sha(bk)}


def part_c4(log):
    samples = make_samples(64, 48, 160)
    out = {"workload": "64 samples, seed 100+s, edges uniform in [48,160]; even: thermal GRF vf 0.4 "
                       "k 1/50; odd: elastic spheres E/nu 1/0.3 and 20/0.2",
           "image_sha256": [sha(img) for _, img, _ in samples],
           "shapes": [list(img.shape) for _, img, _ in samples], "tol": 1e-10, "solved": {}}
    for prop in ("conductivity", "elasticity"):
        idx = [k for k, (p, _, _) in enumerate(samples) if p == prop]
        idx = sorted(idx, key=lambda k: samples[k][1].size)[:4]
        for k in idx:
            _, img, mp = samples[k]
            nz, ny, nx = img.shape
            em = synthetic.elem_map(img, [0, 255])
            if prop == "conductivity":
                m = O.Model(O.THERMAL, (nx, ny, nz), em, [mp[0], mp[255]])
            else:
                m = O.Model(O.ELASTIC, (nx, ny, nz), em, [*mp[0], *mp[255]])
            t0 = time.time()
            C, info = m.homogenize(tol=1e-10, maxit=100000)
            assert all(v[0] for v in info.values())
            out["solved"][str(k)] = {"C": C.tolist(), "iterations": [info[j][1] for j in sorted(info)]}
            log(f"  sample {k} ({prop}, {img.shape}): {time.time() - t0:.1f} s")
    return out


def part_c5(log):
    img = synthetic.spheres((512, 512, 512), rmin=6, rmax=12, fraction=0.3, seed=5)
    return {"workload": "elastic 512^3 spheres r 6-12 fraction 0.3 seed 5", "image_sha256": sha(img),
            "shape": [512, 512, 512]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--parts", default="c1,c2,c3,c4,c5")
    ap.add_argument("--dirs", default="0,3", help="headline load cases to solve")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    if a.threads:
        O.lib().orc_set_threads(a.threads)

    def log(msg):
        print(msg, flush=True)

    for part in a.parts.split(","):
        t0 = time.time()
        log(f"[{part}]")
        if part == "c1":
            d = part_c1(log)
        elif part == "c2":
            d = part_spheres(128, 2, range(6), 1e-9, log, load("c2"))
        elif part == "headline":
            d = part_spheres(256, 3, [int(x) for x in a.dirs.split(",")], 1e-8, log, load("headline"))
        elif part == "c3":
            d = part_c3(log)
        elif part == "c4":
            d = part_c4(log)
        elif part == "c5":
            d = part_c5(log)
        else:
            raise SystemExit(f"unknown part {part}")
        d["oracle_threads"] = int(O.lib().orc_num_threads())
        save(part, d)
        log(f"[{part}] {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
