"""Write the golden fixtures in tests/golden/ by running the reference itself.

Needs oracle/_ref/ (`make -C oracle ref`, which compiles the reference MMSP
headers and example update() loops where they lie under /root/reference).
Each case writes a seeded input .dat (oracle/dat.py), runs the reference's
update for N steps, and keeps the reference's output .dat. manifest.json
records every case with the SHA-256 of both files.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import dat  # noqa: E402

P, N, M, D = O.PERIODIC, O.NEUMANN, O.MIRROR, O.DIRICHLET

# name, model, shape, fields(Q, 0 = scalar), dtype, bc, dx, seed, (lo, hi), steps, variant, params
CASES = [
    ("ac1d_f64", "ac", (256,), 0, "f64", None, None, 11, (-1.0, 1.0), 100, "stock", []),
    ("ac2d_f64", "ac", (64, 48), 0, "f64", None, None, 101, (-0.01, 0.01), 300, "stock", []),
    ("ac3d_f64", "ac", (12, 10, 16), 0, "f64", None, None, 103, (-1.0, 1.0), 60, "stock", []),
    ("ac2d_bc_f64", "ac", (40, 36), 0, "f64", [(N, M), (D, P)], [0.5, 2.0], 107, (-1.0, 1.0), 40, "stock", []),
    ("ac3d_bc_f64", "ac", (9, 8, 10), 0, "f64", [(M, N), (P, D), (N, M)], None, 109, (-1.0, 1.0), 40, "stock", []),
    ("ac2d_kappa_f64", "ac", (32, 40), 0, "f64", None, None, 113, (-0.01, 0.01), 200, "config",
     [1.0, 1.0, 1.7, 1.0, 0.01]),
    ("ac2d_f32ref", "ac", (48, 40), 0, "f32", None, None, 127, (-1.0, 1.0), 100, "stock", []),
    ("ch2d_stock_f64", "ch", (40, 32), 0, "f64", None, None, 131, (-1.0, 1.0), 50, "stock", []),
    ("ch3d_dw_f64", "ch", (16, 12, 20), 0, "f64", None, None, 137, (0.45, 0.55), 100, "config", [1.0, 1.0, 0.005]),
    ("ch2d_dw_f64", "ch", (48, 40), 0, "f64", None, None, 139, (0.45, 0.55), 200, "config", [1.0, 1.0, 0.005]),
    ("ch2d_dw_bc_f64", "ch", (30, 26), 0, "f64", [(N, M), (M, D)], None, 149, (0.45, 0.55), 60, "config",
     [1.0, 1.0, 0.005]),
    ("ch3d_dw_f32ref", "ch", (10, 12, 14), 0, "f32", None, None, 151, (0.45, 0.55), 40, "config", [1.0, 1.0, 0.005]),
    ("gg1d_f64", "gg", (128,), 2, "f64", None, None, 157, (0.0, 1.0), 100, "stock", []),
    ("gg2d_f64", "gg", (24, 20), 6, "f64", None, None, 163, (0.0, 0.01), 60, "config", [0.05]),
    ("gg2d_q32_f64", "gg", (16, 12), 32, "f64", None, None, 167, (0.0, 0.01), 30, "config", [0.05]),
    ("gg3d_f64", "gg", (10, 8, 12), 3, "f64", None, None, 173, (0.0, 1.0), 30, "stock", []),
    ("gg2d_f32ref", "gg", (24, 20), 6, "f32", None, None, 179, (0.0, 0.01), 60, "config", [0.05]),
    ("gg2d_bc_f64", "gg", (14, 18), 4, "f64", [(N, M), (D, N)], [1.0, 0.5], 181, (0.0, 1.0), 20, "stock", []),
    ("zn2d_f64", "zn", (20, 70), 4, "f64", [(N, M), (P, D)], None, 191, (0.0, 1.0), 30, "stock", []),
    ("zn3d_f64", "zn", (8, 10, 12), 3, "f64", None, None, 193, (0.0, 1.0), 20, "stock", []),
    ("ow1d_f64", "ow", (64,), 2, "f64", [(N, D)], None, 197, (0.0, 1.0), 60, "stock", []),
    ("ow2d_f64", "ow", (24, 20), 3, "f64", None, None, 199, (0.0, 1.0), 60, "stock", []),
    ("ow3d_f64", "ow", (8, 9, 10), 3, "f64", [(M, N), (P, P), (D, P)], [0.5, 1.0, 2.0], 211, (0.0, 1.0), 30, "stock",
     []),
    # multi-phase grain growth from one-hot grains (blocks of 4 nodes per axis): nodes inside a grain
    # keep their values, nodes near boundaries evolve
    ("ga1d_f64", "ga", (64,), 3, "f64", [(N, D)], None, 223, (0.0, 1.0), 40, "stock", []),
    ("ga2d_f64", "ga", (24, 28), 4, "f64", None, None, 227, (0.0, 1.0), 40, "stock", []),
    ("ga3d_f64", "ga", (8, 10, 12), 5, "f64", [(M, N), (P, P), (D, P)], [1.0, 0.5, 2.0], 229, (0.0, 1.0), 20,
     "stock", []),
]


def sha256(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def one_hot_grains(shape, q, seed, block=4):
    """phi = 1 for the grain of each node, 0 for the others; grain ids drawn per block of nodes."""
    blocks = tuple((n + block - 1) // block for n in shape)
    ids = np.floor(O.uniform(int(np.prod(blocks)), seed, 0.0, q)).astype(int).reshape(blocks)
    for ax in range(len(shape)):
        ids = np.repeat(ids, block, axis=ax)
    ids = ids[tuple(slice(0, n) for n in shape)]
    return (ids[..., None] == np.arange(q)).astype(np.float64)


def make_input(shape, q, dtype, seed, lo, hi, model=None):
    if model == "ga":
        return one_hot_grains(shape, q, seed)
    n = int(np.prod(shape)) * max(q, 1)
    a = O.uniform(n, seed, lo, hi)
    a = a.reshape(tuple(shape) + ((q,) if q else ()))
    return a.astype(np.float32) if dtype == "f32" else a


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where /root/reference exists")
    manifest = []
    for (name, model, shape, q, dtype, bc, dx, seed, (lo, hi), steps, variant, params) in CASES:
        a = make_input(shape, q, dtype, seed, lo, hi, model)
        g = dat.make_grid(a, vector=q > 0, bc=bc, dx=dx)
        fin = os.path.join(HERE, name + ".in.dat")
        fout = os.path.join(HERE, name + ".out.dat")
        dat.write_dat(fin, g, compress=True)
        O.ref_run(model, fin, fout, steps, variant, params)
        manifest.append(
            dict(name=name, model=model, shape=list(shape), fields=q, dtype=dtype,
                 bc=[list(b) for b in bc] if bc else None, dx=dx, seed=seed, lo=lo, hi=hi,
                 steps=steps, variant=variant, params=params,
                 sha256_in=sha256(fin), sha256_out=sha256(fout))
        )
        print("wrote", name)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
