"""Write tests/golden/fullsize.json: the oracle's states at the benchmark
workloads' full sizes, as SHA-256 digests the GPU tests compare against.

The states are too large to commit (128 MiB to 1 GiB each), so only their
digests and a few scalars travel. Each run here uses the oracle's C
restatement (oracle/liboracle.so, OpenMP node loops with the same bits at any
thread count); the restatement is pinned bitwise to the reference's own
update() by the golden cases in manifest.json.

Workloads (the seeded inputs bench.py and the tests draw; SplitMix64):
  ch3d  3-D Cahn-Hilliard, double-well mu, 256^3 periodic, c0 = U(0.45, 0.55)
        seed 20260102, kappa = M = 1, dt = 0.005, 2000 steps: the FP64 state,
        the FP64 state from the float-rounded start and the FP32 state
        (cahn-hilliard.cpp:65-80 structure)
  gg2d  2-D grain growth, 2048^2 x 32 fields periodic, phi0 = U(0, 0.01)
        seed 20260103 rounded to float, dt = 0.05, 500 steps: FP64 with the
        correctly rounded cube (the device's), FP64 with libm pow (the
        reference's arithmetic) and FP32 (graingrowth.cpp:84-108)
  batch_ac  64 Allen-Cahn members, sides 128 << floor(U(0,5)) seed 20260104,
        kappa = U(0.5, 2) seed 20260105 (seed+1), member m: phi0 = U(-0.01,
        0.01) seed 20260105 + m rounded to float, dt = 0.01, 1000 steps: FP32
        and FP64 (from the float-rounded start) per member
  ac2d  2-D Allen-Cahn 512^2, phi0 = U(-0.01, 0.01) seed 20260101, 1000
        steps: FP64 with the correctly rounded cube and with libm pow
  ac3d  3-D Allen-Cahn 256^3 periodic, phi0 = U(-0.01, 0.01) seed 20260112,
        1000 steps: FP64 with the correctly rounded cube and with libm pow
        (allen-cahn.cpp:64-73)
  gg3d  3-D grain growth 256^3 x 8 fields periodic, phi0 = U(0, 0.01) seed
        20260113 rounded to float, dt = 0.05, 200 steps: FP32 and FP64 (from
        the float-rounded start, correctly rounded cube)
  ostwald3d  3-D Ostwald ripening 256^3 x 2 fields periodic, U(0, 1) seed
        20260114, the example's constants, 200 steps: FP64 (ostwald.cpp:87-122)

Digest convention: SHA-256 of the state's bytes, node-major (the reference's
node order; vector grids: each node's fields contiguous), little-endian.

    python tests/golden/make_fullsize.py [--only ch3d,gg2d,batch_ac,ac2d,ac3d,gg3d,ostwald3d]

Takes about an hour on 8 cores. Errors are normwise per field:
max_i |x_i - ref_i| / max_i |ref_i|.
"""
from __future__ import annotations

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

import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "fullsize.json")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def plane_digests(a: np.ndarray) -> list:
    """Short digests of each axis-0 plane (to locate a difference)."""
    return [hashlib.sha256(np.ascontiguousarray(a[i]).tobytes()).hexdigest()[:16] for i in range(a.shape[0])]


def normwise(x: np.ndarray, ref: np.ndarray, vector: bool) -> list:
    x = x.astype(np.float64)
    if vector:
        q = ref.shape[-1]
        return [float(np.abs(x[..., j] - ref[..., j]).max() / np.abs(ref[..., j]).max()) for j in range(q)]
    return [float(np.abs(x - ref).max() / np.abs(ref).max())]


def log(*a):
    print(time.strftime("%H:%M:%S"), *a, flush=True)


def ch3d():
    shape, steps, params = (256, 256, 256), 2000, [1.0, 1.0, 0.005]
    c0 = O.uniform(int(np.prod(shape)), 20260102, 0.45, 0.55).reshape(shape)
    log("ch3d f64")
    f64 = O.ch(c0, params, steps, mu=O.MU_DW)
    m0, m1 = O.mass(c0)[0], O.mass(f64)[0]
    r32 = c0.astype(np.float32)
    log("ch3d f64 from the float start")
    f64r = O.ch(r32.astype(np.float64), params, steps, mu=O.MU_DW)
    log("ch3d f32")
    f32 = O.ch(r32, params, steps, mu=O.MU_DW, form="f32")
    return dict(shape=list(shape), seed=20260102, lo=0.45, hi=0.55, params=params, steps=steps,
                f64=dict(sha256=digest(f64), planes=plane_digests(f64), mass0=m0, mass1=m1,
                         mass_rel_change=abs(m1 - m0) / abs(m0), maxabs=float(np.abs(f64).max())),
                f64_from_f32_start=dict(sha256=digest(f64r)),
                f32=dict(sha256=digest(f32), planes=plane_digests(f32),
                         err_vs_f64=normwise(f32, f64r, False)[0]))


def gg2d():
    shape, q, steps, dt = (2048, 2048), 32, 500, 0.05
    a0 = O.uniform(int(np.prod(shape)) * q, 20260103, 0.0, 0.01).astype(np.float32).reshape(shape + (q,))
    log("gg2d f64 (correctly rounded cube)")
    cr = O.gg(a0.astype(np.float64), dt, steps, cube=O.CUBE_CR)
    log("gg2d f64 (libm pow)")
    pw = O.gg(a0.astype(np.float64), dt, steps, cube=O.CUBE_POW)
    log("gg2d f32")
    f32 = O.gg(a0, dt, steps, form="f32")
    return dict(shape=list(shape), fields=q, seed=20260103, lo=0.0, hi=0.01, dt=dt, steps=steps,
                f64=dict(sha256=digest(cr), planes=plane_digests(cr),
                         err_vs_reference_pow=normwise(cr, pw, True)),
                f64_pow=dict(sha256=digest(pw)),
                f32=dict(sha256=digest(f32), planes=plane_digests(f32), err_vs_f64=normwise(f32, cr, True)))


def batch_ac():
    count, steps = 64, 1000
    pick = O.uniform(count, 20260104, 0.0, 5.0)
    kap = O.uniform(count, 20260105, 0.5, 2.0)
    members = []
    for m in range(count):
        side = 128 << int(pick[m])
        p = [1.0, 1.0, float(kap[m]), 1.0, 0.01]
        a0 = O.uniform(side * side, 20260105 + m, -0.01, 0.01).astype(np.float32).reshape(side, side)
        log("batch_ac member", m, side)
        f32 = O.ac(a0, p, steps, form="f32")
        f64 = O.ac(a0.astype(np.float64), p, steps)
        members.append(dict(side=side, kappa=float(kap[m]), seed=20260105 + m, f32=digest(f32), f64=digest(f64),
                            err_vs_f64=normwise(f32, f64, False)[0]))
    return dict(count=count, steps=steps, params="[1, 1, kappa, 1, 0.01]", lo=-0.01, hi=0.01, members=members)


def ac2d():
    shape, steps, p = (512, 512), 1000, [1.0, 1.0, 1.0, 1.0, 0.01]
    a0 = O.uniform(int(np.prod(shape)), 20260101, -0.01, 0.01).reshape(shape)
    cr = O.ac(a0, p, steps, cube=O.CUBE_CR)
    pw = O.ac(a0, p, steps, cube=O.CUBE_POW)
    return dict(shape=list(shape), seed=20260101, lo=-0.01, hi=0.01, params=p, steps=steps,
                f64=dict(sha256=digest(cr), err_vs_reference_pow=normwise(cr, pw, False)[0]),
                f64_pow=dict(sha256=digest(pw)))


def ac3d():
    shape, steps, p = (256, 256, 256), 1000, [1.0, 1.0, 1.0, 1.0, 0.01]
    a0 = O.uniform(int(np.prod(shape)), 20260112, -0.01, 0.01).reshape(shape)
    log("ac3d f64 (correctly rounded cube)")
    cr = O.ac(a0, p, steps, cube=O.CUBE_CR)
    log("ac3d f64 (libm pow)")
    pw = O.ac(a0, p, steps, cube=O.CUBE_POW)
    return dict(shape=list(shape), seed=20260112, lo=-0.01, hi=0.01, params=p, steps=steps,
                f64=dict(sha256=digest(cr), planes=plane_digests(cr), err_vs_reference_pow=normwise(cr, pw, False)[0]),
                f64_pow=dict(sha256=digest(pw)))


def gg3d():
    shape, q, steps, dt = (256, 256, 256), 8, 200, 0.05
    a0 = O.uniform(int(np.prod(shape)) * q, 20260113, 0.0, 0.01).astype(np.float32).reshape(shape + (q,))
    log("gg3d f32")
    f32 = O.gg(a0, dt, steps, form="f32")
    log("gg3d f64 from the float start (correctly rounded cube)")
    cr = O.gg(a0.astype(np.float64), dt, steps, cube=O.CUBE_CR)
    return dict(shape=list(shape), fields=q, seed=20260113, lo=0.0, hi=0.01, dt=dt, steps=steps,
                f64=dict(sha256=digest(cr), planes=plane_digests(cr)),
                f32=dict(sha256=digest(f32), planes=plane_digests(f32), err_vs_f64=normwise(f32, cr, True)))


def ostwald3d():
    shape, q, steps = (256, 256, 256), 2, 200
    p = [0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0]
    a0 = O.uniform(int(np.prod(shape)) * q, 20260114, 0.0, 1.0).reshape(shape + (q,))
    log("ostwald3d f64")
    f64 = O.ostwald(a0, p, steps)
    return dict(shape=list(shape), fields=q, seed=20260114, lo=0.0, hi=1.0, params=p, steps=steps,
                f64=dict(sha256=digest(f64), planes=plane_digests(f64), maxabs=float(np.abs(f64).max())))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="ac2d,ch3d,gg2d,batch_ac")
    args = ap.parse_args()
    O.set_threads(os.cpu_count() or 1)
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    for name in args.only.split(","):
        t0 = time.time()
        out[name] = globals()[name]()
        log(name, "done in %.0f s" % (time.time() - t0))
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
