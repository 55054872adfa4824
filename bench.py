"""Benchmark of the explicit update path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload ch3d|ch3d_f32|ac2d|ac3d|gg2d|gg3d|zener2d|ostwald2d|gganiso2d|model_b3d|batch_ac|serve|
                                ckpt|dropin]
                    [--no-cpu-baseline] [--tune | --tile TY,TZ,CHUNK[,OPT]]
                    [--requests N] [--hbm-budget-gb G]
                    [--graph-steps G] [--no-sequential] [--ckpt-threads T]

--tune runs the autotuner before timing instead of using the default tile;
--tile fixes the fused kernel's tile. --requests and --hbm-budget-gb size
the served stream. --graph-steps and --no-sequential shape batch_ac runs
made for profiling (direct launches; no FP64 one-after-another comparison).
ckpt runs the 256^3 CH loop with a .dat checkpoint every half of the steps,
serial writer against overlapped checkpoints on --ckpt-threads threads.
dropin runs the reference's own driver program with the GPU update() on the
256^3 CH grid, --increment steps per update() call, end to end.
serve routes requests of more than --fp64-above steps to the FP64 form.

Default workload: 3-D Cahn-Hilliard, grid<3,double> 256^3 periodic,
mu = 2c(1-c)(1-2c) - kappa*lap(c), dc/dt = M*lap(mu), kappa = M = dx = 1,
dt = 0.005, c0 = 0.5 + 0.05*U(-1,1) (SplitMix64 seed 20260102), FP64,
2000 steps (one step = one explicit time step over the whole grid, the
state resident in HBM). Prints ONE JSON line.

value = cell-updates/s over the timed steps (one node advanced one step).
roofline.achieved = algorithmic bytes per launch / mean device time of one
launch, timed live with a HIP event pair around every step on the stream the
kernel runs on; algorithmic bytes per cell-update: 16 for FP64 CH/AC (read c,
write c'), 8 per field for FP32 forms.
cpu_baseline = the reference's own update() (oracle/_ref, compiled from the
reference sources) timed serially on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth

WORKLOADS = {
    "ch3d": dict(model="ch_dw", shape=(256, 256, 256), prec="f64", params=[1.0, 1.0, 0.005], steps=2000,
                 seed=20260102, lo=0.45, hi=0.55, fields=1, bytes_per_cell=16,
                 name="cahn-hilliard-3d", cpu_steps=2),
    "ch3d_f32": dict(model="ch_dw", shape=(256, 256, 256), prec="f32", params=[1.0, 1.0, 0.005], steps=2000,
                     seed=20260102, lo=0.45, hi=0.55, fields=1, bytes_per_cell=8,
                     name="cahn-hilliard-3d-fp32", cpu_steps=2),
    "ac2d": dict(model="ac", shape=(512, 512), prec="f64", params=[1.0, 1.0, 1.0, 1.0, 0.01], steps=1000,
                 seed=20260101, lo=-0.01, hi=0.01, fields=1, bytes_per_cell=16,
                 name="allen-cahn-2d", cpu_steps=100),
    "gg2d": dict(model="gg", shape=(2048, 2048), prec="f32", params=[0.05], steps=500, seed=20260103,
                 lo=0.0, hi=0.01, fields=32, bytes_per_cell=256, name="grain-growth-2d-q32", cpu_steps=2),
    # model B with the example's parameters (model_B.cpp:86-92: r = u = K = M = 1,
    # dt = 0.01, kT = 0.01, dV = 1, so sigma = sqrt(2 kT / (dt dV))) on the C2 grid
    # Zener pinning (zener.cpp:115: dt = 0.01), 2048^2 with 16 fields (field 0 the particles), FP32
    "zener2d": dict(model="zener", shape=(2048, 2048), prec="f32", params=[0.01], steps=500, seed=20260109,
                    lo=0.0, hi=1.0, fields=16, bytes_per_cell=2 * 16 * 4, name="zener-2d-q16", cpu_steps=2),
    # Ostwald ripening with the example's constants (ostwald.cpp:68-82), 1024^2, C and 3 order parameters, FP64
    "ostwald2d": dict(model="ostwald", shape=(1024, 1024), prec="f64",
                      params=[0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0], steps=500, seed=20260110,
                      lo=0.0, hi=1.0, fields=4, bytes_per_cell=2 * 4 * 8, name="ostwald-2d-q4", cpu_steps=2),
    # multi-phase grain growth (anisotropic/phase_field/graingrowth.cpp:105-106: dt = 0.01, width = 8;
    # energy and mobility 1), 1024^2 with 8 grains as one-hot blocks of 16 nodes, FP64
    "gganiso2d": dict(model="gg_aniso", shape=(1024, 1024), prec="f64", params=[0.01, 8.0, 1.0, 1.0], steps=500,
                      seed=20260111, lo=0.0, hi=1.0, fields=8, bytes_per_cell=2 * 8 * 8, name="multiphase-gg-2d-q8",
                      cpu_steps=2, init="grains"),
    # 3-D Allen-Cahn (allen-cahn.cpp:58-71 parameters) on a 256^3 periodic grid, FP64
    "ac3d": dict(model="ac", shape=(256, 256, 256), prec="f64", params=[1.0, 1.0, 1.0, 1.0, 0.01], steps=1000,
                 seed=20260112, lo=-0.01, hi=0.01, fields=1, bytes_per_cell=16, name="allen-cahn-3d", cpu_steps=2),
    # 3-D grain growth (graingrowth.cpp:84-105 with dt = 0.05), 256^3 with 8 fields, FP32
    "gg3d": dict(model="gg", shape=(256, 256, 256), prec="f32", params=[0.05], steps=200, seed=20260113,
                 lo=0.0, hi=0.01, fields=8, bytes_per_cell=2 * 8 * 4, name="grain-growth-3d-q8", cpu_steps=1),
    # 3-D Ostwald ripening (ostwald.cpp constants as ostwald2d), 256^3 with C and one order parameter (the
    # example's two fields), FP64: the two passes as plane marches
    "ostwald3d": dict(model="ostwald", shape=(256, 256, 256), prec="f64",
                      params=[0.001, 1.0, 1.0, 0.05, 0.95, 2.0, 1.0, 3.0, 2.0], steps=200, seed=20260114,
                      lo=0.0, hi=1.0, fields=2, bytes_per_cell=2 * 2 * 8, name="ostwald-3d-q2", cpu_steps=1),
    # 3-D multi-phase grain growth, 128^3 with 8 grains as one-hot blocks of 16 nodes, FP64
    "gganiso3d": dict(model="gg_aniso", shape=(128, 128, 128), prec="f64", params=[0.01, 8.0, 1.0, 1.0], steps=200,
                      seed=20260115, lo=0.0, hi=1.0, fields=8, bytes_per_cell=2 * 8 * 8, name="multiphase-gg-3d-q8",
                      cpu_steps=1, init="grains"),
    "model_b3d": dict(model="ch_noise", shape=(256, 256, 256), prec="f64",
                      params=[1.0, 1.0, 1.0, 1.0, 0.01, math.sqrt(2.0 * 0.01 / (0.01 * 1.0)), 20260107], steps=2000,
                      seed=20260108, lo=-0.01, hi=0.01, fields=1, bytes_per_cell=16, name="model-b-3d", cpu_steps=2),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, init):
    """Time the reference's update() serially on a bounded sample of the workload."""
    import oracle as O
    from oracle import dat

    shape, steps = w["shape"], w["cpu_steps"]
    nodes = int(np.prod(shape))
    ref_bin = {"ch_dw": "ch", "ac": "ac", "gg": "gg", "ch_noise": "mb", "zener": "zn", "ostwald": "ow",
               "gg_aniso": "ga"}[w["model"]]
    if O.ref_available() and os.path.exists(os.path.join(O.REF_DIR, "ref_" + ref_bin)):
        with tempfile.TemporaryDirectory() as td:
            fin = os.path.join(td, "in.dat")
            vec = w["fields"] > 1
            data = init if not vec else init.reshape(shape + (w["fields"],))
            if w["prec"] != "f64":
                data = data.astype(np.float32)
            dat.write_dat(fin, dat.make_grid(data, vector=vec), compress=False)
            model = ref_bin
            # model B, Zener, Ostwald: the example's own update (its parameters are the workload's)
            variant = "stock" if model in ("mb", "zn", "ow", "ga") else "config"
            sec = O.ref_run(model, fin, "-", steps, variant, w["params"], timeout=600)
        kind, what = "reference", "oracle/_ref reference update() compiled from the reference sources"
    else:
        t0 = time.perf_counter()
        if w["model"] == "ch_dw":
            O.ch(init.reshape(shape), w["params"], steps, mu=O.MU_DW)
        elif w["model"] == "ac":
            O.ac(init.reshape(shape), w["params"], steps)
        elif w["model"] == "ch_noise":
            O.ch_noise(init.reshape(shape), w["params"], steps)
        elif w["model"] == "zener":
            O.zener(init.reshape(shape + (w["fields"],)), w["params"][0], steps)
        elif w["model"] == "ostwald":
            O.ostwald(init.reshape(shape + (w["fields"],)), w["params"], steps)
        elif w["model"] == "gg_aniso":
            O.gg_aniso(init.reshape(shape + (w["fields"],)), w["params"], steps)
        else:
            O.gg(init.reshape(shape + (w["fields"],)), w["params"][0], steps)
        sec = time.perf_counter() - t0
        kind, what = "port", "oracle/liboracle C restatement"
    return {
        "value": nodes * steps / sec,
        "unit": "cell-updates/s",
        "cores": 1,
        "kind": kind,
        "sample": "%s, %s, %d steps of the full %s grid, serial, update() only (I/O excluded)"
        % (what, w["name"], steps, "x".join(map(str, shape))),
    }


def cpu_multicore(w, init):
    """Time the oracle's C restatement of the same update on all host cores
    (OpenMP node loops; Allen-Cahn and Cahn-Hilliard forms only) on the
    bounded sample cpu_baseline uses, or None for the vector models."""
    import oracle as O

    if w["model"] not in ("ac", "ch_dw", "ch_noise"):
        return None
    shape, steps = w["shape"], max(4, w["cpu_steps"])
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    before = O.set_threads(0)
    try:
        used = O.set_threads(threads)
        a = init.reshape(shape)
        form = "f64" if w["prec"] == "f64" else "f32"
        if form == "f32":
            a = a.astype(np.float32)
        t0 = time.perf_counter()
        if w["model"] == "ch_dw":
            O.ch(a, w["params"], steps, mu=O.MU_DW, form=form)
        elif w["model"] == "ac":
            O.ac(a, w["params"], steps, form=form)
        else:
            O.ch_noise(a, w["params"], steps, form=form)
        sec = time.perf_counter() - t0
    finally:
        O.set_threads(before)
    nodes = int(np.prod(shape))
    return {"value": nodes * steps / sec, "unit": "cell-updates/s", "cores": used, "kind": "port",
            "sample": "oracle/liboracle C restatement (%s form), %s, %d steps of the full %s grid, %d OpenMP threads"
            % (form, w["name"], steps, "x".join(map(str, shape)), used)}


def grains(shape, q, seed, block=16):
    """One-hot grains: each block of `block` nodes per axis belongs to a grain
    drawn from the seed (phi = 1 for it, 0 for the others)."""
    import mmsp_amd as mx

    blocks = tuple((n + block - 1) // block for n in shape)
    ids = np.floor(mx.fill_uniform(int(np.prod(blocks)), seed, 0.0, q)).astype(int).reshape(blocks)
    for ax in range(len(shape)):
        ids = np.repeat(ids, block, axis=ax)
    ids = ids[tuple(slice(0, n) for n in shape)]
    return (ids[..., None] == np.arange(q)).astype(np.float64)


def c4_members(seed=20260104, count=64):
    """64 Allen-Cahn instances: sides from {128..2048} and kappa ~ U(0.5, 2) drawn from the seed."""
    import mmsp_amd as mx

    pick = mx.fill_uniform(count, seed, 0.0, 5.0)
    kap = mx.fill_uniform(count, seed + 1, 0.5, 2.0)
    sides = [128 << int(v) for v in pick]
    return sides, kap


def bench_batch_ac(args):
    """Batched Allen-Cahn: 64 uneven FP32 instances in one persistent launch per
    step, hipGraph of 100 steps, autotuned strips; vs the same 64 solves run one
    after another on the plain FP64 single-grid path."""
    import mmsp_amd as mx

    sides, kap = c4_members()
    K = args.steps if args.steps is not None else 1000
    G = args.graph_steps
    shapes = [(s, s) for s in sides]
    params = np.array([[1.0, 1.0, k, 1.0, 0.01] for k in kap])
    inits = [mx.fill_uniform(s * s, 20260105 + m, -0.01, 0.01).reshape(s, s) for m, s in enumerate(sides)]
    cells = sum(s * s for s in sides)
    b = mx.Batch(shapes, "ac", precision="f32")
    b.set_params(params)
    for m, a in enumerate(inits):
        b.upload(m, a.astype(np.float32))
    report = b.autotune(trial_steps=5)
    log(report)
    b.step(args.warmup, graph_steps=0)
    for m, a in enumerate(inits):
        b.upload(m, a.astype(np.float32))
    # the batch's step counter parity is where the warm-up left it; the solves
    # start from the uploaded state either way
    t0 = time.perf_counter()
    dev_s = b.step_timed(K, graph_steps=G)
    wall_s = time.perf_counter() - t0
    out32 = [b.download(m) for m in range(len(sides))]
    tile = b.get_tile()
    b.close()
    # the same solves, sequentially, plain FP64 single grids
    seq_s = 0.0
    err = 0.0
    for m, s in enumerate([] if args.no_sequential else sides):
        with mx.Grid((s, s), precision="f64") as g:
            g.upload(inits[m])
            g.step("ac", params[m], 2)
            g.upload(inits[m])
            sec, _ = g.step_timed("ac", params[m], K)
            seq_s += sec
            r64 = g.download()
        err = max(err, float(np.abs(out32[m] - r64).max() / np.abs(r64).max()))
    value = cells * K / dev_s
    return {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": dev_s / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched-allen-cahn-64",
            "members": len(sides),
            "sides": sides,
            "cells": cells,
            "graph_steps": G,
            "tile": tile,
            "solves_per_s": len(sides) / dev_s,
            "fp64_sequential_solves_per_s": len(sides) / seq_s if seq_s else None,
            "speedup_vs_fp64_sequential": seq_s / dev_s if seq_s else None,
            "max_normwise_err_vs_fp64": err if seq_s else None,
            "wall_s_timed": wall_s,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": value * 8 / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": value * 8 / 1e9 / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("ac_batch_batch_ac", tile),
            "kernel": "ac_batch",
            "algorithmic_bytes_per_launch": 8 * cells,
        },
    }


def bench_ckpt(args):
    """The reference driver's loop (MMSP.main.hpp:278-350: update(grid,
    increment) then output(grid, file), repeated) on the 256^3 FP64
    Cahn-Hilliard grid, 2000 steps with a .dat checkpoint every 1000: once with
    the serial writer after each increment (compress2 level 9 on one core, the
    reference's output()), once with checkpoints that overlap the next
    increment and deflate on --ckpt-threads host threads."""
    import mmsp_amd as mx

    w = WORKLOADS["ch3d"]
    K = args.steps if args.steps is not None else w["steps"]
    inc = max(1, K // 2)
    shape = w["shape"]
    nodes = int(np.prod(shape))
    init = mx.fill_uniform(nodes, w["seed"], w["lo"], w["hi"]).reshape(shape)
    T = args.ckpt_threads
    out = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as td:
        for mode in ("serial", "overlapped"):
            with mx.Grid(shape, precision="f64") as g:
                g.upload(init)
                g.step(w["model"], w["params"], args.warmup)
                g.upload(init)
                g.sync()
                files, write_s, pending = [], [], None
                t0 = time.perf_counter()
                for k in range(0, K, inc):
                    g.step(w["model"], w["params"], min(inc, K - k))
                    f = os.path.join(td, "%s.%05d.dat" % (mode, k + inc))
                    if mode == "serial":
                        t1 = time.perf_counter()
                        g.write_dat(f)
                        write_s.append(time.perf_counter() - t1)
                    else:
                        if pending is not None:
                            write_s.append(pending.wait())
                        pending = g.checkpoint(f, threads=T)
                    files.append(f)
                if pending is not None:
                    write_s.append(pending.wait())
                g.sync()
                wall = time.perf_counter() - t0
                final = g.download()
            sizes = [os.path.getsize(f) for f in files]
            out[mode] = dict(wall_s=wall, write_s=write_s, file_bytes=sizes)
        # both loops wrote the same states (read back by the library's own reader)
        def read_back(name):
            with mx.Grid.from_dat(os.path.join(td, name)) as gg:
                return gg.download()

        for a, b in zip(sorted(f for f in os.listdir(td) if f.startswith("serial")),
                        sorted(f for f in os.listdir(td) if f.startswith("overlapped"))):
            if not np.array_equal(read_back(a), read_back(b)):
                raise RuntimeError("checkpoint %s differs from %s" % (b, a))
        if not np.array_equal(read_back(sorted(os.listdir(td))[-1]), final):
            raise RuntimeError("the last checkpoint is not the final state")
    payload = nodes * 8
    ser, ovl = out["serial"], out["overlapped"]
    value = nodes * K / ovl["wall_s"]
    return {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ovl["wall_s"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "cahn-hilliard-3d-checkpointed",
            "grid": list(shape),
            "checkpoint_every_steps": inc,
            "checkpoints": len(ovl["file_bytes"]),
            "writer_threads": T,
            "payload_bytes": payload,
            "file_bytes": ovl["file_bytes"],
            "serial_file_bytes": ser["file_bytes"],
            "wall_s_overlapped": ovl["wall_s"],
            "wall_s_serial_writer": ser["wall_s"],
            "speedup_vs_serial_writer": ser["wall_s"] / ovl["wall_s"],
            "checkpoint_s_overlapped": ovl["write_s"],
            "checkpoint_s_serial": ser["write_s"],
            "checkpoint_mb_per_s_overlapped": payload / np.mean(ovl["write_s"]) / 1e6,
            "checkpoint_mb_per_s_serial": payload / np.mean(ser["write_s"]) / 1e6,
            "note": "value counts the wall clock of steps plus checkpoints; the loop is bound by the host's "
                    "deflate, so no device roofline applies",
        },
        "roofline": None,
        "cpu_baseline": {
            "value": payload / np.mean(ser["write_s"]) / 1e6,
            "unit": "MB/s",
            "cores": 1,
            "kind": "port",
            "sample": "the serial writer (compress2 level 9 of the whole 128 MiB payload, the reference's "
                      "output()), %d checkpoints of the same run" % len(ser["write_s"]),
        },
    }


def bench_dropin(args):
    """The reference's driver loop end to end on the 256^3 FP64 Cahn-Hilliard
    grid: the reference's explicit Cahn-Hilliard program with its update()
    replaced by include/mmsp_gpu.hpp (mmsp_amd/dropin/cahn_hilliard_dw_gpu.cpp,
    the reference's main(), reader and output() compiled from its headers with
    the RAW switch), run as `prog start.dat 2000 100`: 20 update(grid, 100)
    calls, each followed by output() of a raw .dat. Wall clock of the whole
    process; the drop-in's own counters split it."""
    import re
    import subprocess

    import mmsp_amd as mx

    w = WORKLOADS["ch3d"]
    K = args.steps if args.steps is not None else w["steps"]
    inc = args.increment
    shape = w["shape"]
    nodes = int(np.prod(shape))
    prog = os.path.join(ROOT, "mmsp_amd", "dropin", "bin", "cahn-hilliard-dw-gpu-raw")
    if not os.path.exists(prog):
        raise SystemExit("bench.py: %s is not built (make -C mmsp_amd/dropin REF=<mmsp checkout>)" % prog)
    init = mx.fill_uniform(nodes, w["seed"], w["lo"], w["hi"]).reshape(shape)
    with mx.Grid(shape) as g:  # the expected final state, in-process
        g.upload(init)
        g.step(w["model"], w["params"], K)
        want = g.digest()
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR")) as td:
        mx.write_dat_host(os.path.join(td, "start.dat"), init, shape, compress=False)
        env = dict(os.environ, MMSP_GPU_STATS="1")
        t0 = time.perf_counter()
        r = subprocess.run([prog, "start.dat", str(K), str(inc)], cwd=td, env=env, capture_output=True, text=True,
                           timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("drop-in program failed: %s" % r.stderr[-2000:])
        m = re.search(r"update\(\) calls (\d+), steps (\d+), uploads (\d+), downloads (\d+); seconds: steps "
                      r"([0-9.]+), uploads ([0-9.]+), downloads ([0-9.]+), digests ([0-9.]+)", r.stderr)
        if not m:
            raise RuntimeError("no drop-in counters in: %s" % r.stderr[-2000:])
        calls, steps_run, uploads, downloads = (int(x) for x in m.groups()[:4])
        step_s, up_s, down_s, dig_s = (float(x) for x in m.groups()[4:])
        files = sorted(f for f in os.listdir(td) if f != "start.dat")
        with mx.Grid.from_dat(os.path.join(td, files[-1])) as g:
            got = g.digest()
        if got != want:
            raise RuntimeError("the drop-in's last checkpoint differs from the in-process run")
    value = nodes * K / wall
    return {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": 0,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "cahn-hilliard-3d-dropin-driver-loop",
            "grid": list(shape),
            "increment": inc,
            "update_calls": calls,
            "uploads": uploads,
            "downloads": downloads,
            "checkpoints": len(files),
            "wall_s": wall,
            "update_steps_s": step_s,
            "upload_s": up_s,
            "download_s": down_s,
            "digest_s": dig_s,
            "rest_s": wall - step_s - up_s - down_s - dig_s,
            "rest_note": "process start, the reference's reader and its output() of raw .dat files",
            "device_only_cell_updates_per_s": nodes * K / step_s,
            "final_checkpoint_equals_in_process_run": True,
        },
        "roofline": None,
        "cpu_baseline": cpu_baseline(w, init) if not args.no_cpu_baseline else None,
    }


def serve_trace(n, seed=20260106, lam=200.0):
    """Poisson arrivals (rate lam); half 2-D side U{64..1024}, half 3-D side
    U{64..256}; steps log-uniform in [100, 10000]."""
    import mmsp_amd as mx

    u = mx.fill_uniform(n * 4, seed, 0.0, 1.0).reshape(n, 4)
    gaps = -np.log1p(-u[:, 0]) / lam
    arrivals = np.cumsum(gaps) - gaps[0]
    out = []
    for k in range(n):
        if u[k, 1] < 0.5:
            side = 64 + int(u[k, 2] * (1024 - 64 + 1))
            shape = (side, side)
        else:
            side = 64 + int(u[k, 2] * (256 - 64 + 1))
            shape = (side, side, side)
        steps = int(round(math.exp(math.log(100) + u[k, 3] * (math.log(10000) - math.log(100)))))
        out.append(dict(t=float(arrivals[k]), shape=shape, steps=steps, seed=seed + 1000 + k))
    return out


def bench_serve(args):
    """Served Cahn-Hilliard stream through the continuous-batching server: FP32
    state with FP64 mass checks, requests of more than --fp64-above steps run
    in FP64 in the same rounds; latency is completion minus submission. Then
    every request runs alone, one after another, on the plain FP64
    single-grid path from the start the server used (float-rounded for the
    FP32 requests): that measures the FP64 one-after-another time and each
    served result's error against the FP64 form."""
    import threading

    import mmsp_amd as mx

    n = args.requests
    trace = serve_trace(n)
    params = [1.0, 1.0, 0.005]
    inits = [mx.fill_uniform(int(np.prod(r["shape"])), r["seed"], 0.45, 0.55).reshape(r["shape"]) for r in trace]
    cells = sum(int(np.prod(r["shape"])) * r["steps"] for r in trace)
    thr = args.fp64_above
    sv = mx.Server("ch_dw", "f32", hbm_budget=args.hbm_budget_gb << 30, max_slots=512, chunk_steps=100,
                   fp64_above_steps=thr)
    ids = [None] * n
    late = [0.0] * n

    def submitter():
        t0 = time.perf_counter()
        for k, r in enumerate(trace):
            dt = r["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            late[k] = max(0.0, (time.perf_counter() - t0) - r["t"])
            ids[k] = sv.submit(inits[k], params, r["steps"])

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    infos = [sv.wait(i, 600) for i in ids]
    st = sv.stats()
    done = [k for k in range(n) if infos[k]["status"] == mx.DONE]
    lat = np.array([infos[k]["t_done"] - infos[k]["t_submit"] for k in done])
    span = max(infos[k]["t_done"] for k in done) - min(i["t_submit"] for i in infos)
    value = len(done) / span
    cu = cells / span
    alg_bytes = sum(int(np.prod(trace[k]["shape"])) * trace[k]["steps"] * (16 if infos[k]["precision"] == "f64" else 8)
                    for k in done)
    # every request alone on the FP64 single-grid path, one after another
    errs32, errs64, mass32, mass64 = [], [], [], []
    seq_s = 0.0
    detail = []
    for k in done:
        r = trace[k]
        served = sv.fetch(ids[k])
        f64 = infos[k]["precision"] == "f64"
        start = inits[k] if f64 else inits[k].astype(np.float32).astype(np.float64)
        with mx.Grid(r["shape"], precision="f64") as g:
            g.upload(start)
            sec, _ = g.step_timed("ch_dw", params, r["steps"])
            r64 = g.download()
        seq_s += sec
        detail.append({"shape": list(r["shape"]), "steps": r["steps"], "fp64_alone_s": sec,
                       "precision": infos[k]["precision"], "t_submit": infos[k]["t_submit"],
                       "t_admit": infos[k].get("t_admit"), "t_done": infos[k]["t_done"]})
        e = float(np.abs(served - r64).max() / np.abs(r64).max())
        m = abs(infos[k]["mass1"] - infos[k]["mass0"]) / abs(infos[k]["mass0"])
        (errs64 if f64 else errs32).append((r["steps"], e))
        (mass64 if f64 else mass32).append(m)
    sv.close()
    if args.detail_out:
        with open(args.detail_out, "w") as f:
            json.dump(detail, f)
    all_err = [e for _, e in errs32 + errs64]
    return {
        "metric": "solves/s",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": 0,
        "ms_per_step": span / max(1, len(done)) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "served-cahn-hilliard-stream",
            "requests": n,
            "arrival_rate_per_s": 200.0,
            "completed": len(done),
            "rejected": st["rejected"],
            "fp64_above_steps": thr,
            "requests_fp32": len(errs32),
            "requests_fp64": len(errs64),
            "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_latency_ms": float(lat.mean() * 1e3),
            "makespan_s": span,
            "cell_updates_per_s": cu,
            "total_cell_updates": cells,
            "max_normwise_err_vs_fp64": max(all_err) if all_err else None,
            "max_normwise_err_vs_fp64_fp32_requests": max(e for _, e in errs32) if errs32 else None,
            "max_normwise_err_vs_fp64_fp64_requests": max(e for _, e in errs64) if errs64 else None,
            "normwise_err_vs_fp64_fp32_requests_by_steps": sorted(errs32)[-10:],
            "max_rel_mass_drift_fp32_requests": float(max(mass32)) if mass32 else None,
            "max_rel_mass_drift_fp64_requests": float(max(mass64)) if mass64 else None,
            "fp64_sequential_s": seq_s,
            "fp64_sequential_solves_per_s": len(done) / seq_s if seq_s else None,
            "speedup_vs_fp64_sequential": seq_s / span if seq_s else None,
            "fp64_sequential_note": "device time of every request run alone on the FP64 single-grid path, "
                                    "one after another (uploads and downloads excluded)",
            "peak_concurrent": st["peak_concurrent"],
            "peak_bytes": st["peak_bytes"],
            "rounds": st["rounds"],
            "launches": st["launches"],
            "max_submit_lateness_ms": float(max(late) * 1e3),
            "hbm_budget_gb": args.hbm_budget_gb,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": alg_bytes / span / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": alg_bytes / span / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "ch_batch",
            "note": "algorithmic bytes (8 per FP32 and 16 per FP64 cell-update) over the makespan",
        },
    }


def traffic_from_profiles(kernel_tag, tile=None):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this
    kernel and workload, or None when there is none or it was measured on
    another tile than `tile` (the one the timed run used)."""
    p = os.path.join(ROOT, "profiles", kernel_tag + ".traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    if tile is not None and d.get("tile") != tile:
        return None
    return d.get("traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=100,
                    help="untimed steps first (100 or more also builds the captured graph that small grids replay)")
    ap.add_argument("--workload", default="ch3d",
                    choices=sorted(WORKLOADS) + ["batch_ac", "serve", "ckpt", "dropin"])
    ap.add_argument("--increment", type=int, default=100, help="dropin: steps per update() call")
    ap.add_argument("--ckpt-threads", type=int, default=16, help="ckpt: host threads deflating a checkpoint")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests in the trace")
    ap.add_argument("--hbm-budget-gb", type=int, default=128, help="serve: device memory the server may hold")
    ap.add_argument("--fp64-above", type=int, default=1000,
                    help="serve: requests of more steps run in FP64 (0: every request in FP32)")
    ap.add_argument("--detail-out", default=None, help="serve: write per-request timings (JSON) to this file")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph-steps", type=int, default=100, help="batch_ac: steps per captured graph (0: direct launches)")
    ap.add_argument("--no-sequential", action="store_true",
                    help="batch_ac: skip the FP64 one-after-another comparison (for profiling the batch kernel)")
    ap.add_argument("--tune", action="store_true", help="run the autotuner first instead of the tuned default tile")
    ap.add_argument("--tile", default=None, help="ty,tz,chunk[,opt] for the tiled kernel (opt 1: register queues)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU only: the path is the per-GPU kernel; sharding over several GPUs "
              "(slabs along axis 0, one halo exchange per stencil pass) is not built", file=sys.stderr)
        sys.exit(2)

    import mmsp_amd as mx

    if args.workload == "batch_ac":
        print(json.dumps(bench_batch_ac(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.workload == "ckpt":
        print(json.dumps(bench_ckpt(args)), flush=True)
        return
    if args.workload == "dropin":
        print(json.dumps(bench_dropin(args)), flush=True)
        return
    w = WORKLOADS[args.workload]
    K = args.steps if args.steps is not None else w["steps"]
    W = args.warmup
    shape = w["shape"]
    nodes = int(np.prod(shape))
    q = w["fields"]
    init = mx.fill_uniform(nodes * q, w["seed"], w["lo"], w["hi"])
    if w.get("init") == "grains":
        init = grains(shape, q, w["seed"]).reshape(-1)

    g = mx.Grid(shape, fields=q, vector=q > 1, precision=w["prec"])
    g.upload(init.astype(g.dtype).reshape(g.host_shape))
    tune_report = ""
    if args.tile:
        v = [int(x) for x in args.tile.split(",")]
        g.set_tile(*v[:3])
        if len(v) > 3:
            g.set_tile_options(body=v[3])
    elif args.tune and w["model"] in ("ch_dw", "ch"):
        tune_report = g.autotune(w["model"], w["params"], trial_steps=10)
        log(tune_report)
        g.upload(init.astype(g.dtype).reshape(g.host_shape))
    m0 = g.mass()

    g.step(w["model"], w["params"], W)
    g.sync()
    t0 = time.perf_counter()
    dev_s, launches = g.step_timed(w["model"], w["params"], K)
    wall_s = time.perf_counter() - t0
    m1 = g.mass()
    tile = g.get_tile()  # what the timed steps ran
    state = g.download()
    finite = bool(np.all(np.isfinite(state)))
    err64 = None
    if w["prec"] != "f64":
        # the same K+W steps in FP64 (bitwise the reference's double semantics
        # with a correctly rounded cube) from the float-rounded start
        with mx.Grid(shape, fields=q, vector=q > 1, precision="f64") as g64:
            g64.upload(init.astype(np.float32).astype(np.float64).reshape(g64.host_shape))
            g64.step(w["model"], w["params"], W + K)
            r64 = g64.download()
            e64 = g64.energy(w["model"], w["params"])
        d = np.abs(state.astype(np.float64) - r64)
        if q > 1:
            err64 = float(max(d[..., j].max() / np.abs(r64[..., j]).max() for j in range(q)))
        else:
            err64 = float(d.max() / np.abs(r64).max())

    # Mean device time per launch over the timed region (HIP events on the
    # state's stream bracket the K back-to-back steps); with the fused path
    # one step is one launch of the dominant kernel. A second measurement
    # brackets every step with its own event pair (adds launch gaps).
    lps = launches / K
    mean_launch_s = dev_s / launches
    reps = min(K, 200)
    mean_s, min_s = g.kernel_time(w["model"], w["params"], reps)
    g.close()

    alg_bytes = w["bytes_per_cell"] * nodes
    per_launch = alg_bytes / lps
    achieved = per_launch / mean_launch_s / 1e9
    value = nodes * K / dev_s
    base = {"ch_dw": "ch", "ch": "ch", "ch_noise": "ch", "ac": "ac", "gg": "gg", "zener": "gg", "ostwald": "ow",
            "gg_aniso": "gga"}[w["model"]]
    if tile["variant"] != 2:
        kname = base + "_generic"
    elif len(shape) == 3 and tile["opt"] == 8:
        kname = "march3d"
    elif len(shape) == 3 and tile["opt"] == 12:
        kname = "march3d_ow" if base == "ow" else "march3d2"
    elif base == "ch":
        kname = "ch_fused"
    elif base == "ow":
        kname = "ow_waves"
    elif tile["opt"] == 2:
        kname = "wave_strips"
    elif tile["opt"] in (4, 5):
        kname = "wave_strips_split"
    elif tile["opt"] in (9, 10, 11):
        kname = "wave_strips_split2"
    elif tile["opt"] == 3 and base == "ac":
        kname = "ac_tblocks"
    elif base == "gg" and tile["opt"] == 1:
        kname = "gg_tile2d"
    else:
        kname = base + "_tiled"
    kernel_tag = "%s_%s" % (kname, args.workload)
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dev_s / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": w["prec"],
        "data": "synthetic",
        "config": {
            "workload": w["name"],
            "grid": list(shape),
            "fields": q,
            "model": w["model"],
            "params": w["params"],
            "precision": w["prec"],
            "boundary": "periodic",
            "seed": w["seed"],
            "tile": tile,
            "launches_per_step": launches / K,
            "wall_s_timed": wall_s,
            "hbm_frac_of_value": value * w["bytes_per_cell"] / (HBM_PEAK_GBS * 1e9),
            # conserved by Cahn-Hilliard only
            "mass_rel_change": float(np.max(np.abs(m1 - m0) / np.abs(m0))) if w["model"] in ("ch", "ch_dw") else None,
            "mass_abs_change": float(np.max(np.abs(m1 - m0))) if w["model"] == "ch_noise" else None,
            "max_normwise_err_vs_fp64": err64,
            "finite": finite,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(kernel_tag, tile),
            "kernel": kernel_tag,
            "kernel_mean_us": mean_launch_s * 1e6,
            "kernel_isolated_mean_us": mean_s / lps * 1e6,
            "kernel_isolated_min_us": min_s / lps * 1e6,
            "algorithmic_bytes_per_launch": per_launch,
        },
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(w, init)
        except Exception as e:  # the baseline is reported, never the measured value
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        try:
            mc = cpu_multicore(w, init)
            if mc:
                out["config"]["cpu_multicore"] = mc
        except Exception as e:
            out["config"]["cpu_multicore"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
