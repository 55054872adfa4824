"""GPU: the continuous-batching server. Results must not depend on what else
is resident: each request equals (bitwise) the oracle's run of it alone, the
single-grid path's run of it, and its own run on an otherwise empty server."""
import sys
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
mx = pytest.importorskip("mmsp_amd")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


P_DW = [1.0, 1.0, 0.005]


def req(shape, seed):
    return O.uniform(int(np.prod(shape)), seed, 0.45, 0.55).reshape(shape)


def oracle_f32(init, params, steps):
    return O.ch(init.astype(np.float32), params, steps, mu=O.MU_DW, form="f32").astype(np.float64)


def test_mixed_stream_alone_vs_in_mix():
    specs = [((64, 64), 130, 1), ((24, 20, 28), 75, 2), ((100, 36), 310, 3), ((16, 16, 16), 40, 4),
             ((70, 130), 95, 5), ((33, 40, 18), 200, 6), ((30, 700), 60, 7), ((20, 70, 150), 45, 8)]
    say("[gpu] server mixed stream")
    params = [[1.0, 1.0, 0.005], [0.8, 1.0, 0.004], [1.2, 1.0, 0.005], [1.0, 0.5, 0.005], [1.0, 1.0, 0.003],
              [1.0, 1.0, 0.005], [1.0, 1.0, 0.005], [0.9, 1.0, 0.005]]
    with mx.Server("ch_dw", "f32", max_slots=8, chunk_steps=16) as sv:
        ids = []
        for k, (shape, steps, seed) in enumerate(specs):
            ids.append(sv.submit(req(shape, seed), params[k], steps))
            time.sleep(0.003 * (k % 3))  # arrivals while earlier requests run
        infos = [sv.wait(i, 120) for i in ids]
        mix = [sv.fetch(i) for i in ids]
        st = sv.stats()
    assert st["completed"] == len(specs) and st["rejected"] == 0 and st["worker_error"] == 0
    assert st["peak_concurrent"] >= 2
    for k, (shape, steps, seed) in enumerate(specs):
        assert infos[k]["status"] == mx.DONE and infos[k]["steps_done"] == steps
        # FP32 state, FP64 mass check
        assert abs(infos[k]["mass1"] - infos[k]["mass0"]) <= 1e-5 * abs(infos[k]["mass0"])
        exp = oracle_f32(req(shape, seed), params[k], steps)
        assert np.array_equal(mix[k], exp), (k, shape)
        with mx.Grid(shape, precision="f32") as g:
            g.upload(req(shape, seed).astype(np.float32))
            g.step("ch_dw", params[k], steps)
            assert np.array_equal(g.download().astype(np.float64), mix[k])
    # one of them alone on a fresh server
    with mx.Server("ch_dw", "f32", max_slots=8, chunk_steps=7) as sv:
        shape, steps, seed = specs[2]
        i = sv.submit(req(shape, seed), params[2], steps)
        sv.wait(i, 60)
        assert np.array_equal(sv.fetch(i), mix[2])


def test_fp32_server_runs_long_requests_in_fp64():
    """An FP32 server with fp64_above_steps = 60 runs the longer requests in
    FP64, in the same rounds as the FP32 ones: each result is bitwise the
    oracle's run of it in its form (and the single-grid path's); the FP64 ones
    conserve mass to 1e-12."""
    specs = [((64, 64), 40, 21), ((24, 20, 28), 75, 22), ((100, 36), 130, 23), ((16, 16, 16), 50, 24),
             ((33, 40, 18), 61, 25), ((30, 70), 60, 26), ((20, 18, 300), 90, 27)]
    say("[gpu] server fp32 with fp64 routing")
    with mx.Server("ch_dw", "f32", max_slots=8, chunk_steps=16, fp64_above_steps=60) as sv:
        ids = []
        for k, (shape, steps, seed) in enumerate(specs):
            ids.append(sv.submit(req(shape, seed), P_DW, steps))
            time.sleep(0.002 * (k % 2))
        infos = [sv.wait(i, 120) for i in ids]
        outs = [sv.fetch(i) for i in ids]
        st = sv.stats()
    assert st["completed"] == len(specs) and st["worker_error"] == 0
    assert st["fp64_routed"] == sum(1 for _, steps, _ in specs if steps > 60)
    for k, (shape, steps, seed) in enumerate(specs):
        long = steps > 60
        assert infos[k]["precision"] == ("f64" if long else "f32"), k
        if long:
            exp = O.ch(req(shape, seed), P_DW, steps, mu=O.MU_DW)
            assert abs(infos[k]["mass1"] - infos[k]["mass0"]) <= 1e-12 * abs(infos[k]["mass0"])
        else:
            exp = oracle_f32(req(shape, seed), P_DW, steps)
        assert np.array_equal(outs[k], exp), (k, shape, steps)


def test_large_3d_members_run_the_single_grid_kernels():
    """3-D requests of 2^21 nodes or more advance on the single-grid kernels
    (two steps per pass) in the same rounds as the batched ones: rounds of
    odd lengths (members finish at odd steps, 16-step chunks), FP32 and the
    FP64 form; each result bitwise the oracle's run of it, and the large FP64
    one equal to the single-grid path's run."""
    specs = [((130, 128, 128), 37, 51), ((128, 130, 129), 23, 52), ((20, 18, 30), 25, 53), ((40, 64), 41, 54),
             ((129, 129, 128), 9, 55)]
    say("[gpu] server with large 3-D members")
    with mx.Server("ch_dw", "f32", max_slots=8, chunk_steps=16, fp64_above_steps=30) as sv:
        ids = [sv.submit(req(shape, seed), P_DW, steps) for shape, steps, seed in specs]
        infos = [sv.wait(i, 120) for i in ids]
        outs = [sv.fetch(i) for i in ids]
        st = sv.stats()
    assert st["completed"] == len(specs) and st["worker_error"] == 0
    for k, (shape, steps, seed) in enumerate(specs):
        long = steps > 30
        assert infos[k]["precision"] == ("f64" if long else "f32"), k
        exp = O.ch(req(shape, seed), P_DW, steps, mu=O.MU_DW) if long else oracle_f32(req(shape, seed), P_DW, steps)
        assert np.array_equal(outs[k], exp), (k, shape, steps)
    with mx.Grid(specs[0][0], precision="f64") as g:
        g.upload(req(specs[0][0], specs[0][2]))
        g.step("ch_dw", P_DW, specs[0][1])
        assert np.array_equal(g.download(), outs[0])


def test_fp32_results_at_the_routing_threshold_hold_1e5():
    """At the benchmark stream's threshold (1000 steps), FP32-served results
    stay within 1e-5 normwise of the FP64 form from the same float-rounded
    start (2-D 512^2 and 3-D 96^3)."""
    say("[gpu] server fp32 error at 1000 steps")
    specs = [((512, 512), 1000, 31), ((96, 96, 96), 1000, 32)]
    with mx.Server("ch_dw", "f32", chunk_steps=100, fp64_above_steps=1000) as sv:
        ids = [sv.submit(req(shape, seed), P_DW, steps) for shape, steps, seed in specs]
        infos = [sv.wait(i, 120) for i in ids]
        outs = [sv.fetch(i) for i in ids]
    for k, (shape, steps, seed) in enumerate(specs):
        assert infos[k]["precision"] == "f32"
        with mx.Grid(shape, precision="f64") as g:
            g.upload(req(shape, seed).astype(np.float32).astype(np.float64))
            g.step("ch_dw", P_DW, steps)
            r64 = g.download()
        err = np.abs(outs[k] - r64).max() / np.abs(r64).max()
        say("[gpu] %s %d steps: fp32 err vs fp64 %.3e" % (shape, steps, err))
        assert err <= 1e-5  # the FP32 bar for served solves


def test_fp64_server_conserves_mass():
    say("[gpu] server fp64 mass")
    with mx.Server("ch_dw", "f64", chunk_steps=50) as sv:
        a = req((48, 40, 36), 11)
        i = sv.submit(a, P_DW, 200)
        info = sv.wait(i, 60)
        got = sv.fetch(i)
    assert abs(info["mass1"] - info["mass0"]) <= 1e-12 * abs(info["mass0"])
    assert np.array_equal(got, O.ch(a, P_DW, 200, mu=O.MU_DW))


def test_admission_by_memory_and_rejection():
    say("[gpu] server admission")
    small = req((64, 64), 21)
    one = 2 * ((64 * 64 * 4 + 255) // 256 * 256)
    with mx.Server("ch_dw", "f32", hbm_budget=one, max_slots=16, chunk_steps=10) as sv:
        ids = [sv.submit(small, P_DW, 30) for _ in range(4)]
        big = sv.submit(req((128, 128), 22), P_DW, 10)  # can never fit the budget
        infos = [sv.wait(i, 60) for i in ids]
        binfo = sv.wait(big, 60)
        st = sv.stats()
        outs = [sv.fetch(i) for i in ids]
    assert binfo["status"] == mx.REJECTED
    assert all(i["status"] == mx.DONE for i in infos)
    assert st["peak_concurrent"] == 1 and st["peak_bytes"] <= one and st["rejected"] == 1
    # admitted one after the other, in submission order
    assert all(infos[k]["t_admit"] >= infos[k - 1]["t_done"] - 1e-3 for k in range(1, 4))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_cancel_queued_request():
    say("[gpu] server cancel")
    a = req((64, 64), 31)
    one = 2 * ((64 * 64 * 4 + 255) // 256 * 256)
    with mx.Server("ch_dw", "f32", hbm_budget=one, max_slots=1, chunk_steps=10) as sv:
        i1 = sv.submit(a, P_DW, 2000)
        i2 = sv.submit(a, P_DW, 10)
        sv.cancel(i2)
        assert sv.wait(i2, 60)["status"] == mx.CANCELLED
        assert sv.wait(i1, 60)["status"] == mx.DONE
        with pytest.raises(mx.MmspxError):
            sv.fetch(i2)


def test_cancel_running_large_member():
    """A large 3-D member (the single-grid kernels on the second stream)
    cancelled while it runs: it ends cancelled, a small request beside it
    completes bitwise, and a large request submitted afterwards runs on the
    freed slot and equals the single-grid path."""
    say("[gpu] server cancel of a running large member")
    big, small = (130, 128, 128), (24, 20, 28)
    with mx.Server("ch_dw", "f64", max_slots=4, chunk_steps=50) as sv:
        i0 = sv.submit(req(big, 61), P_DW, 40000)
        i1 = sv.submit(req(small, 62), P_DW, 30)
        t0 = time.time()
        while sv.info(i0)["status"] == mx.QUEUED and time.time() - t0 < 60:
            time.sleep(0.005)
        assert sv.info(i0)["status"] == mx.RUNNING
        sv.cancel(i0)
        assert sv.wait(i0, 60)["status"] == mx.CANCELLED
        assert sv.wait(i1, 60)["status"] == mx.DONE
        assert np.array_equal(sv.fetch(i1), O.ch(req(small, 62), P_DW, 30, mu=O.MU_DW))
        i2 = sv.submit(req(big, 63), P_DW, 21)
        assert sv.wait(i2, 60)["status"] == mx.DONE
        got = sv.fetch(i2)
        st = sv.stats()
    assert st["worker_error"] == 0
    with mx.Grid(big, precision="f64") as g:
        g.upload(req(big, 63))
        g.step("ch_dw", P_DW, 21)
        assert np.array_equal(g.download(), got)


def test_submit_validation():
    with mx.Server("ch_dw", "f32") as sv:
        with pytest.raises(mx.MmspxError):
            sv.submit(req((8, 8), 1), [1.0, 1.0], 10)  # wrong parameter count
        with pytest.raises(mx.MmspxError):
            sv.submit(req((8, 8), 1), P_DW, 10, bc=[(mx.NEUMANN, mx.NEUMANN), (mx.PERIODIC, mx.PERIODIC)])
