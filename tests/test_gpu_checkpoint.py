"""GPU: .dat checkpoints taken from device state while it keeps stepping
(mmspx_checkpoint_begin / mmspx_checkpoint_wait).

Bars: the file holds exactly the state at the moment of begin (bitwise, read
back by the oracle's reader); steps issued while it is written give the same
bits as steps with no checkpoint; one writing thread gives the same bytes as
mmspx_write_dat (the reference's serial writer, compress2 level 9); several
give one zlib stream with the same payload."""
import sys

import numpy as np
import pytest

import oracle as O
from oracle import dat

pytestmark = pytest.mark.gpu

mx = pytest.importorskip("mmsp_amd")

CH_DW = [1.0, 1.0, 0.005]  # M, kappa, dt


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ch_grid(n, prec="f64", seed=3):
    g = mx.Grid((n, n, n), precision=prec)
    g.upload(0.5 + 0.05 * O.uniform(n ** 3, seed, -1, 1).reshape(n, n, n))
    return g


def body(path):
    buf = open(path, "rb").read()
    g = dat.read_dat(buf)
    return g, buf


@pytest.mark.parametrize("threads", [1, 4, 0])
def test_checkpoint_is_the_state_at_begin_and_stepping_goes_on(tmp_path, threads):
    say("[gpu] checkpoint 64^3 CH FP64 threads=%d" % threads)
    with ch_grid(64) as g, ch_grid(64) as ref:
        g.step("ch_dw", CH_DW, 20)
        at_begin = g.download()
        ck = g.checkpoint(str(tmp_path / "c.dat"), threads=threads)
        g.step("ch_dw", CH_DW, 30)  # overlaps the download and the deflate
        ck.wait()
        got, _ = body(str(tmp_path / "c.dat"))
        assert np.array_equal(got.data, at_begin)
        ref.step("ch_dw", CH_DW, 50)
        assert np.array_equal(g.download(), ref.download())


def test_one_thread_matches_the_serial_writer_bytes(tmp_path):
    say("[gpu] checkpoint threads=1 vs mmspx_write_dat")
    with ch_grid(48, seed=5) as g:
        g.step("ch_dw", CH_DW, 5)
        g.write_dat(str(tmp_path / "serial.dat"))
        g.checkpoint(str(tmp_path / "ck.dat"), threads=1).wait()
        g.checkpoint(str(tmp_path / "par.dat"), threads=8).wait()
        serial = open(str(tmp_path / "serial.dat"), "rb").read()
        assert open(str(tmp_path / "ck.dat"), "rb").read() == serial
        par, _ = body(str(tmp_path / "par.dat"))
        assert np.array_equal(par.data, dat.read_dat(serial).data)
        g.checkpoint(str(tmp_path / "raw.dat"), compress=False, threads=3).wait()
        g.write_dat(str(tmp_path / "serial_raw.dat"), compress=False)
        assert open(str(tmp_path / "raw.dat"), "rb").read() == open(str(tmp_path / "serial_raw.dat"), "rb").read()


def test_vector_grid_checkpoint_is_node_major(tmp_path):
    say("[gpu] checkpoint GG 2-D Q=8 FP32")
    n, q = 160, 8
    a = (0.01 * O.uniform(n * n * q, 9, 0, 1)).reshape(n, n, q).astype(np.float32)
    with mx.Grid((n, n), fields=q, vector=True, precision="f32") as g:
        g.upload(a)
        g.step("gg", [0.05], 3)
        at_begin = g.download()
        ck = g.checkpoint(str(tmp_path / "gg.dat"), threads=4)
        g.step("gg", [0.05], 3)
        ck.wait()
        got, buf = body(str(tmp_path / "gg.dat"))
        assert got.vector and got.data.shape == (n, n, q) and np.array_equal(got.data, at_begin)
        g.write_dat(str(tmp_path / "gg_serial.dat"))
        # headers agree; payloads agree once inflated
        s = open(str(tmp_path / "gg_serial.dat"), "rb").read()
        assert buf.split(b"\n")[:5] == s.split(b"\n")[:5]


def test_several_in_flight_and_destroy_before_wait(tmp_path):
    say("[gpu] checkpoints in flight across steps, state destroyed before wait")
    g = ch_grid(40, prec="f32", seed=7)
    snaps, cks = [], []
    for k in range(3):
        g.step("ch_dw", CH_DW, 10)
        snaps.append(g.download())
        cks.append(g.checkpoint(str(tmp_path / ("k%d.dat" % k)), threads=2))
    g.close()
    for k, ck in enumerate(cks):
        ck.wait()
        got, _ = body(str(tmp_path / ("k%d.dat" % k)))
        assert np.array_equal(got.data, snaps[k])


def test_reader_round_trip_of_a_parallel_file(tmp_path):
    say("[gpu] mmspx_read_dat of a many-thread checkpoint")
    with ch_grid(32, seed=11) as g:
        g.step("ch_dw", CH_DW, 4)
        g.checkpoint(str(tmp_path / "p.dat"), threads=6).wait()
        with mx.Grid.from_dat(str(tmp_path / "p.dat")) as h:
            assert np.array_equal(h.download(), g.download())


def test_bad_arguments(tmp_path):
    with ch_grid(8) as g:
        with pytest.raises(mx.MmspxError):
            g.checkpoint(str(tmp_path / "x.dat"), threads=-2)
        ck = g.checkpoint(str(tmp_path / "no" / "dir" / "x.dat"), threads=2)
        with pytest.raises(mx.MmspxError) as e:
            ck.wait()
        assert e.value.code == -4  # MMSPX_EIO
        with pytest.raises(RuntimeError):
            ck.wait()
