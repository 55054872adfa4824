"""CPU: the .dat writer with the zlib stream deflated by several host threads
(mmspx_write_dat_host; the device checkpoint path shares it).

Pinned by: zlib's own inflate (the block decompresses to the payload the
serial layout prescribes, include/MMSP.grid.cpp:1136-1157 and
MMSP.vector.cpp:100-105), byte equality of the one-thread file with
compress2 level 9 (oracle/dat.py writes that with Python's zlib), and the
reference's reader and writer (oracle/_ref, grid::input then grid::output)
turning the many-thread file into the one-thread file byte for byte."""
import os
import struct
import zlib

import numpy as np
import pytest

import oracle as O
from oracle import dat

mm = pytest.importorskip("mmsp_amd")


def block_of(buf):
    """(header text, size_in_mem, size_on_disk, body) of a one-block .dat."""
    g = dat.read_dat(buf)
    dim = g.dim
    pos = 0
    for _ in range(3 + 2 * dim):
        pos = buf.index(b"\n", pos) + 1
    (nb,) = struct.unpack_from("<i", buf, pos)
    assert nb == 1
    pos += 4 + 16 * dim
    sm, sd = struct.unpack_from("<QQ", buf, pos)
    pos += 16
    body = buf[pos:]
    assert len(body) == sd
    return buf[: pos - 16 - 16 * dim - 4], sm, sd, body


def write(tmp_path, name, a, shape, **kw):
    p = str(tmp_path / name)
    mm.write_dat_host(p, a, shape, **kw)
    with open(p, "rb") as f:
        return p, f.read()


CASES = [
    # (shape, fields, vector, precision); scalar grids carry the header field
    # count 0, as the reference's example programs construct them
    ((700, 600), 0, False, "f64"),       # 3.4 MB: many chunks at 8 threads
    ((40, 50, 60), 0, False, "f32"),
    ((300, 200), 5, True, "f32"),        # vector nodes: int32 count + 5 floats
    ((130, 90), 3, True, "f64"),
    ((7,), 0, False, "f64"),             # smaller than one chunk
]


@pytest.mark.parametrize("shape,fields,vector,prec", CASES)
def test_parallel_stream_inflates_to_the_serial_payload(tmp_path, shape, fields, vector, prec):
    n = int(np.prod(shape)) * (fields if vector else 1)
    dt = np.float64 if prec == "f64" else np.float32
    # smooth plus noise: compressible in parts, like a phase field
    x = np.linspace(0, 40, n)
    a = (0.5 + 0.4 * np.tanh(np.sin(x)) + 1e-3 * O.uniform(n, 7, -1, 1)).astype(dt)
    host_shape = tuple(shape) + ((fields,) if vector else ())
    bc = [(O.PERIODIC, O.NEUMANN)] + [(O.PERIODIC, O.PERIODIC)] * (len(shape) - 1)
    kw = dict(fields=fields, vector=vector, precision=prec, bc=bc, dx=[0.5] * len(shape))
    _, one = write(tmp_path, "one.dat", a, shape, threads=1, **kw)
    _, many = write(tmp_path, "many.dat", a, shape, threads=8, **kw)
    # the one-thread file is compress2 level 9's, as oracle/dat.py writes it
    ref = dat.write_dat(None, dat.make_grid(a.reshape(host_shape), vector=vector, bc=bc, dx=[0.5] * len(shape)))
    assert one == ref
    h1, sm1, sd1, body1 = block_of(one)
    h8, sm8, sd8, body8 = block_of(many)
    assert h1 == h8 and sm1 == sm8
    raw = zlib.decompress(body1)
    assert zlib.decompress(body8) == raw and len(raw) == sm8
    # Adler-32 trailer of the whole payload, big-endian (RFC 1950)
    assert struct.unpack(">I", body8[-4:])[0] == zlib.adler32(raw)
    # chunked level-9 deflate costs little ratio
    assert sd8 <= sd1 * 1.01 + 64
    g = dat.read_dat(many)
    assert np.array_equal(g.data.reshape(-1), a.reshape(-1)) and g.b0 == [O.PERIODIC] * len(shape)


def test_raw_blocks_are_identical_for_any_thread_count(tmp_path):
    a = O.uniform(256 * 300 * 4, 3, 0, 1).astype(np.float32)
    kw = dict(fields=4, vector=True, precision="f32", compress=False)
    _, one = write(tmp_path, "one.dat", a, (256, 300), threads=1, **kw)
    _, many = write(tmp_path, "many.dat", a, (256, 300), threads=6, **kw)
    assert one == many
    _, sm, sd, _ = block_of(one)
    assert sm == sd == 256 * 300 * (4 + 16)


def test_chunk_boundaries_do_not_depend_on_thread_count(tmp_path):
    """Different thread counts cut different chunks; every file inflates to the same payload."""
    a = O.uniform(900 * 500, 11, -1, 1)
    bodies = []
    for t in (2, 3, 5, 16):
        _, b = write(tmp_path, "t%d.dat" % t, a, (900, 500), threads=t)
        bodies.append(zlib.decompress(block_of(b)[3]))
    assert all(b == bodies[0] for b in bodies) and bodies[0] == a.tobytes()


def test_errors(tmp_path):
    a = np.zeros(16)
    with pytest.raises(mm.MmspxError) as e:
        mm.write_dat_host(str(tmp_path / "no" / "such" / "dir.dat"), a, (4, 4), threads=4)
    assert e.value.code == -4  # MMSPX_EIO
    with pytest.raises(mm.MmspxError):
        mm.write_dat_host(str(tmp_path / "x.dat"), a, (4, 4), threads=-1)
    with pytest.raises(ValueError):
        mm.write_dat_host(str(tmp_path / "x.dat"), a, (4, 5))
    assert not os.path.exists(str(tmp_path / "x.dat"))


@pytest.mark.skipif(not O.ref_available(), reason="reference not built here")
@pytest.mark.parametrize("model,shape,fields,vector,prec", [
    ("ac", (640, 480), 0, False, "f64"),
    ("ch", (48, 40, 36), 0, False, "f32"),
    ("gg", (200, 150), 6, True, "f32"),
])
def test_reference_reader_reads_the_parallel_file(tmp_path, model, shape, fields, vector, prec):
    """grid::input reads the many-thread file; grid::output of what it read is
    byte for byte the one-thread file (and so the serial writer's)."""
    n = int(np.prod(shape)) * (fields if vector else 1)
    dt = np.float64 if prec == "f64" else np.float32
    a = (0.5 + 0.05 * O.uniform(n, 5, -1, 1)).astype(dt)
    kw = dict(fields=fields, vector=vector, precision=prec)
    p1, one = write(tmp_path, "one.dat", a, shape, threads=1, **kw)
    p8, _ = write(tmp_path, "many.dat", a, shape, threads=8, **kw)
    back = str(tmp_path / "back.dat")
    O.ref_run(model, p8, back, 0, timeout=120)
    with open(back, "rb") as f:
        assert f.read() == one
