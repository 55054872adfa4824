"""TEST INFRASTRUCTURE: restatement of the reference's serial `.dat` format.

Used by the tests to write seeded inputs for the reference harness and to read
what the reference and the product write. Never imported by product code.

Layout (serial writer, include/MMSP.grid.cpp:1358-1402 and 1818-1909):
  text   name            "grid:double", "grid:float", "grid:vector:double" ...
                         (include/MMSP.grid.h:593-596, MMSP.vector.h:115-117)
  text   dim, fields, "g0 g1" per axis, dx per axis (ostream default: 6 sig. digits)
  int32  number of blocks (1 in serial)
  per block: int32 x0,x1 per axis; int32 b0,b1 per axis; uint64 size_in_mem;
             uint64 size_on_disk; payload (zlib compress2 level 9, or raw when
             the two sizes are equal, MMSP.grid.cpp:1306-1330)
  payload: nodes in row-major order (axis 0 slowest); scalar T as sizeof(T)
           bytes; vector<T> as int32 n then n values (MMSP.vector.cpp:100-105).
The reader (MMSP.grid.cpp:1183-1356) copies each block's overlap into the grid.
"""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass, field

import numpy as np

MIRROR, NEUMANN, PERIODIC, PARALLEL, DIRICHLET = 0, 1, 2, 3, 4

_TYPES = {"double": np.float64, "float": np.float32}


def fmt_dx(v: float) -> str:
    """std::ostream << double with default precision 6 (%g)."""
    return "%g" % v


@dataclass
class DatGrid:
    """A serial grid: `data` has shape extents (+ (fields,) for vectors)."""
    data: np.ndarray
    vector: bool
    g0: list = field(default_factory=list)
    b0: list = field(default_factory=list)
    b1: list = field(default_factory=list)
    dx: list = field(default_factory=list)
    hdr_fields: int = -1  # header "fields" value; scalar grids made by generate() carry 0

    @property
    def dim(self) -> int:
        return self.data.ndim - (1 if self.vector else 0)

    @property
    def fields(self) -> int:
        return self.data.shape[-1] if self.vector else 1

    @property
    def extents(self) -> tuple:
        return self.data.shape[: self.dim]

    @property
    def type_name(self) -> str:
        base = {np.dtype(np.float64): "double", np.dtype(np.float32): "float"}[self.data.dtype]
        return "grid:vector:" + base if self.vector else "grid:" + base


def make_grid(data: np.ndarray, vector: bool = False, bc=None, dx=None, g0=None) -> DatGrid:
    dim = data.ndim - (1 if vector else 0)
    bc = bc if bc is not None else [(PERIODIC, PERIODIC)] * dim
    return DatGrid(
        data=np.ascontiguousarray(data),
        vector=vector,
        g0=list(g0) if g0 is not None else [0] * dim,
        b0=[b[0] for b in bc],
        b1=[b[1] for b in bc],
        dx=list(dx) if dx is not None else [1.0] * dim,
    )


def payload(g: DatGrid) -> bytes:
    if not g.vector:
        return g.data.tobytes(order="C")
    n = int(np.prod(g.extents))
    q = g.fields
    rec = np.empty((n,), dtype=[("n", "<i4"), ("v", g.data.dtype, (q,))])
    rec["n"] = q
    rec["v"] = g.data.reshape(n, q)
    return rec.tobytes()


def write_dat(path: str, g: DatGrid, compress: bool = True) -> bytes:
    nf = g.hdr_fields if g.hdr_fields >= 0 else (g.fields if g.vector else 0)
    hdr = [g.type_name, str(g.dim), str(nf)]
    for d in range(g.dim):
        hdr.append("%d %d" % (g.g0[d], g.g0[d] + g.extents[d]))
    for d in range(g.dim):
        hdr.append(fmt_dx(g.dx[d]))
    out = bytearray(("\n".join(hdr) + "\n").encode())
    out += struct.pack("<i", 1)
    for d in range(g.dim):
        out += struct.pack("<ii", g.g0[d], g.g0[d] + g.extents[d])
    for d in range(g.dim):
        out += struct.pack("<ii", g.b0[d], g.b1[d])
    raw = payload(g)
    body = zlib.compress(raw, 9) if compress else raw
    out += struct.pack("<QQ", len(raw), len(body))
    out += body
    if path:
        with open(path, "wb") as f:
            f.write(out)
    return bytes(out)


def read_dat(path_or_bytes) -> DatGrid:
    if isinstance(path_or_bytes, (bytes, bytearray)):
        buf = bytes(path_or_bytes)
    else:
        with open(path_or_bytes, "rb") as f:
            buf = f.read()
    pos = 0

    def line():
        nonlocal pos
        e = buf.index(b"\n", pos)
        s = buf[pos:e].decode()
        pos = e + 1
        return s

    name = line()
    dim = int(line())
    fields = int(line())
    g0, g1 = [], []
    for _ in range(dim):
        a, b = line().split()
        g0.append(int(a))
        g1.append(int(b))
    dx = [float(line()) for _ in range(dim)]
    parts = name.split(":")
    vector = parts[1] == "vector"
    dt = _TYPES[parts[-1]]
    ext = [g1[d] - g0[d] for d in range(dim)]
    shape = ext + ([fields] if vector else [])
    data = np.zeros(shape, dtype=dt)
    (nblocks,) = struct.unpack_from("<i", buf, pos)
    pos += 4
    b0, b1 = [PERIODIC] * dim, [PERIODIC] * dim
    for _ in range(nblocks):
        lim = struct.unpack_from("<%di" % (2 * dim), buf, pos)
        pos += 8 * dim
        bcs = struct.unpack_from("<%di" % (2 * dim), buf, pos)
        pos += 8 * dim
        size_in_mem, size_on_disk = struct.unpack_from("<QQ", buf, pos)
        pos += 16
        body = buf[pos : pos + size_on_disk]
        pos += size_on_disk
        raw = body if size_in_mem == size_on_disk else zlib.decompress(body)
        lo = [lim[2 * d] for d in range(dim)]
        hi = [lim[2 * d + 1] for d in range(dim)]
        bext = [hi[d] - lo[d] for d in range(dim)]
        if vector:
            rec = np.frombuffer(raw, dtype=[("n", "<i4"), ("v", dt, (fields,))])
            blk = rec["v"].reshape(bext + [fields])
        else:
            blk = np.frombuffer(raw, dtype=dt).reshape(bext)
        sl = tuple(slice(lo[d] - g0[d], hi[d] - g0[d]) for d in range(dim))
        data[sl] = blk
        for d in range(dim):
            if lo[d] == g0[d]:
                b0[d] = bcs[2 * d]
            if hi[d] == g1[d]:
                b1[d] = bcs[2 * d + 1]
    return DatGrid(data=data, vector=vector, g0=g0, b0=b0, b1=b1, dx=dx, hdr_fields=fields)
