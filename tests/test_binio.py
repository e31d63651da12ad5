"""MAOS .bin files through the library (mt_bin_* C ABI, maos_amd.binio).

Pinned against files MAOS itself wrote (tests/golden/maos_bin, copied from
the reference's data by tests/golden/copy_maos_bin.py): their values against
an independent parse with struct, and their bytes against what the library
writes back; and against byte strings built by hand from the layout rules of
sys/bin.c (write_header, write_bin_headerstr) and math/cell.c
(writedata_by_id) for cells, header text and empty elements.
"""
import gzip
import os
import struct

import numpy as np
import pytest

from maos_amd import _abi, binio, configs, tomo

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "maos_bin")
FILES = ["hyst10p500nm.bin", "PSD_TMT_WindShake75mas_rad2.bin", "NapDelta.bin"]
M_DBL, MCC_ANY, M_HEADER, M_SKIP = 0x6402, 0x6421, 0x6500, 0x6600


def raw_bytes(path):
    b = open(path, "rb").read()
    return gzip.decompress(b) if b[:2] == b"\x1f\x8b" else b


def parse_dbl(b):
    """Independent reader for a file holding one double matrix: header
    blocks, skip words, then magic, nx, ny and the data."""
    pos, hdr = 0, ""
    while True:
        (m,) = struct.unpack_from("<I", b, pos)
        if m == M_HEADER:
            (n,) = struct.unpack_from("<Q", b, pos + 4)
            hdr += b[pos + 12:pos + 12 + n].split(b"\0")[0].decode()
            pos += 12 + n + 12
        elif (m & M_SKIP) == M_SKIP:
            pos += 4
        else:
            break
    assert m == M_DBL
    nx, ny = struct.unpack_from("<QQ", b, pos + 4)
    data = np.frombuffer(b, dtype="<f8", count=nx * ny, offset=pos + 20)
    return data.reshape((nx, ny), order="F"), hdr


@pytest.mark.parametrize("name", FILES)
def test_reads_maos_files(name):
    v, hdr = binio.read(os.path.join(GOLD, name), with_header=True)
    ref, rhdr = parse_dbl(raw_bytes(os.path.join(GOLD, name)))
    assert v.shape == ref.shape and np.array_equal(v, ref)
    assert hdr == rhdr and hdr.startswith("Created by write")


@pytest.mark.parametrize("name", FILES)
def test_rewrite_is_byte_identical(name, tmp_path):
    """What MAOS wrote (header text, M_SKIP, M_DBL, data) is what the library
    writes for the same array and header (uncompressed)."""
    v, hdr = binio.read(os.path.join(GOLD, name), with_header=True)
    out = tmp_path / "out.bin"
    binio.write(out, v, hdr)
    assert open(out, "rb").read() == raw_bytes(os.path.join(GOLD, name))


def test_gzip_round_trip(tmp_path):
    v = np.arange(12.0).reshape(3, 4)
    p = tmp_path / "a.bin.gz"
    binio.write(p, v, "k=1")
    assert open(p, "rb").read()[:2] == b"\x1f\x8b"
    w, hdr = binio.read(p, with_header=True)
    assert np.array_equal(w, v) and hdr == "k=1"


def head(hdr, magic, nx, ny):
    out = b""
    if hdr:
        n = (len(hdr) + 1 + 7) // 8 * 8
        out += struct.pack("<IQ", M_HEADER, n) + hdr.encode().ljust(n, b"\0") + struct.pack("<QI", n, M_HEADER)
    return out + struct.pack("<IIQQ", M_SKIP, magic, nx, ny)


def test_cell_layout_bytes(tmp_path):
    """A 3 x 1 cell: a 2 x 1 matrix with header text, an empty element, a
    nested 1 x 1 cell; the cell itself with header text."""
    inner = np.empty((1, 1), dtype=object)
    inner[0, 0] = np.array([[7.0]])
    cell = np.empty((3, 1), dtype=object)
    cell[0, 0] = np.array([[1.5], [-2.0]])
    cell[1, 0] = None
    cell[2, 0] = inner
    p = tmp_path / "c.bin"
    L = _abi.lib()
    # element headers go through the C ABI directly
    import ctypes as C
    h = C.c_void_p()
    _abi.check(L.mt_bin_new_cell(3, 1, b"cell=yes", C.byref(h)), "new_cell")
    e0 = C.c_void_p()
    d = np.array([1.5, -2.0])
    _abi.check(L.mt_bin_new_dbl(2, 1, d.ctypes.data_as(_abi.dp), b"first", C.byref(e0)), "new_dbl")
    _abi.check(L.mt_bin_set(h, 0, e0), "set")
    _abi.check(L.mt_bin_set(h, 2, binio._py_to_node(inner)), "set")
    _abi.check(L.mt_bin_write(str(p).encode(), h), "write")
    L.mt_bin_free(h)
    want = (head("cell=yes", MCC_ANY, 3, 1)
            + head("first", M_DBL, 2, 1) + struct.pack("<2d", 1.5, -2.0)
            + head("", MCC_ANY, 0, 0)
            + head("", MCC_ANY, 1, 1) + head("", M_DBL, 1, 1) + struct.pack("<d", 7.0))
    assert open(p, "rb").read() == want
    back = binio.read(p)
    assert back.shape == (3, 1) and back[1, 0] is None
    assert np.array_equal(back[0, 0], cell[0, 0]) and back[2, 0][0, 0][0, 0] == 7.0


def test_empty_cell_is_written_empty(tmp_path):
    p = tmp_path / "e.bin"
    binio.write(p, [None, None])
    assert open(p, "rb").read() == head("", MCC_ANY, 0, 0)
    assert binio.read(p).shape == (0, 0)


def test_other_element_types(tmp_path):
    """The element types of sys/bin.h:34-51 and the sparse layout of
    math/spbin.c:32-89, from records built byte by byte here: single float,
    complex double, 64- and 32-bit integers, sparse double with 64- and
    32-bit indices, sparse complex; our writer gives back the same bytes."""
    import scipy.sparse as sp
    p = tmp_path / "f.bin"
    # M_FLT 2 x 1
    open(p, "wb").write(head("", 0x6408, 2, 1) + struct.pack("<2f", 1.5, -2))
    a = binio.read(p)
    assert a.dtype == np.float32 and a.shape == (2, 1) and list(a[:, 0]) == [1.5, -2]
    # M_CMP 1 x 2
    open(p, "wb").write(head("", 0x6404, 1, 2) + struct.pack("<4d", 1, 2, -3, 4))
    a = binio.read(p)
    assert a.dtype == np.complex128 and np.array_equal(a, np.array([[1 + 2j, -3 + 4j]]))
    # M_INT64 / M_INT32 3 x 1
    for magic, fmt in ((0x6403, "<3q"), (0x6405, "<3i")):
        open(p, "wb").write(head("", magic, 3, 1) + struct.pack(fmt, 7, -8, 9))
        a = binio.read(p)
        assert a.dtype == np.int64 and list(a[:, 0]) == [7, -8, 9]
    # sparse double 3 x 2: (0,0)=1.5 (2,0)=-2 (1,1)=4
    P, I, X = [0, 2, 3], [0, 2, 1], [1.5, -2.0, 4.0]
    want = sp.csc_matrix((X, I, P), shape=(3, 2))
    rec64 = head("", 0x6401, 3, 2) + struct.pack("<Q3q3q3d", 3, *P, *I, *X)
    open(p, "wb").write(rec64)
    got = binio.read(p)
    assert sp.issparse(got) and (got != want).nnz == 0 and got.dtype == np.float64
    rec32 = head("", 0x6407, 3, 2) + struct.pack("<Q3i3i3d", 3, *P, *I, *X)
    open(p, "wb").write(rec32)
    assert (binio.read(p) != want).nnz == 0
    q = tmp_path / "w.bin"
    binio.write(q, want)
    assert open(q, "rb").read() == rec64
    # sparse complex
    Xc = [1 + 1j, 2, -3j]
    recc = head("", 0x6400, 3, 2) + struct.pack("<Q3q3q6d", 3, *P, *I, 1, 1, 2, 0, 0, -3)
    open(p, "wb").write(recc)
    got = binio.read(p)
    assert got.dtype == np.complex128 and np.array_equal(got.toarray(), sp.csc_matrix((Xc, I, P), shape=(3, 2)).toarray())
    binio.write(q, got)
    assert open(q, "rb").read() == recc
    # round trips of every type, with a header and inside a cell
    cell = [np.arange(4.0).reshape(2, 2), np.array([[1 + 2j]]), np.arange(3, dtype=np.float32),
            np.arange(5).reshape(5, 1), want, None]
    binio.write(tmp_path / "c.bin.gz", cell, header="k=v")
    back, hdr = binio.read(tmp_path / "c.bin.gz", with_header=True)
    assert hdr == "k=v"
    assert np.array_equal(back[0, 0], cell[0]) and np.array_equal(back[1, 0], cell[1])
    assert back[2, 0].dtype == np.float32 and np.array_equal(back[2, 0][:, 0], cell[2])
    assert np.array_equal(back[3, 0], cell[3]) and (back[4, 0] != want).nnz == 0 and back[5, 0] is None


def test_refuses_unknown_types_and_truncation(tmp_path):
    p = tmp_path / "f.bin"
    open(p, "wb").write(struct.pack("<IIQQ", M_SKIP, 0x640A, 2, 1) + struct.pack("<2b", 1, 2))  # M_INT8
    with pytest.raises(_abi.MtError):
        binio.read(p)
    open(p, "wb").write(head("", M_DBL, 4, 1) + struct.pack("<2d", 1, 2))  # 2 of 4 values
    with pytest.raises(_abi.MtError):
        binio.read(p)
    # sparse row index out of range
    open(p, "wb").write(head("", 0x6401, 3, 2) + struct.pack("<Q3q3q3d", 3, 0, 2, 3, 0, 5, 1, 1, 2, 3))
    with pytest.raises(_abi.MtError):
        binio.read(p)
    with pytest.raises(_abi.MtError):
        binio.read(tmp_path / "missing.bin")


def test_load_gp(tmp_path):
    """load.GP: the recon's own GP written as MAOS's GP.bin (a cell of sparse
    matrices, per powfs or per WFS) and loaded back gives the same
    operators; a scaled GP scales the right-hand side accordingly."""
    import scipy.sparse as sp
    from oracle.oracle import Oracle
    cfg = configs.small8()
    r = tomo.Recon(cfg)
    grad = np.random.default_rng(2).standard_normal(r.ng)
    rhs = Oracle(r).tomo_r(grad)
    gps = [r.gp(w) for w in range(r.nwfs)]
    binio.write(tmp_path / "GP.bin", [gps[0]])              # one per powfs
    r2 = tomo.Recon(cfg)
    binio.load_gp(r2, tmp_path / "GP.bin")
    for w in range(r.nwfs):
        assert (r2.gp(w) != gps[w]).nnz == 0
    assert np.array_equal(Oracle(r2).tomo_r(grad), rhs)
    binio.write(tmp_path / "GPw.bin", [2.0 * g for g in gps])  # one per WFS
    r3 = tomo.Recon(cfg)
    binio.load_gp(r3, tmp_path / "GPw.bin")
    assert np.allclose(Oracle(r3).tomo_r(grad), 2.0 * rhs, rtol=1e-12, atol=0)
    # the wrong shape, or WFS of one powfs with different GP, are refused
    binio.write(tmp_path / "bad.bin", [sp.csc_matrix((3, 3))])
    with pytest.raises(_abi.MtError):
        binio.load_gp(tomo.Recon(cfg), tmp_path / "bad.bin")
    binio.write(tmp_path / "mix.bin", [gps[0], 2.0 * gps[1], gps[2], gps[3]])
    with pytest.raises(_abi.MtError):
        binio.load_gp(tomo.Recon(cfg), tmp_path / "mix.bin")


def test_grad_and_opdr_layouts(tmp_path):
    """Gradients as MAOS's gradlastcl (nwfs x 1 cell of 2 nsa x 1) into the
    solve's vector, and a solution into opdr (npsr x 1 cell of xnx x xny)."""
    r = tomo.Recon(configs.small8())
    rng = np.random.default_rng(3)
    g = rng.standard_normal(r.ng)
    offs = r.grad_offsets()
    cells = [g[offs[w]:offs[w + 1]] for w in range(r.nwfs)]
    p = tmp_path / "gradlastcl.bin"
    binio.write(p, cells)
    assert np.array_equal(binio.read_grad(p, r), g)
    x = rng.standard_normal(r.nx)
    q = tmp_path / "opdr.bin.gz"
    binio.write_opdr(q, r, x, "seed=3")
    back, hdr = binio.read(q, with_header=True)
    assert hdr == "seed=3" and back.shape == (r.npsr, 1)
    off = r.layer_offsets()
    for l in range(r.npsr):
        assert back[l, 0].shape == (r.rec.xnx[l], r.rec.xny[l])  # layer vector index ix + iy * xnx
        assert np.array_equal(back[l, 0].reshape(-1, order="F"), x[off[l]:off[l + 1]])
    # a gradient cell of the wrong length for a WFS is refused
    cells[0] = cells[0][:-2]
    binio.write(p, cells)
    with pytest.raises(_abi.MtError):
        binio.read_grad(p, r)
