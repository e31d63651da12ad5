"""MAOS .bin files through the library's C ABI (mt_bin_*).

A dense matrix is returned as a numpy array of shape (nx, ny) (MAOS stores
it column-major: nx rows change fastest): float64, complex128, float32,
complex64 or int64 as the file holds it; a sparse matrix as a
scipy.sparse.csc_matrix (float64 or complex128); a cell as a numpy object
array of shape (nx, ny) whose elements are any of these or None (empty).
Writing takes the same forms; a 1-D array is written as an n x 1 matrix and
a list as an n x 1 cell.  Paths ending in .gz are compressed.
"""
import ctypes as C

import numpy as np

from . import _abi


def _node_to_py(h):
    L = _abi.lib()
    if not h:
        return None
    is_cell, nx, ny = C.c_int(), C.c_long(), C.c_long()
    _abi.check(L.mt_bin_info(h, C.byref(is_cell), C.byref(nx), C.byref(ny)), "mt_bin_info")
    if is_cell.value:
        out = np.empty((nx.value, ny.value), dtype=object)
        for i in range(nx.value * ny.value):
            out[i % nx.value, i // nx.value] = _node_to_py(L.mt_bin_child(h, i))
        return out
    n = nx.value * ny.value
    t = L.mt_bin_type(h)
    if t in (_abi.MT_BIN_SPARSE, _abi.MT_BIN_CSPARSE):
        import scipy.sparse as sp
        pp, ii, xx, nz = _abi.lp(), _abi.lp(), _abi.dp(), C.c_long()
        _abi.check(L.mt_bin_sparse(h, C.byref(pp), C.byref(ii), C.byref(xx), C.byref(nz)), "mt_bin_sparse")
        cplx = t == _abi.MT_BIN_CSPARSE
        nzv = nz.value
        P = np.ctypeslib.as_array(pp, shape=(ny.value + 1,)).copy() if n else np.zeros(ny.value + 1, np.int64)
        I = np.ctypeslib.as_array(ii, shape=(nzv,)).copy() if nzv else np.zeros(0, np.int64)
        X = np.ctypeslib.as_array(xx, shape=((2 if cplx else 1) * nzv,)).copy() if nzv else np.zeros(0)
        if cplx:
            X = X.view(np.complex128)
        return sp.csc_matrix((X, I, P), shape=(nx.value, ny.value))
    if t == _abi.MT_BIN_INT:
        p = L.mt_bin_idata(h)
        flat = np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.int64)
        return flat.astype(np.int64).reshape((nx.value, ny.value), order="F")
    cplx = t in (_abi.MT_BIN_COMPLEX, _abi.MT_BIN_FCOMPLEX)
    p = L.mt_bin_data(h)
    m = 2 * n if cplx else n
    flat = np.ctypeslib.as_array(p, shape=(m,)).copy() if m else np.zeros(0)
    if cplx:
        flat = flat.view(np.complex128)
    if t == _abi.MT_BIN_FLOAT:
        flat = flat.astype(np.float32)
    elif t == _abi.MT_BIN_FCOMPLEX:
        flat = flat.astype(np.complex64)
    return flat.reshape((nx.value, ny.value), order="F")


def _py_to_node(v, header=""):
    L = _abi.lib()
    h = C.c_void_p()
    hb = header.encode() if header else None
    if isinstance(v, (list, tuple)):
        v = np.array(list(v) + [None], dtype=object)[:-1].reshape(-1, 1)
    if isinstance(v, np.ndarray) and v.dtype == object:
        cells = v.reshape(v.shape[0], -1) if v.ndim else v.reshape(1, 1)
        nx, ny = cells.shape
        _abi.check(L.mt_bin_new_cell(nx, ny, hb, C.byref(h)), "mt_bin_new_cell")
        try:
            for j in range(ny):
                for i in range(nx):
                    e = cells[i, j]
                    if e is not None:
                        ch = _py_to_node(e)
                        _abi.check(L.mt_bin_set(h, i + nx * j, ch), "mt_bin_set")
        except Exception:
            L.mt_bin_free(h)
            raise
        return h
    try:
        import scipy.sparse as sp
        is_sp = sp.issparse(v)
    except ImportError:
        is_sp = False
    if is_sp:
        m = v.tocsc()
        m.sort_indices()
        cplx = np.iscomplexobj(m.data)
        P = np.ascontiguousarray(m.indptr, dtype=np.int64)
        I = np.ascontiguousarray(m.indices, dtype=np.int64)
        X = np.ascontiguousarray(m.data, dtype=np.complex128 if cplx else np.float64)
        Xd = X.view(np.float64) if cplx else X
        _abi.check(L.mt_bin_new_sparse(m.shape[0], m.shape[1], P.ctypes.data_as(_abi.lp),
                                       I.ctypes.data_as(_abi.lp) if I.size else None,
                                       Xd.ctypes.data_as(_abi.dp) if Xd.size else None, int(cplx), hb, C.byref(h)),
                   "mt_bin_new_sparse")
        return h
    a = np.asarray(v)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.ndim != 2:
        raise ValueError("a .bin matrix has at most 2 dimensions")
    kind = {np.dtype(np.complex128): _abi.MT_BIN_COMPLEX, np.dtype(np.float32): _abi.MT_BIN_FLOAT,
            np.dtype(np.complex64): _abi.MT_BIN_FCOMPLEX}.get(a.dtype)
    if kind is None and np.issubdtype(a.dtype, np.integer):
        kind = _abi.MT_BIN_INT
    if kind is None:
        a = a.astype(np.float64)
        flat = np.ascontiguousarray(a.reshape(-1, order="F"))
        _abi.check(L.mt_bin_new_dbl(a.shape[0], a.shape[1], flat.ctypes.data_as(_abi.dp) if flat.size else None,
                                    hb, C.byref(h)), "mt_bin_new_dbl")
        return h
    if kind in (_abi.MT_BIN_COMPLEX, _abi.MT_BIN_FCOMPLEX):
        flat = np.ascontiguousarray(a.reshape(-1, order="F").astype(np.complex128)).view(np.float64)
    else:
        flat = np.ascontiguousarray(a.reshape(-1, order="F").astype(np.float64))
    _abi.check(L.mt_bin_new_typed(kind, a.shape[0], a.shape[1], flat.ctypes.data_as(_abi.dp) if flat.size else None,
                                  hb, C.byref(h)), "mt_bin_new_typed")
    return h


def read(path, with_header=False):
    """The array or cell stored in a MAOS .bin / .bin.gz file (and its header
    text when with_header)."""
    L = _abi.lib()
    h = C.c_void_p()
    _abi.check(L.mt_bin_read(str(path).encode(), C.byref(h)), "mt_bin_read")
    try:
        v = _node_to_py(h)
        hdr = (L.mt_bin_header(h) or b"").decode()
    finally:
        L.mt_bin_free(h)
    return (v, hdr) if with_header else v


def write(path, value, header=""):
    """Write an array (or a cell: object array / list) as MAOS does."""
    L = _abi.lib()
    h = _py_to_node(value, header)
    try:
        _abi.check(L.mt_bin_write(str(path).encode(), h), "mt_bin_write")
    finally:
        L.mt_bin_free(h)


def read_grad(path, recon):
    """Gradient vector of gpu_tomo_solve from a MAOS gradient cell (nwfs x 1
    of 2*nsa x 1: x slopes then y slopes)."""
    L = _abi.lib()
    h = C.c_void_p()
    _abi.check(L.mt_bin_read(str(path).encode(), C.byref(h)), "mt_bin_read")
    try:
        g = np.zeros(recon.ng)
        _abi.check(L.mt_bin_grad(h, C.byref(recon.rec), g.ctypes.data_as(_abi.dp)), "mt_bin_grad")
    finally:
        L.mt_bin_free(h)
    return g


def write_opdr(path, recon, x, header=""):
    """A solution as MAOS's GPU path holds opdr: npsr x 1 cell of xnx x xny."""
    L = _abi.lib()
    h = C.c_void_p()
    xx = np.ascontiguousarray(x, dtype=np.float64)
    _abi.check(L.mt_bin_opdr(C.byref(recon.rec), xx.ctypes.data_as(_abi.dp), C.byref(h)), "mt_bin_opdr")
    try:
        if header:
            # the header of the cell itself: rebuild through the Python form
            v = _node_to_py(h)
            L.mt_bin_free(h)
            h = _py_to_node(v, header)
        _abi.check(L.mt_bin_write(str(path).encode(), h), "mt_bin_write")
    finally:
        L.mt_bin_free(h)


def load_gp(recon, path):
    """load.GP: replace the recon's GP operators with the sparse matrices of
    a MAOS GP.bin (one per WFS or per powfs; mt_setup_load_gp)."""
    L = _abi.lib()
    h = C.c_void_p()
    _abi.check(L.mt_bin_read(str(path).encode(), C.byref(h)), "mt_bin_read")
    try:
        _abi.check(L.mt_setup_load_gp(recon._h, h), "mt_setup_load_gp")
    finally:
        L.mt_bin_free(h)
    recon._refresh()
