"""Minimal HDF5 reader/writer over ctypes for the `.up` file format.

The reference writes and reads `.up` files with PyTables (py/upside_config.py,
py/upside_engine.py:186-192) and the HDF5 C API (src/h5_support.cpp).  Neither
h5py nor PyTables is available here, so this module binds the handful of HDF5
C calls the `.up` layout needs: groups, n-d numeric datasets, fixed-length
string datasets and attributes, numeric scalar attributes.

Conventions that the reference's loader depends on and that this writer keeps:
  * node arguments are a 1-d fixed-length string attribute `arguments`
    (read by read_attribute<vector<string>>, src/h5_support.cpp:75-100);
  * integer attributes are int64, float attributes float64, as PyTables writes
    them (reads convert to the requested native type, src/h5_support.cpp:57-73).
"""
import ctypes as ct
import os
import numpy as np

_H5_CANDIDATES = [
    os.environ.get("UPSIDE_HDF5_LIB", ""),
    "/opt/conda/lib/libhdf5.so.103",
    "libhdf5.so.103",
    "libhdf5.so",
]


def _load():
    last = None
    for p in _H5_CANDIDATES:
        if not p:
            continue
        try:
            return ct.CDLL(p)
        except OSError as e:  # try the next candidate
            last = e
    raise OSError("libhdf5 not found (set UPSIDE_HDF5_LIB): %s" % last)


_h = _load()
hid_t = ct.c_int64
herr_t = ct.c_int
hsize_t = ct.c_uint64

for name, res, args in [
    ("H5open", herr_t, []),
    ("H5Fopen", hid_t, [ct.c_char_p, ct.c_uint, hid_t]),
    ("H5Fcreate", hid_t, [ct.c_char_p, ct.c_uint, hid_t, hid_t]),
    ("H5Fclose", herr_t, [hid_t]),
    ("H5Gopen2", hid_t, [hid_t, ct.c_char_p, hid_t]),
    ("H5Gcreate2", hid_t, [hid_t, ct.c_char_p, hid_t, hid_t, hid_t]),
    ("H5Gclose", herr_t, [hid_t]),
    ("H5Dopen2", hid_t, [hid_t, ct.c_char_p, hid_t]),
    ("H5Dcreate2", hid_t, [hid_t, ct.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
    ("H5Dget_space", hid_t, [hid_t]),
    ("H5Dget_type", hid_t, [hid_t]),
    ("H5Dread", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ct.c_void_p]),
    ("H5Dwrite", herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ct.c_void_p]),
    ("H5Dclose", herr_t, [hid_t]),
    ("H5Screate", hid_t, [ct.c_int]),
    ("H5Screate_simple", hid_t, [ct.c_int, ct.POINTER(hsize_t), ct.POINTER(hsize_t)]),
    ("H5Sget_simple_extent_ndims", ct.c_int, [hid_t]),
    ("H5Sget_simple_extent_dims", ct.c_int, [hid_t, ct.POINTER(hsize_t), ct.POINTER(hsize_t)]),
    ("H5Sget_simple_extent_type", ct.c_int, [hid_t]),
    ("H5Sclose", herr_t, [hid_t]),
    ("H5Tcopy", hid_t, [hid_t]),
    ("H5Tset_size", herr_t, [hid_t, ct.c_size_t]),
    ("H5Tget_size", ct.c_size_t, [hid_t]),
    ("H5Tget_class", ct.c_int, [hid_t]),
    ("H5Tget_sign", ct.c_int, [hid_t]),
    ("H5Tis_variable_str", ct.c_int, [hid_t]),
    ("H5Tclose", herr_t, [hid_t]),
    ("H5Aopen_by_name", hid_t, [hid_t, ct.c_char_p, ct.c_char_p, hid_t, hid_t]),
    ("H5Aexists_by_name", ct.c_int, [hid_t, ct.c_char_p, ct.c_char_p, hid_t]),
    ("H5Acreate2", hid_t, [hid_t, ct.c_char_p, hid_t, hid_t, hid_t, hid_t]),
    ("H5Adelete", herr_t, [hid_t, ct.c_char_p]),
    ("H5Aget_space", hid_t, [hid_t]),
    ("H5Aget_type", hid_t, [hid_t]),
    ("H5Aread", herr_t, [hid_t, hid_t, ct.c_void_p]),
    ("H5Awrite", herr_t, [hid_t, hid_t, ct.c_void_p]),
    ("H5Aclose", herr_t, [hid_t]),
    ("H5Lexists", ct.c_int, [hid_t, ct.c_char_p, hid_t]),
    ("H5Lget_name_by_idx", ct.c_ssize_t, [hid_t, ct.c_char_p, ct.c_int, ct.c_int, hsize_t,
                                          ct.c_char_p, ct.c_size_t, hid_t]),
    ("H5Gget_info_by_name", herr_t, [hid_t, ct.c_char_p, ct.c_void_p, hid_t]),
    ("H5Oget_info_by_name", herr_t, [hid_t, ct.c_char_p, ct.c_void_p, hid_t]),
    ("H5Eset_auto2", herr_t, [hid_t, ct.c_void_p, ct.c_void_p]),
]:
    f = getattr(_h, name)
    f.restype = res
    f.argtypes = args

_h.H5open()
_h.H5Eset_auto2(0, None, None)  # errors are reported through return codes


def _g(name):
    return hid_t.in_dll(_h, name).value


T_F32 = _g("H5T_NATIVE_FLOAT_g")
T_F64 = _g("H5T_NATIVE_DOUBLE_g")
T_I32 = _g("H5T_NATIVE_INT_g")
T_I64 = _g("H5T_NATIVE_LLONG_g")
T_U8 = _g("H5T_NATIVE_UCHAR_g")
T_I8 = _g("H5T_NATIVE_SCHAR_g")
T_C_S1 = _g("H5T_C_S1_g")
# file formats: little-endian standard types, as PyTables writes
T_STD_F32 = _g("H5T_IEEE_F32LE_g")
T_STD_F64 = _g("H5T_IEEE_F64LE_g")
T_STD_I32 = _g("H5T_STD_I32LE_g")
T_STD_I64 = _g("H5T_STD_I64LE_g")
T_STD_U8 = _g("H5T_STD_U8LE_g")

_H5F_ACC_RDONLY, _H5F_ACC_RDWR, _H5F_ACC_TRUNC = 0, 1, 2
_H5T_INTEGER, _H5T_FLOAT, _H5T_STRING = 0, 1, 3
_H5S_SCALAR, _H5S_SIMPLE, _H5S_NULL = 0, 1, 2

_NATIVE = {
    np.dtype("float32"): (T_F32, T_STD_F32),
    np.dtype("float64"): (T_F64, T_STD_F64),
    np.dtype("int32"): (T_I32, T_STD_I32),
    np.dtype("int64"): (T_I64, T_STD_I64),
    np.dtype("uint8"): (T_U8, T_STD_U8),
    np.dtype("bool"): (T_U8, T_STD_U8),
}


def _chk(x, what):
    if x < 0:
        raise IOError("HDF5 call failed: %s" % what)
    return x


class _Obj:
    def __init__(self, hid, closer):
        self.hid, self.closer = hid, closer

    def close(self):
        if self.hid is not None and self.hid >= 0:
            self.closer(self.hid)
        self.hid = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dims(space):
    nd = _h.H5Sget_simple_extent_ndims(space)
    if nd <= 0:
        return ()
    d = (hsize_t * nd)()
    _h.H5Sget_simple_extent_dims(space, d, None)
    return tuple(int(x) for x in d)


def _mem_type_for(ftype):
    cls = _h.H5Tget_class(ftype)
    size = _h.H5Tget_size(ftype)
    if cls == _H5T_FLOAT:
        return (T_F32, np.float32) if size == 4 else (T_F64, np.float64)
    if cls == _H5T_INTEGER:
        signed = _h.H5Tget_sign(ftype) != 0
        if size == 1:
            return (T_I8, np.int8) if signed else (T_U8, np.uint8)
        if size <= 4:
            return (T_I32, np.int32)
        return (T_I64, np.int64)
    raise TypeError("unsupported HDF5 type class %d" % cls)


class File(_Obj):
    """HDF5 file handle.  mode 'r' (read), 'a' (read/write), 'w' (truncate)."""

    def __init__(self, path, mode="r"):
        p = os.fsencode(path)
        if mode == "r":
            hid = _h.H5Fopen(p, _H5F_ACC_RDONLY, 0)
        elif mode == "a":
            hid = _h.H5Fopen(p, _H5F_ACC_RDWR, 0)
        elif mode == "w":
            hid = _h.H5Fcreate(p, _H5F_ACC_TRUNC, 0, 0)
        else:
            raise ValueError(mode)
        _chk(hid, "open %s" % path)
        super().__init__(hid, _h.H5Fclose)

    # ---- structure -------------------------------------------------------
    def exists(self, path):
        parts = [x for x in path.strip("/").split("/") if x]
        cur = ""
        for p in parts:
            cur += "/" + p
            if _h.H5Lexists(self.hid, cur.encode(), 0) <= 0:
                return False
        return True

    def children(self, path):
        """Names of the links in a group, in name order (H5_INDEX_NAME)."""
        names = []
        i = 0
        while True:
            n = _h.H5Lget_name_by_idx(self.hid, path.encode(), 0, 0, i, None, 0, 0)
            if n < 0:
                break
            buf = ct.create_string_buffer(n + 1)
            _h.H5Lget_name_by_idx(self.hid, path.encode(), 0, 0, i, buf, n + 1, 0)
            names.append(buf.value.decode())
            i += 1
        return names

    def is_group(self, path):
        g = _h.H5Gopen2(self.hid, path.encode(), 0)
        if g < 0:
            return False
        _h.H5Gclose(g)
        return True

    def create_group(self, path):
        parts = [x for x in path.strip("/").split("/") if x]
        cur = ""
        for p in parts:
            cur += "/" + p
            if _h.H5Lexists(self.hid, cur.encode(), 0) > 0:
                continue
            g = _chk(_h.H5Gcreate2(self.hid, cur.encode(), 0, 0, 0), "create group " + cur)
            _h.H5Gclose(g)

    # ---- datasets --------------------------------------------------------
    def read(self, path):
        d = _chk(_h.H5Dopen2(self.hid, path.encode(), 0), "open dataset " + path)
        try:
            sp = _h.H5Dget_space(d)
            shape = _dims(sp)
            _h.H5Sclose(sp)
            ft = _h.H5Dget_type(d)
            try:
                if _h.H5Tget_class(ft) == _H5T_STRING:
                    if _h.H5Tis_variable_str(ft) > 0:
                        raise TypeError("variable-length strings not supported")
                    size = _h.H5Tget_size(ft)
                    mt = _h.H5Tcopy(T_C_S1)
                    _h.H5Tset_size(mt, size)
                    buf = np.zeros(shape, dtype="S%d" % size)
                    _chk(_h.H5Dread(d, mt, 0, 0, 0, buf.ctypes.data), "read " + path)
                    _h.H5Tclose(mt)
                    return buf
                mt, npt = _mem_type_for(ft)
            finally:
                _h.H5Tclose(ft)
            out = np.zeros(shape, dtype=npt)
            if out.size:
                _chk(_h.H5Dread(d, mt, 0, 0, 0, out.ctypes.data), "read " + path)
            return out
        finally:
            _h.H5Dclose(d)

    def write(self, path, arr):
        """Create dataset `path` holding `arr` (parent groups are created)."""
        parent = path.rsplit("/", 1)[0]
        if parent:
            self.create_group(parent)
        arr = np.asarray(arr)
        if arr.dtype.kind == "S" or arr.dtype.kind == "U":
            arr = np.asarray(arr, dtype="S")
            size = max(arr.dtype.itemsize, 1)
            arr = np.ascontiguousarray(arr.astype("S%d" % size))
            mt = _h.H5Tcopy(T_C_S1)
            _h.H5Tset_size(mt, size)
            ft = mt
        else:
            if arr.dtype == np.float16:
                arr = arr.astype(np.float32)
            arr = np.ascontiguousarray(arr)
            if arr.dtype not in _NATIVE:
                if arr.dtype.kind in "iu":
                    arr = arr.astype(np.int64)
                else:
                    arr = arr.astype(np.float64)
            mt, ft = _NATIVE[arr.dtype]
            if arr.dtype == np.bool_:
                arr = arr.astype(np.uint8)
        dims = (hsize_t * max(arr.ndim, 1))(*(arr.shape if arr.ndim else (1,)))
        sp = _h.H5Screate_simple(arr.ndim, dims, None) if arr.ndim else _h.H5Screate(_H5S_SCALAR)
        d = _chk(_h.H5Dcreate2(self.hid, path.encode(), ft, sp, 0, 0, 0), "create dataset " + path)
        if arr.size:
            _chk(_h.H5Dwrite(d, mt, 0, 0, 0, arr.ctypes.data), "write " + path)
        _h.H5Dclose(d)
        _h.H5Sclose(sp)
        if arr.dtype.kind == "S":
            _h.H5Tclose(mt)

    # ---- attributes ------------------------------------------------------
    def has_attr(self, path, name):
        return _h.H5Aexists_by_name(self.hid, path.encode(), name.encode(), 0) > 0

    def attr(self, path, name, default=None):
        if not self.has_attr(path, name):
            if default is not None:
                return default
            raise KeyError("%s@%s" % (path, name))
        a = _chk(_h.H5Aopen_by_name(self.hid, path.encode(), name.encode(), 0, 0), "attr " + name)
        try:
            sp = _h.H5Aget_space(a)
            stype = _h.H5Sget_simple_extent_type(sp)
            shape = _dims(sp) if stype == _H5S_SIMPLE else ()
            _h.H5Sclose(sp)
            if stype == _H5S_NULL:
                return None
            ft = _h.H5Aget_type(a)
            try:
                if _h.H5Tget_class(ft) == _H5T_STRING:
                    size = _h.H5Tget_size(ft)
                    mt = _h.H5Tcopy(T_C_S1)
                    _h.H5Tset_size(mt, size)
                    buf = np.zeros(shape, dtype="S%d" % size)
                    _chk(_h.H5Aread(a, mt, buf.ctypes.data), "read attr " + name)
                    _h.H5Tclose(mt)
                    if shape == ():
                        return buf[()].decode()
                    return [x.decode() for x in buf.reshape(-1)]
                mt, npt = _mem_type_for(ft)
            finally:
                _h.H5Tclose(ft)
            out = np.zeros(shape, dtype=npt)
            _chk(_h.H5Aread(a, mt, out.ctypes.data), "read attr " + name)
            return out[()] if shape == () else out
        finally:
            _h.H5Aclose(a)

    def set_attr(self, path, name, value):
        if self.has_attr(path, name):
            o = _h.H5Gopen2(self.hid, path.encode(), 0)
            _h.H5Adelete(o, name.encode())
            _h.H5Gclose(o)
        if isinstance(value, (list, tuple)) and all(isinstance(v, str) for v in value):
            size = max([len(v) for v in value] + [1])
            buf = np.array([v.encode() for v in value], dtype="S%d" % size)
            mt = _h.H5Tcopy(T_C_S1)
            _h.H5Tset_size(mt, size)
            dims = (hsize_t * 1)(len(value))
            sp = _h.H5Screate_simple(1, dims, None)
            a = _chk(_h.H5Acreate2(self._obj(path), name.encode(), mt, sp, 0, 0), "attr " + name)
            if len(value):
                _chk(_h.H5Awrite(a, mt, buf.ctypes.data), "write attr " + name)
            _h.H5Aclose(a)
            _h.H5Sclose(sp)
            _h.H5Tclose(mt)
            self._close_obj()
            return
        if isinstance(value, str):
            buf = ct.create_string_buffer(value.encode(), max(len(value), 1))
            mt = _h.H5Tcopy(T_C_S1)
            _h.H5Tset_size(mt, max(len(value), 1))
            sp = _h.H5Screate(_H5S_SCALAR)
            a = _chk(_h.H5Acreate2(self._obj(path), name.encode(), mt, sp, 0, 0), "attr " + name)
            _chk(_h.H5Awrite(a, mt, buf), "write attr " + name)
            _h.H5Aclose(a)
            _h.H5Sclose(sp)
            _h.H5Tclose(mt)
            self._close_obj()
            return
        arr = np.asarray(value)
        if arr.dtype == np.bool_:
            arr = arr.astype(np.uint8)
        elif arr.dtype.kind in "iu":
            arr = arr.astype(np.int64)
        else:
            arr = arr.astype(np.float64)
        arr = np.ascontiguousarray(arr)
        mt, ft = _NATIVE[arr.dtype]
        if arr.ndim:
            dims = (hsize_t * arr.ndim)(*arr.shape)
            sp = _h.H5Screate_simple(arr.ndim, dims, None)
        else:
            sp = _h.H5Screate(_H5S_SCALAR)
        a = _chk(_h.H5Acreate2(self._obj(path), name.encode(), ft, sp, 0, 0), "attr " + name)
        _chk(_h.H5Awrite(a, mt, arr.ctypes.data), "write attr " + name)
        _h.H5Aclose(a)
        _h.H5Sclose(sp)
        self._close_obj()

    def _obj(self, path):
        self._open = _h.H5Gopen2(self.hid, path.encode(), 0)
        if self._open < 0:
            self._open = _chk(_h.H5Dopen2(self.hid, path.encode(), 0), "open " + path)
            self._open_is_dset = True
        else:
            self._open_is_dset = False
        return self._open

    def _close_obj(self):
        if self._open_is_dset:
            _h.H5Dclose(self._open)
        else:
            _h.H5Gclose(self._open)
