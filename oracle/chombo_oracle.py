"""Checker for the Chombo-format HDF5 output: the reference's writer restated over an installed
HDF5 library, and a reader that dumps any HDF5 file through that library.

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product package.  The product
writes its files with its own HDF5 writer (sayakaamr_amd/csrc/host/h5lite.cpp); this module
uses the HDF5 library of the image (libhdf5 1.10 under /opt/conda/lib, found at run time;
absent -> HDF5 is None and the tests that need it skip) as an independent implementation of
the format, both to write the reference's file and to read the product's back.

What it restates (file:line in the reference checkout's src/):
  WriteDataHdf5                 SayakaWriterHdf5.cpp:351-530   levels 1..currentFinestLevel;
                                dx per level from block 0, halved; problem box = union of block
                                boxes; per level one "boxes" hyperslab per block and one
                                "data:datatype=0" hyperslab per block and component
  write_header_hdf5             SayakaWriterHdf5.cpp:80-349    root, /Chombo_global and
                                /level_<l> attributes, dataset creation, data_attributes
  realvect_t / prob_domain_t    SayakaWriterHdf5.cpp:20-77     compound types
  WriterMeta::guessName/reset   SayakaWriter.h:27-47

Pinning: the boxes must map back to the blocks' own corners (box * dx + prob_lo == block lo,
(box_hi + 1) * dx + prob_lo == block hi), the data back to the state's valid cells, and the
attributes to the tree's geometry (tests/test_chombo_output.py) -- checks that a dropped term or
a transposed index in this restatement fails.
"""
from __future__ import annotations

import ctypes as C
import glob
import os

import numpy as np

hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_uint64


def _find_lib():
    for pat in ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.*", "/usr/lib/x86_64-linux-gnu/libhdf5*.so*"):
        for p in sorted(glob.glob(pat)):
            if "_hl" in os.path.basename(p) or "fortran" in p or "cpp" in os.path.basename(p):
                continue
            try:
                return C.CDLL(p)
            except OSError:
                continue
    return None


class _H5:
    """The few libhdf5 calls the writer and the reader need (HDF5 1.10 C API)."""

    P_DEFAULT = 0
    S_ALL = 0
    F_ACC_RDONLY, F_ACC_TRUNC = 0, 2
    S_SCALAR, S_SIMPLE = 0, 1
    S_SELECT_SET = 0
    T_INTEGER, T_FLOAT, T_STRING, T_COMPOUND = 0, 1, 3, 6
    T_COMPOUND_CLASS = 6
    INDEX_NAME, ITER_INC = 0, 0
    I_GROUP, I_DATASET = 2, 5

    def __init__(self, lib):
        self.lib = lib
        L = lib
        sig = {
            "H5open": ([], herr_t),
            "H5Eset_auto2": ([hid_t, C.c_void_p, C.c_void_p], herr_t),
            "H5Fcreate": ([C.c_char_p, C.c_uint, hid_t, hid_t], hid_t),
            "H5Fopen": ([C.c_char_p, C.c_uint, hid_t], hid_t),
            "H5Fclose": ([hid_t], herr_t),
            "H5Gcreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t], hid_t),
            "H5Gopen2": ([hid_t, C.c_char_p, hid_t], hid_t),
            "H5Gclose": ([hid_t], herr_t),
            "H5Screate": ([C.c_int], hid_t),
            "H5Screate_simple": ([C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)], hid_t),
            "H5Sclose": ([hid_t], herr_t),
            "H5Sselect_hyperslab": ([hid_t, C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t), C.POINTER(hsize_t),
                                     C.POINTER(hsize_t)], herr_t),
            "H5Sget_simple_extent_type": ([hid_t], C.c_int),
            "H5Sget_simple_extent_ndims": ([hid_t], C.c_int),
            "H5Sget_simple_extent_dims": ([hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)], C.c_int),
            "H5Acreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t], hid_t),
            "H5Awrite": ([hid_t, hid_t, C.c_void_p], herr_t),
            "H5Aread": ([hid_t, hid_t, C.c_void_p], herr_t),
            "H5Aclose": ([hid_t], herr_t),
            "H5Aget_num_attrs": ([hid_t], C.c_int),
            "H5Aopen_by_idx": ([hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, hid_t, hid_t], hid_t),
            "H5Aget_name": ([hid_t, C.c_size_t, C.c_char_p], C.c_ssize_t),
            "H5Aget_type": ([hid_t], hid_t),
            "H5Aget_space": ([hid_t], hid_t),
            "H5Tcreate": ([C.c_int, C.c_size_t], hid_t),
            "H5Tinsert": ([hid_t, C.c_char_p, C.c_size_t, hid_t], herr_t),
            "H5Tcopy": ([hid_t], hid_t),
            "H5Tset_size": ([hid_t, C.c_size_t], herr_t),
            "H5Tclose": ([hid_t], herr_t),
            "H5Tget_class": ([hid_t], C.c_int),
            "H5Tget_size": ([hid_t], C.c_size_t),
            "H5Tget_sign": ([hid_t], C.c_int),
            "H5Tget_order": ([hid_t], C.c_int),
            "H5Tget_precision": ([hid_t], C.c_size_t),
            "H5Tget_strpad": ([hid_t], C.c_int),
            "H5Tget_cset": ([hid_t], C.c_int),
            "H5Tis_variable_str": ([hid_t], C.c_int),
            "H5Tget_nmembers": ([hid_t], C.c_int),
            "H5Tget_member_name": ([hid_t, C.c_uint], C.c_void_p),
            "H5Tget_member_offset": ([hid_t, C.c_uint], C.c_size_t),
            "H5Tget_member_type": ([hid_t, C.c_uint], hid_t),
            "H5free_memory": ([C.c_void_p], herr_t),
            "H5Dcreate2": ([hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t], hid_t),
            "H5Dopen2": ([hid_t, C.c_char_p, hid_t], hid_t),
            "H5Dclose": ([hid_t], herr_t),
            "H5Dget_space": ([hid_t], hid_t),
            "H5Dget_type": ([hid_t], hid_t),
            "H5Dwrite": ([hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], herr_t),
            "H5Dread": ([hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p], herr_t),
            "H5Gget_num_objs": ([hid_t, C.POINTER(hsize_t)], herr_t),
            "H5Gget_objname_by_idx": ([hid_t, hsize_t, C.c_char_p, C.c_size_t], C.c_ssize_t),
            "H5Oopen": ([hid_t, C.c_char_p, hid_t], hid_t),
            "H5Oclose": ([hid_t], herr_t),
            "H5Iget_type": ([hid_t], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        self._chk(L.H5open(), "H5open")
        L.H5Eset_auto2(0, None, None)   # errors come back as return codes, not printed
        self.NATIVE_INT = hid_t.in_dll(L, "H5T_NATIVE_INT_g").value
        self.NATIVE_DOUBLE = hid_t.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
        self.C_S1 = hid_t.in_dll(L, "H5T_C_S1_g").value

    @staticmethod
    def _chk(rc, what):
        if rc < 0:
            raise RuntimeError(f"HDF5: {what} failed")
        return rc

    def __getattr__(self, name):   # H5Xyz(...) with the return code checked
        f = getattr(self.lib, "H5" + name)

        def call(*a):
            return self._chk(f(*a), "H5" + name)
        return call

    # ------------------------------------------------------------------ reading
    def describe_type(self, t):
        cls = self.lib.H5Tget_class(t)
        size = int(self.lib.H5Tget_size(t))
        if cls == self.T_INTEGER:
            return ("int", size, self.lib.H5Tget_sign(t), self.lib.H5Tget_order(t), int(self.lib.H5Tget_precision(t)))
        if cls == self.T_FLOAT:
            return ("float", size, self.lib.H5Tget_order(t), int(self.lib.H5Tget_precision(t)))
        if cls == self.T_STRING:
            return ("string", size, self.lib.H5Tget_strpad(t), self.lib.H5Tget_cset(t),
                    self.lib.H5Tis_variable_str(t))
        if cls == self.T_COMPOUND:
            mem = []
            for i in range(self.lib.H5Tget_nmembers(t)):
                p = self.lib.H5Tget_member_name(t, i)
                name = C.string_at(p).decode()
                self.lib.H5free_memory(p)
                mt = self._chk(self.lib.H5Tget_member_type(t, i), "H5Tget_member_type")
                mem.append((name, int(self.lib.H5Tget_member_offset(t, i)), self.describe_type(mt)))
                self.lib.H5Tclose(mt)
            return ("compound", size, tuple(mem))
        return ("class", cls, size)

    def _attrs(self, obj):
        out = {}
        for i in range(self.lib.H5Aget_num_attrs(obj)):
            a = self._chk(self.lib.H5Aopen_by_idx(obj, b".", self.INDEX_NAME, self.ITER_INC, i, 0, 0),
                          "H5Aopen_by_idx")
            buf = C.create_string_buffer(256)
            self.lib.H5Aget_name(a, 256, buf)
            t = self.lib.H5Aget_type(a)
            s = self.lib.H5Aget_space(a)
            desc = self.describe_type(t)
            space = self.lib.H5Sget_simple_extent_type(s)
            raw = C.create_string_buffer(int(self.lib.H5Tget_size(t)))
            self._chk(self.lib.H5Aread(a, t, raw), "H5Aread")
            out[buf.value.decode()] = (desc, space, raw.raw)
            self.lib.H5Sclose(s)
            self.lib.H5Tclose(t)
            self.lib.H5Aclose(a)
        return out

    def dump(self, path):
        """{path: ("group", attrs, links) | ("dataset", type, dims, raw bytes)} of every object."""
        f = self._chk(self.lib.H5Fopen(path.encode(), self.F_ACC_RDONLY, 0), f"H5Fopen {path}")
        out = {}
        try:
            self._walk(f, "/", out)
        finally:
            self.lib.H5Fclose(f)
        return out

    def _walk(self, f, path, out):
        g = self._chk(self.lib.H5Gopen2(f, path.encode(), 0), f"H5Gopen2 {path}")
        n = hsize_t()
        self._chk(self.lib.H5Gget_num_objs(g, C.byref(n)), "H5Gget_num_objs")
        names = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            self._chk(self.lib.H5Gget_objname_by_idx(g, i, buf, 256), "H5Gget_objname_by_idx")
            names.append(buf.value.decode())
        out[path] = ("group", self._attrs(g), tuple(names))
        for name in names:
            child = (path.rstrip("/") + "/" + name)
            o = self._chk(self.lib.H5Oopen(f, child.encode(), 0), f"H5Oopen {child}")
            kind = self.lib.H5Iget_type(o)
            self.lib.H5Oclose(o)
            if kind == self.I_GROUP:
                self._walk(f, child, out)
            elif kind == self.I_DATASET:
                d = self._chk(self.lib.H5Dopen2(f, child.encode(), 0), f"H5Dopen2 {child}")
                t = self.lib.H5Dget_type(d)
                s = self.lib.H5Dget_space(d)
                nd = self.lib.H5Sget_simple_extent_ndims(s)
                dims = (hsize_t * max(nd, 1))()
                self.lib.H5Sget_simple_extent_dims(s, dims, None)
                dims = tuple(int(dims[i]) for i in range(nd))
                raw = C.create_string_buffer(max(1, int(np.prod(dims)) * int(self.lib.H5Tget_size(t))))
                self._chk(self.lib.H5Dread(d, t, self.S_ALL, self.S_ALL, 0, raw), "H5Dread")
                out[child] = ("dataset", self.describe_type(t), dims,
                              raw.raw[:int(np.prod(dims)) * int(self.lib.H5Tget_size(t))])
                self.lib.H5Sclose(s)
                self.lib.H5Tclose(t)
                self.lib.H5Dclose(d)
            else:
                out[child] = ("other", kind)
        self.lib.H5Gclose(g)


_lib = _find_lib()
HDF5 = _H5(_lib) if _lib is not None else None


# ---------------------------------------------------------------------- the reference's writer
def _scalar_attr(h, loc, name, tid, buf):
    s = h.Screate(h.S_SCALAR)
    a = h.Acreate2(loc, name.encode(), tid, s, 0, 0)
    h.Awrite(a, tid, buf)
    h.Sclose(s)
    h.Aclose(a)


def _realvect_type(h, nd):
    t = h.Tcreate(h.T_COMPOUND_CLASS, 8 * nd)
    for d, n in enumerate("xyz"[:nd]):
        h.Tinsert(t, n.encode(), 8 * d, h.NATIVE_DOUBLE)
    return t


def _box_type(h, nd):
    t = h.Tcreate(h.T_COMPOUND_CLASS, 8 * nd)
    for d in range(nd):
        h.Tinsert(t, ("lo_" + "ijk"[d]).encode(), 4 * d, h.NATIVE_INT)
    for d in range(nd):
        h.Tinsert(t, ("hi_" + "ijk"[d]).encode(), 4 * (nd + d), h.NATIVE_INT)
    return t


def level_dx(tree, nb, nd):
    """dx_level (SayakaWriterHdf5.cpp:368-377): block 0's cell size, halved per level."""
    finest = int(tree["level"].max())
    dx = [np.array(tree["length"][0][:nd] / nb)]
    for _ in range(1, finest):
        dx.append(dx[-1] / 2.0)
    return dx


def problem_box(tree, nd):
    """problo / probhi (SayakaWriterHdf5.cpp:389-405): union of every block's box."""
    return tree["lo"][:, :nd].min(axis=0), tree["hi"][:, :nd].max(axis=0)


def write_reference(path, tree, nb, nd, state, ngrow, time=0.0, dt=1.0, step=0, names=None):
    """WriteDataHdf5 (SayakaWriterHdf5.cpp:351-530) through libhdf5, call for call: state is
    [comp][block][grown box] (2D: [comp][block][1][gy][gx]); names as WriterMeta::name."""
    h = HDF5
    ncomp = state.shape[0]
    finest = int(tree["level"].max())
    names = list(names or [])
    names += [""] * (ncomp - len(names))
    names = [n if n else f"comp{c}" for c, n in enumerate(names)]   # guessName
    dx = level_dx(tree, nb, nd)
    problo, probhi = problem_box(tree, nd)
    levels = [np.nonzero(tree["level"] == l)[0] for l in range(1, finest + 1)]
    ndata = nb ** nd
    f = h.Fcreate(path.encode(), h.F_ACC_TRUNC, 0, 0)
    # write_header_hdf5
    g = h.Gopen2(f, b"/", 0)
    I = lambda v: C.byref(C.c_int(int(v)))
    D = lambda v: C.byref(C.c_double(float(v)))
    _scalar_attr(h, g, "num_components", h.NATIVE_INT, I(ncomp))
    _scalar_attr(h, g, "num_levels", h.NATIVE_INT, I(finest))
    _scalar_attr(h, g, "iteration", h.NATIVE_INT, I(step))
    _scalar_attr(h, g, "time", h.NATIVE_DOUBLE, D(time))
    rv = _realvect_type(h, nd)
    dh = tree["length"][0] / nb
    ar = np.array([1.0, dh[1] / dh[0], dh[2] / dh[0]][:nd])
    _scalar_attr(h, g, "aspect_ratio", rv, ar.ctypes.data)
    _scalar_attr(h, g, "data_centering", h.NATIVE_INT, I(0))
    pl = np.ascontiguousarray(problo, dtype=np.float64)
    _scalar_attr(h, g, "prob_lo", rv, pl.ctypes.data)
    h.Tclose(rv)
    for c in range(ncomp):
        st = h.Tcopy(h.C_S1)
        h.Tset_size(st, 64)
        buf = C.create_string_buffer(names[c].encode()[:64], 64)
        _scalar_attr(h, g, f"component_{c}", st, buf)
        h.Tclose(st)
    h.Gclose(g)
    g = h.Gcreate2(f, b"/Chombo_global", 0, 0, 0)
    _scalar_attr(h, g, "SpaceDim", h.NATIVE_INT, I(nd))
    _scalar_attr(h, g, "testReal", h.NATIVE_DOUBLE, D(0.0))
    h.Gclose(g)
    for il in range(finest):
        g = h.Gcreate2(f, f"/level_{il}".encode(), 0, 0, 0)
        _scalar_attr(h, g, "ref_ratio", h.NATIVE_INT, I(2))
        _scalar_attr(h, g, "dt", h.NATIVE_DOUBLE, D(dt))
        _scalar_attr(h, g, "dx", h.NATIVE_DOUBLE, D(dx[il][0]))
        bt = _box_type(h, nd)
        dom = np.zeros(2 * nd, np.int32)
        for d in range(nd):
            dom[nd + d] = int((probhi[d] - problo[d] - 0.5 * dx[il][d]) / dx[il][d])
        _scalar_attr(h, g, "prob_domain", bt, dom.ctypes.data)
        dims = (hsize_t * 1)(len(levels[il]))
        s = h.Screate_simple(1, dims, None)
        h.Dclose(h.Dcreate2(g, b"boxes", bt, s, 0, 0, 0))
        h.Sclose(s)
        h.Tclose(bt)
        dims = (hsize_t * 1)(ndata * ncomp * len(levels[il]))
        s = h.Screate_simple(1, dims, None)
        h.Dclose(h.Dcreate2(g, b"data:datatype=0", h.NATIVE_DOUBLE, s, 0, 0, 0))
        h.Sclose(s)
        g2 = h.Gcreate2(g, b"data_attributes", 0, 0, 0)
        _scalar_attr(h, g2, "comps", h.NATIVE_INT, I(ncomp))
        h.Gclose(g2)
        h.Gclose(g)
    # the per-level writes
    bt = _box_type(h, nd)
    gz = ngrow if nd == 3 else 0
    for il in range(finest):
        g = h.Gopen2(f, f"/level_{il}".encode(), 0)
        dh = dx[il]
        for igrid, b in enumerate(levels[il]):
            one = (hsize_t * 1)(1)
            sp = h.Screate_simple(1, one, None)
            ds = h.Dopen2(g, b"boxes", 0)
            fs = h.Dget_space(ds)
            h.Sselect_hyperslab(fs, h.S_SELECT_SET, (hsize_t * 1)(igrid), None, one, None)
            box = np.zeros(2 * nd, np.int32)
            for d in range(nd):
                box[d] = int((tree["lo"][b][d] + 0.5 * dh[d] - problo[d]) / dh[d])
                box[nd + d] = int((tree["hi"][b][d] - 0.5 * dh[d] - problo[d]) / dh[d])
            h.Dwrite(ds, bt, sp, fs, 0, box.ctypes.data)
            h.Sclose(sp)
            h.Dclose(ds)
            h.Sclose(fs)
            cnt = (hsize_t * 1)(ndata)
            off = igrid * ncomp * ndata
            for c in range(ncomp):
                sp = h.Screate_simple(1, cnt, None)
                ds = h.Dopen2(g, b"data:datatype=0", 0)
                fs = h.Dget_space(ds)
                h.Sselect_hyperslab(fs, h.S_SELECT_SET, (hsize_t * 1)(off), None, cnt, None)
                off += ndata
                blk = state[c, b]
                unk = np.ascontiguousarray(blk[gz:gz + (nb if nd == 3 else 1), ngrow:ngrow + nb, ngrow:ngrow + nb],
                                           dtype=np.float64)
                h.Dwrite(ds, h.NATIVE_DOUBLE, sp, fs, 0, unk.ctypes.data)
                h.Sclose(sp)
                h.Dclose(ds)
                h.Sclose(fs)
        h.Gclose(g)
    h.Tclose(bt)
    h.Fclose(f)


def pack_levels(tree, nb, nd, state, ngrow):
    """The data:datatype=0 datasets of all levels back to back: per level, per block of the
    level (ascending id), per component, the valid cells, i fastest."""
    finest = int(tree["level"].max())
    gz = ngrow if nd == 3 else 0
    nz = nb if nd == 3 else 1
    parts = []
    for l in range(1, finest + 1):
        for b in np.nonzero(tree["level"] == l)[0]:
            for c in range(state.shape[0]):
                parts.append(state[c, b, gz:gz + nz, ngrow:ngrow + nb, ngrow:ngrow + nb].ravel())
    return np.concatenate(parts) if parts else np.zeros(0)
