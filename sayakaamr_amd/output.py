"""Chombo-format HDF5 output (WriteDataHdf5, SayakaWriterHdf5.cpp:351-530; WriterMeta,
SayakaWriter.h:13-48) over the host library's own HDF5 writer."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class WriterError(RuntimeError):
    pass


def write_chombo(tree, packed, ncomp, filename, time=0.0, dt=1.0, step=0, names=None):
    """Write cell data already in the file's order (for every level 1..finest, every block of
    the level in ascending id, every component, its valid cells, i fastest; what
    DeviceState.pack_levels returns) on tree (a sayakaamr_amd.tree.AmrTree) to filename.
    names: component names (missing or empty ones become "comp<i>", WriterMeta::guessName)."""
    lib = N.host()
    nb, nd = tree.spec.nb, tree.ndim
    need = tree.numblocks * ncomp * nb ** nd
    a = np.ascontiguousarray(packed, dtype=np.float64).ravel()
    if a.size != need:
        raise ValueError(f"packed data holds {a.size} values, the tree needs {need}")
    arr = None
    if names is not None:
        enc = [(n or "").encode() for n in names] + [b""] * (ncomp - len(names))
        arr = (C.c_char_p * ncomp)(*enc[:ncomp])
    rc = lib.sy_write_chombo(tree.h, ncomp, a.ctypes.data_as(C.c_void_p), str(filename).encode(), float(time),
                             float(dt), int(step), C.cast(arr, C.c_void_p) if arr is not None else None)
    if rc != 0:
        raise WriterError(lib.sy_writer_last_error().decode())
