"""Python mirror of the host tree (AmrTree, SayakaTree.h:195-333) over the C ABI."""
from __future__ import annotations

import numpy as np

from . import _native as N


class AmrTree:
    """Owning handle of a host AmrTree; also a tree backend for ``workloads.build_tree``."""

    def __init__(self, ndim: int = 3, max_blocks: int = 40000):
        self.ndim = ndim
        self.max_blocks = max_blocks
        self.lib = N.host()
        self.h = None
        self.spec = None
        self._key = N.track(self)   # registered states may hold device memory: closed at exit

    def create(self, spec):
        if self.h:
            self.lib.sy_tree_destroy(self.h)
        self.spec = spec
        self.ndim = spec.ndim
        nd = np.array(spec.ndiv, dtype=np.int32)
        lo = np.array(spec.lo, dtype=np.float64)
        hi = np.array(spec.hi, dtype=np.float64)
        bc = np.array(spec.bc, dtype=np.int32)
        self.h = self.lib.sy_tree_create(spec.ndim, spec.nb, spec.ngrow, spec.minlevel, spec.maxlevel,
                                         self.max_blocks, N.ptr(nd), N.ptr(lo), N.ptr(hi), N.ptr(bc))
        if not self.h:
            raise ValueError(f"bad tree spec {spec}")
        return self.numblocks

    def close(self):
        if self.h:
            self.lib.sy_tree_destroy(self.h)
            self.h = None
        N.untrack(getattr(self, "_key", None))

    def __del__(self):
        if N.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    @property
    def numblocks(self) -> int:
        return self.lib.sy_tree_numblocks(self.h)

    @property
    def finest(self) -> int:
        return self.lib.sy_tree_finest(self.h)

    def set_refine_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        if self.lib.sy_tree_set_flags(self.h, N.ptr(f), None, f.size) != 0:
            raise ValueError("flag array larger than the block pool")

    def set_coarsen_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        if self.lib.sy_tree_set_flags(self.h, None, N.ptr(f), f.size) != 0:
            raise ValueError("flag array larger than the block pool")

    def regrid(self) -> int:
        n = self.lib.sy_tree_regrid(self.h)
        if n < 0:
            raise MemoryError("AmrTree::regrid: block pool overflow")
        return n

    def flag_shell(self, centre, r0, factor):
        """Refine / coarsen flags of the current leaves for the shell |x - centre| = r0."""
        c = np.ascontiguousarray(centre, dtype=np.float64)
        return self.lib.sy_tree_flag_shell(self.h, N.ptr(c), float(r0), float(factor))

    def get_flags(self):
        """(refineFlag, coarsenFlag) of the current blocks."""
        n = self.numblocks
        r, c = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.lib.sy_tree_get_flags(self.h, N.ptr(r), N.ptr(c), n)
        return r, c

    def regrid_origin(self):
        """Per block, its id before the last regrid (-1 for blocks that regrid created)."""
        out = np.zeros(self.numblocks, np.int32)
        self.lib.sy_tree_regrid_origin(self.h, N.ptr(out))
        return out

    def build_shell(self, spec, centre, r0, factor):
        """Create the tree and refine around the shell |x - centre| = r0 (native loop; same
        result as workloads.build_tree with workloads.shell_criterion)."""
        self.create(spec)
        c = np.ascontiguousarray(centre, dtype=np.float64)
        if self.lib.sy_tree_build_shell(self.h, N.ptr(c), float(r0), float(factor)) < 0:
            raise MemoryError("AmrTree::regrid: block pool overflow")
        return self.export()

    def export(self) -> dict:
        n = self.numblocks
        t = dict(level=np.zeros(n, np.int32), parent=np.zeros(n, np.int32),
                 which_child=np.zeros(n, np.int32), child=np.zeros((n, 8), np.int32),
                 neighbor=np.zeros((n, 6), np.int32), node_type=np.zeros(n, np.int32),
                 node_type2=np.zeros((n, 6), np.int32), lo=np.zeros((n, 3)), hi=np.zeros((n, 3)),
                 center=np.zeros((n, 3)), length=np.zeros((n, 3)),
                 surr=np.zeros((n, 27), np.int32), surr_type=np.zeros((n, 27), np.int32),
                 surr_bcref=np.zeros((n, 27), np.int32))
        self.lib.sy_tree_export(self.h, *[N.ptr(t[k]) for k in ("level", "parent", "which_child", "child",
                                                                "neighbor", "node_type", "node_type2",
                                                                "lo", "hi", "center", "length", "surr",
                                                                "surr_type", "surr_bcref")])
        return t

    # ---- registered states (TreeStateData; regrid moves them, init_new_data fills new blocks)
    def _tchk(self, rc):
        if rc < 0:
            raise RuntimeError(self.lib.sy_tree_last_error().decode())
        return rc

    def state_register(self, dir, ncomp, ngrow, prolong, physbc=(), mathbc=(), bcval=()):
        """Register a state (cell: dir None or -1; face: direction); mathbc / bcval are
        [comp][physical code] on every surrounding position.  Returns its index."""
        d = -1 if dir is None else int(dir)
        pr = np.ascontiguousarray(prolong, dtype=np.int32)
        nphys = len(physbc)
        pb = np.ascontiguousarray(physbc if nphys else [0], dtype=np.int32)
        mb = np.ascontiguousarray(mathbc if nphys else [0], dtype=np.int32)
        bv = np.ascontiguousarray(bcval if nphys else [0.0], dtype=np.float64)
        idx = self._tchk(self.lib.sy_tree_state_register(self.h, d, ncomp, ngrow, N.ptr(pr), nphys, N.ptr(pb),
                                                         N.ptr(mb), N.ptr(bv)))
        if not hasattr(self, "_states"):
            self._states = {}
        self._states[idx] = (d, ncomp, ngrow)
        return idx

    def _state_shape(self, idx):
        d, ncomp, ngrow = self._states[idx]
        nb = self.spec.nb
        ext = [nb + 2 * ngrow + (d == q) for q in range(3)]
        if self.ndim == 2:
            ext[2] = 1
        return (ncomp, self.numblocks, ext[2], ext[1], ext[0])

    def state_set_grown(self, idx, data):
        a = np.ascontiguousarray(data, dtype=np.float64)
        if a.shape != self._state_shape(idx):
            raise ValueError(f"state shape {a.shape}, expected {self._state_shape(idx)}")
        self._tchk(self.lib.sy_tree_state_set_grown(self.h, idx, N.ptr(a)))

    def state_get_grown(self, idx):
        out = np.zeros(self._state_shape(idx))
        self._tchk(self.lib.sy_tree_state_get_grown(self.h, idx, N.ptr(out)))
        return out

    def init_new_data(self):
        """AmrTree::initNewBlockDataAfterRegrid for every registered state (on the GPU)."""
        self._tchk(self.lib.sy_tree_init_new_data(self.h))

    def state_fill(self, idx, nlayer, scomp=0, ncomp=None):
        """FillPatch::fillBoundary of a registered state (on the GPU)."""
        nc = self._states[idx][1] - scomp if ncomp is None else ncomp
        self._tchk(self.lib.sy_tree_state_fill(self.h, idx, scomp, nc, nlayer))

    def state_set_resident(self, idx, on=True, device=0):
        """Keep a registered state on the device (TreeData::setResident): regrid moves,
        new-block fills and ghost fills then run there with no host transfer."""
        self._tchk(self.lib.sy_tree_state_set_resident(self.h, idx, int(bool(on)), device))

    def transfer_stats(self):
        """Host <-> device traffic of the tree's device context: bytes of state to the device,
        to the host, topology uploads, surroundings uploads."""
        out = np.zeros(4, np.int64)
        self._tchk(self.lib.sy_tree_transfer_stats(self.h, N.ptr(out)))
        return dict(to_device=int(out[0]), to_host=int(out[1]), topology_uploads=int(out[2]),
                    surroundings_uploads=int(out[3]))

    def level_blocks(self, level: int) -> np.ndarray:
        n = self.lib.sy_tree_level_count(self.h, level)
        out = np.zeros(n, np.int32)
        if n:
            self.lib.sy_tree_level_blocks(self.h, level, N.ptr(out))
        return out
