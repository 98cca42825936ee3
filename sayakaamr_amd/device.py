"""Python mirror of the device solver (include/sayaka_amd.h) over ctypes.

``MacSolverGPU`` follows the reference MacSolver call sequence (Test7.h:612-637):
prepare (topology) -> set_bc -> coefficients -> rhs / initial solution -> cg_solve -> solution.
There is no CPU path: every call goes to libsayaka_dev.so and raises on error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .workloads import NEUMANN

i32, f64, vp = C.c_int, C.c_double, C.c_void_p

RHS, SOL, RESID, CORR, TMP, X0, X, R, RHAT, P, PHAT, S, SHAT, V, T = range(15)
FP64, MIXED, FP64_FAST = 0, 1, 2


class SolveParams(C.Structure):
    _fields_ = [("tol_rel", f64), ("tol_abs", f64), ("max_iter", i32), ("precond", i32),
                ("precision", i32), ("pre_smooth", i32), ("post_smooth", i32), ("hist_cap", i32)]


class SolveStats(C.Structure):
    _fields_ = [("status", i32), ("iters", i32), ("half_step", i32), ("bottom_calls", i32),
                ("bottom_iters", C.c_longlong), ("rnorm0", f64), ("rnorm", f64), ("nhist", i32),
                ("hist", C.POINTER(f64))]


_dev = None
_SIG = {
    "sayaka_dev_last_error": ([], C.c_char_p),
    "sayaka_dev_create": ([i32, i32, i32, i32, C.POINTER(vp)], i32),
    "sayaka_dev_destroy": ([vp], i32),
    "sayaka_dev_set_topology": ([vp, i32, i32, i32] + [vp] * 7, i32),
    "sayaka_dev_set_bc": ([vp, vp], i32),
    "sayaka_dev_set_coefs": ([vp, f64, f64, vp, vp, vp], i32),
    "sayaka_dev_get_coefs": ([vp, vp, vp, vp, vp, vp], i32),
    "sayaka_dev_upload": ([vp, i32, vp], i32),
    "sayaka_dev_download": ([vp, i32, vp], i32),
    "sayaka_dev_fix_bc": ([vp, vp, i32, i32], i32),
    "sayaka_dev_apply": ([vp, i32, i32, i32], i32),
    "sayaka_dev_residual": ([vp, i32, i32, i32, i32], i32),
    "sayaka_dev_smooth": ([vp, i32, i32, i32], i32),
    "sayaka_dev_restrict": ([vp, i32, i32], i32),
    "sayaka_dev_prolong": ([vp, i32, i32], i32),
    "sayaka_dev_dot": ([vp, i32, i32, i32, C.POINTER(f64), C.POINTER(f64)], i32),
    "sayaka_dev_dot_noalloc": ([vp, i32, vp, vp, vp], i32),
    "sayaka_dev_precond": ([vp, i32, i32, i32], i32),
    "sayaka_dev_add": ([vp, i32, i32, i32], i32),
    "sayaka_dev_bottom": ([vp, i32, i32, f64, f64, i32, C.POINTER(i32), C.POINTER(i32)], i32),
    "sayaka_dev_solve": ([vp, C.POINTER(SolveParams), C.POINTER(SolveStats)], i32),
    "sayaka_dev_sync": ([vp], i32),
    "sayaka_dev_fill": ([vp, i32, f64], i32),
    "sayaka_dev_set_mixed_bottom": ([vp, f64, f64, i32], i32),
    "sayaka_dev_set_mixed_coarse": ([vp, i32], i32),
    "sayaka_dev_flux": ([vp, i32, i32, i32, vp], i32),
    "sayaka_dev_remap": ([vp, vp, i32, vp, i32], i32),
    "sayaka_dev_copy": ([vp, i32, i32], i32),
    "sayaka_dev_set_surroundings": ([vp, i32, vp, vp], i32),
    "sayaka_dev_alloc": ([vp, C.c_size_t, C.POINTER(vp)], i32),
    "sayaka_dev_free": ([vp, vp], i32),
    "sayaka_dev_memcpy": ([vp, vp, vp, C.c_size_t, i32], i32),
    "sayaka_dev_state_fill": ([vp, vp, i32, i32, i32, i32, i32, i32, vp, i32, vp, vp, vp], i32),
    "sayaka_dev_face_fill": ([vp, vp, i32, i32, i32, i32, i32, i32, i32, vp], i32),
    "sayaka_dev_pack_levels": ([vp, vp, i32, i32, i32, vp], i32),
    "sayaka_dev_state_remap": ([vp, vp, i32, i32, vp, i32, i32, i32, i32, vp, vp, i32, vp, vp, vp], i32),
    "sayaka_dev_set_mixed_smoother": ([vp, i32], i32),
    "sayaka_cnn_setup": ([vp, i32], i32),
    "sayaka_cnn_set_params": ([vp, vp], i32),
    "sayaka_cnn_get_params": ([vp, vp], i32),
    "sayaka_cnn_store_reset": ([vp, C.c_longlong], i32),
    "sayaka_cnn_collect": ([vp, vp, i32, f64, f64, f64], C.c_longlong),
    "sayaka_cnn_train_step": ([vp, vp, i32, C.c_float, C.POINTER(C.c_float)], i32),
    "sayaka_cnn_loss_grad": ([vp, vp, i32, C.POINTER(C.c_float), vp, vp], i32),
    "sayaka_cnn_store_read": ([vp, C.c_longlong, i32, vp, vp], i32),
    "sayaka_cnn_guess": ([vp, i32, f64, f64], i32),
    "sayaka_gen_sphere_blobs": ([vp, vp, vp, vp, f64, f64, f64, i32, i32, vp, vp, vp, C.POINTER(f64)], i32),
    "sayaka_gen_sphere": ([vp, vp, vp, vp, f64, f64, f64, i32], i32),
    "sayaka_gen_velocity": ([vp, vp, vp, vp, C.c_ulonglong, i32], i32),
    "sayaka_gen_divergence_rhs": ([vp, C.POINTER(f64)], i32),
    "sayaka_dev_project": ([vp, i32, C.POINTER(f64), C.POINTER(f64)], i32),
    "sayaka_dev_get_velocity": ([vp, vp], i32),
    "sayaka_dev_profile": ([vp, i32, i32], i32),
    "sayaka_dev_set_bottom_placement": ([vp, i32], i32),
    "sayaka_dev_set_half_step_defer": ([vp, f64], i32),
    "sayaka_dev_bottom_placement": ([vp, vp], i32),
    "sayaka_dev_profile_read": ([vp, i32, C.POINTER(f64), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                 C.POINTER(f64)], i32),
}

# kernel classes of sayaka_dev_profile (include/sayaka_amd.h)
K_APPLY, K_SMOOTH, K_SMOOTH32, K_BOTTOM, K_VEC, K_TRANSFER = range(6)
KCLASS_NAMES = ("apply", "smooth", "smooth32", "bottom", "vec", "transfer")


def dev():
    """libsayaka_dev.so; raises when it was not built (no fallback exists)."""
    global _dev
    if _dev is None:
        if not os.path.exists(N.DEV_LIB):
            raise N.NativeLibraryMissing(f"{N.DEV_LIB} not built; run __graft_entry__.build()")
        lib = C.CDLL(N.DEV_LIB)
        N.register(lib, _SIG)
        _dev = lib
    return _dev


class DeviceError(RuntimeError):
    pass


def _chk(rc, what):
    if rc < 0:
        raise DeviceError(f"{what}: {dev().sayaka_dev_last_error().decode()}")
    return rc


def _parr(arrs):
    keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
    pa = (vp * 3)(*([None if a is None else a.ctypes.data for a in keep] + [None] * (3 - len(keep))))
    return keep, pa


class DeviceState:
    """A state with ghost layers in device memory, [ncomp][blocks][grown box] doubles (the
    device counterpart of a grown TreeData, SayakaTreeData.h:86-306): cell-centred, or face-
    centred in direction fdir (one more face than cells along fdir)."""

    def __init__(self, solver, ncomp, ngrow, blocks, fdir=None):
        self.s, self.ncomp, self.ngrow, self.blocks, self.fdir = solver, ncomp, ngrow, blocks, fdir
        nb, nd = solver.nb, solver.ndim
        ext = [nb + 2 * ngrow + (fdir == d) for d in range(3)]
        if nd == 2:
            ext[2] = 1 + (fdir == 2)
        self.shape = (ncomp, blocks, ext[2], ext[1], ext[0])
        self.nbytes = int(np.prod(self.shape)) * 8
        p = vp()
        _chk(solver.lib.sayaka_dev_alloc(solver.h, self.nbytes, C.byref(p)), "alloc")
        self.ptr = p.value
        self._key = N.track(self)   # registered after its context: closed before it at exit

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64).reshape(self.shape)
        _chk(self.s.lib.sayaka_dev_memcpy(self.s.h, C.c_void_p(self.ptr), a.ctypes.data_as(vp), self.nbytes, 1),
             "memcpy")

    def download(self):
        out = np.empty(self.shape)
        _chk(self.s.lib.sayaka_dev_memcpy(self.s.h, out.ctypes.data_as(vp), C.c_void_p(self.ptr), self.nbytes, 0),
             "memcpy")
        return out

    def fill(self, nlayer, prolong, bcmap=None, scomp=0, nc=None):
        """FillPatch::fillBoundary(scomp, nc, nlayer) on this state (face data: no BC map, the
        tree must be periodic)."""
        if self.fdir is None:
            self.s.state_fill(self.ptr, self.ncomp, self.blocks, self.ngrow, nlayer, prolong, bcmap, scomp, nc)
        else:
            self.s.face_fill(self.ptr, self.ncomp, self.blocks, self.fdir, self.ngrow, nlayer, prolong, scomp, nc)

    def remap_from(self, old, origin, prolong, bcmap=None):
        """This state (sized for the tree after a regrid) from `old` (the same state on the tree
        before it): AmrTree::regrid's packing and initNewBlockDataAfterRegrid
        (SayakaRegrid.cpp:530-591, 987-1105).  The solver must hold the new tree's topology and
        surroundings; origin = AmrTree.regrid_origin(); prolong[c] per component; bcmap as for
        fill (cell data)."""
        if old.fdir != self.fdir or old.ncomp != self.ncomp or old.ngrow != self.ngrow:
            raise ValueError("old and new state differ in location, components or ghost layers")
        org = np.ascontiguousarray(origin, dtype=np.int32)
        pr = np.ascontiguousarray(prolong[:self.ncomp], dtype=np.int32)
        if self.fdir is None:
            nphys, pc, mb, bv = self.s._bc_tables(bcmap or [{}] * self.ncomp, self.ncomp)
        else:
            nphys, pc, mb, bv = 0, np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1)
        _chk(self.s.lib.sayaka_dev_state_remap(self.s.h, C.c_void_p(old.ptr), old.blocks, old.blocks,
                                               C.c_void_p(self.ptr), self.ncomp, self.blocks, self.ngrow,
                                               -1 if self.fdir is None else self.fdir, N.ptr(org), N.ptr(pr),
                                               nphys, N.ptr(pc), N.ptr(mb), N.ptr(bv)), "state_remap")

    def pack_levels(self):
        """Valid cells in the Chombo file's order (per level, block, component), on the host:
        packed on the device, then one copy."""
        if self.fdir is not None:
            raise ValueError("Chombo output takes cell data only (the reference exits otherwise)")
        nb, nd = self.s.nb, self.s.ndim
        n = self.s.n * self.ncomp * nb ** nd
        p = vp()
        _chk(self.s.lib.sayaka_dev_alloc(self.s.h, n * 8, C.byref(p)), "alloc")
        try:
            _chk(self.s.lib.sayaka_dev_pack_levels(self.s.h, C.c_void_p(self.ptr), self.ncomp, self.blocks,
                                                   self.ngrow, p), "pack_levels")
            out = np.empty(n)
            _chk(self.s.lib.sayaka_dev_memcpy(self.s.h, out.ctypes.data_as(vp), p, n * 8, 0), "memcpy")
        finally:
            self.s.lib.sayaka_dev_free(self.s.h, p)
        return out

    def write_chombo(self, tree, filename, time=0.0, dt=1.0, step=0, names=None):
        """WriteDataHdf5 (SayakaWriterHdf5.cpp:351-530) of this state on tree (the host
        AmrTree it was filled on)."""
        from . import output
        output.write_chombo(tree, self.pack_levels(), self.ncomp, filename, time, dt, step, names)

    def close(self):
        if self.ptr and getattr(self.s, "h", None):
            self.s.lib.sayaka_dev_free(self.s.h, C.c_void_p(self.ptr))
        self.ptr = None
        N.untrack(getattr(self, "_key", None))

    def __del__(self):
        if N.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class MacSolverGPU:
    """Device MAC solver for one tree (MacSolver, SayakaMacSolver.h:25-425)."""

    def __init__(self, ndim: int, nb: int, max_blocks: int, device: int = 0):
        self.lib = dev()
        self.ndim, self.nb = ndim, nb
        self.ncell = nb ** ndim
        self.nface = (nb + 1) * nb ** (ndim - 1)
        h = vp()
        _chk(self.lib.sayaka_dev_create(device, ndim, nb, max_blocks, C.byref(h)), "create")
        self.h = h
        self.n = 0
        self.finest = 0
        self._key = N.track(self)

    def close(self):
        if getattr(self, "h", None):
            self.lib.sayaka_dev_destroy(self.h)
            self.h = None
        N.untrack(getattr(self, "_key", None))

    def __del__(self):
        # at interpreter finalization the atexit hook has closed every live context already;
        # nothing may reach the HIP runtime from a finalizer then
        if N.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    # --- MacSolver::prepareSolver (tables of the tree) ---------------------------------
    def prepare(self, tree: dict, spec):
        self.n = int(tree["level"].shape[0])
        self.finest = int(tree["level"].max())
        self._tree = {k: np.ascontiguousarray(tree[k], dtype=np.int32)
                      for k in ("level", "parent", "which_child", "child", "neighbor", "node_type")}
        ln = np.ascontiguousarray(tree["length"], dtype=np.float64)
        t = self._tree
        _chk(self.lib.sayaka_dev_set_topology(self.h, self.n, spec.minlevel, spec.maxlevel,
                                              N.ptr(t["level"]), N.ptr(t["parent"]), N.ptr(t["which_child"]),
                                              N.ptr(t["child"]), N.ptr(t["neighbor"]), N.ptr(t["node_type"]),
                                              N.ptr(ln)), "set_topology")
        self._nbr = t["neighbor"]

    def set_bc(self, physbc, mathbc):
        """MacSolver::setBC with one phys->math map for every wall."""
        bc = np.full((self.n, 6), -1, np.int32)
        for pb, mb in zip(physbc, mathbc):
            bc[self._nbr == pb] = mb
        _chk(self.lib.sayaka_dev_set_bc(self.h, N.ptr(bc)), "set_bc")

    def set_coefs(self, vf=None, af=None, b=None, beta=1.0):
        v = None if vf is None else np.ascontiguousarray(vf, dtype=np.float64)
        ka, pa = _parr(af) if af is not None else (None, None)
        kb, pb = _parr(b) if b is not None else (None, None)
        _chk(self.lib.sayaka_dev_set_coefs(self.h, 0.0, beta, N.ptr(v), pa, pb), "set_coefs")

    def get_coefs(self):
        vf = np.zeros((self.n, self.ncell))
        b = [np.zeros((self.n, self.nface)) for _ in range(self.ndim)]
        a = [np.zeros((self.n, self.nface)) for _ in range(self.ndim)]
        w = np.zeros((self.n, self.ndim, self.nface))
        wcf = np.zeros((self.n, 2 * self.ndim, self.nb ** (self.ndim - 1)))
        kb, pb = _parr(b)
        ka, pa = _parr(a)
        _chk(self.lib.sayaka_dev_get_coefs(self.h, N.ptr(vf), pa, pb, N.ptr(w), N.ptr(wcf)), "get_coefs")
        return vf, kb, ka, w, wcf

    def get_vf(self):
        """Volume fractions of every block (the solver's coefficient state)."""
        vf = np.zeros((self.n, self.ncell))
        _chk(self.lib.sayaka_dev_get_coefs(self.h, N.ptr(vf), None, None, None, None), "get_coefs")
        return vf

    # --- fields ----------------------------------------------------------------------
    def upload(self, field, x):
        _chk(self.lib.sayaka_dev_upload(self.h, field, N.ptr(np.ascontiguousarray(x, dtype=np.float64))), "upload")

    def download(self, field):
        out = np.zeros((self.n, self.ncell))
        _chk(self.lib.sayaka_dev_download(self.h, field, N.ptr(out)), "download")
        return out

    def fix_bc(self, phi_grown, ngrow, rhs_field=RHS):
        _chk(self.lib.sayaka_dev_fix_bc(self.h, N.ptr(np.ascontiguousarray(phi_grown, dtype=np.float64)), ngrow,
                                        rhs_field), "fix_bc")

    # --- operators -------------------------------------------------------------------
    def apply(self, m, out, inp):
        _chk(self.lib.sayaka_dev_apply(self.h, m, out, inp), "apply")

    def residual(self, m, r, x, rhs):
        _chk(self.lib.sayaka_dev_residual(self.h, m, r, x, rhs), "residual")

    def smooth(self, m, sol, rhs):
        _chk(self.lib.sayaka_dev_smooth(self.h, m, sol, rhs), "smooth")

    def restrict(self, cl, field):
        _chk(self.lib.sayaka_dev_restrict(self.h, cl, field), "restrict")

    def prolong(self, fl, field):
        _chk(self.lib.sayaka_dev_prolong(self.h, fl, field), "prolong")

    def remap(self, origin, nold, fields):
        """After prepare() with a regridded tree: move fields to the new numbering (origin
        from AmrTree.regrid_origin) and inject the new blocks from their parents."""
        o = np.ascontiguousarray(origin, dtype=np.int32)
        f = np.ascontiguousarray(fields, dtype=np.int32)
        _chk(self.lib.sayaka_dev_remap(self.h, N.ptr(o), int(nold), N.ptr(f), f.size), "remap")

    # --- FillPatch::fillBoundary on a device-resident cell state ------------------------
    def set_surroundings(self, tree: dict):
        """The tree's SurroundingBlocks tables (after prepare, for the current tree)."""
        s = np.ascontiguousarray(tree["surr"], dtype=np.int32)
        r = np.ascontiguousarray(tree["surr_bcref"], dtype=np.int32)
        _chk(self.lib.sayaka_dev_set_surroundings(self.h, self.n, N.ptr(s), N.ptr(r)), "set_surroundings")

    @staticmethod
    def _bc_tables(bcmap, nc):
        """BC register arrays of the C ABI from per-component {physical code: (math BC, value)}."""
        codes = sorted({int(k) for m in bcmap[:nc] for k in m})
        nphys = len(codes)
        mb = np.full((nc, 27, max(nphys, 1)), -1, np.int32)
        bv = np.zeros((nc, 27, max(nphys, 1)))
        for c in range(nc):
            for h, code in enumerate(codes):
                if code in bcmap[c]:
                    mb[c, :, h], bv[c, :, h] = bcmap[c][code]
        pc = np.array(codes if codes else [0], dtype=np.int32)
        return nphys, pc, mb, bv

    def state_fill(self, state_ptr, ncomp, blocks_alloc, ngrow, nlayer, prolong, bcmap, scomp=0, nc=None):
        """Ghost fill of the device state at state_ptr ([ncomp][blocks_alloc][grown box] doubles,
        e.g. DeviceState.ptr).
        prolong[c]: method of component scomp+c (0 injection, 1 centred linear, 2 limited);
        bcmap[c]: {physical code: (math BC, value)} of component scomp+c, the same on every
        surrounding position (BCRegister filled position by position the way the drivers do)."""
        nc = ncomp - scomp if nc is None else nc
        nphys, pc, mb, bv = self._bc_tables(bcmap, nc)
        pr = np.ascontiguousarray(prolong[:nc], dtype=np.int32)
        _chk(self.lib.sayaka_dev_state_fill(self.h, C.c_void_p(int(state_ptr)), ncomp, blocks_alloc, ngrow, scomp,
                                            nc, nlayer, N.ptr(pr), nphys, N.ptr(pc), N.ptr(mb), N.ptr(bv)),
             "state_fill")

    def face_fill(self, state_ptr, ncomp, blocks_alloc, fdir, ngrow, nlayer, prolong, scomp=0, nc=None):
        """Ghost fill of device face data of direction fdir at state_ptr ([ncomp][blocks_alloc]
        [grown face box] doubles); prolong[c] as for state_fill.  Periodic trees only."""
        nc = ncomp - scomp if nc is None else nc
        pr = np.ascontiguousarray(prolong[:nc], dtype=np.int32)
        _chk(self.lib.sayaka_dev_face_fill(self.h, C.c_void_p(int(state_ptr)), ncomp, blocks_alloc, fdir, ngrow,
                                           scomp, nc, nlayer, N.ptr(pr)), "face_fill")

    def flux(self, m, d, phi):
        """MacSolver::mg_flux_level: face flux of direction d on tower m from field phi,
        [blk][face] (0 outside the tower)."""
        out = np.zeros((self.n, self.nface))
        _chk(self.lib.sayaka_dev_flux(self.h, m, d, phi, N.ptr(out)), "flux")
        return out

    def dot(self, m, a, b):
        d = f64()
        mx = f64()
        _chk(self.lib.sayaka_dev_dot(self.h, m, a, b, C.byref(d), C.byref(mx)), "dot")
        return d.value, mx.value

    def precond(self, m, sol, rhs):
        _chk(self.lib.sayaka_dev_precond(self.h, m, sol, rhs), "precond")

    def bottom(self, sol, rhs, tol_rel=1e-6, tol_abs=1e-6, max_iter=2000):
        it = i32()
        st = i32()
        _chk(self.lib.sayaka_dev_bottom(self.h, sol, rhs, tol_rel, tol_abs, max_iter, C.byref(it), C.byref(st)),
             "bottom")
        return it.value, st.value

    def solve(self, tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=FP64, pre_smooth=2,
              post_smooth=2, hist_cap=4096):
        """MacSolver::cg_solve from the SOL field; returns the solve statistics."""
        p = SolveParams(tol_rel, tol_abs, max_iter, precond, precision, pre_smooth, post_smooth, hist_cap)
        hist = np.zeros(max(hist_cap, 1))
        st = SolveStats()
        st.hist = hist.ctypes.data_as(C.POINTER(f64))
        rc = _chk(self.lib.sayaka_dev_solve(self.h, C.byref(p), C.byref(st)), "solve")
        return dict(status=st.status, rc=rc, iters=st.iters, half_step=bool(st.half_step),
                    bottom_calls=st.bottom_calls, bottom_iters=int(st.bottom_iters), rnorm0=st.rnorm0,
                    rnorm=st.rnorm, hist=hist[:min(st.nhist, hist_cap)].copy())

    def sync(self):
        _chk(self.lib.sayaka_dev_sync(self.h), "sync")

    def copy(self, dst, src):
        """TreeData::Copy on the device: field dst = field src."""
        _chk(self.lib.sayaka_dev_copy(self.h, dst, src), "copy")

    def zero(self, field):
        """setZeroSol on the device for any field."""
        _chk(self.lib.sayaka_dev_fill(self.h, field, 0.0), "fill")

    def set_mixed_bottom(self, tol_rel, tol_abs, max_iter):
        _chk(self.lib.sayaka_dev_set_mixed_bottom(self.h, tol_rel, tol_abs, max_iter), "set_mixed_bottom")

    def set_mixed_coarse(self, two_level: bool):
        """Two-level preconditioner of the mixed solve's coarsest BiCGSTAB on / off."""
        _chk(self.lib.sayaka_dev_set_mixed_coarse(self.h, int(bool(two_level))), "set_mixed_coarse")

    def set_mixed_smoother(self, mode: int):
        """Mixed V-cycle smoother: 2 = hybrid sweeps (one launch per sweep, the default),
        1 = exact red-black with one launch per colour over the whole tower (double-buffered),
        0 = exact red-black with one launch per sub-level and colour in place."""
        _chk(self.lib.sayaka_dev_set_mixed_smoother(self.h, int(mode)), "set_mixed_smoother")

    # --- learned initial guess (include/sayaka_amd.h, sayaka_cnn_*) -------------------
    CNN_NP = 4 * (27 * 16 * 16 + 16)

    def cnn_setup(self, max_windows):
        _chk(self.lib.sayaka_cnn_setup(self.h, max_windows), "cnn_setup")

    def cnn_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.CNN_NP
        _chk(self.lib.sayaka_cnn_set_params(self.h, N.ptr(p)), "cnn_set_params")

    def cnn_get_params(self):
        p = np.zeros(self.CNN_NP, np.float32)
        _chk(self.lib.sayaka_cnn_get_params(self.h, N.ptr(p)), "cnn_get_params")
        return p

    def cnn_store_reset(self, cap):
        _chk(self.lib.sayaka_cnn_store_reset(self.h, int(cap)), "cnn_store_reset")

    def cnn_collect(self, leaf_ids, s_rhs, shift, t_scale):
        ids = np.ascontiguousarray(leaf_ids, dtype=np.int32)
        return _chk(self.lib.sayaka_cnn_collect(self.h, N.ptr(ids), ids.size, s_rhs, shift, t_scale), "cnn_collect")

    def cnn_train_step(self, idx, lr):
        i = np.ascontiguousarray(idx, dtype=np.int64)
        loss = C.c_float()
        _chk(self.lib.sayaka_cnn_train_step(self.h, N.ptr(i), i.size, lr, C.byref(loss)), "cnn_train_step")
        return loss.value

    def cnn_loss_grad(self, idx, want_out=False):
        i = np.ascontiguousarray(idx, dtype=np.int64)
        loss = C.c_float()
        g = np.zeros(self.CNN_NP, np.float32)
        out = np.zeros((i.size, 12 ** 3), np.float32) if want_out else None
        _chk(self.lib.sayaka_cnn_loss_grad(self.h, N.ptr(i), i.size, C.byref(loss), N.ptr(g), N.ptr(out)),
             "cnn_loss_grad")
        return loss.value, g, out

    def cnn_store_read(self, first, n):
        w = np.zeros((n, 12 ** 3, 16), np.float32)
        t = np.zeros((n, 512), np.float32)
        _chk(self.lib.sayaka_cnn_store_read(self.h, first, n, N.ptr(w), N.ptr(t)), "cnn_store_read")
        return w, t

    def cnn_guess(self, field, s_rhs, out_scale):
        _chk(self.lib.sayaka_cnn_guess(self.h, field, s_rhs, out_scale), "cnn_guess")

    def gen_sphere_blobs(self, tree, centre, radius, ratio, eps, nsub, blobs):
        """Cut-cell sphere coefficients and a zero-mean Gaussian-blob rhs on the device."""
        lo = np.ascontiguousarray(tree["lo"], dtype=np.float64)
        dx = np.ascontiguousarray(tree["length"][:, 0] / self.nb, dtype=np.float64)
        c = np.ascontiguousarray(centre, dtype=np.float64)
        bc = np.ascontiguousarray(blobs["c"], dtype=np.float64)
        bw = np.ascontiguousarray(blobs["w"], dtype=np.float64)
        ba = np.ascontiguousarray(blobs["a"], dtype=np.float64)
        mean = f64()
        _chk(self.lib.sayaka_gen_sphere_blobs(self.h, N.ptr(lo), N.ptr(dx), N.ptr(c), radius, ratio, eps, nsub,
                                              bw.size, N.ptr(bc), N.ptr(bw), N.ptr(ba), C.byref(mean)), "gen")
        return mean.value

    def gen_sphere(self, tree, centre, radius, ratio, eps, nsub):
        """The sphere's cut-cell coefficients alone (gen_sphere_blobs without the rhs)."""
        lo = np.ascontiguousarray(tree["lo"], dtype=np.float64)
        dx = np.ascontiguousarray(tree["length"][:, 0] / self.nb, dtype=np.float64)
        c = np.ascontiguousarray(centre, dtype=np.float64)
        _chk(self.lib.sayaka_gen_sphere(self.h, N.ptr(lo), N.ptr(dx), N.ptr(c), radius, ratio, eps, nsub), "gen")

    def gen_velocity(self, tree, spec, seed, zero_walls=True):
        """Seeded face velocity of the current tree on the device (workloads.face_velocity's
        definition; 0 on the domain walls when zero_walls)."""
        lo = np.ascontiguousarray(tree["lo"], dtype=np.float64)
        dlo = np.zeros(3)
        root = np.ones(3)
        for a in range(spec.ndim):
            dlo[a] = spec.lo[a]
            root[a] = (spec.hi[a] - spec.lo[a]) / max(spec.ndiv[a], 1)
        _chk(self.lib.sayaka_gen_velocity(self.h, N.ptr(lo), N.ptr(dlo), N.ptr(root), int(seed),
                                          int(bool(zero_walls))), "gen_velocity")

    def gen_divergence_rhs(self):
        """RHS = div(a u) / (dv vf) minus its mean over the fluid leaves, after the coarse/fine
        face sync of u; returns the removed mean."""
        mean = f64()
        _chk(self.lib.sayaka_gen_divergence_rhs(self.h, C.byref(mean)), "gen_divergence_rhs")
        return mean.value

    def project(self, phi=SOL):
        """u -= B flux(phi) on the finest tower; (max |div u| before, after), means removed."""
        b, a = f64(), f64()
        _chk(self.lib.sayaka_dev_project(self.h, phi, C.byref(b), C.byref(a)), "project")
        return b.value, a.value

    def get_velocity(self):
        """Face velocity, a list of ndim arrays [nblk][faces]."""
        out = np.empty((self.n, self.ndim, self.nface))
        _chk(self.lib.sayaka_dev_get_velocity(self.h, N.ptr(out)), "get_velocity")
        return [np.ascontiguousarray(out[:, d]) for d in range(self.ndim)]

    def profile(self, kclass, enable=True):
        _chk(self.lib.sayaka_dev_profile(self.h, kclass, int(enable)), "profile")

    def set_half_step_defer(self, factor=100.0):
        """Queue a half-iteration before reading whether the solve stops (the second half before
        |s|; in FP64 modes also the next first half before |r|) while the last tested norm is at
        least factor x the stopping threshold (0: always, negative: never; results identical)."""
        _chk(self.lib.sayaka_dev_set_half_step_defer(self.h, float(factor)), "set_half_step_defer")

    def set_bottom_placement(self, one_xcd=True):
        """The coarsest solve on one XCD when it fits (True, the default) or spread over the device."""
        _chk(self.lib.sayaka_dev_set_bottom_placement(self.h, int(bool(one_xcd))), "set_bottom_placement")

    def bottom_placement(self):
        """Coarsest-solve launches so far: all, placed on one XCD, and of those the launches that
        found their workgroups on several XCDs (device-wide barriers kept)."""
        out = np.zeros(3, np.int64)
        _chk(self.lib.sayaka_dev_bottom_placement(self.h, N.ptr(out)), "bottom_placement")
        return dict(launches=int(out[0]), one_xcd=int(out[1]), fallen_back=int(out[2]))

    def profile_read(self, kclass):
        t = f64()
        n = C.c_longlong()
        cells = C.c_longlong()
        by = f64()
        _chk(self.lib.sayaka_dev_profile_read(self.h, kclass, C.byref(t), C.byref(n), C.byref(cells), C.byref(by)),
             "profile_read")
        return dict(total_ms=t.value, launches=n.value, cells=cells.value, alg_bytes=by.value)
