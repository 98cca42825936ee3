"""Python mirror of the host MacSolver (include/sayaka/mac_solver.h) over the flat C ABI
(include/sayaka_host.h): the reference's call surface and call order (SayakaMacSolver.h:25-425,
Test7.h:612-653) for callers in Python.

    mac = MacSolver(tree)            # tree: sayakaamr_amd.tree.AmrTree
    mac.prepareSolver()
    mac.setBC([WALL], [NEUMANN])
    mac.setIBFrac(vf, af)            # or setIBFracByLS(ls_grown, ngrow)
    mac.setCoefB(b)
    mac.setRhs(rhs); mac.setZeroSol()
    mac.setTolerance(1e-10, 0.0); mac.setMaxIteration(2000)
    ret = mac.cg_solve()             # 0 converged or not, 1 breakdown (as the reference)
    sol = mac.getSol()

Arrays are packed valid-only blocks: cells [blk][k][j][i], faces [dir][blk][k][j][i].  There
is no CPU path: every call runs on the GPU and raises on error.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

FP64, MIXED, FP64_FAST = 0, 1, 2


class MacSolverError(RuntimeError):
    pass


class MacSolver:
    def __init__(self, tree, device: int = 0):
        self.lib = N.host()
        self.tree = tree
        self.h = self.lib.sy_mac_create(tree.h, device)
        if not self.h:
            raise MacSolverError(f"MacSolver: {self.lib.sy_mac_last_error().decode()}")
        self.ndim, self.nb = tree.ndim, tree.spec.nb
        self.ncell = self.nb ** self.ndim
        self._params = dict(tol_rel=1e-10, tol_abs=0.0, max_iter=2000, precond=1, precision=FP64, pre=2, post=2)
        self.stats = {}
        self._key = N.track(self)

    def _chk(self, rc, what):
        if rc < 0:
            raise MacSolverError(f"{what}: {self.lib.sy_mac_last_error().decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.sy_mac_destroy(self.h)
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
    def numBlocks(self):
        return self.tree.numblocks

    def prepareSolver(self):
        self._chk(self.lib.sy_mac_prepare(self.h), "prepareSolver")

    def setBC(self, phys_bc, math_bc):
        p = np.ascontiguousarray(phys_bc, dtype=np.int32)
        m = np.ascontiguousarray(math_bc, dtype=np.int32)
        self._chk(self.lib.sy_mac_set_bc(self.h, p.size, N.ptr(p), N.ptr(m)), "setBC")

    def setIBFracByLS(self, ls_grown, ngrow):
        g = np.ascontiguousarray(ls_grown, dtype=np.float64)
        self._chk(self.lib.sy_mac_set_ib_ls(self.h, N.ptr(g), ngrow), "setIBFracByLS")

    def setIBFrac(self, vf, af):
        v = np.ascontiguousarray(vf, dtype=np.float64)
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in af]
        pa = (C.c_void_p * 3)(*([a.ctypes.data for a in keep] + [None] * (3 - len(keep))))
        self._chk(self.lib.sy_mac_set_ib_frac(self.h, N.ptr(v), pa), "setIBFrac")

    def setCoefB(self, b):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in b]
        pb = (C.c_void_p * 3)(*([a.ctypes.data for a in keep] + [None] * (3 - len(keep))))
        self._chk(self.lib.sy_mac_set_coefb(self.h, pb), "setCoefB")

    def setRhs(self, rhs):
        r = np.ascontiguousarray(rhs, dtype=np.float64)
        self._chk(self.lib.sy_mac_set_rhs(self.h, N.ptr(r)), "setRhs")

    def setInitSol(self, sol):
        s = np.ascontiguousarray(sol, dtype=np.float64)
        self._chk(self.lib.sy_mac_set_init_sol(self.h, N.ptr(s)), "setInitSol")

    def fixBC(self, phi_grown, ngrow, rhs):
        """MacSolver::fixBC: phi over the grown box [blk][grown] (wall ghosts carry the BC
        values); returns rhs [blk][cell] with the wall terms added."""
        g = np.ascontiguousarray(phi_grown, dtype=np.float64)
        r = np.ascontiguousarray(rhs, dtype=np.float64).copy()
        self._chk(self.lib.sy_mac_fix_bc(self.h, N.ptr(g), int(ngrow), N.ptr(r)), "fixBC")
        return r

    def setZeroSol(self):
        self._chk(self.lib.sy_mac_set_zero_sol(self.h), "setZeroSol")

    def getSol(self):
        out = np.zeros((self.numBlocks, self.ncell))
        self._chk(self.lib.sy_mac_get_sol(self.h, N.ptr(out)), "getSol")
        return out

    def mg_flux_level(self, mg_level, d, phi):
        """MacSolver::mg_flux_level: face flux of direction d on tower mg_level from phi
        [blk][cell]; returns [blk][face] (0 outside the tower)."""
        out = np.zeros((self.numBlocks, (self.nb + 1) * self.nb ** (self.ndim - 1)))
        p = np.ascontiguousarray(phi, dtype=np.float64)
        self._chk(self.lib.sy_mac_flux(self.h, mg_level, d, N.ptr(p), N.ptr(out)), "mg_flux_level")
        return out

    def setTolerance(self, rel, abs_):
        self._params.update(tol_rel=rel, tol_abs=abs_)

    def setMaxIteration(self, n):
        self._params["max_iter"] = n

    def setCGUseMGPrecond(self, on):
        self._params["precond"] = int(on)

    def setPreSmoothNum(self, n):
        self._params["pre"] = n

    def setPostSmoothNum(self, n):
        self._params["post"] = n

    def setPrecision(self, p):
        self._params["precision"] = p

    def _iter_solve(self, fn, name, *pre):
        p = self._params
        self._chk(self.lib.sy_mac_set_params(self.h, p["tol_rel"], p["tol_abs"], p["max_iter"], p["precond"],
                                             p["precision"], p["pre"], p["post"]), "setParams")
        it, st, r0, r = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        ret = self._chk(fn(self.h, *pre, C.byref(it), C.byref(st), C.byref(r0), C.byref(r)), name)
        self.stats = dict(iters=it.value, status=st.value, rnorm0=r0.value, rnorm=r.value)
        return ret

    def mg_solve(self, fixed_iter=0):
        """MacSolver::mg_solve (SayakaMacSolver.cpp:43-162); fixed_iter as setMGFixedIter."""
        return self._iter_solve(self.lib.sy_mac_mg_solve, "mg_solve", int(fixed_iter))

    def relax_solve(self):
        """MacSolver::relax_solve (SayakaMacSolver.cpp:377-452)."""
        return self._iter_solve(self.lib.sy_mac_relax_solve, "relax_solve")

    def cg_solve(self):
        p = self._params
        self._chk(self.lib.sy_mac_set_params(self.h, p["tol_rel"], p["tol_abs"], p["max_iter"], p["precond"],
                                             p["precision"], p["pre"], p["post"]), "setParams")
        it = C.c_int()
        st = C.c_int()
        r0 = C.c_double()
        r = C.c_double()
        ret = self._chk(self.lib.sy_mac_cg_solve(self.h, C.byref(it), C.byref(st), C.byref(r0), C.byref(r)),
                        "cg_solve")
        self.stats = dict(iters=it.value, status=st.value, rnorm0=r0.value, rnorm=r.value)
        return ret
