"""ctypes binding of the C oracle (oracle/lib/libsayaka_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libsayaka_oracle.so")
_lib = None
_i, _d, _p = C.c_int, C.c_double, C.c_void_p

# buffer components (SayakaMacSolver.h:34-53)
RHS, SOL, RESID, CORR, TMP, X0, X, R, RHAT, P, PHAT, S, SHAT, V, T = range(15)


def build(verbose: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["gcc", "-std=c99", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
           os.path.join(HERE, "sayaka_oracle.c"), "-o", LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    sig = {
        "oc_create": ([_i] * 5 + [_p] * 7, _p),
        "oc_destroy": ([_p], None),
        "oc_tower_size": ([_p, _i], _i),
        "oc_tower_export": ([_p, _i] + [_p] * 9, None),
        "oc_prepare": ([_p], None),
        "oc_set_coefb": ([_p, _p], None),
        "oc_set_ib_frac": ([_p, _p, _p], None),
        "oc_set_ib_ls": ([_p, _p, _i], None),
        "oc_get_coefs": ([_p, _p, _p, _p], None),
        "oc_set_bc": ([_p, _i, _p, _p, _p], None),
        "oc_fix_bc": ([_p, _p, _i, _p], None),
        "oc_set_params": ([_p, _d, _d, _i, _i], None),
        "oc_set_rhs": ([_p, _p], None),
        "oc_set_init_sol": ([_p, _p], None),
        "oc_set_zero_sol": ([_p], None),
        "oc_get_sol": ([_p, _p], None),
        "oc_buf_set": ([_p, _i, _p], None),
        "oc_buf_get": ([_p, _i, _p], None),
        "oc_save_get": ([_p, _i, _p], None),
        "oc_fill": ([_p, _i, _i], None),
        "oc_flux": ([_p, _i, _i, _p, _p], None),
        "oc_apply": ([_p, _i, _i, _i], None),
        "oc_residual": ([_p, _i, _i, _i, _i], None),
        "oc_smooth": ([_p, _i, _i, _i], None),
        "oc_restrict": ([_p, _i, _i], None),
        "oc_prolong": ([_p, _i, _i], None),
        "oc_dot": ([_p, _i, _i, _i], _d),
        "oc_norm": ([_p, _i, _i], _d),
        "oc_precond": ([_p, _i, _i, _i], _i),
        "oc_cg_level": ([_p, _i, _i, _i, _d, _d, _i, _i, _p, _p, _p, _i, _p], _i),
        "oc_cg_solve": ([_p, _p, _p, _p, _i, _p], _i),
        "oc_bottom_stats": ([_p, _i, _p, _p], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


class Oracle:
    """Reference-faithful CPU MacSolver over a tree given as exported arrays."""

    def __init__(self, tree: dict, spec):
        self.lib = load()
        self.spec = spec
        self.tree = tree
        self.n = int(tree["level"].shape[0])
        self.ndim, self.nb = spec.ndim, spec.nb
        self.ncell = self.nb ** self.ndim
        self.ngrown = (self.nb + 2) ** 2 * ((self.nb + 2) if self.ndim == 3 else 1)
        self.nface = [(self.nb + 1) * self.nb ** (self.ndim - 1)] * self.ndim
        self._keep = [np.ascontiguousarray(tree[k], dtype=np.int32) for k in
                      ("level", "parent", "which_child", "child", "neighbor", "node_type")]
        self._len = np.ascontiguousarray(tree["length"], dtype=np.float64)
        self.h = self.lib.oc_create(self.ndim, self.nb, self.n, spec.minlevel, spec.maxlevel,
                                    *[_ptr(a) for a in self._keep], _ptr(self._len))
        if not self.h:
            raise ValueError("oc_create failed")

    def __del__(self):
        try:
            if self.h:
                self.lib.oc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def prepare(self):
        self.lib.oc_prepare(self.h)

    def tower(self, m):
        n = self.lib.oc_tower_size(self.h, m)
        out = dict(blocks=np.zeros(n, np.int32), begin=np.zeros(m + 1, np.int32),
                   end=np.zeros(m + 1, np.int32), face_type=np.zeros((n, 6), np.int32),
                   face_neigh=np.zeros((n, 6), np.int32), fc_off=np.zeros((n, 6, 3), np.int32),
                   cf_neigh=np.zeros((n, 6, 4), np.int32), cf_off=np.zeros((n, 6, 4, 3), np.int32),
                   cf_box=np.zeros((n, 6, 4, 6), np.int32))
        self.lib.oc_tower_export(self.h, m, *[_ptr(out[k]) for k in ("blocks", "begin", "end", "face_type",
                                                                     "face_neigh", "fc_off", "cf_neigh",
                                                                     "cf_off", "cf_box")])
        return out

    def _parr(self, arrs):
        keep = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
        pa = (C.c_void_p * 3)(*([a.ctypes.data for a in keep] + [None] * (3 - len(keep))))
        return keep, pa

    def set_coefb(self, b):
        keep, pa = self._parr(b)
        self.lib.oc_set_coefb(self.h, pa)

    def set_ib_frac(self, vf, af):
        keep, pa = self._parr(af)
        v = np.ascontiguousarray(vf, dtype=np.float64)
        self.lib.oc_set_ib_frac(self.h, _ptr(v), pa)

    def set_ib_ls(self, ls_grown, ngrow):
        self.lib.oc_set_ib_ls(self.h, _ptr(np.ascontiguousarray(ls_grown, dtype=np.float64)), ngrow)

    def get_coefs(self):
        vf = np.zeros((self.n, self.ncell))
        b = [np.zeros((self.n, f)) for f in self.nface]
        a = [np.zeros((self.n, f)) for f in self.nface]
        kb, pb = self._parr(b)
        ka, pa = self._parr(a)
        self.lib.oc_get_coefs(self.h, _ptr(vf), pb, pa)
        return vf, kb, ka

    def set_bc(self, physbc, mathbc, vals):
        pb = np.array(physbc, np.int32)
        mb = np.array(mathbc, np.int32)
        vv = np.array(vals, np.float64)
        self.lib.oc_set_bc(self.h, pb.size, _ptr(pb), _ptr(mb), _ptr(vv))

    def fix_bc(self, phi_grown, ngrow, rhs):
        r = np.ascontiguousarray(rhs, dtype=np.float64).copy()
        self.lib.oc_fix_bc(self.h, _ptr(np.ascontiguousarray(phi_grown, dtype=np.float64)), ngrow, _ptr(r))
        return r

    def set_params(self, tol_rel, tol_abs, max_iter, precond=1):
        self.lib.oc_set_params(self.h, tol_rel, tol_abs, max_iter, precond)

    def set_rhs(self, rhs):
        self.lib.oc_set_rhs(self.h, _ptr(np.ascontiguousarray(rhs, dtype=np.float64)))

    def set_init_sol(self, sol):
        self.lib.oc_set_init_sol(self.h, _ptr(np.ascontiguousarray(sol, dtype=np.float64)))

    def set_zero_sol(self):
        self.lib.oc_set_zero_sol(self.h)

    def get_sol(self):
        out = np.zeros((self.n, self.ncell))
        self.lib.oc_get_sol(self.h, _ptr(out))
        return out

    def buf_set(self, comp, grown):
        self.lib.oc_buf_set(self.h, comp, _ptr(np.ascontiguousarray(grown, dtype=np.float64)))

    def buf_get(self, comp):
        out = np.zeros((self.n, self.ngrown))
        self.lib.oc_buf_get(self.h, comp, _ptr(out))
        return out

    def save_get(self, d):
        out = np.zeros((self.n, self.nface[d]))
        self.lib.oc_save_get(self.h, d, _ptr(out))
        return out

    def flux(self, m, d, phi):
        """MacSolver::mg_flux_level: face flux of direction d on tower m, [blk][face]."""
        out = np.zeros((self.n, self.nface[d]))
        self.lib.oc_flux(self.h, m, d, _ptr(np.ascontiguousarray(phi, dtype=np.float64)), _ptr(out))
        return out

    def fill(self, m, comp):
        self.lib.oc_fill(self.h, m, comp)

    def apply(self, m, outc, inc):
        self.lib.oc_apply(self.h, m, outc, inc)

    def residual(self, m, rc, xc, rhsc):
        self.lib.oc_residual(self.h, m, rc, xc, rhsc)

    def smooth(self, m, solc, rhsc):
        self.lib.oc_smooth(self.h, m, solc, rhsc)

    def restrict(self, cl, comp):
        self.lib.oc_restrict(self.h, cl, comp)

    def prolong(self, fl, comp):
        self.lib.oc_prolong(self.h, fl, comp)

    def dot(self, m, a, b):
        return self.lib.oc_dot(self.h, m, a, b)

    def norm(self, m, a):
        return self.lib.oc_norm(self.h, m, a)

    def precond(self, m, solc, rhsc):
        return self.lib.oc_precond(self.h, m, solc, rhsc)

    def cg_level(self, m, solc, rhsc, tol_rel, tol_abs, max_iter, precond, hist_cap=0):
        it = np.zeros(1, np.int32)
        conv = np.zeros(1, np.int32)
        nh = np.zeros(1, np.int32)
        hist = np.zeros(max(hist_cap, 1))
        stat = self.lib.oc_cg_level(self.h, m, solc, rhsc, tol_rel, tol_abs, max_iter, precond,
                                    _ptr(it), _ptr(conv), _ptr(hist), hist_cap, _ptr(nh))
        return dict(stat=stat, iters=int(it[0]), converged=bool(conv[0]),
                    hist=hist[:min(int(nh[0]), hist_cap)].copy())

    def cg_solve(self, hist_cap=20000):
        it = np.zeros(1, np.int32)
        conv = np.zeros(1, np.int32)
        nh = np.zeros(1, np.int32)
        hist = np.zeros(hist_cap)
        stat = self.lib.oc_cg_solve(self.h, _ptr(it), _ptr(conv), _ptr(hist), hist_cap, _ptr(nh))
        return dict(stat=stat, iters=int(it[0]), converged=bool(conv[0]),
                    hist=hist[:min(int(nh[0]), hist_cap)].copy())

    def bottom_stats(self, reset=False):
        calls = np.zeros(1, np.int32)
        its = np.zeros(1, np.int64)
        self.lib.oc_bottom_stats(self.h, int(reset), _ptr(calls), _ptr(its))
        return int(calls[0]), int(its[0])
