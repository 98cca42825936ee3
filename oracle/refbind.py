"""ctypes binding of oracle/_ref/libsayaka_ref{2,3}.so (the compiled reference + harness).

TEST INFRASTRUCTURE ONLY: imported by tests/ (pinning the oracle, writing golden vectors)
and by the cpu_baseline leg of bench.py.  Never imported by the product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_i = C.c_int
_d = C.c_double
_p = C.c_void_p


def lib_path(ndim: int) -> str:
    return os.path.join(HERE, "_ref", f"libsayaka_ref{ndim}.so")


def available(ndim: int = 3) -> bool:
    return os.path.exists(lib_path(ndim))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def load(ndim: int):
    if ndim in _LIBS:
        return _LIBS[ndim]
    lib = C.CDLL(lib_path(ndim))
    sig = {
        "ref_ndim": ([], _i),
        "ref_tree_create": ([_i, _i, _i, _i, _p, _p, _p, _p], _i),
        "ref_tree_numblocks": ([], _i),
        "ref_tree_finest": ([], _i),
        "ref_tree_set_flags": ([_p, _p, _i], None),
        "ref_tree_regrid": ([], _i),
        "ref_tree_export": ([_p] * 14, None),
        "ref_tree_level_count": ([_i], _i),
        "ref_tree_level_blocks": ([_i, _p], None),
        "ref_mac_create": ([_i], None),
        "ref_mac_prepare": ([], None),
        "ref_tower_size": ([_i], _i),
        "ref_tower_export": ([_i] + [_p] * 9, None),
        "ref_mac_set_coefb": ([_p, _p, _p], None),
        "ref_mac_set_ib_frac": ([_p, _p, _p, _p], None),
        "ref_mac_set_ib_ls": ([_p, _i], None),
        "ref_mac_get_coefs": ([_p] * 7, None),
        "ref_mac_set_bc": ([_i, _p, _p, _p], None),
        "ref_mac_fix_bc": ([_p, _i, _p], None),
        "ref_mac_set_rhs": ([_p], None),
        "ref_mac_set_init_sol": ([_p], None),
        "ref_mac_set_zero_sol": ([], None),
        "ref_mac_get_sol": ([_p], None),
        "ref_mac_set_params": ([_d, _d, _i, _i], None),
        "ref_mac_cg_solve": ([_p, _p, _p, _i, _p], _i),
        "ref_buf_ncomp": ([], _i),
        "ref_buf_set": ([_i, _p], None),
        "ref_buf_get": ([_i, _p], None),
        "ref_save_get": ([_i, _p], None),
        "ref_apply": ([_i, _i, _i], None),
        "ref_residual": ([_i, _i, _i, _i], None),
        "ref_smooth": ([_i, _i, _i], None),
        "ref_fill": ([_i, _i], None),
        "ref_mac_flux": ([_i, _i, _p, _p], None),
        "ref_state_create": ([_i, _i], None),
        "ref_state_set": ([_p], None),
        "ref_state_get": ([_p], None),
        "ref_tree_init_new_data": ([], None),
        "ref_mac_mg_solve": ([_i], None),
        "ref_mac_relax_solve": ([], None),
        "ref_state_create_general": ([_i, _i, _i, _p, _i, _p, _p, _p], None),
        "ref_state_set_grown": ([_p], None),
        "ref_state_get_grown": ([_p], None),
        "ref_fill_state": ([_i, _i, _p, _p, _i, _p, _p, _p, _i, _i, _i], None),
        "ref_fill_face_state": ([_i, _i, _i, _p, _p, _i, _i, _i], None),
        "ref_restrict": ([_i, _i], None),
        "ref_prolong": ([_i, _i], None),
        "ref_dot": ([_i, _i, _i], _d),
        "ref_norm": ([_i, _i], _d),
        "ref_precond": ([_i, _i, _i], _i),
        "ref_cg_level": ([_i, _i, _i, _d, _d, _i, _i, _p, _p], _i),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _LIBS[ndim] = lib
    return lib


# Solver buffer components (SayakaMacSolver.h:34-53)
RHS, SOL, RESID, CORR, TMP, X0, X, R, RHAT, P, PHAT, S, SHAT, V, T = range(15)


class RefTree:
    """Tree backend over the reference's AmrTree (see sayakaamr_amd.workloads)."""

    def __init__(self, ndim: int):
        self.ndim = ndim
        self.lib = load(ndim)
        self.spec = None

    def create(self, spec):
        self.spec = spec
        nd = np.array(spec.ndiv, dtype=np.int32)
        lo = np.array(spec.lo, dtype=np.float64)
        hi = np.array(spec.hi, dtype=np.float64)
        bc = np.array(spec.bc, dtype=np.int32)
        return self.lib.ref_tree_create(spec.nb, spec.ngrow, spec.minlevel, spec.maxlevel,
                                        _ptr(nd), _ptr(lo), _ptr(hi), _ptr(bc))

    def set_refine_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        self.lib.ref_tree_set_flags(_ptr(f), None, f.size)

    def set_coarsen_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        self.lib.ref_tree_set_flags(None, _ptr(f), f.size)

    def regrid(self):
        return self.lib.ref_tree_regrid()

    # a registered one-component cell state (injection FillPatch, Neumann walls)
    def state_create(self, ngrow, physbc):
        self.lib.ref_state_create(ngrow, physbc)

    def state_set(self, v):
        self.lib.ref_state_set(_ptr(np.ascontiguousarray(v, dtype=np.float64)))

    def state_get(self, ncell):
        out = np.zeros((self.lib.ref_tree_numblocks(), ncell))
        self.lib.ref_state_get(_ptr(out))
        return out

    def init_new_data(self):
        self.lib.ref_tree_init_new_data()

    # a general registered state: cell (dir < 0) or face data, per-component prolongation,
    # BC register (cell data) as for fill_state; accessed as [comp][block][grown box]
    def state_create_general(self, dir, ncomp, ngrow, prolong, physbc=(), mathbc=(), bcval=()):
        pr = np.ascontiguousarray(prolong, dtype=np.int32)
        pb = np.ascontiguousarray(physbc if len(physbc) else [0], dtype=np.int32)
        nphys = len(physbc)
        mb = np.ascontiguousarray(mathbc if nphys else [0], dtype=np.int32)
        bv = np.ascontiguousarray(bcval if nphys else [0.0], dtype=np.float64)
        self._gstate = (dir, ncomp, ngrow)
        self.lib.ref_state_create_general(dir, ncomp, ngrow, _ptr(pr), nphys, _ptr(pb), _ptr(mb), _ptr(bv))

    def _gshape(self, nb):
        dir, ncomp, ngrow = self._gstate
        ext = [nb + 2 * ngrow + (dir == d) for d in range(3)]
        if self.ndim == 2:
            ext[2] = 1
        return (ncomp, self.lib.ref_tree_numblocks(), ext[2], ext[1], ext[0])

    def state_set_grown(self, data, nb):
        d = np.ascontiguousarray(data, dtype=np.float64)
        assert d.shape == self._gshape(nb), (d.shape, self._gshape(nb))
        self.lib.ref_state_set_grown(_ptr(d))

    def state_get_grown(self, nb):
        out = np.zeros(self._gshape(nb))
        self.lib.ref_state_get_grown(_ptr(out))
        return out

    # FillPatch::fillBoundary on a [comp][block][grown] state (a fresh copy is returned)
    def fill_state(self, data, ngrow, prolong, physbc, mathbc, bcval, scomp, nc, nlayer):
        d = np.array(data, dtype=np.float64, order="C", copy=True)
        ncomp = d.shape[0]
        pr = np.ascontiguousarray(prolong, dtype=np.int32)
        pb = np.ascontiguousarray(physbc, dtype=np.int32)
        mb = np.ascontiguousarray(mathbc, dtype=np.int32).reshape(ncomp, pb.size)
        bv = np.ascontiguousarray(bcval, dtype=np.float64).reshape(ncomp, pb.size)
        self.lib.ref_fill_state(ncomp, ngrow, _ptr(d), _ptr(pr), pb.size, _ptr(pb), _ptr(mb), _ptr(bv),
                                scomp, nc, nlayer)
        return d

    # FillPatch::fillBoundary on a face state of direction dir, [comp][block][grown face box]
    def fill_face_state(self, dir, data, ngrow, prolong, scomp, nc, nlayer):
        d = np.array(data, dtype=np.float64, order="C", copy=True)
        pr = np.ascontiguousarray(prolong, dtype=np.int32)
        self.lib.ref_fill_face_state(dir, d.shape[0], ngrow, _ptr(d), _ptr(pr), scomp, nc, nlayer)
        return d

    def export(self):
        n = self.lib.ref_tree_numblocks()
        t = dict(level=np.zeros(n, np.int32), parent=np.zeros(n, np.int32),
                 which_child=np.zeros(n, np.int32), child=np.zeros((n, 8), np.int32),
                 neighbor=np.zeros((n, 6), np.int32), node_type=np.zeros(n, np.int32),
                 node_type2=np.zeros((n, 6), np.int32), lo=np.zeros((n, 3)), hi=np.zeros((n, 3)),
                 center=np.zeros((n, 3)), length=np.zeros((n, 3)),
                 surr=np.zeros((n, 27), np.int32), surr_type=np.zeros((n, 27), np.int32),
                 surr_bcref=np.zeros((n, 27), np.int32))
        self.lib.ref_tree_export(*[_ptr(t[k]) for k in ("level", "parent", "which_child", "child",
                                                        "neighbor", "node_type", "node_type2", "lo",
                                                        "hi", "center", "length", "surr",
                                                        "surr_type", "surr_bcref")])
        return t

    def level_blocks(self, l):
        n = self.lib.ref_tree_level_count(l)
        out = np.zeros(n, np.int32)
        if n:
            self.lib.ref_tree_level_blocks(l, _ptr(out))
        return out


class RefMac:
    """The reference MacSolver over the current RefTree."""

    def __init__(self, tree: RefTree, verbose: int = 0):
        self.t = tree
        self.lib = tree.lib
        self.lib.ref_mac_create(verbose)

    def prepare(self):
        self.lib.ref_mac_prepare()

    def tower(self, m):
        n = self.lib.ref_tower_size(m)
        out = dict(blocks=np.zeros(n, np.int32), begin=np.zeros(m + 1, np.int32),
                   end=np.zeros(m + 1, np.int32), face_type=np.zeros((n, 6), np.int32),
                   face_neigh=np.zeros((n, 6), np.int32), fc_off=np.zeros((n, 6, 3), np.int32),
                   cf_neigh=np.zeros((n, 6, 4), np.int32), cf_off=np.zeros((n, 6, 4, 3), np.int32),
                   cf_box=np.zeros((n, 6, 4, 6), np.int32))
        self.lib.ref_tower_export(m, *[_ptr(out[k]) for k in ("blocks", "begin", "end", "face_type",
                                                              "face_neigh", "fc_off", "cf_neigh",
                                                              "cf_off", "cf_box")])
        return out

    def set_coefb(self, b):
        bs = [np.ascontiguousarray(x, dtype=np.float64) for x in b] + [None] * (3 - len(b))
        self.lib.ref_mac_set_coefb(*[_ptr(x) for x in bs])

    def set_ib_frac(self, vf, af):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in af] + [None] * (3 - len(af))
        self.lib.ref_mac_set_ib_frac(_ptr(np.ascontiguousarray(vf, dtype=np.float64)), *[_ptr(x) for x in a])

    def set_ib_ls(self, ls_grown, ngrow):
        self.lib.ref_mac_set_ib_ls(_ptr(np.ascontiguousarray(ls_grown, dtype=np.float64)), ngrow)

    def get_coefs(self, nb):
        t = self.t.export()
        n = t["level"].shape[0]
        nd = self.t.ndim
        nc = nb ** nd
        vf = np.zeros((n, nc))
        fsz = [(nb + 1) * nb ** (nd - 1)] * nd
        b = [np.zeros((n, fsz[d])) for d in range(nd)] + [None] * (3 - nd)
        a = [np.zeros((n, fsz[d])) for d in range(nd)] + [None] * (3 - nd)
        self.lib.ref_mac_get_coefs(_ptr(vf), *[_ptr(x) for x in b], *[_ptr(x) for x in a])
        return vf, b[:nd], a[:nd]

    def set_bc(self, physbc, mathbc, vals):
        pb = np.array(physbc, np.int32)
        mb = np.array(mathbc, np.int32)
        vv = np.array(vals, np.float64)
        self.lib.ref_mac_set_bc(pb.size, _ptr(pb), _ptr(mb), _ptr(vv))

    def fix_bc(self, phi_grown, ngrow, rhs):
        r = np.ascontiguousarray(rhs, dtype=np.float64).copy()
        self.lib.ref_mac_fix_bc(_ptr(np.ascontiguousarray(phi_grown, dtype=np.float64)), ngrow, _ptr(r))
        return r

    def set_rhs(self, rhs):
        self.lib.ref_mac_set_rhs(_ptr(np.ascontiguousarray(rhs, dtype=np.float64)))

    def set_init_sol(self, sol):
        self.lib.ref_mac_set_init_sol(_ptr(np.ascontiguousarray(sol, dtype=np.float64)))

    def set_zero_sol(self):
        self.lib.ref_mac_set_zero_sol()

    def get_sol(self, n, ncell):
        out = np.zeros((n, ncell))
        self.lib.ref_mac_get_sol(_ptr(out))
        return out

    def set_params(self, tol_rel, tol_abs, max_iter, precond=1):
        self.lib.ref_mac_set_params(tol_rel, tol_abs, max_iter, precond)

    def mg_solve(self, fixed_iter=0):
        self.lib.ref_mac_mg_solve(fixed_iter)

    def relax_solve(self):
        self.lib.ref_mac_relax_solve()

    def cg_solve(self, hist_cap=20000):
        it = np.zeros(1, np.int32)
        conv = np.zeros(1, np.int32)
        nh = np.zeros(1, np.int32)
        hist = np.zeros(hist_cap)
        stat = self.lib.ref_mac_cg_solve(_ptr(it), _ptr(conv), _ptr(hist), hist_cap, _ptr(nh))
        return dict(stat=stat, iters=int(it[0]), converged=bool(conv[0]), hist=hist[:min(nh[0], hist_cap)].copy())

    # component-level access (grown layout, one ghost)
    def buf_set(self, comp, grown):
        self.lib.ref_buf_set(comp, _ptr(np.ascontiguousarray(grown, dtype=np.float64)))

    def buf_get(self, comp, n, ngrown):
        out = np.zeros((n, ngrown))
        self.lib.ref_buf_get(comp, _ptr(out))
        return out

    def flux(self, m, d, phi, nb):
        """MacSolver::mg_flux_level on tower m, direction d: [blk][face]."""
        n = phi.shape[0]
        nd = self.t.ndim
        out = np.zeros((n, (nb + 1) * nb ** (nd - 1)))
        self.lib.ref_mac_flux(m, d, _ptr(np.ascontiguousarray(phi, dtype=np.float64)), _ptr(out))
        return out

    def save_get(self, d, n, nface):
        out = np.zeros((n, nface))
        self.lib.ref_save_get(d, _ptr(out))
        return out

    def cg_level(self, m, solcomp, rhscomp, tol_rel, tol_abs, max_iter, precond):
        it = np.zeros(1, np.int32)
        conv = np.zeros(1, np.int32)
        stat = self.lib.ref_cg_level(m, solcomp, rhscomp, tol_rel, tol_abs, max_iter, precond,
                                     _ptr(it), _ptr(conv))
        return dict(stat=stat, iters=int(it[0]), converged=bool(conv[0]))
