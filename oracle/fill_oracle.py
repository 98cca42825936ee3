"""CPU restatement of the reference's general ghost fill of a cell-centred state.

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product package.  It checks
the device state fill (sayaka_dev_state_fill) on the same tree and the same data.

What it restates (file:line in the reference checkout's src/):
  FillPatch::fillBoundary / fillLevelBoundary / fillBlockBoundary
      SayakaFillPatch.cpp:17-89      levels 1..maxAmrLevel, every block of a level in
                                     ascending id; per block: same-level copy, coarse
                                     prolongation, physical BC
  FillPatch::fillBlockBndryAtSameLevel   SayakaFillPatch.cpp:92-190
  FillPatch::fillBlockBndryFromCoarseLevel  SayakaFillPatch.cpp:192-299
  FillPatch::fillBlockBndryCond     SayakaFillPatch.cpp:301-368 (sweep order, bcref)
  FillPatch::cacheRanges            SayakaFillPatch.cpp:620-720 (cell data)
  InterpPatch::prolongCellBlockData_Injection / _CenteredLinear / _CenteredLinearLimited
      SayakaInterpPatch.cpp:157-189, 228-391; limited_slope SayakaInterpPatch.h:214-233
  BCPatch::fillBasicCellBlockBC     SayakaBoundaryPatch.cpp:139-229 (reflection over the
                                     data's full ngrow layers, whatever nlayer is)
  IndexMapping::FineToCrse / CrseToFine  SayakaTreeData.h:34-44
Face data (FaceFill):
  FillPatch::cacheRanges            SayakaFillPatch.cpp:660-700 (staggered same-level ranges,
                                     non-overlapping coarse ranges)
  InterpPatch::prolongFaceBlockData_Injection / _CenteredLinear (with or without limiter)
      SayakaInterpPatch.cpp:58-85, 447-591
  fillBlockBndryCond exits for face data next to a wall (SayakaFillPatch.cpp:318-324): the
  restatement covers periodic trees only and raises otherwise.

Pinning: both classes reproduce the reference's own fillBoundary output bit for bit
(tests/golden/fill{2,3}.npz, fface{2,3}.npz, written by the reference compiled from its
sources, oracle/_ref); the cell prolongations also reproduce linear fields and constants
(tests/test_fill_oracle.py).

Layout: state[comp] is [block][gz][gy][gx] over the block grown by ngrow (gz = 1 in 2D);
cell (i, j, k) of the reference's index space sits at [k+g][j+g][i+g].  Face data of
direction fd has one more entry along fd (faces 0..nb of the valid box).
"""
from __future__ import annotations

import numpy as np

PROLONG_INJECTION, PROLONG_CENTERED_LINEAR, PROLONG_CENTERED_LINEAR_LIMITED = 0, 1, 2
NEUMANN, DIRICHLET, SIMPLEFILL = 0, 1, 2
PHYS_BNDRY = -20
LARGE = 64


def fine_to_crse(i, off):
    """IndexMapping::FineToCrse (SayakaTreeData.h:41-44)."""
    return (i + LARGE) // 2 - LARGE // 2 + off


def crse_to_fine(ic, sub, off):
    """IndexMapping::CrseToFine (SayakaTreeData.h:34-36)."""
    return (ic - off) * 2 + sub


def surr_flat(ii, jj, kk):
    """Position of surrounding (ii, jj, kk) in the exported [ii+1][jj+1][kk+1] arrays."""
    return (ii + 1) * 9 + (jj + 1) * 3 + (kk + 1)


def surr_decode(idx, ndim):
    """SurroundingIndex -> (ipos, jpos, kpos) (SayakaSurroundingIndex.h:44-58)."""
    return idx % 3 - 1, (idx // 3) % 3 - 1, (idx // 9 - 1) if ndim == 3 else 0


def std_min(a, b):
    return b if b < a else a


def std_max(a, b):
    return b if a < b else a


def limited_slope(ul, uc, ur):
    """MC-limited centred slope (SayakaInterpPatch.h:214-233)."""
    gradc = 0.5 * (ur - ul)
    gradl = uc - ul
    gradr = ur - uc
    signc = 0.0
    if gradc > 0:
        signc = 1.0
    elif gradc < 0:
        signc = -1.0
    grad = abs(gradc)
    grad = std_min(grad, 2.0 * signc * gradl)
    grad = std_min(grad, 2.0 * signc * gradr)
    grad = std_max(grad, 0.0)
    grad *= signc
    return grad


class StateFill:
    """FillPatch + BCPatch over one tree (tree dict as exported by the tree backends)."""

    def __init__(self, tree, nb, ndim, ngrow, maxlevel):
        self.t = tree
        self.nb, self.nd, self.g, self.maxlevel = nb, ndim, ngrow, maxlevel
        self.nz = nb if ndim == 3 else 1
        self.gz = ngrow if ndim == 3 else 0

    # cell box helpers: (lo[3], hi[3]) inclusive
    def _valid(self):
        return [0, 0, 0], [self.nb - 1, self.nb - 1, self.nz - 1]

    def _extend(self, box, e):
        lo, hi = box
        ez = e if self.nd == 3 else 0
        ext = [e, e, ez]
        return [lo[d] - ext[d] for d in range(3)], [hi[d] + ext[d] for d in range(3)]

    @staticmethod
    def _shift(box, v):
        lo, hi = box
        return [lo[d] + v[d] for d in range(3)], [hi[d] + v[d] for d in range(3)]

    @staticmethod
    def _inter(a, b):
        return [max(a[0][d], b[0][d]) for d in range(3)], [min(a[1][d], b[1][d]) for d in range(3)]

    def ranges(self, ii, jj, kk, nlayer):
        """cacheRanges (SayakaFillPatch.cpp:620-690): ghost box of surrounding (ii,jj,kk) and the
        neighbour's valid box it copies from."""
        vb = self._valid()
        layer = self._extend(vb, nlayer)
        v = [ii * self.nb, jj * self.nb, kk * self.nz]
        dst = self._inter(self._shift(vb, v), layer)
        src = self._inter(self._shift(layer, [-x for x in v]), vb)
        return dst, src

    def _view(self, arr, b, box):
        lo, hi = box
        g, gz = self.g, self.gz
        return arr[b, lo[2] + gz:hi[2] + gz + 1, lo[1] + g:hi[1] + g + 1, lo[0] + g:hi[0] + g + 1]

    def _at(self, arr, b, i, j, k):
        return arr[b, k + self.gz, j + self.g, i + self.g]

    def _set(self, arr, b, i, j, k, v):
        arr[b, k + self.gz, j + self.g, i + self.g] = v

    def surr(self, b, ii, jj, kk):
        return int(self.t["surr"][b, surr_flat(ii, jj, kk)])

    def _surr_iter(self):
        zr = range(-1, 2) if self.nd == 3 else range(0, 1)
        for kk in zr:
            for jj in range(-1, 2):
                for ii in range(-1, 2):
                    if ii == 0 and jj == 0 and kk == 0:
                        continue
                    yield ii, jj, kk

    # -------------------------------------------------------------------- prolongation
    def _prolong_box(self, arr, b, p, box, off, method):
        lo, hi = box
        nd = self.nd
        clo = [fine_to_crse(lo[d], off[d]) if d < nd else lo[d] for d in range(3)]
        chi = [fine_to_crse(hi[d], off[d]) if d < nd else hi[d] for d in range(3)]
        if method == PROLONG_INJECTION:
            for k in range(lo[2], hi[2] + 1):
                kc = fine_to_crse(k, off[2]) if nd == 3 else k
                for j in range(lo[1], hi[1] + 1):
                    jc = fine_to_crse(j, off[1])
                    for i in range(lo[0], hi[0] + 1):
                        ic = fine_to_crse(i, off[0])
                        self._set(arr, b, i, j, k, self._at(arr, p, ic, jc, kc))
            return
        zsubs = (-1, 1) if nd == 3 else (0,)
        for kc in range(clo[2], chi[2] + 1):
            for jc in range(clo[1], chi[1] + 1):
                for ic in range(clo[0], chi[0] + 1):
                    c = self._at(arr, p, ic, jc, kc)
                    grad = [0.0, 0.0, 0.0]
                    for d in range(nd):
                        e = [0, 0, 0]
                        e[d] = 1
                        up = self._at(arr, p, ic + e[0], jc + e[1], kc + e[2])
                        dn = self._at(arr, p, ic - e[0], jc - e[1], kc - e[2])
                        if method == PROLONG_CENTERED_LINEAR:
                            grad[d] = 0.5 * (up - dn)
                        else:
                            grad[d] = limited_slope(dn, c, up)
                    for kk in zsubs:
                        k = crse_to_fine(kc, (kk + 1) // 2, off[2]) if nd == 3 else kc
                        krel = 0.25 * float(kk)
                        if k < lo[2] or k > hi[2]:
                            continue
                        for jj in (-1, 1):
                            j = crse_to_fine(jc, (jj + 1) // 2, off[1])
                            jrel = 0.25 * float(jj)
                            if j < lo[1] or j > hi[1]:
                                continue
                            for ii in (-1, 1):
                                i = crse_to_fine(ic, (ii + 1) // 2, off[0])
                                irel = 0.25 * float(ii)
                                if i < lo[0] or i > hi[0]:
                                    continue
                                v = c + irel * grad[0] + jrel * grad[1] + krel * grad[2]
                                self._set(arr, b, i, j, k, v)

    # ------------------------------------------------------------------------- BC fill
    def _bc_box(self, ii, jj, kk):
        """BCPatch::calcRanges (SayakaBoundaryPatch.cpp:52-78): over the data's ngrow layers."""
        vb = self._valid()
        layer = self._extend(vb, self.g)
        return self._inter(self._shift(vb, [ii * self.nb, jj * self.nb, kk * self.nz]), layer)

    def _bc_one(self, arr, b, comp, sweepdir, xb, yb, zb, bcmap):
        phys = self.surr(b, xb, yb, zb)
        mathbc, bcval = bcmap[comp][phys]
        nb, nz = self.nb, self.nz
        n3 = [nb, nb, nz]
        bnd = [xb, yb, zb]
        rdir = sweepdir
        ref = int(self.t["surr_bcref"][b, surr_flat(xb, yb, zb)])
        if ref >= 0:
            rp = surr_decode(ref, self.nd)
            rdir = -1
            for d in range(3):
                if bnd[d] != rp[d]:
                    rdir = d
        sign = [(-1 if rdir == d else 1) for d in range(3)]
        off = [0 if rdir != d else (-1 if bnd[d] == -1 else 2 * n3[d] - 1) for d in range(3)]
        lo, hi = self._bc_box(xb, yb, zb)
        for k in range(lo[2], hi[2] + 1):
            kr = sign[2] * k + off[2]
            for j in range(lo[1], hi[1] + 1):
                jr = sign[1] * j + off[1]
                for i in range(lo[0], hi[0] + 1):
                    ir = sign[0] * i + off[0]
                    vref = self._at(arr, b, ir, jr, kr)
                    if mathbc == NEUMANN:
                        v = vref
                    elif mathbc == DIRICHLET:
                        v = 2.0 * bcval - vref
                    elif mathbc == SIMPLEFILL:
                        v = bcval
                    else:
                        continue
                    self._set(arr, b, i, j, k, v)

    def bc_order(self):
        """The (sweep, sweepdir, surrounding) sequence of fillBlockBndryCond
        (SayakaFillPatch.cpp:326-360)."""
        nd = self.nd
        out = []
        for sweep in range(nd):
            sweepdir = sweep % nd
            lo, hi, inc = [0, 0, 0], [0, 0, 0], [1, 1, 1]
            for d in range(sweep + 1):
                lo[(d + sweep) % nd] = -1
                hi[(d + sweep) % nd] = 1
            inc[sweepdir] = 2
            for kk in range(lo[2], hi[2] + 1, inc[2]):
                for jj in range(lo[1], hi[1] + 1, inc[1]):
                    for ii in range(lo[0], hi[0] + 1, inc[0]):
                        out.append((sweep, sweepdir, ii, jj, kk))
        return out

    # ---------------------------------------------------------------------- the fill
    def fill(self, state, scomp, nc, nlayer, prolong, bcmap):
        """fillBoundary(scomp, nc, nlayer) on state (list of per-comp arrays, modified in place).
        prolong[c]: method of comp c; bcmap[c][physcode] = (math BC, value)."""
        if nlayer <= 0:
            return
        t = self.t
        for level in range(1, self.maxlevel + 1):
            blocks = np.nonzero(t["level"] == level)[0]
            for b in blocks:
                self.fill_block(state, int(b), scomp, nc, nlayer, prolong, bcmap)

    def fill_block(self, state, b, scomp, nc, nlayer, prolong, bcmap, same_level_only=False):
        """fillBlockBoundary (SayakaFillPatch.cpp:54-89) with onlySameLevel as given, physical BC
        applied and corners filled."""
        t = self.t
        for c in range(scomp, scomp + nc):
            arr = state[c]
            # same level (SayakaFillPatch.cpp:125-188)
            for ii, jj, kk in self._surr_iter():
                nbk = self.surr(b, ii, jj, kk)
                if nbk >= 0:
                    dst, src = self.ranges(ii, jj, kk, nlayer)
                    self._view(arr, b, dst)[...] = self._view(arr, nbk, src)
            # coarse level (SayakaFillPatch.cpp:225-297)
            coarse = [s for s in self._surr_iter() if PHYS_BNDRY < self.surr(b, *s) < 0]
            if coarse and not same_level_only:
                p = int(t["parent"][b])
                off = self._offset(b)
                for ii, jj, kk in coarse:
                    dst, _ = self.ranges(ii, jj, kk, nlayer)
                    self._prolong_box(arr, b, p, dst, off, prolong[c])
            # physical boundary (SayakaFillPatch.cpp:301-368)
            if any(self.surr(b, *s) <= PHYS_BNDRY for s in self._surr_iter()):
                for sweep, sweepdir, ii, jj, kk in self.bc_order():
                    if self.surr(b, ii, jj, kk) <= PHYS_BNDRY:
                        self._bc_one(arr, b, c, sweepdir, ii, jj, kk, bcmap)

    def _offset(self, b):
        """The block's offset in its parent's cells (ChildIndexOffset)."""
        wc = int(self.t["which_child"][b])
        return [((wc >> d) & 1) * (self.nb // 2) if d < self.nd else 0 for d in range(3)]

    def prolong_block(self, state, b, prolong):
        """FillPatch::prolongBlock (SayakaFillPatch.cpp:464-483): the valid box from the parent."""
        p = int(self.t["parent"][b])
        for c, method in enumerate(prolong):
            self._prolong_box(state[c], b, p, self._valid(), self._offset(b), method)


class FaceFill(StateFill):
    """FillPatch::fillBoundary on face data of direction fdir (staggered in fdir): the
    reference's same-level copy over staggered ranges (they overlap on shared faces, the
    surrounding later in loop order wins, SayakaFillPatch.cpp:125-188, 660-700), coarse
    prolongation over the non-overlapping ranges with prolongFaceBlockData_Injection /
    _CenteredLinear (SayakaInterpPatch.cpp:447-591, one surrounding after the other: the middle
    faces of the second sweep can read faces outside the box), and no physical BC: the reference
    exits for face data next to a wall (SayakaFillPatch.cpp:318-324), so the tree is periodic.

    state[comp]: [block][gz][gy][gx] over the face box grown by ngrow (one more face in fdir);
    face (i, j, k) sits at [k+gz][j+g][i+g]."""

    def __init__(self, tree, nb, ndim, ngrow, maxlevel, fdir):
        super().__init__(tree, nb, ndim, ngrow, maxlevel)
        self.fd = fdir

    def _stagger(self, box):
        lo, hi = list(box[0]), list(box[1])
        hi[self.fd] += 1
        return lo, hi

    def ranges(self, ii, jj, kk, nlayer):
        dst, src = super().ranges(ii, jj, kk, nlayer)
        return self._stagger(dst), self._stagger(src)

    def nonoverlap(self, ii, jj, kk, nlayer):
        dst, _ = self.ranges(ii, jj, kk, nlayer)
        lo, hi = dst
        o = (ii, jj, kk)[self.fd]
        if o == -1:
            hi[self.fd] -= 1
        elif o == 1:
            lo[self.fd] += 1
        return lo, hi

    def _prolong_face(self, arr, b, p, box, off, method):
        lo, hi = box
        nd, fd = self.nd, self.fd
        e = [0, 0, 0]
        e[fd] = 1
        ftc = lambda v, d: fine_to_crse(v, off[d]) if d < nd else v
        if method == PROLONG_INJECTION:
            for k in range(lo[2], hi[2] + 1):
                for j in range(lo[1], hi[1] + 1):
                    for i in range(lo[0], hi[0] + 1):
                        ic, jc, kc = ftc(i, 0), ftc(j, 1), ftc(k, 2)
                        side = ((i, j, k)[fd] + LARGE) % 2
                        vl = self._at(arr, p, ic, jc, kc)
                        if side == 0:
                            v = vl
                        else:
                            v = 0.5 * (vl + self._at(arr, p, ic + e[0], jc + e[1], kc + e[2]))
                        self._set(arr, b, i, j, k, v)
            return
        clo = [ftc(lo[d], d) for d in range(3)]
        chi = [ftc(hi[d], d) for d in range(3)]
        rng = [(-1, 1) if (d != fd and d < nd) else (0,) for d in range(3)]
        for kc in range(clo[2], chi[2] + 1):
            for jc in range(clo[1], chi[1] + 1):
                for ic in range(clo[0], chi[0] + 1):
                    c = self._at(arr, p, ic, jc, kc)
                    grad = [0.0, 0.0, 0.0]
                    for d in range(nd):
                        if d == fd:
                            continue
                        u = [0, 0, 0]
                        u[d] = 1
                        dn = self._at(arr, p, ic - u[0], jc - u[1], kc - u[2])
                        up = self._at(arr, p, ic + u[0], jc + u[1], kc + u[2])
                        grad[d] = limited_slope(dn, c, up) if method == PROLONG_CENTERED_LINEAR_LIMITED \
                            else 0.5 * (up - dn)
                    cc = (ic, jc, kc)
                    for kind in rng[2]:
                        k = crse_to_fine(kc, (kind + 1) // 2, off[2]) if nd == 3 else kc
                        if k < lo[2] or k > hi[2]:
                            continue
                        krel = 0.25 * float(kind)
                        for jind in rng[1]:
                            j = crse_to_fine(jc, (jind + 1) // 2, off[1])
                            if j < lo[1] or j > hi[1]:
                                continue
                            jrel = 0.25 * float(jind)
                            for iind in rng[0]:
                                i = crse_to_fine(ic, (iind + 1) // 2, off[0])
                                if i < lo[0] or i > hi[0]:
                                    continue
                                irel = 0.25 * float(iind)
                                self._set(arr, b, i, j, k, c + irel * grad[0] + jrel * grad[1] + krel * grad[2])
        # second sweep: faces between two coarse faces, the average of their two fine neighbours
        for k in range(lo[2], hi[2] + 1):
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    if ((i, j, k)[fd] + LARGE) % 2 == 1:
                        vl = self._at(arr, b, i - e[0], j - e[1], k - e[2])
                        vr = self._at(arr, b, i + e[0], j + e[1], k + e[2])
                        self._set(arr, b, i, j, k, 0.5 * (vl + vr))

    def fill(self, state, scomp, nc, nlayer, prolong, bcmap=None):
        if nlayer <= 0:
            return
        t = self.t
        for level in range(1, self.maxlevel + 1):
            for b in np.nonzero(t["level"] == level)[0]:
                self.fill_block(state, int(b), scomp, nc, nlayer, prolong)

    def fill_block(self, state, b, scomp, nc, nlayer, prolong, bcmap=None, same_level_only=False):
        if any(self.surr(b, *s) <= PHYS_BNDRY for s in self._surr_iter()):
            raise ValueError("face data next to a wall (the reference exits)")
        for c in range(scomp, scomp + nc):
            arr = state[c]
            for ii, jj, kk in self._surr_iter():
                nbk = self.surr(b, ii, jj, kk)
                if nbk >= 0:
                    dst, src = self.ranges(ii, jj, kk, nlayer)
                    self._view(arr, b, dst)[...] = self._view(arr, nbk, src)
            coarse = [s for s in self._surr_iter() if PHYS_BNDRY < self.surr(b, *s) < 0]
            if coarse and not same_level_only:
                p = int(self.t["parent"][b])
                for ii, jj, kk in coarse:
                    self._prolong_face(arr, b, p, self.nonoverlap(ii, jj, kk, nlayer), self._offset(b), prolong[c])

    def prolong_block(self, state, b, prolong):
        p = int(self.t["parent"][b])
        for c, method in enumerate(prolong):
            self._prolong_face(state[c], b, p, self._stagger(self._valid()), self._offset(b), method)

    def correct_faces(self, state, b, is_new):
        """The faces of new block b shared with existing same-level face neighbours take the
        neighbour's copy (correctFaceBlockData, SayakaFillPatch.cpp:485-513, as called by
        SayakaRegrid.cpp:1078-1100)."""
        fd = self.fd
        for side in (0, 1):
            j = int(self.t["neighbor"][b, 2 * fd + side])
            if j < 0 or is_new[j]:
                continue
            lo, hi = self._stagger(self._valid())
            dlo, dhi, slo, shi = list(lo), list(hi), list(lo), list(hi)
            dlo[fd] = dhi[fd] = lo[fd] if side == 0 else hi[fd]
            slo[fd] = shi[fd] = hi[fd] if side == 0 else lo[fd]
            for arr in state:
                self._view(arr, b, (dlo, dhi))[...] = self._view(arr, j, (slo, shi))


def remap_state(tree, origin, pre, nb, ndim, ngrow, maxlevel, prolong, bcmap=None, fdir=None):
    """A registered state after AmrTree::regrid and initNewBlockDataAfterRegrid.

    pre: [comp][old block][grown box] (cell, or face data of direction fdir); origin: per block
    of the new tree its id before the regrid, -1 for a block the regrid created; tree: the new
    tree.  Packing (SayakaRegrid.cpp:530-591) moves each surviving block's whole grown box; a new
    block starts at 0.  Then, level by level from the coarsest level holding a new block
    (SayakaRegrid.cpp:1013-1103): every non-leaf block of the level below with a new child,
    in ascending id, is filled in-level over all ngrow layers with the physical BC
    (fillBlockBoundary(.., onlySameLevel = 1)); every new block of the level, in ascending id,
    is prolonged from its parent over its valid box (prolongBlock) and, for face data, takes
    the shared faces of its existing same-level face neighbours (correctFaceBlockData)."""
    n = origin.shape[0]
    ncomp = pre.shape[0]
    post = np.zeros((ncomp, n) + pre.shape[2:])
    keep = origin >= 0
    post[:, keep] = pre[:, origin[keep]]
    new = ~keep
    if not new.any():
        return post
    ff = (StateFill(tree, nb, ndim, ngrow, maxlevel) if fdir is None
          else FaceFill(tree, nb, ndim, ngrow, maxlevel, fdir))
    st = [post[c] for c in range(ncomp)]
    lv = tree["level"]
    for level in range(int(lv[new].min()), int(lv[new].max()) + 1):
        for p in np.nonzero(lv == level - 1)[0]:
            if tree["node_type"][p] == 1 or not new[tree["child"][p][: 1 << ndim]].any():
                continue
            ff.fill_block(st, int(p), 0, ncomp, ngrow, prolong, bcmap, same_level_only=True)
        for b in np.nonzero((lv == level) & new)[0]:
            ff.prolong_block(st, int(b), prolong)
            if fdir is not None:
                ff.correct_faces(st, int(b), new)
    return post
