"""Synthetic workloads for the MAC Poisson solve: trees, coefficients and right-hand sides.

The same functions drive every tree backend (the product's host tree, the oracle tables and,
in tests, the compiled reference), so every backend sees bit-identical refine flags and inputs.

A backend is any object with
    create(spec)              -> AmrTree::init + initUniformRootLevel + initRefineToMinLevel
    export()                  -> dict of numpy arrays (see ``TREE_FIELDS``)
    set_refine_flags(flags)   -> AmrTree::refineFlag[0:n]
    regrid()                  -> AmrTree::regrid
Driver loop follows the reference demos (Test6.h:190-205): for each level, flag the leaves
and parents-of-leaves of that level whose cell centres meet the criterion, then regrid.

Workloads (sizes and choices):
  C0  2D quadtree on [0,1]^2, 8x8 cells/block, root 4x4, minAmrLevel 2, maxAmrLevel 3; refine
      where |r-0.25| < 0.05 at any cell centre, r = |x-(0.5,0.5)|; homogeneous Dirichlet
      walls; rhs = 2 pi^2 sin(pi x) sin(pi y) (the reference operator is -div(B grad));
      exact p = sin(pi x) sin(pi y).
  C1  3D octree on [0,1]^3, 8^3 cells/block, root 4^3, minAmrLevel 2, maxAmrLevel 4 (four
      tree levels: 4264 leaves, 2.18M cells); refine where | |x-c| - r | < SHELL_FACTOR * (block
      edge at that level) at any cell centre, c=(0.5,0.5,0.5), r=0.2.  The sphere is an
      immersed solid: cut-cell volume fractions vf and face apertures a_f are fluid
      fractions from 4^d sub-samples per cell (face); cells with vf < VF_MIN become solid
      and every face touching a solid cell is closed.  B = 1/rho at face centres with
      rho = 1 + 999 H_eps(-phi), phi = |x-c| - r, eps = 1.5 dx_finest.  Neumann on all walls
      (and, through a_f = 0, on the body).  rhs = discrete divergence of a face velocity
      u ~ U(-1,1) from a counter-based hash of (seed, level, global face index, dir),
      projected to zero volume-weighted mean over fluid cells.  The null space is fixed
      after the solve by subtracting the value of the pinned cell (the leaf cell that
      contains PIN_POINT).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

TREE_FIELDS = ("level", "parent", "which_child", "child", "neighbor", "node_type", "node_type2",
               "lo", "hi", "center", "length", "surr", "surr_type", "surr_bcref")

# Physical-boundary neighbour code used for every wall (any value <= -20 is a wall,
# SayakaTree.h:38-40); Test7.h:42 uses the same value.
WALL = -100
# Math BC types (SayakaBoundaryPatch.h:53-58)
NEUMANN, DIRICHLET, SIMPLEFILL = 0, 1, 2

SHELL_FACTOR = 0.5
VF_MIN = 0.02
PIN_POINT = (0.01, 0.01, 0.01)


@dataclasses.dataclass(frozen=True)
class TreeSpec:
    ndim: int
    nb: int
    minlevel: int
    maxlevel: int
    ndiv: tuple
    lo: tuple
    hi: tuple
    ngrow: int = 2
    bc: tuple = (WALL,) * 6


C0_SPEC = TreeSpec(ndim=2, nb=8, minlevel=2, maxlevel=3, ndiv=(4, 4, 1),
                   lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 0.0))
C1_SPEC = TreeSpec(ndim=3, nb=8, minlevel=2, maxlevel=4, ndiv=(4, 4, 4),
                   lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0))
# Large snapshot: 16^3-cell blocks, root 4^3, tree levels 1-6 refined around an r = 0.15
# sphere with the C1 shell rule (38 848 blocks, 34 000 leaves, 139 M leaf cells).
C4_SPEC = TreeSpec(ndim=3, nb=16, minlevel=2, maxlevel=6, ndiv=(4, 4, 4),
                   lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0))
C4_RADIUS = 0.15


def small_spec(ndim: int, maxlevel: int = 3, nb: int = 8) -> TreeSpec:
    """Small trees for tests (root 2^d, minAmrLevel 2)."""
    if ndim == 2:
        return TreeSpec(ndim=2, nb=nb, minlevel=2, maxlevel=maxlevel, ndiv=(2, 2, 1),
                        lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 0.0))
    return TreeSpec(ndim=3, nb=nb, minlevel=2, maxlevel=maxlevel, ndiv=(2, 2, 2),
                    lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0))


# ----------------------------------------------------------------------------- geometry
def cell_centres(tree: dict, nb: int, ndim: int, blocks=None):
    """Cell-centre coordinates (x, y, z) as arrays [nblk, nz, ny, nx].

    Same expression as the reference demos: x = lo + (i + 0.5) * dx with dx = length / nb
    (Test6.h:690-693, SayakaTree.h:277-284)."""
    lo = tree["lo"] if blocks is None else tree["lo"][blocks]
    ln = tree["length"] if blocks is None else tree["length"][blocks]
    nz = nb if ndim == 3 else 1
    i = np.arange(nb, dtype=np.float64) + 0.5
    k = np.arange(nz, dtype=np.float64) + 0.5
    dx = ln / nb
    x = lo[:, 0, None, None, None] + i[None, None, None, :] * dx[:, 0, None, None, None]
    y = lo[:, 1, None, None, None] + i[None, None, :, None] * dx[:, 1, None, None, None]
    if ndim == 3:
        z = lo[:, 2, None, None, None] + k[None, :, None, None] * dx[:, 2, None, None, None]
    else:
        z = np.zeros((lo.shape[0], 1, 1, 1))
    shape = (lo.shape[0], nz, nb, nb)
    return np.broadcast_to(x, shape), np.broadcast_to(y, shape), np.broadcast_to(z, shape)


def face_centres(tree: dict, nb: int, ndim: int, d: int, blocks=None):
    """Face-centre coordinates of the faces normal to ``d``: arrays over the staggered box
    [nblk, nz', ny', nx'] (i fastest, IndexBox::offset order)."""
    lo = tree["lo"] if blocks is None else tree["lo"][blocks]
    ln = tree["length"] if blocks is None else tree["length"][blocks]
    dx = ln / nb
    n = [nb, nb, nb if ndim == 3 else 1]
    n[d] += 1
    idx = []
    for a in range(3):
        if a >= ndim:
            idx.append(np.zeros(1))
        elif a == d:
            idx.append(np.arange(n[a], dtype=np.float64))
        else:
            idx.append(np.arange(n[a], dtype=np.float64) + 0.5)
    x = lo[:, 0, None, None, None] + idx[0][None, None, None, :] * dx[:, 0, None, None, None]
    y = lo[:, 1, None, None, None] + idx[1][None, None, :, None] * dx[:, 1, None, None, None]
    if ndim == 3:
        z = lo[:, 2, None, None, None] + idx[2][None, :, None, None] * dx[:, 2, None, None, None]
    else:
        z = np.zeros((lo.shape[0], 1, 1, 1))
    shape = (lo.shape[0], n[2], n[1], n[0])
    return np.broadcast_to(x, shape), np.broadcast_to(y, shape), np.broadcast_to(z, shape)


def face_shape(nb: int, ndim: int, d: int):
    n = [nb, nb, nb if ndim == 3 else 1]
    n[d] += 1
    return (n[2], n[1], n[0])


# ------------------------------------------------------------------------- tree builders
def ring_criterion(centre=(0.5, 0.5), r0=0.25, half=0.05):
    def crit(tree, nb, ndim, blocks):
        x, y, _ = cell_centres(tree, nb, ndim, blocks)
        r = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2)
        return np.any(np.abs(r - r0) < half, axis=(1, 2, 3))
    return crit


def shell_criterion(centre=(0.5, 0.5, 0.5), r0=0.2, factor=SHELL_FACTOR):
    def crit(tree, nb, ndim, blocks):
        x, y, z = cell_centres(tree, nb, ndim, blocks)
        r = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2)
        half = factor * tree["length"][blocks, 0]
        return np.any(np.abs(r - r0) < half[:, None, None, None], axis=(1, 2, 3))
    return crit


def build_tree(backend, spec: TreeSpec, criterion, extra_passes: int = 1):
    """Create the root level, refine to minAmrLevel, then refine level by level."""
    backend.create(spec)
    for level in range(spec.minlevel, spec.maxlevel + extra_passes):
        t = backend.export()
        n = t["level"].shape[0]
        cand = np.nonzero((t["level"] == level) & ((t["node_type"] == 1) | (t["node_type"] == 2)))[0]
        flags = np.zeros(n, dtype=np.int32)
        if level < spec.maxlevel and cand.size:
            hit = criterion(t, spec.nb, spec.ndim, cand)
            flags[cand[hit]] = 1
        backend.set_refine_flags(flags)
        backend.regrid()
    return backend.export()


def leaf_blocks(tree: dict) -> np.ndarray:
    return np.nonzero(tree["node_type"] == 1)[0]


# ---------------------------------------------------------------------------- C0 inputs
def c0_rhs_exact(tree: dict, spec: TreeSpec):
    x, y, _ = cell_centres(tree, spec.nb, spec.ndim)
    p = np.sin(np.pi * x) * np.sin(np.pi * y)
    rhs = 2.0 * np.pi ** 2 * p
    n = tree["level"].shape[0]
    return rhs.reshape(n, -1).copy(), p.reshape(n, -1).copy()


# ---------------------------------------------------------------------------- C1 inputs
def smoothed_heaviside(s, eps):
    h = 0.5 * (1.0 + s / eps + np.sin(np.pi * s / eps) / np.pi)
    return np.where(s < -eps, 0.0, np.where(s > eps, 1.0, h))


def _sample_offsets(nsub):
    return (np.arange(nsub) + 0.5) / nsub


def sphere_fractions(tree: dict, spec: TreeSpec, centre, radius, nsub=4):
    """Fluid fractions outside a solid sphere: vf [nblk, ncell], a_f[d] [nblk, nface_d].

    Sub-sampled (nsub^d points per cell, nsub^(d-1) per face).  Coordinates are exact
    dyadic rationals on these trees, so neighbouring blocks compute identical values for
    their shared faces."""
    nb, ndim = spec.nb, spec.ndim
    n = tree["level"].shape[0]
    lo, ln = tree["lo"], tree["length"]
    dx = ln / nb
    off = _sample_offsets(nsub)
    nz = nb if ndim == 3 else 1
    # volume fractions
    vf = np.zeros((n, nz, nb, nb))
    cnt = 0
    kk = off if ndim == 3 else np.array([0.0])
    for sz in kk:
        for sy in off:
            for sx in off:
                i = np.arange(nb) + sx
                j = np.arange(nb) + sy
                k = (np.arange(nz) + sz) if ndim == 3 else np.zeros(1)
                x = lo[:, 0, None, None, None] + i[None, None, None, :] * dx[:, 0, None, None, None]
                y = lo[:, 1, None, None, None] + j[None, None, :, None] * dx[:, 1, None, None, None]
                z = (lo[:, 2, None, None, None] + k[None, :, None, None] * dx[:, 2, None, None, None]) if ndim == 3 else 0.0
                r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + ((z - centre[2]) ** 2 if ndim == 3 else 0.0)
                vf += (r2 > radius * radius)
                cnt += 1
    vf /= cnt
    solid = vf < VF_MIN
    vf[solid] = 0.0
    af = []
    for d in range(ndim):
        fs = face_shape(nb, ndim, d)
        a = np.zeros((n,) + fs)
        cnt = 0
        tang = [a_ for a_ in range(ndim) if a_ != d]
        grids = [off] * len(tang)
        for combo in np.array(np.meshgrid(*grids, indexing="ij")).reshape(len(tang), -1).T:
            idx = []
            for ax in range(3):
                m = fs[2 - ax]
                if ax >= ndim:
                    idx.append(np.zeros(1))
                elif ax == d:
                    idx.append(np.arange(m, dtype=np.float64))
                else:
                    idx.append(np.arange(m) + combo[tang.index(ax)])
            x = lo[:, 0, None, None, None] + idx[0][None, None, None, :] * dx[:, 0, None, None, None]
            y = lo[:, 1, None, None, None] + idx[1][None, None, :, None] * dx[:, 1, None, None, None]
            z = (lo[:, 2, None, None, None] + idx[2][None, :, None, None] * dx[:, 2, None, None, None]) if ndim == 3 else 0.0
            r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + ((z - centre[2]) ** 2 if ndim == 3 else 0.0)
            a += (r2 > radius * radius)
            cnt += 1
        a /= cnt
        # close every face that touches a solid cell (inside the block) ...
        sl = [slice(None)] * 4
        s_lo = solid.copy()
        pad = [(0, 0)] * 4
        pad[3 - d] = (1, 0)
        left = np.pad(s_lo, pad, constant_values=False)      # cell on the low side of face
        pad[3 - d] = (0, 1)
        right = np.pad(s_lo, pad, constant_values=False)     # cell on the high side
        a[left | right] = 0.0
        af.append(a)
    # ... and faces whose outside neighbour cell is solid (block boundary faces): the
    # neighbour cell of the boundary face has the same size on these trees where the body
    # is cut (the body is covered by the finest level), so evaluate it directly.
    for d in range(ndim):
        for side in (0, 1):
            ghost_vf = _ghost_layer_vf(tree, spec, centre, radius, nsub, d, side)
            sel = [slice(None)] * 4
            sel[3 - d] = 0 if side == 0 else af[d].shape[3 - d] - 1
            face = af[d][tuple(sel)]
            face[ghost_vf < VF_MIN] = 0.0
            af[d][tuple(sel)] = face
    return vf.reshape(n, -1), [a.reshape(n, -1) for a in af]


def _ghost_layer_vf(tree, spec, centre, radius, nsub, d, side):
    nb, ndim = spec.nb, spec.ndim
    lo, ln = tree["lo"], tree["length"]
    dx = ln / nb
    off = _sample_offsets(nsub)
    n = lo.shape[0]
    dims = [nb, nb, nb if ndim == 3 else 1]
    dims[d] = 1
    acc = np.zeros((n, dims[2], dims[1], dims[0]))
    cnt = 0
    kk = off if ndim == 3 else np.array([0.0])
    for sz in kk:
        for sy in off:
            for sx in off:
                sub = [sx, sy, sz]
                idx = []
                for ax in range(3):
                    if ax >= ndim:
                        idx.append(np.zeros(1))
                    elif ax == d:
                        idx.append(np.array([-1.0 + sub[ax]]) if side == 0 else np.array([nb + sub[ax]]))
                    else:
                        idx.append(np.arange(nb) + sub[ax])
                x = lo[:, 0, None, None, None] + idx[0][None, None, None, :] * dx[:, 0, None, None, None]
                y = lo[:, 1, None, None, None] + idx[1][None, None, :, None] * dx[:, 1, None, None, None]
                z = (lo[:, 2, None, None, None] + idx[2][None, :, None, None] * dx[:, 2, None, None, None]) if ndim == 3 else 0.0
                r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + ((z - centre[2]) ** 2 if ndim == 3 else 0.0)
                acc += (r2 > radius * radius)
                cnt += 1
    sq = [slice(None)] * 4
    sq[3 - d] = 0
    return (acc / cnt)[tuple(sq)]


def density_coef(tree: dict, spec: TreeSpec, centre, radius, ratio, eps):
    """B_f = 1/rho(face centre), rho = 1 + (ratio-1) H_eps(-phi), phi = |x-c| - r."""
    out = []
    n = tree["level"].shape[0]
    for d in range(spec.ndim):
        x, y, z = face_centres(tree, spec.nb, spec.ndim, d)
        r = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + ((z - centre[2]) ** 2 if spec.ndim == 3 else 0.0))
        rho = 1.0 + (ratio - 1.0) * smoothed_heaviside(-(r - radius), eps)
        out.append((1.0 / rho).reshape(n, -1).copy())
    return out


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def face_velocity(tree: dict, spec: TreeSpec, seed: int):
    """u_f ~ U(-1,1) keyed on (seed, level, global face index, dir): a face shared by two
    blocks of one level gets one value, independent of block numbering."""
    nb, ndim = spec.nb, spec.ndim
    n = tree["level"].shape[0]
    out = []
    root_len = (np.asarray(spec.hi) - np.asarray(spec.lo)) / np.maximum(np.asarray(spec.ndiv), 1)
    for d in range(ndim):
        fs = face_shape(nb, ndim, d)
        lvl = tree["level"].astype(np.int64)
        scale = (2 ** (lvl - 1)) * nb                             # cells per root block edge
        org = []
        for a in range(3):
            if a < ndim:
                org.append(np.rint((tree["lo"][:, a] - spec.lo[a]) / (root_len[a] / scale)).astype(np.int64))
            else:
                org.append(np.zeros(n, dtype=np.int64))
        ii = org[0][:, None, None, None] + np.arange(fs[2])[None, None, None, :]
        jj = org[1][:, None, None, None] + np.arange(fs[1])[None, None, :, None]
        kk = org[2][:, None, None, None] + np.arange(fs[0])[None, :, None, None]
        key = (np.uint64(seed) * np.uint64(0x100000001B3)
               ^ (lvl[:, None, None, None].astype(np.uint64) << np.uint64(56))
               ^ (np.uint64(d) << np.uint64(61))
               ^ (kk.astype(np.uint64) << np.uint64(36))
               ^ (jj.astype(np.uint64) << np.uint64(18))
               ^ ii.astype(np.uint64))
        h = _splitmix64(key)
        u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        out.append((2.0 * u - 1.0).reshape(n, -1))
    return out


def divergence_rhs(tree: dict, spec: TreeSpec, u, vf, af):
    """rhs = sum_d (a u ds)|_hi - (a u ds)|_lo / (dv vf) on fluid cells, then minus its
    volume-weighted mean over fluid leaf cells."""
    nb, ndim = spec.nb, spec.ndim
    n = tree["level"].shape[0]
    dx = tree["length"][:, :ndim] / nb
    dv = np.prod(dx, axis=1)
    nz = nb if ndim == 3 else 1
    div = np.zeros((n, nz, nb, nb))
    for d in range(ndim):
        fs = face_shape(nb, ndim, d)
        flux = (af[d] * u[d]).reshape((n,) + fs) * (dv / dx[:, d])[:, None, None, None]
        hi = [slice(None)] * 4
        lo_ = [slice(None)] * 4
        hi[3 - d] = slice(1, None)
        lo_[3 - d] = slice(0, -1)
        div += flux[tuple(hi)] - flux[tuple(lo_)]
    vfr = vf.reshape(n, nz, nb, nb)
    fluid = vfr > 0
    rhs = np.where(fluid, div / (dv[:, None, None, None] * np.where(fluid, vfr, 1.0)), 0.0)
    leaves = tree["node_type"] == 1
    w = (dv[:, None, None, None] * vfr)[leaves]
    mean = np.sum(rhs[leaves] * w) / np.sum(w)
    rhs = np.where(fluid, rhs - mean, 0.0)
    return rhs.reshape(n, -1)


def c1_inputs(tree: dict, spec: TreeSpec, seed: int = 0, centre=(0.5, 0.5, 0.5), radius=0.2,
              ratio=1000.0):
    dx_fine = (spec.hi[0] - spec.lo[0]) / spec.ndiv[0] / spec.nb / 2 ** (spec.maxlevel - 1)
    eps = 1.5 * dx_fine
    vf, af = sphere_fractions(tree, spec, centre, radius)
    b = density_coef(tree, spec, centre, radius, ratio, eps)
    return dict(vf=vf, af=af, b=b, rhs=c1_rhs(tree, spec, seed, vf, af))


def c1_rhs(tree: dict, spec: TreeSpec, seed: int, vf, af):
    """The seeded right-hand side alone (the geometry terms of c1_inputs are seed-free)."""
    return divergence_rhs(tree, spec, face_velocity(tree, spec, seed), vf, af)


def fluid_rel_l2(a, b, tree: dict, spec: TreeSpec, vf):
    """Relative L2 difference of two solutions of the pure-Neumann problem over the fluid leaf
    cells (vf > 0), each less its volume-weighted mean there: the null space removed, and the
    solid cells -- outside the system, 0 in both -- left out (with a gauge applied to every
    cell they would carry the two solutions' different constants)."""
    fluid = (tree["node_type"] == 1)[:, None] & (vf > 0)
    dv = np.broadcast_to(((tree["length"][:, 0] / spec.nb) ** spec.ndim)[:, None], vf.shape)
    w = (dv * vf)[fluid]
    x = a[fluid] - np.sum(w * a[fluid]) / np.sum(w)
    y = b[fluid] - np.sum(w * b[fluid]) / np.sum(w)
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def pin_cell(tree: dict, spec: TreeSpec, point=PIN_POINT):
    """(block, cell) of the leaf cell containing ``point``."""
    lo, hi = tree["lo"], tree["hi"]
    leaves = leaf_blocks(tree)
    p = np.asarray(point[:spec.ndim])
    inside = np.all((lo[leaves, :spec.ndim] <= p) & (p < hi[leaves, :spec.ndim]), axis=1)
    b = int(leaves[np.nonzero(inside)[0][0]])
    dx = tree["length"][b, :spec.ndim] / spec.nb
    ijk = np.floor((p - lo[b, :spec.ndim]) / dx).astype(int)
    cell = int(ijk[0] + spec.nb * ijk[1] + (spec.nb * spec.nb * ijk[2] if spec.ndim == 3 else 0))
    return b, cell
