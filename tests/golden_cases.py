"""Rebuild the inputs of the golden cases (tests/golden/*.npz) from their recorded arrays."""
from __future__ import annotations

import numpy as np

from sayakaamr_amd import workloads as W

TREE_KEYS = ("level", "parent", "which_child", "child", "neighbor", "node_type", "node_type2", "lo", "hi",
             "center", "length", "surr", "surr_type", "surr_bcref")

# name -> (spec, math BC on every wall)
CASES = {
    "c0_ops": (W.C0_SPEC, W.DIRICHLET),
    "s3_ops": (W.small_spec(3, 3, nb=4), W.NEUMANN),
}


def tree_of(g, prefix="tree_"):
    return {k: g[prefix + k] for k in TREE_KEYS if prefix + k in g}


def inputs_of(g, ndim):
    inp = dict(rhs=g["in_rhs"], vf=g.get("in_vf"), af=None, b=None)
    if "in_af0" in g:
        inp["af"] = [g[f"in_af{d}"] for d in range(ndim)]
    if "in_b0" in g:
        inp["b"] = [g[f"in_b{d}"] for d in range(ndim)]
    return inp


def tower_blocks(tree, m):
    lv, leaf = tree["level"], tree["node_type"] == 1
    return np.nonzero(((lv < m) & leaf) | (lv == m))[0]


# ------------------------------------------------------------ general state fill cases
# FillPatch::fillBoundary on a multi-component cell state with two ghost layers.  Every wall
# has its own physical code so that the BC register, the sweep order and the edge / corner
# references (bcref) all matter; each component has its own prolongation method and BC mix.
FILL_WALLS = (-100, -101, -102, -103, -104, -105)
FILL_NGROW = 2
FILL_CASES = {
    # name -> (ndim, nb, maxlevel, criterion args, components, nlayer values)
    "fill2": (2, 8, 4, ("ring", (0.5, 0.5), 0.2, 0.04), 3, (2, 1)),
    "fill3": (3, 4, 3, ("shell", (0.4, 0.5, 0.5), 0.18), 2, (2,)),
}
# per component: prolongation (0 injection, 1 centred linear, 2 limited) and the math BC /
# value of each wall code
FILL_COMPS = {
    "fill2": dict(prolong=(0, 1, 2),
                  mathbc=((0,) * 6, (1, 1, 2, 1, 0, 1), (1, 2, 0, 1, 1, 2)),
                  bcval=((0.0,) * 6, (0.5, -1.5, 2.0, 0.25, 0.0, 3.0), (1.0, -2.0, 0.0, 0.75, -0.5, 1.5))),
    "fill3": dict(prolong=(1, 2),
                  mathbc=((1, 0, 2, 1, 1, 0), (0, 1, 1, 2, 0, 1)),
                  bcval=((0.5, 0.0, -1.0, 2.0, 0.25, 0.0), (0.0, 1.5, -0.5, 3.0, 0.0, -2.0))),
}


def fill_spec(name):
    import dataclasses
    nd, nb, maxlevel, crit, _, _ = FILL_CASES[name]
    return dataclasses.replace(W.small_spec(nd, maxlevel, nb=nb), bc=FILL_WALLS, ngrow=FILL_NGROW)


def fill_criterion(name):
    kind, centre, r0, *rest = FILL_CASES[name][3]
    if kind == "ring":
        return W.ring_criterion(centre=centre, r0=r0, half=rest[0])
    return W.shell_criterion(centre=centre, r0=r0)


def fill_input(name, nblk):
    """[comp][block][gz][gy][gx]: valid cells U(-1,1) from a counter-based hash of
    (component, block, cell), ghost cells a distinct sentinel each (so a ghost the fill
    leaves alone is checked as well)."""
    nd, nb, _, _, ncomp, _ = FILL_CASES[name]
    g = FILL_NGROW
    nz, gz = (nb, g) if nd == 3 else (1, 0)
    shp = (ncomp, nblk, nz + 2 * gz, nb + 2 * g, nb + 2 * g)
    idx = np.arange(np.prod(shp), dtype=np.uint64).reshape(shp)
    h = W._splitmix64(idx + np.uint64(0x5EED0000))
    u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0
    out = -1000.0 - idx.astype(np.float64) * 1e-3
    sl = (slice(None), slice(None), slice(gz, gz + nz), slice(g, g + nb), slice(g, g + nb))
    out[sl] = u[sl]
    return out


def fill_bcmaps(name):
    comps = FILL_COMPS[name]
    return [{code: (comps["mathbc"][c][q], comps["bcval"][c][q]) for q, code in enumerate(FILL_WALLS)}
            for c in range(len(comps["prolong"]))]


def ghost_mask(nd, nb, g):
    nz, gz = (nb, g) if nd == 3 else (1, 0)
    m = np.ones((nz + 2 * gz, nb + 2 * g, nb + 2 * g), bool)
    m[gz:gz + nz, g:g + nb, g:g + nb] = False
    return m
