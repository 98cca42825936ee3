"""CPU restatement of the regrid data movement (test infrastructure only).

After AmrTree::regrid the reference's registered cell states (TreeStateData) hold, for every
block that existed before the regrid, that block's data at its new position (the survivors
are packed to the front of the pool in order, SayakaRegrid.cpp:530-591), and
AmrTree::initNewBlockDataAfterRegrid (SayakaRegrid.cpp:987-1105) fills the blocks the regrid
created, level by level from the coarsest, from their parents with the state's FillPatch --
by default PROLONG_INJECTION (SayakaInterpPatch.h:78-82, 109): fine cell (i, j, k) takes the
parent cell that contains it (prolongCellBlockData_Injection, SayakaInterpPatch.cpp:155-187).

Pinned against the reference's own output (tests/golden/regrid_remap{2,3}.npz, written from
the compiled reference by tests/golden/make_golden.py): tests/test_remap.py.
"""
from __future__ import annotations

import numpy as np


def inject(parent_cells, which_child, nb, ndim):
    """Fine values of child `which_child` from the parent's [cell] array by injection."""
    nz = nb if ndim == 3 else 1
    p = parent_cells.reshape(nz, nb, nb)
    h = nb // 2
    ox, oy = (which_child & 1) * h, ((which_child >> 1) & 1) * h
    oz = ((which_child >> 2) & 1) * h if ndim == 3 else 0
    k, j, i = np.meshgrid(np.arange(nz), np.arange(nb), np.arange(nb), indexing="ij")
    kc = k // 2 + oz if ndim == 3 else k
    return p[kc, j // 2 + oy, i // 2 + ox].reshape(-1)


def remap(pre, origin, tree, nb, ndim):
    """State after regrid + initNewBlockDataAfterRegrid.  pre: [old block][cell]; origin: per new
    block its old id or -1; tree: the exported tree after the regrid."""
    n = origin.shape[0]
    post = np.zeros((n, pre.shape[1]))
    keep = origin >= 0
    post[keep] = pre[origin[keep]]
    new = np.nonzero(~keep)[0]
    for lev in sorted(set(tree["level"][new].tolist())):
        for b in new[tree["level"][new] == lev]:
            post[b] = inject(post[tree["parent"][b]], int(tree["which_child"][b]), nb, ndim)
    return post
