"""The projection step on the GPU (sayaka_gen_velocity, sayaka_gen_divergence_rhs,
sayaka_dev_project): the projection of the reference's flow demo (Test7.h:580-657) with dt = 1 and
the pressure's sign flipped -- rhs = div u, solve, u -= B flux, coarse/fine faces synchronised.

  * the seeded face velocity is bit-identical to workloads.face_velocity (numpy) in 2D and 3D;
  * its coarse/fine synchronisation is bit-identical to the solver's coefficient synchronisation
    of the same values (sayaka_dev_set_coefs with B = u, itself checked against the oracle);
  * the rhs equals workloads.divergence_rhs of the synchronised velocity (numpy, 1e-12: the
    mean is summed in another order);
  * after the solve and the correction, the divergence left on the fluid cells of the leaves
    equals the solve's final residual, by the device's own measure and by numpy's divergence of
    the downloaded velocity, for unit coefficients and for a cut-cell sphere with density ratio
    1000.
A wrong sign, a missing B or a wrong coarse/fine flux leaves an O(1) divergence."""
from __future__ import annotations

import numpy as np
import pytest

from sayakaamr_amd import device as D
from sayakaamr_amd import workloads as W
from sayakaamr_amd.tree import AmrTree

pytestmark = pytest.mark.gpu


def _tree(ndim):
    if ndim == 3:
        spec = W.C1_SPEC
        tb = AmrTree(3, max_blocks=6000)
        t = W.build_tree(tb, spec, W.shell_criterion())
    else:
        spec = W.C0_SPEC
        tb = AmrTree(2, max_blocks=6000)
        t = W.build_tree(tb, spec, W.ring_criterion())
    return spec, t


@pytest.mark.parametrize("ndim", [3, 2])
def test_velocity_and_rhs_match_numpy(ndim):
    spec, t = _tree(ndim)
    n = t["level"].shape[0]
    print(f"[gpu] projection velocity {ndim}D: {n} blocks", flush=True)
    dev = D.MacSolverGPU(ndim, spec.nb, 6000, device=0)
    try:
        dev.prepare(t, spec)
        dev.set_bc([W.WALL], [W.NEUMANN])
        dev.gen_velocity(t, spec, seed=11, zero_walls=False)
        raw = dev.get_velocity()
        ref = W.face_velocity(t, spec, 11)
        for d in range(ndim):
            assert np.array_equal(raw[d], ref[d]), d
        dev.gen_divergence_rhs()
        synced = dev.get_velocity()
        rhs = dev.download(D.RHS)
        # the same synchronisation as the solver's B coefficient
        dev.set_coefs(b=raw)
        _, b_sync, _, _, _ = dev.get_coefs()
        for d in range(ndim):
            leaf = t["node_type"] == 1
            assert np.array_equal(synced[d][leaf], b_sync[d][leaf]), d
        vf, _, af, _, _ = dev.get_coefs()
        want = W.divergence_rhs(t, spec, synced, vf, af)
        leaf = t["node_type"] == 1
        scale = np.abs(want[leaf]).max()
        assert np.abs(rhs[leaf] - want[leaf]).max() <= 1e-12 * scale
    finally:
        dev.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["unit3", "unit2", "sphere3"])
def test_projection_removes_the_divergence(case, precision):
    """The divergence left after the correction is the solve's residual: u' = u - B flux(p) has
    div u' = rhs - A p on every leaf fluid cell, so max |div u'| equals the final residual norm
    the solver reports (to rounding).  The reference-order FP64 solve is run to 1e-9 (on the
    all-Neumann unit-coefficient problems its BiCGSTAB wanders between 1e-9 and 1e-10 and does not
    reach 1e-10 in 400 iterations, in 2D and 3D), the mixed solve to 1e-10."""
    ndim = 2 if case == "unit2" else 3
    spec, t = _tree(ndim)
    L = int(t["level"].max())
    print(f"[gpu] projection {case} {precision}", flush=True)
    dev = D.MacSolverGPU(ndim, spec.nb, 6000, device=0)
    try:
        dev.prepare(t, spec)
        dev.set_bc([W.WALL], [W.NEUMANN])
        if case == "sphere3":
            dxf = (spec.hi[0] - spec.lo[0]) / spec.ndiv[0] / spec.nb / 2 ** (spec.maxlevel - 1)
            dev.gen_sphere(t, (0.5, 0.5, 0.5), 0.2, 1000.0, 1.5 * dxf, 4)
        dev.gen_velocity(t, spec, seed=5, zero_walls=True)
        dev.gen_divergence_rhs()
        bmax = dev.dot(L, D.RHS, D.RHS)[1]
        dev.zero(D.SOL)
        tol = 1e-9 if precision == "fp64" else 1e-10
        r = dev.solve(tol_rel=0.0, tol_abs=tol * bmax, max_iter=400,
                      precision=D.FP64 if precision == "fp64" else D.MIXED)
        assert r["status"] == 0
        before, after = dev.project(D.SOL)
        u = dev.get_velocity()
        vf, _, af, _, _ = dev.get_coefs()
    finally:
        dev.close()
    leaf = t["node_type"] == 1
    div_np = np.abs(W.divergence_rhs(t, spec, u, vf, af)[leaf]).max()
    print(f"[gpu] projection {case} {precision}: |div| {before:.3e} -> {after:.3e} (numpy {div_np:.3e}), "
          f"residual {r['rnorm']:.3e}, {r['iters']} iterations", flush=True)
    assert abs(before - bmax) <= 1e-12 * bmax
    assert after <= tol * before
    assert abs(after - r["rnorm"]) <= 1e-3 * r["rnorm"] + 1e-13 * before
    assert abs(div_np - after) <= 1e-3 * after + 1e-13 * before
