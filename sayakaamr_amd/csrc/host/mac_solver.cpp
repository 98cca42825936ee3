// mac_solver.cpp -- host MacSolver: the reference's call surface over the device C ABI.
#include "mac_solver.h"

#include <algorithm>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "dev_tree.h"
#include "sayaka_amd.h"

namespace sayaka {
inline namespace amd {

MacSolver::MacSolver(const AmrTree& tree, int device) : tree_(&tree) {
    ncell_ = tree.cellsPerBlock();
    nface_ = (tree.nb + 1) * tree.nb * (tree.ndim == 3 ? tree.nb : 1);
    // the tree's context: one per tree, shared with its FillPatches (the solver's device fields
    // live there, so one MacSolver per tree solves at a time)
    ctx_ = tree.deviceContext(device);
    towers_.assign(tree.maxAmrLevel + 1, MGLevelTower{});
    for (int m = 0; m <= tree.maxAmrLevel; m++) {
        towers_[m].level = m;
        towers_[m].level_begin.assign(tree.maxAmrLevel + 1, 0);
        towers_[m].level_end.assign(tree.maxAmrLevel + 1, 0);
    }
}

// MGLevelTower::update (SayakaTower.cpp:28-68) for every multigrid level
static void build_towers(const AmrTree& t, std::vector<MGLevelTower>& towers) {
    for (int m = 1; m <= t.maxAmrLevel; m++) {
        MGLevelTower& tw = towers[m];
        tw.clear();
        std::fill(tw.level_begin.begin(), tw.level_begin.end(), 0);
        std::fill(tw.level_end.begin(), tw.level_end.end(), 0);
        for (int l = m; l >= 1; l--) {
            tw.level_begin[l] = tw.numLevelBlock;
            for (int b : t.getTreeLevel(l).blockIds())
                if (l == m || t[b].isLeaf()) tw.addBlock(b);
            tw.level_end[l] = tw.numLevelBlock;
        }
    }
}

MacSolver::~MacSolver() = default;   // the context belongs to the tree

void MacSolver::check(int rc, const char* what) const {
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + sayaka_dev_last_error());
}

void upload_tree_topology(sayaka_dev_ctx* ctx, const AmrTree& t) {
    const int n = t.numBlocks;
    std::vector<int> level(n), parent(n), wc(n), child(8 * n), nbr(6 * n), type(n);
    std::vector<double> len(3 * n);
    for (int b = 0; b < n; b++) {
        const AmrTreeNode& nd = t[b];
        level[b] = nd.refineLevel;
        parent[b] = nd.parent;
        wc[b] = nd.whichChild;
        type[b] = nd.nodeType;
        for (int q = 0; q < 8; q++) child[8 * b + q] = nd.child[q];
        for (int f = 0; f < 6; f++) nbr[6 * b + f] = nd.neighbor[f];
        for (int d = 0; d < 3; d++) len[3 * b + d] = nd.blockLength(d);
    }
    if (sayaka_dev_set_topology(ctx, n, t.minAmrLevel, t.maxAmrLevel, level.data(), parent.data(), wc.data(),
                                child.data(), nbr.data(), type.data(), len.data()) < 0)
        throw std::runtime_error(std::string("tree topology: ") + sayaka_dev_last_error());
}

void upload_tree_surroundings(sayaka_dev_ctx* ctx, const AmrTree& t) {
    const int n = t.numBlocks;
    std::vector<int> surr((size_t)n * 27), bcref((size_t)n * 27);
    for (int b = 0; b < n; b++) {
        const SurroundingBlocks& s = t.surrBlocks[b];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    surr[(size_t)b * 27 + i * 9 + j * 3 + k] = s.blocks[i][j][k];
                    bcref[(size_t)b * 27 + i * 9 + j * 3 + k] = s.bcrefs[i][j][k];
                }
    }
    if (sayaka_dev_set_surroundings(ctx, n, surr.data(), bcref.data()) < 0)
        throw std::runtime_error(std::string("tree surroundings: ") + sayaka_dev_last_error());
}

void MacSolver::prepareSolver() {
    const AmrTree& t = *tree_;
    nblk_ = t.numBlocks;
    // always: the upload also resets the device coefficients and BC, as prepareSolver does
    upload_tree_topology(ctx_, t);
    t.markTopologyUploaded();
    build_towers(t, towers_);
    // prepareSolver resets the coefficients to unit values (SayakaMacSolver.cpp:12-26)
    unitIB_ = unitB_ = true;
    alpha_ = 0.0;
    beta_ = 1.0;
    vf_.clear();
    for (int d = 0; d < 3; d++) {
        af_[d].clear();
        b_[d].clear();
    }
}

void MacSolver::uploadCoefs() {
    const double* af[3] = {nullptr, nullptr, nullptr};
    const double* bb[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < tree_->ndim; d++) {
        if (!unitIB_) af[d] = af_[d].data();
        if (!unitB_) bb[d] = b_[d].data();
    }
    check(sayaka_dev_set_coefs(ctx_, alpha_, beta_, unitIB_ ? nullptr : vf_.data(), unitIB_ ? nullptr : af,
                               unitB_ ? nullptr : bb),
          "MacSolver: coefficients");
}

void MacSolver::setUnitIBFrac() {
    unitIB_ = true;
    uploadCoefs();
}

// setIBFracByLS -> calcBlockIBVolumeFrac / calcBlockIBAreaFrac (SayakaMacIB.cpp:34-154):
// on the blocks of the finest tower a cell is solid where ls <= 0 and a face is closed when
// either adjacent cell is solid; the device then synchronises coarse/fine faces and averages
// the fractions down to the parent blocks.
void MacSolver::setIBFracByLS(const TreeData& ls, int lscomp) {
    if (ls.numGrow() < 1 || ls.faceDir() >= 0) throw std::invalid_argument("setIBFracByLS: cell data with one ghost layer");
    std::vector<double> g((std::size_t)nblk_ * ls.extent(0) * ls.extent(1) * ls.extent(2));
    ls.packGrown(lscomp, nblk_, g.data());
    setIBFracByLSPacked(g.data(), ls.numGrow());
}

void MacSolver::setIBFracByLSPacked(const double* lsg, int ng) {
    const AmrTree& t = *tree_;
    if (ng < 1) throw std::invalid_argument("setIBFracByLS: the level set needs one ghost layer");
    const int nb = t.nb, nbz = t.ndim == 3 ? nb : 1, finest = t.currentFinestLevel();
    const int kg = t.ndim == 3 ? ng : 0, lx = nb + 2 * ng, ly = nb + 2 * ng, lz = nbz + 2 * kg;
    auto ls = [&](int b, int i, int j, int k) {
        return lsg[(std::size_t)b * lx * ly * lz + (i + ng) + lx * ((j + ng) + ly * (k + kg))];
    };
    vf_.assign((std::size_t)nblk_ * ncell_, 1.0);
    for (int d = 0; d < t.ndim; d++) af_[d].assign((std::size_t)nblk_ * nface_, 1.0);
    for (int b = 0; b < nblk_; b++) {
        if (!(t[b].isLeaf() || t[b].refineLevel == finest)) continue;
        for (int k = 0; k < nbz; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++)
                    vf_[(std::size_t)b * ncell_ + i + nb * (j + nb * k)] = ls(b, i, j, k) <= 0 ? 0.0 : 1.0;
        for (int d = 0; d < t.ndim; d++) {
            const int fx = nb + (d == 0), fy = nb + (d == 1), fz = t.ndim == 3 ? nb + (d == 2) : 1;
            const int di = d == 0, dj = d == 1, dk = d == 2;
            for (int k = 0; k < fz; k++)
                for (int j = 0; j < fy; j++)
                    for (int i = 0; i < fx; i++)
                        af_[d][(std::size_t)b * nface_ + i + fx * (j + fy * k)] =
                            (ls(b, i - di, j - dj, k - dk) <= 0 || ls(b, i, j, k) <= 0) ? 0.0 : 1.0;
        }
    }
    unitIB_ = false;
    uploadCoefs();
}

void MacSolver::setIBFracPacked(const double* vf, const double* const* af) {
    vf_.assign(vf, vf + (std::size_t)nblk_ * ncell_);
    for (int d = 0; d < tree_->ndim; d++) af_[d].assign(af[d], af[d] + (std::size_t)nblk_ * nface_);
    unitIB_ = false;
    uploadCoefs();
}

void MacSolver::setCoefBPacked(const double* const* bcoef) {
    for (int d = 0; d < tree_->ndim; d++) b_[d].assign(bcoef[d], bcoef[d] + (std::size_t)nblk_ * nface_);
    unitB_ = false;
    uploadCoefs();
}

void MacSolver::setRhsPacked(const double* rhs) { check(sayaka_dev_upload(ctx_, SAYAKA_RHS, rhs), "MacSolver::setRhs"); }

void MacSolver::setInitSolPacked(const double* sol) {
    check(sayaka_dev_upload(ctx_, SAYAKA_SOL, sol), "MacSolver::setInitSol");
}

void MacSolver::getSolPacked(double* sol) const { check(sayaka_dev_download(ctx_, SAYAKA_SOL, sol), "MacSolver::getSol"); }

void MacSolver::setIBFrac(const TreeData& vf, int vfcomp, const std::vector<const TreeData*>& af, int afcomp) {
    if ((int)af.size() < tree_->ndim) throw std::invalid_argument("setIBFrac: one face array per direction");
    vf_.resize((std::size_t)nblk_ * ncell_);
    vf.packValid(vfcomp, nblk_, vf_.data());
    for (int d = 0; d < tree_->ndim; d++) {
        af_[d].resize((std::size_t)nblk_ * nface_);
        af[d]->packValid(afcomp, nblk_, af_[d].data());
    }
    unitIB_ = false;
    uploadCoefs();
}

void MacSolver::setDefaultCoefs() {
    unitB_ = true;
    alpha_ = 0.0;
    beta_ = 1.0;
    uploadCoefs();
}

void MacSolver::setUniformCoefB(double bval) {
    for (int d = 0; d < tree_->ndim; d++) b_[d].assign((std::size_t)nblk_ * nface_, bval);
    unitB_ = false;
    uploadCoefs();
}

void MacSolver::setCoefB(const std::vector<TreeData*>& bcoef, int bcomp) {
    if ((int)bcoef.size() < tree_->ndim) throw std::invalid_argument("setCoefB: one face array per direction");
    for (int d = 0; d < tree_->ndim; d++) {
        b_[d].resize((std::size_t)nblk_ * nface_);
        bcoef[d]->packValid(bcomp, nblk_, b_[d].data());
    }
    unitB_ = false;
    uploadCoefs();
}

void MacSolver::setCoefB(int dir, const TreeData& bcoef, int bcomp) {
    if (unitB_)
        for (int d = 0; d < tree_->ndim; d++) b_[d].assign((std::size_t)nblk_ * nface_, 1.0);
    b_[dir].resize((std::size_t)nblk_ * nface_);
    bcoef.packValid(bcomp, nblk_, b_[dir].data());
    unitB_ = false;
    uploadCoefs();
}

void MacSolver::setScalars(double alpha, double beta) {
    alpha_ = alpha;
    beta_ = beta;
    uploadCoefs();
}

// setBC -> BlockBCRecord::setBlock (SayakaMacOperator.cpp:1462-1469, SayakaBoundaryPatch.h:148-167)
void MacSolver::setBC(const BCRegister& reg) {
    const AmrTree& t = *tree_;
    std::vector<int> bc(6 * (std::size_t)nblk_, -1);
    for (int b = 0; b < nblk_; b++)
        for (int f = 0; f < t.numFace(); f++) {
            const int nbr = t[b].neighbor[f];
            if (nbr > PHYS_BNDRY) continue;
            int o[3] = {0, 0, 0};
            o[face_dir(f)] = face_side(f) ? 1 : -1;
            const int m = reg.getBCMap(o[0], o[1], o[2], nbr);
            if (m < 0) throw std::runtime_error("MacSolver::setBC: unknown BC mapping for phys_bc=" + std::to_string(nbr));
            bc[6 * b + f] = m;
        }
    check(sayaka_dev_set_bc(ctx_, bc.data()), "MacSolver::setBC");
}

// fixBC / fixBlockBC (SayakaMacOperator.cpp:1470-1594); the reference exits when phi has no
// ghost layer, this throws
void MacSolver::fixBC(const TreeData& phi, TreeData& rhs, int phicomp, int rhscomp) const {
    std::vector<double> g((std::size_t)nblk_ * phi.extent(0) * phi.extent(1) * phi.extent(2));
    phi.packGrown(phicomp, nblk_, g.data());
    std::vector<double> r((std::size_t)nblk_ * ncell_);
    rhs.packValid(rhscomp, nblk_, r.data());
    fixBCPacked(g.data(), phi.numGrow(), r.data());
    rhs.unpackValid(rhscomp, nblk_, r.data());
}

void MacSolver::fixBCPacked(const double* phi_grown, int ngrow, double* rhs) const {
    if (ngrow < 1) throw std::runtime_error("MacSolver::fixBC: BC value NGROW<1");
    check(sayaka_dev_upload(ctx_, SAYAKA_TMP, rhs), "MacSolver::fixBC");
    check(sayaka_dev_fix_bc(ctx_, phi_grown, ngrow, SAYAKA_TMP), "MacSolver::fixBC");
    check(sayaka_dev_download(ctx_, SAYAKA_TMP, rhs), "MacSolver::fixBC");
}

void MacSolver::setRhs(const TreeData& rhs, int comp) {
    std::vector<double> r((std::size_t)nblk_ * ncell_);
    rhs.packValid(comp, nblk_, r.data());
    setRhsPacked(r.data());
}

void MacSolver::setInitSol(const TreeData& sol, int comp) {
    std::vector<double> r((std::size_t)nblk_ * ncell_);
    sol.packValid(comp, nblk_, r.data());
    setInitSolPacked(r.data());
}

void MacSolver::setZeroSol() { check(sayaka_dev_fill(ctx_, SAYAKA_SOL, 0.0), "MacSolver::setZeroSol"); }

void MacSolver::fluxPacked(int mg_level, int dir, const double* phi, double* flux) {
    check(sayaka_dev_upload(ctx_, SAYAKA_X, phi), "MacSolver::mg_flux_level");
    check(sayaka_dev_flux(ctx_, mg_level, dir, SAYAKA_X, flux), "MacSolver::mg_flux_level");
}

// mg_flux_level (SayakaMacOperator.cpp:161-239) with everything it writes into the caller's data:
// the flux of the tower's blocks (their grown boxes zeroed first, then the valid faces), the
// parents' faces the coarse/fine synchronisation restricts the fine fluxes onto, and phi's first
// ghost layer as the solver's fill leaves it.  The arithmetic runs on the device (flux, ghost
// gather); the host places the values.
void MacSolver::mg_flux_level(int mg_level, int dir, TreeData& flux, TreeData& phi, int fluxcomp, int phicomp) {
    if (flux.faceDir() != dir) throw std::invalid_argument("mg_flux_level: flux must be face data of direction dir");
    if (phi.numGrow() < 1) throw std::invalid_argument("mg_flux_level: phi needs a ghost layer (its fill writes one)");
    const AmrTree& t = *tree_;
    if (mg_level < 1 || mg_level > t.maxAmrLevel || t.getTreeLevel(mg_level).isEmptyLevel()) return;
    std::vector<double> p((std::size_t)nblk_ * ncell_), f((std::size_t)nblk_ * nface_);
    phi.packValid(phicomp, nblk_, p.data());
    fluxPacked(mg_level, dir, p.data(), f.data());

    const int nb = t.nb, nz = t.ndim == 3 ? nb : 1;
    const int fx = nb + (dir == 0), fy = nb + (dir == 1);
    auto face_at = [&](int b, int i, int j, int k) { return f[(std::size_t)b * nface_ + i + fx * (j + fy * k)]; };
    const MGLevelTower& tw = towers_[mg_level];
    // mg_zero_level on the tower, then the tower's faces (SayakaMacOperator.cpp:175-204)
    for (int b : tw.blockIds()) {
        auto& d = flux[b];
        d.setValue(fluxcomp, 1, 0.0);
        const IndexBox& vb = flux.validBox();
        BEGIN_FOR_BOX_RANGE(vb, i, j, k);
        { d(i, j, k, fluxcomp) = face_at(b, i, j, k); }
        END_FOR_BOX_RANGE;
    }
    // syncSubLevelCrseFineFlux (SayakaTower.cpp:360-413): the parent of a fine block with a
    // coarse neighbour across a face of direction dir receives, on that face, the average of the
    // fine fluxes -- the value its coarse neighbour's facing face holds
    for (int l = mg_level; l >= 2; l--)
        for (int q = tw.begin(l); q < tw.end(l); q++) {
            const int b = tw[q];
            for (int s = 0; s <= 1; s++) {
                if (t[b].neighbor[2 * dir + s] != -1) continue;
                const int par = t[b].parent, cn = t[par].neighbor[2 * dir + s];
                if (par < 0 || cn < 0) continue;
                // the parent's face plane on side s, the neighbour's on the other side
                const int a1 = dir == 0 ? 1 : 0, a2 = dir == 2 ? 1 : 2;
                auto& dp = flux[par];
                for (int v = 0; v < (a2 == 2 ? nz : nb); v++)
                    for (int u = 0; u < nb; u++) {
                        int cp[3] = {0, 0, 0}, cc[3] = {0, 0, 0};
                        cp[a1] = cc[a1] = u;
                        cp[a2] = cc[a2] = v;
                        cp[dir] = s ? nb : 0;
                        cc[dir] = s ? 0 : nb;
                        dp(cp[0], cp[1], cp[2], fluxcomp) = face_at(cn, cc[0], cc[1], cc[2]);
                    }
            }
        }
    // mg_fillbndry_level(mg_level, phi, phicomp, 1, 1) (SayakaMacOperator.cpp:181-182)
    const int nf = 2 * t.ndim, nslab = nb * nz;
    std::vector<double> g((std::size_t)nblk_ * nf * nslab);
    check(sayaka_dev_solver_ghosts(ctx_, mg_level, SAYAKA_X, g.data()), "MacSolver::mg_flux_level");
    for (int l = 1; l <= mg_level; l++) {
        if (l < mg_level && tw.isSubLevelEmpty(l + 1)) continue;
        for (int b : t.getTreeLevel(l).blockIds()) {
            auto& d = phi[b];
            for (int fc = 0; fc < nf; fc++) {
                if (t[b].neighbor[fc] <= PHYS_BNDRY) continue;
                const int dd = fc / 2, s = fc % 2;
                const int a1 = dd == 0 ? 1 : 0, a2 = dd == 2 ? 1 : 2;
                for (int pq = 0; pq < nslab; pq++) {
                    int c[3] = {0, 0, 0};
                    c[a1] = pq % nb;
                    c[a2] = pq / nb;
                    c[dd] = s ? (dd == 2 ? nz : nb) : -1;
                    d(c[0], c[1], c[2], phicomp) = g[((std::size_t)b * nf + fc) * nslab + pq];
                }
            }
        }
    }
}

void MacSolver::getSol(TreeData& sol, int comp) const {
    std::vector<double> r((std::size_t)nblk_ * ncell_);
    getSolPacked(r.data());
    sol.unpackValid(comp, nblk_, r.data());
}

int MacSolver::cg_solve() {
    sayaka_solve_params p{};
    p.tol_rel = tolRel_;
    p.tol_abs = tolAbs_;
    p.max_iter = maxIter_;
    p.precond = precond_;
    p.precision = precision_ == Mixed ? SAYAKA_MIXED : precision_ == FP64Fast ? SAYAKA_FP64_FAST : SAYAKA_FP64;
    p.pre_smooth = preSmooth_;
    p.post_smooth = postSmooth_;
    p.hist_cap = 2 * maxIter_ + 2;
    stats_.hist.assign(p.hist_cap, 0.0);
    sayaka_solve_stats st{};
    st.hist = stats_.hist.data();
    const int rc = sayaka_dev_solve(ctx_, &p, &st);
    check(rc, "MacSolver::cg_solve");
    stats_.status = st.status;
    stats_.iters = st.iters;
    stats_.halfStep = st.half_step != 0;
    stats_.bottomCalls = st.bottom_calls;
    stats_.bottomIters = st.bottom_iters;
    stats_.rnorm0 = st.rnorm0;
    stats_.rnorm = st.rnorm;
    stats_.hist.resize(st.nhist < p.hist_cap ? st.nhist : p.hist_cap);
    if (verbose_)
        std::printf("MacSolver::cg_solve: %s, iter=%d, rnorm=%e, rnorm0=%e\n",
                    st.status == 0 ? "converged" : (st.status == 1 ? "breakdown" : "not converged"), st.iters,
                    st.rnorm, st.rnorm0);
    // the reference returns 1 on breakdown and 0 otherwise, converged or not
    return st.status == 1 ? 1 : 0;
}

int MacSolver::mg_solve() {
    const int fin = tree_->currentFinestLevel();
    const int R = SAYAKA_RESID, X = SAYAKA_SOL, B = SAYAKA_RHS, CORR = SAYAKA_CORR;
    double dot = 0.0, rnorm0 = 0.0, rnorm = 0.0;
    check(sayaka_dev_residual(ctx_, fin, R, X, B), "MacSolver::mg_solve");
    check(sayaka_dev_dot(ctx_, fin, R, R, &dot, &rnorm0), "MacSolver::mg_solve");
    const int max_iter = mgUseFixedIter() ? mgFixedIter_ : maxIter_;
    bool conv = false;
    int iter = 0;
    for (iter = 1; iter <= max_iter; iter++) {
        check(sayaka_dev_fill(ctx_, CORR, 0.0), "MacSolver::mg_solve");
        for (int l = fin; l >= 2; l--) check(sayaka_dev_restrict(ctx_, l - 1, R), "MacSolver::mg_solve");
        int bit = 0, bst = 0;
        // mg_solve_bottom (SayakaMacSolver.cpp:164-242): level-1 BiCGSTAB to 1e-6 / 1e-6, 2000 its
        check(sayaka_dev_bottom(ctx_, CORR, R, 1.0e-6, 1.0e-6, 2000, &bit, &bst), "MacSolver::mg_solve");
        if (bst == 1) throw std::runtime_error("MacSolver::mg_solve: MG: Bottom break down");
        for (int l = 2; l <= fin; l++) {
            check(sayaka_dev_prolong(ctx_, l, CORR), "MacSolver::mg_solve");
            for (int q = 0; q < postSmooth_; q++) check(sayaka_dev_smooth(ctx_, l, CORR, R), "MacSolver::mg_solve");
        }
        check(sayaka_dev_add(ctx_, fin, X, CORR), "MacSolver::mg_solve");
        check(sayaka_dev_residual(ctx_, fin, R, X, B), "MacSolver::mg_solve");
        check(sayaka_dev_dot(ctx_, fin, R, R, &dot, &rnorm), "MacSolver::mg_solve");
        if (rnorm < rnorm0 * tolRel_ + tolAbs_) conv = true;
        if (!mgUseFixedIter() && conv) break;
    }
    stats_ = SolveStats{};
    stats_.status = conv ? 0 : 2;
    stats_.iters = iter > max_iter ? max_iter : iter;
    stats_.rnorm0 = rnorm0;
    stats_.rnorm = rnorm;
    if (verbose_ || (!mgUseFixedIter() && !conv))
        std::printf("MG: %s, iter=%d, |resid|=%e, |resid0|=%e\n", conv ? "converge" : "failure", stats_.iters, rnorm,
                    rnorm0);
    return 0;
}

int MacSolver::relax_solve() {
    const int fin = tree_->currentFinestLevel();
    const int R = SAYAKA_RESID, X = SAYAKA_SOL, B = SAYAKA_RHS;
    double dot = 0.0, rnorm0 = 0.0, rnorm = 0.0;
    check(sayaka_dev_fill(ctx_, SAYAKA_CORR, 0.0), "MacSolver::relax_solve");
    check(sayaka_dev_residual(ctx_, fin, R, X, B), "MacSolver::relax_solve");
    check(sayaka_dev_dot(ctx_, fin, R, R, &dot, &rnorm0), "MacSolver::relax_solve");
    bool conv = false;
    int iter = 0;
    for (iter = 1; iter <= maxIter_; iter++) {
        check(sayaka_dev_smooth(ctx_, fin, X, B), "MacSolver::relax_solve");
        check(sayaka_dev_residual(ctx_, fin, R, X, B), "MacSolver::relax_solve");
        check(sayaka_dev_dot(ctx_, fin, R, R, &dot, &rnorm), "MacSolver::relax_solve");
        if (rnorm < rnorm0 * tolRel_ + tolAbs_) conv = true;
        if (conv) break;
    }
    stats_ = SolveStats{};
    stats_.status = conv ? 0 : 2;
    stats_.iters = iter > maxIter_ ? maxIter_ : iter;
    stats_.rnorm0 = rnorm0;
    stats_.rnorm = rnorm;
    if (verbose_ || !conv)
        std::printf("relax_solve: %s, iter=%d, |resid|=%e\n", conv ? "converge" : "failure", stats_.iters, rnorm);
    return 0;
}

}  // namespace amd
}  // namespace sayaka
