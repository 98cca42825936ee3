// mac_solver.h -- MAC Poisson solver over an AmrTree, solving on the GPU.
//
// The call surface of the reference MacSolver (SayakaMacSolver.h:25-425) for the solve path,
// with the same call order as its drivers (Test7.h:612-653):
//
//     MacSolver mac(tree);
//     mac.prepareSolver();            // after every regrid
//     mac.setBC(bcreg);
//     mac.setIBFracByLS(ls, 0);       // or setIBFrac / setUnitIBFrac
//     mac.setCoefB(bcoefs, 0);
//     mac.setRhs(rhs, 0);  mac.setZeroSol();
//     mac.setTolerance(1e-10, 0.0);  mac.setMaxIteration(2000);
//     int stat = mac.cg_solve();      // 0 solved or not converged, 1 breakdown (as the reference)
//     mac.getSol(phi, 0);
//
// The tree stays on the host and is owned by the caller (it must outlive the solver); the
// solver owns a device context (include/sayaka_amd.h) that holds every field as packed
// valid-only blocks.  prepareSolver() sends the tree's tables; the coefficient, BC, rhs and
// solution calls copy the valid region of the caller's TreeData to or from the device.
// Errors the reference treats with exit(1) throw std::runtime_error here.
#pragma once

#include <string>
#include <vector>

#include "amr_tree.h"
#include "bc_register.h"
#include "tree_data.h"

struct sayaka_dev_ctx;

namespace sayaka {
inline namespace amd {

struct SolveStats {
    int status = 0;          // 0 converged, 1 breakdown, 2 iteration limit
    int iters = 0;
    bool halfStep = false;
    int bottomCalls = 0;
    long long bottomIters = 0;
    double rnorm0 = 0.0, rnorm = 0.0;
    std::vector<double> hist;   // -|s| at half steps, +|r| at full steps
};

// The blocks of one multigrid level's tower (MGLevel / MGLevelTower, SayakaTower.h:50-224):
// for sub-levels l = m .. 1, the level-l blocks that are leaves or have l == m, each sub-level in
// ascending id; [begin(l), end(l)) is sub-level l's range of the list.  Rebuilt by
// prepareSolver.  The face registers stay on the device (its index tables).
struct MGLevelTower : AmrTreeLevel {
    enum TowerFluxSyncMode { FLUX_SYNC_AVG, FLUX_SYNC_SUM };
    std::vector<int> level_begin, level_end;
    int getMGLevel() const { return level; }
    int begin(int l) const { return level_begin[l]; }
    int end(int l) const { return level_end[l]; }
    int size(int l) const { return end(l) - begin(l); }
    int coversTreeLevel(int l) const { return l <= level; }
    int hasTreeLevel(int l) const { return coversTreeLevel(l) && size(l) > 0; }
    int isSubLevelEmpty(int l) const { return size(l) <= 0; }
    using AmrTreeLevel::begin;
    using AmrTreeLevel::end;
    using AmrTreeLevel::size;
};

class MacSolver {
   public:
    enum Precision { FP64 = 0, Mixed = 1, FP64Fast = 2 };

    // Works in the tree's device context (AmrTree::deviceContext, shared with the tree's
    // FillPatches); the solver's device fields live there, so of several MacSolvers made on one
    // tree (a driver that makes a new one every step) only one is used at a time.
    explicit MacSolver(const AmrTree& tree, int device = 0);
    ~MacSolver();
    MacSolver(const MacSolver&) = delete;
    MacSolver& operator=(const MacSolver&) = delete;

    void setVerbose(int v) { verbose_ = v; }
    // MacSolver::prepareSolver (SayakaMacSolver.cpp:12-26): tables of the current tree, unit
    // coefficients.  Call again after every regrid.
    void prepareSolver();
    void clearSolver() {}

    // coefficients (SayakaMacIB.cpp:12-154, SayakaMacCoef.cpp:8-76, 319-367)
    void setUnitIBFrac();
    void setIBFracByLS(const TreeData& ls, int lscomp);
    // fractional volume / area fractions given directly (cut cells); af[dir] face data
    void setIBFrac(const TreeData& vf, int vfcomp, const std::vector<const TreeData*>& af, int afcomp);
    void setDefaultCoefs();
    void setUniformCoefB(double b);
    void setCoefB(const std::vector<TreeData*>& bcoef, int bcomp);
    void setCoefB(int dir, const TreeData& bcoef, int bcomp);
    void setScalars(double alpha, double beta);
    void setAlpha(double alpha) { setScalars(alpha, beta_); }
    void setBeta(double beta) { setScalars(alpha_, beta); }

    // boundary conditions (SayakaMacOperator.cpp:1462-1594)
    void setBC(const BCRegister& reg);
    void fixBC(const TreeData& phi, TreeData& rhs, int phicomp, int rhscomp) const;

    // data in / out (SayakaMacSolver.h:117-132)
    void setRhs(const TreeData& rhs, int comp);
    void setInitSol(const TreeData& sol, int comp);
    void setZeroSol();
    void getSol(TreeData& sol, int comp) const;

    // parameters (SayakaMacSolver.h:349-358)
    void setTolerance(double rel, double abs) { tolRel_ = rel; tolAbs_ = abs; }
    void setMaxIteration(int n) { maxIter_ = n; }
    void setPreSmoothNum(int n) { preSmooth_ = n; }
    void setPostSmoothNum(int n) { postSmooth_ = n; }
    void setCGUseMGPrecond(int on) { precond_ = on; }
    // mg_solve runs exactly n V-cycles when n > 0 (SayakaMacSolver.h:355-356)
    void setMGFixedIter(int n) { mgFixedIter_ = n; }
    bool mgUseFixedIter() const { return mgFixedIter_ > 0; }
    // the A coefficient array (SayakaMacCoef.cpp:15-18): kept, never read -- the operator
    // requires alpha == 0, as the reference's does
    void setUniformCoefA(double a) { coefA_ = a; }
    void setPrecision(Precision p) { precision_ = p; }

    // The same calls on packed valid-only arrays ([blk][k][j][i] over the tree's blocks, face
    // arrays per direction; the level set over the grown box with ngrow >= 1 ghost layers):
    // the form the flat C ABI (include/sayaka_host.h) passes through.
    void setIBFracByLSPacked(const double* ls_grown, int ngrow);
    void setIBFracPacked(const double* vf, const double* const* af);
    void setCoefBPacked(const double* const* bcoef);
    void setRhsPacked(const double* rhs);
    void setInitSolPacked(const double* sol);
    void getSolPacked(double* sol) const;
    // fixBC on phi over the grown box ([blk][grown box], ngrow >= 1 layers, i fastest) and the
    // packed rhs, updated in place
    void fixBCPacked(const double* phi_grown, int ngrow, double* rhs) const;
    int numBlocks() const { return nblk_; }

    // MacSolver::cg_solve (SayakaMacSolver.cpp:455-680)
    int cg_solve();
    // MacSolver::mg_solve (SayakaMacSolver.cpp:43-162): V-cycles on the finest tower from the
    // current solution (no pre-smooth, restriction of the residual, coarsest-level BiCGSTAB,
    // prolongation + post-smooths, correction) until |r| < |r0| tol_rel + tol_abs, or exactly
    // setMGFixedIter cycles; every operation in FP64 in the reference's order on the GPU.
    // Returns 0 as the reference does (lastStats() holds the cycles and the residuals).
    int mg_solve();
    // MacSolver::relax_solve (SayakaMacSolver.cpp:377-452): red-black Gauss-Seidel passes on the
    // finest tower until the same test holds or max_iter passes; returns 0.
    int relax_solve();

    // MacSolver::mg_flux_level (SayakaMacSolver.h:275, SayakaMacOperator.cpp:161-239): the face
    // flux -d(phi)/dx of direction dir on tower mg_level into flux (face data of dir): in-block
    // and same-level faces over dx, the fine side of coarse/fine faces over 1.5 dx, the coarse
    // side the average of the fine fluxes, walls 0.  Like the reference it also zeroes the
    // tower blocks' flux ghosts, writes the restricted fine fluxes onto the faces of their
    // parents, and leaves phi's first ghost layer as its fill sets it (injection, no corners, no
    // walls); phi needs at least one ghost layer.  The device context's X field carries phi.
    void mg_flux_level(int mg_level, int dir, TreeData& flux, TreeData& phi, int fluxcomp, int phicomp);
    // the flux of the tower only, packed [blk][face] (0 outside the tower)
    void fluxPacked(int mg_level, int dir, const double* phi, double* flux);
    const SolveStats& lastStats() const { return stats_; }
    // the tower of multigrid level l (SayakaMacSolver.h:187-188), l = 1 .. maxAmrLevel; valid
    // from construction on, filled by prepareSolver
    const MGLevelTower& getTower(int l) const { return towers_.at(l); }
    MGLevelTower& getTower(int l) { return towers_.at(l); }

    const AmrTree& tree() const { return *tree_; }
    sayaka_dev_ctx* deviceContext() const { return ctx_; }

   private:
    void uploadCoefs();
    void check(int rc, const char* what) const;

    const AmrTree* tree_;
    sayaka_dev_ctx* ctx_ = nullptr;
    int nblk_ = 0, ncell_ = 0, nface_ = 0;
    int verbose_ = 0;
    double alpha_ = 0.0, beta_ = 1.0;
    double tolRel_ = 1e-10, tolAbs_ = 0.0;
    int maxIter_ = 2000, preSmooth_ = 2, postSmooth_ = 2, precond_ = 1, mgFixedIter_ = 0;
    double coefA_ = 0.0;
    Precision precision_ = FP64;
    bool unitIB_ = true, unitB_ = true;
    std::vector<double> vf_, af_[3], b_[3];   // leaf values, packed valid [blk][...]
    SolveStats stats_;
    std::vector<MGLevelTower> towers_;
};

}  // namespace amd
}  // namespace sayaka
