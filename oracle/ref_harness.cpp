// ref_harness.cpp -- C-ABI harness around the *reference* SayakaAmr solver.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the reference
// sources (where they lie under the reference checkout) by oracle/build_ref.sh
// into oracle/_ref/libsayaka_ref{2,3}.so.  It never ships in the product and is
// only loaded by tests/ (golden-vector generation, oracle pinning) and by the
// cpu_baseline leg of bench.py.  It contains no reference code: it only calls
// the reference's public API (AmrTree, TreeData, MacSolver, MGLevelTower,
// FillPatch) and copies data in and out of flat arrays.
//
// Flat layouts used across the whole project:
//   cell field, valid only : [block][k][j][i]  (nb^d values per block)
//   cell field, grown      : [block][k][j][i]  over the block grown by g
//   face field dir d       : [block][k][j][i]  over the cell box staggered in d
//                            (i fastest, exactly IndexBox::offset order,
//                             SayakaBox.h:214-217)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>
#include <fcntl.h>

#include "SayakaMacSolver.h"
#include "SayakaBoundaryPatch.h"

using namespace sayaka;

namespace {

// MacSolver keeps its IB fractions protected; a subclass reaches them so that the
// harness can inject fractional cut-cell weights (the reference only has a binary
// setter, SayakaMacIB.cpp:99-154).
struct HarnessMac : public MacSolver {
    explicit HarnessMac(const AmrTree& t) : MacSolver(t) {}
    TreeData& volfrac() { return *m_ib_volfrac; }
    std::vector<TreeData*>& areafrac() { return m_ib_areafrac; }
    std::vector<TreeData*>& bcoef() { return m_bcoef; }
};

AmrTree* g_tree = nullptr;
HarnessMac* g_mac = nullptr;
int g_nb = 0;

IndexBox valid_box() { return g_tree->validBlockCellBox(); }

IndexBox face_box(int dir) {
    IndexBox b = valid_box();
    b.staggerInDir(dir);
    return b;
}

// --- stdout silencing (the reference prints through LOGPRINTF to stdout) -------
int g_saved_fd = -1;
void quiet_begin() {
    fflush(stdout);
    g_saved_fd = dup(1);
    int dn = open("/dev/null", O_WRONLY);
    dup2(dn, 1);
    close(dn);
}
void quiet_end() {
    fflush(stdout);
    if (g_saved_fd >= 0) {
        dup2(g_saved_fd, 1);
        close(g_saved_fd);
        g_saved_fd = -1;
    }
}

template <class F>
void for_box(const IndexBox& b, F f) {
    for (int k = b.klo(); k <= b.khi(); k++)
        for (int j = b.jlo(); j <= b.jhi(); j++)
            for (int i = b.ilo(); i <= b.ihi(); i++) f(i, j, k);
}

void copy_in_box(TreeData& td, int comp, const IndexBox& b, const double* src, int nblk) {
    const int n = b.stride();
    for (int ib = 0; ib < nblk; ib++) {
        const double* s = src + (size_t)ib * n;
        DoubleBlockData& d = td[ib];
        int c = 0;
        for_box(b, [&](int i, int j, int k) { d(i, j, k, comp) = s[c++]; });
    }
}
void copy_out_box(const TreeData& td, int comp, const IndexBox& b, double* dst, int nblk) {
    const int n = b.stride();
    for (int ib = 0; ib < nblk; ib++) {
        double* o = dst + (size_t)ib * n;
        const DoubleBlockData& d = td[ib];
        int c = 0;
        for_box(b, [&](int i, int j, int k) { o[c++] = d(i, j, k, comp); });
    }
}

IndexBox grown_cell_box(int g) {
    IndexBox b = valid_box();
    b.extend(g);
    return b;
}

}  // namespace

extern "C" {

int ref_ndim() { return NDIM; }

// AmrTree::init / initUniformRootLevel / initRefineToMinLevel (SayakaTree.cpp:34-233).
int ref_tree_create(int nb, int ngrow, int minlevel, int maxlevel, const int* ndiv,
                    const double* plo, const double* phi, const int* bc) {
    if (g_mac) { delete g_mac; g_mac = nullptr; }
    if (g_tree) { g_tree = nullptr; }  // the reference never frees its pools; leak on purpose
    g_tree = new AmrTree;
    g_nb = nb;
    const int nbz = NDIM == 3 ? nb : 1;
    const int gz = NDIM == 3 ? ngrow : 0;
    g_tree->minAmrLevel = minlevel;
    g_tree->maxAmrLevel = maxlevel;
    g_tree->blockValidIndex = IndexBox(Vector3i::VecMake(0, 0, 0), Vector3i::VecMake(nb - 1, nb - 1, nbz - 1));
    g_tree->blockGrownIndex = IndexBox(g_tree->blockValidIndex).extend(ngrow, ngrow, gz);
    g_tree->blockNumGrow = ngrow;
    quiet_begin();
    g_tree->init();
    RealBox pb(Vector3d::VecMake(plo[0], plo[1], plo[2]), Vector3d::VecMake(phi[0], phi[1], phi[2]));
    int bct[MAX_FACE];
    bool per[MAX_FACE];
    // bc[f] == 1 marks a periodic face (initUniformRootLevel's is_periodic), otherwise the wall code
    for (int f = 0; f < MAX_FACE; f++) { bct[f] = bc[f] == 1 ? -100 : bc[f]; per[f] = bc[f] == 1; }
    g_tree->initUniformRootLevel(Vector3i::VecMake(ndiv[0], ndiv[1], ndiv[2]), pb, bct, per);
    g_tree->initRefineToMinLevel();
    quiet_end();
    return g_tree->numBlocks;
}

int ref_tree_numblocks() { return g_tree->numBlocks; }
int ref_tree_finest() { return g_tree->currentFinestLevel(); }

void ref_tree_set_flags(const int* refine, const int* coarsen, int n) {
    for (int i = 0; i < n; i++) {
        if (refine) g_tree->refineFlag[i] = refine[i];
        if (coarsen) g_tree->coarsenFlag[i] = coarsen[i];
    }
}

// AmrTree::regrid (SayakaRegrid.cpp:26-125)
int ref_tree_regrid() {
    quiet_begin();
    g_tree->regrid();
    quiet_end();
    return g_tree->numBlocks;
}

// Flat export of every per-node field the regrid produces (SayakaTree.h:52-152,
// 336-526).  Surrounding arrays keep the reference's [ii+1][jj+1][kk+1] order.
void ref_tree_export(int* level, int* parent, int* which_child, int* child8, int* nbr6,
                     int* node_type, int* node_type2, double* lo3, double* hi3,
                     double* center3, double* length3, int* surr27, int* surrtype27,
                     int* surrbcref27) {
    const int n = g_tree->numBlocks;
    for (int b = 0; b < n; b++) {
        const AmrTreeNode& nd = (*g_tree)[b];
        level[b] = nd.refineLevel;
        parent[b] = nd.parent;
        which_child[b] = nd.whichChild;
        node_type[b] = nd.nodeType;
        for (int c = 0; c < 8; c++) child8[b * 8 + c] = nd.child[c];
        for (int f = 0; f < 6; f++) {
            nbr6[b * 6 + f] = nd.neighbor[f];
            node_type2[b * 6 + f] = nd.nodeType2[f];
        }
        for (int d = 0; d < 3; d++) {
            lo3[b * 3 + d] = nd.boundBox.vlo(d);
            hi3[b * 3 + d] = nd.boundBox.vhi(d);
            center3[b * 3 + d] = nd.blockCenter(d);
            length3[b * 3 + d] = nd.blockLength(d);
        }
        const SurroundingBlocks& s = g_tree->surrBlocks[b];
        for (int q = 0; q < 27; q++) {
            surr27[b * 27 + q] = (&s.blocks[0][0][0])[q];
            surrtype27[b * 27 + q] = (&s.types[0][0][0])[q];
            surrbcref27[b * 27 + q] = (&s.bcrefs[0][0][0])[q];
        }
    }
}

int ref_tree_level_count(int l) { return g_tree->getTreeLevel(l).numLevelBlock; }
void ref_tree_level_blocks(int l, int* out) {
    const AmrTreeLevel& lv = g_tree->getTreeLevel(l);
    for (int i = 0; i < lv.numLevelBlock; i++) out[i] = lv.levelBlock[i];
}

// ---------------------------------------------------------------- MacSolver ---
// MacSolver(tree) + prepareSolver (SayakaMacSolver.h:56-87, SayakaMacSolver.cpp:12-26)
void ref_mac_create(int verbose) {
    if (g_mac) delete g_mac;
    quiet_begin();
    g_mac = new HarnessMac(*g_tree);
    g_mac->setVerbose(verbose);
    quiet_end();
}
void ref_mac_prepare() {
    quiet_begin();
    g_mac->prepareSolver();
    quiet_end();
}

int ref_tower_size(int m) { return g_mac->getTower(m).numLevelBlock; }

// Tower lists and face registers (SayakaTower.cpp:28-187).
void ref_tower_export(int m, int* blocks, int* begin, int* end, int* face_type6,
                      int* face_neigh6, int* fc_off18, int* cf_neigh24, int* cf_off72,
                      int* cf_box144) {
    const MGLevelTower& t = g_mac->getTower(m);
    for (int l = 1; l <= m; l++) { begin[l] = t.begin(l); end[l] = t.end(l); }
    for (int q = 0; q < t.numLevelBlock; q++) {
        const int b = t[q];
        blocks[q] = b;
        for (int f = 0; f < NUM_FACE; f++) {
            const BlockFaceRegister& r = t.getFaceReg(b, f);
            face_type6[q * 6 + f] = r.face_type;
            face_neigh6[q * 6 + f] = r.face_neigh;
            for (int d = 0; d < 3; d++) fc_off18[(q * 6 + f) * 3 + d] = r.face_type == FACE_FINE_CRSE ? r.fine_crse_offset(d) : 0;
            for (int c = 0; c < 4; c++) {
                const bool cf = r.face_type == FACE_CRSE_FINE && c < ChildIndex::NumChildOnFace;
                cf_neigh24[(q * 6 + f) * 4 + c] = cf ? r.crse_fine_neigh[c] : -1;
                for (int d = 0; d < 3; d++) {
                    cf_off72[((q * 6 + f) * 4 + c) * 3 + d] = cf ? r.crse_fine_offset[c](d) : 0;
                    cf_box144[((q * 6 + f) * 4 + c) * 6 + d] = cf ? r.crse_fine_subbox[c].lo(d) : 0;
                    cf_box144[((q * 6 + f) * 4 + c) * 6 + 3 + d] = cf ? r.crse_fine_subbox[c].hi(d) : 0;
                }
            }
        }
    }
}

// setCoefB on every direction (SayakaMacCoef.cpp:319-367).  b[dir] : face fields.
void ref_mac_set_coefb(const double* bx, const double* by, const double* bz) {
    const double* src[3] = {bx, by, bz};
    const int n = g_tree->numBlocks;
    std::vector<TreeData*> tmp = TreeData::CreateFaceDataPArray(*g_tree, 1, 0);
    for (int d = 0; d < NDIM; d++) copy_in_box(*tmp[d], 0, face_box(d), src[d], n);
    g_mac->setCoefB(tmp, 0);
    TreeData::ReleaseDataPArray(tmp);
}

// Fractional IB weights: leaf values, then exactly the sync + average-down steps of
// calcIBFracByLS (SayakaMacIB.cpp:69-96), driven through the reference's public API.
void ref_mac_set_ib_frac(const double* vf, const double* ax, const double* ay, const double* az) {
    const double* af[3] = {ax, ay, az};
    const AmrTree& tree = *g_tree;
    const int finest = tree.currentFinestLevel();
    const MGLevelTower& tower = g_mac->getTower(finest);
    TreeData& vfr = g_mac->volfrac();
    std::vector<TreeData*>& afr = g_mac->areafrac();
    const IndexBox vb = valid_box();
    const int nc = vb.stride();
    for (int q = 0; q < tower.numLevelBlock; q++) {
        const int b = tower[q];
        int c = 0;
        for_box(vb, [&](int i, int j, int k) { vfr[b](i, j, k, 0) = vf[(size_t)b * nc + c++]; });
        for (int d = 0; d < NDIM; d++) {
            const IndexBox fb = face_box(d);
            const int nf = fb.stride();
            int cc = 0;
            for_box(fb, [&](int i, int j, int k) { (*afr[d])[b](i, j, k, 0) = af[d][(size_t)b * nf + cc++]; });
        }
    }
    for (int l = finest; l >= 2; l--)
        for (int d = 0; d < NDIM; d++)
            tower.syncSubLevelCrseFineFlux(l, l - 1, d, *afr[d], 0, MGLevelTower::FLUX_SYNC_AVG);
    {
        FillPatch fill(const_cast<AmrTree&>(tree), vfr);
        fill.setRestriction(0, InterpPatch::RESTRICT_AVERAGE);
        for (int l = finest - 1; l >= 1; l--) fill.restrictLevel(l, 0, 1);
    }
    for (int d = 0; d < NDIM; d++) {
        FillPatch fill(const_cast<AmrTree&>(tree), *afr[d]);
        fill.setRestriction(0, InterpPatch::RESTRICT_AVERAGE);
        for (int l = finest - 1; l >= 1; l--) fill.restrictLevel(l, 0, 1);
    }
}

// Binary IB from a level set with ghosts (SayakaMacIB.cpp:34-154).
void ref_mac_set_ib_ls(const double* ls_grown, int ngrow) {
    TreeData* ls = TreeData::CreateCellData(*g_tree, 1, ngrow);
    copy_in_box(*ls, 0, grown_cell_box(ngrow), ls_grown, g_tree->numBlocks);
    g_mac->setIBFracByLS(*ls, 0);
    delete ls;
}

// Read back the solver's coefficient state for every block.
void ref_mac_get_coefs(double* vf, double* bx, double* by, double* bz, double* ax, double* ay, double* az) {
    const int n = g_tree->numBlocks;
    if (vf) copy_out_box(g_mac->volfrac(), 0, valid_box(), vf, n);
    double* bo[3] = {bx, by, bz};
    double* ao[3] = {ax, ay, az};
    for (int d = 0; d < NDIM; d++) {
        if (bo[d]) copy_out_box(*g_mac->bcoef()[d], 0, face_box(d), bo[d], n);
        if (ao[d]) copy_out_box(*g_mac->areafrac()[d], 0, face_box(d), ao[d], n);
    }
}

// setBC with one (phys_bc -> math_bc, value) map applied to every surrounding
// position (BCRegister::setBCMap/setBCVal, SayakaBoundaryPatch.h:75-98).
void ref_mac_set_bc(int n, const int* physbc, const int* mathbc, const double* val) {
    BCRegister reg;
    for (int q = 0; q < SurroundingIndex::NumSurr; q++) {
        SurroundingIndex s(q);
        for (int t = 0; t < n; t++) {
            reg.setBCMap(s.ipos(), s.jpos(), s.kpos(), physbc[t], mathbc[t]);
            reg.setBCVal(s.ipos(), s.jpos(), s.kpos(), physbc[t], val[t]);
        }
    }
    g_mac->setBC(reg);
}

// fixBC (SayakaMacOperator.cpp:1470-1594): phi carries BC values in its ghosts.
void ref_mac_fix_bc(const double* phi_grown, int ngrow, double* rhs_valid) {
    const int n = g_tree->numBlocks;
    TreeData* phi = TreeData::CreateCellData(*g_tree, 1, ngrow);
    TreeData* rhs = TreeData::CreateCellData(*g_tree, 1, 0);
    copy_in_box(*phi, 0, grown_cell_box(ngrow), phi_grown, n);
    copy_in_box(*rhs, 0, valid_box(), rhs_valid, n);
    g_mac->fixBC(*phi, *rhs, 0, 0);
    copy_out_box(*rhs, 0, valid_box(), rhs_valid, n);
    delete phi;
    delete rhs;
}

void ref_mac_set_rhs(const double* rhs_valid) {
    TreeData* rhs = TreeData::CreateCellData(*g_tree, 1, 0);
    copy_in_box(*rhs, 0, valid_box(), rhs_valid, g_tree->numBlocks);
    g_mac->setRhs(*rhs, 0);
    delete rhs;
}
void ref_mac_set_init_sol(const double* sol_valid) {
    TreeData* s = TreeData::CreateCellData(*g_tree, 1, 0);
    copy_in_box(*s, 0, valid_box(), sol_valid, g_tree->numBlocks);
    g_mac->setInitSol(*s, 0);
    delete s;
}
void ref_mac_set_zero_sol() { g_mac->setZeroSol(); }
void ref_mac_get_sol(double* sol_valid) {
    TreeData* s = TreeData::CreateCellData(*g_tree, 1, 0);
    g_mac->getSol(*s, 0);
    copy_out_box(*s, 0, valid_box(), sol_valid, g_tree->numBlocks);
    delete s;
}
void ref_mac_set_params(double tol_rel, double tol_abs, int max_iter, int precond) {
    g_mac->setTolerance(tol_rel, tol_abs);
    g_mac->setMaxIteration(max_iter);
    g_mac->setCGUseMGPrecond(precond);
}

// mg_solve / relax_solve (SayakaMacSolver.cpp:43-162, 377-452): the multigrid and the
// relaxation solve of the finest tower from the SOL buffer as initial guess; fixed_iter > 0
// is setMGFixedIter (mg_solve runs exactly that many V-cycles).
void ref_mac_mg_solve(int fixed_iter) {
    g_mac->setMGFixedIter(fixed_iter);
    quiet_begin();
    g_mac->mg_solve();
    quiet_end();
}
void ref_mac_relax_solve() {
    quiet_begin();
    g_mac->relax_solve();
    quiet_end();
}

// cg_solve (SayakaMacSolver.cpp:455-680).  The reference returns only a breakdown
// flag; the iteration count and the residual history are recovered from its
// verbose log lines (SayakaMacSolver.cpp:768, 836, 887), as the survey prescribes.
// hist receives, per logged line, +|r| for a full step and -|s| for a half step.
int ref_mac_cg_solve(int* iters, int* converged, double* hist, int hist_cap, int* nhist) {
    char* buf = nullptr;
    size_t len = 0;
    FILE* mem = open_memstream(&buf, &len);
    FILE* old = CLogPrintf::ms_fp;
    CLogPrintf::ms_fp = mem;
    const int vsave = g_mac->Verbose();
    g_mac->setVerbose(2);
    quiet_begin();
    int stat = g_mac->cg_solve();
    quiet_end();
    g_mac->setVerbose(vsave);
    CLogPrintf::ms_fp = old;
    fclose(mem);
    int it = 0, conv = 0, nh = 0;
    const char* p = buf;
    while (p && *p) {
        const char* e = strchr(p, '\n');
        std::string line(p, e ? (size_t)(e - p) : strlen(p));
        int a, b;
        double v;
        if (sscanf(line.c_str(), "cg_solve_level: iter=%d/%d", &a, &b) == 2) it = a;
        else if (sscanf(line.c_str(), "cg_solve_level: half |resid|=%lf", &v) == 1) { if (nh < hist_cap) hist[nh] = -v; nh++; }
        else if (sscanf(line.c_str(), "cg_solve_level: |resid|=%lf", &v) == 1) { if (nh < hist_cap) hist[nh] = v; nh++; }
        else if (line.find("cg_solve_level: converge") != std::string::npos) conv = 1;
        p = e ? e + 1 : nullptr;
    }
    free(buf);
    if (iters) *iters = it;
    if (converged) *converged = conv;
    if (nhist) *nhist = nh;
    return stat;
}

// ------------------------------------------------ component-level operations ---
// The solver buffer has NUM_COMP_BUF components with one ghost layer
// (SayakaMacSolver.h:34-53); these move one component in/out in grown layout.
int ref_buf_ncomp() { return MacSolver::NUM_COMP_BUF; }
void ref_buf_set(int comp, const double* grown) {
    copy_in_box(g_mac->getDataBuf(), comp, grown_cell_box(1), grown, g_tree->numBlocks);
}
void ref_buf_get(int comp, double* grown) {
    copy_out_box(g_mac->getDataBuf(), comp, grown_cell_box(1), grown, g_tree->numBlocks);
}
void ref_save_get(int dir, double* faces) {
    copy_out_box(g_mac->getFluxSaveDataBuf(dir), 0, face_box(dir), faces, g_tree->numBlocks);
}
void ref_apply(int m, int outcomp, int incomp) {
    TreeData& d = g_mac->getDataBuf();
    g_mac->mg_apply_level(m, d, d, outcomp, incomp);
}
void ref_residual(int m, int rcomp, int xcomp, int rhscomp) {
    TreeData& d = g_mac->getDataBuf();
    g_mac->mg_residual_level(m, d, d, d, rcomp, xcomp, rhscomp);
}
void ref_smooth(int m, int solcomp, int rhscomp) {
    TreeData& d = g_mac->getDataBuf();
    g_mac->mg_smooth_level(m, d, d, solcomp, rhscomp);
}
void ref_fill(int m, int comp) { g_mac->mg_fillbndry_level(m, g_mac->getDataBuf(), comp, 1, 1); }

// A one-component cell state registered with the tree the way the reference's demos do it
// (Test6.h:596-660): a FillPatch with its default injection prolongation and a BCPatch with
// Neumann on the wall code physbc, so that regrid packs the data along with the blocks
// (SayakaRegrid.cpp:530-591) and initNewBlockDataAfterRegrid (987-1105) fills new blocks.
void ref_state_create(int ngrow, int physbc) {
    TreeData* st = TreeData::CreateCellData(*g_tree, 1, ngrow);
    BCPatch* bc = new BCPatch();
    bc->defineByData(*st);
    for (int q = 0; q < SurroundingIndex::NumSurr; q++) {
        SurroundingIndex s(q);
        bc->boundaryRegister(0).setBCMap(s.ipos(), s.jpos(), s.kpos(), physbc, 0);
    }
    FillPatch* fp = new FillPatch(*g_tree, *st);
    fp->setBC(*bc);
    TreeStateData tsd;
    tsd.new_data = st;
    tsd.bc_patch = bc;
    tsd.fill_new_patch = fp;
    g_tree->treeStateData.push_back(tsd);
    IndexBox gb = valid_box();
    gb.extend(ngrow);
    std::vector<double> zero((size_t)g_tree->maxBlockNum * gb.stride(), 0.0);
    copy_in_box(*st, 0, gb, zero.data(), g_tree->maxBlockNum);
}
void ref_state_set(const double* valid) {
    copy_in_box(*g_tree->treeStateData.back().new_data, 0, valid_box(), valid, g_tree->numBlocks);
}
void ref_state_get(double* valid) {
    copy_out_box(*g_tree->treeStateData.back().new_data, 0, valid_box(), valid, g_tree->numBlocks);
}
// A general registered state, as the demos register theirs (cell states, Test5.h:590-600;
// face velocity states per direction, Test4.h:455-470): ncomp components with ngrow ghost
// layers, cell data (dir < 0) or face data staggered in dir, a FillPatch with one prolongation
// per component and, for cell data, a BCPatch mapping the nphys physical codes to
// (mathbc, bcval)[comp][h] on every surrounding position.  Every slot of the pool starts at 0.
static IndexBox state_box(int dir, int ngrow) {
    IndexBox gb = grown_cell_box(ngrow);
    if (dir >= 0) gb.staggerInDir(dir);
    return gb;
}
static int g_state_dir = -1, g_state_ngrow = 0;
void ref_state_create_general(int dir, int ncomp, int ngrow, const int* prolong, int nphys, const int* physbc,
                              const int* mathbc, const double* bcval) {
    TreeData* st = dir < 0 ? TreeData::CreateCellData(*g_tree, ncomp, ngrow)
                           : TreeData::CreateFaceData(*g_tree, dir, ncomp, ngrow);
    FillPatch* fp = new FillPatch(*g_tree, *st);
    for (int c = 0; c < ncomp; c++) fp->setProlongation(c, prolong[c]);
    TreeStateData tsd;
    if (dir < 0) {
        BCPatch* bc = new BCPatch();
        bc->defineByData(*st);
        for (int c = 0; c < ncomp; c++)
            for (int q = 0; q < SurroundingIndex::NumSurr; q++) {
                SurroundingIndex sidx(q);
                for (int t = 0; t < nphys; t++) {
                    bc->boundaryRegister(c).setBCMap(sidx.ipos(), sidx.jpos(), sidx.kpos(), physbc[t],
                                                     mathbc[c * nphys + t]);
                    bc->boundaryRegister(c).setBCVal(sidx.ipos(), sidx.jpos(), sidx.kpos(), physbc[t],
                                                     bcval[c * nphys + t]);
                }
            }
        fp->setBC(*bc);
        tsd.bc_patch = bc;
    }
    tsd.new_data = st;
    tsd.fill_new_patch = fp;
    g_tree->treeStateData.push_back(tsd);
    g_state_dir = dir;
    g_state_ngrow = ngrow;
    const IndexBox gb = state_box(dir, ngrow);
    std::vector<double> zero((size_t)g_tree->maxBlockNum * gb.stride(), 0.0);
    for (int c = 0; c < ncomp; c++) copy_in_box(*st, c, gb, zero.data(), g_tree->maxBlockNum);
}
// The whole grown box of every component of the tree's current blocks: [comp][block][box].
void ref_state_set_grown(const double* data) {
    TreeData& st = *g_tree->treeStateData.back().new_data;
    const IndexBox gb = state_box(g_state_dir, g_state_ngrow);
    const size_t per = (size_t)g_tree->numBlocks * gb.stride();
    for (int c = 0; c < st.numComp(); c++) copy_in_box(st, c, gb, data + c * per, g_tree->numBlocks);
}
void ref_state_get_grown(double* data) {
    const TreeData& st = *g_tree->treeStateData.back().new_data;
    const IndexBox gb = state_box(g_state_dir, g_state_ngrow);
    const size_t per = (size_t)g_tree->numBlocks * gb.stride();
    for (int c = 0; c < st.numComp(); c++) copy_out_box(st, c, gb, data + c * per, g_tree->numBlocks);
}

void ref_tree_init_new_data() {
    quiet_begin();
    g_tree->initNewBlockDataAfterRegrid();
    quiet_end();
}

// MacSolver::mg_flux_level (SayakaMacOperator.cpp:161-239): phi (valid values, ghosts zero
// before the reference fills them) in, the face flux of direction dir out for every block
// (the flux array starts at zero).
void ref_mac_flux(int m, int dir, const double* phi_valid, double* flux_faces) {
    const int n = g_tree->numBlocks;
    TreeData* phi = TreeData::CreateCellData(*g_tree, 1, 1);
    std::vector<TreeData*> fl = TreeData::CreateFaceDataPArray(*g_tree, 1, 0);
    const IndexBox gb = grown_cell_box(1);
    std::vector<double> zero((size_t)n * gb.stride(), 0.0);
    copy_in_box(*phi, 0, gb, zero.data(), n);
    copy_in_box(*phi, 0, valid_box(), phi_valid, n);
    std::vector<double> fz((size_t)n * face_box(dir).stride(), 0.0);
    copy_in_box(*fl[dir], 0, face_box(dir), fz.data(), n);
    quiet_begin();
    g_mac->mg_flux_level(m, dir, *fl[dir], *phi, 0, 0);
    quiet_end();
    copy_out_box(*fl[dir], 0, face_box(dir), flux_faces, n);
    delete phi;
    TreeData::ReleaseDataPArray(fl);
}
// FillPatch::fillBoundary (SayakaFillPatch.cpp:17-30) on an ncomp-component cell state with
// ngrow ghost layers.  data: [comp][block][grown box], in and out.  Per component: a
// prolongation method (InterpPatch::ProlongMethod, SayakaInterpPatch.h) and a BC register
// mapping each of the nphys physical codes to (math BC, value) on every surrounding position
// (mathbc / bcval: [comp][nphys]).
void ref_fill_state(int ncomp, int ngrow, double* data, const int* prolong, int nphys, const int* physbc,
                    const int* mathbc, const double* bcval, int scomp, int nc, int nlayer) {
    const int n = g_tree->numBlocks;
    TreeData* st = TreeData::CreateCellData(*g_tree, ncomp, ngrow);
    const IndexBox gb = grown_cell_box(ngrow);
    const size_t per = (size_t)n * gb.stride();
    for (int c = 0; c < ncomp; c++) copy_in_box(*st, c, gb, data + c * per, n);
    BCPatch bc;
    bc.defineByData(*st);
    for (int c = 0; c < ncomp; c++)
        for (int q = 0; q < SurroundingIndex::NumSurr; q++) {
            SurroundingIndex s(q);
            for (int t = 0; t < nphys; t++) {
                bc.boundaryRegister(c).setBCMap(s.ipos(), s.jpos(), s.kpos(), physbc[t], mathbc[c * nphys + t]);
                bc.boundaryRegister(c).setBCVal(s.ipos(), s.jpos(), s.kpos(), physbc[t], bcval[c * nphys + t]);
            }
        }
    FillPatch fp(*g_tree, *st);
    fp.setBC(bc);
    for (int c = 0; c < ncomp; c++) fp.setProlongation(c, prolong[c]);
    quiet_begin();
    fp.fillBoundary(scomp, nc, nlayer);
    quiet_end();
    for (int c = 0; c < ncomp; c++) copy_out_box(*st, c, gb, data + c * per, n);
    delete st;
}

// FillPatch::fillBoundary on an ncomp-component face state of direction dir (staggered in dir)
// with ngrow ghost layers: data [comp][block][grown face box], in and out; prolong per component.
// No BCPatch: the reference fills face data only on domains without walls (fillBlockBndryCond
// exits for face data next to a wall), so the tree must be periodic.
void ref_fill_face_state(int dir, int ncomp, int ngrow, double* data, const int* prolong, int scomp, int nc,
                         int nlayer) {
    const int n = g_tree->numBlocks;
    TreeData* st = TreeData::CreateFaceData(*g_tree, dir, ncomp, ngrow);
    IndexBox gb = grown_cell_box(ngrow);
    gb.staggerInDir(dir);
    const size_t per = (size_t)n * gb.stride();
    for (int c = 0; c < ncomp; c++) copy_in_box(*st, c, gb, data + c * per, n);
    FillPatch fp(*g_tree, *st);
    for (int c = 0; c < ncomp; c++) fp.setProlongation(c, prolong[c]);
    quiet_begin();
    fp.fillBoundary(scomp, nc, nlayer);
    quiet_end();
    for (int c = 0; c < ncomp; c++) copy_out_box(*st, c, gb, data + c * per, n);
    delete st;
}

void ref_restrict(int c, int comp) { g_mac->mg_restict_wgt(c, g_mac->getDataBuf(), comp); }
void ref_prolong(int l, int comp) { g_mac->mg_prolong_wgt(l, g_mac->getDataBuf(), comp); }
double ref_dot(int m, int a, int b) {
    TreeData& d = g_mac->getDataBuf();
    return g_mac->mg_dotprod_level(m, d, d, a, b);
}
double ref_norm(int m, int a) { return g_mac->mg_norm_level(m, g_mac->getDataBuf(), a); }
int ref_precond(int m, int solcomp, int rhscomp) {
    TreeData& d = g_mac->getDataBuf();
    quiet_begin();
    int r = g_mac->mg_precond_level(m, d, d, solcomp, rhscomp);
    quiet_end();
    return r;
}
// cg_solve_level on one MG level with an explicit parameter set; the iteration
// count comes from the verbose log as in ref_mac_cg_solve.
int ref_cg_level(int m, int solcomp, int rhscomp, double tol_rel, double tol_abs, int max_iter,
                 int precond, int* iters, int* converged) {
    char* buf = nullptr;
    size_t len = 0;
    FILE* mem = open_memstream(&buf, &len);
    FILE* old = CLogPrintf::ms_fp;
    CLogPrintf::ms_fp = mem;
    const int vsave = g_mac->Verbose();
    g_mac->setVerbose(1);
    TreeData& d = g_mac->getDataBuf();
    quiet_begin();
    int stat = g_mac->cg_solve_level(m, d, d, solcomp, rhscomp, tol_rel, tol_abs, max_iter, precond,
                                     MacSolver::BCMODE_HOMO);
    quiet_end();
    g_mac->setVerbose(vsave);
    CLogPrintf::ms_fp = old;
    fclose(mem);
    int it = 0, conv = 0;
    const char* p = buf;
    while (p && *p) {
        const char* e = strchr(p, '\n');
        std::string line(p, e ? (size_t)(e - p) : strlen(p));
        int a, b;
        if (sscanf(line.c_str(), "cg_solve_level: iter=%d/%d", &a, &b) == 2) it = a;
        else if (line.find("cg_solve_level: converge") != std::string::npos) conv = 1;
        p = e ? e + 1 : nullptr;
    }
    free(buf);
    if (iters) *iters = it;
    if (converged) *converged = conv;
    return stat;
}

}  // extern "C"
