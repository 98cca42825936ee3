// c1_solve.cpp -- a driver in the style of the reference's projection demo (Test7.h:411-418,
// 612-653), written only against the AmrTree / TreeData / BCRegister / MacSolver headers:
// a 3D octree refined around a sphere, cut-cell volume / area fractions of the sphere
// (sub-sampled, 4 points per cell edge), variable coefficient 1/rho (density ratio 1000),
// Neumann walls, right-hand side = divergence of a seeded random face velocity (the
// projection step's source) with zero mean, then the plain FP64 solve and the
// mixed-precision solve on the GPU.  (With the binary fractions of setIBFracByLS the
// reference's MG-BiCGSTAB stagnates near 1e-7 relative residual on this problem; the
// fractional ones reach 1e-10.)  With an output path, the FP64 solution is written as a
// Chombo HDF5 file (WriteDataHdf5, as the reference's demos write theirs), and beside it
// <path>.blocks (per block: level as a double, lo[3], hi[3], blockLength[3]) and <path>.phi
// (per block its valid cells, TreeData::packValid order) for checking the file; then the
// solution is copied into a state with 2 ghost layers and filled with FillPatch (centred-linear
// prolongation, Neumann walls on every surrounding, as the demos fill their states,
// Test5.h:194-214), written as <path>.ghost (per block its grown box).
//
//   sayaka_c1_solve [maxlevel=4] [seed=0] [output.h5]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "amr_tree.h"
#include "bc_register.h"
#include "mac_solver.h"
#include "fill_patch.h"
#include "tree_data.h"
#include "writer.h"

using namespace sayaka;

static const double C[3] = {0.5, 0.5, 0.5};
static const double R0 = 0.2;

static const int NSUB = 4;
static const double VF_MIN = 0.02;

// fraction of the cell with lower corner (x0, y0, z0) and edge dx outside the sphere
static double cell_fluid(double x0, double y0, double z0, double dx) {
    int out = 0;
    for (int c = 0; c < NSUB; c++)
        for (int b = 0; b < NSUB; b++)
            for (int a = 0; a < NSUB; a++) {
                const double x = x0 + (a + 0.5) / NSUB * dx, y = y0 + (b + 0.5) / NSUB * dx, z = z0 + (c + 0.5) / NSUB * dx;
                out += (x - C[0]) * (x - C[0]) + (y - C[1]) * (y - C[1]) + (z - C[2]) * (z - C[2]) > R0 * R0;
            }
    const double v = (double)out / (NSUB * NSUB * NSUB);
    return v < VF_MIN ? 0.0 : v;
}

// seeded face velocity U(-1,1) keyed on (seed, level, global face index, dir): one value per
// face of a level whatever the block numbering (splitmix64 of the key)
static double face_u(unsigned seed, int level, int d, long long ii, long long jj, long long kk) {
    uint64_t key = (uint64_t)seed * 0x100000001B3ull ^ ((uint64_t)level << 56) ^ ((uint64_t)d << 61) ^
                   ((uint64_t)kk << 36) ^ ((uint64_t)jj << 18) ^ (uint64_t)ii;
    uint64_t z = key + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

static double sphere_ls(double x, double y, double z) {
    return std::sqrt((x - C[0]) * (x - C[0]) + (y - C[1]) * (y - C[1]) + (z - C[2]) * (z - C[2])) - R0;
}

// refine blocks of `level` that have a cell centre within half a block length of the shell
static void refine_level(AmrTree& t, int level) {
    const int nb = t.nb;
    std::fill(t.refineFlag.begin(), t.refineFlag.end(), 0);
    for (int b = 0; b < t.numBlocks; b++) {
        const AmrTreeNode& n = t[b];
        if (n.refineLevel != level || !(n.nodeType == NodeIsLeaf || n.nodeType == NodeHasLeafChild)) continue;
        if (level >= t.maxAmrLevel) continue;
        const double dx = n.blockLength(0) / nb, half = 0.5 * n.blockLength(0);
        bool hit = false;
        for (int k = 0; k < nb && !hit; k++)
            for (int j = 0; j < nb && !hit; j++)
                for (int i = 0; i < nb && !hit; i++) {
                    const double x = n.boundBox.vlo(0) + (i + 0.5) * dx, y = n.boundBox.vlo(1) + (j + 0.5) * dx,
                                 z = n.boundBox.vlo(2) + (k + 0.5) * dx;
                    hit = std::fabs(sphere_ls(x, y, z)) < half;
                }
        t.refineFlag[b] = hit ? 1 : 0;
    }
    if (t.regrid() != 0) {
        std::fprintf(stderr, "regrid: block pool overflow\n");
        std::exit(1);
    }
}

static int write_filled(const AmrTree& tree, const TreeData& phi, const char* path, int wall) {
    TreeData st = TreeData::MakeCellData(tree, 1, 2);
    const int nb = tree.nb;
    for (int b = 0; b < tree.numBlocks; b++)
        for (int k = 0; k < nb; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) st(b, i, j, k, 0) = phi(b, i, j, k, 0);
    BCPatch bc;
    bc.defineByData(st);
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) bc.boundaryRegister(0).setBCMap(i, j, k, wall, Neumann);
    FillPatch fp(tree, st);
    fp.setBC(bc);
    fp.setProlongation(0, InterpPatch::PROLONG_CENTERED_LINEAR);
    fp.fillBoundary(0, 1, 2);
    std::vector<double> grown((size_t)tree.numBlocks * st.grownSize());
    st.packGrown(0, tree.numBlocks, grown.data());
    const std::string p = std::string(path) + ".ghost";
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f || std::fwrite(grown.data(), sizeof(double), grown.size(), f) != grown.size()) {
        std::fprintf(stderr, "cannot write %s\n", p.c_str());
        if (f) std::fclose(f);
        return -1;
    }
    std::fclose(f);
    return 0;
}

static int write_output(const AmrTree& tree, const TreeData& phi, const char* path) {
    WriterMeta meta;
    meta.name = {"phi"};
    std::string err;
    if (WriteDataHdf5(phi, path, meta, &err) != 0) {
        std::fprintf(stderr, "WriteDataHdf5: %s\n", err.c_str());
        return -1;
    }
    std::vector<double> blocks, cells((size_t)tree.numBlocks * tree.cellsPerBlock());
    for (int b = 0; b < tree.numBlocks; b++) {
        blocks.push_back(tree[b].refineLevel);
        for (int d = 0; d < 3; d++) blocks.push_back(tree[b].boundBox.vlo(d));
        for (int d = 0; d < 3; d++) blocks.push_back(tree[b].boundBox.vhi(d));
        for (int d = 0; d < 3; d++) blocks.push_back(tree[b].blockLength(d));
    }
    phi.packValid(0, tree.numBlocks, cells.data());
    const std::string p(path);
    for (const auto& [suffix, v] : {std::make_pair(std::string(".blocks"), &blocks), std::make_pair(std::string(".phi"), &cells)}) {
        FILE* f = std::fopen((p + suffix).c_str(), "wb");
        if (!f || std::fwrite(v->data(), sizeof(double), v->size(), f) != v->size()) {
            std::fprintf(stderr, "cannot write %s%s\n", path, suffix.c_str());
            if (f) std::fclose(f);
            return -1;
        }
        std::fclose(f);
    }
    std::printf("wrote %s\n", path);
    return 0;
}

int main(int argc, char** argv) {
    const int maxlevel = argc > 1 ? std::atoi(argv[1]) : 4;
    const unsigned seed = argc > 2 ? (unsigned)std::atoi(argv[2]) : 0u;
    const int WALL = -100;

    AmrTree tree;
    tree.ndim = 3;
    tree.nb = 8;
    tree.blockNumGrow = 2;
    tree.minAmrLevel = 2;
    tree.maxAmrLevel = maxlevel;
    tree.maxBlockNum = 8000;
    tree.init();
    const int ndiv[3] = {4, 4, 4};
    RealBox box;
    box.vlo = Vector3d::VecMake(0, 0, 0);
    box.vhi = Vector3d::VecMake(1, 1, 1);
    const int bct[6] = {WALL, WALL, WALL, WALL, WALL, WALL};
    const bool per[6] = {false, false, false, false, false, false};
    tree.initUniformRootLevel(ndiv, box, bct, per);
    tree.initRefineToMinLevel();
    for (int l = tree.minAmrLevel; l <= tree.maxAmrLevel; l++) refine_level(tree, l);
    int nleaf = 0;
    for (int b = 0; b < tree.numBlocks; b++) nleaf += tree[b].isLeaf();
    std::printf("blocks=%d leaves=%d finest=%d\n", tree.numBlocks, nleaf, tree.currentFinestLevel());
    std::fflush(stdout);

    const int nb = tree.nb;
    const double dxf = 1.0 / (ndiv[0] * nb * (1 << (maxlevel - 1)));
    const double eps = 1.5 * dxf;
    auto heavi = [&](double s) {
        if (s < -eps) return 0.0;
        if (s > eps) return 1.0;
        return 0.5 * (1.0 + s / eps + std::sin(M_PI * s / eps) / M_PI);
    };

    // fractions, face coefficients B = 1/rho, seeded rhs
    TreeData vf = TreeData::MakeCellData(tree, 1, 0);
    std::vector<TreeData> bface, aface;
    for (int d = 0; d < 3; d++) bface.push_back(TreeData::MakeFaceData(tree, d, 1, 0));
    for (int d = 0; d < 3; d++) aface.push_back(TreeData::MakeFaceData(tree, d, 1, 0));
    TreeData rhs = TreeData::MakeCellData(tree, 1, 0);
    TreeData phi = TreeData::MakeCellData(tree, 1, 0);
    double sum = 0.0, vol = 0.0;
    for (int b = 0; b < tree.numBlocks; b++) {
        const AmrTreeNode& n = tree[b];
        const double dx = n.blockLength(0) / nb;
        const double x0 = n.boundBox.vlo(0), y0 = n.boundBox.vlo(1), z0 = n.boundBox.vlo(2);
        for (int k = 0; k < nb; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) vf(b, i, j, k, 0) = cell_fluid(x0 + i * dx, y0 + j * dx, z0 + k * dx, dx);
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < nb + (d == 2); k++)
                for (int j = 0; j < nb + (d == 1); j++)
                    for (int i = 0; i < nb + (d == 0); i++) {
                        const double x = x0 + (i + (d == 0 ? 0.0 : 0.5)) * dx;
                        const double y = y0 + (j + (d == 1 ? 0.0 : 0.5)) * dx;
                        const double z = z0 + (k + (d == 2 ? 0.0 : 0.5)) * dx;
                        const double rho = 1.0 + 999.0 * heavi(-sphere_ls(x, y, z));
                        bface[d](b, i, j, k, 0) = 1.0 / rho;
                        // open fraction of the face, closed next to a solid cell on either side
                        int out = 0;
                        for (int q = 0; q < NSUB; q++)
                            for (int p = 0; p < NSUB; p++) {
                                double c[3] = {x0 + i * dx, y0 + j * dx, z0 + k * dx};
                                const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                                c[t1] += (p + 0.5) / NSUB * dx;
                                c[t2] += (q + 0.5) / NSUB * dx;
                                out += (c[0] - C[0]) * (c[0] - C[0]) + (c[1] - C[1]) * (c[1] - C[1]) +
                                           (c[2] - C[2]) * (c[2] - C[2]) > R0 * R0;
                            }
                        double lo3[3] = {x0 + i * dx, y0 + j * dx, z0 + k * dx};
                        double hi3[3] = {lo3[0], lo3[1], lo3[2]};
                        lo3[d] -= dx;
                        const bool closed = cell_fluid(lo3[0], lo3[1], lo3[2], dx) == 0.0 ||
                                            cell_fluid(hi3[0], hi3[1], hi3[2], dx) == 0.0;
                        aface[d](b, i, j, k, 0) = closed ? 0.0 : (double)out / (NSUB * NSUB);
                    }
        // rhs = sum_d (a u ds)|hi - (a u ds)|lo / (dv vf) on fluid cells
        const int lvl = n.refineLevel;
        const long long scale = (1ll << (lvl - 1)) * nb;
        const long long org[3] = {std::llround(x0 / (0.25 / scale)), std::llround(y0 / (0.25 / scale)),
                                  std::llround(z0 / (0.25 / scale))};
        const double dv = dx * dx * dx, ds = dx * dx;
        for (int k = 0; k < nb; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++) {
                    const double f = vf(b, i, j, k, 0);
                    double div = 0.0;
                    for (int d = 0; d < 3; d++) {
                        const int di = d == 0, dj = d == 1, dk = d == 2;
                        const double uh = face_u(seed, lvl, d, org[0] + i + di, org[1] + j + dj, org[2] + k + dk);
                        const double ul = face_u(seed, lvl, d, org[0] + i, org[1] + j, org[2] + k);
                        div += aface[d](b, i + di, j + dj, k + dk, 0) * uh * ds - aface[d](b, i, j, k, 0) * ul * ds;
                    }
                    const double v = f > 0 ? div / (dv * f) : 0.0;
                    rhs(b, i, j, k, 0) = v;
                    if (n.isLeaf() && f > 0) {
                        sum += v * f * dv;
                        vol += f * dv;
                    }
                }
    }
    for (int b = 0; b < tree.numBlocks; b++)
        for (int k = 0; k < nb; k++)
            for (int j = 0; j < nb; j++)
                for (int i = 0; i < nb; i++)
                    if (vf(b, i, j, k, 0) > 0) rhs(b, i, j, k, 0) -= sum / vol;

    BCRegister bcreg;
    bcreg.setBCMapAllFaces(WALL, Neumann);
    MacSolver mac(tree);
    mac.setVerbose(1);
    mac.prepareSolver();
    mac.setBC(bcreg);
    std::vector<const TreeData*> aptr = {&aface[0], &aface[1], &aface[2]};
    mac.setIBFrac(vf, 0, aptr, 0);
    std::vector<TreeData*> bptr = {&bface[0], &bface[1], &bface[2]};
    mac.setCoefB(bptr, 0);
    mac.setRhs(rhs, 0);
    mac.setTolerance(1e-10, 0.0);
    mac.setMaxIteration(2000);

    double ms[2];
    int iters[2], status[2];
    std::vector<std::vector<double>> sols(2);
    std::vector<double> wts;
    for (int pass = 0; pass < 2; pass++) {
        mac.setPrecision(pass == 0 ? MacSolver::FP64 : MacSolver::Mixed);
        mac.setZeroSol();
        mac.cg_solve();   // warm-up
        mac.setZeroSol();
        const auto t0 = std::chrono::steady_clock::now();
        const int ret = mac.cg_solve();
        const auto t1 = std::chrono::steady_clock::now();
        ms[pass] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        iters[pass] = mac.lastStats().iters;
        status[pass] = mac.lastStats().status;
        if (ret != 0) std::printf("cg_solve returned %d\n", ret);
        mac.getSol(phi, 0);
        if (pass == 0 && argc > 3 && (write_output(tree, phi, argv[3]) != 0 || write_filled(tree, phi, argv[3], WALL) != 0))
            return 3;
        auto& s = sols[pass];
        wts.clear();
        for (int b = 0; b < tree.numBlocks; b++)
            if (tree[b].isLeaf())
                for (int k = 0; k < nb; k++)
                    for (int j = 0; j < nb; j++)
                        for (int i = 0; i < nb; i++)
                            if (vf(b, i, j, k, 0) > 0) {   // fluid cells: the solid ones are outside the system
                                s.push_back(phi(b, i, j, k, 0));
                                wts.push_back(tree.getBlockCellVolume(b) * vf(b, i, j, k, 0));
                            }
    }
    // gauge: the Neumann problem fixes phi up to a constant -> compare solutions less their
    // volume-weighted fluid means
    double m0 = 0, m1 = 0, wsum = 0;
    for (size_t q = 0; q < sols[0].size(); q++) {
        m0 += wts[q] * sols[0][q];
        m1 += wts[q] * sols[1][q];
        wsum += wts[q];
    }
    m0 /= wsum;
    m1 /= wsum;
    double num = 0, den = 0;
    for (size_t q = 0; q < sols[0].size(); q++) {
        const double a = sols[0][q] - m0, c = sols[1][q] - m1;
        num += (a - c) * (a - c);
        den += a * a;
    }
    std::printf("fp64_status=%d fp64_iters=%d fp64_ms=%.3f\n", status[0], iters[0], ms[0]);
    std::printf("mixed_status=%d mixed_iters=%d mixed_ms=%.3f\n", status[1], iters[1], ms[1]);
    std::printf("rel_l2_mixed_vs_fp64=%.3e\n", std::sqrt(num / den));
    return (status[0] == 0 && status[1] == 0) ? 0 : 2;
}
