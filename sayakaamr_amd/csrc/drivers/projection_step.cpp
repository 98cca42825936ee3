// projection_step.cpp -- one pressure projection of a face velocity, written the way the
// reference's flow demo writes it: tree set-up by hand (Test7.h:285-366), refinement by flags
// (368-394), data, BC and fill set-up (396-480), then init_solver / calc_pres_rhs / solve_pres /
// corr_vel (411-418, 580-653).  It uses the reference's call surface only (the Sayaka*.h
// headers, namespace sayaka), so the same file compiles against the reference's own sources and
// against this build's include/sayaka; each build writes what it computed -- the pressure, the
// corrected velocity and the pressure-gradient flux, every block's grown box -- to the file
// named on the command line, and the two files can be compared value for value.
//
//     projection_step OUT [maxlevel]
//
// Differences from the demo: the velocity starts from a seeded field instead of diffusion and
// forcing steps (each face takes the average of a hash over the finest-level faces it covers,
// so the two copies of a shared face agree and coarse faces are the average of fine ones; wall
// faces are 0), the elbow has no periodic wrap (the demo's commented-out variant), and the
// velocity's own ghost fill after the correction (fill_vel_all) is left out.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <tuple>
#include <vector>

#include "SayakaBoundaryPatch.h"
#include "SayakaCommons.h"
#include "SayakaFillPatch.h"
#include "SayakaMacSolver.h"
#include "SayakaTree.h"
#include "SayakaTreeData.h"

using namespace sayaka;

namespace {

// SplitMix64 finaliser: a well-mixed 64-bit value of a counter
uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// U(-1, 1) value of the finest-level face (dir, gi, gj, gk)
double hash_face(int dir, long gi, long gj, long gk) {
    const uint64_t key = ((uint64_t)(gi + 4096) << 42) ^ ((uint64_t)(gj + 4096) << 28) ^
                         ((uint64_t)(gk + 4096) << 14) ^ (uint64_t)dir;
    return (double)(mix64(key) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

}  // namespace

class ProjectionStep {
   protected:
    enum { Exterior = -100 };

    double m_dt = 1.0e-3;
    std::unique_ptr<AmrTree> m_tree;
    int m_min_level = 2;
    int m_max_level = 4;
    int m_finest_level = 0;
    int m_ng = 2;
    IndexBox m_valid_box;
    IndexBox m_grown_box;

    TreeData m_pres;
    TreeData m_vel[NDIM];
    TreeData m_gradp[NDIM];
    TreeData m_rhs;

    std::unique_ptr<BCPatch> m_bc_pres;
    std::unique_ptr<FillPatch> m_fill_pres;
    std::unique_ptr<MacSolver> m_solver;
    MGLevelTower* m_tower = nullptr;

   public:
    int run(const char* out, int maxlevel) {
        m_max_level = maxlevel;
        const int nbx = 4, nby = 4, nbz = 4;
        m_valid_box = IndexBox({0, 0, 0}, {nbx - 1, nby - 1, nbz - 1});
        m_grown_box = IndexBox(m_valid_box).extend(m_ng);

        init_tree();
        refine_tree();
        init_data();
        init_solver();
        init_velocity();

        calc_pres_rhs();
        solve_pres();
        corr_vel();
        return write(out);
    }

   protected:
    void init_tree() {
        m_tree.reset(new AmrTree());
        m_tree->minAmrLevel = m_min_level;
        m_tree->maxAmrLevel = m_max_level;
        m_tree->blockValidIndex = m_valid_box;
        m_tree->blockGrownIndex = m_grown_box;
        m_tree->blockNumGrow = m_ng;
        m_tree->init();

        // the demo's elbow of seven unit blocks, walls all round
        const int root_level = 1;
        const int nroot = 7;
        const double xlo[][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {2, 1, 0}, {3, 1, 0}, {3, 0, 0}, {4, 0, 0}};
        const double xhi[][3] = {{1, 1, 1}, {2, 1, 1}, {2, 2, 1}, {3, 2, 1}, {4, 2, 1}, {4, 1, 1}, {5, 1, 1}};
        const int conn[][6] = {
            {-100, 1, -100, -100, -100, -100}, {0, -100, -100, 2, -100, -100}, {-100, 3, 1, -100, -100, -100},
            {2, 4, -100, -100, -100, -100},    {3, -100, 5, -100, -100, -100}, {-100, 6, -100, 4, -100, -100},
            {5, -100, -100, -100, -100, -100},
        };
        m_tree->numBlocks = nroot;
        for (int iroot = 0; iroot < nroot; iroot++) {
            auto& root = m_tree->blocks[iroot];
            root.setBound(xlo[iroot], xhi[iroot]);
            root.nodeType = TreeNodeType_IsLeaf;
            root.setLevel(root_level);
            for (FaceIndex iface = 0; iface < FaceIndex::NumFace; ++iface) root.neighbor[iface] = conn[iroot][iface];
        }
    }

    void refine_tree() {
        m_tree->initRefineToMinLevel();
        for (int ilevel = m_min_level; ilevel <= m_max_level; ilevel++) {
            for (int iblock = 0; iblock < m_tree->numBlocks; iblock++)
                if ((*m_tree)[iblock].getLevel() == ilevel) test_block_refine(iblock);
            m_tree->regrid();
        }
        m_finest_level = m_tree->currentFinestLevel();
        std::printf("projection_step: %d blocks, finest level %d\n", m_tree->numBlocks, m_finest_level);
    }

    // the demo's refinement rule (Test7.h:422-470): cells within two cell widths of the lines
    // x = 1, 2, 3, 4 at y = 1
    void test_block_refine(int iblock) {
        auto& block = (*m_tree)[iblock];
        const int level = block.getLevel();
        const IndexBox& validbox = m_tree->validBlockCellBox();
        if (block.isLeaf() || block.isParentOfLeafChild()) {
            const Vector3d blo = block.boundBox.lo();
            const Vector3d dh = m_tree->getBlockCellSize(iblock);
            const double dx = dh.x;
            int need_refine = 0;
            BEGIN_FOR_BOX_RANGE(validbox, i, j, k);
            {
                const double y = blo.y + ((double)j + 0.5) * dh.y;
                const double x = blo.x + ((double)i + 0.5) * dh.x;
                for (int c = 1; c <= 4; c++)
                    if (std::hypot(x - c, y - 1) < dx * 2) need_refine = 1;
            }
            END_FOR_BOX_RANGE;
            if (level < m_tree->maxRefineLevel() && need_refine) m_tree->refineFlag[iblock] = 1;
        }
    }

    void init_data() {
        const int ng = m_ng;
        m_pres = TreeData::MakeCellData(*m_tree, 1, ng);
        m_pres.setValue(0);
        for (int dir = 0; dir < NDIM; dir++) {
            m_vel[dir] = TreeData::MakeFaceData(*m_tree, dir, 1, ng);
            m_vel[dir].setValue(0);
            m_gradp[dir] = TreeData::MakeFaceData(*m_tree, dir, 1, ng);
            m_gradp[dir].setValue(0);
        }
        m_rhs = TreeData::MakeCellData(*m_tree, 1, 0);
        m_rhs.setValue(0);

        m_bc_pres.reset(new BCPatch(*m_tree, m_pres, m_pres.indexType()));
        {
            BCRegister& bcreg = m_bc_pres->boundaryRegister(0);
            for (SurroundingIndex isurr = 0; isurr < SurroundingIndex::NumSurr; ++isurr)
                bcreg.setBCMap(isurr, Exterior, BCType::Neumann);
        }
        m_fill_pres.reset(new FillPatch(*m_tree, m_pres));
        m_fill_pres->setProlongation(0, InterpPatch::PROLONG_CENTERED_LINEAR);
        m_fill_pres->setRestriction(0, InterpPatch::RESTRICT_AVERAGE);
        m_fill_pres->setBC(*m_bc_pres);
    }

    void init_solver() {
        m_solver.reset(new MacSolver(*m_tree));
        m_solver->setVerbose(1);
        m_tower = &m_solver->getTower(m_finest_level);
    }

    // every valid face of every block: the average of the hash over the finest-level faces it
    // covers (the tangential ones, 2^(finest-level) per direction); 0 on walls
    void init_velocity() {
        for (int iblock = 0; iblock < m_tree->numBlocks; iblock++) {
            auto& block = (*m_tree)[iblock];
            const Vector3d blo = block.boundBox.lo();
            const Vector3d dh = m_tree->getBlockCellSize(iblock);
            const int n = 1 << (m_finest_level - block.getLevel());
            for (int dir = 0; dir < NDIM; dir++) {
                auto& u = m_vel[dir][iblock];
                const IndexBox& b = m_vel[dir].validBox();
                BEGIN_FOR_BOX_RANGE(b, i, j, k);
                {
                    const long gi = std::lround(blo.x / dh.x) + i, gj = std::lround(blo.y / dh.y) + j,
                               gk = std::lround(blo.z / dh.z) + k;
                    const int idx[3] = {i, j, k};
                    const bool wall = (idx[dir] == b.lo(dir) && block.neighbor[2 * dir] == Exterior) ||
                                      (idx[dir] == b.hi(dir) && block.neighbor[2 * dir + 1] == Exterior);
                    double s = 0.0;
                    if (!wall) {
                        for (int q2 = 0; q2 < (NDIM == 3 ? n : 1); q2++)
                            for (int q1 = 0; q1 < n; q1++) {
                                long f[3] = {gi * n, gj * n, gk * n};
                                const int a1 = dir == 0 ? 1 : 0, a2 = dir == 2 ? 1 : 2;
                                f[a1] += q1;
                                f[a2] += q2;
                                s += hash_face(dir, f[0], f[1], f[2]);
                            }
                        s /= (double)(NDIM == 3 ? n * n : n);
                    }
                    u(i, j, k, 0) = s;
                }
                END_FOR_BOX_RANGE;
            }
        }
    }

    // Test7.h:580-604
    void calc_pres_rhs() {
        for (int iblock = 0; iblock < m_tree->numBlocks; iblock++) calc_pres_rhs_block(iblock);
    }

    void calc_pres_rhs_block(int iblock) {
        Vector3d dh = m_tree->getBlockCellSize(iblock);
        const IndexBox& b = m_rhs.validBox();
        auto& d = m_rhs[iblock];
        d.setValue(0);
        const double dt_inv = 1.0 / m_dt;
        for (int dir = 0; dir < NDIM; dir++) {
            int ii, jj, kk;
            std::tie(ii, jj, kk) = sayaka::SelectStaggerIncr(dir);
            BEGIN_FOR_BOX_RANGE(b, i, j, k);
            {
                double ul = m_vel[dir][iblock](i, j, k, 0);
                double ur = m_vel[dir][iblock](i + ii, j + jj, k + kk, 0);
                d(i, j, k, 0) += -dt_inv / dh(dir) * (ur - ul);
            }
            END_FOR_BOX_RANGE;
        }
    }

    // Test7.h:606-634
    void solve_pres() {
        MacSolver* mac = m_solver.get();
        mac->prepareSolver();
        mac->setBC(m_bc_pres->boundaryRegister(0));
        mac->setRhs(m_rhs, 0);
        mac->setZeroSol();
        mac->setTolerance(1.0e-7, 1.0e-8);
        mac->setCGUseMGPrecond(1);
        mac->cg_solve();
        mac->getSol(m_pres, 0);
        mac->clearSolver();
        m_fill_pres->fillBoundary(0, 1, 1);
    }

    // Test7.h:636-653, without the velocity's ghost fill
    void corr_vel() {
        const int finest_level = m_finest_level;
        for (int dir = 0; dir < NDIM; dir++) {
            m_solver->mg_flux_level(finest_level, dir, m_gradp[dir], m_pres, 0, 0);
            auto& tower = m_solver->getTower(finest_level);
            TreeData::AddEqual(m_vel[dir], m_gradp[dir], m_dt, 0, 0, 1, 0, tower.blockIds());
        }
    }

    // header (int32): number of blocks, block edge, dimension, finest level, ghost layers, tower
    // size; then, every block in turn, the grown box of the pressure, of the velocity of each
    // direction and of the flux of each direction (i fastest)
    int write(const char* path) {
        FILE* f = std::fopen(path, "wb");
        if (!f) return 1;
        const int32_t hdr[6] = {m_tree->numBlocks, m_valid_box.size(0), NDIM, m_finest_level, m_ng,
                                (int32_t)m_tower->blockIds().size()};
        std::fwrite(hdr, sizeof(hdr), 1, f);
        std::fwrite(m_tower->blockIds().data(), sizeof(int), m_tower->blockIds().size(), f);
        auto dump = [&](TreeData& td) {
            for (int iblock = 0; iblock < m_tree->numBlocks; iblock++) {
                auto& d = td[iblock];
                const IndexBox& gb = d.box();
                BEGIN_FOR_BOX_RANGE(gb, i, j, k);
                {
                    const double v = d(i, j, k, 0);
                    std::fwrite(&v, sizeof(double), 1, f);
                }
                END_FOR_BOX_RANGE;
            }
        };
        dump(m_pres);
        for (int dir = 0; dir < NDIM; dir++) dump(m_vel[dir]);
        for (int dir = 0; dir < NDIM; dir++) dump(m_gradp[dir]);
        return std::fclose(f) == 0 ? 0 : 1;
    }
};

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT [maxlevel]\n", argv[0]);
        return 2;
    }
    ProjectionStep step;
    return step.run(argv[1], argc > 2 ? std::atoi(argv[2]) : 4);
}
