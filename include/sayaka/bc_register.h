// bc_register.h -- physical-to-mathematical boundary condition map (the caller's BC setup).
//
// Mirrors BCRegister (SayakaBoundaryPatch.h:62-140): for each of the 3^3 surrounding
// positions of a block, a map from the physical boundary id stored in the tree's neighbour
// table (<= -20) to a math BC type (Neumann / Dirichlet / SimpleFill) and a value.  The
// reference exits on an unknown mapping; getBCMap here returns -1 and the solver reports it.
#pragma once

#include <map>

#include "index_box.h"
#include "sayaka_types.h"

namespace sayaka {
inline namespace amd {

struct BCRegister {
    static constexpr int NumSurr = 27;
    static int surr(int i, int j, int k) { return (i + 1) + 3 * (j + 1) + 9 * (k + 1); }

    std::map<int, int> bcmaps[NumSurr];
    std::map<int, double> bcvals[NumSurr];

    void setBCMap(int isurr, int jsurr, int ksurr, int phys_bc, int math_bc) {
        const int s = surr(isurr, jsurr, ksurr);
        bcmaps[s][phys_bc] = math_bc;
        bcvals[s][phys_bc] = 0.0;
    }
    void setBCVal(int isurr, int jsurr, int ksurr, int phys_bc, double bc_val) {
        bcvals[surr(isurr, jsurr, ksurr)][phys_bc] = bc_val;
    }
    // the same by a surrounding position (SayakaBoundaryPatch.h:91-99)
    void setBCMap(const SurroundingIndex& s, int phys_bc, int math_bc) {
        setBCMap(s.ipos(), s.jpos(), s.kpos(), phys_bc, math_bc);
    }
    void setBCVal(const SurroundingIndex& s, int phys_bc, double bc_val) {
        setBCVal(s.ipos(), s.jpos(), s.kpos(), phys_bc, bc_val);
    }
    // the same map on every face position (the common case of the reference's drivers)
    void setBCMapAllFaces(int phys_bc, int math_bc) {
        for (int d = 0; d < 3; d++)
            for (int s = -1; s <= 1; s += 2) {
                int o[3] = {0, 0, 0};
                o[d] = s;
                setBCMap(o[0], o[1], o[2], phys_bc, math_bc);
            }
    }
    int getBCMap(int isurr, int jsurr, int ksurr, int phys_bc) const {
        const auto& m = bcmaps[surr(isurr, jsurr, ksurr)];
        auto it = m.find(phys_bc);
        return it == m.end() ? -1 : it->second;
    }
    double getBCVal(int isurr, int jsurr, int ksurr, int phys_bc) const {
        const auto& m = bcvals[surr(isurr, jsurr, ksurr)];
        auto it = m.find(phys_bc);
        return it == m.end() ? 0.0 : it->second;
    }
};

}  // namespace amd
}  // namespace sayaka
