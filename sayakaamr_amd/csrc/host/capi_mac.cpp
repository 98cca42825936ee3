// capi_mac.cpp -- extern "C" entry points of the host MacSolver (for ctypes and other FFIs).
// Exceptions never cross the ABI: they become a negative return and a message.
#include <exception>
#include <stdexcept>
#include <new>
#include <string>

#include "bc_register.h"
#include "mac_solver.h"
#include "sayaka_host.h"

using namespace sayaka;

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

MacSolver* M(sy_mac* h) { return reinterpret_cast<MacSolver*>(h); }
}  // namespace

extern "C" {

const char* sy_mac_last_error(void) { return g_err.c_str(); }

sy_mac* sy_mac_create(sy_tree* tree, int device) {
    if (!tree) {
        g_err = "null tree";
        return nullptr;
    }
    try {
        return reinterpret_cast<sy_mac*>(new MacSolver(*reinterpret_cast<AmrTree*>(tree), device));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void sy_mac_destroy(sy_mac* h) { delete M(h); }

int sy_mac_prepare(sy_mac* h) {
    return guarded([&] {
        M(h)->prepareSolver();
        return 0;
    });
}

int sy_mac_set_bc(sy_mac* h, int n, const int* phys_bc, const int* math_bc) {
    return guarded([&] {
        BCRegister reg;
        for (int q = 0; q < n; q++) reg.setBCMapAllFaces(phys_bc[q], math_bc[q]);
        M(h)->setBC(reg);
        return 0;
    });
}

int sy_mac_set_ib_ls(sy_mac* h, const double* ls_grown, int ngrow) {
    return guarded([&] {
        M(h)->setIBFracByLSPacked(ls_grown, ngrow);
        return 0;
    });
}

int sy_mac_set_ib_frac(sy_mac* h, const double* vf, const double* const* af) {
    return guarded([&] {
        M(h)->setIBFracPacked(vf, af);
        return 0;
    });
}

int sy_mac_set_coefb(sy_mac* h, const double* const* b) {
    return guarded([&] {
        M(h)->setCoefBPacked(b);
        return 0;
    });
}

int sy_mac_set_rhs(sy_mac* h, const double* rhs) {
    return guarded([&] {
        M(h)->setRhsPacked(rhs);
        return 0;
    });
}

int sy_mac_set_init_sol(sy_mac* h, const double* sol) {
    return guarded([&] {
        M(h)->setInitSolPacked(sol);
        return 0;
    });
}

int sy_mac_fix_bc(sy_mac* h, const double* phi_grown, int ngrow, double* rhs) {
    return guarded([&] {
        if (!phi_grown || !rhs) throw std::invalid_argument("sy_mac_fix_bc: null array");
        M(h)->fixBCPacked(phi_grown, ngrow, rhs);
        return 0;
    });
}

int sy_mac_set_zero_sol(sy_mac* h) {
    return guarded([&] {
        M(h)->setZeroSol();
        return 0;
    });
}

int sy_mac_get_sol(sy_mac* h, double* sol) {
    return guarded([&] {
        M(h)->getSolPacked(sol);
        return 0;
    });
}

int sy_mac_flux(sy_mac* h, int mg_level, int dir, const double* phi, double* flux) {
    return guarded([&] {
        if (!phi || !flux) throw std::invalid_argument("sy_mac_flux: null array");
        M(h)->fluxPacked(mg_level, dir, phi, flux);
        return 0;
    });
}

int sy_mac_set_params(sy_mac* h, double tol_rel, double tol_abs, int max_iter, int precond, int precision,
                      int pre_smooth, int post_smooth) {
    return guarded([&] {
        if (max_iter < 1 || tol_rel < 0 || tol_abs < 0 || (precision < 0 || precision > 2))
            throw std::invalid_argument("bad solver parameters");
        MacSolver& m = *M(h);
        m.setTolerance(tol_rel, tol_abs);
        m.setMaxIteration(max_iter);
        m.setCGUseMGPrecond(precond);
        m.setPrecision(static_cast<MacSolver::Precision>(precision));
        m.setPreSmoothNum(pre_smooth);
        m.setPostSmoothNum(post_smooth);
        return 0;
    });
}

int sy_mac_cg_solve(sy_mac* h, int* iters, int* status, double* rnorm0, double* rnorm) {
    return guarded([&] {
        const int r = M(h)->cg_solve();
        const SolveStats& s = M(h)->lastStats();
        if (iters) *iters = s.iters;
        if (status) *status = s.status;
        if (rnorm0) *rnorm0 = s.rnorm0;
        if (rnorm) *rnorm = s.rnorm;
        return r;
    });
}

int sy_mac_mg_solve(sy_mac* h, int fixed_iter, int* iters, int* status, double* rnorm0, double* rnorm) {
    return guarded([&] {
        M(h)->setMGFixedIter(fixed_iter);
        const int r = M(h)->mg_solve();
        const SolveStats& s = M(h)->lastStats();
        if (iters) *iters = s.iters;
        if (status) *status = s.status;
        if (rnorm0) *rnorm0 = s.rnorm0;
        if (rnorm) *rnorm = s.rnorm;
        return r;
    });
}

int sy_mac_relax_solve(sy_mac* h, int* iters, int* status, double* rnorm0, double* rnorm) {
    return guarded([&] {
        const int r = M(h)->relax_solve();
        const SolveStats& s = M(h)->lastStats();
        if (iters) *iters = s.iters;
        if (status) *status = s.status;
        if (rnorm0) *rnorm0 = s.rnorm0;
        if (rnorm) *rnorm = s.rnorm;
        return r;
    });
}

}  // extern "C"
