// Host build of the backend's per-cell kernel bodies (test infrastructure).
//
// The HIP kernels are thin loops over the bodies in capaz_bodies.h; this file
// runs the same bodies in plain loops on the CPU, so the tests can compare
// the kernels' arithmetic with the oracle bit for bit without a GPU.  The
// only difference from the device build is the math library behind pow/exp
// (glibc here, the ROCm device library on the GPU).
#include <cstdint>
#include <vector>

#include "capaz_adjoint.h"
#include "capaz_bodies.h"
#include "capaz_derive.h"

extern "C" {

const char* capaz_emul_validate(const capaz_params* u) {
    static thread_local std::string why;
    why = capaz::validate(*u);
    return why.c_str();
}

// entry: 0 S, 1 P_Xi_ro, 2 u, 3 roS, 4 HE, 5 E, 6 Newton, 7 boundary;
// 8 the energy model's Newton in its two passes (returns the cells redone).
int capaz_emul_entry(const capaz_params* u, double** fields, int entry) {
    if (!capaz::validate(*u).empty()) return -1;
    capaz::DevParams p = capaz::derive(*u);
    capaz::DevFields F;
    for (int f = 0; f < capaz::F_COUNT; ++f) F.f[f] = fields[f];
    if (entry == 8) {  // the energy Newton as the kernels run it: no-pivot pass, then the redo pass
        if (!p.energy) return -1;
        std::vector<unsigned char> redo(p.n);
        int redone = 0;
        for (int64_t l = 0; l < p.n; ++l) {
            int i, j, k;
            capaz::decompose(p, l, i, j, k);
            redo[l] = p.model == 2 ? capaz::interior_newton2(p, i, j, k) && !capaz::body_newton_energy2_nopivot(p, F, l)
                                   : capaz::interior3(p, i, j, k) && !capaz::body_newton_energy_nopivot(p, F, l);
        }
        for (int64_t l = 0; l < p.n; ++l) {
            if (!redo[l]) continue;
            capaz::body_newton(p, F, l);
            ++redone;
        }
        return redone;
    }
    for (int64_t l = 0; l < p.n; ++l) {
        switch (entry) {
            case 0: capaz::body_S(p, F, l); break;
            case 1: capaz::body_P_Xi_ro(p, F, l); break;
            case 2: capaz::body_u(p, F, l); break;
            case 3: capaz::body_roS(p, F, l); break;
            case 4: capaz::body_HE(p, F, l); break;
            case 5: capaz::body_E(p, F, l); break;
            case 6: capaz::body_newton(p, F, l); break;
            case 7: capaz::body_boundary(p, F, l); break;
            default: return -1;
        }
    }
    return 0;
}

// One fused step: cur = {Pw, Sw, Sn, Pn, Pg, m, K}, next = {Pw, Sw, Sn, Pn, Pg}.
int capaz_emul_fused_step(const capaz_params* u, double** cur, double** next) {
    if (!capaz::validate(*u).empty()) return -1;
    capaz::DevParams p = capaz::derive(*u);
    capaz::StepIO io;
    io.Pw = cur[0]; io.Sw = cur[1]; io.Sn = cur[2]; io.Pn = cur[3]; io.Pg = cur[4]; io.m = cur[5]; io.K = cur[6];
    io.Pw1 = next[0]; io.Sw1 = next[1]; io.Sn1 = next[2]; io.Pn1 = next[3]; io.Pg1 = next[4];
    for (int64_t l = 0; l < p.n; ++l) capaz::body_fused_step(p, io, l);
    return 0;
}

// One adjoint step (capaz_adjoint.h) on host arrays:
// state = {Pw, Sw, Sn, Pn, Pg, m, K, dtm} at t, lam1 = lambda_{t+1}
// {Pw, Sw, Sn, Pn, Pg}, work = {mu_w, mu_n, mu_g, dPw, dSw, dSn} (scratch),
// lam = lambda_t (written), grad = {g_lnK, g_m} (accumulated).
int capaz_emul_adjoint_step(const capaz_params* u, double** state, double** lam1, double** work, double** lam,
                            double** grad) {
    if (!capaz::validate(*u).empty()) return -1;
    capaz::DevParams p = capaz::derive(*u);
    capaz::AdjIO a;
    a.Pw = state[0]; a.Sw = state[1]; a.Sn = state[2]; a.Pn = state[3]; a.Pg = state[4];
    a.m = state[5]; a.K = state[6]; a.dtm = state[7];
    a.lPw1 = lam1[0]; a.lSw1 = lam1[1]; a.lSn1 = lam1[2]; a.lPn1 = lam1[3]; a.lPg1 = lam1[4];
    a.mu_w = work[0]; a.mu_n = work[1]; a.mu_g = work[2]; a.dPw = work[3]; a.dSw = work[4]; a.dSn = work[5];
    a.lPw = lam[0]; a.lSw = lam[1]; a.lSn = lam[2]; a.lPn = lam[3]; a.lPg = lam[4];
    a.g_lnK = grad[0]; a.g_m = grad[1];
    if (p.model == 2) {  // two phases: state (P_w, S_n); the other slots are unused
        for (int k = 0; k < p.Nz; ++k)
            for (int j = 0; j < p.Ny; ++j)
                for (int i = 0; i < p.Nx; ++i)
                    if (capaz::interior_newton2(p, i, j, k))
                        capaz::adj2_newton_cell<0>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
        for (int k = 0; k < p.Nz; ++k)
            for (int j = 0; j < p.Ny; ++j)
                for (int i = 0; i < p.Nx; ++i)
                    capaz::adj2_coef_cell(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
        return 0;
    }
    for (int k = 0; k < p.Nz; ++k)
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i)
                if (capaz::interior3(p, i, j, k))
                    capaz::adj_newton_cell<0>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
    for (int k = 0; k < p.Nz; ++k)
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i)
                capaz::adj_coef_cell(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
    return 0;
}

// One tangent-linear step (capaz_adjoint.h) on host arrays: state = {Pw, Sw,
// Sn, Pn, Pg, m, K, dtm} at t, dstate = dx_t {Pw, Sw, Sn, Pn, Pg}, dtheta =
// {dlnK, dm}, dnext = dx_{t+1} (written).
int capaz_emul_tangent_step(const capaz_params* u, double** state, double** dstate, double** dtheta, double** dnext) {
    if (!capaz::validate(*u).empty()) return -1;
    capaz::DevParams p = capaz::derive(*u);
    capaz::TanIO a;
    a.Pw = state[0]; a.Sw = state[1]; a.Sn = state[2]; a.Pn = state[3]; a.Pg = state[4];
    a.m = state[5]; a.K = state[6]; a.dtm = state[7];
    a.dPw = dstate[0]; a.dSw = dstate[1]; a.dSn = dstate[2]; a.dPn = dstate[3]; a.dPg = dstate[4];
    a.dlnK = dtheta[0]; a.dm = dtheta[1];
    a.dPw1 = dnext[0]; a.dSw1 = dnext[1]; a.dSn1 = dnext[2]; a.dPn1 = dnext[3]; a.dPg1 = dnext[4];
    for (int k = 0; k < p.Nz; ++k)
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i)
                if (capaz::interior3(p, i, j, k))
                    capaz::tan_cell<0>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
    for (int k = 0; k < p.Nz; ++k)
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i)
                if (!capaz::interior3(p, i, j, k)) capaz::tan_boundary(p, a, i, j, k);
    return 0;
}

// vg_d (capaz_adjoint.h) at effective saturations (x, y): k = {k_w, k_n, k_g},
// d = {dk_w/dx, dk_n/dx, dk_n/dy, dk_g/dz}.
int capaz_emul_vg_d(const capaz_params* u, double x, double y, double* k, double* d) {
    if (!capaz::validate(*u).empty()) return -1;
    const capaz::DevParams p = capaz::derive(*u);
    capaz::vg_d(p, x, y, k[0], k[1], k[2], d[0], d[1], d[2], d[3]);
    return 0;
}

// coef_d (capaz_adjoint.h) of one cell state x = {Pw, Sw, Sn, Pn, Pg, K}:
// out = {Xi[3], dXi_dSw[3], dXi_dSn[3]}.
int capaz_emul_coef_d(const capaz_params* u, const double* x, int inner, double* out) {
    if (!capaz::validate(*u).empty()) return -1;
    const capaz::DevParams p = capaz::derive(*u);
    capaz::AdjIO a{};
    a.Pw = x; a.Sw = x + 1; a.Sn = x + 2; a.Pn = x + 3; a.Pg = x + 4; a.K = x + 5;
    const capaz::CoefD c = capaz::coef_d(p, a, 0, inner != 0);
    for (int ph = 0; ph < 3; ++ph) {
        out[ph] = c.Xi[ph];
        out[3 + ph] = c.dXi_dSw[ph];
        out[6 + ph] = c.dXi_dSn[ph];
    }
    return 0;
}

int capaz_emul_reverse_matrix(double* a, int n) {
    if (n == -3) return capaz::reverse3(a);  // the no-pivot fast path with fallback
    if (n == -4) return capaz::reverse3<true>(a);  // the same for a[2] == +0
    if (n == 3) return capaz::reverse_matrix<3>(a);
    if (n == 5) return capaz::reverse_matrix<5>(a);
    if (n == -5) return capaz::reverse_nopivot<5>(a) ? 0 : 1;  // 1: refused, `a` untouched
    if (n == 4) return capaz::reverse_matrix<4>(a);
    if (n == -14) return capaz::reverse_nopivot<4>(a) ? 0 : 1;  // the 4x4 no-pivot path
    return -2;
}

}  // extern "C"
