// Sanitizer driver (test infrastructure): the oracle (oracle/capaz_oracle.c)
// and the host build of the kernels' per-cell bodies (capaz_bodies.h) run
// side by side on small, thin and one-cell-wide grids of every model, built
// with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitize.py.  Every neighbour read of the stencils, the boundary
// and the dense solves is then bounds-checked on the CPU, and the two
// implementations must still agree bit for bit (both use glibc's pow/exp).
// Exit status 0 and "sanitize: ok" on success.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capaz_bodies.h"
#include "capaz_derive.h"

extern "C" {
#include "capaz_oracle.h"
}

namespace {

struct Case {
    const char* name;
    int model, energy, closure, nx, ny, nz, source, steps;
};

// The reference decks' values (Two-phase/defines.ini, Three-phase/defines.ini).
capaz_params deck(const Case& c) {
    capaz_params u;
    std::memset(&u, 0, sizeof(u));
    u.struct_size = sizeof(capaz_params);
    u.corey_n = 2.;
    u.model = c.model;
    u.energy = c.energy;
    u.closure = c.closure;
    u.Nx = c.nx;
    u.Ny = c.ny;
    u.Nz = c.nz;
    u.source = c.source;
    u.hx = u.hy = u.hz = 5e-2;
    u.dt = 1e-4;
    u.P_atm = 1e5;
    u.g_const = 9.807;
    if (c.model == 2) {
        u.newton_iterations = 10;
        u.mu_w = 1e-3;
        u.mu_n = 9e-4;
        u.ro0_w = 1000.;
        u.ro0_n = 1460.;
        u.beta_w = 4.4e-7;
        u.beta_n = 1e-6;
        u.lambda = 2.7;
        u.S_wr = 0.09;
        u.P_d = 755.;
        u.S_n_gr = 0.2;
    } else {
        u.newton_iterations = c.energy ? 3 : 7;
        u.mu_w = 1e-3;
        u.mu_n = 1e-3;
        u.mu_g = 1.84e-5;
        u.ro0_w = 1000.;
        u.ro0_n = 850.;
        u.ro0_g = 1.4;
        u.beta_w = 4.4e-7;
        u.beta_n = 1e-6;
        u.lambda = 3.25;
        u.S_w_gr = 0.4;
        u.S_n_gr = 0.3;
    }
    if (c.closure == 1) {
        u.S_wr = 0.1;
        u.S_nr = 0.05;
    }
    return u;
}

oracle_params to_oracle(const capaz_params& u) {
    oracle_params p;
    std::memset(&p, 0, sizeof(p));
    p.model = u.model;
    p.energy = u.energy;
    p.closure = u.closure;
    p.Nx = u.Nx;
    p.Ny = u.Ny;
    p.Nz = u.Nz;
    p.source = u.source;
    p.iterations = u.newton_iterations;
    p.hx = u.hx;
    p.hy = u.hy;
    p.hz = u.hz;
    p.dt = u.dt;
    p.P_atm = u.P_atm;
    p.g_const = u.g_const;
    p.mu_w = u.mu_w;
    p.mu_n = u.mu_n;
    p.mu_g = u.mu_g;
    p.ro0_w = u.ro0_w;
    p.ro0_n = u.ro0_n;
    p.ro0_g = u.ro0_g;
    p.beta_w = u.beta_w;
    p.beta_n = u.beta_n;
    p.lambda = u.lambda;
    p.S_wr = u.S_wr;
    p.S_nr = u.S_nr;
    p.S_gr = u.S_gr;
    p.P_d = u.P_d;
    p.S_w_gr = u.S_w_gr;
    p.S_n_gr = u.S_n_gr;
    p.corey_n = u.corey_n;
    return p;
}

// A reproducible uniform deviate in [0, 1) (64-bit LCG, top 53 bits).
struct Lcg {
    uint64_t s;
    double next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * 0x1p-53;
    }
};

bool has_field(const capaz_params& u, int f) {
    if (f >= capaz::F_T) return u.energy == 1;
    if (f >= capaz::F_S_g) return u.model == 3;
    return true;
}

struct Fields {
    std::vector<std::vector<double>> a;  // one array per field id, empty if absent
    explicit Fields(const capaz_params& u) : a(capaz::F_COUNT) {
        const size_t n = (size_t)u.Nx * u.Ny * u.Nz;
        for (int f = 0; f < capaz::F_COUNT; ++f)
            if (has_field(u, f)) a[f].assign(n, 0.);
    }
    double* p(int f) { return a[f].empty() ? nullptr : a[f].data(); }
};

// Hydrostatic pressure plus noise, saturations inside (0, 1), porosity noise.
void initial_state(const capaz_params& u, Fields& F, uint64_t seed) {
    Lcg r{seed};
    for (int k = 0; k < u.Nz; ++k)
        for (int j = 0; j < u.Ny; ++j)
            for (int i = 0; i < u.Nx; ++i) {
                const size_t l = (size_t)i + (size_t)j * u.Nx + (size_t)k * u.Nx * u.Ny;
                const double P = u.P_atm + 1000. * u.g_const * u.hy * j + 100. * (r.next() - 0.5);
                F.a[capaz::F_P_w][l] = P;
                F.a[capaz::F_m][l] = 0.38 + 0.04 * r.next();
                F.a[capaz::F_K][l] = 6.64e-11;
                if (u.model == 2) {
                    F.a[capaz::F_S_n][l] = 0.05 + 0.35 * r.next();
                } else {
                    F.a[capaz::F_S_w][l] = 0.25 + 0.25 * r.next();
                    F.a[capaz::F_S_n][l] = 0.15 + 0.2 * r.next();
                    F.a[capaz::F_P_n][l] = P + 10. * (r.next() - 0.5);
                    F.a[capaz::F_P_g][l] = P + 10. * (r.next() - 0.5);
                    if (u.energy) F.a[capaz::F_T][l] = 282. + 6. * r.next();
                }
            }
}

oracle_fields oracle_view(Fields& F) {
    oracle_fields o;
    double** slot = reinterpret_cast<double**>(&o);  // members in capaz_field_id order
    static_assert(sizeof(oracle_fields) == capaz::F_COUNT * sizeof(double*), "field order");
    for (int f = 0; f < capaz::F_COUNT; ++f) slot[f] = F.p(f);
    return o;
}

capaz::DevFields body_view(Fields& F) {
    capaz::DevFields d;
    for (int f = 0; f < capaz::F_COUNT; ++f) d.f[f] = F.p(f);
    return d;
}

void body_step(const capaz::DevParams& p, const capaz::DevFields& d) {
    using namespace capaz;
    for (int64_t l = 0; l < p.n; ++l) body_S(p, d, l);
    for (int64_t l = 0; l < p.n; ++l) body_P_Xi_ro(p, d, l);
    for (int64_t l = 0; l < p.n; ++l) body_u(p, d, l);
    for (int64_t l = 0; l < p.n; ++l) body_roS(p, d, l);
    if (p.energy) {
        for (int64_t l = 0; l < p.n; ++l) body_HE(p, d, l);
        for (int64_t l = 0; l < p.n; ++l) body_E(p, d, l);
    }
    for (int64_t l = 0; l < p.n; ++l) body_newton(p, d, l);
    for (int64_t l = 0; l < p.n; ++l) body_boundary(p, d, l);
}

bool same(double a, double b) { return a == b || (std::isnan(a) && std::isnan(b)); }

int compare(const char* what, const Case& c, const capaz_params& u, Fields& A, Fields& B, bool primary_only) {
    int bad = 0;
    for (int f = 0; f < capaz::F_COUNT; ++f) {
        if (!has_field(u, f)) continue;
        if (primary_only && !(f == capaz::F_P_w || f == capaz::F_S_n || (u.model == 3 && f == capaz::F_S_w)))
            continue;
        for (size_t l = 0; l < A.a[f].size(); ++l)
            if (!same(A.a[f][l], B.a[f][l])) {
                std::printf("sanitize: %s %s: field %d differs at cell %zu (%.17g vs %.17g)\n", c.name, what, f, l,
                            A.a[f][l], B.a[f][l]);
                ++bad;
                break;
            }
    }
    return bad;
}

}  // namespace

int main() {
    const Case cases[] = {
        {"2ph 9x8x7", 2, 0, 0, 9, 8, 7, 2, 6},        {"2ph 3x5x1", 2, 0, 0, 3, 5, 1, 1, 6},
        {"2ph corey 6x5x4", 2, 0, 1, 6, 5, 4, 1, 6},  {"3ph 8x9x7", 3, 0, 0, 8, 9, 7, 0, 6},
        {"3ph 1x15x1", 3, 0, 0, 1, 15, 1, 0, 6},      {"3ph corey 7x3x1", 3, 0, 1, 7, 3, 1, 2, 6},
        {"3ph corey 3x3x3", 3, 0, 1, 3, 3, 3, 1, 6},  {"3phe 6x7x5", 3, 1, 0, 6, 7, 5, 0, 4},
        {"3phe 1x9x1", 3, 1, 0, 1, 9, 1, 0, 4},       {"3phe 9x8x1", 3, 1, 0, 9, 8, 1, 2, 4},
        {"3phe 1x12x9", 3, 1, 0, 1, 12, 9, 0, 4},
    };
    int bad = 0;
    uint64_t seed = 1;
    for (const Case& c : cases) {
        const capaz_params u = deck(c);
        const std::string why = capaz::validate(u);
        if (!why.empty()) {
            std::printf("sanitize: %s: %s\n", c.name, why.c_str());
            return 2;
        }
        const capaz::DevParams p = capaz::derive(u);
        const oracle_params op = to_oracle(u);
        Fields O(u), S(u);
        initial_state(u, O, seed);
        initial_state(u, S, seed);
        ++seed;
        oracle_fields of = oracle_view(O);
        const capaz::DevFields sf = body_view(S);
        // entry-point sequence of the bodies against the oracle, every field
        for (int s = 1; s <= c.steps; ++s) {
            oracle_time_step(&op, &of, s * u.dt);
            body_step(p, sf);
        }
        bad += compare("sequence", c, u, O, S, false);
        if (u.energy) continue;
        // the fused step's body (one thread per cell) with ping-pong state
        Fields A(u), B(u);
        initial_state(u, A, seed - 1);
        initial_state(u, B, seed - 1);
        for (int s = 1; s <= c.steps; ++s) {
            capaz::StepIO io;
            io.Pw = A.p(capaz::F_P_w);
            io.Sw = A.p(capaz::F_S_w);
            io.Sn = A.p(capaz::F_S_n);
            io.Pn = A.p(capaz::F_P_n);
            io.Pg = A.p(capaz::F_P_g);
            io.m = A.p(capaz::F_m);
            io.K = A.p(capaz::F_K);
            io.Pw1 = B.p(capaz::F_P_w);
            io.Sw1 = B.p(capaz::F_S_w);
            io.Sn1 = B.p(capaz::F_S_n);
            io.Pn1 = B.p(capaz::F_P_n);
            io.Pg1 = B.p(capaz::F_P_g);
            for (int64_t l = 0; l < p.n; ++l) capaz::body_fused_step(p, io, l);
            std::swap(A.a, B.a);
            for (int f : {capaz::F_m, capaz::F_K}) B.a[f] = A.a[f];
        }
        bad += compare("fused", c, u, O, A, true);
    }
    if (bad) return 1;
    std::printf("sanitize: ok\n");
    return 0;
}
