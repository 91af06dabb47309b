// The fused time-step kernel for gfx950: one pass over the primary state per
// time step (two- and three-phase models without the energy equation).
//
// Mapping.  A workgroup owns a TX x TY block of interior columns (x fastest,
// so each wavefront row streams TX consecutive doubles of every field) and a
// chunk of planes along z, which it walks plane by plane.  For the current
// plane the closure coefficients of each cell -- phase pressures, Xi and
// densities, the values the upwind flux reads from a neighbour -- are computed
// once and staged in LDS together with a one-cell halo ring, so the x and y
// neighbours come from LDS; the z neighbours stay in registers from the
// previous and next plane, and the z face differences (P_c - P_zm)/h are
// carried over from the previous plane (the same bits as recomputing them).
// After the divergence each thread solves its cell's Newton system in
// registers and writes the new state.  The boundary cells are written
// afterwards by k_fused_faces from their inward neighbours' new state
// (Border_S/Border_P, two-phase.cpp:163-215, three-phase.cpp:345-427): an
// O(N^2) pass that keeps the face logic out of the interior kernel's
// registers.
//
// Arithmetic is that of capaz_cell.h / capaz_bodies.h, so the kernel gives
// the entry-point sequence's bits.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "capaz_bodies.h"

namespace capaz {

template <int NPH>
struct CoefT {
    double P[NPH], Xi[NPH], ro[NPH];
    double Sw, Sn;  // the cell's saturations, as read
};

// Corey k_r with the exponent fixed at 2 (CLOSURE 2) or read from p (1).
template <int CLOSURE>
__device__ __forceinline__ double corey_t(const DevParams& p, double s) {
    if constexpr (CLOSURE == 2) {
        s = clamp01(s);
        return s * s;
    } else {
        return corey_kr(p, s);
    }
}

// assign_P_Xi + assign_ro of one cell with the model and closure fixed at
// compile time (cpu.cpp:143-169; two-phase.cpp:4-23; three-phase.cpp:198-240).
// CLOSURE 3 (three phases): the values the closure pass of this step has
// already stored (P_n, P_g, Xi, ro arrays).
template <int MODEL, int CLOSURE>
__device__ __forceinline__ CoefT<MODEL> coef_at(const DevParams& p, const StepIO& io, int64_t l, bool inner) {
    CoefT<MODEL> c;
    if constexpr (CLOSURE == 3) {
        static_assert(MODEL == 3, "stored coefficients: three phases");
        (void)p;
        (void)inner;
        c.Sw = io.Sw[l];
        c.Sn = io.Sn[l];
        c.P[0] = io.Pw[l];
        c.P[1] = io.Pn[l];
        c.P[2] = io.Pg[l];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            c.Xi[a] = io.Xi[a][l];
            c.ro[a] = io.ro[a][l];
        }
        return c;
    }
    const double Pw = io.Pw[l], K = io.K[l];
    if constexpr (MODEL == 2) {
        double kw, kn, Pk;
        const double Sn = io.Sn[l];
        c.Sn = Sn;
        c.Sw = 0.;
        if constexpr (CLOSURE >= 1) {
            double se = exact_div(1. - Sn - p.S_wr, p.den2c, p.r_den2c);
            kw = corey_t<CLOSURE>(p, se);
            kn = corey_t<CLOSURE>(p, 1. - se);
            Pk = 0.;
        } else {
            relperm2(p, Sn, kw, kn, Pk);
        }
        const double Pn = Pw + Pk;
        c.P[0] = Pw;
        c.P[1] = Pn;
        c.Xi[0] = xi_const(K, kw, p.mu_w, p.r_mu_w);
        c.Xi[1] = xi_const(K, kn, p.mu_n, p.r_mu_n);
        c.ro[0] = ro_w_of(p, Pw);
        c.ro[1] = ro_n_of(p, Pn);
    } else {
        double kw, kn, kg;
        const double Sw = io.Sw[l], Sn = io.Sn[l];
        c.Sw = Sw;
        c.Sn = Sn;
        if constexpr (CLOSURE >= 1) {
            const double Swe = exact_div(Sw - p.S_wr, p.den3, p.r_den3);
            const double Sne = exact_div(Sn - p.S_nr, p.den3, p.r_den3);
            const double Sge = 1. - Swe - Sne;
            kw = corey_t<CLOSURE>(p, Swe);
            kn = corey_t<CLOSURE>(p, Sne);
            kg = corey_t<CLOSURE>(p, Sge);
        } else {
            relperm3(p, Sw, Sn, kw, kn, kg);
        }
        const double Pn = inner ? Pw : io.Pn[l];
        const double Pg = inner ? Pw : io.Pg[l];
        c.P[0] = Pw;
        c.P[1] = Pn;
        c.P[2] = Pg;
        c.Xi[0] = xi_const(K, kw, p.mu_w, p.r_mu_w);
        c.Xi[1] = xi_const(K, kn, p.mu_n, p.r_mu_n);
        c.Xi[2] = xi_const(K, kg, p.mu_g, p.r_mu_g);
        c.ro[0] = ro_w_of(p, Pw);
        c.ro[1] = ro_n_of(p, Pn);
        c.ro[2] = ro_g_of(p, Pg);
    }
    return c;
}

template <int MODEL>
__device__ __forceinline__ bool inner_cell(const DevParams& p, int i, int j, int k) {
    return MODEL == 2 ? interior_newton2(p, i, j, k) : interior3(p, i, j, k);
}

// inner_cell split into its (i, j) and k factors: inner_cell(p, i, j, k) ==
// inner_ij(p, i, j) && inner_k(p, k).  A tile column evaluates the (i, j)
// factor once, so only one lane mask per column stays live across the planes.
template <int MODEL>
__device__ __forceinline__ bool inner_ij(const DevParams& p, int i, int j) {
    const bool ii = (i != 0) && (i != p.Nx - 1);
    return (MODEL == 2 ? ii : (ii || p.Nx < 2)) && (j != 0) && (j != p.Ny - 1);
}
__device__ __forceinline__ bool inner_k(const DevParams& p, int k) {
    return ((k != 0) && (k != p.Nz - 1)) || (p.Nz < 2);
}

// The new state (P_w, S_w, S_n, P_n, P_g) of boundary cell (bi, bj, bk) from
// its inward neighbour's new state (two phases: v[0] = P_w, v[2] = S_n).
template <int MODEL>
__device__ __forceinline__ void boundary_values(const DevParams& p, int bi, int bj, int bk, double Pw, double Sw,
                                                double Sn, double ro_eff, double* v) {
    if constexpr (MODEL == 2) {
        double Snb = Sn;
        if (bj == 0 && p.source > 0 && in_source_window(p, bi, bk)) Snb = p.S_n_gr;
        double Pb;
        if ((bj != 0) && (bj != p.Ny - 1))
            Pb = Pw;
        else if (bj == 0)
            Pb = p.P_atm;
        else if (p.P_bottom > 0)  // the configurations' fixed bottom pressure
            Pb = p.P_bottom;
        else
            Pb = Pw + ro_eff;
        v[0] = Pb;
        v[2] = Snb;
    } else {
        (void)bi;
        (void)bk;
        if (bj == 0 && p.source > 0) {
            v[1] = p.S_w_gr;
            v[2] = p.S_n_gr;
        } else {
            v[1] = Sw;
            v[2] = Sn;
        }
        border_p3(p, bj, Pw, v[0], v[3], v[4]);
    }
}

// Boundary cell (bi, bj, bk) from its inward neighbour's new state.
template <int MODEL>
__device__ __forceinline__ void write_boundary(const DevParams& p, const StepIO& io, int bi, int bj, int bk,
                                               double Pw, double Sw, double Sn, double ro_eff) {
    const int64_t lb = (int64_t)bi + (int64_t)bj * p.sy + (int64_t)bk * p.sz;
    double v[5];
    boundary_values<MODEL>(p, bi, bj, bk, Pw, Sw, Sn, ro_eff, v);
    io.Pw1[lb] = v[0];
    io.Sn1[lb] = v[2];
    if constexpr (MODEL == 3) {
        io.Sw1[lb] = v[1];
        io.Pn1[lb] = v[3];
        io.Pg1[lb] = v[4];
    }
}

// An epilogue with kWantsU = true also receives the cell's Darcy velocities
// u[phase][axis] (assign_u's central differences, cpu.cpp:346-477), from the
// neighbours' pressures the tile already holds.
template <class E, class = void>
struct EpiWantsU : std::false_type {};
template <class E>
struct EpiWantsU<E, std::void_t<decltype(E::kWantsU)>> : std::bool_constant<E::kWantsU> {};

template <int NV, int TX, int TY>
struct FusedLds {
    double v[NV][TY + 2][TX + 2];
};

// One tile of the fused step: interior columns i0 .. i0+TX-1 (i0 = first
// interior i + TX*bx), j0 .. j0+TY-1 (j0 = 1 + TY*by), planes k0 .. k0+kchunk-1
// (k0 = first interior k + kchunk*bz), on a TX x TY workgroup (TY >= 3,
// TX >= 2*TY: rows 0-2 of the workgroup also load the halo ring).
//
// CARRY: keep the next plane's coefficients of the own column in registers
// (computed for the z flux) as the next centre, or recompute the centre at
// the top of every plane (fewer live registers across the Newton solve).
// Epi: what happens at an interior cell once its coefficients c and flux
// divergence div are known -- the Newton solve and store of the step
// (FusedStore) or the adjoint's backward Newton (AdjNewtonStore).
template <int MODEL, int CLOSURE, int TX, int TY, bool CARRY, class Epi>
__device__ __forceinline__ void fused_z_tile(const DevParams& p, const StepIO& io, int kchunk, int bx, int by, int bz,
                                             FusedLds<3 * MODEL, TX, TY>& lds, const Epi& epi) {
    static_assert(TY >= 3 && TX >= 2 * TY, "halo ring needs TY >= 3 and TX >= 2*TY");
    constexpr int NPH = MODEL;

    const int tx = threadIdx.x, ty = threadIdx.y;
    const bool hasx = p.Nx >= 3, hasz = p.Nz >= 3;
    const int ilo = hasx ? 1 : 0, ihi = hasx ? p.Nx - 2 : 0;
    const int klo = hasz ? 1 : 0, khi = hasz ? p.Nz - 2 : 0;
    const int i0 = ilo + TX * bx, j0 = 1 + TY * by;
    const int k0 = klo + kchunk * bz;
    const int k1 = min(khi, k0 + kchunk - 1);
    const int i = i0 + tx, j = j0 + ty;
    const bool own_in_grid = i < p.Nx && j < p.Ny;
    const bool active = i <= ihi && j <= p.Ny - 2;

    // This thread's halo cell of each plane, if any (ty 0 / 1: the y rows
    // below / above the block, ty 2: the x columns left / right of it).
    int hi = -1, hj = -1, hx = 0, hy = 0;
    if (ty == 0) { hi = i; hj = j0 - 1; hx = tx + 1; hy = 0; }
    else if (ty == 1) { hi = i; hj = j0 + TY; hx = tx + 1; hy = TY + 1; }
    else if (ty == 2 && tx < TY) { hi = i0 - 1; hj = j0 + tx; hx = 0; hy = tx + 1; }
    else if (ty == 2 && tx >= TY && tx < 2 * TY) { hi = i0 + TX; hj = j0 + tx - TY; hx = TX + 1; hy = tx - TY + 1; }
    const bool halo_in_grid = hi >= 0 && hi < p.Nx && hj >= 0 && hj < p.Ny && (hasx || hi == i0 + tx);

    const int64_t col = (int64_t)i + (int64_t)j * p.sy;
    const int64_t hcol = (int64_t)hi + (int64_t)hj * p.sy;
    const bool own_ij = inner_ij<MODEL>(p, i, j), halo_ij = inner_ij<MODEL>(p, hi, hj);

    // The z window: with ZPARTS the minus z face is carried between planes
    // as its two parts of directed(), zneg = neg_part(x1) and zt2 = t2 (x1 =
    // -dz; computed from the same operands as at the plane itself, one plane
    // earlier): two values per phase instead of dz and the plane below's Xi
    // and ro, and no register copies of those.  That frees registers in the
    // Corey-2 kernels (the headline's: 543 -> 536 us at 256^3), but adds
    // spill to the kernels already held to 128 VGPRs (the closures with
    // powers), which keep dz, Xi and ro.
    constexpr bool ZPARTS = CLOSURE == 2;
    CoefT<NPH> zm{}, c{}, zp{};
    double zneg[NPH], zt2[NPH], dz[NPH];
#pragma unroll
    for (int a = 0; a < NPH; ++a) zneg[a] = zt2[a] = dz[a] = 0.;
    if (own_in_grid) {
        c = coef_at<MODEL, CLOSURE>(p, io, col + (int64_t)k0 * p.sz, own_ij && inner_k(p, k0));
        if (hasz) {
            zm = coef_at<MODEL, CLOSURE>(p, io, col + (int64_t)(k0 - 1) * p.sz, own_ij && inner_k(p, k0 - 1));
#pragma unroll
            for (int a = 0; a < NPH; ++a) {
                const double d = exact_div(c.P[a] - zm.P[a], p.hz, p.r_hz);
                if constexpr (ZPARTS) {
                    zneg[a] = neg_part(-1. * d);
                    zt2[a] = directed_t2(-1. * d, zm.Xi[a], zm.ro[a]);
                } else {
                    dz[a] = d;
                }
            }
        }
    }

    for (int k = k0; k <= k1; ++k) {
        if (!CARRY && own_in_grid && k != k0)
            c = coef_at<MODEL, CLOSURE>(p, io, col + (int64_t)k * p.sz, own_ij && inner_k(p, k));
        // stage plane k: own cell and halo cell
        if (own_in_grid) {
#pragma unroll
            for (int a = 0; a < NPH; ++a) {
                lds.v[a][ty + 1][tx + 1] = c.P[a];
                lds.v[NPH + a][ty + 1][tx + 1] = c.Xi[a];
                lds.v[2 * NPH + a][ty + 1][tx + 1] = c.ro[a];
            }
        }
        if (hi >= 0 && halo_in_grid) {
            CoefT<NPH> h = coef_at<MODEL, CLOSURE>(p, io, hcol + (int64_t)k * p.sz, halo_ij && inner_k(p, k));
#pragma unroll
            for (int a = 0; a < NPH; ++a) {
                lds.v[a][hy][hx] = h.P[a];
                lds.v[NPH + a][hy][hx] = h.Xi[a];
                lds.v[2 * NPH + a][hy][hx] = h.ro[a];
            }
        }
        __syncthreads();

        if (own_in_grid && hasz)
            zp = coef_at<MODEL, CLOSURE>(p, io, col + (int64_t)(k + 1) * p.sz, own_ij && inner_k(p, k + 1));
        // the plus z face's difference; the next plane's dz
        double zd[NPH];
#pragma unroll
        for (int a = 0; a < NPH; ++a) zd[a] = hasz ? exact_div(zp.P[a] - c.P[a], p.hz, p.r_hz) : 0.;

        if (active) {
            const int64_t l = col + (int64_t)k * p.sz;
            double div[NPH];
#pragma unroll
            for (int a = 0; a < NPH; ++a) {
                // order as cpu.cpp:651-708: z, x, y, then (-q + fx + fy + fz)
                double fz = 0., fx = 0.;
                if (hasz) {
                    if constexpr (ZPARTS)
                        fz = directed_parts(zneg[a], zt2[a], -1. * zd[a], c.Xi[a], c.ro[a], zp.Xi[a], zp.ro[a], p.hz,
                                            p.r_hz);
                    else
                        fz = flux_x(dz[a], zd[a], c.Xi[a], c.ro[a], zm.Xi[a], zm.ro[a], zp.Xi[a], zp.ro[a], p.hz,
                                    p.r_hz);
                }
                if (hasx) {
                    const double Pm = lds.v[a][ty + 1][tx], Pp = lds.v[a][ty + 1][tx + 2];
                    const double dm = exact_div(c.P[a] - Pm, p.hx, p.r_hx);
                    const double dp = exact_div(Pp - c.P[a], p.hx, p.r_hx);
                    fx = flux_x(dm, dp, c.Xi[a], c.ro[a], lds.v[NPH + a][ty + 1][tx], lds.v[2 * NPH + a][ty + 1][tx],
                                lds.v[NPH + a][ty + 1][tx + 2], lds.v[2 * NPH + a][ty + 1][tx + 2], p.hx, p.r_hx);
                }
                const double Pm = lds.v[a][ty][tx + 1], Pp = lds.v[a][ty + 2][tx + 1];
                const double dm = exact_div(c.P[a] - Pm, p.hy, p.r_hy);
                const double dp = exact_div(Pp - c.P[a], p.hy, p.r_hy);
                const double fy = flux_y(p, dm, dp, c.Xi[a], c.ro[a], lds.v[NPH + a][ty][tx + 1],
                                         lds.v[2 * NPH + a][ty][tx + 1], lds.v[NPH + a][ty + 2][tx + 1],
                                         lds.v[2 * NPH + a][ty + 2][tx + 1]);
                double s = (a == 2 ? 0. + fx : fx) + fy;
                div[a] = s + fz;
            }
            // the plus z face's parts, the next plane's minus face's (only
            // active cells read them)
            if constexpr (ZPARTS) {
#pragma unroll
                for (int a = 0; a < NPH; ++a) {
                    zneg[a] = neg_part(-1. * zd[a]);
                    zt2[a] = directed_t2(-1. * zd[a], c.Xi[a], c.ro[a]);
                }
            }
            if constexpr (EpiWantsU<Epi>::value) {
                // interior cell: central differences on every axis longer than 2
                double u[NPH][3];
#pragma unroll
                for (int a = 0; a < NPH; ++a) {
                    u[a][0] = hasx ? c.Xi[a] * exact_div(lds.v[a][ty + 1][tx + 2] - lds.v[a][ty + 1][tx], p.two_hx,
                                                         p.r_two_hx)
                                   : 0.;
                    u[a][1] = c.Xi[a] * (exact_div(lds.v[a][ty + 2][tx + 1] - lds.v[a][ty][tx + 1], p.two_hy,
                                                   p.r_two_hy) -
                                         c.ro[a] * p.g);
                    u[a][2] = hasz ? c.Xi[a] * exact_div(zp.P[a] - zm.P[a], p.two_hz, p.r_two_hz) : 0.;
                }
                epi.store_u(l, u);
            }
            epi(p, i, j, k, l, c, div);
        }
        __syncthreads();
        // roll the z window
#pragma unroll
        for (int a = 0; a < NPH; ++a) {
            if constexpr (!ZPARTS) {
                dz[a] = zd[a];
                zm.Xi[a] = c.Xi[a];
                zm.ro[a] = c.ro[a];
            }
            if constexpr (EpiWantsU<Epi>::value) zm.P[a] = c.P[a];
        }
        if (CARRY) c = zp;
    }
}

// Tile of workgroup (blockIdx): workgroups go to the 8 XCDs round robin, so
// when the grid divides by 8 each XCD takes one contiguous run of tiles (x
// fastest, then y, then z) and the halo rows its tiles share stay in its L2.
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t nb = gx * gy * gridDim.z;
    uint32_t id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (nb % 8u == 0u) id = (id % 8u) * (nb / 8u) + id / 8u;
    bx = (int)(id % gx);
    const uint32_t r = id / gx;
    by = (int)(r % gy);
    bz = (int)(r / gy);
}

// The boundary cells whose inward neighbour (set_boundary_basic_coordinate,
// cpu.cpp:107-141) is the interior cell (i, j, k), from its new state: the
// cell's offsets along each axis are 0 and -1 next to the low face, +1 next to
// the high face.  On the supported grids every boundary cell has exactly one
// interior inward neighbour, so this writes each boundary cell once, with the
// values k_fused_faces would write.
template <int MODEL>
__device__ __forceinline__ void store_faces(const DevParams& p, const StepIO& io, int i, int j, int k, int64_t l,
                                            const CoefT<MODEL>& c, double Pw, double Sw, double Sn) {
    const bool hasx = p.Nx >= 3, hasz = p.Nz >= 3;
    const bool xl = hasx && i == 1, xh = hasx && i == p.Nx - 2;
    const bool yl = j == 1, yh = j == p.Ny - 2;
    const bool zl = hasz && k == 1, zh = hasz && k == p.Nz - 2;
    if (!(xl || xh || yl || yh || zl || zh)) return;
    double ro_eff = 0.;
    if constexpr (MODEL == 2) {
        // ro_eff_gdy (main.cpp:255-267) for the bottom face: this step's
        // densities of the cell (from its previous state), its new saturation
        if (yh) ro_eff = (c.ro[1] * Sn + c.ro[0] * (1 - Sn)) * io.m[l] * p.g * p.hy;
    }
    int xo[3] = {0, 0, 0}, yo[3] = {0, 0, 0}, zo[3] = {0, 0, 0};
    int nx = 1, ny = 1, nz = 1;
    if (xl) xo[nx++] = -1;
    if (xh) xo[nx++] = 1;
    if (yl) yo[ny++] = -1;
    if (yh) yo[ny++] = 1;
    if (zl) zo[nz++] = -1;
    if (zh) zo[nz++] = 1;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b)
            for (int d = 0; d < nz; ++d)
                if (a || b || d) write_boundary<MODEL>(p, io, i + xo[a], j + yo[b], k + zo[d], Pw, Sw, Sn, ro_eff);
}

// The step's epilogue: A = roS - dt/m * div, the Newton solve, the new state;
// FACES: also the boundary cells this cell is the inward neighbour of (else
// k_fused_faces writes them after the kernel).  The two-phase kernels write
// their faces here (the strided x faces cost a separate pass more than the
// epilogue costs); in the three-phase kernels the extra registers would cost
// a wave per SIMD, so they keep the face pass.  CLOSURE: the kernel's
// (the two-phase Newton then carries only its closure's code).
template <int MODEL, bool FACES = false, int CLOSURE = -1>
struct FusedStore {
    StepIO io;
    __device__ __forceinline__ void operator()(const DevParams& p, int i, int j, int k, int64_t l,
                                               const CoefT<MODEL>& c, const double* div) const {
        (void)i; (void)j; (void)k;
        const double dtm = io.dtm[l];  // dt / m, computed once per porosity field
        double Pw = c.P[0], Sn = c.Sn, Sw = c.Sw;
        if constexpr (MODEL == 2) {
            const double A1 = c.ro[0] * (1. - Sn) - dtm * div[0];
            const double A2 = c.ro[1] * Sn - dtm * div[1];
            newton2<CLOSURE>(p, Pw, Sn, A1, A2);
            io.Pw1[l] = Pw;
            io.Sn1[l] = Sn;
            if constexpr (FACES) store_faces<2>(p, io, i, j, k, l, c, Pw, 0., Sn);
        } else {
            const double A1 = c.ro[0] * Sw - dtm * div[0];
            const double A2 = c.ro[1] * Sn - dtm * div[1];
            const double A3 = c.ro[2] * (1. - Sw - Sn) - dtm * div[2];
            newton3(p, Pw, Sw, Sn, A1, A2, A3);
            io.Pw1[l] = Pw;
            io.Sw1[l] = Sw;
            io.Sn1[l] = Sn;
            if constexpr (FACES) store_faces<3>(p, io, i, j, k, l, c, Pw, Sw, Sn);
        }
    }
};

// Two phases: boundary cells included (kFusedFaces); three phases: followed
// by k_fused_faces.
template <int MODEL>
constexpr bool kFusedFaces = MODEL == 2;

// FusedStore that also records the cell's flux divergence div_a (field a at
// div + a*stride), for the adjoint's backward Newton.
template <int MODEL, int CLOSURE = -1>
struct FusedStoreRecord {
    FusedStore<MODEL, kFusedFaces<MODEL>, CLOSURE> step;
    double* div;
    int64_t stride;
    __device__ __forceinline__ void operator()(const DevParams& p, int i, int j, int k, int64_t l,
                                               const CoefT<MODEL>& c, const double* d) const {
#pragma unroll
        for (int a = 0; a < MODEL; ++a) div[a * stride + l] = d[a];
        step(p, i, j, k, l, c, d);
    }
};

// assign_roS_nr (cpu.cpp:619-731) on the z tiles from the stored
// coefficients (CLOSURE 3, three phases): roS_old and roS of the cell and,
// with U, its velocities (assign_u) from the same neighbours -- the roS entry
// point, and the energy step's roS + u pass.
template <bool U>
struct RoSStore {
    static constexpr bool kWantsU = U;
    const double* dtm;
    double *roS[3], *roS_old[3];
    double* u[3][3];  // [phase][axis]: ux_w, uy_w, uz_w, ux_n, ... (U only)
    __device__ __forceinline__ void store_u(int64_t l, const double (&v)[3][3]) const {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int x = 0; x < 3; ++x) u[a][x][l] = v[a][x];
    }
    __device__ __forceinline__ void operator()(const DevParams& p, int i, int j, int k, int64_t l,
                                               const CoefT<3>& c, const double* div) const {
        (void)p; (void)i; (void)j; (void)k;
        const double old[3] = {c.ro[0] * c.Sw, c.ro[1] * c.Sn, c.ro[2] * (1. - c.Sw - c.Sn)};
        const double d = dtm[l];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            roS_old[a][l] = old[a];
            roS[a][l] = old[a] - d * div[a];
        }
    }
};

template <int TX, int TY, bool CARRY, bool U>
__global__ void __launch_bounds__(TX * TY) k_roS_z(DevParams p, StepIO io, int kchunk, RoSStore<U> st) {
    __shared__ FusedLds<9, TX, TY> lds;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    fused_z_tile<3, 3, TX, TY, CARRY>(p, io, kchunk, bx, by, bz, lds, st);
}

// The energy step's E pass (assign_E_new, energy.cpp:235-340, as body_E_with)
// on z-marching TX x TY tiles: per plane, each cell's lambda_eff * T and its
// advected products (ro_a * H_a) * ux_a and (ro_a * H_a) * uy_a are staged in
// LDS with a one-cell halo ring; the z neighbours' lambda_eff * T and
// (ro_a * H_a) * uy_a (the z term advects with uy, energy.cpp:300-305) are
// computed per column.  Each term is the product the per-cell body forms, so
// E_new has its bits.
// lam: lambda_eff per cell as the step's first pass stored it, or NULL (the
// entry point) to evaluate it from m, S_w, S_n, S_g and T.
// uzz: the velocities the z term advects with -- uy (energy.cpp:300-305), or
// uz in the configurations' energy closure (p.ecfg), whose conductivity is
// the constant p.lam_eff and whose z conduction adds to the x and y terms.
struct EnergyEIO {
    const double *lam, *T, *E;
    const double *m, *Sw, *Sn, *Sg;
    const double *ro[3], *H[3], *ux[3], *uy[3], *uzz[3];
    double* E_new;
    double lam_cfg;  // > 0: the constant conductivity of p.ecfg
};

__device__ __forceinline__ double energy_lam(const EnergyEIO& a, int64_t l) {
    if (a.lam_cfg > 0) return a.lam_cfg;
    return a.lam ? a.lam[l] : lambda_eff(a.m[l], a.Sw[l], a.Sn[l], a.Sg[l], a.T[l]);
}

// v[0] = lambda_eff * T, v[1 + a] = (ro_a H_a) ux_a, v[4 + a] = (ro_a H_a) uy_a
__device__ __forceinline__ void energy_e_vals(const EnergyEIO& a, int64_t l, double* v) {
    v[0] = energy_lam(a, l) * a.T[l];
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        const double rh = a.ro[ph][l] * a.H[ph][l];
        v[1 + ph] = rh * a.ux[ph][l];
        v[4 + ph] = rh * a.uy[ph][l];
    }
}

// lambda_eff * T and the uzz products of a z neighbour: w[0], w[1 + a]
__device__ __forceinline__ void energy_e_zvals(const EnergyEIO& a, int64_t l, double* w) {
    w[0] = energy_lam(a, l) * a.T[l];
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) w[1 + ph] = a.ro[ph][l] * a.H[ph][l] * a.uzz[ph][l];
}

// advection's (((X_w+ - X_w-) + X_n+) - X_n-) + X_g+) - X_g- (the body's order)
__device__ __forceinline__ double energy_adv(double wp, double wm, double np_, double nm, double gp, double gm) {
    return wp - wm + np_ - nm + gp - gm;
}

template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_energy_e_z(DevParams p, EnergyEIO a, int kchunk) {
    static_assert(TY >= 3 && TX >= 2 * TY, "halo ring needs TY >= 3 and TX >= 2*TY");
    __shared__ double lds[7][TY + 2][TX + 2];
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const bool hasx = p.Nx >= 3, hasz = p.Nz >= 3;
    const int ilo = hasx ? 1 : 0, ihi = hasx ? p.Nx - 2 : 0;
    const int klo = hasz ? 1 : 0, khi = hasz ? p.Nz - 2 : 0;
    const int i0 = ilo + TX * bx, j0 = 1 + TY * by;
    const int k0 = klo + kchunk * bz;
    const int k1 = min(khi, k0 + kchunk - 1);
    const int i = i0 + tx, j = j0 + ty;
    const bool own_in_grid = i < p.Nx && j < p.Ny;
    const bool active = i <= ihi && j <= p.Ny - 2;
    // halo cell of this thread, as in fused_z_tile
    int hi = -1, hj = -1, hx = 0, hy = 0;
    if (ty == 0) { hi = i; hj = j0 - 1; hx = tx + 1; hy = 0; }
    else if (ty == 1) { hi = i; hj = j0 + TY; hx = tx + 1; hy = TY + 1; }
    else if (ty == 2 && tx < TY) { hi = i0 - 1; hj = j0 + tx; hx = 0; hy = tx + 1; }
    else if (ty == 2 && tx >= TY && tx < 2 * TY) { hi = i0 + TX; hj = j0 + tx - TY; hx = TX + 1; hy = tx - TY + 1; }
    const bool halo_in_grid = hi >= 0 && hi < p.Nx && hj >= 0 && hj < p.Ny && (hasx || hi == i0 + tx);
    const int64_t col = (int64_t)i + (int64_t)j * p.sy;
    const int64_t hcol = (int64_t)hi + (int64_t)hj * p.sy;

    double zm[4] = {}, zp[4] = {};
    if (own_in_grid && hasz) energy_e_zvals(a, col + (int64_t)(k0 - 1) * p.sz, zm);
    for (int k = k0; k <= k1; ++k) {
        const int64_t l = col + (int64_t)k * p.sz;
        double v[7] = {};
        if (own_in_grid) {
            energy_e_vals(a, l, v);
#pragma unroll
            for (int q = 0; q < 7; ++q) lds[q][ty + 1][tx + 1] = v[q];
        }
        if (hi >= 0 && halo_in_grid) {
            double h[7];
            energy_e_vals(a, hcol + (int64_t)k * p.sz, h);
#pragma unroll
            for (int q = 0; q < 7; ++q) lds[q][hy][hx] = h[q];
        }
        __syncthreads();
        if (own_in_grid && hasz) energy_e_zvals(a, l + p.sz, zp);
        if (active) {
            const double lc = energy_lam(a, l), Tc = a.T[l];
            double T_flow = 0;
            if (p.Nx > 2)
                T_flow += exact_div(lds[0][ty + 1][tx + 2] - 2 * lc * Tc + lds[0][ty + 1][tx], p.hx2, p.r_hx2);
            if (p.Ny > 2)
                T_flow += exact_div(lds[0][ty + 2][tx + 1] - 2 * lc * Tc + lds[0][ty][tx + 1], p.hy2, p.r_hy2);
            if (p.Nz > 2) {
                const double tz = exact_div(zp[0] - 2 * lc * Tc + zm[0], p.hz2, p.r_hz2);
                if (p.ecfg)
                    T_flow += tz;
                else
                    T_flow = tz;  // '=' as energy.cpp:257
            }
            double E_flow = 0;
            if (p.Nx > 2)
                E_flow += exact_div(energy_adv(lds[1][ty + 1][tx + 2], lds[1][ty + 1][tx], lds[2][ty + 1][tx + 2],
                                               lds[2][ty + 1][tx], lds[3][ty + 1][tx + 2], lds[3][ty + 1][tx]),
                                    p.two_hx, p.r_two_hx);
            if (p.Ny > 2)
                E_flow += exact_div(energy_adv(lds[4][ty + 2][tx + 1], lds[4][ty][tx + 1], lds[5][ty + 2][tx + 1],
                                               lds[5][ty][tx + 1], lds[6][ty + 2][tx + 1], lds[6][ty][tx + 1]),
                                    p.two_hy, p.r_two_hy);
            if (p.Nz > 2)
                E_flow += exact_div(energy_adv(zp[1], zm[1], zp[2], zm[2], zp[3], zm[3]), p.two_hz, p.r_two_hz);
            double Q_hw = 0, Q_hr = 0;
            a.E_new[l] = a.E[l] + p.dt * (T_flow + Q_hw + Q_hr - E_flow);
        }
        __syncthreads();
        // roll the z window: the centre's lambda_eff * T and uzz products
        if (p.ecfg) {
            if (own_in_grid) energy_e_zvals(a, l, zm);
        } else {
            zm[0] = v[0];
#pragma unroll
            for (int ph = 0; ph < 3; ++ph) zm[1 + ph] = v[4 + ph];
        }
    }
}

template <int MODEL, int CLOSURE, int TX, int TY, bool CARRY>
__global__ void __launch_bounds__(TX * TY) k_fused_z_record(DevParams p, StepIO io, int kchunk, double* div,
                                                           int64_t stride) {
    __shared__ FusedLds<3 * MODEL, TX, TY> lds;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    fused_z_tile<MODEL, CLOSURE, TX, TY, CARRY>(p, io, kchunk, bx, by, bz, lds,
                                                FusedStoreRecord<MODEL, CLOSURE>{
                                                    FusedStore<MODEL, kFusedFaces<MODEL>, CLOSURE>{io}, div, stride});
}


// Occupancy bound: the kernels with a closure that raises powers (the
// reference's van Genuchten or Brooks-Corey k_r, a Corey exponent read at run
// time) need 130-160 VGPRs, which leaves one 512-thread workgroup per CU; held
// to 128 (4 waves per SIMD) they spill at most some 32 VGPRs and still run
// faster (the reference's three-phase deck at 256^3: 1.76 against 1.84 ms per
// step).
template <int MODEL, int CLOSURE>
constexpr int kFusedWaves = CLOSURE == 0 || CLOSURE == 1 ? 4 : 1;
// The two-phase Corey-2 tiles of 64 x 4 (256 threads) fit in 94 VGPRs
// without a spill at 5 waves per SIMD: 330 us per step at 256^3 against 374
// for 64 x 8 tiles at 4 (the automatic tile_y of that step).  The
// three-phase ones would spill 28 VGPRs there and take 871 us against 543.
template <int MODEL, int CLOSURE, int TY>
constexpr int kFusedWavesTY = MODEL == 2 && CLOSURE == 2 && TY == 4 ? 5 : kFusedWaves<MODEL, CLOSURE>;

template <int MODEL, int CLOSURE, int TX, int TY, bool CARRY>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(kFusedWavesTY<MODEL, CLOSURE, TY>)))
k_fused_z(DevParams p, StepIO io, int kchunk) {
    __shared__ FusedLds<3 * MODEL, TX, TY> lds;
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    fused_z_tile<MODEL, CLOSURE, TX, TY, CARRY>(p, io, kchunk, bx, by, bz, lds,
                                                FusedStore<MODEL, kFusedFaces<MODEL>, CLOSURE>{io});
}

// Cell number t of a box of ni x nj cells per plane starting at (i0, j0, k0),
// x fastest (32-bit index arithmetic: grids of < 2^31 cells).
__device__ __forceinline__ void box_cell(uint32_t t, uint32_t ni, uint32_t nj, int i0, int j0, int k0, int& i, int& j,
                                         int& k) {
    const uint32_t a = t / ni, b = a / nj;
    i = i0 + (int)(t - a * ni);
    j = j0 + (int)(a - b * nj);
    k = k0 + (int)b;
}

// Coordinates of boundary cell number t (t enumerates the j faces (all i,
// k), then the x faces (j interior, all k), then the z faces (i, j
// interior)); false for t past the last face cell.  32-bit index
// arithmetic: the face cells of a grid of < 2^31 cells.
__device__ __forceinline__ bool face_cell(const DevParams& p, int64_t t64, int& i, int& j, int& k) {
    const bool hasx = p.Nx >= 3, hasz = p.Nz >= 3;
    const uint32_t Nx = (uint32_t)p.Nx, Ny2 = (uint32_t)(p.Ny - 2);
    const uint32_t nj = 2u * Nx * (uint32_t)p.Nz;
    const uint32_t nx = hasx ? 2u * Ny2 * (uint32_t)p.Nz : 0u;
    const uint32_t ni = hasx ? Nx - 2u : Nx;
    const uint32_t nzf = hasz ? 2u * ni * Ny2 : 0u;
    if (t64 >= (int64_t)nj + nx + nzf) return false;
    uint32_t t = (uint32_t)t64;
    if (t < nj) {
        j = (t & 1u) ? p.Ny - 1 : 0;
        t >>= 1;
        const uint32_t q = t / Nx;
        i = (int)(t - q * Nx);
        k = (int)q;
    } else if ((t -= nj) < nx) {
        i = (t & 1u) ? p.Nx - 1 : 0;
        t >>= 1;
        const uint32_t q = t / Ny2;
        j = 1 + (int)(t - q * Ny2);
        k = (int)q;
    } else {
        t -= nx;
        k = (t & 1u) ? p.Nz - 1 : 0;
        t >>= 1;
        const uint32_t q = t / ni;
        i = (hasx ? 1 : 0) + (int)(t - q * ni);
        j = 1 + (int)q;
    }
    return true;
}

// Boundary cell number t of the step after the interior tiles (face_cell's
// numbering); nothing for t past the last face cell.
template <int MODEL, int CLOSURE>
__device__ __forceinline__ void fused_face(const DevParams& p, const StepIO& io, int64_t t64) {
    int i, j, k;
    if (!face_cell(p, t64, i, j, k)) return;
    const int64_t l1 = inward(p, i, j, k);
    const double Pw = io.Pw1[l1], Sn = io.Sn1[l1];
    const double Sw = MODEL == 3 ? io.Sw1[l1] : 0.;
    double ro_eff = 0.;
    if constexpr (MODEL == 2) {
        // ro_eff_gdy (main.cpp:255-267), read by the bottom face only: this
        // step's densities of the inward cell (from its previous state) with
        // its new saturation
        if (j == p.Ny - 1) {
            const CoefT<2> c1 = coef_at<2, CLOSURE>(p, io, l1, true);
            ro_eff = (c1.ro[1] * Sn + c1.ro[0] * (1 - Sn)) * io.m[l1] * p.g * p.hy;
        }
    }
    write_boundary<MODEL>(p, io, i, j, k, Pw, Sw, Sn, ro_eff);
}

template <int MODEL, int CLOSURE>
__global__ void __launch_bounds__(256) k_fused_faces(DevParams p, StepIO io) {
    fused_face<MODEL, CLOSURE>(p, io, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

inline int64_t fused_face_cells(const DevParams& p) {
    const bool hasx = p.Nx >= 3, hasz = p.Nz >= 3;
    const int64_t ni = hasx ? p.Nx - 2 : p.Nx;
    return 2LL * p.Nx * p.Nz + (hasx ? 2LL * (p.Ny - 2) * p.Nz : 0) + (hasz ? 2LL * ni * (p.Ny - 2) : 0);
}

// The halo pack of a z-slab exchange written by the step itself: cell c of
// side s (plane 1, or plane Nz - 2) of the new state into io.pack at
// (f * 2 + s) * Nx * Ny + c for the state fields f (three phases: P_w, S_w,
// S_n, P_n, P_g; two: P_w, S_n) -- capaz_load_faces' layout for axis 2.
// Interior cells give their new values (P_n = P_g = P_w, as the step takes
// them), the plane's x / y boundary cells the Border values of their inward
// cell, computed here rather than read, so the pack needs no second launch
// after the boundary writes.
template <int MODEL, int CLOSURE>
__device__ __forceinline__ void fused_pack(const DevParams& p, const StepIO& io, int64_t t) {
    const int64_t plane = (int64_t)p.Nx * p.Ny;
    if (t >= 2 * plane) return;
    const int s = t >= plane ? 1 : 0;
    const int64_t c = t - s * plane;
    const int k = s ? p.Nz - 2 : 1;
    const int j = (int)(c / p.Nx), i = (int)(c - (int64_t)j * p.Nx);
    const bool interior = (p.Nx < 3 || (i != 0 && i != p.Nx - 1)) && j != 0 && j != p.Ny - 1;
    const int64_t l = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
    double v[5];
    if (interior) {
        v[0] = io.Pw1[l];
        v[2] = io.Sn1[l];
        if constexpr (MODEL == 3) {
            v[1] = io.Sw1[l];
            v[3] = v[4] = v[0];
        }
    } else {
        const int64_t l1 = inward(p, i, j, k);
        const double Pw = io.Pw1[l1], Sn = io.Sn1[l1];
        const double Sw = MODEL == 3 ? io.Sw1[l1] : 0.;
        double ro_eff = 0.;
        if constexpr (MODEL == 2) {
            if (j == p.Ny - 1) {
                const CoefT<2> c1 = coef_at<2, CLOSURE>(p, io, l1, true);
                ro_eff = (c1.ro[1] * Sn + c1.ro[0] * (1 - Sn)) * io.m[l1] * p.g * p.hy;
            }
        }
        boundary_values<MODEL>(p, i, j, k, Pw, Sw, Sn, ro_eff, v);
    }
    if constexpr (MODEL == 3) {
#pragma unroll
        for (int f = 0; f < 5; ++f) io.pack[(f * 2 + s) * plane + c] = v[f];
    } else {
        io.pack[s * plane + c] = v[0];
        io.pack[(2 + s) * plane + c] = v[2];
    }
}

// k_fused_faces followed, in the same launch, by the halo pack (threads past
// the face cells).
template <int MODEL, int CLOSURE>
__global__ void __launch_bounds__(256) k_fused_faces_pack(DevParams p, StepIO io, int64_t nfaces) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < nfaces)
        fused_face<MODEL, CLOSURE>(p, io, t);
    else
        fused_pack<MODEL, CLOSURE>(p, io, t - nfaces);
}

}  // namespace capaz
