// The two-phase energy model's fused time step for gfx950 (the
// configurations' energy closure): one pass over the primary state (P_w,
// S_n, T) per step, reading the current arrays and writing the second ones.
//
// Per interior cell the step of the entry-point sequence is
//   roS (the 7-point upwind mass flux, cpu.cpp:619-731) from the closure
//       coefficients (P, Xi, ro) of the cell and its six neighbours,
//   E   (assign_E_current) of the cell,
//   E_new = E + dt (T_flow - E_flow): the conduction of T over the six
//       neighbours and the central difference of (ro H) u at them, where u
//       at a neighbour is its Darcy velocity (assign_u), a difference of P
//       at the cells on either side of it -- two cells from the centre;
//   the 4 x 4 Newton in (S_w, S_n, P_w, T) in registers,
// and the boundary cells take their new state from their inward cell
// (Border_S, Border_P, Border_T) in the same thread.
//
// Mapping.  A workgroup owns a 64 x 8 block of interior columns and marches a
// chunk of planes along z.  Per plane each thread stages its cell's
// coefficients in LDS; the ring of cells around the block (x and y
// neighbours of the edge cells) is staged too, and a second ring of
// pressures only (the far points of the neighbours' velocities).  Along z
// each thread keeps the coefficients of its column's planes k-1, k, k+1 in
// registers and carries the (ro H) u_z products of planes k-1 and k, so each
// plane's u_z is computed once.  Every value is formed by the per-cell
// bodies' operations in their order, so the step has the bits of the
// entry-point sequence (capaz_bodies.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "capaz_fused.h"
#include "capaz_launch.h"

namespace capaz {

namespace {

constexpr int TX = 64, TY = 8;

struct E2IO {
    const double *Pw, *Sn, *T, *m, *K;  // current state and parameters
    double *Pw1, *Sn1, *T1;             // next state
};

// Coefficients of a cell that a neighbour reads.
struct E2Coef {
    double P[2], Xi[2], ro[2], roH[2];  // phase pressure, Xi, density, density * enthalpy
    double T;
};

// P_n of a cell: P_w + P_k (two-phase.cpp:4-23; 0 for the Corey closures).
template <int CLOSURE>
__device__ __forceinline__ double e2_pn(const DevParams& p, double Pw, double Sn) {
    if constexpr (CLOSURE >= 1) {
        (void)p;
        (void)Sn;
        return Pw;
    } else {
        double kw, kn, Pk;
        relperm2(p, Sn, kw, kn, Pk);
        return Pw + Pk;
    }
}

// assign_P_Xi + assign_ro + assign_H of a cell (body_P_Xi_ro, body_HE).
template <int CLOSURE>
__device__ __forceinline__ E2Coef e2_coef(const DevParams& p, double Pw, double Sn, double T, double K) {
    double kw, kn, Pk;
    if constexpr (CLOSURE >= 1) {
        const double se = exact_div(1. - Sn - p.S_wr, p.den2c, p.r_den2c);
        if constexpr (CLOSURE == 2) {
            const double s = clamp01(se), t = clamp01(1. - se);
            kw = s * s;
            kn = t * t;
        } else {
            kw = corey_kr(p, se);
            kn = corey_kr(p, 1. - se);
        }
        Pk = 0.;
    } else {
        relperm2(p, Sn, kw, kn, Pk);
    }
    E2Coef c;
    c.P[0] = Pw;
    c.P[1] = Pw + Pk;
    c.Xi[0] = xi_var(K, kw, mu_cfg(p, p.mu_w, T));
    c.Xi[1] = xi_var(K, kn, mu_cfg(p, p.mu_n, T));
    c.ro[0] = ro_w_T(p, c.P[0], T);
    c.ro[1] = ro_n_T(p, c.P[1], T);
    c.roH[0] = c.ro[0] * H_w_of(p, T);
    c.roH[1] = c.ro[1] * H_n_T(p, T);
    c.T = T;
    return c;
}

// The difference of P that assign_u takes at a cell of index idx on an axis
// of n >= 3 cells, from P at the cell (pc) and at its minus / plus
// neighbours (cpu.cpp:185-207, 346-477).
__device__ __forceinline__ double e2_dP(double pm, double pc, double pp, int idx, int n, double h, double rh,
                                        double h2, double rh2) {
    if (idx == 0) return exact_div(pp - pc, h, rh);
    if (idx == n - 1) return exact_div(pc - pm, h, rh);
    return exact_div(pp - pm, h2, rh2);
}

// The general Newton for a cell whose inverse needs the pivot search, out of
// line: x = {S_w, S_n, P, T, H_w, H_n, H_r}.  It reads the parameters from
// the workgroup's LDS copy: a reference to the kernel's own parameter block
// would make the compiler copy that block to every lane's scratch memory.
__device__ __attribute__((noinline)) void e2_newton_general(const DevParams* p, double* x, double m, double roSw,
                                                           double roSn, double E_new) {
    newton_energy2_t<false>(*p, x[0], x[1], x[2], x[3], x[4], x[5], x[6], m, roSw, roSn, E_new);
}

// Boundary cell (bi, bj, bk) from its inward cell's new state (Border_S and
// Border_P, two-phase.cpp:163-215, with the fixed bottom pressure; Border_T
// with the configuration's face temperatures).
__device__ __forceinline__ void e2_write_boundary(const DevParams& p, const E2IO& io, int bi, int bj, int bk,
                                                  double Pw, double Sn, double T, double ro_eff) {
    const int64_t lb = (int64_t)bi + (int64_t)bj * p.sy + (int64_t)bk * p.sz;
    double Snb = Sn;
    if (bj == 0 && p.source > 0 && in_source_window(p, bi, bk)) Snb = p.S_n_gr;
    io.Sn1[lb] = Snb;
    double Pb;
    if ((bj != 0) && (bj != p.Ny - 1))
        Pb = Pw;
    else if (bj == 0)
        Pb = p.P_atm;
    else if (p.P_bottom > 0)
        Pb = p.P_bottom;
    else
        Pb = Pw + ro_eff;
    io.Pw1[lb] = Pb;
    io.T1[lb] = border_T_cfg(p, bj, T);
}

__device__ __forceinline__ void e2_store_faces(const DevParams& p, const E2IO& io, int i, int j, int k, int64_t l,
                                               double ro_w, double ro_n, double Pw, double Sn, double T) {
    const bool hasz = p.Nz >= 3;
    const bool xl = i == 1, xh = i == p.Nx - 2;
    const bool yl = j == 1, yh = j == p.Ny - 2;
    const bool zl = hasz && k == 1, zh = hasz && k == p.Nz - 2;
    if (!(xl || xh || yl || yh || zl || zh)) return;
    // ro_eff_gdy (main.cpp:255-267): this step's densities of the cell (from
    // its previous state), its new saturation
    const double ro_eff = yh ? (ro_n * Sn + ro_w * (1 - Sn)) * io.m[l] * p.g * p.hy : 0.;
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
                if (a || b || d) e2_write_boundary(p, io, i + xo[a], j + yo[b], k + zo[d], Pw, Sn, T, ro_eff);
}

// LDS of one plane: pressures of the block and two rings, the other
// coefficients of the block and one ring.
struct E2Lds {
    double P[2][TY + 4][TX + 4];
    double C[7][TY + 2][TX + 2];  // Xi_w, Xi_n, ro_w, ro_n, roH_w, roH_n, T
};

template <int CLOSURE>
__device__ __forceinline__ void e2_tile(const DevParams& p, const E2IO& io, int kchunk, E2Lds& lds, DevParams& lp) {
    int bx, by, bz;
    xcd_tile(bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    if (tx == 0 && ty == 0) lp = p;  // for the out-of-line Newton; read after the first barrier
    const bool hasz = p.Nz >= 3;
    const int ihi = p.Nx - 2;
    const int klo = hasz ? 1 : 0, khi = hasz ? p.Nz - 2 : 0;
    const int i0 = 1 + TX * bx, j0 = 1 + TY * by;
    const int k0 = klo + kchunk * bz;
    const int k1 = min(khi, k0 + kchunk - 1);
    const int i = i0 + tx, j = j0 + ty;
    const bool own_in_grid = i < p.Nx && j < p.Ny;
    const bool active = i <= ihi && j <= p.Ny - 2;

    // This thread's ring cell of each plane, if any: ring 1 (coefficients)
    // for ty 0-2, ring 2 (pressures) for ty 3-5; (hx, hy) its place in the
    // pressure plane, whose origin is (i0 - 2, j0 - 2).
    int hi = -1, hj = -1, hx = 0, hy = 0;
    bool ring1 = false;
    if (ty == 0) { hi = i; hj = j0 - 1; ring1 = true; }
    else if (ty == 1) { hi = i; hj = j0 + TY; ring1 = true; }
    else if (ty == 2 && tx < TY) { hi = i0 - 1; hj = j0 + tx; ring1 = true; }
    else if (ty == 2 && tx >= TY && tx < 2 * TY) { hi = i0 + TX; hj = j0 + tx - TY; ring1 = true; }
    else if (ty == 3) { hi = i; hj = j0 - 2; }
    else if (ty == 4) { hi = i; hj = j0 + TY + 1; }
    else if (ty == 5 && tx < TY) { hi = i0 - 2; hj = j0 + tx; }
    else if (ty == 5 && tx >= TY && tx < 2 * TY) { hi = i0 + TX + 1; hj = j0 + tx - TY; }
    const bool halo = hi >= 0 && hi < p.Nx && hj >= 0 && hj < p.Ny;
    hx = hi - i0 + 2;
    hy = hj - j0 + 2;
    const int64_t col = (int64_t)i + (int64_t)j * p.sy;
    const int64_t hcol = (int64_t)hi + (int64_t)hj * p.sy;

    // own column along z: coefficients of planes k-1, k, k+1 and the
    // (ro H) u_z products of planes k-1 (gm) and k (gc)
    E2Coef zm{}, c{}, zp{};
    double Sn_c = 0., m_c = 0.;
    double gm[2] = {0., 0.}, gc[2] = {0., 0.};
    auto load = [&](int64_t l) { return e2_coef<CLOSURE>(p, io.Pw[l], io.Sn[l], io.T[l], io.K[l]); };
    auto pres = [&](int kk, double* P) {
        const int64_t l = col + (int64_t)kk * p.sz;
        P[0] = io.Pw[l];
        P[1] = e2_pn<CLOSURE>(p, P[0], io.Sn[l]);
    };
    if (own_in_grid) {
        const int64_t l = col + (int64_t)k0 * p.sz;
        c = load(l);
        Sn_c = io.Sn[l];
        m_c = io.m[l];
        if (hasz) {
            zm = load(l - p.sz);
            double P2[2] = {0., 0.}, Pn1[2];
            if (k0 - 1 != 0) pres(k0 - 2, P2);
            pres(k0 + 1, Pn1);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                gm[a] = zm.roH[a] * (zm.Xi[a] * e2_dP(P2[a], zm.P[a], c.P[a], k0 - 1, p.Nz, p.hz, p.r_hz, p.two_hz,
                                                      p.r_two_hz));
                gc[a] = c.roH[a] * (c.Xi[a] * e2_dP(zm.P[a], c.P[a], Pn1[a], k0, p.Nz, p.hz, p.r_hz, p.two_hz,
                                                    p.r_two_hz));
            }
        }
    }

    for (int k = k0; k <= k1; ++k) {
        // stage plane k
        if (own_in_grid) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                lds.P[a][ty + 2][tx + 2] = c.P[a];
                lds.C[a][ty + 1][tx + 1] = c.Xi[a];
                lds.C[2 + a][ty + 1][tx + 1] = c.ro[a];
                lds.C[4 + a][ty + 1][tx + 1] = c.roH[a];
            }
            lds.C[6][ty + 1][tx + 1] = c.T;
        }
        if (halo) {
            const int64_t lh = hcol + (int64_t)k * p.sz;
            if (ring1) {
                const E2Coef h = load(lh);
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    lds.P[a][hy][hx] = h.P[a];
                    lds.C[a][hy - 1][hx - 1] = h.Xi[a];
                    lds.C[2 + a][hy - 1][hx - 1] = h.ro[a];
                    lds.C[4 + a][hy - 1][hx - 1] = h.roH[a];
                }
                lds.C[6][hy - 1][hx - 1] = h.T;
            } else {
                const double Pw = io.Pw[lh];
                lds.P[0][hy][hx] = Pw;
                lds.P[1][hy][hx] = e2_pn<CLOSURE>(p, Pw, io.Sn[lh]);
            }
        }
        __syncthreads();

        double gp[2] = {0., 0.};
        if (own_in_grid && hasz) {
            const int64_t lp = col + (int64_t)(k + 1) * p.sz;
            zp = load(lp);
            double P2[2] = {0., 0.};
            if (k + 1 != p.Nz - 1) pres(k + 2, P2);
#pragma unroll
            for (int a = 0; a < 2; ++a)
                gp[a] = zp.roH[a] * (zp.Xi[a] * e2_dP(c.P[a], zp.P[a], P2[a], k + 1, p.Nz, p.hz, p.r_hz, p.two_hz,
                                                      p.r_two_hz));
        }

        const int cx = tx + 1, cy = ty + 1;  // place in C; in P: (tx + 2, ty + 2)
        if (active) {
            const int64_t l = col + (int64_t)k * p.sz;
            // mass: roS = ro S - dt/m * div (body_roS; order z, x, y)
            double div[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                double fz = 0.;
                if (hasz) {
                    const double dm = exact_div(c.P[a] - zm.P[a], p.hz, p.r_hz);
                    const double dp = exact_div(zp.P[a] - c.P[a], p.hz, p.r_hz);
                    fz = flux_x(dm, dp, c.Xi[a], c.ro[a], zm.Xi[a], zm.ro[a], zp.Xi[a], zp.ro[a], p.hz, p.r_hz);
                }
                const double Pxm = lds.P[a][ty + 2][tx + 1], Pxp = lds.P[a][ty + 2][tx + 3];
                double dm = exact_div(c.P[a] - Pxm, p.hx, p.r_hx), dp = exact_div(Pxp - c.P[a], p.hx, p.r_hx);
                const double fx = flux_x(dm, dp, c.Xi[a], c.ro[a], lds.C[a][cy][cx - 1], lds.C[2 + a][cy][cx - 1],
                                         lds.C[a][cy][cx + 1], lds.C[2 + a][cy][cx + 1], p.hx, p.r_hx);
                const double Pym = lds.P[a][ty + 1][tx + 2], Pyp = lds.P[a][ty + 3][tx + 2];
                dm = exact_div(c.P[a] - Pym, p.hy, p.r_hy);
                dp = exact_div(Pyp - c.P[a], p.hy, p.r_hy);
                const double fy = flux_y(p, dm, dp, c.Xi[a], c.ro[a], lds.C[a][cy - 1][cx], lds.C[2 + a][cy - 1][cx],
                                         lds.C[a][cy + 1][cx], lds.C[2 + a][cy + 1][cx]);
                div[a] = (fx + fy) + fz;
            }
            const double dtm = p.dt / m_c;
            const double Sw_c = 1. - Sn_c;  // S_calculation
            const double roSw = c.ro[0] * (1. - Sn_c) - dtm * div[0];
            const double roSn = c.ro[1] * Sn_c - dtm * div[1];

            // energy: E (assign_E_current) and E_new (assign_E_new)
            const double Hw = H_w_of(p, c.T), Hn = H_n_T(p, c.T), Hr = H_r_of(p, c.T);
            const double E = (m_c * (Sw_c * (c.roH[0] - c.P[0]) + Sn_c * (c.roH[1] - c.P[1])) +
                              (1. - m_c) * (p.ro_r * Hr - c.P[0]));
            const double lam = p.lam_eff;
            double T_flow = 0;
            T_flow += exact_div(lam * lds.C[6][cy][cx + 1] - 2 * lam * c.T + lam * lds.C[6][cy][cx - 1], p.hx2,
                                p.r_hx2);
            T_flow += exact_div(lam * lds.C[6][cy + 1][cx] - 2 * lam * c.T + lam * lds.C[6][cy - 1][cx], p.hy2,
                                p.r_hy2);
            if (hasz) T_flow += exact_div(lam * zp.T - 2 * lam * c.T + lam * zm.T, p.hz2, p.r_hz2);
            // (ro H) u of the x and y neighbours: u from P two cells out
            double ax[2][2], ay[2][2];  // [phase][minus, plus]
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double Pc = c.P[a];
                const double Px1m = lds.P[a][ty + 2][tx + 1], Px1p = lds.P[a][ty + 2][tx + 3];
                const double dxm = e2_dP(lds.P[a][ty + 2][tx], Px1m, Pc, i - 1, p.Nx, p.hx, p.r_hx, p.two_hx,
                                         p.r_two_hx);
                const double dxp = e2_dP(Pc, Px1p, lds.P[a][ty + 2][tx + 4], i + 1, p.Nx, p.hx, p.r_hx, p.two_hx,
                                         p.r_two_hx);
                ax[a][0] = lds.C[4 + a][cy][cx - 1] * (lds.C[a][cy][cx - 1] * dxm);
                ax[a][1] = lds.C[4 + a][cy][cx + 1] * (lds.C[a][cy][cx + 1] * dxp);
                const double Py1m = lds.P[a][ty + 1][tx + 2], Py1p = lds.P[a][ty + 3][tx + 2];
                const double dym = e2_dP(lds.P[a][ty][tx + 2], Py1m, Pc, j - 1, p.Ny, p.hy, p.r_hy, p.two_hy,
                                         p.r_two_hy);
                const double dyp = e2_dP(Pc, Py1p, lds.P[a][ty + 4][tx + 2], j + 1, p.Ny, p.hy, p.r_hy, p.two_hy,
                                         p.r_two_hy);
                ay[a][0] = lds.C[4 + a][cy - 1][cx] * (lds.C[a][cy - 1][cx] * (dym - lds.C[2 + a][cy - 1][cx] * p.g));
                ay[a][1] = lds.C[4 + a][cy + 1][cx] * (lds.C[a][cy + 1][cx] * (dyp - lds.C[2 + a][cy + 1][cx] * p.g));
            }
            double E_flow = 0;
            E_flow += exact_div(ax[0][1] - ax[0][0] + ax[1][1] - ax[1][0], p.two_hx, p.r_two_hx);
            E_flow += exact_div(ay[0][1] - ay[0][0] + ay[1][1] - ay[1][0], p.two_hy, p.r_two_hy);
            if (hasz) E_flow += exact_div(gp[0] - gm[0] + gp[1] - gm[1], p.two_hz, p.r_two_hz);
            double Q_hw = 0, Q_hr = 0;
            const double E_new = E + p.dt * (T_flow + Q_hw + Q_hr - E_flow);

            // the 4 x 4 Newton (energy.cpp:370-463 without the gas row and column)
            double Sw = Sw_c, Sn = Sn_c, P = c.P[0], T = c.T, hw = Hw, hn = Hn, hr = Hr;
            if (!newton_energy2_t<true>(p, Sw, Sn, P, T, hw, hn, hr, m_c, roSw, roSn, E_new)) {
                // the array's stores stay on this rare branch
                double x[7] = {Sw_c, Sn_c, c.P[0], c.T, Hw, Hn, Hr};
                e2_newton_general(&lp, x, m_c, roSw, roSn, E_new);
                Sn = x[1];
                P = x[2];
                T = x[3];
            }
            io.Pw1[l] = P;
            io.Sn1[l] = Sn;
            io.T1[l] = T;
            e2_store_faces(p, io, i, j, k, l, c.ro[0], c.ro[1], P, Sn, T);
        }
        __syncthreads();
        // roll the z window
        zm = c;
        if (own_in_grid && hasz && k < k1) {
            c = zp;
            const int64_t l = col + (int64_t)(k + 1) * p.sz;
            Sn_c = io.Sn[l];
            m_c = io.m[l];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            gm[a] = gc[a];
            gc[a] = gp[a];
        }
    }
}

template <int CLOSURE>
__global__ void __launch_bounds__(TX * TY) k_energy2_z(DevParams p, E2IO io, int kchunk) {
    __shared__ E2Lds lds;
    __shared__ DevParams lp;
    e2_tile<CLOSURE>(p, io, kchunk, lds, lp);
}

template <auto Kernel>
int resident(int threads) {
    static int cached = 0;
    if (cached) return cached;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, threads, 0) != hipSuccess || per_cu < 1)
        return 512;
    cached = cus * per_cu;
    return cached;
}

template <int CLOSURE>
hipError_t launch_e2(const DevParams& p, const E2IO& io, int kchunk, hipStream_t s, const KernelEvents* ev) {
    const int nx = p.Nx - 2, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    constexpr auto kernel = k_energy2_z<CLOSURE>;
    kchunk = fused_kchunk(p, kchunk, TY, resident<kernel>(TX * TY));
    const dim3 grid((nx + TX - 1) / TX, (ny + TY - 1) / TY, (nz + kchunk - 1) / kchunk);
    hipError_t e;
    if (ev && (e = hipEventRecord(ev->begin, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, grid, dim3(TX, TY), 0, s, p, io, kchunk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return ev ? hipEventRecord(ev->end, s) : hipSuccess;
}

}  // namespace

hipError_t launch_energy2_step(const DevParams& p, const double* const* cur, double* const* next, int kchunk,
                               hipStream_t s, const KernelEvents* ev) {
    if (p.model != 2 || !p.energy || p.Nx < 3) return hipErrorInvalidValue;
    E2IO io{cur[0], cur[1], cur[2], cur[3], cur[4], next[0], next[1], next[2]};
    if (p.closure == 1 && p.corey_n == 2.) return launch_e2<2>(p, io, kchunk, s, ev);
    if (p.closure == 1) return launch_e2<1>(p, io, kchunk, s, ev);
    return launch_e2<0>(p, io, kchunk, s, ev);
}

}  // namespace capaz
