// sifv_stage.hpp - fused cons->prim + ghost mapping + reconstruction +
// Riemann flux + flux divergence + SSP-RK stage update for 1D and 2D grids.
//
// Replaces, per RK stage, the reference's chain
//   convertConservativeToPrimitiveVariables (src/SolutionState.cpp:187-205)
//   -> applyBoundaryConditions(PRIM)          (src/RectilinearMesh.cpp:498-524)
//   -> Reconstructor::reconstructX/Y          (src/Reconstruction.cpp:204-432)
//   -> computeRHS face loops                   (src/ExplicitSolver.cpp:270-419)
//   -> stage update                            (src/ExplicitSolver.cpp:127-188)
// with one kernel that reads each conservative value once from HBM.
//
// Work decomposition (wave64): one wave owns a strip of 62 consecutive
// columns plus one halo column on each side (lane l <-> column X0-1+l) and
// marches upward through `rows` consecutive rows.  The y-stencil of each
// lane's column lives in a register window that slides one row per
// iteration; the x-stencil is exchanged through LDS.  Each lane reconstructs
// its own cell once per direction (both face states from one stencil), so
// every face flux is computed exactly once per strip.  Ghost cells are never
// materialised: stencil indices outside the domain are mapped to their
// source cells with the boundary rule and velocity sign flips applied.
#pragma once

#include "sifv_device.hpp"

namespace sifv {

constexpr int kOwn = 62;  // owned columns per wave

template <typename T>
struct StageParams {
    Grid g;
    Eos eos;
    const T* src[5];      // stencil source: rho, mom[DIM], E
    const T* u0[5];       // RK backup, read pointwise (unused when c2 == 0)
    T* dst[5];            // stage output
    const double* sigma;  // entropic pressure (LF with IGR only)
    double c1, c2, c3, c4;
    int rk1;              // RKOrder == 1: U += dt * R
    const double* dt;     // device-resident time step
    const int* active;    // device flag; 0 skips the stage (time loop finished)
    double* dtPartial;    // per-block min of min_d dx_d/(|u_d|+c) of the new state, or null
    double eps;
    int incP, igr, weno;
    int rows;             // rows marched per wave
    int nseg;             // y segments per strip
};

// Logical block id with all blocks of one XCD contiguous (blocks are
// dispatched round-robin over the 8 XCDs of an MI355X), so neighbouring
// strips/segments that share halo data share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int nx = 8;
    int per = nb / nx, rem = nb % nx;
    int x = b % nx, q = b / nx;
    return (x < rem) ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

// Load one cell's conservative state (through the ghost mapping) and return
// primitive (rho, u[DIM], p) in FP64.
template <int DIM, typename T>
__device__ __forceinline__ void load_prim(const StageParams<T>& P, int i, int j, double* w) {
    const Grid& g = P.g;
    bool fnx = false, ftx = false, fny = false, fty = false;
    int ci = i < -g.ngx ? -g.ngx : (i > g.nx - 1 + g.ngx ? g.nx - 1 + g.ngx : i);
    int si = map_axis(ci, g.nx, g.bc[0], g.bc[1], fnx, ftx);
    int sj = 0;
    if (DIM >= 2) {
        int cj = j < -g.ngy ? -g.ngy : (j > g.ny - 1 + g.ngy ? g.ny - 1 + g.ngy : j);
        sj = map_axis(cj, g.ny, g.bc[2], g.bc[3], fny, fty);
    }
    size_t id = cell_index(g, si, sj, 0);
    double rho = (double)P.src[0][id];
    double mom[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) mom[d] = (double)P.src[1 + d][id];
    double E = (double)P.src[DIM + 1][id];
    double u[DIM], p;
    cons_to_prim<DIM>(P.eos, rho, mom, E, u, p);
    w[0] = rho;
    // velocity sign flips of wall rules: component d flips with the normal
    // flag of axis d and the tangential flags of the other axes
    bool fl[2] = {fnx != fty, ftx != fny};
#pragma unroll
    for (int d = 0; d < DIM; ++d) w[1 + d] = fl[d] ? -u[d] : u[d];
    w[DIM + 1] = p;
}

template <int DIM, typename T>
__device__ __forceinline__ double load_sigma(const StageParams<T>& P, int i, int j) {
    const Grid& g = P.g;
    int ci = i < -g.ngx ? -g.ngx : (i > g.nx - 1 + g.ngx ? g.nx - 1 + g.ngx : i);
    // sigma ghosts are refreshed by the IGR solve's scalar fill
    int si = map_axis_scalar(ci, g.nx, g.sbc[0], g.sbc[1]);
    int cj = j < -g.ngy ? -g.ngy : (j > g.ny - 1 + g.ngy ? g.ny - 1 + g.ngy : j);
    int sj = DIM >= 2 ? map_axis_scalar(cj, g.ny, g.sbc[2], g.sbc[3]) : 0;
    return P.sigma[cell_index(g, si, sj, 0)];
}

template <int DIM, int R, int RS, typename T>
__global__ void __launch_bounds__(64) stage_xy_kernel(StageParams<T> P) {
    constexpr int NP = DIM + 2;  // primitive variables rho, u[DIM], p
    constexpr int NC = DIM + 2;  // conservative variables
    constexpr int E = R - 1;     // extra LDS columns per side
    constexpr int S = 2 * R - 1; // stencil width
    constexpr int WLO = (R >= 2) ? (2 - R) : 0;          // window start, relative to r
    constexpr int WS = R - WLO + 1;                      // window rows
    constexpr int WROW = -WLO;                           // window slot of row r
    constexpr int WST = (R >= 2) ? 0 : 1;                // first stencil slot of row r+1

    if (P.active && *P.active == 0) return;
    const Grid& g = P.g;
    const int lane = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int strip = blk / P.nseg;
    const int seg = blk - strip * P.nseg;
    const int X0 = strip * kOwn;
    const int col = X0 - 1 + lane;
    const int Y0 = (DIM >= 2) ? seg * P.rows : 0;
    const int Y1 = (DIM >= 2) ? ((Y0 + P.rows < g.ny) ? Y0 + P.rows : g.ny) : 1;
    const bool owned = lane >= 1 && lane <= kOwn && col < g.nx;
    const bool useSig = (RS == RS_LF) && P.igr;

    __shared__ T sp[NP + 1][kWave + 2 * E];

    // extra column handled by this lane (lanes 0..2E-1)
    int ecol = 0;
    if (lane < E) ecol = X0 - 1 - E + lane;
    else if (lane < 2 * E) ecol = X0 + kWave - 1 + (lane - E);

    const T eps = (T)P.eps;
    const T gam = (T)P.eos.gamma;
    const T pinf = (T)(P.eos.stiffened ? P.eos.pInf : 0.0);
    const bool incP = P.incP != 0;
    const bool weno = P.weno != 0;
    const double dt = *P.dt;

    // geometry of the lane's column
    const int cc = col < g.nx ? col : g.nx - 1;
    const double dxc = g.dxw[(cc < 0 ? 0 : cc) + g.ngx];

    double dtMin = 1.0e300;

    // y machinery (DIM == 2)
    T win[WS > 0 ? WS : 1][NP];
    T sigwin[WS > 0 ? WS : 1];
    FaceState<DIM, T> plusPrev;
    T Glo[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) Glo[q] = T(0);

    int rStart = Y0;
    if (DIM >= 2) {
        rStart = Y0 - 2;
        // preload the window as it stands after iteration rStart-1 (slot 0
        // is slid out before it is read, so it is left empty)
#pragma unroll
        for (int w = 1; w < WS; ++w) {
            double pr[NP];
            const int row = rStart - 1 + WLO + w;
            load_prim<DIM>(P, col, row, pr);
#pragma unroll
            for (int v = 0; v < NP; ++v) win[w][v] = (T)pr[v];
            sigwin[w] = useSig ? (T)load_sigma<DIM>(P, col, row) : T(0);
        }
    }

    for (int r = rStart; r < Y1; ++r) {
        T rhs[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) rhs[q] = T(0);
        double cy = 0.0;
        if (DIM >= 2) {
            // slide the window and load row r+R
#pragma unroll
            for (int w = 0; w < WS - 1; ++w) {
#pragma unroll
                for (int v = 0; v < NP; ++v) win[w][v] = win[w + 1][v];
                sigwin[w] = sigwin[w + 1];
            }
            {
                double pr[NP];
                load_prim<DIM>(P, col, r + R, pr);
#pragma unroll
                for (int v = 0; v < NP; ++v) win[WS - 1][v] = (T)pr[v];
                sigwin[WS - 1] = useSig ? (T)load_sigma<DIM>(P, col, r + R) : T(0);
            }
            // y reconstruction of row r+1
            FaceState<DIM, T> mY, pY;
            {
                T st[S], mn, pl;
#pragma unroll
                for (int v = 0; v < NP; ++v) {
#pragma unroll
                    for (int s = 0; s < S; ++s) st[s] = win[WST + s][v];
                    Recon<R, T>::run(st, eps, weno, mn, pl);
                    if (v == 0) { mY.rho = mn; pY.rho = pl; }
                    else if (v == NP - 1) { mY.p = mn; pY.p = pl; }
                    else { mY.u[v - 1] = mn; pY.u[v - 1] = pl; }
                }
                if (useSig) {
#pragma unroll
                    for (int s = 0; s < S; ++s) st[s] = sigwin[WST + s];
                    Recon<R, T>::run(st, eps, weno, mn, pl);
                    mY.sig = mn;
                    pY.sig = pl;
                } else {
                    mY.sig = T(0);
                    pY.sig = T(0);
                }
            }
            T Ghi[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) Ghi[q] = T(0);
            if (r >= Y0 - 1) {
                T F[DIM + 4];
                riemann<RS, 1, DIM, T>(plusPrev, mY, gam, pinf, incP, P.igr != 0, F);
#pragma unroll
                for (int q = 0; q < NC; ++q) Ghi[q] = F[q];
            }
            plusPrev = pY;
            if (r < Y0) {
#pragma unroll
                for (int q = 0; q < NC; ++q) Glo[q] = Ghi[q];
                continue;  // warm-up rows: no x phase, no output
            }
            // y contribution (added after the x contribution below)
            const double dyr = g.dyw[r + g.ngy];
            cy = (dxc * 1.0) / ((dxc * dyr) * 1.0);
#pragma unroll
            for (int q = 0; q < NC; ++q) rhs[q] = Ghi[q];  // stash G_{r+1/2}
        }

        // ---------------- x phase on row r ----------------
        double w0[NP];
        T sig0 = T(0);
        if (DIM >= 2) {
#pragma unroll
            for (int v = 0; v < NP; ++v) w0[v] = (double)win[WROW][v];
            sig0 = sigwin[WROW];
        } else {
            load_prim<DIM>(P, col, r, w0);
            if (useSig) sig0 = (T)load_sigma<DIM>(P, col, r);
        }
        __syncthreads();  // previous iteration's LDS reads are complete
#pragma unroll
        for (int v = 0; v < NP; ++v) sp[v][lane + E] = (T)w0[v];
        if (useSig) sp[NP][lane + E] = sig0;
        if (E > 0 && lane < 2 * E) {
            double we[NP];
            load_prim<DIM>(P, ecol, r, we);
            int pos = lane < E ? lane : kWave + lane;
#pragma unroll
            for (int v = 0; v < NP; ++v) sp[v][pos] = (T)we[v];
            if (useSig) sp[NP][pos] = (T)load_sigma<DIM>(P, ecol, r);
        }
        __syncthreads();
        FaceState<DIM, T> mX, pX;
        {
            T st[S], mn, pl;
#pragma unroll
            for (int v = 0; v < NP; ++v) {
#pragma unroll
                for (int s = 0; s < S; ++s) st[s] = sp[v][lane + s];
                Recon<R, T>::run(st, eps, weno, mn, pl);
                if (v == 0) { mX.rho = mn; pX.rho = pl; }
                else if (v == NP - 1) { mX.p = mn; pX.p = pl; }
                else { mX.u[v - 1] = mn; pX.u[v - 1] = pl; }
            }
            if (useSig) {
#pragma unroll
                for (int s = 0; s < S; ++s) st[s] = sp[NP][lane + s];
                Recon<R, T>::run(st, eps, weno, mn, pl);
                mX.sig = mn;
                pX.sig = pl;
            } else {
                mX.sig = T(0);
                pX.sig = T(0);
            }
        }
        // face col+1/2: left = plus of this lane, right = minus of lane+1
        FaceState<DIM, T> rX;
        rX.rho = __shfl_down(mX.rho, 1, kWave);
#pragma unroll
        for (int d = 0; d < DIM; ++d) rX.u[d] = __shfl_down(mX.u[d], 1, kWave);
        rX.p = __shfl_down(mX.p, 1, kWave);
        rX.sig = __shfl_down(mX.sig, 1, kWave);
        T Fx[DIM + 4];
        riemann<RS, 0, DIM, T>(pX, rX, gam, pinf, incP, P.igr != 0, Fx);
        T Fl[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) Fl[q] = __shfl_up(Fx[q], 1, kWave);

        if (owned) {
            const double dyr = (DIM >= 2) ? g.dyw[r + g.ngy] : g.dyw[0];
            const double cx = (dyr * 1.0) / ((dxc * dyr) * 1.0);
            T res[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                T acc = (T)cx * Fl[q];
                acc = acc - (T)cx * Fx[q];
                if (DIM >= 2) {
                    acc = acc + (T)cy * Glo[q];
                    acc = acc - (T)cy * rhs[q];
                }
                res[q] = acc;
            }
            size_t id = cell_index(g, col, r, 0);
            double un[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                double U = (double)P.src[q][id];
                double Rq = (double)res[q];
                if (P.rk1) {
                    un[q] = U + dt * Rq;
                } else {
                    double a = P.c1 * U;
                    if (P.c2 != 0.0) a = a + P.c2 * (double)P.u0[q][id];
                    un[q] = (a + (P.c3 * dt) * Rq) / P.c4;
                }
                P.dst[q][id] = (T)un[q];
            }
            if (P.dtPartial) {
                double rho = (double)(T)un[0];
                double mom[DIM], u[DIM], p;
#pragma unroll
                for (int d = 0; d < DIM; ++d) mom[d] = (double)(T)un[1 + d];
                cons_to_prim<DIM>(P.eos, rho, mom, (double)(T)un[DIM + 1], u, p);
                double pinfd = P.eos.stiffened ? P.eos.pInf : 0.0;
                double c = sqrt(P.eos.gamma * dmax_ref(p + pinfd, 1e-14) / dmax_ref(rho, 1e-14));
                double dc = dxc / (fabs(u[0]) + c);
                if (DIM >= 2) dc = dmin_ref(dc, dyr / (fabs(u[1]) + c));
                dtMin = dmin_ref(dtMin, dc);
            }
        }
        if (DIM >= 2) {
#pragma unroll
            for (int q = 0; q < NC; ++q) Glo[q] = rhs[q];
        }
    }
    if (P.dtPartial) {
        double m = wave_min(dtMin);
        if (lane == 0) P.dtPartial[blockIdx.x] = m;
    }
}

}  // namespace sifv
