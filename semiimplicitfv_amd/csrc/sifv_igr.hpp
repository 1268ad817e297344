// sifv_igr.hpp - information geometric regularisation (IGR) on the GPU:
// the entropic pressure sigma solves
//     sigma/rho - alpha div(grad(sigma)/rho) = alpha [tr((grad u)^2) + tr(grad u)^2]
// by in-place lexicographic Gauss-Seidel sweeps with a scalar ghost refresh
// after each sweep (IGRSolver::solveEntropicPressure, src/IGR.cpp:37-158;
// computeIGRRhs :18-35; velocity gradients src/ExplicitSolver.cpp:215-267).
//
// igr_pre_kernel   per interior cell: the right-hand side from the central
//                  velocity gradients, the diagonal (1/rho + sum of face
//                  terms) and rho itself, all from the stage's conservative
//                  buffer through the ghost mapping.
// igr_sweep_kernel one Gauss-Seidel sweep in exactly the reference's order.
//                  Cell (i,j,k) needs the new values of (i-1,j,k), (i,j-1,k),
//                  (i,j,k-1) and the old values of the +1 neighbours, so the
//                  sweep is a wavefront.  The domain is cut into tiles of 64
//                  columns x tj (<= kIgrTj) rows of one plane; a wave owns a
//                  tile and walks its anti-diagonals (lane l <-> column, row =
//                  step - l: the left neighbour's new value comes from lane
//                  l-1's previous step, the lower neighbour's from the lane's
//                  own previous step).  Before the walk the wave stages in LDS
//                  every operand that does not depend on this sweep's new
//                  values (face densities, the old +x/+y quotients, rhs,
//                  diagonal) -- before it waits for its upstream tiles -- and,
//                  after the wait, the new values of the left column, the row
//                  below and the plane below, so the walk's dependent chain
//                  touches LDS only.  Tiles start when their three upstream
//                  tiles have published the current epoch in a flag array;
//                  every block handles tiles in topological order and the grid
//                  never exceeds the co-resident block count, so the lowest
//                  unfinished tile can always proceed.  Waits are bounded: a
//                  wait that exceeds its budget raises an error flag and every
//                  wave exits.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

template <typename T>
struct IgrPreArgs {
    StageParams<T> P;   // grid, EOS, mixture, source buffer (P.src), total
    double alpha;
    double* rhs;
    double* diag;
    double* rho;        // padded; interior written here, ghosts by the scalar fill
};

template <int DIM, int NPH, typename T>
__global__ void __launch_bounds__(256) igr_pre_kernel(IgrPreArgs<T> A) {
    const StageParams<T>& P = A.P;
    const Grid& g = P.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const double alpha = A.alpha;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        Cell<DIM, NPH> c, m, p;
        load_cell<DIM, NPH>(P, P.src, i, j, k, c);
        // velocity gradient g[a][b] = du_a/dx_b (central, PRIM ghosts)
        double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double rhoM[3], rhoP[3];
        const double h[3] = {g.dxw[i + g.ngx], DIM >= 2 ? g.dyw[j + g.ngy] : 1.0, DIM >= 3 ? g.dzw[k + g.ngz] : 1.0};
#pragma unroll
        for (int b = 0; b < DIM; ++b) {
            load_cell<DIM, NPH>(P, P.src, i - (b == 0), j - (b == 1), k - (b == 2), m);
            load_cell<DIM, NPH>(P, P.src, i + (b == 0), j + (b == 1), k + (b == 2), p);
            const double inv = 0.5 / h[b];
#pragma unroll
            for (int a = 0; a < DIM; ++a) G[a][b] = (p.u[a] - m.u[a]) * inv;
            rhoM[b] = m.rho;
            rhoP[b] = p.rho;
        }
        double trSq = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = 0; b < DIM; ++b) trSq += G[a][b] * G[b][a];
        double tr = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; ++a) tr += G[a][a];
        tr *= tr;
        A.rhs[id] = alpha * (trSq + tr);
        const double ri = c.rho;
        double diag = 1.0 / ri;
#pragma unroll
        for (int b = 0; b < DIM; ++b) {
            const double fl = 0.5 * (ri + rhoM[b]), fr = 0.5 * (ri + rhoP[b]);
            const double d2 = 1.0 / (h[b] * h[b]);
            diag += alpha * d2 * (1.0 / fl + 1.0 / fr);
        }
        A.diag[id] = diag;
        A.rho[id] = ri;
    }
}

struct IgrSweepArgs {
    Grid g;
    double* sg;           // sigma, padded; ghosts hold the previous sweep's fill
    const double* rho;    // padded, 1-deep ghosts by the scalar rule
    const double* rhs;
    const double* diag;
    double alpha;
    int ntx, nty, tj;     // tiles: ntx x nty x nz
    int ntiles;
    const int* order;     // tile ids in topological (anti-diagonal) order
    int* flags;           // per tile: last finished epoch
    int epoch;
    int* err;             // set when a wait exceeds its budget
    const int* active;    // device time-loop flag (null: always run)
};

__device__ __forceinline__ bool igr_wait(const IgrSweepArgs& A, int t) {
    if (t < 0) return true;
    int spins = 0;
    while (__hip_atomic_load(&A.flags[t], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < A.epoch) {
        if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (++spins > (1 << 25)) {
            __hip_atomic_store(A.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(8);
    }
    return true;
}

constexpr int kIgrTj = 16;   // most rows per sweep tile (LDS staging)
// The walk multiplies by the staged reciprocals of the face densities and of
// the diagonal instead of dividing (one rounding per operand more, within the
// parity bars; the divisions sat on the dependent chain).
#ifndef SIFV_IGR_RCP
#define SIFV_IGR_RCP 1
#endif

template <int DIM>
__global__ void __launch_bounds__(64) igr_sweep_kernel(IgrSweepArgs A) {
    if (A.active && *A.active == 0) return;
    // per-cell operands of the tile, [row][lane]
    __shared__ double sFl[kIgrTj][kWave];    // x face density to the left
    __shared__ double sTx[kIgrTj][kWave];    // old sigma(+x) / right face density
    __shared__ double sFly[DIM >= 2 ? kIgrTj : 1][kWave];
    __shared__ double sTy[DIM >= 2 ? kIgrTj : 1][kWave];
    __shared__ double sQz[DIM >= 3 ? kIgrTj : 1][kWave];   // both z quotients
    __shared__ double sRhs[kIgrTj][kWave];
    __shared__ double sDiag[kIgrTj][kWave];
    __shared__ double sLeft[kIgrTj];         // new sigma of the column left of the tile
    __shared__ double sBot[kWave];           // new sigma of the row below the tile
    __shared__ double sD2y[kIgrTj];
    const Grid& g = A.g;
    const int lane = threadIdx.x;
    const size_t sy = (size_t)g.nxT, sz = (size_t)g.nxT * g.nyT;
    const double alpha = A.alpha;
    for (int q = blockIdx.x; q < A.ntiles; q += gridDim.x) {
        const int t = A.order[q];
        const int I = t % A.ntx;
        const int rest = t / A.ntx;
        const int J = rest % A.nty;
        const int K = rest / A.nty;
        const int i0 = I * kWave;
        const int i = i0 + lane;
        const int j0 = J * A.tj;
        const int nrows = (j0 + A.tj <= g.ny) ? A.tj : g.ny - j0;
        const bool colOk = i < g.nx;
        const int ci = colOk ? i : g.nx - 1;
        const double dx = g.dxw[ci + g.ngx];
        const double d2x = 1.0 / (dx * dx);
        const double d2z = DIM >= 3 ? 1.0 / (g.dzw[K + g.ngz] * g.dzw[K + g.ngz]) : 0.0;
        // operands that this sweep's new values do not change: the +x / +y
        // neighbours are in this tile or in tiles that wait for it
#pragma unroll 4
        for (int jj = 0; jj < kIgrTj; ++jj) {
            if (jj < nrows && colOk) {
                const size_t id = cell_index(g, i, j0 + jj, K);
                const double ri = A.rho[id];
#if SIFV_IGR_RCP
                sFl[jj][lane] = 1.0 / (0.5 * (ri + A.rho[id - 1]));
#else
                sFl[jj][lane] = 0.5 * (ri + A.rho[id - 1]);
#endif
                sTx[jj][lane] = A.sg[id + 1] / (0.5 * (ri + A.rho[id + 1]));
                if (DIM >= 2) {
#if SIFV_IGR_RCP
                    sFly[jj][lane] = 1.0 / (0.5 * (ri + A.rho[id - sy]));
#else
                    sFly[jj][lane] = 0.5 * (ri + A.rho[id - sy]);
#endif
                    sTy[jj][lane] = A.sg[id + sy] / (0.5 * (ri + A.rho[id + sy]));
                }
                sRhs[jj][lane] = A.rhs[id];
#if SIFV_IGR_RCP
                sDiag[jj][lane] = 1.0 / A.diag[id];
#else
                sDiag[jj][lane] = A.diag[id];
#endif
            }
        }
        if (DIM >= 2 && lane < nrows) {
            const double hy = g.dyw[j0 + lane + g.ngy];
            sD2y[lane] = 1.0 / (hy * hy);
        }
        bool ok = igr_wait(A, I > 0 ? t - 1 : -1);
        ok = ok && igr_wait(A, J > 0 ? t - A.ntx : -1);
        ok = ok && igr_wait(A, K > 0 ? t - A.ntx * A.nty : -1);
        if (!ok) return;
        // new values published by the upstream tiles (ghosts at domain edges)
        if (lane < nrows) sLeft[lane] = A.sg[cell_index(g, i0 - 1, j0 + lane, K)];
        if (DIM >= 2 && colOk) sBot[lane] = A.sg[cell_index(g, i, j0 - 1, K)];
        if (DIM >= 3) {
#pragma unroll 4
            for (int jj = 0; jj < kIgrTj; ++jj) {
                if (jj < nrows && colOk) {
                    const size_t id = cell_index(g, i, j0 + jj, K);
                    const double ri = A.rho[id];
                    const double fl = 0.5 * (ri + A.rho[id - sz]), fr = 0.5 * (ri + A.rho[id + sz]);
                    sQz[jj][lane] = A.sg[id - sz] / fl + A.sg[id + sz] / fr;
                }
            }
        }
        __syncthreads();
        double last = 0.0;
        for (int s = 0; s < nrows + kWave - 1; ++s) {
            const int jj = s - lane;
            const double fromLeft = __shfl_up(last, 1);
            if (colOk && jj >= 0 && jj < nrows) {
                double off = 0.0;
                {
                    const double sl = lane == 0 ? sLeft[jj] : fromLeft;
#if SIFV_IGR_RCP
                    off += alpha * d2x * (sl * sFl[jj][lane] + sTx[jj][lane]);
#else
                    off += alpha * d2x * (sl / sFl[jj][lane] + sTx[jj][lane]);
#endif
                }
                if (DIM >= 2) {
                    const double sl = jj == 0 ? sBot[lane] : last;
#if SIFV_IGR_RCP
                    off += alpha * sD2y[jj] * (sl * sFly[jj][lane] + sTy[jj][lane]);
#else
                    off += alpha * sD2y[jj] * (sl / sFly[jj][lane] + sTy[jj][lane]);
#endif
                }
                if (DIM >= 3) off += alpha * d2z * sQz[jj][lane];
#if SIFV_IGR_RCP
                last = (sRhs[jj][lane] + off) * sDiag[jj][lane];
#else
                last = (sRhs[jj][lane] + off) / sDiag[jj][lane];
#endif
                A.sg[cell_index(g, i, j0 + jj, K)] = last;
            }
        }
        // publish the tile: data before flag
        __threadfence();
        __syncthreads();
        if (lane == 0) __hip_atomic_store(&A.flags[t], A.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace sifv
