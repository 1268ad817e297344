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
//                  columns x tj rows of one plane; a wave owns a tile and
//                  walks its anti-diagonals (lane l <-> column, row = step - l:
//                  the left neighbour's new value comes from lane l-1's previous
//                  step, the lower neighbour's from the lane's own previous
//                  step).  Tiles start when their three upstream tiles have
//                  published the current epoch in a flag array; every block
//                  handles tiles in topological order, so the lowest unfinished
//                  tile can always proceed.  Waits are bounded: a wait that
//                  exceeds its budget raises an error flag and every wave exits.
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

template <int DIM>
__global__ void __launch_bounds__(64) igr_sweep_kernel(IgrSweepArgs A) {
    if (A.active && *A.active == 0) return;
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
        bool ok = igr_wait(A, I > 0 ? t - 1 : -1);
        ok = ok && igr_wait(A, J > 0 ? t - A.ntx : -1);
        ok = ok && igr_wait(A, K > 0 ? t - A.ntx * A.nty : -1);
        if (!ok) return;
        const int i = I * kWave + lane;
        const int j0 = J * A.tj;
        const int nrows = (j0 + A.tj <= g.ny) ? A.tj : g.ny - j0;
        const bool colOk = i < g.nx;
        const int ci = colOk ? i : g.nx - 1;
        const double dx = g.dxw[ci + g.ngx];
        const double d2x = 1.0 / (dx * dx);
        const double d2z = DIM >= 3 ? 1.0 / (g.dzw[K + g.ngz] * g.dzw[K + g.ngz]) : 0.0;
        double last = 0.0;
        for (int s = 0; s < nrows + kWave - 1; ++s) {
            const int jj = s - lane;
            const double fromLeft = __shfl_up(last, 1);
            if (colOk && jj >= 0 && jj < nrows) {
                const int j = j0 + jj;
                const size_t id = cell_index(g, i, j, K);
                const double ri = A.rho[id];
                double off = 0.0;
                {
                    const double sl = lane == 0 ? A.sg[id - 1] : fromLeft;
                    const double sr = A.sg[id + 1];
                    const double fl = 0.5 * (ri + A.rho[id - 1]), fr = 0.5 * (ri + A.rho[id + 1]);
                    off += alpha * d2x * (sl / fl + sr / fr);
                }
                if (DIM >= 2) {
                    const double sl = jj == 0 ? A.sg[id - sy] : last;
                    const double sr = A.sg[id + sy];
                    const double fl = 0.5 * (ri + A.rho[id - sy]), fr = 0.5 * (ri + A.rho[id + sy]);
                    const double hy = g.dyw[j + g.ngy];
                    const double d2 = 1.0 / (hy * hy);
                    off += alpha * d2 * (sl / fl + sr / fr);
                }
                if (DIM >= 3) {
                    const double sl = A.sg[id - sz];
                    const double sr = A.sg[id + sz];
                    const double fl = 0.5 * (ri + A.rho[id - sz]), fr = 0.5 * (ri + A.rho[id + sz]);
                    off += alpha * d2z * (sl / fl + sr / fr);
                }
                last = (A.rhs[id] + off) / A.diag[id];
                A.sg[id] = last;
            }
        }
        // publish the tile: data before flag
        __threadfence();
        __syncthreads();
        if (lane == 0) __hip_atomic_store(&A.flags[t], A.epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace sifv
