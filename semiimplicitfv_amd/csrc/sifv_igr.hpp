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
// igr_sweep_kernel the Gauss-Seidel sweeps of one solve in exactly the
//                  reference's order.
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
//                  touches LDS only.  All the sweeps of a solve run in one
//                  launch, pipelined: work item (s, tile) starts when the
//                  tile's three upstream tiles have finished sweep s, the
//                  tile itself and its three downstream tiles sweep s-1, and,
//                  on a periodic axis, the far-end tile of its line sweep s-1
//                  (each tile publishes its finished sweep count in a flag
//                  array).  Boundary tiles keep the 1-deep scalar ghosts as the
//                  reference's fill after every sweep leaves them: a zero-
//                  gradient ghost is written by the tile that owns its source
//                  cell as it computes it; on a periodic axis the high tile
//                  writes the low ghost as it computes it and the low tile
//                  writes its new values into the parity slot of a side array
//                  that the high tile reads in the next sweep.  Every block
//                  takes work items in (sweep, topological tile) order and the
//                  grid never exceeds the
//                  co-resident block count; every dependency is on an earlier
//                  item, so the lowest unfinished item can always proceed.
//                  Waits are bounded: a wait that exceeds its budget raises an
//                  error flag and every wave exits.
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
    int* flags;           // per tile: sweeps finished (cumulative over launches)
    int epoch;            // flag value of every tile when the launch starts
    int nsweeps;          // sweeps run by this launch
    int* err;             // set when a wait exceeds its budget
    int* next;            // work-item counter, zero at launch (SIFV_IGR_DYN)
    // wrapping high faces: the first cells of each line as the previous sweep
    // left them, two slots by sweep parity ([2][nz][ny], [2][nz][nx], [2][ny][nx])
    double* wrapX;
    double* wrapY;
    double* wrapZ;
    const int* active;    // device time-loop flag (null: always run)
};

// Flags are polled with relaxed loads and each group of waits ends in one
// acquire fence: an acquire at agent scope invalidates the XCD's L2, which an
// acquiring poll loop would do on every iteration (measured 1.8x slower).
// Blocks take work items by a fixed stride of the grid (0) or from a counter
// in item order (1; 3 % slower on the 2D vortex at 1024^2 and 2048^2).
#ifndef SIFV_IGR_DYN
#define SIFV_IGR_DYN 0
#endif

__device__ __forceinline__ void igr_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ bool igr_wait(const IgrSweepArgs& A, int t, int epoch) {
    if (t < 0) return true;
    int spins = 0;
    while (__hip_atomic_load(&A.flags[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
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
    const int ntz = DIM >= 3 ? g.nz : 1;
    const int plane = A.ntx * A.nty;
    // per face: does the scalar fill wrap (ghost from the far end of the line)?
    const bool loX = g.sbc[0] == BK_WRAP, hiX = g.sbc[1] == BK_WRAP;
    const bool loY = DIM >= 2 && g.sbc[2] == BK_WRAP, hiY = DIM >= 2 && g.sbc[3] == BK_WRAP;
    const bool loZ = DIM >= 3 && g.sbc[4] == BK_WRAP, hiZ = DIM >= 3 && g.sbc[5] == BK_WRAP;
    const int nitems = A.nsweeps * A.ntiles;
#if SIFV_IGR_DYN
    for (;;) {
        // every lane takes part in the grab (lane 0 adds 1, the others 0) and
        // in the flag store below: no lane-dependent branch may separate one
        // item from the next, or the compiler may run the lanes' loops apart
        // and read the item from a lane that did not take it
        const int q = __builtin_amdgcn_readfirstlane(
            __hip_atomic_fetch_add(A.next, lane == 0 ? 1 : 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (q >= nitems) break;
#else
    for (int q = blockIdx.x; q < nitems; q += gridDim.x) {
#endif
        const int sw = q / A.ntiles;
        const int t = A.order[q - sw * A.ntiles];
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
        const int now = A.epoch + sw + 1, prev = A.epoch + sw;
        // sweep sw > 0 follows the tile's own sweep sw-1 (another block may
        // have run it), reads the downstream tiles' values of sweep sw-1, and a
        // low tile on a wrapping axis replaces the high ghost the far-end tile
        // read in sweep sw-1 (sweep 0 is ordered by the launch itself)
        bool ok = true;
        if (sw > 0) {
            ok = igr_wait(A, t, prev);
            ok = ok && igr_wait(A, I < A.ntx - 1 ? t + 1 : -1, prev);
            ok = ok && igr_wait(A, DIM >= 2 && J < A.nty - 1 ? t + A.ntx : -1, prev);
            ok = ok && igr_wait(A, DIM >= 3 && K < ntz - 1 ? t + plane : -1, prev);
            ok = ok && igr_wait(A, (loX || hiX) && I == 0 && A.ntx > 1 ? t + A.ntx - 1 : -1, prev);
            ok = ok && igr_wait(A, (loY || hiY) && J == 0 && A.nty > 1 ? t + (A.nty - 1) * A.ntx : -1, prev);
            ok = ok && igr_wait(A, (loZ || hiZ) && K == 0 && ntz > 1 ? t + (ntz - 1) * plane : -1, prev);
            if (!ok) return;
            igr_acquire();
        }
        // wrapping high faces read the first cells of the line as the previous
        // sweep left them: from the ghost the fill before the launch wrote, and
        // later from the side slot the low tile filled in sweep sw-1
        const int par = (A.epoch + sw) & 1, npar = par ^ 1;
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
                double sr = A.sg[id + 1];
                if (hiX && sw > 0 && i == g.nx - 1) sr = A.wrapX[((size_t)par * ntz + K) * g.ny + j0 + jj];
                sTx[jj][lane] = sr / (0.5 * (ri + A.rho[id + 1]));
                if (DIM >= 2) {
#if SIFV_IGR_RCP
                    sFly[jj][lane] = 1.0 / (0.5 * (ri + A.rho[id - sy]));
#else
                    sFly[jj][lane] = 0.5 * (ri + A.rho[id - sy]);
#endif
                    double st = A.sg[id + sy];
                    if (hiY && sw > 0 && j0 + jj == g.ny - 1) st = A.wrapY[((size_t)par * ntz + K) * g.nx + i];
                    sTy[jj][lane] = st / (0.5 * (ri + A.rho[id + sy]));
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
        ok = igr_wait(A, I > 0 ? t - 1 : -1, now);
        ok = ok && igr_wait(A, DIM >= 2 && J > 0 ? t - A.ntx : -1, now);
        ok = ok && igr_wait(A, DIM >= 3 && K > 0 ? t - plane : -1, now);
        if (!ok) return;
        igr_acquire();
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
                    double su = A.sg[id + sz];
                    if (hiZ && sw > 0 && K == g.nz - 1) su = A.wrapZ[((size_t)par * g.ny + j0 + jj) * g.nx + i];
                    sQz[jj][lane] = A.sg[id - sz] / fl + su / fr;
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
                const size_t id = cell_index(g, i, j0 + jj, K);
                A.sg[id] = last;
                // ghosts of the boundary cells as the scalar fill after this
                // sweep leaves them: zero-gradient ghosts copy the cell, the
                // low ghost of a periodic axis the last cell of the line
                if (i == 0) {
                    if (!loX) A.sg[id - 1] = last;
                    if (hiX) A.wrapX[((size_t)npar * ntz + K) * g.ny + j0 + jj] = last;
                }
                if (i == g.nx - 1) {
                    if (loX) A.sg[id - (size_t)g.nx] = last;
                    if (!hiX) A.sg[id + 1] = last;
                }
                if (DIM >= 2) {
                    const int j = j0 + jj;
                    if (j == 0) {
                        if (!loY) A.sg[id - sy] = last;
                        if (hiY) A.wrapY[((size_t)npar * ntz + K) * g.nx + i] = last;
                    }
                    if (j == g.ny - 1) {
                        if (loY) A.sg[id - (size_t)g.ny * sy] = last;
                        if (!hiY) A.sg[id + sy] = last;
                    }
                }
                if (DIM >= 3) {
                    if (K == 0) {
                        if (!loZ) A.sg[id - sz] = last;
                        if (hiZ) A.wrapZ[((size_t)npar * g.ny + j0 + jj) * g.nx + i] = last;
                    }
                    if (K == g.nz - 1) {
                        if (loZ) A.sg[id - (size_t)g.nz * sz] = last;
                        if (!hiZ) A.sg[id + sz] = last;
                    }
                }
            }
        }
        // publish the tile: data before flag (every lane stores the same value)
        __threadfence();
        __syncthreads();
        __hip_atomic_store(&A.flags[t], now, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Strip-streaming sweep (2D).  A wave owns a strip of 64 columns and walks
// every row of every sweep of the solve: at schedule step s lane l updates
// row r = s - l, where r runs over the sweeps with period P = ny + gap (rows
// ny..P-1 of a period are idle).  The left neighbour's new value is lane
// l-1's result of the previous step (DPP shift), the lower one the lane's
// own previous result, the +x neighbour's old value is what lane l+1 holds
// as its +y operand (DPP shift), so no tile waits for another tile to finish.
// Per cell the pack kernel stores the five coefficients of the update
//     new = c0 + cR*right + cT*top + cB*bottom + cL*left,
// c0 = rhs/diag, cX = alpha/h^2 / rho_face / diag, and sigma, in a cyclic
// skew: cell (l, j) of a strip sits in slot (j + l) mod ny, so at any step
// all lanes read one contiguous slot of 64 x 6 doubles (each lane's six
// values adjacent: three 16-byte loads per lane and step).  A wave may have
// 63 vector memory operations outstanding, stores included: two groups of
// kStripGroup steps are loaded ahead (48 loads) and the edge values a chunk
// produces are gathered in registers and written once per chunk.  Lane 0's left
// values and the last lane's right values cross strips through edge arrays
// (by sweep parity), read per chunk of kStripChunk steps after the strip
// has polled its neighbours' progress (rows completed by their last lane),
// which every strip publishes at the end of each chunk.  On a wrapping x
// axis sweep q of strip 0 needs sweep q-1 of the last strip on the same
// row, while a strip's last lane trails its lane 0 by 63 rows and each strip
// trails its left neighbour by up to 64 + kStripChunk steps: the gap between
// sweeps absorbs that skew (gap = W (64 + C) + 64 + C - ny when positive).
// Ghost semantics are those of the per-sweep scalar fill: zero-gradient
// faces read the cell's own previous-sweep value, wrapping y faces the
// previous result (bottom) and the saved previous-sweep row-0 value (top).
constexpr int kStripGroup = 8;     // steps per prefetch group
constexpr int kStripBufs = 3;      // groups held in registers: two in flight while one is used
constexpr int kStripChunk = kStripGroup * kStripBufs;   // steps per progress hand-off
constexpr int kStripSlot = 6 * kWave;   // doubles per record slot: [lane][c0, cL, cR, cB, cT, sigma]

struct IgrStripPackArgs {
    Grid g;
    const double* sigma;   // padded
    const double* rho;     // padded, ghosts filled
    const double* rhs;
    const double* diag;
    double alpha;
    double* rec;           // [W][ny][6][64]
    double* edgeL;         // [2][W][ny]
    double* edgeR;         // [2][W][ny]
    int W;
};

// Threads run in record order (strip, slot, lane): each wave writes one
// contiguous 3 KB slot, and reads the slot's diagonal of cells (row = slot -
// lane), whose lines the neighbouring slots share in L2.
template <int DIM>
__global__ void __launch_bounds__(256) igr_strip_pack_kernel(IgrStripPackArgs A) {
    const Grid& g = A.g;
    const size_t sy = (size_t)g.nxT;
    const size_t total = (size_t)A.W * g.ny * kWave;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int l = (int)(e % kWave);
        const size_t ws = e / kWave;
        const int t = (int)(ws % g.ny), w = (int)(ws / g.ny);
        const int i = w * kWave + l;
        if (i >= g.nx) continue;
        int j = t - l;
        if (j < 0) j += g.ny;
        const size_t id = cell_index(g, i, j, 0);
        const double ri = A.rho[id];
        const double rd = 1.0 / A.diag[id];
        const double hx = g.dxw[i + g.ngx], hy = g.dyw[j + g.ngy];
        const double ax = A.alpha / (hx * hx), ay = A.alpha / (hy * hy);
        double2* r = reinterpret_cast<double2*>(A.rec + ws * kStripSlot + 6 * l);
        const double sg = A.sigma[id];
        r[0] = make_double2(A.rhs[id] * rd, ax / (0.5 * (ri + A.rho[id - 1])) * rd);
        r[1] = make_double2(ax / (0.5 * (ri + A.rho[id + 1])) * rd, ay / (0.5 * (ri + A.rho[id - sy])) * rd);
        r[2] = make_double2(ay / (0.5 * (ri + A.rho[id + sy])) * rd, sg);
        // edge values of "sweep -1" (parity slot 1): the initial iterate
        const int L = min(kWave - 1, g.nx - 1 - w * kWave);
        if (l == 0) A.edgeL[((size_t)A.W + w) * g.ny + j] = sg;
        if (l == L) A.edgeR[((size_t)A.W + w) * g.ny + j] = sg;
    }
}

template <int DIM>
__global__ void __launch_bounds__(256) igr_strip_unpack_kernel(Grid g, const double* rec, double* sigma) {
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, 1); it_.next(i, j, k);) {
        const int w = i / kWave, l = i - w * kWave;
        int t = j + l;
        if (t >= g.ny) t -= g.ny;
        sigma[cell_index(g, i, j, 0)] = rec[((size_t)w * g.ny + t) * kStripSlot + 6 * l + 5];
    }
}

struct IgrStripArgs {
    int nx, ny;
    double* rec;
    double* edgeL;
    double* edgeR;
    int* prog;             // [W] schedule rows completed by the strip's last lane; zero at launch
    int* err;
    const int* active;
    int W, P, nsweeps;
    int perX, perY;
};

__device__ __forceinline__ double dpp_shr1(double v) {   // lane l <- lane l-1
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double dpp_shl1(double v) {   // lane l <- lane l+1
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double read_lane(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Largest schedule row in [a, b] that is a real row (not a gap row, inside
// the launch's [0, Rt)), or -1.
__device__ __forceinline__ int strip_last_row(int a, int b, int P, int ny, int Rt) {
    b = min(b, Rt - 1);
    if (b < a || b < 0) return -1;
    const int jb = b % P;
    const int r = jb < ny ? b : b - jb + ny - 1;
    return r >= a && r >= 0 ? r : -1;
}

// Two waves per strip.  Wave 0 computes; wave 1 carries the strip's traffic
// with its neighbours through LDS: it publishes each chunk wave 0 finishes
// (edge values to the edge arrays, release, progress flag) and loads the
// neighbours' edge values of upcoming chunks as soon as their progress
// allows (acquire), up to kStripRing chunks ahead.  Wave 0 thus never waits
// on a device-scope fence, which would drain its prefetch pipeline.  Wave 1
// never blocks on one duty: each pass of its loop publishes what it can and
// loads what it can.  Neither wave branches on its lane index at a loop
// boundary (stores of lanes without a value go to a sink element).
constexpr int kStripRing = 8;

template <int DIM>
__global__ void __launch_bounds__(128) igr_strip_kernel(IgrStripArgs A) {
    if (A.active && *A.active == 0) return;
    constexpr int U = kStripGroup, C = kStripChunk, NR = kStripRing;
    __shared__ double sIn[NR][2][kWave];    // neighbours' values for wave 0: [slot][left/right][chunk step]
    __shared__ double sOut[NR][C][kWave];   // wave 0's results: [slot][chunk step][lane]
    __shared__ int sFlag[4];                // chunks done by wave 0, loaded for it, published; abort
    const int w = blockIdx.x, lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);   // wave-uniform
    const int ny = A.ny, P = A.P, W = A.W;
    const int L = min(kWave - 1, A.nx - 1 - w * kWave);
    const int Rt = A.nsweeps * P;
    const int nsteps = Rt + kWave - 1;
    const int nchunks = (nsteps + C - 1) / C;
    const bool leftWall = w == 0 && !A.perX, rightWall = w == W - 1 && !A.perX;
    const int leftSrc = w > 0 ? w - 1 : W - 1, rightSrc = w < W - 1 ? w + 1 : 0;
    if (threadIdx.x < 4) sFlag[threadIdx.x] = 0;
    __syncthreads();
    auto lds_load = [](int* p) __attribute__((always_inline)) { return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP); };
    auto lds_store = [](int* p, int v) __attribute__((always_inline)) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP); };
    // wave 1 watches the device error flag and raises the block's abort flag,
    // which wave 0 polls in LDS (a device-scope load would make wave 0 wait
    // for its whole prefetch pipeline)
    auto aborted = [&]() __attribute__((always_inline)) { return __hip_atomic_load(&sFlag[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0; };
    if (wave == 1) {
        double* sink = A.edgeL + 4 * (size_t)W * ny;
        int nextPub = 0, nextIn = 0, knownL = 0, knownR = 0, idle = 0;
        while (nextPub < nchunks) {
            bool moved = false;
            // publish every chunk wave 0 has finished: edge values, then one
            // release and the progress of the last of them
            const int aDone = lds_load(&sFlag[0]);
            if (aDone > nextPub) {
                for (int c = nextPub; c < aDone; ++c) {
                    const int s0 = c * C, slot = c % NR;
                    int r = s0 + lane;
                    bool v = lane < C && r < Rt && r % P < ny;
                    double* dl = v ? A.edgeL + ((size_t)((r / P) & 1) * W + w) * ny + r % P : sink;
                    r = s0 + lane - L;
                    v = lane < C && r >= 0 && r < Rt && r % P < ny;
                    double* dr = v ? A.edgeR + ((size_t)((r / P) & 1) * W + w) * ny + r % P : sink;
                    *dl = sOut[slot][min(lane, C - 1)][0];
                    *dr = sOut[slot][min(lane, C - 1)][L];
                }
                const int done = min(max(min(aDone * C, nsteps) - L, 0), Rt);
                __threadfence();
                __hip_atomic_store(&A.prog[w], done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                lds_store(&sFlag[2], aDone);
                nextPub = aDone;
                moved = true;
            }
            // load the inputs of every chunk whose slot is free and whose rows
            // the neighbours have published: one acquire for all of them
            const int upto = min(nchunks, aDone + NR);
            if (nextIn < upto) {
                int nlL = knownL, nlR = knownR;
                auto need = [&](int c, int& nL, int& nR) __attribute__((always_inline)) {
                    const int s0 = c * C;
                    nL = 0;
                    nR = 0;
                    if (!leftWall) {
                        const int r = strip_last_row(s0, s0 + C - 1, P, ny, Rt);
                        if (r >= 0) nL = w > 0 ? r + 1 : r - P + 1;
                    }
                    if (!rightWall) {
                        const int r = strip_last_row(s0 - L, s0 + C - 1 - L, P, ny, Rt);
                        if (r >= 0) nR = r - P + 1;
                    }
                };
                int nL, nR;
                need(nextIn, nL, nR);
                if (nL > nlL) nlL = __hip_atomic_load(&A.prog[leftSrc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (nR > nlR) nlR = __hip_atomic_load(&A.prog[rightSrc], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                knownL = nlL;
                knownR = nlR;
                int last = nextIn;
                while (last < upto) {
                    need(last, nL, nR);
                    if (nL > knownL || nR > knownR) break;
                    ++last;
                }
                if (last > nextIn) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    for (int c = nextIn; c < last; ++c) {
                        const int s0 = c * C, slot = c % NR;
                        int r = s0 + lane;
                        bool v = lane < C && !leftWall && r < Rt && r % P < ny;
                        int q = v ? r / P : 0, j = v ? r % P : 0;
                        const double a = A.edgeR[v ? ((size_t)((w > 0 ? q : q - 1) & 1) * W + leftSrc) * ny + j : 0];
                        r = s0 + lane - L;
                        const bool vr = lane < C && !rightWall && r >= 0 && r < Rt && r % P < ny;
                        q = vr ? r / P : 0;
                        j = vr ? r % P : 0;
                        const double b = A.edgeL[vr ? ((size_t)((q - 1) & 1) * W + rightSrc) * ny + j : 0];
                        sIn[slot][0][lane] = v ? a : 0.0;
                        sIn[slot][1][lane] = vr ? b : 0.0;
                    }
                    lds_store(&sFlag[1], last);
                    nextIn = last;
                    moved = true;
                }
            }
            if (moved) {
                idle = 0;
            } else {
                const bool budget = ++idle > (1 << 25);
                if (budget) __hip_atomic_store(A.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (budget || __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || aborted()) {
                    __hip_atomic_store(&sFlag[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        return;
    }
    // wave 0: the sweeps
    const bool act = lane <= L;
    double* rec = A.rec + (size_t)w * ny * kStripSlot + 6 * lane;
    // lane state: row 0 of the current and of the previous sweep, last result
    double row0 = rec[(size_t)lane * kStripSlot + 5];           // slot of (lane, 0) = lane
    int tl = ny - 1 + lane;
    if (tl >= ny) tl -= ny;
    double last = A.perY ? rec[(size_t)tl * kStripSlot + 5] : row0;
    double row0old = row0;
    // schedule position of the lane at the next compute step and at the
    // next prefetch step: row in the period and sweep
    int jc = lane == 0 ? 0 : P - lane;
    int jp = jc;
    double buf[kStripBufs][6][U];
    auto prefetch = [&](double (&bf)[6][U]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < U; ++k) {
            int t = jp < ny ? jp + lane : lane;
            if (t >= ny) t -= ny;
            const double2* r = reinterpret_cast<const double2*>(rec + (size_t)t * kStripSlot);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const double2 v = r[f];
                bf[2 * f][k] = v.x;
                bf[2 * f + 1][k] = v.y;
            }
            if (++jp == P) jp = 0;
        }
    };
    // lane k holds lane 0's left value and the last lane's right value of
    // chunk step k (read once per chunk from LDS, picked per step by readlane:
    // an LDS read per step would wait behind the previous step's LDS write)
    double lchunk = 0.0, rchunk = 0.0;
    // a group in which every lane's row is an interior row (1 .. ny-2) of one
    // sweep: all lanes update, none meets a boundary row, and all lanes write
    // the same record slot (s - qP mod ny); lanes past the mesh edge write
    // their padding records
    auto group = [&](int g0, int s0, int slot, const double (&cur)[6][U], const double (&nxt)[6][U]) __attribute__((always_inline)) {
        const int lo = g0 - (kWave - 1), hi = g0 + U - 1;
        const int q = lo >= 0 ? lo / P : -1;
        if (lo >= 0 && hi < Rt && lo - q * P >= 1 && hi - q * P <= ny - 2) {
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int kk = g0 + k - s0;
                const double sgT = k < U - 1 ? cur[5][k + 1] : nxt[5][0];
                const double cellOld = cur[5][k];
                const double rightS = dpp_shl1(sgT);
                const double leftS = dpp_shr1(last);
                const double lv = read_lane(lchunk, kk), rv = read_lane(rchunk, kk);
                const double right = lane < L ? rightS : (rightWall ? cellOld : rv);
                const double left = lane > 0 ? leftS : (leftWall ? cellOld : lv);
                const double base = cur[0][k] + cur[2][k] * right + cur[4][k] * sgT;
                const double nw = base + cur[3][k] * last + cur[1][k] * left;
                int t = g0 + k - q * P;
                if (t >= ny) t -= ny;
                rec[(size_t)t * kStripSlot + 5] = nw;
                sOut[slot][kk][lane] = nw;
                last = nw;
            }
            jc += U;
        } else {
            // general steps: any row position (first or last row of a sweep,
            // gap rows, rows outside the launch)
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const int s = g0 + k, kk = s - s0;
                if (s >= nsteps) break;
                const double sgT = k < U - 1 ? cur[5][k + 1] : nxt[5][0];
                const int r = s - lane;
                const bool valid = act && r >= 0 && r < Rt && jc < ny;
                const double sendR = jc + 1 == P ? row0 : sgT;
                const double rightS = dpp_shl1(sendR);
                const double leftS = dpp_shr1(last);
                const double cellOld = cur[5][k];
                const double top = jc < ny - 1 ? sgT : (A.perY ? row0old : cellOld);
                const double bot = jc > 0 || A.perY ? last : row0;
                const double lv = read_lane(lchunk, kk), rv = read_lane(rchunk, kk);
                const double right = lane < L ? rightS : (rightWall ? cellOld : rv);
                const double left = lane > 0 ? leftS : (leftWall ? cellOld : lv);
                const double base = cur[0][k] + cur[2][k] * right + cur[4][k] * top;
                const double nw = base + cur[3][k] * bot + cur[1][k] * left;
                if (valid) {
                    if (jc == 0) {
                        row0old = row0;
                        row0 = nw;
                    }
                    int t = jc + lane;
                    if (t >= ny) t -= ny;
                    rec[(size_t)t * kStripSlot + 5] = nw;
                    last = nw;
                }
                sOut[slot][kk][lane] = nw;
                if (++jc == P) jc = 0;
            }
        }
    };
    auto wait_flag = [&](int* f, int v) __attribute__((always_inline)) {
        int spins = 0;
        while (lds_load(f) < v) {
            if (aborted()) return false;
            if (++spins > (1 << 27)) {
                __hip_atomic_store(A.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&sFlag[3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return true;
    };
    prefetch(buf[0]);
    prefetch(buf[1]);
    for (int c = 0; c < nchunks; ++c) {
        const int s0 = c * C, slot = c % NR;
        // inputs of the chunk loaded, and the output slot published by wave 1
        if (!wait_flag(&sFlag[1], c + 1)) return;
        if (!wait_flag(&sFlag[2], c + 1 - NR)) return;
        lchunk = sIn[slot][0][lane];
        rchunk = sIn[slot][1][lane];
        prefetch(buf[2]);
        group(s0, s0, slot, buf[0], buf[1]);
        prefetch(buf[0]);
        group(s0 + U, s0, slot, buf[1], buf[2]);
        prefetch(buf[1]);
        group(s0 + 2 * U, s0, slot, buf[2], buf[0]);
        lds_store(&sFlag[0], c + 1);
    }
}

}  // namespace sifv
