// sifv_ensemble.hpp - many independent 1D explicit runs (shock tubes of
// uneven sizes) advanced to their end time in one persistent launch.
//
// Each tube is an ExplicitSolver run (src/ExplicitSolver.cpp:68-203) on a
// 1D mesh with outflow ends, driven by the runTimeLoop dt rule without
// output (targetDt = endTime - time, src/RKTimeStepping.cpp:241-356).
// A workgroup takes the next tube from a global queue (tubes ordered by
// estimated cost, costliest first), keeps the tube's time loop on chip
// (block-wide dt reduction, three conservative buffers in global memory that
// stay L2-resident), and sweeps the tube in tiles of BLOCK cells per RK
// stage:
//   1. BLOCK+6 cells (3-cell halo, outflow = clamped index) -> FP64 cons->prim
//      -> LDS;
//   2. WENO5 of cells -1..BLOCK (both face states of a cell from one stencil);
//   3. Riemann flux of faces -1/2..BLOCK-1/2;
//   4. flux divergence + SSP-RK update of the owned cell (+ the acoustic dt
//      of the new state in the last stage).
// BLOCK is the tile shape the autotuner chooses.  A tube whose estimated cost
// exceeds a workgroup's fair share is cut into P contiguous parts (work items)
// taken by P workgroups that meet at a per-tube barrier after every RK stage
// and share the dt reduction through per-part slots.  Items are queued with
// the parts of a tube adjacent and the launch never exceeds the number of
// co-resident workgroups, so every started group completes; barrier waits are
// bounded (an error flag ends all waits), so every wave ends.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

template <typename T>
struct EnsArgs {
    Eos eos;
    int ntubes;
    const int* order;        // work items (tube * 32 + part), costliest first
    int nitems;
    const int* nparts;       // parts per tube
    int* bar;                // per tube: barrier arrival counter
    double* dtSlot;          // per tube x 32: per-part dt partials
    int* err;                // set when a barrier wait exceeds its budget
    const int* ncell;        // cells per tube
    const long long* off;    // first cell of each tube in the packed arrays
    const double* dxw;       // packed cell widths
    T* U[3];                 // conservative buffers; field q of buffer b at U[b] + q*ntot
    long long ntot;
    const double* velIC;     // initial velocity / pressure (dt of the first step)
    const double* presIC;
    double cfl, maxDt, constDt, tEnd, eps;
    int maxSteps, rk;
    int* queue;              // next tube position
    int* steps;              // per tube: steps taken
    double* time;            // per tube: final time
    int* status;             // set on a non-finite dt
    int cur;                 // buffer holding U^n on entry
};

// Barrier of the P workgroups of one tube (monotone counter; target = P x
// number of barriers passed so far).  Data written before it by any member
// is visible after it to every member.
__device__ __forceinline__ void tube_barrier(int* cnt, int target, int* err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(cnt, 1);
        int spins = 0;
        while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++spins > (1 << 25)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ double block_min_d(double v, double* sh) {
    v = wave_min(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double m = 1.0e300;
    const int nw = blockDim.x / kWave;
    for (int q = 0; q < nw; ++q) m = dmin_ref(m, sh[q]);
    return m;
}

template <int BLOCK, int RS, typename T>
__global__ void __launch_bounds__(BLOCK) ensemble_kernel(EnsArgs<T> A) {
    constexpr int NV = 3;           // rho, u, p
    constexpr int W = BLOCK + 6;    // loaded cells per tile
    __shared__ T X[NV][W];          // primitives, then face fluxes
    __shared__ T Mn[NV][BLOCK + 2]; // minus state of cells -1..BLOCK
    __shared__ T Pl[NV][BLOCK + 2]; // plus state
    __shared__ double red[BLOCK / kWave];
    __shared__ int tubeSlot;
    const int tid = threadIdx.x;
    const T eps = (T)A.eps;
    const double g = A.eos.gamma;
    const double pinf = A.eos.stiffened ? A.eos.pInf : 0.0;
    const T gT = (T)g, piT = (T)pinf, gm1iT = (T)(1.0 / (g - 1.0));
    for (;;) {
        __syncthreads();
        if (tid == 0) tubeSlot = atomicAdd(A.queue, 1);
        __syncthreads();
        const int slot = tubeSlot;
        if (slot >= A.nitems) return;
        const int item = A.order[slot];
        const int tube = item >> 5, part = item & 31;
        const int N = A.ncell[tube];
        const int P = A.nparts[tube];
        const long long o = A.off[tube];
        const double* dx = A.dxw + o;
        // cells [c0, c1) of this part, in whole tiles
        const int tilesAll = (N + BLOCK - 1) / BLOCK;
        const int tpp = (tilesAll + P - 1) / P;
        const int c0 = part * tpp * BLOCK;
        const int c1 = (c0 + tpp * BLOCK < N) ? c0 + tpp * BLOCK : N;
        const int ntile = (c1 - c0 + BLOCK - 1) / BLOCK;
        int phase = 0;
        int* bar = A.bar + tube;
        double* dts = A.dtSlot + (size_t)tube * 32;
        // dt shared by the parts: own partial -> slot, barrier, min over slots
        auto group_min = [&](double v) {
            if (P == 1) return v;
            if (tid == 0) dts[part] = v;
            ++phase;
            tube_barrier(bar, P * phase, A.err);
            double m = 1.0e300;
            for (int q = 0; q < P; ++q) m = dmin_ref(m, __hip_atomic_load(&dts[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            return m;
        };
        // acoustic dt of the initial state (computeAcousticTimeStep,
        // src/RKTimeStepping.cpp:43-76, on the state's primitive arrays)
        double dtc = 1.0e300;
        for (int c = c0 + tid; c < c1; c += BLOCK) {
            const T* Ub = A.U[A.cur];
            const double rho = (double)Ub[o + c];
            const double p = A.presIC[o + c], u = A.velIC[o + c];
            const double cs = sqrt(g * dmax_ref(p + pinf, 1e-14) / dmax_ref(rho, 1e-14));
            dtc = dmin_ref(dtc, dx[c] / (fabs(u) + cs));
        }
        dtc = group_min(block_min_d(dtc, red));
        double t = 0.0;
        int step = 0;
        int a = A.cur, b = (A.cur + 1) % 3, cbuf = (A.cur + 2) % 3;
        bool bad = false;
        while (t < A.tEnd && step < A.maxSteps) {
            const double target = A.tEnd - t;
            double dt = A.constDt > 0.0 ? A.constDt : dmin_ref(A.maxDt, A.cfl * dtc);
            if (target > 0.0) dt = dmin_ref(dt, target);
            if (!(dt > 0.0) || !isfinite(dt)) {
                bad = true;
                break;
            }
            // SSP-RK2/3 (src/ExplicitSolver.cpp:100-103); both end in buffer a
            int src[3], u0[3], dst[3];
            if (A.rk == 2) { src[0] = a; u0[0] = -1; dst[0] = b; src[1] = b; u0[1] = a; dst[1] = a; }
            else { src[0] = a; u0[0] = -1; dst[0] = b; src[1] = b; u0[1] = a; dst[1] = cbuf; src[2] = cbuf; u0[2] = a; dst[2] = a; }
            double dtNew = 1.0e300;
            for (int s = 0; s < A.rk; ++s) {
                double co[4];
                if (s == 0) { co[0] = 1; co[1] = 0; co[2] = 1; co[3] = 1; }
                else if (A.rk == 2) { co[0] = 1; co[1] = 1; co[2] = 1; co[3] = 2; }
                else if (s == 1) { co[0] = 1; co[1] = 3; co[2] = 1; co[3] = 4; }
                else { co[0] = 2; co[1] = 1; co[2] = 2; co[3] = 3; }
                const bool last = s == A.rk - 1;
                const T* S = A.U[src[s]] + o;
                const T* S0 = u0[s] >= 0 ? A.U[u0[s]] + o : nullptr;
                T* D = A.U[dst[s]] + o;
                const long long nt = A.ntot;
                for (int k = 0; k < ntile; ++k) {
                    const int base = c0 + k * BLOCK;
                    // 1. primitives of cells base-3 .. base+BLOCK+2 (outflow ends)
                    for (int m = tid; m < W; m += BLOCK) {
                        int c = base - 3 + m;
                        c = c < 0 ? 0 : (c >= N ? N - 1 : c);
                        double mom[1] = {(double)S[nt + c]}, u[1], p;
                        const double rho = (double)S[c];
                        cons_to_prim<1>(A.eos, rho, mom, (double)S[2 * nt + c], u, p);
                        X[0][m] = (T)rho;
                        X[1][m] = (T)u[0];
                        X[2][m] = (T)p;
                    }
                    __syncthreads();
                    // 2. face states of cells base-1 .. base+BLOCK
                    for (int jj = tid; jj < BLOCK + 2; jj += BLOCK) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            T st[5];
#pragma unroll
                            for (int q = 0; q < 5; ++q) st[q] = X[v][jj + q];
                            T mn, pl;
                            Recon<3, T>::run(st, eps, true, mn, pl);
                            Mn[v][jj] = mn;
                            Pl[v][jj] = pl;
                        }
                    }
                    __syncthreads();
                    // 3. fluxes of faces base-1/2 .. base+BLOCK-1/2 (face f: cells f-1 | f)
                    for (int f = tid; f < BLOCK + 1; f += BLOCK) {
                        FaceState<1, 0, T> L, R;
                        L.rho = Pl[0][f]; L.u[0] = Pl[1][f]; L.p = Pl[2][f];
                        R.rho = Mn[0][f + 1]; R.u[0] = Mn[1][f + 1]; R.p = Mn[2][f + 1];
                        L.sig = R.sig = T(0);
                        L.g = R.g = gT;
                        L.pi = R.pi = piT;
                        L.gm1i = R.gm1i = gm1iT;
                        T F[1 + 4];
                        riemann<RS, 0, 1, 0, T>(L, R, true, false, F);
                        X[0][f] = F[0];
                        X[1][f] = F[1];
                        X[2][f] = F[2];
                    }
                    __syncthreads();
                    // 4. update of the owned cell
                    const int c = base + tid;
                    if (c < c1) {
                        const double cx = (1.0 * 1.0) / ((dx[c] * 1.0) * 1.0);
                        const T cT = (T)cx;
                        double un[3];
#pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            const T lo = X[q][tid], hi = X[q][tid + 1];
                            const T accq = cT * lo - cT * hi;
                            const double U = (double)S[q * nt + c];
                            const double Rq = (double)accq;
                            double av = co[0] * U;
                            if (co[1] != 0.0) av = av + co[1] * (double)S0[q * nt + c];
                            const double bv = av + (co[2] * dt) * Rq;
                            un[q] = co[3] == 3.0 ? bv / 3.0 : bv * (1.0 / co[3]);
                            D[q * nt + c] = (T)un[q];
                        }
                        if (last && A.constDt <= 0.0) {
                            double mom[1] = {(double)(T)un[1]}, u[1], p;
                            const double rho = (double)(T)un[0];
                            cons_to_prim<1>(A.eos, rho, mom, (double)(T)un[2], u, p);
                            const double cs = sqrt(g * dmax_ref(p + pinf, 1e-14) / dmax_ref(rho, 1e-14));
                            dtNew = dmin_ref(dtNew, dx[c] / (fabs(u[0]) + cs));
                        }
                    }
                    __syncthreads();  // X is rewritten by the next tile
                }
                // the next stage reads the whole tube: wait for the other parts
                if (P > 1 && !last) {
                    ++phase;
                    tube_barrier(bar, P * phase, A.err);
                }
            }
            dtc = group_min(block_min_d(dtNew, red));
            t += dt;
            ++step;
        }
        if (tid == 0 && part == 0) {
            A.steps[tube] = step;
            A.time[tube] = t;
            if (bad) atomicOr(A.status, 1);
        }
    }
}

}  // namespace sifv
