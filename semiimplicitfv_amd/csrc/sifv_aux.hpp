// sifv_aux.hpp - time-step reductions, cons->prim and ghost-fill kernels.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

constexpr int kRedThreads = 256;

// Block-wide minimum; every thread gets the result.
__device__ __forceinline__ double block_min(double v, double* sh) {
    v = wave_min(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = v;
    __syncthreads();
    const int nw = blockDim.x / kWave;
    double m = 1.0e300;
    for (int q = 0; q < nw; ++q) m = dmin_ref(m, sh[q]);
    __syncthreads();
    return m;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = v;
    __syncthreads();
    const int nw = blockDim.x / kWave;
    double m = 0.0;
    for (int q = 0; q < nw; ++q) m = (m < sh[q]) ? sh[q] : m;
    __syncthreads();
    return m;
}

// Per-block minimum of the acoustic (advective when `advective`) cell time
// step over interior cells from the primitive arrays:
// computeAcousticTimeStep (src/RKTimeStepping.cpp:43-123) /
// computeAdvectiveTimeStep (src/RKTimeStepping.cpp:15-41).
struct DtArgs {
    Grid g;
    Eos eos;
    Mix mix;
    const double* rho;
    const double* vel[3];
    const double* pres;
    const double* alpha[kMaxPhases];
    const double* alphaRho[kMaxPhases];
    const float* alpha32[kMaxPhases];     // FP32 state: volume fractions and partial
    const float* alphaRho32[kMaxPhases];  // densities are read from these instead
    int f32 = 0;
    int advective;
    double viscMu = 0.0;    // > 0: include the viscous limit
    double capSigma = 0.0;  // > 0: include the capillary limit
    double* partial;
};

// DIM fixed at compile time: a cell's velocity loads are issued together
// (the cell limits and their minimum are the same operations either way).
template <int DIM>
__global__ void __launch_bounds__(kRedThreads) dt_cells_kernel(DtArgs A) {
    __shared__ double sh[kRedThreads / kWave];
    const Grid& g = A.g;
    double m = 1.0e300;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        size_t id = cell_index(g, i, j, k);
        const double h[3] = {g.dxw[i + g.ngx], g.dyw[j + g.ngy], g.dzw[k + g.ngz]};
        double v[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = A.vel[d][id];
        double dc;
        if (A.advective) {
            dc = h[0] / dmax_ref(fabs(v[0]), 1e-14);
#pragma unroll
            for (int d = 1; d < DIM; ++d) dc = dmin_ref(dc, h[d] / dmax_ref(fabs(v[d]), 1e-14));
        } else {
            double c;
            if (A.mix.nph > 0) {
                double a[kMaxPhases], ar[kMaxPhases];
                for (int ph = 0; ph < A.mix.nph; ++ph) {
                    a[ph] = A.f32 ? (double)A.alpha32[ph][id] : A.alpha[ph][id];
                    ar[ph] = A.f32 ? (double)A.alphaRho32[ph][id] : A.alphaRho[ph][id];
                }
                c = mix_sound(A.mix, A.rho[id], A.pres[id], a, ar, A.mix.nph);
            } else {
                double pinf = A.eos.stiffened ? A.eos.pInf : 0.0;
                c = sqrt(A.eos.gamma * dmax_ref(A.pres[id] + pinf, 1e-14) / dmax_ref(A.rho[id], 1e-14));
            }
            dc = h[0] / (fabs(v[0]) + c);
#pragma unroll
            for (int d = 1; d < DIM; ++d) dc = dmin_ref(dc, h[d] / (fabs(v[d]) + c));
        }
        if (A.viscMu > 0.0) dc = dmin_ref(dc, visc_cell_dt(h, g.dim, A.viscMu, A.rho[id]));
        if (A.capSigma > 0.0) dc = dmin_ref(dc, cap_cell_dt(h, g.dim, A.capSigma, A.rho[id]));
        m = dmin_ref(m, dc);
    }
    m = block_min(m, sh);
    if (threadIdx.x == 0) A.partial[blockIdx.x] = m;
}

// Per-block minimum of the viscous and capillary cell limits over interior
// cells from the conservative density of a new state (computeViscousDt /
// computeCapillaryDt read state.rho, which cons->prim leaves as the
// conservative density).  Launched after the last RK stage into the partial
// slots after the stage kernel's fused acoustic minima, so that the next
// dt_finalize takes the minimum over all three limits.
template <typename T>
struct StressDtArgs {
    Grid g;
    const T* rho;
    const int* active;    // device flag; 0 skips (time loop finished)
    double viscMu;        // > 0: viscous limit
    double capSigma;      // > 0: capillary limit
    double* partial;
};

template <typename T>
__global__ void __launch_bounds__(kRedThreads) stress_dt_kernel(StressDtArgs<T> A) {
    if (A.active && *A.active == 0) return;
    __shared__ double sh[kRedThreads / kWave];
    const Grid& g = A.g;
    double m = 1.0e300;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double h[3] = {g.dxw[i + g.ngx], g.dyw[j + g.ngy], g.dzw[k + g.ngz]};
        const double rho = (double)A.rho[id];
        double dc = 1.0e300;
        if (A.viscMu > 0.0) dc = dmin_ref(dc, visc_cell_dt(h, g.dim, A.viscMu, rho));
        if (A.capSigma > 0.0) dc = dmin_ref(dc, cap_cell_dt(h, g.dim, A.capSigma, rho));
        m = dmin_ref(m, dc);
    }
    m = block_min(m, sh);
    if (threadIdx.x == 0) A.partial[blockIdx.x] = m;
}

// Final time step of one step, on device:
//   dt = constDt > 0 ? constDt : min(maxDt, cfl * min(partials))
//   dt = min(dt, target) when target > 0       (ExplicitSolver::step,
//                                              src/ExplicitSolver.cpp:72-90)
//   semi-implicit: clamp to [minDt, maxDt]     (src/SemiImplicitSolver.cpp:110-123)
// In time-loop mode the target is endTime - time and the loop's `active`
// flag is cleared once time >= endTime (runTimeLoop, src/RKTimeStepping.cpp:259-274).
struct DtFinalize {
    const double* partial;
    int npartial;
    double cfl, maxDt, minDt, constDt, target;
    int clampMin;      // semi-implicit clamp
    int loop;          // time-loop mode
    double endTime;
    double time0;      // config time of the step (not time-loop mode)
    int bodyForce;     // hasBodyForce: accel = bfA + bfB cos(bfC t + bfD) per axis
    double bfA[3], bfB[3], bfC[3], bfD[3];
    double* scal;      // [0] dt, [1] time, [2] time at the start of the step, [4..6] accel
    int* flags;        // [0] active, [1] steps taken
};

__global__ void __launch_bounds__(kRedThreads) dt_finalize_kernel(DtFinalize F) {
    __shared__ double sh[kRedThreads / kWave];
    if (F.loop && F.flags[0] == 0) return;
    double m = 1.0e300;
    if (F.constDt <= 0.0)
        for (int t = threadIdx.x; t < F.npartial; t += blockDim.x) m = dmin_ref(m, F.partial[t]);
    m = block_min(m, sh);
    if (threadIdx.x != 0) return;
    double target = F.target;
    // the config time the step's stages see (runTimeLoop sets config.time
    // before calling the step, src/RKTimeStepping.cpp:265)
    F.scal[2] = F.loop ? F.scal[1] : F.time0;
    // body-force acceleration at that time (src/ExplicitSolver.cpp:500-503)
    if (F.bodyForce)
        for (int d = 0; d < 3; ++d) F.scal[4 + d] = F.bfA[d] + F.bfB[d] * cos(F.bfC[d] * F.scal[2] + F.bfD[d]);
    if (F.loop) {
        double time = F.scal[1];
        if (!(time < F.endTime)) {
            F.flags[0] = 0;
            return;
        }
        target = F.endTime - time;
    }
    double dt = F.constDt > 0.0 ? F.constDt : dmin_ref(F.maxDt, F.cfl * m);
    if (target > 0.0) dt = dmin_ref(dt, target);
    if (F.clampMin) {
        if (dt < F.minDt) dt = F.minDt;
        else if (F.maxDt < dt) dt = F.maxDt;
    }
    F.scal[0] = dt;
    if (F.loop) {
        F.scal[1] = F.scal[1] + dt;
        F.flags[1] = F.flags[1] + 1;
    }
}

__global__ void loop_reset_kernel(double* scal, int* flags, double time) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        scal[1] = time;
        flags[0] = 1;
        flags[1] = 0;
    }
}

// cons -> prim on interior cells, writing the primitive arrays
// (convertConservativeToPrimitiveVariables / MixtureEOS::convertConservativeToPrimitive).
template <typename T>
struct C2PArgs {
    Grid g;
    Eos eos;
    Mix mix;
    const T* cons[5 + 2 * kMaxPhases];  // rho, mom[dim], E, alphaRho[nph], alpha[nph]
    double* rho;                         // SolutionState::rho seen by the primitive side
    double* vel[3];
    double* pres;
};

template <typename T>
__global__ void __launch_bounds__(256) cons_to_prim_kernel(C2PArgs<T> A) {
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const int dim = g.dim;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        size_t id = cell_index(g, i, j, k);
        if (A.mix.nph > 0) {
            double r = 0.0;
            for (int ph = 0; ph < A.mix.nph; ++ph) r += (double)A.cons[dim + 2 + ph][id];
            double rs = dmax_ref(r, 1e-14);
            A.rho[id] = rs;
            double u[3];
            for (int d = 0; d < dim; ++d) {
                u[d] = (double)A.cons[1 + d][id] / rs;
                A.vel[d][id] = u[d];
            }
            double ke = 0.5 * rs * u[0] * u[0];
            for (int d = 1; d < dim; ++d) ke += 0.5 * rs * u[d] * u[d];
            double sInv = 0.0, sPi = 0.0;
            for (int ph = 0; ph < A.mix.nph; ++ph) {
                double a = (double)A.cons[dim + 2 + A.mix.nph + ph][id];
                double gm1 = A.mix.gamma[ph] - 1.0;
                sInv += a / gm1;
                sPi += a * A.mix.gamma[ph] * A.mix.pInf[ph] / gm1;
            }
            A.pres[id] = A.mix.eoff ? ((double)A.cons[dim + 1][id] - ke) / sInv
                                    : (((double)A.cons[dim + 1][id] - ke) - sPi) / sInv;
        } else {
            double mom[3], u[3], p;
            for (int d = 0; d < dim; ++d) mom[d] = (double)A.cons[1 + d][id];
            double rho = (double)A.cons[0][id];
            A.rho[id] = rho;
            if (dim == 1) cons_to_prim<1>(A.eos, rho, mom, (double)A.cons[2][id], u, p);
            else if (dim == 2) cons_to_prim<2>(A.eos, rho, mom, (double)A.cons[3][id], u, p);
            else cons_to_prim<3>(A.eos, rho, mom, (double)A.cons[4][id], u, p);
            for (int d = 0; d < dim; ++d) A.vel[d][id] = u[d];
            A.pres[id] = p;
        }
    }
}

// Stored <-> total energy of an FP32 state whose energy field excludes the
// stiffened-gas constant (Eos::eoff / Mix::eoff), over every padded cell:
// toStored: E (FP64, e64) -> E - offset (FP32, into the state);
// otherwise: E_stored + offset -> e64.  The offset uses the volume fractions
// stored in the same state.
struct EnergyArgs {
    size_t total;
    float* E;                          // stored energy field
    const float* alpha[kMaxPhases];    // stored volume fractions (multiphase)
    int nph;
    double eC[kMaxPhases];             // multiphase: gamma_k pInf_k / (gamma_k - 1)
    double c;                          // single phase: gamma pInf / (gamma - 1)
    double* e64;
    int toStored;
};

__global__ void __launch_bounds__(256) energy_offset_kernel(EnergyArgs A) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.total; t += (size_t)gridDim.x * blockDim.x) {
        double o = A.c;
        for (int ph = 0; ph < A.nph; ++ph) o += (double)A.alpha[ph][t] * A.eC[ph];
        if (A.toStored) A.E[t] = (float)(A.e64[t] - o);
        else A.e64[t] = (double)A.E[t] + o;
    }
}

// Ghost fill of one direction for a set of FP64 fields, onion-peel order
// (x, then y over all x-ghost columns, then z over all x/y ghosts), as
// RectilinearMesh::fillGhostX/Y/Z (src/RectilinearMesh.cpp:137-418) and
// fillScalarGhosts (:424-624).  comp[f] = 0 for scalars, 1+d for the d-th
// velocity/momentum component (sign flips of wall rules).
constexpr int kMaxFill = 24;
struct FillArgs {
    Grid g;
    int dir;
    int scalar;  // 1: scalar rule (periodic wraps, everything else zero-gradient)
    int nf;
    double* f[kMaxFill];
    int comp[kMaxFill];
};

__global__ void __launch_bounds__(256) fill_dir_kernel(FillArgs A) {
    const Grid& g = A.g;
    int n0, n1, ngd, nd;
    if (A.dir == 0) { n0 = g.ny; n1 = g.nz; ngd = g.ngx; nd = g.nx; }
    else if (A.dir == 1) { n0 = g.nx + 2 * g.ngx; n1 = g.nz; ngd = g.ngy; nd = g.ny; }
    else { n0 = g.nx + 2 * g.ngx; n1 = g.ny + 2 * g.ngy; ngd = g.ngz; nd = g.nz; }
    const size_t total = (size_t)n0 * n1 * 2 * ngd;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (size_t)gridDim.x * blockDim.x) {
        int gg = (int)(t % (2 * ngd));
        size_t rest = t / (2 * ngd);
        int a = (int)(rest % n0), b = (int)(rest / n0);
        int side = gg / ngd, gi = gg % ngd + 1;  // ghost layer 1..ng
        int c = side ? nd - 1 + gi : -gi;
        bool fn = false, ft = false;
        int s;
        if (A.scalar) s = map_axis_scalar(c, nd, g.sbc[2 * A.dir], g.sbc[2 * A.dir + 1]);
        else s = map_axis(c, nd, g.bc[2 * A.dir], g.bc[2 * A.dir + 1], fn, ft);
        int i, j, k, si, sj, sk;
        if (A.dir == 0) { j = a; k = b; i = c; si = s; sj = j; sk = k; }
        else if (A.dir == 1) { i = a - g.ngx; k = b; j = c; si = i; sj = s; sk = k; }
        else { i = a - g.ngx; j = b - g.ngy; k = c; si = i; sj = j; sk = s; }
        size_t dst = cell_index(g, i, j, k), src = cell_index(g, si, sj, sk);
        for (int q = 0; q < A.nf; ++q) {
            double v = A.f[q][src];
            int cp = A.comp[q];
            if (cp > 0) {
                bool flip = (cp - 1 == A.dir) ? fn : ft;
                if (flip) v = -v;
            }
            A.f[q][dst] = v;
        }
    }
}

}  // namespace sifv
