// sifv_semi.hpp - kernels of the semi-implicit pressure step:
//   semi_setup_kernel    rho c^2, div(u*), rhs = p_adv - rho c^2 dt div(u*),
//                        initial guess (SemiImplicitSolver::solvePressure,
//                        src/SemiImplicitSolver.cpp:533-575, computeDivergence :680-717)
//   jacobi_kernel        one Jacobi sweep of (I + dt^2 rho c^2 L) p = rhs with the
//                        variable-coefficient face Laplacian and the max-norm update
//                        (JacobiPressureSolver::solve, src/JacobiPressureSolver.cpp:57-108;
//                        pressureLaplacian, src/PressureLaplacian.cpp:6-60)
//   semi_correct_kernel  momentum correction, velocity and total energy from the solved
//                        pressure (SemiImplicitSolver::correctionStep, :590-678)
// All work on interior cells; 1-deep neighbours of p and sigma go through the
// scalar ghost rule (fillScalarGhosts), rho and the star velocity are read
// from arrays whose ghosts the PRIM fill has refreshed.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

struct SemiArgs {
    Grid g;
    Eos eos;
    Mix mix;
    int dim, nph;
    size_t total;
    const double* rho;        // star density (ghosts filled)
    const double* ar[kMaxPhases];
    const double* al[kMaxPhases];
    const double* pres;       // stage-start pressure
    const double* vel[3];     // star velocity (ghosts filled) / corrected velocity (out)
    double* velOut[3];
    const double* pAdv;
    double* rhoc2;
    double* rhs;
    double* p0;               // Jacobi initial guess
    double* resid;            // per-iteration max |pNew - p| (zeroed here)
    int maxIter;
    const double* dt;
};

__device__ __forceinline__ void interior_ijk(const Grid& g, size_t t, int& i, int& j, int& k) {
    i = (int)(t % g.nx);
    j = (int)((t / g.nx) % g.ny);
    k = (int)(t / ((size_t)g.nx * g.ny));
}

__global__ void __launch_bounds__(256) semi_setup_kernel(SemiArgs A) {
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const double dt = *A.dt;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)A.maxIter; t += stride)
        A.resid[t] = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double rho = A.rho[id];
        const double p = A.pres[id];
        double c;
        if (A.nph > 0) {
            double a[kMaxPhases], ar[kMaxPhases];
            for (int ph = 0; ph < A.nph; ++ph) {
                a[ph] = A.al[ph][id];
                ar[ph] = A.ar[ph][id];
            }
            c = mix_sound(A.mix, rho, p, a, ar, A.nph);
        } else {
            double pinf = A.eos.stiffened ? A.eos.pInf : 0.0;
            c = sqrt(A.eos.gamma * dmax_ref(p + pinf, 1e-14) / dmax_ref(rho, 1e-14));
        }
        const double rc2 = rho * c * c;
        // central divergence of the star velocity (ghosts filled by the PRIM rule)
        double div = 0.0;
        const size_t sx = 1, sy = (size_t)g.nxT, sz = (size_t)g.nxT * g.nyT;
        const size_t st[3] = {sx, sy, sz};
        const double h[3] = {g.dxw[i + g.ngx], g.dyw[j + g.ngy], g.dzw[k + g.ngz]};
        for (int d = 0; d < A.dim; ++d) {
            const double* v = A.vel[d];
            double fl = 0.5 * (v[id - st[d]] + v[id]);
            double fr = 0.5 * (v[id] + v[id + st[d]]);
            div += (fr - fl) / h[d];
        }
        A.rhoc2[id] = rc2;
        A.rhs[id] = A.pAdv[id] - rc2 * dt * div;
        A.p0[id] = p;
    }
}

struct JacobiArgs {
    Grid g;
    int dim;
    const double* rho;    // ghosts as given (PRIM-filled)
    const double* rhoc2;
    const double* rhs;
    const double* pin;    // current iterate (ghosts through the scalar rule)
    double* pout;
    double* resid;        // resid[iter] accumulates max |pNew - p|
    int iter;
    double tol;
    const double* dt;
};

__device__ __forceinline__ double scalar_at(const Grid& g, const double* a, int i, int j, int k) {
    int si = map_axis_scalar(i, g.nx, g.sbc[0], g.sbc[1]);
    int sj = g.dim >= 2 ? map_axis_scalar(j, g.ny, g.sbc[2], g.sbc[3]) : j;
    int sk = g.dim >= 3 ? map_axis_scalar(k, g.nz, g.sbc[4], g.sbc[5]) : k;
    return a[cell_index(g, si, sj, sk)];
}

__global__ void __launch_bounds__(256) jacobi_kernel(JacobiArgs A) {
    // converged at an earlier iteration: nothing left to do
    if (A.iter > 0 && A.resid[A.iter - 1] < A.tol) return;
    __shared__ double sh[256 / kWave];
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const double dt = *A.dt;
    const double dt2 = dt * dt;
    double mr = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double co = A.rhoc2[id] * dt2;
        const double pc = A.pin[id];
        const double rc = A.rho[id];
        double off = 0.0, diag = 0.0;
        // X, then Y, then Z (pressureLaplacian, src/PressureLaplacian.cpp:13-53)
        {
            const double hl = g.dxw[i - 1 + g.ngx], hc = g.dxw[i + g.ngx], hr = g.dxw[i + 1 + g.ngx];
            const double rL = 0.5 * (rc + A.rho[id - 1]), rR = 0.5 * (rc + A.rho[id + 1]);
            const double dL = 0.5 * (hl + hc), dR = 0.5 * (hc + hr);
            const double cL = 1.0 / (dmax_ref(rL, 1e-14) * dL * hc);
            const double cR = 1.0 / (dmax_ref(rR, 1e-14) * dR * hc);
            off += cL * scalar_at(g, A.pin, i - 1, j, k) + cR * scalar_at(g, A.pin, i + 1, j, k);
            diag += cL + cR;
        }
        if (A.dim >= 2) {
            const size_t s = (size_t)g.nxT;
            const double hl = g.dyw[j - 1 + g.ngy], hc = g.dyw[j + g.ngy], hr = g.dyw[j + 1 + g.ngy];
            const double rL = 0.5 * (rc + A.rho[id - s]), rR = 0.5 * (rc + A.rho[id + s]);
            const double dL = 0.5 * (hl + hc), dR = 0.5 * (hc + hr);
            const double cL = 1.0 / (dmax_ref(rL, 1e-14) * dL * hc);
            const double cR = 1.0 / (dmax_ref(rR, 1e-14) * dR * hc);
            off += cL * scalar_at(g, A.pin, i, j - 1, k) + cR * scalar_at(g, A.pin, i, j + 1, k);
            diag += cL + cR;
        }
        if (A.dim >= 3) {
            const size_t s = (size_t)g.nxT * g.nyT;
            const double hl = g.dzw[k - 1 + g.ngz], hc = g.dzw[k + g.ngz], hr = g.dzw[k + 1 + g.ngz];
            const double rL = 0.5 * (rc + A.rho[id - s]), rR = 0.5 * (rc + A.rho[id + s]);
            const double dL = 0.5 * (hl + hc), dR = 0.5 * (hc + hr);
            const double cL = 1.0 / (dmax_ref(rL, 1e-14) * dL * hc);
            const double cR = 1.0 / (dmax_ref(rR, 1e-14) * dR * hc);
            off += cL * scalar_at(g, A.pin, i, j, k - 1) + cR * scalar_at(g, A.pin, i, j, k + 1);
            diag += cL + cR;
        }
        const double den = 1.0 + co * diag;
        const double pn = (A.rhs[id] + co * off) / den;
        A.pout[id] = pn;
        mr = dmax_ref(mr, fabs(pn - pc));
    }
    // block max, then one atomic per block on the bit pattern (non-negative
    // doubles order like their unsigned 64-bit patterns)
    double w = wave_max(mr);
    const int wv = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[wv] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int q = 0; q < (int)(blockDim.x / kWave); ++q) m = dmax_ref(m, sh[q]);
        atomicMax(reinterpret_cast<unsigned long long*>(&A.resid[A.iter]),
                  (unsigned long long)__double_as_longlong(m));
    }
}

// Writes interior cells of `src` into `dst` (ghosts of dst untouched).
__global__ void __launch_bounds__(256) copy_interior_kernel(Grid g, const double* src, double* dst) {
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        dst[id] = src[id];
    }
}

struct CorrectArgs {
    Grid g;
    Eos eos;
    Mix mix;
    int dim, nph;
    size_t total;
    double* cons;             // stage output buffer: star momentum in, corrected momentum / energy out
    const double* p;          // solved pressure (Jacobi iterate)
    const double* sigma;
    double* vel[3];
    double* pres;
    const double* dt;
};

__global__ void __launch_bounds__(256) semi_correct_kernel(CorrectArgs A) {
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const double dt = *A.dt;
    const int dim = A.dim;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double pc = A.p[id] + A.sigma[id];
        const double rho = A.cons[id];
        const double rs = dmax_ref(rho, 1e-14);
        const double h[3] = {g.dxw[i + g.ngx], g.dyw[j + g.ngy], g.dzw[k + g.ngz]};
        double u[3];
        for (int d = 0; d < dim; ++d) {
            const int di = d == 0, dj = d == 1, dk = d == 2;
            const double pm = scalar_at(g, A.p, i - di, j - dj, k - dk) + scalar_at(g, A.sigma, i - di, j - dj, k - dk);
            const double pp = scalar_at(g, A.p, i + di, j + dj, k + dk) + scalar_at(g, A.sigma, i + di, j + dj, k + dk);
            const double pl = 0.5 * (pm + pc);
            const double pr = 0.5 * (pc + pp);
            double* mom = A.cons + (size_t)(1 + d) * A.total;
            const double m = mom[id] - dt * (pr - pl) / h[d];
            mom[id] = m;
            u[d] = m / rs;
            A.vel[d][id] = u[d];
        }
        double ke = 0.5 * rho * u[0] * u[0];
        for (int d = 1; d < dim; ++d) ke += 0.5 * rho * u[d] * u[d];
        const double p = A.p[id];
        double E;
        if (A.nph > 0) {
            // MixtureEOS::mixtureTotalEnergy (src/MixtureEOS.cpp:65-74)
            E = ke;
            for (int ph = 0; ph < A.nph; ++ph) {
                const double a = A.cons[(size_t)(dim + 2 + A.nph + ph) * A.total + id];
                E += a * (p + A.mix.gamma[ph] * A.mix.pInf[ph]) / (A.mix.gamma[ph] - 1.0);
            }
        } else {
            const double pinf = A.eos.stiffened ? A.eos.pInf : 0.0;
            E = (p + A.eos.gamma * pinf) / (A.eos.gamma - 1.0) + ke;
        }
        A.cons[(size_t)(dim + 1) * A.total + id] = E;
        A.pres[id] = p;
    }
}

}  // namespace sifv

namespace sifv {

// ---- Jacobi with FP32 sweeps and FP64 residual refinement ----------------------
// The system (I + dt^2 rho c^2 L) p = rhs is written as den_i p_i - co_i sum_j c_ij p_j = rhs_i
// (den = 1 + co sum_j c_ij, co = dt^2 rho c^2).  An FP64 pass from the current
// iterate p computes the reference's Jacobi update pn = (rhs + co off(p)) / den,
// its size max |pn - p| (the reference's convergence measure) and the residual
// r = den (pn - p) = rhs - A p; FP32 Jacobi sweeps then solve A d = r from d = 0,
// and the next FP64 pass starts from p + d.  When the FP64 update size drops
// below tol, pn is the answer, as in the reference.
struct JacobiMixArgs {
    Grid g;
    int dim;
    const double* rho;      // ghosts as given (PRIM-filled)
    const double* rhoc2;
    const double* rhs;
    const double* pin;      // iterate before the correction
    const float* din;       // correction from the FP32 sweeps (null: none)
    double* pout;           // pin + din (interior)
    double* pn;             // FP64 Jacobi update of pout
    float* r32;             // residual rhs - A pout
    float* co32;            // dt^2 rho c^2 (FP32, written by every pass)
    float* rho32;           // rho (FP32, padded, written by every pass incl. ghosts it reads)
    double* resid;          // resid[slot] accumulates max |pn - pout|
    int slot;
    const double* dt;
};

__device__ __forceinline__ double pd_at(const Grid& g, const double* p, const float* d, int i, int j, int k) {
    int si = map_axis_scalar(i, g.nx, g.sbc[0], g.sbc[1]);
    int sj = g.dim >= 2 ? map_axis_scalar(j, g.ny, g.sbc[2], g.sbc[3]) : j;
    int sk = g.dim >= 3 ? map_axis_scalar(k, g.nz, g.sbc[4], g.sbc[5]) : k;
    const size_t q = cell_index(g, si, sj, sk);
    return d ? p[q] + (double)d[q] : p[q];
}

__global__ void __launch_bounds__(256) jacobi_mix_pass_kernel(JacobiMixArgs A) {
    __shared__ double sh[256 / kWave];
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    const double dt = *A.dt;
    const double dt2 = dt * dt;
    double mr = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double co = A.rhoc2[id] * dt2;
        const double pc = pd_at(g, A.pin, A.din, i, j, k);
        const double rc = A.rho[id];
        double off = 0.0, diag = 0.0;
        const size_t st[3] = {1, (size_t)g.nxT, (size_t)g.nxT * g.nyT};
        const double* hw[3] = {g.dxw, g.dyw, g.dzw};
        const int ng[3] = {g.ngx, g.ngy, g.ngz};
        const int ii[3] = {i, j, k};
        for (int d = 0; d < A.dim; ++d) {
            const double hl = hw[d][ii[d] - 1 + ng[d]], hc = hw[d][ii[d] + ng[d]], hr = hw[d][ii[d] + 1 + ng[d]];
            const double rL = 0.5 * (rc + A.rho[id - st[d]]), rR = 0.5 * (rc + A.rho[id + st[d]]);
            const double dL = 0.5 * (hl + hc), dR = 0.5 * (hc + hr);
            const double cL = 1.0 / (dmax_ref(rL, 1e-14) * dL * hc);
            const double cR = 1.0 / (dmax_ref(rR, 1e-14) * dR * hc);
            const double pm = pd_at(g, A.pin, A.din, i - (d == 0), j - (d == 1), k - (d == 2));
            const double pp = pd_at(g, A.pin, A.din, i + (d == 0), j + (d == 1), k + (d == 2));
            off += cL * pm + cR * pp;
            diag += cL + cR;
        }
        const double den = 1.0 + co * diag;
        const double pn = (A.rhs[id] + co * off) / den;
        A.pout[id] = pc;
        A.pn[id] = pn;
        A.r32[id] = (float)(den * (pn - pc));
        A.co32[id] = (float)co;
        mr = dmax_ref(mr, fabs(pn - pc));
    }
    double w = wave_max(mr);
    const int wv = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[wv] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int q = 0; q < (int)(blockDim.x / kWave); ++q) m = dmax_ref(m, sh[q]);
        atomicMax(reinterpret_cast<unsigned long long*>(&A.resid[A.slot]),
                  (unsigned long long)__double_as_longlong(m));
    }
}

// rho as FP32 over the padded array (ghosts included, as the sweeps read them).
__global__ void __launch_bounds__(256) to_f32_kernel(const double* a, float* b, size_t n) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        b[t] = (float)a[t];
}

struct Jacobi32Args {
    Grid g;
    int dim;
    const float* rho32;
    const float* co32;
    const float* r32;
    const float* din;       // null on the first sweep (d = 0)
    float* dout;
};

__device__ __forceinline__ float d_at(const Grid& g, const float* d, int i, int j, int k) {
    if (!d) return 0.0f;
    int si = map_axis_scalar(i, g.nx, g.sbc[0], g.sbc[1]);
    int sj = g.dim >= 2 ? map_axis_scalar(j, g.ny, g.sbc[2], g.sbc[3]) : j;
    int sk = g.dim >= 3 ? map_axis_scalar(k, g.nz, g.sbc[4], g.sbc[5]) : k;
    return d[cell_index(g, si, sj, sk)];
}

__global__ void __launch_bounds__(256) jacobi32_kernel(Jacobi32Args A) {
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const float co = A.co32[id];
        const float rc = A.rho32[id];
        float off = 0.0f, diag = 0.0f;
        const size_t st[3] = {1, (size_t)g.nxT, (size_t)g.nxT * g.nyT};
        const double* hw[3] = {g.dxw, g.dyw, g.dzw};
        const int ng[3] = {g.ngx, g.ngy, g.ngz};
        const int ii[3] = {i, j, k};
        for (int d = 0; d < A.dim; ++d) {
            const float hl = (float)hw[d][ii[d] - 1 + ng[d]], hc = (float)hw[d][ii[d] + ng[d]];
            const float hr = (float)hw[d][ii[d] + 1 + ng[d]];
            const float rL = 0.5f * (rc + A.rho32[id - st[d]]), rR = 0.5f * (rc + A.rho32[id + st[d]]);
            const float cL = __builtin_amdgcn_rcpf(fmaxf(rL, 1e-14f) * (0.5f * (hl + hc)) * hc);
            const float cR = __builtin_amdgcn_rcpf(fmaxf(rR, 1e-14f) * (0.5f * (hc + hr)) * hc);
            off += cL * d_at(g, A.din, i - (d == 0), j - (d == 1), k - (d == 2)) +
                   cR * d_at(g, A.din, i + (d == 0), j + (d == 1), k + (d == 2));
            diag += cL + cR;
        }
        A.dout[id] = (A.r32[id] + co * off) * __builtin_amdgcn_rcpf(1.0f + co * diag);
    }
}

}  // namespace sifv
