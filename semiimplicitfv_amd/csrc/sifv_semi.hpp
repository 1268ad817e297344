// sifv_semi.hpp - kernels of the semi-implicit pressure step:
//   semi_setup_kernel    rho c^2 (unless the stage update wrote it), div(u*),
//                        rhs = p_adv - rho c^2 dt div(u*),
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
    const double* pres;       // stage-start pressure (p0 itself when the stage update wrote it there)
    const double* vel[3];     // star velocity (ghosts filled) / corrected velocity (out)
    double* velOut[3];
    const double* pAdv;
    double* rhoc2;            // written here unless the stage update wrote it (rc2Given)
    double* rhs;
    double* p0;               // Jacobi initial guess
    int rc2Given;
    double* resid;            // per-iteration max |pNew - p| (zeroed here)
    int maxIter;
    const double* dt;
};

__device__ __forceinline__ void interior_ijk(const Grid& g, size_t t, int& i, int& j, int& k) {
    i = (int)(t % g.nx);
    j = (int)((t / g.nx) % g.ny);
    k = (int)(t / ((size_t)g.nx * g.ny));
}

// Brick walk of the set-up and the correction: a block of 64 x kSemiBrickRows
// threads owns 64 columns x kSemiBrickRows rows and marches kSemiBrickPlanes
// planes of z, one cell per thread and plane, so that the y and z neighbours
// a cell's stencil reads were just read by the same block (512^3 bubble:
// set-up 1945 -> 1641 us, correction 3911 -> 3358 us against a grid-stride
// walk over rows).  Blocks go to the 8 XCDs in contiguous runs, so
// neighbouring bricks share an L2.  semi_brick gives the thread's column and
// plane range; false when the column lies outside the mesh.
constexpr int kSemiBrickRows = 4;
constexpr int kSemiBrickPlanes = 8;

__device__ __forceinline__ bool semi_brick(const Grid& g, int& i, int& j, int& k0, int& k1) {
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    const int gx = gridDim.x, gy = gridDim.y, nb = gx * gy * gridDim.z;
    if (nb % 8 == 0) {
        const int L = bx + gx * (by + gy * bz);
        const int t = (L % 8) * (nb / 8) + L / 8;
        bx = t % gx;
        by = (t / gx) % gy;
        bz = t / (gx * gy);
    }
    i = bx * 64 + (int)threadIdx.x;
    j = by * kSemiBrickRows + (int)threadIdx.y;
    k0 = bz * kSemiBrickPlanes;
    k1 = min(g.nz, k0 + kSemiBrickPlanes);
    return i < g.nx && j < g.ny;
}

__global__ void __launch_bounds__(64 * kSemiBrickRows) semi_setup_kernel(SemiArgs A) {
    const Grid& g = A.g;
    const double dt = *A.dt;
    {
        const size_t nthr = (size_t)blockDim.x * blockDim.y;
        const size_t lin = ((size_t)blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z)) *
                               nthr + threadIdx.y * blockDim.x + threadIdx.x;
        const size_t stride = (size_t)gridDim.x * gridDim.y * gridDim.z * nthr;
        for (size_t t = lin; t < (size_t)A.maxIter; t += stride) A.resid[t] = 0.0;
    }
    int i, j, k0, k1;
    if (!semi_brick(g, i, j, k0, k1)) return;
    for (int k = k0; k < k1; ++k) {
        const size_t id = cell_index(g, i, j, k);
        double rc2;
        if (A.rc2Given) {
            rc2 = A.rhoc2[id];
        } else {
            const double p = A.pres[id];
            const double rho = A.rho[id];
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
            rc2 = rho * c * c;
            A.rhoc2[id] = rc2;
        }
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
        A.rhs[id] = A.pAdv[id] - rc2 * dt * div;
        // the initial guess is the stage-start pressure
        if (A.p0 != A.pres) A.p0[id] = A.pres[id];
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
    // 0: the velocity and pressure arrays are left as they are (a stage that
    // is not the step's last: the next stage's update overwrites both before
    // anything reads them)
    int storePrim;
    const double* dt;
};

__global__ void __launch_bounds__(64 * kSemiBrickRows) semi_correct_kernel(CorrectArgs A) {
    const Grid& g = A.g;
    const double dt = *A.dt;
    const int dim = A.dim;
    // sigma is zero without IGR (A.sigma null): p + sigma is p then
    const bool sig = A.sigma != nullptr;
    const size_t st[3] = {1, (size_t)g.nxT, (size_t)g.nxT * g.nyT};
    int i, j, k0, k1;
    if (!semi_brick(g, i, j, k0, k1)) return;
    for (int k = k0; k < k1; ++k) {
        const size_t id = cell_index(g, i, j, k);
        const double pc = sig ? A.p[id] + A.sigma[id] : A.p[id];
        const double rho = A.cons[id];
        const double rs = dmax_ref(rho, 1e-14);
        const double h[3] = {g.dxw[i + g.ngx], g.dyw[j + g.ngy], g.dzw[k + g.ngz]};
        // neighbours of an interior cell are read directly, the others through
        // the scalar ghost rule
        const bool inner = i > 0 && i < g.nx - 1 && (dim < 2 || (j > 0 && j < g.ny - 1)) &&
                           (dim < 3 || (k > 0 && k < g.nz - 1));
        double u[3];
        for (int d = 0; d < dim; ++d) {
            const int di = d == 0, dj = d == 1, dk = d == 2;
            double pm, pp;
            if (inner) {
                pm = sig ? A.p[id - st[d]] + A.sigma[id - st[d]] : A.p[id - st[d]];
                pp = sig ? A.p[id + st[d]] + A.sigma[id + st[d]] : A.p[id + st[d]];
            } else {
                pm = scalar_at(g, A.p, i - di, j - dj, k - dk);
                pp = scalar_at(g, A.p, i + di, j + dj, k + dk);
                if (sig) {
                    pm += scalar_at(g, A.sigma, i - di, j - dj, k - dk);
                    pp += scalar_at(g, A.sigma, i + di, j + dj, k + dk);
                }
            }
            const double pl = 0.5 * (pm + pc);
            const double pr = 0.5 * (pc + pp);
            double* mom = A.cons + (size_t)(1 + d) * A.total;
            const double m = mom[id] - dt * (pr - pl) / h[d];
            mom[id] = m;
            u[d] = m / rs;
            if (A.storePrim) A.vel[d][id] = u[d];
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
        if (A.storePrim) A.pres[id] = p;
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

// ---- Tiled Jacobi sweeps (2D/3D) ----------------------------------------------------
// The Jacobi kernels above (the FP64 update with its max-norm, the FP32 correction
// sweep and the FP64 pass of the mixed solve) organised for HBM: a block of
// 64 x kTileRows threads owns a tile of 64 columns x kTileRows rows and marches it
// through up to kTilePlanes planes of z.  The z-neighbours of p and rho stay in
// registers; each plane's p and rho tile with a one-cell x/y halo is staged in LDS,
// double-buffered by plane parity (one barrier per plane).  Lanes on the halo
// column i == nx / row j == ny load it; the tile's outer halo comes from extra
// loads by lanes 0 / 63 and rows 0 / kTileRows-1.  p goes through the scalar ghost
// rule, rho through its filled ghosts.  Per cell the arithmetic is the elementwise
// kernels' expression for expression, so both forms give identical bits.
#ifndef SIFV_JACOBI_TILE_ROWS
#define SIFV_JACOBI_TILE_ROWS 4
#endif
#ifndef SIFV_JACOBI_TILE_PLANES
#define SIFV_JACOBI_TILE_PLANES 16
#endif
constexpr int kTileRows = SIFV_JACOBI_TILE_ROWS;
constexpr int kTilePlanes = SIFV_JACOBI_TILE_PLANES;

enum TileMode : int { TILE_SWEEP64 = 0, TILE_SWEEP32 = 1, TILE_PASS = 2 };

template <typename T>
struct JacobiTileArgs {
    Grid g;
    int dim;
    const T* rho;        // padded, ghosts filled (rho / rho32)
    const T* co;         // FP64: rho c^2 (times dt^2 here); FP32: dt^2 rho c^2
    const T* rhs;        // FP64: rhs; FP32: residual r32
    const T* pin;        // iterate; FP32: null on the first sweep (zero)
    T* pout;             // sweep: the new iterate; pass: pin + din
    double* resid;       // FP64: resid[iter] accumulates max |pNew - p|
    int iter;
    double tol;
    const double* dt;    // FP64 only
    // pass of the mixed solve (jacobi_mix_pass_kernel's outputs)
    const float* din;    // correction from the FP32 sweeps (null: none)
    double* pn;          // FP64 Jacobi update of pin + din
    float* r32;          // residual rhs - A (pin + din)
    float* co32;         // dt^2 rho c^2 in FP32
};

template <typename T>
__device__ __forceinline__ T tile_coef(T r, T dL, T hc) {
    if constexpr (sizeof(T) == 8) return 1.0 / (dmax_ref(r, 1e-14) * dL * hc);
    else return __builtin_amdgcn_rcpf(fmaxf(r, 1e-14f) * dL * hc);
}

template <typename T, int MODE>
__global__ void __launch_bounds__(64 * kTileRows) jacobi_tile_kernel(JacobiTileArgs<T> A) {
    constexpr bool F64 = sizeof(T) == 8;
    constexpr bool PASS = MODE == TILE_PASS;
    static_assert(F64 == (MODE != TILE_SWEEP32), "FP64 sweeps and passes, FP32 sweeps");
    constexpr int TY = kTileRows;
    if constexpr (MODE == TILE_SWEEP64) {
        // converged at an earlier iteration: nothing left to do
        if (A.iter > 0 && A.resid[A.iter - 1] < A.tol) return;
    }
    __shared__ T sp[2][TY + 2][66];
    __shared__ T sr[2][TY + 2][66];
    __shared__ double sh[TY];
    const Grid& g = A.g;
    const bool z = A.dim >= 3;
    // launch order deals blocks round-robin to the 8 XCDs: hand each XCD a
    // contiguous run of tiles so that neighbouring tiles share its L2
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    {
        const int gx = gridDim.x, gy = gridDim.y, nb = gx * gy * gridDim.z;
        if (nb % 8 == 0) {
            const int L = bx + gx * (by + gy * bz);
            const int t = (L % 8) * (nb / 8) + L / 8;
            bx = t % gx;
            by = (t / gx) % gy;
            bz = t / (gx * gy);
        }
    }
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = bx * 64, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int k0 = z ? bz * kTilePlanes : 0, k1 = z ? min(g.nz, k0 + kTilePlanes) : 1;
    const long sy = g.nxT, sz = (long)g.nxT * g.nyT;
    auto raw = [&](int a, int b, int c) {
        return (long)(a + g.ngx) + sy * (long)(b + g.ngy) + sz * (long)(c + g.ngz);
    };
    const bool active = i < g.nx && j < g.ny;
    const bool loads = i <= g.nx && j <= g.ny;
    const int ic = min(i, g.nx), jc = min(j, g.ny);
    const long pofs = (long)(map_axis_scalar(ic, g.nx, g.sbc[0], g.sbc[1]) - ic) +
                      (long)(map_axis_scalar(jc, g.ny, g.sbc[2], g.sbc[3]) - jc) * sy;
    const bool xh = (tx == 0 || (tx == 63 && i0 + 64 <= g.nx)) && j < g.ny;
    const bool yh = (ty == 0 || (ty == TY - 1 && j0 + TY <= g.ny)) && i < g.nx;
    const int hi = tx == 0 ? i0 - 1 : i0 + 64, hj = ty == 0 ? j0 - 1 : j0 + TY;
    const int xs = tx == 0 ? 0 : 65, ys = ty == 0 ? 0 : TY + 1;
    const long hx = raw(hi, jc, k0), hy = raw(ic, hj, k0);
    const long hxP = (long)(map_axis_scalar(hi, g.nx, g.sbc[0], g.sbc[1]) - hi);
    const long hyP = (long)(map_axis_scalar(hj, g.ny, g.sbc[2], g.sbc[3]) - hj) * sy;
    const bool hasP = A.pin != nullptr;
    const bool hasD = PASS && A.din != nullptr;
    // p of the sweep: the iterate; of the pass: iterate + FP32 correction
    auto ldp = [&](long q) -> T {
        if constexpr (PASS) return hasD ? A.pin[q] + (double)A.din[q] : A.pin[q];
        else return A.pin[q];
    };
    auto zofs = [&](int k) { return (long)(map_axis_scalar(k, g.nz, g.sbc[4], g.sbc[5]) - k) * sz; };
    T hcx = 0, dLx = 0, dRx = 0, hcy = 0, dLy = 0, dRy = 0;
    if (active) {
        const T hlx = (T)g.dxw[i - 1 + g.ngx], hrx = (T)g.dxw[i + 1 + g.ngx];
        hcx = (T)g.dxw[i + g.ngx];
        dLx = T(0.5) * (hlx + hcx);
        dRx = T(0.5) * (hcx + hrx);
        const T hly = (T)g.dyw[j - 1 + g.ngy], hry = (T)g.dyw[j + 1 + g.ngy];
        hcy = (T)g.dyw[j + g.ngy];
        dLy = T(0.5) * (hly + hcy);
        dRy = T(0.5) * (hcy + hry);
    }
    double dt2 = 0.0;
    if constexpr (F64) {
        const double dt = *A.dt;
        dt2 = dt * dt;
    }
    long id = raw(ic, jc, k0);
    T pzm = 0, pc = 0, pzp = 0, rzm = 0, rc = 0, rzp = 0;
    T hpx = 0, hrx = 0, hpy = 0, hry = 0;
    if (loads) {
        rc = A.rho[id];
        if (hasP) pc = ldp(id + pofs);
        if (z) {
            rzm = A.rho[id - sz];
            if (hasP) pzm = ldp(id + pofs - sz + zofs(k0 - 1));
        }
    }
    if (xh) {
        hrx = A.rho[hx];
        if (hasP) hpx = ldp(hx + hxP);
    }
    if (yh) {
        hry = A.rho[hy];
        if (hasP) hpy = ldp(hy + hyP);
    }
    double mr = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int b = k & 1;
        sp[b][ty + 1][tx + 1] = pc;
        sr[b][ty + 1][tx + 1] = rc;
        if (xh) {
            sp[b][ty + 1][xs] = hpx;
            sr[b][ty + 1][xs] = hrx;
        }
        if (yh) {
            sp[b][ys][tx + 1] = hpy;
            sr[b][ys][tx + 1] = hry;
        }
        // operands of this plane not staged in LDS, and the next plane's halo
        if (loads && z) {
            rzp = A.rho[id + sz];
            if (hasP) pzp = ldp(id + pofs + sz + zofs(k + 1));
        }
        T q2 = 0, qs = 0;
        if (active) {
            q2 = A.co[id];
            qs = A.rhs[id];
        }
        const long dk = (long)(k + 1 - k0) * sz;
        T nhpx = 0, nhrx = 0, nhpy = 0, nhry = 0;
        if (k + 1 < k1) {
            if (xh) {
                nhrx = A.rho[hx + dk];
                if (hasP) nhpx = ldp(hx + dk + hxP);
            }
            if (yh) {
                nhry = A.rho[hy + dk];
                if (hasP) nhpy = ldp(hy + dk + hyP);
            }
        }
        __syncthreads();
        if (active) {
            T off = 0, diag = 0;
            {
                const T rL = T(0.5) * (rc + sr[b][ty + 1][tx]), rR = T(0.5) * (rc + sr[b][ty + 1][tx + 2]);
                const T cL = tile_coef(rL, dLx, hcx), cR = tile_coef(rR, dRx, hcx);
                off += cL * sp[b][ty + 1][tx] + cR * sp[b][ty + 1][tx + 2];
                diag += cL + cR;
            }
            {
                const T rL = T(0.5) * (rc + sr[b][ty][tx + 1]), rR = T(0.5) * (rc + sr[b][ty + 2][tx + 1]);
                const T cL = tile_coef(rL, dLy, hcy), cR = tile_coef(rR, dRy, hcy);
                off += cL * sp[b][ty][tx + 1] + cR * sp[b][ty + 2][tx + 1];
                diag += cL + cR;
            }
            if (z) {
                const T hl = (T)g.dzw[k - 1 + g.ngz], hc = (T)g.dzw[k + g.ngz], hr = (T)g.dzw[k + 1 + g.ngz];
                const T rL = T(0.5) * (rc + rzm), rR = T(0.5) * (rc + rzp);
                const T cL = tile_coef(rL, T(0.5) * (hl + hc), hc), cR = tile_coef(rR, T(0.5) * (hc + hr), hc);
                off += cL * pzm + cR * pzp;
                diag += cL + cR;
            }
            if constexpr (PASS) {
                const double co = q2 * dt2;
                const double den = 1.0 + co * diag;
                const double pn = (qs + co * off) / den;
                A.pout[id] = pc;
                A.pn[id] = pn;
                A.r32[id] = (float)(den * (pn - pc));
                A.co32[id] = (float)co;
                mr = dmax_ref(mr, fabs(pn - pc));
            } else if constexpr (F64) {
                const double co = q2 * dt2;
                const double den = 1.0 + co * diag;
                const double pn = (qs + co * off) / den;
                A.pout[id] = pn;
                mr = dmax_ref(mr, fabs(pn - pc));
            } else {
                A.pout[id] = (qs + q2 * off) * __builtin_amdgcn_rcpf(1.0f + q2 * diag);
            }
        }
        pzm = pc;
        pc = pzp;
        rzm = rc;
        rc = rzp;
        hpx = nhpx;
        hrx = nhrx;
        hpy = nhpy;
        hry = nhry;
        id += sz;
    }
    if constexpr (F64) {
        const double w = wave_max(mr);
        if (tx == 0) sh[ty] = w;
        __syncthreads();
        if (tx == 0 && ty == 0) {
            double m = 0.0;
            for (int q = 0; q < TY; ++q) m = dmax_ref(m, sh[q]);
            atomicMax(reinterpret_cast<unsigned long long*>(&A.resid[A.iter]),
                      (unsigned long long)__double_as_longlong(m));
        }
    }
}

}  // namespace sifv
