// sifv_pcg.hpp - preconditioned conjugate-gradient pressure solve on the GPU.
//
// The system is the one every PressureSolver of the reference solves,
// (1 + c L) p = rhs with c = rhoc2 dt^2 and L the variable-coefficient face
// Laplacian of pressureLaplacian (src/PressureLaplacian.cpp:6-60), the face
// coefficients of non-periodic domain faces zeroed as
// HyprePressureSolver::assembleSystem does (src/HyprePressureSolver.cpp:161-290).
// Each row scaled by V/c (V the cell volume) gives
//     S_ii = V/c + sum_f w_f,   S_i,nb(f) = -w_f,
//     w_f  = area_f / (max(0.5 (rho_i + rho_nb), 1e-14) * 0.5 (h_i + h_nb)),
// symmetric positive definite (both rows of a face compute the same w_f; the
// host refuses periodic axes whose ghost widths differ from the wrapped
// cells').  Conjugate gradients on S p = (V/c) rhs with the diagonal of S as
// preconditioner, in the HYPRE_StructPCG role (src/HyprePressureSolver.cpp:
// 119-138, 350-358); stop when ||rhs - (1 + c L) p||_2 <= tol ||rhs||_2,
// compared as squares, the original residual being r / (V/c).
//
// Layout: every vector is a padded field like the state's; only interior
// cells are read or written (a periodic neighbour is read at its wrapped
// interior index, so no ghost refresh is needed between kernels).  The face
// weights are recomputed from rho in every product instead of stored: a
// stored set would cost 2*DIM more doubles per cell per product.  Scalars
// (alpha, beta, the dot products) stay on the device; each product and update
// kernel writes per-block partial sums that a one-block kernel folds in fixed
// order, so a solve is deterministic run to run.  A device flag ends the
// iteration: the host enqueues iterations in chunks and polls the flag.
//
// Per iteration, per interior cell: pcg_matvec (read p with its 2*DIM
// neighbours, rho likewise, s; write q: 32 B), pcg_update (read x p r q d s,
// write x r z: 72 B), pcg_dir (read z p, write p: 24 B) -- 128 B/cell,
// HBM-bound.
#pragma once

#include "sifv_aux.hpp"

namespace sifv {

struct PcgArgs {
    Grid g;
    const double* rho;     // padded; ghosts of periodic faces hold the wrapped values
    const double* rhoc2;
    const double* rhs;
    const double* dt;      // device-resident time step
    double* x;             // solution (initial guess in, answer out)
    double *r, *z, *p, *q, *d, *s;
    double* part;          // 3 x nblk per-block partial sums
    int nblk;
    double* sc;            // [0] rz [1] alpha [2] beta [3] bb [4] rr
    int* st;               // [0] done [1] iterations [2] iteration counter [3] zero x
    double tol2;           // tol^2
    int maxIter;
};

// Interior index of the neighbour across the low (side 0) / high (side 1)
// face along dir; *w the face's weight in this cell's row (0 on a
// non-periodic domain face, whose neighbour index is then unused).
template <int DIM>
__device__ __forceinline__ size_t pcg_face(const PcgArgs& A, int dir, int side, int i, int j, int k, double& w) {
    const Grid& g = A.g;
    const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
    const int c = dir == 0 ? i : dir == 1 ? j : k;
    const int sg = side ? 1 : -1;
    const bool edge = side ? c == n - 1 : c == 0;
    const bool wrap = g.sbc[2 * dir + side] == BK_WRAP;
    if (edge && !wrap) {
        w = 0.0;
        return cell_index(g, i, j, k);
    }
    const int ei = dir == 0 ? sg : 0, ej = dir == 1 ? sg : 0, ek = dir == 2 ? sg : 0;
    // rho and the widths at the neighbour slot (ghost copies at a wrap face)
    const size_t id = cell_index(g, i, j, k);
    const size_t gnb = cell_index(g, i + ei, j + ej, k + ek);
    const double rf = 0.5 * (A.rho[id] + A.rho[gnb]);
    double hf;
    if (dir == 0) hf = side ? 0.5 * (g.dxw[i + g.ngx] + g.dxw[i + 1 + g.ngx]) : 0.5 * (g.dxw[i - 1 + g.ngx] + g.dxw[i + g.ngx]);
    else if (dir == 1) hf = side ? 0.5 * (g.dyw[j + g.ngy] + g.dyw[j + 1 + g.ngy]) : 0.5 * (g.dyw[j - 1 + g.ngy] + g.dyw[j + g.ngy]);
    else hf = side ? 0.5 * (g.dzw[k + g.ngz] + g.dzw[k + 1 + g.ngz]) : 0.5 * (g.dzw[k - 1 + g.ngz] + g.dzw[k + g.ngz]);
    const double hx = g.dxw[i + g.ngx], hy = g.dyw[j + g.ngy], hz = g.dzw[k + g.ngz];
    const double area = dir == 0 ? hy * hz : dir == 1 ? hx * hz : hx * hy;
    w = area / (dmax_ref(rf, 1e-14) * hf);
    if (!edge) return gnb;
    // wrapped interior neighbour of a periodic face
    const int wi = dir == 0 ? (side ? 0 : g.nx - 1) : i;
    const int wj = dir == 1 ? (side ? 0 : g.ny - 1) : j;
    const int wk = dir == 2 ? (side ? 0 : g.nz - 1) : k;
    return cell_index(g, wi, wj, wk);
}

// (S v)_i = (V/c)_i v_i + sum_f w_f (v_i - v_nb): neighbour differences first,
// so a smooth v (pressures ~1e5 with small variations) keeps its precision
// where the face terms outweigh V/c by many orders (fine meshes, large dt).
template <int DIM>
__device__ __forceinline__ double pcg_row(const PcgArgs& A, const double* v, double si, int i, int j, int k,
                                          size_t id) {
    const double vi = v[id];
    double acc = si * vi;
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            double w;
            const size_t nb = pcg_face<DIM>(A, dir, side, i, j, k, w);
            acc += w * (vi - v[nb]);
        }
    }
    return acc;
}

// Block sums of up to three values, written to part[q * nblk + blockIdx.x].
__device__ __forceinline__ void pcg_partials(const PcgArgs& A, double* sh, int nv, double a0, double a1, double a2) {
    const double v[3] = {a0, a1, a2};
    for (int q = 0; q < nv; ++q) {
        double t = wave_sum(v[q]);
        const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
        if (l == 0) sh[w] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0.0;
            for (int u = 0; u < (int)(blockDim.x / kWave); ++u) s += sh[u];
            A.part[(size_t)q * A.nblk + blockIdx.x] = s;
        }
        __syncthreads();
    }
}

// Fixed-order fold of the partials of slot q (one block of 256 threads).
__device__ __forceinline__ double pcg_fold(const PcgArgs& A, double* sh, int q) {
    double t = 0.0;
    for (int b = threadIdx.x; b < A.nblk; b += blockDim.x) t += A.part[(size_t)q * A.nblk + b];
    t = wave_sum(t);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = t;
    __syncthreads();
    double s = 0.0;
    for (int u = 0; u < (int)(blockDim.x / kWave); ++u) s += sh[u];
    __syncthreads();
    return s;
}

// Scale, diagonal, initial residual r = (V/c) rhs - S x, z = r / d, p = z;
// partials bb = sum rhs^2, rr = sum (r / s)^2, rz = sum r z.
template <int DIM>
__global__ void __launch_bounds__(256) pcg_setup_kernel(PcgArgs A) {
    __shared__ double sh[256 / kWave];
    const Grid& g = A.g;
    const double dt = *A.dt;
    const double dt2 = dt * dt;
    double bb = 0.0, rr = 0.0, rz = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double vol = g.dxw[i + g.ngx] * g.dyw[j + g.ngy] * g.dzw[k + g.ngz];
        const double sc = vol / (A.rhoc2[id] * dt2);
        double dd = sc;
#pragma unroll
        for (int dir = 0; dir < DIM; ++dir) {
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                double w;
                pcg_face<DIM>(A, dir, side, i, j, k, w);
                dd += w;
            }
        }
        A.s[id] = sc;
        A.d[id] = dd;
        const double rhs = A.rhs[id];
        bb += rhs * rhs;
        const double r = sc * rhs - pcg_row<DIM>(A, A.x, sc, i, j, k, id);
        const double ro = r / sc;
        rr += ro * ro;
        const double z = r / dd;
        rz += r * z;
        A.r[id] = r;
        A.z[id] = z;
        A.p[id] = z;
    }
    pcg_partials(A, sh, 3, bb, rr, rz);
}

__global__ void __launch_bounds__(256) pcg_setup_fold_kernel(PcgArgs A) {
    __shared__ double sh[256 / kWave];
    const double bb = pcg_fold(A, sh, 0);
    const double rr = pcg_fold(A, sh, 1);
    const double rz = pcg_fold(A, sh, 2);
    if (threadIdx.x == 0) {
        A.sc[0] = rz;
        A.sc[3] = bb;
        A.sc[4] = rr;
        A.st[2] = 0;
        A.st[3] = 0;
        if (bb == 0.0) {
            A.st[0] = 1;
            A.st[1] = 0;
            A.st[3] = 1;
        } else if (rr <= A.tol2 * bb) {
            A.st[0] = 1;
            A.st[1] = 0;
        } else {
            A.st[0] = 0;
        }
    }
}

// rhs = 0: the answer is p = 0 (interior).
__global__ void __launch_bounds__(256) pcg_zero_kernel(PcgArgs A) {
    if (A.st[3] == 0) return;
    const Grid& g = A.g;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) A.x[cell_index(g, i, j, k)] = 0.0;
}

// q = S p, partial p.q
template <int DIM>
__global__ void __launch_bounds__(256) pcg_matvec_kernel(PcgArgs A) {
    if (A.st[0]) return;
    __shared__ double sh[256 / kWave];
    const Grid& g = A.g;
    double pq = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double qv = pcg_row<DIM>(A, A.p, A.s[id], i, j, k, id);
        A.q[id] = qv;
        pq += A.p[id] * qv;
    }
    pcg_partials(A, sh, 1, pq, 0.0, 0.0);
}

__global__ void __launch_bounds__(256) pcg_alpha_kernel(PcgArgs A) {
    if (A.st[0]) return;
    __shared__ double sh[256 / kWave];
    const double pq = pcg_fold(A, sh, 0);
    if (threadIdx.x == 0) {
        if (pq > 0.0) {
            A.sc[1] = A.sc[0] / pq;
        } else {
            // breakdown (rounding has exhausted the search directions): stop,
            // reported as not converged
            A.st[0] = 1;
            A.st[1] = A.maxIter;
        }
    }
}

// x += alpha p, r -= alpha q, z = r / d; partials rr = sum (r / s)^2, rz = sum r z
__global__ void __launch_bounds__(256) pcg_update_kernel(PcgArgs A) {
    if (A.st[0]) return;
    __shared__ double sh[256 / kWave];
    const Grid& g = A.g;
    const double alpha = A.sc[1];
    double rr = 0.0, rz = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        A.x[id] += alpha * A.p[id];
        const double r = A.r[id] - alpha * A.q[id];
        A.r[id] = r;
        const double ro = r / A.s[id];
        rr += ro * ro;
        const double z = r / A.d[id];
        A.z[id] = z;
        rz += r * z;
    }
    pcg_partials(A, sh, 2, rr, rz, 0.0);
}

__global__ void __launch_bounds__(256) pcg_beta_kernel(PcgArgs A) {
    if (A.st[0]) return;
    __shared__ double sh[256 / kWave];
    const double rr = pcg_fold(A, sh, 0);
    const double rzn = pcg_fold(A, sh, 1);
    if (threadIdx.x == 0) {
        const int it = A.st[2];
        A.sc[4] = rr;
        if (rr <= A.tol2 * A.sc[3]) {
            A.st[0] = 1;
            A.st[1] = it + 1;
        } else if (it + 1 >= A.maxIter) {
            A.st[0] = 1;
            A.st[1] = A.maxIter;
        } else {
            A.sc[2] = rzn / A.sc[0];
            A.sc[0] = rzn;
            A.st[2] = it + 1;
        }
    }
}

// p = z + beta p
__global__ void __launch_bounds__(256) pcg_dir_kernel(PcgArgs A) {
    if (A.st[0]) return;
    const Grid& g = A.g;
    const double beta = A.sc[2];
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        A.p[id] = A.z[id] + beta * A.p[id];
    }
}

}  // namespace sifv
