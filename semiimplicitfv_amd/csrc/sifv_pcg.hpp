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
// weights are computed once per solve (pcg_weights_kernel) into DIM padded
// fields W[d]: W[d] at a cell holds the weight of its low face along d, and
// the slot one past the last cell of a line holds the line's high boundary
// face (0 on a wall, the wrap face of a periodic axis), so a row reads its
// high weight from the next slot.  Each weight is the expression of the row
// that owns the slot (the other row's operands in the same order), so
// products equal those of recomputing the weights per row, up to the
// compiler's choice of fused multiply-adds, without the 2*DIM FP64 divisions
// per cell.  Scalars
// (alpha, beta, the dot products) stay on the device; each product and update
// kernel writes per-block partial sums that a one-block kernel folds in fixed
// order, so a solve is deterministic run to run.  A device flag ends the
// iteration: the host enqueues iterations in chunks and polls the flag.
//
// Per iteration, per interior cell: pcg_matvec (read p with its 2*DIM
// neighbours, W, s; write q: 24 + 8 DIM B), pcg_update (read x p r q d s,
// write x r z: 72 B), pcg_dir (read z p, write p: 24 B) -- 144 B/cell in 3D,
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
    double* W[3];          // face weights (see above)
    double* part;          // 3 x nblk per-block partial sums
    int nblk;
    double* sc;            // [0] rz [1] alpha [2] beta [3] bb [4] rr
    int* st;               // [0] done [1] iterations [2] iteration counter [3] zero x
    double tol2;           // tol^2
    int maxIter;
    double mgOmega = 0.0;  // > 0: multigrid form, the update writes the V-cycle's pre-smooth q = om r / d
    float* q32 = nullptr;  // multigrid form with an FP32 V-cycle: the pre-smooth goes here, rounded
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

// Interior index of the neighbour across the low (side 0) / high (side 1)
// face along dir: the wrapped interior cell on a periodic domain face, the
// cell itself on a wall (whose weight is 0).
__device__ __forceinline__ size_t pcg_nb(const Grid& g, int dir, int side, int i, int j, int k, size_t id,
                                         size_t stride) {
    const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
    const int c = dir == 0 ? i : dir == 1 ? j : k;
    if (side == 0) {
        if (c > 0) return id - stride;
        return g.sbc[2 * dir] == BK_WRAP ? id + (size_t)(n - 1) * stride : id;
    }
    if (c < n - 1) return id + stride;
    return g.sbc[2 * dir + 1] == BK_WRAP ? id - (size_t)(n - 1) * stride : id;
}

// (S v)_i = (V/c)_i v_i + sum_f w_f (v_i - v_nb): neighbour differences first,
// so a smooth v (pressures ~1e5 with small variations) keeps its precision
// where the face terms outweigh V/c by many orders (fine meshes, large dt).
template <int DIM>
__device__ __forceinline__ double pcg_row(const PcgArgs& A, const double* v, double si, int i, int j, int k,
                                          size_t id) {
    const Grid& g = A.g;
    const double vi = v[id];
    double acc = si * vi;
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
        const size_t stride = dir == 0 ? 1 : dir == 1 ? (size_t)g.nxT : (size_t)g.nxT * g.nyT;
        const double wl = A.W[dir][id], wh = A.W[dir][id + stride];
        acc += wl * (vi - v[pcg_nb(g, dir, 0, i, j, k, id, stride)]);
        acc += wh * (vi - v[pcg_nb(g, dir, 1, i, j, k, id, stride)]);
    }
    return acc;
}

// Coordinate along dir of the neighbour pcg_nb returns.
__device__ __forceinline__ int pcg_nbc(const Grid& g, int dir, int side, int c) {
    const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
    if (side == 0) return c > 0 ? c - 1 : (g.sbc[2 * dir] == BK_WRAP ? n - 1 : c);
    return c < n - 1 ? c + 1 : (g.sbc[2 * dir + 1] == BK_WRAP ? 0 : c);
}

// pcg_row with the vector given by val(index, i, j, k) at the cell and its
// neighbours instead of an array (a vector the kernel forms on the fly from
// the arrays it is made of), over weights W and mass term si of type R, in
// R arithmetic; same operand order as pcg_row.  vi: the value at the cell.
template <int DIM, typename R, typename V>
__device__ __forceinline__ R face_row(const Grid& g, const R* const* W, V&& val, R si, int i, int j, int k, size_t id,
                                      R& vi) {
    vi = val(id, i, j, k);
    R acc = si * vi;
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
        const size_t stride = dir == 0 ? 1 : dir == 1 ? (size_t)g.nxT : (size_t)g.nxT * g.nyT;
        const R wl = W[dir][id], wh = W[dir][id + stride];
        const int c = dir == 0 ? i : dir == 1 ? j : k;
        const int cl = pcg_nbc(g, dir, 0, c), ch = pcg_nbc(g, dir, 1, c);
        const size_t nl = pcg_nb(g, dir, 0, i, j, k, id, stride), nh = pcg_nb(g, dir, 1, i, j, k, id, stride);
        const R vl = dir == 0 ? val(nl, cl, j, k) : dir == 1 ? val(nl, i, cl, k) : val(nl, i, j, cl);
        const R vh = dir == 0 ? val(nh, ch, j, k) : dir == 1 ? val(nh, i, ch, k) : val(nh, i, j, ch);
        acc += wl * (vi - vl);
        acc += wh * (vi - vh);
    }
    return acc;
}

template <int DIM, typename V>
__device__ __forceinline__ double pcg_row_f(const PcgArgs& A, V&& val, double si, int i, int j, int k, size_t id,
                                            double& vi) {
    return face_row<DIM, double>(A.g, A.W, val, si, i, j, k, id, vi);
}

// Traversal of the interior by bricks of 64 columns x 4 rows x kBrickPlanes
// planes for the stencil kernels (2D/3D; blocks of 256 threads = 64 x 4): a
// thread marches its column (i, j) up the brick's planes, so the z
// neighbours of a cell were read by the same thread one plane earlier and
// its y neighbours by the sibling waves of the block, both still in L1/L2.
// Blocks resident on one XCD (block ids are dealt round-robin over the 8
// XCDs) take a contiguous range of bricks, y fastest, so that neighbouring
// bricks share that XCD's L2.  The block count is the elementwise kernels'
// (their per-block partial sums keep one layout).
#ifndef SIFV_PCG_BRICKS
#define SIFV_PCG_BRICKS 1
#endif
constexpr int kBrickPlanes = 16;

template <int DIM, typename F>
__device__ __forceinline__ void pcg_bricks(const Grid& g, F&& body) {
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int bx = (g.nx + 63) / 64, by = (g.ny + 3) / 4;
    const int bz = DIM == 3 ? (g.nz + kBrickPlanes - 1) / kBrickPlanes : 1;
    const long nbr = (long)bx * by * bz;
    const int G = (int)gridDim.x;
    long lo = 0, hi = nbr, t0 = blockIdx.x, step = G;
    if (G >= 8) {
        // every XCD holds at least one block: it sweeps its own brick range
        const int x = (int)blockIdx.x % 8;
        lo = (long)x * nbr / 8;
        hi = (long)(x + 1) * nbr / 8;
        t0 = lo + (int)blockIdx.x / 8;
        step = (G - x + 7) / 8;
    }
    for (long t = t0; t < hi; t += step) {
        const int cy = (int)(t % by);
        const long rest = t / by;
        const int cx = (int)(rest % bx), cz = (int)(rest / bx);
        const int i = cx * 64 + tx, j = cy * 4 + ty;
        if (i >= g.nx || j >= g.ny) continue;
        const int k0 = DIM == 3 ? cz * kBrickPlanes : 0;
        const int k1 = DIM == 3 ? min(g.nz, k0 + kBrickPlanes) : 1;
        for (int k = k0; k < k1; ++k) body(i, j, k);
    }
}

// Face weights W (low face of each cell; high face of the last cell of a
// line in the slot after it), mass term s = V/c and diagonal d, from rho and
// the widths with each row's own operand order.
template <int DIM>
__global__ void __launch_bounds__(256) pcg_weights_kernel(PcgArgs A) {
    const Grid& g = A.g;
    const double dt = *A.dt;
    const double dt2 = dt * dt;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double vol = g.dxw[i + g.ngx] * g.dyw[j + g.ngy] * g.dzw[k + g.ngz];
        const double sc = vol / (A.rhoc2[id] * dt2);
        double dd = sc;
#pragma unroll
        for (int dir = 0; dir < DIM; ++dir) {
            const int n = dir == 0 ? g.nx : dir == 1 ? g.ny : g.nz;
            const int c = dir == 0 ? i : dir == 1 ? j : k;
            const size_t stride = dir == 0 ? 1 : dir == 1 ? (size_t)g.nxT : (size_t)g.nxT * g.nyT;
            double wl, wh;
            pcg_face<DIM>(A, dir, 0, i, j, k, wl);
            pcg_face<DIM>(A, dir, 1, i, j, k, wh);
            dd += wl;
            dd += wh;
            A.W[dir][id] = wl;
            if (c == n - 1) A.W[dir][id + stride] = wh;
        }
        A.s[id] = sc;
        A.d[id] = dd;
    }
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
    double bb = 0.0, rr = 0.0, rz = 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double sc = A.s[id];
        const double dd = A.d[id];
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
    auto cell = [&](int i, int j, int k) {
        const size_t id = cell_index(g, i, j, k);
        const double qv = pcg_row<DIM>(A, A.p, A.s[id], i, j, k, id);
        A.q[id] = qv;
        pq += A.p[id] * qv;
    };
    if (DIM >= 2 && SIFV_PCG_BRICKS) {
        pcg_bricks<DIM>(g, cell);
    } else {
        int i, j, k;
        for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) cell(i, j, k);
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
        if (A.mgOmega > 0.0) {
            // the preconditioner's first fine-level pass (mg_pre_fine_kernel);
            // q is free again until the next product
            const double qv = A.mgOmega * r / A.d[id];
            if (A.q32) A.q32[id] = (float)qv;
            else A.q[id] = qv;
        } else {
            const double z = r / A.d[id];
            A.z[id] = z;
            rz += r * z;
        }
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
        } else if (it + 1 >= A.maxIter || !(rzn > 0.0)) {
            // a preconditioned residual with r.z <= 0 (rounding at the end of
            // a solve) stops the solve as not converged
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

// ---- Multigrid V-cycle preconditioner --------------------------------------
// One V-cycle of geometric multigrid on S replaces z = r / d when the solver
// runs in its multigrid form (the role of the PFMG V-cycle that the reference's
// HyprePressureSolver attaches to StructPCG: weighted-Jacobi relaxation, one
// pre- and one post-sweep, src/HyprePressureSolver.cpp:123-138).  Full 2:1
// coarsening of every axis with piecewise-constant transfers; a coarse cell C
// owns the children {2c, 2c+1} per axis that exist on the finer level and
//   s_C   = sum of the children's mass terms V/c,
//   w_d,C = f_d * (sum of the finer low-face weights along d of the children
//           on C's low side), f_d = 1/2 when the axis was coarsened, else 1,
//   d_C   = s_C + sum_d (w_d,C + w_d,C+e_d (wrapped)).
// Coarse levels are compact x-fastest arrays without ghosts (a low-face
// weight at index 0 is the wrap face of a periodic axis, 0 otherwise, so the
// wrapped neighbour index is always safe to read).  V-cycle on a level with
// right-hand side b: x = om b / d; restrict b - A x by child sums; V-cycle on
// the next level; x += t(parent); out = x + om (b - A x) / d; on the coarsest
// level kMgCoarseSweeps weighted-Jacobi sweeps from zero.  om = 2/3, 4/5, 6/7
// for 1D, 2D, 3D.  Each step is linear and symmetric, so the preconditioner
// is a fixed SPD operator and conjugate gradients stay valid.
// Fine-level work per V-cycle and cell: pre (q = om r / d, folded into
// pcg_update_kernel inside the iteration), restrict (read q with neighbours,
// W, s, r: 48 B in 3D), prolong and post in one pass (read q W s r d and the
// parent's correction, write z, with the r.z partial sums: 64 B).
constexpr int kMgMaxLevels = 24;
constexpr int kMgCoarseSweeps = 16;

__host__ __device__ constexpr double mg_omega(int dim) {
    return dim == 1 ? 2.0 / 3.0 : dim == 2 ? 0.8 : 6.0 / 7.0;
}

// Mixed-precision form (R = float): the V-cycle runs on FP32 copies of the
// fine operator (face weights, mass term, diagonal: rounded once per solve),
// an FP32 pre-smoothed vector written by the CG update and FP32 coarse levels,
// in FP32 arithmetic, while the conjugate-gradient recurrences, the product,
// the dot products and the stopping test stay FP64.  The preconditioner is
// then the V-cycle of the rounded operator evaluated in FP32: symmetric and
// positive definite up to the FP32 rounding of each application.  Its
// fine-level passes move 28 (restrict) and 40 (post) B per cell instead of 48
// and 64.
template <typename R>
struct MgLevelT {
    int n[3];
    int cells;
    R *s, *w[3], *d, *b, *x, *t;
};
using MgLevel = MgLevelT<double>;

// FP32 copies of the fine level (padded fields like PcgArgs'): W, s, d and the
// pre-smoothed q (null in the FP64 form).
struct MgFine32 {
    float* W[3];
    float *s, *d, *q;
};

template <typename R>
struct MgArgsT {
    PcgArgs A;            // fine level (padded fields)
    MgLevelT<R> F;        // finer compact level (unused when fine is the padded grid)
    MgLevelT<R> C;        // this / the coarser compact level
    const int* gate;      // skip when *gate != 0 (solve finished); null: always run
    const R* src;         // smoothing source
    R* dst;               // smoothing destination
    int slot;             // partial-sum slot of the r.z sums (mg_post_fine_kernel)
    int init;             // mg_post_fine_kernel: also p = z
    MgFine32 f32;         // R = float: the FP32 fine level
};
using MgArgs = MgArgsT<double>;

template <typename R>
__device__ __forceinline__ bool mg_skip(const MgArgsT<R>& M) { return M.gate && *M.gate; }

template <typename R>
__device__ __forceinline__ int mg_idx(const MgLevelT<R>& L, int i, int j, int k) {
    return i + L.n[0] * (j + L.n[1] * k);
}

template <int DIM, typename R>
__device__ __forceinline__ R mg_row(const MgLevelT<R>& L, const R* v, int i, int j, int k) {
    const int id = mg_idx(L, i, j, k);
    const R vi = v[id];
    R acc = L.s[id] * vi;
#pragma unroll
    for (int dir = 0; dir < DIM; ++dir) {
        const int n = L.n[dir];
        const int c = dir == 0 ? i : dir == 1 ? j : k;
        const int cm = c == 0 ? n - 1 : c - 1, cp = c == n - 1 ? 0 : c + 1;
        const int lo = dir == 0 ? mg_idx(L, cm, j, k) : dir == 1 ? mg_idx(L, i, cm, k) : mg_idx(L, i, j, cm);
        const int hi = dir == 0 ? mg_idx(L, cp, j, k) : dir == 1 ? mg_idx(L, i, cp, k) : mg_idx(L, i, j, cp);
        acc += L.w[dir][id] * (vi - v[lo]);
        acc += L.w[dir][hi] * (vi - v[hi]);
    }
    return acc;
}

template <typename R>
__device__ __forceinline__ void mg_coords(const MgLevelT<R>& L, int c, int& i, int& j, int& k) {
    i = c % L.n[0];
    const int r = c / L.n[0];
    j = r % L.n[1];
    k = r / L.n[1];
}

// The fine FP64 operator rounded to FP32 (every padded slot, so the high
// boundary weights in the slot after each line come along).
__global__ void __launch_bounds__(256) mg_fine32_kernel(PcgArgs A, MgFine32 F, int dim, size_t n) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        F.s[c] = (float)A.s[c];
        F.d[c] = (float)A.d[c];
        for (int d = 0; d < dim; ++d) F.W[d][c] = (float)A.W[d][c];
    }
}

// Coarse operator of level C from the fine grid (F unused) or from level F,
// summed in FP64 and stored in R.
template <int DIM, bool FROM_FINE, typename R>
__global__ void __launch_bounds__(256) mg_build_kernel(MgArgsT<R> M) {
    const MgLevelT<R>& C = M.C;
    const Grid& g = M.A.g;
    const int nf0 = FROM_FINE ? g.nx : M.F.n[0], nf1 = FROM_FINE ? g.ny : M.F.n[1], nf2 = FROM_FINE ? g.nz : M.F.n[2];
    const double f0 = nf0 != C.n[0] ? 0.5 : 1.0, f1 = nf1 != C.n[1] ? 0.5 : 1.0, f2 = nf2 != C.n[2] ? 0.5 : 1.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C.cells; c += gridDim.x * blockDim.x) {
        int ci, cj, ck;
        mg_coords(C, c, ci, cj, ck);
        double ss = 0.0, w0 = 0.0, w1 = 0.0, w2 = 0.0;
        for (int k = 2 * ck; k < 2 * ck + 2 && k < nf2; ++k)
            for (int j = 2 * cj; j < 2 * cj + 2 && j < nf1; ++j)
                for (int i = 2 * ci; i < 2 * ci + 2 && i < nf0; ++i) {
                    if constexpr (FROM_FINE) {
                        const size_t id = cell_index(g, i, j, k);
                        ss += M.A.s[id];
                        if (i == 2 * ci) w0 += M.A.W[0][id];
                        if (DIM >= 2 && j == 2 * cj) w1 += M.A.W[1][id];
                        if (DIM >= 3 && k == 2 * ck) w2 += M.A.W[2][id];
                    } else {
                        const int f = mg_idx(M.F, i, j, k);
                        ss += M.F.s[f];
                        if (i == 2 * ci) w0 += M.F.w[0][f];
                        if (DIM >= 2 && j == 2 * cj) w1 += M.F.w[1][f];
                        if (DIM >= 3 && k == 2 * ck) w2 += M.F.w[2][f];
                    }
                }
        C.s[c] = (R)ss;
        C.w[0][c] = (R)(f0 * w0);
        C.w[1][c] = (R)(DIM >= 2 ? f1 * w1 : 0.0);
        C.w[2][c] = (R)(DIM >= 3 ? f2 * w2 : 0.0);
    }
}

template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_diag_kernel(MgArgsT<R> M) {
    const MgLevelT<R>& C = M.C;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C.cells; c += gridDim.x * blockDim.x) {
        int i, j, k;
        mg_coords(C, c, i, j, k);
        double dd = C.s[c];
#pragma unroll
        for (int dir = 0; dir < DIM; ++dir) {
            const int n = C.n[dir];
            const int cc = dir == 0 ? i : dir == 1 ? j : k;
            const int cp = cc == n - 1 ? 0 : cc + 1;
            const int up = dir == 0 ? mg_idx(C, cp, j, k) : dir == 1 ? mg_idx(C, i, cp, k) : mg_idx(C, i, j, cp);
            dd += C.w[dir][c];
            dd += C.w[dir][up];
        }
        C.d[c] = (R)dd;
    }
}

// fine pre-smooth: q = om r / d (FP32 form: rounded into the FP32 q)
template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_pre_fine_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    const PcgArgs& A = M.A;
    const Grid& g = A.g;
    constexpr double om = mg_omega(DIM);
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double qv = om * A.r[id] / A.d[id];
        if constexpr (sizeof(R) == 4) M.f32.q[id] = (float)qv;
        else A.q[id] = qv;
    }
}

// coarse right-hand side from the fine residual r - S q
template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_restrict_fine_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    const PcgArgs& A = M.A;
    const Grid& g = A.g;
    const MgLevelT<R>& C = M.C;
    const float* __restrict__ q32 = M.f32.q;
    auto q32v = [&](size_t id, int, int, int) __attribute__((always_inline)) { return q32[id]; };
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C.cells; c += gridDim.x * blockDim.x) {
        int ci, cj, ck;
        mg_coords(C, c, ci, cj, ck);
        R acc = R(0);
        for (int k = 2 * ck; k < 2 * ck + 2 && k < g.nz; ++k)
            for (int j = 2 * cj; j < 2 * cj + 2 && j < g.ny; ++j)
                for (int i = 2 * ci; i < 2 * ci + 2 && i < g.nx; ++i) {
                    const size_t id = cell_index(g, i, j, k);
                    if constexpr (sizeof(R) == 4) {
                        float qi;
                        acc += (float)A.r[id] - face_row<DIM, float>(g, M.f32.W, q32v, M.f32.s[id], i, j, k, id, qi);
                    } else {
                        acc += A.r[id] - pcg_row<DIM>(A, A.q, A.s[id], i, j, k, id);
                    }
                }
        C.b[c] = acc;
    }
}

// fine prolongation and post-smooth: with q' = q + t(parent) (the coarse
// correction added to the pre-smoothed q, formed for the cell and its
// neighbours where the product needs it, so q' is never stored),
// z = q' + om (r - S q') / d, with the partial sums of r.z into slot M.slot
// (and p = z when M.init)
template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_post_fine_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    __shared__ double sh[256 / kWave];
    const PcgArgs& A = M.A;
    const Grid& g = A.g;
    const MgLevelT<R>& C = M.C;
    constexpr double om = mg_omega(DIM);
    const R* __restrict__ tc = C.t;
    double rz = 0.0;
    auto cell = [&](int i, int j, int k) {
        const size_t id = cell_index(g, i, j, k);
        const double r = A.r[id];
        double z;
        if constexpr (sizeof(R) == 4) {
            const float* __restrict__ q = M.f32.q;
            auto qp = [&](size_t n, int a, int b, int c) __attribute__((always_inline)) {
                return q[n] + tc[mg_idx(C, a >> 1, b >> 1, c >> 1)];
            };
            float qi;
            const float a = face_row<DIM, float>(g, M.f32.W, qp, M.f32.s[id], i, j, k, id, qi);
            z = (double)(qi + (float)om * ((float)r - a) / M.f32.d[id]);
        } else {
            const double* __restrict__ q = A.q;
            auto qp = [&](size_t n, int a, int b, int c) __attribute__((always_inline)) {
                return q[n] + tc[mg_idx(C, a >> 1, b >> 1, c >> 1)];
            };
            double qi;
            const double a = pcg_row_f<DIM>(A, qp, A.s[id], i, j, k, id, qi);
            z = qi + om * (r - a) / A.d[id];
        }
        A.z[id] = z;
        rz += r * z;
        if (M.init) A.p[id] = z;
    };
    if (DIM >= 2 && SIFV_PCG_BRICKS) {
        pcg_bricks<DIM>(g, cell);
    } else {
        int i, j, k;
        for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) cell(i, j, k);
    }
    const double t = wave_sum(rz);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int u = 0; u < (int)(blockDim.x / kWave); ++u) s += sh[u];
        A.part[(size_t)M.slot * A.nblk + blockIdx.x] = s;
    }
}

// Compact-level passes over the cells c0, c0 + cs, ...: the bodies of the
// per-level kernels below and of mg_tail_kernel.
// pre-smooth: x = om b / d
template <int DIM, typename R>
__device__ __forceinline__ void mg_pre_cells(const MgLevelT<R>& C, int c0, int cs) {
    constexpr R om = (R)mg_omega(DIM);
    for (int c = c0; c < C.cells; c += cs) C.x[c] = om * C.b[c] / C.d[c];
}

// coarser right-hand side from level F's residual b - A x
template <int DIM, typename R>
__device__ __forceinline__ void mg_restrict_cells(const MgLevelT<R>& F, const MgLevelT<R>& C, int c0, int cs) {
    for (int c = c0; c < C.cells; c += cs) {
        int ci, cj, ck;
        mg_coords(C, c, ci, cj, ck);
        R acc = R(0);
        for (int k = 2 * ck; k < 2 * ck + 2 && k < F.n[2]; ++k)
            for (int j = 2 * cj; j < 2 * cj + 2 && j < F.n[1]; ++j)
                for (int i = 2 * ci; i < 2 * ci + 2 && i < F.n[0]; ++i)
                    acc += F.b[mg_idx(F, i, j, k)] - mg_row<DIM, R>(F, F.x, i, j, k);
        C.b[c] = acc;
    }
}

// level F: x += t(parent) from level C
template <typename R>
__device__ __forceinline__ void mg_prolong_cells(const MgLevelT<R>& F, const MgLevelT<R>& C, int f0, int fs) {
    for (int f = f0; f < F.cells; f += fs) {
        int i, j, k;
        mg_coords(F, f, i, j, k);
        F.x[f] += C.t[mg_idx(C, i >> 1, j >> 1, k >> 1)];
    }
}

// dst = src + om (b - A src) / d on level C
template <int DIM, typename R>
__device__ __forceinline__ void mg_smooth_cells(const MgLevelT<R>& C, const R* src, R* dst, int c0, int cs) {
    constexpr R om = (R)mg_omega(DIM);
    for (int c = c0; c < C.cells; c += cs) {
        int i, j, k;
        mg_coords(C, c, i, j, k);
        dst[c] = src[c] + om * (C.b[c] - mg_row<DIM, R>(C, src, i, j, k)) / C.d[c];
    }
}

// coarse pre-smooth: x = om b / d
template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_pre_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    mg_pre_cells<DIM, R>(M.C, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// coarser right-hand side from level F's residual b - A x
template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_restrict_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    mg_restrict_cells<DIM, R>(M.F, M.C, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Coarsest level (at most 4 cells per axis, so at most kMgCoarsestCells):
// x = om b / d, then kMgCoarseSweeps - 1 weighted-Jacobi sweeps alternating
// x -> t, t -> x and ending in t, in one workgroup with the iterates in LDS;
// the same arithmetic as mg_pre_kernel followed by mg_smooth_kernel launches.
constexpr int kMgCoarsestCells = 64;

// Called by every thread of the workgroup (blockDim.x >= C.cells).
template <int DIM, typename R>
__device__ __forceinline__ void mg_coarsest_body(const MgLevelT<R>& C) {
    __shared__ R xs[kMgCoarsestCells], ts[kMgCoarsestCells];
    constexpr R om = (R)mg_omega(DIM);
    const int c = threadIdx.x;
    const bool on = c < C.cells;
    int i = 0, j = 0, k = 0;
    R b = R(0), d = R(1);
    if (on) {
        mg_coords(C, c, i, j, k);
        b = C.b[c];
        d = C.d[c];
        xs[c] = om * b / d;
    }
    __syncthreads();
    for (int sw = 1; sw < kMgCoarseSweeps; ++sw) {
        const R* src = (sw & 1) ? xs : ts;
        R* dst = (sw & 1) ? ts : xs;
        if (on) dst[c] = src[c] + om * (b - mg_row<DIM, R>(C, src, i, j, k)) / d;
        __syncthreads();
    }
    if (on) {
        C.x[c] = xs[c];
        C.t[c] = ts[c];
    }
}

// level F: x += t(parent) from level C
template <typename R>
__global__ void __launch_bounds__(256) mg_prolong_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    mg_prolong_cells<R>(M.F, M.C, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// dst = src + om (b - A src) / d on level C
template <int DIM, typename R>
__global__ void __launch_bounds__(256) mg_smooth_kernel(MgArgsT<R> M) {
    if (mg_skip(M)) return;
    mg_smooth_cells<DIM, R>(M.C, M.src, M.dst, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The small levels of a V-cycle in one workgroup: from lv[0] (its right-hand
// side b already restricted) down to the coarsest lv[nlev - 1] and back up to
// lv[0]'s answer in t, the same passes in the same order as the per-level
// launches (pre, restrict, ..., coarsest, ..., prolong, smooth), with a
// workgroup barrier between passes instead of a kernel boundary.  Levels of at
// most SIFV_MG_TAIL_CELLS cells take this path; they are too small to fill
// the GPU, so each launch of theirs costs its fixed start-up time.
#ifndef SIFV_MG_TAIL_CELLS
#define SIFV_MG_TAIL_CELLS 4096
#endif
constexpr int kMgTailThreads = 1024;
constexpr int kMgTailMaxLevels = 8;

template <typename R>
struct MgTailArgsT {
    const int* gate;  // skip when *gate != 0; null: always run
    int nlev;
    MgLevelT<R> lv[kMgTailMaxLevels];
};

template <int DIM, typename R>
__global__ void __launch_bounds__(kMgTailThreads) mg_tail_kernel(MgTailArgsT<R> T) {
    if (T.gate && *T.gate) return;
    const int n = T.nlev;
    const int c0 = threadIdx.x, cs = blockDim.x;
    for (int l = 0; l + 1 < n; ++l) {
        mg_pre_cells<DIM, R>(T.lv[l], c0, cs);
        __syncthreads();
        mg_restrict_cells<DIM, R>(T.lv[l], T.lv[l + 1], c0, cs);
        __syncthreads();
    }
    mg_coarsest_body<DIM, R>(T.lv[n - 1]);
    __syncthreads();
    for (int l = n - 2; l >= 0; --l) {
        mg_prolong_cells<R>(T.lv[l], T.lv[l + 1], c0, cs);
        __syncthreads();
        mg_smooth_cells<DIM, R>(T.lv[l], T.lv[l].x, T.lv[l].t, c0, cs);
        __syncthreads();
    }
}

}  // namespace sifv
