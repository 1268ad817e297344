// sifv_visc.hpp - face stresses of the explicit and semi-implicit solvers on
// the GPU: viscous stress (addViscousFluxes, src/ViscousFlux.cpp:25-303) and
// capillary stress (addSurfaceTensionFluxes, src/SurfaceTension.cpp:7-255),
// called from src/ExplicitSolver.cpp:523-535 and
// src/SemiImplicitSolver.cpp:509-520.
//
// visc_kernel writes, per interior cell, the stress part of the momentum and
// energy right-hand sides of the RK stage's state: the cell's 2*DIM faces in
// the reference's accumulation order (x low face -, x high face +, then y,
// then z), each face's stress row and work formed exactly as the reference
// forms them.  Both stresses use the same face stencil: normal derivatives
// from the two cells, transverse derivatives the mean of the two cells'
// central differences (transverse spacing of the high-side cell), so one set
// of cell loads serves both.
//   viscous:   tau = mu (grad u + grad u^T) - 2/3 mu div(u) I (face row),
//              mu scalar or the mean of the two cells' sum_k alpha_k mu_k;
//   capillary: from g = grad alpha_0, T_n = sigma (|g|^2 - g_n^2) / max(|g|, eps),
//              T_a = -sigma g_n g_a / max(|g|, eps);
//   work = (tau or T) . u_face, weighted by faceArea / cellVolume of the cell.
// The reference adds all viscous faces, then all capillary faces; here each
// face's two rows are added in that order into one running sum per cell
// (a reassociation of the same terms).  Velocities and volume fractions come
// from the stage's conservative buffer through the ghost mapping
// (load_cell), as the stage kernels read them.  The stage update adds the
// result after the flux terms and the body force.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

template <typename T>
struct ViscArgs {
    StageParams<T> P;       // grid, EOS, mixture, source buffer P.src
    int viscous;            // hasViscosity
    double mu;              // scalar viscosity
    int nPhaseMu;           // > 0: face viscosity from alpha-weighted phase values
    double phaseMu[kMaxPhases];
    double sigma;           // > 0: capillary stress (multiphase states only)
    double capEps;          // epsGradAlpha
    double* out;            // DIM + 1 padded fields: momentum components, energy
};

template <int DIM, int NPH, typename T>
__device__ __forceinline__ double visc_cell_mu(const ViscArgs<T>& A, const Cell<DIM, NPH>& c) {
    double m = 0.0;
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph)
        if (ph < A.nPhaseMu) m += c.a[ph] * A.phaseMu[ph];
    return m;
}

// Stress rows and work of the face between L = (i,j,k) - e_dir and
// R = (i,j,k), given as cell offsets of the high-side cell R: tau/workV
// (viscous, when A.viscous) and cap/workC (capillary, when sigma > 0).
template <int DIM, int NPH, typename T>
__device__ __forceinline__ void visc_face(const ViscArgs<T>& A, int dir, int i, int j, int k, double tau[3],
                                          double& workV, double cap[3], double& workC) {
    const StageParams<T>& P = A.P;
    const Grid& g = P.g;
    const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
    constexpr bool kCap = NPH > 0;
    const bool cap_on = kCap && A.sigma > 0.0;
    Cell<DIM, NPH> L, R;
    load_cell<DIM, NPH>(P, P.src, i - ei, j - ej, k - ek, L);
    load_cell<DIM, NPH>(P, P.src, i, j, k, R);
    const double hw[3] = {g.dxw[i + g.ngx], DIM >= 2 ? g.dyw[j + g.ngy] : 1.0, DIM >= 3 ? g.dzw[k + g.ngz] : 1.0};
    const double hl = dir == 0 ? g.dxw[i - 1 + g.ngx] : dir == 1 ? g.dyw[j - 1 + g.ngy] : g.dzw[k - 1 + g.ngz];
    const double hn = 0.5 * (hl + hw[dir]);
    double grad[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double ga[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) grad[a][dir] = (R.u[a] - L.u[a]) / hn;
    if constexpr (kCap) ga[dir] = (R.a[0] - L.a[0]) / hn;
#pragma unroll
    for (int b = 0; b < DIM; ++b) {
        if (b == dir) continue;
        const int bi = b == 0, bj = b == 1, bk = b == 2;
        Cell<DIM, NPH> Lm, Lp, Rm, Rp;
        load_cell<DIM, NPH>(P, P.src, i - ei - bi, j - ej - bj, k - ek - bk, Lm);
        load_cell<DIM, NPH>(P, P.src, i - ei + bi, j - ej + bj, k - ek + bk, Lp);
        load_cell<DIM, NPH>(P, P.src, i - bi, j - bj, k - bk, Rm);
        load_cell<DIM, NPH>(P, P.src, i + bi, j + bj, k + bk, Rp);
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
            if (!(a == b || a == dir)) continue;
            const double dL = (Lp.u[a] - Lm.u[a]) / (2.0 * hw[b]);
            const double dR = (Rp.u[a] - Rm.u[a]) / (2.0 * hw[b]);
            grad[a][b] = 0.5 * (dL + dR);
        }
        if constexpr (kCap) {
            const double dL = (Lp.a[0] - Lm.a[0]) / (2.0 * hw[b]);
            const double dR = (Rp.a[0] - Rm.a[0]) / (2.0 * hw[b]);
            ga[b] = 0.5 * (dL + dR);
        }
    }
    double uf[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) uf[a] = 0.5 * (L.u[a] + R.u[a]);
    if (A.viscous) {
        double divU = grad[0][0];
        if (DIM >= 2) divU += grad[1][1];
        if (DIM >= 3) divU += grad[2][2];
        const double mu =
            A.nPhaseMu > 0 ? 0.5 * (visc_cell_mu<DIM, NPH>(A, L) + visc_cell_mu<DIM, NPH>(A, R)) : A.mu;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            tau[a] = a == dir ? mu * (2.0 * grad[dir][dir] - (2.0 / 3.0) * divU) : mu * (grad[a][dir] + grad[dir][a]);
        workV = tau[0] * uf[0] + tau[1] * uf[1] + tau[2] * uf[2];
    }
    if (cap_on) {
        const double nr = dmax_ref(sqrt(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2]), A.capEps);
        const int b0 = dir == 0 ? 1 : 0, b1 = dir == 2 ? 1 : 2;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            cap[a] = a == dir ? A.sigma * (ga[b0] * ga[b0] + ga[b1] * ga[b1]) / nr : -A.sigma * ga[dir] * ga[a] / nr;
        workC = cap[0] * uf[0] + cap[1] * uf[1] + cap[2] * uf[2];
    }
}

template <int DIM, int NPH, typename T>
__global__ void __launch_bounds__(256) visc_kernel(ViscArgs<T> A) {
    const StageParams<T>& P = A.P;
    if (P.active && *P.active == 0) return;
    const Grid& g = P.g;
    const bool cap_on = NPH > 0 && A.sigma > 0.0;
    int i, j, k;
    for (InteriorIter it_(g.nx, g.ny, g.nz); it_.next(i, j, k);) {
        const size_t id = cell_index(g, i, j, k);
        const double h[3] = {g.dxw[i + g.ngx], DIM >= 2 ? g.dyw[j + g.ngy] : 1.0, DIM >= 3 ? g.dzw[k + g.ngz] : 1.0};
        // faceArea / cellVolume with the reference's products
        const double vol = h[0] * h[1] * h[2];
        double acc[DIM + 1];
#pragma unroll
        for (int q = 0; q <= DIM; ++q) acc[q] = 0.0;
#pragma unroll
        for (int dir = 0; dir < DIM; ++dir) {
            const double area = dir == 0 ? h[1] * h[2] : dir == 1 ? h[0] * h[2] : h[0] * h[1];
            const double coeff = area / vol;
            const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                // side 0: low face, this cell is its high-side cell R (-);
                // side 1: high face, this cell is its low-side cell L (+)
                double tau[3], cap[3], wv = 0.0, wc = 0.0;
                visc_face<DIM, NPH>(A, dir, i + side * ei, j + side * ej, k + side * ek, tau, wv, cap, wc);
                if (A.viscous) {
#pragma unroll
                    for (int a = 0; a < DIM; ++a) acc[a] = side ? acc[a] + coeff * tau[a] : acc[a] - coeff * tau[a];
                    acc[DIM] = side ? acc[DIM] + coeff * wv : acc[DIM] - coeff * wv;
                }
                if (cap_on) {
#pragma unroll
                    for (int a = 0; a < DIM; ++a) acc[a] = side ? acc[a] + coeff * cap[a] : acc[a] - coeff * cap[a];
                    acc[DIM] = side ? acc[DIM] + coeff * wc : acc[DIM] - coeff * wc;
                }
            }
        }
#pragma unroll
        for (int q = 0; q <= DIM; ++q) A.out[(size_t)q * P.total + id] = acc[q];
    }
}

}  // namespace sifv
