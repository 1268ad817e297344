// sifv_visc.hpp - viscous stress of the explicit and semi-implicit solvers on the GPU
// (addViscousFluxes, src/ViscousFlux.cpp:25-303; called from
// src/ExplicitSolver.cpp:523-527 and src/SemiImplicitSolver.cpp:509-514).
//
// visc_kernel writes, per interior cell, the viscous part of the momentum and
// energy right-hand sides of the RK stage's state: the cell's 2*DIM faces in
// the reference's accumulation order (x low face -, x high face +, then y,
// then z), each face's stress row and work formed exactly as the reference
// forms them: normal derivatives from the two cells, transverse derivatives
// the mean of the two cells' central differences (transverse spacing of the
// high-side cell), tau = mu (grad u + grad u^T) - 2/3 mu div(u) I, work =
// tau . u_face, weighted by faceArea / cellVolume of the cell.  Velocities
// and volume fractions come from the stage's conservative buffer through the
// ghost mapping (load_cell), as the stage kernels read them.  The stage
// update adds the result after the flux terms and the body force.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

template <typename T>
struct ViscArgs {
    StageParams<T> P;       // grid, EOS, mixture, source buffer P.src
    double mu;              // scalar viscosity
    int nPhaseMu;           // > 0: face viscosity from alpha-weighted phase values
    double phaseMu[kMaxPhases];
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

// Stress row tau[.] and work of the face between L = (i,j,k) - e_dir and
// R = (i,j,k), given as cell offsets of the high-side cell R.
template <int DIM, int NPH, typename T>
__device__ __forceinline__ void visc_face(const ViscArgs<T>& A, int dir, int i, int j, int k, double tau[3],
                                          double& work) {
    const StageParams<T>& P = A.P;
    const Grid& g = P.g;
    const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
    Cell<DIM, NPH> L, R;
    load_cell<DIM, NPH>(P, P.src, i - ei, j - ej, k - ek, L);
    load_cell<DIM, NPH>(P, P.src, i, j, k, R);
    const double hw[3] = {g.dxw[i + g.ngx], DIM >= 2 ? g.dyw[j + g.ngy] : 1.0, DIM >= 3 ? g.dzw[k + g.ngz] : 1.0};
    const double hl = dir == 0 ? g.dxw[i - 1 + g.ngx] : dir == 1 ? g.dyw[j - 1 + g.ngy] : g.dzw[k - 1 + g.ngz];
    const double hn = 0.5 * (hl + hw[dir]);
    double grad[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int a = 0; a < DIM; ++a) grad[a][dir] = (R.u[a] - L.u[a]) / hn;
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
    }
    double divU = grad[0][0];
    if (DIM >= 2) divU += grad[1][1];
    if (DIM >= 3) divU += grad[2][2];
    const double mu = A.nPhaseMu > 0 ? 0.5 * (visc_cell_mu<DIM, NPH>(A, L) + visc_cell_mu<DIM, NPH>(A, R)) : A.mu;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        tau[a] = a == dir ? mu * (2.0 * grad[dir][dir] - (2.0 / 3.0) * divU) : mu * (grad[a][dir] + grad[dir][a]);
    double uf[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) uf[a] = 0.5 * (L.u[a] + R.u[a]);
    work = tau[0] * uf[0] + tau[1] * uf[1] + tau[2] * uf[2];
}

template <int DIM, int NPH, typename T>
__global__ void __launch_bounds__(256) visc_kernel(ViscArgs<T> A) {
    const StageParams<T>& P = A.P;
    if (P.active && *P.active == 0) return;
    const Grid& g = P.g;
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
            double tau[3], work;
            // low face: this cell is its high-side cell R
            visc_face<DIM, NPH>(A, dir, i, j, k, tau, work);
#pragma unroll
            for (int a = 0; a < DIM; ++a) acc[a] -= coeff * tau[a];
            acc[DIM] -= coeff * work;
            // high face: this cell is its low-side cell L
            const int ei = dir == 0, ej = dir == 1, ek = dir == 2;
            visc_face<DIM, NPH>(A, dir, i + ei, j + ej, k + ek, tau, work);
            const double area2 = dir == 0 ? h[1] * h[2] : dir == 1 ? h[0] * h[2] : h[0] * h[1];
            const double coeff2 = area2 / vol;
#pragma unroll
            for (int a = 0; a < DIM; ++a) acc[a] += coeff2 * tau[a];
            acc[DIM] += coeff2 * work;
        }
#pragma unroll
        for (int q = 0; q <= DIM; ++q) A.out[(size_t)q * P.total + id] = acc[q];
    }
}

}  // namespace sifv
