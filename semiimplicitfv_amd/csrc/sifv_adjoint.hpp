// sifv_adjoint.hpp - discrete adjoint (vector-Jacobian product) of the
// explicit right-hand side R(U) on doubly periodic 2D meshes: WENO5
// reconstruction of (rho, u, v, p) (Reconstructor::weno5Left/Right,
// src/Reconstruction.cpp:119-168), Rusanov face flux (computeRusanovFlux,
// src/RusanovSolver.cpp:7-83), flux divergence with faceArea/cellVolume
// coefficients (ExplicitSolver::computeRHS, src/ExplicitSolver.cpp:270-351),
// ideal-gas cons->prim (src/SolutionState.cpp:187-205, src/IdealGasEOS.cpp:46-64).
//
// vjp(U, mu) = (dR/dU)^T mu and d(mu . R)/d(gamma) are computed by
//   adj_prim_kernel   W = prim(U) per cell;
//   adj_face_kernel   per face: both reconstructed states (forward), the face
//                     adjoint Fbar = c_R mu_R - c_L mu_L, the Rusanov VJP ->
//                     adjoints of the two face states (one plus state and one
//                     minus state per cell and direction, so no write conflicts);
//   adj_recon_kernel  per cell: WENO5 VJP of its plus and minus states -> the
//                     contributions to its 5 stencil cells;
//   adj_gather_kernel per cell: sums the contributions of the 2 x 5 stencils
//                     it belongs to, applies the cons->prim VJP and writes
//                     out = sum_m a_m X_m + s * vjp (the SSP-RK combination).
// Max/abs branches take the forward branch.  gamma-adjoint partials are
// summed per block and reduced in a fixed order.
#pragma once

#include "sifv_stage.hpp"

namespace sifv {

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct AdjGrid {
    Grid g;
    size_t total;
};

// ---- WENO5 (left-biased value at the cell's high face) and its VJP ----------
__device__ __forceinline__ double weno5_plus_val(const double* v, double eps, double* om, double* pk, double* a,
                                                 double* e, double& S) {
    const double t0 = v[0] - 2.0 * v[1] + v[2], s0 = v[0] - 4.0 * v[1] + 3.0 * v[2];
    const double t1 = v[1] - 2.0 * v[2] + v[3], s1 = v[1] - v[3];
    const double t2 = v[2] - 2.0 * v[3] + v[4], s2 = 3.0 * v[2] - 4.0 * v[3] + v[4];
    const double b[3] = {(13.0 / 12.0) * t0 * t0 + 0.25 * s0 * s0, (13.0 / 12.0) * t1 * t1 + 0.25 * s1 * s1,
                         (13.0 / 12.0) * t2 * t2 + 0.25 * s2 * s2};
    const double d[3] = {0.1, 0.6, 0.3};
    pk[0] = (1.0 / 3.0) * v[0] - (7.0 / 6.0) * v[1] + (11.0 / 6.0) * v[2];
    pk[1] = -(1.0 / 6.0) * v[1] + (5.0 / 6.0) * v[2] + (1.0 / 3.0) * v[3];
    pk[2] = (1.0 / 3.0) * v[2] + (5.0 / 6.0) * v[3] - (1.0 / 6.0) * v[4];
    S = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e[k] = eps + b[k];
        a[k] = d[k] / (e[k] * e[k]);
        S += a[k];
    }
    double val = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        om[k] = a[k] / S;
        val += om[k] * pk[k];
    }
    return val;
}

// vb[0..4] += gbar * d(weno5_plus(v))/dv
__device__ __forceinline__ void weno5_plus_vjp(const double* v, double eps, double gbar, double* vb) {
    double om[3], pk[3], a[3], e[3], S;
    const double val = weno5_plus_val(v, eps, om, pk, a, e, S);
    const double t0 = v[0] - 2.0 * v[1] + v[2], s0 = v[0] - 4.0 * v[1] + 3.0 * v[2];
    const double t1 = v[1] - 2.0 * v[2] + v[3], s1 = v[1] - v[3];
    const double t2 = v[2] - 2.0 * v[3] + v[4], s2 = 3.0 * v[2] - 4.0 * v[3] + v[4];
    double pb[3], bb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pb[k] = gbar * om[k];
        const double ab = gbar * (pk[k] - val) / S;
        bb[k] = -2.0 * ab * a[k] / e[k];
    }
    // candidates
    vb[0] += pb[0] * (1.0 / 3.0);
    vb[1] += pb[0] * (-7.0 / 6.0) + pb[1] * (-1.0 / 6.0);
    vb[2] += pb[0] * (11.0 / 6.0) + pb[1] * (5.0 / 6.0) + pb[2] * (1.0 / 3.0);
    vb[3] += pb[1] * (1.0 / 3.0) + pb[2] * (5.0 / 6.0);
    vb[4] += pb[2] * (-1.0 / 6.0);
    // smoothness indicators: beta = 13/12 t^2 + 1/4 s^2
    {
        const double gt = bb[0] * (13.0 / 6.0) * t0, gs = bb[0] * 0.5 * s0;
        vb[0] += gt + gs;
        vb[1] += -2.0 * gt - 4.0 * gs;
        vb[2] += gt + 3.0 * gs;
    }
    {
        const double gt = bb[1] * (13.0 / 6.0) * t1, gs = bb[1] * 0.5 * s1;
        vb[1] += gt + gs;
        vb[2] += -2.0 * gt;
        vb[3] += gt - gs;
    }
    {
        const double gt = bb[2] * (13.0 / 6.0) * t2, gs = bb[2] * 0.5 * s2;
        vb[2] += gt + 3.0 * gs;
        vb[3] += -2.0 * gt - 4.0 * gs;
        vb[4] += gt + gs;
    }
}

__device__ __forceinline__ double weno5_plus_only(const double* v, double eps) {
    double om[3], pk[3], a[3], e[3], S;
    return weno5_plus_val(v, eps, om, pk, a, e, S);
}

// ---- kernels ------------------------------------------------------------------
struct AdjArgs {
    Grid g;
    size_t total;
    double gamma;
    double eps;
    const double* U;     // state at the linearisation point (4 fields)
    const double* mu;    // adjoint to transpose through R (4 fields)
    double* W;           // primitives (4 fields)
    double* pbP;         // [dir][var] adjoint of each cell's plus state (8 fields)
    double* pbM;         // [dir][var] adjoint of each cell's minus state (8 fields)
    double* contrib;     // [dir][var][k] stencil contributions (40 fields)
    double* gpart;       // per-block gamma-adjoint partials
    int gpartOff;        // first partial slot of this launch
    double s;            // scale of the VJP in the output combination
    double* out;         // output (4 fields)
    const double* X[3];  // addends of the output combination (may alias out)
    double a[3];
    int nX;
};

__device__ __forceinline__ void adj_ijk(const Grid& g, size_t t, int& i, int& j) {
    i = (int)(t % g.nx);
    j = (int)(t / g.nx);
}

__device__ __forceinline__ double block_sum256(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < (int)(blockDim.x / kWave); ++q) s += sh[q];
    return s;
}

__global__ void __launch_bounds__(256) adj_prim_kernel(AdjArgs A) {
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny;
    Eos e;
    e.gamma = A.gamma;
    e.pInf = 0.0;
    e.stiffened = 0;
    int i, j, k_;
    for (InteriorIter it_(g.nx, g.ny, 1); it_.next(i, j, k_);) {
        const size_t id = cell_index(g, i, j, 0);
        double mom[2] = {A.U[A.total + id], A.U[2 * A.total + id]}, u[2], p;
        cons_to_prim<2>(e, A.U[id], mom, A.U[3 * A.total + id], u, p);
        A.W[id] = A.U[id];
        A.W[A.total + id] = u[0];
        A.W[2 * A.total + id] = u[1];
        A.W[3 * A.total + id] = p;
    }
}

// Rusanov VJP for normal axis n (2D ideal gas).
__device__ __forceinline__ void rusanov_vjp(const double* WL, const double* WR, int n, double g, const double* Fb,
                                            double* bL, double* bR, double& gb) {
    const double h = 0.5;
    const double rL = WL[0], rR = WR[0];
    const double uL[2] = {WL[1], WL[2]}, uR[2] = {WR[1], WR[2]};
    const double pL = WL[3], pR = WR[3];
    const double gm1i = 1.0 / (g - 1.0);
    const double qL = uL[0] * uL[0] + uL[1] * uL[1], qR = uR[0] * uR[0] + uR[1] * uR[1];
    const double EL = pL * gm1i + 0.5 * rL * qL, ER = pR * gm1i + 0.5 * rR * qR;
    const double PL = dmax_ref(pL, 1e-14), PR = dmax_ref(pR, 1e-14);
    const double RL = dmax_ref(rL, 1e-14), RR = dmax_ref(rR, 1e-14);
    const double cL = sqrt(g * PL / RL), cR = sqrt(g * PR / RR);
    const double UnL = uL[n], UnR = uR[n];
    const double sL = fabs(UnL) + cL, sR = fabs(UnR) + cR;
    const bool rightWins = sL < sR;
    const double S = rightWins ? sR : sL;
    double rbL = 0, rbR = 0, pbL = 0, pbR = 0, EbL = 0, EbR = 0, Sb = 0;
    double ubL[2] = {0, 0}, ubR[2] = {0, 0};
    // mass
    rbL += Fb[0] * (h * UnL + h * S);
    ubL[n] += Fb[0] * h * rL;
    rbR += Fb[0] * (h * UnR - h * S);
    ubR[n] += Fb[0] * h * rR;
    Sb += -h * Fb[0] * (rR - rL);
    // momentum
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double F = Fb[1 + i];
        rbL += F * (h * uL[i] * UnL + h * S * uL[i]);
        ubL[i] += F * (h * rL * UnL + h * S * rL);
        ubL[n] += F * h * rL * uL[i];
        rbR += F * (h * uR[i] * UnR - h * S * uR[i]);
        ubR[i] += F * (h * rR * UnR - h * S * rR);
        ubR[n] += F * h * rR * uR[i];
        if (i == n) {
            pbL += F * h;
            pbR += F * h;
        }
        Sb += -h * F * (rR * uR[i] - rL * uL[i]);
    }
    // energy
    {
        const double F = Fb[3];
        EbL += F * (h * UnL + h * S);
        pbL += F * h * UnL;
        ubL[n] += F * h * (EL + pL);
        EbR += F * (h * UnR - h * S);
        pbR += F * h * UnR;
        ubR[n] += F * h * (ER + pR);
        Sb += -h * F * (ER - EL);
    }
    // E = p/(g-1) + 0.5 rho |u|^2
    pbL += EbL * gm1i;
    rbL += EbL * 0.5 * qL;
    ubL[0] += EbL * rL * uL[0];
    ubL[1] += EbL * rL * uL[1];
    pbR += EbR * gm1i;
    rbR += EbR * 0.5 * qR;
    ubR[0] += EbR * rR * uR[0];
    ubR[1] += EbR * rR * uR[1];
    gb += -(EbL * pL + EbR * pR) * gm1i * gm1i;
    // S = max(|un| + c) on the forward branch; c = sqrt(g P / R)
    double cb;
    if (rightWins) {
        ubR[n] += Sb * (UnR < 0.0 ? -1.0 : 1.0);
        cb = Sb;
        if (!(pR < 1e-14)) pbR += cb * g / (2.0 * cR * RR);
        if (!(rR < 1e-14)) rbR += -cb * g * PR / (2.0 * cR * RR * RR);
        gb += cb * PR / (2.0 * cR * RR);
    } else {
        ubL[n] += Sb * (UnL < 0.0 ? -1.0 : 1.0);
        cb = Sb;
        if (!(pL < 1e-14)) pbL += cb * g / (2.0 * cL * RL);
        if (!(rL < 1e-14)) rbL += -cb * g * PL / (2.0 * cL * RL * RL);
        gb += cb * PL / (2.0 * cL * RL);
    }
    bL[0] = rbL; bL[1] = ubL[0]; bL[2] = ubL[1]; bL[3] = pbL;
    bR[0] = rbR; bR[1] = ubR[0]; bR[2] = ubR[1]; bR[3] = pbR;
}

template <int DIR>
__global__ void __launch_bounds__(256) adj_face_kernel(AdjArgs A) {
    __shared__ double sh[4];
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny;
    const size_t T = A.total;
    double gsum = 0.0;
    int i, j, k_;
    for (InteriorIter it_(g.nx, g.ny, 1); it_.next(i, j, k_);) {
        // face at the low side of cell R = (i, j) in direction DIR; L = R - e_DIR
        const int iL = DIR == 0 ? wrapi(i - 1, g.nx) : i;
        const int jL = DIR == 1 ? wrapi(j - 1, g.ny) : j;
        const size_t idR = cell_index(g, i, j, 0), idL = cell_index(g, iL, jL, 0);
        double WL[4], WR[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double sp[5], sm[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const int off = k - 2;
                const size_t a = DIR == 0 ? cell_index(g, wrapi(iL + off, g.nx), jL, 0)
                                          : cell_index(g, iL, wrapi(jL + off, g.ny), 0);
                const size_t b = DIR == 0 ? cell_index(g, wrapi(i + off, g.nx), j, 0)
                                          : cell_index(g, i, wrapi(j + off, g.ny), 0);
                sp[k] = A.W[v * T + a];
                sm[4 - k] = A.W[v * T + b];  // minus state = plus state of the mirrored stencil
            }
            WL[v] = weno5_plus_only(sp, A.eps);
            WR[v] = weno5_plus_only(sm, A.eps);
        }
        // coefficients faceArea / cellVolume of the two cells (2D, dz = 1)
        const double dxL = g.dxw[iL + g.ngx], dyL = g.dyw[jL + g.ngy];
        const double dxR = g.dxw[i + g.ngx], dyR = g.dyw[j + g.ngy];
        const double cL = DIR == 0 ? (dyL * 1.0) / ((dxL * dyL) * 1.0) : (dxL * 1.0) / ((dxL * dyL) * 1.0);
        const double cR = DIR == 0 ? (dyR * 1.0) / ((dxR * dyR) * 1.0) : (dxR * 1.0) / ((dxR * dyR) * 1.0);
        double Fb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) Fb[q] = cR * A.mu[q * T + idR] - cL * A.mu[q * T + idL];
        double bL[4], bR[4], gb = 0.0;
        rusanov_vjp(WL, WR, DIR, A.gamma, Fb, bL, bR, gb);
        gsum += gb;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            A.pbP[(DIR * 4 + v) * T + idL] = bL[v];
            A.pbM[(DIR * 4 + v) * T + idR] = bR[v];
        }
    }
    const double bs = block_sum256(gsum, sh);
    if (threadIdx.x == 0) A.gpart[A.gpartOff + blockIdx.x] = A.s * bs;
}

__global__ void __launch_bounds__(256) adj_recon_kernel(AdjArgs A) {
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny;
    const size_t T = A.total;
    int i, j, k_;
    for (InteriorIter it_(g.nx, g.ny, 1); it_.next(i, j, k_);) {
        const size_t id = cell_index(g, i, j, 0);
#pragma unroll
        for (int dir = 0; dir < 2; ++dir)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double st[5], rs[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const int off = k - 2;
                    const size_t a = dir == 0 ? cell_index(g, wrapi(i + off, g.nx), j, 0)
                                              : cell_index(g, i, wrapi(j + off, g.ny), 0);
                    st[k] = A.W[v * T + a];
                    rs[4 - k] = st[k];
                }
                double vb[5] = {0, 0, 0, 0, 0}, vr[5] = {0, 0, 0, 0, 0};
                weno5_plus_vjp(st, A.eps, A.pbP[(dir * 4 + v) * T + id], vb);
                weno5_plus_vjp(rs, A.eps, A.pbM[(dir * 4 + v) * T + id], vr);
#pragma unroll
                for (int k = 0; k < 5; ++k) A.contrib[((dir * 4 + v) * 5 + k) * T + id] = vb[k] + vr[4 - k];
            }
    }
}

__global__ void __launch_bounds__(256) adj_gather_kernel(AdjArgs A) {
    __shared__ double sh[4];
    const Grid& g = A.g;
    const size_t n = (size_t)g.nx * g.ny;
    const size_t T = A.total;
    const double gam = A.gamma;
    double gsum = 0.0;
    int i, j, k_;
    for (InteriorIter it_(g.nx, g.ny, 1); it_.next(i, j, k_);) {
        const size_t id = cell_index(g, i, j, 0);
        double Wb[4] = {0, 0, 0, 0};
#pragma unroll
        for (int dir = 0; dir < 2; ++dir)
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                // this cell is entry k of the stencil of the cell at offset 2 - k
                const int off = 2 - k;
                const size_t c = dir == 0 ? cell_index(g, wrapi(i + off, g.nx), j, 0)
                                          : cell_index(g, i, wrapi(j + off, g.ny), 0);
#pragma unroll
                for (int v = 0; v < 4; ++v) Wb[v] += A.contrib[((dir * 4 + v) * 5 + k) * T + c];
            }
        // cons -> prim VJP (u = m / max(rho, 1e-14), ke = 0.5 rs |u|^2,
        // p = max((g - 1)(E - ke), 1e-14))
        const double rho = A.U[id], m0 = A.U[T + id], m1 = A.U[2 * T + id], E = A.U[3 * T + id];
        const double rs = dmax_ref(rho, 1e-14);
        const double ri = 1.0 / rs;
        const double u0 = m0 * ri, u1 = m1 * ri;
        const double q = u0 * u0 + u1 * u1;
        const double ke = q * (0.5 * rs);
        const double pr = (gam - 1.0) * (E - ke);
        double Eb = 0.0, keb = 0.0;
        if (!(pr < 1e-14)) {
            Eb = Wb[3] * (gam - 1.0);
            keb = -Wb[3] * (gam - 1.0);
            gsum += Wb[3] * (E - ke);
        }
        double ub0 = Wb[1] + keb * rs * u0;
        double ub1 = Wb[2] + keb * rs * u1;
        double rsb = keb * 0.5 * q;
        const double mb0 = ub0 * ri, mb1 = ub1 * ri;
        rsb += -(ub0 * u0 + ub1 * u1) * ri;
        double rb = Wb[0];
        if (!(rho < 1e-14)) rb += rsb;
        const double vj[4] = {rb, mb0, mb1, Eb};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double o = A.s * vj[f];
            for (int m = 0; m < A.nX; ++m) o += A.a[m] * A.X[m][f * T + id];
            A.out[f * T + id] = o;
        }
    }
    const double bs = block_sum256(gsum, sh);
    if (threadIdx.x == 0) A.gpart[A.gpartOff + blockIdx.x] = A.s * bs;
}

// gbar[0] += sum of partials[0..n): one block, fixed strided order per
// thread and a fixed tree, so the result does not depend on timing.
__global__ void __launch_bounds__(256) adj_gsum_kernel(const double* part, int n, double* gbar) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) s += part[q];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) gbar[0] += s;
}

}  // namespace sifv
