// sifv_device.hpp - device building blocks of the fused finite-volume kernels
// (gfx950, wave64).  Everything here is __device__ inline and templated on
// the arithmetic type T of the bulk reconstruction/flux work (double for the
// plain method, float for the mixed-precision form).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace sifv {

constexpr int kMaxPhases = 8;
constexpr int kWave = 64;

// Ghost-cell rule of one domain face after resolving periodicity.
enum BcKind : int { BK_WRAP = 0, BK_MIRROR = 1, BK_SLIP = 2, BK_NOSLIP = 3, BK_OUTFLOW = 4 };

// Geometry and boundary description shared by all kernels.
struct Grid {
    int dim;
    int nx, ny, nz;
    int ngx, ngy, ngz;
    int nxT, nyT, nzT;
    int bc[6];            // BcKind per face (PRIM/CONS fills)
    int sbc[6];           // BcKind per face for scalar fills (wrap or outflow)
    const double* dxw;    // dx(i) at dxw[i + ngx], i in [-ngx, nx+ngx)
    const double* dyw;
    const double* dzw;
};

__device__ __forceinline__ size_t cell_index(const Grid& g, int i, int j, int k) {
    return (size_t)(i + g.ngx) + (size_t)g.nxT * ((size_t)(j + g.ngy) + (size_t)g.nyT * (size_t)(k + g.ngz));
}

// Source index along one axis for a (possibly ghost) index c, following
// RectilinearMesh::fillGhostX/Y/Z (src/RectilinearMesh.cpp:137-418) and the
// periodic self exchange (src/HaloExchange.cpp:414-465).  flipN/flipT report
// the normal / tangential velocity sign flips of wall rules.
// Grid-stride walk over the interior cells of a block's share: work items are
// (row, chunk of up to kIterChunk columns) so that consecutive threads take
// consecutive x, and the 32-bit index split happens once per item, not per cell.
constexpr int kIterChunk = 2048;
struct InteriorIter {
    int nx, ny, C, items, item, i, iend, j, k;
    __device__ __forceinline__ explicit InteriorIter(int nx_, int ny_, int nz_)
        : nx(nx_), ny(ny_), C((nx_ + kIterChunk - 1) / kIterChunk), items(ny_ * nz_ * ((nx_ + kIterChunk - 1) / kIterChunk)),
          item((int)blockIdx.x - (int)gridDim.x), i(0), iend(0), j(0), k(0) {}
    __device__ __forceinline__ bool next(int& oi, int& oj, int& ok) {
        i += blockDim.x;
        while (i >= iend) {
            item += gridDim.x;
            if (item >= items) return false;
            const int row = item / C, xc = item - row * C;
            j = row % ny;
            k = row / ny;
            i = xc * kIterChunk + (int)threadIdx.x;
            iend = (xc + 1) * kIterChunk < nx ? (xc + 1) * kIterChunk : nx;
        }
        oi = i;
        oj = j;
        ok = k;
        return true;
    }
};

__device__ __forceinline__ int map_axis(int c, int n, int lo, int hi, bool& flipN, bool& flipT) {
    if (c < 0) {
        switch (lo) {
            case BK_WRAP: return c + n;
            case BK_OUTFLOW: return 0;
            case BK_SLIP: flipN = !flipN; return -c - 1;
            case BK_NOSLIP: flipN = !flipN; flipT = !flipT; return -c - 1;
            default: return -c - 1;
        }
    }
    if (c >= n) {
        switch (hi) {
            case BK_WRAP: return c - n;
            case BK_OUTFLOW: return n - 1;
            case BK_SLIP: flipN = !flipN; return 2 * n - 1 - c;
            case BK_NOSLIP: flipN = !flipN; flipT = !flipT; return 2 * n - 1 - c;
            default: return 2 * n - 1 - c;
        }
    }
    return c;
}

// Scalar ghost rule (fillScalarGhosts, src/RectilinearMesh.cpp:424-624):
// periodic wraps, every other rule is zero-gradient.
__device__ __forceinline__ int map_axis_scalar(int c, int n, int lo, int hi) {
    if (c < 0) return lo == BK_WRAP ? c + n : 0;
    if (c >= n) return hi == BK_WRAP ? c - n : n - 1;
    return c;
}

// SIFV_FAST_DIV = 1: reciprocals, divisions and square roots whose operands
// are clamped into the normal range (densities and pressures floored at
// 1e-14, sound speeds > 0) use the hardware approximations refined by Newton
// steps instead of the IEEE sequences with range scaling and special-case
// fix-ups (within one ulp of the IEEE result), and wave-speed bounds use the
// hardware min/max (equal to the reference's std::min/std::max for every
// non-NaN operand).
#ifndef SIFV_FAST_DIV
#define SIFV_FAST_DIV 1
#endif

__device__ __forceinline__ double dmax_ref(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin_ref(double a, double b) { return (b < a) ? b : a; }
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return (a < b) ? b : a; }
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return (b < a) ? b : a; }
// min/max of operands that are never NaN in a finite state (wave speeds,
// clamps of densities and pressures)
template <typename T>
__device__ __forceinline__ T hmax(T a, T b) {
    if constexpr (!SIFV_FAST_DIV) return tmax(a, b);
    else if constexpr (sizeof(T) == 8) return __builtin_fmax(a, b);
    else return __builtin_fmaxf(a, b);
}
template <typename T>
__device__ __forceinline__ T hmin(T a, T b) {
    if constexpr (!SIFV_FAST_DIV) return tmin(a, b);
    else if constexpr (sizeof(T) == 8) return __builtin_fmin(a, b);
    else return __builtin_fminf(a, b);
}

// 1/x for a normal x in [1e-290, 1e290]: the hardware reciprocal refined by two
// Newton-Raphson steps r <- r + r (1 - x r) (error well below one ulp of r).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// a / b and 1 / b, sqrt(x) for normal operands (see SIFV_FAST_DIV); the
// quotient gets one residual correction q + r (a - b q).
__device__ __forceinline__ double div_fast(double a, double b) {
    if constexpr (SIFV_FAST_DIV) {
        const double r = rcp_nr(b);
        const double q = a * r;
        return fma(fma(-b, q, a), r, q);
    } else {
        return a / b;
    }
}
__device__ __forceinline__ double rcp_fast(double b) {
    if constexpr (SIFV_FAST_DIV) return rcp_nr(b);
    else return 1.0 / b;
}
// sqrt of a positive normal x: hardware reciprocal square root refined by the
// coupled Newton iteration for g = sqrt(x), h = 1 / (2 sqrt(x)), then two
// residual corrections of g.
__device__ __forceinline__ double sqrt_fast(double x) {
    if constexpr (SIFV_FAST_DIV) {
        const double y = __builtin_amdgcn_rsq(x);
        double g = x * y, h = 0.5 * y;
        const double r = fma(-g, h, 0.5);
        g = fma(g, r, g);
        h = fma(h, r, h);
        double d = fma(-g, g, x);
        g = fma(d, h, g);
        d = fma(-g, g, x);
        return fma(d, h, g);
    } else {
        return sqrt(x);
    }
}

// Equation of state constants of the single-phase path.  eoff: the stored
// total energy excludes the constant gamma*pInf/(gamma-1) (FP32 storage of a
// stiffened gas, where the stored value would otherwise be dominated by that
// constant and lose the pressure to rounding); the conversion then reads
// p = (gamma-1)(E_stored - ke), which is p = (gamma-1)(E - ke) - gamma*pInf.
struct Eos {
    double gamma, pInf;
    int stiffened;
    int eoff = 0;
};

// cons -> prim of one cell in FP64: IdealGasEOS::toPrimitive
// (src/IdealGasEOS.cpp:46-64) / StiffenedGasEOS::toPrimitive
// (src/StiffenedGasEOS.cpp:47-65).  rho is passed through unclamped because
// the reconstruction reads SolutionState::rho, which the conversion does not
// modify (src/SolutionState.cpp:187-205).
template <int DIM>
__device__ __forceinline__ void cons_to_prim(const Eos& e, double rho, const double* mom, double E,
                                             double* u, double& p) {
    double rs = dmax_ref(rho, 1e-14);
    const double ri = rcp_fast(rs);
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        u[d] = mom[d] * ri;
        ke += u[d] * u[d];
    }
    ke *= 0.5 * rs;
    double ein = E - ke;
    if (e.stiffened)
        p = e.eoff ? (e.gamma - 1.0) * ein : (e.gamma - 1.0) * ein - e.gamma * e.pInf;
    else
        p = dmax_ref((e.gamma - 1.0) * ein, 1e-14);
}

// ---------------------------------------------------------------------------
// Reconstruction of one cell: returns the right state of its low face
// ("minus") and the left state of its high face ("plus") from the 2R-1 cell
// stencil v[0..2R-2] centred on the cell.  Formulas: Reconstructor::weno3*,
// weno5*, upwind3*, upwind5* (src/Reconstruction.cpp:83-200).  For WENO5 the
// three smoothness indicators are shared by both sides (the right-biased
// stencil is the mirror image of the left-biased one).
// ---------------------------------------------------------------------------
template <int R, typename T>
struct Recon;

template <typename T>
struct Recon<1, T> {
    __device__ __forceinline__ static void run(const T* v, T, int, T& minus, T& plus) {
        minus = v[0];
        plus = v[0];
    }
};

// WENO weights a_k = d_k / (eps + beta_k)^2, normalised.  FP64: the weights
// are cleared of their denominators, w_k = d_k prod_{m != k} Q_m / sum, with
// Q_m = (eps + beta_m)^2, and both faces share one division
// 1/(den_minus * den_plus) (a second form with one division per face takes
// over should that product leave the FP64 range).  FP32: hardware reciprocals.
__device__ __forceinline__ void weno_finish(double nP, double dP, double nM, double dM, double& plus,
                                            double& minus) {
    const double pr = dP * dM;
    if (pr <= 1.0e290) {
        const double inv = 1.0 / pr;
        plus = (nP * dM) * inv;
        minus = (nM * dP) * inv;
    } else {
        plus = nP / dP;
        minus = nM / dM;
    }
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Division / reciprocal / square root in the working precision: IEEE FP64,
// hardware (1-ulp) FP32 instructions.
// FAST = false keeps the IEEE FP64 operations (two-phase states, whose
// water/air rounding floor is set by the pInf cancellation).
template <typename T, bool FAST = true>
__device__ __forceinline__ T t_rcp(T x) {
    if constexpr (sizeof(T) == 4) return frcp(x);
    else if constexpr (FAST) return rcp_fast(x);
    else return T(1.0) / x;
}
template <typename T, bool FAST = true>
__device__ __forceinline__ T t_div(T a, T b) {
    if constexpr (sizeof(T) == 4) return a * frcp(b);
    else if constexpr (FAST) return div_fast(a, b);
    else return a / b;
}
// square root of a positive normal operand (sound speeds)
template <typename T, bool FAST = true>
__device__ __forceinline__ T t_sqrt(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_sqrtf(x);
    else if constexpr (FAST) return sqrt_fast(x);
    else return sqrt(x);
}
// square root that may see 0 (velocity magnitudes)
template <typename T>
__device__ __forceinline__ T t_sqrt0(T x) {
    if constexpr (sizeof(T) == 8) return sqrt(x);
    else return __builtin_amdgcn_sqrtf(x);
}

template <typename T>
struct Recon<2, T> {
    __device__ __forceinline__ static void run(const T* v, T eps, int weno, T& minus, T& plus) {
        const T h = T(0.5), th = T(1.5);
        T pl0 = -h * v[0] + th * v[1];
        T pl1 = h * v[1] + h * v[2];
        T pr0 = h * v[0] + h * v[1];
        T pr1 = th * v[1] - h * v[2];
        const T c13 = T(1.0) / T(3.0), c23 = T(2.0) / T(3.0);
        if (!weno) {
            plus = c13 * pl0 + c23 * pl1;
            minus = c23 * pr0 + c13 * pr1;
            return;
        }
        T b0 = (v[1] - v[0]) * (v[1] - v[0]);
        T b1 = (v[2] - v[1]) * (v[2] - v[1]);
        T Q0 = (eps + b0) * (eps + b0);
        T Q1 = (eps + b1) * (eps + b1);
        if constexpr (sizeof(T) == 8) {
            // plus: a = (1/3)/Q0, (2/3)/Q1; minus: (2/3)/Q0, (1/3)/Q1
            const T nP = c13 * Q1 * pl0 + c23 * Q0 * pl1, dP = c13 * Q1 + c23 * Q0;
            const T nM = c23 * Q1 * pr0 + c13 * Q0 * pr1, dM = c23 * Q1 + c13 * Q0;
            weno_finish(nP, dP, nM, dM, plus, minus);
        } else {
            T q0 = frcp(Q0), q1 = frcp(Q1);
            T aL0 = c13 * q0, aL1 = c23 * q1;
            T aR0 = c23 * q0, aR1 = c13 * q1;
            plus = (aL0 * pl0 + aL1 * pl1) * frcp(aL0 + aL1);
            minus = (aR0 * pr0 + aR1 * pr1) * frcp(aR0 + aR1);
        }
    }
};

// WENO5 of cell c = v[2].  With the first differences a..d of the stencil, the
// quadratic of sub-stencil k (cells k..k+2) has second difference t_k and slope
// term S_k (S_0 = 3b - a, S_1 = b + c, S_2 = 3c - d); its values at the cell's
// high / low face are v[2] + (t_k +- 3 S_k) / 12, and its Jiang-Shu smoothness
// indicator is beta_k = 13/12 t_k^2 + 1/4 S_k^2.  These are the candidate
// polynomials and indicators of Reconstructor::weno5* (src/Reconstruction.cpp:
// 119-200) rewritten around the cell value, so that both faces share the
// differences and each face is v[2] plus a weighted correction.  The high face
// weights sub-stencils 0,1,2 with d = 1/10, 6/10, 3/10; the low face, the mirror
// image, with 3/10, 6/10, 1/10.
template <typename T>
struct Recon<3, T> {
    __device__ __forceinline__ static void run(const T* v, T eps, int weno, T& minus, T& plus) {
        const T a = v[1] - v[0], b = v[2] - v[1], c = v[3] - v[2], d = v[4] - v[3];
        const T t0 = b - a, t1 = c - b, t2 = d - c;
        const T S0 = T(3.0) * b - a, S1 = b + c, S2 = T(3.0) * c - d;
        const T Gp0 = t0 + T(3.0) * S0, Gp1 = t1 + T(3.0) * S1, Gp2 = t2 + T(3.0) * S2;
        const T Gm0 = t0 - T(3.0) * S0, Gm1 = t1 - T(3.0) * S1, Gm2 = t2 - T(3.0) * S2;
        const T d0 = T(1.0) / T(10.0), d1 = T(6.0) / T(10.0), d2 = T(3.0) / T(10.0);
        const T c112 = T(1.0) / T(12.0);
        if (!weno) {
            plus = v[2] + (d0 * Gp0 + d1 * Gp1 + d2 * Gp2) * c112;
            minus = v[2] + (d2 * Gm0 + d1 * Gm1 + d0 * Gm2) * c112;
            return;
        }
        const T c1312 = T(13.0) / T(12.0), c14 = T(1.0) / T(4.0);
        const T b0 = c1312 * t0 * t0 + c14 * S0 * S0;
        const T b1 = c1312 * t1 * t1 + c14 * S1 * S1;
        const T b2 = c1312 * t2 * t2 + c14 * S2 * S2;
        const T e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
        const T Q0 = e0 * e0, Q1 = e1 * e1, Q2 = e2 * e2;
        if constexpr (sizeof(T) == 8) {
            // weights d_k / Q_k cleared of their denominators: high face
            // (d0 Q1Q2, d1 Q0Q2, d2 Q0Q1), low face (d2 Q1Q2, d1 Q0Q2, d0 Q0Q1)
            const T P12 = Q1 * Q2, P02 = Q0 * Q2, P01 = Q0 * Q1;
            const T wP0 = d0 * P12, w1 = d1 * P02, wP2 = d2 * P01;
            const T wM0 = d2 * P12, wM2 = d0 * P01;
            const T nP = wP0 * Gp0 + w1 * Gp1 + wP2 * Gp2, dP = wP0 + w1 + wP2;
            const T nM = wM0 * Gm0 + w1 * Gm1 + wM2 * Gm2, dM = wM0 + w1 + wM2;
            const T pr = dP * dM;
            if (pr <= 1.0e290 && pr >= 1.0e-290) {
                const T inv = rcp_nr(pr) * c112;
                plus = v[2] + (nP * dM) * inv;
                minus = v[2] + (nM * dP) * inv;
            } else {
                // products out of range (eps + beta below ~1e-36 or above
                // ~1e36): the weights d_k / Q_k of the reference, computed
                // from the ratios e_min / e_k in (0, 1] so that neither the
                // squares nor their reciprocals leave the FP64 range
                const T em = tmin(tmin(e0, e1), e2);
                const T r0 = em / e0, r1 = em / e1, r2 = em / e2;
                const T q0 = r0 * r0, q1 = r1 * r1, q2 = r2 * r2;
                const T aP0 = d0 * q0, a1 = d1 * q1, aP2 = d2 * q2;
                const T aM0 = d2 * q0, aM2 = d0 * q2;
                plus = v[2] + (aP0 * Gp0 + a1 * Gp1 + aP2 * Gp2) / ((aP0 + a1 + aP2) * T(12.0));
                minus = v[2] + (aM0 * Gm0 + a1 * Gm1 + aM2 * Gm2) / ((aM0 + a1 + aM2) * T(12.0));
            }
        } else {
            // weno == 2: the indicators are scaled by their minimum first
            // (weights depend only on the ratios of eps + beta_k), so that a
            // pressure jump of 1e9..1e19 (stiffened gas) cannot overflow the
            // FP32 square; the smallest indicator keeps the weight d_k
            T q0, q1, q2;
            if (weno == 2) {
                const T em = tmin(tmin(e0, e1), e2);
                const T t0 = frcp(e0) * em, t1 = frcp(e1) * em, t2 = frcp(e2) * em;
                q0 = t0 * t0;
                q1 = t1 * t1;
                q2 = t2 * t2;
            } else {
                q0 = frcp(Q0);
                q1 = frcp(Q1);
                q2 = frcp(Q2);
            }
            const T aP0 = d0 * q0, a1 = d1 * q1, aP2 = d2 * q2;
            const T aM0 = d2 * q0, aM2 = d0 * q2;
            plus = v[2] + (aP0 * Gp0 + a1 * Gp1 + aP2 * Gp2) * (frcp(aP0 + a1 + aP2) * c112);
            minus = v[2] + (aM0 * Gm0 + a1 * Gm1 + aM2 * Gm2) * (frcp(aM0 + a1 + aM2) * c112);
        }
    }
};

// ---------------------------------------------------------------------------
// Riemann fluxes along axis D of a DIM-dimensional problem.  Face states are
// primitive (rho, u[DIM], p, sigma) with their effective gamma / piInf (the
// EOS constants for a single phase, the mixture values of the reconstructed
// volume fractions otherwise, MixtureEOS::effectiveGammaAndPiInf) and NPH
// reconstructed volume fractions.  Formulas: computeHLLCFlux
// (src/HLLCSolver.cpp:7-161), computeRusanovFlux (src/RusanovSolver.cpp:7-83),
// computeLFFlux (src/LFSolver.cpp:8-75), helpers include/RiemannSolver.hpp:58-78.
// Output F[0]=mass, F[1..DIM]=momentum, F[DIM+1]=energy, F[DIM+2]=pressure
// flux, F[DIM+3]=face velocity, F[DIM+4+ph]=volume-fraction flux.
// ---------------------------------------------------------------------------
template <int DIM, int NPH, typename T>
struct FaceState {
    T rho, u[DIM], p, sig, g, pi;
    T gm1i;  // 1 / (g - 1)
    T a[NPH > 0 ? NPH : 1];
};

template <int NPH>
constexpr int flux_len(int dim) { return dim + 4 + NPH; }

template <int DIM, int NPH, typename T>
__device__ __forceinline__ T sound_direct(const FaceState<DIM, NPH, T>& w) {
    constexpr bool F = NPH == 0;
    return t_sqrt<T, F>(t_div<T, F>(w.g * hmax(w.p + w.pi, T(1e-14)), hmax(w.rho, T(1e-14))));
}

// Both sound speeds of a face with one division.
template <int DIM, int NPH, typename T>
__device__ __forceinline__ void sound_pair(const FaceState<DIM, NPH, T>& L, const FaceState<DIM, NPH, T>& R,
                                           T& cL, T& cR) {
    constexpr bool F = NPH == 0;
    const T rL = hmax(L.rho, T(1e-14)), rR = hmax(R.rho, T(1e-14));
    if constexpr (sizeof(T) == 8) {
        const T inv = t_rcp<T, F>(rL * rR);
        cL = t_sqrt<T, F>(L.g * hmax(L.p + L.pi, T(1e-14)) * (rR * inv));
        cR = t_sqrt<T, F>(R.g * hmax(R.p + R.pi, T(1e-14)) * (rL * inv));
    } else {
        // the product of two FP32 densities can leave the normal range
        cL = t_sqrt(t_div(L.g * hmax(L.p + L.pi, T(1e-14)), rL));
        cR = t_sqrt(t_div(R.g * hmax(R.p + R.pi, T(1e-14)), rR));
    }
}

template <int DIM, int NPH, typename T>
__device__ __forceinline__ T rhoE_state(const FaceState<DIM, NPH, T>& w) {
    T q = w.u[0] * w.u[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) q += w.u[d] * w.u[d];
    T ke = T(0.5) * w.rho * q;
    return (w.p + w.g * w.pi) * w.gm1i + ke;
}

enum { RS_LF = 0, RS_RUSANOV = 1, RS_HLLC = 2 };

// Element-wise choice between two face states.  Selecting values (not a
// reference to one of two stack objects) keeps both states in registers.
template <int DIM, int NPH, typename T>
__device__ __forceinline__ FaceState<DIM, NPH, T> pick(bool first, const FaceState<DIM, NPH, T>& A,
                                                       const FaceState<DIM, NPH, T>& B) {
    FaceState<DIM, NPH, T> W;
    W.rho = first ? A.rho : B.rho;
#pragma unroll
    for (int d = 0; d < DIM; ++d) W.u[d] = first ? A.u[d] : B.u[d];
    W.p = first ? A.p : B.p;
    W.sig = first ? A.sig : B.sig;
    W.g = first ? A.g : B.g;
    W.pi = first ? A.pi : B.pi;
    W.gm1i = first ? A.gm1i : B.gm1i;
#pragma unroll
    for (int ph = 0; ph < (NPH > 0 ? NPH : 1); ++ph) W.a[ph] = first ? A.a[ph] : B.a[ph];
    return W;
}

template <int RS, int D, int DIM, int NPH, typename T>
__device__ __forceinline__ void riemann(const FaceState<DIM, NPH, T>& L, const FaceState<DIM, NPH, T>& R,
                                        bool incP, bool igr, T* F) {
    T uL = L.u[D], uR = R.u[D];
    if (RS == RS_HLLC) {
        T EL = rhoE_state(L), ER = rhoE_state(R);
        T sL, sR, sS;
        if (incP) {
            T cL, cR;
            sound_pair(L, R, cL, cR);
            sL = hmin(uL - cL, uR - cR);
            sR = hmax(uL + cL, uR + cR);
            sS = t_div<T, NPH == 0>(R.p - L.p + L.rho * uL * (sL - uL) - R.rho * uR * (sR - uR),
                                    L.rho * (sL - uL) - R.rho * (sR - uR));
        } else {
            sL = hmin(uL, uR);
            sR = hmax(uL, uR);
            sS = T(0.5) * (uL + uR);
        }
        if (sL >= T(0) || sR <= T(0)) {
            const bool lft = sL >= T(0);
            const FaceState<DIM, NPH, T> W = pick(lft, L, R);
            T un = lft ? uL : uR;
            T E = lft ? EL : ER;
            F[0] = W.rho * un;
#pragma unroll
            for (int i = 0; i < DIM; ++i) F[1 + i] = W.rho * W.u[i] * un;
            F[DIM + 1] = E * un;
            if (incP) {
                F[1 + D] += W.p;
                F[DIM + 1] += W.p * un;
            }
            F[DIM + 2] = W.p * un;
            F[DIM + 3] = un;
#pragma unroll
            for (int ph = 0; ph < NPH; ++ph) F[DIM + 4 + ph] = W.a[ph] * un;
        } else {
            const bool left = sS >= T(0);
            const FaceState<DIM, NPH, T> W = pick(left, L, R);
            T un = left ? uL : uR;
            T s = left ? sL : sR;
            T E = left ? EL : ER;
            const T A = s - un, B = s - sS;
            T rS, e, pOverRhoA = T(0);
            if (incP && sizeof(T) == 8) {
                // s - un and s - sS are nonzero here (sound speed > 0, sS strictly
                // between the outer waves): one division for all three quotients
                const T rA = W.rho * A;
                const T inv = t_rcp<T, NPH == 0>(rA * B);
                rS = rA * (rA * inv);
                e = E * (A * B * inv);
                pOverRhoA = W.p * (B * inv);
            } else {
                rS = t_div<T, NPH == 0>(W.rho * A, B);
                e = t_div<T, NPH == 0>(E, W.rho);
                if (incP) pOverRhoA = t_div<T, NPH == 0>(W.p, W.rho * A);
            }
            F[0] = W.rho * un + s * (rS - W.rho);
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                T ui = W.u[i];
                T ruS = (i == D) ? rS * (ui + (sS - un)) : rS * ui;
                T base = W.rho * ui * un;
                if (incP && i == D) base = base + W.p;
                F[1 + i] = base + s * (ruS - W.rho * ui);
            }
            if (incP) {
                T ES = rS * (e + (sS - un) * (sS + pOverRhoA));
                F[DIM + 1] = (E + W.p) * un + s * (ES - E);
            } else {
                T ES = rS * (e + (sS - un) * sS);
                F[DIM + 1] = E * un + s * (ES - E);
            }
            F[DIM + 2] = W.p * sS;
            F[DIM + 3] = sS;
#pragma unroll
            for (int ph = 0; ph < NPH; ++ph) F[DIM + 4 + ph] = W.a[ph] * sS;
        }
    } else {
        // Rusanov and LF share the central + dissipation form.
        T EL = rhoE_state(L), ER = rhoE_state(R);
        T S;
        if (RS == RS_RUSANOV) {
            T aL = fabs(uL), aR = fabs(uR);
            if (incP) {
                T cL, cR;
                sound_pair(L, R, cL, cR);
                S = hmax(aL + cL, aR + cR);
            } else {
                S = hmax(aL, aR);
            }
            T mL = L.rho * uL, mR = R.rho * uR;
            F[0] = T(0.5) * (mL + mR) - T(0.5) * S * (R.rho - L.rho);
#pragma unroll
            for (int i = 0; i < DIM; ++i) {
                T pL = L.rho * L.u[i] * uL, pR = R.rho * R.u[i] * uR;
                if (incP && i == D) {
                    pL += L.p;
                    pR += R.p;
                }
                F[1 + i] = T(0.5) * (pL + pR) - T(0.5) * S * (R.rho * R.u[i] - L.rho * L.u[i]);
            }
            T eL = EL * uL, eR = ER * uR;
            if (incP) {
                eL += L.p * uL;
                eR += R.p * uR;
            }
            F[DIM + 1] = T(0.5) * (eL + eR) - T(0.5) * S * (ER - EL);
        } else {
            T qL = L.u[0] * L.u[0], qR = R.u[0] * R.u[0];
#pragma unroll
            for (int d = 1; d < DIM; ++d) {
                qL += L.u[d] * L.u[d];
                qR += R.u[d] * R.u[d];
            }
            qL = t_sqrt0(qL);
            qR = t_sqrt0(qR);
            if (incP) {
                T cL, cR;
                sound_pair(L, R, cL, cR);
                S = hmax(qL, qR) + hmax(cL, cR);
            } else {
                S = hmax(qL, qR);
            }
            F[0] = T(0.5) * (L.rho * uL + R.rho * uR) - T(0.5) * S * (R.rho - L.rho);
#pragma unroll
            for (int i = 0; i < DIM; ++i)
                F[1 + i] = T(0.5) * (L.rho * L.u[i] * uL + R.rho * R.u[i] * uR) -
                           T(0.5) * S * (R.rho * R.u[i] - L.rho * L.u[i]);
            F[DIM + 1] = T(0.5) * (EL * uL + ER * uR) - T(0.5) * S * (ER - EL);
            if (incP) {
                F[1 + D] += T(0.5) * (L.p + R.p);
                F[DIM + 1] += T(0.5) * (L.p * uL + R.p * uR);
            }
            if (igr) {
                F[1 + D] += T(0.5) * (L.sig + R.sig);
                F[DIM + 1] += T(0.5) * (L.sig * uL + R.sig * uR);
            }
        }
        F[DIM + 3] = T(0.5) * (uL + uR);
        F[DIM + 2] = T(0.5) * (L.p * uL + R.p * uR) - T(0.5) * S * (R.p - L.p);
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph)
            F[DIM + 4 + ph] = T(0.5) * (L.a[ph] * uL + R.a[ph] * uR) - T(0.5) * S * (R.a[ph] - L.a[ph]);
    }
}

// Wave-wide minimum of a double (wave64, DPP/bpermute through __shfl_xor).
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dmin_ref(v, __shfl_xor(v, o, kWave));
    return v;
}
// Wave-wide sum (butterfly: every lane forms the same pairwise tree).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, kWave);
        v = (v < w) ? w : v;
    }
    return v;
}

}  // namespace sifv
