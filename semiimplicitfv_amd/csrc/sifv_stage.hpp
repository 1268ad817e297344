// sifv_stage.hpp - fused cons->prim + ghost mapping + reconstruction +
// Riemann flux + flux divergence + RK stage update.
//
// Replaces, per RK stage, the reference's chain
//   convertConservativeToPrimitiveVariables (src/SolutionState.cpp:187-205)
//     / MixtureEOS::convertConservativeToPrimitive (src/MixtureEOS.cpp:99-143)
//   -> applyBoundaryConditions(PRIM)          (src/RectilinearMesh.cpp:498-524)
//   -> Reconstructor::reconstructX/Y/Z        (src/Reconstruction.cpp:204-541)
//   -> computeRHS face loops and sources      (src/ExplicitSolver.cpp:270-496,
//                                              src/SemiImplicitSolver.cpp:238-482)
//   -> stage update                            (src/ExplicitSolver.cpp:127-188,
//                                              src/SemiImplicitSolver.cpp:159-224)
//
// Work decomposition (wave64).  1D/2D (`stage_xy_kernel`, MODE_FULL): one
// wave owns a strip of 62 consecutive columns plus one halo column on each
// side (lane l <-> column X0-1+l) and marches upward through `rows` rows.
// The y-stencil of each lane's column lives in a register window that slides
// one row per iteration; the x-stencil is exchanged through LDS.  Each lane
// reconstructs its own cell once per direction (both face states from one
// stencil), so every face flux is computed once per strip.
// 3D: the same kernel runs on each z-plane (MODE_XY) and writes the x+y part
// of the right-hand side; `stage_z_kernel` then marches 64-column strips in z
// with a z-window, adds the z part in the reference's accumulation order
// (x, then y, then z) and applies the stage update.
// Ghost cells are never materialised: stencil indices outside the domain are
// mapped to their source cells with the boundary rule and velocity sign
// flips applied.
#pragma once

#include <type_traits>

#include "sifv_device.hpp"

#ifndef SIFV_PREFETCH_ROW_D
#define SIFV_PREFETCH_ROW_D true
#endif
#ifndef SIFV_PREFETCH_Z
#define SIFV_PREFETCH_Z false
#endif
#ifndef SIFV_PRELOAD_PART_Z
#define SIFV_PRELOAD_PART_Z false
#endif
#ifndef SIFV_PRELOAD_UPDATE_D
#define SIFV_PRELOAD_UPDATE_D false
#endif
#ifndef SIFV_PREFETCH_ROW_F
#define SIFV_PREFETCH_ROW_F true
#endif
#ifndef SIFV_PRELOAD_UPDATE_F
#define SIFV_PRELOAD_UPDATE_F true
#endif

namespace sifv {

constexpr int kOwn = 62;  // owned columns per wave in the x-marching kernels

enum { MODE_FULL = 0, MODE_XY = 1 };

// Mixture parameters (MultiPhaseParams, include/SimulationConfig.hpp:43-52).
// eoff: the stored total energy excludes sum_k alpha_k eC_k with
// eC_k = gamma_k pInf_k / (gamma_k - 1), the stiffened-gas part of the mixture
// energy (FP32 storage; the conversion then reads p = (E_stored - ke) / sInv).
struct Mix {
    int nph;
    double gamma[kMaxPhases], pInf[kMaxPhases];
    double alphaMin;
    int eoff = 0;
    double eC[kMaxPhases] = {};   // gamma_k pInf_k / (gamma_k - 1)
    double gi[kMaxPhases] = {};   // 1 / (gamma_k - 1)
};

// Per-phase terms alpha/(gamma-1) and alpha gamma pInf/(gamma-1) of the mixture
// relations, as the reference's quotients.  Two-phase water/air states sit at
// a rounding floor set by the pInf cancellation, so their relations keep the
// reference's operation forms (no reciprocal products or Newton quotients).
template <typename T>
__device__ __forceinline__ T mix_inv(const Mix& m, int ph, T a) {
    return a / ((T)m.gamma[ph] - T(1.0));
}
template <typename T>
__device__ __forceinline__ T mix_pi(const Mix& m, int ph, T a) {
    return a * (T)m.gamma[ph] * (T)m.pInf[ph] / ((T)m.gamma[ph] - T(1.0));
}

// Wood sound speed, MixtureEOS::mixtureSoundSpeed (src/MixtureEOS.cpp:48-63).
__device__ __forceinline__ double mix_sound(const Mix& m, double rho, double p, const double* a,
                                            const double* ar, int nph) {
    double s = 0.0;
    for (int ph = 0; ph < nph; ++ph) {
        double rk = dmax_ref(ar[ph], 1e-14) / dmax_ref(a[ph], 1e-14);
        double ck2 = m.gamma[ph] * (p + m.pInf[ph]) / rk;
        s += a[ph] / (rk * dmax_ref(ck2, 1e-14));
    }
    double c2 = 1.0 / (rho * dmax_ref(s, 1e-30));
    return sqrt(dmax_ref(c2, 0.0));
}

template <typename T>
struct StageParams {
    Grid g;
    Eos eos;
    Mix mix;
    size_t total;         // padded cells per field
    const T* src;         // stencil source buffer (field q at q*total)
    const T* u0;          // RK backup buffer, read pointwise when c2 != 0
    T* dst;               // stage output buffer
    T* part;              // 3D: x+y part of the right-hand side (NCV+2 fields)
    const double* sigma;  // entropic pressure (LF with IGR)
    double c1, c2, c3, c4;
    double c4inv;         // 1/c4, used when c4 is a power of two (then x*c4inv == x/c4 exactly)
    int c4pow2;
    int rk1;              // RKOrder == 1 form (U += dt R)
    const double* dt;     // device-resident time step
    const int* active;    // device flag; 0 skips the stage (time loop finished)
    double* dtPartial;    // per-block min of the new state's acoustic cell dt, or null
    double eps;
    int igr;
    int weno;             // 0 linear weights, 1 WENO, 2 WENO with scaled indicators (FP32, stiffened gas)
    int sigPrim;          // sigma ghosts follow the PRIM rule (no IGR sweep ran this stage)
    int unifY;            // every interior row (and plane) width is the same double
    int rows;             // rows (y) or planes (z) marched per wave
    int nseg;             // segments per strip
    int nstrip;           // strips per row / plane
    int kOff = 0;         // 3D x+y pass: first plane of this launch
    int segOff = 0;       // 3D z pass: first z segment of this launch
    double* pres;         // semi-implicit: stage-start pressure (written)
    double* vel[3];       // semi-implicit: star velocity (written)
    double* pAdv;         // semi-implicit: advected pressure (written)
    double* rhoc2;        // semi-implicit: rho c^2 of the updated state (written; null: not)
    int bodyForce = 0;              // hasBodyForce
    const double* accel = nullptr;  // device-resident body-force acceleration of the step
    const double* visc = nullptr;   // viscous momentum/energy right-hand side (DIM+1 fields) or null
};

// Viscous limit of one cell (computeViscousDt, src/RKTimeStepping.cpp:161-188),
// without the cfl factor that the dt finalize applies.
__device__ __forceinline__ double visc_cell_dt(const double* h, int dim, double mu, double rho) {
    double hm = h[0];
    if (dim >= 2) hm = dmin_ref(hm, h[1]);
    if (dim >= 3) hm = dmin_ref(hm, h[2]);
    const double nu = mu / dmax_ref(rho, 1e-14);
    return hm * hm / (2.0 * dim * nu);
}

// Capillary limit of one cell (computeCapillaryDt, src/RKTimeStepping.cpp:201-228),
// without the cfl factor.
__device__ __forceinline__ double cap_cell_dt(const double* h, int dim, double sigma, double rho) {
    double hm = h[0];
    if (dim >= 2) hm = dmin_ref(hm, h[1]);
    if (dim >= 3) hm = dmin_ref(hm, h[2]);
    return sqrt(dmax_ref(rho, 1e-14) * hm * hm * hm / sigma);
}

// Logical block id with all blocks of one XCD contiguous (blocks are
// dispatched round-robin over the 8 XCDs of an MI355X), so neighbouring
// strips/segments that share halo data share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int nx = 8;
    int per = nb / nx, rem = nb % nx;
    int x = b % nx, q = b / nx;
    return (x < rem) ? x * (per + 1) + q : rem * (per + 1) + (x - rem) * per + q;
}

// Primitive state of one cell (FP64) plus its partial densities.
// C: storage type of the primitive values (double, or float in the march
// windows of single-phase FP32 states).
template <int DIM, int NPH, typename C = double>
struct Cell {
    C rho, u[DIM], p;
    C a[NPH > 0 ? NPH : 1];
    C ar[NPH > 0 ? NPH : 1];
};

// Storage of the march windows: FP32 for single-phase explicit states held
// in FP32 (their reconstruction reads FP32 values anyway), FP64 otherwise
// (two-phase states keep FP64 primitives for the alpha-rho upwinding and the
// semi-implicit update).
template <typename T, int NPH, bool SEMI>
using CellT = typename std::conditional<sizeof(T) == 4 && NPH == 0 && !SEMI, float, double>::type;

__device__ __forceinline__ int clampi(int c, int lo, int hi) { return c < lo ? lo : (c > hi ? hi : c); }

// cons -> prim of the conservative values of one cell.  Single phase:
// toPrimitive of the ideal/stiffened gas (rho passed through unclamped, as
// the reconstruction reads SolutionState::rho).  Multiphase:
// MixtureEOS::convertConservativeToPrimitive (rho = max(sum alphaRho, 1e-14)).
template <int DIM, int NPH>
__device__ __forceinline__ void prim_of(const Eos& e, const Mix& m, const double* U, Cell<DIM, NPH>& c) {
    if constexpr (NPH == 0) {
        double u[DIM], mom[DIM], p;
#pragma unroll
        for (int d = 0; d < DIM; ++d) mom[d] = U[1 + d];
        cons_to_prim<DIM>(e, U[0], mom, U[DIM + 1], u, p);
        c.rho = U[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d) c.u[d] = u[d];
        c.p = p;
    } else {
        double r = 0.0;
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) r += U[DIM + 2 + ph];
        double rs = dmax_ref(r, 1e-14);
        c.rho = rs;
#pragma unroll
        for (int d = 0; d < DIM; ++d) c.u[d] = U[1 + d] / rs;
        double ke = 0.5 * rs * c.u[0] * c.u[0];
#pragma unroll
        for (int d = 1; d < DIM; ++d) ke += 0.5 * rs * c.u[d] * c.u[d];
        double sInv = 0.0, sPi = 0.0;
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) {
            double a = U[DIM + 2 + NPH + ph];
            sInv += mix_inv<double>(m, ph, a);
            sPi += mix_pi<double>(m, ph, a);
            c.a[ph] = a;
            c.ar[ph] = U[DIM + 2 + ph];
        }
        c.p = m.eoff ? (U[DIM + 1] - ke) / sInv : ((U[DIM + 1] - ke) - sPi) / sInv;
    }
}

// cons -> prim of one single-phase ideal-gas cell in FP32 (IdealGasEOS::toPrimitive,
// src/IdealGasEOS.cpp:46-64, in single precision: no pInf cancellation).
template <int DIM>
__device__ __forceinline__ void prim_of_f32(const Eos& e, const float* U, Cell<DIM, 0, float>& c) {
    const float rs = __builtin_fmaxf(U[0], 1e-14f);
    const float ri = frcp(rs);
    float ke = 0.0f;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        c.u[d] = U[1 + d] * ri;
        ke += c.u[d] * c.u[d];
    }
    ke *= 0.5f * rs;
    c.rho = U[0];
    c.p = __builtin_fmaxf((float)(e.gamma - 1.0) * (U[DIM + 1] - ke), 1e-14f);
}

// cons -> prim of raw values of type T into a cell of storage C.  FP32 cells of
// an ideal gas convert in FP32; a stiffened gas (pInf cancellation) and FP64
// cells convert in FP64.
template <int DIM, int NPH, typename T, typename C>
__device__ __forceinline__ void convert_cell(const Eos& e, const Mix& m, const T* raw, Cell<DIM, NPH, C>& c) {
    constexpr int NCV = DIM + 2 + 2 * NPH;
    if constexpr (sizeof(C) == 4 && NPH == 0) {
        if (!e.stiffened) {
            float U[NCV];
#pragma unroll
            for (int q = 0; q < NCV; ++q) U[q] = (float)raw[q];
            prim_of_f32<DIM>(e, U, c);
            return;
        }
    }
    double U[NCV];
#pragma unroll
    for (int q = 0; q < NCV; ++q) U[q] = (double)raw[q];
    if constexpr (sizeof(C) == 8) {
        prim_of<DIM, NPH>(e, m, U, c);
    } else {
        Cell<DIM, NPH, double> d;
        prim_of<DIM, NPH>(e, m, U, d);
        c.rho = (C)d.rho;
#pragma unroll
        for (int q = 0; q < DIM; ++q) c.u[q] = (C)d.u[q];
        c.p = (C)d.p;
    }
}

// Stiffened-gas part of the stored energy of one cell (Mix::eoff form).
template <int DIM, int NPH>
__device__ __forceinline__ double energy_offset(const Mix& m, const double* U) {
    double o = 0.0;
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph) o += U[DIM + 2 + NPH + ph] * m.eC[ph];
    return o;
}

// Raw conservative values of one mapped cell plus its velocity sign flips
// (bit d: negate u[d]); split from the conversion so that the load of the
// next row can be issued one march iteration ahead of its use.
template <int DIM, int NPH, typename T>
struct RawCell {
    T U[DIM + 2 + 2 * NPH];
    unsigned flip;
};

template <int DIM, int NPH, typename T>
__device__ __forceinline__ void load_raw(const StageParams<T>& P, const T* buf, int i, int j, int k,
                                         RawCell<DIM, NPH, T>& rc) {
    constexpr int NCV = DIM + 2 + 2 * NPH;
    const Grid& g = P.g;
    bool fn[3] = {false, false, false}, ft[3] = {false, false, false};
    int si = map_axis(clampi(i, -g.ngx, g.nx - 1 + g.ngx), g.nx, g.bc[0], g.bc[1], fn[0], ft[0]);
    int sj = 0, sk = 0;
    if (DIM >= 2) sj = map_axis(clampi(j, -g.ngy, g.ny - 1 + g.ngy), g.ny, g.bc[2], g.bc[3], fn[1], ft[1]);
    if (DIM >= 3) sk = map_axis(clampi(k, -g.ngz, g.nz - 1 + g.ngz), g.nz, g.bc[4], g.bc[5], fn[2], ft[2]);
    size_t id = cell_index(g, si, sj, sk);
#pragma unroll
    for (int q = 0; q < NCV; ++q) rc.U[q] = buf[(size_t)q * P.total + id];
    unsigned fl = 0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        bool f = fn[d];
#pragma unroll
        for (int e = 0; e < DIM; ++e)
            if (e != d) f = f != ft[e];
        fl |= f ? (1u << d) : 0u;
    }
    rc.flip = fl;
}

template <int DIM, int NPH, typename T, typename C>
__device__ __forceinline__ void raw_to_cell(const StageParams<T>& P, const RawCell<DIM, NPH, T>& rc,
                                            Cell<DIM, NPH, C>& c) {
    convert_cell<DIM, NPH, T, C>(P.eos, P.mix, rc.U, c);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        if (rc.flip & (1u << d)) c.u[d] = -c.u[d];
}

// Load one cell (i, j, k) through the ghost mapping (with velocity flips).
template <int DIM, int NPH, typename T, typename C>
__device__ __forceinline__ void load_cell(const StageParams<T>& P, const T* buf, int i, int j, int k,
                                          Cell<DIM, NPH, C>& c) {
    constexpr int NCV = DIM + 2 + 2 * NPH;
    const Grid& g = P.g;
    bool fn[3] = {false, false, false}, ft[3] = {false, false, false};
    int si = map_axis(clampi(i, -g.ngx, g.nx - 1 + g.ngx), g.nx, g.bc[0], g.bc[1], fn[0], ft[0]);
    int sj = 0, sk = 0;
    if (DIM >= 2) sj = map_axis(clampi(j, -g.ngy, g.ny - 1 + g.ngy), g.ny, g.bc[2], g.bc[3], fn[1], ft[1]);
    if (DIM >= 3) sk = map_axis(clampi(k, -g.ngz, g.nz - 1 + g.ngz), g.nz, g.bc[4], g.bc[5], fn[2], ft[2]);
    size_t id = cell_index(g, si, sj, sk);
    T U[NCV];
#pragma unroll
    for (int q = 0; q < NCV; ++q) U[q] = buf[(size_t)q * P.total + id];
    convert_cell<DIM, NPH, T, C>(P.eos, P.mix, U, c);
    // velocity component d flips with the normal flag of axis d and the
    // tangential flags of the other axes (onion-peel composition)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        bool f = fn[d];
#pragma unroll
        for (int e = 0; e < DIM; ++e)
            if (e != d) f = f != ft[e];
        if (f) c.u[d] = -c.u[d];
    }
}

template <int DIM, typename T>
__device__ __forceinline__ double load_sigma(const StageParams<T>& P, int i, int j, int k) {
    const Grid& g = P.g;
    i = clampi(i, -g.ngx, g.nx - 1 + g.ngx);
    if (DIM >= 2) j = clampi(j, -g.ngy, g.ny - 1 + g.ngy);
    if (DIM >= 3) k = clampi(k, -g.ngz, g.nz - 1 + g.ngz);
    int si, sj = 0, sk = 0;
    if (P.sigPrim) {
        // ghosts as the stage-start PRIM fill leaves them (no sign flips for a scalar)
        bool fn = false, ft = false;
        si = map_axis(i, g.nx, g.bc[0], g.bc[1], fn, ft);
        if (DIM >= 2) sj = map_axis(j, g.ny, g.bc[2], g.bc[3], fn, ft);
        if (DIM >= 3) sk = map_axis(k, g.nz, g.bc[4], g.bc[5], fn, ft);
    } else {
        // ghosts refreshed by the IGR solve's scalar fill
        si = map_axis_scalar(i, g.nx, g.sbc[0], g.sbc[1]);
        if (DIM >= 2) sj = map_axis_scalar(j, g.ny, g.sbc[2], g.sbc[3]);
        if (DIM >= 3) sk = map_axis_scalar(k, g.nz, g.sbc[4], g.sbc[5]);
    }
    return P.sigma[cell_index(g, si, sj, sk)];
}

// Reconstructed variables of a cell: rho, u[DIM], p, alpha[NPH].
template <int DIM, int NPH>
constexpr int n_recon() { return DIM + 2 + NPH; }

template <int DIM, int NPH, typename T, typename C>
__device__ __forceinline__ void recon_vars(const Cell<DIM, NPH, C>& c, T* v) {
    v[0] = (T)c.rho;
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[1 + d] = (T)c.u[d];
    v[DIM + 1] = (T)c.p;
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph) v[DIM + 2 + ph] = (T)c.a[ph];
}

// Face state from reconstructed values; effective gamma/piInf from the
// reconstructed volume fractions (MixtureEOS::effectiveGammaAndPiInf,
// src/MixtureEOS.cpp:18-31) or the single-phase EOS constants
// (src/Reconstruction.cpp:309-314).
template <int DIM, int NPH, typename T>
__device__ __forceinline__ void face_state(const StageParams<T>& P, const T* v, T sig,
                                           FaceState<DIM, NPH, T>& w) {
    w.rho = v[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) w.u[d] = v[1 + d];
    w.p = v[DIM + 1];
    w.sig = sig;
    if constexpr (NPH == 0) {
        w.g = (T)P.eos.gamma;
        w.pi = (T)(P.eos.stiffened ? P.eos.pInf : 0.0);
        w.gm1i = (T)(1.0 / (P.eos.gamma - 1.0));
    } else {
        T sInv = T(0), sPi = T(0);
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) {
            w.a[ph] = v[DIM + 2 + ph];
            sInv += mix_inv<T>(P.mix, ph, w.a[ph]);
            sPi += mix_pi<T>(P.mix, ph, w.a[ph]);
        }
        w.g = T(1.0) + T(1.0) / sInv;
        w.pi = (w.g - T(1.0)) / w.g * sPi;
        w.gm1i = sInv;
    }
}

// Face contribution vector: [mass, mom[DIM], energy, alphaRho flux[NPH],
// alpha flux[NPH], pressure flux, face velocity] (conservative order, then
// the advected-pressure flux and the divergence velocity).
template <int DIM, int NPH>
constexpr int n_fc() { return DIM + 2 + 2 * NPH + 2; }

// Slots of the 3D x+y partial right-hand side the z pass's update reads: the
// semi-implicit update rebuilds the total energy in its correction step, and
// a multiphase update rebuilds the density as the sum of the partial
// densities, so neither slot is stored (nor its arithmetic kept alive).
template <int DIM, int NPH, bool SEMI>
__device__ __forceinline__ constexpr bool part_slot_used(int q) {
    return !((SEMI && q == DIM + 1) || (NPH > 0 && q == 0));
}

template <int DIM, int NPH, typename T, typename C>
__device__ __forceinline__ void face_contrib(const T* F, C rhoL, C rhoR, const C* arL, const C* arR, T* fc) {
    fc[0] = F[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) fc[1 + d] = F[1 + d];
    fc[DIM + 1] = F[DIM + 1];
    if constexpr (NPH > 0) {
        // upwind partial-density flux (src/ExplicitSolver.cpp:307-326)
        const bool up = F[0] >= T(0);
        double ru = dmax_ref((double)(up ? rhoL : rhoR), 1e-14);
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph)
            fc[DIM + 2 + ph] = (T)(((double)(up ? arL[ph] : arR[ph]) / ru) * (double)F[0]);
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) fc[DIM + 2 + NPH + ph] = F[DIM + 4 + ph];
    }
    fc[DIM + 2 + 2 * NPH] = F[DIM + 2];
    fc[DIM + 3 + 2 * NPH] = F[DIM + 3];
}

// acc += c*lo - c*hi over conservative + pressure slots; the divergence slot
// takes -c*lo + c*hi (src/ExplicitSolver.cpp:312-351: divU[L] += c v, divU[R] -= c v).
template <int DIM, int NPH, typename T>
__device__ __forceinline__ void accumulate(T* acc, const T* lo, const T* hi, T c, bool first) {
    constexpr int NF = n_fc<DIM, NPH>();
#pragma unroll
    for (int q = 0; q < NF - 1; ++q) {
        T a = first ? c * lo[q] : acc[q] + c * lo[q];
        acc[q] = a - c * hi[q];
    }
    T d = first ? -(c * lo[NF - 1]) : acc[NF - 1] - c * lo[NF - 1];
    acc[NF - 1] = d + c * hi[NF - 1];
}

// Body force and face stresses of one cell (src/ExplicitSolver.cpp:498-535,
// src/SemiImplicitSolver.cpp:484-520), added to the flux part in FP64.
template <int DIM, typename T, typename C>
__device__ __forceinline__ void add_sources(const StageParams<T>& P, size_t id, T* acc, C rhoC, const C* uC) {
    if (P.bodyForce) {
        const double a[3] = {P.accel[0], P.accel[1], P.accel[2]};
        const double rho = (double)rhoC;
        double work = (double)uC[0] * a[0];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            acc[1 + d] = (T)((double)acc[1 + d] + rho * a[d]);
            if (d > 0) work += (double)uC[d] * a[d];
        }
        acc[DIM + 1] = (T)((double)acc[DIM + 1] + rho * work);
    }
    if (P.visc) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[1 + d] = (T)((double)acc[1 + d] + P.visc[(size_t)d * P.total + id]);
        acc[DIM + 1] = (T)((double)acc[DIM + 1] + P.visc[(size_t)DIM * P.total + id]);
    }
}

// Stage update of one cell of a single-phase explicit FP32 state, in FP32
// (the stored state is FP32, so the FP64 combination would only add one
// final rounding back to FP32); the dt of the new state converts in FP32 for
// an ideal gas and in FP64 for a stiffened gas.
// Store of the new state in the 1D/2D stage update: non-temporal (the
// values are next read by the following stage launch, not by this one;
// measured 0.5 % faster per launch on the 4096^2 vortex); the 3D z pass
// stores plainly.
template <int DIM, typename T>
__device__ __forceinline__ void store_state(T* a, T v) {
    if constexpr (DIM <= 2) __builtin_nontemporal_store(v, a);
    else *a = v;
}

template <int DIM, typename T, bool SRC>
__device__ __forceinline__ double finish_cell_f32(const StageParams<T>& P, size_t id, T* acc,
                                                  const Cell<DIM, 0, float>& cell, const double* h, double dt,
                                                  bool wantDt, const T* Upre, const T* U0pre) {
    constexpr int NCV = DIM + 2;
    if (SRC) add_sources<DIM, T, float>(P, id, acc, cell.rho, cell.u);
    float un[NCV];
    const bool use0 = P.c2 != 0.0;
    const float c1 = (float)P.c1, c2 = (float)P.c2, c3dt = (float)(P.c3 * dt);
    const float c4 = (float)P.c4, c4i = (float)P.c4inv;
    const float dtf = (float)dt;
#pragma unroll
    for (int q = 0; q < NCV; ++q) {
        const float U = (float)(Upre ? Upre[q] : P.src[(size_t)q * P.total + id]);
        const float R = (float)acc[q];
        if (P.rk1) {
            un[q] = fmaf(dtf, R, U);
        } else {
            float a = c1 * U;
            if (use0) a = fmaf(c2, (float)(U0pre ? U0pre[q] : P.u0[(size_t)q * P.total + id]), a);
            const float b = fmaf(c3dt, R, a);
            // b / c4: exact product for c4 = 1, 2, 4; for c4 = 3 the product with
            // the rounded 1/3 is biased by +3e-8 (a drift of the whole state over
            // many steps), so the quotient gets one residual correction
            float qt = b * c4i;
            if (!P.c4pow2) qt = fmaf(fmaf(-c4, qt, b), c4i, qt);
            un[q] = qt;
        }
        store_state<DIM>(&P.dst[(size_t)q * P.total + id], (T)un[q]);
    }
    if (!wantDt) return 1.0e300;
    if (P.eos.stiffened) {
        double Uq[NCV];
#pragma unroll
        for (int q = 0; q < NCV; ++q) Uq[q] = (double)un[q];
        Cell<DIM, 0, double> c;
        prim_of<DIM, 0>(P.eos, P.mix, Uq, c);
        const double cs = sqrt_fast(div_fast(P.eos.gamma * dmax_ref(c.p + P.eos.pInf, 1e-14), dmax_ref(c.rho, 1e-14)));
        double dc = div_fast(h[0], fabs(c.u[0]) + cs);
#pragma unroll
        for (int d = 1; d < DIM; ++d) dc = dmin_ref(dc, div_fast(h[d], fabs(c.u[d]) + cs));
        return dc;
    }
    Cell<DIM, 0, float> c;
    prim_of_f32<DIM>(P.eos, un, c);
    const float cs = __builtin_amdgcn_sqrtf((float)P.eos.gamma * __builtin_fmaxf(c.p, 1e-14f) *
                                            frcp(__builtin_fmaxf(c.rho, 1e-14f)));
    float dc = (float)h[0] * frcp(fabsf(c.u[0]) + cs);
#pragma unroll
    for (int d = 1; d < DIM; ++d) dc = __builtin_fminf(dc, (float)h[d] * frcp(fabsf(c.u[d]) + cs));
    return (double)dc;
}

// Sources and stage update of one owned cell; returns the acoustic cell dt
// of the new state when requested.  acc holds the complete flux part.
// SRC = false: an instantiation for configurations without body force and
// face stresses, which carries none of their code (P.bodyForce and P.visc
// must then be off).
template <int DIM, int NPH, bool SEMI, typename T, bool SRC = true, typename C = double>
__device__ __forceinline__ double finish_cell(const StageParams<T>& P, size_t id, T* acc,
                                              const Cell<DIM, NPH, C>& cell, const double* h, double dt, bool wantDt,
                                              const T* Upre = nullptr, const T* U0pre = nullptr) {
    if constexpr (sizeof(C) == 4) {
        return finish_cell_f32<DIM, T, SRC>(P, id, acc, cell, h, dt, wantDt, Upre, U0pre);
    }
    constexpr int NCV = DIM + 2 + 2 * NPH;
    constexpr int QP = NCV, QD = NCV + 1;
    // sources: alpha div(u) (src/ExplicitSolver.cpp:486-496), p div(u)
    // (src/SemiImplicitSolver.cpp:472-482)
#pragma unroll
    for (int ph = 0; ph < NPH; ++ph) acc[DIM + 2 + NPH + ph] = acc[DIM + 2 + NPH + ph] + (T)cell.a[ph] * acc[QD];
    if constexpr (SEMI) acc[QP] = acc[QP] + (T)cell.p * acc[QD];
    if (SRC) add_sources<DIM, T, C>(P, id, acc, cell.rho, cell.u);
    double un[NCV];
    double U0v[NCV];
    double Uv[NCV];
    const bool use0 = P.c2 != 0.0;
#pragma unroll
    for (int q = 0; q < NCV; ++q)
        U0v[q] = use0 ? (double)(U0pre ? U0pre[q] : P.u0[(size_t)q * P.total + id]) : 0.0;
#pragma unroll
    for (int q = 0; q < NCV; ++q) Uv[q] = (double)(Upre ? Upre[q] : P.src[(size_t)q * P.total + id]);
    // stored energies relative to the volume fractions they were stored with:
    // the update runs on the total energies (offsets added back here and the
    // new state's offset taken off after the clamps)
    const bool eo = NPH > 0 && !SEMI && P.mix.eoff;
    if (eo) {
        Uv[DIM + 1] += energy_offset<DIM, NPH>(P.mix, Uv);
        if (use0) U0v[DIM + 1] += energy_offset<DIM, NPH>(P.mix, U0v);
    }
#pragma unroll
    for (int q = 0; q < NCV; ++q) {
        double U = Uv[q];
        double Rq = (double)acc[q];
        if (SEMI && q == DIM + 1) {
            un[q] = U;  // total energy is rebuilt by the correction step
            continue;
        }
        if (P.rk1 && (!SEMI || q >= DIM + 2)) {
            un[q] = U + dt * Rq;
        } else {
            double a = P.c1 * U;
            if (use0) a = a + P.c2 * U0v[q];
            const double b = a + (P.c3 * dt) * Rq;
            un[q] = P.c4pow2 ? b * P.c4inv : b / P.c4;
        }
    }
    if constexpr (NPH > 0) {
        // post-update clamp / normalisation (src/ExplicitSolver.cpp:169-185)
        double rs = 0.0;
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) {
            un[DIM + 2 + ph] = dmax_ref(un[DIM + 2 + ph], 1e-14);
            rs += un[DIM + 2 + ph];
        }
        un[0] = rs;
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) un[DIM + 2 + NPH + ph] = dmax_ref(un[DIM + 2 + NPH + ph], P.mix.alphaMin);
        double as = 0.0;
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) as += un[DIM + 2 + NPH + ph];
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) un[DIM + 2 + NPH + ph] /= as;
        if (eo) {
            double st[NCV];
#pragma unroll
            for (int q = 0; q < NCV; ++q) st[q] = (double)(T)un[q];
            un[DIM + 1] -= energy_offset<DIM, NPH>(P.mix, st);
        }
    }
#pragma unroll
    for (int q = 0; q < NCV; ++q)
        if (!(SEMI && q == DIM + 1)) store_state<DIM>(&P.dst[(size_t)q * P.total + id], (T)un[q]);
    if constexpr (SEMI) {
        // advected pressure (src/SemiImplicitSolver.cpp:168,185-186); the
        // stage-0 pressure is the conversion of the backup state
        double p0 = 0.0;
        if (use0) {
            Cell<DIM, NPH> c0;
            prim_of<DIM, NPH>(P.eos, P.mix, U0v, c0);
            p0 = c0.p;
        }
        double a = P.c1 * cell.p + P.c2 * p0;
        P.pAdv[id] = (a + (P.c3 * dt) * (double)acc[QP]) / P.c4;
        P.pres[id] = cell.p;
        double rs = dmax_ref(un[0], 1e-14);
#pragma unroll
        for (int d = 0; d < DIM; ++d) P.vel[d][id] = un[1 + d] / rs;
        if (P.rhoc2) {
            // rho c^2 of the updated state at the stage-start pressure, as the
            // pressure system's set-up forms it (semi_setup_kernel)
            const double rho = (double)(T)un[0];
            double c;
            if constexpr (NPH > 0) {
                double a[kMaxPhases], ar[kMaxPhases];
#pragma unroll
                for (int ph = 0; ph < NPH; ++ph) {
                    a[ph] = (double)(T)un[DIM + 2 + NPH + ph];
                    ar[ph] = (double)(T)un[DIM + 2 + ph];
                }
                c = mix_sound(P.mix, rho, cell.p, a, ar, NPH);
            } else {
                const double pinf = P.eos.stiffened ? P.eos.pInf : 0.0;
                c = sqrt(P.eos.gamma * dmax_ref(cell.p + pinf, 1e-14) / dmax_ref(rho, 1e-14));
            }
            P.rhoc2[id] = rho * c * c;
        }
        return 1.0e300;
    }
    if (!wantDt) return 1.0e300;
    Cell<DIM, NPH> c;
    double Uq[NCV];
#pragma unroll
    for (int q = 0; q < NCV; ++q) Uq[q] = (double)(T)un[q];
    prim_of<DIM, NPH>(P.eos, P.mix, Uq, c);
    double cs;
    if constexpr (NPH == 0) {
        double pinf = P.eos.stiffened ? P.eos.pInf : 0.0;
        cs = sqrt_fast(div_fast(P.eos.gamma * dmax_ref(c.p + pinf, 1e-14), dmax_ref(c.rho, 1e-14)));
    } else {
        cs = mix_sound(P.mix, c.rho, c.p, c.a, c.ar, NPH);
    }
    double dc = t_div<double, NPH == 0>(h[0], fabs(c.u[0]) + cs);
#pragma unroll
    for (int d = 1; d < DIM; ++d) dc = dmin_ref(dc, t_div<double, NPH == 0>(h[d], fabs(c.u[d]) + cs));
    return dc;
}

template <int DIM, int NPH, int R, typename T, typename C>
__device__ __forceinline__ void recon_window(const StageParams<T>& P, const Cell<DIM, NPH, C>* win, const T* sigwin,
                                             int first, bool useSig, T eps, int weno,
                                             FaceState<DIM, NPH, T>& minusS, FaceState<DIM, NPH, T>& plusS) {
    constexpr int NV = n_recon<DIM, NPH>();
    constexpr int S = 2 * R - 1;
    T st[S], mn[NV], pl[NV], vv[S][NV];
#pragma unroll
    for (int s = 0; s < S; ++s) recon_vars<DIM, NPH, T>(win[first + s], vv[s]);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int s = 0; s < S; ++s) st[s] = vv[s][v];
        Recon<R, T>::run(st, eps, weno, mn[v], pl[v]);
    }
    T sm = T(0), sl = T(0);
    if (useSig) {
#pragma unroll
        for (int s = 0; s < S; ++s) st[s] = sigwin[first + s];
        Recon<R, T>::run(st, eps, weno, sm, sl);
    }
    face_state<DIM, NPH, T>(P, mn, sm, minusS);
    face_state<DIM, NPH, T>(P, pl, sl, plusS);
}

// Latency hiding in the y-march: the next window row can be loaded one
// iteration ahead (kPrefetchRow) and the update operands of the current row
// (U and U0 at the cell) at the top of the iteration (kPreloadUpdate).  Each
// costs registers; they are enabled where the kernel stays within its
// occupancy step (measured on the 4096^2 vortex).
template <typename T>
struct MarchTuning {
    static constexpr bool kPrefetchRow = SIFV_PREFETCH_ROW_D;
    static constexpr bool kPreloadUpdate = SIFV_PRELOAD_UPDATE_D;
};
template <>
struct MarchTuning<float> {
    static constexpr bool kPrefetchRow = SIFV_PREFETCH_ROW_F;
    static constexpr bool kPreloadUpdate = SIFV_PRELOAD_UPDATE_F;
};

// The two-phase kernels hold nine fields per window row; asking for two
// waves per SIMD costs them a few spilled registers but doubles the latency
// hiding of a one-wave schedule.
#ifndef SIFV_STAGE_WAVES_2D
#define SIFV_STAGE_WAVES_2D 3
#endif
#ifndef SIFV_STAGE_WAVES_NPH
#define SIFV_STAGE_WAVES_NPH 2
#endif
#ifndef SIFV_STAGE_WAVES_Z
#define SIFV_STAGE_WAVES_Z 1
#endif
#ifndef SIFV_STAGE_WAVES_2D_F
#define SIFV_STAGE_WAVES_2D_F 4
#endif
template <int DIM, int NPH, typename T>
constexpr int stage_min_waves() {
    return NPH > 0 ? SIFV_STAGE_WAVES_NPH : (DIM == 2 ? (sizeof(T) == 8 ? SIFV_STAGE_WAVES_2D : SIFV_STAGE_WAVES_2D_F) : 1);
}

// ===========================================================================
// x + y kernel (1D/2D full stage; 3D x-y part on one plane)
// ===========================================================================
template <int DIM, int R, int RS, int NPH, bool SEMI, int MODE, typename T, bool SRC = true>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(stage_min_waves<DIM, NPH, T>(), 8)))
stage_xy_kernel(StageParams<T> P) {
    constexpr int NV = n_recon<DIM, NPH>();   // reconstructed variables
    constexpr int NF = n_fc<DIM, NPH>();      // face contribution length
    constexpr int NFL = DIM + 4 + NPH;        // Riemann output length
    constexpr int E = R - 1;                  // extra LDS columns per side
    constexpr int S = 2 * R - 1;              // stencil width
    constexpr int WLO = (R >= 2) ? (2 - R) : 0;
    constexpr int WS = R - WLO + 1;           // window rows
    constexpr int WROW = -WLO;                // window slot of row r
    constexpr int WST = (R >= 2) ? 0 : 1;     // first stencil slot of row r+1
    constexpr bool YDIR = DIM >= 2;

    if (P.active && *P.active == 0) return;
    const Grid& g = P.g;
    const int lane = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int perPlane = P.nstrip * P.nseg;
    const int kzl = (MODE == MODE_XY) ? blk / perPlane : 0;
    const int rem = blk - kzl * perPlane;
    const int kz = (MODE == MODE_XY) ? P.kOff + kzl : 0;
    const int strip = rem / P.nseg;
    const int seg = rem - strip * P.nseg;
    const int X0 = strip * kOwn;
    const int col = X0 - 1 + lane;
    const int Y0 = YDIR ? seg * P.rows : 0;
    const int Y1 = YDIR ? ((Y0 + P.rows < g.ny) ? Y0 + P.rows : g.ny) : 1;
    const bool owned = lane >= 1 && lane <= kOwn && col < g.nx;
    const bool useSig = (RS == RS_LF) && P.igr;
    const bool incP = !SEMI;
    const bool igr = P.igr != 0;

    __shared__ T sp[NV + 1][kWave + 2 * E];

    int ecol = 0;
    if (lane < E) ecol = X0 - 1 - E + lane;
    else if (lane < 2 * E) ecol = X0 + kWave - 1 + (lane - E);

    const T eps = (T)P.eps;
    const int weno = P.weno;
    const double dt = *P.dt;
    const int cc = col < g.nx ? (col < 0 ? 0 : col) : g.nx - 1;
    const double dxc = g.dxw[cc + g.ngx];
    const double dzk = (DIM >= 3) ? g.dzw[kz + g.ngz] : 1.0;
    double dtMin = 1.0e300;
    // face coefficients faceArea / cellVolume (src/ExplicitSolver.cpp:305-340);
    // constant along the march when every row has the same width
    const double dyU = YDIR ? g.dyw[g.ngy] : g.dyw[0];
    const double cxU = (dyU * dzk) / ((dxc * dyU) * dzk);
    const double cyU = (dxc * dzk) / ((dxc * dyU) * dzk);

    using C = CellT<T, NPH, SEMI>;
    // interior fast path: when every column the wave reads (its 64 lanes and
    // the E halo columns either side) lies inside the mesh, a load of an
    // interior row needs no ghost mapping, clamp or velocity flip
    const bool xin = X0 - 1 - E >= 0 && X0 + kWave - 1 + E < g.nx;
    auto loadRaw = [&](int c, int row, RawCell<DIM, NPH, T>& rc) __attribute__((always_inline)) {
        if (xin && row >= 0 && row < g.ny) {
            const size_t id = cell_index(g, c, row, kz);
#pragma unroll
            for (int q = 0; q < DIM + 2 + 2 * NPH; ++q) rc.U[q] = P.src[(size_t)q * P.total + id];
            rc.flip = 0;
        } else {
            load_raw<DIM, NPH, T>(P, P.src, c, row, kz, rc);
        }
    };
    auto loadCell = [&](int c, int row, Cell<DIM, NPH, C>& out) __attribute__((always_inline)) {
        if (xin && row >= 0 && row < g.ny) {
            const size_t id = cell_index(g, c, row, kz);
            T U[DIM + 2 + 2 * NPH];
#pragma unroll
            for (int q = 0; q < DIM + 2 + 2 * NPH; ++q) U[q] = P.src[(size_t)q * P.total + id];
            convert_cell<DIM, NPH, T, C>(P.eos, P.mix, U, out);
        } else {
            load_cell<DIM, NPH>(P, P.src, c, row, kz, out);
        }
    };
    Cell<DIM, NPH, C> win[WS];
    T sigwin[WS];
    FaceState<DIM, NPH, T> plusPrev;
    T Glo[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) Glo[q] = T(0);

    int rStart = Y0;
    if (YDIR) {
        rStart = Y0 - 2;
        // the window as it stands after iteration rStart-1 (slot 0 is slid
        // out before it is read)
#pragma unroll
        for (int w = 1; w < WS; ++w) {
            const int row = rStart - 1 + WLO + w;
            load_cell<DIM, NPH>(P, P.src, col, row, kz, win[w]);
            sigwin[w] = useSig ? (T)load_sigma<DIM>(P, col, row, kz) : T(0);
        }
    }
    // raw values of the row entering the window next iteration
    constexpr bool PF = MarchTuning<T>::kPrefetchRow;
    constexpr bool PU = MarchTuning<T>::kPreloadUpdate && MODE == MODE_FULL;
    RawCell<DIM, NPH, T> nextRow;
    if (YDIR && PF) loadRaw(col, rStart + R, nextRow);
    T preU[DIM + 2 + 2 * NPH], preU0[DIM + 2 + 2 * NPH];
    // U of the rows between entering the window and their update (R + 1 rows)
    // kept in LDS, each lane its own column: the update reads the cell's U
    // from there instead of from HBM (the row entered R iterations earlier).
    // Single-phase full stages whose update does not preload U (LDS budget:
    // 8 KB per wave at FP64; the FP32 kernel, which preloads U and U0 into
    // registers at the top of the row, measured 5 % slower with the ring).
    constexpr bool RING = YDIR && PF && !PU && MODE == MODE_FULL && NPH == 0;
    constexpr int RN = R + 1;
    constexpr int NCVR = DIM + 2 + 2 * NPH;
    __shared__ T sUr[RING ? RN : 1][RING ? NCVR : 1][kWave];

    for (int r = rStart; r < Y1; ++r) {
        T Ghi[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) Ghi[q] = T(0);
        double cy = 0.0, dyr = 1.0;
        Cell<DIM, NPH, C> cur;
        T sigc = T(0);
        if (YDIR) {
#pragma unroll
            for (int w = 0; w < WS - 1; ++w) {
                win[w] = win[w + 1];
                sigwin[w] = sigwin[w + 1];
            }
            if constexpr (PF) {
                raw_to_cell<DIM, NPH, T>(P, nextRow, win[WS - 1]);
                if constexpr (RING) {
                    const int slot = (r + R) % RN;
#pragma unroll
                    for (int q = 0; q < NCVR; ++q) sUr[slot][q][lane] = nextRow.U[q];
                }
                if (r + 1 < Y1) loadRaw(col, r + 1 + R, nextRow);
            } else {
                loadCell(col, r + R, win[WS - 1]);
            }
            sigwin[WS - 1] = useSig ? (T)load_sigma<DIM>(P, col, r + R, kz) : T(0);
            // y reconstruction of row r+1
            FaceState<DIM, NPH, T> mY, pY;
            recon_window<DIM, NPH, R, T>(P, win, sigwin, WST, useSig, eps, weno, mY, pY);
            if (r >= Y0 - 1) {
                T F[NFL];
                riemann<RS, 1, DIM, NPH, T>(plusPrev, mY, incP, igr, F);
                // cells of face r+1/2: rows r (slot WROW) and r+1 (slot WROW+1)
                face_contrib<DIM, NPH, T>(F, win[WROW].rho, win[WROW + 1].rho, win[WROW].ar,
                                          win[WROW + 1].ar, Ghi);
            }
            plusPrev = pY;
            if (r < Y0) {
#pragma unroll
                for (int q = 0; q < NF; ++q) Glo[q] = Ghi[q];
                continue;  // warm-up rows: no x phase, no output
            }
            dyr = g.dyw[r + g.ngy];
            if (PU && owned) {
                const size_t pid = cell_index(g, col, r, kz);
#pragma unroll
                for (int q = 0; q < DIM + 2 + 2 * NPH; ++q) {
                    preU[q] = P.src[(size_t)q * P.total + pid];
                    preU0[q] = P.c2 != 0.0 ? P.u0[(size_t)q * P.total + pid] : T(0);
                }
            }
            cy = P.unifY ? cyU : (dxc * dzk) / ((dxc * dyr) * dzk);
            cur = win[WROW];
            sigc = sigwin[WROW];
        } else {
            load_cell<DIM, NPH>(P, P.src, col, r, kz, cur);
            if (useSig) sigc = (T)load_sigma<DIM>(P, col, r, kz);
            dyr = g.dyw[0];
        }

        // ---------------- x phase on row r ----------------
        __syncthreads();  // previous iteration's LDS reads are complete
        {
            T vv[NV];
            recon_vars<DIM, NPH, T>(cur, vv);
#pragma unroll
            for (int v = 0; v < NV; ++v) sp[v][lane + E] = vv[v];
            if (useSig) sp[NV][lane + E] = sigc;
        }
        if (E > 0 && lane < 2 * E) {
            Cell<DIM, NPH, C> ce;
            loadCell(ecol, r, ce);
            T vv[NV];
            recon_vars<DIM, NPH, T>(ce, vv);
            const int pos = lane < E ? lane : kWave + lane;
#pragma unroll
            for (int v = 0; v < NV; ++v) sp[v][pos] = vv[v];
            if (useSig) sp[NV][pos] = (T)load_sigma<DIM>(P, ecol, r, kz);
        }
        __syncthreads();
        FaceState<DIM, NPH, T> mX, pX;
        {
            T st[S], mn[NV], pl[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
#pragma unroll
                for (int s = 0; s < S; ++s) st[s] = sp[v][lane + s];
                Recon<R, T>::run(st, eps, weno, mn[v], pl[v]);
            }
            T sm = T(0), sl = T(0);
            if (useSig) {
#pragma unroll
                for (int s = 0; s < S; ++s) st[s] = sp[NV][lane + s];
                Recon<R, T>::run(st, eps, weno, sm, sl);
            }
            face_state<DIM, NPH, T>(P, mn, sm, mX);
            face_state<DIM, NPH, T>(P, pl, sl, pX);
        }
        // face col+1/2: left = plus of this lane, right = minus of lane+1
        FaceState<DIM, NPH, T> rX;
        rX.rho = __shfl_down(mX.rho, 1, kWave);
#pragma unroll
        for (int d = 0; d < DIM; ++d) rX.u[d] = __shfl_down(mX.u[d], 1, kWave);
        rX.p = __shfl_down(mX.p, 1, kWave);
        rX.sig = __shfl_down(mX.sig, 1, kWave);
        if constexpr (NPH == 0) {
            // single phase: the EOS constants are the same in every lane
            rX.g = mX.g;
            rX.pi = mX.pi;
            rX.gm1i = mX.gm1i;
        } else {
            rX.g = __shfl_down(mX.g, 1, kWave);
            rX.pi = __shfl_down(mX.pi, 1, kWave);
            rX.gm1i = __shfl_down(mX.gm1i, 1, kWave);
        }
#pragma unroll
        for (int ph = 0; ph < NPH; ++ph) rX.a[ph] = __shfl_down(mX.a[ph], 1, kWave);
        T Fx[NFL];
        riemann<RS, 0, DIM, NPH, T>(pX, rX, incP, igr, Fx);
        T FCx[NF];
        {
            C rhoR = __shfl_down(cur.rho, 1, kWave);
            C arR[NPH > 0 ? NPH : 1];
#pragma unroll
            for (int ph = 0; ph < NPH; ++ph) arR[ph] = __shfl_down(cur.ar[ph], 1, kWave);
            face_contrib<DIM, NPH, T>(Fx, cur.rho, rhoR, cur.ar, arR, FCx);
        }
        T FCl[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) FCl[q] = __shfl_up(FCx[q], 1, kWave);

        if (owned) {
            const double cx = P.unifY ? cxU : (dyr * dzk) / ((dxc * dyr) * dzk);
            T acc[NF];
            accumulate<DIM, NPH, T>(acc, FCl, FCx, (T)cx, true);
            if (YDIR) accumulate<DIM, NPH, T>(acc, Glo, Ghi, (T)cy, false);
            const size_t id = cell_index(g, col, r, kz);
            if (MODE == MODE_XY) {
#pragma unroll
                for (int q = 0; q < NF; ++q)
                    if (part_slot_used<DIM, NPH, SEMI>(q)) P.part[(size_t)q * P.total + id] = acc[q];
            } else {
                const double h[3] = {dxc, dyr, dzk};
                double dc;
                if constexpr (YDIR && PU) {
                    dc = finish_cell<DIM, NPH, SEMI, T, SRC, C>(P, id, acc, cur, h, dt, P.dtPartial != nullptr, preU,
                                                                preU0);
                } else if constexpr (RING) {
                    // the segment's first row entered the window before the
                    // march (not through the ring): read from HBM
                    T uring[NCVR];
                    if (r > Y0) {
#pragma unroll
                        for (int q = 0; q < NCVR; ++q) uring[q] = sUr[r % RN][q][lane];
                    } else {
#pragma unroll
                        for (int q = 0; q < NCVR; ++q) uring[q] = P.src[(size_t)q * P.total + id];
                    }
                    dc = finish_cell<DIM, NPH, SEMI, T, SRC, C>(P, id, acc, cur, h, dt, P.dtPartial != nullptr, uring);
                } else {
                    dc = finish_cell<DIM, NPH, SEMI, T, SRC, C>(P, id, acc, cur, h, dt, P.dtPartial != nullptr);
                }
                dtMin = dmin_ref(dtMin, dc);
            }
        }
        if (YDIR) {
#pragma unroll
            for (int q = 0; q < NF; ++q) Glo[q] = Ghi[q];
        }
    }
    if (MODE == MODE_FULL && P.dtPartial) {
        double m = wave_min(dtMin);
        if (lane == 0) P.dtPartial[blockIdx.x] = m;
    }
}

// ===========================================================================
// 3D z pass: 64 columns per wave, marching in z with a register window
// ===========================================================================
template <int R, int RS, int NPH, bool SEMI, typename T>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SIFV_STAGE_WAVES_Z, 8)))
stage_z_kernel(StageParams<T> P) {
    constexpr int DIM = 3;
    constexpr int NF = n_fc<DIM, NPH>();
    constexpr int NFL = DIM + 4 + NPH;
    constexpr int WLO = (R >= 2) ? (2 - R) : 0;
    constexpr int WS = R - WLO + 1;
    constexpr int WROW = -WLO;
    constexpr int WST = (R >= 2) ? 0 : 1;

    if (P.active && *P.active == 0) return;
    const Grid& g = P.g;
    const int lane = threadIdx.x;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    // blk -> (row j, strip, segment), segment fastest
    const int perRow = P.nstrip * P.nseg;
    const int j = blk / perRow;
    const int rem = blk - j * perRow;
    const int strip = rem / P.nseg;
    const int seg = P.segOff + rem - strip * P.nseg;
    const int col = strip * kWave + lane;
    const int Z0 = seg * P.rows;
    const int Z1 = (Z0 + P.rows < g.nz) ? Z0 + P.rows : g.nz;
    const bool owned = col < g.nx;
    const bool useSig = (RS == RS_LF) && P.igr;
    const bool incP = !SEMI;
    const bool igr = P.igr != 0;
    const T eps = (T)P.eps;
    const int weno = P.weno;
    const double dt = *P.dt;
    const int cc = owned ? col : g.nx - 1;
    const double dxc = g.dxw[cc + g.ngx];
    const double dyj = g.dyw[j + g.ngy];
    double dtMin = 1.0e300;
    const double czU = (dxc * dyj) / ((dxc * dyj) * g.dzw[g.ngz]);

    using C = CellT<T, NPH, SEMI>;
    Cell<DIM, NPH, C> win[WS];
    T sigwin[WS];
    FaceState<DIM, NPH, T> plusPrev;
    T Hlo[NF];
#pragma unroll
    for (int q = 0; q < NF; ++q) Hlo[q] = T(0);
    const int rStart = Z0 - 2;
    // raw values of the plane entering the window next iteration, and the
    // x+y part of the current cell, are loaded one step ahead of their use
    constexpr bool PF = SIFV_PREFETCH_Z;
    constexpr bool PP = SIFV_PRELOAD_PART_Z;
    // the conservative values of each plane as they entered the window, kept
    // in LDS until the plane reaches the window centre R iterations later and
    // the update reads them (instead of a second read from HBM)
    constexpr int NCV = DIM + 2 + 2 * NPH;
    constexpr int RING = R + 1;
    __shared__ T sU[RING][NCV][kWave];
#pragma unroll
    for (int w = 1; w < WS; ++w) {
        const int kk = rStart - 1 + WLO + w;
        if (w == WS - 1) {
            RawCell<DIM, NPH, T> rc;
            load_raw<DIM, NPH, T>(P, P.src, cc, j, kk, rc);
#pragma unroll
            for (int q = 0; q < NCV; ++q) sU[(kk + RING) % RING][q][lane] = rc.U[q];
            raw_to_cell<DIM, NPH, T>(P, rc, win[w]);
        } else {
            load_cell<DIM, NPH>(P, P.src, cc, j, kk, win[w]);
        }
        sigwin[w] = useSig ? (T)load_sigma<DIM>(P, cc, j, kk) : T(0);
    }
    RawCell<DIM, NPH, T> nextRow;
    if (PF) load_raw<DIM, NPH, T>(P, P.src, cc, j, rStart + R, nextRow);
    for (int r = rStart; r < Z1; ++r) {
#pragma unroll
        for (int w = 0; w < WS - 1; ++w) {
            win[w] = win[w + 1];
            sigwin[w] = sigwin[w + 1];
        }
        if constexpr (PF) {
#pragma unroll
            for (int q = 0; q < NCV; ++q) sU[(r + R + RING) % RING][q][lane] = nextRow.U[q];
            raw_to_cell<DIM, NPH, T>(P, nextRow, win[WS - 1]);
            if (r + 1 < Z1) load_raw<DIM, NPH, T>(P, P.src, cc, j, r + 1 + R, nextRow);
        } else {
            RawCell<DIM, NPH, T> rc;
            load_raw<DIM, NPH, T>(P, P.src, cc, j, r + R, rc);
#pragma unroll
            for (int q = 0; q < NCV; ++q) sU[(r + R + RING) % RING][q][lane] = rc.U[q];
            raw_to_cell<DIM, NPH, T>(P, rc, win[WS - 1]);
        }
        sigwin[WS - 1] = useSig ? (T)load_sigma<DIM>(P, cc, j, r + R) : T(0);
        T accPre[NF];
        if (PP && r >= Z0 && owned) {
            const size_t pid = cell_index(g, col, j, r);
#pragma unroll
            for (int q = 0; q < NF; ++q)
                accPre[q] = part_slot_used<DIM, NPH, SEMI>(q) ? P.part[(size_t)q * P.total + pid] : T(0);
        }
        FaceState<DIM, NPH, T> mZ, pZ;
        recon_window<DIM, NPH, R, T>(P, win, sigwin, WST, useSig, eps, weno, mZ, pZ);
        T Hhi[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) Hhi[q] = T(0);
        if (r >= Z0 - 1) {
            T F[NFL];
            riemann<RS, 2, DIM, NPH, T>(plusPrev, mZ, incP, igr, F);
            face_contrib<DIM, NPH, T>(F, win[WROW].rho, win[WROW + 1].rho, win[WROW].ar, win[WROW + 1].ar, Hhi);
        }
        plusPrev = pZ;
        if (r >= Z0 && owned) {
            const double dzr = g.dzw[r + g.ngz];
            const double cz = P.unifY ? czU : (dxc * dyj) / ((dxc * dyj) * dzr);
            const size_t id = cell_index(g, col, j, r);
            T acc[NF];
#pragma unroll
            for (int q = 0; q < NF; ++q)
                acc[q] = PP ? accPre[q] : (part_slot_used<DIM, NPH, SEMI>(q) ? P.part[(size_t)q * P.total + id] : T(0));
            accumulate<DIM, NPH, T>(acc, Hlo, Hhi, (T)cz, false);
            const double h[3] = {dxc, dyj, dzr};
            T Upre[NCV];
#pragma unroll
            for (int q = 0; q < NCV; ++q) Upre[q] = sU[(r + RING) % RING][q][lane];
            double dc = finish_cell<DIM, NPH, SEMI, T, true, C>(P, id, acc, win[WROW], h, dt, P.dtPartial != nullptr,
                                                                Upre);
            dtMin = dmin_ref(dtMin, dc);
        }
#pragma unroll
        for (int q = 0; q < NF; ++q) Hlo[q] = Hhi[q];
    }
    if (P.dtPartial) {
        double m = wave_min(dtMin);
        if (lane == 0) P.dtPartial[blockIdx.x] = m;
    }
}

}  // namespace sifv
