// sifv_api.hip - C ABI of the MI355X hot path (include/sifv_gpu.h).
//
// A handle owns the device-resident state of one solver instance:
//   three conservative buffers (stage ping-pong; RK stage 3 writes back into
//   the buffer holding U^n, which is read only pointwise in that stage),
//   FP64 primitive arrays (rho, vel, pres, sigma) mirroring SolutionState's
//   primitive members, and small scratch (dt partials, device scalars).
// All work is enqueued on the handle's own HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sifv_gpu.h"
#include "sifv_aux.hpp"
#include "sifv_kernels.hpp"
#include "sifv_semi.hpp"
#include "sifv_adjoint.hpp"
#include "sifv_pcg.hpp"

using namespace sifv;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(SIFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// buildExtendedNodes, src/RectilinearMesh.cpp:91-114
std::vector<double> extended_nodes(const double* phys, int n, int ng) {
    std::vector<double> e(n + 2 * ng + 1);
    for (int i = 0; i <= n; ++i) e[ng + i] = phys[i];
    for (int g = 1; g <= ng; ++g) {
        int m = std::min(g - 1, n - 1);
        e[ng - g] = e[ng - g + 1] - (phys[m + 1] - phys[m]);
    }
    for (int g = 1; g <= ng; ++g) {
        int m = std::max(n - g, 0);
        e[ng + n + g] = e[ng + n + g - 1] + (phys[m + 1] - phys[m]);
    }
    return e;
}

int required_ghosts(int order) {
    switch (order) {
        case SIFV_WENO3:
        case SIFV_UPWIND3: return 2;
        case SIFV_WENO5:
        case SIFV_UPWIND5: return 3;
        default: return 1;
    }
}

// SimulationConfig::validate, include/SimulationConfig.hpp:123-209
int validate_config(const sifv_config* c) {
    if (c->dim < 1 || c->dim > 3) return fail(SIFV_ERR_INVALID, "dim must be 1, 2, or 3 (got " + std::to_string(c->dim) + ")");
    if (c->RKOrder < 1 || c->RKOrder > 3) return fail(SIFV_ERR_INVALID, "RKOrder must be 1, 2, or 3 (got " + std::to_string(c->RKOrder) + ")");
    if (c->reconOrder < 0 || c->reconOrder > 5) return fail(SIFV_ERR_INVALID, "unknown reconOrder");
    int req = required_ghosts(c->reconOrder);
    if (c->nGhost < req)
        return fail(SIFV_ERR_INVALID, "nGhost=" + std::to_string(c->nGhost) +
                                          " is too small for chosen reconOrder (need >= " + std::to_string(req) + ")");
    if (c->semiImplicit) {
        const auto& p = c->semiImplicitParams;
        if (p.cfl <= 0) return fail(SIFV_ERR_INVALID, "semiImplicitParams.cfl must be > 0");
        if (p.minDt <= 0) return fail(SIFV_ERR_INVALID, "semiImplicitParams.minDt must be > 0");
        if (p.maxDt <= p.minDt) return fail(SIFV_ERR_INVALID, "semiImplicitParams.maxDt must be > minDt");
        if (p.pressureTol <= 0) return fail(SIFV_ERR_INVALID, "semiImplicitParams.pressureTol must be > 0");
        if (p.maxPressureIters <= 0) return fail(SIFV_ERR_INVALID, "semiImplicitParams.maxPressureIters must be > 0");
    } else {
        const auto& p = c->explicitParams;
        if (p.cfl <= 0) return fail(SIFV_ERR_INVALID, "explicitParams.cfl must be > 0");
        if (p.minDt <= 0) return fail(SIFV_ERR_INVALID, "explicitParams.minDt must be > 0");
        if (p.maxDt <= p.minDt) return fail(SIFV_ERR_INVALID, "explicitParams.maxDt must be > minDt");
    }
    if (c->useIGR) {
        const auto& p = c->igrParams;
        if (p.alphaCoeff <= 0) return fail(SIFV_ERR_INVALID, "igrParams.alphaCoeff must be > 0");
        if (p.IGRIters <= 0) return fail(SIFV_ERR_INVALID, "igrParams.IGRIters must be > 0");
        if (p.IGRWarmStartIters < 0) return fail(SIFV_ERR_INVALID, "igrParams.IGRWarmStartIters must be >= 0");
    }
    const auto& mp = c->multiPhaseParams;
    if (mp.nPhases < 0 || mp.nPhases > SIFV_MAX_PHASES) return fail(SIFV_ERR_INVALID, "multiPhaseParams.nPhases out of range");
    if (mp.nPhases > 0) {
        for (int ph = 0; ph < mp.nPhases; ++ph) {
            if (mp.phases[ph].gamma <= 1.0)
                return fail(SIFV_ERR_INVALID, "multiPhaseParams.phases[" + std::to_string(ph) + "].gamma must be > 1");
            if (mp.phases[ph].pInf < 0.0)
                return fail(SIFV_ERR_INVALID, "multiPhaseParams.phases[" + std::to_string(ph) + "].pInf must be >= 0");
        }
        if (mp.alphaMin <= 0.0) return fail(SIFV_ERR_INVALID, "multiPhaseParams.alphaMin must be > 0");
    }
    if (c->viscousMu < 0.0) return fail(SIFV_ERR_INVALID, "viscousParams.mu must be >= 0");
    if (c->surfaceTensionSigma < 0.0) return fail(SIFV_ERR_INVALID, "surfaceTensionParams.sigma must be >= 0");
    if (c->surfaceTensionEps <= 0.0) return fail(SIFV_ERR_INVALID, "surfaceTensionParams.epsGradAlpha must be > 0");
    if (c->viscousNPhaseMu < 0 || c->viscousNPhaseMu > SIFV_MAX_PHASES)
        return fail(SIFV_ERR_INVALID, "viscousParams.phaseMu: bad phase count");
    // terms outside the GPU path
    if (c->viscousNPhaseMu > 0 && c->viscousNPhaseMu < c->multiPhaseParams.nPhases)
        return fail(SIFV_ERR_UNSUPPORTED, "viscousParams.phaseMu needs one value per phase of a multiphase state");
    // the capillary stress reads alpha[0] (src/SurfaceTension.cpp:19), which a
    // single-phase state does not have
    if (c->surfaceTensionSigma > 0.0 && c->multiPhaseParams.nPhases == 0)
        return fail(SIFV_ERR_UNSUPPORTED, "surface tension needs a multiphase state");

    return SIFV_OK;
}

// upper bound of the slab count of a pipelined 3D stage (events per handle)
constexpr int kMaxSlabs = 8;

struct Timing {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double ms = 0.0;
    long launches = 0;
};

}  // namespace

namespace sifv {
// Error message of the last failing call, shared with the other C-ABI units.
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace sifv

struct sifv_handle {
    int dim, nx, ny, nz, ng, ngx, ngy, ngz, nxT, nyT, nzT;
    size_t total;
    int bc[6], sbc[6];
    int periods[3];
    sifv_config cfg;
    int riemann;
    sifv_eos eos;
    int precision;
    int nph, nc;
    hipStream_t stream = nullptr;
    // 3D: the z passes of a stage's slabs run on this stream while the x+y
    // passes of the following slabs run on `stream` (fork / join by events)
    hipStream_t side = nullptr;
    hipEvent_t slabEv[kMaxSlabs] = {};
    hipEvent_t joinEv = nullptr;
    double *d_dx = nullptr, *d_dy = nullptr, *d_dz = nullptr;
    void* cons[3] = {nullptr, nullptr, nullptr};
    int cur = 0;
    double* rhoP = nullptr;   // SolutionState::rho as seen by the primitive side
    double* vel[3] = {nullptr, nullptr, nullptr};
    double* pres = nullptr;
    double* sigma = nullptr;
    bool sigmaUploaded = false;   // sigma may be nonzero without IGR (a caller uploaded it)
    double* partial = nullptr;
    int npartialCap = 0;
    int npartialValid = 0;
    double* scal = nullptr;   // [0] dt, [1] time
    int* flags = nullptr;     // [0] active, [1] steps
    double* hostPinned = nullptr;
    bool dtNextValid = false;
    bool primFresh = true;
    int rowsOverride = 0;
    int slabsOverride = 0;   // 3D stage slabs (0: the built-in choice)
    std::vector<double> nodes[3];   // physical node coordinates (VTK output)
    double* visc = nullptr;         // viscous right-hand side of the stage (dim+1 fields)
    Timing timing;
    // sifv_explicit_advance: one RK step (dt finalize + stages) captured as a HIP
    // graph and replayed per step (off by default: measured slower than plain
    // launches, which the device-resident time loop already keeps back to back);
    // rebuilt when a launch argument changes
    bool useGraphs = false;
    // captured chunks of Jacobi sweeps (useGraphs): one per chunk of the
    // sweep schedule and density buffer, replayed while the arguments repeat
    struct JacGraph {
        int start, end;
        const double* rho;
        const double* dt;
        double tol;
        hipGraphExec_t exec;
    };
    std::vector<JacGraph> jacGraphs;
    hipGraph_t stepGraph = nullptr;
    hipGraphExec_t stepExec = nullptr;
    double graphEnd = 0.0;
    int graphCur = -1, graphRows = -1, graphNpart = -1;
    long graphReplays = 0;
    // 3D: x+y part of the right-hand side (n_fc fields)
    void* part = nullptr;
    int nfc = 0;
    // Jacobi / semi-implicit scratch (allocated on first use):
    // [0],[1] pressure ping-pong, [2] rho c^2, [3] rhs, [4] advected pressure, [5] density copy
    double* jac[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double* resid = nullptr;
    int residCap = 0;
    double* hostResid = nullptr;
    int lastIters = 0;
    // IGR: right-hand side, diagonal, padded rho; wavefront tiles and flags
    double *igrRhs = nullptr, *igrDiag = nullptr, *igrRho = nullptr;
    int* igrOrder = nullptr;
    double* igrWrap = nullptr;  // wrap side slots of the sweep: x, y, z parts
    // 2D strip-streaming sweep (used when igrStrips > 0)
    int igrStrips = 0, igrPeriod = 0;
    double *igrRec = nullptr, *igrEdge = nullptr;
    int* igrProg = nullptr;
    int* igrFlags = nullptr;   // [ntiles] tile epochs, [ntiles] error flag
    int igrNtx = 0, igrNty = 0, igrTj = 1, igrTiles = 0;
    int igrEpoch = 0;
    int igrBlocks = 1;          // sweep grid: at most the co-resident block count
    bool sigPrim = false;
    bool unifY = false;       // all interior y (and z) widths are the same double
    bool jacMixed = false;    // semi-implicit pressure solve: FP32 sweeps + FP64 refinement
    int psolver = SIFV_PS_JACOBI;  // pressure solver of the semi-implicit step
    bool pcgSym = true;       // no periodic axis with stretched wrap widths (PCG usable)
    double* pcgv[7] = {};     // PCG vectors r z p q d s and the partial sums
    double* pcgW[3] = {};     // PCG face weights (one padded field per active axis)
    double* pcgSc = nullptr;  // PCG device scalars
    int* pcgSt = nullptr;     // PCG device flags
    int* pcgStHost = nullptr; // pinned copy of the flags
    int pcgNblk = 0;
    int mgLevels = 0;         // coarse levels of the multigrid preconditioner (0: not built)
    MgLevel mg[kMgMaxLevels] = {};
    double* mgBuf = nullptr;  // storage of every coarse level
    int mgLevels32 = 0;       // the same levels in FP32 (mixed-precision V-cycle)
    MgLevelT<float> mg32[kMgMaxLevels] = {};
    float* mgBuf32 = nullptr;
    float* pcg32[6] = {};     // FP32 fine level of that V-cycle: W x3, s, d, pre-smoothed q
    double* jacPn = nullptr;  // FP64 Jacobi update of the refined iterate
    float* jac32 = nullptr;   // rho, co, r, d0, d1 (FP32)
    // discrete adjoint: checkpoints of the forward run and work arrays
    std::vector<double*> adjCk;   // states at steps 0, k, 2k, ...
    std::vector<double*> adjSeg;  // states of one segment during the backward sweep
    double* adjWork = nullptr;    // W (4), pbP (8), pbM (8), contrib (40), U1, U2, L1, L2, lam (4 each)
    double* adjGpart = nullptr;
    double* adjGbar = nullptr;
    int adjN = 0, adjK = 0, adjNb = 0;
    double adjDt = 0.0;
    // FP32 state of a stiffened gas (single or two phase, explicit): the
    // stored energy excludes the stiffened-gas constant (Eos::eoff / Mix::eoff);
    // uploads of the total energy wait in ePend (FP64) until the volume
    // fractions they are taken relative to are known (settle_energy)
    bool eoff = false;
    double* ePend = nullptr;
    bool ePending = false;
    // device the handle was created on, and the allocation-free mode of
    // sifv_explicit_advance_noalloc (dt partials in the library's static
    // device array instead of `partial`; nothing may allocate)
    int device = 0;
    bool noalloc = false;

    size_t tsize() const { return precision == SIFV_MIXED ? sizeof(float) : sizeof(double); }
    Grid grid() const {
        Grid g;
        g.dim = dim;
        g.nx = nx; g.ny = ny; g.nz = nz;
        g.ngx = ngx; g.ngy = ngy; g.ngz = ngz;
        g.nxT = nxT; g.nyT = nyT; g.nzT = nzT;
        for (int f = 0; f < 6; ++f) { g.bc[f] = bc[f]; g.sbc[f] = sbc[f]; }
        g.dxw = d_dx; g.dyw = d_dy; g.dzw = d_dz;
        return g;
    }
    Eos eosd() const {
        Eos e;
        e.gamma = eos.gamma;
        e.pInf = eos.kind == SIFV_EOS_STIFFENED ? eos.pInf : 0.0;
        e.stiffened = eos.kind == SIFV_EOS_STIFFENED;
        e.eoff = (eoff && nph == 0 && e.stiffened) ? 1 : 0;
        return e;
    }
    // single-phase offset gamma pInf / (gamma - 1) of the stored energy
    double eoffConst() const {
        return (eoff && nph == 0 && eos.kind == SIFV_EOS_STIFFENED) ? eos.gamma * eos.pInf / (eos.gamma - 1.0) : 0.0;
    }
    Mix mix() const {
        Mix m;
        m.nph = nph;
        for (int p = 0; p < kMaxPhases; ++p) {
            m.gamma[p] = p < nph ? cfg.multiPhaseParams.phases[p].gamma : 1.4;
            m.pInf[p] = p < nph ? cfg.multiPhaseParams.phases[p].pInf : 0.0;
        }
        m.alphaMin = cfg.multiPhaseParams.alphaMin;
        m.eoff = (eoff && nph > 0) ? 1 : 0;
        for (int p = 0; p < kMaxPhases; ++p) {
            m.eC[p] = p < nph ? m.gamma[p] * m.pInf[p] / (m.gamma[p] - 1.0) : 0.0;
            m.gi[p] = p < nph ? 1.0 / (m.gamma[p] - 1.0) : 0.0;
        }
        return m;
    }
    template <typename T>
    T* consField(int buf, int q) const { return static_cast<T*>(cons[buf]) + (size_t)q * total; }
};

namespace {

int run_igr_any(sifv_handle* h, int src, const int* active);

int ensure_partials(sifv_handle* h, int n) {
    if (n <= h->npartialCap) return SIFV_OK;
    if (h->noalloc) return fail(SIFV_ERR_INVALID, "dt partials exceed the allocation-free scratch");
    if (h->partial) HIPCHK(hipFree(h->partial));
    h->partial = nullptr;
    HIPCHK(hipMalloc(&h->partial, sizeof(double) * n));
    h->npartialCap = n;
    return SIFV_OK;
}

// ---- stage launch ---------------------------------------------------------
int recon_radius(int order) { return required_ghosts(order); }

void stage_shape(const sifv_handle* h, int& nstrip, int& nseg, int& rows) {
    nstrip = (h->nx + kOwn - 1) / kOwn;
    if (h->dim == 1) {
        nseg = 1;
        rows = 1;
        return;
    }
    long planes = h->dim == 3 ? h->nz : 1;
    if (h->rowsOverride > 0) {
        rows = h->rowsOverride;
    } else {
        // aim for >= ~8k waves so every CU holds several waves (2D: ~16k,
        // 16 rows per wave at 4096^2, the fastest of 8..128 rows measured for
        // both the FP64 and the FP32-bulk kernels); meshes too small for 8 rows
        // per wave at that count get about one wave per SIMD slot (3072 at three
        // waves per SIMD) with 4..8 rows, trading the 2-row warm-up against
        // filling the chip (1024^2: 6 rows, measured fastest of 4..16)
        long target = h->dim == 2 ? 16384 : 8192;
        const long cols = (long)h->ny * nstrip * planes;
        long r = cols / target;
        if (r >= 8) {
            rows = 8;
            while (rows < 64 && rows * 2 <= r) rows *= 2;
        } else {
            rows = (int)std::max<long>(4, std::min<long>(8, (cols + 3071) / 3072));
        }
    }
    rows = std::min(rows, h->ny);
    nseg = (h->ny + rows - 1) / rows;
}

void zpass_shape(const sifv_handle* h, int& nstrip, int& nseg, int& rows) {
    nstrip = (h->nx + kWave - 1) / kWave;
    if (h->rowsOverride > 0) {
        rows = h->rowsOverride;
    } else {
        long target = 8192;
        long r = (long)h->nz * nstrip * h->ny / target;
        rows = 8;
        while (rows < 64 && rows * 2 <= r) rows *= 2;
    }
    rows = std::min(rows, h->nz);
    nseg = (h->nz + rows - 1) / rows;
}

bool stage_lean(const sifv_config& c);

template <typename T>
StageKernels<T> stage_kernels(const sifv_handle* h, bool semi);

template <>
StageKernels<double> stage_kernels<double>(const sifv_handle* h, bool semi) {
    int R = recon_radius(h->cfg.reconOrder);
    if (h->dim == 1) return stage_kernels_1d(R, h->riemann, h->nph, semi);
    if (h->dim == 2) return stage_kernels_2d(R, h->riemann, h->nph, semi, stage_lean(h->cfg));
    return semi ? stage_kernels_3d_semi(R, h->riemann, h->nph) : stage_kernels_3d(R, h->riemann, h->nph);
}

template <>
StageKernels<float> stage_kernels<float>(const sifv_handle* h, bool semi) {
    if (semi) return StageKernels<float>();
    if (h->nph == 2) return stage_kernels_fp32_mp(h->dim, recon_radius(h->cfg.reconOrder), h->riemann);
    if (h->nph) return StageKernels<float>();
    return stage_kernels_fp32(h->dim, recon_radius(h->cfg.reconOrder), h->riemann, stage_lean(h->cfg));
}

// SimulationConfig::hasBodyForce (include/SimulationConfig.hpp:101-106)
bool has_body_force(const sifv_config& c) {
    for (int d = 0; d < 3; ++d)
        if (c.bodyForceParams.a[d] != 0.0 || c.bodyForceParams.b[d] != 0.0) return true;
    return false;
}

// SimulationConfig::hasViscosity (include/SimulationConfig.hpp:96-98)
bool has_viscosity(const sifv_config& c) { return c.viscousMu > 0.0 || c.viscousNPhaseMu > 0; }
// SimulationConfig::hasSurfaceTension (include/SimulationConfig.hpp:99)
bool has_surface_tension(const sifv_config& c) { return c.surfaceTensionSigma > 0.0; }
bool has_face_stress(const sifv_config& c) { return has_viscosity(c) || has_surface_tension(c); }
// No source term of the stage update beyond the flux terms (SRC = false kernels).
bool stage_lean(const sifv_config& c) { return !has_body_force(c) && !has_face_stress(c); }

int timed_launch(sifv_handle* h, const void* fn, dim3 grid, dim3 block, void** args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (h->timing.on) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, h->stream));
    }
    HIPCHK(hipLaunchKernel(fn, grid, block, args, 0, h->stream));
    if (h->timing.on) {
        HIPCHK(hipEventRecord(e1, h->stream));
        h->timing.pending.push_back({e0, e1});
    }
    return SIFV_OK;
}

// One RK stage of the fused kernel(s).  semi: advective fluxes + star update
// (writes pres / vel / pAdv) instead of the explicit update.
// Slabs of z segments per 3D stage: the x+y pass of slab k+1 (VALU-bound)
// runs beside the z pass of slab k (memory-bound), which needs only slab k's
// x+y part.  Built-in choice: SIFV_STAGE_SLABS slabs when each slab's z launch
// still holds kSlabMinWaves waves (four rounds of the chip's 2048 slots at the
// z pass's two waves per SIMD; the 512^3 water/air bubble: 4 slabs of 8192
// waves, 86.9 -> 83.3 ms per step; 256^3 stays at one launch, where 4 slabs of
// 2048 waves measured 0.8 % slower).  One slab (a
// single launch of each pass on the handle's stream) under stream capture,
// live kernel timing and the allocation-free entry point.
#ifndef SIFV_STAGE_SLABS
#define SIFV_STAGE_SLABS 4
#endif
static_assert(SIFV_STAGE_SLABS >= 1 && SIFV_STAGE_SLABS <= kMaxSlabs, "slab count out of range");
constexpr long kSlabMinWaves = 8192;

int stage_slabs(sifv_handle* h) {
    if (h->dim != 3 || !h->side || h->noalloc || h->timing.on) return 1;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return 1;
    int zs, zn, zr;
    zpass_shape(h, zs, zn, zr);
    long s = h->slabsOverride;
    if (s <= 0) s = std::min<long>(SIFV_STAGE_SLABS, (long)zs * zn * h->ny / kSlabMinWaves);
    return (int)std::max<long>(1, std::min<long>(s, zn));
}

// Both passes of a 3D stage over `slabs` slabs of z segments.  The x+y
// launches stay on h->stream in slab order; the z launch of slab k waits for
// the event recorded after x+y launch k and runs on h->side; h->stream then
// waits for the last z launch.  The passes of different slabs write disjoint
// cells (the x+y part, the new state and the semi-implicit outputs of their
// own planes) and both only read the stage's source state.
template <typename T>
int launch_stage_slabs(sifv_handle* h, const StageKernels<T>& k, const StageParams<T>& P, int slabs, bool wantDt,
                       int nbs) {
    StageParams<T> Z = P;
    int zs, zn, zr;
    zpass_shape(h, zs, zn, zr);
    Z.rows = zr;
    Z.nstrip = zs;
    Z.dtPartial = nullptr;
    if (wantDt) {
        int rc = ensure_partials(h, zs * zn * h->ny + nbs);
        if (rc) return rc;
    }
    int zoff = 0;
    hipError_t e = hipSuccess;
    for (int s = 0; s < slabs && e == hipSuccess; ++s) {
        const int segLo = (int)((long)s * zn / slabs), segHi = (int)((long)(s + 1) * zn / slabs);
        const int pLo = segLo * zr, pHi = std::min(segHi * zr, h->nz);
        StageParams<T> X = P;
        X.kOff = pLo;
        void* xargs[] = {&X};
        e = hipLaunchKernel(reinterpret_cast<const void*>(k.xy), dim3(P.nstrip * P.nseg * (pHi - pLo)), dim3(kWave),
                            xargs, 0, h->stream);
        if (e == hipSuccess) e = hipEventRecord(h->slabEv[s], h->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(h->side, h->slabEv[s], 0);
        StageParams<T> Zs = Z;
        Zs.segOff = segLo;
        Zs.nseg = segHi - segLo;
        const int nbz = zs * Zs.nseg * h->ny;
        if (wantDt) Zs.dtPartial = h->partial + zoff;
        zoff += nbz;
        void* zargs[] = {&Zs};
        if (e == hipSuccess)
            e = hipLaunchKernel(reinterpret_cast<const void*>(k.z), dim3(nbz), dim3(kWave), zargs, 0, h->side);
    }
    // the handle's stream waits for every z launch already queued, also after a
    // failed launch, so that nothing read through it races the second stream
    const hipError_t ej = hipEventRecord(h->joinEv, h->side);
    const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(h->stream, h->joinEv, 0) : ej;
    HIPCHK(e);
    HIPCHK(ew);
    if (wantDt) h->npartialValid = zoff;
    return SIFV_OK;
}

template <typename T>
int launch_stage_p(sifv_handle* h, const T* srcp, const T* u0p, T* dstp, const double* coef, bool rk1, bool last,
                   const int* active, bool semi) {
    StageParams<T> P;
    P.g = h->grid();
    P.eos = h->eosd();
    P.mix = h->mix();
    P.total = h->total;
    P.src = srcp;
    P.u0 = u0p ? u0p : srcp;
    P.dst = dstp;
    P.part = static_cast<T*>(h->part);
    P.sigma = h->sigma;
    P.c1 = coef[0]; P.c2 = coef[1]; P.c3 = coef[2]; P.c4 = coef[3];
    P.c4inv = 1.0 / coef[3];
    P.c4pow2 = (coef[3] == 1.0 || coef[3] == 2.0 || coef[3] == 4.0) ? 1 : 0;
    P.unifY = h->unifY ? 1 : 0;
    P.rk1 = rk1 ? 1 : 0;
    P.dt = h->scal;
    P.active = active;
    P.eps = h->cfg.wenoEps;
    P.igr = h->cfg.useIGR ? 1 : 0;
    // FP32 WENO of a stiffened gas: indicators scaled before squaring
    P.weno = (h->cfg.reconOrder <= SIFV_WENO5) ? (h->eoff ? 2 : 1) : 0;
    P.sigPrim = h->sigPrim ? 1 : 0;
    // semi-implicit stages write the stage-start pressure straight into the
    // pressure solve's initial guess (the set-up then has no copy to make)
    P.pres = semi ? h->jac[0] : h->pres;
    for (int d = 0; d < 3; ++d) P.vel[d] = h->vel[d] ? h->vel[d] : h->vel[0];
    P.pAdv = h->jac[4];
    // semi-implicit stages also leave rho c^2 of the updated state for the
    // pressure system (semi_setup_kernel reads it then)
    P.rhoc2 = semi ? h->jac[2] : nullptr;
    P.bodyForce = has_body_force(h->cfg) ? 1 : 0;
    P.accel = h->scal + 4;
    P.visc = has_face_stress(h->cfg) ? h->visc : nullptr;
    StageKernels<T> k = stage_kernels<T>(h, semi);
    if (!k.xy || (h->dim == 3 && !k.z)) return fail(SIFV_ERR_UNSUPPORTED, "no stage kernel for this configuration");
    const bool wantDt = last && !semi && h->cfg.explicitParams.constDt <= 0.0;
    int nstrip, nseg, rows;
    stage_shape(h, nstrip, nseg, rows);
    P.rows = rows;
    P.nseg = nseg;
    P.nstrip = nstrip;
    int nb = nstrip * nseg * (h->dim == 3 ? h->nz : 1);
    // viscous / capillary limits of the new state: their own reduction after
    // the stage, into the partial slots after the fused acoustic minima
    const double viscMu = h->cfg.viscousMu > 0.0 ? h->cfg.viscousMu : 0.0;
    const double capSigma = has_surface_tension(h->cfg) ? h->cfg.surfaceTensionSigma : 0.0;
    const bool stressDt = wantDt && (viscMu > 0.0 || capSigma > 0.0);
    const size_t ncell = (size_t)h->nx * h->ny * h->nz;
    const int nbs = stressDt ? (int)std::min<size_t>((ncell + kRedThreads - 1) / kRedThreads, 2048) : 0;
    P.dtPartial = nullptr;
    if (wantDt && h->dim < 3) {
        int rc = ensure_partials(h, nb + nbs);
        if (rc) return rc;
        P.dtPartial = h->partial;
    }
    void* args[] = {&P};
    int rc = SIFV_OK;
    const int slabs = h->dim == 3 ? stage_slabs(h) : 1;
    if (slabs > 1) {
        rc = launch_stage_slabs<T>(h, k, P, slabs, wantDt, nbs);
        if (rc) return rc;
    } else {
        rc = timed_launch(h, reinterpret_cast<const void*>(k.xy), dim3(nb), dim3(kWave), args);
        if (rc) return rc;
    }
    if (wantDt && h->dim < 3) h->npartialValid = nb;
    if (h->dim == 3 && slabs == 1) {
        StageParams<T> Z = P;
        int zs, zn, zr;
        zpass_shape(h, zs, zn, zr);
        Z.rows = zr;
        Z.nseg = zn;
        Z.nstrip = zs;
        int nbz = zs * zn * h->ny;
        Z.dtPartial = nullptr;
        if (wantDt) {
            rc = ensure_partials(h, nbz + nbs);
            if (rc) return rc;
            Z.dtPartial = h->partial;
        }
        void* zargs[] = {&Z};
        rc = timed_launch(h, reinterpret_cast<const void*>(k.z), dim3(nbz), dim3(kWave), zargs);
        if (rc) return rc;
        if (wantDt) h->npartialValid = nbz;
    }
    if (stressDt) {
        // room for the stage's partials plus these was ensured above
        const int base = h->npartialValid;
        if (base + nbs > h->npartialCap) return fail(SIFV_ERR_INVALID, "partial buffer too small");
        StressDtArgs<T> S;
        S.g = h->grid();
        S.rho = P.dst;
        S.active = active;
        S.viscMu = viscMu;
        S.capSigma = capSigma;
        S.partial = h->partial + base;
        hipLaunchKernelGGL(stress_dt_kernel<T>, dim3(nbs), dim3(kRedThreads), 0, h->stream, S);
        HIPCHK(hipGetLastError());
        h->npartialValid = base + nbs;
    }
    return SIFV_OK;
}

template <typename T>
int launch_stage(sifv_handle* h, int src, int u0, int dst, const double* coef, bool rk1, bool last,
                 const int* active, bool semi) {
    return launch_stage_p<T>(h, static_cast<const T*>(h->cons[src]),
                             u0 >= 0 ? static_cast<const T*>(h->cons[u0]) : nullptr, static_cast<T*>(h->cons[dst]),
                             coef, rk1, last, active, semi);
}

int harvest_timing(sifv_handle* h) {
    for (auto& pr : h->timing.pending) {
        float ms = 0.0f;
        HIPCHK(hipEventSynchronize(pr.second));
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        h->timing.ms += ms;
        h->timing.launches += 1;
        HIPCHK(hipEventDestroy(pr.first));
        HIPCHK(hipEventDestroy(pr.second));
    }
    h->timing.pending.clear();
    return SIFV_OK;
}

const double kRK[3][3][4] = {
    {{1, 0, 1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}},
    {{1, 0, 1, 1}, {1, 1, 1, 2}, {0, 0, 0, 0}},
    {{1, 0, 1, 1}, {1, 3, 1, 4}, {2, 1, 2, 3}},
};

// withStress: the step's dt, where computeViscousDt and computeCapillaryDt
// join the acoustic (src/ExplicitSolver.cpp:78-85) or advective
// (src/SemiImplicitSolver.cpp:112-119) limit; the standalone dt queries leave
// them out.
int launch_dt_cells(sifv_handle* h, bool advective, const double* rho = nullptr, bool withStress = false) {
    DtArgs A;
    A.g = h->grid();
    A.eos = h->eosd();
    A.mix = h->mix();
    A.rho = rho ? rho : h->rhoP;
    for (int d = 0; d < 3; ++d) A.vel[d] = h->vel[d] ? h->vel[d] : h->vel[0];
    A.pres = h->pres;
    A.f32 = h->precision == SIFV_MIXED ? 1 : 0;
    for (int p = 0; p < kMaxPhases; ++p) {
        int pp = p < h->nph ? p : 0;
        const bool f = A.f32 != 0;
        A.alphaRho[p] = (h->nph && !f) ? h->consField<double>(h->cur, h->dim + 2 + pp) : nullptr;
        A.alpha[p] = (h->nph && !f) ? h->consField<double>(h->cur, h->dim + 2 + h->nph + pp) : nullptr;
        A.alphaRho32[p] = (h->nph && f) ? h->consField<float>(h->cur, h->dim + 2 + pp) : nullptr;
        A.alpha32[p] = (h->nph && f) ? h->consField<float>(h->cur, h->dim + 2 + h->nph + pp) : nullptr;
    }
    A.advective = advective ? 1 : 0;
    A.viscMu = (withStress && h->cfg.viscousMu > 0.0) ? h->cfg.viscousMu : 0.0;
    A.capSigma = (withStress && has_surface_tension(h->cfg)) ? h->cfg.surfaceTensionSigma : 0.0;
    size_t n = (size_t)h->nx * h->ny * h->nz;
    int nb = (int)std::min<size_t>((n + kRedThreads - 1) / kRedThreads, 2048);
    int rc = ensure_partials(h, nb);
    if (rc) return rc;
    A.partial = h->partial;
    void (*fn)(DtArgs) = h->dim == 1 ? dt_cells_kernel<1> : h->dim == 2 ? dt_cells_kernel<2> : dt_cells_kernel<3>;
    hipLaunchKernelGGL(fn, dim3(nb), dim3(kRedThreads), 0, h->stream, A);
    HIPCHK(hipGetLastError());
    h->npartialValid = nb;
    return SIFV_OK;
}

int launch_finalize(sifv_handle* h, double cfl, double maxDt, double minDt, double constDt,
                    double target, bool clampMin, bool loop, double endTime) {
    DtFinalize F;
    F.partial = h->partial;
    F.npartial = h->npartialValid;
    F.cfl = cfl;
    F.maxDt = maxDt;
    F.minDt = minDt;
    F.constDt = constDt;
    F.target = target;
    F.clampMin = clampMin ? 1 : 0;
    F.loop = loop ? 1 : 0;
    F.endTime = endTime;
    F.time0 = h->cfg.time;
    F.bodyForce = has_body_force(h->cfg) ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
        F.bfA[d] = h->cfg.bodyForceParams.a[d];
        F.bfB[d] = h->cfg.bodyForceParams.b[d];
        F.bfC[d] = h->cfg.bodyForceParams.c[d];
        F.bfD[d] = h->cfg.bodyForceParams.d[d];
    }
    F.scal = h->scal;
    F.flags = h->flags;
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(kRedThreads), 0, h->stream, F);
    HIPCHK(hipGetLastError());
    return SIFV_OK;
}

// Prim refresh after a step: cons->prim on interior cells plus the PRIM
// ghost fill (the finalize block of ExplicitSolver::step, src/ExplicitSolver.cpp:192-200).
int fill_dirs(sifv_handle* h, double** f, const int* comp, int nf, bool scalar,
              const double* skipOnPeriods = nullptr);

template <typename T>
int launch_c2p(sifv_handle* h) {
    C2PArgs<T> A;
    A.g = h->grid();
    A.eos = h->eosd();
    A.mix = h->mix();
    for (int q = 0; q < 5 + 2 * kMaxPhases; ++q) A.cons[q] = h->consField<T>(h->cur, q < h->nc ? q : 0);
    A.rho = h->rhoP;
    for (int d = 0; d < 3; ++d) A.vel[d] = h->vel[d] ? h->vel[d] : h->vel[0];
    A.pres = h->pres;
    size_t n = (size_t)h->nx * h->ny * h->nz;
    int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(cons_to_prim_kernel<T>, dim3(nb), dim3(256), 0, h->stream, A);
    HIPCHK(hipGetLastError());
    return SIFV_OK;
}

int prim_field_set(sifv_handle* h, double** f, int* comp) {
    int n = 0;
    f[n] = h->rhoP; comp[n++] = 0;
    for (int d = 0; d < h->dim; ++d) { f[n] = h->vel[d]; comp[n++] = 1 + d; }
    f[n] = h->pres; comp[n++] = 0;
    f[n] = h->sigma; comp[n++] = 0;
    if (h->precision == SIFV_FP64)
        for (int q = h->dim + 2; q < h->nc; ++q) { f[n] = h->consField<double>(h->cur, q); comp[n++] = 0; }
    return n;
}

// FP32 stiffened-gas state (sifv_handle::eoff): total energies of the current
// state into ePend (toStored false) or the stored energies from ePend.
int energy_convert(sifv_handle* h, bool toStored) {
    if (!h->ePend) {
        if (h->noalloc) return fail(SIFV_ERR_INVALID, "energy staging buffer missing");
        HIPCHK(hipMalloc(&h->ePend, sizeof(double) * h->total));
    }
    EnergyArgs A;
    const Mix m = h->mix();
    A.total = h->total;
    A.E = h->consField<float>(h->cur, h->dim + 1);
    A.nph = h->nph;
    for (int p = 0; p < kMaxPhases; ++p) {
        A.alpha[p] = p < h->nph ? h->consField<float>(h->cur, h->dim + 2 + h->nph + p) : nullptr;
        A.eC[p] = m.eC[p];
    }
    A.c = h->eoffConst();
    A.e64 = h->ePend;
    A.toStored = toStored ? 1 : 0;
    const int nb = (int)std::min<size_t>((h->total + 255) / 256, 8192);
    hipLaunchKernelGGL(energy_offset_kernel, dim3(nb), dim3(256), 0, h->stream, A);
    HIPCHK(hipGetLastError());
    return SIFV_OK;
}

// Stores an uploaded total energy relative to the volume fractions now in the
// state (called before anything reads the state).
int settle_energy(sifv_handle* h) {
    if (!h->ePending) return SIFV_OK;
    int rc = energy_convert(h, true);
    if (rc) return rc;
    h->ePending = false;
    return SIFV_OK;
}

int refresh_prim(sifv_handle* h) {
    int rc = settle_energy(h);
    if (rc) return rc;
    if (h->primFresh) return SIFV_OK;
    if (h->cfg.semiImplicit) {
        // the semi-implicit step leaves the solved pressure and the corrected
        // velocity in the primitive arrays (no closing conversion,
        // src/SemiImplicitSolver.cpp:216-233); only the density copy and the
        // PRIM ghost fill are outstanding
        HIPCHK(hipMemcpyAsync(h->rhoP, h->consField<double>(h->cur, 0), h->total * sizeof(double),
                              hipMemcpyDeviceToDevice, h->stream));
    } else {
        rc = h->precision == SIFV_MIXED ? launch_c2p<float>(h) : launch_c2p<double>(h);
        if (rc) return rc;
    }
    double* f[kMaxFill];
    int comp[kMaxFill];
    int nf = prim_field_set(h, f, comp);
    rc = fill_dirs(h, f, comp, nf, false);
    if (rc) return rc;
    h->primFresh = true;
    return SIFV_OK;
}

int fill_dirs(sifv_handle* h, double** f, const int* comp, int nf, bool scalar, const double* skipOnPeriods) {
    for (int dir = 0; dir < h->dim; ++dir) {
        FillArgs A;
        A.g = h->grid();
        A.dir = dir;
        A.scalar = scalar ? 1 : 0;
        int m = 0;
        for (int q = 0; q < nf; ++q) {
            // the halo pack of a periods axis leaves this field out
            if (skipOnPeriods && f[q] == skipOnPeriods && h->periods[dir]) continue;
            A.f[m] = f[q];
            A.comp[m] = comp[q];
            ++m;
        }
        A.nf = m;
        for (int q = m; q < kMaxFill; ++q) {
            A.f[q] = nullptr;
            A.comp[q] = 0;
        }
        size_t n0, n1, ngd;
        if (dir == 0) { n0 = h->ny; n1 = h->nz; ngd = h->ngx; }
        else if (dir == 1) { n0 = h->nxT; n1 = h->nz; ngd = h->ngy; }
        else { n0 = h->nxT; n1 = h->nyT; ngd = h->ngz; }
        size_t tot = n0 * n1 * 2 * ngd;
        int nb = (int)std::max<size_t>(1, std::min<size_t>((tot + 255) / 256, 4096));
        hipLaunchKernelGGL(fill_dir_kernel, dim3(nb), dim3(256), 0, h->stream, A);
        HIPCHK(hipGetLastError());
    }
    return SIFV_OK;
}

// Enqueue one explicit step (dt finalize + RK stages).  loop: time-loop mode
// with device-resident time and the active flag.
int elementwise_blocks(const sifv_handle* h);

// Viscous right-hand side of the stage whose input is conservative buffer
// `src` (visc_kernel), into h->visc.
template <typename T>
int launch_visc(sifv_handle* h, int src, const int* active) {
    if (!h->visc) {
        if (h->noalloc) return fail(SIFV_ERR_INVALID, "face-stress buffer missing");
        HIPCHK(hipMalloc(&h->visc, sizeof(double) * (h->dim + 1) * h->total));
        HIPCHK(hipMemset(h->visc, 0, sizeof(double) * (h->dim + 1) * h->total));
    }
    ViscFn<T> fn = visc_kernel_for<T>(h->dim, h->nph);
    if (!fn) return fail(SIFV_ERR_UNSUPPORTED, "no viscous kernel for this configuration");
    ViscArgs<T> A;
    A.P.g = h->grid();
    A.P.eos = h->eosd();
    A.P.mix = h->mix();
    A.P.total = h->total;
    A.P.src = h->consField<T>(src, 0);
    A.P.active = active;
    A.viscous = has_viscosity(h->cfg) ? 1 : 0;
    A.mu = h->cfg.viscousMu;
    A.nPhaseMu = h->cfg.viscousNPhaseMu;
    for (int p = 0; p < kMaxPhases; ++p) A.phaseMu[p] = p < A.nPhaseMu ? h->cfg.viscousPhaseMu[p] : 0.0;
    A.sigma = h->cfg.surfaceTensionSigma;
    A.capEps = h->cfg.surfaceTensionEps;
    A.out = h->visc;
    void* args[] = {&A};
    HIPCHK(hipLaunchKernel(reinterpret_cast<const void*>(fn), dim3(elementwise_blocks(h)), dim3(256), args, 0,
                           h->stream));
    return SIFV_OK;
}

int enqueue_explicit_step(sifv_handle* h, double targetDt, bool loop, double endTime) {
    const sifv_config& c = h->cfg;
    int rc;
    if (c.explicitParams.constDt <= 0.0 && !h->dtNextValid) {
        rc = launch_dt_cells(h, false, nullptr, true);
        if (rc) return rc;
    }
    rc = launch_finalize(h, c.explicitParams.cfl, c.explicitParams.maxDt, c.explicitParams.minDt,
                         c.explicitParams.constDt, targetDt, false, loop, endTime);
    if (rc) return rc;
    const int ord = c.RKOrder;
    const int A = h->cur, B = (h->cur + 1) % 3, C = (h->cur + 2) % 3;
    const int* active = loop ? h->flags : nullptr;
    auto stage = [&](int s, int src, int u0, int dst) {
        bool last = s == ord - 1;
        int irc = run_igr_any(h, src, active);
        if (irc) return irc;
        if (has_face_stress(c)) {
            irc = h->precision == SIFV_MIXED ? launch_visc<float>(h, src, active) : launch_visc<double>(h, src, active);
            if (irc) return irc;
        }
        return h->precision == SIFV_MIXED
                   ? launch_stage<float>(h, src, u0, dst, kRK[ord - 1][s], ord == 1, last, active, false)
                   : launch_stage<double>(h, src, u0, dst, kRK[ord - 1][s], ord == 1, last, active, false);
    };
    if (ord == 1) {
        rc = stage(0, A, -1, B);
        if (rc) return rc;
        h->cur = B;
    } else if (ord == 2) {
        rc = stage(0, A, -1, B);
        if (!rc) rc = stage(1, B, A, A);
        if (rc) return rc;
    } else {
        rc = stage(0, A, -1, B);
        if (!rc) rc = stage(1, B, A, C);
        if (!rc) rc = stage(2, C, A, A);
        if (rc) return rc;
    }
    h->dtNextValid = c.explicitParams.constDt <= 0.0;
    h->primFresh = false;
    return SIFV_OK;
}

int elementwise_blocks(const sifv_handle* h) {
    size_t n = (size_t)h->nx * h->ny * h->nz;
    return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192));
}

void drop_jac_graphs(sifv_handle* h) {
    for (auto& g : h->jacGraphs) hipGraphExecDestroy(g.exec);
    h->jacGraphs.clear();
}

int ensure_jacobi(sifv_handle* h, int maxIter) {
    for (int q = 0; q < 6; ++q) {
        if (h->jac[q]) continue;
        HIPCHK(hipMalloc(&h->jac[q], sizeof(double) * h->total));
        HIPCHK(hipMemset(h->jac[q], 0, sizeof(double) * h->total));
    }
    if (maxIter > h->residCap) {
        HIPCHK(hipStreamSynchronize(h->stream));
        drop_jac_graphs(h);
        hipFree(h->resid);
        if (h->hostResid) hipHostFree(h->hostResid);
        h->resid = nullptr;
        h->hostResid = nullptr;
        h->residCap = 0;
        HIPCHK(hipMalloc(&h->resid, sizeof(double) * maxIter));
        HIPCHK(hipHostMalloc(&h->hostResid, sizeof(double) * maxIter, hipHostMallocDefault));
        h->residCap = maxIter;
    }
    return SIFV_OK;
}

int launch_copy_interior(sifv_handle* h, const double* src, double* dst) {
    hipLaunchKernelGGL(copy_interior_kernel, dim3(elementwise_blocks(h)), dim3(256), 0, h->stream, h->grid(), src,
                       dst);
    HIPCHK(hipGetLastError());
    return SIFV_OK;
}

// Jacobi sweeps from h->jac[0] (iterate 0) ping-ponging through jac[0]/jac[1]
// until max |p_new - p| < tol or maxIter sweeps (JacobiPressureSolver::solve,
// src/JacobiPressureSolver.cpp:57-108).  The sweeps are enqueued in chunks;
// between chunks the host reads the per-sweep residuals, and a sweep whose
// predecessor converged returns at once, so a chunk may overshoot harmlessly.
// *n = sweeps done (the reference's iteration count); iterate k lives in
// jac[k % 2] and h->hostResid[0..n) holds the residual history.
// Grid of the set-up and correction bricks (semi_brick): 64 columns x
// kSemiBrickRows rows x kSemiBrickPlanes planes per block.
dim3 semi_brick_grid(const sifv_handle* h) {
    return dim3((h->nx + 63) / 64, (h->ny + kSemiBrickRows - 1) / kSemiBrickRows,
                (h->nz + kSemiBrickPlanes - 1) / kSemiBrickPlanes);
}

// Grid of the tiled sweeps (jacobi_tile_kernel): 64-column x kTileRows-row tiles,
// kTilePlanes planes per block in 3D.
dim3 jacobi_tile_grid(const sifv_handle* h) {
    return dim3((h->nx + 63) / 64, (h->ny + kTileRows - 1) / kTileRows,
                h->dim >= 3 ? (h->nz + kTilePlanes - 1) / kTilePlanes : 1);
}

int run_jacobi(sifv_handle* h, const double* rho, const double* dt, double tol, int maxIter, int* n) {
    // 2D/3D: tiled z-march sweeps; 1D: the elementwise sweep
    const bool tiled = h->dim >= 2;
    JacobiTileArgs<double> Jt;
    Jt.g = h->grid();
    Jt.dim = h->dim;
    Jt.rho = rho;
    Jt.co = h->jac[2];
    Jt.rhs = h->jac[3];
    Jt.resid = h->resid;
    Jt.tol = tol;
    Jt.dt = dt;
    JacobiArgs J;
    J.g = h->grid();
    J.dim = h->dim;
    J.rho = rho;
    J.rhoc2 = h->jac[2];
    J.rhs = h->jac[3];
    J.resid = h->resid;
    J.tol = tol;
    J.dt = dt;
    const int nb = elementwise_blocks(h);
    int it = 0, seen = 0, chunk = 4;
    *n = maxIter;
    // with graphs on, each chunk of tiled sweeps is captured once per density
    // buffer and replayed (a converged sweep's successors return at once)
    const bool graphs = tiled && h->useGraphs && !h->timing.on && !h->noalloc;
    while (it < maxIter) {
        const int end = std::min(maxIter, it + chunk);
        if (graphs) {
            hipGraphExec_t exec = nullptr;
            for (auto& g : h->jacGraphs)
                if (g.start == it && g.end == end && g.rho == rho && g.dt == dt && g.tol == tol) exec = g.exec;
            if (!exec) {
                HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
                hipError_t e = hipSuccess;
                for (int k = it; k < end && e == hipSuccess; ++k) {
                    Jt.iter = k;
                    Jt.pin = h->jac[k % 2];
                    Jt.pout = h->jac[(k + 1) % 2];
                    void* cargs[] = {&Jt};
                    e = hipLaunchKernel(reinterpret_cast<const void*>(jacobi_tile_kernel<double, TILE_SWEEP64>),
                                        jacobi_tile_grid(h), dim3(64, kTileRows), cargs, 0, h->stream);
                }
                hipGraph_t gr = nullptr;
                const hipError_t ec = hipStreamEndCapture(h->stream, &gr);
                if (e == hipSuccess) e = ec;
                if (e == hipSuccess) e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
                if (gr) hipGraphDestroy(gr);
                HIPCHK(e);
                h->jacGraphs.push_back({it, end, rho, dt, tol, exec});
            }
            HIPCHK(hipGraphLaunch(exec, h->stream));
            it = end;
        }
        for (; it < end; ++it) {
            int rc;
            if (tiled) {
                Jt.iter = it;
                Jt.pin = h->jac[it % 2];
                Jt.pout = h->jac[(it + 1) % 2];
                void* args[] = {&Jt};
                rc = timed_launch(h, reinterpret_cast<const void*>(jacobi_tile_kernel<double, TILE_SWEEP64>), jacobi_tile_grid(h),
                                  dim3(64, kTileRows), args);
            } else {
                J.iter = it;
                J.pin = h->jac[it % 2];
                J.pout = h->jac[(it + 1) % 2];
                void* args[] = {&J};
                rc = timed_launch(h, reinterpret_cast<const void*>(jacobi_kernel), dim3(nb), dim3(256), args);
            }
            if (rc) return rc;
        }
        HIPCHK(hipMemcpyAsync(h->hostResid + seen, h->resid + seen, sizeof(double) * (end - seen),
                              hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int t = seen; t < end; ++t)
            if (h->hostResid[t] < tol) {
                *n = t + 1;
                return SIFV_OK;
            }
        seen = end;
        chunk = std::min(chunk * 2, 64);
    }
    return SIFV_OK;
}

// Pressure solve with FP32 Jacobi sweeps and FP64 residual refinement (see
// jacobi_mix_pass_kernel).  It opens with the FP64 solver's first chunk of
// kJacFirstChunk sweeps: a solve that ends there (a flow near rest: 2-3 sweeps)
// is the FP64 solve, bit for bit, and never pays for the FP32 copy of the
// system.  From the chunk's last iterate on, an FP64 pass measures the
// reference's update size max |pn - p| and forms the residual, and FP32 sweeps
// solve for the correction (2 sweeps after the first pass, doubling up to 16),
// repeat.  *n counts FP64 sweeps, passes and FP32 sweeps together, capped at
// maxIter; *pfinal is the last FP64 sweep's output or the FP64 update of the
// last pass (the reference's answer when the update size fell below tol).
constexpr int kJacFirstChunk = 4;

int run_jacobi_mixed(sifv_handle* h, const double* rho, const double* dt, double tol, int maxIter, int* n,
                     const double** pfinal) {
    const int first = std::min(kJacFirstChunk, maxIter);
    int n64 = 0;
    int rc0 = run_jacobi(h, rho, dt, tol, first, &n64);
    if (rc0) return rc0;
    if (h->hostResid[n64 - 1] < tol || first >= maxIter) {
        *n = n64;
        *pfinal = h->jac[n64 % 2];
        return SIFV_OK;
    }
    int inner = 2;
    const size_t T = h->total;
    if (!h->jacPn) {
        HIPCHK(hipMalloc(&h->jacPn, sizeof(double) * T));
        HIPCHK(hipMemset(h->jacPn, 0, sizeof(double) * T));
        HIPCHK(hipMalloc(&h->jac32, sizeof(float) * 5 * T));
        HIPCHK(hipMemset(h->jac32, 0, sizeof(float) * 5 * T));
    }
    float* rho32 = h->jac32;
    float* co32 = h->jac32 + T;
    float* r32 = h->jac32 + 2 * T;
    float* dbuf[2] = {h->jac32 + 3 * T, h->jac32 + 4 * T};
    const int nb = elementwise_blocks(h);
    hipLaunchKernelGGL(to_f32_kernel, dim3(std::min<size_t>((T + 255) / 256, 8192)), dim3(256), 0, h->stream, rho,
                       rho32, T);
    HIPCHK(hipGetLastError());
    JacobiMixArgs M;
    M.g = h->grid();
    M.dim = h->dim;
    M.rho = rho;
    M.rhoc2 = h->jac[2];
    M.rhs = h->jac[3];
    M.pn = h->jacPn;
    M.r32 = r32;
    M.co32 = co32;
    M.rho32 = rho32;
    M.resid = h->resid;
    M.dt = dt;
    JacobiTileArgs<double> Mt;
    Mt.g = h->grid();
    Mt.dim = h->dim;
    Mt.rho = rho;
    Mt.co = h->jac[2];
    Mt.rhs = h->jac[3];
    Mt.resid = h->resid;
    Mt.tol = 0.0;
    Mt.dt = dt;
    Mt.pn = h->jacPn;
    Mt.r32 = r32;
    Mt.co32 = co32;
    Jacobi32Args S;
    S.g = h->grid();
    S.dim = h->dim;
    S.rho32 = rho32;
    S.co32 = co32;
    S.r32 = r32;
    const bool tiled = h->dim >= 2;
    JacobiTileArgs<float> St;
    St.g = h->grid();
    St.dim = h->dim;
    St.rho = rho32;
    St.co = co32;
    St.rhs = r32;
    St.resid = nullptr;
    St.iter = 0;
    St.tol = 0.0;
    St.dt = nullptr;
    // continue from the iterate the FP64 chunk left (sweep k writes jac[(k+1) % 2])
    double* pin = h->jac[first % 2];
    double* pout = h->jac[(first + 1) % 2];
    const float* din = nullptr;
    int total = first, slot = first;
    *n = maxIter;
    while (true) {
        int rc;
        if (tiled) {
            Mt.pin = pin;
            Mt.din = din;
            Mt.pout = pout;
            Mt.iter = slot;
            void* args[] = {&Mt};
            rc = timed_launch(h, reinterpret_cast<const void*>(jacobi_tile_kernel<double, TILE_PASS>),
                              jacobi_tile_grid(h), dim3(64, kTileRows), args);
        } else {
            M.pin = pin;
            M.din = din;
            M.pout = pout;
            M.slot = slot;
            void* args[] = {&M};
            rc = timed_launch(h, reinterpret_cast<const void*>(jacobi_mix_pass_kernel), dim3(nb), dim3(256), args);
        }
        if (rc) return rc;
        ++total;
        HIPCHK(hipMemcpyAsync(h->hostResid + slot, h->resid + slot, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (h->hostResid[slot] < tol) {
            *n = total;
            break;
        }
        if (total >= maxIter) break;
        const int k = std::min(inner, maxIter - total);
        inner = std::min(inner * 2, 16);
        const float* d = nullptr;
        for (int q = 0; q < k; ++q) {
            if (tiled) {
                St.pin = d;
                St.pout = dbuf[q % 2];
                void* sargs[] = {&St};
                rc = timed_launch(h, reinterpret_cast<const void*>(jacobi_tile_kernel<float, TILE_SWEEP32>), jacobi_tile_grid(h),
                                  dim3(64, kTileRows), sargs);
            } else {
                S.din = d;
                S.dout = dbuf[q % 2];
                void* sargs[] = {&S};
                rc = timed_launch(h, reinterpret_cast<const void*>(jacobi32_kernel), dim3(nb), dim3(256), sargs);
            }
            if (rc) return rc;
            d = dbuf[q % 2];
        }
        total += k;
        din = d;
        std::swap(pin, pout);
        ++slot;
    }
    *pfinal = h->jacPn;
    return SIFV_OK;
}

// PCG pressure solve (sifv_pcg.hpp) of the system on rho / jac[2] (rhoc2) /
// jac[3] (rhs) with the device time step dt, x = jac[0] in place (initial
// guess in, answer out; interior cells only).  *n = iterations (0 when the
// guess already passes, maxIter when not converged).
// Coarse levels of the multigrid preconditioner (sifv_pcg.hpp): every axis
// halved (rounding up) while the finest has no coarse level yet or the
// coarsest still has an axis longer than 4 cells; 8 compact arrays per level.
template <typename R>
int ensure_mg_levels(sifv_handle* h, MgLevelT<R>* out, R** bufOut, int* levelsOut) {
    if (*levelsOut) return SIFV_OK;
    static_assert(kMgCoarsestCells == 64, "the coarsest level holds at most 4 x 4 x 4 cells");
    int n[3] = {h->nx, h->ny, h->nz};
    int L = 0;
    size_t total = 0;
    MgLevelT<R> lv[kMgMaxLevels] = {};
    while (L < kMgMaxLevels) {
        const int mx = std::max(n[0], std::max(n[1], n[2]));
        if (L > 0 && mx <= 4) break;
        for (int d = 0; d < 3; ++d) n[d] = (n[d] + 1) / 2;
        for (int d = 0; d < 3; ++d) lv[L].n[d] = n[d];
        lv[L].cells = n[0] * n[1] * n[2];
        total += 8 * (size_t)lv[L].cells;
        ++L;
    }
    if (lv[L - 1].cells > kMgCoarsestCells) return fail(SIFV_ERR_HIP, "multigrid coarsest level too large");
    R* buf = nullptr;
    HIPCHK(hipMalloc(&buf, sizeof(R) * total));
    HIPCHK(hipMemset(buf, 0, sizeof(R) * total));
    R* q = buf;
    for (int l = 0; l < L; ++l) {
        const size_t c = lv[l].cells;
        lv[l].s = q;
        q += c;
        for (int d = 0; d < 3; ++d) {
            lv[l].w[d] = q;
            q += c;
        }
        lv[l].d = q;
        q += c;
        lv[l].b = q;
        q += c;
        lv[l].x = q;
        q += c;
        lv[l].t = q;
        q += c;
    }
    *bufOut = buf;
    for (int l = 0; l < L; ++l) out[l] = lv[l];
    *levelsOut = L;
    return SIFV_OK;
}

int ensure_mg(sifv_handle* h, bool fp32) {
    return fp32 ? ensure_mg_levels<float>(h, h->mg32, &h->mgBuf32, &h->mgLevels32)
                : ensure_mg_levels<double>(h, h->mg, &h->mgBuf, &h->mgLevels);
}

template <typename F1, typename F2, typename F3>
const void* by_dim(int dim, F1 f1, F2 f2, F3 f3) {
    return dim == 1 ? reinterpret_cast<const void*>(f1)
           : dim == 2 ? reinterpret_cast<const void*>(f2) : reinterpret_cast<const void*>(f3);
}

// Multigrid V-cycle of run_pcg over the coarse levels lv[0..L-1] in R
// (double: the FP64 preconditioner; float: the FP32 one of the mixed form):
// build() forms the coarse operators, run() applies z = V(r) on the fine level,
// its kernels skipped once *M.gate is set (null gate during the set-up).
template <typename R>
struct MgCycle {
    sifv_handle* h;
    MgArgsT<R> M;
    const MgLevelT<R>* lv;
    int L, nb, dim;

    int launch(const void* fn, int grid) {
        void* a[] = {&M};
        return timed_launch(h, fn, dim3(grid), dim3(256), a);
    }
    static int cgrid(int cells) { return std::max(1, std::min((cells + 255) / 256, 8192)); }

    int build() {
        int rc = SIFV_OK;
        for (int l = 0; l < L && !rc; ++l) {
            M.C = lv[l];
            if (l == 0) {
                rc = launch(by_dim(dim, mg_build_kernel<1, true, R>, mg_build_kernel<2, true, R>,
                                   mg_build_kernel<3, true, R>),
                            cgrid(M.C.cells));
            } else {
                M.F = lv[l - 1];
                rc = launch(by_dim(dim, mg_build_kernel<1, false, R>, mg_build_kernel<2, false, R>,
                                   mg_build_kernel<3, false, R>),
                            cgrid(M.C.cells));
            }
            if (!rc)
                rc = launch(by_dim(dim, mg_diag_kernel<1, R>, mg_diag_kernel<2, R>, mg_diag_kernel<3, R>),
                            cgrid(M.C.cells));
        }
        return rc;
    }

    // pre: run the fine pre-smooth (inside the iteration pcg_update_kernel
    // has already written it)
    int run(bool pre) {
        int r = pre ? launch(by_dim(dim, mg_pre_fine_kernel<1, R>, mg_pre_fine_kernel<2, R>, mg_pre_fine_kernel<3, R>),
                             nb)
                    : SIFV_OK;
        M.C = lv[0];
        if (!r)
            r = launch(by_dim(dim, mg_restrict_fine_kernel<1, R>, mg_restrict_fine_kernel<2, R>,
                              mg_restrict_fine_kernel<3, R>),
                       cgrid(M.C.cells));
        // levels m..L-1 (at most SIFV_MG_TAIL_CELLS cells each) run in one
        // workgroup, the coarsest included
        int m = L - 1;
        while (m > 0 && lv[m - 1].cells <= SIFV_MG_TAIL_CELLS && L - (m - 1) <= kMgTailMaxLevels) --m;
        for (int l = 0; l < m && !r; ++l) {
            M.C = lv[l];
            r = launch(by_dim(dim, mg_pre_kernel<1, R>, mg_pre_kernel<2, R>, mg_pre_kernel<3, R>), cgrid(M.C.cells));
            if (!r) {
                M.F = lv[l];
                M.C = lv[l + 1];
                r = launch(by_dim(dim, mg_restrict_kernel<1, R>, mg_restrict_kernel<2, R>, mg_restrict_kernel<3, R>),
                           cgrid(M.C.cells));
            }
        }
        if (!r) {
            MgTailArgsT<R> T;
            T.gate = M.gate;
            T.nlev = L - m;
            for (int q = 0; q < T.nlev; ++q) T.lv[q] = lv[m + q];
            void* targs[] = {&T};
            r = timed_launch(h, by_dim(dim, mg_tail_kernel<1, R>, mg_tail_kernel<2, R>, mg_tail_kernel<3, R>),
                             dim3(1), dim3(kMgTailThreads), targs);
        }
        for (int l = m - 1; l >= 0 && !r; --l) {
            M.F = lv[l];
            M.C = lv[l + 1];
            r = launch(reinterpret_cast<const void*>(mg_prolong_kernel<R>), cgrid(M.F.cells));
            if (!r) {
                M.C = lv[l];
                M.src = M.C.x;
                M.dst = M.C.t;
                r = launch(by_dim(dim, mg_smooth_kernel<1, R>, mg_smooth_kernel<2, R>, mg_smooth_kernel<3, R>),
                           cgrid(M.C.cells));
            }
        }
        // fine level: prolongation folded into the post-smooth
        M.C = lv[0];
        if (!r)
            r = launch(by_dim(dim, mg_post_fine_kernel<1, R>, mg_post_fine_kernel<2, R>, mg_post_fine_kernel<3, R>),
                       nb);
        return r;
    }

    // set-up V-cycle (z and p of the initial residual, r.z partials in slot
    // 2), then the gate and slot of the iteration
    int first() {
        M.slot = 2;
        M.init = 1;
        const int rc = run(true);
        M.gate = M.A.st;
        M.slot = 1;
        M.init = 0;
        return rc;
    }
};

int run_pcg(sifv_handle* h, const double* rho, const double* dt, double tol, int maxIter, int* n, bool mg) {
    if (!h->pcgSym)
        return fail(SIFV_ERR_UNSUPPORTED,
                    "PCG needs a symmetric system: a periodic axis with stretched wrap widths (use Jacobi)");
    // the face weights keep the high boundary face of each line in the ghost
    // slot after its last cell
    if (h->ng < 1) return fail(SIFV_ERR_UNSUPPORTED, "the PCG pressure solve needs at least one ghost layer");
    const int nb = elementwise_blocks(h);
    if (!h->pcgv[0] || h->pcgNblk < nb) {
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int q = 0; q < 7; ++q) {
            hipFree(h->pcgv[q]);
            h->pcgv[q] = nullptr;
        }
        for (int q = 0; q < 6; ++q) {
            HIPCHK(hipMalloc(&h->pcgv[q], sizeof(double) * h->total));
            HIPCHK(hipMemset(h->pcgv[q], 0, sizeof(double) * h->total));
        }
        for (int d = 0; d < h->dim; ++d) {
            if (h->pcgW[d]) continue;
            HIPCHK(hipMalloc(&h->pcgW[d], sizeof(double) * h->total));
            HIPCHK(hipMemset(h->pcgW[d], 0, sizeof(double) * h->total));
        }
        HIPCHK(hipMalloc(&h->pcgv[6], sizeof(double) * 3 * nb));
        if (!h->pcgSc) HIPCHK(hipMalloc(&h->pcgSc, sizeof(double) * 8));
        if (!h->pcgSt) HIPCHK(hipMalloc(&h->pcgSt, sizeof(int) * 4));
        if (!h->pcgStHost) HIPCHK(hipHostMalloc(&h->pcgStHost, sizeof(int) * 4, hipHostMallocDefault));
        h->pcgNblk = nb;
    }
    PcgArgs A;
    A.g = h->grid();
    A.rho = rho;
    A.rhoc2 = h->jac[2];
    A.rhs = h->jac[3];
    A.dt = dt;
    A.x = h->jac[0];
    A.r = h->pcgv[0];
    A.z = h->pcgv[1];
    A.p = h->pcgv[2];
    A.q = h->pcgv[3];
    A.d = h->pcgv[4];
    A.s = h->pcgv[5];
    A.part = h->pcgv[6];
    for (int d = 0; d < 3; ++d) A.W[d] = h->pcgW[d];
    A.nblk = nb;
    A.sc = h->pcgSc;
    A.st = h->pcgSt;
    A.tol2 = tol * tol;
    A.maxIter = maxIter;
    A.mgOmega = mg ? mg_omega(h->dim) : 0.0;
    const dim3 blk(256);
    void* args[] = {&A};
    auto launch = [&](const void* fn, int grid) {
        return timed_launch(h, fn, dim3(grid), blk, args);
    };
    const void* setup = h->dim == 1 ? reinterpret_cast<const void*>(pcg_setup_kernel<1>)
                        : h->dim == 2 ? reinterpret_cast<const void*>(pcg_setup_kernel<2>)
                                      : reinterpret_cast<const void*>(pcg_setup_kernel<3>);
    const void* matvec = h->dim == 1 ? reinterpret_cast<const void*>(pcg_matvec_kernel<1>)
                         : h->dim == 2 ? reinterpret_cast<const void*>(pcg_matvec_kernel<2>)
                                       : reinterpret_cast<const void*>(pcg_matvec_kernel<3>);
    // multigrid preconditioner: FP64, or FP32 levels in the mixed form
    const bool mg32 = mg && h->jacMixed;
    MgCycle<double> V64;
    MgCycle<float> V32;
    int rc = SIFV_OK;
    if (mg32) {
        for (int q = 0; q < 6 && !rc; ++q)
            if (!h->pcg32[q] && (q >= 3 || q < h->dim)) {
                HIPCHK(hipMalloc(&h->pcg32[q], sizeof(float) * h->total));
                HIPCHK(hipMemset(h->pcg32[q], 0, sizeof(float) * h->total));
            }
        A.q32 = h->pcg32[5];
        rc = ensure_mg(h, true);
        MgFine32 F32;
        for (int d = 0; d < 3; ++d) F32.W[d] = h->pcg32[d];
        F32.s = h->pcg32[3];
        F32.d = h->pcg32[4];
        F32.q = h->pcg32[5];
        V32 = MgCycle<float>{h, MgArgsT<float>{}, h->mg32, h->mgLevels32, nb, h->dim};
        V32.M.f32 = F32;
    } else if (mg) {
        rc = ensure_mg(h, false);
        V64 = MgCycle<double>{h, MgArgs{}, h->mg, h->mgLevels, nb, h->dim};
    }
    V64.M.A = A;
    V32.M.A = A;
    if (!rc) rc = launch(by_dim(h->dim, pcg_weights_kernel<1>, pcg_weights_kernel<2>, pcg_weights_kernel<3>), nb);
    if (!rc) rc = launch(setup, nb);
    if (!rc && mg32) {
        MgFine32 F32 = V32.M.f32;
        void* cargs[] = {&A, &F32, (void*)&h->dim, (void*)&h->total};
        rc = timed_launch(h, reinterpret_cast<const void*>(mg_fine32_kernel), dim3(nb), blk, cargs);
        if (!rc) rc = V32.build();
        if (!rc) rc = V32.first();
    } else if (!rc && mg) {
        rc = V64.build();
        if (!rc) rc = V64.first();
    }
    if (!rc) rc = launch(reinterpret_cast<const void*>(pcg_setup_fold_kernel), 1);
    if (!rc) rc = launch(reinterpret_cast<const void*>(pcg_zero_kernel), nb);
    if (rc) return rc;
    int it = 0, chunk = 4;
    *n = maxIter;
    while (true) {
        HIPCHK(hipMemcpyAsync(h->pcgStHost, h->pcgSt, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (h->pcgStHost[0]) {
            *n = h->pcgStHost[1];
            return SIFV_OK;
        }
        if (it >= maxIter) return fail(SIFV_ERR_HIP, "PCG iteration count out of step with the device");
        const int end = std::min(maxIter, it + chunk);
        for (; it < end; ++it) {
            rc = launch(matvec, nb);
            if (!rc) rc = launch(reinterpret_cast<const void*>(pcg_alpha_kernel), 1);
            if (!rc) rc = launch(reinterpret_cast<const void*>(pcg_update_kernel), nb);
            if (!rc && mg32) rc = V32.run(false);
            else if (!rc && mg) rc = V64.run(false);
            if (!rc) rc = launch(reinterpret_cast<const void*>(pcg_beta_kernel), 1);
            if (!rc) rc = launch(reinterpret_cast<const void*>(pcg_dir_kernel), nb);
            if (rc) return rc;
        }
        chunk = std::min(chunk * 2, 64);
    }
}

// One semi-implicit step (SemiImplicitSolver::step, src/SemiImplicitSolver.cpp:106-236):
// advective dt, then per RK stage the fused advective stage kernel (star state,
// advected pressure), PRIM ghosts of the star density / velocity, the pressure
// system set-up, Jacobi sweeps and the correction.
int enqueue_semi_step(sifv_handle* h, double targetDt) {
    const sifv_config& c = h->cfg;
    const auto& sp = c.semiImplicitParams;
    int rc = ensure_jacobi(h, sp.maxPressureIters);
    if (rc) return rc;
    rc = launch_dt_cells(h, true, h->consField<double>(h->cur, 0), true);
    if (!rc) rc = launch_finalize(h, sp.cfl, sp.maxDt, sp.minDt, -1.0, targetDt, true, false, 0.0);
    if (rc) return rc;
    const int ord = c.RKOrder;
    const int A = h->cur, B = (h->cur + 1) % 3, C = (h->cur + 2) % 3;
    int src[3], u0[3], dst[3];
    if (ord == 1) { src[0] = A; u0[0] = -1; dst[0] = B; }
    else if (ord == 2) { src[0] = A; u0[0] = -1; dst[0] = B; src[1] = B; u0[1] = A; dst[1] = A; }
    else {
        src[0] = A; u0[0] = -1; dst[0] = B;
        src[1] = B; u0[1] = A; dst[1] = C;
        src[2] = C; u0[2] = A; dst[2] = A;
    }
    const int nb = elementwise_blocks(h);
    for (int s = 0; s < ord; ++s) {
        rc = run_igr_any(h, src[s], nullptr);
        if (rc) return rc;
        if (has_face_stress(c)) {
            rc = launch_visc<double>(h, src[s], nullptr);
            if (rc) return rc;
        }
        rc = launch_stage<double>(h, src[s], u0[s], dst[s], kRK[ord - 1][s], ord == 1, false, nullptr, true);
        if (rc) return rc;
        double* rho = h->consField<double>(dst[s], 0);
        double* f[4];
        int comp[4];
        int nf = 0;
        f[nf] = rho; comp[nf++] = 0;
        for (int d = 0; d < h->dim; ++d) { f[nf] = h->vel[d]; comp[nf++] = 1 + d; }
        rc = fill_dirs(h, f, comp, nf, false);
        if (rc) return rc;
        SemiArgs S;
        S.g = h->grid();
        S.eos = h->eosd();
        S.mix = h->mix();
        S.dim = h->dim;
        S.nph = h->nph;
        S.total = h->total;
        S.rho = rho;
        for (int ph = 0; ph < kMaxPhases; ++ph) {
            int pp = ph < h->nph ? ph : 0;
            S.ar[ph] = h->nph ? h->consField<double>(dst[s], h->dim + 2 + pp) : nullptr;
            S.al[ph] = h->nph ? h->consField<double>(dst[s], h->dim + 2 + h->nph + pp) : nullptr;
        }
        S.pres = h->jac[0];   // the stage update's stage-start pressure
        for (int d = 0; d < 3; ++d) {
            S.vel[d] = h->vel[d] ? h->vel[d] : h->vel[0];
            S.velOut[d] = h->vel[d] ? h->vel[d] : h->vel[0];
        }
        S.pAdv = h->jac[4];
        S.rhoc2 = h->jac[2];
        S.rc2Given = 1;
        S.rhs = h->jac[3];
        S.p0 = h->jac[0];
        S.resid = h->resid;
        S.maxIter = sp.maxPressureIters;
        S.dt = h->scal;
        hipLaunchKernelGGL(semi_setup_kernel, semi_brick_grid(h), dim3(64, kSemiBrickRows), 0, h->stream, S);
        HIPCHK(hipGetLastError());
        int n = 0;
        const double* pfinal = nullptr;
        if (h->psolver == SIFV_PS_PCG || h->psolver == SIFV_PS_MGPCG) {
            rc = run_pcg(h, rho, h->scal, sp.pressureTol, sp.maxPressureIters, &n, h->psolver == SIFV_PS_MGPCG);
            if (rc) return rc;
            // fillScalarGhosts(pressure_) after the solve (src/SemiImplicitSolver.cpp:581)
            double* pf[1] = {h->jac[0]};
            int pc[1] = {0};
            rc = fill_dirs(h, pf, pc, 1, true);
            pfinal = h->jac[0];
        } else if (h->jacMixed) {
            rc = run_jacobi_mixed(h, rho, h->scal, sp.pressureTol, sp.maxPressureIters, &n, &pfinal);
        } else {
            rc = run_jacobi(h, rho, h->scal, sp.pressureTol, sp.maxPressureIters, &n);
            pfinal = h->jac[n % 2];
        }
        if (rc) return rc;
        h->lastIters = n;
        CorrectArgs K;
        K.g = h->grid();
        K.eos = h->eosd();
        K.mix = h->mix();
        K.dim = h->dim;
        K.nph = h->nph;
        K.total = h->total;
        K.cons = h->consField<double>(dst[s], 0);
        K.p = pfinal;
        K.sigma = h->cfg.useIGR || h->sigmaUploaded ? h->sigma : nullptr;   // otherwise zero
        for (int d = 0; d < 3; ++d) K.vel[d] = h->vel[d] ? h->vel[d] : h->vel[0];
        K.pres = h->pres;
        K.storePrim = s == ord - 1 ? 1 : 0;
        K.dt = h->scal;
        hipLaunchKernelGGL(semi_correct_kernel, semi_brick_grid(h), dim3(64, kSemiBrickRows), 0, h->stream, K);
        HIPCHK(hipGetLastError());
    }
    h->cur = dst[ord - 1];
    h->dtNextValid = false;
    h->primFresh = false;
    return SIFV_OK;
}

// Wavefront tiling of the IGR sweep: 64 columns x tj rows of one plane.
// tj minimises the critical path (tile diagonals x steps per tile).
int setup_igr(sifv_handle* h) {
    const int ntx = (h->nx + kWave - 1) / kWave;
    long best = -1;
    int bestTj = 1;
    for (int tj = 1; tj <= kIgrTj; tj *= 2) {
        if (h->dim == 1 && tj > 1) break;
        const int nty = (h->ny + tj - 1) / tj;
        long cost = (long)(ntx + nty + h->nz - 2) * (tj + kWave - 1);
        if (best < 0 || cost < best) { best = cost; bestTj = tj; }
        if (tj >= h->ny) break;
    }
    h->igrTj = bestTj;
    h->igrNtx = ntx;
    h->igrNty = (h->ny + bestTj - 1) / bestTj;
    h->igrTiles = ntx * h->igrNty * h->nz;
    std::vector<int> order(h->igrTiles);
    for (int t = 0; t < h->igrTiles; ++t) order[t] = t;
    auto diag = [&](int t) { return t % ntx + (t / ntx) % h->igrNty + t / (ntx * h->igrNty); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return diag(a) < diag(b); });
    const size_t nb = sizeof(double) * h->total;
    HIPCHK(hipMalloc(&h->igrRhs, nb));
    HIPCHK(hipMalloc(&h->igrDiag, nb));
    HIPCHK(hipMalloc(&h->igrRho, nb));
    HIPCHK(hipMemset(h->igrRhs, 0, nb));
    HIPCHK(hipMemset(h->igrDiag, 0, nb));
    HIPCHK(hipMemset(h->igrRho, 0, nb));
    {
        const size_t nw = 2 * ((size_t)h->ny * h->nz + (size_t)h->nx * h->nz + (size_t)h->nx * h->ny);
        HIPCHK(hipMalloc(&h->igrWrap, sizeof(double) * nw));
        HIPCHK(hipMemset(h->igrWrap, 0, sizeof(double) * nw));
    }
    HIPCHK(hipMalloc(&h->igrOrder, sizeof(int) * h->igrTiles));
    HIPCHK(hipMemcpy(h->igrOrder, order.data(), sizeof(int) * h->igrTiles, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&h->igrFlags, sizeof(int) * (h->igrTiles + 2)));
    HIPCHK(hipMemset(h->igrFlags, 0, sizeof(int) * (h->igrTiles + 2)));
    h->igrEpoch = 0;
    // every sweep block must be co-resident (blocks wait on tiles of other blocks)
    int perCu = 0, dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &perCu, reinterpret_cast<const void*>(igr_sweep_kernel_for(h->dim)), kWave, 0));
    if (perCu < 1 || cus < 1) return fail(SIFV_ERR_HIP, "IGR sweep kernel cannot be resident");
    h->igrBlocks = std::min(h->igrTiles, perCu * cus);
    // 2D: one wave per 64-column strip when every strip can be resident and
    // the rows outnumber a strip's lane skew plus two progress chunks (the
    // wrapping faces of each axis must match: both wrap or neither)
    h->igrStrips = 0;
    const int W = (h->nx + kWave - 1) / kWave;
    const bool wrapPair = (h->sbc[0] == BK_WRAP) == (h->sbc[1] == BK_WRAP) &&
                          (h->sbc[2] == BK_WRAP) == (h->sbc[3] == BK_WRAP);
    if (h->dim == 2 && h->ny >= 2 * kStripChunk + kWave + kStripChunk + 16 && wrapPair) {
        int perCuS = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &perCuS, reinterpret_cast<const void*>(igr_strip_kernels().sweep), 2 * kWave, 0));
        if (W <= perCuS * cus) {
            const bool perX = h->sbc[0] == BK_WRAP;
            const int gap = perX ? std::max(0, W * (kWave + kStripChunk) + kWave + kStripChunk - h->ny) : 0;
            h->igrStrips = W;
            h->igrPeriod = h->ny + gap;
            HIPCHK(hipMalloc(&h->igrRec, sizeof(double) * (size_t)W * h->ny * kStripSlot));
            // edge values [2][W][ny] per side, then a sink for lanes without a value
            HIPCHK(hipMalloc(&h->igrEdge, sizeof(double) * (4 * (size_t)W * h->ny + kWave)));
            HIPCHK(hipMalloc(&h->igrProg, sizeof(int) * W));
        }
    }
    return SIFV_OK;
}

// IGR solve at the start of an RK stage on conservative buffer `src`
// (ExplicitSolver::step, src/ExplicitSolver.cpp:117-121 -> IGRSolver::
// solveEntropicPressure, src/IGR.cpp:102-158): IGRWarmStartIters sweeps at
// config.step == 0, IGRIters otherwise.
template <typename T>
int run_igr(sifv_handle* h, int src, const int* active) {
    const auto& ip = h->cfg.igrParams;
    const int iters = h->cfg.step == 0 ? ip.IGRWarmStartIters : ip.IGRIters;
    h->sigPrim = iters == 0;
    if (iters == 0) return SIFV_OK;
    IgrPreArgs<T> A;
    A.P.g = h->grid();
    A.P.eos = h->eosd();
    A.P.mix = h->mix();
    A.P.total = h->total;
    A.P.src = static_cast<const T*>(h->cons[src]);
    // alpha = alphaCoeff * dx(0)^2 (src/IGR.cpp:45)
    double dx0 = 0.0;
    HIPCHK(hipMemcpy(&dx0, h->d_dx + h->ngx, sizeof(double), hipMemcpyDeviceToHost));
    A.alpha = ip.alphaCoeff * dx0 * dx0;
    A.rhs = h->igrRhs;
    A.diag = h->igrDiag;
    A.rho = h->igrRho;
    IgrPreFn<T> pre = igr_pre_kernel_for<T>(h->dim, h->nph);
    if (!pre) return fail(SIFV_ERR_UNSUPPORTED, "no IGR kernel for this configuration");
    void* pargs[] = {&A};
    int rc = timed_launch(h, reinterpret_cast<const void*>(pre), dim3(elementwise_blocks(h)), dim3(256), pargs);
    if (rc) return rc;
    double* f[1];
    int comp[1] = {0};
    f[0] = h->igrRho;
    rc = fill_dirs(h, f, comp, 1, true);
    if (rc) return rc;
    // 1-deep sigma ghosts as the stage-start fill leaves them
    f[0] = h->sigma;
    rc = fill_dirs(h, f, comp, 1, true);
    if (rc) return rc;
    if (h->igrStrips > 0) {
        const IgrStripKernels K = igr_strip_kernels();
        IgrStripPackArgs PA;
        PA.g = h->grid();
        PA.sigma = h->sigma;
        PA.rho = h->igrRho;
        PA.rhs = h->igrRhs;
        PA.diag = h->igrDiag;
        PA.alpha = A.alpha;
        PA.rec = h->igrRec;
        PA.edgeL = h->igrEdge;
        PA.edgeR = h->igrEdge + 2 * (size_t)h->igrStrips * h->ny;
        PA.W = h->igrStrips;
        void* pargs2[] = {&PA};
        rc = timed_launch(h, reinterpret_cast<const void*>(K.pack), dim3(elementwise_blocks(h)), dim3(256), pargs2);
        if (rc) return rc;
        IgrStripArgs SA;
        SA.nx = h->nx;
        SA.ny = h->ny;
        SA.rec = h->igrRec;
        SA.edgeL = PA.edgeL;
        SA.edgeR = PA.edgeR;
        SA.prog = h->igrProg;
        SA.err = h->igrFlags + h->igrTiles;
        SA.active = active;
        SA.W = h->igrStrips;
        SA.P = h->igrPeriod;
        SA.nsweeps = iters;
        SA.perX = h->sbc[0] == BK_WRAP;
        SA.perY = h->sbc[2] == BK_WRAP;
        HIPCHK(hipMemsetAsync(h->igrProg, 0, sizeof(int) * h->igrStrips, h->stream));
        void* sargs2[] = {&SA};
        rc = timed_launch(h, reinterpret_cast<const void*>(K.sweep), dim3(h->igrStrips), dim3(2 * kWave), sargs2);
        if (rc) return rc;
        Grid g = h->grid();
        const double* recc = h->igrRec;
        double* sig = h->sigma;
        void* uargs[] = {&g, &recc, &sig};
        rc = timed_launch(h, reinterpret_cast<const void*>(K.unpack), dim3(elementwise_blocks(h)), dim3(256), uargs);
        if (rc) return rc;
        return fill_dirs(h, f, comp, 1, true);
    }
    IgrSweepArgs S;
    S.g = h->grid();
    S.sg = h->sigma;
    S.rho = h->igrRho;
    S.rhs = h->igrRhs;
    S.diag = h->igrDiag;
    S.alpha = A.alpha;
    S.ntx = h->igrNtx;
    S.nty = h->igrNty;
    S.tj = h->igrTj;
    S.ntiles = h->igrTiles;
    S.order = h->igrOrder;
    S.flags = h->igrFlags;
    S.err = h->igrFlags + h->igrTiles;
    S.next = h->igrFlags + h->igrTiles + 1;
    S.wrapX = h->igrWrap;
    S.wrapY = S.wrapX + 2 * (size_t)h->ny * h->nz;
    S.wrapZ = S.wrapY + 2 * (size_t)h->nx * h->nz;
    S.active = active;
    IgrSweepFn sweep = igr_sweep_kernel_for(h->dim);
    const int nblk = h->igrBlocks;
    // all sweeps in one launch; the kernel keeps the 1-deep ghosts as the
    // fill after every sweep would, the full fill follows the last sweep
    S.epoch = h->igrEpoch;
    S.nsweeps = iters;
    h->igrEpoch += iters;
#if SIFV_IGR_DYN
    HIPCHK(hipMemsetAsync(S.next, 0, sizeof(int), h->stream));
#endif
    void* sargs[] = {&S};
    rc = timed_launch(h, reinterpret_cast<const void*>(sweep), dim3(nblk), dim3(kWave), sargs);
    if (rc) return rc;
    return fill_dirs(h, f, comp, 1, true);
}

int run_igr_any(sifv_handle* h, int src, const int* active) {
    if (!h->cfg.useIGR) return SIFV_OK;
    return h->precision == SIFV_MIXED ? run_igr<float>(h, src, active) : run_igr<double>(h, src, active);
}

// The IGR wavefront raises its error flag only when a wait ran out of budget.
int check_igr(sifv_handle* h) {
    if (!h->cfg.useIGR) return SIFV_OK;
    int e = 0;
    HIPCHK(hipMemcpy(&e, h->igrFlags + h->igrTiles, sizeof(int), hipMemcpyDeviceToHost));
    if (e) return fail(SIFV_ERR_HIP, "IGR wavefront wait exceeded its budget");
    return SIFV_OK;
}

// ---- discrete adjoint ------------------------------------------------------
int adj_check(const sifv_handle* h) {
    const sifv_config& c = h->cfg;
    if (h->dim != 2 || !(h->bc[0] == BK_WRAP && h->bc[2] == BK_WRAP))
        return fail(SIFV_ERR_UNSUPPORTED, "the adjoint covers doubly periodic 2D meshes");
    if (c.reconOrder != SIFV_WENO5 || h->riemann != SIFV_RUSANOV || c.RKOrder != 3)
        return fail(SIFV_ERR_UNSUPPORTED, "the adjoint covers WENO5 + Rusanov + SSP-RK3");
    if (h->nph != 0 || h->eos.kind != SIFV_EOS_IDEAL || h->precision != SIFV_FP64 || c.useIGR || c.semiImplicit)
        return fail(SIFV_ERR_UNSUPPORTED, "the adjoint covers the single-phase ideal-gas FP64 explicit path");
    if (!(c.explicitParams.constDt > 0.0))
        return fail(SIFV_ERR_UNSUPPORTED, "the adjoint needs a fixed time step (explicitParams.constDt > 0)");
    for (int d = 0; d < 3; ++d)
        if (c.bodyForceParams.a[d] != 0.0 || c.bodyForceParams.b[d] != 0.0)
            return fail(SIFV_ERR_UNSUPPORTED, "the adjoint covers the flux terms only (no body force)");
    if (has_face_stress(c))
        return fail(SIFV_ERR_UNSUPPORTED, "the adjoint covers the flux terms only (no viscous or capillary stress)");
    return SIFV_OK;
}

double* adj_slot(sifv_handle* h, int k) { return h->adjWork + (size_t)k * 4 * h->total; }

// vjp of R at U against mu: out = sum_m a_m X_m + s * (dR/dU)^T mu, gbar += s * d(mu.R)/dgamma
int adj_vjp(sifv_handle* h, const double* U, const double* mu, double s, double* out, const double* const* X,
            const double* a, int nX) {
    const size_t T = h->total;
    AdjArgs A;
    A.g = h->grid();
    A.total = T;
    A.gamma = h->eos.gamma;
    A.eps = h->cfg.wenoEps;
    A.U = U;
    A.mu = mu;
    A.W = h->adjWork;
    A.pbP = h->adjWork + 4 * T;
    A.pbM = h->adjWork + 12 * T;
    A.contrib = h->adjWork + 20 * T;
    A.gpart = h->adjGpart;
    A.s = s;
    A.out = out;
    for (int m = 0; m < 3; ++m) {
        A.X[m] = m < nX ? X[m] : nullptr;
        A.a[m] = m < nX ? a[m] : 0.0;
    }
    A.nX = nX;
    const int nb = h->adjNb;
    A.gpartOff = 0;
    hipLaunchKernelGGL(adj_prim_kernel, dim3(nb), dim3(256), 0, h->stream, A);
    hipLaunchKernelGGL(adj_face_kernel<0>, dim3(nb), dim3(256), 0, h->stream, A);
    A.gpartOff = nb;
    hipLaunchKernelGGL(adj_face_kernel<1>, dim3(nb), dim3(256), 0, h->stream, A);
    hipLaunchKernelGGL(adj_recon_kernel, dim3(nb), dim3(256), 0, h->stream, A);
    A.gpartOff = 2 * nb;
    hipLaunchKernelGGL(adj_gather_kernel, dim3(nb), dim3(256), 0, h->stream, A);
    hipLaunchKernelGGL(adj_gsum_kernel, dim3(1), dim3(256), 0, h->stream, (const double*)h->adjGpart, 3 * nb,
                       h->adjGbar);
    HIPCHK(hipGetLastError());
    return SIFV_OK;
}

// One forward RK3 step from `un` into `next` on explicit pointers, keeping the
// stage states in the adjoint work slots (U1 = slot 15, U2 = slot 16).
int adj_stages(sifv_handle* h, const double* un, double* next) {
    double* U1 = adj_slot(h, 15);
    double* U2 = adj_slot(h, 16);
    int rc = launch_stage_p<double>(h, un, nullptr, U1, kRK[2][0], false, false, nullptr, false);
    if (!rc) rc = launch_stage_p<double>(h, U1, un, U2, kRK[2][1], false, false, nullptr, false);
    if (!rc && next) rc = launch_stage_p<double>(h, U2, un, next, kRK[2][2], false, false, nullptr, false);
    return rc;
}

void drop_step_graph(sifv_handle* h) {
    if (h->stepExec) hipGraphExecDestroy(h->stepExec);
    if (h->stepGraph) hipGraphDestroy(h->stepGraph);
    h->stepExec = nullptr;
    h->stepGraph = nullptr;
}

// Whether a step of sifv_explicit_advance can be replayed from a graph: the
// launch arguments of a step are then the same every step (RK2/RK3 keep the
// state in buffer A; RK1 rotates it), no dt_cells pass is due, no per-launch
// events are recorded, and IGR (whose warm start reads the step counter and
// whose sweeps carry epochs) is off.
bool step_graph_usable(const sifv_handle* h) {
    return h->useGraphs && !h->noalloc && !h->timing.on && h->cfg.RKOrder >= 2 && !h->cfg.useIGR &&
           (h->cfg.explicitParams.constDt > 0.0 || h->dtNextValid);
}

// Enqueues one explicit RK step of the device-resident time loop through the
// step graph, capturing it first when none matches the current arguments.
// The captured work is exactly enqueue_explicit_step's launches (no memset
// or copy nodes).
int launch_step_graph(sifv_handle* h, double endTime) {
    if (!h->stepExec || h->graphEnd != endTime || h->graphCur != h->cur || h->graphRows != h->rowsOverride ||
        h->graphNpart != h->npartialValid) {
        drop_step_graph(h);
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
        int rc = enqueue_explicit_step(h, -1.0, true, endTime);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(h->stream, &gr);
        if (rc) {
            if (gr) hipGraphDestroy(gr);
            return rc;
        }
        HIPCHK(e);
        h->stepGraph = gr;
        HIPCHK(hipGraphInstantiate(&h->stepExec, gr, nullptr, nullptr, 0));
        h->graphEnd = endTime;
        h->graphCur = h->cur;
        h->graphRows = h->rowsOverride;
        h->graphNpart = h->npartialValid;
    }
    HIPCHK(hipGraphLaunch(h->stepExec, h->stream));
    ++h->graphReplays;
    h->dtNextValid = h->cfg.explicitParams.constDt <= 0.0;
    h->primFresh = false;
    return SIFV_OK;
}

int check_step_config(sifv_handle* h, const sifv_config* cfg) {
    if (cfg) {
        // per-call fields of SimulationConfig (step/time, and parameters the
        // reference reads on every call)
        h->cfg.step = cfg->step;
        h->cfg.time = cfg->time;
    }
    return SIFV_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* sifv_last_error(void) { return g_err.c_str(); }
int sifv_abi_version(void) { return SIFV_ABI_VERSION; }

int sifv_create(const sifv_mesh* mesh, const sifv_config* cfg, int riemann, const sifv_eos* eos,
                int precision, sifv_handle** out) {
    if (!mesh || !cfg || !eos || !out) return fail(SIFV_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = validate_config(cfg);
    if (rc) return rc;
    if (mesh->dim != cfg->dim) return fail(SIFV_ERR_INVALID, "mesh.dim != config.dim");
    if (mesh->nGhost != cfg->nGhost) return fail(SIFV_ERR_INVALID, "mesh.nGhost != config.nGhost");
    if (mesh->nx < 1 || (mesh->dim >= 2 && mesh->ny < 1) || (mesh->dim >= 3 && mesh->nz < 1))
        return fail(SIFV_ERR_INVALID, "mesh needs at least one cell per active direction");
    if (!mesh->xNodes || (mesh->dim >= 2 && !mesh->yNodes) || (mesh->dim >= 3 && !mesh->zNodes))
        return fail(SIFV_ERR_INVALID, "missing node array");
    if (riemann < SIFV_LF || riemann > SIFV_HLLC) return fail(SIFV_ERR_INVALID, "unknown Riemann solver");
    if (precision != SIFV_FP64 && precision != SIFV_MIXED) return fail(SIFV_ERR_INVALID, "unknown precision");
    if (eos->kind != SIFV_EOS_IDEAL && eos->kind != SIFV_EOS_STIFFENED) return fail(SIFV_ERR_INVALID, "unknown EOS");
    for (int f = 0; f < 6; ++f)
        if (mesh->bc[f] < 0 || mesh->bc[f] > 4) return fail(SIFV_ERR_INVALID, "unknown boundary condition");
    if (cfg->multiPhaseParams.nPhases != 0 && cfg->multiPhaseParams.nPhases != 2)
        return fail(SIFV_ERR_UNSUPPORTED, "the GPU path carries single-phase or two-phase states");

    auto* h = new sifv_handle();
    h->dim = mesh->dim;
    h->nx = mesh->nx;
    h->ny = mesh->dim >= 2 ? mesh->ny : 1;
    h->nz = mesh->dim >= 3 ? mesh->nz : 1;
    h->ng = mesh->nGhost;
    h->ngx = h->ng;
    h->ngy = h->dim >= 2 ? h->ng : 0;
    h->ngz = h->dim >= 3 ? h->ng : 0;
    h->nxT = h->nx + 2 * h->ngx;
    h->nyT = h->ny + 2 * h->ngy;
    h->nzT = h->nz + 2 * h->ngz;
    h->total = (size_t)h->nxT * h->nyT * h->nzT;
    h->cfg = *cfg;
    h->riemann = riemann;
    h->eos = *eos;
    // semi-implicit + SIFV_MIXED: FP64 state, Jacobi sweeps in FP32 with FP64 refinement
    h->jacMixed = cfg->semiImplicit && precision == SIFV_MIXED;
    h->precision = h->jacMixed ? SIFV_FP64 : precision;
    h->nph = cfg->multiPhaseParams.nPhases;
    h->nc = h->dim + 2 + 2 * h->nph;
    h->eoff = h->precision == SIFV_MIXED && (h->nph > 0 || eos->kind == SIFV_EOS_STIFFENED);
    const int nn[3] = {h->nx, h->ny, h->nz};
    for (int d = 0; d < 3; ++d) h->periods[d] = d < h->dim && mesh->periods[d] != 0;
    for (int f = 0; f < 6; ++f) {
        int d = f / 2;
        bool periodic = d < h->dim && mesh->periods[d] != 0;
        int b = periodic ? SIFV_BC_PERIODIC : mesh->bc[f];
        switch (b) {
            case SIFV_BC_PERIODIC: h->bc[f] = BK_WRAP; break;
            case SIFV_BC_SYMMETRY: h->bc[f] = BK_MIRROR; break;
            case SIFV_BC_SLIPWALL: h->bc[f] = BK_SLIP; break;
            case SIFV_BC_NOSLIPWALL: h->bc[f] = BK_NOSLIP; break;
            default: h->bc[f] = BK_OUTFLOW; break;
        }
        h->sbc[f] = h->bc[f] == BK_WRAP ? BK_WRAP : BK_OUTFLOW;
        if (d < h->dim && h->bc[f] != BK_OUTFLOW && nn[d] < h->ng) {
            delete h;
            return fail(SIFV_ERR_INVALID, "mirror/periodic ghost rules need at least nGhost cells per direction");
        }
    }
    static const double unit[2] = {0.0, 1.0};
    h->nodes[0].assign(mesh->xNodes, mesh->xNodes + h->nx + 1);
    h->nodes[1] = h->dim >= 2 ? std::vector<double>(mesh->yNodes, mesh->yNodes + h->ny + 1) : std::vector<double>(unit, unit + 2);
    h->nodes[2] = h->dim >= 3 ? std::vector<double>(mesh->zNodes, mesh->zNodes + h->nz + 1) : std::vector<double>(unit, unit + 2);
    std::vector<double> xe = extended_nodes(mesh->xNodes, h->nx, h->ngx);
    std::vector<double> ye = extended_nodes(h->dim >= 2 ? mesh->yNodes : unit, h->ny, h->ngy);
    std::vector<double> ze = extended_nodes(h->dim >= 3 ? mesh->zNodes : unit, h->nz, h->ngz);
    std::vector<double> dx(h->nxT), dy(h->nyT), dz(h->nzT);
    for (int i = 0; i < h->nxT; ++i) dx[i] = xe[i + 1] - xe[i];
    for (int j = 0; j < h->nyT; ++j) dy[j] = ye[j + 1] - ye[j];
    for (int k = 0; k < h->nzT; ++k) dz[k] = ze[k + 1] - ze[k];
    // PCG needs the symmetric form: a periodic axis's first ghost widths must
    // match the wrapped cells' widths up to the rounding of the nodes
    {
        const std::vector<double>* w[3] = {&dx, &dy, &dz};
        const int n[3] = {h->nx, h->ny, h->nz}, ng[3] = {h->ngx, h->ngy, h->ngz};
        for (int d = 0; d < h->dim; ++d) {
            if (h->bc[2 * d] != BK_WRAP) continue;
            const double lo = (*w[d])[ng[d] - 1], last = (*w[d])[ng[d] + n[d] - 1];
            const double hi = (*w[d])[ng[d] + n[d]], first = (*w[d])[ng[d]];
            if (std::fabs(lo - last) > 1e-10 * last || std::fabs(hi - first) > 1e-10 * first) h->pcgSym = false;
        }
    }
    h->unifY = true;
    for (int j = 1; j < h->ny; ++j) h->unifY = h->unifY && dy[h->ngy + j] == dy[h->ngy];
    for (int k = 1; k < h->nz; ++k) h->unifY = h->unifY && dz[h->ngz + k] == dz[h->ngz];

    auto bail = [&](int code) {
        sifv_destroy(h);
        return code;
    };
#define CK(expr)                                                                                   \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return bail(fail(SIFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); \
    } while (0)
    CK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    if (h->dim == 3) {
        CK(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
        for (int k = 0; k < kMaxSlabs; ++k) CK(hipEventCreateWithFlags(&h->slabEv[k], hipEventDisableTiming));
        CK(hipEventCreateWithFlags(&h->joinEv, hipEventDisableTiming));
    }
    CK(hipMalloc(&h->d_dx, sizeof(double) * h->nxT));
    CK(hipMalloc(&h->d_dy, sizeof(double) * h->nyT));
    CK(hipMalloc(&h->d_dz, sizeof(double) * h->nzT));
    CK(hipMemcpy(h->d_dx, dx.data(), sizeof(double) * h->nxT, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->d_dy, dy.data(), sizeof(double) * h->nyT, hipMemcpyHostToDevice));
    CK(hipMemcpy(h->d_dz, dz.data(), sizeof(double) * h->nzT, hipMemcpyHostToDevice));
    size_t cb = h->tsize() * h->total * h->nc;
    for (int b = 0; b < 3; ++b) {
        CK(hipMalloc(&h->cons[b], cb));
        CK(hipMemset(h->cons[b], 0, cb));
    }
    CK(hipMalloc(&h->rhoP, sizeof(double) * h->total));
    CK(hipMemset(h->rhoP, 0, sizeof(double) * h->total));
    for (int d = 0; d < h->dim; ++d) {
        CK(hipMalloc(&h->vel[d], sizeof(double) * h->total));
        CK(hipMemset(h->vel[d], 0, sizeof(double) * h->total));
    }
    CK(hipMalloc(&h->pres, sizeof(double) * h->total));
    CK(hipMemset(h->pres, 0, sizeof(double) * h->total));
    CK(hipMalloc(&h->sigma, sizeof(double) * h->total));
    CK(hipMemset(h->sigma, 0, sizeof(double) * h->total));
    CK(hipMalloc(&h->scal, sizeof(double) * 8));
    CK(hipMemset(h->scal, 0, sizeof(double) * 8));
    CK(hipMalloc(&h->flags, sizeof(int) * 4));
    CK(hipMemset(h->flags, 0, sizeof(int) * 4));
    CK(hipHostMalloc(&h->hostPinned, sizeof(double) * 4, hipHostMallocDefault));
    if (h->dim == 3) {
        h->nfc = h->dim + 2 + 2 * h->nph + 2;
        CK(hipMalloc(&h->part, h->tsize() * h->total * h->nfc));
    }
    // buffers the time loop would otherwise allocate on first use
    if (has_face_stress(*cfg)) {
        CK(hipMalloc(&h->visc, sizeof(double) * (h->dim + 1) * h->total));
        CK(hipMemset(h->visc, 0, sizeof(double) * (h->dim + 1) * h->total));
    }
    if (h->eoff) CK(hipMalloc(&h->ePend, sizeof(double) * h->total));
    CK(hipGetDevice(&h->device));
#undef CK
    if (cfg->useIGR) {
        rc = setup_igr(h);
        if (rc) {
            sifv_destroy(h);
            return rc;
        }
    }
    h->dtNextValid = false;
    h->primFresh = true;
    *out = h;
    return SIFV_OK;
}

int sifv_destroy(sifv_handle* h) {
    if (!h) return SIFV_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->side) hipStreamSynchronize(h->side);
    drop_step_graph(h);
    drop_jac_graphs(h);
    for (auto& pr : h->timing.pending) {
        hipEventDestroy(pr.first);
        hipEventDestroy(pr.second);
    }
    hipFree(h->d_dx);
    hipFree(h->d_dy);
    hipFree(h->d_dz);
    for (int b = 0; b < 3; ++b) hipFree(h->cons[b]);
    hipFree(h->rhoP);
    hipFree(h->ePend);
    for (int d = 0; d < 3; ++d) hipFree(h->vel[d]);
    hipFree(h->pres);
    hipFree(h->sigma);
    hipFree(h->partial);
    hipFree(h->scal);
    hipFree(h->flags);
    hipFree(h->part);
    hipFree(h->visc);
    for (int q = 0; q < 7; ++q) hipFree(h->pcgv[q]);
    for (int d = 0; d < 3; ++d) hipFree(h->pcgW[d]);
    hipFree(h->mgBuf);
    hipFree(h->mgBuf32);
    for (int q = 0; q < 6; ++q) hipFree(h->pcg32[q]);
    hipFree(h->pcgSc);
    hipFree(h->pcgSt);
    if (h->pcgStHost) hipHostFree(h->pcgStHost);
    hipFree(h->jacPn);
    hipFree(h->jac32);
    for (double* c : h->adjCk) hipFree(c);
    for (double* c : h->adjSeg) hipFree(c);
    hipFree(h->adjWork);
    hipFree(h->adjGpart);
    hipFree(h->adjGbar);
    hipFree(h->igrRhs);
    hipFree(h->igrDiag);
    hipFree(h->igrRho);
    hipFree(h->igrOrder);
    hipFree(h->igrWrap);
    hipFree(h->igrRec);
    hipFree(h->igrEdge);
    hipFree(h->igrProg);
    hipFree(h->igrFlags);
    for (int q = 0; q < 6; ++q) hipFree(h->jac[q]);
    hipFree(h->resid);
    if (h->hostResid) hipHostFree(h->hostResid);
    if (h->hostPinned) hipHostFree(h->hostPinned);
    if (h->stream) hipStreamDestroy(h->stream);
    if (h->side) hipStreamDestroy(h->side);
    for (int k = 0; k < kMaxSlabs; ++k)
        if (h->slabEv[k]) hipEventDestroy(h->slabEv[k]);
    if (h->joinEv) hipEventDestroy(h->joinEv);
    delete h;
    return SIFV_OK;
}

int sifv_total_cells(const sifv_handle* h, size_t* n) {
    if (!h || !n) return fail(SIFV_ERR_INVALID, "null argument");
    *n = h->total;
    return SIFV_OK;
}

namespace {
// Resolve a field id to (cons buffer slot q, or prim array); returns -1 on
// an unknown/absent field.
int cons_slot(const sifv_handle* h, int field) {
    if (field == SIFV_F_RHO) return 0;
    if (field >= SIFV_F_RHOU && field <= SIFV_F_RHOW) {
        int d = field - SIFV_F_RHOU;
        return d < h->dim ? 1 + d : -1;
    }
    if (field == SIFV_F_RHOE) return h->dim + 1;
    if (field >= SIFV_F_ALPHARHO && field < SIFV_F_ALPHARHO + SIFV_MAX_PHASES) {
        int p = field - SIFV_F_ALPHARHO;
        return p < h->nph ? h->dim + 2 + p : -1;
    }
    if (field >= SIFV_F_ALPHA && field < SIFV_F_ALPHA + SIFV_MAX_PHASES) {
        int p = field - SIFV_F_ALPHA;
        return p < h->nph ? h->dim + 2 + h->nph + p : -1;
    }
    return -1;
}
double* prim_ptr(sifv_handle* h, int field) {
    if (field >= SIFV_F_VELU && field <= SIFV_F_VELW) {
        int d = field - SIFV_F_VELU;
        return d < h->dim ? h->vel[d] : nullptr;
    }
    if (field == SIFV_F_PRES) return h->pres;
    if (field == SIFV_F_SIGMA) return h->sigma;
    return nullptr;
}
}  // namespace

int sifv_upload(sifv_handle* h, int field, const double* host, size_t n) {
    if (!h || !host) return fail(SIFV_ERR_INVALID, "null argument");
    if (n != h->total) return fail(SIFV_ERR_INVALID, "array length != totalCells");
    HIPCHK(hipStreamSynchronize(h->stream));
    int q = cons_slot(h, field);
    if (h->eoff && field == SIFV_F_RHOE) {
        // total energy: kept in FP64 until stored relative to the volume fractions
        if (!h->ePend) HIPCHK(hipMalloc(&h->ePend, sizeof(double) * h->total));
        HIPCHK(hipMemcpy(h->ePend, host, n * sizeof(double), hipMemcpyHostToDevice));
        h->ePending = true;
        h->dtNextValid = false;
        return SIFV_OK;
    }
    if (h->eoff && h->nph > 0 && !h->ePending && field >= SIFV_F_ALPHA && field < SIFV_F_ALPHA + SIFV_MAX_PHASES) {
        // new volume fractions keep the total energy: restate it first
        int rc = energy_convert(h, false);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        h->ePending = true;
    }
    if (q >= 0) {
        if (h->precision == SIFV_MIXED) {
            std::vector<float> tmp(host, host + n);
            HIPCHK(hipMemcpy(h->consField<float>(h->cur, q), tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemcpy(h->consField<double>(h->cur, q), host, n * sizeof(double), hipMemcpyHostToDevice));
        }
        if (field == SIFV_F_RHO) HIPCHK(hipMemcpy(h->rhoP, host, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        double* p = prim_ptr(h, field);
        if (!p) return fail(SIFV_ERR_INVALID, "field not present for this configuration");
        if (field == SIFV_F_SIGMA) h->sigmaUploaded = true;
        HIPCHK(hipMemcpy(p, host, n * sizeof(double), hipMemcpyHostToDevice));
    }
    h->dtNextValid = false;
    return SIFV_OK;
}

int sifv_download(sifv_handle* h, int field, double* host, size_t n) {
    if (!h || !host) return fail(SIFV_ERR_INVALID, "null argument");
    if (n != h->total) return fail(SIFV_ERR_INVALID, "array length != totalCells");
    int rc = refresh_prim(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    int q = cons_slot(h, field);
    if (field == SIFV_F_RHO) {
        HIPCHK(hipMemcpy(host, h->rhoP, n * sizeof(double), hipMemcpyDeviceToHost));
    } else if (h->eoff && field == SIFV_F_RHOE) {
        // stored energy + stiffened-gas offset, in FP64
        rc = energy_convert(h, false);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(host, h->ePend, n * sizeof(double), hipMemcpyDeviceToHost));
    } else if (q >= 0) {
        if (h->precision == SIFV_MIXED) {
            std::vector<float> tmp(n);
            HIPCHK(hipMemcpy(tmp.data(), h->consField<float>(h->cur, q), n * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t t = 0; t < n; ++t) host[t] = tmp[t];
        } else {
            HIPCHK(hipMemcpy(host, h->consField<double>(h->cur, q), n * sizeof(double), hipMemcpyDeviceToHost));
        }
    } else {
        double* p = prim_ptr(h, field);
        if (!p) return fail(SIFV_ERR_INVALID, "field not present for this configuration");
        HIPCHK(hipMemcpy(host, p, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    return SIFV_OK;
}

int sifv_explicit_step(sifv_handle* h, const sifv_config* cfg, double targetDt, double* dt_out) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (h->cfg.semiImplicit) return fail(SIFV_ERR_INVALID, "handle was created for the semi-implicit solver");
    check_step_config(h, cfg);
    int rc = settle_energy(h);
    if (rc) return rc;
    rc = enqueue_explicit_step(h, targetDt, false, 0.0);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->hostPinned, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    rc = harvest_timing(h);
    if (rc) return rc;
    rc = check_igr(h);
    if (rc) return rc;
    if (dt_out) *dt_out = h->hostPinned[0];
    if (!std::isfinite(h->hostPinned[0])) return fail(SIFV_ERR_NONFINITE, "non-finite time step");
    return SIFV_OK;
}

namespace {
// Body of sifv_explicit_advance (shared with the allocation-free form).
int advance_impl(sifv_handle* h, const sifv_config* cfg, int nsteps, double endTime, double* time_inout,
                 int* steps_done) {
    check_step_config(h, cfg);
    int src = settle_energy(h);
    if (src) return src;
    hipLaunchKernelGGL(loop_reset_kernel, dim3(1), dim3(1), 0, h->stream, h->scal, h->flags, *time_inout);
    HIPCHK(hipGetLastError());
    // runTimeLoop advances config.step once per step (src/RKTimeStepping.cpp:266);
    // only the IGR warm start reads it
    const int step0 = h->cfg.step;
    for (int s = 0; s < nsteps; ++s) {
        h->cfg.step = step0 + s;
        int rc = step_graph_usable(h) ? launch_step_graph(h, endTime) : enqueue_explicit_step(h, -1.0, true, endTime);
        if (rc) {
            h->cfg.step = step0;
            return rc;
        }
    }
    h->cfg.step = step0;
    HIPCHK(hipMemcpyAsync(h->hostPinned, h->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    int* hflags = reinterpret_cast<int*>(h->hostPinned + 2);
    HIPCHK(hipMemcpyAsync(hflags, h->flags, 2 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    int rc = harvest_timing(h);
    if (rc) return rc;
    rc = check_igr(h);
    if (rc) return rc;
    *time_inout = h->hostPinned[1];
    if (steps_done) *steps_done = hflags[1];
    if (!std::isfinite(h->hostPinned[1])) return fail(SIFV_ERR_NONFINITE, "non-finite time");
    return SIFV_OK;
}

// ---- allocation-free time loop ---------------------------------------------
// The dt block partials of sifv_explicit_advance_noalloc live in this array:
// one copy per GPU for the life of the process, shared by every call on that
// GPU (overlapping calls on one GPU race on it).  Nothing has to start at
// zero: every partial the dt finalize reads was written earlier in the same
// step.  Host code uses only the per-device address from hipGetSymbolAddress.
constexpr int kScratchPartials = 65536;  // 512 KiB of FP64 block minima
static __device__ __align__(256) double scratch[kScratchPartials];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on the handle's device (looked up once per device,
// with that device current).
int scratch_for(const sifv_handle* h, double** out) {
    if (h->device < 0 || h->device >= kMaxDevices) return fail(SIFV_ERR_INVALID, "device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != h->device) HIPCHK(hipSetDevice(h->device));
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != h->device) HIPCHK(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(SIFV_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[h->device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[h->device];
    return SIFV_OK;
}

// Largest number of dt partials one step of this handle writes.
int partials_needed(const sifv_handle* h) {
    int need = elementwise_blocks(h);  // dt from the primitive arrays
    int nstrip, nseg, rows;
    stage_shape(h, nstrip, nseg, rows);
    int stage = nstrip * nseg;
    if (h->dim == 3) {
        int zs, zn, zr;
        zpass_shape(h, zs, zn, zr);
        stage = zs * zn * h->ny;
    }
    const bool stress = h->cfg.viscousMu > 0.0 || has_surface_tension(h->cfg);
    const size_t ncell = (size_t)h->nx * h->ny * h->nz;
    const int nbs = stress ? (int)std::min<size_t>((ncell + kRedThreads - 1) / kRedThreads, 2048) : 0;
    return std::max(need, stage + nbs);
}
}  // namespace

int sifv_explicit_advance(sifv_handle* h, const sifv_config* cfg, int nsteps, double endTime,
                          double* time_inout, int* steps_done) {
    if (!h || !time_inout) return fail(SIFV_ERR_INVALID, "null argument");
    if (h->cfg.semiImplicit) return fail(SIFV_ERR_INVALID, "handle was created for the semi-implicit solver");
    if (nsteps < 0) return fail(SIFV_ERR_INVALID, "nsteps < 0");
    return advance_impl(h, cfg, nsteps, endTime, time_inout, steps_done);
}

int sifv_explicit_advance_noalloc(sifv_handle* h, const sifv_config* cfg, int nsteps, double endTime,
                                  double* time_inout, int* steps_done) {
    if (!h || !time_inout) return fail(SIFV_ERR_INVALID, "null argument");
    if (h->cfg.semiImplicit) return fail(SIFV_ERR_INVALID, "handle was created for the semi-implicit solver");
    if (nsteps < 0) return fail(SIFV_ERR_INVALID, "nsteps < 0");
    if (h->timing.on) return fail(SIFV_ERR_UNSUPPORTED, "kernel timing creates events: switch it off for this entry point");
    const int need = partials_needed(h);
    if (need > kScratchPartials)
        return fail(SIFV_ERR_INVALID, "mesh needs " + std::to_string(need) + " dt partials per step; the allocation-free "
                                      "entry point holds at most " + std::to_string(kScratchPartials));
    double* sp = nullptr;
    int rc = scratch_for(h, &sp);
    if (rc) return rc;
    double* ownP = h->partial;
    const int ownCap = h->npartialCap;
    h->noalloc = true;
    // the next step's dt partials of an earlier call sit in the handle's own
    // buffer: start from the primitive arrays instead (same cell minima)
    if (h->dtNextValid) {
        rc = refresh_prim(h);
        h->dtNextValid = false;
    }
    h->partial = sp;
    h->npartialCap = kScratchPartials;
    h->npartialValid = 0;
    if (!rc) rc = advance_impl(h, cfg, nsteps, endTime, time_inout, steps_done);
    // leave no reference to the shared array behind: the next call on this
    // handle takes its dt from the refreshed primitive arrays
    if (!rc) rc = refresh_prim(h);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(SIFV_ERR_HIP, "stream synchronize");
    h->dtNextValid = false;
    h->partial = ownP;
    h->npartialCap = ownCap;
    h->npartialValid = 0;
    h->noalloc = false;
    return rc;
}

int sifv_fill_ghosts(sifv_handle* h, int varset) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (h->precision != SIFV_FP64) return fail(SIFV_ERR_UNSUPPORTED, "ghost fill API needs an FP64 handle");
    if (varset < SIFV_VARSET_CONS || varset > SIFV_VARSET_ALL) return fail(SIFV_ERR_INVALID, "unknown varset");
    // The conservative buffer holds the authoritative density; sync the
    // primitive-side copy first so both sides hold the same values.
    HIPCHK(hipMemcpyAsync(h->rhoP, h->consField<double>(h->cur, 0), h->total * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
    double* f[kMaxFill];
    int comp[kMaxFill];
    int n = 0;
    f[n] = h->rhoP; comp[n++] = 0;
    for (int d = 0; d < h->dim; ++d) {
        if (varset != SIFV_VARSET_PRIM) { f[n] = h->consField<double>(h->cur, 1 + d); comp[n++] = 1 + d; }
        if (varset != SIFV_VARSET_CONS) { f[n] = h->vel[d]; comp[n++] = 1 + d; }
    }
    if (varset != SIFV_VARSET_PRIM) { f[n] = h->consField<double>(h->cur, h->dim + 1); comp[n++] = 0; }
    if (varset != SIFV_VARSET_CONS) { f[n] = h->pres; comp[n++] = 0; }
    f[n] = h->sigma; comp[n++] = 0;
    for (int q = h->dim + 2; q < h->nc; ++q) { f[n] = h->consField<double>(h->cur, q); comp[n++] = 0; }
    int rc = fill_dirs(h, f, comp, n, false, varset == SIFV_VARSET_CONS ? h->sigma : nullptr);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->consField<double>(h->cur, 0), h->rhoP, h->total * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SIFV_OK;
}

int sifv_fill_scalar_ghosts(sifv_handle* h, int field) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    double* p = nullptr;
    if (field == SIFV_F_RHO) p = h->rhoP;
    else p = prim_ptr(h, field);
    if (!p) return fail(SIFV_ERR_INVALID, "scalar fill supports rho, vel*, pres, sigma");
    double* f[1] = {p};
    int comp[1] = {0};
    int rc = fill_dirs(h, f, comp, 1, true);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return SIFV_OK;
}

int sifv_acoustic_dt(sifv_handle* h, double cfl, double maxDt, double* dt) {
    if (!h || !dt) return fail(SIFV_ERR_INVALID, "null argument");
    int rc = launch_dt_cells(h, false);
    if (!rc) rc = launch_finalize(h, cfl, maxDt, 0.0, -1.0, -1.0, false, false, 0.0);
    if (rc) return rc;
    h->dtNextValid = false;
    HIPCHK(hipMemcpyAsync(h->hostPinned, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *dt = h->hostPinned[0];
    return SIFV_OK;
}

int sifv_advective_dt(sifv_handle* h, double cfl, double maxDt, double* dt) {
    if (!h || !dt) return fail(SIFV_ERR_INVALID, "null argument");
    int rc = launch_dt_cells(h, true);
    if (!rc) rc = launch_finalize(h, cfl, maxDt, 0.0, -1.0, -1.0, false, false, 0.0);
    if (rc) return rc;
    h->dtNextValid = false;
    HIPCHK(hipMemcpyAsync(h->hostPinned, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *dt = h->hostPinned[0];
    return SIFV_OK;
}

int sifv_cons_to_prim(sifv_handle* h) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    h->primFresh = false;
    int rc = refresh_prim(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return SIFV_OK;
}

int sifv_timing_enable(sifv_handle* h, int on) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    h->timing.on = on != 0;
    return SIFV_OK;
}

int sifv_timing_read(sifv_handle* h, double* kernel_ms, long* launches, int reset) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    int rc = harvest_timing(h);
    if (rc) return rc;
    if (kernel_ms) *kernel_ms = h->timing.ms;
    if (launches) *launches = h->timing.launches;
    if (reset) {
        h->timing.ms = 0.0;
        h->timing.launches = 0;
    }
    return SIFV_OK;
}

int sifv_synchronize(sifv_handle* h) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->side) HIPCHK(hipStreamSynchronize(h->side));
    return SIFV_OK;
}

int sifv_set_tile(sifv_handle* h, int rows_per_wave) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (rows_per_wave < 0 || rows_per_wave > 4096) return fail(SIFV_ERR_INVALID, "rows_per_wave out of range");
    h->rowsOverride = rows_per_wave;
    return SIFV_OK;
}

int sifv_set_stage_slabs(sifv_handle* h, int slabs) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (slabs < 0 || slabs > kMaxSlabs) return fail(SIFV_ERR_INVALID, "slabs out of range");
    h->slabsOverride = slabs;
    return SIFV_OK;
}

int sifv_set_graphs(sifv_handle* h, int on) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    h->useGraphs = on != 0;
    if (!h->useGraphs) {
        HIPCHK(hipStreamSynchronize(h->stream));
        drop_step_graph(h);
        drop_jac_graphs(h);
    }
    return SIFV_OK;
}

// ---- VTK output of the resident state -------------------------------------
// The reference's ASCII XML writer (VTKWriter::writeVTR / writePVTR / writePVD,
// src/VTKWriter.cpp:21-293) as VTKSession drives it on one rank
// (src/VTKSession.cpp:23-56): the same arrays in the same order, 15
// significant digits through the same stream formatting.

namespace {

void vtk_parent_dir(const std::string& path) {
    auto parent = std::filesystem::path(path).parent_path();
    if (!parent.empty()) std::filesystem::create_directories(parent);
}

}  // namespace

int sifv_write_vtr(sifv_handle* h, const char* filename, int rank) {
    if (!h || !filename) return fail(SIFV_ERR_INVALID, "null argument");
    const size_t n = h->total;
    const int nph = h->nph, dim = h->dim;
    const bool igr = h->cfg.useIGR != 0;
    auto get = [&](int fid, std::vector<double>& v) {
        v.resize(n);
        return sifv_download(h, fid, v.data(), n);
    };
    std::vector<double> pres, sigma, rhoE, rho, vel[3], mom[3];
    std::vector<std::vector<double>> ar(nph), al(nph);
    int rc = get(SIFV_F_PRES, pres);
    if (!rc && igr) rc = get(SIFV_F_SIGMA, sigma);
    if (!rc) rc = get(SIFV_F_RHOE, rhoE);
    for (int d = 0; d < dim && !rc; ++d) {
        rc = get(SIFV_F_VELU + d, vel[d]);
        if (!rc) rc = get(SIFV_F_RHOU + d, mom[d]);
    }
    if (!rc && nph == 0) rc = get(SIFV_F_RHO, rho);
    for (int ph = 0; ph < nph && !rc; ++ph) {
        rc = get(SIFV_F_ALPHARHO + ph, ar[ph]);
        if (!rc) rc = get(SIFV_F_ALPHA + ph, al[ph]);
    }
    if (rc) return rc;
    try {
        vtk_parent_dir(filename);
        std::ofstream file(filename);
        if (!file.is_open()) return fail(SIFV_ERR_INVALID, std::string("cannot open ") + filename);
        const int nx = h->nx, ny = h->ny, nz = h->nz;
        auto at = [&](int i, int j, int k) {
            return (size_t)(i + h->ngx) + (size_t)h->nxT * ((size_t)(j + h->ngy) + (size_t)h->nyT * (size_t)(k + h->ngz));
        };
        const std::vector<double>& X = h->nodes[0];
        const std::vector<double>& Y = h->nodes[1];
        const std::vector<double>& Z = h->nodes[2];
        file << std::setprecision(15);
        file << "<?xml version=\"1.0\"?>\n";
        file << "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n";
        file << "  <RectilinearGrid WholeExtent=\"0 " << nx << " 0 " << ny << " 0 " << nz << "\">\n";
        file << "    <Piece Extent=\"0 " << nx << " 0 " << ny << " 0 " << nz << "\">\n";
        file << "      <Coordinates>\n";
        file << "        <DataArray type=\"Float64\" Name=\"X\" format=\"ascii\">\n         ";
        for (int i = 0; i <= nx; ++i) file << " " << X[i];
        file << "\n        </DataArray>\n";
        file << "        <DataArray type=\"Float64\" Name=\"Y\" format=\"ascii\">\n         ";
        const double dx0 = X[1] - X[0], dy0 = Y[1] - Y[0];
        if (ny == 1) file << " 0.0 " << dx0;
        else
            for (int j = 0; j <= ny; ++j) file << " " << Y[j];
        file << "\n        </DataArray>\n";
        file << "        <DataArray type=\"Float64\" Name=\"Z\" format=\"ascii\">\n         ";
        if (nz == 1) file << " 0.0 " << std::min(dx0, std::min(dy0, 1.0));
        else
            for (int k = 0; k <= nz; ++k) file << " " << Z[k];
        file << "\n        </DataArray>\n";
        file << "      </Coordinates>\n";
        file << "      <CellData>\n";
        auto scalar = [&](const std::string& name, const std::vector<double>& f) {
            file << "        <DataArray type=\"Float64\" Name=\"" << name << "\" format=\"ascii\">\n";
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) {
                    file << "         ";
                    for (int i = 0; i < nx; ++i) file << " " << f[at(i, j, k)];
                    file << "\n";
                }
            file << "        </DataArray>\n";
        };
        auto vector = [&](const std::string& name, const std::vector<double>* f) {
            file << "        <DataArray type=\"Float64\" Name=\"" << name
                 << "\" NumberOfComponents=\"3\" format=\"ascii\">\n";
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) {
                    file << "         ";
                    for (int i = 0; i < nx; ++i) {
                        const size_t q = at(i, j, k);
                        file << " " << f[0][q] << " " << (dim >= 2 ? f[1][q] : 0.0) << " "
                             << (dim >= 3 ? f[2][q] : 0.0);
                    }
                    file << "\n";
                }
            file << "        </DataArray>\n";
        };
        scalar("Pressure", pres);
        if (igr) scalar("Sigma", sigma);
        scalar("TotalEnergy", rhoE);
        vector("Velocity", vel);
        vector("Momentum", mom);
        if (nph == 0) scalar("Density", rho);
        for (int ph = 0; ph < nph; ++ph) scalar("AlphaRho_" + std::to_string(ph), ar[ph]);
        for (int ph = 0; ph < nph; ++ph) scalar("Alpha_" + std::to_string(ph), al[ph]);
        if (rank >= 0) {
            file << "        <DataArray type=\"Int32\" Name=\"Rank\" format=\"ascii\">\n";
            for (int k = 0; k < nz; ++k)
                for (int j = 0; j < ny; ++j) {
                    file << "         ";
                    for (int i = 0; i < nx; ++i) file << " " << rank;
                    file << "\n";
                }
            file << "        </DataArray>\n";
        }
        file << "      </CellData>\n";
        file << "    </Piece>\n";
        file << "  </RectilinearGrid>\n";
        file << "</VTKFile>\n";
        if (!file) return fail(SIFV_ERR_INVALID, std::string("write failed: ") + filename);
    } catch (const std::exception& e) {
        return fail(SIFV_ERR_INVALID, e.what());
    }
    return SIFV_OK;
}

int sifv_write_pvtr(const sifv_handle* h, const char* filename, const char* piece_file) {
    if (!h || !filename || !piece_file) return fail(SIFV_ERR_INVALID, "null argument");
    try {
        vtk_parent_dir(filename);
        std::ofstream file(filename);
        if (!file.is_open()) return fail(SIFV_ERR_INVALID, std::string("cannot open ") + filename);
        file << "<?xml version=\"1.0\"?>\n";
        file << "<VTKFile type=\"PRectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\">\n";
        file << "  <PRectilinearGrid WholeExtent=\"0 " << h->nx << " 0 " << h->ny << " 0 " << h->nz
             << "\" GhostLevel=\"0\">\n";
        file << "    <PCoordinates>\n";
        for (const char* c : {"X", "Y", "Z"}) file << "      <PDataArray type=\"Float64\" Name=\"" << c << "\"/>\n";
        file << "    </PCoordinates>\n";
        file << "    <PCellData>\n";
        auto decl = [&](const std::string& name, bool vec) {
            file << "      <PDataArray type=\"Float64\" Name=\"" << name << "\"" << (vec ? " NumberOfComponents=\"3\"" : "")
                 << "/>\n";
        };
        decl("Pressure", false);
        if (h->cfg.useIGR) decl("Sigma", false);
        decl("TotalEnergy", false);
        decl("Velocity", true);
        decl("Momentum", true);
        const int nph = h->cfg.multiPhaseParams.nPhases;
        if (nph <= 0) decl("Density", false);
        for (int ph = 0; ph < nph; ++ph) decl("AlphaRho_" + std::to_string(ph), false);
        for (int ph = 0; ph < nph; ++ph) decl("Alpha_" + std::to_string(ph), false);
        file << "      <PDataArray type=\"Int32\" Name=\"Rank\"/>\n";
        file << "    </PCellData>\n";
        file << "    <Piece Extent=\"0 " << h->nx << " 0 " << h->ny << " 0 " << h->nz << "\" Source=\"" << piece_file
             << "\"/>\n";
        file << "  </PRectilinearGrid>\n";
        file << "</VTKFile>\n";
        if (!file) return fail(SIFV_ERR_INVALID, std::string("write failed: ") + filename);
    } catch (const std::exception& e) {
        return fail(SIFV_ERR_INVALID, e.what());
    }
    return SIFV_OK;
}

int sifv_write_pvd(const char* filename, const char* mode, double time, const char* data_file) {
    if (!filename || !mode) return fail(SIFV_ERR_INVALID, "null argument");
    static const std::string footer = "  </Collection>\n</VTKFile>\n";
    try {
        vtk_parent_dir(filename);
        const std::string m(mode);
        if (m == "w") {
            std::ofstream file(filename);
            if (!file.is_open()) return fail(SIFV_ERR_INVALID, std::string("cannot open ") + filename);
            file << "<?xml version=\"1.0\"?>\n";
            file << "<VTKFile type=\"Collection\" version=\"1.0\" byte_order=\"LittleEndian\">\n";
            file << "  <Collection>\n";
            file << footer;
        } else if (m == "a") {
            if (!data_file) return fail(SIFV_ERR_INVALID, "null data file");
            const auto size = std::filesystem::file_size(filename);
            if (size < footer.size()) return fail(SIFV_ERR_INVALID, std::string("not a collection file: ") + filename);
            std::filesystem::resize_file(filename, size - footer.size());
            std::ofstream file(filename, std::ios::app);
            if (!file.is_open()) return fail(SIFV_ERR_INVALID, std::string("cannot open ") + filename);
            file << std::setprecision(15);
            file << "    <DataSet timestep=\"" << time << "\" file=\"" << data_file << "\"/>\n";
            file << footer;
        } else if (m != "close") {
            return fail(SIFV_ERR_INVALID, "mode must be w, a or close");
        }
    } catch (const std::exception& e) {
        return fail(SIFV_ERR_INVALID, e.what());
    }
    return SIFV_OK;
}

int sifv_graph_replays(const sifv_handle* h, long* replays) {
    if (!h || !replays) return fail(SIFV_ERR_INVALID, "null argument");
    *replays = h->graphReplays;
    return SIFV_OK;
}

int sifv_semi_implicit_step(sifv_handle* h, const sifv_config* cfg, double targetDt, double* dt_out,
                            int* pressure_iters) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (!h->cfg.semiImplicit) return fail(SIFV_ERR_INVALID, "handle was created for the explicit solver");
    check_step_config(h, cfg);
    int rc = enqueue_semi_step(h, targetDt);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->hostPinned, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    rc = harvest_timing(h);
    if (rc) return rc;
    rc = check_igr(h);
    if (rc) return rc;
    if (dt_out) *dt_out = h->hostPinned[0];
    if (pressure_iters) *pressure_iters = h->lastIters;
    if (!std::isfinite(h->hostPinned[0])) return fail(SIFV_ERR_NONFINITE, "non-finite time step");
    return SIFV_OK;
}

int sifv_jacobi_solve(sifv_handle* h, const double* rho, const double* rhoc2, const double* rhs,
                      double* pressure, double dt, double tolerance, int maxIter, int* iters,
                      double* residual_history) {
    if (!h || !rho || !rhoc2 || !rhs || !pressure) return fail(SIFV_ERR_INVALID, "null argument");
    if (maxIter <= 0) return fail(SIFV_ERR_INVALID, "maxIter must be > 0");
    if (!(tolerance > 0.0)) return fail(SIFV_ERR_INVALID, "tolerance must be > 0");
    int rc = ensure_jacobi(h, maxIter);
    if (rc) return rc;
    const size_t nb = h->total * sizeof(double);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->jac[5], rho, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->jac[2], rhoc2, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->jac[3], rhs, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->jac[0], pressure, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->scal + 3, &dt, sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(h->resid, 0, sizeof(double) * maxIter, h->stream));
    int n = 0;
    rc = run_jacobi(h, h->jac[5], h->scal + 3, tolerance, maxIter, &n);
    if (rc) return rc;
    // the reference's array leaves the loop holding the scalar-ghost fill of
    // the penultimate iterate around the interior of the final one
    // (src/JacobiPressureSolver.cpp:64-104)
    double* prev = h->jac[(n - 1) % 2];
    double* fin = h->jac[n % 2];
    double* f[1] = {prev};
    int comp[1] = {0};
    rc = fill_dirs(h, f, comp, 1, true);
    if (rc) return rc;
    rc = launch_copy_interior(h, fin, prev);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(pressure, prev, nb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (iters) *iters = n;
    if (residual_history)
        for (int t = 0; t < n; ++t) residual_history[t] = h->hostResid[t];
    return SIFV_OK;
}

int sifv_set_pressure_solver(sifv_handle* h, int kind) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (kind != SIFV_PS_JACOBI && kind != SIFV_PS_PCG && kind != SIFV_PS_MGPCG)
        return fail(SIFV_ERR_INVALID, "unknown pressure solver");
    if (kind != SIFV_PS_JACOBI && !h->pcgSym)
        return fail(SIFV_ERR_UNSUPPORTED,
                    "PCG needs a symmetric system: a periodic axis with stretched wrap widths (use Jacobi)");
    if (kind == SIFV_PS_PCG && h->jacMixed)
        return fail(SIFV_ERR_UNSUPPORTED,
                    "the diagonal PCG pressure solve runs in FP64 (no mixed-precision form; use SIFV_PS_MGPCG)");
    h->psolver = kind;
    return SIFV_OK;
}

static int pcg_solve_host(sifv_handle* h, const double* rho, const double* rhoc2, const double* rhs,
                          double* pressure, double dt, double tolerance, int maxIter, int* iters, bool mg) {
    if (!h || !rho || !rhoc2 || !rhs || !pressure) return fail(SIFV_ERR_INVALID, "null argument");
    if (maxIter <= 0) return fail(SIFV_ERR_INVALID, "maxIter must be > 0");
    if (!(tolerance > 0.0)) return fail(SIFV_ERR_INVALID, "tolerance must be > 0");
    int rc = ensure_jacobi(h, 1);
    if (rc) return rc;
    const size_t nb = h->total * sizeof(double);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->jac[5], rho, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->jac[2], rhoc2, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->jac[3], rhs, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->jac[0], pressure, nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->scal + 3, &dt, sizeof(double), hipMemcpyHostToDevice));
    int n = 0;
    rc = run_pcg(h, h->jac[5], h->scal + 3, tolerance, maxIter, &n, mg);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(pressure, h->jac[0], nb, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (iters) *iters = n;
    return SIFV_OK;
}

int sifv_pcg_solve(sifv_handle* h, const double* rho, const double* rhoc2, const double* rhs, double* pressure,
                   double dt, double tolerance, int maxIter, int* iters) {
    return pcg_solve_host(h, rho, rhoc2, rhs, pressure, dt, tolerance, maxIter, iters, false);
}

int sifv_mgpcg_solve(sifv_handle* h, const double* rho, const double* rhoc2, const double* rhs, double* pressure,
                     double dt, double tolerance, int maxIter, int* iters) {
    return pcg_solve_host(h, rho, rhoc2, rhs, pressure, dt, tolerance, maxIter, iters, true);
}

// Runs nsteps explicit steps with the fixed dt from the resident state and
// keeps the state every checkpoint_every steps for the backward sweep.
int sifv_adjoint_forward(sifv_handle* h, const sifv_config* cfg, int nsteps, int checkpoint_every) {
    if (!h) return fail(SIFV_ERR_INVALID, "null handle");
    if (nsteps < 1 || checkpoint_every < 1) return fail(SIFV_ERR_INVALID, "nsteps and checkpoint_every must be >= 1");
    check_step_config(h, cfg);
    int rc = adj_check(h);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * 4 * h->total;
    const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
    while ((int)h->adjCk.size() < nck) {
        double* p = nullptr;
        HIPCHK(hipMalloc(&p, bytes));
        h->adjCk.push_back(p);
    }
    h->adjN = nsteps;
    h->adjK = checkpoint_every;
    h->adjDt = h->cfg.explicitParams.constDt;
    for (int n = 0; n < nsteps; ++n) {
        if (n % checkpoint_every == 0)
            HIPCHK(hipMemcpyAsync(h->adjCk[n / checkpoint_every], h->cons[h->cur], bytes, hipMemcpyDeviceToDevice,
                                  h->stream));
        rc = enqueue_explicit_step(h, -1.0, false, 0.0);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return harvest_timing(h);
}

// Transposes the recorded run: grad = (dU^N/dU^0)^T seed, dgamma = seed .
// dU^N/dgamma.  seed and grad are device arrays of 4 fields x totalCells
// (rho, rhoU, rhoV, rhoE; ghosts ignored, written as 0).  The segment between
// two checkpoints is recomputed forward, then each step is transposed stage
// by stage (SSP-RK3: U1 = U + dt R(U); U2 = (U1 + 3U + dt R(U1))/4;
// U' = (2U2 + U + 2dt R(U2))/3).
int sifv_adjoint_backward(sifv_handle* h, const double* seed, double* grad, double* dgamma) {
    if (!h || !seed || !grad) return fail(SIFV_ERR_INVALID, "null argument");
    if (h->adjN < 1) return fail(SIFV_ERR_INVALID, "no recorded forward run (call sifv_adjoint_forward first)");
    int rc = adj_check(h);
    if (rc) return rc;
    const size_t T = h->total;
    const size_t bytes = sizeof(double) * 4 * T;
    h->adjNb = (int)std::min<size_t>(((size_t)h->nx * h->ny + 255) / 256, 4096);
    if (!h->adjWork) {
        HIPCHK(hipMalloc(&h->adjWork, sizeof(double) * 80 * T));
        HIPCHK(hipMemset(h->adjWork, 0, sizeof(double) * 80 * T));
        HIPCHK(hipMalloc(&h->adjGpart, sizeof(double) * 3 * 4096));
        HIPCHK(hipMalloc(&h->adjGbar, sizeof(double)));
    }
    while ((int)h->adjSeg.size() < h->adjK) {
        double* p = nullptr;
        HIPCHK(hipMalloc(&p, bytes));
        h->adjSeg.push_back(p);
    }
    HIPCHK(hipMemsetAsync(h->adjGbar, 0, sizeof(double), h->stream));
    HIPCHK(hipMemcpyAsync(h->scal, &h->adjDt, sizeof(double), hipMemcpyHostToDevice, h->stream));
    double* lam = adj_slot(h, 19);
    double* L1 = adj_slot(h, 17);
    double* L2 = adj_slot(h, 18);
    HIPCHK(hipMemcpyAsync(lam, seed, bytes, hipMemcpyDeviceToDevice, h->stream));
    const double dt = h->adjDt;
    const int nseg = (h->adjN + h->adjK - 1) / h->adjK;
    for (int sgm = nseg - 1; sgm >= 0; --sgm) {
        const int n0 = sgm * h->adjK, n1 = std::min(h->adjN, n0 + h->adjK);
        HIPCHK(hipMemcpyAsync(h->adjSeg[0], h->adjCk[sgm], bytes, hipMemcpyDeviceToDevice, h->stream));
        for (int n = n0 + 1; n < n1; ++n) {
            rc = adj_stages(h, h->adjSeg[n - 1 - n0], h->adjSeg[n - n0]);
            if (rc) return rc;
        }
        for (int n = n1 - 1; n >= n0; --n) {
            const double* Un = h->adjSeg[n - n0];
            rc = adj_stages(h, Un, nullptr);
            if (rc) return rc;
            const double* U1 = adj_slot(h, 15);
            const double* U2 = adj_slot(h, 16);
            {
                const double* X[1] = {lam};
                const double a[1] = {2.0 / 3.0};
                rc = adj_vjp(h, U2, lam, 2.0 * dt / 3.0, L2, X, a, 1);
            }
            if (!rc) {
                const double* X[1] = {L2};
                const double a[1] = {0.25};
                rc = adj_vjp(h, U1, L2, 0.25 * dt, L1, X, a, 1);
            }
            if (!rc) {
                const double* X[3] = {lam, L2, L1};
                const double a[3] = {1.0 / 3.0, 0.75, 1.0};
                rc = adj_vjp(h, Un, L1, dt, lam, X, a, 3);
            }
            if (rc) return rc;
        }
    }
    HIPCHK(hipMemsetAsync(grad, 0, bytes, h->stream));
    // interior cells carry the gradient; ghosts stay 0
    for (int q = 0; q < 4; ++q)
        hipLaunchKernelGGL(copy_interior_kernel, dim3(elementwise_blocks(h)), dim3(256), 0, h->stream, h->grid(),
                           (const double*)(lam + q * T), grad + q * T);
    HIPCHK(hipGetLastError());
    double gb = 0.0;
    HIPCHK(hipMemcpyAsync(h->hostPinned, h->adjGbar, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    gb = h->hostPinned[0];
    if (dgamma) *dgamma = gb;
    return SIFV_OK;
}

// Host-array form of sifv_adjoint_backward.
int sifv_adjoint_backward_host(sifv_handle* h, const double* seed, double* grad, double* dgamma) {
    if (!h || !seed || !grad) return fail(SIFV_ERR_INVALID, "null argument");
    const size_t bytes = sizeof(double) * 4 * h->total;
    double *ds = nullptr, *dg = nullptr;
    HIPCHK(hipMalloc(&ds, bytes));
    if (hipMalloc(&dg, bytes) != hipSuccess) {
        hipFree(ds);
        return fail(SIFV_ERR_NOMEM, "adjoint staging buffers");
    }
    int rc = SIFV_OK;
    if (hipMemcpy(ds, seed, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = fail(SIFV_ERR_HIP, "seed upload");
    if (!rc) rc = sifv_adjoint_backward(h, ds, dg, dgamma);
    if (!rc && hipMemcpy(grad, dg, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(SIFV_ERR_HIP, "grad download");
    hipFree(ds);
    hipFree(dg);
    return rc;
}

// Device pointer of the resident conservative state (4 fields x totalCells
// for 2D), for callers that keep tensors on the device.
int sifv_state_device_ptr(sifv_handle* h, double** ptr) {
    if (!h || !ptr) return fail(SIFV_ERR_INVALID, "null argument");
    if (h->precision != SIFV_FP64) return fail(SIFV_ERR_UNSUPPORTED, "device state pointer needs an FP64 handle");
    *ptr = static_cast<double*>(h->cons[h->cur]);
    return SIFV_OK;
}

}  // extern "C"
