// sifv_ensemble.hip - C ABI of the batched 1D run ensemble
// (include/sifv_gpu.h, sifv_ensemble_*): one persistent launch advances every
// tube to the end time; the tile width (workgroup size) is picked by a
// timing autotuner over the resident state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/sifv_gpu.h"
#include "sifv_ensemble.hpp"

namespace sifv {
int set_error(int code, const std::string& msg);
}
using namespace sifv;

namespace {

#define EHIP(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return set_error(SIFV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
using EnsFn = void (*)(EnsArgs<T>);

template <typename T, int RS>
EnsFn<T> ens_kernel_rs(int block) {
    switch (block) {
        case 64: return ensemble_kernel<64, RS, T>;
        case 128: return ensemble_kernel<128, RS, T>;
        case 256: return ensemble_kernel<256, RS, T>;
        case 512: return ensemble_kernel<512, RS, T>;
        case 1024: return ensemble_kernel<1024, RS, T>;
        default: return nullptr;
    }
}

template <typename T>
EnsFn<T> ens_kernel(int block, int rs) {
    if (rs == RS_HLLC) return ens_kernel_rs<T, RS_HLLC>(block);
    if (rs == RS_RUSANOV) return ens_kernel_rs<T, RS_RUSANOV>(block);
    return ens_kernel_rs<T, RS_LF>(block);
}

template <typename T>
__global__ void __launch_bounds__(256) ens_c2p_kernel(Eos e, const T* U, long long nt, double* vel, double* pres) {
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < nt; c += (long long)gridDim.x * blockDim.x) {
        double mom[1] = {(double)U[nt + c]}, u[1], p;
        cons_to_prim<1>(e, (double)U[c], mom, (double)U[2 * nt + c], u, p);
        vel[c] = u[0];
        pres[c] = p;
    }
}

}  // namespace

struct sifv_ensemble {
    int ntubes = 0;
    long long ntot = 0;
    int precision = SIFV_FP64;
    int riemann = SIFV_HLLC;
    sifv_config cfg;
    Eos eos;
    std::vector<int> ncell;
    std::vector<long long> off;
    std::vector<double> cost;   // relative cost estimate (cells x wave speed / dx)
    hipStream_t stream = nullptr;
    int *dNcell = nullptr, *dOrder = nullptr, *dQueue = nullptr, *dSteps = nullptr, *dStatus = nullptr;
    int *dNparts = nullptr, *dBar = nullptr;
    double* dDtSlot = nullptr;
    int nitems = 0;
    int lastGrid = 0;
    long long* dOff = nullptr;
    double *dDx = nullptr, *dTime = nullptr, *dVel = nullptr, *dPres = nullptr;
    void* U[3] = {nullptr, nullptr, nullptr};
    void* saved = nullptr;      // autotuner: copy of the state before a trial
    int block = 256;
    int wgPerCu = 4;
    int cus = 256;
    size_t tsize() const { return precision == SIFV_MIXED ? sizeof(float) : sizeof(double); }
};

namespace {

template <typename T>
int ens_launch(sifv_ensemble* e, double endTime, int maxSteps, float* ms) {
    EnsFn<T> fn = ens_kernel<T>(e->block, e->riemann == SIFV_HLLC ? RS_HLLC : e->riemann == SIFV_RUSANOV ? RS_RUSANOV : RS_LF);
    if (!fn) return set_error(SIFV_ERR_INVALID, "unsupported ensemble block size");
    EnsArgs<T> A;
    A.eos = e->eos;
    A.ntubes = e->ntubes;
    A.order = e->dOrder;
    A.nitems = e->nitems;
    A.nparts = e->dNparts;
    A.bar = e->dBar;
    A.dtSlot = e->dDtSlot;
    A.err = e->dStatus + 1;
    A.ncell = e->dNcell;
    A.off = e->dOff;
    A.dxw = e->dDx;
    for (int b = 0; b < 3; ++b) A.U[b] = static_cast<T*>(e->U[b]);
    A.ntot = e->ntot;
    A.velIC = e->dVel;
    A.presIC = e->dPres;
    A.cfl = e->cfg.explicitParams.cfl;
    A.maxDt = e->cfg.explicitParams.maxDt;
    A.constDt = e->cfg.explicitParams.constDt;
    A.tEnd = endTime;
    A.eps = e->cfg.wenoEps;
    A.maxSteps = maxSteps;
    A.rk = e->cfg.RKOrder;
    A.queue = e->dQueue;
    A.steps = e->dSteps;
    A.time = e->dTime;
    A.status = e->dStatus;
    A.cur = 0;
    EHIP(hipMemsetAsync(e->dQueue, 0, sizeof(int), e->stream));
    EHIP(hipMemsetAsync(e->dStatus, 0, 2 * sizeof(int), e->stream));
    EHIP(hipMemsetAsync(e->dBar, 0, sizeof(int) * e->ntubes, e->stream));
    const int grid = e->lastGrid;
    hipEvent_t e0, e1;
    EHIP(hipEventCreate(&e0));
    EHIP(hipEventCreate(&e1));
    EHIP(hipEventRecord(e0, e->stream));
    void* args[] = {&A};
    EHIP(hipLaunchKernel(reinterpret_cast<const void*>(fn), dim3(grid), dim3(e->block), args, 0, e->stream));
    EHIP(hipEventRecord(e1, e->stream));
    EHIP(hipEventSynchronize(e1));
    float t = 0.0f;
    EHIP(hipEventElapsedTime(&t, e0, e1));
    EHIP(hipEventDestroy(e0));
    EHIP(hipEventDestroy(e1));
    if (ms) *ms = t;
    int st[2] = {0, 0};
    EHIP(hipMemcpy(st, e->dStatus, 2 * sizeof(int), hipMemcpyDeviceToHost));
    if (st[1]) return set_error(SIFV_ERR_HIP, "ensemble part barrier exceeded its wait budget");
    if (st[0]) return set_error(SIFV_ERR_NONFINITE, "non-finite time step in an ensemble member");
    return SIFV_OK;
}

// Work items for the current shape: the persistent grid is capped at the
// number of co-resident workgroups; a tube whose cost exceeds the fair share
// of one workgroup is cut into up to 32 parts (each at least one tile, never
// more parts than workgroups); items are queued costliest part first with
// the parts of a tube adjacent (longest-processing-time order).
template <typename T>
int ens_plan(sifv_ensemble* e) {
    EnsFn<T> fn = ens_kernel<T>(e->block, e->riemann == SIFV_HLLC ? RS_HLLC : e->riemann == SIFV_RUSANOV ? RS_RUSANOV : RS_LF);
    if (!fn) return set_error(SIFV_ERR_INVALID, "unsupported ensemble block size");
    int perCu = 0;
    EHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void*>(fn), e->block, 0));
    perCu = std::max(1, std::min(perCu, e->wgPerCu));
    const int G = std::max(1, e->cus * perCu);
    double total = 0.0;
    for (int t = 0; t < e->ntubes; ++t) total += e->cost[t];
    const double share = total / G;
    std::vector<int> parts(e->ntubes, 1);
    std::vector<std::pair<double, int>> items;
    for (int t = 0; t < e->ntubes; ++t) {
        const int tiles = (e->ncell[t] + e->block - 1) / e->block;
        int p = share > 0.0 ? (int)std::ceil(e->cost[t] / share) : 1;
        p = std::max(1, std::min({p, tiles, 32, G}));
        parts[t] = p;
    }
    std::vector<int> tubes(e->ntubes);
    std::iota(tubes.begin(), tubes.end(), 0);
    std::stable_sort(tubes.begin(), tubes.end(),
                     [&](int a, int b) { return e->cost[a] / parts[a] > e->cost[b] / parts[b]; });
    std::vector<int> order;
    for (int t : tubes)
        for (int p = 0; p < parts[t]; ++p) order.push_back(t * 32 + p);
    e->nitems = (int)order.size();
    if (!e->dNparts) {
        EHIP(hipMalloc(&e->dNparts, sizeof(int) * e->ntubes));
        EHIP(hipMalloc(&e->dBar, sizeof(int) * e->ntubes));
        EHIP(hipMalloc(&e->dDtSlot, sizeof(double) * 32 * e->ntubes));
    }
    hipFree(e->dOrder);
    e->dOrder = nullptr;
    EHIP(hipMalloc(&e->dOrder, sizeof(int) * order.size()));
    EHIP(hipMemcpy(e->dOrder, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    EHIP(hipMemcpy(e->dNparts, parts.data(), sizeof(int) * e->ntubes, hipMemcpyHostToDevice));
    e->lastGrid = std::min(G, e->nitems);
    return SIFV_OK;
}

int ens_run(sifv_ensemble* e, double endTime, int maxSteps, float* ms) {
    int rc = e->precision == SIFV_MIXED ? ens_plan<float>(e) : ens_plan<double>(e);
    if (rc) return rc;
    return e->precision == SIFV_MIXED ? ens_launch<float>(e, endTime, maxSteps, ms)
                                      : ens_launch<double>(e, endTime, maxSteps, ms);
}

int field_slot(int field) {
    if (field == SIFV_F_RHO) return 0;
    if (field == SIFV_F_RHOU) return 1;
    if (field == SIFV_F_RHOE) return 2;
    return -1;
}

}  // namespace

extern "C" {

int sifv_ensemble_create(int ntubes, const int* ncells, const double* nodes, const sifv_config* cfg, int riemann,
                         const sifv_eos* eos, int precision, sifv_ensemble** out) {
    if (!ncells || !nodes || !cfg || !eos || !out || ntubes <= 0) return set_error(SIFV_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->dim != 1) return set_error(SIFV_ERR_INVALID, "ensemble members are 1D runs");
    if (cfg->reconOrder != SIFV_WENO5) return set_error(SIFV_ERR_UNSUPPORTED, "ensemble runs use WENO5");
    if (cfg->RKOrder != 2 && cfg->RKOrder != 3) return set_error(SIFV_ERR_UNSUPPORTED, "ensemble runs use SSP-RK2 or RK3");
    if (cfg->semiImplicit || cfg->useIGR || cfg->multiPhaseParams.nPhases != 0)
        return set_error(SIFV_ERR_UNSUPPORTED, "ensemble runs are single-phase explicit without IGR");
    if (cfg->explicitParams.cfl <= 0 || cfg->explicitParams.maxDt <= 0)
        return set_error(SIFV_ERR_INVALID, "explicitParams.cfl and maxDt must be > 0");
    if (riemann < SIFV_LF || riemann > SIFV_HLLC) return set_error(SIFV_ERR_INVALID, "unknown Riemann solver");
    if (precision != SIFV_FP64 && precision != SIFV_MIXED) return set_error(SIFV_ERR_INVALID, "unknown precision");
    if (precision == SIFV_MIXED && eos->kind == SIFV_EOS_STIFFENED)
        return set_error(SIFV_ERR_UNSUPPORTED,
                         "ensemble tubes store their FP32 state without the stiffened-gas energy offset: "
                         "stiffened gases run in FP64");
    for (int d = 0; d < 3; ++d)
        if (cfg->bodyForceParams.a[d] != 0.0 || cfg->bodyForceParams.b[d] != 0.0 || cfg->viscousMu > 0.0 ||
            cfg->viscousNPhaseMu > 0 || cfg->surfaceTensionSigma > 0.0)
            return set_error(SIFV_ERR_UNSUPPORTED, "ensemble runs carry the flux terms only");
    auto* e = new sifv_ensemble();
    e->ntubes = ntubes;
    e->precision = precision;
    e->riemann = riemann;
    e->cfg = *cfg;
    e->eos.gamma = eos->gamma;
    e->eos.pInf = eos->kind == SIFV_EOS_STIFFENED ? eos->pInf : 0.0;
    e->eos.stiffened = eos->kind == SIFV_EOS_STIFFENED;
    e->ncell.assign(ncells, ncells + ntubes);
    e->off.resize(ntubes);
    std::vector<double> dx;
    long long o = 0, on = 0;
    for (int t = 0; t < ntubes; ++t) {
        if (ncells[t] < 1) {
            delete e;
            return set_error(SIFV_ERR_INVALID, "every tube needs at least one cell");
        }
        e->off[t] = o;
        for (int i = 0; i < ncells[t]; ++i) dx.push_back(nodes[on + i + 1] - nodes[on + i]);
        o += ncells[t];
        on += ncells[t] + 1;
    }
    e->ntot = o;
    e->cost.assign(ntubes, 0.0);
    for (int t = 0; t < ntubes; ++t) e->cost[t] = (double)ncells[t] * ncells[t];
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0) e->cus = prop.multiProcessorCount;
    auto bail = [&](hipError_t err, const char* what) {
        sifv_ensemble_destroy(e);
        return set_error(SIFV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
    };
#define ECK(expr)                                  \
    do {                                           \
        hipError_t e_ = (expr);                    \
        if (e_ != hipSuccess) return bail(e_, #expr); \
    } while (0)
    ECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    ECK(hipMalloc(&e->dNcell, sizeof(int) * ntubes));
    ECK(hipMalloc(&e->dOrder, sizeof(int) * ntubes));
    ECK(hipMalloc(&e->dOff, sizeof(long long) * ntubes));
    ECK(hipMalloc(&e->dSteps, sizeof(int) * ntubes));
    ECK(hipMalloc(&e->dTime, sizeof(double) * ntubes));
    ECK(hipMalloc(&e->dQueue, sizeof(int)));
    ECK(hipMalloc(&e->dStatus, 2 * sizeof(int)));
    ECK(hipMalloc(&e->dDx, sizeof(double) * e->ntot));
    ECK(hipMalloc(&e->dVel, sizeof(double) * e->ntot));
    ECK(hipMalloc(&e->dPres, sizeof(double) * e->ntot));
    for (int b = 0; b < 3; ++b) {
        ECK(hipMalloc(&e->U[b], e->tsize() * 3 * e->ntot));
        ECK(hipMemset(e->U[b], 0, e->tsize() * 3 * e->ntot));
    }
    ECK(hipMemset(e->dVel, 0, sizeof(double) * e->ntot));
    ECK(hipMemset(e->dPres, 0, sizeof(double) * e->ntot));
    ECK(hipMemcpy(e->dNcell, ncells, sizeof(int) * ntubes, hipMemcpyHostToDevice));
    ECK(hipMemcpy(e->dOff, e->off.data(), sizeof(long long) * ntubes, hipMemcpyHostToDevice));
    ECK(hipMemcpy(e->dDx, dx.data(), sizeof(double) * e->ntot, hipMemcpyHostToDevice));
#undef ECK
    *out = e;
    return SIFV_OK;
}

int sifv_ensemble_destroy(sifv_ensemble* e) {
    if (!e) return SIFV_OK;
    if (e->stream) hipStreamSynchronize(e->stream);
    hipFree(e->dNcell);
    hipFree(e->dOrder);
    hipFree(e->dOff);
    hipFree(e->dSteps);
    hipFree(e->dTime);
    hipFree(e->dQueue);
    hipFree(e->dStatus);
    hipFree(e->dNparts);
    hipFree(e->dBar);
    hipFree(e->dDtSlot);
    hipFree(e->dDx);
    hipFree(e->dVel);
    hipFree(e->dPres);
    for (int b = 0; b < 3; ++b) hipFree(e->U[b]);
    hipFree(e->saved);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return SIFV_OK;
}

int sifv_ensemble_total_cells(const sifv_ensemble* e, size_t* n) {
    if (!e || !n) return set_error(SIFV_ERR_INVALID, "null argument");
    *n = (size_t)e->ntot;
    return SIFV_OK;
}

int sifv_ensemble_upload(sifv_ensemble* e, int field, const double* host, size_t n) {
    if (!e || !host) return set_error(SIFV_ERR_INVALID, "null argument");
    if ((long long)n != e->ntot) return set_error(SIFV_ERR_INVALID, "array length != total ensemble cells");
    EHIP(hipStreamSynchronize(e->stream));
    if (field == SIFV_F_VELU || field == SIFV_F_PRES) {
        EHIP(hipMemcpy(field == SIFV_F_VELU ? e->dVel : e->dPres, host, n * sizeof(double), hipMemcpyHostToDevice));
        if (field == SIFV_F_PRES) {
            // cost estimate for the queue order: cells x fastest initial signal / dx
            // (steps scale with it), refreshed once the pressure is known
            std::vector<double> rho(n), u(n);
            std::vector<double> p(host, host + n);
            // density and velocity as uploaded so far (the order only steers load balance)
            if (e->precision == SIFV_MIXED) {
                std::vector<float> tmp(n);
                EHIP(hipMemcpy(tmp.data(), e->U[0], n * sizeof(float), hipMemcpyDeviceToHost));
                for (size_t c = 0; c < n; ++c) rho[c] = tmp[c];
            } else {
                EHIP(hipMemcpy(rho.data(), e->U[0], n * sizeof(double), hipMemcpyDeviceToHost));
            }
            EHIP(hipMemcpy(u.data(), e->dVel, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int t = 0; t < e->ntubes; ++t) {
                double smax = 0.0;
                for (int i = 0; i < e->ncell[t]; ++i) {
                    const long long c = e->off[t] + i;
                    const double r = std::max(rho[c], 1e-14);
                    const double cs = std::sqrt(e->eos.gamma * std::max(p[c] + e->eos.pInf, 1e-14) / r);
                    smax = std::max(smax, std::fabs(u[c]) + cs);
                }
                e->cost[t] = (double)e->ncell[t] * e->ncell[t] * std::max(smax, 1e-30);
            }
        }
        return SIFV_OK;
    }
    const int q = field_slot(field);
    if (q < 0) return set_error(SIFV_ERR_INVALID, "ensemble fields: rho, rhoU, rhoE, velU, pres");
    if (e->precision == SIFV_MIXED) {
        std::vector<float> tmp(host, host + n);
        EHIP(hipMemcpy(static_cast<float*>(e->U[0]) + q * e->ntot, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
    } else {
        EHIP(hipMemcpy(static_cast<double*>(e->U[0]) + q * e->ntot, host, n * sizeof(double), hipMemcpyHostToDevice));
    }
    return SIFV_OK;
}

int sifv_ensemble_download(sifv_ensemble* e, int field, double* host, size_t n) {
    if (!e || !host) return set_error(SIFV_ERR_INVALID, "null argument");
    if ((long long)n != e->ntot) return set_error(SIFV_ERR_INVALID, "array length != total ensemble cells");
    EHIP(hipStreamSynchronize(e->stream));
    if (field == SIFV_F_VELU || field == SIFV_F_PRES) {
        // closing conversion of ExplicitSolver::step (src/ExplicitSolver.cpp:192-200)
        const int nb = (int)std::min<long long>((e->ntot + 255) / 256, 8192);
        if (e->precision == SIFV_MIXED)
            hipLaunchKernelGGL(ens_c2p_kernel<float>, dim3(nb), dim3(256), 0, e->stream, e->eos,
                               static_cast<const float*>(e->U[0]), e->ntot, e->dVel, e->dPres);
        else
            hipLaunchKernelGGL(ens_c2p_kernel<double>, dim3(nb), dim3(256), 0, e->stream, e->eos,
                               static_cast<const double*>(e->U[0]), e->ntot, e->dVel, e->dPres);
        EHIP(hipGetLastError());
        EHIP(hipStreamSynchronize(e->stream));
        EHIP(hipMemcpy(host, field == SIFV_F_VELU ? e->dVel : e->dPres, n * sizeof(double), hipMemcpyDeviceToHost));
        return SIFV_OK;
    }
    const int q = field_slot(field);
    if (q < 0) return set_error(SIFV_ERR_INVALID, "ensemble fields: rho, rhoU, rhoE, velU, pres");
    if (e->precision == SIFV_MIXED) {
        std::vector<float> tmp(n);
        EHIP(hipMemcpy(tmp.data(), static_cast<float*>(e->U[0]) + q * e->ntot, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t c = 0; c < n; ++c) host[c] = tmp[c];
    } else {
        EHIP(hipMemcpy(host, static_cast<double*>(e->U[0]) + q * e->ntot, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    return SIFV_OK;
}

int sifv_ensemble_set_shape(sifv_ensemble* e, int block, int wg_per_cu) {
    if (!e) return set_error(SIFV_ERR_INVALID, "null argument");
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
        return set_error(SIFV_ERR_INVALID, "block must be 64, 128, 256, 512 or 1024");
    if (wg_per_cu < 1 || wg_per_cu > 32) return set_error(SIFV_ERR_INVALID, "wg_per_cu out of range");
    e->block = block;
    e->wgPerCu = wg_per_cu;
    return SIFV_OK;
}

int sifv_ensemble_run(sifv_ensemble* e, double endTime, int maxSteps, int* steps, double* times, double* kernel_ms) {
    if (!e) return set_error(SIFV_ERR_INVALID, "null argument");
    if (maxSteps < 1) return set_error(SIFV_ERR_INVALID, "maxSteps must be >= 1");
    float ms = 0.0f;
    int rc = ens_run(e, endTime, maxSteps, &ms);
    if (kernel_ms) *kernel_ms = ms;
    if (steps) EHIP(hipMemcpy(steps, e->dSteps, sizeof(int) * e->ntubes, hipMemcpyDeviceToHost));
    if (times) EHIP(hipMemcpy(times, e->dTime, sizeof(double) * e->ntubes, hipMemcpyDeviceToHost));
    return rc;
}

// Times one run per candidate shape (block size x workgroups per CU) from the
// same resident state, restores the state, and keeps the fastest shape.
int sifv_ensemble_autotune(sifv_ensemble* e, double endTime, int maxSteps, int* block_out, int* wg_out,
                           double* best_ms) {
    if (!e) return set_error(SIFV_ERR_INVALID, "null argument");
    const size_t bytes = e->tsize() * 3 * e->ntot;
    if (!e->saved) EHIP(hipMalloc(&e->saved, bytes));
    EHIP(hipMemcpy(e->saved, e->U[0], bytes, hipMemcpyDeviceToDevice));
    const int blocks[] = {64, 128, 256, 512, 1024};
    const int wgs[] = {1, 2, 4, 8};
    int bestB = e->block, bestW = e->wgPerCu;
    float bestT = 1e30f;
    for (int b : blocks)
        for (int w : wgs) {
            if (b * w > 2048) continue;  // at most 32 waves per CU
            e->block = b;
            e->wgPerCu = w;
            float ms = 0.0f;
            int rc = ens_run(e, endTime, maxSteps, &ms);
            EHIP(hipMemcpy(e->U[0], e->saved, bytes, hipMemcpyDeviceToDevice));
            if (rc) return rc;
            if (ms < bestT) {
                bestT = ms;
                bestB = b;
                bestW = w;
            }
        }
    e->block = bestB;
    e->wgPerCu = bestW;
    if (block_out) *block_out = bestB;
    if (wg_out) *wg_out = bestW;
    if (best_ms) *best_ms = bestT;
    return SIFV_OK;
}

}  // extern "C"
