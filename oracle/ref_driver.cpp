// ref_driver.cpp - thin C-ABI shim over the UNMODIFIED reference sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
// src/*.cpp (read in place under the reference checkout) together with this
// file into oracle/_ref/libsifv_ref.so.  It exposes the same entry points as
// the C restatement (oracle/sifv_oracle.h, prefixed ref_ instead of or_) so
// tests can check the restatement against the real reference call for call.
// Only tests/ and bench.py's cpu_baseline leg load it.
//
// The shim owns an MPI singleton (MPI_Init on first use, never finalized),
// builds MPIContext/RectilinearMesh/HaloExchange the way Runtime does
// (src/Runtime.cpp:39-139) and copies field arrays in and out of a
// SolutionState around each reference call.
#include "ExplicitSolver.hpp"
#include "HLLCSolver.hpp"
#include "HaloExchange.hpp"
#include "IGR.hpp"
#include "IdealGasEOS.hpp"
#include "JacobiPressureSolver.hpp"
#include "LFSolver.hpp"
#include "MPIContext.hpp"
#include "MixtureEOS.hpp"
#include "RKTimeStepping.hpp"
#include "Reconstruction.hpp"
#include "RectilinearMesh.hpp"
#include "RiemannSolver.hpp"
#include "RusanovSolver.hpp"
#include "SemiImplicitSolver.hpp"
#include "SimulationConfig.hpp"
#include "SolutionState.hpp"
#include "StiffenedGasEOS.hpp"
#include "VTKWriter.hpp"

#include <mpi.h>

#include <cstring>
#include <memory>
#include <string>
#include <vector>

extern "C" {
#include "sifv_gpu.h"
}

using namespace SemiImplicitFV;

namespace {

std::string g_err;

void ensure_mpi() {
    int init = 0;
    MPI_Initialized(&init);
    if (!init) MPI_Init(nullptr, nullptr);
}

SimulationConfig to_config(const sifv_config* c) {
    SimulationConfig s;
    s.dim = c->dim;
    s.nGhost = c->nGhost;
    s.RKOrder = c->RKOrder;
    s.useIGR = c->useIGR != 0;
    s.semiImplicit = c->semiImplicit != 0;
    s.reconOrder = static_cast<ReconstructionOrder>(c->reconOrder);
    s.wenoEps = c->wenoEps;
    s.step = c->step;
    s.time = c->time;
    s.explicitParams.cfl = c->explicitParams.cfl;
    s.explicitParams.constDt = c->explicitParams.constDt;
    s.explicitParams.maxDt = c->explicitParams.maxDt;
    s.explicitParams.minDt = c->explicitParams.minDt;
    s.semiImplicitParams.cfl = c->semiImplicitParams.cfl;
    s.semiImplicitParams.maxDt = c->semiImplicitParams.maxDt;
    s.semiImplicitParams.minDt = c->semiImplicitParams.minDt;
    s.semiImplicitParams.maxPressureIters = c->semiImplicitParams.maxPressureIters;
    s.semiImplicitParams.pressureTol = c->semiImplicitParams.pressureTol;
    s.igrParams.alphaCoeff = c->igrParams.alphaCoeff;
    s.igrParams.IGRIters = c->igrParams.IGRIters;
    s.igrParams.IGRWarmStartIters = c->igrParams.IGRWarmStartIters;
    s.multiPhaseParams.nPhases = c->multiPhaseParams.nPhases;
    s.multiPhaseParams.alphaMin = c->multiPhaseParams.alphaMin;
    s.multiPhaseParams.phases.clear();
    for (int p = 0; p < c->multiPhaseParams.nPhases; ++p) {
        PhaseEOS e;
        e.gamma = c->multiPhaseParams.phases[p].gamma;
        e.pInf = c->multiPhaseParams.phases[p].pInf;
        s.multiPhaseParams.phases.push_back(e);
    }
    s.viscousParams.mu = c->viscousMu;
    s.viscousParams.phaseMu.assign(c->viscousPhaseMu, c->viscousPhaseMu + c->viscousNPhaseMu);
    s.surfaceTensionParams.sigma = c->surfaceTensionSigma;
    s.surfaceTensionParams.epsGradAlpha = c->surfaceTensionEps;
    for (int q = 0; q < 3; ++q) {
        s.bodyForceParams.a[q] = c->bodyForceParams.a[q];
        s.bodyForceParams.b[q] = c->bodyForceParams.b[q];
        s.bodyForceParams.c[q] = c->bodyForceParams.c[q];
        s.bodyForceParams.d[q] = c->bodyForceParams.d[q];
    }
    return s;
}

struct RefSolver {
    SimulationConfig config;
    std::unique_ptr<MPIContext> mpi;
    std::unique_ptr<RectilinearMesh> mesh;
    std::unique_ptr<HaloExchange> halo;
    SolutionState state;
    std::shared_ptr<EquationOfState> eos;
    std::shared_ptr<RiemannSolver> rs;
    std::shared_ptr<IGRSolver> igr;
    std::shared_ptr<JacobiPressureSolver> jac;
    std::unique_ptr<ExplicitSolver> ex;
    std::unique_ptr<SemiImplicitSolver> si;
};

std::vector<double>* field_vec(SolutionState& st, int id) {
    switch (id) {
        case SIFV_F_RHO: return &st.rho;
        case SIFV_F_RHOU: return &st.rhoU;
        case SIFV_F_RHOV: return &st.rhoV;
        case SIFV_F_RHOW: return &st.rhoW;
        case SIFV_F_RHOE: return &st.rhoE;
        case SIFV_F_VELU: return &st.velU;
        case SIFV_F_VELV: return &st.velV;
        case SIFV_F_VELW: return &st.velW;
        case SIFV_F_PRES: return &st.pres;
        case SIFV_F_SIGMA: return &st.sigma;
        default: break;
    }
    if (id >= SIFV_F_ALPHARHO && id < SIFV_F_ALPHARHO + SIFV_MAX_PHASES) {
        std::size_t p = id - SIFV_F_ALPHARHO;
        return p < st.alphaRho.size() ? &st.alphaRho[p] : nullptr;
    }
    if (id >= SIFV_F_ALPHA && id < SIFV_F_ALPHA + SIFV_MAX_PHASES) {
        std::size_t p = id - SIFV_F_ALPHA;
        return p < st.alpha.size() ? &st.alpha[p] : nullptr;
    }
    return nullptr;
}

void load(RefSolver* s, double* const* f) {
    for (int id = 0; id < 40; ++id) {
        if (!f[id]) continue;
        std::vector<double>* v = field_vec(s->state, id);
        if (v && !v->empty()) std::memcpy(v->data(), f[id], v->size() * sizeof(double));
    }
}

void store(RefSolver* s, double* const* f) {
    for (int id = 0; id < 40; ++id) {
        if (!f[id]) continue;
        std::vector<double>* v = field_vec(s->state, id);
        if (v && !v->empty()) std::memcpy(f[id], v->data(), v->size() * sizeof(double));
    }
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

void* ref_create(const sifv_mesh* m, const sifv_config* c, int riemann, const sifv_eos* e) {
    try {
        ensure_mpi();
        auto* s = new RefSolver();
        s->config = to_config(c);
        s->config.validate();
        std::vector<double> xn(m->xNodes, m->xNodes + m->nx + 1);
        std::vector<double> yn = {0.0, 1.0}, zn = {0.0, 1.0};
        if (m->dim >= 2) yn.assign(m->yNodes, m->yNodes + m->ny + 1);
        if (m->dim >= 3) zn.assign(m->zNodes, m->zNodes + m->nz + 1);
        int gny = m->dim >= 2 ? m->ny : 1, gnz = m->dim >= 3 ? m->nz : 1;
        std::array<int, 3> periods = {m->periods[0], m->periods[1], m->periods[2]};
        s->mpi = std::make_unique<MPIContext>(
            MPIContext::create(m->nx, gny, gnz, xn, yn, zn, m->dim, periods));
        s->mesh = std::make_unique<RectilinearMesh>(s->config, s->mpi->localXNodes(),
                                                    s->mpi->localYNodes(), s->mpi->localZNodes());
        for (int face = 0; face < 6; ++face)
            if (s->mpi->isPhysicalBoundary(face))
                s->mesh->setBoundaryCondition(face, static_cast<BoundaryCondition>(m->bc[face]));
        s->state.allocate(s->mesh->totalCells(), s->config);
        if (e->kind == SIFV_EOS_STIFFENED)
            s->eos = std::make_shared<StiffenedGasEOS>(e->gamma, e->pInf, 1.0, s->config);
        else
            s->eos = std::make_shared<IdealGasEOS>(e->gamma, 1.0, s->config);
        if (riemann == SIFV_HLLC) s->rs = std::make_shared<HLLCSolver>(s->eos, s->config);
        else if (riemann == SIFV_RUSANOV) s->rs = std::make_shared<RusanovSolver>(s->eos, s->config);
        else s->rs = std::make_shared<LFSolver>(s->eos, s->config);
        if (s->config.useIGR) s->igr = std::make_shared<IGRSolver>(s->config.igrParams);
        s->halo = std::make_unique<HaloExchange>(*s->mpi, *s->mesh);
        if (s->config.semiImplicit) {
            s->jac = std::make_shared<JacobiPressureSolver>();
            s->si = std::make_unique<SemiImplicitSolver>(*s->mesh, s->rs, s->jac, s->eos, s->igr,
                                                         s->config);
            s->si->setHaloExchange(s->halo.get());
        } else {
            s->ex = std::make_unique<ExplicitSolver>(*s->mesh, s->rs, s->eos, s->igr, s->config);
            s->ex->setHaloExchange(s->halo.get());
        }
        return s;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return nullptr;
    }
}

void ref_destroy(void* p) { delete static_cast<RefSolver*>(p); }

size_t ref_total_cells(void* p) { return static_cast<RefSolver*>(p)->mesh->totalCells(); }

double ref_explicit_step(void* p, double* const* f, const sifv_config* c, double targetDt) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    s->config.step = c->step;
    s->config.time = c->time;
    double dt = s->ex->step(s->config, *s->mesh, s->state, targetDt);
    store(s, f);
    return dt;
}

double ref_semi_step(void* p, double* const* f, const sifv_config* c, double targetDt, int* iters) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    s->config.step = c->step;
    s->config.time = c->time;
    double dt = s->si->step(s->config, *s->mesh, s->state, targetDt);
    if (iters) *iters = s->si->lastPressureIterations();
    store(s, f);
    return dt;
}

int ref_jacobi(void* p, const double* rho, const double* rhoc2, const double* rhs, double* pr,
               double dt, double tol, int maxIter) {
    auto* s = static_cast<RefSolver*>(p);
    std::size_t n = s->mesh->totalCells();
    std::vector<double> r(rho, rho + n), c2(rhoc2, rhoc2 + n), b(rhs, rhs + n), x(pr, pr + n);
    JacobiPressureSolver j;
    int it = j.solve(*s->mesh, r, c2, b, x, dt, tol, maxIter, *s->halo);
    std::memcpy(pr, x.data(), n * sizeof(double));
    return it;
}

void ref_fill_ghosts(void* p, double* const* f, int varset) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    s->mesh->applyBoundaryConditions(s->state, static_cast<VarSet>(varset), *s->halo);
    store(s, f);
}

void ref_fill_scalar(void* p, double* a) {
    auto* s = static_cast<RefSolver*>(p);
    std::size_t n = s->mesh->totalCells();
    std::vector<double> v(a, a + n);
    s->mesh->fillScalarGhosts(v, *s->halo);
    std::memcpy(a, v.data(), n * sizeof(double));
}

void ref_cons_to_prim(void* p, double* const* f, const sifv_config* c) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    if (s->config.isMultiPhase())
        MixtureEOS::convertConservativeToPrimitive(*s->mesh, s->state, s->config.multiPhaseParams);
    else
        s->state.convertConservativeToPrimitiveVariables(*s->mesh, s->eos);
    s->mesh->applyBoundaryConditions(s->state, VarSet::PRIM, *s->halo);
    store(s, f);
    (void)c;
}

double ref_acoustic_dt(void* p, double* const* f, const sifv_config* c, double cfl, double maxDt) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    (void)c;
    return computeAcousticTimeStep(*s->mesh, s->state, *s->eos, s->config, cfl, maxDt,
                                   s->halo->mpi().comm());
}

double ref_advective_dt(void* p, double* const* f, double cfl, double maxDt) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    return computeAdvectiveTimeStep(*s->mesh, s->state, cfl, maxDt, s->halo->mpi().comm());
}

void ref_igr(void* p, double* const* f, const sifv_config* c) {
    auto* s = static_cast<RefSolver*>(p);
    load(s, f);
    s->config.step = c->step;
    const RectilinearMesh& m = *s->mesh;
    int dim = m.dim();
    std::vector<GradientTensor> g(m.totalCells());
    const SolutionState& st = s->state;
    for (int k = 0; k < m.nz(); ++k)
        for (int j = 0; j < m.ny(); ++j)
            for (int i = 0; i < m.nx(); ++i) {
                std::size_t id = m.index(i, j, k);
                auto vel = [&](std::size_t q) {
                    return std::array<double, 3>{st.velU[q], dim >= 2 ? st.velV[q] : 0.0,
                                                 dim >= 3 ? st.velW[q] : 0.0};
                };
                auto uxm = vel(m.index(i - 1, j, k)), uxp = vel(m.index(i + 1, j, k));
                std::array<double, 3> uym, uyp, uzm, uzp;
                double dy = 1.0, dz = 1.0;
                if (dim >= 2) {
                    uym = vel(m.index(i, j - 1, k));
                    uyp = vel(m.index(i, j + 1, k));
                    dy = m.dy(j);
                } else {
                    uym = uyp = {st.velU[id], 0.0, 0.0};
                }
                if (dim >= 3) {
                    uzm = vel(m.index(i, j, k - 1));
                    uzp = vel(m.index(i, j, k + 1));
                    dz = m.dz(k);
                } else {
                    uzm = uzp = {st.velU[id], dim >= 2 ? st.velV[id] : 0.0, 0.0};
                }
                g[id] = IGRSolver::computeVelocityGradient(uxm, uxp, uym, uyp, uzm, uzp, m.dx(i),
                                                           dy, dz, dim);
            }
    s->igr->solveEntropicPressure(s->config, m, s->state, g, *s->halo);
    store(s, f);
}

// Face states of a 1D reconstruction through the public Reconstructor API:
// v holds nT = nx + 2*ng values written straight into state.rho (ghosts
// included); outputs are the left/right rho states of the nx+1 x-faces.
int ref_recon_line(int order, double eps, int nx, int ng, const double* v, double* outL,
                   double* outR) {
    ensure_mpi();
    SimulationConfig c;
    c.dim = 1;
    c.nGhost = ng;
    c.reconOrder = static_cast<ReconstructionOrder>(order);
    c.wenoEps = eps;
    std::vector<double> xn(nx + 1);
    for (int i = 0; i <= nx; ++i) xn[i] = i;
    RectilinearMesh m(c, xn);
    SolutionState st;
    st.allocate(m.totalCells(), c);
    for (std::size_t q = 0; q < m.totalCells(); ++q) st.rho[q] = v[q];
    Reconstructor r(c.reconOrder, eps, 1.4, 0.0);
    r.allocate(m);
    r.reconstruct(c, m, st);
    for (int f = 0; f <= nx; ++f) {
        outL[f] = r.xFaceLeft(f).rho;
        outR[f] = r.xFaceRight(f).rho;
    }
    return 0;
}

void ref_riemann_flux(int type, const double* Lp, const double* Rp, int dir, int dim, int incP,
                      int igr, int nph, double* out) {
    auto unpack = [](const double* a) {
        PrimitiveState w;
        w.rho = a[0];
        w.u = {a[1], a[2], a[3]};
        w.p = a[4];
        w.sigma = a[5];
        w.gammaEff = a[6];
        w.piInfEff = a[7];
        for (int q = 0; q < 8; ++q) w.alpha[q] = a[8 + q];
        return w;
    };
    PrimitiveState L = unpack(Lp), R = unpack(Rp);
    FluxConfig fc{dim, incP != 0, igr != 0, nph};
    std::array<double, 3> n = {0.0, 0.0, 0.0};
    n[dir] = 1.0;
    RiemannFlux F = computeFluxDirect(static_cast<RiemannSolverType>(type), L, R, n, fc);
    out[0] = F.massFlux;
    out[1] = F.momentumFlux[0];
    out[2] = F.momentumFlux[1];
    out[3] = F.momentumFlux[2];
    out[4] = F.energyFlux;
    for (int q = 0; q < 8; ++q) out[5 + q] = F.alphaFlux[q];
    out[13] = F.pressureFlux;
    out[14] = F.faceVelocity;
}

// VTKWriter::writeVTR of the given fields with this rank's extent and rank
// number, as VTKSession::write calls it (src/VTKSession.cpp:23-31); rank < 0:
// the serial form (no extent, no Rank array).
int ref_write_vtr(void* p, double* const* f, const char* filename, int rank) {
    try {
        auto* s = static_cast<RefSolver*>(p);
        load(s, f);
        if (rank >= 0)
            VTKWriter::writeVTR(filename, *s->mesh, s->state, s->config, s->mpi->localExtent(), rank);
        else
            VTKWriter::writeVTR(filename, *s->mesh, s->state, s->config);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// VTKWriter::writePVTR for the one-rank decomposition (src/VTKSession.cpp:33-49).
int ref_write_pvtr(void* p, const char* filename, const char* pieceFile) {
    try {
        auto* s = static_cast<RefSolver*>(p);
        const auto& e = s->mpi->localExtent();
        VTKWriter::writePVTR(filename, e[1] - e[0], e[3] - e[2], e[5] - e[4], {e}, {std::string(pieceFile)},
                             s->config);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

int ref_write_pvd(const char* filename, const char* mode, double time, const char* dataFile) {
    try {
        VTKWriter::writePVD(filename, mode, time, dataFile);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

}  // extern "C"
