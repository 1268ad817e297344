// sifv_gpu_solvers.hpp - drop-in C++ solver classes for the reference's own
// driver code, over the C ABI of sifv_gpu.h (link with libsifv_hip.so).
//
// Include after the reference headers.  The classes keep the reference's
// constructor and step()/solve() signatures:
//   ExplicitSolver(mesh, riemannSolver, eos, igrSolver, config)   include/ExplicitSolver.hpp:23-29
//   double step(config, mesh, state, targetDt)                   include/ExplicitSolver.hpp:33-36
//   SemiImplicitSolver(mesh, riemannSolver, pressureSolver, eos, igrSolver, config)
//                                                                 include/SemiImplicitSolver.hpp:24-31
//   int PressureSolver::solve(mesh, rho, rhoc2, rhs, pressure, dt, tol, maxIter[, halo])
//                                                                 include/PressureSolver.hpp:17-40
// By default step() uploads the SolutionState, advances it on the GPU and
// writes every field back, as the reference mutates its SolutionState in
// place.  setResident(true) keeps the state on the device between steps
// instead: step() uploads only a SolutionState it has not seen (or one marked
// with markHostModified()) and writes the state back only after a step that
// reaches its targetDt (targetDt > 0 and dt within 1e-6 of it), which are
// exactly the steps after which the reference's runTimeLoop
// (src/RKTimeStepping.cpp:241-356) writes VTK output, checks for NaN or ends;
// syncToHost() writes it back on demand.  So the reference's own
// runTimeLoop drives a resident solver unchanged, with transfers only at
// output times.  (Its per-step console line reads state.sigma for IGR runs,
// which a resident solver refreshes only at those syncs.)
// setHaloExchange() only records which faces the MPI topology makes
// periodic (one device: the periodic self exchange becomes a wrap).
#ifndef SIFV_GPU_SOLVERS_HPP
#define SIFV_GPU_SOLVERS_HPP

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sifv_gpu.h"

namespace SemiImplicitFV {
namespace gpu {

inline void check_status(int rc) {
    if (rc == SIFV_OK) return;
    if (rc == SIFV_ERR_INVALID) throw std::invalid_argument(sifv_last_error());
    throw std::runtime_error(sifv_last_error());
}

inline sifv_config to_c(const SimulationConfig& c) {
    sifv_config o{};
    o.dim = c.dim;
    o.nGhost = c.nGhost;
    o.RKOrder = c.RKOrder;
    o.useIGR = c.useIGR;
    o.semiImplicit = c.semiImplicit;
    o.reconOrder = static_cast<int>(c.reconOrder);
    o.wenoEps = c.wenoEps;
    o.step = c.step;
    o.time = c.time;
    o.explicitParams.cfl = c.explicitParams.cfl;
    o.explicitParams.constDt = c.explicitParams.constDt;
    o.explicitParams.maxDt = c.explicitParams.maxDt;
    o.explicitParams.minDt = c.explicitParams.minDt;
    o.semiImplicitParams.cfl = c.semiImplicitParams.cfl;
    o.semiImplicitParams.maxDt = c.semiImplicitParams.maxDt;
    o.semiImplicitParams.minDt = c.semiImplicitParams.minDt;
    o.semiImplicitParams.maxPressureIters = c.semiImplicitParams.maxPressureIters;
    o.semiImplicitParams.pressureTol = c.semiImplicitParams.pressureTol;
    o.igrParams.alphaCoeff = c.igrParams.alphaCoeff;
    o.igrParams.IGRIters = c.igrParams.IGRIters;
    o.igrParams.IGRWarmStartIters = c.igrParams.IGRWarmStartIters;
    o.multiPhaseParams.nPhases = c.multiPhaseParams.nPhases;
    for (int p = 0; p < c.multiPhaseParams.nPhases && p < SIFV_MAX_PHASES; ++p) {
        o.multiPhaseParams.phases[p].gamma = c.multiPhaseParams.phases[p].gamma;
        o.multiPhaseParams.phases[p].pInf = c.multiPhaseParams.phases[p].pInf;
    }
    o.multiPhaseParams.alphaMin = c.multiPhaseParams.alphaMin;
    for (int d = 0; d < 3; ++d) {
        o.bodyForceParams.a[d] = c.bodyForceParams.a[d];
        o.bodyForceParams.b[d] = c.bodyForceParams.b[d];
        o.bodyForceParams.c[d] = c.bodyForceParams.c[d];
        o.bodyForceParams.d[d] = c.bodyForceParams.d[d];
    }
    o.viscousMu = c.viscousParams.mu;
    o.viscousNPhaseMu = (int)c.viscousParams.phaseMu.size();
    for (int p = 0; p < o.viscousNPhaseMu && p < SIFV_MAX_PHASES; ++p)
        o.viscousPhaseMu[p] = c.viscousParams.phaseMu[p];
    o.surfaceTensionSigma = c.surfaceTensionParams.sigma;
    o.surfaceTensionEps = c.surfaceTensionParams.epsGradAlpha;
    return o;
}

// Field ids of the SolutionState members present for this state.
inline std::vector<std::pair<int, std::vector<double>*>> state_fields(SolutionState& s) {
    std::vector<std::pair<int, std::vector<double>*>> f = {
        {SIFV_F_RHO, &s.rho}, {SIFV_F_RHOU, &s.rhoU}, {SIFV_F_RHOE, &s.rhoE},
        {SIFV_F_VELU, &s.velU}, {SIFV_F_PRES, &s.pres}, {SIFV_F_SIGMA, &s.sigma}};
    if (s.dim() >= 2) {
        f.push_back({SIFV_F_RHOV, &s.rhoV});
        f.push_back({SIFV_F_VELV, &s.velV});
    }
    if (s.dim() >= 3) {
        f.push_back({SIFV_F_RHOW, &s.rhoW});
        f.push_back({SIFV_F_VELW, &s.velW});
    }
    for (std::size_t p = 0; p < s.alphaRho.size(); ++p) f.push_back({SIFV_F_ALPHARHO + (int)p, &s.alphaRho[p]});
    for (std::size_t p = 0; p < s.alpha.size(); ++p) f.push_back({SIFV_F_ALPHA + (int)p, &s.alpha[p]});
    return f;
}

class GpuSolverBase {
public:
    GpuSolverBase(std::shared_ptr<RiemannSolver> rs, std::shared_ptr<EquationOfState> eos,
                  const SimulationConfig& config, int precision)
        : rs_(std::move(rs)), eos_(std::move(eos)), precision_(precision) {
        config.validate();
        const std::string n = rs_->name();  // name() -> enum, src/ExplicitSolver.cpp:31-34
        riemann_ = n == "HLLC" ? SIFV_HLLC : (n == "Rusanov" ? SIFV_RUSANOV : SIFV_LF);
    }
    virtual ~GpuSolverBase() { sifv_destroy(h_); }
    GpuSolverBase(const GpuSolverBase&) = delete;
    GpuSolverBase& operator=(const GpuSolverBase&) = delete;

    void setHaloExchange(HaloExchange* halo) { halo_ = halo; }
    sifv_handle* handle() const { return h_; }

    // Resident mode (see the file comment); off by default.
    void setResident(bool on) {
        resident_ = on;
        bound_ = nullptr;
    }
    bool resident() const { return resident_; }
    // Writes the device state into `state` (every SolutionState member the GPU holds).
    void syncToHost(SolutionState& state) {
        if (h_) download(state);
    }
    // The host changed the bound state in place: the next step uploads it.
    void markHostModified() { bound_ = nullptr; }
    // Whole-state uploads / downloads so far (each moves every field).
    long uploads() const { return uploads_; }
    long downloads() const { return downloads_; }

protected:
    void stage_in(SolutionState& s) {
        if (!resident_ || bound_ != &s || boundData_ != s.rho.data()) {
            upload(s);
            bound_ = &s;
            boundData_ = s.rho.data();
        }
    }
    void stage_out(SolutionState& s, double dt, double targetDt) {
        if (!resident_ || (targetDt > 0.0 && dt >= targetDt * (1.0 - 1e-6))) download(s);
    }
    void ensure(const SimulationConfig& config, const RectilinearMesh& mesh) {
        if (h_) return;
        std::vector<double> x(mesh.nx() + 1), y(mesh.ny() + 1), z(mesh.nz() + 1);
        for (int i = 0; i <= mesh.nx(); ++i) x[i] = mesh.nodeX(i);
        for (int j = 0; j <= mesh.ny(); ++j) y[j] = mesh.nodeY(j);
        for (int k = 0; k <= mesh.nz(); ++k) z[k] = mesh.nodeZ(k);
        sifv_mesh m{};
        m.dim = mesh.dim();
        m.nx = mesh.nx();
        m.ny = mesh.ny();
        m.nz = mesh.nz();
        m.nGhost = mesh.nGhost();
        m.xNodes = x.data();
        m.yNodes = y.data();
        m.zNodes = z.data();
        for (int f = 0; f < 6; ++f) m.bc[f] = static_cast<int>(mesh.boundaryCondition(f));
        for (int d = 0; d < 3; ++d)
            m.periods[d] = halo_ && d < mesh.dim() && !halo_->mpi().isPhysicalBoundary(2 * d) ? 1 : 0;
        sifv_config c = to_c(config);
        sifv_eos e{};
        e.kind = eos_->pInf() != 0.0 ? SIFV_EOS_STIFFENED : SIFV_EOS_IDEAL;
        e.gamma = eos_->gamma();
        e.pInf = eos_->pInf();
        check_status(sifv_create(&m, &c, riemann_, &e, precision_, &h_));
    }
    void upload(SolutionState& s) {
        for (auto& f : state_fields(s)) check_status(sifv_upload(h_, f.first, f.second->data(), f.second->size()));
        ++uploads_;
    }
    void download(SolutionState& s) {
        for (auto& f : state_fields(s)) check_status(sifv_download(h_, f.first, f.second->data(), f.second->size()));
        ++downloads_;
    }

    std::shared_ptr<RiemannSolver> rs_;
    std::shared_ptr<EquationOfState> eos_;
    int precision_;
    int riemann_ = SIFV_HLLC;
    HaloExchange* halo_ = nullptr;
    sifv_handle* h_ = nullptr;
    bool resident_ = false;
    const SolutionState* bound_ = nullptr;  // state the device copy belongs to
    const double* boundData_ = nullptr;
    long uploads_ = 0, downloads_ = 0;
};

// Drop-in for ExplicitSolver (include/ExplicitSolver.hpp:21-82).
class ExplicitSolver : public GpuSolverBase {
public:
    ExplicitSolver(const RectilinearMesh&, std::shared_ptr<RiemannSolver> riemannSolver,
                   std::shared_ptr<EquationOfState> eos, std::shared_ptr<IGRSolver>,
                   const SimulationConfig& config, int precision = SIFV_FP64)
        : GpuSolverBase(std::move(riemannSolver), std::move(eos), config, precision) {}

    double step(const SimulationConfig& config, const RectilinearMesh& mesh, SolutionState& state,
                double targetDt = -1.0) {
        ensure(config, mesh);
        stage_in(state);
        sifv_config c = to_c(config);
        double dt = 0.0;
        check_status(sifv_explicit_step(h_, &c, targetDt, &dt));
        stage_out(state, dt, targetDt);
        return dt;
    }
};

// Drop-in for SemiImplicitSolver (include/SemiImplicitSolver.hpp:22-96) with
// its pressure solve on the GPU, chosen by the pressure solver's name() as
// the reference dispatches on the object passed in: "PCG" -> conjugate
// gradients, "MGPCG" or the reference's "HyprePCG+PFMG" -> the multigrid
// preconditioned form, anything else -> Jacobi.
class SemiImplicitSolver : public GpuSolverBase {
public:
    SemiImplicitSolver(const RectilinearMesh&, std::shared_ptr<RiemannSolver> riemannSolver,
                       std::shared_ptr<PressureSolver> pressureSolver, std::shared_ptr<EquationOfState> eos,
                       std::shared_ptr<IGRSolver>, const SimulationConfig& config, int precision = SIFV_FP64)
        : GpuSolverBase(std::move(riemannSolver), std::move(eos), config, precision) {
        const std::string n = pressureSolver ? pressureSolver->name() : std::string("Jacobi");
        psKind_ = n == "PCG" ? SIFV_PS_PCG : (n == "MGPCG" || n == "HyprePCG+PFMG") ? SIFV_PS_MGPCG : SIFV_PS_JACOBI;
    }

    double step(const SimulationConfig& config, const RectilinearMesh& mesh, SolutionState& state,
                double targetDt = -1.0) {
        ensure(config, mesh);
        if (!psSet_) {
            if (psKind_ != SIFV_PS_JACOBI) check_status(sifv_set_pressure_solver(h_, psKind_));
            psSet_ = true;
        }
        stage_in(state);
        sifv_config c = to_c(config);
        double dt = 0.0;
        check_status(sifv_semi_implicit_step(h_, &c, targetDt, &dt, &lastIters_));
        stage_out(state, dt, targetDt);
        return dt;
    }
    int lastPressureIterations() const { return lastIters_; }

private:
    int lastIters_ = 0;
    int psKind_ = SIFV_PS_JACOBI;
    bool psSet_ = false;
};

// PressureSolver backed by the GPU Jacobi sweep (src/JacobiPressureSolver.cpp:9-108):
// drops into the reference's CPU SemiImplicitSolver through its vtable.
// The GPU pressure solvers share one handle-owning base; kind selects the
// entry point: SIFV_PS_JACOBI (sifv_jacobi_solve), SIFV_PS_PCG
// (sifv_pcg_solve) or SIFV_PS_MGPCG (sifv_mgpcg_solve).
class GpuPressureSolver : public PressureSolver {
public:
    GpuPressureSolver(int kind, const SimulationConfig& config, std::shared_ptr<EquationOfState> eos,
                      std::shared_ptr<RiemannSolver> rs)
        : kind_(kind), cfg_(config), eos_(std::move(eos)), rs_(std::move(rs)) {}
    ~GpuPressureSolver() override { sifv_destroy(h_); }

    int solve(const RectilinearMesh& mesh, const std::vector<double>& rho, const std::vector<double>& rhoc2,
              const std::vector<double>& rhs, std::vector<double>& pressure, double dt, double tolerance,
              int maxIter) override {
        ensure(mesh, nullptr);
        return run(rho, rhoc2, rhs, pressure, dt, tolerance, maxIter);
    }
    int solve(const RectilinearMesh& mesh, const std::vector<double>& rho, const std::vector<double>& rhoc2,
              const std::vector<double>& rhs, std::vector<double>& pressure, double dt, double tolerance,
              int maxIter, HaloExchange& halo) override {
        ensure(mesh, &halo);
        return run(rho, rhoc2, rhs, pressure, dt, tolerance, maxIter);
    }
    int kind() const { return kind_; }

private:
    int run(const std::vector<double>& rho, const std::vector<double>& rhoc2, const std::vector<double>& rhs,
            std::vector<double>& pressure, double dt, double tolerance, int maxIter) {
        int it = 0;
        if (kind_ == SIFV_PS_PCG)
            check_status(sifv_pcg_solve(h_, rho.data(), rhoc2.data(), rhs.data(), pressure.data(), dt, tolerance,
                                        maxIter, &it));
        else if (kind_ == SIFV_PS_MGPCG)
            check_status(sifv_mgpcg_solve(h_, rho.data(), rhoc2.data(), rhs.data(), pressure.data(), dt, tolerance,
                                          maxIter, &it));
        else
            check_status(sifv_jacobi_solve(h_, rho.data(), rhoc2.data(), rhs.data(), pressure.data(), dt,
                                           tolerance, maxIter, &it, nullptr));
        return it;
    }
    void ensure(const RectilinearMesh& mesh, HaloExchange* halo) {
        if (h_) return;
        std::vector<double> x(mesh.nx() + 1), y(mesh.ny() + 1), z(mesh.nz() + 1);
        for (int i = 0; i <= mesh.nx(); ++i) x[i] = mesh.nodeX(i);
        for (int j = 0; j <= mesh.ny(); ++j) y[j] = mesh.nodeY(j);
        for (int k = 0; k <= mesh.nz(); ++k) z[k] = mesh.nodeZ(k);
        sifv_mesh m{};
        m.dim = mesh.dim();
        m.nx = mesh.nx();
        m.ny = mesh.ny();
        m.nz = mesh.nz();
        m.nGhost = mesh.nGhost();
        m.xNodes = x.data();
        m.yNodes = y.data();
        m.zNodes = z.data();
        for (int f = 0; f < 6; ++f) m.bc[f] = static_cast<int>(mesh.boundaryCondition(f));
        for (int d = 0; d < 3; ++d)
            m.periods[d] = halo && d < mesh.dim() && !halo->mpi().isPhysicalBoundary(2 * d) ? 1 : 0;
        sifv_config c = to_c(cfg_);
        sifv_eos e{};
        e.kind = eos_->pInf() != 0.0 ? SIFV_EOS_STIFFENED : SIFV_EOS_IDEAL;
        e.gamma = eos_->gamma();
        e.pInf = eos_->pInf();
        const std::string n = rs_->name();
        int rt = n == "HLLC" ? SIFV_HLLC : (n == "Rusanov" ? SIFV_RUSANOV : SIFV_LF);
        check_status(sifv_create(&m, &c, rt, &e, SIFV_FP64, &h_));
    }
    int kind_;
    SimulationConfig cfg_;
    std::shared_ptr<EquationOfState> eos_;
    std::shared_ptr<RiemannSolver> rs_;
    sifv_handle* h_ = nullptr;
};

// Drop-in for JacobiPressureSolver (src/JacobiPressureSolver.cpp:9-108).
class JacobiPressureSolver : public GpuPressureSolver {
public:
    JacobiPressureSolver(const SimulationConfig& config, std::shared_ptr<EquationOfState> eos,
                         std::shared_ptr<RiemannSolver> rs)
        : GpuPressureSolver(SIFV_PS_JACOBI, config, std::move(eos), std::move(rs)) {}
    std::string name() const override { return "Jacobi"; }
};

// In the HyprePressureSolver role (src/HyprePressureSolver.cpp:312-408):
// conjugate gradients with a diagonal preconditioner.
class PCGPressureSolver : public GpuPressureSolver {
public:
    PCGPressureSolver(const SimulationConfig& config, std::shared_ptr<EquationOfState> eos,
                      std::shared_ptr<RiemannSolver> rs)
        : GpuPressureSolver(SIFV_PS_PCG, config, std::move(eos), std::move(rs)) {}
    std::string name() const override { return "PCG"; }
};

// In the HyprePressureSolver role with its StructPCG + PFMG pairing
// (src/HyprePressureSolver.cpp:123-138): conjugate gradients preconditioned
// by one geometric-multigrid V-cycle.
class MGPCGPressureSolver : public GpuPressureSolver {
public:
    MGPCGPressureSolver(const SimulationConfig& config, std::shared_ptr<EquationOfState> eos,
                        std::shared_ptr<RiemannSolver> rs)
        : GpuPressureSolver(SIFV_PS_MGPCG, config, std::move(eos), std::move(rs)) {}
    std::string name() const override { return "MGPCG"; }
};

}  // namespace gpu
}  // namespace SemiImplicitFV

#endif  // SIFV_GPU_SOLVERS_HPP
