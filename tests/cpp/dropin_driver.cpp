// Drop-in check: the reference's own classes and time loop driving the GPU
// solvers through include/sifv_gpu_solvers.hpp.
//   1. Sod tube (WENO5 + HLLC + RK3, CFL 0.6, t = 0.2) stepped by the
//      reference's CPU ExplicitSolver and by gpu::ExplicitSolver, both through
//      the reference's runTimeLoop-style targetDt clamping;
//   2. prints the max relative difference of rho, rhoU, rhoE;
//   3. a 2D pressure pulse (ideal gas, 64^2, acoustic CFL ~4) stepped by the
//      reference's CPU SemiImplicitSolver with its CPU JacobiPressureSolver,
//      by the same CPU SemiImplicitSolver with gpu::MGPCGPressureSolver
//      plugged in through the PressureSolver vtable, and by
//      gpu::SemiImplicitSolver given a gpu::MGPCGPressureSolver (both solves
//      tight), and prints the max relative differences of the three states.
// Built by oracle/Makefile (target dropin) against the reference sources.
#include "ExplicitSolver.hpp"
#include "HLLCSolver.hpp"
#include "HaloExchange.hpp"
#include "IdealGasEOS.hpp"
#include "JacobiPressureSolver.hpp"
#include "MPIContext.hpp"
#include "RectilinearMesh.hpp"
#include "SemiImplicitSolver.hpp"
#include "SimulationConfig.hpp"
#include "SolutionState.hpp"
#include "sifv_gpu_solvers.hpp"

#include <mpi.h>

#include <cmath>
#include <cstdio>

using namespace SemiImplicitFV;

static void sod(const RectilinearMesh& mesh, SolutionState& st, const IdealGasEOS& eos) {
    for (int i = 0; i < mesh.nx(); ++i) {
        std::size_t id = mesh.index(i, 0, 0);
        PrimitiveState W;
        bool left = mesh.cellCentroidX(i) < 0.5;
        W.rho = left ? 1.0 : 0.125;
        W.p = left ? 1.0 : 0.1;
        st.setPrimitiveState(id, W);
        st.setConservativeState(id, eos.toConservative(W));
    }
}

// Part 3: semi-implicit pressure pulse; returns the largest relative
// difference between the three runs.
static double semi_pulse(int, char**) {
    SimulationConfig cfg;
    cfg.dim = 2;
    cfg.nGhost = 3;
    cfg.RKOrder = 3;
    cfg.reconOrder = ReconstructionOrder::WENO5;
    cfg.semiImplicit = true;
    cfg.semiImplicitParams.maxDt = 0.05;
    cfg.semiImplicitParams.pressureTol = 1e-12;
    cfg.semiImplicitParams.maxPressureIters = 20000;
    cfg.validate();
    const int n = 64;
    std::vector<double> xn(n + 1);
    for (int i = 0; i <= n; ++i) xn[i] = double(i) / n;
    MPIContext mpi = MPIContext::create(n, n, 1, xn, xn, {0.0, 1.0}, 2, {0, 0, 0});
    RectilinearMesh mesh(cfg, mpi.localXNodes(), mpi.localYNodes(), mpi.localZNodes());
    HaloExchange halo(mpi, mesh);
    auto eos = std::make_shared<IdealGasEOS>(1.4, 287.0, cfg);
    auto rs = std::make_shared<HLLCSolver>(eos, cfg);
    SolutionState s0;
    s0.allocate(mesh.totalCells(), cfg);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double x = mesh.cellCentroidX(i) - 0.5, y = mesh.cellCentroidY(j) - 0.5;
            PrimitiveState W;
            W.rho = 1.0;
            W.p = 1.0 + 0.1 * std::exp(-(x * x + y * y) / 0.01);
            std::size_t id = mesh.index(i, j, 0);
            s0.setPrimitiveState(id, W);
            s0.setConservativeState(id, eos->toConservative(W));
        }
    SolutionState a = s0, b = s0, c = s0;
    SemiImplicitSolver cpuJ(mesh, rs, std::make_shared<JacobiPressureSolver>(), eos, nullptr, cfg);
    cpuJ.setHaloExchange(&halo);
    auto mg = std::make_shared<gpu::MGPCGPressureSolver>(cfg, eos, rs);
    SemiImplicitSolver cpuM(mesh, rs, mg, eos, nullptr, cfg);
    cpuM.setHaloExchange(&halo);
    gpu::SemiImplicitSolver dev(mesh, rs, std::make_shared<gpu::MGPCGPressureSolver>(cfg, eos, rs), eos, nullptr,
                                cfg);
    dev.setHaloExchange(&halo);
    int itJ = 0, itM = 0, itD = 0;
    for (int k = 0; k < 5; ++k) {
        cfg.step = k;
        cpuJ.step(cfg, mesh, a);
        itJ = cpuJ.lastPressureIterations();
        cpuM.step(cfg, mesh, b);
        itM = cpuM.lastPressureIterations();
        dev.step(cfg, mesh, c);
        itD = dev.lastPressureIterations();
    }
    double eM = 0.0, eD = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            std::size_t id = mesh.index(i, j, 0);
            eM = std::fmax(eM, std::fabs(a.rhoE[id] - b.rhoE[id]) / std::fabs(a.rhoE[id]));
            eD = std::fmax(eD, std::fabs(a.rhoE[id] - c.rhoE[id]) / std::fabs(a.rhoE[id]));
            eM = std::fmax(eM, std::fabs(a.rhoU[id] - b.rhoU[id]) / 0.1);
            eD = std::fmax(eD, std::fabs(a.rhoU[id] - c.rhoU[id]) / 0.1);
        }
    std::printf("{\"semi_iters_cpu_jacobi\": %d, \"semi_iters_cpu_gpu_mgpcg\": %d, \"semi_iters_gpu_mgpcg\": %d, "
                "\"max_rel_diff_cpu_solver_gpu_ps\": %.3e, \"max_rel_diff_gpu_solver\": %.3e}\n",
                itJ, itM, itD, eM, eD);
    return std::fmax(eM, eD);
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    SimulationConfig cfg;
    cfg.dim = 1;
    cfg.nGhost = 3;
    cfg.RKOrder = 3;
    cfg.reconOrder = ReconstructionOrder::WENO5;
    cfg.explicitParams.cfl = 0.6;
    cfg.validate();
    const int n = 1024;
    std::vector<double> xn(n + 1);
    for (int i = 0; i <= n; ++i) xn[i] = 0.0 + (1.0 - 0.0) * i / n;
    MPIContext mpi = MPIContext::create(n, 1, 1, xn, {0.0, 1.0}, {0.0, 1.0}, 1, {0, 0, 0});
    RectilinearMesh mesh(cfg, mpi.localXNodes(), mpi.localYNodes(), mpi.localZNodes());
    HaloExchange halo(mpi, mesh);
    auto eos = std::make_shared<IdealGasEOS>(1.4, 287.0, cfg);
    auto rs = std::make_shared<HLLCSolver>(eos, cfg);
    SolutionState a, b;
    a.allocate(mesh.totalCells(), cfg);
    b.allocate(mesh.totalCells(), cfg);
    sod(mesh, a, *eos);
    sod(mesh, b, *eos);

    ExplicitSolver cpu(mesh, rs, eos, nullptr, cfg);
    cpu.setHaloExchange(&halo);
    gpu::ExplicitSolver dev(mesh, rs, eos, nullptr, cfg);
    dev.setHaloExchange(&halo);

    double ta = 0.0, tb = 0.0, T = 0.2;
    int steps = 0;
    while (ta < T) {
        ta += cpu.step(cfg, mesh, a, T - ta);
        tb += dev.step(cfg, mesh, b, T - tb);
        cfg.step++;
        steps++;
    }
    double err = 0.0;
    for (int i = 0; i < mesh.nx(); ++i) {
        std::size_t id = mesh.index(i, 0, 0);
        err = std::fmax(err, std::fabs(a.rho[id] - b.rho[id]) / 1.0);
        err = std::fmax(err, std::fabs(a.rhoU[id] - b.rhoU[id]) / 1.0);
        err = std::fmax(err, std::fabs(a.rhoE[id] - b.rhoE[id]) / 2.5);
    }
    std::printf("{\"steps\": %d, \"time_cpu\": %.17g, \"time_gpu\": %.17g, \"max_rel_diff\": %.3e}\n", steps, ta, tb,
                err);
    const double errSemi = semi_pulse(argc, argv);
    MPI_Finalize();
    return (err <= 1e-12 && errSemi <= 1e-8) ? 0 : 1;
}
