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
//      tight), and prints the max relative differences of the three states;
//   4. the reference's own runTimeLoop with VTK output every 0.05 (Runtime,
//      VTKSession) driving the reference's CPU ExplicitSolver and a resident
//      gpu::ExplicitSolver over Sod to t = 0.2: final states within 1e-12 and
//      the GPU state transferred only at the four output steps;
//   5. ms per step of the 4096^2 isentropic vortex (WENO5 + HLLC + RK3)
//      through the resident drop-in, through the copying drop-in (every field
//      up and down per step) and through sifv_explicit_advance on the same
//      handle; the resident drop-in must stay within 10 % of the last.
// Built by oracle/Makefile (target dropin) against the reference sources.
#include "ExplicitSolver.hpp"
#include "HLLCSolver.hpp"
#include "HaloExchange.hpp"
#include "IdealGasEOS.hpp"
#include "JacobiPressureSolver.hpp"
#include "MPIContext.hpp"
#include "RKTimeStepping.hpp"
#include "RectilinearMesh.hpp"
#include "Runtime.hpp"
#include "VTKSession.hpp"
#include "SemiImplicitSolver.hpp"
#include "SimulationConfig.hpp"
#include "SolutionState.hpp"
#include "sifv_gpu_solvers.hpp"

#include <mpi.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>

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

// Part 4: the reference's runTimeLoop with output every 0.05 driving the CPU
// solver and a resident GPU solver over Sod-1024 to t = 0.2.
static bool resident_time_loop(Runtime& rt) {
    SimulationConfig cfg;
    cfg.dim = 1;
    cfg.nGhost = 3;
    cfg.RKOrder = 3;
    cfg.reconOrder = ReconstructionOrder::WENO5;
    cfg.explicitParams.cfl = 0.6;
    cfg.validate();
    RectilinearMesh mesh = rt.createUniformMesh(cfg, 1024, 0.0, 1.0);
    rt.setBoundaryCondition(mesh, RectilinearMesh::XLow, BoundaryCondition::Outflow);
    rt.setBoundaryCondition(mesh, RectilinearMesh::XHigh, BoundaryCondition::Outflow);
    auto eos = std::make_shared<IdealGasEOS>(1.4, 287.0, cfg);
    auto rs = std::make_shared<HLLCSolver>(eos, cfg);
    SolutionState a, b;
    a.allocate(mesh.totalCells(), cfg);
    b.allocate(mesh.totalCells(), cfg);
    sod(mesh, a, *eos);
    sod(mesh, b, *eos);
    ExplicitSolver cpu(mesh, rs, eos, nullptr, cfg);
    rt.attachSolver(cpu, mesh);
    gpu::ExplicitSolver dev(mesh, rs, eos, nullptr, cfg);
    dev.setHaloExchange(rt.haloExchange());
    dev.setResident(true);
    const std::string dir = (std::filesystem::temp_directory_path() / "sifv_dropin_vtk").string();
    TimeLoopParams params;
    params.endTime = 0.2;
    params.outputInterval = 0.05;
    params.printInterval = 1 << 30;
    SimulationConfig ca = cfg, cb = cfg;
    {
        VTKSession vtk(rt, "sod_cpu", mesh, ca, dir + "/cpu");
        runTimeLoop(rt, ca, mesh, a, vtk, [&](double t) { return cpu.step(ca, mesh, a, t); }, params);
    }
    {
        VTKSession vtk(rt, "sod_gpu", mesh, cb, dir + "/gpu");
        runTimeLoop(rt, cb, mesh, b, vtk, [&](double t) { return dev.step(cb, mesh, b, t); }, params);
    }
    double err = 0.0;
    for (int i = 0; i < mesh.nx(); ++i) {
        std::size_t id = mesh.index(i, 0, 0);
        err = std::fmax(err, std::fabs(a.rho[id] - b.rho[id]) / 1.0);
        err = std::fmax(err, std::fabs(a.rhoU[id] - b.rhoU[id]) / 1.0);
        err = std::fmax(err, std::fabs(a.rhoE[id] - b.rhoE[id]) / 2.5);
        err = std::fmax(err, std::fabs(a.pres[id] - b.pres[id]) / 1.0);
    }
    std::printf("{\"resident_loop_steps_cpu\": %d, \"resident_loop_steps_gpu\": %d, \"uploads\": %ld, "
                "\"downloads\": %ld, \"max_rel_diff\": %.3e}\n",
                ca.step, cb.step, dev.uploads(), dev.downloads(), err);
    std::filesystem::remove_all(dir);
    return err <= 1e-12 && ca.step == cb.step && dev.uploads() == 1 && dev.downloads() == 4;
}

// Part 5: ms per step of the 4096^2 vortex through the drop-in (resident and
// copying) and through sifv_explicit_advance on the same handle.
static bool resident_timing(Runtime& rt) {
    SimulationConfig cfg;
    cfg.dim = 2;
    cfg.nGhost = 3;
    cfg.RKOrder = 3;
    cfg.reconOrder = ReconstructionOrder::WENO5;
    cfg.explicitParams.cfl = 0.6;
    cfg.explicitParams.maxDt = 10.0;
    cfg.validate();
    const int n = 4096;
    RectilinearMesh mesh = rt.createUniformMesh(cfg, n, 0.0, 10.0, n, 0.0, 10.0, {1, 1, 0});
    auto eos = std::make_shared<IdealGasEOS>(1.4, 1.0, cfg);
    auto rs = std::make_shared<HLLCSolver>(eos, cfg);
    SolutionState st;
    st.allocate(mesh.totalCells(), cfg);
    const double g = 1.4, beta = 5.0, pi = 3.14159265358979323846;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double dx = mesh.cellCentroidX(i) - 5.0, dy = mesh.cellCentroidY(j) - 5.0;
            const double r2 = dx * dx + dy * dy;
            const double T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * pi * pi) * std::exp(1.0 - r2);
            const double f = std::exp(0.5 * (1.0 - r2));
            PrimitiveState W;
            W.rho = std::pow(T, 1.0 / (g - 1.0));
            W.p = std::pow(T, g / (g - 1.0));
            W.u[0] = 1.0 - beta / (2.0 * pi) * dy * f;
            W.u[1] = 1.0 + beta / (2.0 * pi) * dx * f;
            std::size_t id = mesh.index(i, j, 0);
            st.setPrimitiveState(id, W);
            st.setConservativeState(id, eos->toConservative(W));
        }
    gpu::ExplicitSolver dev(mesh, rs, eos, nullptr, cfg);
    dev.setHaloExchange(rt.haloExchange());
    dev.setResident(true);
    using clk = std::chrono::steady_clock;
    for (int k = 0; k < 3; ++k) dev.step(cfg, mesh, st);  // first call uploads
    const int K = 60;
    auto t0 = clk::now();
    for (int k = 0; k < K; ++k) dev.step(cfg, mesh, st);
    const double resident = std::chrono::duration<double, std::milli>(clk::now() - t0).count() / K;
    sifv_config c = gpu::to_c(cfg);
    double time = 0.0;
    int done = 0;
    t0 = clk::now();
    gpu::check_status(sifv_explicit_advance(dev.handle(), &c, K, 1e30, &time, &done));
    const double advance = std::chrono::duration<double, std::milli>(clk::now() - t0).count() / K;
    dev.syncToHost(st);
    dev.setResident(false);
    const int Kc = 3;
    t0 = clk::now();
    for (int k = 0; k < Kc; ++k) dev.step(cfg, mesh, st);
    const double copying = std::chrono::duration<double, std::milli>(clk::now() - t0).count() / Kc;
    bool finite = true;
    for (int j = 0; j < n && finite; j += 64)
        for (int i = 0; i < n; i += 64) finite = finite && std::isfinite(st.rho[mesh.index(i, j, 0)]);
    std::printf("{\"vortex\": %d, \"dropin_resident_ms_per_step\": %.4f, \"advance_ms_per_step\": %.4f, "
                "\"dropin_copying_ms_per_step\": %.2f, \"resident_over_advance\": %.4f, \"finite\": %d}\n",
                n, resident, advance, copying, resident / advance, finite ? 1 : 0);
    return finite && done == K && resident <= 1.10 * advance;
}

int main(int argc, char** argv) {
    Runtime rt(argc, argv);
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
    const bool loopOk = resident_time_loop(rt);
    const bool timingOk = resident_timing(rt);
    return (err <= 1e-12 && errSemi <= 1e-8 && loopOk && timingOk) ? 0 : 1;
}
