// Drop-in check: the reference's own classes and time loop driving the GPU
// solvers through include/sifv_gpu_solvers.hpp.
//   1. Sod tube (WENO5 + HLLC + RK3, CFL 0.6, t = 0.2) stepped by the
//      reference's CPU ExplicitSolver and by gpu::ExplicitSolver, both through
//      the reference's runTimeLoop-style targetDt clamping;
//   2. prints the max relative difference of rho, rhoU, rhoE.
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
    MPI_Finalize();
    return err <= 1e-12 ? 0 : 1;
}
