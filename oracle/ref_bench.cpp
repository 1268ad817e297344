// ref_bench.cpp - times the UNMODIFIED reference solver (MPI ranks).
//
// TEST/MEASUREMENT INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg runs it
// (through mpiexec when several host cores are used) to time the reference
// step loop beside the GPU figure.  Built by oracle/Makefile together with
// the reference sources; never linked into the product library.
//
// Usage: ref_bench <case> <n> <steps> [riemann]
//   case = vortex2d : 2D isentropic vortex, periodic [0,10]^2, n x n cells,
//                     rho=p=u=v=1 freestream, beta=5 at (5,5),
//                     WENO5 + RK3, CFL 0.6 (same IC as the GPU bench)
//          sod1d    : Sod tube, n cells, outflow, WENO5 + RK3, CFL 0.6
//   riemann = hllc (default) | rusanov | lf
// Prints one JSON line from rank 0: interior cell-updates per second of the
// timed step loop (one untimed warm-up step first), ranks, seconds.
#include "ExplicitSolver.hpp"
#include "HLLCSolver.hpp"
#include "HaloExchange.hpp"
#include "IdealGasEOS.hpp"
#include "LFSolver.hpp"
#include "MPIContext.hpp"
#include "RectilinearMesh.hpp"
#include "RusanovSolver.hpp"
#include "SimulationConfig.hpp"
#include "SolutionState.hpp"

#include <mpi.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace SemiImplicitFV;

static std::vector<double> nodes(double a, double b, int n) {
    std::vector<double> v(n + 1);
    for (int i = 0; i <= n; ++i) v[i] = a + (b - a) * i / n;
    return v;
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc < 4) {
        if (rank == 0) std::fprintf(stderr, "usage: ref_bench <vortex2d|sod1d> <n> <steps> [hllc|rusanov|lf]\n");
        MPI_Finalize();
        return 2;
    }
    std::string cs = argv[1];
    int n = std::atoi(argv[2]);
    int steps = std::atoi(argv[3]);
    std::string rsn = argc > 4 ? argv[4] : "hllc";

    SimulationConfig cfg;
    cfg.nGhost = 3;
    cfg.RKOrder = 3;
    cfg.reconOrder = ReconstructionOrder::WENO5;
    cfg.explicitParams.cfl = 0.6;
    bool vortex = cs == "vortex2d";
    cfg.dim = vortex ? 2 : 1;
    cfg.explicitParams.maxDt = vortex ? 10.0 : 1e-2;
    cfg.validate();

    std::unique_ptr<MPIContext> mpi;
    std::vector<double> xn, yn = {0.0, 1.0};
    if (vortex) {
        xn = nodes(0.0, 10.0, n);
        yn = nodes(0.0, 10.0, n);
        mpi = std::make_unique<MPIContext>(MPIContext::create(n, n, 1, xn, yn, {0.0, 1.0}, 2, {1, 1, 0}));
    } else {
        xn = nodes(0.0, 1.0, n);
        mpi = std::make_unique<MPIContext>(MPIContext::create(n, 1, 1, xn, yn, {0.0, 1.0}, 1, {0, 0, 0}));
    }
    RectilinearMesh mesh(cfg, mpi->localXNodes(), mpi->localYNodes(), mpi->localZNodes());
    SolutionState st;
    st.allocate(mesh.totalCells(), cfg);
    auto eos = std::make_shared<IdealGasEOS>(1.4, 1.0, cfg);
    std::shared_ptr<RiemannSolver> rs;
    if (rsn == "rusanov") rs = std::make_shared<RusanovSolver>(eos, cfg);
    else if (rsn == "lf") rs = std::make_shared<LFSolver>(eos, cfg);
    else rs = std::make_shared<HLLCSolver>(eos, cfg);
    ExplicitSolver solver(mesh, rs, eos, nullptr, cfg);
    HaloExchange halo(*mpi, mesh);
    solver.setHaloExchange(&halo);

    const double g = 1.4, beta = 5.0, pi = std::acos(-1.0);
    for (int j = 0; j < mesh.ny(); ++j)
        for (int i = 0; i < mesh.nx(); ++i) {
            std::size_t id = mesh.index(i, j, 0);
            PrimitiveState W;
            if (vortex) {
                double dx = mesh.cellCentroidX(i) - 5.0, dy = mesh.cellCentroidY(j) - 5.0;
                double r2 = dx * dx + dy * dy;
                double T = 1.0 - (g - 1.0) * beta * beta / (8.0 * g * pi * pi) * std::exp(1.0 - r2);
                W.rho = std::pow(T, 1.0 / (g - 1.0));
                W.p = std::pow(T, g / (g - 1.0));
                W.u = {1.0 - beta / (2.0 * pi) * dy * std::exp(0.5 * (1.0 - r2)),
                       1.0 + beta / (2.0 * pi) * dx * std::exp(0.5 * (1.0 - r2)), 0.0};
            } else {
                bool left = mesh.cellCentroidX(i) < 0.5;
                W.rho = left ? 1.0 : 0.125;
                W.p = left ? 1.0 : 0.1;
                W.u = {0.0, 0.0, 0.0};
            }
            st.setPrimitiveState(id, W);
            st.setConservativeState(id, eos->toConservative(W));
        }
    mesh.applyBoundaryConditions(st, VarSet::PRIM, halo);

    solver.step(cfg, mesh, st, -1.0);  // warm-up
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    for (int s = 0; s < steps; ++s) solver.step(cfg, mesh, st, -1.0);
    MPI_Barrier(MPI_COMM_WORLD);
    double el = MPI_Wtime() - t0, mx = 0.0;
    MPI_Allreduce(&el, &mx, 1, MPI_DOUBLE, MPI_MAX, MPI_COMM_WORLD);
    double cells = vortex ? double(n) * n : double(n);
    if (rank == 0)
        std::printf("{\"case\": \"%s\", \"n\": %d, \"steps\": %d, \"ranks\": %d, \"riemann\": \"%s\", "
                    "\"seconds\": %.6f, \"cell_updates_per_s\": %.6e}\n",
                    cs.c_str(), n, steps, size, rsn.c_str(), mx, cells * steps / mx);
    MPI_Finalize();
    return 0;
}
