// ref_flow_driver.cpp -- TEST INFRASTRUCTURE: the reference's own mesh
// generator (load_mesh, src/examples/mesh_examples.cpp:120-571, compiled from
// its source under the reference tree by oracle/Makefile) feeding the
// drop-in FlowSolver (minicombust_app_amd/host/FlowSolver.hpp) the way
// src/minicombust.cpp:93-174 and :221 do at one flow rank: MPI_Config set up
// as there, `Mesh<double> *mesh = load_mesh(...)`, then
// `new FlowSolver<double>(&mpi_config, mesh, delta, stdout)` and timestep()
// calls, with `using namespace minicombust::flow` in force.  The fields after
// the last timestep go to a binary file for the tests.
//
// Usage: ref_flow_driver <modifier> <timesteps> <compat|simple> <out.bin>
// Layout (little endian): int64 N, F; then for the 11 transported fields
// (U V W P TE ED TEM FUL PRO VARF VARP) N doubles each; then the face mass
// fluxes, F doubles.
#include "examples/mesh_examples.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "FlowSolver.hpp"

using namespace minicombust::geometry;
using namespace minicombust::utils;
using namespace minicombust::flow;

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <modifier> <timesteps> <compat|simple> <out.bin>\n", argv[0]);
        return 2;
    }
    const uint64_t modifier = strtoull(argv[1], nullptr, 10);
    const int timesteps = atoi(argv[2]);
    MPI_Init(&argc, &argv);
    MPI_Config mpi_config = {};
    mpi_config.world = MPI_COMM_WORLD;
    MPI_Comm_rank(mpi_config.world, &mpi_config.rank);
    MPI_Comm_size(mpi_config.world, &mpi_config.world_size);
    mpi_config.solver_type = FLOW;   // one flow rank (src/minicombust.cpp:104)
    MPI_Comm_split(mpi_config.world, mpi_config.solver_type, mpi_config.rank, &mpi_config.particle_flow_world);
    MPI_Comm_rank(mpi_config.particle_flow_world, &mpi_config.particle_flow_rank);
    MPI_Comm_size(mpi_config.particle_flow_world, &mpi_config.particle_flow_world_size);

    const double delta = 1.0e-5;   // src/minicombust.cpp:120
    vec<double> box_dim = {0.05 * modifier * 2, 0.05 * modifier, 0.05 * modifier};
    vec<uint64_t> elements_per_dim = {modifier * 2, modifier * 1, modifier * 1};
    FILE *devnull = fopen("/dev/null", "w");
    Mesh<double> *mesh = load_mesh(&mpi_config, box_dim, elements_per_dim, 1, devnull);

    // the reference's defaults but for the pressure semantics; solves tight,
    // zero-started, so the fields can be held against the oracle's
    mc_options opt;
    mc_options_default(&opt);
    opt.mode = strcmp(argv[3], "simple") == 0 ? MC_MODE_SIMPLE : MC_MODE_COMPAT;
    opt.mom_rtol = 1e-12;
    opt.p_rtol = 1e-12;
    opt.p_maxit = 5000;
    opt.warm_start = 0;
    FlowSolver<double> *flow_solver = new FlowSolver<double>(&mpi_config, mesh, delta, stdout, &opt);
    for (int t = 0; t < timesteps; t++) flow_solver->timestep();

    FILE *fp = fopen(argv[4], "wb");
    if (!fp) { perror("fopen"); return 1; }
    const int64_t N = (int64_t)mesh->mesh_size, F = (int64_t)mesh->faces_size;
    const int64_t hdr[2] = {N, F};
    fwrite(hdr, sizeof(int64_t), 2, fp);
    const int fields[11] = {MC_U, MC_V, MC_W, MC_P, MC_TE, MC_ED, MC_TEM, MC_FUL, MC_PRO, MC_VARF, MC_VARP};
    for (int f : fields) {
        const std::vector<double> v = flow_solver->field(f);
        fwrite(v.data(), sizeof(double), (size_t)N, fp);
    }
    std::vector<double> mflux((size_t)F);
    if (mc_array_get(flow_solver->context(), MC_ARR_MFLUX, mflux.data(), F) != MC_OK) return 1;
    fwrite(mflux.data(), sizeof(double), (size_t)F, fp);
    fclose(fp);
    flow_solver->print_logger_stats((uint64_t)timesteps, flow_solver->compute_time);
    delete flow_solver;
    fclose(devnull);
    MPI_Finalize();
    return 0;
}
