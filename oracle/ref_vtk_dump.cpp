// ref_vtk_dump.cpp -- TEST INFRASTRUCTURE: writes a flow VTK file with the
// reference's own writer, VisitWriter<double>::write_flow
// (include/visit/VisitWriter.hpp:127-234, compiled from the reference tree by
// oracle/Makefile), on the reference's own box mesh at one flow rank
// (src/examples/mesh_examples.cpp, called as src/minicombust.cpp:123-137), so
// that tests can compare the library's VTK output with it byte for byte.
//
// Usage: ref_vtk_dump <modifier> <fields.bin> <prefix> <timestep>
// fields.bin (little endian): int64 n, then f64[11 n]: U V W P TE ED TEM FUL
// PRO VARF VARP, n values each.
#include "examples/mesh_examples.hpp"
#include "particles/Particle.hpp"
// VisitWriter.hpp's particle writers name Particle unqualified, as the
// reference's own translation units see it after ParticleSolver.hpp
using namespace minicombust::particles;
#include "visit/VisitWriter.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace minicombust::geometry;
using namespace minicombust::utils;
using namespace minicombust::visit;

int main(int argc, char **argv)
{
    if (argc < 5) { fprintf(stderr, "usage: %s <modifier> <fields.bin> <prefix> <timestep>\n", argv[0]); return 2; }
    const uint64_t modifier = strtoull(argv[1], nullptr, 10);
    MPI_Init(&argc, &argv);
    MPI_Config cfg = {};
    cfg.world = MPI_COMM_WORLD;
    MPI_Comm_rank(cfg.world, &cfg.rank);
    MPI_Comm_size(cfg.world, &cfg.world_size);
    cfg.solver_type = FLOW;
    MPI_Comm_split(cfg.world, cfg.solver_type, cfg.rank, &cfg.particle_flow_world);
    MPI_Comm_rank(cfg.particle_flow_world, &cfg.particle_flow_rank);
    MPI_Comm_size(cfg.particle_flow_world, &cfg.particle_flow_world_size);
    vec<double> box_dim = {0.05 * modifier * 2, 0.05 * modifier, 0.05 * modifier};
    vec<uint64_t> elements_per_dim = {modifier * 2, modifier * 1, modifier * 1};
    FILE *devnull = fopen("/dev/null", "w");
    Mesh<double> *mesh = load_mesh(&cfg, box_dim, elements_per_dim, 1, devnull);

    FILE *fp = fopen(argv[2], "rb");
    if (!fp) { perror("fopen"); return 1; }
    int64_t n = 0;
    if (fread(&n, sizeof(n), 1, fp) != 1 || n != (int64_t)mesh->local_mesh_size) { fprintf(stderr, "bad n\n"); return 1; }
    std::vector<double> v(11 * n);
    if (fread(v.data(), sizeof(double), v.size(), fp) != v.size()) { fprintf(stderr, "short read\n"); return 1; }
    fclose(fp);
    phi_vector<double> phi = {};
    double *f = v.data();
    phi.U = f; phi.V = f + n; phi.W = f + 2 * n; phi.P = f + 3 * n; phi.TE = f + 4 * n; phi.ED = f + 5 * n;
    phi.TEM = f + 6 * n; phi.FUL = f + 7 * n; phi.PRO = f + 8 * n; phi.VARF = f + 9 * n; phi.VARP = f + 10 * n;
    VisitWriter<double> writer(mesh, &cfg);
    writer.write_flow(argv[3], atoi(argv[4]), &phi);
    MPI_Finalize();
    return 0;
}
