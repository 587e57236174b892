// ref_mesh_dump.cpp -- TEST INFRASTRUCTURE: drives the reference's own mesh
// generator (src/examples/mesh_examples.cpp:120-571, compiled from its source
// under the reference tree by oracle/Makefile) at one flow rank, exactly as
// src/minicombust.cpp:123-137 calls it, and writes the mesh arrays to a binary
// file so tests can compare them with the oracle and the HIP mesh generator.
//
// Usage: ref_mesh_dump <modifier> <out.bin>
// Layout (little endian): int64 m, N, F, NP; then points f64[3NP], cells
// u64[8N], cell_neighbours u64[6N], faces u64[2F], cell_faces u64[6N],
// cells_per_point u8[NP], cell_centers f64[3N], boundary_types u64[6].
#include "examples/mesh_examples.hpp"

#include <cstdio>
#include <cstdlib>

using namespace minicombust::geometry;
using namespace minicombust::utils;

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: %s <modifier> <out.bin>\n", argv[0]); return 2; }
    const uint64_t modifier = strtoull(argv[1], nullptr, 10);
    MPI_Init(&argc, &argv);
    MPI_Config cfg = {};
    cfg.world = MPI_COMM_WORLD;
    MPI_Comm_rank(cfg.world, &cfg.rank);
    MPI_Comm_size(cfg.world, &cfg.world_size);
    cfg.solver_type = FLOW;  // one flow rank (src/minicombust.cpp:104)
    MPI_Comm_split(cfg.world, cfg.solver_type, cfg.rank, &cfg.particle_flow_world);
    MPI_Comm_rank(cfg.particle_flow_world, &cfg.particle_flow_rank);
    MPI_Comm_size(cfg.particle_flow_world, &cfg.particle_flow_world_size);

    // src/minicombust.cpp:123-124
    vec<double> box_dim = {0.05 * modifier * 2, 0.05 * modifier, 0.05 * modifier};
    vec<uint64_t> elements_per_dim = {modifier * 2, modifier * 1, modifier * 1};
    FILE *devnull = fopen("/dev/null", "w");
    Mesh<double> *mesh = load_mesh(&cfg, box_dim, elements_per_dim, 1, devnull);

    FILE *fp = fopen(argv[2], "wb");
    if (!fp) { perror("fopen"); return 1; }
    int64_t hdr[4] = {(int64_t)modifier, (int64_t)mesh->mesh_size, (int64_t)mesh->faces_size, (int64_t)mesh->points_size};
    fwrite(hdr, sizeof(int64_t), 4, fp);
    fwrite(mesh->points, sizeof(vec<double>), mesh->points_size, fp);
    fwrite(mesh->cells, sizeof(uint64_t), 8 * mesh->mesh_size, fp);
    fwrite(mesh->cell_neighbours, sizeof(uint64_t), 6 * mesh->mesh_size, fp);
    fwrite(mesh->faces, sizeof(Face<uint64_t>), mesh->faces_size, fp);
    fwrite(mesh->cell_faces, sizeof(uint64_t), 6 * mesh->mesh_size, fp);
    fwrite(mesh->cells_per_point, sizeof(uint8_t), mesh->points_size, fp);
    fwrite(mesh->cell_centers, sizeof(vec<double>), mesh->mesh_size, fp);
    fwrite(mesh->boundary_types, sizeof(uint64_t), 6, fp);
    fclose(fp);
    fclose(devnull);
    MPI_Finalize();
    return 0;
}
