// FlowSolver.hpp -- C++ host side of the MI355X flow solver, keeping the call
// surface of the reference's FlowSolver<T> (include/flow/FlowSolver.hpp:8-1049,
// include/flow/gpu/FlowSolver.hpp:45-1489) as src/minicombust.cpp uses it:
//
//   FlowSolver<double>(MPI_Config*, Mesh<double>*, double delta, FILE*)  (minicombust.cpp:174)
//   timestep()                                                          (:221)
//   output_data(t)                                                      (:226)
//   print_logger_stats(ntimesteps, program_time)                        (:249)
//   public compute_time                                                 (:277-282)
//
// Everything below the class goes through the C-ABI of include/mcflow.h (one
// context = one GPU).  The reference constructor takes the reference's own
// MPI_Config and Mesh<T> objects as they are (template parameters, so this
// header needs none of the reference's headers); one flow rank owns the whole
// mesh, so a mesh split over several flow ranks is refused.  Two more
// constructors take the reference's box (modifier m: 2m x m x m cells,
// generated on the device) or host arrays in the Mesh<T> layout
// (mc_mesh_desc).  The alias minicombust::flow::FlowSolver at the end makes
// the class the one src/minicombust.cpp names after its
// `using namespace minicombust::flow`.  As in the reference (gpuErrchk,
// include/flow/gpu/gpu_kernels.cu:24-32), a failed call prints the error and
// exits.
#pragma once

#include <chrono>
#include <filesystem>
#include <initializer_list>
#include <utility>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mcflow.h"

namespace minicombust_amd {

template <typename T>
class FlowSolver {
    static_assert(sizeof(T) == 8, "the flow solver computes in double precision");

  public:
    double compute_time = 0.0;   // seconds inside timestep(), as the reference's

    // FlowSolver(MPI_Config *, Mesh<T> *, double delta, FILE *)
    // (include/flow/gpu/FlowSolver.hpp:274, include/flow/FlowSolver.hpp:147),
    // as src/minicombust.cpp:174 calls it after load_mesh (:134).  The mesh
    // members read are those of include/geometry/Mesh.hpp:70-107: sizes,
    // points (vec<T>, three doubles), cells (8 vertex ids), cell_neighbours,
    // faces (Face<uint64_t>: cell0, cell1), cell_faces, boundary_types.  The
    // arrays are copied to the device here; the mesh object stays the caller's.
    template <class MPIConfig, class MeshT>
    FlowSolver(MPIConfig *mpi_config, MeshT *mesh, double delta, FILE *output_file, const mc_options *opts = nullptr)
        : out_(output_file ? output_file : stdout)
    {
        static_assert(sizeof(*mesh->points) == 3 * sizeof(double), "Mesh<T>::points must be vec<double>");
        static_assert(sizeof(*mesh->faces) == 2 * sizeof(uint64_t), "Mesh<T>::faces must be Face<uint64_t>");
        if (mpi_config && mpi_config->particle_flow_world_size != 1)
            fail("FlowSolver: one flow rank owns the whole mesh on one GPU (particle_flow_world_size must be 1)");
        if (mesh->local_mesh_size != mesh->mesh_size || mesh->cell_size != 8 || mesh->faces_per_cell != 6)
            fail("FlowSolver: the mesh must be the whole hexahedral mesh of one flow rank");
        mc_mesh_desc d{};
        d.n_cells = (int64_t)mesh->mesh_size;
        d.n_faces = (int64_t)mesh->faces_size;
        d.n_points = (int64_t)mesh->points_size;
        d.points = reinterpret_cast<const double *>(mesh->points);
        d.cells = mesh->cells;
        d.cell_neighbours = mesh->cell_neighbours;
        d.faces = reinterpret_cast<const uint64_t *>(mesh->faces);
        d.cell_faces = mesh->cell_faces;
        for (int s = 0; s < 6; s++) d.boundary_types[s] = (uint64_t)mesh->boundary_types[s];
        mc_options o = options(delta, opts);
        check(mc_ctx_create(&d, &o, &ctx_), "mc_ctx_create");
        check(mc_stage_enable(ctx_, 1), "mc_stage_enable");
    }

    // box mesh of the given modifier (src/minicombust.cpp:123-124)
    FlowSolver(uint64_t modifier, double delta, FILE *output_file, const mc_options *opts = nullptr)
        : out_(output_file ? output_file : stdout)
    {
        mc_options o = options(delta, opts);
        check(mc_ctx_create_box((int64_t)modifier, &o, &ctx_), "mc_ctx_create_box");
        check(mc_stage_enable(ctx_, 1), "mc_stage_enable");
    }

    // a host mesh in the reference Mesh<T> layout (include/geometry/Mesh.hpp:70-107)
    FlowSolver(const mc_mesh_desc &mesh, double delta, FILE *output_file, const mc_options *opts = nullptr)
        : out_(output_file ? output_file : stdout)
    {
        mc_options o = options(delta, opts);
        check(mc_ctx_create(&mesh, &o, &ctx_), "mc_ctx_create");
        check(mc_stage_enable(ctx_, 1), "mc_stage_enable");
    }

    ~FlowSolver()
    {
        if (ctx_) mc_ctx_destroy(ctx_);
    }

    FlowSolver(const FlowSolver &) = delete;
    FlowSolver &operator=(const FlowSolver &) = delete;

    // FlowSolver<T>::timestep() (FlowSolver.inl:3148-3356): three SIMPLE passes
    void timestep()
    {
        const auto t0 = std::chrono::steady_clock::now();
        check(mc_timestep(ctx_), "mc_timestep");
        check(mc_sync(ctx_), "mc_sync");
        compute_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        timesteps_++;
    }

    // FlowSolver<T>::output_data (FlowSolver.inl:18-23): the flow fields as the
    // reference's VTK file, out/flow/minicombust_flow0_timestep<t>.vtk
    void output_data(uint64_t timestep)
    {
        std::error_code ec;
        std::filesystem::create_directories("out/flow", ec);
        check(mc_write_vtk(ctx_, "out/flow/minicombust", (int64_t)timestep, 0), "mc_write_vtk");
    }

    // FlowSolver<T>::print_logger_stats (FlowSolver.inl:3079): solver work and time
    void print_logger_stats(uint64_t timesteps, double runtime)
    {
        mc_stats s{};
        check(mc_get_stats(ctx_, &s), "mc_get_stats");
        fprintf(out_, "Flow Solver Stats:\n");
        fprintf(out_, "\tTimesteps:                 %" PRIu64 "\n", timesteps);
        fprintf(out_, "\tSIMPLE iterations:         %" PRId64 "\n", s.simple_iterations);
        fprintf(out_, "\tMomentum BiCGStab iters:   %" PRId64 "\n", s.mom_iters);
        fprintf(out_, "\tScalar BiCGStab iters:     %" PRId64 "\n", s.scalar_iters);
        fprintf(out_, "\tPressure solves / iters:   %" PRId64 " / %" PRId64 "\n", s.p_solves, s.p_iters);
        fprintf(out_, "\tPressure solve time:       %.6fs\n", s.p_solve_ms * 1e-3);
        fprintf(out_, "\tAMG setups / time:         %" PRId64 " / %.6fs\n", s.amg_setups, s.amg_setup_ms * 1e-3);
        fprintf(out_, "\tMax pressure rel. residual %.3e\n", s.p_relres_max);
        fprintf(out_, "\tFlow Compute Time:         %.6fs\n", compute_time);
        fprintf(out_, "\tProgram time:              %.6fs\n", runtime);
        print_stage_times();
    }

    // the per-stage breakdown the reference prints with FLOW_SOLVER_FINE_TIME
    // (velocity: FlowSolver.inl:1768-1806; pressure and scalars alike): GPU
    // time of each stage from HIP events, seconds and share of its equation
    // group (flux and matrix set-up are one fused pass here)
    void print_stage_times()
    {
        double ms[MC_ST_COUNT] = {};
        int64_t calls[MC_ST_COUNT] = {};
        check(mc_stage_read(ctx_, ms, calls), "mc_stage_read");
        auto group = [&](const char *title, std::initializer_list<std::pair<const char *, int>> rows) {
            double tot = 0.0;
            for (auto &r : rows) tot += ms[r.second];
            fprintf(out_, "\n%s:\n", title);
            for (auto &r : rows)
                fprintf(out_, "%s: %.3f (%.2f %%)\n", r.first, ms[r.second] * 1e-3,
                        tot > 0.0 ? 100.0 * ms[r.second] / tot : 0.0);
            fprintf(out_, "Total time: %f\n", tot * 1e-3);
        };
        group("Gradient Timings", {{"Gradients (11 fields) time", MC_ST_GRAD}});
        group("Velocity Field Solve Timings", {{"Compute Flux + Matrix Setup time", MC_ST_VEL_FLUX},
                                               {"Matrix Solve time", MC_ST_VEL_SOLVE}});
        group("Pressure Solve Timings", {{"Compute Flux + Matrix Setup time", MC_ST_PRES_FLUX},
                                         {"Matrix Solve time", MC_ST_PRES_SOLVE},
                                         {"Correction time", MC_ST_PRES_CORRECT}});
        group("Turbulence (TE, ED) Solve Timings", {{"Compute Flux + Matrix Setup time", MC_ST_TURB_FLUX},
                                                   {"Matrix Solve time", MC_ST_TURB_SOLVE}});
        group("Scalar (TEM, FUL, PRO, VARF, VARP) Solve Timings",
              {{"Compute Flux + Matrix Setup time", MC_ST_SCA_FLUX}, {"Matrix Solve time", MC_ST_SCA_SOLVE}});
    }

    // host copy of a field's N cell values (field ids: include/mcflow.h)
    std::vector<double> field(int f)
    {
        std::vector<double> v(n_cells_ + 6);
        check(mc_array_get(ctx_, MC_ARR_PHI + f, v.data(), (int64_t)v.size()), "mc_array_get");
        v.resize(n_cells_);
        return v;
    }

    uint64_t cells() const { return n_cells_; }
    mc_ctx *context() { return ctx_; }

  private:
    mc_ctx *ctx_ = nullptr;
    FILE *out_;
    uint64_t n_cells_ = 0;
    uint64_t timesteps_ = 0;

    mc_options options(double delta, const mc_options *opts)
    {
        mc_options o;
        if (opts) o = *opts; else mc_options_default(&o);
        o.delta = delta;
        return o;
    }

    [[noreturn]] void fail(const char *msg)
    {
        fprintf(stderr, "%s\n", msg);
        exit(EXIT_FAILURE);
    }

    void check(int rc, const char *what)
    {
        if (rc != MC_OK) {
            fprintf(stderr, "mcflow error %d in %s: %s\n", rc, what, ctx_ ? mc_last_error(ctx_) : "");
            exit(EXIT_FAILURE);
        }
        if (!n_cells_ && ctx_) {
            int64_t sizes[4] = {0, 0, 0, 0};
            if (mc_ctx_sizes(ctx_, &sizes[0], &sizes[1], &sizes[2], &sizes[3]) == MC_OK) n_cells_ = (uint64_t)sizes[0];
        }
    }
};

}  // namespace minicombust_amd

// The name src/minicombust.cpp uses (`using namespace minicombust::flow`,
// `new FlowSolver<double>(&mpi_config, mesh, delta, stdout)` at :174).
namespace minicombust {
namespace flow {
template <typename T>
using FlowSolver = ::minicombust_amd::FlowSolver<T>;
}  // namespace flow
}  // namespace minicombust
