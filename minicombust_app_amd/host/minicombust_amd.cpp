// minicombust_amd.cpp -- the flow-solver half of the reference's driver
// (src/minicombust.cpp:55-330) on one MI355X: the same environment variables,
// the same calls on FlowSolver, the same timing report.  Particle ranks are
// outside this build, so MINICOMBUST_FRANKS/PRANKS/PARTICLES are ignored.
//
//   MINICOMBUST_CELLS        box modifier m (2m x m x m cells)       (:78-79)
//   MINICOMBUST_ITERS        timesteps                               (:69-70)
//   MINICOMBUST_OUTPUT_ITER  output every k timesteps, -1 = never    (:72-73)
//   MINICOMBUST_MODE         compat (default) | simple  (pressure semantics)
//   MINICOMBUST_PSOLVE       fp64 (default) | mixed     (pressure-solve form)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "FlowSolver.hpp"

static long env_long(const char *name, long dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? strtol(v, nullptr, 10) : dflt;
}

int main()
{
    const uint64_t modifier = (uint64_t)env_long("MINICOMBUST_CELLS", 20);
    const uint64_t ntimesteps = (uint64_t)env_long("MINICOMBUST_ITERS", 10);
    const int64_t output_iteration = env_long("MINICOMBUST_OUTPUT_ITER", -1);
    const char *mode = getenv("MINICOMBUST_MODE");
    const char *psolve = getenv("MINICOMBUST_PSOLVE");

    mc_options opt;
    mc_options_default(&opt);
    if (mode && !strcmp(mode, "simple")) opt.mode = MC_MODE_SIMPLE;
    if (psolve && !strcmp(psolve, "mixed")) opt.psolve = MC_PSOLVE_MIXED;
    const double delta = 1.0e-5;   // src/minicombust.cpp:120

    printf("Starting miniCOMBUST (MI355X flow solver)..\n");
    printf("Mesh: box modifier %llu -> %llu cells; %llu timesteps; mode %s; pressure %s\n",
           (unsigned long long)modifier, (unsigned long long)(2 * modifier * modifier * modifier),
           (unsigned long long)ntimesteps, opt.mode == MC_MODE_SIMPLE ? "simple" : "compat",
           opt.psolve == MC_PSOLVE_MIXED ? "mixed" : "fp64");

    const auto t_setup = std::chrono::steady_clock::now();
    minicombust_amd::FlowSolver<double> flow_solver(modifier, delta, stdout, &opt);
    const double setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_setup).count();

    double output_time = 0.0;
    const auto t_prog = std::chrono::steady_clock::now();
    for (uint64_t t = 0; t < ntimesteps; t++) {
        flow_solver.timestep();
        if (output_iteration > 0 && (int64_t)(t % output_iteration) == output_iteration - 1) {
            const auto t0 = std::chrono::steady_clock::now();
            flow_solver.output_data(t + 1);
            output_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    const double program_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_prog).count();
    flow_solver.print_logger_stats(ntimesteps, program_time);
    printf("Setup Time:            %.6fs\n", setup_time);
    printf("Program Time:          %.6fs\n", program_time);
    printf("Output Time:           %.6fs\n", output_time);
    printf("Flow Compute Time:     %.6fs\n", flow_solver.compute_time);
    return 0;
}
