// main.cpp -- mpm_v3_hip: run a configuration file with the HIP solver.
//
//   mpm_v3_hip -c CFGFILE [--steps K] [--write-points PREFIX] [--check]
//
// -c as the reference's mpm_v3 (main.cpp:26-110): configure, init, driver->run.
// --steps K        run exactly K solver steps instead of the driver's stop time
// --write-points   after the run, write each body's points as PREFIX<body name>.points
// --check          configure and init only (no GPU work): validates a configuration
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "mpm_host.hpp"

using namespace mpmhost;

static void usage(const char *prog) {
    std::cout << prog << " -c CFGFILE [--steps K] [--write-points PREFIX] [--check]" << std::endl;
}

int main(int argc, char **argv) {
    std::string cfg, prefix;
    long long steps = -1;
    bool check_only = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "-c") && i + 1 < argc) cfg = argv[++i];
        else if (!std::strcmp(argv[i], "--steps") && i + 1 < argc) steps = std::atoll(argv[++i]);
        else if (!std::strcmp(argv[i], "--write-points") && i + 1 < argc) prefix = argv[++i];
        else if (!std::strcmp(argv[i], "--check")) check_only = true;
        else {
            std::cout << "Unexpected argument \"" << argv[i] << "\". Exiting." << std::endl;
            usage(argv[0]);
            return 0;
        }
    }
    if (cfg.empty()) {
        std::cout << "No arguments given. Exiting." << std::endl;
        usage(argv[0]);
        return 0;
    }
    Job job;
    Configurator config(cfg);
    if (!config.configureJob(&job)) return 0;
    job.init();
    std::cout << "Job Initialized." << std::endl;
    if (check_only) {
        for (auto &b : job.bodies)
            std::cout << "body " << b->id << " " << b->name << ": " << b->points->size() << " points" << std::endl;
        return 0;
    }
    if (steps < 0) {
        job.driver->run(&job);
    } else {
        job.driver->applyGravity(&job);
        for (long long k = 0; k < steps; k++) {
            job.solver->step(&job);
            job.t += job.dt;
        }
        job.solver->syncHost(&job);
    }
    if (!prefix.empty())
        for (auto &b : job.bodies) b->points->writeToFile(&job, prefix + b->name + ".points");
    return 0;
}
