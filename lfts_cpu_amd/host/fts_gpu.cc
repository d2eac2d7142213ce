// `lfts-gpu <inputFile>`: drop-in for the reference's `lfts-cpu <inputFile>`
// (src/fts_cpu.cc:20-72) with the Langevin step on the GPU.
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

#include "lfts_simulation.h"

// Options after the input file (none of them changes the arithmetic):
//   --checkpoint FILE      checkpoint file name (default checkpoint.lfts)
//   --checkpoint-every K   write a lossless checkpoint every K Langevin steps
//   --stop-after K         stop with a checkpoint once K Langevin steps have run in total
//   --restart FILE         resume from a checkpoint
//   --vtk                  also write <phi file>.vtk for each phi file
int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cout << "Please supply an input file name as first command-line argument." << std::endl << std::endl;
        std::exit(1);
    }
    std::string inputFile(argv[1]);
    run_options opt;
    for (int a = 2; a < argc; a++) {
        const std::string s(argv[a]);
        const bool more = a + 1 < argc;
        if (s == "--vtk") opt.vtk = true;
        else if (s == "--checkpoint" && more) opt.checkpoint = argv[++a];
        else if (s == "--checkpoint-every" && more) opt.checkpoint_every = std::atol(argv[++a]);
        else if (s == "--stop-after" && more) opt.stop_after = std::atol(argv[++a]);
        else if (s == "--restart" && more) opt.restart = argv[++a];
        else {
            std::cout << "ERROR => unknown option " << s << std::endl;
            std::exit(1);
        }
    }

    lfts_simulation* sim = new lfts_simulation(inputFile, opt);

    std::cout << "Starting Equilibration..." << std::endl;
    auto start = std::chrono::steady_clock::now();
    sim->equilibrate();
    auto end = std::chrono::steady_clock::now();
    auto dur = std::chrono::duration_cast<std::chrono::seconds>(end - start);
    std::cout << "Equlibration time = " << dur.count() << "secs" << std::endl << std::endl;

    std::cout << "Starting Statistics..." << std::endl;
    start = std::chrono::steady_clock::now();
    sim->statistics();
    end = std::chrono::steady_clock::now();
    dur = std::chrono::duration_cast<std::chrono::seconds>(end - start);
    std::cout << "Statistics time = " << dur.count() << "secs" << std::endl << std::endl;

    if (sim->stopped()) {
        std::cout << "Stopped after " << sim->langevin_steps() << " Langevin steps; checkpoint written to "
                  << opt.checkpoint << std::endl;
        delete sim;
        return 0;
    }
    std::cout.precision(6);
    std::cout << sim->getH() << std::endl;

    delete sim;
    return 0;
}
