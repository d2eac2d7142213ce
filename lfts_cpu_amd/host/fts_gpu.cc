// `lfts-gpu <inputFile>`: drop-in for the reference's `lfts-cpu <inputFile>`
// (src/fts_cpu.cc:20-72) with the Langevin step on the GPU.
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

#include "lfts_simulation.h"

int main(int argc, char* argv[]) {
    if (argc != 2) {
        std::cout << "Please supply an input file name as first command-line argument." << std::endl << std::endl;
        std::exit(1);
    }
    std::string inputFile(argv[1]);

    lfts_simulation* sim = new lfts_simulation(inputFile);

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

    std::cout.precision(6);
    std::cout << sim->getH() << std::endl;

    delete sim;
    return 0;
}
