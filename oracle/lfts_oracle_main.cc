// ORACLE — test infrastructure only: command-line front end of the CPU restatement,
// mirroring `lfts-cpu <inputFile>` (src/fts_cpu.cc:20-72). Used as the CPU baseline
// and to produce reference-format output files for comparison with the HIP path.
#include <chrono>
#include <cstdio>
#include <iostream>
#include <string>

extern "C" {
void* orc_sim_create(const char* inputFile, const char* outdir, int write_files);
void orc_sim_destroy(void*);
void orc_sim_equilibrate(void*);
void orc_sim_statistics(void*);
double orc_sim_getH(void*);
}

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cout << "Please supply an input file name as first command-line argument." << std::endl << std::endl;
        return 1;
    }
    const char* outdir = (argc > 2) ? argv[2] : "";
    void* s = orc_sim_create(argv[1], outdir, 1);
    if (!s) {
        std::cout << "ERROR => Cannot open one of the input files." << std::endl;
        return 1;
    }
    std::cout << "Starting Equilibration..." << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    orc_sim_equilibrate(s);
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "Equlibration time = " << std::chrono::duration_cast<std::chrono::seconds>(t1 - t0).count() << "secs" << std::endl << std::endl;
    std::cout << "Starting Statistics..." << std::endl;
    t0 = std::chrono::steady_clock::now();
    orc_sim_statistics(s);
    t1 = std::chrono::steady_clock::now();
    std::cout << "Statistics time = " << std::chrono::duration_cast<std::chrono::seconds>(t1 - t0).count() << "secs" << std::endl << std::endl;
    std::cout.precision(6);
    std::cout << orc_sim_getH(s) << std::endl;
    orc_sim_destroy(s);
    return 0;
}
