// The L-FTS driver: same construction order, random-number consumption order, stdout
// lines and output files as the reference's lfts_simulation (src/lfts_simulation.h:19-54),
// with the propagators, Anderson mixing, the Langevin update and S(k) on the GPU
// through the C ABI of include/lfts_gpu.h.
#pragma once
#include <future>
#include <memory>
#include <string>
#include <vector>

#include "lfts_gpu.h"
#include "lfts_params.h"

class lfts_simulation {
    std::unique_ptr<lfts_params> P_;
    lfts_ctx* ctx_ = nullptr;
    lfts_rng* rng_ = nullptr;
    std::vector<double> w_;          // host copy of [W- | w+ | phi- | phi+] for saves
    std::vector<double> eta_[2];     // host noise, double-buffered
    std::future<void> next_eta_;     // draw of the next step's noise, overlapped with the GPU
    int eta_cur_ = 0;
    long M_ = 0;
    long long nAM_total_ = 0, steps_total_ = 0;

  public:
    explicit lfts_simulation(const std::string& inputFile);
    ~lfts_simulation();

    void equilibrate();
    void statistics();
    double getH();

    long long anderson_iterations() const { return nAM_total_; }
    long long langevin_steps() const { return steps_total_; }

  private:
    void check(int rc, const char* what);
    void saveStdOutputFile(const std::string& fileName);
    void generate_field(double* w, int loadType);
    void prefetch_noise();
    double langevin_step();
    void save_fields(const std::string& tag, int it, bool with_sk);
};
