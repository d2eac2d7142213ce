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

// Options beyond the reference's command line (all off by default).
struct run_options {
    std::string checkpoint = "checkpoint.lfts";  // file written at checkpoints
    long checkpoint_every = 0;                   // Langevin steps between checkpoints (0: never)
    long stop_after = 0;                         // stop, with a checkpoint, once this many steps ran
    std::string restart;                         // resume from this checkpoint
    bool vtk = false;                            // write <phi file>.vtk beside each phi file
};

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
    run_options opt_;
    int phase_ = 0;          // 0: equilibration, 1: statistics, 2: done
    long long resume_it_ = 0;
    bool stopped_ = false;

  public:
    explicit lfts_simulation(const std::string& inputFile, const run_options& opt = run_options());
    ~lfts_simulation();

    void equilibrate();
    void statistics();
    double getH();

    long long anderson_iterations() const { return nAM_total_; }
    long long langevin_steps() const { return steps_total_; }
    bool stopped() const { return stopped_; }

  private:
    void check(int rc, const char* what);
    void saveStdOutputFile(const std::string& fileName);
    void generate_field(double* w, int loadType);
    void prefetch_noise();
    double langevin_step();
    void save_fields(const std::string& tag, int it, bool with_sk);
    void save_phi(const std::string& fileName);
    void write_checkpoint(int phase, long long it);
    void restore_checkpoint(const std::string& file);
    bool after_step(int phase, long long it);
};
