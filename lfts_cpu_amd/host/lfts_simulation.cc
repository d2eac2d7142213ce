// L-FTS driver on the GPU (src/lfts_simulation.cc:8-157).
#include "lfts_simulation.h"

#include <cstdlib>
#include <iostream>

#include "field_generator.h"
#include "file_io.h"

void lfts_simulation::check(int rc, const char* what) {
    if (rc == LFTS_OK) return;
    std::cout << "ERROR => " << what << ": " << lfts_last_error(ctx_) << std::endl;
    std::exit(1);
}

// src/lfts_simulation.cc:8-57
lfts_simulation::lfts_simulation(const std::string& inputFile) {
    if (!file_IO::isValidFile(inputFile)) {
        std::cout << "ERROR => Cannot open one of the input files." << std::endl;
        std::exit(1);
    }
    std::cout << "Creating lfts_params object..." << std::endl;
    P_ = std::make_unique<lfts_params>(inputFile);
    if (!P_->ok()) {
        std::cout << "ERROR => Cannot parse the input file header (n_smpl and save_freq must be >= 1)." << std::endl;
        std::exit(1);
    }
    P_->outputParameters();
    M_ = P_->M();

    // Random number generator, seeded as the reference (src/lfts_simulation.cc:77)
    lfts_rng_create(123456789ULL, P_->sigma(), &rng_);

    w_.assign(4 * M_, 0.0);

    std::cout << "Creating diblock object..." << std::endl;
    lfts_desc d;
    d.N = P_->N();
    d.NA = P_->NA();
    for (int k = 0; k < 3; k++) {
        d.m[k] = P_->m()[k];
        d.L[k] = P_->L()[k];
    }
    d.nhMax = 10;
    int rc = lfts_create(&d, &ctx_);
    if (rc != LFTS_OK) {
        std::cout << "ERROR => cannot create the GPU context (code " << rc << ")" << std::endl;
        std::exit(1);
    }
    std::cout << "Creating anderson object..." << std::endl;

    // langevin constructor: the first M normals are the "previous" noise (src/langevin.cc:16)
    std::cout << "Creating langevin object..." << std::endl;
    eta_[0].assign(M_, 0.0);
    eta_[1].assign(M_, 0.0);
    lfts_rng_normals(rng_, eta_[0].data(), M_);
    check(lfts_langevin_init(ctx_, eta_[0].data()), "langevin init");

    // Fields: from the file, or generated (src/lfts_simulation.cc:99-103)
    if (P_->loadType() == 1) {
        std::cout << "Loading input field..." << std::endl;
        file_IO::readArray(w_.data(), inputFile, 2 * M_, 3);
    } else {
        generate_field(w_.data(), P_->loadType());
    }
    check(lfts_set_fields(ctx_, w_.data(), 2), "set fields");

    // From here on the random stream only feeds Langevin noise, so the next draw can
    // run on a host thread while the GPU works.
    prefetch_noise();

    int k = 0;
    rc = lfts_anderson_mix(ctx_, 200, 1e-4, &k, nullptr);
    if (rc != LFTS_ENUM) check(rc, "anderson mix");

    saveStdOutputFile("w_0");
    file_IO::saveArray(w_.data() + 2 * M_, "phi_0", 2 * M_);
}

lfts_simulation::~lfts_simulation() {
    if (next_eta_.valid()) next_eta_.wait();
    if (ctx_) lfts_destroy(ctx_);
    if (rng_) lfts_rng_destroy(rng_);
}

void lfts_simulation::prefetch_noise() {
    const int slot = 1 - eta_cur_;
    double* dst = eta_[slot].data();
    lfts_rng* r = rng_;
    const long M = M_;
    next_eta_ = std::async(std::launch::async, [r, dst, M]() { lfts_rng_normals(r, dst, M); });
}

// step_wm + mix + calc_concs of one Langevin step (src/lfts_simulation.cc:72-76)
double lfts_simulation::langevin_step() {
    next_eta_.wait();
    eta_cur_ = 1 - eta_cur_;
    check(lfts_langevin_step(ctx_, eta_[eta_cur_].data(), P_->XbN(), P_->dt()), "langevin step");
    prefetch_noise();
    int k = 0;
    int rc = lfts_anderson_mix(ctx_, 200, 1e-4, &k, nullptr);
    if (rc != LFTS_ENUM) check(rc, "anderson mix");
    nAM_total_ += k - 1;
    steps_total_++;
    double lnQ = 0.0;
    rc = lfts_calc_concs(ctx_, &lnQ);
    if (rc != LFTS_ENUM) check(rc, "calc_concs");
    return lnQ;
}

void lfts_simulation::save_fields(const std::string& tag, int it, bool with_sk) {
    saveStdOutputFile("w_" + tag + std::to_string(it));
    file_IO::saveArray(w_.data() + 2 * M_, "phi_" + tag + std::to_string(it), 2 * M_);
    if (with_sk) {
        std::vector<double> S(P_->Mk());
        int ns = 0;
        check(lfts_sk_get(ctx_, S.data(), &ns), "S(k) read");
        lfts_sk_save(("struct_st_" + std::to_string(it)).c_str(), P_->m(), P_->L(), S.data(), ns, P_->n(),
                     P_->XbN(), 1e-5, 8);
    }
}

// src/lfts_simulation.cc:67-87
void lfts_simulation::equilibrate() {
    int it;
    for (it = 1; it <= P_->equil_its(); it++) {
        std::cout << "lnQ = " << langevin_step() << std::endl;
        if (it % P_->save_freq() == 0) save_fields("eq_", it, false);
    }
    save_fields("eq_", it - 1, false);
}

// src/lfts_simulation.cc:91-118
void lfts_simulation::statistics() {
    int it;
    for (it = 1; it <= P_->sim_its(); it++) {
        std::cout << "lnQ = " << langevin_step() << std::endl;
        if (it % P_->sample_freq() == 0) check(lfts_sk_sample(ctx_), "S(k) sample");
        if (it % P_->save_freq() == 0) save_fields("st_", it, true);
    }
    save_fields("st_", it - 1, true);
}

// src/lfts_simulation.cc:122-135
double lfts_simulation::getH() {
    double H = 0.0;
    int rc = lfts_hamiltonian(ctx_, P_->XbN(), &H);
    if (rc != LFTS_ENUM) check(rc, "hamiltonian");
    return H;
}

// src/lfts_simulation.cc:139-142: header, then W- and w+ (fields read back from the GPU)
void lfts_simulation::saveStdOutputFile(const std::string& fileName) {
    check(lfts_get_fields(ctx_, w_.data(), 4), "get fields");
    P_->saveOutputParams(fileName);
    file_IO::saveArray(w_.data(), fileName, 2 * M_, true);
}

// src/lfts_simulation.cc:145-157
void lfts_simulation::generate_field(double* w, int loadType) {
    switch (loadType) {
        case 2:
            field_generator::create_lamellar(w, P_->XbN(), P_->m());
            break;
        default: {
            // disordered start: W- = XbN (U - 1/2), w+ = 0, U from the same random stream
            std::vector<double> u(M_);
            lfts_rng_uniforms(rng_, u.data(), M_);
            for (long r = 0; r < M_; r++) {
                w[r] = P_->XbN() * (u[r] - 0.5);
                w[r + M_] = 0.0;
            }
            break;
        }
    }
}
