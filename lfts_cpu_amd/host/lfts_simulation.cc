// L-FTS driver on the GPU (src/lfts_simulation.cc:8-157).
#include "lfts_simulation.h"

#include <cstdlib>
#include <iostream>

#include "checkpoint.h"
#include "field_generator.h"
#include "file_io.h"
#include "vtk.h"

void lfts_simulation::check(int rc, const char* what) {
    if (rc == LFTS_OK) return;
    std::cout << "ERROR => " << what << ": " << lfts_last_error(ctx_) << std::endl;
    std::exit(1);
}

// src/lfts_simulation.cc:8-57
lfts_simulation::lfts_simulation(const std::string& inputFile, const run_options& opt) : opt_(opt) {
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

    // Random number generator, seeded as the reference (src/lfts_simulation.cc:23)
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
        std::cout << "ERROR => cannot create the GPU context: " << create_error_text(rc) << std::endl;
        std::exit(1);
    }
    std::cout << "Creating anderson object..." << std::endl;

    // langevin constructor: the first M normals are the "previous" noise (src/langevin.cc:16)
    std::cout << "Creating langevin object..." << std::endl;
    eta_[0].assign(M_, 0.0);
    eta_[1].assign(M_, 0.0);
    if (!opt_.restart.empty()) {
        restore_checkpoint(opt_.restart);
        return;
    }
    lfts_rng_normals(rng_, eta_[0].data(), M_);
    check(lfts_langevin_init(ctx_, eta_[0].data()), "langevin init");

    // Fields: from the file, or generated (src/lfts_simulation.cc:45-49)
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
    save_phi("phi_0");
}

void lfts_simulation::save_phi(const std::string& fileName) {
    file_IO::saveArray(w_.data() + 2 * M_, fileName, 2 * M_);
    if (opt_.vtk) lfts_vtk::write_vtk_values(fileName + ".vtk", w_.data() + 2 * M_, P_->m(), false);
}

// Everything the next step depends on, at full precision (the pending noise draw is
// waited for and saved too, so the random stream continues exactly).
void lfts_simulation::write_checkpoint(int phase, long long it) {
    if (next_eta_.valid()) next_eta_.wait();
    lfts_ckpt::State st;
    st.phase = phase;
    st.it = it;
    st.steps_total = steps_total_;
    st.nAM_total = nAM_total_;
    st.N = P_->N();
    st.NA = P_->NA();
    for (int k = 0; k < 3; k++) {
        st.m[k] = P_->m()[k];
        st.L[k] = P_->L()[k];
    }
    st.w.assign(4 * M_, 0.0);
    check(lfts_get_fields(ctx_, st.w.data(), 4), "checkpoint: fields");
    st.eta_prev.assign(M_, 0.0);
    check(lfts_langevin_get_prev(ctx_, st.eta_prev.data()), "checkpoint: noise");
    st.has_next = 1;
    st.eta_next = eta_[1 - eta_cur_];
    st.S.assign(P_->Mk(), 0.0);
    int ns = 0;
    check(lfts_sk_get(ctx_, st.S.data(), &ns), "checkpoint: S(k)");
    st.nsamples = ns;
    long long n = 0;
    lfts_rng_save(rng_, nullptr, 0, &n);
    st.rng.resize((size_t)n);
    lfts_rng_save(rng_, &st.rng[0], n, &n);
    lfts_slab_shapes(nullptr, 0, &n);
    st.shapes.resize((size_t)n);
    if (n) lfts_slab_shapes(&st.shapes[0], n, &n);
    if (!lfts_ckpt::save(opt_.checkpoint, st)) {
        std::cout << "ERROR => cannot write the checkpoint " << opt_.checkpoint << std::endl;
        std::exit(1);
    }
}

void lfts_simulation::restore_checkpoint(const std::string& file) {
    lfts_ckpt::State st;
    if (!lfts_ckpt::load(file, st)) {
        std::cout << "ERROR => cannot read the checkpoint " << file << std::endl;
        std::exit(1);
    }
    bool same = st.N == P_->N() && st.NA == P_->NA() && (long)st.w.size() == 4 * M_ &&
                (long)st.eta_prev.size() == M_ && (long)st.S.size() == P_->Mk();
    for (int k = 0; k < 3; k++) same = same && st.m[k] == P_->m()[k] && st.L[k] == P_->L()[k];
    if (!same) {
        std::cout << "ERROR => the checkpoint " << file << " belongs to another system" << std::endl;
        std::exit(1);
    }
    std::cout << "Restarting from checkpoint " << file << " (Langevin step " << st.steps_total << ")" << std::endl;
    if (lfts_rng_load(rng_, st.rng.data(), (long long)st.rng.size()) != LFTS_OK) {
        std::cout << "ERROR => the checkpoint's random-stream state is unreadable" << std::endl;
        std::exit(1);
    }
    // the factorisations the stopped run used (fixed before this run's first evaluation)
    if (!st.shapes.empty()) setenv("LFTS_SLAB_SHAPES", st.shapes.c_str(), 1);
    check(lfts_langevin_init(ctx_, st.eta_prev.data()), "restart: noise");
    w_ = st.w;
    check(lfts_set_fields(ctx_, w_.data(), 4), "restart: fields");
    check(lfts_sk_set(ctx_, st.S.data(), st.nsamples), "restart: S(k)");
    steps_total_ = st.steps_total;
    nAM_total_ = st.nAM_total;
    phase_ = st.phase;
    resume_it_ = st.it;
    eta_cur_ = 0;
    if (st.has_next && (long)st.eta_next.size() == M_) {
        eta_[1] = st.eta_next;
        std::promise<void> ready;
        ready.set_value();
        next_eta_ = ready.get_future();
    } else {
        prefetch_noise();
    }
}

// Checkpoint and stop rules after each Langevin step; true when the run stops here.
bool lfts_simulation::after_step(int phase, long long it) {
    const bool stop = opt_.stop_after > 0 && steps_total_ >= opt_.stop_after;
    if (stop || (opt_.checkpoint_every > 0 && steps_total_ % opt_.checkpoint_every == 0)) write_checkpoint(phase, it);
    if (stop) stopped_ = true;
    return stop;
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
    save_phi("phi_" + tag + std::to_string(it));
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
    if (phase_ > 0 || stopped_) return;
    int it;
    for (it = (int)resume_it_ + 1; it <= P_->equil_its(); it++) {
        std::cout << "lnQ = " << langevin_step() << std::endl;
        if (it % P_->save_freq() == 0) save_fields("eq_", it, false);
        if (after_step(0, it)) return;
    }
    save_fields("eq_", it - 1, false);
    phase_ = 1;
    resume_it_ = 0;
}

// src/lfts_simulation.cc:91-118
void lfts_simulation::statistics() {
    if (phase_ > 1 || stopped_) return;
    int it;
    for (it = (int)resume_it_ + 1; it <= P_->sim_its(); it++) {
        std::cout << "lnQ = " << langevin_step() << std::endl;
        if (it % P_->sample_freq() == 0) check(lfts_sk_sample(ctx_), "S(k) sample");
        if (it % P_->save_freq() == 0) save_fields("st_", it, true);
        if (after_step(1, it)) return;
    }
    save_fields("st_", it - 1, true);
    phase_ = 2;
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
