// `lfts-gpu-batch <inputFile> [<inputFile> ...]`: many runs of the reference's
// lfts_simulation (src/lfts_simulation.cc:8-157, driven as in src/fts_cpu.cc:20-72) in one
// process, advanced together as one replica set (one launch per contour step for every
// mesh size present). Each replica keeps the single-run behaviour: its input file in the
// reference format, its own random stream seeded 123456789 as the reference seeds every
// run, its own output files written into the directory of its input file, identical
// byte for byte to what `lfts-gpu` writes for that input alone. Each replica's next noise
// is drawn on its own host thread while the GPU runs the current step.
//
// All inputs must share equil_its, sim_its, sample_freq and save_freq (the replicas
// step in lockstep).
#include <chrono>
#include <climits>
#include <cstdlib>
#include <future>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "field_generator.h"
#include "file_io.h"
#include "lfts_gpu.h"
#include "lfts_params.h"

namespace {

struct Replica {
    std::unique_ptr<lfts_params> P;
    lfts_ctx* ctx = nullptr;
    lfts_rng* rng = nullptr;
    std::string dir;
    std::vector<double> w;
    std::vector<double> eta[2];
    std::future<void> next;
    int cur = 0;
    long M = 0;

    std::string path(const std::string& f) const { return dir + f; }
    void prefetch() {
        double* dst = eta[1 - cur].data();
        lfts_rng* r = rng;
        const long m = M;
        next = std::async(std::launch::async, [r, dst, m]() { lfts_rng_normals(r, dst, m); });
    }
};

void die(const std::string& msg) {
    std::cout << "ERROR => " << msg << std::endl;
    std::exit(1);
}

void check(int rc, lfts_ctx* c, const char* what) {
    if (rc != LFTS_OK && rc != LFTS_ENUM) die(std::string(what) + ": " + lfts_last_error(c));
}

std::string dir_of(const std::string& f) {
    const size_t p = f.find_last_of('/');
    return p == std::string::npos ? std::string() : f.substr(0, p + 1);
}

// The directory a replica writes its files into, resolved (symbolic links, "..", the
// working directory for a bare file name), so that two spellings of one directory compare
// equal. Empty if it does not exist.
std::string output_dir(const std::string& f) {
    const std::string d = dir_of(f);
    char* r = realpath(d.empty() ? "." : d.c_str(), nullptr);
    if (!r) return std::string();
    std::string out(r);
    std::free(r);
    return out;
}

// src/lfts_simulation.cc:139-142 and the phi file beside it
void save_fields(Replica& r, const std::string& tag) {
    check(lfts_get_fields(r.ctx, r.w.data(), 4), r.ctx, "get fields");
    r.P->saveOutputParams(r.path("w_" + tag));
    file_IO::saveArray(r.w.data(), r.path("w_" + tag), 2 * r.M, true);
    file_IO::saveArray(r.w.data() + 2 * r.M, r.path("phi_" + tag), 2 * r.M);
}

void save_sk(Replica& r, const std::string& name) {
    std::vector<double> S(r.P->Mk());
    int ns = 0;
    check(lfts_sk_get(r.ctx, S.data(), &ns), r.ctx, "S(k) read");
    lfts_sk_save(r.path(name).c_str(), r.P->m(), r.P->L(), S.data(), ns, r.P->n(), r.P->XbN(), 1e-5, 8);
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cout << "Please supply one or more input file names as command-line arguments." << std::endl << std::endl;
        return 1;
    }
    const int n = argc - 1;
    // Every replica writes the same file names (w_0, phi_0, w_eq_*, ...) into its input's
    // directory, so two inputs in one directory would overwrite each other's outputs:
    // refuse that before any work starts.
    {
        std::vector<std::string> dirs(n);
        for (int i = 0; i < n; i++) {
            const std::string in(argv[i + 1]);
            if (!file_IO::isValidFile(in)) die("Cannot open one of the input files.");
            dirs[i] = output_dir(in);
            for (int j = 0; j < i; j++)
                if (dirs[j] == dirs[i])
                    die("inputs " + std::string(argv[j + 1]) + " and " + in + " are in the same directory (" +
                        dirs[i] + "); each replica writes its output files into its input's directory, "
                        "so every input needs a directory of its own");
        }
    }
    std::vector<Replica> rep(n);
    for (int i = 0; i < n; i++) {
        Replica& r = rep[i];
        const std::string in(argv[i + 1]);
        if (!file_IO::isValidFile(in)) die("Cannot open one of the input files.");
        std::cout << "== replica " << i << ": " << in << std::endl;
        r.P = std::make_unique<lfts_params>(in);
        if (!r.P->ok()) die("Cannot parse the input file header of " + in);
        r.P->outputParameters();
        const lfts_params& P0 = *rep[0].P;
        if (r.P->equil_its() != P0.equil_its() || r.P->sim_its() != P0.sim_its() ||
            r.P->sample_freq() != P0.sample_freq() || r.P->save_freq() != P0.save_freq())
            die("the replicas must share equil_its, sim_its, sample_freq and save_freq");
        r.dir = dir_of(in);
        r.M = r.P->M();
        lfts_rng_create(123456789ULL, r.P->sigma(), &r.rng);  // src/lfts_simulation.cc:23
        lfts_desc d;
        d.N = r.P->N();
        d.NA = r.P->NA();
        for (int k = 0; k < 3; k++) {
            d.m[k] = r.P->m()[k];
            d.L[k] = r.P->L()[k];
        }
        d.nhMax = 10;
        if (const int rc = lfts_create(&d, &r.ctx)) die("cannot create the GPU context of " + in + ": " + create_error_text(rc));
        r.eta[0].assign(r.M, 0.0);
        r.eta[1].assign(r.M, 0.0);
        lfts_rng_normals(r.rng, r.eta[0].data(), r.M);  // src/langevin.cc:16
        check(lfts_langevin_init(r.ctx, r.eta[0].data()), r.ctx, "langevin init");
        r.w.assign(4 * r.M, 0.0);
        if (r.P->loadType() == 1) {
            file_IO::readArray(r.w.data(), in, 2 * r.M, 3);
        } else if (r.P->loadType() == 2) {
            field_generator::create_lamellar(r.w.data(), r.P->XbN(), r.P->m());
        } else {
            std::vector<double> u(r.M);
            lfts_rng_uniforms(r.rng, u.data(), r.M);  // src/lfts_simulation.cc:155-157
            for (long p = 0; p < r.M; p++) {
                r.w[p] = r.P->XbN() * (u[p] - 0.5);
                r.w[p + r.M] = 0.0;
            }
        }
        check(lfts_set_fields(r.ctx, r.w.data(), 2), r.ctx, "set fields");
        r.prefetch();
    }
    std::vector<lfts_ctx*> ctxs(n);
    for (int i = 0; i < n; i++) ctxs[i] = rep[i].ctx;
    lfts_batch* b = nullptr;
    if (lfts_batch_create(ctxs.data(), n, &b) != LFTS_OK) die("cannot create the replica set");
    std::vector<int> k(n);
    std::vector<double> err(n), lnQ(n), XbN(n), dt(n);
    for (int i = 0; i < n; i++) {
        XbN[i] = rep[i].P->XbN();
        dt[i] = rep[i].P->dt();
    }
    int rc = lfts_batch_anderson_mix(b, 200, 1e-4, k.data(), err.data());
    if (rc != LFTS_OK && rc != LFTS_ENUM) die(lfts_batch_last_error(b));
    for (auto& r : rep) save_fields(r, "0");

    std::vector<const double*> eta(n);
    auto step = [&]() {
        for (int i = 0; i < n; i++) {
            Replica& r = rep[i];
            r.next.wait();
            r.cur = 1 - r.cur;
            eta[i] = r.eta[r.cur].data();
        }
        const int rc2 = lfts_batch_full_step(b, eta.data(), XbN.data(), dt.data(), 200, 1e-4, lnQ.data(), k.data());
        if (rc2 != LFTS_OK && rc2 != LFTS_ENUM) die(lfts_batch_last_error(b));
        for (auto& r : rep) r.prefetch();  // the copies of this step's noise are done
    };
    const lfts_params& P = *rep[0].P;

    std::cout << "Starting Equilibration..." << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    int it;
    for (it = 1; it <= P.equil_its(); it++) {
        step();
        for (int i = 0; i < n; i++) std::cout << "[" << i << "] lnQ = " << lnQ[i] << std::endl;
        if (it % P.save_freq() == 0)
            for (auto& r : rep) save_fields(r, "eq_" + std::to_string(it));
    }
    for (auto& r : rep) save_fields(r, "eq_" + std::to_string(it - 1));
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "Equlibration time = " << std::chrono::duration_cast<std::chrono::seconds>(t1 - t0).count() << "secs"
              << std::endl << std::endl;

    std::cout << "Starting Statistics..." << std::endl;
    t0 = std::chrono::steady_clock::now();
    for (it = 1; it <= P.sim_its(); it++) {
        step();
        for (int i = 0; i < n; i++) std::cout << "[" << i << "] lnQ = " << lnQ[i] << std::endl;
        if (it % P.sample_freq() == 0)
            for (auto& r : rep) check(lfts_sk_sample(r.ctx), r.ctx, "S(k) sample");
        if (it % P.save_freq() == 0)
            for (auto& r : rep) {
                save_fields(r, "st_" + std::to_string(it));
                save_sk(r, "struct_st_" + std::to_string(it));
            }
    }
    for (auto& r : rep) {
        save_fields(r, "st_" + std::to_string(it - 1));
        save_sk(r, "struct_st_" + std::to_string(it - 1));
    }
    t1 = std::chrono::steady_clock::now();
    std::cout << "Statistics time = " << std::chrono::duration_cast<std::chrono::seconds>(t1 - t0).count() << "secs"
              << std::endl << std::endl;

    lfts_batch_destroy(b);
    std::cout.precision(6);
    for (int i = 0; i < n; i++) {
        double H = 0.0;
        check(lfts_hamiltonian(rep[i].ctx, rep[i].P->XbN(), &H), rep[i].ctx, "hamiltonian");
        std::cout << "[" << i << "] H = " << H << std::endl;
    }
    for (auto& r : rep) {
        r.next.wait();
        lfts_destroy(r.ctx);
        lfts_rng_destroy(r.rng);
    }
    return 0;
}
