// Lossless checkpoint of a Langevin run (no reference counterpart: the reference resumes
// from its 6-significant-digit field files, src/file_IO.h:35 and src/lfts_params.cc:79,
// and re-seeds its random stream, src/lfts_simulation.cc:23).
//
// A checkpoint holds everything the next step depends on, at full precision: the four
// fields, the noise langevin keeps between steps, the next step's noise if already drawn,
// the random stream's state (engine and normal distribution), the S(k) accumulator and
// its sample count, the loop position and the counters. A run restarted from it prints
// and writes exactly what the uninterrupted run would have.
//
// It also holds the contour-step line factorisations the library's autotuner chose
// (lfts_slab_shapes): a restart fixes them (LFTS_SLAB_SHAPES) before its first
// evaluation, so that it rounds exactly as the run it continues.
//
// Layout (native byte order): magic "LFTSCKP2", then the fields below in order, then a
// 64-bit FNV-1a checksum of every preceding byte. Written to <file>.tmp and renamed.
// Files of the first layout ("LFTSCKP1", no factorisations) still load.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

namespace lfts_ckpt {

struct State {
    int32_t phase = 0;        // 0: equilibration, 1: statistics, 2: both done
    int64_t it = 0;           // last completed iteration of the phase
    int64_t steps_total = 0;  // Langevin steps run since the start
    int64_t nAM_total = 0;    // Anderson iterations run since the start
    int32_t N = 0, NA = 0, m[3] = {0, 0, 0};
    double L[3] = {0, 0, 0};
    std::vector<double> w;         // 4M
    std::vector<double> eta_prev;  // M
    int32_t has_next = 0;
    std::vector<double> eta_next;  // M when has_next
    std::vector<double> S;         // Mk
    int32_t nsamples = 0;
    std::string rng;               // lfts_rng_save text
    std::string shapes;            // lfts_slab_shapes text (may be empty)
};

inline uint64_t fnv1a(const unsigned char* p, size_t n, uint64_t h = 1469598103934665603ULL) {
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ULL;
    }
    return h;
}

class Writer {
    std::vector<unsigned char> buf_;

  public:
    void raw(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        buf_.insert(buf_.end(), c, c + n);
    }
    template <class T>
    void put(const T& v) { raw(&v, sizeof v); }
    void vec(const std::vector<double>& v) {
        put<int64_t>((int64_t)v.size());
        raw(v.data(), sizeof(double) * v.size());
    }
    bool save(const std::string& file) {
        const uint64_t h = fnv1a(buf_.data(), buf_.size());
        const std::string tmp = file + ".tmp";
        {
            std::ofstream o(tmp, std::ios::binary);
            if (!o) return false;
            o.write(reinterpret_cast<const char*>(buf_.data()), (std::streamsize)buf_.size());
            o.write(reinterpret_cast<const char*>(&h), sizeof h);
            if (!o) return false;
        }
        return std::rename(tmp.c_str(), file.c_str()) == 0;
    }
};

class Reader {
    std::vector<unsigned char> buf_;
    size_t pos_ = 0;
    bool ok_ = true;

  public:
    explicit Reader(const std::string& file) {
        std::ifstream in(file, std::ios::binary);
        if (!in) { ok_ = false; return; }
        buf_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
        if (buf_.size() < 16) { ok_ = false; return; }
        uint64_t h;
        std::memcpy(&h, buf_.data() + buf_.size() - 8, 8);
        buf_.resize(buf_.size() - 8);
        if (fnv1a(buf_.data(), buf_.size()) != h) ok_ = false;
    }
    bool ok() const { return ok_; }
    void raw(void* p, size_t n) {
        if (!ok_ || pos_ + n > buf_.size()) { ok_ = false; return; }
        std::memcpy(p, buf_.data() + pos_, n);
        pos_ += n;
    }
    template <class T>
    T get() { T v{}; raw(&v, sizeof v); return v; }
    std::vector<double> vec() {
        const int64_t n = get<int64_t>();
        std::vector<double> v;
        if (!ok_ || n < 0 || (size_t)n > (buf_.size() - pos_) / sizeof(double)) { ok_ = false; return v; }
        v.resize((size_t)n);
        raw(v.data(), sizeof(double) * (size_t)n);
        return v;
    }
};

inline bool save(const std::string& file, const State& s) {
    Writer w;
    w.raw("LFTSCKP2", 8);
    w.put(s.phase);
    w.put(s.it);
    w.put(s.steps_total);
    w.put(s.nAM_total);
    w.put(s.N);
    w.put(s.NA);
    for (int k = 0; k < 3; k++) w.put(s.m[k]);
    for (int k = 0; k < 3; k++) w.put(s.L[k]);
    w.vec(s.w);
    w.vec(s.eta_prev);
    w.put(s.has_next);
    w.vec(s.eta_next);
    w.vec(s.S);
    w.put(s.nsamples);
    w.put<int64_t>((int64_t)s.rng.size());
    w.raw(s.rng.data(), s.rng.size());
    w.put<int64_t>((int64_t)s.shapes.size());
    w.raw(s.shapes.data(), s.shapes.size());
    return w.save(file);
}

inline bool load(const std::string& file, State& s) {
    Reader r(file);
    char magic[8];
    r.raw(magic, 8);
    const bool v2 = std::memcmp(magic, "LFTSCKP2", 8) == 0;
    if (!r.ok() || (!v2 && std::memcmp(magic, "LFTSCKP1", 8) != 0)) return false;
    s.phase = r.get<int32_t>();
    s.it = r.get<int64_t>();
    s.steps_total = r.get<int64_t>();
    s.nAM_total = r.get<int64_t>();
    s.N = r.get<int32_t>();
    s.NA = r.get<int32_t>();
    for (int k = 0; k < 3; k++) s.m[k] = r.get<int32_t>();
    for (int k = 0; k < 3; k++) s.L[k] = r.get<double>();
    s.w = r.vec();
    s.eta_prev = r.vec();
    s.has_next = r.get<int32_t>();
    s.eta_next = r.vec();
    s.S = r.vec();
    s.nsamples = r.get<int32_t>();
    const int64_t n = r.get<int64_t>();
    if (!r.ok() || n <= 0 || n > (1 << 20)) return false;
    s.rng.resize((size_t)n);
    r.raw(&s.rng[0], (size_t)n);
    s.shapes.clear();
    if (v2) {
        const int64_t ns = r.get<int64_t>();
        if (!r.ok() || ns < 0 || ns > 4096) return false;
        s.shapes.resize((size_t)ns);
        if (ns) r.raw(&s.shapes[0], (size_t)ns);
    }
    return r.ok();
}

}  // namespace lfts_ckpt
