// Simulation parameters read from the input file header (src/lfts_params.h:23-93).
#pragma once
#include <string>

#include "lfts_gpu.h"

class lfts_params {
    int N_ = 0, NA_ = 0, NB_ = 0;
    double chi_b_ = 0, C_ = 0, dt_ = 0;
    int isXeN_ = 0;
    int m_[3] = {0, 0, 0};
    double L_[3] = {0, 0, 0};
    int equil_its_ = 0, sim_its_ = 0, sample_freq_ = 0, save_freq_ = 0, loadType_ = 0;
    long M_ = 0, Mk_ = 0;
    double V_ = 0, sigma_ = 0, XeN_ = 0;
    bool ok_ = false;

  public:
    explicit lfts_params(const std::string& inputFile);

    bool ok() const { return ok_; }
    void outputParameters() const;
    void saveOutputParams(const std::string& fileName, bool append = false) const;

    int N() const { return N_; }
    int NA() const { return NA_; }
    int NB() const { return NB_; }
    double XbN() const { return chi_b_; }
    double XeN() const { return XeN_; }
    double C() const { return C_; }
    double dt() const { return dt_; }
    int isXeN() const { return isXeN_; }
    const int* m() const { return m_; }
    const double* L() const { return L_; }
    int equil_its() const { return equil_its_; }
    int sim_its() const { return sim_its_; }
    int sample_freq() const { return sample_freq_; }
    int save_freq() const { return save_freq_; }
    int loadType() const { return loadType_; }
    long M() const { return M_; }
    long Mk() const { return Mk_; }
    double V() const { return V_; }
    double sigma() const { return sigma_; }
    double n() const { return C_ * V_; }  // number of polymers, C*V
};

// Why lfts_create refused a run's parameters, for the drivers' error line.
inline std::string create_error_text(int rc) {
    if (rc == LFTS_EINVAL)
        return "the input cannot run on the GPU: a mesh size with a prime factor above 7 may have at most 2047 "
               "points, N >= 2, 0 <= NA <= N and box edges > 0";
    if (rc == LFTS_EHIP) return "no usable GPU (HIP error)";
    return "error code " + std::to_string(rc);
}
