// Input-file header parsing and derived parameters (src/lfts_params.cc:46-119).
#include "lfts_params.h"

#include <cmath>
#include <fstream>
#include <iostream>

#include "lfts_gpu.h"

lfts_params::lfts_params(const std::string& inputFile) {
    std::ifstream in(inputFile);
    if (in.fail()) return;
    in >> N_ >> NA_ >> chi_b_ >> C_ >> dt_ >> isXeN_;
    in >> m_[0] >> m_[1] >> m_[2] >> L_[0] >> L_[1] >> L_[2];
    in >> equil_its_ >> sim_its_ >> sample_freq_ >> save_freq_ >> loadType_;
    if (in.fail()) return;
    // The reference divides by both (src/lfts_params.cc:97-98, src/lfts_simulation.cc:79);
    // zero or negative values are rejected here instead of raising SIGFPE.
    if (sample_freq_ <= 0 || save_freq_ <= 0) return;
    // whole numbers of sampling periods (src/lfts_params.cc:97-98)
    equil_its_ = (equil_its_ / sample_freq_) * sample_freq_;
    sim_its_ = (sim_its_ / sample_freq_) * sample_freq_;
    // bare <-> effective chi through z_inf (src/lfts_params.cc:101-107)
    const double z_inf = lfts_z_inf(L_, m_, N_, C_ * C_, 100);
    if (isXeN_ == 1) {
        XeN_ = chi_b_;
        chi_b_ = chi_b_ / z_inf;
    } else {
        XeN_ = chi_b_ * z_inf;
    }
    M_ = (long)m_[0] * m_[1] * m_[2];
    Mk_ = (long)m_[0] * m_[1] * (m_[2] / 2 + 1);
    V_ = L_[0] * L_[1] * L_[2];
    sigma_ = std::sqrt(2.0 * M_ * dt_ / (C_ * V_));
    NB_ = N_ - NA_;
    ok_ = true;
}

void lfts_params::outputParameters() const {
    std::cout << "N = " << N_ << std::endl;
    std::cout << "NA = " << NA_ << std::endl;
    std::cout << "NB = " << NB_ << std::endl;
    std::cout << "chi_b = " << chi_b_ << std::endl;
    std::cout << "C = " << C_ << std::endl;
    std::cout << "dt = " << dt_ << std::endl;
    std::cout << "XeN = " << XeN_ << std::endl;
    std::cout << "isXeN = " << isXeN_ << std::endl;
    std::cout << "m[0] = " << m_[0] << std::endl;
    std::cout << "m[1] = " << m_[1] << std::endl;
    std::cout << "m[2] = " << m_[2] << std::endl;
    std::cout << "L[0] = " << L_[0] << std::endl;
    std::cout << "L[1] = " << L_[1] << std::endl;
    std::cout << "L[2] = " << L_[2] << std::endl;
    std::cout << "equil_its = " << equil_its_ << std::endl;
    std::cout << "sim_its = " << sim_its_ << std::endl;
    std::cout << "sample_freq = " << sample_freq_ << std::endl;
    std::cout << "save_freq_ = " << save_freq_ << std::endl;
    std::cout << "loadType_ = " << loadType_ << std::endl;
    std::cout << "M_ = " << M_ << std::endl;
    std::cout << "Mk_ = " << Mk_ << std::endl;
    std::cout << "V_ = " << V_ << std::endl;
    std::cout << "sigma_ = " << sigma_ << std::endl;
}

// Output-file header: loadType rewritten to 1, XN as given (src/lfts_params.cc:73-83)
void lfts_params::saveOutputParams(const std::string& fileName, bool append) const {
    const double XN_out = (isXeN_ == 1) ? XeN_ : chi_b_;
    std::ofstream out;
    if (append) out.open(fileName, std::ios_base::app);
    else out.open(fileName);
    out << N_ << " " << NA_ << " " << XN_out << " " << C_ << " " << dt_ << " " << isXeN_ << std::endl;
    out << m_[0] << " " << m_[1] << " " << m_[2] << " " << L_[0] << " " << L_[1] << " " << L_[2] << std::endl;
    out << equil_its_ << " " << sim_its_ << " " << sample_freq_ << " " << save_freq_ << " " << 1 << std::endl;
}
