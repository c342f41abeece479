// HipManager.h -- BaseManager backed by the gfx950 HIP library (gexs_hip.h).
//
// The drop-in sibling of the reference's GpuManager
// (Managers/Accel/GPU/GPUManager.h): BaseProfile constructs it where it
// constructs MultiGpuManager (Profiles/BaseProfile.cpp:60-61) and calls the
// same virtuals.  Differences a caller can see:
//   - GetNtrans() is capped at 2^31-1 (BaseProfile holds it in an int,
//     BaseProfile.h:38) and at the device budget, never 95% of HBM / 44 B
//     per line (GPUManager.cpp:12,98-101);
//   - line data are not modified in place between (T,P) executes;
//   - errors throw gexs::Error, or print and exit(-1) like the reference's
//     CheckCudaError (compute/CUDA/cuda_utils.cu:8-17) when the manager is
//     made with exit_on_error = true.
// Conventions the reference leaves implicit are explicit setters (cutoff,
// width convention, pressure scaling, precision).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

#include "BaseManager.h"

struct gexs_ctx;

namespace gexs {

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

struct Conventions {
    double cutoff = 0.0;     // window half width cm^-1; <= 0: min(25 P, 100) (OpenMPManager.cpp:386)
    int width_conv = 0;      // 0: 1/(nu dpwcoef) (OpenMPManager.cpp:404); 1: sqrt(ln2)/(nu dpwcoef)
    int pressure_mode = 1;   // 0: gammaL without P (OpenMPManager.cpp:398); 1: times P*ref_press
    int precision = 0;       // 0: FP32 + compensated sums; 1: plain FP64
    double gauss_hwhm = -1;  // GAUSSIAN HWHM; < 0: half_width of InitializeConstants
    int64_t mem_budget = 0;  // device bytes for lines; 0: 90% of free memory
};

}  // namespace gexs

class HipManager : public BaseManager {
   public:
    explicit HipManager(ProfileType pprofile, int device = 0, bool exit_on_error = false);
    ~HipManager() override;

    void SetConventions(const gexs::Conventions& c) { conv_ = c; }
    const gexs::Conventions& GetConventions() const { return conv_; }

    void InitializeVectors(int Npoints) override;
    void InitializeVectors(int Npoints, int num_cs) override;
    void InitializeConstants(double half_width, double temperature, double partition,
                             double dfreq, double meanmass, double pressure = 10.0,
                             double ref_temp = 296.0, double ref_press = 1.0) override;
    void TransferVectors(size_t Nener, double* h_energies, double* h_nu, double* h_aif,
                         int* h_gns, double* h_gamma = nullptr, double* h_n = nullptr) override;
    void TransferFreq(double* h_freq, double* h_intens, int N) override;
    void ExecuteCrossSection(int N, int N_ener, int start_idx) override;
    void ExecuteCrossSection(int N, int N_ener, int start_idx, double temperature,
                             double pressure, double partition) override;
    void TransferResults(double* h_freq, double* h_intens, int N) override;
    void TransferResults(double* h_freq, double* h_intens, int N, int I) override;
    void Cleanup() override;
    bool ReadyForWork() override;

    gexs_ctx* context() { return ctx_; }
    int64_t LastNeval();

   private:
    void check(int rc, const char* what);
    gexs_ctx* ctx_ = nullptr;
    bool exit_on_error_;
    gexs::Conventions conv_;
    double half_width_ = 100.0, temperature_ = 296.0, partition_ = 1.0, pressure_ = 1.0;
    int npoints_ = 0;
};
