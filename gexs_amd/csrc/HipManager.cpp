// HipManager.cpp -- BaseManager over the gexs C ABI (see include/gexs/HipManager.h).
#include "gexs/HipManager.h"

#include <cstdio>
#include <cstdlib>

#include "gexs_hip.h"

namespace {
int profile_id(ProfileType p) {
    switch (p) {
        case GAUSSIAN: return GEXS_GAUSSIAN;
        case DOPPLER: return GEXS_DOPPLER;
        case VOIGT: return GEXS_VOIGT;
        case VOIGT_QUAD: return GEXS_VOIGT_QUAD;
        default: return -1;
    }
}
}  // namespace

HipManager::HipManager(ProfileType pprofile, int device, bool exit_on_error)
    : BaseManager(pprofile), exit_on_error_(exit_on_error) {
    int id = profile_id(pprofile);
    if (id < 0) check(GEXS_EINVAL, "HipManager: profile not implemented");
    check(gexs_create(device, id, &ctx_), "gexs_create");
}

HipManager::~HipManager() { Cleanup(); }

void HipManager::check(int rc, const char* what) {
    if (rc == GEXS_OK) return;
    std::string msg = std::string(what) + ": " + (ctx_ ? gexs_last_error(ctx_) : "no device context");
    if (exit_on_error_) {
        std::fprintf(stderr, "HipManager error (%d) %s\n", rc, msg.c_str());
        std::exit(-1);
    }
    throw gexs::Error(rc, msg);
}

void HipManager::InitializeVectors(int Npoints) { InitializeVectors(Npoints, 1); }

void HipManager::InitializeVectors(int Npoints, int num_cs) {
    int64_t cap = 0;
    check(gexs_init_vectors(ctx_, Npoints, num_cs, conv_.mem_budget, &cap), "InitializeVectors");
    N_trans = (size_t)cap;
    npoints_ = Npoints;
    InitializeMemory((size_t)cap);
}

void HipManager::InitializeConstants(double half_width, double temperature, double partition,
                                     double dfreq, double meanmass, double pressure,
                                     double ref_temp, double ref_press) {
    half_width_ = half_width;
    temperature_ = temperature;
    partition_ = partition;
    pressure_ = pressure;
    gexs_constants k{};
    k.half_width = half_width;
    k.temperature = temperature;
    k.partition = partition;
    k.dfreq = dfreq;
    k.mean_mass = meanmass;
    k.pressure = pressure;
    k.ref_temp = ref_temp;
    k.ref_press = ref_press;
    k.cutoff = conv_.cutoff;
    k.gauss_hwhm = conv_.gauss_hwhm > 0 ? conv_.gauss_hwhm : half_width;
    k.width_conv = conv_.width_conv;
    k.pressure_mode = conv_.pressure_mode;
    k.precision = conv_.precision;
    check(gexs_init_constants(ctx_, &k), "InitializeConstants");
}

void HipManager::TransferVectors(size_t Nener, double* h_energies, double* h_nu, double* h_aif,
                                 int* h_gns, double* h_gamma, double* h_n) {
    check(gexs_transfer_lines(ctx_, (int64_t)Nener, h_energies, h_nu, h_aif, h_gns, h_gamma, h_n),
          "TransferVectors");
}

void HipManager::TransferFreq(double* h_freq, double* /*h_intens*/, int N) {
    // the reference copies h_intens (zeros) into its result grids; the device
    // grids are zeroed directly
    check(gexs_transfer_freq(ctx_, h_freq, N), "TransferFreq");
}

void HipManager::ExecuteCrossSection(int N, int N_ener, int start_idx) {
    ExecuteCrossSection(N, N_ener, start_idx, temperature_, pressure_, partition_);
}

void HipManager::ExecuteCrossSection(int N, int N_ener, int start_idx, double temperature,
                                     double pressure, double partition) {
    check(gexs_execute(ctx_, N, N_ener, start_idx, temperature, pressure, partition),
          "ExecuteCrossSection");
}

void HipManager::TransferResults(double* h_freq, double* h_intens, int N) {
    TransferResults(h_freq, h_intens, N, 0);
}

void HipManager::TransferResults(double* /*h_freq*/, double* h_intens, int N, int I) {
    // adds into h_intens, which the caller zeroes (BaseProfile.cpp:279-281), as
    // MultiGpuManager sums device grids into it (MultiGPUManager.cpp:104-126)
    check(gexs_transfer_results(ctx_, h_intens, N, I, 1), "TransferResults");
}

void HipManager::Cleanup() {
    if (ctx_) {
        gexs_destroy(ctx_);
        ctx_ = nullptr;
    }
}

bool HipManager::ReadyForWork() { return ctx_ && gexs_ready(ctx_) == 1; }

int64_t HipManager::LastNeval() {
    gexs_stats st{};
    check(gexs_last_stats(ctx_, &st), "LastNeval");
    return st.n_eval;
}
