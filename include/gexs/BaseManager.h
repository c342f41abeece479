// BaseManager.h -- the lower plugin API of the cross-section path.
//
// Same virtual interface as the reference's Managers/BaseManager.h:6-38 (the
// interface is the drop-in contract, so its names, signatures and defaults are
// kept).  When HipManager is compiled inside the reference tree, the
// reference's own BaseManager.h is used instead of this one; the two agree
// method for method.  ProfileType values follow common/Input.h:10-17.
#pragma once
#include <cstddef>

#ifndef GEXS_HAVE_PROFILETYPE
#define GEXS_HAVE_PROFILETYPE
enum ProfileType { GAUSSIAN, DOPPLER, VOIGT, PSUEDO_VOIGT, VOIGT_QUAD };
#endif

class BaseManager {
   private:
    size_t available_memory = 0;
    size_t total_memory = 0;

   protected:
    size_t N_trans = 0;
    ProfileType profile;
    void InitializeMemory(size_t bytes) { available_memory = total_memory = bytes; }
    void TrackMemory(size_t bytes) { available_memory -= bytes; }
    void FreeMemory(size_t bytes) { available_memory += bytes; }
    size_t GetAvailableMemory() { return available_memory; }
    explicit BaseManager(ProfileType p) : profile(p) {}

   public:
    // public and virtual so that a driver may own a manager through this base
    virtual ~BaseManager() {}
    virtual void InitializeVectors(int Npoints) = 0;
    virtual void InitializeVectors(int Npoints, int num_cs) = 0;
    virtual void InitializeConstants(double half_width, double temperature, double partition,
                                     double dfreq, double meanmass, double pressure = 10.0,
                                     double ref_temp = 296.0, double ref_press = 1.0) = 0;
    virtual void TransferVectors(size_t Nener, double* h_energies, double* h_nu, double* h_aif,
                                 int* h_gns, double* h_gamma = nullptr,
                                 double* h_n = nullptr) = 0;
    virtual void TransferFreq(double* h_freq, double* h_intens, int N) = 0;
    virtual void ExecuteCrossSection(int N, int N_ener, int start_idx) = 0;
    virtual void ExecuteCrossSection(int N, int N_ener, int start_idx, double temperature,
                                     double pressure, double partition) = 0;
    virtual void TransferResults(double* h_freq, double* h_intens, int N) = 0;
    virtual void TransferResults(double* h_freq, double* h_intens, int N, int I) = 0;
    virtual void Cleanup() = 0;
    virtual bool ReadyForWork() { return true; }
    size_t GetNtrans() { return N_trans; }
};
