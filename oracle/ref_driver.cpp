// ref_driver.cpp -- C entry points around the reference's own CPU path.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `ref`) compiles this file
// together with four reference sources, read in place from the reference
// tree (never copied):
//   Managers/CPU/OpenMPManager.cpp, Managers/BaseManager.cpp,
//   compute/CPU/profiles.cpp, Readers/Exomol/exomol_functions.cpp
// into oracle/_ref/libgexs_ref.so.  Tests use it to pin the FP64 oracle
// (oracle/gexs_oracle.c) against the reference arithmetic; bench.py may time
// it as the CPU baseline (kind "reference").
//
// The reference's (T,P) CPU route is unreachable through its public API: the
// job thread exits at once because the constructor sets master_thread_done
// before starting it (OpenMPManager.cpp:14,17,209), and the Doppler worker
// never merges its private grid (OpenMPManager.cpp:280-323).  So this driver
// reaches the worker functions ComputeVoigtTP / ComputeDoppler directly, which
// requires opening the class's private section below.  Per-thread work is the
// reference's own contiguous line partition; merges run in thread order.

#include <algorithm>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <mutex>
#include <queue>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#define private public
#define protected public
#include "OpenMPManager.h"
#include "exomol_functions.h"
#include "exomol_objects.h"
#include "profiles.h"
#undef private
#undef protected

namespace {

OpenMPManager* make_manager(ProfileType prof, int nthreads, int npoints, long nlines) {
    // InitializeVectors(N, num_cs) sizes its line buffers from the memory
    // budget: budget/2 must cover the grids plus 44 B per line.
    size_t grids = size_t(npoints) * sizeof(double) * size_t(2 + nthreads);
    size_t lines = size_t(nlines + 16) * 64;
    size_t total = 2 * (grids + lines + (1u << 20));
    OpenMPManager* m = new OpenMPManager(prof, nthreads, total);
    return m;
}

void free_manager(OpenMPManager* m, int nthreads) {
    // Cleanup() is empty in the reference; release what InitializeVectors made.
    delete[] m->g_freq;
    for (size_t i = 0; i < m->g_tp_intens.size(); ++i) delete[] m->g_tp_intens[i];
    for (int i = 0; i < nthreads; ++i) delete[] m->t_intens[i];
    delete[] m->g_energies;
    delete[] m->g_nu;
    delete[] m->g_aif;
    delete[] m->g_gns;
    if (m->profile == VOIGT) {
        delete[] m->g_gamma;
        delete[] m->g_n;
    }
    // the manager object itself (with its detached job thread) is leaked on
    // purpose: destroying an unjoined std::thread would terminate the process
}

}  // namespace

extern "C" {

double gxr_humlic(double x, double y) { return humlic(x, y); }

double gxr_partition(const double* E, const int* g, long n, double T) {
    exomol_states st;
    st.N = int(n);
    st.states = new state[n];
    for (long i = 0; i < n; ++i) {
        st.states[i].id = int(i);
        st.states[i].energy = E[i];
        st.states[i].gns = g[i];
        st.states[i].J = 0;
    }
    double q = compute_partition(&st, T);
    delete[] st.states;
    return q;
}

// One (T,P) cross-section by OpenMPManager::ComputeVoigtTP.  out receives the
// full grid (npoints); with parallel != 0 the workers run as threads (the
// reference's own concurrency), else one after another in thread-id order.
void gxr_voigt_tp(int nthreads, int parallel, const double* freq, int npoints, double dnu,
                  double T, double P, double Q, double mass, double t_ref, double p_ref,
                  long n, const double* E, const double* nu, const double* A, const int* g,
                  const double* gamma, const double* nexp, int start_idx, int nwin,
                  double* out) {
    OpenMPManager* m = make_manager(VOIGT, nthreads, npoints, n);
    m->InitializeVectors(npoints, 1);
    std::vector<double> zeros(size_t(npoints), 0.0);
    std::memset(m->g_tp_intens[0], 0, sizeof(double) * size_t(npoints));  // unzeroed in the reference
    m->TransferFreq(const_cast<double*>(freq), zeros.data(), npoints);
    m->InitializeConstants(100.0, T, Q, dnu, mass, P, t_ref, p_ref);
    m->TransferVectors(size_t(n), const_cast<double*>(E), const_cast<double*>(nu),
                       const_cast<double*>(A), const_cast<int*>(g), const_cast<double*>(gamma),
                       const_cast<double*>(nexp));
    if (parallel && nthreads > 1) {
        std::vector<std::thread> ws;
        for (int i = 0; i < nthreads; ++i)
            ws.emplace_back(&OpenMPManager::ComputeVoigtTP, m, i, nwin, int(n), start_idx, 0, T, P, Q);
        for (auto& w : ws) w.join();
    } else {
        for (int i = 0; i < nthreads; ++i) m->ComputeVoigtTP(i, nwin, int(n), start_idx, 0, T, P, Q);
    }
    std::memcpy(out, m->g_tp_intens[0], sizeof(double) * size_t(npoints));
    free_manager(m, nthreads);
}

// OpenMPManager::ComputeDoppler (cutoff DOPPLER_CUTOFF = 100, constants from
// InitializeConstants).  Returns the sum of the workers' private grids, i.e.
// sum_lines S*(erf(xp)-erf(xm)) without the 0.5/dnu factor.
void gxr_doppler(int nthreads, const double* freq, int npoints, double dnu, double T, double Q,
                 double mass, long n, const double* E, const double* nu, const double* A,
                 const int* g, int start_idx, int nwin, double* out) {
    OpenMPManager* m = make_manager(DOPPLER, nthreads, npoints, n);
    m->InitializeVectors(npoints, 1);
    std::vector<double> zeros(size_t(npoints), 0.0);
    m->TransferFreq(const_cast<double*>(freq), zeros.data(), npoints);
    m->InitializeConstants(100.0, T, Q, dnu, mass, 1.0, 296.0, 1.0);
    m->TransferVectors(size_t(n), const_cast<double*>(E), const_cast<double*>(nu),
                       const_cast<double*>(A), const_cast<int*>(g), nullptr, nullptr);
    for (int i = 0; i < nthreads; ++i) m->ComputeDoppler(i, nwin, int(n), start_idx);
    for (int j = 0; j < npoints; ++j) {
        double s = 0.0;
        for (int i = 0; i < nthreads; ++i) s += m->t_intens[i][j];
        out[j] = s;
    }
    free_manager(m, nthreads);
}

}  // extern "C"
