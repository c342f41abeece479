// readers.cpp -- ExoMol, HITRAN and in-memory line sources (include/gexs/StateReader.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>

#include "gexs/StateReader.h"

namespace gexs {

namespace {
// h, c, k of Readers/Exomol/exomol_functions.cpp:55-58 (equal to common/defines.h)
constexpr double kPlanck = 6.6260693e-27;
constexpr double kVellgt = 2.99792458e10;
constexpr double kBoltz = 1.380658e-16;
constexpr double kPi = 3.14159265358979323846;

FILE* open_or_throw(const std::string& f) {
    FILE* fp = std::fopen(f.c_str(), "r");
    if (!fp) throw std::runtime_error("cannot open " + f);
    return fp;
}
}  // namespace

double PartitionSum(const std::vector<double>& E, const std::vector<int>& g, double T) {
    const double beta = kPlanck * kVellgt / (kBoltz * T);
    double q = 0.0;
    for (size_t i = 0; i < E.size(); ++i) q += g[i] * std::exp(-beta * E[i]);
    return q;
}

// ------------------------------------------------------------------ ExoMol
ExomolStateReader::ExomolStateReader(const std::string& states_file,
                                     const std::string& broadener_file, double default_gamma,
                                     double default_n)
    : def_gamma_(default_gamma), def_n_(default_n) {
    FILE* fp = open_or_throw(states_file);
    char buf[4096];
    while (std::fgets(buf, sizeof buf, fp)) {
        char* p = buf;
        char* q = nullptr;
        std::strtol(p, &q, 10);  // id (states are indexed by file position)
        if (q == p) continue;    // blank line
        p = q;
        double e = std::strtod(p, &q);
        if (q == p) throw std::runtime_error("unreadable energy in " + states_file + ": " + buf);
        p = q;
        long g = std::strtol(p, &q, 10);
        if (q == p) throw std::runtime_error("unreadable degeneracy in " + states_file + ": " + buf);
        p = q;
        long j = std::strtol(p, &q, 10);
        E_.push_back(e);
        g_.push_back((int)g);
        J_.push_back((int)j);
    }
    std::fclose(fp);
    if (!broadener_file.empty()) {
        FILE* fb = std::fopen(broadener_file.c_str(), "r");
        if (fb) {
            while (std::fgets(buf, sizeof buf, fb)) {
                std::istringstream ss(buf);
                std::string code;
                double gam, n;
                int ji, jf;
                if (!(ss >> code >> gam >> n >> ji)) continue;
                if (code == "a0") {
                    a0_[ji] = {gam, n};
                } else if (code == "a1" && (ss >> jf)) {
                    a1_[{ji, jf}] = {gam, n};
                }
            }
            std::fclose(fb);
        }
    }
}

bool ExomolStateReader::OpenFile(const std::string& filename) {
    CloseFile();
    fp_ = open_or_throw(filename);
    return true;
}

bool ExomolStateReader::CloseFile() {
    if (fp_) std::fclose((FILE*)fp_);
    fp_ = nullptr;
    return true;
}

bool ExomolStateReader::ReadNextState(double& nu, double& gns, double& e_i, double& aif,
                                      double& gam, double& n) {
    if (!fp_) return false;
    char buf[512];
    for (;;) {
        if (!std::fgets(buf, sizeof buf, (FILE*)fp_)) return false;
        char* p = buf;
        char* q = nullptr;
        long idf = std::strtol(p, &q, 10);
        if (q == p) continue;
        p = q;
        long idi = std::strtol(p, &q, 10);
        p = q;
        double a = std::strtod(p, &q);
        if (q == p) throw std::runtime_error(std::string("unreadable transition record: ") + buf);
        if (idf < 1 || idi < 1 || idf > (long)E_.size() || idi > (long)E_.size())
            throw std::runtime_error("transition refers to a state outside the .states file");
        const size_t f = (size_t)(idf - 1), i = (size_t)(idi - 1);
        nu = E_[f] - E_[i];
        e_i = E_[i];
        gns = g_[f];
        aif = a;
        gam = def_gamma_;
        n = def_n_;
        auto it1 = a1_.find({J_[i], J_[f]});
        if (it1 != a1_.end()) {
            gam = it1->second.first;
            n = it1->second.second;
        } else {
            auto it0 = a0_.find(J_[i]);
            if (it0 != a0_.end()) {
                gam = it0->second.first;
                n = it0->second.second;
            }
        }
        return true;
    }
}

double ExomolStateReader::ComputePartition(double temperature) {
    return PartitionSum(E_, g_, temperature);
}

// ------------------------------------------------------------------ HITRAN
namespace {
const int kWidths[] = {2, 1, 12, 10, 10, 5, 5, 10, 4, 8, 15, 15, 15, 15, 6, 12, 1, 7, 7};

std::string field(const char* rec, size_t len, int k) {
    size_t start = 0;
    for (int i = 0; i < k; ++i) start += (size_t)kWidths[i];
    if (start >= len) return std::string();
    size_t w = std::min((size_t)kWidths[k], len - start);
    return std::string(rec + start, w);
}
double fnum(const char* rec, size_t len, int k) { return std::atof(field(rec, len, k).c_str()); }
}  // namespace

HitranStateReader::HitranStateReader(double mixture, double q_ref_296,
                                     std::vector<std::pair<double, double>> q_table)
    : mix_(mixture), q296_(q_ref_296), qtab_(std::move(q_table)) {
    std::sort(qtab_.begin(), qtab_.end());
}

bool HitranStateReader::OpenFile(const std::string& filename) {
    CloseFile();
    fp_ = open_or_throw(filename);
    char buf[1024];
    if (std::fgets(buf, sizeof buf, (FILE*)fp_)) {
        size_t len = std::strlen(buf);
        mol_ = std::atoi(field(buf, len, 0).c_str());
        iso_ = std::atoi(field(buf, len, 1).c_str());
    }
    std::rewind((FILE*)fp_);
    return true;
}

bool HitranStateReader::CloseFile() {
    if (fp_) std::fclose((FILE*)fp_);
    fp_ = nullptr;
    return true;
}

bool HitranStateReader::ReadNextState(double& nu, double& gns, double& e_i, double& aif,
                                      double& gam, double& n) {
    if (!fp_) return false;
    char buf[1024];
    for (;;) {
        if (!std::fgets(buf, sizeof buf, (FILE*)fp_)) return false;
        size_t len = std::strlen(buf);
        while (len && (buf[len - 1] == '\n' || buf[len - 1] == '\r')) buf[--len] = 0;
        if (len < 100) continue;  // not a 160-character record
        nu = fnum(buf, len, 2);
        const double S = fnum(buf, len, 3);
        aif = fnum(buf, len, 4);
        const double g_air = fnum(buf, len, 5);
        const double g_self = fnum(buf, len, 6);
        e_i = fnum(buf, len, 7);
        n = fnum(buf, len, 8);
        gns = fnum(buf, len, 17);
        gam = g_air * mix_ + g_self * (1.0 - mix_);
        if (aif == 0.0) {
            // A from the 296 K intensity (HITRANStateReader.cpp:222-247)
            const double beta = kPlanck * kVellgt / (kBoltz * 296.0);
            aif = S * 8.0 * kPi * kVellgt * q296_ * nu * nu /
                  (std::exp(-beta * e_i) * (1.0 - std::exp(-beta * nu)));
            gns = 1.0;
        }
        return true;
    }
}

double HitranStateReader::ComputePartition(double T) {
    if (qtab_.empty()) return q296_;
    if (T <= qtab_.front().first) return qtab_.front().second;
    if (T >= qtab_.back().first) return qtab_.back().second;
    auto hi = std::lower_bound(qtab_.begin(), qtab_.end(), std::make_pair(T, -1e300));
    auto lo = hi - 1;
    const double w = (T - lo->first) / (hi->first - lo->first);
    return lo->second + w * (hi->second - lo->second);
}

// ------------------------------------------------------------------ arrays
ArrayStateReader::ArrayStateReader(std::vector<double> nu, std::vector<double> E,
                                   std::vector<double> A, std::vector<int> g,
                                   std::vector<double> gamma, std::vector<double> nexp)
    : nu_(std::move(nu)), E_(std::move(E)), A_(std::move(A)), gamma_(std::move(gamma)),
      nexp_(std::move(nexp)), g_(std::move(g)) {}

bool ArrayStateReader::ReadNextState(double& nu, double& gns, double& e_i, double& aif,
                                     double& gam, double& n) {
    if (pos_ >= nu_.size()) return false;
    nu = nu_[pos_];
    e_i = E_[pos_];
    aif = A_[pos_];
    gns = g_[pos_];
    gam = pos_ < gamma_.size() ? gamma_[pos_] : 0.0;
    n = pos_ < nexp_.size() ? nexp_[pos_] : 0.0;
    ++pos_;
    return true;
}

double ArrayStateReader::ComputePartition(double T) {
    if (!sE_.empty()) return PartitionSum(sE_, sg_, T);
    return q_;
}

}  // namespace gexs
