// cross_section.cpp -- keyword input and host driver (include/gexs/CrossSection.h).
#include "gexs/CrossSection.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "gexs/HipManager.h"

namespace gexs {

namespace {
std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return std::string();
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::vector<std::string> split(const std::string& s) {
    std::istringstream ss(s);
    std::vector<std::string> out;
    std::string w;
    while (ss >> w) out.push_back(w);
    return out;
}

ProfileType to_profile(const std::string& s) {  // common/Input.cpp:85-96
    if (s == "GAUSSIAN") return GAUSSIAN;
    if (s == "DOPPLER") return DOPPLER;
    if (s == "VOIGT") return VOIGT;
    if (s == "VOIGT-QUAD" || s == "VOIGT_QUAD") return VOIGT_QUAD;
    throw std::runtime_error("PROFILE " + s + " is not supported");
}

// CM_COEF and the Boltzmann factor of BaseProfile.cpp:255-260
constexpr double kCmCoef = 1.327209365e-12;
constexpr double kPlanck = 6.6260693e-27;
constexpr double kVellgt = 2.99792458e+10;
constexpr double kBoltz = 1.380658e-16;
}  // namespace

Input ParseInput(std::istream& is) {
    Input in;
    std::string line;
    if (!std::getline(is, line)) throw std::runtime_error("empty input");
    const std::string kind = trim(line);
    if (kind == "HITRAN")
        in.file_type = FileType::HITRAN;
    else if (kind == "EXOMOL")
        in.file_type = FileType::EXOMOL;
    else
        throw std::runtime_error("File type: " + kind + " not supported");
    while (std::getline(is, line)) {
        auto w = split(trim(line));
        if (w.empty()) continue;
        const std::string& k = w[0];
        auto num = [&](size_t i) {
            if (w.size() <= i) throw std::runtime_error("keyword " + k + " needs a value");
            return std::atof(w[i].c_str());
        };
        auto block = [&](std::vector<std::string>& dst) {
            while (std::getline(is, line)) {
                std::string t = trim(line);
                if (t == "END") break;
                if (!t.empty()) dst.push_back(t);
            }
        };
        if (k == "PRESSURE") in.pressure = num(1);
        else if (k == "TEMPERATURE") in.temperature = num(1);
        else if (k == "TEMPERATURE-PRESSURE-GRID" || k == "TEMPERATURE-HALFWIDTH-GRID" ||
                 k == "TEMPERATURES") {
            std::vector<std::string> rows;
            block(rows);
            for (auto& r : rows) {
                auto c = split(r);
                TemperaturePressure tp;
                tp.temperature = std::atof(c[0].c_str());
                tp.pressure = c.size() > 1 ? std::atof(c[1].c_str()) : 0.1;  // Input.cpp:152
                in.tp.push_back(tp);
            }
        } else if (k == "MEAN-MASS") in.mean_mass = num(1);
        else if (k == "HITRAN-MIXTURE-AIR") in.mixture = num(1);
        else if (k == "PARTITION") in.partition = num(1);
        else if (k == "PARTITION-FILE") in.partition_file = w.size() > 1 ? w[1] : "";
        else if (k == "HALF-WIDTH") in.half_width = num(1);
        else if (k == "GAMMA") in.gamma = num(1);
        else if (k == "GAMMA-N") in.gamma_n = num(1);
        else if (k == "MAX-POINTS") in.max_points = (int)num(1);
        else if (k == "STATE-FILE") in.state_file = w.size() > 1 ? w[1] : "";
        else if (k == "TRANS-FILES") block(in.trans_files);
        else if (k == "BROADENERS") block(in.broadener_files);
        else if (k == "PROFILE") in.profile = to_profile(w.size() > 1 ? w[1] : "");
        else if (k == "N-POINTS") in.npoints = (int)num(1);
        else if (k == "NU-RANGE") {
            in.nu_start = num(1);
            in.nu_end = num(2);
        } else if (k == "MEMORY") in.memory = (int64_t)(num(1) * 1e9);
        else if (k == "NUM-THREADS") in.num_threads = (int)num(1);
        else if (k == "CUTOFF") in.cutoff = num(1);
        else if (k == "T-REF") in.t_ref = num(1);
        else if (k == "PRESSURE-SCALE") in.p_scale = num(1);
        else if (k == "WIDTH-CONVENTION") in.width_conv = (w.size() > 1 && w[1] == "textbook") ? 1 : 0;
        else if (k == "PRESSURE-MODE") in.pressure_mode = (w.size() > 1 && w[1] == "ref_cpu") ? 0 : 1;
        else if (k == "PRECISION") in.precision = (w.size() > 1 && w[1] == "fp64") ? 1 : 0;
        else if (k == "CHUNK-LINES") in.chunk_lines = (int64_t)num(1);
        else throw std::runtime_error("unknown keyword " + k);
    }
    if (in.tp.empty()) {
        TemperaturePressure tp;
        tp.temperature = in.temperature;
        tp.pressure = in.pressure;
        in.tp.push_back(tp);
    }
    if (in.p_scale < 0) in.p_scale = in.file_type == FileType::HITRAN ? 1.0 : 1.0 / 0.986923267;
    return in;
}

CrossSection::CrossSection(const Input& in, StateReader* reader, BaseManager* manager)
    : in_(in), reader_(reader), manager_(manager) {
    start_ = in_.nu_start;
    end_ = in_.nu_end;
    npoints_ = in_.npoints;
    if (npoints_ <= 0 || !(end_ > start_)) throw std::runtime_error("bad grid");
    dfreq_ = (end_ - start_) / double(npoints_);
    freq_.resize((size_t)npoints_);
    for (int i = 0; i < npoints_; ++i) freq_[(size_t)i] = start_ + double(i) * dfreq_;
    tp_ = in_.tp;
    // selection half width m_half_width of the profile plugins
    // (VoigtProfile.h:15, DopplerProfile.h:11, GaussianProfile.h:12)
    hw_ = in_.profile == GAUSSIAN ? 10.0 : 100.0;
    pressure_ = in_.profile == VOIGT || in_.profile == VOIGT_QUAD;
}

CrossSection::~CrossSection() {}

std::unique_ptr<StateReader> MakeReader(const Input& in) {
    if (in.file_type == FileType::EXOMOL)
        return std::unique_ptr<StateReader>(new ExomolStateReader(
            in.state_file, in.broadener_files.empty() ? "" : in.broadener_files[0], in.gamma,
            in.gamma_n));
    std::vector<std::pair<double, double>> tab;
    if (!in.partition_file.empty()) {
        std::ifstream f(in.partition_file);
        double t, q;
        while (f >> t >> q) tab.push_back({t, q});
    }
    return std::unique_ptr<StateReader>(new HitranStateReader(in.mixture, in.partition, tab));
}

void CrossSection::Initialize() {
    if (!reader_) {
        own_reader_ = MakeReader(in_);
        reader_ = own_reader_.get();
    }
    for (auto& tp : tp_) tp.partition = reader_->ComputePartition(tp.temperature);  // :91-98
    if (!manager_) {
        HipManager* hm = new HipManager(in_.profile);
        Conventions c;
        double cut = in_.cutoff;
        if (cut < 0) cut = in_.profile == DOPPLER ? 100.0 : in_.profile == GAUSSIAN ? 10.0 * in_.half_width : 0.0;
        c.cutoff = cut;
        c.width_conv = in_.width_conv;
        c.pressure_mode = in_.pressure_mode;
        c.precision = in_.precision;
        c.gauss_hwhm = in_.half_width;
        c.mem_budget = 0;
        hm->SetConventions(c);
        own_manager_.reset(hm);
        manager_ = hm;
    }
    manager_->InitializeVectors(npoints_, (int)tp_.size());
    std::vector<double> zeros((size_t)npoints_, 0.0);
    manager_->TransferFreq(freq_.data(), zeros.data(), npoints_);
    capacity_ = (int64_t)manager_->GetNtrans();
    if (in_.chunk_lines > 0) capacity_ = std::min<int64_t>(capacity_, in_.chunk_lines);
    if (capacity_ <= 0) throw std::runtime_error("manager has no line capacity");
    manager_->InitializeConstants(in_.half_width, tp_[0].temperature, tp_[0].partition, dfreq_,
                                  in_.mean_mass, tp_[0].pressure, in_.t_ref, in_.p_scale);
    const size_t cap = (size_t)std::min<int64_t>(capacity_, 1 << 24);
    h_E_.reserve(cap);
    h_nu_.reserve(cap);
    h_A_.reserve(cap);
    h_g_.reserve(cap);
    if (pressure_) {
        h_gamma_.reserve(cap);
        h_n_.reserve(cap);
    }
}

void CrossSection::PerformCrossSection() {
    ntrans_ = (int64_t)h_nu_.size();
    if (ntrans_ == 0) return;
    manager_->TransferVectors((size_t)ntrans_, h_E_.data(), h_nu_.data(), h_A_.data(), h_g_.data(),
                              pressure_ ? h_gamma_.data() : nullptr, pressure_ ? h_n_.data() : nullptr);
    const int ib = (int)std::max(std::round((min_nu_ - hw_ - start_) / dfreq_), 0.0);
    const int ie = (int)std::min(std::round((max_nu_ + hw_ - start_) / dfreq_), double(npoints_));
    const int points = ie - ib;
    if (points > 0)
        for (const auto& tp : tp_)
            manager_->ExecuteCrossSection(points, (int)ntrans_, ib, tp.temperature, tp.pressure,
                                          tp.partition);
    ComputeTotalIntensity();
    ++chunks_;
}

void CrossSection::ComputeTotalIntensity() {
    // first (T,P) only, CM_COEF without 1/sqrt(pi) (BaseProfile.cpp:250-267)
    const double T = tp_[0].temperature, Q = tp_[0].partition;
    const double beta = kPlanck * kVellgt / (kBoltz * T);
    double s = 0.0;
    for (int64_t i = 0; i < ntrans_; ++i) {
        const double v = h_nu_[(size_t)i];
        s += kCmCoef * h_A_[(size_t)i] * h_g_[(size_t)i] * std::exp(-beta * h_E_[(size_t)i]) *
             (1.0 - std::exp(-beta * v)) / (v * v * Q);
    }
    total_intens_ += s;
}

void CrossSection::ExecuteCrossSection() {
    double nu, gns, e_i, aif, gam, n;
    auto reset = [&]() {
        h_E_.clear();
        h_nu_.clear();
        h_A_.clear();
        h_g_.clear();
        h_gamma_.clear();
        h_n_.clear();
        min_nu_ = 99999.0;  // BaseProfile.cpp:126-127
        max_nu_ = 0.0;
    };
    reset();
    std::vector<std::string> files = in_.trans_files;
    if (files.empty()) files.push_back(std::string());  // readers without files
    for (const auto& f : files) {
        reader_->OpenFile(f);
        while (reader_->ReadNextState(nu, gns, e_i, aif, gam, n)) {
            ++read_;
            if (nu < start_ - hw_) continue;
            if (nu > end_ + hw_) continue;
            if (nu == 0.0) continue;
            ++selected_;
            h_E_.push_back(e_i);
            h_nu_.push_back(nu);
            h_g_.push_back((int)gns);
            h_A_.push_back(aif);
            if (pressure_) {
                h_gamma_.push_back(gam);
                h_n_.push_back(n);
            }
            min_nu_ = std::min(min_nu_, nu);
            max_nu_ = std::max(max_nu_, nu);
            if ((int64_t)h_nu_.size() >= capacity_) {
                PerformCrossSection();
                reset();
            }
        }
        reader_->CloseFile();
    }
    if (!h_nu_.empty()) PerformCrossSection();
}

void CrossSection::OutputProfile(FILE* out) {
    results_.assign(tp_.size(), std::vector<double>((size_t)npoints_, 0.0));
    for (size_t i = 0; i < tp_.size(); ++i) {
        std::vector<double>& intens = results_[i];
        manager_->TransferResults(freq_.data(), intens.data(), npoints_, (int)i);
        if (i == 0) {
            double cross = 0.0;
            for (double v : intens) cross += v;
            cross *= dfreq_;
            integrated_ = cross;
        }
        if (!out) continue;
        if (i == 0) {
            const double cross = integrated_;
            std::fprintf(out,
                         "Our summed intensity = %16.6E compared to the integrated intensity = "
                         "%16.6E with a difference of %12.6f %%\n",
                         total_intens_, cross, (total_intens_ - cross) * 100.0 / total_intens_);
        }
        std::fprintf(out, "-------T=%12.6f K--------P=%12.6f atm---------\n\n", tp_[i].temperature,
                     tp_[i].pressure);
        for (int j = 0; j < npoints_; ++j)
            std::fprintf(out, "[%i] %12.6f %13.8E\n", (int)i, freq_[(size_t)j], intens[(size_t)j]);
        std::fprintf(out, "\n");
    }
}

}  // namespace gexs
