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

// ------------------------------------------------------------ line sources
// bzlib 1.0 decompression API (the system libbz2.so.1 is linked; the image
// has no bzlib.h, so the public declarations are restated here)
extern "C" {
struct gx_bz_stream {
    char* next_in;
    unsigned int avail_in;
    unsigned int total_in_lo32, total_in_hi32;
    char* next_out;
    unsigned int avail_out;
    unsigned int total_out_lo32, total_out_hi32;
    void* state;
    void* (*bzalloc)(void*, int, int);
    void (*bzfree)(void*, void*);
    void* opaque;
};
int BZ2_bzDecompressInit(gx_bz_stream* strm, int verbosity, int small);
int BZ2_bzDecompress(gx_bz_stream* strm);
int BZ2_bzDecompressEnd(gx_bz_stream* strm);
}
namespace {
constexpr int kBzOk = 0, kBzStreamEnd = 4;
}

class LineSource {
   public:
    virtual ~LineSource() {}
    // the next line without its newline; false at the end of the file
    virtual bool next(std::string& line) = 0;
};

namespace {
class TextLines : public LineSource {
   public:
    explicit TextLines(const std::string& f) : fp_(open_or_throw(f)) {}
    ~TextLines() override { std::fclose(fp_); }
    bool next(std::string& line) override {
        line.clear();
        char buf[512];
        while (std::fgets(buf, sizeof buf, fp_)) {
            const size_t n = std::strlen(buf);
            if (n && buf[n - 1] == '\n') {
                line.append(buf, n - 1);
                return true;
            }
            line.append(buf, n);
        }
        return !line.empty();
    }

   private:
    FILE* fp_;
};

// bzip2 file -> text lines.  A file may hold several concatenated bzip2
// streams (parallel compressors write them): at the end of one stream the
// decoder restarts on the remaining input.
class Bz2Lines : public LineSource {
   public:
    explicit Bz2Lines(const std::string& f) : name_(f), in_(1 << 16), out_(1 << 16) {
        fp_ = std::fopen(f.c_str(), "rb");
        if (!fp_) throw std::runtime_error("cannot open " + f);
        std::memset(&s_, 0, sizeof s_);
    }
    ~Bz2Lines() override {
        if (active_) BZ2_bzDecompressEnd(&s_);
        std::fclose(fp_);
    }
    bool next(std::string& line) override {
        line.clear();
        for (;;) {
            const char* b = out_.data() + opos_;
            const char* nl = (const char*)std::memchr(b, '\n', olen_ - opos_);
            if (nl) {
                line.append(b, (size_t)(nl - b));
                opos_ = (size_t)(nl - out_.data()) + 1;
                return true;
            }
            line.append(b, olen_ - opos_);
            opos_ = olen_;
            if (!refill()) return !line.empty();
        }
    }

   private:
    bool refill() {
        for (;;) {
            if (s_.avail_in == 0 && !eof_) {
                const size_t n = std::fread(in_.data(), 1, in_.size(), fp_);
                if (n == 0) eof_ = true;
                s_.next_in = in_.data();
                s_.avail_in = (unsigned)n;
            }
            if (!active_) {
                if (s_.avail_in == 0 && eof_) return false;
                char* ni = s_.next_in;
                const unsigned ai = s_.avail_in;
                std::memset(&s_, 0, sizeof s_);
                if (BZ2_bzDecompressInit(&s_, 0, 0) != kBzOk)
                    throw std::runtime_error("bzip2 decoder init failed for " + name_);
                s_.next_in = ni;
                s_.avail_in = ai;
                active_ = true;
            }
            s_.next_out = out_.data();
            s_.avail_out = (unsigned)out_.size();
            const int rc = BZ2_bzDecompress(&s_);
            const size_t produced = out_.size() - s_.avail_out;
            if (rc == kBzStreamEnd) {
                char* ni = s_.next_in;
                const unsigned ai = s_.avail_in;
                BZ2_bzDecompressEnd(&s_);
                active_ = false;
                s_.next_in = ni;
                s_.avail_in = ai;
            } else if (rc != kBzOk) {
                throw std::runtime_error("corrupt bzip2 data in " + name_);
            } else if (produced == 0 && s_.avail_in == 0 && eof_) {
                throw std::runtime_error("truncated bzip2 stream in " + name_);
            }
            if (produced > 0) {
                opos_ = 0;
                olen_ = produced;
                return true;
            }
        }
    }

    std::string name_;
    FILE* fp_ = nullptr;
    gx_bz_stream s_;
    bool active_ = false, eof_ = false;
    std::vector<char> in_, out_;
    size_t opos_ = 0, olen_ = 0;
};

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
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

ExomolStateReader::~ExomolStateReader() {}

bool ExomolStateReader::OpenFile(const std::string& filename) {
    CloseFile();
    if (ends_with(filename, ".bz2"))
        src_.reset(new Bz2Lines(filename));
    else
        src_.reset(new TextLines(filename));
    return true;
}

bool ExomolStateReader::CloseFile() {
    src_.reset();
    return true;
}

bool ExomolStateReader::ReadNextState(double& nu, double& gns, double& e_i, double& aif,
                                      double& gam, double& n) {
    if (!src_) return false;
    std::string line;
    for (;;) {
        if (!src_->next(line)) return false;
        const char* buf = line.c_str();
        char* p = (char*)buf;
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
