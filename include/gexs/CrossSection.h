// CrossSection.h -- the host driver of the line-by-line path.
//
// The analogue of the reference's Profiles/BaseProfile (the upper plugin API)
// plus the keyword input of common/Input.cpp:98-240.  Its per-step semantics:
//   grid        f[i] = start + i*dfreq, dfreq = (end - start)/N   (BaseProfile.cpp:10,19)
//   Q(T)        reader->ComputePartition(T) per grid (T,P)         (:91-98)
//   selection   keep start-HW <= nu <= end+HW, nu != 0 -- a filter on both
//               sides; the reference breaks at nu > end+HW, which drops
//               lines of unsorted files                             (:152-157)
//   chunks      lines are sent in chunks of at most the manager's capacity;
//               chunk window ib = max(round((min_nu-HW-start)/dfreq), 0),
//               ie = min(round((max_nu+HW-start)/dfreq), N)          (:106-108)
//   (T,P) loop  one ExecuteCrossSection per (T,P) and chunk          (:111-112)
//   check       sum sigma*dfreq against sum of line intensities      (:250-267, 283-291)
//   output      "[i] nu sigma" per point                            (:292-295)
// The reference also flushes a chunk whenever its window exceeds MAX-POINTS,
// which with the default flushes after every line (BaseProfile.cpp:133-137,
// 174-181); here chunks are bounded by capacity only.
#pragma once
#include <cstdint>
#include <cstdio>
#include <istream>
#include <memory>
#include <string>
#include <vector>

#include "BaseManager.h"
#include "StateReader.h"

namespace gexs {

enum class FileType { HITRAN, EXOMOL };

struct TemperaturePressure {  // common/Input.h:24-29
    double temperature = 296.0;
    double pressure = 1.0;
    double partition = 0.0;
};

// Keyword input (common/Input.cpp:98-240) plus explicit conventions.
struct Input {
    FileType file_type = FileType::EXOMOL;
    ProfileType profile = VOIGT;
    double nu_start = 0.0, nu_end = 100.0;
    int npoints = 1000;
    std::vector<TemperaturePressure> tp;  // TEMPERATURE-PRESSURE-GRID ... END
    double temperature = 296.0, pressure = 1.0;
    double mean_mass = 10.0;
    double half_width = 0.1;  // HALF-WIDTH: the Gaussian HWHM
    double gamma = 0.05, gamma_n = 0.4;
    double mixture = 1.0;
    double partition = -1.0;
    std::string state_file, partition_file;
    std::vector<std::string> trans_files, broadener_files;
    int64_t memory = 1000000000LL;  // MEMORY (GB in the file)
    int num_threads = 1;
    int max_points = 500;
    // conventions the reference leaves implicit (new keywords)
    double cutoff = -1.0;  // CUTOFF; < 0: profile default (Voigt min(25P,100), Doppler 100, Gaussian 10 HW)
    double t_ref = 296.0;  // T-REF (the reference uses 256 K for ExoMol, BaseProfile.cpp:53)
    double p_scale = -1.0; // PRESSURE-SCALE; < 0: HITRAN 1, ExoMol 1/0.986923267 (BaseProfile.cpp:50,54)
    int width_conv = 0;    // WIDTH-CONVENTION ref_cpu | textbook
    int pressure_mode = 1; // PRESSURE-MODE ref_cpu | scaled
    int precision = 0;     // PRECISION fp32c | fp64
    int64_t chunk_lines = 0;  // CHUNK-LINES: cap on lines per chunk (0: manager capacity)
};

Input ParseInput(std::istream& in);

// the reader the input names (ExoMol or HITRAN), as the driver builds it
std::unique_ptr<StateReader> MakeReader(const Input& in);

class CrossSection {
   public:
    // reader / manager may be supplied (not owned); otherwise they are made
    // from the input (ExoMol or HITRAN reader, HipManager on device 0)
    explicit CrossSection(const Input& in, StateReader* reader = nullptr,
                          BaseManager* manager = nullptr);
    ~CrossSection();

    void Initialize();           // BaseProfile::Initialize
    void ExecuteCrossSection();  // BaseProfile::ExecuteCrossSection
    void OutputProfile(FILE* out);  // BaseProfile::OutputProfile (nullptr: results only)

    int Npoints() const { return npoints_; }
    double Dfreq() const { return dfreq_; }
    const std::vector<double>& Freq() const { return freq_; }
    const std::vector<TemperaturePressure>& TP() const { return tp_; }
    const std::vector<double>& Result(int i) const { return results_.at((size_t)i); }
    double TotalIntensity() const { return total_intens_; }
    // sum(sigma) * dfreq of the first (T,P), set by OutputProfile
    double IntegratedIntensity() const { return integrated_; }
    int64_t LinesSelected() const { return selected_; }
    int64_t LinesRead() const { return read_; }
    int Chunks() const { return chunks_; }
    double HalfWidth() const { return hw_; }

   private:
    void PerformCrossSection();
    void ComputeTotalIntensity();
    Input in_;
    StateReader* reader_ = nullptr;
    BaseManager* manager_ = nullptr;
    std::unique_ptr<StateReader> own_reader_;
    std::unique_ptr<BaseManager> own_manager_;
    std::vector<TemperaturePressure> tp_;
    std::vector<double> freq_;
    std::vector<std::vector<double>> results_;
    std::vector<double> h_E_, h_nu_, h_A_, h_gamma_, h_n_;
    std::vector<int> h_g_;
    int npoints_ = 0;
    double dfreq_ = 0.0, start_ = 0.0, end_ = 0.0, hw_ = 100.0;
    double min_nu_ = 0.0, max_nu_ = 0.0, total_intens_ = 0.0, integrated_ = 0.0;
    int64_t capacity_ = 0, ntrans_ = 0, selected_ = 0, read_ = 0;
    int chunks_ = 0;
    bool pressure_ = true;
};

}  // namespace gexs
