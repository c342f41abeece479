// StateReader.h -- streaming line source of the cross-section driver.
//
// Same interface as the reference's Readers/StateReader.h:6-19: one
// ReadNextState call yields one line (nu, g', E'', A, gamma_ref, n) and
// ComputePartition(T) gives Q(T).  Implementations here:
//   ExomolStateReader  .states table + .trans (text, or bzip2-compressed when
//                      the name ends in .bz2) + .broad (Readers/Exomol/*)
//   HitranStateReader  160-character .par records (Readers/HITRAN/*)
//   ArrayStateReader   in-memory arrays (tests, embedding)
#pragma once
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace gexs {

class LineSource;  // text lines of a plain or bzip2-compressed file (readers.cpp)

class StateReader {
   public:
    virtual ~StateReader() {}
    virtual bool OpenFile(const std::string& filename) = 0;
    virtual bool CloseFile() = 0;
    virtual bool ReadNextState(double& nu, double& gns, double& e_i, double& aif, double& gam,
                               double& n) = 0;
    virtual bool ReadNextState(double& nu, double& gns, double& e_i, double& aif) {
        double g = 0.0, n = 0.0;
        return ReadNextState(nu, gns, e_i, aif, g, n);
    }
    virtual double ComputePartition(double temperature) = 0;
};

// Q(T) = sum g exp(-c2 E / T) in state order (Readers/Exomol/exomol_functions.cpp:50-73)
double PartitionSum(const std::vector<double>& E, const std::vector<int>& g, double T);

// ExoMol: states "id E g J ..." (1-based ids in file order,
// exomol_functions.cpp:8-48); transitions "id_upper id_lower A"
// (ExomolStateReader.cpp:147); line = (E_u - E_l, g_u, E_l, A)
// (ExomolStateReader.cpp:158-162).  Broadeners: "a0 gamma n J''" keyed by the
// lower J and "a1 gamma n J'' J'" keyed by both J; a1 is looked up first,
// then a0, else the defaults.  (The reference never stores a1 entries and
// leaves gamma/n unset without a broadener file, ExomolStateReader.cpp:21,64,163.)
class ExomolStateReader : public StateReader {
   public:
    ExomolStateReader(const std::string& states_file, const std::string& broadener_file,
                      double default_gamma, double default_n);
    ~ExomolStateReader() override;
    // a name ending in ".bz2" is read through bzip2 (concatenated streams
    // included); the reference reads such files as 37-byte "id_f id_i A"
    // records (Readers/Exomol/bz2_compression/read_compress_trans.c:48-90),
    // which are text lines of 36 characters plus the newline
    bool OpenFile(const std::string& filename) override;
    bool CloseFile() override;
    bool ReadNextState(double& nu, double& gns, double& e_i, double& aif, double& gam,
                       double& n) override;
    double ComputePartition(double temperature) override;
    const std::vector<double>& energies() const { return E_; }
    const std::vector<int>& degeneracies() const { return g_; }
    const std::vector<int>& J() const { return J_; }

   private:
    std::vector<double> E_;
    std::vector<int> g_, J_;
    std::map<int, std::pair<double, double>> a0_;
    std::map<std::pair<int, int>, std::pair<double, double>> a1_;
    double def_gamma_, def_n_;
    std::unique_ptr<LineSource> src_;
};

// HITRAN 160-character records; field widths {2,1,12,10,10,5,5,10,4,8,15,15,
// 15,15,6,12,1,7,7} (HITRANStateReader.cpp:3-18): nu = field 2, S = 3, A = 4,
// gamma_air = 5, gamma_self = 6, E'' = 7, n_air = 8, g' = 17.
// gamma_ref = gamma_air*mix + gamma_self*(1-mix); pressure is applied once by
// the manager (the reference also multiplies here, HITRANStateReader.cpp:218).
// A == 0 records take A from S at 296 K with g' = 1 (HITRANStateReader.cpp:222-247).
// Q(T): the TIPS-2011 tables are not available, so Q comes from a (T, Q)
// table with linear interpolation, or a constant.
class HitranStateReader : public StateReader {
   public:
    HitranStateReader(double mixture, double q_ref_296, std::vector<std::pair<double, double>> q_table);
    bool OpenFile(const std::string& filename) override;
    bool CloseFile() override;
    bool ReadNextState(double& nu, double& gns, double& e_i, double& aif, double& gam,
                       double& n) override;
    double ComputePartition(double temperature) override;
    int mol_id() const { return mol_; }
    int iso() const { return iso_; }

   private:
    double mix_, q296_;
    std::vector<std::pair<double, double>> qtab_;
    void* fp_ = nullptr;
    int mol_ = 0, iso_ = 0;
};

// In-memory lines (one "file"); Q from a state table if given, else constant.
class ArrayStateReader : public StateReader {
   public:
    ArrayStateReader(std::vector<double> nu, std::vector<double> E, std::vector<double> A,
                     std::vector<int> g, std::vector<double> gamma, std::vector<double> nexp);
    void SetStates(std::vector<double> E, std::vector<int> g) {
        sE_ = std::move(E);
        sg_ = std::move(g);
    }
    void SetPartition(double q) { q_ = q; }
    bool OpenFile(const std::string&) override {
        pos_ = 0;
        return true;
    }
    bool CloseFile() override { return true; }
    bool ReadNextState(double& nu, double& gns, double& e_i, double& aif, double& gam,
                       double& n) override;
    double ComputePartition(double temperature) override;

   private:
    std::vector<double> nu_, E_, A_, gamma_, nexp_, sE_;
    std::vector<int> g_, sg_;
    size_t pos_ = 0;
    double q_ = 1.0;
};

}  // namespace gexs
