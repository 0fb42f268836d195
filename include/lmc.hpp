// C++ mirror of the reference's predictor and driver interface over the C ABI (lmc_capi.h).
//
// Names, argument meanings and exceptions follow zhucongx/LatticeMonteCarlo so that host code
// written against it compiles with few changes:
//   cfg::Config, cfg::GenerateFCC                  lmc/cfg/include/Config.h:15-115 (with
//                                                  ReadConfig/WriteConfig and the map files)
//   pred::EnergyPredictor                          lmc/pred/include/EnergyPredictor.h:16-25
//   pred::VacancyMigrationPredictorE0/Quartic      VacancyMigrationPredictorE0.h:15-24,
//                                                  VacancyMigrationPredictorQuartic.h:18-27
//   pred::EnergyChangePredictorPairSite            EnergyChangePredictorPairSite.h:16-25
//   mc::KineticMcFirstBatch, mc::CanonicalMcBatch  KineticMcFirstOmp.h:11-25,
//                                                  CanonicalMcSerial.h:11-20 (R replicas each)
//   mc::KineticMcChainBatch                        KineticMcChainOmpi.h (second-order steps)
//   mc::SimulatedAnnealingBatch                    SimulatedAnnealing.h (adaptive schedule)
//   pred::Surrogate                                no counterpart (learned jump evaluation)
//   ansys::ExitTime                                lmc/ansys/include/ExitTime.h:12-63 (barrier
//                                                  lists and exit times)
// Elements are strings ("Al", "X"); species codes are indices into sorted(element_set + "X").
// Errors: LMC_ERR_IO -> std::runtime_error, LMC_ERR_MODEL -> std::out_of_range,
// LMC_ERR_ARG -> std::invalid_argument, LMC_ERR_HIP -> std::runtime_error, each carrying
// lmc_last_error(). Every computation runs on the GPU; there is no CPU path.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "lmc_capi.h"

// Element (lmc/cfg/include/Element.hpp:7-100): the reference's element names, ordered by name
// (std::set<Element> is string-sorted); an unknown string is the vacancy X.
class Element {
 public:
  Element() = default;
  explicit Element(const std::string &element_string)
      : name_(Known(element_string) ? element_string : "X") {}
  [[nodiscard]] std::string GetString() const { return name_; }
  [[nodiscard]] double GetMass() const {
    const std::string b = name_.size() > 1 && name_[0] == 'p' ? name_.substr(1) : name_;
    if (b == "Al") return 26.98;
    if (b == "Mg") return 24.31;
    if (b == "Zn") return 65.38;
    if (b == "Cu") return 63.55;
    if (b == "Sn") return 118.71;
    return 0.0;
  }
  bool operator==(const Element &rhs) const { return name_ == rhs.name_; }
  bool operator!=(const Element &rhs) const { return name_ != rhs.name_; }
  bool operator<(const Element &rhs) const { return name_ < rhs.name_; }

 private:
  static bool Known(const std::string &s) {
    for (const char *k : {"X", "Al", "Mg", "Zn", "Cu", "Sn", "pAl", "pMg", "pZn", "pCu", "pSn"})
      if (s == k) return true;
    return false;
  }
  std::string name_ = "X";
};

namespace lmc {

// Element names of a reference-style element set.
inline std::set<std::string> ElementNames(const std::set<Element> &element_set) {
  std::set<std::string> out;
  for (const auto &e : element_set) out.insert(e.GetString());
  return out;
}

inline void Check(int rc) {
  if (rc == LMC_OK) return;
  const std::string msg = lmc_last_error();
  if (rc == LMC_ERR_ARG) throw std::invalid_argument(msg);
  if (rc == LMC_ERR_MODEL) throw std::out_of_range(msg);
  throw std::runtime_error(msg);
}

// Owning device allocation (lmc_device_alloc / lmc_device_free).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { Check(lmc_device_alloc(bytes, &p_)); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : p_(o.p_) { o.p_ = nullptr; }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    std::swap(p_, o.p_);
    return *this;
  }
  ~DeviceBuffer() { (void)lmc_device_free(p_); }
  template <typename T>
  T *as() const { return static_cast<T *>(p_); }
  void Upload(const void *src, size_t bytes) { Check(lmc_copy_to_device(p_, src, bytes)); }
  void Download(void *dst, size_t bytes) const { Check(lmc_copy_to_host(dst, p_, bytes)); }

 private:
  void *p_ = nullptr;
};

// Model handle shared by the predictor classes.
class ModelHandle {
 public:
  ModelHandle(const std::string &json, const std::set<std::string> &element_set, int kind) {
    std::vector<const char *> els;
    for (const auto &e : element_set)
      if (e != "X") els.push_back(e.c_str());
    lmc_model *m = nullptr;
    Check(lmc_model_load(json.c_str(), els.data(), static_cast<int32_t>(els.size()), kind, &m));
    m_.reset(m);
  }
  lmc_model *get() const { return m_.get(); }
  int Code(const std::string &element) const {
    const int c = lmc_model_species_code(m_.get(), element.c_str());
    return c >= 0 ? c : lmc_model_species_code(m_.get(), "X");  // unknown strings are X
  }

 private:
  struct Free {
    void operator()(lmc_model *m) const { lmc_model_free(m); }
  };
  std::unique_ptr<lmc_model, Free> m_;
};

}  // namespace lmc

namespace cfg {

namespace detail {
struct FreeConfig {
  void operator()(lmc_config *c) const { lmc_config_free(c); }
};
using ConfigHandle = std::unique_ptr<lmc_config, FreeConfig>;
inline std::vector<const char *> Names(const std::vector<std::string> &v) {
  std::vector<const char *> out;
  for (const auto &s : v) out.push_back(s.c_str());
  return out;
}
// Every element name the reference knows (Element.hpp:7), sorted.
inline std::vector<std::string> AllElements() {
  std::vector<std::string> v{"Al", "Cu", "Mg", "Sn", "X", "Zn", "pAl", "pCu", "pMg", "pSn", "pZn"};
  std::sort(v.begin(), v.end());
  return v;
}
}  // namespace detail

// An n x n x n-cell FCC box with one atom per lattice site, lattice ids in the reference's
// x-major order (ReassignLatticeVector, Config.cpp:466-552). It keeps the element of every site
// (species codes of sorted(element_set + "X"), which index the engine's tables), the atom on
// every site and every atom's map_shift. A device copy of the codes is made on first use and
// refreshed after every change.
class Config {
 public:
  // GenerateFCC (Config.cpp:1060-1095): basis 4.046 n, atoms numbered as GenerateFCC creates them.
  Config(int n, std::set<std::string> element_set_without_vacancy, const std::string &element)
      : n_(n) {
    SetSpecies(std::move(element_set_without_vacancy));
    lmc_config *c = nullptr;
    lmc::Check(lmc_config_generate(n, element.c_str(), &c));
    detail::ConfigHandle h(c);
    codes_.assign(static_cast<size_t>(lmc_fcc_num_sites(n)), CodeOf(element));
    LoadAtoms(h.get());
  }
  // Copies get their own device buffer.
  Config(const Config &o)
      : n_(o.n_), basis_(o.basis_), species_(o.species_), codes_(o.codes_), atoms_(o.atoms_),
        shifts_(o.shifts_) {}
  Config &operator=(const Config &o) {
    n_ = o.n_;
    basis_ = o.basis_;
    species_ = o.species_;
    codes_ = o.codes_;
    atoms_ = o.atoms_;
    shifts_ = o.shifts_;
    dirty_ = true;
    dev_.reset();
    return *this;
  }
  Config(Config &&) = default;
  Config &operator=(Config &&) = default;

  // An empty configuration, to be assigned from ReadConfig / ReadMap (api/src/Home.cpp:163-169).
  Config() = default;
  // Config::ReassignLatticeVector (Config.cpp:466-552): the lattice ids of every configuration
  // here are already the x-major ones it assigns (ReadConfig and ReadMap sort on reading), so
  // this keeps the configuration as it is.
  void ReassignLatticeVector() {}
  // The same configuration with its species codes taken from sorted(element_set + "X"): the
  // code order of the models built on that element set. Elements outside the set become X.
  [[nodiscard]] Config WithElementSet(std::set<std::string> element_set) const {
    Config out(*this);
    out.SetSpecies(std::move(element_set));
    for (size_t i = 0; i < codes_.size(); ++i) out.codes_[i] = out.CodeOf(species_.at(codes_[i]));
    out.dirty_ = true;
    return out;
  }

  // Config::ReadConfig(filename) + ReassignLatticeVector (Config.cpp:554-661; gzip by the .gz
  // extension). The species list is sorted(element_set + "X"); an empty element_set takes the
  // elements found in the file.
  static Config ReadConfig(const std::string &filename, std::set<std::string> element_set = {}) {
    lmc_config *c = nullptr;
    lmc::Check(lmc_config_read(filename.c_str(), &c));
    return FromHandle(detail::ConfigHandle(c), std::move(element_set));
  }
  // Config::ReadMap(lattice_filename, element_filename, map_filename) (Config.cpp:814-885).
  static Config ReadMap(const std::string &lattice_filename, const std::string &element_filename,
                        const std::string &map_filename, std::set<std::string> element_set = {}) {
    lmc_config *c = nullptr;
    lmc::Check(lmc_config_read_map(lattice_filename.c_str(), element_filename.c_str(),
                                   map_filename.c_str(), &c));
    return FromHandle(detail::ConfigHandle(c), std::move(element_set));
  }
  // Config::WriteConfig (Config.cpp:663-710): the reference's text, gzip by the .gz extension.
  void WriteConfig(const std::string &filename) const {
    lmc::Check(lmc_config_write(Handle().get(), filename.c_str()));
  }
  // Config::WriteExtendedXyz(filename, {}, {}) (Config.cpp:712-809).
  void WriteExtendedXyz(const std::string &filename) const {
    lmc::Check(lmc_config_write_xyz(Handle().get(), filename.c_str()));
  }
  // Config::WriteLattice / WriteElement / WriteMap (Config.cpp:887-922).
  void WriteLattice(const std::string &filename) const {
    lmc::Check(lmc_config_write_map(Handle().get(), filename.c_str(), nullptr, nullptr));
  }
  void WriteElement(const std::string &filename) const {
    lmc::Check(lmc_config_write_map(Handle().get(), nullptr, filename.c_str(), nullptr));
  }
  void WriteMap(const std::string &filename) const {
    lmc::Check(lmc_config_write_map(Handle().get(), nullptr, nullptr, filename.c_str()));
  }

  [[nodiscard]] size_t GetNumAtoms() const { return codes_.size(); }
  [[nodiscard]] int GetNumCells() const { return n_; }
  [[nodiscard]] const std::vector<double> &GetBasis() const { return basis_; }
  [[nodiscard]] std::string GetElementAtLatticeId(size_t lattice_id) const {
    return species_.at(codes_.at(lattice_id));
  }
  [[nodiscard]] size_t GetAtomIdFromLatticeId(size_t lattice_id) const {
    return static_cast<size_t>(atoms_.at(lattice_id));
  }
  // Config::GetLatticeIdFromAtomId (Config.cpp:124-126): a search over the lattice.
  [[nodiscard]] size_t GetLatticeIdFromAtomId(size_t atom_id) const {
    for (size_t s = 0; s < atoms_.size(); ++s)
      if (static_cast<size_t>(atoms_[s]) == atom_id) return s;
    throw std::out_of_range("atom id " + std::to_string(atom_id));
  }
  // lattice id -> atom id, and map_shift [atom id * 3 + axis]
  [[nodiscard]] const std::vector<int64_t> &GetLatticeToAtomVector() const { return atoms_; }
  [[nodiscard]] const std::vector<int32_t> &GetMapShiftList() const { return shifts_; }
  void SetAtomElementTypeAtLattice(size_t lattice_id, const std::string &element) {
    codes_.at(lattice_id) = CodeOf(element);
    dirty_ = true;
  }
  // Config::LatticeJump (Config.cpp:431-456): exchanges the two sites' atoms and moves the image
  // change of the displacement lhs -> rhs from the lhs atom's map_shift to the rhs atom's.
  void LatticeJump(const std::pair<size_t, size_t> &lattice_id_jump_pair) {
    const size_t l = lattice_id_jump_pair.first, r = lattice_id_jump_pair.second;
    std::swap(codes_.at(l), codes_.at(r));
    auto xyz = [this](size_t id, int *c) {
      const int64_t plane = 2ll * n_ * n_;
      c[0] = static_cast<int>(static_cast<int64_t>(id) / plane);
      const int64_t rest = static_cast<int64_t>(id) - c[0] * plane;
      c[1] = static_cast<int>(rest / n_);
      c[2] = 2 * static_cast<int>(rest - static_cast<int64_t>(c[1]) * n_) + ((c[0] + c[1]) & 1);
    };
    int a[3], b[3];
    xyz(l, a);
    xyz(r, b);
    const int64_t al = atoms_.at(l), ar = atoms_.at(r);
    for (int d = 0; d < 3; ++d) {
      const int dd = b[d] - a[d];
      const int ic = (dd > n_) - (dd < -n_);  // floor(shift + 0.5) for |shift| > 0.5, shift = dd/2n
      shifts_[3 * al + d] -= ic;
      shifts_[3 * ar + d] += ic;
    }
    std::swap(atoms_[l], atoms_[r]);
    dirty_ = true;
  }
  // ---- host-side queries of the reference's Config (Config.cpp:87-420), on lattice and atom ids
  [[nodiscard]] std::string GetElementAtAtomId(size_t atom_id) const {
    return GetElementAtLatticeId(GetLatticeIdFromAtomId(atom_id));
  }
  [[nodiscard]] std::vector<int64_t> GetAtomToLatticeVector() const {
    std::vector<int64_t> v(atoms_.size());
    for (size_t s = 0; s < atoms_.size(); ++s) v[static_cast<size_t>(atoms_[s])] = static_cast<int64_t>(s);
    return v;
  }
  [[nodiscard]] std::map<std::string, size_t> GetElementCountMap() const {
    std::map<std::string, size_t> m;
    for (uint8_t c : codes_) ++m[species_[c]];
    return m;
  }
  // atom ids by element, ascending (Config.cpp:204-210)
  [[nodiscard]] std::map<std::string, std::vector<size_t>> GetElementAtomIdVectorMap() const {
    std::map<std::string, std::vector<size_t>> m;
    const auto a2l = GetAtomToLatticeVector();
    for (size_t a = 0; a < a2l.size(); ++a) m[species_[codes_[static_cast<size_t>(a2l[a])]]].push_back(a);
    return m;
  }
  [[nodiscard]] size_t GetVacancyAtomId() const { return GetAtomIdFromLatticeId(GetVacancyLatticeId()); }
  [[nodiscard]] double GetVacancyConcentration() const {
    const auto m = GetElementCountMap();
    const auto it = m.find("X");
    return static_cast<double>(it == m.end() ? 0 : it->second) / static_cast<double>(GetNumAtoms());
  }
  [[nodiscard]] double GetSoluteConcentration(const std::string &solvent_element) const {
    size_t k = 0;
    for (const auto &[e, c] : GetElementCountMap())
      if (e != solvent_element && e != "X") k += c;
    return static_cast<double>(k) / static_cast<double>(GetNumAtoms());
  }
  // the first element (string order) with the largest count (Config.cpp:411-421)
  [[nodiscard]] std::string GetSolventElement() const {
    std::string best = "X";
    size_t mx = 0;
    for (const auto &[e, c] : GetElementCountMap())
      if (c > mx) {
        mx = c;
        best = e;
      }
    return best;
  }
  // 1, 2 or 3 for neighbours in that shell, else -1 (Config.cpp:354-371)
  [[nodiscard]] int FindDistanceLabelBetweenLattice(size_t lattice_id1, size_t lattice_id2) const {
    for (int shell = 1; shell <= 3; ++shell) {
      const auto nb = GetNeighborLatticeIdVector(lattice_id1, shell);
      if (std::find(nb.begin(), nb.end(), lattice_id2) != nb.end()) return shell;
    }
    return -1;
  }
  [[nodiscard]] std::vector<size_t> GetFirstNeighborsAtomIdVectorOfAtom(size_t atom_id) const {
    return NeighbourAtoms(atom_id, 1);
  }
  [[nodiscard]] std::vector<size_t> GetSecondNeighborsAtomIdVectorOfAtom(size_t atom_id) const {
    return NeighbourAtoms(atom_id, 2);
  }
  [[nodiscard]] std::vector<size_t> GetThirdNeighborsAtomIdVectorOfAtom(size_t atom_id) const {
    return NeighbourAtoms(atom_id, 3);
  }
  // element counts of a shell around a site, every element of the set listed (Config.cpp:393-410)
  [[nodiscard]] std::map<std::string, size_t> GetLocalInfoOfLatticeId(size_t lattice_id, int shell) const {
    std::map<std::string, size_t> m;
    for (const auto &e : GetElementSetWithoutVacancy()) m.emplace(e, 0);
    for (size_t nb : GetNeighborLatticeIdVector(lattice_id, shell)) ++m[GetElementAtLatticeId(nb)];
    return m;
  }
  // Element masses summed over the atoms that are not the solvent (Config.cpp:148-160)
  [[nodiscard]] double GetTotalSoluteMass() const {
    const std::string solvent = GetSolventElement();
    double m = 0.0;
    for (uint8_t c : codes_)
      if (species_[c] != solvent) m += Mass(species_[c]);
    return m;
  }
  // (relative position + map_shift) times the basis (Config.cpp:266-291)
  [[nodiscard]] std::array<double, 3> GetUnwrappedCartesianPositionOfLattice(size_t lattice_id) const {
    const auto rel = RelativePosition(lattice_id);
    const int64_t a = atoms_.at(lattice_id);
    double u[3];
    for (int d = 0; d < 3; ++d) u[d] = rel[d] + static_cast<double>(shifts_[3 * a + d]);
    return Times(u);
  }
  [[nodiscard]] std::array<double, 3> GetUnwrappedCartesianPositionOfAtom(size_t atom_id) const {
    return GetUnwrappedCartesianPositionOfLattice(GetLatticeIdFromAtomId(atom_id));
  }
  // Config::AtomJump / SetAtomElementTypeAtAtom on atom ids.
  void AtomJump(const std::pair<size_t, size_t> &atom_id_jump_pair) {
    LatticeJump({GetLatticeIdFromAtomId(atom_id_jump_pair.first),
                 GetLatticeIdFromAtomId(atom_id_jump_pair.second)});
  }
  void SetAtomElementTypeAtAtom(size_t atom_id, const std::string &element) {
    SetAtomElementTypeAtLattice(GetLatticeIdFromAtomId(atom_id), element);
  }
  // Relative position of a site: X / 2n, Y / 2n, Z / 2n in a/2 units.
  [[nodiscard]] std::array<double, 3> RelativePosition(size_t lattice_id) const {
    const int64_t plane = 2ll * n_ * n_;
    const int64_t id = static_cast<int64_t>(lattice_id);
    const int64_t X = id / plane, r = id - X * plane, Y = r / n_;
    const int64_t Z = 2 * (r - Y * n_) + ((X + Y) & 1);
    const double m = 2.0 * n_;
    return {static_cast<double>(X) / m, static_cast<double>(Y) / m, static_cast<double>(Z) / m};
  }

  [[nodiscard]] size_t GetVacancyLatticeId() const {
    const uint8_t x = CodeOf("X");
    for (size_t i = 0; i < codes_.size(); ++i)
      if (codes_[i] == x) return i;
    throw std::runtime_error("vacancy not found");
  }
  [[nodiscard]] std::vector<size_t> GetNeighborLatticeIdVector(size_t lattice_id, int shell) const {
    int64_t out[24];
    const int32_t k = lmc_fcc_neighbors(n_, static_cast<int64_t>(lattice_id), shell, out);
    if (k < 0) lmc::Check(k);
    return std::vector<size_t>(out, out + k);
  }
  [[nodiscard]] std::set<std::string> GetElementSetWithoutVacancy() const {
    std::set<std::string> s(species_.begin(), species_.end());
    s.erase("X");
    return s;
  }
  // Species codes by lattice id, in the order of sorted(element_set + "X").
  [[nodiscard]] const std::vector<uint8_t> &GetCodes() const { return codes_; }
  [[nodiscard]] const std::vector<std::string> &GetSpecies() const { return species_; }
  // The state after a run on the GPU: codes by lattice id and, when the run tracked atoms, the
  // atom ids and map_shift (empty vectors keep the current ones).
  void SetState(const uint8_t *codes, const std::vector<int64_t> &atoms,
                const std::vector<int32_t> &shifts) {
    std::copy(codes, codes + codes_.size(), codes_.begin());
    if (!atoms.empty()) atoms_ = atoms;
    if (!shifts.empty()) shifts_ = shifts;
    dirty_ = true;
  }
  [[nodiscard]] const uint8_t *DeviceOccupancy() const {
    if (dirty_ || !dev_) {
      if (!dev_) dev_ = std::make_unique<lmc::DeviceBuffer>(codes_.size());
      dev_->Upload(codes_.data(), codes_.size());
      dirty_ = false;
    }
    return dev_->as<uint8_t>();
  }
  [[nodiscard]] uint8_t CodeOf(const std::string &element) const {
    for (size_t i = 0; i < species_.size(); ++i)
      if (species_[i] == element) return static_cast<uint8_t>(i);
    for (size_t i = 0; i < species_.size(); ++i)  // unknown strings are X (Element.hpp:25-26)
      if (species_[i] == "X") return static_cast<uint8_t>(i);
    return 0;
  }

 private:
  std::vector<size_t> NeighbourAtoms(size_t atom_id, int shell) const {
    std::vector<size_t> out;
    for (size_t nb : GetNeighborLatticeIdVector(GetLatticeIdFromAtomId(atom_id), shell))
      out.push_back(GetAtomIdFromLatticeId(nb));
    return out;
  }
  // Vector_d * Matrix_d (VectorMatrix.hpp:275-280)
  std::array<double, 3> Times(const double *u) const {
    std::array<double, 3> o{};
    for (int i = 0; i < 3; ++i) o[i] = u[0] * basis_[i] + u[1] * basis_[3 + i] + u[2] * basis_[6 + i];
    return o;
  }
  static double Mass(const std::string &e) {  // Element::GetMass (Element.hpp:71-88)
    const std::string b = e.size() > 1 && e[0] == 'p' ? e.substr(1) : e;
    if (b == "Al") return 26.98;
    if (b == "Mg") return 24.31;
    if (b == "Zn") return 65.38;
    if (b == "Cu") return 63.55;
    if (b == "Sn") return 118.71;
    return 0.0;
  }
  void SetSpecies(std::set<std::string> element_set_without_vacancy) {
    element_set_without_vacancy.erase("X");
    species_.assign(element_set_without_vacancy.begin(), element_set_without_vacancy.end());
    species_.push_back("X");
    std::sort(species_.begin(), species_.end());
  }
  void LoadAtoms(lmc_config *c) {
    int32_t n = 0;
    int64_t N = 0;
    basis_.assign(9, 0.0);
    lmc::Check(lmc_config_info(c, &n, &N, basis_.data()));
    n_ = n;
    atoms_.resize(static_cast<size_t>(N));
    shifts_.resize(static_cast<size_t>(3 * N));
    lmc::Check(lmc_config_atoms(c, atoms_.data(), shifts_.data()));
  }
  static Config FromHandle(detail::ConfigHandle h, std::set<std::string> element_set) {
    Config out;
    out.LoadAtoms(h.get());
    if (element_set.empty()) {  // the elements present in the file
      const auto all = detail::AllElements();
      const auto names = detail::Names(all);
      std::vector<uint8_t> c(out.atoms_.size());
      lmc::Check(lmc_config_codes(h.get(), static_cast<int32_t>(names.size()), names.data(), c.data()));
      for (uint8_t v : c) element_set.insert(all[v]);
    }
    out.SetSpecies(std::move(element_set));
    const auto names = detail::Names(out.species_);
    out.codes_.resize(out.atoms_.size());
    lmc::Check(lmc_config_codes(h.get(), static_cast<int32_t>(names.size()), names.data(),
                                out.codes_.data()));
    return out;
  }
  // A configuration handle of the current state (the box edge from the basis).
  detail::ConfigHandle Handle() const {
    lmc_config *c = nullptr;
    lmc::Check(lmc_config_generate(n_, "X", &c));
    detail::ConfigHandle h(c);
    const auto names = detail::Names(species_);
    lmc::Check(lmc_config_update(h.get(), static_cast<int32_t>(names.size()), names.data(),
                                 codes_.data(), atoms_.data(), shifts_.data()));
    lmc::Check(lmc_config_set_basis(h.get(), basis_.data()));
    return h;
  }

  int n_ = 0;
  std::vector<double> basis_;
  std::vector<std::string> species_;
  std::vector<uint8_t> codes_;
  std::vector<int64_t> atoms_;
  std::vector<int32_t> shifts_;
  mutable bool dirty_ = true;
  mutable std::unique_ptr<lmc::DeviceBuffer> dev_;
};

// GenerateFCC (Config.cpp:1060-1095) for a cubic box of factors {n, n, n}.
inline Config GenerateFCC(int n, const std::set<std::string> &element_set, const std::string &element) {
  return Config(n, element_set, element);
}

}  // namespace cfg

namespace pred {

namespace detail {
// Device copies of (offset 0, ncell) for single-replica calls.
struct One {
  lmc::DeviceBuffer off{sizeof(int64_t)}, ncell{sizeof(int32_t)}, idx{4 * sizeof(int64_t)},
      out{sizeof(double)}, code{8};
  explicit One(int n) {
    const int64_t zero = 0;
    off.Upload(&zero, sizeof(zero));
    ncell.Upload(&n, sizeof(n));
  }
};
}  // namespace detail

class EnergyPredictor {
 public:
  EnergyPredictor(const std::string &predictor_filename, const std::set<std::string> &element_set)
      : m_(predictor_filename, element_set, LMC_MODEL_ENERGY) {}
  EnergyPredictor(const std::string &predictor_filename, const std::set<Element> &element_set)
      : EnergyPredictor(predictor_filename, lmc::ElementNames(element_set)) {}
  [[nodiscard]] double GetEnergy(const cfg::Config &config) const {
    detail::One one(config.GetNumCells());
    lmc::Check(lmc_total_energy(m_.get(), config.DeviceOccupancy(), one.off.as<int64_t>(),
                                one.ncell.as<int32_t>(), 1, one.out.as<double>(), nullptr));
    double e = 0;
    one.out.Download(&e, sizeof(e));
    lmc::Check(lmc_check_errors(m_.get()));
    return e;
  }
  // GetEncode (EnergyPredictor.cpp:40-96): counts / normalisers in ECI key order.
  [[nodiscard]] std::vector<double> GetEncode(const cfg::Config &config) const {
    detail::One one(config.GetNumCells());
    int32_t n_eci = 0;
    lmc::Check(lmc_model_layout(m_.get(), &n_eci, nullptr, nullptr));
    lmc::DeviceBuffer out(sizeof(double) * static_cast<size_t>(n_eci));
    lmc::Check(lmc_total_encode(m_.get(), config.DeviceOccupancy(), one.off.as<int64_t>(),
                                one.ncell.as<int32_t>(), 1, out.as<double>(), nullptr));
    std::vector<double> enc(static_cast<size_t>(n_eci));
    out.Download(enc.data(), sizeof(double) * enc.size());
    lmc::Check(lmc_check_errors(m_.get()));
    return enc;
  }
  // GetEnergyOfCluster (EnergyPredictor.cpp:179-187): the clusters inside the atoms' three
  // neighbour shells.
  [[nodiscard]] double GetEnergyOfCluster(const cfg::Config &config,
                                          const std::vector<size_t> &atom_id_list) const {
    const auto &l2a = config.GetLatticeToAtomVector();
    std::vector<int64_t> site_of(l2a.size());
    for (size_t s = 0; s < l2a.size(); ++s) site_of[static_cast<size_t>(l2a[s])] = static_cast<int64_t>(s);
    std::vector<int64_t> sites;
    for (size_t a : atom_id_list) sites.push_back(site_of.at(a));
    double e = 0;
    lmc::Check(lmc_cluster_energy(m_.get(), config.DeviceOccupancy(), config.GetNumCells(),
                                  sites.data(), static_cast<int32_t>(sites.size()), nullptr, &e));
    lmc::Check(lmc_check_errors(m_.get()));
    return e;
  }
  // GetChemicalPotential(solvent) (EnergyPredictor.cpp:194-212), by element string.
  [[nodiscard]] std::map<std::string, double> GetChemicalPotential(const std::string &solvent) const {
    const int32_t S = lmc_model_num_species(m_.get());
    std::vector<double> mu(static_cast<size_t>(S));
    lmc::Check(lmc_chemical_potential(m_.get(), solvent.c_str(), mu.data()));
    lmc::Check(lmc_check_errors(m_.get()));
    std::map<std::string, double> out;
    for (int32_t c = 0; c < S; ++c) out[lmc_model_species_name(m_.get(), c)] = mu[static_cast<size_t>(c)];
    return out;
  }
  [[nodiscard]] lmc_model *handle() const { return m_.get(); }

 private:
  lmc::ModelHandle m_;
};

// Base of the two barrier models: GetBarrierAndDiffFromLatticeIdPair returns {Ea, dE} of the
// jump of the vacancy at pair.first to its first neighbour pair.second.
class VacancyMigrationPredictor {
 public:
  VacancyMigrationPredictor(const std::string &predictor_filename, const cfg::Config & /*reference_config*/,
                            const std::set<std::string> &element_set, int kind)
      : m_(predictor_filename, element_set, kind) {}
  virtual ~VacancyMigrationPredictor() = default;
  [[nodiscard]] virtual std::pair<double, double> GetBarrierAndDiffFromLatticeIdPair(
      const cfg::Config &config, const std::pair<size_t, size_t> &lattice_id_jump_pair) const {
    double ea = 0, de = 0;
    lmc::Check(lmc_eval_pair(m_.get(), config.DeviceOccupancy(), config.GetNumCells(),
                             static_cast<int64_t>(lattice_id_jump_pair.first),
                             static_cast<int64_t>(lattice_id_jump_pair.second), &ea, &de));
    lmc::Check(lmc_check_errors(m_.get()));
    return {ea, de};
  }
  // GetBarrierAndDiffFromAtomIdPair: the same jump named by the atoms on the two sites.
  [[nodiscard]] std::pair<double, double> GetBarrierAndDiffFromAtomIdPair(
      const cfg::Config &config, const std::pair<size_t, size_t> &atom_id_jump_pair) const {
    return GetBarrierAndDiffFromLatticeIdPair(
        config, {config.GetLatticeIdFromAtomId(atom_id_jump_pair.first),
                 config.GetLatticeIdFromAtomId(atom_id_jump_pair.second)});
  }
  [[nodiscard]] lmc_model *handle() const { return m_.get(); }

 private:
  lmc::ModelHandle m_;
};

class VacancyMigrationPredictorE0 : public VacancyMigrationPredictor {
 public:
  VacancyMigrationPredictorE0(const std::string &predictor_filename, const cfg::Config &reference_config,
                              const std::set<std::string> &element_set)
      : VacancyMigrationPredictor(predictor_filename, reference_config, element_set, LMC_MODEL_E0) {}
  VacancyMigrationPredictorE0(const std::string &predictor_filename, const cfg::Config &reference_config,
                              const std::set<Element> &element_set)
      : VacancyMigrationPredictorE0(predictor_filename, reference_config,
                                    lmc::ElementNames(element_set)) {}
};

class VacancyMigrationPredictorQuartic : public VacancyMigrationPredictor {
 public:
  VacancyMigrationPredictorQuartic(const std::string &predictor_filename,
                                   const cfg::Config &reference_config,
                                   const std::set<std::string> &element_set)
      : VacancyMigrationPredictor(predictor_filename, reference_config, element_set,
                                  LMC_MODEL_QUARTIC) {}
  VacancyMigrationPredictorQuartic(const std::string &predictor_filename,
                                   const cfg::Config &reference_config,
                                   const std::set<Element> &element_set)
      : VacancyMigrationPredictorQuartic(predictor_filename, reference_config,
                                         lmc::ElementNames(element_set)) {}
};

// The ...Lru predictors (VacancyMigrationPredictorE0Lru.h, VacancyMigrationPredictorQuarticLru.h)
// take a cache size; every jump is recomputed on the GPU, so it is accepted and unused.
class VacancyMigrationPredictorE0Lru : public VacancyMigrationPredictorE0 {
 public:
  VacancyMigrationPredictorE0Lru(const std::string &predictor_filename, const cfg::Config &reference_config,
                                 const std::set<std::string> &element_set, size_t /*cache_size*/)
      : VacancyMigrationPredictorE0(predictor_filename, reference_config, element_set) {}
  VacancyMigrationPredictorE0Lru(const std::string &predictor_filename, const cfg::Config &reference_config,
                                 const std::set<Element> &element_set, size_t cache_size)
      : VacancyMigrationPredictorE0Lru(predictor_filename, reference_config,
                                       lmc::ElementNames(element_set), cache_size) {}
};
class VacancyMigrationPredictorQuarticLru : public VacancyMigrationPredictorQuartic {
 public:
  VacancyMigrationPredictorQuarticLru(const std::string &predictor_filename,
                                      const cfg::Config &reference_config,
                                      const std::set<std::string> &element_set, size_t /*cache_size*/)
      : VacancyMigrationPredictorQuartic(predictor_filename, reference_config, element_set) {}
  VacancyMigrationPredictorQuarticLru(const std::string &predictor_filename,
                                      const cfg::Config &reference_config,
                                      const std::set<Element> &element_set, size_t cache_size)
      : VacancyMigrationPredictorQuarticLru(predictor_filename, reference_config,
                                            lmc::ElementNames(element_set), cache_size) {}
};

class EnergyChangePredictorPairSite {
 public:
  EnergyChangePredictorPairSite(const std::string &predictor_filename,
                                const cfg::Config & /*reference_config*/,
                                const std::set<std::string> &element_set)
      : m_(predictor_filename, element_set, LMC_MODEL_PAIRSITE) {}
  EnergyChangePredictorPairSite(const std::string &predictor_filename,
                                const cfg::Config &reference_config,
                                const std::set<Element> &element_set)
      : EnergyChangePredictorPairSite(predictor_filename, reference_config,
                                      lmc::ElementNames(element_set)) {}
  [[nodiscard]] double GetDeFromLatticeIdPair(const cfg::Config &config,
                                              const std::pair<size_t, size_t> &lattice_id_jump_pair) const {
    detail::One one(config.GetNumCells());
    const int64_t p[2] = {static_cast<int64_t>(lattice_id_jump_pair.first),
                          static_cast<int64_t>(lattice_id_jump_pair.second)};
    one.idx.Upload(p, sizeof(p));
    lmc::Check(lmc_swap_de(m_.get(), config.DeviceOccupancy(), one.off.as<int64_t>(),
                           one.ncell.as<int32_t>(), one.idx.as<int64_t>(), 1, one.out.as<double>(),
                           nullptr));
    return Result(one);
  }
  [[nodiscard]] double GetDeFromLatticeIdSite(const cfg::Config &config, size_t lattice_id,
                                              const std::string &new_element) const {
    detail::One one(config.GetNumCells());
    const int64_t s = static_cast<int64_t>(lattice_id);
    const uint8_t c = config.CodeOf(new_element);
    one.idx.Upload(&s, sizeof(s));
    one.code.Upload(&c, 1);
    lmc::Check(lmc_site_de(m_.get(), config.DeviceOccupancy(), one.off.as<int64_t>(),
                           one.ncell.as<int32_t>(), one.idx.as<int64_t>(), one.code.as<uint8_t>(), 1,
                           one.out.as<double>(), nullptr));
    return Result(one);
  }
  // GetDeFromAtomIdPair / GetDeFromAtomIdSite (EnergyChangePredictorPairSite.h:20-25): the same
  // changes named by atom ids.
  [[nodiscard]] double GetDeFromAtomIdPair(const cfg::Config &config,
                                           const std::pair<size_t, size_t> &atom_id_jump_pair) const {
    return GetDeFromLatticeIdPair(config, {config.GetLatticeIdFromAtomId(atom_id_jump_pair.first),
                                           config.GetLatticeIdFromAtomId(atom_id_jump_pair.second)});
  }
  [[nodiscard]] double GetDeFromAtomIdSite(const cfg::Config &config, size_t atom_id,
                                           const std::string &new_element) const {
    return GetDeFromLatticeIdSite(config, config.GetLatticeIdFromAtomId(atom_id), new_element);
  }
  [[nodiscard]] lmc_model *handle() const { return m_.get(); }

 private:
  double Result(const detail::One &one) const {
    double de = 0;
    one.out.Download(&de, sizeof(de));
    lmc::Check(lmc_check_errors(m_.get()));
    return de;
  }
  lmc::ModelHandle m_;
};

// Learned stand-in for a barrier predictor's jump evaluation (no reference counterpart): an
// ensemble of BF16 MLPs trained on the GPU on exact samples, with the predictor as fallback.
// Device pointers are from DeviceBuffer, hipMalloc or any HIP allocation.
class Surrogate {
 public:
  Surrogate(const VacancyMigrationPredictor &barrier, int members = 4, uint64_t seed = 0) {
    lmc_surrogate *s = nullptr;
    lmc::Check(lmc_surrogate_create(barrier.handle(), members, seed, &s));
    s_.reset(s);
  }
  // Adam on (environment rows [n][64], E_kra, dE) device arrays; returns the last epoch's mean
  // squared error per member and output (normalised units).
  std::vector<double> Train(const uint8_t *d_env, const double *d_ekra, const double *d_de,
                            int32_t n_rows, int32_t epochs, int32_t batch = 2048,
                            double lr = 3e-3, uint64_t shuffle_seed = 0) {
    int32_t m = 0, k = 0, p = 0, c = 0, codes[4], log_ekra = 0;
    lmc::Check(lmc_surrogate_layout(s_.get(), &m, &k, &p, &c, codes, &log_ekra));
    std::vector<double> mse(static_cast<size_t>(m) * 2);
    lmc::Check(lmc_surrogate_train(s_.get(), d_env, d_ekra, d_de, n_rows, epochs, batch, lr,
                                   shuffle_seed, mse.data()));
    return mse;
  }
  [[nodiscard]] lmc_surrogate *handle() const { return s_.get(); }

 private:
  struct Free {
    void operator()(lmc_surrogate *s) const { lmc_surrogate_free(s); }
  };
  std::unique_ptr<lmc_surrogate, Free> s_;
};

}  // namespace pred

namespace mc {

// KineticMcFirstOmp for R independent replicas advanced by one persistent kernel. Per-replica
// parameters follow the reference constructor (temperature, restart_*, maximum_steps,
// rate_corrector, early_stop, solute_disp); Simulate() runs every replica to maximum_steps.
namespace detail {
// Atom ids and map_shift of a batch of configurations, concatenated (the drivers' layout).
inline void ConcatAtoms(const std::vector<cfg::Config> &configs, std::vector<int64_t> &atoms,
                        std::vector<int32_t> &shifts) {
  for (const auto &c : configs) {
    atoms.insert(atoms.end(), c.GetLatticeToAtomVector().begin(), c.GetLatticeToAtomVector().end());
    shifts.insert(shifts.end(), c.GetMapShiftList().begin(), c.GetMapShiftList().end());
  }
}
// Splits a batch's occupancy (and tracked atoms) back into its configurations.
inline void SplitState(std::vector<cfg::Config> &configs, const std::vector<uint8_t> &occ,
                       const std::vector<int64_t> &atoms, const std::vector<int32_t> &shifts) {
  size_t at = 0;
  for (auto &c : configs) {
    const size_t N = c.GetNumAtoms();
    std::vector<int64_t> a;
    std::vector<int32_t> sh;
    if (!atoms.empty()) {
      a.assign(atoms.begin() + at, atoms.begin() + at + N);
      sh.assign(shifts.begin() + 3 * at, shifts.begin() + 3 * (at + N));
    }
    c.SetState(occ.data() + at, a, sh);
    at += N;
  }
}
}  // namespace detail

class KineticMcFirstBatch {
 public:
  KineticMcFirstBatch(const std::vector<cfg::Config> &configs, const std::vector<lmc_kmc_replica> &params,
                      const pred::VacancyMigrationPredictor &barrier, const pred::EnergyPredictor *energy) {
    if (configs.empty() || configs.size() != params.size())
      throw std::invalid_argument("one parameter set per configuration");
    std::vector<int32_t> ncell;
    std::vector<uint8_t> occ;
    for (const auto &c : configs) {
      ncell.push_back(c.GetNumCells());
      occ.insert(occ.end(), c.GetCodes().begin(), c.GetCodes().end());
    }
    lmc_kmc *k = nullptr;
    lmc::Check(lmc_kmc_create(barrier.handle(), energy ? energy->handle() : nullptr,
                              static_cast<int32_t>(configs.size()), ncell.data(), occ.data(),
                              params.data(), &k));
    k_.reset(k);
    R_ = configs.size();
    total_ = occ.size();
    for (const auto &p : params) max_steps_ = std::max<uint64_t>(max_steps_, p.maximum_steps);
  }
  // Track atom ids and map_shift from these configurations on (Config::LatticeJump).
  void TrackAtoms(const std::vector<cfg::Config> &configs) {
    std::vector<int64_t> a;
    std::vector<int32_t> sh;
    detail::ConcatAtoms(configs, a, sh);
    if (a.size() != total_) throw std::invalid_argument("configurations differ from the batch");
    lmc::Check(lmc_kmc_set_atoms(k_.get(), a.data(), sh.data()));
    tracking_ = true;
  }
  // The batch's current state back into its configurations: elements, and atom ids and
  // map_shift when tracking.
  void UpdateConfigs(std::vector<cfg::Config> &configs) const {
    std::vector<uint8_t> occ(total_);
    lmc::Check(lmc_kmc_occupancy(k_.get(), occ.data()));
    std::vector<int64_t> a;
    std::vector<int32_t> sh;
    if (tracking_) {
      a.resize(total_);
      sh.resize(3 * total_);
      lmc::Check(lmc_kmc_atoms(k_.get(), a.data(), sh.data()));
    }
    detail::SplitState(configs, occ, a, sh);
  }
  // kmc_log.txt records at the reference's dump cadence, written by WriteLog (one file per
  // replica, appended; see lmc_kmc_set_log).
  void SetLog(uint64_t log_dump_steps, uint64_t capacity = 100000) {
    lmc::Check(lmc_kmc_set_log(k_.get(), log_dump_steps, capacity, nullptr));
  }
  void WriteLog(const std::vector<std::string> &filenames) {
    if (filenames.size() != R_) throw std::invalid_argument("one log file per replica");
    std::vector<const char *> f;
    for (const auto &s : filenames) f.push_back(s.c_str());
    lmc::Check(lmc_kmc_write_log(k_.get(), f.data()));
  }
  void SetTimeTemperature(const std::vector<double> &time, const std::vector<double> &temperature) {
    lmc::Check(lmc_kmc_set_schedule(k_.get(), static_cast<int32_t>(time.size()), time.data(),
                                    temperature.data()));
  }
  int Autotune(int steps = 200) {
    int32_t chosen = 0;
    double a = 0, b = 0;
    lmc::Check(lmc_kmc_autotune(k_.get(), steps, &chosen, &a, &b));
    return chosen;
  }
  void Run(uint64_t steps, void *stream = nullptr) { lmc::Check(lmc_kmc_run(k_.get(), steps, stream)); }
  // Kernel shape (see lmc_kmc_set_formulation).
  void SetFormulation(int32_t formulation) { lmc::Check(lmc_kmc_set_formulation(k_.get(), formulation)); }
  // `launches` launches of `steps_per_launch` steps replayed from one captured HIP graph.
  void RunGraph(uint64_t steps_per_launch, int32_t launches, void *stream = nullptr) {
    lmc::Check(lmc_kmc_run_graph(k_.get(), steps_per_launch, launches, stream));
  }
  void Simulate() { Run(max_steps_ + 1); }
  // Jump evaluation by `s` (nullptr: the cluster expansion) from the next run on.
  void SetSurrogate(const pred::Surrogate *s, double threshold = 0.01, double window = 0.0) {
    lmc::Check(lmc_kmc_set_surrogate(k_.get(), s ? s->handle() : nullptr, threshold, window));
  }
  // (events evaluated, events that fell back to the cluster expansion)
  [[nodiscard]] std::pair<uint64_t, uint64_t> SurrogateStats() const {
    uint64_t ev = 0, fb = 0;
    lmc::Check(lmc_kmc_surrogate_stats(k_.get(), &ev, &fb));
    return {ev, fb};
  }
  // Exact training samples of the 12 jumps of every replica in its current state: device
  // arrays d_env[12 R][64], d_ekra[12 R], d_de[12 R].
  void Samples(const pred::VacancyMigrationPredictor &barrier, uint8_t *d_env, double *d_ekra,
               double *d_de, void *stream = nullptr) {
    uint8_t *occ = nullptr;
    const int64_t *off = nullptr, *vac = nullptr;
    const int32_t *nc = nullptr;
    lmc::Check(lmc_kmc_device_views(k_.get(), &occ, &off, &nc, &vac));
    lmc::Check(lmc_jump_samples(barrier.handle(), occ, off, nc, vac, static_cast<int32_t>(R_),
                                d_env, d_ekra, d_de, stream));
  }
  struct State {
    std::vector<uint64_t> steps;
    std::vector<double> time, energy, absolute_energy, temperature;
    std::vector<int64_t> vacancy;
  };
  [[nodiscard]] State GetState() const {
    State s;
    s.steps.resize(R_);
    s.time.resize(R_);
    s.energy.resize(R_);
    s.absolute_energy.resize(R_);
    s.temperature.resize(R_);
    s.vacancy.resize(R_);
    lmc::Check(lmc_kmc_state(k_.get(), s.steps.data(), s.time.data(), s.energy.data(),
                             s.absolute_energy.data(), s.temperature.data(), s.vacancy.data(),
                             nullptr, nullptr, nullptr));
    return s;
  }
  [[nodiscard]] std::vector<uint8_t> GetOccupancy(size_t total_sites) const {
    std::vector<uint8_t> occ(total_sites);
    lmc::Check(lmc_kmc_occupancy(k_.get(), occ.data()));
    return occ;
  }
  [[nodiscard]] lmc_kmc *handle() const { return k_.get(); }

 private:
  struct Free {
    void operator()(lmc_kmc *k) const { lmc_kmc_free(k); }
  };
  std::unique_ptr<lmc_kmc, Free> k_;
  size_t R_ = 0, total_ = 0;
  uint64_t max_steps_ = 0;
  bool tracking_ = false;
};

// KineticMcChainOmpi for R independent replicas: second-order residence-time steps.
class KineticMcChainBatch : public KineticMcFirstBatch {
 public:
  KineticMcChainBatch(const std::vector<cfg::Config> &configs, const std::vector<lmc_kmc_replica> &params,
                      const pred::VacancyMigrationPredictor &barrier, const pred::EnergyPredictor *energy)
      : KineticMcFirstBatch(configs, params, barrier, energy) {
    SetFormulation(3);
  }
};

// CanonicalMcSerial for R independent replicas (swap trials, Metropolis acceptance).
class CanonicalMcBatch {
 public:
  CanonicalMcBatch(const std::vector<cfg::Config> &configs, const std::vector<lmc_cmc_replica> &params,
                   const pred::EnergyChangePredictorPairSite &pairsite,
                   const pred::EnergyPredictor *energy) {
    if (configs.empty() || configs.size() != params.size())
      throw std::invalid_argument("one parameter set per configuration");
    std::vector<int32_t> ncell;
    std::vector<uint8_t> occ;
    for (const auto &c : configs) {
      ncell.push_back(c.GetNumCells());
      occ.insert(occ.end(), c.GetCodes().begin(), c.GetCodes().end());
    }
    lmc_cmc *c = nullptr;
    lmc::Check(lmc_cmc_create(pairsite.handle(), energy ? energy->handle() : nullptr,
                              static_cast<int32_t>(configs.size()), ncell.data(), occ.data(),
                              params.data(), &c));
    c_.reset(c);
    R_ = configs.size();
    total_ = occ.size();
  }
  void Run(uint64_t trials, void *stream = nullptr) { lmc::Check(lmc_cmc_run(c_.get(), trials, stream)); }
  // Track atom ids and map_shift through accepted swaps (Config::LatticeJump).
  void TrackAtoms(const std::vector<cfg::Config> &configs) {
    std::vector<int64_t> a;
    std::vector<int32_t> sh;
    detail::ConcatAtoms(configs, a, sh);
    if (a.size() != total_) throw std::invalid_argument("configurations differ from the batch");
    lmc::Check(lmc_cmc_set_atoms(c_.get(), a.data(), sh.data()));
    tracking_ = true;
  }
  void UpdateConfigs(std::vector<cfg::Config> &configs) const {
    std::vector<uint8_t> occ(total_);
    lmc::Check(lmc_cmc_occupancy(c_.get(), occ.data()));
    std::vector<int64_t> a;
    std::vector<int32_t> sh;
    if (tracking_) {
      a.resize(total_);
      sh.resize(3 * total_);
      lmc::Check(lmc_cmc_atoms(c_.get(), a.data(), sh.data()));
    }
    detail::SplitState(configs, occ, a, sh);
  }
  // Occupancy in LDS during runs where it fits (default), or in HBM; results are identical.
  void SetLdsOccupancy(bool on) { lmc::Check(lmc_cmc_set_lds_occupancy(c_.get(), on ? 1 : 0)); }
  // SimulatedAnnealing's adaptive schedule from t_start (see lmc_cmc_set_annealing).
  void SetAnnealing(bool on) { lmc::Check(lmc_cmc_set_annealing(c_.get(), on ? 1 : 0)); }
  // cmc_log.txt records (CanonicalMcAbstract::Dump) with ThermodynamicAveraging over the last
  // thermodynamic_averaging_steps energies; with annealing on, the SimulatedAnnealing::Dump log.
  void SetLog(uint64_t log_dump_steps, uint64_t thermodynamic_averaging_steps = 0,
              uint64_t capacity = 100000) {
    lmc::Check(lmc_cmc_set_log(c_.get(), log_dump_steps, capacity, thermodynamic_averaging_steps));
  }
  void WriteLog(const std::vector<std::string> &filenames) {
    if (filenames.size() != R_) throw std::invalid_argument("one log file per replica");
    std::vector<const char *> f;
    for (const auto &s : filenames) f.push_back(s.c_str());
    lmc::Check(lmc_cmc_write_log(c_.get(), f.data()));
  }
  [[nodiscard]] std::vector<double> GetAnnealingTemperature() const {
    std::vector<double> T(R_);
    lmc::Check(lmc_cmc_annealing_state(c_.get(), T.data(), nullptr, nullptr));
    return T;
  }
  [[nodiscard]] std::vector<double> GetEnergy() const {
    std::vector<double> e(R_);
    lmc::Check(lmc_cmc_state(c_.get(), nullptr, e.data(), nullptr, nullptr));
    return e;
  }
  [[nodiscard]] std::vector<uint64_t> GetAccepted() const {
    std::vector<uint64_t> a(R_);
    lmc::Check(lmc_cmc_state(c_.get(), nullptr, nullptr, nullptr, a.data()));
    return a;
  }
  [[nodiscard]] lmc_cmc *handle() const { return c_.get(); }

 private:
  struct Free {
    void operator()(lmc_cmc *c) const { lmc_cmc_free(c); }
  };
  std::unique_ptr<lmc_cmc, Free> c_;
  size_t R_ = 0, total_ = 0;
  bool tracking_ = false;
};

// SimulatedAnnealing (SimulatedAnnealing.h) for R independent replicas: swap trials with the
// reference's adaptive schedule from params[r].t_start over params[r].maximum_steps trials.
class SimulatedAnnealingBatch : public CanonicalMcBatch {
 public:
  SimulatedAnnealingBatch(const std::vector<cfg::Config> &configs,
                          const std::vector<lmc_cmc_replica> &params,
                          const pred::EnergyChangePredictorPairSite &pairsite,
                          const pred::EnergyPredictor *energy)
      : CanonicalMcBatch(configs, params, pairsite, energy) {
    SetAnnealing(true);
  }
};

// ---- Drop-in drivers: the reference's constructors and Simulate() for one configuration ----
namespace detail {
// TimeTemperatureInterpolator(filename) (lmc/pred/src/TimeTemperatureInterpolator.cpp:10-26):
// the lines before the first one that starts with '0' are skipped, then "time temperature"
// pairs are read to the end; the points are sorted by time (SortPoints, :32-43). An empty name
// gives no schedule.
inline void ReadTimeTemperature(const std::string &filename, std::vector<double> &time,
                                std::vector<double> &temperature) {
  time.clear();
  temperature.clear();
  if (filename.empty()) return;
  std::ifstream ifs(filename);
  if (!ifs.is_open()) throw std::runtime_error("Cannot open " + filename);
  std::string line;
  std::streampos at = ifs.tellg();
  bool found = false;
  while (std::getline(ifs, line)) {  // the reference loops on peek() != '0'
    if (!line.empty() && line[0] == '0') {
      found = true;
      break;
    }
    at = ifs.tellg();
  }
  if (!found) throw std::runtime_error("no line starting with 0 in " + filename);
  ifs.clear();
  ifs.seekg(at);
  std::vector<std::pair<double, double>> pts;
  double t = 0, T = 0;
  while (ifs >> t >> T) {
    pts.emplace_back(t, T);
    ifs.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  }
  std::sort(pts.begin(), pts.end());
  for (const auto &p : pts) {
    time.push_back(p.first);
    temperature.push_back(p.second);
  }
}
// McAbstract seeds its generator from the clock (lmc/mc/src/McAbstract.cpp:32).
inline uint64_t ClockSeed() {
  return static_cast<uint64_t>(std::chrono::system_clock::now().time_since_epoch().count());
}
// The steps at which McAbstract-derived drivers write a configuration before the step's trial:
// d.cfg.gz where d % config_dump_steps == 0, end.cfg.gz at maximum_steps
// (KineticMcAbstract.cpp:59-77, CanonicalMcAbstract.cpp:53-68). The first step of a restarted
// run writes nothing (is_restarted_). Returns the first such step after `s`, or max + 1.
inline uint64_t NextDump(uint64_t s, uint64_t config_dump_steps, uint64_t maximum_steps) {
  uint64_t next = maximum_steps + 1;
  if (s < maximum_steps) next = maximum_steps;
  if (config_dump_steps) {
    const uint64_t d = (s / config_dump_steps + 1) * config_dump_steps;
    next = std::min(next, d);
  }
  return next;
}
inline void WriteDumps(const cfg::Config &config, uint64_t step, uint64_t config_dump_steps,
                       uint64_t maximum_steps) {
  if (config_dump_steps && step % config_dump_steps == 0)
    config.WriteConfig(std::to_string(step) + ".cfg.gz");
  if (step == maximum_steps) config.WriteConfig("end.cfg.gz");
}
constexpr uint64_t kDriverChunk = 100000;  // steps per launch of the drop-in drivers
}  // namespace detail

// KineticMcFirstOmp (lmc/mc/include/KineticMcFirstOmp.h:11-25): the reference's constructor and
// Simulate() (KineticMcAbstract.cpp:140-190) on one configuration, its 12 events per step
// evaluated on the GPU. As in the reference, the barriers come from the JSON file's Quartic
// model (KineticMcAbstract.h:50) and the absolute energy from its EnergyPredictor
// (McAbstract.cpp:26). Simulate() writes kmc_log.txt, d.cfg.gz every config_dump_steps steps
// and end.cfg.gz in the working directory, and with is_early_stop escaped.cfg.gz plus the
// "t_exit: " and "steps: " lines on std::cout. thermodynamic_averaging_steps is accepted and
// unused, as the first-order driver never prints its average.
class KineticMcFirstOmp {
 public:
  KineticMcFirstOmp(cfg::Config config, unsigned long long int log_dump_steps,
                    unsigned long long int config_dump_steps, unsigned long long int maximum_steps,
                    unsigned long long int thermodynamic_averaging_steps,
                    unsigned long long int restart_steps, double restart_energy,
                    double restart_time, double temperature, const std::set<Element> &element_set,
                    const std::string &json_coefficients_filename,
                    const std::string &time_temperature_filename, bool is_rate_corrector,
                    bool is_early_stop, bool is_solute_disp)
      : config_(config.WithElementSet(lmc::ElementNames(element_set))),
        barrier_(json_coefficients_filename, config_, element_set, 100000),
        energy_(json_coefficients_filename, element_set),
        log_dump_steps_(log_dump_steps),
        config_dump_steps_(config_dump_steps),
        maximum_steps_(maximum_steps),
        seed_(detail::ClockSeed()) {
    (void)thermodynamic_averaging_steps;
    detail::ReadTimeTemperature(time_temperature_filename, tt_time_, tt_temp_);
    p_.temperature = temperature;
    p_.replica_id = 0;
    p_.restart_steps = restart_steps;
    p_.restart_energy = restart_energy;
    p_.restart_time = restart_time;
    p_.maximum_steps = maximum_steps;
    p_.rate_corrector = is_rate_corrector ? 1 : 0;
    p_.early_stop = is_early_stop ? 1 : 0;
    p_.solute_disp = is_solute_disp ? 1 : 0;
  }
  virtual ~KineticMcFirstOmp() = default;
  KineticMcFirstOmp(KineticMcFirstOmp &&) = default;

  void Simulate() {
    p_.seed = seed_;
    KineticMcFirstBatch b({config_}, {p_}, barrier_, &energy_);
    b.SetFormulation(formulation_);
    b.TrackAtoms({config_});
    if (!tt_time_.empty()) b.SetTimeTemperature(tt_time_, tt_temp_);
    if (log_dump_steps_) b.SetLog(log_dump_steps_, detail::kDriverChunk + 64);
    const uint64_t s0 = p_.restart_steps, mx = maximum_steps_;
    std::vector<cfg::Config> cfgs{config_};
    uint64_t s = s0;
    if (s0 == 0 && s <= mx) detail::WriteDumps(config_, s, config_dump_steps_, mx);
    while (s <= mx) {
      const uint64_t target = detail::NextDump(s, config_dump_steps_, mx);
      while (s < target) {
        b.Run(std::min<uint64_t>(target - s, detail::kDriverChunk));
        if (log_dump_steps_) b.WriteLog({"kmc_log.txt"});
        const auto st = b.GetState();
        const uint64_t before = s;
        s = st.steps[0];
        if (Escaped(b)) {
          b.UpdateConfigs(cfgs);
          config_ = cfgs[0];
          OnEscape(b, st.vacancy[0]);
          return;
        }
        if (s == before) throw std::runtime_error("KMC run advanced no step");
      }
      if (s <= mx) {
        b.UpdateConfigs(cfgs);
        config_ = cfgs[0];
        detail::WriteDumps(config_, s, config_dump_steps_, mx);
      }
    }
    b.UpdateConfigs(cfgs);
    config_ = cfgs[0];
    state_ = b.GetState();
  }
  // No reference counterpart: a fixed seed for the Philox stream in place of the clock, so that
  // a run can be repeated.
  void SetSeed(uint64_t seed) { seed_ = seed; }
  // The configuration after Simulate() (before it: the initial one).
  [[nodiscard]] const cfg::Config &GetConfig() const { return config_; }
  // steps, time, energy, absolute energy, temperature and vacancy after Simulate().
  [[nodiscard]] const KineticMcFirstBatch::State &GetState() const { return state_; }

 protected:
  int formulation_ = 2;  // one replica: the wide shape, a CU per replica

 private:
  static bool Escaped(KineticMcFirstBatch &b) {
    int32_t esc = 0;
    lmc::Check(lmc_kmc_state(b.handle(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, &esc));
    return esc != 0;
  }
  // IsEscaped (KineticMcAbstract.cpp:192-223): the configuration the check saw (before the
  // step's jump) goes to escaped.cfg.gz, and to (maximum_steps + 1).cfg.gz when Dump's cadence
  // names that step; the final configuration keeps the jump.
  void OnEscape(KineticMcFirstBatch &b, int64_t vacancy) {
    double t_exit = 0;
    uint64_t step = 0;
    int64_t before = -1;
    lmc::Check(lmc_kmc_escape_info(b.handle(), &t_exit, &step, &before));
    state_ = b.GetState();
    cfg::Config at_check = config_;
    at_check.LatticeJump({static_cast<size_t>(vacancy), static_cast<size_t>(before)});
    at_check.WriteConfig("escaped.cfg.gz");
    std::cout << "t_exit: " << t_exit << std::endl;
    std::cout << "steps: " << step << std::endl;
    if (config_dump_steps_ && (maximum_steps_ + 1) % config_dump_steps_ == 0)
      at_check.WriteConfig(std::to_string(maximum_steps_ + 1) + ".cfg.gz");
  }

  cfg::Config config_;
  pred::VacancyMigrationPredictorQuarticLru barrier_;
  pred::EnergyPredictor energy_;
  uint64_t log_dump_steps_, config_dump_steps_, maximum_steps_, seed_;
  lmc_kmc_replica p_{};
  std::vector<double> tt_time_, tt_temp_;
  KineticMcFirstBatch::State state_;
};

// KineticMcChainOmpi (lmc/mc/include/KineticMcChainOmpi.h): the same constructor and outputs,
// second-order residence-time steps (KineticMcChainOmpi.cpp:56-152).
class KineticMcChainOmpi : public KineticMcFirstOmp {
 public:
  template <typename... A>
  explicit KineticMcChainOmpi(A &&...args) : KineticMcFirstOmp(std::forward<A>(args)...) {
    formulation_ = 3;
  }
};

// CanonicalMcSerial (lmc/mc/include/CanonicalMcSerial.h:11-20) and Simulate()
// (CanonicalMcSerial.cpp:40-51) on one configuration at constant temperature: cmc_log.txt with
// the thermodynamic average over the last thermodynamic_averaging_steps energies, d.cfg.gz every
// config_dump_steps trials and end.cfg.gz, in the working directory.
class CanonicalMcSerial {
 public:
  CanonicalMcSerial(cfg::Config config, unsigned long long int log_dump_steps,
                    unsigned long long int config_dump_steps, unsigned long long int maximum_steps,
                    unsigned long long int thermodynamic_averaging_steps,
                    unsigned long long int restart_steps, double restart_energy,
                    double temperature, const std::set<Element> &element_set,
                    const std::string &json_coefficients_filename)
      : config_(config.WithElementSet(lmc::ElementNames(element_set))),
        pairsite_(json_coefficients_filename, config_, element_set),
        energy_(json_coefficients_filename, element_set),
        log_dump_steps_(log_dump_steps),
        config_dump_steps_(config_dump_steps),
        maximum_steps_(maximum_steps),
        averaging_(thermodynamic_averaging_steps),
        seed_(detail::ClockSeed()) {
    p_.t_start = p_.t_end = temperature;
    p_.replica_id = 0;
    p_.restart_steps = restart_steps;
    p_.restart_energy = restart_energy;
    p_.maximum_steps = maximum_steps;
  }
  virtual ~CanonicalMcSerial() = default;
  CanonicalMcSerial(CanonicalMcSerial &&) = default;

  void Simulate() {
    p_.seed = seed_;
    CanonicalMcBatch b({config_}, {p_}, pairsite_, &energy_);
    Configure(b);
    b.TrackAtoms({config_});
    if (log_dump_steps_) b.SetLog(log_dump_steps_, averaging_, detail::kDriverChunk + 64);
    const uint64_t s0 = p_.restart_steps, mx = maximum_steps_, cds = config_dump_steps_;
    std::vector<cfg::Config> cfgs{config_};
    uint64_t s = s0;
    auto run_to = [&](uint64_t target) {
      while (s < target) {
        b.Run(std::min<uint64_t>(target - s, detail::kDriverChunk));
        if (log_dump_steps_) b.WriteLog({"cmc_log.txt"});
        const uint64_t before = s;
        lmc::Check(lmc_cmc_state(b.handle(), &s, nullptr, nullptr, nullptr));
        if (s == before) throw std::runtime_error("CMC run advanced no trial");
      }
    };
    auto dump = [&]() {
      b.UpdateConfigs(cfgs);
      config_ = cfgs[0];
      detail::WriteDumps(config_, s, cds, mx);
    };
    if (s0 == 0 && s <= mx) detail::WriteDumps(config_, s, cds, mx);
    while (s <= mx) {
      run_to(detail::NextDump(s, cds, mx));
      if (s <= mx) dump();
    }
    // a started CanonicalMcOmp round runs to its end past maximum_steps, dumping on the way
    for (uint64_t left = RoundLeft(b); left > 0; left = RoundLeft(b)) {
      if (cds && s % cds == 0) dump();  // the trial of step s runs
      const uint64_t next = cds ? (s / cds + 1) * cds : std::numeric_limits<uint64_t>::max();
      run_to(std::min(s + left, next));
    }
    b.UpdateConfigs(cfgs);
    config_ = cfgs[0];
    energies_ = b.GetEnergy();
  }
  // No reference counterpart: a fixed seed for the Philox stream in place of the clock.
  void SetSeed(uint64_t seed) { seed_ = seed; }
  [[nodiscard]] const cfg::Config &GetConfig() const { return config_; }
  // The energy relative to the start (restart_energy included) after Simulate().
  [[nodiscard]] double GetEnergy() const { return energies_.empty() ? p_.restart_energy : energies_[0]; }

 protected:
  virtual void Configure(CanonicalMcBatch &) {}
  // Trials of the current round still to run (serial trials: none).
  virtual uint64_t RoundLeft(CanonicalMcBatch &) { return 0; }
  cfg::Config config_;

 private:
  pred::EnergyChangePredictorPairSite pairsite_;
  pred::EnergyPredictor energy_;
  uint64_t log_dump_steps_, config_dump_steps_, maximum_steps_, averaging_, seed_;
  lmc_cmc_replica p_{};
  std::vector<double> energies_;
};

// CanonicalMcOmp (lmc/mc/include/CanonicalMcOmp.h:9-29): the same constructor and outputs,
// trials in rounds of up to K pairs under the 3NN exclusion set (CanonicalMcOmp.cpp:40-92). The
// reference takes K from its OpenMP thread count; here K is OMP_NUM_THREADS when that is set
// (1..64), else 64, and SetParallelTrials changes it.
class CanonicalMcOmp : public CanonicalMcSerial {
 public:
  template <typename... A>
  explicit CanonicalMcOmp(A &&...args) : CanonicalMcSerial(std::forward<A>(args)...) {
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
      const long v = std::strtol(e, nullptr, 10);
      if (v >= 1) k_ = static_cast<int32_t>(std::min<long>(v, 64));
    }
  }
  // No reference counterpart: the number of pairs per round (1..64).
  void SetParallelTrials(int32_t k) {
    if (k < 1 || k > 64) throw std::invalid_argument("trials per round must be in [1, 64]");
    k_ = k;
  }

 protected:
  void Configure(CanonicalMcBatch &b) override {
    lmc::Check(lmc_cmc_set_parallel_trials(b.handle(), k_));
  }
  uint64_t RoundLeft(CanonicalMcBatch &b) override {
    int32_t left = 0;
    lmc::Check(lmc_cmc_round_left(b.handle(), &left));
    return static_cast<uint64_t>(left);
  }

 private:
  int32_t k_ = 64;
};

}  // namespace mc

namespace ansys {

// ExitTime over one configuration (lmc/ansys/include/ExitTime.h:12-63): every migration barrier
// of a vacancy placed at a site is evaluated on the GPU (lmc_vacancy_scan) in the configuration
// whose own vacancy holds the solvent; the derived quantities are host arithmetic.
class ExitTime {
 public:
  ExitTime(const cfg::Config &config, const std::string &solvent_element, double escape_temperature,
           const pred::VacancyMigrationPredictor &vacancy_migration_predictor)
      : config_(config),
        solvent_(config.CodeOf(solvent_element)),
        beta_escape_(1.0 / 8.617333262145e-5 / escape_temperature),
        predictor_(vacancy_migration_predictor) {}
  // The reference's constructor (ExitTime.cpp:6-23), elements as strings; the pair-site
  // predictor and chemical potentials serve GetProfileEnergy and GetBindingEnergy.
  ExitTime(const cfg::Config &config, const std::string &solvent_element,
           const std::set<std::string> & /*element_set*/, double escape_temperature,
           double /*current_temperature*/, const pred::EnergyPredictor & /*energy_predictor*/,
           const pred::VacancyMigrationPredictor &vacancy_migration_predictor,
           const pred::EnergyChangePredictorPairSite &energy_change_predictor_site,
           std::map<std::string, double> chemical_potential)
      : ExitTime(config, solvent_element, escape_temperature, vacancy_migration_predictor) {
    pairsite_ = &energy_change_predictor_site;
    mu_ = std::move(chemical_potential);
  }

  // GetProfileEnergy (ExitTime.cpp:473-512), by lattice id.
  [[nodiscard]] std::vector<double> GetProfileEnergy() const {
    const size_t N = config_.GetNumAtoms();
    const cfg::Config c = WithoutVacancy();
    const uint8_t x = config_.CodeOf("X");
    std::vector<int64_t> sites(N);
    std::vector<uint8_t> from(c.GetCodes()), to(N, x);
    for (size_t i = 0; i < N; ++i) sites[i] = static_cast<int64_t>(i);
    std::vector<double> de = SiteChange(c, sites, from, to);
    std::vector<double> prof(N);
    for (size_t i = 0; i < N; ++i) prof[i] = de[i] + Mu(c.GetSpecies()[from[i]]) - Mu("X");
    const int64_t v = Vacancy();
    if (v >= 0) {
      std::vector<int64_t> vs;
      std::vector<uint8_t> vf, vt;
      for (size_t e = 0; e < c.GetSpecies().size(); ++e)
        if (c.GetSpecies()[e] != "X") {
          vs.push_back(v);
          vf.push_back(static_cast<uint8_t>(e));
          vt.push_back(x);
        }
      const std::vector<double> dv = SiteChange(c, vs, vf, vt);
      double best = -std::numeric_limits<double>::infinity();
      for (size_t j = 0; j < dv.size(); ++j)
        best = std::max(best, dv[j] + Mu(c.GetSpecies()[vf[j]]) - Mu("X"));
      prof[static_cast<size_t>(v)] = best;
    }
    return prof;
  }
  // GetBindingEnergy (ExitTime.cpp:401-429): per element of the chemical potentials other than
  // X, by lattice id.
  [[nodiscard]] std::map<std::string, std::vector<double>> GetBindingEnergy() const {
    const size_t N = config_.GetNumAtoms();
    const cfg::Config c = WithoutVacancy();
    const uint8_t x = config_.CodeOf("X");
    std::vector<std::string> els;
    for (const auto &[e, m] : mu_)
      if (e != "X") els.push_back(e);
    std::vector<int64_t> sites;
    std::vector<uint8_t> from, to;
    for (const auto &e : els)
      for (size_t i = 0; i < N; ++i) {
        sites.push_back(static_cast<int64_t>(i));
        from.push_back(c.CodeOf(e));
        to.push_back(x);
      }
    const std::vector<double> de = SiteChange(c, sites, from, to);
    std::map<std::string, std::vector<double>> out;
    for (size_t k = 0; k < els.size(); ++k) {
      const double pot = Mu("X") - Mu(els[k]);
      auto &v = out[els[k]];
      for (size_t i = 0; i < N; ++i) v.push_back(de[k * N + i] - pot);
    }
    return out;
  }

  // GetBarrierListAndExitTime (ExitTime.cpp:216-244): per lattice site, the 12 migration
  // barriers of a vacancy placed there (ascending), their mean, and the exit time
  // 1 / (nu sum_j exp(-Ea_j beta_escape)).
  [[nodiscard]] std::tuple<std::vector<std::vector<double>>, std::vector<double>, std::vector<double>>
  GetBarrierListAndExitTime() const {
    const size_t N = config_.GetNumAtoms();
    std::vector<int64_t> sites(N);
    for (size_t i = 0; i < N; ++i) sites[i] = static_cast<int64_t>(i);
    int64_t vac = -1;
    const uint8_t x = config_.CodeOf("X");
    for (size_t i = 0; i < N; ++i)
      if (config_.GetCodes()[i] == x) vac = static_cast<int64_t>(i);
    lmc::DeviceBuffer d_sites(N * sizeof(int64_t)), nb(12 * N * sizeof(int64_t)),
        ea(12 * N * sizeof(double)), de(12 * N * sizeof(double));
    d_sites.Upload(sites.data(), N * sizeof(int64_t));
    lmc::Check(lmc_vacancy_scan(predictor_.handle(), config_.DeviceOccupancy(),
                                config_.GetNumCells(), vac, solvent_, d_sites.as<int64_t>(),
                                static_cast<int32_t>(N), nb.as<int64_t>(), ea.as<double>(),
                                de.as<double>(), nullptr));
    std::vector<double> h(12 * N);
    ea.Download(h.data(), h.size() * sizeof(double));
    lmc::Check(lmc_check_errors(predictor_.handle()));
    std::vector<std::vector<double>> lists(N);
    std::vector<double> mean(N), exit_time(N);
    for (size_t i = 0; i < N; ++i) {
      lists[i].assign(h.begin() + static_cast<std::ptrdiff_t>(12 * i),
                      h.begin() + static_cast<std::ptrdiff_t>(12 * i + 12));
      std::sort(lists[i].begin(), lists[i].end());
      double sum = 0.0, rate = 0.0;
      for (double b : lists[i]) {
        sum += b;
        rate += std::exp(-b * beta_escape_);
      }
      mean[i] = sum / 12.0;
      exit_time[i] = 1.0 / rate / 1e13;
    }
    return {lists, mean, exit_time};
  }

 private:
  int64_t Vacancy() const {
    const uint8_t x = config_.CodeOf("X");
    for (size_t i = 0; i < config_.GetNumAtoms(); ++i)
      if (config_.GetCodes()[i] == x) return static_cast<int64_t>(i);
    return -1;
  }
  cfg::Config WithoutVacancy() const {
    cfg::Config c = config_;
    const int64_t v = Vacancy();
    if (v >= 0) c.SetAtomElementTypeAtLattice(static_cast<size_t>(v), config_.GetSpecies()[solvent_]);
    return c;
  }
  double Mu(const std::string &e) const {
    const auto it = mu_.find(e);
    if (it == mu_.end()) throw std::invalid_argument("no chemical potential for " + e);
    return it->second;
  }
  // lmc_site_change_de over the shared occupancy of c.
  std::vector<double> SiteChange(const cfg::Config &c, const std::vector<int64_t> &sites,
                                 const std::vector<uint8_t> &from, const std::vector<uint8_t> &to) const {
    if (!pairsite_) throw std::invalid_argument("this analysis needs the pair-site predictor");
    const size_t K = sites.size();
    std::vector<int64_t> off(K, 0);
    std::vector<int32_t> nc(K, c.GetNumCells());
    lmc::DeviceBuffer d_s(K * 8), d_o(K * 8), d_n(K * 4), d_f(K), d_t(K), d_de(K * 8);
    d_s.Upload(sites.data(), K * 8);
    d_o.Upload(off.data(), K * 8);
    d_n.Upload(nc.data(), K * 4);
    d_f.Upload(from.data(), K);
    d_t.Upload(to.data(), K);
    lmc::Check(lmc_site_change_de(pairsite_->handle(), c.DeviceOccupancy(), d_o.as<int64_t>(),
                                  d_n.as<int32_t>(), d_s.as<int64_t>(), d_f.as<uint8_t>(),
                                  d_t.as<uint8_t>(), static_cast<int32_t>(K), d_de.as<double>(),
                                  nullptr));
    std::vector<double> de(K);
    d_de.Download(de.data(), K * 8);
    lmc::Check(lmc_check_errors(pairsite_->handle()));
    return de;
  }

  const cfg::Config &config_;
  int32_t solvent_;
  double beta_escape_;
  const pred::VacancyMigrationPredictor &predictor_;
  const pred::EnergyChangePredictorPairSite *pairsite_ = nullptr;
  std::map<std::string, double> mu_;
};

}  // namespace ansys
