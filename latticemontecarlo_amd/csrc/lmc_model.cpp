// Coefficient loading and derivation of the kernels' weight tables.
//
// The barrier models are log-linear in orbit-averaged one-hot features
// (VacancyMigrationPredictorE0.cpp:128-151, VacancyMigrationPredictorQuartic.cpp:167-246):
//   logit = theta . (U ((x - mu) / sigma)),   x_f = n_f / |orbit(f)|
// which is rewritten, exactly in real arithmetic, as
//   logit = c0 + sum over clusters C of b[f(C)],
//   g = U^T theta,  b_f = g_f / (sigma_f |orbit(f)|),  c0 = -sum_f g_f mu_f / sigma_f,
// so a kernel adds one table entry per cluster. dE uses w = theta / N_label per cluster type
// (VacancyMigrationPredictorQuartic.cpp:13-14, :156-165).
#include "lmc_model.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace lmc {

namespace {

// Small JSON reader: objects, arrays, numbers (strtod, correctly rounded), strings, literals.
struct JValue {
  enum Type { kNull, kNumber, kString, kArray, kObject, kBool } type = kNull;
  double number = 0;
  std::string text;
  std::vector<JValue> items;
  std::vector<std::string> keys;

  const JValue *Get(const std::string &k) const {
    for (size_t i = 0; i < keys.size(); ++i)
      if (keys[i] == k) return &items[i];
    return nullptr;
  }
  const JValue &At(const std::string &k) const {
    const JValue *v = Get(k);
    if (!v) throw std::runtime_error("[json.exception.out_of_range.403] key '" + k + "' not found");
    return *v;
  }
  std::vector<double> Vector() const {
    if (type != kArray) throw std::runtime_error("expected a JSON array of numbers");
    std::vector<double> v;
    v.reserve(items.size());
    for (const auto &x : items) {
      if (x.type != kNumber) throw std::runtime_error("expected a JSON number");
      v.push_back(x.number);
    }
    return v;
  }
  std::vector<std::vector<double>> Matrix() const {
    if (type != kArray) throw std::runtime_error("expected a JSON array of arrays");
    std::vector<std::vector<double>> m;
    for (const auto &row : items) m.push_back(row.Vector());
    return m;
  }
  double Number() const {
    if (type != kNumber) throw std::runtime_error("expected a JSON number");
    return number;
  }
};

class JReader {
 public:
  explicit JReader(const std::string &s) : s_(s) {}
  JValue Parse() {
    JValue v = Value();
    Skip();
    if (i_ != s_.size()) Fail("trailing characters");
    return v;
  }

 private:
  [[noreturn]] void Fail(const char *what) {
    throw std::runtime_error(std::string("JSON parse error: ") + what + " at offset " +
                             std::to_string(i_));
  }
  void Skip() {
    while (i_ < s_.size() && std::strchr(" \t\r\n", s_[i_])) ++i_;
  }
  JValue Value() {
    Skip();
    if (i_ >= s_.size()) Fail("unexpected end");
    const char c = s_[i_];
    JValue v;
    if (c == '{') {
      v.type = JValue::kObject;
      ++i_;
      Skip();
      if (i_ < s_.size() && s_[i_] == '}') {
        ++i_;
        return v;
      }
      for (;;) {
        Skip();
        JValue k = Value();
        if (k.type != JValue::kString) Fail("object key must be a string");
        Skip();
        if (i_ >= s_.size() || s_[i_] != ':') Fail("expected ':'");
        ++i_;
        v.keys.push_back(k.text);
        v.items.push_back(Value());
        Skip();
        if (i_ < s_.size() && s_[i_] == ',') {
          ++i_;
          continue;
        }
        if (i_ < s_.size() && s_[i_] == '}') {
          ++i_;
          break;
        }
        Fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.type = JValue::kArray;
      ++i_;
      Skip();
      if (i_ < s_.size() && s_[i_] == ']') {
        ++i_;
        return v;
      }
      for (;;) {
        v.items.push_back(Value());
        Skip();
        if (i_ < s_.size() && s_[i_] == ',') {
          ++i_;
          continue;
        }
        if (i_ < s_.size() && s_[i_] == ']') {
          ++i_;
          break;
        }
        Fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.type = JValue::kString;
      ++i_;
      while (i_ < s_.size() && s_[i_] != '"') {
        if (s_[i_] == '\\' && i_ + 1 < s_.size()) ++i_;
        v.text += s_[i_++];
      }
      if (i_ >= s_.size()) Fail("unterminated string");
      ++i_;
    } else if (s_.compare(i_, 4, "true") == 0 || s_.compare(i_, 5, "false") == 0) {
      v.type = JValue::kBool;
      i_ += s_[i_] == 't' ? 4 : 5;
    } else if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4;
    } else {
      const char *b = s_.c_str() + i_;
      char *e = nullptr;
      v.type = JValue::kNumber;
      v.number = std::strtod(b, &e);
      if (e == b) Fail("bad number");
      i_ += static_cast<size_t>(e - b);
    }
    return v;
  }
  const std::string &s_;
  size_t i_ = 0;
};

// Element strings the reference knows (lmc/cfg/include/Element.hpp:10-24); others become X.
std::string CanonicalElement(const std::string &s) {
  static const char *known[] = {"X", "Al", "Mg", "Zn", "Cu", "Sn"};
  for (const char *k : known)
    if (s == k) return s;
  return "X";
}

}  // namespace

Model *LoadModel(const std::string &json_path, const std::vector<std::string> &elements,
                 int kind) {
  if (kind < kEnergy || kind > kPairSite) throw std::runtime_error("unknown model kind");
  auto m = std::make_unique<Model>();
  m->kind = kind;
  std::vector<std::string> el;
  for (const auto &e : elements) {
    const std::string c = CanonicalElement(e);
    if (c == "X") continue;  // the element set is given without the vacancy
    if (std::find(el.begin(), el.end(), c) == el.end()) el.push_back(c);
  }
  std::sort(el.begin(), el.end());
  if (el.empty()) throw std::runtime_error("element set is empty");
  m->elements = el;
  m->eci = BuildEciSpace(el);
  m->S = m->eci.S();
  m->ne = static_cast<int>(el.size());
  m->x_code = m->eci.x_code;
  if (m->S > kMaxSpecies) throw std::runtime_error("too many elements");
  m->onehot.assign(m->S, -1);
  for (int c = 0, r = 0; c < m->S; ++c)
    if (c != m->x_code) m->onehot[c] = r++;

  std::ifstream ifs(json_path);
  if (!ifs) throw std::runtime_error("Cannot open " + json_path);
  std::stringstream ss;
  ss << ifs.rdbuf();
  const std::string text = ss.str();
  const JValue root = JReader(text).Parse();
  if (root.type != JValue::kObject) throw std::runtime_error("coefficient file is not an object");

  const JumpTables &jt = GetJumpTables();
  jt.mmm.Layout(m->ne, m->base_mmm, m->len_mmm, m->F_mmm);
  jt.mm2.Layout(m->ne, m->base_mm2, m->len_mm2, m->F_mm2);

  m->bar.assign(m->S, {});
  bool have_theta = false;
  for (size_t i = 0; i < root.keys.size(); ++i) {
    const std::string &key = root.keys[i];
    const JValue &val = root.items[i];
    if (key == "Base") {
      m->theta = val.At("theta").Vector();
      have_theta = true;
      continue;
    }
    if (kind != kE0 && kind != kQuartic) continue;
    const std::string elem = CanonicalElement(key);
    int c = -1;
    for (int s = 0; s < m->S; ++s)
      if (m->eci.species[s] == elem) c = s;
    BarrierParams p;
    p.present = true;
    p.mu_mmm = val.At("mu_x_mmm").Vector();
    if (kind == kQuartic) p.mu_mm2 = val.At("mu_x_mm2").Vector();
    p.sigma_mmm = val.At("sigma_x_mmm").Vector();
    if (kind == kQuartic) p.sigma_mm2 = val.At("sigma_x_mm2").Vector();
    p.U_mmm = val.At("U_mmm").Matrix();
    if (kind == kQuartic) p.U_mm2 = val.At("U_mm2").Matrix();
    if (kind == kE0) {
      p.theta_a = val.At("theta_e0").Vector();
      p.mu_a = val.At("mu_e0").Number();
      p.sigma_a = val.At("sigma_e0").Number();
    } else {
      p.theta_a = val.At("theta_D").Vector();
      p.theta_b = val.At("theta_Ks").Vector();
      p.mu_a = val.At("mu_D").Number();
      p.mu_b = val.At("mu_Ks").Number();
      p.sigma_a = val.At("sigma_D").Number();
      p.sigma_b = val.At("sigma_Ks").Number();
    }
    auto check = [](const std::vector<double> &v, int n, const char *what) {
      if (static_cast<int>(v.size()) != n)
        throw std::runtime_error(std::string(what) + " has length " + std::to_string(v.size()) +
                                 ", expected " + std::to_string(n));
    };
    check(p.mu_mmm, m->F_mmm, "mu_x_mmm");
    check(p.sigma_mmm, m->F_mmm, "sigma_x_mmm");
    for (auto &row : p.U_mmm) check(row, m->F_mmm, "U_mmm row");
    check(p.theta_a, static_cast<int>(p.U_mmm.size()), kind == kE0 ? "theta_e0" : "theta_D");
    if (kind == kQuartic) {
      check(p.mu_mm2, m->F_mm2, "mu_x_mm2");
      check(p.sigma_mm2, m->F_mm2, "sigma_x_mm2");
      for (auto &row : p.U_mm2) check(row, m->F_mm2, "U_mm2 row");
      check(p.theta_b, static_cast<int>(p.U_mm2.size()), "theta_Ks");
    }
    if (c >= 0) m->bar[c] = std::move(p);
  }
  if (!have_theta) throw std::runtime_error("[json.exception.out_of_range.403] key 'Base' not found");
  if (static_cast<int>(m->theta.size()) != static_cast<int>(m->eci.keys.size()))
    throw std::runtime_error("Base.theta has length " + std::to_string(m->theta.size()) +
                             ", expected " + std::to_string(m->eci.keys.size()));

  // dE weights per cluster type: theta / N_label (ΔE normalisers, ...Quartic.cpp:13-14).
  static const double kDeNorm[8] = {256, 1536, 768, 3072, 2048, 3072, 6144, 6144};
  static const double kTotNorm[8] = {256, 3072, 1536, 6144, 12288, 6144, 12288, 6144};
  const int S = m->S;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  m->w.assign(S + 3 * S * S + 4 * S * S * S, nan);
  auto wof = [&](int idx) { return m->theta[idx] / kDeNorm[m->eci.keys[idx].label]; };
  for (int c = 0; c < S; ++c)
    if (m->eci.t1[c] >= 0) m->w[c] = wof(m->eci.t1[c]);
  for (int i = 0; i < 3 * S * S; ++i)
    if (m->eci.t2[i] >= 0) m->w[S + i] = wof(m->eci.t2[i]);
  for (int i = 0; i < 4 * S * S * S; ++i)
    if (m->eci.t3[i] >= 0) m->w[S + 3 * S * S + i] = wof(m->eci.t3[i]);
  for (const auto &k : m->eci.keys) m->norm_total.push_back(kTotNorm[k.label]);

  // One-hot type index by (kind, codes) (GetOneHotEncodeHashmap, EnergyUtility.cpp:7-43).
  m->type_tab.assign(3 * S * S, -1);
  const int ne = m->ne;
  for (int a = 0; a < S; ++a)
    for (int b = 0; b < S; ++b) {
      const int ia = m->onehot[a], ib = m->onehot[b];
      if (ia < 0) continue;
      m->type_tab[(0 * S + a) * S + b] = static_cast<int8_t>(ia);
      if (ib < 0) continue;
      m->type_tab[(1 * S + a) * S + b] = static_cast<int8_t>(ia * ne + ib);
      const int lo = std::min(ia, ib), hi = std::max(ia, ib);
      const int tri = lo * ne - lo * (lo - 1) / 2 + (hi - lo);
      m->type_tab[(2 * S + a) * S + b] = static_cast<int8_t>(tri);
    }

  // Barrier weights: b_f = g_f / (sigma_f |o_f|), c0 = -sum g_f mu_f / sigma_f, g = U^T theta.
  m->b_mmm.assign(static_cast<size_t>(S) * m->F_mmm, 0.0);
  m->b_mm2.assign(static_cast<size_t>(S) * m->F_mm2, 0.0);
  m->par.assign(static_cast<size_t>(S) * 6, 0.0);
  auto derive = [](const std::vector<std::vector<double>> &U, const std::vector<double> &th,
                   const std::vector<double> &mu, const std::vector<double> &sg,
                   const std::vector<int> &base, const std::vector<int> &len,
                   const std::vector<int> &osize, double *b, double &c0) {
    const size_t F = mu.size();
    long double acc0 = 0;
    for (size_t o = 0; o < base.size(); ++o)
      for (int t = 0; t < len[o]; ++t) {
        const size_t f = static_cast<size_t>(base[o] + t);
        long double g = 0;
        for (size_t r = 0; r < U.size(); ++r)
          g += static_cast<long double>(th[r]) * static_cast<long double>(U[r][f]);
        b[f] = static_cast<double>(g / (static_cast<long double>(sg[f]) * osize[o]));
        acc0 -= g * static_cast<long double>(mu[f]) / static_cast<long double>(sg[f]);
      }
    (void)F;
    c0 = static_cast<double>(acc0);
  };
  for (int c = 0; c < S; ++c) {
    const BarrierParams &p = m->bar[c];
    if (!p.present) continue;
    double *pr = &m->par[static_cast<size_t>(c) * 6];
    derive(p.U_mmm, p.theta_a, p.mu_mmm, p.sigma_mmm, m->base_mmm, m->len_mmm,
           jt.mmm.orbit_size, &m->b_mmm[static_cast<size_t>(c) * m->F_mmm], pr[0]);
    pr[1] = p.mu_a;
    pr[2] = p.sigma_a;
    if (kind == kQuartic) {
      derive(p.U_mm2, p.theta_b, p.mu_mm2, p.sigma_mm2, m->base_mm2, m->len_mm2,
             jt.mm2.orbit_size, &m->b_mm2[static_cast<size_t>(c) * m->F_mm2], pr[3]);
      pr[4] = p.mu_b;
      pr[5] = p.sigma_b;
    }
  }
  return m.release();
}

}  // namespace lmc
