// Configuration files of the reference: .cfg[.gz] (Config::ReadConfig / WriteConfig /
// WriteExtendedConfig, lmc/cfg/src/Config.cpp:554-710) and the lattice / element / map triple
// (ReadMap, WriteLattice, WriteElement, WriteMap, Config.cpp:814-922), on the cubic
// n x n x n-cell FCC box the engine simulates. Lattice ids are the x-major order that
// ReassignLatticeVector (Config.cpp:466-552) gives; atom ids are the file's line order, as in the
// reference.
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lmc_capi.h"
#include "lmc_element.h"
#include "lmc_error.h"

struct lmc_config {
  int32_t n = 0;
  double basis[9] = {};
  std::vector<std::string> element;  // by atom id
  std::vector<int64_t> atom;         // lattice id -> atom id
  std::vector<int32_t> shift;        // [atom id][3] periodic image (map_shift)
};

namespace {

using lmc::Guard;

constexpr double kLatticeConstant = 4.046;  // Constants.hpp:6

bool EndsWithGz(const std::string &f) {
  return f.size() >= 3 && f.compare(f.size() - 3, 3, ".gz") == 0;
}

// Whole file, through zlib when the name ends in .gz.
std::string Slurp(const std::string &filename) {
  std::string out;
  if (EndsWithGz(filename)) {
    gzFile g = gzopen(filename.c_str(), "rb");
    if (!g) throw std::runtime_error("Cannot open " + filename);
    char buf[1 << 16];
    int k;
    while ((k = gzread(g, buf, sizeof(buf))) > 0) out.append(buf, static_cast<size_t>(k));
    const bool bad = k < 0;
    gzclose(g);
    if (bad) throw std::runtime_error("Cannot decompress " + filename);
    return out;
  }
  std::FILE *f = std::fopen(filename.c_str(), "rb");
  if (!f) throw std::runtime_error("Cannot open " + filename);
  char buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, k);
  std::fclose(f);
  return out;
}

void Spill(const std::string &filename, const std::string &text) {
  if (EndsWithGz(filename)) {
    gzFile g = gzopen(filename.c_str(), "wb");
    if (!g) throw std::runtime_error("Cannot open " + filename);
    size_t at = 0;
    while (at < text.size()) {
      const unsigned chunk = static_cast<unsigned>(std::min<size_t>(text.size() - at, 1u << 30));
      if (gzwrite(g, text.data() + at, chunk) != static_cast<int>(chunk)) {
        gzclose(g);
        throw std::runtime_error("Cannot write " + filename);
      }
      at += chunk;
    }
    if (gzclose(g) != Z_OK) throw std::runtime_error("Cannot write " + filename);
    return;
  }
  std::FILE *f = std::fopen(filename.c_str(), "wb");
  if (!f) throw std::runtime_error("Cannot open " + filename);
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  if (std::fclose(f) != 0 || !ok) throw std::runtime_error("Cannot write " + filename);
}

// Token reader with the istream operations the reference uses: ignore(max, c), >> number,
// >> string, peek.
struct Reader {
  const std::string &t;
  const std::string &name;
  size_t i = 0;
  bool AtEnd() const { return i >= t.size(); }
  void IgnorePast(char c) {
    const size_t k = t.find(c, i);
    i = k == std::string::npos ? t.size() : k + 1;
  }
  void SkipSpace() {
    while (i < t.size() && std::isspace(static_cast<unsigned char>(t[i]))) ++i;
  }
  [[noreturn]] void Bad(const char *what) const {
    throw std::runtime_error("Malformed " + name + ": expected " + what + " at byte " +
                             std::to_string(i));
  }
  double Double() {
    SkipSpace();
    const char *b = t.c_str() + i;
    char *e = nullptr;
    const double v = std::strtod(b, &e);
    if (e == b) Bad("a number");
    i += static_cast<size_t>(e - b);
    return v;
  }
  long long Int() {
    SkipSpace();
    const char *b = t.c_str() + i;
    char *e = nullptr;
    const long long v = std::strtoll(b, &e, 10);
    if (e == b) Bad("an integer");
    i += static_cast<size_t>(e - b);
    return v;
  }
  std::string Word() {
    SkipSpace();
    const size_t b = i;
    while (i < t.size() && !std::isspace(static_cast<unsigned char>(t[i]))) ++i;
    if (i == b) Bad("a word");
    return t.substr(b, i - b);
  }
};

// n of a box of num_sites = 4 n^3 sites.
int32_t CellsOf(long long num_sites, const std::string &name) {
  const int32_t n = static_cast<int32_t>(std::llround(std::cbrt(static_cast<double>(num_sites) / 4)));
  if (num_sites <= 0 || 4ll * n * n * n != num_sites || n < 4 || n > 1000)
    throw std::invalid_argument(name + ": " + std::to_string(num_sites) +
                                " sites is not an n x n x n-cell FCC box with 4 <= n <= 1000");
  return n;
}

// The cubic box's basis: diagonal, equal edges.
void CheckBasis(const double *b, const std::string &name) {
  const double L = b[0];
  bool ok = L > 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      ok = ok && (r == c ? std::fabs(b[3 * r + c] - L) <= 1e-9 * L : std::fabs(b[3 * r + c]) <= 1e-9 * L);
  if (!ok) throw std::invalid_argument(name + ": the basis is not a cubic box");
}

// Lattice id of a relative position on the n-cell box (a/2 grid, x-major ids). Positions within
// 1e-3 grid steps of a grid point are accepted; 1.0 wraps to 0.
int64_t SiteOfPosition(const double *rel, int32_t n, const std::string &name) {
  const int m = 2 * n;
  int c[3];
  for (int d = 0; d < 3; ++d) {
    const double x = rel[d] * m;
    const long long k = std::llround(x);
    if (!(std::fabs(x - static_cast<double>(k)) <= 1e-3))
      throw std::invalid_argument(name + ": a position is off the FCC lattice");
    c[d] = static_cast<int>(((k % m) + m) % m);
  }
  if ((c[0] + c[1] + c[2]) & 1) throw std::invalid_argument(name + ": a position is off the FCC lattice");
  return static_cast<int64_t>(c[0]) * 2 * n * n + static_cast<int64_t>(c[1]) * n + (c[2] >> 1);
}

void PositionOfSite(int64_t id, int32_t n, double *rel) {
  const int64_t plane = 2ll * n * n;
  const int X = static_cast<int>(id / plane);
  const int64_t r = id - X * plane;
  const int Y = static_cast<int>(r / n);
  const int Z = 2 * static_cast<int>(r - static_cast<int64_t>(Y) * n) + ((X + Y) & 1);
  const double m = 2.0 * n;
  rel[0] = X / m;
  rel[1] = Y / m;
  rel[2] = Z / m;
}

void Append(std::string &s, const char *fmt, double v) {
  char buf[64];
  const int k = std::snprintf(buf, sizeof(buf), fmt, v);
  s.append(buf, static_cast<size_t>(k));
}

// Config::ReadConfig (Config.cpp:554-661) followed by ReassignLatticeVector.
std::unique_ptr<lmc_config> ReadConfigText(const std::string &text, const std::string &name) {
  Reader rd{text, name};
  rd.IgnorePast('=');  // "Number of particles = %i"
  const long long num_atoms = rd.Int();
  rd.IgnorePast('=');  // "A = 1.0 Angstrom (basic length-scale)"
  auto c = std::make_unique<lmc_config>();
  for (double &b : c->basis) {  // "H0(r,c) = %lf A"
    rd.IgnorePast('=');
    b = rd.Double();
  }
  rd.IgnorePast('\n');
  bool no_velocity = false;
  if (!rd.AtEnd() && text[rd.i] == '.') {  // ".NO_VELOCITY."
    no_velocity = true;
    rd.IgnorePast('\n');
  }
  rd.IgnorePast('=');  // "entry_count = 3" or 6 or more
  const long long entry_count = rd.Int();
  rd.IgnorePast('\n');
  while (!rd.AtEnd()) {  // header lines up to the first one that starts with a number
    while (!rd.AtEnd() && std::isspace(static_cast<unsigned char>(text[rd.i])) && text[rd.i] != '\n')
      ++rd.i;
    if (rd.AtEnd()) break;
    const char ch = text[rd.i];
    if (std::isdigit(static_cast<unsigned char>(ch)) || ch == '.' || ch == '-') break;
    rd.IgnorePast('\n');
  }
  c->n = CellsOf(num_atoms, name);
  CheckBasis(c->basis, name);
  const size_t N = static_cast<size_t>(num_atoms);
  c->element.resize(N);
  c->atom.assign(N, -1);
  c->shift.assign(3 * N, 0);
  for (size_t a = 0; a < N; ++a) {
    (void)rd.Double();  // mass (implied by the element)
    c->element[a] = lmc::CanonicalElement(rd.Word());
    double rel[3];
    for (double &x : rel) x = rd.Double();
    if (entry_count >= 6 && no_velocity)
      for (int d = 0; d < 3; ++d) c->shift[3 * a + d] = static_cast<int32_t>(rd.Int());
    rd.IgnorePast('\n');
    const int64_t s = SiteOfPosition(rel, c->n, name);
    if (c->atom[s] >= 0) throw std::invalid_argument(name + ": two atoms on one lattice site");
    c->atom[s] = static_cast<int64_t>(a);
  }
  return c;
}

// Config::WriteConfig / WriteExtendedConfig without auxiliary lists (Config.cpp:663-710). precision(16) throughout; the first position
// written switches the stream to std::fixed, so only the first atom's mass is in the general
// format.
std::string WriteConfigText(const lmc_config &c) {
  const size_t N = c.atom.size();
  std::vector<int64_t> site_of(N);
  for (size_t s = 0; s < N; ++s) site_of[static_cast<size_t>(c.atom[s])] = static_cast<int64_t>(s);
  std::string o;
  o.reserve(N * 96 + 512);
  o += "Number of particles = " + std::to_string(N) + "\n";
  o += "A = 1.0 Angstrom (basic length-scale)\n";
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) {
      o += "H0(" + std::to_string(r + 1) + "," + std::to_string(q + 1) + ") = ";
      Append(o, "%.16g", c.basis[3 * r + q]);
      o += " A\n";
    }
  o += ".NO_VELOCITY.\nentry_count = 6\nauxiliary[0] = ix\nauxiliary[1] = iy\nauxiliary[2] = iz\n";
  char buf[96];
  for (size_t a = 0; a < N; ++a) {
    Append(o, a == 0 ? "%.16g" : "%.16f", lmc::ElementMass(c.element[a]));
    o += '\n';
    o += c.element[a];
    o += '\n';
    double rel[3];
    PositionOfSite(site_of[a], c.n, rel);
    const int k = std::snprintf(buf, sizeof(buf), "%.16f %.16f %.16f %d %d %d\n", rel[0], rel[1],
                                rel[2], c.shift[3 * a], c.shift[3 * a + 1], c.shift[3 * a + 2]);
    o.append(buf, static_cast<size_t>(k));
  }
  return o;
}

// The 12 / 6 / 24 neighbour ids of site s in ascending order (UpdateNeighbors,
// Config.cpp:955-1045).
void Neighbours(int32_t n, int64_t s, std::string &o) {
  for (int shell = 1; shell <= 3; ++shell) {
    int64_t nb[24];
    const int k = lmc_fcc_neighbors(n, s, shell, nb);
    if (k < 0) throw std::invalid_argument(lmc::g_last_error);
    for (int j = 0; j < k; ++j) {
      o += std::to_string(nb[j]);
      o += ' ';
    }
  }
}

lmc_config *Release(std::unique_ptr<lmc_config> c) { return c.release(); }

}  // namespace

int lmc_config_generate(int32_t n, const char *element, lmc_config **out) {
  return Guard([&] {
    if (!element || !out) throw std::invalid_argument("null argument");
    if (n < 4 || n > 1000) throw std::invalid_argument("lattice size n must be in [4, 1000]");
    auto c = std::make_unique<lmc_config>();
    c->n = n;
    for (int d = 0; d < 3; ++d) c->basis[4 * d] = kLatticeConstant * n;
    const size_t N = 4ull * n * n * n;
    c->element.assign(N, lmc::CanonicalElement(element));
    c->atom.resize(N);
    c->shift.assign(3 * N, 0);
    // GenerateFCC (Config.cpp:1060-1095) numbers atoms cell by cell (z outermost, x innermost),
    // four basis sites per cell: (0,0,0), (1,1,0), (1,0,1), (0,1,1) in a/2 units.
    for (size_t s = 0; s < N; ++s) {
      double rel[3];
      PositionOfSite(static_cast<int64_t>(s), n, rel);
      int X[3];
      for (int d = 0; d < 3; ++d) X[d] = static_cast<int>(std::llround(rel[d] * 2 * n));
      const int b = (X[0] & 1) ? ((X[1] & 1) ? 1 : 2) : ((X[1] & 1) ? 3 : 0);
      c->atom[s] = ((static_cast<int64_t>(X[2] >> 1) * n + (X[1] >> 1)) * n + (X[0] >> 1)) * 4 + b;
    }
    *out = Release(std::move(c));
    return LMC_OK;
  });
}

int lmc_config_read(const char *filename, lmc_config **out) {
  return Guard([&] {
    if (!filename || !out) throw std::invalid_argument("null argument");
    const std::string text = Slurp(filename);
    *out = Release(ReadConfigText(text, filename));
    return LMC_OK;
  });
}

int lmc_config_read_map(const char *lattice_filename, const char *element_filename,
                        const char *map_filename, lmc_config **out) {
  return Guard([&] {
    if (!lattice_filename || !element_filename || !map_filename || !out)
      throw std::invalid_argument("null argument");
    const std::string lname = lattice_filename;
    const std::string lt = Slurp(lname);
    Reader rl{lt, lname};
    const long long num_atoms = rl.Int();
    rl.IgnorePast('\n');
    auto c = std::make_unique<lmc_config>();
    for (double &b : c->basis) b = rl.Double();
    c->n = CellsOf(num_atoms, lname);
    CheckBasis(c->basis, lname);
    const size_t N = static_cast<size_t>(num_atoms);
    for (size_t s = 0; s < N; ++s) {  // positions must be in the engine's x-major order
      double rel[3];
      for (double &x : rel) x = rl.Double();
      if (SiteOfPosition(rel, c->n, lname) != static_cast<int64_t>(s))
        throw std::invalid_argument(lname + ": lattice ids are not in x-major position order");
      rl.IgnorePast('\n');  // the neighbour lists follow from the positions
    }
    const std::string ename = element_filename;
    const std::string et = Slurp(ename);
    Reader re{et, ename};
    c->element.resize(N);
    for (size_t a = 0; a < N; ++a) {
      c->element[a] = lmc::CanonicalElement(re.Word());
      re.IgnorePast('\n');
    }
    const std::string mname = map_filename;
    const std::string mt = Slurp(mname);
    Reader rm{mt, mname};
    c->atom.assign(N, -1);
    for (size_t a = 0; a < N; ++a) {
      const long long s = rm.Int();
      if (s < 0 || static_cast<size_t>(s) >= N)
        throw std::invalid_argument("Lattice id out of bounds in map file: " + std::to_string(s));
      if (c->atom[static_cast<size_t>(s)] >= 0)
        throw std::invalid_argument("Duplicate lattice id in map file: " + std::to_string(s));
      c->atom[static_cast<size_t>(s)] = static_cast<int64_t>(a);
      rm.IgnorePast('\n');
    }
    c->shift.assign(3 * N, 0);
    *out = Release(std::move(c));
    return LMC_OK;
  });
}

void lmc_config_free(lmc_config *c) { delete c; }

int lmc_config_info(const lmc_config *c, int32_t *ncell, int64_t *num_sites, double *basis) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null config");
    if (ncell) *ncell = c->n;
    if (num_sites) *num_sites = static_cast<int64_t>(c->atom.size());
    if (basis) std::memcpy(basis, c->basis, sizeof(c->basis));
    return LMC_OK;
  });
}

int lmc_config_codes(const lmc_config *c, int32_t n_species, const char *const *species,
                     uint8_t *codes) {
  return Guard([&] {
    if (!c || !species || !codes || n_species <= 0 || n_species > 255)
      throw std::invalid_argument("bad argument");
    std::vector<std::string> sp(species, species + n_species);
    for (size_t s = 0; s < c->atom.size(); ++s) {
      const std::string &e = c->element[static_cast<size_t>(c->atom[s])];
      const auto it = std::find(sp.begin(), sp.end(), e);
      if (it == sp.end())
        throw std::invalid_argument("element " + e + " of the configuration is not in the species list");
      codes[s] = static_cast<uint8_t>(it - sp.begin());
    }
    return LMC_OK;
  });
}

int lmc_config_atoms(const lmc_config *c, int64_t *lattice_to_atom, int32_t *map_shift) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null config");
    if (lattice_to_atom) std::copy(c->atom.begin(), c->atom.end(), lattice_to_atom);
    if (map_shift) std::copy(c->shift.begin(), c->shift.end(), map_shift);
    return LMC_OK;
  });
}

int lmc_config_update(lmc_config *c, int32_t n_species, const char *const *species,
                      const uint8_t *codes, const int64_t *lattice_to_atom,
                      const int32_t *map_shift) {
  return Guard([&] {
    if (!c || !species || !codes || n_species <= 0) throw std::invalid_argument("bad argument");
    const size_t N = c->atom.size();
    std::vector<int64_t> atom = c->atom;
    if (lattice_to_atom) {
      std::vector<char> seen(N, 0);
      for (size_t s = 0; s < N; ++s) {
        const int64_t a = lattice_to_atom[s];
        if (a < 0 || static_cast<size_t>(a) >= N || seen[static_cast<size_t>(a)])
          throw std::invalid_argument("lattice_to_atom is not a permutation of the atom ids");
        seen[static_cast<size_t>(a)] = 1;
      }
      atom.assign(lattice_to_atom, lattice_to_atom + N);
    }
    std::vector<std::string> element(N);
    for (size_t s = 0; s < N; ++s) {
      if (codes[s] >= n_species) throw std::invalid_argument("species code out of range");
      element[static_cast<size_t>(atom[s])] = lmc::CanonicalElement(species[codes[s]]);
    }
    c->atom = std::move(atom);
    c->element = std::move(element);
    if (map_shift) c->shift.assign(map_shift, map_shift + 3 * N);
    return LMC_OK;
  });
}

int lmc_config_set_basis(lmc_config *c, const double *basis) {
  return Guard([&] {
    if (!c || !basis) throw std::invalid_argument("null argument");
    CheckBasis(basis, "basis");
    std::memcpy(c->basis, basis, sizeof(c->basis));
    return LMC_OK;
  });
}

int lmc_config_write(const lmc_config *c, const char *filename) {
  return Guard([&] {
    if (!c || !filename) throw std::invalid_argument("null argument");
    Spill(filename, WriteConfigText(*c));
    return LMC_OK;
  });
}

int lmc_config_write_xyz(const lmc_config *c, const char *filename) {
  return Guard([&] {
    if (!c || !filename) throw std::invalid_argument("null argument");
    // Config::WriteExtendedXyz without auxiliary or global lists (Config.cpp:712-809):
    // precision(16), every double through a vector (std::fixed), atoms in atom-id order with
    // their Cartesian position (relative position times the basis) and map_shift.
    const size_t N = c->atom.size();
    std::vector<int64_t> site_of(N);
    for (size_t s = 0; s < N; ++s) site_of[static_cast<size_t>(c->atom[s])] = static_cast<int64_t>(s);
    std::string o = std::to_string(N) + "\nLattice=\"";
    char buf[160];
    for (int r = 0; r < 3; ++r) {
      const int k = std::snprintf(buf, sizeof(buf), "%s%.16f %.16f %.16f", r ? " " : "",
                                  c->basis[3 * r], c->basis[3 * r + 1], c->basis[3 * r + 2]);
      o.append(buf, static_cast<size_t>(k));
    }
    o += "\" pbc=\"T T T\" Properties=species:S:1:pos:R:3:map_shift:I:3\n";
    for (size_t a = 0; a < N; ++a) {
      double rel[3], cart[3];
      PositionOfSite(site_of[a], c->n, rel);
      for (int d = 0; d < 3; ++d)  // Vector_d * Matrix_d (VectorMatrix.hpp:275-280)
        cart[d] = rel[0] * c->basis[d] + rel[1] * c->basis[3 + d] + rel[2] * c->basis[6 + d];
      o += c->element[a];
      const int k = std::snprintf(buf, sizeof(buf), " %.16f %.16f %.16f %d %d %d \n", cart[0], cart[1],
                                  cart[2], c->shift[3 * a], c->shift[3 * a + 1], c->shift[3 * a + 2]);
      o.append(buf, static_cast<size_t>(k));
    }
    Spill(filename, o);
    return LMC_OK;
  });
}

int lmc_config_write_map(const lmc_config *c, const char *lattice_filename,
                         const char *element_filename, const char *map_filename) {
  return Guard([&] {
    if (!c) throw std::invalid_argument("null config");
    const size_t N = c->atom.size();
    if (lattice_filename) {  // WriteLattice (Config.cpp:887-906)
      std::string o = std::to_string(N) + " positions in total\n";
      char buf[96];
      for (int r = 0; r < 3; ++r) {
        const int k = std::snprintf(buf, sizeof(buf), "%.16f %.16f %.16f\n", c->basis[3 * r],
                                    c->basis[3 * r + 1], c->basis[3 * r + 2]);
        o.append(buf, static_cast<size_t>(k));
      }
      for (size_t s = 0; s < N; ++s) {
        double rel[3];
        PositionOfSite(static_cast<int64_t>(s), c->n, rel);
        const int k = std::snprintf(buf, sizeof(buf), "%.16f %.16f %.16f # ", rel[0], rel[1], rel[2]);
        o.append(buf, static_cast<size_t>(k));
        Neighbours(c->n, static_cast<int64_t>(s), o);
        o += '\n';
      }
      Spill(lattice_filename, o);
    }
    if (element_filename) {  // WriteElement (Config.cpp:908-914)
      std::string o;
      for (const auto &e : c->element) o += e + "\n";
      Spill(element_filename, o);
    }
    if (map_filename) {  // WriteMap (Config.cpp:916-922)
      std::vector<int64_t> site_of(N);
      for (size_t s = 0; s < N; ++s) site_of[static_cast<size_t>(c->atom[s])] = static_cast<int64_t>(s);
      std::string o;
      for (int64_t s : site_of) o += std::to_string(s) + "\n";
      Spill(map_filename, o);
    }
    return LMC_OK;
  });
}
