// Host program written against the C++ mirror (include/lmc.hpp) the way reference host code
// uses pred::/mc:: objects. Driven by tests/test_cpp_mirror.py.
//
//   mirror_check cpu <json> <dir>                             host-only checks, prints "ok"
//   mirror_check gpu <n> <occ.bin> <e0.json> <pairsite.json> <dir>  one JSON line per result
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <set>
#include <string>
#include <vector>

#include "lmc.hpp"

namespace {

int Fail(const char *what) {
  std::printf("FAIL %s\n", what);
  return 1;
}

int CpuChecks(const std::string &json, const std::string &dir) {
  const std::set<std::string> els{"Al", "Mg", "Zn"};
  cfg::Config c = cfg::GenerateFCC(5, els, "Al");
  if (c.GetNumAtoms() != 500) return Fail("site count");
  if (c.GetNeighborLatticeIdVector(7, 1).size() != 12) return Fail("1NN count");
  if (c.GetNeighborLatticeIdVector(7, 2).size() != 6) return Fail("2NN count");
  if (c.GetNeighborLatticeIdVector(7, 3).size() != 24) return Fail("3NN count");
  try {
    (void)c.GetVacancyLatticeId();
    return Fail("missing vacancy accepted");
  } catch (const std::runtime_error &e) {
    if (std::string(e.what()) != "vacancy not found") return Fail("vacancy message");
  }
  c.SetAtomElementTypeAtLattice(3, "X");
  c.SetAtomElementTypeAtLattice(4, "Unobtainium");  // unknown strings are X (Element.hpp:25-26)
  if (c.GetElementAtLatticeId(4) != "X" || c.GetVacancyLatticeId() != 3) return Fail("elements");
  c.LatticeJump({3, 9});
  if (c.GetVacancyLatticeId() != 4 || c.GetElementAtLatticeId(9) != "X") return Fail("jump");
  try {
    pred::EnergyPredictor p("/nonexistent/coefficients.json", els);
    return Fail("missing file accepted");
  } catch (const std::runtime_error &e) {
    if (std::string(e.what()).find("Cannot open") == std::string::npos) return Fail("io message");
  }
  pred::VacancyMigrationPredictorE0 e0(json, c, els);  // loads and derives tables on the host
  (void)e0;
  // atoms and map_shift: GenerateFCC numbering, a jump across the x boundary
  // lattice 1 is (X, Y, Z) = (0, 0, 2): cell (0, 0, 1), basis site 0, atom 4 * 25
  if (c.GetAtomIdFromLatticeId(0) != 0 || c.GetAtomIdFromLatticeId(1) != 100) return Fail("atom ids");
  const size_t across = 9 * 2 * 5 * 5 + 1 * 5;  // (X, Y, Z) = (9, 1, 0): a first neighbour of 0
  const size_t a0 = c.GetAtomIdFromLatticeId(0), a1 = c.GetAtomIdFromLatticeId(across);
  c.LatticeJump({0, across});
  const auto &sh = c.GetMapShiftList();
  if (c.GetAtomIdFromLatticeId(across) != a0 || sh[3 * a0] != -1 || sh[3 * a1] != 1 ||
      sh[3 * a0 + 1] != 0 || sh[3 * a1 + 2] != 0)
    return Fail("map_shift");
  // WriteConfig / ReadConfig and the map files
  c.WriteConfig(dir + "/c.cfg.gz");
  const cfg::Config r = cfg::Config::ReadConfig(dir + "/c.cfg.gz", els);
  if (r.GetCodes() != c.GetCodes() || r.GetLatticeToAtomVector() != c.GetLatticeToAtomVector() ||
      r.GetMapShiftList() != c.GetMapShiftList() || r.GetSpecies() != c.GetSpecies() ||
      r.GetBasis() != c.GetBasis())
    return Fail("cfg round trip");
  const cfg::Config r2 = cfg::Config::ReadConfig(dir + "/c.cfg.gz", {"Al", "Cu", "Mg", "Zn"});
  if (r2.GetSpecies().size() != 5 || r2.GetElementAtLatticeId(4) != "X") return Fail("species set");
  const cfg::Config r3 = cfg::Config::ReadConfig(dir + "/c.cfg.gz");  // the elements in the file
  if (r3.GetSpecies() != std::vector<std::string>{"Al", "X"}) return Fail("file species");
  c.WriteLattice(dir + "/lattice.txt");
  c.WriteElement(dir + "/element.txt");
  c.WriteMap(dir + "/map.txt");
  const cfg::Config m = cfg::Config::ReadMap(dir + "/lattice.txt", dir + "/element.txt", dir + "/map.txt", els);
  if (m.GetCodes() != c.GetCodes() || m.GetLatticeToAtomVector() != c.GetLatticeToAtomVector())
    return Fail("map round trip");
  try {
    (void)cfg::Config::ReadConfig(dir + "/none.cfg");
    return Fail("missing cfg accepted");
  } catch (const std::runtime_error &e) {
    if (std::string(e.what()).find("Cannot open") == std::string::npos) return Fail("cfg io message");
  }
  {  // host queries of the reference's Config
    cfg::Config q = cfg::GenerateFCC(4, els, "Al");
    q.SetAtomElementTypeAtLattice(5, "X");
    q.SetAtomElementTypeAtLattice(9, "Mg");
    q.SetAtomElementTypeAtLattice(10, "Zn");
    const auto cnt = q.GetElementCountMap();
    if (cnt.at("Al") != 253 || cnt.at("X") != 1 || q.GetSolventElement() != "Al") return Fail("counts");
    if (q.GetVacancyConcentration() != 1.0 / 256 || q.GetSoluteConcentration("Al") != 2.0 / 256)
      return Fail("concentrations");
    if (q.GetTotalSoluteMass() != 24.31 + 65.38) return Fail("solute mass");
    const size_t nb = q.GetNeighborLatticeIdVector(5, 2)[0];
    if (q.FindDistanceLabelBetweenLattice(5, nb) != 2 || q.FindDistanceLabelBetweenLattice(5, 5) != -1)
      return Fail("distance label");
    if (q.GetElementAtAtomId(q.GetVacancyAtomId()) != "X") return Fail("vacancy atom");
    const auto m = q.GetElementAtomIdVectorMap();
    if (m.at("Mg").size() != 1 || q.GetLatticeIdFromAtomId(m.at("Mg")[0]) != 9) return Fail("atom map");
    const auto local = q.GetLocalInfoOfLatticeId(5, 1);
    if (local.size() != 3) return Fail("local info");
    const auto u0 = q.GetUnwrappedCartesianPositionOfLattice(5);
    const size_t v5 = q.GetVacancyAtomId();
    q.AtomJump({v5, q.GetAtomIdFromLatticeId(q.GetNeighborLatticeIdVector(5, 1)[0])});
    const auto u1 = q.GetUnwrappedCartesianPositionOfAtom(v5);
    const double d2 = (u1[0] - u0[0]) * (u1[0] - u0[0]) + (u1[1] - u0[1]) * (u1[1] - u0[1]) +
                      (u1[2] - u0[2]) * (u1[2] - u0[2]);
    if (std::fabs(d2 - 4.046 * 4.046 / 2) > 1e-9) return Fail("unwrapped jump length");
  }
  std::printf("ok\n");
  return 0;
}

int GpuRun(int n, const std::string &occ_path, const std::string &e0_json, const std::string &ps_json,
           const std::string &dir) {
  const std::set<std::string> els{"Al", "Mg", "Zn"};
  std::ifstream f(occ_path, std::ios::binary);
  std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  cfg::Config c = cfg::GenerateFCC(n, els, "Al");
  if (raw.size() != c.GetNumAtoms()) return Fail("occupancy size");
  for (size_t i = 0; i < raw.size(); ++i)
    c.SetAtomElementTypeAtLattice(i, c.GetSpecies().at(static_cast<uint8_t>(raw[i])));
  const pred::VacancyMigrationPredictorE0 barrier(e0_json, c, els);
  const pred::EnergyPredictor energy(e0_json, els);
  const pred::EnergyChangePredictorPairSite pairsite(ps_json, c, els);
  const size_t v = c.GetVacancyLatticeId();
  for (size_t nb : c.GetNeighborLatticeIdVector(v, 1)) {
    const auto [ea, de] = barrier.GetBarrierAndDiffFromLatticeIdPair(c, {v, nb});
    std::printf("{\"jump\": [%zu, %zu], \"ea\": %.17g, \"de\": %.17g}\n", v, nb, ea, de);
  }
  std::printf("{\"energy\": %.17g}\n", energy.GetEnergy(c));
  {
    const auto enc = energy.GetEncode(c);
    const auto mu = energy.GetChemicalPotential("Al");
    const double ec = energy.GetEnergyOfCluster(c, {c.GetAtomIdFromLatticeId(v)});
    std::printf("{\"encode_len\": %zu, \"encode0\": %.17g, \"mu_Mg\": %.17g, \"mu_X\": %.17g, "
                "\"cluster_energy\": %.17g}\n",
                enc.size(), enc[0], mu.at("Mg"), mu.at("X"), ec);
  }
  const size_t p = (v + 11) % c.GetNumAtoms(), q = (v + 137) % c.GetNumAtoms();
  std::printf("{\"swap\": [%zu, %zu], \"de\": %.17g}\n", p, q,
              pairsite.GetDeFromLatticeIdPair(c, {p, q}));
  std::printf("{\"site\": %zu, \"element\": \"Mg\", \"de\": %.17g}\n", p,
              pairsite.GetDeFromLatticeIdSite(c, p, "Mg"));
  {  // the atom-id forms name the same sites
    const size_t ap = c.GetAtomIdFromLatticeId(p), aq = c.GetAtomIdFromLatticeId(q);
    const size_t av = c.GetAtomIdFromLatticeId(v);
    const size_t nb0 = c.GetNeighborLatticeIdVector(v, 1)[0];
    const pred::VacancyMigrationPredictorE0Lru lru(e0_json, c, els, 100000);
    const bool same =
        pairsite.GetDeFromAtomIdPair(c, {ap, aq}) == pairsite.GetDeFromLatticeIdPair(c, {p, q}) &&
        pairsite.GetDeFromAtomIdSite(c, ap, "Mg") == pairsite.GetDeFromLatticeIdSite(c, p, "Mg") &&
        lru.GetBarrierAndDiffFromAtomIdPair(c, {av, c.GetAtomIdFromLatticeId(nb0)}) ==
            barrier.GetBarrierAndDiffFromLatticeIdPair(c, {v, nb0});
    std::printf("{\"atom_forms\": %s}\n", same ? "true" : "false");
  }
  lmc_kmc_replica prm{};
  prm.temperature = 600.0;
  prm.seed = 99;
  prm.replica_id = 0;
  prm.maximum_steps = 200;
  mc::KineticMcFirstBatch kmc({c}, {prm}, barrier, &energy);
  kmc.Simulate();
  const auto s = kmc.GetState();
  std::printf("{\"kmc_steps\": %llu, \"kmc_time\": %.17g, \"kmc_vacancy\": %lld}\n",
              static_cast<unsigned long long>(s.steps[0]), s.time[0], static_cast<long long>(s.vacancy[0]));
  // the same 201 steps with atom tracking and the log, the end state written as a .cfg.gz
  {
    std::vector<cfg::Config> cs{c};
    mc::KineticMcFirstBatch t(cs, {prm}, barrier, &energy);
    t.TrackAtoms(cs);
    t.SetLog(10, 1000);
    t.Simulate();
    t.UpdateConfigs(cs);
    t.WriteLog({dir + "/kmc_log.txt"});
    cs[0].WriteConfig(dir + "/end.cfg.gz");
    std::printf("{\"tracked\": \"%s\", \"log\": \"%s\"}\n", (dir + "/end.cfg.gz").c_str(),
                (dir + "/kmc_log.txt").c_str());
  }
  // the same 201 steps replayed from a captured graph of three launches
  mc::KineticMcFirstBatch g({c}, {prm}, barrier, &energy);
  g.RunGraph(67, 3);
  const auto sg = g.GetState();
  std::printf("{\"graph_steps\": %llu, \"graph_time\": %.17g, \"graph_vacancy\": %lld}\n",
              static_cast<unsigned long long>(sg.steps[0]), sg.time[0], static_cast<long long>(sg.vacancy[0]));
  // a surrogate trained briefly on the batch's own samples, with every event falling back
  mc::KineticMcFirstBatch u({c}, {prm}, barrier, &energy);
  pred::Surrogate sur(barrier, 4, 1);
  lmc::DeviceBuffer env(12 * 64), ekra(12 * sizeof(double)), dde(12 * sizeof(double));
  u.Samples(barrier, env.as<uint8_t>(), ekra.as<double>(), dde.as<double>());
  lmc::Check(lmc_synchronize(nullptr));
  const auto mse = sur.Train(env.as<uint8_t>(), ekra.as<double>(), dde.as<double>(), 12, 2, 12);
  u.SetSurrogate(&sur, -1.0, 20.0);
  u.Simulate();
  const auto su = u.GetState();
  const auto [events, fallbacks] = u.SurrogateStats();
  std::printf("{\"sur_steps\": %llu, \"sur_time\": %.17g, \"sur_vacancy\": %lld, \"sur_events\": %llu, "
              "\"sur_fallbacks\": %llu, \"sur_mse_size\": %zu}\n",
              static_cast<unsigned long long>(su.steps[0]), su.time[0], static_cast<long long>(su.vacancy[0]),
              static_cast<unsigned long long>(events), static_cast<unsigned long long>(fallbacks), mse.size());
  {  // profile and binding energies through the reference's ExitTime constructor
    const ansys::ExitTime full(c, "Al", els, 500.0, 500.0, energy, barrier, pairsite,
                               energy.GetChemicalPotential("Al"));
    const auto prof = full.GetProfileEnergy();
    const auto bind = full.GetBindingEnergy();
    std::printf("{\"profile_vac\": %.17g, \"profile_7\": %.17g, \"binding_Mg_7\": %.17g}\n",
                prof[v], prof[7], bind.at("Mg")[7]);
  }
  // exit-time analysis of the starting configuration: three sites' barrier lists
  const ansys::ExitTime et(c, "Al", 500.0, barrier);
  const auto [lists, mean, exit_time] = et.GetBarrierListAndExitTime();
  for (size_t site : {v, (v + 7) % c.GetNumAtoms(), (v + 301) % c.GetNumAtoms()}) {
    std::printf("{\"exit_site\": %zu, \"mean\": %.17g, \"exit_time\": %.17g, \"barriers\": [", site,
                mean[site], exit_time[site]);
    for (size_t j = 0; j < 12; ++j) std::printf("%s%.17g", j ? ", " : "", lists[site][j]);
    std::printf("]}\n");
  }
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "cpu" && argc == 4) return CpuChecks(argv[2], argv[3]);
    if (mode == "gpu" && argc == 7) return GpuRun(std::stoi(argv[2]), argv[3], argv[4], argv[5], argv[6]);
    std::printf("usage: mirror_check cpu <json> <dir> | gpu <n> <occ.bin> <e0.json> <pairsite.json> <dir>\n");
    return 2;
  } catch (const std::exception &e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
}
