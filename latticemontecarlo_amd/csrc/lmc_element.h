// Element names and masses (lmc/cfg/include/Element.hpp:7-88).
#pragma once
#include <string>

namespace lmc {

// Element(const std::string &) (Element.hpp:13-27): a name outside the list is the vacancy X.
inline std::string CanonicalElement(const std::string &e) {
  static const char *const kNames[] = {"X",   "Al",  "Mg",  "Zn",  "Cu", "Sn",
                                       "pAl", "pMg", "pZn", "pCu", "pSn"};
  for (const char *k : kNames)
    if (e == k) return e;
  return "X";
}

// Element::GetMass (Element.hpp:71-88)
inline double ElementMass(const std::string &e) {
  const std::string b = e.size() > 1 && e[0] == 'p' ? e.substr(1) : e;
  if (b == "Al") return 26.98;
  if (b == "Mg") return 24.31;
  if (b == "Zn") return 65.38;
  if (b == "Cu") return 63.55;
  if (b == "Sn") return 118.71;
  return 0.0;
}

}  // namespace lmc
