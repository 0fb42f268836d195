"""C ABI checks that need no GPU: the library loads, exports every function the header declares,
and its host-side entry points (model loading, layout, geometry, Philox) behave like the
reference's constructors and helpers, errors included."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import latticemontecarlo_amd as L
from latticemontecarlo_amd import capi, synth
from conftest import ROOT, TERNARY

HEADER = os.path.join(ROOT, "include", "lmc_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lmc_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(capi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_engine_philox_known_answers():
    with open(os.path.join(ROOT, "tests", "golden", "philox4x32_10_kat.json")) as f:
        kat = json.load(f)["vectors"]
    for v in kat:
        c = [int(x, 16) for x in v["ctr"]]
        k = [int(x, 16) for x in v["key"]]
        out = (C.c_uint32 * 4)()
        capi.lib().lmc_philox4x32(k[0] | (k[1] << 32), c[0] | (c[1] << 32), c[2], c[3], out)
        assert list(out) == [int(x, 16) for x in v["out"]]


def test_model_load_errors(tmp_path):
    with pytest.raises(capi.LmcError, match="Cannot open"):
        L.Model(str(tmp_path / "missing.json"), TERNARY, L.E0)
    bad = tmp_path / "nobase.json"
    bad.write_text(json.dumps({"Al": {}}))
    with pytest.raises(capi.LmcError, match="Base"):
        L.Model(str(bad), TERNARY, L.ENERGY)
    short = tmp_path / "short.json"
    short.write_text(json.dumps({"Base": {"theta": [0.0] * 10}}))
    with pytest.raises(capi.LmcError, match="theta"):
        L.Model(str(short), TERNARY, L.ENERGY)
    e0 = tmp_path / "e0.json"
    synth.write_coefficients(str(e0), TERNARY, "E0", seed=1)
    with pytest.raises(capi.LmcError, match="mu_x_mm2"):  # Quartic needs the MM2 keys
        L.Model(str(e0), TERNARY, L.QUARTIC)
    broken = tmp_path / "broken.json"
    broken.write_text('{"Base": {"theta": [1, 2,')
    with pytest.raises(capi.LmcError):
        L.Model(str(broken), TERNARY, L.ENERGY)
    m = L.Model(str(e0), TERNARY, L.E0)
    assert m.species == ["Al", "Mg", "X", "Zn"]
    assert [m.code(e) for e in m.species] == [0, 1, 2, 3]
    assert m.code("Cu") == -1
    assert m.layout() == (95, 711, 1401)


def test_unknown_elements_become_vacancy(tmp_path):
    """Element strings the reference does not know map to X (Element.hpp:25-26)."""
    p = tmp_path / "z.json"
    p.write_text(json.dumps({"Base": {"theta": [0.0] * synth.eci_count(["Al", "Mg"])}}))
    m = L.Model(str(p), ["Al", "Mg", "Unobtainium"], L.ENERGY)
    assert m.species == ["Al", "Mg", "X"]


def test_geometry_entry_points():
    assert L.fcc_num_sites(10) == 4000
    assert capi.lib().lmc_fcc_num_sites(3) < 0
    nb = L.fcc_neighbors(6, 0, 1)
    assert list(nb) == [6, 11, 66, 71, 72, 77, 78, 138, 792, 797, 798, 858]
    out = np.zeros(24, np.int64)
    rc = capi.lib().lmc_fcc_neighbors(6, 10 ** 6, 1, capi.ptr(out, capi._i64p))
    assert rc == capi.LMC_ERR_ARG if hasattr(capi, "LMC_ERR_ARG") else rc < 0
    with pytest.raises(capi.LmcError, match="first neighbours"):
        L.jump_state_clusters(6, 0, 1)


def test_orbit_layout_is_contiguous():
    lay = synth.layout(TERNARY)
    for orb, F in ((lay["mmm"], 711), (lay["mm2"], 1401)):
        assert orb["base"][0] == 0
        assert np.all(orb["base"][1:] == orb["base"][:-1] + orb["len"][:-1])
        assert orb["base"][-1] + orb["len"][-1] == F
        assert orb["size"].sum() == 614  # 58 singlets + 556 pairs over the 58 non-endpoint sites
