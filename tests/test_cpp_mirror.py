"""The C++ mirror (include/lmc.hpp) compiled into a host program with g++ and linked against
liblmc.so: host-side behaviour on the CPU, and its GPU results against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, TERNARY, random_alloc

LIBDIR = os.path.join(ROOT, "latticemontecarlo_amd")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("cpp") / "mirror_check"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "mirror_check.cpp"), "-L", LIBDIR, "-llmc",
           f"-Wl,-rpath,{LIBDIR}", "-o", str(out)]
    subprocess.run(cmd, check=True)
    return str(out)


def test_mirror_host_behaviour(exe, coeffs, tmp_path):
    path, _ = coeffs["e0_3"]
    r = subprocess.run([exe, "cpu", path, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.gpu
def test_mirror_gpu_matches_oracle(exe, coeffs, tmp_path):
    import oracle as O
    path, els = coeffs["e0_3n"]
    n = 6
    species = sorted(TERNARY + ["X"])
    occ = random_alloc(n, species, 0.1, seed=31)
    occ_file = tmp_path / "occ.bin"
    occ.astype(np.uint8).tofile(occ_file)
    print(f"mirror_check gpu n={n} seed=31", flush=True)
    r = subprocess.run([exe, "gpu", str(n), str(occ_file), path, path, str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [json.loads(line) for line in r.stdout.splitlines()]
    lat = O.Lattice(n)
    ob = O.Model(lat, path, els, O.E0)
    oe = O.Model(lat, path, els, O.ENERGY)
    op = O.Model(lat, path, els, O.PAIRSITE)
    jumps = [x for x in rows if "jump" in x]
    assert len(jumps) == 12
    for x in jumps:
        ea, de = ob.barrier_de(occ, *x["jump"])
        assert abs(x["ea"] - ea) <= 1e-9 and abs(x["de"] - de) <= 1e-9
    assert next(x for x in rows if "atom_forms" in x)["atom_forms"] is True
    from latticemontecarlo_amd import ansys as A
    pr = next(x for x in rows if "profile_vac" in x)
    import latticemontecarlo_amd as L
    gm_, gp_ = L.Model(path, els, L.E0), L.Model(path, els, L.PAIRSITE)
    mu = L.chemical_potential(L.Model(path, els, L.ENERGY), "Al")
    et = A.ExitTime(gm_, occ, n, species.index("Al"), 500.0, pairsite_model=gp_)
    prof, bind = et.profile_energy(mu), et.binding_energy(mu)
    v0 = int(np.flatnonzero(occ == species.index("X"))[0])
    assert pr["profile_vac"] == prof[v0] and pr["profile_7"] == prof[7]
    assert pr["binding_Mg_7"] == bind["Mg"][7]
    x_enc = next(x for x in rows if "encode_len" in x)  # EnergyPredictor extras, C++ mirror
    import latticemontecarlo_amd as L
    import torch
    gm = L.Model(path, els, L.ENERGY)
    ref_enc = L.total_encode(gm, L.ReplicaBatch([n], [occ])).cpu().numpy()[0]
    assert x_enc["encode_len"] == ref_enc.size and x_enc["encode0"] == ref_enc[0]
    mu = L.chemical_potential(gm, "Al")
    assert x_enc["mu_Mg"] == mu["Mg"] and x_enc["mu_X"] == mu["X"]
    v0 = int(np.flatnonzero(occ == species.index("X"))[0])
    assert x_enc["cluster_energy"] == L.cluster_energy(gm, torch.from_numpy(occ), n, [v0])[0]
    e = next(x for x in rows if "energy" in x)["energy"]
    assert abs(e - oe.total_energy(occ)) <= 1e-9 * max(1.0, abs(e))
    sw = next(x for x in rows if "swap" in x)
    assert abs(sw["de"] - op.swap_de(occ, *sw["swap"])) <= 1e-9
    si = next(x for x in rows if "site" in x)
    assert abs(si["de"] - op.site_de(occ, si["site"], species.index("Mg"))) <= 1e-9
    k = next(x for x in rows if "kmc_steps" in x)
    ref = O.Kmc(ob, oe, lat, occ, 600.0, 99, replica=0)
    ref.run(201, maximum_steps=200)
    st = ref.state()
    assert k["kmc_steps"] == st["steps"] == 201
    assert k["kmc_vacancy"] == st["vacancy"]
    assert abs(k["kmc_time"] - st["time"]) <= 1e-12 * abs(st["time"])
    # tracked run written through cfg::Config::WriteConfig: atom ids and map_shift against the
    # oracle's LatticeJump replayed over the oracle's own trajectory; the log's line count
    import cfg_oracle as CO
    t = next(x for x in rows if "tracked" in x)
    ref = O.Kmc(ob, oe, lat, occ, 600.0, 99, replica=0)
    otr = ref.run(201, maximum_steps=200, trace=True)
    o = CO.generate_fcc(n, "Al")
    o.reassign()
    for s_, a in enumerate(o.lattice_to_atom):
        o.elements[a] = species[occ[s_]]
    vac = int(np.flatnonzero(occ == species.index("X"))[0])
    for s_ in otr["sel"]:
        o.lattice_jump(vac, int(s_))
        vac = int(s_)
    w = CO.read_file(t["tracked"])
    w.reassign()
    assert w.lattice_to_atom == o.lattice_to_atom and w.map_shift == o.map_shift
    assert w.elements == o.elements
    lines = open(t["log"]).read().split("\n")
    assert lines[0].startswith("steps\ttime") and lines[-1] == ""
    # every step below 99, then every 10 (steps 100, 110, ..., 200)
    assert len(lines) - 2 == 99 + 11
    g = next(x for x in rows if "graph_steps" in x)  # graph replay: bit-identical to eager
    assert (g["graph_steps"], g["graph_vacancy"], g["graph_time"]) == (
        k["kmc_steps"], k["kmc_vacancy"], k["kmc_time"])
    # surrogate with full fallback: the plain run (its exact sums take another lane order, so
    # the clock agrees to rounding)
    u = next(x for x in rows if "sur_steps" in x)
    assert (u["sur_steps"], u["sur_vacancy"]) == (k["kmc_steps"], k["kmc_vacancy"])
    assert abs(u["sur_time"] - k["kmc_time"]) <= 1e-12 * abs(k["kmc_time"])
    assert u["sur_events"] == u["sur_fallbacks"] == 12 * 201 and u["sur_mse_size"] == 8
    vac = int(np.flatnonzero(occ == species.index("X"))[0])
    for x in (x for x in rows if "exit_site" in x):  # ExitTime barrier lists
        o = occ.copy()
        o[vac] = species.index("Al")
        o[x["exit_site"]] = species.index("X")
        ref = sorted(ob.barrier_de(o, x["exit_site"], int(nb))[0]
                     for nb in lat.neighbors(x["exit_site"], 1))
        np.testing.assert_allclose(x["barriers"], ref, rtol=0, atol=1e-9)
        assert abs(x["mean"] - np.mean(ref)) <= 1e-9
        kT = 8.617333262145e-5 * 500.0
        t = 1 / (1e13 * np.exp(-np.array(ref) / kT).sum())
        assert abs(x["exit_time"] - t) <= 1e-8 * t
