import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

TERNARY = ["Al", "Mg", "Zn"]
QUATERNARY = ["Al", "Cu", "Mg", "Zn"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def coeffs(tmp_path_factory):
    """Synthetic coefficient files in the reference JSON schema, keyed by name."""
    from latticemontecarlo_amd import synth
    d = tmp_path_factory.mktemp("coeffs")
    specs = {
        "e0_3": (TERNARY, "E0", dict(seed=1)),
        "e0_3n": (TERNARY, "E0", dict(seed=2, randomize_norm=True, n_rows=2)),
        "e0_4n": (QUATERNARY, "E0", dict(seed=3, randomize_norm=True, n_rows=2)),
        "q_3n": (TERNARY, "Quartic", dict(seed=4, randomize_norm=True)),
        "q_4": (QUATERNARY, "Quartic", dict(seed=5)),
    }
    out = {}
    for name, (els, kind, kw) in specs.items():
        p = str(d / f"{name}.json")
        synth.write_coefficients(p, els, kind, **kw)
        out[name] = (p, els)
    return out


def random_alloc(n, species, solute_frac, seed, vacancy=True):
    """Random occupancy codes: each non-solvent element gets solute_frac of the sites."""
    from latticemontecarlo_amd import synth
    N = 4 * n ** 3
    solutes = [e for e in species if e not in ("Al", "X")]
    counts = {e: int(solute_frac * N) for e in solutes}
    return synth.random_occupancy(n, species, counts, seed, vacancy=vacancy)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
