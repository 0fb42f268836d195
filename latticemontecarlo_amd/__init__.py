"""MI355X-native engine for the LatticeMC event-evaluation hot path.

The compute path is liblmc.so (hand-written gfx950 HIP kernels behind the C ABI of
include/lmc_capi.h); this package is the Python mirror of the reference's predictor and driver
interface over that ABI.
"""
from . import cfg  # noqa: F401,E402
from .api import (E0, ENERGY, PAIRSITE, QUARTIC, CanonicalMc, KineticMcChain, KineticMcFirst, Model,  # noqa: F401,E501
                  ReplicaBatch, SimulatedAnnealing, Surrogate, augment, eval_jumps, fcc_neighbors, fcc_num_sites,
                  jump_environment, jump_orbit_clusters, jump_samples, jump_state_clusters, jump_symmetries,
                  site_de, swap_de, total_energy, total_encode, cluster_energy, chemical_potential,
                  vacancy_scan)
from . import ansys  # noqa: F401,E402
