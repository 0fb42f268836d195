"""Configuration files of the reference (cfg::Config I/O, lmc/cfg/src/Config.cpp:554-922) over
the engine's C ABI (lmc_config_*): .cfg[.gz] read and write, the lattice / element / map triple,
GenerateFCC's atom numbering, atom ids and map_shift. Lattice ids are the engine's x-major ids
(ReassignLatticeVector order); atom ids are the file's line order.

    c = read_config("start.cfg.gz")
    occ = c.codes(model.species)                 # species code by lattice id for a model
    lat2atom, shift = c.atoms()
    ...                                          # run; read back codes, atom ids, map_shift
    c.update(model.species, occ, lat2atom, shift)
    c.write("end.cfg.gz")
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import check, lib, ptr


def _names(species):
    arr = (C.c_char_p * len(species))(*[s.encode() for s in species])
    return arr


class Config:
    """A configuration of the n x n x n-cell FCC box (owns an lmc_config handle)."""

    def __init__(self, handle):
        self.h = handle
        n, N = C.c_int32(), C.c_int64()
        basis = np.zeros(9)
        check(lib().lmc_config_info(self.h, C.byref(n), C.byref(N), ptr(basis, capi._f64p)))
        self.n = n.value
        self.num_sites = N.value
        self.basis = basis.reshape(3, 3)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().lmc_config_free(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def codes(self, species):
        """uint8 species code of every lattice id for the species list (a model's .species)."""
        out = np.zeros(self.num_sites, np.uint8)
        check(lib().lmc_config_codes(self.h, len(species), _names(species), ptr(out, capi._u8p)))
        return out

    def atoms(self):
        """(lattice_to_atom int64 [N], map_shift int32 [N, 3] by atom id)."""
        l2a = np.zeros(self.num_sites, np.int64)
        shift = np.zeros((self.num_sites, 3), np.int32)
        check(lib().lmc_config_atoms(self.h, ptr(l2a, capi._i64p), ptr(shift, capi._i32p)))
        return l2a, shift

    def elements(self, species):
        """Element string of every lattice id (species: a list holding every element present)."""
        sp = list(species)
        return [sp[c] for c in self.codes(sp)]

    def update(self, species, codes, lattice_to_atom=None, map_shift=None):
        """The state after a run: species codes by lattice id, and optionally the atom on every
        site and map_shift by atom id (None keeps the current ones)."""
        codes = np.ascontiguousarray(codes, np.uint8)
        if codes.size != self.num_sites:
            raise ValueError(f"codes has {codes.size} entries, expected {self.num_sites}")
        l2a = None if lattice_to_atom is None else np.ascontiguousarray(lattice_to_atom, np.int64)
        sh = None if map_shift is None else np.ascontiguousarray(map_shift, np.int32)
        if l2a is not None and l2a.size != self.num_sites:
            raise ValueError("lattice_to_atom has the wrong size")
        if sh is not None and sh.size != 3 * self.num_sites:
            raise ValueError("map_shift has the wrong size")
        check(lib().lmc_config_update(self.h, len(species), _names(species),
                                      ptr(codes, capi._u8p),
                                      None if l2a is None else ptr(l2a, capi._i64p),
                                      None if sh is None else ptr(sh, capi._i32p)))

    def write(self, filename):
        """Config::WriteConfig: .cfg, or gzip when the name ends in .gz."""
        check(lib().lmc_config_write(self.h, str(filename).encode()))

    def write_xyz(self, filename):
        """Config::WriteExtendedXyz without auxiliary lists (.gz by extension)."""
        check(lib().lmc_config_write_xyz(self.h, str(filename).encode()))

    def write_map(self, lattice_filename=None, element_filename=None, map_filename=None):
        """Config::WriteLattice / WriteElement / WriteMap (None skips a file)."""
        enc = lambda f: None if f is None else str(f).encode()  # noqa: E731
        check(lib().lmc_config_write_map(self.h, enc(lattice_filename), enc(element_filename),
                                         enc(map_filename)))


def generate_fcc(n, element):
    """cfg::GenerateFCC({n, n, n}, element) with x-major lattice ids."""
    h = C.c_void_p()
    check(lib().lmc_config_generate(int(n), element.encode(), C.byref(h)))
    return Config(h)


def read_config(filename):
    """Config::ReadConfig + ReassignLatticeVector (.gz by extension)."""
    h = C.c_void_p()
    check(lib().lmc_config_read(str(filename).encode(), C.byref(h)))
    return Config(h)


def read_map(lattice_filename, element_filename, map_filename):
    """Config::ReadMap."""
    h = C.c_void_p()
    check(lib().lmc_config_read_map(str(lattice_filename).encode(), str(element_filename).encode(),
                                    str(map_filename).encode(), C.byref(h)))
    return Config(h)
