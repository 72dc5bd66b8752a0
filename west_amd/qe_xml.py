"""Ground-state description from a Quantum ESPRESSO run (``data-file-schema.xml``).

Wstat starts from a pw.x run: ``Tools/init_pw_arrays.f90`` reads the run's
``data-file-schema.xml`` (cell, atoms, cutoffs, FFT grid, bands) and its
wave functions.  This module reads the XML's ``<output>`` section -- the
same schema the reference's test-suite keeps as ``test-suite/*/ref/pw.xml``
-- and builds from it the Gamma basis and the :class:`~west_amd.systems.System`
the library runs on:

* the basis (FFT grid, wave-function and density spheres) from the run's
  cell and cutoffs, checked against the run's own ``fft_grid``, ``ngm`` and
  ``npwx``/``npw``;
* the atoms with this package's pseudopotentials for their species (the
  run's UPF files are not read); the valence charge must equal the run's
  ``nelec``.

The wave functions are computed by :mod:`west_amd.groundstate` on that
system, not read from the run's ``wfc*.dat`` files.  QE's XML is in Hartree
atomic units; everything returned here is in Rydberg (energies) and bohr.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from .gsphere import count_sphere, make_gamma_basis
from .systems import H_HGH, SI_HGH, System

HA_TO_RY = 2.0


@dataclass
class KPoint:
    weight: float
    xk: np.ndarray            # 2 pi / alat units
    npw: int
    et: np.ndarray            # Ry
    occ: np.ndarray


@dataclass
class PwOutput:
    alat: float
    at: np.ndarray            # (3, 3) lattice vectors, rows, bohr
    species: list             # names in ATOMIC_SPECIES order
    pseudo_files: list
    atoms: list               # species name per atom
    tau: np.ndarray           # (nat, 3) bohr
    gamma_only: bool
    ecutwfc: float            # Ry
    ecutrho: float            # Ry
    fft_grid: tuple
    ngm: int
    npwx: int
    lsda: bool
    nbnd: int                 # bands (per spin channel when lsda)
    nelec: float
    kpoints: list = field(default_factory=list)

    @property
    def nat(self) -> int:
        return len(self.atoms)


def _floats(text: str) -> np.ndarray:
    return np.array([float(x) for x in text.split()])


def read_pw_xml(path: str) -> PwOutput:
    """The ``<output>`` section of a pw.x ``data-file-schema.xml``."""
    root = ET.parse(path).getroot()
    out = root.find("output")
    if out is None:
        raise ValueError(f"{path}: no <output> section (not a finished pw.x run)")
    st = out.find("atomic_structure")
    alat = float(st.get("alat"))
    at = np.array([_floats(st.find(f"cell/a{i}").text) for i in (1, 2, 3)])
    atoms, tau = [], []
    for a in st.find("atomic_positions").findall("atom"):
        atoms.append(a.get("name"))
        tau.append(_floats(a.text))
    sp = out.find("atomic_species")
    species = [s.get("name") for s in sp.findall("species")]
    pseudo = [s.findtext("pseudo_file", "").strip() for s in sp.findall("species")]
    bs = out.find("basis_set")
    fg = bs.find("fft_grid")
    band = out.find("band_structure")
    kpts = []
    for ks in band.findall("ks_energies"):
        kp = ks.find("k_point")
        kpts.append(KPoint(float(kp.get("weight")), _floats(kp.text), int(ks.findtext("npw")),
                           HA_TO_RY * _floats(ks.findtext("eigenvalues")), _floats(ks.findtext("occupations"))))
    return PwOutput(alat=alat, at=at, species=species, pseudo_files=pseudo, atoms=atoms, tau=np.array(tau),
                    gamma_only=bs.findtext("gamma_only").strip() == "true",
                    ecutwfc=HA_TO_RY * float(bs.findtext("ecutwfc")), ecutrho=HA_TO_RY * float(bs.findtext("ecutrho")),
                    fft_grid=(int(fg.get("nr1")), int(fg.get("nr2")), int(fg.get("nr3"))),
                    ngm=int(bs.findtext("ngm")), npwx=int(bs.findtext("npwx")),
                    lsda=band.findtext("lsda").strip() == "true",
                    # spin-polarised runs give nbnd_up / nbnd_dw (one k point's
                    # eigenvalues then hold both channels)
                    nbnd=int(band.findtext("nbnd") or band.findtext("nbnd_up")),
                    nelec=float(band.findtext("nelec")), kpoints=kpts)


SPECIES = {"H": H_HGH, "Si": SI_HGH}


def system_from_pw_xml(run: PwOutput, name: str | None = None) -> System:
    """The run's Gamma-point system on this package's pseudopotentials.

    Raises ValueError for runs the path does not cover (k points, spin
    polarisation) and when the basis built here disagrees with the run's
    (FFT grid, number of G vectors, plane waves); KeyError for a species
    without a pseudopotential here."""
    if not run.gamma_only:
        raise ValueError("the Wstat path runs at Gamma only; this run uses k points")
    if run.lsda:
        raise ValueError("the Wstat path runs with nspin = 1; this run is spin polarised")
    basis = make_gamma_basis(run.at, run.ecutwfc, run.ecutrho, alat=run.alat)
    if tuple(basis.nr) != tuple(run.fft_grid):
        raise ValueError(f"FFT grid {basis.nr} differs from the run's {run.fft_grid}")
    # at Gamma QE keeps half spheres: ngm and npwx count G and not -G
    if count_sphere(run.at, run.ecutrho, half=True) != run.ngm or basis.npw != run.npwx:
        raise ValueError("G spheres differ from the run's ngm / npwx")
    kinds = [SPECIES[s] for s in run.species]
    ityp = np.array([run.species.index(a) for a in run.atoms], dtype=int)
    zval = sum(kinds[t].zion for t in ityp)
    if abs(zval - run.nelec) > 1e-8:
        raise ValueError(f"valence charge {zval} of the pseudopotentials here differs from the run's {run.nelec}")
    formula = "".join(f"{s}{n if n > 1 else ''}" for s in run.species if (n := run.atoms.count(s)))
    return System(name or f"{formula}_{run.ecutwfc:g}Ry", basis, kinds, ityp, run.tau.copy(),
                  int(round(run.nelec)), {"source": "pw.x data-file-schema.xml", "pseudo_files": run.pseudo_files})
