"""PDEP database: the on-disk result of Wstat, in the reference's format.

Wstat ends by saving the converged eigenpairs of the dielectric matrix
(``Wstat/wstat.f90`` -> ``pdep_db_write``), and Wfreq starts by reading them
back (``pdep_db_read``).  The format:

* ``<save_dir>/summary.json`` (``IO_kernel/pdep_db.f90:124-180``)::

    {"dielectric_matrix": {
        "domain": {"a1": [..], "a2": [..], "a3": [..],      # lattice vectors, bohr
                   "b1": [..], "b2": [..], "b3": [..]},     # reciprocal vectors, bohr^-1
        "pdep": [{"iq": 1, "q": [0, 0, 0],
                  "eigenval": [ev_1, ..],
                  "eigenvec": ["Q000000001E000000001.dat", ..]}]}}

  One ``pdep`` record per q point; writing a q point that is already listed
  replaces its record (``pdep_db.f90:160-171``).
* one file per eigenvector, named ``Q<iq:09d>E<j:09d>.dat``
  (``pdep_db.f90:23-48``): a header of 32 default (4-byte) integers --
  (1) the format version 210405, (2) 1 for little endian, 0 for big endian,
  (3) the number of complex entries -- then the coefficients as complex128
  (``IO_kernel/mod_west_io.f90:22-35``, ``IO_kernel/pdep_io.f90:115-141``).
  At Gamma the entries are the npwq coefficients of the half sphere, in
  the order of the global G list (on one process: the order of ``dvg``).

Reading checks version, dimension and endianness and fails with the
reference's messages (``pdep_io.f90:254-280``).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HD_LENGTH = 32
HD_VERSION = 210405
HD_ID_VERSION = 0          # 0-based positions of the header fields
HD_ID_LITTLE_ENDIAN = 1
HD_ID_DIMENSION = 2


def pdep_fname(j: int, iq: int = 1) -> str:
    """``generate_pdep_fname`` (``pdep_db.f90:23-48``); j and iq are 1-based."""
    if not (0 <= j < 10**9 and 0 <= iq < 10**9):
        raise ValueError(f"pdep index out of range: j={j}, iq={iq}")
    return f"Q{iq:09d}E{j:09d}.dat"


def write_vector(fname: str, pdepg: np.ndarray) -> None:
    """One eigenvector file (``pdep_merge_and_write_G``, Gamma branch,
    ``pdep_io.f90:115-141``): header, then ``pdepg`` as complex128."""
    v = np.ascontiguousarray(pdepg, dtype=np.complex128).reshape(-1)
    little = sys.byteorder == "little"
    header = np.zeros(HD_LENGTH, dtype=np.int32)
    header[HD_ID_VERSION] = HD_VERSION
    header[HD_ID_DIMENSION] = v.size
    header[HD_ID_LITTLE_ENDIAN] = 1 if little else 0
    with open(fname, "wb") as f:
        f.write(header.tobytes())
        f.write(v.tobytes())


def read_vector(fname: str, ndim: int) -> np.ndarray:
    """``pdep_read_G_and_distribute`` (Gamma branch, ``pdep_io.f90:250-285``):
    returns the ``ndim`` complex128 entries of one eigenvector file."""
    try:
        f = open(fname, "rb")
    except OSError as e:
        raise OSError(f"pdep_read: Cannot read file: {fname}") from e
    with f:
        header = np.frombuffer(f.read(4 * HD_LENGTH), dtype=np.int32)
        if header.size != HD_LENGTH or header[HD_ID_VERSION] != HD_VERSION:
            raise ValueError(f"pdep_read: Unknown file format: {fname}")
        if header[HD_ID_DIMENSION] != ndim:
            raise ValueError(f"pdep_read: Dimension mismatch: {fname}")
        little = sys.byteorder == "little"
        if (little and header[HD_ID_LITTLE_ENDIAN] == 0) or (not little and header[HD_ID_LITTLE_ENDIAN] == 1):
            raise ValueError(f"pdep_read: Endianness mismatch: {fname}")
        data = np.frombuffer(f.read(16 * ndim), dtype=np.complex128)
    if data.size != ndim:
        raise ValueError(f"pdep_read: Cannot read file: {fname}")
    return data.copy()


def _domain(at: np.ndarray, bg: np.ndarray) -> dict:
    # celldm(1)*at(:,i) and tpiba*bg(:,i) are the lattice and reciprocal
    # vectors in bohr and bohr^-1 (pdep_db.f90:129-134); here the rows of
    # at (bohr) and bg (a_i . b_j = 2 pi delta_ij)
    return {**{f"a{i + 1}": [float(x) for x in at[i]] for i in range(3)},
            **{f"b{i + 1}": [float(x) for x in bg[i]] for i in range(3)}}


def pdep_db_write(save_dir: str, ev, dvg, npwq: int, at: np.ndarray, bg: np.ndarray, iq: int = 1,
                  q=(0.0, 0.0, 0.0)) -> list:
    """``pdep_db_write`` (``pdep_db.f90:55-205``): the summary record of q
    point ``iq`` and one file per eigenvector.

    ``ev`` (n_pdep_eigen,) eigenvalues; ``dvg`` (n_pdep_eigen, npwqx) complex
    eigenvectors (rows; the reference's ``dvg(npwqx, j)``); ``npwq`` the
    number of plane waves written per vector (``npwq_g``).  Returns the
    file names written."""
    ev = np.asarray(ev, dtype=np.float64).reshape(-1)
    dvg = np.asarray(dvg)
    n = ev.size
    if dvg.ndim != 2 or dvg.shape[0] < n or dvg.shape[1] < npwq:
        raise ValueError(f"pdep_db_write: dvg of shape {dvg.shape} does not hold {n} vectors of {npwq} entries")
    os.makedirs(save_dir, exist_ok=True)
    names = [pdep_fname(j + 1, iq) for j in range(n)]
    summary = os.path.join(save_dir, "summary.json")
    if os.path.exists(summary):
        with open(summary) as f:
            doc = json.load(f)
    else:
        doc = {"dielectric_matrix": {"domain": _domain(np.asarray(at), np.asarray(bg)), "pdep": []}}
    recs = doc["dielectric_matrix"].setdefault("pdep", [])
    rec = {"iq": int(iq), "q": [float(x) for x in q], "eigenval": [float(x) for x in ev], "eigenvec": names}
    for k, r in enumerate(recs):
        if r.get("iq") == iq:
            recs[k] = rec
            break
    else:
        recs.append(rec)
    with open(summary, "w") as f:
        json.dump(doc, f, indent=2)
    for j, name in enumerate(names):
        write_vector(os.path.join(save_dir, name), dvg[j, :npwq])
    return names


def pdep_db_read(save_dir: str, npwq: int, npwqx: int | None = None, nglob_to_be_read: int = 0, iq: int = 1):
    """``pdep_db_read`` (``pdep_db.f90:211-369``): eigenvalues and
    eigenvectors of q point ``iq``.  ``nglob_to_be_read`` = 0 reads all.
    Returns ``(ev, dvg)`` with ``dvg`` (n, npwqx) complex, zero beyond npwq."""
    summary = os.path.join(save_dir, "summary.json")
    try:
        with open(summary) as f:
            doc = json.load(f)
    except OSError as e:
        raise OSError(f"pdep_db_read: Cannot open file: {summary}") from e
    rec = next((r for r in doc["dielectric_matrix"].get("pdep", []) if r.get("iq") == iq), None)
    if rec is None:
        raise ValueError(f"pdep_db_read: no record for iq = {iq} in {summary}")
    navail = len(rec["eigenval"])
    n = navail if nglob_to_be_read == 0 else min(navail, nglob_to_be_read)
    ev = np.asarray(rec["eigenval"][:n], dtype=np.float64)
    dvg = np.zeros((n, npwqx or npwq), dtype=np.complex128)
    for j in range(n):
        dvg[j, :npwq] = read_vector(os.path.join(save_dir, rec["eigenvec"][j]), npwq)
    return ev, dvg
