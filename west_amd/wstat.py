"""Host-side mirror of the reference's Wstat call surfaces, running on the
MI355X library.

Names, argument meaning and error behaviour follow the reference:

* :func:`apply_sternheimerop_to_m_wfcs` -- ``DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:14``
* :func:`linsolve_sternheimer_m_wfcts`  -- ``DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:14``
  (``b`` may be destroyed, ``x`` is refined in place, returns ``ierr``)
* :func:`dfpt`                          -- ``DFPT_kernel/dfpt_module.f90:22``
* :func:`apply_operator`                -- ``Wstat/apply_operator.f90:14``
* :func:`davidson_diago`                -- ``Wstat/davidson_diago.f90:31``
* :func:`davidson_diago_k`              -- ``Wstat/davidson_diago.f90:441`` (one q point),
  after :func:`set_k_ground_state`; :func:`dfpt` / :func:`apply_operator` then take ``iq``
* :func:`set_davidson_restart`          -- ``Wstat/davidson_restart.f90`` (inputs of ``Modules/westcom.f90``)

Arrays are numpy ``(m, npwx)`` complex128 (the memory layout of the
reference's Fortran ``(npwx, m)``) or torch tensors of the same shape on the
device.  Non-convergence is reported like the reference: a warning count
(``ierr``), not an exception.
"""
from __future__ import annotations

import ctypes as C
import threading
import warnings

import numpy as np

from .lib import DavidsonParams, DavidsonStep, check, load, ptr

_state = {"device": None}
# per system: the registered ground state and the grid registered by gpu_h_factory;
# the system a thread works on is the one it selected last (0 by default)
_gs_by_sys: dict = {}
_grid_key_by_sys: dict = {}
_tls = threading.local()


def _cur() -> int:
    return getattr(_tls, "handle", 0)


def init(device: int = 0):
    lib = load()
    check(lib.wh_init(int(device)))
    _state["device"] = device
    return lib


def system_create() -> int:
    """A new, empty system (own stream and workspaces); returns its handle."""
    h = C.c_int(-1)
    check(load().wh_system_create(C.addressof(h)))
    return int(h.value)


def system_select(handle: int):
    """Target the system ``handle`` with every later call from this thread."""
    check(load().wh_system_select(int(handle)))
    _tls.handle = int(handle)


def current_ground_state():
    """The ground state registered on this thread's system."""
    return _gs_by_sys[_cur()]


def system_destroy(handle: int):
    check(load().wh_system_destroy(int(handle)))


def set_ground_state(gs, sqvc=None, n_dfpt_maxiter: int = 250, l_kinetic_only: bool = False):
    """Register a :class:`west_amd.groundstate.GroundState` (the module globals
    that the reference's dfpt reads: dffts, g2kin, vrs, vkb, evc, et ...)."""
    lib = load()
    b = gs.basis
    nl = np.ascontiguousarray(b.nl, dtype=np.int32)
    nlm = np.ascontiguousarray(b.nlm, dtype=np.int32)
    g2 = np.zeros(gs.npwx)
    g2[:b.npw] = b.g2kin
    check(lib.wh_set_grid(b.nr[0], b.nr[1], b.nr[2], b.npw, gs.npwx, b.gstart, ptr(nl), ptr(nlm),
                          float(b.omega), float(b.tpiba2), ptr(g2)))
    check(lib.wh_set_params(int(n_dfpt_maxiter), int(bool(l_kinetic_only))))
    check(lib.wh_set_potential(ptr(np.ascontiguousarray(gs.vrs, dtype=np.float64))))
    vkb = np.ascontiguousarray(gs.vkb, dtype=np.complex128)
    check(lib.wh_set_nonlocal(int(vkb.shape[0]), ptr(vkb) if vkb.size else None,
                              ptr(np.ascontiguousarray(gs.dvan, dtype=np.float64)) if vkb.size else None))
    nbv = gs.nbndval
    evc = np.ascontiguousarray(gs.evc[:nbv], dtype=np.complex128)
    et = np.ascontiguousarray(np.asarray(gs.et)[:nbv], dtype=np.float64)
    occ = np.ascontiguousarray(np.asarray(gs.occ)[:nbv], dtype=np.float64)
    check(lib.wh_set_states(nbv, ptr(evc), ptr(et), ptr(occ), float(gs.weight), float(gs.alpha_pv)))
    if sqvc is not None:
        sq = np.zeros(gs.npwx)
        sq[:len(sqvc)] = sqvc
        check(lib.wh_set_coulomb(ptr(sq)))
    _gs_by_sys[_cur()] = gs
    _grid_key_by_sys[_cur()] = None


def _out_like(a, shape):
    if isinstance(a, np.ndarray):
        return np.zeros(shape, dtype=np.complex128)
    import torch
    return torch.zeros(shape, dtype=torch.complex128, device=a.device)


def apply_sternheimerop_to_m_wfcs(nbndval: int, psi, e, alpha, m: int | None = None, hpsi=None):
    """hpsi = H psi - e psi + alpha P_v psi."""
    lib = load()
    m = psi.shape[0] if m is None else m
    if hpsi is None:
        hpsi = _out_like(psi, psi.shape)
    e = np.ascontiguousarray(np.asarray(e, dtype=np.float64)) if not hasattr(e, "data_ptr") else e
    al = np.array([complex(alpha).real, complex(alpha).imag])
    check(lib.wh_apply_sternheimerop_to_m_wfcs(int(nbndval), ptr(psi), ptr(hpsi), ptr(e), ptr(al), int(m)))
    return hpsi


def linsolve_sternheimer_m_wfcts(nbndval: int, m: int, b, x, e, eprec, tr2: float) -> int:
    """Solves (H - e_i + alpha_pv P_v) x_i = b_i in place of x; returns ierr."""
    lib = load()
    e = np.ascontiguousarray(np.asarray(e, dtype=np.float64))
    eprec = np.ascontiguousarray(np.asarray(eprec, dtype=np.float64))
    ierr = C.c_int(0)
    rc = check(lib.wh_linsolve_sternheimer_m_wfcts(int(nbndval), int(m), ptr(b), ptr(x), ptr(e), ptr(eprec),
                                                    float(tr2), C.addressof(ierr)), allow_positive=True)
    return int(ierr.value) if rc >= 0 else rc


def dfpt(m: int, dvg, tr2: float, iq: int = 1, dng=None, return_ierr: bool = False):
    """dng = chi0 dvg (m perturbations)."""
    lib = load()
    if dng is None:
        dng = _out_like(dvg, dvg.shape)
    ierr = np.zeros(max(m, 1), dtype=np.int32)
    rc = check(lib.wh_dfpt(int(m), ptr(dvg), ptr(dng), float(tr2), int(iq), ptr(ierr)), allow_positive=True)
    for p in np.flatnonzero(ierr[:m]):
        warnings.warn(f"** WARNING : PERT {p + 1} not converged, ierr = {ierr[p]}")
    return (dng, ierr[:m]) if return_ierr else dng


def apply_operator(m: int, dvg, tr2: float, iq: int = 1, dng=None):
    """dng = sqvc chi0 (sqvc dvg)."""
    lib = load()
    if dng is None:
        dng = _out_like(dvg, dvg.shape)
    rc = check(lib.wh_apply_operator(int(m), ptr(dvg), ptr(dng), float(tr2), int(iq)), allow_positive=True)
    if rc > 0:
        warnings.warn(f"** WARNING : {rc} Sternheimer systems not converged")
    return dng


def davidson_diago(n_pdep_eigen: int, n_pdep_times: int = 4, n_pdep_maxiter: int = 100, trev_pdep: float = 1e-3,
                   trev_pdep_rel: float = 0.1, tr2_dfpt: float = 1e-16, want_vectors: bool = True):
    """Wstat Davidson; returns (ev, conv, dvg or None, davitr records)."""
    lib = load()
    gs = _gs_by_sys[_cur()]
    prm = DavidsonParams(n_pdep_eigen, n_pdep_times, n_pdep_maxiter, trev_pdep, trev_pdep_rel, tr2_dfpt)
    ev = np.zeros(n_pdep_eigen)
    conv = np.zeros(n_pdep_eigen, dtype=np.int32)
    dvg = np.zeros((n_pdep_eigen, gs.npwx), dtype=np.complex128) if want_vectors else None
    cap = n_pdep_maxiter + 2
    log = (DavidsonStep * cap)()
    ev_log = np.zeros((cap, n_pdep_eigen))
    conv_log = np.zeros((cap, n_pdep_eigen), dtype=np.int32)
    nlog = C.c_int(0)
    rc = check(lib.wh_davidson_diago(C.addressof(prm), ptr(ev), ptr(conv), ptr(dvg) if dvg is not None else None,
                                     C.addressof(log), cap, C.addressof(nlog), ptr(ev_log), ptr(conv_log)),
               allow_positive=True)
    recs = [{"dav_iter": s.dav_iter, "ev": ev_log[i].tolist(), "conv": [bool(v) for v in conv_log[i]],
             "notcnv": s.notcnv, "time_elap:sec": s.time_elap, "time_iter:sec": s.time_iter,
             "sternop_ncalls": s.sternop_ncalls, "dfpt_tr2": s.dfpt_tr2, "dfpt_dim": s.dfpt_dim,
             "diago_dim": s.diago_dim} for i, s in enumerate(log[:nlog.value])]
    if rc > 0:
        warnings.warn(f"WARNING : {rc} eigenvalues not converged in chidiago")
    return ev, conv.astype(bool), dvg, recs


# ---------------------------------------------------------------------------
# q points (Wstat away from Gamma)
# ---------------------------------------------------------------------------
_kgs_by_sys: dict = {}


def set_k_ground_state(kgs, n_dfpt_maxiter: int = 250, l_kinetic_only: bool = False):
    """Register a :class:`west_amd.kpoint.KGroundState` (the inputs of the
    non-Gamma branches of dfpt: per k the sphere, g2kin, evc, et and vkb; per q
    the |q+G| sphere, the Coulomb factor and the k - q map).  The system's
    operator, CG and chi0 entry points then act at k and q points
    (:func:`set_current_k`; ``iq`` of :func:`dfpt` / :func:`apply_operator`)."""
    from .coulomb import sqvc_init
    lib = load()
    nr = kgs.nr
    nks, nqs = kgs.nks, len(kgs.qpts)
    wk = np.ascontiguousarray([k.weight for k in kgs.kpts], dtype=np.float64)
    dvan = np.ascontiguousarray(kgs.dvan, dtype=np.float64)
    check(lib.wh_set_kgrid(nr[0], nr[1], nr[2], float(kgs.omega), ptr(np.ascontiguousarray(kgs.vrs, dtype=np.float64)),
                           nks, ptr(wk), int(kgs.npwx), nqs, int(kgs.npwqx), int(kgs.nbndval), int(kgs.nkb),
                           ptr(dvan) if kgs.nkb else None, float(kgs.alpha_pv)))
    check(lib.wh_set_params(int(n_dfpt_maxiter), int(bool(l_kinetic_only))))
    nbv = kgs.nbndval
    for ik, kp in enumerate(kgs.kpts):
        b = kp.basis
        evc = np.ascontiguousarray(kp.evc[:nbv], dtype=np.complex128)
        et = np.ascontiguousarray(kp.et[:nbv], dtype=np.float64)
        vkb = np.ascontiguousarray(kp.vkb, dtype=np.complex128)
        check(lib.wh_set_kpoint(ik + 1, b.npw, ptr(np.ascontiguousarray(b.nl, dtype=np.int32)),
                                ptr(np.ascontiguousarray(b.g2kin, dtype=np.float64)), ptr(evc), ptr(et),
                                ptr(vkb) if kgs.nkb else None))
    for iq, qp in enumerate(kgs.qpts):
        b = qp.basis
        sq = sqvc_init(b)
        check(lib.wh_set_qpoint(iq + 1, b.npw, ptr(np.ascontiguousarray(b.nl, dtype=np.int32)),
                                ptr(np.ascontiguousarray(b.g2kin, dtype=np.float64)),
                                ptr(np.ascontiguousarray(qp.ig_global, dtype=np.int32)), ptr(sq),
                                ptr(np.ascontiguousarray(qp.ikq + 1, dtype=np.int32)),
                                ptr(np.ascontiguousarray(qp.g0, dtype=np.int32))))
    _kgs_by_sys[_cur()] = kgs
    _gs_by_sys.pop(_cur(), None)
    _grid_key_by_sys[_cur()] = None


def set_current_k(ik: int):
    """k point (1-based, QE's ``current_k``) of the operator and CG calls."""
    check(load().wh_set_current_k(int(ik)))


def davidson_diago_k(iq: int, n_pdep_eigen: int, n_pdep_times: int = 4, n_pdep_maxiter: int = 100,
                     trev_pdep: float = 1e-3, trev_pdep_rel: float = 0.1, tr2_dfpt: float = 1e-16,
                     want_vectors: bool = True):
    """Wstat Davidson at q point ``iq`` (1-based; ``davidson_diago_k``); returns
    (ev, conv, dvg or None, davitr records)."""
    lib = load()
    kgs = _kgs_by_sys[_cur()]
    prm = DavidsonParams(n_pdep_eigen, n_pdep_times, n_pdep_maxiter, trev_pdep, trev_pdep_rel, tr2_dfpt)
    ev = np.zeros(n_pdep_eigen)
    conv = np.zeros(n_pdep_eigen, dtype=np.int32)
    dvg = np.zeros((n_pdep_eigen, kgs.npwqx), dtype=np.complex128) if want_vectors else None
    cap = n_pdep_maxiter + 2
    log = (DavidsonStep * cap)()
    ev_log = np.zeros((cap, n_pdep_eigen))
    conv_log = np.zeros((cap, n_pdep_eigen), dtype=np.int32)
    nlog = C.c_int(0)
    rc = check(lib.wh_davidson_diago_k(int(iq), C.addressof(prm), ptr(ev), ptr(conv),
                                       ptr(dvg) if dvg is not None else None, C.addressof(log), cap,
                                       C.addressof(nlog), ptr(ev_log), ptr(conv_log)), allow_positive=True)
    recs = [{"dav_iter": s.dav_iter, "ev": ev_log[i].tolist(), "conv": [bool(v) for v in conv_log[i]],
             "notcnv": s.notcnv, "time_elap:sec": s.time_elap, "time_iter:sec": s.time_iter,
             "sternop_ncalls": s.sternop_ncalls, "dfpt_tr2": s.dfpt_tr2, "dfpt_dim": s.dfpt_dim,
             "diago_dim": s.diago_dim} for i, s in enumerate(log[:nlog.value])]
    if rc > 0:
        warnings.warn(f"WARNING : {rc} eigenvalues not converged in chidiago")
    return ev, conv.astype(bool), dvg, recs


def _human_time(sec: float) -> str:
    """Elapsed time in the reference's format (``Tools/human_readable_time.f90``):
    "< 00.1s", "SS.ds", "MMm-SS.ds", "HHh-MMm-SS.ds" or "Dd-HHh-MMm-SS.ds"
    (tenths truncated)."""
    t = float(sec)
    if t < 0.1:
        return "< 00.1s"
    d = int(t / 86400.0) if t >= 86400.0 else 0
    h = int((t - d * 86400.0) / 3600.0) if t >= 3600.0 else 0
    m = int((t - d * 86400.0 - h * 3600.0) / 60.0) if t >= 60.0 else 0
    s = t - d * 86400.0 - h * 3600.0 - m * 60.0
    out = f"{int(s):02d}.{int(s * 10.0 - 10.0 * int(s)):01d}s"
    if t >= 60.0:
        out = f"{m:02d}m-" + out
    if t >= 3600.0:
        out = f"{h:02d}h-" + out
    if t >= 86400.0:
        out = f"{d}d-" + out
    return out


def write_davitr(logfile, records, append: bool = True):
    """Appends the Davidson step records of :func:`davidson_diago` to the
    ``exec.davitr`` array of a JSON log in the layout ``output_ev_and_time``
    writes (``Wstat/davidson_diago.f90:1296-1332``): ``exec.ndav`` counts
    the steps, each entry has dav_iter, ev, conv, notcnv, time_elap:sec/hum,
    time_iter:sec/hum, sternop_ncalls, dfpt_tr2, dfpt_dim and diago_dim.  The
    reference's test harness reads ``exec.davitr[-1].ev``
    (``test-suite/main_test.py:36-51``)."""
    import json
    from pathlib import Path
    path = Path(logfile)
    doc = json.loads(path.read_text()) if append and path.exists() else {}
    ex = doc.setdefault("exec", {})
    itr = ex.setdefault("davitr", [])
    for r in records:
        itr.append({"dav_iter": int(r["dav_iter"]), "ev": [float(v) for v in r["ev"]],
                    "conv": [bool(v) for v in r["conv"]], "notcnv": int(r["notcnv"]),
                    "time_elap:sec": float(r["time_elap:sec"]), "time_elap:hum": _human_time(r["time_elap:sec"]),
                    "time_iter:sec": float(r["time_iter:sec"]), "time_iter:hum": _human_time(r["time_iter:sec"]),
                    "sternop_ncalls": int(r["sternop_ncalls"]), "dfpt_tr2": float(r["dfpt_tr2"]),
                    "dfpt_dim": int(r["dfpt_dim"]), "diago_dim": int(r["diago_dim"])})
    ex["ndav"] = len(itr)
    path.write_text(json.dumps(doc, indent=2))
    return doc


def set_davidson_restart(wstat_restart_dir: str | None, n_steps_write_restart: int = 0,
                         restart_from: bool = False):
    """Restart files of :func:`davidson_diago` in the reference's format
    (``Wstat/davidson_restart.f90``): written to ``wstat_restart_dir`` every
    ``n_steps_write_restart`` steps (and after both starting passes when 1),
    removed on convergence; ``restart_from`` starts the next Davidson of the
    selected system from them (``wstat_calculation = 'R'``; that run only)."""
    lib = load()
    d = None if wstat_restart_dir is None else str(wstat_restart_dir).encode()
    check(lib.wh_set_davidson_restart(d, int(n_steps_write_restart), 1 if restart_from else 0))


def davidson_batch(jobs, concurrent: bool = True, max_workers: int | None = None):
    """Wstat Davidson on several registered systems.

    ``jobs``: list of (system handle, keyword arguments of :func:`davidson_diago`).
    With ``concurrent`` each system runs in its own host thread on its own
    stream, so the GPU works on all of them at once; otherwise one after the
    other.  Returns the list of :func:`davidson_diago` results, in order."""
    def run(job):
        handle, kw = job
        system_select(handle)
        return davidson_diago(**kw)

    if not concurrent:
        return [run(j) for j in jobs]
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max_workers or len(jobs)) as ex:
        return list(ex.map(run, jobs))


def gpu_h_factory(gs):
    """``apply_h_factory`` for :func:`west_amd.groundstate.make_ground_state`:
    registers the grid (once), the potential and the projectors of the partial
    ground state ``gs`` and returns ``psi -> H psi`` computed by the library
    (the Sternheimer operator with no registered states, e = 0)."""
    lib = load()
    b = gs.basis
    if _grid_key_by_sys.get(_cur()) != id(b):
        nl = np.ascontiguousarray(b.nl, dtype=np.int32)
        nlm = np.ascontiguousarray(b.nlm, dtype=np.int32)
        g2 = np.zeros(gs.npwx)
        g2[:b.npw] = b.g2kin
        check(lib.wh_set_grid(b.nr[0], b.nr[1], b.nr[2], b.npw, gs.npwx, b.gstart, ptr(nl), ptr(nlm),
                              float(b.omega), float(b.tpiba2), ptr(g2)))
        vkb = np.ascontiguousarray(gs.vkb, dtype=np.complex128)
        check(lib.wh_set_nonlocal(int(vkb.shape[0]), ptr(vkb) if vkb.size else None,
                                  ptr(np.ascontiguousarray(gs.dvan, dtype=np.float64)) if vkb.size else None))
        check(lib.wh_set_params(250, 0))
        _grid_key_by_sys[_cur()] = id(b)
    check(lib.wh_set_potential(ptr(np.ascontiguousarray(gs.vrs, dtype=np.float64))))

    def h_psi(psi):
        if not hasattr(psi, "data_ptr"):
            psi = np.ascontiguousarray(psi, dtype=np.complex128)
        return apply_sternheimerop_to_m_wfcs(0, psi, np.zeros(psi.shape[0]), 0.0)

    return h_psi


def set_precision(mode: str = "fp64", refine_delta: float = 0.0):
    """Arithmetic of the Sternheimer operator and solve: ``"fp64"`` (the
    reference's), ``"mixed"`` (FP32 operator with 3-product split-BF16
    projector GEMMs, CG in FP32 on an increment of the FP64 solution, FP64
    residual recomputed when rho has dropped by ``refine_delta`` (0: library
    default, 1e-6 for "mixed" and 1e-8 for "mixed6"); convergence is tested
    on the FP64 residual) or
    ``"mixed6"`` (the same with 6-product split-BF16 GEMMs, about FP32
    accurate)."""
    modes = {"fp64": 0, "mixed": 1, "mixed6": 2}
    if mode not in modes:
        raise ValueError(f"precision mode {mode!r} is not one of {sorted(modes)}")
    check(load().wh_set_precision(modes[mode], float(refine_delta)))


def counters():
    """(operator applications in band columns, CG iterations) so far."""
    lib = load()
    a, b = C.c_long(0), C.c_long(0)
    check(lib.wh_get_operator_columns(C.addressof(a)))
    check(lib.wh_get_counters(None, C.addressof(b)))
    return int(a.value), int(b.value)


def stern_calls():
    """Operator calls so far as the reference's 'stern' clock counts them
    (one per perturbation's Sternheimer solve per CG step)."""
    lib = load()
    a = C.c_long(0)
    check(lib.wh_get_counters(C.addressof(a), None))
    return int(a.value)


def precision_counters():
    """(FP32 operator applications, FP64 refinement residuals, bands finished
    by the FP64 CG) so far."""
    lib = load()
    a, b, f = C.c_long(0), C.c_long(0), C.c_long(0)
    check(lib.wh_get_precision_counters(C.addressof(a), C.addressof(b), C.addressof(f)))
    return int(a.value), int(b.value), int(f.value)


def profile(enable: bool):
    check(load().wh_profile(int(bool(enable))))


def profile_query(name: str):
    ms, cnt = C.c_double(0), C.c_long(0)
    check(load().wh_profile_query(name.encode(), C.addressof(ms), C.addressof(cnt)))
    return float(ms.value), int(cnt.value)
