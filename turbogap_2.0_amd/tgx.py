"""Python host mirror of the tgx C-ABI (include/tgx.h).

This module is the Python-side caller of the HIP library: it loads
``libtgx.so`` from this directory with ctypes and exposes the same entry
points the Fortran shim binds (``fortran/tgx_mod.f90``).  It does no
arithmetic of its own: when the library is missing it raises, it never falls
back to a CPU path.

The argument meaning mirrors the reference's Fortran interfaces:
``build_neighbors_list`` (neighbors/neighbors_interface.f90:144),
``get_gap_soap`` / ``get_soap_energy_and_forces`` (gap/gap_interface.f90:41,
gap/gap.f90:39), ``get_2b_energy_and_forces`` (gap/gap.f90:329) and
``velocity_verlet`` (md/md_utils.f90:486).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtgx.so")

MAX_SPECIES = 8
MAX_LMAX = 10
MAX_NR = 32

TGX_OK = 0
TGX_ERR_ARG = 1
TGX_ERR_HIP = 2
TGX_ERR_CAPACITY = 3
TGX_ERR_UNSUPPORTED = 4
TGX_ERR_STATE = 5

FP64, FP32, MIXED = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "mixed": MIXED}

# amu -> eV fs^2 / A^2 (reference read/read_files.f90:333)
AMU = 103.6426965268
KB_MD = 8.6173303e-5  # md/md_interface.f90:111-118


class CSoapHypers(C.Structure):
    """ctypes mirror of ``tgx_soap_hypers``."""

    _fields_ = [
        ("n_species", C.c_int32),
        ("l_max", C.c_int32),
        ("radial_enhancement", C.c_int32),
        ("central_species", C.c_int32),
        ("basis", C.c_int32),
        ("pad_", C.c_int32),
        ("alpha_max", C.c_int32 * MAX_SPECIES),
        ("rcut_hard", C.c_double * MAX_SPECIES),
        ("rcut_soft", C.c_double * MAX_SPECIES),
        ("atom_sigma_r", C.c_double * MAX_SPECIES),
        ("atom_sigma_t", C.c_double * MAX_SPECIES),
        ("atom_sigma_r_scaling", C.c_double * MAX_SPECIES),
        ("atom_sigma_t_scaling", C.c_double * MAX_SPECIES),
        ("amplitude_scaling", C.c_double * MAX_SPECIES),
        ("central_weight", C.c_double * MAX_SPECIES),
        ("global_scaling", C.c_double * MAX_SPECIES),
        ("nf", C.c_double * MAX_SPECIES),
        ("compress_P_nonzero", C.c_int32),
        ("pad2_", C.c_int32),
        ("compress_P_i", C.POINTER(C.c_int32)),
        ("compress_P_j", C.POINTER(C.c_int32)),
        ("compress_P_el", C.POINTER(C.c_double)),
    ]


class DaOptions(C.Structure):
    """tgx_da_options (delayed-acceptance Monte Carlo)."""
    _fields_ = [("n_train_trials", C.c_int32), ("train_every", C.c_int32),
                ("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_double),
                ("max_rmse", C.c_double)]


class McStats(C.Structure):
    """tgx_mc_stats."""
    _fields_ = [("n_accepted", C.c_int64), ("n_stage1_rejected", C.c_int64),
                ("n_gap_evaluations", C.c_int64), ("n_rebuilds", C.c_int64),
                ("train_rmse", C.c_double), ("da_rmse", C.c_double), ("fallback", C.c_int32),
                ("pad_", C.c_int32)]


SURROGATE_HIDDEN = 128


def surrogate_nparams(dim: int) -> int:
    H = SURROGATE_HIDDEN
    return H * dim + H * H + 3 * H + 1


@dataclass
class SoapHypers:
    """One soap_turbo block.  Defaults are the reference's
    (gap/read_gap.f90:164-180, types/types.f90:252-312); ``buffer`` is the
    .gap keyword, rcut_soft = rcut_hard - buffer (read_gap.f90:425-432)."""

    n_species: int = 1
    alpha_max: Sequence[int] = (8,)
    l_max: int = 8
    rcut_hard: Sequence[float] = (3.7,)
    buffer: Sequence[float] = (0.5,)
    atom_sigma_r: Sequence[float] = (0.5,)
    atom_sigma_t: Sequence[float] = (0.5,)
    atom_sigma_r_scaling: Sequence[float] = (0.0,)
    atom_sigma_t_scaling: Sequence[float] = (0.0,)
    amplitude_scaling: Sequence[float] = (1.0,)
    central_weight: Sequence[float] = (1.0,)
    global_scaling: Sequence[float] = (1.0,)
    nf: Sequence[float] = (4.0,)
    radial_enhancement: int = 0
    central_species: int = 0
    basis: str = "poly3"  # or "poly3gauss"
    # compress_soap: (P_i, P_j, P_el) 1-based, and the compressed dimension
    compress_P: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None
    compress_dim: int = 0

    def __post_init__(self):
        if self.compress_P is not None:
            pi, pj, pe = self.compress_P
            self.compress_P = (np.ascontiguousarray(pi, np.int32),
                               np.ascontiguousarray(pj, np.int32),
                               np.ascontiguousarray(pe, np.float64))
            if not self.compress_dim:
                self.compress_dim = int(self.compress_P[0].max())

    @classmethod
    def uniform(cls, n_species: int, **kw) -> "SoapHypers":
        """Same per-species values for every species."""
        base = cls()
        args = {"n_species": n_species}
        for name in ("alpha_max", "rcut_hard", "buffer", "atom_sigma_r",
                     "atom_sigma_t", "atom_sigma_r_scaling",
                     "atom_sigma_t_scaling", "amplitude_scaling",
                     "central_weight", "global_scaling", "nf"):
            v = kw.pop(name, getattr(base, name)[0])
            args[name] = tuple([v] * n_species) if np.isscalar(v) else tuple(v)
        args.update(kw)
        return cls(**args)

    @property
    def n_r(self) -> int:
        return int(sum(self.alpha_max))

    @property
    def full_dim(self) -> int:
        """Dimension of the uncompressed power spectrum."""
        nr = self.n_r
        return nr * (nr + 1) // 2 * (self.l_max + 1)

    @property
    def dim(self) -> int:
        """Descriptor dimension (the compressed one with compress_soap)."""
        return self.compress_dim if self.compress_P is not None else self.full_dim

    @property
    def rcut_max(self) -> float:
        return float(max(self.rcut_hard))

    def to_c(self) -> CSoapHypers:
        h = CSoapHypers()
        h.n_species = self.n_species
        h.l_max = self.l_max
        h.radial_enhancement = self.radial_enhancement
        h.central_species = self.central_species
        h.basis = {"poly3": 0, "poly3gauss": 1}[self.basis]
        if self.compress_P is not None:
            pi, pj, pe = self.compress_P
            h.compress_P_nonzero = len(pi)
            h.compress_P_i = pi.ctypes.data_as(C.POINTER(C.c_int32))
            h.compress_P_j = pj.ctypes.data_as(C.POINTER(C.c_int32))
            h.compress_P_el = pe.ctypes.data_as(C.POINTER(C.c_double))
        for s in range(self.n_species):
            h.alpha_max[s] = int(self.alpha_max[s])
            h.rcut_hard[s] = float(self.rcut_hard[s])
            b = float(self.buffer[s])
            h.rcut_soft[s] = h.rcut_hard[s] if b == 0.0 else h.rcut_hard[s] - b
            h.atom_sigma_r[s] = float(self.atom_sigma_r[s])
            h.atom_sigma_t[s] = float(self.atom_sigma_t[s])
            h.atom_sigma_r_scaling[s] = float(self.atom_sigma_r_scaling[s])
            h.atom_sigma_t_scaling[s] = float(self.atom_sigma_t_scaling[s])
            h.amplitude_scaling[s] = float(self.amplitude_scaling[s])
            h.central_weight[s] = float(self.central_weight[s])
            h.global_scaling[s] = float(self.global_scaling[s])
            h.nf[s] = float(self.nf[s])
        return h


@dataclass
class TwoBody:
    """One distance_2b block (gap_2b_t, types/types.f90:314-321)."""

    sp1: int
    sp2: int
    rcut: float
    delta: float
    sigma: float
    Qs: np.ndarray
    alphas: np.ndarray
    cutoff: np.ndarray
    buffer: float = 0.5


@dataclass
class Angle3b:
    """One angle_3b block (gap_3b_t; read_gap.f90:599-643): centre species
    sp0, neighbour species sp1/sp2 (1-based), sparse descriptors Qs (S, 3)
    of q = (r12 + r13, (r12 - r13)^2, r23), kernel "exp" or "pol"."""

    sp0: int
    sp1: int
    sp2: int
    rcut: float
    delta: float
    sigma: Sequence[float]
    Qs: np.ndarray
    alphas: np.ndarray
    cutoff: np.ndarray
    kernel: str = "exp"


@dataclass
class CorePot:
    """One core_pot block (gap_core_pot_t): tabulated V(x), ascending x,
    end slopes yp1/ypn (above 0.99e30: natural end; read_gap.f90:644-752).
    The table is taken as given (after any core_pot_cutoff taper)."""

    sp1: int
    sp2: int
    x: np.ndarray
    V: np.ndarray
    yp1: float = 1e31
    ypn: float = 1e31


@dataclass
class GapModel:
    """A soap_turbo GAP block plus optional 2b blocks and e0 per species."""

    hypers: SoapHypers
    Qs: np.ndarray          # (n_sparse, dim) row-major == Qs(dim, n_sparse)
    alphas: np.ndarray      # (n_sparse,)
    delta: float = 1.0
    zeta: float = 4.0
    e0: Optional[np.ndarray] = None
    two_body: list = field(default_factory=list)
    # further soap_turbo blocks (GapModel objects whose hypers, Qs, alphas,
    # delta and zeta are used), evaluated after this one and summed
    more_soap: list = field(default_factory=list)
    angle_3b: list = field(default_factory=list)
    core_pot: list = field(default_factory=list)

    @property
    def rcut_max(self) -> float:
        r = self.hypers.rcut_max
        for b in self.more_soap:
            r = max(r, b.hypers.rcut_max)
        for b in self.two_body:
            r = max(r, b.rcut)
        for b in self.angle_3b:
            r = max(r, b.rcut)
        for b in self.core_pot:
            r = max(r, float(np.max(b.x)))
        return r


def _dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_int))


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libtgx.so; raise when it is not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP library {path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.c_void_p
    d, i, i64 = C.c_double, C.c_int, C.c_int64
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "tgx_create": (i, [i, i, C.POINTER(P)]),
        "tgx_destroy": (i, [P]),
        "tgx_last_error": (C.c_char_p, [P]),
        "tgx_set_precision": (i, [P, i]),
        "tgx_set_autotune": (i, [P, i, C.c_char_p]),
        "tgx_set_soap_model": (i, [P, C.POINTER(CSoapHypers), i, i, dp, dp, d, d]),
        "tgx_set_e0": (i, [P, i, dp]),
        "tgx_add_soap_model": (i, [P, C.POINTER(CSoapHypers), i, i, dp, dp, d, d]),
        "tgx_clear_soap_models": (i, [P]),
        "tgx_n_soap_models": (i, [P]),
        "tgx_add_2b_model": (i, [P, i, i, d, d, d, d, i, dp, dp, dp]),
        "tgx_clear_2b_models": (i, [P]),
        "tgx_add_angle_3b": (i, [P, i, i, i, d, d, dp, i, i, dp, dp, dp]),
        "tgx_add_core_pot": (i, [P, i, i, i, dp, dp, d, d]),
        "tgx_clear_extra_terms": (i, [P]),
        "tgx_set_mc_batch": (i, [P, i]),
        "tgx_add_local_property": (i, [P, i, i, i, dp, dp, d, d, d, ip]),
        "tgx_clear_local_properties": (i, [P]),
        "tgx_set_vdw_ts": (i, [P, i, i, dp, dp, dp, d, d, d, d, d, d]),
        "tgx_clear_vdw": (i, [P]),
        "tgx_vdw_energy_forces": (i, [P, i, dp, ip, dp, i, dp, dp, dp]),
        "tgx_local_property": (i, [P, i, i, dp, ip, dp, d, dp, i64, dp, C.POINTER(i64)]),
        "tgx_neighbors_build": (i, [P, i, dp, dp, d, i64, ip, ip, dp, dp, dp, dp,
                                    C.POINTER(i64), ip]),
        "tgx_soap_descriptors": (i, [P, i, dp, ip, dp, dp, ip, ip]),
        "tgx_gap_kernel_step": (i, [P, i, dp, dp, dp]),
        "tgx_energy_forces": (i, [P, i, dp, ip, dp, i, dp, dp, dp]),
        "tgx_energy_forces_noalloc": (i, [P, i, dp, ip, dp, i, dp, dp, dp]),
        "tgx_energy_forces_batched": (i, [P, i, ip, dp, ip, dp, i, dp, dp, dp]),
        "tgx_md_run": (i, [P, i, i, d, d, ip, dp, dp, dp, dp, dp, dp, dp, ip]),
        "tgx_stage_times": (i, [P, dp]),
        "tgx_kernel_times": (i, [P, dp]),
        "tgx_set_list_buffer": (i, [P, d]),
        "tgx_set_timing": (i, [P, i]),
        "tgx_mc_run_da": (i, [P, i, ip, dp, dp, i, d, d, d, C.c_uint64,
                              C.POINTER(DaOptions), dp, C.POINTER(McStats)]),
        "tgx_surrogate_init": (i, [P, i, C.c_uint64]),
        "tgx_surrogate_set": (i, [P, i, dp, d, d]),
        "tgx_surrogate_get": (i, [P, i, dp, dp, dp]),
        "tgx_surrogate_predict": (i, [P, i, dp, dp]),
        "tgx_surrogate_gradient": (i, [P, i, dp, dp, dp, dp]),
        "tgx_surrogate_fit": (i, [P, i, dp, dp, i, i, d, C.c_uint64, dp]),
        "tgx_mc_run": (i, [P, i, ip, dp, dp, i, d, d, d, C.c_uint64, dp, ip, ip]),
        "tgx_md_begin": (i, [P, i, ip, dp, dp, dp, dp, d, d, i]),
        "tgx_md_steps": (i, [P, i]),
        "tgx_md_read": (i, [P, dp, dp, dp, dp, dp, ip, ip]),
        "tgx_md_end": (i, [P]),
        "tgx_sync": (i, [P]),
        "tgx_poly3_orthonormalization": (i, [i, dp, dp]),
        "tgx_poly3gauss_orthonormalization": (i, [i, d, d, dp, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def poly3_orthonormalization(alpha_max: int):
    """S and W = S^(-1/2) of the poly3 basis (host-only library call)."""
    lib = load_library()
    S = np.zeros((alpha_max, alpha_max))
    W = np.zeros((alpha_max, alpha_max))
    rc = lib.tgx_poly3_orthonormalization(alpha_max, _dptr(S), _dptr(W))
    if rc != TGX_OK:
        raise TgxError(rc, "tgx_poly3_orthonormalization")
    return S.T.copy(), W.T.copy()


def poly3gauss_orthonormalization(alpha_max: int, atom_sigma: float, rcut_hard: float):
    """S and W = S^(-1/2) of the poly3gauss basis (host-only library call)."""
    lib = load_library()
    S = np.zeros((alpha_max, alpha_max))
    W = np.zeros((alpha_max, alpha_max))
    rc = lib.tgx_poly3gauss_orthonormalization(alpha_max, atom_sigma, rcut_hard, _dptr(S),
                                               _dptr(W))
    if rc != TGX_OK:
        raise TgxError(rc, "tgx_poly3gauss_orthonormalization")
    return S.T.copy(), W.T.copy()


class TgxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tgx error {code}: {msg}")
        self.code = code


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class Engine:
    """One library context on one GPU."""

    def __init__(self, device: int = 0, precision: str = "fp64"):
        self.lib = load_library()
        self.ctx = C.c_void_p()
        self._check(self.lib.tgx_create(device, PRECISIONS[precision], C.byref(self.ctx)))
        self.model: Optional[GapModel] = None

    # -- plumbing --------------------------------------------------------
    def _check(self, rc: int):
        if rc != TGX_OK:
            msg = ""
            if self.ctx:
                m = self.lib.tgx_last_error(self.ctx)
                msg = m.decode() if m else ""
            raise TgxError(rc, msg)

    def close(self):
        if self.ctx:
            self.lib.tgx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: str):
        self._check(self.lib.tgx_set_precision(self.ctx, PRECISIONS[precision]))

    def set_autotune(self, on: bool, cache_path=None):
        """Time the candidate launch shapes on first use of each shape
        (tgx_set_autotune); decisions go to cache_path when given."""
        path = None if cache_path is None else os.fsencode(cache_path)
        self._check(self.lib.tgx_set_autotune(self.ctx, int(on), path))

    # -- models ------------------------------------------------------------
    def set_model(self, model: GapModel):
        h = model.hypers.to_c()
        Qs = _f64(model.Qs)
        al = _f64(model.alphas)
        n_sparse, dim = Qs.shape
        self._check(self.lib.tgx_set_soap_model(self.ctx, C.byref(h), dim, n_sparse,
                                                _dptr(Qs), _dptr(al), model.delta, model.zeta))
        for blk in model.more_soap:
            hb = blk.hypers.to_c()
            Qb, ab = _f64(blk.Qs), _f64(blk.alphas)
            self._check(self.lib.tgx_add_soap_model(self.ctx, C.byref(hb), Qb.shape[1],
                                                    Qb.shape[0], _dptr(Qb), _dptr(ab),
                                                    blk.delta, blk.zeta))
        e0 = _f64(model.e0 if model.e0 is not None else np.zeros(model.hypers.n_species))
        self._check(self.lib.tgx_set_e0(self.ctx, len(e0), _dptr(e0)))
        self._check(self.lib.tgx_clear_2b_models(self.ctx))
        for b in model.two_body:
            q, a, c = _f64(b.Qs), _f64(b.alphas), _f64(b.cutoff)
            self._check(self.lib.tgx_add_2b_model(self.ctx, b.sp1, b.sp2, b.rcut, b.buffer,
                                                  b.delta, b.sigma, len(q), _dptr(q), _dptr(a),
                                                  _dptr(c)))
        self._check(self.lib.tgx_clear_extra_terms(self.ctx))
        for b in model.angle_3b:
            q = _f64(np.asarray(b.Qs, float).T.copy())  # Qs(n_sparse, 3) column-major
            a, c = _f64(b.alphas), _f64(b.cutoff)
            sg = _f64(b.sigma)
            self._check(self.lib.tgx_add_angle_3b(self.ctx, b.sp0, b.sp1, b.sp2, b.rcut, b.delta,
                                                  _dptr(sg), {"exp": 0, "pol": 1}[b.kernel],
                                                  len(a), _dptr(q), _dptr(a), _dptr(c)))
        for b in model.core_pot:
            x, v = _f64(b.x), _f64(b.V)
            self._check(self.lib.tgx_add_core_pot(self.ctx, b.sp1, b.sp2, len(x), _dptr(x),
                                                  _dptr(v), b.yp1, b.ypn))
        self.model = model

    # -- hot path ----------------------------------------------------------
    def neighbors(self, pos, lattice, rcut: float):
        pos, lat = _f64(pos), _f64(lattice)
        n = pos.shape[0]
        need = C.c_int64(0)
        idx = np.zeros(3, np.int32)
        cap = max(64 * n, 1024)
        while True:
            nn = np.zeros(n, np.int32)
            lst = np.zeros(cap, np.int32)
            rjs = np.zeros(cap)
            xyz = np.zeros((cap, 3))
            th = np.zeros(cap)
            ph = np.zeros(cap)
            rc = self.lib.tgx_neighbors_build(self.ctx, n, _dptr(pos), _dptr(lat), rcut, cap,
                                              _iptr(nn), _iptr(lst), _dptr(rjs), _dptr(xyz),
                                              _dptr(th), _dptr(ph), C.byref(need), _iptr(idx))
            if rc == TGX_ERR_CAPACITY:
                cap = int(need.value)
                continue
            self._check(rc)
            m = int(need.value)
            return dict(n_neigh=nn, list=lst[:m], rjs=rjs[:m], xyz=xyz[:m], thetas=th[:m],
                        phis=ph[:m], indices=idx)

    def soap(self, pos, species, lattice):
        pos, lat, sp = _f64(pos), _f64(lattice), _i32(species)
        n = pos.shape[0]
        dim = self.model.hypers.dim
        out = np.zeros((n, dim))
        ns = C.c_int(0)
        site = np.zeros(n, np.int32)
        self._check(self.lib.tgx_soap_descriptors(self.ctx, n, _dptr(pos), _iptr(sp), _dptr(lat),
                                                  _dptr(out), C.byref(ns), _iptr(site)))
        return out[: ns.value], site[: ns.value] - 1

    def kernel_step(self, soap, want_qss: bool = True):
        q = _f64(soap)
        ns, dim = q.shape
        e = np.zeros(ns)
        qss = np.zeros((dim, ns)) if want_qss else None
        self._check(self.lib.tgx_gap_kernel_step(self.ctx, ns, _dptr(q), _dptr(e), _dptr(qss)))
        return e, (qss.T.copy() if want_qss else None)

    def energy_forces(self, pos, species, lattice, do_forces: bool = True):
        pos, lat, sp = _f64(pos), _f64(lattice), _i32(species)
        n = pos.shape[0]
        e = np.zeros(n)
        f = np.zeros((n, 3))
        v = np.zeros((3, 3))
        self._check(self.lib.tgx_energy_forces(self.ctx, n, _dptr(pos), _iptr(sp), _dptr(lat),
                                               int(do_forces), _dptr(e), _dptr(f), _dptr(v)))
        return e, f, v.T.copy()

    def energy_forces_noalloc(self, pos, species, lattice, do_forces: bool = True):
        """energy_forces with every work buffer in the library's per-GPU
        scratch array (tgx_energy_forces_noalloc): no device allocation."""
        pos, lat, sp = _f64(pos), _f64(lattice), _i32(species)
        n = pos.shape[0]
        e = np.zeros(n)
        f = np.zeros((n, 3))
        v = np.zeros((3, 3))
        self._check(self.lib.tgx_energy_forces_noalloc(self.ctx, n, _dptr(pos), _iptr(sp),
                                                       _dptr(lat), int(do_forces), _dptr(e),
                                                       _dptr(f), _dptr(v)))
        return e, f, v.T.copy()

    def energy_forces_batched(self, structures, do_forces: bool = True):
        """structures: list of (pos, species, lattice)."""
        offs = np.zeros(len(structures) + 1, np.int32)
        for s, (p, _, _) in enumerate(structures):
            offs[s + 1] = offs[s] + len(p)
        pos = _f64(np.concatenate([p for p, _, _ in structures]))
        sp = _i32(np.concatenate([s for _, s, _ in structures]))
        lat = _f64(np.stack([np.asarray(l, np.float64).reshape(9) for _, _, l in structures]))
        n = int(offs[-1])
        e = np.zeros(n)
        f = np.zeros((n, 3))
        v = np.zeros((len(structures), 3, 3))
        self._check(self.lib.tgx_energy_forces_batched(self.ctx, len(structures), _iptr(offs),
                                                       _dptr(pos), _iptr(sp), _dptr(lat),
                                                       int(do_forces), _dptr(e), _dptr(f),
                                                       _dptr(v)))
        return offs, e, f, np.transpose(v, (0, 2, 1)).copy()

    def mc_run(self, pos, species, lattice, n_trials: int, temp: float, move_max: float,
               skin: float, seed: int):
        """Single-atom-move Metropolis MC (energies only).  Returns dict with
        the final positions, the energy trace (n_trials + 1), accepted and
        list-rebuild counts."""
        pos = _f64(pos).copy()
        sp, lat = _i32(species), _f64(lattice)
        tr = np.zeros(n_trials + 1)
        acc, nreb = C.c_int(0), C.c_int(0)
        self._check(self.lib.tgx_mc_run(self.ctx, len(pos), _iptr(sp), _dptr(lat), _dptr(pos),
                                        n_trials, temp, move_max, skin, seed, _dptr(tr),
                                        C.byref(acc), C.byref(nreb)))
        return dict(pos=pos, e_trace=tr, n_accepted=acc.value, n_rebuilds=nreb.value)

    def mc_run_da(self, pos, species, lattice, n_trials: int, temp: float, move_max: float,
                  skin: float, seed: int, n_train_trials: int = 20000, train_every: int = 1024,
                  epochs: int = 4, batch: int = 512, lr: float = 1e-3, max_rmse: float = 0.5):
        """Delayed-acceptance MC with the surrogate (tgx_surrogate_init first)."""
        pos = _f64(pos).copy()
        sp, lat = _i32(species), _f64(lattice)
        tr = np.zeros(n_trials + 1)
        opt = DaOptions(n_train_trials, train_every, epochs, batch, lr, max_rmse)
        st = McStats()
        self._check(self.lib.tgx_mc_run_da(self.ctx, len(pos), _iptr(sp), _dptr(lat),
                                           _dptr(pos), n_trials, temp, move_max, skin, seed,
                                           C.byref(opt), _dptr(tr), C.byref(st)))
        out = {f[0]: getattr(st, f[0]) for f in McStats._fields_ if f[0] != "pad_"}
        out.update(pos=pos, e_trace=tr)
        return out

    # surrogate network
    def surrogate_init(self, dim: int, seed: int = 0):
        self._check(self.lib.tgx_surrogate_init(self.ctx, dim, seed))
        self._sg_dim = dim

    def surrogate_set(self, P, mu: float = 0.0, sigma: float = 1.0):
        P = _f64(P)
        self._check(self.lib.tgx_surrogate_set(self.ctx, len(P), _dptr(P), mu, sigma))

    def surrogate_get(self):
        P = np.zeros(surrogate_nparams(self._sg_dim))
        mu, sg = C.c_double(0), C.c_double(0)
        self._check(self.lib.tgx_surrogate_get(self.ctx, len(P), _dptr(P), C.byref(mu),
                                               C.byref(sg)))
        return P, mu.value, sg.value

    def surrogate_predict(self, q):
        q = _f64(q)
        e = np.zeros(len(q))
        self._check(self.lib.tgx_surrogate_predict(self.ctx, len(q), _dptr(q), _dptr(e)))
        return e

    def surrogate_gradient(self, q, y):
        q, y = _f64(q), _f64(y)
        g = np.zeros(surrogate_nparams(q.shape[1]))
        loss = C.c_double(0)
        self._check(self.lib.tgx_surrogate_gradient(self.ctx, len(q), _dptr(q), _dptr(y),
                                                    _dptr(g), C.byref(loss)))
        return g, loss.value

    def surrogate_fit(self, q, y, epochs: int = 10, batch: int = 256, lr: float = 1e-3,
                      seed: int = 0):
        q, y = _f64(q), _f64(y)
        r = C.c_double(0)
        self._check(self.lib.tgx_surrogate_fit(self.ctx, len(q), _dptr(q), _dptr(y), epochs,
                                               batch, lr, seed, C.byref(r)))
        return r.value

    def md_run(self, pos, vel, species, lattice, masses_amu, n_steps: int, dt: float,
               skin: float):
        pos, vel = _f64(pos).copy(), _f64(vel).copy()
        lat, sp, m = _f64(lattice), _i32(species), _f64(masses_amu)
        n = pos.shape[0]
        f = np.zeros((n, 3))
        ep = np.zeros(n_steps + 1)
        ek = np.zeros(n_steps + 1)
        nreb = C.c_int(0)
        self._check(self.lib.tgx_md_run(self.ctx, n, n_steps, dt, skin, _iptr(sp), _dptr(lat),
                                        _dptr(m), _dptr(pos), _dptr(vel), _dptr(f), _dptr(ep),
                                        _dptr(ek), C.byref(nreb)))
        return dict(pos=pos, vel=vel, forces=f, e_pot=ep, e_kin=ek, n_rebuilds=nreb.value)

    def md_begin(self, pos, vel, species, lattice, masses_amu, dt: float, skin: float,
                 history: int = 1024):
        pos, vel = _f64(pos), _f64(vel)
        lat, sp, m = _f64(lattice), _i32(species), _f64(masses_amu)
        self._md_n = pos.shape[0]
        self._md_hist = history
        self._check(self.lib.tgx_md_begin(self.ctx, self._md_n, _iptr(sp), _dptr(lat), _dptr(m),
                                          _dptr(pos), _dptr(vel), dt, skin, history))

    def md_steps(self, n_steps: int):
        self._check(self.lib.tgx_md_steps(self.ctx, n_steps))

    def md_read(self):
        n = self._md_n
        pos, vel, f = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        ep, ek = np.zeros(self._md_hist), np.zeros(self._md_hist)
        done, nreb = C.c_int(0), C.c_int(0)
        self._check(self.lib.tgx_md_read(self.ctx, _dptr(pos), _dptr(vel), _dptr(f), _dptr(ep),
                                         _dptr(ek), C.byref(done), C.byref(nreb)))
        m = min(done.value + 1, self._md_hist)
        return dict(pos=pos, vel=vel, forces=f, e_pot=ep[:m], e_kin=ek[:m], steps=done.value,
                    n_rebuilds=nreb.value)

    def md_end(self):
        self._check(self.lib.tgx_md_end(self.ctx))

    def sync(self):
        self._check(self.lib.tgx_sync(self.ctx))

    def stage_times(self):
        out = np.zeros(7)
        self._check(self.lib.tgx_stage_times(self.ctx, _dptr(out)))
        return dict(zip(["neighbors", "soap_fwd", "gemm", "soap_bwd", "two_body", "integrator",
                         "total"], out))

    def add_local_property(self, block: int, Qs, alphas, delta: float, zeta: float,
                           V0: float = 0.0) -> int:
        """A local-property model (e.g. Hirshfeld volumes) on SOAP block
        `block`; Qs (n_sparse, dim).  Returns its index."""
        q, a = _f64(Qs), _f64(alphas)
        idx = C.c_int(-1)
        self._check(self.lib.tgx_add_local_property(self.ctx, block, q.shape[1], q.shape[0],
                                                    _dptr(q), _dptr(a), delta, zeta, V0,
                                                    C.byref(idx)))
        return idx.value

    def clear_local_properties(self):
        self._check(self.lib.tgx_clear_local_properties(self.ctx))

    def local_property(self, index: int, pos, species, lattice, rcut_list: float = 0.0,
                       derivatives: bool = False):
        """V (n,) and, with derivatives, dV_i/dr_j (n_pairs, 3) in the order
        of neighbors(pos, lattice, rcut_list)."""
        pos, lat, sp = _f64(pos), _f64(lattice), _i32(species)
        n = pos.shape[0]
        V = np.zeros(n)
        if not derivatives:
            self._check(self.lib.tgx_local_property(self.ctx, index, n, _dptr(pos), _iptr(sp),
                                                    _dptr(lat), rcut_list, _dptr(V), 0, None,
                                                    None))
            return V, None
        need = C.c_int64(0)
        cap = max(64 * n, 1024)
        while True:
            der = np.zeros((cap, 3))
            rc = self.lib.tgx_local_property(self.ctx, index, n, _dptr(pos), _iptr(sp),
                                             _dptr(lat), rcut_list, _dptr(V), cap, _dptr(der),
                                             C.byref(need))
            if rc == TGX_ERR_CAPACITY:
                cap = int(need.value)
                continue
            self._check(rc)
            return V, der[: need.value]

    def set_vdw_ts(self, local_property: int, c6_ref, r0_ref, alpha0_ref, rcut=10.0,
                   buffer=1.0, rcut_inner=0.5, buffer_inner=0.5, sR=0.94, d=20.0):
        """Tkatchenko-Scheffler dispersion on the Hirshfeld volumes of local
        property `local_property`; per-species references (eV A^6, A, A^3)."""
        c6, r0, a0 = _f64(c6_ref), _f64(r0_ref), _f64(alpha0_ref)
        self._check(self.lib.tgx_set_vdw_ts(self.ctx, int(local_property), len(c6), _dptr(c6),
                                            _dptr(r0), _dptr(a0), rcut, buffer, rcut_inner,
                                            buffer_inner, sR, d))

    def clear_vdw(self):
        self._check(self.lib.tgx_clear_vdw(self.ctx))

    def vdw_energy_forces(self, pos, species, lattice, do_forces=True):
        """TS energies (n,), forces (n, 3) and virial (3, 3) of one structure
        (forces and virial None without do_forces)."""
        pos, lat, sp = _f64(pos), _f64(lattice), _i32(species)
        n = pos.shape[0]
        e = np.zeros(n)
        f = np.zeros((n, 3))
        v = np.zeros((3, 3))
        self._check(self.lib.tgx_vdw_energy_forces(self.ctx, n, _dptr(pos), _iptr(sp),
                                                   _dptr(lat), int(do_forces), _dptr(e),
                                                   _dptr(f), _dptr(v)))
        if not do_forces:
            return e, None, None
        return e, f, v.T.copy()

    def set_mc_batch(self, trials_per_batch: int):
        """Monte Carlo trials per speculative batch (0 automatic, 1 one at a
        time)."""
        self._check(self.lib.tgx_set_mc_batch(self.ctx, int(trials_per_batch)))

    def set_list_buffer(self, buffer: float):
        self._check(self.lib.tgx_set_list_buffer(self.ctx, float(buffer)))

    def set_timing(self, on: bool = True):
        """Stage and kernel timing events on or off (off by default)."""
        self._check(self.lib.tgx_set_timing(self.ctx, int(bool(on))))

    def kernel_times(self):
        """Event-bracketed device time (ms) of single kernels in the last
        energy_forces call."""
        out = np.zeros(4)
        self._check(self.lib.tgx_kernel_times(self.ctx, _dptr(out)))
        return dict(zip(["soap_fwd", "gemm", "soap_bwd", "force_gather"], out))
