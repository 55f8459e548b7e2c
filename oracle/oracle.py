"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  Each function names the
reference routine it restates (see tgx_oracle.c for what pins it).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "turbogap_2.0_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from tgx import CSoapHypers, GapModel, SoapHypers  # noqa: E402

LIB = os.path.join(HERE, "liboracle.so")
_lib = None


class C2b(C.Structure):
    _fields_ = [("sp1", C.c_int32), ("sp2", C.c_int32), ("n_sparse", C.c_int32),
                ("pad_", C.c_int32), ("rcut", C.c_double), ("buffer", C.c_double),
                ("delta", C.c_double), ("sigma", C.c_double),
                ("Qs", C.POINTER(C.c_double)), ("alphas", C.POINTER(C.c_double)),
                ("cutoff", C.POINTER(C.c_double))]


class C3b(C.Structure):
    _fields_ = [("sp0", C.c_int32), ("sp1", C.c_int32), ("sp2", C.c_int32),
                ("n_sparse", C.c_int32), ("kernel", C.c_int32), ("pad_", C.c_int32),
                ("rcut", C.c_double), ("delta", C.c_double), ("sigma", C.c_double * 3),
                ("Qs", C.POINTER(C.c_double)), ("alphas", C.POINTER(C.c_double)),
                ("cutoff", C.POINTER(C.c_double))]


class CCore(C.Structure):
    _fields_ = [("sp1", C.c_int32), ("sp2", C.c_int32), ("n", C.c_int32), ("pad_", C.c_int32),
                ("yp1", C.c_double), ("ypn", C.c_double), ("x", C.POINTER(C.c_double)),
                ("V", C.POINTER(C.c_double)), ("d2", C.POINTER(C.c_double))]


class CModel(C.Structure):
    pass


CModel._fields_ = [("h", CSoapHypers), ("has_soap", C.c_int32), ("dim", C.c_int32),
                   ("n_sparse", C.c_int32), ("n2b", C.c_int32),
                   ("Qs", C.POINTER(C.c_double)), ("alphas", C.POINTER(C.c_double)),
                   ("delta", C.c_double), ("zeta", C.c_double),
                   ("e0", C.c_double * 8), ("b2", C.POINTER(C2b)),
                   ("n_more", C.c_int32), ("pad_", C.c_int32), ("more", C.POINTER(CModel)),
                   ("n3b", C.c_int32), ("ncore", C.c_int32), ("b3", C.POINTER(C3b)),
                   ("core", C.POINTER(CCore))]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's parallel loops (omp_set_num_threads of
    the libgomp the oracle links); returns the count in effect."""
    L = lib()
    L.omp_set_num_threads(int(n))
    L.omp_get_max_threads.restype = C.c_int
    return int(L.omp_get_max_threads())


def threads() -> int:
    L = lib()
    L.omp_get_max_threads.restype = C.c_int
    return int(L.omp_get_max_threads())


def _d(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mic(pi, pj, L):
    """get_distance, orthorhombic branch (neighbors_interface.f90:547-581)."""
    out = np.zeros(3)
    d = C.c_double()
    lib().orc_mic(_d(_f64(pi)), _d(_f64(pj)), _d(_f64(L)), _d(out), C.byref(d))
    return out, d.value


def cells_for_cutoff(lattice, rcut):
    idx = np.zeros(3, np.int32)
    lib().orc_cells_for_cutoff(_d(_f64(lattice)), C.c_double(rcut), _i(idx))
    return idx


def neighbors(pos, lattice, rcut, algo=0):
    """build_neighbors_list (neighbors_interface.f90:144-524), sorted lists."""
    pos, lat = _f64(pos), _f64(lattice)
    n = len(pos)
    L = lib()
    np_ = C.c_int64(0)
    L.orc_neighbors.restype = C.c_int
    rc = L.orc_neighbors(n, _d(pos), _d(lat), C.c_double(rcut), algo, C.c_int64(0), None, None,
                         None, None, None, None, C.byref(np_), None)
    if rc not in (0, 3):
        raise RuntimeError(f"orc_neighbors failed {rc}")
    m = np_.value
    nn = np.zeros(n, np.int32)
    lst = np.zeros(m, np.int32)
    rjs = np.zeros(m)
    xyz = np.zeros((m, 3))
    th = np.zeros(m)
    ph = np.zeros(m)
    idx = np.zeros(3, np.int32)
    rc = L.orc_neighbors(n, _d(pos), _d(lat), C.c_double(rcut), algo, C.c_int64(m), _i(nn),
                         _i(lst), _d(rjs), _d(xyz), _d(th), _d(ph), C.byref(np_), _i(idx))
    assert rc == 0
    return dict(n_neigh=nn, list=lst, rjs=rjs, xyz=xyz, thetas=th, phis=ph, indices=idx)


def poly3_overlap(nmax):
    S = np.zeros((nmax, nmax))
    W = np.zeros((nmax, nmax))
    lib().orc_poly3_overlap(nmax, _d(S), _d(W))
    return S, W


def poly3gauss_overlap(nmax, rh, sigma):
    """Overlap S and W = S^(-1/2) of the poly3gauss basis (a Gaussian of
    width sigma at the origin, then nmax - 1 poly3 functions)."""
    S = np.zeros((nmax, nmax))
    W = np.zeros((nmax, nmax))
    lib().orc_poly3gauss_overlap(nmax, C.c_double(rh), C.c_double(sigma), _d(S), _d(W))
    return S, W


def radial(h: SoapHypers, s: int, r):
    """R_n(r), dR_n/dr of species s (0-based)."""
    r = _f64(np.atleast_1d(r))
    nmax = h.alpha_max[s]
    R = np.zeros((len(r), nmax))
    dR = np.zeros((len(r), nmax))
    ch = h.to_c()
    lib().orc_radial(C.byref(ch), s, len(r), _d(r), _d(R), _d(dR))
    return R, dR


def radial_raw(r, rh, sig, p, nk):
    out = np.zeros(nk)
    lib().orc_radial_raw(C.c_double(r), C.c_double(rh), C.c_double(sig), p, nk, _d(out))
    return out


def sph(lmax, u):
    L2 = (lmax + 1) ** 2
    Y = np.zeros(L2)
    dF = np.zeros((L2, 3))
    lib().orc_sph(lmax, _d(_f64(u)), _d(Y), _d(dF))
    return Y, dF


def soap_coeffs(h: SoapHypers, si: int, x, sj):
    """c_nlm of one site (central species si, neighbour vectors x (np, 3)
    with species sj, 0-based) as (n_r, (l_max+1)^2), lm = l*l + l + m."""
    x = _f64(np.atleast_2d(x)) if len(x) else np.zeros((0, 3))
    sj = np.ascontiguousarray(sj, dtype=np.int32)
    c = np.zeros((h.n_r, (h.l_max + 1) ** 2))
    ch = h.to_c()
    rc = lib().orc_soap_coeffs(C.byref(ch), si, len(x), _d(x), _i(sj), _d(c))
    assert rc == 0, rc
    return c


def soap(h: SoapHypers, pos, species, lattice, derivatives=False):
    """get_soap: descriptors (n_sites, dim) and optionally soap_cart_der as
    (n_pairs, dim, 3), with pair_atom / site_atom / site_n_neigh."""
    pos, lat = _f64(pos), _f64(lattice)
    sp = np.ascontiguousarray(species, dtype=np.int32)
    n = len(pos)
    ch = h.to_c()
    L = lib()
    ns = C.c_int(0)
    npairs = C.c_int64(0)
    out = np.zeros((n, h.dim))
    site = np.zeros(n, np.int32)
    nn = np.zeros(n, np.int32)
    if not derivatives:
        rc = L.orc_soap(C.byref(ch), n, _d(pos), _d(lat), _i(sp), _d(out), None, C.c_int64(0),
                        C.byref(ns), C.byref(npairs), _i(site), None, None, _i(nn))
        if rc != 0:
            raise RuntimeError(f"orc_soap failed {rc}")
        return dict(soap=out[: ns.value], site_atom=site[: ns.value])
    rc = L.orc_soap(C.byref(ch), n, _d(pos), _d(lat), _i(sp), _d(out), None, C.c_int64(0),
                    C.byref(ns), C.byref(npairs), None, None, None, None)
    m = npairs.value
    der = np.zeros((m, h.dim, 3))
    pa = np.zeros(m, np.int32)
    pxyz = np.zeros((m, 3))
    rc = L.orc_soap(C.byref(ch), n, _d(pos), _d(lat), _i(sp), _d(out), _d(der), C.c_int64(m),
                    C.byref(ns), C.byref(npairs), _i(site), _i(pa), _d(pxyz), _i(nn))
    assert rc == 0, rc
    k = ns.value
    return dict(soap=out[:k], soap_der=der, site_atom=site[:k], pair_atom=pa, pair_xyz=pxyz,
                site_n_neigh=nn[:k])


def local_property(h: SoapHypers, Qs, alphas, delta, zeta, V0, pos, species, lattice,
                   rcut_list=None, derivatives=False):
    """local_property_predict (local_properties.f90:46-146) on get_soap's
    descriptors: V_i = delta^2 sum_s alphas_s K_is^zeta + V0 clipped at 0
    (zeta integer when |zeta - nint(zeta)| < 1e-10); with derivatives
    V_der = Qss(i,:) . soap_cart_der(:,:,pair), zero where V_i = 0, scattered
    onto the full list at rcut_list (get_local_properties,
    gap_interface.f90:342-351), pairs matched by atom and xyz."""
    pos, lat = _f64(pos), _f64(lattice)
    sp = np.ascontiguousarray(species, np.int32)
    n = len(pos)
    s = soap(h, pos, sp, lat, derivatives=derivatives)
    q = s["soap"]
    Qs, alphas = _f64(Qs), _f64(alphas)
    K = q @ Qs.T
    zi = int(round(zeta))
    integer = abs(zi - zeta) < 1e-10
    Kz = K ** zi if integer else K ** zeta
    Vs = delta ** 2 * (Kz @ alphas) + V0
    Vs = np.where(Vs < 0.0, 0.0, Vs)
    V = np.zeros(n)
    V[s["site_atom"]] = Vs
    if not derivatives:
        return V, None
    Kd = zi * K ** (zi - 1) if integer else zeta * K ** (zeta - 1.0)
    Qss = delta ** 2 * (Kd * alphas[None, :]) @ Qs
    full = neighbors(pos, lat, rcut_list)
    off = np.concatenate([[0], np.cumsum(full["n_neigh"])])
    # entries are matched by (i, j) and the vector; the vectors of the two
    # lists agree to rounding (a list at a larger cutoff may use a larger
    # supercell), images of one j are far apart
    lookup = {}
    for i in range(n):
        for k in range(off[i], off[i + 1]):
            lookup.setdefault((i, (full["list"][k] - 1) % n), []).append(k)
    der = np.zeros((len(full["list"]), 3))
    soff = np.concatenate([[0], np.cumsum(s["site_n_neigh"])])
    for si, i in enumerate(s["site_atom"]):
        if Vs[si] == 0.0:
            continue
        for k in range(soff[si], soff[si + 1]):
            cand = lookup[(int(i), int(s["pair_atom"][k]))]
            dist = [np.abs(full["xyz"][kk] - s["pair_xyz"][k]).max() for kk in cand]
            kk = cand[int(np.argmin(dist))]
            assert min(dist) < 1e-9
            der[kk] = Qss[si] @ s["soap_der"][k]
    return V, der


def gap_kernel_step(soap_, Qs, alphas, delta, zeta):
    """gap.f90:118-171: energies and Qss (n_sites, dim)."""
    q = _f64(soap_)
    ns, dim = q.shape
    Qs = _f64(Qs)
    e = np.zeros(ns)
    qss = np.zeros((dim, ns))
    lib().orc_gap_kernel_step(ns, dim, len(Qs), _d(q), _d(Qs), _d(_f64(alphas)),
                              C.c_double(delta), C.c_double(zeta), _d(e), _d(qss))
    return e, qss.T.copy()


def gap_soap_ef(soap_, soap_der, alphas, delta, zeta, e0, Qs, n_neigh, pair_atom, xyz,
                n_atoms, do_forces=True):
    """get_soap_energy_and_forces (gap.f90:39-244)."""
    q = _f64(soap_)
    ns, dim = q.shape
    e = np.zeros(ns)
    f = np.zeros((n_atoms, 3))
    v = np.zeros((3, 3))
    lib().orc_gap_soap_ef(ns, dim, len(Qs), _d(q), _d(_f64(soap_der)), _d(_f64(alphas)),
                          C.c_double(delta), C.c_double(zeta), C.c_double(e0), _d(_f64(Qs)),
                          _i(np.ascontiguousarray(n_neigh, np.int32)),
                          _i(np.ascontiguousarray(pair_atom, np.int32)), _d(_f64(xyz)),
                          n_atoms, int(do_forces), _d(e), _d(f), _d(v))
    return e, f, v.T.copy()


def _cmodel(model: GapModel):
    keep = []
    cm = CModel()
    cm.h = model.hypers.to_c()
    cm.has_soap = 1
    Qs = _f64(model.Qs)
    al = _f64(model.alphas)
    keep += [Qs, al]
    cm.dim = Qs.shape[1]
    cm.n_sparse = Qs.shape[0]
    cm.Qs = _d(Qs)
    cm.alphas = _d(al)
    cm.delta = model.delta
    cm.zeta = model.zeta
    e0 = model.e0 if model.e0 is not None else np.zeros(model.hypers.n_species)
    for s, v in enumerate(e0):
        cm.e0[s] = float(v)
    arr = (C2b * max(1, len(model.two_body)))()
    for k, b in enumerate(model.two_body):
        q, a, c = _f64(b.Qs), _f64(b.alphas), _f64(b.cutoff)
        keep += [q, a, c]
        arr[k].sp1, arr[k].sp2, arr[k].n_sparse = b.sp1, b.sp2, len(q)
        arr[k].rcut, arr[k].buffer, arr[k].delta, arr[k].sigma = b.rcut, b.buffer, b.delta, b.sigma
        arr[k].Qs, arr[k].alphas, arr[k].cutoff = _d(q), _d(a), _d(c)
    keep.append(arr)
    cm.n2b = len(model.two_body)
    cm.b2 = C.cast(arr, C.POINTER(C2b))
    more = list(getattr(model, "more_soap", []) or [])
    marr = (CModel * max(1, len(more)))()
    for k, blk in enumerate(more):
        q, a = _f64(blk.Qs), _f64(blk.alphas)
        keep += [q, a]
        marr[k].h = blk.hypers.to_c()
        marr[k].has_soap = 1
        marr[k].dim, marr[k].n_sparse = q.shape[1], q.shape[0]
        marr[k].Qs, marr[k].alphas = _d(q), _d(a)
        marr[k].delta, marr[k].zeta = blk.delta, blk.zeta
    keep.append(marr)
    cm.n_more = len(more)
    cm.more = C.cast(marr, C.POINTER(CModel))
    a3 = list(getattr(model, "angle_3b", []) or [])
    arr3 = (C3b * max(1, len(a3)))()
    for k, b in enumerate(a3):
        q = _f64(np.asarray(b.Qs, float).T.copy())  # Qs(n_sparse, 3) column-major
        a, c = _f64(b.alphas), _f64(b.cutoff)
        keep += [q, a, c]
        arr3[k].sp0, arr3[k].sp1, arr3[k].sp2 = b.sp0, b.sp1, b.sp2
        arr3[k].n_sparse, arr3[k].kernel = len(a), {"exp": 0, "pol": 1}[b.kernel]
        arr3[k].rcut, arr3[k].delta = b.rcut, b.delta
        for t in range(3):
            arr3[k].sigma[t] = float(b.sigma[t])
        arr3[k].Qs, arr3[k].alphas, arr3[k].cutoff = _d(q), _d(a), _d(c)
    keep.append(arr3)
    cm.n3b = len(a3)
    cm.b3 = C.cast(arr3, C.POINTER(C3b))
    cp = list(getattr(model, "core_pot", []) or [])
    arrc = (CCore * max(1, len(cp)))()
    for k, b in enumerate(cp):
        x, v, d2, y1, yn = core_pot_setup(b.x, b.V, b.yp1, b.ypn)
        keep += [x, v, d2]
        arrc[k].sp1, arrc[k].sp2, arrc[k].n = b.sp1, b.sp2, len(x)
        arrc[k].yp1, arrc[k].ypn = y1, yn
        arrc[k].x, arrc[k].V, arrc[k].d2 = _d(x), _d(v), _d(d2)
    keep.append(arrc)
    cm.ncore = len(cp)
    cm.core = C.cast(arrc, C.POINTER(CCore))
    return cm, keep


def core_pot_setup(x, V, yp1, ypn):
    """read_gap.f90:701-752: spline second derivatives and end slopes."""
    x, V = _f64(x), _f64(V)
    d2 = np.zeros(len(x))
    y1, yn = C.c_double(yp1), C.c_double(ypn)
    lib().orc_core_pot_setup(len(x), _d(x), _d(V), C.byref(y1), C.byref(yn), _d(d2))
    return x, V, d2, y1.value, yn.value


def spline(x, V, d2, yp1, ypn, r, rcut, derivative=False):
    """spline / spline_der (utils/splines.f90:33-156) at each r."""
    fn = lib().orc_spline
    fn.restype = C.c_double
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                   C.c_double, C.c_double, C.c_int]
    x, V, d2 = _f64(x), _f64(V), _f64(d2)
    return np.array([fn(len(x), _d(x), _d(V), _d(d2), yp1, ypn, float(t), rcut, int(derivative))
                     for t in np.atleast_1d(r)])


def gap_2b(b, pos, species, lattice, do_forces=True):
    """get_2b_energy_and_forces (gap.f90:329-460) on the oracle's list."""
    nb = neighbors(pos, lattice, b.rcut)
    n = len(pos)
    arr = C2b()
    q, a, c = _f64(b.Qs), _f64(b.alphas), _f64(b.cutoff)
    arr.sp1, arr.sp2, arr.n_sparse = b.sp1, b.sp2, len(q)
    arr.rcut, arr.buffer, arr.delta, arr.sigma = b.rcut, b.buffer, b.delta, b.sigma
    arr.Qs, arr.alphas, arr.cutoff = _d(q), _d(a), _d(c)
    e = np.zeros(n)
    f = np.zeros((n, 3))
    v = np.zeros((3, 3))
    sp = np.ascontiguousarray(species, np.int32)
    lib().orc_gap_2b_ef(C.byref(arr), n, _i(sp), _i(nb["n_neigh"]), _i(nb["list"]),
                        _d(nb["rjs"]), _d(_f64(nb["xyz"])), int(do_forces), _d(e), _d(f), _d(v))
    return e, f, v.T.copy()


class CTs(C.Structure):
    _fields_ = [("n_species", C.c_int32), ("pad_", C.c_int32),
                ("c6_ref", C.POINTER(C.c_double)), ("r0_ref", C.POINTER(C.c_double)),
                ("alpha0_ref", C.POINTER(C.c_double)), ("rcut", C.c_double),
                ("buffer", C.c_double), ("rcut_inner", C.c_double),
                ("buffer_inner", C.c_double), ("sR", C.c_double), ("d", C.c_double)]


def ts_vdw(c6_ref, r0_ref, alpha0_ref, pos, species, lattice, V, V_der, rcut=10.0, buffer=1.0,
           rcut_inner=0.5, buffer_inner=0.5, sR=0.94, d=20.0, rcut_list=None, do_forces=True):
    """get_ts_energy_and_forces (vdw/vdw_interface.f90:127-385) on the full
    list at rcut_list (default rcut).  V (n,) Hirshfeld volumes, V_der
    (n_pairs, 3) dV_i/dr_j in that list's order (local_property with the same
    rcut_list), or None when do_forces is false.  Returns energies, forces
    (n, 3) and the virial (3, 3)."""
    pos, lat = _f64(pos), _f64(lattice)
    sp = np.ascontiguousarray(species, np.int32)
    n = len(pos)
    nb = neighbors(pos, lat, rcut if rcut_list is None else rcut_list)
    c6, r0, a0 = _f64(c6_ref), _f64(r0_ref), _f64(alpha0_ref)
    t = CTs()
    t.n_species = len(c6)
    t.c6_ref, t.r0_ref, t.alpha0_ref = _d(c6), _d(r0), _d(a0)
    t.rcut, t.buffer, t.rcut_inner, t.buffer_inner, t.sR, t.d = (
        rcut, buffer, rcut_inner, buffer_inner, sR, d)
    V = _f64(V)
    der = _f64(np.zeros((len(nb["list"]), 3)) if V_der is None else V_der)
    assert der.shape == (len(nb["list"]), 3)
    e = np.zeros(n)
    f = np.zeros((n, 3))
    v = np.zeros((3, 3))
    lib().orc_ts_ef(C.byref(t), n, _i(sp), _i(nb["n_neigh"]), _i(nb["list"]), _d(nb["rjs"]),
                    _d(_f64(nb["xyz"])), _d(V), _d(der), int(do_forces), _d(e), _d(f), _d(v))
    return e, f, v.T.copy()


def energy_forces(model: GapModel, pos, species, lattice, do_forces=True, chunk_sites=0,
                  soap=True):
    """Whole path for one structure (calculate_gap_soap + calculate_gap_2b + e0)."""
    pos, lat = _f64(pos), _f64(lattice)
    sp = np.ascontiguousarray(species, np.int32)
    n = len(pos)
    cm, keep = _cmodel(model)
    cm.has_soap = 1 if soap else 0
    e = np.zeros(n)
    f = np.zeros((n, 3))
    v = np.zeros((3, 3))
    rc = lib().orc_energy_forces(C.byref(cm), n, _d(pos), _i(sp), _d(lat), int(do_forces),
                                 chunk_sites, _d(e), _d(f), _d(v))
    if rc != 0:
        raise RuntimeError(f"orc_energy_forces failed {rc}")
    return e, f, v.T.copy()


def velocity_verlet(pos, pos_prev, vel, forces, forces_prev, masses_ev, dt, dt_prev,
                    first_step):
    """velocity_verlet (md_utils.f90:486-532), arrays updated in place."""
    dtp = C.c_double(dt_prev)
    lib().orc_velocity_verlet(len(pos), _d(pos), _d(pos_prev), _d(vel), _d(_f64(forces)),
                              _d(forces_prev), _d(_f64(masses_ev)), C.c_double(dt),
                              C.byref(dtp), int(first_step))
    return dtp.value


def mc_accept(e_new, e_prev, temp):
    fn = lib().orc_mc_accept
    fn.restype = C.c_double
    return fn(C.c_double(e_new), C.c_double(e_prev), C.c_double(temp))


def mc_uniform(seed, t, k):
    fn = lib().orc_mc_uniform
    fn.restype = C.c_double
    fn.argtypes = [C.c_uint64, C.c_int64, C.c_int]
    return fn(seed, t, k)


def mc_run(model: GapModel, pos, species, lattice, n_trials, temp, move_max, seed):
    """Sequential single-atom-move Metropolis MC with a full-system energy per
    trial (mc_interface.f90:183-447).  Returns final positions, the energy
    trace (n_trials + 1) and the accepted count."""
    pos = _f64(pos).copy()
    lat = _f64(lattice)
    sp = np.ascontiguousarray(species, np.int32)
    cm, keep = _cmodel(model)
    tr = np.zeros(n_trials + 1)
    acc = C.c_int(0)
    fn = lib().orc_mc_run
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                   C.c_double, C.c_uint64, C.c_void_p, C.c_void_p]
    rc = fn(C.addressof(cm), len(pos), pos.ctypes.data, sp.ctypes.data, lat.ctypes.data,
            n_trials, temp, move_max, seed, tr.ctypes.data, C.addressof(acc))
    if rc != 0:
        raise RuntimeError(f"orc_mc_run failed {rc}")
    return pos, tr, acc.value


def mlp(P, q, mu, sigma, H=128, round_bf16=True, y=None, grad=False):
    """Surrogate MLP energies (and the loss gradient when y is given)."""
    q = _f64(q)
    n, D = q.shape
    P = _f64(P)
    e = np.zeros(n)
    g = np.zeros_like(P) if grad else None
    fn = lib().orc_mlp
    fn.restype = C.c_double
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                   C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    yy = None if y is None else _f64(y)
    loss = fn(n, D, H, P.ctypes.data, q.ctypes.data, mu, sigma, int(round_bf16),
              None if yy is None else yy.ctypes.data, e.ctypes.data,
              None if g is None else g.ctypes.data)
    return (e, loss, g) if grad else e


def mc_run_da(model: GapModel, pos, species, lattice, n_trials, temp, move_max, seed, P, mu,
              sigma, max_rmse=0.5, with_fallback=False):
    """Delayed-acceptance MC with a fixed surrogate (P, mu, sigma), falling
    back to plain Metropolis once the running RMS of the stage-2 energy
    difference exceeds max_rmse (max_rmse < 0: never).  Returns positions,
    trace, accepted and stage-1 rejection counts (and, with_fallback, the
    first plain trial or -1)."""
    pos = _f64(pos).copy()
    lat = _f64(lattice)
    sp = np.ascontiguousarray(species, np.int32)
    cm, keep = _cmodel(model)
    P = _f64(P)
    tr = np.zeros(n_trials + 1)
    acc, r1, fb = C.c_int(0), C.c_int(0), C.c_int(-1)
    fn = lib().orc_mc_run_da
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                   C.c_double, C.c_uint64, C.c_void_p, C.c_double, C.c_double, C.c_double,
                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    rc = fn(C.addressof(cm), len(pos), pos.ctypes.data, sp.ctypes.data, lat.ctypes.data,
            n_trials, temp, move_max, seed, P.ctypes.data, mu, sigma, max_rmse, tr.ctypes.data,
            C.addressof(acc), C.addressof(r1), C.addressof(fb))
    if rc != 0:
        raise RuntimeError(f"orc_mc_run_da failed {rc}")
    if with_fallback:
        return pos, tr, acc.value, r1.value, fb.value
    return pos, tr, acc.value, r1.value


def mc_toy(n, prm, x0, n_steps, temp, move_max, seed, use_surrogate=False, variant=0):
    """Independent chains (one per row of x0, n particles each) of the shared
    Metropolis / delayed-acceptance driver on the toy potential
    V = sum a (x^2-1)^2 + c x0 x1, surrogate V + eps sum sin(w x);
    prm = (a, c, eps, w).  Returns the final x (n_chains, n) and the
    accepted and stage-1-rejected counts."""
    x0 = _f64(x0).reshape(-1, n)
    prm = _f64(prm)
    out = np.zeros_like(x0)
    acc, r1 = C.c_int64(0), C.c_int64(0)
    fn = lib().orc_mc_toy
    fn.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64,
                   C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    rc = fn(n, prm.ctypes.data, len(x0), n_steps, temp, move_max, seed, int(use_surrogate),
            variant, x0.ctypes.data, out.ctypes.data, C.addressof(acc), C.addressof(r1))
    if rc != 0:
        raise RuntimeError(f"orc_mc_toy failed {rc}")
    return out, acc.value, r1.value
