"""Pins the FP64 CPU oracle against mathematics and independent numpy code.

The reference ships no golden vectors for this path and its SOAP library
(soap_turbo) is absent, so each oracle function is checked against what the
mathematics fixes: quadrature, library special functions, invariances,
finite differences, closed forms, and independent brute-force restatements.
"""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import sph_harm_y

import oracle
import structures as S
import tgx
import workloads as W

KB_SINGLE = float(np.float32(8.617333262e-5))


def _fd5(f, x, h):
    """Fourth-order central difference."""
    return (-f(x + 2 * h) + 8 * f(x + h) - 8 * f(x - h) + f(x - 2 * h)) / (12 * h)


# --------------------------------------------------------------------------
# radial basis
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nmax", [1, 4, 8, 12])
def test_poly3_overlap_vs_quadrature(nmax):
    S_, W_ = oracle.poly3_overlap(nmax)
    rh = 3.7
    for a in range(1, nmax + 1):
        for b in range(1, nmax + 1):
            Na = math.sqrt((2 * a + 5) / rh ** (2 * a + 5))
            Nb = math.sqrt((2 * b + 5) / rh ** (2 * b + 5))
            v = quad(lambda r: Na * Nb * (rh - r) ** (a + 2) * (rh - r) ** (b + 2), 0, rh,
                     epsrel=1e-13)[0]
            assert abs(v - S_[a - 1, b - 1]) < 1e-12


@pytest.mark.parametrize("nmax,rh,sg", [(1, 3.7, 0.5), (4, 3.7, 0.5), (8, 3.7, 0.5),
                                         (8, 5.0, 0.3), (12, 4.0, 0.45)])
def test_poly3gauss_overlap_vs_quadrature(nmax, rh, sg):
    """The poly3gauss overlap: a Gaussian of width sigma at the origin and
    n_max - 1 poly3 functions, each normalised on [0, rh] by quadrature."""
    h = tgx.SoapHypers.uniform(1, alpha_max=nmax, rcut_hard=rh, atom_sigma_r=sg)
    h = tgx.SoapHypers(**{**h.__dict__, "basis": "poly3gauss"})
    S_, W_ = oracle.poly3gauss_overlap(nmax, rh, sg)
    fs = _basis_functions(h, 0)
    for a in range(nmax):
        for b in range(nmax):
            v = quad(lambda u: fs[a](u) * fs[b](u), 0, rh, epsabs=1e-15, epsrel=1e-13)[0]
            assert abs(v - S_[a, b]) < 2e-12, (a, b, v, S_[a, b])


@pytest.mark.parametrize("nmax,tol", [(4, 1e-13), (8, 1e-9), (12, 1e-6)])
def test_poly3gauss_W_is_inverse_sqrt(nmax, tol):
    """W S W = I with S evaluated exactly (50 digits, mpmath's erf and
    quadrature) and the FP64 W of the oracle."""
    mp.mp.dps = 50
    rh, sg = mp.mpf("3.7"), mp.mpf("0.5")
    _, W_ = oracle.poly3gauss_overlap(nmax, 3.7, 0.5)
    assert np.array_equal(W_, W_.T)
    g = lambda u: mp.exp(-u * u / (2 * sg * sg))
    fs = [g] + [lambda u, a=a: (rh - u) ** (a + 2) for a in range(1, nmax)]
    nrm = [1 / mp.sqrt(mp.quad(lambda u: f(u) ** 2, [0, rh])) for f in fs]
    Sm = mp.matrix(nmax, nmax)
    for a in range(nmax):
        for b in range(nmax):
            Sm[a, b] = nrm[a] * nrm[b] * mp.quad(lambda u: fs[a](u) * fs[b](u), [0, rh])
    R = mp.matrix(W_.tolist()) * Sm * mp.matrix(W_.tolist()) - mp.eye(nmax)
    assert float(max(abs(x) for x in R)) < tol
    assert np.all(np.linalg.eigvalsh(W_) > 0)


@pytest.mark.parametrize("p,sigs", [(0, 0.0), (1, 0.0), (2, 0.0), (0, 0.08), (2, 0.1)])
def test_poly3gauss_radial_vs_quadrature_and_derivative(p, sigs):
    """R_n of the poly3gauss basis (W times the integrals of each basis
    function against the smeared density u^p G) against adaptive
    quadrature, with r-dependent width sigma_r + sigma_r_scaling r; dR/dr by
    five-point differences."""
    h = tgx.SoapHypers.uniform(1, alpha_max=8, rcut_hard=3.7, buffer=0.0, atom_sigma_r=0.5,
                               atom_sigma_r_scaling=sigs, amplitude_scaling=0.0)
    h = tgx.SoapHypers(**{**h.__dict__, "radial_enhancement": p, "basis": "poly3gauss"})
    r = np.array([0.05, 0.6, 1.3, 2.2, 3.0, 3.6])
    R, dR = oracle.radial(h, 0, r)
    for k, rk in enumerate(r):
        ref = _radial_quad(h, 0, rk, 0.5 + sigs * rk)
        assert np.abs(R[k] - ref).max() < 1e-8 * max(1.0, np.abs(ref).max()), (rk, R[k], ref)
    fd = _fd5(lambda x: oracle.radial(h, 0, x)[0], r, 1e-3)
    assert np.abs(fd - dR).max() < 1e-6 * max(1.0, np.abs(dR).max())


@pytest.mark.parametrize("nmax,tol", [(4, 1e-13), (8, 1e-9), (12, 1e-6)])
def test_W_is_inverse_sqrt(nmax, tol):
    """W S W = I evaluated exactly (50 digits) with the FP64 W; W symmetric."""
    mp.mp.dps = 50
    _, W_ = oracle.poly3_overlap(nmax)
    assert np.array_equal(W_, W_.T)
    Sm = mp.matrix(nmax, nmax)
    for a in range(1, nmax + 1):
        for b in range(1, nmax + 1):
            Sm[a - 1, b - 1] = mp.sqrt((2 * a + 5) * (2 * b + 5)) / (a + b + 5)
    R = mp.matrix(W_.tolist()) * Sm * mp.matrix(W_.tolist()) - mp.eye(nmax)
    assert float(max(abs(x) for x in R)) < tol
    # positive definite
    assert np.all(np.linalg.eigvalsh(W_) > 0)


@pytest.mark.parametrize("p", [0, 1, 2])
def test_radial_raw_vs_quadrature(p):
    rh, sig = 3.7, 0.5
    for r in [0.0, 0.4, 1.2, 2.5, 3.3, 3.69]:
        out = oracle.radial_raw(r, rh, sig, p, 8)
        for k in range(8):
            ref = quad(lambda x: x ** p * (rh - x) ** (k + 3) * math.exp(-(x - r) ** 2 / (2 * sig * sig)),
                       0, rh, epsabs=1e-15, epsrel=1e-13)[0] / (math.sqrt(2 * math.pi) * sig)
            assert abs(out[k] - ref) <= 1e-11 * max(abs(ref), 1e-6), (r, k, out[k], ref)


def test_radial_derivative_and_limits():
    h = S.carbon_hypers()
    r = np.linspace(0.3, 3.19, 41)  # below the soft radius (C2 kink there)
    R, dR = oracle.radial(h, 0, r)
    assert np.abs(_fd5(lambda x: oracle.radial(h, 0, x)[0], r, 1e-3) - dR).max() < 1e-7
    r2 = np.linspace(3.21, 3.68, 11)
    _, dR2 = oracle.radial(h, 0, r2)
    assert np.abs(_fd5(lambda x: oracle.radial(h, 0, x)[0], r2, 1e-4) - dR2).max() < 1e-6
    # zero at and beyond the hard cutoff, continuous with zero slope there
    R1, dR1 = oracle.radial(h, 0, np.array([3.7, 4.0]))
    assert np.all(R1 == 0) and np.all(dR1 == 0)
    R2, dR2 = oracle.radial(h, 0, np.array([3.7 - 1e-6]))
    assert np.abs(R2).max() < 1e-9 and np.abs(dR2).max() < 1e-4


def test_radial_sigma_scaling_derivative():
    h = S.SoapHypers.uniform(1, atom_sigma_r_scaling=0.1, atom_sigma_t_scaling=0.05,
                             amplitude_scaling=2.0, radial_enhancement=1)
    r = np.linspace(0.5, 3.6, 17)
    r = np.linspace(0.5, 3.15, 17)
    _, dR = oracle.radial(h, 0, r)
    fd = _fd5(lambda x: oracle.radial(h, 0, x)[0], r, 1e-3)
    assert np.abs(fd - dR).max() < 1e-6 * max(1.0, np.abs(dR).max())


# --------------------------------------------------------------------------
# spherical harmonics
# --------------------------------------------------------------------------
def test_real_harmonics_vs_scipy():
    rng = np.random.default_rng(0)
    L = 10
    for _ in range(20):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        Y, _ = oracle.sph(L, u)
        th, ph = math.acos(u[2]), math.atan2(u[1], u[0])
        for l in range(L + 1):
            for m in range(-l, l + 1):
                yc = sph_harm_y(l, abs(m), th, ph)
                if m == 0:
                    ref = yc.real
                elif m > 0:
                    ref = math.sqrt(2) * (-1) ** m * yc.real
                else:
                    ref = math.sqrt(2) * (-1) ** m * yc.imag
                assert abs(Y[l * l + l + m] - ref) < 1e-13


def test_real_harmonics_orthonormal():
    """Quadrature on the sphere (Lebedev-free: product Gauss-Legendre x
    uniform phi) gives the identity Gram matrix."""
    L = 8
    xg, wg = np.polynomial.legendre.leggauss(24)
    nphi = 48
    G = np.zeros(((L + 1) ** 2, (L + 1) ** 2))
    for z, w in zip(xg, wg):
        s = math.sqrt(1 - z * z)
        for k in range(nphi):
            ph = 2 * math.pi * k / nphi
            Y, _ = oracle.sph(L, [s * math.cos(ph), s * math.sin(ph), z])
            G += w * (2 * math.pi / nphi) * np.outer(Y, Y)
    assert np.abs(G - np.eye(len(G))).max() < 1e-12


def test_harmonic_gradient_fd():
    rng = np.random.default_rng(1)
    L = 8
    for _ in range(5):
        v = rng.normal(size=3)
        r = np.linalg.norm(v)
        u = v / r
        Y, dF = oracle.sph(L, u)
        # grad_r Y(r/|r|) = (dF - (u.dF) u)/r
        grad = (dF - (dF @ u)[:, None] * u[None, :]) / r
        e = 1e-6
        for k in range(3):
            vp, vm = v.copy(), v.copy()
            vp[k] += e
            vm[k] -= e
            Yp, _ = oracle.sph(L, vp / np.linalg.norm(vp))
            Ym, _ = oracle.sph(L, vm / np.linalg.norm(vm))
            assert np.abs((Yp - Ym) / (2 * e) - grad[:, k]).max() < 1e-7


# --------------------------------------------------------------------------
# SOAP descriptor
# --------------------------------------------------------------------------
def _cluster(n=12, seed=5):
    rng = np.random.default_rng(seed)
    pts = [np.zeros(3)]
    while len(pts) < n:
        p = rng.uniform(-3.3, 3.3, 3)
        if np.linalg.norm(p) < 3.6 and all(np.linalg.norm(p - q) > 1.1 for q in pts):
            pts.append(p)
    return np.array(pts) + 20.0


def test_soap_rotation_permutation_invariance():
    h = S.carbon_hypers()
    lat = S.cubic_lattice(40.0)
    pos = _cluster()
    sp = np.ones(len(pos), np.int32)
    q0 = oracle.soap(h, pos, sp, lat)["soap"][0]
    assert abs(np.linalg.norm(q0) - 1) < 1e-14
    rng = np.random.default_rng(2)
    A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    rot = (pos - 20.0) @ A.T + 20.0
    perm = np.concatenate([[0], 1 + rng.permutation(len(pos) - 1)])
    q1 = oracle.soap(h, rot[perm], sp, lat)["soap"][0]
    assert np.abs(q1 - q0).max() < 1e-11
    # reflection too (power spectrum is parity-even)
    q2 = oracle.soap(h, (pos - 20.0) * [-1, 1, 1] + 20.0, sp, lat)["soap"][0]
    assert np.abs(q2 - q0).max() < 1e-11


def test_soap_isolated_atom_and_dimer_closed_form():
    """A lone atom has only c_n00 = w R_n(0) Y00; a dimer adds
    R_n(r) A_l Y_lm(u) - so p_{nn'l} for l > 0 is R_n R_n' A_l^2 (2l+1)/(4 pi)."""
    h = S.carbon_hypers()
    lat = S.cubic_lattice(30.0)
    r = 2.1
    pos = np.array([[10.0, 10.0, 10.0], [10.0, 10.0 + r, 10.0]])
    sp = np.ones(2, np.int32)
    q = oracle.soap(h, pos, sp, lat)["soap"][0]
    R, _ = oracle.radial(h, 0, [r])
    R = R[0]
    A = np.exp(-0.5 * np.arange(9) * np.arange(1, 10) * 0.25 / r ** 2)
    nr, L = 8, 8
    p = []
    for n in range(nr):
        for n2 in range(n, nr):
            w = 1.0 if n == n2 else math.sqrt(2)
            for l in range(L + 1):
                if l == 0:
                    continue
                p.append((n, n2, l, w * R[n] * R[n2] * A[l] ** 2 * (2 * l + 1) / (4 * math.pi)))
    # compare l > 0 components up to the common normalisation
    idx = lambda n, n2, l: (n * nr - n * (n - 1) // 2 + (n2 - n)) * (L + 1) + l
    ref = np.array([v for *_, v in p])
    got = np.array([q[idx(n, n2, l)] for n, n2, l, _ in p])
    scale = got[np.argmax(np.abs(ref))] / ref[np.argmax(np.abs(ref))]
    assert np.abs(got - scale * ref).max() < 1e-10 * np.abs(got).max()


def test_soap_derivatives_fd():
    pos, sp, lat, _ = W.carbon216()
    pos, sp = pos[:80].copy(), sp[:80]
    lat = S.cubic_lattice(12.915)  # keep density structure, box unchanged
    h = S.carbon_hypers()
    d = oracle.soap(h, pos, sp, lat, derivatives=True)
    off = np.concatenate([[0], np.cumsum(d["site_n_neigh"])])
    i = 3
    pa = d["pair_atom"][off[i]:off[i + 1]]
    der = d["soap_der"][off[i]:off[i + 1]]
    e = 1e-4
    for jj in [0, 1, 2, len(pa) - 1]:
        j = pa[jj]
        for k in range(3):
            vals = []
            for s in (2, 1, -1, -2):
                pp = pos.copy()
                pp[j, k] += s * e
                vals.append(oracle.soap(h, pp, sp, lat)["soap"][i])
            fd = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * e)
            assert np.abs(fd - der[jj, :, k]).max() < 2e-7


# --------------------------------------------------------------------------
# kernel step, full path
# --------------------------------------------------------------------------
@pytest.mark.parametrize("zeta", [4.0, 2.5, 1.0])
def test_kernel_step_closed_form(zeta):
    rng = np.random.default_rng(3)
    q = rng.random((30, 45))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    Qs = rng.random((20, 45))
    Qs /= np.linalg.norm(Qs, axis=1, keepdims=True)
    al = rng.normal(0, 0.01, 20)
    delta = 1.3
    e, qss = oracle.gap_kernel_step(q, Qs, al, delta, zeta)
    K = np.einsum("id,sd->is", q, Qs)
    e_ref = delta ** 2 * (K ** zeta) @ al
    qss_ref = -zeta * delta ** 2 * (K ** (zeta - 1) * al) @ Qs
    assert np.abs(e - e_ref).max() < 1e-15
    assert np.abs(qss - qss_ref).max() < 1e-15


def _strained(pos, lat, eps):
    F = np.eye(3) + eps
    return pos @ F.T, (lat.reshape(3, 3) @ F.T).reshape(9)


@pytest.fixture(scope="module")
def c64():
    pos, sp, lat = S.carbon_cell(64, 2.0, 1.2, 21)
    h = S.carbon_hypers()
    q = oracle.soap(h, pos, sp, lat)["soap"]
    model = S.model_from_soaps(h, q, 80, seed=21, two_body=[S.carbon_2b(21)])
    return pos, sp, lat, model


def test_forces_are_minus_gradient(c64):
    pos, sp, lat, model = c64
    e, f, v = oracle.energy_forces(model, pos, sp, lat)
    h = 1e-4
    for a in [0, 7, 33]:
        for k in range(3):
            Es = []
            for s in (2, 1, -1, -2):
                p = pos.copy()
                p[a, k] += s * h
                Es.append(oracle.energy_forces(model, p, sp, lat, do_forces=False)[0].sum())
            fd = -(-Es[0] + 8 * Es[1] - 8 * Es[2] + Es[3]) / (12 * h)
            assert abs(fd - f[a, k]) < 1e-8 * max(1.0, abs(f[a, k])) + 1e-9


def test_virial_is_minus_strain_derivative(c64):
    """virial = -dE/d(strain) for SOAP and 2-body terms together (the
    reference's conventions at gap.f90:201-207 and 437-447)."""
    pos, sp, lat, model = c64
    _, _, v = oracle.energy_forces(model, pos, sp, lat)
    h = 1e-5
    # orthorhombic cells only: diagonal strains; off-diagonal via the
    # symmetric Cauchy-Born pair rule on a big cell is not orthorhombic.
    for a in range(3):
        Es = []
        for s in (1, -1):
            eps = np.zeros((3, 3))
            eps[a, a] = s * h
            p, l = _strained(pos, lat, eps)
            Es.append(oracle.energy_forces(model, p, sp, l, do_forces=False)[0].sum())
        fd = -(Es[0] - Es[1]) / (2 * h)
        assert abs(fd - v[a, a]) < 1e-6 * max(1.0, abs(v[a, a]))
    assert np.abs(v - v.T).max() < 1e-12


def test_virial_offdiagonal_pair_identity(c64):
    """Off-diagonal virial equals sum over pairs of sym(F_ij x r_ij) computed
    independently from the materialised soap_cart_der (gap.f90:191-209)."""
    pos, sp, lat, model = c64
    d = oracle.soap(model.hypers, pos, sp, lat, derivatives=True)
    e, f, v = oracle.gap_soap_ef(d["soap"], d["soap_der"], model.alphas, model.delta,
                                 model.zeta, 0.0, model.Qs, d["site_n_neigh"], d["pair_atom"],
                                 d["pair_xyz"], len(pos))
    K = d["soap"] @ model.Qs.T
    Qss = -model.zeta * model.delta ** 2 * (K ** (model.zeta - 1) * model.alphas) @ model.Qs
    off = np.concatenate([[0], np.cumsum(d["site_n_neigh"])])
    vv = np.zeros((3, 3))
    ff = np.zeros_like(f)
    for i in range(len(d["soap"])):
        for l in range(off[i], off[i + 1]):
            F = Qss[i] @ d["soap_der"][l]
            ff[d["pair_atom"][l]] += F
            x = d["pair_xyz"][l]
            vv += 0.5 * (np.outer(F, x) + np.outer(x, F))
    assert np.abs(vv - v).max() < 1e-12
    assert np.abs(ff - f).max() < 1e-12


def test_extensivity_supercell(c64):
    pos, sp, lat, model = c64
    e1, f1, v1 = oracle.energy_forces(model, pos, sp, lat)
    L = lat[0]
    shifts = np.array([[i, j, k] for i in range(2) for j in range(2) for k in range(2)]) * L
    pos2 = np.concatenate([pos + s for s in shifts])
    sp2 = np.tile(sp, 8)
    e2, f2, v2 = oracle.energy_forces(model, pos2, sp2, S.cubic_lattice(2 * L))
    assert abs(e2.sum() - 8 * e1.sum()) < 1e-9
    assert np.abs(f2[:64] - f1).max() < 1e-9
    assert np.abs(v2 - 8 * v1).max() < 1e-8


def test_chunked_equals_unchunked(c64):
    pos, sp, lat, model = c64
    a = oracle.energy_forces(model, pos, sp, lat, chunk_sites=0)
    b = oracle.energy_forces(model, pos, sp, lat, chunk_sites=7)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() < 1e-13


def test_translation_sum_of_forces(c64):
    pos, sp, lat, model = c64
    _, f, _ = oracle.energy_forces(model, pos, sp, lat)
    assert np.abs(f.sum(0)).max() < 1e-11


# --------------------------------------------------------------------------
# 2-body
# --------------------------------------------------------------------------
def test_2b_dimer_closed_form():
    b = S.carbon_2b(3)
    lat = S.cubic_lattice(30.0)
    for r in (1.5, 3.1, 3.45, 3.69):  # plain, and inside the 0.5 buffer
        pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
        sp = np.ones(2, np.int32)
        e, f, v = oracle.gap_2b(b, pos, sp, lat)
        fc = 1.0 if r < b.rcut - b.buffer else (math.cos(math.pi * (r - b.rcut + b.buffer) / b.buffer) + 1) / 2
        phi = b.delta ** 2 * np.sum(b.alphas * b.cutoff * fc * np.exp(-0.5 * (r - b.Qs) ** 2 / b.sigma ** 2))
        assert abs(e[0] - phi) < 1e-15 and abs(e[1] - phi) < 1e-15
        # F = -dE_total/dx with E_total = 2 phi
        hstep = 1e-6
        def etot(rr):
            p = np.array([[5.0, 5.0, 5.0], [5.0 + rr, 5.0, 5.0]])
            return oracle.gap_2b(b, p, sp, lat, do_forces=False)[0].sum()
        dE = (etot(r + hstep) - etot(r - hstep)) / (2 * hstep)
        assert abs(f[1, 0] - (-dE)) < 1e-8 and abs(f[0, 0] - dE) < 1e-8
        # virial = -dE/dstrain_xx = -dE/dr * r
        assert abs(v[0, 0] - (-dE * r)) < 1e-8


# --------------------------------------------------------------------------
# neighbours
# --------------------------------------------------------------------------
def _numpy_neighbors(pos, L, rc):
    """Independent brute force with explicit periodic images (any box)."""
    n = len(pos)
    reps = [int(math.ceil(rc / L)) + 1] * 3
    out = []
    for i in range(n):
        s = set()
        for a in range(-reps[0], reps[0] + 1):
            for b in range(-reps[1], reps[1] + 1):
                for c in range(-reps[2], reps[2] + 1):
                    d = pos + np.array([a, b, c]) * L - pos[i]
                    r = np.linalg.norm(d, axis=1)
                    for j in np.nonzero(r < rc)[0]:
                        if not (j == i and a == b == c == 0):
                            s.add((j, round(r[j], 9)))
        out.append(sorted(s))
    return out


@pytest.mark.parametrize("n,rho,rc", [(216, 2.0, 3.7), (32, 3.2, 3.7), (40, 1.5, 4.2)])
def test_neighbors_vs_numpy(n, rho, rc):
    pos, sp, lat = S.carbon_cell(n, rho, 1.2, 10 + n)
    nb = oracle.neighbors(pos, lat, rc)
    ref = _numpy_neighbors(pos, lat[0], rc)
    off = np.concatenate([[0], np.cumsum(nb["n_neigh"])])
    for i in range(n):
        lst = nb["list"][off[i]:off[i + 1]]
        assert lst[0] == i + 1 and nb["rjs"][off[i]] == 0.0
        got = sorted(((j - 1) % n, round(r, 9)) for j, r in zip(lst[1:], nb["rjs"][off[i] + 1:off[i + 1]]))
        assert got == ref[i]
        assert np.all(np.diff(lst[1:]) > 0)
        # xyz consistent with rjs and theta/phi
        x = nb["xyz"][off[i] + 1:off[i + 1]]
        r = nb["rjs"][off[i] + 1:off[i + 1]]
        assert np.abs(np.linalg.norm(x, axis=1) - r).max() < 1e-12
        th = nb["thetas"][off[i] + 1:off[i + 1]]
        assert np.abs(np.cos(th) * r - x[:, 2]).max() < 1e-12


def test_neighbors_cells_equal_brute():
    pos, sp, lat = S.carbon_cell(500, 2.0, 1.2, 77)
    a = oracle.neighbors(pos, lat, 3.7, algo=0)
    b = oracle.neighbors(pos, lat, 3.7, algo=1)
    for k in a:
        assert np.array_equal(a[k], b[k])


def test_mic_matches_gfortran_modulo():
    """modulo(a,p) = fmod(a,p) (+p when a<0 and remainder non-zero); fold
    at L/2; d = sqrt(dx^2+dy^2+dz^2) in that order."""
    L = np.array([10.0, 11.0, 12.5])
    rng = np.random.default_rng(4)
    for _ in range(200):
        pi, pj = rng.uniform(-25, 25, 3), rng.uniform(-25, 25, 3)
        d, r = oracle.mic(pi, pj, L)
        ref = []
        for k in range(3):
            a = pj[k] - pi[k]
            t = math.fmod(a, L[k])
            if t != 0.0 and (a < 0) != (L[k] < 0):
                t = t + L[k]
            if t > L[k] / 2.0:
                t = t - L[k]
            ref.append(t)
        assert d.tolist() == ref
        assert r == math.sqrt((ref[0] * ref[0] + ref[1] * ref[1]) + ref[2] * ref[2])


# --------------------------------------------------------------------------
# MD / MC
# --------------------------------------------------------------------------
def test_velocity_verlet_harmonic():
    """Two unit masses on a spring: VV conserves energy to O(dt^2) and
    matches the analytic oscillation to O(dt^2)."""
    k, m = 2.0, 1.0
    def forces(x):
        d = x[1, 0] - x[0, 0] - 1.0
        return np.array([[k * d, 0, 0], [-k * d, 0, 0]], float)
    errs = []
    for dt in (0.02, 0.01):
        x = np.array([[0.0, 0, 0], [1.2, 0, 0]])
        v = np.zeros((2, 3))
        xp, fp = np.zeros_like(x), np.zeros_like(x)
        masses = np.full(2, m)
        f = forces(x)
        dtp = dt
        t = 0.0
        nsteps = int(round(2.0 / dt))
        for s in range(nsteps):
            dtp = oracle.velocity_verlet(x, xp, v, f, fp, masses, dt, dtp, s == 0)
            f = forces(x)
            t += dt
        w = math.sqrt(2 * k / m)
        ext = (x[1, 0] - x[0, 0]) - 1.0
        errs.append(abs(ext - 0.2 * math.cos(w * t)))
    assert errs[1] < errs[0] / 3.0  # second order


def test_mc_acceptance_formula():
    for de in (-0.1, 0.0, 0.01, 0.05, 0.3):
        p = oracle.mc_accept(1.0 + de, 1.0, 1500.0)
        assert p == min(1.0, math.exp(-((1.0 + de) - 1.0) / (KB_SINGLE * 1500.0)))


# ---------------------------------------------------------------------------
# Monte Carlo trial stream and the sequential Metropolis restatement
# ---------------------------------------------------------------------------

def test_mc_uniform_is_splitmix64():
    """u(0, 0, k) are the first outputs of splitmix64 from state 0 (published
    values 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F), top
    53 bits over 2^53."""
    for k, v in enumerate([0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]):
        assert oracle.mc_uniform(0, 0, k) == (v >> 11) / 2.0 ** 53
    # trial t continues the same counter: u(0, t, k) = output 6t + k + 1
    assert oracle.mc_uniform(0, 1, 0) == oracle.mc_uniform(0, 0, 6)
    u = np.array([oracle.mc_uniform(7, t, k) for t in range(2000) for k in range(5)])
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def _mc_system(n=24, seed=11):
    pos, sp, lat = S.carbon_cell(n, 2.0, 1.2, seed)
    _, _, _, model = W.carbon216()
    return pos, sp, lat, model


def test_mc_hot_limit_accepts_everything():
    pos, sp, lat, model = _mc_system()
    p, tr, acc = oracle.mc_run(model, pos, sp, lat, 12, 1e12, 0.6, 5)
    assert acc == 12
    # every move applied: atom floor(u0 n) moved by move_max (u - 1/2)
    q = pos.copy()
    for t in range(12):
        m = int(np.floor(oracle.mc_uniform(5, t, 0) * len(q)))
        q[m] += [0.6 * (oracle.mc_uniform(5, t, 1 + k) - 0.5) for k in range(3)]
    assert np.array_equal(p, q)
    e, _, _ = oracle.energy_forces(model, p, sp, lat, do_forces=False)
    assert abs(e.sum() - tr[-1]) < 1e-12


def test_mc_cold_limit_only_downhill():
    pos, sp, lat, model = _mc_system()
    p, tr, acc = oracle.mc_run(model, pos, sp, lat, 20, 1e-9, 0.6, 9)
    d = np.diff(tr)
    assert np.all(d <= 0.0)
    assert acc == int(np.sum(d < 0))
    e, _, _ = oracle.energy_forces(model, p, sp, lat, do_forces=False)
    assert abs(e.sum() - tr[-1]) < 1e-12


# ---------------------------------------------------------------------------
# Surrogate network and delayed acceptance
# ---------------------------------------------------------------------------

def _mlp_case(n=6, D=20, seed=3):
    rng = np.random.default_rng(seed)
    H = 128
    NP = H * D + H * H + 3 * H + 1
    P = rng.normal(0, 0.2, NP)
    q = rng.normal(0, 1, (n, D))
    y = rng.normal(0, 1, n)
    return P, q, y


def test_mlp_gradient_matches_finite_differences():
    """orc_mlp's loss gradient (exact arithmetic, no BF16 rounding) against
    central differences on a spread of parameters of every block."""
    P, q, y = _mlp_case()
    mu, sigma = 0.3, 1.7
    e, loss, g = oracle.mlp(P, q, mu, sigma, round_bf16=False, y=y, grad=True)
    assert loss == pytest.approx(np.mean(((e - y) / sigma) ** 2), rel=1e-12)
    D, H = q.shape[1], 128
    idx = [0, 7, H * D - 1, H * D + 5, H * D + H + 11, H * D + H + H * H + 3,
           H * D + H + H * H + H + 9, H * D + H * H + 3 * H]
    for k in idx:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        _, lp, _ = oracle.mlp(Pp, q, mu, sigma, round_bf16=False, y=y, grad=True)
        _, lm, _ = oracle.mlp(Pm, q, mu, sigma, round_bf16=False, y=y, grad=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_mlp_closed_form_single_unit():
    """A network whose only path is W1[0][0] -> unit 0 -> w3[0] computes
    mu + sigma (w3 s(W2 s(W1 q + b1) + b2) + b3) by hand."""
    D, H = 3, 128
    P = np.zeros(H * D + H * H + 3 * H + 1)
    P[0] = 0.75                      # W1[0][0]
    P[H * D] = -0.2                  # b1[0]
    P[H * D + H] = 1.5               # W2[0][0]
    P[H * D + H + H * H] = 0.1       # b2[0]
    P[H * D + H + H * H + H] = 2.0   # w3[0]
    P[-1] = 0.05                     # b3
    s = lambda x: x / (1 + np.exp(-x))  # noqa: E731
    q = np.array([[0.4, -1.0, 2.0]])
    z1 = 0.75 * 0.4 - 0.2
    want = 1.0 + 0.5 * (2.0 * s(1.5 * s(z1) + 0.1) + 0.05 + 127 * 0.0)
    # units 1..127 have s(0) = 0 in layer 1 and s(b2 = 0) = 0 in layer 2
    got = oracle.mlp(P, q, 1.0, 0.5, round_bf16=False)
    assert got[0] == pytest.approx(want, rel=1e-14)


def test_delayed_acceptance_flat_surrogate_is_plain_metropolis():
    """A zero network and no 2-body term make stage 1 pass every trial; the
    chain is then plain Metropolis drawn with u_5 (composition of the pinned
    energy, acceptance and stream)."""
    pos, sp, lat, model = _mc_system()
    m = tgx.GapModel(model.hypers, model.Qs, model.alphas, model.delta, model.zeta)
    D = model.hypers.dim
    P = np.zeros(128 * D + 128 * 128 + 3 * 128 + 1)
    p_da, tr_da, acc, r1 = oracle.mc_run_da(m, pos, sp, lat, 10, 1500.0, 0.6, 17, P, 0.0, 1.0)
    assert r1 == 0
    q = pos.copy()
    e_cur = oracle.energy_forces(m, q, sp, lat, do_forces=False)[0].sum()
    n_acc = 0
    for t in range(10):
        k = int(np.floor(oracle.mc_uniform(17, t, 0) * len(q)))
        old = q[k].copy()
        q[k] = old + [0.6 * (oracle.mc_uniform(17, t, 1 + j) - 0.5) for j in range(3)]
        e_new = oracle.energy_forces(m, q, sp, lat, do_forces=False)[0].sum()
        if oracle.mc_accept(e_new, e_cur, 1500.0) > oracle.mc_uniform(17, t, 5):
            e_cur, n_acc = e_new, n_acc + 1
        else:
            q[k] = old
        assert abs(tr_da[t + 1] - e_cur) < 1e-12
    assert n_acc == acc and np.array_equal(q, p_da)


def test_non_default_nf_is_refused():
    """nf is read by read_gap.f90:164-165,192-194 (default 4) and handed to
    soap_turbo, whose use of it is not in the reference tree: the oracle
    refuses other values instead of treating them as 4."""
    import dataclasses

    pos, sp, lat = S.carbon_cell(32, 2.0, 1.2, 3)
    h = dataclasses.replace(S.carbon_hypers(), nf=(4.5,))
    with pytest.raises(RuntimeError, match="failed 4"):
        oracle.soap(h, pos, sp, lat)
    model = S.model_from_soaps(S.carbon_hypers(), oracle.soap(S.carbon_hypers(), pos, sp, lat)["soap"], 8, seed=3)
    model = dataclasses.replace(model, hypers=h)
    with pytest.raises(RuntimeError, match="failed 4"):
        oracle.energy_forces(model, pos, sp, lat)


# ---------------------------------------------------------------------------
# Sampling statistics of the Metropolis / delayed-acceptance driver
# (mc_utils.f90:172-180, 215-277; Christen & Fox 2005).  orc_mc_run and
# orc_mc_run_da decide moves with the same driver the toy chains use here.
# Chains start from the exact Boltzmann distribution, so after any number of
# steps their end points are independent draws of it when the acceptance
# rule is right: a chi-squared test on those end points checks that.

KB_SINGLE = float(np.float32(8.617333262e-5))


# Bin edges of the tests (outer bins catch the tails) and a midpoint-rule
# grid whose cells tile every bin exactly, so that the expected bin masses
# and the cell-wise start draws carry no edge error.
_STEP1, _STEP2 = 0.15, 0.4
_EDGES1 = np.concatenate([[-9.0], np.arange(-12, 13) * _STEP1, [9.0]])
_EDGES2 = np.concatenate([[-9.0], np.arange(-4, 5) * _STEP2, [9.0]])


def _boltzmann_grid(prm, n, temp):
    a, c = prm[0], prm[1]
    step = _STEP1 if n == 1 else _STEP2
    lo = (-12 - 6) * _STEP1 if n == 1 else (-4 - 3) * _STEP2
    h = step / 32
    x = lo + (np.arange(int(round(-2 * lo / h))) + 0.5) * h
    kT = KB_SINGLE * temp
    if n == 1:
        v = a * (x * x - 1) ** 2
        return x, h, np.exp(-(v - v.min()) / kT)
    X0, X1 = np.meshgrid(x, x, indexing="ij")
    v = a * (X0 ** 2 - 1) ** 2 + a * (X1 ** 2 - 1) ** 2 + c * X0 * X1
    return x, h, np.exp(-(v - v.min()) / kT)


def _draw(prm, n, temp, k, seed):
    """k draws from the Boltzmann density: a grid cell by its mass, then a
    uniform point in it."""
    x, h, w = _boltzmann_grid(prm, n, temp)
    rng = np.random.default_rng(seed)
    p = (w / w.sum()).ravel()
    idx = rng.choice(p.size, size=k, p=p)
    if n == 1:
        return (x[idx] + (rng.random(k) - 0.5) * h).reshape(k, 1)
    i0, i1 = np.unravel_index(idx, w.shape)
    return np.stack([x[i0], x[i1]], 1) + (rng.random((k, 2)) - 0.5) * h


def _chi2_pvalue(samples, prm, n, temp, edges):
    from scipy.stats import chi2

    x, _, w = _boltzmann_grid(prm, n, temp)
    w = w / w.sum()
    if n == 1:
        obs, _ = np.histogram(samples[:, 0], bins=edges)
        exp, _ = np.histogram(x, bins=edges, weights=w)
    else:
        obs, _, _ = np.histogram2d(samples[:, 0], samples[:, 1], bins=[edges, edges])
        X0, X1 = np.meshgrid(x, x, indexing="ij")
        exp, _, _ = np.histogram2d(X0.ravel(), X1.ravel(), bins=[edges, edges], weights=w.ravel())
        obs, exp = obs.ravel(), exp.ravel()
    exp = exp / exp.sum() * obs.sum()
    keep = exp >= 5
    stat = np.sum((obs[keep] - exp[keep]) ** 2 / exp[keep])
    return float(chi2.sf(stat, keep.sum() - 1)), stat, int(keep.sum())


@pytest.mark.parametrize("n,da", [(1, False), (1, True), (2, False), (2, True)])
def test_mc_driver_samples_boltzmann(n, da):
    """Plain Metropolis and delayed acceptance with a surrogate that is
    neither exact nor flat (V + 0.03 sin 5x) sample exp(-V/kT): 1-D double
    well and two coupled wells."""
    prm = (0.1, 0.04, 0.03, 5.0)
    T = 1000.0
    k = 20000
    x0 = _draw(prm, n, T, k, seed=10 + n)
    xe, acc, r1 = oracle.mc_toy(n, prm, x0, 200, T, 1.0, 1234 + n, use_surrogate=da)
    p, stat, nb = _chi2_pvalue(xe, prm, n, T, _EDGES1 if n == 1 else _EDGES2)
    print(f"n={n} da={da} acc={acc / (k * 200):.3f} rej1={r1} chi2={stat:.1f}/{nb - 1} p={p:.3g}")
    assert p > 1e-3
    if da:
        assert r1 > 0  # the surrogate stage does reject trials


@pytest.mark.parametrize("n", [1, 2])
def test_mc_driver_chi2_rejects_wrong_stage2_ratio(n):
    """The same test has the power to see a plausible mistake: stage 2 with
    the plain ratio exp(-dE/kT) instead of exp(-(dE - dS)/kT) (planted in the
    oracle's driver, variant 1) fails it."""
    prm = (0.1, 0.04, 0.03, 5.0)
    T = 1000.0
    k = 20000
    x0 = _draw(prm, n, T, k, seed=20 + n)
    xe, _, _ = oracle.mc_toy(n, prm, x0, 200, T, 1.0, 99 + n, use_surrogate=True, variant=1)
    p, stat, nb = _chi2_pvalue(xe, prm, n, T, _EDGES1 if n == 1 else _EDGES2)
    print(f"n={n} wrong ratio chi2={stat:.1f}/{nb - 1} p={p:.3g}")
    assert p < 1e-6


def test_two_soap_blocks_add_up():
    """Several soap_turbo blocks (calculate_gap_soap.f90:324-447): the model
    of two blocks gives e0 plus the sum of each block's energies, the sum of
    their forces and virials -- checked against the two one-block models."""
    import dataclasses

    pos, sp, lat = S.cho_cell(60, 1.5, 1.0, 21)
    h1 = dataclasses.replace(S.cho_hypers(), central_species=1)
    h2 = dataclasses.replace(S.cho_hypers(), central_species=2, l_max=5, alpha_max=(4, 5, 3),
                             rcut_hard=(3.4, 3.6, 3.5))
    m1 = S.model_from_soaps(h1, oracle.soap(h1, pos, sp, lat)["soap"], 20, seed=1)
    m2 = S.model_from_soaps(h2, oracle.soap(h2, pos, sp, lat)["soap"], 15, seed=2, zeta=3.0,
                            delta=0.7)
    e0 = np.array([-1.0, -0.5, -2.0])
    m1.e0 = e0
    m2.e0 = np.zeros(3)
    both = dataclasses.replace(m1, more_soap=[m2])
    e, f, v = oracle.energy_forces(both, pos, sp, lat)
    e1, f1, v1 = oracle.energy_forces(m1, pos, sp, lat)
    e2, f2, v2 = oracle.energy_forces(m2, pos, sp, lat)
    assert np.abs(e - (e1 + e2)).max() < 1e-12
    assert np.abs(f - (f1 + f2)).max() < 1e-12
    assert np.abs(v - (v1 + v2)).max() < 1e-12
    # each block's sites: block 2 contributes only to its central species' energies
    assert np.all(e2[sp != 2] == 0.0) and np.any(e2[sp == 2] != 0.0)


def _numpy_neighbors_general(pos, lat, rc, reach=4):
    """Independent brute force for any cell: every image r_j + T M (T in
    [-reach, reach]^3, M rows = lattice vectors) within rc of r_i."""
    M = np.asarray(lat, float).reshape(3, 3)
    T = np.array([[a, b, c] for a in range(-reach, reach + 1) for b in range(-reach, reach + 1)
                  for c in range(-reach, reach + 1)], float)
    sh = T @ M
    out = []
    for i in range(len(pos)):
        d = (pos[None, :, :] + sh[:, None, :]) - pos[i]
        r = np.linalg.norm(d, axis=2)
        hits = []
        for t, j in zip(*np.nonzero(r < rc)):
            if not (j == i and np.all(T[t] == 0)):
                hits.append((j, round(r[t, j], 9)))
        out.append(sorted(hits))
    return out


@pytest.mark.parametrize("n,rho,rc,skew,rot", [
    (40, 2.0, 3.0, (0.3, -0.2, 0.25), False),
    (24, 2.0, 3.7, (0.45, 0.1, -0.3), False),   # supercell images
    (12, 1.6, 3.5, (-0.2, 0.35, 0.4), True),    # rotated, supercell
    (60, 2.4, 2.5, (0.5, 0.5, 0.5), True),
])
def test_neighbors_general_cell_vs_numpy(n, rho, rc, skew, rot):
    """General cells (get_distance's non-orthorhombic branch,
    neighbors_interface.f90:582-629; brute force :390-421): the pair set and
    distances against an exhaustive image search; self first, then ascending
    supercell id; xyz, rjs, theta consistent."""
    pos, sp, lat = S.triclinic_cell(n, rho, 1.1, 300 + n, skew=skew, rotate=rot)
    nb = oracle.neighbors(pos, lat, rc)
    ref = _numpy_neighbors_general(pos, lat, rc)
    off = np.concatenate([[0], np.cumsum(nb["n_neigh"])])
    for i in range(n):
        lst = nb["list"][off[i]:off[i + 1]]
        assert lst[0] == i + 1 and nb["rjs"][off[i]] == 0.0
        got = sorted(((j - 1) % n, round(r, 9)) for j, r in zip(lst[1:], nb["rjs"][off[i] + 1:off[i + 1]]))
        assert got == ref[i]
        assert np.all(np.diff(lst[1:]) > 0)
        x = nb["xyz"][off[i] + 1:off[i + 1]]
        r = nb["rjs"][off[i] + 1:off[i + 1]]
        assert np.abs(np.linalg.norm(x, axis=1) - r).max() < 1e-12
        th = nb["thetas"][off[i] + 1:off[i + 1]]
        assert np.abs(np.cos(th) * r - x[:, 2]).max() < 1e-12
    assert np.prod(nb["indices"]) >= 1


def test_neighbors_general_cell_rotation_invariant():
    """Rotating the whole system leaves pair set and distances unchanged."""
    pos, sp, lat = S.triclinic_cell(30, 2.0, 1.1, 41, skew=(0.2, 0.3, -0.1))
    q, _ = np.linalg.qr(np.random.default_rng(5).normal(size=(3, 3)))
    a = oracle.neighbors(pos, lat, 3.2)
    b = oracle.neighbors(pos @ q.T, (lat.reshape(3, 3) @ q.T).reshape(9), 3.2)
    assert np.array_equal(a["n_neigh"], b["n_neigh"])
    assert np.array_equal(a["list"], b["list"])
    assert np.abs(a["rjs"] - b["rjs"]).max() < 1e-12


# --------------------------------------------------------------------------
# c_nlm against a 3-D quadrature of the smeared density
# --------------------------------------------------------------------------
def _real_sph_scipy(L, th, ph):
    """Real spherical harmonics from scipy's complex ones, lm = l*l + l + m
    (the convention pinned by test_real_harmonics_vs_scipy)."""
    out = np.zeros(((L + 1) ** 2,) + np.shape(th))
    for l in range(L + 1):
        for m in range(-l, l + 1):
            yc = sph_harm_y(l, abs(m), th, ph)
            if m == 0:
                v = yc.real
            elif m > 0:
                v = math.sqrt(2) * (-1) ** m * yc.real
            else:
                v = math.sqrt(2) * (-1) ** m * yc.imag
            out[l * l + l + m] = v
    return out


def _sphere_grid(nz=200, nphi=400):
    z, wz = np.polynomial.legendre.leggauss(nz)
    ph = 2 * math.pi * np.arange(nphi) / nphi
    Z, PH = np.meshgrid(z, ph, indexing="ij")
    W = np.repeat(wz[:, None], nphi, axis=1) * (2 * math.pi / nphi)
    th = np.arccos(Z)
    U = np.stack([np.sin(th) * np.cos(PH), np.sin(th) * np.sin(PH), Z], axis=-1)
    return th, PH, U, W


def _basis_functions(h, s):
    """The unnormalised radial basis functions of species s, normalised here
    by quadrature: poly3 (rh - u)^(a+2), a = 1..n_max; poly3gauss the
    Gaussian exp(-u^2/(2 sigma_r^2)) then (rh - u)^(a+2), a = 1..n_max-1."""
    rh, nmax = h.rcut_hard[s], h.alpha_max[s]
    if h.basis == "poly3gauss":
        sg = h.atom_sigma_r[s]
        fs = [lambda u: math.exp(-u * u / (2 * sg * sg))]
        fs += [lambda u, a=a: (rh - u) ** (a + 2) for a in range(1, nmax)]
    else:
        fs = [lambda u, a=a: (rh - u) ** (a + 2) for a in range(1, nmax + 1)]
    norms = [1 / math.sqrt(quad(lambda u: f(u) ** 2, 0, rh, epsabs=1e-16, epsrel=1e-13)[0])
             for f in fs]
    return [lambda u, f=f, c=c: c * f(u) for f, c in zip(fs, norms)]


def _basis_W(h, s):
    if h.basis == "poly3gauss":
        return oracle.poly3gauss_overlap(h.alpha_max[s], h.rcut_hard[s], h.atom_sigma_r[s])[1]
    return oracle.poly3_overlap(h.alpha_max[s])[1]


def _radial_quad(h, s, r, sig):
    """sum_a W_na int_0^rh phi_a(u) u^p G_sig(u - r) du by adaptive
    quadrature (G the unit-area Gaussian), phi_a the normalised basis."""
    rh = h.rcut_hard[s]
    p = h.radial_enhancement
    W = _basis_W(h, s)
    I = np.zeros(h.alpha_max[s])
    for a, phi in enumerate(_basis_functions(h, s)):
        f = lambda u: (phi(u) * u ** p
                       * math.exp(-(u - r) ** 2 / (2 * sig * sig)) / (math.sqrt(2 * math.pi) * sig))
        I[a] = quad(f, 0.0, rh, points=[min(max(r, 1e-9), rh)], epsabs=1e-15, epsrel=1e-13,
                    limit=200)[0]
    return W @ I


def _cnlm_by_quadrature(h, si, x, sj, grid):
    """c_nlm of the density: central atom (isotropic, radial Gaussian at 0,
    weight central_weight) plus, per neighbour, a radial Gaussian at r_j
    times the normalised angular profile exp(kappa (u.u_j - 1)) of a 3-D
    Gaussian of width st on the sphere of radius r_j, kappa = r_j^2/st^2.
    Radial integrals by adaptive quadrature, angular ones by product
    Gauss-Legendre x trapezoid quadrature on the sphere with scipy's
    harmonics: nothing shared with the oracle's closed forms but W."""
    th, PH, U, Wq = grid
    L = h.l_max
    Y = _real_sph_scipy(L, th, PH)
    off = np.concatenate([[0], np.cumsum(h.alpha_max)])
    c = np.zeros((h.n_r, (L + 1) ** 2))
    c[off[si]:off[si + 1], 0] = (h.central_weight[si] * _radial_quad(h, si, 0.0, h.atom_sigma_r[si])
                                 / math.sqrt(4 * math.pi))
    for v, s in zip(x, sj):
        r = np.linalg.norm(v)
        kap = r * r / h.atom_sigma_t[s] ** 2
        f = np.exp(kap * (U @ (v / r) - 1.0))
        ang = np.einsum("kij,ij->k", Y, f * Wq) / np.sum(f * Wq)
        c[off[s]:off[s + 1]] += np.outer(_radial_quad(h, s, r, h.atom_sigma_r[s]), ang)
    return c


def _cnlm_cases():
    rng = np.random.default_rng(77)

    def dirs(k):
        d = rng.normal(size=(k, 3))
        return d / np.linalg.norm(d, axis=1)[:, None]

    cases = []
    for st, p in ((0.3, 0), (0.3, 2), (0.5, 1), (0.25, 0)):
        for rs in ((2.1,), (1.7,), (2.9,), (1.8, 2.6), (1.6, 2.2, 3.0)):
            x = dirs(len(rs)) * np.array(rs)[:, None]
            cases.append((st, p, 1, x, [0] * len(rs), "poly3"))
    # two species: neighbours of both, central species 1
    x = dirs(3) * np.array([1.9, 2.4, 2.8])[:, None]
    cases.append((0.3, 0, 2, x, [0, 1, 1], "poly3"))
    # the poly3gauss basis
    for st, p, rs in ((0.3, 0, (2.1,)), (0.3, 1, (1.8, 2.6)), (0.5, 2, (1.6, 2.2, 3.0))):
        x = dirs(len(rs)) * np.array(rs)[:, None]
        cases.append((st, p, 1, x, [0] * len(rs), "poly3gauss"))
    x = dirs(3) * np.array([1.9, 2.4, 2.8])[:, None]
    cases.append((0.3, 0, 2, x, [0, 1, 1], "poly3gauss"))
    return cases


@pytest.fixture(scope="module")
def sphere_grid():
    return _sphere_grid()


@pytest.mark.parametrize("case", _cnlm_cases())
def test_cnlm_vs_3d_quadrature(case, sphere_grid):
    """Expansion coefficients of 1-3 neighbours equal the 3-D quadrature of
    the smeared density up to the heat-kernel approximation of the angular
    profile: exp(-l(l+1)/(2 kappa)) against the exact e^-k i_l(k)/(e^-k
    i_0(k)), whose log differs by -l(l+1)/(4 kappa^2) + O(kappa^-3); the
    test allows 0.6 l(l+1)/kappa^2 (tabulated for kappa >= 10, l <= 8).
    Amplitude scaling off and every r_j inside rcut_soft, so the envelope
    is 1; sigma_r 0.4 so the radial Gaussian is not a delta."""
    st, p, nsp, x, sj, basis = case
    h = tgx.SoapHypers.uniform(nsp, alpha_max=(6 if nsp == 1 else 5), l_max=8, rcut_hard=4.0,
                           buffer=0.5, atom_sigma_r=0.4, atom_sigma_t=st, amplitude_scaling=0.0,
                           central_weight=0.7)
    h = tgx.SoapHypers(**{**h.__dict__, "radial_enhancement": p, "basis": basis})
    si = nsp - 1
    got = oracle.soap_coeffs(h, si, x, sj)
    ref = _cnlm_by_quadrature(h, si, x, sj, sphere_grid)
    # bound: per neighbour |R_n Y_lm A_l| * 0.6 l(l+1)/kappa^2
    th, PH, U, Wq = sphere_grid
    off = np.concatenate([[0], np.cumsum(h.alpha_max)])
    bound = np.zeros_like(got)
    for v, s in zip(x, sj):
        r = np.linalg.norm(v)
        kap = r * r / st ** 2
        u = v / r
        Yj = _real_sph_scipy(h.l_max, np.arccos(u[2]), math.atan2(u[1], u[0]))
        Rj = _radial_quad(h, s, r, h.atom_sigma_r[s])
        for l in range(h.l_max + 1):
            A = math.exp(-l * (l + 1) / (2 * kap))
            tol = 0.6 * l * (l + 1) / kap ** 2
            sl = slice(l * l, (l + 1) ** 2)
            bound[off[s]:off[s + 1], sl] += np.abs(np.outer(Rj, Yj[sl])) * A * tol
    floor = 1e-9 * np.abs(ref).max()
    assert np.all(np.abs(got - ref) <= bound + floor)
    # the l = 0 column carries no angular approximation: tight
    assert np.abs(got[:, 0] - ref[:, 0]).max() <= floor


def test_cnlm_quadrature_rejects_planted_mistakes(sphere_grid):
    """The bound above fails for plausible mistakes in the angular factor
    or the harmonic convention, applied to the oracle's coefficients on the
    CPU: l^2 for l(l+1), sigma_t instead of sigma_t/r, Y_l,m and Y_l,-m
    swapped."""
    st, r = 0.3, 2.2
    h = tgx.SoapHypers.uniform(1, alpha_max=6, l_max=8, rcut_hard=4.0, buffer=0.5, atom_sigma_r=0.4,
                           atom_sigma_t=st, amplitude_scaling=0.0, central_weight=0.7)
    u = np.array([0.3, -0.5, 0.81])
    u /= np.linalg.norm(u)
    x = (r * u)[None, :]
    got = oracle.soap_coeffs(h, 0, x, [0])
    ref = _cnlm_by_quadrature(h, 0, x, [0], sphere_grid)
    kap = r * r / st ** 2
    Rj = _radial_quad(h, 0, r, 0.4)
    Yj = _real_sph_scipy(8, math.acos(u[2]), math.atan2(u[1], u[0]))
    cen = np.zeros_like(got)
    cen[:, 0] = got[:, 0] - Rj * Yj[0]

    def with_A(Afun, Y=Yj):
        c = cen.copy()
        for l in range(9):
            sl = slice(l * l, (l + 1) ** 2)
            c[:, sl] += np.outer(Rj, Y[sl]) * Afun(l)
        return c

    def within(c):
        b = np.zeros_like(c)
        for l in range(9):
            sl = slice(l * l, (l + 1) ** 2)
            b[:, sl] = np.abs(np.outer(Rj, Yj[sl])) * math.exp(-l * (l + 1) / (2 * kap)) * (
                0.6 * l * (l + 1) / kap ** 2)
        return np.all(np.abs(c - ref) <= b + 1e-9 * np.abs(ref).max())

    # the oracle's own factor, rebuilt from its pieces, passes
    assert np.abs(with_A(lambda l: math.exp(-l * (l + 1) * st * st / (2 * r * r))) - got).max() < 1e-12
    assert within(got)
    assert not within(with_A(lambda l: math.exp(-l * l * st * st / (2 * r * r))))
    assert not within(with_A(lambda l: math.exp(-l * (l + 1) * st * st / 2)))
    Ysw = Yj.copy()
    for l in range(1, 9):
        for m in range(1, l + 1):
            Ysw[l * l + l + m], Ysw[l * l + l - m] = Yj[l * l + l - m], Yj[l * l + l + m]
    assert not within(with_A(lambda l: math.exp(-l * (l + 1) * st * st / (2 * r * r)), Ysw))


@pytest.mark.parametrize("rot", [None, 4])
def test_general_cell_energy_forces_same_crystal(rot):
    """The oracle's energies, forces and virial on the reference's
    general-cell list equal those of the orthorhombic path for the same
    crystal in another basis (a, a + b, c - 2b), rotated or not."""
    pos, sp, lat, model = W.carbon216()
    L = np.asarray(lat).reshape(3, 3)
    M = np.array([L[0], L[1] + L[0], L[2] - 2 * L[1]])
    A = np.eye(3)
    if rot is not None:
        A = np.linalg.qr(np.random.default_rng(rot).normal(size=(3, 3)))[0]
    e1, f1, v1 = oracle.energy_forces(model, pos, sp, lat)[:3]
    e2, f2, v2 = oracle.energy_forces(model, pos @ A.T, sp, (M @ A.T).reshape(9))[:3]
    assert oracle.neighbors(pos @ A.T, (M @ A.T).reshape(9), 4.2)["indices"].prod() > 1
    assert np.abs(e1 - e2).max() < 1e-10
    assert np.abs(f1 @ A.T - f2).max() < 1e-9
    assert np.abs(A @ v1 @ A.T - v2).max() < 1e-8


# --------------------------------------------------------------------------
# core_pot and angle_3b terms
# --------------------------------------------------------------------------
def _core_table(n=120, lo=0.3, hi=2.1):
    """A screened-Coulomb-like repulsion tabulated on a non-uniform grid."""
    x = lo + (hi - lo) * np.linspace(0, 1, n) ** 1.3
    V = 30.0 * np.exp(-x / 0.35) / x
    dV0 = -30.0 * np.exp(-lo / 0.35) * (1 / (0.35 * lo) + 1 / lo ** 2)
    return x, V, dV0


@pytest.mark.parametrize("ends", ["clamped", "natural", "mixed"])
def test_core_pot_spline_vs_scipy(ends):
    """orc_core_pot_setup + spline/spline_der equal scipy's cubic spline with
    the same end conditions; below x(1) the continuation is linear with
    slope yp1; natural ends get the spline's own end slope."""
    from scipy.interpolate import CubicSpline
    x, V, dV0 = _core_table()
    yp1 = dV0 if ends != "natural" else 1e31
    ypn = 0.0 if ends == "clamped" else 1e31
    xs, Vs, d2, y1, yn = oracle.core_pot_setup(x, V, yp1, ypn)
    bc = ((1, yp1) if yp1 < 1e30 else (2, 0.0), (1, ypn) if ypn < 1e30 else (2, 0.0))
    cs = CubicSpline(x, V, bc_type=bc)
    assert np.abs(d2 - cs(x, 2)).max() < 1e-7 * np.abs(d2).max()
    r = np.random.default_rng(0).uniform(x[0], x[-1], 200)
    s = oracle.spline(x, V, d2, y1, yn, r, x[-1] + 1.0)
    ds = oracle.spline(x, V, d2, y1, yn, r, x[-1] + 1.0, derivative=True)
    assert np.abs(s - cs(r)).max() < 1e-10 * np.abs(V).max()
    assert np.abs(ds - cs(r, 1)).max() < 1e-8 * np.abs(cs(r, 1)).max()
    assert abs(y1 - cs(x[0], 1)) < 1e-8 * abs(y1) and abs(yn - cs(x[-1], 1)) < 1e-8 * max(1, abs(yn))
    lo = x[0] - np.array([0.05, 0.1])
    assert np.allclose(oracle.spline(x, V, d2, y1, yn, lo, 5.0), V[0] + (lo - x[0]) * y1,
                       rtol=1e-14)
    assert np.all(oracle.spline(x, V, d2, y1, yn, [x[-1], x[-1] + 0.1], x[-1]) == 0.0)


def _terms_model(kernel="exp", sp=(1, 1, 1), seed=3):
    """A model of angle_3b + core_pot (no SOAP) for carbon cells: sparse
    points spread over the descriptor range of neighbour pairs within
    2.6 A."""
    rng = np.random.default_rng(seed)
    S3 = 40
    q1 = rng.uniform(2.4, 5.0, S3)
    q2 = rng.uniform(0.0, 1.0, S3)
    q3 = rng.uniform(1.2, 4.5, S3)
    sig = (0.5, 0.4, 0.5) if kernel == "exp" else (2.0, 1.2, 2.0)
    a3 = tgx.Angle3b(sp0=sp[0], sp1=sp[1], sp2=sp[2], rcut=2.6, delta=0.8, sigma=sig,
                     Qs=np.stack([q1, q2, q3], 1), alphas=rng.normal(0, 0.05, S3),
                     cutoff=rng.uniform(0.5, 1.0, S3), kernel=kernel)
    x, V, dV0 = _core_table()
    cp = tgx.CorePot(sp1=1, sp2=1, x=x, V=V, yp1=dV0, ypn=1e31)
    h = S.carbon_hypers()
    return tgx.GapModel(hypers=h, Qs=np.zeros((1, h.dim)), alphas=np.zeros(1),
                        angle_3b=[a3], core_pot=[cp])


def _ef_terms(model, pos, sp, lat, do_forces=True):
    return oracle.energy_forces(model, pos, sp, lat, do_forces=do_forces, soap=False)[:3]


def test_core_pot_dimer_closed_form():
    x, V, dV0 = _core_table()
    m = _terms_model()
    m.angle_3b = []
    lat = S.cubic_lattice(30.0)
    for r in (0.2, 0.75, 1.33, 2.0, 2.2):
        pos = np.array([[5.0, 5.0, 5.0], [5.0 + r * 0.6, 5.0 + r * 0.8, 5.0]])
        e, f, v = _ef_terms(m, pos, np.ones(2, np.int32), lat)
        xs, Vs, d2, y1, yn = oracle.core_pot_setup(x, V, dV0, 1e31)
        Vr = oracle.spline(x, V, d2, y1, yn, [r], x[-1])[0]
        dVr = oracle.spline(x, V, d2, y1, yn, [r], x[-1], derivative=True)[0]
        assert np.allclose(e, [0.5 * Vr, 0.5 * Vr], rtol=1e-14, atol=0)
        u = np.array([0.6, 0.8, 0.0])
        assert np.abs(f[0] - dVr * u).max() <= 1e-12 * max(1, abs(dVr))
        assert np.abs(f[0] + f[1]).max() <= 1e-12 * max(1, abs(dVr))
        # pair virial: -r (x) F_j summed = -dV/dr r u u^T
        assert np.abs(v - (-dVr * r * np.outer(u, u))).max() <= 1e-12 * max(1, abs(dVr * r))


def _trimer_energy(a3, p0, p1, p2):
    """Independent numpy form of one angle_3b centre with two neighbours."""
    x12, x13 = p1 - p0, p2 - p0
    r12, r13, r23 = np.linalg.norm(x12), np.linalg.norm(x13), np.linalg.norm(p2 - p1)
    if not (r12 < a3.rcut and r13 < a3.rcut):
        return 0.0
    fc = 1.0
    for r in (r12, r13):
        if r > a3.rcut - 0.5:
            fc *= 0.5 * (1 + math.cos(math.pi * (r - a3.rcut + 0.5) / 0.5))
    q = np.array([r12 + r13, (r12 - r13) ** 2, r23])
    d = np.sqrt((((q - a3.Qs) / np.asarray(a3.sigma)) ** 2).sum(1))
    if a3.kernel == "exp":
        k = np.exp(-0.5 * d * d)
    else:
        k = np.where(d < 1, (1 - np.minimum(d, 1)) ** 4 * (4 * d + 1), 0.0)
    return fc * np.sum(2 * a3.delta ** 2 * a3.alphas * a3.cutoff * k)


@pytest.mark.parametrize("kernel", ["exp", "pol"])
def test_angle_3b_trimer_closed_form(kernel):
    m = _terms_model(kernel)
    m.core_pot = []
    a3 = m.angle_3b[0]
    lat = S.cubic_lattice(30.0)
    rng = np.random.default_rng(8)
    for _ in range(6):
        p = np.array([10.0, 10.0, 10.0]) + rng.uniform(-1.3, 1.3, (3, 3))
        e, f, v = _ef_terms(m, p, np.ones(3, np.int32), lat)
        ref = [_trimer_energy(a3, p[0], p[1], p[2]), _trimer_energy(a3, p[1], p[0], p[2]),
               _trimer_energy(a3, p[2], p[0], p[1])]
        assert np.allclose(e, ref, rtol=1e-12, atol=1e-15)
        assert np.abs(f.sum(0)).max() < 1e-12


@pytest.mark.parametrize("kernel", ["exp", "pol"])
def test_angle_3b_core_pot_forces_and_virial_fd(kernel):
    """Forces = -dE/dr and virial = -dE/d(strain), all nine components (the
    strained cells are triclinic), for angle_3b + core_pot together."""
    m = _terms_model(kernel)
    pos, sp, lat = S.carbon_cell(48, 2.2, 1.05, 41)
    e, f, v = _ef_terms(m, pos, sp, lat)
    assert abs(e.sum()) > 1e-3
    h = 1e-5
    rng = np.random.default_rng(1)
    for i in rng.choice(len(pos), 6, replace=False):
        for a in range(3):
            Es = []
            for s in (1, -1):
                p = pos.copy()
                p[i, a] += s * h
                Es.append(_ef_terms(m, p, sp, lat, False)[0].sum())
            fd = -(Es[0] - Es[1]) / (2 * h)
            assert abs(fd - f[i, a]) < 1e-6 * max(1.0, abs(f[i, a])), (i, a, fd, f[i, a])
    for a in range(3):
        for b in range(3):
            Es = []
            for s in (1, -1):
                eps = np.zeros((3, 3))
                eps[a, b] = s * h
                p, l = _strained(pos, lat, eps)
                Es.append(_ef_terms(m, p, sp, l, False)[0].sum())
            fd = -(Es[0] - Es[1]) / (2 * h)
            assert abs(fd - v[a, b]) < 1e-6 * max(1.0, abs(v[a, b])), (a, b, fd, v[a, b])


def test_angle_3b_species_filter_and_invariance():
    """Centre / neighbour species filters (a C-centred C-H block on a C/H
    cell) and rotation + permutation invariance."""
    pos, sp, lat = S.cho_cell(80, 1.5, 1.0, 5)
    sp = np.where(sp == 3, 1, sp).astype(np.int32)  # C/H only
    m = _terms_model("exp", sp=(1, 1, 2))
    m.core_pot = []
    e, f, v = _ef_terms(m, pos, sp, lat)
    assert np.all(e[sp != 1] == 0.0) and np.abs(e).max() > 0
    rng = np.random.default_rng(3)
    A = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    perm = rng.permutation(len(pos))
    e2, f2, v2 = _ef_terms(m, pos[perm] @ A.T, sp[perm], (lat.reshape(3, 3) @ A.T).reshape(9))
    assert np.abs(e2 - e[perm]).max() < 1e-12
    assert np.abs(f2 - f[perm] @ A.T).max() < 1e-11
    assert np.abs(v2 - A @ v @ A.T).max() < 1e-10


# --------------------------------------------------------------------------
# local-property models
# --------------------------------------------------------------------------
def _lp_model(h, q, seed=5, n_sparse=60):
    rng = np.random.default_rng(seed)
    Qs = q[rng.choice(len(q), n_sparse, replace=len(q) < n_sparse)]
    return Qs, rng.normal(0.0, 0.5, n_sparse)


def test_local_property_values_and_derivatives():
    """V equals the C kernel step's site energies plus V0, clipped at 0;
    dV_i/dr_j summed over a unit atom's entries equals the central finite
    difference of V_i in that atom's position (images included: small cell);
    entries of clipped sites are zero."""
    pos, sp, lat = S.carbon_cell(24, 2.6, 1.2, 61)  # 8.1 A box: supercell images
    h = S.carbon_hypers()
    q = oracle.soap(h, pos, sp, lat)["soap"]
    Qs, al = _lp_model(h, q)
    e, _ = oracle.gap_kernel_step(q, Qs, al, 0.9, 3.0)
    V0 = -float(np.median(e))  # about half the sites clip
    V, der = oracle.local_property(h, Qs, al, 0.9, 3.0, V0, pos, sp, lat, 4.2, derivatives=True)
    assert np.allclose(V, np.maximum(e + V0, 0.0), rtol=0, atol=1e-13)
    assert 0 < np.sum(V == 0.0) < len(V)
    full = oracle.neighbors(pos, lat, 4.2)
    off = np.concatenate([[0], np.cumsum(full["n_neigh"])])
    hh = 1e-5
    for i in range(len(pos)):
        rows = der[off[i]:off[i + 1]]
        if V[i] == 0.0:
            assert np.all(rows == 0.0)
            continue
        if i % 3:
            continue
        units = (full["list"][off[i]:off[i + 1]] - 1) % len(pos)
        for j in sorted(set(units.tolist()))[:4]:
            an = rows[units == j].sum(0)
            for a in range(3):
                vs = []
                for sgn in (1, -1):
                    p = pos.copy()
                    p[j, a] += sgn * hh
                    vs.append(oracle.local_property(h, Qs, al, 0.9, 3.0, V0, p, sp, lat)[0][i])
                fd = (vs[0] - vs[1]) / (2 * hh)
                assert abs(fd - an[a]) < 1e-6 * max(1.0, abs(an[a])), (i, j, a, fd, an[a])


# --------------------------------------------------------------------------
# Tkatchenko-Scheffler dispersion
# --------------------------------------------------------------------------
HARTREE, BOHR = 27.211386024367243, 0.5291772105638411
# free-atom references (C6 eV A^6, R0 A, alpha0 A^3): the reference's
# defaults for C and H (get_vdw_ref_params); O from the TS tables
TS_REF = np.array([[46.6 * HARTREE * BOHR ** 6, 1.900, 12.0 * BOHR ** 3],
                   [6.5 * HARTREE * BOHR ** 6, 1.64, 4.5 * BOHR ** 3],
                   [15.6 * HARTREE * BOHR ** 6, 1.71, 5.4 * BOHR ** 3]])


def _ts_pair_energy(r, sa, sb, Va, Vb, sR=0.94, d=20.0, mpm=False):
    """Total TS energy of one pair (Tkatchenko & Scheffler, PRL 102, 073005
    (2009)): -f_damp(r) C6_AB / r^6 with the combination rule C6_AB =
    2 C6_A C6_B / (a_B/a_A C6_A + a_A/a_B C6_B) and R0_AB = R0_A + R0_B, the
    free-atom references scaled by V^2 (C6), V (alpha) and V^(1/3) (R0)."""
    ex = mp.exp if mpm else math.exp
    c6a, r0a, aa = TS_REF[sa - 1]
    c6b, r0b, ab = TS_REF[sb - 1]
    C6A, C6B = c6a * Va ** 2, c6b * Vb ** 2
    aA, aB = aa * Va, ab * Vb
    C6 = 2 * C6A * C6B / (aB / aA * C6A + aA / aB * C6B)
    R0 = r0a * Va ** (1.0 / 3.0) + r0b * Vb ** (1.0 / 3.0)
    return -C6 / (1 + ex(-d * (r / (sR * R0) - 1))) / r ** 6, C6


def _ts(pos, sp, lat, V, der=None, do_forces=True, **kw):
    return oracle.ts_vdw(TS_REF[:, 0], TS_REF[:, 1], TS_REF[:, 2], pos, sp, lat, V, der,
                         do_forces=do_forces, **kw)


@pytest.mark.parametrize("sa,sb,r,Va,Vb", [(1, 1, 3.0, 0.8, 0.8), (1, 2, 4.5, 0.7, 1.1),
                                           (2, 3, 7.3, 0.9, 0.6), (1, 1, 1.9, 1.2, 0.5)])
def test_ts_dimer_closed_form(sa, sb, r, Va, Vb):
    """Energy of a dimer outside the switching buffers equals the published
    TS pair energy, half on each atom; with fixed volumes the force on A is
    E'(r) along the bond (mpmath derivative of the closed form)."""
    rng = np.random.default_rng(int(100 * r))
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    pos = np.array([[20.0, 20.0, 20.0], [20.0, 20.0, 20.0] + r * u])
    sp = np.array([sa, sb], np.int32)
    lat = np.diag([40.0] * 3).reshape(9)
    V = np.array([Va, Vb])
    e, f, v = _ts(pos, sp, lat, V)
    E, _ = _ts_pair_energy(r, sa, sb, Va, Vb)
    assert np.allclose(e, [E / 2, E / 2], rtol=1e-13, atol=0)
    mp.mp.dps = 30
    dE = float(mp.diff(lambda t: _ts_pair_energy(t, sa, sb, Va, Vb, mpm=True)[0], mp.mpf(r)))
    assert np.allclose(f[0], dE * u, rtol=1e-10, atol=1e-14)
    assert np.allclose(f[1], -dE * u, rtol=1e-10, atol=1e-14)
    # virial of a pair force: -F_A (x) d
    assert np.allclose(v, -np.outer(f[0], r * u) * 1.0, rtol=1e-10, atol=1e-14)


def test_ts_buffers_switch_and_reference_slope():
    """In the buffers the energy is the pair energy times 1 - 3x^2 + 2x^3
    (x from 0 at the buffer's inner edge to 1 at the cutoff); the force uses
    the reference's slope, the exact derivative plus 6 C6 f_damp (1 - sw)
    / r^7 along the bond (vdw_interface.f90:199-211 differentiates the
    unswitched r^-6)."""
    lat = np.diag([40.0] * 3).reshape(9)
    sp = np.array([1, 2], np.int32)
    V = np.array([0.9, 1.05])
    kw = dict(rcut=10.0, buffer=1.0, rcut_inner=0.5, buffer_inner=0.5)

    def sw(x):
        return 1 - 3 * x ** 2 + 2 * x ** 3

    def exact(t, mpm=False):
        E, C6 = _ts_pair_energy(t, 1, 2, V[0], V[1], mpm=mpm)
        x = (t - 9.0) / 1.0 if t > 5 else (1.0 - t) / 0.5
        return E * sw(x), E, C6, sw(x)

    mp.mp.dps = 30
    for r in (9.4, 9.97, 0.8, 0.55):
        pos = np.array([[20.0, 20.0, 20.0], [20.0 + r, 20.0, 20.0]])
        e, f, _ = _ts(pos, sp, lat, V, **kw)
        Es, E, C6, s = exact(r)
        assert abs(e.sum() - Es) < 1e-13 * abs(E)
        dEs = float(mp.diff(lambda t: exact(t, mpm=True)[0], mp.mpf(r)))
        fd = 1.0 / (1.0 + math.exp(-20.0 * (r / (0.94 * (TS_REF[0, 1] * V[0] ** (1 / 3) +
                                                         TS_REF[1, 1] * V[1] ** (1 / 3))) - 1)))
        slope = dEs + 6.0 * C6 * fd * (1.0 - s) / r ** 7
        assert abs(f[0, 0] - slope) < 1e-9 * max(abs(slope), abs(dEs)), (r, f[0, 0], slope, dEs)
        assert abs(slope - dEs) > 1e-6 * abs(dEs)  # the convention is visible here
    # beyond the cutoff and inside rcut_inner: nothing
    for r in (10.0, 10.3, 0.5, 0.3):
        pos = np.array([[20.0, 20.0, 20.0], [20.0 + r, 20.0, 20.0]])
        e, f, v = _ts(pos, sp, lat, V, **kw)
        assert np.all(e == 0) and np.all(f == 0) and np.all(v == 0)


def test_ts_fixed_volumes_forces_and_virial_fd():
    """Fixed volumes (no Hirshfeld term), no buffers: forces = -dE/dr and
    virial = -dE/d(strain) in all nine components on a C/H/O cell."""
    pos, sp, lat = S.cho_cell(40, 1.5, 1.0, 5)
    V = np.random.default_rng(2).uniform(0.5, 1.2, len(pos))
    kw = dict(rcut=5.0, buffer=0.0, rcut_inner=0.5, buffer_inner=0.0)
    e, f, v = _ts(pos, sp, lat, V, **kw)
    assert e.sum() < 0
    h = 1e-5
    for i in (0, 9, 27):
        for a in range(3):
            Es = []
            for s in (1, -1):
                p = pos.copy()
                p[i, a] += s * h
                Es.append(_ts(p, sp, lat, V, do_forces=False, **kw)[0].sum())
            fdv = -(Es[0] - Es[1]) / (2 * h)
            assert abs(fdv - f[i, a]) < 1e-7 * max(1.0, abs(f[i, a])), (i, a, fdv, f[i, a])
    for a in range(3):
        for b in range(3):
            Es = []
            for s in (1, -1):
                eps = np.zeros((3, 3))
                eps[a, b] = s * h
                p, l = _strained(pos, lat, eps)
                Es.append(_ts(p, sp, l, V, do_forces=False, **kw)[0].sum())
            fdv = -(Es[0] - Es[1]) / (2 * h)
            assert abs(fdv - v[a, b]) < 1e-7 * max(1.0, abs(v[a, b])), (a, b, fdv, v[a, b])


def test_ts_hirshfeld_term_fd():
    """Volumes from a local-property model (supercell images): the TS force
    plus the self entry's term the reference leaves out, (-dE/dV_a)
    dV_a/dr_a with dE/dV_a by finite differences of the energy in V, equals
    -dE/dr_a of the full energy E(V(r), r); the virial (the self entry's
    vector is zero) equals -dE/d(strain)."""
    pos, sp, lat = S.carbon_cell(16, 2.0, 1.2, 7)
    h = S.carbon_hypers()
    q = oracle.soap(h, pos, sp, lat)["soap"]
    Qs, al = _lp_model(h, q, n_sparse=16)
    al = 0.05 * al
    rl = 4.5
    kw = dict(rcut=rl, buffer=0.0, rcut_inner=0.5, buffer_inner=0.0)

    def vol(p, l):
        return oracle.local_property(h, Qs, al, 0.9, 3.0, 1.0, p, sp, l)[0]

    def energy(p, l):
        return _ts(p, sp, l, vol(p, l), do_forces=False, **kw)[0].sum()

    V, der = oracle.local_property(h, Qs, al, 0.9, 3.0, 1.0, pos, sp, lat, rl, derivatives=True)
    assert np.all(V > 0.5)
    e, f, v = _ts(pos, sp, lat, V, der, **kw)
    off = np.concatenate([[0], np.cumsum(oracle.neighbors(pos, lat, rl)["n_neigh"])])
    hh = 1e-5
    for a_ in (0, 5, 11):
        Es = []
        for s in (1, -1):
            Vp = V.copy()
            Vp[a_] += s * hh
            Es.append(_ts(pos, sp, lat, Vp, do_forces=False, **kw)[0].sum())
        mdEdV = -(Es[0] - Es[1]) / (2 * hh)
        self_term = mdEdV * der[off[a_]]
        for a in range(3):
            Es = []
            for s in (1, -1):
                p = pos.copy()
                p[a_, a] += s * hh
                Es.append(energy(p, lat))
            fdv = -(Es[0] - Es[1]) / (2 * hh)
            assert abs(fdv - (f[a_, a] + self_term[a])) < 1e-7 * max(1.0, abs(fdv)), \
                (a_, a, fdv, f[a_, a], self_term[a])
        assert np.abs(self_term).max() > 1e-4
    for a in range(3):
        for b in range(3):
            Es = []
            for s in (1, -1):
                eps = np.zeros((3, 3))
                eps[a, b] = s * hh
                p, l = _strained(pos, lat, eps)
                Es.append(energy(p, l))
            fdv = -(Es[0] - Es[1]) / (2 * hh)
            assert abs(fdv - v[a, b]) < 1e-7 * max(1.0, abs(v[a, b])), (a, b, fdv, v[a, b])


# --------------------------------------------------------------------------
# compress_soap
# --------------------------------------------------------------------------
def _compressed(h, P, dim=0):
    import dataclasses
    return dataclasses.replace(h, compress_P=P, compress_dim=dim)


def _random_P(h, dim, seed, per_row=3):
    rng = np.random.default_rng(seed)
    pi, pj, pe = [], [], []
    for i in range(1, dim + 1):
        for j in rng.choice(h.full_dim, per_row, replace=False):
            pi.append(i)
            pj.append(j + 1)
            pe.append(rng.normal())
    return np.array(pi), np.array(pj), np.array(pe)


def test_compress_identity_and_selection():
    """The identity transformation gives the uncompressed descriptors and
    derivatives bit for bit; a selection of components (the old file format:
    one kept index per line, weight 1) gives q[sel] / |q[sel]|."""
    pos, sp, lat = S.carbon_cell(20, 2.0, 1.2, 3)
    h = S.carbon_hypers()
    ref = oracle.soap(h, pos, sp, lat, derivatives=True)
    d = h.full_dim
    ident = _compressed(h, (np.arange(1, d + 1), np.arange(1, d + 1), np.ones(d)))
    got = oracle.soap(ident, pos, sp, lat, derivatives=True)
    assert np.array_equal(got["soap"], ref["soap"])
    assert np.array_equal(got["soap_der"], ref["soap_der"])
    sel = np.sort(np.random.default_rng(4).choice(d, 100, replace=False))
    hs = _compressed(h, (np.arange(1, 101), sel + 1, np.ones(100)))
    got = oracle.soap(hs, pos, sp, lat)
    want = ref["soap"][:, sel]
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.abs(got["soap"] - want).max() < 1e-14


def test_compress_general_values_derivatives_and_forces():
    """A general sparse transformation (3 weighted entries per row):
    q = P p / |P p| (numpy on the uncompressed descriptors: P p is P q times
    |p|), dq/dr by finite differences, and forces = -dE/dr, virial =
    -dE/d(strain) for a GAP model on the compressed descriptors."""
    pos, sp, lat = S.carbon_cell(24, 2.0, 1.2, 9)
    h = S.carbon_hypers()
    P = _random_P(h, 90, 2)
    hc = _compressed(h, P, 90)
    full = oracle.soap(h, pos, sp, lat)["soap"]
    M = np.zeros((90, h.full_dim))
    np.add.at(M, (P[0] - 1, P[1] - 1), P[2])
    want = full @ M.T
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    got = oracle.soap(hc, pos, sp, lat, derivatives=True)
    assert got["soap"].shape == (24, 90)
    assert np.abs(got["soap"] - want).max() < 1e-13
    # dq_i/dr_j of one pair by central differences
    k = 5
    i, j = int(np.repeat(np.arange(24), got["site_n_neigh"])[k]), int(got["pair_atom"][k])
    assert i != j
    hh = 1e-5
    for a in range(3):
        qs = []
        for s in (1, -1):
            p = pos.copy()
            p[j, a] += s * hh
            qs.append(oracle.soap(hc, p, sp, lat)["soap"][i])
        fd = (qs[0] - qs[1]) / (2 * hh)
        # every entry of row i pointing at j (images included)
        rows = np.arange(np.sum(got["site_n_neigh"][:i]), np.sum(got["site_n_neigh"][:i + 1]))
        an = got["soap_der"][rows[got["pair_atom"][rows] == j], :, a].sum(0)
        assert np.abs(fd - an).max() < 1e-7, a
    model = S.model_from_soaps(hc, got["soap"], 30, seed=3)
    e, f, v = oracle.energy_forces(model, pos, sp, lat)
    for at in (0, 11):
        for a in range(3):
            Es = []
            for s in (1, -1):
                p = pos.copy()
                p[at, a] += s * hh
                Es.append(oracle.energy_forces(model, p, sp, lat, do_forces=False)[0].sum())
            fd = -(Es[0] - Es[1]) / (2 * hh)
            assert abs(fd - f[at, a]) < 1e-7 * max(1.0, abs(f[at, a])), (at, a, fd, f[at, a])
    for a in range(3):
        def strained_energy(t):
            eps = np.zeros((3, 3))
            eps[a, a] = t
            p, l = _strained(pos, lat, eps)
            return oracle.energy_forces(model, p, sp, l, do_forces=False)[0].sum()
        fd = -_fd5(strained_energy, 0.0, 1e-4)
        assert abs(fd - v[a, a]) < 5e-8 * max(1.0, abs(v[a, a])), (a, fd, v[a, a])
