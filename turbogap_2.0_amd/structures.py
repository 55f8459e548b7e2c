"""Seeded synthetic structures and GAP models for the workloads this engine
is measured and tested on.

All structures are random periodic cubic cells built by sequential random
insertion with a minimum-distance rejection, at a given mass density.  The
workloads (sizes, densities, species mix, SOAP settings) are:

  carbon216   216 C, 2.0 g/cm^3, min 1.2 A; SOAP rcut 3.7 (buffer 0.5),
              n_max 8, l_max 8, sigma 0.5; zeta 4, delta 1, 500 sparse
              vectors drawn from the structure's own SOAPs, alphas ~ N(0, 0.01)
              plus a C-C 2-body block.
  cho4096     4,096 atoms C/H/O 60/30/10 %, 1.5 g/cm^3, min 1.0 A; rcut 4.0,
              n_max 8 per species, l_max 8; 4,000 sparse, zeta 4.
  batch256    256 carbon cells, 32-1,024 atoms (log-uniform), 1.5-3.2 g/cm^3;
              carbon216 SOAP settings, 2,000 sparse.
  carbon32k   32,768 C, 2.0 g/cm^3, min 1.2 A, Maxwell-Boltzmann velocities
              at 3,000 K; carbon216 model (NVE MD, dt 1 fs, skin 0.5 A).
  carbon512   512 C, 2.0 g/cm^3 (Metropolis MC at 1,500 K).
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np

from tgx import AMU, GapModel, SoapHypers, TwoBody

N_A = 6.02214076e23
MASS = {"C": 12.011, "H": 1.008, "O": 15.999}
KB_EV = 8.617333262e-5


def box_length(n: int, mean_mass: float, density_gcc: float) -> float:
    """Cubic box edge (A) for n atoms of mean mass (amu) at density (g/cm^3)."""
    vol_a3 = n * mean_mass / (density_gcc * N_A) * 1e24
    return vol_a3 ** (1.0 / 3.0)


def random_cell(n: int, L: float, min_dist: float, seed: int,
                max_tries: int = 2000) -> np.ndarray:
    """n positions in [0, L)^3 with all minimum-image distances >= min_dist,
    by sequential uniform insertion with rejection (cell grid for speed)."""
    rng = np.random.default_rng(seed)
    m = max(1, int(L / min_dist))
    h = L / m
    grid: dict = {}
    pos = np.zeros((n, 3))
    for i in range(n):
        for _ in range(max_tries):
            p = rng.random(3) * L
            c = np.floor(p / h).astype(int) % m
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        key = ((c[0] + dx) % m, (c[1] + dy) % m, (c[2] + dz) % m)
                        lst = grid.get(key)
                        if not lst:
                            continue
                        d = pos[lst] - p
                        d -= L * np.round(d / L)
                        if np.any(np.einsum("ij,ij->i", d, d) < min_dist * min_dist):
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[i] = p
                grid.setdefault((c[0], c[1], c[2]), []).append(i)
                break
        else:
            raise RuntimeError("random_cell: could not insert atom %d" % i)
    return pos


def triclinic_cell(n: int, density: float, min_dist: float, seed: int,
                   skew=(0.3, -0.2, 0.25), rotate: bool = False, mass: float = MASS["C"]):
    """n atoms in a non-orthorhombic periodic cell of the given density:
    lattice a = (L,0,0), b = (s0 L, L, 0), c = (s1 L, s2 L, L) (volume L^3),
    optionally rotated by a random rotation; atoms at uniform fractional
    coordinates with every periodic distance >= min_dist (exhaustive image
    search; for small cells).  Returns pos, species (all 1), lattice (9)."""
    rng = np.random.default_rng(seed)
    L = box_length(n, mass, density)
    M = np.array([[L, 0, 0], [skew[0] * L, L, 0], [skew[1] * L, skew[2] * L, L]])
    if rotate:
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        M = M @ q.T
    shifts = np.array([[a, b, c] for a in range(-2, 3) for b in range(-2, 3)
                       for c in range(-2, 3)], float) @ M
    pos = np.zeros((n, 3))
    for i in range(n):
        for _ in range(20000):
            p = rng.random(3) @ M
            if i == 0:
                break
            d = (pos[:i, None, :] + shifts[None, :, :]) - p
            if np.min(np.einsum("ijk,ijk->ij", d, d)) >= min_dist * min_dist:
                break
        else:
            raise RuntimeError("triclinic_cell: could not insert atom %d" % i)
        pos[i] = p
    return pos, np.ones(n, np.int32), M.reshape(9).copy()


def cubic_lattice(L: float) -> np.ndarray:
    return np.array([L, 0, 0, 0, L, 0, 0, 0, L], dtype=np.float64)


def carbon_cell(n: int, density: float, min_dist: float, seed: int):
    L = box_length(n, MASS["C"], density)
    pos = random_cell(n, L, min_dist, seed)
    return pos, np.ones(n, np.int32), cubic_lattice(L)


def cho_cell(n: int, density: float, min_dist: float, seed: int,
             fractions=(0.6, 0.3, 0.1)):
    counts = [int(round(f * n)) for f in fractions]
    counts[0] += n - sum(counts)
    rng = np.random.default_rng(seed)
    species = np.concatenate([np.full(c, s + 1, np.int32) for s, c in enumerate(counts)])
    rng.shuffle(species)
    masses = np.array([MASS["C"], MASS["H"], MASS["O"]])
    mean_mass = float(np.mean(masses[species - 1]))
    L = box_length(n, mean_mass, density)
    pos = random_cell(n, L, min_dist, seed + 7)
    return pos, species, cubic_lattice(L)


def carbon_hypers(rcut: float = 3.7) -> SoapHypers:
    return SoapHypers.uniform(1, alpha_max=8, l_max=8, rcut_hard=rcut, buffer=0.5,
                              atom_sigma_r=0.5, atom_sigma_t=0.5)


def cho_hypers() -> SoapHypers:
    return SoapHypers.uniform(3, alpha_max=8, l_max=8, rcut_hard=4.0, buffer=0.5,
                              atom_sigma_r=0.5, atom_sigma_t=0.5)


def carbon_2b(seed: int, n_sparse: int = 50, rcut: float = 3.7) -> TwoBody:
    """C-C distance_2b block: rcut 3.7, buffer 0.5, sigma 0.5, delta 1,
    sparse distances uniform in [1.0, rcut], cutoff 1, alphas ~ N(0, 0.01)."""
    rng = np.random.default_rng(seed)
    return TwoBody(sp1=1, sp2=1, rcut=rcut, delta=1.0, sigma=0.5,
                   Qs=np.sort(rng.uniform(1.0, rcut, n_sparse)),
                   alphas=rng.normal(0.0, 0.01, n_sparse),
                   cutoff=np.ones(n_sparse))


def model_from_soaps(hypers: SoapHypers, soaps: np.ndarray, n_sparse: int, seed: int,
                     zeta: float = 4.0, delta: float = 1.0, replace=None,
                     two_body: Sequence[TwoBody] = ()) -> GapModel:
    """Sparse set drawn from the given (normalised) SOAP vectors; alphas ~
    N(0, 0.01).  Drawn with replacement when there are fewer vectors than
    n_sparse."""
    rng = np.random.default_rng(seed)
    if replace is None:
        replace = soaps.shape[0] < n_sparse
    idx = rng.choice(soaps.shape[0], size=n_sparse, replace=replace)
    Qs = np.ascontiguousarray(soaps[idx])
    alphas = rng.normal(0.0, 0.01, n_sparse)
    return GapModel(hypers=hypers, Qs=Qs, alphas=alphas, delta=delta, zeta=zeta,
                    e0=np.zeros(hypers.n_species), two_body=list(two_body))


def maxwell_boltzmann(n: int, masses_amu: np.ndarray, T: float, seed: int) -> np.ndarray:
    """Velocities (A/fs) ~ N(0, kB T/m), centre-of-mass motion removed, then
    rescaled to exactly T with 3(N-1) degrees of freedom
    (md/md_interface.f90:94, 111-118)."""
    rng = np.random.default_rng(seed)
    m = masses_amu * AMU
    v = rng.normal(size=(n, 3)) * np.sqrt(KB_EV * T / m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    ek = 0.5 * np.sum(m[:, None] * v * v)
    t_now = 2.0 * ek / (3.0 * (n - 1) * 8.6173303e-5)
    return v * math.sqrt(T / t_now)


def batch_cells(n_struct: int, seed: int, n_lo: int = 32, n_hi: int = 1024,
                rho_lo: float = 1.5, rho_hi: float = 3.2):
    rng = np.random.default_rng(seed)
    ns = np.rint(np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), n_struct))).astype(int)
    rhos = rng.uniform(rho_lo, rho_hi, n_struct)
    out = []
    for k, (n, rho) in enumerate(zip(ns, rhos)):
        out.append(carbon_cell(int(n), float(rho), 1.2, seed * 1000 + k))
    return out
