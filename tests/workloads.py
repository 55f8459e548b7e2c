"""Seeded workloads shared by the tests (small enough for the oracle)."""
import functools

import numpy as np

import oracle
import structures as S


@functools.lru_cache(maxsize=None)
def carbon216():
    """216 C at 2.0 g/cm^3, min 1.2 A, seed 1000, with its GAP model
    (500 sparse SOAPs drawn with replacement from its own descriptors,
    alphas ~ N(0, 0.01), zeta 4, plus the C-C 2-body block)."""
    pos, sp, lat = S.carbon_cell(216, 2.0, 1.2, 1000)
    h = S.carbon_hypers()
    q = oracle.soap(h, pos, sp, lat)["soap"]
    model = S.model_from_soaps(h, q, 500, seed=1000, two_body=[S.carbon_2b(1000)])
    return pos, sp, lat, model


@functools.lru_cache(maxsize=None)
def cho(n=300, seed=1001):
    """Small C/H/O 60/30/10 cell (1.5 g/cm^3, min 1.0 A, rcut 4.0, n_max 8
    per species) with a model from its own SOAPs."""
    pos, sp, lat = S.cho_cell(n, 1.5, 1.0, seed)
    h = S.cho_hypers()
    q = oracle.soap(h, pos, sp, lat)["soap"]
    model = S.model_from_soaps(h, q, min(200, n), seed=seed)
    return pos, sp, lat, model


@functools.lru_cache(maxsize=None)
def small_box(n=32, density=3.2, seed=1002):
    """A cell thinner than 2 rcut (supercell images needed)."""
    pos, sp, lat = S.carbon_cell(n, density, 1.2, seed)
    _, _, _, model = carbon216()
    return pos, sp, lat, model
