"""Seeded synthetic workloads: the ensemble of uneven lattices, the
full-elasticity lattice, and helpers shared by bench.py, the profiling runs,
the warm-start data pipeline and the tests.

Ensemble: lattices whose edge triples are drawn uniformly from the 65 triples
of {24, 32, 48, 64, 96}^3 with max/min <= 2, each with 0-3 spheres (radius 4-16,
capped at a third of the shortest edge, uniform random centres on the periodic
lattice), homeotropic or degenerate-planar anchoring with W log-uniform in
[0.5, 20], and its own director seed.
"""
import itertools
import math

import numpy as np

EDGES = (24, 32, 48, 64, 96)


def phase_constants(a=-0.172, b=-2.12, c=1.73):
    """openQmin's rescaled bulk constants (openQmin.cpp:173-175) and the
    uniaxial S0 they fix."""
    A, B, C = -1.0, -b / a, -c / a
    S0 = (-B + math.sqrt(B * B - 24 * A * C)) / (6 * C)
    return A, B, C, S0


def admissible_triples(edges=EDGES, max_aspect=2.0):
    return [t for t in itertools.product(edges, repeat=3) if max(t) / min(t) <= max_aspect]


def ensemble_specs(count=1024, seed=2024, edges=EDGES):
    """List of dicts {dims, seed, spheres: [(centre, radius, btype, W)]}."""
    rng = np.random.default_rng(seed)
    triples = admissible_triples(edges)
    specs = []
    for k in range(count):
        dims = triples[int(rng.integers(len(triples)))]
        spheres = []
        for _ in range(int(rng.integers(0, 4))):
            r = float(min(rng.uniform(4.0, 16.0), min(dims) / 3.0))
            centre = tuple(float(rng.uniform(0, d)) for d in dims)
            btype = int(rng.integers(0, 2))
            W = float(math.exp(rng.uniform(math.log(0.5), math.log(20.0))))
            spheres.append((centre, r, btype, W))
        specs.append({"dims": dims, "seed": int(rng.integers(1, 2 ** 31 - 1)), "spheres": spheres})
    return specs


def build_lattice(qmin, spec, precision, S0):
    lat = qmin.Lattice(*spec["dims"], precision=precision)
    lat.random_init(spec["seed"], S0)
    for centre, r, btype, W in spec["spheres"]:
        lat.sphere(centre, r, btype, W, S0)
    return lat


def full_elasticity(qmin, L, precision, seed=1234, geometry_seed=384):
    """The full-elasticity workload on an L^3 periodic lattice: L1..L4 and L6 =
    (2.32, 1.0, 0.5, 0.2, 0.3) (open-Qmin has no L5 term), a uniform E field
    (0, 0, 0.5) with delta eps = 1 (eps = eps0 = 1), and four homeotropic
    spheres of radius 12-24 (scaled by L/384) with W uniform in [1, 10] at
    seeded uniform positions; seeded random director start.  Returns
    (lattice, force); the reference GPU metric semantics."""
    A, B, C, S0 = phase_constants()
    rng = np.random.default_rng(geometry_seed)
    lat = qmin.Lattice(L, L, L, precision=precision)
    lat.random_init(seed, S0)
    for _ in range(4):
        c = tuple(float(v) for v in rng.uniform(0, L, 3))
        lat.sphere(c, float(rng.integers(12, 25)) * L / 384, qmin.HOMEOTROPIC, float(rng.uniform(1, 10)), S0)
    f = qmin.LandauDeGennesLC(A, B, C, 2.32, 1.0, 0.5, 0.2, 0.3, semantics=qmin.SEMANTICS_REF_GPU)
    f.set_efield((0.0, 0.0, 0.5), 1.0, 1.0, 1.0)
    return lat, f
