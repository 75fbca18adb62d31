"""Benchmark systems: the geometries of the J/K workloads, built from
their textual definitions with fixed seeds.

* ``h2o``              O at the origin, H at (0, +-0.757, 0.587) Angstrom
* ``water_cluster``    16 waters on the first sites of a 3.0 Angstrom cubic
                       lattice (4 x 2 x 2 block), uniform jitter +-0.3
                       Angstrom and a uniform random rotation about O per
                       molecule, numpy.random.default_rng(seed)
* ``alkane``           all-trans CnH(2n+2) zig-zag, C-C 1.54, C-H 1.09
                       Angstrom, tetrahedral angles
* ``random_chain``     seeded random walk of 3-12 heavy atoms (C/N/O),
                       bonds 1.40-1.50 Angstrom, angles 105-115 degrees,
                       uniform dihedrals, capped with H at 1.09 Angstrom
"""

import math

import numpy

TETRA = math.acos(-1.0 / 3.0)    # 109.47 degrees

H2O = [('O', (0.0, 0.0, 0.0)), ('H', (0.0, -0.757, 0.587)), ('H', (0.0, 0.757, 0.587))]


def atoms_to_str(atoms):
    return '; '.join('%s %.10f %.10f %.10f' % (s, x, y, z) for s, (x, y, z) in atoms)


def h2o():
    return atoms_to_str(H2O)


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= numpy.linalg.norm(q)
    w, x, y, z = q
    return numpy.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(n=16, spacing=3.0, jitter=0.3, seed=1, shape=(4, 2, 2)):
    rng = numpy.random.default_rng(seed)
    sites = [(i, j, k) for k in range(shape[2]) for j in range(shape[1]) for i in range(shape[0])][:n]
    mono = numpy.array([c for _, c in H2O])
    atoms = []
    for site in sites:
        shift = numpy.array(site, dtype=float) * spacing + rng.uniform(-jitter, jitter, 3)
        rot = _random_rotation(rng)
        for (sym, _), r in zip(H2O, mono):
            atoms.append((sym, tuple(rot @ r + shift)))
    return atoms_to_str(atoms)


def _unit(v):
    return v / numpy.linalg.norm(v)


def alkane(n=30, dcc=1.54, dch=1.09):
    half = TETRA / 2
    c = [numpy.array([i * dcc * math.sin(half), (i % 2) * dcc * math.cos(half), 0.0]) for i in range(n)]
    atoms = [('C', tuple(x)) for x in c]
    zhat = numpy.array([0.0, 0.0, 1.0])
    alpha = TETRA / 2
    for i in range(n):
        nbr = [c[j] for j in (i - 1, i + 1) if 0 <= j < n]
        if len(nbr) == 2:
            b = -_unit(_unit(nbr[0] - c[i]) + _unit(nbr[1] - c[i]))
            dirs = [math.cos(alpha) * b + s * math.sin(alpha) * zhat for s in (1, -1)]
        else:
            a = _unit(c[i] - nbr[0])        # away from the single neighbour
            perp = _unit(numpy.cross(a, zhat))
            inplane = _unit(numpy.cross(zhat, a))
            dirs = []
            for k in range(3):
                phi = 2 * math.pi * k / 3
                side = math.cos(phi) * inplane + math.sin(phi) * zhat
                if i == n - 1:
                    side = math.cos(phi) * -inplane + math.sin(phi) * zhat
                dirs.append(math.cos(math.pi - TETRA) * a + math.sin(math.pi - TETRA) * side)
            del perp
        for d in dirs:
            atoms.append(('H', tuple(c[i] + dch * _unit(d))))
    return atoms_to_str(atoms)


def random_chain(seed, nmin=3, nmax=12, bond=(1.40, 1.50), angle=(105.0, 115.0), dch=1.09):
    """Random-walk chain of heavy atoms capped with hydrogens.

    Heavy atoms: uniform over C/N/O; each heavy atom gets 4 - (neighbours)
    valences filled with H (C), 3 - n (N), 2 - n (O), so the electron count
    stays even; returns an atom string."""
    rng = numpy.random.default_rng(seed)
    n = int(rng.integers(nmin, nmax + 1))
    elems = [str(rng.choice(['C', 'N', 'O'])) for _ in range(n)]
    pos = [numpy.zeros(3)]
    if n > 1:
        pos.append(numpy.array([rng.uniform(*bond), 0.0, 0.0]))
    for i in range(2, n):
        b = rng.uniform(*bond)
        th = math.radians(rng.uniform(*angle))
        ph = rng.uniform(0, 2 * math.pi)
        a, bb = pos[i - 2], pos[i - 1]
        u = _unit(bb - a)
        ref = numpy.array([0.0, 0.0, 1.0]) if abs(u[2]) < 0.9 else numpy.array([1.0, 0.0, 0.0])
        v = _unit(numpy.cross(u, ref))
        w = numpy.cross(u, v)
        d = -math.cos(th) * u + math.sin(th) * (math.cos(ph) * v + math.sin(ph) * w)
        pos.append(bb + b * d)
    valence = {'C': 4, 'N': 3, 'O': 2}
    atoms = [(e, tuple(p)) for e, p in zip(elems, pos)]
    nH = 0
    for i in range(n):
        nb = [pos[j] for j in (i - 1, i + 1) if 0 <= j < n]
        k = valence[elems[i]] - len(nb)
        if k <= 0:
            continue
        if nb:
            axis = -_unit(sum(_unit(x - pos[i]) for x in nb))
        else:
            axis = numpy.array([1.0, 0.0, 0.0])
        ref = numpy.array([0.0, 0.0, 1.0]) if abs(axis[2]) < 0.9 else numpy.array([1.0, 0.0, 0.0])
        v = _unit(numpy.cross(axis, ref))
        w = numpy.cross(axis, v)
        spread = 0.0 if k == 1 else math.radians(55.0)
        for h in range(k):
            phi = 2 * math.pi * h / k + rng.uniform(0, 0.2)
            d = math.cos(spread) * axis + math.sin(spread) * (math.cos(phi) * v + math.sin(phi) * w)
            atoms.append(('H', tuple(pos[i] + dch * _unit(d))))
            nH += 1
    return atoms_to_str(atoms)
