"""Benchmark systems: the geometries of the J/K workloads, built from
their textual definitions with fixed seeds.

* ``h2o``              O at the origin, H at (0, +-0.757, 0.587) Angstrom
* ``water_cluster``    16 waters on the first sites of a 3.0 Angstrom cubic
                       lattice (4 x 2 x 2 block), uniform jitter +-0.3
                       Angstrom and a uniform random rotation about O per
                       molecule, numpy.random.default_rng(seed)
* ``alkane``           all-trans CnH(2n+2) zig-zag, C-C 1.54, C-H 1.09
                       Angstrom, tetrahedral angles
* ``random_chain``     seeded self-avoiding random walk of 3-12 heavy
                       atoms (C/N/O), bonds 1.40-1.50 Angstrom, angles
                       105-115 degrees, uniform dihedrals, capped with H at
                       1.09 Angstrom on tetrahedral sites
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


def _chain_positions(rng, n, bond, angle, tries=200):
    """Heavy-atom positions of one self-avoiding random walk: bond lengths
    and angles uniform in their ranges, dihedrals uniform; a step that comes
    closer than 2.2 Angstrom to an earlier atom, or 2.6 Angstrom to one at
    least three bonds back (room for the hydrogens), is drawn again.
    Returns None when a step finds no place in `tries` draws."""
    pos = [numpy.zeros(3)]
    if n > 1:
        pos.append(numpy.array([rng.uniform(*bond), 0.0, 0.0]))
    for i in range(2, n):
        a, bb = pos[i - 2], pos[i - 1]
        u = _unit(bb - a)
        ref = numpy.array([0.0, 0.0, 1.0]) if abs(u[2]) < 0.9 else numpy.array([1.0, 0.0, 0.0])
        v = _unit(numpy.cross(u, ref))
        w = numpy.cross(u, v)
        earlier = numpy.array(pos[:i - 1])
        lim = numpy.where(numpy.arange(i - 1) <= i - 3, 2.6, 2.2)
        for _ in range(tries):
            b = rng.uniform(*bond)
            th = math.radians(rng.uniform(*angle))
            ph = rng.uniform(0, 2 * math.pi)
            d = -math.cos(th) * u + math.sin(th) * (math.cos(ph) * v + math.sin(ph) * w)
            cand = bb + b * d
            if (numpy.linalg.norm(earlier - cand, axis=1) >= lim).all():
                pos.append(cand)
                break
        else:
            return None
    return pos


def _cap_directions(rng, nbr_units, k):
    """Unit vectors of k hydrogens on a heavy atom whose bonds to its heavy
    neighbours point along nbr_units: tetrahedral sites away from those
    bonds (two neighbours: the two sites off the heavy-atom plane, or the
    bisector for one H; one neighbour: sites at 109.47 degrees from the
    bond, 120 degrees apart, with a random twist)."""
    if k <= 0:
        return []
    if len(nbr_units) == 2:
        u1, u2 = nbr_units
        a = -_unit(u1 + u2)
        nrm = numpy.cross(u1, u2)
        if numpy.linalg.norm(nrm) < 1e-8:      # collinear neighbours
            ref = numpy.array([0.0, 0.0, 1.0]) if abs(a[2]) < 0.9 else numpy.array([1.0, 0.0, 0.0])
            nrm = numpy.cross(a, ref)
        nrm = _unit(nrm)
        if k == 1:
            return [a]
        half = TETRA / 2
        return [math.cos(half) * a + s * math.sin(half) * nrm for s in (1.0, -1.0)][:k]
    if len(nbr_units) == 1:
        axis = -nbr_units[0]
    else:
        axis = numpy.array([1.0, 0.0, 0.0])
    ref = numpy.array([0.0, 0.0, 1.0]) if abs(axis[2]) < 0.9 else numpy.array([1.0, 0.0, 0.0])
    v = _unit(numpy.cross(axis, ref))
    w = numpy.cross(axis, v)
    tilt = math.pi - TETRA
    twist = rng.uniform(0, 2 * math.pi)
    dirs = []
    for h in range(k):
        if not nbr_units and h == 0:
            dirs.append(axis)
            continue
        phi = twist + 2 * math.pi * h / 3
        dirs.append(math.cos(tilt) * axis + math.sin(tilt) * (math.cos(phi) * v + math.sin(phi) * w))
    return dirs


# closest approach allowed between atoms that are not bonded to each other
MIN_SEPARATION = {('H', 'H'): 1.5, ('H', 'X'): 1.8, ('X', 'X'): 2.2}


def _clash(atoms, bonded):
    xyz = numpy.array([p for _, p in atoms])
    kind = ['H' if s == 'H' else 'X' for s, _ in atoms]
    d = numpy.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
    for i in range(len(atoms)):
        for j in range(i):
            if (j, i) in bonded:
                continue
            lim = MIN_SEPARATION[tuple(sorted((kind[i], kind[j])))]
            if d[i, j] < lim:
                return True
    return False


def random_chain(seed, nmin=3, nmax=12, bond=(1.40, 1.50), angle=(105.0, 115.0), dch=1.09, tries=1000):
    """Self-avoiding random-walk chain of heavy atoms capped with hydrogens.

    Heavy atoms: uniform over C/N/O; each heavy atom gets 4 - (neighbours)
    valences filled with H (C), 3 - n (N), 2 - n (O) on tetrahedral sites,
    so the electron count stays even.  Steps of the walk and the twist of a
    chain end's hydrogens are drawn again while they bring two atoms that
    are not bonded closer than MIN_SEPARATION (a walk that cannot continue
    starts over), all from the one generator, so the chain is a function of
    the seed alone.  Returns an atom string."""
    rng = numpy.random.default_rng(seed)
    n = int(rng.integers(nmin, nmax + 1))
    elems = [str(rng.choice(['C', 'N', 'O'])) for _ in range(n)]
    valence = {'C': 4, 'N': 3, 'O': 2}
    lim_h = {'H': MIN_SEPARATION[('H', 'H')], 'X': MIN_SEPARATION[('H', 'X')]}
    for _ in range(tries):
        pos = _chain_positions(rng, n, bond, angle)
        if pos is None:
            continue
        atoms = [(e, tuple(p)) for e, p in zip(elems, pos)]
        bonded = {(i, i + 1) for i in range(n - 1)}
        ok = True
        for i in range(n):
            nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
            units = [_unit(pos[j] - pos[i]) for j in nb]
            k = valence[elems[i]] - len(nb)
            if k <= 0:
                continue
            others = [(s_, numpy.array(p)) for idx, (s_, p) in enumerate(atoms) if idx != i]
            for _twist in range(50):
                hs = [pos[i] + dch * _unit(d) for d in _cap_directions(rng, units, k)]
                if all(numpy.linalg.norm(h - p) >= lim_h['H' if s_ == 'H' else 'X']
                       for h in hs for s_, p in others):
                    break
                if len(nb) == 2:        # fixed sites: no other twist to try
                    hs = None
                    break
            else:
                hs = None
            if hs is None:
                ok = False
                break
            for h in hs:
                bonded.add((i, len(atoms)))
                atoms.append(('H', tuple(h)))
        if ok and not _clash(atoms, bonded):
            return atoms_to_str(atoms)
    raise RuntimeError('no self-avoiding chain for seed %r in %d draws' % (seed, tries))
