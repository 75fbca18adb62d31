"""Gaussian basis-set data and the NWChem-format reader.

Parsing follows pyscf/gto/basis/parse_nwchem.py:105-154 and 298-314:
shells are grouped by ascending angular momentum (file order within each
l), ``SP`` shells are split into an S and a P shell, primitives whose
coefficients are all zero are dropped, and contractions are kept as
written (no contraction optimisation, pyscf/gto/basis/__init__.py:475).

Only the elements H, He, C, N, O, F, Ne of cc-pVDZ, def2-SVP, def2-TZVP
and 6-31G, and of the fitting sets cc-pVDZ-JKFIT, def2-universal-JFIT
('weigend') and def2-universal-JKFIT are shipped (data files beside this
module).
"""

import os

_DIR = os.path.dirname(os.path.abspath(__file__))
MAPSPDF = {'S': 0, 'P': 1, 'D': 2, 'F': 3, 'G': 4, 'H': 5, 'I': 6}
MAXL = 15

ALIAS = {
    'ccpvdz': 'cc-pvdz.dat',
    'def2svp': 'def2-svp.dat',
    'def2tzvp': 'def2-tzvp.dat',
    '631g': '6-31g.dat',
    # auxiliary (density fitting) sets, names of pyscf/gto/basis/__init__.py:67-197
    'ccpvdzjkfit': 'cc-pvdz-jkfit.dat',
    'def2universaljfit': 'def2-universal-jfit.dat',
    'weigend': 'def2-universal-jfit.dat',
    'weigendcfit': 'def2-universal-jfit.dat',
    'def2universaljkfit': 'def2-universal-jkfit.dat',
    'def2svpjkfit': 'def2-universal-jkfit.dat',
    'def2tzvpjkfit': 'def2-universal-jkfit.dat',
}


def _format_name(name):
    return ''.join(c for c in name.lower() if c not in '-_ ()')


def parse(text, symb=None):
    """Parse NWChem-format basis text; if ``symb`` is given keep only the
    shells of that element.  Returns ``[[l, [exp, c1, c2, ...], ...], ...]``."""
    per_l = [[] for _ in range(MAXL)]
    key = None
    keep = symb is None
    cur = None
    for line in text.splitlines():
        dat = line.split('#')[0].strip()
        if not dat or dat.upper().startswith(('END', 'BASIS')):
            continue
        if dat[0].isalpha():
            toks = dat.split()
            el = toks[0] if len(toks) > 1 else None
            key = (toks[1] if len(toks) > 1 else toks[0]).upper()
            keep = symb is None or (el is not None and el.upper() == symb.upper())
            if not keep:
                continue
            if key == 'SP':
                per_l[0].append([0])
                per_l[1].append([1])
            elif key in MAPSPDF:
                cur = [MAPSPDF[key]]
                per_l[MAPSPDF[key]].append(cur)
            else:
                raise ValueError('not basis data: %r' % line)
        else:
            if not keep:
                continue
            vals = [float(x) for x in dat.replace('D', 'e').replace('d', 'e').split()]
            if key == 'SP':
                per_l[0][-1].append([vals[0], vals[1]])
                per_l[1][-1].append([vals[0], vals[2]])
            else:
                cur.append(vals)
    shells = [b for bs in per_l for b in bs]
    out = []
    for b in shells:
        rows = [r for r in b[1:] if any(c != 0 for c in r[1:])]
        if rows:
            out.append([b[0]] + rows)
    if not out:
        raise KeyError('basis not found for %s' % symb)
    return out


def load(name, symb):
    """Basis of element ``symb`` from the named set (e.g. 'cc-pVDZ')."""
    key = _format_name(name)
    if key not in ALIAS:
        if os.path.isfile(name):
            with open(name) as f:
                return parse(f.read(), symb)
        raise KeyError('basis %r is not shipped (available: %s)' % (name, sorted(ALIAS)))
    with open(os.path.join(_DIR, ALIAS[key])) as f:
        return parse(f.read(), symb)
