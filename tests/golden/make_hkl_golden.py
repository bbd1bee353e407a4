"""Writes the hand-derived reflection lists of tests/golden/hkl_*.json.

Each list is the one js/rietveld_core.js:247-276 (generateReflections) must produce for a
cell, a centring/absence rule and a dmin, derived here without the oracle, the product or
the space-group generators of powder6_amd/symmetry.py:

  * the Laue group acting on Miller indices as row vectors (h' = h R, the reference's
    hR, :189) is closed from two or three hand-written integer matrices per crystal
    family -- the cubic 4-fold and 3-fold, the hexagonal 6-fold and the 2-fold along
    [110], the trigonal 3-fold and the same 2-fold -- plus the inversion (Friedel pairs
    merged, :182-188);
  * the systematic absences are the textbook centring conditions (F: h, k, l unmixed;
    R obverse on hexagonal axes: -h + k + l = 3n; P: none);
  * 1/d^2 is exact rational arithmetic (fractions.Fraction) from the closed forms
    (h^2 + k^2 + l^2)/a^2 and 4 (h^2 + h k + k^2)/(3 a^2) + l^2/c^2, cut at
    (1/dmin^2)(1 + 1e-9) (:250);
  * the canonical member of each orbit is the one with the most non-negative indices,
    then the largest (h, k, l) in lexicographic order (:257-259: "(1 1 0) rather than
    (2 -1 0) in a hexagonal cell, (0 1 2) rather than (1 0 -2)");
  * the multiplicity is the orbit size; the order is decreasing d, then h, k, l
    decreasing (:274) -- here on exact values, so tied orbits are ordered by the
    comparator's tie rule alone.

Run: python tests/golden/make_hkl_golden.py   (writes the JSON files next to it)
"""
import json
import math
import os
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def _mul(A, B):
    return tuple(tuple(sum(A[i][k] * B[k][j] for k in range(3)) for j in range(3)) for i in range(3))


def closure(gens):
    ident = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    inv = ((-1, 0, 0), (0, -1, 0), (0, 0, -1))
    group, frontier = {ident}, [ident]
    gens = list(gens) + [inv]
    while frontier:
        nxt = []
        for a in frontier:
            for g in gens:
                c = _mul(a, g)
                if c not in group:
                    group.add(c)
                    nxt.append(c)
        frontier = nxt
    return sorted(group)


def row_times(h, R):
    return tuple(sum(h[i] * R[i][j] for i in range(3)) for j in range(3))


# direct-space rotation matrices (x' = R x); Miller indices transform as h' = h R
CUBIC = closure([((0, -1, 0), (1, 0, 0), (0, 0, 1)),      # 4-fold about c: -y, x, z
                 ((0, 0, 1), (1, 0, 0), (0, 1, 0))])     # 3-fold about [111]: z, x, y
HEXAGONAL = closure([((1, -1, 0), (1, 0, 0), (0, 0, 1)),  # 6-fold about c: x-y, x, z
                     ((0, 1, 0), (1, 0, 0), (0, 0, -1))])  # 2-fold along [110]: y, x, -z
TRIGONAL = closure([((0, -1, 0), (1, -1, 0), (0, 0, 1)),  # 3-fold about c: -y, x-y, z
                    ((0, 1, 0), (1, 0, 0), (0, 0, -1))])   # 2-fold along [110]: y, x, -z
assert len(CUBIC) == 48 and len(HEXAGONAL) == 24 and len(TRIGONAL) == 12


def canonical(orbit):
    best = None
    for v in orbit:
        key = (sum(1 for c in v if c >= 0), v)
        if best is None or key > best:
            best = key
    return best[1]


def reflection_list(laue, qform, allowed, dmin, hmax):
    lim = Fraction(1) / (Fraction(dmin) * Fraction(dmin)) * (1 + Fraction(1, 10 ** 9))
    rows = []
    for h in range(-hmax[0], hmax[0] + 1):
        for k in range(-hmax[1], hmax[1] + 1):
            for l in range(-hmax[2], hmax[2] + 1):
                if h == k == l == 0:
                    continue
                q = qform(h, k, l)
                if q > lim:
                    continue
                orbit = {row_times((h, k, l), R) for R in laue}
                if canonical(orbit) != (h, k, l) or not allowed(h, k, l):
                    continue
                rows.append((q, h, k, l, len(orbit)))
    # decreasing d = increasing q; ties: h, k, l decreasing
    rows.sort(key=lambda r: (r[0], -r[1], -r[2], -r[3]))
    return [[h, k, l, m] for (q, h, k, l, m) in rows], [str(q) for (q, *_r) in rows]


def dmin_of(lam, tth_max):
    """reflectionDmin (rietveld_core.js:238-245) with zero, disp and transp at 0."""
    upper = min(179.999999, tth_max + 10 + 0.001)
    return lam / (2 * math.sin(upper * (math.pi / 180) / 2))


def main():
    cases = {}
    # NaCl Fm-3m a = 5.64, Cu Ka1 1.5406, 2theta to 120 deg
    a = Fraction('5.64')
    dmin = dmin_of(1.5406, 120.0)
    hm = int(math.floor(5.64 / dmin))
    rows, q = reflection_list(CUBIC, lambda h, k, l: Fraction(h * h + k * k + l * l) / (a * a),
                              lambda h, k, l: (h % 2 == k % 2 == l % 2), dmin, (hm, hm, hm))
    cases['hkl_nacl_fm3m.json'] = {'group': 225, 'system': 'cubic', 'cell': [5.64, 5.64, 5.64, 90, 90, 90],
                                   'dmin': dmin, 'rows': rows, 'inv_dsq_exact': q}
    # hexagonal P6/mmm a = 3.1, c = 5.05, Cu Ka1 to 100 deg
    a, c = Fraction('3.1'), Fraction('5.05')
    dmin = dmin_of(1.5406, 100.0)
    qh = lambda h, k, l: Fraction(4, 3) * (h * h + h * k + k * k) / (a * a) + Fraction(l * l) / (c * c)
    hm = (int(math.floor(3.1 / dmin)), int(math.floor(3.1 / dmin)), int(math.floor(5.05 / dmin)))
    rows, q = reflection_list(HEXAGONAL, qh, lambda h, k, l: True, dmin, hm)
    cases['hkl_hexagonal_p6mmm.json'] = {'group': 191, 'system': 'hexagonal', 'cell': [3.1, 3.1, 5.05, 90, 90, 120],
                                         'dmin': dmin, 'rows': rows, 'inv_dsq_exact': q}
    # trigonal R-3m (hexagonal axes) a = 4.9, c = 13.9, Cu Ka1 to 80 deg
    a, c = Fraction('4.9'), Fraction('13.9')
    dmin = dmin_of(1.5406, 80.0)
    qh = lambda h, k, l: Fraction(4, 3) * (h * h + h * k + k * k) / (a * a) + Fraction(l * l) / (c * c)
    hm = (int(math.floor(4.9 / dmin)), int(math.floor(4.9 / dmin)), int(math.floor(13.9 / dmin)))
    rows, q = reflection_list(TRIGONAL, qh, lambda h, k, l: (-h + k + l) % 3 == 0, dmin, hm)
    cases['hkl_trigonal_r3m.json'] = {'group': 166, 'system': 'trigonal', 'cell': [4.9, 4.9, 13.9, 90, 90, 120],
                                      'dmin': dmin, 'rows': rows, 'inv_dsq_exact': q}
    for name, data in cases.items():
        data['source'] = 'tests/golden/make_hkl_golden.py (exact rational arithmetic, hand-written Laue generators)'
        with open(os.path.join(HERE, name), 'w') as f:
            json.dump(data, f, indent=0)
        print(name, len(data['rows']), 'rows')


if __name__ == '__main__':
    main()
