"""Space-group operators and Wyckoff sites for the synthetic PT workloads.

Operators are generated here by closing a few generators from the International Tables
for Crystallography, Vol. A (coordinate triplets of the standard settings) under
composition, then crossing them with the centring translations -- the list the
reference keeps per setting in sg/all.json and expands with expandOps
(js/rietveld_core.js:139-153).  Translations are held in 1/24ths so that the closure is
exact; they are handed to the library as doubles.

site_symmetry() restates js/rietveld_core.js:305-370: the stabiliser of a point, the
multiplicity N/|stabiliser|, the position projected onto the site, and the free
directions as the null space of the stacked (R - I) in reduced row-echelon form with
the columns taken z, y, x.
"""

import math

import numpy as np

DEN = 24   # translation denominator


def _parse_triplet(text):
    """'-x+1/2,y+1/2,-z' -> (R 3x3 int, t[3] in 1/24ths)."""
    rows, ts = [], []
    for comp in text.replace(' ', '').split(','):
        row, t, i, sign = [0, 0, 0], 0, 0, 1
        while i < len(comp):
            ch = comp[i]
            if ch in '+-':
                sign = 1 if ch == '+' else -1
                i += 1
                continue
            if ch in 'xyz':
                row['xyz'.index(ch)] += sign
                sign = 1
                i += 1
                continue
            j = i
            while j < len(comp) and (comp[j].isdigit() or comp[j] == '/'):
                j += 1
            num = comp[i:j]
            if '/' in num:
                a, b = num.split('/')
                t += sign * int(a) * DEN // int(b)
            else:
                t += sign * int(num) * DEN
            sign = 1
            i = j
        rows.append(row)
        ts.append(t % DEN)
    return (tuple(tuple(r) for r in rows), tuple(ts))


def _compose(a, b):
    """(Ra, ta) o (Rb, tb): x -> Ra (Rb x + tb) + ta."""
    Ra, ta = np.array(a[0]), np.array(a[1])
    Rb, tb = np.array(b[0]), np.array(b[1])
    R = Ra @ Rb
    t = (Ra @ tb + ta) % DEN
    return (tuple(tuple(int(v) for v in r) for r in R), tuple(int(v) for v in t))


def _closure(gens):
    ident = (((1, 0, 0), (0, 1, 0), (0, 0, 1)), (0, 0, 0))
    ops = [ident]
    seen = {ident}
    frontier = [ident]
    gens = [_parse_triplet(g) for g in gens]
    while frontier:
        nxt = []
        for a in frontier:
            for g in gens:
                c = _compose(g, a)
                if c not in seen:
                    seen.add(c)
                    ops.append(c)
                    nxt.append(c)
        frontier = nxt
        if len(ops) > 192:
            raise ValueError('generators do not close on a crystallographic group')
    return ops


CENTRING = {
    'P': ['0,0,0'],
    'I': ['0,0,0', '1/2,1/2,1/2'],
    'F': ['0,0,0', '0,1/2,1/2', '1/2,0,1/2', '1/2,1/2,0'],
    'C': ['0,0,0', '1/2,1/2,0'],
    'A': ['0,0,0', '0,1/2,1/2'],
    'B': ['0,0,0', '1/2,0,1/2'],
    'R': ['0,0,0', '2/3,1/3,1/3', '1/3,2/3,2/3'],      # obverse, hexagonal axes
}

# ITA Vol. A generators (standard settings; origin choice 1 where there are two)
GROUPS = {
    # number: (symbol, system, centring, generators)
    1: ('P1', 'triclinic', 'P', []),
    2: ('P-1', 'triclinic', 'P', ['-x,-y,-z']),
    4: ('P2_1', 'monoclinic', 'P', ['-x,y+1/2,-z']),
    10: ('P2/m', 'monoclinic', 'P', ['-x,y,-z', '-x,-y,-z']),
    14: ('P2_1/c', 'monoclinic', 'P', ['-x,y+1/2,-z+1/2', '-x,-y,-z']),
    15: ('C2/c', 'monoclinic', 'C', ['-x,y,-z+1/2', '-x,-y,-z']),
    19: ('P2_12_12_1', 'orthorhombic', 'P', ['-x+1/2,-y,z+1/2', '-x,y+1/2,-z+1/2']),
    62: ('Pnma', 'orthorhombic', 'P', ['-x+1/2,-y,z+1/2', '-x,y+1/2,-z', '-x,-y,-z']),
    123: ('P4/mmm', 'tetragonal', 'P', ['-y,x,z', 'x,-y,-z', '-x,-y,-z']),
    136: ('P4_2/mnm', 'tetragonal', 'P', ['-y+1/2,x+1/2,z+1/2', '-x+1/2,y+1/2,-z+1/2', '-x,-y,-z']),
    139: ('I4/mmm', 'tetragonal', 'I', ['-y,x,z', 'x,-y,-z', '-x,-y,-z']),
    164: ('P-3m1', 'trigonal', 'P', ['-y,x-y,z', 'y,x,-z', '-x,-y,-z']),
    166: ('R-3m', 'trigonal', 'R', ['-y,x-y,z', 'y,x,-z', '-x,-y,-z']),
    191: ('P6/mmm', 'hexagonal', 'P', ['-y,x-y,z', '-x,-y,z', 'y,x,-z', '-x,-y,-z']),
    221: ('Pm-3m', 'cubic', 'P', ['-y,x,z', 'z,x,y', '-x,-y,-z']),
    216: ('F-43m', 'cubic', 'F', ['y,-x,-z', 'z,x,y']),
    225: ('Fm-3m', 'cubic', 'F', ['-y,x,z', 'z,x,y', '-x,-y,-z']),
    229: ('Im-3m', 'cubic', 'I', ['-y,x,z', 'z,x,y', '-x,-y,-z']),
}


def group_ops(number):
    """Every operator of the group (point operations x centrings), deduplicated on
    1/24ths, in the order expandOps would produce it.  Returns (R[n,9] int32,
    t[n,3] float64)."""
    sym, system, cen, gens = GROUPS[number]
    base = _closure(gens)
    cens = [_parse_triplet(_cen_triplet(c))[1] for c in CENTRING[cen]]
    out, seen = [], set()
    for R, t in base:
        for c in cens:
            tt = tuple((t[i] + c[i]) % DEN for i in range(3))
            key = (R, tt)
            if key not in seen:
                seen.add(key)
                out.append(key)
    r = np.array([[v for row in R for v in row] for R, _ in out], dtype=np.int32)
    tt = np.array([[v / DEN for v in t] for _, t in out], dtype=np.float64)
    return r, tt


def _cen_triplet(c):
    a, b, d = c.split(',')
    return 'x+%s,y+%s,z+%s' % (a, b, d)


def group_system(number):
    return GROUPS[number][1]


def _frac_distance(G, d):
    return math.sqrt(max(0.0, float(d @ G @ d)))


def metric_tensor(cell):
    a, b, c, al, be, ga = cell
    ca, cb, cg = (math.cos(math.radians(v)) for v in (al, be, ga))
    return np.array([[a * a, a * b * cg, a * c * cb],
                     [a * b * cg, b * b, b * c * ca],
                     [a * c * cb, b * c * ca, c * c]])


def site_symmetry(op_r, op_t, xyz, cell, tol=0.02):
    """Wyckoff site of a point (js/rietveld_core.js:305-370): returns dict with r0[3],
    basis[3][3] (row f = direction of free coordinate f), free (list), mult, pos."""
    G = metric_tensor(cell)
    x = np.array(xyz, dtype=float)
    stab, acc = [], np.zeros(3)
    for R9, t in zip(op_r, op_t):
        R = np.array(R9, dtype=float).reshape(3, 3)
        y = R @ x + t
        d = y - x
        L = np.round(d)
        if _frac_distance(G, d - L) < tol:
            stab.append(R)
            acc += y - L
    ns = max(1, len(stab))
    pos = acc / ns
    rows = []
    for R in stab:
        for i in range(3):
            row = [R[i, 2] - (1 if i == 2 else 0), R[i, 1] - (1 if i == 1 else 0), R[i, 0] - (1 if i == 0 else 0)]
            if any(abs(v) > 1e-12 for v in row):
                rows.append(row)
    piv, rr = [], 0
    for c in range(3):
        if rr >= len(rows):
            break
        best, bv = -1, 1e-9
        for i in range(rr, len(rows)):
            if abs(rows[i][c]) > bv:
                bv, best = abs(rows[i][c]), i
        if best < 0:
            continue
        rows[rr], rows[best] = rows[best], rows[rr]
        p = rows[rr][c]
        rows[rr] = [v / p for v in rows[rr]]
        for i in range(len(rows)):
            if i != rr and abs(rows[i][c]) > 1e-12:
                f = rows[i][c]
                rows[i] = [rows[i][j] - f * rows[rr][j] for j in range(3)]
        piv.append((c, rr))
        rr += 1
    pivcols = [c for c, _ in piv]
    col_to_coord = [2, 1, 0]
    free, basis = [], {}
    for col in (2, 1, 0):
        if col in pivcols:
            continue
        b = [0.0, 0.0, 0.0]
        b[col_to_coord[col]] = 1.0
        for pc, prow in piv:
            b[col_to_coord[pc]] = -rows[prow][col]
        b = [0.0 if abs(v) < 1e-12 else v for v in b]
        free.append(col_to_coord[col])
        basis[col_to_coord[col]] = b
    r0 = list(pos)
    for f in free:
        u = pos[f]
        for i in range(3):
            r0[i] -= u * basis[f][i]
    r0 = [0.0 if abs(v) < 1e-12 else v for v in r0]
    B = [[0.0] * 3 for _ in range(3)]
    for f in free:
        B[f] = list(basis[f])
    return {'r0': r0, 'basis': B, 'free': free, 'mult': int(round(len(op_r) / ns)), 'pos': list(pos)}
