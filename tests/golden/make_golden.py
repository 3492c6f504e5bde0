"""Writes tests/golden/ref_cubature.npz from the reference's own C++ sources.

Run in a container that has the reference tree and oracle/_ref built
(`make -C oracle`). The fixture holds inputs and outputs only:
  * cubature_nop 6x6 blocks (cubature_nop.cpp) at the orders 4/10/20 the
    assembly uses (assembly_nop.m:17-19), for several offsets and (dx, f);
  * the five self-term integrals from DIRECTFN and DEMCEM at Np = 20
    (singular_st/ea/va.m, assembly_nop.m:16);
  * Gauss-Legendre nodes/weights quoted from gauss_1d.m for orders 4, 10, 20.
and tests/golden/ref_cubature_kop.npz:
  * cubature_kop 3-vectors (cubature_kop.cpp) at the orders 4/10 of
    assembly_kop.m:32-66 for offsets without face contact.
and tests/golden/ref_singular.npz:
  * the five self-term integrals from DIRECTFN and DEMCEM at Np = 10, 20, 30
    for the four (dx, f) cases (getOPERATORS.m:40-42 defaults to DIRECTFN;
    singular_st.m:13-19, singular_ea.m:16-29, singular_va.m);
  * the C and P tables of nearby_correct.m:5-133, per offset: C as codes
    0 = zero, 1 = ST, 2 = EAc, 3 = EAo, 4 = VAc, 5 = VAo, P as 0/1.
"""
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import operators as oop  # noqa: E402
from oracle import ref  # noqa: E402

CASES = [(6.25e-3, 298.2e6), (2e-3, 128e6), (2e-3, 298e6), (1.5e-3, 298e6)]
OFFSETS = [(0, 0, 0), (1, 0, 0), (1, 1, 1), (2, 1, 0), (3, 1, 2), (4, 4, 4), (7, 0, 0), (10, 3, 5)]
# the K cubature is evaluated on face pairs' Gauss points directly, so offsets
# whose cubes share a face (coincident points) are left out
OFFSETS_K = [(1, 1, 0), (1, 1, 1), (2, 1, 0), (3, 1, 2), (4, 4, 4), (7, 0, 0), (10, 3, 5)]
C0 = 299792458.0


def gauss_table(order, path='/root/reference/src_vie/src_operators/cubatures/gauss_1d.m'):
    txt = open(path).read()
    blk = re.split(r'order == %d \)' % order, txt)[1]
    blk = re.split(r'elseif|else\s', blk)[0]
    xs = {int(i): float(v.replace(' ', '')) for i, v in re.findall(r'xtab\((\d+)\)\s*=\s*([-+ 0-9.Ee]+);', blk)}
    ws = {int(i): float(v.replace(' ', '')) for i, v in re.findall(r'weight\((\d+)\)\s*=\s*([-+ 0-9.Ee]+);', blk)}
    n = len(xs)
    return np.array([xs[i + 1] for i in range(n)]), np.array([ws[i + 1] for i in range(n)])


def nearby_tables_from_reference(path='/root/reference/src_vie/src_operators/nearby_correct.m'):
    """(offsets 8 x 3 one-based, C codes 8 x 6 x 6, P 8 x 6 x 6) parsed from nearby_correct.m."""
    txt = open(path).read()
    names = {'0.0': 0, 'ST': 1, 'EAc': 2, 'EAo': 3, 'VAc': 4, 'VAo': 5}
    offs, cc, pp = [], [], []
    for b in re.split(r'isequal\(m,\s*\[', txt)[1:]:
        offs.append([int(v) for v in re.match(r'([0-9,\s]+)\]', b).group(1).split(',')])
        C = re.search(r'C\s*=\s*\[(.*?)\];', b, re.S).group(1)
        P = re.search(r'P\s*=\s*\[(.*?)\];', b, re.S).group(1)
        cc.append([[names[t] for t in row.split()] for row in C.strip().split('\n') if row.strip()])
        pp.append([[int(float(t)) for t in row.split()] for row in P.strip().split('\n') if row.strip()])
    return np.array(offs, np.int64), np.array(cc, np.int8), np.array(pp, np.int8)


def singular_sweep(R, nps=(10, 20, 30)):
    dfn = np.array([[R.singular_terms(2 * np.pi / C0 * f, dx, 'DIRECTFN', n) for n in nps] for dx, f in CASES])
    dem = np.array([[R.singular_terms(2 * np.pi / C0 * f, dx, 'DEMCEM', n) for n in nps] for dx, f in CASES])
    offs, cc, pp = nearby_tables_from_reference()
    np.savez_compressed(os.path.join(os.path.dirname(__file__), 'ref_singular.npz'), np=np.array(nps),
                        cases=np.array(CASES), directfn=dfn, demcem=dem, nearby_offsets=offs, nearby_C=cc,
                        nearby_P=pp)
    print('wrote self-terms at Np', nps, 'and', len(offs), 'nearby tables')


def main():
    R = ref.Ref()
    if len(sys.argv) > 1 and sys.argv[1] == 'singular':
        singular_sweep(R)
        return
    out = {}
    blocks, meta = [], []
    for dx, f in CASES:
        ko = 2 * np.pi / C0 * f
        Rf = oop.cube_faces(dx)
        for order in (4, 10, 20):
            Np, w, u, v = oop.gauss_2d(order)
            for m in OFFSETS:
                rm = np.array(m, float) * dx
                blocks.append(R.cubature_nop(Np, w, u, v, rm, np.zeros(3), Rf, ko, dx))
                meta.append([dx, f, order, m[0], m[1], m[2]])
        out_d = np.array(R.singular_terms(ko, dx, 'DIRECTFN', 20))
        out_m = np.array(R.singular_terms(ko, dx, 'DEMCEM', 20))
        out.setdefault('sing_directfn', []).append(out_d)
        out.setdefault('sing_demcem', []).append(out_m)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), 'ref_cubature.npz'),
                        cub_blocks=np.array(blocks), cub_meta=np.array(meta),
                        sing_cases=np.array(CASES), sing_directfn=np.array(out['sing_directfn']),
                        sing_demcem=np.array(out['sing_demcem']),
                        **{'gauss_x_%d' % o: gauss_table(o)[0] for o in (4, 10, 20)},
                        **{'gauss_w_%d' % o: gauss_table(o)[1] for o in (4, 10, 20)})
    print('wrote', len(blocks), 'cubature blocks')
    kb, km = [], []
    for dx, f in CASES:
        ko = 2 * np.pi / C0 * f
        Rf = oop.cube_faces(dx)
        for order in (4, 10):
            Np, w, u, v = oop.gauss_2d(order)
            for m in OFFSETS_K:
                rm = np.array(m, float) * dx
                kb.append(R.cubature_kop(Np, w, u, v, rm, np.zeros(3), Rf, ko))
                km.append([dx, f, order, m[0], m[1], m[2]])
    np.savez_compressed(os.path.join(os.path.dirname(__file__), 'ref_cubature_kop.npz'),
                        kop_blocks=np.array(kb), kop_meta=np.array(km))
    print('wrote', len(kb), 'K cubature vectors')


if __name__ == '__main__':
    main()
