"""Known answers of the reference's translation-symmetry tests, as one check that runs against any
implementation of the primitives:

- tests/test_symmetry_translation_stage1.cpp:14-76: site indexing x + 4y on 4x4, periodicity,
  the site / occupation / raw-state translation homomorphisms, single-particle moves and wraps;
- tests/test_symmetry_stage2.cpp:30-112: orbit sizes, stabilisers and primitive periods of a single
  particle, a full row stripe, the vacuum and a two-particle state; |orbit| |Stab| = |G| and
  #compatible momenta = |orbit|;
- tests/test_symmetry_stage3.cpp:23-97: momentum compatibility of a stabiliser containing (2,0),
  the Bloch phase on trivial / non-trivial characters, and the K-basis dimensions 3 / 1 / 2 of
  {vacuum, single particle, row stripe} at K = (0,0) / (1,0) / (0,1);
- tests/test_momentum_partition_2x2.cpp:26-67: on every 2x2 sector, sum_K dim K-basis = dim and
  sum over orbits of #compatible momenta = dim.

`translate(lx, ly, up, dn, dx, dy) -> (packed, sign)`, `canonical(lx, ly, up, dn) -> (packed,
|Stab|)`, `bloch(lx, ly, kx, ky, dx, dy) -> complex` and `sector(n, nu, nd) -> (ups, dns)` are the
implementation under test.  Packed states are (up << 16) | dn.
"""
from __future__ import annotations

import math


def _pack(u, d):
    return (int(u) << 16) | int(d)


def _stabiliser(translate, lx, ly, u, d):
    s = _pack(u, d)
    return [(dx, dy) for dy in range(ly) for dx in range(lx) if translate(lx, ly, u, d, dx, dy)[0] == s]


def _orbit(translate, lx, ly, u, d):
    return {translate(lx, ly, u, d, dx, dy)[0] for dy in range(ly) for dx in range(lx)}


def _compatible(kx, ky, lx, ly, stab):
    # the bosonic character exp(-2 pi i (kx dx / lx + ky dy / ly)) is 1 on every stabiliser element
    return all((kx * dx * ly + ky * dy * lx) % (lx * ly) == 0 for dx, dy in stab)


def _period(translate, lx, ly, u, d, axis):
    s, n = _pack(u, d), (lx if axis == 0 else ly)
    for p in range(1, n + 1):
        if translate(lx, ly, u, d, p if axis == 0 else 0, p if axis == 1 else 0)[0] == s:
            return p
    raise AssertionError("no period")


def check_translation_stage1(translate):
    one = lambda site: (1 << site, 0)  # noqa: E731
    # indexing site = x + 4y: (1,0) -> 1, (0,1) -> 4, (3,3) -> 15, and site 6 is (2,1)
    assert translate(4, 4, *one(0), 1, 0)[0] == _pack(1 << 1, 0)
    assert translate(4, 4, *one(0), 0, 1)[0] == _pack(1 << 4, 0)
    assert translate(4, 4, *one(0), 3, 3)[0] == _pack(1 << 15, 0)
    assert translate(4, 4, *one(0), 2, 1)[0] == _pack(1 << 6, 0)
    # periodicity: a full cycle is the identity, -1 wraps, (1,1) from site 15 reaches site 0
    assert translate(4, 4, *one(0), 4, 0)[0] == _pack(1, 0)
    assert translate(4, 4, *one(0), -1, 0)[0] == _pack(1 << 3, 0)
    assert translate(4, 4, *one(15), 1, 1)[0] == _pack(1, 0)
    assert translate(4, 4, *one(3), 1, 0)[0] == _pack(1, 0)
    assert translate(4, 4, *one(12), 0, 1)[0] == _pack(1, 0)
    # site homomorphism: (2,-3) then (-5,8) equals (-3,5)
    s1 = translate(4, 4, *one(7), 2, -3)[0]
    s2 = translate(4, 4, s1 >> 16, s1 & 0xFFFF, -5, 8)[0]
    assert s2 == translate(4, 4, *one(7), -3, 5)[0]
    # occupation bits: a particle moves by one site, and wraps at the row end
    assert translate(4, 4, 1, 0, 1, 0)[0] == _pack(1 << 1, 0)
    assert translate(4, 4, 1 << 3, 0, 1, 0)[0] == _pack(1, 0)
    # group action on a three-particle pattern
    bits = (1 << 0) | (1 << 5) | (1 << 15)
    once = translate(4, 4, bits, 0, 2, -7)[0]
    assert translate(4, 4, once >> 16, once & 0xFFFF, 2, -7)[0] == translate(4, 4, bits, 0, 4, -14)[0]
    # raw states move both strings
    t = translate(4, 4, 1, 1 << 4, 1, 0)[0]
    assert t == _pack(1 << 1, 1 << 5)
    assert translate(4, 4, t >> 16, t & 0xFFFF, -1, 0)[0] == _pack(1, 1 << 4)
    assert translate(4, 4, 1, 1 << 4, 0, 1)[0] == _pack(1 << 4, 1 << 8)


def check_symmetry_stage2(translate, canonical):
    # single up particle: orbit 16, trivial stabiliser, periods 4 and 4
    assert len(_orbit(translate, 4, 4, 1, 0)) == 16
    rep, ss = canonical(4, 4, 1, 0)
    assert canonical(4, 4, rep >> 16, rep & 0xFFFF)[0] == rep and ss == 1
    assert _stabiliser(translate, 4, 4, rep >> 16, rep & 0xFFFF) == [(0, 0)]
    assert _period(translate, 4, 4, 1, 0, 0) == 4 and _period(translate, 4, 4, 1, 0, 1) == 4
    # row stripe 0x000F: orbit 4, stabiliser = the 4 x-translations, periods 1 and 4
    assert len(_orbit(translate, 4, 4, 0x000F, 0)) == 4
    rep, ss = canonical(4, 4, 0x000F, 0)
    stab = _stabiliser(translate, 4, 4, rep >> 16, rep & 0xFFFF)
    assert ss == 4 and len(stab) == 4 and all(dy == 0 for _, dy in stab)
    assert _period(translate, 4, 4, rep >> 16, rep & 0xFFFF, 0) == 1
    assert _period(translate, 4, 4, rep >> 16, rep & 0xFFFF, 1) == 4
    # |orbit| * |Stab| = |G| and #compatible momenta = |orbit| for each seed
    for u, d in [(1, 0), (0x000F, 0), (1, 1 << 1), (0, 0)]:
        orb = _orbit(translate, 4, 4, u, d)
        stab = _stabiliser(translate, 4, 4, u, d)
        assert len(orb) * len(stab) == 16
        assert sum(_compatible(kx, ky, 4, 4, stab) for ky in range(4) for kx in range(4)) == len(orb)
    # orbit members share the representative
    a = (1, 1 << 4)
    b = translate(4, 4, *a, 1, 0)[0]
    assert canonical(4, 4, *a)[0] == canonical(4, 4, b >> 16, b & 0xFFFF)[0]
    # vacuum: orbit 1, periods 1, its own representative
    assert _orbit(translate, 4, 4, 0, 0) == {0}
    assert canonical(4, 4, 0, 0) == (0, 16)
    assert _period(translate, 4, 4, 0, 0, 0) == 1 and _period(translate, 4, 4, 0, 0, 1) == 1
    # periods divide 4 and fix the representative
    rep = canonical(4, 4, 1, 1 << 5)[0]
    for axis in (0, 1):
        p = _period(translate, 4, 4, rep >> 16, rep & 0xFFFF, axis)
        assert p in (1, 2, 4)


def check_symmetry_stage3(translate, canonical, bloch):
    # a stabiliser containing (2,0) admits exactly the even kx
    for kx in range(4):
        for ky in range(4):
            assert _compatible(kx, ky, 4, 4, [(0, 0), (2, 0)]) == (kx % 2 == 0)
    # Bloch phase: 1 where the character is trivial, unit modulus elsewhere
    for dy in range(4):
        for dx in range(4):
            z = bloch(4, 4, 1, 2, dx, dy)
            if _compatible(1, 2, 4, 4, [(dx, dy)]):
                assert abs(z.real - 1.0) < 1e-14 and abs(z.imag) < 1e-14
            else:
                assert abs(abs(z) - 1.0) < 1e-14
    # K-basis dimensions of the seeds {vacuum, single up, row stripe}
    seeds = [(0, 0), (1, 0), (0x000F, 0)]
    stabs = []
    for u, d in seeds:
        rep = canonical(4, 4, u, d)[0]
        stabs.append(_stabiliser(translate, 4, 4, rep >> 16, rep & 0xFFFF))
    dims = {K: sum(_compatible(*K, 4, 4, s) for s in stabs) for K in [(0, 0), (1, 0), (0, 1)]}
    assert dims == {(0, 0): 3, (1, 0): 1, (0, 1): 2}, dims
    assert not _compatible(0, 1, 4, 4, stabs[0])  # the vacuum is Gamma-only


def check_momentum_partition_2x2(translate, canonical, sector):
    for nu in range(5):
        for nd in range(5):
            ups, dns = sector(4, nu, nd)
            dim = len(ups)
            if dim == 0:
                continue
            reps = {canonical(2, 2, int(u), int(d))[0] for u, d in zip(ups, dns)}
            stabs = [_stabiliser(translate, 2, 2, r >> 16, r & 0xFFFF) for r in reps]
            sum_k = sum(_compatible(kx, ky, 2, 2, s) for ky in range(2) for kx in range(2) for s in stabs)
            sum_orb = sum(len(_orbit(translate, 2, 2, r >> 16, r & 0xFFFF)) for r in reps)
            assert sum_k == dim and sum_orb == dim, (nu, nd, dim, sum_k, sum_orb)


def dense_translation(oracle, lx, ly, up, dn, dx, dy):
    """Dense fermionic translation T(dx, dy) on one (N_up, N_dn) sector (column i holds
    eta * |T s_i>), from the oracle's translate / translation_sign."""
    import numpy as np

    index = {(int(u) << 16) | int(d): i for i, (u, d) in enumerate(zip(up, dn))}
    T = np.zeros((len(up), len(up)))
    for i, (u, d) in enumerate(zip(up, dn)):
        j = index[int(oracle.translate(int(u), int(d), lx, ly, dx, dy))]
        T[j, i] = oracle.translation_sign(int(u), int(d), lx, ly, dx, dy)
    return T


def projector_range(oracle, lx, ly, nu, nd, kx, ky):
    """Orthonormal basis (columns) of the range of P_K = (1/N) sum_t e^{-iK.t} T(t), the dense
    translation projector of tests/test_k_gram_mismatch_scan.cpp:105-114 (eigenvalues > 0.5)."""
    import numpy as np

    up, dn = oracle.basis_masks(lx * ly, nu, nd)
    P = sum(np.exp(-2j * math.pi * (kx * dx / lx + ky * dy / ly)) * dense_translation(oracle, lx, ly, up, dn, dx, dy)
            for dy in range(ly) for dx in range(lx)) / (lx * ly)
    w, Q = np.linalg.eigh((P + P.conj().T) / 2)
    return Q[:, w > 0.5]
