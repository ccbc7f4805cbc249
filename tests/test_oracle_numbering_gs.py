"""Pins the oracle's setvert3d numbering (navier8.f:2004-2360) and its
gather-scatter (dssum.f), and checks the product's host numbering
(nkh_setvert3d in libnekamd.so, no GPU needed) against it bit for bit.

Independent routes:
  * two local nodes share a nonzero id  <=>  they sit at the same physical
    point (periodic images identified); element interiors have id 0;
  * the counting rules re-derived from coordinates: on every edge the ids
    grow away from the smaller vertex id; on every face they start next to the
    smallest corner and run fastest toward its smaller neighbour corner;
  * the edge and face blocks are ordered lexicographically by their vertex
    tuples (np.unique on geometric tuples), vertices keep their ids;
  * dssum(u) = Q Q^T u with Q the dense id-incidence matrix; dssum(1) is the
    geometric multiplicity.
"""
import numpy as np
import pytest

from nek5000_amd.mesh import BoxMesh, setvert3d as product_setvert3d
from oracle import oracle as O

CASES = [
    dict(nel=(2, 2, 2), lx1=4),
    dict(nel=(3, 2, 2), lx1=5),
    dict(nel=(4, 3, 2), lx1=6, periodic_x=True),
    dict(nel=(3, 3, 3), lx1=8),
    dict(nel=(2, 2, 3), lx1=3),
    dict(nel=(2, 2, 2), lx1=2),
    # every element's frame turned by a seeded rotation of the cube and the
    # vertex ids permuted: the smallest corner of a face lands anywhere, so
    # every orientation branch of the edge and face numbering is taken
    # (navier8.f:2113-2312, including ind(1) = 3, 4 at :2226-2238)
    dict(nel=(3, 3, 2), lx1=5, rotate=1),
    dict(nel=(3, 3, 3), lx1=6, rotate=2, jitter=0.1),
    dict(nel=(2, 3, 2), lx1=4, rotate=3, relabel_only=True),
]


def make(case):
    nel = case["nel"]
    kw = {k: v for k, v in case.items() if k not in ("nel", "lx1", "rotate", "relabel_only")}
    b = BoxMesh(nel[0], nel[1], nel[2], case["lx1"], **kw)
    if "rotate" in case:
        from meshes import RotatedBox
        b = RotatedBox(b, seed=case["rotate"], rotate=not case.get("relabel_only", False))
    return b


def key_points(b):
    x, y, z = b.xyz()
    L = b.lengths
    xx = np.mod(x, L[0]) if b.periodic_x else x
    q = np.round(np.stack([xx, y, z], -1).reshape(-1, 3) * 1e9).astype(np.int64)
    if b.periodic_x:
        q[:, 0] = np.where(q[:, 0] == int(round(L[0] * 1e9)), 0, q[:, 0])
    return q


@pytest.mark.parametrize("case", CASES)
def test_ids_match_geometry(case):
    b = make(case)
    glo, ngv = O.setvert3d(b.vertex, b.lx1)
    n = b.lx1
    loc = glo.reshape(b.nel, n, n, n)
    interior = np.zeros((n, n, n), bool)
    if n > 2:
        interior[1:-1, 1:-1, 1:-1] = True
    assert np.all(loc[:, interior] == 0)
    assert np.all(loc[:, ~interior] > 0)
    q = key_points(b)
    bnd = (glo != 0)
    _, pt_group = np.unique(q[bnd], axis=0, return_inverse=True)
    _, id_group = np.unique(glo[bnd], return_inverse=True)
    # the two partitions are identical: a bijection between group labels
    pairs = np.unique(np.stack([pt_group.ravel(), id_group.ravel()], 1), axis=0)
    assert len(pairs) == pt_group.max() + 1 == id_group.max() + 1
    assert glo.max() == ngv


@pytest.mark.parametrize("case", CASES)
def test_product_numbering_bit_exact(case):
    b = make(case)
    g1, n1 = O.setvert3d(b.vertex, b.lx1)
    g2, n2 = product_setvert3d(b.vertex, b.lx1)
    assert n1 == n2
    assert np.array_equal(g1, g2)


def _re2(name, lx1):
    import os
    from nek5000_amd.mesh import Re2Mesh
    return Re2Mesh(os.path.join(os.path.dirname(__file__), "golden", "re2", name), lx1)


@pytest.mark.parametrize("case", [c for c in CASES if c["lx1"] > 2]
                         + [dict(re2="ethier.re2", lx1=5), dict(re2="lpm_one_uniform.re2",
                                                                 lx1=4)])
def test_counting_rules(case):
    """(On the reference's own short-test meshes too: their element frames
    and vertex ids are whatever genbox/prenek wrote.)"""
    b = _re2(case["re2"], case["lx1"]) if "re2" in case else make(case)
    n = b.lx1
    glo, ngv = O.setvert3d(b.vertex, n)
    loc = glo.reshape(b.nel, n, n, n)   # [e, k, j, i]
    V = b.vertex                        # [e, k, j, i] in {0,1}^3
    nv = int(V.max())
    # vertices keep their ids
    for (k, j, i) in np.ndindex(2, 2, 2):
        assert np.array_equal(loc[:, k * (n - 1), j * (n - 1), i * (n - 1)], V[:, k, j, i])
    noe = n - 2
    # edges: collect geometric edges as (min id, max id) with the id path
    edges = {}
    lin = np.arange(1, n - 1)
    for e in range(b.nel):
        for a, bb in (((0, 0, 0), (0, 0, 1)), ((0, 1, 0), (0, 1, 1)), ((1, 0, 0), (1, 0, 1)),
                      ((1, 1, 0), (1, 1, 1)), ((0, 0, 0), (0, 1, 0)), ((0, 0, 1), (0, 1, 1)),
                      ((1, 0, 0), (1, 1, 0)), ((1, 0, 1), (1, 1, 1)), ((0, 0, 0), (1, 0, 0)),
                      ((0, 0, 1), (1, 0, 1)), ((0, 1, 0), (1, 1, 0)), ((0, 1, 1), (1, 1, 1))):
            va, vb = V[e][a], V[e][bb]
            # node path from corner a to corner bb
            pa = np.array(a) * (n - 1)
            pb = np.array(bb) * (n - 1)
            path = [tuple(pa + (pb - pa) * t // (n - 1)) for t in lin]
            ids = np.array([loc[e][p] for p in path])
            if va > vb:
                ids = ids[::-1]
                va, vb = vb, va
            # grows away from the smaller vertex id by 1 per node
            assert np.array_equal(np.diff(ids), np.ones(noe - 1, dtype=np.int64))
            edges.setdefault((int(va), int(vb)), set()).add(int(ids[0]))
    keys = sorted(edges)
    for r, kk in enumerate(keys):
        assert edges[kk] == {nv + noe * r + 1}  # lexicographic rank, block of n-2
    ngv_e = nv + noe * len(keys)
    # faces: start next to the min corner, run fastest toward its smaller neighbour
    faces = {}
    for e in range(b.nel):
        for axis in range(3):
            for side in (0, 1):
                # corners of this face as (k,j,i) tuples in {0,1}
                cs = [c for c in np.ndindex(2, 2, 2) if c[2 - axis] == side]
                ids = {c: int(V[e][c]) for c in cs}
                c0 = min(cs, key=lambda c: ids[c])
                nb = [c for c in cs if sum(abs(np.array(c) - np.array(c0))) == 1]
                c1, c2 = sorted(nb, key=lambda c: ids[c])
                d1 = np.array(c1) - np.array(c0)
                d2 = np.array(c2) - np.array(c0)
                start = np.array(c0) * (n - 1)
                got = np.zeros((noe, noe), dtype=np.int64)
                for a_ in range(noe):
                    for b_ in range(noe):
                        # move along c0->c1 by a_+1 and along c0->c2 by b_+1
                        p = start + (d1 * (a_ + 1)) + (d2 * (b_ + 1))
                        got[b_, a_] = loc[e][tuple(p)]
                flat = got.reshape(-1)
                assert np.array_equal(np.diff(flat), np.ones(noe * noe - 1, dtype=np.int64))
                tup = tuple(sorted(ids.values())[:3])
                faces.setdefault(tup, set()).add(int(flat[0]))
    fk = sorted(faces)
    for r, kk in enumerate(fk):
        assert faces[kk] == {ngv_e + noe * noe * r + 1}
    assert ngv == ngv_e + noe * noe * len(fk)


@pytest.mark.parametrize("case", CASES[:4])
def test_gs_dense_and_multiplicity(case):
    b = make(case)
    glo, _ = O.setvert3d(b.vertex, b.lx1)
    gs = O.GS(glo)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(glo.size)
    ids, inv = np.unique(glo, return_inverse=True)
    Q = np.zeros((glo.size, len(ids)))
    Q[np.arange(glo.size), inv] = 1.0
    nz = glo != 0
    ref = np.where(nz, Q @ (Q.T @ u), u)
    got = gs.op(u.copy(), "+")
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-14)
    one = gs.op(np.ones(glo.size), "+")
    q = key_points(b)
    _, inv_pt, cnt = np.unique(q, axis=0, return_inverse=True, return_counts=True)
    geo_mult = np.where(nz, cnt[inv_pt.ravel()], 1)
    assert np.array_equal(one.astype(np.int64), geo_mult)
    # min/max/product agree with dense group reductions
    for op, fn in (("m", np.minimum), ("M", np.maximum), ("*", np.multiply)):
        got = gs.op(u.copy(), op)
        grp = np.full(len(ids), np.nan)
        for gi in range(len(ids)):
            sel = inv == gi
            grp[gi] = fn.reduce(u[sel])
        ref = np.where(nz, grp[inv], u)
        np.testing.assert_allclose(got, ref, rtol=1e-13)
