"""Test meshes beyond the lexicographic box: the same box with every
element's local frame turned by one of the 24 proper rotations of the cube
and its global vertex ids relabelled by a permutation.  The physical mesh is
unchanged; only the local (r,s,t) axes of each element and the numbers on
its vertices change, so the orientation branches of the numbering
(navier8.f:2113-2312) and every face/edge/vertex alignment of the
gather-scatter and of the Schwarz overlap (hsmg.f:440-494) are exercised.

Each element also reports the lattice position of any of its local nodes at
a given number of points per direction: this is the independent global
numbering the dense multigrid model in dense_mg.py uses (no setvert3d).
"""
import itertools

import numpy as np

from nek5000_amd.mesh import BoxMesh, _HexMesh


def proper_rotations():
    """The 24 signed permutations of the axes with determinant +1, as
    (perm, sign): new local axis b runs along old axis perm[b], in the
    direction sign[b]."""
    out = []
    for perm in itertools.permutations(range(3)):
        for sign in itertools.product((1, -1), repeat=3):
            M = np.zeros((3, 3), dtype=int)
            for b in range(3):
                M[perm[b], b] = sign[b]
            if round(np.linalg.det(M)) == 1:
                out.append((tuple(perm), tuple(sign)))
    assert len(out) == 24
    return out


def old_index(new_idx, perm, sign, n):
    """Old local index (i, j, k) of the node with new local index new_idx
    (each in 0..n-1, or -1 / n for the ghost layer of a Schwarz patch)."""
    old = [0, 0, 0]
    for b in range(3):
        v = new_idx[b]
        old[perm[b]] = v if sign[b] > 0 else (n - 1) - v
    return tuple(old)


def _reorient(a, perm, sign):
    """a: (..., n, n, n) indexed [k, j, i] in the old frame -> the same array
    indexed [k', j', i'] in the new frame."""
    n = a.shape[-1]
    idx = np.indices((n, n, n))  # [0]=k', [1]=j', [2]=i'
    new = [idx[2], idx[1], idx[0]]  # new axis b index arrays (r, s, t)
    old = [None, None, None]
    for b in range(3):
        old[perm[b]] = new[b] if sign[b] > 0 else (n - 1) - new[b]
    return a[..., old[2], old[1], old[0]]


class RotatedBox(_HexMesh):
    """A BoxMesh whose elements have rotated local frames (rotation index per
    element from a seed, or all 0 = identity) and whose vertex ids are
    relabelled by a seeded permutation (relabel=False keeps them)."""

    def __init__(self, box: BoxMesh, seed=0, rotate=True, relabel=True):
        assert not box.periodic_x, "periodic boxes are not rotated here"
        self.box = box
        self.periodic_x = False
        self.lx1 = box.lx1
        self.nel = box.nel
        self.h = box.h
        self.lengths = box.lengths
        rng = np.random.default_rng(seed)
        rots = proper_rotations()
        self.rot = rng.integers(0, 24, box.nel) if rotate else np.zeros(box.nel, dtype=int)
        self.perm = [rots[r][0] for r in self.rot]
        self.sign = [rots[r][1] for r in self.rot]
        v = box.vertex
        nv = int(v.max())
        lab = np.arange(nv + 1, dtype=np.int64)
        if relabel:
            lab[1:] = 1 + rng.permutation(nv)
        self._vertex = np.stack([_reorient(lab[v[e]], self.perm[e], self.sign[e])
                                 for e in range(self.nel)]).astype(np.int64)
        self._xyz = None
        e = np.arange(self.nel)
        self.ex = e % box.nelx
        self.ey = (e // box.nelx) % box.nely
        self.ez = e // (box.nelx * box.nely)

    @property
    def n(self):
        return self.nel * self.lx1 ** 3

    @property
    def vertex(self):
        return self._vertex

    def xyz(self):
        if self._xyz is None:
            x, y, z = self.box.xyz()
            self._xyz = tuple(
                np.ascontiguousarray(np.stack([_reorient(a[e], self.perm[e], self.sign[e])
                                               for e in range(self.nel)]))
                for a in (x, y, z))
        return self._xyz

    def dirichlet_faces(self):
        """Preprocessor face order, in each element's new frame."""
        old = self.box.dirichlet_faces()  # (nel, 6) preprocessor order
        sym_old = old[:, list(self.SYM_TO_PREP)]  # symmetric order r-, r+, s-, s+, t-, t+
        sym_new = np.zeros_like(sym_old)
        for e in range(self.nel):
            for b in range(3):
                for side in (0, 1):
                    a = self.perm[e][b]
                    s_old = side if self.sign[e][b] > 0 else 1 - side
                    sym_new[e, 2 * b + side] = sym_old[e, 2 * a + s_old]
        prep = np.zeros_like(old)
        prep[:, list(self.SYM_TO_PREP)] = sym_new
        return prep

    def face_codes(self, codes_old):
        """Per element symmetric-order face codes in the new frame, from codes
        given in the box frame (nel, 6) symmetric order."""
        out = np.zeros_like(codes_old)
        for e in range(self.nel):
            for b in range(3):
                for side in (0, 1):
                    a = self.perm[e][b]
                    s_old = side if self.sign[e][b] > 0 else 1 - side
                    out[e, 2 * b + side] = codes_old[e, 2 * a + s_old]
        return out

    def lattice(self, e, new_idx, nh):
        """Global lattice point (ix, iy, iz) of element e's node new_idx at nh
        points per direction (ghost indices -1 / nh reach into neighbours)."""
        o = old_index(new_idx, self.perm[e], self.sign[e], nh)
        return (self.ex[e] * (nh - 1) + o[0], self.ey[e] * (nh - 1) + o[1],
                self.ez[e] * (nh - 1) + o[2])

    def lattice_dims(self, nh):
        b = self.box
        return (b.nelx * (nh - 1) + 1, b.nely * (nh - 1) + 1, b.nelz * (nh - 1) + 1)

    def axis_global(self, e, b):
        """(global axis, sign) of element e's local axis b."""
        return self.perm[e][b], self.sign[e][b]
