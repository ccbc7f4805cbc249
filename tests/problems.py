"""Test problems built on the host with the oracle (no GPU involved).

The mesh arrays (vertex ids, GLL coordinates, Dirichlet faces) come from
nek5000_amd.mesh.BoxMesh -- they are the *inputs*; everything derived from
them here (geometry, numbering, gather-scatter, masks) is the oracle's.
"""
import numpy as np

from nek5000_amd.mesh import BoxMesh, hash_normal, hash_uniform
from oracle import oracle as O


def mask_from_fbc(fbc, n):
    """facev (connect1.f:1058) of the faces whose symmetric-order code is 1:
    ones with zeros on those faces, shape (nel, n, n, n) as [e, k, j, i]."""
    fbc = np.asarray(fbc).reshape(-1, 6)
    m = np.ones((fbc.shape[0], n, n, n))
    d = fbc == 1
    m[d[:, 0], :, :, 0] = 0.0
    m[d[:, 1], :, :, n - 1] = 0.0
    m[d[:, 2], :, 0, :] = 0.0
    m[d[:, 3], :, n - 1, :] = 0.0
    m[d[:, 4], 0, :, :] = 0.0
    m[d[:, 5], n - 1, :, :] = 0.0
    return m


class Problem:
    def __init__(self, box: BoxMesh, fbc=None):
        """fbc: optional (nel, 6) symmetric-order face codes (get_fast_bc);
        the Dirichlet mask then zeros the code-1 faces instead of every
        domain-boundary face."""
        self.box = box
        self.fbc = fbc
        n = box.lx1
        self.z, self.w = O.zwgll(n)
        self.D = O.dgll(self.z)
        self.x, self.y, self.zc = box.xyz()
        self.g, self.bm1, self.jac = O.geom(self.x, self.y, self.zc, self.D, self.w)
        self.glo, self.ngv = O.setvert3d(box.vertex, n)
        gs = O.GS(self.glo)
        mask = np.ascontiguousarray(box.face_mask() if fbc is None else mask_from_fbc(fbc, n))
        gs.op(mask, "*")
        self.mask = mask
        self.field = O.Field(self.D, self.g, self.bm1, self.glo, mask)
        self.mult = self.field.mult
        self.uid = box.unique_ids(self.glo, self.ngv)

    @property
    def shape(self):
        return self.bm1.shape

    def uniform_field(self, seed, stream=0):
        """U(-1,1) per unique node (same value at every local copy)."""
        return 2.0 * hash_uniform(seed, self.uid.reshape(-1), stream).reshape(self.shape) - 1.0

    def rhs_config1(self, seed=1, noise=1e-3):
        """f = 3 pi^2 sin(pi x) sin(pi y) sin(pi z) + noise*N(0,1) per unique node."""
        x, y, z = self.x, self.y, self.zc
        f = 3 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)
        if noise:
            f = f + noise * hash_normal(seed, self.uid.reshape(-1)).reshape(self.shape)
        return f

    def free_dof(self):
        return float(np.sum(self.mask * self.mult))


def box(nel, lx1, **kw):
    if isinstance(nel, int):
        nel = (nel, nel, nel)
    return BoxMesh(nel[0], nel[1], nel[2], lx1, **kw)


def mg_for(P, code=1):
    """Oracle multigrid hierarchy for a Problem (Dirichlet via outflow faces,
    or the Problem's own face codes)."""
    fbc = O.fbc_from_box(P.box, code) if P.fbc is None else np.ascontiguousarray(P.fbc, np.int32)
    return O.MG(P.field, P.box.vertex, fbc, P.x, P.y, P.zc, P.w)
