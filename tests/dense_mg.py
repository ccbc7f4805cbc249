"""Dense model of the H1 multigrid preconditioner (hsmg.f:1855-1949
h1mg_solve with sigma = 1) on small box meshes, built from its definition
with nothing of the oracle's multigrid code (test infrastructure).

What the reference computes, restated as global linear algebra:

  * level l carries the continuous piecewise polynomials with nh_l GLL
    points per direction (h1mg_setup_mg_nx, hsmg.f:2272-2323); global
    nodes are lattice points of the box, found from each element's place in
    the box and its local frame (meshes.RotatedBox.lattice), not from
    setvert3d;
  * Dirichlet mask of a level: a node is 0 when it lies on a face of code 1
    of any element (h1mg_setup_mask, hsmg.f:2426-2508: facev + dsprod);
  * Schwarz patch of an element (h1mg_schwarz_part1, hsmg.f:440-494): its own
    nodes plus, across every face with a neighbour, the neighbour's first
    interior layer at every tangential position of the face (the ghost
    faces of the (nh+2)^3 extended block; ghost edges and corners carry
    nothing);
  * local solve of a patch: the inverse of the separable extended operator
    A_t(x)B_s(x)B_r + B_t(x)A_s(x)B_r + B_t(x)B_s(x)A_r restricted to the kept
    nodes, where each 1-D pair (A, B) is the GLL stiffness and mass of the
    element and its two neighbours (lengths ll, lm, lr) assembled and
    restricted to the extended node set; a face of code 1 drops its ghost
    and the face node, a face of code 2 (Neumann) drops the ghost
    (hsmg_setup_fast1d(_a,_b), hsmg.f:783-882, with row_zero);
  * element lengths: the weighted harmonic mean of the squared distance
    between the first and last GLL planes over the face-interior nodes
    (plane_space, fast3d.f:306-423), the weight of 0-based node j being the
    1-based wxm1(j), i.e. the GLL weight of node j-1 (fast3d.f:326-333 read
    w(j) with j = 1..lx1-2 over an array indexed from 0); the neighbour's
    length is that of the element across the face along the same global
    axis (swap_lengths, fast3d.f:1542-1617);
  * Schwarz sum: every patch adds its local solution at the global nodes it
    covers; the sum is masked and each element-local copy multiplied by
    (1/c)^v, c = the number of patches covering the node, v = how many times
    the loops of hsmg_schwarz_wt3d (hsmg.f:1285-1319) visit that local index
    (twice for the middle index at nh = 3);
  * restriction = transpose of the continuous interpolation between
    levels; coarse level: vertices, the Galerkin matrix Phi^T A Phi of the
    fine stiffness (h1 = 1, h2 = 0) with the trilinear interpolation Phi
    (navier8.f:1648-1687), solved exactly on the unmasked vertices;
  * prolongation = interpolation of each element's local copy; the result
    is averaged over copies (dsavg, ic.f:1871-1897).

GLL points, weights, derivative and interpolation matrices come from
numpy.polynomial (roots of P'_N, 2/(N(N+1)P_N^2), barycentric formulas).
"""
import numpy as np
from numpy.polynomial import legendre as npleg


def gll(npts):
    N = npts - 1
    if N == 1:
        z = np.array([-1.0, 1.0])
    else:
        inner = np.sort(npleg.Legendre.basis(N).deriv().roots().real)
        z = np.concatenate([[-1.0], inner, [1.0]])
    PN = npleg.legval(z, [0] * N + [1])
    w = 2.0 / (N * (N + 1) * PN ** 2)
    return z, w


def _bary_w(z):
    n = len(z)
    w = np.ones(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                w[j] /= (z[j] - z[k])
    return w


def interp(zs, zd):
    """I[i, j] = l_j(zd[i]) for the Lagrange basis on zs."""
    w = _bary_w(zs)
    I = np.zeros((len(zd), len(zs)))
    for i, x in enumerate(zd):
        d = x - zs
        hit = np.where(np.abs(d) < 1e-15)[0]
        if len(hit):
            I[i, hit[0]] = 1.0
            continue
        t = w / d
        I[i] = t / t.sum()
    return I


def deriv(z):
    """D[i, j] = l_j'(z_i)."""
    n = len(z)
    w = _bary_w(z)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = (w[j] / w[i]) / (z[i] - z[j])
        D[i, i] = -D[i].sum()
    return D


def levels_nh(lx1):
    """Points per direction of every level, coarsest first (hsmg.f:2272-2323)."""
    mgn2 = [1, 2, 2, 2, 2, 3, 3, 5, 5, 5]
    lmax = 2 if lx1 == 4 else 3
    nx = [0] * (lmax + 1)
    nx[1] = 1
    lx2 = lx1
    m2 = 2 * (lx2 // 4) + 1
    if lx1 == 5:
        m2 = 3
    if lx1 <= 10:
        m2 = mgn2[min(lx1, 10) - 1]
    if lx1 == 8:
        m2 = 3
    m2 = min(3, m2)
    if lmax >= 3:
        nx[2] = m2
    nx[lmax] = lx1 - 1
    return [v + 1 for v in nx[1:]]


def wt_visits(n):
    """How many times hsmg_schwarz_wt3d (hsmg.f:1285-1319) multiplies local
    node (i, j, k) (0-based, array [k, j, i]) by its weight."""
    v = np.zeros((n, n, n), dtype=int)
    lay = [0, 1, n - 2, n - 1]
    mid = list(range(2, n - 2))
    for i in lay:
        v[:, :, i] += 1
    for j in lay:
        for i in mid:
            v[:, j, i] += 1
    for kk in lay:
        for j in mid:
            for i in mid:
                v[kk, j, i] += 1
    return v


class Level:
    def __init__(self, mesh, nh, codes):
        self.nh = nh
        nel = mesh.nel
        dims = mesh.lattice_dims(nh)
        self.dims = dims
        self.N = int(np.prod(dims))
        gid = np.zeros((nel, nh, nh, nh), dtype=np.int64)
        for e in range(nel):
            for k in range(nh):
                for j in range(nh):
                    for i in range(nh):
                        p = mesh.lattice(e, (i, j, k), nh)
                        gid[e, k, j, i] = p[0] + dims[0] * (p[1] + dims[1] * p[2])
        self.gid = gid
        mask = np.ones(self.N)
        for e in range(nel):
            for f in range(6):
                if codes[e, f] != 1:
                    continue
                b, side = divmod(f, 2)
                sl = [slice(None)] * 3
                sl[2 - b] = 0 if side == 0 else nh - 1
                mask[gid[e][tuple(sl)]] = 0.0
        self.mask = mask

    def gather(self, u):
        """global -> element-local copies"""
        return u[self.gid]

    def assemble(self, uloc):
        out = np.zeros(self.N)
        np.add.at(out, self.gid.reshape(-1), uloc.reshape(-1))
        return out

    def mult(self):
        return self.assemble(np.ones(self.gid.shape))


class DenseMG:
    def __init__(self, mesh, codes, xyz, axhelm, lengths_only=False):
        """mesh: RotatedBox (or anything with its lattice API); codes: (nel, 6)
        symmetric-order face codes in each element's frame (0 interior, 1
        Dirichlet, 2 Neumann); xyz: fine GLL coordinates (new frames);
        axhelm(u_local, h1, h2) -> the element-local fine operator."""
        self.mesh = mesh
        self.codes = np.asarray(codes)
        nel = mesh.nel
        lx1 = mesh.lx1
        self.nhs = levels_nh(lx1)
        if lengths_only:
            self._lengths(xyz)
            return
        self.lev = [Level(mesh, nh, self.codes) for nh in self.nhs]
        self.L = len(self.nhs)
        # interpolation from level l to l+1 (global) and per element
        self.Jloc = []
        self.P = []
        for l in range(self.L - 1):
            zc, _ = gll(self.nhs[l])
            zf, _ = gll(self.nhs[l + 1])
            J = interp(zc, zf)
            Jk = np.kron(np.kron(J, J), J)  # [(k,j,i) fine, (k,j,i) coarse]
            self.Jloc.append(Jk)
            Lc, Lf = self.lev[l], self.lev[l + 1]
            P = np.zeros((Lf.N, Lc.N))
            for e in range(nel):
                P[np.ix_(Lf.gid[e].reshape(-1), Lc.gid[e].reshape(-1))] = Jk
            self.P.append(P)
        self._lengths(xyz)
        self._patches()
        self._coarse(axhelm)

    # ---- element lengths (plane_space / swap_lengths)
    def _lengths(self, xyz):
        x, y, z = xyz
        n = self.mesh.lx1
        _, w = gll(n)
        nel = self.mesh.nel
        lm = np.zeros((nel, 3))
        inner = np.arange(1, n - 1)
        wq = w[inner - 1]  # weight of 0-based node j is wxm1(j) (1-based)
        W = np.outer(wq, wq)  # [b_idx, a_idx] over the two tangential indices
        for e in range(nel):
            X = np.stack([x[e], y[e], z[e]])  # [c, k, j, i]
            # r: planes i = 0 and i = n-1, tangential (j, k) interior
            d = X[:, 1:-1, 1:-1, n - 1] - X[:, 1:-1, 1:-1, 0]  # [c, k, j]
            lm[e, 0] = 1 / np.sqrt(np.sum(W / np.sum(d * d, 0)) / W.sum())
            d = X[:, 1:-1, n - 1, 1:-1] - X[:, 1:-1, 0, 1:-1]  # [c, k, i]
            lm[e, 1] = 1 / np.sqrt(np.sum(W / np.sum(d * d, 0)) / W.sum())
            d = X[:, n - 1, 1:-1, 1:-1] - X[:, 0, 1:-1, 1:-1]  # [c, j, i]
            lm[e, 2] = 1 / np.sqrt(np.sum(W / np.sum(d * d, 0)) / W.sum())
        self.lm = lm
        m = self.mesh
        # neighbour across each face along the same global axis
        pos = {(int(m.ex[e]), int(m.ey[e]), int(m.ez[e])): e for e in range(nel)}
        ll = np.zeros((nel, 3))
        lr = np.zeros((nel, 3))
        for e in range(nel):
            for b in range(3):
                g, sg = m.axis_global(e, b)
                for side, dst in ((0, ll), (1, lr)):
                    step = (-1 if side == 0 else 1) * sg
                    c = [int(m.ex[e]), int(m.ey[e]), int(m.ez[e])]
                    c[g] += step
                    e2 = pos.get(tuple(c))
                    if e2 is None:
                        dst[e, b] = 0.0
                        continue
                    b2 = [bb for bb in range(3) if m.axis_global(e2, bb)[0] == g][0]
                    dst[e, b] = lm[e2, b2]
        self.ll, self.lr = ll, lr

    # ---- Schwarz patches and local solves
    def _ext1d(self, N, ah, bh, ll, lm, lr, lbc, rbc):
        nl = N + 3
        A = np.zeros((nl, nl))
        B = np.zeros(nl)
        A[1:N + 2, 1:N + 2] += (2.0 / lm) * ah
        B[1:N + 2] += (lm / 2.0) * bh
        if lbc == 0:
            A[0:2, 0:2] += (2.0 / ll) * ah[N - 1:N + 1, N - 1:N + 1]
            B[0:2] += (ll / 2.0) * bh[N - 1:N + 1]
        if rbc == 0:
            A[N + 1:N + 3, N + 1:N + 3] += (2.0 / lr) * ah[0:2, 0:2]
            B[N + 1:N + 3] += (lr / 2.0) * bh[0:2]
        keep = np.ones(nl, bool)
        if lbc != 0:
            keep[0] = False
        if lbc == 1:
            keep[1] = False
        if rbc != 0:
            keep[N + 2] = False
        if rbc == 1:
            keep[N + 1] = False
        return A, np.diag(B), keep

    def local_inverse(self, lev, e):
        nh = self.nhs[lev]
        N = nh - 1
        z, w = gll(nh)
        D = deriv(z)
        ah = D.T @ np.diag(w) @ D
        bh = w
        mats = []
        for b in range(3):
            lbc, rbc = self.codes[e, 2 * b], self.codes[e, 2 * b + 1]
            mats.append(self._ext1d(N, ah, bh, self.ll[e, b], self.lm[e, b], self.lr[e, b],
                                    lbc, rbc))
        (Ar, Br, kr), (As, Bs, ks), (At, Bt, kt) = mats
        kron3 = lambda T, S, R: np.kron(np.kron(T, S), R)  # index (t, s, r), r fastest
        Aext = kron3(Bt, Bs, Ar) + kron3(Bt, As, Br) + kron3(At, Bs, Br)
        keep = kron3(kt.astype(float), ks.astype(float), kr.astype(float)) > 0
        F = np.zeros_like(Aext)
        F[np.ix_(keep, keep)] = np.linalg.inv(Aext[np.ix_(keep, keep)])
        return F

    def _patches(self):
        m = self.mesh
        self.patch = {}   # (level, e) -> (global ids (-1 = none), F)
        self.count = {}
        for lev in range(1, self.L):  # Schwarz on levels 2..lmax
            nh = self.nhs[lev]
            Lv = self.lev[lev]
            dims = Lv.dims
            nl = nh + 2
            cnt = np.zeros(Lv.N)
            for e in range(m.nel):
                ids = -np.ones((nl, nl, nl), dtype=np.int64)
                for c in range(nl):
                    for bq in range(nl):
                        for a in range(nl):
                            v = (a - 1, bq - 1, c - 1)
                            nvirt = sum(1 for t in v if t < 0 or t > nh - 1)
                            if nvirt >= 2:
                                continue
                            p = m.lattice(e, v, nh)
                            if not all(0 <= p[d] < dims[d] for d in range(3)):
                                continue
                            ids[c, bq, a] = p[0] + dims[0] * (p[1] + dims[1] * p[2])
                ids = ids.reshape(-1)
                F = self.local_inverse(lev, e)
                self.patch[(lev, e)] = (ids, F)
                np.add.at(cnt, ids[ids >= 0], 1.0)
            self.count[lev] = cnt

    def schwarz(self, lev, r):
        """Level lev (index into nhs) Schwarz of the global vector r (already
        masked): element-local copies of the weighted, masked sum."""
        Lv = self.lev[lev]
        acc = np.zeros(Lv.N)
        for e in range(self.mesh.nel):
            ids, F = self.patch[(lev, e)]
            rp = np.where(ids >= 0, r[np.maximum(ids, 0)], 0.0)
            out = F @ rp
            sel = ids >= 0
            np.add.at(acc, ids[sel], out[sel])
        acc *= Lv.mask
        loc = Lv.gather(acc)
        wl = Lv.gather(1.0 / self.count[lev])
        return loc * wl ** wt_visits(Lv.nh)[None]

    # ---- coarse Galerkin problem
    def _coarse(self, axhelm):
        n = self.mesh.lx1
        nel = self.mesh.nel
        z, _ = gll(n)
        ph = np.stack([(1 - z) / 2, (1 + z) / 2])  # [c, node]
        L1 = self.lev[0]
        Gc = np.zeros((L1.N, L1.N))
        basis = []
        for c in range(8):
            ci, cj, ck = c & 1, (c >> 1) & 1, (c >> 2) & 1
            b = np.einsum("k,j,i->kji", ph[ck], ph[cj], ph[ci])
            basis.append(np.broadcast_to(b, (nel, n, n, n)).copy())
        one = np.ones((nel, n, n, n))
        zero = np.zeros((nel, n, n, n))
        Ab = [axhelm(bb, one, zero) for bb in basis]
        for e in range(nel):
            G = np.array([[np.sum(basis[c2][e] * Ab[c][e]) for c in range(8)]
                          for c2 in range(8)])
            gi = L1.gid[e].reshape(-1)  # (k,j,i) in {0,1}: index c = i + 2j + 4k
            Gc[np.ix_(gi, gi)] += G
        free = L1.mask > 0
        # no Dirichlet face: the constant is in the null space; one unknown is
        # held at zero and the solution's mean removed after the solve
        # (crs_xxt.c:883-955).  The unknown held is the vertex whose id comes
        # last in order of first appearance over (element, corner) -- the
        # numbering of the coarse unknowns.
        self.null_space = not np.any(self.codes == 1)
        if self.null_space:
            vtx = np.asarray(self.mesh.vertex).reshape(nel, 8)
            seen, last = set(), None
            for e in range(nel):
                for c in range(8):
                    if vtx[e, c] not in seen:
                        seen.add(vtx[e, c])
                        last = (e, c)
            self.pin = int(L1.gid[last[0]].reshape(-1)[last[1]])
            free = free.copy()
            free[self.pin] = False
        self.free = free
        self.Ac = Gc
        self.Acinv = np.linalg.inv(Gc[np.ix_(free, free)])

    def solve(self, rhs_global):
        """z = M rhs (global continuous vectors)."""
        L = self.L
        Lf = self.lev[L - 1]
        r = {L - 1: rhs_global * Lf.mask}
        zloc = self.schwarz(L - 1, r[L - 1])
        eloc = {}
        for lev in range(L - 2, 0, -1):
            rl = self.P[lev].T @ r[lev + 1]
            rl = rl * self.lev[lev].mask
            r[lev] = rl
            eloc[lev] = self.schwarz(lev, rl)
        r1 = (self.P[0].T @ r[1]) * self.lev[0].mask
        e1 = np.zeros(self.lev[0].N)
        e1[self.free] = self.Acinv @ r1[self.free]
        if self.null_space:
            e1 = e1 - np.mean(e1)
        e1 = e1 * self.lev[0].mask
        eloc[0] = self.lev[0].gather(e1)
        for lev in range(1, L):
            nel = self.mesh.nel
            nhc, nhf = self.nhs[lev - 1], self.nhs[lev]
            up = (self.Jloc[lev - 1] @ eloc[lev - 1].reshape(nel, -1).T).T.reshape(
                nel, nhf, nhf, nhf)
            if lev == L - 1:
                zloc = zloc + up
            else:
                eloc[lev] = eloc[lev] + up
        # dsavg
        return Lf.assemble(zloc) / Lf.mult()
