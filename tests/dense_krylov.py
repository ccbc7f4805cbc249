"""The reference's Krylov solvers restated on dense global matrices (test
infrastructure): cggo (hmholtz.f:611-846), chktcg1 (hmholtz.f:527-609),
hmh_flex_cg (hmholtz.f:2164-2281) and hmh_gmres (gmres.f:304-545).

Global vectors live on the lattice numbering of dense_mg.Level (not
setvert3d).  Nek's multiplicity-weighted sums over element-local copies are
plain sums over global nodes: for continuous a, b, sum_local a b mult =
sum_global a b; binvm1 = 1/B with B the assembled mass diagonal; the
operator is the assembled matrix A = sum_e Q_e A_e Q_e^T of the element
operator (axhelm, pinned on its own), the mask a diagonal 0/1 matrix.
"""
import numpy as np


def element_matrices(axhelm, nel, n, h1, h2):
    """A_e[e] (n^3 x n^3) from the element-local operator: column q is the
    operator applied to the q-th unit vector in every element at once."""
    n3 = n ** 3
    A = np.zeros((nel, n3, n3))
    u = np.zeros((nel, n3))
    for q in range(n3):
        u[:] = 0.0
        u[:, q] = 1.0
        A[:, :, q] = axhelm(u.reshape(nel, n, n, n), h1, h2).reshape(nel, n3)
    return A


def assemble(lev, Ae):
    N = lev.N
    A = np.zeros((N, N))
    for e in range(Ae.shape[0]):
        g = lev.gid[e].reshape(-1)
        np.add.at(A, (g[:, None], g[None, :]), Ae[e])
    return A


class System:
    """Dense global Helmholtz system of one field."""

    def __init__(self, lev, axhelm, bm1, h1, h2, mask_global):
        n = lev.nh
        nel = lev.gid.shape[0]
        self.lev = lev
        self.Ae = element_matrices(axhelm, nel, n, h1, h2)
        self.A = assemble(lev, self.Ae)
        self.B = lev.assemble(bm1)
        self.h1, self.h2 = h1, h2
        self.vol = float(np.sum(bm1))
        self.mask = mask_global
        self.h2max = float(np.max(h2))
        self.bm1 = bm1

    def setprec(self, g):
        """hmholtz.f:380-524 SETPREC on the element geometric factors g
        (g1..g6): the g1..g3 part of the diagonal of the unit-coefficient
        stiffness, the g4..g6 cross terms only at the element corners (the
        loops 600/700/800 visit r, s, t in {1, lx1} pairs there), times h1 at
        the node, plus h2 bm1; assembled and inverted."""
        from dense_mg import deriv, gll
        n = self.lev.nh
        D = deriv(gll(n)[0])
        g1, g2, g3, g4, g5, g6 = g
        d = (np.einsum("ekjq,qi->ekji", g1, D * D) + np.einsum("ekqi,qj->ekji", g2, D * D)
             + np.einsum("eqji,qk->ekji", g3, D * D))
        N = n - 1
        for iy in (0, N):
            for iz in (0, N):
                for ix in (0, N):
                    d[:, iz, iy, ix] += (g4[:, iz, iy, ix] * D[ix, ix] * D[iy, iy]
                                         + g5[:, iz, iy, ix] * D[ix, ix] * D[iz, iz])
        for ix in (0, N):
            for iz in (0, N):
                for iy in (0, N):
                    d[:, iz, iy, ix] += (g4[:, iz, iy, ix] * D[iy, iy] * D[ix, ix]
                                         + g6[:, iz, iy, ix] * D[iy, iy] * D[iz, iz])
        for ix in (0, N):
            for iy in (0, N):
                for iz in (0, N):
                    d[:, iz, iy, ix] += (g5[:, iz, iy, ix] * D[iz, iz] * D[ix, ix]
                                         + g6[:, iz, iy, ix] * D[iz, iz] * D[iy, iy])
        d = d * self.h1 + self.h2 * self.bm1
        self.dinv = 1.0 / self.lev.assemble(d)
        return self.dinv

    def Am(self, p):
        """mask * dssum(axhelm(p)) on a global vector"""
        return self.mask * (self.A @ p)


def chktcg1(S, tol, res):
    """hmholtz.f:527-609 on global vectors (imesh 1, EIGAA = 0 so the
    condition-number estimate is 10)."""
    eps, acondno = 1e-13, 10.0
    rinit = np.sqrt(np.sum(res * res / S.B) / S.vol)
    tol = max(tol, eps * rinit)
    bctest = abs(np.sum(S.mask) - np.sum(np.ones_like(S.mask)))
    one = np.ones(S.Ae.shape[1])
    a1 = np.einsum("eij,j->ei", S.Ae, one)  # element-local A_e 1, unassembled
    bcrob = np.sqrt(np.sum(a1 * a1 * S.bm1.reshape(a1.shape)) / S.vol)
    if bctest < 0.1 and bcrob < eps * acondno:
        tol = max(tol, rinit * eps * 10.0)
    return tol


def cggo(S, f, tin, maxit, istep=1, param18=0.0, param22=1e-7, restol=1e-7):
    """Jacobi PCG as cggo (non-PRES).  Returns (x, niter, history, rbn0);
    history rows (rbn2, rtz1, rho, alpha, beta), rho/alpha/beta 0 on the
    exit iteration."""
    tol = abs(tin)
    if restol != 0:
        tol = restol
    if tin < 0:
        tol = abs(tin)
    niter = min(maxit, 900)
    dinv = S.dinv
    r = f.copy()
    x = np.zeros_like(f)
    p = np.zeros_like(f)
    ifmcor = np.min(S.mask) > 0 and S.h2max == 0
    if ifmcor:
        smean = -1.0 / S.vol
        r = r + smean * np.sum(r) * S.B
    rtz1, rho = 1.0, 0.0
    hist = []
    rbn0 = 0.0
    it = 0
    for it in range(1, niter + 1):
        z = r * dinv
        if ifmcor:
            z = z + smean * np.sum(z * S.B)
        rtz2 = rtz1
        rtz1 = float(np.sum(z * r))
        s2 = np.sum(r * r) if param18 == 1 else np.sum(r * r / S.B)
        rbn2 = np.sqrt(s2 / S.vol)
        if it == 1:
            rbn0 = rbn2
        if param22 < 0:
            tol = abs(param22) * rbn0
        if tin < 0:
            tol = abs(tin) * rbn0
        if rbn2 <= tol and (it > 1 or istep <= 5):
            hist.append((rbn2, rtz1, 0.0, 0.0, 0.0))
            return x, it - 1, np.array(hist), rbn0
        beta = 0.0 if it == 1 else rtz1 / rtz2
        p = beta * p + z
        w = S.Am(p)
        rho = float(np.sum(w * p))
        alpha = rtz1 / rho
        x = x + alpha * p
        r = r - alpha * w
        hist.append((rbn2, rtz1, rho, alpha, beta))
    return x, it, np.array(hist), rbn0


def ortho(x, count):
    """navier1.f:223-255 ORTHO on a continuous vector held globally: the
    local-node mean, sum(count * x) / sum(count) with count the number of
    local copies of each global node, is subtracted."""
    return x - float(np.sum(count * x)) / float(np.sum(count))


def flex_cg(S, M, res, maxit, param21, istep=1, null_count=None):
    """hmh_flex_cg with the preconditioner M (a function of a global
    vector): (solution, iterations, rnorm history, div0).  null_count: the
    local-copy count of each global node when the constant null space is
    present (then ortho at the end, hmholtz.f:2272)."""
    tolps = chktcg1(S, 0.0, res)
    if param21 > 0 and tolps > abs(param21):
        tolps = abs(param21)
    if istep == 0:
        tolps = 1e-4
    tolpss = tolps
    r = res.copy()
    r1 = np.zeros_like(r)
    p = np.zeros_like(r)
    x = np.zeros_like(r)
    rho1 = 1.0
    div0 = np.sqrt(np.sum(r * r) / S.vol)
    if param21 < 0:
        tolpss = abs(param21) * div0
    hist = []
    it = 0
    for _ in range(maxit):
        z = M(r)
        r1 = r1 - r
        rho0 = rho1
        rho1 = float(np.sum(z * r))
        rho2 = -float(np.sum(z * r1))
        beta = rho2 / rho0
        r1 = r.copy()
        p = beta * p + z
        w = S.Am(p)
        den = float(np.sum(w * p))
        alpha = rho1 / den
        x = x + alpha * p
        r = r - alpha * w
        rnorm = np.sqrt(np.sum(r * r) / S.vol)
        it += 1
        hist.append(rnorm)
        if rnorm < tolpss:
            break
    if null_count is not None:
        x = ortho(x, null_count)
    return x, it, np.array(hist), div0


def gmres(S, M, res, maxit, param21, m=30, istep=1, null_count=None):
    """hmh_gmres (ml = mu = 1, as uzawa_gmres_split sets them):
    (solution, iterations, rnorm history, div0)."""
    norm_fac = 1.0 / np.sqrt(S.vol)
    tolps = chktcg1(S, 0.0, res)
    if param21 > 0 and tolps > abs(param21):
        tolps = abs(param21)
    if istep == 0:
        tolps = 1e-4
    tolpss = tolps
    x = np.zeros_like(res)
    it = 0
    hist = []
    div0 = None
    conv = False
    while not conv:
        if it == 0:
            r = res.copy()
        else:
            r = res - S.Am(x)
        gamma = np.zeros(m + 1)
        gamma[0] = np.sqrt(np.sum(r * r))
        if it == 0:
            div0 = gamma[0] * norm_fac
            if param21 < 0:
                tolpss = abs(param21) * div0
        if gamma[0] == 0:
            break
        V = [r / gamma[0]]
        Z = []
        H = np.zeros((m + 1, m))
        c = np.zeros(m)
        s = np.zeros(m)
        j = 0
        for j in range(m):
            it += 1
            z = M(V[j])
            if null_count is not None:
                z = ortho(z, null_count)  # gmres.f:427
            Z.append(z)
            w = S.Am(z)
            for i in range(j + 1):
                H[i, j] = float(np.sum(w * V[i]))
            for i in range(j + 1):
                w = w - H[i, j] * V[i]
            for i in range(j):
                t = H[i, j]
                H[i, j] = c[i] * t + s[i] * H[i + 1, j]
                H[i + 1, j] = -s[i] * t + c[i] * H[i + 1, j]
            alpha = np.sqrt(np.sum(w * w))
            if alpha == 0:
                conv = True
                break
            lj = np.sqrt(H[j, j] ** 2 + alpha ** 2)
            c[j] = H[j, j] / lj
            s[j] = alpha / lj
            H[j, j] = lj
            gamma[j + 1] = -s[j] * gamma[j]
            gamma[j] = c[j] * gamma[j]
            rnorm = abs(gamma[j + 1]) * norm_fac
            hist.append(rnorm)
            if it + 1 > maxit or rnorm < tolpss:
                conv = True
                break
            if j == m - 1:
                break
            V.append(w / alpha)
        jj = j + 1
        y = np.zeros(jj)
        for k in range(jj - 1, -1, -1):
            t = gamma[k]
            for i in range(jj - 1, k, -1):
                t -= H[k, i] * y[i]
            y[k] = t / H[k, k]
        for i in range(jj):
            x = x + y[i] * Z[i]
    if null_count is not None:
        x = ortho(x, null_count)  # gmres.f:534
    return x, it, np.array(hist), div0
