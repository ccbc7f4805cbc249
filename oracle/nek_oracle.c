/*
 * nek_oracle.c -- CPU restatement of Nek5000's Helmholtz/Poisson solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * implementation in nek5000_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Every routine restates one reference routine (file:line given at each
 * function, paths relative to the Nek5000 source tree).  Arrays follow Nek's
 * layout: a field is (lx1,ly1,lz1,nel) column-major, i fastest, so node
 * (i,j,k) of element e sits at i + n*(j + n*(k + n*e)).  Small matrices
 * (D, interpolation) are stored column-major too: A(i,j) at A[i + ld*j].
 *
 * Parity status (what pins each function; the tests in tests/ implement it):
 *   zwgll/dgll/pnleg/igllm  : flang build of the reference speclib.f
 *                             (oracle/_ref), closed forms (P'_N roots, weights
 *                             2/(N(N+1)P_N^2), D exact on x^k).
 *   geom (xyzrst/glmapm1/geodat1): closed form on affine boxes; independent
 *                             inverse-Jacobian quadrature on deformed meshes.
 *   axhelm                  : Kronecker-sum closed form (affine), independent
 *                             quadrature of grad u . grad v (deformed),
 *                             symmetry, A*1 = 0.
 *   setvert3d               : geometric coincidence of GLL nodes; the
 *                             counting rules on edges and faces re-derived
 *                             geometrically, on boxes with rotated element
 *                             frames and permuted vertex ids and on the
 *                             reference's own short-test meshes.
 *   gs (dssum/dsop)         : dense Q Q^T on tiny meshes; dssum(1)=multiplicity.
 *   setprec                 : diagonal of the brute-force assembled matrix
 *                             (uniform coefficients); its h1-at-the-node and
 *                             corner-only cross terms restated densely.
 *   chktcg1                 : its three floors restated on global sums
 *                             (Dirichlet, Neumann with and without h2).
 *   cggo                    : iterate by iterate (rbn2, rtz1, rho, alpha,
 *                             beta) against Jacobi PCG on the dense assembled
 *                             global matrix with the same stop rule, norm and
 *                             null-space correction; convergence to the
 *                             analytic solution.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

extern int g_mutation; /* planted mistakes for tests (nek_oracle_hsmg.c) */
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX3(i, j, k, n) ((i) + (n) * ((j) + (n) * (k)))

/* ------------------------------------------------------------------ */
/* speclib.f : Gauss-Lobatto-Legendre points, weights, derivatives      */
/* ------------------------------------------------------------------ */

/* speclib.f:479-518 JACOBF: Jacobi polynomial of degree n and derivative at x */
static void jacobf(double *poly, double *pder, double *polym1, double *pderm1,
                   double *polym2, double *pderm2, int n, double alp,
                   double bet, double x) {
  double apb = alp + bet;
  double p = 1.0, pd = 0.0;
  double pl, pdl, psave = 0.0, pdsave = 0.0;
  *poly = p;
  *pder = pd;
  if (n == 0) return;
  pl = p;
  pdl = pd;
  p = (alp - bet + (apb + 2.0) * x) / 2.0;
  pd = (apb + 2.0) / 2.0;
  if (n == 1) {
    *poly = p;
    *pder = pd;
    return;
  }
  for (int k = 2; k <= n; k++) {
    double dk = (double)k;
    double a1 = 2.0 * dk * (dk + apb) * (2.0 * dk + apb - 2.0);
    double a2 = (2.0 * dk + apb - 1.0) * (alp * alp - bet * bet);
    double b3 = (2.0 * dk + apb - 2.0);
    double a3 = b3 * (b3 + 1.0) * (b3 + 2.0);
    double a4 = 2.0 * (dk + alp - 1.0) * (dk + bet - 1.0) * (2.0 * dk + apb);
    double polyn = ((a2 + a3 * x) * p - a4 * pl) / a1;
    double pdern = ((a2 + a3 * x) * pd - a4 * pdl + a3 * p) / a1;
    psave = pl;
    pdsave = pdl;
    pl = p;
    p = polyn;
    pdl = pd;
    pd = pdern;
  }
  *poly = p;
  *pder = pd;
  *polym1 = pl;
  *pderm1 = pdl;
  *polym2 = psave;
  *pderm2 = pdsave;
}

/* speclib.f:371-393 GAMMAF (table of the arguments the GLL code needs) */
static double gammaf(double x) {
  double pi = 4.0 * atan(1.0);
  double g = 1.0;
  if (x == -0.5) g = -2.0 * sqrt(pi);
  if (x == 0.5) g = sqrt(pi);
  if (x == 1.0) g = 1.0;
  if (x == 2.0) g = 1.0;
  if (x == 1.5) g = sqrt(pi) / 2.0;
  if (x == 2.5) g = 1.5 * sqrt(pi) / 2.0;
  if (x == 3.5) g = 0.5 * (2.5 * (1.5 * sqrt(pi)));
  if (x == 3.0) g = 2.0;
  if (x == 4.0) g = 6.0;
  if (x == 5.0) g = 24.0;
  if (x == 6.0) g = 120.0;
  return g;
}

/* speclib.f:395-419 PNORMJ */
static double pnormj(int n, double alpha, double beta) {
  double dn = (double)n;
  double cnst = alpha + beta + 1.0;
  double prod;
  if (n <= 1) {
    prod = gammaf(dn + alpha) * gammaf(dn + beta);
    prod = prod / (gammaf(dn) * gammaf(dn + alpha + beta));
    return prod * pow(2.0, cnst) / (2.0 * dn + cnst);
  }
  prod = gammaf(alpha + 1.0) * gammaf(beta + 1.0);
  prod = prod / (2.0 * (1.0 + cnst) * gammaf(cnst + 1.0));
  prod = prod * (1.0 + alpha) * (2.0 + alpha);
  prod = prod * (1.0 + beta) * (2.0 + beta);
  for (int i = 3; i <= n; i++) {
    double di = (double)i;
    double frac = (di + alpha) * (di + beta) / (di * (di + alpha + beta));
    prod = prod * frac;
  }
  return prod * pow(2.0, cnst) / (2.0 * dn + cnst);
}

/* speclib.f:421-477 JACG: zeros of the Jacobi polynomial, Newton with deflation */
static void jacg(double *xjac, int np, double alpha, double beta) {
  const int kstop = 10;
  const double eps = 1.0e-12;
  int n = np - 1;
  double dth = 4.0 * atan(1.0) / (2.0 * (double)n + 2.0);
  double xlast = 0.0;
  double p, pd, pm1, pdm1, pm2, pdm2;
  for (int j = 1; j <= np; j++) {
    double x;
    if (j == 1) {
      x = cos((2.0 * ((double)j - 1.0) + 1.0) * dth);
    } else {
      double x1 = cos((2.0 * ((double)j - 1.0) + 1.0) * dth);
      double x2 = xlast;
      x = (x1 + x2) / 2.0;
    }
    for (int k = 1; k <= kstop; k++) {
      jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, np, alpha, beta, x);
      double recsum = 0.0;
      for (int i = 1; i <= j - 1; i++) recsum = recsum + 1.0 / (x - xjac[np - i]);
      double delx = -p / (pd - recsum * p);
      x = x + delx;
      if (fabs(delx) < eps) break;
    }
    xjac[np - j] = x;
    xlast = x;
  }
  /* selection sort, ascending */
  for (int i = 0; i < np; i++) {
    double xmin = 2.0;
    int jmin = i;
    for (int j = i; j < np; j++) {
      if (xjac[j] < xmin) {
        xmin = xjac[j];
        jmin = j;
      }
    }
    if (jmin != i) {
      double s = xjac[i];
      xjac[i] = xjac[jmin];
      xjac[jmin] = s;
    }
  }
}

/* speclib.f:155-205 ZWGJD: Gauss-Jacobi points and weights */
static void zwgjd(double *z, double *w, int np, double alpha, double beta) {
  int n = np - 1;
  double apb = alpha + beta;
  double p, pd, pm1, pdm1, pm2, pdm2;
  if (np == 1) {
    z[0] = (beta - alpha) / (apb + 2.0);
    w[0] = gammaf(alpha + 1.0) * gammaf(beta + 1.0) / gammaf(apb + 2.0) *
           pow(2.0, apb + 1.0);
    return;
  }
  jacg(z, np, alpha, beta);
  int np1 = n + 1, np2 = n + 2;
  double dnp1 = (double)np1, dnp2 = (double)np2;
  double fac1 = dnp1 + alpha + beta + 1.0;
  double fac2 = fac1 + dnp1;
  double fac3 = fac2 + 1.0;
  double fnorm = pnormj(np1, alpha, beta);
  double rcoef = (fnorm * fac2 * fac3) / (2.0 * fac1 * dnp2);
  for (int i = 0; i < np; i++) {
    jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, np2, alpha, beta, z[i]);
    w[i] = -rcoef / (p * pdm1);
  }
}

/* speclib.f:283-325 ENDW1 and :327-369 ENDW2 (alpha=beta=0 paths are the ones used) */
static double endw(int n, double alpha, double beta, int which) {
  double apb = alpha + beta;
  if (n == 0) return 0.0;
  double f1, f2, f3 = 0.0, fint1, fint2;
  if (which == 1) {
    f1 = gammaf(alpha + 2.0) * gammaf(beta + 1.0) / gammaf(apb + 3.0);
  } else {
    f1 = gammaf(alpha + 1.0) * gammaf(beta + 2.0) / gammaf(apb + 3.0);
  }
  f1 = f1 * (apb + 2.0) * pow(2.0, apb + 2.0) / 2.0;
  if (n == 1) return f1;
  if (which == 1) {
    fint1 = gammaf(alpha + 2.0) * gammaf(beta + 1.0) / gammaf(apb + 3.0);
  } else {
    fint1 = gammaf(alpha + 1.0) * gammaf(beta + 2.0) / gammaf(apb + 3.0);
  }
  fint1 = fint1 * pow(2.0, apb + 2.0);
  fint2 = gammaf(alpha + 2.0) * gammaf(beta + 2.0) / gammaf(apb + 4.0);
  fint2 = fint2 * pow(2.0, apb + 3.0);
  if (which == 1)
    f2 = (-2.0 * (beta + 2.0) * fint1 + (apb + 4.0) * fint2) * (apb + 3.0) / 4.0;
  else
    f2 = (2.0 * (alpha + 2.0) * fint1 - (apb + 4.0) * fint2) * (apb + 3.0) / 4.0;
  if (n == 2) return f2;
  for (int i = 3; i <= n; i++) {
    double di = (double)(i - 1);
    double abn = alpha + beta + di;
    double abnn = abn + di;
    double a1 = -(2.0 * (di + alpha) * (di + beta)) / (abn * abnn * (abnn + 1.0));
    double a2 = (2.0 * (alpha - beta)) / (abnn * (abnn + 2.0));
    double a3 = (2.0 * (abn + 1.0)) / ((abnn + 2.0) * (abnn + 1.0));
    f3 = -(a2 * f2 + a1 * f1) / a3;
    f1 = f2;
    f2 = f3;
  }
  return f3;
}

/* speclib.f:238-281 ZWGLJD; speclib.f:107-122 ZWGLL (alpha=beta=0) */
void nko_zwgll(double *z, double *w, int np) {
  double alpha = 0.0, beta = 0.0;
  int n = np - 1, nm1 = n - 1;
  double p, pd, pm1, pdm1, pm2, pdm2;
  if (nm1 > 0) zwgjd(z + 1, w + 1, nm1, alpha + 1.0, beta + 1.0);
  z[0] = -1.0;
  z[np - 1] = 1.0;
  for (int i = 1; i < np - 1; i++) w[i] = w[i] / (1.0 - z[i] * z[i]);
  jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, n, alpha, beta, z[0]);
  w[0] = endw(n, alpha, beta, 1) / (2.0 * pd);
  jacobf(&p, &pd, &pm1, &pdm1, &pm2, &pdm2, n, alpha, beta, z[np - 1]);
  w[np - 1] = endw(n, alpha, beta, 2) / (2.0 * pd);
}

/* speclib.f:875-905 PNLEG: Legendre polynomial by three-term recursion */
double nko_pnleg(double z, int n) {
  if (fabs(z) < 1.0e-25) z = 0.0;
  double p1 = 1.0;
  if (n == 0) return p1;
  double p2 = z, p3 = p2;
  for (int k = 1; k <= n - 1; k++) {
    double fk = (double)k;
    p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

/* speclib.f:907-932 PNDLEG */
static double pndleg(double z, int n) {
  double p1 = 1.0, p2 = z, p1d = 0.0, p2d = 1.0, p3d = 1.0, p3;
  for (int k = 1; k <= n - 1; k++) {
    double fk = (double)k;
    p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p3d = ((2.0 * fk + 1.0) * p2 + (2.0 * fk + 1.0) * z * p2d - fk * p1d) / (fk + 1.0);
    p1 = p2;
    p2 = p3;
    p1d = p2d;
    p2d = p3d;
  }
  if (n == 0) return 0.0;
  return p3d;
}

/* speclib.f:800-833 DGLL: D(i,j) = phi_j'(z_i); column-major, ld = nz */
void nko_dgll(double *d, double *dt, const double *z, int nz) {
  int n = nz - 1;
  if (nz == 1) {
    d[0] = 0.0;
    dt[0] = 0.0;
    return;
  }
  double fn = (double)n;
  double d0 = fn * (fn + 1.0) / 4.0;
  for (int i = 0; i < nz; i++)
    for (int j = 0; j < nz; j++) {
      double v = 0.0;
      if (i != j) v = nko_pnleg(z[i], n) / (nko_pnleg(z[j], n) * (z[i] - z[j]));
      if (i == j && i == 0) v = -d0;
      if (i == j && i == nz - 1) v = d0;
      d[i + nz * j] = v;
      dt[j + nz * i] = v;
    }
}

/* speclib.f:835-854 HGLL: Lagrange interpolant through GLL points at z */
static double hgll(int i, double z, const double *zgll, int nz) {
  double eps = 1.0e-5;
  double dz = z - zgll[i];
  if (fabs(dz) < eps) return 1.0;
  int n = nz - 1;
  double alfan = (double)n * ((double)n + 1.0);
  return -(1.0 - z * z) * pndleg(z, n) / (alfan * nko_pnleg(zgll[i], n) * (z - zgll[i]));
}

/* speclib.f:1101-1124 IGLLM: I12(i,j) = h_j(z2_i), column-major ld = nz2 */
void nko_igllm(double *i12, double *it12, const double *z1, const double *z2,
               int nz1, int nz2) {
  if (nz1 == 1) {
    i12[0] = 1.0;
    it12[0] = 1.0;
    return;
  }
  for (int i = 0; i < nz2; i++)
    for (int j = 0; j < nz1; j++) {
      double v = hgll(j, z2[i], z1, nz1);
      i12[i + nz2 * j] = v;
      it12[j + nz1 * i] = v;
    }
}

/* ------------------------------------------------------------------ */
/* mxm_std.f:2-66 mxmf2: C(n1,n3) = A(n1,n2) B(n2,n3), column-major     */
/* ------------------------------------------------------------------ */
static void mxm(const double *a, int n1, const double *b, int n2, double *c, int n3) {
  for (int j = 0; j < n3; j++)
    for (int i = 0; i < n1; i++) {
      double s = 0.0;
      for (int k = 0; k < n2; k++) s += a[i + n1 * k] * b[k + n2 * j];
      c[i + n1 * j] = s;
    }
}

/* ------------------------------------------------------------------ */
/* coef.f : geometric factors                                          */
/* ------------------------------------------------------------------ */

/*
 * coef.f:879-927 XYZRST + coef.f:555-632 GLMAPM1 + coef.f:633-785 GEODAT1
 * (3-D, non-axisymmetric).  Outputs g1..g6 (g[0..5]), bm1, jacm1.
 * dxm1 is D (column-major), w is the 1-D GLL weight vector.
 */
void nko_geom(int n, int nel, const double *xm1, const double *ym1, const double *zm1,
              const double *dxm1, const double *w, double *g1, double *g2, double *g3,
              double *g4, double *g5, double *g6, double *bm1, double *jacm1) {
  int nxyz = n * n * n, nxy = n * n;
  double *dt = malloc(sizeof(double) * n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) dt[j + n * i] = dxm1[i + n * j];
  double *w3 = malloc(sizeof(double) * nxyz);
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) w3[IDX3(i, j, k, n)] = w[i] * w[j] * w[k];
#pragma omp parallel
  {
    double *xr = malloc(sizeof(double) * 9 * nxyz);
    double *yr = xr + nxyz, *zr = yr + nxyz, *xs = zr + nxyz, *ys = xs + nxyz,
           *zs = ys + nxyz, *xt = zs + nxyz, *yt = xt + nxyz, *zt = yt + nxyz;
#pragma omp for
    for (int e = 0; e < nel; e++) {
      const double *x = xm1 + (size_t)e * nxyz, *y = ym1 + (size_t)e * nxyz,
                   *z = zm1 + (size_t)e * nxyz;
      mxm(dxm1, n, x, n, xr, n * n);
      mxm(dxm1, n, y, n, yr, n * n);
      mxm(dxm1, n, z, n, zr, n * n);
      for (int iz = 0; iz < n; iz++) {
        mxm(x + iz * nxy, n, dt, n, xs + iz * nxy, n);
        mxm(y + iz * nxy, n, dt, n, ys + iz * nxy, n);
        mxm(z + iz * nxy, n, dt, n, zs + iz * nxy, n);
      }
      mxm(x, nxy, dt, n, xt, n);
      mxm(y, nxy, dt, n, yt, n);
      mxm(z, nxy, dt, n, zt, n);
      for (int q = 0; q < nxyz; q++) {
        size_t l = (size_t)e * nxyz + q;
        /* glmapm1: addcol4/subcol4 sequence starting from zero */
        double jac = 0.0;
        jac = jac + xr[q] * ys[q] * zt[q];
        jac = jac + xt[q] * yr[q] * zs[q];
        jac = jac + xs[q] * yt[q] * zr[q];
        jac = jac - xr[q] * yt[q] * zs[q];
        jac = jac - xs[q] * yr[q] * zt[q];
        jac = jac - xt[q] * ys[q] * zr[q];
        /* ascol5: cofactors (J times the inverse-map derivatives) */
        double rx = ys[q] * zt[q] - yt[q] * zs[q];
        double ry = xt[q] * zs[q] - xs[q] * zt[q];
        double rz = xs[q] * yt[q] - xt[q] * ys[q];
        double sx = yt[q] * zr[q] - yr[q] * zt[q];
        double sy = xr[q] * zt[q] - xt[q] * zr[q];
        double sz = xt[q] * yr[q] - xr[q] * yt[q];
        double tx = yr[q] * zs[q] - ys[q] * zr[q];
        double ty = xs[q] * zr[q] - xr[q] * zs[q];
        double tz = xr[q] * ys[q] - xs[q] * yr[q];
        double wj = 1.0 / jac; /* invers2 */
        double a1 = rx * rx + ry * ry + rz * rz;
        double a2 = sx * sx + sy * sy + sz * sz;
        double a3 = tx * tx + ty * ty + tz * tz;
        double a4 = rx * sx + ry * sy + rz * sz;
        double a5 = rx * tx + ry * ty + rz * tz;
        double a6 = sx * tx + sy * ty + sz * tz;
        g1[l] = a1 * wj * w3[q];
        g2[l] = a2 * wj * w3[q];
        g3[l] = a3 * wj * w3[q];
        g4[l] = a4 * wj * w3[q];
        g5[l] = a5 * wj * w3[q];
        g6[l] = a6 * wj * w3[q];
        bm1[l] = jac * w3[q];
        if (jacm1) jacm1[l] = jac;
      }
    }
    free(xr);
  }
  free(dt);
  free(w3);
}

/* ------------------------------------------------------------------ */
/* navier8.f : global numbering (setvert3d), serial rank semantics     */
/* ------------------------------------------------------------------ */

/* navier8.f:1131-1179 I8RANK: heap-sort index (Numerical Recipes indexx); 0-based out */
static void i8rank(const int64_t *a, int *ind, int n) {
  /* works on 1-based ind values internally to mirror the reference */
  int l, ir, i, j, indx;
  int64_t q;
  for (j = 1; j <= n; j++) ind[j - 1] = j;
  if (n <= 1) return;
  l = n / 2 + 1;
  ir = n;
  for (;;) {
    if (l > 1) {
      l = l - 1;
      indx = ind[l - 1];
      q = a[indx - 1];
    } else {
      indx = ind[ir - 1];
      q = a[indx - 1];
      ind[ir - 1] = ind[0];
      ir = ir - 1;
      if (ir == 1) {
        ind[0] = indx;
        return;
      }
    }
    i = l;
    j = l + l;
    while (j <= ir) {
      if (j < ir) {
        if (a[ind[j - 1] - 1] < a[ind[j] - 1]) j = j + 1;
      }
      if (q < a[ind[j - 1] - 1]) {
        ind[i - 1] = ind[j - 1];
        i = j;
        j = j + j;
      } else {
        j = ir + 1;
      }
    }
    ind[i - 1] = indx;
  }
}

typedef struct {
  int64_t k[3];
  int64_t pos;
} tup3;

static int cmp_tup3(const void *pa, const void *pb) {
  const tup3 *a = pa, *b = pb;
  for (int m = 0; m < 3; m++) {
    if (a->k[m] < b->k[m]) return -1;
    if (a->k[m] > b->k[m]) return 1;
  }
  return (a->pos < b->pos) ? -1 : (a->pos > b->pos);
}

/*
 * navier8.f:1934-2002 GBTUPLE_RANK8 at np=1 (crystal-router transfers are
 * the identity on one rank): rank = position of the tuple in the
 * lexicographically sorted list of distinct tuples, 1-based.
 */
static void tuple_rank(tup3 *t, int64_t n, int64_t *rank_out) {
  qsort(t, (size_t)n, sizeof(tup3), cmp_tup3);
  int64_t r = 0;
  for (int64_t i = 0; i < n; i++) {
    if (i == 0 || t[i].k[0] != t[i - 1].k[0] || t[i].k[1] != t[i - 1].k[1] ||
        t[i].k[2] != t[i - 1].k[2])
      r++;
    rank_out[t[i].pos] = r;
  }
}

/* connect1.f:546- DSSET skpdat (1-based indices as in the reference) */
static void dsset_skpdat(int nx, int skp[6][6]) {
  int ny = nx, nz = nx;
  int s[6][6] = {
      {1, nx * (ny - 1) + 1, nx, 1, ny * (nz - 1) + 1, ny},
      {1 + (nx - 1), nx * (ny - 1) + 1 + (nx - 1), nx, 1, ny * (nz - 1) + 1, ny},
      {1, nx, 1, 1, ny * (nz - 1) + 1, ny},
      {1 + nx * (ny - 1), nx + nx * (ny - 1), 1, 1, ny * (nz - 1) + 1, ny},
      {1, nx, 1, 1, ny, 1},
      {1 + nx * ny * (nz - 1), nx + nx * ny * (nz - 1), 1, 1, ny, 1}};
  memcpy(skp, s, sizeof(s));
}

/* TOPOL:37-42 ICFACE (3-D part), symmetric face notation */
static const int icface[6][4] = {{1, 3, 5, 7}, {2, 4, 6, 8}, {1, 2, 5, 6},
                                 {3, 4, 7, 8}, {1, 2, 3, 4}, {5, 6, 7, 8}};

/*
 * navier8.f:2004-2360 SETVERT3D (ifcenter=.false.), serial.
 * vertex: (2,2,2,nel) int64 per element, hypercube order.  glo_num: n^3*nel.
 * Returns ngv (number of unique ids excluding interiors).
 */
int64_t nko_setvert3d(int64_t *glo_num, int nx, int nel, const int64_t *vertex) {
  int nxyz = nx * nx * nx;
  int64_t ngvv = 0;
  for (int64_t i = 0; i < (int64_t)8 * nel; i++)
    if (vertex[i] > ngvv) ngvv = vertex[i];
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          int il = 1 + (nx - 1) * i + nx * (nx - 1) * j + nx * nx * (nx - 1) * k;
          glo_num[il - 1 + (size_t)nxyz * e] = vertex[i + 2 * j + 4 * k + 8 * e];
        }
  int64_t ngv = ngvv;
  if (nx == 2) return ngv;

  /* edges: edge(i,j,k,dir,e) labelled by bounding vertices */
  int64_t *edge = malloc(sizeof(int64_t) * 24 * nel); /* (0:1,0:1,0:1,3,nel) */
  for (int e = 0; e < nel; e++)
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) {
          int64_t v = vertex[i + 2 * j + 4 * k + 8 * e];
          edge[i + 2 * j + 4 * k + 8 * 0 + 24 * e] = v; /* r-edge (i,j,k) */
          edge[j + 2 * i + 4 * k + 8 * 1 + 24 * e] = v; /* s-edge (j,i,k) */
          edge[k + 2 * i + 4 * j + 8 * 2 + 24 * e] = v; /* t-edge (k,i,j) */
        }
  int64_t nedge = (int64_t)12 * nel;
  tup3 *et = malloc(sizeof(tup3) * nedge);
  for (int64_t i = 0; i < nedge; i++) {
    int64_t a = edge[2 * i], b = edge[2 * i + 1];
    if (a > b) {
      int64_t s = a;
      a = b;
      b = s;
    }
    et[i].k[0] = a;
    et[i].k[1] = b;
    et[i].k[2] = 0;
    et[i].pos = i;
  }
  int64_t *enumr = malloc(sizeof(int64_t) * nedge);
  tuple_rank(et, nedge, enumr);
  free(et);
  int64_t n_unique_edges = 0;
  for (int64_t i = 0; i < nedge; i++)
    if (enumr[i] > n_unique_edges) n_unique_edges = enumr[i];
  int n_on_edge = nx - 2;
  int64_t ngve = n_unique_edges * n_on_edge;
  for (int e = 0; e < nel; e++) {
    int64_t *g = glo_num + (size_t)nxyz * e; /* g[l-1] == glo_num(l) local */
    int iedg = 0;
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++) { /* edges 1-4 */
        int64_t igv = ngv + (int64_t)n_on_edge * (enumr[iedg + 12 * e] - 1);
        int i0 = nx * (nx - 1) * j + nx * nx * (nx - 1) * k;
        if (g[i0 + 1 - 1] < g[i0 + nx - 1]) {
          for (int i = 2; i <= nx - 1; i++) g[i0 + i - 1] = igv + i - 1;
        } else {
          for (int i = 2; i <= nx - 1; i++) g[i0 + i - 1] = igv + 1 + n_on_edge - (i - 1);
        }
        iedg++;
      }
    for (int k = 0; k < 2; k++)
      for (int i = 0; i < 2; i++) { /* edges 5-8 */
        int64_t igv = ngv + (int64_t)n_on_edge * (enumr[iedg + 12 * e] - 1);
        int i0 = 1 + (nx - 1) * i + nx * nx * (nx - 1) * k;
        if (g[i0 - 1] < g[i0 + nx * (nx - 1) - 1]) {
          for (int j = 2; j <= nx - 1; j++) g[i0 + (j - 1) * nx - 1] = igv + j - 1;
        } else {
          for (int j = 2; j <= nx - 1; j++)
            g[i0 + (j - 1) * nx - 1] = igv + 1 + n_on_edge - (j - 1);
        }
        iedg++;
      }
    for (int j = 0; j < 2; j++)
      for (int i = 0; i < 2; i++) { /* edges 9-12 */
        int64_t igv = ngv + (int64_t)n_on_edge * (enumr[iedg + 12 * e] - 1);
        int i0 = 1 + (nx - 1) * i + nx * (nx - 1) * j;
        if (g[i0 - 1] < g[i0 + nx * nx * (nx - 1) - 1]) {
          for (int k = 2; k <= nx - 1; k++) g[i0 + (k - 1) * nx * nx - 1] = igv + k - 1;
        } else {
          for (int k = 2; k <= nx - 1; k++)
            g[i0 + (k - 1) * nx * nx - 1] = igv + 1 + n_on_edge - (k - 1);
        }
        iedg++;
      }
  }
  free(edge);
  free(enumr);
  ngv = ngv + ngve;

  /* faces: 3 smallest of the 4 sorted corner ids */
  int64_t nface = (int64_t)6 * nel;
  tup3 *ft = malloc(sizeof(tup3) * nface);
  for (int e = 0; e < nel; e++)
    for (int f = 0; f < 6; f++) {
      int64_t facet[4];
      for (int c = 0; c < 4; c++) facet[c] = vertex[icface[f][c] - 1 + 8 * e];
      /* i8sort: ascending */
      for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++)
          if (facet[b] < facet[a]) {
            int64_t s = facet[a];
            facet[a] = facet[b];
            facet[b] = s;
          }
      tup3 *t = &ft[f + 6 * e];
      t->k[0] = facet[0];
      t->k[1] = facet[1];
      t->k[2] = facet[2];
      t->pos = f + 6 * e;
    }
  int64_t *fnum = malloc(sizeof(int64_t) * nface);
  tuple_rank(ft, nface, fnum);
  free(ft);
  int64_t n_unique_faces = 0;
  for (int64_t i = 0; i < nface; i++)
    if (fnum[i] > n_unique_faces) n_unique_faces = fnum[i];
  int skp[6][6];
  dsset_skpdat(nx, skp);
  int nxx = nx * nx;
  int n_on_face = (nx - 2) * (nx - 2);
  for (int e = 0; e < nel; e++) {
    int64_t *g = glo_num + (size_t)nxyz * e;
    for (int f = 0; f < 6; f++) {
      int i0 = skp[f][0], i1 = skp[f][1], is = skp[f][2];
      int j0 = skp[f][3], j1 = skp[f][4], js = skp[f][5];
      int64_t gvf[4];
      int ind[4];
      gvf[0] = g[i0 + nx * (j0 - 1) - 1];
      gvf[1] = g[i1 + nx * (j0 - 1) - 1];
      gvf[2] = g[i0 + nx * (j1 - 1) - 1];
      gvf[3] = g[i1 + nx * (j1 - 1) - 1];
      i8rank(gvf, ind, 4);
      int ifij = 0, idir = 1, jdir = 1;
      if (ind[0] == 1) {
        idir = 1;
        jdir = 1;
        if (gvf[1] < gvf[2]) ifij = 1;
      } else if (ind[0] == 2) {
        idir = -1;
        jdir = 1;
        if (gvf[0] < gvf[3]) ifij = 1;
      } else if (ind[0] == 3) {
        idir = 1;
        jdir = -1;
        if (gvf[3] < gvf[0]) ifij = 1;
      } else if (ind[0] == 4) {
        idir = -1;
        jdir = -1;
        if (gvf[2] < gvf[1]) ifij = 1;
      }
      if (idir < 0) {
        int t = i0;
        i0 = i1;
        i1 = t;
        is = -is;
      }
      if (jdir < 0) {
        int t = j0;
        j0 = j1;
        j1 = t;
        js = -js;
      }
      int64_t ig0 = ngv + (int64_t)n_on_face * (fnum[f + 6 * e] - 1);
      int k = 0;
      int64_t l = 0;
      if (ifij) {
        for (int j = j0; (js > 0) ? (j <= j1) : (j >= j1); j += js)
          for (int i = i0; (is > 0) ? (i <= i1) : (i >= i1); i += is) {
            k++;
            if (k > nx && k < nxx - nx && (k % nx) != 1 && (k % nx) != 0) {
              l++;
              g[i + nx * (j - 1) - 1] = l + ig0;
            }
          }
      } else {
        for (int i = i0; (is > 0) ? (i <= i1) : (i >= i1); i += is)
          for (int j = j0; (js > 0) ? (j <= j1) : (j >= j1); j += js) {
            k++;
            if (k > nx && k < nxx - nx && (k % nx) != 1 && (k % nx) != 0) {
              l++;
              g[i + nx * (j - 1) - 1] = l + ig0;
            }
          }
      }
    }
  }
  free(fnum);
  ngv = ngv + n_unique_faces * n_on_face;
  /* element interiors are not shared: id 0 */
  for (int e = 0; e < nel; e++) {
    int64_t *g = glo_num + (size_t)nxyz * e;
    for (int k = 2; k <= nx - 1; k++)
      for (int j = 2; j <= nx - 1; j++)
        for (int i = 2; i <= nx - 1; i++) g[i - 1 + nx * (j - 1) + nxx * (k - 1)] = 0;
  }
  return ngv;
}

/* ------------------------------------------------------------------ */
/* gather-scatter (gslib gs_op semantics used by dssum.f:33-161)       */
/* ------------------------------------------------------------------ */

typedef struct {
  int64_t n;      /* local size */
  int64_t nseg;   /* groups with >= 2 members */
  int64_t *off;   /* nseg+1 */
  int64_t *idx;   /* members, grouped */
} nko_gs;

typedef struct {
  int64_t g, i;
} gpair;

static int cmp_gpair(const void *pa, const void *pb) {
  const gpair *a = pa, *b = pb;
  if (a->g != b->g) return (a->g < b->g) ? -1 : 1;
  return (a->i < b->i) ? -1 : (a->i > b->i);
}

/* dssum.f:1-31 SETUPDS -> fgslib_gs_setup: group nonzero ids */
nko_gs *nko_gs_setup(const int64_t *glo_num, int64_t n) {
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++)
    if (glo_num[i] != 0) m++;
  gpair *p = malloc(sizeof(gpair) * (m ? m : 1));
  m = 0;
  for (int64_t i = 0; i < n; i++)
    if (glo_num[i] != 0) {
      p[m].g = glo_num[i];
      p[m].i = i;
      m++;
    }
  qsort(p, (size_t)m, sizeof(gpair), cmp_gpair);
  nko_gs *h = calloc(1, sizeof(nko_gs));
  h->n = n;
  h->off = malloc(sizeof(int64_t) * (m + 1));
  h->idx = malloc(sizeof(int64_t) * (m ? m : 1));
  int64_t nseg = 0, w = 0;
  for (int64_t a = 0; a < m;) {
    int64_t b = a;
    while (b < m && p[b].g == p[a].g) b++;
    if (b - a >= 2) {
      h->off[nseg] = w;
      for (int64_t c = a; c < b; c++) h->idx[w++] = p[c].i;
      nseg++;
    }
    a = b;
  }
  h->off[nseg] = w;
  h->nseg = nseg;
  free(p);
  return h;
}

void nko_gs_free(nko_gs *h) {
  if (!h) return;
  free(h->off);
  free(h->idx);
  free(h);
}

/* op: 0 '+', 1 '*', 2 'm' (min), 3 'M' (max) -- dssum.f:33-161 */
void nko_gs_op(const nko_gs *h, double *u, int op) {
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < h->nseg; s++) {
    int64_t a = h->off[s], b = h->off[s + 1];
    double acc = u[h->idx[a]];
    for (int64_t c = a + 1; c < b; c++) {
      double v = u[h->idx[c]];
      if (op == 0) acc += v;
      else if (op == 1) acc *= v;
      else if (op == 2) acc = (v < acc) ? v : acc;
      else acc = (v > acc) ? v : acc;
    }
    for (int64_t c = a; c < b; c++) u[h->idx[c]] = acc;
  }
}

/* ------------------------------------------------------------------ */
/* hmholtz.f : operator, preconditioner, CG                            */
/* ------------------------------------------------------------------ */

typedef struct {
  int n;         /* lx1 */
  int64_t nel;
  const double *dxm1, *dxtm1; /* n x n column-major */
  const double *g1, *g2, *g3, *g4, *g5, *g6, *bm1;
  const double *binv, *mask, *mult;
  double vol;
  const nko_gs *gs;
} nko_field;

/* hmholtz.f:262-307 SETFAST: ifh2 = vlamax(h2) > 0 */
static int setfast_ifh2(const double *h2, int64_t ntot) {
  double m = 0.0;
  for (int64_t i = 0; i < ntot; i++) {
    double a = fabs(h2[i]);
    if (a > m) m = a;
  }
  return m > 0.0;
}

/*
 * hmholtz.f:72-259 AXHELM, 3-D deformed branch (:188-218); ifh2 term (:225).
 * Default param(59)=1 marks every element deformed, so the 6-factor path is
 * the one the reference executes.
 */
void nko_axhelm(const nko_field *f, double *au, const double *u, const double *h1,
                const double *h2) {
  int n = f->n, nxy = n * n, nxyz = n * n * n;
  int64_t ntot = (int64_t)nxyz * f->nel;
  int ifh2 = setfast_ifh2(h2, ntot);
#pragma omp parallel
  {
    double *buf = malloc(sizeof(double) * 9 * nxyz);
    double *dudr = buf, *duds = buf + nxyz, *dudt = buf + 2 * nxyz, *tmp1 = buf + 3 * nxyz,
           *tmp2 = buf + 4 * nxyz, *tmp3 = buf + 5 * nxyz, *tm1 = buf + 6 * nxyz,
           *tm2 = buf + 7 * nxyz, *tm3 = buf + 8 * nxyz;
#pragma omp for schedule(static)
    for (int64_t e = 0; e < f->nel; e++) {
      size_t o = (size_t)e * nxyz;
      const double *ue = u + o;
      mxm(f->dxm1, n, ue, n, dudr, nxy);
      for (int iz = 0; iz < n; iz++) mxm(ue + iz * nxy, n, f->dxtm1, n, duds + iz * nxy, n);
      mxm(ue, nxy, f->dxtm1, n, dudt, n);
      for (int q = 0; q < nxyz; q++) {
        size_t l = o + q;
        double t1 = dudr[q] * f->g1[l];
        double t2 = duds[q] * f->g2[l];
        double t3 = dudt[q] * f->g3[l];
        t1 = t1 + duds[q] * f->g4[l];
        t1 = t1 + dudt[q] * f->g5[l];
        t2 = t2 + dudr[q] * f->g4[l];
        t2 = t2 + dudt[q] * f->g6[l];
        t3 = t3 + dudr[q] * f->g5[l];
        t3 = t3 + duds[q] * f->g6[l];
        tmp1[q] = t1 * h1[l];
        tmp2[q] = t2 * h1[l];
        tmp3[q] = t3 * h1[l];
      }
      mxm(f->dxtm1, n, tmp1, n, tm1, nxy);
      for (int iz = 0; iz < n; iz++) mxm(tmp2 + iz * nxy, n, f->dxm1, n, tm2 + iz * nxy, n);
      mxm(tmp3, nxy, f->dxm1, n, tm3, n);
      for (int q = 0; q < nxyz; q++) {
        double a = 0.0;
        a = a + tm1[q];
        a = a + tm2[q];
        a = a + tm3[q];
        au[o + q] = a;
      }
    }
    free(buf);
  }
  if (ifh2) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ntot; i++) au[i] = au[i] + h2[i] * f->bm1[i] * u[i];
  }
}

/*
 * Adjoint coefficient gradient of the Helmholtz bilinear form (no reference
 * counterpart; built from the reference's own pieces): with a(u,v) =
 * sum_e (D v)^T (h1 G) (D u) + h2 bm1 u v as axhelm applies it, the nodal
 * derivative of lam^T A(h1,h2) u with respect to the local coefficient value
 * at node l is (D lam)_l^T G_l (D u)_l (D = local_grad3, navier5.f:428-445;
 * G = g1..g6 as axhelm pairs them, hmholtz.f:188-218), and with respect to h2
 * it is lam_l bm1_l u_l.  Outputs the gradients of J with A u = f and
 * A lam = dJ/du: g1 = -(D lam)^T G (D u), g2 = -lam bm1 u (g2 may be NULL).
 * Pinned by sum_l g1_l h1_l = -lam . A_L(h1, 0) u (the form is linear in
 * h1) and by central finite differences of J.
 */
void nko_adj_grad(const nko_field *f, double *g1o, double *g2o, const double *u,
                  const double *lam) {
  int n = f->n, nxy = n * n, nxyz = n * n * n;
#pragma omp parallel
  {
    double *buf = malloc(sizeof(double) * 6 * nxyz);
    double *ur = buf, *us = buf + nxyz, *ut = buf + 2 * nxyz, *lr = buf + 3 * nxyz,
           *ls = buf + 4 * nxyz, *lt = buf + 5 * nxyz;
#pragma omp for schedule(static)
    for (int64_t e = 0; e < f->nel; e++) {
      size_t o = (size_t)e * nxyz;
      const double *ue = u + o, *le = lam + o;
      mxm(f->dxm1, n, ue, n, ur, nxy);
      for (int iz = 0; iz < n; iz++) mxm(ue + iz * nxy, n, f->dxtm1, n, us + iz * nxy, n);
      mxm(ue, nxy, f->dxtm1, n, ut, n);
      mxm(f->dxm1, n, le, n, lr, nxy);
      for (int iz = 0; iz < n; iz++) mxm(le + iz * nxy, n, f->dxtm1, n, ls + iz * nxy, n);
      mxm(le, nxy, f->dxtm1, n, lt, n);
      for (int q = 0; q < nxyz; q++) {
        size_t l = o + q;
        double wr = ur[q] * f->g1[l] + us[q] * f->g4[l] + ut[q] * f->g5[l];
        double ws = ur[q] * f->g4[l] + us[q] * f->g2[l] + ut[q] * f->g6[l];
        double wt = ur[q] * f->g5[l] + us[q] * f->g6[l] + ut[q] * f->g3[l];
        g1o[l] = -(lr[q] * wr + ls[q] * ws + lt[q] * wt);
        if (g2o) g2o[l] = -(le[q] * f->bm1[l] * ue[q]);
      }
    }
    free(buf);
  }
}

/* hmholtz.f:380-524 SETPREC (3-D, every element deformed) */
void nko_setprec(const nko_field *f, double *d, const double *h1, const double *h2) {
  int n = f->n, nxyz = n * n * n;
  int64_t ntot = (int64_t)nxyz * f->nel;
  const double *dt = f->dxtm1; /* DXTM1(ix,iq) = dt[ix + n*iq] */
#define DT(a, b) dt[(a) + n * (b)]
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < f->nel; e++) {
    size_t o = (size_t)e * nxyz;
    const double *g1 = f->g1 + o, *g2 = f->g2 + o, *g3 = f->g3 + o, *g4 = f->g4 + o,
                 *g5 = f->g5 + o, *g6 = f->g6 + o;
    double *de = d + o;
    for (int q = 0; q < nxyz; q++) de[q] = 0.0;
    for (int iq = 0; iq < n; iq++)
      for (int iz = 0; iz < n; iz++)
        for (int iy = 0; iy < n; iy++)
          for (int ix = 0; ix < n; ix++)
            de[IDX3(ix, iy, iz, n)] += g1[IDX3(iq, iy, iz, n)] * (DT(ix, iq) * DT(ix, iq));
    for (int iq = 0; iq < n; iq++)
      for (int iz = 0; iz < n; iz++)
        for (int iy = 0; iy < n; iy++)
          for (int ix = 0; ix < n; ix++)
            de[IDX3(ix, iy, iz, n)] += g2[IDX3(ix, iq, iz, n)] * (DT(iy, iq) * DT(iy, iq));
    for (int iq = 0; iq < n; iq++)
      for (int iz = 0; iz < n; iz++)
        for (int iy = 0; iy < n; iy++)
          for (int ix = 0; ix < n; ix++)
            de[IDX3(ix, iy, iz, n)] += g3[IDX3(ix, iy, iq, n)] * (DT(iz, iq) * DT(iz, iq));
    /* cross terms, loops 600/700/800: the strides (ly1-1), (lz1-1) visit
       only the two end values of each index, i.e. the 8 element corners */
    int L = n - 1;
    for (int iy = 0; iy <= L; iy += L)
      for (int iz = 0; iz <= L; iz += L) {
        de[IDX3(0, iy, iz, n)] = de[IDX3(0, iy, iz, n)] +
                                 g4[IDX3(0, iy, iz, n)] * DT(0, 0) * DT(iy, iy) +
                                 g5[IDX3(0, iy, iz, n)] * DT(0, 0) * DT(iz, iz);
        de[IDX3(L, iy, iz, n)] = de[IDX3(L, iy, iz, n)] +
                                 g4[IDX3(L, iy, iz, n)] * DT(L, L) * DT(iy, iy) +
                                 g5[IDX3(L, iy, iz, n)] * DT(L, L) * DT(iz, iz);
      }
    for (int ix = 0; ix <= L; ix += L)
      for (int iz = 0; iz <= L; iz += L) {
        de[IDX3(ix, 0, iz, n)] = de[IDX3(ix, 0, iz, n)] +
                                 g4[IDX3(ix, 0, iz, n)] * DT(0, 0) * DT(ix, ix) +
                                 g6[IDX3(ix, 0, iz, n)] * DT(0, 0) * DT(iz, iz);
        de[IDX3(ix, L, iz, n)] = de[IDX3(ix, L, iz, n)] +
                                 g4[IDX3(ix, L, iz, n)] * DT(L, L) * DT(ix, ix) +
                                 g6[IDX3(ix, L, iz, n)] * DT(L, L) * DT(iz, iz);
      }
    for (int ix = 0; ix <= L; ix += L)
      for (int iy = 0; iy <= L; iy += L) {
        de[IDX3(ix, iy, 0, n)] = de[IDX3(ix, iy, 0, n)] +
                                 g5[IDX3(ix, iy, 0, n)] * DT(0, 0) * DT(ix, ix) +
                                 g6[IDX3(ix, iy, 0, n)] * DT(0, 0) * DT(iy, iy);
        de[IDX3(ix, iy, L, n)] = de[IDX3(ix, iy, L, n)] +
                                 g5[IDX3(ix, iy, L, n)] * DT(L, L) * DT(ix, ix) +
                                 g6[IDX3(ix, iy, L, n)] * DT(L, L) * DT(iy, iy);
      }
  }
#undef DT
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ntot; i++) {
    d[i] = d[i] * h1[i];            /* col2 */
    d[i] = d[i] + h2[i] * f->bm1[i]; /* addcol3 */
  }
  nko_gs_op(f->gs, d, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ntot; i++) d[i] = 1.0 / d[i];
}

static double glsc3(const double *a, const double *b, const double *c, int64_t n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i] * c[i];
  return s;
}

/* hmholtz.f:527-609 CHKTCG1 (EIGAA=0 so ACONDNO=10); returns adjusted tol */
double nko_chktcg1(const nko_field *f, double tol, const double *res, const double *h1,
                   const double *h2) {
  int64_t n = (int64_t)f->n * f->n * f->n * f->nel;
  double eps = 1.0e-13, acondno = 10.0;
  double *w1 = malloc(sizeof(double) * n), *w2 = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {
    w1[i] = res[i];
    w2[i] = f->binv[i] * w1[i];
  }
  double rinit = sqrt(glsc3(w2, w1, f->mult, n) / f->vol);
  double rmin = eps * rinit;
  if (tol < rmin) tol = rmin;
  for (int64_t i = 0; i < n; i++) w1[i] = 1.0;
  double bcneu1 = glsc3(w1, f->mask, f->mult, n);
  double bcneu2 = glsc3(w1, w1, f->mult, n);
  double bctest = fabs(bcneu1 - bcneu2);
  nko_axhelm(f, w2, w1, h1, h2);
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) {
    w2[i] = w2[i] * w2[i];
    w2[i] = w2[i] * f->bm1[i];
    s += w2[i];
  }
  double bcrob = sqrt(s / f->vol);
  if (bctest < 0.1 && bcrob < eps * acondno) {
    double tolmin = rinit * eps * 10.0;
    if (tol < tolmin) tol = tolmin;
  }
  free(w1);
  free(w2);
  return tol;
}

/* per-iteration record of the CG scalars (for iterate-by-iterate parity) */
typedef struct {
  double rbn2, rtz1, rho, alpha, beta;
} nko_cgrec;

typedef struct {
  int istep;       /* time step index (affects first-iteration exit) */
  double param18;  /* 0: B^{-1}-weighted residual norm */
  double param22;  /* helmholtz tolerance knob (<0: relative) */
  double restol;   /* restol(ifield) */
} nko_params;

/*
 * hmholtz.f:611-846 CGGO, Jacobi branch (kfldfdm<0) for a non-PRES field.
 * x: solution (output), f: rhs (already dssum'd and masked; input, left
 * untouched), binv = binvm1 or bintm1.  Returns niter; rec[0..niter] holds the
 * scalars of each iteration (rbn2 also for the exit iteration).
 */
int nko_cggo(const nko_field *fl, double *x, const double *f, const double *h1,
             const double *h2, double tin, int maxit, const nko_params *prm,
             nko_cgrec *rec, double *rbn0_out) {
  int64_t n = (int64_t)fl->n * fl->n * fl->n * fl->nel;
  const int maxcg = 900;
  double *d = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n),
         *w = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
         *z = malloc(sizeof(double) * n);
  double vol = fl->vol;
  double tol = fabs(tin);
  if (prm->restol != 0.0) tol = prm->restol;
  if (tin < 0) tol = fabs(tin);
  int niter = maxit < maxcg ? maxit : maxcg;
  nko_setprec(fl, d, h1, h2);
  for (int64_t i = 0; i < n; i++) {
    r[i] = f[i];
    x[i] = 0.0;
    p[i] = 0.0;
  }
  double fmax = 0.0;
  for (int64_t i = 0; i < n; i++)
    if (fabs(f[i]) > fmax) fmax = fabs(f[i]);
  if (fmax == 0.0) {
    free(d); free(r); free(w); free(p); free(z);
    if (rbn0_out) *rbn0_out = 0.0;
    return 0;
  }
  /* null-space check (:705-720) */
  double h2max = -INFINITY, skmin = INFINITY;
  for (int64_t i = 0; i < n; i++) {
    if (h2[i] > h2max) h2max = h2[i];
    if (fl->mask[i] < skmin) skmin = fl->mask[i];
  }
  int ifmcor = (skmin > 0 && h2max == 0);
  double smean = 0.0;
  if (ifmcor) {
    double sb = 0.0, sr = 0.0;
    for (int64_t i = 0; i < n; i++) {
      sb += fl->bm1[i];
      sr += r[i] * fl->mult[i];
    }
    smean = -1.0 / sb;
    double rmean = smean * sr;
    for (int64_t i = 0; i < n; i++) x[i] = fl->bm1[i];
    nko_gs_op(fl->gs, x, 0);
    for (int64_t i = 0; i < n; i++) {
      r[i] = r[i] + rmean * x[i];
      x[i] = 0.0;
    }
  }
  double rtz1 = 1.0, rtz2, rho = 0.0, rho0, rbn0 = 0.0, rbn2, beta, alpha;
  int iter;
  for (iter = 1; iter <= niter; iter++) {
    for (int64_t i = 0; i < n; i++) z[i] = r[i] * d[i];
    if (ifmcor) {
      double s = 0.0;
      for (int64_t i = 0; i < n; i++) s += z[i] * fl->bm1[i];
      double rmean = smean * s;
      for (int64_t i = 0; i < n; i++) z[i] += rmean;
    }
    rtz2 = rtz1;
    double s1 = 0.0, s2 = 0.0;
#pragma omp parallel for reduction(+ : s1, s2) schedule(static)
    for (int64_t i = 0; i < n; i++) {
      s1 += z[i] * r[i] * fl->mult[i];
      if (prm->param18 == 1.0) s2 += r[i] * r[i] * fl->mult[i];
      else if (g_mutation & 128) s2 += fl->mult[i] * r[i] * r[i];
      else s2 += fl->binv[i] * fl->mult[i] * r[i] * r[i]; /* vlsc32 :848-856 */
    }
    rtz1 = s1;
    rbn2 = sqrt(s2 / vol);
    if (iter == 1) rbn0 = rbn2;
    if (prm->param22 < 0) tol = fabs(prm->param22) * rbn0;
    if (tin < 0) tol = fabs(tin) * rbn0;
    if (rec) {
      rec[iter - 1].rbn2 = rbn2;
      rec[iter - 1].rtz1 = rtz1;
    }
    if (rbn2 <= tol && (iter > 1 || prm->istep <= 5)) {
      niter = iter - 1;
      goto done;
    }
    beta = rtz1 / rtz2;
    if (iter == 1) beta = 0.0;
    for (int64_t i = 0; i < n; i++) p[i] = beta * p[i] + z[i]; /* add2s1 */
    nko_axhelm(fl, w, p, h1, h2);
    nko_gs_op(fl->gs, w, 0);
    for (int64_t i = 0; i < n; i++) w[i] = w[i] * fl->mask[i];
    rho0 = rho;
    (void)rho0;
    rho = glsc3(w, p, fl->mult, n);
    alpha = rtz1 / rho;
    double alphm = -alpha;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] + alphm * w[i];
    }
    if (rec) {
      rec[iter - 1].rho = rho;
      rec[iter - 1].alpha = alpha;
      rec[iter - 1].beta = beta;
    }
  }
  niter = iter - 1;
done:
  if (rbn0_out) *rbn0_out = rbn0;
  free(d); free(r); free(w); free(p); free(z);
  return niter;
}

/*
 * hmholtz.f:2-69 HMHOLTZ for a non-PRES field: dssum(rhs), rhs*=mask,
 * chktcg1 when param(22)==0 or istep<=10, then cggo.  rhs is modified in place.
 */
int nko_hmholtz(const nko_field *fl, double *u, double *rhs, const double *h1,
                const double *h2, double tli, int maxit, const nko_params *prm,
                nko_cgrec *rec, double *rbn0_out) {
  int64_t n = (int64_t)fl->n * fl->n * fl->n * fl->nel;
  double tol = fabs(tli);
  nko_gs_op(fl->gs, rhs, 0);
  for (int64_t i = 0; i < n; i++) rhs[i] = rhs[i] * fl->mask[i];
  if (prm->param22 == 0 || prm->istep <= 10) tol = nko_chktcg1(fl, tol, rhs, h1, h2);
  if (tli < 0) tol = tli;
  return nko_cggo(fl, u, rhs, h1, h2, tol, maxit, prm, rec, rbn0_out);
}

int nko_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void nko_set_num_threads(int t) {
#ifdef _OPENMP
  omp_set_num_threads(t);
#else
  (void)t;
#endif
}
