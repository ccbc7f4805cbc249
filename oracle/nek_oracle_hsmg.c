/*
 * nek_oracle_hsmg.c -- CPU restatement of Nek5000's H1 multigrid
 * preconditioner (hsmg.f h1mg_*, fast3d.f) and the flexible CG that uses it
 * (hmholtz.f:2164-2281 hmh_flex_cg).
 *
 * TEST INFRASTRUCTURE ONLY (see nek_oracle.c).
 *
 * Layout: level-l fields are (nh,nh,nh,nel) column-major; extended arrays
 * (nh+2)^3 per element.  Element faces use Nek's symmetric notation inside
 * (1: r=-1, 2: r=+1, 3: s=-1, 4: s=+1, 5: t=-1, 6: t=+1) and the BC codes of
 * get_fast_bc (fast3d.f:802-877): 0 interior/periodic, 1 Dirichlet
 * (outflow 'O'), 2 Neumann.
 *
 * Parity status: fd_weights_full/semhat pinned by closed forms (exactness on
 * polynomials, A 1 = 0, B = GLL weights); generalev by the defining identities
 * A S = B S Lambda, S^T B S = I and scipy.linalg.eigh; the FDM local solve by
 * (S D S^T) A_ext = I on unmasked rows; the coarse matrix by brute-force
 * Galerkin assembly; swap_lengths/plane_space by their definition evaluated
 * on each element's GLL planes and the box lattice; the whole of h1mg_solve
 * (weights, overlap exchange, masks, level transfers, coarse solve, dsavg)
 * by a dense model built from the definition of each piece on boxes with
 * rotated element frames (tests/dense_mg.py); hmh_flex_cg and hmh_gmres
 * iterate by iterate against dense restatements preconditioned by that
 * model (tests/dense_krylov.py).  Mistakes planted through nko_set_mutation
 * make those comparisons fail.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct nko_gs nko_gs;
nko_gs *nko_gs_setup(const int64_t *glo_num, int64_t n);
void nko_gs_free(nko_gs *h);
void nko_gs_op(const nko_gs *h, double *u, int op);
int64_t nko_setvert3d(int64_t *glo_num, int nx, int nel, const int64_t *vertex);
void nko_zwgll(double *z, double *w, int np);

typedef struct {
  int n;
  int64_t nel;
  const double *dxm1, *dxtm1;
  const double *g1, *g2, *g3, *g4, *g5, *g6, *bm1;
  const double *binv, *mask, *mult;
  double vol;
  const nko_gs *gs;
} nko_field;
void nko_axhelm(const nko_field *f, double *au, const double *u, const double *h1,
                const double *h2);
double nko_chktcg1(const nko_field *f, double tol, const double *res, const double *h1,
                   const double *h2);

#define MAXL 4

/* ------------------------------------------------------------------ */
/* fast3d.f:1294-1349 FD_WEIGHTS_FULL (Fornberg): c(0:n,0:m) col-major   */
/* ------------------------------------------------------------------ */
void nko_fd_weights_full(double xx, const double *x, int n, int m, double *c) {
#define C(j, k) c[(j) + (n + 1) * (k)]
  double c1 = 1.0, c4 = x[0] - xx;
  for (int k = 0; k <= m; k++)
    for (int j = 0; j <= n; j++) C(j, k) = 0.0;
  C(0, 0) = 1.0;
  for (int i = 1; i <= n; i++) {
    int mn = i < m ? i : m;
    double c2 = 1.0, c5 = c4;
    c4 = x[i] - xx;
    for (int j = 0; j <= i - 1; j++) {
      double c3 = x[i] - x[j];
      c2 = c2 * c3;
      for (int k = mn; k >= 1; k--) C(i, k) = c1 * (k * C(i - 1, k - 1) - c5 * C(i - 1, k)) / c2;
      C(i, 0) = -c1 * c5 * C(i - 1, 0) / c2;
      for (int k = mn; k >= 1; k--) C(j, k) = (c4 * C(j, k) - k * C(j, k - 1)) / c3;
      C(j, 0) = c4 * C(j, 0) / c3;
    }
    c1 = c2;
  }
#undef C
}

/* fast3d.f:1215-1292 SEMHAT (a = D^T B D, b, d, z): a, d col-major (n+1)^2 */
void nko_semhat(int n, double *a, double *b, double *d, double *z) {
  int np = n + 1;
  double w[2 * 32];
  nko_zwgll(z, b, np);
  for (int i = 0; i <= n; i++) {
    nko_fd_weights_full(z[i], z, n, 1, w);
    for (int j = 0; j <= n; j++) d[i + np * j] = w[j + np];
  }
  for (int j = 0; j <= n; j++)
    for (int i = 0; i <= n; i++) {
      double s = 0.0;
      for (int k = 0; k <= n; k++) s = s + d[k + np * i] * b[k] * d[k + np * j];
      a[i + np * j] = s;
    }
}

/* ------------------------------------------------------------------ */
/* hmholtz.f:1369-1418 GENERALEV: A x = lam B x (A sym, B SPD).          */
/* The reference calls LAPACK dsygv; this is Cholesky B = L L^T, cyclic  */
/* Jacobi on L^-1 A L^-T, S = L^-T Q, so S^T B S = I.  Eigenpairs are    */
/* returned in ascending order.  a (n x n col-major) is overwritten by S. */
/* ------------------------------------------------------------------ */
void nko_generalev(double *a, const double *b, double *lam, int n) {
  double L[32 * 32], C[32 * 32], Q[32 * 32], T[32 * 32];
#define M(A, i, j) A[(i) + n * (j)]
  /* Cholesky (lower) */
  for (int j = 0; j < n; j++) {
    double s = M(b, j, j);
    for (int k = 0; k < j; k++) s -= M(L, j, k) * M(L, j, k);
    M(L, j, j) = sqrt(s);
    for (int i = j + 1; i < n; i++) {
      double t = M(b, i, j);
      for (int k = 0; k < j; k++) t -= M(L, i, k) * M(L, j, k);
      M(L, i, j) = t / M(L, j, j);
    }
    for (int i = 0; i < j; i++) M(L, i, j) = 0.0;
  }
  /* T = L^-1 A (forward substitution on columns), C = T L^-T = L^-1 T^T (A sym) */
  for (int c = 0; c < n; c++)
    for (int i = 0; i < n; i++) {
      double s = M(a, i, c);
      for (int k = 0; k < i; k++) s -= M(L, i, k) * M(T, k, c);
      M(T, i, c) = s / M(L, i, i);
    }
  for (int c = 0; c < n; c++)
    for (int i = 0; i < n; i++) {
      double s = M(T, c, i); /* (T^T)(i,c) */
      for (int k = 0; k < i; k++) s -= M(L, i, k) * M(C, k, c);
      M(C, i, c) = s / M(L, i, i);
    }
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      double s = 0.5 * (M(C, i, j) + M(C, j, i));
      M(C, i, j) = s;
      M(C, j, i) = s;
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) M(Q, i, j) = (i == j) ? 1.0 : 0.0;
  /* cyclic Jacobi sweeps */
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += M(C, i, j) * M(C, i, j);
        if (i != j) off += M(C, i, j) * M(C, i, j);
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = M(C, p, q);
        if (fabs(apq) < 1e-300) continue;
        double app = M(C, p, p), aqq = M(C, q, q);
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; k++) {
          double ckp = M(C, k, p), ckq = M(C, k, q);
          M(C, k, p) = cs * ckp - sn * ckq;
          M(C, k, q) = sn * ckp + cs * ckq;
        }
        for (int k = 0; k < n; k++) {
          double cpk = M(C, p, k), cqk = M(C, q, k);
          M(C, p, k) = cs * cpk - sn * cqk;
          M(C, q, k) = sn * cpk + cs * cqk;
        }
        for (int k = 0; k < n; k++) {
          double qkp = M(Q, k, p), qkq = M(Q, k, q);
          M(Q, k, p) = cs * qkp - sn * qkq;
          M(Q, k, q) = sn * qkp + cs * qkq;
        }
      }
  }
  /* sort ascending */
  int ord[32];
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      if (M(C, ord[j], ord[j]) < M(C, ord[i], ord[i])) {
        int t = ord[i];
        ord[i] = ord[j];
        ord[j] = t;
      }
  /* S = L^-T Q (back substitution) */
  for (int c = 0; c < n; c++) {
    int oc = ord[c];
    lam[c] = M(C, oc, oc);
    for (int i = n - 1; i >= 0; i--) {
      double s = M(Q, i, oc);
      for (int k = i + 1; k < n; k++) s -= M(L, k, i) * M(a, k, c);
      M(a, i, c) = s / M(L, i, i);
    }
  }
#undef M
}

/* ------------------------------------------------------------------ */
/* Multigrid hierarchy                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
  int lmax, nel;
  int nx[MAXL + 1], nh[MAXL + 1];
  double *ah[MAXL + 1], *bh[MAXL + 1], *zh[MAXL + 1];
  double *jh[MAXL + 1]; /* l -> l+1: (nh[l+1] x nh[l]) col-major */
  nko_gs *gs[MAXL + 1], *gsx[MAXL + 1];
  int64_t *glo[MAXL + 1], *glox[MAXL + 1];
  double *mask[MAXL + 1]; /* 0/1 per level node */
  double *rwt[MAXL + 1];  /* restriction weights (1/multiplicity on borders) */
  double *swt[MAXL + 1];  /* Schwarz weights */
  double *fs[MAXL + 1];   /* per element: S(nl,nl,2,3) */
  double *fd[MAXL + 1];   /* per element: D(nl^3) */
  double *lr, *ll, *lm;   /* per element, 3 directions: [3*e + dir] */
  int *fbc;               /* per element 6 faces (symmetric) */
  /* coarse */
  int ncrs;
  int64_t *crs_slot; /* per level-1 local node: dense index or -1 */
  double *Lc;        /* dense Cholesky factor ncrs x ncrs */
  int null_space;    /* no Dirichlet face: the pressure's constant null space
                        (the reference's ifvcor, navier8.f:201-205) */
  double *vmult;     /* fine multiplicity inverse */
  const nko_gs *gs_fine;
} nko_mg;

static int64_t npow3(int n) { return (int64_t)n * n * n; }

/* Planted mistakes for the tests that prove the pins can fail (never set
   outside tests): bit 1 no dsavg, 2 no Schwarz weights, 4 no ghost-layer
   exchange, 8 no overlap sum of the neighbours' ghost solutions, 16
   plane_space weights w(j) read unshifted, 32 neighbour lengths replaced by
   the element's own, 64 each Schwarz weight applied once (not as many times
   as the weight loops visit the node). */
int g_mutation = 0;
void nko_set_mutation(int m) { g_mutation = m; }

/* hsmg.f:2272-2323 H1MG_SETUP_MG_NX */
static void mg_levels(int lx1, int *lmax, int *nx) {
  static const int mgn2[10] = {1, 2, 2, 2, 2, 3, 3, 5, 5, 5};
  int lx2 = lx1; /* PnPn */
  *lmax = (lx1 == 4) ? 2 : 3;
  nx[1] = 1;
  int mglx2 = 2 * (lx2 / 4) + 1;
  if (lx1 == 5) mglx2 = 3;
  if (lx1 <= 10) mglx2 = mgn2[(lx1 < 10 ? lx1 : 10) - 1];
  if (lx1 == 8) mglx2 = 4;
  if (lx1 == 8) mglx2 = 3;
  if (mglx2 > 3) mglx2 = 3;
  nx[2] = mglx2;
  nx[3] = mglx2 + 1;
  nx[*lmax] = lx1 - 1;
}

/* connect1.f:1009-1024 FACIND in symmetric notation: index ranges of face f */
static void face_range(int f, int n, int *k1, int *k2) {
  /* k1/k2: [3][2] ranges for i,j,k (0-based inclusive) */
  for (int d = 0; d < 3; d++) {
    k1[d] = 0;
    k2[d] = n - 1;
  }
  int dir = (f - 1) / 2, side = (f - 1) % 2;
  k1[dir] = k2[dir] = side ? n - 1 : 0;
}

static void facev0(double *a, int n, int64_t e, int f) {
  int k1[3], k2[3];
  face_range(f, n, k1, k2);
  int64_t o = npow3(n) * e;
  for (int k = k1[2]; k <= k2[2]; k++)
    for (int j = k1[1]; j <= k2[1]; j++)
      for (int i = k1[0]; i <= k2[0]; i++) a[o + i + n * (j + n * k)] = 0.0;
}

/* fast3d.f:306-423 PLANE_SPACE (3-D) as called by swap_lengths with
   i1=0, i2=lx1-1: harmonic-type mean of the squared distance between the
   first and last GLL planes over the face-interior nodes 1..lx1-2, weighted
   by w(j)*w(k) where w is the 1-based wxm1 (so node j gets wxm1(j), i.e. the
   weight of node j-1 in 0-based numbering, exactly as the reference reads). */
static void plane_space(int n, int64_t nel, const double *w, const double *x,
                        const double *y, const double *z, double *lm) {
  int n2 = n - 1, nx = n - 2;
#define X(a, i, j, k, e) a[(i) + n * ((j) + n * ((k) + (int64_t)n * (e)))]
  for (int64_t e = 0; e < nel; e++) {
    double s = 0, ws = 0;
    for (int k = 1; k <= nx; k++)
      for (int j = 1; j <= nx; j++) {
        double wt = (g_mutation & 16) ? w[j] * w[k] : w[j - 1] * w[k - 1];
        double dx = X(x, n2, j, k, e) - X(x, 0, j, k, e), dy = X(y, n2, j, k, e) - X(y, 0, j, k, e),
               dz = X(z, n2, j, k, e) - X(z, 0, j, k, e);
        s = s + wt / (dx * dx + dy * dy + dz * dz);
        ws = ws + wt;
      }
    s = s / ws;
    lm[3 * e + 0] = 1.0 / sqrt(s);
    s = 0;
    ws = 0;
    for (int k = 1; k <= nx; k++)
      for (int i = 1; i <= nx; i++) {
        double wt = (g_mutation & 16) ? w[i] * w[k] : w[i - 1] * w[k - 1];
        double dx = X(x, i, n2, k, e) - X(x, i, 0, k, e), dy = X(y, i, n2, k, e) - X(y, i, 0, k, e),
               dz = X(z, i, n2, k, e) - X(z, i, 0, k, e);
        s = s + wt / (dx * dx + dy * dy + dz * dz);
        ws = ws + wt;
      }
    s = s / ws;
    lm[3 * e + 1] = 1.0 / sqrt(s);
    s = 0;
    ws = 0;
    for (int j = 1; j <= nx; j++)
      for (int i = 1; i <= nx; i++) {
        double wt = (g_mutation & 16) ? w[i] * w[j] : w[i - 1] * w[j - 1];
        double dx = X(x, i, j, n2, e) - X(x, i, j, 0, e), dy = X(y, i, j, n2, e) - X(y, i, j, 0, e),
               dz = X(z, i, j, n2, e) - X(z, i, j, 0, e);
        s = s + wt / (dx * dx + dy * dy + dz * dz);
        ws = ws + wt;
      }
    s = s / ws;
    lm[3 * e + 2] = 1.0 / sqrt(s);
  }
#undef X
}

/* fast3d.f:1542-1617 SWAP_LENGTHS (3-D): neighbour lengths via dssum of the
   element lengths written on face-interior nodes */
static void swap_lengths(int n, int64_t nel, const double *w, const double *x, const double *y,
                         const double *z, const nko_gs *gs, double *ll, double *lm,
                         double *lr) {
  plane_space(n, nel, w, x, y, z, lm);
  int64_t nt = npow3(n) * nel;
  double *L = calloc((size_t)nt, sizeof(double));
#define LL(i, j, k, e) L[(i) + n * ((j) + n * ((k) + (int64_t)n * (e)))]
  for (int64_t e = 0; e < nel; e++)
    for (int j = 1; j <= n - 2; j++)
      for (int k = 1; k <= n - 2; k++) {
        LL(0, k, j, e) = lm[3 * e];
        LL(n - 1, k, j, e) = lm[3 * e];
        LL(k, 0, j, e) = lm[3 * e + 1];
        LL(k, n - 1, j, e) = lm[3 * e + 1];
        LL(k, j, 0, e) = lm[3 * e + 2];
        LL(k, j, n - 1, e) = lm[3 * e + 2];
      }
  nko_gs_op(gs, L, 0);
  for (int64_t e = 0; e < nel; e++) {
    ll[3 * e + 0] = LL(0, 1, 1, e) - lm[3 * e];
    lr[3 * e + 0] = LL(n - 1, 1, 1, e) - lm[3 * e];
    ll[3 * e + 1] = LL(1, 0, 1, e) - lm[3 * e + 1];
    lr[3 * e + 1] = LL(1, n - 1, 1, e) - lm[3 * e + 1];
    ll[3 * e + 2] = LL(1, 1, 0, e) - lm[3 * e + 2];
    lr[3 * e + 2] = LL(1, 1, n - 1, e) - lm[3 * e + 2];
    if (g_mutation & 32)
      for (int d = 0; d < 3; d++) ll[3 * e + d] = lr[3 * e + d] = lm[3 * e + d];
  }
#undef LL
  free(L);
}

/* hsmg.f:749-820 HSMG_SETUP_FAST1D_A/_B and :733-747 HSMG_SETUP_FAST1D */
static void setup_fast1d(double *s /* nl*nl*2 */, double *lam, int nl, int lbc, int rbc,
                         double ll, double lm, double lr, const double *ah, const double *bh,
                         int n) {
  double a[32 * 32], b[32 * 32];
#define A(i, j) a[(i) + nl * (j)]
#define B(i, j) b[(i) + nl * (j)]
#define AH(i, j) ah[(i) + (n + 1) * (j)]
  int i0 = (lbc == 1) ? 1 : 0, i1 = (rbc == 1) ? n - 1 : n;
  for (int q = 0; q < nl * nl; q++) a[q] = b[q] = 0.0;
  double fac = 2.0 / lm;
  A(1, 1) = 1.0;
  A(n + 1, n + 1) = 1.0;
  for (int j = i0; j <= i1; j++)
    for (int i = i0; i <= i1; i++) A(i + 1, j + 1) = fac * AH(i, j);
  if (lbc == 0) {
    fac = 2.0 / ll;
    A(0, 0) = fac * AH(n - 1, n - 1);
    A(1, 0) = fac * AH(n, n - 1);
    A(0, 1) = fac * AH(n - 1, n);
    A(1, 1) = A(1, 1) + fac * AH(n, n);
  } else {
    A(0, 0) = 1.0;
  }
  if (rbc == 0) {
    fac = 2.0 / lr;
    A(n + 1, n + 1) = A(n + 1, n + 1) + fac * AH(0, 0);
    A(n + 2, n + 1) = fac * AH(1, 0);
    A(n + 1, n + 2) = fac * AH(0, 1);
    A(n + 2, n + 2) = fac * AH(1, 1);
  } else {
    A(n + 2, n + 2) = 1.0;
  }
  fac = 0.5 * lm;
  B(1, 1) = 1.0;
  B(n + 1, n + 1) = 1.0;
  for (int i = i0; i <= i1; i++) B(i + 1, i + 1) = fac * bh[i];
  if (lbc == 0) {
    fac = 0.5 * ll;
    B(0, 0) = fac * bh[n - 1];
    B(1, 1) = B(1, 1) + fac * bh[n];
  } else {
    B(0, 0) = 1.0;
  }
  if (rbc == 0) {
    fac = 0.5 * lr;
    B(n + 1, n + 1) = B(n + 1, n + 1) + fac * bh[0];
    B(n + 2, n + 2) = fac * bh[1];
  } else {
    B(n + 2, n + 2) = 1.0;
  }
#undef A
#undef B
#undef AH
  nko_generalev(a, b, lam, nl);
  /* row_zero (fast3d.f:1619) of ghost / Dirichlet rows */
  if (lbc > 0)
    for (int j = 0; j < nl; j++) a[0 + nl * j] = 0.0;
  if (lbc == 1)
    for (int j = 0; j < nl; j++) a[1 + nl * j] = 0.0;
  if (rbc > 0)
    for (int j = 0; j < nl; j++) a[(nl - 1) + nl * j] = 0.0;
  if (rbc == 1)
    for (int j = 0; j < nl; j++) a[(nl - 2) + nl * j] = 0.0;
  for (int q = 0; q < nl * nl; q++) s[q] = a[q];
  for (int i = 0; i < nl; i++)
    for (int j = 0; j < nl; j++) s[nl * nl + j + nl * i] = a[i + nl * j];
}

/* hsmg.f:368-423 HSMG_EXTRUDE (3-D): on the nx^3 extended arrays */
static void extrude(double *a1, int l1, double f1, const double *a2, int l2, double f2, int nx,
                    int64_t nel) {
  int i0 = 1, i1 = nx - 2; /* 0-based 2..nx-1 */
#define AX(a, i, j, k, e) a[(i) + nx * ((j) + nx * ((k) + (int64_t)nx * (e)))]
  for (int64_t e = 0; e < nel; e++) {
    for (int k = i0; k <= i1; k++)
      for (int j = i0; j <= i1; j++) {
        AX(a1, l1, j, k, e) = f1 * AX(a1, l1, j, k, e) + f2 * AX(a2, l2, j, k, e);
        AX(a1, nx - 1 - l1, j, k, e) =
            f1 * AX(a1, nx - 1 - l1, j, k, e) + f2 * AX(a2, nx - 1 - l2, j, k, e);
      }
    for (int k = i0; k <= i1; k++)
      for (int i = i0; i <= i1; i++) {
        AX(a1, i, l1, k, e) = f1 * AX(a1, i, l1, k, e) + f2 * AX(a2, i, l2, k, e);
        AX(a1, i, nx - 1 - l1, k, e) =
            f1 * AX(a1, i, nx - 1 - l1, k, e) + f2 * AX(a2, i, nx - 1 - l2, k, e);
      }
    for (int j = i0; j <= i1; j++)
      for (int i = i0; i <= i1; i++) {
        AX(a1, i, j, l1, e) = f1 * AX(a1, i, j, l1, e) + f2 * AX(a2, i, j, l2, e);
        AX(a1, i, j, nx - 1 - l1, e) =
            f1 * AX(a1, i, j, nx - 1 - l1, e) + f2 * AX(a2, i, j, nx - 1 - l2, e);
      }
  }
#undef AX
}

static void toext(double *a, const double *b, int n, int64_t nel) {
  int m = n + 2;
  memset(a, 0, sizeof(double) * (size_t)(npow3(m) * nel));
  for (int64_t e = 0; e < nel; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
          a[(i + 1) + m * ((j + 1) + m * ((k + 1) + (int64_t)m * e))] =
              b[i + n * (j + n * (k + (int64_t)n * e))];
}

static void toreg(double *b, const double *a, int n, int64_t nel) {
  int m = n + 2;
  for (int64_t e = 0; e < nel; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
          b[i + n * (j + n * (k + (int64_t)n * e))] =
              a[(i + 1) + m * ((j + 1) + m * ((k + 1) + (int64_t)m * e))];
}

/* mxm-based tensor apply: v(nv^3) = [C (x) B (x) A] u(nu^3), A,B,C (nv x nu) col-major */
static void tnsr3(double *v, int nv, const double *u, int nu, const double *A, const double *B,
                  const double *Cm) {
  double w1[32 * 32 * 32], w2[32 * 32 * 32];
  for (int k = 0; k < nu; k++)
    for (int j = 0; j < nu; j++)
      for (int i = 0; i < nv; i++) {
        double s = 0;
        for (int m = 0; m < nu; m++) s += A[i + nv * m] * u[m + nu * (j + nu * k)];
        w1[i + nv * (j + nu * k)] = s;
      }
  for (int k = 0; k < nu; k++)
    for (int j = 0; j < nv; j++)
      for (int i = 0; i < nv; i++) {
        double s = 0;
        for (int m = 0; m < nu; m++) s += w1[i + nv * (m + nu * k)] * B[j + nv * m];
        w2[i + nv * (j + nv * k)] = s;
      }
  for (int k = 0; k < nv; k++)
    for (int j = 0; j < nv; j++)
      for (int i = 0; i < nv; i++) {
        double s = 0;
        for (int m = 0; m < nu; m++) s += w2[i + nv * (j + nv * m)] * Cm[k + nv * m];
        v[i + nv * (j + nv * k)] = s;
      }
}

/* hsmg.f:897-929 HSMG_DO_FAST: e = (S_t (x) S_s (x) S_r) D (S_t^T (x) ...) r */
static void do_fast(double *e, double *r, const double *s, const double *d, int nl,
                    int64_t nel) {
  int64_t nn = npow3(nl);
  int n2 = nl * nl;
  double St[3][32 * 32];
  for (int64_t ie = 0; ie < nel; ie++) {
    const double *se = s + (int64_t)ie * n2 * 2 * 3;
    /* first apply: A = S_r^T, B = S_s^T, C = S_t^T (the stored transposes) */
    for (int dir = 0; dir < 3; dir++) memcpy(St[dir], se + n2 * (2 * dir + 1), sizeof(double) * n2);
    tnsr3(e + nn * ie, nl, r + nn * ie, nl, St[0], St[1], St[2]);
    for (int64_t i = 0; i < nn; i++) r[nn * ie + i] = d[nn * ie + i] * e[nn * ie + i];
    tnsr3(e + nn * ie, nl, r + nn * ie, nl, se + n2 * 0, se + n2 * 2, se + n2 * 4);
  }
}

static void gen_fbc(nko_mg *mg, const int *fbc) {
  mg->fbc = malloc(sizeof(int) * 6 * mg->nel);
  memcpy(mg->fbc, fbc, sizeof(int) * 6 * mg->nel);
}

/* level-l mask (hsmg.f:2426-2508 H1MG_SETUP_MASK): faces with BC 1 zeroed, dsprod */
static void level_mask(nko_mg *mg, int l) {
  int n = mg->nh[l];
  int64_t nt = npow3(n) * mg->nel;
  double *m = malloc(sizeof(double) * nt);
  for (int64_t i = 0; i < nt; i++) m[i] = 1.0;
  for (int64_t e = 0; e < mg->nel; e++)
    for (int f = 1; f <= 6; f++)
      if (mg->fbc[6 * e + f - 1] == 1) facev0(m, n, e, f);
  nko_gs_op(mg->gs[l], m, 1);
  mg->mask[l] = m;
}

/* hsmg.f:987-1067 HSMG_SETUP_RSTR_WT as a node array */
static void level_rstr_wt(nko_mg *mg, int l) {
  int n = mg->nh[l];
  int64_t nt = npow3(n) * mg->nel;
  double *w = calloc((size_t)nt, sizeof(double));
  double *wt = malloc(sizeof(double) * nt);
  for (int64_t e = 0; e < mg->nel; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
          if (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 || k == n - 1)
            w[i + n * (j + n * (k + (int64_t)n * e))] = 1.0;
  nko_gs_op(mg->gs[l], w, 0);
  for (int64_t q = 0; q < nt; q++) wt[q] = 1.0;
  for (int64_t e = 0; e < mg->nel; e++)
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          int64_t q = i + n * (j + n * (k + (int64_t)n * e));
          if (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 || k == n - 1)
            wt[q] = 1.0 / w[q];
        }
  free(w);
  mg->rwt[l] = wt;
}

/* the overlap sum of h1mg_schwarz_part1 (hsmg.f:475-491): from the local
   solutions W2 (extended) to level-l values e */
static void overlap_sum(nko_mg *mg, int l, double *W1, double *W2, double *e) {
  int n = mg->nh[l], m = n + 2;
  int64_t nel = mg->nel;
  extrude(W1, 0, 0.0, W2, 0, 1.0, m, nel);
  nko_gs_op(mg->gsx[l], W2, 0);
  extrude(W2, 0, 1.0, W1, 0, -1.0, m, nel);
  if (!(g_mutation & 8)) extrude(W2, 2, 1.0, W2, 0, 1.0, m, nel);
  toreg(e, W2, n, nel);
  nko_gs_op(mg->gs[l], e, 0);
}

/* hsmg.f:3048-3104 H1MG_SETUP_SCHWARZ_WT_1 (+ _3D_2) folded with the loops of
   hsmg.f:1285-1319 HSMG_SCHWARZ_WT3D into one node array: every statement of
   those loops multiplies its node by 1/count (a node hit twice, as i=2 and
   i=n-1 at n=3, is multiplied twice, as in the reference). */
static void level_schwarz_wt(nko_mg *mg, int l) {
  int n = mg->nh[l], m = n + 2;
  int64_t nel = mg->nel, ns = npow3(m) * nel, nt = npow3(n) * nel;
  double *W1 = calloc((size_t)ns, sizeof(double)), *W2 = malloc(sizeof(double) * ns);
  double *cnt = malloc(sizeof(double) * nt), *wt = malloc(sizeof(double) * nt);
  for (int64_t q = 0; q < ns; q++) W2[q] = 1.0;
  overlap_sum(mg, l, W1, W2, cnt);
  for (int64_t q = 0; q < nt; q++) wt[q] = 1.0;
  const int lay[4] = {0, 1, n - 2, n - 1};
#define Q(i, j, k) ((i) + n * ((j) + n * ((k) + (int64_t)n * e)))
  if (g_mutation & 64) {
    for (int64_t q = 0; q < nt; q++) wt[q] = 1.0 / cnt[q];
    free(W1);
    free(W2);
    free(cnt);
    mg->swt[l] = wt;
    return;
  }
  for (int64_t e = 0; e < nel; e++) {
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int t = 0; t < 4; t++) wt[Q(lay[t], j, k)] *= 1.0 / cnt[Q(lay[t], j, k)];
    for (int k = 0; k < n; k++)
      for (int i = 2; i <= n - 3; i++)
        for (int t = 0; t < 4; t++) wt[Q(i, lay[t], k)] *= 1.0 / cnt[Q(i, lay[t], k)];
    for (int j = 2; j <= n - 3; j++)
      for (int i = 2; i <= n - 3; i++)
        for (int t = 0; t < 4; t++) wt[Q(i, j, lay[t])] *= 1.0 / cnt[Q(i, j, lay[t])];
  }
#undef Q
  free(W1);
  free(W2);
  free(cnt);
  mg->swt[l] = wt;
}

/* coarse space (navier8.f:83-233 SET_UP_H1_CRS with get_local_crs_galerkin
   :1648-1687): vertex ids with Dirichlet vertices dropped, Galerkin element
   matrices b_i^T A_e b_j from the fine axhelm (h1=1, h2=0), dense Cholesky of
   the assembled matrix (the reference's XXT is an exact solve). */
static int coarse_setup(nko_mg *mg, const nko_field *fine, const int64_t *vertex) {
  int64_t nel = mg->nel;
  int n = fine->n;
  int64_t nxyz = npow3(n), nt = nxyz * nel;
  /* mask on coarse vertices: ones, zero Dirichlet faces, dsprod over vertex ids */
  int64_t *ids = malloc(sizeof(int64_t) * 8 * nel);
  memcpy(ids, vertex, sizeof(int64_t) * 8 * nel);
  double *cm = malloc(sizeof(double) * 8 * nel);
  for (int64_t q = 0; q < 8 * nel; q++) cm[q] = 1.0;
  for (int64_t e = 0; e < nel; e++)
    for (int f = 1; f <= 6; f++)
      if (mg->fbc[6 * e + f - 1] == 1) facev0(cm, 2, e, f);
  nko_gs *g = nko_gs_setup(ids, 8 * nel);
  nko_gs_op(g, cm, 1);
  nko_gs_free(g);
  for (int64_t q = 0; q < 8 * nel; q++)
    if (cm[q] < 0.1) ids[q] = 0;
  free(cm);
  /* dense index of the free vertex ids */
  int64_t vmax = 0;
  for (int64_t q = 0; q < 8 * nel; q++)
    if (ids[q] > vmax) vmax = ids[q];
  int64_t *slot_of = malloc(sizeof(int64_t) * (vmax + 1));
  for (int64_t v = 0; v <= vmax; v++) slot_of[v] = -1;
  int ncrs = 0;
  for (int64_t q = 0; q < 8 * nel; q++)
    if (ids[q] != 0 && slot_of[ids[q]] < 0) slot_of[ids[q]] = ncrs++;
  mg->crs_slot = malloc(sizeof(int64_t) * 8 * nel);
  for (int64_t q = 0; q < 8 * nel; q++) mg->crs_slot[q] = ids[q] ? slot_of[ids[q]] : -1;
  free(slot_of);
  free(ids);
  mg->ncrs = ncrs;
  if ((double)ncrs * ncrs > 4.0e8) {
    fprintf(stderr, "oracle coarse problem too large for a dense factor (%d)\n", ncrs);
    return 1;
  }
  /* Galerkin element matrices */
  double z[32], w[32];
  nko_zwgll(z, w, n);
  double *bas = malloc(sizeof(double) * 8 * nxyz);
  for (int j = 1; j <= 8; j++) { /* gen_crs_basis */
    double zr[32], zs[32], zt[32];
    for (int i = 0; i < n; i++) zr[i] = zs[i] = zt[i] = 0.5 * (1 - z[i]);
    if (j % 2 == 0)
      for (int i = 0; i < n; i++) zr[i] = 0.5 * (1 + z[i]);
    if (j == 3 || j == 4 || j == 7 || j == 8)
      for (int i = 0; i < n; i++) zs[i] = 0.5 * (1 + z[i]);
    if (j > 4)
      for (int i = 0; i < n; i++) zt[i] = 0.5 * (1 + z[i]);
    for (int r = 0; r < n; r++)
      for (int q = 0; q < n; q++)
        for (int p = 0; p < n; p++) bas[(j - 1) * nxyz + p + n * (q + n * r)] = zr[p] * zs[q] * zt[r];
  }
  double *w1 = malloc(sizeof(double) * nt), *w2 = malloc(sizeof(double) * nt);
  double *h1 = malloc(sizeof(double) * nt), *h2 = calloc((size_t)nt, sizeof(double));
  for (int64_t q = 0; q < nt; q++) h1[q] = 1.0;
  double *Ae = malloc(sizeof(double) * 64 * nel);
  for (int j = 0; j < 8; j++) {
    for (int64_t e = 0; e < nel; e++) memcpy(w1 + nxyz * e, bas + j * nxyz, sizeof(double) * nxyz);
    nko_axhelm(fine, w2, w1, h1, h2);
    for (int64_t e = 0; e < nel; e++)
      for (int i = 0; i < 8; i++) {
        double s = 0.0;
        for (int64_t q = 0; q < nxyz; q++) s = s + bas[i * nxyz + q] * w2[nxyz * e + q];
        Ae[64 * e + i + 8 * j] = s;
      }
  }
  free(w1);
  free(w2);
  free(h1);
  free(h2);
  free(bas);
  double *A = calloc((size_t)ncrs * ncrs, sizeof(double));
  for (int64_t e = 0; e < nel; e++)
    for (int j = 0; j < 8; j++)
      for (int i = 0; i < 8; i++) {
        int64_t si = mg->crs_slot[8 * e + i], sj = mg->crs_slot[8 * e + j];
        if (si >= 0 && sj >= 0) A[si + (int64_t)ncrs * sj] += Ae[64 * e + i + 8 * j];
      }
  free(Ae);
  /* with the null space (crs_xxt.c:883-900, :932-955): one unknown, here the
     last, is held at zero -- its row and column dropped, unit diagonal -- and
     the solution's mean is removed after the solve */
  if (mg->null_space && ncrs > 0) {
    int pn = ncrs - 1;
    for (int i = 0; i < ncrs; i++) {
      A[pn + (int64_t)ncrs * i] = 0.0;
      A[i + (int64_t)ncrs * pn] = 0.0;
    }
    A[pn + (int64_t)ncrs * pn] = 1.0;
  }
  /* Cholesky, lower, in place (column-major) */
  for (int j = 0; j < ncrs; j++) {
    double s = A[j + (int64_t)ncrs * j];
    for (int k = 0; k < j; k++) s -= A[j + (int64_t)ncrs * k] * A[j + (int64_t)ncrs * k];
    if (s <= 0) {
      fprintf(stderr, "oracle coarse matrix not SPD at %d\n", j);
      return 1;
    }
    double d = sqrt(s);
    A[j + (int64_t)ncrs * j] = d;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < ncrs; i++) {
      double t = A[i + (int64_t)ncrs * j];
      for (int k = 0; k < j; k++) t -= A[i + (int64_t)ncrs * k] * A[j + (int64_t)ncrs * k];
      A[i + (int64_t)ncrs * j] = t / d;
    }
  }
  mg->Lc = A;
  return 0;
}

/* fcrs.c:77-85 crs_solve semantics: assemble local values by vertex id,
   solve, scatter; Dirichlet vertices get 0 */
static void coarse_solve(const nko_mg *mg, double *x, const double *b) {
  int nc = mg->ncrs;
  double *v = calloc((size_t)nc + 1, sizeof(double));
  for (int64_t q = 0; q < 8 * (int64_t)mg->nel; q++)
    if (mg->crs_slot[q] >= 0) v[mg->crs_slot[q]] += b[q];
  if (mg->null_space && nc > 0) v[nc - 1] = 0.0;
  const double *L = mg->Lc;
  for (int i = 0; i < nc; i++) {
    double s = v[i];
    for (int k = 0; k < i; k++) s -= L[i + (int64_t)nc * k] * v[k];
    v[i] = s / L[i + (int64_t)nc * i];
  }
  for (int i = nc - 1; i >= 0; i--) {
    double s = v[i];
    for (int k = i + 1; k < nc; k++) s -= L[k + (int64_t)nc * i] * v[k];
    v[i] = s / L[i + (int64_t)nc * i];
  }
  if (mg->null_space && nc > 0) { /* crs_xxt.c:950-955, one rank: plain mean */
    double s = 0.0;
    for (int i = 0; i < nc; i++) s += v[i];
    s = s / nc;
    for (int i = 0; i < nc; i++) v[i] -= s;
  }
  for (int64_t q = 0; q < 8 * (int64_t)mg->nel; q++)
    x[q] = mg->crs_slot[q] >= 0 ? v[mg->crs_slot[q]] : 0.0;
  free(v);
}

/*
 * hsmg.f:2234-2270 H1MG_SETUP for a fine field.  fbc: per element 6 BC codes
 * in symmetric face order.  xyz: fine GLL coordinates; w: fine 1-D weights.
 */
nko_mg *nko_mg_setup(const nko_field *fine, const int64_t *vertex, const int *fbc,
                     const double *x, const double *y, const double *z, const double *w) {
  nko_mg *mg = calloc(1, sizeof(nko_mg));
  int lx1 = fine->n;
  int64_t nel = fine->nel;
  mg->nel = (int)nel;
  mg->gs_fine = fine->gs;
  mg->vmult = (double *)fine->mult;
  mg_levels(lx1, &mg->lmax, mg->nx);
  gen_fbc(mg, fbc);
  mg->null_space = 1;
  for (int64_t q = 0; q < 6 * nel; q++)
    if (fbc[q] == 1) mg->null_space = 0;
  for (int l = 1; l <= mg->lmax; l++) {
    int n = mg->nx[l];
    mg->nh[l] = n + 1;
    mg->ah[l] = malloc(sizeof(double) * (n + 1) * (n + 1));
    mg->bh[l] = malloc(sizeof(double) * (n + 1));
    mg->zh[l] = malloc(sizeof(double) * (n + 1));
    double *dh = malloc(sizeof(double) * (n + 1) * (n + 1));
    nko_semhat(n, mg->ah[l], mg->bh[l], dh, mg->zh[l]);
    free(dh);
  }
  for (int l = 1; l < mg->lmax; l++) { /* hsmg.f:90-113 HSMG_SETUP_INTP */
    int nf = mg->nh[l + 1], nc = mg->nh[l];
    mg->jh[l] = malloc(sizeof(double) * nf * nc);
    double wk[2 * 64];
    for (int i = 0; i < nf; i++) {
      nko_fd_weights_full(mg->zh[l + 1][i], mg->zh[l], nc - 1, 1, wk);
      for (int j = 0; j < nc; j++) mg->jh[l][i + nf * j] = wk[j];
    }
  }
  for (int l = 1; l <= mg->lmax; l++) { /* hsmg.f:2361-2394 H1MG_SETUP_DSSUM */
    int n = mg->nh[l];
    mg->glo[l] = malloc(sizeof(int64_t) * npow3(n) * nel);
    nko_setvert3d(mg->glo[l], n, (int)nel, vertex);
    mg->gs[l] = nko_gs_setup(mg->glo[l], npow3(n) * nel);
    mg->glox[l] = malloc(sizeof(int64_t) * npow3(n + 2) * nel);
    nko_setvert3d(mg->glox[l], n + 2, (int)nel, vertex);
    mg->gsx[l] = nko_gs_setup(mg->glox[l], npow3(n + 2) * nel);
  }
  for (int l = 1; l <= mg->lmax; l++) {
    level_mask(mg, l);
    level_rstr_wt(mg, l);
  }
  /* lengths (fast3d.f:1542 swap_lengths, called before h1mg_setup) */
  mg->ll = malloc(sizeof(double) * 3 * nel);
  mg->lm = malloc(sizeof(double) * 3 * nel);
  mg->lr = malloc(sizeof(double) * 3 * nel);
  swap_lengths(lx1, nel, w, x, y, z, fine->gs, mg->ll, mg->lm, mg->lr);
  for (int l = 2; l <= mg->lmax; l++) { /* hsmg.f:633-665 H1MG_SETUP_FDM */
    int n = mg->nx[l], nl = mg->nh[l] + 2;
    mg->fs[l] = malloc(sizeof(double) * (int64_t)nl * nl * 2 * 3 * nel);
    mg->fd[l] = malloc(sizeof(double) * npow3(nl) * nel);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < nel; e++) {
      double lam[3][32];
      const int *bc = mg->fbc + 6 * e;
      for (int dir = 0; dir < 3; dir++)
        setup_fast1d(mg->fs[l] + (int64_t)nl * nl * 2 * (3 * e + dir), lam[dir], nl,
                     bc[2 * dir], bc[2 * dir + 1], mg->ll[3 * e + dir], mg->lm[3 * e + dir],
                     mg->lr[3 * e + dir], mg->ah[l], mg->bh[l], n);
      double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int dir = 0; dir < 3; dir++)
        for (int i = 1; i <= nl - 2; i++)
          if (lam[dir][i] > mx[dir]) mx[dir] = lam[dir][i];
      double eps = 1.e-5 * (mx[0] + mx[1] + mx[2]);
      double *d = mg->fd[l] + npow3(nl) * e;
      int il = 0;
      for (int k = 0; k < nl; k++)
        for (int j = 0; j < nl; j++)
          for (int i = 0; i < nl; i++) {
            double diag = lam[0][i] + lam[1][j] + lam[2][k];
            d[il++] = (diag > eps) ? 1.0 / diag : 0.0;
          }
    }
  }
  for (int l = 2; l <= mg->lmax; l++) level_schwarz_wt(mg, l);
  if (coarse_setup(mg, fine, vertex)) return NULL;
  return mg;
}

int nko_mg_levels(const nko_mg *mg, int *nh) {
  for (int l = 1; l <= mg->lmax; l++) nh[l - 1] = mg->nh[l];
  return mg->lmax;
}

/* hsmg.f:440-494 H1MG_SCHWARZ_PART1 then :425-438 weights and sigma */
static void schwarz(nko_mg *mg, int l, double *e, double *r, double sigma) {
  int n = mg->nh[l], m = n + 2, nl = m;
  int64_t nel = mg->nel, nt = npow3(n) * nel, ns = npow3(m) * nel;
  for (int64_t q = 0; q < nt; q++) r[q] = r[q] * mg->mask[l][q]; /* h1mg_mask */
  double *W1 = malloc(sizeof(double) * ns), *W2 = malloc(sizeof(double) * ns);
  toext(W1, r, n, nel);
  if (!(g_mutation & 4)) {
    extrude(W1, 0, 0.0, W1, 2, 1.0, m, nel);
    nko_gs_op(mg->gsx[l], W1, 0);
    extrude(W1, 0, 1.0, W1, 2, -1.0, m, nel);
  }
  do_fast(W2, W1, mg->fs[l], mg->fd[l], nl, nel);
  overlap_sum(mg, l, W1, W2, e);
  for (int64_t q = 0; q < nt; q++) e[q] = e[q] * mg->mask[l][q];
  free(W1);
  free(W2);
  if (!(g_mutation & 2))
    for (int64_t q = 0; q < nt; q++) e[q] = e[q] * mg->swt[l][q];
  for (int64_t q = 0; q < nt; q++) e[q] = sigma * e[q];
}

/* in-place tensor restriction/prolongation per element */
static void tnsr_levels(double *v, int nv, const double *u, int nu, const double *A, int64_t nel) {
  /* A: (nv x nu) col-major, same matrix in all three directions */
  double *tmp = malloc(sizeof(double) * npow3(nv) * nel);
  for (int64_t e = 0; e < nel; e++) tnsr3(tmp + npow3(nv) * e, nv, u + npow3(nu) * e, nu, A, A, A);
  memcpy(v, tmp, sizeof(double) * npow3(nv) * nel);
  free(tmp);
}

/* hsmg.f:2216-2232 H1MG_RSTR: r_l = J^T (wt .* r_{l+1}) (+ dssum) */
static void rstr(nko_mg *mg, int l, const double *rf, double *rc, int ifdssum) {
  int nf = mg->nh[l + 1], nc = mg->nh[l];
  int64_t nel = mg->nel, ntf = npow3(nf) * nel;
  double *t = malloc(sizeof(double) * ntf);
  for (int64_t q = 0; q < ntf; q++) t[q] = rf[q] * mg->rwt[l + 1][q];
  double *jt = malloc(sizeof(double) * nf * nc);
  for (int i = 0; i < nf; i++)
    for (int j = 0; j < nc; j++) jt[j + nc * i] = mg->jh[l][i + nf * j];
  tnsr_levels(rc, nc, t, nf, jt, nel);
  free(jt);
  free(t);
  if (ifdssum) nko_gs_op(mg->gs[l], rc, 0);
}

/* hsmg.f:1855-1949 H1MG_SOLVE (additive, sigma = 1) */
void nko_mg_solve(nko_mg *mg, double *z, double *rhs) {
  int L = mg->lmax;
  int64_t nel = mg->nel;
  double *e[MAXL + 1], *r[MAXL + 1];
  for (int l = 1; l <= L; l++) {
    e[l] = calloc((size_t)(npow3(mg->nh[l]) * nel), sizeof(double));
    r[l] = calloc((size_t)(npow3(mg->nh[l]) * nel), sizeof(double));
  }
  schwarz(mg, L, z, rhs, 1.0);
  memcpy(r[L], rhs, sizeof(double) * npow3(mg->nh[L]) * nel);
  for (int l = L - 1; l >= 2; l--) {
    rstr(mg, l, r[l + 1], r[l], 1);
    schwarz(mg, l, e[l], r[l], 1.0);
  }
  rstr(mg, 1, r[2], r[1], 0);
  int64_t n1 = npow3(mg->nh[1]) * nel;
  for (int64_t q = 0; q < n1; q++) r[1][q] *= mg->mask[1][q];
  coarse_solve(mg, e[1], r[1]);
  for (int64_t q = 0; q < n1; q++) e[1][q] *= mg->mask[1][q];
  for (int l = 2; l <= L; l++) { /* prolong and add */
    int nf = mg->nh[l], nc = mg->nh[l - 1];
    int64_t nt = npow3(nf) * nel;
    double *w = malloc(sizeof(double) * nt);
    tnsr_levels(w, nf, e[l - 1], nc, mg->jh[l - 1], nel);
    double *tgt = (l == L) ? z : e[l];
    for (int64_t q = 0; q < nt; q++) tgt[q] = tgt[q] + w[q];
    free(w);
  }
  /* ic.f:1871-1897 DSAVG */
  int64_t nt = npow3(mg->nh[L]) * nel;
  if (!(g_mutation & 1)) {
    nko_gs_op(mg->gs_fine, z, 0);
    for (int64_t q = 0; q < nt; q++) z[q] = z[q] * mg->vmult[q];
  }
  for (int l = 1; l <= L; l++) {
    free(e[l]);
    free(r[l]);
  }
}

/* gmres.f:284-302 AX: w = mask * dssum(axhelm(x)) */
static void ax(const nko_field *f, double *w, const double *x, const double *h1,
               const double *h2) {
  int64_t n = (int64_t)f->n * f->n * f->n * f->nel;
  nko_axhelm(f, w, x, h1, h2);
  nko_gs_op(f->gs, w, 0);
  for (int64_t i = 0; i < n; i++) w[i] = w[i] * f->mask[i];
}

/* navier1.f:223-255 ORTHO (ifvcor): v <- v - sum(v)/n over the local nodes */
static void ortho(double *v, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += v[i];
  const double rlam = s / (double)n;
  for (int64_t i = 0; i < n; i++) v[i] -= rlam;
}

/* the pressure's null space is present: no Dirichlet face and h2 = 0 */
static int null_space_on(const nko_mg *mg, const double *h2, int64_t n) {
  if (!mg->null_space) return 0;
  for (int64_t i = 0; i < n; i++)
    if (h2[i] != 0.0) return 0;
  return 1;
}

static double glsc3(const double *a, const double *b, const double *c, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i] * c[i];
  return s;
}

/*
 * hmholtz.f:2164-2281 HMH_FLEX_CG with h1mg_solve.  res: rhs in, solution
 * out.  Tolerance: chktcg1 value, capped by param(21) when > 0, relative
 * (|param21| * div0) when param21 < 0; istep == 0 forces 1e-4.  Returns the
 * iteration count; rnorm_hist[k] = residual norm after iteration k+1.
 */
int nko_flex_cg(const nko_field *f, nko_mg *mg, double *res, const double *h1,
                const double *h2, int maxit, double param21, int istep,
                double *rnorm_hist, double *div0_out) {
  int64_t n = (int64_t)f->n * f->n * f->n * f->nel;
  const double *wt = f->mult;
  double *r = malloc(sizeof(double) * n), *r1 = calloc((size_t)n, sizeof(double)),
         *p = calloc((size_t)n, sizeof(double)), *z = calloc((size_t)n, sizeof(double)),
         *w = malloc(sizeof(double) * n);
  double tolps = nko_chktcg1(f, 0.0, res, h1, h2);
  if (param21 > 0 && tolps > fabs(param21)) tolps = fabs(param21);
  if (istep == 0) tolps = 1.e-4;
  double tolpss = tolps;
  memcpy(r, res, sizeof(double) * n);
  memset(res, 0, sizeof(double) * n);
  double rho1 = 1.0;
  double div0 = sqrt(glsc3(r, wt, r, n) / f->vol);
  if (param21 < 0) tolpss = fabs(param21) * div0;
  if (div0_out) *div0_out = div0;
  int iter = 0;
  for (int k = 1; k <= maxit; k++) {
    nko_mg_solve(mg, z, r);
    for (int64_t i = 0; i < n; i++) r1[i] = r1[i] - r[i];
    double rho0 = rho1;
    rho1 = glsc3(z, wt, r, n);
    double rho2 = (g_mutation & 256) ? glsc3(z, wt, r1, n) : -glsc3(z, wt, r1, n);
    double beta = rho2 / rho0;
    memcpy(r1, r, sizeof(double) * n);
    for (int64_t i = 0; i < n; i++) p[i] = beta * p[i] + z[i];
    ax(f, w, p, h1, h2);
    double den = glsc3(w, wt, p, n);
    double alpha = rho1 / den;
    double rnorm = 0.0;
    for (int64_t i = 0; i < n; i++) {
      res[i] = res[i] + alpha * p[i];
      r[i] = r[i] - alpha * w[i];
      rnorm = rnorm + r[i] * r[i] * wt[i];
    }
    rnorm = sqrt(rnorm / f->vol);
    iter = iter + 1;
    if (rnorm_hist) rnorm_hist[k - 1] = rnorm;
    if (rnorm < tolpss) break;
  }
  if (null_space_on(mg, h2, n)) ortho(res, n); /* hmholtz.f:2272 */
  free(r);
  free(r1);
  free(p);
  free(z);
  free(w);
  return iter;
}

/*
 * gmres.f:304-545 HMH_GMRES with h1mg_solve (ifmgrid), restart m = lgmres.
 * ml = mu = 1 (uzawa_gmres_split, gmres.f:252-271), so no scaling.  res:
 * rhs in, solution out.  Tolerance as nko_flex_cg.  Returns the iteration
 * count; rnorm_hist[k] = |gamma_{j+1}| * norm_fac after Arnoldi step k+1.
 */
int nko_gmres(const nko_field *f, nko_mg *mg, double *res, const double *h1, const double *h2,
              int maxit, int m, double param21, int istep, double *rnorm_hist,
              double *div0_out) {
  int64_t n = (int64_t)f->n * f->n * f->n * f->nel;
  const double *wt = f->mult;
  double *x = calloc((size_t)n, sizeof(double)), *r = malloc(sizeof(double) * n),
         *w = malloc(sizeof(double) * n);
  double *v = malloc(sizeof(double) * n * (m + 1)), *z = malloc(sizeof(double) * n * m);
  double *H = calloc((size_t)m * m, sizeof(double)), *gam = calloc((size_t)m + 1, sizeof(double)),
         *c = calloc((size_t)m, sizeof(double)), *sn = calloc((size_t)m, sizeof(double));
#define HH(i, j) H[(i) + m * (j)]
  const double norm_fac = 1.0 / sqrt(f->vol);
  double tolps = nko_chktcg1(f, 0.0, res, h1, h2);
  if (param21 > 0 && tolps > fabs(param21)) tolps = fabs(param21);
  if (istep == 0) tolps = 1.e-4;
  double tolpss = tolps, div0 = 0.0, rnorm = 0.0;
  int iter = 0, iconv = 0, j = 0;
  const int nullsp = null_space_on(mg, h2, n);
  while (!iconv) {
    if (iter == 0) {
      memcpy(r, res, sizeof(double) * n);
    } else {
      ax(f, w, x, h1, h2);
      for (int64_t i = 0; i < n; i++) r[i] = res[i] - w[i];
    }
    gam[0] = sqrt(glsc3(r, r, wt, n));
    if (iter == 0) {
      div0 = gam[0] * norm_fac;
      if (param21 < 0) tolpss = fabs(param21) * div0;
    }
    rnorm = 0.0;
    if (gam[0] == 0.0) break;
    double temp = 1.0 / gam[0];
    for (int64_t i = 0; i < n; i++) v[i] = r[i] * temp;
    int restart = 0;
    for (j = 1; j <= m; j++) {
      iter = iter + 1;
      double *vj = v + n * (j - 1), *zj = z + n * (j - 1);
      memcpy(w, vj, sizeof(double) * n);
      nko_mg_solve(mg, zj, w);
      if (nullsp) ortho(zj, n); /* gmres.f:427 */
      ax(f, w, zj, h1, h2);
      for (int i = 1; i <= j; i++) { /* classical Gram-Schmidt, one pass */
        double sum = 0.0;
        const double *vi = v + n * (i - 1);
        for (int64_t q = 0; q < n; q++) sum = sum + w[q] * vi[q] * wt[q];
        HH(i - 1, j - 1) = sum;
      }
      for (int i = 1; i <= j; i++) {
        const double *vi = v + n * (i - 1);
        const double hij = HH(i - 1, j - 1);
        for (int64_t q = 0; q < n; q++) w[q] = w[q] + (-hij) * vi[q];
      }
      for (int i = 1; i <= ((g_mutation & 512) ? 0 : j - 1); i++) { /* previous Givens rotations */
        temp = HH(i - 1, j - 1);
        HH(i - 1, j - 1) = c[i - 1] * temp + sn[i - 1] * HH(i, j - 1);
        HH(i, j - 1) = -sn[i - 1] * temp + c[i - 1] * HH(i, j - 1);
      }
      double alpha = sqrt(glsc3(w, w, wt, n));
      rnorm = 0.0;
      if (alpha == 0.0) {
        iconv = 1;
        break;
      }
      double l = sqrt(HH(j - 1, j - 1) * HH(j - 1, j - 1) + alpha * alpha);
      temp = 1.0 / l;
      c[j - 1] = HH(j - 1, j - 1) * temp;
      sn[j - 1] = alpha * temp;
      HH(j - 1, j - 1) = l;
      gam[j] = -sn[j - 1] * gam[j - 1];
      gam[j - 1] = c[j - 1] * gam[j - 1];
      rnorm = fabs(gam[j]) * norm_fac;
      if (rnorm_hist) rnorm_hist[iter - 1] = rnorm;
      if (iter + 1 > maxit || rnorm < tolpss) {
        iconv = 1;
        break;
      }
      if (j == m) {
        restart = 1;
        break;
      }
      temp = 1.0 / alpha;
      double *vn = v + n * j;
      for (int64_t q = 0; q < n; q++) vn[q] = w[q] * temp;
    }
    (void)restart;
    if (j > m) j = m;
    /* back substitution c = H^-1 gamma, then x += sum c_i z_i */
    for (int k = j; k >= 1; k--) {
      temp = gam[k - 1];
      for (int i = j; i >= k + 1; i--) temp = temp - HH(k - 1, i - 1) * c[i - 1];
      c[k - 1] = temp / HH(k - 1, k - 1);
    }
    for (int i = 1; i <= j; i++) {
      const double *zi = z + n * (i - 1);
      for (int64_t q = 0; q < n; q++) x[q] = x[q] + c[i - 1] * zi[q];
    }
  }
#undef HH
  memcpy(res, x, sizeof(double) * n);
  if (nullsp) ortho(res, n); /* gmres.f:534 */
  if (div0_out) *div0_out = div0;
  free(x); free(r); free(w); free(v); free(z); free(H); free(gam); free(c); free(sn);
  return iter;
}

/* accessors for tests */
void nko_mg_apply_fdm(nko_mg *mg, int l, double *e_ext, double *r_ext) {
  do_fast(e_ext, r_ext, mg->fs[l], mg->fd[l], mg->nh[l] + 2, mg->nel);
}
const double *nko_mg_fdm_s(const nko_mg *mg, int l) { return mg->fs[l]; }
const double *nko_mg_fdm_d(const nko_mg *mg, int l) { return mg->fd[l]; }
const double *nko_mg_lengths(const nko_mg *mg, int which) {
  return which == 0 ? mg->ll : (which == 1 ? mg->lm : mg->lr);
}
const double *nko_mg_mask(const nko_mg *mg, int l) { return mg->mask[l]; }
const double *nko_mg_swt(const nko_mg *mg, int l) { return mg->swt[l]; }
const double *nko_mg_rwt(const nko_mg *mg, int l) { return mg->rwt[l]; }
const double *nko_mg_jh(const nko_mg *mg, int l) { return mg->jh[l]; }
int nko_mg_ncrs(const nko_mg *mg) { return mg->ncrs; }
void nko_mg_coarse_solve(const nko_mg *mg, double *x, const double *b) { coarse_solve(mg, x, b); }
void nko_mg_schwarz(nko_mg *mg, int l, double *e, double *r) { schwarz(mg, l, e, r, 1.0); }
int nko_mg_null_space(const nko_mg *mg) { return mg->null_space; }
