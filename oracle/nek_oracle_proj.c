/*
 * nek_oracle_proj.c -- CPU restatement of Nek5000's residual projection
 * around the Helmholtz solve (navier4.f:562-1325: hsolve, project1,
 * project1_a, iproj_chk, proj_matvec, proj_ortho, proj_ortho_full_cgs2,
 * project2, project2_a, proj_get_ivar, givens_rotation).
 *
 * TEST INFRASTRUCTURE ONLY (see nek_oracle.c).
 *
 * The projection space holds mmx = (mxprev-4)/2 solution vectors X and their
 * images B = mask*dssum(A X), orthonormal in the weighted inner product
 * <x,b>_w = sum x*w*b, plus the h1/h2 they were formed with.  bbar (A xbar)
 * is not kept: the reference stores it in its work array but never reads it.
 *
 * Parity status: pinned by what the mathematics fixes -- after every
 * project2, X^T W B = I; a right-hand side equal to A x for x in span(X)
 * projects to zero with xbar = x; after the Givens sweep the first column is
 * a multiple of the newest solution; a repeated solution is rejected as
 * rank deficient; a change of h1 rebuilds B = A X and re-orthonormalises.
 * hypot is the C library's (the reference's Fortran resolves hypot to the
 * Fortran 2008 intrinsic), which agrees to the last bit or one ulp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct nko_gs nko_gs;
void nko_gs_op(const nko_gs *h, double *u, int op);
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

#define MXPJ 32

typedef struct {
  int64_t n;
  int mmx, m;
  double *h1old, *h2old, *xbar, *xx, *bb;
} nko_proj;

/* navier4.f:1179-1206 proj_get_ivar: mmx = (mxprev-4)/2; arrays start zeroed
   (Nek's approx arrays live in zero-initialised commons) */
nko_proj *nko_proj_create(int64_t n, int mxprev) {
  int mmx = (mxprev - 4) / 2;
  if (mmx < 1 || mmx > MXPJ) return NULL;
  nko_proj *P = calloc(1, sizeof(nko_proj));
  P->n = n;
  P->mmx = mmx;
  P->h1old = calloc((size_t)n, sizeof(double));
  P->h2old = calloc((size_t)n, sizeof(double));
  P->xbar = calloc((size_t)n, sizeof(double));
  P->xx = calloc((size_t)n * mmx, sizeof(double));
  P->bb = calloc((size_t)n * mmx, sizeof(double));
  return P;
}

void nko_proj_free(nko_proj *P) {
  if (!P) return;
  free(P->h1old); free(P->h2old); free(P->xbar); free(P->xx); free(P->bb);
  free(P);
}

int nko_proj_m(const nko_proj *P) { return P->m; }
const double *nko_proj_x(const nko_proj *P, int k) { return P->xx + (int64_t)k * P->n; }
const double *nko_proj_b(const nko_proj *P, int k) { return P->bb + (int64_t)k * P->n; }
const double *nko_proj_xbar(const nko_proj *P) { return P->xbar; }

/* vlsc3: sum x*w*b */
static double vlsc3(const double *x, const double *w, const double *b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s = s + x[i] * w[i] * b[i];
  return s;
}

/* add2s2: a = a + c*b */
static void add2s2(double *a, const double *b, double c, int64_t n) {
  for (int64_t i = 0; i < n; i++) a[i] = a[i] + c * b[i];
}

static void cmult(double *a, double c, int64_t n) {
  for (int64_t i = 0; i < n; i++) a[i] = a[i] * c;
}

/* navier4.f:845-866 proj_matvec: b = mask * dssum(A x), isd = 1 */
static void proj_matvec(const nko_field *f, double *b, const double *x, const double *h1,
                        const double *h2, const double *msk, int64_t n) {
  nko_axhelm(f, b, x, h1, h2);
  nko_gs_op(f->gs, b, 0);
  for (int64_t i = 0; i < n; i++) b[i] = b[i] * msk[i];
}

/* navier4.f:1308-1325 givens_rotation */
static void givens_rotation(double a, double b, double *c, double *s, double *r) {
  if (b != 0.0) {
    double h = hypot(a, b);
    double d = 1.0 / h;
    *c = fabs(a) * d;
    *s = copysign(d, a) * b;
    *r = copysign(1.0, a) * h;
  } else {
    *c = 1.0;
    *s = 0.0;
    *r = a;
  }
}

/* navier4.f:809-843 iproj_chk: 1 when h1/h2 differ from the saved ones (and
   then saves them) or the mesh moves */
static int iproj_chk(nko_proj *P, const double *h1, const double *h2, int ifmvbd) {
  if (ifmvbd) return 1;
  double dh1 = 0.0, dh2 = 0.0;
  for (int64_t i = 0; i < P->n; i++) {
    dh1 = fmax(dh1, fabs(h1[i] - P->h1old[i]));
    dh2 = fmax(dh2, fabs(h2[i] - P->h2old[i]));
  }
  double dh = fmax(dh1, dh2);
  if (dh > 0) {
    memcpy(P->h1old, h1, sizeof(double) * P->n);
    memcpy(P->h2old, h2, sizeof(double) * P->n);
    return 1;
  }
  return 0;
}

/* navier4.f:1053-1121 proj_ortho_full_cgs2 */
static void proj_ortho_full_cgs2(nko_proj *P, const double *w) {
  int m = P->m;
  int64_t n = P->n;
  if (m <= 0) return;
  const double tol = 1.e-7;
  int flag[MXPJ];
  double alpha[MXPJ];
#define XX(k) (P->xx + (int64_t)((k) - 1) * n)
#define BB(k) (P->bb + (int64_t)((k) - 1) * n)
  for (int it = 1; it <= 2; it++) {
    for (int k = m; k >= 1; k--) {
      for (int j = m; j >= k; j--)
        alpha[j - 1] = .5 * (vlsc3(XX(j), w, BB(k), n) + vlsc3(BB(j), w, XX(k), n));
      for (int j = m; j >= k + 1; j--) {
        add2s2(XX(k), XX(j), -alpha[j - 1], n);
        add2s2(BB(k), BB(j), -alpha[j - 1], n);
      }
      double normp = sqrt(alpha[k - 1]);
      double normk = sqrt(vlsc3(XX(k), w, BB(k), n));
      if (normk > tol * normp) {
        double scl1 = 1.0 / normk;
        cmult(XX(k), scl1, n);
        cmult(BB(k), scl1, n);
        flag[k - 1] = 1;
      } else {
        flag[k - 1] = 0;
      }
    }
  }
  int k = 0;
  for (int j = 1; j <= m; j++) {
    if (flag[j - 1] == 1) {
      k = k + 1;
      if (k < j) {
        memcpy(XX(k), XX(j), sizeof(double) * n);
        memcpy(BB(k), BB(j), sizeof(double) * n);
      }
    }
  }
  P->m = k;
}

/* navier4.f:868-975 proj_ortho (O(nm) update with the new column m) */
static void proj_ortho(nko_proj *P, const double *w) {
  int m = P->m;
  int64_t n = P->n;
  if (m <= 0) return;
  double alpha[MXPJ], beta[MXPJ];
  for (int k = 1; k <= m; k++)
    alpha[k - 1] = 0.5 * (vlsc3(XX(k), w, BB(m), n) + vlsc3(BB(k), w, XX(m), n));
  double nrm = sqrt(alpha[m - 1]);
  for (int k = 1; k <= m - 1; k++) {
    add2s2(XX(m), XX(k), -alpha[k - 1], n);
    add2s2(BB(m), BB(k), -alpha[k - 1], n);
  }
  for (int k = 1; k <= m - 1; k++)
    beta[k - 1] = 0.5 * (vlsc3(XX(k), w, BB(m), n) + vlsc3(BB(k), w, XX(m), n));
  for (int k = 1; k <= m - 1; k++) {
    add2s2(XX(m), XX(k), -beta[k - 1], n);
    add2s2(BB(m), BB(k), -beta[k - 1], n);
    alpha[k - 1] = alpha[k - 1] + beta[k - 1];
  }
  alpha[m - 1] = sqrt(vlsc3(XX(m), w, BB(m), n));
  const double tol = 1.e-7;
  if (alpha[m - 1] > tol * nrm) {
    double scl1 = 1.0 / alpha[m - 1];
    cmult(XX(m), scl1, n);
    cmult(BB(m), scl1, n);
    for (int k = m; k >= 2; k--) {
      int h = k - 1;
      double c, s;
      givens_rotation(alpha[h - 1], alpha[k - 1], &c, &s, &nrm);
      alpha[h - 1] = nrm;
      double *xh = XX(h), *xk = XX(k), *bh = BB(h), *bk = BB(k);
      for (int64_t i = 0; i < n; i++) {
        double t1 = c * xh[i] + s * xk[i];
        xk[i] = -s * xh[i] + c * xk[i];
        xh[i] = t1;
        double t2 = c * bh[i] + s * bk[i];
        bk[i] = -s * bh[i] + c * bk[i];
        bh[i] = t2;
      }
    }
  } else {
    P->m = m - 1;
  }
}

/*
 * navier4.f:636-735 project1 + :737-807 project1_a: b <- b - B X^T W b (two
 * rounds of classical Gram-Schmidt), xbar = X X^T W b.  norms[0..1] (when
 * given and m > 0) = sqrt(b.w.b) before and after.  Returns 1 when the space
 * was rebuilt (h1/h2 changed).
 */
int nko_project1(const nko_field *f, nko_proj *P, double *b, const double *h1,
                 const double *h2, const double *msk, const double *w, int ifmvbd,
                 double *norms) {
  int64_t n = P->n;
  if (P->m <= 0) return 0;
  int ireset = iproj_chk(P, h1, h2, ifmvbd);
  double bb4 = sqrt(vlsc3(b, w, b, n));
  if (ireset == 1) {
    for (int j = 1; j <= P->m; j++) proj_matvec(f, BB(j), XX(j), h1, h2, msk, n);
    proj_ortho_full_cgs2(P, w);
  }
  int m = P->m;
  if (m > 0) {
    double alpha[MXPJ];
    for (int k = 1; k <= m; k++) alpha[k - 1] = vlsc3(XX(k), w, b, n);
    for (int64_t i = 0; i < n; i++) P->xbar[i] = XX(1)[i] * alpha[0];
    add2s2(b, BB(1), -alpha[0], n);
    for (int k = 2; k <= m; k++) {
      add2s2(P->xbar, XX(k), alpha[k - 1], n);
      add2s2(b, BB(k), -alpha[k - 1], n);
    }
    for (int k = 1; k <= m; k++) alpha[k - 1] = vlsc3(XX(k), w, b, n);
    for (int k = 1; k <= m; k++) {
      add2s2(P->xbar, XX(k), alpha[k - 1], n);
      add2s2(b, BB(k), -alpha[k - 1], n);
    }
  }
  double baf = sqrt(vlsc3(b, w, b, n));
  if (norms) {
    norms[0] = bb4;
    norms[1] = baf;
  }
  return ireset;
}

/*
 * navier4.f:1123-1177 project2 + project2_a: x <- x + xbar (when m > 0),
 * then x enters the space as column min(m+1, mmx) with its image, and
 * proj_ortho restores the orthonormality.
 */
void nko_project2(const nko_field *f, nko_proj *P, double *x, const double *h1,
                  const double *h2, const double *msk, const double *w) {
  int64_t n = P->n;
  if (P->m > 0)
    for (int64_t i = 0; i < n; i++) x[i] = x[i] + P->xbar[i];
  P->m = P->m + 1 < P->mmx ? P->m + 1 : P->mmx;
  memcpy(XX(P->m), x, sizeof(double) * n);
  proj_matvec(f, BB(P->m), XX(P->m), h1, h2, msk, n);
  proj_ortho(P, w);
}
#undef XX
#undef BB
