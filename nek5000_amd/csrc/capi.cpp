// C ABI of libnekamd: lifetime, meshes, host-side basis and numbering, and
// the nkh_/nkd_ entry points declared in include/nekamd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {

namespace {
thread_local char g_err[1024] = "";
std::mutex g_mu;
bool g_init = false;
int g_device = -1;
hipStream_t g_lib_stream = nullptr;
hipStream_t g_stream = nullptr;
std::vector<Mesh *> g_meshes;
}  // namespace

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

hipStream_t cur_stream() { return g_stream; }

Timers &timers() {
  static Timers t;
  return t;
}

Mesh *get_mesh(int id) {
  if (id < 0 || id >= (int)g_meshes.size()) return nullptr;
  return g_meshes[id];
}


// ---------------------------------------------------------------------------
// speclib.f restatement for the host-side basis (GLL points, weights, D).
// ---------------------------------------------------------------------------
namespace spec {

// speclib.f:479-518 JACOBF
void jacobf(double &poly, double &pder, double &polym1, double &pderm1, int n, double alp,
            double bet, double x) {
  const double apb = alp + bet;
  double p = 1.0, pd = 0.0, pl = 0.0, pdl = 0.0;
  poly = p;
  pder = pd;
  polym1 = 0.0;
  pderm1 = 0.0;
  if (n == 0) return;
  pl = p;
  pdl = pd;
  p = (alp - bet + (apb + 2.0) * x) / 2.0;
  pd = (apb + 2.0) / 2.0;
  if (n == 1) {
    poly = p;
    pder = pd;
    polym1 = pl;
    pderm1 = pdl;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const double dk = k;
    const double a1 = 2.0 * dk * (dk + apb) * (2.0 * dk + apb - 2.0);
    const double a2 = (2.0 * dk + apb - 1.0) * (alp * alp - bet * bet);
    const double b3 = (2.0 * dk + apb - 2.0);
    const double a3 = b3 * (b3 + 1.0) * (b3 + 2.0);
    const double a4 = 2.0 * (dk + alp - 1.0) * (dk + bet - 1.0) * (2.0 * dk + apb);
    const double pn = ((a2 + a3 * x) * p - a4 * pl) / a1;
    const double pdn = ((a2 + a3 * x) * pd - a4 * pdl + a3 * p) / a1;
    pl = p;
    p = pn;
    pdl = pd;
    pd = pdn;
  }
  poly = p;
  pder = pd;
  polym1 = pl;
  pderm1 = pdl;
}

// speclib.f:371-393 GAMMAF (only the arguments GLL needs)
double gammaf(double x) {
  const double pi = 4.0 * std::atan(1.0);
  if (x == 0.5) return std::sqrt(pi);
  if (x == 1.5) return std::sqrt(pi) / 2.0;
  if (x == 2.5) return 1.5 * std::sqrt(pi) / 2.0;
  if (x == 3.5) return 0.5 * (2.5 * (1.5 * std::sqrt(pi)));
  if (x == 3.0) return 2.0;
  if (x == 4.0) return 6.0;
  if (x == 5.0) return 24.0;
  if (x == 6.0) return 120.0;
  if (x == -0.5) return -2.0 * std::sqrt(pi);
  return 1.0;  // x == 1, 2 and anything else (as in the reference table)
}

// speclib.f:395-419 PNORMJ
double pnormj(int n, double a, double b) {
  const double dn = n, c = a + b + 1.0;
  if (n <= 1) {
    double prod = gammaf(dn + a) * gammaf(dn + b);
    prod = prod / (gammaf(dn) * gammaf(dn + a + b));
    return prod * std::pow(2.0, c) / (2.0 * dn + c);
  }
  double prod = gammaf(a + 1.0) * gammaf(b + 1.0);
  prod = prod / (2.0 * (1.0 + c) * gammaf(c + 1.0));
  prod = prod * (1.0 + a) * (2.0 + a);
  prod = prod * (1.0 + b) * (2.0 + b);
  for (int i = 3; i <= n; ++i) {
    const double di = i;
    prod = prod * ((di + a) * (di + b) / (di * (di + a + b)));
  }
  return prod * std::pow(2.0, c) / (2.0 * dn + c);
}

// speclib.f:421-477 JACG: Newton with deflation from Chebyshev-like guesses
void jacg(double *xj, int np, double a, double b) {
  const int n = np - 1;
  const double dth = 4.0 * std::atan(1.0) / (2.0 * n + 2.0);
  double xlast = 0.0, p, pd, pm1, pdm1;
  for (int j = 1; j <= np; ++j) {
    double x = std::cos((2.0 * (j - 1.0) + 1.0) * dth);
    if (j > 1) x = (x + xlast) / 2.0;
    for (int k = 1; k <= 10; ++k) {
      jacobf(p, pd, pm1, pdm1, np, a, b, x);
      double rs = 0.0;
      for (int i = 1; i <= j - 1; ++i) rs = rs + 1.0 / (x - xj[np - i]);
      const double dx = -p / (pd - rs * p);
      x = x + dx;
      if (std::fabs(dx) < 1.0e-12) break;
    }
    xj[np - j] = x;
    xlast = x;
  }
  for (int i = 0; i < np; ++i) {  // ascending selection sort
    int jm = i;
    double xm = 2.0;
    for (int j = i; j < np; ++j)
      if (xj[j] < xm) {
        xm = xj[j];
        jm = j;
      }
    std::swap(xj[i], xj[jm]);
  }
}

// speclib.f:283-369 ENDW1 / ENDW2 for alpha = beta = 0
double endw(int n, int which) {
  const double a = 0.0, b = 0.0, apb = 0.0;
  if (n == 0) return 0.0;
  double f1 = (which == 1) ? gammaf(a + 2.0) * gammaf(b + 1.0) / gammaf(apb + 3.0)
                           : gammaf(a + 1.0) * gammaf(b + 2.0) / gammaf(apb + 3.0);
  f1 = f1 * (apb + 2.0) * std::pow(2.0, apb + 2.0) / 2.0;
  if (n == 1) return f1;
  double fint1 = (which == 1) ? gammaf(a + 2.0) * gammaf(b + 1.0) / gammaf(apb + 3.0)
                              : gammaf(a + 1.0) * gammaf(b + 2.0) / gammaf(apb + 3.0);
  fint1 = fint1 * std::pow(2.0, apb + 2.0);
  double fint2 = gammaf(a + 2.0) * gammaf(b + 2.0) / gammaf(apb + 4.0);
  fint2 = fint2 * std::pow(2.0, apb + 3.0);
  double f2 = (which == 1)
                  ? (-2.0 * (b + 2.0) * fint1 + (apb + 4.0) * fint2) * (apb + 3.0) / 4.0
                  : (2.0 * (a + 2.0) * fint1 - (apb + 4.0) * fint2) * (apb + 3.0) / 4.0;
  if (n == 2) return f2;
  double f3 = 0.0;
  for (int i = 3; i <= n; ++i) {
    const double di = i - 1;
    const double abn = a + b + di, abnn = abn + di;
    const double a1 = -(2.0 * (di + a) * (di + b)) / (abn * abnn * (abnn + 1.0));
    const double a2 = (2.0 * (a - b)) / (abnn * (abnn + 2.0));
    const double a3 = (2.0 * (abn + 1.0)) / ((abnn + 2.0) * (abnn + 1.0));
    f3 = -(a2 * f2 + a1 * f1) / a3;
    f1 = f2;
    f2 = f3;
  }
  return f3;
}

// speclib.f:155-205 ZWGJD with np >= 2
void zwgjd(double *z, double *w, int np, double a, double b) {
  double p, pd, pm1, pdm1;
  if (np == 1) {
    z[0] = (b - a) / (a + b + 2.0);
    w[0] = gammaf(a + 1.0) * gammaf(b + 1.0) / gammaf(a + b + 2.0) * std::pow(2.0, a + b + 1.0);
    return;
  }
  jacg(z, np, a, b);
  const int n = np - 1, np1 = n + 1, np2 = n + 2;
  const double dnp1 = np1, dnp2 = np2;
  const double fac1 = dnp1 + a + b + 1.0, fac2 = fac1 + dnp1, fac3 = fac2 + 1.0;
  const double fnorm = pnormj(np1, a, b);
  const double rcoef = (fnorm * fac2 * fac3) / (2.0 * fac1 * dnp2);
  for (int i = 0; i < np; ++i) {
    jacobf(p, pd, pm1, pdm1, np2, a, b, z[i]);
    w[i] = -rcoef / (p * pdm1);
  }
}

// speclib.f:238-281 ZWGLJD (alpha = beta = 0)
void zwgll(double *z, double *w, int np) {
  const int n = np - 1;
  double p, pd, pm1, pdm1;
  if (n - 1 > 0) zwgjd(z + 1, w + 1, n - 1, 1.0, 1.0);
  z[0] = -1.0;
  z[np - 1] = 1.0;
  for (int i = 1; i < np - 1; ++i) w[i] = w[i] / (1.0 - z[i] * z[i]);
  jacobf(p, pd, pm1, pdm1, n, 0.0, 0.0, z[0]);
  w[0] = endw(n, 1) / (2.0 * pd);
  jacobf(p, pd, pm1, pdm1, n, 0.0, 0.0, z[np - 1]);
  w[np - 1] = endw(n, 2) / (2.0 * pd);
}

// speclib.f:875-905 PNLEG
double pnleg(double z, int n) {
  if (std::fabs(z) < 1.0e-25) z = 0.0;
  if (n == 0) return 1.0;
  double p1 = 1.0, p2 = z, p3 = z;
  for (int k = 1; k <= n - 1; ++k) {
    const double fk = k;
    p3 = ((2.0 * fk + 1.0) * z * p2 - fk * p1) / (fk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

// speclib.f:800-833 DGLL
void dgll(double *d, double *dt, const double *z, int nz) {
  const int n = nz - 1;
  if (nz == 1) {
    d[0] = 0.0;
    if (dt) dt[0] = 0.0;
    return;
  }
  const double fn = n, d0 = fn * (fn + 1.0) / 4.0;
  for (int i = 0; i < nz; ++i)
    for (int j = 0; j < nz; ++j) {
      double v = 0.0;
      if (i != j) v = pnleg(z[i], n) / (pnleg(z[j], n) * (z[i] - z[j]));
      if (i == j && i == 0) v = -d0;
      if (i == j && i == nz - 1) v = d0;
      d[i + nz * j] = v;
      if (dt) dt[j + nz * i] = v;
    }
}

}  // namespace spec

// ---------------------------------------------------------------------------
// navier8.f:2004-2360 SETVERT3D, one rank.  Edge and face tuples are ranked
// by their lexicographic position among the distinct tuples (gbtuple_rank8 at
// np = 1, navier8.f:1934-2002).
// ---------------------------------------------------------------------------
namespace numbering {

// navier8.f:1131-1179 I8RANK heap-sort; returns the 1-based position of the
// first index of the ranked order (the smallest entry under the heap's
// tie-breaking).
int i8rank_first(const int64_t *a, int n) {
  int ind[4];
  for (int j = 1; j <= n; ++j) ind[j - 1] = j;
  if (n <= 1) return 1;
  int l = n / 2 + 1, ir = n;
  for (;;) {
    int indx;
    int64_t q;
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
        return ind[0];
      }
    }
    int i = l, j = l + l;
    while (j <= ir) {
      if (j < ir && a[ind[j - 1] - 1] < a[ind[j] - 1]) j = j + 1;
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

struct Key3 {
  int64_t a, b, c;
  int64_t pos;
};

void rank_tuples(std::vector<Key3> &t, std::vector<int64_t> &rank) {
  rank.assign(t.size(), 0);
  std::sort(t.begin(), t.end(), [](const Key3 &x, const Key3 &y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    if (x.c != y.c) return x.c < y.c;
    return x.pos < y.pos;
  });
  int64_t r = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    if (i == 0 || t[i].a != t[i - 1].a || t[i].b != t[i - 1].b || t[i].c != t[i - 1].c) ++r;
    rank[(size_t)t[i].pos] = r;
  }
}

int64_t setvert3d(int64_t *glo, int nx, int64_t nel, const int64_t *vtx) {
  const int64_t nxyz = (int64_t)nx * nx * nx;
  int64_t ngv = 0;
  for (int64_t i = 0; i < 8 * nel; ++i) ngv = std::max(ngv, vtx[i]);
  for (int64_t e = 0; e < nel; ++e) {
    int64_t *g = glo + nxyz * e;
    for (int q = 0; q < (int)nxyz; ++q) g[q] = 0;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i)
          g[(nx - 1) * i + nx * (nx - 1) * j + nx * nx * (nx - 1) * k] =
              vtx[i + 2 * j + 4 * k + 8 * e];
  }
  if (nx == 2) return ngv;
  // Edge labels: r-edges (j,k), s-edges (i,k), t-edges (i,j), each listed in
  // the order of the reference's edge(0:1, ..., dir, e) array.
  std::vector<Key3> et((size_t)12 * nel);
  for (int64_t e = 0; e < nel; ++e) {
    const int64_t *v = vtx + 8 * e;
    int64_t lab[12][2];
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          const int64_t x = v[i + 2 * j + 4 * k];
          lab[0 + j + 2 * k][i] = x;  // r-edge: edge(i,j,k,1)
          lab[4 + i + 2 * k][j] = x;  // s-edge: edge(j,i,k,2)
          lab[8 + i + 2 * j][k] = x;  // t-edge: edge(k,i,j,3)
        }
    for (int q = 0; q < 12; ++q) {
      int64_t a = lab[q][0], b = lab[q][1];
      if (a > b) std::swap(a, b);
      et[(size_t)(12 * e + q)] = Key3{a, b, 0, 12 * e + q};
    }
  }
  std::vector<int64_t> erank;
  rank_tuples(et, erank);
  int64_t nue = 0;
  for (int64_t r : erank) nue = std::max(nue, r);
  const int noe = nx - 2;
  for (int64_t e = 0; e < nel; ++e) {
    int64_t *g = glo + nxyz * e;
    int q = 0;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j, ++q) {  // edges 1-4 along r
        const int64_t igv = ngv + (int64_t)noe * (erank[(size_t)(12 * e + q)] - 1);
        const int64_t i0 = (int64_t)nx * (nx - 1) * j + (int64_t)nx * nx * (nx - 1) * k;
        const bool fwd = g[i0] < g[i0 + nx - 1];
        for (int i = 2; i <= nx - 1; ++i)
          g[i0 + i - 1] = fwd ? igv + i - 1 : igv + 1 + noe - (i - 1);
      }
    for (int k = 0; k < 2; ++k)
      for (int i = 0; i < 2; ++i, ++q) {  // edges 5-8 along s
        const int64_t igv = ngv + (int64_t)noe * (erank[(size_t)(12 * e + q)] - 1);
        const int64_t i0 = (int64_t)(nx - 1) * i + (int64_t)nx * nx * (nx - 1) * k;
        const bool fwd = g[i0] < g[i0 + (int64_t)nx * (nx - 1)];
        for (int j = 2; j <= nx - 1; ++j)
          g[i0 + (int64_t)(j - 1) * nx] = fwd ? igv + j - 1 : igv + 1 + noe - (j - 1);
      }
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i, ++q) {  // edges 9-12 along t
        const int64_t igv = ngv + (int64_t)noe * (erank[(size_t)(12 * e + q)] - 1);
        const int64_t i0 = (int64_t)(nx - 1) * i + (int64_t)nx * (nx - 1) * j;
        const bool fwd = g[i0] < g[i0 + (int64_t)nx * nx * (nx - 1)];
        for (int k = 2; k <= nx - 1; ++k)
          g[i0 + (int64_t)(k - 1) * nx * nx] = fwd ? igv + k - 1 : igv + 1 + noe - (k - 1);
      }
  }
  ngv += nue * noe;
  // Faces (symmetric notation; TOPOL icface) ranked by their 3 smallest
  // corner ids.
  static const int icf[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
  std::vector<Key3> ft((size_t)6 * nel);
  for (int64_t e = 0; e < nel; ++e)
    for (int f = 0; f < 6; ++f) {
      int64_t c[4];
      for (int q = 0; q < 4; ++q) c[q] = vtx[8 * e + icf[f][q]];
      std::sort(c, c + 4);
      ft[(size_t)(6 * e + f)] = Key3{c[0], c[1], c[2], 6 * e + f};
    }
  std::vector<int64_t> frank;
  rank_tuples(ft, frank);
  int64_t nuf = 0;
  for (int64_t r : frank) nuf = std::max(nuf, r);
  // connect1.f DSSET skpdat: face f spans i0..i1 step is, j0..j1 step js with
  // node index i + nx*(j-1) (1-based)
  const int ny = nx, nz = nx;
  const int skp[6][6] = {{1, nx * (ny - 1) + 1, nx, 1, ny * (nz - 1) + 1, ny},
                         {nx, nx * (ny - 1) + nx, nx, 1, ny * (nz - 1) + 1, ny},
                         {1, nx, 1, 1, ny * (nz - 1) + 1, ny},
                         {1 + nx * (ny - 1), nx * ny, 1, 1, ny * (nz - 1) + 1, ny},
                         {1, nx, 1, 1, ny, 1},
                         {1 + nx * ny * (nz - 1), nx + nx * ny * (nz - 1), 1, 1, ny, 1}};
  const int nxx = nx * nx, nof = (nx - 2) * (nx - 2);
  for (int64_t e = 0; e < nel; ++e) {
    int64_t *g = glo + nxyz * e;
    for (int f = 0; f < 6; ++f) {
      int i0 = skp[f][0], i1 = skp[f][1], is = skp[f][2];
      int j0 = skp[f][3], j1 = skp[f][4], js = skp[f][5];
      const int64_t gvf[4] = {g[i0 + nx * (j0 - 1) - 1], g[i1 + nx * (j0 - 1) - 1],
                              g[i0 + nx * (j1 - 1) - 1], g[i1 + nx * (j1 - 1) - 1]};
      const int first = i8rank_first(gvf, 4);
      bool ifij = false;
      int idir = 1, jdir = 1;
      if (first == 1) {
        ifij = gvf[1] < gvf[2];
      } else if (first == 2) {
        idir = -1;
        ifij = gvf[0] < gvf[3];
      } else if (first == 3) {
        jdir = -1;
        ifij = gvf[3] < gvf[0];
      } else {
        idir = -1;
        jdir = -1;
        ifij = gvf[2] < gvf[1];
      }
      if (idir < 0) {
        std::swap(i0, i1);
        is = -is;
      }
      if (jdir < 0) {
        std::swap(j0, j1);
        js = -js;
      }
      const int64_t ig0 = ngv + (int64_t)nof * (frank[(size_t)(6 * e + f)] - 1);
      int k = 0;
      int64_t l = 0;
      auto visit = [&](int i, int j) {
        ++k;
        if (k > nx && k < nxx - nx && (k % nx) != 1 && (k % nx) != 0) {
          ++l;
          g[i + nx * (j - 1) - 1] = l + ig0;
        }
      };
      if (ifij) {
        for (int j = j0; js > 0 ? j <= j1 : j >= j1; j += js)
          for (int i = i0; is > 0 ? i <= i1 : i >= i1; i += is) visit(i, j);
      } else {
        for (int i = i0; is > 0 ? i <= i1 : i >= i1; i += is)
          for (int j = j0; js > 0 ? j <= j1 : j >= j1; j += js) visit(i, j);
      }
    }
  }
  ngv += nuf * nof;
  return ngv;
}

}  // namespace numbering

}  // namespace nekamd

using namespace nekamd;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *nkh_last_error(void) { return g_err; }

int nkh_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) {
    NK_CHECK(device == g_device, NKH_ERR_STATE, "nkh_init: already initialised on device %d",
             g_device);
    return NKH_OK;
  }
  int count = 0;
  NK_HIP(hipGetDeviceCount(&count));
  NK_CHECK(device >= 0 && device < count, NKH_ERR_ARG, "nkh_init: device %d of %d", device,
           count);
  NK_HIP(hipSetDevice(device));
  NK_HIP(hipStreamCreateWithFlags(&g_lib_stream, hipStreamNonBlocking));
  // device address of the allocation-free solve's static scratch
  NK_TRY(noalloc_scratch(nullptr));
  g_stream = g_lib_stream;
  g_device = device;
  g_init = true;
  return NKH_OK;
}

static void mesh_release(Mesh *m) {
  mg_free(*m);
  gs_free(m->gs);
  for (void *p : {(void *)m->Dd, (void *)m->g, (void *)m->bm1, (void *)m->binv, (void *)m->ws,
                  (void *)m->st, (void *)m->red, (void *)m->g32, (void *)m->bm1_32,
                  (void *)m->undeformed})
    if (p) (void)hipFree(p);
  delete m;
}

int nkh_finalize(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_init) return NKH_OK;
  (void)hipDeviceSynchronize();
  for (size_t i = 0; i < g_meshes.size(); ++i)
    if (g_meshes[i]) mesh_release(g_meshes[i]);
  g_meshes.clear();
  if (g_lib_stream) (void)hipStreamDestroy(g_lib_stream);
  g_lib_stream = g_stream = nullptr;
  g_init = false;
  g_device = -1;
  return NKH_OK;
}

int nkh_set_stream(void *s) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkh_set_stream before nkh_init");
  g_stream = (hipStream_t)s;  // NULL is HIP's null stream
  return NKH_OK;
}

int nkh_use_library_stream(void) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkh_use_library_stream before nkh_init");
  g_stream = g_lib_stream;
  return NKH_OK;
}

int nkh_synchronize(void) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkh_synchronize before nkh_init");
  NK_HIP(hipStreamSynchronize(g_stream));
  return NKH_OK;
}

int nkh_zwgll(double *z, double *w, int np) {
  NK_CHECK(z && w && np >= 2 && np <= 84, NKH_ERR_ARG, "nkh_zwgll: np=%d", np);
  spec::zwgll(z, w, np);
  return NKH_OK;
}

int nkh_dgll(double *d, double *dt, const double *z, int nz) {
  NK_CHECK(d && z && nz >= 1 && nz <= 84, NKH_ERR_ARG, "nkh_dgll: nz=%d", nz);
  spec::dgll(d, dt, z, nz);
  return NKH_OK;
}

int nkh_setvert3d(int64_t *glo_num, int64_t *ngv, int nx, int nel, const int64_t *vertex) {
  NK_CHECK(glo_num && vertex && nx >= 2 && nel >= 0, NKH_ERR_ARG, "nkh_setvert3d: bad args");
  const int64_t g = numbering::setvert3d(glo_num, nx, nel, vertex);
  if (ngv) *ngv = g;
  return NKH_OK;
}

int nkh_mesh_create(int lx1, int nel, int *mesh) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkh_mesh_create before nkh_init");
  NK_CHECK(lx1 >= 2 && lx1 <= kMaxNP && nel > 0 && mesh, NKH_ERR_ARG,
           "nkh_mesh_create: lx1=%d nel=%d", lx1, nel);
  Mesh *m = new Mesh();
  m->np = lx1;
  m->nel = nel;
  m->n = (int64_t)lx1 * lx1 * lx1 * nel;
  spec::zwgll(m->z, m->w, lx1);
  spec::dgll(m->D, nullptr, m->z, lx1);
  std::vector<double> basis((size_t)kMaxNP * kMaxNP + kMaxNP, 0.0);
  for (int t = 0; t < lx1 * lx1; ++t) basis[(size_t)t] = m->D[t];
  for (int t = 0; t < lx1; ++t) basis[(size_t)kMaxNP * kMaxNP + t] = m->w[t];
  if (hipMalloc(&m->Dd, sizeof(double) * basis.size()) != hipSuccess) {
    delete m;
    set_error("nkh_mesh_create: allocation failed");
    return NKH_ERR_NOMEM;
  }
  NK_HIP(hipMemcpy(m->Dd, basis.data(), sizeof(double) * basis.size(), hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> lk(g_mu);
  g_meshes.push_back(m);
  *mesh = (int)g_meshes.size() - 1;
  return NKH_OK;
}

int nkh_mesh_destroy(int id) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m, NKH_ERR_ARG, "nkh_mesh_destroy: bad mesh %d", id);
  NK_HIP(hipStreamSynchronize(g_stream));
  mesh_release(m);
  g_meshes[(size_t)id] = nullptr;
  return NKH_OK;
}

static int alloc_geom(Mesh &m) {
  if (!m.g) {
    if (hipMalloc(&m.g, sizeof(double) * 6 * m.n) != hipSuccess) {
      set_error("geometry allocation failed (%lld doubles)", (long long)(6 * m.n));
      return NKH_ERR_NOMEM;
    }
  }
  if (!m.bm1) NK_HIP(hipMalloc(&m.bm1, sizeof(double) * m.n));
  if (!m.binv) NK_HIP(hipMalloc(&m.binv, sizeof(double) * m.n));
  return NKH_OK;
}

// binvm1 = 1/dssum(bm1), volvm1 = sum(bm1) (coef.f:1500-1543; drive2.f geom)
static int finish_mass(Mesh &m) {
  if (!m.has_geom || !m.has_gs) return NKH_OK;
  hipStream_t s = g_stream;
  NK_HIP(hipMemcpyAsync(m.binv, m.bm1, sizeof(double) * m.n, hipMemcpyDeviceToDevice, s));
  NK_TRY(gs_op(m.gs, m.binv, 0, s));
  std::vector<double> h((size_t)m.n);
  NK_HIP(hipMemcpyAsync(h.data(), m.bm1, sizeof(double) * m.n, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  double vol = 0.0;
  for (double v : h) vol += v;
  m.vol = vol;
  NK_HIP(hipMemcpyAsync(h.data(), m.binv, sizeof(double) * m.n, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  for (double &v : h) v = 1.0 / v;
  NK_HIP(hipMemcpyAsync(m.binv, h.data(), sizeof(double) * m.n, hipMemcpyHostToDevice, s));
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

int nkd_mesh_geom_from_xyz(int id, const double *x, const double *y, const double *z) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m && x && y && z, NKH_ERR_ARG, "nkd_mesh_geom_from_xyz: bad args");
  NK_TRY(alloc_geom(*m));
  NK_TRY(launch_geom(*m, x, y, z, g_stream));
  NK_TRY(apply_undeformed(*m, g_stream));
  NK_HIP(hipStreamSynchronize(g_stream));
  m->has_geom = true;
  m->geom_version += 1;
  return finish_mass(*m);
}

int nkh_mesh_geom_from_xyz(int id, const double *x, const double *y, const double *z) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m && x && y && z, NKH_ERR_ARG, "nkh_mesh_geom_from_xyz: bad args");
  double *d = nullptr;
  NK_HIP(hipMalloc(&d, sizeof(double) * 3 * m->n));
  hipStream_t s = g_stream;
  int rc = NKH_OK;
  if (hipMemcpyAsync(d, x, sizeof(double) * m->n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d + m->n, y, sizeof(double) * m->n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d + 2 * m->n, z, sizeof(double) * m->n, hipMemcpyHostToDevice, s) !=
          hipSuccess) {
    set_error("nkh_mesh_geom_from_xyz: upload failed");
    rc = NKH_ERR_HIP;
  }
  if (rc == NKH_OK) rc = nkd_mesh_geom_from_xyz(id, d, d + m->n, d + 2 * m->n);
  (void)hipFree(d);
  return rc;
}

int nkh_mesh_set_geom(int id, const double *g1, const double *g2, const double *g3,
                      const double *g4, const double *g5, const double *g6, const double *bm1) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m && g1 && g2 && g3 && g4 && g5 && g6 && bm1, NKH_ERR_ARG,
           "nkh_mesh_set_geom: bad args");
  NK_TRY(alloc_geom(*m));
  const double *src[6] = {g1, g2, g3, g4, g5, g6};
  for (int f = 0; f < 6; ++f)
    NK_HIP(hipMemcpy(m->g + (size_t)f * m->n, src[f], sizeof(double) * m->n,
                     hipMemcpyHostToDevice));
  NK_HIP(hipMemcpy(m->bm1, bm1, sizeof(double) * m->n, hipMemcpyHostToDevice));
  NK_TRY(apply_undeformed(*m, g_stream));
  NK_HIP(hipStreamSynchronize(g_stream));
  m->has_geom = true;
  m->geom_version += 1;
  return finish_mass(*m);
}

int nkh_mesh_set_deformed(int id, const int *ifdfrm) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m, NKH_ERR_ARG, "nkh_mesh_set_deformed: bad mesh %d", id);
  if (!ifdfrm) {
    if (m->undeformed) {
      NK_HIP(hipStreamSynchronize(g_stream));
      (void)hipFree(m->undeformed);
      m->undeformed = nullptr;
    }
    return NKH_OK;  // factors already uploaded keep their cross terms
  }
  std::vector<uint8_t> f((size_t)m->nel);
  bool any = false;
  for (int64_t e = 0; e < m->nel; ++e) {
    f[(size_t)e] = ifdfrm[e] == 0 ? 1 : 0;
    any = any || f[(size_t)e];
  }
  if (!any) return nkh_mesh_set_deformed(id, nullptr);
  if (!m->undeformed) NK_HIP(hipMalloc(&m->undeformed, (size_t)m->nel));
  NK_HIP(hipMemcpy(m->undeformed, f.data(), (size_t)m->nel, hipMemcpyHostToDevice));
  if (m->has_geom) {
    NK_TRY(apply_undeformed(*m, g_stream));
    NK_HIP(hipStreamSynchronize(g_stream));
    m->geom_version += 1;
  }
  return NKH_OK;
}

int nkh_mesh_gs_setup(int id, const int64_t *glo_num) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m && glo_num, NKH_ERR_ARG, "nkh_mesh_gs_setup: bad args");
  mg_free(*m);  // the hierarchy shares the fine map
  NK_TRY(gs_build(m->gs, glo_num, m->n, m->np, m->nel));
  m->has_gs = true;
  return finish_mass(*m);
}

int nkh_mesh_info(int id, int64_t *n, double *vol, int64_t *nseg, int64_t *nsingle,
                  int *max_mult) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m, NKH_ERR_ARG, "nkh_mesh_info: bad mesh %d", id);
  if (n) *n = m->n;
  if (vol) *vol = m->vol;
  if (nseg) *nseg = m->gs.nseg;
  if (nsingle) *nsingle = m->gs.nsingle;
  if (max_mult) *max_mult = m->gs.max_mult;
  return NKH_OK;
}

int nkh_mesh_ptrs(int id, const double **bm1, const double **binvm1) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m && m->has_geom, NKH_ERR_STATE, "nkh_mesh_ptrs: mesh %d has no geometry", id);
  if (bm1) *bm1 = m->bm1;
  if (binvm1) *binvm1 = m->binv;
  return NKH_OK;
}

static int ready(Mesh *m, int id, bool need_gs) {
  NK_CHECK(g_init, NKH_ERR_STATE, "library not initialised");
  NK_CHECK(m, NKH_ERR_ARG, "bad mesh handle %d", id);
  NK_CHECK(m->has_geom, NKH_ERR_STATE, "mesh %d has no geometry", id);
  NK_CHECK(!need_gs || m->has_gs, NKH_ERR_STATE, "mesh %d has no gather-scatter map", id);
  return NKH_OK;
}

int nkd_axhelm(int id, double *au, const double *u, const double *h1, const double *h2,
               int isd) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, false));
  NK_CHECK(au && u && h1 && h2, NKH_ERR_ARG, "nkd_axhelm: null array");
  NK_CHECK(isd == 1 || isd == 2, NKH_ERR_ARG, "nkd_axhelm: isd=%d", isd);
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, g_stream));
  NK_TRY(launch_axhelm(*m, au, u, c1, c2, g_stream));
  return NKH_OK;
}

int nkh_axhelm_tuning(int lx1, int *variant, int *grid, int *tuned) {
  return axhelm_tuning(lx1, variant, grid, tuned);
}

int nkh_axhelm_set_tuning(int lx1, int variant, int grid) {
  return axhelm_set_tuning(lx1, variant, grid);
}

int nkh_axhelm_tuning_fused(int lx1, int *variant_cg, int *variant_flex) {
  return axhelm_tuning_fused(lx1, variant_cg, variant_flex);
}

int nkh_axhelm32_set_variant(int lx1, int variant) { return axhelm32_set_variant(lx1, variant); }

int nkh_axhelm32_variant(int lx1, int *variant) {
  NK_CHECK(variant, NKH_ERR_ARG, "nkh_axhelm32_variant: null pointer");
  return axhelm32_variant(lx1, variant);
}

int nkh_fdm_set_variant(int variant) { return fdm_set_variant(variant); }

int nkh_fdm_variant(int *variant) {
  NK_CHECK(variant, NKH_ERR_ARG, "nkh_fdm_variant: null pointer");
  *variant = fdm_variant();
  return NKH_OK;
}

int nkd_dssum(int id, double *u) {
  Mesh *m = get_mesh(id);
  NK_CHECK(g_init && m && m->has_gs, NKH_ERR_STATE, "nkd_dssum: mesh %d not ready", id);
  NK_CHECK(u, NKH_ERR_ARG, "nkd_dssum: null array");
  return gs_op(m->gs, u, 0, g_stream);
}

int nkd_dsop(int id, double *u, char op) {
  Mesh *m = get_mesh(id);
  NK_CHECK(g_init && m && m->has_gs, NKH_ERR_STATE, "nkd_dsop: mesh %d not ready", id);
  NK_CHECK(u, NKH_ERR_ARG, "nkd_dsop: null array");
  int k;
  switch (op) {
    case '+': k = 0; break;
    case '*': k = 1; break;
    case 'm': k = 2; break;
    case 'M': k = 3; break;
    default: set_error("nkd_dsop: op '%c'", op); return NKH_ERR_ARG;
  }
  return gs_op(m->gs, u, k, g_stream);
}

int nkd_adjoint_coef_grad(int id, double *g1, double *g2, const double *u, const double *lam) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, false));
  NK_CHECK(g1 && u && lam, NKH_ERR_ARG, "nkd_adjoint_coef_grad: null array");
  return launch_adj_grad(*m, g1, g2, u, lam, g_stream);
}

int nkd_setprec(int id, double *d, const double *h1, const double *h2) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(d && h1 && h2, NKH_ERR_ARG, "nkd_setprec: null array");
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, g_stream));
  return run_setprec(*m, d, c1, c2, g_stream);
}

// Nek's defaults (reader_par.f:26-27,76-79,86-89,159-162): param(22) =
// restol = 1e-7, param(42) = 0, param(21) = 0, param(93) = mxprev = 20,
// param(94) = 5, param(95) = 0, ifprojfld = .false.
static const nkh_params kDefaultParams = {1, 0, 1.0e-7, 1.0e-7, 0, 0, 0.0, 0,
                                          20, 5, 0, 0, 1, 1, 0};

// cggo from its head (hmholtz.f:611-653): the PRES branches copy f into x and
// run hmh_gmres (param42 = 0) or hmh_flex_cg (param42 = 2) in place; other
// fields take the Jacobi PCG.
static int cggo_branches(Mesh &m, bool pres, double *x, const double *f, Coef c1, Coef c2,
                         const double *mask, const double *mult, double tin, int maxit,
                         const nkh_params &p, nkh_result *res, double *history, hipStream_t s) {
  if (pres) {
    NK_HIP(hipMemcpyAsync(x, f, sizeof(double) * m.n, hipMemcpyDeviceToDevice, s));
    if (p.param42 == 0) return run_gmres(m, x, c1, c2, mult, mask, maxit, p, res, history);
    if (p.precision == 2)
      return run_flex_cg_mixed(m, x, c1, c2, mult, mask, maxit, p, res, history);
    return run_flex_cg(m, x, c1, c2, mult, mask, maxit, p, res, history);
  }
  return run_cggo(m, x, f, c1, c2, mask, mult, tin, maxit, p, res, history);
}

static int check_solver_params(const nkh_params &p, bool pres, const char *who) {
  NK_CHECK(p.precision == 0 || (pres && p.precision == 1) ||
               (p.precision == 2 && (!pres || p.param42 == 2)),
           NKH_ERR_UNSUPPORTED,
           "%s: precision=%d (1 for the multigrid pressure solvers, 2 for the Jacobi PCG and "
           "the flexible CG)",
           who, p.precision);
  NK_CHECK(!pres || p.param42 == 2 || p.param42 == 0, NKH_ERR_UNSUPPORTED,
           "%s: PRES with param42=%d (0 GMRES and 2 flexible CG with h1mg are built)", who,
           p.param42);
  return NKH_OK;
}

int nkd_cggo(int id, double *x, const double *f, const double *h1, const double *h2,
             const double *mask, const double *mult, double tin, int maxit,
             const nkh_params *prm, nkh_result *res, double *history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(x && f && h1 && h2 && mask && mult, NKH_ERR_ARG, "nkd_cggo: null array");
  NK_CHECK(maxit >= 0, NKH_ERR_ARG, "nkd_cggo: maxit=%d", maxit);
  const nkh_params &p = prm ? *prm : kDefaultParams;
  NK_CHECK(p.precision == 0 || p.precision == 2, NKH_ERR_UNSUPPORTED,
           "nkd_cggo: precision=%d", p.precision);
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, g_stream));
  return run_cggo(*m, x, f, c1, c2, mask, mult, tin, maxit, p, res, history);
}

int nkd_cggo_noalloc(int id, double *x, const double *f, const double *h1, const double *h2,
                     const double *mask, const double *mult, double tin, int maxit,
                     const nkh_params *prm, nkh_result *res, double *history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(x && f && h1 && h2 && mask && mult, NKH_ERR_ARG, "nkd_cggo_noalloc: null array");
  NK_CHECK(maxit >= 0, NKH_ERR_ARG, "nkd_cggo_noalloc: maxit=%d", maxit);
  const nkh_params &p = prm ? *prm : kDefaultParams;
  NK_CHECK(p.precision == 0, NKH_ERR_UNSUPPORTED,
           "nkd_cggo_noalloc: precision=%d (the allocation-free solve is FP64 only)",
           p.precision);
  return run_cggo_noalloc(*m, x, f, h1, h2, mask, mult, tin, maxit, p, res, history);
}

int nkd_hmholtz(int id, const char *name, double *u, double *rhs, const double *h1,
                const double *h2, const double *mask, const double *mult, double tli,
                int maxit, int isd, const nkh_params *prm, nkh_result *res, double *history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(u && rhs && h1 && h2 && mask && mult, NKH_ERR_ARG, "nkd_hmholtz: null array");
  NK_CHECK(isd == 1 || isd == 2, NKH_ERR_ARG, "nkd_hmholtz: isd=%d", isd);
  const nkh_params &p = prm ? *prm : kDefaultParams;
  const bool pres = name && std::strncmp(name, "PRES", 4) == 0;
  NK_TRY(check_solver_params(p, pres, "nkd_hmholtz"));
  hipStream_t s = g_stream;
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, s));
  // hmholtz.f:50-51: dssum(rhs); rhs *= mask
  NK_TRY(gs_op(m->gs, rhs, 0, s));
  NK_TRY(mul_inplace(rhs, mask, m->n, s));
  // hmholtz.f:52-55's chktcg1 only sets cggo's tolerance, which the PRES
  // branches (hmholtz.f:639-653) do not read
  if (pres) return cggo_branches(*m, true, u, rhs, c1, c2, mask, mult, 0.0, maxit, p, res, history, s);
  double tol = std::fabs(tli);
  if (p.param22 == 0.0 || p.istep <= 10) NK_TRY(run_chktcg1(*m, &tol, rhs, c1, c2, mask, mult));
  if (tli < 0) tol = tli;
  return run_cggo(*m, u, rhs, c1, c2, mask, mult, tol, maxit, p, res, history);
}

int nkd_hsolve(int id, const char *name, double *u, double *r, const double *h1,
               const double *h2, const double *vmk, const double *vml, double tol, int maxit,
               int isd, int proj, const nkh_params *prm, nkh_result *res, double *history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(u && r && h1 && h2 && vmk && vml, NKH_ERR_ARG, "nkd_hsolve: null array");
  const nkh_params &p = prm ? *prm : kDefaultParams;
  // navier4.f:585-606: capit(name), then the switches
  char cname[5] = {0, 0, 0, 0, 0};
  for (int i = 0; i < 4 && name && name[i]; ++i)
    cname[i] = (char)std::toupper((unsigned char)name[i]);
  bool ifstdh = !p.ifprojfld;
  int p945 = p.param94;
  if (std::strncmp(cname, "PRES", 4) == 0) {
    ifstdh = false;
    p945 = p.param95;
  }
  if (p.ifield > p.ldimt_proj + 1) ifstdh = true;
  if (p.param93 == 0) ifstdh = true;
  if (p945 == 0) ifstdh = true;
  if (p.istep < p945) ifstdh = true;
  if (ifstdh) return nkd_hmholtz(id, name, u, r, h1, h2, vmk, vml, tol, maxit, isd, prm, res, history);
  // cggo's PRES test compares the name as given (hmholtz.f:639)
  const bool pres = name && std::strncmp(name, "PRES", 4) == 0;
  NK_CHECK(proj >= 0, NKH_ERR_ARG,
           "nkd_hsolve: the switches select projection but no projection space was given");
  NK_CHECK(isd == 1 || isd == 2, NKH_ERR_ARG, "nkd_hsolve: isd=%d", isd);
  NK_TRY(check_solver_params(p, pres, "nkd_hsolve"));
  hipStream_t s = g_stream;
  // navier4.f:615-616: r *= vmk; dssum(r)
  NK_TRY(mul_inplace(r, vmk, m->n, s));
  NK_TRY(gs_op(m->gs, r, 0, s));
  NK_TRY(proj_project1(proj, r, h1, h2, vmk, vml, p.ifmvbd, nullptr, nullptr));
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, s));
  // hmhzpf (navier4.f:513-560): tol = tli or |param22|, chktcg1, cggo
  double t = tol;
  if (p.param22 != 0.0) t = std::fabs(p.param22);
  NK_TRY(run_chktcg1(*m, &t, r, c1, c2, vmk, vml));
  NK_TRY(cggo_branches(*m, pres, u, r, c1, c2, vmk, vml, t, maxit, p, res, history, s));
  return proj_project2(proj, u, h1, h2, vmk, vml);
}

int nkh_batch_create(int nb, const int *meshes, int *batch) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkh_batch_create: library not initialised");
  return batch_create(nb, meshes, batch);
}
int nkh_batch_destroy(int batch) { return batch_destroy(batch); }

int nkd_hmholtz_batch(int batch, const nkh_batch_solve *solves, int maxit,
                      const nkh_params *prm, nkh_result *res, int *graph_launches) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkd_hmholtz_batch: library not initialised");
  NK_CHECK(solves && res && maxit >= 0, NKH_ERR_ARG, "nkd_hmholtz_batch: bad arguments");
  const nkh_params &p = prm ? *prm : kDefaultParams;
  NK_CHECK(p.precision == 0 || p.precision == 2, NKH_ERR_UNSUPPORTED,
           "nkd_hmholtz_batch: precision=%d (0 FP64, 2 mixed)", p.precision);
  int nb = 0;
  std::vector<int> meshes;
  NK_TRY(batch_meshes(batch, &meshes));
  nb = (int)meshes.size();
  hipStream_t s = g_stream;
  std::vector<const double *> x(nb), f(nb), mk(nb), mu(nb);
  std::vector<Coef> c1(nb), c2(nb);
  std::vector<double> tin(nb);
  for (int k = 0; k < nb; ++k) {
    Mesh *m = get_mesh(meshes[k]);
    NK_TRY(ready(m, meshes[k], true));
    const nkh_batch_solve &q = solves[k];
    NK_CHECK(q.u && q.rhs && q.h1 && q.h2 && q.mask && q.mult, NKH_ERR_ARG,
             "nkd_hmholtz_batch: null array in solve %d", k);
    NK_TRY(classify_coefs(*m, q.h1, q.h2, &c1[k], &c2[k], s));
    // hmholtz.f:50-57: dssum(rhs); rhs *= mask; chktcg1; tli < 0 relative
    NK_TRY(gs_op(m->gs, q.rhs, 0, s));
    NK_TRY(mul_inplace(q.rhs, q.mask, m->n, s));
    double tol = std::fabs(q.tli);
    if (p.param22 == 0.0 || p.istep <= 10)
      NK_TRY(run_chktcg1(*m, &tol, q.rhs, c1[k], c2[k], q.mask, q.mult));
    if (q.tli < 0) tol = q.tli;
    x[k] = q.u;
    f[k] = q.rhs;
    mk[k] = q.mask;
    mu[k] = q.mult;
    tin[k] = tol;
  }
  return run_cggo_batch(batch, x.data(), f.data(), c1.data(), c2.data(), mk.data(), mu.data(),
                        tin.data(), maxit, p, res, graph_launches);
}

int nkh_re2_header(const char *path, int *nelgt, int *ndim, int *nelgv, int *nfields) {
  return re2_header(path, nelgt, ndim, nelgv, nfields);
}
int nkh_re2_read(const char *path, int nelgt, int ndim, int nfields, int *group, double *xc,
                 double *yc, double *zc, double *curve, char *ccurve, char *cbc, double *bc) {
  return re2_read(path, nelgt, ndim, nfields, group, xc, yc, zc, curve, ccurve, cbc, bc);
}
int nkh_map_header(const char *path, int *nel, int *nactive) {
  return map_header(path, nel, nactive);
}
int nkh_map_read(const char *path, int nel, int ndim, int *part, int64_t *vertex) {
  return map_read(path, nel, ndim, part, vertex);
}
int nkh_mesh_vertex_ids(int nel, int ndim, const double *xc, const double *yc, const double *zc,
                        const char *cbc, const double *bc, double tol_rel, int64_t *vertex,
                        int64_t *nvert) {
  return mesh_vertex_ids(nel, ndim, xc, yc, zc, cbc, bc, tol_rel, vertex, nvert);
}
int nkh_mesh_xyz_from_corners(int lx1, int nel, const double *xc, const double *yc,
                              const double *zc, const char *ccurve, const double *curve,
                              double *xm1, double *ym1, double *zm1) {
  return mesh_xyz_from_corners(lx1, nel, xc, yc, zc, ccurve, curve, xm1, ym1, zm1);
}

int nkh_proj_create(int mesh, int mxprev, int *proj) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkh_proj_create: library not initialised");
  return proj_create(mesh, mxprev, proj);
}
int nkh_proj_destroy(int proj) { return proj_destroy(proj); }
int nkh_proj_info(int proj, int *m, int *mmx) { return proj_info(proj, m, mmx); }
int nkh_proj_column(int proj, int k, int which, const double **ptr) {
  return proj_column(proj, k, which, ptr);
}
int nkd_project1(int proj, double *b, const double *h1, const double *h2, const double *msk,
                 const double *w, int ifmvbd, double *norms, int *ireset) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkd_project1: library not initialised");
  return proj_project1(proj, b, h1, h2, msk, w, ifmvbd, norms, ireset);
}
int nkd_project2(int proj, double *x, const double *h1, const double *h2, const double *msk,
                 const double *w) {
  NK_CHECK(g_init, NKH_ERR_STATE, "nkd_project2: library not initialised");
  return proj_project2(proj, x, h1, h2, msk, w);
}

int nkh_mg_setup(int id, const int *fbc, const int64_t *vertex, const double *xm1,
                 const double *ym1, const double *zm1) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(fbc && vertex && xm1 && ym1 && zm1, NKH_ERR_ARG, "nkh_mg_setup: null array");
  return mg_setup(*m, fbc, vertex, xm1, ym1, zm1);
}

int nkd_mg_solve(int id, double *z, double *rhs) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(z && rhs, NKH_ERR_ARG, "nkd_mg_solve: null array");
  NK_TRY(mg_mask_fine(*m, rhs, g_stream));  // h1mg_mask of the fine level, in place
  return mg_solve(*m, z, rhs, g_stream);
}

int nkd_hmh_flex_cg(int id, double *res, const double *h1, const double *h2, const double *wt,
                    const double *pmask, int maxit, const nkh_params *prm, nkh_result *out,
                    double *rnorm_history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(res && h1 && h2 && wt && pmask, NKH_ERR_ARG, "nkd_hmh_flex_cg: null array");
  NK_CHECK(maxit >= 0, NKH_ERR_ARG, "nkd_hmh_flex_cg: maxit=%d", maxit);
  const nkh_params &p = prm ? *prm : kDefaultParams;
  NK_CHECK(p.precision >= 0 && p.precision <= 2, NKH_ERR_UNSUPPORTED,
           "nkd_hmh_flex_cg: precision=%d", p.precision);
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, g_stream));
  if (p.precision == 2)
    return run_flex_cg_mixed(*m, res, c1, c2, wt, pmask, maxit, p, out, rnorm_history);
  return run_flex_cg(*m, res, c1, c2, wt, pmask, maxit, p, out, rnorm_history);
}

int nkd_hmh_gmres(int id, double *res, const double *h1, const double *h2, const double *wt,
                  const double *pmask, int maxit, const nkh_params *prm, nkh_result *out,
                  double *rnorm_history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  NK_CHECK(res && h1 && h2 && wt && pmask, NKH_ERR_ARG, "nkd_hmh_gmres: null array");
  NK_CHECK(maxit >= 0, NKH_ERR_ARG, "nkd_hmh_gmres: maxit=%d", maxit);
  const nkh_params &p = prm ? *prm : kDefaultParams;
  NK_CHECK(p.precision == 0 || p.precision == 1, NKH_ERR_UNSUPPORTED,
           "nkd_hmh_gmres: precision=%d", p.precision);
  Coef c1, c2;
  NK_TRY(classify_coefs(*m, h1, h2, &c1, &c2, g_stream));
  return run_gmres(*m, res, c1, c2, wt, pmask, maxit, p, out, rnorm_history);
}

int nkh_mg_info(int id, int *lmax, int *nh, int *ncrs, double *crs_avg_iters) {
  Mesh *m = get_mesh(id);
  NK_CHECK(m, NKH_ERR_ARG, "nkh_mg_info: bad mesh %d", id);
  return mg_info(*m, lmax, nh, ncrs, crs_avg_iters);
}

int nkh_timers_enable(int on) {
  timers().on = on != 0;
  return NKH_OK;
}

int nkh_timers_reset(void) {
  const bool on = timers().on;
  timers() = Timers();
  timers().on = on;
  return NKH_OK;
}

int nkh_timers_read(const char *name, double *seconds, int64_t *calls) {
  NK_CHECK(name, NKH_ERR_ARG, "nkh_timers_read: null name");
  const Timers &t = timers();
  double s;
  int64_t c;
  if (std::strcmp(name, "axhm") == 0) {
    s = t.taxhm;
    c = t.naxhm;
  } else if (std::strcmp(name, "dsum") == 0) {
    s = t.tdsum;
    c = t.ndsum;
  } else if (std::strcmp(name, "hmhz") == 0) {
    s = t.thmhz;
    c = t.nhmhz;
  } else if (std::strcmp(name, "fdm") == 0) {
    s = t.tfdm;
    c = t.nfdm;
  } else if (std::strcmp(name, "mgv") == 0) {
    s = t.tmgv;
    c = t.nmgv;
  } else {
    set_error("nkh_timers_read: unknown timer '%s'", name);
    return NKH_ERR_ARG;
  }
  if (seconds) *seconds = s;
  if (calls) *calls = c;
  return NKH_OK;
}

// ---- host-pointer wrappers: copy in, run, copy out ----
struct DevBuf {
  double *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int nkh_axhelm(int id, double *au, const double *u, const double *h1, const double *h2,
               int isd) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, false));
  const size_t b = sizeof(double) * m->n;
  DevBuf d;
  NK_HIP(hipMalloc(&d.p, 4 * b));
  hipStream_t s = g_stream;
  NK_HIP(hipMemcpyAsync(d.p + m->n, u, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(d.p + 2 * m->n, h1, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(d.p + 3 * m->n, h2, b, hipMemcpyHostToDevice, s));
  NK_TRY(nkd_axhelm(id, d.p, d.p + m->n, d.p + 2 * m->n, d.p + 3 * m->n, isd));
  NK_HIP(hipMemcpyAsync(au, d.p, b, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

int nkh_dssum(int id, double *u) {
  Mesh *m = get_mesh(id);
  NK_CHECK(g_init && m && m->has_gs, NKH_ERR_STATE, "nkh_dssum: mesh %d not ready", id);
  const size_t b = sizeof(double) * m->n;
  DevBuf d;
  NK_HIP(hipMalloc(&d.p, b));
  hipStream_t s = g_stream;
  NK_HIP(hipMemcpyAsync(d.p, u, b, hipMemcpyHostToDevice, s));
  NK_TRY(nkd_dssum(id, d.p));
  NK_HIP(hipMemcpyAsync(u, d.p, b, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

int nkh_hmholtz(int id, const char *name, double *u, double *rhs, const double *h1,
                const double *h2, const double *mask, const double *mult, double tli,
                int maxit, int isd, const nkh_params *prm, nkh_result *res, double *history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  const int64_t n = m->n;
  const size_t b = sizeof(double) * n;
  DevBuf d;
  NK_HIP(hipMalloc(&d.p, 6 * b));
  hipStream_t s = g_stream;
  double *du = d.p, *dr = d.p + n, *d1 = d.p + 2 * n, *d2 = d.p + 3 * n, *dk = d.p + 4 * n,
         *dm = d.p + 5 * n;
  NK_HIP(hipMemcpyAsync(dr, rhs, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(d1, h1, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(d2, h2, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(dk, mask, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(dm, mult, b, hipMemcpyHostToDevice, s));
  NK_TRY(nkd_hmholtz(id, name, du, dr, d1, d2, dk, dm, tli, maxit, isd, prm, res, history));
  NK_HIP(hipMemcpyAsync(u, du, b, hipMemcpyDeviceToHost, s));
  NK_HIP(hipMemcpyAsync(rhs, dr, b, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

int nkh_hsolve(int id, const char *name, double *u, double *r, const double *h1,
               const double *h2, const double *vmk, const double *vml, double tol, int maxit,
               int isd, int proj, const nkh_params *prm, nkh_result *res, double *history) {
  Mesh *m = get_mesh(id);
  NK_TRY(ready(m, id, true));
  const int64_t n = m->n;
  const size_t b = sizeof(double) * n;
  DevBuf d;
  NK_HIP(hipMalloc(&d.p, 6 * b));
  hipStream_t s = g_stream;
  double *du = d.p, *dr = d.p + n, *d1 = d.p + 2 * n, *d2 = d.p + 3 * n, *dk = d.p + 4 * n,
         *dm = d.p + 5 * n;
  NK_HIP(hipMemcpyAsync(dr, r, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(d1, h1, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(d2, h2, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(dk, vmk, b, hipMemcpyHostToDevice, s));
  NK_HIP(hipMemcpyAsync(dm, vml, b, hipMemcpyHostToDevice, s));
  NK_TRY(nkd_hsolve(id, name, du, dr, d1, d2, dk, dm, tol, maxit, isd, proj, prm, res, history));
  NK_HIP(hipMemcpyAsync(u, du, b, hipMemcpyDeviceToHost, s));
  NK_HIP(hipMemcpyAsync(r, dr, b, hipMemcpyDeviceToHost, s));
  NK_HIP(hipStreamSynchronize(s));
  return NKH_OK;
}

}  // extern "C"
