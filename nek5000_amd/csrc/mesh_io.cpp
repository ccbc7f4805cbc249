// Mesh ingestion for real Nek cases (host code): the .re2 binary mesh reader
// (reader_re2.f), the .map/.ma2 partition and vertex file (map2.f read_map),
// element connectivity by corner coincidence (what the map or the parallel
// connectivity step supplies to Nek as `vertex`, for cases without a map),
// and the GLL coordinates of straight-sided, midside-curved ('m'),
// circular-arc ('C') and spherical-face ('s') elements (genxyz.f).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "nekamd_internal.h"

namespace nekamd {
namespace {

struct Re2File {
  std::vector<unsigned char> data;
  int version = 0;  // 1..4
  int ws = 8;       // word size of the records
  bool swap = false;
  int nelgt = 0, ndim = 0, nelgv = 0, nbc_hdr = 0;
  size_t mesh_off = 84;
};

void bswap(unsigned char *p, int w) { std::reverse(p, p + w); }

int load(const char *path, Re2File &f) {
  NK_CHECK(path, NKH_ERR_ARG, "re2: null path");
  FILE *fp = std::fopen(path, "rb");
  NK_CHECK(fp, NKH_ERR_ARG, "re2: cannot open %s", path);
  std::fseek(fp, 0, SEEK_END);
  const long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (sz < 84) {
    std::fclose(fp);
    set_error("re2: %s is too short for a header", path);
    return NKH_ERR_ARG;
  }
  f.data.resize((size_t)sz);
  const size_t got = std::fread(f.data.data(), 1, (size_t)sz, fp);
  std::fclose(fp);
  NK_CHECK(got == (size_t)sz, NKH_ERR_ARG, "re2: short read on %s", path);
  // reader_re2.f:543-620 read_re2_hdr: 80-character header, then the
  // float 6.54321 that reveals the byte order
  char hdr[81];
  std::memcpy(hdr, f.data.data(), 80);
  hdr[80] = 0;
  NK_CHECK(hdr[0] == '#' && hdr[1] == 'v' && hdr[2] == '0' && hdr[3] == '0' && hdr[4] >= '1' &&
               hdr[4] <= '4',
           NKH_ERR_ARG, "re2: %s has no #v00N header", path);
  f.version = hdr[4] - '0';
  if (f.version == 4) {
    char ver[8];
    NK_CHECK(std::sscanf(hdr, "%5s %d %d %d %d", ver, &f.nelgt, &f.ndim, &f.nelgv,
                         &f.nbc_hdr) == 5,
             NKH_ERR_ARG, "re2: bad #v004 header");
  } else {
    // format (a5,i9,i3,i9)
    auto field = [&](int a, int w) {
      std::string s(hdr + a, hdr + a + w);
      return std::atoi(s.c_str());
    };
    f.nelgt = field(5, 9);
    f.ndim = field(14, 3);
    f.nelgv = field(17, 9);
  }
  f.ws = f.version >= 2 ? 8 : 4;
  float test;
  std::memcpy(&test, f.data.data() + 80, 4);
  if (std::fabs(test - 6.54321f) > 1e-5f) {
    unsigned char b[4];
    std::memcpy(b, f.data.data() + 80, 4);
    bswap(b, 4);
    std::memcpy(&test, b, 4);
    NK_CHECK(std::fabs(test - 6.54321f) <= 1e-5f, NKH_ERR_ARG, "re2: byte-order test failed");
    f.swap = true;
  }
  NK_CHECK(f.nelgt > 0 && f.nelgv > 0 && f.nelgv <= f.nelgt, NKH_ERR_ARG,
           "re2: invalid element counts %d/%d", f.nelgt, f.nelgv);
  NK_CHECK(f.ndim == 2 || f.ndim == 3, NKH_ERR_ARG, "re2: invalid dimension %d", f.ndim);
  return NKH_OK;
}

// one record word as a double (8-byte words) or an int/float (4-byte words)
double word_real(const Re2File &f, size_t off) {
  unsigned char b[8];
  std::memcpy(b, f.data.data() + off, (size_t)f.ws);
  if (f.swap) bswap(b, f.ws);
  if (f.ws == 8) {
    double d;
    std::memcpy(&d, b, 8);
    return d;
  }
  float x;
  std::memcpy(&x, b, 4);
  return x;
}

int word_int(const Re2File &f, size_t off) {
  unsigned char b[8];
  std::memcpy(b, f.data.data() + off, (size_t)f.ws);
  if (f.swap) bswap(b, f.ws);
  if (f.ws == 8) {
    double d;
    std::memcpy(&d, b, 8);
    return (int)d;
  }
  int32_t i;
  std::memcpy(&i, b, 4);
  return i;
}

size_t mesh_bytes(const Re2File &f) {
  const int rec = 1 + f.ndim * (1 << f.ndim);
  return (size_t)f.nelgt * rec * f.ws;
}

// offsets of the curve section and of each boundary-condition section
int sections(const Re2File &f, size_t *curve_off, std::vector<size_t> &bc_off) {
  size_t off = f.mesh_off + mesh_bytes(f);
  const size_t n = f.data.size();
  NK_CHECK(off + f.ws <= n, NKH_ERR_ARG, "re2: truncated before the curve section");
  *curve_off = off;
  const int ncurve = word_int(f, off);
  NK_CHECK(ncurve >= 0 && ncurve <= 12 * f.nelgt, NKH_ERR_ARG, "re2: bad curve count %d",
           ncurve);
  off += f.ws + (size_t)ncurve * 8 * f.ws;
  bc_off.clear();
  while (off + f.ws <= n) {
    const int nbc = word_int(f, off);
    NK_CHECK(nbc >= 0 && nbc <= 6 * f.nelgt, NKH_ERR_ARG, "re2: bad boundary count %d", nbc);
    const size_t end = off + f.ws + (size_t)nbc * 8 * f.ws;
    NK_CHECK(end <= n, NKH_ERR_ARG, "re2: truncated boundary section");
    bc_off.push_back(off);
    off = end;
  }
  NK_CHECK(off == n, NKH_ERR_ARG, "re2: %zu trailing bytes", n - off);
  return NKH_OK;
}

}  // namespace

int re2_header(const char *path, int *nelgt, int *ndim, int *nelgv, int *nfields) {
  Re2File f;
  NK_TRY(load(path, f));
  size_t coff;
  std::vector<size_t> boffs;
  NK_TRY(sections(f, &coff, boffs));
  if (nelgt) *nelgt = f.nelgt;
  if (ndim) *ndim = f.ndim;
  if (nelgv) *nelgv = f.nelgv;
  if (nfields) *nfields = (int)boffs.size();
  return NKH_OK;
}

// reader_re2.f:68-157 (mesh), :159-286 (curves), :288-428 (boundary
// conditions), with buf_to_xyz / buf_to_curve / buf_to_bc (:430-541).
int re2_read(const char *path, int nelgt, int ndim, int nfields, int *group, double *xc,
             double *yc, double *zc, double *curve, char *ccurve, char *cbc, double *bc) {
  Re2File f;
  NK_TRY(load(path, f));
  NK_CHECK(nelgt == f.nelgt && ndim == f.ndim, NKH_ERR_ARG,
           "re2_read: sizes %d/%d do not match the file's %d/%d", nelgt, ndim, f.nelgt, f.ndim);
  size_t coff;
  std::vector<size_t> boffs;
  NK_TRY(sections(f, &coff, boffs));
  NK_CHECK(nfields >= 0 && nfields <= (int)boffs.size(), NKH_ERR_ARG,
           "re2_read: %d boundary fields requested, %zu present", nfields, boffs.size());
  const int nv = 1 << ndim;
  const int rec = 1 + ndim * nv;
  for (int e = 0; e < nelgt; ++e) {
    const size_t o = f.mesh_off + (size_t)e * rec * f.ws;
    if (group) group[e] = word_int(f, o);
    for (int c = 0; c < nv; ++c) {
      if (xc) xc[(size_t)nv * e + c] = word_real(f, o + (size_t)(1 + c) * f.ws);
      if (yc) yc[(size_t)nv * e + c] = word_real(f, o + (size_t)(1 + nv + c) * f.ws);
      if (zc && ndim == 3) zc[(size_t)nv * e + c] = word_real(f, o + (size_t)(1 + 2 * nv + c) * f.ws);
    }
  }
  // curves: curve(5,12,nel), ccurve(12,nel)
  if (curve) std::memset(curve, 0, sizeof(double) * 5 * 12 * (size_t)nelgt);
  if (ccurve) std::memset(ccurve, ' ', 12 * (size_t)nelgt);
  const int ncurve = word_int(f, coff);
  for (int q = 0; q < ncurve; ++q) {
    const size_t o = coff + f.ws + (size_t)q * 8 * f.ws;
    const int eg = word_int(f, o), ed = word_int(f, o + f.ws);
    NK_CHECK(eg >= 1 && eg <= nelgt && ed >= 1 && ed <= 12, NKH_ERR_ARG,
             "re2: curve record %d names element %d edge %d", q, eg, ed);
    const size_t slot = (size_t)(eg - 1) * 12 + (ed - 1);
    if (curve)
      for (int k = 0; k < 5; ++k) curve[slot * 5 + k] = word_real(f, o + (size_t)(2 + k) * f.ws);
    if (ccurve) ccurve[slot] = (char)f.data[o + 7 * f.ws];
  }
  // boundary conditions per field: cbc(3 chars, 6, nel), bc(5, 6, nel);
  // faces without a record are 'E  ' (reader_re2.f:388-392)
  for (int fld = 0; fld < nfields; ++fld) {
    char *cb = cbc ? cbc + (size_t)fld * 18 * nelgt : nullptr;
    double *bv = bc ? bc + (size_t)fld * 30 * nelgt : nullptr;
    if (cb)
      for (size_t q = 0; q < 6 * (size_t)nelgt; ++q) {
        cb[3 * q] = 'E';
        cb[3 * q + 1] = ' ';
        cb[3 * q + 2] = ' ';
      }
    if (bv) std::memset(bv, 0, sizeof(double) * 30 * (size_t)nelgt);
    const size_t bo = boffs[(size_t)fld];
    const int nbc = word_int(f, bo);
    for (int q = 0; q < nbc; ++q) {
      const size_t o = bo + f.ws + (size_t)q * 8 * f.ws;
      const int eg = word_int(f, o), fc = word_int(f, o + f.ws);
      NK_CHECK(eg >= 1 && eg <= nelgt && fc >= 1 && fc <= 6, NKH_ERR_ARG,
               "re2: boundary record %d names element %d face %d", q, eg, fc);
      const size_t slot = (size_t)(eg - 1) * 6 + (fc - 1);
      if (bv)
        for (int k = 0; k < 5; ++k) bv[slot * 5 + k] = word_real(f, o + (size_t)(2 + k) * f.ws);
      if (cb) std::memcpy(cb + 3 * slot, f.data.data() + o + 7 * f.ws, 3);
    }
  }
  return NKH_OK;
}

// ---- map files -----------------------------------------------------------
// map2.f:712-942 read_map: the case's .map (text) or .ma2 (binary) file
// from genmap: header (version, element count, active count), then per
// element its partition (0-based) and its 2^ndim vertex ids in the
// hypercube (lexicographic) corner order (genmap.f:1641-1712 dmp_mapfile,
// genmap.f:362 h2s).  .ma2: a 132-byte text header read as (a5, 2i12), the
// real*4 6.54321 that gives the byte order, then (1 + nv) int32 per element.
namespace {

bool has_suffix(const char *path, const char *suf) {
  const size_t n = std::strlen(path), m = std::strlen(suf);
  return n >= m && std::strcmp(path + n - m, suf) == 0;
}

struct MapFile {
  int nel = 0, nactive = 0;
  bool binary = false, swap = false;
  std::vector<unsigned char> data;  // .ma2 bytes
  std::vector<long long> words;     // .map integers after the header line
};

int map_load(const char *path, MapFile &m) {
  NK_CHECK(path, NKH_ERR_ARG, "map: null path");
  m.binary = has_suffix(path, ".ma2");
  FILE *fp = std::fopen(path, "rb");
  NK_CHECK(fp, NKH_ERR_ARG, "map: cannot open %s", path);
  if (m.binary) {
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    m.data.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = m.data.empty() ? 0 : std::fread(m.data.data(), 1, m.data.size(), fp);
    std::fclose(fp);
    NK_CHECK(got == m.data.size() && m.data.size() >= 136, NKH_ERR_ARG,
             "ma2: %s is shorter than its header", path);
    char ver[6] = {0}, a[13] = {0}, b[13] = {0};
    std::memcpy(ver, m.data.data(), 5);
    std::memcpy(a, m.data.data() + 5, 12);
    std::memcpy(b, m.data.data() + 17, 12);
    char *end = nullptr;
    m.nel = (int)std::strtol(a, &end, 10);
    NK_CHECK(end != a && m.nel > 0, NKH_ERR_ARG, "ma2: bad header '%.29s'",
             (const char *)m.data.data());
    m.nactive = (int)std::strtol(b, nullptr, 10);
    float t;
    std::memcpy(&t, m.data.data() + 132, 4);
    if (std::fabs(t - 6.54321f) > 0.0002f) {
      unsigned char *q = m.data.data() + 132;
      std::reverse(q, q + 4);
      std::memcpy(&t, q, 4);
      std::reverse(q, q + 4);
      NK_CHECK(std::fabs(t - 6.54321f) <= 0.0002f, NKH_ERR_ARG,
               "ma2: byte-order word is neither order of 6.54321");
      m.swap = true;
    }
    return NKH_OK;
  }
  std::string text;
  char buf[65536];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, k);
  std::fclose(fp);
  const size_t eol = text.find('\n');
  NK_CHECK(eol != std::string::npos, NKH_ERR_ARG, "map: %s has no header line", path);
  long long h0 = 0, h1 = 0;
  NK_CHECK(std::sscanf(text.c_str(), "%lld %lld", &h0, &h1) == 2 && h0 > 0, NKH_ERR_ARG,
           "map: bad header line in %s", path);
  m.nel = (int)h0;
  m.nactive = (int)h1;
  const char *s = text.c_str() + eol + 1;
  while (*s) {
    char *end = nullptr;
    const long long v = std::strtoll(s, &end, 10);
    if (end == s) {
      NK_CHECK(std::isspace((unsigned char)*s), NKH_ERR_ARG, "map: stray '%c' in %s", *s, path);
      ++s;
      continue;
    }
    m.words.push_back(v);
    s = end;
  }
  return NKH_OK;
}

}  // namespace

int map_header(const char *path, int *nel, int *nactive) {
  MapFile m;
  NK_TRY(map_load(path, m));
  if (nel) *nel = m.nel;
  if (nactive) *nactive = m.nactive;
  return NKH_OK;
}

int map_read(const char *path, int nel, int ndim, int *part, int64_t *vertex) {
  NK_CHECK(ndim == 2 || ndim == 3, NKH_ERR_ARG, "map: ndim %d", ndim);
  MapFile m;
  NK_TRY(map_load(path, m));
  NK_CHECK(nel == m.nel, NKH_ERR_ARG,
           "map: %d elements requested, the file has %d (the .map and .re2 must agree)", nel,
           m.nel);
  const int nv = 1 << ndim, rec = 1 + nv;
  if (m.binary) {
    const size_t need = 136 + (size_t)nel * rec * 4;
    NK_CHECK(m.data.size() == need, NKH_ERR_ARG, "ma2: %zu bytes, %zu expected for %d elements",
             m.data.size(), need, nel);
    for (int e = 0; e < nel; ++e)
      for (int c = 0; c < rec; ++c) {
        unsigned char w[4];
        std::memcpy(w, m.data.data() + 136 + ((size_t)e * rec + c) * 4, 4);
        if (m.swap) std::reverse(w, w + 4);
        int32_t v;
        std::memcpy(&v, w, 4);
        if (c == 0) {
          if (part) part[e] = v;
        } else if (vertex) {
          vertex[(size_t)nv * e + (c - 1)] = v;
        }
      }
  } else {
    NK_CHECK(m.words.size() == (size_t)nel * rec, NKH_ERR_ARG,
             "map: %zu numbers after the header, %zu expected for %d elements", m.words.size(),
             (size_t)nel * rec, nel);
    for (int e = 0; e < nel; ++e) {
      if (part) part[e] = (int)m.words[(size_t)e * rec];
      for (int c = 0; c < nv && vertex; ++c)
        vertex[(size_t)nv * e + c] = m.words[(size_t)e * rec + 1 + c];
    }
  }
  return NKH_OK;
}

namespace {

// preprocessor corner c (0-based) of lexicographic corner ix: the
// reference's indx = (1,2,4,3,5,6,8,7) (genxyz.f:1289-1291)
constexpr int kPrepOfLex[8] = {0, 1, 3, 2, 4, 5, 7, 6};
// lexicographic corners of each preprocessor face (faces 1..6: s-, r+, s+,
// r-, t-, t+)
constexpr int kFaceCorners[6][4] = {{0, 1, 4, 5}, {1, 3, 5, 7}, {2, 3, 6, 7},
                                    {0, 2, 4, 6}, {0, 1, 2, 3}, {4, 5, 6, 7}};

struct DSU {
  std::vector<int> p;
  explicit DSU(int n) : p((size_t)n) { std::iota(p.begin(), p.end(), 0); }
  int find(int a) {
    while (p[(size_t)a] != a) a = p[(size_t)a] = p[(size_t)p[(size_t)a]];
    return a;
  }
  void unite(int a, int b) {
    a = find(a);
    b = find(b);
    if (a != b) p[(size_t)std::max(a, b)] = std::min(a, b);
  }
};

}  // namespace

// Vertex ids by corner coincidence.  Corners closer than tol_rel times the
// shortest element edge are one vertex; periodic faces ('P  ' with bc(1) =
// the connected element, bc(2) = its face) join their corners after the
// translation between the two faces.  Ids are 1.. in order of first
// appearance; vertex is written in lexicographic corner order (corner
// (i,j,k) at i + 2j + 4k), the layout setvert3d reads.
int mesh_vertex_ids(int nel, int ndim, const double *xc, const double *yc, const double *zc,
                    const char *cbc, const double *bc, double tol_rel, int64_t *vertex,
                    int64_t *nvert) {
  NK_CHECK(ndim == 3, NKH_ERR_UNSUPPORTED, "vertex ids: 3-D meshes only");
  NK_CHECK(nel > 0 && xc && yc && zc && vertex, NKH_ERR_ARG, "vertex ids: bad arguments");
  const int nv = 8, np = nv * nel;
  auto P = [&](int q, int d) {
    const double *a = d == 0 ? xc : d == 1 ? yc : zc;
    const int e = q / nv, c = q % nv;
    return a[(size_t)nv * e + kPrepOfLex[c]];  // q indexes lexicographic corners
  };
  // shortest edge
  double hmin = INFINITY;
  for (int e = 0; e < nel; ++e)
    for (int a = 0; a < 8; ++a)
      for (int d = 0; d < 3; ++d) {
        const int b = a ^ (1 << d);
        if (b < a) continue;
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
          const double t = P(e * nv + a, k) - P(e * nv + b, k);
          s += t * t;
        }
        hmin = std::min(hmin, std::sqrt(s));
      }
  NK_CHECK(hmin > 0.0 && std::isfinite(hmin), NKH_ERR_ARG, "vertex ids: degenerate element");
  const double tol = tol_rel * hmin;
  DSU dsu(np);
  std::vector<int> ord((size_t)np);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return P(a, 0) < P(b, 0); });
  for (size_t s = 0; s < ord.size(); ++s) {
    const int a = ord[s];
    for (size_t t = s + 1; t < ord.size() && P(ord[t], 0) - P(a, 0) <= tol; ++t) {
      const int b = ord[t];
      if (std::fabs(P(a, 1) - P(b, 1)) <= tol && std::fabs(P(a, 2) - P(b, 2)) <= tol)
        dsu.unite(a, b);
    }
  }
  if (cbc && bc) {
    for (int e = 0; e < nel; ++e)
      for (int fc = 0; fc < 6; ++fc) {
        const char *c = cbc + 3 * ((size_t)6 * e + fc);
        if (!(c[0] == 'P' && c[1] == ' ' && c[2] == ' ')) continue;
        const int e2 = (int)bc[((size_t)6 * e + fc) * 5] - 1;
        const int f2 = (int)bc[((size_t)6 * e + fc) * 5 + 1] - 1;
        NK_CHECK(e2 >= 0 && e2 < nel && f2 >= 0 && f2 < 6, NKH_ERR_ARG,
                 "vertex ids: periodic face %d of element %d points at element %d face %d",
                 fc + 1, e + 1, e2 + 1, f2 + 1);
        double c1[3] = {0, 0, 0}, c2[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k)
          for (int d = 0; d < 3; ++d) {
            c1[d] += 0.25 * P(e * nv + kFaceCorners[fc][k], d);
            c2[d] += 0.25 * P(e2 * nv + kFaceCorners[f2][k], d);
          }
        for (int k = 0; k < 4; ++k) {
          const int a = e * nv + kFaceCorners[fc][k];
          int best = -1;
          double bd = INFINITY;
          for (int k2 = 0; k2 < 4; ++k2) {
            const int b = e2 * nv + kFaceCorners[f2][k2];
            double s = 0.0;
            for (int d = 0; d < 3; ++d) {
              const double t = (P(a, d) - c1[d]) - (P(b, d) - c2[d]);
              s += t * t;
            }
            if (s < bd) {
              bd = s;
              best = b;
            }
          }
          NK_CHECK(std::sqrt(bd) <= tol, NKH_ERR_ARG,
                   "vertex ids: periodic faces of elements %d and %d do not match", e + 1,
                   e2 + 1);
          dsu.unite(a, best);
        }
      }
  }
  std::vector<int64_t> id((size_t)np, 0);
  int64_t next = 0;
  for (int q = 0; q < np; ++q) {
    const int r = dsu.find(q);
    if (id[(size_t)r] == 0) id[(size_t)r] = ++next;
    vertex[q] = id[(size_t)r];
  }
  if (nvert) *nvert = next;
  return NKH_OK;
}

namespace {

// 1-D Lagrange weights of the nodes zn at the point x (degree m-1)
void lagrange(double x, const double *zn, int m, double *w) {
  for (int j = 0; j < m; ++j) {
    double p = 1.0;
    for (int k = 0; k < m; ++k)
      if (k != j) p *= (x - zn[k]) / (zn[j] - zn[k]);
    w[j] = p;
  }
}

// navier5.f:2422-2560 gh_face_extend_3d with gh_type 2 (vertices and
// edges) on a 3x3x3 array, zg = (-1, 0, 1)
void gh_extend_edges(double *x) {
  const int n = 3;
  const double zg[3] = {-1.0, 0.0, 1.0};
  auto X = [&](int i, int j, int k) -> double & { return x[i + n * (j + n * k)]; };
  auto h = [&](int ii, int i) {  // linear blend toward end ii (0 or n-1)
    return 0.5 * ((n - 1 - ii) * (1 - zg[i]) + ii * (1 + zg[i])) / (n - 1);
  };
  double v[27] = {0}, e[27] = {0};
  for (int kk = 0; kk < n; kk += n - 1)
    for (int jj = 0; jj < n; jj += n - 1)
      for (int ii = 0; ii < n; ii += n - 1)
        for (int k = 0; k < n; ++k)
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
              v[i + n * (j + n * k)] += h(ii, i) * h(jj, j) * h(kk, k) * X(ii, jj, kk);
  auto V = [&](int i, int j, int k) { return v[i + n * (j + n * k)]; };
  for (int kk = 0; kk < n; kk += n - 1)
    for (int jj = 0; jj < n; jj += n - 1)
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i)
            e[i + n * (j + n * k)] += h(jj, j) * h(kk, k) * (X(i, jj, kk) - V(i, jj, kk));
  for (int kk = 0; kk < n; kk += n - 1)
    for (int ii = 0; ii < n; ii += n - 1)
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i)
            e[i + n * (j + n * k)] += h(ii, i) * h(kk, k) * (X(ii, j, kk) - V(ii, j, kk));
  for (int jj = 0; jj < n; jj += n - 1)
    for (int ii = 0; ii < n; ii += n - 1)
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i)
            e[i + n * (j + n * k)] += h(ii, i) * h(jj, j) * (X(ii, jj, k) - V(ii, jj, k));
  for (int q = 0; q < 27; ++q) x[q] = e[q] + v[q];
}

// edge (preprocessor numbering 1..12) -> node of the 3x3x3 array
// (genxyz.f:1349-1353 eindx), 0-based
constexpr int kEdgeNode[12] = {1, 5, 7, 3, 19, 23, 25, 21, 9, 11, 17, 15};

// preprocessor face -> symmetric face (connect1.f:200-205 eface1), 1-based
constexpr int kEface1[6] = {3, 2, 4, 1, 5, 6};

double dot3(const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i) s += a[i] * b[i];
  return s;
}
void cross3(double *v1, const double *v2, const double *v3) {
  v1[0] = v2[1] * v3[2] - v2[2] * v3[1];
  v1[1] = v2[2] * v3[0] - v2[0] * v3[2];
  v1[2] = v2[0] * v3[1] - v2[1] * v3[0];
}
void norm3(double *v) {
  const double l = std::sqrt(dot3(v, v));
  if (l > 0) {
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
  }
}

// genxyz.f:846-882 edg3d: n points on the great circle from x1 towards x2
// (unit sphere), at the GLL angles; p(q) = point q, stride st in pts.
void edg3d(double *pts, int st, int n, const double *x1, const double *x2, const double *z) {
  double u1[3] = {x1[0], x1[1], x1[2]}, u2[3] = {x2[0], x2[1], x2[2]}, vn[3], b[3];
  norm3(u1);
  norm3(u2);
  cross3(vn, x1, x2);
  cross3(b, vn, x1);
  norm3(vn);
  norm3(b);
  const double theta = std::acos(dot3(u1, u2));
  for (int q = 0; q < n; ++q) {
    const double tp = 0.5 * theta * (z[q] + 1.0);
    const double c = std::cos(tp), s = std::sin(tp);
    for (int d = 0; d < 3; ++d) pts[(size_t)st * q + d] = c * u1[d] + s * b[d];
  }
}

// genxyz.f:677-845 sphsrf with crn3d (:934-978): face ifce (preprocessor,
// 0-based) of element e moved onto the sphere curve[0..2] (centre),
// curve[3] (radius); the deflection of each face point is blended linearly
// to zero at the opposite face.  X[d] = the element's coordinates (n^3,
// i fastest); xc/yc/zc its preprocessor corners.
int sphsrf(double *X[3], int n, const double *z, const double *xc, const double *yc,
           const double *zc, const double *curve, int ifce, int e) {
  static const int indvtx[6][4] = {{1, 5, 3, 7}, {2, 4, 6, 8}, {1, 2, 5, 6},
                                   {3, 7, 4, 8}, {1, 3, 2, 4}, {5, 6, 7, 8}};
  static const int indx[8] = {1, 2, 4, 3, 5, 6, 8, 7};
  static const int cface[6][2] = {{1, 4}, {2, 1}, {3, 2}, {4, 3}, {1, 5}, {5, 1}};
  const double ctr[3] = {curve[0], curve[1], curve[2]}, radius = curve[3];
  const int iface = kEface1[ifce];
  NK_CHECK(radius > 0.0, NKH_ERR_ARG, "sphere of radius %g on element %d face %d", radius,
           e + 1, ifce + 1);
  double xcv[4][3];
  for (int i = 0; i < 4; ++i) {
    const int k = indx[indvtx[iface - 1][i] - 1] - 1;
    xcv[i][0] = xc[k] - ctr[0];
    xcv[i][1] = yc[k] - ctr[1];
    xcv[i][2] = zc[k] - ctr[2];
    const double radt = std::sqrt(dot3(xcv[i], xcv[i]));
    NK_CHECK(std::fabs(radt - radius) / radius <= 1.0e-4, NKH_ERR_ARG,
             "element %d face %d: vertex not on the requested sphere", e + 1, ifce + 1);
  }
  // points (a, b) of the face on the unit sphere, a along corner 1 -> 2,
  // b along corner 1 -> 3
  std::vector<double> p((size_t)3 * n * n);
  auto P = [&](int a, int b) { return &p[(size_t)3 * (a + (size_t)n * b)]; };
  edg3d(P(0, 0), 3 * n, n, xcv[0], xcv[2], z);
  edg3d(P(n - 1, 0), 3 * n, n, xcv[1], xcv[3], z);
  edg3d(P(0, 0), 3, n, xcv[0], xcv[1], z);
  edg3d(P(0, n - 1), 3, n, xcv[2], xcv[3], z);
  const int iv = cface[ifce][0] - 1, io = cface[ifce][1] - 1;
  const double vout[3] = {xc[iv] - xc[io], yc[iv] - yc[io], zc[iv] - zc[io]};
  const double vsph[3] = {xc[iv] - ctr[0], yc[iv] - ctr[1], zc[iv] - ctr[2]};
  const bool concave = !(dot3(vsph, vout) > 0);
  for (int b = 1; b < n - 1; ++b) {
    double vn1[3], vn2[3];
    cross3(vn1, P(0, b), P(n - 1, b));
    for (int a = 1; a < n - 1; ++a) {
      cross3(vn2, P(a, 0), P(a, n - 1));
      if (concave) cross3(P(a, b), vn2, vn1);
      else cross3(P(a, b), vn1, vn2);
    }
  }
  for (int q = 0; q < n * n; ++q) {
    double *v = &p[(size_t)3 * q];
    norm3(v);
    for (int d = 0; d < 3; ++d) v[d] = v[d] * radius;
    for (int d = 0; d < 3; ++d) v[d] = v[d] + ctr[d];
  }
  if (iface == 1 || iface == 4 || iface == 5)
    for (int b = 0; b < n; ++b)
      for (int a = b + 1; a < n; ++a)
        for (int d = 0; d < 3; ++d) std::swap(P(a, b)[d], P(b, a)[d]);
  // connect1.f:586-629 skpdat of the symmetric face: the face nodes as
  // (j1, j2) of the array viewed as (n*n, n), 1-based
  const int nx = n, ny = n, nz = n;
  int sk[6];
  switch (iface) {
    case 1: sk[0] = 1; sk[1] = nx * (ny - 1) + 1; sk[2] = nx; sk[3] = 1;
            sk[4] = ny * (nz - 1) + 1; sk[5] = ny; break;
    case 2: sk[0] = nx; sk[1] = nx * ny; sk[2] = nx; sk[3] = 1;
            sk[4] = ny * (nz - 1) + 1; sk[5] = ny; break;
    case 3: sk[0] = 1; sk[1] = nx; sk[2] = 1; sk[3] = 1; sk[4] = ny * (nz - 1) + 1;
            sk[5] = ny; break;
    case 4: sk[0] = 1 + nx * (ny - 1); sk[1] = nx * ny; sk[2] = 1; sk[3] = 1;
            sk[4] = ny * (nz - 1) + 1; sk[5] = ny; break;
    case 5: sk[0] = 1; sk[1] = nx; sk[2] = 1; sk[3] = 1; sk[4] = ny; sk[5] = 1; break;
    default: sk[0] = 1 + nx * ny * (nz - 1); sk[1] = nx + nx * ny * (nz - 1); sk[2] = 1;
             sk[3] = 1; sk[4] = ny; sk[5] = 1; break;
  }
  const int iopp[3] = {nx - 1, nx * (ny - 1), nx * ny * (nz - 1)};
  const int idir = 2 * (iface % 2) - 1, ifc2 = (iface + 1) / 2;
  int q = 0;
  for (int j2 = sk[3]; j2 <= sk[4]; j2 += sk[5])
    for (int j1 = sk[0]; j1 <= sk[1]; j1 += sk[2]) {
      const double *xs = &p[(size_t)3 * q++];
      const int jopp = j1 + iopp[ifc2 - 1] * idir;
      const size_t node = (size_t)(j1 - 1) + (size_t)nx * (j2 - 1);
      const double dx[3] = {xs[0] - X[0][node], xs[1] - X[1][node], xs[2] - X[2][node]};
      const int joff = (j1 - jopp) / (n - 1);
      for (int ix = 2; ix <= n; ++ix) {
        const int j = jopp + joff * (ix - 1);
        const double zeta = 0.5 * (z[ix - 1] + 1.0);
        const size_t o = (size_t)(j - 1) + (size_t)nx * (j2 - 1);
        for (int d = 0; d < 3; ++d) X[d][o] = X[d][o] + dx[d] * zeta;
      }
    }
  return NKH_OK;
}

// genxyz.f:2-103 arcsrf in 3-D: edge isid (1..8, preprocessor; the edges of
// the faces t = -1, +1) bent into the circular arc of radius curve[0] in
// its x-y plane (centre on the side given by the sign of the radius); the
// arc's deviation from the chord is blended into the element with the
// linear weights of the two faces the edge bounds.
int arcsrf(double *X[3], int n, const double *z, const double *xc, const double *yc,
           double radius, int isid, int e) {
  const int i1 = isid - 1;
  const int i2 = isid == 4 ? 0 : isid == 8 ? 4 : isid;
  const double p1x = xc[i1], p1y = yc[i1], p2x = xc[i2], p2y = yc[i2];
  const double gap = std::sqrt((p1x - p2x) * (p1x - p2x) + (p1y - p2y) * (p1y - p2y));
  NK_CHECK(!(std::fabs(2.0 * radius) <= gap * 1.00001), NKH_ERR_ARG,
           "too small a radius (%g) on side %d of element %d: chord %g", radius, isid, e + 1,
           gap);
  const double xs = p2y - p1y, ys = p1x - p2x;
  const double xys = std::sqrt(xs * xs + ys * ys);
  const double dtheta = std::fabs(std::asin(0.5 * gap / radius));
  const double p12x = (p1x + p2x) / 2.0, p12y = (p1y + p2y) / 2.0;
  const double xcen = p12x - xs / xys * radius * std::cos(dtheta);
  const double ycen = p12y - ys / xys * radius * std::cos(dtheta);
  const double theta0 = std::atan2(p12y - ycen, p12x - xcen);
  const int s1 = (isid - 1) % 4 + 1;
  std::vector<double> xcrv((size_t)n), ycrv((size_t)n), h1((size_t)n), h2((size_t)n);
  for (int ix = 0; ix < n; ++ix) {
    h1[(size_t)ix] = (1.0 - z[ix]) * 0.5;
    h2[(size_t)ix] = (1.0 + z[ix]) * 0.5;
  }
  for (int ix = 0; ix < n; ++ix) {
    const int ixt = s1 > 2 ? n - 1 - ix : ix;
    double r = z[ix];
    if (radius < 0.0) r = -r;
    xcrv[(size_t)ixt] = xcen + std::fabs(radius) * std::cos(theta0 + r * dtheta) -
                        (h1[(size_t)ix] * p1x + h2[(size_t)ix] * p2x);
    ycrv[(size_t)ixt] = ycen + std::fabs(radius) * std::sin(theta0 + r * dtheta) -
                        (h1[(size_t)ix] * p1y + h2[(size_t)ix] * p2y);
  }
  // math.f:476-492 addtnsr with the blending weights of the symmetric face
  // the edge lies on and of the t face
  const int sf = kEface1[s1 - 1];
  const double *ht = (isid - 1) / 4 == 0 ? h1.data() : h2.data();
  for (int d = 0; d < 2; ++d) {
    const double *crv = d == 0 ? xcrv.data() : ycrv.data();
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) {
          double f1, f2;
          if (sf <= 2) {  // varies along s, weight in r
            f1 = (sf == 1 ? h1 : h2)[(size_t)i];
            f2 = crv[j];
          } else {  // varies along r, weight in s
            f1 = crv[i];
            f2 = (sf == 3 ? h1 : h2)[(size_t)j];
          }
          const double hh = f2 * ht[k];
          X[d][(size_t)i + (size_t)n * (j + (size_t)n * k)] += hh * f1;
        }
      }
  }
  return NKH_OK;
}

}  // namespace

// genxyz.f:562-605 genxyz in 3-D: GLL coordinates from the corners by
// linquad (:1235-1400 xyzlin / xyzquad, 'm' midside-node edges), then per
// element the 's' (sphere) faces in face order and the 'C' (circular arc)
// edges 1..8 in edge order.  'e' and other curve types are not built.
int mesh_xyz_from_corners(int lx1, int nel, const double *xc, const double *yc,
                          const double *zc, const char *ccurve, const double *curve,
                          double *xm1, double *ym1, double *zm1) {
  NK_CHECK(lx1 >= 2 && lx1 <= kMaxNP && nel > 0 && xc && yc && zc && xm1 && ym1 && zm1,
           NKH_ERR_ARG, "xyz from corners: bad arguments");
  std::vector<double> z((size_t)lx1), w((size_t)lx1);
  NK_TRY(nkh_zwgll(z.data(), w.data(), lx1));
  const double zq[3] = {-1.0, 0.0, 1.0}, zl[2] = {-1.0, 1.0};
  std::vector<double> jq((size_t)lx1 * 3), jl((size_t)lx1 * 2);
  for (int i = 0; i < lx1; ++i) {
    lagrange(z[(size_t)i], zq, 3, &jq[(size_t)3 * i]);
    lagrange(z[(size_t)i], zl, 2, &jl[(size_t)2 * i]);
  }
  const int n3 = lx1 * lx1 * lx1;
  for (int e = 0; e < nel; ++e) {
    bool quad = false;
    for (int k = 0; k < 12 && ccurve; ++k) {
      const char c = ccurve[(size_t)12 * e + k];
      if (c == 'm') quad = true;
      else
        NK_CHECK(c == ' ' || c == '\0' || (c == 'C' && k < 8) || (c == 's' && k < 6) ||
                     (c == 'm'),
                 NKH_ERR_UNSUPPORTED, "xyz from corners: curve type '%c' on element %d edge %d",
                 c, e + 1, k + 1);
    }
    if (lx1 == 2) quad = false;  // genxyz.f:1258
    const double *src[3] = {xc + (size_t)8 * e, yc + (size_t)8 * e, zc + (size_t)8 * e};
    double *dst[3] = {xm1 + (size_t)n3 * e, ym1 + (size_t)n3 * e, zm1 + (size_t)n3 * e};
    for (int d = 0; d < 3; ++d) {
      double cb[8];
      for (int c = 0; c < 8; ++c) cb[c] = src[d][kPrepOfLex[c]];
      const int m = quad ? 3 : 2;
      double q[27];
      if (quad) {
        // xyzlin onto the 3x3x3 nodes, then the midside nodes, then the
        // Gordon-Hall extension of the edges (genxyz.f:1356-1380)
        for (int k = 0; k < 3; ++k)
          for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
              double lw[3][2];
              for (int t = 0; t < 3; ++t) lagrange(zq[t == 0 ? i : t == 1 ? j : k], zl, 2, lw[t]);
              double s = 0.0;
              for (int c = 0; c < 8; ++c)
                s += lw[0][c & 1] * lw[1][(c >> 1) & 1] * lw[2][c >> 2] * cb[c];
              q[i + 3 * (j + 3 * k)] = s;
            }
        for (int k = 0; k < 12; ++k)
          if (ccurve[(size_t)12 * e + k] == 'm')
            q[kEdgeNode[k]] = curve[((size_t)12 * e + k) * 5 + d];
        gh_extend_edges(q);
      } else {
        for (int c = 0; c < 8; ++c) q[c] = cb[c];
      }
      const double *J = quad ? jq.data() : jl.data();
      for (int k = 0; k < lx1; ++k)
        for (int j = 0; j < lx1; ++j)
          for (int i = 0; i < lx1; ++i) {
            double s = 0.0;
            for (int c = 0; c < m; ++c)
              for (int b = 0; b < m; ++b)
                for (int a = 0; a < m; ++a)
                  s += J[(size_t)m * i + a] * J[(size_t)m * j + b] * J[(size_t)m * k + c] *
                       q[a + m * (b + m * c)];
            dst[d][i + lx1 * (j + lx1 * k)] = s;
          }
    }
    if (!ccurve || lx1 < 3) continue;
    const double *xe = xc + (size_t)8 * e, *ye = yc + (size_t)8 * e, *ze = zc + (size_t)8 * e;
    for (int f = 0; f < 6; ++f)
      if (ccurve[(size_t)12 * e + f] == 's')
        NK_TRY(sphsrf(dst, lx1, z.data(), xe, ye, ze, &curve[((size_t)12 * e + f) * 5], f, e));
    for (int k = 0; k < 8; ++k)
      if (ccurve[(size_t)12 * e + k] == 'C')
        NK_TRY(arcsrf(dst, lx1, z.data(), xe, ye, curve[((size_t)12 * e + k) * 5], k + 1, e));
  }
  return NKH_OK;
}

}  // namespace nekamd
