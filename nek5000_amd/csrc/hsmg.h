// Internal declarations of the H1 multigrid (hsmg_setup.hip builds the
// hierarchy, hsmg.hip applies it and runs the flexible CG).
#pragma once
#include <vector>

#include "coarse.h"
#include "nekamd_internal.h"

namespace nekamd {

// Scalars of one hmh_flex_cg solve, resident on the device.
struct FcgState {
  double rho0, rho1, rho2, beta, den, alpha, rnorm, div0, tolpss;
  int iter, done;
};

// One multigrid level.  Fields are (nh,nh,nh,nel) element-major.  On the
// Schwarz levels (l >= 2) the overlapping subdomain of an element is its
// (nh+2)^3 extension; its ghost faces are exchanged through two gather maps
// built from the extended numbering (hsmg.f:2361-2394):
//   gin[(e*6+f)*nh^2 + a + nh*b]  = index (in a level field) of the node of
//        the face neighbour that lies one layer inside it, behind ghost node
//        (f,a,b) of element e; -1 where the face has no neighbour;
//   gout[...] = the neighbour's own ghost slot that coincides with (f,a,b).
// Faces f in symmetric order r-, r+, s-, s+, t-, t+; (a,b) run over the two
// other directions, lower direction first.
struct MgLevel {
  int nx = 0, nh = 0;
  int64_t n = 0;  // nh^3 * nel
  std::vector<double> ah, bh, zh;
  double *J = nullptr;   // interpolation to level l+1: (nh[l+1] x nh[l]) col-major
  double *Jt = nullptr;  // its transpose
  GsMap gs;
  bool own_gs = false;
  double *mask = nullptr;  // 0/1
  double *rwt = nullptr;   // restriction weights
  double *mswt = nullptr;  // mask * Schwarz weight (l >= 2)
  double *fs = nullptr;    // per element S_r, S_s, S_t ((nh+2)^2 each)
  double *fl = nullptr;    // per element lam_r, lam_s, lam_t (nh+2 each), eps
  int32_t *gin = nullptr, *gout = nullptr;
  double *wint = nullptr;  // local solutions: interior (nh^3 per element)
  // local solutions of the neighbours at this element's ghost slots (6 nh^2
  // per element; 0 where a face has no neighbour)
  double *wgh = nullptr;
  double *e = nullptr, *r = nullptr;
  // single-precision copies for the mixed-precision smoother (fine level)
  float *fs32 = nullptr, *fl32 = nullptr, *wint32 = nullptr, *wgh32 = nullptr, *e32 = nullptr;
  // fine level: mask*Schwarz weights as byte codes into Mg::wlut, and
  // restriction weights as counts (CountCode: weight 1/count), when every
  // value has one
  uint8_t *mswt8 = nullptr, *rwt8 = nullptr;
  float *mswt32 = nullptr;  // mask * Schwarz weight, single precision
  float *rwt32 = nullptr;   // restriction weights, single precision
};

struct Mg {
  int lmax = 0;
  int64_t nel = 0;
  MgLevel lev[4];
  int *fbc = nullptr;  // device, 6 per element
  hipStream_t sb = nullptr;  // second stream: coarse levels beside the fine one
  hipStream_t sc = nullptr;  // third: intermediate-level Schwarz beside the coarse solve
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_lr[4] = {}, ev_le[4] = {};  // per level: r ready, e ready
  // timing events (nkh_timers): fine local solves start/end, V-cycle start/end
  hipEvent_t tev[4] = {};
  // coarse problem (navier8.f:83-233): unknowns = free vertex ids
  int ncrs = 0;
  // no Dirichlet face: the constant is in the null space of the pressure
  // operator (the reference's ifvcor); the coarse unknown ncrs-1 is held at
  // zero and the coarse solution's mean removed (crs_xxt.c:883-955), and the
  // Krylov solvers apply ortho (navier1.f:223-255)
  bool null_space = false;
  double *opart = nullptr;  // ortho partial sums
  int32_t *slot = nullptr;                   // per level-1 node: coarse unknown or -1
  int32_t *goff = nullptr, *gidx = nullptr;  // unknown -> level-1 nodes (CSR)
  double *cb = nullptr, *cx = nullptr;       // coarse rhs and solution
  CoarseSolver crs;
  // flexible CG work (4 fine fields) and scalars
  double *fr = nullptr, *fp = nullptr, *fz = nullptr, *fw = nullptr;
  double *fpart = nullptr;
  FcgState *fst = nullptr;
  int *fflag = nullptr;
  // the multiplicity weights as counts (wt = 1/count, count <= 255), plus a flag
  uint8_t *wt8 = nullptr;
  // the 256 weights the fine level's byte codes index, then the same 256
  // rounded to FP32
  double *wlut = nullptr;
  // GMRES work: V (kGmresM+1 fields), Z (kGmresM fields), x, r, w, partials
  double *gV = nullptr, *gZ = nullptr, *gx = nullptr, *gr = nullptr, *gw = nullptr;
  double *gpart = nullptr, *gsm = nullptr;
  // mixed-precision solver: single-precision inner vectors, partial sums
  // and scalars (the FP32 geometric factors belong to the mesh)
  float *xr = nullptr, *xp = nullptr, *xz = nullptr, *xw = nullptr, *xx = nullptr,
        *xwt = nullptr;
  double *xpart = nullptr;
  void *xst = nullptr;
};

constexpr int kGmresM = 30;  // SIZE lgmres: Krylov dimension before a restart

int mg_grid(int64_t n);
// allocate and fill the single-precision smoother data of the fine level
int mg_prepare_fp32(Mg &mg, hipStream_t s, bool inner);
// navier1.f:223-255 ORTHO: v <- v - sum(v)/n over the n local nodes (device)
int mg_ortho(Mg &mg, double *v, int64_t n, hipStream_t s);
// the multiplicity weights wt as count bytes (wt = 1/count), into mg.wt8;
// *codes = nullptr when some weight is not 1/count for a count in 1..255
int mg_wt_codes(Mg &mg, const double *wt, int64_t n, hipStream_t s, const uint8_t **codes);
// the constant null space is present for this solve (no Dirichlet face, h2 = 0)
inline bool mg_null_space_on(const Mg &mg, Coef h2) {
  return mg.null_space && h2.uniform() && h2.s == 0.0;
}
int mg_dalloc(double **p, int64_t n);

// Schwarz solve on level l of r (masked): writes the level's e (not yet
// summed) from the local FDM solutions; wint/wgh as scratch.
int mg_schwarz_local(Mg &mg, int l, const double *r, double *e, hipStream_t s);
// The overlap sum alone: e = interior of wint + the neighbours' solutions
// held in wgh, on the first interior layer.
int mg_overlap_out(Mg &mg, int l, double *e, hipStream_t s);

}  // namespace nekamd
