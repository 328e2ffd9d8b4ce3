// lepic_mg.hip -- V-cycle multigrid with red-black SOR smoothing (a8):
// pdesolver (modules/legacy/pdesolver.f90:1-135), mg (mg.f90:1-246),
// restriction (mg.f90:249-806), getres (mg.f90:809-848), prolongation
// (mg.f90:850-940) and sor_rb (modules/legacy/sors.f90:1-535), one rank.
//
// Each smoother sweep is one launch per colour over that colour's nodes
// (x fastest), followed by a deterministic block-tree L1 residual and a
// one-block kernel that applies sor_rb's exit rules on the device (res <= eps
// on the coarsest level, res/res_s <= 0.6 elsewhere, at most 4 sweeps) and
// clears the level's active flag.  Sweeps queued after the exit return at once, so
// a V-cycle needs no host round trip; the host reads the fine residual once
// per V-cycle for pdesolver's convergence test.  Levels below a few thousand
// nodes run inside one workgroup (k_mg_coarse).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lepic_internal.h"

namespace {

constexpr int BLOCK = 256;
constexpr int MAX_LEVELS = 12;

struct MGCtl {
  int64_t active;
  int64_t ksor;
  double res;
  double res_s;
  int64_t total;
  int64_t pad[11];
};
static_assert(sizeof(MGCtl) == 16 * sizeof(int64_t), "MGCtl layout");

// One multigrid level: geometry, stencil coefficients and the level's u, b
// and labels.  T is the storage and arithmetic type of u and b (double for
// the reference's solver, float for the smoother of the mixed-precision one).
// floor(x / d) for x < 2^32 by a 64-bit multiply-high: m = ceil(2^64 / d),
// and x m / 2^64 exceeds x / d by less than x / 2^64 < 2^-32 <= 1/d, so the
// floor is exact.  Decodes thread numbers into (i, j, k) without a division.
struct MDiv {
  uint64_t m;
  uint32_t d;
};

inline MDiv mdiv(uint32_t d) {
  MDiv v;
  v.d = d;
  v.m = d <= 1 ? 0 : (~0ull) / d + 1;
  return v;
}

__device__ inline uint32_t mdiv_q(uint32_t x, const MDiv &v) {
  return v.d <= 1 ? x : (uint32_t)__umul64hi((uint64_t)x, v.m);
}

// B: the label type (int32_t in global memory; int16_t in the coarse
// workgroup's LDS, labels saturated to that range, which keeps every test
// the solver makes of them: sign, 0, -1, -2).
template <class T, class B = int32_t>
struct KLevelT {
  KGrid g;
  T aw, ae, an, as, at, ab, ac;
  T *u;
  T *b;
  B *bc;
  // thread-number decoders: n1/2+1, n2+1, n1+1, n1+3, n1+2, n2/2+1
  MDiv d_h1, d_n2p1, d_n1p1, d_n1p3, d_n1p2, d_h2;
  double acy;   // RN(1 / ac) for the FP64 levels' division (0: plain division)
};

// x / ac of the SOR update: for FP64 levels by the reciprocal and two FMA
// corrections (lepic_qdiv.h: the correctly rounded quotient, identical to
// the division), the plain division otherwise.
template <class T, class B>
__device__ inline T div_ac(T x, const KLevelT<T, B> &L) {
  return x / L.ac;
}
template <class B>
__device__ inline double div_ac(double x, const KLevelT<double, B> &L) {
  if (L.acy == 0.0) return x / L.ac;
  QDiv d;
  d.h = L.ac;
  d.y = L.acy;
  return qdiv(x, d);
}

template <class T>
inline void set_level_divs(KLevelT<T> &L) {
  const KGrid &g = L.g;
  L.d_h1 = mdiv((uint32_t)(g.n1 / 2 + 1));
  L.d_n2p1 = mdiv((uint32_t)(g.n2 + 1));
  L.d_n1p1 = mdiv((uint32_t)(g.n1 + 1));
  L.d_n1p3 = mdiv((uint32_t)(g.n1 + 3));
  L.d_n1p2 = mdiv((uint32_t)(g.n1 + 2));
  L.d_h2 = mdiv((uint32_t)(g.n2 / 2 + 1));
}
using KLevel = KLevelT<double>;

template <class T, class B>
__device__ inline T getres(const KLevelT<T, B> &L, int ir, int jr, int kr, int flag_nmn) {
  const KGrid &g = L.g;
  if (ir < 0 || ir > g.n1 + 1 || jr < 0 || jr > g.n2 + 1 || kr < 0 || kr > g.n3 + 1) return T(0);
  const T *u = L.u;
  T aw = L.aw, ae = L.ae;
  const int64_t c = gidx(g, ir, jr, kr);
  T uij = u[c];
  T un = u[c + g.sy];
  T us = u[c - g.sy];
  T ue = u[c + 1];
  T ut = u[c + g.sz];
  T ub = kr >= 1 ? u[c - g.sz] : T(0);
  T uw;
  if (ir == 1 && flag_nmn == 1) {
    uw = T(0); aw = T(0); ae = T(2) * ae;
  } else {
    uw = u[c - 1];
  }
  T urhs = L.b[c];
  return urhs - (aw * uw + ae * ue + L.an * un + L.as * us + L.ac * uij + L.at * ut + L.ab * ub);
}

// sor_rb's start (the level active, no sweep counted yet).
__device__ inline void ctl_begin(MGCtl *ctl) {
  ctl->active = 1;
  ctl->ksor = 0;
  ctl->res = 0.0;
  ctl->res_s = 0.0;
}

__global__ void k_reset_totals(MGCtl *ctl, int n) {
  int t = threadIdx.x;
  if (t < n) ctl[t].total = 0;
}

// One colour of sor_rb (sors.f90:87-140 red, 249-302 black).
// colour 0 = red = (i+j+k) odd, 1 = black.  Threads over (i2, j, k).
template <class T, class B>
__device__ inline void sor_node(const KLevelT<T, B> &L, int colour, T omega, int64_t t) {
  const KGrid &g = L.g;
  const int ni2 = g.n1 / 2 + 1;
  int64_t total = (int64_t)ni2 * (g.n2 + 1) * (g.n3 + 1);
  if (t >= total) return;
  const uint32_t tt = (uint32_t)t, r = mdiv_q(tt, L.d_h1), rk = mdiv_q(r, L.d_n2p1);
  int i2 = (int)(tt - r * (uint32_t)ni2);
  int j = 1 + (int)(r - rk * (uint32_t)(g.n2 + 1));
  int k = 1 + (int)rk;
  int ks = (k % 2 == 0) ? 0 : 1;
  int is;
  if (colour == 0) is = (j % 2 == 0) ? 1 : 2;
  else is = (j % 2 == 0) ? 2 : 1;
  is += ks;
  if (is > 2) is = 1;
  int i = is + 2 * i2;
  if (i > g.n1 + 1) return;
  const int64_t c = gidx(g, i, j, k);
  const int bcv = L.bc[c];
  if (bcv > 0) return;
  if (bcv == 0 && (k == g.n3 + 1 || j == g.n2 + 1)) return;
  T *u = L.u;
  // every operand read before any arithmetic, so the reads overlap (the
  // ghost layer keeps u[c - 1] in bounds where the -2 rows do not use it)
  const T uc = u[c], uw = u[c - 1], ue = u[c + 1], un = u[c + g.sy], us = u[c - g.sy],
          ut = u[c + g.sz], ub = u[c - g.sz], bb = L.b[c];
  if (bcv == -2) {
    u[c] = (T(1) - omega) * uc +
           div_ac(omega * (bb - L.an * un - L.as * us - T(2) * L.ae * ue - L.at * ut - L.ab * ub),
                  L);
    return;
  }
  u[c] = (T(1) - omega) * uc +
         div_ac(omega * (bb - L.aw * uw - L.an * un - L.as * us - L.ae * ue - L.at * ut -
                         L.ab * ub),
                L);
}

// ctl == nullptr: a fixed sweep count (the level is always active).
// begin != nullptr: the first red sweep of an sor_rb, which always runs and
// arms the level's control block for the kernels after it.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_sor(KLevelT<T> L, int colour, T omega,
                                               const MGCtl *ctl, MGCtl *begin) {
  if (begin) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_begin(begin);
  } else if (ctl && ctl->active == 0) {
    return;
  }
  sor_node(L, colour, omega, blockIdx.x * (int64_t)BLOCK + threadIdx.x);
}

// Periodic ghost copies (sors.f90:184-239, single rank): y planes for every
// (i, k), then z planes for every (i, j).
template <class T, class B>
__device__ inline void per_y_node(const KLevelT<T, B> &L, int64_t t) {
  const KGrid &g = L.g;
  int64_t ni = g.n1 + 3, nk = g.n3 + 3;
  if (t >= ni * nk) return;
  const uint32_t tt = (uint32_t)t, kk = mdiv_q(tt, L.d_n1p3);
  int i = (int)(tt - kk * (uint32_t)ni), k = (int)kk;
  int b1 = L.bc[gidx(g, i, 1, k)], b2 = L.bc[gidx(g, i, g.n2 + 1, k)];
  if (b1 == 0 || b1 == -2) L.u[gidx(g, i, 0, k)] = L.u[gidx(g, i, g.n2, k)];
  if (b2 == 0 || b2 == -2) L.u[gidx(g, i, g.n2 + 1, k)] = L.u[gidx(g, i, 1, k)];
}

template <class T, class B>
__device__ inline void per_z_node(const KLevelT<T, B> &L, int64_t t) {
  const KGrid &g = L.g;
  int64_t ni = g.n1 + 3, nj = g.n2 + 3;
  if (t >= ni * nj) return;
  const uint32_t tt = (uint32_t)t, jj = mdiv_q(tt, L.d_n1p3);
  int i = (int)(tt - jj * (uint32_t)ni), j = (int)jj;
  const int m = g.n3;
  int b1 = L.bc[gidx(g, i, j, 1)], b2 = L.bc[gidx(g, i, j, m + 1)];
  if (b1 == 0 || b1 == -2) L.u[gidx(g, i, j, 0)] = L.u[gidx(g, i, j, m)];
  if (b2 == 0 || b2 == -2) L.u[gidx(g, i, j, m + 1)] = L.u[gidx(g, i, j, 1)];
}

template <class T>
__global__ void k_per_y(KLevelT<T> L, const MGCtl *ctl) {
  if (ctl && ctl->active == 0) return;
  per_y_node(L, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

template <class T>
__global__ void k_per_z(KLevelT<T> L, const MGCtl *ctl) {
  if (ctl && ctl->active == 0) return;
  per_z_node(L, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

// L1 residual over bcnd == -1 nodes (sors.f90:409-431), block partials.
// |b - A u| at interior node t of the residual range (0 outside plasma).
template <class T, class B>
__device__ inline double resid_node(const KLevelT<T, B> &L, int64_t t) {
  const KGrid &g = L.g;
  int64_t nx = g.n1 + 1, ny = g.n2 + 1, nz = g.n3 + 1;
  double a = 0.0;
  if (t < nx * ny * nz) {
    const uint32_t tt = (uint32_t)t, r = mdiv_q(tt, L.d_n1p1), rk = mdiv_q(r, L.d_n2p1);
    int i = 1 + (int)(tt - r * (uint32_t)nx);
    int j = 1 + (int)(r - rk * (uint32_t)ny);
    int k = 1 + (int)rk;
    const int64_t c = gidx(g, i, j, k);
    if (L.bc[c] == -1) {
      const T *u = L.u;
      const T uc = u[c], uw = u[c - 1], ue = u[c + 1], un = u[c + g.sy], us = u[c - g.sy],
              ut = u[c + g.sz], ub = u[c - g.sz], bb = L.b[c];
      T rij = bb - (L.aw * uw + L.an * un + L.as * us + L.ae * ue + L.ac * uc + L.at * ut +
                    L.ab * ub);
      a = fabs((double)rij);
    }
  }
  return a;
}

// At most RESID_BLOCKS blocks, each thread striding over the nodes, so that
// the one-block decision sums few partials.  The order of the sum is fixed
// for a level (deterministic from run to run).
constexpr int RESID_BLOCKS = 1024;

template <class T>
__global__ __launch_bounds__(BLOCK) void k_resid(KLevelT<T> L, double *partials, const MGCtl *ctl,
                                                 int64_t nres) {
  if (ctl->active == 0) return;
  __shared__ double red[BLOCK / LEPIC_WAVE];
  double a = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < nres; t += stride)
    a = a + resid_node(L, t);
  for (int d = LEPIC_WAVE / 2; d > 0; d >>= 1) a = a + __shfl_down(a, d);
  if (__lane_id() == 0) red[threadIdx.x / LEPIC_WAVE] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < BLOCK / LEPIC_WAVE; ++w) b = b + red[w];
    partials[blockIdx.x] = b;
  }
}

// Sum the partials in a fixed order and apply sor_rb's exit rules
// (sors.f90:446-454).  A separate one-block launch: the kernel boundary
// makes every block's partial visible without a fence per block (a
// last-block-done reduction inside k_resid, with an agent-scope
// acquire-release count per block, measured slower than this launch).
__global__ __launch_bounds__(BLOCK) void k_decide(const double *partials, int64_t nparts,
                                                  MGCtl *ctl, double eps, int dig, int64_t nr) {
  if (ctl->active == 0) return;
  __shared__ double red[BLOCK];
  double a = 0.0;
  for (int64_t p = threadIdx.x; p < nparts; p += BLOCK) a = a + partials[p];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double res = red[0];
    ctl->res = res;
    ctl->ksor += 1;
    ctl->total += 1;
    bool stop;
    if (dig == 0) stop = (res <= eps);
    else stop = (ctl->res_s > 0.0 && (res / ctl->res_s) <= 0.6);
    ctl->res_s = res;
    if (stop || ctl->ksor >= nr) ctl->active = 0;
  }
}


// r_pbc: residual of plane k = m, the periodic image of k = 0
// (mg.f90:322-392).  The reference loops over coarse (j, i) and each
// iteration writes up to nine overlapping entries, the later iteration
// winning; here one thread per entry finds that last writer (largest jr,
// then largest ir) and applies its branch: getres when bcnd(ir,jr,m+1) <= 0,
// else 0.  The two branches guard the (ir, jr-1) entry differently.
template <class T, class B>
__device__ inline void rpbc_entry(const KLevelT<T, B> &F, T *rp, int flag_nmn, int64_t t) {
  const KGrid &g = F.g;
  const int64_t rs = g.n1 + 2;
  const int64_t total = rs * (g.n2 + 2);
  if (t >= total) return;
  const uint32_t tt = (uint32_t)t, bb = mdiv_q(tt, F.d_n1p2);
  const int a = (int)(tt - bb * (uint32_t)rs), b = (int)bb;
  const int m = g.n3, N1 = g.n1 + 1, N2 = g.n2 + 1;
  T val = T(0);
  bool found = false;
  for (int jr = b + 1; jr >= b - 1 && !found; --jr) {
    if (jr < 1 || jr > N2 || (jr % 2) == 0) continue;
    const int db = b - jr;
    for (int ir = a + 1; ir >= a - 1 && !found; --ir) {
      if (ir < 1 || ir > N1 || (ir % 2) == 0) continue;
      const int da = a - ir;
      const bool live = F.bc[gidx(g, ir, jr, m + 1)] <= 0;
      bool writes = true;
      if (da == -1 && ir == 1) writes = false;
      if (da == 1 && ir == N1) writes = false;
      if (db == 1 && jr == N2) writes = false;
      if (db == -1) {
        if (jr == 1) writes = false;
        if (live && da == 0 && ir == 1) writes = false;  // getres branch also needs ir /= 1
      }
      if (!writes) continue;
      found = true;
      val = live ? getres(F, a, b, m, flag_nmn) : T(0);
    }
  }
  rp[t] = val;
}

template <class T>
__global__ void k_rpbc(KLevelT<T> F, T *rp, int flag_nmn) {
  rpbc_entry(F, rp, flag_nmn, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

// Restriction (mg.f90:413-718): coarse labels injected from the fine grid and
// the 27-point full-weighting residual with the reference's neighbour rules.
// R != nullptr: getres of every node of the fine residual range, formed
// beforehand by k_getres (the same expression, so the same bits).
template <class T, class LB>
__device__ inline void restrict_node(const KLevelT<T, LB> &F, const KLevelT<T, LB> &C, const T *rp,
                                     int flag_nmn, int64_t t, const T *R = nullptr) {
  const KGrid &g = F.g;
  const KGrid &gc = C.g;
  const int m = g.n3;
  const int kp = m / 2 + 1;
  int64_t ni = g.n1 / 2 + 1, nj = g.n2 / 2 + 1;
  if (t >= ni * nj * kp) return;
  const uint32_t tt = (uint32_t)t, r = mdiv_q(tt, F.d_h1), rk = mdiv_q(r, F.d_h2);
  int i = 1 + (int)(tt - r * (uint32_t)ni);
  int j = 1 + (int)(r - rk * (uint32_t)nj);
  int k = 1 + (int)rk;
  const int ir = 2 * i - 1, jr = 2 * j - 1, kr = 2 * k - 1;
  const int N1 = g.n1 + 1, N2 = g.n2 + 1, M1 = m + 1;
  const LB *bc = F.bc;
  LB *bc2 = C.bc;
  const LB lab = bc[gidx(g, ir, jr, kr)];
  bc2[gidx(gc, i, j, k)] = lab;
  if (ir == 1) bc2[gidx(gc, 0, j, k)] = bc[gidx(g, 0, jr, kr)];
  if (ir == N1) bc2[gidx(gc, g.n1 / 2 + 2, j, k)] = bc[gidx(g, g.n1 + 2, jr, kr)];
  if (jr == 1) bc2[gidx(gc, i, 0, k)] = bc[gidx(g, ir, 0, kr)];
  if (jr == N2) bc2[gidx(gc, i, g.n2 / 2 + 2, k)] = bc[gidx(g, ir, g.n2 + 2, kr)];
  if (kr == 1) bc2[gidx(gc, i, j, 0)] = bc[gidx(g, ir, jr, 0)];
  if (kr == M1) bc2[gidx(gc, i, j, m / 2 + 2)] = bc[gidx(g, ir, jr, m + 2)];
  T re = T(0);
  if (lab <= 0) {
    const int n2 = g.n2;
    const int64_t rs = g.n1 + 2;
    auto B = [&](int a, int b_, int c) { return bc[gidx(g, a, b_, c)]; };
    auto GR = [&](int a, int b_, int c) {
      const bool in = a >= 1 && a <= N1 && b_ >= 1 && b_ <= N2 && c >= 1 && c <= M1;
      return (R && in) ? R[gidx(g, a, b_, c)] : getres(F, a, b_, c, flag_nmn);
    };
    auto RP = [&](int a, int b_) { return rp ? rp[(int64_t)a + rs * b_] : T(0); };
    const bool kb = (kr == 1);
    T r000 = 0, rpp0 = 0, rp00 = 0, rpm0 = 0, r0m0 = 0, rmm0 = 0, rm00 = 0, rmp0 = 0,
      r0p0 = 0, r00p = 0, rppp = 0, rp0p = 0, rpmp = 0, r0mp = 0, rmmp = 0, rm0p = 0,
      rmpp = 0, r0pp = 0, r00m = 0, rppm = 0, rp0m = 0, rpmm = 0, r0mm = 0, rmmm = 0,
      rm0m = 0, rmpm = 0, r0pm = 0;
    r000 = GR(ir, jr, kr);
    if (ir != N1 && jr != N2) { if (B(ir + 1, jr + 1, kr) <= 0) rpp0 = GR(ir + 1, jr + 1, kr); }
    if (ir != N1) { if (B(ir + 1, jr, kr) <= 0) rp00 = GR(ir + 1, jr, kr); }
    if (B(ir + 1, jr - 1, kr) <= 0) {
      if (ir != N1 && jr != 1) rpm0 = GR(ir + 1, jr - 1, kr);
      else if (jr == 1) rpm0 = GR(ir + 1, n2, kr);
    }
    if (B(ir, jr - 1, kr) <= 0) {
      if (jr == 1) r0m0 = GR(ir, n2, kr);
      else r0m0 = GR(ir, jr - 1, kr);
    }
    if (B(ir - 1, jr - 1, kr) <= 0 && ir != 1) {
      if (jr != 1) rmm0 = GR(ir - 1, jr - 1, kr);
      else rmm0 = GR(ir - 1, n2, kr);
    }
    if (ir != 1) { if (B(ir - 1, jr, kr) <= 0) rm00 = GR(ir - 1, jr, kr); }
    if (ir != 1 && jr != N2) { if (B(ir - 1, jr + 1, kr) <= 0) rmp0 = GR(ir - 1, jr + 1, kr); }
    if (jr != N2) { if (B(ir, jr + 1, kr) <= 0) r0p0 = GR(ir, jr + 1, kr); }
    if (kr != M1) { if (B(ir, jr, kr + 1) <= 0) r00p = GR(ir, jr, kr + 1); }
    if (ir != N1 && jr != N2 && kr != M1) {
      if (B(ir + 1, jr + 1, kr + 1) <= 0) rppp = GR(ir + 1, jr + 1, kr + 1);
    }
    if (ir != N1 && kr != M1) { if (B(ir + 1, jr, kr + 1) <= 0) rp0p = GR(ir + 1, jr, kr + 1); }
    if (B(ir + 1, jr - 1, kr + 1) <= 0) {
      if (ir != N1 && jr != 1 && kr != M1) rpmp = GR(ir + 1, jr - 1, kr + 1);
      else if (jr == 1 && kr != M1) rpmp = GR(ir + 1, n2, kr + 1);
    }
    if (B(ir, jr - 1, kr + 1) <= 0) {
      if (jr != 1 && kr != M1) r0mp = GR(ir, jr - 1, kr + 1);
      else if (jr == 1 && kr != M1) r0mp = GR(ir, n2, kr + 1);
    }
    if (B(ir - 1, jr - 1, kr + 1) <= 0 && ir != 1) {
      if (jr != 1 && kr != M1) rmmp = GR(ir - 1, jr - 1, kr + 1);
      else if (jr == 1 && kr != M1) rmmp = GR(ir - 1, n2, kr + 1);
    }
    if (ir != 1 && kr != M1) { if (B(ir - 1, jr, kr + 1) <= 0) rm0p = GR(ir - 1, jr, kr + 1); }
    if (ir != 1 && jr != N2 && kr != M1) {
      if (B(ir - 1, jr + 1, kr + 1) <= 0) rmpp = GR(ir - 1, jr + 1, kr + 1);
    }
    if (jr != N2 && kr != M1) { if (B(ir, jr + 1, kr + 1) <= 0) r0pp = GR(ir, jr + 1, kr + 1); }
    if (B(ir, jr, kr - 1) <= 0) {
      if (kb) r00m = RP(ir, jr);
      else r00m = GR(ir, jr, kr - 1);
    }
    if (B(ir + 1, jr + 1, kr - 1) <= 0) {
      if (kb) { if (ir != N1 && jr != N2) rppm = RP(ir + 1, jr + 1); }
      else { if (ir != N1 && jr != N2) rppm = GR(ir + 1, jr + 1, kr - 1); }
    }
    if (B(ir + 1, jr, kr - 1) <= 0) {
      if (kb) { if (ir != N1) rp0m = RP(ir + 1, jr); }
      else { if (ir != N1) rp0m = GR(ir + 1, jr, kr - 1); }
    }
    if (B(ir + 1, jr - 1, kr - 1) <= 0) {
      if (kb) { if (ir != N1 && jr != 1) rpmm = RP(ir + 1, jr - 1); }
      else {
        if (ir != N1 && jr != 1) rpmm = GR(ir + 1, jr - 1, kr - 1);
        if (jr == 1) rpmm = GR(ir + 1, n2, kr - 1);
      }
    }
    if (B(ir, jr - 1, kr - 1) <= 0) {
      if (kb) { if (jr != 1) r0mm = RP(ir, jr - 1); }
      else {
        if (jr != 1) r0mm = GR(ir, jr - 1, kr - 1);
        if (jr == 1) r0mm = GR(ir, n2, kr - 1);
      }
    }
    if (B(ir - 1, jr - 1, kr - 1) <= 0 && ir != 1) {
      if (kb) { if (jr != 1) rmmm = RP(ir - 1, jr - 1); }
      else {
        if (jr != 1) rmmm = GR(ir - 1, jr - 1, kr - 1);
        if (jr == 1) rmmm = GR(ir - 1, n2, kr - 1);
      }
    }
    if (B(ir - 1, jr, kr - 1) <= 0 && ir != 1) {
      if (kb) rm0m = RP(ir - 1, jr);
      else rm0m = GR(ir - 1, jr, kr - 1);
    }
    if (B(ir - 1, jr + 1, kr - 1) <= 0 && ir != 1) {
      if (kb) { if (jr != N2) rmpm = RP(ir - 1, jr + 1); }
      else { if (jr != N2) rmpm = GR(ir - 1, jr + 1, kr - 1); }
    }
    if (B(ir, jr + 1, kr - 1) <= 0) {
      if (kb) { if (jr != N2) r0pm = RP(ir, jr + 1); }
      else { if (jr != N2) r0pm = GR(ir, jr + 1, kr - 1); }
    }
    re = (T(8) * r000 + T(4) * (rp00 + rm00 + r0p0 + r0m0 + r00p + r00m) +
          T(2) * (rpp0 + rmp0 + rpm0 + rmm0 + rp0p + rm0p + r0pp + r0mp + rp0m + rm0m + r0pm + r0mm) +
          rppp + rmpp + rpmp + rmmp + rppm + rmpm + rpmm + rmmm) /
         T(64);
  }
  C.b[gidx(gc, i, j, k)] = re;
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_restrict(KLevelT<T> F, KLevelT<T> C, const T *rp,
                                                    int flag_nmn, const T *R) {
  restrict_node(F, C, rp, flag_nmn, blockIdx.x * (int64_t)BLOCK + threadIdx.x, R);
}

// The restriction's first half: getres (mg.f90:809-848) at every node of
// the fine level's residual range (1..n+1 in each direction) into R, once
// per node instead of once per coarse neighbour that reads it, and the
// coarse correction u zeroed (nc nodes).  The restriction's few getres
// calls outside that range (the n2 wrap at i = n1 + 2) still go to getres.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_getres(KLevelT<T> F, T *R, T *cu, int64_t nc,
                                                  int flag_nmn) {
  const int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (t < nc) cu[t] = T(0);
  const KGrid &g = F.g;
  const int64_t nx = g.n1 + 1, ny = g.n2 + 1, nz = g.n3 + 1;
  if (t >= nx * ny * nz) return;
  const uint32_t tt = (uint32_t)t, r = mdiv_q(tt, F.d_n1p1), rk = mdiv_q(r, F.d_n2p1);
  const int i = 1 + (int)(tt - r * (uint32_t)nx);
  const int j = 1 + (int)(r - rk * (uint32_t)ny);
  const int k = 1 + (int)rk;
  R[gidx(g, i, j, k)] = getres(F, i, j, k, flag_nmn);
}

// Prolongation (mg.f90:886-935): one thread per coarse node updates its 8
// fine nodes (disjoint between threads).
template <class T, class B>
__device__ inline void prolong_node(const KLevelT<T, B> &F, const KLevelT<T, B> &C, int64_t t) {
  const KGrid &g = F.g;
  const KGrid &gc = C.g;
  const int kp = g.n3 / 2 + 1;
  int64_t ni = g.n1 / 2 + 1, nj = g.n2 / 2 + 1;
  if (t >= ni * nj * kp) return;
  const uint32_t tt = (uint32_t)t, r = mdiv_q(tt, F.d_h1), rk = mdiv_q(r, F.d_h2);
  int i = 1 + (int)(tt - r * (uint32_t)ni);
  int j = 1 + (int)(r - rk * (uint32_t)nj);
  int k = 1 + (int)rk;
  const T *e = C.u;
  T r000 = e[gidx(gc, i, j, k)], r0m0 = e[gidx(gc, i, j - 1, k)],
         rmm0 = e[gidx(gc, i - 1, j - 1, k)], rm00 = e[gidx(gc, i - 1, j, k)],
         r00m = e[gidx(gc, i, j, k - 1)], r0mm = e[gidx(gc, i, j - 1, k - 1)],
         rmmm = e[gidx(gc, i - 1, j - 1, k - 1)], rm0m = e[gidx(gc, i - 1, j, k - 1)];
  const int ir = 2 * i - 1, jr = 2 * j - 1, kr = 2 * k - 1;
  T *u = F.u;
  const B *bc = F.bc;
  int64_t c;
  c = gidx(g, ir, jr, kr);
  if (bc[c] <= 0) u[c] = u[c] + r000;
  c = gidx(g, ir, jr - 1, kr);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r0m0) / T(2);
  c = gidx(g, ir - 1, jr - 1, kr);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r0m0 + rmm0 + rm00) / T(4);
  c = gidx(g, ir - 1, jr, kr);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + rm00) / T(2);
  c = gidx(g, ir, jr, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m) / T(2);
  c = gidx(g, ir, jr - 1, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + r0mm + r0m0) / T(4);
  c = gidx(g, ir - 1, jr, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + rm0m + rm00) / T(4);
  c = gidx(g, ir - 1, jr - 1, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + rm0m + rm00 + r0m0 + r0mm + rmmm + rmm0) / T(8);
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_prolong(KLevelT<T> F, KLevelT<T> C) {
  prolong_node(F, C, blockIdx.x * (int64_t)BLOCK + threadIdx.x);
}

// --------------------------------------------------------------------------
// Coarse levels in one workgroup.  Below a few thousand nodes a level's
// sweep is far shorter than a kernel dispatch, so the part of a V-cycle that
// only touches levels lc..ng-1 (sor_rb down from lc, the coarsest solve and
// sor_rb back up to lc, with the restrictions and prolongations between
// them) runs as one launch: the same per-node routines as the multi-block
// kernels, phases separated by workgroup barriers, the exit rules of sor_rb
// evaluated in the workgroup.
// --------------------------------------------------------------------------
constexpr int CBLOCK = 1024;
// u, b and the labels of every level the workgroup handles live in its LDS
// during the launch (2 x 8704 doubles and 8704 int16 labels, 153 KB): each
// barrier-separated phase then waits on LDS instead of L2.  lc is the first level from which the remaining
// levels fit.
constexpr int64_t COARSE_LDS_NODES = 8704;

template <class T>
struct CoarseArgs {
  KLevelT<T> L[MAX_LEVELS];   // u and b point to global memory; the kernel works on LDS copies
  int lc, ng;
  T omega;
  double eps;
  int flag_pbc, flag_nmn;
  int first_cycle;
  int fixed;                  // 1: every sor_rb runs its sweep count, no residual exit
  int nr_fixed, nr_coarsest;  // those counts (fixed = 1)
  T *rp;
  MGCtl *ctl;
  int lds_off[MAX_LEVELS];    // level l's u at lds + lds_off[l], b lds_nn[l] after it
  int lds_nn[MAX_LEVELS];
  int lab_off[MAX_LEVELS];    // level l's labels at labs + lab_off[l]
};

// A label as the coarse workgroup keeps it: saturated to int16, which keeps
// its sign and the values 0, -1 and -2, all the solver tests.
__device__ inline int16_t label16(int32_t v) {
  return (int16_t)(v > 32767 ? 32767 : (v < -32768 ? -32768 : v));
}

// Level l as the workgroup sees it: u and b in LDS.  The pointers are formed
// from the workgroup's LDS array itself, never stored and reloaded, so the
// compiler keeps them in the LDS address space (ds_read/ds_write rather than
// flat accesses, which wait on the vector memory counter too).
template <class T>
__device__ inline KLevelT<T, int16_t> lds_level(const CoarseArgs<T> &A, T *lds, int16_t *labs,
                                                int l) {
  const KLevelT<T> &G = A.L[l];
  KLevelT<T, int16_t> v;
  v.g = G.g;
  v.aw = G.aw; v.ae = G.ae; v.an = G.an; v.as = G.as; v.at = G.at; v.ab = G.ab; v.ac = G.ac;
  v.u = lds + A.lds_off[l];
  v.b = lds + A.lds_off[l] + A.lds_nn[l];
  v.bc = labs + A.lab_off[l];
  v.d_h1 = G.d_h1; v.d_n2p1 = G.d_n2p1; v.d_n1p1 = G.d_n1p1;
  v.d_n1p3 = G.d_n1p3; v.d_n1p2 = G.d_n1p2; v.d_h2 = G.d_h2;
  v.acy = G.acy;
  return v;
}

// Workgroup sum in a fixed order: butterfly within each wave, then the 16
// wave sums in index order.  Every thread gets the result.
__device__ inline double wg_sum(double a, double *red) {
  for (int d = 32; d > 0; d >>= 1) a += __shfl_xor(a, d);
  const int wave = threadIdx.x / LEPIC_WAVE;
  if (__lane_id() == 0) red[wave] = a;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < CBLOCK / LEPIC_WAVE; ++w) r = r + red[w];
  __syncthreads();
  return r;
}

// sor_rb (sors.f90:1-535) on level l: at most nr sweeps, exit rules as k_decide
// (with A.fixed, exactly nr sweeps and no residual).
template <class T>
__device__ void wg_sor_rb(const CoarseArgs<T> &A, T *S, int16_t *Lb, int l, int dig, int64_t nr,
                          double *red) {
  const KLevelT<T, int16_t> Lv = lds_level(A, S, Lb, l);
  const KGrid &g = Lv.g;
  const int64_t nsor = (int64_t)(g.n1 / 2 + 1) * (g.n2 + 1) * (g.n3 + 1);
  const int64_t nres = (int64_t)(g.n1 + 1) * (g.n2 + 1) * (g.n3 + 1);
  const int64_t npy = (int64_t)(g.n1 + 3) * (g.n3 + 3), npz = (int64_t)(g.n1 + 3) * (g.n2 + 3);
  int64_t ksor = 0;
  double res = 0.0, res_s = 0.0;
  for (;;) {
    for (int colour = 0; colour < 2; ++colour) {
      for (int64_t t = threadIdx.x; t < nsor; t += CBLOCK) sor_node(Lv, colour, A.omega, t);
      __syncthreads();
      if (A.flag_pbc) {
        for (int64_t t = threadIdx.x; t < npy; t += CBLOCK) per_y_node(Lv, t);
        __syncthreads();
        for (int64_t t = threadIdx.x; t < npz; t += CBLOCK) per_z_node(Lv, t);
        __syncthreads();
      }
    }
    ++ksor;
    if (A.fixed) {
      if (ksor >= nr) break;
      continue;
    }
    double a = 0.0;
    for (int64_t t = threadIdx.x; t < nres; t += CBLOCK) a = a + resid_node(Lv, t);
    res = wg_sum(a, red);
    bool stop;
    if (dig == 0) stop = (res <= A.eps);
    else stop = (res_s > 0.0 && (res / res_s) <= 0.6);
    res_s = res;
    if (stop || ksor >= nr) break;
  }
  if (threadIdx.x == 0 && !A.fixed) {
    MGCtl *c = &A.ctl[l];
    c->active = 0;
    c->ksor = ksor;
    c->res = res;
    c->res_s = res_s;
    c->total += ksor;
  }
}

template <class T>
__device__ void wg_restriction(const CoarseArgs<T> &A, T *S, int16_t *Lb, int l) {
  const KLevelT<T, int16_t> F = lds_level(A, S, Lb, l);
  const KLevelT<T, int16_t> C = lds_level(A, S, Lb, l + 1);
  const KGrid &g = F.g;
  const int64_t nc = ((int64_t)C.g.n1 + 3) * ((int64_t)C.g.n2 + 3) * ((int64_t)C.g.n3 + 3);
  for (int64_t t = threadIdx.x; t < nc; t += CBLOCK) C.u[t] = T(0);
  const T *rpp = nullptr;
  if (A.flag_pbc) {
    const int64_t nent = ((int64_t)g.n1 + 2) * ((int64_t)g.n2 + 2);
    for (int64_t t = threadIdx.x; t < nent; t += CBLOCK) rpbc_entry(F, A.rp, A.flag_nmn, t);
    rpp = A.rp;
  }
  __syncthreads();
  const int64_t nt = (int64_t)(g.n1 / 2 + 1) * (g.n2 / 2 + 1) * (g.n3 / 2 + 1);
  for (int64_t t = threadIdx.x; t < nt; t += CBLOCK) restrict_node(F, C, rpp, A.flag_nmn, t);
  __syncthreads();
}

template <class T>
__device__ void wg_prolongation(const CoarseArgs<T> &A, T *S, int16_t *Lb, int l) {  // L[l] += P(L[l+1])
  const KLevelT<T, int16_t> F = lds_level(A, S, Lb, l);
  const KLevelT<T, int16_t> C = lds_level(A, S, Lb, l + 1);
  const KGrid &g = F.g;
  const int64_t nt = (int64_t)(g.n1 / 2 + 1) * (g.n2 / 2 + 1) * (g.n3 / 2 + 1);
  for (int64_t t = threadIdx.x; t < nt; t += CBLOCK) prolong_node(F, C, t);
  __syncthreads();
}

// The V-cycle of mg (mg.f90:1-246) restricted to levels lc..ng-1; the
// prolongation into level lc-1 is left to the caller.
// The host sizes lc so that levels lc..ng-1 hold at most COARSE_LDS_NODES
// nodes together.  u and b of every level are copied in at the start (the
// global copies of the levels below lc hold what the global-memory version
// would read: zeros at the ghost nodes no restriction writes, values that
// the V-cycle overwrites before reading elsewhere), and u of level lc, the
// only array read after the launch, is copied back at the end.
template <class T>
__global__ __launch_bounds__(CBLOCK) void k_mg_coarse(CoarseArgs<T> A) {
  __shared__ double red[CBLOCK / LEPIC_WAVE];
  __shared__ T lds[2 * COARSE_LDS_NODES];
  __shared__ int16_t labs[COARSE_LDS_NODES];
  T *const S = lds;
  int16_t *const Lb = labs;
  const int ng = A.ng;
  // level lc's labels are the ones the last global restriction wrote; the
  // coarser levels' global labels are never written (0, as at allocation),
  // and the workgroup's restrictions write their values into LDS before any
  // read, leaving 0 where none writes
  for (int l = A.lc; l < ng; ++l) {
    T *p = lds + A.lds_off[l];
    int16_t *q = labs + A.lab_off[l];
    const int nn = A.lds_nn[l];
    for (int t = threadIdx.x; t < nn; t += CBLOCK) {
      p[t] = A.L[l].u[t];
      p[nn + t] = A.L[l].b[t];
      q[t] = label16(A.L[l].bc[t]);
    }
  }
  __syncthreads();
  const int64_t nr = A.fixed ? A.nr_fixed : 4;
  for (int l = A.lc; l <= ng - 2; ++l) {
    wg_sor_rb(A, S, Lb, l, l + 1 - ng, nr, red);
    __syncthreads();
    wg_restriction(A, S, Lb, l);
  }
  wg_sor_rb(A, S, Lb, ng - 1, 0, A.fixed ? A.nr_coarsest : (A.first_cycle ? 1000000 : 4), red);
  __syncthreads();
  if (ng - 2 >= A.lc) wg_prolongation(A, S, Lb, ng - 2);
  for (int l = ng - 2; l >= A.lc; --l) {
    wg_sor_rb(A, S, Lb, l, ng - 1 - l, nr, red);
    __syncthreads();
    if (l - 1 >= A.lc) wg_prolongation(A, S, Lb, l - 1);
  }
  __syncthreads();
  {
    const KGrid &g = A.L[A.lc].g;
    const int64_t nn = ((int64_t)g.n1 + 3) * ((int64_t)g.n2 + 3) * ((int64_t)g.n3 + 3);
    for (int64_t t = threadIdx.x; t < nn; t += CBLOCK) A.L[A.lc].u[t] = lds[A.lds_off[A.lc] + t];
  }
}

template <class T>
__global__ void k_fill_d(T *a, int64_t n, T v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

template <class T>
struct Solver {
  lepic_ctx *ctx;
  KLevelT<T> L[MAX_LEVELS];
  int ng;
  T omega;
  int flag_pbc, flag_nmn;
  double *partials;
  T *rp;
  MGCtl *ctl;  // device, MAX_LEVELS entries
  int lc;      // first level handled by k_mg_coarse (ng: none)
  // fixed sweep counts (the inner smoother of the mixed-precision solve):
  // nr_fixed per sor_rb, nr_coarsest on the coarsest level, no residuals
  int fixed = 0, nr_fixed = 2, nr_coarsest = 64;
  // the fine residual of a restriction (the finest level's ghosted size), or
  // nullptr: the restriction forms it per coarse node
  T *resbuf = nullptr;

  void coarse(double eps, int kmg) {
    CoarseArgs<T> A;
    for (int l = 0; l < MAX_LEVELS; ++l) A.L[l] = L[l];
    A.lc = lc;
    A.ng = ng;
    A.omega = omega;
    A.eps = eps;
    A.flag_pbc = flag_pbc;
    A.flag_nmn = flag_nmn;
    A.first_cycle = kmg == 1 ? 1 : 0;
    A.fixed = fixed;
    A.nr_fixed = nr_fixed;
    A.nr_coarsest = nr_coarsest;
    A.rp = rp;
    A.ctl = ctl;
    int off = 0;
    for (int l = 0; l < MAX_LEVELS; ++l) {
      A.lds_off[l] = 0;
      A.lds_nn[l] = 0;
      A.lab_off[l] = 0;
      if (l < lc || l >= ng) continue;
      A.lds_off[l] = off;
      A.lds_nn[l] = (int)grid_nodes_k(L[l].g);
      A.lab_off[l] = off / 2;
      off += 2 * A.lds_nn[l];
    }
    hipLaunchKernelGGL(k_mg_coarse<T>, dim3(1), dim3(CBLOCK), 0, ctx->stream, A);
  }

  // One sweep; first: the sor_rb's first, whose red half arms ctl[l].
  void sor_sweep(int l, double eps, int dig, int64_t nr, bool first) {
    KLevelT<T> &Lv = L[l];
    const KGrid &g = Lv.g;
    int64_t nsor = (int64_t)(g.n1 / 2 + 1) * (g.n2 + 1) * (g.n3 + 1);
    int64_t nres = (int64_t)(g.n1 + 1) * (g.n2 + 1) * (g.n3 + 1);
    unsigned bs = grid_blocks(nsor, BLOCK);
    const unsigned br = std::min<unsigned>(grid_blocks(nres, BLOCK), RESID_BLOCKS);
    hipStream_t s = ctx->stream;
    const MGCtl *c = fixed ? nullptr : &ctl[l];   // fixed counts: no exit, no control block
    for (int colour = 0; colour < 2; ++colour) {
      MGCtl *begin = (first && colour == 0 && !fixed) ? &ctl[l] : nullptr;
      hipLaunchKernelGGL(k_sor<T>, dim3(bs), dim3(BLOCK), 0, s, Lv, colour, omega, c, begin);
      if (flag_pbc) {
        hipLaunchKernelGGL(k_per_y<T>, dim3(grid_blocks((int64_t)(g.n1 + 3) * (g.n3 + 3), BLOCK)),
                           dim3(BLOCK), 0, s, Lv, c);
        hipLaunchKernelGGL(k_per_z<T>, dim3(grid_blocks((int64_t)(g.n1 + 3) * (g.n2 + 3), BLOCK)),
                           dim3(BLOCK), 0, s, Lv, c);
      }
    }
    if (fixed) return;   // a fixed sweep count: no residual, the level stays active
    hipLaunchKernelGGL(k_resid<T>, dim3(br), dim3(BLOCK), 0, s, Lv, partials, &ctl[l], nres);
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(BLOCK), 0, s, partials, (int64_t)br, &ctl[l], eps,
                       dig, nr);
  }

  // sor_rb: returns LEPIC_OK; sweeps are counted on the device.
  int sor_rb(int l, double eps, int kmg, int dig) {
    if (fixed) {
      for (int c = 0; c < (dig == 0 ? nr_coarsest : nr_fixed); ++c) sor_sweep(l, eps, dig, 4, false);
      return LEPIC_OK;
    }
    if (kmg == 1 && dig == 0) {
      // first V-cycle's coarsest level: up to 1e6 sweeps to res <= eps.
      const int64_t nr = 1000000;
      int64_t done = 0;
      while (done < nr) {
        int chunk = 32;
        for (int c = 0; c < chunk; ++c) sor_sweep(l, eps, dig, nr, done == 0 && c == 0);
        done += chunk;
        LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, &ctl[l], sizeof(MGCtl), hipMemcpyDeviceToHost,
                                      ctx->stream));
        LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->h_pinned[0] == 0) break;
      }
    } else {
      for (int c = 0; c < 4; ++c) sor_sweep(l, eps, dig, 4, c == 0);
    }
    return LEPIC_OK;
  }

  void restriction(int l) {
    KLevelT<T> &F = L[l];
    KLevelT<T> &C = L[l + 1];
    const KGrid &g = F.g;
    int64_t nc = grid_nodes_k(C.g);
    if (resbuf) {
      const int64_t nres = (int64_t)(g.n1 + 1) * (g.n2 + 1) * (g.n3 + 1);
      hipLaunchKernelGGL(k_getres<T>, dim3(grid_blocks(std::max(nc, nres), BLOCK)), dim3(BLOCK), 0,
                         ctx->stream, F, resbuf, C.u, nc, flag_nmn);
    } else {
      hipLaunchKernelGGL(k_fill_d<T>, dim3(grid_blocks(nc, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                         C.u, nc, T(0));
    }
    const T *rpp = nullptr;
    if (flag_pbc) {
      int64_t nent = (int64_t)(g.n1 + 2) * (g.n2 + 2);
      hipLaunchKernelGGL(k_rpbc<T>, dim3(grid_blocks(nent, BLOCK)), dim3(BLOCK), 0, ctx->stream, F,
                         rp, flag_nmn);
      rpp = rp;
    }
    int64_t nt = (int64_t)(g.n1 / 2 + 1) * (g.n2 / 2 + 1) * (g.n3 / 2 + 1);
    hipLaunchKernelGGL(k_restrict<T>, dim3(grid_blocks(nt, BLOCK)), dim3(BLOCK), 0, ctx->stream, F,
                       C, rpp, flag_nmn, (const T *)resbuf);
  }

  void prolongation(int l) {  // L[l] += P(L[l+1])
    const KGrid &g = L[l].g;
    int64_t nt = (int64_t)(g.n1 / 2 + 1) * (g.n2 / 2 + 1) * (g.n3 / 2 + 1);
    hipLaunchKernelGGL(k_prolong<T>, dim3(grid_blocks(nt, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       L[l], L[l + 1]);
  }

  static int64_t grid_nodes_k(const KGrid &g) {
    return ((int64_t)g.n1 + 3) * ((int64_t)g.n2 + 3) * ((int64_t)g.n3 + 3);
  }

  // One V-cycle of mg (mg.f90:99-239).  Levels lc..ng-1 run inside
  // k_mg_coarse: the stretch from the down-leg sor_rb on lc to the up-leg
  // sor_rb on lc touches no level above lc and is issued as one launch.
  int vcycle(int kmg, double epsa) {
    for (int ig = 1; ig <= 2 * ng - 1; ++ig) {
      for (int il = 1; il <= ng - 1 && il <= 11; ++il) {
        const int l = il - 1;
        if (ig < ng && ig == il) {
          if (l < lc) {
            int st = sor_rb(l, epsa, kmg, ig - ng);
            if (st) return st;
            restriction(l);
          } else if (l == lc) {
            coarse(epsa, kmg);
          }
        }
        if (ig > ng && ig == 2 * ng - il) {
          if (l < lc) {
            int st = sor_rb(l, epsa, kmg, ig - ng);
            if (st) return st;
          }
          if (l <= lc && il >= 2) prolongation(il - 2);
        }
      }
      if (ig == ng) {
        if (lc >= ng) {
          int st = sor_rb(ng - 1, epsa, kmg, 0);
          if (st) return st;
        } else if (lc == ng - 1) {
          coarse(epsa, kmg);
        }
        if (ng >= 2 && ng - 2 < lc) prolongation(ng - 2);
      }
    }
    return LEPIC_OK;
  }

  // lc: the first level from which the remaining levels fit the coarse
  // workgroup's LDS (ng: none do).
  void choose_lc() {
    lc = ng;
    for (int l = 1; l < ng; ++l) {
      int64_t tot = 0;
      for (int m = l; m < ng; ++m) tot += grid_nodes_k(L[m].g);
      if (tot <= COARSE_LDS_NODES) {
        lc = l;
        break;
      }
    }
  }
};

// mg.h:9-15, formed in FP64 (and rounded once for an FP32 level).
template <class T>
void set_coeffs(KLevelT<T> &L, double eps0) {
  const KGrid &g = L.g;
  const double aw = eps0 / (g.h1 * g.h1);
  const double ae = aw;
  const double an = eps0 / (g.h2 * g.h2);
  const double as = an;
  const double at = eps0 / (g.h3 * g.h3);
  const double ab = at;
  const double ac = -(an + as + aw + ae + at + ab);
  L.aw = (T)aw; L.ae = (T)ae; L.an = (T)an; L.as = (T)as; L.at = (T)at; L.ab = (T)ab;
  L.ac = (T)ac;
  L.acy = qdiv_divisor_ok(ac) ? 1.0 / ac : 0.0;
}

}  // namespace

extern "C" int lepic_poisson_solve(lepic_ctx *ctx, const lepic_grid *g, double *phi,
                                   const double *rho, const int32_t *bcnd,
                                   const lepic_mg_params *mp, lepic_mg_stats *stats) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!g || !phi || !rho || !bcnd || !mp) return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson: null argument");
  const int ng = mp->ng;
  if (ng < 1 || ng > MAX_LEVELS) return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson: ng outside [1,12]");
  if (grid_nodes(g) > INT32_MAX)   // thread numbers are decoded in 32 bits
    return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson: more than 2^31 nodes");
  Solver<double> S;
  S.ctx = ctx;
  S.ng = ng;
  S.omega = mp->omega;
  S.flag_pbc = mp->flag_pbc;
  S.flag_nmn = mp->flag_nmn;
  // level geometry: n/2^(l-1), h*2^(l-1) (pdesolver.f90:42-51)
  int64_t coarse_nodes = 0, max_parts = 0;
  for (int l = 0; l < ng; ++l) {
    lepic_grid gl;
    for (int d = 0; d < 3; ++d) {
      gl.n[d] = g->n[d] >> l;
      gl.h[d] = g->h[d] * (double)(1 << l);
    }
    if (l < ng - 1)
      for (int d = 0; d < 3; ++d)
        if (gl.n[d] < 2 || (gl.n[d] % 2) != 0)
          return lepic_fail(ctx, LEPIC_ERR_ARG,
                            "poisson: every level above the coarsest needs even n >= 2");
    if (gl.n[0] < 1 || gl.n[1] < 1 || gl.n[2] < 1)
      return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson: too many levels for this grid");
    S.L[l].g = make_kgrid(&gl);
    set_level_divs(S.L[l]);
    set_coeffs(S.L[l], mp->eps0);
    if (l > 0) coarse_nodes += grid_nodes(&gl);
    int64_t nres = (int64_t)(gl.n[0] + 1) * (gl.n[1] + 1) * (gl.n[2] + 1);
    max_parts = std::max(max_parts, (nres + BLOCK - 1) / BLOCK);
  }
  // coarse-level storage (kept across calls; labels zeroed once)
  const int64_t rp_len = ((int64_t)g->n[0] + 2) * ((int64_t)g->n[1] + 2);
  const int64_t fine_nodes = ng > 1 ? grid_nodes(g) : 0;   // the restriction's residual
  size_t need = sizeof(double) * (size_t)(2 * coarse_nodes + max_parts + rp_len + fine_nodes) +
                sizeof(int32_t) * (size_t)coarse_nodes + 4096;
  bool same = ctx->mg_mem && ctx->mg_ng == ng && ctx->mg_n[0] == g->n[0] &&
              ctx->mg_n[1] == g->n[1] && ctx->mg_n[2] == g->n[2] && ctx->mg_bytes >= need;
  if (!same) {
    if (ctx->mg_mem) {
      LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      LEPIC_HIP(ctx, hipFree(ctx->mg_mem));
      ctx->mg_mem = nullptr;
    }
    LEPIC_HIP(ctx, hipMalloc(&ctx->mg_mem, need));
    LEPIC_HIP(ctx, hipMemsetAsync(ctx->mg_mem, 0, need, ctx->stream));
    ctx->mg_bytes = need;
    ctx->mg_ng = ng;
    for (int d = 0; d < 3; ++d) ctx->mg_n[d] = g->n[d];
  }
  char *mem = reinterpret_cast<char *>(ctx->mg_mem);
  double *dmem = reinterpret_cast<double *>(mem);
  S.partials = dmem;
  S.rp = dmem + max_parts;
  double *lv = S.rp + rp_len;
  S.resbuf = fine_nodes ? lv + 2 * coarse_nodes : nullptr;
  int32_t *lb = reinterpret_cast<int32_t *>(lv + 2 * coarse_nodes + fine_nodes);
  S.L[0].u = phi;
  S.L[0].b = const_cast<double *>(rho);
  S.L[0].bc = const_cast<int32_t *>(bcnd);
  for (int l = 1; l < ng; ++l) {
    int64_t nn = Solver<double>::grid_nodes_k(S.L[l].g);
    S.L[l].u = lv;
    S.L[l].b = lv + nn;
    lv += 2 * nn;
    S.L[l].bc = lb;
    lb += nn;
  }
  S.ctl = reinterpret_cast<MGCtl *>(ctx->d_ctrl + CTRL_MG);
  hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(64), 0, ctx->stream, S.ctl, MAX_LEVELS);

  // pdesolver: alpha with the fine-grid ac (pdesolver.f90:33-34)
  const double alpha =
      std::fabs((double)((g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1)) * S.L[0].ac) * 1.0;
  const double epsa = alpha * mp->eps;
  double res = 0.0, res_s = 0.0;
  int k = 1;
  int status = LEPIC_OK;
  S.choose_lc();
  auto vcycle = [&](int kmg) -> int { return S.vcycle(kmg, epsa); };
  // With the coarsest level inside k_mg_coarse a V-cycle's launches are fixed
  // (every decision is taken on the device), so they are captured once per
  // solve configuration -- the first V-cycle and the later ones differ in one
  // argument -- and replayed: one submission per V-cycle instead of ~100.
  const bool replay = mp->graph && S.lc < ng;
  if (replay) {
    const int32_t ints[9] = {ng, g->n[0], g->n[1], g->n[2], S.flag_pbc, S.flag_nmn, S.lc, 0, 0};
    const void *ptrs[5] = {phi, rho, bcnd, ctx->mg_mem, ctx->d_ctrl};
    const double dbl[5] = {g->h[0], g->h[1], g->h[2], S.omega, epsa};
    std::vector<unsigned char> key(sizeof(ints) + sizeof(ptrs) + sizeof(dbl));
    std::memcpy(key.data(), ints, sizeof(ints));
    std::memcpy(key.data() + sizeof(ints), ptrs, sizeof(ptrs));
    std::memcpy(key.data() + sizeof(ints) + sizeof(ptrs), dbl, sizeof(dbl));
    if (key != ctx->mg_key || !ctx->mg_exec[0] || !ctx->mg_exec[1]) {
      for (hipGraphExec_t &e : ctx->mg_exec)
        if (e) {
          LEPIC_HIP(ctx, hipGraphExecDestroy(e));
          e = nullptr;
        }
      ctx->mg_key.clear();
      if (!ctx->capture_stream)
        LEPIC_HIP(ctx, hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
      const hipStream_t eager = ctx->stream;
      ctx->stream = ctx->capture_stream;   // the Solver launches on ctx->stream
      int st = LEPIC_OK;
      for (int which = 0; which < 2 && st == LEPIC_OK; ++which) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
          st = vcycle(which == 0 ? 1 : 2);
          e = hipStreamEndCapture(ctx->capture_stream, &graph);
        }
        if (e == hipSuccess && st == LEPIC_OK)
          e = hipGraphInstantiate(&ctx->mg_exec[which], graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess && st == LEPIC_OK) st = lepic_fail(ctx, LEPIC_ERR_HIP, hipGetErrorString(e));
      }
      ctx->stream = eager;
      if (st) return st;
      ctx->mg_key = key;
    }
  }
  for (;;) {
    if (replay) {
      LEPIC_HIP(ctx, hipGraphLaunch(ctx->mg_exec[k == 1 ? 0 : 1], ctx->stream));
    } else {
      int st = vcycle(k);
      if (st) return st;
    }
    LEPIC_LAUNCH_CHECK(ctx, "multigrid launch");
    // the V-cycle's residual is the last sor_rb's: level 1 (or the only level)
    // every level's counters come back with the fine residual (no extra
    // round trip for the statistics)
    LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, S.ctl, sizeof(MGCtl) * MAX_LEVELS,
                                  hipMemcpyDeviceToHost, ctx->stream));
    LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double r;
    std::memcpy(&r, &ctx->h_pinned[2], sizeof(double));
    res = r / alpha;
    if (k == 1) res_s = res;
    else if (std::fabs(res / res_s) > 10.0) {
      status = lepic_fail(ctx, LEPIC_ERR_DIVERGED, "poisson: res/res_s > 10 (diverging)");
      break;
    }
    if (res < mp->eps || k == mp->ncycl) break;
    k = k + 1;
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    double ktot = 0.0;
    for (int l = 0; l < ng; ++l) {
      int64_t tot = ctx->h_pinned[16 * l + 4];
      stats->sweeps[l] = tot;
      ktot = ktot + (double)tot / std::pow(2.0, 2.0 * l);
    }
    stats->cycles = k;
    stats->ktot = ktot;
    stats->res = res;
  }
  return status;
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 red-black smoothing V-cycles inside an FP64 iterative
// refinement.  Each refinement step forms the FP64 residual r = b - A phi at
// every node the smoother updates (the smoother's own row rules: Neumann rows
// with 2 ae and no west term, periodic copies left to the ghost update),
// rounds it to FP32, runs a few fixed-sweep FP32 V-cycles of the same
// multigrid on A e = r with e = 0 on electrodes, and adds e to phi in FP64.
// Convergence is the reference's: L1 residual over plasma nodes / alpha
// below eps (pdesolver.f90:97-115), measured on the FP64 residual.
// ---------------------------------------------------------------------------
namespace {

constexpr int MIX_BLOCKS = 1024;

// FP64 residual of the refinement step at ghosted node t; r32, e32 (the
// FP32 right-hand side and correction) written for every node, the L1 sum
// over plasma nodes (in the order of sors.f90:409-431's residual) to
// partials.
__global__ __launch_bounds__(BLOCK) void k_mix_resid(KLevel L, float *r32, float *e32,
                                                     double *partials, int64_t nodes) {
  __shared__ double red[BLOCK / LEPIC_WAVE];
  const KGrid &g = L.g;
  double a = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < nodes; t += stride) {
    const int i = (int)(t % g.sy);
    const int64_t q = t / g.sy;
    const int j = (int)(q % (g.n2 + 3)), k = (int)(q / (g.n2 + 3));
    double r = 0.0;
    if (i >= 1 && i <= g.n1 + 1 && j >= 1 && j <= g.n2 + 1 && k >= 1 && k <= g.n3 + 1) {
      const int bcv = L.bc[t];
      const double *u = L.u;
      if (bcv == -2) {
        r = L.b[t] - (L.an * u[t + g.sy] + L.as * u[t - g.sy] + 2.0 * L.ae * u[t + 1] +
                      L.at * u[t + g.sz] + L.ab * u[t - g.sz] + L.ac * u[t]);
      } else if (bcv == -1 || (bcv == 0 && j != g.n2 + 1 && k != g.n3 + 1)) {
        r = L.b[t] - (L.aw * u[t - 1] + L.an * u[t + g.sy] + L.as * u[t - g.sy] +
                      L.ae * u[t + 1] + L.ac * u[t] + L.at * u[t + g.sz] + L.ab * u[t - g.sz]);
      }
      if (bcv == -1) a = a + fabs(r);
    }
    r32[t] = (float)r;
    e32[t] = 0.0f;
  }
  for (int d = LEPIC_WAVE / 2; d > 0; d >>= 1) a = a + __shfl_down(a, d);
  if (__lane_id() == 0) red[threadIdx.x / LEPIC_WAVE] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0.0;
    for (int w = 0; w < BLOCK / LEPIC_WAVE; ++w) b = b + red[w];
    partials[blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(BLOCK) void k_mix_sum(const double *partials, int64_t nparts,
                                                   double *out) {
  __shared__ double red[BLOCK];
  double a = 0.0;
  for (int64_t p = threadIdx.x; p < nparts; p += BLOCK) a = a + partials[p];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// phi += e where the smoother may change phi (bcnd <= 0).
__global__ void k_mix_correct(double *phi, const float *e32, const int32_t *bc, int64_t nodes) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < nodes && bc[t] <= 0) phi[t] = phi[t] + (double)e32[t];
}

}  // namespace

// ---------------------------------------------------------------------------
// Set-up shared by the two solves that smooth in FP32: the FP32 correction
// hierarchy and the FP64 work arrays on ctx->mgx_mem, laid out as
// [MIX_COLS x MIX_BLOCKS partial sums][8 scalars][nf64 FP64 node arrays]
// [e32, r32][periodic residual plane][coarse u, b][coarse labels].
// ---------------------------------------------------------------------------
namespace {

constexpr int MIX_COLS = 4;

template <class T>
struct MixSetup {
  Solver<T> S;         // the correction hierarchy (T: FP32, or FP64 for the CG preconditioner)
  KLevel F64;          // the FP64 fine level: phi, rho, labels
  double *partials = nullptr, *scal = nullptr, *f64 = nullptr;
  T *e32 = nullptr, *r32 = nullptr;
  int64_t nodes = 0;
  double alpha = 0.0;  // pdesolver's residual normalisation
  unsigned rb = 0;     // blocks of the FP64 residual kernels
};

template <class T>
int mix_setup(lepic_ctx *ctx, const lepic_grid *g, double *phi, const double *rho,
              const int32_t *bcnd, const lepic_mg_params *mp, const lepic_mgmix_params *xp,
              int nf64, const char *who, MixSetup<T> &M) {
  const std::string w(who);
  if (!g || !phi || !rho || !bcnd || !mp || !xp)
    return lepic_fail(ctx, LEPIC_ERR_ARG, w + ": null argument");
  const int ng = mp->ng;
  if (ng < 1 || ng > MAX_LEVELS) return lepic_fail(ctx, LEPIC_ERR_ARG, w + ": ng outside [1,12]");
  if (xp->inner_cycles < 1 || xp->inner_cycles > 64 || xp->nu < 1 || xp->nu > 16 ||
      xp->nu_coarsest < 1 || xp->nu_coarsest > 100000 || !(xp->omega > 0.0 && xp->omega < 2.0))
    return lepic_fail(ctx, LEPIC_ERR_ARG, w + ": inner V-cycle parameters out of range");
  if (grid_nodes(g) > INT32_MAX)   // thread numbers are decoded in 32 bits
    return lepic_fail(ctx, LEPIC_ERR_ARG, w + ": more than 2^31 nodes");
  Solver<T> &S = M.S;
  S.ctx = ctx;
  S.ng = ng;
  S.omega = (T)xp->omega;
  S.flag_pbc = mp->flag_pbc;
  S.flag_nmn = mp->flag_nmn;
  S.fixed = 1;
  S.nr_fixed = xp->nu;
  S.nr_coarsest = xp->nu_coarsest;
  int64_t coarse_nodes = 0;
  for (int l = 0; l < ng; ++l) {
    lepic_grid gl;
    for (int d = 0; d < 3; ++d) {
      gl.n[d] = g->n[d] >> l;
      gl.h[d] = g->h[d] * (double)(1 << l);
    }
    if (l < ng - 1)
      for (int d = 0; d < 3; ++d)
        if (gl.n[d] < 2 || (gl.n[d] % 2) != 0)
          return lepic_fail(ctx, LEPIC_ERR_ARG,
                            w + ": every level above the coarsest needs even n >= 2");
    if (gl.n[0] < 1 || gl.n[1] < 1 || gl.n[2] < 1)
      return lepic_fail(ctx, LEPIC_ERR_ARG, w + ": too many levels for this grid");
    S.L[l].g = make_kgrid(&gl);
    set_level_divs(S.L[l]);
    set_coeffs(S.L[l], mp->eps0);
    if (l == 0) {
      M.F64.g = S.L[0].g;
      set_level_divs(M.F64);
      set_coeffs(M.F64, mp->eps0);
    } else {
      coarse_nodes += grid_nodes(&gl);
    }
  }
  const int64_t nodes = grid_nodes(g);
  const int64_t rp_len = ((int64_t)g->n[0] + 2) * ((int64_t)g->n[1] + 2);
  const size_t need = sizeof(double) * (size_t)(MIX_COLS * MIX_BLOCKS + 8) +
                      sizeof(double) * (size_t)nf64 * (size_t)nodes +
                      sizeof(T) * (size_t)(2 * nodes + 2 * coarse_nodes + rp_len) +
                      sizeof(T) * (size_t)(ng > 1 ? nodes : 0) +   // the restriction's residual
                      sizeof(int32_t) * (size_t)coarse_nodes + 4096;
  if (!ctx->mgx_mem || ctx->mgx_bytes < need) {
    if (ctx->mgx_mem) {
      LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      LEPIC_HIP(ctx, hipFree(ctx->mgx_mem));
      ctx->mgx_mem = nullptr;
      ctx->mgx_bytes = 0;
      ctx->mgx_key.clear();
      ctx->pcg_key.clear();
    }
    LEPIC_HIP(ctx, hipMalloc(&ctx->mgx_mem, need));
    LEPIC_HIP(ctx, hipMemsetAsync(ctx->mgx_mem, 0, need, ctx->stream));
    ctx->mgx_bytes = need;
  }
  M.partials = reinterpret_cast<double *>(ctx->mgx_mem);
  M.scal = M.partials + MIX_COLS * MIX_BLOCKS;
  M.f64 = M.scal + 8;
  T *fmem = reinterpret_cast<T *>(M.f64 + (size_t)nf64 * (size_t)nodes);
  M.e32 = fmem;
  M.r32 = fmem + nodes;
  T *lv = M.r32 + nodes;
  S.rp = lv;
  lv += rp_len;
  S.resbuf = ng > 1 ? lv + 2 * coarse_nodes : nullptr;
  int32_t *lb = reinterpret_cast<int32_t *>(lv + 2 * coarse_nodes + (ng > 1 ? nodes : 0));
  S.L[0].u = M.e32;
  S.L[0].b = M.r32;
  S.L[0].bc = const_cast<int32_t *>(bcnd);
  for (int l = 1; l < ng; ++l) {
    const int64_t nn = Solver<T>::grid_nodes_k(S.L[l].g);
    S.L[l].u = lv;
    S.L[l].b = lv + nn;
    lv += 2 * nn;
    S.L[l].bc = lb;
    lb += nn;
  }
  S.partials = M.partials;   // unused with fixed sweeps
  S.ctl = reinterpret_cast<MGCtl *>(ctx->d_ctrl + CTRL_MG);
  S.choose_lc();
  M.F64.u = phi;
  M.F64.b = const_cast<double *>(rho);
  M.F64.bc = const_cast<int32_t *>(bcnd);
  M.nodes = nodes;
  M.alpha =
      std::fabs((double)((g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1)) * M.F64.ac) * 1.0;
  M.rb = (unsigned)std::min<int64_t>(grid_blocks(nodes, BLOCK), MIX_BLOCKS);
  return LEPIC_OK;
}

}  // namespace

extern "C" int lepic_poisson_solve_mixed(lepic_ctx *ctx, const lepic_grid *g, double *phi,
                                         const double *rho, const int32_t *bcnd,
                                         const lepic_mg_params *mp, const lepic_mgmix_params *xp,
                                         lepic_mg_stats *stats) {
  if (!ctx) return LEPIC_ERR_ARG;
  MixSetup<float> M;
  int st0 = mix_setup(ctx, g, phi, rho, bcnd, mp, xp, 0, "poisson_mixed", M);
  if (st0) return st0;
  Solver<float> &S = M.S;
  KLevel &F64 = M.F64;
  const int ng = mp->ng;
  const int64_t nodes = M.nodes;
  double *partials = M.partials;
  double *res_dev = M.scal;
  float *e32 = M.e32, *r32 = M.r32;
  const double alpha = M.alpha;
  const unsigned rb = M.rb;
  // the inner V-cycles and the correction of one refinement step: one
  // captured graph per solve configuration
  auto inner = [&]() -> int {
    for (int c = 0; c < xp->inner_cycles; ++c) {
      int st = S.vcycle(2, 0.0);
      if (st) return st;
    }
    hipLaunchKernelGGL(k_mix_correct, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                       phi, e32, bcnd, nodes);
    if (mp->flag_pbc) {   // phi's periodic ghosts from their images (sors.f90:184-239)
      const KGrid &gg = F64.g;
      hipLaunchKernelGGL(k_per_y<double>, dim3(grid_blocks((int64_t)(gg.n1 + 3) * (gg.n3 + 3), BLOCK)),
                         dim3(BLOCK), 0, ctx->stream, F64, &S.ctl[0]);
      hipLaunchKernelGGL(k_per_z<double>, dim3(grid_blocks((int64_t)(gg.n1 + 3) * (gg.n2 + 3), BLOCK)),
                         dim3(BLOCK), 0, ctx->stream, F64, &S.ctl[0]);
    }
    return LEPIC_OK;
  };
  const bool replay = mp->graph != 0;
  if (replay) {
    const int32_t ints[10] = {ng, g->n[0], g->n[1], g->n[2], mp->flag_pbc, mp->flag_nmn,
                              xp->inner_cycles, xp->nu, xp->nu_coarsest, S.lc};
    const void *ptrs[4] = {phi, bcnd, ctx->mgx_mem, ctx->d_ctrl};
    const double dbl[5] = {g->h[0], g->h[1], g->h[2], xp->omega, mp->eps0};
    std::vector<unsigned char> key(sizeof(ints) + sizeof(ptrs) + sizeof(dbl));
    std::memcpy(key.data(), ints, sizeof(ints));
    std::memcpy(key.data() + sizeof(ints), ptrs, sizeof(ptrs));
    std::memcpy(key.data() + sizeof(ints) + sizeof(ptrs), dbl, sizeof(dbl));
    if (key != ctx->mgx_key || !ctx->mgx_exec) {
      if (ctx->mgx_exec) {
        LEPIC_HIP(ctx, hipGraphExecDestroy(ctx->mgx_exec));
        ctx->mgx_exec = nullptr;
      }
      ctx->mgx_key.clear();
      if (!ctx->capture_stream)
        LEPIC_HIP(ctx, hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
      const hipStream_t eager = ctx->stream;
      ctx->stream = ctx->capture_stream;   // the Solver launches on ctx->stream
      int st = LEPIC_OK;
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        st = inner();
        e = hipStreamEndCapture(ctx->capture_stream, &graph);
      }
      if (e == hipSuccess && st == LEPIC_OK)
        e = hipGraphInstantiate(&ctx->mgx_exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      ctx->stream = eager;
      if (st) return st;
      if (e != hipSuccess) return lepic_fail(ctx, LEPIC_ERR_HIP, hipGetErrorString(e));
      ctx->mgx_key = key;
    }
  }
  hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(64), 0, ctx->stream, S.ctl, MAX_LEVELS);
  double res = 0.0, res_s = 0.0;
  int k = 0;
  int status = LEPIC_OK;
  for (;;) {
    hipLaunchKernelGGL(k_mix_resid, dim3(rb), dim3(BLOCK), 0, ctx->stream, F64, r32, e32,
                       partials, nodes);
    hipLaunchKernelGGL(k_mix_sum, dim3(1), dim3(BLOCK), 0, ctx->stream, partials, (int64_t)rb,
                       res_dev);
    LEPIC_LAUNCH_CHECK(ctx, "mixed residual");
    LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, res_dev, sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
    LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double r;
    std::memcpy(&r, &ctx->h_pinned[0], sizeof(double));
    res = r / alpha;
    if (k == 0) res_s = res;
    else if (std::fabs(res / res_s) > 10.0) {
      status = lepic_fail(ctx, LEPIC_ERR_DIVERGED, "poisson_mixed: res/res_s > 10 (diverging)");
      break;
    }
    if (res < mp->eps || k == mp->ncycl) break;
    ++k;
    if (replay) {
      LEPIC_HIP(ctx, hipGraphLaunch(ctx->mgx_exec, ctx->stream));
    } else {
      int st = inner();
      if (st) return st;
    }
    LEPIC_LAUNCH_CHECK(ctx, "mixed refinement step");
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    double ktot = 0.0;
    for (int l = 0; l < ng; ++l) {
      // fixed sweep counts: nu per leg on every level above the coarsest
      const int64_t per = (l == ng - 1) ? xp->nu_coarsest : 2 * xp->nu;
      const int64_t tot = per * (int64_t)xp->inner_cycles * k;
      stats->sweeps[l] = tot;
      ktot = ktot + (double)tot / std::pow(2.0, 2.0 * l);
    }
    stats->cycles = k;
    stats->ktot = ktot;
    stats->res = res;
  }
  return status;
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradients: the FP32 V-cycles of the mixed solve
// used as the preconditioner of an FP64 flexible CG (Notay's FCG with one
// previous direction kept) on the plasma nodes.  With plasma (-1) and
// electrode (> 0) labels and no periodic or Neumann walls the operator is
// the symmetric 7-point Dirichlet operator of sors.f90 restricted to plasma
// nodes; the V-cycle is applied to A z = r from z = 0 with fixed sweeps.  An
// iteration: z = M(r); p = z - (<z, A p_old> / <p_old, A p_old>) p_old;
// q = A p; phi += a p, r -= a q with a = <p, r> / <p, q>.  It stops on the
// reference's criterion (L1 residual over plasma nodes / alpha < eps),
// confirmed on the residual recomputed from phi.
// ---------------------------------------------------------------------------
namespace {

__device__ inline double block_sum_d(double a, double *red) {
  for (int d = LEPIC_WAVE / 2; d > 0; d >>= 1) a = a + __shfl_down(a, d);
  if (__lane_id() == 0) red[threadIdx.x / LEPIC_WAVE] = a;
  __syncthreads();
  double b = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BLOCK / LEPIC_WAVE; ++w) b = b + red[w];
  __syncthreads();
  return b;
}

// The residual from phi (start and restarts): r = b - A phi on plasma nodes,
// 0 elsewhere; the directions p, q cleared; r32 = r, e32 = 0; the L1 sum to
// partials column 0; a physical node labelled 0 or -2, or a ghost node
// labelled -1, raises *bad.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_pcg_resid(KLevel L, MDiv dy, double *r, double *p,
                                                     double *q, T *r32, T *e32,
                                                     double *partials, int *bad, int64_t nodes) {
  __shared__ double red[BLOCK / LEPIC_WAVE];
  const KGrid &g = L.g;
  double a = 0.0;
  int flag = 0;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < nodes; t += stride) {
    const uint32_t tt = (uint32_t)t, row = mdiv_q(tt, L.d_n1p3), kk = mdiv_q(row, dy);
    const int i = (int)(tt - row * (uint32_t)(g.n1 + 3));
    const int j = (int)(row - kk * (uint32_t)(g.n2 + 3)), k = (int)kk;
    double rv = 0.0;
    if (i >= 1 && i <= g.n1 + 1 && j >= 1 && j <= g.n2 + 1 && k >= 1 && k <= g.n3 + 1) {
      const int bcv = L.bc[t];
      const double *u = L.u;
      if (bcv == -1) {
        rv = L.b[t] - (L.aw * u[t - 1] + L.an * u[t + g.sy] + L.as * u[t - g.sy] +
                       L.ae * u[t + 1] + L.ac * u[t] + L.at * u[t + g.sz] + L.ab * u[t - g.sz]);
        a = a + fabs(rv);
      } else if (bcv == 0 || bcv == -2) {
        flag = 1;
      }
    } else if (L.bc[t] == -1) {
      flag = 1;   // a plasma label on a ghost node: its stencil would leave the array
    }
    r[t] = rv;
    p[t] = 0.0;
    q[t] = 0.0;
    r32[t] = (T)rv;
    e32[t] = T(0);
  }
  const double s = block_sum_d(a, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
  if (__syncthreads_or(flag) && threadIdx.x == 0) atomicOr(bad, 1);
}

// <z, q_old> (z = e32, zero off the plasma nodes) to partials column 1.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_pcg_zq(const T *e32, const double *q,
                                                  double *partials, int64_t nodes) {
  __shared__ double red[BLOCK / LEPIC_WAVE];
  double a = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < nodes; t += stride)
    a = a + (double)e32[t] * q[t];
  const double s = block_sum_d(a, red);
  if (threadIdx.x == 0) partials[MIX_BLOCKS + blockIdx.x] = s;
}

// The new direction p_new = z - beta p (beta = scal[1] / scal[2], the
// previous <p, q>) on plasma nodes, 0 elsewhere, and q = A p_new from the
// neighbours' p_new formed on the fly (z and p are 0 off the plasma nodes);
// <p_new, q> and <p_new, r> to partials columns 2 and 3.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_pcg_dir(KLevel L, const T *e32, const double *p,
                                                   double *pn_out, double *q, const double *r,
                                                   const double *scal, double *partials,
                                                   int64_t nodes) {
  __shared__ double red[BLOCK / LEPIC_WAVE];
  const KGrid &g = L.g;
  const double beta = scal[2] != 0.0 ? scal[1] / scal[2] : 0.0;
  double apq = 0.0, apr = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < nodes; t += stride) {
    double pn = 0.0, qv = 0.0;
    if (L.bc[t] == -1) {
      auto P = [&](int64_t u) { return (double)e32[u] - beta * p[u]; };
      pn = P(t);
      qv = L.aw * P(t - 1) + L.an * P(t + g.sy) + L.as * P(t - g.sy) + L.ae * P(t + 1) +
           L.ac * pn + L.at * P(t + g.sz) + L.ab * P(t - g.sz);
      apq = apq + pn * qv;
      apr = apr + pn * r[t];
    }
    pn_out[t] = pn;
    q[t] = qv;
  }
  const double s1 = block_sum_d(apq, red);
  const double s2 = block_sum_d(apr, red);
  if (threadIdx.x == 0) {
    partials[2 * MIX_BLOCKS + blockIdx.x] = s1;
    partials[3 * MIX_BLOCKS + blockIdx.x] = s2;
  }
}

// phi += a p_new, r -= a q (a = scal[3] / scal[2]) on plasma nodes; r32 = r,
// e32 = 0, p = p_new; the L1 sum of r to partials column 0.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_pcg_update(double *phi, const int32_t *bc,
                                                      const double *pn, const double *q,
                                                      double *p, double *r, T *r32,
                                                      T *e32, const double *scal,
                                                      double *partials, int64_t nodes) {
  __shared__ double red[BLOCK / LEPIC_WAVE];
  const double a = scal[2] != 0.0 ? scal[3] / scal[2] : 0.0;
  double l1 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < nodes; t += stride) {
    const double pv = pn[t];
    if (bc[t] == -1) {
      phi[t] = phi[t] + a * pv;
      const double rv = r[t] - a * q[t];
      r[t] = rv;
      r32[t] = (T)rv;
      l1 = l1 + fabs(rv);
    }
    p[t] = pv;
    e32[t] = T(0);
  }
  const double s = block_sum_d(l1, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// scal[first + c] = the sum of partials column first + c (one block each).
__global__ __launch_bounds__(BLOCK) void k_pcg_sum(const double *partials, int64_t nparts,
                                                   int first, double *scal) {
  __shared__ double red[BLOCK];
  const int col = first + (int)blockIdx.x;
  const double *pc = partials + (int64_t)col * MIX_BLOCKS;
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < nparts; i += BLOCK) a = a + pc[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) scal[col] = red[0];
}

__global__ void k_pcg_clear_flag(int *bad) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *bad = 0;
}

}  // namespace

namespace {

template <class T>
int pcg_solve(lepic_ctx *ctx, const lepic_grid *g, double *phi, const double *rho,
              const int32_t *bcnd, const lepic_mg_params *mp, const lepic_mgmix_params *xp,
              lepic_mg_stats *stats) {
  MixSetup<T> M;
  // FP64 arrays: r, p, p_new, q
  int st0 = mix_setup(ctx, g, phi, rho, bcnd, mp, xp, 4, "poisson_pcg", M);
  if (st0) return st0;
  Solver<T> &S = M.S;
  KLevel &F64 = M.F64;
  const int ng = mp->ng;
  const int64_t nodes = M.nodes;
  double *r = M.f64, *p = r + nodes, *pn = p + nodes, *q = pn + nodes;
  double *scal = M.scal;
  int *bad = reinterpret_cast<int *>(scal + 4);
  const unsigned rb = M.rb;
  const MDiv dy = mdiv((uint32_t)(g->n[1] + 3));
  hipStream_t &s = ctx->stream;
  auto iteration = [&]() -> int {
    for (int c = 0; c < xp->inner_cycles; ++c) {
      int st = S.vcycle(2, 0.0);
      if (st) return st;
    }
    hipLaunchKernelGGL(k_pcg_zq<T>, dim3(rb), dim3(BLOCK), 0, s, M.e32, q, M.partials, nodes);
    hipLaunchKernelGGL(k_pcg_sum, dim3(1), dim3(BLOCK), 0, s, M.partials, (int64_t)rb, 1, scal);
    hipLaunchKernelGGL(k_pcg_dir<T>, dim3(rb), dim3(BLOCK), 0, s, F64, M.e32, p, pn, q, r, scal,
                       M.partials, nodes);
    hipLaunchKernelGGL(k_pcg_sum, dim3(2), dim3(BLOCK), 0, s, M.partials, (int64_t)rb, 2, scal);
    hipLaunchKernelGGL(k_pcg_update<T>, dim3(rb), dim3(BLOCK), 0, s, phi, bcnd, pn, q, p, r, M.r32,
                       M.e32, scal, M.partials, nodes);
    hipLaunchKernelGGL(k_pcg_sum, dim3(1), dim3(BLOCK), 0, s, M.partials, (int64_t)rb, 0, scal);
    return LEPIC_OK;
  };
  const bool replay = mp->graph != 0;
  if (replay) {
    const int32_t ints[10] = {ng, g->n[0], g->n[1], g->n[2], 0, (int32_t)sizeof(T),
                              xp->inner_cycles, xp->nu, xp->nu_coarsest, S.lc};
    const void *ptrs[4] = {phi, bcnd, ctx->mgx_mem, ctx->d_ctrl};
    const double dbl[5] = {g->h[0], g->h[1], g->h[2], xp->omega, mp->eps0};
    std::vector<unsigned char> key(sizeof(ints) + sizeof(ptrs) + sizeof(dbl));
    std::memcpy(key.data(), ints, sizeof(ints));
    std::memcpy(key.data() + sizeof(ints), ptrs, sizeof(ptrs));
    std::memcpy(key.data() + sizeof(ints) + sizeof(ptrs), dbl, sizeof(dbl));
    if (key != ctx->pcg_key || !ctx->pcg_exec) {
      if (ctx->pcg_exec) {
        LEPIC_HIP(ctx, hipGraphExecDestroy(ctx->pcg_exec));
        ctx->pcg_exec = nullptr;
      }
      ctx->pcg_key.clear();
      if (!ctx->capture_stream)
        LEPIC_HIP(ctx, hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
      const hipStream_t eager = ctx->stream;
      ctx->stream = ctx->capture_stream;   // the Solver and the lambda launch on ctx->stream
      int st = LEPIC_OK;
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        st = iteration();
        e = hipStreamEndCapture(ctx->capture_stream, &graph);
      }
      if (e == hipSuccess && st == LEPIC_OK)
        e = hipGraphInstantiate(&ctx->pcg_exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      ctx->stream = eager;
      if (st) return st;
      if (e != hipSuccess) return lepic_fail(ctx, LEPIC_ERR_HIP, hipGetErrorString(e));
      ctx->pcg_key = key;
    }
  }
  // r from phi (and the directions cleared), read back with the label flag
  auto restart = [&]() -> int {
    hipLaunchKernelGGL(k_pcg_clear_flag, dim3(1), dim3(64), 0, s, bad);
    hipLaunchKernelGGL(k_pcg_resid<T>, dim3(rb), dim3(BLOCK), 0, s, F64, dy, r, p, q, M.r32, M.e32,
                       M.partials, bad, nodes);
    hipLaunchKernelGGL(k_pcg_sum, dim3(1), dim3(BLOCK), 0, s, M.partials, (int64_t)rb, 0, scal);
    LEPIC_LAUNCH_CHECK(ctx, "pcg residual");
    return LEPIC_OK;
  };
  auto read_back = [&](double &res, int &flag) -> int {
    LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, scal, 5 * sizeof(double), hipMemcpyDeviceToHost,
                                  s));
    LEPIC_HIP(ctx, hipStreamSynchronize(s));
    double v;
    std::memcpy(&v, &ctx->h_pinned[0], sizeof(double));
    std::memcpy(&flag, &ctx->h_pinned[4], sizeof(int));
    res = v / M.alpha;
    return LEPIC_OK;
  };
  hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(64), 0, s, S.ctl, MAX_LEVELS);
  int st = restart();
  if (st) return st;
  double res = 0.0;
  int flag = 0;
  if ((st = read_back(res, flag))) return st;
  if (flag)
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      "poisson_pcg: labels 0 or -2 (the operator is not symmetric), or plasma "
                      "on a ghost node");
  const double res_s = res;
  int k = 0;
  bool fresh = true;   // r was just formed from phi
  int status = LEPIC_OK;
  for (;;) {
    if (res < mp->eps) {
      if (fresh) break;
      if ((st = restart())) return st;   // confirm on the residual from phi
      if ((st = read_back(res, flag))) return st;
      fresh = true;
      if (res < mp->eps) break;
    }
    if (k == mp->ncycl) break;
    ++k;
    if (replay) {
      LEPIC_HIP(ctx, hipGraphLaunch(ctx->pcg_exec, s));
    } else if ((st = iteration())) {
      return st;
    }
    LEPIC_LAUNCH_CHECK(ctx, "pcg iteration");
    if ((st = read_back(res, flag))) return st;
    fresh = false;
    // the residual of conjugate gradients need not fall at every iteration:
    // only a non-finite or exploding one stops the solve
    if (!(res <= 1.0e6 * res_s)) {
      status = lepic_fail(ctx, LEPIC_ERR_DIVERGED, "poisson_pcg: res/res_s > 1e6 (diverging)");
      break;
    }
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    double ktot = 0.0;
    for (int l = 0; l < ng; ++l) {
      const int64_t per = (l == ng - 1) ? xp->nu_coarsest : 2 * xp->nu;
      const int64_t tot = per * (int64_t)xp->inner_cycles * k;
      stats->sweeps[l] = tot;
      ktot = ktot + (double)tot / std::pow(2.0, 2.0 * l);
    }
    stats->cycles = k;
    stats->ktot = ktot;
    stats->res = res;
  }
  return status;
}

}  // namespace

extern "C" int lepic_poisson_solve_pcg(lepic_ctx *ctx, const lepic_grid *g, double *phi,
                                       const double *rho, const int32_t *bcnd,
                                       const lepic_mg_params *mp, const lepic_mgmix_params *xp,
                                       lepic_mg_stats *stats) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!mp || !xp) return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson_pcg: null argument");
  if (mp->flag_pbc || mp->flag_nmn)
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      "poisson_pcg: periodic or Neumann walls (the operator is not symmetric)");
  if (xp->precond_fp64 != 0 && xp->precond_fp64 != 1)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson_pcg: precond_fp64 must be 0 or 1");
  return xp->precond_fp64 ? pcg_solve<double>(ctx, g, phi, rho, bcnd, mp, xp, stats)
                          : pcg_solve<float>(ctx, g, phi, rho, bcnd, mp, xp, stats);
}
