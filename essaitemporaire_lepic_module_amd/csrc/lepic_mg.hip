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

struct KLevel {
  KGrid g;
  double aw, ae, an, as, at, ab, ac;
  double *u;
  double *b;
  int32_t *bc;
};

__device__ inline double getres(const KLevel &L, int ir, int jr, int kr, int flag_nmn) {
  const KGrid &g = L.g;
  if (ir < 0 || ir > g.n1 + 1 || jr < 0 || jr > g.n2 + 1 || kr < 0 || kr > g.n3 + 1) return 0.0;
  const double *u = L.u;
  double aw = L.aw, ae = L.ae;
  const int64_t c = gidx(g, ir, jr, kr);
  double uij = u[c];
  double un = u[c + g.sy];
  double us = u[c - g.sy];
  double ue = u[c + 1];
  double ut = u[c + g.sz];
  double ub = kr >= 1 ? u[c - g.sz] : 0.0;
  double uw;
  if (ir == 1 && flag_nmn == 1) {
    uw = 0.0; aw = 0.0; ae = 2.0 * ae;
  } else {
    uw = u[c - 1];
  }
  double urhs = L.b[c];
  return urhs - (aw * uw + ae * ue + L.an * un + L.as * us + L.ac * uij + L.at * ut + L.ab * ub);
}

__global__ void k_begin(MGCtl *ctl) {
  if (threadIdx.x == 0) {
    ctl->active = 1;
    ctl->ksor = 0;
    ctl->res = 0.0;
    ctl->res_s = 0.0;
  }
}

__global__ void k_reset_totals(MGCtl *ctl, int n) {
  int t = threadIdx.x;
  if (t < n) ctl[t].total = 0;
}

// One colour of sor_rb (sors.f90:87-140 red, 249-302 black).
// colour 0 = red = (i+j+k) odd, 1 = black.  Threads over (i2, j, k).
__device__ inline void sor_node(const KLevel &L, int colour, double omega, int64_t t) {
  const KGrid &g = L.g;
  const int ni2 = g.n1 / 2 + 1;
  int64_t total = (int64_t)ni2 * (g.n2 + 1) * (g.n3 + 1);
  if (t >= total) return;
  int i2 = (int)(t % ni2);
  int64_t r = t / ni2;
  int j = 1 + (int)(r % (g.n2 + 1));
  int k = 1 + (int)(r / (g.n2 + 1));
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
  double *u = L.u;
  if (bcv == -2) {
    u[c] = (1.0 - omega) * u[c] +
           omega * (L.b[c] - L.an * u[c + g.sy] - L.as * u[c - g.sy] - 2.0 * L.ae * u[c + 1] -
                    L.at * u[c + g.sz] - L.ab * u[c - g.sz]) /
               L.ac;
    return;
  }
  if (bcv == 0) {
    if (k == g.n3 + 1) return;
    if (j == g.n2 + 1) return;
  }
  u[c] = (1.0 - omega) * u[c] +
         omega * (L.b[c] - L.aw * u[c - 1] - L.an * u[c + g.sy] - L.as * u[c - g.sy] -
                  L.ae * u[c + 1] - L.at * u[c + g.sz] - L.ab * u[c - g.sz]) /
             L.ac;
}

__global__ __launch_bounds__(BLOCK) void k_sor(KLevel L, int colour, double omega,
                                               const MGCtl *ctl) {
  if (ctl->active == 0) return;
  sor_node(L, colour, omega, blockIdx.x * (int64_t)BLOCK + threadIdx.x);
}

// Periodic ghost copies (sors.f90:184-239, single rank): y planes for every
// (i, k), then z planes for every (i, j).
__device__ inline void per_y_node(const KLevel &L, int64_t t) {
  const KGrid &g = L.g;
  int64_t ni = g.n1 + 3, nk = g.n3 + 3;
  if (t >= ni * nk) return;
  int i = (int)(t % ni), k = (int)(t / ni);
  int b1 = L.bc[gidx(g, i, 1, k)], b2 = L.bc[gidx(g, i, g.n2 + 1, k)];
  if (b1 == 0 || b1 == -2) L.u[gidx(g, i, 0, k)] = L.u[gidx(g, i, g.n2, k)];
  if (b2 == 0 || b2 == -2) L.u[gidx(g, i, g.n2 + 1, k)] = L.u[gidx(g, i, 1, k)];
}

__device__ inline void per_z_node(const KLevel &L, int64_t t) {
  const KGrid &g = L.g;
  int64_t ni = g.n1 + 3, nj = g.n2 + 3;
  if (t >= ni * nj) return;
  int i = (int)(t % ni), j = (int)(t / ni);
  const int m = g.n3;
  int b1 = L.bc[gidx(g, i, j, 1)], b2 = L.bc[gidx(g, i, j, m + 1)];
  if (b1 == 0 || b1 == -2) L.u[gidx(g, i, j, 0)] = L.u[gidx(g, i, j, m)];
  if (b2 == 0 || b2 == -2) L.u[gidx(g, i, j, m + 1)] = L.u[gidx(g, i, j, 1)];
}

__global__ void k_per_y(KLevel L, const MGCtl *ctl) {
  if (ctl->active == 0) return;
  per_y_node(L, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

__global__ void k_per_z(KLevel L, const MGCtl *ctl) {
  if (ctl->active == 0) return;
  per_z_node(L, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

// L1 residual over bcnd == -1 nodes (sors.f90:409-431), block partials.
// |b - A u| at interior node t of the residual range (0 outside plasma).
__device__ inline double resid_node(const KLevel &L, int64_t t) {
  const KGrid &g = L.g;
  int64_t nx = g.n1 + 1, ny = g.n2 + 1, nz = g.n3 + 1;
  double a = 0.0;
  if (t < nx * ny * nz) {
    int i = 1 + (int)(t % nx);
    int64_t r = t / nx;
    int j = 1 + (int)(r % ny);
    int k = 1 + (int)(r / ny);
    const int64_t c = gidx(g, i, j, k);
    if (L.bc[c] == -1) {
      const double *u = L.u;
      double rij = L.b[c] - (L.aw * u[c - 1] + L.an * u[c + g.sy] + L.as * u[c - g.sy] +
                             L.ae * u[c + 1] + L.ac * u[c] + L.at * u[c + g.sz] +
                             L.ab * u[c - g.sz]);
      a = fabs(rij);
    }
  }
  return a;
}

// At most RESID_BLOCKS blocks, each thread striding over the nodes, so that
// the one-block decision sums few partials.  The order of the sum is fixed
// for a level (deterministic from run to run).
constexpr int RESID_BLOCKS = 1024;

__global__ __launch_bounds__(BLOCK) void k_resid(KLevel L, double *partials, const MGCtl *ctl,
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
// makes every block's partial visible without a fence per block.
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
__device__ inline void rpbc_entry(const KLevel &F, double *rp, int flag_nmn, int64_t t) {
  const KGrid &g = F.g;
  const int64_t rs = g.n1 + 2;
  const int64_t total = rs * (g.n2 + 2);
  if (t >= total) return;
  const int a = (int)(t % rs), b = (int)(t / rs);
  const int m = g.n3, N1 = g.n1 + 1, N2 = g.n2 + 1;
  double val = 0.0;
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
      val = live ? getres(F, a, b, m, flag_nmn) : 0.0;
    }
  }
  rp[t] = val;
}

__global__ void k_rpbc(KLevel F, double *rp, int flag_nmn) {
  rpbc_entry(F, rp, flag_nmn, blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
}

// Restriction (mg.f90:413-718): coarse labels injected from the fine grid and
// the 27-point full-weighting residual with the reference's neighbour rules.
__device__ inline void restrict_node(const KLevel &F, const KLevel &C, const double *rp,
                                     int flag_nmn, int64_t t) {
  const KGrid &g = F.g;
  const KGrid &gc = C.g;
  const int m = g.n3;
  const int kp = m / 2 + 1;
  int64_t ni = g.n1 / 2 + 1, nj = g.n2 / 2 + 1;
  if (t >= ni * nj * kp) return;
  int i = 1 + (int)(t % ni);
  int64_t r = t / ni;
  int j = 1 + (int)(r % nj);
  int k = 1 + (int)(r / nj);
  const int ir = 2 * i - 1, jr = 2 * j - 1, kr = 2 * k - 1;
  const int N1 = g.n1 + 1, N2 = g.n2 + 1, M1 = m + 1;
  const int32_t *bc = F.bc;
  int32_t *bc2 = C.bc;
  int32_t lab = bc[gidx(g, ir, jr, kr)];
  bc2[gidx(gc, i, j, k)] = lab;
  if (ir == 1) bc2[gidx(gc, 0, j, k)] = bc[gidx(g, 0, jr, kr)];
  if (ir == N1) bc2[gidx(gc, g.n1 / 2 + 2, j, k)] = bc[gidx(g, g.n1 + 2, jr, kr)];
  if (jr == 1) bc2[gidx(gc, i, 0, k)] = bc[gidx(g, ir, 0, kr)];
  if (jr == N2) bc2[gidx(gc, i, g.n2 / 2 + 2, k)] = bc[gidx(g, ir, g.n2 + 2, kr)];
  if (kr == 1) bc2[gidx(gc, i, j, 0)] = bc[gidx(g, ir, jr, 0)];
  if (kr == M1) bc2[gidx(gc, i, j, m / 2 + 2)] = bc[gidx(g, ir, jr, m + 2)];
  double re = 0.0;
  if (lab <= 0) {
    const int n2 = g.n2;
    const int64_t rs = g.n1 + 2;
    auto B = [&](int a, int b_, int c) { return bc[gidx(g, a, b_, c)]; };
    auto GR = [&](int a, int b_, int c) { return getres(F, a, b_, c, flag_nmn); };
    auto RP = [&](int a, int b_) { return rp ? rp[(int64_t)a + rs * b_] : 0.0; };
    const bool kb = (kr == 1);
    double r000 = 0, rpp0 = 0, rp00 = 0, rpm0 = 0, r0m0 = 0, rmm0 = 0, rm00 = 0, rmp0 = 0,
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
    re = (8.0 * r000 + 4.0 * (rp00 + rm00 + r0p0 + r0m0 + r00p + r00m) +
          2.0 * (rpp0 + rmp0 + rpm0 + rmm0 + rp0p + rm0p + r0pp + r0mp + rp0m + rm0m + r0pm + r0mm) +
          rppp + rmpp + rpmp + rmmp + rppm + rmpm + rpmm + rmmm) /
         64.0;
  }
  C.b[gidx(gc, i, j, k)] = re;
}

__global__ __launch_bounds__(BLOCK) void k_restrict(KLevel F, KLevel C, const double *rp,
                                                    int flag_nmn) {
  restrict_node(F, C, rp, flag_nmn, blockIdx.x * (int64_t)BLOCK + threadIdx.x);
}

// Prolongation (mg.f90:886-935): one thread per coarse node updates its 8
// fine nodes (disjoint between threads).
__device__ inline void prolong_node(const KLevel &F, const KLevel &C, int64_t t) {
  const KGrid &g = F.g;
  const KGrid &gc = C.g;
  const int kp = g.n3 / 2 + 1;
  int64_t ni = g.n1 / 2 + 1, nj = g.n2 / 2 + 1;
  if (t >= ni * nj * kp) return;
  int i = 1 + (int)(t % ni);
  int64_t r = t / ni;
  int j = 1 + (int)(r % nj);
  int k = 1 + (int)(r / nj);
  const double *e = C.u;
  double r000 = e[gidx(gc, i, j, k)], r0m0 = e[gidx(gc, i, j - 1, k)],
         rmm0 = e[gidx(gc, i - 1, j - 1, k)], rm00 = e[gidx(gc, i - 1, j, k)],
         r00m = e[gidx(gc, i, j, k - 1)], r0mm = e[gidx(gc, i, j - 1, k - 1)],
         rmmm = e[gidx(gc, i - 1, j - 1, k - 1)], rm0m = e[gidx(gc, i - 1, j, k - 1)];
  const int ir = 2 * i - 1, jr = 2 * j - 1, kr = 2 * k - 1;
  double *u = F.u;
  const int32_t *bc = F.bc;
  int64_t c;
  c = gidx(g, ir, jr, kr);
  if (bc[c] <= 0) u[c] = u[c] + r000;
  c = gidx(g, ir, jr - 1, kr);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r0m0) / 2.0;
  c = gidx(g, ir - 1, jr - 1, kr);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r0m0 + rmm0 + rm00) / 4.0;
  c = gidx(g, ir - 1, jr, kr);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + rm00) / 2.0;
  c = gidx(g, ir, jr, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m) / 2.0;
  c = gidx(g, ir, jr - 1, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + r0mm + r0m0) / 4.0;
  c = gidx(g, ir - 1, jr, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + rm0m + rm00) / 4.0;
  c = gidx(g, ir - 1, jr - 1, kr - 1);
  if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + rm0m + rm00 + r0m0 + r0mm + rmmm + rmm0) / 8.0;
}

__global__ __launch_bounds__(BLOCK) void k_prolong(KLevel F, KLevel C) {
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
// u and b of every level the workgroup handles live in its LDS during the
// launch (2 x 8704 doubles, 136 KB): each barrier-separated phase then waits
// on LDS instead of L2.  lc is the first level from which the remaining
// levels fit.
constexpr int64_t COARSE_LDS_NODES = 8704;

struct CoarseArgs {
  KLevel L[MAX_LEVELS];   // u and b point to global memory; the kernel works on LDS copies
  int lc, ng;
  double omega, eps;
  int flag_pbc, flag_nmn;
  int first_cycle;
  double *rp;
  MGCtl *ctl;
};

// Level l as the workgroup sees it: u and b in LDS.
struct CoarseLDS {
  double *u[MAX_LEVELS];
  double *b[MAX_LEVELS];
};

__device__ inline KLevel lds_level(const CoarseArgs &A, const CoarseLDS &S, int l) {
  KLevel v = A.L[l];
  v.u = S.u[l];
  v.b = S.b[l];
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

// sor_rb (sors.f90:1-535) on level l: at most nr sweeps, exit rules as k_decide.
__device__ void wg_sor_rb(const CoarseArgs &A, const CoarseLDS &S, int l, int dig, int64_t nr,
                          double *red) {
  const KLevel Lv = lds_level(A, S, l);
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
    double a = 0.0;
    for (int64_t t = threadIdx.x; t < nres; t += CBLOCK) a = a + resid_node(Lv, t);
    res = wg_sum(a, red);
    ++ksor;
    bool stop;
    if (dig == 0) stop = (res <= A.eps);
    else stop = (res_s > 0.0 && (res / res_s) <= 0.6);
    res_s = res;
    if (stop || ksor >= nr) break;
  }
  if (threadIdx.x == 0) {
    MGCtl *c = &A.ctl[l];
    c->active = 0;
    c->ksor = ksor;
    c->res = res;
    c->res_s = res_s;
    c->total += ksor;
  }
}

__device__ void wg_restriction(const CoarseArgs &A, const CoarseLDS &S, int l) {
  const KLevel F = lds_level(A, S, l);
  const KLevel C = lds_level(A, S, l + 1);
  const KGrid &g = F.g;
  const int64_t nc = ((int64_t)C.g.n1 + 3) * ((int64_t)C.g.n2 + 3) * ((int64_t)C.g.n3 + 3);
  for (int64_t t = threadIdx.x; t < nc; t += CBLOCK) C.u[t] = 0.0;
  const double *rpp = nullptr;
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

__device__ void wg_prolongation(const CoarseArgs &A, const CoarseLDS &S, int l) {  // L[l] += P(L[l+1])
  const KLevel F = lds_level(A, S, l);
  const KLevel C = lds_level(A, S, l + 1);
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
__global__ __launch_bounds__(CBLOCK) void k_mg_coarse(CoarseArgs A) {
  __shared__ double red[CBLOCK / LEPIC_WAVE];
  __shared__ double lds[2 * COARSE_LDS_NODES];
  __shared__ CoarseLDS S;   // level pointers into lds, uniform
  const int ng = A.ng;
  {
    double *p = lds;
    for (int l = A.lc; l < ng; ++l) {
      const KGrid &g = A.L[l].g;
      const int64_t nn = ((int64_t)g.n1 + 3) * ((int64_t)g.n2 + 3) * ((int64_t)g.n3 + 3);
      if (threadIdx.x == 0) {
        S.u[l] = p;
        S.b[l] = p + nn;
      }
      for (int64_t t = threadIdx.x; t < nn; t += CBLOCK) {
        p[t] = A.L[l].u[t];
        p[nn + t] = A.L[l].b[t];
      }
      p += 2 * nn;
    }
  }
  __syncthreads();
  for (int l = A.lc; l <= ng - 2; ++l) {
    wg_sor_rb(A, S, l, l + 1 - ng, 4, red);
    __syncthreads();
    wg_restriction(A, S, l);
  }
  wg_sor_rb(A, S, ng - 1, 0, A.first_cycle ? 1000000 : 4, red);
  __syncthreads();
  if (ng - 2 >= A.lc) wg_prolongation(A, S, ng - 2);
  for (int l = ng - 2; l >= A.lc; --l) {
    wg_sor_rb(A, S, l, ng - 1 - l, 4, red);
    __syncthreads();
    if (l - 1 >= A.lc) wg_prolongation(A, S, l - 1);
  }
  __syncthreads();
  {
    const KGrid &g = A.L[A.lc].g;
    const int64_t nn = ((int64_t)g.n1 + 3) * ((int64_t)g.n2 + 3) * ((int64_t)g.n3 + 3);
    for (int64_t t = threadIdx.x; t < nn; t += CBLOCK) A.L[A.lc].u[t] = S.u[A.lc][t];
  }
}

__global__ void k_fill_d(double *a, int64_t n, double v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

struct Solver {
  lepic_ctx *ctx;
  KLevel L[MAX_LEVELS];
  int ng;
  double omega;
  int flag_pbc, flag_nmn;
  double *partials;
  double *rp;
  MGCtl *ctl;  // device, MAX_LEVELS entries
  int lc;      // first level handled by k_mg_coarse (ng: none)

  void coarse(double eps, int kmg) {
    CoarseArgs A;
    for (int l = 0; l < MAX_LEVELS; ++l) A.L[l] = L[l];
    A.lc = lc;
    A.ng = ng;
    A.omega = omega;
    A.eps = eps;
    A.flag_pbc = flag_pbc;
    A.flag_nmn = flag_nmn;
    A.first_cycle = kmg == 1 ? 1 : 0;
    A.rp = rp;
    A.ctl = ctl;
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(CBLOCK), 0, ctx->stream, A);
  }

  void sor_sweep(int l, double eps, int dig, int64_t nr) {
    KLevel &Lv = L[l];
    const KGrid &g = Lv.g;
    int64_t nsor = (int64_t)(g.n1 / 2 + 1) * (g.n2 + 1) * (g.n3 + 1);
    int64_t nres = (int64_t)(g.n1 + 1) * (g.n2 + 1) * (g.n3 + 1);
    unsigned bs = grid_blocks(nsor, BLOCK);
    const unsigned br = std::min<unsigned>(grid_blocks(nres, BLOCK), RESID_BLOCKS);
    hipStream_t s = ctx->stream;
    for (int colour = 0; colour < 2; ++colour) {
      hipLaunchKernelGGL(k_sor, dim3(bs), dim3(BLOCK), 0, s, Lv, colour, omega, &ctl[l]);
      if (flag_pbc) {
        hipLaunchKernelGGL(k_per_y, dim3(grid_blocks((int64_t)(g.n1 + 3) * (g.n3 + 3), BLOCK)),
                           dim3(BLOCK), 0, s, Lv, &ctl[l]);
        hipLaunchKernelGGL(k_per_z, dim3(grid_blocks((int64_t)(g.n1 + 3) * (g.n2 + 3), BLOCK)),
                           dim3(BLOCK), 0, s, Lv, &ctl[l]);
      }
    }
    hipLaunchKernelGGL(k_resid, dim3(br), dim3(BLOCK), 0, s, Lv, partials, &ctl[l], nres);
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(BLOCK), 0, s, partials, (int64_t)br, &ctl[l], eps,
                       dig, nr);
  }

  // sor_rb: returns LEPIC_OK; sweeps are counted on the device.
  int sor_rb(int l, double eps, int kmg, int dig) {
    hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, ctx->stream, &ctl[l]);
    if (kmg == 1 && dig == 0) {
      // first V-cycle's coarsest level: up to 1e6 sweeps to res <= eps.
      const int64_t nr = 1000000;
      int64_t done = 0;
      while (done < nr) {
        int chunk = 32;
        for (int c = 0; c < chunk; ++c) sor_sweep(l, eps, dig, nr);
        done += chunk;
        LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, &ctl[l], sizeof(MGCtl), hipMemcpyDeviceToHost,
                                      ctx->stream));
        LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->h_pinned[0] == 0) break;
      }
    } else {
      for (int c = 0; c < 4; ++c) sor_sweep(l, eps, dig, 4);
    }
    return LEPIC_OK;
  }

  void restriction(int l) {
    KLevel &F = L[l];
    KLevel &C = L[l + 1];
    const KGrid &g = F.g;
    int64_t nc = grid_nodes_k(C.g);
    hipLaunchKernelGGL(k_fill_d, dim3(grid_blocks(nc, BLOCK)), dim3(BLOCK), 0, ctx->stream, C.u,
                       nc, 0.0);
    const double *rpp = nullptr;
    if (flag_pbc) {
      int64_t nent = (int64_t)(g.n1 + 2) * (g.n2 + 2);
      hipLaunchKernelGGL(k_rpbc, dim3(grid_blocks(nent, BLOCK)), dim3(BLOCK), 0, ctx->stream, F,
                         rp, flag_nmn);
      rpp = rp;
    }
    int64_t nt = (int64_t)(g.n1 / 2 + 1) * (g.n2 / 2 + 1) * (g.n3 / 2 + 1);
    hipLaunchKernelGGL(k_restrict, dim3(grid_blocks(nt, BLOCK)), dim3(BLOCK), 0, ctx->stream, F,
                       C, rpp, flag_nmn);
  }

  void prolongation(int l) {  // L[l] += P(L[l+1])
    const KGrid &g = L[l].g;
    int64_t nt = (int64_t)(g.n1 / 2 + 1) * (g.n2 / 2 + 1) * (g.n3 / 2 + 1);
    hipLaunchKernelGGL(k_prolong, dim3(grid_blocks(nt, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       L[l], L[l + 1]);
  }

  static int64_t grid_nodes_k(const KGrid &g) {
    return ((int64_t)g.n1 + 3) * ((int64_t)g.n2 + 3) * ((int64_t)g.n3 + 3);
  }
};

void set_coeffs(KLevel &L, double eps0) {
  // mg.h:9-15
  const KGrid &g = L.g;
  L.aw = eps0 / (g.h1 * g.h1);
  L.ae = L.aw;
  L.an = eps0 / (g.h2 * g.h2);
  L.as = L.an;
  L.at = eps0 / (g.h3 * g.h3);
  L.ab = L.at;
  L.ac = -(L.an + L.as + L.aw + L.ae + L.at + L.ab);
}

}  // namespace

extern "C" int lepic_poisson_solve(lepic_ctx *ctx, const lepic_grid *g, double *phi,
                                   const double *rho, const int32_t *bcnd,
                                   const lepic_mg_params *mp, lepic_mg_stats *stats) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!g || !phi || !rho || !bcnd || !mp) return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson: null argument");
  const int ng = mp->ng;
  if (ng < 1 || ng > MAX_LEVELS) return lepic_fail(ctx, LEPIC_ERR_ARG, "poisson: ng outside [1,12]");
  Solver S;
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
    set_coeffs(S.L[l], mp->eps0);
    if (l > 0) coarse_nodes += grid_nodes(&gl);
    int64_t nres = (int64_t)(gl.n[0] + 1) * (gl.n[1] + 1) * (gl.n[2] + 1);
    max_parts = std::max(max_parts, (nres + BLOCK - 1) / BLOCK);
  }
  // coarse-level storage (kept across calls; labels zeroed once)
  const int64_t rp_len = ((int64_t)g->n[0] + 2) * ((int64_t)g->n[1] + 2);
  size_t need = sizeof(double) * (size_t)(2 * coarse_nodes + max_parts + rp_len) +
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
  int32_t *lb = reinterpret_cast<int32_t *>(lv + 2 * coarse_nodes);
  S.L[0].u = phi;
  S.L[0].b = const_cast<double *>(rho);
  S.L[0].bc = const_cast<int32_t *>(bcnd);
  for (int l = 1; l < ng; ++l) {
    int64_t nn = Solver::grid_nodes_k(S.L[l].g);
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
  // Levels lc..ng-1 of every V-cycle run inside k_mg_coarse.  The sequence of
  // operations is mg.f90's; the stretch from the down-leg sor_rb on lc to the
  // up-leg sor_rb on lc touches no level above lc and is issued as one launch.
  S.lc = ng;
  for (int l = 1; l < ng; ++l) {
    int64_t tot = 0;
    for (int m = l; m < ng; ++m) tot += Solver::grid_nodes_k(S.L[m].g);
    if (tot <= COARSE_LDS_NODES) {
      S.lc = l;
      break;
    }
  }
  // one V-cycle, mg (mg.f90:99-239)
  auto vcycle = [&](int kmg) -> int {
    for (int ig = 1; ig <= 2 * ng - 1; ++ig) {
      for (int il = 1; il <= ng - 1 && il <= 11; ++il) {
        const int l = il - 1;
        if (ig < ng && ig == il) {
          if (l < S.lc) {
            int st = S.sor_rb(l, epsa, kmg, ig - ng);
            if (st) return st;
            S.restriction(l);
          } else if (l == S.lc) {
            S.coarse(epsa, kmg);
          }
        }
        if (ig > ng && ig == 2 * ng - il) {
          if (l < S.lc) {
            int st = S.sor_rb(l, epsa, kmg, ig - ng);
            if (st) return st;
          }
          if (l <= S.lc && il >= 2) S.prolongation(il - 2);
        }
      }
      if (ig == ng) {
        if (S.lc >= ng) {
          int st = S.sor_rb(ng - 1, epsa, kmg, 0);
          if (st) return st;
        } else if (S.lc == ng - 1) {
          S.coarse(epsa, kmg);
        }
        if (ng >= 2 && ng - 2 < S.lc) S.prolongation(ng - 2);
      }
    }
    return LEPIC_OK;
  };
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
