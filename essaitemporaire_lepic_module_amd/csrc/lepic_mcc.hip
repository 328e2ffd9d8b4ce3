// lepic_mcc.hip -- null-collision Monte Carlo collisions (a10):
// perform_collisions_step (modules/particles/mod_collisions.f90:214-739).
//
// Two modes share one trial routine:
//  * REPLAY_RAN2 -- the reference's sequential Markov chain (trial order, ran2
//    consumption, n growing with appended products, immediate kills) run by a
//    single device thread, for exact parity with the reference;
//  * PHILOX -- every particle takes k ~ Poisson(Pmax) trials (the large-N
//    law of Nc = N Pmax picks with replacement), products appended with an
//    atomic slot counter.  Statistically the reference's process, fully
//    parallel, in two passes: a select pass walks each chunk of particles by
//    geometric gaps (the particles with k >= 1 form a Bernoulli(1 - e^-Pmax)
//    sequence) and lists (particle, k) for the live ones; the trial pass runs
//    one thread per listed particle, its draws from Philox-4x32-10 keyed by
//    (key, counter, species, particle).
#include <cmath>
#include <cstring>
#include <vector>

#include "lepic_internal.h"
#include "lepic_philox.h"

namespace {

constexpr int BLOCK = 256;
constexpr int MAXR = 16;   // reactions per projectile handled by the kernels
constexpr int MAXT = 4;    // distinct targets per projectile
constexpr int SEL_CHUNK = 64;   // particles per thread in the select pass
constexpr double QE_SI = 1.60217646e-19;
constexpr double PI_F = 3.14159265358979311600;

struct KParts {
  double *x[LEPIC_MAX_SPECIES], *y[LEPIC_MAX_SPECIES], *z[LEPIC_MAX_SPECIES];
  double *vx[LEPIC_MAX_SPECIES], *vy[LEPIC_MAX_SPECIES], *vz[LEPIC_MAX_SPECIES];
  int8_t *dead[LEPIC_MAX_SPECIES];
  int64_t cap[LEPIC_MAX_SPECIES];
};

// device counters in the control block (int64 words from CTRL_MCC)
struct MccCounters {
  int64_t n[LEPIC_MAX_SPECIES];          // live count per species
  int64_t accepted[LEPIC_MAX_SPECIES];
  int64_t err_coll[LEPIC_MAX_SPECIES];
  int64_t skipped_dead[LEPIC_MAX_SPECIES];
  int64_t appended[LEPIC_MAX_SPECIES];
  int64_t overflow;
  int64_t seed;
  int64_t nsel[LEPIC_MAX_SPECIES];       // select-pass list lengths
};

// Reaction plan of a projectile whose reactions all have neutral targets
// (at most FR reactions, FT distinct targets), built on the host from
// sig_list and col_info; ok == 0 sends the projectile to the general trial.
constexpr int FR = 8;
constexpr int FT = 2;
struct FastPlan {
  int32_t ok, nr, ntg;
  int32_t tgt[FT];    // target species of each slot, first-seen order
  int32_t col[FR];    // reactions in sig_list order
  int32_t slot[FR];   // target slot of each reaction
};

struct KChem {
  lepic_chem c;
  double np_mx_all[LEPIC_MAX_SPECIES];
  double nu_max_local[LEPIC_MAX_SPECIES];
  FastPlan fp[4];
  // density already deposited for this step ([ntype_tracked][nodes]; null:
  // none): products are added to it and killed particles taken out
  double *np_dep;
  const double *kq;
  double kdep[4];   // Nm / (h1 h2 h3) per tracked species
};

// Particle access of the register-only trial, one per storage layout.
// PA64: the FP64 sets (positions absolute; kills set dead = 1 and, when a
// deposit already ran, take the particle's FP64 deposit out of np_dep).
struct PA64 {
  KParts P;
  __device__ bool dead(int pt, int64_t ip) const { return P.dead[pt][ip] != 0; }
  __device__ void vel(int pt, int64_t ip, double &x, double &y, double &z) const {
    x = P.vx[pt][ip]; y = P.vy[pt][ip]; z = P.vz[pt][ip];
  }
  __device__ void set_vel(int pt, int64_t ip, double x, double y, double z) const {
    P.vx[pt][ip] = x; P.vy[pt][ip] = y; P.vz[pt][ip] = z;
  }
  __device__ int64_t cap(int t) const { return P.cap[t]; }
  __device__ void append(const KGrid &g, const KChem &K, int64_t nodes, int bt, int64_t ib,
                         int pt, int64_t ip, double vx, double vy, double vz) const {
    const double x0 = P.x[pt][ip], y0 = P.y[pt][ip], z0 = P.z[pt][ip];
    P.x[bt][ib] = x0; P.y[bt][ib] = y0; P.z[bt][ib] = z0;
    P.vx[bt][ib] = vx; P.vy[bt][ib] = vy; P.vz[bt][ib] = vz;
    P.dead[bt][ib] = 0;
    if (K.np_dep) deposit_one(g, K.np_dep + (int64_t)bt * nodes, K.kq, K.kdep[bt], x0, y0, z0, 1.0);
  }
  __device__ void kill(const KGrid &g, const KChem &K, int64_t nodes, int pt, int64_t ip) const {
    if (K.np_dep)
      deposit_one(g, K.np_dep + (int64_t)pt * nodes, K.kq, K.kdep[pt], P.x[pt][ip], P.y[pt][ip],
                  P.z[pt][ip], -1.0);
    P.dead[pt][ip] = 1;
  }
};

// PA32: the FP32 sets (cell id + in-cell offsets; cell = -1 is no particle).
// Products take the parent's cell and offsets; with dep set, products add
// and kills remove their fixed-point deposit (the split of the deposit
// kernels) in the int64 node sums of their species, by device-scope integer
// atomics -- exact, whatever the order.
struct KParts32 {
  int32_t *cell[4];
  float *fx[4], *fy[4], *fz[4], *vx[4], *vy[4], *vz[4];
  int64_t cap[4];
  long long *dep[4];
  double U[4];
  int neg[4];
};

struct PA32 {
  KParts32 P;
  __device__ bool dead(int pt, int64_t ip) const { return P.cell[pt][ip] < 0; }
  __device__ void vel(int pt, int64_t ip, double &x, double &y, double &z) const {
    x = (double)P.vx[pt][ip]; y = (double)P.vy[pt][ip]; z = (double)P.vz[pt][ip];
  }
  __device__ void set_vel(int pt, int64_t ip, double x, double y, double z) const {
    P.vx[pt][ip] = (float)x; P.vy[pt][ip] = (float)y; P.vz[pt][ip] = (float)z;
  }
  __device__ int64_t cap(int t) const { return P.cap[t]; }
  __device__ void dep(const KGrid &g, int t, int c, float fx, float fy, float fz, int sign) const {
    if (!P.dep[t] || c < 0 || (int64_t)c >= (int64_t)g.n1 * g.n2 * g.n3) return;
    const int cx = c % g.n1, r = c / g.n1, cy = r % g.n2, cz = r / g.n2;
    const int64_t b = gidx(g, cx + 1, cy + 1, cz + 1);
    const int64_t nd[8] = {b, b + 1, b + g.sy + 1, b + g.sy, b + g.sz, b + g.sz + 1,
                           b + g.sz + g.sy + 1, b + g.sz + g.sy};
    int a[8];
    fixed_split32((float)P.U[t], fx, fy, fz, a);
    const bool flip = (sign < 0) != (P.neg[t] != 0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (a[q] != 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(&P.dep[t][nd[q]]),
                  (unsigned long long)(long long)(flip ? -a[q] : a[q]));
  }
  __device__ void append(const KGrid &g, const KChem &, int64_t, int bt, int64_t ib, int pt,
                         int64_t ip, double vx, double vy, double vz) const {
    const int c = P.cell[pt][ip];
    const float fx = P.fx[pt][ip], fy = P.fy[pt][ip], fz = P.fz[pt][ip];
    P.cell[bt][ib] = c;
    P.fx[bt][ib] = fx; P.fy[bt][ib] = fy; P.fz[bt][ib] = fz;
    P.vx[bt][ib] = (float)vx; P.vy[bt][ib] = (float)vy; P.vz[bt][ib] = (float)vz;
    dep(g, bt, c, fx, fy, fz, 1);
  }
  __device__ void kill(const KGrid &g, const KChem &, int64_t, int pt, int64_t ip) const {
    dep(g, pt, P.cell[pt][ip], P.fx[pt][ip], P.fy[pt][ip], P.fz[pt][ip], -1);
    P.cell[pt][ip] = -1;
  }
};

struct Ran2Stream {
  int32_t s;
  __device__ double uniform() {
    const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
    const double am = 1.0 / (double)im;
    int32_t k = s / iq;
    s = ia * (s - k * iq) - ir * k;
    if (s < 0) s += im;
    return am * (double)s;
  }
};

__device__ inline void box_muller(double &v1, double &v2, double vt, double r1_in, double r2_in) {
  // mod_collisions.f90:742-753
  double r1 = r1_in > 1.0e-14 ? r1_in : 1.0e-14;
  double fac = vt * sqrt(-log(1.0 - r1));
  v1 = fac * cos(2.0 * PI_F * r2_in);
  v2 = fac * sin(2.0 * PI_F * r2_in);
}

__device__ inline void scatter_vector(double &vx1, double &vy1, double &vz1, double vx, double vy,
                                      double vz, double costheta, double phi) {
  // mod_collisions.f90:64-92
  double s = 1.0 - costheta * costheta;
  double sintheta = sqrt(s > 0.0 ? s : 0.0);
  double sinphi = sin(phi), cosphi = cos(phi);
  double v = sqrt(vx * vx + vy * vy + vz * vz);
  vx1 = vx; vy1 = vy; vz1 = vz;
  if (v == 0.0) return;
  if (fabs(vy) > fabs(vz)) {
    double vv = sqrt(vx * vx + vy * vy);
    vx1 = vx * costheta + (vy * v * sinphi + vx * vz * cosphi) / vv * sintheta;
    vy1 = vy * costheta + (-vx * v * sinphi + vy * vz * cosphi) / vv * sintheta;
    vz1 = vz * costheta - vv * cosphi * sintheta;
  } else {
    double vv = sqrt(vx * vx + vz * vz);
    vx1 = vx * costheta + (vz * v * sinphi - vy * vx * cosphi) / vv * sintheta;
    vy1 = vy * costheta + vv * cosphi * sintheta;
    vz1 = vz * costheta - (vx * v * sinphi + vy * vz * cosphi) / vv * sintheta;
  }
}

// First bracket Er[ip-1] < E <= Er[ip] (mod_collisions.f90:512-546 returns
// this unique bracket on a strictly increasing grid).
__device__ inline int find_bracket(const double *Er, int npt, double E) {
  if (npt < 2) return -1;
  if (!(E > Er[0]) || E > Er[npt - 1]) return -1;
  int lo = 0, hi = npt - 1;
  while (hi - lo > 1) {
    int mid = lo + (hi - lo) / 2;
    if (E <= Er[mid]) hi = mid;
    else lo = mid;
  }
  return hi;
}

// Appending a product: the replay chain bumps n in order, the parallel mode
// takes an atomic slot.
struct SerialAppend {
  MccCounters *cnt;
  __device__ int64_t slot(int bt) { return cnt->n[bt]; }
  __device__ void commit(int bt, int64_t ib) {
    if (ib >= cnt->n[bt]) cnt->n[bt] = ib + 1;
  }
  __device__ void add(double *a, double v) { *a = *a + v; }
  __device__ void count(int64_t *a) { *a += 1; }
};

// Slots for appended products: the active lanes of the calling wave, grouped
// by the species they append to, take consecutive values of ctr[species], one
// global atomic per species present in the wave.
__device__ inline int64_t wave_slot(unsigned long long *ctr, int sp) {
  unsigned long long todo = __ballot(1);
  int64_t mine = 0;
  while (todo) {
    const int leader = __ffsll(todo) - 1;
    const int sl = __shfl(sp, leader);
    const unsigned long long grp = __ballot(sp == sl) & todo;
    unsigned long long b = 0;
    if (__lane_id() == leader) b = atomicAdd(&ctr[sl], (unsigned long long)__popcll(grp));
    b = __shfl(b, leader);
    if (sp == sl) mine = (int64_t)(b + (unsigned long long)__popcll(grp & ((1ull << __lane_id()) - 1ull)));
    todo &= ~grp;
  }
  return mine;
}

// Event counters of the trial pass live in LDS (one uint32 per int64 word of
// MccCounters) and reach the device counters once per block.
constexpr int CNT_WORDS = (int)(sizeof(MccCounters) / sizeof(int64_t));

struct BlockAppend {
  MccCounters *cnt;
  unsigned *lds;     // CNT_WORDS block counters
  double pcoll[4];   // tracked species (at most 4); summed per block at the end
  double *pcoll_base;
  __device__ int64_t slot(int bt) {
    return wave_slot(reinterpret_cast<unsigned long long *>(cnt->n), bt);
  }
  __device__ void commit(int, int64_t) {}
  __device__ void add(double *a, double v) {
    const int t = (int)(a - pcoll_base);
#pragma unroll
    for (int i = 0; i < 4; ++i)   // branch-free, static indices: pcoll stays in registers
      pcoll[i] += i == t ? v : 0.0;
  }
  // active lanes may count different words: one LDS atomic per distinct word
  __device__ void count(int64_t *a) {
    const int w = (int)(a - reinterpret_cast<int64_t *>(cnt));
    unsigned long long todo = __ballot(1);
    while (todo) {
      const int leader = __ffsll(todo) - 1;
      const int wl = __shfl(w, leader);
      const unsigned long long grp = __ballot(w == wl) & todo;
      if (__lane_id() == leader) atomicAdd(&lds[wl], (unsigned)__popcll(grp));
      todo &= ~grp;
    }
  }
};

// Exclusive offset of this thread's `count` items in a block-wide run taken
// from *ctr with one atomic per block.  Every thread of the block calls it.
__device__ inline int64_t block_reserve(int count, unsigned long long *ctr) {
  __shared__ int wsum[BLOCK / 64];
  __shared__ unsigned long long base;
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  int v = count;
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d);
    if (lane >= d) v += o;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int i = 0; i < BLOCK / 64; ++i) {
      const int s = wsum[i];
      wsum[i] = tot;
      tot += s;
    }
    base = tot ? atomicAdd(ctr, (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  return (int64_t)(base + (unsigned long long)(wsum[w] + v - count));
}

// One null-collision trial on particle ip of species pt
// (mod_collisions.f90:406-726).
template <class RNG, class APP>
__device__ void trial(const KGrid &g, const KParts &P, const KChem &K, const double *np_red,
                      int64_t nodes, int pt, int64_t ip, RNG &rng, APP &app, MccCounters *cnt,
                      double *Pcoll) {
  const lepic_chem &c = K.c;
  const int NT = c.ntype_tracked, NA = c.ntype_total;
  if (P.dead[pt][ip] != 0) {
    app.count(&cnt->skipped_dead[pt]);
    return;
  }
  int ix = (int)qdiv(P.x[pt][ip], g.q1) + 1;
  int iy = (int)qdiv(P.y[pt][ip], g.q2) + 1;
  int iz = (int)qdiv(P.z[pt][ip], g.q3) + 1;
  ix = imax(1, imin(g.n1, ix));
  iy = imax(1, imin(g.n2, iy));
  iz = imax(1, imin(g.n3, iz));
  const double vx1 = P.vx[pt][ip], vy1 = P.vy[pt][ip], vz1 = P.vz[pt][ip];
  const int ncl = c.p_ncol[pt] < MAXR ? c.p_ncol[pt] : MAXR;
  double nu[MAXR];
  double sum_nu = 0.0;
  unsigned tsel = 0, tfound = 0;
  // relative-motion data of each sampled neutral target, in first-seen slots
  int tslot[MAXT];
  int ntg = 0;
  double svr[MAXT], sEkr[MAXT], smu[MAXT], scx[MAXT], scy[MAXT], scz[MAXT];
  auto slot_of = [&](int tt) {
    int s = 0;
    while (s < ntg - 1 && tslot[s] != tt) ++s;
    return s;
  };
  for (int a = 0; a < ncl; ++a) nu[a] = 0.0;
  for (int ic2 = 0; ic2 < ncl; ++ic2) {
    const int col = c.sig_list[pt * LEPIC_MAX_COL + ic2];
    const int n_re = c.col_info[col * 10 + 0];
    const int tt = c.col_info[col * 10 + 1 + n_re];
    if (tt < 0 || tt >= NA) continue;
    if (!(tsel & (1u << tt))) {
      tsel |= 1u << tt;
      if (tt >= NT) {
        double vx2, vy2, vz2, dummy;
        double r1 = rng.uniform(), r2 = rng.uniform();
        box_muller(vx2, vy2, c.vt0[tt], r1, r2);
        r1 = rng.uniform(); r2 = rng.uniform();
        box_muller(vz2, dummy, c.vt0[tt], r1, r2);
        const double mp = fabs(c.mass[pt]), mt = fabs(c.mass[tt]);
        const int sl = ntg < MAXT ? ntg : MAXT - 1;   // the host caps targets at MAXT
        tslot[sl] = tt;
        ntg = sl + 1;
        smu[sl] = mp * mt / (mp + mt);
        const double dx = vx1 - vx2, dy = vy1 - vy2, dz = vz1 - vz2;
        svr[sl] = sqrt(dx * dx + dy * dy + dz * dz);
        sEkr[sl] = 0.5 * smu[sl] * (svr[sl] * svr[sl]) / QE_SI;
        scx[sl] = (mp * vx1 + mt * vx2) / (mp + mt);
        scy[sl] = (mp * vy1 + mt * vy2) / (mp + mt);
        scz[sl] = (mp * vz1 + mt * vz2) / (mp + mt);
        tfound |= 1u << tt;
      } else {
        // charged target: the cell tables are cleared before MCC runs
        // (mod_particleMover.f90:153-155, mod_particleBC.f90:255-257), so the
        // neighbour-cell pick always fails (mod_collisions.f90:473-478)
        app.count(&cnt->err_coll[pt]);
      }
    }
    if (!(tfound & (1u << tt))) continue;
    const int sl = slot_of(tt);
    const double vr = svr[sl], Ekr = sEkr[sl];
    const int ipt = find_bracket(c.sig_Er, c.npt, Ekr);
    if (ipt > 0) {
      const double EkL = c.sig_Er[ipt - 1], EkR = c.sig_Er[ipt];
      const double sL = c.sig[(int64_t)col * c.npt + ipt - 1];
      const double sR = c.sig[(int64_t)col * c.npt + ipt];
      const double sig_p = sL + (Ekr - EkL) * (sR - sL) / (EkR - EkL);
      double np_t;
      if (tt < NT) {
        // density_average_at_particle_cell on a 1-based dummy: nodes ix-1..ix
        const double *d = np_red + (int64_t)tt * nodes;
        np_t = 0.125 * (d[gidx(g, ix - 1, iy - 1, iz - 1)] + d[gidx(g, ix, iy - 1, iz - 1)] +
                        d[gidx(g, ix - 1, iy, iz - 1)] + d[gidx(g, ix, iy, iz - 1)] +
                        d[gidx(g, ix - 1, iy - 1, iz)] + d[gidx(g, ix, iy - 1, iz)] +
                        d[gidx(g, ix - 1, iy, iz)] + d[gidx(g, ix, iy, iz)]);
      } else {
        np_t = K.np_mx_all[tt];
      }
      nu[ic2] = np_t * sig_p * vr / K.nu_max_local[pt];
      sum_nu = sum_nu + nu[ic2];
    }
  }
  // indexx_small: ascending exchange sort (mod_collisions.f90:756-776)
  int indx[MAXR];
  for (int a = 0; a < ncl; ++a) indx[a] = a;
  for (int a = 0; a < ncl - 1; ++a)
    for (int b = a + 1; b < ncl; ++b)
      if (nu[indx[b]] < nu[indx[a]]) {
        int t = indx[a]; indx[a] = indx[b]; indx[b] = t;
      }
  const double rnd = rng.uniform();
  int c_ind = -1;
  if (rnd <= sum_nu) {
    double s = 0.0;
    for (int a = 0; a < ncl; ++a) {
      s = s + nu[indx[a]];
      if (rnd <= s) {
        c_ind = c.sig_list[pt * LEPIC_MAX_COL + indx[a]];
        break;
      }
    }
  }
  if (c_ind < 0) return;
  app.count(&cnt->accepted[pt]);
  const int n_re = c.col_info[c_ind * 10 + 0];
  const int n_by = c.col_info[c_ind * 10 + 1];
  const int tt = c.col_info[c_ind * 10 + 1 + n_re];
  if (tt < 0 || tt >= NA) return;
  if (!(tfound & (1u << tt))) return;
  const int sl = slot_of(tt);
  const double mu = smu[sl], vr = svr[sl];
  const double vxc = scx[sl], vyc = scy[sl], vzc = scz[sl];
  const double Eth = c.sig_Eth[c_ind];
  const double Ee = 0.5 * mu * vr * vr - Eth * QE_SI;
  if (Ee < 0.0) return;
  double smi = 0.0;
  for (int ib = 0; ib < n_by; ++ib) {
    const int bt = c.col_info[c_ind * 10 + 2 + n_re + ib];
    if (bt >= 0 && bt < NA && c.mass[bt] > 0.0) smi = smi + 1.0 / fabs(c.mass[bt]);
  }
  if (smi <= 0.0) return;
  const double costh = 1.0 - 2.0 * rng.uniform();
  const double phi = 2.0 * PI_F * rng.uniform();
  const double costh_s = 1.0 - 2.0 * rng.uniform();
  const double phi_s = 2.0 * PI_F * rng.uniform();
  const double th_add = 2.0 * PI_F / (double)(n_by > 1 ? n_by : 1);
  const double x0 = P.x[pt][ip], y0 = P.y[pt][ip], z0 = P.z[pt][ip];
  int flag_add[10];
  for (int a = 0; a < 10; ++a) flag_add[a] = a < n_re ? -1 : 0;
  for (int ib = 0; ib < n_by; ++ib) {
    const int bt = c.col_info[c_ind * 10 + 2 + n_re + ib];
    if (bt < 0 || bt >= NT) continue;
    if (c.mass[bt] <= 0.0) continue;
    const double s2 = 1.0 - costh * costh;
    const double st = sqrt(0.0 > s2 ? 0.0 : s2);
    const double ex = costh;
    const double ey = st * sin(phi + th_add * (double)ib);
    const double ez = st * cos(phi + th_add * (double)ib);
    double ex1, ey1, ez1;
    scatter_vector(ex1, ey1, ez1, ex, ey, ez, costh_s, phi_s);
    const double mb = fabs(c.mass[bt]);
    const double vp = sqrt(2.0 * (Ee / smi) / (mb * mb));
    int64_t ibx;
    bool fresh;
    if (ib == 0 && bt == pt) {
      ibx = ip;
      flag_add[0] = 0;
      fresh = false;
    } else if (bt == tt && tt < NT) {
      flag_add[1] = 0;
      continue;  // target index of a failed charged pick (never reached)
    } else {
      ibx = app.slot(bt);
      flag_add[n_re + ib] = 1;
      fresh = true;
    }
    if (ibx >= P.cap[bt]) {
      atomicExch((unsigned long long *)&cnt->overflow, 1ull);
      continue;
    }
    double v2old = 0.0;
    if (!fresh)
      v2old = P.vx[bt][ibx] * P.vx[bt][ibx] + P.vy[bt][ibx] * P.vy[bt][ibx] +
              P.vz[bt][ibx] * P.vz[bt][ibx];
    app.commit(bt, ibx);
    P.x[bt][ibx] = x0; P.y[bt][ibx] = y0; P.z[bt][ibx] = z0;
    const double nvx = vxc + vp * ex1, nvy = vyc + vp * ey1, nvz = vzc + vp * ez1;
    P.vx[bt][ibx] = nvx; P.vy[bt][ibx] = nvy; P.vz[bt][ibx] = nvz;
    P.dead[bt][ibx] = 0;
    if (fresh) {
      app.count(&cnt->appended[bt]);
      if (K.np_dep) deposit_one(g, K.np_dep + (int64_t)bt * nodes, K.kq, K.kdep[bt], x0, y0, z0, 1.0);
    }
    const double v2new = nvx * nvx + nvy * nvy + nvz * nvz;
    app.add(&Pcoll[bt], 0.5 * c.Nm[bt] * c.mass[bt] * (v2new - v2old));
  }
  for (int ir = 0; ir < n_re; ++ir) {
    const int ct = c.col_info[c_ind * 10 + 2 + ir];
    if (ct < 0 || ct >= NT) continue;
    if (c.mass[ct] <= 0.0) continue;
    if (flag_add[ir] == -1 && ct == pt && P.dead[pt][ip] == 0) {
      if (K.np_dep) deposit_one(g, K.np_dep + (int64_t)pt * nodes, K.kq, K.kdep[pt], x0, y0, z0, -1.0);
      P.dead[pt][ip] = 1;
    }
  }
}

// The trial above for a projectile with a FastPlan, PHILOX mode only.  Same
// process, arranged for registers: each neutral target is drawn once (first-
// seen order, 4 draws) and its cross-section bracket found once; the reaction
// is the first, in ascending nu (ties in sig_list order), whose running sum
// reaches rnd -- the exchange sort's choice -- found by ranking instead of
// sorting; rnd is drawn after the targets, as above.
template <class PA, class APP>
__device__ void trial_fast(const KGrid &g, const PA &P, const KChem &K, int64_t nodes, int pt,
                           int64_t ip, PhiloxStream &rng, APP &app, MccCounters *cnt) {
  const lepic_chem &c = K.c;
  const FastPlan &fp = K.fp[pt];
  const int NT = c.ntype_tracked, NA = c.ntype_total;
  if (P.dead(pt, ip)) {
    app.count(&cnt->skipped_dead[pt]);
    return;
  }
  double vx1, vy1, vz1;
  P.vel(pt, ip, vx1, vy1, vz1);
  const double mp = fabs(c.mass[pt]);
  double tvx[FT], tvy[FT], tvz[FT], tvr[FT], tEk[FT];
  int tip[FT];
#pragma unroll
  for (int sl = 0; sl < FT; ++sl) {
    tvx[sl] = tvy[sl] = tvz[sl] = tvr[sl] = tEk[sl] = 0.0;
    tip[sl] = -1;
    if (sl < fp.ntg) {
      const int tt = fp.tgt[sl];
      double vx2, vy2, vz2, dummy;
      double r1 = rng.uniform(), r2 = rng.uniform();
      box_muller(vx2, vy2, c.vt0[tt], r1, r2);
      r1 = rng.uniform(); r2 = rng.uniform();
      box_muller(vz2, dummy, c.vt0[tt], r1, r2);
      const double mt = fabs(c.mass[tt]);
      const double mu = mp * mt / (mp + mt);
      const double dx = vx1 - vx2, dy = vy1 - vy2, dz = vz1 - vz2;
      const double vr = sqrt(dx * dx + dy * dy + dz * dz);
      tvx[sl] = vx2; tvy[sl] = vy2; tvz[sl] = vz2;
      tvr[sl] = vr;
      tEk[sl] = 0.5 * mu * (vr * vr) / QE_SI;
      tip[sl] = find_bracket(c.sig_Er, c.npt, tEk[sl]);
    }
  }
  double nu[FR];
  double sum_nu = 0.0;
#pragma unroll
  for (int r = 0; r < FR; ++r) {
    nu[r] = 0.0;
    if (r < fp.nr) {
      const bool s1 = FT > 1 && fp.slot[r] == 1;
      const int ipt = s1 ? tip[1] : tip[0];
      if (ipt > 0) {
        const int col = fp.col[r];
        const double Ekr = s1 ? tEk[1] : tEk[0];
        const double vr = s1 ? tvr[1] : tvr[0];
        const double EkL = c.sig_Er[ipt - 1], EkR = c.sig_Er[ipt];
        const double sL = c.sig[(int64_t)col * c.npt + ipt - 1];
        const double sR = c.sig[(int64_t)col * c.npt + ipt];
        const double sig_p = sL + (Ekr - EkL) * (sR - sL) / (EkR - EkL);
        nu[r] = K.np_mx_all[fp.tgt[s1 ? 1 : 0]] * sig_p * vr / K.nu_max_local[pt];
        sum_nu = sum_nu + nu[r];
      }
    }
  }
  const double rnd = rng.uniform();
  if (!(rnd <= sum_nu)) return;
  int pick = -1;
  double best = INFINITY;
#pragma unroll
  for (int a = 0; a < FR; ++a) {
    if (a < fp.nr) {
      double run = nu[a];
#pragma unroll
      for (int b = 0; b < FR; ++b)
        if (b < fp.nr && b != a && (nu[b] < nu[a] || (nu[b] == nu[a] && b < a))) run = run + nu[b];
      if (rnd <= run && run < best) {
        best = run;
        pick = a;
      }
    }
  }
  if (pick < 0) return;
  app.count(&cnt->accepted[pt]);
  const int c_ind = fp.col[pick];
  const bool s1 = FT > 1 && fp.slot[pick] == 1;
  const int tt = fp.tgt[s1 ? 1 : 0];
  const double vx2 = s1 ? tvx[1] : tvx[0], vy2 = s1 ? tvy[1] : tvy[0], vz2 = s1 ? tvz[1] : tvz[0];
  const double vr = s1 ? tvr[1] : tvr[0];
  const double mt = fabs(c.mass[tt]);
  const double mu = mp * mt / (mp + mt);
  const double vxc = (mp * vx1 + mt * vx2) / (mp + mt);
  const double vyc = (mp * vy1 + mt * vy2) / (mp + mt);
  const double vzc = (mp * vz1 + mt * vz2) / (mp + mt);
  const int n_re = c.col_info[c_ind * 10 + 0];
  const int n_by = c.col_info[c_ind * 10 + 1];
  const double Eth = c.sig_Eth[c_ind];
  const double Ee = 0.5 * mu * vr * vr - Eth * QE_SI;
  if (Ee < 0.0) return;
  double smi = 0.0;
  for (int ib = 0; ib < n_by; ++ib) {
    const int bt = c.col_info[c_ind * 10 + 2 + n_re + ib];
    if (bt >= 0 && bt < NA && c.mass[bt] > 0.0) smi = smi + 1.0 / fabs(c.mass[bt]);
  }
  if (smi <= 0.0) return;
  const double costh = 1.0 - 2.0 * rng.uniform();
  const double phi = 2.0 * PI_F * rng.uniform();
  const double costh_s = 1.0 - 2.0 * rng.uniform();
  const double phi_s = 2.0 * PI_F * rng.uniform();
  const double th_add = 2.0 * PI_F / (double)(n_by > 1 ? n_by : 1);
  bool reused = false;
  for (int ib = 0; ib < n_by; ++ib) {
    const int bt = c.col_info[c_ind * 10 + 2 + n_re + ib];
    if (bt < 0 || bt >= NT) continue;
    if (c.mass[bt] <= 0.0) continue;
    const double s2 = 1.0 - costh * costh;
    const double st = sqrt(0.0 > s2 ? 0.0 : s2);
    const double ex = costh;
    const double ey = st * sin(phi + th_add * (double)ib);
    const double ez = st * cos(phi + th_add * (double)ib);
    double ex1, ey1, ez1;
    scatter_vector(ex1, ey1, ez1, ex, ey, ez, costh_s, phi_s);
    const double mb = fabs(c.mass[bt]);
    const double vp = sqrt(2.0 * (Ee / smi) / (mb * mb));
    const double nvx = vxc + vp * ex1, nvy = vyc + vp * ey1, nvz = vzc + vp * ez1;
    const double v2new = nvx * nvx + nvy * nvy + nvz * nvz;
    if (ib == 0 && bt == pt) {
      // the projectile itself carries the first same-species byproduct
      reused = true;
      P.set_vel(pt, ip, nvx, nvy, nvz);
      const double v2old = vx1 * vx1 + vy1 * vy1 + vz1 * vz1;
      app.add(&app.pcoll_base[bt], 0.5 * c.Nm[bt] * c.mass[bt] * (v2new - v2old));
      continue;
    }
    const int64_t ibx = app.slot(bt);
    if (ibx >= P.cap(bt)) {
      atomicExch((unsigned long long *)&cnt->overflow, 1ull);
      continue;
    }
    P.append(g, K, nodes, bt, ibx, pt, ip, nvx, nvy, nvz);
    app.count(&cnt->appended[bt]);
    app.add(&app.pcoll_base[bt], 0.5 * c.Nm[bt] * c.mass[bt] * v2new);
  }
  for (int ir = 0; ir < n_re; ++ir) {
    const int ct = c.col_info[c_ind * 10 + 2 + ir];
    if (ct < 0 || ct >= NT) continue;
    if (c.mass[ct] <= 0.0) continue;
    if (ct == pt && !(ir == 0 && reused) && !P.dead(pt, ip)) P.kill(g, K, nodes, pt, ip);
  }
}

// REPLAY_RAN2: the whole chain on one device thread (trial order and ran2
// consumption of mod_collisions.f90:391-728 for one virtual thread).
__global__ void k_mcc_replay(KGrid g, KParts P, KChem K, const double *np_red, int64_t nodes,
                             int32_t Nc0, int32_t Nc1, int32_t Nc2, int32_t Nc3, MccCounters *cnt,
                             double *Pcoll) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  Ran2Stream rng;
  rng.s = (int32_t)cnt->seed;
  SerialAppend app{cnt};
  const int32_t Ncs[4] = {Nc0, Nc1, Nc2, Nc3};
  const lepic_chem &c = K.c;
  for (int pt = 0; pt < c.ntype_tracked && pt < 4; ++pt) {
    if (c.p_ncol[pt] == 0 || c.mass[pt] <= 0.0) continue;
    if (Ncs[pt] == 0) continue;
    if (cnt->n[pt] == 0) continue;
    if (K.nu_max_local[pt] <= 0.0) continue;
    for (int32_t ic = 0; ic < Ncs[pt]; ++ic) {
      const double rnd = rng.uniform();
      const int64_t n = cnt->n[pt];
      int64_t ip = (int64_t)((double)n * rnd) + 1;
      if (ip > n) ip = n;
      trial(g, P, K, np_red, nodes, pt, ip - 1, rng, app, cnt, Pcoll);
    }
  }
  cnt->seed = rng.s;
}

// PHILOX select pass: thread t walks particles [t*SEL_CHUNK, (t+1)*SEL_CHUNK)
// of species pt.  Particles with k >= 1 trials form a Bernoulli(p) sequence,
// p = 1 - e^-lambda, so the gap to the next one is G = floor(-ln(1-u)/lambda)
// (geometric, since ln(1-p) = -lambda); k given k >= 1 is zero-truncated
// Poisson(lambda), drawn by inversion from P(1) = lambda / (e^lambda - 1).
// Live particles are listed as (index | k << 40); holes and MCC-killed
// particles take no trials.
struct LiveDead {   // FP64 sets: dead flags
  const int8_t *d;
  __device__ bool operator()(int64_t j) const { return d[j] == 0; }
};
struct LiveCell {   // FP32 sets: cell id >= 0
  const int32_t *c;
  __device__ bool operator()(int64_t j) const { return c[j] >= 0; }
};

template <class LIVE>
__global__ __launch_bounds__(BLOCK) void k_mcc_select(LIVE live, int pt, int64_t n0,
                                                      double lambda, double p1, uint64_t key,
                                                      uint64_t counter, uint64_t *list,
                                                      int64_t cap, unsigned long long *nsel) {
  const int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  const int64_t begin = t * SEL_CHUNK;
  const int64_t end = begin + SEL_CHUNK < n0 ? begin + SEL_CHUNK : n0;
  PhiloxStream rng0;
  rng0.init(key ^ ((uint64_t)pt << 56) ^ 0x5E1EC7A11ull, (uint32_t)t, (uint32_t)(t >> 32),
            (uint32_t)counter);
  rng0.c3 ^= (uint32_t)(counter >> 32) * 0x9E3779B9u;
  // two walks over the same draws: count the listed particles, take the
  // block's run of the list, then write them in particle order
  int64_t slot = 0;
  for (int pass = 0; pass < 2; ++pass) {
    PhiloxStream rng = rng0;
    int count = 0;
    int64_t j = begin - 1;
    while (begin < end) {
      const double gap = -log(1.0 - rng.uniform()) / lambda;   // u in [0,1): finite, >= 0
      if (gap >= (double)(end - 1 - j)) break;
      j += 1 + (int64_t)gap;
      const double u = rng.uniform();
      double pk = p1, cdf = p1;
      int k = 1;
      while (u > cdf && k < 32) {
        ++k;
        pk = pk * lambda / (double)k;
        cdf = cdf + pk;
      }
      if (!live(j)) continue;
      if (pass == 1 && slot < cap) list[slot] = (uint64_t)j | ((uint64_t)k << 40);
      if (pass == 1) ++slot;
      ++count;
    }
    if (pass == 0) slot = block_reserve(count, nsel);
  }
}

// PHILOX trial pass: one thread per listed particle, k trials in a row on the
// particle's own stream.
template <bool FAST, class PA>
__global__ __launch_bounds__(BLOCK) void k_mcc_trials(KGrid g, PA P, KChem K,
                                                      const double *np_red, int64_t nodes, int pt,
                                                      const uint64_t *list,
                                                      const unsigned long long *nsel, int64_t cap,
                                                      uint64_t key, uint64_t counter,
                                                      MccCounters *cnt, double *Pcoll) {
  __shared__ unsigned lcnt[CNT_WORDS];
  __shared__ double lp[BLOCK / 64][4];
  for (int i = threadIdx.x; i < CNT_WORDS; i += BLOCK) lcnt[i] = 0u;
  __syncthreads();
  const int64_t e = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  BlockAppend app;
  app.cnt = cnt;
  app.lds = lcnt;
  app.pcoll_base = Pcoll;
  for (int t = 0; t < 4; ++t) app.pcoll[t] = 0.0;
  const int64_t m = (int64_t)*nsel < cap ? (int64_t)*nsel : cap;
  if (e < m) {
    const uint64_t w = list[e];
    const int64_t ip = (int64_t)(w & ((1ull << 40) - 1));
    const int k = (int)(w >> 40);
    PhiloxStream rng;
    rng.init(key ^ ((uint64_t)pt << 56), (uint32_t)ip, (uint32_t)(ip >> 32), (uint32_t)counter);
    rng.c3 ^= (uint32_t)(counter >> 32) * 0x9E3779B9u;
    for (int t = 0; t < k; ++t) {
      if constexpr (FAST) trial_fast(g, P, K, nodes, pt, ip, rng, app, cnt);
      else trial(g, P.P, K, np_red, nodes, pt, ip, rng, app, cnt, Pcoll);
    }
  }
  // Pcoll: wave sums, then one atomic per block and species
  const int NT = K.c.ntype_tracked;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (t >= NT) break;
    double v = app.pcoll[t];
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    if (__lane_id() == 0) lp[threadIdx.x >> 6][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < NT) {
    double v = 0.0;
    for (int i = 0; i < BLOCK / 64; ++i) v += lp[i][threadIdx.x];
    if (v != 0.0) atomicAdd(&Pcoll[threadIdx.x], v);
  }
  for (int i = threadIdx.x; i < CNT_WORDS; i += BLOCK)
    if (lcnt[i] != 0u)
      atomicAdd(reinterpret_cast<unsigned long long *>(cnt) + i, (unsigned long long)lcnt[i]);
}

__global__ void k_fill_u64(uint64_t *a, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// np_mx of the reference: max over stored nodes 0..n of each dimension
// (maxval(np_red(1:n+1,...)) through a 1-based dummy, mod_collisions.f90:293).
__global__ __launch_bounds__(BLOCK) void k_npmax(KGrid g, const double *np, int64_t nodes,
                                                 double *partials) {
  __shared__ double red[BLOCK];
  int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  int64_t nx = g.n1 + 1, ny = g.n2 + 1, nz = g.n3 + 1;
  double m = -INFINITY;
  if (t < nx * ny * nz) {
    int i = (int)(t % nx);
    int64_t r = t / nx;
    int j = (int)(r % ny), k = (int)(r / ny);
    m = np[gidx(g, i, j, k)];
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

}  // namespace

// The particle sets of one MCC call: FP64 sets (parts) or FP32 sets
// (parts32, PHILOX mode, neutral targets only), with the FP32 sets' int64
// deposit sums (dep, [ntype_tracked][nodes], may be null) and their fraction
// bits.
struct MccSets {
  lepic_pset *parts = nullptr;
  lepic_pset32 *parts32 = nullptr;
  int64_t *dep = nullptr;
  const int32_t *frac_bits = nullptr;
  int64_t n(int t) const { return parts ? parts[t].n : parts32[t].n; }
  int64_t cap(int t) const { return parts ? parts[t].cap : parts32[t].cap; }
  void set_n(int t, int64_t v) {
    if (parts) parts[t].n = v;
    else parts32[t].n = v;
  }
};

static int collisions_impl(lepic_ctx *ctx, const lepic_grid *g, MccSets sets,
                           const double *np_red, const lepic_chem *chem, int32_t ns_coll,
                           double dt, int32_t mode, int32_t *iseed, uint64_t philox_key,
                           uint64_t philox_counter, double *Pcoll, lepic_mcc_stats *stats,
                           double *np_dep, const double *kq) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!g || (!sets.parts && !sets.parts32) || !np_red || !chem || !Pcoll ||
      (mode == LEPIC_MCC_REPLAY_RAN2 && !iseed))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: null argument");
  const bool fp32 = sets.parts32 != nullptr;
  if (fp32 && mode != LEPIC_MCC_PHILOX)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions32: FP32 sets run the PHILOX mode only");
  const lepic_chem &c = *chem;
  const int NT = c.ntype_tracked, NA = c.ntype_total;
  if (NT < 1 || NT > 4 || NA < NT || NA > LEPIC_MAX_SPECIES)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: species counts out of range");
  for (int t = 0; t < NA; ++t)
    if (c.p_ncol[t] > MAXR)
      return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: more than 16 reactions per projectile");
  if (!c.sig_list || !c.sig_Er || !c.sig || !c.sig_Eth || !c.col_info || !c.sigv_mx)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: null reaction table");
  KGrid kg = make_kgrid(g);
  const int64_t nodes = grid_nodes(g);
  KChem K;
  std::memset(&K, 0, sizeof(K));
  K.c = c;
  K.np_dep = np_dep;
  K.kq = kq;
  for (int t = 0; t < NT; ++t) K.kdep[t] = c.Nm[t] / (g->h[0] * g->h[1] * g->h[2]);
  // np_mx for tracked species (device max), neutral densities for the rest
  {
    int64_t cnt = (int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1);
    int64_t nb = (cnt + BLOCK - 1) / BLOCK;
    double *part = reinterpret_cast<double *>(lepic_scratch(ctx, sizeof(double) * (size_t)(nb * NT), 1));
    if (!part) return lepic_fail(ctx, LEPIC_ERR_HIP, "collisions: scratch");
    for (int t = 0; t < NT; ++t)
      hipLaunchKernelGGL(k_npmax, dim3((unsigned)nb), dim3(BLOCK), 0, ctx->stream, kg,
                         np_red + (int64_t)t * nodes, nodes, part + t * nb);
    std::vector<double> h((size_t)(nb * NT));
    LEPIC_HIP(ctx, hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                  ctx->stream));
    LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int t = 0; t < NT; ++t) {
      double m = -INFINITY;
      for (int64_t b = 0; b < nb; ++b) m = std::max(m, h[(size_t)(t * nb + b)]);
      K.np_mx_all[t] = m;
    }
    for (int t = NT; t < NA; ++t) K.np_mx_all[t] = c.neutral_density[t - NT];
  }
  // nu_max, Pmax, Nc (mod_collisions.f90:317-372); sigv_mx is a small device table
  std::vector<int32_t> hsl((size_t)NA * LEPIC_MAX_COL), hci((size_t)c.ncol * 10);
  std::vector<double> hsv((size_t)NA * c.ncol);
  LEPIC_HIP(ctx, hipMemcpyAsync(hsl.data(), c.sig_list, sizeof(int32_t) * hsl.size(),
                                hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(ctx, hipMemcpyAsync(hci.data(), c.col_info, sizeof(int32_t) * hci.size(),
                                hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(ctx, hipMemcpyAsync(hsv.data(), c.sigv_mx, sizeof(double) * hsv.size(),
                                hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // validate the tables before any kernel indexes them
  for (int t = 0; t < NA; ++t)
    for (int ic = 0; ic < c.p_ncol[t]; ++ic) {
      int col = hsl[(size_t)t * LEPIC_MAX_COL + ic];
      if (col < 0 || col >= c.ncol)
        return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: sig_list entry out of range");
    }
  for (int col = 0; col < c.ncol; ++col) {
    int n_re = hci[(size_t)col * 10], n_by = hci[(size_t)col * 10 + 1];
    if (n_re < 1 || n_by < 0 || n_re + n_by > 8)
      return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: col_info reaction shape out of range");
  }
  if (c.npt < 2) return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: energy grid too short");
  for (int t = 0; t < NA; ++t) {
    unsigned seen = 0;
    int ntg = 0;
    for (int ic = 0; ic < c.p_ncol[t]; ++ic) {
      int col = hsl[(size_t)t * LEPIC_MAX_COL + ic];
      int tt = hci[(size_t)col * 10 + 1 + hci[(size_t)col * 10]];
      if (tt >= NT && tt < NA && !(seen & (1u << tt))) {
        seen |= 1u << tt;
        ++ntg;
      }
    }
    if (ntg > MAXT)
      return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: more than 4 neutral targets per projectile");
  }
  // register-friendly plans for projectiles whose targets are all neutral
  for (int t = 0; t < NT; ++t) {
    FastPlan &fp = K.fp[t];
    fp.ok = c.p_ncol[t] <= FR;
    for (int ic = 0; ic < c.p_ncol[t] && fp.ok; ++ic) {
      const int col = hsl[(size_t)t * LEPIC_MAX_COL + ic];
      const int tt = hci[(size_t)col * 10 + 1 + hci[(size_t)col * 10]];
      if (tt < NT || tt >= NA) {   // charged or invalid target: general trial
        fp.ok = 0;
        break;
      }
      int sl = 0;
      while (sl < fp.ntg && fp.tgt[sl] != tt) ++sl;
      if (sl == fp.ntg) {
        if (fp.ntg == FT) {
          fp.ok = 0;
          break;
        }
        fp.tgt[fp.ntg++] = tt;
      }
      fp.col[fp.nr] = col;
      fp.slot[fp.nr] = sl;
      ++fp.nr;
    }
  }
  int32_t Nc[4] = {0, 0, 0, 0};
  double lambda[4] = {0, 0, 0, 0};
  for (int t = 0; t < NT; ++t) {
    if (c.p_ncol[t] == 0 || c.mass[t] <= 0.0) continue;
    double nu_max = 0.0;
    for (int ic = 0; ic < c.p_ncol[t]; ++ic) {
      int col = hsl[(size_t)t * LEPIC_MAX_COL + ic];
      int n_re = hci[(size_t)col * 10];
      int tt = hci[(size_t)col * 10 + 1 + n_re];
      if (tt < 0 || tt >= NA) continue;
      nu_max = nu_max + K.np_mx_all[tt] * hsv[(size_t)t * c.ncol + col];
    }
    if (nu_max > c.nu_uplim[t]) nu_max = c.nu_uplim[t];
    double Pmax = nu_max * (double)ns_coll * dt;
    if (Pmax > 1.0) return lepic_fail(ctx, LEPIC_ERR_PMAX, "collisions: Pmax > 1");
    if (mode == LEPIC_MCC_REPLAY_RAN2) {
      double dNc = (double)sets.n(t) * Pmax / 1.0;
      int32_t nct = (int32_t)dNc;
      double rnd = lepic_ran2(&iseed[0]);
      if (rnd <= (dNc - (double)nct)) nct = nct + 1;
      Nc[t] = nct;
      int64_t nn = sets.n(t) > 1 ? sets.n(t) : 1;
      K.nu_max_local[t] = (double)nct / (double)nn / ((double)ns_coll * dt);
    } else {
      lambda[t] = Pmax;
      Nc[t] = (int32_t)std::llround((double)sets.n(t) * Pmax);
      K.nu_max_local[t] = Pmax / ((double)ns_coll * dt);
    }
  }
  if (fp32)
    for (int t = 0; t < NT; ++t)
      if (c.p_ncol[t] > 0 && c.mass[t] > 0.0 && !K.fp[t].ok)
        return lepic_fail(ctx, LEPIC_ERR_ARG,
                          "collisions32: FP32 sets need every reaction's target to be a neutral");
  PA64 P;
  PA32 P32;
  std::memset(&P, 0, sizeof(P));
  std::memset(&P32, 0, sizeof(P32));
  for (int t = 0; t < NT; ++t) {
    if (fp32) {
      const lepic_pset32 &q = sets.parts32[t];
      P32.P.cell[t] = q.cell;
      P32.P.fx[t] = q.fx; P32.P.fy[t] = q.fy; P32.P.fz[t] = q.fz;
      P32.P.vx[t] = q.vx; P32.P.vy[t] = q.vy; P32.P.vz[t] = q.vz;
      P32.P.cap[t] = q.cap;
      if (sets.dep) {
        const double k1 = c.Nm[t] / (g->h[0] * g->h[1] * g->h[2]);
        const double unit = std::ldexp(std::fabs(k1), sets.frac_bits[t]);
        if (k1 != 0.0 && !(unit >= 1.0 && unit <= LEPIC_FIXED32_MAX))
          return lepic_fail(ctx, LEPIC_ERR_ARG,
                            "collisions32: |Nm/(h1 h2 h3)| * 2^frac_bits must lie in [1, 2^21]");
        P32.P.dep[t] = reinterpret_cast<long long *>(sets.dep + (int64_t)t * nodes);
        P32.P.U[t] = std::rint(unit);
        P32.P.neg[t] = k1 < 0.0 ? 1 : 0;
      }
    } else {
      const lepic_pset &q = sets.parts[t];
      P.P.x[t] = q.x; P.P.y[t] = q.y; P.P.z[t] = q.z;
      P.P.vx[t] = q.vx; P.P.vy[t] = q.vy; P.P.vz[t] = q.vz;
      P.P.dead[t] = q.dead;
      P.P.cap[t] = q.cap;
    }
  }
  MccCounters hc;
  std::memset(&hc, 0, sizeof(hc));
  for (int t = 0; t < NT; ++t) hc.n[t] = sets.n(t);
  hc.seed = iseed ? iseed[0] : 0;
  MccCounters *dc = reinterpret_cast<MccCounters *>(ctx->d_ctrl + CTRL_MCC);
  static_assert(sizeof(MccCounters) <= 256 * sizeof(int64_t), "MCC counters fit the control block");
  LEPIC_HIP(ctx, hipMemcpyAsync(dc, &hc, sizeof(hc), hipMemcpyHostToDevice, ctx->stream));
  if (mode == LEPIC_MCC_REPLAY_RAN2) {
    hipLaunchKernelGGL(k_mcc_replay, dim3(1), dim3(64), 0, ctx->stream, kg, P.P, K, np_red, nodes,
                       Nc[0], Nc[1], Nc[2], Nc[3], dc, Pcoll);
    LEPIC_LAUNCH_CHECK(ctx, "k_mcc_replay");
  } else if (mode == LEPIC_MCC_PHILOX) {
    for (int t = 0; t < NT; ++t) {
      if (c.p_ncol[t] == 0 || c.mass[t] <= 0.0 || lambda[t] <= 0.0 || sets.n(t) == 0) continue;
      const int64_t n0 = sets.n(t);
      const double lam = lambda[t];
      const double p1 = lam / std::expm1(lam);
      const double mean = (double)n0 * -std::expm1(-lam);
      int64_t cap = (int64_t)(mean + 16.0 * std::sqrt(mean)) + 65536;
      if (cap > n0) cap = n0;
      unsigned long long *nsel = reinterpret_cast<unsigned long long *>(&dc->nsel[t]);
      const int64_t nthreads = (n0 + SEL_CHUNK - 1) / SEL_CHUNK;
      for (int pass = 0; pass < 2; ++pass) {
        uint64_t *list = reinterpret_cast<uint64_t *>(lepic_scratch(ctx, sizeof(uint64_t) * (size_t)cap, 0));
        if (!list) return lepic_fail(ctx, LEPIC_ERR_HIP, "collisions: selection list allocation");
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream,
                           reinterpret_cast<uint64_t *>(nsel), (int64_t)1);
        if (fp32)
          hipLaunchKernelGGL(k_mcc_select<LiveCell>, dim3(grid_blocks(nthreads, BLOCK)),
                             dim3(BLOCK), 0, ctx->stream, LiveCell{P32.P.cell[t]}, t, n0, lam, p1,
                             philox_key, philox_counter, list, cap, nsel);
        else
          hipLaunchKernelGGL(k_mcc_select<LiveDead>, dim3(grid_blocks(nthreads, BLOCK)),
                             dim3(BLOCK), 0, ctx->stream, LiveDead{P.P.dead[t]}, t, n0, lam, p1,
                             philox_key, philox_counter, list, cap, nsel);
        LEPIC_LAUNCH_CHECK(ctx, "k_mcc_select");
        LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, nsel, sizeof(int64_t), hipMemcpyDeviceToHost,
                                      ctx->stream));
        LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const int64_t m = ctx->h_pinned[0];
        if (m > cap) {   // an unlikely tail: list again with room for all (same draws)
          cap = m;
          continue;
        }
        if (m > 0) {
          if (fp32)
            hipLaunchKernelGGL((k_mcc_trials<true, PA32>), dim3(grid_blocks(m, BLOCK)), dim3(BLOCK),
                               0, ctx->stream, kg, P32, K, np_red, nodes, t, list, nsel, cap,
                               philox_key, philox_counter, dc, Pcoll);
          else if (K.fp[t].ok)
            hipLaunchKernelGGL((k_mcc_trials<true, PA64>), dim3(grid_blocks(m, BLOCK)), dim3(BLOCK),
                               0, ctx->stream, kg, P, K, np_red, nodes, t, list, nsel, cap,
                               philox_key, philox_counter, dc, Pcoll);
          else
            hipLaunchKernelGGL((k_mcc_trials<false, PA64>), dim3(grid_blocks(m, BLOCK)),
                               dim3(BLOCK), 0, ctx->stream, kg, P, K, np_red, nodes, t, list, nsel,
                               cap, philox_key, philox_counter, dc, Pcoll);
        }
        LEPIC_LAUNCH_CHECK(ctx, "k_mcc_trials");
        break;
      }
    }
  } else {
    return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions: unknown mode");
  }
  LEPIC_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int t = 0; t < NT; ++t) sets.set_n(t, hc.n[t] < sets.cap(t) ? hc.n[t] : sets.cap(t));
  if (mode == LEPIC_MCC_REPLAY_RAN2) iseed[0] = (int32_t)hc.seed;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    for (int t = 0; t < NT; ++t) {
      stats->Nc[t] = Nc[t];
      stats->accepted[t] = hc.accepted[t];
      stats->err_coll[t] = hc.err_coll[t];
      stats->skipped_dead[t] = hc.skipped_dead[t];
      stats->appended[t] = hc.appended[t];
    }
  }
  if (hc.overflow) return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "collisions: product capacity exceeded");
  return LEPIC_OK;
}

extern "C" int lepic_collisions(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *parts,
                                const double *np_red, const lepic_chem *chem, int32_t ns_coll,
                                double dt, int32_t mode, int32_t *iseed, uint64_t philox_key,
                                uint64_t philox_counter, double *Pcoll, lepic_mcc_stats *stats) {
  MccSets sets;
  sets.parts = parts;
  return collisions_impl(ctx, g, sets, np_red, chem, ns_coll, dt, mode, iseed, philox_key,
                         philox_counter, Pcoll, stats, nullptr, nullptr);
}

extern "C" int lepic_collisions_deposited(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *parts,
                                          const double *np_red, const lepic_chem *chem,
                                          int32_t ns_coll, double dt, int32_t mode, int32_t *iseed,
                                          uint64_t philox_key, uint64_t philox_counter,
                                          double *Pcoll, lepic_mcc_stats *stats, double *np_dep,
                                          const double *kq) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!np_dep || !kq) return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions_deposited: null density");
  MccSets sets;
  sets.parts = parts;
  return collisions_impl(ctx, g, sets, np_red, chem, ns_coll, dt, mode, iseed, philox_key,
                         philox_counter, Pcoll, stats, np_dep, kq);
}

extern "C" int lepic_collisions32(lepic_ctx *ctx, const lepic_grid *g, lepic_pset32 *parts,
                                  const double *np_red, const lepic_chem *chem, int32_t ns_coll,
                                  double dt, uint64_t philox_key, uint64_t philox_counter,
                                  double *Pcoll, lepic_mcc_stats *stats, int64_t *dep,
                                  const int32_t *frac_bits) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!parts || (dep && !frac_bits))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions32: null argument");
  if (chem)
    for (int t = 0; t < chem->ntype_tracked && t < 4; ++t) {
      const void *a[7] = {parts[t].cell, parts[t].fx, parts[t].fy, parts[t].fz,
                          parts[t].vx, parts[t].vy, parts[t].vz};
      for (const void *q : a)
        if (parts[t].cap > 0 && !q) return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions32: bad set");
      if (parts[t].n < 0 || parts[t].n > parts[t].cap)
        return lepic_fail(ctx, LEPIC_ERR_ARG, "collisions32: bad set");
    }
  MccSets sets;
  sets.parts32 = parts;
  sets.dep = dep;
  sets.frac_bits = frac_bits;
  return collisions_impl(ctx, g, sets, np_red, chem, ns_coll, dt, LEPIC_MCC_PHILOX, nullptr,
                         philox_key, philox_counter, Pcoll, stats, nullptr, nullptr);
}
