// lepic_particles.hip -- particle-phase kernels of the PIC step on gfx950:
// movers (a3/a4), wall BC with stable compaction (a5), deposition (a6),
// the fused gather+push+BC+deposit kernel, the stable cell sort (a2) and the
// device loader (a13).
//
// Layout: one species per lepic_pset, structure of arrays in HBM
// (x,y,z,vx,vy,vz FP64, cell int32, dead int8).  Stable compaction and the
// particle order of the sort are those of the reference; both are produced
// out of place (src -> dst) so that every particle is read once and written
// once per pass.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <cmath>
#include <vector>

#include "lepic_internal.h"
#include "lepic_partutil.h"
#include "lepic_philox.h"

namespace {

constexpr int BLOCK = 256;
constexpr int NWAVES = BLOCK / LEPIC_WAVE;
// Doubles of the per-block deposit box.  k_push holds 153 VGPRs (3 waves per
// SIMD), which leaves 53 KB of LDS per block: its box covers the tiles that
// straddle a row of cells after electrons drifted for a few steps since the
// last sort (131 x 7 x 7 nodes); the other deposit kernels keep 32 KB.
constexpr int LDS_BOX = 4096;
constexpr int LDS_BOX_PUSH = 6656;
constexpr long SPIN_LIMIT = 1L << 26;    // bounded look-back wait

struct KPset {
  double *x, *y, *z, *vx, *vy, *vz;
  int32_t *cell;
  int8_t *dead;
};

inline KPset kp(const lepic_pset *p) {
  KPset k{p->x, p->y, p->z, p->vx, p->vy, p->vz, p->cell, p->dead};
  return k;
}

// --------------------------------------------------------------------------
// a3: move_particles_electrostatic (mod_particleMover.f90:72-156)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_move_es(KGrid g, KPset p, int64_t n,
                                                   const double4 *__restrict__ E4, double qmdt,
                                                   double dt) {
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  double xp = p.x[i], yp = p.y[i], zp = p.z[i];
  Stencil s;
  make_stencil(g, xp, yp, zp, s);
  double Ex, Ey, Ez;
  gather3(E4, g, s, Ex, Ey, Ez);
  double vx = p.vx[i] + qmdt * Ex;
  double vy = p.vy[i] + qmdt * Ey;
  double vz = p.vz[i] + qmdt * Ez;
  p.vx[i] = vx; p.vy[i] = vy; p.vz[i] = vz;
  p.x[i] = xp + dt * vx; p.y[i] = yp + dt * vy; p.z[i] = zp + dt * vz;
}

// --------------------------------------------------------------------------
// a4: part_mover Boris/leapfrog velocity update (Src/part_expmover.f90:51-246)
// --------------------------------------------------------------------------
template <int BM>
__global__ __launch_bounds__(BLOCK) void k_move_boris(KGrid g, KPset p, int64_t n,
                                                      const double4 *__restrict__ E4, KBField bf,
                                                      double k1, double dt) {
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  double xp = p.x[i], yp = p.y[i], zp = p.z[i];
  double vx = p.vx[i], vy = p.vy[i], vz = p.vz[i];
  Stencil s;
  make_stencil(g, xp, yp, zp, s);
  double Ex, Ey, Ez;
  gather3(E4, g, s, Ex, Ey, Ez);
  velocity_update<BM>(g, bf, s, k1, Ex, Ey, Ez, vx, vy, vz, xp, yp, zp);
  p.vx[i] = vx; p.vy[i] = vy; p.vz[i] = vz;
  p.x[i] = xp + dt * vx; p.y[i] = yp + dt * vy; p.z[i] = zp + dt * vz;
}

// --------------------------------------------------------------------------
// Wall test of apply_particle_bc_legacy (mod_particleBC.f90:96-116).
// Returns 0 survive, 1 lost on walls, 2 lost negative ion at a Neumann wall.
// --------------------------------------------------------------------------
__device__ inline int wall_test(const KGrid &g, const int *__restrict__ bcnd, const KBC &bp,
                                double xp, double yp, double zp, int &label) {
  int ix = (int)floor(qdiv(xp, g.q1)) + 1;
  int iy = (int)floor(qdiv(yp, g.q2)) + 1;
  int iz = (int)floor(qdiv(zp, g.q3)) + 1;
  if (ix < 0) ix = 0;
  if (ix > g.n1 + 1) ix = g.n1 + 1;
  if (iy < 0) iy = 0;
  if (iy > g.n2 + 1) iy = g.n2 + 1;
  if (iz < 0) iz = 0;
  if (iz > g.n3 + 1) iz = g.n3 + 1;
  int64_t b = gidx(g, ix, iy, iz);
  int c0 = bcnd[b], c1 = bcnd[b + 1], c2 = bcnd[b + g.sy + 1], c3 = bcnd[b + g.sy];
  int c4 = bcnd[b + g.sz], c5 = bcnd[b + g.sz + 1], c6 = bcnd[b + g.sz + g.sy + 1],
      c7 = bcnd[b + g.sz + g.sy];
  int lost = (c0 >= 1 && c1 >= 1 && c2 >= 1 && c3 >= 1 && c4 >= 1 && c5 >= 1 && c6 >= 1 &&
              c7 >= 1) ? 1 : 0;
  if (bp.ptype == bp.tag_neg && xp < 0.0 && bp.flag_nmn == 1) lost = 2;
  label = c0;
  if (lost == 2) label = 0;
  if (label < 0) label = 0;
  return lost;
}

// The same test from a per-node table: wt[node] = bcnd[node] when all 8
// nodes of the cell whose lower corner is node have bcnd >= 1 (a particle
// there is lost), -1 otherwise.  One load instead of eight.
__device__ inline int wall_test_tab(const KGrid &g, const int *__restrict__ wt, const KBC &bp,
                                    double xp, double yp, double zp, int &label) {
  int ix = (int)floor(qdiv(xp, g.q1)) + 1;
  int iy = (int)floor(qdiv(yp, g.q2)) + 1;
  int iz = (int)floor(qdiv(zp, g.q3)) + 1;
  ix = imax(0, imin(g.n1 + 1, ix));
  iy = imax(0, imin(g.n2 + 1, iy));
  iz = imax(0, imin(g.n3 + 1, iz));
  const int c = wt[gidx(g, ix, iy, iz)];
  int lost = c >= 0 ? 1 : 0;
  label = c >= 0 ? c : 0;
  if (bp.ptype == bp.tag_neg && xp < 0.0 && bp.flag_nmn == 1) {
    lost = 2;
    label = 0;
  }
  return lost;
}

// wall_test_tab from the quotients x/h1, y/h2, z/h3 already formed.
__device__ inline int wall_test_tab_q(const KGrid &g, const int *__restrict__ wt, const KBC &bp,
                                      double xp, double qx, double qy, double qz, int &label) {
  int ix = (int)floor(qx) + 1;
  int iy = (int)floor(qy) + 1;
  int iz = (int)floor(qz) + 1;
  ix = imax(0, imin(g.n1 + 1, ix));
  iy = imax(0, imin(g.n2 + 1, iy));
  iz = imax(0, imin(g.n3 + 1, iz));
  const int c = wt[gidx(g, ix, iy, iz)];
  int lost = c >= 0 ? 1 : 0;
  label = c >= 0 ? c : 0;
  if (bp.ptype == bp.tag_neg && xp < 0.0 && bp.flag_nmn == 1) {
    lost = 2;
    label = 0;
  }
  return lost;
}

// Surface charge of a particle lost on a dielectric
// (mod_particleBC.f90:139-188): igrid is the lost particle's wall label
// (> 0), x, y, z its position after the push.  The position is wrapped
// along z for periodic runs before the charge is spread; the particle is
// lost either way.  Called for lost particles only.
__device__ inline void die_charge(const KGrid &g, const KBC &bp, int igrid, double xp, double yp,
                                  double zp) {
  if (igrid > bp.ngrid + 1) return;     // outside the dtype table
  const int dt = bp.dtype[igrid - 1];   // the routine's 1-based dtype(:) dummy
  if (dt <= 1) return;
  int ix = (int)floor(qdiv(xp, g.q1)) + 1;
  int iy = (int)floor(qdiv(yp, g.q2)) + 1;
  int iz = (int)floor(qdiv(zp, g.q3)) + 1;
  if (bp.flag_pbc == 1) {
    if (zp >= bp.zmax) {
      zp = zp - bp.zmax;
      iz = (int)floor(qdiv(zp, g.q3)) + 1;
    }
    if (zp <= 0.0) {
      zp = bp.zmax + zp;
      iz = (int)floor(qdiv(zp, g.q3)) + 1;
    }
  }
  ix = imax(0, imin(g.n1 + 1, ix));
  iy = imax(0, imin(g.n2 + 1, iy));
  iz = imax(0, imin(g.n3 + 1, iz));
  const double pz = qdiv((double)iz * g.h3 - zp, g.q3);
  const double q = bp.qmacro;
  if (dt == 2) {
    const double px = qdiv((double)ix * g.h1 - xp, g.q1);
    const int64_t s = g.n1 + 3;
    atomicAdd(&bp.sum_q_xz[ix + s * iz], q * px * pz);
    atomicAdd(&bp.sum_q_xz[(ix + 1) + s * iz], q * (1.0 - px) * pz);
    atomicAdd(&bp.sum_q_xz[(ix + 1) + s * (iz + 1)], q * (1.0 - px) * (1.0 - pz));
    atomicAdd(&bp.sum_q_xz[ix + s * (iz + 1)], q * px * (1.0 - pz));
  } else if (dt == 3 || dt == 4) {
    const double py = qdiv((double)iy * g.h2 - yp, g.q2);
    const int64_t s = g.n2 + 3;
    const int c = dt - 3;
    atomicAdd(&bp.sum_q_yz[c + 2 * (iy + s * iz)], q * py * pz);
    atomicAdd(&bp.sum_q_yz[c + 2 * ((iy + 1) + s * iz)], q * (1.0 - py) * pz);
    atomicAdd(&bp.sum_q_yz[c + 2 * ((iy + 1) + s * (iz + 1))], q * (1.0 - py) * (1.0 - pz));
    atomicAdd(&bp.sum_q_yz[c + 2 * (iy + s * (iz + 1))], q * py * (1.0 - pz));
  }
}


__device__ inline void survivor_wrap(const KBC &bp, double &xp, double &yp, double &zp,
                                     double &vx) {
  if (bp.flag_nmn == 1 && xp <= 0.0) {
    xp = -xp;
    vx = -vx;
  }
  if (bp.flag_pbc == 1) {
    if (yp >= bp.ymax) yp = yp - bp.ymax;
    if (yp <= 0.0) yp = bp.ymax + yp;
    if (zp >= bp.zmax) zp = zp - bp.zmax;
    if (zp <= 0.0) zp = bp.zmax + zp;
  }
}

// --------------------------------------------------------------------------
// Block-wide stable compaction offsets with decoupled look-back.
// Status word: epoch(16) | state(2) | value(46); state 1 = tile aggregate,
// state 2 = inclusive prefix.  Tiles are numbered in the order blocks start,
// so every wait is on a block that is already running.
// --------------------------------------------------------------------------
struct LBShared {
  int wave_tot[NWAVES];
  int64_t prefix;
  int tile;
};

__device__ inline int lane_id() { return __lane_id(); }

__device__ inline int64_t lookback(uint64_t *status, int64_t tile, uint32_t epoch, int64_t agg,
                                   int64_t *spin_fail) {
  const uint64_t E = (uint64_t)epoch << 48;
  const uint64_t VMASK = (1ull << 46) - 1;
  if (tile == 0) {
    __hip_atomic_store(&status[0], E | (2ull << 46) | (uint64_t)agg, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  __hip_atomic_store(&status[tile], E | (1ull << 46) | (uint64_t)agg, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  int64_t prefix = 0;
  int64_t j = tile - 1;
  long spins = 0;
  while (j >= 0) {
    uint64_t w = __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t ep = (uint32_t)(w >> 48);
    uint32_t st = (uint32_t)(w >> 46) & 3u;
    if (ep != epoch || st == 0) {
      if (++spins > SPIN_LIMIT) {
        atomicExch((unsigned long long *)spin_fail, 1ull);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += (int64_t)(w & VMASK);
    if (st == 2) break;
    --j;
  }
  __hip_atomic_store(&status[tile], E | (2ull << 46) | (uint64_t)(prefix + agg),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// Returns this thread's output slot (valid when keep) and, through sh, the
// tile's exclusive prefix.  Must be called by every thread of the block.
__device__ inline int64_t compact_slot(bool keep, LBShared &sh, uint64_t *status, int64_t tile,
                                       uint32_t epoch, int64_t *spin_fail, int64_t *count_out,
                                       int64_t ntiles) {
  const int lane = lane_id();
  const int wave = threadIdx.x / LEPIC_WAVE;
  uint64_t mask = __ballot(keep);
  int rank_in_wave = __popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) sh.wave_tot[wave] = __popcll(mask);
  __syncthreads();
  int wave_off = 0, tot = 0;
  for (int w = 0; w < NWAVES; ++w) {
    if (w < wave) wave_off += sh.wave_tot[w];
    tot += sh.wave_tot[w];
  }
  if (threadIdx.x == 0) {
    int64_t pre = lookback(status, tile, epoch, tot, spin_fail);
    sh.prefix = pre;
    if (tile == ntiles - 1) *count_out = pre + tot;
  }
  __syncthreads();
  return sh.prefix + wave_off + rank_in_wave;
}

__device__ inline int64_t take_tile(int *counter, int64_t ntiles, LBShared &sh) {
  if (threadIdx.x == 0) {
    int t = atomicAdd(counter, 1);
    if (t == ntiles - 1) atomicExch(counter, 0);  // last tile handed out: reset
    sh.tile = t;
  }
  __syncthreads();
  return sh.tile;
}

// --------------------------------------------------------------------------
// a5: apply_particle_bc_legacy (mod_particleBC.f90:42-259) out of place.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_apply_bc(KGrid g, KPset src, KPset dst, int64_t n,
                                                    const int *__restrict__ bcnd, KBC bp,
                                                    double *p_mac, double *P_loss,
                                                    uint64_t *status, uint32_t epoch,
                                                    int *counter, int64_t *spin_fail,
                                                    int64_t *count_out, int64_t ntiles) {
  __shared__ LBShared sh;
  int64_t tile = take_tile(counter, ntiles, sh);
  int64_t i = tile * BLOCK + threadIdx.x;
  bool keep = false, lost = false;
  int label = 0;
  double xp = 0, yp = 0, zp = 0, vx = 0, vy = 0, vz = 0;
  if (i < n) {
    if (src.dead[i] != 1) {
      xp = src.x[i]; yp = src.y[i]; zp = src.z[i];
      vx = src.vx[i]; vy = src.vy[i]; vz = src.vz[i];
      lost = wall_test(g, bcnd, bp, xp, yp, zp, label) != 0;
      if (!lost) {
        survivor_wrap(bp, xp, yp, zp, vx);
        keep = true;
      } else if (bp.flag_die == 1 && label > 0) {
        die_charge(g, bp, label, xp, yp, zp);
      }
    }
  }
  wall_diag_wave(bp, lost, label, vx, vy, vz, p_mac, P_loss);
  int64_t o = compact_slot(keep, sh, status, tile, epoch, spin_fail, count_out, ntiles);
  if (keep) {
    dst.x[o] = xp; dst.y[o] = yp; dst.z[o] = zp;
    dst.vx[o] = vx; dst.vy[o] = vy; dst.vz[o] = vz;
    dst.dead[o] = 0;
  }
}

// --------------------------------------------------------------------------
// a6: deposition helpers.  The 8 weights of deposit_particle_set_to_np_thread
// (mod_chargeDeposition.f90:90-131); ix = int(x/h)+1 without clamping, after
// the [-1e-12, L+1e-12] range guard (:67-69).
// --------------------------------------------------------------------------
__device__ inline bool deposit_weights(const KGrid &g, double k1, double x, double y, double z,
                                       int &ix, int &iy, int &iz, double w[8]) {
  const double xmax = (double)g.n1 * g.h1, ymax = (double)g.n2 * g.h2,
               zmax = (double)g.n3 * g.h3;
  // written as negated ranges so that NaN positions are skipped as well
  if (!(x >= -1.0e-12 && x <= xmax + 1.0e-12)) return false;
  if (!(y >= -1.0e-12 && y <= ymax + 1.0e-12)) return false;
  if (!(z >= -1.0e-12 && z <= zmax + 1.0e-12)) return false;
  ix = (int)qdiv(x, g.q1) + 1;
  iy = (int)qdiv(y, g.q2) + 1;
  iz = (int)qdiv(z, g.q3) + 1;
  double px = qdiv((double)ix * g.h1 - x, g.q1);
  double py = qdiv((double)iy * g.h2 - y, g.q2);
  double pz = qdiv((double)iz * g.h3 - z, g.q3);
  double wx2 = 1.0 - px, wy2 = 1.0 - py, wz2 = 1.0 - pz;
  w[0] = k1 * px * py * pz;
  w[1] = k1 * wx2 * py * pz;
  w[2] = k1 * wx2 * wy2 * pz;
  w[3] = k1 * px * wy2 * pz;
  w[4] = k1 * px * py * wz2;
  w[5] = k1 * wx2 * py * wz2;
  w[6] = k1 * wx2 * wy2 * wz2;
  w[7] = k1 * px * wy2 * wz2;
  return true;
}

// Block deposition.  Each block owns an LDS box of nodes covering its
// particles' deposit stencils.  Within a wave, lanes whose particles share a
// lower node (runs of one cell in sorted data) are first summed with a
// segmented shuffle scan, so one lane per run issues the 8 LDS atomics; the
// box is then flushed with one L2-resident atomic per touched node into this
// XCD's replica.  Particles whose stencil leaves the box go straight to the
// replica.  The replicas hold sum(w); k_rep_flush multiplies each node's
// total by kq once.  kq is 1 or 2, so kq*sum(w) equals sum(kq*w) term for
// term.
template <int NB>
struct DepSharedT {
  static constexpr int CAP = NB;
  // extents of a cut box (nodes): x, then y and z
  static constexpr int CUT_X = 36;
  static constexpr int CUT_YZ = NB >= 36 * 13 * 13 ? 13 : 10;
  static_assert(CUT_X * CUT_YZ * CUT_YZ <= NB, "cut box exceeds the box");
  int mn[3], mx[3];
  int x0, y0, z0, bx, by, bz, vol;
  uint64_t mbx, mby;   // small_div multipliers of bx and by
  double box[NB];
};
using DepShared = DepSharedT<LDS_BOX>;
using DepSharedPush = DepSharedT<LDS_BOX_PUSH>;

// Inclusive segmented sum over runs of equal keys in lane order; returns true
// on the last lane of each run, which then holds the run's totals.
__device__ inline bool seg_sum8(int key, double v[8]) {
  const int lane = __lane_id();
  const int prev = __shfl_up(key, 1);
  const int next = __shfl_down(key, 1);
  const bool head = (lane == 0) || (prev != key);
  const bool tail = (lane == LEPIC_WAVE - 1) || (next != key);
  int start = head ? lane : 0;
  for (int d = 1; d < LEPIC_WAVE; d <<= 1) {
    int s2 = __shfl_up(start, d);
    if (lane >= d) start = max(start, s2);
  }
  for (int d = 1; d < LEPIC_WAVE; d <<= 1) {
    const bool take = (lane - d) >= start;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double u = __shfl_up(v[q], d);
      if (take) v[q] += u;
    }
  }
  return tail;
}

// Pass 1: the block's cell range from the particles' current positions,
// widened by one cell below and two above for motion within the step.  When
// that range exceeds the box, the box is cut to BOXCUT nodes around the
// block's mean lower node (sm / cnt: the lane's sum of lower nodes and its
// particle count): a cell-sorted tile spans a row of cells along x and one
// cell in y and z, and the few particles that drifted beyond the cut box
// deposit into the replica one by one instead of sending the whole block
// there.
template <class DS>
__device__ inline void box_begin(const KGrid &g, DS &ds, int lo[3], int hi[3], const int sm[3],
                                 int cnt) {
  __shared__ int bsum[4];
  if (threadIdx.x == 0) {
    ds.mn[0] = ds.mn[1] = ds.mn[2] = 1 << 30;
    ds.mx[0] = ds.mx[1] = ds.mx[2] = -(1 << 30);
    bsum[0] = bsum[1] = bsum[2] = bsum[3] = 0;
  }
  __syncthreads();
  int a0 = wave_min(lo[0]), a1 = wave_min(lo[1]), a2 = wave_min(lo[2]);
  int b0 = wave_max(hi[0]), b1 = wave_max(hi[1]), b2 = wave_max(hi[2]);
  int s0 = sm[0], s1 = sm[1], s2 = sm[2], c = cnt;
  for (int d = 32; d > 0; d >>= 1) {
    s0 += __shfl_xor(s0, d); s1 += __shfl_xor(s1, d); s2 += __shfl_xor(s2, d);
    c += __shfl_xor(c, d);
  }
  if (__lane_id() == 0) {
    atomicMin(&ds.mn[0], a0); atomicMin(&ds.mn[1], a1); atomicMin(&ds.mn[2], a2);
    atomicMax(&ds.mx[0], b0); atomicMax(&ds.mx[1], b1); atomicMax(&ds.mx[2], b2);
    atomicAdd(&bsum[0], s0); atomicAdd(&bsum[1], s1); atomicAdd(&bsum[2], s2);
    atomicAdd(&bsum[3], c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int vol = 0;
    if (ds.mx[0] >= ds.mn[0]) {
      const int top[3] = {g.n1 + 2, g.n2 + 2, g.n3 + 2};
      int a[3], e[3];
      for (int d = 0; d < 3; ++d) {
        a[d] = max(ds.mn[d] - 1, 0);
        e[d] = min(ds.mx[d] + 2, top[d]) - a[d] + 1;
      }
      if (e[0] * e[1] * e[2] > DS::CAP) {
        const int lim[3] = {DS::CUT_X, DS::CUT_YZ, DS::CUT_YZ};
        for (int d = 0; d < 3; ++d) {
          if (e[d] <= lim[d]) continue;
          const int mean = bsum[d] / max(bsum[3], 1);
          int st = mean - lim[d] / 2;
          st = min(max(st, a[d]), a[d] + e[d] - lim[d]);
          a[d] = st;
          e[d] = lim[d];
        }
      }
      ds.x0 = a[0]; ds.y0 = a[1]; ds.z0 = a[2];
      ds.bx = e[0]; ds.by = e[1]; ds.bz = e[2];
      ds.mbx = small_div_magic((uint32_t)e[0]);
      ds.mby = small_div_magic((uint32_t)e[1]);
      vol = e[0] * e[1] * e[2];
    }
    ds.vol = vol;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ds.vol; t += BLOCK) ds.box[t] = 0.0;
  __syncthreads();
}

// Deposit one particle per lane (has = false lanes take part in the shuffles
// with a unique key and no weight).
template <class DS>
__device__ inline void box_add(const KGrid &g, DS &ds, bool has, int ix, int iy, int iz,
                               double w[8], double *rep) {
  int l = -1;
  if (has && ds.vol > 0) {
    int lx = ix - ds.x0, ly = iy - ds.y0, lz = iz - ds.z0;
    if (lx >= 0 && ly >= 0 && lz >= 0 && lx + 1 < ds.bx && ly + 1 < ds.by && lz + 1 < ds.bz)
      l = lx + ds.bx * (ly + ds.by * lz);
  }
  if (has && l < 0) {  // outside the box: straight to this XCD's replica
    int64_t b = gidx(g, ix, iy, iz);
    const int64_t nd[8] = {b, b + 1, b + g.sy + 1, b + g.sy, b + g.sz, b + g.sz + 1,
                           b + g.sz + g.sy + 1, b + g.sz + g.sy};
#pragma unroll
    for (int q = 0; q < 8; ++q) l2_atomic_add(&rep[nd[q]], w[q]);
  }
  if (ds.vol <= 0) return;  // uniform across the block
  const int key = (l >= 0) ? l : -1 - __lane_id();
  if (l < 0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) w[q] = 0.0;
  }
  const bool tail = seg_sum8(key, w);
  if (tail && l >= 0) {
    const int bx = ds.bx, bxy = ds.bx * ds.by;
    const int ld[8] = {l, l + 1, l + bx + 1, l + bx, l + bxy, l + bxy + 1, l + bxy + bx + 1,
                       l + bxy + bx};
#pragma unroll
    for (int q = 0; q < 8; ++q) atomicAdd(&ds.box[ld[q]], w[q]);
  }
}

template <class DS>
__device__ inline void box_flush(const KGrid &g, DS &ds, double *rep) {
  __syncthreads();
  for (int t = threadIdx.x; t < ds.vol; t += BLOCK) {
    double v = ds.box[t];
    if (v != 0.0) {
      const int r = (int)small_div((uint32_t)t, ds.mbx), lz = (int)small_div((uint32_t)r, ds.mby);
      const int lx = t - r * ds.bx, ly = r - lz * ds.by;
      int64_t node = gidx(g, ds.x0 + lx, ds.y0 + ly, ds.z0 + lz);
      l2_atomic_add(&rep[node], v);
    }
  }
}

// Cell of a position for the LDS box bounds only: x * RN(1/h) without the
// corrections, so the index can be one off next to a cell face.  The bounds
// decide which deposits go through LDS and which straight to the replica
// (every deposit tests its own stencil against the box), never what is
// deposited, so an approximate cell costs at most a few global atomics.
// Clamped in floating point first so that any input, NaN included, converts.
__device__ inline int approx_index(double x, const QDiv &d, int n) {
  const double q = fmin(fmax(x * d.y, -1.0), (double)n + 1.0);
  return imax(0, imin(n + 1, (int)q + 1));
}

__device__ inline void cell_of(const KGrid &g, double x, double y, double z, int c[3]) {
  c[0] = approx_index(x, g.q1, g.n1);
  c[1] = approx_index(y, g.q2, g.n2);
  c[2] = approx_index(z, g.q3, g.n3);
}

// --------------------------------------------------------------------------
// Deposit in runs.  Cell-sorted particles come in runs of one cell; a lane
// sums the 8 weights of consecutive particles with the same key in registers
// and hands a run to the block's LDS box (or, outside the box, to this XCD's
// replica) with one set of 8 atomics.
// --------------------------------------------------------------------------
// deposit_fracs given the quotients x/h1, y/h2, z/h3 (ok: the particle's
// qdiv_coord_ok).
__device__ inline bool deposit_fracs_q(const KGrid &g, double x, double y, double z, double qx,
                                       double qy, double qz, bool ok, int &ix, int &iy, int &iz,
                                       double &px, double &py, double &pz) {
  const double xmax = (double)g.n1 * g.h1, ymax = (double)g.n2 * g.h2,
               zmax = (double)g.n3 * g.h3;
  if (!(x >= -1.0e-12 && x <= xmax + 1.0e-12)) return false;
  if (!(y >= -1.0e-12 && y <= ymax + 1.0e-12)) return false;
  if (!(z >= -1.0e-12 && z <= zmax + 1.0e-12)) return false;
  ix = (int)qx + 1;
  iy = (int)qy + 1;
  iz = (int)qz + 1;
  px = qdiv_pre((double)ix * g.h1 - x, g.q1, ok);
  py = qdiv_pre((double)iy * g.h2 - y, g.q2, ok);
  pz = qdiv_pre((double)iz * g.h3 - z, g.q3, ok);
  return true;
}


// Adds a run's 8 sums into the LDS box.  Neighbouring lanes of sorted data
// flush the same cell; each lane starts at node (lane mod 8) of the stencil,
// so the lanes of one instruction hit different addresses instead of
// serialising on the same one.
template <class DS>
__device__ inline void box_atomic8(DS &ds, int l, const double a[8]) {
  const int bx = ds.bx, bxy = ds.bx * ds.by;
  const int rot = __lane_id() & 7;
  double b[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) b[j] = a[j];
#pragma unroll
  for (int stage = 1; stage < 8; stage <<= 1) {   // b[j] <- b[(j + stage) & 7] where bit is set
    const bool take = (rot & stage) != 0;
    double t[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) t[j] = b[(j + stage) & 7];
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = take ? t[j] : b[j];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) atomicAdd(&ds.box[l + node_off((j + rot) & 7, bx, bxy)], b[j]);
}

// A run with a box key goes to LDS, a run with a global key to this XCD's
// replica.
template <class DS>
__device__ inline void flush_run(const KGrid &g, DS &ds, int key, const double a[8],
                                 double *rep) {
  if (key >= 0) {
    box_atomic8(ds, key, a);
  } else if (key <= -2) {
    const int64_t b = (int64_t)(-(key + 2));
    const int64_t nd[8] = {b, b + 1, b + g.sy + 1, b + g.sy, b + g.sz, b + g.sz + 1,
                           b + g.sz + g.sy + 1, b + g.sz + g.sy};
#pragma unroll
    for (int q = 0; q < 8; ++q) l2_atomic_add(&rep[nd[q]], a[q]);
  }
}

// Key of one particle's stencil for the run deposit: the box-local
// lower node (>= 0) when the stencil lies in the block's LDS box, otherwise
// -(global lower node) - 2, so that runs outside the box are summed in
// registers too before their global atomics.
template <class DS>
__device__ inline int stage_key(const KGrid &g, const DS &ds, int ix, int iy, int iz) {
  if (ds.vol > 0) {
    const int lx = ix - ds.x0, ly = iy - ds.y0, lz = iz - ds.z0;
    if (lx >= 0 && ly >= 0 && lz >= 0 && lx + 1 < ds.bx && ly + 1 < ds.by && lz + 1 < ds.bz)
      return lx + ds.bx * (ly + ds.by * lz);
  }
  return -(int)gidx(g, ix, iy, iz) - 2;
}

// --------------------------------------------------------------------------
// Lane-contiguous particle runs: lane t of a block owns particles
// [tile + CPT*t, tile + CPT*t + CPT).
// --------------------------------------------------------------------------
constexpr int CPT = 4;
constexpr int CTILE = BLOCK * CPT;
static_assert(CPT == 4, "the vector accesses below move particles in pairs");

__device__ inline void load_pair(const double *a, int64_t i, double &v0, double &v1) {
  const double2 t = *reinterpret_cast<const double2 *>(a + i);
  v0 = t.x;
  v1 = t.y;
}

__device__ inline void store_pair(double *a, int64_t i, double v0, double v1) {
  *reinterpret_cast<double2 *>(a + i) = make_double2(v0, v1);
}

// Loads a lane's run; dead and out-of-range slots come back not live.
__device__ inline void load_run(const KPset &p, int64_t n, int64_t i0, bool full, double x[CPT],
                                double y[CPT], double z[CPT], double vx[CPT], double vy[CPT],
                                double vz[CPT], bool live[CPT]) {
  if (full) {
    const uint32_t d4 = *reinterpret_cast<const uint32_t *>(p.dead + i0);
    load_pair(p.x, i0, x[0], x[1]); load_pair(p.x, i0 + 2, x[2], x[3]);
    load_pair(p.y, i0, y[0], y[1]); load_pair(p.y, i0 + 2, y[2], y[3]);
    load_pair(p.z, i0, z[0], z[1]); load_pair(p.z, i0 + 2, z[2], z[3]);
    load_pair(p.vx, i0, vx[0], vx[1]); load_pair(p.vx, i0 + 2, vx[2], vx[3]);
    load_pair(p.vy, i0, vy[0], vy[1]); load_pair(p.vy, i0 + 2, vy[2], vy[3]);
    load_pair(p.vz, i0, vz[0], vz[1]); load_pair(p.vz, i0 + 2, vz[2], vz[3]);
#pragma unroll
    for (int q = 0; q < CPT; ++q) live[q] = ((d4 >> (8 * q)) & 0xFFu) == 0u;
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int64_t i = i0 + q;
      live[q] = i < n && p.dead[i] == 0;
      x[q] = y[q] = z[q] = vx[q] = vy[q] = vz[q] = 0.0;
      if (live[q]) {
        x[q] = p.x[i]; y[q] = p.y[i]; z[q] = p.z[i];
        vx[q] = p.vx[i]; vy[q] = p.vy[i]; vz[q] = p.vz[i];
      }
    }
  }
}

// Writes a lane's run back: survivors' new state, dead = 2 for the lost.
// Slots that were not live are written with the values they were read with.
__device__ inline void store_run(const KPset &p, int64_t n, int64_t i0, bool full,
                                 const double x[CPT], const double y[CPT], const double z[CPT],
                                 const double vx[CPT], const double vy[CPT], const double vz[CPT],
                                 const bool live[CPT], const bool lost[CPT]) {
  if (full) {
    store_pair(p.x, i0, x[0], x[1]); store_pair(p.x, i0 + 2, x[2], x[3]);
    store_pair(p.y, i0, y[0], y[1]); store_pair(p.y, i0 + 2, y[2], y[3]);
    store_pair(p.z, i0, z[0], z[1]); store_pair(p.z, i0 + 2, z[2], z[3]);
    store_pair(p.vx, i0, vx[0], vx[1]); store_pair(p.vx, i0 + 2, vx[2], vx[3]);
    store_pair(p.vy, i0, vy[0], vy[1]); store_pair(p.vy, i0 + 2, vy[2], vy[3]);
    store_pair(p.vz, i0, vz[0], vz[1]); store_pair(p.vz, i0 + 2, vz[2], vz[3]);
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int64_t i = i0 + q;
      if (live[q] && !lost[q]) {
        p.x[i] = x[q]; p.y[i] = y[q]; p.z[i] = z[q];
        p.vx[i] = vx[q]; p.vy[i] = vy[q]; p.vz[i] = vz[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (lost[q]) p.dead[i0 + q] = 2;
}

// In-register run of equal deposit keys.
struct RunAcc {
  int cur;
  double a[8];
  __device__ void reset() {
    cur = -1;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = 0.0;
  }
};

template <class DS>
__device__ inline void run_add(const KGrid &g, DS &ds, RunAcc &ra, int key, double k1,
                               double px, double py, double pz, double *rep) {
  if (key != ra.cur) {
    flush_run(g, ds, ra.cur, ra.a, rep);
    ra.reset();
    ra.cur = key;
  }
  const double wx2 = 1.0 - px, wy2 = 1.0 - py, wz2 = 1.0 - pz;
  ra.a[0] += k1 * px * py * pz;
  ra.a[1] += k1 * wx2 * py * pz;
  ra.a[2] += k1 * wx2 * wy2 * pz;
  ra.a[3] += k1 * px * wy2 * pz;
  ra.a[4] += k1 * px * py * wz2;
  ra.a[5] += k1 * wx2 * py * wz2;
  ra.a[6] += k1 * wx2 * wy2 * wz2;
  ra.a[7] += k1 * px * wy2 * wz2;
}

__global__ __launch_bounds__(BLOCK) void k_deposit(KGrid g, KPset p, int64_t n, double k1,
                                                   double *rep_base, int64_t nodes) {
  __shared__ DepShared ds;
  double *rep = rep_base + (int64_t)xcc_id() * nodes;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  double x[CPT], y[CPT], z[CPT];
  bool live[CPT];
  if (i0 + CPT <= n) {
    const uint32_t d4 = *reinterpret_cast<const uint32_t *>(p.dead + i0);
    load_pair(p.x, i0, x[0], x[1]); load_pair(p.x, i0 + 2, x[2], x[3]);
    load_pair(p.y, i0, y[0], y[1]); load_pair(p.y, i0 + 2, y[2], y[3]);
    load_pair(p.z, i0, z[0], z[1]); load_pair(p.z, i0 + 2, z[2], z[3]);
#pragma unroll
    for (int q = 0; q < CPT; ++q) live[q] = ((d4 >> (8 * q)) & 0xFFu) == 0u;
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int64_t i = i0 + q;
      live[q] = i < n && p.dead[i] == 0;
      x[q] = live[q] ? p.x[i] : 0.0;
      y[q] = live[q] ? p.y[i] : 0.0;
      z[q] = live[q] ? p.z[i] : 0.0;
    }
  }
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  int sm[3] = {0, 0, 0}, cnt = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (live[q]) {
      int c[3];
      cell_of(g, x[q], y[q], z[q], c);
      for (int d = 0; d < 3; ++d) {
        lo[d] = min(lo[d], c[d]); hi[d] = max(hi[d], c[d]); sm[d] += c[d];
      }
      ++cnt;
    }
  box_begin(g, ds, lo, hi, sm, cnt);
  RunAcc ra;
  ra.reset();
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    int ix, iy, iz;
    double px, py, pz;
    if (live[q] && deposit_fracs(g, x[q], y[q], z[q], ix, iy, iz, px, py, pz))
      run_add(g, ds, ra, stage_key(g, ds, ix, iy, iz), k1, px, py, pz, rep);
  }
  flush_run(g, ds, ra.cur, ra.a, rep);
  box_flush(g, ds, rep);
}

// np_local += sum of the replicas; replicas cleared for the next deposit.
__global__ void k_rep_flush(double *np, double *rep, int64_t nodes,
                            const double *__restrict__ kq) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < LEPIC_XCDS; ++r) {
    s = s + rep[(int64_t)r * nodes + i];
    rep[(int64_t)r * nodes + i] = 0.0;
  }
  np[i] = np[i] + kq[i] * s;
}

// The particle phase of a collision step, in place: gather + push + wall BC
// only (the deposit follows MCC), one particle per thread, no LDS.
template <int BM>
__global__ __launch_bounds__(BLOCK) void k_push_nodep(
    KGrid g, KPset p, int64_t n, const double4 *__restrict__ E4, KBField bf, double k1, double dt,
    const int *__restrict__ wt, KBC bp, double *p_mac, double *P_loss,
    unsigned long long *lost_ctr) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  bool lost = false;
  int label = 0;
  double vx = 0.0, vy = 0.0, vz = 0.0;
  if (i < n && p.dead[i] == 0) {
    double xp = p.x[i], yp = p.y[i], zp = p.z[i];
    vx = p.vx[i]; vy = p.vy[i]; vz = p.vz[i];
    Stencil s;
    make_stencil(g, xp, yp, zp, s);
    double Ex, Ey, Ez;
    gather3(E4, g, s, Ex, Ey, Ez);
    velocity_update<BM>(g, bf, s, k1, Ex, Ey, Ez, vx, vy, vz, xp, yp, zp);
    xp = xp + dt * vx; yp = yp + dt * vy; zp = zp + dt * vz;
    if (wall_test_tab(g, wt, bp, xp, yp, zp, label)) {
      p.dead[i] = 2;
      lost = true;
      if (bp.flag_die == 1 && label > 0) die_charge(g, bp, label, xp, yp, zp);
    } else {
      survivor_wrap(bp, xp, yp, zp, vx);
      p.x[i] = xp; p.y[i] = yp; p.z[i] = zp;
      p.vx[i] = vx; p.vy[i] = vy; p.vz[i] = vz;
    }
  }
  wall_diag_wave(bp, lost, label, vx, vy, vz, p_mac, P_loss);
  const uint64_t m = __ballot(lost);
  if (m && __lane_id() == __ffsll((unsigned long long)m) - 1)
    atomicAdd(lost_ctr, (unsigned long long)__popcll(m));
}

// Production particle phase, in place: gather + push + wall BC (+ deposit).
// Lost particles become holes (dead = 2) instead of being compacted; the
// next sort drops holes, keeping the survivors' relative order, which is the
// order the reference's per-step stable compaction produces.
//
// Each lane owns CPT consecutive particles, read and written with 16-byte
// vector accesses (a wave still covers one contiguous 2 KB span per array).
// In cell-sorted data a lane's particles are mostly in one cell, so the
// deposit sums them in registers and issues one set of 8 atomics per run.
template <int BM>
__global__ __launch_bounds__(BLOCK) void k_push(
    KGrid g, KPset p, int64_t n, const double4 *__restrict__ E4, KBField bf, double k1, double dt,
    const int *__restrict__ wt, KBC bp, double kdep,
    double *rep_base, int64_t nodes, int do_dep, double *p_mac, double *wdrep,
    unsigned long long *lost_ctr) {
  __shared__ DepSharedPush ds;
  __shared__ WallDiagShared wd;
  wd_begin(wd);
  __syncthreads();
  double *rep = rep_base ? rep_base + (int64_t)xcc_id() * nodes : nullptr;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  const bool full = i0 + CPT <= n;
  double x[CPT], y[CPT], z[CPT], vx[CPT], vy[CPT], vz[CPT];
  bool live[CPT];
  load_run(p, n, i0, full, x, y, z, vx, vy, vz, live);
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  if (do_dep) {
    int sm[3] = {0, 0, 0}, cnt = 0;
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (live[q]) {
        int c[3];
        cell_of(g, x[q], y[q], z[q], c);
        for (int d = 0; d < 3; ++d) {
          lo[d] = min(lo[d], c[d]); hi[d] = max(hi[d], c[d]); sm[d] += c[d];
        }
        ++cnt;
      }
    box_begin(g, ds, lo, hi, sm, cnt);
  }
  RunAcc ra;
  ra.reset();
  // E at the 8 nodes of the last gathered cell: in cell-sorted runs the next
  // particle is usually in the same cell and reuses them
  int64_t eb = -1;
  double4 f1, f2, f3, f4, f5, f6, f7, f8;
  bool lostq[CPT];
  int lost_local = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    bool lost = false;
    int label = 0;
    if (live[q]) {
      double xp = x[q], yp = y[q], zp = z[q];
      double ux = vx[q], uy = vy[q], uz = vz[q];
      Stencil s;
      make_stencil_pre(g, xp, yp, zp,
                       qdiv_coords_ok(xp, yp, zp), s);
      if (s.base != eb) {
        eb = s.base;
        const int64_t b1 = eb + g.sy, b2 = eb + g.sz, b3 = eb + g.sz + g.sy;
        f1 = E4[eb]; f2 = E4[eb + 1]; f3 = E4[b1 + 1]; f4 = E4[b1];
        f5 = E4[b2]; f6 = E4[b2 + 1]; f7 = E4[b3 + 1]; f8 = E4[b3];
      }
      // gather3's nodes, weights and summation order
      const double Ex = s.w[0] * f1.x + s.w[1] * f2.x + s.w[2] * f3.x + s.w[3] * f4.x +
                        s.w[4] * f5.x + s.w[5] * f6.x + s.w[6] * f7.x + s.w[7] * f8.x;
      const double Ey = s.w[0] * f1.y + s.w[1] * f2.y + s.w[2] * f3.y + s.w[3] * f4.y +
                        s.w[4] * f5.y + s.w[5] * f6.y + s.w[6] * f7.y + s.w[7] * f8.y;
      const double Ez = s.w[0] * f1.z + s.w[1] * f2.z + s.w[2] * f3.z + s.w[3] * f4.z +
                        s.w[4] * f5.z + s.w[5] * f6.z + s.w[6] * f7.z + s.w[7] * f8.z;
      velocity_update<BM>(g, bf, s, k1, Ex, Ey, Ez, ux, uy, uz, xp, yp, zp);
      xp = xp + dt * ux; yp = yp + dt * uy; zp = zp + dt * uz;
      vx[q] = ux; vy[q] = uy; vz[q] = uz;
      // one set of quotients of the new position serves the wall test and
      // the deposit (formed again only if the periodic wrap moved it)
      bool okn = qdiv_coords_ok(xp, yp, zp);
      double qx = qdiv_pre(xp, g.q1, okn), qy = qdiv_pre(yp, g.q2, okn),
             qz = qdiv_pre(zp, g.q3, okn);
      if (wall_test_tab_q(g, wt, bp, xp, qx, qy, qz, label)) {
        lost = true;
        ++lost_local;
        if (bp.flag_die == 1 && label > 0) die_charge(g, bp, label, xp, yp, zp);
      } else {
        const double xw = xp, yw = yp, zw = zp;
        survivor_wrap(bp, xp, yp, zp, vx[q]);
        if (xp != xw || yp != yw || zp != zw) {
          okn = qdiv_coords_ok(xp, yp, zp);
          qx = qdiv_pre(xp, g.q1, okn); qy = qdiv_pre(yp, g.q2, okn); qz = qdiv_pre(zp, g.q3, okn);
        }
        x[q] = xp; y[q] = yp; z[q] = zp;
        int ix, iy, iz;
        double px, py, pz;
        if (do_dep && deposit_fracs_q(g, xp, yp, zp, qx, qy, qz, okn, ix, iy, iz, px, py, pz))
          run_add(g, ds, ra, stage_key(g, ds, ix, iy, iz), kdep, px, py, pz, rep);
      }
    }
    if (lost) wd_lost(wd, bp, label, vx[q], vy[q], vz[q], p_mac);
    lostq[q] = lost;
  }
  store_run(p, n, i0, full, x, y, z, vx, vy, vz, live, lostq);
  if (do_dep) {
    flush_run(g, ds, ra.cur, ra.a, rep);
    box_flush(g, ds, rep);
  }
  wd_finish(wd, wdrep);
  // losses are rare: one atomic per wave that saw any
  int tot = lost_local;
  for (int d = 32; d > 0; d >>= 1) tot += __shfl_xor(tot, d);
  if (__lane_id() == 0 && tot > 0) atomicAdd(lost_ctr, (unsigned long long)tot);
}

// --------------------------------------------------------------------------
// Fused particle phase: gather + push + wall BC + stable compaction + deposit.
// --------------------------------------------------------------------------
template <int BM>
__global__ __launch_bounds__(BLOCK) void k_push_bc_deposit(
    KGrid g, KPset src, KPset dst, int64_t n, const double4 *__restrict__ E4, KBField bf,
    double k1, double dt, const int *__restrict__ bcnd, KBC bp,
    double kdep, double *rep_base, int64_t nodes, double *p_mac, double *P_loss,
    uint64_t *status, uint32_t epoch, int *counter, int64_t *spin_fail, int64_t *count_out,
    int64_t ntiles) {
  __shared__ LBShared sh;
  __shared__ DepShared ds;
  int64_t tile = take_tile(counter, ntiles, sh);
  int64_t i = tile * BLOCK + threadIdx.x;
  bool keep = false, lost = false;
  int label = 0;
  double xp = 0, yp = 0, zp = 0, vx = 0, vy = 0, vz = 0;
  if (i < n && src.dead[i] != 1) {
    xp = src.x[i]; yp = src.y[i]; zp = src.z[i];
    vx = src.vx[i]; vy = src.vy[i]; vz = src.vz[i];
    Stencil s;
    make_stencil(g, xp, yp, zp, s);
    double Ex, Ey, Ez;
    gather3(E4, g, s, Ex, Ey, Ez);
    velocity_update<BM>(g, bf, s, k1, Ex, Ey, Ez, vx, vy, vz, xp, yp, zp);
    xp = xp + dt * vx; yp = yp + dt * vy; zp = zp + dt * vz;
    lost = wall_test(g, bcnd, bp, xp, yp, zp, label) != 0;
    if (!lost) {
      survivor_wrap(bp, xp, yp, zp, vx);
      keep = true;
    } else if (bp.flag_die == 1 && label > 0) {
      die_charge(g, bp, label, xp, yp, zp);
    }
  }
  wall_diag_wave(bp, lost, label, vx, vy, vz, p_mac, P_loss);
  int64_t o = compact_slot(keep, sh, status, tile, epoch, spin_fail, count_out, ntiles);
  if (keep) {
    dst.x[o] = xp; dst.y[o] = yp; dst.z[o] = zp;
    dst.vx[o] = vx; dst.vy[o] = vy; dst.vz[o] = vz;
    dst.dead[o] = 0;
  }
  if (rep_base) {
    double *rep = rep_base + (int64_t)xcc_id() * nodes;
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    if (keep) {
      cell_of(g, xp, yp, zp, lo);
      hi[0] = lo[0]; hi[1] = lo[1]; hi[2] = lo[2];
    }
    box_begin(g, ds, lo, hi, lo, keep ? 1 : 0);
    bool has = false;
    int ix = 0, iy = 0, iz = 0;
    double w[8];
    if (keep) has = deposit_weights(g, kdep, xp, yp, zp, ix, iy, iz, w);
    box_add(g, ds, has, ix, iy, iz, w, rep);
    box_flush(g, ds, rep);
  }
}

// --------------------------------------------------------------------------
// a2: cell keys (cell_index_from_position, mod_particle_sorting.f90:16-48)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_cell_keys(KGrid g, KPset p, int64_t n,
                                                     uint32_t *keys, uint32_t *vals,
                                                     int32_t *count, int drop_dead,
                                                     unsigned long long *ndead) {
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  if (drop_dead) {
    const bool d = p.dead[i] != 0;
    const uint64_t m = __ballot(d);
    if (d) {
      // holes sort behind every cell; one atomic per wave counts them
      keys[i] = (uint32_t)g.n1 * (uint32_t)g.n2 * (uint32_t)g.n3;
      vals[i] = (uint32_t)i;
      if (__lane_id() == __ffsll((unsigned long long)m) - 1)
        atomicAdd(ndead, (unsigned long long)__popcll(m));
      return;
    }
  }
  int ix = (int)qdiv(p.x[i], g.q1) + 1;
  int iy = (int)qdiv(p.y[i], g.q2) + 1;
  int iz = (int)qdiv(p.z[i], g.q3) + 1;
  ix = imax(1, imin(g.n1, ix));
  iy = imax(1, imin(g.n2, iy));
  iz = imax(1, imin(g.n3, iz));
  uint32_t c = (uint32_t)((ix - 1) + g.n1 * ((iy - 1) + g.n2 * (iz - 1)));
  keys[i] = c;
  vals[i] = (uint32_t)i;
  if (count) atomicAdd(&count[c], 1);
}

__global__ __launch_bounds__(BLOCK) void k_gather_sorted(KPset src, KPset dst, int64_t n,
                                                         const uint32_t *__restrict__ keys,
                                                         const uint32_t *__restrict__ perm) {
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  int64_t j = perm[i];
  dst.x[i] = src.x[j]; dst.y[i] = src.y[j]; dst.z[i] = src.z[j];
  dst.vx[i] = src.vx[j]; dst.vy[i] = src.vy[j]; dst.vz[i] = src.vz[j];
  dst.cell[i] = (int32_t)keys[i];
  dst.dead[i] = src.dead[j];
}

// --------------------------------------------------------------------------
// Counting sort by cell (production re-sort, holes dropped, order within a
// cell unspecified).  Lane t of a block owns particles tile + 4t .. 4t+3, as
// in k_push.  A tile's cells form a small box (particles drift less than a
// cell per step since the last sort), so the block counts its particles per
// cell in LDS and touches each of the tile's cells once in global memory:
// one atomic per (tile, cell) instead of one per particle, and the particles
// of a cell leave the tile as one run of consecutive slots.  Tiles whose box
// exceeds BIN_BOX cells (none in sorted data) take one atomic per particle.
// --------------------------------------------------------------------------
constexpr int BIN_BOX = 8192;   // cells (int counters, 32 KB)

struct BinShared {
  int mn[3], mx[3];
  int x0, y0, z0, bx, bxy, vol;   // vol < 0: per-particle atomics
  int cnt[BIN_BOX];
};

// 1-based cell of a live particle, clamped to [1, n] (cell_index_from_position,
// mod_particle_sorting.f90:16-48)
__device__ inline void bin_cell(const KGrid &g, double x, double y, double z, int c[3]) {
  const bool ok = qdiv_coords_ok(x, y, z);
  c[0] = imax(1, imin(g.n1, (int)qdiv_pre(x, g.q1, ok) + 1));
  c[1] = imax(1, imin(g.n2, (int)qdiv_pre(y, g.q2, ok) + 1));
  c[2] = imax(1, imin(g.n3, (int)qdiv_pre(z, g.q3, ok) + 1));
}

__device__ inline uint32_t bin_key(const KGrid &g, const int c[3]) {
  return (uint32_t)((c[0] - 1) + g.n1 * ((c[1] - 1) + g.n2 * (c[2] - 1)));
}

// The block's cell box from its lanes' cells; every thread calls it.
__device__ inline void bin_box(BinShared &bs, const int lo[3], const int hi[3]) {
  if (threadIdx.x == 0) {
    bs.mn[0] = bs.mn[1] = bs.mn[2] = 1 << 30;
    bs.mx[0] = bs.mx[1] = bs.mx[2] = -(1 << 30);
  }
  __syncthreads();
  const int a0 = wave_min(lo[0]), a1 = wave_min(lo[1]), a2 = wave_min(lo[2]);
  const int b0 = wave_max(hi[0]), b1 = wave_max(hi[1]), b2 = wave_max(hi[2]);
  if (__lane_id() == 0) {
    atomicMin(&bs.mn[0], a0); atomicMin(&bs.mn[1], a1); atomicMin(&bs.mn[2], a2);
    atomicMax(&bs.mx[0], b0); atomicMax(&bs.mx[1], b1); atomicMax(&bs.mx[2], b2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int vol = 0;
    if (bs.mx[0] >= bs.mn[0]) {
      bs.x0 = bs.mn[0]; bs.y0 = bs.mn[1]; bs.z0 = bs.mn[2];
      bs.bx = bs.mx[0] - bs.mn[0] + 1;
      const int by = bs.mx[1] - bs.mn[1] + 1, bz = bs.mx[2] - bs.mn[2] + 1;
      bs.bxy = bs.bx * by;
      const int64_t v = (int64_t)bs.bxy * bz;
      vol = v > BIN_BOX ? -1 : (int)v;
    }
    bs.vol = vol;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < bs.vol; t += BLOCK) bs.cnt[t] = 0;
  __syncthreads();
}

__device__ inline int bin_local(const BinShared &bs, const int c[3]) {
  return (c[0] - bs.x0) + bs.bx * (c[1] - bs.y0) + bs.bxy * (c[2] - bs.z0);
}

__device__ inline uint32_t bin_key_of_local(const KGrid &g, const BinShared &bs, int t) {
  const int lz = t / bs.bxy, r = t - lz * bs.bxy;
  const int ly = r / bs.bx, lx = r - ly * bs.bx;
  const int c[3] = {bs.x0 + lx, bs.y0 + ly, bs.z0 + lz};
  return bin_key(g, c);
}

// Loads positions and the live mask of a lane's run.
__device__ inline void load_pos_run(const KPset &p, int64_t n, int64_t i0, bool full,
                                    double x[CPT], double y[CPT], double z[CPT], bool live[CPT]) {
  if (full) {
    const uint32_t d4 = *reinterpret_cast<const uint32_t *>(p.dead + i0);
    load_pair(p.x, i0, x[0], x[1]); load_pair(p.x, i0 + 2, x[2], x[3]);
    load_pair(p.y, i0, y[0], y[1]); load_pair(p.y, i0 + 2, y[2], y[3]);
    load_pair(p.z, i0, z[0], z[1]); load_pair(p.z, i0 + 2, z[2], z[3]);
#pragma unroll
    for (int q = 0; q < CPT; ++q) live[q] = ((d4 >> (8 * q)) & 0xFFu) == 0u;
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int64_t i = i0 + q;
      live[q] = i < n && p.dead[i] == 0;
      x[q] = y[q] = z[q] = 0.0;
      if (live[q]) { x[q] = p.x[i]; y[q] = p.y[i]; z[q] = p.z[i]; }
    }
  }
}

// Fixed-point deposit.  Every particle carries the same full weight U =
// rint(|k1| 2^frac_bits) units (k1 = Nm / (h1 h2 h3)), split over its 8
// stencil nodes by fixed_split with the trilinear fractions of
// deposit_particle_set_to_np_thread (mod_chargeDeposition.f90:90-131): the
// parts are integers summing to U exactly, so the charge on the grid is N U
// units for N deposited particles, and int64 node sums do not depend on the
// order in which particles arrive.  A lane sums its run of one cell in
// registers and adds it to the block's LDS box (or, outside it, to this
// XCD's int64 replica); two's complement wrap-around makes unsigned adds of
// negative values exact.
// The block's cells +-1 get an LDS box of int64 sums (the FP64 deposit's
// box, 32 KB); stencils outside it go to the replica directly.
struct DepSharedFixed {
  static constexpr int CAP = LDS_BOX;
  static constexpr int CUT_X = 36;
  static constexpr int CUT_YZ = 10;
  int mn[3], mx[3];
  int x0, y0, z0, bx, by, bz, vol;
  uint64_t mbx, mby;   // small_div multipliers of bx and by
  unsigned long long box[LDS_BOX];
};

__device__ inline void flush_fixed(const KGrid &g, DepSharedFixed &ds, int key,
                                   const long long a[8], unsigned long long *rep) {
  if (key >= 0) {
    const int bx = ds.bx, bxy = ds.bx * ds.by;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (a[q] != 0) atomicAdd(&ds.box[key + node_off(q, bx, bxy)], (unsigned long long)a[q]);
  } else if (key <= -2) {
    const int64_t b = (int64_t)(-(key + 2));
    const int64_t nd[8] = {b, b + 1, b + g.sy + 1, b + g.sy, b + g.sz, b + g.sz + 1,
                           b + g.sz + g.sy + 1, b + g.sz + g.sy};
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (a[q] != 0)
        __hip_atomic_fetch_add(&rep[nd[q]], (unsigned long long)a[q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// U: a full weight in units (rint(|k1| 2^S)); neg: k1 < 0.
__global__ __launch_bounds__(BLOCK) void k_deposit_fixed(KGrid g, KPset p, int64_t n, double U,
                                                         int neg,
                                                         unsigned long long *rep_base,
                                                         int64_t nodes) {
  __shared__ DepSharedFixed ds;
  unsigned long long *rep = rep_base + (int64_t)xcc_id() * nodes;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  const bool full = i0 + CPT <= n;
  double x[CPT], y[CPT], z[CPT];
  bool live[CPT];
  load_pos_run(p, n, i0, full, x, y, z, live);
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  int sm[3] = {0, 0, 0}, cnt = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (live[q]) {
      int c[3];
      cell_of(g, x[q], y[q], z[q], c);
      for (int d = 0; d < 3; ++d) {
        lo[d] = min(lo[d], c[d]); hi[d] = max(hi[d], c[d]); sm[d] += c[d];
      }
      ++cnt;
    }
  box_begin(g, ds, lo, hi, sm, cnt);
  int cur = -1;
  long long a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    int ix, iy, iz;
    double px, py, pz;
    if (!(live[q] && deposit_fracs(g, x[q], y[q], z[q], ix, iy, iz, px, py, pz))) continue;
    const int key = stage_key(g, ds, ix, iy, iz);
    if (key != cur) {
      flush_fixed(g, ds, cur, a, rep);
      cur = key;
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] = 0;
    }
    const double wx2 = 1.0 - px, wy2 = 1.0 - py, wz2 = 1.0 - pz;
    long long parts[8];
    fixed_split(U, wx2, wy2, wz2, parts);
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] += neg ? -parts[j] : parts[j];
  }
  flush_fixed(g, ds, cur, a, rep);
  __syncthreads();
  for (int t = threadIdx.x; t < ds.vol; t += BLOCK) {
    const unsigned long long v = ds.box[t];
    if (v != 0) {
      const int r = (int)small_div((uint32_t)t, ds.mbx), lz = (int)small_div((uint32_t)r, ds.mby);
      const int lx = t - r * ds.bx, ly = r - lz * ds.by;
      const int64_t node = gidx(g, ds.x0 + lx, ds.y0 + ly, ds.z0 + lz);
      __hip_atomic_fetch_add(&rep[node], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
}

// np += kq * (sum of the replicas) * 2^-frac_bits, clearing the replicas.
__global__ void k_fixed_flush(double *np, unsigned long long *rep, int64_t nodes,
                              const double *__restrict__ kq, double inv_scale) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  unsigned long long s = 0;
#pragma unroll
  for (int r = 0; r < LEPIC_XCDS; ++r) {
    s += rep[(int64_t)r * nodes + i];
    rep[(int64_t)r * nodes + i] = 0;
  }
  np[i] = np[i] + kq[i] * ((double)(long long)s * inv_scale);
}

__global__ __launch_bounds__(BLOCK) void k_bin_count(KGrid g, KPset p, int64_t n,
                                                     int32_t *count) {
  __shared__ BinShared bs;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  const bool full = i0 + CPT <= n;
  double x[CPT], y[CPT], z[CPT];
  bool live[CPT];
  load_pos_run(p, n, i0, full, x, y, z, live);
  int c[CPT][3];
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (live[q]) {
      bin_cell(g, x[q], y[q], z[q], c[q]);
      for (int d = 0; d < 3; ++d) { lo[d] = min(lo[d], c[q][d]); hi[d] = max(hi[d], c[q][d]); }
    }
  bin_box(bs, lo, hi);
  if (bs.vol > 0) {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (live[q]) atomicAdd(&bs.cnt[bin_local(bs, c[q])], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < bs.vol; t += BLOCK) {
      const int k = bs.cnt[t];
      if (k) atomicAdd(&count[bin_key_of_local(g, bs, t)], k);
    }
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (live[q]) atomicAdd(&count[bin_key(g, c[q])], 1);
  }
}

// cursor[c] starts at the first slot of cell c and is advanced by the tiles.
// The tile's particles are written through LDS one array at a time, in the
// order of their destinations (cells of the box in order, each cell's run
// contiguous), so that neighbouring lanes store to neighbouring addresses.
constexpr int BIN_STAGE = 4096;   // cells of the staged path (2 int arrays, 32 KB)

struct BinStage {
  int mn[3], mx[3];
  int x0, y0, z0, bx, bxy, vol;    // vol < 0: per-particle atomics, direct stores
  int wsum[NWAVES];
  int total;
  int cnt[BIN_STAGE];    // tile count per cell, then first staged slot of the cell
  int dlt[BIN_STAGE];    // destination of the cell's run minus its first slot
  int dsti[CTILE];       // destination of each staged slot
  double buf[CTILE];     // one array of the tile at a time
};

__global__ __launch_bounds__(BLOCK) void k_bin_scatter(KGrid g, KPset src, KPset dst, int64_t n,
                                                       int32_t *cursor) {
  __shared__ BinStage bs;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  const bool full = i0 + CPT <= n;
  double x[CPT], y[CPT], z[CPT], vx[CPT], vy[CPT], vz[CPT];
  bool live[CPT];
  load_run(src, n, i0, full, x, y, z, vx, vy, vz, live);
  int c[CPT][3];
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (live[q]) {
      bin_cell(g, x[q], y[q], z[q], c[q]);
      for (int d = 0; d < 3; ++d) { lo[d] = min(lo[d], c[q][d]); hi[d] = max(hi[d], c[q][d]); }
    }
  // the block's cell box
  if (threadIdx.x == 0) {
    bs.mn[0] = bs.mn[1] = bs.mn[2] = 1 << 30;
    bs.mx[0] = bs.mx[1] = bs.mx[2] = -(1 << 30);
  }
  __syncthreads();
  {
    const int a0 = wave_min(lo[0]), a1 = wave_min(lo[1]), a2 = wave_min(lo[2]);
    const int b0 = wave_max(hi[0]), b1 = wave_max(hi[1]), b2 = wave_max(hi[2]);
    if (__lane_id() == 0) {
      atomicMin(&bs.mn[0], a0); atomicMin(&bs.mn[1], a1); atomicMin(&bs.mn[2], a2);
      atomicMax(&bs.mx[0], b0); atomicMax(&bs.mx[1], b1); atomicMax(&bs.mx[2], b2);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int vol = 0;
    if (bs.mx[0] >= bs.mn[0]) {
      bs.x0 = bs.mn[0]; bs.y0 = bs.mn[1]; bs.z0 = bs.mn[2];
      bs.bx = bs.mx[0] - bs.mn[0] + 1;
      const int by = bs.mx[1] - bs.mn[1] + 1, bz = bs.mx[2] - bs.mn[2] + 1;
      bs.bxy = bs.bx * by;
      const int64_t v = (int64_t)bs.bxy * bz;
      vol = v > BIN_STAGE ? -1 : (int)v;
    }
    bs.vol = vol;
  }
  __syncthreads();
  const int vol = bs.vol;
  if (vol < 0) {   // wide tile: one atomic per particle, stores straight to dst
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (live[q]) {
        const uint32_t key = bin_key(g, c[q]);
        const int64_t d = atomicAdd(&cursor[key], 1);
        dst.x[d] = x[q]; dst.y[d] = y[q]; dst.z[d] = z[q];
        dst.vx[d] = vx[q]; dst.vy[d] = vy[q]; dst.vz[d] = vz[q];
        dst.cell[d] = (int32_t)key;
        dst.dead[d] = 0;
      }
    return;   // uniform across the block
  }
  for (int t = threadIdx.x; t < vol; t += BLOCK) bs.cnt[t] = 0;
  __syncthreads();
  int l[CPT], rank[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    l[q] = live[q] ? (c[q][0] - bs.x0) + bs.bx * (c[q][1] - bs.y0) + bs.bxy * (c[q][2] - bs.z0) : 0;
    rank[q] = live[q] ? atomicAdd(&bs.cnt[l[q]], 1) : 0;
  }
  __syncthreads();
  // exclusive scan of the counts over the box (thread t owns a chunk), one
  // global atomic per nonempty cell for its run of destinations
  {
    const int ch = (vol + BLOCK - 1) / BLOCK;
    const int b = min((int)threadIdx.x * ch, vol), e = min(b + ch, vol);
    int sum = 0;
    for (int t = b; t < e; ++t) sum += bs.cnt[t];
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    int inc = sum;
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(inc, d);
      if (lane >= d) inc += o;
    }
    if (lane == 63) bs.wsum[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
      int run = 0;
      for (int i = 0; i < NWAVES; ++i) {
        const int v = bs.wsum[i];
        bs.wsum[i] = run;
        run += v;
      }
      bs.total = run;
    }
    __syncthreads();
    int pre = bs.wsum[w] + inc - sum;
    for (int t = b; t < e; ++t) {
      const int k = bs.cnt[t];
      int base = 0;
      if (k) {
        const int lz = t / bs.bxy, r = t - lz * bs.bxy;
        const int ly = r / bs.bx, lx = r - ly * bs.bx;
        const int cc[3] = {bs.x0 + lx, bs.y0 + ly, bs.z0 + lz};
        base = atomicAdd(&cursor[bin_key(g, cc)], k);
      }
      bs.cnt[t] = pre;
      bs.dlt[t] = base - pre;
      pre += k;
    }
  }
  __syncthreads();
  const int total = bs.total;
  int slot[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    slot[q] = live[q] ? bs.cnt[l[q]] + rank[q] : 0;
    if (live[q]) bs.dsti[slot[q]] = slot[q] + bs.dlt[l[q]];
  }
  // one array at a time through LDS, stored in destination order
  for (int f = 0; f < 7; ++f) {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (live[q]) {
        double v;
        switch (f) {
          case 0: v = x[q]; break;
          case 1: v = y[q]; break;
          case 2: v = z[q]; break;
          case 3: v = vx[q]; break;
          case 4: v = vy[q]; break;
          case 5: v = vz[q]; break;
          default: v = __longlong_as_double((long long)bin_key(g, c[q]));
        }
        bs.buf[slot[q]] = v;
      }
    __syncthreads();
    double *out = f == 0 ? dst.x : f == 1 ? dst.y : f == 2 ? dst.z : f == 3 ? dst.vx
                : f == 4 ? dst.vy : dst.vz;
    for (int s = threadIdx.x; s < total; s += BLOCK) {
      const int d = bs.dsti[s];
      if (f < 6) {
        out[d] = bs.buf[s];
      } else {
        dst.cell[d] = (int32_t)__double_as_longlong(bs.buf[s]);
        dst.dead[d] = 0;
      }
    }
    __syncthreads();
  }
}

__global__ void k_add_u64(unsigned long long *acc, const unsigned long long *v) {
  if (threadIdx.x == 0) *acc = *acc + *v;
}

__global__ void k_fill_i32(int32_t *a, int64_t n, int32_t v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// --------------------------------------------------------------------------
// a13: device loader -- np_cell particles in every loadable cell, uniform
// within the cell, Maxwellian velocities by Box-Muller as load_gauss
// (mod_particle_loader.f90:249-260), Philox-4x32-10 draws.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_load_uniform(KGrid g, KPset p, int64_t n,
                                                        const int32_t *__restrict__ cells,
                                                        int np_cell, double vt, uint64_t key,
                                                        uint32_t species) {
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  const double pi = 3.14159265358979311600;
  int32_t c = cells[i / np_cell];
  int ix = c % g.n1, r = c / g.n1, iy = r % g.n2, iz = r / g.n2;
  // positions: counter (i, 0, 0xC0FFEE) -- shared by all species
  philox4 a = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 0u, 0xC0FFEEu, (uint32_t)key,
                            (uint32_t)(key >> 32));
  philox4 b = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 1u, 0xC0FFEEu, (uint32_t)key,
                            (uint32_t)(key >> 32));
  double ux = u01_53(a.v[0], a.v[1]), uy = u01_53(a.v[2], a.v[3]), uz = u01_53(b.v[0], b.v[1]);
  p.x[i] = ((double)ix + ux) * g.h1;
  p.y[i] = ((double)iy + uy) * g.h2;
  p.z[i] = ((double)iz + uz) * g.h3;
  philox4 v1 = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 2u + species, 0xBEEFu,
                             (uint32_t)key, (uint32_t)(key >> 32));
  philox4 v2 = philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), 2u + species, 0xFACEu,
                             (uint32_t)key, (uint32_t)(key >> 32));
  double r1 = u01_53(v1.v[0], v1.v[1]), r2 = u01_53(v1.v[2], v1.v[3]);
  double r3 = u01_53(v2.v[0], v2.v[1]), r4 = u01_53(v2.v[2], v2.v[3]);
  double vp = vt * sqrt(-log(1.0 - r1));
  double th = 2.0 * pi * r2;
  p.vx[i] = vp * cos(th);
  p.vy[i] = vp * sin(th);
  double vq = vt * sqrt(-log(1.0 - r3));
  p.vz[i] = vq * cos(2.0 * pi * r4);
  p.dead[i] = 0;
  p.cell[i] = c;
}

int check_pset(lepic_ctx *ctx, const lepic_pset *p, const char *what) {
  if (!p) return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(what) + ": null particle set");
  if (p->n < 0 || p->n > p->cap)
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(what) + ": n outside [0, cap]");
  if (p->n > 0 && (!p->x || !p->y || !p->z || !p->vx || !p->vy || !p->vz || !p->dead))
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(what) + ": null particle array");
  if (p->n > ((int64_t)1 << 40))
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(what) + ": too many particles");
  return LEPIC_OK;
}

// The run kernels read and write particles with 16-byte vector accesses and
// the dead flags 4 at a time.
static bool aligned_runs(const lepic_pset *p) {
  const void *a[6] = {p->x, p->y, p->z, p->vx, p->vy, p->vz};
  for (const void *q : a)
    if (reinterpret_cast<uintptr_t>(q) % 16 != 0) return false;
  return reinterpret_cast<uintptr_t>(p->dead) % 4 == 0;
}


// Run a compaction kernel launch: sets up the look-back epoch, reads back
// the output count and the spin-failure flag.
template <typename Launch>
int run_compaction(lepic_ctx *ctx, int64_t n, Launch launch, int64_t *n_out) {
  int64_t ntiles = (n + BLOCK - 1) / BLOCK;
  if (ntiles == 0) {
    *n_out = 0;
    return LEPIC_OK;
  }
  uint32_t epoch = lepic_next_epoch(ctx, ntiles);
  if (!ctx->d_status) return lepic_fail(ctx, LEPIC_ERR_HIP, "look-back status allocation failed");
  int *counter = reinterpret_cast<int *>(ctx->d_ctrl + CTRL_TILE_COUNTER);
  int64_t *count_out = ctx->d_ctrl + CTRL_COUNT_OUT;
  int64_t *spin = ctx->d_ctrl + CTRL_SPIN_FAIL;
  launch(dim3((unsigned)ntiles), epoch, counter, spin, count_out, ntiles);
  LEPIC_LAUNCH_CHECK(ctx, "compaction kernel launch");
  LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, ctx->d_ctrl, sizeof(int64_t) * 32,
                                hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->h_pinned[CTRL_SPIN_FAIL] != 0) {
    int64_t zero = 0;
    (void)hipMemcpy(spin, &zero, sizeof(zero), hipMemcpyHostToDevice);
    return lepic_fail(ctx, LEPIC_ERR_SPIN, "compaction look-back wait exceeded its bound");
  }
  *n_out = ctx->h_pinned[CTRL_COUNT_OUT];
  return LEPIC_OK;
}

}  // namespace

double *lepic_replicas(lepic_ctx *ctx, int64_t nodes) {
  if (ctx->rep && ctx->rep_nodes >= nodes) return ctx->rep;
  if (ctx->rep) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->rep);
    ctx->rep = nullptr;
    ctx->rep_nodes = 0;
  }
  size_t bytes = sizeof(double) * (size_t)LEPIC_XCDS * (size_t)nodes;
  if (hipMalloc((void **)&ctx->rep, bytes) != hipSuccess) return nullptr;
  if (hipMemsetAsync(ctx->rep, 0, bytes, ctx->stream) != hipSuccess) return nullptr;
  ctx->rep_nodes = nodes;
  return ctx->rep;
}

int lepic_replica_flush(lepic_ctx *ctx, int64_t nodes, double *np_local, const double *kq) {
  // the deposit kernels stride the replicas by this grid's node count; the
  // buffer stays all-zero between calls because every written entry is
  // cleared here
  if (!ctx->rep || nodes > ctx->rep_nodes)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "replica flush: replicas smaller than the grid");
  hipLaunchKernelGGL(k_rep_flush, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     np_local, ctx->rep, nodes, kq);
  LEPIC_LAUNCH_CHECK(ctx, "k_rep_flush");
  return LEPIC_OK;
}

// A map on its own grid needs a sane grid.
static int check_bgrid(lepic_ctx *ctx, const lepic_bfield *bf, const char *who) {
  if (!bf || bf->mode != 3) return LEPIC_OK;
  for (int d = 0; d < 3; ++d)
    if (bf->nB[d] < 1 || !qdiv_divisor_ok(bf->hB[d]))
      return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": bad B grid");
  return LEPIC_OK;
}

// Dielectric charging needs its tables (flag_die = 1).
static int check_die(lepic_ctx *ctx, const lepic_bc_params *bp, const char *who) {
  if (bp->flag_die != 1) return LEPIC_OK;
  if (!bp->dtype || !bp->sum_q_xz || !bp->sum_q_yz || bp->ngrid < 0)
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      std::string(who) + ": flag_die needs dtype, sum_q_xz and sum_q_yz");
  return LEPIC_OK;
}

extern "C" {

int lepic_move_electrostatic(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p,
                             const double *E4, double q, double m, double dt) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, p, "move_electrostatic"))) return st;
  if (!E4) return lepic_fail(ctx, LEPIC_ERR_ARG, "move_electrostatic: null E");
  if (p->n == 0) return LEPIC_OK;
  double qmdt = dt * q / m;
  hipLaunchKernelGGL(k_move_es, dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, reinterpret_cast<const double4 *>(E4), qmdt, dt);
  LEPIC_LAUNCH_CHECK(ctx, "k_move_es");
  return LEPIC_OK;
}

int lepic_move_boris(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p, const double *E4,
                     const lepic_bfield *bf, double q, double m, double dt) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, p, "move_boris"))) return st;
  if (!E4) return lepic_fail(ctx, LEPIC_ERR_ARG, "move_boris: null E");
  if (bf && (bf->mode == 2 || bf->mode == 3) && !bf->B4)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "move_boris: null B map");
  if ((st = check_bgrid(ctx, bf, "move_boris"))) return st;
  if (bf && (bf->mode < 0 || bf->mode > 3)) return lepic_fail(ctx, LEPIC_ERR_ARG, "move_boris: bad mode");
  if (p->n == 0) return LEPIC_OK;
  double k1 = dt * q / (2.0 * m);
  const int bmode = bf ? bf->mode : 0;
  switch (bmode) {
    case 0: hipLaunchKernelGGL((k_move_boris<0>), dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                     make_kbfield(bf, k1), k1, dt); break;
    case 1: hipLaunchKernelGGL((k_move_boris<1>), dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                     make_kbfield(bf, k1), k1, dt); break;
    default: hipLaunchKernelGGL((k_move_boris<2>), dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                     make_kbfield(bf, k1), k1, dt); break;
  }
  LEPIC_LAUNCH_CHECK(ctx, "k_move_boris");
  return LEPIC_OK;
}

int lepic_apply_bc(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src, lepic_pset *dst,
                   const int32_t *bcnd, const lepic_bc_params *bp, double *p_mac,
                   double *P_loss, int64_t *np_lost) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, src, "apply_bc src")) ||
      (st = check_pset(ctx, dst, "apply_bc dst")))
    return st;
  if (!bcnd || !bp || !p_mac || !P_loss)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "apply_bc: null argument");
  if ((st = check_die(ctx, bp, "apply_bc"))) return st;
  if (dst->cap < src->n) return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "apply_bc: dst too small");
  if (src == dst || src->x == dst->x)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "apply_bc: src and dst must differ");
  int64_t n = src->n, nout = 0;
  KGrid kg = make_kgrid(g);
  KBC kb = kbc(bp);
  st = run_compaction(
      ctx, n,
      [&](dim3 grid, uint32_t epoch, int *counter, int64_t *spin, int64_t *count_out,
          int64_t ntiles) {
        hipLaunchKernelGGL(k_apply_bc, grid, dim3(BLOCK), 0, ctx->stream, kg, kp(src), kp(dst),
                           n, bcnd, kb, p_mac, P_loss, ctx->d_status, epoch, counter, spin,
                           count_out, ntiles);
      },
      &nout);
  if (st) return st;
  dst->n = nout;
  if (np_lost) *np_lost = n - nout;
  return LEPIC_OK;
}

int lepic_deposit(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *p, const double *kq,
                  double Nm, double *np_local) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, p, "deposit"))) return st;
  if (!aligned_runs(p))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit: particle arrays must be 16-byte aligned");
  if (!kq || !np_local) return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit: null argument");
  if (p->n == 0) return LEPIC_OK;
  const int64_t nodes = grid_nodes(g);
  double *rep = lepic_replicas(ctx, nodes);
  if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "deposit: replica allocation failed");
  double k1 = Nm / (g->h[0] * g->h[1] * g->h[2]);
  hipLaunchKernelGGL(k_deposit, dim3(grid_blocks(p->n, CTILE)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, k1, rep, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_deposit");
  return lepic_replica_flush(ctx, nodes, np_local, kq);
}

int lepic_deposit_fixed(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *p,
                        const double *kq, double Nm, int32_t frac_bits, double *np_local) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, p, "deposit_fixed"))) return st;
  if (!aligned_runs(p))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit_fixed: particle arrays must be 16-byte aligned");
  if (!kq || !np_local) return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit_fixed: null argument");
  const double k1 = Nm / (g->h[0] * g->h[1] * g->h[2]);
  if (k1 == 0.0) return LEPIC_OK;   // every contribution is zero
  const double unit = std::ldexp(std::fabs(k1), frac_bits);   // one full weight, in units
  if (frac_bits < -1000 || frac_bits > 1000 || !(unit >= 1.0 && unit <= 17592186044416.0))
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      "deposit_fixed: |Nm/(h1 h2 h3)| * 2^frac_bits must lie in [1, 2^44]");
  const double U = std::rint(unit);
  if (p->n == 0) return LEPIC_OK;
  const int64_t nodes = grid_nodes(g);
  double *rep = lepic_replicas(ctx, nodes);
  if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "deposit_fixed: replica allocation failed");
  if (nodes > ctx->rep_nodes)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit_fixed: replicas smaller than the grid");
  unsigned long long *irep = reinterpret_cast<unsigned long long *>(rep);   // zero bits either way
  hipLaunchKernelGGL(k_deposit_fixed, dim3(grid_blocks(p->n, CTILE)), dim3(BLOCK), 0,
                     ctx->stream, make_kgrid(g), kp(p), p->n, U, k1 < 0.0 ? 1 : 0, irep, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_deposit_fixed");
  hipLaunchKernelGGL(k_fixed_flush, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     np_local, irep, nodes, kq, std::ldexp(1.0, -frac_bits));
  LEPIC_LAUNCH_CHECK(ctx, "k_fixed_flush");
  return LEPIC_OK;
}

int lepic_push_bc_deposit_inplace(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p,
                                  const double *E4, const lepic_bfield *bf, double q, double m,
                                  double dt, const int32_t *bcnd, const lepic_bc_params *bp,
                                  const double *kq, double *np_local, double *p_mac,
                                  double *P_loss, uint64_t *lost_acc, int64_t *np_lost) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, p, "push_bc_deposit_inplace"))) return st;
  if (!aligned_runs(p))
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      "push_bc_deposit_inplace: particle arrays must be 16-byte aligned");
  if (!E4 || !bcnd || !bp || !p_mac || !P_loss || (np_local && !kq))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit_inplace: null argument");
  if ((st = check_die(ctx, bp, "push_bc_deposit_inplace"))) return st;
  if (bf && (bf->mode == 2 || bf->mode == 3) && !bf->B4)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit_inplace: null B map");
  if (bf && (bf->mode < 0 || bf->mode > 3))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit_inplace: bad B mode");
  if ((st = check_bgrid(ctx, bf, "push_bc_deposit_inplace"))) return st;
  const int bmode = bf ? bf->mode : 0;
  const int64_t nodes = grid_nodes(g);
  double *rep = nullptr;
  if (np_local) {
    rep = lepic_replicas(ctx, nodes);
    if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "push_bc_deposit_inplace: replicas");
  }
  // losses go to the context counter when the host asks for this call's
  // count, straight to the caller's accumulator otherwise (no host sync)
  unsigned long long *own = reinterpret_cast<unsigned long long *>(ctx->d_ctrl + CTRL_COUNT_OUT);
  unsigned long long *ctr = own;
  if (np_lost || !lost_acc) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream,
                       reinterpret_cast<int32_t *>(own), (int64_t)2, 0);
  } else {
    ctr = reinterpret_cast<unsigned long long *>(lost_acc);
  }
  if (p->n > 0) {
    const double k1 = dt * q / (2.0 * m);
    const KGrid kg = make_kgrid(g);
    const KBField kbf = make_kbfield(bf, k1);
    // per-node wall table, rebuilt from bcnd on every call (a few microseconds)
    int32_t *wt = reinterpret_cast<int32_t *>(lepic_scratch(ctx, sizeof(int32_t) * (size_t)nodes, 1));
    if (!wt) return lepic_fail(ctx, LEPIC_ERR_HIP, "push_bc_deposit_inplace: wall table");
    hipLaunchKernelGGL(k_wall_table, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream, kg,
                       bcnd, wt, nodes);
    bcnd = wt;
    if (np_local) {
      const double kdep = bp->Nm / (g->h[0] * g->h[1] * g->h[2]);
      switch (bmode) {
        case 0: hipLaunchKernelGGL((k_push<0>), dim3(grid_blocks(p->n, CTILE)), dim3(BLOCK), 0, ctx->stream, kg,
                         kp(p), p->n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                         kbc(bp), kdep, rep, nodes, 1, p_mac, ctx->wdiag, ctr); break;
        case 1: hipLaunchKernelGGL((k_push<1>), dim3(grid_blocks(p->n, CTILE)), dim3(BLOCK), 0, ctx->stream, kg,
                         kp(p), p->n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                         kbc(bp), kdep, rep, nodes, 1, p_mac, ctx->wdiag, ctr); break;
        default: hipLaunchKernelGGL((k_push<2>), dim3(grid_blocks(p->n, CTILE)), dim3(BLOCK), 0, ctx->stream, kg,
                         kp(p), p->n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                         kbc(bp), kdep, rep, nodes, 1, p_mac, ctx->wdiag, ctr); break;
      }
      LEPIC_LAUNCH_CHECK(ctx, "k_push");
      hipLaunchKernelGGL(k_wd_flush, dim3(1), dim3(64), 0, ctx->stream, ctx->wdiag, p_mac, P_loss,
                         std::min(bp->ngrid + 1, LEPIC_WD_LABELS));
      LEPIC_LAUNCH_CHECK(ctx, "k_wd_flush");
    } else {
      switch (bmode) {
        case 0: hipLaunchKernelGGL((k_push_nodep<0>), dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0,
                         ctx->stream, kg, kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                         kbf, k1, dt, bcnd, kbc(bp), p_mac, P_loss, ctr); break;
        case 1: hipLaunchKernelGGL((k_push_nodep<1>), dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0,
                         ctx->stream, kg, kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                         kbf, k1, dt, bcnd, kbc(bp), p_mac, P_loss, ctr); break;
        default: hipLaunchKernelGGL((k_push_nodep<2>), dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0,
                         ctx->stream, kg, kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                         kbf, k1, dt, bcnd, kbc(bp), p_mac, P_loss, ctr); break;
      }
      LEPIC_LAUNCH_CHECK(ctx, "k_push_nodep");
    }
  }
  if (np_local && (st = lepic_replica_flush(ctx, nodes, np_local, kq))) return st;
  if (np_lost) {
    if (lost_acc)
      hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(64), 0, ctx->stream,
                         reinterpret_cast<unsigned long long *>(lost_acc), own);
    LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, own, sizeof(int64_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
    LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *np_lost = ctx->h_pinned[0];
  }
  return LEPIC_OK;
}

int lepic_push_bc_deposit(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                          lepic_pset *dst, const double *E4, const lepic_bfield *bf, double q,
                          double m, double dt, const int32_t *bcnd, const lepic_bc_params *bp,
                          const double *kq, double *np_local, double *p_mac, double *P_loss,
                          int64_t *np_lost) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, src, "push_bc_deposit src")) ||
      (st = check_pset(ctx, dst, "push_bc_deposit dst")))
    return st;
  if (!E4 || !bcnd || !bp || !p_mac || !P_loss || (np_local && !kq))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit: null argument");
  if ((st = check_die(ctx, bp, "push_bc_deposit"))) return st;
  if (bf && (bf->mode == 2 || bf->mode == 3) && !bf->B4)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit: null B map");
  if (bf && (bf->mode < 0 || bf->mode > 3))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit: bad B mode");
  if ((st = check_bgrid(ctx, bf, "push_bc_deposit"))) return st;
  const int bmode = bf ? bf->mode : 0;
  if (dst->cap < src->n)
    return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "push_bc_deposit: dst too small");
  if (src == dst || src->x == dst->x)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit: src and dst must differ");
  int64_t n = src->n, nout = 0;
  KGrid kg = make_kgrid(g);
  KBC kb = kbc(bp);
  double k1 = dt * q / (2.0 * m);
  KBField kbf = make_kbfield(bf, k1);
  double kdep = np_local ? bp->Nm / (g->h[0] * g->h[1] * g->h[2]) : 0.0;
  const int64_t nodes = grid_nodes(g);
  double *rep = nullptr;
  if (np_local) {
    rep = lepic_replicas(ctx, nodes);
    if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "push_bc_deposit: replicas");
  }
  st = run_compaction(
      ctx, n,
      [&](dim3 grid, uint32_t epoch, int *counter, int64_t *spin, int64_t *count_out,
          int64_t ntiles) {
        switch (bmode) {
          case 0: hipLaunchKernelGGL((k_push_bc_deposit<0>), grid, dim3(BLOCK), 0, ctx->stream, kg, kp(src),
                           kp(dst), n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                           kb, kdep, rep, nodes, p_mac, P_loss, ctx->d_status, epoch,
                           counter, spin, count_out, ntiles); break;
          case 1: hipLaunchKernelGGL((k_push_bc_deposit<1>), grid, dim3(BLOCK), 0, ctx->stream, kg, kp(src),
                           kp(dst), n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                           kb, kdep, rep, nodes, p_mac, P_loss, ctx->d_status, epoch,
                           counter, spin, count_out, ntiles); break;
          default: hipLaunchKernelGGL((k_push_bc_deposit<2>), grid, dim3(BLOCK), 0, ctx->stream, kg, kp(src),
                           kp(dst), n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                           kb, kdep, rep, nodes, p_mac, P_loss, ctx->d_status, epoch,
                           counter, spin, count_out, ntiles); break;
        }
      },
      &nout);
  if (st) return st;
  if (np_local && (st = lepic_replica_flush(ctx, nodes, np_local, kq))) return st;
  dst->n = nout;
  if (np_lost) *np_lost = n - nout;
  return LEPIC_OK;
}

static int sort_impl(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src, lepic_pset *dst,
                     int32_t *cell_count, int32_t *cell_start, int drop_dead) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, src, "sort src")) ||
      (st = check_pset(ctx, dst, "sort dst")))
    return st;
  if (!src->cell && !dst->cell) return lepic_fail(ctx, LEPIC_ERR_ARG, "sort: null cell array");
  if (dst->cap < src->n) return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "sort: dst too small");
  if (src->x == dst->x) return lepic_fail(ctx, LEPIC_ERR_ARG, "sort: src and dst must differ");
  const int64_t n = src->n;
  const int64_t ncells = (int64_t)g->n[0] * g->n[1] * g->n[2];
  if (ncells >= ((int64_t)1 << 31) || n >= ((int64_t)1 << 32))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "sort: grid or particle count too large");
  KGrid kg = make_kgrid(g);
  int32_t *count = cell_count;
  if (cell_start && !count) {
    count = reinterpret_cast<int32_t *>(lepic_scratch(ctx, sizeof(int32_t) * (size_t)ncells, 1));
    if (!count) return lepic_fail(ctx, LEPIC_ERR_HIP, "sort: scratch allocation failed");
  }
  if (count)
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_blocks(ncells, 256)), dim3(256), 0, ctx->stream,
                       count, ncells, 0);
  dst->n = n;
  if (n > 0) {
    int nbits = 1;
    while (((int64_t)1 << nbits) <= ncells) ++nbits;   // keys 0..ncells (holes)
    size_t temp_bytes = 0;
    uint32_t *nul = nullptr;
    LEPIC_HIP(ctx, rocprim::radix_sort_pairs(nullptr, temp_bytes, nul, nul, nul, nul, (size_t)n,
                                             0, nbits, ctx->stream));
    size_t arr = sizeof(uint32_t) * (size_t)n;
    size_t off = (arr + 255) & ~(size_t)255;
    char *base = reinterpret_cast<char *>(lepic_scratch(ctx, 4 * off + temp_bytes, 0));
    if (!base) return lepic_fail(ctx, LEPIC_ERR_HIP, "sort: scratch allocation failed");
    uint32_t *keys = reinterpret_cast<uint32_t *>(base);
    uint32_t *vals = reinterpret_cast<uint32_t *>(base + off);
    uint32_t *keys_out = reinterpret_cast<uint32_t *>(base + 2 * off);
    uint32_t *vals_out = reinterpret_cast<uint32_t *>(base + 3 * off);
    void *temp = base + 4 * off;
    unsigned long long *ndead = reinterpret_cast<unsigned long long *>(ctx->d_ctrl + CTRL_COUNT_OUT);
    if (drop_dead)
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, ctx->stream,
                         reinterpret_cast<int32_t *>(ndead), (int64_t)2, 0);
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_blocks(n, BLOCK)), dim3(BLOCK), 0, ctx->stream, kg,
                       kp(src), n, keys, vals, count, drop_dead, ndead);
    LEPIC_LAUNCH_CHECK(ctx, "k_cell_keys");
    LEPIC_HIP(ctx, rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_out, vals, vals_out,
                                             (size_t)n, 0, nbits, ctx->stream));
    KPset d = kp(dst);
    if (!d.cell) return lepic_fail(ctx, LEPIC_ERR_ARG, "sort: dst has no cell array");
    int64_t nkeep = n;
    if (drop_dead) {
      LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, ndead, sizeof(int64_t), hipMemcpyDeviceToHost,
                                    ctx->stream));
      LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      nkeep = n - ctx->h_pinned[0];
      dst->n = nkeep;
    }
    if (nkeep > 0)
      hipLaunchKernelGGL(k_gather_sorted, dim3(grid_blocks(nkeep, BLOCK)), dim3(BLOCK), 0,
                         ctx->stream, kp(src), d, nkeep, keys_out, vals_out);
    LEPIC_LAUNCH_CHECK(ctx, "k_gather_sorted");
  }
  if (cell_start) {
    size_t temp_bytes = 0;
    LEPIC_HIP(ctx, rocprim::exclusive_scan(nullptr, temp_bytes, count, cell_start, 0,
                                           (size_t)ncells, rocprim::plus<int32_t>(),
                                           ctx->stream));
    // scratch 0 held the sort keys; the gather above is ordered before the scan
    void *temp = lepic_scratch(ctx, temp_bytes + 1024, 0);
    if (!temp) return lepic_fail(ctx, LEPIC_ERR_HIP, "sort: scan scratch allocation failed");
    LEPIC_HIP(ctx, rocprim::exclusive_scan(temp, temp_bytes, count, cell_start, 0, (size_t)ncells,
                                           rocprim::plus<int32_t>(), ctx->stream));
  }
  return LEPIC_OK;
}

int lepic_sort_particles_by_cell(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                                 lepic_pset *dst, int32_t *cell_count, int32_t *cell_start) {
  return sort_impl(ctx, g, src, dst, cell_count, cell_start, 0);
}

int lepic_sort_compact(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                       lepic_pset *dst) {
  return sort_impl(ctx, g, src, dst, nullptr, nullptr, 1);
}

int lepic_bin_compact(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                      lepic_pset *dst) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_pset(ctx, src, "bin_compact src")) ||
      (st = check_pset(ctx, dst, "bin_compact dst")))
    return st;
  if (!dst->cell) return lepic_fail(ctx, LEPIC_ERR_ARG, "bin_compact: dst has no cell array");
  if (dst->cap < src->n) return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "bin_compact: dst too small");
  if (src->x == dst->x) return lepic_fail(ctx, LEPIC_ERR_ARG, "bin_compact: src and dst must differ");
  const int64_t n = src->n;
  const int64_t ncells = (int64_t)g->n[0] * g->n[1] * g->n[2];
  if (ncells >= ((int64_t)1 << 31) - 1 || n >= ((int64_t)1 << 31))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "bin_compact: grid or particle count too large");
  if (n == 0) {
    dst->n = 0;
    return LEPIC_OK;
  }
  if (!aligned_runs(src) || !aligned_runs(dst))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "bin_compact: particle arrays must be 16-byte aligned");
  // scratch 0: count[ncells + 1] | start[ncells + 1] | scan temp
  size_t temp_bytes = 0;
  int32_t *nul = nullptr;
  LEPIC_HIP(ctx, rocprim::exclusive_scan(nullptr, temp_bytes, nul, nul, 0, (size_t)ncells + 1,
                                         rocprim::plus<int32_t>(), ctx->stream));
  const size_t cb = (sizeof(int32_t) * (size_t)(ncells + 1) + 255) & ~(size_t)255;
  char *base = reinterpret_cast<char *>(lepic_scratch(ctx, 2 * cb + temp_bytes + 256, 0));
  if (!base) return lepic_fail(ctx, LEPIC_ERR_HIP, "bin_compact: scratch allocation failed");
  int32_t *count = reinterpret_cast<int32_t *>(base);
  int32_t *start = reinterpret_cast<int32_t *>(base + cb);
  void *temp = base + 2 * cb;
  const KGrid kg = make_kgrid(g);
  const unsigned tiles = (unsigned)((n + CTILE - 1) / CTILE);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_blocks(ncells + 1, 256)), dim3(256), 0, ctx->stream,
                     count, ncells + 1, 0);
  hipLaunchKernelGGL(k_bin_count, dim3(tiles), dim3(BLOCK), 0, ctx->stream, kg, kp(src), n, count);
  LEPIC_LAUNCH_CHECK(ctx, "k_bin_count");
  // start[c] = particles in cells before c; start[ncells] = live total
  LEPIC_HIP(ctx, rocprim::exclusive_scan(temp, temp_bytes, count, start, 0, (size_t)ncells + 1,
                                         rocprim::plus<int32_t>(), ctx->stream));
  LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, start + ncells, sizeof(int32_t),
                                hipMemcpyDeviceToHost, ctx->stream));
  hipLaunchKernelGGL(k_bin_scatter, dim3(tiles), dim3(BLOCK), 0, ctx->stream, kg, kp(src),
                     kp(dst), n, start);
  LEPIC_LAUNCH_CHECK(ctx, "k_bin_scatter");
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  dst->n = (int64_t)reinterpret_cast<int32_t *>(ctx->h_pinned)[0];
  return LEPIC_OK;
}

int lepic_load_uniform_device(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                              int32_t np_cell, int32_t ntype, const double *vt0, uint64_t key,
                              lepic_pset *dp) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g))) return st;
  if (!bcnd || !vt0 || !dp || np_cell < 1 || ntype < 1 || ntype > LEPIC_MAX_SPECIES)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "load_uniform_device: bad argument");
  // loadable cells: any of the 8 corner nodes is plasma (mod_particle_loader.f90:75-91)
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const int64_t nodes = grid_nodes(g);
  std::vector<int32_t> hb((size_t)nodes);
  LEPIC_HIP(ctx, hipMemcpyAsync(hb.data(), bcnd, sizeof(int32_t) * (size_t)nodes,
                                hipMemcpyDeviceToHost, ctx->stream));
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t sy = n1 + 3, sz = (int64_t)(n1 + 3) * (n2 + 3);
  std::vector<int32_t> cells;
  cells.reserve((size_t)n1 * n2 * n3);
  for (int iz = 1; iz <= n3; ++iz)
    for (int iy = 1; iy <= n2; ++iy)
      for (int ix = 1; ix <= n1; ++ix) {
        int64_t b = ix + sy * iy + sz * iz;
        bool any = hb[b] == -1 || hb[b + 1] == -1 || hb[b + sy + 1] == -1 || hb[b + sy] == -1 ||
                   hb[b + sz] == -1 || hb[b + sz + 1] == -1 || hb[b + sz + sy + 1] == -1 ||
                   hb[b + sz + sy] == -1;
        if (any) cells.push_back((ix - 1) + n1 * ((iy - 1) + n2 * (iz - 1)));
      }
  const int64_t n = (int64_t)cells.size() * np_cell;
  for (int t = 0; t < ntype; ++t) {
    if ((st = check_pset(ctx, &dp[t], "load_uniform_device"))) return st;
    if (dp[t].cap < n) return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "load_uniform_device: capacity");
    if (!dp[t].x || !dp[t].cell) return lepic_fail(ctx, LEPIC_ERR_ARG, "load_uniform_device: null array");
  }
  int32_t *dcells = reinterpret_cast<int32_t *>(lepic_scratch(ctx, sizeof(int32_t) * cells.size() + 4, 1));
  if (!dcells) return lepic_fail(ctx, LEPIC_ERR_HIP, "load_uniform_device: scratch");
  LEPIC_HIP(ctx, hipMemcpyAsync(dcells, cells.data(), sizeof(int32_t) * cells.size(),
                                hipMemcpyHostToDevice, ctx->stream));
  for (int t = 0; t < ntype; ++t) {
    if (n > 0)
      hipLaunchKernelGGL(k_load_uniform, dim3(grid_blocks(n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                         make_kgrid(g), kp(&dp[t]), n, dcells, np_cell, vt0[t], key, (uint32_t)t);
    LEPIC_LAUNCH_CHECK(ctx, "k_load_uniform");
    dp[t].n = n;
  }
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LEPIC_OK;
}

}  // extern "C"
