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
#include "lepic_philox.h"

namespace {

constexpr int BLOCK = 256;
constexpr int NWAVES = BLOCK / LEPIC_WAVE;
constexpr int LDS_BOX = 4096;            // doubles of the per-block deposit box (32 KB)
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

struct KBC {
  double xmax, ymax, zmax;
  int flag_pbc, flag_nmn, ptype, tag_neg;
  double qmacro, mass, Nm;
  int ngrid;
};

inline KBC kbc(const lepic_bc_params *b) {
  KBC k;
  k.xmax = b->xmax; k.ymax = b->ymax; k.zmax = b->zmax;
  k.flag_pbc = b->flag_pbc; k.flag_nmn = b->flag_nmn;
  k.ptype = b->ptype; k.tag_neg = b->tag_neg;
  k.qmacro = b->qmacro; k.mass = b->mass; k.Nm = b->Nm; k.ngrid = b->ngrid;
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
  velocity_update(g, bf, s, k1, Ex, Ey, Ez, vx, vy, vz);
  p.vx[i] = vx; p.vy[i] = vy; p.vz[i] = vz;
  p.x[i] = xp + dt * vx; p.y[i] = yp + dt * vy; p.z[i] = zp + dt * vz;
}

// --------------------------------------------------------------------------
// Wall test of apply_particle_bc_legacy (mod_particleBC.f90:96-116).
// Returns 0 survive, 1 lost on walls, 2 lost negative ion at a Neumann wall.
// --------------------------------------------------------------------------
__device__ inline int wall_test(const KGrid &g, const int *__restrict__ bcnd, const KBC &bp,
                                double xp, double yp, double zp, int &label) {
  int ix = (int)floor(xp / g.h1) + 1;
  int iy = (int)floor(yp / g.h2) + 1;
  int iz = (int)floor(zp / g.h3) + 1;
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

__device__ inline void wall_diagnostics(const KBC &bp, int label, double vx, double vy,
                                        double vz, double *p_mac, double *P_loss) {
  double v2 = vx * vx + vy * vy + vz * vz;
  double Ek_eV = 0.5 * bp.mass * v2 / 1.60217646e-19;
  if (label <= bp.ngrid) {
    atomicAdd(&p_mac[2 * label + 0], bp.qmacro);
    atomicAdd(&p_mac[2 * label + 1], fabs(bp.qmacro) * Ek_eV);
  }
  double Ek_J = 0.5 * bp.mass * v2;
  atomicAdd(P_loss, bp.Nm * Ek_J);
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
  bool keep = false;
  double xp = 0, yp = 0, zp = 0, vx = 0, vy = 0, vz = 0;
  if (i < n) {
    if (src.dead[i] != 1) {
      xp = src.x[i]; yp = src.y[i]; zp = src.z[i];
      vx = src.vx[i]; vy = src.vy[i]; vz = src.vz[i];
      int label;
      int lost = wall_test(g, bcnd, bp, xp, yp, zp, label);
      if (lost) {
        wall_diagnostics(bp, label, vx, vy, vz, p_mac, P_loss);
      } else {
        survivor_wrap(bp, xp, yp, zp, vx);
        keep = true;
      }
    }
  }
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
  ix = (int)(x / g.h1) + 1;
  iy = (int)(y / g.h2) + 1;
  iz = (int)(z / g.h3) + 1;
  double px = ((double)ix * g.h1 - x) / g.h1;
  double py = ((double)iy * g.h2 - y) / g.h2;
  double pz = ((double)iz * g.h3 - z) / g.h3;
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

struct DepShared {
  int mn[3], mx[3];
  double box[LDS_BOX];
};

// Block deposit: accumulate the block's weights in an LDS box spanning its
// node range and flush each touched node with one global atomic, or fall back
// to per-particle global atomics when the box does not fit.  kq is 1 or 2, so
// kq*sum(w) equals sum(kq*w) term for term.
__device__ inline void block_deposit(const KGrid &g, DepShared &ds, bool has, int ix, int iy,
                                     int iz, const double w[8], const double *__restrict__ kq,
                                     double *np) {
  if (threadIdx.x == 0) {
    ds.mn[0] = ds.mn[1] = ds.mn[2] = 1 << 30;
    ds.mx[0] = ds.mx[1] = ds.mx[2] = -(1 << 30);
  }
  __syncthreads();
  if (has) {
    atomicMin(&ds.mn[0], ix); atomicMin(&ds.mn[1], iy); atomicMin(&ds.mn[2], iz);
    atomicMax(&ds.mx[0], ix); atomicMax(&ds.mx[1], iy); atomicMax(&ds.mx[2], iz);
  }
  __syncthreads();
  const int x0 = ds.mn[0], y0 = ds.mn[1], z0 = ds.mn[2];
  if (ds.mx[0] < x0) return;  // nothing to deposit (uniform branch)
  const int bx = ds.mx[0] - x0 + 2, by = ds.mx[1] - y0 + 2, bz = ds.mx[2] - z0 + 2;
  const int vol = bx * by * bz;
  if (vol > LDS_BOX) {
    if (has) {
      int64_t b = gidx(g, ix, iy, iz);
      const int64_t nd[8] = {b, b + 1, b + g.sy + 1, b + g.sy, b + g.sz, b + g.sz + 1,
                             b + g.sz + g.sy + 1, b + g.sz + g.sy};
#pragma unroll
      for (int q = 0; q < 8; ++q) atomicAdd(&np[nd[q]], kq[nd[q]] * w[q]);
    }
    return;
  }
  for (int t = threadIdx.x; t < vol; t += BLOCK) ds.box[t] = 0.0;
  __syncthreads();
  if (has) {
    int l = (ix - x0) + bx * ((iy - y0) + by * (iz - z0));
    const int ld[8] = {l, l + 1, l + bx + 1, l + bx, l + bx * by, l + bx * by + 1,
                       l + bx * by + bx + 1, l + bx * by + bx};
#pragma unroll
    for (int q = 0; q < 8; ++q) atomicAdd(&ds.box[ld[q]], w[q]);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < vol; t += BLOCK) {
    double v = ds.box[t];
    if (v != 0.0) {
      int lx = t % bx, r = t / bx, ly = r % by, lz = r / by;
      int64_t node = gidx(g, x0 + lx, y0 + ly, z0 + lz);
      atomicAdd(&np[node], kq[node] * v);
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_deposit(KGrid g, KPset p, int64_t n,
                                                   const double *__restrict__ kq, double k1,
                                                   double *np) {
  __shared__ DepShared ds;
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  bool has = false;
  int ix = 0, iy = 0, iz = 0;
  double w[8];
  if (i < n && p.dead[i] == 0) has = deposit_weights(g, k1, p.x[i], p.y[i], p.z[i], ix, iy, iz, w);
  block_deposit(g, ds, has, ix, iy, iz, w, kq, np);
}

// --------------------------------------------------------------------------
// Fused particle phase: gather + push + wall BC + stable compaction + deposit.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_push_bc_deposit(
    KGrid g, KPset src, KPset dst, int64_t n, const double4 *__restrict__ E4, KBField bf,
    double k1, double dt, const int *__restrict__ bcnd, KBC bp, const double *__restrict__ kq,
    double kdep, double *np, double *p_mac, double *P_loss, uint64_t *status, uint32_t epoch,
    int *counter, int64_t *spin_fail, int64_t *count_out, int64_t ntiles) {
  __shared__ LBShared sh;
  __shared__ DepShared ds;
  int64_t tile = take_tile(counter, ntiles, sh);
  int64_t i = tile * BLOCK + threadIdx.x;
  bool keep = false;
  double xp = 0, yp = 0, zp = 0, vx = 0, vy = 0, vz = 0;
  if (i < n && src.dead[i] != 1) {
    xp = src.x[i]; yp = src.y[i]; zp = src.z[i];
    vx = src.vx[i]; vy = src.vy[i]; vz = src.vz[i];
    Stencil s;
    make_stencil(g, xp, yp, zp, s);
    double Ex, Ey, Ez;
    gather3(E4, g, s, Ex, Ey, Ez);
    velocity_update(g, bf, s, k1, Ex, Ey, Ez, vx, vy, vz);
    xp = xp + dt * vx; yp = yp + dt * vy; zp = zp + dt * vz;
    int label;
    int lost = wall_test(g, bcnd, bp, xp, yp, zp, label);
    if (lost) {
      wall_diagnostics(bp, label, vx, vy, vz, p_mac, P_loss);
    } else {
      survivor_wrap(bp, xp, yp, zp, vx);
      keep = true;
    }
  }
  int64_t o = compact_slot(keep, sh, status, tile, epoch, spin_fail, count_out, ntiles);
  if (keep) {
    dst.x[o] = xp; dst.y[o] = yp; dst.z[o] = zp;
    dst.vx[o] = vx; dst.vy[o] = vy; dst.vz[o] = vz;
    dst.dead[o] = 0;
  }
  if (np) {
    bool has = false;
    int ix = 0, iy = 0, iz = 0;
    double w[8];
    if (keep) has = deposit_weights(g, kdep, xp, yp, zp, ix, iy, iz, w);
    block_deposit(g, ds, has, ix, iy, iz, w, kq, np);
  }
}

// --------------------------------------------------------------------------
// a2: cell keys (cell_index_from_position, mod_particle_sorting.f90:16-48)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_cell_keys(KGrid g, KPset p, int64_t n,
                                                     uint32_t *keys, uint32_t *vals,
                                                     int32_t *count) {
  int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  int ix = (int)(p.x[i] / g.h1) + 1;
  int iy = (int)(p.y[i] / g.h2) + 1;
  int iz = (int)(p.z[i] / g.h3) + 1;
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

int check_grid(lepic_ctx *ctx, const lepic_grid *g) {
  if (!g || g->n[0] < 1 || g->n[1] < 1 || g->n[2] < 1 || !(g->h[0] > 0) || !(g->h[1] > 0) ||
      !(g->h[2] > 0))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "bad grid");
  return LEPIC_OK;
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
  if (bf && bf->mode == 2 && !bf->B4) return lepic_fail(ctx, LEPIC_ERR_ARG, "move_boris: null B map");
  if (bf && (bf->mode < 0 || bf->mode > 2)) return lepic_fail(ctx, LEPIC_ERR_ARG, "move_boris: bad mode");
  if (p->n == 0) return LEPIC_OK;
  double k1 = dt * q / (2.0 * m);
  hipLaunchKernelGGL(k_move_boris, dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, reinterpret_cast<const double4 *>(E4),
                     make_kbfield(bf), k1, dt);
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
  if (!kq || !np_local) return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit: null argument");
  if (p->n == 0) return LEPIC_OK;
  double k1 = Nm / (g->h[0] * g->h[1] * g->h[2]);
  hipLaunchKernelGGL(k_deposit, dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), kp(p), p->n, kq, k1, np_local);
  LEPIC_LAUNCH_CHECK(ctx, "k_deposit");
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
  if (bf && bf->mode == 2 && !bf->B4)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit: null B map");
  if (dst->cap < src->n)
    return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "push_bc_deposit: dst too small");
  if (src == dst || src->x == dst->x)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push_bc_deposit: src and dst must differ");
  int64_t n = src->n, nout = 0;
  KGrid kg = make_kgrid(g);
  KBC kb = kbc(bp);
  KBField kbf = make_kbfield(bf);
  double k1 = dt * q / (2.0 * m);
  double kdep = np_local ? bp->Nm / (g->h[0] * g->h[1] * g->h[2]) : 0.0;
  st = run_compaction(
      ctx, n,
      [&](dim3 grid, uint32_t epoch, int *counter, int64_t *spin, int64_t *count_out,
          int64_t ntiles) {
        hipLaunchKernelGGL(k_push_bc_deposit, grid, dim3(BLOCK), 0, ctx->stream, kg, kp(src),
                           kp(dst), n, reinterpret_cast<const double4 *>(E4), kbf, k1, dt, bcnd,
                           kb, kq, kdep, np_local, p_mac, P_loss, ctx->d_status, epoch, counter,
                           spin, count_out, ntiles);
      },
      &nout);
  if (st) return st;
  dst->n = nout;
  if (np_lost) *np_lost = n - nout;
  return LEPIC_OK;
}

int lepic_sort_particles_by_cell(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                                 lepic_pset *dst, int32_t *cell_count, int32_t *cell_start) {
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
    while (((int64_t)1 << nbits) < ncells) ++nbits;
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
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_blocks(n, BLOCK)), dim3(BLOCK), 0, ctx->stream, kg,
                       kp(src), n, keys, vals, count);
    LEPIC_LAUNCH_CHECK(ctx, "k_cell_keys");
    LEPIC_HIP(ctx, rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_out, vals, vals_out,
                                             (size_t)n, 0, nbits, ctx->stream));
    KPset d = kp(dst);
    if (!d.cell) return lepic_fail(ctx, LEPIC_ERR_ARG, "sort: dst has no cell array");
    hipLaunchKernelGGL(k_gather_sorted, dim3(grid_blocks(n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       kp(src), d, n, keys_out, vals_out);
    LEPIC_LAUNCH_CHECK(ctx, "k_gather_sorted");
  }
  if (cell_start) {
    size_t temp_bytes = 0;
    LEPIC_HIP(ctx, rocprim::exclusive_scan(nullptr, temp_bytes, count, cell_start, 0,
                                           (size_t)ncells, rocprim::plus<int32_t>(),
                                           ctx->stream));
    void *temp = lepic_scratch(ctx, temp_bytes + 1024, 0);
    if (n > 0) {
      // scratch 0 holds the sort keys, which are no longer needed
    }
    if (!temp) return lepic_fail(ctx, LEPIC_ERR_HIP, "sort: scan scratch allocation failed");
    LEPIC_HIP(ctx, rocprim::exclusive_scan(temp, temp_bytes, count, cell_start, 0, (size_t)ncells,
                                           rocprim::plus<int32_t>(), ctx->stream));
  }
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
