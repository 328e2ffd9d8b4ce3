// lepic_particles32.hip -- the FP32 particle phase (the precision form of the
// step): particles held as an int32 cell id plus FP32 in-cell offsets and FP32
// velocities (28 B per particle), the FP32 gather + Boris/leapfrog push + wall
// BC in place, a fixed-point int64 deposit, the stable sort by cell, and the
// conversions to and from the FP64 layout.
//
// Position: x = ((ix - 1) + fx) h1 with ix the cell's 1-based x index and fx in
// [0, 1), so the stencil's lower node is the cell's own lower corner and its
// trilinear weights are products of fx and 1 - fx (px = ix - x/h1 = 1 - fx in
// the notation of mod_particleMover.f90:104-127).  Offsets keep 24 bits of a
// cell wherever the particle is, where an absolute FP32 coordinate would lose
// precision with the distance from the origin.
//
// Every FP32 operation is one IEEE single operation in a fixed order
// (-ffp-contract=off; sqrt and division correctly rounded), so the oracle's C
// restatement reproduces cells, offsets and velocities bit for bit.
#include <cmath>

#include "lepic_internal.h"
#include "lepic_partutil.h"

namespace {

constexpr int BLOCK = 256;
constexpr int CPT = 4;              // particles per lane, moved with 16-byte accesses
constexpr int CTILE = BLOCK * CPT;
constexpr int BOX32 = 8000;   // uint32 nodes of the per-block deposit box (with its header and the wall sums, 32 KB of LDS: 5 blocks per CU)

struct KP32 {
  int32_t *cell;
  float *fx, *fy, *fz, *vx, *vy, *vz;
};

inline KP32 kp32(const lepic_pset32 *p) {
  KP32 k{p->cell, p->fx, p->fy, p->fz, p->vx, p->vy, p->vz};
  return k;
}

// Exact x / d for x, d < 2^32: m = ceil(2^64 / d) (2^64 / d for powers of
// two), q = high 64 bits of x * m.  x m / 2^64 exceeds x / d by less than
// x / 2^64 < 2^-32 <= 1/d, so the floor is exact.
struct UDiv {
  uint64_t m;
  uint32_t d;
};

inline UDiv make_udiv(uint32_t d) {
  UDiv u;
  u.d = d;
  u.m = d <= 1 ? 0 : (~0ull) / d + 1;
  return u;
}

__device__ inline uint32_t udiv(uint32_t x, const UDiv &u) {
  return u.d <= 1 ? x : (uint32_t)__umul64hi((uint64_t)x, u.m);
}

struct KG32 {
  int n1, n2, n3;
  int ncells;            // cell ids outside [0, ncells) are treated as holes
  int64_t sy, sz;        // node strides (the FP64 grid's)
  UDiv d1, d2;           // cell id decode
  float dtx, dty, dtz;   // dt / h per axis, in cell widths per (m/s)
};

inline KG32 make_kg32(const lepic_grid *g, double dt) {
  KG32 k;
  k.n1 = g->n[0]; k.n2 = g->n[1]; k.n3 = g->n[2];
  k.ncells = g->n[0] * g->n[1] * g->n[2];
  k.sy = (int64_t)g->n[0] + 3;
  k.sz = ((int64_t)g->n[0] + 3) * ((int64_t)g->n[1] + 3);
  k.d1 = make_udiv((uint32_t)g->n[0]);
  k.d2 = make_udiv((uint32_t)g->n[1]);
  k.dtx = (float)(dt / g->h[0]);
  k.dty = (float)(dt / g->h[1]);
  k.dtz = (float)(dt / g->h[2]);
  return k;
}

// 0-based cell coordinates of cell id c >= 0.
__device__ inline void decode(const KG32 &g, int c, int &cx, int &cy, int &cz) {
  const uint32_t r = udiv((uint32_t)c, g.d1);
  cx = c - (int)r * g.n1;
  const uint32_t s = udiv(r, g.d2);
  cy = (int)r - (int)s * g.n2;
  cz = (int)s;
}

__device__ inline int64_t node_of(const KG32 &g, int ix, int iy, int iz) {
  return (int64_t)ix + g.sy * iy + g.sz * iz;
}

struct KBF32 {
  int mode;        // 0 none, 1 uniform, 2 map on the E grid (|B| gate at the lower node)
  int magnetized;
  float bx, by, bz;
  float k2u;       // uniform B: k1*2/(1+(k1*|B|)^2) formed in FP64 and rounded once
  const float4 *B4;
};

inline KBF32 make_kbf32(const lepic_bfield32 *bf, double k1d) {
  KBF32 k;
  k.mode = bf ? bf->mode : 0;
  k.magnetized = bf ? bf->magnetized : 1;
  const double bx = bf ? bf->B[0] : 0.0, by = bf ? bf->B[1] : 0.0, bz = bf ? bf->B[2] : 0.0;
  k.bx = (float)bx; k.by = (float)by; k.bz = (float)bz;
  const double Bt = std::sqrt(bx * bx + by * by + bz * bz);
  k.k2u = (float)(k1d * (2.0 / (1.0 + (k1d * Bt) * (k1d * Bt))));
  k.B4 = bf ? reinterpret_cast<const float4 *>(bf->B4) : nullptr;
  return k;
}

// Trilinear weights w1..w8 (mod_particleMover.f90:120-142 order) from the
// in-cell offsets: the lower node's share in x is 1 - fx, the upper's fx.
__device__ inline void weights32(float fx, float fy, float fz, float w[8]) {
  const float ax = 1.0f - fx, ay = 1.0f - fy, az = 1.0f - fz;
  w[0] = ax * ay * az;
  w[1] = fx * ay * az;
  w[2] = fx * fy * az;
  w[3] = ax * fy * az;
  w[4] = ax * ay * fz;
  w[5] = fx * ay * fz;
  w[6] = fx * fy * fz;
  w[7] = ax * fy * fz;
}

__device__ inline void sum8(const float w[8], const float4 &f1, const float4 &f2, const float4 &f3,
                            const float4 &f4, const float4 &f5, const float4 &f6, const float4 &f7,
                            const float4 &f8, float &ox, float &oy, float &oz) {
  ox = w[0] * f1.x + w[1] * f2.x + w[2] * f3.x + w[3] * f4.x + w[4] * f5.x + w[5] * f6.x +
       w[6] * f7.x + w[7] * f8.x;
  oy = w[0] * f1.y + w[1] * f2.y + w[2] * f3.y + w[3] * f4.y + w[4] * f5.y + w[5] * f6.y +
       w[6] * f7.y + w[7] * f8.y;
  oz = w[0] * f1.z + w[1] * f2.z + w[2] * f3.z + w[3] * f4.z + w[4] * f5.z + w[5] * f6.z +
       w[6] * f7.z + w[7] * f8.z;
}

// The velocity update of part_mover (Src/part_expmover.f90:139-240) in FP32.
template <int BM>
__device__ inline void velocity_update32(const KG32 &g, const KBF32 &bf, const float w[8],
                                         int64_t base, float k1, float Ex, float Ey, float Ez,
                                         float &vx, float &vy, float &vz) {
  const int mode = BM >= 0 ? BM : bf.mode;
  int use_b = 0;
  float Bx = 0.0f, By = 0.0f, Bz = 0.0f;
  if (mode == 1) {
    Bx = bf.bx; By = bf.by; Bz = bf.bz;
    use_b = 1;
  } else if (mode == 2) {
    if (bf.B4[base].w > 1.0e-5f) {
      const int64_t b1 = base + g.sy, b2 = base + g.sz, b3 = base + g.sz + g.sy;
      sum8(w, bf.B4[base], bf.B4[base + 1], bf.B4[b1 + 1], bf.B4[b1], bf.B4[b2], bf.B4[b2 + 1],
           bf.B4[b3 + 1], bf.B4[b3], Bx, By, Bz);
      use_b = 1;
    }
  }
  if (!bf.magnetized) use_b = 0;
  if (use_b) {
    float k2;
    if (mode == 1) {
      k2 = bf.k2u;
    } else {
      const float Bt = sqrtf(Bx * Bx + By * By + Bz * Bz);
      k2 = k1 * (2.0f / (1.0f + (k1 * Bt) * (k1 * Bt)));
    }
    const float v1x = vx + k1 * Ex, v1y = vy + k1 * Ey, v1z = vz + k1 * Ez;
    const float v3x = v1x + k1 * (v1y * Bz - v1z * By);
    const float v3y = v1y + k1 * (v1z * Bx - v1x * Bz);
    const float v3z = v1z + k1 * (v1x * By - v1y * Bx);
    const float v2x = v1x + k2 * (v3y * Bz - v3z * By);
    const float v2y = v1y + k2 * (v3z * Bx - v3x * Bz);
    const float v2z = v1z + k2 * (v3x * By - v3y * Bx);
    vx = v2x + k1 * Ex; vy = v2y + k1 * Ey; vz = v2z + k1 * Ez;
  } else {
    vx = vx + 2.0f * k1 * Ex; vy = vy + 2.0f * k1 * Ey; vz = vz + 2.0f * k1 * Ez;
  }
}

// Moves an offset by d cell widths: the new offset in [0, 1) and the whole
// cells crossed.  f + d - floor(f + d) rounds to 1 only for tiny negative
// sums; that position is the next cell's lower face.
__device__ inline int advance(float &f, float d) {
  const float s = f + d;
  const float fl = floorf(s);
  float r = s - fl;
  int k = (int)fl;
  if (r >= 1.0f) {
    r = 0.0f;
    k += 1;
  }
  f = r;
  return k;
}

// Reflection x -> -x of a position in cell ix (1-based) <= 0 with offset f:
// -x/h = -ix + (1 - f) cells, i.e. cell 1 - ix with offset 1 - f (the next
// cell with offset 0 when 1 - f rounds to 1).
__device__ inline void reflect(int &ix, float &f) {
  const float r = 1.0f - f;
  ix = 1 - ix;
  if (r >= 1.0f) {
    f = 0.0f;
    ix += 1;
  } else {
    f = r;
  }
}

// The wall BC of apply_particle_bc_legacy (mod_particleBC.f90:96-227) on the
// moved particle: lower node (ix, iy, iz) (1-based, unclamped) and offsets.
// Returns 1 lost on a wall (label set), 2 a negative ion lost at the Neumann
// face, 3 a survivor outside the cells (x, y, z off [0, L) after the wrap),
// 0 otherwise; survivors come back wrapped (Neumann reflection at x <= 0,
// periodic y and z).  A survivor exactly on y = 0 or z = 0 stays there (the
// reference moves it to the far face y = ymax, which is no cell of the
// offset layout).
__device__ inline int wall_entry32(const KG32 &g, const int *__restrict__ wt, int ix, int iy,
                                   int iz) {
  const int cx = imax(0, imin(g.n1 + 1, ix));
  const int cy = imax(0, imin(g.n2 + 1, iy));
  const int cz = imax(0, imin(g.n3 + 1, iz));
  return wt[node_of(g, cx, cy, cz)];
}

// c: the wall table's entry at the clamped lower node (wall_entry32).
__device__ inline int wall32(const KG32 &g, int c, const KBC &bp, int &ix, int &iy, int &iz,
                             float &fx, float &fy, float &fz, float &vx, int &label) {
  int lost = c >= 0 ? 1 : 0;
  label = c >= 0 ? c : 0;
  if (bp.ptype == bp.tag_neg && ix <= 0 && bp.flag_nmn == 1) {   // x < 0
    lost = 2;
    label = 0;
  }
  if (lost) return lost;
  if (bp.flag_nmn == 1) {
    if (ix <= 0) {                         // x < 0: x -> -x, vx -> -vx
      reflect(ix, fx);
      vx = -vx;
    } else if (ix == 1 && fx == 0.0f) {    // x = 0
      vx = -vx;
    }
  }
  if (bp.flag_pbc == 1) {
    if (iy >= g.n2 + 1) iy -= g.n2;        // y >= ymax
    if (iy <= 0) iy += g.n2;               // y < 0
    if (iz >= g.n3 + 1) iz -= g.n3;
    if (iz <= 0) iz += g.n3;
  }
  if (ix < 1 || ix > g.n1 || iy < 1 || iy > g.n2 || iz < 1 || iz > g.n3) return 3;
  return 0;
}

// ---------------------------------------------------------------------------
// Fixed-point deposit through a per-block int64 LDS box (the block's cells,
// widened by one node each way for motion within the step) and per-XCD int64
// replicas for stencils outside it.
// ---------------------------------------------------------------------------
// The box holds magnitudes: a species' parts all carry the sign of its
// charge (neg), applied when the box goes out to the replicas.  A node of the
// box sums at most one full weight per particle of the block, 1024 x 2^21 =
// 2^31 units, so uint32 holds it, and LDS atomics are 32-bit.
struct Box32 {
  int mn[3], mx[3], sum[3], cnt;
  int x0, y0, z0, bx, by, bz, vol;   // vol <= 0: no box (everything to the replicas)
  int neg;
  uint64_t mbx, mby;                 // small_div multipliers of bx and by
  uint32_t box[BOX32];
};

// Box extents (nodes) used when the block's full range does not fit BOX32:
// a cell-sorted tile spans a row of cells along x and one cell in y and z,
// so the box keeps 36 x 14 x 14 nodes around the tile's mean lower node and
// the few particles that drifted beyond go to the replicas.
constexpr int BOXLIM_X = 36, BOXLIM_Y = 14, BOXLIM_Z = 14;
static_assert(BOXLIM_X * BOXLIM_Y * BOXLIM_Z <= BOX32, "box limits exceed the box");

// lo / hi: the lane's lower nodes (1-based) before the move; sm / cnt: the
// sum of those nodes over the lane's particles and their number.  The box
// spans the block's lower nodes widened by one node below and two above
// (motion within the step); when that exceeds BOX32 nodes, it is cut to
// BOXLIM around the mean lower node, and particles outside it deposit into
// the replicas.
__device__ inline void box32_begin(const KG32 &g, Box32 &ds, const int lo[3], const int hi[3],
                                   const int sm[3], int cnt, int neg) {
  if (threadIdx.x == 0) {
    ds.neg = neg;
    ds.mn[0] = ds.mn[1] = ds.mn[2] = 1 << 30;
    ds.mx[0] = ds.mx[1] = ds.mx[2] = -(1 << 30);
    ds.sum[0] = ds.sum[1] = ds.sum[2] = 0;
    ds.cnt = 0;
  }
  __syncthreads();
  const int a0 = wave_min(lo[0]), a1 = wave_min(lo[1]), a2 = wave_min(lo[2]);
  const int b0 = wave_max(hi[0]), b1 = wave_max(hi[1]), b2 = wave_max(hi[2]);
  int s0 = sm[0], s1 = sm[1], s2 = sm[2], c = cnt;
  for (int d = 32; d > 0; d >>= 1) {
    s0 += __shfl_xor(s0, d); s1 += __shfl_xor(s1, d); s2 += __shfl_xor(s2, d);
    c += __shfl_xor(c, d);
  }
  if (__lane_id() == 0) {
    atomicMin(&ds.mn[0], a0); atomicMin(&ds.mn[1], a1); atomicMin(&ds.mn[2], a2);
    atomicMax(&ds.mx[0], b0); atomicMax(&ds.mx[1], b1); atomicMax(&ds.mx[2], b2);
    atomicAdd(&ds.sum[0], s0); atomicAdd(&ds.sum[1], s1); atomicAdd(&ds.sum[2], s2);
    atomicAdd(&ds.cnt, c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int vol = 0;
    if (ds.mx[0] >= ds.mn[0]) {
      const int top[3] = {g.n1 + 2, g.n2 + 2, g.n3 + 2};
      const int lim[3] = {BOXLIM_X, BOXLIM_Y, BOXLIM_Z};
      int a[3], e[3];
      for (int d = 0; d < 3; ++d) {
        a[d] = max(ds.mn[d] - 1, 0);
        e[d] = min(ds.mx[d] + 2, top[d]) - a[d] + 1;
      }
      if ((int64_t)e[0] * e[1] * e[2] > BOX32) {
        for (int d = 0; d < 3; ++d) {
          if (e[d] <= lim[d]) continue;
          const int mean = ds.sum[d] / max(ds.cnt, 1);
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
  for (int t = threadIdx.x; t < ds.vol; t += BLOCK) ds.box[t] = 0u;
  __syncthreads();
}

// Run key of a stencil with lower node (ix, iy, iz): the box-local node
// (>= 0) inside the box, -(global node) - 2 outside it.
__device__ inline int key32(const KG32 &g, const Box32 &ds, int ix, int iy, int iz) {
  if (ds.vol > 0) {
    const int lx = ix - ds.x0, ly = iy - ds.y0, lz = iz - ds.z0;
    if (lx >= 0 && ly >= 0 && lz >= 0 && lx + 1 < ds.bx && ly + 1 < ds.by && lz + 1 < ds.bz)
      return lx + ds.bx * (ly + ds.by * lz);
  }
  return -(int)node_of(g, ix, iy, iz) - 2;
}

// A run's 8 node sums (magnitudes below 2^23: a lane's run holds at most CPT
// particles of at most 2^21 units) added to the box or, outside it, with the
// species' sign to this XCD's int64 replica.  In the box, lane l starts at
// stencil node l mod 8, so the lanes of one instruction that flush the same
// cell (sorted data) hit different addresses; off[j] is the box offset of the
// lane's j-th node (box32_offsets), and the sums are rotated to match by three
// conditional shifts.
__device__ inline void box32_offsets(const Box32 &ds, int off[8]) {
  const int rot = __lane_id() & 7;
#pragma unroll
  for (int j = 0; j < 8; ++j) off[j] = node_off((j + rot) & 7, ds.bx, ds.bx * ds.by);
}

__device__ inline void flush32(const KG32 &g, Box32 &ds, int key, const int a[8], const int off[8],
                               unsigned long long *rep) {
  if (key >= 0) {
    const int rot = __lane_id() & 7;
    int b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = a[j];
#pragma unroll
    for (int stage = 1; stage < 8; stage <<= 1) {   // b[j] <- b[(j + stage) & 7] where bit is set
      const bool take = (rot & stage) != 0;
      int t[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) t[j] = b[(j + stage) & 7];
#pragma unroll
      for (int j = 0; j < 8; ++j) b[j] = take ? t[j] : b[j];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) atomicAdd(&ds.box[key + off[j]], (uint32_t)b[j]);
  } else if (key <= -2) {
    const int64_t b = (int64_t)(-(key + 2));
    const int64_t nd[8] = {b, b + 1, b + g.sy + 1, b + g.sy, b + g.sz, b + g.sz + 1,
                           b + g.sz + g.sy + 1, b + g.sz + g.sy};
    const bool neg = ds.neg != 0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (a[q] != 0)
        __hip_atomic_fetch_add(&rep[nd[q]], (unsigned long long)(long long)(neg ? -a[q] : a[q]),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

__device__ inline void box32_flush(const KG32 &g, Box32 &ds, unsigned long long *rep) {
  __syncthreads();
  for (int t = threadIdx.x; t < ds.vol; t += BLOCK) {
    const uint32_t m = ds.box[t];
    if (m != 0u) {
      const unsigned long long v =
          (unsigned long long)(ds.neg ? -(long long)m : (long long)m);
      const int r = (int)small_div((uint32_t)t, ds.mbx), lz = (int)small_div((uint32_t)r, ds.mby);
      const int lx = t - r * ds.bx, ly = r - lz * ds.by;
      __hip_atomic_fetch_add(&rep[node_of(g, ds.x0 + lx, ds.y0 + ly, ds.z0 + lz)], v,
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
}

// In-register run of one stencil key.
struct Run32 {
  int cur;
  int a[8];
  int off[8];   // box offsets of the lane's rotated stencil (box32_offsets)
  __device__ void reset() {
    cur = -1;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = 0;
  }
  __device__ void add(const KG32 &g, Box32 &ds, int key, float U, float fx, float fy, float fz,
                      unsigned long long *rep) {
    if (key != cur) {
      flush32(g, ds, cur, a, off, rep);
      reset();
      cur = key;
    }
    int parts[8];
    fixed_split32(U, fx, fy, fz, parts);
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] += parts[q];
  }
};

__device__ inline void load32(const KP32 &p, int64_t n, int64_t i0, bool full, int c[CPT],
                              float fx[CPT], float fy[CPT], float fz[CPT], float vx[CPT],
                              float vy[CPT], float vz[CPT]) {
  if (full) {
    const int4 cc = *reinterpret_cast<const int4 *>(p.cell + i0);
    const float4 a = *reinterpret_cast<const float4 *>(p.fx + i0);
    const float4 b = *reinterpret_cast<const float4 *>(p.fy + i0);
    const float4 d = *reinterpret_cast<const float4 *>(p.fz + i0);
    const float4 e = *reinterpret_cast<const float4 *>(p.vx + i0);
    const float4 f = *reinterpret_cast<const float4 *>(p.vy + i0);
    const float4 h = *reinterpret_cast<const float4 *>(p.vz + i0);
    c[0] = cc.x; c[1] = cc.y; c[2] = cc.z; c[3] = cc.w;
    fx[0] = a.x; fx[1] = a.y; fx[2] = a.z; fx[3] = a.w;
    fy[0] = b.x; fy[1] = b.y; fy[2] = b.z; fy[3] = b.w;
    fz[0] = d.x; fz[1] = d.y; fz[2] = d.z; fz[3] = d.w;
    vx[0] = e.x; vx[1] = e.y; vx[2] = e.z; vx[3] = e.w;
    vy[0] = f.x; vy[1] = f.y; vy[2] = f.z; vy[3] = f.w;
    vz[0] = h.x; vz[1] = h.y; vz[2] = h.z; vz[3] = h.w;
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int64_t i = i0 + q;
      const bool in = i < n;
      c[q] = in ? p.cell[i] : -1;
      fx[q] = in ? p.fx[i] : 0.0f; fy[q] = in ? p.fy[i] : 0.0f; fz[q] = in ? p.fz[i] : 0.0f;
      vx[q] = in ? p.vx[i] : 0.0f; vy[q] = in ? p.vy[i] : 0.0f; vz[q] = in ? p.vz[i] : 0.0f;
    }
  }
}

__device__ inline void store32(const KP32 &p, int64_t n, int64_t i0, bool full, const int c[CPT],
                               const float fx[CPT], const float fy[CPT], const float fz[CPT],
                               const float vx[CPT], const float vy[CPT], const float vz[CPT]) {
  if (full) {
    *reinterpret_cast<int4 *>(p.cell + i0) = make_int4(c[0], c[1], c[2], c[3]);
    *reinterpret_cast<float4 *>(p.fx + i0) = make_float4(fx[0], fx[1], fx[2], fx[3]);
    *reinterpret_cast<float4 *>(p.fy + i0) = make_float4(fy[0], fy[1], fy[2], fy[3]);
    *reinterpret_cast<float4 *>(p.fz + i0) = make_float4(fz[0], fz[1], fz[2], fz[3]);
    *reinterpret_cast<float4 *>(p.vx + i0) = make_float4(vx[0], vx[1], vx[2], vx[3]);
    *reinterpret_cast<float4 *>(p.vy + i0) = make_float4(vy[0], vy[1], vy[2], vy[3]);
    *reinterpret_cast<float4 *>(p.vz + i0) = make_float4(vz[0], vz[1], vz[2], vz[3]);
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
      const int64_t i = i0 + q;
      if (i < n) {
        p.cell[i] = c[q];
        p.fx[i] = fx[q]; p.fy[i] = fy[q]; p.fz[i] = fz[q];
        p.vx[i] = vx[q]; p.vy[i] = vy[q]; p.vz[i] = vz[q];
      }
    }
  }
}

// The FP32 particle phase in place: gather + push + wall BC (+ fixed-point
// deposit).  Lost particles become holes (cell = -1).  ctr[0] counts losses,
// ctr[1] survivors that left the cells (an error the host reports).
// 5 waves per SIMD: at most 96 VGPRs and 32 KB of LDS per block (4 with a
// B map, whose gather needs more registers).
template <int BM>
__global__ __launch_bounds__(BLOCK, BM == 2 ? 4 : 5) void k_push32(
    KG32 g, KP32 p, int64_t n, const float4 *__restrict__ E4, KBF32 bf, float k1,
    const int *__restrict__ wt, KBC bp, float U, int neg, unsigned long long *rep_base,
    int64_t nodes, int do_dep, double *p_mac, double *wdrep, unsigned long long *lost_ctr,
    unsigned long long *esc_ctr) {
  __shared__ Box32 ds;
  __shared__ WallDiagShared wd;
  wd_begin(wd);
  __syncthreads();
  unsigned long long *rep = do_dep ? rep_base + (int64_t)xcc_id() * nodes : nullptr;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  const bool full = i0 + CPT <= n;
  int c[CPT];
  float fx[CPT], fy[CPT], fz[CPT], vx[CPT], vy[CPT], vz[CPT];
  load32(p, n, i0, full, c, fx, fy, fz, vx, vy, vz);
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (c[q] >= g.ncells) c[q] = -1;
  int ix[CPT], iy[CPT], iz[CPT];
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  int sm[3] = {0, 0, 0}, cnt = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    ix[q] = iy[q] = iz[q] = 0;
    if (c[q] >= 0) {
      decode(g, c[q], ix[q], iy[q], iz[q]);
      ++ix[q]; ++iy[q]; ++iz[q];
      lo[0] = min(lo[0], ix[q]); lo[1] = min(lo[1], iy[q]); lo[2] = min(lo[2], iz[q]);
      hi[0] = max(hi[0], ix[q]); hi[1] = max(hi[1], iy[q]); hi[2] = max(hi[2], iz[q]);
      sm[0] += ix[q]; sm[1] += iy[q]; sm[2] += iz[q];
      ++cnt;
    }
  }
  Run32 ra;
  ra.reset();
  if (do_dep) {
    box32_begin(g, ds, lo, hi, sm, cnt, neg);
    box32_offsets(ds, ra.off);
  }
  int64_t eb = -1;
  float4 f1, f2, f3, f4, f5, f6, f7, f8;
  int lost_local = 0, esc_local = 0;
  int nx[CPT], ny[CPT], nz[CPT];
  // gather + velocity update + move of the lane's particles
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    nx[q] = ny[q] = nz[q] = 0;
    if (c[q] >= 0) {
      float w[8];
      weights32(fx[q], fy[q], fz[q], w);
      const int64_t base = node_of(g, ix[q], iy[q], iz[q]);
      if (base != eb) {   // E of the last gathered cell is reused by the next particle in it
        eb = base;
        const int64_t b1 = eb + g.sy, b2 = eb + g.sz, b3 = eb + g.sz + g.sy;
        f1 = E4[eb]; f2 = E4[eb + 1]; f3 = E4[b1 + 1]; f4 = E4[b1];
        f5 = E4[b2]; f6 = E4[b2 + 1]; f7 = E4[b3 + 1]; f8 = E4[b3];
      }
      float Ex, Ey, Ez;
      sum8(w, f1, f2, f3, f4, f5, f6, f7, f8, Ex, Ey, Ez);
      velocity_update32<BM>(g, bf, w, base, k1, Ex, Ey, Ez, vx[q], vy[q], vz[q]);
      nx[q] = ix[q] + advance(fx[q], vx[q] * g.dtx);
      ny[q] = iy[q] + advance(fy[q], vy[q] * g.dty);
      nz[q] = iz[q] + advance(fz[q], vz[q] * g.dtz);
    }
  }
  // the wall table at the new lower nodes, all loads issued together
  int wc[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) wc[q] = c[q] >= 0 ? wall_entry32(g, wt, nx[q], ny[q], nz[q]) : -1;
  // wall BC, new cell, deposit
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    bool lost = false;
    int label = 0;
    if (c[q] >= 0) {
      const int r = wall32(g, wc[q], bp, nx[q], ny[q], nz[q], fx[q], fy[q], fz[q], vx[q], label);
      if (r == 1 || r == 2) {
        lost = true;
        ++lost_local;
        c[q] = -1;
      } else if (r == 3) {
        ++esc_local;
        c[q] = -1;
      } else {
        c[q] = (nx[q] - 1) + g.n1 * ((ny[q] - 1) + g.n2 * (nz[q] - 1));
        if (do_dep)
          ra.add(g, ds, key32(g, ds, nx[q], ny[q], nz[q]), U, fx[q], fy[q], fz[q], rep);
      }
    }
    if (lost) wd_lost(wd, bp, label, (double)vx[q], (double)vy[q], (double)vz[q], p_mac);
  }
  store32(p, n, i0, full, c, fx, fy, fz, vx, vy, vz);
  if (do_dep) {
    flush32(g, ds, ra.cur, ra.a, ra.off, rep);
    box32_flush(g, ds, rep);
  }
  wd_finish(wd, wdrep);
  int tot = lost_local, esc = esc_local;
  for (int d = 32; d > 0; d >>= 1) {
    tot += __shfl_xor(tot, d);
    esc += __shfl_xor(esc, d);
  }
  if (__lane_id() == 0 && tot > 0) atomicAdd(lost_ctr, (unsigned long long)tot);
  if (__lane_id() == 0 && esc > 0) atomicAdd(esc_ctr, (unsigned long long)esc);
}

// Fixed-point deposit of a set without moving it.
__global__ __launch_bounds__(BLOCK) void k_deposit32(KG32 g, KP32 p, int64_t n, float U, int neg,
                                                     unsigned long long *rep_base,
                                                     int64_t nodes) {
  __shared__ Box32 ds;
  unsigned long long *rep = rep_base + (int64_t)xcc_id() * nodes;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  const bool full = i0 + CPT <= n;
  int c[CPT];
  float fx[CPT], fy[CPT], fz[CPT], vx[CPT], vy[CPT], vz[CPT];
  load32(p, n, i0, full, c, fx, fy, fz, vx, vy, vz);
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (c[q] >= g.ncells) c[q] = -1;
  int ix[CPT], iy[CPT], iz[CPT];
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
  int sm[3] = {0, 0, 0}, cnt = 0;
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    ix[q] = iy[q] = iz[q] = 0;
    if (c[q] >= 0) {
      decode(g, c[q], ix[q], iy[q], iz[q]);
      ++ix[q]; ++iy[q]; ++iz[q];
      lo[0] = min(lo[0], ix[q]); lo[1] = min(lo[1], iy[q]); lo[2] = min(lo[2], iz[q]);
      hi[0] = max(hi[0], ix[q]); hi[1] = max(hi[1], iy[q]); hi[2] = max(hi[2], iz[q]);
      sm[0] += ix[q]; sm[1] += iy[q]; sm[2] += iz[q];
      ++cnt;
    }
  }
  box32_begin(g, ds, lo, hi, sm, cnt, neg);
  Run32 ra;
  ra.reset();
  box32_offsets(ds, ra.off);
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (c[q] >= 0)
      ra.add(g, ds, key32(g, ds, ix[q], iy[q], iz[q]), U, fx[q], fy[q], fz[q], rep);
  flush32(g, ds, ra.cur, ra.a, ra.off, rep);
  box32_flush(g, ds, rep);
}

// dep += sum of the int64 replicas; replicas cleared for the next deposit.
__global__ void k_rep_flush_i64(long long *dep, unsigned long long *rep, int64_t nodes) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  unsigned long long s = (unsigned long long)dep[i];
#pragma unroll
  for (int r = 0; r < LEPIC_XCDS; ++r) {
    s += rep[(int64_t)r * nodes + i];
    rep[(int64_t)r * nodes + i] = 0ull;
  }
  dep[i] = (long long)s;
}

// np += kq * dep * 2^-S.
__global__ void k_fixed_density(double *np, const long long *__restrict__ dep,
                                const double *__restrict__ kq, int64_t nodes, double inv) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  np[i] = np[i] + kq[i] * ((double)dep[i] * inv);
}

// FP64 -> FP32 layout.  The cell is floor(x / h) (correctly rounded
// quotient) and the offset x/h - floor(x/h), exact in FP64, rounded once to
// FP32 (rounding up to 1 moves the particle to the next cell's lower face).
// Dead particles and particles outside [0, L) become holes; ctr counts the
// latter.
__global__ __launch_bounds__(BLOCK) void k_to32(KGrid g, lepic_pset src, KP32 dst, int64_t n,
                                                unsigned long long *ctr) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  int cell = -1;
  float f[3] = {0.0f, 0.0f, 0.0f};
  bool outside = false;
  if (src.dead[i] == 0) {
    const double x[3] = {src.x[i], src.y[i], src.z[i]};
    const QDiv *qd[3] = {&g.q1, &g.q2, &g.q3};
    const int nn[3] = {g.n1, g.n2, g.n3};
    int c[3];
    for (int d = 0; d < 3; ++d) {
      const double q = qdiv(x[d], *qd[d]);
      const double fl = floor(q);
      c[d] = -1;
      if (q >= 0.0 && fl < (double)nn[d]) {
        c[d] = (int)fl;
        f[d] = (float)(q - fl);
        if (f[d] >= 1.0f) {
          f[d] = 0.0f;
          c[d] += 1;
        }
        if (c[d] >= nn[d]) c[d] = -1;
      }
      if (c[d] < 0) outside = true;
    }
    if (outside) f[0] = f[1] = f[2] = 0.0f;   // a hole carries no offsets
    else cell = c[0] + g.n1 * (c[1] + g.n2 * c[2]);
  }
  if (outside) atomicAdd(ctr, 1ull);
  dst.cell[i] = cell;
  dst.fx[i] = f[0]; dst.fy[i] = f[1]; dst.fz[i] = f[2];
  dst.vx[i] = (float)src.vx[i]; dst.vy[i] = (float)src.vy[i]; dst.vz[i] = (float)src.vz[i];
}

// FP32 -> FP64 layout: x = ((ix - 1) + fx) h1 in FP64; holes get dead = 2.
__global__ __launch_bounds__(BLOCK) void k_from32(KG32 g, double h1, double h2, double h3, KP32 src,
                                                  lepic_pset dst, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n) return;
  const int c = src.cell[i];
  int cx = 0, cy = 0, cz = 0;
  if (c >= 0) decode(g, c, cx, cy, cz);
  dst.x[i] = ((double)cx + (double)src.fx[i]) * h1;
  dst.y[i] = ((double)cy + (double)src.fy[i]) * h2;
  dst.z[i] = ((double)cz + (double)src.fz[i]) * h3;
  dst.vx[i] = (double)src.vx[i];
  dst.vy[i] = (double)src.vy[i];
  dst.vz[i] = (double)src.vz[i];
  dst.cell[i] = c;
  dst.dead[i] = c >= 0 ? 0 : 2;
}

__global__ void k_round4(const double4 *__restrict__ a, float4 *b, int64_t nodes) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  const double4 v = a[i];
  b[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}

// ---------------------------------------------------------------------------
// Stable sort by cell (mod_particle_sorting.f90:88-181) that drops holes:
//   count  -- cells per tile in LDS, one global atomic per (tile, cell);
//   scan   -- exclusive prefix of the counts (first slot of every cell);
//   claim  -- every tile takes a run of slots per cell with one atomic and
//             writes its particles' source indices there (order within a
//             cell is the atomics' order at this point);
//   rank   -- one wave per cell puts the cell's source indices in ascending
//             order, which is the stable order;
//   gather -- every slot copies its particle from the source index.
// ---------------------------------------------------------------------------
constexpr int SBOX = 8192;   // cells of a tile's counting box (32 KB)

struct SortShared {
  int mn[3], mx[3];
  int x0, y0, z0, bx, bxy, vol;   // vol < 0: one global atomic per particle
  int cnt[SBOX];
};

__device__ inline void sort_box(SortShared &bs, const int lo[3], const int hi[3]) {
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
      vol = v > SBOX ? -1 : (int)v;
    }
    bs.vol = vol;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < bs.vol; t += BLOCK) bs.cnt[t] = 0;
  __syncthreads();
}

__device__ inline int sort_local(const SortShared &bs, int cx, int cy, int cz) {
  return (cx - bs.x0) + bs.bx * (cy - bs.y0) + bs.bxy * (cz - bs.z0);
}

__device__ inline int sort_cell_of_local(const KG32 &g, const SortShared &bs, int t) {
  const int lz = t / bs.bxy, r = t - lz * bs.bxy;
  const int ly = r / bs.bx, lx = r - ly * bs.bx;
  return (bs.x0 + lx) + g.n1 * ((bs.y0 + ly) + g.n2 * (bs.z0 + lz));
}

__device__ inline void load_cells(const KG32 &g, const int32_t *cell, int64_t n, int64_t i0,
                                  int c[CPT]) {
  if (i0 + CPT <= n) {
    const int4 v = *reinterpret_cast<const int4 *>(cell + i0);
    c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q) c[q] = i0 + q < n ? cell[i0 + q] : -1;
  }
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (c[q] >= g.ncells) c[q] = -1;
}

__global__ __launch_bounds__(BLOCK) void k_sort_count(KG32 g, const int32_t *__restrict__ cell,
                                                      int64_t n, int32_t *count) {
  __shared__ SortShared bs;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  int c[CPT], cx[CPT], cy[CPT], cz[CPT];
  load_cells(g, cell, n, i0, c);
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    cx[q] = cy[q] = cz[q] = 0;
    if (c[q] >= 0) {
      decode(g, c[q], cx[q], cy[q], cz[q]);
      lo[0] = min(lo[0], cx[q]); lo[1] = min(lo[1], cy[q]); lo[2] = min(lo[2], cz[q]);
      hi[0] = max(hi[0], cx[q]); hi[1] = max(hi[1], cy[q]); hi[2] = max(hi[2], cz[q]);
    }
  }
  sort_box(bs, lo, hi);
  if (bs.vol > 0) {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (c[q] >= 0) atomicAdd(&bs.cnt[sort_local(bs, cx[q], cy[q], cz[q])], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < bs.vol; t += BLOCK) {
      const int k = bs.cnt[t];
      if (k) atomicAdd(&count[sort_cell_of_local(g, bs, t)], k);
    }
  } else {
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (c[q] >= 0) atomicAdd(&count[c[q]], 1);
  }
}

// Exclusive scan of count[0..m) into start, in three launches: per-block
// sums, one block over the block sums, then every block adds its offset.
// The total goes to *total.
constexpr int SCAN_B = 1024;   // elements per block (4 per thread)

__device__ inline int block_excl_scan(int v, int *wsum, int &total) {
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  int inc = v;
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(inc, d);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < BLOCK / 64; ++i) {
      const int t = wsum[i];
      wsum[i] = run;
      run += t;
    }
    wsum[BLOCK / 64] = run;
  }
  __syncthreads();
  total = wsum[BLOCK / 64];
  const int r = wsum[w] + inc - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(BLOCK) void k_scan_sums(const int32_t *count, int64_t m,
                                                     int32_t *bsum) {
  __shared__ int wsum[BLOCK / 64 + 1];
  const int64_t b0 = (int64_t)blockIdx.x * SCAN_B + threadIdx.x * 4;
  int s = 0;
  for (int q = 0; q < 4; ++q)
    if (b0 + q < m) s += count[b0 + q];
  int total;
  block_excl_scan(s, wsum, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(BLOCK) void k_scan_top(int32_t *bsum, int64_t nb, int64_t *total_out) {
  __shared__ int wsum[BLOCK / 64 + 1];
  int carry = 0;
  for (int64_t base = 0; base < nb; base += BLOCK) {
    const int64_t i = base + threadIdx.x;
    const int v = i < nb ? bsum[i] : 0;
    int total;
    const int e = block_excl_scan(v, wsum, total);
    if (i < nb) bsum[i] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(BLOCK) void k_scan_apply(const int32_t *count, int64_t m,
                                                      const int32_t *bsum, int32_t *start) {
  __shared__ int wsum[BLOCK / 64 + 1];
  const int64_t b0 = (int64_t)blockIdx.x * SCAN_B + threadIdx.x * 4;
  int v[4], s = 0;
  for (int q = 0; q < 4; ++q) {
    v[q] = b0 + q < m ? count[b0 + q] : 0;
    s += v[q];
  }
  int total;
  int e = block_excl_scan(s, wsum, total) + bsum[blockIdx.x];
  for (int q = 0; q < 4; ++q) {
    if (b0 + q < m) start[b0 + q] = e;
    e += v[q];
  }
}

// Claim: cursor[c] starts at start[c]; each tile takes one run per cell.
__global__ __launch_bounds__(BLOCK) void k_sort_claim(KG32 g, const int32_t *__restrict__ cell,
                                                      int64_t n, int32_t *cursor, int32_t *idx) {
  __shared__ SortShared bs;
  const int64_t i0 = (int64_t)blockIdx.x * CTILE + (int64_t)threadIdx.x * CPT;
  int c[CPT], cx[CPT], cy[CPT], cz[CPT];
  load_cells(g, cell, n, i0, c);
  int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    cx[q] = cy[q] = cz[q] = 0;
    if (c[q] >= 0) {
      decode(g, c[q], cx[q], cy[q], cz[q]);
      lo[0] = min(lo[0], cx[q]); lo[1] = min(lo[1], cy[q]); lo[2] = min(lo[2], cz[q]);
      hi[0] = max(hi[0], cx[q]); hi[1] = max(hi[1], cy[q]); hi[2] = max(hi[2], cz[q]);
    }
  }
  sort_box(bs, lo, hi);
  if (bs.vol < 0) {   // wide tile: one atomic per particle
#pragma unroll
    for (int q = 0; q < CPT; ++q)
      if (c[q] >= 0) idx[atomicAdd(&cursor[c[q]], 1)] = (int32_t)(i0 + q);
    return;   // uniform across the block
  }
  int l[CPT], rank[CPT];
#pragma unroll
  for (int q = 0; q < CPT; ++q) {
    l[q] = c[q] >= 0 ? sort_local(bs, cx[q], cy[q], cz[q]) : 0;
    rank[q] = c[q] >= 0 ? atomicAdd(&bs.cnt[l[q]], 1) : 0;
  }
  __syncthreads();
  // one global atomic per nonempty cell of the box: its run's first slot
  const int vol = bs.vol;
  for (int t = threadIdx.x; t < vol; t += BLOCK) {
    const int k = bs.cnt[t];
    if (k) bs.cnt[t] = atomicAdd(&cursor[sort_cell_of_local(g, bs, t)], k);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < CPT; ++q)
    if (c[q] >= 0) idx[bs.cnt[l[q]] + rank[q]] = (int32_t)(i0 + q);
}

// Rank: one wave per cell sorts the cell's source indices ascending (they are
// distinct).  Up to 256: a bitonic sort of the cell padded to 64 x R
// entries, R per lane (position q*64 + lane in register q: partners 64 or
// more apart in the same lane, nearer ones across lanes); larger cells:
// each index's rank counted over the cell.
template <int R>
__device__ inline void wave_bitonic(int v[R], int lane) {
#pragma unroll
  for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int r = j / 64;
#pragma unroll
        for (int q = 0; q < R; ++q) {
          if ((q & r) != 0) continue;
          const int p = q * 64 + lane;
          const bool up = (p & k) == 0;
          const int a = v[q], b = v[q | r];
          v[q] = up ? min(a, b) : max(a, b);
          v[q | r] = up ? max(a, b) : min(a, b);
        }
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int p = q * 64 + lane;
          const int o = __shfl_xor(v[q], j);
          const bool up = (p & k) == 0;
          const bool low = (lane & j) == 0;
          v[q] = (low == up) ? min(v[q], o) : max(v[q], o);
        }
      }
    }
  }
}

template <int R>
__device__ inline void rank_small(const int32_t *__restrict__ idx, int32_t *sidx, int64_t b, int s,
                                  int lane) {
  int v[R];
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = q * 64 + lane < s ? idx[b + q * 64 + lane] : 0x7FFFFFFF;
  wave_bitonic<R>(v, lane);
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (q * 64 + lane < s) sidx[b + q * 64 + lane] = v[q];
}

__global__ __launch_bounds__(BLOCK) void k_sort_rank(const int32_t *__restrict__ count,
                                                     const int32_t *__restrict__ start,
                                                     int64_t ncells, const int32_t *__restrict__ idx,
                                                     int32_t *sidx) {
  const int64_t c = (int64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
  if (c >= ncells) return;
  const int lane = __lane_id();
  const int s = count[c];
  const int64_t b = start[c];
  if (s == 0) return;
  if (s <= 64) {
    rank_small<1>(idx, sidx, b, s, lane);
  } else if (s <= 128) {
    rank_small<2>(idx, sidx, b, s, lane);
  } else if (s <= 256) {
    rank_small<4>(idx, sidx, b, s, lane);
  } else {
    for (int e = lane; e < s; e += 64) {
      const int v = idx[b + e];
      int r = 0;
      for (int j = 0; j < s; ++j) r += idx[b + j] < v ? 1 : 0;
      sidx[b + r] = v;
    }
  }
}

// Four destination slots per thread: one 16-byte read of the source slots,
// the 28 scattered reads in flight together, 16-byte stores (sidx and the
// destination arrays are 16-byte aligned at multiples of 4 slots).
__global__ __launch_bounds__(BLOCK) void k_sort_gather(KP32 src, KP32 dst,
                                                       const int32_t *__restrict__ sidx,
                                                       int64_t m) {
  const int64_t d0 = (blockIdx.x * (int64_t)BLOCK + threadIdx.x) * 4;
  if (d0 >= m) return;
  if (d0 + 4 <= m) {
    const int4 j = *reinterpret_cast<const int4 *>(sidx + d0);
    const int4 c = make_int4(src.cell[j.x], src.cell[j.y], src.cell[j.z], src.cell[j.w]);
    const float4 fx = make_float4(src.fx[j.x], src.fx[j.y], src.fx[j.z], src.fx[j.w]);
    const float4 fy = make_float4(src.fy[j.x], src.fy[j.y], src.fy[j.z], src.fy[j.w]);
    const float4 fz = make_float4(src.fz[j.x], src.fz[j.y], src.fz[j.z], src.fz[j.w]);
    const float4 vx = make_float4(src.vx[j.x], src.vx[j.y], src.vx[j.z], src.vx[j.w]);
    const float4 vy = make_float4(src.vy[j.x], src.vy[j.y], src.vy[j.z], src.vy[j.w]);
    const float4 vz = make_float4(src.vz[j.x], src.vz[j.y], src.vz[j.z], src.vz[j.w]);
    *reinterpret_cast<int4 *>(dst.cell + d0) = c;
    *reinterpret_cast<float4 *>(dst.fx + d0) = fx;
    *reinterpret_cast<float4 *>(dst.fy + d0) = fy;
    *reinterpret_cast<float4 *>(dst.fz + d0) = fz;
    *reinterpret_cast<float4 *>(dst.vx + d0) = vx;
    *reinterpret_cast<float4 *>(dst.vy + d0) = vy;
    *reinterpret_cast<float4 *>(dst.vz + d0) = vz;
    return;
  }
  for (int64_t d = d0; d < m; ++d) {
    const int64_t j = sidx[d];
    dst.cell[d] = src.cell[j];
    dst.fx[d] = src.fx[j]; dst.fy[d] = src.fy[j]; dst.fz[d] = src.fz[j];
    dst.vx[d] = src.vx[j]; dst.vy[d] = src.vy[j]; dst.vz[d] = src.vz[j];
  }
}

__global__ void k_zero_i32(int32_t *a, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_copy_i32(const int32_t *a, int32_t *b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_zero_u64_2(unsigned long long *a) {
  if (threadIdx.x < 2) a[threadIdx.x] = 0ull;
}

int check_pset32(lepic_ctx *ctx, const lepic_pset32 *p, const char *who) {
  if (!p || p->n < 0 || p->n > p->cap || (p->n > 0 && (!p->cell || !p->fx || !p->fy || !p->fz ||
                                                       !p->vx || !p->vy || !p->vz)))
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": bad particle set");
  const void *a[7] = {p->cell, p->fx, p->fy, p->fz, p->vx, p->vy, p->vz};
  for (const void *q : a)
    if (reinterpret_cast<uintptr_t>(q) % 16 != 0)
      return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": arrays must be 16-byte aligned");
  if (p->n >= ((int64_t)1 << 31) - CTILE)
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": 2^31 particles or more");
  return LEPIC_OK;
}

int check_grid32(lepic_ctx *ctx, const lepic_grid *g, const char *who) {
  int st = check_grid(ctx, g);
  if (st) return st;
  if ((int64_t)g->n[0] * g->n[1] * g->n[2] >= ((int64_t)1 << 31))
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": 2^31 cells or more");
  return LEPIC_OK;
}

// Full weight in units of 2^-S: rint(|Nm / (h1 h2 h3)| 2^S) within [1, 2^21]
// (the FP32 split's exact range).
int fixed_unit(lepic_ctx *ctx, const lepic_grid *g, double Nm, int32_t frac_bits, double &U,
               int &neg, const char *who) {
  const double k1 = Nm / (g->h[0] * g->h[1] * g->h[2]);
  neg = k1 < 0.0 ? 1 : 0;
  if (k1 == 0.0) {
    U = 0.0;
    return LEPIC_OK;
  }
  const double unit = std::ldexp(std::fabs(k1), frac_bits);
  if (frac_bits < -1000 || frac_bits > 1000 || !(unit >= 1.0 && unit <= LEPIC_FIXED32_MAX))
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      std::string(who) + ": |Nm/(h1 h2 h3)| * 2^frac_bits must lie in [1, 2^21]");
  U = std::rint(unit);
  return LEPIC_OK;
}

unsigned long long *int_replicas(lepic_ctx *ctx, int64_t nodes) {
  // the FP64 replicas, zero bits either way
  return reinterpret_cast<unsigned long long *>(lepic_replicas(ctx, nodes));
}

}  // namespace

extern "C" {

int lepic_pset_to32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                    lepic_pset32 *dst, int64_t *n_outside) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid32(ctx, g, "pset_to32")) || (st = check_pset32(ctx, dst, "pset_to32 dst")))
    return st;
  if (!src || src->n < 0 || src->n > dst->cap)
    return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "pset_to32: destination smaller than the source");
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(ctx->d_ctrl + CTRL_COUNT_OUT);
  hipLaunchKernelGGL(k_zero_u64_2, dim3(1), dim3(64), 0, ctx->stream, ctr);
  if (src->n > 0)
    hipLaunchKernelGGL(k_to32, dim3(grid_blocks(src->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       make_kgrid(g), *src, kp32(dst), src->n, ctr);
  LEPIC_LAUNCH_CHECK(ctx, "k_to32");
  dst->n = src->n;
  if (n_outside) {
    LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, ctr, sizeof(int64_t), hipMemcpyDeviceToHost,
                                  ctx->stream));
    LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *n_outside = ctx->h_pinned[0];
  }
  return LEPIC_OK;
}

int lepic_pset_from32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset32 *src,
                      lepic_pset *dst) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid32(ctx, g, "pset_from32")) || (st = check_pset32(ctx, src, "pset_from32")))
    return st;
  if (!dst || dst->cap < src->n || (src->n > 0 && (!dst->x || !dst->dead || !dst->cell)))
    return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "pset_from32: destination smaller than the source");
  if (src->n > 0)
    hipLaunchKernelGGL(k_from32, dim3(grid_blocks(src->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       make_kg32(g, 0.0), g->h[0], g->h[1], g->h[2], kp32(src), *dst, src->n);
  LEPIC_LAUNCH_CHECK(ctx, "k_from32");
  dst->n = src->n;
  return LEPIC_OK;
}

int lepic_round_field4(lepic_ctx *ctx, const double *F4, int64_t nodes, float *F4f) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!F4 || !F4f || nodes < 0) return lepic_fail(ctx, LEPIC_ERR_ARG, "round_field4: bad argument");
  if (nodes > 0)
    hipLaunchKernelGGL(k_round4, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const double4 *>(F4), reinterpret_cast<float4 *>(F4f),
                       nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_round4");
  return LEPIC_OK;
}

int lepic_push32(lepic_ctx *ctx, const lepic_grid *g, lepic_pset32 *p, const float *E4f,
                 const lepic_bfield32 *bf, double q, double m, double dt, const int32_t *bcnd,
                 const lepic_bc_params *bp, int32_t frac_bits, int64_t *dep, double *p_mac,
                 double *P_loss, uint64_t *lost_acc, uint64_t *escaped) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid32(ctx, g, "push32")) || (st = check_pset32(ctx, p, "push32"))) return st;
  if (!E4f || !bcnd || !bp || !p_mac || !P_loss || !lost_acc || !escaped)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push32: null argument");
  if (bp->flag_die == 1)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "push32: dielectric charging runs on the FP64 path only");
  const int bmode = bf ? bf->mode : 0;
  if (bmode < 0 || bmode > 2) return lepic_fail(ctx, LEPIC_ERR_ARG, "push32: B mode must be 0, 1 or 2");
  if (bmode == 2 && !bf->B4) return lepic_fail(ctx, LEPIC_ERR_ARG, "push32: null B map");
  double U = 0.0;
  int neg = 0;
  if (dep && (st = fixed_unit(ctx, g, bp->Nm, frac_bits, U, neg, "push32"))) return st;
  if (p->n == 0) return LEPIC_OK;
  const int64_t nodes = grid_nodes(g);
  unsigned long long *rep = nullptr;
  if (dep) {
    rep = int_replicas(ctx, nodes);
    if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "push32: replicas");
  }
  int32_t *wt = reinterpret_cast<int32_t *>(lepic_scratch(ctx, sizeof(int32_t) * (size_t)nodes, 1));
  if (!wt) return lepic_fail(ctx, LEPIC_ERR_HIP, "push32: wall table");
  hipLaunchKernelGGL(k_wall_table, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     make_kgrid(g), bcnd, wt, nodes);
  const double k1d = dt * q / (2.0 * m);
  const KG32 kg = make_kg32(g, dt);
  const KBF32 kbf = make_kbf32(bf, k1d);
  const float k1 = (float)k1d;
  const int dd = dep ? 1 : 0;
  unsigned long long *lc = reinterpret_cast<unsigned long long *>(lost_acc);
  unsigned long long *ec = reinterpret_cast<unsigned long long *>(escaped);
  const dim3 grid(grid_blocks(p->n, CTILE)), block(BLOCK);
  switch (bmode) {
    case 0: hipLaunchKernelGGL((k_push32<0>), grid, block, 0, ctx->stream, kg, kp32(p), p->n,
                               reinterpret_cast<const float4 *>(E4f), kbf, k1, wt, kbc(bp), (float)U, neg,
                               rep, nodes, dd, p_mac, ctx->wdiag, lc, ec); break;
    case 1: hipLaunchKernelGGL((k_push32<1>), grid, block, 0, ctx->stream, kg, kp32(p), p->n,
                               reinterpret_cast<const float4 *>(E4f), kbf, k1, wt, kbc(bp), (float)U, neg,
                               rep, nodes, dd, p_mac, ctx->wdiag, lc, ec); break;
    default: hipLaunchKernelGGL((k_push32<2>), grid, block, 0, ctx->stream, kg, kp32(p), p->n,
                                reinterpret_cast<const float4 *>(E4f), kbf, k1, wt, kbc(bp), (float)U, neg,
                                rep, nodes, dd, p_mac, ctx->wdiag, lc, ec); break;
  }
  LEPIC_LAUNCH_CHECK(ctx, "k_push32");
  hipLaunchKernelGGL(k_wd_flush, dim3(1), dim3(64), 0, ctx->stream, ctx->wdiag, p_mac, P_loss,
                     std::min(bp->ngrid + 1, LEPIC_WD_LABELS));
  LEPIC_LAUNCH_CHECK(ctx, "k_wd_flush");
  if (dep) {
    hipLaunchKernelGGL(k_rep_flush_i64, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<long long *>(dep), rep, nodes);
    LEPIC_LAUNCH_CHECK(ctx, "k_rep_flush_i64");
  }
  return LEPIC_OK;
}

int lepic_deposit32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset32 *p, double Nm,
                    int32_t frac_bits, int64_t *dep) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid32(ctx, g, "deposit32")) || (st = check_pset32(ctx, p, "deposit32"))) return st;
  if (!dep) return lepic_fail(ctx, LEPIC_ERR_ARG, "deposit32: null accumulator");
  double U = 0.0;
  int neg = 0;
  if ((st = fixed_unit(ctx, g, Nm, frac_bits, U, neg, "deposit32"))) return st;
  if (p->n == 0 || U == 0.0) return LEPIC_OK;
  const int64_t nodes = grid_nodes(g);
  unsigned long long *rep = int_replicas(ctx, nodes);
  if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "deposit32: replicas");
  hipLaunchKernelGGL(k_deposit32, dim3(grid_blocks(p->n, CTILE)), dim3(BLOCK), 0, ctx->stream,
                     make_kg32(g, 0.0), kp32(p), p->n, (float)U, neg, rep, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_deposit32");
  hipLaunchKernelGGL(k_rep_flush_i64, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<long long *>(dep), rep, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_rep_flush_i64");
  return LEPIC_OK;
}

int lepic_fixed_density(lepic_ctx *ctx, const lepic_grid *g, const int64_t *dep, const double *kq,
                        int32_t frac_bits, double *np_local) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g))) return st;
  if (!dep || !kq || !np_local || frac_bits < -1000 || frac_bits > 1000)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "fixed_density: bad argument");
  const int64_t nodes = grid_nodes(g);
  hipLaunchKernelGGL(k_fixed_density, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     np_local, reinterpret_cast<const long long *>(dep), kq, nodes,
                     std::ldexp(1.0, -frac_bits));
  LEPIC_LAUNCH_CHECK(ctx, "k_fixed_density");
  return LEPIC_OK;
}

int lepic_sort32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset32 *src, lepic_pset32 *dst) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid32(ctx, g, "sort32")) || (st = check_pset32(ctx, src, "sort32 src")) ||
      (st = check_pset32(ctx, dst, "sort32 dst")))
    return st;
  if (dst->cap < src->n)
    return lepic_fail(ctx, LEPIC_ERR_CAPACITY, "sort32: destination smaller than the source");
  const int64_t n = src->n;
  if (n == 0) {
    dst->n = 0;
    return LEPIC_OK;
  }
  const int64_t ncells = (int64_t)g->n[0] * g->n[1] * g->n[2];
  const int64_t nb = (ncells + SCAN_B - 1) / SCAN_B;
  // scratch: count, start, cursor (ncells each), block sums, idx and sidx (n each)
  const size_t words = (size_t)(3 * ncells + nb + 2 * n + 64);
  int32_t *w = reinterpret_cast<int32_t *>(lepic_scratch(ctx, sizeof(int32_t) * words, 0));
  if (!w) return lepic_fail(ctx, LEPIC_ERR_HIP, "sort32: scratch");
  int32_t *count = w, *start = w + ncells, *cursor = w + 2 * ncells, *bsum = w + 3 * ncells;
  int32_t *idx = bsum + nb + 16;
  idx += (16 - (reinterpret_cast<uintptr_t>(idx) / 4) % 16) % 16;
  int32_t *sidx = idx + ((n + 15) & ~(int64_t)15);   // 64-byte aligned, like idx
  int64_t *total = ctx->d_ctrl + CTRL_COUNT_OUT;
  const KG32 kg = make_kg32(g, 0.0);
  const hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_blocks(ncells, 256)), dim3(256), 0, s, count, ncells);
  hipLaunchKernelGGL(k_sort_count, dim3(grid_blocks(n, CTILE)), dim3(BLOCK), 0, s, kg, src->cell,
                     n, count);
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(BLOCK), 0, s, count, ncells, bsum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BLOCK), 0, s, bsum, nb, total);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(BLOCK), 0, s, count, ncells, bsum,
                     start);
  hipLaunchKernelGGL(k_copy_i32, dim3(grid_blocks(ncells, 256)), dim3(256), 0, s, start, cursor,
                     ncells);
  hipLaunchKernelGGL(k_sort_claim, dim3(grid_blocks(n, CTILE)), dim3(BLOCK), 0, s, kg, src->cell,
                     n, cursor, idx);
  hipLaunchKernelGGL(k_sort_rank, dim3(grid_blocks(ncells, BLOCK / 64)), dim3(BLOCK), 0, s, count,
                     start, ncells, idx, sidx);
  LEPIC_LAUNCH_CHECK(ctx, "sort32 passes");
  LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  LEPIC_HIP(ctx, hipStreamSynchronize(s));
  const int64_t m = ctx->h_pinned[0];
  if (m < 0 || m > n) return lepic_fail(ctx, LEPIC_ERR_HIP, "sort32: inconsistent count");
  if (m > 0)
    hipLaunchKernelGGL(k_sort_gather, dim3(grid_blocks((m + 3) / 4, BLOCK)), dim3(BLOCK), 0, s,
                       kp32(src), kp32(dst), sidx, m);
  LEPIC_LAUNCH_CHECK(ctx, "k_sort_gather");
  dst->n = m;
  return LEPIC_OK;
}

}  // extern "C"
