// lepic_partutil.h -- helpers shared by the particle kernels of both
// precisions: wall-BC parameters, wave reductions, the wall diagnostics of
// apply_particle_bc_legacy and the stencil node offsets.
#pragma once

#include "lepic_internal.h"

namespace {

struct KBC {
  double xmax, ymax, zmax;
  int flag_pbc, flag_nmn, ptype, tag_neg;
  double qmacro, mass, Nm;
  int ngrid;
  int flag_die;
  const int32_t *dtype;
  double *sum_q_xz, *sum_q_yz;
};

inline KBC kbc(const lepic_bc_params *b) {
  KBC k;
  k.xmax = b->xmax; k.ymax = b->ymax; k.zmax = b->zmax;
  k.flag_pbc = b->flag_pbc; k.flag_nmn = b->flag_nmn;
  k.ptype = b->ptype; k.tag_neg = b->tag_neg;
  k.qmacro = b->qmacro; k.mass = b->mass; k.Nm = b->Nm; k.ngrid = b->ngrid;
  k.flag_die = b->flag_die;
  k.dtype = b->dtype;
  k.sum_q_xz = b->sum_q_xz;
  k.sum_q_yz = b->sum_q_yz;
  return k;
}

__device__ inline int wave_min(int v) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d));
  return v;
}
__device__ inline int wave_max(int v) {
  for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d));
  return v;
}

__device__ inline double wave_sum_d(double v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Wall diagnostics of apply_particle_bc_legacy (mod_particleBC.f90:195-206):
// p_mac(1,label) += Nm q, p_mac(2,label) += |Nm q| Ek[eV], P_loss += Nm Ek[J]
// for each lost particle.  Done for a whole wave: lanes with lost set
// contribute, and the wave's sums per electrode label go out with one atomic
// each instead of every lost particle hitting the same words.  Every lane of
// the wave calls it.
__device__ inline void wall_diag_wave(const KBC &bp, bool lost, int label, double vx, double vy,
                                      double vz, double *p_mac, double *P_loss) {
  const uint64_t any = __ballot(lost);
  if (any == 0) return;
  const int lane = __lane_id();
  double Ek_J = 0.0, Ek_eV = 0.0;
  if (lost) {
    const double v2 = vx * vx + vy * vy + vz * vz;
    Ek_eV = 0.5 * bp.mass * v2 / 1.60217646e-19;
    Ek_J = 0.5 * bp.mass * v2;
  }
  const double pl = wave_sum_d(lost ? bp.Nm * Ek_J : 0.0);
  if (lane == __ffsll((unsigned long long)any) - 1) atomicAdd(P_loss, pl);
  uint64_t rem = __ballot(lost && label <= bp.ngrid);
  while (rem) {
    const int leader = __ffsll((unsigned long long)rem) - 1;
    const int L = __shfl(label, leader);
    const bool mine = lost && label <= bp.ngrid && label == L;
    const double q = wave_sum_d(mine ? bp.qmacro : 0.0);
    const double e = wave_sum_d(mine ? fabs(bp.qmacro) * Ek_eV : 0.0);
    if (lane == leader) {
      atomicAdd(&p_mac[2 * L + 0], q);
      atomicAdd(&p_mac[2 * L + 1], e);
    }
    rem &= ~__ballot(mine);
  }
}

// Wall diagnostics summed per block (the production pushes): each lost
// particle adds its p_mac and P_loss terms to the block's LDS sums (labels
// below WD_LABELS; higher labels go straight to p_mac), and the block adds
// its sums once into this XCD's replica of the diagnostics with L2-scope
// atomics; k_wd_flush adds the replicas into p_mac / P_loss and clears them.
// (Per-wave device-scope atomics on p_mac / P_loss serialised on those few
// words and dominated the electron push, which loses ~10^5 particles per
// step at 128^3.)
constexpr int WD_LABELS = LEPIC_WD_LABELS;
constexpr int WD_STRIDE = 2 * WD_LABELS + 1;   // p_mac pairs, then P_loss

struct WallDiagShared {
  double s[WD_STRIDE];
};

// Zeroes the block's sums; a __syncthreads must follow before wd_lost.
__device__ inline void wd_begin(WallDiagShared &wd) {
  for (int t = threadIdx.x; t < WD_STRIDE; t += blockDim.x) wd.s[t] = 0.0;
}

// One lost particle (mod_particleBC.f90:195-206): p_mac(1,label) += Nm q,
// p_mac(2,label) += |Nm q| Ek[eV], P_loss += Nm Ek[J].
__device__ inline void wd_lost(WallDiagShared &wd, const KBC &bp, int label, double vx, double vy,
                               double vz, double *p_mac) {
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double Ek_eV = 0.5 * bp.mass * v2 / 1.60217646e-19;
  const double Ek_J = 0.5 * bp.mass * v2;
  atomicAdd(&wd.s[2 * WD_LABELS], bp.Nm * Ek_J);
  if (label > bp.ngrid) return;
  if (label < WD_LABELS) {
    atomicAdd(&wd.s[2 * label + 0], bp.qmacro);
    atomicAdd(&wd.s[2 * label + 1], fabs(bp.qmacro) * Ek_eV);
  } else {
    atomicAdd(&p_mac[2 * label + 0], bp.qmacro);
    atomicAdd(&p_mac[2 * label + 1], fabs(bp.qmacro) * Ek_eV);
  }
}

// The block's sums into this XCD's replica (rep: LEPIC_XCDS x WD_STRIDE).
// Every thread calls it after its last wd_lost.
__device__ inline void wd_finish(WallDiagShared &wd, double *rep) {
  __syncthreads();
  double *r = rep + (int64_t)xcc_id() * WD_STRIDE;
  for (int t = threadIdx.x; t < WD_STRIDE; t += blockDim.x) {
    const double v = wd.s[t];
    if (v != 0.0) l2_atomic_add(&r[t], v);
  }
}

__global__ void k_wd_flush(double *rep, double *p_mac, double *P_loss, int npairs) {
  const int t = threadIdx.x;
  if (t >= WD_STRIDE) return;
  double v = 0.0;
  for (int x = 0; x < LEPIC_XCDS; ++x) {
    v += rep[x * WD_STRIDE + t];
    rep[x * WD_STRIDE + t] = 0.0;
  }
  if (t == 2 * WD_LABELS) {
    if (v != 0.0) P_loss[0] += v;
  } else if (t < 2 * npairs && v != 0.0) {
    p_mac[t] += v;
  }
}

// floor(t / d) for t, d < 2^16 by one multiply with m = ceil(2^32 / d): with
// m d = 2^32 + e, 0 <= e < d, t m / 2^32 = t / d + t e / (d 2^32) and
// t e < t d < 2^32, so the floor is t's quotient.  Used to decode box-local
// node numbers without a division per node.
__host__ __device__ inline uint64_t small_div_magic(uint32_t d) {
  return ((1ull << 32) + d - 1) / d;
}
__device__ inline uint32_t small_div(uint32_t t, uint64_t m) {
  return (uint32_t)(((uint64_t)t * m) >> 32);
}

// Offset of stencil node q (the reference's order w1..w8) from the lower node.
__device__ inline int node_off(int q, int bx, int bxy) {
  const int r = q & 3;
  const int dx = (r == 1 || r == 2) ? 1 : 0, dy = r >> 1;
  return dx + dy * bx + (q >> 2) * bxy;
}

// Per-node wall table of the wall test (mod_particleBC.f90:96-116):
// wt[node] = bcnd[node] when all 8 nodes of the cell whose lower corner is
// node have bcnd >= 1 (a particle there is lost), -1 otherwise.
__global__ void k_wall_table(KGrid g, const int *__restrict__ bcnd, int *wt, int64_t nodes) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  const int i = (int)(t % g.sy);
  const int64_t r = t / g.sy;
  const int j = (int)(r % (g.n2 + 3)), k = (int)(r / (g.n2 + 3));
  int v = -1;
  if (i <= g.n1 + 1 && j <= g.n2 + 1 && k <= g.n3 + 1) {
    const int c0 = bcnd[t], c1 = bcnd[t + 1], c2 = bcnd[t + g.sy + 1], c3 = bcnd[t + g.sy];
    const int c4 = bcnd[t + g.sz], c5 = bcnd[t + g.sz + 1], c6 = bcnd[t + g.sz + g.sy + 1],
              c7 = bcnd[t + g.sz + g.sy];
    if (c0 >= 1 && c1 >= 1 && c2 >= 1 && c3 >= 1 && c4 >= 1 && c5 >= 1 && c6 >= 1 && c7 >= 1)
      v = c0;
  }
  wt[t] = v;
}

inline int check_grid(lepic_ctx *ctx, const lepic_grid *g) {
  if (!g || g->n[0] < 1 || g->n[1] < 1 || g->n[2] < 1 || !(g->h[0] > 0) || !(g->h[1] > 0) ||
      !(g->h[2] > 0))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "bad grid");
  if (!qdiv_divisor_ok(g->h[0]) || !qdiv_divisor_ok(g->h[1]) || !qdiv_divisor_ok(g->h[2]))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "grid step outside [1e-200, 1e200]");
  if (grid_nodes(g) >= ((int64_t)1 << 31) - 2)   // node ids travel as int32 in the deposit
    return lepic_fail(ctx, LEPIC_ERR_ARG, "grid has 2^31 nodes or more");
  return LEPIC_OK;
}

}  // namespace
