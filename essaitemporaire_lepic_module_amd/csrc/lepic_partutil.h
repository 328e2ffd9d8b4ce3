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
