// lepic_adjoint.hip -- the adjoint (reverse-mode) kernels of the
// differentiable step: the Boris push with the E gather, the E field, the
// charge deposit, the Dirichlet potential coupling and the wall-flux loss.
//
// The forward step they differentiate is the FP64 production step on
// particles held in place (holes instead of compaction, so a particle keeps
// its slot through the run): deposit of x_k -> rho -> multigrid phi with the
// electrodes at their voltages -> E = -grad phi (mod_electricField.f90) ->
// v_{k+1} = Boris(v_k, E(x_k), B) (Src/part_expmover.f90:214-246) ->
// x_{k+1} = x_k + dt v_{k+1} -> wall test.  Discrete events (which particles
// a wall absorbs, and when) are frozen: the adjoint is the derivative of the
// step with that event sequence held fixed.
//
// Each kernel is the transpose of its forward operator, element for
// element: the gather's transpose scatters dL/dE_p with the same trilinear
// weights (the deposit's form), the deposit's transpose gathers dL/drho with
// the weights' position derivatives, the E field's transpose is applied as
// wall-node scatter then plasma-node gather.
#include <cmath>

#include "lepic_internal.h"
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function"   // the wall-table kernel is not used here
#include "lepic_partutil.h"
#pragma clang diagnostic pop
#include "lepic_ewall.h"

namespace {

constexpr int BLOCK = 256;

struct KAdj {
  double *lx, *ly, *lz, *lvx, *lvy, *lvz;
};

inline KAdj kadj(const lepic_adjoint_set *a) {
  KAdj k{a->lx, a->ly, a->lz, a->lvx, a->lvy, a->lvz};
  return k;
}

// Position derivatives of the trilinear weights, contracted with per-node
// values e[0..7] (node order w1..w8): returns dL/dpx, dL/dpy, dL/dpz where
// L = sum_c w_c e_c and px = ix - x/h (so dL/dx = -dL/dpx / h).
__device__ inline void weight_grads(double px, double py, double pz, const double e[8],
                                    double &gx, double &gy, double &gz) {
  const double qx = 1.0 - px, qy = 1.0 - py, qz = 1.0 - pz;
  gx = py * pz * (e[0] - e[1]) + qy * pz * (e[3] - e[2]) + py * qz * (e[4] - e[5]) +
       qy * qz * (e[7] - e[6]);
  gy = px * pz * (e[0] - e[3]) + qx * pz * (e[1] - e[2]) + px * qz * (e[4] - e[7]) +
       qx * qz * (e[5] - e[6]);
  gz = px * py * (e[0] - e[4]) + qx * py * (e[1] - e[5]) + qx * qy * (e[2] - e[6]) +
       px * qy * (e[3] - e[7]);
}

// The mover's stencil lower node (make_stencil): ix = int(x/h) + 1 clamped
// to [0, n + 1] per axis.
__device__ inline void stencil_lo(const KGrid &g, double xp, double yp, double zp, int &ix,
                                  int &iy, int &iz) {
  ix = imax(0, imin(g.n1 + 1, (int)qdiv(xp, g.q1) + 1));
  iy = imax(0, imin(g.n2 + 1, (int)qdiv(yp, g.q2) + 1));
  iz = imax(0, imin(g.n3 + 1, (int)qdiv(zp, g.q3) + 1));
}

// The mover's stencil (make_stencil) with its fractions kept.
__device__ inline void stencil_fracs(const KGrid &g, double xp, double yp, double zp,
                                     int64_t &base, double &px, double &py, double &pz,
                                     int &ix, int &iy, int &iz) {
  stencil_lo(g, xp, yp, zp, ix, iy, iz);
  px = qdiv((double)ix * g.h1 - xp, g.q1);
  py = qdiv((double)iy * g.h2 - yp, g.q2);
  pz = qdiv((double)iz * g.h3 - zp, g.q3);
  base = gidx(g, ix, iy, iz);
}

__device__ inline void node8(const KGrid &g, int64_t b, int64_t nd[8]) {
  nd[0] = b; nd[1] = b + 1; nd[2] = b + g.sy + 1; nd[3] = b + g.sy;
  nd[4] = b + g.sz; nd[5] = b + g.sz + 1; nd[6] = b + g.sz + g.sy + 1; nd[7] = b + g.sz + g.sy;
}

struct KBAdj {
  int use_b;
  double B[3];
  double k2;        // k1 * 2 / (1 + (k1 |B|)^2), formed as the forward forms it
  double dk2;       // dk2 / d(|B|^2)
  const double4 *B4;   // map on the E grid (MAP kernels): 4 doubles per node
  int magnetized;      // map: 0 when this species ignores B
};

// Adjoint of the push of one particle: given dL/dx', dL/dv' of the state
// after the step (x' = x + dt v', v' = Boris(v, E_p, B)), returns dL/dx and
// dL/dv of the state before it (the gather's position term included),
// scatters dL/dE to the node replicas and accumulates dL/dB.  A lane takes
// APT consecutive particles and sums the dL/dE of consecutive particles with
// the same stencil in registers (one cell of cell-sorted particles), so a run
// costs one set of 24 atomics instead of one per particle, and those go to
// the block's LDS box (AdjBox) when the run's stencil lies in it.
constexpr int APT = 4;

__device__ inline void adj_flush(const KGrid &g, double *dst, int64_t b, const double acc[8][3]) {
  if (b < 0) return;
  int64_t nd[8];
  node8(g, b, nd);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double *d = dst + 4 * nd[c];
    l2_atomic_add(d + 0, acc[c][0]);
    l2_atomic_add(d + 1, acc[c][1]);
    l2_atomic_add(d + 2, acc[c][2]);
  }
}

// Per-block LDS box of dL/dE (3 doubles per node) over the block's stencils.
// Global atomics on the replicas, not arithmetic, bound the push adjoint
// (it runs 6x faster with them removed); cell-sorted particles keep a
// block's stencils within a few cells, so their runs add into LDS and each
// box node goes to the replica once.  Runs outside the box (a block spread
// wider than the box, drifted particles) go to the replica directly.
constexpr int ABOX = 1536;                            // nodes: 36 KB of LDS
constexpr int ABOX_X = 24, ABOX_Y = 8, ABOX_Z = 8;    // cut extents, product <= ABOX
static_assert(ABOX_X * ABOX_Y * ABOX_Z <= ABOX, "adjoint box limits exceed the box");
struct AdjBox {
  int mn[3], mx[3], sum[3], cnt;
  int x0, y0, z0, ex, ey, ez, vol;   // vol 0: no box
  double v[3 * ABOX];
};

// lo / hi / sm / cnt: the lane's lower nodes' bounds, sum and count.  The
// box spans the block's lower nodes and one node above (the stencils), cut
// to ABOX_* around the mean lower node when that exceeds ABOX nodes.
__device__ inline void adj_box_begin(const KGrid &g, AdjBox &bx, const int lo[3], const int hi[3],
                                     const int sm[3], int cnt) {
  if (threadIdx.x == 0) {
    bx.mn[0] = bx.mn[1] = bx.mn[2] = 1 << 30;
    bx.mx[0] = bx.mx[1] = bx.mx[2] = -(1 << 30);
    bx.sum[0] = bx.sum[1] = bx.sum[2] = 0;
    bx.cnt = 0;
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
    atomicMin(&bx.mn[0], a0); atomicMin(&bx.mn[1], a1); atomicMin(&bx.mn[2], a2);
    atomicMax(&bx.mx[0], b0); atomicMax(&bx.mx[1], b1); atomicMax(&bx.mx[2], b2);
    atomicAdd(&bx.sum[0], s0); atomicAdd(&bx.sum[1], s1); atomicAdd(&bx.sum[2], s2);
    atomicAdd(&bx.cnt, c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int vol = 0;
    if (bx.cnt > 0) {
      const int top[3] = {g.n1 + 2, g.n2 + 2, g.n3 + 2};
      const int lim[3] = {ABOX_X, ABOX_Y, ABOX_Z};
      int a[3], e[3];
      for (int d = 0; d < 3; ++d) {
        a[d] = bx.mn[d];
        e[d] = min(bx.mx[d] + 1, top[d]) - a[d] + 1;
      }
      if ((int64_t)e[0] * e[1] * e[2] > ABOX) {
        for (int d = 0; d < 3; ++d) {
          if (e[d] <= lim[d]) continue;
          int st = bx.sum[d] / bx.cnt - lim[d] / 2;
          st = min(max(st, a[d]), a[d] + e[d] - lim[d]);
          a[d] = st;
          e[d] = lim[d];
        }
      }
      bx.x0 = a[0]; bx.y0 = a[1]; bx.z0 = a[2];
      bx.ex = e[0]; bx.ey = e[1]; bx.ez = e[2];
      vol = e[0] * e[1] * e[2];
    }
    bx.vol = vol;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * bx.vol; t += BLOCK) bx.v[t] = 0.0;
  __syncthreads();
}

// A run's 8 nodes x 3 components: into the box when its stencil lies inside
// it, else to the replica.
__device__ inline void adj_flush_box(const KGrid &g, AdjBox &bx, double *dst, int64_t b, int ix,
                                     int iy, int iz, const double acc[8][3]) {
  if (b < 0) return;
  const int lx = ix - bx.x0, ly = iy - bx.y0, lz = iz - bx.z0;
  if (bx.vol > 0 && lx >= 0 && ly >= 0 && lz >= 0 && lx + 1 < bx.ex && ly + 1 < bx.ey &&
      lz + 1 < bx.ez) {
    const int l = lx + bx.ex * (ly + bx.ey * lz);
    const int sy = bx.ex, sz = bx.ex * bx.ey;
    const int o[8] = {0, 1, sy + 1, sy, sz, sz + 1, sz + sy + 1, sz + sy};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      double *d = &bx.v[3 * (l + o[c])];
      atomicAdd(d + 0, acc[c][0]);
      atomicAdd(d + 1, acc[c][1]);
      atomicAdd(d + 2, acc[c][2]);
    }
  } else {
    adj_flush(g, dst, b, acc);
  }
}

// The box's nonzero sums to this XCD's replica.
__device__ inline void adj_box_flush(const KGrid &g, AdjBox &bx, double *dst) {
  __syncthreads();
  const int exy = bx.ex * bx.ey;
  for (int t = threadIdx.x; t < bx.vol; t += BLOCK) {
    const int lz = t / exy, r = t - lz * exy, ly = r / bx.ex, lx = r - ly * bx.ex;
    double *d = dst + 4 * gidx(g, bx.x0 + lx, bx.y0 + ly, bx.z0 + lz);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double v = bx.v[3 * t + k];
      if (v != 0.0) l2_atomic_add(d + k, v);
    }
  }
}

// MAP: B is a map on the E grid gathered with the particle's stencil and
// gated at its lower node (|B| > 1e-5, Src/part_expmover.f90:158), so B,
// k2 and dk2 are per particle and the B gather adds its own position term;
// dL/dB of a map is not accumulated (the map is not a parameter here).
template <bool MAP>
__global__ __launch_bounds__(BLOCK) void k_adj_push(KGrid g, lepic_pset st, KAdj a, int64_t n,
                                                    const double4 *__restrict__ E4, KBAdj bf,
                                                    double k1, double dt, double *rep,
                                                    int64_t nodes, double *lamB) {
  __shared__ AdjBox bx;
  const int64_t i0 = (blockIdx.x * (int64_t)BLOCK + threadIdx.x) * APT;
  double *dst = rep + (int64_t)xcc_id() * 4 * nodes;
  {
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    int sm[3] = {0, 0, 0}, cnt = 0;
#pragma unroll
    for (int q = 0; q < APT; ++q) {
      const int64_t i = i0 + q;
      if (i >= n || st.dead[i] != 0) continue;
      int jx, jy, jz;
      stencil_lo(g, st.x[i], st.y[i], st.z[i], jx, jy, jz);
      lo[0] = min(lo[0], jx); lo[1] = min(lo[1], jy); lo[2] = min(lo[2], jz);
      hi[0] = max(hi[0], jx); hi[1] = max(hi[1], jy); hi[2] = max(hi[2], jz);
      sm[0] += jx; sm[1] += jy; sm[2] += jz;
      ++cnt;
    }
    adj_box_begin(g, bx, lo, hi, sm, cnt);
  }
  double lb0 = 0.0, lb1 = 0.0, lb2 = 0.0;
  double acc[8][3];
  int64_t cur = -1;
  int cx = 0, cy = 0, cz = 0;   // the run's lower node
#pragma unroll
  for (int q = 0; q < APT; ++q) {
    const int64_t i = i0 + q;
    if (i >= n || st.dead[i] != 0) continue;
    const double x = st.x[i], y = st.y[i], z = st.z[i];
    const double v[3] = {st.vx[i], st.vy[i], st.vz[i]};
    double lx = a.lx[i], ly = a.ly[i], lz = a.lz[i];
    double lam[3] = {a.lvx[i] + dt * lx, a.lvy[i] + dt * ly, a.lvz[i] + dt * lz};
    int64_t b;
    double px, py, pz;
    int jx, jy, jz;
    stencil_fracs(g, x, y, z, b, px, py, pz, jx, jy, jz);
    int64_t nd[8];
    node8(g, b, nd);
    const double qx = 1.0 - px, qy = 1.0 - py, qz = 1.0 - pz;
    const double w[8] = {px * py * pz, qx * py * pz, qx * qy * pz, px * qy * pz,
                         px * py * qz, qx * py * qz, qx * qy * qz, px * qy * qz};
    double4 Ec[8];
    double E[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      Ec[c] = E4[nd[c]];
      E[0] += w[c] * Ec[c].x;
      E[1] += w[c] * Ec[c].y;
      E[2] += w[c] * Ec[c].z;
    }
    double lv[3], lE[3];
    bool use_b = bf.use_b != 0;
    double Bm[3] = {0.0, 0.0, 0.0}, k2 = bf.k2, dk2 = bf.dk2;
    double4 Bc[8];
    if (MAP) {
      use_b = bf.magnetized && bf.B4[nd[0]].w > 1.0e-5;
      if (use_b) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          Bc[c] = bf.B4[nd[c]];
          Bm[0] += w[c] * Bc[c].x;
          Bm[1] += w[c] * Bc[c].y;
          Bm[2] += w[c] * Bc[c].z;
        }
        const double B2 = Bm[0] * Bm[0] + Bm[1] * Bm[1] + Bm[2] * Bm[2];
        const double Bt = sqrt(B2);
        k2 = k1 * (2.0 / (1.0 + (k1 * Bt) * (k1 * Bt)));
        const double den = 1.0 + k1 * k1 * B2;
        dk2 = -2.0 * k1 * k1 * k1 / (den * den);
      }
    }
    double pb0 = 0.0, pb1 = 0.0, pb2 = 0.0;   // dL/dB at this particle
    if (use_b) {
      const double *B = MAP ? Bm : bf.B;
      const double v1[3] = {v[0] + k1 * E[0], v[1] + k1 * E[1], v[2] + k1 * E[2]};
      const double v3[3] = {v1[0] + k1 * (v1[1] * B[2] - v1[2] * B[1]),
                            v1[1] + k1 * (v1[2] * B[0] - v1[0] * B[2]),
                            v1[2] + k1 * (v1[0] * B[1] - v1[1] * B[0])};
      // v2 = v1 + k2 (v3 x B), v' = v2 + k1 E
      const double *l2 = lam;
      const double l3[3] = {k2 * (B[1] * l2[2] - B[2] * l2[1]),
                            k2 * (B[2] * l2[0] - B[0] * l2[2]),
                            k2 * (B[0] * l2[1] - B[1] * l2[0])};
      const double lk2 = l2[0] * (v3[1] * B[2] - v3[2] * B[1]) +
                         l2[1] * (v3[2] * B[0] - v3[0] * B[2]) +
                         l2[2] * (v3[0] * B[1] - v3[1] * B[0]);
      // dL/dB from v3 x B (through k2's factor) and from k2(|B|^2)
      pb0 = k2 * (l2[1] * v3[2] - l2[2] * v3[1]) + lk2 * dk2 * 2.0 * B[0];
      pb1 = k2 * (l2[2] * v3[0] - l2[0] * v3[2]) + lk2 * dk2 * 2.0 * B[1];
      pb2 = k2 * (l2[0] * v3[1] - l2[1] * v3[0]) + lk2 * dk2 * 2.0 * B[2];
      // v3 = v1 + k1 (v1 x B)
      const double l1[3] = {l2[0] + l3[0] + k1 * (B[1] * l3[2] - B[2] * l3[1]),
                            l2[1] + l3[1] + k1 * (B[2] * l3[0] - B[0] * l3[2]),
                            l2[2] + l3[2] + k1 * (B[0] * l3[1] - B[1] * l3[0])};
      pb0 += k1 * (l3[1] * v1[2] - l3[2] * v1[1]);
      pb1 += k1 * (l3[2] * v1[0] - l3[0] * v1[2]);
      pb2 += k1 * (l3[0] * v1[1] - l3[1] * v1[0]);
      if (!MAP) {
        lb0 += pb0;
        lb1 += pb1;
        lb2 += pb2;
      }
      for (int d = 0; d < 3; ++d) {
        lv[d] = l1[d];
        lE[d] = k1 * lam[d] + k1 * l1[d];
      }
    } else {
      for (int d = 0; d < 3; ++d) {
        lv[d] = lam[d];
        lE[d] = 2.0 * k1 * lam[d];
      }
    }
    // gather transpose: dL/dE_c += w_c dL/dE_p, summed over the run;
    // position term of the weights
    if (b != cur) {
      adj_flush_box(g, bx, dst, cur, cx, cy, cz, acc);
      cur = b;
      cx = jx; cy = jy; cz = jz;
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c][0] = acc[c][1] = acc[c][2] = 0.0;
    }
    double e[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      e[c] = Ec[c].x * lE[0] + Ec[c].y * lE[1] + Ec[c].z * lE[2];
      if (MAP && use_b) e[c] += Bc[c].x * pb0 + Bc[c].y * pb1 + Bc[c].z * pb2;
      acc[c][0] += w[c] * lE[0];
      acc[c][1] += w[c] * lE[1];
      acc[c][2] += w[c] * lE[2];
    }
    double gx, gy, gz;
    weight_grads(px, py, pz, e, gx, gy, gz);
    lx -= gx / g.h1;
    ly -= gy / g.h2;
    lz -= gz / g.h3;
    a.lx[i] = lx; a.ly[i] = ly; a.lz[i] = lz;
    a.lvx[i] = lv[0]; a.lvy[i] = lv[1]; a.lvz[i] = lv[2];
  }
  adj_flush_box(g, bx, dst, cur, cx, cy, cz, acc);
  adj_box_flush(g, bx, dst);
  if (!MAP && bf.use_b) {
    __shared__ double red[3][BLOCK / LEPIC_WAVE];
    lb0 = wave_sum_d(lb0); lb1 = wave_sum_d(lb1); lb2 = wave_sum_d(lb2);
    const int w = threadIdx.x / LEPIC_WAVE;
    if (__lane_id() == 0) { red[0][w] = lb0; red[1][w] = lb1; red[2][w] = lb2; }
    __syncthreads();
    if (threadIdx.x < 3) {
      double s = 0.0;
      for (int q = 0; q < BLOCK / LEPIC_WAVE; ++q) s += red[threadIdx.x][q];
      atomicAdd(&lamB[threadIdx.x], s);
    }
  }
}

__global__ void k_adj_rep_flush(double *lamE, double *rep, int64_t words) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= words) return;
  double s = lamE[t];
#pragma unroll
  for (int r = 0; r < LEPIC_XCDS; ++r) {
    s += rep[(int64_t)r * words + t];
    rep[(int64_t)r * words + t] = 0.0;
  }
  lamE[t] = s;
}

__device__ inline bool physical_node(const KGrid &g, int i, int j, int k) {
  return i >= 1 && i <= g.n1 + 1 && j >= 1 && j <= g.n2 + 1 && k >= 1 && k <= g.n3 + 1;
}

__device__ inline void ijk_of(const KGrid &g, int64_t c, int &i, int &j, int &k) {
  i = (int)(c % g.sy);
  const int64_t r = c / g.sy;
  j = (int)(r % (g.n2 + 3));
  k = (int)(r / (g.n2 + 3));
}

// E-field transpose, part 1: electrode nodes next to plasma
// (lepic_ewall.h): the tangential components are central differences of
// phi, the normal one the chain 2 a0 - (2 a1 - (... - term)) over pre-pass
// values, so dL/dE of the node goes to lamphi and, with alternating signs,
// to the pre-pass adjoint lamE1 of the chain's nodes.
__global__ void k_adj_e_walls(KGrid g, const int32_t *__restrict__ bcnd,
                              const double4 *__restrict__ lamE, double *lamE1, double *lamphi,
                              int64_t nodes) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  if (!ew_physical(g, c) || bcnd[c] < 1) return;
  const int last = ew_last(g, bcnd, c);
  if (last < 0) return;
  EChain ch;
  ew_chain(g, bcnd, c, last, nodes, ch);
  const double4 l = lamE[c];
  const double lc[3] = {l.x, l.y, l.z};
  const double hh[3] = {g.h1, g.h2, g.h3};
  const int64_t s[3] = {1, g.sy, g.sz};
  for (int d = 0; d < 3; ++d) {
    if (d == ch.ax) continue;
    const double f = lc[d] / (2.0 * hh[d]);   // E_d = -(phi(c + s) - phi(c - s)) / (2 h)
    atomicAdd(&lamphi[c + s[d]], -f);
    atomicAdd(&lamphi[c - s[d]], f);
  }
  const double ln = lc[ch.ax];
  double sign = 1.0;
  for (int j = 0; j < ch.k; ++j) {
    atomicAdd(&lamE1[4 * ch.a[j] + ch.ax], 2.0 * sign * ln);
    sign = -sign;
  }
  if (ch.term_kind == 1) {
    atomicAdd(&lamE1[4 * ch.t + ch.ax], sign * ln);
  } else if (ch.term_kind == 2) {
    const double f = sign * ln / (2.0 * hh[ch.ax]);
    atomicAdd(&lamphi[ch.t + s[ch.ax]], -f);
    atomicAdd(&lamphi[ch.t - s[ch.ax]], f);
  }
}

// E-field transpose, part 2: the central differences of pass 1 at plasma
// nodes (bcnd <= 0, physical), gathered per phi node from its six
// neighbours; their dL/dE is the final one (pass 2 never writes them) plus
// what the wall extrapolation read from them.
__global__ void k_adj_e_interior(KGrid g, const int32_t *__restrict__ bcnd,
                                 const double4 *__restrict__ lamE,
                                 const double4 *__restrict__ lamE1, double *lamphi,
                                 int64_t nodes) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  int i, j, k;
  ijk_of(g, c, i, j, k);
  double acc = 0.0;
  const double hh[3] = {g.h1, g.h2, g.h3};
  const int64_t s[3] = {1, g.sy, g.sz};
  const int lo[3] = {i, j, k};
  const int hi[3] = {g.n1 + 2, g.n2 + 2, g.n3 + 2};
  for (int d = 0; d < 3; ++d) {
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      // neighbour m = c + sgn s: E_d(m) = -(phi(m + s) - phi(m - s)) / (2h)
      const int idx = lo[d] + sgn;
      if (idx < 0 || idx > hi[d]) continue;
      const int64_t m = c + sgn * s[d];
      int mi, mj, mk;
      ijk_of(g, m, mi, mj, mk);
      if (!physical_node(g, mi, mj, mk) || bcnd[m] > 0) continue;
      const double4 a = lamE[m], b = lamE1[m];
      const double lm = (d == 0 ? a.x + b.x : d == 1 ? a.y + b.y : a.z + b.z);
      // c = m - s (sgn = +1): dE/dphi(c) = +1/(2h); c = m + s: -1/(2h)
      acc += (sgn > 0 ? 1.0 : -1.0) * lm / (2.0 * hh[d]);
    }
  }
  lamphi[c] += acc;
}

// dL/dV of an electrode at voltage V on every node labelled `label`:
// sum over those nodes of dL/dphi there minus the coupling of their plasma
// neighbours, sum_p a_pe mu_p (phi_P = A_PP^-1 (b_P - A_PE phi_E)).  Also
// returns, in rhs, dL/dphi on plasma nodes and 0 elsewhere (the adjoint
// solve's right-hand side).
__global__ void k_adj_bias(KGrid g, const int32_t *__restrict__ bcnd, int label,
                           const double *__restrict__ lamphi, const double *__restrict__ mu,
                           double aw, double an, double at, int64_t nodes, double *out) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double v = 0.0;
  if (c < nodes && bcnd[c] == label) {
    v = lamphi[c];
    int i, j, k;
    ijk_of(g, c, i, j, k);
    const int64_t s[3] = {1, g.sy, g.sz};
    const double a[3] = {aw, an, at};
    const int lo[3] = {i, j, k};
    const int hi[3] = {g.n1 + 2, g.n2 + 2, g.n3 + 2};
    for (int d = 0; d < 3; ++d)
      for (int sgn = -1; sgn <= 1; sgn += 2) {
        const int idx = lo[d] + sgn;
        if (idx < 0 || idx > hi[d]) continue;
        const int64_t m = c + sgn * s[d];
        int mi, mj, mk;
        ijk_of(g, m, mi, mj, mk);
        if (physical_node(g, mi, mj, mk) && bcnd[m] <= 0) v -= a[d] * mu[m];
      }
  }
  __shared__ double red[BLOCK / LEPIC_WAVE];
  v = wave_sum_d(v);
  if (__lane_id() == 0) red[threadIdx.x / LEPIC_WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < BLOCK / LEPIC_WAVE; ++q) s += red[q];
    if (s != 0.0) atomicAdd(out, s);
  }
}

__global__ void k_adj_rhs(KGrid g, const int32_t *__restrict__ bcnd,
                          const double *__restrict__ lamphi, double scale, double *rhs,
                          int64_t nodes) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  int i, j, k;
  ijk_of(g, c, i, j, k);
  rhs[c] = (physical_node(g, i, j, k) && bcnd[c] <= 0) ? lamphi[c] * scale : 0.0;
}

// Adjoint of the deposit of one species: dL/dx_p += coef sum_c kq_c lamb_c
// dw_c/dx (coef = -q Nm / (h1 h2 h3): rho = -sum_s q_s np_s, np = kq k1 w).
__global__ __launch_bounds__(BLOCK) void k_adj_deposit(KGrid g, lepic_pset st, KAdj a, int64_t n,
                                                       const double *__restrict__ kq,
                                                       const double *__restrict__ lamb,
                                                       double coef) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  if (i >= n || st.dead[i] != 0) return;
  int ix, iy, iz;
  double px, py, pz;
  if (!deposit_fracs(g, st.x[i], st.y[i], st.z[i], ix, iy, iz, px, py, pz)) return;
  int64_t nd[8];
  node8(g, gidx(g, ix, iy, iz), nd);
  double e[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) e[c] = kq[nd[c]] * lamb[nd[c]];
  double gx, gy, gz;
  weight_grads(px, py, pz, e, gx, gy, gz);
  a.lx[i] -= coef * gx / g.h1;
  a.ly[i] -= coef * gy / g.h2;
  a.lz[i] -= coef * gz / g.h3;
}

// The loss term of one step: coef * sum over live particles of v_a * S(x_a),
// a the axis (0 x, 1 y, 2 z), S the cloud-in-cell hat of node plane ip
// (x_a = (ip - 1) h_a) over h_a; with a != NULL also dL/dx_a, dL/dv_a +=
// scale * its derivatives.
__global__ __launch_bounds__(BLOCK) void k_plane_flux(KGrid g, lepic_pset st, int64_t n, int axis,
                                                      int ip, double coef, double *acc, KAdj a,
                                                      int adj, double scale) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  const double *pos = axis == 0 ? st.x : axis == 1 ? st.y : st.z;
  const double *vel = axis == 0 ? st.vx : axis == 1 ? st.vy : st.vz;
  const double h = axis == 0 ? g.h1 : axis == 1 ? g.h2 : g.h3;
  double v = 0.0;
  if (i < n && st.dead[i] == 0) {
    const double u = pos[i] / h - (double)(ip - 1);
    const double au = fabs(u);
    if (au < 1.0) {
      const double hat = 1.0 - au;
      const double va = vel[i];
      v = coef * va * hat;
      if (adj) {
        double *lp = axis == 0 ? a.lx : axis == 1 ? a.ly : a.lz;
        double *lv = axis == 0 ? a.lvx : axis == 1 ? a.lvy : a.lvz;
        lv[i] += scale * coef * hat;
        lp[i] += scale * coef * va * (u > 0.0 ? -1.0 : 1.0) / h;
      }
    }
  }
  if (!acc) return;
  __shared__ double red[BLOCK / LEPIC_WAVE];
  v = wave_sum_d(v);
  if (__lane_id() == 0) red[threadIdx.x / LEPIC_WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < BLOCK / LEPIC_WAVE; ++q) s += red[q];
    if (s != 0.0) atomicAdd(acc, s);
  }
}

// Frozen wall events: the slots whose recorded loss step is `step` become
// holes (dead = 2), whatever their position.
__global__ void k_apply_deaths(int8_t *dead, const int64_t *__restrict__ deaths, int64_t n,
                               int64_t step) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n && deaths[i] == step) dead[i] = 2;
}

int check_adj(lepic_ctx *ctx, const lepic_pset *p, const lepic_adjoint_set *a, const char *who) {
  if (!p || !a || p->n < 0 || a->n < p->n)
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": adjoint set shorter than the state");
  if (p->n > 0 && (!p->x || !p->dead || !a->lx || !a->ly || !a->lz || !a->lvx || !a->lvy ||
                   !a->lvz))
    return lepic_fail(ctx, LEPIC_ERR_ARG, std::string(who) + ": null array");
  return LEPIC_OK;
}

}  // namespace

extern "C" {

int lepic_adj_push(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                   lepic_adjoint_set *adj, const double *E4, const double *Bunif, double q,
                   double m, double dt, double *lamE4, double *lamB) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_adj(ctx, state, adj, "adj_push"))) return st;
  if (!E4 || !lamE4 || (Bunif && !lamB))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "adj_push: null argument");
  const int64_t n = state->n;
  if (n == 0) return LEPIC_OK;
  const int64_t nodes = grid_nodes(g);
  const double k1 = dt * q / (2.0 * m);
  KBAdj bf;
  bf.use_b = Bunif ? 1 : 0;
  bf.B[0] = Bunif ? Bunif[0] : 0.0;
  bf.B[1] = Bunif ? Bunif[1] : 0.0;
  bf.B[2] = Bunif ? Bunif[2] : 0.0;
  const double B2 = bf.B[0] * bf.B[0] + bf.B[1] * bf.B[1] + bf.B[2] * bf.B[2];
  const double Bt = std::sqrt(B2);
  bf.k2 = k1 * (2.0 / (1.0 + (k1 * Bt) * (k1 * Bt)));
  const double den = 1.0 + k1 * k1 * B2;
  bf.dk2 = -2.0 * k1 * k1 * k1 / (den * den);
  bf.B4 = nullptr;
  bf.magnetized = 1;
  double *rep = lepic_replicas(ctx, 4 * nodes);
  if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "adj_push: replicas");
  hipLaunchKernelGGL(k_adj_push<false>, dim3(grid_blocks(n, BLOCK * APT)), dim3(BLOCK), 0,
                     ctx->stream, make_kgrid(g), *state, kadj(adj), n,
                     reinterpret_cast<const double4 *>(E4), bf, k1, dt, rep, nodes, lamB);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_push");
  hipLaunchKernelGGL(k_adj_rep_flush, dim3(grid_blocks(4 * nodes, 256)), dim3(256), 0,
                     ctx->stream, lamE4, rep, 4 * nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_rep_flush");
  return LEPIC_OK;
}

int lepic_adj_push_map(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                       lepic_adjoint_set *adj, const double *E4, const double *B4,
                       int32_t magnetized, double q, double m, double dt, double *lamE4) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_adj(ctx, state, adj, "adj_push_map"))) return st;
  if (!E4 || !B4 || !lamE4) return lepic_fail(ctx, LEPIC_ERR_ARG, "adj_push_map: null argument");
  const int64_t n = state->n;
  if (n == 0) return LEPIC_OK;
  const int64_t nodes = grid_nodes(g);
  const double k1 = dt * q / (2.0 * m);
  KBAdj bf;
  bf.use_b = 0;
  bf.B[0] = bf.B[1] = bf.B[2] = 0.0;
  bf.k2 = 0.0;
  bf.dk2 = 0.0;
  bf.B4 = reinterpret_cast<const double4 *>(B4);
  bf.magnetized = magnetized ? 1 : 0;
  double *rep = lepic_replicas(ctx, 4 * nodes);
  if (!rep) return lepic_fail(ctx, LEPIC_ERR_HIP, "adj_push_map: replicas");
  hipLaunchKernelGGL(k_adj_push<true>, dim3(grid_blocks(n, BLOCK * APT)), dim3(BLOCK), 0,
                     ctx->stream, make_kgrid(g), *state, kadj(adj), n,
                     reinterpret_cast<const double4 *>(E4), bf, k1, dt, rep, nodes, nullptr);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_push_map");
  hipLaunchKernelGGL(k_adj_rep_flush, dim3(grid_blocks(4 * nodes, 256)), dim3(256), 0,
                     ctx->stream, lamE4, rep, 4 * nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_rep_flush");
  return LEPIC_OK;
}

int lepic_adj_efield(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                     const double *lamE4, double *lamphi) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g))) return st;
  if (!bcnd || !lamE4 || !lamphi) return lepic_fail(ctx, LEPIC_ERR_ARG, "adj_efield: null argument");
  const int64_t nodes = grid_nodes(g);
  double *lamE1 = reinterpret_cast<double *>(lepic_scratch(ctx, sizeof(double) * 4 * nodes, 1));
  if (!lamE1) return lepic_fail(ctx, LEPIC_ERR_HIP, "adj_efield: scratch");
  LEPIC_HIP(ctx, hipMemsetAsync(lamE1, 0, sizeof(double) * 4 * nodes, ctx->stream));
  LEPIC_HIP(ctx, hipMemsetAsync(lamphi, 0, sizeof(double) * nodes, ctx->stream));
  const KGrid kg = make_kgrid(g);
  hipLaunchKernelGGL(k_adj_e_walls, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream, kg,
                     bcnd, reinterpret_cast<const double4 *>(lamE4), lamE1, lamphi, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_e_walls");
  hipLaunchKernelGGL(k_adj_e_interior, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     kg, bcnd, reinterpret_cast<const double4 *>(lamE4),
                     reinterpret_cast<const double4 *>(lamE1), lamphi, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_e_interior");
  return LEPIC_OK;
}

int lepic_adj_poisson_rhs(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                          const double *lamphi, double scale, double *rhs) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g))) return st;
  if (!bcnd || !lamphi || !rhs) return lepic_fail(ctx, LEPIC_ERR_ARG, "adj_poisson_rhs: null argument");
  const int64_t nodes = grid_nodes(g);
  hipLaunchKernelGGL(k_adj_rhs, dim3(grid_blocks(nodes, 256)), dim3(256), 0, ctx->stream,
                     make_kgrid(g), bcnd, lamphi, scale, rhs, nodes);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_rhs");
  return LEPIC_OK;
}

int lepic_adj_bias(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd, int32_t label,
                   const double *lamphi, const double *mu, double eps0, double *out) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g))) return st;
  if (!bcnd || !lamphi || !mu || !out || label < 1)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "adj_bias: bad argument");
  const int64_t nodes = grid_nodes(g);
  const double aw = eps0 / (g->h[0] * g->h[0]), an = eps0 / (g->h[1] * g->h[1]),
               at = eps0 / (g->h[2] * g->h[2]);
  hipLaunchKernelGGL(k_adj_bias, dim3(grid_blocks(nodes, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), bcnd, label, lamphi, mu, aw, an, at, nodes, out);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_bias");
  return LEPIC_OK;
}

int lepic_adj_deposit(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                      lepic_adjoint_set *adj, const double *kq, const double *lamb, double q,
                      double Nm) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g)) || (st = check_adj(ctx, state, adj, "adj_deposit"))) return st;
  if (!kq || !lamb) return lepic_fail(ctx, LEPIC_ERR_ARG, "adj_deposit: null argument");
  if (state->n == 0) return LEPIC_OK;
  const double coef = -q * (Nm / (g->h[0] * g->h[1] * g->h[2]));
  hipLaunchKernelGGL(k_adj_deposit, dim3(grid_blocks(state->n, BLOCK)), dim3(BLOCK), 0,
                     ctx->stream, make_kgrid(g), *state, kadj(adj), state->n, kq, lamb, coef);
  LEPIC_LAUNCH_CHECK(ctx, "k_adj_deposit");
  return LEPIC_OK;
}

int lepic_plane_flux_axis(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                          int32_t axis, int32_t plane, double coef, double *acc,
                          lepic_adjoint_set *adj, double scale) {
  if (!ctx) return LEPIC_ERR_ARG;
  int st;
  if ((st = check_grid(ctx, g))) return st;
  if (axis < 0 || axis > 2) return lepic_fail(ctx, LEPIC_ERR_ARG, "plane_flux: axis");
  if (!state || state->n < 0 ||
      (state->n > 0 && (!state->x || !state->y || !state->z || !state->vx || !state->vy ||
                        !state->vz || !state->dead)))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "plane_flux: bad particle set");
  if (plane < 1 || plane > g->n[axis] + 1)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "plane_flux: plane");
  if (adj && (st = check_adj(ctx, state, adj, "plane_flux"))) return st;
  if (!acc && !adj) return LEPIC_OK;
  if (state->n == 0) return LEPIC_OK;
  KAdj a = adj ? kadj(adj) : KAdj{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hipLaunchKernelGGL(k_plane_flux, dim3(grid_blocks(state->n, BLOCK)), dim3(BLOCK), 0,
                     ctx->stream, make_kgrid(g), *state, state->n, (int)axis, plane, coef, acc, a,
                     adj ? 1 : 0, scale);
  LEPIC_LAUNCH_CHECK(ctx, "k_plane_flux");
  return LEPIC_OK;
}

int lepic_plane_flux(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state, int32_t plane,
                     double coef, double *acc, lepic_adjoint_set *adj, double scale) {
  return lepic_plane_flux_axis(ctx, g, state, 0, plane, coef, acc, adj, scale);
}

int lepic_apply_deaths(lepic_ctx *ctx, lepic_pset *p, const int64_t *deaths, int64_t step) {
  if (!ctx) return LEPIC_ERR_ARG;
  if (!p || p->n < 0 || (p->n > 0 && (!p->dead || !deaths)))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "apply_deaths: bad argument");
  if (p->n == 0) return LEPIC_OK;
  hipLaunchKernelGGL(k_apply_deaths, dim3(grid_blocks(p->n, 256)), dim3(256), 0, ctx->stream,
                     p->dead, deaths, p->n, step);
  LEPIC_LAUNCH_CHECK(ctx, "k_apply_deaths");
  return LEPIC_OK;
}

}  // extern "C"
