// lepic_ewall.h -- E at electrode nodes next to the plasma
// (calc_Efield_modular's wall pass, mod_electricField.f90:110-152), shared by
// the field kernel and its adjoint.
//
// The reference visits electrode nodes in its (iz, iy, ix) loop order and,
// at each, applies the W, E, S, N, B, T blocks whose neighbour is plasma;
// every block sets all three components, so the last applicable block alone
// defines the node's E: its normal component extrapolated from the plasma
// side, 2 E(c + o) - E(c + 2 o), its tangential components central
// differences of phi.  E(c + o) is a plasma node's (untouched by the pass);
// E(c + 2 o) is the value the loop sees there: final when c + 2 o is an
// electrode node the loop visited before c (o pointing to lower indices),
// its value before the pass otherwise.  Electrodes one plasma node apart
// along an axis thus form chains, followed here to their end and evaluated
// from there, as the loop does -- with no read of a node another thread
// writes, so the GPU pass is race-free and bit-exact for any geometry.
#pragma once

#include "lepic_internal.h"

constexpr int E_CHAIN_MAX = 32;

__device__ inline void ew_ijk(const KGrid &g, int64_t t, int &i, int &j, int &k) {
  i = (int)(t % g.sy);
  const int64_t r = t / g.sy;
  j = (int)(r % (g.n2 + 3));
  k = (int)(r / (g.n2 + 3));
}

__device__ inline bool ew_physical(const KGrid &g, int64_t t) {
  int i, j, k;
  ew_ijk(g, t, i, j, k);
  return i >= 1 && i <= g.n1 + 1 && j >= 1 && j <= g.n2 + 1 && k >= 1 && k <= g.n3 + 1;
}

__device__ inline int64_t ew_off(const KGrid &g, int q) {
  switch (q) {
    case 0: return 1;
    case 1: return -1;
    case 2: return g.sy;
    case 3: return -g.sy;
    case 4: return g.sz;
    default: return -g.sz;
  }
}

// Last applicable block (0..5 = W, E, S, N, B, T: neighbour at +x, -x, +y,
// -y, +z, -z is plasma) of a physical electrode node, -1 if none.
__device__ inline int ew_last(const KGrid &g, const int32_t *__restrict__ bcnd, int64_t c) {
  int last = -1;
  for (int q = 0; q < 6; ++q)
    if (bcnd[c + ew_off(g, q)] <= 0) last = q;
  return last;
}

__device__ inline double ew_comp(const double4 &e, int ax) {
  return ax == 0 ? e.x : (ax == 1 ? e.y : e.z);
}

// The normal component's chain from electrode node c with last block
// `last`: link j reads the pre-pass E of node a[j]; the chain ends on
// term_kind 0 (outside the array: 0), 1 (the pre-pass E of node t) or 2 (the
// central difference of phi at node t along the axis, an electrode whose
// last block has another axis).  value = 2 a0 - (2 a1 - (... - term)).
struct EChain {
  int k;
  int ax;
  int term_kind;
  int overflow;
  int64_t t;
  int64_t a[E_CHAIN_MAX];
};

__device__ inline void ew_chain(const KGrid &g, const int32_t *__restrict__ bcnd, int64_t c,
                                int last, int64_t nodes, EChain &ch) {
  ch.k = 0;
  ch.ax = last / 2;
  ch.overflow = 0;
  int64_t cur = c;
  int lq = last;
  for (;;) {
    const int64_t o = ew_off(g, lq);
    ch.a[ch.k++] = cur + o;
    const int64_t w = cur + 2 * o;
    if (w < 0 || w >= nodes) {
      ch.term_kind = 0;
      return;
    }
    if (o < 0 && ew_physical(g, w) && bcnd[w] >= 1) {
      const int lw = ew_last(g, bcnd, w);
      if (lw < 0) {
        ch.term_kind = 1;
        ch.t = w;
        return;
      }
      if (lw / 2 != ch.ax) {
        ch.term_kind = 2;
        ch.t = w;
        return;
      }
      if (ch.k == E_CHAIN_MAX) {
        ch.overflow = 1;
        ch.term_kind = 0;
        return;
      }
      cur = w;
      lq = lw;
      continue;
    }
    ch.term_kind = 1;
    ch.t = w;
    return;
  }
}

__device__ inline double ew_central(const KGrid &g, const double *__restrict__ phi, int64_t c,
                                    int ax) {
  const int64_t s = ax == 0 ? 1 : (ax == 1 ? g.sy : g.sz);
  const double h = ax == 0 ? g.h1 : (ax == 1 ? g.h2 : g.h3);
  return -(phi[c + s] - phi[c - s]) / (2.0 * h);
}
