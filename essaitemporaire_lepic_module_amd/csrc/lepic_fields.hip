// lepic_fields.hip -- node-field kernels: kq, fills, density reduction with
// periodic stitching (a7), rho (a7) and the electric field (a9).
// One thread per node (x fastest), so every pass is a coalesced stream.
#include <mutex>
#include <string>

#include "lepic_internal.h"
#include "lepic_ewall.h"

namespace {

constexpr int BLOCK = 256;

__global__ void k_fill(double *a, int64_t n, double v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// build_kq (mod_charge_weights.f90:9-33)
__global__ void k_kq(const int32_t *__restrict__ bcnd, double *kq, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) kq[i] = (bcnd[i] != -1) ? 2.0 : 1.0;
}

// apply_periodic_density_bc, y stitching (mod_density.f90:192-213):
// one thread per (ix, iz) column, ix in 1..n1+1, iz in 1..n3+1.
__global__ void k_stitch_y(KGrid g, const int32_t *__restrict__ bcnd, double *np) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t nx = g.n1 + 1, nzz = g.n3 + 1;
  if (t >= nx * nzz) return;
  int ix = 1 + (int)(t % nx), iz = 1 + (int)(t / nx);
  const int n2 = g.n2;
  if (bcnd[gidx(g, ix, 1, iz)] == 0) {
    np[gidx(g, ix, 1, iz)] = 0.5 * (np[gidx(g, ix, 1, iz)] + np[gidx(g, ix, n2 + 1, iz)]);
    np[gidx(g, ix, 0, iz)] = np[gidx(g, ix, n2, iz)];
  }
  if (bcnd[gidx(g, ix, n2 + 1, iz)] == 0) {
    np[gidx(g, ix, n2 + 2, iz)] = np[gidx(g, ix, 2, iz)];
    np[gidx(g, ix, n2 + 1, iz)] = np[gidx(g, ix, 1, iz)];
  }
}

// z stitching (mod_density.f90:219-240): one thread per (ix, iy).
__global__ void k_stitch_z(KGrid g, const int32_t *__restrict__ bcnd, double *np) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t nx = g.n1 + 1, ny = g.n2 + 1;
  if (t >= nx * ny) return;
  int ix = 1 + (int)(t % nx), iy = 1 + (int)(t / nx);
  const int n3 = g.n3;
  if (bcnd[gidx(g, ix, iy, 1)] == 0) {
    np[gidx(g, ix, iy, 1)] = 0.5 * (np[gidx(g, ix, iy, 1)] + np[gidx(g, ix, iy, n3 + 1)]);
    np[gidx(g, ix, iy, 0)] = np[gidx(g, ix, iy, n3)];
  }
  if (bcnd[gidx(g, ix, iy, n3 + 1)] == 0) {
    np[gidx(g, ix, iy, n3 + 2)] = np[gidx(g, ix, iy, 2)];
    np[gidx(g, ix, iy, n3 + 1)] = np[gidx(g, ix, iy, 1)];
  }
}

__device__ inline void node_ijk(const KGrid &g, int64_t t, int &i, int &j, int &k) {
  i = (int)(t % g.sy);
  int64_t r = t / g.sy;
  j = (int)(r % (g.n2 + 3));
  k = (int)(r / (g.n2 + 3));
}

__device__ inline bool physical(const KGrid &g, int i, int j, int k) {
  return i >= 1 && i <= g.n1 + 1 && j >= 1 && j <= g.n2 + 1 && k >= 1 && k <= g.n3 + 1;
}

// np_red = sum over the one virtual thread on nodes 1..n+1, 0 elsewhere
// (mod_density.f90:50-66).
__global__ void k_reduce(KGrid g, const double *__restrict__ np_thread, double *np_red,
                         int64_t nodes, int ntype) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nodes * ntype) return;
  int64_t c = t % nodes;
  int i, j, k;
  node_ijk(g, c, i, j, k);
  np_red[t] = physical(g, i, j, k) ? (0.0 + np_thread[t]) : 0.0;
}

struct Charges {
  double q[LEPIC_MAX_SPECIES];
};

// build_rho_from_np (mod_density.f90:91-104): rho = -sum_s q_s np_s.
__global__ void k_rho(KGrid g, const double *__restrict__ np_red, Charges ch, int ntype,
                      int64_t nodes, double *rho) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  int i, j, k;
  node_ijk(g, c, i, j, k);
  if (!physical(g, i, j, k)) {
    rho[c] = 0.0;
    return;
  }
  double r = 0.0;
  for (int s = 0; s < ntype; ++s) r = r - ch.q[s] * np_red[(int64_t)s * nodes + c];
  rho[c] = r;
}

// calc_Efield_modular pass 1 (mod_electricField.f90:18-42): E = 0, then
// central differences at nodes with bcnd <= 0.
__global__ void k_e_interior(KGrid g, const double *__restrict__ phi,
                             const int32_t *__restrict__ bcnd, double4 *E, int64_t nodes) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  int i, j, k;
  node_ijk(g, c, i, j, k);
  double4 e = make_double4(0.0, 0.0, 0.0, 0.0);
  if (physical(g, i, j, k) && bcnd[c] <= 0) {
    e.x = -(phi[c + 1] - phi[c - 1]) / (2.0 * g.h1);
    e.y = -(phi[c + g.sy] - phi[c - g.sy]) / (2.0 * g.h2);
    e.z = -(phi[c + g.sz] - phi[c - g.sz]) / (2.0 * g.h3);
  }
  E[c] = e;
}

// Pass 2a: Neumann and periodic nodes (mod_electricField.f90:58-105).  The
// reference runs one loop in (iz, iy, ix) order and its y and z ghost copies
// read each other's results; the phases below reproduce that order:
//   phase 0  Neumann LHS nodes (x = 0 plane): Ex = 0, interior ones done;
//   phase 1  y-branch of the iy = 1 nodes on the iz = 1 plane;
//   phase 2  z-branch of the iz = 1 nodes;
//   phase 3  y-branch of the iy = 1 nodes with iz >= 2.
__device__ inline void e_ybranch(const KGrid &g, const double *phi, double4 *E, int ix, int iz) {
  const int n2 = g.n2;
  auto P = [&](int i, int j, int k) { return phi[gidx(g, i, j, k)]; };
  auto A = [&](int j) -> double4 & { return E[gidx(g, ix, j, iz)]; };
  A(1).y = -(P(ix, 2, iz) - P(ix, n2, iz)) / (2.0 * g.h2);
  A(0).y = A(n2).y;
  A(n2 + 1).y = A(1).y;
  A(n2 + 2).y = A(2).y;
  A(0).x = A(n2).x;
  A(n2 + 2).x = A(2).x;
  A(0).z = A(n2).z;
  A(n2 + 2).z = A(2).z;
}

__device__ inline void e_zbranch(const KGrid &g, const double *phi, double4 *E, int ix, int iy) {
  const int n3 = g.n3;
  auto P = [&](int i, int j, int k) { return phi[gidx(g, i, j, k)]; };
  auto A = [&](int k) -> double4 & { return E[gidx(g, ix, iy, k)]; };
  A(1).z = -(P(ix, iy, 2) - P(ix, iy, n3)) / (2.0 * g.h3);
  A(0).z = A(n3).z;
  A(n3 + 1).z = A(1).z;
  A(n3 + 2).z = A(2).z;
  A(0).x = A(n3).x;
  A(n3 + 2).x = A(2).x;
  A(0).y = A(n3).y;
  A(n3 + 2).y = A(2).y;
}

__device__ inline bool neumann_interior(const KGrid &g, int iy, int iz) {
  return iy > 1 && iy < g.n2 + 1 && iz > 1 && iz < g.n3 + 1;
}

__global__ void k_e_periodic(KGrid g, const double *__restrict__ phi,
                             const int32_t *__restrict__ bcnd, double4 *E, int phase) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (phase == 0) {  // Neumann nodes: threads over (ix, iy, iz) of the physical range
    int64_t nx = g.n1 + 1, ny = g.n2 + 1, nz = g.n3 + 1;
    if (t >= nx * ny * nz) return;
    int ix = 1 + (int)(t % nx);
    int64_t r = t / nx;
    int iy = 1 + (int)(r % ny), iz = 1 + (int)(r / ny);
    if (bcnd[gidx(g, ix, iy, iz)] != -2) return;
    E[gidx(g, 1, iy, iz)].x = 0.0;
    if (neumann_interior(g, iy, iz)) {
      E[gidx(g, 1, iy, iz)].y = -(phi[gidx(g, 1, iy + 1, iz)] - phi[gidx(g, 1, iy - 1, iz)]) / (2.0 * g.h2);
      E[gidx(g, 1, iy, iz)].z = -(phi[gidx(g, 1, iy, iz + 1)] - phi[gidx(g, 1, iy, iz - 1)]) / (2.0 * g.h3);
    }
    return;
  }
  if (phase == 1 || phase == 3) {  // iy = 1 plane, threads over (ix, iz)
    int64_t nx = g.n1 + 1, nz = g.n3 + 1;
    if (t >= nx * nz) return;
    int ix = 1 + (int)(t % nx), iz = 1 + (int)(t / nx);
    if ((phase == 1) != (iz == 1)) return;
    int b = bcnd[gidx(g, ix, 1, iz)];
    if (b != 0 && b != -2) return;
    if (b == -2 && neumann_interior(g, 1, iz)) return;
    e_ybranch(g, phi, E, ix, iz);
    return;
  }
  // phase 2: iz = 1 plane, threads over (ix, iy)
  int64_t nx = g.n1 + 1, ny = g.n2 + 1;
  if (t >= nx * ny) return;
  int ix = 1 + (int)(t % nx), iy = 1 + (int)(t / nx);
  int b = bcnd[gidx(g, ix, iy, 1)];
  if (b != 0 && b != -2) return;
  e_zbranch(g, phi, E, ix, iy);
}

// Pass 2b: electrode nodes next to the plasma (mod_electricField.f90:110-152):
// the last applicable block's normal component (lepic_ewall.h), computed
// from the pass-1/2a field into `outv` (one double per node, written only at
// the nodes the pass sets) so that no thread reads a node another one
// writes; k_e_walls_store then sets those nodes' E from it and the central
// differences of phi.  A chain longer than E_CHAIN_MAX links (electrodes one
// plasma node apart, 32 in a row) gives NaN.
__global__ void k_e_walls(KGrid g, const double *__restrict__ phi,
                          const int32_t *__restrict__ bcnd, const double4 *__restrict__ E,
                          double *outv, int64_t nodes) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  if (!ew_physical(g, c) || bcnd[c] < 1) return;
  const int last = ew_last(g, bcnd, c);
  if (last < 0) return;
  EChain ch;
  ew_chain(g, bcnd, c, last, nodes, ch);
  double v;
  if (ch.term_kind == 0) v = 0.0;
  else if (ch.term_kind == 1) v = ew_comp(E[ch.t], ch.ax);
  else v = ew_central(g, phi, ch.t, ch.ax);
  for (int j = ch.k - 1; j >= 0; --j) v = 2.0 * ew_comp(E[ch.a[j]], ch.ax) - v;
  if (ch.overflow) v = __builtin_nan("");
  outv[c] = v;
}

__global__ void k_e_walls_store(KGrid g, const double *__restrict__ phi,
                                const int32_t *__restrict__ bcnd,
                                const double *__restrict__ outv, double4 *E, int64_t nodes) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nodes) return;
  if (!ew_physical(g, c) || bcnd[c] < 1) return;
  const int last = ew_last(g, bcnd, c);
  if (last < 0) return;
  const int ax = last / 2;
  double e[3];
  for (int d = 0; d < 3; ++d) e[d] = d == ax ? outv[c] : ew_central(g, phi, c, d);
  E[c] = make_double4(e[0], e[1], e[2], E[c].w);
}

// apply_dielectric_bc_to_phi (mod_state.f90:378-430): on every node of a
// label whose type is a dielectric, phi takes each species' collected
// surface charge, in species order: the x-z sum of its column over
// 2 Ca (type 2), the y-z sums of the first or second plane over Ca (3, 4).
__global__ void k_die_phi(KGrid g, const int32_t *__restrict__ bcnd,
                          const int32_t *__restrict__ dtype, int ngrid, int ntype, int64_t sxz_n,
                          int64_t syz_n, const double *__restrict__ sxz,
                          const double *__restrict__ syz, double Ca, int64_t nodes, double *phi) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  const int igrid = bcnd[t];
  if (igrid <= 0 || igrid > ngrid) return;   // labels above the table are not walls of it
  const int dt = dtype[igrid];
  if (dt < 2 || dt > 4) return;
  int i, j, k;
  node_ijk(g, t, i, j, k);
  double v = phi[t];
  for (int s = 0; s < ntype; ++s) {
    // qsum = 0 + the one thread's sum (mod_state.f90:403-406), then / Ca
    if (dt == 2) {
      v = v + (0.0 + sxz[s * sxz_n + i + (int64_t)(g.n1 + 3) * k]) / (2.0 * Ca);
    } else {
      const int c = dt - 3;
      v = v + (0.0 + syz[s * syz_n + c + 2 * (j + (int64_t)(g.n2 + 3) * k)]) / Ca;
    }
  }
  phi[t] = v;
}

}  // namespace

// Scratch of the wall pass for lepic_efield_noalloc: one double per node,
// in a file-scope device array that every GPU the library runs on holds
// until the process ends (LEPIC_EFIELD_NOALLOC_MAX_NODES doubles, 18 MiB).
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address); it uses the address hipGetSymbolAddress
// gives for the device, resolved once per device when a context is created.
static __device__ __align__(256) double scratch[LEPIC_EFIELD_NOALLOC_MAX_NODES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};
bool g_scratch_tried[MAX_DEVICES] = {};

double *scratch_ptr(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// E = -grad(phi) (calc_Efield_modular, mod_electricField.f90:9-164): the
// interior and periodic passes, then the electrode-wall pass through outv.
int efield_launch(lepic_ctx *ctx, const lepic_grid *g, const double *phi, const int32_t *bcnd,
                  double *E4, double *outv) {
  KGrid kg = make_kgrid(g);
  const int64_t nodes = grid_nodes(g);
  double4 *E = reinterpret_cast<double4 *>(E4);
  dim3 gr(grid_blocks(nodes, BLOCK));
  hipLaunchKernelGGL(k_e_interior, gr, dim3(BLOCK), 0, ctx->stream, kg, phi, bcnd, E, nodes);
  const int64_t nphys = (int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1);
  const int64_t nxz = (int64_t)(g->n[0] + 1) * (g->n[2] + 1);
  const int64_t nxy = (int64_t)(g->n[0] + 1) * (g->n[1] + 1);
  hipLaunchKernelGGL(k_e_periodic, dim3(grid_blocks(nphys, BLOCK)), dim3(BLOCK), 0, ctx->stream, kg,
                     phi, bcnd, E, 0);
  hipLaunchKernelGGL(k_e_periodic, dim3(grid_blocks(nxz, BLOCK)), dim3(BLOCK), 0, ctx->stream, kg,
                     phi, bcnd, E, 1);
  hipLaunchKernelGGL(k_e_periodic, dim3(grid_blocks(nxy, BLOCK)), dim3(BLOCK), 0, ctx->stream, kg,
                     phi, bcnd, E, 2);
  hipLaunchKernelGGL(k_e_periodic, dim3(grid_blocks(nxz, BLOCK)), dim3(BLOCK), 0, ctx->stream, kg,
                     phi, bcnd, E, 3);
  hipLaunchKernelGGL(k_e_walls, gr, dim3(BLOCK), 0, ctx->stream, kg, phi, bcnd, E, outv, nodes);
  hipLaunchKernelGGL(k_e_walls_store, gr, dim3(BLOCK), 0, ctx->stream, kg, phi, bcnd, outv, E,
                     nodes);
  LEPIC_LAUNCH_CHECK(ctx, "efield");
  return LEPIC_OK;
}

}  // namespace

// Called by lepic_create with the context's device current.
void lepic_efield_scratch_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_tried[device]) return;
  g_scratch_tried[device] = true;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess && ptr)
    g_scratch_ptr[device] = reinterpret_cast<double *>(ptr);
  else
    (void)hipGetLastError();
}

extern "C" {

int lepic_fill(lepic_ctx *ctx, double *a, int64_t count, double value) {
  if (!ctx || (!a && count > 0) || count < 0) return LEPIC_ERR_ARG;
  if (count == 0) return LEPIC_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_blocks(count, BLOCK)), dim3(BLOCK), 0, ctx->stream, a,
                     count, value);
  LEPIC_LAUNCH_CHECK(ctx, "k_fill");
  return LEPIC_OK;
}

int lepic_build_kq(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd, double *kq) {
  if (!ctx || !g || !bcnd || !kq) return LEPIC_ERR_ARG;
  int64_t n = grid_nodes(g);
  hipLaunchKernelGGL(k_kq, dim3(grid_blocks(n, BLOCK)), dim3(BLOCK), 0, ctx->stream, bcnd, kq, n);
  LEPIC_LAUNCH_CHECK(ctx, "k_kq");
  return LEPIC_OK;
}

int lepic_reduce_density(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                         int32_t ntype, double *np_thread, double *np_red) {
  if (!ctx || !g || !bcnd || !np_thread || !np_red || ntype < 1) return LEPIC_ERR_ARG;
  if (np_thread == np_red) return lepic_fail(ctx, LEPIC_ERR_ARG, "reduce_density: aliasing");
  KGrid kg = make_kgrid(g);
  const int64_t nodes = grid_nodes(g);
  for (int s = 0; s < ntype; ++s) {
    double *np = np_thread + (int64_t)s * nodes;
    hipLaunchKernelGGL(k_stitch_y, dim3(grid_blocks((int64_t)(g->n[0] + 1) * (g->n[2] + 1), BLOCK)),
                       dim3(BLOCK), 0, ctx->stream, kg, bcnd, np);
    hipLaunchKernelGGL(k_stitch_z, dim3(grid_blocks((int64_t)(g->n[0] + 1) * (g->n[1] + 1), BLOCK)),
                       dim3(BLOCK), 0, ctx->stream, kg, bcnd, np);
  }
  hipLaunchKernelGGL(k_reduce, dim3(grid_blocks(nodes * ntype, BLOCK)), dim3(BLOCK), 0,
                     ctx->stream, kg, np_thread, np_red, nodes, ntype);
  LEPIC_LAUNCH_CHECK(ctx, "reduce_density");
  return LEPIC_OK;
}

int lepic_build_rho(lepic_ctx *ctx, const lepic_grid *g, int32_t ntype, const double *np_red,
                    const double *charge, double *rho) {
  if (!ctx || !g || !np_red || !charge || !rho || ntype < 1 || ntype > LEPIC_MAX_SPECIES)
    return LEPIC_ERR_ARG;
  Charges ch;
  for (int s = 0; s < LEPIC_MAX_SPECIES; ++s) ch.q[s] = s < ntype ? charge[s] : 0.0;
  const int64_t nodes = grid_nodes(g);
  hipLaunchKernelGGL(k_rho, dim3(grid_blocks(nodes, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), np_red, ch, ntype, nodes, rho);
  LEPIC_LAUNCH_CHECK(ctx, "k_rho");
  return LEPIC_OK;
}

int lepic_dielectric_phi(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                         const int32_t *dtype, int32_t ngrid, int32_t ntype,
                         const double *sum_q_xz, const double *sum_q_yz, double Ca_cell,
                         double *phi) {
  if (!ctx || !g || !bcnd || !dtype || !sum_q_xz || !sum_q_yz || !phi || ntype < 1 || ngrid < 0)
    return LEPIC_ERR_ARG;
  const int64_t nodes = grid_nodes(g);
  const int64_t sxz_n = (int64_t)(g->n[0] + 3) * (g->n[2] + 3);
  const int64_t syz_n = 2 * (int64_t)(g->n[1] + 3) * (g->n[2] + 3);
  hipLaunchKernelGGL(k_die_phi, dim3(grid_blocks(nodes, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     make_kgrid(g), bcnd, dtype, (int)ngrid, (int)ntype, sxz_n, syz_n, sum_q_xz, sum_q_yz,
                     Ca_cell, nodes, phi);
  LEPIC_LAUNCH_CHECK(ctx, "k_die_phi");
  return LEPIC_OK;
}

int lepic_efield(lepic_ctx *ctx, const lepic_grid *g, const double *phi, const int32_t *bcnd,
                 double *E4) {
  if (!ctx || !g || !phi || !bcnd || !E4) return LEPIC_ERR_ARG;
  if (g->n[0] < 2 || g->n[1] < 2 || g->n[2] < 2)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "efield: need n >= 2 in every dimension");
  const int64_t nodes = grid_nodes(g);
  double *outv = reinterpret_cast<double *>(lepic_scratch(ctx, sizeof(double) * (size_t)nodes, 1));
  if (!outv) return lepic_fail(ctx, LEPIC_ERR_HIP, "efield: scratch");
  return efield_launch(ctx, g, phi, bcnd, E4, outv);
}

int lepic_efield_noalloc(lepic_ctx *ctx, const lepic_grid *g, const double *phi,
                         const int32_t *bcnd, double *E4) {
  if (!ctx || !g || !phi || !bcnd || !E4) return LEPIC_ERR_ARG;
  if (g->n[0] < 2 || g->n[1] < 2 || g->n[2] < 2)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "efield_noalloc: need n >= 2 in every dimension");
  const int64_t nodes = grid_nodes(g);
  if (nodes > LEPIC_EFIELD_NOALLOC_MAX_NODES)
    return lepic_fail(ctx, LEPIC_ERR_ARG,
                      "efield_noalloc: " + std::to_string(nodes) + " nodes exceed the static "
                      "scratch of " + std::to_string(LEPIC_EFIELD_NOALLOC_MAX_NODES) + " nodes");
  double *outv = scratch_ptr(ctx->device);
  if (!outv)
    return lepic_fail(ctx, LEPIC_ERR_HIP, "efield_noalloc: the static scratch of device " +
                                              std::to_string(ctx->device) +
                                              " was not resolved when the context was created");
  return efield_launch(ctx, g, phi, bcnd, E4, outv);
}

}  // extern "C"
