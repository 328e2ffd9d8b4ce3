// lepic_diag.hip -- plane moments of a particle species
// (compute_particle_plane_moments_species, modules/particles/mod_planeMoments.f90:13-308).
//
// One pass over the particles: each live particle, taken half a step back
// along its velocity, adds bilinear weights to the xy, xz and yz planes it
// lies within one cell of.  Only particles in those thin slabs contribute, so
// the pass is a read of the particle arrays (49 bytes per particle) with
// device-scope FP64 atomics for the few in a slab.  The sums go to a scratch
// block (count, sum w v, sum w v^2 per plane node); a second kernel per plane
// applies the reference's closing update to data (5 doubles per node).
#include "lepic_internal.h"

namespace {

constexpr int BLOCK = 256;
constexpr int NACC = 7;   // count, sum w v (3), sum w v^2 (3)

struct Plane {
  double *acc;       // NACC arrays of the plane's node count, one after another
  int64_t stride;    // first index stride (nodes along the plane's first axis)
  int64_t nodes;
};

__global__ void k_zero(double *a, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// Nodes (a,b), (a+1,b), (a+1,b+1), (a,b+1) with the weights of
// mod_planeMoments.f90:96-141 (w11, w21, w22, w12).
__device__ inline void plane_add(const Plane &P, int a, int b, double pplane, double wa1,
                                 double wa2, double wb1, double wb2, double vx, double vy,
                                 double vz) {
  const double w[4] = {pplane * wa1 * wb1, pplane * wa2 * wb1, pplane * wa2 * wb2,
                       pplane * wa1 * wb2};
  const int64_t nd[4] = {a + P.stride * b, (a + 1) + P.stride * b,
                         (a + 1) + P.stride * (b + 1), a + P.stride * (b + 1)};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double *c = P.acc + nd[q];
    atomicAdd(c, w[q]);
    atomicAdd(c + P.nodes, w[q] * vx);
    atomicAdd(c + 2 * P.nodes, w[q] * vy);
    atomicAdd(c + 3 * P.nodes, w[q] * vz);
    atomicAdd(c + 4 * P.nodes, w[q] * vx * vx);
    atomicAdd(c + 5 * P.nodes, w[q] * vy * vy);
    atomicAdd(c + 6 * P.nodes, w[q] * vz * vz);
  }
}

__global__ __launch_bounds__(BLOCK) void k_plane_acc(KGrid g, lepic_pset p, double dt,
                                                     int ixp, int iyp, int izp, Plane pxy,
                                                     Plane pxz, Plane pyz) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= p.n || p.dead[i] != 0) return;
  const double vx = p.vx[i], vy = p.vy[i], vz = p.vz[i];
  const double x = p.x[i] - vx * dt * 0.5;
  const double y = p.y[i] - vy * dt * 0.5;
  const double z = p.z[i] - vz * dt * 0.5;
  const double qx = qdiv(x, g.q1), qy = qdiv(y, g.q2), qz = qdiv(z, g.q3);
  // int(q) + 1 lies in [1, n] exactly when q lies in (-1, n)
  if (!(qx > -1.0 && qx < (double)g.n1)) return;
  if (!(qy > -1.0 && qy < (double)g.n2)) return;
  if (!(qz > -1.0 && qz < (double)g.n3)) return;
  const int ix = (int)qx + 1, iy = (int)qy + 1, iz = (int)qz + 1;
  const bool in_xy = iz == izp || iz == izp - 1;
  const bool in_xz = iy == iyp || iy == iyp - 1;
  const bool in_yz = ix == ixp || ix == ixp - 1;
  if (!(in_xy || in_xz || in_yz)) return;
  double px = qdiv((double)ix * g.h1 - x, g.q1);
  double py = qdiv((double)iy * g.h2 - y, g.q2);
  double pz = qdiv((double)iz * g.h3 - z, g.q3);
  px = fmax(0.0, fmin(1.0, px));
  py = fmax(0.0, fmin(1.0, py));
  pz = fmax(0.0, fmin(1.0, pz));
  const double wx1 = px, wx2 = 1.0 - px, wy1 = py, wy2 = 1.0 - py, wz1 = pz, wz2 = 1.0 - pz;
  if (in_xy) plane_add(pxy, ix, iy, iz == izp ? pz : 1.0 - pz, wx1, wx2, wy1, wy2, vx, vy, vz);
  if (in_xz) plane_add(pxz, ix, iz, iy == iyp ? py : 1.0 - py, wx1, wx2, wz1, wz2, vx, vy, vz);
  if (in_yz) plane_add(pyz, iy, iz, ix == ixp ? px : 1.0 - px, wy1, wy2, wz1, wz2, vx, vy, vz);
}

// The closing loops of mod_planeMoments.f90:246-302 for one plane: the
// particle sums are added to data(u1..u3) as the reference adds each
// particle's term to what data already held.
__global__ void k_plane_close(Plane P, double *data, double mass) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= P.nodes) return;
  double *d = data + 5 * t;
  const double c = P.acc[t];
  const double u1 = d[2] + P.acc[t + P.nodes];
  const double u2 = d[3] + P.acc[t + 2 * P.nodes];
  const double u3 = d[4] + P.acc[t + 3 * P.nodes];
  if (c > 0.0) {
    const double a1 = u1 / c, a2 = u2 / c, a3 = u3 / c;
    d[0] = c;
    d[2] = a1;
    d[3] = a2;
    d[4] = a3;
    const double uu = a1 * a1 + a2 * a2 + a3 * a3;
    const double v2mean =
        (P.acc[t + 4 * P.nodes] + P.acc[t + 5 * P.nodes] + P.acc[t + 6 * P.nodes]) / c;
    const double th = fmax(0.0, v2mean - uu);
    d[1] = d[1] + mass * th / (3.0 * 1.60217646e-19);
  } else {
    d[2] = u1;
    d[3] = u2;
    d[4] = u3;
  }
}

}  // namespace

extern "C" {

int lepic_plane_moments(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *p, double dt,
                        double mass, int32_t ix_plane, int32_t iy_plane, int32_t iz_plane,
                        double *data_xy, double *data_xz, double *data_yz) {
  if (!ctx || !g || !p || !data_xy || !data_xz || !data_yz) return LEPIC_ERR_ARG;
  if (p->n > 0 && (!p->x || !p->y || !p->z || !p->vx || !p->vy || !p->vz || !p->dead))
    return LEPIC_ERR_ARG;
  if (g->n[0] < 1 || g->n[1] < 1 || g->n[2] < 1) return LEPIC_ERR_ARG;
  // nodes (plane-1 .. plane+1) of each plane must lie in 0..n+2
  if (ix_plane < 1 || ix_plane > g->n[0] + 1 || iy_plane < 1 || iy_plane > g->n[1] + 1 ||
      iz_plane < 1 || iz_plane > g->n[2] + 1)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "plane_moments: plane index outside 1..n+1");
  if (!qdiv_divisor_ok(g->h[0]) || !qdiv_divisor_ok(g->h[1]) || !qdiv_divisor_ok(g->h[2]))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "plane_moments: cell size outside [1e-200, 1e200]");
  KGrid kg = make_kgrid(g);
  const int64_t sx = g->n[0] + 3, sy = g->n[1] + 3, sz = g->n[2] + 3;
  const int64_t nxy = sx * sy, nxz = sx * sz, nyz = sy * sz;
  const int64_t total = NACC * (nxy + nxz + nyz);
  double *acc = reinterpret_cast<double *>(lepic_scratch(ctx, sizeof(double) * (size_t)total, 0));
  if (!acc) return lepic_fail(ctx, LEPIC_ERR_HIP, "plane_moments: scratch allocation failed");
  Plane pxy{acc, sx, nxy};
  Plane pxz{acc + NACC * nxy, sx, nxz};
  Plane pyz{acc + NACC * (nxy + nxz), sy, nyz};
  hipLaunchKernelGGL(k_zero, dim3(grid_blocks(total, BLOCK)), dim3(BLOCK), 0, ctx->stream, acc,
                     total);
  if (p->n > 0)
    hipLaunchKernelGGL(k_plane_acc, dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                       kg, *p, dt, (int)ix_plane, (int)iy_plane, (int)iz_plane, pxy, pxz, pyz);
  hipLaunchKernelGGL(k_plane_close, dim3(grid_blocks(nxy, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     pxy, data_xy, mass);
  hipLaunchKernelGGL(k_plane_close, dim3(grid_blocks(nxz, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     pxz, data_xz, mass);
  hipLaunchKernelGGL(k_plane_close, dim3(grid_blocks(nyz, BLOCK)), dim3(BLOCK), 0, ctx->stream,
                     pyz, data_yz, mass);
  LEPIC_LAUNCH_CHECK(ctx, "plane_moments");
  return LEPIC_OK;
}

}  // extern "C"
