// lepic_ctx.hip -- context, errors, host RNG and the host loader.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lepic_internal.h"

int lepic_fail(lepic_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int lepic_hip_check(lepic_ctx *ctx, hipError_t e, const char *what) {
  return lepic_fail(ctx, LEPIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void *lepic_scratch(lepic_ctx *ctx, size_t bytes, int which) {
  void **p = which == 0 ? &ctx->scratch : &ctx->scratch2;
  size_t *cur = which == 0 ? &ctx->scratch_bytes : &ctx->scratch2_bytes;
  if (bytes <= *cur) return *p;
  if (*p) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(*p);
    *p = nullptr;
    *cur = 0;
  }
  size_t want = bytes + bytes / 8 + 4096;
  if (hipMalloc(p, want) != hipSuccess) return nullptr;
  *cur = want;
  return *p;
}

__global__ void k_zero_u64(uint64_t *a, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Epoch-tagged look-back status words need no clearing between launches;
// the array is cleared when it grows and when the 16-bit epoch wraps.
uint32_t lepic_next_epoch(lepic_ctx *ctx, int64_t ntiles) {
  bool clear = false;
  if (ntiles > ctx->status_len) {
    if (ctx->d_status) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(ctx->d_status);
    }
    int64_t len = ntiles + ntiles / 4 + 1024;
    if (hipMalloc((void **)&ctx->d_status, sizeof(uint64_t) * (size_t)len) != hipSuccess) {
      ctx->d_status = nullptr;
      ctx->status_len = 0;
      return 0;
    }
    ctx->status_len = len;
    clear = true;
  }
  ctx->epoch = (ctx->epoch + 1) & 0xFFFFu;
  if (ctx->epoch == 0) {
    ctx->epoch = 1;
    clear = true;
  }
  if (clear)
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_blocks(ctx->status_len, 256)), dim3(256), 0,
                       ctx->stream, ctx->d_status, ctx->status_len);
  return ctx->epoch;
}

extern "C" {

const char *lepic_version(void) { return "lepic-hip 0.1 gfx950"; }

int lepic_create(int32_t device, lepic_ctx **out) {
  if (!out) return LEPIC_ERR_ARG;
  *out = nullptr;
  lepic_ctx *ctx = new lepic_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete ctx;
    return LEPIC_ERR_HIP;
  }
  const size_t wd_bytes = sizeof(double) * LEPIC_XCDS * (2 * LEPIC_WD_LABELS + 1);
  if (hipMalloc((void **)&ctx->d_ctrl, sizeof(int64_t) * CTRL_WORDS) != hipSuccess ||
      hipHostMalloc((void **)&ctx->h_pinned, sizeof(int64_t) * CTRL_WORDS) != hipSuccess ||
      hipMemset(ctx->d_ctrl, 0, sizeof(int64_t) * CTRL_WORDS) != hipSuccess ||
      hipMalloc((void **)&ctx->wdiag, wd_bytes) != hipSuccess ||
      hipMemset(ctx->wdiag, 0, wd_bytes) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    delete ctx;
    return LEPIC_ERR_HIP;
  }
  lepic_efield_scratch_setup(device);
  *out = ctx;
  return LEPIC_OK;
}

int lepic_destroy(lepic_ctx *ctx) {
  if (!ctx) return LEPIC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->scratch2) (void)hipFree(ctx->scratch2);
  if (ctx->d_ctrl) (void)hipFree(ctx->d_ctrl);
  if (ctx->wdiag) (void)hipFree(ctx->wdiag);
  if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
  if (ctx->d_status) (void)hipFree(ctx->d_status);
  if (ctx->mg_mem) (void)hipFree(ctx->mg_mem);
  for (hipGraphExec_t &e : ctx->mg_exec)
    if (e) (void)hipGraphExecDestroy(e);
  if (ctx->mgx_mem) (void)hipFree(ctx->mgx_mem);
  if (ctx->mgx_exec) (void)hipGraphExecDestroy(ctx->mgx_exec);
  if (ctx->pcg_exec) (void)hipGraphExecDestroy(ctx->pcg_exec);
  if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
  delete ctx;
  return LEPIC_OK;
}

const char *lepic_last_error(lepic_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int lepic_set_stream(lepic_ctx *ctx, void *hip_stream) {
  if (!ctx) return LEPIC_ERR_ARG;
  ctx->stream = reinterpret_cast<hipStream_t>(hip_stream);
  return LEPIC_OK;
}

int lepic_synchronize(lepic_ctx *ctx) {
  if (!ctx) return LEPIC_ERR_ARG;
  LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LEPIC_OK;
}

// ran2: Park-Miller minimal standard, Schrage factorisation
// (modules/core/mod_RNG.f90:28-45).
double lepic_ran2(int32_t *irand) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  const double am = 1.0 / (double)im;
  int32_t k = *irand / iq;
  *irand = ia * (*irand - k * iq) - ir * k;
  if (*irand < 0) *irand += im;
  return am * (double)(*irand);
}

// 123456*iproc*(10*rank+1) with int32 wrap-around (mod_simParams.f90:202).
int32_t lepic_seed(int32_t iproc, int32_t mpi_rank) {
  uint32_t s = (uint32_t)123456 * (uint32_t)iproc;
  s = s * (uint32_t)(10 * mpi_rank + 1);
  int32_t r = (int32_t)s;
  return r == 0 ? 1 : r;
}

// load_particles_modular (modules/particles/mod_particle_loader.f90:14-246)
// for one virtual thread: rejection of all-wall cells, ni0 test, Box-Muller
// velocities with the cached second Gaussian for vz, initial kq-weighted
// deposit into np_thread.  Host arrays.
int lepic_load_particles_host(const lepic_grid *g, const int32_t *bcnd, const double *kq,
                              double x_load, int32_t np_cell, int32_t ntype, const double *vt0,
                              const double *Nm, const double *ni0, int32_t *iseed,
                              lepic_pset *hp, double *np_thread) {
  if (!g || !bcnd || !kq || !iseed || !hp || !np_thread || ntype < 1 ||
      ntype > LEPIC_MAX_SPECIES)
    return LEPIC_ERR_ARG;
  const int n1 = g->n[0], n2 = g->n[1], n3 = g->n[2];
  const double h1 = g->h[0], h2 = g->h[1], h3 = g->h[2];
  const int64_t sy = n1 + 3, sz = (int64_t)(n1 + 3) * (n2 + 3);
  const int64_t nodes = sz * (n3 + 3);
  auto at = [&](int i, int j, int k) { return (int64_t)i + sy * j + sz * k; };
  const double pi = 3.14159265358979311600;
  if (x_load <= 0.0) x_load = (double)n1 * h1;
  else if (x_load > (double)n1 * h1) x_load = (double)n1 * h1;
  double ymax = (double)n2 * h2, zmax = (double)n3 * h3;
  int ix_load = (int)std::floor(x_load / h1) + 1;
  if (ix_load < 1) ix_load = 1;
  if (ix_load > n1 + 1) ix_load = n1 + 1;
  auto any_plasma = [&](int ix, int iy, int iz) {
    return bcnd[at(ix, iy, iz)] == -1 || bcnd[at(ix + 1, iy, iz)] == -1 ||
           bcnd[at(ix + 1, iy + 1, iz)] == -1 || bcnd[at(ix, iy + 1, iz)] == -1 ||
           bcnd[at(ix, iy, iz + 1)] == -1 || bcnd[at(ix + 1, iy, iz + 1)] == -1 ||
           bcnd[at(ix + 1, iy + 1, iz + 1)] == -1 || bcnd[at(ix, iy + 1, iz + 1)] == -1;
  };
  auto all_wall = [&](int ix, int iy, int iz) {
    return bcnd[at(ix, iy, iz)] >= 1 && bcnd[at(ix + 1, iy, iz)] >= 1 &&
           bcnd[at(ix + 1, iy + 1, iz)] >= 1 && bcnd[at(ix, iy + 1, iz)] >= 1 &&
           bcnd[at(ix, iy, iz + 1)] >= 1 && bcnd[at(ix + 1, iy, iz + 1)] >= 1 &&
           bcnd[at(ix + 1, iy + 1, iz + 1)] >= 1 && bcnd[at(ix, iy + 1, iz + 1)] >= 1;
  };
  int64_t n_cell = 0;
  for (int iz = 1; iz <= n3; ++iz)
    for (int iy = 1; iy <= n2; ++iy)
      for (int ix = 1; ix <= ix_load - 1; ++ix)
        if (any_plasma(ix, iy, iz)) n_cell++;
  int64_t jmax = std::llround((double)((int64_t)np_cell * n_cell));
  std::memset(np_thread, 0, sizeof(double) * (size_t)(nodes * ntype));
  std::vector<double> vz_sav(ntype, 0.0);
  for (int t = 0; t < ntype; ++t) hp[t].n = 0;
  int32_t seed = iseed[0];
  auto gauss = [&](double vt, double r1, double r2, double &a, double &b) {
    double vp = vt * std::sqrt(-std::log(1.0 - r1));
    double th = 2.0 * pi * r2;
    a = vp * std::cos(th);
    b = vp * std::sin(th);
  };
  for (int64_t j = 0; j < jmax; ++j) {
    double x, y, z;
    int ix, iy, iz;
    for (;;) {
      x = lepic_ran2(&seed) * x_load;
      y = lepic_ran2(&seed) * ymax;
      z = lepic_ran2(&seed) * zmax;
      ix = (int)(x / h1) + 1;
      iy = (int)(y / h2) + 1;
      iz = (int)(z / h3) + 1;
      if (!all_wall(ix, iy, iz)) break;
    }
    for (int t = 0; t < ntype; ++t) {
      double r = lepic_ran2(&seed);
      if (r > ni0[t]) continue;
      lepic_pset &p = hp[t];
      if (p.n >= p.cap) {
        iseed[0] = seed;
        return LEPIC_ERR_CAPACITY;
      }
      int64_t k = p.n++;
      double a, b;
      p.x[k] = x; p.y[k] = y; p.z[k] = z;
      double r1 = lepic_ran2(&seed), r2 = lepic_ran2(&seed);
      gauss(vt0[t], r1, r2, a, b);
      p.vx[k] = a; p.vy[k] = b;
      if (vz_sav[t] == 0.0) {
        r1 = lepic_ran2(&seed); r2 = lepic_ran2(&seed);
        gauss(vt0[t], r1, r2, a, b);
        p.vz[k] = a;
        vz_sav[t] = b;
      } else {
        p.vz[k] = vz_sav[t];
        vz_sav[t] = 0.0;
      }
      if (p.dead) p.dead[k] = 0;
      if (p.cell) p.cell[k] = 0;
      double px = ((double)ix * h1 - x) / h1;
      double py = ((double)iy * h2 - y) / h2;
      double pz = ((double)iz * h3 - z) / h3;
      double kp = Nm[t] / (h1 * h2 * h3);
      double ki[8] = {kp * px * py * pz,
                      kp * (1.0 - px) * py * pz,
                      kp * (1.0 - px) * (1.0 - py) * pz,
                      kp * px * (1.0 - py) * pz,
                      kp * px * py * (1.0 - pz),
                      kp * (1.0 - px) * py * (1.0 - pz),
                      kp * (1.0 - px) * (1.0 - py) * (1.0 - pz),
                      kp * px * (1.0 - py) * (1.0 - pz)};
      int64_t c[8] = {at(ix, iy, iz),         at(ix + 1, iy, iz),     at(ix + 1, iy + 1, iz),
                      at(ix, iy + 1, iz),     at(ix, iy, iz + 1),     at(ix + 1, iy, iz + 1),
                      at(ix + 1, iy + 1, iz + 1), at(ix, iy + 1, iz + 1)};
      double *np = np_thread + (int64_t)t * nodes;
      for (int q = 0; q < 8; ++q) np[c[q]] = np[c[q]] + kq[c[q]] * ki[q];
    }
  }
  iseed[0] = seed;
  return LEPIC_OK;
}

}  // extern "C"
