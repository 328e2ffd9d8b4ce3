// lepic_heating.hip -- electron heating of the modular step:
// compute_heating_region_moments and update_heating_vt
// (modules/core/mod_state.f90:524-612) and apply_electron_heating
// (modules/particles/mod_heating.f90:12-82, driven by mod_state.f90:341-373).
//
// 1. k_heat_moments: electrons in the heating window are counted and their
//    kinetic energy summed (wave sums, one atomic per wave).
// 2. k_heat_vt (one thread): Te from the sums and the absorbed power, vt of
//    the re-thermalised electrons; stays on the device for step 3.
// 3. Re-thermalisation with probability nudt per electron in the window:
//    REPLAY_RAN2 runs the reference's ran2 chain on one device thread (draw
//    order, the vz value carried over to the next heated electron);
//    PHILOX gives every electron its own counter-based draws.
// The absorbed energy 0.5 Nm m (v^2 new - v^2 old) accumulates into P_heat.
#include <cmath>
#include <cstring>

#include "lepic_internal.h"
#include "lepic_philox.h"

namespace {

constexpr int BLOCK = 256;
constexpr double PI_H = 3.14159265358979311600;   // 4 atan(1) in double
constexpr double QE_H = 1.60217646e-19;

struct KHeat {
  double nudt, Pabs, nu_h;
  double xl_q, xr_q;         // xl_pow / h1, xr_pow / h1 (the moments' window)
  int ixl_pow, ixr_pow;      // int(x_pow / h1) + 1 (the heating window)
  int flag_circxh, flag_ahp;
  double R_mom, R_app;       // R_ahp seen by the moments and by the heating
  double ymax, zmax, Nm, mass, vt0;
};

struct HeatCtl {
  unsigned long long nh;
  double sum_ek;
  double vt;
  int64_t seed;
};

// Outside the circular heating section (flag_circxh): the particle is
// skipped (mod_heating.f90:47-56, mod_state.f90:551-560).
__device__ inline bool circle_skip(const KHeat &k, double R, double y, double z) {
  if (k.flag_circxh != 1) return false;
  const double dy = y - k.ymax / 2.0, dz = z - k.zmax / 2.0;
  const double r2 = dy * dy + dz * dz;
  return k.flag_ahp == 0 ? r2 > R * R : r2 < R * R;
}

// dead: 0 live, 1 killed by MCC this step (still in the reference's arrays
// when heating runs), 2 hole.  all_present selects which are visited.
__device__ inline bool visited(int8_t d, bool all_present) {
  return all_present ? d != 2 : d == 0;
}

__global__ __launch_bounds__(BLOCK) void k_heat_moments(KGrid g, lepic_pset p, KHeat k,
                                                        int all_present, HeatCtl *ctl) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  double e = 0.0;
  int c = 0;
  if (i < p.n && visited(p.dead[i], all_present)) {
    const int ix = (int)fmax(fmin(qdiv(p.x[i], g.q1), 2.0e9), -2.0e9) + 1;
    const double dix = (double)ix;
    if (!(dix < k.xl_q || dix > k.xr_q) && !circle_skip(k, k.R_mom, p.y[i], p.z[i])) {
      const double vx = p.vx[i], vy = p.vy[i], vz = p.vz[i];
      const double v2 = vx * vx + vy * vy + vz * vz;
      e = 0.5 * k.Nm * k.mass * v2;
      c = 1;
    }
  }
  for (int d = 32; d > 0; d >>= 1) {
    e += __shfl_xor(e, d);
    c += __shfl_xor(c, d);
  }
  if (__lane_id() == 0 && c > 0) {
    atomicAdd(&ctl->nh, (unsigned long long)c);
    atomicAdd(&ctl->sum_ek, e);
  }
}

// update_heating_vt (mod_state.f90:576-612) for one rank.
__global__ void k_heat_vt(KHeat k, HeatCtl *ctl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double vt = k.vt0;
  if (ctl->nh > 0) {
    const double Te = (2.0 / 3.0) * (ctl->sum_ek + k.Pabs / k.nu_h) /
                      (k.Nm * QE_H * (double)(int32_t)ctl->nh);
    if (Te > 0.0) vt = sqrt(2.0 * QE_H * Te / fabs(k.mass));
  }
  ctl->vt = vt;
}

// load_gauss_local (mod_heating.f90:85-97).
__device__ inline void load_gauss(double &v1, double &v2, double vt, double r1, double r2) {
  const double vp = vt * sqrt(-log(1.0 - r1));
  const double theta = 2.0 * PI_H * r2;
  v1 = vp * cos(theta);
  v2 = vp * sin(theta);
}

__device__ inline bool in_window(const KGrid &g, const KHeat &k, double x, double y, double z) {
  const int ix = (int)fmax(fmin(qdiv(x, g.q1), 2.0e9), -2.0e9) + 1;
  if (ix < k.ixl_pow || ix > k.ixr_pow) return false;
  return !circle_skip(k, k.R_app, y, z);
}

struct Ran2 {
  int32_t s;
  __device__ double uniform() {
    const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
    const double am = 1.0 / (double)im;
    const int32_t kk = s / iq;
    s = ia * (s - kk * iq) - ir * kk;
    if (s < 0) s += im;
    return am * (double)s;
  }
};

// apply_electron_heating's loop, in order, on one thread.
__global__ void k_heat_replay(KGrid g, lepic_pset p, KHeat k, HeatCtl *ctl, double *P_heat) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ran2 rng{(int32_t)ctl->seed};
  const double vt = ctl->vt;
  double vz_sav = 0.0;
  double P = *P_heat;
  for (int64_t i = 0; i < p.n; ++i) {
    if (p.dead[i] == 2) continue;
    if (!in_window(g, k, p.x[i], p.y[i], p.z[i])) continue;
    double r1 = rng.uniform();
    if (r1 > k.nudt) continue;
    double vx = p.vx[i], vy = p.vy[i], vz = p.vz[i];
    const double v2old = vx * vx + vy * vy + vz * vz;
    r1 = rng.uniform();
    double r2 = rng.uniform();
    load_gauss(vx, vy, vt, r1, r2);
    if (vz_sav == 0.0) {
      r1 = rng.uniform();
      r2 = rng.uniform();
      load_gauss(vz, vz_sav, vt, r1, r2);
    } else {
      vz = vz_sav;
      vz_sav = 0.0;
    }
    p.vx[i] = vx; p.vy[i] = vy; p.vz[i] = vz;
    P = P + 0.5 * k.Nm * k.mass * (vx * vx + vy * vy + vz * vz - v2old);
  }
  *P_heat = P;
  ctl->seed = rng.s;
}

// Counter-based heating: electron i draws from its own Philox stream, the
// acceptance first, then (vx, vy) and vz from two Box-Muller pairs.
__global__ __launch_bounds__(BLOCK) void k_heat_philox(KGrid g, lepic_pset p, KHeat k,
                                                       const HeatCtl *ctl, uint64_t key,
                                                       uint64_t counter, double *P_heat) {
  const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
  double dE = 0.0;
  if (i < p.n && p.dead[i] == 0 && in_window(g, k, p.x[i], p.y[i], p.z[i])) {
    PhiloxStream rng;
    rng.init(key ^ 0x4EA7ED0E1EC7A11ull, (uint32_t)i, (uint32_t)(i >> 32), (uint32_t)counter);
    rng.c3 ^= (uint32_t)(counter >> 32) * 0x9E3779B9u;
    if (rng.uniform() <= k.nudt) {
      const double vt = ctl->vt;
      double vx = p.vx[i], vy = p.vy[i], vz = p.vz[i], unused;
      const double v2old = vx * vx + vy * vy + vz * vz;
      const double r1 = rng.uniform(), r2 = rng.uniform();
      load_gauss(vx, vy, vt, r1, r2);
      const double r3 = rng.uniform(), r4 = rng.uniform();
      load_gauss(vz, unused, vt, r3, r4);
      p.vx[i] = vx; p.vy[i] = vy; p.vz[i] = vz;
      dE = 0.5 * k.Nm * k.mass * (vx * vx + vy * vy + vz * vz - v2old);
    }
  }
  for (int d = 32; d > 0; d >>= 1) dE += __shfl_xor(dE, d);
  if (__lane_id() == 0 && dE != 0.0) atomicAdd(P_heat, dE);
}

__global__ void k_heat_reset(HeatCtl *ctl, int64_t seed) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctl->nh = 0;
  ctl->sum_ek = 0.0;
  ctl->vt = 0.0;
  ctl->seed = seed;
}

}  // namespace

extern "C" {

int lepic_electron_heating(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p,
                           const lepic_heating *hp, int32_t mode, int32_t *iseed,
                           uint64_t philox_key, uint64_t philox_counter, double *P_heat,
                           double *out) {
  if (!ctx || !g || !p || !hp || !P_heat) return LEPIC_ERR_ARG;
  if (mode == LEPIC_MCC_REPLAY_RAN2 && !iseed) return LEPIC_ERR_ARG;
  if (mode != LEPIC_MCC_REPLAY_RAN2 && mode != LEPIC_MCC_PHILOX)
    return lepic_fail(ctx, LEPIC_ERR_ARG, "electron_heating: unknown mode");
  if (p->n > 0 && (!p->x || !p->y || !p->z || !p->vx || !p->vy || !p->vz || !p->dead))
    return LEPIC_ERR_ARG;
  if (!qdiv_divisor_ok(g->h[0]))
    return lepic_fail(ctx, LEPIC_ERR_ARG, "electron_heating: cell size outside [1e-200, 1e200]");
  KGrid kg = make_kgrid(g);
  KHeat k;
  k.nudt = hp->nudt; k.Pabs = hp->Pabs; k.nu_h = hp->nu_h;
  k.xl_q = hp->xl_pow / g->h[0];
  k.xr_q = hp->xr_pow / g->h[0];
  // int(x_pow / h1) + 1 with the window's ends far outside clamped first
  const double ql = std::fmax(std::fmin(k.xl_q, 2.0e9), -2.0e9);
  const double qr = std::fmax(std::fmin(k.xr_q, 2.0e9), -2.0e9);
  k.ixl_pow = (int)ql + 1;
  k.ixr_pow = (int)qr + 1;
  k.flag_circxh = hp->flag_circxh; k.flag_ahp = hp->flag_ahp;
  k.R_mom = hp->R_ahp_moments; k.R_app = hp->R_ahp;
  k.ymax = hp->ymax; k.zmax = hp->zmax;
  k.Nm = hp->Nm; k.mass = hp->mass; k.vt0 = hp->vt0;
  const bool replay = mode == LEPIC_MCC_REPLAY_RAN2;
  HeatCtl *ctl = reinterpret_cast<HeatCtl *>(ctx->d_ctrl + CTRL_HEAT);
  static_assert(sizeof(HeatCtl) <= 8 * sizeof(int64_t), "heating control fits its words");
  hipLaunchKernelGGL(k_heat_reset, dim3(1), dim3(64), 0, ctx->stream, ctl,
                     (int64_t)(replay ? iseed[0] : 0));
  if (p->n > 0)
    hipLaunchKernelGGL(k_heat_moments, dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0,
                       ctx->stream, kg, *p, k, replay ? 1 : 0, ctl);
  hipLaunchKernelGGL(k_heat_vt, dim3(1), dim3(64), 0, ctx->stream, k, ctl);
  if (p->n > 0) {
    if (replay)
      hipLaunchKernelGGL(k_heat_replay, dim3(1), dim3(64), 0, ctx->stream, kg, *p, k, ctl, P_heat);
    else
      hipLaunchKernelGGL(k_heat_philox, dim3(grid_blocks(p->n, BLOCK)), dim3(BLOCK), 0,
                         ctx->stream, kg, *p, k, ctl, philox_key, philox_counter, P_heat);
  }
  LEPIC_LAUNCH_CHECK(ctx, "electron_heating");
  if (replay || out) {
    LEPIC_HIP(ctx, hipMemcpyAsync(ctx->h_pinned, ctl, sizeof(HeatCtl), hipMemcpyDeviceToHost,
                                  ctx->stream));
    LEPIC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HeatCtl h;
    std::memcpy(&h, ctx->h_pinned, sizeof(h));
    if (replay) iseed[0] = (int32_t)h.seed;
    if (out) {
      out[0] = (double)h.nh;
      out[1] = h.sum_ek;
      out[2] = h.vt;
    }
  }
  return LEPIC_OK;
}

}  // extern "C"
