/*
 * lepic_oracle.c -- CPU restatement of the LePIC 3D PIC-MCC step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product package.  Each function names the reference routine it restates
 * (paths relative to the reference repository root).  Expressions keep the
 * Fortran left-to-right evaluation order; the file is compiled with
 * -ffp-contract=off and without fast-math so every product and sum rounds
 * exactly once, as in a strict build of the reference.
 */
#include "lepic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NX(g) ((int64_t)(g)->n[0] + 3)
#define NY(g) ((int64_t)(g)->n[1] + 3)
#define NZ(g) ((int64_t)(g)->n[2] + 3)
#define NODES(g) (NX(g) * NY(g) * NZ(g))
#define IDX(g, i, j, k) ((int64_t)(i) + NX(g) * ((int64_t)(j) + NY(g) * (int64_t)(k)))

/* The multigrid loops below are marked for OpenMP.  The parity library is
 * built without -fopenmp, so they run serially in source order; the host
 * baseline's copy (liblepic_oracle_omp.so) runs them on every core, with the
 * residual summed per thread (a different rounding order, used for timing
 * only). */
#ifdef _OPENMP
#define LO_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#define LO_PARALLEL_SUM_RES _Pragma("omp parallel for schedule(static) reduction(+ : res)")
#else
#define LO_PARALLEL_FOR
#define LO_PARALLEL_SUM_RES
#endif

static const double QE = 1.60217646e-19;           /* mod_constants.f90 qe */
static const double PI_F = 3.14159265358979311600; /* 4*atan(1) in double  */

static inline int32_t imax32(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin32(int32_t a, int32_t b) { return a < b ? a : b; }

/* ===================================================================== */
/* RNG: modules/core/mod_RNG.f90:13-45                                   */
/* ===================================================================== */

/* Park-Miller minimal standard with Schrage factorisation (mod_RNG.f90:28-45). */
double lo_ran2(int32_t *irand) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  const double am = 1.0 / (double)im;
  int32_t k = *irand / iq;
  *irand = ia * (*irand - k * iq) - ir * k;
  if (*irand < 0) *irand += im;
  return am * (double)(*irand);
}

/* Seed formula 123456*iproc*(10*rank+1) (mod_RNG.f90:19, mod_simParams.f90:202),
 * int32 wrap-around made explicit. */
int32_t lo_seed(int32_t iproc, int32_t mpi_rank) {
  int64_t s = (int64_t)123456 * (int64_t)iproc;
  s = (int64_t)(int32_t)(uint32_t)(uint64_t)s;
  s = s * (int64_t)(10 * mpi_rank + 1);
  int32_t r = (int32_t)(uint32_t)(uint64_t)s;
  if (r == 0) r = 1;
  return r;
}

void lo_ran2_fill(int32_t *irand, double *out, int64_t count) {
  for (int64_t i = 0; i < count; ++i) out[i] = lo_ran2(irand);
}

/* ===================================================================== */
/* Loader: modules/particles/mod_particle_loader.f90:14-260 (nproc = 1)  */
/* ===================================================================== */

static void load_gauss(double *vx, double *vy, double vt, double r1, double r2) {
  /* mod_particle_loader.f90:249-260 */
  double vp = vt * sqrt(-log(1.0 - r1));
  double theta = 2.0 * PI_F * r2;
  *vx = vp * cos(theta);
  *vy = vp * sin(theta);
}

static int all_walls(const lo_grid *g, const int32_t *bcnd, int ix, int iy, int iz) {
  return bcnd[IDX(g, ix, iy, iz)] >= 1 && bcnd[IDX(g, ix + 1, iy, iz)] >= 1 &&
         bcnd[IDX(g, ix + 1, iy + 1, iz)] >= 1 && bcnd[IDX(g, ix, iy + 1, iz)] >= 1 &&
         bcnd[IDX(g, ix, iy, iz + 1)] >= 1 && bcnd[IDX(g, ix + 1, iy, iz + 1)] >= 1 &&
         bcnd[IDX(g, ix + 1, iy + 1, iz + 1)] >= 1 && bcnd[IDX(g, ix, iy + 1, iz + 1)] >= 1;
}

int lo_load_particles(const lo_grid *g, const int32_t *bcnd, const double *kq,
                      double x_load, int32_t np_cell, int32_t ntype,
                      const double *vt0, const double *Nm, const double *ni0,
                      int32_t *iseed, lo_pset *parts, double *np_thread) {
  const double *h = g->h;
  const int32_t *n = g->n;
  const int64_t nodes = NODES(g);
  if (x_load <= 0.0) x_load = (double)n[0] * h[0];
  else if (x_load > (double)n[0] * h[0]) x_load = (double)n[0] * h[0];
  double ymax = (double)n[1] * h[1];
  double zmax = (double)n[2] * h[2];
  int32_t ix_load = (int32_t)floor(x_load / h[0]) + 1;
  if (ix_load < 1) ix_load = 1;
  if (ix_load > n[0] + 1) ix_load = n[0] + 1;

  /* loadable cells (mod_particle_loader.f90:75-91) */
  int64_t n_cell = 0;
  for (int iz = 1; iz <= n[2]; ++iz)
    for (int iy = 1; iy <= n[1]; ++iy)
      for (int ix = 1; ix <= ix_load - 1; ++ix) {
        if (bcnd[IDX(g, ix, iy, iz)] == -1 || bcnd[IDX(g, ix + 1, iy, iz)] == -1 ||
            bcnd[IDX(g, ix + 1, iy + 1, iz)] == -1 || bcnd[IDX(g, ix, iy + 1, iz)] == -1 ||
            bcnd[IDX(g, ix, iy, iz + 1)] == -1 || bcnd[IDX(g, ix + 1, iy, iz + 1)] == -1 ||
            bcnd[IDX(g, ix + 1, iy + 1, iz + 1)] == -1 || bcnd[IDX(g, ix, iy + 1, iz + 1)] == -1)
          n_cell++;
      }
  int64_t jmax = llround((double)(np_cell * n_cell) / 1.0);

  memset(np_thread, 0, sizeof(double) * (size_t)(nodes * ntype));
  double vz_sav[16];
  for (int t = 0; t < ntype; ++t) { vz_sav[t] = 0.0; parts[t].n = 0; }
  int32_t seed = iseed[0];

  for (int64_t j = 0; j < jmax; ++j) {
    double x, y, z;
    int ix, iy, iz;
    for (;;) { /* label 70: rejection (mod_particle_loader.f90:128-149) */
      x = lo_ran2(&seed) * x_load;
      y = lo_ran2(&seed) * ymax;
      z = lo_ran2(&seed) * zmax;
      ix = (int32_t)(x / h[0]) + 1;
      iy = (int32_t)(y / h[1]) + 1;
      iz = (int32_t)(z / h[2]) + 1;
      if (!all_walls(g, bcnd, ix, iy, iz)) break;
    }
    for (int t = 0; t < ntype; ++t) {
      double r = lo_ran2(&seed);
      if (r > ni0[t]) continue;
      lo_pset *p = &parts[t];
      if (p->n >= p->cap) { iseed[0] = seed; return LO_ERR_CAPACITY; }
      int64_t k = p->n++;
      double vt = vt0[t], vx, vy;
      p->x[k] = x; p->y[k] = y; p->z[k] = z;
      double r1 = lo_ran2(&seed), r2 = lo_ran2(&seed);
      load_gauss(&vx, &vy, vt, r1, r2);
      p->vx[k] = vx; p->vy[k] = vy;
      if (vz_sav[t] == 0.0) {
        r1 = lo_ran2(&seed); r2 = lo_ran2(&seed);
        load_gauss(&vx, &vy, vt, r1, r2);
        p->vz[k] = vx;
        vz_sav[t] = vy;
      } else {
        p->vz[k] = vz_sav[t];
        vz_sav[t] = 0.0;
      }
      p->dead[k] = 0;
      p->cell[k] = 0;
      double px = ((double)ix * h[0] - x) / h[0];
      double py = ((double)iy * h[1] - y) / h[1];
      double pz = ((double)iz * h[2] - z) / h[2];
      double kp = Nm[t] / (h[0] * h[1] * h[2]);
      double ki[8];
      ki[0] = kp * px * py * pz;
      ki[1] = kp * (1.0 - px) * py * pz;
      ki[2] = kp * (1.0 - px) * (1.0 - py) * pz;
      ki[3] = kp * px * (1.0 - py) * pz;
      ki[4] = kp * px * py * (1.0 - pz);
      ki[5] = kp * (1.0 - px) * py * (1.0 - pz);
      ki[6] = kp * (1.0 - px) * (1.0 - py) * (1.0 - pz);
      ki[7] = kp * px * (1.0 - py) * (1.0 - pz);
      double *np = np_thread + (int64_t)t * nodes;
      const int64_t c[8] = {IDX(g, ix, iy, iz),         IDX(g, ix + 1, iy, iz),
                            IDX(g, ix + 1, iy + 1, iz), IDX(g, ix, iy + 1, iz),
                            IDX(g, ix, iy, iz + 1),     IDX(g, ix + 1, iy, iz + 1),
                            IDX(g, ix + 1, iy + 1, iz + 1), IDX(g, ix, iy + 1, iz + 1)};
      for (int q = 0; q < 8; ++q) np[c[q]] = np[c[q]] + kq[c[q]] * ki[q];
    }
  }
  iseed[0] = seed;
  return LO_OK;
}

/* ===================================================================== */
/* Sort: modules/particles/mod_particle_sorting.f90:16-181               */
/* ===================================================================== */

/* cell_index_from_position (mod_particle_sorting.f90:16-48), returned 0-based. */
int32_t lo_cell_index(double x, double y, double z, const lo_grid *g) {
  int32_t ix = (int32_t)(x / g->h[0]) + 1;
  int32_t iy = (int32_t)(y / g->h[1]) + 1;
  int32_t iz = (int32_t)(z / g->h[2]) + 1;
  ix = imax32(1, imin32(g->n[0], ix));
  iy = imax32(1, imin32(g->n[1], iy));
  iz = imax32(1, imin32(g->n[2], iz));
  return (ix - 1) + g->n[0] * ((iy - 1) + g->n[1] * (iz - 1));
}

/* compute_particle_cell_ids (:51-85) + stable scatter (:88-181). */
void lo_sort_particles_by_cell(const lo_grid *g, lo_pset *p, int32_t *cell_count,
                               int32_t *cell_start) {
  const int64_t ncells = (int64_t)g->n[0] * g->n[1] * g->n[2];
  memset(cell_count, 0, sizeof(int32_t) * (size_t)ncells);
  memset(cell_start, 0, sizeof(int32_t) * (size_t)ncells);
  const int64_t np = p->n;
  if (np <= 0) return;
  for (int64_t i = 0; i < np; ++i) {
    int32_t ic = lo_cell_index(p->x[i], p->y[i], p->z[i], g);
    p->cell[i] = ic;
    cell_count[ic]++;
  }
  cell_start[0] = 0;
  for (int64_t c = 1; c < ncells; ++c) cell_start[c] = cell_start[c - 1] + cell_count[c - 1];
  if (np <= 1) return;

  double *buf = (double *)malloc(sizeof(double) * (size_t)np * 6);
  int32_t *cbuf = (int32_t *)malloc(sizeof(int32_t) * (size_t)np);
  int8_t *dbuf = (int8_t *)malloc((size_t)np);
  int32_t *next = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncells);
  memcpy(next, cell_start, sizeof(int32_t) * (size_t)ncells);
  double *xs = buf, *ys = buf + np, *zs = buf + 2 * np, *vxs = buf + 3 * np, *vys = buf + 4 * np,
         *vzs = buf + 5 * np;
  for (int64_t i = 0; i < np; ++i) {
    int32_t ic = p->cell[i];
    int32_t pos = next[ic]++;
    xs[pos] = p->x[i]; ys[pos] = p->y[i]; zs[pos] = p->z[i];
    vxs[pos] = p->vx[i]; vys[pos] = p->vy[i]; vzs[pos] = p->vz[i];
    cbuf[pos] = ic;
    dbuf[pos] = p->dead[i];
  }
  memcpy(p->x, xs, sizeof(double) * (size_t)np);
  memcpy(p->y, ys, sizeof(double) * (size_t)np);
  memcpy(p->z, zs, sizeof(double) * (size_t)np);
  memcpy(p->vx, vxs, sizeof(double) * (size_t)np);
  memcpy(p->vy, vys, sizeof(double) * (size_t)np);
  memcpy(p->vz, vzs, sizeof(double) * (size_t)np);
  memcpy(p->cell, cbuf, sizeof(int32_t) * (size_t)np);
  memcpy(p->dead, dbuf, (size_t)np);
  free(buf); free(cbuf); free(dbuf); free(next);
}

/* ===================================================================== */
/* Movers                                                                */
/* ===================================================================== */

/* Shared trilinear weights and 8-node gather, in the node order and product
 * order of mod_particleMover.f90:104-142 (ix clamped to [0,n+1]). */
typedef struct {
  int64_t node[8];
  double w[8];
} stencil8;

static inline void stencil_at(const lo_grid *g, double xp, double yp, double zp, stencil8 *s) {
  const double *h = g->h;
  int32_t ix = (int32_t)(xp / h[0]) + 1;
  int32_t iy = (int32_t)(yp / h[1]) + 1;
  int32_t iz = (int32_t)(zp / h[2]) + 1;
  ix = imax32(0, imin32(g->n[0] + 1, ix));
  iy = imax32(0, imin32(g->n[1] + 1, iy));
  iz = imax32(0, imin32(g->n[2] + 1, iz));
  double px = ((double)ix * h[0] - xp) / h[0];
  double py = ((double)iy * h[1] - yp) / h[1];
  double pz = ((double)iz * h[2] - zp) / h[2];
  double wx2 = 1.0 - px, wy2 = 1.0 - py, wz2 = 1.0 - pz;
  s->w[0] = px * py * pz;
  s->w[1] = wx2 * py * pz;
  s->w[2] = wx2 * wy2 * pz;
  s->w[3] = px * wy2 * pz;
  s->w[4] = px * py * wz2;
  s->w[5] = wx2 * py * wz2;
  s->w[6] = wx2 * wy2 * wz2;
  s->w[7] = px * wy2 * wz2;
  s->node[0] = IDX(g, ix, iy, iz);
  s->node[1] = IDX(g, ix + 1, iy, iz);
  s->node[2] = IDX(g, ix + 1, iy + 1, iz);
  s->node[3] = IDX(g, ix, iy + 1, iz);
  s->node[4] = IDX(g, ix, iy, iz + 1);
  s->node[5] = IDX(g, ix + 1, iy, iz + 1);
  s->node[6] = IDX(g, ix + 1, iy + 1, iz + 1);
  s->node[7] = IDX(g, ix, iy + 1, iz + 1);
}

static inline double gather4(const double *F4, const stencil8 *s, int comp) {
  double acc = s->w[0] * F4[4 * s->node[0] + comp];
  for (int q = 1; q < 8; ++q) acc = acc + s->w[q] * F4[4 * s->node[q] + comp];
  return acc;
}

/* move_particles_electrostatic (mod_particleMover.f90:72-156). */
void lo_move_electrostatic(const lo_grid *g, lo_pset *p, const double *E4, double q,
                           double m, double dt) {
  double qmdt = dt * q / m;
  for (int64_t i = 0; i < p->n; ++i) {
    double xp = p->x[i], yp = p->y[i], zp = p->z[i];
    stencil8 s;
    stencil_at(g, xp, yp, zp, &s);
    double Ex = gather4(E4, &s, 0), Ey = gather4(E4, &s, 1), Ez = gather4(E4, &s, 2);
    p->vx[i] = p->vx[i] + qmdt * Ex;
    p->vy[i] = p->vy[i] + qmdt * Ey;
    p->vz[i] = p->vz[i] + qmdt * Ez;
    p->x[i] = xp + dt * p->vx[i];
    p->y[i] = yp + dt * p->vy[i];
    p->z[i] = zp + dt * p->vz[i];
    p->cell[i] = 0;
  }
}

/* Boris velocity update of part_mover (Src/part_expmover.f90:51,139-246) on
 * the modular gather above.  b_mode 1 = constant B (n_B == 1, :142-148);
 * b_mode 2 = map on the E grid (flag_gridB = 0) gated by Bi(4) > 1e-5 T at the
 * lower node (:151-207).  magnetized == 0 reproduces flag_B_pos for ions (:212). */
static void move_boris(const lo_grid *g, lo_pset *p, const double *E4, int b_mode,
                       const double *Bconst, const double *B4, const lo_grid *gb, double q,
                       double m, double dt, int magnetized) {
  double k1 = dt * q / (2.0 * m);
  for (int64_t i = 0; i < p->n; ++i) {
    double xp = p->x[i], yp = p->y[i], zp = p->z[i];
    double vx = p->vx[i], vy = p->vy[i], vz = p->vz[i];
    stencil8 s;
    stencil_at(g, xp, yp, zp, &s);
    double Ex = gather4(E4, &s, 0), Ey = gather4(E4, &s, 1), Ez = gather4(E4, &s, 2);
    int use_b = 0;
    double Bx = 0.0, By = 0.0, Bz = 0.0;
    if (b_mode == 1) {
      Bx = Bconst[0]; By = Bconst[1]; Bz = Bconst[2];
      use_b = 1;
    } else if (b_mode == 2) {
      if (B4[4 * s.node[0] + 3] > 1.0e-5) {
        Bx = gather4(B4, &s, 0); By = gather4(B4, &s, 1); Bz = gather4(B4, &s, 2);
        use_b = 1;
      }
    } else if (b_mode == 3) {
      /* the map on its own grid (part_expmover.f90:149-175, flag_gridB = 1):
       * lower node int(x/hB) + 1 -- inside the domain the same as the
       * clamped stencil of that grid */
      stencil8 sb;
      stencil_at(gb, xp, yp, zp, &sb);
      if (B4[4 * sb.node[0] + 3] > 1.0e-5) {
        Bx = gather4(B4, &sb, 0); By = gather4(B4, &sb, 1); Bz = gather4(B4, &sb, 2);
        use_b = 1;
      }
    }
    if (!magnetized) use_b = 0;
    if (use_b) {
      double Bt = sqrt(Bx * Bx + By * By + Bz * Bz);
      double k2 = k1 * (2.0 / (1.0 + (k1 * Bt) * (k1 * Bt)));
      double v1x = vx + k1 * Ex, v1y = vy + k1 * Ey, v1z = vz + k1 * Ez;
      double v3x = v1x + k1 * (v1y * Bz - v1z * By);
      double v3y = v1y + k1 * (v1z * Bx - v1x * Bz);
      double v3z = v1z + k1 * (v1x * By - v1y * Bx);
      double v2x = v1x + k2 * (v3y * Bz - v3z * By);
      double v2y = v1y + k2 * (v3z * Bx - v3x * Bz);
      double v2z = v1z + k2 * (v3x * By - v3y * Bx);
      vx = v2x + k1 * Ex; vy = v2y + k1 * Ey; vz = v2z + k1 * Ez;
    } else {
      vx = vx + 2.0 * k1 * Ex; vy = vy + 2.0 * k1 * Ey; vz = vz + 2.0 * k1 * Ez;
    }
    p->vx[i] = vx; p->vy[i] = vy; p->vz[i] = vz;
    p->x[i] = xp + dt * vx; p->y[i] = yp + dt * vy; p->z[i] = zp + dt * vz;
    p->cell[i] = 0;
  }
}

void lo_move_boris(const lo_grid *g, lo_pset *p, const double *E4, int b_mode,
                   const double *Bconst, const double *B4, double q, double m,
                   double dt, int magnetized) {
  move_boris(g, p, E4, b_mode, Bconst, B4, NULL, q, m, dt, magnetized);
}

void lo_move_boris_bgrid(const lo_grid *g, lo_pset *p, const double *E4, const double *B4,
                         const lo_grid *gb, double q, double m, double dt, int magnetized) {
  move_boris(g, p, E4, 3, NULL, B4, gb, q, m, dt, magnetized);
}

/* ===================================================================== */
/* Wall BC: modules/particles/mod_particleBC.f90:25-259 (flag_die = 0)   */
/* ===================================================================== */

int lo_apply_bc(const lo_grid *g, lo_pset *p, const int32_t *bcnd, double xmax,
                double ymax, double zmax, int flag_pbc, int flag_nmn, int ptype,
                int tag_neg, double qmacro, double mass, double Nm, int32_t ngrid,
                double *p_mac, double *P_loss, int64_t *np_lost_out) {
  return lo_apply_bc_die(g, p, bcnd, xmax, ymax, zmax, flag_pbc, flag_nmn, ptype, tag_neg,
                         qmacro, mass, Nm, ngrid, 0, NULL, NULL, NULL, p_mac, P_loss,
                         np_lost_out);
}

/* The dielectric branch of apply_particle_bc_legacy (mod_particleBC.f90:
 * 139-188): dtype is read through the routine's 1-based dtype(:) dummy, so
 * label igrid's type is the domain's dtype[igrid - 1]. */
static void die_charge(const lo_grid *g, const int32_t *dtype, int flag_pbc, double zmax,
                       double qmacro, int32_t igrid, int32_t ix, int32_t iy, int32_t iz, double xp,
                       double yp, double zp, double *sxz, double *syz) {
  const double *h = g->h;
  const int32_t *n = g->n;
  const int32_t dt = dtype[igrid - 1];
  if (dt <= 1) return;
  if (flag_pbc == 1) {
    if (zp >= zmax) {
      zp = zp - zmax;
      iz = (int32_t)floor(zp / h[2]) + 1;
    }
    if (zp <= 0.0) {
      zp = zmax + zp;
      iz = (int32_t)floor(zp / h[2]) + 1;
    }
  }
  if (ix < 0) ix = 0;
  if (ix > n[0] + 1) ix = n[0] + 1;
  if (iy < 0) iy = 0;
  if (iy > n[1] + 1) iy = n[1] + 1;
  if (iz < 0) iz = 0;
  if (iz > n[2] + 1) iz = n[2] + 1;
  const double pz = ((double)iz * h[2] - zp) / h[2];
  if (dt == 2) {
    const double px = ((double)ix * h[0] - xp) / h[0];
    const int64_t s = n[0] + 3;
    sxz[ix + s * iz] = sxz[ix + s * iz] + qmacro * px * pz;
    sxz[(ix + 1) + s * iz] = sxz[(ix + 1) + s * iz] + qmacro * (1.0 - px) * pz;
    sxz[(ix + 1) + s * (iz + 1)] = sxz[(ix + 1) + s * (iz + 1)] + qmacro * (1.0 - px) * (1.0 - pz);
    sxz[ix + s * (iz + 1)] = sxz[ix + s * (iz + 1)] + qmacro * px * (1.0 - pz);
  }
  if (dt == 3 || dt == 4) {
    const double py = ((double)iy * h[1] - yp) / h[1];
    const int64_t s = n[1] + 3;
    const int c = dt - 3;
    syz[c + 2 * (iy + s * iz)] = syz[c + 2 * (iy + s * iz)] + qmacro * py * pz;
    syz[c + 2 * ((iy + 1) + s * iz)] = syz[c + 2 * ((iy + 1) + s * iz)] + qmacro * (1.0 - py) * pz;
    syz[c + 2 * ((iy + 1) + s * (iz + 1))] =
        syz[c + 2 * ((iy + 1) + s * (iz + 1))] + qmacro * (1.0 - py) * (1.0 - pz);
    syz[c + 2 * (iy + s * (iz + 1))] = syz[c + 2 * (iy + s * (iz + 1))] + qmacro * py * (1.0 - pz);
  }
}

/* apply_dielectric_bc_to_phi (mod_state.f90:378-430), one rank and thread. */
void lo_dielectric_phi(const lo_grid *g, const int32_t *bcnd, const int32_t *dtype,
                       int32_t ngrid, int32_t ntype, const double *sxz, const double *syz,
                       double Ca, double *phi) {
  const int32_t *n = g->n;
  const int64_t sxz_n = (int64_t)(n[0] + 3) * (n[2] + 3);
  const int64_t syz_n = 2 * (int64_t)(n[1] + 3) * (n[2] + 3);
  for (int32_t k = 0; k <= n[2] + 2; ++k)
    for (int32_t j = 0; j <= n[1] + 2; ++j)
      for (int32_t i = 0; i <= n[0] + 2; ++i) {
        const int64_t c = IDX(g, i, j, k);
        const int32_t igrid = bcnd[c];
        if (igrid <= 0 || igrid > ngrid) continue;
        for (int32_t s = 0; s < ntype; ++s) {
          if (dtype[igrid] == 2)
            phi[c] = phi[c] + (0.0 + sxz[s * sxz_n + i + (int64_t)(n[0] + 3) * k]) / (2.0 * Ca);
          if (dtype[igrid] == 3)
            phi[c] = phi[c] + (0.0 + syz[s * syz_n + 2 * (j + (int64_t)(n[1] + 3) * k)]) / Ca;
          if (dtype[igrid] == 4)
            phi[c] = phi[c] + (0.0 + syz[s * syz_n + 1 + 2 * (j + (int64_t)(n[1] + 3) * k)]) / Ca;
        }
      }
}

int lo_apply_bc_die(const lo_grid *g, lo_pset *p, const int32_t *bcnd, double xmax,
                    double ymax, double zmax, int flag_pbc, int flag_nmn, int ptype,
                    int tag_neg, double qmacro, double mass, double Nm, int32_t ngrid,
                    int flag_die, const int32_t *dtype, double *sum_q_xz, double *sum_q_yz,
                    double *p_mac, double *P_loss, int64_t *np_lost_out) {
  (void)xmax;
  const double *h = g->h;
  const int32_t *n = g->n;
  int64_t np_lost = 0;
  for (int64_t i = 0; i < p->n; ++i) {
    double xp = p->x[i], yp = p->y[i], zp = p->z[i];
    double vx = p->vx[i], vy = p->vy[i], vz = p->vz[i];
    if (p->dead[i] == 1) { np_lost++; continue; }
    int32_t ix = (int32_t)floor(xp / h[0]) + 1;
    int32_t iy = (int32_t)floor(yp / h[1]) + 1;
    int32_t iz = (int32_t)floor(zp / h[2]) + 1;
    if (ix < 0) ix = 0;
    if (ix > n[0] + 1) ix = n[0] + 1;
    if (iy < 0) iy = 0;
    if (iy > n[1] + 1) iy = n[1] + 1;
    if (iz < 0) iz = 0;
    if (iz > n[2] + 1) iz = n[2] + 1;
    int flag_lost = all_walls(g, bcnd, ix, iy, iz) ? 1 : 0;
    if (ptype == tag_neg && xp < 0.0 && flag_nmn == 1) flag_lost = 2;
    if (flag_lost >= 1) {
      int32_t igrid = bcnd[IDX(g, ix, iy, iz)];
      if (flag_lost == 2 && ptype == tag_neg) igrid = 0;
      if (flag_die == 1 && igrid > 0 && igrid <= ngrid + 1)
        die_charge(g, dtype, flag_pbc, zmax, qmacro, igrid, ix, iy, iz, xp, yp, zp, sum_q_xz,
                   sum_q_yz);
      if (igrid < 0) igrid = 0;
      double v2 = vx * vx + vy * vy + vz * vz;
      double Ek_eV = 0.5 * mass * v2 / QE;
      if (igrid <= ngrid) {
        p_mac[2 * igrid + 0] = p_mac[2 * igrid + 0] + qmacro;
        p_mac[2 * igrid + 1] = p_mac[2 * igrid + 1] + fabs(qmacro) * Ek_eV;
      }
      double Ek_J = 0.5 * mass * v2;
      *P_loss = *P_loss + Nm * Ek_J;
      np_lost++;
      continue;
    }
    if (flag_nmn == 1 && xp <= 0.0) { xp = -xp; vx = -vx; }
    if (flag_pbc == 1) {
      if (yp >= ymax) yp = yp - ymax;
      if (yp <= 0.0) yp = ymax + yp;
      if (zp >= zmax) zp = zp - zmax;
      if (zp <= 0.0) zp = zmax + zp;
    }
    int64_t is = i - np_lost;
    p->x[is] = xp; p->y[is] = yp; p->z[is] = zp;
    p->vx[is] = vx; p->vy[is] = vy; p->vz[is] = vz;
    p->dead[is] = 0;
  }
  p->n -= np_lost;
  if (p->n < 0) p->n = 0;
  for (int64_t i = p->n; i < p->cap; ++i) p->dead[i] = 0;
  for (int64_t i = 0; i < p->n; ++i) p->cell[i] = 0;
  *np_lost_out = np_lost;
  return LO_OK;
}

/* ===================================================================== */
/* Deposition / density / rho: modules/fields                            */
/* ===================================================================== */

/* build_kq (mod_charge_weights.f90:9-33). */
void lo_build_kq(const lo_grid *g, const int32_t *bcnd, double *kq) {
  for (int64_t c = 0; c < NODES(g); ++c) kq[c] = (bcnd[c] != -1) ? 2.0 : 1.0;
}

/* deposit_particle_set_to_np_thread (mod_chargeDeposition.f90:25-133). */
void lo_deposit(const lo_grid *g, const lo_pset *p, const double *kq, double Nm,
                double *np_local) {
  const double *h = g->h;
  double xmax = (double)g->n[0] * h[0], ymax = (double)g->n[1] * h[1],
         zmax = (double)g->n[2] * h[2];
  double k1 = Nm / (h[0] * h[1] * h[2]);
  for (int64_t i = 0; i < p->n; ++i) {
    if (p->dead[i] != 0) continue;
    double x = p->x[i], y = p->y[i], z = p->z[i];
    if (x < -1.0e-12 || x > xmax + 1.0e-12) continue;
    if (y < -1.0e-12 || y > ymax + 1.0e-12) continue;
    if (z < -1.0e-12 || z > zmax + 1.0e-12) continue;
    int32_t ix = (int32_t)(x / h[0]) + 1;
    int32_t iy = (int32_t)(y / h[1]) + 1;
    int32_t iz = (int32_t)(z / h[2]) + 1;
    double px = ((double)ix * h[0] - x) / h[0];
    double py = ((double)iy * h[1] - y) / h[1];
    double pz = ((double)iz * h[2] - z) / h[2];
    double wx2 = 1.0 - px, wy2 = 1.0 - py, wz2 = 1.0 - pz;
    double w[8];
    w[0] = k1 * px * py * pz;
    w[1] = k1 * wx2 * py * pz;
    w[2] = k1 * wx2 * wy2 * pz;
    w[3] = k1 * px * wy2 * pz;
    w[4] = k1 * px * py * wz2;
    w[5] = k1 * wx2 * py * wz2;
    w[6] = k1 * wx2 * wy2 * wz2;
    w[7] = k1 * px * wy2 * wz2;
    const int64_t c[8] = {IDX(g, ix, iy, iz),         IDX(g, ix + 1, iy, iz),
                          IDX(g, ix + 1, iy + 1, iz), IDX(g, ix, iy + 1, iz),
                          IDX(g, ix, iy, iz + 1),     IDX(g, ix + 1, iy, iz + 1),
                          IDX(g, ix + 1, iy + 1, iz + 1), IDX(g, ix, iy + 1, iz + 1)};
    for (int q = 0; q < 8; ++q) np_local[c[q]] = np_local[c[q]] + kq[c[q]] * w[q];
  }
}

/* The split of one particle's full weight U (integer units held in a double)
 * over its 8 stencil nodes, w1..w8 order: divided along x, then y, then z,
 * the upper node taking rint(parent * share) and the lower node the rest, so
 * the parts are integers summing to U exactly.  ux, uy, uz: upper shares. */
void lo_fixed_split(double U, double ux, double uy, double uz, int64_t a[8]) {
  const double x1 = rint(U * ux), x0 = U - x1;
  const double x0y1 = rint(x0 * uy), x0y0 = x0 - x0y1;
  const double x1y1 = rint(x1 * uy), x1y0 = x1 - x1y1;
  const double c[4] = {x0y0, x1y0, x1y1, x0y1};
  for (int q = 0; q < 4; ++q) {
    const double z1 = rint(c[q] * uz);
    a[q] = (int64_t)(c[q] - z1);
    a[q + 4] = (int64_t)z1;
  }
}

/* lo_deposit in fixed point (the contract of lepic_deposit_fixed): with the
 * fractions lo_deposit forms (same guard, same arithmetic), each particle's
 * full weight U = rint(|k1| 2^frac_bits) is split by lo_fixed_split (upper
 * shares 1 - px, 1 - py, 1 - pz) and the parts are summed per node as
 * wrapping 64-bit integers (negated when k1 < 0); then
 * np += kq * ((double)sum * 2^-frac_bits) at every node. */
void lo_deposit_fixed(const lo_grid *g, const lo_pset *p, const double *kq, double Nm,
                      int32_t frac_bits, double *np_local) {
  const double *h = g->h;
  const int64_t nodes = (int64_t)(g->n[0] + 3) * (g->n[1] + 3) * (g->n[2] + 3);
  uint64_t *acc = (uint64_t *)calloc((size_t)nodes, sizeof(uint64_t));
  double xmax = (double)g->n[0] * h[0], ymax = (double)g->n[1] * h[1],
         zmax = (double)g->n[2] * h[2];
  double k1 = Nm / (h[0] * h[1] * h[2]);
  const double U = rint(ldexp(fabs(k1), frac_bits)), inv = ldexp(1.0, -frac_bits);
  const int neg = k1 < 0.0;
  for (int64_t i = 0; i < p->n; ++i) {
    if (p->dead[i] != 0) continue;
    double x = p->x[i], y = p->y[i], z = p->z[i];
    if (x < -1.0e-12 || x > xmax + 1.0e-12) continue;
    if (y < -1.0e-12 || y > ymax + 1.0e-12) continue;
    if (z < -1.0e-12 || z > zmax + 1.0e-12) continue;
    int32_t ix = (int32_t)(x / h[0]) + 1;
    int32_t iy = (int32_t)(y / h[1]) + 1;
    int32_t iz = (int32_t)(z / h[2]) + 1;
    double px = ((double)ix * h[0] - x) / h[0];
    double py = ((double)iy * h[1] - y) / h[1];
    double pz = ((double)iz * h[2] - z) / h[2];
    double wx2 = 1.0 - px, wy2 = 1.0 - py, wz2 = 1.0 - pz;
    int64_t a[8];
    lo_fixed_split(U, wx2, wy2, wz2, a);
    const int64_t c[8] = {IDX(g, ix, iy, iz),         IDX(g, ix + 1, iy, iz),
                          IDX(g, ix + 1, iy + 1, iz), IDX(g, ix, iy + 1, iz),
                          IDX(g, ix, iy, iz + 1),     IDX(g, ix + 1, iy, iz + 1),
                          IDX(g, ix + 1, iy + 1, iz + 1), IDX(g, ix, iy + 1, iz + 1)};
    for (int q = 0; q < 8; ++q) acc[c[q]] += (uint64_t)(neg ? -a[q] : a[q]);
  }
  for (int64_t t = 0; t < nodes; ++t)
    np_local[t] = np_local[t] + kq[t] * ((double)(int64_t)acc[t] * inv);
  free(acc);
}

/* load_gauss_local (mod_heating.f90:85-97). */
static void heat_gauss(double *v1, double *v2, double vt, double r1, double r2) {
  const double vp = vt * sqrt(-log(1.0 - r1));
  const double theta = 2.0 * PI_F * r2;
  *v1 = vp * cos(theta);
  *v2 = vp * sin(theta);
}

static int heat_circle_skip(const lo_heating *k, double R, double y, double z) {
  if (k->flag_circxh != 1) return 0;
  const double dy = y - k->ymax / 2.0, dz = z - k->zmax / 2.0;
  const double r2 = dy * dy + dz * dz;
  return k->flag_ahp == 0 ? r2 > R * R : r2 < R * R;
}

/* compute_heating_region_moments (mod_state.f90:524-573), update_heating_vt
 * (:576-612) and apply_electron_heating (mod_heating.f90:12-82). */
void lo_electron_heating(const lo_grid *g, lo_pset *p, const lo_heating *k,
                         int32_t *iseed, double *P_heat, double *out) {
  const double h1 = g->h[0];
  int32_t Nh = 0;
  double sum_ek = 0.0;
  for (int64_t i = 0; i < p->n; ++i) {
    const int32_t ix = (int32_t)(p->x[i] / h1) + 1;
    if ((double)ix < k->xl_pow / h1 || (double)ix > k->xr_pow / h1) continue;
    if (heat_circle_skip(k, k->R_ahp_moments, p->y[i], p->z[i])) continue;
    const double v2 = p->vx[i] * p->vx[i] + p->vy[i] * p->vy[i] + p->vz[i] * p->vz[i];
    sum_ek = sum_ek + 0.5 * k->Nm * k->mass * v2;
    Nh = Nh + 1;
  }
  double vt = k->vt0;
  if (Nh > 0) {
    const double Te = (2.0 / 3.0) * (sum_ek + k->Pabs / k->nu_h) / (k->Nm * QE * (double)Nh);
    if (Te > 0.0) vt = sqrt(2.0 * QE * Te / fabs(k->mass));
  }
  if (out) {
    out[0] = (double)Nh;
    out[1] = sum_ek;
    out[2] = vt;
  }
  const int32_t ixl = (int32_t)(k->xl_pow / h1) + 1, ixr = (int32_t)(k->xr_pow / h1) + 1;
  double vz_sav = 0.0;
  for (int64_t i = 0; i < p->n; ++i) {
    const int32_t ix = (int32_t)(p->x[i] / h1) + 1;
    if (ix < ixl || ix > ixr) continue;
    if (heat_circle_skip(k, k->R_ahp, p->y[i], p->z[i])) continue;
    double r1 = lo_ran2(iseed);
    if (r1 > k->nudt) continue;
    const double v2old = p->vx[i] * p->vx[i] + p->vy[i] * p->vy[i] + p->vz[i] * p->vz[i];
    r1 = lo_ran2(iseed);
    double r2 = lo_ran2(iseed);
    heat_gauss(&p->vx[i], &p->vy[i], vt, r1, r2);
    if (vz_sav == 0.0) {
      r1 = lo_ran2(iseed);
      r2 = lo_ran2(iseed);
      heat_gauss(&p->vz[i], &vz_sav, vt, r1, r2);
    } else {
      p->vz[i] = vz_sav;
      vz_sav = 0.0;
    }
    *P_heat = *P_heat + 0.5 * k->Nm * k->mass *
              (p->vx[i] * p->vx[i] + p->vy[i] * p->vy[i] + p->vz[i] * p->vz[i] - v2old);
  }
}

/* One plane of compute_particle_plane_moments_species: the bilinear
 * deposit of mod_planeMoments.f90:96-141 (and :146-191, :196-241) into cnt,
 * data(u1..u3) and the <v^2> sums, nodes (a,b), (a+1,b), (a+1,b+1), (a,b+1). */
static void plane_add(double *data, double *acc, int64_t stride, int32_t a, int32_t b,
                      double pplane, double wa1, double wa2, double wb1, double wb2,
                      double vx, double vy, double vz) {
  const double w[4] = {pplane * wa1 * wb1, pplane * wa2 * wb1, pplane * wa2 * wb2,
                       pplane * wa1 * wb2};
  const int64_t nd[4] = {a + stride * b, (a + 1) + stride * b, (a + 1) + stride * (b + 1),
                         a + stride * (b + 1)};
  for (int q = 0; q < 4; ++q) acc[4 * nd[q] + 0] = acc[4 * nd[q] + 0] + w[q];
  for (int q = 0; q < 4; ++q) data[5 * nd[q] + 2] = data[5 * nd[q] + 2] + w[q] * vx;
  for (int q = 0; q < 4; ++q) data[5 * nd[q] + 3] = data[5 * nd[q] + 3] + w[q] * vy;
  for (int q = 0; q < 4; ++q) data[5 * nd[q] + 4] = data[5 * nd[q] + 4] + w[q] * vz;
  for (int q = 0; q < 4; ++q) acc[4 * nd[q] + 1] = acc[4 * nd[q] + 1] + w[q] * vx * vx;
  for (int q = 0; q < 4; ++q) acc[4 * nd[q] + 2] = acc[4 * nd[q] + 2] + w[q] * vy * vy;
  for (int q = 0; q < 4; ++q) acc[4 * nd[q] + 3] = acc[4 * nd[q] + 3] + w[q] * vz * vz;
}

/* The closing loops of mod_planeMoments.f90:246-302 over one plane. */
static void plane_close(double *data, const double *acc, int64_t nodes, double mass) {
  for (int64_t t = 0; t < nodes; ++t) {
    const double c = acc[4 * t + 0];
    if (c > 0.0) {
      double *d = data + 5 * t;
      d[0] = c;
      d[2] = d[2] / c;
      d[3] = d[3] / c;
      d[4] = d[4] / c;
      const double u2 = d[2] * d[2] + d[3] * d[3] + d[4] * d[4];
      const double v2mean = (acc[4 * t + 1] + acc[4 * t + 2] + acc[4 * t + 3]) / c;
      const double th = v2mean - u2 > 0.0 ? v2mean - u2 : 0.0;
      d[1] = d[1] + mass * th / (3.0 * QE);
    }
  }
}

/* compute_particle_plane_moments_species (mod_planeMoments.f90:13-308). */
void lo_plane_moments(const lo_grid *g, const lo_pset *p, double dt, double mass,
                      int32_t ix_plane, int32_t iy_plane, int32_t iz_plane,
                      double *data_xy, double *data_xz, double *data_yz) {
  const int32_t *n = g->n;
  const double *h = g->h;
  const int64_t sx = n[0] + 3, sy = n[1] + 3, sz = n[2] + 3;
  double *axy = (double *)calloc((size_t)(4 * sx * sy), sizeof(double));
  double *axz = (double *)calloc((size_t)(4 * sx * sz), sizeof(double));
  double *ayz = (double *)calloc((size_t)(4 * sy * sz), sizeof(double));
  for (int64_t i = 0; i < p->n; ++i) {
    if (p->dead[i] != 0) continue;
    const double x = p->x[i] - p->vx[i] * dt * 0.5;
    const double y = p->y[i] - p->vy[i] * dt * 0.5;
    const double z = p->z[i] - p->vz[i] * dt * 0.5;
    const double vx = p->vx[i], vy = p->vy[i], vz = p->vz[i];
    const double qx = x / h[0], qy = y / h[1], qz = z / h[2];
    /* int(q) + 1 in [1, n] exactly when q lies in (-1, n) (truncation) */
    if (!(qx > -1.0 && qx < (double)n[0])) continue;
    if (!(qy > -1.0 && qy < (double)n[1])) continue;
    if (!(qz > -1.0 && qz < (double)n[2])) continue;
    const int32_t ix = (int32_t)qx + 1, iy = (int32_t)qy + 1, iz = (int32_t)qz + 1;
    double px = ((double)ix * h[0] - x) / h[0];
    double py = ((double)iy * h[1] - y) / h[1];
    double pz = ((double)iz * h[2] - z) / h[2];
    px = px < 1.0 ? px : 1.0; px = px > 0.0 ? px : 0.0;
    py = py < 1.0 ? py : 1.0; py = py > 0.0 ? py : 0.0;
    pz = pz < 1.0 ? pz : 1.0; pz = pz > 0.0 ? pz : 0.0;
    const double wx1 = px, wx2 = 1.0 - px, wy1 = py, wy2 = 1.0 - py, wz1 = pz, wz2 = 1.0 - pz;
    if (iz == iz_plane || iz == iz_plane - 1)
      plane_add(data_xy, axy, sx, ix, iy, iz == iz_plane ? pz : 1.0 - pz, wx1, wx2, wy1, wy2,
                vx, vy, vz);
    if (iy == iy_plane || iy == iy_plane - 1)
      plane_add(data_xz, axz, sx, ix, iz, iy == iy_plane ? py : 1.0 - py, wx1, wx2, wz1, wz2,
                vx, vy, vz);
    if (ix == ix_plane || ix == ix_plane - 1)
      plane_add(data_yz, ayz, sy, iy, iz, ix == ix_plane ? px : 1.0 - px, wy1, wy2, wz1, wz2,
                vx, vy, vz);
  }
  plane_close(data_xy, axy, sx * sy, mass);
  plane_close(data_xz, axz, sx * sz, mass);
  plane_close(data_yz, ayz, sy * sz, mass);
  free(axy);
  free(axz);
  free(ayz);
}

/* apply_periodic_density_bc (mod_density.f90:180-242), y then z stitching. */
static void periodic_density(const lo_grid *g, const int32_t *bcnd, double *np) {
  const int32_t *n = g->n;
  for (int iz = 1; iz <= n[2] + 1; ++iz)
    for (int ix = 1; ix <= n[0] + 1; ++ix) {
      if (bcnd[IDX(g, ix, 1, iz)] == 0) {
        np[IDX(g, ix, 1, iz)] = 0.5 * (np[IDX(g, ix, 1, iz)] + np[IDX(g, ix, n[1] + 1, iz)]);
        np[IDX(g, ix, 0, iz)] = np[IDX(g, ix, n[1], iz)];
      }
      if (bcnd[IDX(g, ix, n[1] + 1, iz)] == 0) {
        np[IDX(g, ix, n[1] + 2, iz)] = np[IDX(g, ix, 2, iz)];
        np[IDX(g, ix, n[1] + 1, iz)] = np[IDX(g, ix, 1, iz)];
      }
    }
  for (int iy = 1; iy <= n[1] + 1; ++iy)
    for (int ix = 1; ix <= n[0] + 1; ++ix) {
      if (bcnd[IDX(g, ix, iy, 1)] == 0) {
        np[IDX(g, ix, iy, 1)] = 0.5 * (np[IDX(g, ix, iy, 1)] + np[IDX(g, ix, iy, n[2] + 1)]);
        np[IDX(g, ix, iy, 0)] = np[IDX(g, ix, iy, n[2])];
      }
      if (bcnd[IDX(g, ix, iy, n[2] + 1)] == 0) {
        np[IDX(g, ix, iy, n[2] + 2)] = np[IDX(g, ix, iy, 2)];
        np[IDX(g, ix, iy, n[2] + 1)] = np[IDX(g, ix, iy, 1)];
      }
    }
}

/* reduce_species_density (mod_density.f90:31-73), single MPI rank.
 * np_thread is [nproc][ntype][nodes]; np_red is [ntype][nodes]. */
void lo_reduce_density(const lo_grid *g, const int32_t *bcnd, int32_t ntype, int32_t nproc,
                       double *np_thread, double *np_red) {
  const int64_t nodes = NODES(g);
  const int32_t *n = g->n;
  for (int ip = 0; ip < nproc; ++ip)
    for (int t = 0; t < ntype; ++t)
      periodic_density(g, bcnd, np_thread + ((int64_t)ip * ntype + t) * nodes);
  memset(np_red, 0, sizeof(double) * (size_t)(nodes * ntype));
  for (int t = 0; t < ntype; ++t)
    for (int iz = 1; iz <= n[2] + 1; ++iz)
      for (int iy = 1; iy <= n[1] + 1; ++iy)
        for (int ix = 1; ix <= n[0] + 1; ++ix) {
          int64_t c = IDX(g, ix, iy, iz);
          double acc = 0.0;
          for (int ip = 0; ip < nproc; ++ip) acc = acc + np_thread[((int64_t)ip * ntype + t) * nodes + c];
          np_red[(int64_t)t * nodes + c] = acc;
        }
}

/* build_rho_from_np (mod_density.f90:76-107). */
void lo_build_rho(const lo_grid *g, int32_t ntype, const double *np_red, const double *charge,
                  double *rho) {
  const int64_t nodes = NODES(g);
  const int32_t *n = g->n;
  memset(rho, 0, sizeof(double) * (size_t)nodes);
  for (int iz = 1; iz <= n[2] + 1; ++iz)
    for (int iy = 1; iy <= n[1] + 1; ++iy)
      for (int ix = 1; ix <= n[0] + 1; ++ix) {
        int64_t c = IDX(g, ix, iy, iz);
        double r = 0.0;
        for (int t = 0; t < ntype; ++t) r = r - charge[t] * np_red[(int64_t)t * nodes + c];
        rho[c] = r;
      }
}

/* ===================================================================== */
/* Poisson: modules/legacy/pdesolver.f90, mg.f90, sors.f90 (1 rank)      */
/* ===================================================================== */

typedef struct {
  lo_grid g;     /* this level's n and h */
  double *u;     /* solution / correction, (n+3)^3 (plane k=-1 never used) */
  double *b;     /* right-hand side (indices 0..n+1 used)                */
  int32_t *bcnd; /* labels                                               */
  double aw, ae, an, as, at, ab, ac;
} mg_level;

static void coeffs(mg_level *L, double eps0) {
  /* mg.h:9-15 */
  const double *h = L->g.h;
  L->aw = eps0 / (h[0] * h[0]);
  L->ae = L->aw;
  L->an = eps0 / (h[1] * h[1]);
  L->as = L->an;
  L->at = eps0 / (h[2] * h[2]);
  L->ab = L->at;
  L->ac = -(L->an + L->as + L->aw + L->ae + L->at + L->ab);
}

/* getres (mg.f90:809-848). */
static double getres(const mg_level *L, int ir, int jr, int kr, int flag_nmn) {
  const lo_grid *g = &L->g;
  /* guard the out-of-range read the reference can make at ir = n+2 */
  if (ir < 0 || ir > g->n[0] + 1 || jr < 0 || jr > g->n[1] + 1 || kr < 0 || kr > g->n[2] + 1)
    return 0.0;
  const double *u = L->u;
  double aw = L->aw, ae = L->ae;
  double uij = u[IDX(g, ir, jr, kr)];
  double un = u[IDX(g, ir, jr + 1, kr)];
  double us = u[IDX(g, ir, jr - 1, kr)];
  double ue = u[IDX(g, ir + 1, jr, kr)];
  double ut = u[IDX(g, ir, jr, kr + 1)];
  double ub = kr >= 1 ? u[IDX(g, ir, jr, kr - 1)] : 0.0;
  double uw;
  if (ir == 1 && flag_nmn == 1) {
    uw = 0.0; aw = 0.0; ae = 2.0 * ae;
  } else {
    uw = u[IDX(g, ir - 1, jr, kr)];
  }
  double urhs = L->b[IDX(g, ir, jr, kr)];
  return urhs - (aw * uw + ae * ue + L->an * un + L->as * us + L->ac * uij + L->at * ut +
                 L->ab * ub);
}

static void periodic_u(const mg_level *L) {
  /* sors.f90:184-239 (single rank: u(:,:,0)=u(:,:,m), u(:,:,m+1)=u(:,:,1)) */
  const lo_grid *g = &L->g;
  const int32_t *n = g->n;
  double *u = L->u;
  const int32_t *bc = L->bcnd;
  const int m = n[2];
  for (int k = 0; k <= m + 2; ++k)
    for (int i = 0; i <= n[0] + 2; ++i) {
      int32_t b1 = bc[IDX(g, i, 1, k)], b2 = bc[IDX(g, i, n[1] + 1, k)];
      if (b1 == 0 || b1 == -2) u[IDX(g, i, 0, k)] = u[IDX(g, i, n[1], k)];
      if (b2 == 0 || b2 == -2) u[IDX(g, i, n[1] + 1, k)] = u[IDX(g, i, 1, k)];
    }
  for (int j = 0; j <= n[1] + 2; ++j)
    for (int i = 0; i <= n[0] + 2; ++i) {
      int32_t b1 = bc[IDX(g, i, j, 1)], b2 = bc[IDX(g, i, j, m + 1)];
      if (b1 == 0 || b1 == -2) u[IDX(g, i, j, 0)] = u[IDX(g, i, j, m)];
      if (b2 == 0 || b2 == -2) u[IDX(g, i, j, m + 1)] = u[IDX(g, i, j, 1)];
    }
}

/* One colour of the red-black SOR (sors.f90:87-140 red, :249-302 black).
 * colour 0 = red = (i+j+k) odd, colour 1 = black = (i+j+k) even. */
static void sor_colour(const mg_level *L, int colour, double omega) {
  const lo_grid *g = &L->g;
  const int32_t *n = g->n;
  double *u = L->u;
  const double *b = L->b;
  const int32_t *bc = L->bcnd;
  const int m = n[2];
  LO_PARALLEL_FOR
  for (int k = 1; k <= m + 1; ++k)
    for (int j = 1; j <= n[1] + 1; ++j) {
      int ks = (k % 2 == 0) ? 0 : 1;
      int is;
      if (colour == 0) is = (j % 2 == 0) ? 1 : 2;
      else is = (j % 2 == 0) ? 2 : 1;
      is = is + ks;
      if (is > 2) is = 1;
      for (int i = is; i <= n[0] + 1; i += 2) {
        int64_t c = IDX(g, i, j, k);
        int32_t bcv = bc[c];
        if (bcv > 0) continue;
        if (bcv == -2) {
          u[c] = (1.0 - omega) * u[c] +
                 omega * (b[c] - L->an * u[IDX(g, i, j + 1, k)] - L->as * u[IDX(g, i, j - 1, k)] -
                          2.0 * L->ae * u[IDX(g, i + 1, j, k)] - L->at * u[IDX(g, i, j, k + 1)] -
                          L->ab * u[IDX(g, i, j, k - 1)]) /
                     L->ac;
          continue;
        }
        if (bcv == 0) {
          if (k == m + 1) continue;
          if (j == n[1] + 1) continue;
        }
        u[c] = (1.0 - omega) * u[c] +
               omega * (b[c] - L->aw * u[IDX(g, i - 1, j, k)] - L->an * u[IDX(g, i, j + 1, k)] -
                        L->as * u[IDX(g, i, j - 1, k)] - L->ae * u[IDX(g, i + 1, j, k)] -
                        L->at * u[IDX(g, i, j, k + 1)] - L->ab * u[IDX(g, i, j, k - 1)]) /
                   L->ac;
      }
    }
}

/* L1 residual over bcnd == -1 nodes (sors.f90:409-431). */
static double sor_residual(const mg_level *L) {
  const lo_grid *g = &L->g;
  const int32_t *n = g->n;
  const double *u = L->u, *b = L->b;
  double res = 0.0;
  LO_PARALLEL_SUM_RES
  for (int k = 1; k <= n[2] + 1; ++k)
    for (int j = 1; j <= n[1] + 1; ++j)
      for (int i = 1; i <= n[0] + 1; ++i) {
        int64_t c = IDX(g, i, j, k);
        double rij = 0.0;
        if (L->bcnd[c] == -1)
          rij = b[c] - (L->aw * u[IDX(g, i - 1, j, k)] + L->an * u[IDX(g, i, j + 1, k)] +
                        L->as * u[IDX(g, i, j - 1, k)] + L->ae * u[IDX(g, i + 1, j, k)] +
                        L->ac * u[c] + L->at * u[IDX(g, i, j, k + 1)] +
                        L->ab * u[IDX(g, i, j, k - 1)]);
        res = res + fabs(rij);
      }
  return res;
}

/* sor_rb (sors.f90:1-535), single rank (no concatenation/splitting). */
static int sor_rb(const mg_level *L, double *res_out, double omega, double eps, int kmg, int dig,
                  int flag_pbc) {
  double res_s = 0.0, res = 0.0;
  int ksor = 0;
  long nr = 4;
  if (kmg == 1 && dig == 0) nr = 1000000;
  const double eta = 0.6;
  for (long l = 1; l <= nr; ++l) {
    sor_colour(L, 0, omega);
    if (flag_pbc) periodic_u(L);
    sor_colour(L, 1, omega);
    if (flag_pbc) periodic_u(L);
    res = sor_residual(L);
    ksor = ksor + 1;
    if (dig == 0) {
      if (res <= eps) break;
    } else {
      if (res_s > 0.0 && (res / res_s) <= eta) break;
    }
    res_s = res;
  }
  *res_out = res;
  return ksor;
}

/* restriction (mg.f90:249-806), single rank, flag = 1 (even n3). */
static void restriction(const mg_level *F, mg_level *C, int flag_pbc, int flag_nmn) {
  const lo_grid *g = &F->g;
  const lo_grid *gc = &C->g;
  const int32_t *n = g->n;
  const int m = n[2];
  const int km = 1, kp = m / 2 + 1;
  const int32_t *bc = F->bcnd;
  int32_t *bc2 = C->bcnd;
  /* r_pbc(n1+1, n2+1): residual of plane k = m, periodic image of k = 0 */
  double *r_pbc = (double *)calloc((size_t)(n[0] + 2) * (size_t)(n[1] + 2), sizeof(double));
#define RP(i, j) r_pbc[(int64_t)(i) + (int64_t)(n[0] + 2) * (int64_t)(j)]
  if (flag_pbc) {
    for (int j = 1; j <= n[1] / 2 + 1; ++j)
      for (int i = 1; i <= n[0] / 2 + 1; ++i) {
        int ir = 2 * i - 1, jr = 2 * j - 1;
        if (bc[IDX(g, ir, jr, m + 1)] <= 0) {
          if (ir != 1 && jr != 1) RP(ir - 1, jr - 1) = getres(F, ir - 1, jr - 1, m, flag_nmn);
          if (ir != 1) RP(ir - 1, jr) = getres(F, ir - 1, jr, m, flag_nmn);
          if (ir != 1 && jr != n[1] + 1) RP(ir - 1, jr + 1) = getres(F, ir - 1, jr + 1, m, flag_nmn);
          if (ir != 1 && jr != 1) RP(ir, jr - 1) = getres(F, ir, jr - 1, m, flag_nmn);
          RP(ir, jr) = getres(F, ir, jr, m, flag_nmn);
          if (jr != n[1] + 1) RP(ir, jr + 1) = getres(F, ir, jr + 1, m, flag_nmn);
          if (ir != n[0] + 1 && jr != 1) RP(ir + 1, jr - 1) = getres(F, ir + 1, jr - 1, m, flag_nmn);
          if (ir != n[0] + 1) RP(ir + 1, jr) = getres(F, ir + 1, jr, m, flag_nmn);
          if (ir != n[0] + 1 && jr != n[1] + 1)
            RP(ir + 1, jr + 1) = getres(F, ir + 1, jr + 1, m, flag_nmn);
        } else {
          if (ir != 1 && jr != 1) RP(ir - 1, jr - 1) = 0.0;
          if (ir != 1) RP(ir - 1, jr) = 0.0;
          if (ir != 1 && jr != n[1] + 1) RP(ir - 1, jr + 1) = 0.0;
          if (jr != 1) RP(ir, jr - 1) = 0.0;
          RP(ir, jr) = 0.0;
          if (jr != n[1] + 1) RP(ir, jr + 1) = 0.0;
          if (ir != n[0] + 1 && jr != 1) RP(ir + 1, jr - 1) = 0.0;
          if (ir != n[0] + 1) RP(ir + 1, jr) = 0.0;
          if (ir != n[0] + 1 && jr != n[1] + 1) RP(ir + 1, jr + 1) = 0.0;
        }
      }
  }
  memset(C->u, 0, sizeof(double) * (size_t)NODES(gc));
  const int N1 = n[0] + 1, N2 = n[1] + 1, M1 = m + 1;
  LO_PARALLEL_FOR
  for (int k = km; k <= kp; ++k)
    for (int j = 1; j <= n[1] / 2 + 1; ++j)
      for (int i = 1; i <= n[0] / 2 + 1; ++i) {
        int ir = 2 * i - 1, jr = 2 * j - 1, kr = 2 * k - 1;
        bc2[IDX(gc, i, j, k)] = bc[IDX(g, ir, jr, kr)];
        if (ir == 1) bc2[IDX(gc, 0, j, k)] = bc[IDX(g, 0, jr, kr)];
        if (ir == N1) bc2[IDX(gc, n[0] / 2 + 2, j, k)] = bc[IDX(g, n[0] + 2, jr, kr)];
        if (jr == 1) bc2[IDX(gc, i, 0, k)] = bc[IDX(g, ir, 0, kr)];
        if (jr == N2) bc2[IDX(gc, i, n[1] / 2 + 2, k)] = bc[IDX(g, ir, n[1] + 2, kr)];
        if (kr == 1) bc2[IDX(gc, i, j, 0)] = bc[IDX(g, ir, jr, 0)];
        if (kr == M1) bc2[IDX(gc, i, j, m / 2 + 2)] = bc[IDX(g, ir, jr, m + 2)];
        double re = 0.0;
        if (bc2[IDX(gc, i, j, k)] <= 0) {
          double r000 = 0, rpp0 = 0, rp00 = 0, rpm0 = 0, r0m0 = 0, rmm0 = 0, rm00 = 0, rmp0 = 0,
                 r0p0 = 0, r00p = 0, rppp = 0, rp0p = 0, rpmp = 0, r0mp = 0, rmmp = 0, rm0p = 0,
                 rmpp = 0, r0pp = 0, r00m = 0, rppm = 0, rp0m = 0, rpmm = 0, r0mm = 0, rmmm = 0,
                 rm0m = 0, rmpm = 0, r0pm = 0;
          const int kb = (kr == 1); /* (rank == 0 .or. flag == 0) is true on one rank */
#define B(a, b_, c) bc[IDX(g, (a), (b_), (c))]
#define GR(a, b_, c) getres(F, (a), (b_), (c), flag_nmn)
          r000 = GR(ir, jr, kr);
          if (ir != N1 && jr != N2) { if (B(ir + 1, jr + 1, kr) <= 0) rpp0 = GR(ir + 1, jr + 1, kr); }
          if (ir != N1) { if (B(ir + 1, jr, kr) <= 0) rp00 = GR(ir + 1, jr, kr); }
          if (B(ir + 1, jr - 1, kr) <= 0) {
            if (ir != N1 && jr != 1) rpm0 = GR(ir + 1, jr - 1, kr);
            else if (jr == 1) rpm0 = GR(ir + 1, n[1], kr);
          }
          if (B(ir, jr - 1, kr) <= 0) {
            if (jr == 1) r0m0 = GR(ir, n[1], kr);
            else r0m0 = GR(ir, jr - 1, kr);
          }
          if (B(ir - 1, jr - 1, kr) <= 0 && ir != 1) {
            if (jr != 1) rmm0 = GR(ir - 1, jr - 1, kr);
            else rmm0 = GR(ir - 1, n[1], kr);
          }
          if (ir != 1) { if (B(ir - 1, jr, kr) <= 0) rm00 = GR(ir - 1, jr, kr); }
          if (ir != 1 && jr != N2) { if (B(ir - 1, jr + 1, kr) <= 0) rmp0 = GR(ir - 1, jr + 1, kr); }
          if (jr != N2) { if (B(ir, jr + 1, kr) <= 0) r0p0 = GR(ir, jr + 1, kr); }
          if (kr != M1) { if (B(ir, jr, kr + 1) <= 0) r00p = GR(ir, jr, kr + 1); }
          if (ir != N1 && jr != N2 && kr != M1) {
            if (B(ir + 1, jr + 1, kr + 1) <= 0) rppp = GR(ir + 1, jr + 1, kr + 1);
          }
          if (ir != N1 && kr != M1) { if (B(ir + 1, jr, kr + 1) <= 0) rp0p = GR(ir + 1, jr, kr + 1); }
          if (B(ir + 1, jr - 1, kr + 1) <= 0) {
            if (ir != N1 && jr != 1 && kr != M1) rpmp = GR(ir + 1, jr - 1, kr + 1);
            else if (jr == 1 && kr != M1) rpmp = GR(ir + 1, n[1], kr + 1);
          }
          if (B(ir, jr - 1, kr + 1) <= 0) {
            if (jr != 1 && kr != M1) r0mp = GR(ir, jr - 1, kr + 1);
            else if (jr == 1 && kr != M1) r0mp = GR(ir, n[1], kr + 1);
          }
          if (B(ir - 1, jr - 1, kr + 1) <= 0 && ir != 1) {
            if (jr != 1 && kr != M1) rmmp = GR(ir - 1, jr - 1, kr + 1);
            else if (jr == 1 && kr != M1) rmmp = GR(ir - 1, n[1], kr + 1);
          }
          if (ir != 1 && kr != M1) { if (B(ir - 1, jr, kr + 1) <= 0) rm0p = GR(ir - 1, jr, kr + 1); }
          if (ir != 1 && jr != N2 && kr != M1) {
            if (B(ir - 1, jr + 1, kr + 1) <= 0) rmpp = GR(ir - 1, jr + 1, kr + 1);
          }
          if (jr != N2 && kr != M1) { if (B(ir, jr + 1, kr + 1) <= 0) r0pp = GR(ir, jr + 1, kr + 1); }
          if (B(ir, jr, kr - 1) <= 0) {
            if (kb) r00m = RP(ir, jr);
            else r00m = GR(ir, jr, kr - 1);
          }
          if (B(ir + 1, jr + 1, kr - 1) <= 0) {
            if (kb) { if (ir != N1 && jr != N2) rppm = RP(ir + 1, jr + 1); }
            else { if (ir != N1 && jr != N2) rppm = GR(ir + 1, jr + 1, kr - 1); }
          }
          if (B(ir + 1, jr, kr - 1) <= 0) {
            if (kb) { if (ir != N1) rp0m = RP(ir + 1, jr); }
            else { if (ir != N1) rp0m = GR(ir + 1, jr, kr - 1); }
          }
          if (B(ir + 1, jr - 1, kr - 1) <= 0) {
            if (kb) { if (ir != N1 && jr != 1) rpmm = RP(ir + 1, jr - 1); }
            else {
              if (ir != N1 && jr != 1) rpmm = GR(ir + 1, jr - 1, kr - 1);
              if (jr == 1) rpmm = GR(ir + 1, n[1], kr - 1);
            }
          }
          if (B(ir, jr - 1, kr - 1) <= 0) {
            if (kb) { if (jr != 1) r0mm = RP(ir, jr - 1); }
            else {
              if (jr != 1) r0mm = GR(ir, jr - 1, kr - 1);
              if (jr == 1) r0mm = GR(ir, n[1], kr - 1);
            }
          }
          if (B(ir - 1, jr - 1, kr - 1) <= 0 && ir != 1) {
            if (kb) { if (jr != 1) rmmm = RP(ir - 1, jr - 1); }
            else {
              if (jr != 1) rmmm = GR(ir - 1, jr - 1, kr - 1);
              if (jr == 1) rmmm = GR(ir - 1, n[1], kr - 1);
            }
          }
          if (B(ir - 1, jr, kr - 1) <= 0 && ir != 1) {
            if (kb) rm0m = RP(ir - 1, jr);
            else rm0m = GR(ir - 1, jr, kr - 1);
          }
          if (B(ir - 1, jr + 1, kr - 1) <= 0 && ir != 1) {
            if (kb) { if (jr != N2) rmpm = RP(ir - 1, jr + 1); }
            else { if (jr != N2) rmpm = GR(ir - 1, jr + 1, kr - 1); }
          }
          if (B(ir, jr + 1, kr - 1) <= 0) {
            if (kb) { if (jr != N2) r0pm = RP(ir, jr + 1); }
            else { if (jr != N2) r0pm = GR(ir, jr + 1, kr - 1); }
          }
#undef B
#undef GR
          re = (8.0 * r000 + 4.0 * (rp00 + rm00 + r0p0 + r0m0 + r00p + r00m) +
                2.0 * (rpp0 + rmp0 + rpm0 + rmm0 + rp0p + rm0p + r0pp + r0mp + rp0m + rm0m + r0pm +
                       r0mm) +
                rppp + rmpp + rpmp + rmmp + rppm + rmpm + rpmm + rmmm) /
               64.0;
        }
        C->b[IDX(gc, i, j, k)] = re;
      }
#undef RP
  free(r_pbc);
}

/* prolongation (mg.f90:850-940). F is the finer level receiving the correction. */
static void prolongation(mg_level *F, const mg_level *C) {
  const lo_grid *g = &F->g;
  const lo_grid *gc = &C->g;
  const int32_t *n = g->n;
  const int kp = n[2] / 2 + 1;
  double *u = F->u;
  const double *e = C->u;
  const int32_t *bc = F->bcnd;
  LO_PARALLEL_FOR
  for (int k = 1; k <= kp; ++k)
    for (int j = 1; j <= n[1] / 2 + 1; ++j)
      for (int i = 1; i <= n[0] / 2 + 1; ++i) {
        double r000 = e[IDX(gc, i, j, k)], r0m0 = e[IDX(gc, i, j - 1, k)],
               rmm0 = e[IDX(gc, i - 1, j - 1, k)], rm00 = e[IDX(gc, i - 1, j, k)],
               r00m = e[IDX(gc, i, j, k - 1)], r0mm = e[IDX(gc, i, j - 1, k - 1)],
               rmmm = e[IDX(gc, i - 1, j - 1, k - 1)], rm0m = e[IDX(gc, i - 1, j, k - 1)];
        int ir = 2 * i - 1, jr = 2 * j - 1, kr = 2 * k - 1;
        int64_t c;
        c = IDX(g, ir, jr, kr);
        if (bc[c] <= 0) u[c] = u[c] + r000;
        c = IDX(g, ir, jr - 1, kr);
        if (bc[c] <= 0) u[c] = u[c] + (r000 + r0m0) / 2.0;
        c = IDX(g, ir - 1, jr - 1, kr);
        if (bc[c] <= 0) u[c] = u[c] + (r000 + r0m0 + rmm0 + rm00) / 4.0;
        c = IDX(g, ir - 1, jr, kr);
        if (bc[c] <= 0) u[c] = u[c] + (r000 + rm00) / 2.0;
        c = IDX(g, ir, jr, kr - 1);
        if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m) / 2.0;
        c = IDX(g, ir, jr - 1, kr - 1);
        if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + r0mm + r0m0) / 4.0;
        c = IDX(g, ir - 1, jr, kr - 1);
        if (bc[c] <= 0) u[c] = u[c] + (r000 + r00m + rm0m + rm00) / 4.0;
        c = IDX(g, ir - 1, jr - 1, kr - 1);
        if (bc[c] <= 0)
          u[c] = u[c] + (r000 + r00m + rm0m + rm00 + r0m0 + r0mm + rmmm + rmm0) / 8.0;
      }
}

/* The multigrid's pieces one at a time, on caller-owned arrays, so that the
 * tests can pin each against what the mathematics fixes: one colour of the
 * SOR, the periodic ghost copy, sor_rb with its exit rules, the L1
 * residual, one restriction and one prolongation. */
static mg_level hook_level(const lo_grid *g, double *u, const double *b, const int32_t *bcnd,
                           double eps0) {
  mg_level L;
  memset(&L, 0, sizeof(L));
  L.g = *g;
  L.u = u;
  L.b = (double *)b;
  L.bcnd = (int32_t *)bcnd;
  coeffs(&L, eps0);
  return L;
}

void lo_mg_sor_colour(const lo_grid *g, double *u, const double *b, const int32_t *bcnd,
                      double eps0, int colour, double omega) {
  mg_level L = hook_level(g, u, b, bcnd, eps0);
  sor_colour(&L, colour, omega);
}

void lo_mg_periodic(const lo_grid *g, double *u, const int32_t *bcnd) {
  mg_level L = hook_level(g, u, NULL, bcnd, 1.0);
  periodic_u(&L);
}

double lo_mg_residual(const lo_grid *g, const double *u, const double *b, const int32_t *bcnd,
                      double eps0) {
  mg_level L = hook_level(g, (double *)u, b, bcnd, eps0);
  return sor_residual(&L);
}

int lo_mg_sor_rb(const lo_grid *g, double *u, const double *b, const int32_t *bcnd, double eps0,
                 double omega, double eps, int kmg, int dig, int flag_pbc, double *res_out) {
  mg_level L = hook_level(g, u, b, bcnd, eps0);
  return sor_rb(&L, res_out, omega, eps, kmg, dig, flag_pbc);
}

void lo_mg_restrict(const lo_grid *gf, const double *uf, const double *bf, const int32_t *bcf,
                    double eps0, int flag_pbc, int flag_nmn, double *b_c, int32_t *bc_c) {
  mg_level F = hook_level(gf, (double *)uf, bf, bcf, eps0);
  lo_grid gc;
  for (int d = 0; d < 3; ++d) {
    gc.n[d] = gf->n[d] / 2;
    gc.h[d] = gf->h[d] * 2.0;
  }
  double *uc = (double *)calloc((size_t)NODES(&gc), sizeof(double));
  mg_level C = hook_level(&gc, uc, b_c, bc_c, eps0);
  restriction(&F, &C, flag_pbc, flag_nmn);
  free(uc);
}

void lo_mg_prolong(const lo_grid *gf, double *uf, const int32_t *bcf, const double *uc) {
  mg_level F = hook_level(gf, uf, NULL, bcf, 1.0);
  lo_grid gc;
  for (int d = 0; d < 3; ++d) {
    gc.n[d] = gf->n[d] / 2;
    gc.h[d] = gf->h[d] * 2.0;
  }
  mg_level C = hook_level(&gc, (double *)uc, NULL, NULL, 1.0);
  prolongation(&F, &C);
}

/* pdesolver (pdesolver.f90:1-135) + mg V-cycle (mg.f90:1-246). */
int lo_pdesolver(const lo_grid *g, double *phi, const double *rho, const int32_t *bcnd,
                 int flag_pbc, int flag_nmn, double eps0, int32_t ncycl, double eps, double omega,
                 int32_t ng, lo_mg_stats *stats) {
  if (ng < 1 || ng > 12) return LO_ERR_ARG;
  mg_level L[12];
  memset(L, 0, sizeof(L));
  for (int l = 0; l < ng; ++l) {
    int32_t div = 1 << l;
    for (int d = 0; d < 3; ++d) {
      L[l].g.n[d] = g->n[d] / div;
      L[l].g.h[d] = g->h[d] * (double)div;
    }
    if (l < ng - 1)
      for (int d = 0; d < 3; ++d)
        if (L[l].g.n[d] < 2 || (L[l].g.n[d] % 2) != 0) return LO_ERR_ARG;
    coeffs(&L[l], eps0);
  }
  L[0].u = phi;
  L[0].b = (double *)rho;
  L[0].bcnd = (int32_t *)bcnd;
  for (int l = 1; l < ng; ++l) {
    int64_t nn = NODES(&L[l].g);
    L[l].u = (double *)calloc((size_t)nn, sizeof(double));
    L[l].b = (double *)calloc((size_t)nn, sizeof(double));
    L[l].bcnd = (int32_t *)calloc((size_t)nn, sizeof(int32_t));
  }
  /* alpha = |(n1+1)(n2+1)(n3+1) ac| * 1V with the fine-grid ac (pdesolver.f90:33-34) */
  double alpha = fabs((double)((g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1)) * L[0].ac) * 1.0;
  double epsa = alpha * eps;
  double res = 0.0, res_s = 0.0, ktot = 0.0;
  int status = LO_OK;
  int k = 1;
  if (stats) memset(stats, 0, sizeof(*stats));
  for (;;) {
    int64_t iter[12] = {0};
    for (int ig = 1; ig <= 2 * ng - 1; ++ig) {
      for (int il = 1; il <= ng - 1 && il <= 11; ++il) {
        if ((ig < ng && ig == il) || (ig > ng && ig == 2 * ng - il)) {
          iter[il - 1] += sor_rb(&L[il - 1], &res, omega, epsa, k, ig - ng, flag_pbc);
          if (ig < ng) restriction(&L[il - 1], &L[il], flag_pbc, flag_nmn);
          if (ig > ng && il >= 2) prolongation(&L[il - 2], &L[il - 1]);
        }
      }
      if (ig == ng) {
        iter[ng - 1] += sor_rb(&L[ng - 1], &res, omega, epsa, k, 0, flag_pbc);
        if (ng >= 2) prolongation(&L[ng - 2], &L[ng - 1]);
      }
    }
    for (int i = 0; i < ng; ++i) {
      ktot = ktot + (double)iter[i] / pow(2.0, 2.0 * i);
      if (stats) stats->sweeps[i] += iter[i];
    }
    res = res / alpha;
    if (k == 1) res_s = res;
    else if (fabs(res / res_s) > 10.0) { status = LO_ERR_DIVERGED; break; }
    if (res < eps || k == ncycl) break;
    k = k + 1;
  }
  for (int l = 1; l < ng; ++l) { free(L[l].u); free(L[l].b); free(L[l].bcnd); }
  if (stats) { stats->cycles = k; stats->ktot = ktot; stats->res = res; }
  return status;
}

/* ===================================================================== */
/* E field: modules/fields/mod_electricField.f90:9-164                    */
/* ===================================================================== */

void lo_efield(const lo_grid *g, const double *phi, const int32_t *bcnd, double *E4) {
  const int32_t *n = g->n;
  const double *h = g->h;
  memset(E4, 0, sizeof(double) * 4 * (size_t)NODES(g));
#define P(i, j, k) phi[IDX(g, i, j, k)]
#define E(c, i, j, k) E4[4 * IDX(g, i, j, k) + (c)]
#define BC(i, j, k) bcnd[IDX(g, i, j, k)]
  /* second neighbours outside the array read as 0 (the reference indexes
   * past its bounds there) */
#define E2(c, i, j, k) \
  ((IDX(g, i, j, k) >= 0 && IDX(g, i, j, k) < NODES(g)) ? E4[4 * IDX(g, i, j, k) + (c)] : 0.0)
  for (int iz = 1; iz <= n[2] + 1; ++iz)
    for (int iy = 1; iy <= n[1] + 1; ++iy)
      for (int ix = 1; ix <= n[0] + 1; ++ix)
        if (BC(ix, iy, iz) <= 0) {
          E(0, ix, iy, iz) = -(P(ix + 1, iy, iz) - P(ix - 1, iy, iz)) / (2.0 * h[0]);
          E(1, ix, iy, iz) = -(P(ix, iy + 1, iz) - P(ix, iy - 1, iz)) / (2.0 * h[1]);
          E(2, ix, iy, iz) = -(P(ix, iy, iz + 1) - P(ix, iy, iz - 1)) / (2.0 * h[2]);
        }
  for (int iz = 1; iz <= n[2] + 1; ++iz)
    for (int iy = 1; iy <= n[1] + 1; ++iy)
      for (int ix = 1; ix <= n[0] + 1; ++ix) {
        int32_t b = BC(ix, iy, iz);
        if (b == -1) continue;
        if (b == -2) {
          E(0, 1, iy, iz) = 0.0;
          if (iy > 1 && iy < n[1] + 1 && iz > 1 && iz < n[2] + 1) {
            E(1, 1, iy, iz) = -(P(1, iy + 1, iz) - P(1, iy - 1, iz)) / (2.0 * h[1]);
            E(2, 1, iy, iz) = -(P(1, iy, iz + 1) - P(1, iy, iz - 1)) / (2.0 * h[2]);
            continue;
          }
        }
        if (b == 0 || b == -2) {
          if (iy == 1) {
            E(1, ix, 1, iz) = -(P(ix, 2, iz) - P(ix, n[1], iz)) / (2.0 * h[1]);
            E(1, ix, 0, iz) = E(1, ix, n[1], iz);
            E(1, ix, n[1] + 1, iz) = E(1, ix, 1, iz);
            E(1, ix, n[1] + 2, iz) = E(1, ix, 2, iz);
            E(0, ix, 0, iz) = E(0, ix, n[1], iz);
            E(0, ix, n[1] + 2, iz) = E(0, ix, 2, iz);
            E(2, ix, 0, iz) = E(2, ix, n[1], iz);
            E(2, ix, n[1] + 2, iz) = E(2, ix, 2, iz);
          }
          if (iz == 1) {
            E(2, ix, iy, 1) = -(P(ix, iy, 2) - P(ix, iy, n[2])) / (2.0 * h[2]);
            E(2, ix, iy, 0) = E(2, ix, iy, n[2]);
            E(2, ix, iy, n[2] + 1) = E(2, ix, iy, 1);
            E(2, ix, iy, n[2] + 2) = E(2, ix, iy, 2);
            E(0, ix, iy, 0) = E(0, ix, iy, n[2]);
            E(0, ix, iy, n[2] + 2) = E(0, ix, iy, 2);
            E(1, ix, iy, 0) = E(1, ix, iy, n[2]);
            E(1, ix, iy, n[2] + 2) = E(1, ix, iy, 2);
          }
          continue;
        }
        if (b >= 1) {
          if (BC(ix + 1, iy, iz) <= 0) {
            E(0, ix, iy, iz) = 2.0 * E(0, ix + 1, iy, iz) - E2(0, ix + 2, iy, iz);
            E(1, ix, iy, iz) = -(P(ix, iy + 1, iz) - P(ix, iy - 1, iz)) / (2.0 * h[1]);
            E(2, ix, iy, iz) = -(P(ix, iy, iz + 1) - P(ix, iy, iz - 1)) / (2.0 * h[2]);
          }
          if (BC(ix - 1, iy, iz) <= 0) {
            E(0, ix, iy, iz) = 2.0 * E(0, ix - 1, iy, iz) - E2(0, ix - 2, iy, iz);
            E(1, ix, iy, iz) = -(P(ix, iy + 1, iz) - P(ix, iy - 1, iz)) / (2.0 * h[1]);
            E(2, ix, iy, iz) = -(P(ix, iy, iz + 1) - P(ix, iy, iz - 1)) / (2.0 * h[2]);
          }
          if (BC(ix, iy + 1, iz) <= 0) {
            E(0, ix, iy, iz) = -(P(ix + 1, iy, iz) - P(ix - 1, iy, iz)) / (2.0 * h[0]);
            E(1, ix, iy, iz) = 2.0 * E(1, ix, iy + 1, iz) - E2(1, ix, iy + 2, iz);
            E(2, ix, iy, iz) = -(P(ix, iy, iz + 1) - P(ix, iy, iz - 1)) / (2.0 * h[2]);
          }
          if (BC(ix, iy - 1, iz) <= 0) {
            E(0, ix, iy, iz) = -(P(ix + 1, iy, iz) - P(ix - 1, iy, iz)) / (2.0 * h[0]);
            E(1, ix, iy, iz) = 2.0 * E(1, ix, iy - 1, iz) - E2(1, ix, iy - 2, iz);
            E(2, ix, iy, iz) = -(P(ix, iy, iz + 1) - P(ix, iy, iz - 1)) / (2.0 * h[2]);
          }
          if (BC(ix, iy, iz + 1) <= 0) {
            E(0, ix, iy, iz) = -(P(ix + 1, iy, iz) - P(ix - 1, iy, iz)) / (2.0 * h[0]);
            E(1, ix, iy, iz) = -(P(ix, iy + 1, iz) - P(ix, iy - 1, iz)) / (2.0 * h[1]);
            E(2, ix, iy, iz) = 2.0 * E(2, ix, iy, iz + 1) - E2(2, ix, iy, iz + 2);
          }
          if (BC(ix, iy, iz - 1) <= 0) {
            E(0, ix, iy, iz) = -(P(ix + 1, iy, iz) - P(ix - 1, iy, iz)) / (2.0 * h[0]);
            E(1, ix, iy, iz) = -(P(ix, iy + 1, iz) - P(ix, iy - 1, iz)) / (2.0 * h[1]);
            E(2, ix, iy, iz) = 2.0 * E(2, ix, iy, iz - 1) - E2(2, ix, iy, iz - 2);
          }
        }
      }
#undef P
#undef E
#undef BC
#undef E2
}

/* ===================================================================== */
/* MCC: modules/particles/mod_collisions.f90:64-776 (nproc = 1)          */
/* ===================================================================== */

static void scatter_vector(double *vx1, double *vy1, double *vz1, double vx, double vy,
                           double vz, double costheta, double phi) {
  /* mod_collisions.f90:64-92 */
  double s = 1.0 - costheta * costheta;
  double sintheta = sqrt(s > 0.0 ? s : 0.0);
  double sinphi = sin(phi), cosphi = cos(phi);
  double v = sqrt(vx * vx + vy * vy + vz * vz);
  *vx1 = vx; *vy1 = vy; *vz1 = vz;
  if (v == 0.0) return;
  if (fabs(vy) > fabs(vz)) {
    double vv = sqrt(vx * vx + vy * vy);
    *vx1 = vx * costheta + (vy * v * sinphi + vx * vz * cosphi) / vv * sintheta;
    *vy1 = vy * costheta + (-vx * v * sinphi + vy * vz * cosphi) / vv * sintheta;
    *vz1 = vz * costheta - vv * cosphi * sintheta;
  } else {
    double vv = sqrt(vx * vx + vz * vz);
    *vx1 = vx * costheta + (vz * v * sinphi - vy * vx * cosphi) / vv * sintheta;
    *vy1 = vy * costheta + vv * cosphi * sintheta;
    *vz1 = vz * costheta - (vx * v * sinphi + vy * vz * cosphi) / vv * sintheta;
  }
}

static void box_muller(double *v1, double *v2, double vt, double r1_in, double r2_in) {
  /* mod_collisions.f90:742-753 */
  double r1 = r1_in > 1.0e-14 ? r1_in : 1.0e-14;
  double fac = vt * sqrt(-log(1.0 - r1));
  *v1 = fac * cos(2.0 * PI_F * r2_in);
  *v2 = fac * sin(2.0 * PI_F * r2_in);
}

/* First bracket Er[ip-1] < E <= Er[ip] on the common energy grid.  The
 * reference's 5-step bisection + linear scan over a zero-padded table
 * (mod_collisions.f90:512-546) returns this same unique bracket for a strictly
 * increasing grid.  Returns ip (1..npt-1) or -1. */
static int find_bracket(const double *Er, int npt, double E) {
  if (npt < 2) return -1;
  if (!(E > Er[0]) || E > Er[npt - 1]) return -1;
  int lo = 0, hi = npt - 1; /* Er[lo] < E <= Er[hi] */
  while (hi - lo > 1) {
    int mid = lo + (hi - lo) / 2;
    if (E <= Er[mid]) hi = mid;
    else lo = mid;
  }
  return hi;
}

/* Optional log of accepted events (tests only): one LO_MCC_REC-double record
 * per accepted event, written while a log is set. */
static double *mcc_log = NULL;
static int64_t mcc_log_cap = 0, mcc_log_n = 0;

void lo_mcc_set_log(double *buf, int64_t cap) {
  mcc_log = buf;
  mcc_log_cap = buf ? cap : 0;
  mcc_log_n = 0;
}

int64_t lo_mcc_log_count(void) { return mcc_log_n; }

static double *mcc_log_rec(void) {
  if (!mcc_log || mcc_log_n >= mcc_log_cap) return NULL;
  double *r = mcc_log + mcc_log_n * LO_MCC_REC;
  for (int q = 0; q < LO_MCC_REC; ++q) r[q] = -1.0;
  ++mcc_log_n;
  return r;
}

int lo_collisions(const lo_grid *g, lo_pset *parts, const double *np_red, const lo_chem *c,
                  int32_t ns_coll, double dt, int32_t *iseed, double *Pcoll,
                  lo_mcc_stats *stats) {
  const int32_t *n = g->n;
  const double *h = g->h;
  const int NT = c->ntype_tracked, NA = c->ntype_total;
  const int64_t nodes = NODES(g);
  if (NT > 16 || NA > 16) return LO_ERR_ARG;
  double np_mx_all[16];
  int32_t Nc[16];
  memset(Nc, 0, sizeof(Nc));
  if (stats) memset(stats, 0, sizeof(*stats));
  /* np_mx: the reference reads np_red through an assumed-shape dummy whose
   * lower bound is 1, so maxval(np_red(1:n+1,...)) covers stored nodes 0..n
   * (mod_collisions.f90:222,293). */
  for (int t = 0; t < NT; ++t) {
    double mx = -INFINITY;
    for (int k = 0; k <= n[2]; ++k)
      for (int j = 0; j <= n[1]; ++j)
        for (int i = 0; i <= n[0]; ++i) {
          double v = np_red[(int64_t)t * nodes + IDX(g, i, j, k)];
          if (v > mx) mx = v;
        }
    np_mx_all[t] = mx;
  }
  for (int t = NT; t < NA; ++t) np_mx_all[t] = c->neutral_density[t - NT];

  /* nu_max and Nc (mod_collisions.f90:317-372) */
  for (int t = 0; t < NT; ++t) {
    if (c->p_ncol[t] == 0) continue;
    if (c->mass[t] <= 0.0) continue;
    double nu_max = 0.0;
    int ncl = c->p_ncol[t] < LO_MAX_COL ? c->p_ncol[t] : LO_MAX_COL;
    for (int ic = 0; ic < ncl; ++ic) {
      int col = c->sig_list[t * LO_MAX_COL + ic];
      int n_re = c->col_info[col * 10 + 0];
      int tt = c->col_info[col * 10 + 1 + n_re];
      if (tt < 0 || tt >= NA) continue;
      nu_max = nu_max + np_mx_all[tt] * c->sigv_mx[t * c->ncol + col];
    }
    if (nu_max > c->nu_uplim[t]) nu_max = c->nu_uplim[t];
    double Pmax = nu_max * (double)ns_coll * dt;
    if (Pmax > 1.0) return LO_ERR_PMAX;
    double dNc = (double)parts[t].n * Pmax / 1.0;
    int32_t nct = (int32_t)dNc;
    double rnd = lo_ran2(&iseed[0]);
    if (rnd <= (dNc - (double)nct)) nct = nct + 1;
    Nc[t] = nct;
  }
  if (stats) for (int t = 0; t < NT; ++t) stats->Nc[t] = Nc[t];

  /* trials (mod_collisions.f90:391-728), single virtual thread */
  for (int pt = 0; pt < NT; ++pt) {
    if (c->p_ncol[pt] == 0 || c->mass[pt] <= 0.0) continue;
    if (Nc[pt] == 0) continue;
    lo_pset *P = &parts[pt];
    if (P->n == 0) continue;
    double nu_max_local = (double)Nc[pt] / (double)(P->n > 1 ? P->n : 1) / ((double)ns_coll * dt);
    if (nu_max_local <= 0.0) continue;
    const int ncl = c->p_ncol[pt] < LO_MAX_COL ? c->p_ncol[pt] : LO_MAX_COL;
    for (int32_t ic = 0; ic < Nc[pt]; ++ic) {
      double rnd = lo_ran2(&iseed[0]);
      int64_t ip = (int64_t)((double)P->n * rnd) + 1;
      if (ip > P->n) ip = P->n;
      ip -= 1; /* 0-based */
      if (P->dead[ip] != 0) { if (stats) stats->skipped_dead[pt]++; continue; }
      int32_t ix = (int32_t)(P->x[ip] / h[0]) + 1;
      int32_t iy = (int32_t)(P->y[ip] / h[1]) + 1;
      int32_t iz = (int32_t)(P->z[ip] / h[2]) + 1;
      ix = imax32(1, imin32(n[0], ix));
      iy = imax32(1, imin32(n[1], iy));
      iz = imax32(1, imin32(n[2], iz));
      double vx1 = P->vx[ip], vy1 = P->vy[ip], vz1 = P->vz[ip];
      double nu[LO_MAX_COL], sum_nu = 0.0;
      int tsel[16] = {0}, tfound[16] = {0};
      double svr[16], sEkr[16], smu[16], scx[16], scy[16], scz[16], stv[16][3];
      for (int ic2 = 0; ic2 < ncl; ++ic2) nu[ic2] = 0.0;
      for (int ic2 = 0; ic2 < ncl; ++ic2) {
        int col = c->sig_list[pt * LO_MAX_COL + ic2];
        int n_re = c->col_info[col * 10 + 0];
        int tt = c->col_info[col * 10 + 1 + n_re];
        if (tt < 0 || tt >= NA) continue;
        if (!tsel[tt]) {
          double vx2, vy2, vz2;
          if (tt >= NT) {
            double r1 = lo_ran2(&iseed[0]), r2 = lo_ran2(&iseed[0]), dummy;
            box_muller(&vx2, &vy2, c->vt0[tt], r1, r2);
            r1 = lo_ran2(&iseed[0]); r2 = lo_ran2(&iseed[0]);
            box_muller(&vz2, &dummy, c->vt0[tt], r1, r2);
          } else {
            /* Charged target: the cell tables are zeroed by the mover and the
             * BC before MCC runs (mod_particleMover.f90:153-155,
             * mod_particleBC.f90:255-257), so the pick always fails. */
            if (stats) stats->err_coll[pt]++;
            tsel[tt] = 1; tfound[tt] = 0;
            continue;
          }

          double mp = fabs(c->mass[pt]), mt = fabs(c->mass[tt]);
          smu[tt] = mp * mt / (mp + mt);
          double dx = vx1 - vx2, dy = vy1 - vy2, dz = vz1 - vz2;
          svr[tt] = sqrt(dx * dx + dy * dy + dz * dz);
          sEkr[tt] = 0.5 * smu[tt] * (svr[tt] * svr[tt]) / QE;
          scx[tt] = (mp * vx1 + mt * vx2) / (mp + mt);
          scy[tt] = (mp * vy1 + mt * vy2) / (mp + mt);
          scz[tt] = (mp * vz1 + mt * vz2) / (mp + mt);
          stv[tt][0] = vx2; stv[tt][1] = vy2; stv[tt][2] = vz2;
          tsel[tt] = 1; tfound[tt] = 1;
        }
        if (!tfound[tt]) continue;
        double vr = svr[tt], Ekr = sEkr[tt];
        int ipt = find_bracket(c->sig_Er, c->npt, Ekr);
        if (ipt > 0) {
          double EkL = c->sig_Er[ipt - 1], EkR = c->sig_Er[ipt];
          double sL = c->sig[(int64_t)col * c->npt + ipt - 1], sR = c->sig[(int64_t)col * c->npt + ipt];
          double sig_p = sL + (Ekr - EkL) * (sR - sL) / (EkR - EkL);
          double np_t;
          if (tt < NT) {
            /* density_average_at_particle_cell on the 1-based dummy: stored nodes ix-1..ix */
            const double *d = np_red + (int64_t)tt * nodes;
            np_t = 0.125 * (d[IDX(g, ix - 1, iy - 1, iz - 1)] + d[IDX(g, ix, iy - 1, iz - 1)] +
                            d[IDX(g, ix - 1, iy, iz - 1)] + d[IDX(g, ix, iy, iz - 1)] +
                            d[IDX(g, ix - 1, iy - 1, iz)] + d[IDX(g, ix, iy - 1, iz)] +
                            d[IDX(g, ix - 1, iy, iz)] + d[IDX(g, ix, iy, iz)]);
          } else {
            np_t = np_mx_all[tt];
          }
          nu[ic2] = np_t * sig_p * vr / nu_max_local;
          sum_nu = sum_nu + nu[ic2];
        }
      }
      /* indexx_small: ascending exchange sort (mod_collisions.f90:756-776) */
      int indx[LO_MAX_COL];
      for (int a = 0; a < ncl; ++a) indx[a] = a;
      for (int a = 0; a < ncl - 1; ++a)
        for (int b2 = a + 1; b2 < ncl; ++b2)
          if (nu[indx[b2]] < nu[indx[a]]) { int tmp = indx[a]; indx[a] = indx[b2]; indx[b2] = tmp; }
      rnd = lo_ran2(&iseed[0]);
      int chosen = -1, c_ind = -1;
      if (rnd <= sum_nu) {
        double s = 0.0;
        for (int a = 0; a < ncl; ++a) {
          s = s + nu[indx[a]];
          if (rnd <= s) { chosen = indx[a]; c_ind = c->sig_list[pt * LO_MAX_COL + chosen]; break; }
        }
      }
      if (chosen < 0) continue;
      if (stats) stats->accepted[pt]++;
      int n_re = c->col_info[c_ind * 10 + 0];
      int n_by = c->col_info[c_ind * 10 + 1];
      int tt = c->col_info[c_ind * 10 + 1 + n_re];
      double *rec = mcc_log_rec();
      if (rec) {
        rec[0] = pt; rec[1] = (double)ip; rec[2] = c_ind; rec[3] = tt;
        rec[4] = vx1; rec[5] = vy1; rec[6] = vz1;
        rec[11] = 0.0;   /* executed */
      }
      if (tt < 0 || tt >= NA) continue;
      if (!tfound[tt]) continue;
      if (rec) {
        rec[7] = stv[tt][0]; rec[8] = stv[tt][1]; rec[9] = stv[tt][2];
        rec[10] = n_by;
      }
      double mu = smu[tt], vr = svr[tt];
      double vxc = scx[tt], vyc = scy[tt], vzc = scz[tt];
      double Eth = c->sig_Eth[c_ind];
      double Ee = 0.5 * mu * vr * vr - Eth * QE;
      if (Ee < 0.0) continue;
      double smi = 0.0;
      for (int ib = 0; ib < n_by; ++ib) {
        int bt = c->col_info[c_ind * 10 + 2 + n_re + ib];
        if (bt >= 0 && bt < NA && c->mass[bt] > 0.0) smi = smi + 1.0 / fabs(c->mass[bt]);
      }
      if (smi <= 0.0) continue;
      if (stats) stats->executed[pt]++;
      if (rec) rec[11] = 1.0;
      double costh = 1.0 - 2.0 * lo_ran2(&iseed[0]);
      double phi = 2.0 * PI_F * lo_ran2(&iseed[0]);
      double costh_s = 1.0 - 2.0 * lo_ran2(&iseed[0]);
      double phi_s = 2.0 * PI_F * lo_ran2(&iseed[0]);
      double th_add = 2.0 * PI_F / (double)(n_by > 1 ? n_by : 1);
      double x0 = P->x[ip], y0 = P->y[ip], z0 = P->z[ip];
      int flag_add[10] = {0};
      for (int a = 0; a < n_re; ++a) flag_add[a] = -1;
      for (int ib = 0; ib < n_by; ++ib) {
        int bt = c->col_info[c_ind * 10 + 2 + n_re + ib];
        if (bt < 0 || bt >= NT) continue;
        if (c->mass[bt] <= 0.0) continue;
        double s2 = 1.0 - costh * costh;
        double st = sqrt(0.0 > s2 ? 0.0 : s2);
        double ex = costh;
        double ey = st * sin(phi + th_add * (double)ib);
        double ez = st * cos(phi + th_add * (double)ib);
        double ex1, ey1, ez1;
        scatter_vector(&ex1, &ey1, &ez1, ex, ey, ez, costh_s, phi_s);
        double mb = fabs(c->mass[bt]);
        double vp = sqrt(2.0 * (Ee / smi) / (mb * mb));
        int64_t ib_idx;
        if (ib == 0 && bt == pt) {
          ib_idx = ip;
          flag_add[0] = 0;
        } else if (bt == tt && tt < NT) {
          ib_idx = -1; /* saved_target_ip is never set for a failed charged pick */
          flag_add[1] = 0;
        } else {
          ib_idx = parts[bt].n;
          flag_add[n_re + ib] = 1;
        }
        if (ib_idx < 0) continue;
        lo_pset *B = &parts[bt];
        double v2old = 0.0;
        if (ib_idx < B->n) v2old = B->vx[ib_idx] * B->vx[ib_idx] + B->vy[ib_idx] * B->vy[ib_idx] +
                                   B->vz[ib_idx] * B->vz[ib_idx];
        if (ib_idx >= B->cap) return LO_ERR_CAPACITY;
        if (stats && ib_idx >= B->n) stats->appended[bt]++;
        if (ib_idx >= B->n) B->n = ib_idx + 1;
        B->x[ib_idx] = x0; B->y[ib_idx] = y0; B->z[ib_idx] = z0;
        B->vx[ib_idx] = vxc + vp * ex1;
        B->vy[ib_idx] = vyc + vp * ey1;
        B->vz[ib_idx] = vzc + vp * ez1;
        B->dead[ib_idx] = 0;
        B->cell[ib_idx] = 0;
        if (rec && ib < 3) {
          double *o = rec + 12 + 5 * ib;
          o[0] = bt; o[1] = (double)ib_idx;
          o[2] = B->vx[ib_idx]; o[3] = B->vy[ib_idx]; o[4] = B->vz[ib_idx];
        }
        double v2new = B->vx[ib_idx] * B->vx[ib_idx] + B->vy[ib_idx] * B->vy[ib_idx] +
                       B->vz[ib_idx] * B->vz[ib_idx];
        Pcoll[bt] = Pcoll[bt] + 0.5 * c->Nm[bt] * c->mass[bt] * (v2new - v2old);
      }
      for (int ir = 0; ir < n_re; ++ir) {
        int ct = c->col_info[c_ind * 10 + 2 + ir];
        if (ct < 0 || ct >= NT) continue;
        if (c->mass[ct] <= 0.0) continue;
        if (flag_add[ir] == -1) {
          if (ct == pt) P->dead[ip] = 1;
        }
      }
    }
  }
  return LO_OK;
}

/* ===================================================================== */
/* One step: modules/simulation/mod_simulation.f90:340-472                */
/* ===================================================================== */

int lo_advance_one_step(lo_state *s, int32_t istep) {
  const lo_grid *g = &s->g;
  const int64_t nodes = NODES(g);
  int st;
  if (istep % s->nb_step_sort == 1)
    for (int t = 0; t < s->ntype; ++t)
      if (s->parts[t].n > 1) lo_sort_particles_by_cell(g, &s->parts[t], s->cell_count, s->cell_start);
  lo_reduce_density(g, s->bcnd, s->ntype, 1, s->np_thread, s->np_red);
  lo_build_rho(g, s->ntype, s->np_red, s->charge, s->rho);
  st = lo_pdesolver(g, s->phi, s->rho, s->bcnd, s->flag_pbc, s->flag_nmn, s->eps0, s->ncycl,
                    s->eps, s->omega, s->ng, &s->last_mg);
  if (st != LO_OK) return st;
  lo_efield(g, s->phi, s->bcnd, s->E4);
  double xmax = (double)g->n[0] * g->h[0], ymax = (double)g->n[1] * g->h[1],
         zmax = (double)g->n[2] * g->h[2];
  for (int t = 0; t < s->ntype; ++t) {
    if (s->mass[t] <= 0.0) continue;
    if (s->b_mode == 0)
      lo_move_electrostatic(g, &s->parts[t], s->E4, s->charge[t], s->mass[t], s->dt);
    else
      lo_move_boris(g, &s->parts[t], s->E4, s->b_mode, s->Bconst, s->B4, s->charge[t], s->mass[t],
                    s->dt, (s->charge[t] > 0.0 && !s->ions_magnetized) ? 0 : 1);
  }
  for (int t = 0; t < s->ntype; ++t) {
    int64_t lost = 0;
    lo_apply_bc(g, &s->parts[t], s->bcnd, xmax, ymax, zmax, s->flag_pbc, s->flag_nmn, t + 1,
                s->tag_neg, s->Nm[t] * s->charge[t], s->mass[t], s->Nm[t], s->ngrid,
                s->p_mac + (int64_t)t * 2 * (s->ngrid + 1), &s->P_loss[t], &lost);
  }
  if (s->chem && s->nb_step_collisions > 0 && istep % s->nb_step_collisions == 0) {
    st = lo_collisions(g, s->parts, s->np_red, s->chem, s->nb_step_collisions, s->dt, s->iseed,
                       s->Pcoll, NULL);
    if (st != LO_OK) return st;
  }
  memset(s->np_thread, 0, sizeof(double) * (size_t)(nodes * s->ntype));
  for (int t = 0; t < s->ntype; ++t)
    lo_deposit(g, &s->parts[t], s->kq, s->Nm[t], s->np_thread + (int64_t)t * nodes);
  lo_reduce_density(g, s->bcnd, s->ntype, 1, s->np_thread, s->np_red);
  return LO_OK;
}

/* ===================================================================== */
/* The FP32 particle phase (the precision form; no reference counterpart */
/* beyond the FP64 routines it restates in single precision):            */
/* cell id + FP32 in-cell offsets + FP32 velocities.                     */
/* ===================================================================== */

/* F4f = F4 rounded to FP32. */
void lo32_round4(const double *F4, int64_t count, float *F4f) {
  for (int64_t i = 0; i < count; ++i) F4f[i] = (float)F4[i];
}

/* FP64 -> FP32 set, slot for slot: cell floor(x/h), offset x/h - floor(x/h)
 * (exact in FP64) rounded once to FP32, a rounding up to 1 moving the
 * particle to the next cell's lower face.  Dead particles and particles
 * outside [0, L) become holes (cell -1); returns the count of the latter. */
int64_t lo32_from64(const lo_grid *g, const lo_pset *src, lo_pset32 *dst) {
  int64_t outside = 0;
  for (int64_t i = 0; i < src->n; ++i) {
    int32_t cell = -1;
    float f[3] = {0.0f, 0.0f, 0.0f};
    if (src->dead[i] == 0) {
      const double x[3] = {src->x[i], src->y[i], src->z[i]};
      int c[3];
      int out = 0;
      for (int d = 0; d < 3; ++d) {
        const double q = x[d] / g->h[d];
        const double fl = floor(q);
        c[d] = -1;
        if (q >= 0.0 && fl < (double)g->n[d]) {
          c[d] = (int)fl;
          f[d] = (float)(q - fl);
          if (f[d] >= 1.0f) { f[d] = 0.0f; c[d] += 1; }
          if (c[d] >= g->n[d]) c[d] = -1;
        }
        if (c[d] < 0) out = 1;
      }
      if (out) {
        outside++;
        f[0] = f[1] = f[2] = 0.0f;
      } else {
        cell = c[0] + g->n[0] * (c[1] + g->n[1] * c[2]);
      }
    }
    dst->cell[i] = cell;
    dst->fx[i] = f[0]; dst->fy[i] = f[1]; dst->fz[i] = f[2];
    dst->vx[i] = (float)src->vx[i]; dst->vy[i] = (float)src->vy[i]; dst->vz[i] = (float)src->vz[i];
  }
  dst->n = src->n;
  return outside;
}

static void decode32(const lo_grid *g, int32_t c, int *cx, int *cy, int *cz) {
  *cx = c % g->n[0];
  const int r = c / g->n[0];
  *cy = r % g->n[1];
  *cz = r / g->n[1];
}

/* FP32 -> FP64 set: x = ((ix-1) + fx) h1 in FP64; holes -> dead = 2. */
void lo32_to64(const lo_grid *g, const lo_pset32 *src, lo_pset *dst) {
  for (int64_t i = 0; i < src->n; ++i) {
    const int32_t c = src->cell[i];
    int cx = 0, cy = 0, cz = 0;
    if (c >= 0) decode32(g, c, &cx, &cy, &cz);
    dst->x[i] = ((double)cx + (double)src->fx[i]) * g->h[0];
    dst->y[i] = ((double)cy + (double)src->fy[i]) * g->h[1];
    dst->z[i] = ((double)cz + (double)src->fz[i]) * g->h[2];
    dst->vx[i] = (double)src->vx[i];
    dst->vy[i] = (double)src->vy[i];
    dst->vz[i] = (double)src->vz[i];
    dst->cell[i] = c;
    dst->dead[i] = c >= 0 ? 0 : 2;
  }
  dst->n = src->n;
}

static void weights32(float fx, float fy, float fz, float w[8]) {
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

/* 4-float node field at the 8 stencil nodes of lower node b, component k,
 * summed w1..w8 left to right in FP32. */
static float gather32(const lo_grid *g, const float *F4, int64_t b, const float w[8], int k) {
  const int64_t sy = NX(g), sz = NX(g) * NY(g);
  const int64_t nd[8] = {b, b + 1, b + sy + 1, b + sy, b + sz, b + sz + 1, b + sz + sy + 1,
                         b + sz + sy};
  float s = w[0] * F4[4 * nd[0] + k];
  for (int q = 1; q < 8; ++q) s = s + w[q] * F4[4 * nd[q] + k];
  return s;
}

static int advance32(float *f, float d) {
  const float s = *f + d;
  const float fl = floorf(s);
  float r = s - fl;
  int k = (int)fl;
  if (r >= 1.0f) { r = 0.0f; k += 1; }
  *f = r;
  return k;
}

/* Fixed-point parts of a particle at lower node (ix,iy,iz), offsets f, added
 * (sign +1) or removed (-1) from dep. */
/* The FP32 phase's split of a full weight U (an integer of at most 2^22
 * units) over the 8 stencil nodes, in FP32: along x, then y, then z the
 * upper node takes rint(parent * offset) and the lower one the rest; every
 * product stays below 2^22 and every difference is of integers below 2^23,
 * so the parts are exact integers summing to U. */
void lo32_fixed_split(float U, float ux, float uy, float uz, int32_t a[8]) {
  const float x1 = rintf(U * ux), x0 = U - x1;
  const float x0y1 = rintf(x0 * uy), x0y0 = x0 - x0y1;
  const float x1y1 = rintf(x1 * uy), x1y0 = x1 - x1y1;
  const float c[4] = {x0y0, x1y0, x1y1, x0y1};
  for (int q = 0; q < 4; ++q) {
    const float z1 = rintf(c[q] * uz);
    a[q] = (int32_t)(c[q] - z1);
    a[q + 4] = (int32_t)z1;
  }
}

static void dep32(const lo_grid *g, int64_t *dep, double U, int neg, int ix, int iy, int iz,
                  float fx, float fy, float fz, int sign) {
  int32_t a[8];
  lo32_fixed_split((float)U, fx, fy, fz, a);
  const int64_t c[8] = {IDX(g, ix, iy, iz),         IDX(g, ix + 1, iy, iz),
                        IDX(g, ix + 1, iy + 1, iz), IDX(g, ix, iy + 1, iz),
                        IDX(g, ix, iy, iz + 1),     IDX(g, ix + 1, iy, iz + 1),
                        IDX(g, ix + 1, iy + 1, iz + 1), IDX(g, ix, iy + 1, iz + 1)};
  const int flip = (sign < 0) != (neg != 0);
  for (int q = 0; q < 8; ++q)
    dep[c[q]] = (int64_t)((uint64_t)dep[c[q]] + (uint64_t)(int64_t)(flip ? -a[q] : a[q]));
}

static double unit32(const lo_grid *g, double Nm, int32_t frac_bits, int *neg) {
  const double k1 = Nm / (g->h[0] * g->h[1] * g->h[2]);
  *neg = k1 < 0.0;
  return rint(ldexp(fabs(k1), frac_bits));
}

/* The FP32 particle phase in place: gather (E4f, FP32), Boris/leapfrog
 * update (Src/part_expmover.f90:139-240 in FP32), offsets advanced by v dt/h,
 * the wall test / losses / Neumann reflection / periodic wrap of
 * apply_particle_bc_legacy (mod_particleBC.f90:96-227) in cell terms, and
 * the fixed-point deposit of survivors (dep may be NULL).  Lost particles
 * become holes; so do survivors left outside the cells (counted in
 * *escaped).  Constants: k1 = dt q/(2m), dt/h, the uniform-B factor k2,
 * formed in FP64 and rounded once. */
void lo32_push(const lo_grid *g, lo_pset32 *p, const float *E4f, int bmode, const double *Bc,
               const float *B4f, int magnetized, double q, double m, double dt,
               const int32_t *bcnd, int flag_pbc, int flag_nmn, int ptype, int tag_neg,
               double qmacro, double mass, double Nm, int32_t ngrid, int32_t frac_bits,
               int64_t *dep, double *p_mac, double *P_loss, int64_t *lost, int64_t *escaped) {
  const int32_t *n = g->n;
  const double k1d = dt * q / (2.0 * m);
  const float k1 = (float)k1d;
  const float dtx = (float)(dt / g->h[0]), dty = (float)(dt / g->h[1]),
              dtz = (float)(dt / g->h[2]);
  const float bx = (float)Bc[0], by = (float)Bc[1], bz = (float)Bc[2];
  const double Bt = sqrt(Bc[0] * Bc[0] + Bc[1] * Bc[1] + Bc[2] * Bc[2]);
  const float k2u = (float)(k1d * (2.0 / (1.0 + (k1d * Bt) * (k1d * Bt))));
  int neg = 0;
  const double U = dep ? unit32(g, Nm, frac_bits, &neg) : 0.0;
  for (int64_t i = 0; i < p->n; ++i) {
    const int32_t c = p->cell[i];
    if (c < 0) continue;
    int cx, cy, cz;
    decode32(g, c, &cx, &cy, &cz);
    const int ix = cx + 1, iy = cy + 1, iz = cz + 1;
    float fx = p->fx[i], fy = p->fy[i], fz = p->fz[i];
    float vx = p->vx[i], vy = p->vy[i], vz = p->vz[i];
    float w[8];
    weights32(fx, fy, fz, w);
    const int64_t b = IDX(g, ix, iy, iz);
    const float Ex = gather32(g, E4f, b, w, 0), Ey = gather32(g, E4f, b, w, 1),
                Ez = gather32(g, E4f, b, w, 2);
    int use_b = 0;
    float Bx = 0.0f, By = 0.0f, Bz = 0.0f;
    if (bmode == 1) {
      Bx = bx; By = by; Bz = bz;
      use_b = 1;
    } else if (bmode == 2 && B4f[4 * b + 3] > 1.0e-5f) {
      Bx = gather32(g, B4f, b, w, 0); By = gather32(g, B4f, b, w, 1);
      Bz = gather32(g, B4f, b, w, 2);
      use_b = 1;
    }
    if (!magnetized) use_b = 0;
    if (use_b) {
      float k2;
      if (bmode == 1) {
        k2 = k2u;
      } else {
        const float Btf = sqrtf(Bx * Bx + By * By + Bz * Bz);
        k2 = k1 * (2.0f / (1.0f + (k1 * Btf) * (k1 * Btf)));
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
    int nx = ix + advance32(&fx, vx * dtx);
    int ny = iy + advance32(&fy, vy * dty);
    int nz = iz + advance32(&fz, vz * dtz);
    /* wall test at the clamped lower node */
    int ax = nx < 0 ? 0 : (nx > n[0] + 1 ? n[0] + 1 : nx);
    int ay = ny < 0 ? 0 : (ny > n[1] + 1 ? n[1] + 1 : ny);
    int az = nz < 0 ? 0 : (nz > n[2] + 1 ? n[2] + 1 : nz);
    int flag_lost = all_walls(g, bcnd, ax, ay, az) ? 1 : 0;
    int label = flag_lost ? bcnd[IDX(g, ax, ay, az)] : 0;
    if (ptype == tag_neg && nx <= 0 && flag_nmn == 1) { flag_lost = 2; label = 0; }
    if (flag_lost) {
      if (label < 0) label = 0;
      const double v2 = (double)vx * (double)vx + (double)vy * (double)vy + (double)vz * (double)vz;
      const double Ek_eV = 0.5 * mass * v2 / QE;
      if (label <= ngrid) {
        p_mac[2 * label + 0] = p_mac[2 * label + 0] + qmacro;
        p_mac[2 * label + 1] = p_mac[2 * label + 1] + fabs(qmacro) * Ek_eV;
      }
      *P_loss = *P_loss + Nm * (0.5 * mass * v2);
      (*lost)++;
      p->cell[i] = -1;
      p->fx[i] = fx; p->fy[i] = fy; p->fz[i] = fz;
      p->vx[i] = vx; p->vy[i] = vy; p->vz[i] = vz;
      continue;
    }
    if (flag_nmn == 1) {
      if (nx <= 0) {
        const float r = 1.0f - fx;
        nx = 1 - nx;
        if (r >= 1.0f) { fx = 0.0f; nx += 1; } else { fx = r; }
        vx = -vx;
      } else if (nx == 1 && fx == 0.0f) {
        vx = -vx;
      }
    }
    if (flag_pbc == 1) {
      if (ny >= n[1] + 1) ny -= n[1];
      if (ny <= 0) ny += n[1];
      if (nz >= n[2] + 1) nz -= n[2];
      if (nz <= 0) nz += n[2];
    }
    p->fx[i] = fx; p->fy[i] = fy; p->fz[i] = fz;
    p->vx[i] = vx; p->vy[i] = vy; p->vz[i] = vz;
    if (nx < 1 || nx > n[0] || ny < 1 || ny > n[1] || nz < 1 || nz > n[2]) {
      (*escaped)++;
      p->cell[i] = -1;
      continue;
    }
    p->cell[i] = (nx - 1) + n[0] * ((ny - 1) + n[1] * (nz - 1));
    if (dep) dep32(g, dep, U, neg, nx, ny, nz, fx, fy, fz, 1);
  }
}

/* The fixed-point deposit of an FP32 set (lepic_deposit_fixed's split with
 * the offsets as upper shares), added to dep. */
void lo32_deposit(const lo_grid *g, const lo_pset32 *p, double Nm, int32_t frac_bits,
                  int64_t *dep) {
  int neg = 0;
  const double U = unit32(g, Nm, frac_bits, &neg);
  for (int64_t i = 0; i < p->n; ++i) {
    const int32_t c = p->cell[i];
    if (c < 0) continue;
    int cx, cy, cz;
    decode32(g, c, &cx, &cy, &cz);
    dep32(g, dep, U, neg, cx + 1, cy + 1, cz + 1, p->fx[i], p->fy[i], p->fz[i], 1);
  }
}

/* np += kq * dep * 2^-frac_bits. */
void lo32_fixed_density(const lo_grid *g, const int64_t *dep, const double *kq,
                        int32_t frac_bits, double *np) {
  const double inv = ldexp(1.0, -frac_bits);
  for (int64_t t = 0; t < NODES(g); ++t) np[t] = np[t] + kq[t] * ((double)dep[t] * inv);
}

/* Stable counting sort by cell of an FP32 set, holes dropped
 * (sort_particles_by_cell, mod_particle_sorting.f90:88-181). */
void lo32_sort(const lo_grid *g, const lo_pset32 *src, lo_pset32 *dst) {
  const int64_t ncells = (int64_t)g->n[0] * g->n[1] * g->n[2];
  int64_t *start = (int64_t *)calloc((size_t)ncells + 1, sizeof(int64_t));
  for (int64_t i = 0; i < src->n; ++i)
    if (src->cell[i] >= 0) start[src->cell[i] + 1]++;
  for (int64_t c = 0; c < ncells; ++c) start[c + 1] += start[c];
  for (int64_t i = 0; i < src->n; ++i) {
    const int32_t c = src->cell[i];
    if (c < 0) continue;
    const int64_t d = start[c]++;
    dst->cell[d] = c;
    dst->fx[d] = src->fx[i]; dst->fy[d] = src->fy[i]; dst->fz[d] = src->fz[i];
    dst->vx[d] = src->vx[i]; dst->vy[d] = src->vy[i]; dst->vz[d] = src->vz[i];
  }
  dst->n = start[ncells - 1 >= 0 ? ncells - 1 : 0];
  free(start);
}
