/*
 * lepic.h -- C ABI of the MI355X (gfx950) PIC-MCC step library, liblepic_hip.so.
 *
 * Plain C: device pointers, sizes and small POD structs; no torch or HIP types
 * in the signatures (streams travel as void*).  Every entry point returns an
 * int status (LEPIC_OK or an LEPIC_ERR_* code, message in lepic_last_error())
 * and never exits the process.  Calls are enqueued on the context's stream;
 * the few that must return a count to the host (apply_bc, the fused step,
 * poisson_solve, collisions) synchronise that stream before returning.
 *
 * Each entry point replaces one routine of the reference (Fortran, cited as
 * path:line relative to the reference repository).  The reference has no FFI
 * of its own; the Python package essaitemporaire_lepic_module_amd binds these
 * symbols with ctypes under the reference's module/procedure names.
 *
 * Index conventions (identical to the reference's index space, 0-based here):
 *  - grid n = (n1,n2,n3) cells, nodes 0..n+2 per dimension (Fortran 0:n+2),
 *    node (i,j,k) at ((i-1)h1,(j-1)h2,(k-1)h3); arrays are x-fastest with
 *    idx = i + (n1+3)*(j + (n2+3)*k).  b/rho use the same strides.
 *  - cell id c = (ix-1) + n1*((iy-1) + n2*(iz-1)), ix in [1,n1] (the
 *    reference stores c+1).  Particle indices are 0-based.
 *  - E fields are 4 doubles per node (Ex,Ey,Ez,pad); B maps are 4 doubles per
 *    node (Bx,By,Bz,|B|) like the reference's Bi(4,...).
 *  - bcnd labels: -1 plasma, 0 periodic, -2 Neumann, >= 1 electrode label.
 */
#ifndef LEPIC_H
#define LEPIC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LEPIC_OK = 0,
  LEPIC_ERR_ARG = 1,       /* bad argument / unsupported shape          */
  LEPIC_ERR_CAPACITY = 2,  /* destination too small; see lepic_last_error */
  LEPIC_ERR_DIVERGED = 3,  /* Poisson res/res_s > 10 (pdesolver.f90:103)  */
  LEPIC_ERR_PMAX = 4,      /* MCC Pmax > 1 (mod_collisions.f90:354)        */
  LEPIC_ERR_HIP = 5,       /* HIP runtime error                           */
  LEPIC_ERR_SPIN = 6       /* a bounded device wait gave up                */
};

typedef struct lepic_ctx lepic_ctx;

typedef struct {
  int32_t n[3];
  double h[3];
} lepic_grid;

/* Device-resident structure-of-arrays particle set of one species (FP64).
 * Replaces ParticleSet (modules/particles/mod_particles.f90:7-34); the
 * reference's w, sp and flag_cex carry no information on this path. */
typedef struct {
  double *x, *y, *z, *vx, *vy, *vz;
  int32_t *cell;   /* cell id, valid after lepic_sort_particles_by_cell */
  int8_t *dead;    /* flag_dead: 1 = killed by MCC                      */
  int64_t n;       /* live particles (host copy)                        */
  int64_t cap;     /* length of every array                             */
} lepic_pset;

/* Wall-BC parameters (apply_particle_bc_legacy arguments,
 * modules/particles/mod_particleBC.f90:42-62).  With flag_die = 1 a particle
 * lost on a label whose type is a dielectric (> 1) adds its charge Nm*q to
 * the species' surface charge (mod_particleBC.f90:139-188): bilinear in x-z
 * into sum_q_xz ((n1+3) x (n3+3), x fastest) for type 2, in y-z into
 * sum_q_yz (2 x (n2+3) x (n3+3), the plane index 0 or 1 fastest) for types
 * 3 and 4.  dtype is the domain's table over labels 0..ngrid, read as the
 * routine's 1-based dummy reads it: label L's type is dtype[L-1]. */
typedef struct {
  double xmax, ymax, zmax;
  int32_t flag_pbc, flag_nmn;
  int32_t ptype, tag_neg;  /* 1-based species id, negative-ion species or -1 */
  double qmacro;           /* Nm*q */
  double mass, Nm;
  int32_t ngrid;           /* highest electrode label                 */
  int32_t flag_die;        /* 1: dielectric surface charging           */
  const int32_t *dtype;    /* device, ngrid+1 (flag_die = 1)           */
  double *sum_q_xz;        /* device (flag_die = 1)                    */
  double *sum_q_yz;        /* device (flag_die = 1)                    */
} lepic_bc_params;

/* Magnetic field of the Boris push (Src/part_expmover.f90:139-212). */
typedef struct {
  int32_t mode;            /* 0 none (leapfrog), 1 uniform, 2 map on the E grid,
                              3 map on its own grid (flag_gridB = 1)   */
  double B[3];             /* uniform field (mode 1)                  */
  const double *B4;        /* device map, 4 doubles per node (modes 2, 3); mode 3:
                              (nB1+3)(nB2+3)(nB3+3) nodes, x fastest  */
  int32_t magnetized;      /* 0: this species ignores B (flag_B_pos)  */
  int32_t nB[3];           /* mode 3: cells of the B grid             */
  double hB[3];            /* mode 3: its cell size                   */
} lepic_bfield;

typedef struct {
  int32_t flag_pbc, flag_nmn;
  double eps0;
  int32_t ncycl;           /* V-cycle cap (20000 in mod_simulation.f90:392) */
  double eps;              /* tolerance on L1 residual / alpha            */
  double omega;            /* SOR over-relaxation                          */
  int32_t ng;              /* multigrid levels (<= 12)                     */
  int32_t graph;           /* 1: every V-cycle replays one captured HIP graph
                              of the same launches (0: launched one by one) */
} lepic_mg_params;

typedef struct {
  int32_t cycles;          /* V-cycles executed                        */
  double ktot;             /* equivalent fine-grid sweeps              */
  double res;              /* final L1 residual / alpha                */
  int64_t sweeps[12];      /* SOR sweeps per level                     */
} lepic_mg_stats;

/* MCC reaction tables, device pointers, layout as documented per field
 * (built on the host as modules/particles/mod_reactions.f90:414-558 does). */
#define LEPIC_MAX_COL 128
#define LEPIC_MAX_SPECIES 16
typedef struct {
  int32_t ntype_tracked, ntype_total;
  double mass[LEPIC_MAX_SPECIES];       /* neutrals negative              */
  double vt0[LEPIC_MAX_SPECIES];
  double Nm[LEPIC_MAX_SPECIES];
  double neutral_density[LEPIC_MAX_SPECIES];
  double nu_uplim[LEPIC_MAX_SPECIES];
  int32_t p_ncol[LEPIC_MAX_SPECIES];
  int32_t ncol, npt;
  const int32_t *sig_list;  /* device [ntype_total][LEPIC_MAX_COL]          */
  const double *sig_Er;     /* device [npt], strictly increasing (eV)       */
  const double *sig;        /* device [ncol][npt] (m^2)                      */
  const double *sig_Eth;    /* device [ncol] threshold (eV)                  */
  const int32_t *col_info;  /* device [ncol][10]: n_re, n_by, species...    */
  const double *sigv_mx;    /* device [ntype_total][ncol]                    */
} lepic_chem;

enum { LEPIC_MCC_REPLAY_RAN2 = 0, LEPIC_MCC_PHILOX = 1 };

typedef struct {
  int32_t Nc[LEPIC_MAX_SPECIES];
  int64_t accepted[LEPIC_MAX_SPECIES];
  int64_t err_coll[LEPIC_MAX_SPECIES];
  int64_t skipped_dead[LEPIC_MAX_SPECIES];
  int64_t appended[LEPIC_MAX_SPECIES];
} lepic_mcc_stats;

/* ---------------- context ---------------- */
int lepic_create(int32_t device, lepic_ctx **out);
int lepic_destroy(lepic_ctx *ctx);
const char *lepic_last_error(lepic_ctx *ctx);
int lepic_set_stream(lepic_ctx *ctx, void *hip_stream); /* NULL = null stream */
int lepic_synchronize(lepic_ctx *ctx);
const char *lepic_version(void);

/* ---------------- a11: RNG (modules/core/mod_RNG.f90:13-45) ---------------- */
double lepic_ran2(int32_t *irand);                          /* host, one draw   */
int32_t lepic_seed(int32_t iproc, int32_t mpi_rank);        /* mod_simParams.f90:202 */

/* ---------------- a13: loaders ---------------- */
/* load_particles_modular (mod_particle_loader.f90:14-246) for one virtual
 * thread, on the host with ran2, written to host SoA arrays given in hp[]
 * (ntype sets); np_thread (host, ntype*nodes) receives the initial deposit. */
int lepic_load_particles_host(const lepic_grid *g, const int32_t *bcnd_host,
                              const double *kq_host, double x_load, int32_t np_cell,
                              int32_t ntype, const double *vt0, const double *Nm,
                              const double *ni0, int32_t *iseed, lepic_pset *hp,
                              double *np_thread_host);
/* Device loader for large synthetic plasmas: np_cell particles per loadable
 * cell, uniform in the cell, Maxwellian (vt) velocities, counter-based
 * Philox-4x32-10 stream (key, species) -- statistically the reference's
 * loader, not its ran2 sequence.  Species share positions when
 * share_positions != 0 (as the reference loader does). */
int lepic_load_uniform_device(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                              int32_t np_cell, int32_t ntype, const double *vt0,
                              uint64_t key, lepic_pset *dp);

/* ---------------- a2: sort (mod_particle_sorting.f90:16-181) ---------------- */
/* Stable counting sort by cell of src into dst (dst->cap >= src->n).
 * cell_count/cell_start (device, ncells int32, 0-based start) may be NULL. */
int lepic_sort_particles_by_cell(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                                 lepic_pset *dst, int32_t *cell_count, int32_t *cell_start);

/* ---------------- a3/a4: movers ---------------- */
/* move_particles_electrostatic (mod_particleMover.f90:72-156), in place. */
int lepic_move_electrostatic(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p,
                             const double *E4, double q, double m, double dt);
/* part_mover's Boris/leapfrog velocity update (Src/part_expmover.f90:51-246)
 * on the modular gather, in place. */
int lepic_move_boris(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p, const double *E4,
                     const lepic_bfield *bf, double q, double m, double dt);

/* ---------------- a5: wall BC + stable compaction ---------------- */
/* apply_particle_bc_legacy (mod_particleBC.f90:42-259): src -> dst, survivors
 * in order.  p_mac (device, 2*(ngrid+1), indexed by label) and P_loss
 * (device, 1) accumulate; *np_lost (host) receives the removed count. */
int lepic_apply_bc(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src, lepic_pset *dst,
                   const int32_t *bcnd, const lepic_bc_params *bp, double *p_mac,
                   double *P_loss, int64_t *np_lost);

/* ---------------- a6/a7: deposition, density, rho ---------------- */
int lepic_build_kq(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd, double *kq);
/* clear_np_thread (mod_chargeDeposition.f90:16-22) */
int lepic_fill(lepic_ctx *ctx, double *a, int64_t count, double value);
/* deposit_particle_set_to_np_thread (mod_chargeDeposition.f90:25-133):
 * np_local += kq*Nm/(h1 h2 h3)*w over the 8 nodes (FP64 atomics). */
int lepic_deposit(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *p, const double *kq,
                  double Nm, double *np_local);

/* deposit_particle_set_to_np_thread (mod_chargeDeposition.f90:25-133) in
 * fixed point: every particle's full weight U = rint(|k1| 2^frac_bits) units
 * (k1 = Nm/(h1 h2 h3)) is split over its 8 stencil nodes with the
 * reference's trilinear fractions -- divided along x, then y, then z, the
 * upper node of each division taking rint(parent * share) and the lower one
 * the rest -- so the 8 parts are integers that sum to U exactly; they are
 * summed per node as int64, then np_local += kq * sum * 2^-frac_bits (parts
 * negated when k1 < 0).  The result does not depend on particle order or
 * atomic timing, the charge on the grid is exactly N U units for N
 * deposited particles, and each part is within 1.5 units of U times its
 * trilinear weight.  |k1| * 2^frac_bits must lie in [1, 2^44] (LEPIC_ERR_ARG
 * otherwise; Nm = 0 deposits nothing and succeeds).  No part exceeds U, so a
 * node's int64 sum cannot overflow while fewer than 2^19 particles touch it;
 * past that the node total wraps without an error. */
int lepic_deposit_fixed(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *p,
                        const double *kq, double Nm, int32_t frac_bits, double *np_local);
/* reduce_species_density (mod_density.f90:31-73), one rank, one virtual
 * thread: periodic stitching in place on np_thread [ntype][nodes], then
 * np_red = np_thread on nodes 1..n+1 and 0 elsewhere. */
int lepic_reduce_density(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                         int32_t ntype, double *np_thread, double *np_red);
/* build_rho_from_np (mod_density.f90:76-107); charge[] is host memory. */
int lepic_build_rho(lepic_ctx *ctx, const lepic_grid *g, int32_t ntype, const double *np_red,
                    const double *charge, double *rho);

/* ---------------- a8: Poisson ---------------- */
/* solve_poisson_legacy -> pdesolver -> mg -> sor_rb/restriction/prolongation
 * (mod_PoissonSolver_legacy.f90:27-101, pdesolver.f90:1-135, mg.f90:1-940,
 * sors.f90:1-535), one rank.  phi is updated in place (warm start). */
int lepic_poisson_solve(lepic_ctx *ctx, const lepic_grid *g, double *phi, const double *rho,
                        const int32_t *bcnd, const lepic_mg_params *mp, lepic_mg_stats *stats);

/* Precision form of the solve: FP32 red-black smoothing V-cycles inside an
 * FP64 iterative refinement.  Each refinement step forms the FP64 residual
 * r = rho - A phi at every node the smoother updates, rounds it to FP32,
 * runs inner_cycles V-cycles of the multigrid above in FP32 on A e = r (e = 0
 * on electrodes) with fixed sweep counts -- nu red-black sweeps per level on
 * each leg at over-relaxation omega, nu_coarsest on the coarsest level --
 * and adds e to phi in FP64.  It stops, as pdesolver does, when the FP64 L1
 * residual over plasma nodes / alpha falls below mp->eps (or after
 * mp->ncycl refinement steps); stats->cycles counts refinement steps and
 * stats->sweeps the FP32 sweeps per level.  mp->omega is not used.
 * (No counterpart in the reference, whose solve is FP64 throughout.) */
typedef struct {
  int32_t inner_cycles;    /* FP32 V-cycles per refinement step (1..64)  */
  int32_t nu;              /* sweeps per level per leg (1..16)           */
  int32_t nu_coarsest;     /* sweeps on the coarsest level               */
  double omega;            /* over-relaxation of the FP32 smoother       */
  int32_t precond_fp64;    /* lepic_poisson_solve_pcg only: 1 runs the
                              preconditioner's V-cycles in FP64 (0: FP32) */
} lepic_mgmix_params;

int lepic_poisson_solve_mixed(lepic_ctx *ctx, const lepic_grid *g, double *phi, const double *rho,
                              const int32_t *bcnd, const lepic_mg_params *mp,
                              const lepic_mgmix_params *xp, lepic_mg_stats *stats);

/* Preconditioned conjugate gradients in FP64 with the FP32 V-cycles above
 * as the preconditioner (a flexible CG keeping one previous direction), or
 * with the same V-cycles in FP64 (xp->precond_fp64 = 1: FP64 throughout):
 * each iteration applies inner_cycles fixed-sweep V-cycles to A z = r from
 * z = 0, makes the new direction A-orthogonal to the previous one, and
 * updates phi and r along it.  Stops on the same criterion as the solves
 * above (FP64 L1 residual over plasma nodes / alpha < mp->eps, confirmed on
 * the residual recomputed from phi) or after mp->ncycl iterations;
 * stats->cycles counts iterations.  The operator must be symmetric: plasma
 * (-1) and electrode (> 0) labels only, no periodic or Neumann walls
 * (LEPIC_ERR_ARG otherwise).  Fewer V-cycles than the refinement above where
 * the electrodes' shapes slow the multigrid (curved or internal
 * electrodes).  (No counterpart in the reference.) */
int lepic_poisson_solve_pcg(lepic_ctx *ctx, const lepic_grid *g, double *phi, const double *rho,
                            const int32_t *bcnd, const lepic_mg_params *mp,
                            const lepic_mgmix_params *xp, lepic_mg_stats *stats);

/* ---------------- a9: E field ---------------- */
/* calc_Efield_modular (mod_electricField.f90:9-164). */
int lepic_efield(lepic_ctx *ctx, const lepic_grid *g, const double *phi, const int32_t *bcnd,
                 double *E4);

/* lepic_efield for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  Same arguments and the same E,
 * bit for bit; the wall pass works in a static device array of
 * LEPIC_EFIELD_NOALLOC_MAX_NODES doubles (18 MiB, held by every GPU the
 * library runs on until the process ends) instead of context scratch, and
 * the call allocates nothing.  Grids with more nodes ((n1+3)(n2+3)(n3+3))
 * give LEPIC_ERR_ARG.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (two streams, two contexts) race. */
#define LEPIC_EFIELD_NOALLOC_MAX_NODES 2359296
int lepic_efield_noalloc(lepic_ctx *ctx, const lepic_grid *g, const double *phi,
                         const int32_t *bcnd, double *E4);

/* ---------------- a10: null-collision MCC ---------------- */
/* perform_collisions_step (mod_collisions.f90:214-739) for one rank and one
 * virtual thread.  parts[ntype_tracked] are updated in place (velocities,
 * flag_dead, appended products up to cap).  np_red is [ntype][nodes].
 * mode LEPIC_MCC_REPLAY_RAN2 replays the reference's sequential chain with
 * the ran2 stream iseed (host in/out); LEPIC_MCC_PHILOX runs all particles in
 * parallel with a Poisson(Pmax) trial count each, drawing from Philox-4x32-10
 * (key, counter).  Pcoll (device, ntype_total) accumulates. */
int lepic_collisions(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *parts,
                     const double *np_red, const lepic_chem *chem, int32_t ns_coll, double dt,
                     int32_t mode, int32_t *iseed, uint64_t philox_key, uint64_t philox_counter,
                     double *Pcoll, lepic_mcc_stats *stats);
/* lepic_collisions for a step whose deposit already ran before MCC (the
 * production collision step deposits in the push kernel): np_dep (device,
 * [ntype_tracked][nodes], as deposit_particle_set_to_np_thread fills it) gets
 * the deposit of every product MCC appends and loses that of every particle
 * MCC kills, with the deposit's weights (kq, Nm / (h1 h2 h3)).  It then holds
 * what the reference's deposit after MCC (mod_simulation.f90:425-451) gives,
 * summed in another order. */
int lepic_collisions_deposited(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *parts,
                               const double *np_red, const lepic_chem *chem, int32_t ns_coll,
                               double dt, int32_t mode, int32_t *iseed, uint64_t philox_key,
                               uint64_t philox_counter, double *Pcoll, lepic_mcc_stats *stats,
                               double *np_dep, const double *kq);

/* ---------------- f4: dielectric surface charging ---------------- */
/* apply_dielectric_bc_to_phi (mod_state.f90:378-430): every node whose label
 * (bcnd > 0) has dtype 2, 3 or 4 takes, species by species, the surface
 * charge the wall BC collected: sum_q_xz (ntype blocks of (n1+3)(n3+3), the
 * node's x-z column) / (2 Ca_cell) for type 2, sum_q_yz (ntype blocks of
 * 2(n2+3)(n3+3), plane 0 for type 3, 1 for type 4) / Ca_cell.  dtype is the
 * domain's table over labels 0..ngrid (indexed by the label itself);
 * Ca_cell = Ca h1 h3 (mod_simParams.f90:166-167).  Labels above ngrid are
 * left alone. */
int lepic_dielectric_phi(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                         const int32_t *dtype, int32_t ngrid, int32_t ntype,
                         const double *sum_q_xz, const double *sum_q_yz, double Ca_cell,
                         double *phi);

/* ---------------- f4: electron heating ---------------- */
/* Parameters of the modular step's electron heating (mod_state.f90:341-373,
 * 524-612; mod_heating.f90:12-82). */
typedef struct {
  double nudt;             /* nu_h * nb_step_heating * dt (mod_simParams.f90:150)  */
  double Pabs, nu_h;       /* absorbed power (W), heating frequency (conditions)   */
  double xl_pow, xr_pow;   /* heating window along x (m)                           */
  int32_t flag_circxh, flag_ahp;
  double R_ahp_moments;    /* R_ahp the window moments see (0 before the first
                              heating sets it, mod_state.f90:350)                 */
  double R_ahp;            /* R_ahp the heating sees                               */
  double ymax, zmax;       /* domain extents (m)                                   */
  double Nm, mass;         /* electron macro-particle weight and mass              */
  double vt0;              /* vt used when the window is empty or Te <= 0          */
} lepic_heating;

/* compute_heating_region_moments + update_heating_vt + apply_electron_heating
 * on the electron set p (device): Te from the window's kinetic energy and
 * Pabs/nu_h, then each electron in the window re-drawn from a Maxwellian of
 * that temperature with probability nudt.  mode LEPIC_MCC_REPLAY_RAN2: the
 * reference's ran2 chain on one device thread, iseed (host) read and
 * advanced, every particle but holes (dead = 2) visited in order;
 * LEPIC_MCC_PHILOX: live particles only, each with its own draws from
 * (philox_key, philox_counter, index).  P_heat (device, 1 double)
 * accumulates 0.5 Nm m (v^2 new - v^2 old) (the reference's P_loss(2)).
 * out (host, optional): window count, window kinetic energy, vt used. */
int lepic_electron_heating(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p,
                           const lepic_heating *hp, int32_t mode, int32_t *iseed,
                           uint64_t philox_key, uint64_t philox_counter, double *P_heat,
                           double *out);

/* ---------------- f3: plane moments ---------------- */
/* compute_particle_plane_moments_species (mod_planeMoments.f90:13-308) for
 * one species.  Each live particle, moved back half a step (x - v dt/2),
 * adds bilinear weights to the xy plane at node iz_plane, the xz plane at
 * iy_plane and the yz plane at ix_plane (1-based node indices in 1..n+1) when
 * it lies within one cell of it.  data_xy, data_xz, data_yz (device) hold 5
 * doubles per plane node, component fastest (np, T, u1, u2, u3), then the
 * plane's first axis, over the ghosted 0..n+2 index space ((n1+3)(n2+3),
 * (n1+3)(n3+3) and (n2+3)(n3+3) nodes).  They are updated as the reference
 * updates them: where a node has weight, np and u1..u3 are overwritten by
 * this call's count and (previous u + sum w v) / count, and T accumulates
 * mass <v^2 - u^2> / (3 qe); elsewhere u1..u3 keep their value.  The sums go
 * through FP64 atomics (summation order is the hardware's). */
int lepic_plane_moments(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *p, double dt,
                        double mass, int32_t ix_plane, int32_t iy_plane, int32_t iz_plane,
                        double *data_xy, double *data_xz, double *data_yz);

/* ---------------- fused particle phase ---------------- */
/* One launch per species of gather + (Boris|leapfrog) push + wall BC +
 * stable compaction + deposit: the mover, apply_particle_bc_legacy and
 * deposit_particle_set_to_np_thread of one modular step
 * (mod_simulation.f90:425-451) without the intermediate passes over the
 * particles.  Valid when no MCC runs between the BC and the deposit.
 * np_local may be NULL (no deposit). */
int lepic_push_bc_deposit(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                          lepic_pset *dst, const double *E4, const lepic_bfield *bf, double q,
                          double m, double dt, const int32_t *bcnd, const lepic_bc_params *bp,
                          const double *kq, double *np_local, double *p_mac, double *P_loss,
                          int64_t *np_lost);

/* The same particle phase in place for production runs: lost particles (and
 * MCC-killed ones) become holes (dead = 2) instead of being compacted, so the
 * kernel streams each particle once with no cross-block ordering.  Holes are
 * skipped by every later pass and removed by lepic_sort_compact, which keeps
 * the survivors' relative order -- the order the reference's per-step stable
 * compaction gives.  np_local NULL skips the deposit (collision steps).
 * lost_acc (device, may be NULL) is incremented by this call's removals
 * without a host synchronisation; np_lost (host, may be NULL) makes the call
 * synchronise and return them.  The particle arrays must be 16-byte aligned
 * (dead: 4): the kernel moves them with vector accesses. */
int lepic_push_bc_deposit_inplace(lepic_ctx *ctx, const lepic_grid *g, lepic_pset *p,
                                  const double *E4, const lepic_bfield *bf, double q, double m,
                                  double dt, const int32_t *bcnd, const lepic_bc_params *bp,
                                  const double *kq, double *np_local, double *p_mac,
                                  double *P_loss, uint64_t *lost_acc, int64_t *np_lost);

/* Stable sort by cell (as lepic_sort_particles_by_cell) that drops every
 * particle with dead != 0; dst->n is the surviving count. */
int lepic_sort_compact(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                       lepic_pset *dst);

/* Production binning by cell (the cell order of sort_particles_by_cell,
 * mod_particle_sorting.f90:88-181, with the order inside a cell left open):
 * a counting sort that drops every particle with dead != 0 and sets dst->cell.
 * Two streaming passes over tiles of particles; for runs that the production
 * path re-sorts more often than the reference, where the order inside a cell
 * is immaterial.  The particle arrays must be 16-byte aligned (dead: 4). */
int lepic_bin_compact(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                      lepic_pset *dst);

/* ---------------- precision form: the FP32 particle phase ---------------- */
/* FP32 particle set of one species: the reference's ParticleSet
 * (modules/particles/mod_particles.f90:7-34) held as an int32 cell id (as in
 * lepic_pset.cell; -1: no particle -- lost, killed or never set) plus the
 * position inside the cell in cell widths (x = ((ix-1) + fx) h1, fx in
 * [0, 1)) and the velocity, all FP32: 28 bytes per particle.  Every array
 * must be 16-byte aligned; kernels move them with vector accesses. */
typedef struct {
  int32_t *cell;
  float *fx, *fy, *fz;
  float *vx, *vy, *vz;
  int64_t n;       /* slots in use (holes included)                  */
  int64_t cap;     /* length of every array                          */
} lepic_pset32;

/* Magnetic field of the FP32 push (Src/part_expmover.f90:139-212): mode 0
 * none, 1 uniform B, 2 a map on the E grid, 4 floats per node (Bx, By, Bz,
 * |B|) with the |B| > 1e-5 gate at the lower node. */
typedef struct {
  int32_t mode;
  double B[3];
  const float *B4;
  int32_t magnetized;
} lepic_bfield32;

/* FP64 set -> FP32 set, slot for slot: cell floor(x/h), offset x/h - cell
 * rounded once to FP32.  Dead particles and particles outside [0, L) become
 * holes; *n_outside (host, may be NULL: no synchronisation) receives the
 * latter's count.  (The layout conversion of ParticleSet, no reference
 * counterpart.) */
int lepic_pset_to32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *src,
                    lepic_pset32 *dst, int64_t *n_outside);
/* FP32 set -> FP64 set: x = ((ix-1) + fx) h1 in FP64, holes -> dead = 2. */
int lepic_pset_from32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset32 *src,
                      lepic_pset *dst);
/* F4f = F4 rounded to FP32 (4 values per node: E fields, B maps). */
int lepic_round_field4(lepic_ctx *ctx, const double *F4, int64_t nodes, float *F4f);

/* The particle phase of one modular step in FP32, in place: gather of the
 * FP32 E (E4f, 4 floats per node) + part_mover's Boris/leapfrog update
 * (Src/part_expmover.f90:51-246) + apply_particle_bc_legacy's wall test,
 * losses, Neumann reflection and periodic wrap (mod_particleBC.f90:42-259)
 * + the deposit (mod_chargeDeposition.f90:25-133) in fixed point: every
 * survivor's full weight rint(|Nm/(h1 h2 h3)| 2^frac_bits) (in [1, 2^21],
 * else LEPIC_ERR_ARG) split into 8 integer node parts as lepic_deposit_fixed
 * splits it (upper shares = the new offsets) and added to dep (device int64,
 * one per node, NULL: no deposit).  Constants are formed in FP64 and rounded once: k1 = dt q/(2m),
 * dt/h per axis, the uniform-B rotation factor.  Lost particles become holes
 * (cell = -1) and are added to *lost_acc (device); a survivor left outside
 * the cells (no wall around the domain) also becomes a hole and is added to
 * *escaped (device) -- an error for the caller to report.  Wall diagnostics
 * p_mac / P_loss accumulate in FP64 as lepic_push_bc_deposit_inplace's.
 * Dielectric charging (flag_die) is not supported here. */
int lepic_push32(lepic_ctx *ctx, const lepic_grid *g, lepic_pset32 *p, const float *E4f,
                 const lepic_bfield32 *bf, double q, double m, double dt, const int32_t *bcnd,
                 const lepic_bc_params *bp, int32_t frac_bits, int64_t *dep, double *p_mac,
                 double *P_loss, uint64_t *lost_acc, uint64_t *escaped);
/* The fixed-point deposit (deposit_particle_set_to_np_thread,
 * mod_chargeDeposition.f90:25-133) of an FP32 set, added to dep. */
int lepic_deposit32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset32 *p, double Nm,
                    int32_t frac_bits, int64_t *dep);
/* np_local += kq * dep * 2^-frac_bits: the FP64 density of fixed-point sums. */
int lepic_fixed_density(lepic_ctx *ctx, const lepic_grid *g, const int64_t *dep, const double *kq,
                        int32_t frac_bits, double *np_local);
/* sort_particles_by_cell (mod_particle_sorting.f90:88-181) of an FP32 set
 * into dst: stable (order within a cell = order in src), holes dropped,
 * dst->n = particles kept.  Synchronises once. */
int lepic_sort32(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset32 *src,
                 lepic_pset32 *dst);
/* perform_collisions_step (mod_collisions.f90:214-739) on FP32 sets, PHILOX
 * mode (as lepic_collisions), every reaction's target a neutral.  Products
 * take the parent's cell and offsets; killed particles become holes.  dep
 * (device, [ntype_tracked][nodes] int64, may be NULL) holds the step's
 * fixed-point deposit: appended products add theirs and killed particles
 * remove theirs, with frac_bits[t] (host) per species. */
int lepic_collisions32(lepic_ctx *ctx, const lepic_grid *g, lepic_pset32 *parts,
                       const double *np_red, const lepic_chem *chem, int32_t ns_coll, double dt,
                       uint64_t philox_key, uint64_t philox_counter, double *Pcoll,
                       lepic_mcc_stats *stats, int64_t *dep, const int32_t *frac_bits);

/* ---------------- differentiation: adjoint kernels ---------------- */
/* The reverse-mode derivative of the FP64 step run in place (holes instead
 * of compaction, so a particle keeps its slot): deposit -> rho -> multigrid
 * phi with electrode voltages -> E (mod_electricField.f90) -> Boris push
 * with the E gather (Src/part_expmover.f90:214-246, uniform B) -> x += dt v
 * -> wall test.  Which particles the walls absorb, and when, is frozen: the
 * adjoint is the derivative of the step with that event sequence fixed.
 * (No reference counterpart: the reference has no derivative of its step.)
 * An adjoint set holds dL/dx and dL/dv per particle slot (device FP64). */
typedef struct {
  double *lx, *ly, *lz, *lvx, *lvy, *lvz;
  int64_t n;
} lepic_adjoint_set;

/* Push adjoint of one species: state holds x_k, v_k, dead of the step's
 * start; adj holds dL/dx_{k+1}, dL/dv_{k+1} on entry and dL/dx_k, dL/dv_k
 * (push and gather terms) on return.  dL/dE (4 doubles per node, the E4
 * layout) accumulates the gather's transpose: dL/dE_p spread with the
 * gather's trilinear weights.  Bunif: the uniform B (3 doubles, host; NULL:
 * no B, the leapfrog update), whose dL/dB accumulates into lamB (device, 3
 * doubles). */
int lepic_adj_push(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                   lepic_adjoint_set *adj, const double *E4, const double *Bunif, double q,
                   double m, double dt, double *lamE4, double *lamB);
/* lepic_adj_push for a B map on the E grid (lepic_bfield mode 2: B4, 4
 * doubles per node, the |B| > 1e-5 gate at the particle's lower node,
 * Src/part_expmover.f90:149-212): B, the rotation factor and its
 * derivative are the particle's own, and the B gather's position term is
 * added to dL/dx.  magnetized = 0: this species ignores B.  The map is not
 * a parameter, so no dL/dB is returned. */
int lepic_adj_push_map(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                       lepic_adjoint_set *adj, const double *E4, const double *B4,
                       int32_t magnetized, double q, double m, double dt, double *lamE4);
/* lamphi = D^T lamE4, D the E field of calc_Efield_modular
 * (mod_electricField.f90:9-164) on a grid of plasma (-1) and electrode
 * (>= 1) labels (no periodic or Neumann labels); lamphi is overwritten. */
int lepic_adj_efield(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                     const double *lamE4, double *lamphi);
/* rhs = scale * lamphi on plasma nodes, 0 elsewhere: the right-hand side of
 * the adjoint Poisson solve, A mu = rhs with mu = 0 on electrodes, solved by
 * lepic_poisson_solve (the Dirichlet operator is symmetric). */
int lepic_adj_poisson_rhs(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd,
                          const double *lamphi, double scale, double *rhs);
/* *out (device) += dL/dV of the electrode with this label: sum over its
 * nodes of lamphi minus sum over their plasma neighbours of a mu, with a the
 * Laplacian's coefficient eps0/h^2 of that direction. */
int lepic_adj_bias(lepic_ctx *ctx, const lepic_grid *g, const int32_t *bcnd, int32_t label,
                   const double *lamphi, const double *mu, double eps0, double *out);
/* Deposit adjoint of one species of charge q and weight Nm: dL/dx_p +=
 * -q Nm/(h1 h2 h3) sum_c kq_c lamb_c dw_c/dx (the transpose of
 * deposit_particle_set_to_np_thread + build_rho_from_np taken as an
 * operator of the positions), lamb = dL/drho on nodes. */
int lepic_adj_deposit(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                      lepic_adjoint_set *adj, const double *kq, const double *lamb, double q,
                      double Nm);
/* The wall-flux loss term of one step: *acc (device, may be NULL) +=
 * coef * sum over live particles of vx * hat(x / h1 - (plane - 1)), the
 * cloud-in-cell weight of node plane `plane` (1-based); with adj, dL/dx and
 * dL/dv += scale * the term's derivatives. */
int lepic_plane_flux(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state, int32_t plane,
                     double coef, double *acc, lepic_adjoint_set *adj, double scale);
/* The same term through a node plane normal to any axis (0 x, 1 y, 2 z):
 * coef * sum over live particles of v_a * hat(x_a / h_a - (plane - 1)),
 * 1 <= plane <= n_a + 1 (the ion current into an electrode face normal to
 * y or z, e.g. the anode at z = 0 of the annular channel). */
int lepic_plane_flux_axis(lepic_ctx *ctx, const lepic_grid *g, const lepic_pset *state,
                          int32_t axis, int32_t plane, double coef, double *acc,
                          lepic_adjoint_set *adj, double scale);

/* Frozen wall events for replays of a differentiable run: the slots whose
 * recorded loss step (deaths, device int64 per slot, 0: never) equals `step`
 * become holes (dead = 2).  With lepic_move_boris and lepic_deposit in place
 * of the fused push it evaluates the run whose derivative the adjoint
 * computes. */
int lepic_apply_deaths(lepic_ctx *ctx, lepic_pset *p, const int64_t *deaths, int64_t step);

#ifdef __cplusplus
}
#endif
#endif
