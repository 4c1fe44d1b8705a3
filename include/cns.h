/*
 * cns.h -- C ABI of the MI355X (gfx950) explicit compressible Navier-Stokes
 * time advance: right-hand side, ghost fills, sponge and RK stage updates of
 * simone-silvestri/CudaNavierStokes, as a drop-in for its solver entry
 * points.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Fields are float64 host arrays in the reference layout
 *    idx = i + j*mx + k*mx*my, x unit-stride (src/globals.h:60-61).
 *  - State variables are the reference's (rho, u, v, w, rhoE): density,
 *    velocity, total energy per volume (src/init.cpp:139-146).
 *  - Axes follow the reference: x may be stretched and carry isothermal walls,
 *    y and z are periodic, the forcing acts along z (src/cuda_rhs.cu:382-394).
 *  - Names avoid the reference's case macros (mx, Lx, Re, forcing, nsteps,
 *    ...), so this header can be included after src/globals.h.
 *  - Every function returns CNS_OK (0) or a negative error code; the message
 *    of the last failure is cns_last_error(handle) (or cns_last_error(NULL)
 *    when no handle exists).  Nothing calls exit().
 *  - The library owns device memory; the caller owns host buffers.  A handle
 *    is used by one host thread at a time; handles are independent.
 */
#ifndef CNS_H_
#define CNS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CNS_OK 0
#define CNS_ERR_ARG (-1)
#define CNS_ERR_HIP (-2)
#define CNS_ERR_STATE (-3)
#define CNS_ERR_UNSUPPORTED (-4)
#define CNS_ERR_NOMEM (-5)
#define CNS_ERR_IO (-6)

/* time-advance schemes */
#define CNS_RK3_LOWSTORAGE 0 /* Spalart-Moser-Rogers RK3, src/cuda_main.cu:109-186 */
#define CNS_RK3_KUTTA 1      /* Kutta RK3, src/cuda_main.cu:44-107 */
#define CNS_RK4 2            /* classic RK4 (no reference counterpart) */

/* storage precision of the conserved variables */
#define CNS_FP64 0
#define CNS_FP32 1

/* Run-time form of the reference's compile-time case macros
 * (src/globals.h:14-58, src/sponge.h:4-17). */
typedef struct cns_config {
    int nx, ny, nz;       /* grid points (mx_tot, my_tot, mz_tot) */
    int stencil_size;     /* advective half width s: order 2s (stencilSize) */
    int stencil_visc;     /* viscous half width sv <= s (stencilVisc) */
    int periodic_x;       /* periodicX */
    int nonuniform_x;     /* nonUniformX: tanh-stretched x with metrics */
    int forcing_z;        /* forcing: dp/dz controller driving bulk velocity 1 */
    int sponge;           /* sponge relaxation (boundaryLayer's addSponge) */
    int scheme;           /* CNS_RK3_LOWSTORAGE (lowStorage), CNS_RK3_KUTTA, CNS_RK4 */
    int check_cfl;        /* checkCFLcondition */
    int check_bulk;       /* checkBulk */
    int fix_y_work;       /* 0: y viscous work exactly as src/cuda_rhs.cu:175; 1: corrected */
    int precision;        /* CNS_FP64 or CNS_FP32 */
    int compensated;      /* CNS_FP32: carry a compensation (low) part of the state through the RK sums */
    int device;           /* HIP device ordinal */
    double len_x, len_y, len_z;  /* Lx, Ly, Lz */
    double reynolds, prandtl, mach, gamma; /* Re, Pr, Ma, gam */
    double visc_exp;      /* viscexp: mu = T^visc_exp / Re */
    double cfl;           /* CFL */
    double stretch;       /* tanh stretching of x (stretch) */
    double twall_bot, twall_top; /* wall temperatures (TwallBot, TwallTop) */
    double dpdz0;         /* initial forcing (src/cuda_utils.cu:68-70) */
    int boundary_layer;   /* boundaryLayer: with periodic_x 0, an adiabatic no-slip wall at x = 0, linear
                             extrapolation at the x top and both z ends, the wall-clustered half-tanh grid
                             (src/boundary_condition_{x,z}.h, src/init.cpp:41-50); FP64 storage, lone
                             solvers (not in batches), nx a multiple of 16, nz > 2*stencil_size */
    int wall_blowing;     /* perturbed: wall-normal blowing at x = 0 on z planes 90..130
                             (src/perturbation.h:25-53); named apart from the reference's macro */
} cns_config;

typedef struct cns_solver cns_solver;

/* Defaults: 64^3 periodic box of side 2*pi, s = sv = 2, SMR RK3, Re 1600,
 * Pr 0.71, Ma 0.1, gamma 1.4, viscexp 1, CFL 0.5, checks every 10 steps. */
int cns_config_init(cns_config *cfg);

/* replaces setGPUParameters + initSolver (src/cuda_utils.cu:49,415) */
int cns_create(const cns_config *cfg, cns_solver **out);
/* replaces clearSolver (src/cuda_utils.cu:521) */
int cns_destroy(cns_solver *s);
const char *cns_last_error(const cns_solver *s);

/* grid of initGrid/derivGrid (src/init.cpp:32-92,258-277); any pointer may
 * be NULL.  x, xp (dxi/dx), xpp (d2x/dxi2), dxv (cell widths) have mx
 * entries, y my, z mz. */
int cns_get_grid(const cns_solver *s, double *x, double *y, double *z, double *xp, double *xpp, double *dxv);

/* Initial conditions written into host arrays (src/init.cpp:94-148):
 * Taylor-Green vortex of initCHIT; compressible channel with a laminar
 * Poiseuille profile and SplitMix64 uniform perturbations of amplitude
 * amp*ubulk on every velocity component. */
int cns_fill_taylor_green(const cns_solver *s, double *r, double *u, double *v, double *w, double *e);
int cns_fill_channel(const cns_solver *s, uint64_t seed, double ubulk, double amp, double *r, double *u,
                     double *v, double *w, double *e);

/* The same initial conditions evaluated on the device and written straight
 * into the state (no host arrays, no transfer); like cns_set_state they reset
 * time.  They agree with the host fills to rounding (device sin/cos). */
int cns_init_taylor_green(cns_solver *s);
int cns_init_channel(cns_solver *s, uint64_t seed, double ubulk, double amp);

/* replaces copyField(0): host (rho,u,v,w,rhoE) -> device; resets time */
int cns_set_state(cns_solver *s, const double *r, const double *u, const double *v, const double *w,
                  const double *e);
/* replaces copyField(1): device -> host (rho,u,v,w,rhoE) */
int cns_get_state(cns_solver *s, double *r, double *u, double *v, double *w, double *e);

/* one right-hand-side evaluation of the current state (calcRHS,
 * src/cuda_main.cu:15-42), five host arrays of the conserved-variable RHS */
int cns_rhs(cns_solver *s, double *rr, double *ru, double *rv, double *rw, double *re);

/* No counterpart in the reference: cns_rhs for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid allocation).  One right-hand side of the state q into rhs, both
 * device arrays the caller owns in the solver's device layout (five fields
 * of nx*ny*nz values: rho, rho u, rho v, rho w, rho E; FP64, or for FP32
 * storage the FP32 offsets from the solver's reference state), queued on the
 * solver's stream (cns_sync waits).  Allocates nothing: the dilatation field
 * goes to a file-scope device scratch array of 16 MiB per GPU (held until
 * the process ends), so grids above 2^21 FP64 or 2^22 FP32 points are
 * CNS_ERR_ARG.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads) race. */
int cns_rhs_noalloc(cns_solver *s, const void *q, void *rhs);

/* replaces runSimulation[LowStorage] (src/cuda_main.cu:44-186): n_steps
 * steps.  Optional outputs of length n_steps, host or device arrays (the
 * reference passes device arrays, src/main.h:35): kin/enst receive the
 * reference's par1/par2 bulk statistics at steps with istep % check_bulk == 0
 * (NaN elsewhere; par2 is NaN without forcing), time the simulation time
 * after each step. */
int cns_step(cns_solver *s, int n_steps, double *kin, double *enst, double *time);

/* sponge strengths and reference state of calculateSponge
 * (src/sponge.cu:83-168) from a boundary-layer profile on y / delta
 * (n points, increasing: the layout of blasius1D/{x,r,u,w}Prof.bin: density,
 * wall-normal and streamwise velocity), interpolated by a natural cubic
 * spline at x / sqrt(1 + z/Re) on the grid x_coord[n_x], z_coord[n_z].  zones: 9 values
 * (strength, length, exponent) of the top, inlet and outlet zones, or NULL
 * for src/sponge.h:4-17.  Outputs sigma_x[n_x], sigma_z[n_z] and the
 * reference conserved fields [n_z][n_x] for cns_set_sponge.  Host
 * computation only. */
int cns_sponge_from_profile(int n, const double *y, const double *rho, const double *u_wallnormal,
                            const double *w_streamwise, int n_x, const double *x_coord, int n_z,
                            const double *z_coord, double len_x, double len_z, double reynolds, double gamma,
                            double mach, const double *zones, double *sigma_x, double *sigma_z, double *rref,
                            double *ruref, double *rvref, double *rwref, double *reref);

/* sponge of addSponge (src/sponge.cu:31-41): strengths sigma_x[mx],
 * sigma_z[mz] and reference conserved fields [mz][mx] */
int cns_set_sponge(cns_solver *s, const double *sigma_x, const double *sigma_z, const double *rref,
                   const double *ruref, const double *rvref, const double *rwref, const double *reref);

/* current dt, dp/dz and simulation time */
int cns_get_scalars(cns_solver *s, double *dt, double *dpdz, double *time);

/* Diagnostics of the current state, evaluated on the device in FP64 (FP32
 * storage is read as stored value + low part + offset).
 *
 * cns_channel_profiles: the wall-normal profiles of calcAvgChan
 * (src/init.cpp:150-208), averaged over y-z planes; prof receives 11*nx
 * values, prof[c*nx + i] for the columns of its prof.txt: 0 x, 1 mean rho,
 * 2-4 Favre means <rho u>/<rho>, <rho v>/<rho>, <rho w>/<rho>, 5 mean rhoE,
 * 6-10 mean squared deviations of rho, u, v, w, rhoE from those means.
 *
 * cns_friction_reynolds: the average friction Reynolds number printRes
 * prints (src/init.cpp:210-240) -- per y-z row, dw/dx at the first node from
 * the advective-order stencil on w mirrored evenly about the bottom face,
 * u_tau = sqrt(mu_w |dw/dx| / rho) with mu_w = 1/Re, averaged u_tau rho/mu_w.
 *
 * cns_flow_stats (no reference counterpart): out[0] kinetic energy
 * int(rho |u|^2/2) / int(rho), out[1] enstrophy int(rho |curl u|^2/2) /
 * int(rho), out[2] volume mean of (div u)^2, out[3] volume mean density;
 * velocity derivatives are the RHS's viscous-order central differences and
 * ghost rules, integrals use the cell widths in x.
 * All three overwrite the first right-hand-side buffer (free between steps). */
int cns_channel_profiles(cns_solver *s, double *prof);
int cns_friction_reynolds(cns_solver *s, double *re_tau);
int cns_flow_stats(cns_solver *s, double *out);

/* The statistics of cns_flow_stats folded into the step loop: n_steps steps
 * taken as n_steps/every successive cns_step(every) calls (a CFL check opens
 * each), each followed by the statistics of the state, reduced on the
 * device with nothing read back until the end; stats receives 4 values per
 * sample (n_steps/every samples; n_steps must be a multiple of every). */
int cns_step_stats(cns_solver *s, int n_steps, int every, double *stats);

/* Field files of writeField / initField (src/init.cpp:13-30), written and
 * read like saveFileMPI / readFileMPI (src/comm.cpp:205-279):
 * <dir>/<c>.<timestep, %07d>.bin for c = r, u, v, w, e, each the whole grid
 * as native-endian float64 in [nz][ny][nx] order (x unit-stride), variables
 * (rho, u, v, w, rhoE).  dir is the reference's "fields" directory and must
 * exist.  cns_read_fields sets the state as cns_set_state does (time restarts
 * at zero, as in the reference's restart); a missing file or one whose size
 * is not nx*ny*nz*8 bytes is CNS_ERR_IO and leaves the solver unchanged. */
int cns_write_fields(cns_solver *s, const char *dir, int timestep);
int cns_read_fields(cns_solver *s, const char *dir, int timestep);

/* grid files of initGrid (src/init.cpp:78-88): <dir>/x.bin, y.bin, z.bin,
 * the node coordinates as float64 (nx, ny, nz values) */
int cns_write_grid(const cns_solver *s, const char *dir);

/* Time-advance control besides the state: with it a run resumed from field
 * files continues as the uninterrupted run would (to the rounding of the
 * u = rho u / rho round trip of the file format).  No reference counterpart:
 * the reference's restart starts time at zero and dp/dz at its initial value
 * (src/cuda_main.cu:286-291, src/cuda_utils.cu:68-70).  Set it after the
 * state, which resets it. */
typedef struct cns_control {
    double time;          /* simulation time */
    double dt;            /* time step set by the last CFL check */
    double dpdz;          /* forcing */
    double visc_max;      /* viscous maximum gathered for the next CFL check */
    int64_t steps;        /* batched steps since the last cns_step / state set (CFL check phase) */
    int visc_from_state;  /* 1: the next CFL check takes the viscous limit from the state */
    int reserved;
} cns_control;
int cns_get_control(cns_solver *s, cns_control *c);
int cns_set_control(cns_solver *s, const cns_control *c);

/* Wait for all work queued on the solver's stream. */
int cns_sync(cns_solver *s);

/* Bracket every fused RHS/stage-update kernel launch with HIP events on the
 * solver's stream (enable=1); cns_profile_read returns the summed kernel
 * time in ms and the number of launches since the last enable. */
int cns_profile_enable(cns_solver *s, int enable);
int cns_profile_read(cns_solver *s, double *total_ms, int *launches);

/* n_steps steps exactly as cns_step queues them with all outputs (CFL and
 * forcing checks, the bulk statistics every check_bulk steps, the time
 * array), inputs resident on the device, timed with HIP events on the
 * solver's stream; *ms receives the elapsed time.  Only the copies of the
 * per-step arrays to the host are left out. */
int cns_time_steps(cns_solver *s, int n_steps, double *ms);

/* Launch shape of the stage kernels: each workgroup marches a pencil of
 * this many z planes (1..nz; results depend on it only through rounding:
 * every point follows the same arithmetic, but the compiled loop copies may
 * fuse multiply-adds differently).  cns_tune times
 * a few steps per candidate length (8..256) on the current state, keeps the
 * fastest and restores the state, time and step control bit for bit;
 * *pencil (may be NULL) receives the choice.  No reference counterpart (the
 * reference's z chunking, nDivZ, is a compile-time constant). */
int cns_set_pencil(cns_solver *s, int pencil);
int cns_tune(cns_solver *s, int *pencil);

/* bytes of device memory held by the solver */
int cns_device_bytes(const cns_solver *s, uint64_t *bytes);

/* free and total device memory of a device (the query of checkGpuMem,
 * src/cuda_utils.cu:780-813) */
int cns_mem_info(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* No counterpart in the reference (one solver per process there).  Device
 * blocks that cns_destroy and cns_batch_destroy release stay cached per
 * device and size for the next solver or batch of the same sizes (freeing
 * device memory waits for the whole device); at most CNS_CACHE_BYTES bytes
 * (environment, default 16 GiB) per device.  This frees the cached blocks of
 * one device, or of every device for device < 0. */
int cns_release_cache(int device);

/* ---- Batched advance of many solvers (ensembles, the serving loop) ----
 * No counterpart in the reference, which runs one flow per process.  A batch
 * advances its members together: members that share a kernel template
 * (stencil widths, x boundary and grid kind, viscosity law, storage
 * precision) and time scheme form a group, and each group advances with one
 * launch per RK stage covering every member's tiles (plus one dilatation
 * launch per stage and four small control launches per step).  A step of the
 * batch is captured once as a hipGraph and replayed.  Every member follows
 * the arithmetic of cns_step on its own (same tiles, operators and step
 * control); the batched kernels are compiled separately and fuse a few
 * multiply-adds differently, so a member's result after n batched steps
 * equals its result after cns_step(n) to rounding.  Per member,
 * steps are counted from its last cns_set_state or cns_step, across
 * successive batches (the CFL check and forcing controller run on steps 0,
 * check_cfl, 2 check_cfl, ...), so a member may move between batches.
 * Members must have a state, live on one device and not be used on their
 * own while the batch exists. */
typedef struct cns_batch cns_batch;
int cns_batch_create(cns_solver *const *members, int n_members, cns_batch **out);
/* Pick per group the z-pencil length, the persistent workgroup count and the
 * XCD-contiguous tile order by timing RHS launches whose outputs go to
 * scratch buffers (the members' states are not changed); then re-captures. */
int cns_batch_tune(cns_batch *b);
/* n_steps steps of every member (graph replays on the batch stream) */
int cns_batch_step(cns_batch *b, int n_steps);
/* the same without waiting: cns_batch_launch queues n_steps and returns;
 * cns_batch_wait waits for everything queued.  Members must not be read
 * before the wait (cns_batch_step, _tune, _time_steps and _destroy wait
 * first). */
int cns_batch_launch(cns_batch *b, int n_steps);
int cns_batch_wait(cns_batch *b);
/* the same, timed with HIP events on the batch stream */
int cns_batch_time_steps(cns_batch *b, int n_steps, double *ms);
/* tuned launch shape of group g: pencil length, workgroups, XCD ordering, work items */
int cns_batch_group_info(const cns_batch *b, int g, int *n_members, int *pencil, int *grid, int *remap,
                         int *work_items);
int cns_batch_groups(const cns_batch *b);
/* set group g's launch shape by hand (pencil length >= 1, persistent
 * workgroups >= 1 or 0 for one workgroup per work item, XCD-contiguous order
 * 0/1) instead of cns_batch_tune */
int cns_batch_set_shape(cns_batch *b, int g, int pencil, int grid, int remap);
int cns_batch_destroy(cns_batch *b);
const char *cns_batch_last_error(const cns_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* CNS_H_ */
