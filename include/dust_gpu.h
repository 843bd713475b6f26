/*
 * dust_gpu.h -- C ABI of libdust_gpu.so, the MI355X (gfx950) implementation of
 * DUST's per-timestep hot path: the octree FMM evaluation of the vortex-particle
 * wake (src/fmm) and the dense AIC assembly plus LU solve (src/linsys).
 *
 * Plain C types only (pointers, sizes, int status codes) so that Fortran binds
 * it with ISO_C_BINDING and Python with ctypes.  Every function returns an int
 * status: 0 on success, > 0 a LAPACK-style `info`, < 0 one of DUST_E_*.
 *
 * Array conventions follow the Fortran side: 3-vectors per particle or panel
 * are stored as Fortran real(wp) :: x(3,n), i.e. interleaved (x1,y1,z1,x2,...);
 * matrices are column-major with a leading dimension; pivot indices are
 * 1-based as returned by LAPACK xGETRF.
 *
 * Host-pointer entry points are synchronous: outputs are on the host when the
 * call returns.  Entry points with the _dev suffix take device pointers, are
 * asynchronous on the handle's stream (dust_fmm_stream) and leave results in
 * device memory.  A handle is single-threaded and not re-entrant.
 */
#ifndef DUST_GPU_H
#define DUST_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DUST_OK 0
#define DUST_E_ARG (-1)        /* invalid argument / API misuse              */
#define DUST_E_HIP (-2)        /* HIP runtime error                          */
#define DUST_E_OUTSIDE (-3)    /* particle outside the octree box            */
#define DUST_E_CAPACITY (-4)   /* size beyond what the handle was built for  */
#define DUST_E_NOMEM (-5)      /* device allocation failed                   */
#define DUST_E_UNSUPPORTED (-6)

/* Human-readable text of the last error raised on the calling thread. */
const char *dust_last_error(void);

/* ======================================================================== */
/* FMM: src/fmm/mod_octree.f90 + src/fmm/mod_multipole.f90                   */
/* ======================================================================== */

/* Particle core (dust_fmm_params.kernel): kernel_coeffs,
 * src/aeroel/mod_vortpart.f90:260-299.  Rosenhead is the reference's active
 * branch (269-271); Gaussian is its commented alternative "from Alvarez"
 * (283-291), c = -erf(r/(sqrt2 s))/r^3 + 2 exp(-r^2/(2 s^2))/(r^2 sqrt(2 pi) s),
 * d = (1/r) dc/dr, zero at r <= 1e-13.  Only the particle-particle terms
 * (P2P, direct sum) depend on it; the expansions use the Rosenhead-regularised
 * Green's function of `delta` either way, as in the reference. */
#define DUST_KERNEL_ROSENHEAD 0
#define DUST_KERNEL_GAUSSIAN 1

/* Evaluation forms (dust_fmm_params.precision). */
#define DUST_FMM_FP64 0   /* plain FP64 path: the drop-in                    */
#define DUST_FMM_TUNED 1  /* FP32 P2P with FP64 near pairs + GEMM M2L        */

typedef struct {
  /* initialize_octree arguments, mod_octree.f90:251-260 */
  double box_length;     /* edge of a level-1 box                           */
  int nbox[3];           /* level-1 boxes per direction                      */
  double origin[3];      /* octree origin (xmin)                             */
  int nlevels;           /* uniform tree depth                               */
  int min_part;          /* min_octree_part                                  */
  int degree;            /* multipole expansion order p                      */
  double delta;          /* regularisation radius of the M2L kernel          */
                         /* derivatives (sim_param%RankineRad, dust.f90:374)  */
  /* sim_param flags read by the reference FMM (mod_octree.f90:1190-1408) */
  int use_vs;            /* vortex stretching                               */
  int use_vd;            /* PSE vortex diffusion                             */
  int use_tv;            /* turbulent viscosity (needs use_vd)               */
  double nu_inf;         /* kinematic viscosity                              */
  double u_inf[3];       /* uniform free stream; gusts: dust_fmm_set_wind   */
  int precision;         /* DUST_FMM_FP64 or DUST_FMM_TUNED                  */
  int kernel;            /* DUST_KERNEL_ROSENHEAD (0) or DUST_KERNEL_GAUSSIAN */
  int use_divfilt;       /* with use_vs: apply_multipole also forms rotu, the */
                         /* divergence filter's curl of the velocity          */
                         /* (mod_octree.f90:1207-1212,1274-1336), read with   */
                         /* dust_fmm_rotu                                     */
  int reserved[5];
} dust_fmm_params;

typedef struct dust_fmm_s *dust_fmm_t;

/* initialize_octree (mod_octree.f90:251): builds the per-level kernel
 * derivative tables (compute_der_ker, mod_multipole.f90:388) and the device
 * state.  `capacity` is the maximum particle count the handle accepts. */
int dust_fmm_create(const dust_fmm_params *prm, int capacity, dust_fmm_t *out);

/* destroy_octree (mod_octree.f90:694). */
int dust_fmm_destroy(dust_fmm_t h);

/* Particle input, the SoA image of t_vortpart (mod_vortpart.f90:85-97):
 * cen(3,n), dir(3,n), mag(n), r_Vortex(n).  Host pointers. */
int dust_fmm_set_particles(dust_fmm_t h, int n, const double *cen, const double *dir,
                           const double *mag, const double *r_vortex);
/* Same with device pointers; the data are copied on the handle's stream. */
int dust_fmm_set_particles_dev(dust_fmm_t h, int n, const double *cen, const double *dir,
                               const double *mag, const double *r_vortex);

/* sort_particles (mod_octree.f90:703): bins the particles on the finest
 * level, flags leaves/branches, builds the P2P lists.  Returns the number of
 * leaves in *nleaves (may be NULL).  DUST_E_OUTSIDE if a particle lies
 * outside the box (mod_octree.f90:757-762). */
int dust_fmm_sort_particles(dust_fmm_t h, int *nleaves);

/* Particle redistribution of sort_particles (mod_octree.f90:827-894; the
 * sim_param switches particles_redistribution, octree_level_solid and
 * particles_redistribution_ratio of mod_sim_param.f90:462-467,792-800).
 * While on, every later sort frees, in each finest cell not near a wall, the
 * particles whose magnitude falls below the cell's mean / ratio, spreading
 * each one's vorticity over the cell's other particles; freed particles are
 * left out of the tree and get zero results from apply_multipole.  A level
 * lvl_solid (1-based) cell is near a wall when it or a neighbour holds one of
 * the m element centres elem_cen(3,m) (host or device memory, copied here;
 * centres outside the box are skipped).  on = 0 switches it off.
 * DUST_E_ARG for lvl_solid outside 1..nlevels or ratio <= 0. */
int dust_fmm_set_redistribution(dust_fmm_t h, int on, int lvl_solid, double ratio, int m,
                                const double *elem_cen);
/* Outcome of the last sort in input order: keep[i] = 0 for a particle the
 * redistribution freed (the reference sets its free flag), the strengths
 * dir(3,n), mag(n) after it (the input ones without a redistribution), and
 * *n_kept = the reference's n_prt after sort_particles.  Host pointers, any
 * may be NULL. */
int dust_fmm_redistributed(dust_fmm_t h, int *n_kept, unsigned char *keep, double *dir,
                           double *mag);
/* Same into device memory, on the handle's stream. */
int dust_fmm_redistributed_dev(dust_fmm_t h, unsigned char *keep, double *dir, double *mag);

/* calculate_multipole (mod_octree.f90:1002): P2M, M2M, M2L, L2L. */
int dust_fmm_calculate_multipole(dust_fmm_t h);

/* apply_multipole (mod_octree.f90:1139), particle terms plus free stream:
 * L2P + P2P, writes vel(3,n), stretch(3,n) (zero unless use_vs/use_vd) and
 * turbvisc(n) in input particle order.  Host pointers (any may be NULL). */
int dust_fmm_apply_multipole(dust_fmm_t h, double *vel, double *stretch, double *turbvisc);
/* Asynchronous form: results stay on the device (see dust_fmm_results_dev). */
int dust_fmm_apply_multipole_dev(dust_fmm_t h);
/* rotu(3,n) of the last apply_multipole, input particle order (the
 * t_vortpart%rotu the reference writes with use_divfilt: the curl of the
 * local expansion's velocity plus compute_rotu / (4 pi) of every P2P source,
 * mod_vortpart.f90:236-256; zero for particles freed by a redistribution).
 * DUST_E_ARG unless the handle was created with use_vs and use_divfilt.
 * Host pointer (synchronous) / device pointer on the handle's stream. */
int dust_fmm_rotu(dust_fmm_t h, double *rotu);
int dust_fmm_rotu_dev(dust_fmm_t h, double *rotu);
/* Device-to-device copy of the last results (input particle order). */
int dust_fmm_results_dev(dust_fmm_t h, double *vel, double *stretch, double *turbvisc);

/* apply_multipole_panels (mod_octree.f90:1434-1510) with the panel binning
 * of sort_particles (777-820, the use_fmm_pan path): the particle velocity at
 * m panel centres cen(3,m) from the tree and expansions of the last
 * sort_particles + calculate_multipole, added into uvort(3,m): L2P of the
 * leaf that contains the centre's finest cell plus the particles of that
 * leaf's P2P list (compute_vel / 4 pi).  DUST_E_OUTSIDE for a centre outside the
 * box.  Host pointers (synchronous) / device pointers on the handle's
 * stream (synchronises to report DUST_E_OUTSIDE). */
int dust_fmm_apply_multipole_panels(dust_fmm_t h, int m, const double *cen, double *uvort);
int dust_fmm_apply_multipole_panels_dev(dust_fmm_t h, int m, const double *cen,
                                        double *uvort);

/* One full pass (sort + multipole + apply) on device-resident particles. */
int dust_fmm_pass_dev(dust_fmm_t h, int *nleaves);

/* Several independent cases, one pass each (sort + multipole + apply on the
 * device-resident particles of every handle), advanced in lockstep: every
 * phase of the pass is launched for all handles, each on its own stream,
 * before one synchronisation of all of them, so the cases' kernels overlap
 * and the host round trips are shared.  Returns when all passes are done;
 * nleaves[i] (may be NULL) receives each case's leaf count.  The handles must
 * be distinct.  No counterpart in the reference (its cases are separate
 * runs); for many-case workloads. */
int dust_fmm_pass_batch_dev(int nh, dust_fmm_t *handles, int *nleaves);

/* Graph form of the pass: the sequence of dust_fmm_pass_dev
 * (sort_particles, calculate_multipole, apply_multipole,
 * src/fmm/mod_octree.f90:703-1429, as called once per step from the time
 * loop, src/dust.f90:504-1160) with every count on the device, so that a
 * stream capture can record it and a graph can replay it for a wake whose
 * particle count and tree change from step to step.  No counterpart in the
 * reference; for time loops captured in a HIP graph.
 *
 * dust_fmm_graph_reserve: fixes the capacities from the last completed pass
 * on the handle (its active cells, leaves, cells per level and P2P items,
 * each times `slack` plus 64; particles up to the handle capacity) and
 * allocates for them.  Not with the particle redistribution or a gust
 * (DUST_E_ARG).
 *
 * dust_fmm_pass_graph_dev: enqueues the pass on the handle's stream without
 * host synchronisation, allocation or memset: particle count read from
 * *n_dev, inputs read from cen/dir/mag/rv (device, the layouts of
 * dust_fmm_set_particles), velocity and stretching written to vel/stretch
 * (device, (3, n); stretch may be NULL).  Capturable; the FP64 form gives the
 * same values as dust_fmm_pass_dev bit for bit, the tuned form to its M2L's
 * atomic order.  The host-side entry points that read a pass's lists
 * (dust_fmm_leaves, dust_fmm_results_dev, apply_multipole_panels) need a
 * pass of dust_fmm_pass_dev after it.
 *
 * dust_fmm_graph_capture: captures that sequence into a graph owned by the
 * handle (the pointers are fixed in it: refill the same buffers between
 * launches); dust_fmm_graph_launch replays it on the handle's stream.
 *
 * dust_fmm_graph_status: synchronises the stream and reports the last graph
 * pass: DUST_E_CAPACITY when a list outgrew its capacity or *n_dev was
 * outside 0..capacity (every kernel then did nothing past that point: reserve
 * again after a dust_fmm_pass_dev and capture again), DUST_E_OUTSIDE for a
 * particle outside the box (results invalid).  counts (may be NULL, 4 ints):
 * particles, active cells, leaves, P2P items of that pass. */
int dust_fmm_graph_reserve(dust_fmm_t h, double slack);
int dust_fmm_pass_graph_dev(dust_fmm_t h, const int *n_dev, const double *cen, const double *dir,
                            const double *mag, const double *rv, double *vel, double *stretch);
int dust_fmm_graph_capture(dust_fmm_t h, const int *n_dev, const double *cen, const double *dir,
                           const double *mag, const double *rv, double *vel, double *stretch);
int dust_fmm_graph_launch(dust_fmm_t h);
int dust_fmm_graph_status(dust_fmm_t h, int *counts);

/* Knobs.  Tuned form: "near_factor" -- pairs closer than near_factor times
 * the source core radius are evaluated in FP64 (default 1); "m2l_chunk" --
 * interaction offsets per M2L job, 1..189 (default 21): larger chunks mean
 * fewer atomic additions into the local expansions and fewer, longer jobs.
 * Both forms: "seg_iters" -- P2P source iterations per lane in one work
 * item, 64..65536 (default 1536).  No counterpart in the reference. */
int dust_fmm_set_tuning(dust_fmm_t h, const char *name, double value);
int dust_fmm_get_tuning(dust_fmm_t h, const char *name, double *value);

/* Autotuner of the performance knobs (seg_iters; m2l_chunk in the tuned
 * form): times `reps` full passes (sort + multipole + apply) over the
 * particles already set in the handle for each setting of a small grid and
 * keeps the fastest; *best_ms (may be NULL) receives its time per pass.
 * near_factor, which changes the results, is not touched.  The handle's
 * outputs afterwards are those of the last timed pass: run a pass before
 * reading them.  No counterpart in the reference (a one-off cost per run
 * configuration, e.g. at the first time step). */
int dust_fmm_autotune(dust_fmm_t h, int reps, double *best_ms);

/* HIP stream (hipStream_t) the handle launches on. */
void *dust_fmm_stream(dust_fmm_t h);
int dust_fmm_synchronize(dust_fmm_t h);

/* Leaf membership dump after sort_particles: per leaf its level (1-based),
 * cartesian index (1-based, (3,nleaves)) and particle count; `members` gets
 * the input indices (0-based) of each leaf's particles, leaf after leaf.
 * Any pointer may be NULL. */
int dust_fmm_leaves(dust_fmm_t h, int *level, int *ijk, int *npart, int *members);

/* Finest-level cell index (1-based, (3,n)) of every particle, i.e.
 * ceiling((x - xmin)/cell_size) of mod_octree.f90:755. */
int dust_fmm_cell_index(dust_fmm_t h, int *ijk);

/* Algorithmic work of the last sort: P2P pairs and active M2L pairs. */
int dust_fmm_work(dust_fmm_t h, double *p2p_pairs, double *m2l_pairs);

/* Stage timing with HIP events recorded on the handle's stream around every
 * stage of a pass: enable, then read accumulated milliseconds and launch
 * counts per stage (DUST_FMM_STAGE_*). */
#define DUST_FMM_STAGE_SORT 0
#define DUST_FMM_STAGE_P2M 1
#define DUST_FMM_STAGE_M2M 2
#define DUST_FMM_STAGE_M2L 3
#define DUST_FMM_STAGE_L2L 4
#define DUST_FMM_STAGE_P2P 5   /* P2P items (partial sums per target) */
#define DUST_FMM_STAGE_L2P 6   /* L2P + sum of the partials + free stream */
#define DUST_FMM_NSTAGES 7
int dust_fmm_set_timing(dust_fmm_t h, int enable);
int dust_fmm_stage_times(dust_fmm_t h, double *ms, int *launches);

/* Direct O(N^2) evaluation of the same particle terms, the reference's
 * fmm=F branch (mod_wake.f90:1050-1113).  Host pointers, synchronous. */
int dust_vortpart_direct(int n, const double *cen, const double *dir, const double *mag,
                         const double *r_vortex, int use_vs, int use_vd, double nu_inf,
                         const double u_inf[3], double *vel, double *stretch);
/* Same with the particle core chosen (DUST_KERNEL_*). */
int dust_vortpart_direct_k(int n, const double *cen, const double *dir, const double *mag,
                           const double *r_vortex, int use_vs, int use_vd, double nu_inf,
                           const double u_inf[3], int kernel, double *vel, double *stretch);

/* ======================================================================== */
/* Wind: variable_wind, src/aeroel/mod_wind.f90:66-121                      */
/* ======================================================================== */

/* Gust forms (sim_param%GustType with sim_param%use_gust): the reference's
 * 'ACM' one-minus-cosine gust and its 'linear' test form; any other type,
 * including the option's default string 'AMC', adds nothing (the select
 * case of mod_wind.f90:82-117 matches only 'ACM' and 'linear'). */
#define DUST_GUST_NONE 0
#define DUST_GUST_ACM 1
#define DUST_GUST_LINEAR 2

typedef struct {
  double u_inf[3];               /* sim_param%u_inf                                  */
  int gust_type;                 /* DUST_GUST_*                                      */
  int reserved0;
  double origin[3];              /* gust_origin                                      */
  double front_direction[3];     /* gust_front_direction (normalised here)          */
  double front_speed;            /* gust_front_speed                                 */
  double u_des;                  /* gust_u_des                                       */
  double perturb_direction[3];   /* gust_perturb_direction: read, unused by both    */
                                 /* forms, as in the reference                       */
  double gradient;               /* gust_gradient                                    */
  double start_time;             /* gust_time (gust_start_time)                      */
  int reserved[4];
} dust_wind_params;

/* wind(3,n) = variable_wind(pos(:,i), time) at n points pos(3,n): u_inf,
 * plus for an ACM gust u_des/2 (1 - cos(pi s / gradient)) on every component
 * where the penetration s = -(pos - (origin + front_speed fd (time -
 * start_time))) . fd behind the moving front lies in [0, 2 gradient)
 * (fd = front_direction normalised), or for the linear form 0.1 pos_x on z.
 * Host pointers / device pointers on `stream`. */
int dust_variable_wind(int n, const double *pos, const dust_wind_params *w, double time,
                       double *wind);
int dust_variable_wind_dev(int n, const double *pos, const dust_wind_params *w, double time,
                           double *wind, void *stream);

/* FMM: the free-stream term of apply_multipole as the reference forms it,
 * variable_wind(pos, time) at every particle (mod_octree.f90:1376), for the
 * next passes: w->u_inf replaces the params' u_inf; time = sim_param%time.
 * w = NULL returns to the params' uniform u_inf. */
int dust_fmm_set_wind(dust_fmm_t h, const dust_wind_params *w, double time);

/* ======================================================================== */
/* Element <-> particle direct terms: src/fmm/mod_octree.f90:1350-1401,     */
/* get_vort_vel (src/aeroel/mod_surfpan.f90:962-977, mod_vortlatt.f90:676) */
/* ======================================================================== */

/* Element kinds (the element types behind apply_multipole's lists). */
#define DUST_ELEM_SURFPAN 0 /* t_surfpan: doublet + source (compute_vel_surfpan, mod_surfpan.f90:641) */
#define DUST_ELEM_RING 1    /* t_vortlatt / t_liftlin / wake panels: vortex ring (velocity_calc_doublet) */
#define DUST_ELEM_LINE 2    /* t_vortline end vortex (compute_vel_vortline, mod_vortline.f90:126) */

/* One element list of apply_multipole (elem, wpan, wrin or wvort).
 * Polygons: n_ver(n) in {3,4}, ver(3,4,n) (the 4th vertex of a triangle is
 * unused); lines: ver(3,2,n), n_ver unused.  mag(n): doublet / circulation
 * strength, NULL = no doublet or line term (a line whose strength is not
 * associated induces nothing).  sigma(n), surface panels only: the source
 * strength nor.(ub - wind(cen) - uvort) of compute_vel_surfpan; NULL = none.
 * stretch: 1 if the list takes part in the vs_elems stretching term (elem,
 * wpan, wvort in the reference), 0 if not (wrin). */
typedef struct {
  int kind;
  int n;
  const int *n_ver;
  const double *ver;
  const double *mag;
  const double *sigma;
  int stretch;
  int reserved;
} dust_elem_group;

typedef struct {
  double ff_ratio_doublet; /* sim_param%FarFieldRatioDoublet                        */
  double rankine_rad;      /* sim_param%RankineRad                                  */
  double cutoff_rad;       /* sim_param%CutoffRad                                   */
  double vortex_rad;       /* sim_param%VortexRad, core of gradient_calc_doublet    */
  int precision;           /* DUST_ELEM_FP64 (the drop-in) or DUST_ELEM_TUNED:      */
                           /* far-field element velocities in FP32 (FP64 offsets,  */
                           /* switches and sums; near field and stretching FP64)   */
  int reserved[3];
} dust_elem_params;
#define DUST_ELEM_FP64 0
#define DUST_ELEM_TUNED 1

/* The element loops of apply_multipole (mod_octree.f90:1350-1401) at n
 * points pos(3,n): vel(3,n) += sum over the groups, in order, of
 * compute_vel / (4 pi); if stretch and alpha are given (use_vs and
 * vs_elems), stretch(3,n) += sum over the groups with `stretch` set of
 * transpose(compute_grad) alpha / (4 pi).  Host pointers, synchronous. */
int dust_elem_apply(int ngroups, const dust_elem_group *groups, const dust_elem_params *prm,
                    int n, const double *pos, const double *alpha, double *vel, double *stretch);
/* Same with device pointers (group arrays, pos, alpha, vel, stretch; the
 * group and parameter structs themselves are host memory), ordered on
 * `stream`; n_ver must hold 3 or 4. */
int dust_elem_apply_dev(int ngroups, const dust_elem_group *groups, const dust_elem_params *prm,
                        int n, const double *pos, const double *alpha, double *vel,
                        double *stretch, void *stream);

/* get_vort_vel (mod_surfpan.f90:962-977, mod_vortlatt.f90:676-693) at m
 * centres cen(3,m): uvort(3,m) = sum over the np vortex particles
 * (compute_vel_vortpart, mod_vortpart.f90:139-153: Rosenhead core of the
 * source radius r_vortex) and then the nl end-vortex lines lver(3,2,nl) of
 * strength lmag (NULL: none associated) of compute_vel / (4 pi).
 * Host pointers, synchronous. */
int dust_vort_vel(int m, const double *cen, int np, const double *pcen, const double *pdir,
                  const double *pmag, const double *prv, int nl, const double *lver,
                  const double *lmag, const dust_elem_params *prm, double *uvort);
int dust_vort_vel_dev(int m, const double *cen, int np, const double *pcen, const double *pdir,
                      const double *pmag, const double *prv, int nl, const double *lver,
                      const double *lmag, const dust_elem_params *prm, double *uvort,
                      void *stream);

/* ======================================================================== */
/* Device-resident particle pool: src/aeroel/mod_wake.f90 complete_wake     */
/* (1305-1355, 1888-1967) and prepare_wake (826-835)                         */
/* ======================================================================== */

/* The pool is the reference's wake_parts(nmax) as device SoA: cen(3,nmax),
 * dir(3,nmax), mag(nmax), r_vortex(nmax), vel(3,nmax), stretch(3,nmax),
 * rotu(3,nmax) and free_flag(nmax) (1 = free slot).  n_prt is a device int
 * holding wake%n_prt. */
typedef struct {
  double box_min[3]; /* wake%part_box_min                                       */
  double box_max[3]; /* wake%part_box_max                                       */
  double dt;         /* sim_param%dt * ndt_update_wake                          */
  double filt_eta;   /* sim_param%filt_eta / ndt_update_wake                    */
  int use_vs, use_vd, use_divfilt;
  int reserved[5];
} dust_wake_params;

/* Particle loop of complete_wake (mod_wake.f90:1305-1355) over the non-free
 * slots: move by vel*dt; inside the box, optional divergence filter on
 * stretch and explicit Euler on the direction (mag unchanged); outside, free
 * the slot and decrement *n_prt.  The collision avoidance of use_pa runs
 * before it (dust_wake_avoid_collision_dev).  Device pointers, ordered on
 * `stream`, no host synchronisation. */
int dust_wake_update_dev(int nmax, unsigned char *free_flag, double *cen, double *dir,
                         const double *mag, const double *vel, double *stretch,
                         const double *rotu, const dust_wake_params *prm, int *n_prt,
                         void *stream);

/* The live list of prepare_wake (mod_wake.f90:826-835): live(0..*n_live-1)
 * = the non-free slots in ascending order (0-based); n_live is a device
 * int. */
int dust_wake_live_dev(int nmax, const unsigned char *free_flag, int *live, int *n_live,
                       void *stream);

/* Insertion of k shed particles (mod_wake.f90:1888-1967 and the ring branch):
 * in order, each new particle inside the box takes the lowest free slot:
 * mag = |partvec|, dir = partvec/mag (partvec when mag <= 1e-13), cen,
 * r_vortex, vel as given; *n_prt is incremented.  The geometry that forms
 * partvec (last wake panel row, neighbours, join_te) stays with the caller.
 * Returns DUST_E_CAPACITY, as the reference stops, when the free slots run
 * out.  Device pointers; synchronises `stream` to report that error. */
int dust_wake_shed_dev(int nmax, unsigned char *free_flag, double *cen, double *dir,
                       double *mag, double *r_vortex, double *vel, int k, const double *new_cen,
                       const double *new_partvec, const double *new_rv, const double *new_vel,
                       const dust_wake_params *prm, int *n_prt, void *stream);

/* Collision avoidance of use_pa (avoid_collision, mod_wake.f90:2350-2534,
 * called from complete_wake :1313-1318 before the particle update): for
 * every non-free slot, the velocity is corrected element by element, in list
 * order, against the body's surface panels (DUST_ELEM_SURFPAN lists) and
 * vortex-lattice panels (DUST_ELEM_RING lists; lifting lines take no part,
 * DUST_ELEM_LINE lists are skipped).  ub[k] (device, (3,n) of group k, or
 * NULL) is the panels' velocity; mag and sigma of the groups are unused.
 * The reference's quirks are kept: the vortex-lattice normal distance is
 * |dist*n| elementwise, and the layer thickness (the particle's r_Vortex)
 * doubles each time that branch reaches its test.  Device pointers. */
typedef struct {
  double dt;         /* sim_param%dt                     */
  double dt_wake;    /* sim_param%dt * ndt_update_wake   */
  double u_ref;      /* sim_param%u_ref                  */
  double rad_mult;   /* sim_param%pa_rad_mult            */
  double elrad_mult; /* sim_param%pa_elrad_mult          */
  int reserved[4];
} dust_pa_params;
int dust_wake_avoid_collision_dev(int nmax, const unsigned char *free_flag, const double *cen,
                                  const double *r_vortex, double *vel, int ngroups,
                                  const dust_elem_group *groups, const double *const *ub,
                                  const dust_pa_params *prm, void *stream);

/* The pointer arrays the FMM and the element loops read (the reference
 * passes part_p(:)%p): pcen(3,n), pdir(3,n), pmag(n), prv(n) and optional
 * palpha(3,n) = dir*mag of the live slots live(0..n-1), and the way back of
 * the velocities and stretching (apply_multipole writes them through the
 * same pointers, mod_octree.f90:1377-1397).  Device pointers. */
int dust_wake_gather_dev(int n, const int *live, const double *cen, const double *dir,
                         const double *mag, const double *r_vortex, double *pcen, double *pdir,
                         double *pmag, double *prv, double *palpha, void *stream);
int dust_wake_scatter_dev(int n, const int *live, const double *pvel, const double *pstretch,
                          double *vel, double *stretch, void *stream);

/* ======================================================================== */
/* AIC assembly: src/aeroel/mod_surfpan.f90, mod_vortlatt.f90, mod_doublet   */
/* ======================================================================== */

/* Panels are given by their vertices ver(3,4,n) (triangles leave the 4th
 * vertex unused) and n_ver(n) in {3,4}, as in calc_geo_data_surfpan
 * (mod_surfpan.f90:890) / calc_geo_data_vortlatt (mod_vortlatt.f90:601).   */
typedef struct {
  double ff_ratio_doublet;  /* sim_param%FarFieldRatioDoublet (default 10) */
  double rankine_rad;       /* sim_param%RankineRad                        */
  double cutoff_rad;        /* sim_param%CutoffRad                         */
  int reserved[4];
} dust_aic_params;

/* Static surface-panel system, build_row_static_surfpan with
 * compute_pot_surfpan (mod_surfpan.f90:448-481, 579-601):
 *   A(i,j) = -Omega_j(cen_i), A(i,i) = 2*pi;  B(i,j) = source potential.
 * Column-major A(lda,n), B(ldb,n); B may be NULL.  Host pointers. */
int dust_aic_surfpan(int n, const int *n_ver, const double *ver, const dust_aic_params *prm,
                     double *A, int lda, double *B, int ldb);
int dust_aic_surfpan_dev(int n, const int *n_ver, const double *ver,
                         const dust_aic_params *prm, double *A, int lda, double *B, int ldb,
                         void *stream);

/* Static vortex-lattice system, build_row_static_vortlatt with
 * compute_psi_vortlatt (mod_vortlatt.f90:201-237, 365-391):
 *   A(i,j) = v_ring_j(cen_i).nor_i;  B(i,i) = 4*pi, else 0. */
int dust_aic_vortlatt(int n, const int *n_ver, const double *ver, const dust_aic_params *prm,
                      double *A, int lda, double *B, int ldb);

/* Static part of the right-hand side (mod_surfpan.f90:400-404,
 * mod_vortlatt.f90:158-164): b(i) = sum_j B(i,j) nor_j.(-u_inf - uvort_j)
 * for nrhs free-stream vectors u_inf(3,nrhs); uvort(3,n) may be NULL.
 * b is (ldb_out, nrhs).  Host pointers. */
int dust_aic_rhs(int n, const int *n_ver, const double *ver, const double *B, int ldb,
                 int nrhs, const double *u_inf, const double *uvort, double *b, int ldbo);
/* Same with device pointers (n_ver, ver, B, uvort, b; u_inf(3,nrhs) is host
 * memory), ordered on `stream`.  If sigma (device, (n,nrhs)) is given it
 * receives the coefficients nor_j.(-u_inf - uvort_j), which are also the
 * source strengths of compute_vel_surfpan for a body at rest. */
int dust_aic_rhs_dev(int n, const int *n_ver, const double *ver, const double *B, int ldb,
                     int nrhs, const double *u_inf, const double *uvort, double *b, int ldbo,
                     double *sigma, void *stream);

/* ======================================================================== */
/* LU: the LAPACK call sites of src/linsys/mod_linsys.f90 (236, 435, 473)    */
/* ======================================================================== */

/* dgetrf / dgetrs with LAPACK semantics (column-major, 1-based ipiv, info),
 * plain FP64.  Host pointers, synchronous. */
int dust_dgetrf(int m, int n, double *A, int lda, int *ipiv);
int dust_dgetrs(int n, int nrhs, const double *A, int lda, const int *ipiv, double *B,
                int ldb);
/* Device-pointer forms on `stream` (hipStream_t, NULL = default).  info is
 * returned through the function value after an internal synchronisation. */
int dust_dgetrf_dev(int m, int n, double *dA, int lda, int *dipiv, void *stream);
int dust_dgetrs_dev(int n, int nrhs, const double *dA, int lda, const int *dipiv, double *dB,
                    int ldb, void *stream);
/* dust_dgetrs_dev without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight per-step solve loops, harnesses that forbid hipMalloc).
 * Same kernels and results as dust_dgetrs_dev; its workspace is a static
 * device array of about 12.4 MB that every GPU the library runs on holds
 * until the process ends, shared by all calls on that GPU (calls that
 * overlap in time on one GPU race; each call clears what must start at
 * zero).  n <= 6144, else DUST_E_CAPACITY.  Synchronises `stream`. */
int dust_dgetrs_noalloc_dev(int n, int nrhs, const double *dA, int lda, const int *dipiv,
                            double *dB, int ldb, void *stream);

/* Tuned mixed-precision solve of A X = B (tuned form of the dgetrf 236/435 +
 * dgetrs 473 call sites): the LU runs on the matrix cores in `prec`
 * (DUST_LU_FP32, or DUST_LU_FP16S = row-scaled FP16 operands with FP32
 * accumulation); every 128-wide diagonal block is factored in FP64 with row
 * interchanges among its rows, and the multipliers below it must satisfy
 * |l| <= 10 (threshold pivoting) -- otherwise, or on a zero pivot, or when
 * refinement stalls, the system is solved by the FP64 partial-pivoting path
 * and *iters is set to -1.  FP64 residuals drive iterative refinement until
 * max_k ||b_k - A x_k||_inf / (||A||_inf ||x_k||_inf) <= tol or max_iter
 * sweeps.  DUST_LU_FP64 solves with the FP64 path (*iters = 0).  A is not
 * modified; B is overwritten by X.  Returns 0 or the FP64 path's info > 0 for
 * a singular matrix.  *iters and *resid (may be NULL) report the sweeps and
 * the final backward error.  Process-wide workspace: not re-entrant. */
#define DUST_LU_FP64 0
#define DUST_LU_FP32 1
#define DUST_LU_FP16S 2
int dust_dgesv_mixed(int n, int nrhs, const double *A, int lda, double *B, int ldb, int prec,
                     double tol, int max_iter, int *iters, double *resid);
int dust_dgesv_mixed_dev(int n, int nrhs, const double *dA, int lda, double *dB, int ldb,
                         int prec, double tol, int max_iter, int *iters, double *resid,
                         void *stream);

/* Performance knobs of the LU solvers (no counterpart in the reference):
 * "getrf_reserved_cus" / "mixed_reserved_cus" = compute units kept free of
 * the trailing updates for the diagonal-block chain (CU-masked lookahead
 * stream; 0 = no mask; default 32, at most half the CUs).  Results do not
 * depend on them.  Setting one waits for the device and rebuilds that
 * solver's streams (and, for the mixed solve, its workspaces). */
int dust_lu_set_tuning(const char *name, int value);
int dust_lu_get_tuning(const char *name, int *value);

/* Batch of independent tuned solves (device pointers, one system per entry,
 * sizes may differ), advanced in lockstep: the factorisations of all systems
 * are launched (each system on its own stream and workspace), then each
 * refinement sweep of all unfinished systems, with one synchronisation of all
 * streams per step.  iters[i] / resid[i] as for dust_dgesv_mixed_dev (may be
 * NULL).  Returns 0 or the first failing system's status.  Process-wide
 * workspaces: not re-entrant.  No counterpart in the reference; for many-case
 * runs. */
int dust_dgesv_mixed_batch_dev(int nb, const int *n, const double *const *dA, const int *lda,
                               double *const *dB, const int *ldb, int nrhs, int prec, double tol,
                               int max_iter, int *iters, double *resid);

/* solve_linsys block path (mod_linsys.f90:352-483): the leading nstatic
 * block of A(lda,rank) is already factored (ipiv(1:nstatic) from dgetrf);
 * the moving rows/columns are eliminated with laswp/trsm/gemm, the moving
 * block is factored, the pivots are made global, and res = A^-1 b is
 * solved with one right-hand side.  Host pointers; A and ipiv are updated
 * in place exactly as the reference leaves them. */
int dust_solve_linsys(int rank, int nstatic, double *A, int lda, int *ipiv, const double *b,
                      double *res);

/* ======================================================================== */
/* Assembly of the system with its wake: src/linsys/mod_linsys.f90          */
/* initialize_linsys (109-272) and assemble_linsys (279-344)                 */
/* ======================================================================== */

#define DUST_ROW_SURFPAN 0  /* t_surfpan: compute_pot rows (mod_surfpan.f90:377-481) */
#define DUST_ROW_VORTLATT 1 /* t_vortlatt: compute_psi rows (mod_vortlatt.f90:143-237) */

/* The implicit elements (linsys%rank of them, the nstatic static ones
 * first), the wake panels of wake%pan_p (the n_impl implicit first-row
 * panels first, with pan_gen_elems_id), the explicit elements (lifting-line
 * panels, the nstatic_expl static ones first) and the surface-panel index
 * list idSurfPan of the pressure system.  Vertex lists are (3,4,n) with
 * triangles padded.  For the _dev entry points every pointer is device
 * memory (the struct itself is host memory); for the host entry points,
 * host memory. */
typedef struct dust_linsys_desc {
  int rank, nstatic;
  const int *kind;         /* (rank) DUST_ROW_*                                    */
  const int *n_ver;        /* (rank) 3 or 4                                         */
  const double *ver;       /* (3,4,rank)                                            */
  const double *ub;        /* (3,rank) body velocity at the centres, or NULL (0);
                              read for the moving elements only (zero for static) */
  const double *uvort;     /* (3,rank) get_vort_vel result, or NULL (0)             */
  const double *bern;      /* (rank) bernoulli_source of the surface panels, or NULL */
  double u_inf[3];         /* uniform wind (gusts: fold dust_variable_wind into uvort) */
  int n_sp, nstatic_sp;    /* surface panels, the static ones first                */
  const int *id_sp;        /* (n_sp) 0-based element index of each (idSurfPan - 1)  */
  int n_wake, n_impl;      /* wake panels, the implicit first row first             */
  const int *wake_nver;    /* (n_wake)                                              */
  const double *wake_ver;  /* (3,4,n_wake)                                          */
  const double *wake_mag;  /* (n_wake) doublet intensity (read for the explicit ones) */
  const int *impl_ind;     /* (2,n_impl) pan_gen_elems_id: 1-based, 0 = no lower   */
  int n_expl, nstatic_expl;
  const int *expl_nver;    /* (n_expl)                                              */
  const double *expl_ver;  /* (3,4,n_expl)                                          */
  const double *expl_mag;  /* (n_expl)                                              */
  double ff_ratio_doublet, rankine_rad, cutoff_rad;
} dust_linsys_desc;

/* initialize_linsys (mod_linsys.f90:109-272): the static rows
 * (build_row_static, 174-179) into A(1:nst,1:nst) and the static source
 * coefficients b_static into B(1:nst,1:nst); L_static(nst, nstatic_expl);
 * the pressure matrix A_pres(n_sp,n_sp) copied by idSurfPan before the wake
 * correction (its static block; the moving blocks are written by every
 * assembly); the implicit wake panels on static columns added into A
 * (191-198); then dgetrf of A(1:nst,1:nst) into ipiv(nst) and of
 * A_pres(1:nstatic_sp,1:nstatic_sp) into ipiv_pres (236, 255).  Returns the
 * first nonzero dgetrf info, as the reference stops on it.  B(ldb, rank) is
 * the whole source-coefficient matrix of this library: its static block is
 * b_static and dust_assemble_linsys_dev writes the moving blocks (the
 * reference forms those products inside its row loops). */
int dust_initialize_linsys(const dust_linsys_desc *desc, double *A, int lda, double *B, int ldb,
                           double *L_static, int ldl, double *A_pres, int ldp, int *ipiv,
                           int *ipiv_pres);
int dust_initialize_linsys_dev(const dust_linsys_desc *desc, double *A, int lda, double *B, int ldb,
                               double *L_static, int ldl, double *A_pres, int ldp, int *ipiv,
                               int *ipiv_pres, void *stream);

/* assemble_linsys (mod_linsys.f90:279-344) after get_vort_vel (desc->uvort):
 *   (2) build_row of the static rows over the moving columns and of the
 *       moving rows over all columns: A entries and their B coefficients;
 *       b(ie) = sum_j B(ie,j) nor_j.(ub_j - u_inf - uvort_j), ub taken for
 *       the moving columns only; b_pres(n_sp) (optional, needs desc->bern) =
 *       the Bernoulli right-hand side of build_row_surfpan (405-437);
 *   (3) A_pres blocks of the moving surface panels copied from A (318-324);
 *   (4) add_wake and add_expl: implicit wake panels whose generating
 *       elements are columns of the row's range add +-a into A(ie,ind1/ind2);
 *       explicit wake panels and explicit elements subtract a*mag from b(ie)
 *       (static explicit elements of static rows through L_static).
 * The static block of A keeps its LU factors.  A_pres may be NULL. */
int dust_assemble_linsys(const dust_linsys_desc *desc, double *A, int lda, double *B, int ldb,
                         const double *L_static, int ldl, double *b, double *A_pres, int ldp,
                         double *b_pres);
int dust_assemble_linsys_dev(const dust_linsys_desc *desc, double *A, int lda, double *B, int ldb,
                             const double *L_static, int ldl, double *b, double *A_pres, int ldp,
                             double *b_pres, void *stream);

/* ======================================================================== */
/* Pressure (Bernoulli) integral equation: src/linsys/mod_pressure_equation */
/* ======================================================================== */

/* solve_pressure_sys (mod_pressure_equation.f90:251-390): the same block
 * elimination as dust_solve_linsys on the surface-panel system A_pres
 * (n_sp x n_sp, static block of nstatic_sp rows factored by dgetrf as
 * initialize_linsys does, mod_linsys.f90:247-262), pivots P_pres; A_pres and
 * P_pres are left as the reference leaves them, res_pres = A_pres^-1 b_pres.
 * Host pointers. */
int dust_solve_pressure_sys(int n_sp, int nstatic_sp, double *A_pres, int lda, int *P_pres,
                            const double *b_pres, double *res_pres);

/* Source part (b) of the pressure right-hand side, formed row by row in
 * build_row_surfpan (mod_surfpan.f90:405-437): b_pres = B_sp src, with B_sp
 * (ldb, n_sp) the surface panels' source matrix (b_static_pres for the static
 * columns, the b1 of compute_pot for the moving ones) and src(n_sp) their
 * bernoulli_source (press_normvel_der).  b_pres is overwritten.  Host
 * pointers / device pointers ordered on `stream`. */
int dust_pres_rhs_src(int n_sp, const double *B, int ldb, const double *src, double *b_pres);
int dust_pres_rhs_src_dev(int n_sp, const double *B, int ldb, const double *src, double *b_pres,
                          void *stream);

/* press_normvel_der (mod_pressure_equation.f90:396-462): the source
 * strength bernoulli_source(n_sp) of every surface panel,
 *   n.(ub - surf_vel_old)/dt - GradS_Un . ub + DivS_U (ub . n),
 * GradS_Un = sum_e R (stencil_e n.(U_e - U)), DivS_U = sum_e (R stencil_e).(U_e - U)
 * over the panel's n_ver edges, U = surf_vel, U_e = the neighbour's surf_vel;
 * an edge without neighbour (neigh = -1) takes U_e - U = -2 U, an edge whose
 * neighbour is not a surface panel (neigh = -2) adds nothing.  Arrays over
 * the surface panels in idSurfPan order: nor, ub, surf_vel, surf_vel_old
 * (3,n_sp); stencil (3,4,n_sp) = pot_vel_stencil; n_ver (n_sp); neigh
 * (4,n_sp) 0-based surface-panel index of the neighbour across each edge
 * (or -1 / -2); R (3,3,n_sp) column-major, the R_g the reference applies to
 * panel i (that of the component of elems(i), as the reference indexes it).
 * bernoulli_source is overwritten.  Host pointers (n_ver and neigh checked:
 * DUST_E_ARG) / device pointers ordered on `stream` (not checked on the host:
 * the kernel reads at most 4 edges and treats a neighbour index outside
 * -1..n_sp-1 as -2). */
int dust_press_normvel_der(int n_sp, const double *nor, const double *ub,
                           const double *surf_vel, const double *surf_vel_old,
                           const double *stencil, const int *n_ver, const int *neigh,
                           const double *R, double dt, double *bernoulli_source);
int dust_press_normvel_der_dev(int n_sp, const double *nor, const double *ub,
                               const double *surf_vel, const double *surf_vel_old,
                               const double *stencil, const int *n_ver, const int *neigh,
                               const double *R, double dt, double *bernoulli_source,
                               void *stream);

/* assemble_pressure_sys (mod_pressure_equation.f90:139-236), rotational part
 * (c) of the right-hand side at the m surface-panel centres cen(3,m),
 * accumulated into b_pres(m) in the reference's order (particles, then
 * lines, then wake panels):
 *   + the np particles (live slots only; 171-187): pcen, pdir, pmag,
 *     prv (r_Vortex) and pvel (3,np), d.(dir x vel) mag / (|d|^2+r^2)^1.5,
 *     d = pcen - cen;
 *   - the nl end-vortex lines lver(3,2,nl) of strength lmag (NULL: none
 *     associated) with vertex velocities lvel(3,2,nl) (ver_vel; 191-203);
 *   - the nw wake panels wver(3,4,nw) (wnver vertices each, triangles
 *     padded) of strength wmag with vertex velocities wvel(3,4,nw) in the
 *     vertex order (pan_w_vel at the panel's points; 204-225), edge by edge.
 *     Every wnver entry must be 3 or 4: the host entry point returns
 *     DUST_E_ARG otherwise; the device entry point does not read the list on
 *     the host and treats any value other than 3 as 4.
 * Host pointers, synchronous. */
int dust_pres_rhs_rot(int m, const double *cen, int np, const double *pcen, const double *pdir,
                      const double *pmag, const double *prv, const double *pvel, int nl,
                      const double *lver, const double *lmag, const double *lvel, int nw,
                      const int *wnver, const double *wver, const double *wmag,
                      const double *wvel, double *b_pres);
/* Same with device pointers, ordered on `stream`. */
int dust_pres_rhs_rot_dev(int m, const double *cen, int np, const double *pcen,
                          const double *pdir, const double *pmag, const double *prv,
                          const double *pvel, int nl, const double *lver, const double *lmag,
                          const double *lvel, int nw, const int *wnver, const double *wver,
                          const double *wmag, const double *wvel, double *b_pres,
                          void *stream);

/* ======================================================================== */
/* Panel wake: src/aeroel/mod_wake.f90 initialize_wake (331-579) and the    */
/* panel parts of update_wake (860-1214) and complete_wake (1222-2105)       */
/* ======================================================================== */

/* Static description of the panel wake (host memory; copied at creation). */
typedef struct {
  int n_stripes;            /* wake%n_pan_stripes (trailing-edge segments)        */
  int n_points;             /* wake%n_pan_points (trailing-edge points)           */
  int nmax_pan;             /* sim_param%n_wake_panels (rows of panels)           */
  const int *start_points;  /* (2,n_stripes) i_start_points, 1-based              */
  const int *neigh;         /* (2,n_stripes) pan_neigh, 1-based stripe, 0 = none  */
  const int *neigh_o;       /* (2,n_stripes) pan_neigh_o, +1 / -1                 */
  const int *gen_elems;     /* (2,n_stripes) pan_gen_elems_id: implicit elements, 1-based, 0 = none */
  double dt;                /* sim_param%dt                                        */
  int ndt_update_wake;      /* sim_param%ndt_update_wake                           */
  double min_vel_at_te;     /* sim_param%min_vel_at_te                             */
  int join_te;              /* sim_param%join_te                                   */
  double join_te_factor;    /* sim_param%join_te_factor                            */
  int refine_wake, k_refine;/* sim_param%refine_wake, k_refine                    */
  double k_vortex_rad;      /* sim_param%KVortexRad (< 1e-10: use vortex_rad)      */
  double vortex_rad;        /* sim_param%VortexRad                                 */
  double box_min[3], box_max[3]; /* particles_box_min / max                      */
  int reserved[4];
} dust_panwake_params;

typedef struct dust_panwake_s *dust_panwake_t;

/* Device state of one panel wake: points pan_w_points(3,n_points,nmax_pan+1)
 * and velocities pan_w_vel, intensities, end and last intensities, the
 * junction history of join_te, the panels' vertex lists; pan_wake_len = 1. */
int dust_panwake_create(const dust_panwake_params *prm, dust_panwake_t *out);
int dust_panwake_destroy(dust_panwake_t h);

/* Rows 1 and 2 of points from the trailing edge: first = 1 as in
 * initialize_wake (512-546), first = 0 as in complete_wake (1250-1280, then
 * the junction history moves back one row); join_first_panels (2199-2313)
 * when join_te; then the vertex lists of the pan_wake_len rows of panels.
 * start_points(3,n_points) = w_start_points (mean of the two generating mesh
 * points), gen_dir(3,n_points) = R_g t_hinged, scaling(n_points) =
 * pan_gen_scaling, vel_te(3,n_points) = calc_node_vel at the points, u_inf
 * the wind (host, 3 doubles).  Device pointers, ordered on `stream`. */
int dust_panwake_start_dev(dust_panwake_t h, int first, const double *start_points,
                           const double *gen_dir, const double *scaling, const double *vel_te,
                           const double u_inf[3], void *stream);

/* The wake panels of the assembly (wake%pan_p): n_wake = n_stripes *
 * pan_wake_len panels in (row, stripe) order, the first n_impl = n_stripes
 * implicit, with device pointers to their vertex counts, vertex lists
 * (3,4,n_wake), intensities and pan_gen_elems_id (2,n_impl) -- the fields of
 * dust_linsys_desc. */
int dust_panwake_panels(dust_panwake_t h, int *n_wake, int *n_impl, const int **n_ver,
                        const double **ver, const double **mag, const int **impl_ind);

/* The points at which update_wake evaluates velocities (rows 2 ..
 * pan_wake_len+1 of pan_w_points, (3,n) device memory, n = pan_wake_len *
 * n_points). */
int dust_panwake_points(dust_panwake_t h, int *n, const double **pts);

/* The first-row intensities alone (update_wake 884-894): mu(ind1) -
 * mu(ind2); update_wake evaluates the velocities of the step with them,
 * before dust_panwake_update_dev moves the panels. */
int dust_panwake_first_row_dev(dust_panwake_t h, const double *mu, void *stream);

/* update_wake's panel part: first-row intensities from the solution mu
 * (implicit elements, device) and pan_gen_elems_id (884-894); the points of
 * rows 3.. moved with vel(3,n) = the velocity (compute_vel_from_all + wind)
 * at dust_panwake_points, the end points when the wake is full (910-955);
 * pan_wake_len grown; intensities shifted back one row (1177-1187). */
int dust_panwake_update_dev(dust_panwake_t h, const double *mu, const double *vel, void *stream);

/* complete_wake's shedding when the panel wake is full (classical, or
 * refined with k_refine; 1798-1969): per stripe the vortex vector of the
 * last panel row from its left, right and end sides with the neighbours'
 * end intensities (or the joined intensity of join_te), inserted into the
 * particle pool as dust_wake_shed_dev does; then the end vortices (2089-2103).
 * *status (device int) is or-ed with 1 when the pool's free slots run out.
 * Device pointers, no host synchronisation. */
int dust_panwake_shed_dev(dust_panwake_t h, int nmax, unsigned char *free_flag, double *cen,
                          double *dir, double *mag, double *r_vortex, double *vel,
                          const dust_wake_params *prm, int *n_prt, int *status, void *stream);

/* The end vortices (n = n_stripes once the wake is full, else 0): device
 * vertex lists (3,2,n), intensities (last_pan_idou) and vertex velocities. */
int dust_panwake_end_vortices(dust_panwake_t h, int *n, const double **ver, const double **mag,
                              const double **vel);

/* pan_wake_len, full_panels and the most particles one shedding creates. */
int dust_panwake_state(dust_panwake_t h, int *len, int *full, int *nshed_max);

/* Host copies of the state (each pointer may be NULL): points and
 * velocities (3,n_points,nmax_pan+1), intensities (n_stripes,nmax_pan)
 * stored row by row, end and last intensities (n_stripes), end points
 * (3,n_points).  Synchronises the device. */
int dust_panwake_get(dust_panwake_t h, double *points, double *pvel, double *idou,
                     double *end_idou, double *last_idou, double *points_end);

/* Checkpoint / restart of the panel wake's device state (points, velocities,
 * intensities, junction history, panels, end vortices, row counters), for a
 * restart of the time loop that continues bit for bit.  checkpoint: with
 * buf == NULL writes the size needed to *size; else fills buf (host memory,
 * *size bytes available) and synchronises the device.  restore: the buffer
 * of a checkpoint of a wake with the same tables and dimensions.  No
 * counterpart in the reference's API (its restart reads the HDF5 result
 * files, mod_wake.f90:597-778). */
int dust_panwake_checkpoint(dust_panwake_t h, void *buf, size_t *size);
int dust_panwake_restore(dust_panwake_t h, const void *buf, size_t size);

/* The panel part of load_wake (src/aeroel/mod_wake.f90:611-655, 766-775):
 * the wake of a result file's PanelWake group, host arrays as the file holds
 * them -- points and velocities [len+1][n_points][3] (WakePoints, WakeVels),
 * intensities [len][n_stripes] (WakeVort) -- and ParticleWake/LastPanIdou
 * [n_stripes].  pan_wake_len becomes min(nmax_pan, len) (the leading rows
 * are kept), the panels are rebuilt from the points, and a wake of
 * nmax_pan rows is full with its end vortices on the last row of points.
 * The trailing-edge rows of a start(first = 1) call and the junction tables
 * are kept, as the reference keeps them.  The caller checks StartPoints
 * against its own table (the reference's connectivity check, 627-630).
 * Synchronises `stream`. */
int dust_panwake_load(dust_panwake_t h, int len, const double *points, const double *pvel,
                      const double *idou, const double *last_idou, void *stream);

/* Release every library-global workspace (device buffers, page-locked
 * records, streams and events of the solvers and element kernels) after
 * synchronising the device; they are re-created on the next call that needs
 * them.  FMM handles are not touched (dust_fmm_destroy releases each).  For
 * callers that tear the GPU runtime down before process exit.  No
 * counterpart in the reference (its LAPACK calls hold no state). */
int dust_shutdown(void);

/* Library build/version information. */
const char *dust_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DUST_GPU_H */
