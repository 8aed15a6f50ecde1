/*
 * infraga_gpu.h -- C ABI of the MI355X ray-tracing engine (libinfraga_gpu.so).
 *
 * Plain C types only: no HIP or torch types appear in any signature.  All
 * entry points return GA_OK (0) or a negative GA_E_* code and never print
 * or prompt; ga_last_error() gives the message of the last failure on a
 * context.  One context may be used by one host thread at a time.
 *
 * Units at this boundary follow the reference CLI: km, km/s, g/cm^3, and
 * angles in degrees.  Atmosphere arrays are the values the reference loader
 * holds after conversion (c, u, v already divided by 1000).
 *
 * Reference interfaces replaced (paths relative to the infraGA source tree):
 *   ga_atmo_load_strat  <- set_region(char*, char*, bool)
 *                          src/atmo/atmo_io.sph.strat.cpp:57-144,
 *                          src/atmo/atmo_io.3d.strat.cpp:57-140
 *   ga_fan_run          <- the -prop azimuth x inclination x bounce loop
 *                          src/main.infraga.sph.cpp:418-558 and
 *                          src/main.infraga.3d.cpp:407-548, which drive
 *                          geoac::set_initial / prop_rk4 / travel_time /
 *                          atten / amp / set_refl
 *                          (src/geoac/geoac.eqset.h:5-38, src/util/rk4solver.h:6)
 *   ga_eig_search       <- run_eig_search's per-bounce loop over
 *                          geoac::est_eigenray + geoac::find_eigenray
 *                          src/main.infraga.3d.cpp:944-951,
 *                          src/geoac/geoac.eigenray.h:17-20
 *   ga_fan_run_paths    <- the write_rays / write_caustics branch of the
 *                          -prop loop (<id>.raypaths.dat, <id>.caustics-N.dat)
 *                          src/main.infraga.sph.cpp:388-413, 456-496, 549;
 *                          src/main.infraga.3d.cpp:378-400, 442-487
 *   ga_fan_run(_paths) with GA_GEOM_2D
 *                       <- infraga-2d run_prop (src/main.infraga.2d.cpp:127-400)
 *                          over geoac.eqset.2d.cpp
 *   ga_eig_direct       <- run_eig_direct: geoac::find_eigenray from a
 *                          launch estimate, src/main.infraga.3d.cpp:955-1086,
 *                          src/main.infraga.sph.cpp:998-1139
 *   ga_eig_search_paths <- the same search, plus find_eigenray's
 *                          <id>.eigenray-N.dat rows
 *                          src/geoac/geoac.eigenray.3d.cpp:223-276,
 *                          src/geoac/geoac.eigenray.sph.cpp:228-271
 *   ga_wnl_wvfrm        <- run_wnl_wvfrm (src/main.infraga.sph.cpp:1142-1479)
 *                          and geoac::wnl_wvfrm
 *                          (src/geoac/geoac.eqset.sph.strat.cpp:853-1148,
 *                          src/geoac/geoac.eqset.3d.strat.cpp), one
 *                          waveform per launch
 */
#ifndef INFRAGA_GPU_H_
#define INFRAGA_GPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GA_OK          0
#define GA_E_ARG      -1   /* invalid argument */
#define GA_E_HIP      -2   /* HIP runtime failure */
#define GA_E_NOMEM    -3   /* device or host allocation failed */
#define GA_E_CAP      -4   /* output buffer too small; *n_out holds the size needed */
#define GA_E_NODEV    -5   /* no usable gfx950 device */
#define GA_E_ATMO     -6   /* unknown or released atmosphere id */
#define GA_E_IO       -7   /* file could not be read or parsed */

#define GA_GEOM_3D     0   /* Cartesian x, y, z (infraga-3d) */
#define GA_GEOM_SPH    1   /* spherical r, lat, lon (infraga-sph) */
#define GA_GEOM_SPH_RNGDEP 2   /* spherical, range-dependent atmosphere and
                                  terrain (infraga-sph-rngdep); atmo_id is a
                                  ga_rngdep_load id */
#define GA_GEOM_2D     3   /* 2D effective sound speed in the vertical plane
                              of each launch azimuth (infraga-2d); arrivals:
                              pos0 = range r [km], pos1 = out-of-plane drift
                              of the last leg [km], back_az unused (0) */
#define GA_GEOM_3D_RNGDEP 4  /* 3D Cartesian range-dependent (infraga-3d-rngdep):
                                atmo_id a ga_rngdep_load_xy id; pos0, pos1 =
                                x, y [km]; src[0..1] x, y [km] (NaN: the
                                region centre); bounds x_min, x_max, y_min,
                                y_max [km] (NaN: the grid's) */

typedef struct ga_ctx ga_ctx;

/* Propagation parameters: the -prop / -eig_search key=value set of the
 * reference (src/main.infraga.sph.cpp:208-309).  Fill with
 * ga_prop_params_default() and override. */
typedef struct {
    int    geometry;       /* GA_GEOM_3D or GA_GEOM_SPH */
    int    calc_amp;       /* 1: 18 equations incl. auxiliary (TL), 0: 6 */
    int    bounces;        /* ground reflections to follow */
    double freq;           /* Hz, Sutherland-Bass absorption frequency */
    double abs_coeff;      /* absorption scale (tweak_abs), reference driver default 1.0 */
    double turn_ht_min;    /* km, shadow-zone test height above ground */
    double src[3];         /* sph: lat [deg], lon [deg], alt [km]; 3D: x, y, z [km] */
    double z_grnd;         /* km; NaN = bottom of the profile */
    double s_max, ds_min, ds_max;     /* km */
    double alt_max;        /* km; NaN = top of the profile */
    double rng_max;        /* km */
    double time_max;       /* hours (sph only, as in the reference) */
    double bounds[4];      /* sph: lat_min, lat_max, lon_min, lon_max [deg];
                              3D: x_min, x_max, y_min, y_max [km];
                              range-dependent: NaN = the grid's region
                              (set_limits, atmo_io.sph.rngdep.cpp:27-43) */
    double r_earth;        /* km; the reference hard-codes 6370 */
    int32_t topo_use_blw;  /* range-dependent: boundary-layer vertical wind
                              over terrain (topo_use_BLw, default 0) */
    double refl_alt;       /* 2D: altitude of the elevated (partial) reflection
                              [km]; NaN = none (profile top + 1 km) */
} ga_prop_params;

/* One row of <id>.arrivals.dat (src/main.infraga.sph.cpp:531-544). */
typedef struct {
    int32_t launch_idx;    /* index into the launch list */
    int32_t n_bounce;
    double  incl, az;      /* launch angles [deg] as given */
    double  pos0, pos1;    /* sph: lat, lon [deg]; 3D: x, y [km] */
    double  time;          /* s */
    double  cel;           /* km/s */
    double  turn_ht;       /* km */
    double  incl_arr;      /* deg */
    double  back_az;       /* deg */
    double  tl;            /* dB (0 when calc_amp = 0) */
    double  absorption;    /* dB */
} ga_arrival;

/* One row of <id>.raypaths.dat / <id>.eigenray-N.dat (kind GA_PT_RAY:
 * rows m == 1 or m % 15 == 0 of every leg, main.infraga.sph.cpp:471-482) or
 * of <id>.caustics-<leg>.dat (kind GA_PT_CAUSTIC: the Jacobian determinant
 * changed sign between rows m-1 and m, main.infraga.sph.cpp:485-490). */
typedef struct {
    int32_t launch_idx;    /* ray (index into the launch list) or eigenray (index into the output) */
    int32_t leg;           /* reflections before the row */
    int32_t step;          /* row index m within the leg */
    int32_t kind;          /* GA_PT_RAY or GA_PT_CAUSTIC */
    double  pos0, pos1;    /* sph: lat, lon [deg]; 3D: x, y [km] */
    double  z;             /* sph: altitude r - r0; 3D: max(z, ground) [km] */
    double  tl;            /* 20 log10(amplitude) [dB]; 0 without amplitudes and for caustics */
    double  absorption;    /* -2 x the absorption integral so far [dB] */
    double  time;          /* travel time so far [s] */
} ga_path_pt;

#define GA_PATH_RAYS      1   /* write_rays */
#define GA_PATH_CAUSTICS  2   /* write_caustics (needs calc_amp) */
#define GA_PT_RAY         0
#define GA_PT_CAUSTIC     1

typedef struct {
    int64_t ray_steps;     /* RK4 iterations summed over rays and legs */
    double  kernel_ms;     /* device time of the trace kernels (HIP events) */
    double  wall_ms;       /* host wall time of the call */
    int32_t n_rays;
    int32_t n_launches;    /* kernel launches issued */
} ga_run_stats;

/* Eigenray search controls (src/main.infraga.3d.cpp:773-800 defaults). */
typedef struct {
    double incl_min, incl_max;     /* deg */
    int    bnc_min, bnc_max;
    int    iterations;             /* LM iteration limit (25) */
    double damping;                /* 1e-3 */
    double tolerance;              /* km (0.1) */
    double az_dev_lim;             /* deg (2.0) */
    double incl_step_min, incl_step_max;   /* deg (0.001, 0.1) */
} ga_eig_params;

/* One eigenray: the arrival row written to <id>.arrivals.dat by
 * find_eigenray (src/geoac/geoac.eigenray.3d.cpp:302-315); launch_idx is the
 * receiver index. */
typedef ga_arrival ga_eigenray;

void        ga_prop_params_default(ga_prop_params *p, int geometry);
void        ga_eig_params_default(ga_eig_params *e);

int         ga_ctx_create(ga_ctx **out, int hip_device);
void        ga_ctx_destroy(ga_ctx *ctx);
const char *ga_last_error(const ga_ctx *ctx);

/* Upload a stratified profile (n >= 3 strictly increasing z).  The natural
 * cubic spline slopes are solved on the host exactly as interp::set does
 * (src/util/interpolation.cpp:103-136).  The first run at a frequency and
 * geometry also builds the atmosphere's absorption table (piecewise
 * polynomials of the smooth Sutherland-Bass terms, 176 B per 0.05 km of
 * profile), kept until ga_atmo_release; GA_ABS_TABLE=0 in the environment
 * evaluates absorption directly instead. */
int ga_atmo_load_strat(ga_ctx *ctx, const double *z_km, const double *c_kms,
                       const double *u_kms, const double *v_kms, const double *rho_gcc,
                       int n, int *atmo_id);

/* Parse a .met file (formats zTuvdp, zuvwTdp, zcuvd) with the reference
 * loader's conversions and upload it. */
int ga_atmo_load_file(ga_ctx *ctx, const char *path, const char *format, int reverse_winds,
                      int *atmo_id);

int ga_atmo_release(ga_ctx *ctx, int atmo_id);

/* A stratified profile over terrain, as infraga-sph runs it with topo_file=
 * (set_region(atmo_file, topo_file, format, invert_winds),
 * src/atmo/atmo_io.sph.strat.cpp:147-200, and the stratified medium with
 * terrain of src/atmo/atmo_state.sph.strat.cpp:37-330).  The profile is
 * held as 4 x 4 identical columns spanning the terrain grid, over which the
 * range-dependent splines reduce to the stratified ones exactly, with the
 * stratified sources' boundary-layer top (terrain maximum + 0.01 km, then
 * + 2 km) and the stratified driver's arrival inclination sign; the
 * returned id is a range-dependent id: run it with p->geometry =
 * GA_GEOM_SPH_RNGDEP (region, source and ground from the terrain as the
 * stratified driver sets them).  Terrain heights [nlat][nlon] in km,
 * nodes in degrees.  ga_atmo_load_file_topo reads the reference's files: a
 * .met profile (formats as ga_atmo_load_file) and "lat lon z" terrain rows,
 * longitude fastest.  Release with ga_rngdep_release. */
int ga_atmo_load_strat_topo(ga_ctx *ctx, const double *z_km, const double *c_kms, const double *u_kms,
                            const double *v_kms, const double *rho_gcc, int n, const double *lat_deg, int nlat,
                            const double *lon_deg, int nlon, const double *zg_km, int *atmo_id);
int ga_atmo_load_file_topo(ga_ctx *ctx, const char *path, const char *topo_path, const char *format,
                           int reverse_winds, int *atmo_id);
/* The 3D forms (infraga-3d with topo_file; set_region, src/atmo/
 * atmo_io.3d.strat.cpp:146-257, with atmo_state.3d.strat.cpp): terrain node
 * coordinates x, y in km ("x y z" rows, y fastest, in the file form); the
 * id runs with GA_GEOM_3D_RNGDEP under the stratified 3D semantics (boundary
 * layer top at the terrain maximum + 0.1 + 2 km, the upper-triangle
 * second-derivative contraction of geoac.eqset.3d.strat.cpp:305-310, turning
 * height from the profile bottom, caustic rows floored at the terrain). */
int ga_atmo_load_strat_topo_xy(ga_ctx *ctx, const double *z_km, const double *c_kms, const double *u_kms,
                               const double *v_kms, const double *rho_gcc, int n, const double *x_km, int nx,
                               const double *y_km, int ny, const double *zg_km, int *atmo_id);
int ga_atmo_load_file_topo_xy(ga_ctx *ctx, const char *path, const char *topo_path, const char *format,
                              int reverse_winds, int *atmo_id);

/* Trace a fan.  launches holds n_launch (incl_deg, az_deg) pairs.  Arrivals
 * are returned sorted by (launch_idx, n_bounce).  When out_cap is too small
 * the call returns GA_E_CAP with *n_out set to the count needed. */
int ga_fan_run(ga_ctx *ctx, int atmo_id, const ga_prop_params *p,
               const double *launches, int n_launch,
               ga_arrival *out, int out_cap, int *n_out, ga_run_stats *stats);

/* Device-resident form for callers that keep buffers in HBM: d_launches
 * (2*n doubles), d_out (n * (bounces + 1) slots, slot r*(bounces+1)+b holds
 * bounce b of ray r), d_counts (n ints, arrivals per ray) and d_steps (n
 * ints, RK4 iterations per ray) are device pointers; stream is a hipStream_t
 * or NULL for the context stream.  Plain stratified fans run on the
 * rotating slice scheduler: the call enqueues batches of slices (a captured
 * graph) and waits on the stream after each batch until every ray is
 * finished, then enqueues the arrival kernel and returns without waiting
 * for it; the other cases only enqueue. */
int ga_fan_run_device(ga_ctx *ctx, int atmo_id, const ga_prop_params *p,
                      const double *d_launches, int n_launch,
                      ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, void *stream);

/* Scheduler settings of the plain (write_rays=false) fan: RK4 steps a ray
 * advances per slice of the rotating scheduler, and slices captured per
 * graph launch. */
typedef struct {
    int32_t slice_steps;
    int32_t slices_per_graph;
    double  pass_ms;       /* device time of the fastest candidate pass */
} ga_fan_tuning;

/* Autotuner of the fan scheduler.  No counterpart in the reference, whose
 * MPI fan has no such knobs (rank count aside): it traces the given
 * HBM-resident launch list (d_launches, 2*n doubles) once per candidate
 * (steps per slice 32..512 x slices per graph 8, 32), keeps the fastest in
 * the context for this geometry and equation count (later ga_fan_run /
 * ga_fan_run_device calls use it) and returns it in *best.  Synchronises
 * stream (NULL: the context stream).  Stratified geometries only. */
int ga_fan_autotune(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *d_launches, int n_launch,
                    void *stream, ga_fan_tuning *best);

/* ga_fan_run_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The work queue counter, the per-lane row buffer and the raw
 * arrival records live in a fixed 48 MiB device array that the library
 * holds on each GPU it runs on until the process ends; the call makes no
 * hipMalloc / hipFree and launches the same kernels as ga_fan_run_device.
 * Stratified geometries only (GA_GEOM_SPH, GA_GEOM_3D); n_launch x
 * (bounces + 1) may be at most 157,285 arrival slots (GA_E_ARG beyond, with
 * nothing launched).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other contexts or threads) race, so
 * serialise them.  The caller's buffers are as for ga_fan_run_device.
 * Absorption uses the atmosphere's absorption table for the run's
 * frequency and geometry when an earlier allocating call (ga_fan_run*,
 * ga_eig_search*) has built it, and evaluates Sutherland-Bass directly
 * otherwise; the two agree within 1e-12 relative per step, so results
 * equal ga_fan_run_device's bit for bit once such a call has run. */
int ga_fan_run_device_noalloc(ga_ctx *ctx, int atmo_id, const ga_prop_params *p,
                              const double *d_launches, int n_launch,
                              ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, void *stream);

/* ga_fan_run with the reference's write_rays / write_caustics bookkeeping
 * (path_flags = GA_PATH_RAYS | GA_PATH_CAUSTICS): travel time and
 * absorption are summed segment by segment over rows 1..k-1 of each leg, so
 * the arrivals are those the reference writes with write_rays=true (they
 * differ from ga_fan_run's write_rays=false sums by the segment into the
 * sub-ground row of each leg).  Path rows come back in pts ordered by ray,
 * then leg and step.  pts_cap bounds both the records returned and the
 * device space the call takes (64 B per record plus 2 KB per ray); when the
 * rays produce more records the call returns GA_E_CAP with *n_pts set to the
 * count needed (arrivals are still returned when out_cap allows).  RNGDEP
 * atmospheres are accepted. */
int ga_fan_run_paths(ga_ctx *ctx, int atmo_id, const ga_prop_params *p,
                     const double *launches, int n_launch, int path_flags,
                     ga_arrival *out, int out_cap, int *n_out,
                     ga_path_pt *pts, int64_t pts_cap, int64_t *n_pts, ga_run_stats *stats);

/* Batched eigenray search for n_rcvr receivers, all from p->src: 3D
 * receivers are (x, y) [km] pairs, spherical ones (lat, lon) [deg] pairs.
 * Eigenrays are returned grouped by receiver (launch_idx), then bounce
 * count, in the order the reference finds them; rcvr_latency_ms (may be
 * NULL) receives each receiver's completion time within the call. */
int ga_eig_search(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                  const double *rcvr_xy, int n_rcvr,
                  ga_eigenray *out, int out_cap, int *n_out, double *rcvr_latency_ms,
                  ga_run_stats *stats);

/* -eig_direct: the Levenberg-Marquardt eigenray iteration of find_eigenray
 * alone, from the launch estimate (incl_est_deg, az_est_deg) with `bounces`
 * reflections toward one receiver (rcvr_xy as for ga_eig_search), using
 * e->iterations, e->damping and e->tolerance.  az_est_deg NaN: toward the
 * receiver (sph: great-circle bearing; 3D: straight line), the reference's
 * default.  *found = 1 and out[0] hold the eigenray when the iteration
 * converges, with its eigenray-0.dat rows in pts (GA_E_CAP with *n_pts set
 * when pts_cap is too small); *found = 0 otherwise. */
int ga_eig_direct(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                  const double *rcvr_xy, double incl_est_deg, double az_est_deg, int bounces,
                  ga_eigenray *out, int *found, ga_path_pt *pts, int64_t pts_cap, int64_t *n_pts,
                  ga_run_stats *stats);

/* ga_eig_search plus the path rows of every eigenray found (the rows
 * find_eigenray writes to <id>.eigenray-N.dat, N counting the eigenrays in
 * output order): pts[].launch_idx is the eigenray's index in out.  When
 * pts_cap is too small the call returns GA_E_CAP with *n_pts set to the
 * count needed; the eigenrays are returned either way. */
int ga_eig_search_paths(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                        const double *rcvr_xy, int n_rcvr,
                        ga_eigenray *out, int out_cap, int *n_out, double *rcvr_latency_ms,
                        ga_path_pt *pts, int64_t pts_cap, int64_t *n_pts, ga_run_stats *stats);

/* ---- Weakly non-linear waveforms (-wnl_wvfrm) ---------------------------- */

/* Evolution settings: wvfrm_len, wvfrm_ds, wvfrm_ref and the two filters
 * (defaults of src/util/waveforms.cpp:20-35 and run_wnl_wvfrm: 8192, 0.5,
 * 1.0, (0.5, 4, 0.5), (0.75, 2, 1)). */
typedef struct {
    int32_t len;           /* samples of the source waveform: a power of two, 64 .. 8192 */
    int32_t reserved;
    double  ds;            /* wvfrm_ds (geoac::ds_wvfrm) */
    double  ref;           /* wvfrm_ref: ray length [km] where the source waveform is set */
    double  filt1_g, filt1_n1, filt1_n2;
    double  filt2_g, filt2_n1, filt2_n2;
} ga_wvfrm_params;

typedef struct {
    int64_t ray_steps;     /* RK4 iterations of the ray traces */
    int64_t wv_steps;      /* evolution steps summed over the waveforms */
    double  trace_ms;      /* device time of the ray traces and the per-ray set-up */
    double  evolve_ms;     /* device time of the evolution kernel */
    double  wall_ms;       /* host wall time of the call */
    int32_t n_waveforms;
    int32_t reserved;
} ga_wvfrm_stats;

/* -wnl_wvfrm for n launches (incl, az) [deg] from p->src over a stratified
 * atmosphere (GA_GEOM_SPH, GA_GEOM_3D; GA_GEOM_2D, geoac.eqset.2d.cpp and
 * src/main.infraga.2d.cpp run_wnl_wvfrm, every leg counting toward
 * bounces) or a range-dependent one
 * (GA_GEOM_SPH_RNGDEP, GA_GEOM_3D_RNGDEP, also a stratified profile over
 * terrain; geoac.eqset.{sph,3d}.rngdep.cpp wnl_wvfrm,
 * src/main.infraga.sph.rngdep.cpp:1108-1410,
 * src/main.infraga.3d.rngdep.cpp:1051-1350); p->calc_amp is implied, and,
 * as in run_wnl_wvfrm, no turning-height or time limit stops a ray.  The source
 * waveform t_src[len] [s], p_src[len] [Pa] (a built-in one or a loaded
 * file) is set at ray length w->ref and evolved along every leg to the last
 * ground return.  Per launch i, rows i*len .. i*len+len-1 of t_init / p_init
 * receive <id>.wvfrm_init.dat (travel time to the reference point + t, the
 * source pressure) and of t_out / p_out <id>.wvfrm_out.dat (total travel
 * time + t, the evolved pressure scaled to the arrival).  arrivals[i] is the
 * arrival summary of the last ground return; status[i] is 1 when the ray
 * returned to the ground after every bounce, 0 when it left the region
 * first (the waveform is evolved over the legs traced; the arrival holds
 * only launch angles and travel time) and -1 when the ray never reached the
 * reference point (its rows are NaN).  evolve_steps (may be NULL) receives
 * the evolution steps per launch.  Lengths that are not powers of two are
 * refused (GA_E_ARG). */
int ga_wnl_wvfrm(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_wvfrm_params *w,
                 const double *launches, int n, const double *t_src, const double *p_src,
                 double *t_init, double *p_init, double *t_out, double *p_out,
                 ga_arrival *arrivals, int32_t *status, int64_t *evolve_steps, ga_wvfrm_stats *stats);

/* ga_wnl_wvfrm with the wvfrm_out_step dumps (the sc_spectrum-<L>.dat /
 * sc_waveform-<L>.dat files of wnl_wvfrm, geoac.eqset.sph.strat.cpp:989-1002,
 * and the same in every equation set; out_step from wvfrm_out_step=,
 * main.infraga.sph.cpp:1210): at the top of each evolution step whose
 * previous step size carries the total ray length s + s_prev across a
 * multiple of out_step, the state before the step.  s_prev is ref on the leg
 * from the reference point and the legs' accumulated length after it.
 * Snapshot j of launch i sits at snaps + (i * snap_cap + j) * (1 + len +
 * 2 (len / 2 + 1)): [label L = out_step ceil((s + s_prev) / out_step), the
 * scaled pressure u[len] (sc_waveform's second column; its first is the
 * source time axis), (Re U, Im U) for bins 0 .. len / 2 (sc_spectrum's
 * columns 2-3 at f = bin / (len dt))].  snap_n[i] counts the dumps reached,
 * which may exceed snap_cap (then only the first snap_cap are written).
 * out_step >= 1e4 writes none, as the reference. */
int ga_wnl_wvfrm_dump(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_wvfrm_params *w,
                      const double *launches, int n, const double *t_src, const double *p_src,
                      double *t_init, double *p_init, double *t_out, double *p_out,
                      ga_arrival *arrivals, int32_t *status, int64_t *evolve_steps, ga_wvfrm_stats *stats,
                      double out_step, int snap_cap, double *snaps, int32_t *snap_n);

/* Terrain profile z(r) [km] of the 2D geometry (infraga-2d topo_file=, read
 * by set_region, src/atmo/atmo_io.2d.cpp:138-156, into topo::spline): n >= 3
 * strictly increasing ranges.  Attached to a stratified atmosphere; used by
 * GA_GEOM_2D runs on that id (ground, boundary layer and range limit follow
 * it as set_limits / set_bndlyr do, atmo_io.2d.cpp:23-33,
 * atmo_state.2d.cpp:34-45). */
int ga_topo_load_2d(ga_ctx *ctx, int atmo_id, const double *r_km, const double *z_km, int n);

/* topo::z, dz, ddz of the 2D geometry at range r [km] (atmo_state.2d.cpp:47-69):
 * out = (z, dz/dr, d2z/dr2); without terrain (z0 of the profile bottom, 0, 0). */
int ga_topo_eval_2d(ga_ctx *ctx, int atmo_id, double r_km, double out[3]);

/* ---- Range-dependent atmosphere and topography --------------------------
 *
 * ga_rngdep_load replaces the loader of the range-dependent spherical
 * geometry (set_region, src/atmo/atmo_io.sph.rngdep.cpp:61-200: a
 * lat x lon grid of profiles into interp::hybrid_spline_3D c/u/v/rho
 * splines, src/util/interpolation.cpp:886-1045); node coordinates in degrees
 * (as the node files give them), fields [nlat][nlon][nz] in engine units
 * (km/s, g/cm^3).  ga_topo_load replaces the topography loader (a
 * natural_cubic_spline_2D, interpolation.cpp:266-317) with heights
 * [nlat][nlon] in km.  ga_rngdep_eval evaluates at n points (lat, lon [rad],
 * z [km]) what atmo::c/u/v/rho and their derivatives
 * (atmo_state.sph.rngdep.cpp, interp::eval_all of the hybrid spline) and
 * topo::z/dz/ddz/dddz (atmo_state.sph.rngdep.cpp:59-97) compute:
 * atmo_out[n][4][10] = per field (c, u, v, rho) {f, f_lat, f_lon, f_z,
 * f_latlat, f_lonlon, f_zz, f_latlon, f_latz, f_lonz}; topo_out[n][10] =
 * {z, z_lat, z_lon, z_latlat, z_latlon, z_lonlon, z_latlatlat, z_latlatlon,
 * z_latlonlon, z_lonlonlon} (zeros without topography). */
int ga_rngdep_load(ga_ctx *ctx, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
                   const double *z_km, int nz, const double *c, const double *u, const double *v,
                   const double *rho, int *rd_id);
/* ga_rngdep_load_xy replaces the loader of the 3D Cartesian range-dependent
 * geometry (set_region, src/atmo/atmo_io.3d.rngdep.cpp:56-225): the same
 * node grid with x / y node coordinates in km; a topography loaded on such
 * an id (ga_topo_load) takes its node coordinates in km too.
 * ga_rngdep_load_files_xy reads the reference's x / y node files. */
int ga_rngdep_load_xy(ga_ctx *ctx, const double *x_km, int nx, const double *y_km, int ny, const double *z_km, int nz,
                      const double *c, const double *u, const double *v, const double *rho, int *rd_id);
int ga_rngdep_load_files_xy(ga_ctx *ctx, const char *prefix, const char *x_path, const char *y_path,
                            const char *topo_path, const char *format, int reverse_winds, int *rd_id);
int ga_topo_load(ga_ctx *ctx, int rd_id, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
                 const double *zg_km);
int ga_rngdep_eval(ga_ctx *ctx, int rd_id, const double *pts, int n, double *atmo_out, double *topo_out);
int ga_rngdep_release(ga_ctx *ctx, int rd_id);

/* The reference's file layout: node files of lat / lon degrees (one per
 * line), profiles <prefix><i * nlon + j>.met in format zTuvdp, zuvwTdp or
 * zcuvd, and optionally a terrain file of "lat lon z" rows (lon fastest)
 * (set_region, atmo_io.sph.rngdep.cpp:61-200 / 203-366; file_length and
 * file_2d_dims, src/util/fileIO.cpp:18-85).  topo_path may be NULL. */
int ga_rngdep_load_files(ga_ctx *ctx, const char *prefix, const char *lat_path, const char *lon_path,
                         const char *topo_path, const char *format, int reverse_winds, int *rd_id);

/* The region and ground a loaded range-dependent atmosphere defines
 * (set_limits, atmo_io.sph.rngdep.cpp:27-43; set_bndlyr,
 * atmo_state.sph.rngdep.cpp:42-55): out = {lat_min, lat_max, lon_min,
 * lon_max [rad], alt_max [km], z0 (profile bottom) [km], z_max (terrain
 * maximum + 0.01, or z0) [km], boundary-layer top [km]}. */
int ga_rngdep_info(ga_ctx *ctx, int rd_id, double out[8]);

/* Range-dependent fans run through ga_fan_run / ga_fan_run_device with
 * p->geometry = GA_GEOM_SPH_RNGDEP and atmo_id a ga_rngdep_load id
 * (run_prop, src/main.infraga.sph.rngdep.cpp:412-554).  p->src[0..1] NaN =
 * the region centre; the source sits at max(src[2], terrain).  The launch
 * list is explicit: the reference starts each azimuth's inclinations at the
 * terrain slope along it plus 0.1 degree (main.infraga.sph.rngdep.cpp:
 * 415-428), which a caller reads from ga_rngdep_eval's terrain derivatives. */

/* ---- Request server: continuous batching of fans and eigenray queries ----
 *
 * No counterpart in the reference, whose unit of work is one process per
 * run (run_prop / run_eig_search, src/main.infraga.sph.cpp:200-564,
 * 786-996); it serves callers that submit many requests against resident
 * atmospheres.  A scheduler thread owns the context's GPU work while the
 * server exists (make no other calls on that context meanwhile).  Requests
 * are admitted in arrival order while their device footprint fits the
 * budget; up to n_slots launches run at once, each on its own streams, and
 * their workgroups share the GPU.  Eigenray queries with the same
 * atmosphere, source and parameters join one running search: receivers
 * that arrive while it runs are traced in its next round. */
typedef struct ga_server ga_server;

enum { GA_REQ_FAN = 0, GA_REQ_EIG = 1 };
enum { GA_SERVE_MERGE = 1,        /* merge eigenray queries into running searches */
       GA_SERVE_SERIAL = 2 };     /* one request at a time (one launch per request step) */
#define GA_PENDING 1              /* ga_server_poll: not finished within the timeout */

typedef struct {
    int32_t kind;                 /* GA_REQ_FAN or GA_REQ_EIG */
    int32_t atmo_id;
    ga_prop_params prop;
    ga_eig_params eig;            /* GA_REQ_EIG */
    const double *data;           /* fan: (incl, az) pairs; eigenray: receiver pairs (copied) */
    int32_t n;                    /* launches or receivers */
} ga_request;

typedef struct {
    int32_t status;               /* GA_OK or a GA_E_* code */
    int32_t n_rows;               /* arrivals (fan) or eigenrays */
    double  queue_ms;             /* submit -> admission */
    double  total_ms;             /* submit -> completion */
    int64_t ray_steps;            /* fan requests: RK4 iterations */
} ga_result;

int  ga_server_create(ga_ctx *ctx, int n_slots, double hbm_budget_bytes, int flags, ga_server **out);
/* Thread-safe.  Copies the request's data; *ticket identifies it. */
int  ga_server_submit(ga_server *srv, const ga_request *req, uint64_t *ticket);
/* Thread-safe.  Waits up to timeout_ms (< 0: forever).  Returns GA_PENDING
 * if unfinished; GA_E_CAP if rows_cap is too small (res->n_rows says how
 * many; poll again with a larger buffer); otherwise GA_OK or the request's
 * error, and the request is released. */
int  ga_server_poll(ga_server *srv, uint64_t ticket, int timeout_ms, ga_result *res,
                    ga_arrival *rows, int rows_cap);
/* Device totals since creation. */
void ga_server_stats(ga_server *srv, int64_t *ray_steps, int32_t *launches, double *kernel_ms);
/* Finishes admitted and queued requests, then stops the scheduler. */
void ga_server_destroy(ga_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* INFRAGA_GPU_H_ */
