/*
 * ga_oracle.h -- CPU restatement of infraGA's stratified ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP engine is
 * compared against: only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product path (libinfraga_gpu.so and the
 * infraga-* front ends) never links or calls it.
 *
 * Every function restates the algorithm of the reference file:line named in
 * its comment in ga_oracle.c.  The arithmetic is plain IEEE binary64 in the
 * reference's operation order (compile without -ffast-math and without FMA
 * contraction so it reproduces the reference's rounding).
 *
 * Coverage (geometry x feature):
 *   - 3D Cartesian stratified (src/geoac/geoac.eqset.3d.strat.cpp,
 *     src/atmo/atmo_state.3d.strat.cpp, src/atmo/atmo_state.abs.cart.cpp)
 *   - spherical stratified    (src/geoac/geoac.eqset.sph.strat.cpp,
 *     src/atmo/atmo_state.sph.strat.cpp, src/atmo/atmo_state.abs.sph.cpp)
 *   - flat ground at z0 (no terrain file), boundary-layer vertical wind off
 *     (topo_use_BLw=false, the reference default).
 *   - 18 equations (calc_amp=true) and 6 equations (calc_amp=false).
 *   - -prop fans (write_rays=false bookkeeping) and -eig_search.
 */
#ifndef GA_ORACLE_H_
#define GA_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

#define GA_GEOM_3D  0
#define GA_GEOM_SPH 1

/* Natural cubic spline on a shared abscissa (interpolation.h:36-42). */
typedef struct {
    int n;
    const double *x;
    double *f;
    double *slope;
    int accel;          /* segment cursor (interpolation.h:38) */
} go_spline;

/* Stratified medium: c, u, v, rho splines in z plus ground/boundary layer. */
typedef struct {
    int geom;
    int n;
    double *z;          /* km, shared by all four splines */
    go_spline c, u, v, rho;   /* km/s, km/s, km/s, g/cm^3 */
    double r0;          /* globe radius (sph; globe.cpp:10 uses 6370.0) */
    double z0;          /* ground elevation [km] */
    double z_max;       /* topography maximum (= z0 without terrain) */
    double z_bndlyr;    /* boundary-layer top = z_max + 2 */
    double bndlyr_param;/* 0.1 km (atmo_state.*.cpp) */
    double tweak_abs;   /* Sutherland-Bass scale factor (drivers set 1.0) */
} go_atmo;

/* Solver limits (geoac.params.{sph,3d}.cpp) and configuration. */
typedef struct {
    int calc_amp;       /* 1 -> 18 equations, 0 -> 6 */
    int eq_cnt;
    double s_max, ds_min, ds_max;
    double alt_max, rng_max, time_max;
    double x_min, x_max, y_min, y_max;
    double lat_min, lat_max, lon_min, lon_max;
} go_params;

/* Per-ray launch context (the reference's refs.src_loc/c0/nu0 + theta/phi). */
typedef struct {
    double src[3];      /* sph: r (km from centre), lat, lon [rad]; 3D: x, y, z */
    double c0;
    double nu0;
    double theta, phi;  /* radians, phi = pi/2 - azimuth */
} go_src;

/* One arrival row of <id>.arrivals.dat (main.infraga.sph.cpp:531-544). */
typedef struct {
    int launch_idx;
    int n_bounce;
    double incl_launch, az_launch;   /* degrees, as launched */
    double pos0, pos1;               /* sph: lat, lon [deg]; 3D: x, y [km] */
    double time, cel, turn_ht, incl_arr, back_az, tl, absorption;
} go_arrival;

typedef struct {
    int bounces;
    double freq;
    double turn_ht_min;
    double src0, src1, src_alt;      /* sph: lat, lon [deg], alt [km]; 3D: x, y, z [km] */
} go_prop;

/* One sampled ray-path row (kind 0: <id>.raypaths.dat, <id>.eigenray-N.dat)
 * or caustic point (kind 1: <id>.caustics-N.dat). */
typedef struct {
    int launch_idx;
    int leg;                         /* reflections before the row */
    int step;                        /* row index m within the leg */
    int kind;
    double pos0, pos1;               /* sph: lat, lon [deg]; 3D: x, y [km] */
    double z;                        /* sph: r - r0; 3D: max(z, ground) [km] */
    double tl;                       /* 20 log10(amp) [dB]; 0 without amplitudes or for caustics */
    double absorption;               /* -2 x absorption integral so far [dB] */
    double time;                     /* s */
} go_path_pt;

enum { GO_PATH_RAYS = 1, GO_PATH_CAUSTICS = 2 };

/* ---- atmosphere set-up ------------------------------------------------ */
int    go_atmo_init(go_atmo *a, int geom, const double *z, const double *c,
                    const double *u, const double *v, const double *rho, int n);
void   go_atmo_free(go_atmo *a);
void   go_params_default(go_params *p, int geom, const go_atmo *a, int calc_amp);

/* ---- splines (interpolation.cpp) -------------------------------------- */
void   go_spline_set(go_spline *s);
int    go_find_segment(double x, const double *xv, int n, int *prev);
double go_spline_f(go_spline *s, double x);
double go_spline_df(go_spline *s, double x);
double go_spline_ddf(go_spline *s, double x);
void   go_spline_all(go_spline *s, double x, double *f, double *df, double *ddf);

/* ---- medium (atmo_state.*.cpp) ---------------------------------------- */
double go_c(go_atmo *a, double p0, double p1, double p2);
double go_rho(go_atmo *a, double p0, double p1, double p2);
double go_u(go_atmo *a, double p0, double p1, double p2);
double go_v(go_atmo *a, double p0, double p1, double p2);
double go_dc(go_atmo *a, double p0, double p1, double p2, int n);
double go_du(go_atmo *a, double p0, double p1, double p2, int n);
double go_dv(go_atmo *a, double p0, double p1, double p2, int n);
/* calc_uvw with first (and optionally second) derivatives; dd may be NULL. */
void   go_calc_uvw(go_atmo *a, double p0, double p1, double p2,
                   double *u, double *v, double *w, double du[3], double dv[3], double dw[3],
                   double ddu[6], double ddv[6], double ddw[6]);
double go_sb_alpha(go_atmo *a, double p0, double p1, double p2, double freq);
/* Sutherland-Bass [dB/km] from point values: c [km/s], rho [g/cm^3], height
 * z [km]; sph selects the spherical source file's variant. */
double go_sb_alpha_at(double c, double rho, double z, int sph, double tweak, double freq);
double go_gc_dist(double r0, double lat1, double lon1, double lat2, double lon2);

/* ---- ray equations (geoac.eqset.*.cpp) -------------------------------- */
void   go_set_initial(go_atmo *a, const go_params *p, go_src *s, double *y0);
void   go_rhs(go_atmo *a, const go_params *p, const go_src *s, const double *y, double *dyds);
double go_set_ds(go_atmo *a, const go_params *p, const double *y);
int    go_break_check(go_atmo *a, const go_params *p, const go_src *s, double *y);
int    go_ground_check(go_atmo *a, const double *y);
void   go_set_refl(go_atmo *a, const go_params *p, const go_src *s,
                   const double *row_km2, const double *row_km1, double *y0);
double go_travel_time_seg(go_atmo *a, const double *y0, const double *y1);
double go_atten_seg(go_atmo *a, const double *y0, const double *y1, double freq);
double go_jacobian(go_atmo *a, const double *y);
double go_amp(go_atmo *a, const go_src *s, const double *y);
double go_eikonal_resid(go_atmo *a, const go_src *s, const double *y);

/* One RK4 step y -> y1 (rk4solver.cpp:29-55); returns ds. */
double go_rk4_step(go_atmo *a, const go_params *p, const go_src *s,
                   const double *y, double *y1);

/* Integrate one leg (rk4solver.cpp:14-85) into rows[0..]; rows[0] preset.
 * Returns k (index of the last row written); *check = leave-region flag. */
int    go_prop_rk4(go_atmo *a, const go_params *p, const go_src *s,
                   double *rows, int length, int *check);

/* -prop fan with write_rays=false bookkeeping (main.infraga.sph.cpp:418-558,
 * main.infraga.3d.cpp:407-548).  launches are (incl_deg, az_deg) pairs.
 * Returns 0, or -1 when out_cap was too small (n_out still counts all). */
int    go_fan(go_atmo *a, const go_params *p, const go_prop *pr,
              const double *launches, int n_launch,
              go_arrival *out, int out_cap, int *n_out, long long *ray_steps);

/* -prop fan with write_rays / write_caustics bookkeeping
 * (main.infraga.sph.cpp:450-496, main.infraga.3d.cpp:442-487): per leg the
 * segments of rows 1..k-1 are added one by one to the running travel time
 * and absorption (the segment into the sub-ground row k is not), rows
 * m == 1 or m % 15 == 0 are sampled (flags & GO_PATH_RAYS), and a sign
 * change of the Jacobian determinant between rows m-1 and m is a caustic
 * (flags & GO_PATH_CAUSTICS, needs calc_amp).  Arrivals as go_fan otherwise.
 * Returns 0, or -1 when out_cap or pts_cap was too small (counts still
 * complete). */
int    go_fan_paths(go_atmo *a, const go_params *p, const go_prop *pr,
                    const double *launches, int n_launch, int flags,
                    go_arrival *out, int out_cap, int *n_out,
                    go_path_pt *pts, long long pts_cap, long long *n_pts, long long *ray_steps);

/* The accepted eigenray's path file <id>.eigenray-N.dat
 * (geoac.eigenray.3d.cpp:223-276, geoac.eigenray.sph.cpp:228-271) and its
 * arrival row, for launch angles in degrees.  src as for the searches:
 * 3D (x, y, z) [km]; sph (lat [deg], lon [deg], alt [km]).  Rows m == 1 or
 * m % 15 == 0 of every leg; absorption is stored as -2 x the integral in
 * every leg (the 3D file prints the bare integral after the first
 * reflection: a text-format matter of the writer).  Returns the number of
 * rows (all counted, at most cap stored). */
long long go_eig_path(go_atmo *a, const go_params *p, const double src[3], double incl_deg, double az_deg,
                      int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A);

/* Eigenray search for one receiver, 3D geometry (main.infraga.3d.cpp:944-951,
 * geoac.eigenray.3d.cpp).  Returns the number of eigenrays written to out. */
/* The row find_eigenray writes for an accepted 3D eigenray with these launch
 * angles (degrees) and bounces. */
int    go_eig_row_3d(go_atmo *a, const go_params *p, const double src[3], double incl_deg, double az_deg,
                     int bnc_cnt, double freq, go_arrival *A);
int    go_eig_find_3d(go_atmo *a, const go_params *p, const double src_in[3], const double rcvr[2], double th_est,
                      double ph_est, int bnc, int iterations, double damping, double tolerance, double freq,
                      go_arrival *A, long long *ray_steps);
int    go_eig_find_3d_traj(go_atmo *a, const go_params *p, const double src_in[3], const double rcvr[2],
                           double th_est, double ph_est, int bnc, int iterations, double damping, double tolerance,
                           double freq, go_arrival *A, double *traj, int cap, int *n_traj);
int    go_eig_search_3d(go_atmo *a, const go_params *p, const double src[3],
                        const double rcvr[2], double th_min_deg, double th_max_deg,
                        int bnc_min, int bnc_max, int iterations, double damping,
                        double tolerance, double az_err_lim, double dth_big_deg,
                        double dth_sml_deg, double freq,
                        go_arrival *out, int out_cap, long long *ray_steps);

/* Eigenray search for one receiver, spherical geometry
 * (main.infraga.sph.cpp:962-989, geoac.eigenray.sph.cpp).  src = (lat [deg],
 * lon [deg], alt [km]), rcvr = (lat, lon) [deg]. */
int    go_eig_search_sph(go_atmo *a, const go_params *p, const double src[3],
                         const double rcvr[2], double th_min_deg, double th_max_deg,
                         int bnc_min, int bnc_max, int iterations, double damping,
                         double tolerance, double az_err_lim, double dth_big_deg,
                         double dth_sml_deg, double freq,
                         go_arrival *out, int out_cap, long long *ray_steps);

/* -eig_direct: find_eigenray from the launch estimate (degrees; az_est_deg
 * NaN = toward the receiver) with bnc reflections.  src and rcvr as for
 * the searches of the geometry.  Returns 1 with *A filled when it
 * converges, else 0. */
int    go_eig_direct(go_atmo *a, const go_params *p, const double src[3], const double rcvr[2],
                     double incl_est_deg, double az_est_deg, int bnc, int iterations, double damping,
                     double tolerance, double freq, go_arrival *A, long long *ray_steps);

/* A medium's operations for the eigenray search (the stratified media and
 * the range-dependent ones of ga_oracle_rd.c).  Coordinates are those of the
 * geometry's rows: spherical (r, lat, lon) with topo_z a radius and topo_dz
 * n = 0 lat, 1 lon; 3D (x, y, z) with topo_z a height and topo_dz n = 0 x,
 * 1 y. */
typedef struct go_eig_med {
    void *m;
    double r0;
    void   (*set_initial)(void *m, const go_params *p, go_src *s, double *y0);
    int    (*prop_rk4)(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check);
    void   (*set_refl)(void *m, const go_params *p, const go_src *s, const double *rk2, const double *rk1,
                       double *out);
    double (*tt_seg)(void *m, const double *y0, const double *y1);
    double (*att_seg)(void *m, const double *y0, const double *y1, double freq);
    double (*c)(void *m, double r, double lat, double lon);
    double (*amp)(void *m, const go_src *s, const double *y);
    double (*topo_z)(void *m, double lat, double lon);
    double (*topo_dz)(void *m, double lat, double lon, int n);
    int is_topo;      /* geoac::is_topo: the 3D LM update's terrain-slope form */
} go_eig_med;

/* The 3D search (est_eigenray + find_eigenray of geoac.eigenray.3d.cpp)
 * over a medium: src = (x, y, z [km]) already raised to the ground, rcvr =
 * (x, y) [km], th_* [rad].  Returns the eigenray count. */
int    go_eig_search_3d_med(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                            double th_min, double th_max, int bnc_min, int bnc_max, int iterations,
                            double damping, double tolerance, double az_err_lim, double dth_big_deg,
                            double dth_sml_deg, double freq, go_arrival *out, int out_cap, long long *ray_steps);
/* find_eigenray alone over a medium from (th, ph) [rad]: 1 with *A when it accepts. */
int    go_eig_find_3d_med(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                          double th, double ph, int bnc, int iterations, double damping, double tolerance,
                          double freq, go_arrival *A, long long *ray_steps);
/* find_eigenray's ground-point derivatives at row K (yk):
 * J = {dx/dth, dy/dth, dx/dph, dy/dph}. */
void   go_eig_lm_jac_3d(const go_eig_med *M, const double src[3], const double *yk, double J[4]);
/* Rows k-1 and k of the first leg of an 18-equation trace; k, or -1 on a break. */
int    go_eig_trace_leg_3d(const go_eig_med *M, const go_params *p, const double src[3], double th, double ph,
                           double y_km1[18], double y_k[18]);
/* The accepted 3D eigenray's path rows (pts may be NULL) and row for launch
 * angles th, ph [rad] over a medium. */
long long go_eig_path_3d_med(const go_eig_med *M, const go_params *p, const double src[3], double th, double ph,
                             int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A);

/* The spherical search over a medium: src = (alt [km], lat, lon [rad]) already
 * raised to the ground, rcvr [rad], th_* [rad]; th_grnd is the lowest
 * inclination the driver starts from.  Returns the eigenray count. */
int    go_eig_search_sph_med(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                             double th_min, double th_max, double th_grnd, int bnc_min, int bnc_max,
                             int iterations, double damping, double tolerance, double az_err_lim,
                             double dth_big_deg, double dth_sml_deg, double freq,
                             go_arrival *out, int out_cap, long long *ray_steps);
/* The accepted spherical eigenray's path rows (pts may be NULL) and row for
 * launch angles th, ph [rad] over a medium. */
long long go_eig_path_sph_med(const go_eig_med *M, const go_params *p, const double src[3], double th, double ph,
                              int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A);

/* Work counters of the last go_eig_search_3d / go_eig_search_sph call. */
enum { GO_EIG_EST, GO_EIG_SWEEP_TRACES, GO_EIG_SWEEP_CORR, GO_EIG_SWEEP_ADAPT, GO_EIG_LM_TRACES,
       GO_EIG_FINAL_TRACES, GO_EIG_FINDS, GO_EIG_FOUND, GO_EIG_NCOUNT };
void   go_eig_counts(long long out[GO_EIG_NCOUNT]);

/* ---- 2D effective-sound-speed geometry (infraga-2d) ------------------- */
typedef struct {
    go_atmo *a;          /* stratified splines; a->z0, z_max, z_bndlyr as set here */
    go_spline topo;      /* terrain z(r) when is_topo */
    int is_topo;
    int calc_amp;        /* 8 equations with R, Z, mu_r, mu_z; 4 without */
    double z_reflect, alt_max, rng_max, s_max, ds_min, ds_max;
    double theta, phi;   /* launch inclination, phi = pi/2 - azimuth [rad] */
    double z_src, c_eff0;
} go_2d;

void   go2_init(go_2d *g, go_atmo *a);
int    go2_set_topo(go_2d *g, const double *r, const double *z, int n);
void   go2_set_ground(go_2d *g, double z0);
void   go2_free(go_2d *g);
void   go2_set_initial(go_2d *g, double *y, double r0, double z0);
int    go2_prop_rk4(go_2d *g, double *rows, int length, int *check);
void   go2_set_refl(go_2d *g, const double *rk2, const double *rk1, double *out);
double go2_travel_time_seg(go_2d *g, const double *y0, const double *y1);
double go2_atten_seg(go_2d *g, const double *y0, const double *y1, double freq);
double go2_jacobian(go_2d *g, const double *y);
double go2_amp(go_2d *g, const double *y);
double go2_est_dev(go_2d *g, const double *rows, int k);
double go2_topo_z(go_2d *g, double r);   /* topo::z(r): terrain, or the ground z0 */
double go2_ceff(go_2d *g, double z);     /* c + u cos(phi) + v sin(phi), no boundary-layer scaling in 2D */
double go2_rho(go_2d *g, double z);
int    go2_fan(go_2d *g, const double *launches, int n_launch, int bounces, double src_alt, double freq, int flags,
               go_arrival *out, int out_cap, int *n_out, go_path_pt *pts, long long pts_cap, long long *n_pts,
               long long *ray_steps);
void  *go2_h_new(void *atmo_h);
void   go2_h_free(void *h);
void   go2_h_config(void *h, int calc_amp, double z_reflect, double alt_max, double rng_max, double s_max,
                    double ds_min, double ds_max);

/* ---- opaque-handle helpers for test bindings ------------------------- */
void  *go_h_atmo(int geom, const double *z, const double *c, const double *u,
                 const double *v, const double *rho, int n);
void   go_h_free(void *h);
void   go_h_set_ground(void *h, double z0);
void   go_h_set_tweak(void *h, double t);
void   go_h_set_r0(void *h, double r0);
void   go_h_tables(void *h, int which, double *x, double *f, double *slope);
void   go_h_params(void *h, int calc_amp, go_params *p);
int    go_h_fan_paths(void *h, const go_params *p, const go_prop *pr, const double *launches, int n, int flags,
                      go_arrival *out, int cap, int *n_out, go_path_pt *pts, long long pts_cap, long long *n_pts,
                      long long *steps);
int    go_h_fan(void *h, const go_params *p, const go_prop *pr, const double *launches, int n,
                go_arrival *out, int cap, int *n_out, long long *steps);

#ifdef __cplusplus
}
#endif
#endif
