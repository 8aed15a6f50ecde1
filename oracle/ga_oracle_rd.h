/*
 * ga_oracle_rd.h -- CPU restatement of infraGA's spherical range-dependent
 * propagation path (infraga-sph-rngdep -prop).
 *
 * TEST INFRASTRUCTURE ONLY, like ga_oracle.h: tests/, the smoke check and
 * the bench's CPU baseline may load it; the product path never does.
 *
 * Follows (paths relative to the infraGA source tree):
 *   src/atmo/atmo_state.sph.rngdep.cpp   topography, boundary layer, medium,
 *                                        calc_uvw (both overloads)
 *   src/atmo/atmo_io.sph.rngdep.cpp      set_limits / set_bndlyr
 *   src/atmo/atmo_state.abs.sph.cpp      Sutherland-Bass
 *   src/geoac/geoac.eqset.sph.rngdep.cpp ray equations, reflection off
 *                                        sloped terrain, bookkeeping
 *   src/main.infraga.sph.rngdep.cpp      run_prop (write_rays=false)
 * with the splines of ga_interp.h.
 */
#ifndef GA_ORACLE_RD_H_
#define GA_ORACLE_RD_H_

#include "ga_oracle.h"
#include "ga_interp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    go_h3d c, u, v, rho;      /* hybrid splines over (lat, lon [rad], z [km]) */
    go_s2d topo;              /* terrain height above sea level [km] */
    int is_topo;              /* geoac::is_topo */
    int use_blw;              /* topo::use_BLw (boundary-layer vertical wind) */
    int cart;                 /* 3D Cartesian grid: node axes in km, r0 = 0 */
    int strat;                /* a stratified profile over terrain (infraga-sph /
                                 infraga-3d with topo_file): identical columns */
    int upper;                /* the stratified 3D set's upper-triangle contraction of the
                                 second-derivative tensors (geoac.eqset.3d.strat.cpp:305-310);
                                 set with strat on a Cartesian grid */
    double r0;                /* globe radius, 6370 km */
    double z0, z_max, z_bndlyr;
    double bl_param;          /* topo::bndlyr_param, 0.1 km */
    double vert_wind_grad;    /* topo::vert_wind_grad, 0.5 */
    double tweak_abs;
    /* set_limits: propagation region [rad] and altitude ceiling [km] */
    double lat_min, lat_max, lon_min, lon_max, alt_max;
} go_rd;

/* Fields are [nlat][nlon][nz]: c, u, v in km/s, rho in g/cm^3; node
 * coordinates in degrees (converted as the loader does). */
int    go_rd_init(go_rd *a, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
                  const double *z, int nz, const double *c, const double *u, const double *v,
                  const double *rho);
/* Terrain [nlat][nlon] in km; enables is_topo and re-derives the limits. */
int    go_rd_set_topo(go_rd *a, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
                      const double *zg);
/* z_grnd=... without terrain (main.infraga.sph.rngdep.cpp:294-297). */
void   go_rd_set_ground(go_rd *a, double z0);
void   go_rd_free(go_rd *a);
/* geoac.params.sph.cpp defaults with the region from set_limits. */
void   go_rd_params(const go_rd *a, int calc_amp, go_params *p);
/* Default source: the region centre in degrees (run_prop :242-243). */
void   go_rd_centre(const go_rd *a, double *lat_deg, double *lon_deg);

/* ---- medium, coordinates (r, lat, lon); derivative index 0 r, 1 lat, 2 lon */
double go_rd_topo_z(go_rd *a, double lat, double lon);
double go_rd_topo_dz(go_rd *a, double lat, double lon, int n);
double go_rd_topo_ddz(go_rd *a, double lat, double lon, int n1, int n2);
double go_rd_topo_dddz(go_rd *a, double lat, double lon, int n1, int n2, int n3);
double go_rd_bl_sc(go_rd *a, double r, double lat, double lon);
double go_rd_bl_dsc(go_rd *a, double r, double lat, double lon, int n);
double go_rd_bl_ddsc(go_rd *a, double r, double lat, double lon, int n1, int n2);
double go_rd_c(go_rd *a, double r, double lat, double lon);
double go_rd_rho(go_rd *a, double r, double lat, double lon);
double go_rd_u(go_rd *a, double r, double lat, double lon);
double go_rd_v(go_rd *a, double r, double lat, double lon);
double go_rd_w(go_rd *a, double r, double lat, double lon);
double go_rd_dc(go_rd *a, double r, double lat, double lon, int n);
double go_rd_du(go_rd *a, double r, double lat, double lon, int n);
double go_rd_dv(go_rd *a, double r, double lat, double lon, int n);
double go_rd_dw(go_rd *a, double r, double lat, double lon, int n);
/* ddu == NULL selects the first-derivative overload. */
void   go_rd_calc_uvw(go_rd *a, double r, double lat, double lon, double *u, double *v, double *w,
                      double du[3], double dv[3], double dw[3], double ddu[6], double ddv[6],
                      double ddw[6]);
double go_rd_sb_alpha(go_rd *a, double r, double lat, double lon, double freq);

/* ---- rays -------------------------------------------------------------- */
void   go_rd_set_initial(go_rd *a, const go_params *p, go_src *s, double *y0);
void   go_rd_rhs(go_rd *a, const go_params *p, const double *y, double *dyds);
double go_rd_set_ds(go_rd *a, const go_params *p, const double *y);
int    go_rd_break_check(go_rd *a, const go_params *p, const go_src *s, double *y);
int    go_rd_ground_check(go_rd *a, const double *y);
double go_rd_rk4_step(go_rd *a, const go_params *p, const double *y, double *y1);
int    go_rd_prop_rk4(go_rd *a, const go_params *p, const go_src *s, double *rows, int length,
                      int *check);
void   go_rd_set_refl(go_rd *a, const go_params *p, const go_src *s, const double *row_km2,
                      const double *row_km1, double *y0);
double go_rd_travel_time_seg(go_rd *a, const double *y0, const double *y1);
double go_rd_atten_seg(go_rd *a, const double *y0, const double *y1, double freq);
double go_rd_jacobian(go_rd *a, const double *y);
double go_rd_amp(go_rd *a, const go_src *s, const double *y);
double go_rd_eikonal_resid(go_rd *a, const go_src *s, const double *y);
/* |d/dtheta, d/dphi| of the eikonal residual from the auxiliary variables. */
double go_rd_eikonal_deriv_resid(go_rd *a, const double *y);

/* Lowest inclination [deg] launched at azimuth az_deg from a source at
 * (lat, lon [rad], z_src [km]) (run_prop :415-426). */
double go_rd_theta_grnd(go_rd *a, double lat_src, double lon_src, double z_src, double az_deg);

/* run_prop with write_rays=false (main.infraga.sph.rngdep.cpp:412-554) over
 * explicit (incl_deg, az_deg) launches; pr->src0/src1 in degrees. */
int    go_rd_fan(go_rd *a, const go_params *p, const go_prop *pr, const double *launches,
                 int n_launch, go_arrival *out, int out_cap, int *n_out, long long *ray_steps);

/* run_prop with write_rays / write_caustics (main.infraga.sph.rngdep.cpp:451-490):
 * segments of rows 1..k-1 added one by one, rows m == 1 or m % 15 == 0
 * sampled (flags & GO_PATH_RAYS), Jacobian sign changes recorded
 * (flags & GO_PATH_CAUSTICS, needs calc_amp).  As go_fan_paths otherwise. */
int    go_rd_fan_paths(go_rd *a, const go_params *p, const go_prop *pr, const double *launches,
                       int n_launch, int flags, go_arrival *out, int out_cap, int *n_out,
                       go_path_pt *pts, long long pts_cap, long long *n_pts, long long *ray_steps);

/* infraga-sph-rngdep -eig_search (main.infraga.sph.rngdep.cpp:777-965): the
 * spherical search of geoac.eigenray.sph.cpp over this medium and terrain;
 * src = (lat, lon [deg], alt [km]), rcvr = (lat, lon) [deg]. */
int    go_rd_eig_search(go_rd *a, const go_params *p, const double src_in[3], const double rcvr_in[2],
                        double th_min_deg, double th_max_deg, int bnc_min, int bnc_max, int iterations,
                        double damping, double tolerance, double az_err_lim, double dth_big_deg,
                        double dth_sml_deg, double freq, go_arrival *out, int out_cap, long long *ray_steps);
/* The search's lowest inclination (with terrain: degrees, as the reference computes it). */
double go_rd_eig_theta_grnd(go_rd *a, double src_alt, double lat, double lon);
/* The accepted eigenray's path rows and row for launch angles in degrees. */
long long go_rd_eig_path(go_rd *a, const go_params *p, const double src_in[3], double incl_deg, double az_deg,
                         int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A);

/* ---- opaque handles for ctypes ------------------------------------------ */
/* 3D Cartesian range-dependent geometry (infraga-3d-rngdep) */
int    go_rd_init_xy(go_rd *a, const double *x_km, int nx, const double *y_km, int ny, const double *z, int nz,
                     const double *c, const double *u, const double *v, const double *rho);
void   go_rd3_rhs(go_rd *a, const go_params *p, const double *y, double *f);
void   go_rd3_set_initial(go_rd *a, const go_params *p, go_src *s, double *y);
void   go_rd3_set_refl(go_rd *a, const go_params *p, const go_src *s, const double *rk2, const double *rk1, double *y0);
double go_rd3_jacobian(go_rd *a, const double *y);
int    go_rd3_prop_rk4(go_rd *a, const go_params *p, const go_src *s, double *rows, int length, int *check);
void   go_rd3_seg(go_rd *a, const double *y0, const double *y1, double freq, double *dt, double *datt);
/* infraga-3d-rngdep -eig_search / -eig_direct: the 3D search of
 * geoac.eigenray.3d.cpp over this medium; src = (x, y, z) [km], rcvr = (x, y). */
int    go_rd3_eig_search(go_rd *a, const go_params *p, const double src_in[3], const double rcvr[2],
                         double th_min_deg, double th_max_deg, int bnc_min, int bnc_max, int iterations,
                         double damping, double tolerance, double az_err_lim, double dth_big_deg,
                         double dth_sml_deg, double freq, go_arrival *out, int out_cap, long long *ray_steps);
int    go_rd3_eig_direct(go_rd *a, const go_params *p, const double src_in[3], const double rcvr[2],
                         double incl_est_deg, double az_est_deg, int bnc, int iterations, double damping,
                         double tolerance, double freq, go_arrival *A, long long *ray_steps);
void   go_rd3_lm_jac(go_rd *a, const double src[3], const double *yk, double J[4]);
int    go_rd3_trace_leg(go_rd *a, const go_params *p, const double src[3], double th, double ph, double y_km1[18],
                        double y_k[18]);
long long go_rd3_eig_path(go_rd *a, const go_params *p, const double src_in[3], double incl_deg, double az_deg,
                          int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A);
double go_rd3_amp(go_rd *a, const go_src *s, const double *y);
int    go_rd3_fan(go_rd *a, const go_params *p, const go_prop *pr, const double *launches, int n_launch, int write,
                  int flags, go_arrival *out, int out_cap, int *n_out, go_path_pt *pts, long long pts_cap,
                  long long *n_pts, long long *ray_steps);
void  *go_rd_new_xy(const double *x_km, int nx, const double *y_km, int ny, const double *z, int nz,
                    const double *c, const double *u, const double *v, const double *rho);

void  *go_rd_new(const double *lat_deg, int nlat, const double *lon_deg, int nlon, const double *z,
                 int nz, const double *c, const double *u, const double *v, const double *rho);
void   go_rd_delete(void *h);
void   go_rd_h_config(void *h, int use_blw, double tweak_abs);
void   go_rd_h_limits(void *h, double out[8]);   /* lat/lon limits, alt_max, z0, z_max, z_bndlyr */
/* Stratified-binary semantics over identical columns (see go_rd_h_set_strat). */
void   go_rd_h_set_strat(void *h, int strat);

#ifdef __cplusplus
}
#endif
#endif
