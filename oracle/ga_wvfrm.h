/*
 * ga_wvfrm.h -- CPU restatement of infraGA's weakly non-linear waveform
 * evolution along a ray (-wnl_wvfrm).
 *
 * TEST INFRASTRUCTURE ONLY, like ga_oracle.h: tests/ and the bench's CPU
 * baseline may load it; the product path never does.
 *
 * Follows (paths relative to the infraGA source tree):
 *   src/util/waveforms.cpp               built-in source waveforms, file loader
 *   src/geoac/geoac.eqset.sph.strat.cpp  wnl_wvfrm (sph), :853-1148
 *   src/geoac/geoac.eqset.3d.strat.cpp   wnl_wvfrm (3D)
 *   src/main.infraga.sph.cpp             run_wnl_wvfrm, :1142-1479
 *   src/main.infraga.3d.cpp              run_wnl_wvfrm
 *   src/geoac/geoac.eqset.{sph,3d}.rngdep.cpp, src/main.infraga.{sph,3d}.rngdep.cpp
 *                                        the range-dependent forms
 * The reference transforms with FFTW (r2c / c2r, unnormalised), which this
 * image lacks; go_fft is a radix-2 transform with FFTW's sign convention
 * and scaling, so the waveform lengths here are powers of two.
 */
#ifndef GA_WVFRM_H_
#define GA_WVFRM_H_

#include "ga_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* In-place complex DFT of length n (a power of two): X[k] = sum x[j]
 * exp(-+2 pi i jk / n) (forward: minus, as FFTW_FORWARD), unnormalised. */
int  go_fft(double *re, double *im, int n, int inverse);
/* r2c: the n/2 + 1 bins of a real series; c2r: the real series of a
 * Hermitian spectrum given by its n/2 + 1 bins (the imaginary parts of bins
 * 0 and n/2 ignored), unnormalised (FFTW's plan semantics). */
int  go_rfft(const double *x, int n, double *re, double *im);
int  go_irfft(const double *re, const double *im, int n, double *x);

/* Built-in waveforms (waveforms.cpp:37-145): option 0 impulse, 1 U-wave,
 * 2 N-wave, any other value the unrecognised-option impulse; *alpha is
 * raised to the builder's floor as the reference does.  Fills t[len], p[len]. */
void go_wvfrm_build(int option, double p0, double t0, double *alpha, int len, int n_rise, int n_period,
                    double *t, double *p);

/* Evolution settings (geoac::ds_wvfrm and the wvfrm:: filter parameters). */
typedef struct {
    int len;
    double ds;                       /* geoac::ds_wvfrm */
    double filt1_g, filt1_n1, filt1_n2;
    double filt2_g, filt2_n1, filt2_n2;
} go_wv_params;

/* A medium for the waveform: ray tracing and point functions in the
 * geometry's row coordinates (sph: r, lat, lon; 3D: x, y, z). */
typedef struct {
    void *m;
    int sph;
    double r0;
    void   (*set_initial)(void *m, const go_params *p, go_src *s, double *y0);
    int    (*prop_rk4)(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check);
    void   (*set_refl)(void *m, const go_params *p, const go_src *s, const double *rk2, const double *rk1,
                       double *out);
    double (*tt_seg)(void *m, const double *y0, const double *y1);
    double (*att_seg)(void *m, const double *y0, const double *y1, double freq);
    double (*c)(void *m, double a, double b, double c);
    double (*rho)(void *m, double a, double b, double c);
    void   (*winds)(void *m, double a, double b, double c, double w[3]);   /* u, v, w */
    double (*sb)(void *m, double a, double b, double c, double freq);
    double (*jac)(void *m, const double *y);
    double (*amp)(void *m, const go_src *s, const double *y);
    double (*topo_z)(void *m, double a, double b);      /* ground height under the row position [km] */
    double ds_floor;          /* smallest evolution step: 1e-7 (sph), 1e-5 (3D), 1e-6 (range-dependent sets) */
} go_wv_med;

go_wv_med go_wv_strat_med(go_atmo *a);
/* A range-dependent medium (ga_oracle_rd.h handle; spherical or Cartesian,
 * with the stratified semantics when go_rd_h_set_strat is on). */
go_wv_med go_wv_rd_med(void *rd);

/* wnl_wvfrm over rows k1..k2-1 (18 doubles each) of one leg: evolves the
 * scaled waveform u (time t[len], values in place) and returns the leg
 * length s_tot.  c_src: c at the ray's source (the beta0 normalisation);
 * (c0, cg0, nu0, rho0, D0, p0): the reference point's quantities;
 * *n_steps receives the number of evolution steps (may be NULL). */
/* wvfrm_out_step dumps of the runs that follow (out_step >= 1e4: none);
 * snapshots of 1 + len + 2 (len / 2 + 1) doubles, at most cap written. */
void go_wv_dump_set(double out_step, double *snap, int cap);
int go_wv_dump_count(void);
double go_wnl_evolve(const go_wv_med *M, const go_wv_params *W, const double *rows, int k1, int k2, double c_src,
                     double c0, double cg0, double nu0, double rho0, double D0, double p0, double *t, double *u,
                     long long *n_steps);

/* run_wnl_wvfrm's ray and waveform (sph: src = lat, lon [deg], alt [km];
 * 3D: x, y, z [km]; incl, az [deg]).  In: the source waveform t[len],
 * p[len] [Pa] (a built-in or loaded one).  Out: t_init / p_init (the
 * wvfrm_init.dat rows), t_out / p_out (the wvfrm_out.dat rows), arrival[8]
 * = {pos0, pos1, time, cel, turn_ht, incl_arr, back_az, tl} with
 * absorption in arrival[8], and *ok = 1 when the ray returns to the ground.
 * Returns the total evolution steps, or -1 when no reference point is
 * reached. */
long long go_wnl_run(const go_wv_med *M, const go_params *p, const go_wv_params *W, const double src_in[3],
                     double incl_deg, double az_deg, int bounces, double freq, double wvfrm_ref, const double *t_src,
                     const double *p_src, double *t_init, double *p_init, double *t_out, double *p_out,
                     double arrival[9], int *ok);

/* go_wnl_run over a stratified atmosphere handle (ctypes). */
long long go_h_wnl_run(void *h, const go_params *p, const go_wv_params *W, const double src_in[3], double incl_deg,
                       double az_deg, int bounces, double freq, double wvfrm_ref, const double *t_src,
                       const double *p_src, double *t_init, double *p_init, double *t_out, double *p_out,
                       double arrival[9], int *ok);

/* The 2D geometry (geoac.eqset.2d.cpp wnl_wvfrm, main.infraga.2d.cpp
 * run_wnl_wvfrm) over a go_2d (ga_oracle.h): source at range 0, altitude
 * src_alt; arrival[0] = range, [2] time, [3] celerity, [4] turning height,
 * [7] TL, [8] absorption. */
double go2_wnl_evolve(go_2d *g, const go_wv_params *W, const double *rows, int k1, int k2, double c0, double rho0,
                      double D0, double p0, double *t, double *u, long long *n_steps);
long long go2_wnl_run(go_2d *g, const go_wv_params *W, double src_alt, double incl_deg, double az_deg, int bounces,
                      double freq, double wvfrm_ref, const double *t_src, const double *p_src, double *t_init,
                      double *p_init, double *t_out, double *p_out, double arrival[9], int *ok);
long long go2_h_wnl_run(void *h, const go_wv_params *W, double src_alt, double incl_deg, double az_deg, int bounces,
                        double freq, double wvfrm_ref, const double *t_src, const double *p_src, double *t_init,
                        double *p_init, double *t_out, double *p_out, double arrival[9], int *ok);

/* go_wnl_run over a range-dependent handle (ctypes). */
long long go_h_rd_wnl_run(void *rd, const go_params *p, const go_wv_params *W, const double src_in[3],
                          double incl_deg, double az_deg, int bounces, double freq, double wvfrm_ref,
                          const double *t_src, const double *p_src, double *t_init, double *p_init, double *t_out,
                          double *p_out, double arrival[9], int *ok);

#ifdef __cplusplus
}
#endif

#endif /* GA_WVFRM_H_ */
