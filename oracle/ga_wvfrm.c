/*
 * ga_wvfrm.c -- CPU restatement of infraGA's weakly non-linear waveform
 * evolution (-wnl_wvfrm).  TEST INFRASTRUCTURE ONLY (see ga_wvfrm.h).
 *
 * Plain binary64 in the reference's operation order, except the transforms:
 * the reference calls FFTW, absent here, so go_fft is an iterative radix-2
 * decimation-in-time transform (checked against numpy.fft by the tests)
 * with FFTW's sign convention and no normalisation.
 */
#include "ga_wvfrm.h"
#include "ga_oracle_rd.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define WV_PI 3.141592653589793238462643   /* geoac.params.h */

static double sq(double x) { return pow(x, 2); }

/* ---- transforms ---------------------------------------------------------- */
static int pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

int go_fft(double *re, double *im, int n, int inverse)
{
    if (!pow2(n)) return -1;
    for (int i = 1, j = 0; i < n; i++) {          /* bit reversal */
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    const double sgn = inverse ? 1.0 : -1.0;
    for (int len = 2; len <= n; len <<= 1) {
        const int h = len >> 1;
        for (int k = 0; k < h; k++) {
            /* twiddle from the exact angle, not a recurrence */
            const double ang = sgn * 2.0 * WV_PI * (double)k / (double)len;
            const double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                const int j = i + h;
                const double tr = re[j] * wr - im[j] * wi, ti = re[j] * wi + im[j] * wr;
                re[j] = re[i] - tr; im[j] = im[i] - ti;
                re[i] += tr;        im[i] += ti;
            }
        }
    }
    return 0;
}

int go_rfft(const double *x, int n, double *re, double *im)
{
    double *a = (double *)malloc(sizeof(double) * 2 * (size_t)n), *b = a + n;
    for (int i = 0; i < n; i++) { a[i] = x[i]; b[i] = 0.0; }
    int rc = go_fft(a, b, n, 0);
    for (int k = 0; k <= n / 2; k++) { re[k] = a[k]; im[k] = b[k]; }
    free(a);
    return rc;
}

int go_irfft(const double *re, const double *im, int n, double *x)
{
    double *a = (double *)malloc(sizeof(double) * 2 * (size_t)n), *b = a + n;
    a[0] = re[0]; b[0] = 0.0;
    a[n / 2] = re[n / 2]; b[n / 2] = 0.0;
    for (int k = 1; k < n / 2; k++) {
        a[k] = re[k]; b[k] = im[k];
        a[n - k] = re[k]; b[n - k] = -im[k];
    }
    int rc = go_fft(a, b, n, 1);
    for (int i = 0; i < n; i++) x[i] = a[i];
    free(a);
    return rc;
}

/* ---- built-in waveforms (waveforms.cpp:37-145) --------------------------- */
static double impulse(double t, double t0, double alpha)   /* :40-54 */
{
    if (t >= 0.0) {
        double x = t / t0, x_pk = (alpha + 1.0) - sqrt(alpha + 1.0);
        double r = pow(x, alpha) * (1.0 - x / (1.0 + alpha)) * exp(-x);
        r /= pow(x_pk, alpha) * (1.0 - x_pk / (1.0 + alpha)) * exp(-x_pk);
        return r;
    }
    return 0.0;
}

static double n_wave(double t, double t0) { return fabs(t) <= t0 ? t / t0 : 0.0; }   /* :56-59 */

void go_wvfrm_build(int option, double p0, double t0, double *alpha, int len, int n_rise, int n_period,
                    double *t, double *p)
{
    double duration, a;
    if (option == 0) {                           /* impulse, :97-108 */
        *alpha = *alpha > 0.01 ? *alpha : 0.01;
        a = *alpha;
        duration = t0 * len * fmax(((1.0 + a) - sqrt(1.0 + a)) / n_rise, 2.0 * 3.14159 * sqrt(1 + a) / n_period);
        for (int n = 0; n < len; n++) {
            t[n] = -duration / 2.0 + duration * n / len;
            p[n] = p0 * impulse(t[n], t0, a);
        }
    } else if (option == 1) {                    /* U-wave, :109-122 */
        *alpha = *alpha > 2.0 ? *alpha : 2.0;
        a = *alpha;
        duration = t0 * len * fmax(((1.0 + a) - sqrt(1.0 + a)) / n_rise, 2.0 * 3.14159 * sqrt(1 + a) / n_period);
        const double eps = 1.0e-2;
        for (int n = 0; n < len; n++) {
            t[n] = -duration / 2.0 + duration * n / len;
            p[n] = p0 * (impulse(t[n] + eps, t0, a) - impulse(t[n] - eps, t0, a)) / (2.0 * eps);
            p[n] /= 1.126 * (impulse(eps, t0, a) - impulse(-eps, t0, a)) / (2.0 * eps);
        }
    } else if (option == 2) {                    /* N-wave, :123-132 */
        duration = (t0 / n_period) * len;
        for (int n = 0; n < len; n++) {
            t[n] = -duration / 2.0 + duration * n / len;
            p[n] = -p0 * n_wave(t[n] - t0 / 2.0, t0);
        }
    } else {                                     /* unrecognised option, :133-144 */
        *alpha = *alpha > 0.01 ? *alpha : 0.01;
        a = *alpha;
        duration = t0 * len * fmin(((1.0 + a) - sqrt(1.0 + a)) / n_rise, 2.0 * 3.14159 * sqrt(1 + a) / n_period);
        for (int n = 0; n < len; n++) {
            t[n] = -duration / 2.0 + duration * n / len;
            p[n] = p0 * impulse(t[n], t0, a);
        }
    }
}

/* ---- linear splines (interpolation.cpp:61-89) ----------------------------- */
typedef struct {
    int n, accel;
    double *x, *f, *sl;
} lin1;

static void lin_set(lin1 *s)
{
    for (int i = 0; i < s->n - 1; i++) s->sl[i] = (s->f[i + 1] - s->f[i]) / (s->x[i + 1] - s->x[i]);
    s->sl[s->n - 1] = s->sl[s->n - 2];
}

static double lin_f(lin1 *s, double x)
{
    int k = go_find_segment(x, s->x, s->n, &s->accel);
    return s->f[k] + (x - s->x[k]) * s->sl[k];
}

/* ---- the stratified media --------------------------------------------------- */
static void st_init(void *m, const go_params *p, go_src *s, double *y) { go_set_initial((go_atmo *)m, p, s, y); }
static int st_prop(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    return go_prop_rk4((go_atmo *)m, p, s, rows, length, check);
}
static void st_refl(void *m, const go_params *p, const go_src *s, const double *a, const double *b, double *o)
{
    go_set_refl((go_atmo *)m, p, s, a, b, o);
}
static double st_tt(void *m, const double *a, const double *b) { return go_travel_time_seg((go_atmo *)m, a, b); }
static double st_att(void *m, const double *a, const double *b, double f) { return go_atten_seg((go_atmo *)m, a, b, f); }
static double st_c(void *m, double a, double b, double c) { return go_c((go_atmo *)m, a, b, c); }
static double st_rho(void *m, double a, double b, double c) { return go_rho((go_atmo *)m, a, b, c); }
static void st_winds(void *m, double a, double b, double c, double w[3])
{
    w[0] = go_u((go_atmo *)m, a, b, c);
    w[1] = go_v((go_atmo *)m, a, b, c);
    w[2] = 0.0;                                   /* atmo::w: no vertical wind without terrain */
}
static double st_sb(void *m, double a, double b, double c, double f) { return go_sb_alpha((go_atmo *)m, a, b, c, f); }
static double st_jac(void *m, const double *y) { return go_jacobian((go_atmo *)m, y); }
static double st_amp(void *m, const go_src *s, const double *y) { return go_amp((go_atmo *)m, s, y); }
static double st_tz(void *m, double a, double b)
{
    (void)a; (void)b;
    return ((go_atmo *)m)->z0;
}

go_wv_med go_wv_strat_med(go_atmo *a)
{
    go_wv_med M = {a, a->geom == GA_GEOM_SPH, a->r0, st_init, st_prop, st_refl, st_tt, st_att, st_c, st_rho,
                   st_winds, st_sb, st_jac, st_amp, st_tz, a->geom == GA_GEOM_SPH ? 1.0e-7 : 1.0e-5};
    return M;
}

/* ---- the range-dependent media (ga_oracle_rd.c): sph rows (r, lat, lon);
 * Cartesian rows (x, y, z), whose point functions take (z, x, y) ---------- */
static void rd_init(void *m, const go_params *p, go_src *s, double *y)
{
    go_rd *a = (go_rd *)m;
    if (a->cart) go_rd3_set_initial(a, p, s, y);
    else go_rd_set_initial(a, p, s, y);
}
static int rd_prop(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    go_rd *a = (go_rd *)m;
    return a->cart ? go_rd3_prop_rk4(a, p, s, rows, length, check) : go_rd_prop_rk4(a, p, s, rows, length, check);
}
static void rd_refl(void *m, const go_params *p, const go_src *s, const double *k2, const double *k1, double *o)
{
    go_rd *a = (go_rd *)m;
    if (a->cart) go_rd3_set_refl(a, p, s, k2, k1, o);
    else go_rd_set_refl(a, p, s, k2, k1, o);
}
static double rd_tt(void *m, const double *y0, const double *y1)
{
    go_rd *a = (go_rd *)m;
    if (!a->cart) return go_rd_travel_time_seg(a, y0, y1);
    double dt, da;
    go_rd3_seg(a, y0, y1, 0.1, &dt, &da);
    return dt;
}
static double rd_att(void *m, const double *y0, const double *y1, double f)
{
    go_rd *a = (go_rd *)m;
    if (!a->cart) return go_rd_atten_seg(a, y0, y1, f);
    double dt, da;
    go_rd3_seg(a, y0, y1, f, &dt, &da);
    return da;
}
#define RD_PT(a, p0, p1, p2) ((a)->cart ? (p2) : (p0)), ((a)->cart ? (p0) : (p1)), ((a)->cart ? (p1) : (p2))
static double rd_c(void *m, double p0, double p1, double p2) { go_rd *a = (go_rd *)m; return go_rd_c(a, RD_PT(a, p0, p1, p2)); }
static double rd_rho(void *m, double p0, double p1, double p2) { go_rd *a = (go_rd *)m; return go_rd_rho(a, RD_PT(a, p0, p1, p2)); }
static void rd_winds(void *m, double p0, double p1, double p2, double w[3])
{
    go_rd *a = (go_rd *)m;
    w[0] = go_rd_u(a, RD_PT(a, p0, p1, p2));
    w[1] = go_rd_v(a, RD_PT(a, p0, p1, p2));
    w[2] = go_rd_w(a, RD_PT(a, p0, p1, p2));
}
static double rd_sb(void *m, double p0, double p1, double p2, double f)
{
    go_rd *a = (go_rd *)m;
    if (!a->cart) return go_rd_sb_alpha(a, p0, p1, p2, f);
    return go_sb_alpha_at(go_rd_c(a, p2, p0, p1), go_rd_rho(a, p2, p0, p1), p2, 0, a->tweak_abs, f);
}
#undef RD_PT
static double rd_jac(void *m, const double *y)
{
    go_rd *a = (go_rd *)m;
    return a->cart ? go_rd3_jacobian(a, y) : go_rd_jacobian(a, y);
}
static double rd_amp(void *m, const go_src *s, const double *y)
{
    go_rd *a = (go_rd *)m;
    return a->cart ? go_rd3_amp(a, s, y) : go_rd_amp(a, s, y);
}
static double rd_tz(void *m, double p0, double p1)
{
    go_rd *a = (go_rd *)m;
    return go_rd_topo_z(a, p0, p1) - a->r0;       /* altitude (r0 = 0 on a Cartesian grid) */
}

/* The range-dependent equation sets floor the evolution step at 1e-6 km
 * (geoac.eqset.sph.rngdep.cpp and 3d.rngdep.cpp, wnl_wvfrm); a stratified
 * profile over terrain runs the stratified sets' wnl_wvfrm. */
go_wv_med go_wv_rd_med(void *h)
{
    go_rd *a = (go_rd *)h;
    go_wv_med M = {a, !a->cart, a->r0, rd_init, rd_prop, rd_refl, rd_tt, rd_att, rd_c, rd_rho, rd_winds, rd_sb,
                   rd_jac, rd_amp, rd_tz, a->strat ? (a->cart ? 1.0e-5 : 1.0e-7) : 1.0e-6};
    return M;
}

/* ---- the step loop of wnl_wvfrm, shared by the geometries ---------------
 * Splines over one abscissa (lx, n rows): D, beta, beta_caustic.  absorb(ctx,
 * s, freq) returns -c / c0 * alpha(freq) at arc length s (the geometry's c
 * and point).  strict: the sph / 2D caustic test (s + ds < s_tot and a
 * strict sign change) instead of the 3D inclusive one. */
typedef double (*wv_absorb_fn)(void *ctx, double s, double freq);

/* wvfrm_out_step dumps (geoac.eqset.sph.strat.cpp:989-1002, the same in every
 * equation set): at the top of a step whose previous step size carries the
 * total ray length s + s_prev across a multiple of out_step (the first step
 * uses ds_wvfrm), the state u, U before the step, labelled
 * out_step * ceil((s + s_prev) / out_step).  s_prev is wvfrm_ref on the leg
 * from the reference point and the ray length accumulated before the later
 * legs (main.infraga.sph.cpp:1394-1399).  A snapshot is [label, u[len],
 * (Re U, Im U)[len/2 + 1]]; count keeps counting past cap. */
static struct {
    double out_step, s_prev;
    double *snap;
    int cap, n;
} wv_dump = {1.0e10, 0.0, NULL, 0, 0};

void go_wv_dump_set(double out_step, double *snap, int cap)
{
    wv_dump.out_step = out_step;
    wv_dump.snap = snap;
    wv_dump.cap = cap;
    wv_dump.n = 0;
    wv_dump.s_prev = 0.0;
}
int go_wv_dump_count(void) { return wv_dump.n; }

static void wv_dump_step(int len, double s, double ds_prev, const double *u, const double *Ur, const double *Ui)
{
    const double os = wv_dump.out_step;
    if (!(os < 1.0e4)) return;
    const double lab = os * ceil((s + wv_dump.s_prev) / os);
    if (!(fabs(os * ceil((s + wv_dump.s_prev + ds_prev) / os) - lab) > 0.0)) return;
    if (wv_dump.snap && wv_dump.n < wv_dump.cap) {
        const int fft_len = len / 2 + 1;
        double *S = wv_dump.snap + (size_t)wv_dump.n * (1 + len + 2 * fft_len);
        S[0] = lab;
        for (int m = 0; m < len; m++) S[1 + m] = u[m];
        for (int b = 0; b < fft_len; b++) { S[1 + len + 2 * b] = Ur[b]; S[2 + len + 2 * b] = Ui[b]; }
    }
    wv_dump.n++;
}

static long long wv_evolve_core(const go_wv_params *W, lin1 *LD, lin1 *Lb, lin1 *Lbc, double s_tot, double x_last,
                                double ds_floor, int strict, wv_absorb_fn absorb, void *actx, const double *t,
                                double *u)
{
    const int len = W->len, fft_len = len / 2 + 1;
    double *buf = (double *)malloc(sizeof(double) * ((size_t)len * 4 + (size_t)fft_len * 10));
    double *fft_time = buf, *u_sqr = fft_time + len, *du_sqr = u_sqr + len, *tukey = du_sqr + len;
    double *filter1 = tukey + len, *filter2 = filter1 + fft_len;
    double *Ur = filter2 + fft_len, *Ui = Ur + fft_len, *Ar = Ui + fft_len, *Ai = Ar + fft_len;
    double *Br = Ai + fft_len, *Bi = Br + fft_len, *Cr = Bi + fft_len, *Ci = Cr + fft_len;

    const double dt = t[1] - t[0];
    const double tp = 0.8;
    for (int m = 0; m < len; m++) {
        if (m < tp * (len - 1) / 2.0) tukey[m] = 1.0 / 2.0 * (1.0 + cos(WV_PI * (2.0 * m / (tp * (len - 1)) - 1.0)));
        else if (m > (len - 1) * (1.0 - tp / 2.0))
            tukey[m] = 1.0 / 2.0 * (1.0 + cos(WV_PI * (2.0 * m / (tp * (len - 1)) - 2.0 / tp + 1.0)));
        else tukey[m] = 1.0;
    }
    for (int m = 0; m < fft_len; m++) {
        filter1[m] = pow(1.0 + pow(m / (W->filt1_g * fft_len), W->filt1_n1), -W->filt1_n2);
        filter2[m] = pow(1.0 + pow(m / (W->filt2_g * fft_len), W->filt2_n1), -W->filt2_n2);
    }
    go_rfft(u, len, Ur, Ui);
    double Uf_max = 0.0;
    for (int m = 0; m < fft_len; m++) Uf_max = fmax(Uf_max, (1.0 / dt * m / len) * sqrt(sq(Ur[m]) + sq(Ui[m])));

    long long steps = 0;
    double s = 0.0, ds = W->ds;
    while (s < s_tot) {
        double bc, bn;
        wv_dump_step(len, s, ds, u, Ur, Ui);
        if (fabs(lin_f(LD, s)) >= 0.5) bc = lin_f(Lb, s);
        else bc = lin_f(Lbc, s);
        ds = W->ds / (WV_PI * Uf_max * bc);
        ds = fmin(ds, 1.0);
        ds = fmax(ds, ds_floor);
        const double sn = fmin(s + ds, x_last);
        if (fabs(lin_f(LD, sn)) >= 0.5) bn = lin_f(Lb, sn);
        else bn = lin_f(Lbc, sn);

        for (int m = 0; m < len; m++) u_sqr[m] = sq(u[m]);
        du_sqr[0] = (u_sqr[1] - u_sqr[0]) / (t[1] - t[0]);
        for (int m = 1; m < len - 1; m++) du_sqr[m] = (u_sqr[m + 1] - u_sqr[m - 1]) / (t[m + 1] - t[m - 1]);
        du_sqr[len - 1] = (u_sqr[len - 1] - u_sqr[len - 2]) / (t[len - 1] - t[len - 2]);
        go_rfft(u_sqr, len, Ar, Ai);
        for (int m = 0; m < len; m++) fft_time[m] = u[m] * du_sqr[m];
        go_rfft(fft_time, len, Br, Bi);
        for (int m = 0; m < len; m++) fft_time[m] = sq(du_sqr[m]);
        go_rfft(fft_time, len, Cr, Ci);

        /* caustic between s and s + ds: phase shift by pi / 2 */
        if (strict ? (s + ds < s_tot) : (s + ds <= s_tot)) {
            const double d0 = lin_f(LD, s), d1 = lin_f(LD, s + ds);
            if (strict ? (d0 * d1 < 0.0) : (d0 * d1 <= 0.0))
                for (int m = 0; m < fft_len; m++) {
                    double tmp = Ur[m];
                    Ur[m] = -Ui[m];
                    Ui[m] = tmp;
                }
        }
        for (int m = 0; m < fft_len; m++) {
            const double freq = 1.0 / dt * m / len;
            const double absorp = pow(10.0, absorb(actx, s, freq) * ds / 10.0);
            Ur[m] -= WV_PI * freq * (ds * (bc + bn) / 2.0 * Ai[m] + pow(ds, 2) / 2.0 * bc * bn * Bi[m] +
                                     pow(ds, 3) / 8.0 * pow(bc, 2) * bn * Ci[m]) * filter1[m];
            Ui[m] += WV_PI * freq * (ds * (bc + bn) / 2.0 * Ar[m] + pow(ds, 2) / 2.0 * bc * bn * Br[m] +
                                     pow(ds, 3) / 8.0 * pow(bc, 2) * bn * Cr[m]) * filter1[m];
            Ur[m] *= absorp * filter2[m];
            Ui[m] *= absorp * filter2[m];
        }
        go_irfft(Ur, Ui, len, fft_time);
        for (int m = 0; m < len; m++) u[m] = fft_time[m] * tukey[m] / len;
        Uf_max = 0.0;
        for (int m = 0; m < fft_len; m++) Uf_max = fmax(Uf_max, (1.0 / dt * m / len) * sqrt(sq(Ur[m]) + sq(Ui[m])));
        s += ds;
        steps++;
    }
    free(buf);
    return steps;
}

/* ---- wnl_wvfrm (geoac.eqset.sph.strat.cpp:853-1148; 3D: eqset.3d.strat.cpp) -- */
typedef struct {
    const go_wv_med *M;
    lin1 *L;                 /* ray position splines */
    double c0;
} wv_pos_absorb;

/* -c(pos) / c0 * alpha(pos, freq) at the ray position of s */
static double absorb_at_pos(void *ctx, double s, double freq)
{
    wv_pos_absorb *q = (wv_pos_absorb *)ctx;
    const go_wv_med *M = q->M;
    const double pa = lin_f(&q->L[0], s), pb = lin_f(&q->L[1], s), pc = lin_f(&q->L[2], s);
    const double cpt = M->c(M->m, pa, pb, pc);
    return -cpt / q->c0 * M->sb(M->m, pa, pb, pc, freq);
}

double go_wnl_evolve(const go_wv_med *M, const go_wv_params *W, const double *rows, int k1, int k2, double c_src,
                     double c0, double cg0, double nu0, double rho0, double D0, double p0, double *t, double *u,
                     long long *n_steps)
{
    const int n = k2 - k1;
    const int sph = M->sph;
    if (n_steps) *n_steps = 0;
    if (n < 2) return 0.0;
    double *lx = (double *)malloc(sizeof(double) * (size_t)n * 13);
    lin1 L[6];   /* ray_r, ray_th, ray_ph, D, beta, beta_caustic */
    for (int i = 0; i < 6; i++) {
        L[i].n = n; L[i].accel = 0;
        L[i].x = lx;                          /* one shared abscissa s */
        L[i].f = lx + n * (1 + 2 * i);
        L[i].sl = lx + n * (2 + 2 * i);
    }
    const double beta0 = 1.2 * p0 / ((rho0 * 1.0e3) * pow(c0 * 1.0e3, 2)) * (nu0 * c0) / (cg0 * c_src);
    double s = 0.0;
    for (int k = 0; k < n; k++) {
        const double *y = rows + 18 * (size_t)(k1 + k);
        double c = M->c(M->m, y[0], y[1], y[2]);
        double rho = M->rho(M->m, y[0], y[1], y[2]);
        double nu = sqrt(sq(y[3]) + sq(y[4]) + sq(y[5]));
        lx[k] = s;
        L[0].f[k] = y[0]; L[1].f[k] = y[1]; L[2].f[k] = y[2];
        L[3].f[k] = M->jac(M->m, y);
        L[4].f[k] = beta0 * sqrt(fabs((D0 * rho0 * pow(cg0, 3)) / (L[3].f[k] * rho * pow(c, 3)) * (c * pow(nu, 3)) /
                                      (c0 * pow(nu0, 3))));
        if (k < 1) L[5].f[k] = 0.0;
        else
            L[5].f[k] = beta0 * sqrt(fabs((D0 * rho0 * pow(cg0, 3)) /
                                          ((L[3].f[k] - L[3].f[k - 1]) / (lx[k] - lx[k - 1]) * rho * pow(c, 3)) *
                                          (c * pow(nu, 3)) / (c0 * pow(nu0, 3))));
        if ((k + 1) < n) {
            const double *z = rows + 18 * (size_t)(k1 + k + 1);
            if (sph)
                s += sqrt(pow(z[0] - y[0], 2) + pow(y[0] * (z[1] - y[1]), 2) + pow(y[0] * cos(y[1]) * (z[2] - y[2]), 2));
            else
                s += sqrt(pow(z[0] - y[0], 2) + pow(z[1] - y[1], 2) + pow(z[2] - y[2], 2));
        }
    }
    L[5].f[0] = L[5].f[1];
    const double s_tot = s;
    for (int i = 0; i < 6; i++) lin_set(&L[i]);
    wv_pos_absorb q = {M, L, c0};
    const long long steps = wv_evolve_core(W, &L[3], &L[4], &L[5], s_tot, lx[n - 1], M->ds_floor, sph, absorb_at_pos,
                                           &q, t, u);
    free(lx);
    if (n_steps) *n_steps = steps;
    return s_tot;
}

/* ---- run_wnl_wvfrm (main.infraga.sph.cpp:1142-1479, main.infraga.3d.cpp) --- */
long long go_wnl_run(const go_wv_med *M, const go_params *p_in, const go_wv_params *W, const double src_in[3],
                     double incl_deg, double az_deg, int bounces, double freq, double wvfrm_ref, const double *t_src,
                     const double *p_src, double *t_init, double *p_init, double *t_out, double *p_out,
                     double arrival[9], int *ok)
{
    go_params p = *p_in;
    p.calc_amp = 1;
    p.eq_cnt = 18;
    const int len = W->len;
    const int length = (int)(p.s_max / p.ds_min);
    double *rows = (double *)malloc(sizeof(double) * 18 * (size_t)length);
    double *u = (double *)malloc(sizeof(double) * 2 * (size_t)len), *t = u + len;
    for (int i = 0; i < len; i++) { t[i] = t_src[i]; u[i] = p_src[i]; }
    go_src s;
    memset(&s, 0, sizeof(s));
    double z_src = src_in[2];
    if (M->sph) {
        s.src[1] = src_in[0] * (WV_PI / 180.0);
        s.src[2] = src_in[1] * (WV_PI / 180.0);
        const double zg = M->topo_z(M->m, s.src[1], s.src[2]);   /* z_src = max(z_src, topo::z - r0) */
        z_src = z_src > zg ? z_src : zg;
        s.src[0] = M->r0 + z_src;
    } else {
        const double zg = M->topo_z(M->m, src_in[0], src_in[1]);
        z_src = z_src > zg ? z_src : zg;
        s.src[0] = src_in[0]; s.src[1] = src_in[1]; s.src[2] = z_src;
    }
    s.theta = incl_deg * (WV_PI / 180.0);
    s.phi = WV_PI / 2.0 - az_deg * (WV_PI / 180.0);
    M->set_initial(M->m, &p, &s, rows);
    const double c_src = M->c(M->m, s.src[0], s.src[1], s.src[2]);

    double tt = 0.0, att = 0.0, rmax = 0.0, ray_length = 0.0;
    double c0 = 0, rho0 = 0, nu0 = 0, cg0 = 0, D0 = 0, p0 = 0;
    int started = 0, brk = 0, k = 0;
    long long steps = 0;
    for (int bnc = 0; bnc <= bounces; bnc++) {
        k = M->prop_rk4(M->m, &p, &s, rows, length, &brk);
        for (int m = 1; m < k; m++) {
            tt += M->tt_seg(M->m, rows + 18 * (m - 1), rows + 18 * m);
            att += M->att_seg(M->m, rows + 18 * (m - 1), rows + 18 * m, freq);
            const double h = M->sph ? rows[18 * m] - M->r0 : rows[18 * m + 2];
            rmax = rmax > h ? rmax : h;
        }
        if (!started) {
            for (int kr = 0; kr < k; kr++) {
                const double *a = rows + 18 * kr, *b = rows + 18 * (kr + 1);
                if (M->sph) {
                    const double dr = b[0] - a[0], r = a[0] + dr / 2.0;
                    const double dt_ = b[1] - a[1], th = a[1] + dt_ / 2.0;
                    const double dp = b[2] - a[2];
                    ray_length += sqrt(pow(dr, 2) + pow(r * dt_, 2) + pow(r * cos(th) * dp, 2));
                } else {
                    ray_length += sqrt(pow(b[0] - a[0], 2) + pow(b[1] - a[1], 2) + pow(b[2] - a[2], 2));
                }
                if (ray_length >= wvfrm_ref) {
                    const double *y = b;
                    double w[3];
                    c0 = M->c(M->m, y[0], y[1], y[2]);
                    rho0 = M->rho(M->m, y[0], y[1], y[2]);
                    nu0 = sqrt(pow(y[3], 2) + pow(y[4], 2) + pow(y[5], 2));
                    M->winds(M->m, y[0], y[1], y[2], w);
                    double g[3];
                    if (M->sph) {   /* (r, theta, phi) components: w, v, u */
                        g[0] = c0 * y[3] / nu0 + w[2];
                        g[1] = c0 * y[4] / nu0 + w[1];
                        g[2] = c0 * y[5] / nu0 + w[0];
                    } else {
                        g[0] = c0 * y[3] / nu0 + w[0];
                        g[1] = c0 * y[4] / nu0 + w[1];
                        g[2] = c0 * y[5] / nu0 + w[2];
                    }
                    cg0 = sqrt(pow(g[0], 2) + pow(g[1], 2) + pow(g[2], 2));
                    D0 = M->jac(M->m, y);
                    /* wvfrm_init.dat: travel_time(solution, k_ref + 1) + t */
                    double tref = 0.0;
                    for (int q = 0; q < kr + 1; q++) tref += M->tt_seg(M->m, rows + 18 * q, rows + 18 * (q + 1));
                    for (int i = 0; i < len; i++) { t_init[i] = tref + t[i]; p_init[i] = u[i]; }
                    p0 = 0.0;
                    for (int i = 0; i < len; i++) p0 = fmax(p0, fabs(u[i]));
                    for (int i = 0; i < len; i++) u[i] /= p0;
                    started = 1;
                    long long ns = 0;
                    wv_dump.s_prev = wvfrm_ref;
                    ray_length = go_wnl_evolve(M, W, rows, kr + 1, k, c_src, c0, cg0, nu0, rho0, D0, p0, t, u, &ns);
                    steps += ns;
                    break;
                }
            }
        } else {
            long long ns = 0;
            wv_dump.s_prev = ray_length;
            ray_length += go_wnl_evolve(M, W, rows, 0, k, c_src, c0, cg0, nu0, rho0, D0, p0, t, u, &ns);
            steps += ns;
        }
        if (brk) break;
        M->set_refl(M->m, &p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
    }
    const double *yk = rows + 18 * (size_t)k;
    /* arrival summary and inclination (sph :1399-1403; 3D: c at the terrain under row k) */
    if (M->sph) {
        arrival[0] = yk[1] * (180.0 / WV_PI);
        arrival[1] = yk[2] * (180.0 / WV_PI);
        arrival[2] = tt;
        arrival[3] = go_gc_dist(M->r0, yk[1], yk[2], s.src[1], s.src[2]) / tt;
        arrival[5] = -asin(M->c(M->m, yk[0], yk[1], yk[2]) / M->c(M->m, M->r0 + z_src, s.src[1], s.src[2]) * yk[3]) *
                     (180.0 / WV_PI);
        double baz = 90.0 - atan2(-yk[4], -yk[5]) * (180.0 / WV_PI);
        while (baz < -180.0) baz += 360.0;
        while (baz > 180.0) baz -= 360.0;
        arrival[6] = baz;
    } else {
        arrival[0] = yk[0];
        arrival[1] = yk[1];
        arrival[2] = tt;
        arrival[3] = sqrt(sq(yk[0] - s.src[0]) + sq(yk[1] - s.src[1])) / tt;
        arrival[5] = -asin(M->c(M->m, yk[0], yk[1], M->topo_z(M->m, yk[0], yk[1])) /
                           M->c(M->m, s.src[0], s.src[1], s.src[2]) * yk[5]) * 180.0 / WV_PI;
        double baz = 90.0 - atan2(-yk[4], -yk[3]) * 180.0 / WV_PI;
        while (baz < -180.0) baz += 360.0;
        while (baz > 180.0) baz -= 360.0;
        arrival[6] = baz;
    }
    arrival[4] = rmax;
    arrival[7] = 20.0 * log10(M->amp(M->m, &s, yk));
    arrival[8] = -2.0 * att;
    *ok = !brk;
    /* wvfrm_out.dat (:1455-1461) */
    const double c = M->c(M->m, yk[0], yk[1], yk[2]);
    const double rho = M->rho(M->m, yk[0], yk[1], yk[2]);
    const double nu = sqrt(pow(yk[3], 2) + pow(yk[4], 2) + pow(yk[5], 2));
    double w[3], g[3];
    M->winds(M->m, yk[0], yk[1], yk[2], w);
    if (M->sph) {
        g[0] = c * yk[3] / nu + w[2];
        g[1] = c * yk[4] / nu + w[1];
        g[2] = c * yk[5] / nu + w[0];
    } else {
        g[0] = c * yk[3] / nu + w[0];
        g[1] = c * yk[4] / nu + w[1];
        g[2] = c * yk[5] / nu + w[2];
    }
    const double cg = sqrt(pow(g[0], 2) + pow(g[1], 2) + pow(g[2], 2));
    const double Dk = M->jac(M->m, yk);
    for (int i = 0; i < len; i++) {
        t_out[i] = tt + t[i];
        p_out[i] = u[i] * p0 * sqrt(fabs((rho * pow(c, 3) * nu) / (rho0 * pow(c0, 3) * nu0) * (cg0 * D0) / (cg * Dk)));
    }
    free(rows);
    free(u);
    return started ? steps : -1;
}

/* ---- the 2D geometry (geoac.eqset.2d.cpp wnl_wvfrm, main.infraga.2d.cpp
 * run_wnl_wvfrm): rows (r, z, ...) of 8 doubles; beta from the effective
 * sound speed c + u cos(phi) + v sin(phi) without the ray-direction
 * factors; the absorption scaled by the effective sound speed of the leg's
 * last evolved row (the set-up loop's last c, which the step loop reuses)
 * at the altitude of s; every prop_rk4 leg, ground or elevated reflection,
 * counts toward bounces. ------------------------------------------------ */
typedef struct {
    go_2d *g;
    lin1 *alt;
    double c_last, c0;
} wv2_absorb;

static double absorb_2d(void *ctx, double s, double freq)
{
    wv2_absorb *q = (wv2_absorb *)ctx;
    return -q->c_last / q->c0 * go_sb_alpha(q->g->a, 0.0, 0.0, lin_f(q->alt, s), freq);
}

double go2_wnl_evolve(go_2d *g, const go_wv_params *W, const double *rows, int k1, int k2, double c0, double rho0,
                      double D0, double p0, double *t, double *u, long long *n_steps)
{
    const int n = k2 - k1;
    if (n_steps) *n_steps = 0;
    if (n < 2) return 0.0;
    double *lx = (double *)malloc(sizeof(double) * (size_t)n * 9);
    lin1 L[4];   /* alt, D, beta, beta_caustic */
    for (int i = 0; i < 4; i++) {
        L[i].n = n; L[i].accel = 0;
        L[i].x = lx;
        L[i].f = lx + n * (1 + 2 * i);
        L[i].sl = lx + n * (2 + 2 * i);
    }
    const double beta0 = 1.2 * p0 / ((rho0 * 1.0e3) * pow(c0 * 1.0e3, 2));
    double s = 0.0, c = 0.0;
    for (int k = 0; k < n; k++) {
        const double *y = rows + 8 * (size_t)(k1 + k);
        lx[k] = s;
        L[0].f[k] = y[1];
        L[1].f[k] = go2_jacobian(g, y);
        c = go2_ceff(g, y[1]);
        const double rho = go2_rho(g, y[1]);
        L[2].f[k] = beta0 / c0 * sqrt(fabs((D0 * rho0 * c0) / (L[1].f[k] * rho * c)));
        if (k < 1) L[3].f[k] = 0.0;
        else L[3].f[k] = beta0 / c0 * sqrt(fabs((D0 * rho0 * c0) / ((L[1].f[k] - L[1].f[k - 1]) / (lx[k] - lx[k - 1]) *
                                                                    rho * c)));
        if ((k + 1) < n) {
            const double *z = rows + 8 * (size_t)(k1 + k + 1);
            s += sqrt(pow(z[0] - y[0], 2) + pow(z[1] - y[1], 2));
        }
    }
    L[3].f[0] = L[3].f[1];
    const double s_tot = s;
    for (int i = 0; i < 4; i++) lin_set(&L[i]);
    wv2_absorb q = {g, &L[0], c, c0};
    const long long steps = wv_evolve_core(W, &L[1], &L[2], &L[3], s_tot, lx[n - 1], 1.0e-7, 1, absorb_2d, &q, t, u);
    free(lx);
    if (n_steps) *n_steps = steps;
    return s_tot;
}

long long go2_wnl_run(go_2d *g, const go_wv_params *W, double src_alt, double incl_deg, double az_deg, int bounces,
                      double freq, double wvfrm_ref, const double *t_src, const double *p_src, double *t_init,
                      double *p_init, double *t_out, double *p_out, double arrival[9], int *ok)
{
    const int len = W->len;
    const int length = (int)(g->s_max / g->ds_min);
    double *rows = (double *)malloc(sizeof(double) * 8 * (size_t)length);
    double *u = (double *)malloc(sizeof(double) * 2 * (size_t)len), *t = u + len;
    for (int i = 0; i < len; i++) { t[i] = t_src[i]; u[i] = p_src[i]; }
    g->calc_amp = 1;
    g->theta = incl_deg * (WV_PI / 180.0);
    g->phi = WV_PI / 2.0 - az_deg * (WV_PI / 180.0);
    const double zg0 = go2_topo_z(g, 0.0);
    const double z_src = src_alt > zg0 ? src_alt : zg0;          /* z_src = max(z_src, topo::z(r_src)) */
    go2_set_initial(g, rows, 0.0, z_src);
    double tt = 0.0, att = 0.0, zmax = 0.0, ray_length = 0.0;
    double c0 = 0, rho0 = 0, D0 = 0, p0 = 0;
    int started = 0, brk = 0, k = 0;
    long long steps = 0;
    for (int bnc = 0; bnc <= bounces; bnc++) {
        k = go2_prop_rk4(g, rows, length, &brk);
        for (int m = 1; m < k; m++) {
            tt += go2_travel_time_seg(g, rows + 8 * (m - 1), rows + 8 * m);
            att += go2_atten_seg(g, rows + 8 * (m - 1), rows + 8 * m, freq);
            zmax = zmax > rows[8 * m + 1] ? zmax : rows[8 * m + 1];
        }
        if (!started) {
            for (int kr = 0; kr < k; kr++) {
                const double *a = rows + 8 * kr, *b = rows + 8 * (kr + 1);
                ray_length += sqrt(pow(b[0] - a[0], 2) + pow(b[1] - a[1], 2));
                if (ray_length >= wvfrm_ref) {
                    c0 = go2_ceff(g, b[1]);
                    rho0 = go2_rho(g, b[1]);
                    D0 = go2_jacobian(g, b);
                    double tref = 0.0;
                    for (int q = 0; q < kr + 1; q++) tref += go2_travel_time_seg(g, rows + 8 * q, rows + 8 * (q + 1));
                    for (int i = 0; i < len; i++) { t_init[i] = tref + t[i]; p_init[i] = u[i]; }
                    for (int i = 0; i < len; i++) p0 = fmax(p0, fabs(u[i]));
                    for (int i = 0; i < len; i++) u[i] /= p0;
                    started = 1;
                    long long ns = 0;
                    wv_dump.s_prev = wvfrm_ref;
                    ray_length = go2_wnl_evolve(g, W, rows, kr + 1, k, c0, rho0, D0, p0, t, u, &ns);
                    steps += ns;
                    break;
                }
            }
        } else {
            long long ns = 0;
            wv_dump.s_prev = ray_length;
            ray_length += go2_wnl_evolve(g, W, rows, 0, k, c0, rho0, D0, p0, t, u, &ns);
            steps += ns;
        }
        if (brk) break;
        go2_set_refl(g, rows + 8 * (k - 2), rows + 8 * (k - 1), rows);
    }
    const double *yk = rows + 8 * (size_t)k;
    memset(arrival, 0, sizeof(double) * 9);
    arrival[0] = yk[0];
    arrival[2] = tt;
    arrival[3] = yk[0] / tt;
    arrival[4] = zmax;
    arrival[7] = 20.0 * log10(go2_amp(g, yk));
    arrival[8] = -2.0 * att;
    *ok = !brk;
    const double rho = go2_rho(g, yk[1]), ce = go2_ceff(g, yk[1]);
    for (int i = 0; i < len; i++) {
        t_out[i] = tt + t[i];
        p_out[i] = u[i] * p0 * sqrt(fabs((rho * ce) / (rho0 * c0) * D0 / go2_jacobian(g, yk)));
    }
    free(rows);
    free(u);
    return started ? steps : -1;
}

/* ctypes entry: go2_wnl_run over a 2D handle (go2_h_new). */
long long go2_h_wnl_run(void *h, const go_wv_params *W, double src_alt, double incl_deg, double az_deg, int bounces,
                        double freq, double wvfrm_ref, const double *t_src, const double *p_src, double *t_init,
                        double *p_init, double *t_out, double *p_out, double arrival[9], int *ok)
{
    return go2_wnl_run((go_2d *)h, W, src_alt, incl_deg, az_deg, bounces, freq, wvfrm_ref, t_src, p_src, t_init,
                       p_init, t_out, p_out, arrival, ok);
}

/* ctypes entry: go_wnl_run over a range-dependent handle. */
long long go_h_rd_wnl_run(void *h, const go_params *p, const go_wv_params *W, const double src_in[3], double incl_deg,
                          double az_deg, int bounces, double freq, double wvfrm_ref, const double *t_src,
                          const double *p_src, double *t_init, double *p_init, double *t_out, double *p_out,
                          double arrival[9], int *ok)
{
    go_wv_med M = go_wv_rd_med(h);
    return go_wnl_run(&M, p, W, src_in, incl_deg, az_deg, bounces, freq, wvfrm_ref, t_src, p_src, t_init, p_init,
                      t_out, p_out, arrival, ok);
}

/* ctypes entry: go_wnl_run over a stratified atmosphere handle. */
long long go_h_wnl_run(void *h, const go_params *p, const go_wv_params *W, const double src_in[3], double incl_deg,
                       double az_deg, int bounces, double freq, double wvfrm_ref, const double *t_src,
                       const double *p_src, double *t_init, double *p_init, double *t_out, double *p_out,
                       double arrival[9], int *ok)
{
    go_wv_med M = go_wv_strat_med((go_atmo *)h);
    return go_wnl_run(&M, p, W, src_in, incl_deg, az_deg, bounces, freq, wvfrm_ref, t_src, p_src, t_init, p_init,
                      t_out, p_out, arrival, ok);
}
