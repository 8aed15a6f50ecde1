/*
 * ga_oracle.c -- CPU restatement of infraGA's stratified propagation path.
 *
 * TEST INFRASTRUCTURE ONLY (see ga_oracle.h).  Each function cites the
 * reference file:line whose algorithm it restates.  Operation order follows
 * the reference so that, compiled with -O2 -ffp-contract=off and glibc libm,
 * the oracle reproduces the reference's binary64 rounding.
 *
 * Scope: flat ground (no terrain file), topo_use_BLw=false (vertical wind
 * w == 0 everywhere), 3D Cartesian and spherical stratified geometries.
 */
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ga_oracle.h"

#define GO_PI 3.141592653589793238462643   /* geoac.params.sph.cpp:9 */

static double sq(double x) { return pow(x, 2); }

/* ======================================================================
 * Splines  (src/util/interpolation.cpp)
 * ==================================================================== */

/* interpolation.cpp:14-49 -- cursor search: current segment, then the one
 * above, then the one below, then a two-ended linear scan. */
int go_find_segment(double x, const double *xv, int length, int *prev)
{
    int done = 0;
    if (x > xv[length - 1] || x < xv[0]) {
        return *prev;   /* reference prints a warning and keeps the cursor */
    }
    if (x >= xv[*prev] && x <= xv[*prev + 1]) done = 1;
    if (!done && *prev + 2 <= length - 1) {
        if (x >= xv[*prev + 1] && x <= xv[*prev + 2]) { (*prev)++; done = 1; }
    }
    if (!done && *prev - 1 >= 0) {
        if (x >= xv[*prev - 1] && x <= xv[*prev]) { (*prev)--; done = 1; }
    }
    if (!done) {
        for (int n = 0; n < length; n++) {
            if (x >= xv[n] && x <= xv[n + 1]) { *prev = n; break; }
            else if (x >= xv[(length - 1) - (n + 1)] && x < xv[(length - 1) - n]) {
                *prev = (length - 1) - (n + 1); break;
            }
        }
    }
    return *prev;
}

/* interpolation.cpp:52-54 */
static double in_interval(double x, double lo, double hi)
{
    double t = x > lo ? x : lo;
    return t < hi ? t : hi;
}

/* interpolation.cpp:103-136 -- natural end conditions, Thomas solve of the
 * slope system. */
void go_spline_set(go_spline *s)
{
    int n = s->n;
    const double *x = s->x;
    double *f = s->f;
    double *nc = (double *)malloc(sizeof(double) * (size_t)n);
    double *nd = (double *)malloc(sizeof(double) * (size_t)n);
    double ai, bi, ci, di;

    bi = 2.0 / (x[1] - x[0]);
    ci = 1.0 / (x[1] - x[0]);
    di = 3.0 * (f[1] - f[0]) / sq(x[1] - x[0]);
    nc[0] = ci / bi;
    nd[0] = di / bi;
    for (int i = 1; i < n - 1; i++) {
        double dx1 = x[i] - x[i - 1], dx2 = x[i + 1] - x[i];
        ai = 1.0 / dx1; bi = 2.0 * (1.0 / dx1 + 1.0 / dx2); ci = 1.0 / dx2;
        di = 3.0 * ((f[i] - f[i - 1]) / sq(dx1) + (f[i + 1] - f[i]) / sq(dx2));
        nc[i] = ci / (bi - nc[i - 1] * ai);
        nd[i] = (di - nd[i - 1] * ai) / (bi - nc[i - 1] * ai);
    }
    ai = 1.0 / (x[n - 1] - x[n - 2]);
    bi = 2.0 / (x[n - 1] - x[n - 2]);
    di = 3.0 * (f[n - 1] - f[n - 2]) / sq(x[n - 1] - x[n - 2]);
    nd[n - 1] = (di - nd[n - 2] * ai) / (bi - nc[n - 2] * ai);
    s->slope[n - 1] = nd[n - 1];
    for (int i = n - 2; i > -1; i--) s->slope[i] = nd[i] - nc[i] * s->slope[i + 1];
    free(nc);
    free(nd);
}

/* Hermite (A,B) form shared by interpolation.cpp:144-260. */
#define GO_SEG(s, xq)                                                      \
    int k = go_find_segment((xq), (s)->x, (s)->n, &(s)->accel);            \
    double dx = (s)->x[k + 1] - (s)->x[k];                                 \
    double X = ((xq) - (s)->x[k]) / dx;                                    \
    double df = (s)->f[k + 1] - (s)->f[k];                                 \
    double A = (s)->slope[k] * dx - df;                                    \
    double B = -(s)->slope[k + 1] * dx + df

double go_spline_f(go_spline *s, double x)     /* interpolation.cpp:144-158 */
{
    GO_SEG(s, x);
    return s->f[k] + X * (df + (1.0 - X) * (A * (1.0 - X) + B * X));
}

double go_spline_df(go_spline *s, double x)    /* interpolation.cpp:160-174 */
{
    GO_SEG(s, x);
    return df / dx + (1.0 - 2.0 * X) * (A * (1.0 - X) + B * X) / dx + X * (1.0 - X) * (B - A) / dx;
}

double go_spline_ddf(go_spline *s, double x)   /* interpolation.cpp:176-190 */
{
    GO_SEG(s, x);
    (void)df;
    return 2.0 * (B - 2.0 * A + (A - B) * 3.0 * X) / sq(dx);
}

/* interpolation.cpp:225-241 (f, f', f''); ddf may be NULL (208-223). */
void go_spline_all(go_spline *s, double x, double *f, double *dfdx, double *ddf)
{
    GO_SEG(s, x);
    *f = s->f[k] + X * (df + (1.0 - X) * (A * (1.0 - X) + B * X));
    *dfdx = df / dx + (1.0 - 2.0 * X) * (A * (1.0 - X) + B * X) / dx + X * (1.0 - X) * (B - A) / dx;
    if (ddf) *ddf = 2.0 * (B - 2.0 * A + (A - B) * 3.0 * X) / sq(dx);
}

/* ======================================================================
 * Medium set-up  (atmo_io.{sph,3d}.strat.cpp:27-39 set_limits and
 * atmo_state.*.cpp set_bndlyr)
 * ==================================================================== */
int go_atmo_init(go_atmo *a, int geom, const double *z, const double *c,
                 const double *u, const double *v, const double *rho, int n)
{
    memset(a, 0, sizeof(*a));
    if (n < 3) return -1;
    a->geom = geom;
    a->n = n;
    a->z = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(a->z, z, sizeof(double) * (size_t)n);
    go_spline *sp[4] = {&a->c, &a->u, &a->v, &a->rho};
    const double *src[4] = {c, u, v, rho};
    for (int i = 0; i < 4; i++) {
        sp[i]->n = n;
        sp[i]->x = a->z;
        sp[i]->accel = 0;
        sp[i]->f = (double *)malloc(sizeof(double) * (size_t)n);
        sp[i]->slope = (double *)malloc(sizeof(double) * (size_t)n);
        memcpy(sp[i]->f, src[i], sizeof(double) * (size_t)n);
        go_spline_set(sp[i]);
    }
    a->r0 = (geom == GA_GEOM_SPH) ? 6370.0 : 0.0;   /* globe.cpp:10 */
    a->z0 = z[0];                                    /* set_limits: z0 = bottom of profile */
    a->z_max = a->z0;                                /* set_bndlyr, no terrain */
    a->z_bndlyr = a->z_max + 2.0;
    a->bndlyr_param = 0.1;
    a->tweak_abs = 1.0;                              /* main.infraga.sph.cpp:219 */
    return 0;
}

void go_atmo_free(go_atmo *a)
{
    go_spline *sp[4] = {&a->c, &a->u, &a->v, &a->rho};
    for (int i = 0; i < 4; i++) { free(sp[i]->f); free(sp[i]->slope); }
    free(a->z);
    memset(a, 0, sizeof(*a));
}

/* geoac.params.{sph,3d}.cpp defaults + set_limits (atmo_io.*.strat.cpp). */
void go_params_default(go_params *p, int geom, const go_atmo *a, int calc_amp)
{
    memset(p, 0, sizeof(*p));
    p->calc_amp = calc_amp;
    p->eq_cnt = calc_amp ? 18 : 6;                   /* geoac.interface.sph.cpp:22-25 */
    p->s_max = 1000.0; p->ds_min = 0.001; p->ds_max = 0.05;
    p->alt_max = a->z[a->n - 1];
    p->time_max = 10.0;
    if (geom == GA_GEOM_SPH) {
        p->rng_max = 1000.0;
        p->lat_min = -GO_PI / 2.0; p->lat_max = GO_PI / 2.0;
        p->lon_min = -GO_PI; p->lon_max = GO_PI;
    } else {
        p->rng_max = 2000.0;                         /* atmo_io.3d.strat.cpp:30 */
        p->x_min = -p->rng_max; p->x_max = p->rng_max;
        p->y_min = -p->rng_max; p->y_max = p->rng_max;
    }
}

/* ======================================================================
 * Medium evaluation  (atmo_state.sph.strat.cpp:100-249, 3d.strat.cpp:81-240)
 * Coordinates: sph (r, lat, lon); 3D (x, y, z).
 * ==================================================================== */
static double zc(const go_atmo *a, double p0, double p2)
{
    return a->geom == GA_GEOM_SPH ? p0 - a->r0 : p2;
}
static double ground(const go_atmo *a)           /* topo::z without terrain */
{
    return a->geom == GA_GEOM_SPH ? a->r0 + a->z0 : a->z0;
}
static double zev(const go_atmo *a, double p0, double p2)
{
    return in_interval(zc(a, p0, p2), a->z[0], a->z[a->n - 1]);
}
static int vert(const go_atmo *a) { return a->geom == GA_GEOM_SPH ? 0 : 2; }

/* topo::bndlyr_sc / bndlyr_dsc (sph :100-123, 3d :81-103); flat ground so
 * the horizontal derivative terms carry dz = 0. */
static double bl_sc(const go_atmo *a, double p0, double p2)
{
    double X, result = 1.0;
    double h = a->geom == GA_GEOM_SPH ? p0 : p2;
    if (zc(a, p0, p2) < a->z_bndlyr) {
        X = exp(-(h - ground(a)) / a->bndlyr_param);
        result = (1.0 - X) / (1.0 + X);
    }
    return result;
}
static double bl_dsc(const go_atmo *a, double p0, double p2, int n)
{
    double X, dX, result = 0.0;
    double h = a->geom == GA_GEOM_SPH ? p0 : p2;
    if (zc(a, p0, p2) < a->z_bndlyr) {
        X = exp(-(h - ground(a)) / a->bndlyr_param);
        if (n != vert(a)) dX = X / a->bndlyr_param * 0.0;
        else dX = -X / a->bndlyr_param;
        result = -2.0 * dX / sq(1.0 + X);
    }
    return result;
}

double go_c(go_atmo *a, double p0, double p1, double p2)   /* sph :172-175 */
{
    (void)p1;
    return go_spline_f(&a->c, zev(a, p0, p2));
}
double go_rho(go_atmo *a, double p0, double p1, double p2) /* sph :166-169 */
{
    (void)p1;
    return go_spline_f(&a->rho, zev(a, p0, p2));
}
double go_dc(go_atmo *a, double p0, double p1, double p2, int n) /* sph :176-183 */
{
    (void)p1;
    if (n == vert(a)) return go_spline_df(&a->c, zev(a, p0, p2));
    return 0.0;
}
double go_u(go_atmo *a, double p0, double p1, double p2)   /* sph :193-196 */
{
    (void)p1;
    return go_spline_f(&a->u, zev(a, p0, p2)) * bl_sc(a, p0, p2);
}
double go_v(go_atmo *a, double p0, double p1, double p2)   /* sph :216-219 */
{
    (void)p1;
    return go_spline_f(&a->v, zev(a, p0, p2)) * bl_sc(a, p0, p2);
}
static double dwind(go_atmo *a, go_spline *s, double p0, double p2, int n)
{
    double z = zev(a, p0, p2);
    double result = go_spline_f(s, z) * bl_dsc(a, p0, p2, n);
    if (n == vert(a)) result += go_spline_df(s, z) * bl_sc(a, p0, p2);
    return result;
}
double go_du(go_atmo *a, double p0, double p1, double p2, int n) /* sph :197-205 */
{
    (void)p1;
    return dwind(a, &a->u, p0, p2, n);
}
double go_dv(go_atmo *a, double p0, double p1, double p2, int n) /* sph :220-228 */
{
    (void)p1;
    return dwind(a, &a->v, p0, p2, n);
}

/* atmo::calc_uvw, both overloads: sph :498-610 / :613-803, 3D :462-555 /
 * :556-739.  Flat ground: zg = topo::z, all zg derivatives are zero, so
 * every horizontal entry of dX, ddX (and the scaled winds) vanishes. */
void go_calc_uvw(go_atmo *a, double p0, double p1, double p2,
                 double *u, double *v, double *w, double du[3], double dv[3], double dw[3],
                 double ddu[6], double ddv[6], double ddw[6])
{
    (void)p1;
    int iv = vert(a);
    double bp = a->bndlyr_param;
    double z_eval = zev(a, p0, p2);
    double h = a->geom == GA_GEOM_SPH ? p0 : p2;
    double zg = ground(a);
    double u0, v0, du0v, dv0v, ddu0v = 0.0, ddv0v = 0.0;
    double X, dX[3], ddX[3][3], sc0, dsc0[3], ddsc0[3][3];

    if (ddu) {
        go_spline_all(&a->u, z_eval, &u0, &du0v, &ddu0v);
        go_spline_all(&a->v, z_eval, &v0, &dv0v, &ddv0v);
    } else {
        /* first-derivative overload evaluates the splines twice (sph :527-550) */
        go_spline_all(&a->u, z_eval, &u0, &du0v, NULL);
        go_spline_all(&a->v, z_eval, &v0, &dv0v, NULL);
        if (a->geom == GA_GEOM_SPH) {
            u0 = go_spline_f(&a->u, z_eval);
            v0 = go_spline_f(&a->v, z_eval);
            du0v = go_spline_df(&a->u, z_eval);
            dv0v = go_spline_df(&a->v, z_eval);
        }
    }
    double du0[3] = {0.0, 0.0, 0.0}, dv0[3] = {0.0, 0.0, 0.0};
    du0[iv] = du0v;
    dv0[iv] = dv0v;

    X = exp(-(h - zg) / bp);
    for (int n = 0; n < 3; n++) {
        dX[n] = 0.0;
        for (int m = 0; m < 3; m++) ddX[n][m] = 0.0;
    }
    dX[iv] = -X / bp;
    ddX[iv][iv] = X / sq(bp);

    sc0 = (1.0 - X) / (1.0 + X);
    for (int n = 0; n < 3; n++) {
        dsc0[n] = -2.0 * dX[n] / sq(1.0 + X);
        for (int m = 0; m < 3; m++)
            ddsc0[n][m] = -2.0 * (ddX[n][m] / sq(1.0 + X) - 2.0 * dX[n] * dX[m] / pow(1.0 + X, 3));
    }

    *u = u0 * sc0;
    *v = v0 * sc0;
    *w = 0.0;
    for (int n = 0; n < 3; n++) {
        du[n] = du0[n] * sc0 + u0 * dsc0[n];
        dv[n] = dv0[n] * sc0 + v0 * dsc0[n];
        dw[n] = 0.0;
    }
    if (ddu) {
        double ddu0[3][3] = {{0}}, ddv0[3][3] = {{0}};
        ddu0[iv][iv] = ddu0v;
        ddv0[iv][iv] = ddv0v;
        for (int n = 0; n < 3; n++) {
            ddu[n] = ddu0[n][n] * sc0 + 2.0 * du0[n] * dsc0[n] + u0 * ddsc0[n][n];
            ddv[n] = ddv0[n][n] * sc0 + 2.0 * dv0[n] * dsc0[n] + v0 * ddsc0[n][n];
        }
        ddu[3] = ddu0[0][1] * sc0 + du0[0] * dsc0[1] + du0[1] * dsc0[0] + u0 * ddsc0[0][1];
        ddu[4] = ddu0[0][2] * sc0 + du0[0] * dsc0[2] + du0[2] * dsc0[0] + u0 * ddsc0[0][2];
        ddu[5] = ddu0[1][2] * sc0 + du0[1] * dsc0[2] + du0[2] * dsc0[1] + u0 * ddsc0[1][2];
        ddv[3] = ddv0[0][1] * sc0 + dv0[0] * dsc0[1] + dv0[1] * dsc0[0] + v0 * ddsc0[0][1];
        ddv[4] = ddv0[0][2] * sc0 + dv0[0] * dsc0[2] + dv0[2] * dsc0[0] + v0 * ddsc0[0][2];
        ddv[5] = ddv0[1][2] * sc0 + dv0[1] * dsc0[2] + dv0[2] * dsc0[1] + v0 * ddsc0[1][2];
        for (int n = 0; n < 6; n++) ddw[n] = 0.0;
    }
}

/* globe.cpp:23-28 haversine great-circle distance. */
double go_gc_dist(double r0, double lat1, double lon1, double lat2, double lon2)
{
    double term1 = sq(sin((lat2 - lat1) / 2.0));
    double term2 = cos(lat1) * cos(lat2) * sq(sin((lon2 - lon1) / 2.0));
    return 2.0 * r0 * asin(sqrt(term1 + term2));
}

/* Sutherland & Bass (2004) absorption in dB/km.  Spherical variant:
 * atmo_state.abs.sph.cpp:14-151; Cartesian variant (different rotational
 * and classical terms, pow(z,2) in the N profile): atmo_state.abs.cart.cpp:12-141. */
double go_sb_alpha_at(double c, double rho, double z, int sph, double tweak, double freq)
{
    const double gam = 1.4, R = 287.058;         /* atmo_state.*.strat.cpp:23-27 */
    double mu_o = 18.192e-6, T_o = 293.15, P_o = 101325, S = 117.0;
    double Cv_R[4] = {5.0 / 2.0, 5.0 / 2.0, 3.0, 3.0};
    double Cp_R[4] = {7.0 / 2.0, 7.0 / 2.0, 4.0, 4.0};
    double Theta[4] = {2239.1, 3352.0, 915.0, 1037.0};
    double X[7], X_ON, Z_rot[2], Z_rot_, sigma, nn, chi, cchi, mu, nu;
    double beta_0, beta_1, beta_2, alpha_1, alpha_2, a_cl, a_rot, a_diff = 0.0, a_vib;
    double T_z, P_z, c_snd_z, A1, A2, B, C, D, E, F, G, H, I, J, K, L, ZZ, hu;
    double f_vib[4], C_R, A_max, Tr;

    T_z = sq(c * 1.0e3) / (R * gam);
    P_z = rho * 1.0e3 * sq(c * 1.0e3) / gam;
    c_snd_z = c * 1.0e3;

    mu = mu_o * sqrt(T_z / T_o) * ((1.0 + S / T_o) / (1.0 + S / T_z));
    nu = (8.0 * GO_PI * freq * mu) / (3.0 * P_z);

    if (z > 90.) X[0] = pow(10.0, 49.296 - (1.5524 * z) + (1.8714E-2 * sq(z)) - (1.1069E-4 * pow(z, 3))
                                + (3.199E-7 * pow(z, 4)) - (3.6211E-10 * pow(z, 5)));
    else X[0] = pow(10.0, -0.67887);
    if (z > 76.) X[1] = pow(10.0, (1.3972E-1) - (5.6269E-3 * z) + (3.9407E-5 * sq(z)) - (1.0737E-7 * pow(z, 3)));
    else X[1] = pow(10.0, -0.10744);
    X[2] = pow(10.0, -3.3979);
    if (z > 80.) X[3] = pow(10.0, -4.234 - (3.0975E-2 * z));
    else X[3] = pow(10.0, -19.027 + (1.3093 * z) - (4.6496E-2 * sq(z)) + (7.8543E-4 * pow(z, 3))
                         - (6.5169E-6 * pow(z, 4)) + (2.1343E-8 * pow(z, 5)));
    if (z > 95.) X[4] = pow(10.0, -3.2456 + (4.6642E-2 * z) - (2.6894E-4 * sq(z)) + (5.264E-7 * pow(z, 3)));
    else X[4] = pow(10.0, -11.195 + (1.5408E-1 * z) - (1.4348E-3 * sq(z)) + (1.0166E-5 * pow(z, 3)));
    if (sph)
        X[5] = pow(10.0, -53.746 + (1.5439 * z) - (1.8824E-2 * sq(z)) + (1.1587E-4 * pow(z, 3))
                         - (3.5399E-7 * pow(z, 4)) + (4.2609E-10 * pow(z, 5)));
    else    /* abs.cart.cpp:78 uses pow(z, 2) in the fourth-order term */
        X[5] = pow(10.0, -53.746 + (1.5439 * z) - (1.8824e-2 * sq(z)) + (1.1587e-4 * pow(z, 3))
                         - (3.5399e-7 * sq(z)) + (4.2609e-10 * pow(z, 5)));
    if (z > 30.) X[6] = pow(10.0, -4.2563 + (7.6245E-2 * z) - (2.1824E-3 * sq(z)) - (2.3010E-6 * pow(z, 3))
                               + (2.4265E-7 * pow(z, 4)) - (1.2500E-09 * pow(z, 5)));
    else {
        if (z > 100.) X[6] = pow(10.0, -0.62534 - (8.3665E-2 * z));
        else X[6] = pow(10.0, -1.7491 + (4.4986E-2 * z) - (6.8549E-2 * sq(z)) + (5.4639E-3 * pow(z, 3))
                           - (1.5539E-4 * pow(z, 4)) + (1.5063E-06 * pow(z, 5)));
    }
    X_ON = (X[0] + X[1]) / 0.9903;

    Z_rot[0] = 54.1 * exp(-17.3 * (pow(T_z, -1.0 / 3.0)));
    Z_rot[1] = 63.3 * exp(-16.7 * (pow(T_z, -1.0 / 3.0)));
    Z_rot_ = 1.0 / ((X[1] / Z_rot[1]) + (X[0] / Z_rot[0]));

    sigma = 5.0 / sqrt(21.0);
    nn = (4.0 / 5.0) * sqrt(3.0 / 7.0) * Z_rot_;
    chi = 3.0 * nn * nu / 4.0;
    cchi = 2.36 * chi;

    beta_0 = 2.0 * GO_PI * freq / c_snd_z;
    if (sph) {
        beta_1 = beta_0 * sqrt(0.5 * (sqrt(1.0 + sq(nu)) + 1.0) / (1.0 + sq(nu)));
        beta_2 = beta_0 * sqrt((1.0 + sq(chi)) / (1 + sq((sigma * chi))));
        alpha_1 = beta_0 * sqrt(0.5 * (sqrt(1.0 + sq(nu)) - 1.0) / (1.0 + sq(nu)));
        alpha_2 = beta_0 * (((sigma / 2.0 - 1.0 / (2.0 * sigma)) * chi) / (sqrt((1.0 + sq(chi)) * (1.0 + sq(sigma * chi)))));
        (void)beta_1; (void)beta_2; (void)alpha_1; (void)alpha_2;
        a_cl = (2.0 * GO_PI * freq / c_snd_z) * sqrt(0.5 * (sqrt(1.0 + sq(nu)) - 1.0) * (1.0 + sq(cchi)) / ((1.0 + sq(nu)) * (1.0 + sq(sigma * cchi))));
        a_rot = (2.0 * GO_PI * freq / c_snd_z) * X_ON * ((sq(sigma) - 1.0) * chi / (2 * sigma)) * sqrt(0.5 * (sqrt(1.0 + sq(nu)) + 1.0) / ((1.0 + sq(nu)) * (1.0 + sq(cchi))));
        a_diff = 0.003 * a_cl;
    } else {
        beta_1 = sqrt(0.5 * (sqrt(1.0 + sq(nu)) + 1.0) / (1.0 + sq(nu)));
        beta_2 = sqrt((1.0 + sq(cchi)) / (1.0 + sq((sigma * cchi))));
        alpha_1 = beta_0 * sqrt(0.5 * (sqrt(1.0 + sq(nu)) - 1.0) / (1.0 + sq(nu)));
        alpha_2 = beta_0 * (((sigma / 2.0 - 1.0 / (2.0 * sigma)) * chi) / (sqrt((1.0 + sq(cchi)) * (1.0 + sq(sigma * cchi)))));
        a_cl = alpha_1 * beta_2;
        a_rot = alpha_2 * beta_1 * X_ON;
    }

    Tr = pow(T_z / T_o, -1.0 / 3.0) - 1.0;
    A1 = (X[0] + X[1]) * 24.0 * exp(-9.16 * Tr);
    A2 = (X[4] + X[5]) * 2400.0;
    B = 40400.0 * exp(10.0 * Tr);
    C = 0.02 * exp(-11.2 * Tr);
    D = 0.391 * exp(8.41 * Tr);
    E = 9.0 * exp(-19.9 * Tr);
    F = 60000.0;
    G = 28000.0 * exp(-4.17 * Tr);
    H = 22000.0 * exp(-7.68 * Tr);
    I = 15100.0 * exp(-10.4 * Tr);
    J = 11500.0 * exp(-9.17 * Tr);
    K = (8.48E08) * exp(9.17 * Tr);
    L = exp(-7.72 * Tr);
    ZZ = H * X[2] + I * (X[0] + 0.5 * X[4]) + J * (X[1] + 0.5 * X[5]) + K * (X[6] + X[3]);
    hu = 100.0 * (X[3] + X[6]);
    f_vib[0] = (P_z / P_o) * (mu_o / mu) * (A1 + A2 + B * hu * (C + hu) * (D + hu));
    f_vib[1] = (P_z / P_o) * (mu_o / mu) * (E + F * X[3] + G * X[6]);
    f_vib[2] = (P_z / P_o) * (mu_o / mu) * ZZ;
    f_vib[3] = (P_z / P_o) * (mu_o / mu) * (1.2E5) * L;

    a_vib = 0.0;
    for (int m = 0; m < 4; m++) {
        C_R = ((sq(Theta[m] / T_z)) * exp(-Theta[m] / T_z)) / (sq(1.0 - exp(-Theta[m] / T_z)));
        A_max = (X[m] * (GO_PI / 2) * C_R) / (Cp_R[m] * (Cv_R[m] + C_R));
        a_vib += (A_max / c_snd_z) * ((2 * (sq(freq)) / f_vib[m]) / (1.0 + sq(freq / f_vib[m])));
    }
    if (sph) return (a_cl + a_rot + a_diff + a_vib) * tweak * 8.685889e3;
    return (a_cl * 1.003 + a_rot + a_vib) * tweak * 8.685889e3;
}

/* SB_alpha at a point of the stratified medium: c and rho from the splines,
 * z = r - r0 (sph) or z (3D). */
double go_sb_alpha(go_atmo *a, double p0, double p1, double p2, double freq)
{
    const int sph = (a->geom == GA_GEOM_SPH);
    double z = sph ? (p0 - a->r0) : p2;
    return go_sb_alpha_at(go_c(a, p0, p1, p2), go_rho(a, p0, p1, p2), z, sph, a->tweak_abs, freq);
}

/* ======================================================================
 * Ray equations
 * ==================================================================== */

/* set_initial: sph geoac.eqset.sph.strat.cpp:59-104; 3D 3d.strat.cpp:58-105.
 * s->src, s->theta, s->phi must be set; fills c0, nu0 and y0. */
void go_set_initial(go_atmo *a, const go_params *p, go_src *s, double *y0)
{
    double mach[3], nu0[3], mu_t[3], mu_p[3];
    double th = s->theta, ph = s->phi;
    double *L = s->src;
    s->c0 = go_c(a, L[0], L[1], L[2]);
    if (a->geom == GA_GEOM_SPH) {
        mach[0] = 0.0 / s->c0;                       /* atmo::w == 0 */
        mach[1] = go_v(a, L[0], L[1], L[2]) / s->c0;
        mach[2] = go_u(a, L[0], L[1], L[2]) / s->c0;
        nu0[0] = sin(th); nu0[1] = cos(th) * sin(ph); nu0[2] = cos(th) * cos(ph);
        mu_t[0] = cos(th); mu_t[1] = -sin(th) * sin(ph); mu_t[2] = -sin(th) * cos(ph);
        mu_p[0] = 0.0; mu_p[1] = cos(th) * cos(ph); mu_p[2] = -cos(th) * sin(ph);
    } else {
        mach[0] = go_u(a, L[0], L[1], L[2]) / s->c0;
        mach[1] = go_v(a, L[0], L[1], L[2]) / s->c0;
        mach[2] = 0.0 / s->c0;
        nu0[0] = cos(th) * cos(ph); nu0[1] = cos(th) * sin(ph); nu0[2] = sin(th);
        mu_t[0] = -sin(th) * cos(ph); mu_t[1] = -sin(th) * sin(ph); mu_t[2] = cos(th);
        mu_p[0] = -cos(th) * sin(ph); mu_p[1] = cos(th) * cos(ph); mu_p[2] = 0.0;
    }
    double msc = 1.0 + (nu0[0] * mach[0] + nu0[1] * mach[1] + nu0[2] * mach[2]);
    s->nu0 = 1.0 / msc;
    for (int i = 0; i < 3; i++) y0[i] = L[i];
    for (int i = 0; i < 3; i++) y0[3 + i] = nu0[i] / msc;
    if (p->eq_cnt == 18) {
        for (int i = 0; i < 3; i++) { y0[6 + i] = 0.0; y0[12 + i] = 0.0; }
        double mt = mu_t[0] * mach[0] + mu_t[1] * mach[1] + mu_t[2] * mach[2];
        double mp = mu_p[0] * mach[0] + mu_p[1] * mach[1] + mu_p[2] * mach[2];
        for (int i = 0; i < 3; i++) {
            y0[9 + i] = mu_t[i] / msc - nu0[i] / sq(msc) * mt;
            y0[15 + i] = mu_p[i] / msc - nu0[i] / sq(msc) * mp;
        }
    }
}

/* Per-evaluation cache (src_refs, sph :28-42 / 3D :32-43). */
typedef struct {
    double c, dc[5], ddc[3][2];
    double u, du[5], ddu[3][2];
    double v, dv[5], ddv[3][2];
    double w, dw[5], ddw[3][2];
    double nu_mag, dnu_mag[2];
    double cg[3], cg_mag, dcg[3][2], dcg_mag[2];
    double G[3], dG[3][2], T[3], dT[3][2];
} go_refs;

/* update_refs sph: geoac.eqset.sph.strat.cpp:227-371. */
static void refs_sph(go_atmo *a, const go_params *p, const double *y, go_refs *R)
{
    double r = y[0], t = y[1], pp = y[2];
    const double *nu = y + 3;
    double dd_u_t[6], dd_v_t[6], dd_w_t[6], dd_c[3][3], dd_u[3][3], dd_v[3][3], dd_w[3][3];
    double dc_t, ddc_t;
    memset(R, 0, sizeof(*R));
    R->nu_mag = sqrt(sq(nu[0]) + sq(nu[1]) + sq(nu[2]));
    double z_eval = in_interval(r - a->r0, a->z[0], a->z[a->n - 1]);
    if (!p->calc_amp) {
        go_spline_all(&a->c, z_eval, &R->c, &dc_t, NULL);
        R->dc[0] = dc_t;
        a->u.accel = a->c.accel;
        a->v.accel = a->c.accel;
        go_calc_uvw(a, r, t, pp, &R->u, &R->v, &R->w, R->du, R->dv, R->dw, NULL, NULL, NULL);
    } else {
        go_spline_all(&a->c, z_eval, &R->c, &dc_t, &ddc_t);
        memset(dd_c, 0, sizeof(dd_c));
        R->dc[0] = dc_t;
        dd_c[0][0] = ddc_t;
        a->u.accel = a->c.accel;
        a->v.accel = a->c.accel;
        go_calc_uvw(a, r, t, pp, &R->u, &R->v, &R->w, R->du, R->dv, R->dw, dd_u_t, dd_v_t, dd_w_t);
    }
    R->cg[0] = R->c * nu[0] / R->nu_mag + R->w;
    R->cg[1] = R->c * nu[1] / R->nu_mag + R->v;
    R->cg[2] = R->c * nu[2] / R->nu_mag + R->u;
    R->cg_mag = sqrt(sq(R->cg[0]) + sq(R->cg[1]) + sq(R->cg[2]));

    R->G[0] = 1.0;                R->T[0] = -1.0 / r * (nu[1] * R->cg[1] + nu[2] * R->cg[2]);
    R->G[1] = 1.0 / r;            R->T[1] = (nu[0] * R->v - nu[1] * R->w) + nu[0] * R->cg[1] + nu[2] * R->cg[2] * tan(t);
    R->G[2] = 1.0 / (r * cos(t)); R->T[2] = (nu[0] * R->u - nu[2] * R->w) * cos(t) + (nu[1] * R->u - nu[2] * R->v) * sin(t)
                                            + R->cg[2] * (nu[0] * cos(t) - nu[1] * sin(t));
    if (!p->calc_amp) return;

    for (int n = 0; n < 3; n++) { dd_u[n][n] = dd_u_t[n]; dd_v[n][n] = dd_v_t[n]; dd_w[n][n] = dd_w_t[n]; }
    dd_w[0][1] = dd_w[1][0] = dd_w_t[3]; dd_w[0][2] = dd_w[2][0] = dd_w_t[4]; dd_w[1][2] = dd_w[2][1] = dd_w_t[5];
    dd_v[0][1] = dd_v[1][0] = dd_v_t[3]; dd_v[0][2] = dd_v[2][0] = dd_v_t[4]; dd_v[1][2] = dd_v[2][1] = dd_v_t[5];
    dd_u[0][1] = dd_u[1][0] = dd_u_t[3]; dd_u[0][2] = dd_u[2][0] = dd_u_t[4]; dd_u[1][2] = dd_u[2][1] = dd_u_t[5];

    const double *Rt = y + 6, *mt = y + 9, *Rp = y + 12, *mp = y + 15;
    for (int n = 0; n < 3; n++) {
        R->dc[3] += Rt[n] * R->dc[n];  R->dc[4] += Rp[n] * R->dc[n];
        R->dw[3] += Rt[n] * R->dw[n];  R->dw[4] += Rp[n] * R->dw[n];
        R->dv[3] += Rt[n] * R->dv[n];  R->dv[4] += Rp[n] * R->dv[n];
        R->du[3] += Rt[n] * R->du[n];  R->du[4] += Rp[n] * R->du[n];
        for (int m = 0; m < 3; m++) {
            R->ddc[m][0] += Rt[n] * dd_c[m][n]; R->ddc[m][1] += Rp[n] * dd_c[m][n];
            R->ddw[m][0] += Rt[n] * dd_w[m][n]; R->ddw[m][1] += Rp[n] * dd_w[m][n];
            R->ddv[m][0] += Rt[n] * dd_v[m][n]; R->ddv[m][1] += Rp[n] * dd_v[m][n];
            R->ddu[m][0] += Rt[n] * dd_u[m][n]; R->ddu[m][1] += Rp[n] * dd_u[m][n];
        }
    }
    R->dnu_mag[0] = (nu[0] * mt[0] + nu[1] * mt[1] + nu[2] * mt[2]) / R->nu_mag;
    R->dnu_mag[1] = (nu[0] * mp[0] + nu[1] * mp[1] + nu[2] * mp[2]) / R->nu_mag;

    const double *wind_d[3] = {R->dw, R->dv, R->du};
    for (int i = 0; i < 3; i++) {
        R->dcg[i][0] = nu[i] / R->nu_mag * R->dc[3] + R->c * mt[i] / R->nu_mag - R->c * nu[i] / sq(R->nu_mag) * R->dnu_mag[0] + wind_d[i][3];
    }
    for (int i = 0; i < 3; i++) {
        R->dcg[i][1] = nu[i] / R->nu_mag * R->dc[4] + R->c * mp[i] / R->nu_mag - R->c * nu[i] / sq(R->nu_mag) * R->dnu_mag[1] + wind_d[i][4];
    }
    R->dcg_mag[0] = (R->cg[0] * R->dcg[0][0] + R->cg[1] * R->dcg[1][0] + R->cg[2] * R->dcg[2][0]) / R->cg_mag;
    R->dcg_mag[1] = (R->cg[0] * R->dcg[0][1] + R->cg[1] * R->dcg[1][1] + R->cg[2] * R->dcg[2][1]) / R->cg_mag;

    R->dG[0][0] = 0.0; R->dG[0][1] = 0.0;
    R->dG[1][0] = -Rt[0] / (sq(r));  R->dG[1][1] = -Rp[0] / (sq(r));
    R->dG[2][0] = -Rt[0] / (sq(r) * cos(t)) + sin(t) / (r * sq(cos(t))) * Rt[1];
    R->dG[2][1] = -Rp[0] / (sq(r) * cos(t)) + sin(t) / (r * sq(cos(t))) * Rp[1];

    for (int j = 0; j < 2; j++) {
        const double *Rj = j ? Rp : Rt;
        const double *mj = j ? mp : mt;
        int d = 3 + j;
        R->dT[0][j] = Rj[0] / sq(r) * (nu[1] * R->cg[1] + nu[2] * R->cg[2])
                      - (1.0 / r) * (mj[1] * R->cg[1] + nu[1] * R->dcg[1][j] + mj[2] * R->cg[2] + nu[2] * R->dcg[2][j]);
        R->dT[1][j] = (mj[0] * R->v + nu[0] * R->dv[d] - mj[1] * R->w - nu[1] * R->dw[d]);
        R->dT[1][j] += mj[0] * R->cg[1] + nu[0] * R->dcg[1][j] + (mj[2] * R->cg[2] + nu[2] * R->dcg[2][j]) * tan(t)
                       + nu[2] * R->cg[2] * Rj[1] / sq(cos(t));
        R->dT[2][j] = (mj[0] * R->u + nu[0] * R->du[d] - mj[2] * R->w - nu[2] * R->dw[d]) * cos(t)
                      - (nu[0] * R->u - nu[2] * R->w) * Rj[1] * sin(t);
        R->dT[2][j] += (mj[1] * R->u + nu[1] * R->du[d] - mj[2] * R->v - nu[2] * R->dv[d]) * sin(t)
                       + (nu[1] * R->u - nu[2] * R->v) * Rj[1] * cos(t);
        R->dT[2][j] += R->dcg[2][j] * (nu[0] * cos(t) - nu[1] * sin(t))
                       + R->cg[2] * ((mj[0] - nu[1] * Rj[1]) * cos(t) - (mj[1] + nu[0] * Rj[1]) * sin(t));
    }
}

/* eval_src_eq sph: geoac.eqset.sph.strat.cpp:375-433. */
static void eqs_sph(const go_params *p, const double *y, const go_refs *R, double *f)
{
    const double *nu = y + 3;
    const double *dwind[3] = {R->dw, R->dv, R->du};   /* (w, v, u) <-> (r, t, p) */
    (void)dwind;
    for (int i = 0; i < 3; i++) f[i] = R->G[i] * R->cg[i] / R->cg_mag;
    for (int i = 0; i < 3; i++)
        f[3 + i] = -R->G[i] / R->cg_mag * (R->nu_mag * R->dc[i] + nu[0] * R->dw[i] + nu[1] * R->dv[i] + nu[2] * R->du[i] + R->T[i]);
    if (p->eq_cnt < 18) return;
    for (int j = 0; j < 2; j++) {
        const double *mj = y + (j ? 15 : 9);
        int oR = j ? 12 : 6, om = j ? 15 : 9;
        for (int i = 0; i < 3; i++)
            f[oR + i] = R->dG[i][j] * R->cg[i] / R->cg_mag + R->G[i] * R->dcg[i][j] / R->cg_mag
                        - R->G[i] * R->cg[i] / sq(R->cg_mag) * R->dcg_mag[j];
        for (int i = 0; i < 3; i++)
            f[om + i] = -R->dG[i][j] / R->cg_mag * (R->nu_mag * R->dc[i] + nu[0] * R->dw[i] + nu[1] * R->dv[i] + nu[2] * R->du[i] + R->T[i])
                        + R->G[i] / sq(R->cg_mag) * R->dcg_mag[j] * (R->nu_mag * R->dc[i] + nu[0] * R->dw[i] + nu[1] * R->dv[i] + nu[2] * R->du[i])
                        - R->G[i] / R->cg_mag * (R->dnu_mag[j] * R->dc[i] + R->nu_mag * R->ddc[i][j] + mj[0] * R->dw[i] + mj[1] * R->dv[i] + mj[2] * R->du[i]
                                                 + nu[0] * R->ddw[i][j] + nu[1] * R->ddv[i][j] + nu[2] * R->ddu[i][j] + R->dT[i][j]);
    }
}

/* update_refs 3D: geoac.eqset.3d.strat.cpp:225-328 (keeps the dv[3]
 * accumulation of the azimuth term, line 298, and the upper-triangle second
 * derivative contraction, lines 305-310). */
static void refs_3d(go_atmo *a, const go_params *p, const double *y, go_refs *R)
{
    double x = y[0], yy = y[1], z = y[2];
    const double *nu = y + 3;
    double dd_u_t[6], dd_v_t[6], dd_w_t[6], dd_c[3][3], dd_u[3][3], dd_v[3][3], dd_w[3][3];
    double dc_t, ddc_t;
    memset(R, 0, sizeof(*R));
    R->nu_mag = sqrt(sq(nu[0]) + sq(nu[1]) + sq(nu[2]));
    double z_eval = in_interval(z, a->z[0], a->z[a->n - 1]);
    if (!p->calc_amp) {
        go_spline_all(&a->c, z_eval, &R->c, &dc_t, NULL);
        R->dc[2] = dc_t;
        a->u.accel = a->c.accel;
        a->v.accel = a->c.accel;
        go_calc_uvw(a, x, yy, z, &R->u, &R->v, &R->w, R->du, R->dv, R->dw, NULL, NULL, NULL);
        R->cg[0] = R->c * nu[0] / R->nu_mag + R->u;
        R->cg[1] = R->c * nu[1] / R->nu_mag + R->v;
        R->cg[2] = R->c * nu[2] / R->nu_mag + R->w;
        R->cg_mag = sqrt(sq(R->cg[0]) + sq(R->cg[1]) + sq(R->cg[2]));
        return;
    }
    go_spline_all(&a->c, z_eval, &R->c, &dc_t, &ddc_t);
    memset(dd_c, 0, sizeof(dd_c));
    R->dc[2] = dc_t;
    dd_c[2][2] = ddc_t;
    a->u.accel = a->c.accel;
    a->v.accel = a->c.accel;
    go_calc_uvw(a, x, yy, z, &R->u, &R->v, &R->w, R->du, R->dv, R->dw, dd_u_t, dd_v_t, dd_w_t);
    R->cg[0] = R->c * nu[0] / R->nu_mag + R->u;
    R->cg[1] = R->c * nu[1] / R->nu_mag + R->v;
    R->cg[2] = R->c * nu[2] / R->nu_mag + R->w;
    R->cg_mag = sqrt(sq(R->cg[0]) + sq(R->cg[1]) + sq(R->cg[2]));

    for (int n = 0; n < 3; n++) { dd_u[n][n] = dd_u_t[n]; dd_v[n][n] = dd_v_t[n]; dd_w[n][n] = dd_w_t[n]; }
    dd_u[0][1] = dd_u[1][0] = dd_u_t[3]; dd_u[0][2] = dd_u[2][0] = dd_u_t[4]; dd_u[1][2] = dd_u[2][1] = dd_u_t[5];
    dd_v[0][1] = dd_v[1][0] = dd_v_t[3]; dd_v[0][2] = dd_v[2][0] = dd_v_t[4]; dd_v[1][2] = dd_v[2][1] = dd_v_t[5];
    dd_w[0][1] = dd_w[1][0] = dd_w_t[3]; dd_w[0][2] = dd_w[2][0] = dd_w_t[4]; dd_w[1][2] = dd_w[2][1] = dd_w_t[5];

    const double *Xt = y + 6, *mt = y + 9, *Xp = y + 12, *mp = y + 15;
    for (int n = 0; n < 3; n++) {
        R->dc[3] += Xt[n] * R->dc[n]; R->du[3] += Xt[n] * R->du[n]; R->dv[3] += Xt[n] * R->dv[n]; R->dw[3] += Xt[n] * R->dw[n];
        R->dc[4] += Xp[n] * R->dc[n]; R->du[4] += Xp[n] * R->du[n]; R->dv[3] += Xp[n] * R->dv[n]; R->dw[4] += Xp[n] * R->dw[n];
        R->ddc[n][0] = 0.0; R->ddc[n][1] = 0.0; R->ddu[n][0] = 0.0; R->ddu[n][1] = 0.0;
        R->ddv[n][0] = 0.0; R->ddv[n][1] = 0.0; R->ddw[n][0] = 0.0; R->ddw[n][1] = 0.0;
        for (int m = n; m < 3; m++) {
            R->ddc[n][0] += Xt[m] * dd_c[n][m]; R->ddc[n][1] += Xp[m] * dd_c[n][m];
            R->ddu[n][0] += Xt[m] * dd_u[n][m]; R->ddu[n][1] += Xp[m] * dd_u[n][m];
            R->ddv[n][0] += Xt[m] * dd_v[n][m]; R->ddv[n][1] += Xp[m] * dd_v[n][m];
            R->ddw[n][0] += Xt[m] * dd_w[n][m]; R->ddw[n][1] += Xp[m] * dd_w[n][m];
        }
    }
    R->dnu_mag[0] = (nu[0] * mt[0] + nu[1] * mt[1] + nu[2] * mt[2]) / R->nu_mag;
    R->dnu_mag[1] = (nu[0] * mp[0] + nu[1] * mp[1] + nu[2] * mp[2]) / R->nu_mag;
    for (int n = 0; n < 3; n++) {
        R->dcg[n][0] = nu[n] / R->nu_mag * R->dc[3] + R->c * mt[n] / R->nu_mag - R->c * nu[n] / sq(R->nu_mag) * R->dnu_mag[0];
        R->dcg[n][1] = nu[n] / R->nu_mag * R->dc[4] + R->c * mp[n] / R->nu_mag - R->c * nu[n] / sq(R->nu_mag) * R->dnu_mag[1];
    }
    R->dcg[0][0] += R->du[3]; R->dcg[0][1] += R->du[4];
    R->dcg[1][0] += R->dv[3]; R->dcg[1][1] += R->dv[4];
    R->dcg[2][0] += R->dw[3]; R->dcg[2][1] += R->dw[4];
    R->dcg_mag[0] = (R->cg[0] * R->dcg[0][0] + R->cg[1] * R->dcg[1][0] + R->cg[2] * R->dcg[2][0]) / R->cg_mag;
    R->dcg_mag[1] = (R->cg[0] * R->dcg[0][1] + R->cg[1] * R->dcg[1][1] + R->cg[2] * R->dcg[2][1]) / R->cg_mag;
}

/* eval_src_eq 3D: geoac.eqset.3d.strat.cpp:333-389 (no terrain: the
 * horizontal slowness and their mu terms stay zero). */
static void eqs_3d(const go_params *p, const double *y, const go_refs *R, double *f)
{
    const double *nu = y + 3;
    for (int i = 0; i < 3; i++) f[i] = R->cg[i] / R->cg_mag;
    f[3] = 0.0; f[4] = 0.0;
    f[5] = -1.0 / R->cg_mag * (R->nu_mag * R->dc[2] + nu[0] * R->du[2] + nu[1] * R->dv[2] + nu[2] * R->dw[2]);
    if (p->eq_cnt < 18) return;
    for (int j = 0; j < 2; j++) {
        const double *mj = y + (j ? 15 : 9);
        int oX = j ? 12 : 6, om = j ? 15 : 9;
        for (int i = 0; i < 3; i++)
            f[oX + i] = R->dcg[i][j] / R->cg_mag - R->cg[i] / sq(R->cg_mag) * R->dcg_mag[j];
        f[om + 0] = 0.0; f[om + 1] = 0.0;
        f[om + 2] = R->dcg_mag[j] / sq(R->cg_mag) * (R->nu_mag * R->dc[2] + nu[0] * R->du[2] + nu[1] * R->dv[2] + nu[2] * R->dw[2])
                    - 1.0 / R->cg_mag * (R->dnu_mag[j] * R->dc[2] + R->nu_mag * R->ddc[2][j] + mj[0] * R->du[2] + mj[1] * R->dv[2] + mj[2] * R->dw[2]
                                         + nu[0] * R->ddu[2][j] + nu[1] * R->ddv[2][j] + nu[2] * R->ddw[2][j]);
    }
}

void go_rhs(go_atmo *a, const go_params *p, const go_src *s, const double *y, double *f)
{
    go_refs R;
    (void)s;
    if (a->geom == GA_GEOM_SPH) { refs_sph(a, p, y, &R); eqs_sph(p, y, &R, f); }
    else { refs_3d(a, p, y, &R); eqs_3d(p, y, &R, f); }
}

/* set_ds: sph :216-222, 3D :214-220. */
double go_set_ds(go_atmo *a, const go_params *p, const double *y)
{
    double result = p->ds_max;
    double h = a->geom == GA_GEOM_SPH ? y[0] : y[2];
    if (zc(a, y[0], y[2]) < a->z_bndlyr)
        result -= (p->ds_max - p->ds_min) * exp(-(h - ground(a)) / 0.2);
    return result;
}

/* break_check: sph :484-498 (wraps longitude in place), 3D :438-446. */
int go_break_check(go_atmo *a, const go_params *p, const go_src *s, double *y)
{
    if (a->geom == GA_GEOM_SPH) {
        if (y[2] > GO_PI) y[2] -= 2.0 * GO_PI;
        else if (y[2] < -GO_PI) y[2] += 2.0 * GO_PI;
        if (y[1] < -GO_PI / 2.0 || y[1] > GO_PI / 2.0) return 1;
        if (y[0] - a->r0 > p->alt_max) return 1;
        if (go_gc_dist(a->r0, y[1], y[2], s->src[1], s->src[2]) > p->rng_max) return 1;
        if (y[1] < p->lat_min || y[1] > p->lat_max) return 1;
        if (y[2] < p->lon_min || y[2] > p->lon_max) return 1;
        return 0;
    }
    if (sqrt(sq(y[0] - s->src[0]) + sq(y[1] - s->src[1])) > p->rng_max) return 1;
    if (y[0] > p->x_max || y[0] < p->x_min) return 1;
    if (y[1] > p->y_max || y[1] < p->y_min) return 1;
    if (y[2] > p->alt_max) return 1;
    return 0;
}

/* ground_check: sph :500-507, 3D :448-453. */
int go_ground_check(go_atmo *a, const double *y)
{
    if (a->geom == GA_GEOM_SPH) {
        if (y[0] - a->r0 <= a->z_max)
            if (y[0] < ground(a)) return 1;
        return 0;
    }
    if (y[2] <= a->z_max)
        if (y[2] < ground(a)) return 1;
    return 0;
}

/* Classical RK4 step: rk4solver.cpp:29-55. */
double go_rk4_step(go_atmo *a, const go_params *p, const go_src *s, const double *y, double *y1)
{
    int n = p->eq_cnt;
    double t1[18], t2[18], t3[18], t4[18], pa[18], f[18];
    double ds = go_set_ds(a, p, y);
    go_rhs(a, p, s, y, f);
    for (int i = 0; i < n; i++) { t1[i] = ds * f[i]; pa[i] = y[i] + t1[i] / 2.0; }
    go_rhs(a, p, s, pa, f);
    for (int i = 0; i < n; i++) { t2[i] = ds * f[i]; }
    for (int i = 0; i < n; i++) { pa[i] = y[i] + t2[i] / 2.0; }
    go_rhs(a, p, s, pa, f);
    for (int i = 0; i < n; i++) { t3[i] = ds * f[i]; }
    for (int i = 0; i < n; i++) { pa[i] = y[i] + t3[i]; }
    go_rhs(a, p, s, pa, f);
    for (int i = 0; i < n; i++) {
        t4[i] = ds * f[i];
        y1[i] = y[i] + t1[i] / 6.0 + t2[i] / 3.0 + t3[i] / 3.0 + t4[i] / 6.0;
    }
    return ds;
}

/* prop_rk4: rk4solver.cpp:14-85.  rows is [length][18]. */
int go_prop_rk4(go_atmo *a, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    int k;
    double sum_s = 0.0;
    *check = 0;
    for (k = 0; k < length - 1; k++) {
        double *y = rows + (size_t)k * 18, *y1 = rows + (size_t)(k + 1) * 18;
        double ds = go_rk4_step(a, p, s, y, y1);
        sum_s += ds;
        if (sum_s > p->s_max) { *check = 1; break; }
        if (go_break_check(a, p, s, y1)) { *check = 1; break; }
        if (go_ground_check(a, y1)) { *check = 0; break; }
    }
    return k + 1;
}

/* approx_intercept + set_refl for flat ground: sph :108-211, 3D :110-209.
 * With zero terrain slope C1* = 1, C2* = 0 and every da/db term vanishes;
 * what survives is the specular flip of the vertical components and the
 * ds0 jump term of the vertical mu component. */
void go_set_refl(go_atmo *a, const go_params *p, const go_src *s,
                 const double *rk2, const double *rk1, double *y0)
{
    int n = p->eq_cnt;
    double prev[18] = {0};
    int iv = vert(a);
    double hg = ground(a);
    double dr_grnd = hg - rk1[iv];
    double dr_k = (rk1[iv] - rk2[iv]) - (rk1[a->geom == GA_GEOM_SPH ? 1 : 0] - rk2[a->geom == GA_GEOM_SPH ? 1 : 0]) * 0.0
                  - (rk1[a->geom == GA_GEOM_SPH ? 2 : 1] - rk2[a->geom == GA_GEOM_SPH ? 2 : 1]) * 0.0;
    if (a->geom == GA_GEOM_SPH) {
        /* sph :110-112 divides the slope terms by r and r cos(lat); dz = 0 */
        dr_k = (rk1[0] - rk2[0]) - (rk1[1] - rk2[1]) * 1.0 / rk1[0] * 0.0
               - (rk1[2] - rk2[2]) * 1.0 / (rk1[0] * cos(rk1[1])) * 0.0;
    }
    for (int i = 0; i < n; i++) prev[i] = rk1[i] + (rk1[i] - rk2[i]) / dr_k * dr_grnd;

    if (a->geom == GA_GEOM_SPH) {
        double rg = hg;
        double c_grnd = go_c(a, rg, prev[1], prev[2]);
        double C1r = 1.0, C1th = 1.0, C1ph = 1.0, C2th = 0.0, C2ph = 0.0;
        /* a = b = 0: C* from sph :131-135 evaluate exactly to these values */
        y0[0] = rg;
        y0[1] = prev[1];
        y0[2] = prev[2];
        y0[3] = -C1r * prev[3] + C2th * prev[4] + C2ph * prev[5];
        y0[4] = C1th * prev[4] + C2th * (prev[3] - 0.0 * prev[5]);
        y0[5] = C1ph * prev[5] + C2ph * (prev[3] - 0.0 * prev[4]);
        if (n == 18) {
            double norm2 = prev[3];                                  /* sph :138 */
            double ds0_t = -s->c0 / c_grnd * (prev[6] - rg * 0.0 * prev[7] - rg * cos(prev[1]) * 0.0 * prev[8]) / norm2;
            double ds0_p = -s->c0 / c_grnd * (prev[12] - rg * 0.0 * prev[13] - rg * cos(prev[1]) * 0.0 * prev[14]) / norm2;
            double dnr = s->c0 / c_grnd * go_dc(a, rg, prev[1], prev[2], 0) + prev[3] * 0.0
                         + prev[4] * go_dv(a, rg, prev[1], prev[2], 0) + prev[5] * go_du(a, rg, prev[1], prev[2], 0);
            double dnt = s->c0 / c_grnd * go_dc(a, rg, prev[1], prev[2], 1) + prev[3] * 0.0
                         + prev[4] * go_dv(a, rg, prev[1], prev[2], 1) + prev[5] * go_du(a, rg, prev[1], prev[2], 1);
            double dnp = s->c0 / c_grnd * go_dc(a, rg, prev[1], prev[2], 2) + prev[3] * 0.0
                         + prev[4] * go_dv(a, rg, prev[1], prev[2], 2) + prev[5] * go_du(a, rg, prev[1], prev[2], 2);
            dnr += sq(c_grnd) / s->c0 * 1.0 / rg * (sq(prev[4]) + sq(prev[5]));
            dnt -= sq(c_grnd) / s->c0 * (prev[3] * prev[4] + tan(prev[1]) * sq(prev[5]));
            dnp -= sq(c_grnd) / s->c0 * prev[5] * (prev[3] * cos(prev[1]) + prev[4] * sin(prev[1]));
            dnr *= -1.0 / c_grnd;
            dnt *= -1.0 / (c_grnd * rg);
            dnp *= -1.0 / (c_grnd * rg * cos(prev[1]));
            for (int j = 0; j < 2; j++) {
                int o = j ? 12 : 6;
                double ds0 = j ? ds0_p : ds0_t;
                y0[o + 0] = -C1r * prev[o] + rg * C2th * prev[o + 1] + rg * cos(prev[1]) * C2ph * prev[o + 2];
                y0[o + 1] = C1th * prev[o + 1] + C2th * (prev[o] / rg - 0.0 * cos(prev[1]) * prev[o + 2]);
                y0[o + 2] = C1ph * prev[o + 2] + C2ph * (prev[o] / (rg * cos(prev[1])) - 0.0 / cos(prev[1]) * prev[o + 1]);
                /* mu: the dC and da/db terms are exactly zero on flat ground */
                y0[o + 3] = -C1r * prev[o + 3] + C2th * prev[o + 4] + C2ph * prev[o + 5] - 0.0 * prev[3] + 0.0 * prev[4] + 0.0 * prev[5]
                            + (-(C1r + 1.0) * dnr + C2th * dnt + C2ph * dnp) * ds0;
                y0[o + 4] = C1th * prev[o + 4] + C2th * (prev[o + 3] - 0.0 * prev[o + 5] - 0.0 * prev[5]) + 0.0 * prev[4]
                            + 0.0 * (prev[3] - 0.0 * prev[5]) + ((C1th - 1.0) * dnt + C2th * (dnr - 0.0 * dnp)) * ds0;
                y0[o + 5] = C1ph * prev[o + 5] + C2ph * (prev[o + 3] - 0.0 * prev[o + 4] - 0.0 * prev[4]) + 0.0 * prev[5]
                            + 0.0 * (prev[3] - 0.0 * prev[4]) + ((C1ph - 1.0) * dnp + C2ph * (dnr - 0.0 * dnt)) * ds0;
            }
        }
        return;
    }

    /* 3D: 3d.strat.cpp:114-205 with a = b = 0 */
    double zg = hg;
    double c_grnd = go_c(a, prev[0], prev[1], zg);
    double C1x = 1.0, C1y = 1.0, C1z = 1.0, C2x = 0.0, C2y = 0.0;
    y0[0] = prev[0];
    y0[1] = prev[1];
    y0[2] = zg;
    y0[3] = C1x * prev[3] + C2x * (prev[5] - prev[4] * 0.0);
    y0[4] = C1y * prev[4] + C2y * (prev[5] - prev[3] * 0.0);
    y0[5] = -C1z * prev[5] + prev[3] * C2x + prev[4] * C2y;
    if (n == 18) {
        double norm2 = prev[5];                                      /* 3d :136 */
        double ds0_t = -s->c0 / c_grnd * (prev[8] - 0.0 * prev[6] - 0.0 * prev[7]) / norm2;
        double ds0_p = -s->c0 / c_grnd * (prev[14] - 0.0 * prev[12] - 0.0 * prev[13]) / norm2;
        double dnx = -1.0 / c_grnd * (s->c0 / c_grnd * go_dc(a, prev[0], prev[1], zg, 0) + prev[3] * go_du(a, prev[0], prev[1], zg, 0)
                                      + prev[4] * go_dv(a, prev[0], prev[1], zg, 0) + prev[5] * 0.0);
        double dny = -1.0 / c_grnd * (s->c0 / c_grnd * go_dc(a, prev[0], prev[1], zg, 1) + prev[3] * go_du(a, prev[0], prev[1], zg, 1)
                                      + prev[4] * go_dv(a, prev[0], prev[1], zg, 1) + prev[5] * 0.0);
        double dnz = -1.0 / c_grnd * (s->c0 / c_grnd * go_dc(a, prev[0], prev[1], zg, 2) + prev[3] * go_du(a, prev[0], prev[1], zg, 2)
                                      + prev[4] * go_dv(a, prev[0], prev[1], zg, 2) + prev[5] * 0.0);
        for (int j = 0; j < 2; j++) {
            int o = j ? 12 : 6;
            double ds0 = j ? ds0_p : ds0_t;
            y0[o + 0] = C1x * prev[o] + C2x * (prev[o + 2] - prev[o + 1] * 0.0);
            y0[o + 1] = C1y * prev[o + 1] + C2y * (prev[o + 2] - prev[o] * 0.0);
            y0[o + 2] = -C1z * prev[o + 2] + prev[o] * C2x + prev[o + 1] * C2y;
            y0[o + 3] = C1x * prev[o + 3] + C2x * (prev[o + 5] - 0.0 * prev[o + 4] - prev[4] * 0.0) + 0.0 * prev[3]
                        + 0.0 * (prev[5] - 0.0 * prev[4]) + ((C1x - 1.0) * dnx + C2x * (dnz - 0.0 * dny)) * ds0;
            y0[o + 4] = C1y * prev[o + 4] + C2y * (prev[o + 5] - 0.0 * prev[o + 3] - prev[3] * 0.0) + 0.0 * prev[4]
                        + 0.0 * (prev[5] - 0.0 * prev[3]) + ((C1y - 1.0) * dny + C2y * (dnz - 0.0 * dnx)) * ds0;
            y0[o + 5] = -C1z * prev[o + 5] + C2x * prev[o + 3] + C2y * prev[o + 4] - 0.0 * prev[5] + 0.0 * prev[3] + 0.0 * prev[4]
                        + (-(C1z + 1.0) * dnz + C2x * dnx + C2y * dny) * ds0;
        }
    }
}

/* Midpoint-rule travel-time segment: sph :549-577, 3D :498-513. */
double go_travel_time_seg(go_atmo *a, const double *y0, const double *y1)
{
    double ds, nu[3], nu_mag, cp[3], cp_mag, snd;
    if (a->geom == GA_GEOM_SPH) {
        double dr = y1[0] - y0[0], r = y0[0] + dr / 2.0;
        double dt = y1[1] - y0[1], t = y0[1] + dt / 2.0;
        double dp = y1[2] - y0[2], pp = y0[2] + dp / 2.0;
        if (dp > GO_PI) dp -= 2.0 * GO_PI;
        else if (dp < -GO_PI) dp += 2.0 * GO_PI;
        ds = sqrt(sq(dr) + sq(r * dt) + sq(r * cos(t) * dp));
        for (int i = 0; i < 3; i++) nu[i] = y0[3 + i] + (y1[3 + i] - y0[3 + i]) / 2.0;
        nu_mag = sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
        snd = go_c(a, r, t, pp);
        cp[0] = snd * nu[0] / nu_mag + 0.0;
        cp[1] = snd * nu[1] / nu_mag + go_v(a, r, t, pp);
        cp[2] = snd * nu[2] / nu_mag + go_u(a, r, t, pp);
        cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
        return ds / cp_mag;
    }
    double dx = y1[0] - y0[0], x = y0[0] + dx / 2.0;
    double dy = y1[1] - y0[1], yy = y0[1] + dy / 2.0;
    double dz = y1[2] - y0[2], z = y0[2] + dz / 2.0;
    ds = sqrt(dx * dx + dy * dy + dz * dz);     /* whole-leg form, 3d :469 */
    for (int i = 0; i < 3; i++) nu[i] = y0[3 + i] + (y1[3 + i] - y0[3 + i]) / 2.0;
    nu_mag = sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
    snd = go_c(a, x, yy, z);
    cp[0] = snd * nu[0] / nu_mag + go_u(a, x, yy, z);
    cp[1] = snd * nu[1] / nu_mag + go_v(a, x, yy, z);
    cp[2] = snd * nu[2] / nu_mag + 0.0;
    cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    return ds / cp_mag;
}

/* Sutherland-Bass attenuation segment: sph :774-794, 3D :685-699. */
double go_atten_seg(go_atmo *a, const double *y0, const double *y1, double freq)
{
    if (a->geom == GA_GEOM_SPH) {
        double dr = y1[0] - y0[0], r = y0[0] + dr / 2.0;
        double dt = y1[1] - y0[1], t = y0[1] + dt / 2.0;
        double dp = y1[2] - y0[2], pp = y0[2] + dp / 2.0;
        if (dp > GO_PI) dp -= 2.0 * GO_PI;
        else if (dp < -GO_PI) dp += 2.0 * GO_PI;
        double ds = sqrt(sq(dr) + sq(r * dt) + sq(r * cos(t) * dp));
        return go_sb_alpha(a, r, t, pp, freq) * ds;
    }
    double dx = y1[0] - y0[0], x = y0[0] + dx / 2.0;
    double dy = y1[1] - y0[1], yy = y0[1] + dy / 2.0;
    double dz = y1[2] - y0[2], z = y0[2] + dz / 2.0;
    double ds = sqrt(sq(dx) + sq(dy) + sq(dz));
    return go_sb_alpha(a, x, yy, z, freq) * ds;
}

/* Jacobian determinant: sph :800-823, 3D :704-723. */
double go_jacobian(go_atmo *a, const double *y)
{
    double nu_mag = sqrt(sq(y[3]) + sq(y[4]) + sq(y[5]));
    double c = go_c(a, y[0], y[1], y[2]);
    double cp[3], cp_mag;
    if (a->geom == GA_GEOM_SPH) {
        double r = y[0], t = y[1], p = y[2];
        cp[0] = go_c(a, r, t, p) * y[3] / nu_mag + 0.0;
        cp[1] = go_c(a, r, t, p) * y[4] / nu_mag + go_v(a, r, t, p);
        cp[2] = go_c(a, r, t, p) * y[5] / nu_mag + go_u(a, r, t, p);
        cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
        double dr_ds = cp[0] / cp_mag, dt_ds = 1.0 / r * cp[1] / cp_mag, dp_ds = 1.0 / (r * cos(t)) * cp[2] / cp_mag;
        double dr_t = y[6], dt_t = y[7], dp_t = y[8], dr_p = y[12], dt_p = y[13], dp_p = y[14];
        return sq(r) * cos(t) * (dr_ds * (dt_t * dp_p - dt_p * dp_t) - dr_t * (dt_ds * dp_p - dp_ds * dt_p)
                                 + dr_p * (dt_ds * dp_t - dp_ds * dt_t));
    }
    cp[0] = c * y[3] / nu_mag + go_u(a, y[0], y[1], y[2]);
    cp[1] = c * y[4] / nu_mag + go_v(a, y[0], y[1], y[2]);
    cp[2] = c * y[5] / nu_mag + 0.0;
    cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    double dxds = cp[0] / cp_mag, dyds = cp[1] / cp_mag, dzds = cp[2] / cp_mag;
    return dxds * (y[7] * y[14] - y[13] * y[8]) - y[6] * (dyds * y[14] - dzds * y[13]) + y[12] * (dyds * y[8] - dzds * y[7]);
}

/* Transport-equation amplitude: sph :825-849, 3D :725-748. */
double go_amp(go_atmo *a, const go_src *s, const double *y)
{
    double num, den;
    if (a->geom == GA_GEOM_SPH) {
        double r = y[0], t = y[1], p = y[2];
        double r0 = s->src[0], t0 = s->src[1], p0 = s->src[2];
        double nu0[3] = {s->nu0 * sin(s->theta), s->nu0 * cos(s->theta) * sin(s->phi), s->nu0 * cos(s->theta) * cos(s->phi)};
        double nu_mag = (s->c0 - y[3] * 0.0 - y[4] * go_v(a, r, t, p) - y[5] * go_u(a, r, t, p)) / go_c(a, r, t, p);
        double nu_mag0 = s->nu0;
        double c0p[3], cp[3];
        c0p[0] = s->c0 * nu0[0] / nu_mag0 + 0.0;                      cp[0] = go_c(a, r, t, p) * y[3] / nu_mag + 0.0;
        c0p[1] = s->c0 * nu0[1] / nu_mag0 + go_v(a, r0, t0, p0);      cp[1] = go_c(a, r, t, p) * y[4] / nu_mag + go_v(a, r, t, p);
        c0p[2] = s->c0 * nu0[2] / nu_mag0 + go_u(a, r0, t0, p0);      cp[2] = go_c(a, r, t, p) * y[5] / nu_mag + go_u(a, r, t, p);
        double cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
        double cp_mag0 = sqrt(sq(c0p[0]) + sq(c0p[1]) + sq(c0p[2]));
        num = go_rho(a, r, t, p) * nu_mag * pow(go_c(a, r, t, p), 3) * cp_mag0 * cos(s->theta);
        den = go_rho(a, r0, t0, p0) * nu_mag0 * pow(go_c(a, r0, t0, p0), 3) * cp_mag * go_jacobian(a, y);
        return sqrt(fabs(num / den));
    }
    double snd = go_c(a, y[0], y[1], y[2]);
    double wi[3] = {go_u(a, y[0], y[1], y[2]), go_v(a, y[0], y[1], y[2]), 0.0};
    double w0[3] = {go_u(a, s->src[0], s->src[1], s->src[2]), go_v(a, s->src[0], s->src[1], s->src[2]), 0.0};
    double nu_mag = sqrt(sq(y[3]) + sq(y[4]) + sq(y[5]));
    double nu_mag0 = 1.0 - (y[3] * w0[0] + y[4] * w0[1] + y[5] * w0[2]) / s->c0;
    double cg[3], cg0[3];
    for (int n = 0; n < 3; n++) cg[n] = snd * y[3 + n] / nu_mag + wi[n];
    cg0[0] = s->c0 * cos(s->theta) * cos(s->phi) + w0[0];
    cg0[1] = s->c0 * cos(s->theta) * sin(s->phi) + w0[1];
    cg0[2] = s->c0 * sin(s->theta) + w0[2];
    double cg_mag = sqrt(sq(cg[0]) + sq(cg[1]) + sq(cg[2]));
    double cg_mag0 = sqrt(sq(cg0[0]) + sq(cg0[1]) + sq(cg0[2]));
    num = go_rho(a, y[0], y[1], y[2]) * nu_mag * pow(snd, 3) * cg_mag0 * cos(s->theta);
    den = go_rho(a, s->src[0], s->src[1], s->src[2]) * nu_mag0 * pow(s->c0, 3) * cg_mag * go_jacobian(a, y);
    return sqrt(fabs(num / den));
}

/* Eikonal residual |nu| - c0/c + (v . nu)/c: sph :438-452, 3D :394-401. */
double go_eikonal_resid(go_atmo *a, const go_src *s, const double *y)
{
    double c = go_c(a, y[0], y[1], y[2]);
    double vn;
    if (a->geom == GA_GEOM_SPH)
        vn = 0.0 * y[3] + go_v(a, y[0], y[1], y[2]) * y[4] + go_u(a, y[0], y[1], y[2]) * y[5];
    else
        vn = go_u(a, y[0], y[1], y[2]) * y[3] + go_v(a, y[0], y[1], y[2]) * y[4] + 0.0 * y[5];
    return sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]) - s->c0 / c + vn / c;
}

/* ======================================================================
 * -prop fan driver (write_rays=false): main.infraga.sph.cpp:418-558,
 * main.infraga.3d.cpp:407-548.
 * ==================================================================== */
static double *rows_buf = NULL;
static int rows_len = 0;

static double *get_rows(int length)
{
    if (rows_len < length) {
        free(rows_buf);
        rows_buf = (double *)malloc(sizeof(double) * 18 * (size_t)length);
        rows_len = length;
    }
    return rows_buf;
}

int go_fan(go_atmo *a, const go_params *p, const go_prop *pr,
           const double *launches, int n_launch,
           go_arrival *out, int out_cap, int *n_out, long long *ray_steps)
{
    const int sph = (a->geom == GA_GEOM_SPH);
    int length = (int)(p->s_max / p->ds_min);
    double *rows = get_rows(length);
    go_src s;
    int cnt = 0;
    long long steps = 0;
    double lat_src = 0, lon_src = 0, z_src;

    if (sph) {
        lat_src = pr->src0 * (GO_PI / 180.0);                /* sph :311-312 */
        lon_src = pr->src1 * (GO_PI / 180.0);
        z_src = pr->src_alt;
        double zg = ground(a) - a->r0;
        z_src = z_src > zg ? z_src : zg;                      /* sph :314 */
        s.src[0] = a->r0 + z_src; s.src[1] = lat_src; s.src[2] = lon_src;
    } else {
        z_src = pr->src_alt > ground(a) ? pr->src_alt : ground(a);   /* 3d :306 */
        s.src[0] = pr->src0; s.src[1] = pr->src1; s.src[2] = z_src;
    }

    for (int li = 0; li < n_launch; li++) {
        double incl = launches[2 * li], az = launches[2 * li + 1];
        s.phi = GO_PI / 2.0 - az * (GO_PI / 180.0);
        s.theta = incl * (GO_PI / 180.0);
        go_set_initial(a, p, &s, rows);
        double tt = 0.0, att = 0.0;
        double rmax = sph ? 0.0 : a->z0;
        int brk = 0, k;
        for (int bnc = 0; bnc <= pr->bounces; bnc++) {
            k = go_prop_rk4(a, p, &s, rows, length, &brk);
            steps += k;
            double tleg = 0.0, aleg = 0.0;
            for (int n = 0; n < k; n++) tleg += go_travel_time_seg(a, rows + 18 * n, rows + 18 * (n + 1));
            tt += tleg;
            for (int n = 0; n < k; n++) aleg += go_atten_seg(a, rows + 18 * n, rows + 18 * (n + 1), pr->freq);
            att += aleg;
            for (int m = 0; m < k; m++) {
                double zz = sph ? rows[18 * m] - a->r0 : rows[18 * m + 2];
                rmax = rmax > zz ? rmax : zz;
            }
            if (rmax < (a->z0 + pr->turn_ht_min)) break;           /* sph :515 */
            if (brk || k < 2) break;
            if (sph && (tt / 3600.0) > p->time_max) break;
            double *yk = rows + 18 * k;
            go_arrival A;
            memset(&A, 0, sizeof(A));
            A.launch_idx = li;
            A.n_bounce = bnc;
            A.incl_launch = incl;
            A.az_launch = az;
            double baz;
            if (sph) {
                A.incl_arr = -asin(go_c(a, yk[0], yk[1], yk[2]) / go_c(a, a->r0 + z_src, lat_src, lon_src) * yk[3]) * (180.0 / GO_PI);
                baz = 90.0 - atan2(-yk[4], -yk[5]) * (180.0 / GO_PI);
                A.pos0 = yk[1] * (180.0 / GO_PI);
                A.pos1 = yk[2] * (180.0 / GO_PI);
                A.cel = go_gc_dist(a->r0, yk[1], yk[2], lat_src, lon_src) / tt;
            } else {
                A.incl_arr = -asin(go_c(a, yk[0], yk[1], ground(a)) / go_c(a, s.src[0], s.src[1], s.src[2]) * yk[5]) * 180.0 / GO_PI;
                baz = 90.0 - atan2(-yk[4], -yk[3]) * 180.0 / GO_PI;
                A.pos0 = yk[0];
                A.pos1 = yk[1];
                A.cel = sqrt(sq(yk[0] - s.src[0]) + sq(yk[1] - s.src[1])) / tt;
            }
            if (baz < -180.0) baz += 360.0;
            else if (baz > 180.0) baz -= 360.0;
            A.back_az = baz;
            A.time = tt;
            A.turn_ht = rmax;
            A.tl = p->calc_amp ? 20.0 * log10(go_amp(a, &s, yk)) : 0.0;
            A.absorption = -2.0 * att;
            if (cnt < out_cap) out[cnt] = A;
            cnt++;
            go_set_refl(a, p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
        }
    }
    *n_out = cnt;
    if (ray_steps) *ray_steps = steps;
    return cnt > out_cap ? -1 : 0;
}

/* ======================================================================
 * -prop fan with write_rays / write_caustics: main.infraga.sph.cpp:450-496
 * (3D: main.infraga.3d.cpp:442-487).  Everything else as go_fan.
 * ==================================================================== */
typedef struct {
    go_path_pt *pts;
    long long cap, n;
} path_sink;

static void sink_pt(path_sink *k, int li, int leg, int m, int kind, const go_atmo *a, const double *y,
                    double tl, double att, double tt)
{
    if (!k) return;
    if (k->n < k->cap) {
        go_path_pt *P = k->pts + k->n;
        P->launch_idx = li;
        P->leg = leg;
        P->step = m;
        P->kind = kind;
        if (a->geom == GA_GEOM_SPH) {
            P->pos0 = y[1] * (180.0 / GO_PI);
            P->pos1 = y[2] * (180.0 / GO_PI);
            P->z = y[0] - a->r0;
        } else {
            P->pos0 = y[0];
            P->pos1 = y[1];
            P->z = y[2] > ground(a) ? y[2] : ground(a);     /* max(z, topo::z(x, y)) */
        }
        P->tl = tl;
        P->absorption = -2.0 * att;
        P->time = tt;
    }
    k->n++;
}

int go_fan_paths(go_atmo *a, const go_params *p, const go_prop *pr,
                 const double *launches, int n_launch, int flags,
                 go_arrival *out, int out_cap, int *n_out,
                 go_path_pt *pts, long long pts_cap, long long *n_pts, long long *ray_steps)
{
    const int sph = (a->geom == GA_GEOM_SPH);
    const int rays = (flags & GO_PATH_RAYS) != 0, caus = (flags & GO_PATH_CAUSTICS) != 0 && p->calc_amp;
    int length = (int)(p->s_max / p->ds_min);
    double *rows = get_rows(length);
    go_src s;
    int cnt = 0;
    long long steps = 0;
    double lat_src = 0, lon_src = 0, z_src;
    path_sink sink = {pts, pts_cap, 0};

    if (sph) {
        lat_src = pr->src0 * (GO_PI / 180.0);
        lon_src = pr->src1 * (GO_PI / 180.0);
        z_src = pr->src_alt;
        double zg = ground(a) - a->r0;
        z_src = z_src > zg ? z_src : zg;
        s.src[0] = a->r0 + z_src; s.src[1] = lat_src; s.src[2] = lon_src;
    } else {
        z_src = pr->src_alt > ground(a) ? pr->src_alt : ground(a);
        s.src[0] = pr->src0; s.src[1] = pr->src1; s.src[2] = z_src;
    }

    for (int li = 0; li < n_launch; li++) {
        double incl = launches[2 * li], az = launches[2 * li + 1];
        s.phi = GO_PI / 2.0 - az * (GO_PI / 180.0);
        s.theta = incl * (GO_PI / 180.0);
        go_set_initial(a, p, &s, rows);
        double tt = 0.0, att = 0.0;
        double rmax = sph ? 0.0 : a->z0;
        int brk = 0, k;
        for (int bnc = 0; bnc <= pr->bounces; bnc++) {
            k = go_prop_rk4(a, p, &s, rows, length, &brk);
            steps += k;
            /* sph :456-496 / 3d :442-487: segments of rows 1..k-1 one by one */
            double D = 0.0, D_prev = caus ? go_jacobian(a, rows + 18) : 0.0;
            for (int m = 1; m < k; m++) {
                const double *ym = rows + 18 * m;
                if (caus) D = go_jacobian(a, ym);
                tt += go_travel_time_seg(a, rows + 18 * (m - 1), ym);
                att += go_atten_seg(a, rows + 18 * (m - 1), ym, pr->freq);
                if (rays && (m == 1 || m % 15 == 0))
                    sink_pt(&sink, li, bnc, m, 0, a, ym, p->calc_amp ? 20.0 * log10(go_amp(a, &s, ym)) : 0.0, att, tt);
                if (caus && D * D_prev < 0.0) sink_pt(&sink, li, bnc, m, 1, a, ym, 0.0, att, tt);
                if (caus) D_prev = D;
            }
            for (int m = 0; m < k; m++) {
                double zz = sph ? rows[18 * m] - a->r0 : rows[18 * m + 2];
                rmax = rmax > zz ? rmax : zz;
            }
            if (rmax < (a->z0 + pr->turn_ht_min)) break;
            if (brk || k < 2) break;
            if (sph && (tt / 3600.0) > p->time_max) break;
            double *yk = rows + 18 * k;
            go_arrival A;
            memset(&A, 0, sizeof(A));
            A.launch_idx = li;
            A.n_bounce = bnc;
            A.incl_launch = incl;
            A.az_launch = az;
            double baz;
            if (sph) {
                A.incl_arr = -asin(go_c(a, yk[0], yk[1], yk[2]) / go_c(a, a->r0 + z_src, lat_src, lon_src) * yk[3]) * (180.0 / GO_PI);
                baz = 90.0 - atan2(-yk[4], -yk[5]) * (180.0 / GO_PI);
                A.pos0 = yk[1] * (180.0 / GO_PI);
                A.pos1 = yk[2] * (180.0 / GO_PI);
                A.cel = go_gc_dist(a->r0, yk[1], yk[2], lat_src, lon_src) / tt;
            } else {
                A.incl_arr = -asin(go_c(a, yk[0], yk[1], ground(a)) / go_c(a, s.src[0], s.src[1], s.src[2]) * yk[5]) * 180.0 / GO_PI;
                baz = 90.0 - atan2(-yk[4], -yk[3]) * 180.0 / GO_PI;
                A.pos0 = yk[0];
                A.pos1 = yk[1];
                A.cel = sqrt(sq(yk[0] - s.src[0]) + sq(yk[1] - s.src[1])) / tt;
            }
            if (baz < -180.0) baz += 360.0;
            else if (baz > 180.0) baz -= 360.0;
            A.back_az = baz;
            A.time = tt;
            A.turn_ht = rmax;
            A.tl = p->calc_amp ? 20.0 * log10(go_amp(a, &s, yk)) : 0.0;
            A.absorption = -2.0 * att;
            if (cnt < out_cap) out[cnt] = A;
            cnt++;
            go_set_refl(a, p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
        }
    }
    *n_out = cnt;
    if (n_pts) *n_pts = sink.n;
    if (ray_steps) *ray_steps = steps;
    return (cnt > out_cap || sink.n > pts_cap) ? -1 : 0;
}

#ifndef GO_OPCOUNT
/* ======================================================================
 * Eigenray search, 3D: geoac.eigenray.3d.cpp:34-366 and the driver loop of
 * main.infraga.3d.cpp:944-951.
 * ==================================================================== */
/* A stratified medium (either geometry) as a go_eig_med: topo::z is the flat
 * ground (sph: r0 + z0, 3D: z0) with zero slopes. */
static void sm_init(void *m, const go_params *p, go_src *s, double *y) { go_set_initial((go_atmo *)m, p, s, y); }
static int sm_prop(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    return go_prop_rk4((go_atmo *)m, p, s, rows, length, check);
}
static void sm_refl(void *m, const go_params *p, const go_src *s, const double *a, const double *b, double *o)
{
    go_set_refl((go_atmo *)m, p, s, a, b, o);
}
static double sm_tt(void *m, const double *a, const double *b) { return go_travel_time_seg((go_atmo *)m, a, b); }
static double sm_att(void *m, const double *a, const double *b, double f) { return go_atten_seg((go_atmo *)m, a, b, f); }
static double sm_c(void *m, double r, double t, double p) { return go_c((go_atmo *)m, r, t, p); }
static double sm_amp(void *m, const go_src *s, const double *y) { return go_amp((go_atmo *)m, s, y); }
static double sm_tz(void *m, double t, double p) { (void)t; (void)p; return ground((go_atmo *)m); }
static double sm_tdz(void *m, double t, double p, int n) { (void)m; (void)t; (void)p; (void)n; return 0.0; }

static go_eig_med strat_med(go_atmo *a)
{
    go_eig_med M = {a, a->r0, sm_init, sm_prop, sm_refl, sm_tt, sm_att, sm_c, sm_amp, sm_tz, sm_tdz};
    return M;
}

typedef struct {
    go_atmo *a;
    go_params p6, p18;
    double dth_big, dth_sml, damping, tolerance;
    long long steps;
    const go_eig_med *med;    /* the medium's operations */
} eig_ctx;

/* Work counters of the last go_eig_search_3d call (see go_eig_counts). */
static long long eig_cnt[GO_EIG_NCOUNT];

/* Decision log of the 3D search for diagnosis (GO_EIG_LOG=<path>, appended):
 * "E nb seq th_est ph_est" per accepted estimate, "L nb seq n theta phi dr act"
 * per Levenberg-Marquardt iteration (act: b break, a accept, h halve, s step),
 * "R x y" at the start of a receiver's search; angles in radians. */
static FILE *eig_log(void)
{
    static int tried = 0;
    static FILE *f = NULL;
    if (!tried) {
        tried = 1;
        const char *path = getenv("GO_EIG_LOG");
        if (path && *path) f = fopen(path, "a");
        if (f) setvbuf(f, NULL, _IOLBF, 0);
    }
    return f;
}
static int eig_seq = 0;
/* Trajectory of the last find_eigenray started by go_eig_find_3d (theta,
 * phi, dr per iteration; dr NaN for a break), for conditioning probes. */
static double *eig_traj = NULL;
static int eig_traj_cap = 0, eig_traj_n = 0;

static double mod_dth(const eig_ctx *E, double dr, double dr_dtheta)   /* eigenray.3d.cpp:32 */
{
    return E->dth_big - (E->dth_big - E->dth_sml) * exp(-1.0 / 2.0 * sq(dr / dr_dtheta));
}

/* Trace with up to nb reflections; returns k, rows hold the last leg. */
static int trace_legs(eig_ctx *E, const go_params *p, go_src *s, double *rows, int length, int nb, int *brk,
                      int single_reflection)
{
    const go_eig_med *M = E->med;
    M->set_initial(M->m, p, s, rows);
    int k = M->prop_rk4(M->m, p, s, rows, length, brk);
    E->steps += k;
    if (!*brk) {
        for (int b = 1; b <= nb; b++) {
            /* the reference reads row k-2 < 0 here (undefined); defined as a break */
            if (k < 2) { *brk = 1; break; }
            M->set_refl(M->m, p, s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
            k = M->prop_rk4(M->m, p, s, rows, length, brk);
            E->steps += k;
            if (*brk) break;
            if (single_reflection) break;   /* eigenray.3d.cpp:206-220: unconditional break */
        }
    }
    return k;
}

static int est_eigenray(eig_ctx *E, const double src[3], const double rcvr[2], double th_min, double th_max,
                        double *th_est, double *ph_est, double *th_next, int bncs, double az_err_lim)
{
    const go_params *p = &E->p6;
    double rcvr_rng = sqrt(sq(rcvr[0] - src[0]) + sq(rcvr[1] - src[1]));
    double rcvr_az = atan2(rcvr[1] - src[1], rcvr[0] - src[0]);
    int iterations = 0, k, length = (int)(p->s_max / p->ds_min);
    double r, r_prev, dth = E->dth_big, dph = 100.0;
    int brk, th_max_reached = 0;
    double *rows = get_rows(length);
    go_src s;
    s.src[0] = src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    s.phi = rcvr_az;
    eig_cnt[GO_EIG_EST]++;
    while (fabs(dph * (180.0 / GO_PI)) > az_err_lim && iterations < 5) {
        r = rcvr_rng;
        r_prev = rcvr_rng;
        for (double th = th_min; th <= th_max; th += dth) {
            if (th + dth >= th_max) th_max_reached = 1;
            s.theta = th;
            k = trace_legs(E, p, &s, rows, length, bncs, &brk, 0);
            eig_cnt[GO_EIG_SWEEP_TRACES]++;
            if (iterations >= 1) eig_cnt[GO_EIG_SWEEP_CORR]++;
            if (iterations >= 2) eig_cnt[GO_EIG_SWEEP_ADAPT]++;
            double *yk = rows + 18 * k;
            r = sqrt(sq(yk[0] - src[0]) + sq(yk[1] - src[1]));
            if (brk) r = p->rng_max;
            if (((r - rcvr_rng) * (r_prev - rcvr_rng) <= 0.0) && (r < p->rng_max) && (r_prev < p->rng_max) && (th > th_min)) {
                if (iterations == 0) *th_next = th;
                dph = atan2(yk[1] - src[1], yk[0] - src[0]) * (180.0 / GO_PI);
                dph -= atan2(rcvr[1] - src[1], rcvr[0] - src[0]) * (180.0 / GO_PI);
                while (dph > 180.0) dph -= 360.0;
                while (dph < -180.0) dph += 360.0;
                if (fabs(dph) < az_err_lim) {
                    *th_est = s.theta - dth;
                    *ph_est = s.phi;
                    return 1;
                }
                s.phi -= dph * (GO_PI / 180.0) * 0.9;
                th_min = (th - 5.0 * (GO_PI / 180.0)) > th_min ? (th - 5.0 * (GO_PI / 180.0)) : th_min;
                iterations++;
                break;
            }
            if (iterations >= 2) dth = mod_dth(E, r - rcvr_rng, (r - r_prev) / (2.0 * dth));
            r_prev = r;
        }
        if (th_max_reached) { *th_next = th_max; break; }
    }
    return 0;
}

/* A 3D path row (x, y, z floored at the terrain) into a path sink. */
static void sink_pt_3d(path_sink *k, int leg, int m, const go_eig_med *M, const double *y, double tl, double att,
                       double tt)
{
    if (k->n < k->cap) {
        go_path_pt *P = k->pts + k->n;
        const double zg = M->topo_z(M->m, y[0], y[1]);
        P->launch_idx = 0; P->leg = leg; P->step = m; P->kind = 0;
        P->pos0 = y[0];
        P->pos1 = y[1];
        P->z = y[2] > zg ? y[2] : zg;     /* max(z, topo::z(x, y)) */
        P->tl = tl; P->absorption = -2.0 * att; P->time = tt;
    }
    k->n++;
}

/* The accepted eigenray's row (find_eigenray, geoac.eigenray.3d.cpp:223-276):
 * a re-trace of launch (s->theta, s->phi) with per-step bookkeeping of the
 * segments 1..k-1 of each leg over bnc_cnt reflections, the row taken at
 * row k of the last leg.  Returns the RK4 steps taken. */
static long long eig_accept_3d(const go_eig_med *M, const go_params *p, go_src *s, const double src[3], int bnc_cnt,
                               double freq, double *rows, int length, go_arrival *A, path_sink *ps)
{
    double att = 0.0, tt = 0.0, zmax = 0.0;
    int brk, k;
    long long steps = 0;
    M->set_initial(M->m, p, s, rows);
    k = M->prop_rk4(M->m, p, s, rows, length, &brk);
    steps += k;
    for (int m = 1; m < k; m++) {
        tt += M->tt_seg(M->m, rows + 18 * (m - 1), rows + 18 * m);
        att += M->att_seg(M->m, rows + 18 * (m - 1), rows + 18 * m, freq);
        zmax = zmax > rows[18 * m + 2] ? zmax : rows[18 * m + 2];
        if (ps && (m == 1 || m % 15 == 0))     /* 3d.cpp:240-247 */
            sink_pt_3d(ps, 0, m, M, rows + 18 * m, 20.0 * log10(M->amp(M->m, s, rows + 18 * m)), att, tt);
    }
    for (int b = 1; b <= bnc_cnt; b++) {
        if (k < 2) break;   /* row k-2 < 0: no further reflection */
        M->set_refl(M->m, p, s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
        k = M->prop_rk4(M->m, p, s, rows, length, &brk);
        steps += k;
        for (int m = 1; m < k; m++) {
            tt += M->tt_seg(M->m, rows + 18 * (m - 1), rows + 18 * m);
            att += M->att_seg(M->m, rows + 18 * (m - 1), rows + 18 * m, freq);
            zmax = zmax > rows[18 * m + 2] ? zmax : rows[18 * m + 2];
            /* 3d.cpp:258-265: z printed without the ground floor, which rows
             * 1..k-1 (above the ground by the ground check) never reach */
            if (ps && (m == 1 || m % 15 == 0))
                sink_pt_3d(ps, b, m, M, rows + 18 * m, 20.0 * log10(M->amp(M->m, s, rows + 18 * m)), att, tt);
        }
    }
    const double *yk = rows + 18 * k;
    memset(A, 0, sizeof(*A));
    A->n_bounce = bnc_cnt;
    A->incl_launch = s->theta * (180.0 / GO_PI);
    A->az_launch = 90.0 - s->phi * (180.0 / GO_PI);
    A->pos0 = yk[0];
    A->pos1 = yk[1];
    A->time = tt;
    A->cel = sqrt(sq(yk[0] - src[0]) + sq(yk[1] - src[1])) / tt;
    A->turn_ht = zmax;
    A->incl_arr = -asin(M->c(M->m, yk[0], yk[1], yk[2]) / M->c(M->m, src[0], src[1], src[2]) * yk[5]) * (180.0 / GO_PI);
    double baz = 90.0 - atan2(-yk[4], -yk[3]) * (180.0 / GO_PI);
    while (baz > 180.0) baz -= 360.0;
    while (baz < -180.0) baz += 360.0;
    A->back_az = baz;
    A->tl = 20.0 * log10(M->amp(M->m, s, yk));
    A->absorption = -2.0 * att;
    return steps;
}

/* The accepted-eigenray row for given launch angles (degrees; azimuth
 * clockwise from north as in the output rows): a pure function of the
 * launch, so that a search's rows can be checked one by one. */
int go_eig_row_3d(go_atmo *a, const go_params *p, const double src[3], double incl_deg, double az_deg,
                  int bnc_cnt, double freq, go_arrival *A)
{
    int length = (int)(p->s_max / p->ds_min);
    double *rows = get_rows(length);
    go_src s;
    s.src[0] = src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    s.theta = incl_deg * (GO_PI / 180.0);
    s.phi = (90.0 - az_deg) * (GO_PI / 180.0);
    go_eig_med M = strat_med(a);
    eig_accept_3d(&M, p, &s, src, bnc_cnt, freq, rows, length, A, NULL);
    return 0;
}

/* The arrival's ground-point derivatives d(x, y)/d(theta, phi) that
 * find_eigenray's update uses (geoac.eigenray.3d.cpp:334-348), in long
 * double: the auxiliary parameters at row K moved along the ray to the
 * ground, over flat ground or (is_topo) the terrain's tangent plane.
 * J = {dx/dth, dy/dth, dx/dph, dy/dph}. */
static void lm_jac_3d(const go_eig_med *M, const double src[3], const double *yk, long double J[4])
{
    long double c_src, c_grnd, dzg_dx, dzg_dy, ds_dth, ds_dph;
    c_src = M->c(M->m, src[0], src[1], src[2]);
    c_grnd = M->c(M->m, yk[0], yk[1], M->topo_z(M->m, yk[0], yk[1]));
    if (M->is_topo) {
        dzg_dx = M->topo_dz(M->m, yk[0], yk[1], 0);
        dzg_dy = M->topo_dz(M->m, yk[0], yk[1], 1);
        ds_dth = -c_src / c_grnd * (yk[8] - dzg_dx * yk[6] - dzg_dy * yk[7]) / (yk[5] - dzg_dx * yk[3] - dzg_dy * yk[4]);
        ds_dph = -c_src / c_grnd * (yk[14] - dzg_dx * yk[12] - dzg_dy * yk[13]) / (yk[5] - dzg_dx * yk[3] - dzg_dy * yk[4]);
    } else {
        ds_dth = -c_src / c_grnd * yk[8] / yk[5];
        ds_dph = -c_src / c_grnd * yk[14] / yk[5];
    }
    J[0] = yk[6] + yk[3] * ds_dth;
    J[1] = yk[7] + yk[4] * ds_dth;
    J[2] = yk[12] + yk[3] * ds_dph;
    J[3] = yk[13] + yk[4] * ds_dph;
}

/* lm_jac_3d in doubles, for tests. */
void go_eig_lm_jac_3d(const go_eig_med *M, const double src[3], const double *yk, double J[4])
{
    long double L[4];
    lm_jac_3d(M, src, yk, L);
    for (int i = 0; i < 4; i++) J[i] = (double)L[i];
}

/* One leg of an 18-equation trace over a medium from (th, ph): rows k-1 and
 * k of the first leg (k: the first row below the ground).  Returns k, or -1
 * when the ray leaves the region. */
int go_eig_trace_leg_3d(const go_eig_med *M, const go_params *p, const double src[3], double th, double ph,
                        double y_km1[18], double y_k[18])
{
    go_params p18 = *p;
    p18.calc_amp = 1;
    p18.eq_cnt = 18;
    int length = (int)(p18.s_max / p18.ds_min), brk;
    double *rows = get_rows(length);
    go_src s;
    s.src[0] = src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    s.theta = th;
    s.phi = ph;
    M->set_initial(M->m, &p18, &s, rows);
    int k = M->prop_rk4(M->m, &p18, &s, rows, length, &brk);
    if (brk || k < 1) return -1;
    memcpy(y_km1, rows + 18 * (k - 1), 18 * sizeof(double));
    memcpy(y_k, rows + 18 * k, 18 * sizeof(double));
    return k;
}

static int find_eigenray(eig_ctx *E, const double src[3], const double rcvr[2], double th_est, double ph_est,
                         double freq, int bnc_cnt, int iterate_limit, go_arrival *A)
{
    const go_params *p = &E->p18;
    const go_eig_med *M = E->med;
    int brk, k, length = (int)(p->s_max / p->ds_min);
    double dr, dr_prev = 10000.0, step_sc = 1.0;
    long double x, y, dx, dy, dx_dth, dy_dth, dx_dph, dy_dph;
    long double det, dth = 0.0L, dph = 0.0L;     /* uninitialised in the reference */
    double *rows = get_rows(length);
    go_src s;
    s.src[0] = src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    s.theta = th_est;
    s.phi = ph_est;
    for (int n = 0; n <= iterate_limit; n++) {
        if (n == iterate_limit) break;
        M->set_initial(M->m, p, &s, rows);
        k = M->prop_rk4(M->m, p, &s, rows, length, &brk);
        E->steps += k;
        eig_cnt[GO_EIG_LM_TRACES]++;
        if (!brk && bnc_cnt > 0 && k < 2) brk = 1;   /* row k-2 < 0: defined as a break */
        if (brk) {
            s.theta -= dth * step_sc; s.phi -= dph * step_sc; step_sc /= 2.0;
        }
        if (!brk) {
            for (int b = 1; b <= bnc_cnt; b++) {
                M->set_refl(M->m, p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
                k = M->prop_rk4(M->m, p, &s, rows, length, &brk);
                E->steps += k;
                if (brk) { s.theta -= dth * step_sc; s.phi -= dph * step_sc; step_sc /= 2.0; }
                break;
            }
        }
        if (brk && eig_log()) fprintf(eig_log(), "L %d %d %d %.17g %.17g nan b\n", bnc_cnt, eig_seq, n, s.theta, s.phi);
        if (eig_traj && eig_traj_n < eig_traj_cap) {
            double *t = eig_traj + 3 * eig_traj_n++;
            t[0] = s.theta; t[1] = s.phi; t[2] = NAN;
            if (!brk) {
                const double *yk = rows + 18 * k;
                t[2] = sqrt(sq(rcvr[0] - yk[0]) + sq(rcvr[1] - yk[1]));
            }
        }
        if (!brk) {
            double *yk = rows + 18 * k;
            x = yk[0]; dx = rcvr[0] - x;
            y = yk[1]; dy = rcvr[1] - y;
            dr = (double)sqrtl(powl(dx, 2) + powl(dy, 2));
            FILE *lg = eig_log();
            if (lg) fprintf(lg, "L %d %d %d %.17g %.17g %.17g %c\n", bnc_cnt, eig_seq, n, s.theta, s.phi, dr,
                            dr < E->tolerance ? 'a' : (n > 0 && dr > dr_prev) ? 'h' : 's');
            if (dr < E->tolerance) {
                E->steps += eig_accept_3d(M, p, &s, src, bnc_cnt, freq, rows, length, A, NULL);
                eig_cnt[GO_EIG_FINAL_TRACES]++;
                return 1;
            } else if (n > 0 && dr > dr_prev) {
                s.theta -= dth * step_sc; s.phi -= dph * step_sc; step_sc /= 2.0;
            } else {
                step_sc = (1.0 < step_sc * 1.25) ? 1.0 : step_sc * 1.25;
                long double J[4];
                lm_jac_3d(M, src, yk, J);
                dx_dth = J[0]; dy_dth = J[1]; dx_dph = J[2]; dy_dph = J[3];
                det = pow(1.0 + E->damping, 2) * (dx_dth * dy_dph - dx_dph * dy_dth);
                dth = ((1.0 + E->damping) * dy_dph * dx - dx_dph * dy) / det;
                dph = ((1.0 + E->damping) * dx_dth * dy - dy_dth * dx) / det;
                s.theta += dth * step_sc;
                s.phi += dph * step_sc;
                dr_prev = dr;
            }
        }
        if (sqrtl(dth * dth + dph * dph) * step_sc < 1.0e-12) break;
    }
    return 0;
}

/* find_eigenray alone from a launch estimate in radians (the search's
 * starting points, for sensitivity probes): 1 and the row in *A when it
 * accepts, 0 otherwise. */
int go_eig_find_3d_med(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                       double th_est, double ph_est, int bnc, int iterations, double damping, double tolerance,
                       double freq, go_arrival *A, long long *ray_steps)
{
    eig_ctx E;
    memset(&E, 0, sizeof(E));
    E.med = M;
    E.p6 = *p; E.p6.calc_amp = 0; E.p6.eq_cnt = 6;
    E.p18 = *p; E.p18.calc_amp = 1; E.p18.eq_cnt = 18;
    E.damping = damping;
    E.tolerance = tolerance;
    E.steps = 0;
    int ok = find_eigenray(&E, src, rcvr, th_est, ph_est, freq, bnc, iterations, A);
    if (ray_steps) *ray_steps = E.steps;
    return ok;
}

int go_eig_find_3d(go_atmo *a, const go_params *p, const double src_in[3], const double rcvr[2], double th_est,
                   double ph_est, int bnc, int iterations, double damping, double tolerance, double freq,
                   go_arrival *A, long long *ray_steps)
{
    go_eig_med M = strat_med(a);
    double src[3] = {src_in[0], src_in[1], src_in[2] > ground(a) ? src_in[2] : ground(a)};
    return go_eig_find_3d_med(&M, p, src, rcvr, th_est, ph_est, bnc, iterations, damping, tolerance, freq, A,
                              ray_steps);
}

/* go_eig_find_3d that also records the iterates (theta, phi, dr) into traj
 * (up to cap rows); *n_traj receives the number recorded. */
int go_eig_find_3d_traj(go_atmo *a, const go_params *p, const double src_in[3], const double rcvr[2], double th_est,
                        double ph_est, int bnc, int iterations, double damping, double tolerance, double freq,
                        go_arrival *A, double *traj, int cap, int *n_traj)
{
    eig_traj = traj;
    eig_traj_cap = cap;
    eig_traj_n = 0;
    int ok = go_eig_find_3d(a, p, src_in, rcvr, th_est, ph_est, bnc, iterations, damping, tolerance, freq, A, NULL);
    *n_traj = eig_traj_n;
    eig_traj = NULL;
    eig_traj_cap = 0;
    return ok;
}

int go_eig_search_3d_med(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                         double th_min, double th_max, int bnc_min, int bnc_max, int iterations, double damping,
                         double tolerance, double az_err_lim, double dth_big_deg, double dth_sml_deg, double freq,
                         go_arrival *out, int out_cap, long long *ray_steps)
{
    eig_ctx E;
    memset(&E, 0, sizeof(E));
    E.med = M;
    E.p6 = *p; E.p6.calc_amp = 0; E.p6.eq_cnt = 6;
    E.p18 = *p; E.p18.calc_amp = 1; E.p18.eq_cnt = 18;
    E.dth_big = dth_big_deg * (GO_PI / 180.0);
    E.dth_sml = dth_sml_deg * (GO_PI / 180.0);
    E.damping = damping;
    E.tolerance = tolerance;
    E.steps = 0;
    memset(eig_cnt, 0, sizeof(eig_cnt));
    int cnt = 0;
    if (eig_log()) fprintf(eig_log(), "R %.17g %.17g\n", rcvr[0], rcvr[1]);
    for (int nb = bnc_min; nb <= bnc_max; nb++) {
        double th_start = th_min, th_next = th_min, th_est = 0, ph_est = 0;
        eig_seq = 0;
        while (th_start < th_max) {
            int ok = est_eigenray(&E, src, rcvr, th_start, th_max, &th_est, &ph_est, &th_next, nb, az_err_lim);
            if (ok) {
                go_arrival A;
                eig_cnt[GO_EIG_FINDS]++;
                if (eig_log()) fprintf(eig_log(), "E %d %d %.17g %.17g\n", nb, eig_seq, th_est, ph_est);
                if (find_eigenray(&E, src, rcvr, th_est, ph_est, freq, nb, iterations, &A)) {
                    if (cnt < out_cap) out[cnt] = A;
                    cnt++;
                }
                eig_seq++;
            }
            th_start = th_next;
        }
    }
    if (ray_steps) *ray_steps = E.steps;
    eig_cnt[GO_EIG_FOUND] = cnt;
    return cnt;
}

int go_eig_search_3d(go_atmo *a, const go_params *p, const double src_in[3],
                     const double rcvr[2], double th_min_deg, double th_max_deg,
                     int bnc_min, int bnc_max, int iterations, double damping,
                     double tolerance, double az_err_lim, double dth_big_deg,
                     double dth_sml_deg, double freq,
                     go_arrival *out, int out_cap, long long *ray_steps)
{
    go_eig_med M = strat_med(a);
    double src[3] = {src_in[0], src_in[1], src_in[2] > ground(a) ? src_in[2] : ground(a)};
    return go_eig_search_3d_med(&M, p, src, rcvr, th_min_deg * (GO_PI / 180.0), th_max_deg * (GO_PI / 180.0),
                                bnc_min, bnc_max, iterations, damping, tolerance, az_err_lim, dth_big_deg,
                                dth_sml_deg, freq, out, out_cap, ray_steps);
}

long long go_eig_path_3d_med(const go_eig_med *M, const go_params *p, const double src[3], double th, double ph,
                             int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A)
{
    go_params p18 = *p;
    p18.calc_amp = 1;
    p18.eq_cnt = 18;
    int length = (int)(p18.s_max / p18.ds_min);
    double *rows = get_rows(length);
    path_sink sink = {pts, cap, 0};
    go_src s;
    s.src[0] = src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    s.theta = th;
    s.phi = ph;
    eig_accept_3d(M, &p18, &s, src, bnc_cnt, freq, rows, length, A, &sink);
    return sink.n;
}

/* ======================================================================
 * Eigenray search, spherical: geoac.eigenray.sph.cpp:36-380 and the driver
 * loop of main.infraga.sph.cpp:962-989.  src = (z [km], lat, lon [rad]) and
 * rcvr = (lat, lon [rad]) as run_eig_search holds them.
 * ==================================================================== */
static double go_bearing(double lat1, double lon1, double lat2, double lon2)   /* globe.cpp:16-21 */
{
    double term1 = sin(lon2 - lon1);
    double term2 = cos(lat1) * tan(lat2) - sin(lat1) * cos(lon2 - lon1);
    return atan2(term1, term2);
}

/* trace_legs over the search's medium operations (spherical searches). */
static int trace_legs_m(eig_ctx *E, const go_params *p, go_src *s, double *rows, int length, int nb, int *brk)
{
    const go_eig_med *M = E->med;
    M->set_initial(M->m, p, s, rows);
    int k = M->prop_rk4(M->m, p, s, rows, length, brk);
    E->steps += k;
    if (!*brk) {
        for (int b = 1; b <= nb; b++) {
            if (k < 2) { *brk = 1; break; }
            M->set_refl(M->m, p, s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
            k = M->prop_rk4(M->m, p, s, rows, length, brk);
            E->steps += k;
            if (*brk) break;
        }
    }
    return k;
}

static int est_eigenray_sph(eig_ctx *E, const double src[3], const double rcvr[2], double th_min, double th_max,
                            double *th_est, double *ph_est, double *th_next, int bncs, double az_err_lim)
{
    const go_params *p = &E->p6;
    const double r0 = E->med->r0;
    double rcvr_rng = go_gc_dist(r0, src[1], src[2], rcvr[0], rcvr[1]);
    double rcvr_az = go_bearing(src[1], src[2], rcvr[0], rcvr[1]);
    int iterations = 0, k, length = (int)(p->s_max / p->ds_min);
    double r, r_prev, dth = E->dth_big, dph = 100.0;
    int brk, th_max_reached = 0;
    double *rows = get_rows(length);
    go_src s;
    s.src[0] = r0 + src[0]; s.src[1] = src[1]; s.src[2] = src[2];   /* set_initial, eqset.sph :59-62 */
    s.phi = GO_PI / 2.0 - rcvr_az;
    eig_cnt[GO_EIG_EST]++;
    while (fabs(dph * (180.0 / GO_PI)) > az_err_lim && iterations < 5) {
        r = rcvr_rng;
        r_prev = rcvr_rng;
        for (double th = th_min; th < th_max; th += dth) {          /* sph: strict bound */
            if (th + dth >= th_max) th_max_reached = 1;
            s.theta = th;
            k = trace_legs_m(E, p, &s, rows, length, bncs, &brk);
            eig_cnt[GO_EIG_SWEEP_TRACES]++;
            if (iterations >= 1) eig_cnt[GO_EIG_SWEEP_CORR]++;
            if (iterations >= 3) eig_cnt[GO_EIG_SWEEP_ADAPT]++;
            double *yk = rows + 18 * k;
            r = go_gc_dist(r0, src[1], src[2], yk[1], yk[2]);
            if (brk) r = p->rng_max;
            if (((r - rcvr_rng) * (r_prev - rcvr_rng) <= 0.0) && (r < p->rng_max) && (r_prev < p->rng_max) && (th > th_min)) {
                if (iterations == 0) *th_next = s.theta;
                dph = go_bearing(src[1], src[2], rcvr[0], rcvr[1]) * (180.0 / GO_PI);
                dph -= go_bearing(src[1], src[2], yk[1], yk[2]) * (180.0 / GO_PI);
                while (dph > 180.0) dph -= 360.0;
                while (dph < -180.0) dph += 360.0;
                if (fabs(dph) < az_err_lim) {
                    *th_est = s.theta - dth;
                    *ph_est = s.phi;
                    return 1;
                }
                s.phi -= dph * (GO_PI / 180.0) * 0.9;
                th_min = (s.theta - 5.0 * (GO_PI / 180.0)) > th_min ? (s.theta - 5.0 * (GO_PI / 180.0)) : th_min;
                iterations++;
                break;
            }
            if (iterations >= 3) dth = mod_dth(E, r - rcvr_rng, (r - r_prev) / (2.0 * dth));   /* sph: >= 3 */
            r_prev = r;
        }
        if (th_max_reached) { *th_next = th_max; break; }
    }
    return 0;
}

/* A spherical path row (lat, lon [deg], r - r0) into a path sink. */
static void sink_pt_sph(path_sink *k, int leg, int m, double r0, const double *y, double tl, double att, double tt)
{
    if (k->n < k->cap) {
        go_path_pt *P = k->pts + k->n;
        P->launch_idx = 0; P->leg = leg; P->step = m; P->kind = 0;
        P->pos0 = y[1] * (180.0 / GO_PI);
        P->pos1 = y[2] * (180.0 / GO_PI);
        P->z = y[0] - r0;
        P->tl = tl; P->absorption = -2.0 * att; P->time = tt;
    }
    k->n++;
}

/* The accepted spherical eigenray's row: re-trace with per-step bookkeeping
 * over rows 1..k-1 of every leg (geoac.eigenray.sph.cpp:219-271), path rows
 * m == 1 or m % 15 == 0 into ps when given.  src = (alt [km], lat, lon
 * [rad]).  Returns the RK4 steps taken. */
static long long eig_accept_sph(const go_eig_med *M, const go_params *p, go_src *s, const double src[3],
                                int bnc_cnt, double freq, double *rows, int length, go_arrival *A, path_sink *ps)
{
    const double r0 = M->r0;
    double att = 0.0, tt = 0.0, rmax = 0.0;
    int brk, k;
    long long steps = 0;
    M->set_initial(M->m, p, s, rows);
    k = M->prop_rk4(M->m, p, s, rows, length, &brk);
    steps += k;
    for (int b = 0; b <= bnc_cnt; b++) {
        if (b > 0) {
            if (k < 2) break;   /* row k-2 < 0: no further reflection */
            M->set_refl(M->m, p, s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
            k = M->prop_rk4(M->m, p, s, rows, length, &brk);
            steps += k;
        }
        for (int m = 1; m < k; m++) {
            tt += M->tt_seg(M->m, rows + 18 * (m - 1), rows + 18 * m);
            att += M->att_seg(M->m, rows + 18 * (m - 1), rows + 18 * m, freq);
            rmax = rmax > rows[18 * m] - r0 ? rmax : rows[18 * m] - r0;
            if (ps && (m == 1 || m % 15 == 0))     /* sph.cpp:250-257, 266-273 */
                sink_pt_sph(ps, b, m, r0, rows + 18 * m, 20.0 * log10(M->amp(M->m, s, rows + 18 * m)), att, tt);
        }
    }
    const double *yk = rows + 18 * k;
    memset(A, 0, sizeof(*A));
    A->n_bounce = bnc_cnt;
    A->incl_launch = s->theta * (180.0 / GO_PI);
    A->az_launch = 90.0 - s->phi * (180.0 / GO_PI);
    A->pos0 = yk[1] * 180.0 / GO_PI;
    A->pos1 = yk[2] * 180.0 / GO_PI;
    A->time = tt;
    A->cel = go_gc_dist(r0, yk[1], yk[2], src[1], src[2]) / tt;
    A->turn_ht = rmax;
    /* atmo::c(src[0], ...) takes the source altitude as a radius:
     * the profile's bottom value (sph.cpp:273) */
    A->incl_arr = -asin(M->c(M->m, yk[0], yk[1], yk[2]) / M->c(M->m, src[0], src[1], src[2]) * yk[3]) * 180.0 / GO_PI;
    double baz = 90.0 - atan2(-yk[4], -yk[5]) * (180.0 / GO_PI);
    while (baz < -180.0) baz += 360.0;
    while (baz > 180.0) baz -= 360.0;
    A->back_az = baz;
    A->tl = 20.0 * log10(M->amp(M->m, s, yk));
    A->absorption = -2.0 * att;
    return steps;
}

long long go_eig_path(go_atmo *a, const go_params *p, const double src_in[3], double incl_deg, double az_deg,
                      int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A)
{
    go_params p18 = *p;
    p18.calc_amp = 1;
    p18.eq_cnt = 18;
    int length = (int)(p18.s_max / p18.ds_min);
    double *rows = get_rows(length);
    path_sink sink = {pts, cap, 0};
    go_src s;
    s.theta = incl_deg * (GO_PI / 180.0);
    s.phi = (90.0 - az_deg) * (GO_PI / 180.0);
    if (a->geom == GA_GEOM_SPH) {
        double src[3];
        src[1] = src_in[0] * (GO_PI / 180.0);
        src[2] = src_in[1] * (GO_PI / 180.0);
        src[0] = src_in[2] > ground(a) - a->r0 ? src_in[2] : ground(a) - a->r0;
        s.src[0] = a->r0 + src[0]; s.src[1] = src[1]; s.src[2] = src[2];
        go_eig_med M = strat_med(a);
        eig_accept_sph(&M, &p18, &s, src, bnc_cnt, freq, rows, length, A, &sink);
    } else {
        double src[3] = {src_in[0], src_in[1], src_in[2] > ground(a) ? src_in[2] : ground(a)};
        s.src[0] = src[0]; s.src[1] = src[1]; s.src[2] = src[2];
        go_eig_med M = strat_med(a);
        eig_accept_3d(&M, &p18, &s, src, bnc_cnt, freq, rows, length, A, &sink);
    }
    return sink.n;
}

static int find_eigenray_sph(eig_ctx *E, const double src[3], const double rcvr[2], double th_est, double ph_est,
                             double freq, int bnc_cnt, int iterate_limit, go_arrival *A)
{
    const go_params *p = &E->p18;
    const go_eig_med *M = E->med;
    const double r0 = M->r0;
    int brk, k, length = (int)(p->s_max / p->ds_min);
    double dr, dr_prev = 10000.0, step_sc = 1.0;
    long double r_grnd, c_src, c_grnd, dzg_dlat, dzg_dlon, ds_norm, ds_dth, ds_dph, dlat, dlon;
    long double dlat_dth, dlon_dth, dlat_dph, dlon_dph;
    long double det, dth = 0.0L, dph = 0.0L;     /* uninitialised in the reference */
    double *rows = get_rows(length);
    go_src s;
    s.src[0] = r0 + src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    s.theta = th_est;
    s.phi = ph_est;
    for (int n = 0; n <= iterate_limit; n++) {
        if (n == iterate_limit) break;
        k = trace_legs_m(E, p, &s, rows, length, bnc_cnt, &brk);    /* eigenray.sph.cpp:178-201 */
        eig_cnt[GO_EIG_LM_TRACES]++;
        if (brk) { s.theta -= dth * step_sc; s.phi -= dph * step_sc; step_sc /= 2.0; }
        if (!brk) {
            double *yk = rows + 18 * k;
            dr = go_gc_dist(r0, yk[1], yk[2], rcvr[0], rcvr[1]);
            if (dr < E->tolerance) {
                E->steps += eig_accept_sph(M, p, &s, src, bnc_cnt, freq, rows, length, A, NULL);
                eig_cnt[GO_EIG_FINAL_TRACES]++;
                return 1;
            } else if (n > 0 && dr > dr_prev) {
                s.theta -= dth * step_sc; s.phi -= dph * step_sc; step_sc /= 2.0;
            } else {
                step_sc = (1.0 < step_sc * 1.25) ? 1.0 : step_sc * 1.25;
                dlat = rcvr[0] - yk[1];
                dlon = rcvr[1] - yk[2];
                r_grnd = M->topo_z(M->m, yk[1], yk[2]);            /* topo::z (a radius) */
                c_grnd = M->c(M->m, r_grnd, yk[1], yk[2]);
                c_src = M->c(M->m, src[0] + r0, src[1], src[2]);
                dzg_dlat = M->topo_dz(M->m, yk[1], yk[2], 0) / r_grnd;
                dzg_dlon = M->topo_dz(M->m, yk[1], yk[2], 1) / (r_grnd * cos(yk[1]));
                ds_norm = yk[3] - dzg_dlat * yk[4] - dzg_dlon * yk[5];
                ds_dth = -c_src / c_grnd * (yk[6] - dzg_dlat * r_grnd * yk[7] - dzg_dlon * (r_grnd * cos(yk[1])) * yk[8]) / ds_norm;
                ds_dph = -c_src / c_grnd * (yk[12] - dzg_dlat * r_grnd * yk[13] - dzg_dlon * (r_grnd * cos(yk[1])) * yk[14]) / ds_norm;
                dlat_dth = yk[7] + yk[4] / r_grnd * ds_dth;
                dlon_dth = yk[8] + yk[5] / (r_grnd * cos(yk[1])) * ds_dth;
                dlat_dph = yk[13] + yk[4] / r_grnd * ds_dph;
                dlon_dph = yk[14] + yk[5] / (r_grnd * cos(yk[1])) * ds_dph;
                /* sph.cpp:351: (1 + damping)^2 multiplies the first product only */
                det = pow(1.0 + E->damping, 2) * dlat_dth * dlon_dph - dlat_dph * dlon_dth;
                dth = ((1.0 + E->damping) * dlon_dph * dlat - dlat_dph * dlon) / det;
                dph = ((1.0 + E->damping) * dlat_dth * dlon - dlon_dth * dlat) / det;
                s.theta += dth * step_sc;
                s.phi += dph * step_sc;
                dr_prev = dr;
            }
        }
        if (sqrtl(dth * dth + dph * dph) * step_sc < 1.0e-12) break;
    }
    return 0;
}

/* The spherical search of run_eig_search (main.infraga.sph.cpp:962-989) over
 * a medium: src = (alt [km], lat, lon [rad]) raised to the ground by the
 * caller, rcvr [rad], th_grnd the lowest inclination [rad] (or whatever the
 * caller's driver computes there). */
static int eig_search_sph_core(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                               double th_min, double th_max, double th_grnd, int bnc_min, int bnc_max,
                               int iterations, double damping, double tolerance, double az_err_lim,
                               double dth_big_deg, double dth_sml_deg, double freq,
                               go_arrival *out, int out_cap, long long *ray_steps)
{
    eig_ctx E;
    E.a = NULL;
    E.med = M;
    E.p6 = *p; E.p6.calc_amp = 0; E.p6.eq_cnt = 6;
    E.p18 = *p; E.p18.calc_amp = 1; E.p18.eq_cnt = 18;
    E.dth_big = dth_big_deg * (GO_PI / 180.0);
    E.dth_sml = dth_sml_deg * (GO_PI / 180.0);
    E.damping = damping;
    E.tolerance = tolerance;
    E.steps = 0;
    memset(eig_cnt, 0, sizeof(eig_cnt));
    int cnt = 0;
    for (int nb = bnc_min; nb <= bnc_max; nb++) {
        double th_start = th_min > th_grnd ? th_min : th_grnd;
        double th_next = th_start, th_est = 0, ph_est = 0;
        while (th_start < th_max) {
            int ok = est_eigenray_sph(&E, src, rcvr, th_start, th_max, &th_est, &ph_est, &th_next, nb, az_err_lim);
            if (ok) {
                go_arrival A;
                eig_cnt[GO_EIG_FINDS]++;
                if (find_eigenray_sph(&E, src, rcvr, th_est, ph_est, freq, nb, iterations, &A)) {
                    if (cnt < out_cap) out[cnt] = A;
                    cnt++;
                }
            }
            th_start = th_next;
        }
    }
    if (ray_steps) *ray_steps = E.steps;
    eig_cnt[GO_EIG_FOUND] = cnt;
    return cnt;
}

int go_eig_search_sph(go_atmo *a, const go_params *p, const double src_in[3],
                      const double rcvr_in[2], double th_min_deg, double th_max_deg,
                      int bnc_min, int bnc_max, int iterations, double damping,
                      double tolerance, double az_err_lim, double dth_big_deg,
                      double dth_sml_deg, double freq,
                      go_arrival *out, int out_cap, long long *ray_steps)
{
    /* main.infraga.sph.cpp:906-909 */
    double src[3], rcvr[2];
    src[1] = src_in[0] * (GO_PI / 180.0);
    src[2] = src_in[1] * (GO_PI / 180.0);
    rcvr[0] = rcvr_in[0] * (GO_PI / 180.0);
    rcvr[1] = rcvr_in[1] * (GO_PI / 180.0);
    src[0] = src_in[2] > ground(a) - a->r0 ? src_in[2] : ground(a) - a->r0;
    double th_min = th_min_deg * (GO_PI / 180.0), th_max = th_max_deg * (GO_PI / 180.0);
    /* ground source: theta_grnd = 0.1 deg (sph.cpp:966-979) */
    double th_grnd = (src[0] - (ground(a) - a->r0) < 0.01) ? 0.1 * (GO_PI / 180.0) : -89.9 * (GO_PI / 180.0);
    go_eig_med M = strat_med(a);
    return eig_search_sph_core(&M, p, src, rcvr, th_min, th_max, th_grnd, bnc_min, bnc_max, iterations, damping,
                               tolerance, az_err_lim, dth_big_deg, dth_sml_deg, freq, out, out_cap, ray_steps);
}

int go_eig_search_sph_med(const go_eig_med *M, const go_params *p, const double src[3], const double rcvr[2],
                          double th_min, double th_max, double th_grnd, int bnc_min, int bnc_max, int iterations,
                          double damping, double tolerance, double az_err_lim, double dth_big_deg,
                          double dth_sml_deg, double freq, go_arrival *out, int out_cap, long long *ray_steps)
{
    return eig_search_sph_core(M, p, src, rcvr, th_min, th_max, th_grnd, bnc_min, bnc_max, iterations, damping,
                               tolerance, az_err_lim, dth_big_deg, dth_sml_deg, freq, out, out_cap, ray_steps);
}

long long go_eig_path_sph_med(const go_eig_med *M, const go_params *p, const double src[3], double th, double ph,
                              int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A)
{
    go_params p18 = *p;
    p18.calc_amp = 1;
    p18.eq_cnt = 18;
    int length = (int)(p18.s_max / p18.ds_min);
    double *rows = get_rows(length);
    path_sink sink = {pts, cap, 0};
    go_src s;
    s.theta = th;
    s.phi = ph;
    s.src[0] = M->r0 + src[0]; s.src[1] = src[1]; s.src[2] = src[2];
    eig_accept_sph(M, &p18, &s, src, bnc_cnt, freq, rows, length, A, &sink);
    return sink.n;
}

/* run_eig_direct (main.infraga.3d.cpp:955-1086, main.infraga.sph.cpp:998-1139):
 * find_eigenray from a launch estimate, no sweep. */
int go_eig_direct(go_atmo *a, const go_params *p, const double src_in[3], const double rcvr_in[2],
                  double incl_est_deg, double az_est_deg, int bnc, int iterations, double damping,
                  double tolerance, double freq, go_arrival *A, long long *ray_steps)
{
    eig_ctx E;
    memset(&E, 0, sizeof(E));
    E.a = a;
    E.p6 = *p; E.p6.calc_amp = 0; E.p6.eq_cnt = 6;
    E.p18 = *p; E.p18.calc_amp = 1; E.p18.eq_cnt = 18;
    E.dth_big = E.dth_sml = 0.0;
    E.damping = damping;
    E.tolerance = tolerance;
    E.steps = 0;
    memset(eig_cnt, 0, sizeof(eig_cnt));
    const double th_est = incl_est_deg * (GO_PI / 180.0);
    int ok;
    if (a->geom == GA_GEOM_SPH) {
        double src[3], rcvr[2];
        src[1] = src_in[0] * (GO_PI / 180.0);
        src[2] = src_in[1] * (GO_PI / 180.0);
        rcvr[0] = rcvr_in[0] * (GO_PI / 180.0);
        rcvr[1] = rcvr_in[1] * (GO_PI / 180.0);
        src[0] = src_in[2] > ground(a) - a->r0 ? src_in[2] : ground(a) - a->r0;   /* sph.cpp:1106 */
        double ph_est;
        if (isnan(az_est_deg)) {                                                   /* sph.cpp:1110 */
            double t1 = sin(rcvr[1] - src[2]);
            double t2 = cos(src[1]) * tan(rcvr[0]) - sin(src[1]) * cos(rcvr[1] - src[2]);
            ph_est = GO_PI / 2.0 - atan2(t1, t2);
        } else {
            ph_est = GO_PI / 2.0 - az_est_deg * (GO_PI / 180.0);
        }
        go_eig_med M = strat_med(a);
        E.med = &M;
        ok = find_eigenray_sph(&E, src, rcvr, th_est, ph_est, freq, bnc, iterations, A);
    } else {
        double src[3] = {src_in[0], src_in[1], src_in[2] > ground(a) ? src_in[2] : ground(a)};   /* 3d.cpp:1048 */
        double ph_est = isnan(az_est_deg) ? atan2(rcvr_in[1] - src[1], rcvr_in[0] - src[0])     /* 3d.cpp:1050 */
                                          : GO_PI / 2.0 - az_est_deg * (GO_PI / 180.0);
        go_eig_med M = strat_med(a);
        E.med = &M;
        ok = find_eigenray(&E, src, rcvr_in, th_est, ph_est, freq, bnc, iterations, A);
    }
    if (ray_steps) *ray_steps = E.steps;
    return ok;
}

void go_eig_counts(long long out[GO_EIG_NCOUNT])
{
    for (int i = 0; i < GO_EIG_NCOUNT; i++) out[i] = eig_cnt[i];
}

#endif /* GO_OPCOUNT */

/* ======================================================================
 * Handle helpers for the test bindings (ctypes cannot size go_atmo).
 * ==================================================================== */
void *go_h_atmo(int geom, const double *z, const double *c, const double *u,
                const double *v, const double *rho, int n)
{
    go_atmo *a = (go_atmo *)malloc(sizeof(go_atmo));
    if (go_atmo_init(a, geom, z, c, u, v, rho, n) != 0) { free(a); return NULL; }
    return a;
}
void go_h_free(void *h) { if (h) { go_atmo_free((go_atmo *)h); free(h); } }
void go_h_set_ground(void *h, double z0)
{
    go_atmo *a = (go_atmo *)h;
    a->z0 = z0; a->z_max = z0; a->z_bndlyr = z0 + 2.0;   /* z_grnd= handling, main.infraga.sph.cpp:287-290 */
}
void go_h_set_tweak(void *h, double t) { ((go_atmo *)h)->tweak_abs = t; }
void go_h_set_r0(void *h, double r0) { ((go_atmo *)h)->r0 = r0; }
void go_h_params(void *h, int calc_amp, go_params *p)
{
    go_atmo *a = (go_atmo *)h;
    go_params_default(p, a->geom, a, calc_amp);
}
int go_h_fan_paths(void *h, const go_params *p, const go_prop *pr, const double *launches, int n, int flags,
                   go_arrival *out, int cap, int *n_out, go_path_pt *pts, long long pts_cap, long long *n_pts,
                   long long *steps)
{
    return go_fan_paths((go_atmo *)h, p, pr, launches, n, flags, out, cap, n_out, pts, pts_cap, n_pts, steps);
}

int go_h_fan(void *h, const go_params *p, const go_prop *pr, const double *launches, int n,
             go_arrival *out, int cap, int *n_out, long long *steps)
{
    return go_fan((go_atmo *)h, p, pr, launches, n, out, cap, n_out, steps);
}

void go_h_tables(void *h, int which, double *x, double *f, double *slope)
{
    go_atmo *a = (go_atmo *)h;
    go_spline *sp[4] = {&a->c, &a->u, &a->v, &a->rho};
    for (int i = 0; i < a->n; i++) { x[i] = a->z[i]; f[i] = sp[which]->f[i]; slope[i] = sp[which]->slope[i]; }
}

#ifndef GO_OPCOUNT
/* ======================================================================
 * 2D effective-sound-speed geometry: infraga-2d -prop
 * (geoac.eqset.2d.cpp, atmo_state.2d.cpp, atmo_io.2d.cpp:27-33 set_limits,
 * main.infraga.2d.cpp:127-400).  Equations: r, z, nu_r, nu_z and, with
 * amplitudes, R, Z, mu_r, mu_z; c_eff = c + u cos(phi) + v sin(phi) with no
 * boundary-layer wind scaling.
 * ==================================================================== */
static double g2_topo_z(go_2d *g, double r)           /* atmo_state.2d.cpp:47-53 */
{
    if (!g->is_topo) return g->a->z0;
    return go_spline_f(&g->topo, in_interval(r, g->topo.x[0], g->topo.x[g->topo.n - 1]));
}
static double g2_topo_dz(go_2d *g, double r)          /* :55-61 */
{
    if (!g->is_topo) return 0.0;
    double f, df, ddf;
    go_spline_all(&g->topo, in_interval(r, g->topo.x[0], g->topo.x[g->topo.n - 1]), &f, &df, &ddf);
    return df;
}
static double g2_topo_ddz(go_2d *g, double r)         /* :63-69 */
{
    if (!g->is_topo) return 0.0;
    double f, df, ddf;
    go_spline_all(&g->topo, in_interval(r, g->topo.x[0], g->topo.x[g->topo.n - 1]), &f, &df, &ddf);
    return ddf;
}
static double g2_pt(go_2d *g, go_spline *s, double z)   /* atmo::c/u/v/rho, :75-131 */
{
    go_atmo *a = g->a;
    return go_spline_f(s, in_interval(in_interval(z, a->z[0], a->z[a->n - 1]), a->z[0], a->z[a->n - 1]));
}
static double g2_dpt(go_2d *g, go_spline *s, double z)  /* atmo::dc/du/dv with n = 2 */
{
    go_atmo *a = g->a;
    double f, df, ddf;
    go_spline_all(s, in_interval(in_interval(z, a->z[0], a->z[a->n - 1]), a->z[0], a->z[a->n - 1]), &f, &df, &ddf);
    return df;
}
static double g2_ceff(go_2d *g, double z)
{
    return g2_pt(g, &g->a->c, z) + g2_pt(g, &g->a->u, z) * cos(g->phi) + g2_pt(g, &g->a->v, z) * sin(g->phi);
}

void go2_init(go_2d *g, go_atmo *a)
{
    memset(g, 0, sizeof(*g));
    g->a = a;
    a->z0 = a->z[0];                                  /* set_limits (atmo_io.2d.cpp:27-33) */
    g->alt_max = a->z[a->n - 1];
    g->rng_max = 2000.0;
    g->z_reflect = g->alt_max + 1.0;
    a->z_max = a->z0;                                 /* set_bndlyr without terrain (:34-45) */
    a->z_bndlyr = a->z_max + 2.0;
    g->s_max = 1000.0; g->ds_min = 0.001; g->ds_max = 0.05;   /* geoac.params.2d.cpp */
    g->calc_amp = 1;
}

int go2_set_topo(go_2d *g, const double *r, const double *z, int n)
{
    if (n < 3) return -1;
    g->topo.n = n;
    g->topo.x = (double *)malloc(sizeof(double) * (size_t)n);
    g->topo.f = (double *)malloc(sizeof(double) * (size_t)n);
    g->topo.slope = (double *)malloc(sizeof(double) * (size_t)n);
    g->topo.accel = 0;
    memcpy((double *)g->topo.x, r, sizeof(double) * (size_t)n);
    memcpy(g->topo.f, z, sizeof(double) * (size_t)n);
    go_spline_set(&g->topo);
    g->is_topo = 1;
    g->rng_max = r[n - 1];                            /* set_limits with terrain */
    double zm = 0.0;                                  /* set_bndlyr with terrain */
    for (int i = 0; i < n; i++) zm = zm > z[i] ? zm : z[i];
    g->a->z_max = zm + 0.1;
    g->a->z_bndlyr = g->a->z_max + 2.0;
    return 0;
}

void go2_set_ground(go_2d *g, double z0)            /* main.infraga.2d.cpp z_grnd= */
{
    if (g->is_topo) return;
    g->a->z0 = z0;
    g->a->z_max = z0;
    g->a->z_bndlyr = z0 + 2.0;
}

void go2_free(go_2d *g)
{
    if (g->is_topo) {
        free((double *)g->topo.x);
        free(g->topo.f);
        free(g->topo.slope);
    }
    g->is_topo = 0;
}

/* update_refs + eval_src_eq (geoac.eqset.2d.cpp:176-239) at state y into f. */
static void g2_rhs(go_2d *g, const double *y, double *f)
{
    go_atmo *a = g->a;
    double z_eval = in_interval(y[1], a->z[0], a->z[a->n - 1]);
    double c, dc, ddc, u, du, ddu, v, dv, ddv;
    go_spline_all(&a->c, z_eval, &c, &dc, &ddc);
    go_spline_all(&a->u, z_eval, &u, &du, &ddu);
    go_spline_all(&a->v, z_eval, &v, &dv, &ddv);
    double ce = c + u * cos(g->phi) + v * sin(g->phi);
    double dce = dc + du * cos(g->phi) + dv * sin(g->phi);
    double ddce = ddc + ddu * cos(g->phi) + ddv * sin(g->phi);
    double nu_r = y[2], nu_z = y[3], c0 = g->c_eff0;
    double dc_dr = 0.0, dc_dz = dce;
    f[0] = ce / c0 * nu_r;
    f[1] = ce / c0 * nu_z;
    f[2] = -c0 / pow(ce, 2) * dc_dr;
    f[3] = -c0 / pow(ce, 2) * dc_dz;
    if (g->calc_amp) {
        double R = y[4], Z = y[5], mu_r = y[6], mu_z = y[7];
        double ddc_ddr = 0.0, ddc_ddz = ddce, ddc_drdz = 0.0;
        double dc_dth = dc_dr * R + dc_dz * Z;
        double ddc_drdth = ddc_ddr * R + ddc_drdz * Z;
        double ddc_dzdth = ddc_drdz * R + ddc_ddz * Z;
        f[4] = dc_dth / c0 * nu_r + ce / c0 * mu_r;
        f[5] = dc_dth / c0 * nu_z + ce / c0 * mu_z;
        f[6] = 2.0 * c0 / pow(ce, 3) * dc_dr * dc_dth - c0 / pow(ce, 2) * ddc_drdth;
        f[7] = 2.0 * c0 / pow(ce, 3) * dc_dz * dc_dth - c0 / pow(ce, 2) * ddc_dzdth;
    }
}

static double g2_set_ds(go_2d *g, const double *y)   /* :165-171 */
{
    double result = g->ds_max;
    if (y[1] < g->a->z_bndlyr) result -= (g->ds_max - g->ds_min) * exp(-(y[1] - g2_topo_z(g, y[0])) / 0.2);
    return result;
}

/* rk4solver.cpp:14-85 with the 2D checks (geoac.eqset.2d.cpp:258-281):
 * break r > rng_max or z > alt_max; ground z <= z_max and z < topo::z(r);
 * elevated reflection z > z_reflect. */
int go2_prop_rk4(go_2d *g, double *rows, int length, int *check)
{
    const int n = g->calc_amp ? 8 : 4;
    double t1[8], t2[8], t3[8], t4[8], p1[8], p2[8], p3[8], f[8];
    double s = 0.0, ds;
    int k;
    *check = 0;
    for (k = 0; k < length - 1; k++) {
        const double *y = rows + 8 * k;
        double *y1 = rows + 8 * (k + 1);
        ds = g2_set_ds(g, y);
        s += ds;
        g2_rhs(g, y, f);
        for (int i = 0; i < n; i++) { t1[i] = ds * f[i]; p1[i] = y[i] + t1[i] / 2.0; }
        g2_rhs(g, p1, f);
        for (int i = 0; i < n; i++) { t2[i] = ds * f[i]; p2[i] = y[i] + t2[i] / 2.0; }
        g2_rhs(g, p2, f);
        for (int i = 0; i < n; i++) { t3[i] = ds * f[i]; p3[i] = y[i] + t3[i]; }
        g2_rhs(g, p3, f);
        for (int i = 0; i < n; i++) {
            t4[i] = ds * f[i];
            y1[i] = y[i] + t1[i] / 6.0 + t2[i] / 3.0 + t3[i] / 3.0 + t4[i] / 6.0;
        }
        if (s > g->s_max) { *check = 1; break; }
        if (y1[0] > g->rng_max || y1[1] > g->alt_max) { *check = 1; break; }
        if (y1[1] <= g->a->z_max && y1[1] < g2_topo_z(g, y1[0])) { *check = 0; break; }
        if (y1[1] > g->z_reflect) { *check = 0; break; }
    }
    return k + 1;
}

/* set_initial (:46-68). */
void go2_set_initial(go_2d *g, double *y, double r0, double z0)
{
    g->z_src = z0;
    g->c_eff0 = g2_ceff(g, z0);
    memset(y, 0, sizeof(double) * 8);
    y[0] = r0; y[1] = z0;
    y[2] = cos(g->theta); y[3] = sin(g->theta);
    if (g->calc_amp) { y[4] = 0.0; y[5] = 0.0; y[6] = -sin(g->theta); y[7] = cos(g->theta); }
}

/* approx_intercept + set_refl (:73-160); rk2, rk1 = rows k-2, k-1; out = row 0.
 * dC1 uses the C2 of this reflection (the reference reads C2 before setting
 * it, an uninitialised value). */
void go2_set_refl(go_2d *g, const double *rk2, const double *rk1, double *out)
{
    const int n = g->calc_amp ? 8 : 4;
    double prev[8];
    if (rk1[1] < g->a->z_bndlyr) {
        double dz_grnd = g2_topo_z(g, rk1[0]) - rk1[1];
        double dz_k = (rk1[1] - rk2[1]) - g2_topo_dz(g, rk1[0]) * (rk1[0] - rk2[0]);
        for (int i = 0; i < n; i++) prev[i] = rk1[i] + (rk1[i] - rk2[i]) / dz_k * dz_grnd;
    } else {
        double dz_grnd = g->z_reflect - rk1[1];
        double dz_k = (rk1[1] - rk2[1]);
        for (int i = 0; i < n; i++) prev[i] = rk1[i] + (rk1[i] - rk2[i]) / dz_k * dz_grnd;
    }
    double zg, c_eff, c_eff_diff, C1, C2, dC1, dC2, ds0_dtheta = 0.0;
    go_atmo *a = g->a;
    if (prev[1] < a->z_bndlyr) {
        zg = g2_topo_z(g, prev[0]);
        c_eff = g2_pt(g, &a->c, zg) + g2_pt(g, &a->u, zg) * cos(g->phi) + g2_pt(g, &a->v, zg) * sin(g->phi);
        c_eff_diff = g2_dpt(g, &a->c, zg) + g2_dpt(g, &a->u, zg) * cos(g->phi) + g2_dpt(g, &a->v, zg) * sin(g->phi);
        if (g->is_topo) {
            double ta = g2_topo_dz(g, prev[0]);
            double norm = 1.0 + pow(ta, 2);
            double da_dtheta = (prev[4] * prev[3] - prev[5] * prev[2]) / (prev[3] - ta * prev[2]) * g2_topo_ddz(g, prev[0]);
            C1 = (1.0 - pow(ta, 2)) / norm;
            C2 = 2.0 * ta / norm;
            dC1 = -2.0 * da_dtheta / norm * C2;
            dC2 = 2.0 * da_dtheta / norm * C1;
            ds0_dtheta = -g->c_eff0 / c_eff * (prev[5] - ta * prev[4]) / (prev[3] - ta * prev[2]);
        } else {
            C1 = 1.0; dC1 = 0.0; C2 = 0.0; dC2 = 0.0;
            ds0_dtheta = -g->c_eff0 / c_eff * prev[5] / prev[3];
        }
    } else {
        zg = g->z_reflect;
        c_eff = g2_pt(g, &a->c, zg) + g2_pt(g, &a->u, zg) * cos(g->phi) + g2_pt(g, &a->v, zg) * sin(g->phi);
        c_eff_diff = g2_dpt(g, &a->c, zg) + g2_dpt(g, &a->u, zg) * cos(g->phi) + g2_dpt(g, &a->v, zg) * sin(g->phi);
        C1 = 1.0; dC1 = 0.0; C2 = 0.0; dC2 = 0.0;
        ds0_dtheta = -g->c_eff0 / c_eff * prev[5] / prev[3];
    }
    double dnur_ds = 0.0, dnuz_ds = -g->c_eff0 / pow(c_eff, 2) * c_eff_diff;
    memset(out, 0, sizeof(double) * 8);
    out[0] = prev[0];
    out[1] = zg;
    out[2] = C1 * prev[2] + C2 * prev[3];
    out[3] = -C1 * prev[3] + C2 * prev[2];
    if (g->calc_amp) {
        out[4] = C1 * prev[4] + C2 * prev[5];
        out[5] = -C1 * prev[5] + C2 * prev[4];
        out[6] = C1 * prev[6] + C2 * prev[7] + dC1 * prev[2] + dC2 * prev[3] + ((C1 - 1.0) * dnur_ds + C2 * dnuz_ds) * ds0_dtheta;
        out[7] = -C1 * prev[7] + C2 * prev[6] - dC1 * prev[3] + dC2 * prev[2] + (-(C1 + 1.0) * dnuz_ds + C2 * dnur_ds) * ds0_dtheta;
    }
}

double go2_travel_time_seg(go_2d *g, const double *y0, const double *y1)   /* :283-296 */
{
    double dr = y1[0] - y0[0], dz = y1[1] - y0[1];
    double z_avg = y0[1] + dz / 2.0;
    double c_eff = g2_ceff(g, z_avg);
    double ds = sqrt(pow(dr, 2) + pow(dz, 2));
    return ds / c_eff;
}

double go2_atten_seg(go_2d *g, const double *y0, const double *y1, double freq)   /* :313-326 */
{
    double dr = y1[0] - y0[0], dz = y1[1] - y0[1];
    double z_avg = y0[1] + dz / 2.0;
    double ds = sqrt(pow(dr, 2) + pow(dz, 2));
    go_atmo *a = g->a;
    return go_sb_alpha_at(g2_pt(g, &a->c, z_avg), g2_pt(g, &a->rho, z_avg), z_avg, 0, a->tweak_abs, freq) * ds;
}

double go2_jacobian(go_2d *g, const double *y)        /* :345-351 */
{
    double r = y[0];
    double drds = g2_pt(g, &g->a->c, y[1]) / g->c_eff0 * y[2], drdth = y[4];
    double dzds = g2_pt(g, &g->a->c, y[1]) / g->c_eff0 * y[3], dzdth = y[5];
    return r * (drds * dzdth - dzds * drdth);
}

double go2_amp(go_2d *g, const double *y)             /* :354-361 */
{
    double z = y[1];
    double num = g2_pt(g, &g->a->rho, z) * g2_pt(g, &g->a->c, z) * cos(g->theta);
    double den = g2_pt(g, &g->a->rho, g->z_src) * g->c_eff0 * go2_jacobian(g, y);
    return sqrt(fabs(num / den));
}

/* est_dev (:366-380): out-of-plane drift over rows 0..k of the leg. */
double go2_est_dev(go_2d *g, const double *rows, int k)
{
    double dev = 0.0;
    go_atmo *a = g->a;
    for (int n = 0; n < k; n++) {
        const double *y0 = rows + 8 * n, *y1 = rows + 8 * (n + 1);
        double ds = sqrt(pow(y1[0] - y0[0], 2) + pow(y1[1] - y0[1], 2));
        double z_avg = (y1[1] + y0[1]) / 2.0;
        double c_eff = g2_ceff(g, z_avg);
        double uu = g2_pt(g, &a->u, z_avg), vv = g2_pt(g, &a->v, z_avg);
        double v_perp = -uu * sin(g->phi) + vv * cos(g->phi);
        dev += v_perp / c_eff * 1.0 / sqrt(1.0 + 2.0 * v_perp / c_eff + (pow(uu, 2) + pow(vv, 2)) / pow(c_eff, 2)) * ds;
    }
    return dev;
}

/* run_prop of infraga-2d (main.infraga.2d.cpp:318-395): raypath rows of
 * every leg (always written), caustics when flags & GO_PATH_CAUSTICS,
 * elevated reflections above the boundary layer not counted as bounces.
 * Arrival rows: pos0 = r, pos1 = perpendicular drift of the last leg,
 * back_az = 0. */
int go2_fan(go_2d *g, const double *launches, int n_launch, int bounces, double src_alt, double freq, int flags,
            go_arrival *out, int out_cap, int *n_out, go_path_pt *pts, long long pts_cap, long long *n_pts,
            long long *ray_steps)
{
    const int caus = (flags & GO_PATH_CAUSTICS) != 0 && g->calc_amp;
    int length = (int)(g->s_max / g->ds_min);
    double *rows = get_rows(length);          /* 18 doubles per row: room for 8 */
    double *r8 = rows;
    int cnt = 0;
    long long steps = 0, np = 0;
    const double r_src = 0.0;
    double z_src = src_alt > g2_topo_z(g, r_src) ? src_alt : g2_topo_z(g, r_src);   /* :222 */
    for (int li = 0; li < n_launch; li++) {
        double incl = launches[2 * li], az = launches[2 * li + 1];
        g->theta = incl * GO_PI / 180.0;
        g->phi = GO_PI / 2.0 - az * GO_PI / 180.0;
        go2_set_initial(g, r8, r_src, z_src);
        double tt = 0.0, att = 0.0, zmax = 0.0;
        int brk = 0, k;
        for (int bnc = 0; bnc <= bounces; bnc++) {
            k = go2_prop_rk4(g, r8, length, &brk);
            steps += k;
            double D = 0.0, D_prev = caus ? go2_jacobian(g, r8 + 8) : 0.0;
            for (int m = 1; m < k; m++) {
                const double *ym = r8 + 8 * m;
                tt += go2_travel_time_seg(g, r8 + 8 * (m - 1), ym);
                att += go2_atten_seg(g, r8 + 8 * (m - 1), ym, freq);
                if (caus) D = go2_jacobian(g, ym);
                if (m == 1 || m % 15 == 0) {
                    if (np < pts_cap) {
                        go_path_pt *P = pts + np;
                        double zt = g2_topo_z(g, ym[0]);
                        P->launch_idx = li; P->leg = bnc; P->step = m; P->kind = 0;
                        P->pos0 = ym[0]; P->pos1 = 0.0; P->z = ym[1] > zt ? ym[1] : zt;
                        P->tl = g->calc_amp ? 20.0 * log10(go2_amp(g, ym)) : 0.0;
                        P->absorption = -2.0 * att; P->time = tt;
                    }
                    np++;
                }
                if (caus && D * D_prev < 0.0) {
                    if (np < pts_cap) {
                        go_path_pt *P = pts + np;
                        P->launch_idx = li; P->leg = bnc; P->step = m; P->kind = 1;
                        P->pos0 = ym[0]; P->pos1 = 0.0; P->z = ym[1];
                        P->tl = 0.0; P->absorption = -2.0 * att; P->time = tt;
                    }
                    np++;
                }
                if (caus) D_prev = D;
            }
            for (int m = 0; m < k; m++) zmax = zmax > r8[8 * m + 1] ? zmax : r8[8 * m + 1];
            if (brk || k < 2) break;
            if (r8[8 * (k - 1) + 1] > g->a->z_bndlyr) {
                bnc -= 1;                                     /* elevated reflection (:371-372) */
            } else {
                const double *yk = r8 + 8 * k;
                go_arrival A;
                memset(&A, 0, sizeof(A));
                A.launch_idx = li;
                A.n_bounce = bnc;
                A.incl_launch = incl;
                A.az_launch = az;
                A.pos0 = yk[0];
                A.time = tt;
                A.cel = yk[0] / tt;
                A.turn_ht = zmax;
                A.incl_arr = -asin(g2_pt(g, &g->a->c, g2_topo_z(g, yk[0])) / g2_pt(g, &g->a->c, z_src) * yk[3]) * (180.0 / GO_PI);
                A.tl = g->calc_amp ? 20.0 * log10(go2_amp(g, yk)) : 0.0;
                A.absorption = -2.0 * att;
                A.pos1 = go2_est_dev(g, r8, k);
                A.back_az = 0.0;
                if (cnt < out_cap) out[cnt] = A;
                cnt++;
            }
            double rk2[8], rk1[8];
            memcpy(rk2, r8 + 8 * (k - 2), sizeof(rk2));
            memcpy(rk1, r8 + 8 * (k - 1), sizeof(rk1));
            go2_set_refl(g, rk2, rk1, r8);
        }
    }
    *n_out = cnt;
    if (n_pts) *n_pts = np;
    if (ray_steps) *ray_steps = steps;
    return (cnt > out_cap || np > pts_cap) ? -1 : 0;
}

double go2_topo_z(go_2d *g, double r) { return g2_topo_z(g, r); }
double go2_ceff(go_2d *g, double z) { return g2_ceff(g, z); }
double go2_rho(go_2d *g, double z) { return g2_pt(g, &g->a->rho, z); }

void *go2_h_new(void *atmo_h) { go_2d *g = (go_2d *)malloc(sizeof(go_2d)); go2_init(g, (go_atmo *)atmo_h); return g; }
void go2_h_free(void *h) { if (h) { go2_free((go_2d *)h); free(h); } }
void go2_h_config(void *h, int calc_amp, double z_reflect, double alt_max, double rng_max, double s_max, double ds_min,
                  double ds_max)
{
    go_2d *g = (go_2d *)h;
    g->calc_amp = calc_amp;
    if (!isnan(z_reflect)) g->z_reflect = z_reflect;
    if (!isnan(alt_max)) g->alt_max = alt_max;
    if (!isnan(rng_max)) g->rng_max = rng_max;
    if (!isnan(s_max)) g->s_max = s_max;
    if (!isnan(ds_min)) g->ds_min = ds_min;
    if (!isnan(ds_max)) g->ds_max = ds_max;
}
#endif
