/* ga_oracle_rd.c -- CPU restatement of infraGA's spherical range-dependent
 * propagation (TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of
 * the GPU range-dependent trace; never part of the product path).
 *
 * Operation order follows the reference source expression by expression so
 * that, compiled with gcc -O2 and no FMA contraction, the medium functions
 * are bit-identical to the reference's own compiled atmo_state.sph.rngdep.cpp
 * (tests/test_oracle_rd.py checks that against oracle/_ref).  File:line
 * citations are relative to the infraGA source tree.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ga_oracle_rd.h"

#define RD_PI 3.141592653589793238462643   /* geoac.params.sph.cpp:9 */

static double sq(double x) { return pow(x, 2); }
static double in_iv(double x, double lo, double hi)   /* interpolation.cpp:52-54 */
{
    double m = x > lo ? x : lo;
    return m < hi ? m : hi;
}
static double dmax(double a, double b) { return a < b ? b : a; }   /* std::max */
static double dmin(double a, double b) { return b < a ? b : a; }   /* std::min */

/* ======================================================================
 * Set-up: atmo_io.sph.rngdep.cpp:27-43 (set_limits), 61-200 / 203-366
 * (set_region), atmo_state.sph.rngdep.cpp:42-55 (set_bndlyr)
 * ==================================================================== */
static void set_limits(go_rd *a)
{
    double buffer = a->cart ? 1.0e-3 : 1.0e-3 * (RD_PI / 180.0);   /* atmo_io.3d.rngdep.cpp:25 */
    const go_h3d *c = &a->c;
    a->lat_min = c->x[0] + buffer;  a->lat_max = c->x[c->nx - 1] - buffer;
    a->lon_min = c->y[0] + buffer;  a->lon_max = c->y[c->ny - 1] - buffer;
    if (a->is_topo) {
        const go_s2d *t = &a->topo;
        a->lat_min = dmax(a->lat_min, t->x[0] + buffer); a->lat_max = dmin(a->lat_max, t->x[t->nx - 1] - buffer);
        a->lon_min = dmax(a->lon_min, t->y[0] + buffer); a->lon_max = dmin(a->lon_max, t->y[t->ny - 1] - buffer);
    }
    a->z0 = c->z[0];
    a->alt_max = c->z[c->nz - 1];
}

static void set_bndlyr(go_rd *a)
{
    if (a->is_topo) {
        a->z_max = 0.0;
        for (int i = 0; i < a->topo.nx; i++)
            for (int j = 0; j < a->topo.ny; j++) a->z_max = dmax(a->z_max, a->topo.f[i * a->topo.ny + j]);
        if (a->strat) {
            /* the stratified sources raise z_max first, then add 2 km:
             * atmo_state.sph.strat.cpp:42-54 (+0.01), atmo_state.3d.strat.cpp:36-48 (+0.1) */
            a->z_max += a->cart ? 0.1 : 0.01;
            a->z_bndlyr = a->z_max + 2.0;
            return;
        }
        a->z_bndlyr = a->z_max + 2.0;
        a->z_max += 0.01;
    } else {
        a->z_max = a->z0;
        a->z_bndlyr = a->z_max + 2.0;
    }
}

int go_rd_init(go_rd *a, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
               const double *z, int nz, const double *c, const double *u, const double *v,
               const double *rho)
{
    memset(a, 0, sizeof(*a));
    if (nlat < 2 || nlon < 2 || nz < 3) return -1;
    double *x = (double *)malloc(sizeof(double) * (size_t)nlat);
    double *y = (double *)malloc(sizeof(double) * (size_t)nlon);
    for (int i = 0; i < nlat; i++) { x[i] = lat_deg[i]; x[i] *= RD_PI / 180.0; }   /* :98 */
    for (int j = 0; j < nlon; j++) { y[j] = lon_deg[j]; y[j] *= RD_PI / 180.0; }   /* :109 */
    int rc = go_h3d_init(&a->c, x, nlat, y, nlon, z, nz, c);
    rc |= go_h3d_init(&a->u, x, nlat, y, nlon, z, nz, u);
    rc |= go_h3d_init(&a->v, x, nlat, y, nlon, z, nz, v);
    rc |= go_h3d_init(&a->rho, x, nlat, y, nlon, z, nz, rho);
    free(x);
    free(y);
    if (rc) return -1;
    a->r0 = 6370.0;                /* globe.cpp:10 */
    a->bl_param = 0.1;             /* atmo_state.sph.rngdep.cpp:95 */
    a->vert_wind_grad = 0.5;       /* :96 */
    a->tweak_abs = 1.0;            /* main.infraga.sph.rngdep.cpp:224 */
    a->is_topo = 0;
    a->use_blw = 0;
    set_limits(a);
    set_bndlyr(a);
    return 0;
}

int go_rd_set_topo(go_rd *a, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
                   const double *zg)
{
    double *x = (double *)malloc(sizeof(double) * (size_t)nlat);
    double *y = (double *)malloc(sizeof(double) * (size_t)nlon);
    for (int i = 0; i < nlat; i++) { x[i] = lat_deg[i]; if (!a->cart) x[i] *= RD_PI / 180.0; }   /* :231 */
    for (int j = 0; j < nlon; j++) { y[j] = lon_deg[j]; if (!a->cart) y[j] *= RD_PI / 180.0; }   /* :232 */
    if (a->is_topo) go_s2d_free(&a->topo);
    int rc = go_s2d_init(&a->topo, x, nlat, y, nlon, zg);
    free(x);
    free(y);
    if (rc) return -1;
    a->is_topo = 1;
    set_limits(a);
    set_bndlyr(a);
    return 0;
}

void go_rd_set_ground(go_rd *a, double z0)
{
    if (a->is_topo) return;        /* "cannot adjust ground elevation with topography" */
    a->z0 = z0;
    a->z_max = z0;
    a->z_bndlyr = z0 + 2.0;
}

void go_rd_free(go_rd *a)
{
    go_h3d_free(&a->c); go_h3d_free(&a->u); go_h3d_free(&a->v); go_h3d_free(&a->rho);
    if (a->is_topo) go_s2d_free(&a->topo);
    memset(a, 0, sizeof(*a));
}

void go_rd_params(const go_rd *a, int calc_amp, go_params *p)
{
    memset(p, 0, sizeof(*p));
    p->calc_amp = calc_amp;
    p->eq_cnt = calc_amp ? 18 : 6;
    p->s_max = 1000.0; p->ds_min = 0.001; p->ds_max = 0.05;
    p->alt_max = a->alt_max;
    p->rng_max = 1000.0;
    p->time_max = 10.0;
    p->lat_min = a->lat_min; p->lat_max = a->lat_max;
    p->lon_min = a->lon_min; p->lon_max = a->lon_max;
}

void go_rd_centre(const go_rd *a, double *lat_deg, double *lon_deg)
{
    *lat_deg = (a->lat_max + a->lat_min) / 2.0 * (180.0 / RD_PI);
    *lon_deg = (a->lon_max + a->lon_min) / 2.0 * (180.0 / RD_PI);
}

/* ======================================================================
 * Topography and boundary layer: atmo_state.sph.rngdep.cpp:57-133
 * ==================================================================== */
static double tx(go_rd *a, double lat) { return in_iv(lat, a->topo.x[0], a->topo.x[a->topo.nx - 1]); }
static double ty(go_rd *a, double lon) { return in_iv(lon, a->topo.y[0], a->topo.y[a->topo.ny - 1]); }

/* Without terrain the spherical ground sits at r0 + z0
 * (atmo_state.sph.rngdep.cpp:57-66); the Cartesian topo::z returns 0.0
 * whatever z0 is (atmo_state.3d.rngdep.cpp:45-52), so the 3D geometry
 * reflects at z = 0 while z_max, the boundary-layer top and the
 * turning-height test still follow z0. */
double go_rd_topo_z(go_rd *a, double lat, double lon)
{
    double result = a->r0;
    if (a->is_topo) result += go_s2d_f(&a->topo, tx(a, lat), ty(a, lon));
    else if (!a->cart) result += a->z0;
    return result;
}
double go_rd_topo_dz(go_rd *a, double lat, double lon, int n)
{
    double result = 0.0;
    if (a->is_topo) result += go_s2d_df(&a->topo, tx(a, lat), ty(a, lon), n);
    return result;
}
double go_rd_topo_ddz(go_rd *a, double lat, double lon, int n1, int n2)
{
    double result = 0.0;
    if (a->is_topo) result += go_s2d_ddf(&a->topo, tx(a, lat), ty(a, lon), n1, n2);
    return result;
}
double go_rd_topo_dddz(go_rd *a, double lat, double lon, int n1, int n2, int n3)
{
    if (a->is_topo) return go_s2d_dddf(&a->topo, tx(a, lat), ty(a, lon), n1, n2, n3);
    return 0.0;
}

double go_rd_bl_sc(go_rd *a, double r, double lat, double lon)
{
    double X, result = 1.0;
    if (r - a->r0 < a->z_bndlyr) {
        X = exp(-(r - go_rd_topo_z(a, lat, lon)) / a->bl_param);
        result = (1.0 - X) / (1.0 + X);
    }
    return result;
}
double go_rd_bl_dsc(go_rd *a, double r, double lat, double lon, int n)
{
    double X, dX, result = 0.0;
    if (r - a->r0 < a->z_bndlyr) {
        X = exp(-(r - go_rd_topo_z(a, lat, lon)) / a->bl_param);
        if (n > 0) dX = X / a->bl_param * go_rd_topo_dz(a, lat, lon, n - 1);
        else dX = -X / a->bl_param;
        result = -2.0 * dX / sq(1.0 + X);
    }
    return result;
}
double go_rd_bl_ddsc(go_rd *a, double r, double lat, double lon, int n1, int n2)
{
    double X, dX1, dX2, ddX, result = 0.0, bp = a->bl_param;
    if (r - a->r0 < a->z_bndlyr) {
        X = exp(-(r - go_rd_topo_z(a, lat, lon)) / bp);
        if (n1 > 0) dX1 = X / bp * go_rd_topo_dz(a, lat, lon, n1 - 1); else dX1 = -X / bp;
        if (n2 > 0) dX2 = X / bp * go_rd_topo_dz(a, lat, lon, n2 - 1); else dX2 = -X / bp;
        if (n1 > 0 && n2 > 0)
            ddX = 1.0 / bp * (go_rd_topo_ddz(a, lat, lon, n1 - 1, n2 - 1)
                              + 1.0 / bp * go_rd_topo_dz(a, lat, lon, n1 - 1) * go_rd_topo_dz(a, lat, lon, n2 - 1)) * X;
        else if (n1 > 0 && n2 == 0) ddX = -X / sq(bp) * go_rd_topo_dz(a, lat, lon, n1 - 1);
        else if (n1 == 0 && n2 > 0) ddX = -X / sq(bp) * go_rd_topo_dz(a, lat, lon, n2 - 1);
        else ddX = X / sq(bp);
        result = -2.0 * (ddX / sq(1.0 + X) - 2.0 * dX1 * dX2 / pow(1.0 + X, 3));
    }
    return result;
}

/* ======================================================================
 * Medium point functions: atmo_state.sph.rngdep.cpp:149-403.  The splines
 * run over (lat, lon, z); derivative index n of (r, lat, lon) maps to spline
 * index 2 for n = 0 and n - 1 otherwise.
 * ==================================================================== */
static void ev(go_rd *a, const go_h3d *s, double r, double lat, double lon, double *xe, double *ye, double *ze)
{
    *xe = in_iv(lat, s->x[0], s->x[s->nx - 1]);
    *ye = in_iv(lon, s->y[0], s->y[s->ny - 1]);
    *ze = in_iv(r - a->r0, s->z[0], s->z[s->nz - 1]);
}
static int nmap(int n) { return n == 0 ? 2 : n - 1; }

double go_rd_rho(go_rd *a, double r, double lat, double lon)
{
    double x, y, z;
    ev(a, &a->rho, r, lat, lon, &x, &y, &z);
    return go_h3d_f(&a->rho, x, y, z);
}
double go_rd_c(go_rd *a, double r, double lat, double lon)
{
    double x, y, z;
    ev(a, &a->c, r, lat, lon, &x, &y, &z);
    return go_h3d_f(&a->c, x, y, z);
}
double go_rd_dc(go_rd *a, double r, double lat, double lon, int n)
{
    double x, y, z;
    ev(a, &a->c, r, lat, lon, &x, &y, &z);
    return go_h3d_df(&a->c, x, y, z, nmap(n));
}
static double wind(go_rd *a, go_h3d *s, double r, double lat, double lon)
{
    double x, y, z;
    ev(a, s, r, lat, lon, &x, &y, &z);
    return go_h3d_f(s, x, y, z) * go_rd_bl_sc(a, r, lat, lon);
}
static double dwind(go_rd *a, go_h3d *s, double r, double lat, double lon, int n)
{
    double x, y, z, result;
    ev(a, s, r, lat, lon, &x, &y, &z);
    result = go_h3d_df(s, x, y, z, nmap(n)) * go_rd_bl_sc(a, r, lat, lon);
    result += go_h3d_f(s, x, y, z) * go_rd_bl_dsc(a, r, lat, lon, n);
    return result;
}
double go_rd_u(go_rd *a, double r, double lat, double lon) { return wind(a, &a->u, r, lat, lon); }
double go_rd_v(go_rd *a, double r, double lat, double lon) { return wind(a, &a->v, r, lat, lon); }
double go_rd_du(go_rd *a, double r, double lat, double lon, int n) { return dwind(a, &a->u, r, lat, lon, n); }
double go_rd_dv(go_rd *a, double r, double lat, double lon, int n) { return dwind(a, &a->v, r, lat, lon, n); }

static int blw_on(go_rd *a, double r) { return a->is_topo && (r - a->r0 < a->z_bndlyr) && a->use_blw; }

/* Cartesian atmo::w (atmo_state.3d.strat.cpp:235-256, the same expression as
 * atmo_state.3d.rngdep.cpp:243-267): u pairs with the x slope, v with the y
 * slope, no metric factors.  Arguments in the medium's (z, x, y) order. */
static double rd3_w(go_rd *a, double z, double x, double y)
{
    double xe, ye, ze, X, S, dS[2], w0, sg_ratio, result = 0.0, bp = a->bl_param;
    if (blw_on(a, z)) {
        ev(a, &a->v, z, x, y, &xe, &ye, &ze);
        sg_ratio = bp / a->vert_wind_grad;
        X = exp(-(z - go_rd_topo_z(a, x, y)) / bp);
        S = (1.0 - X) / (1.0 + X);
        dS[0] = -2.0 / bp * (X / sq(1.0 + X)) * go_rd_topo_dz(a, x, y, 0);
        dS[1] = -2.0 / bp * (X / sq(1.0 + X)) * go_rd_topo_dz(a, x, y, 1);
        w0 = go_h3d_f(&a->u, xe, ye, ze) * dS[0];
        w0 += go_h3d_f(&a->v, xe, ye, ze) * dS[1];
        w0 *= -2.0 * bp;
        result = S * X * w0;
        result /= (X + (sg_ratio / 2.0) * (1.0 - sq(X)));
    }
    return result;
}

/* Cartesian atmo::dw (atmo_state.3d.strat.cpp:258-317; rngdep :269-331);
 * n is the medium's derivative index (0 z, 1 x, 2 y). */
static double rd3_dw(go_rd *a, double z, double x, double y, int n)
{
    double xe, ye, ze, X, dX[3], ddX[2], S, dS[3], ddS[2], N, D, sg_ratio, result = 0.0, bp = a->bl_param;
    double dzg[2], u0, v0, w0, dw0, dN, dD;
    const int c = n == 0 ? 2 : n - 1;   /* the reference's index: 0 x, 1 y, 2 z */
    if (blw_on(a, z)) {
        ev(a, &a->v, z, x, y, &xe, &ye, &ze);
        sg_ratio = bp / a->vert_wind_grad;
        double zg = go_rd_topo_z(a, x, y);
        dzg[0] = go_rd_topo_dz(a, x, y, 0);
        dzg[1] = go_rd_topo_dz(a, x, y, 1);
        u0 = go_h3d_f(&a->u, xe, ye, ze);
        v0 = go_h3d_f(&a->v, xe, ye, ze);
        X = exp(-(z - zg) / bp);
        for (int j = 0; j < 2; j++) dX[j] = X / bp * dzg[j];
        dX[2] = -X / bp;
        S = (1.0 - X) / (1.0 + X);
        for (int j = 0; j < 3; j++) dS[j] = -2.0 * dX[j] / sq(1.0 + X);
        if (c < 2) {
            ddX[0] = X / bp * (go_rd_topo_ddz(a, x, y, 0, c) + 1.0 / bp * dzg[0] * dzg[c]);
            ddX[1] = X / bp * (go_rd_topo_ddz(a, x, y, 1, c) + 1.0 / bp * dzg[1] * dzg[c]);
        } else {
            ddX[0] = -X / sq(bp) * dzg[0];
            ddX[1] = -X / sq(bp) * dzg[1];
        }
        ddS[0] = -2.0 * (ddX[0] / sq(1.0 + X) - 2.0 * dX[0] * dX[c] / pow(1.0 + X, 3));
        ddS[1] = -2.0 * (ddX[1] / sq(1.0 + X) - 2.0 * dX[1] * dX[c] / pow(1.0 + X, 3));
        w0 = -2.0 * bp * (u0 * dS[0] + v0 * dS[1]);
        N = S * X * w0;
        D = X + (sg_ratio / 2.0) * (1.0 - sq(X));
        dw0 = u0 * ddS[0] + v0 * ddS[1];
        if (c == 2) {
            dw0 += go_h3d_df(&a->u, xe, ye, ze, 2) * dS[0];
            dw0 += go_h3d_df(&a->v, xe, ye, ze, 2) * dS[1];
        }
        dw0 *= -2.0 * bp;
        dN = (dS[c] * X + S * dX[c]) * w0 + S * X * dw0;
        dD = (1.0 - sg_ratio * X) * dX[c];
        result = (dN * D - N * dD) / sq(D);
    }
    return result;
}

/* atmo::w, :308-338 */
double go_rd_w(go_rd *a, double r, double lat, double lon)
{
    double x, y, z, zg, dzg[2], X, S, dS[2], w0, sg_ratio, result = 0.0, bp = a->bl_param;
    if (a->cart) return rd3_w(a, r, lat, lon);
    if (blw_on(a, r)) {
        ev(a, &a->v, r, lat, lon, &x, &y, &z);
        zg = go_rd_topo_z(a, lat, lon);
        dzg[0] = go_rd_topo_dz(a, lat, lon, 0);
        dzg[1] = go_rd_topo_dz(a, lat, lon, 1);
        sg_ratio = bp / a->vert_wind_grad;
        X = exp(-(r - zg) / bp);
        S = (1.0 - X) / (1.0 + X);
        dS[0] = -2.0 / bp * (X / sq(1.0 + X)) * dzg[0];
        dS[1] = -2.0 / bp * (X / sq(1.0 + X)) * dzg[1];
        w0 = go_h3d_f(&a->v, x, y, z) * dS[0];
        w0 += go_h3d_f(&a->u, x, y, z) * dS[1] / cos(lat);
        w0 *= -2.0 * bp / (a->r0 + zg);
        result = S * X * w0;
        result /= (X + (sg_ratio / 2.0) * (1.0 - sq(X)));
    }
    return result;
}

/* atmo::dw, :340-403 */
double go_rd_dw(go_rd *a, double r, double lat, double lon, int n)
{
    double x, y, z, X, dX[3], ddX[2], S, dS[3], ddS[2], N, D, sg_ratio, result = 0.0, bp = a->bl_param;
    double zg, dzg[2], u0, v0, w0, dw0, dN, dD;
    if (a->cart) return rd3_dw(a, r, lat, lon, n);
    if (blw_on(a, r)) {
        ev(a, &a->v, r, lat, lon, &x, &y, &z);
        sg_ratio = bp / a->vert_wind_grad;
        zg = go_rd_topo_z(a, lat, lon);
        dzg[0] = go_rd_topo_dz(a, lat, lon, 0);
        dzg[1] = go_rd_topo_dz(a, lat, lon, 1);
        u0 = go_h3d_f(&a->u, x, y, z);
        v0 = go_h3d_f(&a->v, x, y, z);
        X = exp(-(r - zg) / bp);
        for (int j = 0; j < 3; j++) {
            if (j > 0) dX[j] = X / bp * dzg[j - 1];
            else dX[j] = -X / bp;
        }
        S = (1.0 - X) / (1.0 + X);
        for (int j = 0; j < 3; j++) dS[j] = -2.0 * dX[j] / sq(1.0 + X);
        if (n > 0) {
            ddX[0] = X / bp * (go_rd_topo_ddz(a, lat, lon, 0, n - 1) + 1.0 / bp * dzg[0] * dzg[n - 1]);
            ddX[1] = X / bp * (go_rd_topo_ddz(a, lat, lon, 1, n - 1) + 1.0 / bp * dzg[1] * dzg[n - 1]);
        } else {
            ddX[0] = -X / sq(bp) * dzg[0];
            ddX[1] = -X / sq(bp) * dzg[1];
        }
        ddS[0] = -2.0 * (ddX[0] / sq(1.0 + X) - 2.0 * dX[1] * dX[n] / pow(1.0 + X, 3));
        ddS[1] = -2.0 * (ddX[1] / sq(1.0 + X) - 2.0 * dX[2] * dX[n] / pow(1.0 + X, 3));
        w0 = v0 * dS[1] + u0 * dS[2] / cos(lat);
        w0 *= -2.0 * bp / (a->r0 + zg);
        N = S * X * w0;
        D = X + (sg_ratio / 2.0) * (1.0 - sq(X));
        dw0 = v0 * ddS[0] + u0 * ddS[1] / cos(lat);
        if (n == 0) {
            dw0 += go_h3d_df(&a->v, x, y, z, 2) * dS[1];
            dw0 += go_h3d_df(&a->u, x, y, z, 2) * dS[2] / cos(lat);
        } else if (n == 1) {
            dw0 += u0 * dS[2] * sin(lat) / sq(cos(lat));
        }
        dw0 *= -2.0 * bp / (a->r0 + zg);
        dN = (dS[n] * X + S * dX[n]) * w0 + S * X * dw0;
        dD = (1.0 - sg_ratio * X) * dX[n];
        result = (dN * D - N * dD) / sq(D);
    }
    return result;
}

/* atmo::calc_uvw, both overloads: :584-691 (first derivatives) and :693-882
 * (second derivatives).  The terrain is evaluated at the unclamped (lat, lon)
 * and its first/second derivatives at the clamped ones in the first overload,
 * as the reference does; inside the region both agree. */
void go_rd_calc_uvw(go_rd *a, double r, double lat, double lon, double *u, double *v, double *w,
                    double du[3], double dv[3], double dw[3], double ddu[6], double ddv[6], double ddw[6])
{
    const double bp = a->bl_param;
    double lat_e, lon_e, z_e, zg, dzg[2], ddzg[2][2], dddzg[2][2][2];
    double u0, v0, du0[3], dv0[3], ddu0[6], ddv0[6], t3[3], t6[6];
    double X, dX[3], ddX[3][3], dddX[2][3][3], sc0, dsc0[3], ddsc0[3][3], dddsc0[2][3][3];
    double w0, sg_ratio, N, D;
    const int second = (ddu != NULL);

    ev(a, &a->v, r, lat, lon, &lat_e, &lon_e, &z_e);
    memset(dddzg, 0, sizeof(dddzg));
    if (a->is_topo && (r - a->r0 < a->z_bndlyr)) {
        if (!second) {
            zg = go_s2d_f(&a->topo, lat, lon);
            zg += a->r0;
            for (int n = 0; n < 2; n++) {
                dzg[n] = go_rd_topo_dz(a, lat, lon, n);
                for (int m = n; m < 2; m++) {
                    ddzg[n][m] = go_rd_topo_ddz(a, lat, lon, n, m);
                    ddzg[m][n] = ddzg[n][m];
                }
            }
        } else {
            zg = go_s2d_f(&a->topo, lat, lon);
            zg += a->r0;
            dzg[0] = go_s2d_df(&a->topo, lat, lon, 0);
            dzg[1] = go_s2d_df(&a->topo, lat, lon, 1);
            for (int n = 0; n < 2; n++)
                for (int m = 0; m < 2; m++) {
                    ddzg[n][m] = go_s2d_ddf(&a->topo, lat, lon, n, m);
                    for (int k = 0; k < 2; k++) dddzg[n][m][k] = go_s2d_dddf(&a->topo, lat, lon, n, m, k);
                }
        }
    } else {
        zg = go_rd_topo_z(a, lat, lon);
        for (int n = 0; n < 2; n++) {
            dzg[n] = 0.0;
            for (int m = 0; m < 2; m++) ddzg[n][m] = 0.0;
        }
    }

    go_h3d_all(&a->u, lat_e, lon_e, z_e, &u0, t3, t6);
    du0[1] = t3[0]; du0[2] = t3[1]; du0[0] = t3[2];
    ddu0[1] = t6[0]; ddu0[2] = t6[1]; ddu0[0] = t6[2]; ddu0[5] = t6[3]; ddu0[3] = t6[4]; ddu0[4] = t6[5];
    go_h3d_all(&a->v, lat_e, lon_e, z_e, &v0, t3, t6);
    dv0[1] = t3[0]; dv0[2] = t3[1]; dv0[0] = t3[2];
    ddv0[1] = t6[0]; ddv0[2] = t6[1]; ddv0[0] = t6[2]; ddv0[5] = t6[3]; ddv0[3] = t6[4]; ddv0[4] = t6[5];

    X = exp(-(r - zg) / bp);
    if (!second) {
        for (int n = 0; n < 3; n++) {
            if (n > 0) dX[n] = X / bp * dzg[n - 1];
            else dX[n] = -X / bp;
            for (int m = 0; m < 3; m++) {
                if ((n > 0) && (m > 0)) {
                    /* the Cartesian calc_uvw fills x <= y and mirrors
                     * (atmo_state.3d.rngdep.cpp:515-521) */
                    const int i = (a->cart && m < n) ? m : n, j = (a->cart && m < n) ? n : m;
                    ddX[n][m] = X / bp * (ddzg[i - 1][j - 1] + 1.0 / bp * dzg[i - 1] * dzg[j - 1]);
                } else if ((n == 0) && (m > 0)) ddX[n][m] = -X / sq(bp) * dzg[m - 1];
                else if ((n > 0) && (m == 0)) ddX[n][m] = -X / sq(bp) * dzg[n - 1];
                else ddX[n][m] = X / sq(bp);
            }
        }
    } else {
        for (int n = 1; n < 3; n++) {
            dX[n] = X / bp * dzg[n - 1];
            for (int m = 0; m < 3; m++) {
                if (m == 0) {
                    ddX[n][m] = -X / sq(bp) * dzg[n - 1];
                    ddX[m][n] = ddX[n][m];
                    for (int k = 0; k < 2; k++) {
                        dddX[k][n][m] = ddzg[n - 1][k] + dzg[n - 1] * dzg[k] / bp;
                        dddX[k][n][m] *= -X / sq(bp);
                        dddX[k][m][n] = dddX[k][n][m];
                    }
                } else if (a->strat && a->cart) {
                    /* atmo_state.3d.strat.cpp:600-625: the reciprocal form,
                     * filled for n <= m and mirrored */
                    const int i = n < m ? n : m, j = n < m ? m : n;
                    ddX[n][m] = X / bp * (ddzg[i - 1][j - 1] + 1.0 / bp * dzg[i - 1] * dzg[j - 1]);
                    for (int k = 0; k < 2; k++) {
                        dddX[k][n][m] = dddzg[k][i - 1][j - 1];
                        dddX[k][n][m] += (ddzg[i - 1][j - 1] * dzg[k] + ddzg[k][j - 1] * dzg[i - 1] + ddzg[k][i - 1] * dzg[j - 1]) / bp;
                        dddX[k][n][m] += dzg[k] * dzg[i - 1] * dzg[j - 1] / sq(bp);
                        dddX[k][n][m] *= X / bp;
                    }
                } else {
                    ddX[n][m] = ddzg[n - 1][m - 1] + dzg[n - 1] * dzg[m - 1] / bp;
                    ddX[n][m] *= X / bp;
                    ddX[m][n] = ddX[n][m];
                    for (int k = 0; k < 2; k++) {
                        dddX[k][n][m] = dddzg[k][n - 1][m - 1];
                        dddX[k][n][m] += (ddzg[n - 1][m - 1] * dzg[k] + ddzg[k][m - 1] * dzg[n - 1] + ddzg[k][n - 1] * dzg[m - 1]) / bp;
                        dddX[k][n][m] += dzg[k] * dzg[n - 1] * dzg[m - 1] / sq(bp);
                        dddX[k][n][m] *= X / bp;
                        dddX[k][m][n] = dddX[k][n][m];
                    }
                }
            }
        }
        dX[0] = -X / bp;
        ddX[0][0] = X / sq(bp);
        dddX[0][0][0] = X / pow(bp, 3) * dzg[0];
        dddX[1][0][0] = X / pow(bp, 3) * dzg[1];
    }

    sc0 = (1.0 - X) / (1.0 + X);
    for (int n = 0; n < 3; n++) {
        dsc0[n] = -2.0 * dX[n] / sq(1.0 + X);
        for (int m = 0; m < 3; m++) {
            ddsc0[n][m] = -2.0 * (ddX[n][m] / sq(1.0 + X) - 2.0 * dX[n] * dX[m] / pow(1.0 + X, 3));
            if (second)
                for (int k = 0; k < 2; k++) {
                    dddsc0[k][n][m] = dddX[k][n][m] / sq(1.0 + X);
                    dddsc0[k][n][m] -= 2.0 * (ddX[n][m] * dX[k + 1] + ddX[n][k + 1] * dX[m] + ddX[k + 1][m] * dX[n]) / pow(1.0 + X, 3);
                    dddsc0[k][n][m] += 6.0 * dX[n] * dX[m] * dX[k + 1] / pow(1.0 + X, 4);
                    dddsc0[k][n][m] *= -2.0;
                }
        }
    }

    *u = u0 * sc0;
    *v = v0 * sc0;
    for (int n = 0; n < 3; n++) {
        du[n] = du0[n] * sc0 + u0 * dsc0[n];
        dv[n] = dv0[n] * sc0 + v0 * dsc0[n];
    }
    if (second) {
        for (int n = 0; n < 3; n++) {
            ddu[n] = ddu0[n] * sc0 + 2.0 * du0[n] * dsc0[n] + u0 * ddsc0[n][n];
            ddv[n] = ddv0[n] * sc0 + 2.0 * dv0[n] * dsc0[n] + v0 * ddsc0[n][n];
        }
        ddu[3] = ddu0[3] * sc0 + du0[0] * dsc0[1] + du0[1] * dsc0[0] + u0 * ddsc0[0][1];
        ddu[4] = ddu0[4] * sc0 + du0[0] * dsc0[2] + du0[2] * dsc0[0] + u0 * ddsc0[0][2];
        ddu[5] = ddu0[5] * sc0 + du0[1] * dsc0[2] + du0[2] * dsc0[1] + u0 * ddsc0[1][2];
        ddv[3] = ddv0[3] * sc0 + dv0[0] * dsc0[1] + dv0[1] * dsc0[0] + v0 * ddsc0[0][1];
        ddv[4] = ddv0[4] * sc0 + dv0[0] * dsc0[2] + dv0[2] * dsc0[0] + v0 * ddsc0[0][2];
        ddv[5] = ddv0[5] * sc0 + dv0[1] * dsc0[2] + dv0[2] * dsc0[1] + v0 * ddsc0[1][2];
    }

    if (a->cart && blw_on(a, r)) {
        /* Cartesian calc_uvw vertical wind: atmo_state.3d.strat.cpp:531-553
         * (first overload) and :650-710 (second), in their operation order.
         * The 3D range-dependent calc_uvw (atmo_state.3d.rngdep.cpp:543-563,
         * :667-720) and its atmo::ddw (:333-468, which its equation set
         * contracts) are the same expressions without the horizontal
         * derivatives of u0 and v0, which a stratified profile does not have:
         * those terms enter for the stratified case only.  The reference's
         * index i (0 x, 1 y, 2 z) sits at medium index I[i]. */
        static const int I[3] = {1, 2, 0};
        double dw0[3], dN[3], dD[3], ddw0, ddN, ddD;
        w0 = -2.0 * bp * (u0 * dsc0[1] + v0 * dsc0[2]);
        sg_ratio = bp / a->vert_wind_grad;
        N = sc0 * X * w0;
        D = X + sg_ratio / 2.0 * (1.0 - sq(X));
        *w = N / D;
        for (int i = 0; i < 3; i++) {
            const int n = I[i];
            dw0[i] = u0 * ddsc0[1][n] + v0 * ddsc0[2][n];
            if (i == 2 || a->strat) dw0[i] += du0[n] * dsc0[1] + dv0[n] * dsc0[2];
            dw0[i] *= -2.0 * bp;
            dN[i] = (dsc0[n] * X + sc0 * dX[n]) * w0 + sc0 * X * dw0[i];
            dD[i] = (1.0 - sg_ratio * X) * dX[n];
            dw[n] = (dN[i] * D - N * dD[i]) / sq(D);
            if (!second) continue;
            if (i < 2) {
                ddw0 = -2.0 * bp * (dddsc0[0][n][n] * u0 + dddsc0[1][n][n] * v0);
            } else {
                ddw0 = dddsc0[0][0][0] * u0 + 2.0 * ddsc0[1][0] * du0[0] + dsc0[1] * ddu0[0];
                ddw0 += dddsc0[1][0][0] * v0 + 2.0 * ddsc0[2][0] * dv0[0] + dsc0[2] * ddv0[0];
                ddw0 *= -2.0 * bp;
            }
            ddN = ddsc0[n][n] * X * w0 + sc0 * ddX[n][n] * w0 + sc0 * X * ddw0;
            ddN += 2.0 * (dsc0[n] * dX[n] * w0 + dsc0[n] * X * dw0[i] + sc0 * dX[n] * dw0[i]);
            ddD = (1.0 - sg_ratio * X) * ddX[n][n] - sg_ratio * sq(dX[n]);
            ddw[n] = ((ddN * D - 2.0 * dN[i] * dD[i] - N * ddD) * sq(D) + 2.0 * N * D * sq(dD[i])) / pow(D, 4);
        }
        if (second) {
            /* the reference's ddw[3..5] = xy, xz, yz: medium packed 5, 3, 4 */
            static const int P[3][3] = {{0, 1, 5}, {0, 2, 3}, {1, 2, 4}};
            for (int k = 0; k < 3; k++) {
                const int i = P[k][0], j = P[k][1], p = I[i], q = I[j];
                if (j < 2) {
                    ddw0 = -2.0 * bp * (dddsc0[0][p][q] * u0 + dddsc0[1][p][q] * v0);
                } else {
                    ddw0 = dddsc0[0][p][q] * u0 + ddsc0[1][p] * du0[0];
                    ddw0 += dddsc0[1][p][q] * v0 + ddsc0[2][p] * dv0[0];
                    ddw0 *= -2.0 * bp;
                }
                ddN = ddsc0[p][q] * X * w0 + sc0 * ddX[p][q] * w0 + sc0 * X * ddw0;
                ddN += dsc0[p] * dX[q] * w0 + dsc0[p] * X * dw0[j] + sc0 * dX[p] * dw0[j];
                ddN += dsc0[q] * dX[p] * w0 + dsc0[q] * X * dw0[i] + sc0 * dX[q] * dw0[i];
                ddD = (1.0 - sg_ratio * X) * ddX[p][q] - sg_ratio * dX[p] * dX[q];
                ddw[P[k][2]] = ((ddN * D - dN[i] * dD[j] - dN[j] * dD[i] - N * ddD) * sq(D) + 2.0 * N * D * dD[i] * dD[j]) / pow(D, 4);
            }
        }
    } else if (a->is_topo && (r - a->r0 < a->z_bndlyr) && a->use_blw) {
        w0 = v0 * dsc0[1] + u0 * dsc0[2] / cos(lat);
        w0 *= -2.0 * bp / (a->r0 + zg);
        sg_ratio = bp / a->vert_wind_grad;
        N = sc0 * X * w0;
        D = X + (sg_ratio / 2.0) * (1.0 - sq(X));
        *w = N / D;
        if (!second) {
            for (int n = 0; n < 3; n++) {
                double dw0, dN, dD;
                if (n == 0) {
                    dw0 = dv0[0] * dsc0[1] + v0 * ddsc0[1][0];
                    dw0 += (du0[0] * dsc0[2] + u0 * ddsc0[2][0]) / cos(lat);
                } else if (n == 1) {
                    dw0 = v0 * ddsc0[1][1];
                    dw0 += u0 / cos(lat) * (ddsc0[1][2] + dsc0[2] * tan(lat));
                } else {
                    dw0 = v0 * ddsc0[1][2];
                    dw0 += u0 * ddsc0[2][2] / cos(lat);
                }
                dw0 *= -2.0 * bp / (a->r0 + zg);
                dN = (dsc0[n] * X + sc0 * dX[n]) * w0 + sc0 * X * dw0;
                dD = (1.0 - sg_ratio * X) * dX[n];
                dw[n] = (dN * D - N * dD) / sq(D);
            }
        } else {
            double dw0[3], dN[3], dD[3], ddw0, ddN, ddD;
            for (int n = 0; n < 3; n++) {
                dw0[n] = v0 * ddsc0[1][n] + u0 * ddsc0[2][n] / cos(lat);
                if (n == 0) dw0[n] += dv0[0] * dsc0[1] + du0[0] * dsc0[2] / cos(lat);
                else if (n == 1) dw0[n] += u0 * dsc0[2] * sin(lat) / sq(cos(lat));
                dw0[n] *= -2.0 * bp / (a->r0 + zg);
                dN[n] = (dsc0[n] * X + sc0 * dX[n]) * w0 + sc0 * X * dw0[n];
                dD[n] = (1.0 - sg_ratio * X) * dX[n];
                dw[n] = (dN[n] * D - N * dD[n]) / sq(D);

                ddw0 = dddsc0[0][n][n] * v0 + dddsc0[1][n][n] * u0 / cos(lat);
                if (n == 0) {
                    ddw0 += 2.0 * ddsc0[1][0] * dv0[0] + dsc0[1] * ddv0[0];
                    ddw0 += (2.0 * ddsc0[2][0] * du0[0] + dsc0[2] * ddu0[0]) / cos(lat);
                } else if (n == 1) {
                    ddw0 += 2.0 * sin(lat) / sq(cos(lat)) * ddsc0[1][2] * u0;
                    ddw0 += (sq(sin(lat)) + 1.0) / pow(cos(lat), 3) * dsc0[2] * u0;
                }
                ddw0 *= -2.0 * bp / (a->r0 + zg);
                ddN = ddsc0[n][n] * X * w0 + sc0 * ddX[n][n] * w0 + sc0 * X * ddw0;
                ddN += 2.0 * (dsc0[n] * dX[n] * w0 + dsc0[n] * X * dw0[n] + sc0 * dX[n] * dw0[n]);
                ddD = -sg_ratio * sq(dX[n]) + (1.0 - sg_ratio * X) * ddX[n][n];
                ddw[n] = ((ddN * D - 2.0 * dN[n] * dD[n] - N * ddD) * sq(D) + 2.0 * N * D * sq(dD[n])) / pow(D, 4);
            }
            /* ddw[3..5]: r-lat, r-lon, lat-lon (:848-876) */
            ddw0 = dddsc0[0][0][1] * v0 + ddsc0[1][1] * dv0[0] + (dddsc0[1][0][1] * u0 + ddsc0[1][2] * du0[0]) / cos(lat);
            ddw0 += sin(lat) / sq(cos(lat)) * (ddsc0[0][2] * u0 + dsc0[2] * du0[0]);
            ddw0 *= -2.0 * bp / (a->r0 + zg);
            ddN = ddsc0[0][1] * X * w0 + sc0 * ddX[0][1] * w0 + sc0 * X * ddw0;
            ddN += dsc0[0] * dX[1] * w0 + dsc0[0] * X * dw0[1] + sc0 * dX[0] * dw0[1];
            ddN += dsc0[1] * dX[0] * w0 + dsc0[1] * X * dw0[0] + sc0 * dX[1] * dw0[0];
            ddD = -sg_ratio * dX[0] * dX[1] + (1.0 - sg_ratio * X) * ddX[0][1];
            ddw[3] = ((ddN * D - dN[0] * dD[1] - dN[1] * dD[0] - N * ddD) * sq(D) + 2.0 * N * D * dD[0] * dD[1]) / pow(D, 4);

            ddw0 = dddsc0[0][0][2] * v0 + ddsc0[1][2] * dv0[0] + (dddsc0[1][0][2] * u0 + ddsc0[2][2] * du0[0]) / cos(lat);
            ddw0 *= -2.0 * bp / (a->r0 + zg);
            ddN = ddsc0[0][2] * X * w0 + sc0 * ddX[0][2] * w0 + sc0 * X * ddw0;
            ddN += dsc0[0] * dX[2] * w0 + dsc0[0] * X * dw0[2] + sc0 * dX[0] * dw0[2];
            ddN += dsc0[2] * dX[0] * w0 + dsc0[2] * X * dw0[0] + sc0 * dX[2] * dw0[0];
            ddD = -sg_ratio * dX[0] * dX[2] + (1.0 - sg_ratio * X) * ddX[0][2];
            ddw[4] = ((ddN * D - dN[0] * dD[2] - dN[2] * dD[0] - N * ddD) * sq(D) + 2.0 * N * D * dD[0] * dD[2]) / pow(D, 4);

            ddw0 = dddsc0[0][1][2] * v0 + dddsc0[1][1][2] * u0 / cos(lat);
            ddw0 += sin(lat) / sq(cos(lat)) * ddsc0[2][2] * u0;
            ddw0 *= -2.0 * bp / (a->r0 + zg);
            ddN = ddsc0[1][2] * X * w0 + sc0 * ddX[1][2] * w0 + sc0 * X * ddw0;
            ddN += dsc0[1] * dX[2] * w0 + dsc0[1] * X * dw0[2] + sc0 * dX[1] * dw0[2];
            ddN += dsc0[2] * dX[1] * w0 + dsc0[2] * X * dw0[1] + sc0 * dX[2] * dw0[1];
            ddD = -sg_ratio * dX[1] * dX[2] + (1.0 - sg_ratio * X) * ddX[1][2];
            ddw[5] = ((ddN * D - dN[1] * dD[2] - dN[2] * dD[1] - N * ddD) * sq(D) + 2.0 * N * D * dD[1] * dD[2]) / pow(D, 4);
        }
    } else {
        *w = 0.0;
        for (int n = 0; n < 3; n++) dw[n] = 0.0;
        if (second) for (int n = 0; n < 6; n++) ddw[n] = 0.0;
    }
}

/* atmo::SB_alpha over the range-dependent c and rho (atmo_state.abs.sph.cpp:14-151). */
double go_rd_sb_alpha(go_rd *a, double r, double lat, double lon, double freq)
{
    return go_sb_alpha_at(go_rd_c(a, r, lat, lon), go_rd_rho(a, r, lat, lon), r - a->r0, 1, a->tweak_abs, freq);
}

/* ======================================================================
 * Ray equations: geoac.eqset.sph.rngdep.cpp
 * ==================================================================== */

/* set_initial, :68-112 */
void go_rd_set_initial(go_rd *a, const go_params *p, go_src *s, double *y0)
{
    double th = s->theta, ph = s->phi, *L = s->src;
    s->c0 = go_rd_c(a, L[0], L[1], L[2]);
    double mach[3] = {go_rd_w(a, L[0], L[1], L[2]) / s->c0, go_rd_v(a, L[0], L[1], L[2]) / s->c0,
                      go_rd_u(a, L[0], L[1], L[2]) / s->c0};
    double nu0[3] = {sin(th), cos(th) * sin(ph), cos(th) * cos(ph)};
    double mt[3] = {cos(th), -sin(th) * sin(ph), -sin(th) * cos(ph)};
    double mp[3] = {0.0, cos(th) * cos(ph), -cos(th) * sin(ph)};
    double msc = 1.0 + (nu0[0] * mach[0] + nu0[1] * mach[1] + nu0[2] * mach[2]);
    s->nu0 = 1.0 / msc;
    for (int i = 0; i < 3; i++) y0[i] = L[i];
    for (int i = 0; i < 3; i++) y0[3 + i] = nu0[i] / msc;
    if (p->eq_cnt == 18) {
        for (int i = 0; i < 3; i++) { y0[6 + i] = 0.0; y0[12 + i] = 0.0; }
        for (int i = 0; i < 3; i++) {
            y0[9 + i] = mt[i] / msc - nu0[i] / sq(msc) * (mt[0] * mach[0] + mt[1] * mach[1] + mt[2] * mach[2]);
            y0[15 + i] = mp[i] / msc - nu0[i] / sq(msc) * (mp[0] * mach[0] + mp[1] * mach[1] + mp[2] * mach[2]);
        }
    }
}

typedef struct {
    double c, dc[5], ddc[3][2];
    double u, du[5], ddu[3][2];
    double v, dv[5], ddv[3][2];
    double w, dw[5], ddw[3][2];
    double nu_mag, dnu_mag[2];
    double cg[3], cg_mag, dcg[3][2], dcg_mag[2];
    double G[3], dG[3][2], T[3], dT[3][2];
} rd_refs;

/* update_refs, :238-373 */
static void refs_rd(go_rd *a, const go_params *p, const double *y, rd_refs *R)
{
    double r = y[0], t = y[1], pp = y[2];
    const double *nu = y + 3;
    double dct[3], ddct[6], dd_u_t[6], dd_v_t[6], dd_w_t[6];
    double dd_c[3][3], dd_u[3][3], dd_v[3][3], dd_w[3][3];
    memset(R, 0, sizeof(*R));
    R->nu_mag = sqrt(sq(nu[0]) + sq(nu[1]) + sq(nu[2]));
    go_h3d_all(&a->c, t, pp, r - a->r0, &R->c, dct, ddct);
    R->dc[1] = dct[0]; R->dc[2] = dct[1]; R->dc[0] = dct[2];
    if (!p->calc_amp) {
        go_rd_calc_uvw(a, r, t, pp, &R->u, &R->v, &R->w, R->du, R->dv, R->dw, NULL, NULL, NULL);
    } else {
        dd_c[1][1] = ddct[0]; dd_c[2][2] = ddct[1]; dd_c[0][0] = ddct[2];
        dd_c[1][2] = ddct[3]; dd_c[2][1] = dd_c[1][2];
        dd_c[0][1] = ddct[4]; dd_c[1][0] = dd_c[0][1];
        dd_c[0][2] = ddct[5]; dd_c[2][0] = dd_c[0][2];
        go_rd_calc_uvw(a, r, t, pp, &R->u, &R->v, &R->w, R->du, R->dv, R->dw, dd_u_t, dd_v_t, dd_w_t);
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

    const double *wd[3] = {R->dw, R->dv, R->du};
    for (int i = 0; i < 3; i++)
        R->dcg[i][0] = nu[i] / R->nu_mag * R->dc[3] + R->c * mt[i] / R->nu_mag - R->c * nu[i] / sq(R->nu_mag) * R->dnu_mag[0] + wd[i][3];
    for (int i = 0; i < 3; i++)
        R->dcg[i][1] = nu[i] / R->nu_mag * R->dc[4] + R->c * mp[i] / R->nu_mag - R->c * nu[i] / sq(R->nu_mag) * R->dnu_mag[1] + wd[i][4];
    R->dcg_mag[0] = (R->cg[0] * R->dcg[0][0] + R->cg[1] * R->dcg[1][0] + R->cg[2] * R->dcg[2][0]) / R->cg_mag;
    R->dcg_mag[1] = (R->cg[0] * R->dcg[0][1] + R->cg[1] * R->dcg[1][1] + R->cg[2] * R->dcg[2][1]) / R->cg_mag;

    R->dG[0][0] = 0.0; R->dG[0][1] = 0.0;
    R->dG[1][0] = -Rt[0] / (sq(r));  R->dG[1][1] = -Rp[0] / (sq(r));
    R->dG[2][0] = -Rt[0] / (sq(r) * cos(t)) + sin(t) / (r * sq(cos(t))) * Rt[1];
    R->dG[2][1] = -Rp[0] / (sq(r) * cos(t)) + sin(t) / (r * sq(cos(t))) * Rp[1];

    for (int j = 0; j < 2; j++) {
        const double *Rj = j ? Rp : Rt, *mj = j ? mp : mt;
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

/* eval_src_eq, :378-436 */
static void eqs_rd(const go_params *p, const double *y, const rd_refs *R, double *f)
{
    const double *nu = y + 3;
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

void go_rd_rhs(go_rd *a, const go_params *p, const double *y, double *f)
{
    rd_refs R;
    refs_rd(a, p, y, &R);
    eqs_rd(p, y, &R, f);
}

/* set_ds, :227-233 */
double go_rd_set_ds(go_rd *a, const go_params *p, const double *y)
{
    double result = p->ds_max;
    if (y[0] - a->r0 < a->z_bndlyr)
        result -= (p->ds_max - p->ds_min) * exp(-(y[0] - go_rd_topo_z(a, y[1], y[2])) / 0.2);
    return result;
}

/* break_check, :483-494 (wraps longitude in place) */
int go_rd_break_check(go_rd *a, const go_params *p, const go_src *s, double *y)
{
    if (y[2] > RD_PI) y[2] -= 2.0 * RD_PI;
    else if (y[2] < -RD_PI) y[2] += 2.0 * RD_PI;
    if (y[1] < -RD_PI / 2.0 || y[1] > RD_PI / 2.0) return 1;
    if (y[0] - a->r0 > p->alt_max) return 1;
    if (go_gc_dist(a->r0, y[1], y[2], s->src[1], s->src[2]) > p->rng_max) return 1;
    if (y[1] < p->lat_min || y[1] > p->lat_max) return 1;
    if (y[2] < p->lon_min || y[2] > p->lon_max) return 1;
    return 0;
}

/* ground_check, :496-503 */
int go_rd_ground_check(go_rd *a, const double *y)
{
    if (y[0] - a->r0 <= a->z_max)
        if (y[0] < go_rd_topo_z(a, y[1], y[2])) return 1;
    return 0;
}

/* rk4solver.cpp:29-55 */
double go_rd_rk4_step(go_rd *a, const go_params *p, const double *y, double *y1)
{
    int n = p->eq_cnt;
    double t1[18], t2[18], t3[18], t4[18], pa[18], f[18];
    double ds = go_rd_set_ds(a, p, y);
    go_rd_rhs(a, p, y, f);
    for (int i = 0; i < n; i++) { t1[i] = ds * f[i]; pa[i] = y[i] + t1[i] / 2.0; }
    go_rd_rhs(a, p, pa, f);
    for (int i = 0; i < n; i++) t2[i] = ds * f[i];
    for (int i = 0; i < n; i++) pa[i] = y[i] + t2[i] / 2.0;
    go_rd_rhs(a, p, pa, f);
    for (int i = 0; i < n; i++) t3[i] = ds * f[i];
    for (int i = 0; i < n; i++) pa[i] = y[i] + t3[i];
    go_rd_rhs(a, p, pa, f);
    for (int i = 0; i < n; i++) {
        t4[i] = ds * f[i];
        y1[i] = y[i] + t1[i] / 6.0 + t2[i] / 3.0 + t3[i] / 3.0 + t4[i] / 6.0;
    }
    return ds;
}

/* rk4solver.cpp:14-85 */
int go_rd_prop_rk4(go_rd *a, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    int k;
    double sum_s = 0.0;
    *check = 0;
    for (k = 0; k < length - 1; k++) {
        double *y = rows + (size_t)k * 18, *y1 = rows + (size_t)(k + 1) * 18;
        double ds = go_rd_rk4_step(a, p, y, y1);
        sum_s += ds;
        if (sum_s > p->s_max) { *check = 1; break; }
        if (go_rd_break_check(a, p, s, y1)) { *check = 1; break; }
        if (go_rd_ground_check(a, y1)) { *check = 0; break; }
    }
    return k + 1;
}

/* approx_intercept + set_refl, :117-221.  Reference expressions kept as they
 * are, including tan(prev[4]) in db_dlth / db_dlph (:161, :164) and db_dlth
 * twice in dC2ph_dlth (:171). */
void go_rd_set_refl(go_rd *a, const go_params *p, const go_src *s, const double *rk2, const double *rk1, double *y0)
{
    int neq = p->eq_cnt;
    double prev[18] = {0};
    double dr_grnd = go_rd_topo_z(a, rk1[1], rk1[2]) - rk1[0];
    double dr_k = (rk1[0] - rk2[0])
                  - (rk1[1] - rk2[1]) * 1.0 / rk1[0] * go_rd_topo_dz(a, rk1[1], rk1[2], 0)
                  - (rk1[2] - rk2[2]) * 1.0 / (rk1[0] * cos(rk1[1])) * go_rd_topo_dz(a, rk1[1], rk1[2], 1);
    for (int i = 0; i < neq; i++) prev[i] = rk1[i] + (rk1[i] - rk2[i]) / dr_k * dr_grnd;

    double rg, c_grnd, aa, b, norm1, norm2, C1r, C1th, C1ph, C2th, C2ph;
    double da_dlth = 0, db_dlth = 0, da_dlph = 0, db_dlph = 0;
    double dC1r_dlth = 0, dC1th_dlth = 0, dC1ph_dlth = 0, dC2th_dlth = 0, dC2ph_dlth = 0, ds0_dlth = 0;
    double dC1r_dlph = 0, dC1th_dlph = 0, dC1ph_dlph = 0, dC2th_dlph = 0, dC2ph_dlph = 0, ds0_dlph = 0;
    double dnu_r_ds = 0, dnu_th_ds = 0, dnu_ph_ds = 0;

    rg = go_rd_topo_z(a, prev[1], prev[2]);
    c_grnd = go_rd_c(a, rg, prev[1], prev[2]);
    aa = 1.0 / rg * go_rd_topo_dz(a, prev[1], prev[2], 0);
    b = 1.0 / (rg * cos(prev[1])) * go_rd_topo_dz(a, prev[1], prev[2], 1);
    norm1 = 1.0 + sq(aa) + sq(b);
    C1r = (1.0 - sq(aa) - sq(b)) / norm1;
    C1th = (1.0 - sq(aa) + sq(b)) / norm1;  C2th = 2.0 * aa / norm1;
    C1ph = (1.0 + sq(aa) - sq(b)) / norm1;  C2ph = 2.0 * b / norm1;

    if (p->calc_amp) {
        double z00 = go_rd_topo_ddz(a, prev[1], prev[2], 0, 0), z01 = go_rd_topo_ddz(a, prev[1], prev[2], 0, 1);
        double z10 = go_rd_topo_ddz(a, prev[1], prev[2], 1, 0), z11 = go_rd_topo_ddz(a, prev[1], prev[2], 1, 1);
        norm2 = prev[3] - aa * prev[4] - b * prev[5];
        da_dlth = -prev[6] / rg * aa + 1.0 / rg * ((prev[7] * (prev[3] - b * prev[5]) - prev[4] * (prev[6] / rg - cos(prev[1]) * b * prev[8])) / norm2 * z00
                                                   + (prev[8] * (prev[3] - aa * prev[4]) - prev[5] * (prev[6] / (rg * cos(prev[1])) - aa / cos(prev[1]) * prev[7])) / norm2 * z01);
        da_dlph = -prev[12] / rg * aa + 1.0 / rg * ((prev[13] * (prev[3] - b * prev[5]) - prev[4] * (prev[12] / rg - cos(prev[1]) * b * prev[14])) / norm2 * z00
                                                    + (prev[14] * (prev[3] - aa * prev[4]) - prev[5] * (prev[12] / (rg * cos(prev[1])) - aa / cos(prev[1]) * prev[13])) / norm2 * z01);
        db_dlth = -(prev[6] / rg - tan(prev[4]) * prev[7]) * b + 1.0 / (rg * cos(prev[1])) * ((prev[7] * (prev[3] - b * prev[5]) - prev[4] * (prev[6] / rg - cos(prev[1]) * b * prev[8])) / norm2 * z10
                                                                                              + (prev[8] * (prev[3] - aa * prev[4]) - prev[5] * (prev[6] / (rg * cos(prev[1])) - aa / cos(prev[1]) * prev[7])) / norm2 * z11);
        db_dlph = -(prev[12] / rg - tan(prev[4]) * prev[13]) * b + 1.0 / (rg * cos(prev[1])) * ((prev[13] * (prev[3] - b * prev[5]) - prev[4] * (prev[12] / rg - cos(prev[1]) * b * prev[14])) / norm2 * z10
                                                                                                + (prev[14] * (prev[3] - aa * prev[4]) - prev[5] * (prev[12] / (rg * cos(prev[1])) - aa / cos(prev[1]) * prev[13])) / norm2 * z11);

        dC1r_dlth = -4.0 / sq(norm1) * (aa * da_dlth + b * db_dlth);
        dC1r_dlph = -4.0 / sq(norm1) * (aa * da_dlph + b * db_dlph);
        dC1th_dlth = -4.0 * aa / sq(norm1) * ((1.0 + sq(b)) * da_dlth - aa * b * db_dlth);
        dC1th_dlph = -4.0 * aa / sq(norm1) * ((1.0 + sq(b)) * da_dlph - aa * b * db_dlph);
        dC1ph_dlth = -4.0 * b / sq(norm1) * ((1.0 + sq(aa)) * db_dlth - aa * b * da_dlth);
        dC1ph_dlph = -4.0 * b / sq(norm1) * ((1.0 + sq(aa)) * db_dlph - aa * b * da_dlph);
        dC2th_dlth = 2.0 / sq(norm1) * ((1.0 - sq(aa) + sq(b)) * da_dlth - 2.0 * aa * b * db_dlth);
        dC2th_dlph = 2.0 / sq(norm1) * ((1.0 - sq(aa) + sq(b)) * da_dlph - 2.0 * aa * b * db_dlph);
        dC2ph_dlth = 2.0 / sq(norm1) * ((1.0 + sq(aa) - sq(b)) * db_dlth - 2.0 * aa * b * db_dlth);
        dC2ph_dlph = 2.0 / sq(norm1) * ((1.0 + sq(aa) - sq(b)) * db_dlph - 2.0 * aa * b * da_dlph);

        ds0_dlth = -s->c0 / c_grnd * (prev[6] - rg * aa * prev[7] - rg * cos(prev[1]) * b * prev[8]) / norm2;
        ds0_dlph = -s->c0 / c_grnd * (prev[12] - rg * aa * prev[13] - rg * cos(prev[1]) * b * prev[14]) / norm2;

        dnu_r_ds = s->c0 / c_grnd * go_rd_dc(a, rg, prev[1], prev[2], 0) + prev[3] * go_rd_dw(a, rg, prev[1], prev[2], 0)
                   + prev[4] * go_rd_dv(a, rg, prev[1], prev[2], 0) + prev[5] * go_rd_du(a, rg, prev[1], prev[2], 0);
        dnu_th_ds = s->c0 / c_grnd * go_rd_dc(a, rg, prev[1], prev[2], 1) + prev[3] * go_rd_dw(a, rg, prev[1], prev[2], 1)
                    + prev[4] * go_rd_dv(a, rg, prev[1], prev[2], 1) + prev[5] * go_rd_du(a, rg, prev[1], prev[2], 1);
        dnu_ph_ds = s->c0 / c_grnd * go_rd_dc(a, rg, prev[1], prev[2], 2) + prev[3] * go_rd_dw(a, rg, prev[1], prev[2], 2)
                    + prev[4] * go_rd_dv(a, rg, prev[1], prev[2], 2) + prev[5] * go_rd_du(a, rg, prev[1], prev[2], 2);
        dnu_r_ds += sq(c_grnd) / s->c0 * 1.0 / rg * (sq(prev[4]) + sq(prev[5]));
        dnu_th_ds -= sq(c_grnd) / s->c0 * (prev[3] * prev[4] + tan(prev[1]) * sq(prev[5]));
        dnu_ph_ds -= sq(c_grnd) / s->c0 * prev[5] * (prev[3] * cos(prev[1]) + prev[4] * sin(prev[1]));
        dnu_r_ds *= -1.0 / c_grnd;
        dnu_th_ds *= -1.0 / (c_grnd * rg);
        dnu_ph_ds *= -1.0 / (c_grnd * rg * cos(prev[1]));
    }

    y0[0] = rg;
    y0[1] = prev[1];
    y0[2] = prev[2];
    y0[3] = -C1r * prev[3] + C2th * prev[4] + C2ph * prev[5];
    y0[4] = C1th * prev[4] + C2th * (prev[3] - b * prev[5]);
    y0[5] = C1ph * prev[5] + C2ph * (prev[3] - aa * prev[4]);
    if (neq < 18) return;
    for (int j = 0; j < 2; j++) {
        int o = j ? 12 : 6;
        y0[o] = -C1r * prev[o] + rg * C2th * prev[o + 1] + rg * cos(prev[1]) * C2ph * prev[o + 2];
        y0[o + 1] = C1th * prev[o + 1] + C2th * (prev[o] / rg - b * cos(prev[1]) * prev[o + 2]);
        y0[o + 2] = C1ph * prev[o + 2] + C2ph * (prev[o] / (rg * cos(prev[1])) - aa / cos(prev[1]) * prev[o + 1]);
    }
    y0[9] = -C1r * prev[9] + C2th * prev[10] + C2ph * prev[11] - dC1r_dlth * prev[3] + dC2th_dlth * prev[4] + dC2ph_dlth * prev[5]
            + (-(C1r + 1.0) * dnu_r_ds + C2th * dnu_th_ds + C2ph * dnu_ph_ds) * ds0_dlth;
    y0[10] = C1th * prev[10] + C2th * (prev[9] - b * prev[11] - db_dlth * prev[5]) + dC1th_dlth * prev[4] + dC2th_dlth * (prev[3] - b * prev[5])
             + ((C1th - 1.0) * dnu_th_ds + C2th * (dnu_r_ds - b * dnu_ph_ds)) * ds0_dlth;
    y0[11] = C1ph * prev[11] + C2ph * (prev[9] - aa * prev[10] - da_dlth * prev[4]) + dC1ph_dlth * prev[5] + dC2ph_dlth * (prev[3] - aa * prev[4])
             + ((C1ph - 1.0) * dnu_ph_ds + C2ph * (dnu_r_ds - aa * dnu_th_ds)) * ds0_dlth;
    y0[15] = -C1r * prev[15] + C2th * prev[16] + C2ph * prev[17] - dC1r_dlph * prev[3] + dC2th_dlph * prev[4] + dC2ph_dlph * prev[5]
             + (-(C1r + 1.0) * dnu_r_ds + C2th * dnu_th_ds + C2ph * dnu_ph_ds) * ds0_dlph;
    y0[16] = C1th * prev[16] + C2th * (prev[15] - b * prev[17] - db_dlph * prev[5]) + dC1th_dlph * prev[4] + dC2th_dlph * (prev[3] - b * prev[5])
             + ((C1th - 1.0) * dnu_th_ds + C2th * (dnu_r_ds - b * dnu_ph_ds)) * ds0_dlph;
    y0[17] = C1ph * prev[17] + C2ph * (prev[15] - aa * prev[16] - da_dlph * prev[4]) + dC1ph_dlph * prev[5] + dC2ph_dlph * (prev[3] - aa * prev[4])
             + ((C1ph - 1.0) * dnu_ph_ds + C2ph * (dnu_r_ds - aa * dnu_th_ds)) * ds0_dlph;
}

/* travel_time (segment form of :547-575) */
double go_rd_travel_time_seg(go_rd *a, const double *y0, const double *y1)
{
    double dr = y1[0] - y0[0], r = y0[0] + dr / 2.0;
    double dt = y1[1] - y0[1], t = y0[1] + dt / 2.0;
    double dp = y1[2] - y0[2], pp = y0[2] + dp / 2.0;
    double nu[3], nu_mag, cp[3], cp_mag, snd, ds;
    if (dp > RD_PI) dp -= 2.0 * RD_PI;
    else if (dp < -RD_PI) dp += 2.0 * RD_PI;
    ds = sqrt(sq(dr) + sq(r * dt) + sq(r * cos(t) * dp));
    for (int i = 0; i < 3; i++) nu[i] = y0[3 + i] + (y1[3 + i] - y0[3 + i]) / 2.0;
    nu_mag = sqrt(sq(nu[0]) + sq(nu[1]) + sq(nu[2]));
    snd = go_rd_c(a, r, t, pp);
    cp[0] = snd * nu[0] / nu_mag + go_rd_w(a, r, t, pp);
    cp[1] = snd * nu[1] / nu_mag + go_rd_v(a, r, t, pp);
    cp[2] = snd * nu[2] / nu_mag + go_rd_u(a, r, t, pp);
    cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    return ds / cp_mag;
}

/* atten (segment form of :775-792) */
double go_rd_atten_seg(go_rd *a, const double *y0, const double *y1, double freq)
{
    double dr = y1[0] - y0[0], r = y0[0] + dr / 2.0;
    double dt = y1[1] - y0[1], t = y0[1] + dt / 2.0;
    double dp = y1[2] - y0[2], pp = y0[2] + dp / 2.0;
    if (dp > RD_PI) dp -= 2.0 * RD_PI;
    else if (dp < -RD_PI) dp += 2.0 * RD_PI;
    double ds = sqrt(sq(dr) + sq(r * dt) + sq(r * cos(t) * dp));
    return go_rd_sb_alpha(a, r, t, pp, freq) * ds;
}

/* jacobian, :798-821 */
double go_rd_jacobian(go_rd *a, const double *y)
{
    double r = y[0], t = y[1], p = y[2];
    double nu_mag = sqrt(sq(y[3]) + sq(y[4]) + sq(y[5]));
    double cp[3], cp_mag;
    cp[0] = go_rd_c(a, r, t, p) * y[3] / nu_mag + go_rd_w(a, r, t, p);
    cp[1] = go_rd_c(a, r, t, p) * y[4] / nu_mag + go_rd_v(a, r, t, p);
    cp[2] = go_rd_c(a, r, t, p) * y[5] / nu_mag + go_rd_u(a, r, t, p);
    cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    double dr_ds = cp[0] / cp_mag, dt_ds = 1.0 / r * cp[1] / cp_mag, dp_ds = 1.0 / (r * cos(t)) * cp[2] / cp_mag;
    double dr_t = y[6], dt_t = y[7], dp_t = y[8], dr_p = y[12], dt_p = y[13], dp_p = y[14];
    return sq(r) * cos(t) * (dr_ds * (dt_t * dp_p - dt_p * dp_t) - dr_t * (dt_ds * dp_p - dp_ds * dt_p)
                             + dr_p * (dt_ds * dp_t - dp_ds * dt_t));
}

/* amp, :823-847 */
double go_rd_amp(go_rd *a, const go_src *s, const double *y)
{
    double r = y[0], t = y[1], p = y[2];
    double r0 = s->src[0], t0 = s->src[1], p0 = s->src[2];
    double nu0[3] = {s->nu0 * sin(s->theta), s->nu0 * cos(s->theta) * sin(s->phi), s->nu0 * cos(s->theta) * cos(s->phi)};
    double nu_mag = (s->c0 - y[3] * go_rd_w(a, r, t, p) - y[4] * go_rd_v(a, r, t, p) - y[5] * go_rd_u(a, r, t, p)) / go_rd_c(a, r, t, p);
    double nu_mag0 = s->nu0;
    double c0p[3], cp[3];
    c0p[0] = s->c0 * nu0[0] / nu_mag0 + go_rd_w(a, r0, t0, p0);  cp[0] = go_rd_c(a, r, t, p) * y[3] / nu_mag + go_rd_w(a, r, t, p);
    c0p[1] = s->c0 * nu0[1] / nu_mag0 + go_rd_v(a, r0, t0, p0);  cp[1] = go_rd_c(a, r, t, p) * y[4] / nu_mag + go_rd_v(a, r, t, p);
    c0p[2] = s->c0 * nu0[2] / nu_mag0 + go_rd_u(a, r0, t0, p0);  cp[2] = go_rd_c(a, r, t, p) * y[5] / nu_mag + go_rd_u(a, r, t, p);
    double cp_mag = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    double cp_mag0 = sqrt(sq(c0p[0]) + sq(c0p[1]) + sq(c0p[2]));
    double num = go_rd_rho(a, r, t, p) * nu_mag * pow(go_rd_c(a, r, t, p), 3) * cp_mag0 * cos(s->theta);
    double den = go_rd_rho(a, r0, t0, p0) * nu_mag0 * pow(go_rd_c(a, r0, t0, p0), 3) * cp_mag * go_rd_jacobian(a, y);
    return sqrt(fabs(num / den));
}

/* eval_eikonal, :441-447 */
double go_rd_eikonal_resid(go_rd *a, const go_src *s, const double *y)
{
    double r = y[0], t = y[1], p = y[2];
    return sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]) - s->c0 / go_rd_c(a, r, t, p)
           + (go_rd_w(a, r, t, p) * y[3] + go_rd_v(a, r, t, p) * y[4] + go_rd_u(a, r, t, p) * y[5]) / go_rd_c(a, r, t, p);
}

/* eval_eikonal_deriv, :457-478: launch-angle derivatives of the eikonal
 * residual carried by the auxiliary variables. */
double go_rd_eikonal_deriv_resid(go_rd *a, const double *y)
{
    double r = y[0], t = y[1], p = y[2];
    const double *nu = y + 3, *Rt = y + 6, *mt = y + 9, *Rp = y + 12, *mp = y + 15;
    double mag_nu = sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
    double dc_t = 0, dc_p = 0, du_t = 0, du_p = 0, dv_t = 0, dv_p = 0, dw_t = 0, dw_p = 0;
    for (int n = 0; n < 3; n++) {
        double dc = go_rd_dc(a, r, t, p, n), du = go_rd_du(a, r, t, p, n);
        double dv = go_rd_dv(a, r, t, p, n), dw = go_rd_dw(a, r, t, p, n);
        dc_t += Rt[n] * dc; dc_p += Rp[n] * dc;
        du_t += Rt[n] * du; du_p += Rp[n] * du;
        dv_t += Rt[n] * dv; dv_p += Rp[n] * dv;
        dw_t += Rt[n] * dw; dw_p += Rp[n] * dw;
    }
    double c = go_rd_c(a, r, t, p), u = go_rd_u(a, r, t, p), v = go_rd_v(a, r, t, p), w = go_rd_w(a, r, t, p);
    double res_t = (nu[0] * mt[0] + nu[1] * mt[1] + nu[2] * mt[2]) / mag_nu + mag_nu / c * dc_t
                   + 1.0 / c * (mt[0] * w + mt[1] * v + mt[2] * u + nu[0] * dw_t + nu[1] * dv_t + nu[2] * du_t);
    double res_p = (nu[0] * mp[0] + nu[1] * mp[1] + nu[2] * mp[2]) / mag_nu + mag_nu / c * dc_p
                   + 1.0 / c * (mp[0] * w + mp[1] * v + mp[2] * u + nu[0] * dw_p + nu[1] * dv_p + nu[2] * du_p);
    return sqrt(pow(res_t, 2) + pow(res_p, 2));
}

/* ======================================================================
 * run_prop: main.infraga.sph.rngdep.cpp:311-554
 * ==================================================================== */
double go_rd_theta_grnd(go_rd *a, double lat_src, double lon_src, double z_src, double az_deg)
{
    double phi = az_deg, dzgdx, dzgdy;
    if (z_src - (go_rd_topo_z(a, lat_src, lon_src) - a->r0) < 0.01) {
        if (a->is_topo) {
            dzgdx = 1.0 / (a->r0 * cos(lat_src)) * go_rd_topo_dz(a, lat_src, lon_src, 1);
            dzgdy = 1.0 / a->r0 * go_rd_topo_dz(a, lat_src, lon_src, 0);
            return atan(dzgdx * sin(phi * RD_PI / 180.0) + dzgdy * cos(phi * RD_PI / 180.0)) * (180.0 / RD_PI) + 0.1;
        }
        return 0.1;
    }
    return -89.9;
}

static double *rd_rows = NULL;
static int rd_rows_len = 0;

int go_rd_fan(go_rd *a, const go_params *p, const go_prop *pr, const double *launches, int n_launch,
              go_arrival *out, int out_cap, int *n_out, long long *ray_steps)
{
    int length = (int)(p->s_max / p->ds_min);
    if (rd_rows_len < length) {
        free(rd_rows);
        rd_rows = (double *)malloc(sizeof(double) * 18 * (size_t)length);
        rd_rows_len = length;
    }
    double *rows = rd_rows;
    go_src s;
    int cnt = 0;
    long long steps = 0;
    double lat_src = pr->src0, lon_src = pr->src1, z_src = pr->src_alt;
    lat_src *= RD_PI / 180.0;                                          /* :311-312 */
    lon_src *= RD_PI / 180.0;
    z_src = dmax(z_src, go_rd_topo_z(a, lat_src, lon_src) - a->r0);     /* :315 */
    s.src[0] = a->r0 + z_src; s.src[1] = lat_src; s.src[2] = lon_src;

    for (int li = 0; li < n_launch; li++) {
        double incl = launches[2 * li], az = launches[2 * li + 1];
        s.phi = RD_PI / 2.0 - az * (RD_PI / 180.0);
        s.theta = incl * (RD_PI / 180.0);
        go_rd_set_initial(a, p, &s, rows);
        double tt = 0.0, att = 0.0, rmax = 0.0;
        int brk = 0, k;
        for (int bnc = 0; bnc <= pr->bounces; bnc++) {
            k = go_rd_prop_rk4(a, p, &s, rows, length, &brk);
            steps += k;
            double tleg = 0.0, aleg = 0.0;
            for (int n = 0; n < k; n++) tleg += go_rd_travel_time_seg(a, rows + 18 * n, rows + 18 * (n + 1));
            tt += tleg;
            for (int n = 0; n < k; n++) aleg += go_rd_atten_seg(a, rows + 18 * n, rows + 18 * (n + 1), pr->freq);
            att += aleg;
            for (int m = 0; m < k; m++) rmax = dmax(rmax, rows[18 * m] - a->r0);
            if (!a->is_topo && rmax < (a->z0 + pr->turn_ht_min)) break;     /* :507 */
            if (brk || k < 2 || (tt / 3600.0) > p->time_max) break;         /* :511 */
            double *yk = rows + 18 * k;
            go_arrival A;
            memset(&A, 0, sizeof(A));
            A.launch_idx = li;
            A.n_bounce = bnc;
            A.incl_launch = incl;
            A.az_launch = az;
            /* :515, no leading minus sign; the stratified driver has one
             * (main.infraga.sph.cpp:523) */
            const double sn = asin(go_rd_c(a, yk[0], yk[1], yk[2]) / go_rd_c(a, a->r0 + z_src, lat_src, lon_src) * yk[3]);
            double baz;
            if (a->strat) {   /* main.infraga.sph.cpp:523-524 */
                A.incl_arr = -sn * (180.0 / RD_PI);
                baz = 90.0 - atan2(-yk[4], -yk[5]) * (180.0 / RD_PI);
            } else {
                A.incl_arr = sn * 180.0 / RD_PI;
                baz = 90.0 - atan2(-yk[4], -yk[5]) * 180.0 / RD_PI;
            }
            if (baz < -180.0) baz += 360.0;
            else if (baz > 180.0) baz -= 360.0;
            A.back_az = baz;
            A.pos0 = yk[1] * (180.0 / RD_PI);
            A.pos1 = yk[2] * (180.0 / RD_PI);
            A.time = tt;
            A.cel = go_gc_dist(a->r0, yk[1], yk[2], lat_src, lon_src) / tt;
            A.turn_ht = rmax;
            A.tl = p->calc_amp ? 20.0 * log10(go_rd_amp(a, &s, yk)) : 0.0;
            A.absorption = -2.0 * att;
            if (cnt < out_cap) out[cnt] = A;
            cnt++;
            go_rd_set_refl(a, p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
        }
    }
    *n_out = cnt;
    if (ray_steps) *ray_steps = steps;
    return cnt > out_cap ? -1 : 0;
}

static void rd_pt(go_path_pt *pts, long long cap, long long *n, const go_rd *a, int li, int leg, int m, int kind,
                  const double *y, double tl, double att, double tt)
{
    if (*n < cap) {
        go_path_pt *P = pts + *n;
        P->launch_idx = li;
        P->leg = leg;
        P->step = m;
        P->kind = kind;
        P->pos0 = y[1] * (180.0 / RD_PI);
        P->pos1 = y[2] * (180.0 / RD_PI);
        P->z = y[0] - a->r0;
        P->tl = tl;
        P->absorption = -2.0 * att;
        P->time = tt;
    }
    (*n)++;
}

int go_rd_fan_paths(go_rd *a, const go_params *p, const go_prop *pr, const double *launches, int n_launch,
                    int flags, go_arrival *out, int out_cap, int *n_out, go_path_pt *pts, long long pts_cap,
                    long long *n_pts, long long *ray_steps)
{
    const int rays = (flags & GO_PATH_RAYS) != 0, caus = (flags & GO_PATH_CAUSTICS) != 0 && p->calc_amp;
    int length = (int)(p->s_max / p->ds_min);
    if (rd_rows_len < length) {
        free(rd_rows);
        rd_rows = (double *)malloc(sizeof(double) * 18 * (size_t)length);
        rd_rows_len = length;
    }
    double *rows = rd_rows;
    go_src s;
    int cnt = 0;
    long long steps = 0, np = 0;
    double lat_src = pr->src0, lon_src = pr->src1, z_src = pr->src_alt;
    lat_src *= RD_PI / 180.0;
    lon_src *= RD_PI / 180.0;
    z_src = dmax(z_src, go_rd_topo_z(a, lat_src, lon_src) - a->r0);
    s.src[0] = a->r0 + z_src; s.src[1] = lat_src; s.src[2] = lon_src;

    for (int li = 0; li < n_launch; li++) {
        double incl = launches[2 * li], az = launches[2 * li + 1];
        s.phi = RD_PI / 2.0 - az * (RD_PI / 180.0);
        s.theta = incl * (RD_PI / 180.0);
        go_rd_set_initial(a, p, &s, rows);
        double tt = 0.0, att = 0.0, rmax = 0.0;
        int brk = 0, k;
        for (int bnc = 0; bnc <= pr->bounces; bnc++) {
            k = go_rd_prop_rk4(a, p, &s, rows, length, &brk);
            steps += k;
            /* :451-490 */
            double D = 0.0, D_prev = caus ? go_rd_jacobian(a, rows + 18) : 0.0;
            for (int m = 1; m < k; m++) {
                const double *ym = rows + 18 * m;
                if (caus) D = go_rd_jacobian(a, ym);
                tt += go_rd_travel_time_seg(a, rows + 18 * (m - 1), ym);
                att += go_rd_atten_seg(a, rows + 18 * (m - 1), ym, pr->freq);
                if (rays && (m == 1 || m % 15 == 0))
                    rd_pt(pts, pts_cap, &np, a, li, bnc, m, 0, ym,
                          p->calc_amp ? 20.0 * log10(go_rd_amp(a, &s, ym)) : 0.0, att, tt);
                if (caus && D * D_prev < 0.0) rd_pt(pts, pts_cap, &np, a, li, bnc, m, 1, ym, 0.0, att, tt);
                if (caus) D_prev = D;
            }
            for (int m = 0; m < k; m++) rmax = dmax(rmax, rows[18 * m] - a->r0);
            if (!a->is_topo && rmax < (a->z0 + pr->turn_ht_min)) break;
            if (brk || k < 2 || (tt / 3600.0) > p->time_max) break;
            double *yk = rows + 18 * k;
            go_arrival A;
            memset(&A, 0, sizeof(A));
            A.launch_idx = li;
            A.n_bounce = bnc;
            A.incl_launch = incl;
            A.az_launch = az;
            const double sn = asin(go_rd_c(a, yk[0], yk[1], yk[2]) / go_rd_c(a, a->r0 + z_src, lat_src, lon_src) * yk[3]);
            double baz;
            if (a->strat) {   /* main.infraga.sph.cpp:523-524 */
                A.incl_arr = -sn * (180.0 / RD_PI);
                baz = 90.0 - atan2(-yk[4], -yk[5]) * (180.0 / RD_PI);
            } else {
                A.incl_arr = sn * 180.0 / RD_PI;
                baz = 90.0 - atan2(-yk[4], -yk[5]) * 180.0 / RD_PI;
            }
            if (baz < -180.0) baz += 360.0;
            else if (baz > 180.0) baz -= 360.0;
            A.back_az = baz;
            A.pos0 = yk[1] * (180.0 / RD_PI);
            A.pos1 = yk[2] * (180.0 / RD_PI);
            A.time = tt;
            A.cel = go_gc_dist(a->r0, yk[1], yk[2], lat_src, lon_src) / tt;
            A.turn_ht = rmax;
            A.tl = p->calc_amp ? 20.0 * log10(go_rd_amp(a, &s, yk)) : 0.0;
            A.absorption = -2.0 * att;
            if (cnt < out_cap) out[cnt] = A;
            cnt++;
            go_rd_set_refl(a, p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
        }
    }
    *n_out = cnt;
    if (n_pts) *n_pts = np;
    if (ray_steps) *ray_steps = steps;
    return (cnt > out_cap || np > pts_cap) ? -1 : 0;
}

/* ---- eigenray search (infraga-sph-rngdep -eig_search / -eig_direct) ---- */
static void rm_init(void *m, const go_params *p, go_src *s, double *y) { go_rd_set_initial((go_rd *)m, p, s, y); }
static int rm_prop(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    return go_rd_prop_rk4((go_rd *)m, p, s, rows, length, check);
}
static void rm_refl(void *m, const go_params *p, const go_src *s, const double *a, const double *b, double *o)
{
    go_rd_set_refl((go_rd *)m, p, s, a, b, o);
}
static double rm_tt(void *m, const double *a, const double *b) { return go_rd_travel_time_seg((go_rd *)m, a, b); }
static double rm_att(void *m, const double *a, const double *b, double f) { return go_rd_atten_seg((go_rd *)m, a, b, f); }
static double rm_c(void *m, double r, double t, double p) { return go_rd_c((go_rd *)m, r, t, p); }
static double rm_amp(void *m, const go_src *s, const double *y) { return go_rd_amp((go_rd *)m, s, y); }
static double rm_tz(void *m, double t, double p) { return go_rd_topo_z((go_rd *)m, t, p); }
static double rm_tdz(void *m, double t, double p, int n) { return go_rd_topo_dz((go_rd *)m, t, p, n); }

static go_eig_med rd_med(go_rd *a)
{
    go_eig_med M = {a, a->r0, rm_init, rm_prop, rm_refl, rm_tt, rm_att, rm_c, rm_amp, rm_tz, rm_tdz};
    return M;
}

/* Source conversion of run_eig_search (main.infraga.sph.rngdep.cpp:877-880):
 * degrees to radians, altitude raised to the terrain. */
static void rd_eig_src(go_rd *a, const double src_in[3], const double rcvr_in[2], double src[3], double rcvr[2])
{
    src[1] = src_in[0] * (RD_PI / 180.0);
    src[2] = src_in[1] * (RD_PI / 180.0);
    rcvr[0] = rcvr_in[0] * (RD_PI / 180.0);
    rcvr[1] = rcvr_in[1] * (RD_PI / 180.0);
    src[0] = dmax(src_in[2], go_rd_topo_z(a, src[1], src[2]) - a->r0);
}

/* theta_grnd of run_eig_search (:938-954): with terrain the largest slope
 * angle over azimuths 0, 3, ..., 357 in DEGREES plus 0.1, compared with the
 * radian inclination bounds as it stands; without terrain 0.1 deg in
 * radians; -89.9 deg in radians for an elevated source. */
double go_rd_eig_theta_grnd(go_rd *a, double src_alt, double lat, double lon)
{
    if (src_alt - (go_rd_topo_z(a, lat, lon) - a->r0) < 0.01) {
        if (a->is_topo) {
            double th = 0.0;
            for (double ph = 0.0; ph < 360.0; ph += 3.0) {
                double dzgdx = 1.0 / (a->r0 * cos(lat)) * go_rd_topo_dz(a, lat, lon, 1);
                double dzgdy = 1.0 / a->r0 * go_rd_topo_dz(a, lat, lon, 0);
                th = dmax(th, atan(dzgdx * sin(ph * RD_PI / 180.0) + dzgdy * cos(ph * RD_PI / 180.0)) * (180.0 / RD_PI) + 0.1);
            }
            return th;
        }
        return 0.1 * (RD_PI / 180.0);
    }
    return -89.9 * (RD_PI / 180.0);
}

int go_rd_eig_search(go_rd *a, const go_params *p, const double src_in[3], const double rcvr_in[2],
                     double th_min_deg, double th_max_deg, int bnc_min, int bnc_max, int iterations,
                     double damping, double tolerance, double az_err_lim, double dth_big_deg, double dth_sml_deg,
                     double freq, go_arrival *out, int out_cap, long long *ray_steps)
{
    double src[3], rcvr[2];
    rd_eig_src(a, src_in, rcvr_in, src, rcvr);
    go_eig_med M = rd_med(a);
    double th_grnd = go_rd_eig_theta_grnd(a, src[0], src[1], src[2]);
    return go_eig_search_sph_med(&M, p, src, rcvr, th_min_deg * (RD_PI / 180.0), th_max_deg * (RD_PI / 180.0),
                                 th_grnd, bnc_min, bnc_max, iterations, damping, tolerance, az_err_lim, dth_big_deg,
                                 dth_sml_deg, freq, out, out_cap, ray_steps);
}

long long go_rd_eig_path(go_rd *a, const go_params *p, const double src_in[3], double incl_deg, double az_deg,
                         int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A)
{
    double src[3], rcvr[2], dummy[2] = {0.0, 0.0};
    rd_eig_src(a, src_in, dummy, src, rcvr);
    go_eig_med M = rd_med(a);
    return go_eig_path_sph_med(&M, p, src, incl_deg * (RD_PI / 180.0), (90.0 - az_deg) * (RD_PI / 180.0), bnc_cnt,
                               freq, pts, cap, A);
}

/* ---- handles ------------------------------------------------------------ */
void *go_rd_new(const double *lat_deg, int nlat, const double *lon_deg, int nlon, const double *z, int nz,
                const double *c, const double *u, const double *v, const double *rho)
{
    go_rd *a = (go_rd *)malloc(sizeof(go_rd));
    if (go_rd_init(a, lat_deg, nlat, lon_deg, nlon, z, nz, c, u, v, rho)) { free(a); return NULL; }
    return a;
}
void go_rd_delete(void *h) { if (h) { go_rd_free((go_rd *)h); free(h); } }
void go_rd_h_config(void *h, int use_blw, double tweak_abs)
{
    go_rd *a = (go_rd *)h;
    a->use_blw = use_blw;
    a->tweak_abs = tweak_abs;
}
/* A stratified profile over terrain, as infraga-sph / infraga-3d run it
 * with topo_file (atmo_io.sph.strat.cpp:147-200, atmo_state.sph.strat.cpp):
 * the same medium as this restatement over identical columns (the
 * eigenray sources, the reflection and the RHS are shared or identical), but
 * z_max / z_bndlyr from the stratified set_bndlyr and the stratified
 * driver's arrival inclination sign. */
void go_rd_h_set_strat(void *h, int strat)
{
    go_rd *a = (go_rd *)h;
    a->strat = strat != 0;
    a->upper = a->strat && a->cart;
    set_bndlyr(a);
}
/* The contraction alone, for checks of its effect on one medium. */
void go_rd_h_set_upper(void *h, int on) { ((go_rd *)h)->upper = on != 0; }
void go_rd_h_limits(void *h, double out[8])
{
    go_rd *a = (go_rd *)h;
    out[0] = a->lat_min; out[1] = a->lat_max; out[2] = a->lon_min; out[3] = a->lon_max;
    out[4] = a->alt_max; out[5] = a->z0; out[6] = a->z_max; out[7] = a->z_bndlyr;
}

/* ======================================================================
 * 3D Cartesian range-dependent geometry (infraga-3d-rngdep):
 * geoac.eqset.3d.rngdep.cpp and run_prop of main.infraga.3d.rngdep.cpp over
 * the medium above with r0 = 0, node coordinates in km, and (z, x, y) in
 * the places of (r, lat, lon) (atmo_state.3d.rngdep.cpp has the same
 * functions with those arguments; its boundary-layer vertical wind, whose
 * geometric factors differ, is rd3_w / rd3_dw and the Cartesian block of
 * go_rd_calc_uvw).  State rows
 * are the reference's (x, y, z, nu, X_th, mu_th, X_ph, mu_ph).
 * ==================================================================== */
static int mi3(int n) { return n == 2 ? 0 : n + 1; }   /* Cartesian -> medium derivative index */

int go_rd_init_xy(go_rd *a, const double *x_km, int nx, const double *y_km, int ny, const double *z, int nz,
                  const double *c, const double *u, const double *v, const double *rho)
{
    memset(a, 0, sizeof(*a));
    if (nx < 2 || ny < 2 || nz < 3) return -1;
    int rc = go_h3d_init(&a->c, x_km, nx, y_km, ny, z, nz, c);   /* atmo_io.3d.rngdep.cpp:93-160 */
    rc |= go_h3d_init(&a->u, x_km, nx, y_km, ny, z, nz, u);
    rc |= go_h3d_init(&a->v, x_km, nx, y_km, ny, z, nz, v);
    rc |= go_h3d_init(&a->rho, x_km, nx, y_km, ny, z, nz, rho);
    if (rc) return -1;
    a->cart = 1;
    a->r0 = 0.0;
    a->bl_param = 0.1;             /* atmo_state.3d.rngdep.cpp:81 */
    a->vert_wind_grad = 1.0;       /* :82 */
    a->tweak_abs = 1.0;
    set_limits(a);
    set_bndlyr(a);
    return 0;
}

typedef struct {
    double c, dc[5], u, du[5], v, dv[5], w, dw[5];
    double ddc[3][2], ddu[3][2], ddv[3][2], ddw[3][2];
    double nu_mag, dnu_mag[2], cg[3], cg_mag, dcg[3][2], dcg_mag[2];
} r3_refs;

/* update_refs, geoac.eqset.3d.rngdep.cpp:226-323 */
static void refs3(go_rd *a, const go_params *p, const double *y, r3_refs *R)
{
    const double *nu = y + 3;
    double xe, ye, ze, du_t[3], dv_t[3], dw_t[3], ddu_t[6], ddv_t[6], ddw_t[6];
    double dd_c[3][3], dd_u[3][3], dd_v[3][3], dd_w[3][3];
    memset(R, 0, sizeof(*R));
    R->nu_mag = sqrt(sq(nu[0]) + sq(nu[1]) + sq(nu[2]));
    ev(a, &a->c, y[2], y[0], y[1], &xe, &ye, &ze);
    R->c = go_h3d_f(&a->c, xe, ye, ze);                                  /* atmo::c, :148-154 */
    for (int n = 0; n < 3; n++) R->dc[n] = go_h3d_df(&a->c, xe, ye, ze, n);
    go_rd_calc_uvw(a, y[2], y[0], y[1], &R->u, &R->v, &R->w, du_t, dv_t, dw_t,
                   p->calc_amp ? ddu_t : NULL, p->calc_amp ? ddv_t : NULL, p->calc_amp ? ddw_t : NULL);
    for (int n = 0; n < 3; n++) { R->du[n] = du_t[mi3(n)]; R->dv[n] = dv_t[mi3(n)]; R->dw[n] = dw_t[mi3(n)]; }
    for (int n = 0; n < 3; n++) R->cg[n] = R->c * nu[n] / R->nu_mag + (n == 0 ? R->u : (n == 1 ? R->v : R->w));
    R->cg_mag = sqrt(sq(R->cg[0]) + sq(R->cg[1]) + sq(R->cg[2]));
    if (!p->calc_amp) return;

    /* packed medium second derivatives: zz, xx, yy, zx, zy, xy */
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) {
            int i = mi3(n), j = mi3(m), q = (i == j) ? i : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5));
            dd_c[n][m] = go_h3d_ddf(&a->c, xe, ye, ze, n, m);
            dd_u[n][m] = ddu_t[q]; dd_v[n][m] = ddv_t[q]; dd_w[n][m] = ddw_t[q];
        }
    const double *Xt = y + 6, *mt = y + 9, *Xp = y + 12, *mp = y + 15;
    for (int n = 0; n < 3; n++) {
        R->dc[3] += Xt[n] * R->dc[n]; R->du[3] += Xt[n] * R->du[n]; R->dv[3] += Xt[n] * R->dv[n]; R->dw[3] += Xt[n] * R->dw[n];
        R->dc[4] += Xp[n] * R->dc[n]; R->du[4] += Xp[n] * R->du[n]; R->dv[3] += Xp[n] * R->dv[n]; R->dw[4] += Xp[n] * R->dw[n];   /* :288 */
        /* the stratified 3D set over terrain contracts the upper triangle only
         * (geoac.eqset.3d.strat.cpp:305-310, m = n..2) */
        for (int m = a->upper ? n : 0; m < 3; m++) {
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
    for (int j = 0; j < 2; j++)
        R->dcg_mag[j] = (R->cg[0] * R->dcg[0][j] + R->cg[1] * R->dcg[1][j] + R->cg[2] * R->dcg[2][j]) / R->cg_mag;
}

/* eval_src_eq, :329-382 */
void go_rd3_rhs(go_rd *a, const go_params *p, const double *y, double *f)
{
    r3_refs R;
    refs3(a, p, y, &R);
    const double *nu = y + 3;
    for (int n = 0; n < 3; n++) {
        f[n] = R.cg[n] / R.cg_mag;
        f[3 + n] = -1.0 / R.cg_mag * (R.nu_mag * R.dc[n] + nu[0] * R.du[n] + nu[1] * R.dv[n] + nu[2] * R.dw[n]);
    }
    if (p->eq_cnt < 18) return;
    for (int j = 0; j < 2; j++) {
        const double *mj = y + (j ? 15 : 9);
        for (int n = 0; n < 3; n++) {
            f[(j ? 12 : 6) + n] = R.dcg[n][j] / R.cg_mag - R.cg[n] / sq(R.cg_mag) * R.dcg_mag[j];
            f[(j ? 15 : 9) + n] = 1.0 / sq(R.cg_mag) * R.dcg_mag[j] * (R.nu_mag * R.dc[n] + nu[0] * R.du[n] + nu[1] * R.dv[n] + nu[2] * R.dw[n])
                                  - 1.0 / R.cg_mag * (R.dnu_mag[j] * R.dc[n] + R.nu_mag * R.ddc[n][j]
                                                      + mj[0] * R.du[n] + mj[1] * R.dv[n] + mj[2] * R.dw[n]
                                                      + nu[0] * R.ddu[n][j] + nu[1] * R.ddv[n][j] + nu[2] * R.ddw[n][j]);
        }
    }
}

/* set_initial, :58-105; s->src = (x, y, z) */
void go_rd3_set_initial(go_rd *a, const go_params *p, go_src *s, double *y)
{
    double x0 = s->src[0], y0 = s->src[1], z0 = s->src[2], th = s->theta, ph = s->phi;
    s->c0 = go_rd_c(a, z0, x0, y0);
    double mach[3] = {go_rd_u(a, z0, x0, y0) / s->c0, go_rd_v(a, z0, x0, y0) / s->c0, go_rd_w(a, z0, x0, y0) / s->c0};
    double nu0[3] = {cos(th) * cos(ph), cos(th) * sin(ph), sin(th)};
    double mt[3] = {-sin(th) * cos(ph), -sin(th) * sin(ph), cos(th)};
    double mp[3] = {-cos(th) * sin(ph), cos(th) * cos(ph), 0.0};
    double msc = 1.0 + (nu0[0] * mach[0] + nu0[1] * mach[1] + nu0[2] * mach[2]);
    y[0] = x0; y[1] = y0; y[2] = z0;
    for (int n = 0; n < 3; n++) y[3 + n] = nu0[n] / msc;
    if (p->eq_cnt < 18) return;
    for (int n = 0; n < 3; n++) {
        y[6 + n] = 0.0;
        y[12 + n] = 0.0;
        y[9 + n] = mt[n] / msc - nu0[n] / sq(msc) * (mt[0] * mach[0] + mt[1] * mach[1] + mt[2] * mach[2]);
        y[15 + n] = mp[n] / msc - nu0[n] / sq(msc) * (mp[0] * mach[0] + mp[1] * mach[1] + mp[2] * mach[2]);
    }
}

static double rd3_set_ds(go_rd *a, const go_params *p, const double *y)   /* :215-221 */
{
    double result = p->ds_max;
    if (y[2] < a->z_bndlyr) result -= (p->ds_max - p->ds_min) * exp(-(y[2] - go_rd_topo_z(a, y[0], y[1])) / 0.2);
    return result;
}

static int rd3_break(go_rd *a, const go_params *p, const go_src *s, const double *y)   /* :434-442 */
{
    if (sqrt(sq(y[0] - s->src[0]) + sq(y[1] - s->src[1])) > p->rng_max) return 1;
    if (y[0] > p->lat_max || y[0] < p->lat_min) return 1;
    if (y[1] > p->lon_max || y[1] < p->lon_min) return 1;
    if (y[2] > p->alt_max) return 1;
    return 0;
}

static int rd3_ground(go_rd *a, const double *y)   /* :444-451 */
{
    if (y[2] <= a->z_max)
        if (y[2] < go_rd_topo_z(a, y[0], y[1])) return 1;
    return 0;
}

static int rd3_prop_rk4(go_rd *a, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    int k, n = p->eq_cnt;
    double sum_s = 0.0;
    *check = 0;
    for (k = 0; k < length - 1; k++) {
        double *y = rows + (size_t)k * 18, *y1 = rows + (size_t)(k + 1) * 18;
        double t1[18], t2[18], t3[18], t4[18], pa[18], f[18];
        double ds = rd3_set_ds(a, p, y);
        go_rd3_rhs(a, p, y, f);
        for (int i = 0; i < n; i++) { t1[i] = ds * f[i]; pa[i] = y[i] + t1[i] / 2.0; }
        go_rd3_rhs(a, p, pa, f);
        for (int i = 0; i < n; i++) t2[i] = ds * f[i];
        for (int i = 0; i < n; i++) pa[i] = y[i] + t2[i] / 2.0;
        go_rd3_rhs(a, p, pa, f);
        for (int i = 0; i < n; i++) t3[i] = ds * f[i];
        for (int i = 0; i < n; i++) pa[i] = y[i] + t3[i];
        go_rd3_rhs(a, p, pa, f);
        for (int i = 0; i < n; i++) {
            t4[i] = ds * f[i];
            y1[i] = y[i] + t1[i] / 6.0 + t2[i] / 3.0 + t3[i] / 3.0 + t4[i] / 6.0;
        }
        sum_s += ds;
        if (sum_s > p->s_max) { *check = 1; break; }
        if (rd3_break(a, p, s, y1)) { *check = 1; break; }
        if (rd3_ground(a, y1)) { *check = 0; break; }
    }
    return k + 1;
}

/* approx_intercept + set_refl, :108-213 */
void go_rd3_set_refl(go_rd *a, const go_params *p, const go_src *s, const double *rk2, const double *rk1, double *y0)
{
    int neq = p->eq_cnt;
    double prev[18];
    double dz_grnd = go_rd_topo_z(a, rk1[0], rk1[1]) - rk1[2];
    double dz_k = (rk1[2] - rk2[2]) - (rk1[0] - rk2[0]) * go_rd_topo_dz(a, rk1[0], rk1[1], 0)
                  - (rk1[1] - rk2[1]) * go_rd_topo_dz(a, rk1[0], rk1[1], 1);
    for (int n = 0; n < neq; n++) prev[n] = rk1[n] + (rk1[n] - rk2[n]) / dz_k * dz_grnd;

    double zg = go_rd_topo_z(a, prev[0], prev[1]);
    double c_grnd = go_rd_c(a, zg, prev[0], prev[1]);
    double A = go_rd_topo_dz(a, prev[0], prev[1], 0), B = go_rd_topo_dz(a, prev[0], prev[1], 1);
    double norm1 = 1.0 + sq(A) + sq(B);
    double C1x = (1.0 - sq(A) + sq(B)) / norm1, C2x = 2.0 * A / norm1;
    double C1y = (1.0 + sq(A) - sq(B)) / norm1, C2y = 2.0 * B / norm1;
    double C1z = (1.0 - sq(A) - sq(B)) / norm1;
    double da[2] = {0, 0}, db[2] = {0, 0}, ds0[2] = {0, 0}, dC1x[2] = {0, 0}, dC1y[2] = {0, 0}, dC1z[2] = {0, 0};
    double dC2x[2] = {0, 0}, dC2y[2] = {0, 0}, dnx = 0, dny = 0, dnz = 0;
    if (p->calc_amp) {
        double norm2 = prev[5] - A * prev[3] - B * prev[4];
        for (int j = 0; j < 2; j++) {
            int o = j ? 12 : 6;
            double k1 = (prev[o] * (prev[5] - B * prev[4]) - prev[3] * (prev[o + 2] - B * prev[o + 1])) / norm2;
            double k2 = (prev[o + 1] * (prev[5] - A * prev[3]) - prev[4] * (prev[o + 2] - A * prev[o])) / norm2;
            da[j] = k1 * go_rd_topo_ddz(a, prev[0], prev[1], 0, 0) + k2 * go_rd_topo_ddz(a, prev[0], prev[1], 0, 1);
            db[j] = k1 * go_rd_topo_ddz(a, prev[0], prev[1], 1, 0) + k2 * go_rd_topo_ddz(a, prev[0], prev[1], 1, 1);
            dC1x[j] = -4.0 * A / sq(norm1) * ((1.0 + sq(B)) * da[j] - A * B * db[j]);
            dC1y[j] = -4.0 * B / sq(norm1) * ((1.0 + sq(A)) * db[j] - A * B * da[j]);
            dC1z[j] = -4.0 / sq(norm1) * (A * da[j] + B * db[j]);
            dC2x[j] = 2.0 / sq(norm1) * ((1.0 - sq(A) + sq(B)) * da[j] - 2.0 * A * B * db[j]);
            dC2y[j] = 2.0 / sq(norm1) * ((1.0 + sq(A) - sq(B)) * db[j] - 2.0 * A * B * da[j]);
            ds0[j] = -s->c0 / c_grnd * (prev[o + 2] - A * prev[o] - B * prev[o + 1]) / norm2;
        }
        /* atmo::dc(x, y, zg, n), n = 0 x, 1 y, 2 z */
        double dn[3];
        for (int n = 0; n < 3; n++)
            dn[n] = -1.0 / c_grnd * (s->c0 / c_grnd * go_rd_dc(a, zg, prev[0], prev[1], mi3(n))
                                     + prev[3] * go_rd_du(a, zg, prev[0], prev[1], mi3(n))
                                     + prev[4] * go_rd_dv(a, zg, prev[0], prev[1], mi3(n))
                                     + prev[5] * go_rd_dw(a, zg, prev[0], prev[1], mi3(n)));
        dnx = dn[0]; dny = dn[1]; dnz = dn[2];
    }
    y0[0] = prev[0];
    y0[1] = prev[1];
    y0[2] = zg;
    for (int j = -1; j < 2; j++) {
        int o = j < 0 ? 3 : (j ? 12 : 6);
        if (j >= 0 && neq < 18) break;
        y0[o] = C1x * prev[o] + C2x * (prev[o + 2] - prev[o + 1] * B);
        y0[o + 1] = C1y * prev[o + 1] + C2y * (prev[o + 2] - prev[o] * A);
        y0[o + 2] = -C1z * prev[o + 2] + prev[o] * C2x + prev[o + 1] * C2y;
    }
    if (neq < 18) return;
    for (int j = 0; j < 2; j++) {
        int o = j ? 15 : 9;
        y0[o] = C1x * prev[o] + C2x * (prev[o + 2] - B * prev[o + 1] - prev[4] * db[j]) + dC1x[j] * prev[3]
                + dC2x[j] * (prev[5] - B * prev[4]) + ((C1x - 1.0) * dnx + C2x * (dnz - B * dny)) * ds0[j];
        y0[o + 1] = C1y * prev[o + 1] + C2y * (prev[o + 2] - A * prev[o] - prev[3] * da[j]) + dC1y[j] * prev[4]
                    + dC2y[j] * (prev[5] - A * prev[3]) + ((C1y - 1.0) * dny + C2y * (dnz - A * dnx)) * ds0[j];
        y0[o + 2] = -C1z * prev[o + 2] + C2x * prev[o] + C2y * prev[o + 1] - dC1z[j] * prev[5] + dC2x[j] * prev[3]
                    + dC2y[j] * prev[4] + (-(C1z + 1.0) * dnz + C2x * dnx + C2y * dny) * ds0[j];
    }
}

/* travel_time / atten of one segment, :460-485 (Cartesian Sutherland-Bass) */
static void rd3_seg(go_rd *a, const double *y0, const double *y1, double freq, double *dt, double *datt)
{
    double dx = y1[0] - y0[0], dy = y1[1] - y0[1], dz = y1[2] - y0[2];
    double x = y0[0] + dx / 2.0, y = y0[1] + dy / 2.0, z = y0[2] + dz / 2.0;
    double ds = sqrt(dx * dx + dy * dy + dz * dz), nu[3], cp[3];
    for (int n = 0; n < 3; n++) nu[n] = y0[3 + n] + (y1[3 + n] - y0[3 + n]) / 2.0;
    double nm = sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
    double c = go_rd_c(a, z, x, y);
    cp[0] = c * nu[0] / nm + go_rd_u(a, z, x, y);
    cp[1] = c * nu[1] / nm + go_rd_v(a, z, x, y);
    cp[2] = c * nu[2] / nm + go_rd_w(a, z, x, y);
    *dt = ds / sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    *datt = go_sb_alpha_at(c, go_rd_rho(a, z, x, y), z, 0, a->tweak_abs, freq) * ds;
}

/* jacobian, :704-721 */
double go_rd3_jacobian(go_rd *a, const double *y)
{
    double c = go_rd_c(a, y[2], y[0], y[1]), nm = sqrt(sq(y[3]) + sq(y[4]) + sq(y[5])), cp[3];
    cp[0] = c * y[3] / nm + go_rd_u(a, y[2], y[0], y[1]);
    cp[1] = c * y[4] / nm + go_rd_v(a, y[2], y[0], y[1]);
    cp[2] = c * y[5] / nm + go_rd_w(a, y[2], y[0], y[1]);
    double cm = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2]));
    double dxs = cp[0] / cm, dys = cp[1] / cm, dzs = cp[2] / cm;
    return dxs * (y[7] * y[14] - y[13] * y[8]) - y[6] * (dys * y[14] - dzs * y[13]) + y[12] * (dys * y[8] - dzs * y[7]);
}

/* amp, :723-746 */
double go_rd3_amp(go_rd *a, const go_src *s, const double *y)
{
    double x0 = s->src[0], y0 = s->src[1], z0 = s->src[2];
    double c = go_rd_c(a, y[2], y[0], y[1]);
    double w[3] = {go_rd_u(a, y[2], y[0], y[1]), go_rd_v(a, y[2], y[0], y[1]), go_rd_w(a, y[2], y[0], y[1])};
    double w0[3] = {go_rd_u(a, z0, x0, y0), go_rd_v(a, z0, x0, y0), go_rd_w(a, z0, x0, y0)};
    double nm = sqrt(sq(y[3]) + sq(y[4]) + sq(y[5]));
    double nm0 = 1.0 - (y[3] * w0[0] + y[4] * w0[1] + y[5] * w0[2]) / s->c0;
    double cp[3], cp0[3];
    for (int n = 0; n < 3; n++) cp[n] = c * y[3 + n] / nm + w[n];
    cp0[0] = s->c0 * cos(s->theta) * cos(s->phi) + w0[0];
    cp0[1] = s->c0 * cos(s->theta) * sin(s->phi) + w0[1];
    cp0[2] = s->c0 * sin(s->theta) + w0[2];
    double cm = sqrt(sq(cp[0]) + sq(cp[1]) + sq(cp[2])), cm0 = sqrt(sq(cp0[0]) + sq(cp0[1]) + sq(cp0[2]));
    double num = go_rd_rho(a, y[2], y[0], y[1]) * nm * pow(c, 3) * cm0 * cos(s->theta);
    double den = go_rd_rho(a, z0, x0, y0) * nm0 * pow(s->c0, 3) * cm * go_rd3_jacobian(a, y);
    return sqrt(fabs(num / den));
}

static void rd3_pt(go_path_pt *pts, long long cap, long long *n, go_rd *a, int li, int leg, int m, int kind,
                   const double *y, double tl, double att, double tt)
{
    if (*n < cap) {
        go_path_pt *P = pts + *n;
        P->launch_idx = li; P->leg = leg; P->step = m; P->kind = kind;
        P->pos0 = y[0]; P->pos1 = y[1];
        /* raypath rows clip z at the terrain (:456), caustic rows do not (:469);
         * the stratified driver clips both (main.infraga.3d.cpp:457, 471) */
        P->z = (kind == 0 || a->strat) ? dmax(y[2], go_rd_topo_z(a, y[0], y[1])) : y[2];
        P->tl = tl; P->absorption = -2.0 * att; P->time = tt;
    }
    (*n)++;
}

/* run_prop, main.infraga.3d.rngdep.cpp:400-540.  write = 0: the
 * write_rays=false sums (travel_time / atten over rows 0..k-1 of a leg);
 * write = 1: the write_rays / write_caustics bookkeeping (rows 1..k-1 one
 * segment at a time, rows m == 1 or m % 15 == 0 and Jacobian sign changes
 * written to pts as flags select).  pr->src0, src1: x, y [km] (NaN: the
 * region centre). */
int go_rd3_fan(go_rd *a, const go_params *p, const go_prop *pr, const double *launches, int n_launch, int write,
               int flags, go_arrival *out, int out_cap, int *n_out, go_path_pt *pts, long long pts_cap,
               long long *n_pts, long long *ray_steps)
{
    int length = (int)(p->s_max / p->ds_min);
    if (rd_rows_len < length) {
        free(rd_rows);
        rd_rows = (double *)malloc(sizeof(double) * 18 * (size_t)length);
        rd_rows_len = length;
    }
    double *rows = rd_rows;
    go_src s;
    int cnt = 0;
    long long steps = 0, np = 0;
    double x_src = pr->src0, y_src = pr->src1;
    if (isnan(x_src)) x_src = (p->lat_max + p->lat_min) / 2.0;           /* :37-38 */
    if (isnan(y_src)) y_src = (p->lon_max + p->lon_min) / 2.0;
    double z_src = dmax(pr->src_alt, go_rd_topo_z(a, x_src, y_src));    /* :305 */
    s.src[0] = x_src; s.src[1] = y_src; s.src[2] = z_src;

    for (int li = 0; li < n_launch; li++) {
        double incl = launches[2 * li], az = launches[2 * li + 1];
        s.phi = (90.0 - az) * (RD_PI / 180.0);
        s.theta = incl * (RD_PI / 180.0);
        go_rd3_set_initial(a, p, &s, rows);
        /* the stratified driver starts the turning height at the profile
         * bottom (main.infraga.3d.cpp:426 with set_limits' z0) */
        double tt = 0.0, att = 0.0, zmax = a->strat ? a->z0 : 0.0;
        int brk = 0, k;
        for (int bnc = 0; bnc <= pr->bounces; bnc++) {
            k = rd3_prop_rk4(a, p, &s, rows, length, &brk);
            steps += k;
            if (write) {
                double d_prev = (p->calc_amp && (flags & GO_PATH_CAUSTICS)) ? go_rd3_jacobian(a, rows + 18) : 0.0;
                for (int m = 1; m < k; m++) {
                    double dt, da, *ym = rows + 18 * m;
                    rd3_seg(a, rows + 18 * (m - 1), ym, pr->freq, &dt, &da);
                    tt += dt;
                    att += da;
                    if ((flags & GO_PATH_RAYS) && (m == 1 || m % 15 == 0))
                        rd3_pt(pts, pts_cap, &np, a, li, bnc, m, 0, ym,
                               p->calc_amp ? 20.0 * log10(go_rd3_amp(a, &s, ym)) : 0.0, att, tt);
                    if (p->calc_amp && (flags & GO_PATH_CAUSTICS)) {
                        double d = go_rd3_jacobian(a, ym);
                        if (d * d_prev < 0.0) rd3_pt(pts, pts_cap, &np, a, li, bnc, m, 1, ym, 0.0, att, tt);
                        d_prev = d;
                    }
                }
            } else {
                double tleg = 0.0, aleg = 0.0, dt, da;
                for (int n = 0; n < k; n++) {
                    rd3_seg(a, rows + 18 * n, rows + 18 * (n + 1), pr->freq, &dt, &da);
                    tleg += dt;
                    aleg += da;
                }
                tt += tleg;
                att += aleg;
            }
            for (int m = 0; m < k; m++) zmax = dmax(zmax, rows[18 * m + 2]);      /* :490-492 */
            if (!a->is_topo && zmax < (a->z0 + pr->turn_ht_min)) break;         /* :494 */
            if (brk || k < 2) break;                                   /* :498, no travel-time limit */
            double *yk = rows + 18 * k;
            go_arrival A;
            memset(&A, 0, sizeof(A));
            A.launch_idx = li;
            A.n_bounce = bnc;
            A.incl_launch = incl;
            A.az_launch = az;
            A.incl_arr = -asin(go_rd_c(a, go_rd_topo_z(a, yk[0], yk[1]), yk[0], yk[1]) / go_rd_c(a, z_src, x_src, y_src)
                               * yk[5]) * 180.0 / RD_PI;                       /* :502 */
            double baz = 90.0 - atan2(-yk[4], -yk[3]) * 180.0 / RD_PI;
            if (baz < -180.0) baz += 360.0;
            else if (baz > 180.0) baz -= 360.0;
            A.back_az = baz;
            A.pos0 = yk[0];
            A.pos1 = yk[1];
            A.time = tt;
            A.cel = sqrt(sq(yk[0] - x_src) + sq(yk[1] - y_src)) / tt;
            A.turn_ht = zmax;
            A.tl = p->calc_amp ? 20.0 * log10(go_rd3_amp(a, &s, yk)) : 0.0;
            A.absorption = -2.0 * att;
            if (cnt < out_cap) out[cnt] = A;
            cnt++;
            go_rd3_set_refl(a, p, &s, rows + 18 * (k - 2), rows + 18 * (k - 1), rows);
        }
    }
    *n_out = cnt;
    if (n_pts) *n_pts = np;
    if (ray_steps) *ray_steps = steps;
    return (cnt > out_cap || np > pts_cap) ? -1 : 0;
}

/* ---- eigenray search (infraga-3d-rngdep -eig_search / -eig_direct) ----
 * The 3D search of geoac.eigenray.3d.cpp over the Cartesian medium: rows are
 * (x, y, z), the medium is read at (z, x, y). */
static void r3_init(void *m, const go_params *p, go_src *s, double *y) { go_rd3_set_initial((go_rd *)m, p, s, y); }
static int r3_prop(void *m, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    return rd3_prop_rk4((go_rd *)m, p, s, rows, length, check);
}
static void r3_refl(void *m, const go_params *p, const go_src *s, const double *a, const double *b, double *o)
{
    go_rd3_set_refl((go_rd *)m, p, s, a, b, o);
}
static double r3_tt(void *m, const double *a, const double *b)
{
    double dt, da;
    rd3_seg((go_rd *)m, a, b, 0.1, &dt, &da);
    return dt;
}
static double r3_att(void *m, const double *a, const double *b, double f)
{
    double dt, da;
    rd3_seg((go_rd *)m, a, b, f, &dt, &da);
    return da;
}
static double r3_c(void *m, double x, double y, double z) { return go_rd_c((go_rd *)m, z, x, y); }
static double r3_amp(void *m, const go_src *s, const double *y) { return go_rd3_amp((go_rd *)m, s, y); }
static double r3_tz(void *m, double x, double y) { return go_rd_topo_z((go_rd *)m, x, y); }
static double r3_tdz(void *m, double x, double y, int n) { return go_rd_topo_dz((go_rd *)m, x, y, n); }

static go_eig_med rd3_med(go_rd *a)
{
    go_eig_med M = {a, 0.0, r3_init, r3_prop, r3_refl, r3_tt, r3_att, r3_c, r3_amp, r3_tz, r3_tdz, a->is_topo};
    return M;
}

/* src = (x, y, z [km]) raised to the terrain (main.infraga.3d.rngdep.cpp:845, 1037). */
static void rd3_eig_src(go_rd *a, const double src_in[3], double src[3])
{
    src[0] = src_in[0];
    src[1] = src_in[1];
    src[2] = dmax(src_in[2], go_rd_topo_z(a, src_in[0], src_in[1]));
}

/* run_eig_search (main.infraga.3d.rngdep.cpp:754-898): the sweep starts at
 * theta_min for every bounce count (no ground-inclination floor). */
int go_rd3_eig_search(go_rd *a, const go_params *p, const double src_in[3], const double rcvr[2], double th_min_deg,
                      double th_max_deg, int bnc_min, int bnc_max, int iterations, double damping, double tolerance,
                      double az_err_lim, double dth_big_deg, double dth_sml_deg, double freq, go_arrival *out,
                      int out_cap, long long *ray_steps)
{
    double src[3];
    rd3_eig_src(a, src_in, src);
    go_eig_med M = rd3_med(a);
    return go_eig_search_3d_med(&M, p, src, rcvr, th_min_deg * (RD_PI / 180.0), th_max_deg * (RD_PI / 180.0),
                                bnc_min, bnc_max, iterations, damping, tolerance, az_err_lim, dth_big_deg,
                                dth_sml_deg, freq, out, out_cap, ray_steps);
}

/* run_eig_direct (:901-1052): find_eigenray from (incl_est, az_est) [deg];
 * az_est NaN aims at the receiver (:1039). */
int go_rd3_eig_direct(go_rd *a, const go_params *p, const double src_in[3], const double rcvr[2], double incl_est_deg,
                      double az_est_deg, int bnc, int iterations, double damping, double tolerance, double freq,
                      go_arrival *A, long long *ray_steps)
{
    double src[3];
    rd3_eig_src(a, src_in, src);
    go_eig_med M = rd3_med(a);
    double ph = isnan(az_est_deg) ? atan2(rcvr[1] - src[1], rcvr[0] - src[0])
                                  : RD_PI / 2.0 - az_est_deg * (RD_PI / 180.0);
    return go_eig_find_3d_med(&M, p, src, rcvr, incl_est_deg * (RD_PI / 180.0), ph, bnc, iterations, damping,
                              tolerance, freq, A, ray_steps);
}

/* go_eig_lm_jac_3d / go_eig_trace_leg_3d over this medium (th, ph in radians,
 * src as given). */
void go_rd3_lm_jac(go_rd *a, const double src[3], const double *yk, double J[4])
{
    go_eig_med M = rd3_med(a);
    go_eig_lm_jac_3d(&M, src, yk, J);
}
int go_rd3_trace_leg(go_rd *a, const go_params *p, const double src[3], double th, double ph, double y_km1[18],
                     double y_k[18])
{
    go_eig_med M = rd3_med(a);
    return go_eig_trace_leg_3d(&M, p, src, th, ph, y_km1, y_k);
}

/* The accepted eigenray's path rows and row for launch angles in degrees. */
long long go_rd3_eig_path(go_rd *a, const go_params *p, const double src_in[3], double incl_deg, double az_deg,
                          int bnc_cnt, double freq, go_path_pt *pts, long long cap, go_arrival *A)
{
    double src[3];
    rd3_eig_src(a, src_in, src);
    go_eig_med M = rd3_med(a);
    return go_eig_path_3d_med(&M, p, src, incl_deg * (RD_PI / 180.0), (90.0 - az_deg) * (RD_PI / 180.0), bnc_cnt,
                              freq, pts, cap, A);
}

void *go_rd_new_xy(const double *x_km, int nx, const double *y_km, int ny, const double *z, int nz,
                   const double *c, const double *u, const double *v, const double *rho)
{
    go_rd *a = (go_rd *)malloc(sizeof(go_rd));
    if (go_rd_init_xy(a, x_km, nx, y_km, ny, z, nz, c, u, v, rho)) { free(a); return NULL; }
    return a;
}

/* Public forms of the Cartesian set's prop_rk4 and segment bookkeeping
 * (the waveform driver, ga_wvfrm.c). */
int go_rd3_prop_rk4(go_rd *a, const go_params *p, const go_src *s, double *rows, int length, int *check)
{
    return rd3_prop_rk4(a, p, s, rows, length, check);
}

void go_rd3_seg(go_rd *a, const double *y0, const double *y1, double freq, double *dt, double *datt)
{
    rd3_seg(a, y0, y1, freq, dt, datt);
}
