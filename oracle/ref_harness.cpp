// ref_harness.cpp -- C entry points into the reference's own compiled
// spline / medium / absorption / globe sources (built from /root/reference
// by oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: it lets
// tests pin the oracle restatement against the reference implementation for
// the parts of the path whose sources build without third-party headers.
//
// Built twice: with -DREF_SPH against the spherical sources and with
// -DREF_3D against the Cartesian ones (each geometry defines the same
// atmo:: / topo:: symbols, so they live in separate libraries).
#include <cstring>
#include <cmath>

#include "atmo/atmo_state.h"
#ifdef REF_SPH
#include "atmo/atmo_io.sph.strat.h"
#else
#include "atmo/atmo_io.3d.strat.h"
#endif
#include "geoac/geoac.params.h"
#include "util/interpolation.h"
#include "util/globe.h"

extern "C" {

// set_region(file, format, invert_winds): atmo_io.*.strat.cpp
int ref_load(const char *path, const char *fmt)
{
    static char pbuf[4096], fbuf[64];
    std::strncpy(pbuf, path, sizeof(pbuf) - 1);
    std::strncpy(fbuf, fmt, sizeof(fbuf) - 1);
    geoac::is_topo = false;
    topo::z0 = 0.0;
    atmo::tweak_abs = 1.0;
    return set_region(pbuf, fbuf, false);
}

// set_region(atmo_file, topo_file, format, invert_winds): the stratified
// profile over terrain (atmo_io.*.strat.cpp), as the drivers run it with
// topo_file=.
int ref_load_topo(const char *path, const char *topo_path, const char *fmt)
{
    static char pbuf[4096], tbuf[4096], fbuf[64];
    std::strncpy(pbuf, path, sizeof(pbuf) - 1);
    std::strncpy(tbuf, topo_path, sizeof(tbuf) - 1);
    std::strncpy(fbuf, fmt, sizeof(fbuf) - 1);
    geoac::is_topo = true;
    topo::z0 = 0.0;
    atmo::tweak_abs = 1.0;
    return set_region(pbuf, tbuf, fbuf, false);
}

void ref_use_blw(int on) { topo::use_BLw = on != 0; }
double ref_z_max() { return topo::z_max; }

// out = {z, dz/dx0, dz/dx1, d2z/dx0^2, d2z/dx0dx1, d2z/dx1^2} of topo::z (which
// includes r0 in the spherical geometry).
void ref_topo(double a, double b, double *out)
{
    out[0] = topo::z(a, b);
    out[1] = topo::dz(a, b, 0);
    out[2] = topo::dz(a, b, 1);
    out[3] = topo::ddz(a, b, 0, 0);
    out[4] = topo::ddz(a, b, 0, 1);
    out[5] = topo::ddz(a, b, 1, 1);
}

void ref_clear() { clear_region(); }

int ref_len() { return atmo::c_spline.length; }

// Copy a spline's node values and Thomas-solved slopes (which: 0 c, 1 u, 2 v, 3 rho).
void ref_spline_tables(int which, double *x, double *f, double *slope)
{
    interp::natural_cubic_spline_1D *s[4] = {&atmo::c_spline, &atmo::u_spline, &atmo::v_spline, &atmo::rho_spline};
    interp::natural_cubic_spline_1D *sp = s[which];
    for (int i = 0; i < sp->length; i++) { x[i] = sp->x_vals[i]; f[i] = sp->f_vals[i]; slope[i] = sp->slopes[i]; }
}

// interp::eval_all with f, f', f'' on one of the four splines.
void ref_spline_eval(int which, double z, double out[3])
{
    interp::natural_cubic_spline_1D *s[4] = {&atmo::c_spline, &atmo::u_spline, &atmo::v_spline, &atmo::rho_spline};
    interp::eval_all(z, *s[which], out[0], out[1], out[2]);
}

double ref_c(double a, double b, double c) { return atmo::c(a, b, c); }
double ref_w(double a, double b, double c) { return atmo::w(a, b, c); }
double ref_rho(double a, double b, double c) { return atmo::rho(a, b, c); }
double ref_u(double a, double b, double c) { return atmo::u(a, b, c); }
double ref_v(double a, double b, double c) { return atmo::v(a, b, c); }
double ref_dc(double a, double b, double c, int n) { return atmo::dc(a, b, c, n); }
double ref_du(double a, double b, double c, int n) { return atmo::du(a, b, c, n); }
double ref_dv(double a, double b, double c, int n) { return atmo::dv(a, b, c, n); }

// calc_uvw: out = u, v, w, du[3], dv[3], dw[3] (+ ddu[6], ddv[6], ddw[6] when second != 0)
void ref_calc_uvw(double a, double b, double c, int second, double *out)
{
    double u, v, w, du[3], dv[3], dw[3], ddu[6], ddv[6], ddw[6];
    if (second) atmo::calc_uvw(a, b, c, u, v, w, du, dv, dw, ddu, ddv, ddw);
    else atmo::calc_uvw(a, b, c, u, v, w, du, dv, dw);
    out[0] = u; out[1] = v; out[2] = w;
    for (int i = 0; i < 3; i++) { out[3 + i] = du[i]; out[6 + i] = dv[i]; out[9 + i] = dw[i]; }
    if (second) for (int i = 0; i < 6; i++) { out[12 + i] = ddu[i]; out[18 + i] = ddv[i]; out[24 + i] = ddw[i]; }
}

double ref_sb_alpha(double a, double b, double c, double freq) { return atmo::SB_alpha(a, b, c, freq); }

void ref_set_tweak(double t) { atmo::tweak_abs = t; }

double ref_z_bndlyr() { return topo::z_bndlyr; }
double ref_z0() { return topo::z0; }
double ref_alt_max() { return geoac::alt_max; }

#ifdef REF_3D
void ref_limits(double *out)
{
    out[0] = geoac::x_min; out[1] = geoac::x_max; out[2] = geoac::y_min; out[3] = geoac::y_max;
    out[4] = geoac::alt_max; out[5] = topo::z0; out[6] = topo::z_max; out[7] = topo::z_bndlyr;
}
#endif

#ifdef REF_SPH
void ref_limits(double *out)
{
    out[0] = geoac::lat_min; out[1] = geoac::lat_max; out[2] = geoac::lon_min; out[3] = geoac::lon_max;
    out[4] = geoac::alt_max; out[5] = topo::z0; out[6] = topo::z_max; out[7] = topo::z_bndlyr;
}
double ref_gc_dist(double lat1, double lon1, double lat2, double lon2) { return globe::gc_dist(lat1, lon1, lat2, lon2); }
double ref_bearing(double lat1, double lon1, double lat2, double lon2) { return globe::bearing(lat1, lon1, lat2, lon2); }
double ref_r0() { return globe::r0; }
#endif

}
