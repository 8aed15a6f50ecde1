// ref_rd_harness.cpp -- C entry points into the reference's own compiled
// range-dependent spherical atmosphere (atmo_state.sph.rngdep.cpp,
// atmo_io.sph.rngdep.cpp) with its splines, absorption and globe sources,
// built from /root/reference by oracle/Makefile into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: it pins the oracle's range-dependent medium
// (ga_oracle_rd.c) against the reference implementation.
#include <cstring>
#include <cmath>

#include "atmo/atmo_state.h"
#include "atmo/atmo_io.sph.rngdep.h"
#include "geoac/geoac.params.h"
#include "util/interpolation.h"
#include "util/globe.h"

extern "C" {

// set_region(prefix, lats, lons[, topo], format, invert_winds): atmo_io.sph.rngdep.cpp
int ref_rd_load(const char *prefix, const char *lats, const char *lons, const char *topo, const char *fmt)
{
    static char b0[4096], b1[4096], b2[4096], b3[4096], bf[64];
    std::strncpy(b0, prefix, sizeof(b0) - 1);
    std::strncpy(b1, lats, sizeof(b1) - 1);
    std::strncpy(b2, lons, sizeof(b2) - 1);
    std::strncpy(bf, fmt, sizeof(bf) - 1);
    topo::z0 = 0.0;
    atmo::tweak_abs = 1.0;
    topo::use_BLw = false;
    if (topo && topo[0]) {
        std::strncpy(b3, topo, sizeof(b3) - 1);
        geoac::is_topo = true;
        return set_region(b0, b1, b2, b3, bf, false);
    }
    geoac::is_topo = false;
    return set_region(b0, b1, b2, bf, false);
}

void ref_rd_clear() { clear_region(); }
void ref_rd_use_blw(int on) { topo::use_BLw = on != 0; }
void ref_rd_set_tweak(double t) { atmo::tweak_abs = t; }

// lat_min, lat_max, lon_min, lon_max [rad], alt_max, z0, z_max, z_bndlyr
void ref_rd_limits(double *out)
{
    out[0] = geoac::lat_min; out[1] = geoac::lat_max;
    out[2] = geoac::lon_min; out[3] = geoac::lon_max;
    out[4] = geoac::alt_max; out[5] = topo::z0; out[6] = topo::z_max; out[7] = topo::z_bndlyr;
}

// point functions; out = {c, rho, u, v, w, dc[3], du[3], dv[3], dw[3]}
void ref_rd_point(double r, double t, double p, double *out)
{
    out[0] = atmo::c(r, t, p);
    out[1] = atmo::rho(r, t, p);
    out[2] = atmo::u(r, t, p);
    out[3] = atmo::v(r, t, p);
    out[4] = atmo::w(r, t, p);
    for (int n = 0; n < 3; n++) {
        out[5 + n] = atmo::dc(r, t, p, n);
        out[8 + n] = atmo::du(r, t, p, n);
        out[11 + n] = atmo::dv(r, t, p, n);
        out[14 + n] = atmo::dw(r, t, p, n);
    }
}

// calc_uvw: out = u, v, w, du[3], dv[3], dw[3] (+ ddu[6], ddv[6], ddw[6] when second != 0)
void ref_rd_calc_uvw(double r, double t, double p, int second, double *out)
{
    double u, v, w, du[3], dv[3], dw[3], ddu[6], ddv[6], ddw[6];
    if (second) atmo::calc_uvw(r, t, p, u, v, w, du, dv, dw, ddu, ddv, ddw);
    else atmo::calc_uvw(r, t, p, u, v, w, du, dv, dw);
    out[0] = u; out[1] = v; out[2] = w;
    for (int i = 0; i < 3; i++) { out[3 + i] = du[i]; out[6 + i] = dv[i]; out[9 + i] = dw[i]; }
    if (second) for (int i = 0; i < 6; i++) { out[12 + i] = ddu[i]; out[18 + i] = ddv[i]; out[24 + i] = ddw[i]; }
}

// topo::z, dz(0..1), ddz(00, 01, 11), dddz(000, 001, 011, 111)
void ref_rd_topo(double lat, double lon, double *out)
{
    out[0] = topo::z(lat, lon);
    out[1] = topo::dz(lat, lon, 0);
    out[2] = topo::dz(lat, lon, 1);
    out[3] = topo::ddz(lat, lon, 0, 0);
    out[4] = topo::ddz(lat, lon, 0, 1);
    out[5] = topo::ddz(lat, lon, 1, 1);
    out[6] = topo::dddz(lat, lon, 0, 0, 0);
    out[7] = topo::dddz(lat, lon, 0, 0, 1);
    out[8] = topo::dddz(lat, lon, 0, 1, 1);
    out[9] = topo::dddz(lat, lon, 1, 1, 1);
}

// bndlyr_sc, bndlyr_dsc(0..2), bndlyr_ddsc (00, 01, 02, 11, 12, 22)
void ref_rd_bl(double r, double t, double p, double *out)
{
    out[0] = topo::bndlyr_sc(r, t, p);
    for (int n = 0; n < 3; n++) out[1 + n] = topo::bndlyr_dsc(r, t, p, n);
    int k = 4;
    for (int n = 0; n < 3; n++)
        for (int m = n; m < 3; m++) out[k++] = topo::bndlyr_ddsc(r, t, p, n, m);
}

double ref_rd_sb_alpha(double r, double t, double p, double freq) { return atmo::SB_alpha(r, t, p, freq); }

}
