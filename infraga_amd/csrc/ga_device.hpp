// ga_device.hpp -- device-side physics of the stratified ray engine (gfx950).
//
// One ray per lane.  The medium is a natural cubic spline table staged in
// LDS; the ray state lives in registers for the whole ray.  The equations
// are the geometric-acoustics ray system with its auxiliary (launch-angle
// derivative) equations for a horizontally stratified moving medium over
// flat ground, in spherical (r, lat, lon) or Cartesian (x, y, z) form.
//
// Relation to the reference (paths relative to the infraGA source tree):
//   spline evaluation      src/util/interpolation.cpp:144-260
//   medium + BL scaling    src/atmo/atmo_state.sph.strat.cpp:100-249, 613-803
//                          src/atmo/atmo_state.3d.strat.cpp:81-240, 556-739
//   Sutherland-Bass        src/atmo/atmo_state.abs.sph.cpp:14-151,
//                          src/atmo/atmo_state.abs.cart.cpp:12-141
//   RHS                    src/geoac/geoac.eqset.sph.strat.cpp:227-433,
//                          src/geoac/geoac.eqset.3d.strat.cpp:225-389
//   RK4                    src/util/rk4solver.cpp:26-76
//   checks / reflection    geoac.eqset.sph.strat.cpp:108-222, 484-507;
//                          geoac.eqset.3d.strat.cpp:110-220, 438-453
//   bookkeeping            geoac.eqset.sph.strat.cpp:516-547, 756-849;
//                          geoac.eqset.3d.strat.cpp:463-489, 676-748
//
// Without terrain every horizontal derivative of the medium is exactly zero
// and the vertical wind is zero, so those terms are omitted: in IEEE
// arithmetic adding an exact zero leaves a sum unchanged, so dropping them
// does not change any result.  Divisions by a quantity used several times
// are replaced by one reciprocal and products, and pow(x, 2) by x * x; these
// change results by a few ulp, which the parity tests bound.
#pragma once

// GA_HD qualifies every function of the physics; a host build of the same
// code (used to debug the kernels on the CPU) defines GA_HD and
// GA_HOST_BUILD before including this header.
#ifndef GA_HOST_BUILD
#include <hip/hip_runtime.h>
#else
#include <cmath>
#include <cstring>
using std::fma; using std::sqrt; using std::rint; using std::fabs; using std::fmin; using std::fmax;
#endif
#ifndef GA_HD
#define GA_HD __device__ __forceinline__
#endif
#include "ga_math.hpp"

namespace ga {

constexpr double PI = 3.141592653589793238462643;
constexpr int NODE = 10;          // doubles per spline node: z, c, c', u, u', v, v', rho, rho', pad
constexpr int GEOM_3D = 0;
constexpr int GEOM_SPH = 1;

struct Medium {
    const double *tab;   // [n][NODE]
    int n;
    double zlo, zhi;     // table range
    double r0;           // globe radius (sph) or 0
    double zg;           // ground coordinate: r0 + z0 (sph) or z0 (3D)
    double z0;           // ground elevation [km]
    double z_max;        // topography maximum (= z0, flat ground)
    double z_bndlyr;     // boundary-layer top
    double tweak;        // Sutherland-Bass scale
    double sbc[3];       // 10^-0.67887, 10^-0.10744, 10^-3.3979 (host libm pow)
    const double *abs;   // absorption table at the run's frequency (nullptr: direct)
    int abs_per;         // table pieces per node interval
};

struct Limits {
    double s_max, ds_min, ds_max;
    double alt_max, rng_max, time_max;
    // rows above this altitude [km] cannot be row K-2 of a ground return
    // (a step moves the position by at most ds <= ds_max), so the trace
    // loops keep their copy of the previous row only below it
    double z_keep;
    double b0, b1, b2, b3;       // sph: lat/lon bounds [rad]; 3D: x/y bounds [km]
    double gc_thresh;            // sph: sin^2(rng_max / (2 r0)) for the range test
};

// --------------------------------------------------------------------------
// Elementary functions with bounded register footprint.
//
// sin/cos: reduction by the nearest multiple of pi/2 with a two-part
// Cody-Waite constant, then the fdlibm polynomial kernels (Sun Microsystems
// fdlibm k_sin.c / k_cos.c coefficients).  Accurate to about 1 ulp for the
// arguments that occur here (|x| <= 2 pi: latitudes, launch angles, angle
// differences); results agree with glibc to within a few ulp.
// rcp: hardware reciprocal estimate refined by two Newton steps (error below
// 1 ulp, not always the correctly rounded quotient).
// --------------------------------------------------------------------------
#ifdef GA_HOST_BUILD
GA_HD double ga_rcp(double x) { return 1.0 / x; }
GA_HD double ga_rcbrt(double x) { return 1.0 / std::cbrt(x); }
GA_HD long long __double_as_longlong(double x) { long long b; std::memcpy(&b, &x, 8); return b; }
GA_HD double __longlong_as_double(long long b) { double x; std::memcpy(&x, &b, 8); return x; }
#else
GA_HD double ga_rcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
GA_HD double ga_rcbrt(double x) { return rcbrt(x); }
#endif

GA_HD double k_sin(double x)
{
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = x * x, v = z * x;
    double r = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    return x + v * (S1 + z * r);
}

GA_HD double k_cos(double x)
{
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    double z = x * x;
    double r = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    double ax = fabs(x);
    if (ax < 0.3) return 1.0 - (0.5 * z - z * r);
    // split 1 - x^2/2 exactly as fdlibm does: qx = x^2/4 with low bits cleared
    double qx;
    if (ax > 0.78125) qx = 0.28125;
    else {
        long long b = __double_as_longlong(ax);
        b = ((b >> 32) - 0x00200000LL) << 32;
        qx = __longlong_as_double(b);
    }
    double hz = 0.5 * z - qx;
    double a = 1.0 - qx;
    return a - (hz - z * r);
}

GA_HD void ga_sincos(double x, double &s, double &c)
{
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    double n = rint(x * invpio2);
    double y = (x - n * pio2_1) - n * pio2_1t;
    double ks = k_sin(y), kc = k_cos(y);
    int q = ((int)n) & 3;
    s = (q == 0) ? ks : (q == 1) ? kc : (q == 2) ? -ks : -kc;
    c = (q == 0) ? kc : (q == 1) ? -ks : (q == 2) ? -kc : ks;
}

GA_HD double ga_sin(double x)
{
    double s, c;
    ga_sincos(x, s, c);
    return s;
}

GA_HD double ga_cos(double x)
{
    double s, c;
    ga_sincos(x, s, c);
    return c;
}

// --------------------------------------------------------------------------
// Spline segment search with a per-lane cursor.  Same segment as the
// reference's find_segment (interpolation.cpp:14-49) for every z strictly
// inside a segment; z is always clamped into the table first.
// --------------------------------------------------------------------------
GA_HD int seg_find(const Medium &M, double z, int cur)
{
    const double *t = M.tab;
    while (cur > 0 && z < t[cur * NODE]) --cur;
    while (cur < M.n - 2 && z > t[(cur + 1) * NODE]) ++cur;
    return cur;
}

struct SegPt {
    const double *a;   // node k
    double X, omX, inv_dx;
};

GA_HD SegPt seg_point(const Medium &M, double z, int &cur)
{
    cur = seg_find(M, z, cur);
    SegPt s;
    s.a = M.tab + cur * NODE;
    double dx = s.a[NODE] - s.a[0];
    s.inv_dx = ga_rcp(dx);
    s.X = (z - s.a[0]) * s.inv_dx;
    s.omX = 1.0 - s.X;
    return s;
}

// Hermite (A, B) form of interp::eval_all; field j in {1 c, 3 u, 5 v, 7 rho}.
GA_HD double herm_f(const SegPt &s, int j)
{
    double fk = s.a[j], fk1 = s.a[NODE + j];
    double dx = s.a[NODE] - s.a[0];
    double d = fk1 - fk;
    double A = s.a[j + 1] * dx - d;
    double B = -s.a[NODE + j + 1] * dx + d;
    return fk + s.X * (d + s.omX * (A * s.omX + B * s.X));
}

GA_HD void herm_fd(const SegPt &s, int j, double &f, double &df)
{
    double fk = s.a[j], fk1 = s.a[NODE + j];
    double dx = s.a[NODE] - s.a[0];
    double d = fk1 - fk;
    double A = s.a[j + 1] * dx - d;
    double B = -s.a[NODE + j + 1] * dx + d;
    double q = A * s.omX + B * s.X;
    f = fk + s.X * (d + s.omX * q);
    df = (d + (1.0 - 2.0 * s.X) * q + s.X * s.omX * (B - A)) * s.inv_dx;
}

GA_HD void herm_fdd(const SegPt &s, int j, double &f, double &df, double &ddf)
{
    double fk = s.a[j], fk1 = s.a[NODE + j];
    double dx = s.a[NODE] - s.a[0];
    double d = fk1 - fk;
    double A = s.a[j + 1] * dx - d;
    double B = -s.a[NODE + j + 1] * dx + d;
    double q = A * s.omX + B * s.X;
    f = fk + s.X * (d + s.omX * q);
    df = (d + (1.0 - 2.0 * s.X) * q + s.X * s.omX * (B - A)) * s.inv_dx;
    ddf = 2.0 * (B - 2.0 * A + (A - B) * 3.0 * s.X) * (s.inv_dx * s.inv_dx);
}

GA_HD double clampz(const Medium &M, double z)
{
    return fmin(fmax(z, M.zlo), M.zhi);
}

// Vertical coordinate and the "height above ground" coordinate.
template <int G> GA_HD double zcoord(const Medium &M, const double *y)
{
    return G == GEOM_SPH ? y[0] - M.r0 : y[2];
}
template <int G> GA_HD double hcoord(const double *y)
{
    return G == GEOM_SPH ? y[0] : y[2];
}

// topo::bndlyr_sc (no-slip scaling of u, v; flat ground).
GA_HD double bl_sc(const Medium &M, double z, double h)
{
    if (z < M.z_bndlyr) {
        double X = ga_exp(-(h - M.zg) * 10.0);
        return (1.0 - X) / (1.0 + X);
    }
    return 1.0;
}

// Medium values used by the RHS: c, u, v with first and second vertical
// derivatives, the winds carrying the boundary-layer factor (1-X)/(1+X),
// X = exp(-(h - zg) / 0.1), as calc_uvw computes it at every height.
struct Atm {
    double c, dc, ddc;
    double u, du, ddu;
    double v, dv, ddv;
};

template <bool SECOND>
GA_HD void medium_rhs(const Medium &M, double z, double h, int &cur, Atm &m)
{
    SegPt s = seg_point(M, clampz(M, z), cur);
    double u0, du0, ddu0 = 0.0, v0, dv0, ddv0 = 0.0;
    if (SECOND) {
        herm_fdd(s, 1, m.c, m.dc, m.ddc);
        herm_fdd(s, 3, u0, du0, ddu0);
        herm_fdd(s, 5, v0, dv0, ddv0);
    } else {
        herm_fd(s, 1, m.c, m.dc);
        herm_fd(s, 3, u0, du0);
        herm_fd(s, 5, v0, dv0);
        m.ddc = 0.0;
    }
    double X = ga_exp(-(h - M.zg) * 10.0);
    double opx = 1.0 + X;
    double iop = ga_rcp(opx);
    double iop2 = iop * iop;
    double sc0 = (1.0 - X) * iop;
    double dX = -X * 10.0;
    double dsc0 = -2.0 * dX * iop2;
    m.u = u0 * sc0;
    m.v = v0 * sc0;
    m.du = du0 * sc0 + u0 * dsc0;
    m.dv = dv0 * sc0 + v0 * dsc0;
    if (SECOND) {
        double ddX = X * 100.0;
        double ddsc0 = -2.0 * (ddX * iop2 - 2.0 * dX * dX * (iop2 * iop));
        m.ddu = ddu0 * sc0 + 2.0 * du0 * dsc0 + u0 * ddsc0;
        m.ddv = ddv0 * sc0 + 2.0 * dv0 * dsc0 + v0 * ddsc0;
    } else {
        m.ddu = m.ddv = 0.0;
    }
}

// --------------------------------------------------------------------------
// Fused RK4 stages.  Each right-hand side consumes its derivative values as
// they are produced:  acc[i] += ha * f_i  and  stg[i] = y[i] + hn * f_i,
// updating a component of the stage state in place once nothing later in the
// evaluation reads it.  (ha, hn) = (ds/6, ds/2), (ds/3, ds/2), (ds/3, ds),
// (ds/6, -) reproduce rk4solver.cpp:34-54:
//   y1 = y + k1/6 + k2/3 + k3/3 + k4/6, summed in that order.
// --------------------------------------------------------------------------
struct Stage {
    double ha, hn;
    bool last;
};

GA_HD void upd(const Stage &S, const double *y, double *stg, double *acc, int i, double f)
{
    acc[i] += S.ha * f;
    if (!S.last) stg[i] = y[i] + S.hn * f;
}

// Spherical (r, lat, lon): state r, t, p, nu_r, nu_t, nu_p, R_lt[3],
// mu_lt[3], R_lp[3], mu_lp[3].  Winds: w (radial) = 0, v (lat), u (lon).
template <int NEQ>
GA_HD void stage_sph(const Medium &M, const Stage &S, const double *y, double *stg, double *acc,
                                          int &cur)
{
    const double r = stg[0], t = stg[1];
    const double nu0 = stg[3], nu1 = stg[4], nu2 = stg[5];
    Atm m;
    medium_rhs<NEQ == 18>(M, r - M.r0, r, cur, m);

    const double nu_mag = sqrt(nu0 * nu0 + nu1 * nu1 + nu2 * nu2);
    const double inu = ga_rcp(nu_mag);
    const double cg0 = m.c * nu0 * inu;
    const double cg1 = m.c * nu1 * inu + m.v;
    const double cg2 = m.c * nu2 * inu + m.u;
    const double cg_mag = sqrt(cg0 * cg0 + cg1 * cg1 + cg2 * cg2);
    const double icg = ga_rcp(cg_mag);

    double sn, cs;
    ga_sincos(t, sn, cs);
    const double ics = ga_rcp(cs);
    const double tn = sn * ics;
    const double ir = ga_rcp(r);
    const double G1 = ir;
    const double G2 = ir * ics;

    const double T0 = -ir * (nu1 * cg1 + nu2 * cg2);
    const double T1 = nu0 * m.v + nu0 * cg1 + nu2 * cg2 * tn;
    const double T2 = nu0 * m.u * cs + (nu1 * m.u - nu2 * m.v) * sn + cg2 * (nu0 * cs - nu1 * sn);
    const double S0 = nu_mag * m.dc + nu1 * m.dv + nu2 * m.du;   // vertical gradient term

    const double f0 = cg0 * icg;
    const double f1 = G1 * cg1 * icg;
    const double f2 = G2 * cg2 * icg;
    const double f3 = -icg * (S0 + T0);
    const double f4 = -G1 * icg * T1;
    const double f5 = -G2 * icg * T2;

    if (NEQ == 18) {
        const double icg2 = icg * icg;
        const double inu2 = inu * inu;
        const double ir2 = ir * ir;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int oR = j ? 12 : 6, om = j ? 15 : 9;
            const double R0 = stg[oR], R1 = stg[oR + 1], R2 = stg[oR + 2];
            const double m0 = stg[om], m1 = stg[om + 1], m2 = stg[om + 2];
            const double dcj = R0 * m.dc;
            const double duj = R0 * m.du;
            const double dvj = R0 * m.dv;
            const double ddcj = R0 * m.ddc;
            const double dduj = R0 * m.ddu;
            const double ddvj = R0 * m.ddv;
            const double dnu = (nu0 * m0 + nu1 * m1 + nu2 * m2) * inu;
            const double cnd = m.c * inu2 * dnu;
            const double dcg0 = nu0 * inu * dcj + m.c * m0 * inu - cnd * nu0;
            const double dcg1 = nu1 * inu * dcj + m.c * m1 * inu - cnd * nu1 + dvj;
            const double dcg2 = nu2 * inu * dcj + m.c * m2 * inu - cnd * nu2 + duj;
            const double dcgm = (cg0 * dcg0 + cg1 * dcg1 + cg2 * dcg2) * icg;

            const double dG1 = -R0 * ir2;
            const double dG2 = -R0 * ir2 * ics + sn * ir * ics * ics * R1;

            const double dT0 = R0 * ir2 * (nu1 * cg1 + nu2 * cg2) - ir * (m1 * cg1 + nu1 * dcg1 + m2 * cg2 + nu2 * dcg2);
            double dT1 = m0 * m.v + nu0 * dvj;
            dT1 += m0 * cg1 + nu0 * dcg1 + (m2 * cg2 + nu2 * dcg2) * tn + nu2 * cg2 * R1 * ics * ics;
            double dT2 = (m0 * m.u + nu0 * duj) * cs - nu0 * m.u * R1 * sn;
            dT2 += (m1 * m.u + nu1 * duj - m2 * m.v - nu2 * dvj) * sn + (nu1 * m.u - nu2 * m.v) * R1 * cs;
            dT2 += dcg2 * (nu0 * cs - nu1 * sn) + cg2 * ((m0 - nu1 * R1) * cs - (m1 + nu0 * R1) * sn);

            const double cgm2 = dcgm * icg2;
            const double fR0 = dcg0 * icg - cg0 * cgm2;
            const double fR1 = dG1 * cg1 * icg + G1 * dcg1 * icg - G1 * cg1 * cgm2;
            const double fR2 = dG2 * cg2 * icg + G2 * dcg2 * icg - G2 * cg2 * cgm2;
            const double fm0 = cgm2 * S0 - icg * (dnu * m.dc + nu_mag * ddcj + m1 * m.dv + m2 * m.du + nu1 * ddvj + nu2 * dduj + dT0);
            const double fm1 = -dG1 * icg * T1 - G1 * icg * dT1;
            const double fm2 = -dG2 * icg * T2 - G2 * icg * dT2;
            (void)R2;
            upd(S, y, stg, acc, oR, fR0);
            upd(S, y, stg, acc, oR + 1, fR1);
            upd(S, y, stg, acc, oR + 2, fR2);
            upd(S, y, stg, acc, om, fm0);
            upd(S, y, stg, acc, om + 1, fm1);
            upd(S, y, stg, acc, om + 2, fm2);
        }
    }
    upd(S, y, stg, acc, 0, f0);
    upd(S, y, stg, acc, 1, f1);
    upd(S, y, stg, acc, 2, f2);
    upd(S, y, stg, acc, 3, f3);
    upd(S, y, stg, acc, 4, f4);
    upd(S, y, stg, acc, 5, f5);
}

// Cartesian (x, y, z): state x, y, z, nu_x, nu_y, nu_z, X_th[3], mu_th[3],
// X_ph[3], mu_ph[3].  Winds u (x), v (y), w = 0.  Keeps the reference's
// accumulation of the azimuth term into dv[3] (geoac.eqset.3d.strat.cpp:298),
// so dv[3] = (X_th,z + X_ph,z) dv/dz and dv[4] = 0.
template <int NEQ>
GA_HD void stage_3d(const Medium &M, const Stage &S, const double *y, double *stg, double *acc,
                                         int &cur)
{
    const double nu0 = stg[3], nu1 = stg[4], nu2 = stg[5];
    Atm m;
    medium_rhs<NEQ == 18>(M, stg[2], stg[2], cur, m);
    const double nu_mag = sqrt(nu0 * nu0 + nu1 * nu1 + nu2 * nu2);
    const double inu = ga_rcp(nu_mag);
    const double cg0 = m.c * nu0 * inu + m.u;
    const double cg1 = m.c * nu1 * inu + m.v;
    const double cg2 = m.c * nu2 * inu;
    const double cg_mag = sqrt(cg0 * cg0 + cg1 * cg1 + cg2 * cg2);
    const double icg = ga_rcp(cg_mag);
    const double S_ = nu_mag * m.dc + nu0 * m.du + nu1 * m.dv;
    const double f0 = cg0 * icg, f1 = cg1 * icg, f2 = cg2 * icg, f5 = -icg * S_;
    if (NEQ == 18) {
        const double icg2 = icg * icg;
        const double inu2 = inu * inu;
        const double dv3 = stg[8] * m.dv + stg[14] * m.dv;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int oX = j ? 12 : 6, om = j ? 15 : 9;
            const double X2 = stg[oX + 2];
            const double m0 = stg[om], m1 = stg[om + 1], m2 = stg[om + 2];
            const double dcj = X2 * m.dc;
            const double duj = X2 * m.du;
            const double dvj = j ? 0.0 : dv3;
            const double ddcj = X2 * m.ddc;
            const double dduj = X2 * m.ddu;
            const double ddvj = X2 * m.ddv;
            const double dnu = (nu0 * m0 + nu1 * m1 + nu2 * m2) * inu;
            const double cnd = m.c * inu2 * dnu;
            const double dcg0 = nu0 * inu * dcj + m.c * m0 * inu - cnd * nu0 + duj;
            const double dcg1 = nu1 * inu * dcj + m.c * m1 * inu - cnd * nu1 + dvj;
            const double dcg2 = nu2 * inu * dcj + m.c * m2 * inu - cnd * nu2;
            const double dcgm = (cg0 * dcg0 + cg1 * dcg1 + cg2 * dcg2) * icg;
            const double cgm2 = dcgm * icg2;
            const double fX0 = dcg0 * icg - cg0 * cgm2;
            const double fX1 = dcg1 * icg - cg1 * cgm2;
            const double fX2 = dcg2 * icg - cg2 * cgm2;
            const double fm2 = cgm2 * S_ - icg * (dnu * m.dc + nu_mag * ddcj + m0 * m.du + m1 * m.dv + nu0 * dduj + nu1 * ddvj);
            upd(S, y, stg, acc, oX, fX0);
            upd(S, y, stg, acc, oX + 1, fX1);
            upd(S, y, stg, acc, oX + 2, fX2);
            upd(S, y, stg, acc, om, 0.0);
            upd(S, y, stg, acc, om + 1, 0.0);
            upd(S, y, stg, acc, om + 2, fm2);
        }
    }
    upd(S, y, stg, acc, 0, f0);
    upd(S, y, stg, acc, 1, f1);
    upd(S, y, stg, acc, 2, f2);
    upd(S, y, stg, acc, 3, 0.0);
    upd(S, y, stg, acc, 4, 0.0);
    upd(S, y, stg, acc, 5, f5);
}

template <int G, int NEQ>
GA_HD void stage(const Medium &M, const Stage &S, const double *y, double *stg, double *acc,
                                      int &cur)
{
    if (G == GEOM_SPH) stage_sph<NEQ>(M, S, y, stg, acc, cur);
    else stage_3d<NEQ>(M, S, y, stg, acc, cur);
}

// set_ds: smaller steps inside the boundary layer.
template <int G>
GA_HD double set_ds(const Medium &M, const Limits &L, const double *y)
{
    double ds = L.ds_max;
    if (zcoord<G>(M, y) < M.z_bndlyr) ds -= (L.ds_max - L.ds_min) * ga_exp(-(hcoord<G>(y) - M.zg) * 5.0);
    return ds;
}

// One classical RK4 step y -> y1 (rk4solver.cpp:29-55); returns ds.  The
// four stages run as a loop that is deliberately not unrolled: unrolled, the
// compiler overlaps the stages' live ranges and the 18-equation system needs
// more than 512 registers per lane; rolled, one stage's working set bounds it.
template <int G, int NEQ>
GA_HD double rk4_step(const Medium &M, const Limits &L, const double *y, double *y1, int &cur)
{
    double stg[NEQ];
    const double ds = set_ds<G>(M, L, y);
    const double h6 = ds * (1.0 / 6.0), h3 = ds * (1.0 / 3.0), h2 = ds * 0.5;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) { stg[i] = y[i]; y1[i] = y[i]; }
#pragma clang loop unroll(disable)
    for (int st = 0; st < 4; ++st) {
        const double ha = (st == 0 || st == 3) ? h6 : h3;
        const double hn = (st == 2) ? ds : h2;
        stage<G, NEQ>(M, Stage{ha, hn, false}, y, stg, y1, cur);
    }
    return ds;
}

// --------------------------------------------------------------------------
// Sutherland & Bass absorption [dB/km] at height z with sound speed c [km/s]
// and density rho [g/cm^3].  SPH selects the spherical variant.  Quotients
// are products with refined reciprocals (ga_rcp), shared where a divisor
// repeats (1/T, 1/c, 1/f_vib): within a few ulp of the divisions, about half
// their instructions.
//
// The sum is written as  (1.003 A Q + R) * tweak * 8.685889e3  with
//   A = beta_0 sqrt((sqrt(1 + nu^2) - 1) / (2 (1 + nu^2)))   (classical),
//   Q = sqrt((1 + cchi^2) / (1 + (sigma cchi)^2)),
//   R = a_rot + a_vib,
// which is the reference's a_cl + a_rot + 0.003 a_cl + a_vib (sph) and
// a_cl * 1.003 + a_rot + a_vib (cart) with a_cl = A Q.  A depends on c and
// rho alone and carries the cancellation of sqrt(1 + nu^2) - 1 (nu ~ 1e-7
// at 65 km: 1 + nu^2 is a few dozen ulp above 1, so A steps with rho and c
// as the reference's does); Q and R are smooth functions of the height.
// --------------------------------------------------------------------------
struct SbGas {
    double cms, icms, T_z, iTz, P_z, mu, imu, nu, nu2p1, snu;
};

// The frequency-dependent members nu, nu2p1, snu of g.
GA_HD void sb_gas_freq(SbGas &g, double freq)
{
    g.nu = (8.0 * PI * freq * g.mu) * ga_rcp(3.0 * g.P_z);
    g.nu2p1 = 1.0 + g.nu * g.nu;
    g.snu = sqrt(g.nu2p1);
}

GA_HD SbGas sb_gas(double c, double rho, double freq)
{
    const double gam = 1.4, Rg = 287.058;
    const double mu_o = 18.192e-6, T_o = 293.15, S = 117.0;
    SbGas g;
    g.cms = c * 1.0e3;
    g.icms = ga_rcp(g.cms);
    g.T_z = g.cms * g.cms * (1.0 / (Rg * gam));
    g.iTz = ga_rcp(g.T_z);
    g.P_z = rho * 1.0e3 * (g.cms * g.cms) * (1.0 / gam);
    g.mu = mu_o * sqrt(g.T_z * (1.0 / T_o)) * ((1.0 + S / T_o) * ga_rcp(1.0 + S * g.iTz));
    g.imu = ga_rcp(g.mu);
    sb_gas_freq(g, freq);
    return g;
}

// A: the classical term without the rotational factor Q.
GA_HD double sb_cl(const SbGas &g, double freq)
{
    return (2.0 * PI * freq * g.icms) * sqrt(0.5 * (g.snu - 1.0) * ga_rcp(g.nu2p1));
}

// The factors of Q and R that depend on the height and the gas state but
// not on the frequency: the composition fits X_ON, the rotational
// collision number (through nn), the vibrational relaxation strengths
// A_max and inverse relaxation frequencies 1 / f_vib of O2, N2, CO2 and O3.
struct SbPoint {
    double X_ON, nn, A_max[4], ifv[4];
};

template <bool SPH>
GA_HD SbPoint sb_point(double z, const SbGas &g, const double *sbc)
{
    const double T_o = 293.15, P_o = 101325.0, mu_o = 18.192e-6;
    const double z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z;
    double X0, X1, X3, X4, X5, X6;
    const double X2 = sbc[2];
    X0 = (z > 90.0) ? ga_exp10(49.296 - (1.5524 * z) + (1.8714E-2 * z2) - (1.1069E-4 * z3) + (3.199E-7 * z4) - (3.6211E-10 * z5))
                    : sbc[0];
    X1 = (z > 76.0) ? ga_exp10((1.3972E-1) - (5.6269E-3 * z) + (3.9407E-5 * z2) - (1.0737E-7 * z3))
                    : sbc[1];
    // one exp10 of the selected exponent per species: lanes of a wave on
    // either side of a branch height then share the call
    X3 = ga_exp10((z > 80.0) ? (-4.234 - (3.0975E-2 * z))
                          : (-19.027 + (1.3093 * z) - (4.6496E-2 * z2) + (7.8543E-4 * z3) - (6.5169E-6 * z4) + (2.1343E-8 * z5)));
    X4 = ga_exp10((z > 95.0) ? (-3.2456 + (4.6642E-2 * z) - (2.6894E-4 * z2) + (5.264E-7 * z3))
                          : (-11.195 + (1.5408E-1 * z) - (1.4348E-3 * z2) + (1.0166E-5 * z3)));
    if (SPH) X5 = ga_exp10(-53.746 + (1.5439 * z) - (1.8824E-2 * z2) + (1.1587E-4 * z3) - (3.5399E-7 * z4) + (4.2609E-10 * z5));
    else     X5 = ga_exp10(-53.746 + (1.5439 * z) - (1.8824e-2 * z2) + (1.1587e-4 * z3) - (3.5399e-7 * z2) + (4.2609e-10 * z5));
    // the reference's "z > 100" branch is unreachable behind "z > 30"
    X6 = ga_exp10((z > 30.0) ? (-4.2563 + (7.6245E-2 * z) - (2.1824E-3 * z2) - (2.3010E-6 * z3) + (2.4265E-7 * z4) - (1.2500E-09 * z5))
                          : (-1.7491 + (4.4986E-2 * z) - (6.8549E-2 * z2) + (5.4639E-3 * z3) - (1.5539E-4 * z4) + (1.5063E-06 * z5)));
    SbPoint P;
    P.X_ON = (X0 + X1) * (1.0 / 0.9903);
    const double Tm13 = ga_rcbrt(g.T_z);
    const double Z_rot = ga_rcp((X1 * ga_rcp(63.3 * ga_exp(-16.7 * Tm13))) + (X0 * ga_rcp(54.1 * ga_exp(-17.3 * Tm13))));
    P.nn = (4.0 / 5.0) * sqrt(3.0 / 7.0) * Z_rot;
    const double Tr = ga_rcbrt(g.T_z * (1.0 / T_o)) - 1.0;
    const double A1 = (X0 + X1) * 24.0 * ga_exp(-9.16 * Tr);
    const double A2 = (X4 + X5) * 2400.0;
    const double B = 40400.0 * ga_exp(10.0 * Tr);
    const double C = 0.02 * ga_exp(-11.2 * Tr);
    const double D = 0.391 * ga_exp(8.41 * Tr);
    const double E = 9.0 * ga_exp(-19.9 * Tr);
    const double F = 60000.0;
    const double G = 28000.0 * ga_exp(-4.17 * Tr);
    const double H = 22000.0 * ga_exp(-7.68 * Tr);
    const double I = 15100.0 * ga_exp(-10.4 * Tr);
    const double J = 11500.0 * ga_exp(-9.17 * Tr);
    const double K = (8.48E08) * ga_exp(9.17 * Tr);
    const double L = ga_exp(-7.72 * Tr);
    const double ZZ = H * X2 + I * (X0 + 0.5 * X4) + J * (X1 + 0.5 * X5) + K * (X6 + X3);
    const double hu = 100.0 * (X3 + X6);
    const double pf = (g.P_z * (1.0 / P_o)) * (mu_o * g.imu);
    const double fv[4] = {pf * (A1 + A2 + B * hu * (C + hu) * (D + hu)), pf * (E + F * X3 + G * X6), pf * ZZ, pf * (1.2E5) * L};
    const double Xm[4] = {X0, X1, X2, X3};
    const double Th[4] = {2239.1, 3352.0, 915.0, 1037.0};
    const double Cv[4] = {2.5, 2.5, 3.0, 3.0};
    const double Cp[4] = {3.5, 3.5, 4.0, 4.0};
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double tt = Th[m] * g.iTz;
        const double e = ga_exp(-tt);
        const double ome = 1.0 - e;
        const double iome = ga_rcp(ome);
        const double C_R = (tt * tt * e) * (iome * iome);
        P.A_max[m] = (Xm[m] * (PI / 2) * C_R) * ga_rcp(Cp[m] * (Cv[m] + C_R));
        P.ifv[m] = ga_rcp(fv[m]);
    }
    return P;
}

// Q and R at frequency freq from the point factors P and the gas state g
// (whose nu, nu2p1, snu belong to freq).
template <bool SPH>
GA_HD void sb_qr(const SbPoint &P, const SbGas &g, double freq, double &Q, double &R)
{
    const double sigma = 5.0 / sqrt(21.0);
    const double chi = 3.0 * P.nn * g.nu / 4.0;
    const double cchi = 2.36 * chi;
    const double beta_0 = 2.0 * PI * freq * g.icms;
    Q = sqrt((1.0 + cchi * cchi) * ga_rcp(1.0 + (sigma * cchi) * (sigma * cchi)));
    double a_rot;
    if (SPH) {
        a_rot = beta_0 * P.X_ON * ((sigma * sigma - 1.0) * (1.0 / (2 * sigma)) * chi) *
                sqrt(0.5 * (g.snu + 1.0) * ga_rcp(g.nu2p1 * (1.0 + cchi * cchi)));
    } else {
        const double beta_1 = sqrt(0.5 * (g.snu + 1.0) * ga_rcp(g.nu2p1));
        const double alpha_2 = beta_0 * (((sigma / 2.0 - 1.0 / (2.0 * sigma)) * chi) *
                                         ga_rcp(sqrt((1.0 + cchi * cchi) * (1.0 + (sigma * cchi) * (sigma * cchi)))));
        a_rot = alpha_2 * beta_1 * P.X_ON;
    }
    double a_vib = 0.0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const double fr = freq * P.ifv[m];
        a_vib += (P.A_max[m] * g.icms) * ((2 * (freq * freq) * P.ifv[m]) * ga_rcp(1.0 + fr * fr));
    }
    R = a_rot + a_vib;
}

// Q and R at height z.
template <bool SPH>
GA_HD void sb_smooth(double z, const SbGas &g, double freq, const double *sbc, double &Q, double &R)
{
    sb_qr<SPH>(sb_point<SPH>(z, g, sbc), g, freq, Q, R);
}

template <bool SPH>
GA_HD double sb_alpha(double z, double c, double rho, double freq, double tweak, const double *sbc)
{
    const SbGas g = sb_gas(c, rho, freq);
    double Q, R;
    sb_smooth<SPH>(z, g, freq, sbc, Q, R);
    return (1.003 * sb_cl(g, freq) * Q + R) * tweak * 8.685889e3;
}

// --------------------------------------------------------------------------
// Absorption table of a stratified atmosphere at one frequency.  Q(z) and
// R(z) are smooth between the spline nodes (c and rho are cubics there) and
// between the heights where the composition fits switch formula (30, 76,
// 80, 90, 95 km), so each node interval is cut into abs_per equal pieces of
// at most ABS_PIECE km and each piece holds the degree-ABS_DEG polynomials
// in t = (z - mid) / half-width that interpolate Q and R at the Chebyshev
// points of the piece (values from sb_smooth itself).  Where a switch height
// h falls inside a piece, the record covers only the piece's lower side and
// an extra record the upper side; where h is a piece's lower end, z == h
// belongs to the piece below (the fits switch for z > h).  The step then
// evaluates A directly and Q, R by Horner's rule: agreement with sb_alpha
// within 1e-12 relative on the synthetic profiles, checked for every piece
// when the table is built (abs_fill_record's error output).  Record:
//   [0] split_lo, [1] split_hi: z <= split_lo or z > split_hi selects the
//       record at index [2] instead (-inf / +inf: no split)
//   [3] piece mid, [4] 1 / half-width, [5] unused
//   [6..13] R coefficients, [14..21] Q coefficients (ascending powers of t)
// --------------------------------------------------------------------------
constexpr int ABS_DEG = 7;
constexpr int ABS_REC = 22;
constexpr double ABS_PIECE = 0.05;
constexpr int ABS_MAX_PER = 64;          // node intervals wider than 3.2 km: no table
constexpr double ABS_TOL = 1.0e-10;      // largest relative error a table may have

struct AbsBasis {
    double t[ABS_DEG + 1];                       // Chebyshev points cos(pi (k + 1/2) / 8)
    double inv[ABS_DEG + 1][ABS_DEG + 1];        // inverse Vandermonde matrix at t
};

GA_HD double abs_poly(const double *c, double t)
{
    double p = c[ABS_DEG];
#pragma unroll
    for (int j = ABS_DEG - 1; j >= 0; --j) p = fma(p, t, c[j]);
    return p;
}

// Absorption [dB/km] from the table (z inside [zlo, zhi]; s: the spline
// point of z in node interval cur).
GA_HD double abs_alpha(const Medium &M, const SegPt &s, int cur, double z, double c, double rho, double freq)
{
    const int per = M.abs_per;
    int j = (int)(s.X * (double)per);
    j = j < 0 ? 0 : (j >= per ? per - 1 : j);
    const double *r = M.abs + ((size_t)cur * per + j) * ABS_REC;
    if (z <= r[0] || z > r[1]) r = M.abs + (size_t)r[2] * ABS_REC;
    const double t = (z - r[3]) * r[4];
    const double R = abs_poly(r + 6, t), Q = abs_poly(r + 14, t);
    const SbGas g = sb_gas(c, rho, freq);
    return (1.003 * sb_cl(g, freq) * Q + R) * M.tweak * 8.685889e3;
}

template <bool SPH>
GA_HD void abs_exact(const Medium &M, double z, double freq, int &cur, double &Q, double &R)
{
    const SegPt s = seg_point(M, z, cur);
    const SbGas g = sb_gas(herm_f(s, 1), herm_f(s, 7), freq);
    sb_smooth<SPH>(z, g, freq, M.sbc, Q, R);
}

// One table record for the piece [lo, hi] of node interval cur: the
// interpolating coefficients, and in *err the largest relative difference
// between the polynomials and sb_smooth at three points between the
// Chebyshev points.
template <bool SPH>
GA_HD void abs_fill_record(const Medium &M, const AbsBasis &B, double lo, double hi, double split_lo, double split_hi,
                           double alt, int cur, double freq, double *rec, double *err)
{
    const double mid = 0.5 * (lo + hi), hw = 0.5 * (hi - lo);
    double vq[ABS_DEG + 1], vr[ABS_DEG + 1];
    for (int k = 0; k <= ABS_DEG; ++k) abs_exact<SPH>(M, mid + hw * B.t[k], freq, cur, vq[k], vr[k]);
    rec[0] = split_lo;
    rec[1] = split_hi;
    rec[2] = alt;
    rec[3] = mid;
    rec[4] = 1.0 / hw;
    rec[5] = 0.0;
    for (int j = 0; j <= ABS_DEG; ++j) {
        double r = 0.0, q = 0.0;
        for (int k = 0; k <= ABS_DEG; ++k) {
            r = fma(B.inv[j][k], vr[k], r);
            q = fma(B.inv[j][k], vq[k], q);
        }
        rec[6 + j] = r;
        rec[14 + j] = q;
    }
    double e = 0.0;
    const double tc[3] = {-0.6, 0.1, 0.9};
    for (int k = 0; k < 3; ++k) {
        double q, r;
        abs_exact<SPH>(M, mid + hw * tc[k], freq, cur, q, r);
        e = fmax(e, fabs(abs_poly(rec + 6, tc[k]) - r) / fabs(r));
        e = fmax(e, fabs(abs_poly(rec + 14, tc[k]) - q) / fabs(q));
    }
    *err = (e == e) ? e : 1.0;
}

// Travel time and absorption of the segment y0 -> y1 evaluated at the
// midpoint (travel_time / atten whole-leg forms).
template <int G>
GA_HD void seg_book(const Medium &M, const double *y0, const double *y1, double freq,
                                         int &cur, double &dt, double &datt)
{
    double m0 = y0[0] + (y1[0] - y0[0]) * 0.5;
    double m1 = y0[1] + (y1[1] - y0[1]) * 0.5;
    double m2 = y0[2] + (y1[2] - y0[2]) * 0.5;
    double ds;
    if (G == GEOM_SPH) {
        double dr = y1[0] - y0[0], dt_ = y1[1] - y0[1], dp = y1[2] - y0[2];
        if (dp > PI) dp -= 2.0 * PI;
        else if (dp < -PI) dp += 2.0 * PI;
        double a = m0 * dt_, b = m0 * ga_cos(m1) * dp;
        ds = sqrt(dr * dr + a * a + b * b);
    } else {
        double dx = y1[0] - y0[0], dy = y1[1] - y0[1], dz = y1[2] - y0[2];
        ds = sqrt(dx * dx + dy * dy + dz * dz);
    }
    double n0 = y0[3] + (y1[3] - y0[3]) * 0.5;
    double n1 = y0[4] + (y1[4] - y0[4]) * 0.5;
    double n2 = y0[5] + (y1[5] - y0[5]) * 0.5;
    double inm = ga_rcp(sqrt(n0 * n0 + n1 * n1 + n2 * n2));
    double z = (G == GEOM_SPH) ? m0 - M.r0 : m2;
    double h = (G == GEOM_SPH) ? m0 : m2;
    SegPt s = seg_point(M, clampz(M, z), cur);
    double c = herm_f(s, 1);
    double u0 = herm_f(s, 3);
    double v0 = herm_f(s, 5);
    double rho = herm_f(s, 7);
    double sc = bl_sc(M, z, h);
    double uu = u0 * sc, vv = v0 * sc;
    double cp0, cp1, cp2;
    if (G == GEOM_SPH) { cp0 = c * n0 * inm; cp1 = c * n1 * inm + vv; cp2 = c * n2 * inm + uu; }
    else               { cp0 = c * n0 * inm + uu; cp1 = c * n1 * inm + vv; cp2 = c * n2 * inm; }
    dt = ds * ga_rcp(sqrt(cp0 * cp0 + cp1 * cp1 + cp2 * cp2));
    const double alpha = (M.abs && z >= M.zlo && z <= M.zhi)
                             ? abs_alpha(M, s, cur, z, c, rho, freq)
                             : sb_alpha<G == GEOM_SPH>(z, c, rho, freq, M.tweak, M.sbc);
    datt = alpha * ds;
}

// Point values c, u, v, rho at a state (atmo::c/u/v/rho with bndlyr_sc).
template <int G>
GA_HD void point_medium(const Medium &M, const double *y, int &cur,
                                             double &c, double &u, double &v, double &rho)
{
    double z = zcoord<G>(M, y);
    SegPt s = seg_point(M, clampz(M, z), cur);
    c = herm_f(s, 1);
    double sc = bl_sc(M, z, hcoord<G>(y));
    u = herm_f(s, 3) * sc;
    v = herm_f(s, 5) * sc;
    rho = herm_f(s, 7);
}

}  // namespace ga
