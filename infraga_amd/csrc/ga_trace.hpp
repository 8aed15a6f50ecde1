// ga_trace.hpp -- per-lane ray tracing loop of the fan engine: launch state,
// region / ground checks, reflection, arrival evaluation and the step loop.
// Included by ga_engine.hip (device) and usable from host-side builds.
#pragma once

#include "ga_device.hpp"
#include "../../include/infraga_gpu.h"

namespace ga {

// Threads per workgroup of the trace kernels.  The spline table takes one
// LDS copy per workgroup (120 KB for 1501 levels), so one workgroup runs per
// CU and its size sets the waves per SIMD; the register budget of the
// equation set bounds it: 18 spherical equations want ~280 registers per
// lane and run two waves per SIMD with a few spilled to scratch (measured
// faster than one wave without spills), 18 Cartesian ones ~215 (two), 6
// equations ~165 (three).
#ifndef GA_WG_SPH18
#define GA_WG_SPH18 512
#endif
#ifndef GA_WG_3D18
#define GA_WG_3D18 512
#endif
#ifndef GA_WG6
#define GA_WG6 768
#endif
template <int G, int NEQ> constexpr int trace_wg()
{
    return NEQ == 18 ? (G == GEOM_SPH ? GA_WG_SPH18 : GA_WG_3D18) : GA_WG6;
}

// Raw arrival record written by the step loop at each ground return:
// row K (18 values), tt, att, rmax, theta, phi, c0, nu0.  The arrival row
// itself is evaluated by arrival_kernel, which keeps the Jacobian/amplitude
// code out of the step loop's register budget.
constexpr int RAW = 25;

// Ray-path output (write_rays / write_caustics of run_prop, the
// eigenray-N.dat files of find_eigenray).  Records of one item (ray or
// eigenray) are appended by the lane tracing it into chunks of PCHUNK
// records taken from a device counter; each chunk carries its item and its
// order within the item, so the host puts the records in sequence without a
// sort.  Per-item record counts are kept also beyond the capacity, so a
// caller can size a second pass exactly.
constexpr int PCHUNK = 32;

// Waveform rows: an internal path mode of the -wnl_wvfrm trace
// (run_wnl_wvfrm, main.infraga.sph.cpp:1336-1401, keeps the whole solution
// array of every leg).  Every row m = 0..K of every leg is written as two
// records: PT_WV_STATE holds the state y[0..5] in (pos0, pos1, z, tl,
// absorption, time); PT_WV_MEDIUM holds the Jacobian determinant (pos0),
// c (pos1), rho (z), the group speed |c nu/|nu| + wind| (tl) and the travel
// time of the segment from row m-1 into row m (absorption; 0 for row 0).
constexpr int PATH_WVROWS = 4;
constexpr int PT_WV_STATE = 2;
constexpr int PT_WV_MEDIUM = 3;

struct PathOut {
    ga_path_pt *pts;           // [max_chunks][PCHUNK]
    int32_t *chunk_item;       // item of each chunk
    int32_t *chunk_seq;        // position of the chunk within its item
    int32_t *chunk_fill;       // records in the chunk
    int32_t *item_pts;         // records per item (all counted)
    unsigned int *ctr;         // chunks handed out
    unsigned int max_chunks;
    int flags;                 // GA_PATH_RAYS | GA_PATH_CAUSTICS
};

// Per-lane cursor into PathOut.
struct PathLane {
    unsigned int chunk;        // current chunk (valid when fill > 0 or n > 0)
    int fill, seq, n;
};

GA_HD void path_reset(PathLane &c)
{
    c.chunk = 0xffffffffu;
    c.fill = PCHUNK;
    c.seq = 0;
    c.n = 0;
}

GA_HD void path_put(const PathOut &P, PathLane &c, int item, const ga_path_pt &r)
{
    if (c.fill == PCHUNK) {
        if (c.chunk < P.max_chunks) P.chunk_fill[c.chunk] = c.fill;
        c.chunk = atomicAdd(P.ctr, 1u);
        c.fill = 0;
        if (c.chunk < P.max_chunks) {
            P.chunk_item[c.chunk] = item;
            P.chunk_seq[c.chunk] = c.seq;
        }
        ++c.seq;
    }
    if (c.chunk < P.max_chunks) P.pts[(size_t)c.chunk * PCHUNK + c.fill] = r;
    ++c.fill;
    ++c.n;
}

GA_HD void path_close(const PathOut &P, PathLane &c, int item)
{
    if (c.n > 0 && c.chunk < P.max_chunks) P.chunk_fill[c.chunk] = c.fill;
    if (P.item_pts) P.item_pts[item] = c.n;
    path_reset(c);
}

struct FanArgs {
    Medium M;
    Limits L;
    double src0, src1, src2;   // sph: r, lat, lon [rad]; 3D: x, y, z [km]
    double c_src;              // c at the source (arrival inclination)
    double freq, turn_ht_min, rmax_init;
    int bounces;
    int n_rays;
    int max_steps;             // prop_rk4's row budget per leg (s_max / ds_min) - 1
    int lds;                   // 1: table staged in LDS
    const double *launches;    // (incl_deg, az_deg) pairs
    double *raw;               // RAW doubles per arrival slot (row K + leg sums)
    double *ypbuf;             // per-lane copy of row k-1, [nlanes][18]
    int nlanes;
    int one_shot;              // 1: lane i traces item i only (grid covers all items);
                               // 0: lanes pull items from the queue counter
    ga_arrival *out;
    int32_t *counts;
    int32_t *steps;
    unsigned int *queue;
    PathOut path;              // path output of the write-mode kernels
};

// ---------------------------------------------------------------------------
// Launch state (set_initial: geoac.eqset.sph.strat.cpp:59-104,
// geoac.eqset.3d.strat.cpp:58-105).
// ---------------------------------------------------------------------------
// th, ph: launch inclination and azimuth in radians (theta, phi of geoac).
template <int G, int NEQ>
GA_HD void init_ray_rad(const FanArgs &A, const Medium &M, double th, double ph, double *y,
                        double &c0, double &nu0s, int &cur)
{
    double src[3] = {A.src0, A.src1, A.src2};
    double c, u, v, rho;
    point_medium<G>(M, src, cur, c, u, v, rho);
    c0 = c;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    double mach[3], n0[3], mt[3], mp[3];
    if (G == GEOM_SPH) {
        mach[0] = 0.0; mach[1] = v / c0; mach[2] = u / c0;
        n0[0] = st; n0[1] = ct * sp; n0[2] = ct * cp;
        mt[0] = ct; mt[1] = -st * sp; mt[2] = -st * cp;
        mp[0] = 0.0; mp[1] = ct * cp; mp[2] = -ct * sp;
    } else {
        mach[0] = u / c0; mach[1] = v / c0; mach[2] = 0.0;
        n0[0] = ct * cp; n0[1] = ct * sp; n0[2] = st;
        mt[0] = -st * cp; mt[1] = -st * sp; mt[2] = ct;
        mp[0] = -ct * sp; mp[1] = ct * cp; mp[2] = 0.0;
    }
    double msc = 1.0 + (n0[0] * mach[0] + n0[1] * mach[1] + n0[2] * mach[2]);
    double imsc = 1.0 / msc;
    nu0s = imsc;
    y[0] = src[0]; y[1] = src[1]; y[2] = src[2];
    for (int i = 0; i < 3; ++i) y[3 + i] = n0[i] * imsc;
    if (NEQ == 18) {
        double a = mt[0] * mach[0] + mt[1] * mach[1] + mt[2] * mach[2];
        double b = mp[0] * mach[0] + mp[1] * mach[1] + mp[2] * mach[2];
        for (int i = 0; i < 3; ++i) {
            y[6 + i] = 0.0;
            y[12 + i] = 0.0;
            y[9 + i] = mt[i] * imsc - n0[i] * (imsc * imsc) * a;
            y[15 + i] = mp[i] * imsc - n0[i] * (imsc * imsc) * b;
        }
    }
}

// incl, az: degrees (inclination above horizontal, azimuth clockwise from north).
template <int G, int NEQ>
GA_HD void init_ray(const FanArgs &A, const Medium &M, double incl, double az, double *y,
                    double &c0, double &nu0s, double &th, double &ph, int &cur)
{
    ph = PI / 2.0 - az * (PI / 180.0);
    th = incl * (PI / 180.0);
    init_ray_rad<G, NEQ>(A, M, th, ph, y, c0, nu0s, cur);
}

// break_check (sph: wraps longitude in place) / ground_check.
template <int G>
GA_HD bool leaves_region(const FanArgs &A, const Medium &M, const Limits &L, double *y,
                                              double sin_half_dummy)
{
    (void)sin_half_dummy;
    if (G == GEOM_SPH) {
        if (y[2] > PI) y[2] -= 2.0 * PI;
        else if (y[2] < -PI) y[2] += 2.0 * PI;
        if (y[1] < -PI / 2.0 || y[1] > PI / 2.0) return true;
        if (y[0] - M.r0 > L.alt_max) return true;
        // haversine range test gc_dist > rng_max as h > sin^2(rng_max / 2 r0).
        // Since sin^2(x) <= x^2 and cos(lat) cos(lat_src) <= 1, h is at most
        // (d_lat^2 + d_lon^2) / 4; while that bound stays clearly below the
        // threshold the ray is inside and the sines need not be evaluated
        // (the margin dwarfs the rounding of the exact form, so both forms
        // decide alike).
        const double d1 = A.src1 - y[1], d2 = A.src2 - y[2];
        if ((d1 * d1 + d2 * d2) * (0.25 * (1.0 + 1.0e-9)) >= L.gc_thresh) {
            double s1 = sin(d1 * 0.5);
            double s2 = sin(d2 * 0.5);
            double h = s1 * s1 + cos(y[1]) * cos(A.src1) * (s2 * s2);
            if (h > L.gc_thresh) return true;
        }
        if (y[1] < L.b0 || y[1] > L.b1) return true;
        if (y[2] < L.b2 || y[2] > L.b3) return true;
        return false;
    } else {
        double dx = y[0] - A.src0, dy = y[1] - A.src1;
        if (sqrt(dx * dx + dy * dy) > L.rng_max) return true;
        if (y[0] > L.b1 || y[0] < L.b0) return true;
        if (y[1] > L.b3 || y[1] < L.b2) return true;
        if (y[2] > L.alt_max) return true;
        return false;
    }
}

template <int G>
GA_HD bool below_ground(const Medium &M, const double *y)
{
    double h = hcoord<G>(y);
    return (zcoord<G>(M, y) <= M.z_max) && (h < M.zg);
}

// Jacobian determinant (geoac.eqset.sph.strat.cpp:800-823, 3d.strat.cpp:704-723).
template <int G>
GA_HD double jacobian(const double *y, double c, double u, double v)
{
    double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    double p0, p1, p2;
    if (G == GEOM_SPH) { p0 = c * y[3] / nm; p1 = c * y[4] / nm + v; p2 = c * y[5] / nm + u; }
    else               { p0 = c * y[3] / nm + u; p1 = c * y[4] / nm + v; p2 = c * y[5] / nm; }
    double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2);
    if (G == GEOM_SPH) {
        double r = y[0], ct = cos(y[1]);
        double drs = p0 / pm, dts = 1.0 / r * p1 / pm, dps = 1.0 / (r * ct) * p2 / pm;
        return r * r * ct * (drs * (y[7] * y[14] - y[13] * y[8]) - y[6] * (dts * y[14] - dps * y[13])
                             + y[12] * (dts * y[8] - dps * y[7]));
    }
    double dxs = p0 / pm, dys = p1 / pm, dzs = p2 / pm;
    return dxs * (y[7] * y[14] - y[13] * y[8]) - y[6] * (dys * y[14] - dzs * y[13]) + y[12] * (dys * y[8] - dzs * y[7]);
}

// 20 log10 of the transport-equation amplitude at row y, whose medium
// values are c, u, v, rho (geoac::amp, geoac.eqset.sph.strat.cpp:825-849,
// geoac.eqset.3d.strat.cpp:725-748).
template <int G>
GA_HD double amp_db(const FanArgs &A, const Medium &M, const double *y, double c, double u, double v, double rho,
                    double c0, double nu0s, double th, double ph, int &cur)
{
    double src[3] = {A.src0, A.src1, A.src2};
    double cs, us, vs, rhos;
    point_medium<G>(M, src, cur, cs, us, vs, rhos);
    double jac = jacobian<G>(y, c, u, v);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    double num, den;
    if (G == GEOM_SPH) {
        double nm = (c0 - y[4] * v - y[5] * u) / c;
        double q0 = c0 * (nu0s * st) / nu0s, q1 = c0 * (nu0s * ct * sp) / nu0s + vs, q2 = c0 * (nu0s * ct * cp) / nu0s + us;
        double p0 = c * y[3] / nm, p1 = c * y[4] / nm + v, p2 = c * y[5] / nm + u;
        double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2), qm = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
        num = rho * nm * (c * c * c) * qm * ct;
        den = rhos * nu0s * (cs * cs * cs) * pm * jac;
    } else {
        double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
        double nm0 = 1.0 - (y[3] * us + y[4] * vs) / c0;
        double p0 = c * y[3] / nm + u, p1 = c * y[4] / nm + v, p2 = c * y[5] / nm;
        double q0 = c0 * ct * cp + us, q1 = c0 * ct * sp + vs, q2 = c0 * st;
        double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2), qm = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
        num = rho * nm * (c * c * c) * qm * ct;
        den = rhos * nm0 * (c0 * c0 * c0) * pm * jac;
    }
    return 20.0 * log10(sqrt(fabs(num / den)));
}

// One ray-path / caustic record of row y (raypath and caustics columns of
// main.infraga.sph.cpp:471-490, main.infraga.3d.cpp:458-478): sph lat, lon
// [deg] and r - r0; 3D x, y and max(z, ground).
template <int G>
GA_HD ga_path_pt path_row(const Medium &M, const double *y, int item, int leg, int m, int kind, double tl,
                          double att, double tt)
{
    ga_path_pt P;
    P.launch_idx = item;
    P.leg = leg;
    P.step = m;
    P.kind = kind;
    if (G == GEOM_SPH) {
        P.pos0 = y[1] * (180.0 / PI);
        P.pos1 = y[2] * (180.0 / PI);
        P.z = y[0] - M.r0;
    } else {
        P.pos0 = y[0];
        P.pos1 = y[1];
        P.z = y[2] > M.zg ? y[2] : M.zg;
    }
    P.tl = tl;
    P.absorption = -2.0 * att;
    P.time = tt;
    return P;
}

// The two waveform-row records of row y (PATH_WVROWS); dt: travel time of
// the segment into the row.
template <int G>
GA_HD void wv_rows_put(const PathOut &P, PathLane &pl, const Medium &M, const double *y, int item, int leg, int m,
                       double dt, int &cur)
{
    double c, u, v, rho;
    point_medium<G>(M, y, cur, c, u, v, rho);
    const double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    double g0, g1, g2;
    if (G == GEOM_SPH) { g0 = c * y[3] / nm; g1 = c * y[4] / nm + v; g2 = c * y[5] / nm + u; }
    else               { g0 = c * y[3] / nm + u; g1 = c * y[4] / nm + v; g2 = c * y[5] / nm; }
    ga_path_pt a, b;
    a.launch_idx = b.launch_idx = item;
    a.leg = b.leg = leg;
    a.step = b.step = m;
    a.kind = PT_WV_STATE;
    a.pos0 = y[0]; a.pos1 = y[1]; a.z = y[2];
    a.tl = y[3]; a.absorption = y[4]; a.time = y[5];
    b.kind = PT_WV_MEDIUM;
    b.pos0 = jacobian<G>(y, c, u, v);
    b.pos1 = c;
    b.z = rho;
    b.tl = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    b.absorption = dt;
    b.time = 0.0;
    path_put(P, pl, item, a);
    path_put(P, pl, item, b);
}

// Arrival record (main.infraga.sph.cpp:523-544, main.infraga.3d.cpp:511-530).
template <int G, int NEQ>
GA_HD ga_arrival arrival_row(const FanArgs &A, const Medium &M, const double *y, int ray, int bnc,
                             double incl, double az, double tt, double att, double rmax,
                             double c0, double nu0s, double th, double ph, int &cur)
{
    ga_arrival R;
    R.launch_idx = ray;
    R.n_bounce = bnc;
    R.incl = incl;
    R.az = az;
    double c, u, v, rho;
    point_medium<G>(M, y, cur, c, u, v, rho);
    double baz;
    if (G == GEOM_SPH) {
        R.incl_arr = -asin(c / A.c_src * y[3]) * (180.0 / PI);
        baz = 90.0 - atan2(-y[4], -y[5]) * (180.0 / PI);
        R.pos0 = y[1] * (180.0 / PI);
        R.pos1 = y[2] * (180.0 / PI);
        double s1 = sin((A.src1 - y[1]) / 2.0), s2 = sin((A.src2 - y[2]) / 2.0);
        double gc = 2.0 * M.r0 * asin(sqrt(s1 * s1 + cos(y[1]) * cos(A.src1) * (s2 * s2)));
        R.cel = gc / tt;
    } else {
        double cg = herm_f(seg_point(M, clampz(M, M.z0), cur), 1);
        R.incl_arr = -asin(cg / A.c_src * y[5]) * 180.0 / PI;
        baz = 90.0 - atan2(-y[4], -y[3]) * 180.0 / PI;
        R.pos0 = y[0];
        R.pos1 = y[1];
        double dx = y[0] - A.src0, dy = y[1] - A.src1;
        R.cel = sqrt(dx * dx + dy * dy) / tt;
    }
    if (baz < -180.0) baz += 360.0;
    else if (baz > 180.0) baz -= 360.0;
    R.back_az = baz;
    R.time = tt;
    R.turn_ht = rmax;
    R.absorption = -2.0 * att;
    R.tl = (NEQ == 18) ? amp_db<G>(A, M, y, c, u, v, rho, c0, nu0s, th, ph, cur) : 0.0;
    return R;
}

template <int G, int NEQ>
GA_HD void eval_arrival(const FanArgs &A, const Medium &M, const double *y, int ray, int bnc,
                        double incl, double az, double tt, double att, double rmax,
                        double c0, double nu0s, double th, double ph, int &cur)
{
    A.out[(size_t)ray * (A.bounces + 1) + bnc] =
        arrival_row<G, NEQ>(A, M, y, ray, bnc, incl, az, tt, att, rmax, c0, nu0s, th, ph, cur);
}

// Specular reflection off flat ground (set_refl with zero terrain slope:
// geoac.eqset.sph.strat.cpp:120-211, geoac.eqset.3d.strat.cpp:114-205).
// yk2, yk1 are rows K-2 and K-1 of the leg; result overwrites y.
template <int G, int NEQ>
GA_HD void reflect(const Medium &M, const double *yk2, const double *yk1, double *y,
                                        double c0, int &cur)
{
    const int iv = (G == GEOM_SPH) ? 0 : 2;
    double dr_k = yk1[iv] - yk2[iv];
    double dr_g = M.zg - yk1[iv];
    double prev[NEQ];
#pragma unroll
    for (int i = 0; i < NEQ; ++i) prev[i] = yk1[i] + (yk1[i] - yk2[i]) / dr_k * dr_g;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) y[i] = prev[i];
    y[iv] = M.zg;
    y[3 + iv] = -prev[3 + iv];
    if (NEQ == 18) {
        y[6 + iv] = -prev[6 + iv];
        y[12 + iv] = -prev[12 + iv];
        // medium at the ground: atmo::c, dc and the wind derivatives with the
        // boundary-layer factor at zero height (X = 1: sc = 0, dsc = 5)
        SegPt s = seg_point(M, clampz(M, M.z0), cur);
        double cg, dcg, u0, du0, v0, dv0;
        herm_fd(s, 1, cg, dcg);
        herm_fd(s, 3, u0, du0);
        herm_fd(s, 5, v0, dv0);
        double X = exp(-(M.zg - M.zg) * 10.0);
        double sc = (1.0 - X) / (1.0 + X);
        double dsc = -2.0 * (-X * 10.0) / ((1.0 + X) * (1.0 + X));
        double duz = u0 * dsc + du0 * sc, dvz = v0 * dsc + dv0 * sc;
        double dnv;
        if (G == GEOM_SPH) {
            double rg = M.zg;
            dnv = c0 / cg * dcg + prev[4] * dvz + prev[5] * duz;
            dnv += cg * cg / c0 * 1.0 / rg * (prev[4] * prev[4] + prev[5] * prev[5]);
            dnv *= -1.0 / cg;
        } else {
            dnv = -1.0 / cg * (c0 / cg * dcg + prev[3] * duz + prev[4] * dvz);
        }
        double ds0_t = -c0 / cg * prev[6 + iv] / prev[3 + iv];
        double ds0_p = -c0 / cg * prev[12 + iv] / prev[3 + iv];
        y[9 + iv] = -prev[9 + iv] + (-2.0 * dnv) * ds0_t;
        y[15 + iv] = -prev[15 + iv] + (-2.0 * dnv) * ds0_p;
    }
}

// Per-lane trace loop: pulls rays from next_ray() until it returns a value
// >= A.n_rays.  Shared by the fan kernel and host-side builds of this code.
// PATHS: the write_rays / write_caustics bookkeeping of run_prop
// (main.infraga.sph.cpp:456-496, main.infraga.3d.cpp:442-487): the segments
// of rows 1..k-1 of each leg are added one by one to the running travel time
// and absorption (the segment into the sub-ground row k is not), rows m == 1
// or m % 15 == 0 are written to A.path (GA_PATH_RAYS), and a sign change of
// the Jacobian determinant between rows m-1 and m is a caustic
// (GA_PATH_CAUSTICS, 18 equations); PATH_WVROWS (18 equations) writes
// every row 0..k of every leg as waveform-row records.  Otherwise each leg
// is summed whole (rows 0..k) and added to the totals at its end
// (write_rays=false).
template <int G, int NEQ, bool PATHS = false, class NextRay>
GA_HD void lane_loop(const FanArgs &A, const Medium &M, int lane, NextRay next_ray)
{
    const Limits L = A.L;

    double y[NEQ];
    double *ypl = A.ypbuf + (size_t)lane * 18;
    double s = 0.0, tt = 0.0, att = 0.0, ttl = 0.0, attl = 0.0, rmax = 0.0;
    double c0 = 0.0, nu0s = 0.0, th = 0.0, ph = 0.0, incl = 0.0, az = 0.0;
    int k = 0, bnc = 0, nst = 0, ray = 0, cur = 0, narr = 0;
    bool fetch = true;
    PathLane pl;
    double d_prev = 0.0;
    if (PATHS) path_reset(pl);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) y[i] = 0.0;

    for (;;) {
        if (fetch) {
            ray = next_ray();
            if (ray >= A.n_rays) break;
            incl = A.launches[2 * ray];
            az = A.launches[2 * ray + 1];
            init_ray<G, NEQ>(A, M, incl, az, y, c0, nu0s, th, ph, cur);
            s = 0.0; tt = 0.0; att = 0.0; ttl = 0.0; attl = 0.0; rmax = A.rmax_init;
            k = 0; bnc = 0; nst = 0; narr = 0;
            fetch = false;
            if constexpr (PATHS && NEQ == 18) {
                if (A.path.flags & PATH_WVROWS) wv_rows_put<G>(A.path, pl, M, y, ray, 0, 0, 0.0, cur);
            }
        }
        double y1[NEQ];
        s += rk4_step<G, NEQ>(M, L, y, y1, cur);
        ++k;
        ++nst;
        bool brk = false, end = false;
        if (s > L.s_max) { brk = true; end = true; }
        else if (leaves_region<G>(A, M, L, y1, 0.0)) { brk = true; end = true; }
        else if (below_ground<G>(M, y1)) { end = true; }
        else if (k >= A.max_steps) { brk = true; end = true; }
        else if (!(y1[0] == y1[0])) { brk = true; end = true; }   // non-finite state: stop the ray

        if (PATHS) {
            const bool wv = NEQ == 18 && (A.path.flags & PATH_WVROWS);
            double dt = 0.0, da = 0.0;
            if (!end || wv) seg_book<G>(M, y, y1, A.freq, cur, dt, da);
            if (!end) {
                tt += dt;
                att += da;
                const bool caus = NEQ == 18 && (A.path.flags & GA_PATH_CAUSTICS);
                const bool samp = (A.path.flags & GA_PATH_RAYS) && (k == 1 || k % 15 == 0);
                if (caus || samp) {
                    double c, u, v, rho;
                    point_medium<G>(M, y1, cur, c, u, v, rho);
                    if (samp) {
                        const double tl = NEQ == 18 ? amp_db<G>(A, M, y1, c, u, v, rho, c0, nu0s, th, ph, cur) : 0.0;
                        path_put(A.path, pl, ray, path_row<G>(M, y1, ray, bnc, k, GA_PT_RAY, tl, att, tt));
                    }
                    if (caus) {
                        const double d = jacobian<G>(y1, c, u, v);
                        if (k == 1) d_prev = d;
                        if (d * d_prev < 0.0)
                            path_put(A.path, pl, ray, path_row<G>(M, y1, ray, bnc, k, GA_PT_CAUSTIC, 0.0, att, tt));
                        d_prev = d;
                    }
                }
            }
            if constexpr (NEQ == 18) {
                if (wv) wv_rows_put<G>(A.path, pl, M, y1, ray, bnc, k, dt, cur);
            }
        } else {
            double dt, da;
            seg_book<G>(M, y, y1, A.freq, cur, dt, da);
            ttl += dt;
            attl += da;
        }
        rmax = fmax(rmax, zcoord<G>(M, y));

        if (!end) {
            if (zcoord<G>(M, y) < L.z_keep) {
#pragma unroll
                for (int i = 0; i < NEQ; ++i) ypl[i] = y[i];
            }
#pragma unroll
            for (int i = 0; i < NEQ; ++i) y[i] = y1[i];
            continue;
        }
        // leg finished at row K = k (y1); rows K-1 = y, K-2 = yp
        tt += ttl;
        att += attl;
        ttl = 0.0;
        attl = 0.0;
        bool done = false;
        if (rmax < M.z0 + A.turn_ht_min) done = true;
        else if (brk || k < 2 || (G == GEOM_SPH && (tt / 3600.0) > L.time_max)) done = true;
        else {
            {
                double *r = A.raw + ((size_t)ray * (A.bounces + 1) + bnc) * RAW;
#pragma unroll
                for (int i = 0; i < NEQ; ++i) r[i] = y1[i];
                r[18] = tt; r[19] = att; r[20] = rmax; r[21] = th; r[22] = ph; r[23] = c0; r[24] = nu0s;
            }
            ++narr;
            if (bnc == A.bounces) done = true;
            else {
                double yr[NEQ], yp[NEQ];
#pragma unroll
                for (int i = 0; i < NEQ; ++i) yp[i] = ypl[i];
                reflect<G, NEQ>(M, yp, y, yr, c0, cur);
#pragma unroll
                for (int i = 0; i < NEQ; ++i) y[i] = yr[i];
                ++bnc;
                s = 0.0;
                k = 0;
                if constexpr (PATHS && NEQ == 18) {
                    if (A.path.flags & PATH_WVROWS) wv_rows_put<G>(A.path, pl, M, y, ray, bnc, 0, 0.0, cur);
                }
            }
        }
        if (done) {
            A.counts[ray] = narr;
            A.steps[ray] = nst;
            if (PATHS) path_close(A.path, pl, ray);
            fetch = true;
        }
    }
}


// ---------------------------------------------------------------------------
// Rotating slice scheduler of the fan (write_rays=false bookkeeping).
//
// Every ray that is not finished sits in a ring of records in HBM.  One slice
// takes up to `lanes` records from the head of the ring, spread evenly over
// the workgroups, advances each ray by at most `slice_steps` RK4 steps and
// appends the records of the rays still running at the tail, packed by a
// wavefront ballot (one atomic per wave).  Rays are therefore visited round
// robin, every lane works on a live ray while at least `lanes` rays are
// left, and once fewer are left the survivors occupy whole waves spread over
// all CUs instead of a few live lanes in every wave.  The per-ray arithmetic
// is that of lane_loop step for step: results do not depend on the slice
// length or on the lane a ray lands in.
// ---------------------------------------------------------------------------
constexpr int RING_SC = 8;     // s, tt, att, ttl, attl, rmax, c0, nu0s
constexpr int RING_IV = 6;     // ray, k (< 0: not started), bnc, nst, narr, cur

struct RayRing {
    double *y;                 // [18][cap] state row of each record
    double *sc;                // [RING_SC][cap]
    int32_t *iv;               // [RING_IV][cap]
    unsigned long long *ctl;   // head, tail, append counter
    unsigned long long cap;    // records the ring holds (>= rays + lanes)
    int lanes;                 // records one slice takes at most
    int slice_steps;           // RK4 steps per ray per slice
};

// One ray's record in registers.
template <int NEQ>
struct RayRec {
    double y[NEQ];
    double s, tt, att, ttl, attl, rmax, c0, nu0s;
    int ray, k, bnc, nst, narr, cur;
};

template <int NEQ>
GA_HD void ring_load(const RayRing &R, unsigned long long q, RayRec<NEQ> &r)
{
    const unsigned long long c = R.cap;
    r.ray = R.iv[q];
    r.k = R.iv[c + q];
    if (r.k < 0) return;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) r.y[i] = R.y[i * c + q];
    r.s = R.sc[q]; r.tt = R.sc[c + q]; r.att = R.sc[2 * c + q]; r.ttl = R.sc[3 * c + q];
    r.attl = R.sc[4 * c + q]; r.rmax = R.sc[5 * c + q]; r.c0 = R.sc[6 * c + q]; r.nu0s = R.sc[7 * c + q];
    r.bnc = R.iv[2 * c + q]; r.nst = R.iv[3 * c + q]; r.narr = R.iv[4 * c + q]; r.cur = R.iv[5 * c + q];
}

template <int NEQ>
GA_HD void ring_store(const RayRing &R, unsigned long long q, const RayRec<NEQ> &r)
{
    const unsigned long long c = R.cap;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) R.y[i * c + q] = r.y[i];
    R.sc[q] = r.s; R.sc[c + q] = r.tt; R.sc[2 * c + q] = r.att; R.sc[3 * c + q] = r.ttl;
    R.sc[4 * c + q] = r.attl; R.sc[5 * c + q] = r.rmax; R.sc[6 * c + q] = r.c0; R.sc[7 * c + q] = r.nu0s;
    R.iv[q] = r.ray; R.iv[c + q] = r.k; R.iv[2 * c + q] = r.bnc; R.iv[3 * c + q] = r.nst;
    R.iv[4 * c + q] = r.narr; R.iv[5 * c + q] = r.cur;
}

// Advances ray record r by at most `budget` RK4 steps with the leg, arrival
// and reflection handling of lane_loop (write_rays=false sums).  Row k-1 of
// a ray near the ground is kept in A.ypbuf at the ray's index.  Returns
// true while the ray is unfinished; a finished ray has its arrival and step
// counts written.
template <int G, int NEQ>
GA_HD bool slice_trace(const FanArgs &A, const Medium &M, RayRec<NEQ> &r, int budget)
{
    const Limits L = A.L;
    const double incl = A.launches[2 * r.ray], az = A.launches[2 * r.ray + 1];
    const double ph = PI / 2.0 - az * (PI / 180.0), th = incl * (PI / 180.0);
    double *ypl = A.ypbuf + (size_t)r.ray * 18;
    if (r.k < 0) {
        r.cur = 0;
        init_ray_rad<G, NEQ>(A, M, th, ph, r.y, r.c0, r.nu0s, r.cur);
        r.s = 0.0; r.tt = 0.0; r.att = 0.0; r.ttl = 0.0; r.attl = 0.0; r.rmax = A.rmax_init;
        r.k = 0; r.bnc = 0; r.nst = 0; r.narr = 0;
    }
    for (int it = 0; it < budget; ++it) {
        double y1[NEQ];
        r.s += rk4_step<G, NEQ>(M, L, r.y, y1, r.cur);
        ++r.k;
        ++r.nst;
        bool brk = false, end = false;
        if (r.s > L.s_max) { brk = true; end = true; }
        else if (leaves_region<G>(A, M, L, y1, 0.0)) { brk = true; end = true; }
        else if (below_ground<G>(M, y1)) { end = true; }
        else if (r.k >= A.max_steps) { brk = true; end = true; }
        else if (!(y1[0] == y1[0])) { brk = true; end = true; }   // non-finite state: stop the ray
        {
            double dt, da;
            seg_book<G>(M, r.y, y1, A.freq, r.cur, dt, da);
            r.ttl += dt;
            r.attl += da;
        }
        r.rmax = fmax(r.rmax, zcoord<G>(M, r.y));
        if (!end) {
            if (zcoord<G>(M, r.y) < L.z_keep) {
#pragma unroll
                for (int i = 0; i < NEQ; ++i) ypl[i] = r.y[i];
            }
#pragma unroll
            for (int i = 0; i < NEQ; ++i) r.y[i] = y1[i];
            continue;
        }
        // leg finished at row K = k (y1); rows K-1 = r.y, K-2 = ypl
        r.tt += r.ttl;
        r.att += r.attl;
        r.ttl = 0.0;
        r.attl = 0.0;
        bool done = false;
        if (r.rmax < M.z0 + A.turn_ht_min) done = true;
        else if (brk || r.k < 2 || (G == GEOM_SPH && (r.tt / 3600.0) > L.time_max)) done = true;
        else {
            {
                double *w = A.raw + ((size_t)r.ray * (A.bounces + 1) + r.bnc) * RAW;
#pragma unroll
                for (int i = 0; i < NEQ; ++i) w[i] = y1[i];
                w[18] = r.tt; w[19] = r.att; w[20] = r.rmax; w[21] = th; w[22] = ph; w[23] = r.c0; w[24] = r.nu0s;
            }
            ++r.narr;
            if (r.bnc == A.bounces) done = true;
            else {
                double yr[NEQ], yp[NEQ];
#pragma unroll
                for (int i = 0; i < NEQ; ++i) yp[i] = ypl[i];
                reflect<G, NEQ>(M, yp, r.y, yr, r.c0, r.cur);
#pragma unroll
                for (int i = 0; i < NEQ; ++i) r.y[i] = yr[i];
                ++r.bnc;
                r.s = 0.0;
                r.k = 0;
            }
        }
        if (done) {
            A.counts[r.ray] = r.narr;
            A.steps[r.ray] = r.nst;
            return false;
        }
    }
    return true;
}

}  // namespace ga
