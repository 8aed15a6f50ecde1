// ga_2d.hpp -- the 2D effective-sound-speed geometry (infraga-2d -prop) on
// the GPU: one ray per lane, the stratified spline table staged in LDS, the
// write_rays bookkeeping of the 2D driver (its path rows are always written).
//
// Equations (geoac.eqset.2d.cpp:176-239): r, z, nu_r, nu_z and, with
// amplitudes, R, Z, mu_r, mu_z, driven by c_eff = c + u cos(phi) + v sin(phi)
// (no boundary-layer wind scaling in 2D, atmo_state.2d.cpp:75-160).
#pragma once

#include "ga_trace.hpp"

namespace ga {

// Terrain profile z(r): natural cubic spline nodes [n][3] = (r, z, slope).
struct Topo1D {
    const double *tab;
    int n;                      // 0: flat ground at M.z0
};

struct Fan2DArgs {
    Medium M;                   // stratified table; z0, z_max, z_bndlyr of the 2D region
    Topo1D T;
    double s_max, ds_min, ds_max;
    double alt_max, rng_max, z_reflect;
    double z_src;               // source altitude (raised to the ground)
    double c_src;               // atmo::c at the source (arrival inclination)
    double rho_src;             // atmo::rho at the source (amplitude)
    double freq;
    int bounces, n_rays, max_steps;
    const double *launches;     // (incl_deg, az_deg)
    ga_arrival *out;            // (bounces + 1) slots per ray
    int32_t *counts, *steps;
    unsigned int *queue;
    PathOut path;               // flags: GA_PATH_RAYS | GA_PATH_CAUSTICS
};

// topo::z / dz / ddz (atmo_state.2d.cpp:47-69): the terrain spline at r
// clamped into its node range.
GA_HD void topo1d(const Fan2DArgs &A, double r, int &cur, double &z, double &dz, double &ddz)
{
    if (A.T.n == 0) {
        z = A.M.z0;
        dz = 0.0;
        ddz = 0.0;
        return;
    }
    const double *t = A.T.tab;
    const int n = A.T.n;
    r = fmin(fmax(r, t[0]), t[(n - 1) * 3]);
    while (cur > 0 && r < t[cur * 3]) --cur;
    while (cur < n - 2 && r > t[(cur + 1) * 3]) ++cur;
    const double *a = t + cur * 3;
    const double dx = a[3] - a[0];
    const double X = (r - a[0]) / dx, omX = 1.0 - X;
    const double d = a[4] - a[1];
    const double Aa = a[2] * dx - d, B = -a[5] * dx + d;
    const double q = Aa * omX + B * X;
    z = a[1] + X * (d + omX * q);
    dz = (d + (1.0 - 2.0 * X) * q + X * omX * (B - Aa)) / dx;
    ddz = 2.0 * (B - 2.0 * Aa + (Aa - B) * 3.0 * X) / (dx * dx);
}

GA_HD double topo1d_z(const Fan2DArgs &A, double r, int &cur)
{
    double z, dz, ddz;
    topo1d(A, r, cur, z, dz, ddz);
    return z;
}

// Point values atmo::c / u / v / rho at height z (clamped into the table).
GA_HD void pt2d(const Medium &M, double z, int &cur, double &c, double &u, double &v, double &rho)
{
    SegPt s = seg_point(M, clampz(M, z), cur);
    c = herm_f(s, 1);
    u = herm_f(s, 3);
    v = herm_f(s, 5);
    rho = herm_f(s, 7);
}

// Right-hand side (update_refs + eval_src_eq, geoac.eqset.2d.cpp:176-239).
template <int NEQ>
GA_HD void rhs2d(const Medium &M, const double *y, double cph, double sph, double c0, int &cur, double *f)
{
    SegPt s = seg_point(M, clampz(M, y[1]), cur);
    double c, dc, ddc, u, du, ddu, v, dv, ddv;
    herm_fdd(s, 1, c, dc, ddc);
    herm_fdd(s, 3, u, du, ddu);
    herm_fdd(s, 5, v, dv, ddv);
    const double ce = c + u * cph + v * sph;
    const double dce = dc + du * cph + dv * sph;
    const double ic = 1.0 / ce;
    const double cc0 = ce / c0;
    const double k2 = c0 * (ic * ic);
    f[0] = cc0 * y[2];
    f[1] = cc0 * y[3];
    f[2] = 0.0;
    f[3] = -k2 * dce;
    if (NEQ == 8) {
        const double ddce = ddc + ddu * cph + ddv * sph;
        const double dcth = dce * y[5];
        f[4] = dcth / c0 * y[2] + cc0 * y[6];
        f[5] = dcth / c0 * y[3] + cc0 * y[7];
        f[6] = 0.0;
        f[7] = 2.0 * k2 * ic * dce * dcth - k2 * (ddce * y[5]);
    }
}

// set_ds (:165-171).
GA_HD double ds2d(const Fan2DArgs &A, const double *y, int &tcur)
{
    double r = A.ds_max;
    if (y[1] < A.M.z_bndlyr) r -= (A.ds_max - A.ds_min) * exp(-(y[1] - topo1d_z(A, y[0], tcur)) / 0.2);
    return r;
}

// Classical RK4 step (rk4solver.cpp:26-55) in the reference's stage order.
template <int NEQ>
GA_HD double rk4_2d(const Fan2DArgs &A, const double *y, double *y1, double cph, double sph, double c0, int &cur,
                    int &tcur)
{
    const double ds = ds2d(A, y, tcur);
    double f[NEQ], t[NEQ], acc[NEQ], p[NEQ];
    rhs2d<NEQ>(A.M, y, cph, sph, c0, cur, f);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
        t[i] = ds * f[i];
        acc[i] = y[i] + t[i] / 6.0;
        p[i] = y[i] + t[i] / 2.0;
    }
    rhs2d<NEQ>(A.M, p, cph, sph, c0, cur, f);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
        t[i] = ds * f[i];
        acc[i] = acc[i] + t[i] / 3.0;
        p[i] = y[i] + t[i] / 2.0;
    }
    rhs2d<NEQ>(A.M, p, cph, sph, c0, cur, f);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) {
        t[i] = ds * f[i];
        acc[i] = acc[i] + t[i] / 3.0;
        p[i] = y[i] + t[i];
    }
    rhs2d<NEQ>(A.M, p, cph, sph, c0, cur, f);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) y1[i] = acc[i] + (ds * f[i]) / 6.0;
    return ds;
}

// 20 log10 of the amplitude at row y (jacobian / amp, :345-361): atmo::c
// (not c_eff) in dr/ds and dz/ds, as the reference writes it.
GA_HD double amp2d_db(const Fan2DArgs &A, const double *y, double th, double c0, int &cur)
{
    double c, u, v, rho;
    pt2d(A.M, y[1], cur, c, u, v, rho);
    const double drds = c / c0 * y[2], dzds = c / c0 * y[3];
    const double jac = y[0] * (drds * y[5] - dzds * y[4]);
    const double num = rho * c * cos(th);
    const double den = A.rho_src * c0 * jac;
    return 20.0 * log10(sqrt(fabs(num / den)));
}

GA_HD double jac2d(const Fan2DArgs &A, const double *y, double c0, int &cur)
{
    double c, u, v, rho;
    pt2d(A.M, y[1], cur, c, u, v, rho);
    return y[0] * (c / c0 * y[2] * y[5] - c / c0 * y[3] * y[4]);
}

// approx_intercept + set_refl (:73-160): ground (flat or terrain) below the
// boundary-layer top, otherwise the elevated reflection altitude.
template <int NEQ>
GA_HD void refl2d(const Fan2DArgs &A, const double *yk2, const double *yk1, double *y, double cph, double sph,
                  double c0, int &cur, int &tcur)
{
    const Medium &M = A.M;
    double prev[NEQ];
    if (yk1[1] < M.z_bndlyr) {
        double tz, tdz, tddz;
        topo1d(A, yk1[0], tcur, tz, tdz, tddz);
        const double dz_grnd = tz - yk1[1];
        const double dz_k = (yk1[1] - yk2[1]) - tdz * (yk1[0] - yk2[0]);
#pragma unroll
        for (int i = 0; i < NEQ; ++i) prev[i] = yk1[i] + (yk1[i] - yk2[i]) / dz_k * dz_grnd;
    } else {
        const double dz_grnd = A.z_reflect - yk1[1];
        const double dz_k = yk1[1] - yk2[1];
#pragma unroll
        for (int i = 0; i < NEQ; ++i) prev[i] = yk1[i] + (yk1[i] - yk2[i]) / dz_k * dz_grnd;
    }
    double zg, C1 = 1.0, C2 = 0.0, dC1 = 0.0, dC2 = 0.0, ta = 0.0;
    const bool grnd = prev[1] < M.z_bndlyr;
    double tddz = 0.0;
    if (grnd) {
        double tz, tdz;
        topo1d(A, prev[0], tcur, tz, tdz, tddz);
        zg = tz;
        ta = tdz;
    } else {
        zg = A.z_reflect;
    }
    SegPt s = seg_point(M, clampz(M, zg), cur);
    double c, dc, u, du, v, dv;
    herm_fd(s, 1, c, dc);
    herm_fd(s, 3, u, du);
    herm_fd(s, 5, v, dv);
    const double ce = c + u * cph + v * sph;
    const double dce = dc + du * cph + dv * sph;
    double ds0 = 0.0;
    if (grnd && A.T.n > 0) {
        const double norm = 1.0 + ta * ta;
        C1 = (1.0 - ta * ta) / norm;
        C2 = 2.0 * ta / norm;
        if (NEQ == 8) {
            const double da = (prev[4] * prev[3] - prev[5] * prev[2]) / (prev[3] - ta * prev[2]) * tddz;
            dC1 = -2.0 * da / norm * C2;
            dC2 = 2.0 * da / norm * C1;
            ds0 = -c0 / ce * (prev[5] - ta * prev[4]) / (prev[3] - ta * prev[2]);
        }
    } else if (NEQ == 8) {
        ds0 = -c0 / ce * prev[5] / prev[3];
    }
    y[0] = prev[0];
    y[1] = zg;
    y[2] = C1 * prev[2] + C2 * prev[3];
    y[3] = -C1 * prev[3] + C2 * prev[2];
    if (NEQ == 8) {
        const double dnur = 0.0, dnuz = -c0 / (ce * ce) * dce;
        y[4] = C1 * prev[4] + C2 * prev[5];
        y[5] = -C1 * prev[5] + C2 * prev[4];
        y[6] = C1 * prev[6] + C2 * prev[7] + dC1 * prev[2] + dC2 * prev[3] + ((C1 - 1.0) * dnur + C2 * dnuz) * ds0;
        y[7] = -C1 * prev[7] + C2 * prev[6] - dC1 * prev[3] + dC2 * prev[2] + (-(C1 + 1.0) * dnuz + C2 * dnur) * ds0;
    }
}

// Per-lane 2D fan driver (main.infraga.2d.cpp:318-395).
// The two waveform-row records of row y (PATH_WVROWS; ga_trace.hpp) in
// 2D: the state (r, z, nu_r, nu_z), then the Jacobian determinant, the
// effective sound speed c + u cos(phi) + v sin(phi) (also in the group-speed
// slot), rho and the travel time dt of the segment into the row.
GA_HD void wv2d_rows_put(const Fan2DArgs &A, PathLane &pl, const double *y, double cph, double sph, double c0,
                         int item, int leg, int m, double dt, int &cur)
{
    double c, u, v, rho;
    pt2d(A.M, y[1], cur, c, u, v, rho);
    const double ce = c + u * cph + v * sph;
    ga_path_pt a, b;
    a.launch_idx = b.launch_idx = item;
    a.leg = b.leg = leg;
    a.step = b.step = m;
    a.kind = PT_WV_STATE;
    a.pos0 = y[0]; a.pos1 = y[1]; a.z = 0.0;
    a.tl = y[2]; a.absorption = y[3]; a.time = 0.0;
    b.kind = PT_WV_MEDIUM;
    b.pos0 = jac2d(A, y, c0, cur);
    b.pos1 = ce;
    b.z = rho;
    b.tl = ce;
    b.absorption = dt;
    b.time = 0.0;
    path_put(A.path, pl, item, a);
    path_put(A.path, pl, item, b);
}

template <int NEQ, class NextRay>
GA_HD void lane_loop_2d(const Fan2DArgs &A, const Medium &M, NextRay next_ray)
{
    double y[NEQ], yp[NEQ];
    double s = 0.0, tt = 0.0, att = 0.0, dev = 0.0, zmax = 0.0, d_prev = 0.0;
    double th = 0.0, cph = 1.0, sph = 0.0, c0 = 1.0, incl = 0.0, az = 0.0;
    int k = 0, bnc = 0, nst = 0, ray = 0, cur = 0, tcur = 0, narr = 0;
    bool fetch = true;
    PathLane pl;
    path_reset(pl);
    const bool caus = NEQ == 8 && (A.path.flags & GA_PATH_CAUSTICS);
    const bool rays = (A.path.flags & GA_PATH_RAYS) != 0;
    // waveform rows: every row of every leg, and (run_wnl_wvfrm's loop,
    // main.infraga.2d.cpp) every leg counts toward bounces, elevated
    // reflections included
    const bool wv = NEQ == 8 && (A.path.flags & PATH_WVROWS);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) { y[i] = 0.0; yp[i] = 0.0; }

    for (;;) {
        if (fetch) {
            ray = next_ray();
            if (ray >= A.n_rays) break;
            incl = A.launches[2 * ray];
            az = A.launches[2 * ray + 1];
            th = incl * PI / 180.0;
            const double ph = PI / 2.0 - az * PI / 180.0;
            cph = cos(ph);
            sph = sin(ph);
            // set_initial (:46-68)
            double c, u, v, rho;
            pt2d(M, A.z_src, cur, c, u, v, rho);
            c0 = c + u * cph + v * sph;
            const double st = sin(th), ct = cos(th);
            y[0] = 0.0; y[1] = A.z_src; y[2] = ct; y[3] = st;
            if (NEQ == 8) { y[4] = 0.0; y[5] = 0.0; y[6] = -st; y[7] = ct; }
            s = 0.0; tt = 0.0; att = 0.0; dev = 0.0; zmax = 0.0;
            k = 0; bnc = 0; nst = 0; narr = 0;
            fetch = false;
            if (wv) wv2d_rows_put(A, pl, y, cph, sph, c0, ray, 0, 0, 0.0, cur);
        }
        double y1[NEQ];
        s += rk4_2d<NEQ>(A, y, y1, cph, sph, c0, cur, tcur);
        ++k;
        ++nst;
        bool brk = false, end = false;
        if (s > A.s_max) { brk = true; end = true; }
        else if (y1[0] > A.rng_max || y1[1] > A.alt_max) { brk = true; end = true; }
        else if (y1[1] <= M.z_max && y1[1] < topo1d_z(A, y1[0], tcur)) { end = true; }
        else if (y1[1] > A.z_reflect) { end = true; }
        else if (k >= A.max_steps) { brk = true; end = true; }
        else if (!(y1[0] == y1[0])) { brk = true; end = true; }

        // segment y -> y1: perpendicular drift of the leg (all its segments,
        // est_dev :366-380); travel time and absorption only for rows 1..k-1
        {
            const double ds = sqrt((y1[0] - y[0]) * (y1[0] - y[0]) + (y1[1] - y[1]) * (y1[1] - y[1]));
            double c, u, v, rho;
            pt2d(M, (y1[1] + y[1]) / 2.0, cur, c, u, v, rho);
            const double ce = c + u * cph + v * sph;
            const double vp = -u * sph + v * cph;
            dev += vp / ce * 1.0 / sqrt(1.0 + 2.0 * vp / ce + (u * u + v * v) / (ce * ce)) * ds;
            if (!end || wv) {
                const double zm = y[1] + (y1[1] - y[1]) / 2.0;
                pt2d(M, zm, cur, c, u, v, rho);
                const double dt = ds / (c + u * cph + v * sph);
                if (!end) {
                    tt += dt;
                    att += sb_alpha<false>(zm, c, rho, A.freq, M.tweak, M.sbc) * ds;
                }
                if (wv) wv2d_rows_put(A, pl, y1, cph, sph, c0, ray, bnc, k, dt, cur);
            }
        }
        zmax = fmax(zmax, y[1]);
        if (!end) {
            const bool samp = rays && (k == 1 || k % 15 == 0);
            if (samp) {
                const double tl = NEQ == 8 ? amp2d_db(A, y1, th, c0, cur) : 0.0;
                ga_path_pt P;
                P.launch_idx = ray; P.leg = bnc; P.step = k; P.kind = GA_PT_RAY;
                P.pos0 = y1[0]; P.pos1 = 0.0;
                P.z = fmax(y1[1], topo1d_z(A, y1[0], tcur));
                P.tl = tl; P.absorption = -2.0 * att; P.time = tt;
                path_put(A.path, pl, ray, P);
            }
            if (caus) {
                const double d = jac2d(A, y1, c0, cur);
                if (k == 1) d_prev = d;
                if (d * d_prev < 0.0) {
                    ga_path_pt P;
                    P.launch_idx = ray; P.leg = bnc; P.step = k; P.kind = GA_PT_CAUSTIC;
                    P.pos0 = y1[0]; P.pos1 = 0.0; P.z = y1[1];
                    P.tl = 0.0; P.absorption = -2.0 * att; P.time = tt;
                    path_put(A.path, pl, ray, P);
                }
                d_prev = d;
            }
#pragma unroll
            for (int i = 0; i < NEQ; ++i) { yp[i] = y[i]; y[i] = y1[i]; }
            continue;
        }
        // leg ended at row k = y1; rows k-1 = y, k-2 = yp
        bool done = false;
        if (brk || k < 2) done = true;
        else {
            if (!(y[1] > M.z_bndlyr)) {
                // ground arrival (:373-386)
                ga_arrival R;
                R.launch_idx = ray;
                R.n_bounce = bnc;
                R.incl = incl;
                R.az = az;
                R.pos0 = y1[0];
                R.pos1 = dev;
                R.time = tt;
                R.cel = y1[0] / tt;
                R.turn_ht = zmax;
                double c, u, v, rho;
                pt2d(M, topo1d_z(A, y1[0], tcur), cur, c, u, v, rho);
                R.incl_arr = -asin(c / A.c_src * y1[3]) * (180.0 / PI);
                R.back_az = 0.0;
                R.tl = NEQ == 8 ? amp2d_db(A, y1, th, c0, cur) : 0.0;
                R.absorption = -2.0 * att;
                A.out[(size_t)ray * (A.bounces + 1) + bnc] = R;
                ++narr;
                if (bnc == A.bounces) done = true;
                else ++bnc;
            } else if (wv) {
                if (bnc == A.bounces) done = true;
                else ++bnc;
            }
            if (!done) {
                double yr[NEQ];
                refl2d<NEQ>(A, yp, y, yr, cph, sph, c0, cur, tcur);
#pragma unroll
                for (int i = 0; i < NEQ; ++i) y[i] = yr[i];
                s = 0.0;
                k = 0;
                dev = 0.0;
                if (wv) wv2d_rows_put(A, pl, y, cph, sph, c0, ray, bnc, 0, 0.0, cur);
            }
        }
        if (done) {
            A.counts[ray] = narr;
            A.steps[ray] = nst;
            path_close(A.path, pl, ray);
            fetch = true;
        }
    }
}

}  // namespace ga
