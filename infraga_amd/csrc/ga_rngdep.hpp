// ga_rngdep.hpp -- device evaluation of the range-dependent atmosphere
// (3D hybrid spline, src/util/interpolation.cpp:886-1793) and of the
// topography (2D multi-stage spline, interpolation.cpp:262-680).
//
// Hybrid spline.  Per field (c, u, v, rho) and node (i, j, k) the table
// holds 5 doubles: f, the z-spline slope of f, the centred x difference of f
// (the node value finite_diff / eval_node_dfdx work from), the z-spline
// slope of that difference, and the centred y difference.  The reference's
// eval_node_dfdy runs its z-spline with the x-difference slopes
// (interpolation.cpp:1119-1120); the table keeps that pairing.  A query
// evaluates the five node quantities (f, f_z, f_zz, x-difference,
// y-difference) on the 4 x 4 node columns around its cell, forms the five
// bicubic patches of eval_all (corner values, centred differences scaled by
// the cell size, 16 x 16 conversion matrix) and sums their polynomials.
//
// Topography.  interp::eval_f on a natural_cubic_spline_2D interpolates the
// y-splines of all x rows at y and solves a natural spline along x through
// those values on every call.  Natural-spline slopes are linear in the data,
// so the x slopes at y equal the Hermite (in y) interpolation of the x
// slopes of the node columns of f and of the y slopes.  The table keeps per
// node f, f_y slope, the x slope of column j's f values and the x slope of
// column j's f_y slopes; a query is a tensor Hermite patch, O(1) instead of
// O(nx) per call, and the same function up to rounding.
#pragma once

#include "ga_device.hpp"

namespace ga {

constexpr int RD_Q = 5;    // doubles per hybrid-spline node
constexpr int TP_Q = 4;    // doubles per topography node

// ix, iy, iz: inverse node spacing of an evenly spaced axis (the segment is
// then found in O(1)), 0 for a binary search.
struct RdGrid {
    const double *x, *y, *z;   // node coordinates (lat, lon [rad]; z [km])
    int nx, ny, nz;
    const double *tab;         // [field][nx][ny][nz][RD_Q]
    double ix, iy, iz;
};

struct TopoGrid {
    const double *x, *y;       // lat, lon [rad]
    int nx, ny;
    const double *tab;         // [nx][ny][TP_Q]
    double ix, iy;
};

GA_HD int clampi(int i, int lo, int hi) { return i < lo ? lo : (i > hi ? hi : i); }

// Segment k with v[k] <= q <= v[k+1] (q already inside [v[0], v[n-1]]).
GA_HD int seg_of(const double *v, int n, double q)
{
    int lo = 0, hi = n - 2;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (v[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The same segment as seg_of: an estimate from the mean spacing, then
// exact steps to the largest k with v[k] <= q (q inside [v[0], v[n-1]]).
GA_HD int seg_at(const double *v, int n, double q, double iv)
{
    if (!(iv > 0.0)) return seg_of(v, n, q);
    int k = (int)((q - v[0]) * iv);
    k = k < 0 ? 0 : (k > n - 2 ? n - 2 : k);
    while (k < n - 2 && v[k + 1] <= q) ++k;
    while (k > 0 && v[k] > q) --k;
    return k;
}

GA_HD double hv(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return f0 + X * (df + (1.0 - X) * (A * (1.0 - X) + B * X));
}
GA_HD double hd1(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return (df + (1.0 - 2.0 * X) * (A * (1.0 - X) + B * X) + X * (1.0 - X) * (B - A)) / d;
}
GA_HD double hd2(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return 2.0 * (B - 2.0 * A + (A - B) * 3.0 * X) / (d * d);
}
GA_HD double hd3(double f0, double f1, double s0, double s1, double d, double)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return 6.0 * (A - B) / (d * d * d);
}
GA_HD double hder(int m, double f0, double f1, double s0, double s1, double d, double X)
{
    return m == 0 ? hv(f0, f1, s0, s1, d, X) : m == 1 ? hd1(f0, f1, s0, s1, d, X)
                  : m == 2 ? hd2(f0, f1, s0, s1, d, X) : hd3(f0, f1, s0, s1, d, X);
}

// Bicubic conversion (interpolation.cpp:686-703) and polynomial sums.
// which: 0 value, 1 d/dx, 2 d/dy, 3 d2/dxdy.
GA_HD void bicubic_coeffs(const double X[16], double A[16])
{
    A[0] = X[0];
    A[1] = X[4];
    A[2] = -3.0 * X[0] + 3.0 * X[1] - 2.0 * X[4] - X[5];
    A[3] = 2.0 * X[0] - 2.0 * X[1] + X[4] + X[5];
    A[4] = X[8];
    A[5] = X[12];
    A[6] = -3.0 * X[8] + 3.0 * X[9] - 2.0 * X[12] - X[13];
    A[7] = 2.0 * X[8] - 2.0 * X[9] + X[12] + X[13];
    A[8] = -3.0 * X[0] + 3.0 * X[2] - 2.0 * X[8] - X[10];
    A[9] = -3.0 * X[4] + 3.0 * X[6] - 2.0 * X[12] - X[14];
    A[10] = 9.0 * X[0] - 9.0 * X[1] - 9.0 * X[2] + 9.0 * X[3] + 6.0 * X[4] + 3.0 * X[5] - 6.0 * X[6] - 3.0 * X[7] +
            6.0 * X[8] - 6.0 * X[9] + 3.0 * X[10] - 3.0 * X[11] + 4.0 * X[12] + 2.0 * X[13] + 2.0 * X[14] + X[15];
    A[11] = -6.0 * X[0] + 6.0 * X[1] + 6.0 * X[2] - 6.0 * X[3] - 3.0 * X[4] - 3.0 * X[5] + 3.0 * X[6] + 3.0 * X[7] -
            4.0 * X[8] + 4.0 * X[9] - 2.0 * X[10] + 2.0 * X[11] - 2.0 * X[12] - 2.0 * X[13] - X[14] - X[15];
    A[12] = 2.0 * X[0] - 2.0 * X[2] + X[8] + X[10];
    A[13] = 2.0 * X[4] - 2.0 * X[6] + X[12] + X[14];
    A[14] = -6.0 * X[0] + 6.0 * X[1] + 6.0 * X[2] - 6.0 * X[3] - 4.0 * X[4] - 2.0 * X[5] + 4.0 * X[6] + 2.0 * X[7] -
            3.0 * X[8] + 3.0 * X[9] - 3.0 * X[10] + 3.0 * X[11] - 2.0 * X[12] - X[13] - 2.0 * X[14] - X[15];
    A[15] = 4.0 * X[0] - 4.0 * X[1] - 4.0 * X[2] + 4.0 * X[3] + 2.0 * X[4] + 2.0 * X[5] - 2.0 * X[6] - 2.0 * X[7] +
            2.0 * X[8] - 2.0 * X[9] + 2.0 * X[10] - 2.0 * X[11] + X[12] + X[13] + X[14] + X[15];
}

GA_HD double bicubic_poly(const double A[16], double xs, double ys, double dx, double dy, int which)
{
    double px[4] = {1.0, xs, xs * xs, xs * xs * xs};
    double py[4] = {1.0, ys, ys * ys, ys * ys * ys};
    double r = 0.0;
    if (which == 0) {
        for (int k1 = 0; k1 < 4; ++k1)
            for (int k2 = 0; k2 < 4; ++k2) r += A[k2 * 4 + k1] * px[k1] * py[k2];
    } else if (which == 1) {
        for (int k1 = 1; k1 < 4; ++k1)
            for (int k2 = 0; k2 < 4; ++k2) r += A[k2 * 4 + k1] * px[k1 - 1] * py[k2] * (double)k1;
        r /= dx;
    } else if (which == 2) {
        for (int k1 = 0; k1 < 4; ++k1)
            for (int k2 = 1; k2 < 4; ++k2) r += A[k2 * 4 + k1] * px[k1] * py[k2 - 1] * (double)k2;
        r /= dy;
    } else {
        for (int k1 = 1; k1 < 4; ++k1)
            for (int k2 = 1; k2 < 4; ++k2) r += A[k2 * 4 + k1] * px[k1 - 1] * py[k2 - 1] * (double)(k1 * k2);
        r /= dx * dy;
    }
    return r;
}

// Cell of a query point, shared by the fields of one grid: segments, the
// clamped and scaled coordinates and the 4 x 4 node columns around the cell
// (indices clamped at the grid edge, as the reference's finite differences
// clamp their neighbours).
struct RdCell {
    int kz;
    int ci[4], cj[4];
    double dx, dy, dz, xs, ys, Xz, Xz_u;
};

GA_HD RdCell rd_cell(const RdGrid &g, double x, double y, double z)
{
    RdCell c;
    const double xe = fmin(fmax(x, g.x[0]), g.x[g.nx - 1]);
    const double ye = fmin(fmax(y, g.y[0]), g.y[g.ny - 1]);
    const double ze = fmin(fmax(z, g.z[0]), g.z[g.nz - 1]);
    const int kx = seg_at(g.x, g.nx, xe, g.ix), ky = seg_at(g.y, g.ny, ye, g.iy);
    c.kz = seg_at(g.z, g.nz, ze, g.iz);
    c.dx = g.x[kx + 1] - g.x[kx];
    c.xs = (xe - g.x[kx]) / c.dx;
    c.dy = g.y[ky + 1] - g.y[ky];
    c.ys = (ye - g.y[ky]) / c.dy;
    c.dz = g.z[c.kz + 1] - g.z[c.kz];
    c.Xz = (ze - g.z[c.kz]) / c.dz;
    c.Xz_u = (z - g.z[c.kz]) / c.dz;   // f_zz patch: unclamped z (interpolation.cpp:1763-1771)
    for (int a = 0; a < 4; ++a) {
        c.ci[a] = clampi(kx - 1 + a, 0, g.nx - 1);
        c.cj[a] = clampi(ky - 1 + a, 0, g.ny - 1);
    }
    return c;
}

// One field of the hybrid spline at a cell (interp::eval_f / eval_all,
// interpolation.cpp:1334-1369, 1544-1788).  MODE 0: f; 1: also
// df = {fx, fy, fz}; 2: also ddf = {fxx, fyy, fzz, fxy, fxz, fyz}.
template <int MODE>
GA_HD void rd_field_eval(const RdGrid &g, const RdCell &cl, int field, double &f, double *df, double *ddf)
{
    constexpr int NQ = MODE == 0 ? 1 : (MODE == 1 ? 2 : 5);
    // node quantities on the 4 x 4 columns: 0 f, 1 f_z, 2 f_zz, 3 x-diff, 4 y-diff
    double Q[NQ][4][4];
    const double *base = g.tab + (size_t)field * g.nx * g.ny * g.nz * RD_Q;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            const double *p = base + (((size_t)cl.ci[a] * g.ny + cl.cj[b]) * g.nz + cl.kz) * RD_Q;
            const double *q = p + RD_Q;
            Q[0][a][b] = hv(p[0], q[0], p[1], q[1], cl.dz, cl.Xz);
            if constexpr (NQ > 1) Q[1][a][b] = hd1(p[0], q[0], p[1], q[1], cl.dz, cl.Xz);
            if constexpr (NQ > 2) {
                Q[2][a][b] = hd2(p[0], q[0], p[1], q[1], cl.dz, cl.Xz_u);
                Q[3][a][b] = hv(p[2], q[2], p[3], q[3], cl.dz, cl.Xz);
                Q[4][a][b] = hv(p[4], q[4], p[3], q[3], cl.dz, cl.Xz);
            }
        }
    // corner (a, b) in {1, 2}^2; differences over the clamped neighbours
    auto ddx = [&](int q, int a, int b) {
        return (Q[q][a + 1][b] - Q[q][a - 1][b]) / (g.x[cl.ci[a + 1]] - g.x[cl.ci[a - 1]]);
    };
    auto ddy = [&](int q, int a, int b) {
        return (Q[q][a][b + 1] - Q[q][a][b - 1]) / (g.y[cl.cj[b + 1]] - g.y[cl.cj[b - 1]]);
    };
    auto ddxy = [&](int q, int a, int b) {
        return (Q[q][a + 1][b + 1] - Q[q][a + 1][b - 1] - Q[q][a - 1][b + 1] + Q[q][a - 1][b - 1]) /
               ((g.x[cl.ci[a + 1]] - g.x[cl.ci[a - 1]]) * (g.y[cl.cj[b + 1]] - g.y[cl.cj[b - 1]]));
    };
    // corner order of the reference's X_vec: (kx,ky), (kx+1,ky), (kx,ky+1), (kx+1,ky+1)
    const int A_[4] = {1, 2, 1, 2}, B_[4] = {1, 1, 2, 2};
    const double dx = cl.dx, dy = cl.dy, xs = cl.xs, ys = cl.ys;
    double X[16], C[16];
    // patch of quantity qv with x slopes from (qx, kind) etc.; kinds: 0 ddx, 1 ddy, 2 ddxy
    auto build = [&](int qv, int qxs, int kxs, int qys, int kys, int qxys) {
        for (int c = 0; c < 4; ++c) {
            const int a = A_[c], b = B_[c];
            X[c] = Q[qv][a][b];
            X[4 + c] = (kxs == 0 ? ddx(qxs, a, b) : ddxy(qxs, a, b)) * dx;
            X[8 + c] = (kys == 1 ? ddy(qys, a, b) : ddxy(qys, a, b)) * dy;
            X[12 + c] = ddxy(qxys, a, b) * dx * dy;
        }
        bicubic_coeffs(X, C);
    };
    build(0, 0, 0, 0, 1, 0);                         // f
    f = bicubic_poly(C, xs, ys, dx, dy, 0);
    if constexpr (MODE >= 1) {
        df[0] = bicubic_poly(C, xs, ys, dx, dy, 1);
        df[1] = bicubic_poly(C, xs, ys, dx, dy, 2);
    }
    if constexpr (MODE == 2) ddf[3] = bicubic_poly(C, xs, ys, dx, dy, 3);
    if constexpr (MODE >= 1) {
        build(1, 1, 0, 1, 1, 1);                     // f_z
        df[2] = bicubic_poly(C, xs, ys, dx, dy, 0);
    }
    if constexpr (MODE == 2) {
        ddf[4] = bicubic_poly(C, xs, ys, dx, dy, 1);
        ddf[5] = bicubic_poly(C, xs, ys, dx, dy, 2);
        build(3, 3, 0, 0, 2, 3);                     // x-difference: slopes ddx(Qx), ddxy(f), ddxy(Qx)
        ddf[0] = bicubic_poly(C, xs, ys, dx, dy, 1);
        build(4, 0, 2, 4, 1, 4);                     // y-difference: slopes ddxy(f), ddy(Qy), ddxy(Qy)
        ddf[1] = bicubic_poly(C, xs, ys, dx, dy, 2);
        build(2, 2, 0, 2, 1, 2);                     // f_zz
        ddf[2] = bicubic_poly(C, xs, ys, dx, dy, 0);
    }
}

// One field of the hybrid spline: f, df = {fx, fy, fz}, ddf = {fxx, fyy, fzz,
// fxy, fxz, fyz} at (x, y, z) (interp::eval_all, interpolation.cpp:1618-1788).
GA_HD void rd_field_all(const RdGrid &g, int field, double x, double y, double z, double &f, double df[3],
                        double ddf[6])
{
    const RdCell c = rd_cell(g, x, y, z);
    rd_field_eval<2>(g, c, field, f, df, ddf);
}

// Topography z_g(x, y) and its derivatives: out = {z, zx, zy, zxx, zxy, zyy,
// zxxx, zxxy, zxyy, zyyy} (topo::z, dz, ddz, dddz with clamped arguments).
GA_HD void topo_all(const TopoGrid &g, double x, double y, double out[10])
{
    const double xe = fmin(fmax(x, g.x[0]), g.x[g.nx - 1]);
    const double ye = fmin(fmax(y, g.y[0]), g.y[g.ny - 1]);
    const int kx = seg_at(g.x, g.nx, xe, g.ix), ky = seg_at(g.y, g.ny, ye, g.iy);
    const double dx = g.x[kx + 1] - g.x[kx], X = (xe - g.x[kx]) / dx;
    const double dy = g.y[ky + 1] - g.y[ky], Y = (ye - g.y[ky]) / dy;
    const double *n00 = g.tab + ((size_t)kx * g.ny + ky) * TP_Q, *n01 = n00 + TP_Q;
    const double *n10 = n00 + (size_t)g.ny * TP_Q, *n11 = n10 + TP_Q;
    // y-derivative m of the row values and of the row x slopes at rows kx, kx+1
    double gv[2][4], gs[2][4];
    for (int m = 0; m < 4; ++m) {
        gv[0][m] = hder(m, n00[0], n01[0], n00[1], n01[1], dy, Y);
        gv[1][m] = hder(m, n10[0], n11[0], n10[1], n11[1], dy, Y);
        gs[0][m] = hder(m, n00[2], n01[2], n00[3], n01[3], dy, Y);
        gs[1][m] = hder(m, n10[2], n11[2], n10[3], n11[3], dy, Y);
    }
    auto d = [&](int q, int m) { return hder(q, gv[0][m], gv[1][m], gs[0][m], gs[1][m], dx, X); };
    out[0] = d(0, 0);
    out[1] = d(1, 0);
    out[2] = d(0, 1);
    out[3] = d(2, 0);
    out[4] = d(1, 1);
    out[5] = d(0, 2);
    out[6] = d(3, 0);
    out[7] = d(2, 1);
    out[8] = d(1, 2);
    out[9] = d(0, 3);
}

}  // namespace ga
