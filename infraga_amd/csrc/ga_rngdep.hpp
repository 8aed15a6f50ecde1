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
// y-difference) on the 4 x 4 node columns around its cell and sums them with
// the weights of the five bicubic patches of eval_all (corner values and
// centred differences scaled by the cell size); the patches are the
// reference's functions, evaluated in another order of operations.
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

// lo, hi: first and last node of each axis (x, y, z); iv: inverse node
// spacing of an evenly spaced axis (its segments are then found in O(1)),
// 0 for a binary search.
struct RdGrid {
    const double *x, *y, *z;   // node coordinates (lat, lon [rad]; z [km])
    int nx, ny, nz;
    const double *tab;         // [field][nx][ny][nz][RD_Q]
    double lo[3], hi[3], iv[3];
};

struct TopoGrid {
    const double *x, *y;       // lat, lon [rad]
    int nx, ny;
    const double *tab;         // [nx][ny][TP_Q]
    double lo[2], hi[2], iv[2];
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

// Segment of q along an axis and the 4 node coordinates around it,
// v[a] = axis[clamp(k - 1 + a)]: the same segment as seg_of.  On an evenly
// spaced axis (iv = 1 / spacing > 0, v0 = first node) the segment is
// estimated from the spacing; axis_guess loads the four nodes of the
// estimate (the loads of all axes of a query go out together), axis_ok
// checks it and axis_fix steps an estimate that rounding put one node off.
struct AxisSeg {
    int k;
    double v[4];
};

GA_HD AxisSeg axis_guess(const double *ax, int n, double q, double iv, double v0)
{
    int k;
    if (iv > 0.0) {
        k = (int)((q - v0) * iv);
        k = k < 0 ? 0 : (k > n - 2 ? n - 2 : k);
    } else {
        k = seg_of(ax, n, q);
    }
    AxisSeg s;
    s.k = k;
    for (int a = 0; a < 4; ++a) s.v[a] = ax[clampi(k - 1 + a, 0, n - 1)];
    return s;
}

GA_HD bool axis_ok(const AxisSeg &s, int n, double q)
{
    return !(s.k > 0 && s.v[1] > q) && !(s.k < n - 2 && s.v[2] <= q);
}

GA_HD void axis_fix(AxisSeg &s, const double *ax, int n, double q)
{
    while (!axis_ok(s, n, q)) {
        s.k += (s.k > 0 && s.v[1] > q) ? -1 : 1;
        for (int a = 0; a < 4; ++a) s.v[a] = ax[clampi(s.k - 1 + a, 0, n - 1)];
    }
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

// Cubic Hermite basis on [0, 1]: h0 weights the end values, h1 the end
// slopes (per unit s), g0 and g1 are their s-derivatives.
GA_HD void herm_basis(double s, double h0[2], double h1[2], double g0[2], double g1[2])
{
    const double s2 = s * s, s3 = s2 * s;
    h0[0] = 1.0 - 3.0 * s2 + 2.0 * s3;
    h0[1] = 3.0 * s2 - 2.0 * s3;
    h1[0] = s - 2.0 * s2 + s3;
    h1[1] = s3 - s2;
    g0[0] = 6.0 * (s2 - s);
    g0[1] = 6.0 * (s - s2);
    g1[0] = 1.0 - 4.0 * s + 3.0 * s2;
    g1[1] = 3.0 * s2 - 2.0 * s;
}

// One horizontal axis of a hybrid-spline patch as weights over the 4 node
// columns a = 0..3 around the cell.  The patch (interpolation.cpp:686-703,
// the bicubic conversion of corner values, slopes times the cell size and
// cross slopes) takes its corner values at a = 1, 2 and its corner slopes as
// the centred differences (Q[a+1] - Q[a-1]) / (x[a+1] - x[a-1])
// (interpolation.cpp:1544-1616), so every patch is a sum of tensor products
// of these weight vectors with the node quantities:
//   V   value weights (corners only)       Vd  their x-derivative
//   S   slope weights (through the centred differences, times the cell size)
//   Sd  their x-derivative
//   T   the centred differences of Vd (cross-slope terms)
// P = V + S is the weight of a patch whose slopes all come from its own
// quantity; d/dx of it is Pd = Vd + Sd.
struct AxisW {
    double V[4], S[4], Vd[4], Sd[4], T[4];
};

GA_HD AxisW axis_weights(double s, double d, double e0, double e1)
{
    // e0 = 1 / (x[2] - x[0]), e1 = 1 / (x[3] - x[1]) over the clamped columns
    double h0[2], h1[2], g0[2], g1[2];
    herm_basis(s, h0, h1, g0, g1);
    AxisW w;
    const double id = 1.0 / d;
    w.V[0] = 0.0; w.V[1] = h0[0]; w.V[2] = h0[1]; w.V[3] = 0.0;
    w.Vd[0] = 0.0; w.Vd[1] = g0[0] * id; w.Vd[2] = g0[1] * id; w.Vd[3] = 0.0;
    const double a0 = h1[0] * d * e0, a1 = h1[1] * d * e1;
    w.S[0] = -a0; w.S[1] = -a1; w.S[2] = a0; w.S[3] = a1;
    const double b0 = g1[0] * e0, b1 = g1[1] * e1;
    w.Sd[0] = -b0; w.Sd[1] = -b1; w.Sd[2] = b0; w.Sd[3] = b1;
    const double t0 = w.Vd[1] * e0, t1 = w.Vd[2] * e1;
    w.T[0] = -t0; w.T[1] = -t1; w.T[2] = t0; w.T[3] = t1;
    return w;
}

// Cell of a query point, shared by the fields of one grid: segments, the
// clamped and scaled coordinates, the 4 x 4 node columns around the cell
// (indices clamped at the grid edge, as the reference's finite differences
// clamp their neighbours) and the inverse spans of their centred differences.
struct RdCell {
    int kz;
    int ci[4], cj[4];
    double dx, dy, dz, xs, ys, Xz, Xz_u;
    double ex[2], ey[2];
};

GA_HD RdCell rd_cell(const RdGrid &g, double x, double y, double z)
{
    RdCell c;
    const double xe = fmin(fmax(x, g.lo[0]), g.hi[0]);
    const double ye = fmin(fmax(y, g.lo[1]), g.hi[1]);
    const double ze = fmin(fmax(z, g.lo[2]), g.hi[2]);
    AxisSeg sx = axis_guess(g.x, g.nx, xe, g.iv[0], g.lo[0]);
    AxisSeg sy = axis_guess(g.y, g.ny, ye, g.iv[1], g.lo[1]);
    AxisSeg sz = axis_guess(g.z, g.nz, ze, g.iv[2], g.lo[2]);
    if (!(axis_ok(sx, g.nx, xe) && axis_ok(sy, g.ny, ye) && axis_ok(sz, g.nz, ze))) {
        axis_fix(sx, g.x, g.nx, xe);
        axis_fix(sy, g.y, g.ny, ye);
        axis_fix(sz, g.z, g.nz, ze);
    }
    c.kz = sz.k;
    c.dz = sz.v[2] - sz.v[1];
    c.Xz = (ze - sz.v[1]) / c.dz;
    c.Xz_u = (z - sz.v[1]) / c.dz;   // f_zz patch: unclamped z (interpolation.cpp:1764-1787)
    const double *xv = sx.v, *yv = sy.v;
    for (int a = 0; a < 4; ++a) {
        c.ci[a] = clampi(sx.k - 1 + a, 0, g.nx - 1);
        c.cj[a] = clampi(sy.k - 1 + a, 0, g.ny - 1);
    }
    c.dx = xv[2] - xv[1];
    c.xs = (xe - xv[1]) / c.dx;
    c.dy = yv[2] - yv[1];
    c.ys = (ye - yv[1]) / c.dy;
    c.ex[0] = 1.0 / (xv[2] - xv[0]);
    c.ex[1] = 1.0 / (xv[3] - xv[1]);
    c.ey[0] = 1.0 / (yv[2] - yv[0]);
    c.ey[1] = 1.0 / (yv[3] - yv[1]);
    return c;
}

// One field of the hybrid spline at a cell (interp::eval_f / eval_all,
// interpolation.cpp:1334-1369, 1544-1788).  MODE 0: f; 1: also
// df = {fx, fy, fz}; 2: also ddf = {fxx, fyy, fzz, fxy, fxz, fyz}.
//
// Node quantities per column, from the z-splines: Q0 f, Q1 f_z, Q2 f_zz (at
// the unclamped z), Q3 the x-difference, Q4 the y-difference.  The
// reference's patches, as weight sums (x weights wx, y weights wy):
//   f, fx, fy, fxy    patch of Q0:  Px Q0 Py, Pdx Q0 Py, Px Q0 Pdy, Pdx Q0 Pdy
//   fz, fxz, fyz      patch of Q1:  Px Q1 Py, Pdx Q1 Py, Px Q1 Pdy
//   fzz               patch of Q2:  Px Q2 Py
//   fxx  d/dx of the Q3 patch with y slopes from the cross differences of
//        Q0 (interpolation.cpp:1714-1737): Pdx Q3 Vy + Sdx Q3 Sy + Tx Q0 Sy
//   fyy  d/dy of the Q4 patch with x slopes from the cross differences of
//        Q0 (:1739-1762):                   Vx Q4 Pdy + Sx Q4 Sdy + Sx Q0 Ty
// The columns are streamed (the sums run over the 16 columns as they are
// loaded); Q0 and Q1 are taken relative to the cell's first corner so that
// the derivative sums keep the accuracy of differences.
template <int MODE>
GA_HD void rd_field_eval(const RdGrid &g, const RdCell &cl, int field, double &f, double *df, double *ddf)
{
    const AxisW wx = axis_weights(cl.xs, cl.dx, cl.ex[0], cl.ex[1]);
    const AxisW wy = axis_weights(cl.ys, cl.dy, cl.ey[0], cl.ey[1]);
    const double *base = g.tab + (size_t)field * g.nx * g.ny * g.nz * RD_Q;
    auto col = [&](int a, int b) {
        return base + (((size_t)cl.ci[a] * g.ny + cl.cj[b]) * g.nz + cl.kz) * RD_Q;
    };
    double ref0, ref1 = 0.0;
    {
        const double *p = col(1, 1), *q = p + RD_Q;
        ref0 = hv(p[0], q[0], p[1], q[1], cl.dz, cl.Xz);
        if constexpr (MODE >= 1) ref1 = hd1(p[0], q[0], p[1], q[1], cl.dz, cl.Xz);
    }
    double F = 0.0, Fx = 0.0, Fy = 0.0, Fz = 0.0;
    double Fxx = 0.0, Fyy = 0.0, Fzz = 0.0, Fxy = 0.0, Fxz = 0.0, Fyz = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double r0P = 0.0, r0Pd = 0.0, r0S = 0.0, r0T = 0.0, r1P = 0.0, r1Pd = 0.0, r2P = 0.0;
        double r3V = 0.0, r3S = 0.0, r4Pd = 0.0, r4Sd = 0.0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double *p = col(a, b), *q = p + RD_Q;
            const double Py = wy.V[b] + wy.S[b];
            const double q0 = hv(p[0], q[0], p[1], q[1], cl.dz, cl.Xz) - ref0;
            r0P += Py * q0;
            if constexpr (MODE >= 1) {
                const double Pdy = wy.Vd[b] + wy.Sd[b];
                const double q1 = hd1(p[0], q[0], p[1], q[1], cl.dz, cl.Xz) - ref1;
                r0Pd += Pdy * q0;
                r1P += Py * q1;
                if constexpr (MODE == 2) {
                    r1Pd += Pdy * q1;
                    r0S += wy.S[b] * q0;
                    r0T += wy.T[b] * q0;
                    r2P += Py * hd2(p[0], q[0], p[1], q[1], cl.dz, cl.Xz_u);
                    const double q3 = hv(p[2], q[2], p[3], q[3], cl.dz, cl.Xz);
                    r3V += wy.V[b] * q3;
                    r3S += wy.S[b] * q3;
                    const double q4 = hv(p[4], q[4], p[3], q[3], cl.dz, cl.Xz);
                    r4Pd += Pdy * q4;
                    r4Sd += wy.Sd[b] * q4;
                }
            }
        }
        const double Px = wx.V[a] + wx.S[a];
        F += Px * r0P;
        if constexpr (MODE >= 1) {
            const double Pdx = wx.Vd[a] + wx.Sd[a];
            Fx += Pdx * r0P;
            Fy += Px * r0Pd;
            Fz += Px * r1P;
            if constexpr (MODE == 2) {
                Fxy += Pdx * r0Pd;
                Fxz += Pdx * r1P;
                Fyz += Px * r1Pd;
                Fzz += Px * r2P;
                Fxx += Pdx * r3V + wx.Sd[a] * r3S + wx.T[a] * r0S;
                Fyy += wx.V[a] * r4Pd + wx.S[a] * r4Sd + wx.S[a] * r0T;
            }
        }
    }
    f = ref0 + F;
    if constexpr (MODE >= 1) {
        df[0] = Fx;
        df[1] = Fy;
        df[2] = ref1 + Fz;
    }
    if constexpr (MODE == 2) {
        ddf[0] = Fxx;
        ddf[1] = Fyy;
        ddf[2] = Fzz;
        ddf[3] = Fxy;
        ddf[4] = Fxz;
        ddf[5] = Fyz;
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
    const double xe = fmin(fmax(x, g.lo[0]), g.hi[0]);
    const double ye = fmin(fmax(y, g.lo[1]), g.hi[1]);
    AxisSeg sx = axis_guess(g.x, g.nx, xe, g.iv[0], g.lo[0]);
    AxisSeg sy = axis_guess(g.y, g.ny, ye, g.iv[1], g.lo[1]);
    if (!(axis_ok(sx, g.nx, xe) && axis_ok(sy, g.ny, ye))) {
        axis_fix(sx, g.x, g.nx, xe);
        axis_fix(sy, g.y, g.ny, ye);
    }
    const int kx = sx.k, ky = sy.k;
    const double dx = sx.v[2] - sx.v[1], X = (xe - sx.v[1]) / dx;
    const double dy = sy.v[2] - sy.v[1], Y = (ye - sy.v[1]) / dy;
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
