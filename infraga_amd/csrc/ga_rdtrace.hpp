// ga_rdtrace.hpp -- device ray tracing through a range-dependent spherical
// atmosphere over terrain (infraga-sph-rngdep -prop).
//
// One ray per lane, as in the stratified engine; the medium is the hybrid
// spline of ga_rngdep.hpp (c, u, v, rho over lat, lon, z) read through the
// cache hierarchy, and the ground is the terrain spline.  The equations are
// the 18 (or 6) spherical ray and auxiliary equations with every horizontal
// medium derivative, the boundary-layer wind scaling over terrain and,
// optionally, the boundary-layer vertical wind.
//
// Relation to the reference (paths relative to the infraGA source tree):
//   medium, calc_uvw     src/atmo/atmo_state.sph.rngdep.cpp:57-882
//   RHS                  src/geoac/geoac.eqset.sph.rngdep.cpp:238-436
//   reflection           geoac.eqset.sph.rngdep.cpp:117-221
//   checks, step size    geoac.eqset.sph.rngdep.cpp:227-233, 483-503
//   bookkeeping          geoac.eqset.sph.rngdep.cpp:547-575, 775-847
//   driver               src/main.infraga.sph.rngdep.cpp:412-554
// Differences from the reference's arithmetic: the terrain is a tensor
// Hermite patch (the reference re-solves a spline per call; the same
// function up to rounding), divisions and pow(x, 2) become reciprocals and
// products.  Every formula keeps the reference's terms, including its
// quirks: tan(nu_theta) and the repeated db_dlth in the reflection, the
// unclamped-z second derivative of c, winds scaled at every height by
// calc_uvw but only below the boundary-layer top by the point functions,
// and no minus sign on the arrival inclination.
#pragma once

#include "ga_rngdep.hpp"
#include "ga_trace.hpp"

namespace ga {

constexpr double RD_BL = 0.1;     // topo::bndlyr_param [km]
// topo::vert_wind_grad is 0.5 in the spherical sources and 1.0 in the 3D
// ones: RdMed::bl_sg holds the ratio.
constexpr int F_C = 0, F_U = 1, F_V = 2, F_RHO = 3;

struct RdMed {
    RdGrid g;
    TopoGrid t;
    int is_topo;
    int use_blw;
    double r0, z0, z_max, z_bndlyr;
    double zflat;       // ground height without terrain: z0 (sph), 0 (3D Cartesian)
    double tweak;
    double sbc[3];
    int strat;          // a stratified profile over terrain (identical columns):
                        // the stratified driver's arrival inclination sign
    int cart;           // the 3D Cartesian medium: (z, x, y) in the places of (r, lat, lon)
    double bl_sg;       // topo::bndlyr_param / topo::vert_wind_grad: 0.1 / 0.5 (sph), 0.1 / 1.0 (3D)
};

struct RdFanArgs {
    RdMed M;
    Limits L;
    double src0, src1, src2;   // r [km], lat, lon [rad]
    double freq, turn_ht_min;
    int bounces;
    int n_rays;
    int max_steps;
    int nlanes;
    int one_shot;
    const double *launches;    // (incl_deg, az_deg)
    double *raw;               // RAW doubles per arrival slot
    double *ypbuf;             // per-lane row k-1
    ga_arrival *out;
    int32_t *counts;
    int32_t *steps;
    unsigned int *queue;
    PathOut path;              // path output of the write-mode kernels
};

// A lane's vector stored lane-interleaved: element i at p[i * STRIDE] with p
// = base + lane.  The kernels keep the RK4 state rows (step start and
// accumulator) in LDS this way (64 lanes read 64 consecutive doubles, no
// bank conflicts), out of the registers the medium evaluation needs; a host
// build uses STRIDE 1 over a plain array.
template <int STRIDE>
struct LaneVec {
    double *p;
    GA_HD double &operator[](int i) const { return p[i * STRIDE]; }
};

// Value fence: the values pass through an empty asm statement, so nothing
// that uses them is scheduled before this point.  Without it the compiler
// interleaves the 18-equation algebra with the spline evaluation of the
// medium and spills most of the registers to scratch.
template <class T>
GA_HD void value_fence(T &v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double *q = reinterpret_cast<double *>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / sizeof(double)); ++i) asm volatile("" : "+v"(q[i]));
#else
    (void)v;
#endif
}

template <class V, class SV>
GA_HD void rd_upd(const Stage &S, const V &y, const SV &stg, const V &acc, int i, double f)
{
    acc[i] += S.ha * f;
    if (!S.last) stg[i] = y[i] + S.hn * f;
}

// --------------------------------------------------------------------------
// Terrain: z_g = r0 + h(lat, lon) (topo::z, dz, ddz, dddz with clamped
// arguments); without a terrain file flat ground at r0 + z0 in the
// spherical geometry (atmo_state.sph.rngdep.cpp:57-66) and at z = 0 in the
// Cartesian one, whose topo::z ignores z0 (atmo_state.3d.rngdep.cpp:45-52).
// --------------------------------------------------------------------------
struct Terr {
    double zg;          // r0 + height
    double d[2];        // dh/dlat, dh/dlon
    double dd[2][2];
    double ddd[2][2][2];
};

GA_HD double rd_ground(const RdMed &M, double lat, double lon)
{
    if (!M.is_topo) return M.r0 + M.zflat;
    double o[10];
    topo_all(M.t, lat, lon, o);
    return M.r0 + o[0];
}

GA_HD void rd_terrain(const RdMed &M, double lat, double lon, Terr &T)
{
    if (!M.is_topo) {
        T.zg = M.r0 + M.zflat;
        T.d[0] = T.d[1] = 0.0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                T.dd[i][j] = 0.0;
                T.ddd[0][i][j] = T.ddd[1][i][j] = 0.0;
            }
        return;
    }
    double o[10];   // z, zx, zy, zxx, zxy, zyy, zxxx, zxxy, zxyy, zyyy
    topo_all(M.t, lat, lon, o);
    T.zg = M.r0 + o[0];
    T.d[0] = o[1];
    T.d[1] = o[2];
    T.dd[0][0] = o[3];
    T.dd[0][1] = T.dd[1][0] = o[4];
    T.dd[1][1] = o[5];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) T.ddd[i][j][k] = o[6 + i + j + k];
}

// --------------------------------------------------------------------------
// Point functions atmo::c, rho, u, v, w and their first derivatives
// (atmo_state.sph.rngdep.cpp:149-403): winds carry the boundary-layer factor
// only below the boundary-layer top; derivative index 0 r, 1 lat, 2 lon.
// --------------------------------------------------------------------------
struct RdPt {
    double c, rho, u, v, w;
};

GA_HD void rd_point(const RdMed &M, double r, double lat, double lon, RdPt &P)
{
    // terrain first (see rd_medium_rhs)
    double sc = 1.0, X = 0.0, td0 = 0.0, td1 = 0.0, zg = 0.0;
    const bool bl = r - M.r0 < M.z_bndlyr;
    if (bl) {
        Terr T;
        rd_terrain(M, lat, lon, T);
        X = exp(-(r - T.zg) / RD_BL);
        sc = (1.0 - X) / (1.0 + X);
        td0 = T.d[0];
        td1 = T.d[1];
        zg = T.zg;
    }
    const RdCell cl = rd_cell(M.g, lat, lon, r - M.r0);
    double u0, v0;
    rd_field_eval<0>(M.g, cl, F_C, P.c, nullptr, nullptr);
    rd_field_eval<0>(M.g, cl, F_RHO, P.rho, nullptr, nullptr);
    rd_field_eval<0>(M.g, cl, F_U, u0, nullptr, nullptr);
    rd_field_eval<0>(M.g, cl, F_V, v0, nullptr, nullptr);
    P.w = 0.0;
    if (bl && M.is_topo && M.use_blw) {
        // atmo::w, :308-338; Cartesian: atmo_state.3d.rngdep.cpp:243-267
        // (u pairs with the x slope, v with the y slope, no metric factors)
        const double sg = M.bl_sg;
        const double q = X / ((1.0 + X) * (1.0 + X));
        const double dS0 = -2.0 / RD_BL * q * td0;
        const double dS1 = -2.0 / RD_BL * q * td1;
        double w0;
        if (M.cart) {
            w0 = -2.0 * RD_BL * (u0 * dS0 + v0 * dS1);
        } else {
            w0 = v0 * dS0 + u0 * dS1 / cos(lat);
            w0 *= -2.0 * RD_BL / (M.r0 + zg);
        }
        P.w = sc * X * w0 / (X + (sg / 2.0) * (1.0 - X * X));
    }
    P.u = u0 * sc;
    P.v = v0 * sc;
}

struct RdPtD {
    double c, dc[3];
    double u, du[3], v, dv[3], w, dw[3];
};

// Values and first derivatives (atmo::dc, du, dv, dw), used at the
// reflection point.
GA_HD void rd_point_d(const RdMed &M, double r, double lat, double lon, RdPtD &P)
{
    const RdCell cl = rd_cell(M.g, lat, lon, r - M.r0);
    double f, df[3], u0, du0[3], v0, dv0[3];
    rd_field_eval<1>(M.g, cl, F_C, f, df, nullptr);
    P.c = f;
    P.dc[0] = df[2]; P.dc[1] = df[0]; P.dc[2] = df[1];
    rd_field_eval<1>(M.g, cl, F_U, u0, df, nullptr);
    du0[0] = df[2]; du0[1] = df[0]; du0[2] = df[1];
    rd_field_eval<1>(M.g, cl, F_V, v0, df, nullptr);
    dv0[0] = df[2]; dv0[1] = df[0]; dv0[2] = df[1];
    double sc = 1.0, dsc[3] = {0.0, 0.0, 0.0};
    P.w = 0.0;
    for (int n = 0; n < 3; ++n) P.dw[n] = 0.0;
    if (r - M.r0 < M.z_bndlyr) {
        Terr T;
        rd_terrain(M, lat, lon, T);
        const double X = exp(-(r - T.zg) / RD_BL);
        const double opx2 = (1.0 + X) * (1.0 + X);
        sc = (1.0 - X) / (1.0 + X);
        double dX[3] = {-X / RD_BL, X / RD_BL * T.d[0], X / RD_BL * T.d[1]};
        for (int n = 0; n < 3; ++n) dsc[n] = -2.0 * dX[n] / opx2;
        if (M.is_topo && M.use_blw) {
            // atmo::w and atmo::dw, :308-403; Cartesian: atmo_state.3d.rngdep.cpp:243-331
            const double sg = M.bl_sg;
            const double cl_ = M.cart ? 1.0 : cos(lat), sl = M.cart ? 0.0 : sin(lat);
            const double gsc = M.cart ? -2.0 * RD_BL : -2.0 * RD_BL / (M.r0 + T.zg);
            double w0 = M.cart ? u0 * dsc[1] + v0 * dsc[2] : v0 * dsc[1] + u0 * dsc[2] / cl_;
            w0 *= gsc;
            const double N = sc * X * w0;
            const double D = X + (sg / 2.0) * (1.0 - X * X);
            P.w = N / D;
            for (int n = 0; n < 3; ++n) {
                double ddX[2];
                if (n > 0) {
                    ddX[0] = X / RD_BL * (T.dd[0][n - 1] + 1.0 / RD_BL * T.d[0] * T.d[n - 1]);
                    ddX[1] = X / RD_BL * (T.dd[1][n - 1] + 1.0 / RD_BL * T.d[1] * T.d[n - 1]);
                } else {
                    ddX[0] = -X / (RD_BL * RD_BL) * T.d[0];
                    ddX[1] = -X / (RD_BL * RD_BL) * T.d[1];
                }
                const double ddS0 = -2.0 * (ddX[0] / opx2 - 2.0 * dX[1] * dX[n] / (opx2 * (1.0 + X)));
                const double ddS1 = -2.0 * (ddX[1] / opx2 - 2.0 * dX[2] * dX[n] / (opx2 * (1.0 + X)));
                double dw0;
                if (M.cart) {
                    dw0 = u0 * ddS0 + v0 * ddS1;
                    if (n == 0) dw0 += du0[0] * dsc[1] + dv0[0] * dsc[2];
                } else {
                    dw0 = v0 * ddS0 + u0 * ddS1 / cl_;
                    if (n == 0) {
                        dw0 += dv0[0] * dsc[1];
                        dw0 += du0[0] * dsc[2] / cl_;
                    } else if (n == 1) {
                        dw0 += u0 * dsc[2] * sl / (cl_ * cl_);
                    }
                }
                dw0 *= gsc;
                const double dN = (dsc[n] * X + sc * dX[n]) * w0 + sc * X * dw0;
                const double dD = (1.0 - sg * X) * dX[n];
                P.dw[n] = (dN * D - N * dD) / (D * D);
            }
        }
    }
    P.u = u0 * sc;
    P.v = v0 * sc;
    for (int n = 0; n < 3; ++n) {
        P.du[n] = du0[n] * sc + u0 * dsc[n];
        P.dv[n] = dv0[n] * sc + v0 * dsc[n];
    }
}

// --------------------------------------------------------------------------
// Medium of the right-hand side: c from interp::eval_all at the raw
// (lat, lon, r - r0) (the spline clamps inside; the f_zz patch does not)
// and the winds of atmo::calc_uvw at the clamped point, scaled by the
// boundary-layer factor at every height.  Indices 0 r, 1 lat, 2 lon;
// second derivatives packed rr, tt, pp, rt, rp, tp.
// --------------------------------------------------------------------------
struct RdAtm {
    double c, dc[3], ddc[6];
    double u, du[3], ddu[6];
    double v, dv[3], ddv[6];
    double w, dw[3], ddw[6];
};

GA_HD int p6(int n, int m)   // packed index of the symmetric pair (n, m)
{
    if (n == m) return n;
    const int s = n + m;   // (0,1) -> 3, (0,2) -> 4, (1,2) -> 5
    return s == 1 ? 3 : (s == 2 ? 4 : 5);
}

// BLW: the boundary-layer vertical wind is on (a separate instantiation:
// its terms keep the terrain derivatives live across the spline fields).
// CART: the 3D Cartesian medium, whose c comes from atmo::c / dc / ddc at
// the clamped point (no f_zz patch).
template <bool SECOND, bool BLW, bool CART = false>
GA_HD void rd_medium_rhs(const RdMed &M, double r, double lat, double lon, RdAtm &A)
{
    constexpr int MODE = SECOND ? 2 : 1;
    // The terrain and the boundary-layer scaling come before the three
    // spline fields: with the terrain evaluated after them the compiler
    // issues all of the fields' loads at once and spills.
    Terr T;
    const bool in_bl = M.is_topo && (r - M.r0 < M.z_bndlyr);
    if (in_bl) {
        rd_terrain(M, lat, lon, T);
    } else {
        T.zg = rd_ground(M, lat, lon);
        T.d[0] = T.d[1] = 0.0;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                T.dd[i][j] = 0.0;
                T.ddd[0][i][j] = T.ddd[1][i][j] = 0.0;
            }
    }
    const double X = exp(-(r - T.zg) / RD_BL);
    const double ib = 1.0 / RD_BL;
    const double opx = 1.0 + X, iopx = 1.0 / opx, iopx2 = iopx * iopx, iopx3 = iopx2 * iopx;
    double dX[3], ddX[3][3];
    dX[0] = -X * ib;
    dX[1] = X * ib * T.d[0];
    dX[2] = X * ib * T.d[1];
    ddX[0][0] = X * ib * ib;
    for (int n = 1; n < 3; ++n) {
        ddX[n][0] = ddX[0][n] = -X * ib * ib * T.d[n - 1];
        for (int m = 1; m < 3; ++m) ddX[n][m] = (T.dd[n - 1][m - 1] + T.d[n - 1] * T.d[m - 1] * ib) * (X * ib);
    }
    const double sc0 = (1.0 - X) * iopx;
    double dsc0[3], ddsc0[3][3];
    for (int n = 0; n < 3; ++n) {
        dsc0[n] = -2.0 * dX[n] * iopx2;
        for (int m = 0; m < 3; ++m) ddsc0[n][m] = -2.0 * (ddX[n][m] * iopx2 - 2.0 * dX[n] * dX[m] * iopx3);
    }
    // c from eval_all at the raw point (the spline clamps inside, the f_zz
    // patch does not); u and v at the clamped point: the same cell with the
    // f_zz patch at the clamped height.
    RdCell cl = rd_cell(M.g, lat, lon, r - M.r0);
    if (CART) cl.Xz_u = cl.Xz;
    double f, df[3], ddf[6];
    rd_field_eval<MODE>(M.g, cl, F_C, f, df, ddf);
    A.c = f;
    A.dc[0] = df[2]; A.dc[1] = df[0]; A.dc[2] = df[1];
    if (SECOND) {
        A.ddc[0] = ddf[2]; A.ddc[1] = ddf[0]; A.ddc[2] = ddf[1];
        A.ddc[3] = ddf[4]; A.ddc[4] = ddf[5]; A.ddc[5] = ddf[3];
    }
    cl.Xz_u = cl.Xz;
    double u0, du0[3], ddu0[6], v0, dv0[3], ddv0[6];
    rd_field_eval<MODE>(M.g, cl, F_U, u0, df, ddf);
    du0[0] = df[2]; du0[1] = df[0]; du0[2] = df[1];
    if (SECOND) { ddu0[0] = ddf[2]; ddu0[1] = ddf[0]; ddu0[2] = ddf[1]; ddu0[3] = ddf[4]; ddu0[4] = ddf[5]; ddu0[5] = ddf[3]; }
    rd_field_eval<MODE>(M.g, cl, F_V, v0, df, ddf);
    dv0[0] = df[2]; dv0[1] = df[0]; dv0[2] = df[1];
    if (SECOND) { ddv0[0] = ddf[2]; ddv0[1] = ddf[0]; ddv0[2] = ddf[1]; ddv0[3] = ddf[4]; ddv0[4] = ddf[5]; ddv0[5] = ddf[3]; }
    A.u = u0 * sc0;
    A.v = v0 * sc0;
    for (int n = 0; n < 3; ++n) {
        A.du[n] = du0[n] * sc0 + u0 * dsc0[n];
        A.dv[n] = dv0[n] * sc0 + v0 * dsc0[n];
    }
    if (SECOND) {
        for (int n = 0; n < 3; ++n) {
            A.ddu[n] = ddu0[n] * sc0 + 2.0 * du0[n] * dsc0[n] + u0 * ddsc0[n][n];
            A.ddv[n] = ddv0[n] * sc0 + 2.0 * dv0[n] * dsc0[n] + v0 * ddsc0[n][n];
        }
        A.ddu[3] = ddu0[3] * sc0 + du0[0] * dsc0[1] + du0[1] * dsc0[0] + u0 * ddsc0[0][1];
        A.ddu[4] = ddu0[4] * sc0 + du0[0] * dsc0[2] + du0[2] * dsc0[0] + u0 * ddsc0[0][2];
        A.ddu[5] = ddu0[5] * sc0 + du0[1] * dsc0[2] + du0[2] * dsc0[1] + u0 * ddsc0[1][2];
        A.ddv[3] = ddv0[3] * sc0 + dv0[0] * dsc0[1] + dv0[1] * dsc0[0] + v0 * ddsc0[0][1];
        A.ddv[4] = ddv0[4] * sc0 + dv0[0] * dsc0[2] + dv0[2] * dsc0[0] + v0 * ddsc0[0][2];
        A.ddv[5] = ddv0[5] * sc0 + dv0[1] * dsc0[2] + dv0[2] * dsc0[1] + v0 * ddsc0[1][2];
    }
    A.w = 0.0;
    for (int n = 0; n < 3; ++n) A.dw[n] = 0.0;
    if (SECOND)
        for (int n = 0; n < 6; ++n) A.ddw[n] = 0.0;
    if (!BLW || !in_bl) return;
    struct {
        double u0, du0[3], ddu0[6], v0, dv0[3], ddv0[6];
    } W;
    W.u0 = u0; W.v0 = v0;
    for (int n = 0; n < 3; ++n) { W.du0[n] = du0[n]; W.dv0[n] = dv0[n]; }
    if (SECOND)
        for (int n = 0; n < 6; ++n) { W.ddu0[n] = ddu0[n]; W.ddv0[n] = ddv0[n]; }
    value_fence(W);
    u0 = W.u0; v0 = W.v0;
    for (int n = 0; n < 3; ++n) { du0[n] = W.du0[n]; dv0[n] = W.dv0[n]; }
    if (SECOND)
        for (int n = 0; n < 6; ++n) { ddu0[n] = W.ddu0[n]; ddv0[n] = W.ddv0[n]; }

    // boundary-layer vertical wind (calc_uvw :658-690, :808-876; Cartesian:
    // atmo_state.3d.rngdep.cpp:543-563 and :667-720, u with the x slope, v
    // with the y slope, no metric factors, no horizontal derivatives of u0
    // and v0)
    const double sg = M.bl_sg;
    const double cl_ = CART ? 1.0 : cos(lat), sl = CART ? 0.0 : sin(lat), icl = 1.0 / cl_;
    const double gsc = CART ? -2.0 * RD_BL : -2.0 * RD_BL / (M.r0 + T.zg);
    double w0 = CART ? u0 * dsc0[1] + v0 * dsc0[2] : v0 * dsc0[1] + u0 * dsc0[2] * icl;
    w0 *= gsc;
    const double N = sc0 * X * w0;
    const double D = X + (sg / 2.0) * (1.0 - X * X);
    A.w = N / D;
    if (!SECOND) {
        for (int n = 0; n < 3; ++n) {
            double dw0;
            if (CART) {
                dw0 = u0 * ddsc0[1][n] + v0 * ddsc0[2][n];
                if (n == 0) dw0 += du0[0] * dsc0[1] + dv0[0] * dsc0[2];
            } else if (n == 0) {
                dw0 = dv0[0] * dsc0[1] + v0 * ddsc0[1][0];
                dw0 += (du0[0] * dsc0[2] + u0 * ddsc0[2][0]) * icl;
            } else if (n == 1) {
                dw0 = v0 * ddsc0[1][1];
                dw0 += u0 * icl * (ddsc0[1][2] + dsc0[2] * (sl * icl));
            } else {
                dw0 = v0 * ddsc0[1][2];
                dw0 += u0 * ddsc0[2][2] * icl;
            }
            dw0 *= gsc;
            const double dN = (dsc0[n] * X + sc0 * dX[n]) * w0 + sc0 * X * dw0;
            const double dD = (1.0 - sg * X) * dX[n];
            A.dw[n] = (dN * D - N * dD) / (D * D);
        }
        return;
    }
    // third derivatives of X and of the scaling (k: terrain direction lat, lon)
    double dddX[2][3][3], dddsc0[2][3][3];
    const double ib2 = ib * ib;
    for (int k = 0; k < 2; ++k) {
        dddX[k][0][0] = X * ib2 * ib * T.d[k];
        for (int n = 1; n < 3; ++n) {
            dddX[k][n][0] = dddX[k][0][n] = (T.dd[n - 1][k] + T.d[n - 1] * T.d[k] * ib) * (-X * ib2);
            for (int m = 1; m < 3; ++m)
                dddX[k][n][m] = (T.ddd[k][n - 1][m - 1] + (T.dd[n - 1][m - 1] * T.d[k] + T.dd[k][m - 1] * T.d[n - 1] +
                                                           T.dd[k][n - 1] * T.d[m - 1]) * ib +
                                 T.d[k] * T.d[n - 1] * T.d[m - 1] * ib2) * (X * ib);
        }
        for (int n = 0; n < 3; ++n)
            for (int m = 0; m < 3; ++m) {
                double t = dddX[k][n][m] * iopx2;
                t -= 2.0 * (ddX[n][m] * dX[k + 1] + ddX[n][k + 1] * dX[m] + ddX[k + 1][m] * dX[n]) * iopx3;
                t += 6.0 * dX[n] * dX[m] * dX[k + 1] * (iopx2 * iopx2);
                dddsc0[k][n][m] = -2.0 * t;
            }
    }
    double dw0[3], dN[3], dD[3];
    if constexpr (CART) {
        for (int n = 0; n < 3; ++n) {
            dw0[n] = u0 * ddsc0[1][n] + v0 * ddsc0[2][n];
            if (n == 0) dw0[n] += du0[0] * dsc0[1] + dv0[0] * dsc0[2];
            dw0[n] *= gsc;
            dN[n] = (dsc0[n] * X + sc0 * dX[n]) * w0 + sc0 * X * dw0[n];
            dD[n] = (1.0 - sg * X) * dX[n];
            A.dw[n] = (dN[n] * D - N * dD[n]) / (D * D);
        }
        // every pair (p, q) of medium indices, packed as p6 (:667-720)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int p = k < 3 ? k : (k == 5 ? 1 : 0), q = k < 3 ? k : (k == 3 ? 1 : 2);
            double ddw0 = u0 * dddsc0[0][p][q] + v0 * dddsc0[1][p][q];
            if (p == 0) ddw0 += ddsc0[1][q] * du0[0] + ddsc0[2][q] * dv0[0];
            if (q == 0) ddw0 += ddsc0[1][p] * du0[0] + ddsc0[2][p] * dv0[0];
            if (p == 0 && q == 0) ddw0 += dsc0[1] * ddu0[0] + dsc0[2] * ddv0[0];
            ddw0 *= gsc;
            double ddN = ddsc0[p][q] * X * w0 + sc0 * ddX[p][q] * w0 + sc0 * X * ddw0;
            ddN += dsc0[p] * dX[q] * w0 + dsc0[p] * X * dw0[q] + sc0 * dX[p] * dw0[q];
            ddN += dsc0[q] * dX[p] * w0 + dsc0[q] * X * dw0[p] + sc0 * dX[q] * dw0[p];
            const double ddD = -sg * dX[p] * dX[q] + (1.0 - sg * X) * ddX[p][q];
            A.ddw[p6(p, q)] =
                ((ddN * D - dN[p] * dD[q] - dN[q] * dD[p] - N * ddD) * (D * D) + 2.0 * N * D * dD[p] * dD[q]) / (D * D * D * D);
        }
        return;
    }
    for (int n = 0; n < 3; ++n) {
        dw0[n] = v0 * ddsc0[1][n] + u0 * ddsc0[2][n] * icl;
        if (n == 0) dw0[n] += dv0[0] * dsc0[1] + du0[0] * dsc0[2] * icl;
        else if (n == 1) dw0[n] += u0 * dsc0[2] * sl * (icl * icl);
        dw0[n] *= gsc;
        dN[n] = (dsc0[n] * X + sc0 * dX[n]) * w0 + sc0 * X * dw0[n];
        dD[n] = (1.0 - sg * X) * dX[n];
        A.dw[n] = (dN[n] * D - N * dD[n]) / (D * D);

        double ddw0 = dddsc0[0][n][n] * v0 + dddsc0[1][n][n] * u0 * icl;
        if (n == 0) {
            ddw0 += 2.0 * ddsc0[1][0] * dv0[0] + dsc0[1] * ddv0[0];
            ddw0 += (2.0 * ddsc0[2][0] * du0[0] + dsc0[2] * ddu0[0]) * icl;
        } else if (n == 1) {
            ddw0 += 2.0 * sl * (icl * icl) * ddsc0[1][2] * u0;
            ddw0 += (sl * sl + 1.0) * (icl * icl * icl) * dsc0[2] * u0;
        }
        ddw0 *= gsc;
        double ddN = ddsc0[n][n] * X * w0 + sc0 * ddX[n][n] * w0 + sc0 * X * ddw0;
        ddN += 2.0 * (dsc0[n] * dX[n] * w0 + dsc0[n] * X * dw0[n] + sc0 * dX[n] * dw0[n]);
        const double ddD = -sg * dX[n] * dX[n] + (1.0 - sg * X) * ddX[n][n];
        A.ddw[n] = ((ddN * D - 2.0 * dN[n] * dD[n] - N * ddD) * (D * D) + 2.0 * N * D * (dD[n] * dD[n])) / (D * D * D * D);
    }
    // r-lat, r-lon, lat-lon (:848-876)
    {
        double ddw0 = dddsc0[0][0][1] * v0 + ddsc0[1][1] * dv0[0] + (dddsc0[1][0][1] * u0 + ddsc0[1][2] * du0[0]) * icl;
        ddw0 += sl * (icl * icl) * (ddsc0[0][2] * u0 + dsc0[2] * du0[0]);
        ddw0 *= gsc;
        double ddN = ddsc0[0][1] * X * w0 + sc0 * ddX[0][1] * w0 + sc0 * X * ddw0;
        ddN += dsc0[0] * dX[1] * w0 + dsc0[0] * X * dw0[1] + sc0 * dX[0] * dw0[1];
        ddN += dsc0[1] * dX[0] * w0 + dsc0[1] * X * dw0[0] + sc0 * dX[1] * dw0[0];
        const double ddD = -sg * dX[0] * dX[1] + (1.0 - sg * X) * ddX[0][1];
        A.ddw[3] = ((ddN * D - dN[0] * dD[1] - dN[1] * dD[0] - N * ddD) * (D * D) + 2.0 * N * D * dD[0] * dD[1]) / (D * D * D * D);
    }
    {
        double ddw0 = dddsc0[0][0][2] * v0 + ddsc0[1][2] * dv0[0] + (dddsc0[1][0][2] * u0 + ddsc0[2][2] * du0[0]) * icl;
        ddw0 *= gsc;
        double ddN = ddsc0[0][2] * X * w0 + sc0 * ddX[0][2] * w0 + sc0 * X * ddw0;
        ddN += dsc0[0] * dX[2] * w0 + dsc0[0] * X * dw0[2] + sc0 * dX[0] * dw0[2];
        ddN += dsc0[2] * dX[0] * w0 + dsc0[2] * X * dw0[0] + sc0 * dX[2] * dw0[0];
        const double ddD = -sg * dX[0] * dX[2] + (1.0 - sg * X) * ddX[0][2];
        A.ddw[4] = ((ddN * D - dN[0] * dD[2] - dN[2] * dD[0] - N * ddD) * (D * D) + 2.0 * N * D * dD[0] * dD[2]) / (D * D * D * D);
    }
    {
        double ddw0 = dddsc0[0][1][2] * v0 + dddsc0[1][1][2] * u0 * icl;
        ddw0 += sl * (icl * icl) * ddsc0[2][2] * u0;
        ddw0 *= gsc;
        double ddN = ddsc0[1][2] * X * w0 + sc0 * ddX[1][2] * w0 + sc0 * X * ddw0;
        ddN += dsc0[1] * dX[2] * w0 + dsc0[1] * X * dw0[2] + sc0 * dX[1] * dw0[2];
        ddN += dsc0[2] * dX[1] * w0 + dsc0[2] * X * dw0[1] + sc0 * dX[2] * dw0[1];
        const double ddD = -sg * dX[1] * dX[2] + (1.0 - sg * X) * ddX[1][2];
        A.ddw[5] = ((ddN * D - dN[1] * dD[2] - dN[2] * dD[1] - N * ddD) * (D * D) + 2.0 * N * D * dD[1] * dD[2]) / (D * D * D * D);
    }
}

// --------------------------------------------------------------------------
// Right-hand side (update_refs + eval_src_eq, :238-436) and the fused RK4
// stage: f[i] is consumed by upd() as it is produced.
// --------------------------------------------------------------------------
template <int NEQ, bool BLW, class V, class SV>
GA_HD void rd_stage(const RdMed &M, const Stage &S, const V &y, const SV &stg, const V &acc)
{
    const double r = stg[0], t = stg[1];
    const double nu0 = stg[3], nu1 = stg[4], nu2 = stg[5];
    RdAtm A;
    rd_medium_rhs<NEQ == 18, BLW>(M, r, t, stg[2], A);
    value_fence(A);

    const double nu_mag = sqrt(nu0 * nu0 + nu1 * nu1 + nu2 * nu2);
    const double inu = 1.0 / nu_mag;
    const double cg0 = A.c * nu0 * inu + A.w;
    const double cg1 = A.c * nu1 * inu + A.v;
    const double cg2 = A.c * nu2 * inu + A.u;
    const double cg_mag = sqrt(cg0 * cg0 + cg1 * cg1 + cg2 * cg2);
    const double icg = 1.0 / cg_mag;
    double sn, cs;
    ga_sincos(t, sn, cs);
    const double ics = 1.0 / cs;
    const double tn = sn * ics;
    const double ir = 1.0 / r;
    const double G[3] = {1.0, ir, ir * ics};
    const double T[3] = {-ir * (nu1 * cg1 + nu2 * cg2),
                         (nu0 * A.v - nu1 * A.w) + nu0 * cg1 + nu2 * cg2 * tn,
                         (nu0 * A.u - nu2 * A.w) * cs + (nu1 * A.u - nu2 * A.v) * sn + cg2 * (nu0 * cs - nu1 * sn)};
    double Sg[3];   // |nu| grad c + nu . grad(w, v, u)
    for (int i = 0; i < 3; ++i) Sg[i] = nu_mag * A.dc[i] + nu0 * A.dw[i] + nu1 * A.dv[i] + nu2 * A.du[i];
    const double cgv[3] = {cg0, cg1, cg2};
    double f[6];
    for (int i = 0; i < 3; ++i) f[i] = G[i] * cgv[i] * icg;
    for (int i = 0; i < 3; ++i) f[3 + i] = -G[i] * icg * (Sg[i] + T[i]);

    if (NEQ == 18) {
        const double icg2 = icg * icg;
        const double inu2 = inu * inu;
        const double ir2 = ir * ir;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int oR = j ? 12 : 6, om = j ? 15 : 9;
            const double Rj[3] = {stg[oR], stg[oR + 1], stg[oR + 2]};
            const double mj[3] = {stg[om], stg[om + 1], stg[om + 2]};
            // launch-angle derivatives of the medium along R (refs.dc[3], ddc[m][j], ...)
            double dcj = 0.0, dwj = 0.0, dvj = 0.0, duj = 0.0, ddc[3], ddw[3], ddv[3], ddu[3];
            for (int n = 0; n < 3; ++n) {
                dcj += Rj[n] * A.dc[n];
                dwj += Rj[n] * A.dw[n];
                dvj += Rj[n] * A.dv[n];
                duj += Rj[n] * A.du[n];
            }
            for (int m = 0; m < 3; ++m) {
                ddc[m] = ddw[m] = ddv[m] = ddu[m] = 0.0;
                for (int n = 0; n < 3; ++n) {
                    const int q = p6(m, n);
                    ddc[m] += Rj[n] * A.ddc[q];
                    ddw[m] += Rj[n] * A.ddw[q];
                    ddv[m] += Rj[n] * A.ddv[q];
                    ddu[m] += Rj[n] * A.ddu[q];
                }
            }
            const double dnu = (nu0 * mj[0] + nu1 * mj[1] + nu2 * mj[2]) * inu;
            const double cnd = A.c * inu2 * dnu;
            const double dcg0 = nu0 * inu * dcj + A.c * mj[0] * inu - cnd * nu0 + dwj;
            const double dcg1 = nu1 * inu * dcj + A.c * mj[1] * inu - cnd * nu1 + dvj;
            const double dcg2 = nu2 * inu * dcj + A.c * mj[2] * inu - cnd * nu2 + duj;
            const double dcgm = (cg0 * dcg0 + cg1 * dcg1 + cg2 * dcg2) * icg;
            const double dG[3] = {0.0, -Rj[0] * ir2, -Rj[0] * ir2 * ics + sn * ir * ics * ics * Rj[1]};
            double dT[3];
            dT[0] = Rj[0] * ir2 * (nu1 * cg1 + nu2 * cg2) - ir * (mj[1] * cg1 + nu1 * dcg1 + mj[2] * cg2 + nu2 * dcg2);
            dT[1] = (mj[0] * A.v + nu0 * dvj - mj[1] * A.w - nu1 * dwj);
            dT[1] += mj[0] * cg1 + nu0 * dcg1 + (mj[2] * cg2 + nu2 * dcg2) * tn + nu2 * cg2 * Rj[1] * ics * ics;
            dT[2] = (mj[0] * A.u + nu0 * duj - mj[2] * A.w - nu2 * dwj) * cs - (nu0 * A.u - nu2 * A.w) * Rj[1] * sn;
            dT[2] += (mj[1] * A.u + nu1 * duj - mj[2] * A.v - nu2 * dvj) * sn + (nu1 * A.u - nu2 * A.v) * Rj[1] * cs;
            dT[2] += dcg2 * (nu0 * cs - nu1 * sn) + cg2 * ((mj[0] - nu1 * Rj[1]) * cs - (mj[1] + nu0 * Rj[1]) * sn);
            const double dcgv[3] = {dcg0, dcg1, dcg2};
            const double cgm2 = dcgm * icg2;
            for (int i = 0; i < 3; ++i) {
                const double fR = dG[i] * cgv[i] * icg + G[i] * dcgv[i] * icg - G[i] * cgv[i] * cgm2;
                const double fm = -dG[i] * icg * (Sg[i] + T[i]) + G[i] * cgm2 * Sg[i] -
                                  G[i] * icg * (dnu * A.dc[i] + nu_mag * ddc[i] + mj[0] * A.dw[i] + mj[1] * A.dv[i] +
                                                mj[2] * A.du[i] + nu0 * ddw[i] + nu1 * ddv[i] + nu2 * ddu[i] + dT[i]);
                rd_upd(S, y, stg, acc, oR + i, fR);
                rd_upd(S, y, stg, acc, om + i, fm);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) rd_upd(S, y, stg, acc, i, f[i]);
}

GA_HD double rd_set_ds(const RdMed &M, const Limits &L, double r, double lat, double lon)
{
    double ds = L.ds_max;
    if (r - M.r0 < M.z_bndlyr) ds -= (L.ds_max - L.ds_min) * exp(-(r - rd_ground(M, lat, lon)) / 0.2);
    return ds;
}

// One RK4 step from y into y1 (rows of V), stg holding the stage point.
template <int NEQ, bool BLW, class V, class SV>
GA_HD double rd_rk4_step(const RdMed &M, const Limits &L, const V &y, const V &y1, SV &stg)
{
    const double ds = rd_set_ds(M, L, y[0], y[1], y[2]);
    const double h6 = ds * (1.0 / 6.0), h3 = ds * (1.0 / 3.0), h2 = ds * 0.5;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) { stg[i] = y[i]; y1[i] = y[i]; }
#pragma clang loop unroll(disable)
    for (int st = 0; st < 4; ++st) {
        const double ha = (st == 0 || st == 3) ? h6 : h3;
        const double hn = (st == 2) ? ds : h2;
        rd_stage<NEQ, BLW>(M, Stage{ha, hn, false}, y, stg, y1);
    }
    return ds;
}

GA_HD bool rd_leaves_region(const RdFanArgs &A, const Limits &L, double *y)
{
    if (y[2] > PI) y[2] -= 2.0 * PI;
    else if (y[2] < -PI) y[2] += 2.0 * PI;
    if (y[1] < -PI / 2.0 || y[1] > PI / 2.0) return true;
    if (y[0] - A.M.r0 > L.alt_max) return true;
    // exact haversine test only when the bound (d_lat^2 + d_lon^2) / 4 >= h
    // is near the threshold (as in the stratified leaves_region)
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
}

GA_HD bool rd_below_ground(const RdMed &M, const double *y)
{
    return (y[0] - M.r0 <= M.z_max) && (y[0] < rd_ground(M, y[1], y[2]));
}

// Travel time and absorption of the segment y0 -> y1 at its midpoint.
GA_HD void rd_seg_book(const RdMed &M, const double *y0, const double *y1, double freq, double &dt, double &datt)
{
    const double dr = y1[0] - y0[0], dth = y1[1] - y0[1];
    double dp = y1[2] - y0[2];
    const double r = y0[0] + dr * 0.5, t = y0[1] + dth * 0.5, p = y0[2] + dp * 0.5;
    if (dp > PI) dp -= 2.0 * PI;
    else if (dp < -PI) dp += 2.0 * PI;
    const double a = r * dth, b = r * cos(t) * dp;
    const double ds = sqrt(dr * dr + a * a + b * b);
    const double n0 = y0[3] + (y1[3] - y0[3]) * 0.5;
    const double n1 = y0[4] + (y1[4] - y0[4]) * 0.5;
    const double n2 = y0[5] + (y1[5] - y0[5]) * 0.5;
    const double inm = 1.0 / sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    RdPt P;
    rd_point(M, r, t, p, P);
    value_fence(P);
    const double c0 = P.c * n0 * inm + P.w, c1 = P.c * n1 * inm + P.v, c2 = P.c * n2 * inm + P.u;
    dt = ds / sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    datt = sb_alpha<true>(r - M.r0, P.c, P.rho, freq, M.tweak, M.sbc) * ds;
}

// Launch state (set_initial, :68-112); th, ph in radians.
template <int NEQ>
GA_HD void rd_init_ray(const RdFanArgs &A, double th, double ph, double *y, double &c0, double &nu0s)
{
    RdPt P;
    rd_point(A.M, A.src0, A.src1, A.src2, P);
    c0 = P.c;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double mach[3] = {P.w / c0, P.v / c0, P.u / c0};
    const double n0[3] = {st, ct * sp, ct * cp};
    const double mt[3] = {ct, -st * sp, -st * cp};
    const double mp[3] = {0.0, ct * cp, -ct * sp};
    const double msc = 1.0 + (n0[0] * mach[0] + n0[1] * mach[1] + n0[2] * mach[2]);
    const double imsc = 1.0 / msc;
    nu0s = imsc;
    y[0] = A.src0; y[1] = A.src1; y[2] = A.src2;
    for (int i = 0; i < 3; ++i) y[3 + i] = n0[i] * imsc;
    if (NEQ == 18) {
        const double a = mt[0] * mach[0] + mt[1] * mach[1] + mt[2] * mach[2];
        const double b = mp[0] * mach[0] + mp[1] * mach[1] + mp[2] * mach[2];
        for (int i = 0; i < 3; ++i) {
            y[6 + i] = 0.0;
            y[12 + i] = 0.0;
            y[9 + i] = mt[i] * imsc - n0[i] * (imsc * imsc) * a;
            y[15 + i] = mp[i] * imsc - n0[i] * (imsc * imsc) * b;
        }
    }
}

// Reflection off the terrain (approx_intercept + set_refl, :117-221): rows
// K-2 and K-1 of the leg in, row 0 of the next leg out.
template <int NEQ>
GA_HD void rd_reflect(const RdMed &M, const double *yk2, const double *yk1, double *y, double c0)
{
    double prev[NEQ];
    {
        Terr T1;
        rd_terrain(M, yk1[1], yk1[2], T1);
        const double dr_grnd = T1.zg - yk1[0];
        const double dr_k = (yk1[0] - yk2[0]) - (yk1[1] - yk2[1]) / yk1[0] * T1.d[0]
                            - (yk1[2] - yk2[2]) / (yk1[0] * cos(yk1[1])) * T1.d[1];
#pragma unroll
        for (int i = 0; i < NEQ; ++i) prev[i] = yk1[i] + (yk1[i] - yk2[i]) / dr_k * dr_grnd;
    }
    Terr T;
    rd_terrain(M, prev[1], prev[2], T);
    const double rg = T.zg;
    const double cl = cos(prev[1]);
    const double aa = 1.0 / rg * T.d[0];
    const double b = 1.0 / (rg * cl) * T.d[1];
    const double norm1 = 1.0 + aa * aa + b * b;
    const double in1 = 1.0 / norm1;
    const double C1r = (1.0 - aa * aa - b * b) * in1;
    const double C1th = (1.0 - aa * aa + b * b) * in1, C2th = 2.0 * aa * in1;
    const double C1ph = (1.0 + aa * aa - b * b) * in1, C2ph = 2.0 * b * in1;
    y[0] = rg;
    y[1] = prev[1];
    y[2] = prev[2];
    y[3] = -C1r * prev[3] + C2th * prev[4] + C2ph * prev[5];
    y[4] = C1th * prev[4] + C2th * (prev[3] - b * prev[5]);
    y[5] = C1ph * prev[5] + C2ph * (prev[3] - aa * prev[4]);
    if (NEQ < 18) return;

    RdPtD P;
    rd_point_d(M, rg, prev[1], prev[2], P);
    value_fence(P);
    const double c_grnd = P.c;
    const double norm2 = prev[3] - aa * prev[4] - b * prev[5];
    const double tn4 = tan(prev[4]);   // the reference's tan(prev[4]) (:161, :164)
    double da[2], db[2], ds0[2];
    for (int j = 0; j < 2; ++j) {
        const int o = j ? 12 : 6;
        const double Ra = prev[o], Rt = prev[o + 1], Rp = prev[o + 2];
        const double k1 = (Rt * (prev[3] - b * prev[5]) - prev[4] * (Ra / rg - cl * b * Rp)) / norm2;
        const double k2 = (Rp * (prev[3] - aa * prev[4]) - prev[5] * (Ra / (rg * cl) - aa / cl * Rt)) / norm2;
        da[j] = -Ra / rg * aa + 1.0 / rg * (k1 * T.dd[0][0] + k2 * T.dd[0][1]);
        db[j] = -(Ra / rg - tn4 * Rt) * b + 1.0 / (rg * cl) * (k1 * T.dd[1][0] + k2 * T.dd[1][1]);
        ds0[j] = -c0 / c_grnd * (Ra - rg * aa * Rt - rg * cl * b * Rp) / norm2;
    }
    const double in12 = in1 * in1;
    double dC1r[2], dC1th[2], dC1ph[2], dC2th[2], dC2ph[2];
    for (int j = 0; j < 2; ++j) {
        dC1r[j] = -4.0 * in12 * (aa * da[j] + b * db[j]);
        dC1th[j] = -4.0 * aa * in12 * ((1.0 + b * b) * da[j] - aa * b * db[j]);
        dC1ph[j] = -4.0 * b * in12 * ((1.0 + aa * aa) * db[j] - aa * b * da[j]);
        dC2th[j] = 2.0 * in12 * ((1.0 - aa * aa + b * b) * da[j] - 2.0 * aa * b * db[j]);
    }
    dC2ph[0] = 2.0 * in12 * ((1.0 + aa * aa - b * b) * db[0] - 2.0 * aa * b * db[0]);   // db_dlth twice (:171)
    dC2ph[1] = 2.0 * in12 * ((1.0 + aa * aa - b * b) * db[1] - 2.0 * aa * b * da[1]);

    const double g2 = c_grnd * c_grnd / c0;
    double dnr = c0 / c_grnd * P.dc[0] + prev[3] * P.dw[0] + prev[4] * P.dv[0] + prev[5] * P.du[0];
    double dnt = c0 / c_grnd * P.dc[1] + prev[3] * P.dw[1] + prev[4] * P.dv[1] + prev[5] * P.du[1];
    double dnp = c0 / c_grnd * P.dc[2] + prev[3] * P.dw[2] + prev[4] * P.dv[2] + prev[5] * P.du[2];
    dnr += g2 / rg * (prev[4] * prev[4] + prev[5] * prev[5]);
    dnt -= g2 * (prev[3] * prev[4] + tan(prev[1]) * (prev[5] * prev[5]));
    dnp -= g2 * prev[5] * (prev[3] * cl + prev[4] * sin(prev[1]));
    dnr *= -1.0 / c_grnd;
    dnt *= -1.0 / (c_grnd * rg);
    dnp *= -1.0 / (c_grnd * rg * cl);
    for (int j = 0; j < 2; ++j) {
        const int o = j ? 12 : 6;
        y[o] = -C1r * prev[o] + rg * C2th * prev[o + 1] + rg * cl * C2ph * prev[o + 2];
        y[o + 1] = C1th * prev[o + 1] + C2th * (prev[o] / rg - b * cl * prev[o + 2]);
        y[o + 2] = C1ph * prev[o + 2] + C2ph * (prev[o] / (rg * cl) - aa / cl * prev[o + 1]);
        y[o + 3] = -C1r * prev[o + 3] + C2th * prev[o + 4] + C2ph * prev[o + 5] - dC1r[j] * prev[3] + dC2th[j] * prev[4] +
                   dC2ph[j] * prev[5] + (-(C1r + 1.0) * dnr + C2th * dnt + C2ph * dnp) * ds0[j];
        y[o + 4] = C1th * prev[o + 4] + C2th * (prev[o + 3] - b * prev[o + 5] - db[j] * prev[5]) + dC1th[j] * prev[4] +
                   dC2th[j] * (prev[3] - b * prev[5]) + ((C1th - 1.0) * dnt + C2th * (dnr - b * dnp)) * ds0[j];
        y[o + 5] = C1ph * prev[o + 5] + C2ph * (prev[o + 3] - aa * prev[o + 4] - da[j] * prev[4]) + dC1ph[j] * prev[5] +
                   dC2ph[j] * (prev[3] - aa * prev[4]) + ((C1ph - 1.0) * dnp + C2ph * (dnr - aa * dnt)) * ds0[j];
    }
}

// Jacobian determinant at row y with medium values P (geoac.eqset.sph.rngdep.cpp:798-821).
GA_HD double rd_jacobian(const double *y, const RdPt &P)
{
    const double r = y[0], ct_ = cos(y[1]);
    const double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    const double p0 = P.c * y[3] / nm + P.w, p1 = P.c * y[4] / nm + P.v, p2 = P.c * y[5] / nm + P.u;
    const double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2);
    const double drs = p0 / pm, dts = 1.0 / r * p1 / pm, dps = 1.0 / (r * ct_) * p2 / pm;
    return r * r * ct_ * (drs * (y[7] * y[14] - y[13] * y[8]) - y[6] * (dts * y[14] - dps * y[13]) +
                          y[12] * (dts * y[8] - dps * y[7]));
}

// 20 log10 of the amplitude at row y (geoac.eqset.sph.rngdep.cpp:823-847);
// P: medium at y, S: medium at the source.
GA_HD double rd_amp_db(const double *y, const RdPt &P, const RdPt &S, double c0, double nu0s, double th, double ph)
{
    const double jac = rd_jacobian(y, P);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double nm = (c0 - y[3] * P.w - y[4] * P.v - y[5] * P.u) / P.c;
    const double q0 = c0 * (nu0s * st) / nu0s + S.w, q1 = c0 * (nu0s * ct * sp) / nu0s + S.v,
                 q2 = c0 * (nu0s * ct * cp) / nu0s + S.u;
    const double p0 = P.c * y[3] / nm + P.w, p1 = P.c * y[4] / nm + P.v, p2 = P.c * y[5] / nm + P.u;
    const double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2);
    const double qm = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
    const double num = P.rho * nm * (P.c * P.c * P.c) * qm * ct;
    const double den = S.rho * nu0s * (S.c * S.c * S.c) * pm * jac;
    return 20.0 * log10(sqrt(fabs(num / den)));
}

// Arrival row (run_prop :515-539): inclination without the stratified
// driver's minus sign; TL from the Jacobian and amplitude (:798-847).
template <int NEQ>
GA_HD ga_arrival rd_arrival_row(const RdFanArgs &A, const double *y, int ray, int bnc, double incl, double az, double tt,
                                double att, double rmax, double c0, double nu0s, double th, double ph)
{
    ga_arrival R;
    R.launch_idx = ray;
    R.n_bounce = bnc;
    R.incl = incl;
    R.az = az;
    RdPt P, S;
    rd_point(A.M, y[0], y[1], y[2], P);
    rd_point(A.M, A.src0, A.src1, A.src2, S);   // atmo::c(r0 + z_src, lat_src, lon_src)
    double baz;
    if (A.M.strat) {
        // main.infraga.sph.cpp:523-524
        R.incl_arr = -asin(P.c / S.c * y[3]) * (180.0 / PI);
        baz = 90.0 - atan2(-y[4], -y[5]) * (180.0 / PI);
    } else {
        R.incl_arr = asin(P.c / S.c * y[3]) * 180.0 / PI;
        baz = 90.0 - atan2(-y[4], -y[5]) * 180.0 / PI;
    }
    if (baz < -180.0) baz += 360.0;
    else if (baz > 180.0) baz -= 360.0;
    R.back_az = baz;
    R.pos0 = y[1] * (180.0 / PI);
    R.pos1 = y[2] * (180.0 / PI);
    const double s1 = sin((A.src1 - y[1]) / 2.0), s2 = sin((A.src2 - y[2]) / 2.0);
    R.cel = 2.0 * A.M.r0 * asin(sqrt(s1 * s1 + cos(y[1]) * cos(A.src1) * (s2 * s2))) / tt;
    R.time = tt;
    R.turn_ht = rmax;
    R.absorption = -2.0 * att;
    R.tl = (NEQ == 18) ? rd_amp_db(y, P, S, c0, nu0s, th, ph) : 0.0;
    return R;
}

// --------------------------------------------------------------------------
// 3D Cartesian range-dependent geometry (infraga-3d-rngdep:
// geoac.eqset.3d.rngdep.cpp over atmo_state.3d.rngdep.cpp).  The medium is
// the same hybrid-spline table with node coordinates in km and r0 = 0, so
// the point functions above take (z, x, y) in the places of (r, lat, lon)
// (medium derivative index 0 z, 1 x, 2 y).  The state row is the
// reference's (x, y, z, nu, X_th, mu_th, X_ph, mu_ph); the source is kept
// as src0 = z, src1 = x, src2 = y.
// --------------------------------------------------------------------------
GA_HD int c3(int n) { return n == 2 ? 0 : n + 1; }   // Cartesian index -> medium index

// update_refs + eval_src_eq (geoac.eqset.3d.rngdep.cpp:226-382): c and its
// derivatives at the clamped point (atmo::c, dc, ddc), winds of calc_uvw,
// the full symmetric second-derivative contraction and the refs.dv[3]
// accumulation of X_ph (:288), which leaves refs.dv[4] at zero.  A
// stratified profile over terrain (M.strat, infraga-3d with topo_file)
// contracts the upper triangle only (geoac.eqset.3d.strat.cpp:305-310): the
// lower terms enter with weight zero, which leaves every sum as the
// reference forms it.  BLW: the boundary-layer vertical wind w and its
// derivatives enter c_g[2], the nu equations and the auxiliary sums
// (refs.w, refs.dw, refs.ddw: :242-317, :349-379); without it they are zero
// and drop out at compile time.
template <int NEQ, bool BLW = false, class V, class SV>
GA_HD void rd3_stage(const RdMed &M, const Stage &S, const V &y, const SV &stg, const V &acc)
{
    const double nu0 = stg[3], nu1 = stg[4], nu2 = stg[5];
    RdAtm A;
    rd_medium_rhs<NEQ == 18, BLW, true>(M, stg[2], stg[0], stg[1], A);
    value_fence(A);
    double dc[3], du[3], dv[3], dw[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        dc[n] = A.dc[c3(n)];
        du[n] = A.du[c3(n)];
        dv[n] = A.dv[c3(n)];
        dw[n] = BLW ? A.dw[c3(n)] : 0.0;
    }
    const double Aw = BLW ? A.w : 0.0;
    const double nu_mag = sqrt(nu0 * nu0 + nu1 * nu1 + nu2 * nu2);
    const double inu = 1.0 / nu_mag;
    const double cg0 = A.c * nu0 * inu + A.u;
    const double cg1 = A.c * nu1 * inu + A.v;
    const double cg2 = A.c * nu2 * inu + Aw;
    const double cg_mag = sqrt(cg0 * cg0 + cg1 * cg1 + cg2 * cg2);
    const double icg = 1.0 / cg_mag;
    double Sg[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) Sg[n] = nu_mag * dc[n] + nu0 * du[n] + nu1 * dv[n] + nu2 * dw[n];
    const double cgv[3] = {cg0, cg1, cg2};

    if (NEQ == 18) {
        const double icg2 = icg * icg;
        const double inu2 = inu * inu;
        const double Xt[3] = {stg[6], stg[7], stg[8]}, Xp[3] = {stg[12], stg[13], stg[14]};
        double dv3 = 0.0;
#pragma unroll
        for (int n = 0; n < 3; ++n) dv3 += Xt[n] * dv[n] + Xp[n] * dv[n];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int oX = j ? 12 : 6, om = j ? 15 : 9;
            const double Xj[3] = {stg[oX], stg[oX + 1], stg[oX + 2]};
            const double mj[3] = {stg[om], stg[om + 1], stg[om + 2]};
            double dcj = 0.0, duj = 0.0, dwj = 0.0;
            double ddc[3], ddu[3], ddv[3], ddw[3];
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                dcj += Xj[n] * dc[n];
                duj += Xj[n] * du[n];
                if (BLW) dwj += Xj[n] * dw[n];
                ddc[n] = ddu[n] = ddv[n] = ddw[n] = 0.0;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int q = p6(c3(n), c3(m));
                    const double xm = (M.strat && m < n) ? 0.0 : Xj[m];
                    ddc[n] += xm * A.ddc[q];
                    ddu[n] += xm * A.ddu[q];
                    ddv[n] += xm * A.ddv[q];
                    if (BLW) ddw[n] += xm * A.ddw[q];
                }
            }
            const double dvj = j ? 0.0 : dv3;
            const double dnu = (nu0 * mj[0] + nu1 * mj[1] + nu2 * mj[2]) * inu;
            const double cnd = A.c * inu2 * dnu;
            const double dcg[3] = {nu0 * inu * dcj + A.c * mj[0] * inu - cnd * nu0 + duj,
                                   nu1 * inu * dcj + A.c * mj[1] * inu - cnd * nu1 + dvj,
                                   nu2 * inu * dcj + A.c * mj[2] * inu - cnd * nu2 + dwj};
            const double dcgm = (cg0 * dcg[0] + cg1 * dcg[1] + cg2 * dcg[2]) * icg;
            const double cgm2 = dcgm * icg2;
#pragma unroll
            for (int n = 0; n < 3; ++n) {
                const double fX = dcg[n] * icg - cgv[n] * cgm2;
                const double fm = cgm2 * Sg[n] - icg * (dnu * dc[n] + nu_mag * ddc[n] + mj[0] * du[n] + mj[1] * dv[n] +
                                                        mj[2] * dw[n] + nu0 * ddu[n] + nu1 * ddv[n] + nu2 * ddw[n]);
                rd_upd(S, y, stg, acc, oX + n, fX);
                rd_upd(S, y, stg, acc, om + n, fm);
            }
        }
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        rd_upd(S, y, stg, acc, n, cgv[n] * icg);
        rd_upd(S, y, stg, acc, 3 + n, -icg * Sg[n]);
    }
}

template <int NEQ, bool BLW = false, class V, class SV>
GA_HD double rd3_rk4_step(const RdMed &M, const Limits &L, const V &y, const V &y1, SV &stg)
{
    const double ds = rd_set_ds(M, L, y[2], y[0], y[1]);
    const double h6 = ds * (1.0 / 6.0), h3 = ds * (1.0 / 3.0), h2 = ds * 0.5;
#pragma unroll
    for (int i = 0; i < NEQ; ++i) { stg[i] = y[i]; y1[i] = y[i]; }
#pragma clang loop unroll(disable)
    for (int st = 0; st < 4; ++st) {
        const double ha = (st == 0 || st == 3) ? h6 : h3;
        const double hn = (st == 2) ? ds : h2;
        rd3_stage<NEQ, BLW>(M, Stage{ha, hn, false}, y, stg, y1);
    }
    return ds;
}

// break_check (:434-442): horizontal range from the source, the x / y
// region (L.b0..b3) and the altitude ceiling.
GA_HD bool rd3_leaves_region(const RdFanArgs &A, const Limits &L, const double *y)
{
    const double dx = y[0] - A.src1, dy = y[1] - A.src2;
    if (sqrt(dx * dx + dy * dy) > L.rng_max) return true;
    if (y[0] > L.b1 || y[0] < L.b0) return true;
    if (y[1] > L.b3 || y[1] < L.b2) return true;
    return y[2] > L.alt_max;
}

GA_HD bool rd3_below_ground(const RdMed &M, const double *y)
{
    return (y[2] <= M.z_max) && (y[2] < rd_ground(M, y[0], y[1]));
}

// travel_time / atten of one segment at its midpoint (:460-485, Cartesian
// Sutherland-Bass absorption).
GA_HD void rd3_seg_book(const RdMed &M, const double *y0, const double *y1, double freq, double &dt, double &datt)
{
    const double dx = y1[0] - y0[0], dy = y1[1] - y0[1], dz = y1[2] - y0[2];
    const double x = y0[0] + dx * 0.5, yy = y0[1] + dy * 0.5, z = y0[2] + dz * 0.5;
    const double ds = sqrt(dx * dx + dy * dy + dz * dz);
    const double n0 = y0[3] + (y1[3] - y0[3]) * 0.5;
    const double n1 = y0[4] + (y1[4] - y0[4]) * 0.5;
    const double n2 = y0[5] + (y1[5] - y0[5]) * 0.5;
    const double inm = 1.0 / sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    RdPt P;
    rd_point(M, z, x, yy, P);
    value_fence(P);
    const double c0 = P.c * n0 * inm + P.u, c1 = P.c * n1 * inm + P.v, c2 = P.c * n2 * inm + P.w;
    dt = ds / sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    datt = sb_alpha<false>(z, P.c, P.rho, freq, M.tweak, M.sbc) * ds;
}

// set_initial (:58-105); th, ph in radians.
template <int NEQ>
GA_HD void rd3_init_ray(const RdFanArgs &A, double th, double ph, double *y, double &c0, double &nu0s)
{
    RdPt P;
    rd_point(A.M, A.src0, A.src1, A.src2, P);
    c0 = P.c;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double mach[3] = {P.u / c0, P.v / c0, P.w / c0};
    const double n0[3] = {ct * cp, ct * sp, st};
    const double mt[3] = {-st * cp, -st * sp, ct};
    const double mp[3] = {-ct * sp, ct * cp, 0.0};
    const double msc = 1.0 + (n0[0] * mach[0] + n0[1] * mach[1] + n0[2] * mach[2]);
    const double imsc = 1.0 / msc;
    nu0s = imsc;
    y[0] = A.src1; y[1] = A.src2; y[2] = A.src0;
    for (int i = 0; i < 3; ++i) y[3 + i] = n0[i] * imsc;
    if (NEQ == 18) {
        const double a = mt[0] * mach[0] + mt[1] * mach[1] + mt[2] * mach[2];
        const double b = mp[0] * mach[0] + mp[1] * mach[1] + mp[2] * mach[2];
        for (int i = 0; i < 3; ++i) {
            y[6 + i] = 0.0;
            y[12 + i] = 0.0;
            y[9 + i] = mt[i] * imsc - n0[i] * (imsc * imsc) * a;
            y[15 + i] = mp[i] * imsc - n0[i] * (imsc * imsc) * b;
        }
    }
}

// approx_intercept + set_refl (:108-213): rows K-2 and K-1 of the leg in,
// row 0 of the next leg out; terrain slopes a = dz/dx, b = dz/dy.
template <int NEQ>
GA_HD void rd3_reflect(const RdMed &M, const double *yk2, const double *yk1, double *y, double c0)
{
    double prev[NEQ];
    {
        Terr T1;
        rd_terrain(M, yk1[0], yk1[1], T1);
        const double dz_grnd = T1.zg - yk1[2];
        const double dz_k = (yk1[2] - yk2[2]) - (yk1[0] - yk2[0]) * T1.d[0] - (yk1[1] - yk2[1]) * T1.d[1];
#pragma unroll
        for (int i = 0; i < NEQ; ++i) prev[i] = yk1[i] + (yk1[i] - yk2[i]) / dz_k * dz_grnd;
    }
    Terr T;
    rd_terrain(M, prev[0], prev[1], T);
    const double zg = T.zg, a = T.d[0], b = T.d[1];
    const double norm1 = 1.0 + a * a + b * b, in1 = 1.0 / norm1;
    const double C1x = (1.0 - a * a + b * b) * in1, C2x = 2.0 * a * in1;
    const double C1y = (1.0 + a * a - b * b) * in1, C2y = 2.0 * b * in1;
    const double C1z = (1.0 - a * a - b * b) * in1;
    y[0] = prev[0];
    y[1] = prev[1];
    y[2] = zg;
    y[3] = C1x * prev[3] + C2x * (prev[5] - prev[4] * b);
    y[4] = C1y * prev[4] + C2y * (prev[5] - prev[3] * a);
    y[5] = -C1z * prev[5] + prev[3] * C2x + prev[4] * C2y;
    if (NEQ < 18) return;

    RdPtD P;
    rd_point_d(M, zg, prev[0], prev[1], P);
    value_fence(P);
    const double cg = P.c;
    const double norm2 = prev[5] - a * prev[3] - b * prev[4];
    double da[2], db[2], ds0[2];
    for (int j = 0; j < 2; ++j) {
        const int o = j ? 12 : 6;
        const double k1 = (prev[o] * (prev[5] - b * prev[4]) - prev[3] * (prev[o + 2] - b * prev[o + 1])) / norm2;
        const double k2 = (prev[o + 1] * (prev[5] - a * prev[3]) - prev[4] * (prev[o + 2] - a * prev[o])) / norm2;
        da[j] = k1 * T.dd[0][0] + k2 * T.dd[0][1];
        db[j] = k1 * T.dd[1][0] + k2 * T.dd[1][1];
        ds0[j] = -c0 / cg * (prev[o + 2] - a * prev[o] - b * prev[o + 1]) / norm2;
    }
    const double in12 = in1 * in1;
    double dC1x[2], dC1y[2], dC1z[2], dC2x[2], dC2y[2];
    for (int j = 0; j < 2; ++j) {
        dC1x[j] = -4.0 * a * in12 * ((1.0 + b * b) * da[j] - a * b * db[j]);
        dC1y[j] = -4.0 * b * in12 * ((1.0 + a * a) * db[j] - a * b * da[j]);
        dC1z[j] = -4.0 * in12 * (a * da[j] + b * db[j]);
        dC2x[j] = 2.0 * in12 * ((1.0 - a * a + b * b) * da[j] - 2.0 * a * b * db[j]);
        dC2y[j] = 2.0 * in12 * ((1.0 + a * a - b * b) * db[j] - 2.0 * a * b * da[j]);
    }
    // medium derivative index: 1 x, 2 y, 0 z
    const double icg = -1.0 / cg, cr = c0 / cg;
    const double dnx = icg * (cr * P.dc[1] + prev[3] * P.du[1] + prev[4] * P.dv[1] + prev[5] * P.dw[1]);
    const double dny = icg * (cr * P.dc[2] + prev[3] * P.du[2] + prev[4] * P.dv[2] + prev[5] * P.dw[2]);
    const double dnz = icg * (cr * P.dc[0] + prev[3] * P.du[0] + prev[4] * P.dv[0] + prev[5] * P.dw[0]);
    for (int j = 0; j < 2; ++j) {
        const int o = j ? 12 : 6;
        y[o] = C1x * prev[o] + C2x * (prev[o + 2] - prev[o + 1] * b);
        y[o + 1] = C1y * prev[o + 1] + C2y * (prev[o + 2] - prev[o] * a);
        y[o + 2] = -C1z * prev[o + 2] + prev[o] * C2x + prev[o + 1] * C2y;
        y[o + 3] = C1x * prev[o + 3] + C2x * (prev[o + 5] - b * prev[o + 4] - prev[4] * db[j]) + dC1x[j] * prev[3] +
                   dC2x[j] * (prev[5] - b * prev[4]) + ((C1x - 1.0) * dnx + C2x * (dnz - b * dny)) * ds0[j];
        y[o + 4] = C1y * prev[o + 4] + C2y * (prev[o + 5] - a * prev[o + 3] - prev[3] * da[j]) + dC1y[j] * prev[4] +
                   dC2y[j] * (prev[5] - a * prev[3]) + ((C1y - 1.0) * dny + C2y * (dnz - a * dnx)) * ds0[j];
        y[o + 5] = -C1z * prev[o + 5] + C2x * prev[o + 3] + C2y * prev[o + 4] - dC1z[j] * prev[5] + dC2x[j] * prev[3] +
                   dC2y[j] * prev[4] + (-(C1z + 1.0) * dnz + C2x * dnx + C2y * dny) * ds0[j];
    }
}

// jacobian (:704-721) at row y with medium values P.
GA_HD double rd3_jacobian(const double *y, const RdPt &P)
{
    const double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    const double p0 = P.c * y[3] / nm + P.u, p1 = P.c * y[4] / nm + P.v, p2 = P.c * y[5] / nm + P.w;
    const double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2);
    const double dxs = p0 / pm, dys = p1 / pm, dzs = p2 / pm;
    return dxs * (y[7] * y[14] - y[13] * y[8]) - y[6] * (dys * y[14] - dzs * y[13]) + y[12] * (dys * y[8] - dzs * y[7]);
}

// 20 log10 of amp (:723-746); P at y, S at the source.
GA_HD double rd3_amp_db(const double *y, const RdPt &P, const RdPt &S, double c0, double th, double ph)
{
    const double jac = rd3_jacobian(y, P);
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    const double nm0 = 1.0 - (y[3] * S.u + y[4] * S.v + y[5] * S.w) / c0;
    const double p0 = P.c * y[3] / nm + P.u, p1 = P.c * y[4] / nm + P.v, p2 = P.c * y[5] / nm + P.w;
    const double q0 = c0 * ct * cp + S.u, q1 = c0 * ct * sp + S.v, q2 = c0 * st + S.w;
    const double pm = sqrt(p0 * p0 + p1 * p1 + p2 * p2), qm = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
    const double num = P.rho * nm * (P.c * P.c * P.c) * qm * ct;
    const double den = S.rho * nm0 * (c0 * c0 * c0) * pm * jac;
    return 20.0 * log10(sqrt(fabs(num / den)));
}

GA_HD double rd3_tl(const RdFanArgs &A, const double *y, const RdPt &P, double c0, double th, double ph)
{
    RdPt S;
    rd_point(A.M, A.src0, A.src1, A.src2, S);
    return rd3_amp_db(y, P, S, c0, th, ph);
}

// Arrival row (main.infraga.3d.rngdep.cpp:494-525): inclination from c on
// the terrain under the arrival over c at the source, with the minus sign.
template <int NEQ>
GA_HD ga_arrival rd3_arrival_row(const RdFanArgs &A, const double *y, int ray, int bnc, double incl, double az,
                                 double tt, double att, double rmax, double c0, double th, double ph)
{
    ga_arrival R;
    R.launch_idx = ray;
    R.n_bounce = bnc;
    R.incl = incl;
    R.az = az;
    RdPt P, G, S;
    rd_point(A.M, y[2], y[0], y[1], P);
    rd_point(A.M, rd_ground(A.M, y[0], y[1]), y[0], y[1], G);
    rd_point(A.M, A.src0, A.src1, A.src2, S);
    R.incl_arr = -asin(G.c / S.c * y[5]) * 180.0 / PI;
    double baz = 90.0 - atan2(-y[4], -y[3]) * 180.0 / PI;
    if (baz < -180.0) baz += 360.0;
    else if (baz > 180.0) baz -= 360.0;
    R.back_az = baz;
    R.pos0 = y[0];
    R.pos1 = y[1];
    const double dx = y[0] - A.src1, dy = y[1] - A.src2;
    R.cel = sqrt(dx * dx + dy * dy) / tt;
    R.time = tt;
    R.turn_ht = rmax;
    R.absorption = -2.0 * att;
    R.tl = (NEQ == 18) ? rd3_amp_db(y, P, S, c0, th, ph) : 0.0;
    return R;
}

// Raypath columns of a 3D row (main.infraga.3d.rngdep.cpp:454-456): x, y
// and max(z, terrain).
GA_HD void rd3_path_xyz(const RdMed &M, const double *y, ga_path_pt &R)
{
    const double zg = rd_ground(M, y[0], y[1]);
    R.pos0 = y[0];
    R.pos1 = y[1];
    R.z = y[2] > zg ? y[2] : zg;
}

// The two waveform-row records of row y (PATH_WVROWS; ga_trace.hpp): the
// state, then the Jacobian determinant, c, rho, the group speed
// |c nu/|nu| + wind| and the travel time dt of the segment into the row.
template <bool CART>
GA_HD void rd_wv_rows_put(const PathOut &Po, PathLane &pl, const RdMed &M, const double *y, int item, int leg, int m,
                          double dt)
{
    RdPt P;
    if (CART) rd_point(M, y[2], y[0], y[1], P);
    else rd_point(M, y[0], y[1], y[2], P);
    const double nm = sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
    double g0, g1, g2;
    if (CART) { g0 = P.c * y[3] / nm + P.u; g1 = P.c * y[4] / nm + P.v; g2 = P.c * y[5] / nm + P.w; }
    else      { g0 = P.c * y[3] / nm + P.w; g1 = P.c * y[4] / nm + P.v; g2 = P.c * y[5] / nm + P.u; }
    ga_path_pt a, b;
    a.launch_idx = b.launch_idx = item;
    a.leg = b.leg = leg;
    a.step = b.step = m;
    a.kind = PT_WV_STATE;
    a.pos0 = y[0]; a.pos1 = y[1]; a.z = y[2];
    a.tl = y[3]; a.absorption = y[4]; a.time = y[5];
    b.kind = PT_WV_MEDIUM;
    b.pos0 = CART ? rd3_jacobian(y, P) : rd_jacobian(y, P);
    b.pos1 = P.c;
    b.z = P.rho;
    b.tl = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    b.absorption = dt;
    b.time = 0.0;
    path_put(Po, pl, item, a);
    path_put(Po, pl, item, b);
}

// Per-lane trace loop (the range-dependent run_prop with write_rays=false).
// BLW: instantiation for M.use_blw on a terrain (rd_medium_rhs).  ya, yb:
// the lane's two state rows (current row and the RK4 accumulator, which
// swap roles after an accepted step); ys: the RK4 stage point.  PATHS: the
// write_rays / write_caustics bookkeeping of run_prop
// (main.infraga.sph.rngdep.cpp:451-490): segments of rows 1..k-1 added one
// by one, rows m == 1 or m % 15 == 0 and Jacobian sign changes written to
// A.path; PATH_WVROWS (18 equations) writes every row 0..k of every leg as
// waveform-row records (run_wnl_wvfrm, main.infraga.sph.rngdep.cpp:1108).
template <int NEQ, bool BLW, bool PATHS = false, bool CART = false, class V, class NextRay>
GA_HD void rd_lane_loop(const RdFanArgs &A, int lane, NextRay next_ray, V ya, V yb, V ys)
{
    const Limits L = A.L;
    const RdMed &M = A.M;
    V y = ya, y1 = yb;
    double *ypl = A.ypbuf + (size_t)lane * 18;
    double s = 0.0, tt = 0.0, att = 0.0, ttl = 0.0, attl = 0.0, rmax = 0.0;
    double c0 = 0.0, nu0s = 0.0, th = 0.0, ph = 0.0;
    int k = 0, bnc = 0, nst = 0, ray = 0, narr = 0;
    bool fetch = true;
    PathLane pl;
    double d_prev = 0.0;
    if (PATHS) path_reset(pl);

    for (;;) {
        if (fetch) {
            ray = next_ray();
            if (ray >= A.n_rays) break;
            ph = PI / 2.0 - A.launches[2 * ray + 1] * (PI / 180.0);
            th = A.launches[2 * ray] * (PI / 180.0);
            double y0[NEQ];
            if (CART) rd3_init_ray<NEQ>(A, th, ph, y0, c0, nu0s);
            else rd_init_ray<NEQ>(A, th, ph, y0, c0, nu0s);
#pragma unroll
            for (int i = 0; i < NEQ; ++i) y[i] = y0[i];
            s = 0.0; tt = 0.0; att = 0.0; ttl = 0.0; attl = 0.0;
            // the stratified 3D driver starts the turning height at the
            // profile bottom (main.infraga.3d.cpp:426)
            rmax = (CART && M.strat) ? M.z0 : 0.0;
            k = 0; bnc = 0; nst = 0; narr = 0;
            fetch = false;
            if constexpr (PATHS && NEQ == 18) {
                if (A.path.flags & PATH_WVROWS) rd_wv_rows_put<CART>(A.path, pl, M, y0, ray, 0, 0, 0.0);
            }
        }
        if (CART) s += rd3_rk4_step<NEQ, BLW>(M, L, y, y1, ys);
        else s += rd_rk4_step<NEQ, BLW>(M, L, y, y1, ys);
        ++k;
        ++nst;
        double n[NEQ];   // the new row
#pragma unroll
        for (int i = 0; i < NEQ; ++i) n[i] = y1[i];
        bool brk = false, end = false;
        if (s > L.s_max) { brk = true; end = true; }
        else if (CART ? rd3_leaves_region(A, L, n) : rd_leaves_region(A, L, n)) { brk = true; end = true; }
        else if (CART ? rd3_below_ground(M, n) : rd_below_ground(M, n)) { end = true; }
        else if (k >= A.max_steps) { brk = true; end = true; }
        else if (!(n[0] == n[0])) { brk = true; end = true; }   // non-finite state: stop the ray
        y1[2] = n[2];   // longitude wrapped by rd_leaves_region

        double o[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) o[i] = y[i];
        if (PATHS) {
            const bool wv = NEQ == 18 && (A.path.flags & PATH_WVROWS);
            double dt = 0.0, da = 0.0;
            if (!end || wv) {
                if (CART) rd3_seg_book(M, o, n, A.freq, dt, da);
                else rd_seg_book(M, o, n, A.freq, dt, da);
            }
            if constexpr (NEQ == 18) {
                if (wv) rd_wv_rows_put<CART>(A.path, pl, M, n, ray, bnc, k, dt);
            }
            if (!end) {
                tt += dt;
                att += da;
                const bool caus = NEQ == 18 && (A.path.flags & GA_PATH_CAUSTICS);
                const bool samp = (A.path.flags & GA_PATH_RAYS) && (k == 1 || k % 15 == 0);
                if (caus || samp) {
                    RdPt P;
                    if (CART) rd_point(M, n[2], n[0], n[1], P);
                    else rd_point(M, n[0], n[1], n[2], P);
                    if (samp) {
                        double tl = 0.0;
                        if (NEQ == 18) {
                            RdPt S;
                            rd_point(M, A.src0, A.src1, A.src2, S);
                            tl = CART ? rd3_amp_db(n, P, S, c0, th, ph) : rd_amp_db(n, P, S, c0, nu0s, th, ph);
                        }
                        ga_path_pt R;
                        R.launch_idx = ray; R.leg = bnc; R.step = k; R.kind = GA_PT_RAY;
                        R.pos0 = n[1] * (180.0 / PI); R.pos1 = n[2] * (180.0 / PI); R.z = n[0] - M.r0;
                        R.tl = tl; R.absorption = -2.0 * att; R.time = tt;
                        if (CART) rd3_path_xyz(M, n, R);
                        path_put(A.path, pl, ray, R);
                    }
                    if (caus) {
                        const double d = CART ? rd3_jacobian(n, P) : rd_jacobian(n, P);
                        if (k == 1) d_prev = d;
                        if (d * d_prev < 0.0) {
                            ga_path_pt R;
                            R.launch_idx = ray; R.leg = bnc; R.step = k; R.kind = GA_PT_CAUSTIC;
                            R.pos0 = n[1] * (180.0 / PI); R.pos1 = n[2] * (180.0 / PI); R.z = n[0] - M.r0;
                            R.tl = 0.0; R.absorption = -2.0 * att; R.time = tt;
                            if (CART) {
                                // caustic rows keep the raw z (main.infraga.3d.rngdep.cpp:467-471);
                                // the stratified driver floors it at the terrain (main.infraga.3d.cpp:469-472)
                                rd3_path_xyz(M, n, R);
                                if (!M.strat) R.z = n[2];
                            }
                            path_put(A.path, pl, ray, R);
                        }
                        d_prev = d;
                    }
                }
            }
        } else {
            double dt, da;
            if (CART) rd3_seg_book(M, o, n, A.freq, dt, da);
            else rd_seg_book(M, o, n, A.freq, dt, da);
            ttl += dt;
            attl += da;
        }
        const double ho = CART ? o[2] : o[0] - M.r0;
        rmax = fmax(rmax, ho);

        if (!end) {
            if (ho < L.z_keep) {
#pragma unroll
                for (int i = 0; i < NEQ; ++i) ypl[i] = y[i];
            }
            const V t = y;
            y = y1;
            y1 = t;
            continue;
        }
        tt += ttl;
        att += attl;
        ttl = 0.0;
        attl = 0.0;
        bool done = false;
        if (!M.is_topo && rmax < M.z0 + A.turn_ht_min) done = true;
        // the 3D range-dependent driver has no travel-time limit
        // (main.infraga.3d.rngdep.cpp:498)
        else if (brk || k < 2 || (!CART && (tt / 3600.0) > L.time_max)) done = true;
        else {
            {
                double *r = A.raw + ((size_t)ray * (A.bounces + 1) + bnc) * RAW;
#pragma unroll
                for (int i = 0; i < NEQ; ++i) r[i] = n[i];
                r[18] = tt; r[19] = att; r[20] = rmax; r[21] = th; r[22] = ph; r[23] = c0; r[24] = nu0s;
            }
            ++narr;
            if (bnc == A.bounces) done = true;
            else {
                double yr[NEQ], yp[NEQ], yk[NEQ];
#pragma unroll
                for (int i = 0; i < NEQ; ++i) {
                    yp[i] = ypl[i];
                    yk[i] = y[i];
                }
                if (CART) rd3_reflect<NEQ>(M, yp, yk, yr, c0);
                else rd_reflect<NEQ>(M, yp, yk, yr, c0);
#pragma unroll
                for (int i = 0; i < NEQ; ++i) y[i] = yr[i];
                ++bnc;
                s = 0.0;
                k = 0;
                if constexpr (PATHS && NEQ == 18) {
                    if (A.path.flags & PATH_WVROWS) rd_wv_rows_put<CART>(A.path, pl, M, yr, ray, bnc, 0, 0.0);
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

}  // namespace ga
