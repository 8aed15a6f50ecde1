// ga_rdeig.hpp -- ray traces of the batched eigenray search over the
// range-dependent media with terrain: spherical (infraga-sph-rngdep
// -eig_search / -eig_direct: geoac.eigenray.sph.cpp over
// geoac.eqset.sph.rngdep.cpp) and, with CART, 3D Cartesian
// (infraga-3d-rngdep: geoac.eigenray.3d.cpp over geoac.eqset.3d.rngdep.cpp).
//
// The jobs are those of ga_eig.hpp (EIG_SWEEP, EIG_LM, EIG_FINAL).  An
// LM / FINAL result also carries what find_eigenray's Levenberg-Marquardt
// update reads at the arrival row K besides the row itself
// (geoac.eigenray.sph.cpp:335-345, geoac.eigenray.3d.cpp:334-342): the
// terrain topo::z (sph: a radius) and its slopes topo::dz under row K, and
// atmo::c at that ground point.
#pragma once

#include "ga_eig.hpp"
#include "ga_rdtrace.hpp"

namespace ga {

struct RdEigArgs {
    RdFanArgs F;               // medium, limits, source (r, lat, lon), freq, max_steps, ypbuf, queue, path
    const EigJob *jobs;
    int n_jobs;
    int stride;                // doubles per result
    double *res;
    int32_t *brk_leg;
    int32_t *steps;
};

template <int NEQ, bool BLW, bool PATHS, bool CART, class V, class NextJob>
GA_HD void rd_eig_lane_loop(const RdEigArgs &E, int lane, NextJob next_job, V ya, V yb, V ys)
{
    const RdFanArgs &A = E.F;
    const Limits L = A.L;
    const RdMed &M = A.M;
    V y = ya, y1 = yb;
    double *ypl = A.ypbuf + (size_t)lane * 18;
    double s = 0.0, tt = 0.0, att = 0.0, rmax = 0.0, c0 = 0.0, nu0s = 0.0, th = 0.0, ph = 0.0;
    int k = 0, leg = 0, nst = 0, job = 0, legs = 0, mode = 0, brk_leg = 0;
    bool fetch = true;
    PathLane pl;
    if (PATHS) path_reset(pl);

    for (;;) {
        if (fetch) {
            job = next_job();
            if (job >= E.n_jobs) break;
            const EigJob J = E.jobs[job];
            th = J.th;
            ph = J.ph;
            double y0[NEQ];
            if (CART) rd3_init_ray<NEQ>(A, th, ph, y0, c0, nu0s);
            else rd_init_ray<NEQ>(A, th, ph, y0, c0, nu0s);
#pragma unroll
            for (int i = 0; i < NEQ; ++i) y[i] = y0[i];
            legs = J.legs;
            mode = J.mode;
            s = 0.0; tt = 0.0; att = 0.0; rmax = 0.0;
            k = 0; leg = 0; nst = 0;
            brk_leg = legs + 1;
            fetch = false;
        }
        if (CART) s += rd3_rk4_step<NEQ, BLW>(M, L, y, y1, ys);
        else s += rd_rk4_step<NEQ, BLW>(M, L, y, y1, ys);
        ++k;
        ++nst;
        double n[NEQ];
#pragma unroll
        for (int i = 0; i < NEQ; ++i) n[i] = y1[i];
        bool brk = false, end = false;
        if (s > L.s_max) { brk = true; end = true; }
        else if (CART ? rd3_leaves_region(A, L, n) : rd_leaves_region(A, L, n)) { brk = true; end = true; }
        else if (CART ? rd3_below_ground(M, n) : rd_below_ground(M, n)) { end = true; }
        else if (k >= A.max_steps) { brk = true; end = true; }
        else if (!(n[0] == n[0])) { brk = true; end = true; }
        y1[2] = n[2];

        if (!end) {
            if (NEQ == 18 && mode == EIG_FINAL) {
                // rows 1..K-1 of every leg (geoac.eigenray.sph.cpp:246-270,
                // geoac.eigenray.3d.cpp:237-268)
                double o[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) o[i] = y[i];
                double dt, da;
                if (CART) rd3_seg_book(M, o, n, A.freq, dt, da);
                else rd_seg_book(M, o, n, A.freq, dt, da);
                tt += dt;
                att += da;
                rmax = fmax(rmax, CART ? n[2] : n[0] - M.r0);
                if (PATHS && (k == 1 || k % 15 == 0)) {
                    RdPt P, S;
                    if (CART) rd_point(M, n[2], n[0], n[1], P);
                    else rd_point(M, n[0], n[1], n[2], P);
                    rd_point(M, A.src0, A.src1, A.src2, S);
                    ga_path_pt R;
                    R.launch_idx = job; R.leg = leg; R.step = k; R.kind = GA_PT_RAY;
                    if (CART) {
                        // x, y and z, floored at the terrain on the first leg
                        // only (3d.cpp:243-245 and 261-263)
                        rd3_path_xyz(M, n, R);
                        if (leg > 0) R.z = n[2];
                        R.tl = rd3_amp_db(n, P, S, c0, th, ph);
                    } else {
                        R.pos0 = n[1] * (180.0 / PI); R.pos1 = n[2] * (180.0 / PI); R.z = n[0] - M.r0;
                        R.tl = rd_amp_db(n, P, S, c0, nu0s, th, ph);
                    }
                    R.absorption = -2.0 * att; R.time = tt;
                    path_put(A.path, pl, job, R);
                }
            }
            if ((CART ? y[2] : y[0] - M.r0) < L.z_keep) {
#pragma unroll
                for (int i = 0; i < NEQ; ++i) ypl[i] = y[i];
            }
            const V t = y;
            y = y1;
            y1 = t;
            continue;
        }
        double *r = E.res + (size_t)job * E.stride;
        if (NEQ == 6) {
            r[2 * leg] = CART ? n[0] : n[1];
            r[2 * leg + 1] = CART ? n[1] : n[2];
        }
        if (brk && brk_leg > legs) brk_leg = leg;
        bool stop = (leg == legs) || (brk && mode != EIG_FINAL);
        if (!stop && k < 2) {
            if (brk_leg > legs) brk_leg = leg + 1;
            stop = true;
        }
        if (stop) {
            if (NEQ == 18) {
#pragma unroll
                for (int i = 0; i < 18; ++i) r[i] = (i < NEQ) ? n[i] : 0.0;
                r[18] = tt; r[19] = att; r[20] = rmax;
                const double h0 = CART ? n[0] : n[1], h1 = CART ? n[1] : n[2];
                Terr T;
                rd_terrain(M, h0, h1, T);
                RdPt G;
                rd_point(M, T.zg, h0, h1, G);
                r[21] = T.zg; r[22] = T.d[0]; r[23] = T.d[1]; r[24] = G.c;
            }
            E.brk_leg[job] = brk_leg;
            E.steps[job] = nst;
            if (PATHS) path_close(A.path, pl, job);
            fetch = true;
            continue;
        }
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
        ++leg;
        s = 0.0;
        k = 0;
    }
}

// Arrival row of an accepted eigenray (geoac.eigenray.sph.cpp:273-319):
// the inclination's divisor is atmo::c at the source altitude read as a
// radius (the bottom of the grid at the source's latitude and longitude);
// inclination and back azimuth are the eigenray source's expressions for
// either spherical binary.
GA_HD ga_arrival rd_eig_arrival_row(const RdEigArgs &E, const EigJob &J, const double *r)
{
    const RdFanArgs &A = E.F;
    double y0[18], c0, nu0s;
    rd_init_ray<18>(A, J.th, J.ph, y0, c0, nu0s);
    double y[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) y[i] = r[i];
    ga_arrival R = rd_arrival_row<18>(A, y, 0, J.legs, J.th * (180.0 / PI), 90.0 - J.ph * (180.0 / PI), r[18], r[19],
                                      r[20], c0, nu0s, J.th, J.ph);
    RdPt P, S;
    rd_point(A.M, y[0], y[1], y[2], P);
    rd_point(A.M, A.src0 - A.M.r0, A.src1, A.src2, S);
    R.incl_arr = -asin(P.c / S.c * y[3]) * 180.0 / PI;
    double baz = 90.0 - atan2(-y[4], -y[5]) * (180.0 / PI);   // :283-289, whatever the driver
    while (baz < -180.0) baz += 360.0;
    while (baz > 180.0) baz -= 360.0;
    R.back_az = baz;
    return R;
}

// Arrival row of an accepted 3D eigenray (geoac.eigenray.3d.cpp:272-319): c
// at row K itself over c at the source for the inclination, the horizontal
// distance from the source for the celerity.
GA_HD ga_arrival rd3_eig_arrival_row(const RdEigArgs &E, const EigJob &J, const double *r)
{
    const RdFanArgs &A = E.F;
    double y[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) y[i] = r[i];
    ga_arrival R;
    R.launch_idx = 0;
    R.n_bounce = J.legs;
    R.incl = J.th * (180.0 / PI);
    R.az = 90.0 - J.ph * (180.0 / PI);
    RdPt P, S;
    rd_point(A.M, y[2], y[0], y[1], P);
    rd_point(A.M, A.src0, A.src1, A.src2, S);
    R.pos0 = y[0];
    R.pos1 = y[1];
    R.time = r[18];
    const double dx = y[0] - A.src1, dy = y[1] - A.src2;
    R.cel = sqrt(dx * dx + dy * dy) / r[18];
    R.turn_ht = r[20];
    R.incl_arr = -asin(P.c / S.c * y[5]) * (180.0 / PI);
    double baz = 90.0 - atan2(-y[4], -y[3]) * (180.0 / PI);
    while (baz > 180.0) baz -= 360.0;
    while (baz < -180.0) baz += 360.0;
    R.back_az = baz;
    R.tl = rd3_amp_db(y, P, S, S.c, J.th, J.ph);
    R.absorption = -2.0 * r[19];
    return R;
}

}  // namespace ga
