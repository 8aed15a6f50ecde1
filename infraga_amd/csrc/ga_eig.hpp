// ga_eig.hpp -- per-lane ray traces of the batched eigenray search.
//
// The search (est_eigenray / find_eigenray, src/geoac/geoac.eigenray.3d.cpp
// and geoac.eigenray.sph.cpp) needs three kinds of ray trace, all from the
// same source:
//   EIG_SWEEP  6 equations, up to `legs` reflections, stopping at the first
//              break (est_eigenray's inclination sweep, 3d.cpp:66-76); the
//              ground position of every leg is recorded so that one trace
//              serves every bounce count up to `legs`.
//   EIG_LM     18 equations, up to `legs` reflections, stopping at a break
//              (find_eigenray's iteration trace, 3d.cpp:180-209).
//   EIG_FINAL  18 equations, `legs` reflections whatever break_check says,
//              with travel time, absorption and turning height summed over
//              rows 1..K-1 of every leg (the accepted eigenray, 3d.cpp:223-276).
// A lane takes jobs from a device queue and writes one result per job.
#pragma once

#include "ga_trace.hpp"

namespace ga {

enum EigMode : int32_t { EIG_SWEEP = 0, EIG_LM = 1, EIG_FINAL = 2 };

struct EigJob {
    double th, ph;      // launch inclination and azimuth [rad] (geoac theta, phi)
    int32_t legs;       // reflections to trace through
    int32_t mode;       // EigMode
};

// Result layout per job (stride doubles):
//   EIG_SWEEP: row K's (x, y) [3D] or (lat, lon) [sph] of legs 0..legs
//                                                     stride 2 * (legs_max + 1)
//   EIG_LM / EIG_FINAL: row K (18), tt, att, zmax, and for the
//              range-dependent medium topo::z, topo::dz (lat, lon) and
//              atmo::c at the ground under row K       stride EIG_FULL
constexpr int EIG_FULL = 25;

struct EigArgs {
    FanArgs F;               // medium, limits, source, freq, max_steps, ypbuf, queue
    const EigJob *jobs;
    int n_jobs;
    int stride;              // doubles per result
    double *res;
    int32_t *brk_leg;        // first leg ended by a break (legs + 1: none)
    int32_t *steps;          // RK4 steps of the job
};

// PATHS (EIG_FINAL jobs, 18 equations): rows m == 1 or m % 15 == 0 of every
// leg go to E.F.path with the job index as item, as find_eigenray writes
// <id>.eigenray-N.dat (geoac.eigenray.3d.cpp:240-266,
// geoac.eigenray.sph.cpp:246-273).
template <int G, int NEQ, bool PATHS = false, class NextJob>
GA_HD void eig_lane_loop(const EigArgs &E, const Medium &M, int lane, NextJob next_job)
{
    const FanArgs &A = E.F;
    const Limits L = A.L;

    double y[NEQ];
    double *ypl = A.ypbuf + (size_t)lane * 18;
    double s = 0.0, tt = 0.0, att = 0.0, zmax = 0.0, c0 = 0.0, nu0s = 0.0, th = 0.0, ph = 0.0;
    int k = 0, leg = 0, nst = 0, job = 0, cur = 0, legs = 0, mode = 0, brk_leg = 0;
    bool fetch = true;
    PathLane pl;
    if (PATHS) path_reset(pl);
#pragma unroll
    for (int i = 0; i < NEQ; ++i) y[i] = 0.0;

    for (;;) {
        if (fetch) {
            job = next_job();
            if (job >= E.n_jobs) break;
            const EigJob J = E.jobs[job];
            th = J.th;
            ph = J.ph;
            init_ray_rad<G, NEQ>(A, M, J.th, J.ph, y, c0, nu0s, cur);
            legs = J.legs;
            mode = J.mode;
            s = 0.0; tt = 0.0; att = 0.0; zmax = 0.0;
            k = 0; leg = 0; nst = 0;
            brk_leg = legs + 1;
            fetch = false;
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
        else if (!(y1[0] == y1[0])) { brk = true; end = true; }

        if (!end) {
            if (NEQ == 18 && mode == EIG_FINAL) {
                double dt, da;
                seg_book<G>(M, y, y1, A.freq, cur, dt, da);
                tt += dt;
                att += da;
                const double zz = zcoord<G>(M, y1);
                zmax = (zmax < zz) ? zz : zmax;
                if (PATHS && (k == 1 || k % 15 == 0)) {
                    double c, u, v, rho;
                    point_medium<G>(M, y1, cur, c, u, v, rho);
                    const double tl = amp_db<G>(A, M, y1, c, u, v, rho, c0, nu0s, th, ph, cur);
                    path_put(A.path, pl, job, path_row<G>(M, y1, job, leg, k, GA_PT_RAY, tl, att, tt));
                }
            }
            if (zcoord<G>(M, y) < L.z_keep) {
#pragma unroll
                for (int i = 0; i < NEQ; ++i) ypl[i] = y[i];
            }
#pragma unroll
            for (int i = 0; i < NEQ; ++i) y[i] = y1[i];
            continue;
        }
        // leg finished at row K = k (y1); rows K-1 = y, K-2 = ypl
        double *r = E.res + (size_t)job * E.stride;
        if (NEQ == 6) {
            const int o = (G == GEOM_SPH) ? 1 : 0;
            r[2 * leg] = y1[o];
            r[2 * leg + 1] = y1[o + 1];
        }
        if (brk && brk_leg > legs) brk_leg = leg;
        bool stop = (leg == legs) || (brk && mode != EIG_FINAL);
        if (!stop && k < 2) {
            // no row K-2 to reflect from: the trace ends here and every later
            // leg counts as broken
            if (brk_leg > legs) brk_leg = leg + 1;
            stop = true;
        }
        if (stop) {
            if (NEQ == 18) {
#pragma unroll
                for (int i = 0; i < 18; ++i) r[i] = (i < NEQ) ? y1[i] : 0.0;
                r[18] = tt; r[19] = att; r[20] = zmax;
            }
            E.brk_leg[job] = brk_leg;
            E.steps[job] = nst;
            if (PATHS) path_close(A.path, pl, job);
            fetch = true;
            continue;
        }
        double yr[NEQ], yp[NEQ];
#pragma unroll
        for (int i = 0; i < NEQ; ++i) yp[i] = ypl[i];
        reflect<G, NEQ>(M, yp, y, yr, c0, cur);
#pragma unroll
        for (int i = 0; i < NEQ; ++i) y[i] = yr[i];
        ++leg;
        s = 0.0;
        k = 0;
    }
}

// Arrival row of an accepted eigenray (geoac.eigenray.3d.cpp:278-315,
// geoac.eigenray.sph.cpp:273-319): the fan's row with the arrival
// inclination taken at row K's own altitude and divided by E.F.c_src.
template <int G>
GA_HD ga_arrival eig_arrival_row(const EigArgs &E, const Medium &M, const EigJob &J, const double *r, int &cur)
{
    double y0[18], c0, nu0s;
    init_ray_rad<G, 18>(E.F, M, J.th, J.ph, y0, c0, nu0s, cur);
    double y[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) y[i] = r[i];
    ga_arrival R = arrival_row<G, 18>(E.F, M, y, 0, J.legs, J.th * (180.0 / PI), 90.0 - J.ph * (180.0 / PI),
                                      r[18], r[19], r[20], c0, nu0s, J.th, J.ph, cur);
    double c, u, v, rho;
    point_medium<G>(M, y, cur, c, u, v, rho);
    if (G == GEOM_SPH) {
        R.incl_arr = -asin(c / E.F.c_src * y[3]) * 180.0 / PI;
        R.pos0 = y[1] * 180.0 / PI;
        R.pos1 = y[2] * 180.0 / PI;
    } else {
        R.incl_arr = -asin(c / E.F.c_src * y[5]) * (180.0 / PI);
    }
    return R;
}

}  // namespace ga
