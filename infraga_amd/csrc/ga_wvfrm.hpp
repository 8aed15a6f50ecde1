// ga_wvfrm.hpp -- weakly non-linear waveform evolution along traced rays
// (geoac::wnl_wvfrm, geoac.eqset.sph.strat.cpp:853-1148 and
// geoac.eqset.3d.strat.cpp; driver run_wnl_wvfrm, main.infraga.sph.cpp:
// 1142-1479).
//
// The waveform u(t) (scaled by its peak p0) obeys Burgers' equation along
// the ray, u_s = beta(s) (u^2)_t / 2 written in the frequency domain:
// U_s(f) = i pi f beta FFT(u^2), plus the Sutherland-Bass absorption and two
// low-pass filters per step.  A step of length ds is Heun's method with
// beta at both ends, bc and bn:
//   U += i pi f [ ds (bc + bn) / 2 A + ds^2 / 2 bc bn B + ds^3 / 8 bc^2 bn C ] F1
//   U *= 10^(-c / c0 alpha(f) ds / 10) F2
// with A = FFT(u^2), B = FFT(u d(u^2)/dt), C = FFT((d(u^2)/dt)^2), and
// u = IFFT(U) Tukey / len afterwards.  ds = ds_wvfrm / (pi max_f(f |U|) bc),
// within [1e-7 (sph) or 1e-5 (3D), 1] km; where the Jacobian determinant D
// changes sign within a step (a caustic) U is turned by pi / 2 first, and
// where |D| < 0.5 beta is taken from the caustic form with dD/ds.
//
// MI355X layout: one workgroup of WV_T threads per waveform, the transform
// in LDS (two arrays of len doubles, 128 KB at the largest length 8192),
// radix-2 decimation in time over bit-reversed input with the twiddles of
// a host-built table, three stages fused per LDS round trip.  Each thread keeps the spectrum U of its bins
// (b = tid + j WV_T <= len / 2) in registers across the steps; u lives in
// LDS between steps.  A step is three complex transforms of length len:
// Z1 = u^2 + i u d(u^2)/dt gives A and B by the even/odd split, Z2 =
// (d(u^2)/dt)^2 gives C, and the Hermitian U the new u.  The step scalars
// (beta, ds, the ray position) are computed by every thread from the leg's
// rows (uniform across the workgroup); the height-dependent factors of the
// absorption by the last wave into LDS, the frequency-dependent ones per bin.
// The rows of a leg are the linear splines of the reference: arc length s,
// ray position (3), D, beta, beta_caustic (WV_ROW doubles per row).
#pragma once

#include "ga_device.hpp"
#include "ga_rdtrace.hpp"
#include "../../include/infraga_gpu.h"

namespace ga {

constexpr int WV_T = 1024;
constexpr int WV_MIN_LEN = 64;
constexpr int WV_MAX_LEN = 8192;
constexpr int WV_BPT = WV_MAX_LEN / 2 / WV_T + 1;   // bins per thread (0..len/2)
constexpr int WV_SPT = WV_MAX_LEN / WV_T;           // samples per thread
constexpr int WV_ROW = 8;                            // s, pos[3], D, beta, beta_c, pad
constexpr int WV_F_S = 0, WV_F_POS = 1, WV_F_D = 4, WV_F_BETA = 5, WV_F_BETAC = 6;

// Per-waveform set-up results of wv_prep_kernel.
struct WvRay {
    double c0, rho0, nu0, cg0, D0;   // at the reference point (run_wnl_wvfrm, main.infraga.sph.cpp:1370-1379)
    double tref;                     // travel time to the reference row
    double tt;                       // travel time of the whole ray (rows 1..K-1 of every leg)
    double scale;                    // wvfrm_out.dat factor at row K of the last leg (:1440-1461)
    int32_t started;                 // 1: the reference point was reached
    int32_t pad;
};

constexpr int WV_PREP_T = 256;
constexpr int WV_MAX_SLOTS = 256;    // legs per waveform (bounces + 1)

// The medium the absorption is evaluated in: a stratified profile, or a
// range-dependent one on the spherical or the Cartesian grid.
constexpr int WV_STRAT = 0, WV_RD_SPH = 1, WV_RD_CART = 2, WV_2D = 3;

struct WvArgs {
    Medium M;                    // WV_STRAT
    RdMed RM;                    // WV_RD_SPH, WV_RD_CART
    double ds_floor;             // smallest evolution step [km]
    int strict;                  // caustic test: sph and 2D strict, 3D inclusive
    int len, lg, n_wv;
    double ds_wv;                // wvfrm_ds
    const double *tw;            // [len/2][2]: cos, sin of -2 pi j / len
    const double *t;             // [len] time axis [s]
    const double *tukey;         // [len] window
    const double *filt;          // [len/2 + 1][2]: filter1, filter2
    const double *u0;            // [len] initial scaled waveform
    double *u;                   // [n_wv][len] evolved scaled waveforms
    int slots;                   // legs per waveform (bounces + 1)
    const int64_t *leg_first;    // [n_wv][slots]: first row of each evolved leg
    const int32_t *leg_n;        // [n_wv][slots]: its rows (< 2: nothing to evolve)
    const double *rows;          // [n_rows][WV_ROW]
    const struct WvRay *ray;     // [n_wv]: reference-point quantities (c0)
    long long *steps;            // [n_wv]: evolution steps
    // wvfrm_out_step dumps (the DUMP instantiation): the state at the top
    // of a step whose previous step size carries the total ray length
    // across a multiple of out_step (geoac.eqset.sph.strat.cpp:989-1002)
    double out_step;
    double ref;                  // wvfrm_ref: the total length before the first leg
    double *snap;                // [n_wv][snap_cap][1 + len + 2 (len/2 + 1)]
    int snap_cap;
    int *snap_n;                 // [n_wv]: dumps reached (may exceed snap_cap)
};

// Segment of the leg's abscissa holding x at or after segment k (queries
// move forward along the leg): the lowest k' >= k with x <= s[k' + 1], as
// the reference's find_segment picks for increasing x.
GA_HD int wv_seek(const double *R, int n, int k, double x)
{
    while (k < n - 2 && x > R[(size_t)(k + 1) * WV_ROW]) ++k;
    return k;
}

// Linear spline of field f at x in segment k (interp::eval_f of the linear
// spline, interpolation.cpp:61-89: f[k] + (x - x[k]) (f[k+1] - f[k]) /
// (x[k+1] - x[k])).
GA_HD double wv_lin(const double *R, int k, int f, double x)
{
    const double *a = R + (size_t)k * WV_ROW, *b = a + WV_ROW;
    return a[f] + (x - a[0]) * ((b[f] - a[f]) / (b[0] - a[0]));
}

GA_HD double wv_beta(const double *R, int k, double x)
{
    return fabs(wv_lin(R, k, WV_F_D, x)) >= 0.5 ? wv_lin(R, k, WV_F_BETA, x) : wv_lin(R, k, WV_F_BETAC, x);
}

// LDS slot of element i: one pad double after every 128, so that the
// bit-reversed scatter of the transforms' inputs (consecutive lanes writing
// elements len/64 apart) and the fused passes' strided accesses spread over
// the banks instead of landing on one.
constexpr int WV_LDS = WV_MAX_LEN + WV_MAX_LEN / 128;
__device__ inline int wv_ix(int i) { return i + (i >> 7); }

__device__ inline int wv_rev(int i, int lg) { return (int)(__brev((unsigned int)i) >> (32 - lg)); }

// Stages st .. st + RL - 1 of the in-place radix-2 decimation-in-time
// transform fused in registers: each thread loads the 2^RL elements
// i0 + m h (h = 2^(st-1) the first fused stage's half size) that those
// stages combine, applies their butterflies with the stage-by-stage
// twiddles (element pair (m, m + 2^q) of stage st + q takes W^k, k = the
// lower bits k0 + (m mod 2^q) h, from the table at k << (lg - st - q)) and
// stores them back: the same arithmetic as one LDS round trip per stage,
// with one per RL stages.  Ends with a barrier.
template <int RL>
__device__ void wv_fft_pass(double *re, double *im, int len, int lg, int st, const double *tw, bool inv, int tid)
{
    constexpr int R = 1 << RL;
    const int h = 1 << (st - 1);
    const int groups = len >> RL;
    for (int g = tid; g < groups; g += WV_T) {
        const int k0 = g & (h - 1);
        const int i0 = ((g >> (st - 1)) << (st - 1 + RL)) + k0;
        double xr[R], xi[R];
#pragma unroll
        for (int m = 0; m < R; ++m) {
            xr[m] = re[wv_ix(i0 + m * h)];
            xi[m] = im[wv_ix(i0 + m * h)];
        }
#pragma unroll
        for (int q = 0; q < RL; ++q) {
            const int sh = lg - (st + q);
#pragma unroll
            for (int m = 0; m < R; ++m) {
                if (m & (1 << q)) continue;
                const int mb = m + (1 << q);
                const int k = k0 + (m & ((1 << q) - 1)) * h;
                const double wr = tw[2 * (k << sh)];
                const double wi = inv ? -tw[2 * (k << sh) + 1] : tw[2 * (k << sh) + 1];
                const double tr = xr[mb] * wr - xi[mb] * wi, ti = xr[mb] * wi + xi[mb] * wr;
                xr[mb] = xr[m] - tr;
                xi[mb] = xi[m] - ti;
                xr[m] += tr;
                xi[m] += ti;
            }
        }
#pragma unroll
        for (int m = 0; m < R; ++m) {
            re[wv_ix(i0 + m * h)] = xr[m];
            im[wv_ix(i0 + m * h)] = xi[m];
        }
    }
    __syncthreads();
}

// In-place transform over LDS: input in bit-reversed order, output in
// natural order; inverse: the conjugate twiddles (FFTW's sign convention,
// unnormalised).  Passes of three fused stages, then one of two or one.
__device__ void wv_fft(double *re, double *im, int len, int lg, const double *tw, bool inv, int tid)
{
    int st = 1;
    for (; st + 2 <= lg; st += 3) wv_fft_pass<3>(re, im, len, lg, st, tw, inv, tid);
    if (st + 1 <= lg) wv_fft_pass<2>(re, im, len, lg, st, tw, inv, tid);
    else if (st <= lg) wv_fft_pass<1>(re, im, len, lg, st, tw, inv, tid);
}

// Largest v over the workgroup (red: WV_T / 64 doubles of LDS).
__device__ double wv_block_max(double v, double *red, int tid)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int i = 1; i < WV_T / 64; ++i) r = fmax(r, red[i]);
    __syncthreads();
    return r;
}

// Point factors of the absorption shared by all bins of a step, held in LDS.
struct WvPoint {
    SbGas g;
    SbPoint P;
    double cpt;        // c at the ray position
};

template <bool SPH, int MED, bool DUMP = false>
__global__ void __launch_bounds__(WV_T) wv_evolve_kernel(WvArgs A)
{
    __shared__ double re[WV_LDS], im[WV_LDS];
    __shared__ double red[WV_T / 64];
    __shared__ WvPoint pt;
    const int tid0 = threadIdx.x, w = blockIdx.x;
    const int tid = tid0;
    if (w >= A.n_wv) return;                      // grid == n_wv: uniform per workgroup
    const int len = A.len, lg = A.lg, half = len >> 1;
    const double *t = A.t;
    const double dt = t[1] - t[0];
    const double fs = 1.0 / dt;
    const double ilen = 1.0 / len;                // exact: len is a power of two
    const double c0 = A.ray[w].c0;
    const Medium &M = A.M;

    for (int m = tid; m < len; m += WV_T) re[wv_ix(m)] = A.u0[m];
    __syncthreads();

    double Ur[WV_BPT], Ui[WV_BPT];
#pragma unroll
    for (int j = 0; j < WV_BPT; ++j) { Ur[j] = 0.0; Ui[j] = 0.0; }
    long long nstep = 0;
    int curz = 0;                                 // spline cursor of the absorption point
    double ray_len = 0.0;                         // the drivers' ray_length before each leg
    bool begun = false;
    int nsnap = 0;

    for (int leg = 0; leg < A.slots; ++leg) {
        const double *R = A.rows + (size_t)A.leg_first[(size_t)w * A.slots + leg] * WV_ROW;
        const int n = A.leg_n[(size_t)w * A.slots + leg];
        if (n < 2) continue;
        const double s_tot = R[(size_t)(n - 1) * WV_ROW];
        // s_prev of the leg: wvfrm_ref from the reference point, then the
        // legs' accumulated length (main.infraga.sph.cpp:1394-1399)
        const double sp = begun ? ray_len : A.ref;
        ray_len += s_tot;
        begun = true;

        // U = FFT(u): u from LDS (natural order) to bit-reversed order
        {
            double us[WV_SPT];
#pragma unroll
            for (int j = 0; j < WV_SPT; ++j) {
                const int m = tid + j * WV_T;
                us[j] = m < len ? re[wv_ix(m)] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < WV_SPT; ++j) {
                const int m = tid + j * WV_T;
                if (m < len) {
                    const int r = wv_rev(m, lg);
                    re[wv_ix(r)] = us[j];
                    im[wv_ix(r)] = 0.0;
                }
            }
            __syncthreads();
            wv_fft(re, im, len, lg, A.tw, false, tid);
            double fm = 0.0;
#pragma unroll
            for (int j = 0; j < WV_BPT; ++j) {
                const int b = tid + j * WV_T;
                if (b <= half) {
                    Ur[j] = re[wv_ix(b)];
                    Ui[j] = im[wv_ix(b)];
                    fm = fmax(fm, (fs * b * ilen) * sqrt(Ur[j] * Ur[j] + Ui[j] * Ui[j]));
                }
            }
            __syncthreads();
            // u back to natural order for the first step
#pragma unroll
            for (int j = 0; j < WV_SPT; ++j) {
                const int m = tid + j * WV_T;
                if (m < len) re[wv_ix(m)] = us[j];
            }
            double Uf_max = wv_block_max(fm, red, tid);   // its barriers publish u
            // step loop
            double s = 0.0;
            int ks = 0;
            double ds_prev = A.ds_wv;
            while (s < s_tot) {
                // the lane index passes through an empty asm each step, so the
                // per-lane addresses and frequencies below are recomputed
                // (a few instructions) instead of being kept live across the
                // loop, which spills at 1024 threads
                int tid = tid0;
                asm volatile("" : "+v"(tid));
                if (DUMP) {
                    // u (LDS, natural order, published) and U (registers)
                    // before the step; label out_step ceil((s + s_prev) / out_step)
                    const double os = A.out_step;
                    const double lab = os * ceil((s + sp) / os);
                    if (fabs(os * ceil((s + sp + ds_prev) / os) - lab) > 0.0) {
                        if (nsnap < A.snap_cap) {
                            double *S = A.snap + ((size_t)w * A.snap_cap + nsnap) * (size_t)(1 + len + 2 * (half + 1));
                            if (tid == 0) S[0] = lab;
                            for (int m = tid; m < len; m += WV_T) S[1 + m] = re[wv_ix(m)];
#pragma unroll
                            for (int j = 0; j < WV_BPT; ++j) {
                                const int b = tid + j * WV_T;
                                if (b <= half) {
                                    S[1 + len + 2 * b] = Ur[j];
                                    S[2 + len + 2 * b] = Ui[j];
                                }
                            }
                        }
                        ++nsnap;
                    }
                }
                ks = wv_seek(R, n, ks, s);
                const double bc = wv_beta(R, ks, s);
                double ds = A.ds_wv / (PI * Uf_max * bc);
                ds = fmin(ds, 1.0);
                ds = fmax(ds, A.ds_floor);
                const double x_last = s_tot;
                const double sn = fmin(s + ds, x_last);
                const double bn = wv_beta(R, wv_seek(R, n, ks, sn), sn);
                bool caustic = false;
                if (A.strict ? (s + ds < s_tot) : (s + ds <= s_tot)) {
                    const double d0 = wv_lin(R, ks, WV_F_D, s);
                    const double d1 = wv_lin(R, wv_seek(R, n, ks, s + ds), WV_F_D, s + ds);
                    caustic = A.strict ? (d0 * d1 < 0.0) : (d0 * d1 <= 0.0);
                }
                // absorption point factors at the ray position of s (the last
                // wave: wave 0 holds the extra Nyquist bin)
                if (tid >= WV_T - 64) {
                    const double pa = wv_lin(R, ks, WV_F_POS, s), pb = wv_lin(R, ks, WV_F_POS + 1, s);
                    const double pc = wv_lin(R, ks, WV_F_POS + 2, s);
                    double z, cpt, rho;
                    const double *sbc;
                    if (MED == WV_STRAT || MED == WV_2D) {
                        z = MED == WV_2D ? pb : (SPH ? pa - M.r0 : pc);
                        const SegPt sp = seg_point(M, clampz(M, z), curz);
                        cpt = herm_f(sp, 1);
                        rho = herm_f(sp, 7);
                        sbc = M.sbc;
                    } else {
                        RdPt Q;
                        if (MED == WV_RD_CART) rd_point(A.RM, pc, pa, pb, Q);
                        else rd_point(A.RM, pa, pb, pc, Q);
                        z = MED == WV_RD_CART ? pc : pa - A.RM.r0;
                        cpt = Q.c;
                        rho = Q.rho;
                        sbc = A.RM.sbc;
                    }
                    const SbGas g = sb_gas(cpt, rho, 0.0);
                    const SbPoint P = sb_point<SPH>(z, g, sbc);
                    if (tid == WV_T - 64) {
                        pt.g = g;
                        pt.P = P;
                        pt.cpt = cpt;
                    }
                }
                // u^2, d(u^2)/dt (centred, one-sided at the ends); Z1's
                // imaginary part u d(u^2)/dt goes to im at once (nothing
                // reads im now), its real part u^2 after the barrier
                double z1r[WV_SPT], z2[WV_SPT];
#pragma unroll
                for (int j = 0; j < WV_SPT; ++j) {
                    const int m = tid + j * WV_T;
                    z1r[j] = z2[j] = 0.0;
                    if (m < len) {
                        const double um = re[wv_ix(m)];
                        const double usq = um * um;
                        double du;
                        if (m == 0) {
                            const double u1 = re[wv_ix(1)];
                            du = (u1 * u1 - usq) / (t[1] - t[0]);
                        } else if (m == len - 1) {
                            const double ua = re[wv_ix(m - 1)];
                            du = (usq - ua * ua) / (t[m] - t[m - 1]);
                        } else {
                            const double ua = re[wv_ix(m - 1)], ub = re[wv_ix(m + 1)];
                            du = (ub * ub - ua * ua) / (t[m + 1] - t[m - 1]);
                        }
                        im[wv_ix(wv_rev(m, lg))] = um * du;
                        z1r[j] = usq;
                        z2[j] = du * du;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < WV_SPT; ++j) {
                    const int m = tid + j * WV_T;
                    if (m < len) re[wv_ix(wv_rev(m, lg))] = z1r[j];
                }
                __syncthreads();
                wv_fft(re, im, len, lg, A.tw, false, tid);
                // per bin: the caustic turn of U, the A and B terms of the
                // update (A = FFT(u^2), B = FFT(u du^2) from Z1 = A + i B),
                // and the absorption-filter factor 10^(-c/c0 alpha ds/10) F2
                const double h1 = ds * (bc + bn) / 2.0;
                const double h2 = (ds * ds) / 2.0 * bc * bn;
                const double h3 = (ds * ds * ds) / 8.0 * (bc * bc) * bn;
                double Pr[WV_BPT], Pi[WV_BPT], gf[WV_BPT];
                {
                    // 2D: the effective sound speed of the leg's last evolved
                    // row, which geoac.eqset.2d.cpp's step loop reuses from its
                    // set-up (kept in the first row's pad)
                    const double cc = -(MED == WV_2D ? R[WV_ROW - 1] : pt.cpt) / c0;
                    const double tweak = (MED == WV_STRAT || MED == WV_2D) ? M.tweak : A.RM.tweak;
#pragma unroll
                    for (int j = 0; j < WV_BPT; ++j) {
                        const int b = tid + j * WV_T;
                        Pr[j] = Pi[j] = gf[j] = 0.0;
                        if (b <= half) {
                            if (caustic) {
                                const double tmp = Ur[j];
                                Ur[j] = -Ui[j];
                                Ui[j] = tmp;
                            }
                            const int nb = (len - b) & (len - 1);
                            const double za = re[wv_ix(b)], zb = im[wv_ix(b)], zc = re[wv_ix(nb)], zd = im[wv_ix(nb)];
                            const double Ar = 0.5 * (za + zc), Ai = 0.5 * (zb - zd);
                            const double Br = 0.5 * (zb + zd), Bi = 0.5 * (zc - za);
                            Pr[j] = h1 * Ai + h2 * Bi;
                            Pi[j] = h1 * Ar + h2 * Br;
                            const double freq = fs * b * ilen;
                            SbGas g = pt.g;
                            sb_gas_freq(g, freq);
                            double Q, Rr;
                            sb_qr<SPH>(pt.P, g, freq, Q, Rr);
                            const double alpha = (1.003 * sb_cl(g, freq) * Q + Rr) * tweak * 8.685889e3;
                            gf[j] = ga_exp10(cc * alpha * ds / 10.0) * A.filt[2 * b + 1];
                        }
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < WV_SPT; ++j) {
                    const int m = tid + j * WV_T;
                    if (m < len) {
                        const int r = wv_rev(m, lg);
                        re[wv_ix(r)] = z2[j];
                        im[wv_ix(r)] = 0.0;
                    }
                }
                __syncthreads();
                wv_fft(re, im, len, lg, A.tw, false, tid);
                // the update with C = FFT((du^2)^2) = Z2
                double fm = 0.0;
#pragma unroll
                for (int j = 0; j < WV_BPT; ++j) {
                    const int b = tid + j * WV_T;
                    if (b <= half) {
                        const double freq = fs * b * ilen;
                        const double f1 = A.filt[2 * b];
                        const double Cr = re[wv_ix(b)], Ci = im[wv_ix(b)];
                        Ur[j] -= PI * freq * (Pr[j] + h3 * Ci) * f1;
                        Ui[j] += PI * freq * (Pi[j] + h3 * Cr) * f1;
                        Ur[j] *= gf[j];
                        Ui[j] *= gf[j];
                        fm = fmax(fm, freq * sqrt(Ur[j] * Ur[j] + Ui[j] * Ui[j]));
                    }
                }
                __syncthreads();
                // Hermitian spectrum (imaginary parts of bins 0 and len/2
                // ignored, as a c2r transform does) to bit-reversed order
#pragma unroll
                for (int j = 0; j < WV_BPT; ++j) {
                    const int b = tid + j * WV_T;
                    if (b <= half) {
                        const int r = wv_rev(b, lg);
                        const bool edge = (b == 0) || (b == half);
                        re[wv_ix(r)] = Ur[j];
                        im[wv_ix(r)] = edge ? 0.0 : Ui[j];
                        if (!edge) {
                            const int r2 = wv_rev(len - b, lg);
                            re[wv_ix(r2)] = Ur[j];
                            im[wv_ix(r2)] = -Ui[j];
                        }
                    }
                }
                __syncthreads();
                wv_fft(re, im, len, lg, A.tw, true, tid);
#pragma unroll
                for (int j = 0; j < WV_SPT; ++j) {
                    const int m = tid + j * WV_T;
                    if (m < len) re[wv_ix(m)] = re[wv_ix(m)] * A.tukey[m] * ilen;
                }
                Uf_max = wv_block_max(fm, red, tid);   // its barriers publish the new u
                s += ds;
                ds_prev = ds;
                ++nstep;
            }
        }
    }
    for (int m = tid; m < len; m += WV_T) A.u[(size_t)w * len + m] = re[wv_ix(m)];
    if (tid == 0) A.steps[w] = nstep;
    if (DUMP && tid == 0) A.snap_n[w] = nsnap;
}


// Set-up of every waveform from its recorded rows (records in ray order,
// two per row: PT_WV_STATE then PT_WV_MEDIUM), one workgroup per ray:
// thread 0 walks the legs as run_wnl_wvfrm does (main.infraga.sph.cpp:
// 1336-1401): the travel time of rows 1..K-1 of each leg, the reference
// point where the ray length (midpoint radius and latitude; 3D chord) first
// reaches ref, and the rows each leg evolves over (from the reference row
// on in its leg, whole legs after it, rows up to K-1), whose arc length,
// position and D it writes; then all threads fill beta and the caustic beta
// of those rows (wnl_wvfrm's set-up, geoac.eqset.sph.strat.cpp:903-941).
template <bool SPH, bool TWO_D = false>
__global__ void __launch_bounds__(WV_PREP_T) wv_prep_kernel(const ga_path_pt *rec, const int64_t *off, int n_wv,
                                                           int slots, double ref, double p0, double *rows,
                                                           int64_t *leg_first, int32_t *leg_n, WvRay *ray)
{
    __shared__ int64_t s_first[WV_MAX_SLOTS];
    __shared__ int32_t s_n[WV_MAX_SLOTS];
    __shared__ WvRay s_ray;
    __shared__ double s_beta0;
    const int w = blockIdx.x, tid = threadIdx.x;
    if (w >= n_wv) return;
    const ga_path_pt *P = rec + off[w];
    const int64_t nr = (off[w + 1] - off[w]) / 2;       // rows of this ray
    const int64_t r0 = off[w] / 2;                      // its first row in the rows array
    for (int j = tid; j < slots; j += WV_PREP_T) { s_first[j] = 0; s_n[j] = 0; }
    __syncthreads();
    if (tid == 0) {
        WvRay R;
        R.c0 = R.rho0 = R.nu0 = R.cg0 = R.D0 = R.tref = R.tt = 0.0;
        R.scale = 0.0;
        R.started = 0;
        R.pad = 0;
        auto y = [&](int64_t r, int k) -> double {
            const ga_path_pt &a = P[2 * r];
            return k == 0 ? a.pos0 : k == 1 ? a.pos1 : k == 2 ? a.z : k == 3 ? a.tl : k == 4 ? a.absorption : a.time;
        };
        auto nu = [&](int64_t r) { return sqrt(y(r, 3) * y(r, 3) + y(r, 4) * y(r, 4) + y(r, 5) * y(r, 5)); };
        double ray_length = 0.0;
        int64_t a = 0;
        while (a < nr) {
            const int leg = P[2 * a].leg;
            int64_t b = a;
            while (b + 1 < nr && P[2 * (b + 1)].leg == leg) ++b;
            const int64_t k = b - a;                     // row K of the leg
            for (int64_t m = 1; m < k; ++m) R.tt += P[2 * (a + m) + 1].absorption;
            int64_t k1 = -1;
            if (!R.started) {
                for (int64_t kr = 0; kr < k; ++kr) {
                    const int64_t p = a + kr, q = p + 1;
                    double dl;
                    if (SPH) {
                        const double dr = y(q, 0) - y(p, 0), rr = y(p, 0) + dr / 2.0;
                        const double dth = y(q, 1) - y(p, 1), th = y(p, 1) + dth / 2.0;
                        const double dph = y(q, 2) - y(p, 2);
                        dl = sqrt(dr * dr + (rr * dth) * (rr * dth) + (rr * cos(th) * dph) * (rr * cos(th) * dph));
                    } else {
                        const double dx = y(q, 0) - y(p, 0), dy = y(q, 1) - y(p, 1), dz = y(q, 2) - y(p, 2);
                        dl = sqrt(dx * dx + dy * dy + dz * dz);
                    }
                    ray_length += dl;
                    if (ray_length >= ref) {
                        const ga_path_pt &md = P[2 * q + 1];
                        R.c0 = md.pos1;
                        R.rho0 = md.z;
                        R.nu0 = nu(q);
                        R.cg0 = md.tl;
                        R.D0 = md.pos0;
                        for (int64_t m = 1; m <= kr + 1; ++m) R.tref += P[2 * (a + m) + 1].absorption;
                        R.started = 1;
                        k1 = kr + 1;
                        break;
                    }
                }
            } else {
                k1 = 0;
            }
            if (k1 >= 0 && k - k1 >= 2 && leg >= 0 && leg < slots) {
                s_first[leg] = r0 + a + k1;
                s_n[leg] = (int32_t)(k - k1);
                double sarc = 0.0;
                for (int64_t m = a + k1; m < a + k; ++m) {
                    double *o = rows + (size_t)(r0 + m) * WV_ROW;
                    o[WV_F_S] = sarc;
                    o[WV_F_POS] = y(m, 0);
                    o[WV_F_POS + 1] = y(m, 1);
                    o[WV_F_POS + 2] = y(m, 2);
                    o[WV_F_D] = P[2 * m + 1].pos0;
                    if (m + 1 < a + k) {
                        const double d0 = y(m + 1, 0) - y(m, 0), d1 = y(m + 1, 1) - y(m, 1), d2 = y(m + 1, 2) - y(m, 2);
                        if (SPH) {
                            const double e1 = y(m, 0) * d1, e2 = y(m, 0) * cos(y(m, 1)) * d2;
                            sarc += sqrt(d0 * d0 + e1 * e1 + e2 * e2);
                        } else {
                            sarc += sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                        }
                    }
                }
            }
            a = b + 1;
        }
        if (R.started && nr > 0) {
            const int64_t kl = nr - 1;                   // row K of the last leg
            const ga_path_pt &md = P[2 * kl + 1];
            const double c = md.pos1, rho = md.z, nuk = nu(kl);
            const double c_src = P[1].pos1;
            if (TWO_D) {
                // main.infraga.2d.cpp run_wnl_wvfrm: the effective sound speed,
                // no ray-direction factors
                R.scale = sqrt(fabs((rho * c) / (R.rho0 * R.c0) * R.D0 / md.pos0));
                s_beta0 = 1.2 * p0 / ((R.rho0 * 1.0e3) * ((R.c0 * 1.0e3) * (R.c0 * 1.0e3)));
            } else {
                R.scale = sqrt(fabs((rho * (c * c * c) * nuk) / (R.rho0 * (R.c0 * R.c0 * R.c0) * R.nu0) *
                                    (R.cg0 * R.D0) / (md.tl * md.pos0)));
                s_beta0 = 1.2 * p0 / ((R.rho0 * 1.0e3) * ((R.c0 * 1.0e3) * (R.c0 * 1.0e3))) * (R.nu0 * R.c0) /
                          (R.cg0 * c_src);
            }
        }
        s_ray = R;
        ray[w] = R;
    }
    __syncthreads();
    if (!s_ray.started) {
        for (int j = tid; j < slots; j += WV_PREP_T) { leg_first[(size_t)w * slots + j] = 0; leg_n[(size_t)w * slots + j] = 0; }
        return;
    }
    const WvRay R = s_ray;
    const double beta0 = s_beta0;
    const double num = R.D0 * R.rho0 * (R.cg0 * R.cg0 * R.cg0), den0 = R.c0 * (R.nu0 * R.nu0 * R.nu0);
    for (int j = 0; j < slots; ++j) {
        const int32_t nn = s_n[j];
        const int64_t f = s_first[j];
        for (int q = tid; q < nn; q += WV_PREP_T) {
            const int64_t m = f + q;                     // global row
            const ga_path_pt &md = rec[2 * m + 1];
            const ga_path_pt &st = rec[2 * m];
            const double c = md.pos1, rho = md.z;
            const double nuq = sqrt(st.tl * st.tl + st.absorption * st.absorption + st.time * st.time);
            double *o = rows + (size_t)m * WV_ROW;
            const double Dq = o[WV_F_D];
            if (TWO_D) {
                // geoac.eqset.2d.cpp wnl_wvfrm: beta0 / c0 sqrt(D0 rho0 c0 / (D rho c))
                const double n2 = R.D0 * R.rho0 * R.c0;
                o[WV_F_BETA] = beta0 / R.c0 * sqrt(fabs(n2 / (Dq * rho * c)));
                if (q > 0) {
                    const double *op = o - WV_ROW;
                    o[WV_F_BETAC] = beta0 / R.c0 * sqrt(fabs(n2 / ((Dq - op[WV_F_D]) / (o[WV_F_S] - op[WV_F_S]) * rho * c)));
                }
            } else {
                o[WV_F_BETA] = beta0 * sqrt(fabs(num / (Dq * rho * (c * c * c)) * (c * (nuq * nuq * nuq)) / den0));
                if (q > 0) {
                    const double *op = o - WV_ROW;
                    o[WV_F_BETAC] = beta0 * sqrt(fabs(num / ((Dq - op[WV_F_D]) / (o[WV_F_S] - op[WV_F_S]) * rho *
                                                             (c * c * c)) * (c * (nuq * nuq * nuq)) / den0));
                }
            }
            // the first row's pad keeps the effective sound speed of the
            // leg's last evolved row (2D absorption scale)
            if (q == 0) o[WV_ROW - 1] = rec[2 * (f + nn - 1) + 1].pos1;
        }
    }
    __syncthreads();
    for (int j = tid; j < slots; j += WV_PREP_T) {
        leg_first[(size_t)w * slots + j] = s_first[j];
        leg_n[(size_t)w * slots + j] = s_n[j];
        if (s_n[j] >= 2) {
            double *o = rows + (size_t)s_first[j] * WV_ROW;
            o[WV_F_BETAC] = o[WV_ROW + WV_F_BETAC];
        }
    }
}

}  // namespace ga
