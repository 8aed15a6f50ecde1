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
// a host-built table.  Each thread keeps the spectrum U of its bins
// (b = tid + j WV_T <= len / 2) in registers across the steps; u lives in
// LDS between steps.  A step is three complex transforms of length len:
// Z1 = u^2 + i u d(u^2)/dt gives A and B by the even/odd split, Z2 =
// (d(u^2)/dt)^2 gives C, and the Hermitian U the new u.  The step scalars
// (beta, ds, the ray position) are computed by every thread from the leg's
// rows (uniform across the workgroup); the height-dependent factors of the
// absorption by wave 0 into LDS, the frequency-dependent ones per bin.
// The rows of a leg are the linear splines of the reference: arc length s,
// ray position (3), D, beta, beta_caustic (WV_ROW doubles per row).
#pragma once

#include "ga_device.hpp"

namespace ga {

constexpr int WV_T = 1024;
constexpr int WV_MIN_LEN = 64;
constexpr int WV_MAX_LEN = 8192;
constexpr int WV_BPT = WV_MAX_LEN / 2 / WV_T + 1;   // bins per thread (0..len/2)
constexpr int WV_SPT = WV_MAX_LEN / WV_T;           // samples per thread
constexpr int WV_ROW = 8;                            // s, pos[3], D, beta, beta_c, pad
constexpr int WV_F_S = 0, WV_F_POS = 1, WV_F_D = 4, WV_F_BETA = 5, WV_F_BETAC = 6;

struct WvArgs {
    Medium M;
    int len, lg, n_wv;
    double ds_wv;                // wvfrm_ds
    const double *tw;            // [len/2][2]: cos, sin of -2 pi j / len
    const double *t;             // [len] time axis [s]
    const double *tukey;         // [len] window
    const double *filt;          // [len/2 + 1][2]: filter1, filter2
    const double *u0;            // [len] initial scaled waveform
    double *u;                   // [n_wv][len] evolved scaled waveforms
    const int32_t *wv_leg;       // [n_wv + 1]: legs of waveform w are wv_leg[w] .. wv_leg[w+1]-1
    const int64_t *leg_row;      // [n_legs + 1]: rows of leg j are leg_row[j] .. leg_row[j+1]-1
    const double *rows;          // [n_rows][WV_ROW]
    const double *c0;            // [n_wv]: c at the reference point
    long long *steps;            // [n_wv]: evolution steps
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

__device__ inline int wv_rev(int i, int lg) { return (int)(__brev((unsigned int)i) >> (32 - lg)); }

// In-place radix-2 DIT transform over LDS: input in bit-reversed order,
// output in natural order; inverse: the conjugate twiddles (FFTW's sign
// convention, unnormalised).  Ends with a barrier.
__device__ void wv_fft(double *re, double *im, int len, int lg, const double *tw, bool inv, int tid)
{
    const int nb = len >> 1;
    for (int st = 1; st <= lg; ++st) {
        const int h = 1 << (st - 1);
        const int sh = lg - st;
        for (int b = tid; b < nb; b += WV_T) {
            const int k = b & (h - 1);
            const int i = ((b >> (st - 1)) << st) | k;
            const int j = i + h;
            const double wr = tw[2 * (k << sh)];
            const double wi = inv ? -tw[2 * (k << sh) + 1] : tw[2 * (k << sh) + 1];
            const double xr = re[j], xi = im[j];
            const double tr = xr * wr - xi * wi, ti = xr * wi + xi * wr;
            const double ar = re[i], ai = im[i];
            re[j] = ar - tr;
            im[j] = ai - ti;
            re[i] = ar + tr;
            im[i] = ai + ti;
        }
        __syncthreads();
    }
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

template <bool SPH>
__global__ void __launch_bounds__(WV_T) wv_evolve_kernel(WvArgs A)
{
    __shared__ double re[WV_MAX_LEN], im[WV_MAX_LEN];
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
    const double c0 = A.c0[w];
    const Medium &M = A.M;

    for (int m = tid; m < len; m += WV_T) re[m] = A.u0[m];
    __syncthreads();

    double Ur[WV_BPT], Ui[WV_BPT];
#pragma unroll
    for (int j = 0; j < WV_BPT; ++j) { Ur[j] = 0.0; Ui[j] = 0.0; }
    long long nstep = 0;
    int curz = 0;                                 // spline cursor of the absorption point

    for (int leg = A.wv_leg[w]; leg < A.wv_leg[w + 1]; ++leg) {
        const double *R = A.rows + (size_t)A.leg_row[leg] * WV_ROW;
        const int n = (int)(A.leg_row[leg + 1] - A.leg_row[leg]);
        if (n < 2) continue;
        const double s_tot = R[(size_t)(n - 1) * WV_ROW];

        // U = FFT(u): u from LDS (natural order) to bit-reversed order
        {
            double us[WV_SPT];
#pragma unroll
            for (int j = 0; j < WV_SPT; ++j) {
                const int m = tid + j * WV_T;
                us[j] = m < len ? re[m] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < WV_SPT; ++j) {
                const int m = tid + j * WV_T;
                if (m < len) {
                    const int r = wv_rev(m, lg);
                    re[r] = us[j];
                    im[r] = 0.0;
                }
            }
            __syncthreads();
            wv_fft(re, im, len, lg, A.tw, false, tid);
            double fm = 0.0;
#pragma unroll
            for (int j = 0; j < WV_BPT; ++j) {
                const int b = tid + j * WV_T;
                if (b <= half) {
                    Ur[j] = re[b];
                    Ui[j] = im[b];
                    fm = fmax(fm, (fs * b * ilen) * sqrt(Ur[j] * Ur[j] + Ui[j] * Ui[j]));
                }
            }
            __syncthreads();
            // u back to natural order for the first step
#pragma unroll
            for (int j = 0; j < WV_SPT; ++j) {
                const int m = tid + j * WV_T;
                if (m < len) re[m] = us[j];
            }
            double Uf_max = wv_block_max(fm, red, tid);   // its barriers publish u
            // step loop
            double s = 0.0;
            int ks = 0;
            while (s < s_tot) {
                // the lane index passes through an empty asm each step, so the
                // per-lane addresses and frequencies below are recomputed
                // (a few instructions) instead of being kept live across the
                // loop, which spills at 1024 threads
                int tid = tid0;
                asm volatile("" : "+v"(tid));
                ks = wv_seek(R, n, ks, s);
                const double bc = wv_beta(R, ks, s);
                double ds = A.ds_wv / (PI * Uf_max * bc);
                ds = fmin(ds, 1.0);
                ds = fmax(ds, SPH ? 1.0e-7 : 1.0e-5);
                const double x_last = s_tot;
                const double sn = fmin(s + ds, x_last);
                const double bn = wv_beta(R, wv_seek(R, n, ks, sn), sn);
                bool caustic = false;
                if (SPH ? (s + ds < s_tot) : (s + ds <= s_tot)) {
                    const double d0 = wv_lin(R, ks, WV_F_D, s);
                    const double d1 = wv_lin(R, wv_seek(R, n, ks, s + ds), WV_F_D, s + ds);
                    caustic = SPH ? (d0 * d1 < 0.0) : (d0 * d1 <= 0.0);
                }
                // absorption point factors at the ray position of s (wave 0)
                if (tid < 64) {
                    const double pa = wv_lin(R, ks, WV_F_POS, s), pb = wv_lin(R, ks, WV_F_POS + 1, s);
                    const double pc = wv_lin(R, ks, WV_F_POS + 2, s);
                    (void)pb;
                    const double z = SPH ? pa - M.r0 : pc;
                    const SegPt sp = seg_point(M, clampz(M, z), curz);
                    const double cpt = herm_f(sp, 1), rho = herm_f(sp, 7);
                    const SbGas g = sb_gas(cpt, rho, 0.0);
                    const SbPoint P = sb_point<SPH>(z, g, M.sbc);
                    if (tid == 0) {
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
                        const double um = re[m];
                        const double usq = um * um;
                        double du;
                        if (m == 0) {
                            const double u1 = re[1];
                            du = (u1 * u1 - usq) / (t[1] - t[0]);
                        } else if (m == len - 1) {
                            const double ua = re[m - 1];
                            du = (usq - ua * ua) / (t[m] - t[m - 1]);
                        } else {
                            const double ua = re[m - 1], ub = re[m + 1];
                            du = (ub * ub - ua * ua) / (t[m + 1] - t[m - 1]);
                        }
                        im[wv_rev(m, lg)] = um * du;
                        z1r[j] = usq;
                        z2[j] = du * du;
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < WV_SPT; ++j) {
                    const int m = tid + j * WV_T;
                    if (m < len) re[wv_rev(m, lg)] = z1r[j];
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
                    const double cc = -pt.cpt / c0;
                    const double tweak = M.tweak;
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
                            const double za = re[b], zb = im[b], zc = re[nb], zd = im[nb];
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
                        re[r] = z2[j];
                        im[r] = 0.0;
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
                        const double Cr = re[b], Ci = im[b];
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
                        re[r] = Ur[j];
                        im[r] = edge ? 0.0 : Ui[j];
                        if (!edge) {
                            const int r2 = wv_rev(len - b, lg);
                            re[r2] = Ur[j];
                            im[r2] = -Ui[j];
                        }
                    }
                }
                __syncthreads();
                wv_fft(re, im, len, lg, A.tw, true, tid);
#pragma unroll
                for (int j = 0; j < WV_SPT; ++j) {
                    const int m = tid + j * WV_T;
                    if (m < len) re[m] = re[m] * A.tukey[m] * ilen;
                }
                Uf_max = wv_block_max(fm, red, tid);   // its barriers publish the new u
                s += ds;
                ++nstep;
            }
        }
    }
    for (int m = tid; m < len; m += WV_T) A.u[(size_t)w * len + m] = re[m];
    if (tid == 0) A.steps[w] = nstep;
}

}  // namespace ga
