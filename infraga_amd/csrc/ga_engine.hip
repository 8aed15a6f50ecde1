// ga_engine.hip -- kernels and C ABI of the MI355X stratified ray engine.
//
// Execution model: one ray per lane in a persistent grid (one workgroup per
// CU, the spline table staged once into LDS).  Each lane pulls ray indices
// from a device queue counter, integrates the ray leg by leg with its state
// in registers, accumulates travel time, absorption and turning height per
// step, writes one arrival record per ground return into the ray's own
// output slots and then pulls the next ray.  Lanes of a wave therefore work
// on rays of different lengths without waiting for each other; only the
// final drain leaves lanes idle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/infraga_gpu.h"
#include "ga_eig_search.hpp"
#include "ga_trace.hpp"

using namespace ga;

namespace {

template <int G, int NEQ, bool LDS>
__global__ void __launch_bounds__((trace_wg<G, NEQ>())) fan_kernel(FanArgs A)
{
    extern __shared__ double lds_tab[];
    Medium M = A.M;
    if constexpr (LDS) {
        const int nd = M.n * NODE;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) lds_tab[i] = A.M.tab[i];
        __syncthreads();
        M.tab = lds_tab;
    }
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    lane_loop<G, NEQ>(A, M, lane, [&]() { return (int)atomicAdd(A.queue, 1u); });
}

template <int G, int NEQ>
__global__ void __launch_bounds__(256) arrival_kernel(FanArgs A)
{
    const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = A.bounces + 1;
    if (slot >= (long long)A.n_rays * nb) return;
    const int ray = (int)(slot / nb), bnc = (int)(slot % nb);
    if (bnc >= A.counts[ray]) return;
    const double *r = A.raw + slot * RAW;
    double y[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) y[i] = (i < NEQ) ? r[i] : 0.0;
    int cur = 0;
    eval_arrival<G, NEQ>(A, A.M, y, ray, bnc, A.launches[2 * ray], A.launches[2 * ray + 1], r[18], r[19], r[20],
                         r[23], r[24], r[21], r[22], cur);
}

__global__ void reset_counter(unsigned int *q) { *q = 0u; }

// Eigenray-search traces: same persistent layout as fan_kernel, jobs instead
// of launch angles.  WGS lanes per workgroup share one LDS copy of the table;
// the 6-equation sweep kernel needs <= 110 registers and runs four waves per
// SIMD.
#ifndef GA_EIG_WG6
#define GA_EIG_WG6 1024
#endif
template <int G, int NEQ, bool LDS, int WGS>
__global__ void __launch_bounds__(WGS) eig_kernel(EigArgs E)
{
    extern __shared__ double lds_tab[];
    Medium M = E.F.M;
    if constexpr (LDS) {
        const int nd = M.n * NODE;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) lds_tab[i] = E.F.M.tab[i];
        __syncthreads();
        M.tab = lds_tab;
    }
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    eig_lane_loop<G, NEQ>(E, M, lane, [&]() { return (int)atomicAdd(E.F.queue, 1u); });
}

// Arrival rows of the EIG_FINAL jobs [first, first + n).
template <int G>
__global__ void __launch_bounds__(256) eig_arrival_kernel(EigArgs E, int first, int n, ga_arrival *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cur = 0;
    const EigJob J = E.jobs[first + i];
    out[i] = eig_arrival_row<G>(E, E.F.M, J, E.res + (size_t)(first + i) * EIG_FULL, cur);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct DevAtmo {
    double *d_tab = nullptr;
    int n = 0;
    std::vector<double> z, c, u, v, rho;
};

}  // namespace

struct ga_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned int *d_queue = nullptr;
    double *d_raw = nullptr;       // raw arrival records (grown on demand)
    double *d_yp = nullptr;        // per-lane row k-1 buffer (grown on demand)
    size_t yp_cap = 0;
    size_t raw_cap = 0;            // in doubles
    int n_cu = 0;
    size_t max_lds = 0;
    std::map<int, DevAtmo> atmos;
    int next_id = 1;
    std::string err;
};

namespace {

int fail(ga_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail((ctx), GA_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Natural cubic spline slopes by the Thomas algorithm, in the operation
// order of interp::set (src/util/interpolation.cpp:103-136).
void spline_slopes(const std::vector<double> &x, const std::vector<double> &f, std::vector<double> &s)
{
    const int n = (int)x.size();
    std::vector<double> nc(n), nd(n);
    s.assign(n, 0.0);
    double ai, bi, ci, di;
    bi = 2.0 / (x[1] - x[0]);
    ci = 1.0 / (x[1] - x[0]);
    di = 3.0 * (f[1] - f[0]) / ((x[1] - x[0]) * (x[1] - x[0]));
    nc[0] = ci / bi;
    nd[0] = di / bi;
    for (int i = 1; i < n - 1; i++) {
        double dx1 = x[i] - x[i - 1], dx2 = x[i + 1] - x[i];
        ai = 1.0 / dx1;
        bi = 2.0 * (1.0 / dx1 + 1.0 / dx2);
        ci = 1.0 / dx2;
        di = 3.0 * ((f[i] - f[i - 1]) / (dx1 * dx1) + (f[i + 1] - f[i]) / (dx2 * dx2));
        nc[i] = ci / (bi - nc[i - 1] * ai);
        nd[i] = (di - nd[i - 1] * ai) / (bi - nc[i - 1] * ai);
    }
    double dl = x[n - 1] - x[n - 2];
    ai = 1.0 / dl;
    bi = 2.0 / dl;
    di = 3.0 * (f[n - 1] - f[n - 2]) / (dl * dl);
    nd[n - 1] = (di - nd[n - 2] * ai) / (bi - nc[n - 2] * ai);
    s[n - 1] = nd[n - 1];
    for (int i = n - 2; i > -1; i--) s[i] = nd[i] - nc[i] * s[i + 1];
}

// c(z) by the spline formula of the device code, on the host (arrival
// inclinations and the eigenray update need c at the source and ground).
double host_c_at(const DevAtmo &a, double z)
{
    double zs = std::min(std::max(z, a.z.front()), a.z.back());
    int k = 0;
    while (k < a.n - 2 && zs > a.z[k + 1]) ++k;
    std::vector<double> sl;
    spline_slopes(a.z, a.c, sl);
    double dx = a.z[k + 1] - a.z[k], X = (zs - a.z[k]) / dx, df = a.c[k + 1] - a.c[k];
    double Ac = sl[k] * dx - df, Bc = -sl[k + 1] * dx + df;
    return a.c[k] + X * (df + (1.0 - X) * (Ac * (1.0 - X) + Bc * X));
}

bool build_medium(ga_ctx *ctx, const DevAtmo &a, const ga_prop_params *p, Medium &M, Limits &L, int &rc)
{
    const bool sph = p->geometry == GA_GEOM_SPH;
    M.tab = a.d_tab;
    M.n = a.n;
    M.zlo = a.z.front();
    M.zhi = a.z.back();
    M.r0 = sph ? p->r_earth : 0.0;
    M.z0 = std::isnan(p->z_grnd) ? a.z.front() : p->z_grnd;
    M.zg = sph ? M.r0 + M.z0 : M.z0;
    M.z_max = M.z0;
    M.z_bndlyr = M.z0 + 2.0;
    M.tweak = std::max(0.0, p->abs_coeff);
    M.sbc[0] = std::pow(10.0, -0.67887);
    M.sbc[1] = std::pow(10.0, -0.10744);
    M.sbc[2] = std::pow(10.0, -3.3979);
    L.s_max = p->s_max;
    L.ds_min = p->ds_min;
    L.ds_max = p->ds_max;
    L.alt_max = std::isnan(p->alt_max) ? a.z.back() : p->alt_max;
    L.rng_max = p->rng_max;
    L.time_max = p->time_max;
    L.z_keep = M.z_max + 4.0 * p->ds_max;
    if (sph) {
        L.b0 = p->bounds[0] * (PI / 180.0);
        L.b1 = p->bounds[1] * (PI / 180.0);
        L.b2 = p->bounds[2] * (PI / 180.0);
        L.b3 = p->bounds[3] * (PI / 180.0);
        double q = p->rng_max / (2.0 * M.r0);
        double sq = q >= PI / 2.0 ? 2.0 : std::sin(q);
        L.gc_thresh = sq >= 1.0 ? 2.0 : sq * sq;
    } else {
        L.b0 = p->bounds[0];
        L.b1 = p->bounds[1];
        L.b2 = p->bounds[2];
        L.b3 = p->bounds[3];
        L.gc_thresh = 0.0;
    }
    if (!(p->ds_min > 0.0) || !(p->ds_max >= p->ds_min) || !(p->s_max > 0.0)) {
        rc = fail(ctx, GA_E_ARG, "invalid step parameters ds_min=%g ds_max=%g s_max=%g", p->ds_min, p->ds_max, p->s_max);
        return false;
    }
    return true;
}

template <int G, int NEQ>
int launch_fan(ga_ctx *ctx, const FanArgs &A, hipStream_t st)
{
    size_t lds = A.lds ? (size_t)A.M.n * NODE * sizeof(double) : 0;
    auto kern = A.lds ? fan_kernel<G, NEQ, true> : fan_kernel<G, NEQ, false>;
    if (lds > 65536) HIPCHK(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    constexpr int W = trace_wg<G, NEQ>();
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, W, lds));
    if (per_cu < 1) per_cu = 1;
    int grid = per_cu * ctx->n_cu;
    int need = (A.n_rays + W - 1) / W;
    grid = std::max(1, std::min(grid, need));
    FanArgs B = A;
    B.nlanes = grid * W;
    size_t ypn = (size_t)B.nlanes * 18;
    if (ypn > ctx->yp_cap) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        HIPCHK(ctx, hipFree(ctx->d_yp));
        ctx->d_yp = nullptr;
        ctx->yp_cap = 0;
        if (hipMalloc(&ctx->d_yp, ypn * sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "lane buffer");
        ctx->yp_cap = ypn;
    }
    B.ypbuf = ctx->d_yp;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, A.queue);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(W), lds, st, B);
    HIPCHK(ctx, hipGetLastError());
    long long slots = (long long)A.n_rays * (A.bounces + 1);
    int agrid = (int)((slots + 255) / 256);
    hipLaunchKernelGGL((arrival_kernel<G, NEQ>), dim3(agrid), dim3(256), 0, st, A);
    HIPCHK(ctx, hipGetLastError());
    return GA_OK;
}

int fan_enqueue(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *d_launches, int n,
                ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, hipStream_t st)
{
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    if (p->geometry != GA_GEOM_SPH && p->geometry != GA_GEOM_3D) return fail(ctx, GA_E_ARG, "bad geometry %d", p->geometry);
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    FanArgs A;
    std::memset(&A, 0, sizeof(A));
    int rc = GA_OK;
    if (!build_medium(ctx, it->second, p, A.M, A.L, rc)) return rc;
    const bool sph = p->geometry == GA_GEOM_SPH;
    const DevAtmo &a = it->second;
    // source position: z_src = max(z_src, ground) (main.infraga.sph.cpp:314, 3d.cpp:306)
    double zsrc = p->src[2];
    if (sph) {
        double zg = A.M.zg - A.M.r0;
        zsrc = std::max(zsrc, zg);
        A.src0 = A.M.r0 + zsrc;
        A.src1 = p->src[0] * (PI / 180.0);
        A.src2 = p->src[1] * (PI / 180.0);
    } else {
        zsrc = std::max(zsrc, A.M.zg);
        A.src0 = p->src[0];
        A.src1 = p->src[1];
        A.src2 = zsrc;
    }
    A.c_src = host_c_at(a, sph ? A.src0 - A.M.r0 : A.src2);
    A.freq = p->freq;
    A.turn_ht_min = p->turn_ht_min;
    A.rmax_init = sph ? 0.0 : A.M.z0;
    A.bounces = p->bounces;
    A.n_rays = n;
    long long rows = (long long)(p->s_max / p->ds_min);
    A.max_steps = (int)std::min<long long>(rows - 1, 2000000000LL);
    A.launches = d_launches;
    {
        size_t need = (size_t)std::max(n, 1) * (size_t)(p->bounces + 1) * RAW;
        if (need > ctx->raw_cap) {
            HIPCHK(ctx, hipStreamSynchronize(st));
            HIPCHK(ctx, hipFree(ctx->d_raw));
            ctx->d_raw = nullptr;
            ctx->raw_cap = 0;
            if (hipMalloc(&ctx->d_raw, need * sizeof(double)) != hipSuccess)
                return fail(ctx, GA_E_NOMEM, "raw arrival buffer (%zu doubles)", need);
            ctx->raw_cap = need;
        }
        A.raw = ctx->d_raw;
    }
    A.out = d_out;
    A.counts = d_counts;
    A.steps = d_steps;
    A.queue = ctx->d_queue;
    size_t lds = (size_t)a.n * NODE * sizeof(double);
    A.lds = lds <= ctx->max_lds ? 1 : 0;
    if (n == 0) return GA_OK;
    if (sph) return p->calc_amp ? launch_fan<GEOM_SPH, 18>(ctx, A, st) : launch_fan<GEOM_SPH, 6>(ctx, A, st);
    return p->calc_amp ? launch_fan<GEOM_3D, 18>(ctx, A, st) : launch_fan<GEOM_3D, 6>(ctx, A, st);
}


// RAII device buffer and event pair for the host entry points.
template <class T> struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    bool alloc(size_t n)
    {
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) { p = nullptr; return false; }
        cap = n;
        return true;
    }
    // grow-only; the caller has synchronised the streams that use the buffer
    bool ensure(size_t n)
    {
        if (n <= cap) return true;
        if (p && hipFree(p) != hipSuccess) return false;
        p = nullptr;
        cap = 0;
        return alloc(std::max(n, (size_t)1));
    }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

struct Events {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = false;
    Events() { ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess; }
    ~Events()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// GPU side of the eigenray search: one round of traces per call.  EIG_SWEEP
// jobs run on the context stream in the 6-equation kernel, EIG_LM/EIG_FINAL
// jobs on a second stream in the 18-equation kernel, followed there by the
// arrival rows of the EIG_FINAL jobs.
struct EigGpu {
    ga_ctx *ctx = nullptr;
    int geom = GA_GEOM_3D;
    FanArgs F{};                 // medium, limits, source, c_src, freq, max_steps, lds
    hipStream_t s2 = nullptr;
    DevBuf<EigJob> jobs6, jobs18;
    DevBuf<double> res6, res18, yp6, yp18;
    DevBuf<int32_t> brk6, brk18, st6, st18;
    DevBuf<ga_arrival> arr;
    DevBuf<unsigned int> q;
    double kernel_ms = 0.0;
    int launches = 0;

    ~EigGpu()
    {
        if (s2) {
            (void)hipStreamSynchronize(s2);
            (void)hipStreamDestroy(s2);
        }
    }

    int init()
    {
        HIPCHK(ctx, hipStreamCreate(&s2));
        if (!q.alloc(2)) return fail(ctx, GA_E_NOMEM, "eigenray queue counters");
        return GA_OK;
    }

    template <int G, int NEQ, int WGS>
    int launch(const EigArgs &E0, DevBuf<double> &yp, unsigned int *queue, hipStream_t st)
    {
        size_t lds = F.lds ? (size_t)F.M.n * NODE * sizeof(double) : 0;
        auto kern = F.lds ? eig_kernel<G, NEQ, true, WGS> : eig_kernel<G, NEQ, false, WGS>;
        if (lds > 65536) HIPCHK(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int per_cu = 0;
        HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WGS, lds));
        per_cu = std::max(per_cu, 1);
        int grid = std::max(1, std::min(per_cu * ctx->n_cu, (E0.n_jobs + WGS - 1) / WGS));
        if (!yp.ensure((size_t)grid * WGS * 18)) return fail(ctx, GA_E_NOMEM, "eigenray lane buffer");
        EigArgs E = E0;
        E.F.ypbuf = yp.p;
        E.F.nlanes = grid * WGS;
        E.F.queue = queue;
        hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, queue);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(WGS), lds, st, E);
        HIPCHK(ctx, hipGetLastError());
        launches += 2;
        return GA_OK;
    }

    int operator()(const EigBatch &B, EigBatchOut &O)
    {
        const size_t n6 = B.sweep.size(), n18 = B.full.size(), nf = n18 - (size_t)B.n_lm;
        const hipStream_t st = ctx->stream;
        HIPCHK(ctx, hipStreamSynchronize(st));
        HIPCHK(ctx, hipStreamSynchronize(s2));
        if (!jobs6.ensure(n6) || !res6.ensure(n6 * B.sweep_stride) || !brk6.ensure(n6) || !st6.ensure(n6) ||
            !jobs18.ensure(n18) || !res18.ensure(n18 * EIG_FULL) || !brk18.ensure(n18) || !st18.ensure(n18) ||
            !arr.ensure(nf))
            return fail(ctx, GA_E_NOMEM, "eigenray batch buffers (%zu + %zu traces)", n6, n18);
        Events ev;
        hipEvent_t mid = nullptr;
        if (!ev.ok || hipEventCreate(&mid) != hipSuccess) return fail(ctx, GA_E_HIP, "hipEventCreate failed");
        struct EvGuard { hipEvent_t e; ~EvGuard() { (void)hipEventDestroy(e); } } guard{mid};
        if (n6) HIPCHK(ctx, hipMemcpyAsync(jobs6.p, B.sweep.data(), n6 * sizeof(EigJob), hipMemcpyHostToDevice, st));
        if (n18) HIPCHK(ctx, hipMemcpyAsync(jobs18.p, B.full.data(), n18 * sizeof(EigJob), hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipEventRecord(ev.e0, st));
        HIPCHK(ctx, hipStreamWaitEvent(s2, ev.e0, 0));
        if (n6) {
            EigArgs E{};
            E.F = F;
            E.jobs = jobs6.p;
            E.n_jobs = (int)n6;
            E.stride = B.sweep_stride;
            E.res = res6.p;
            E.brk_leg = brk6.p;
            E.steps = st6.p;
            int rc = geom == GA_GEOM_SPH ? launch<GEOM_SPH, 6, GA_EIG_WG6>(E, yp6, q.p, st)
                                         : launch<GEOM_3D, 6, GA_EIG_WG6>(E, yp6, q.p, st);
            if (rc != GA_OK) return rc;
        }
        if (n18) {
            EigArgs E{};
            E.F = F;
            E.jobs = jobs18.p;
            E.n_jobs = (int)n18;
            E.stride = EIG_FULL;
            E.res = res18.p;
            E.brk_leg = brk18.p;
            E.steps = st18.p;
            int rc = geom == GA_GEOM_SPH ? launch<GEOM_SPH, 18, trace_wg<GEOM_SPH, 18>()>(E, yp18, q.p + 1, s2)
                                         : launch<GEOM_3D, 18, trace_wg<GEOM_3D, 18>()>(E, yp18, q.p + 1, s2);
            if (rc != GA_OK) return rc;
            if (nf) {
                auto ak = geom == GA_GEOM_SPH ? eig_arrival_kernel<GEOM_SPH> : eig_arrival_kernel<GEOM_3D>;
                hipLaunchKernelGGL(ak, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s2, E, B.n_lm,
                                   (int)nf, arr.p);
                HIPCHK(ctx, hipGetLastError());
                launches += 1;
            }
        }
        HIPCHK(ctx, hipEventRecord(mid, s2));
        HIPCHK(ctx, hipStreamWaitEvent(st, mid, 0));
        HIPCHK(ctx, hipEventRecord(ev.e1, st));
        O.sweep_res.resize(n6 * B.sweep_stride);
        O.sweep_brk.resize(n6);
        O.sweep_steps.resize(n6);
        O.full_res.resize(n18 * EIG_FULL);
        O.full_brk.resize(n18);
        O.full_steps.resize(n18);
        O.final_rows.resize(nf);
        if (n6) {
            HIPCHK(ctx, hipMemcpyAsync(O.sweep_res.data(), res6.p, O.sweep_res.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(O.sweep_brk.data(), brk6.p, n6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(O.sweep_steps.data(), st6.p, n6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        }
        if (n18) {
            HIPCHK(ctx, hipMemcpyAsync(O.full_res.data(), res18.p, O.full_res.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(O.full_brk.data(), brk18.p, n18 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(O.full_steps.data(), st18.p, n18 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        }
        if (nf) HIPCHK(ctx, hipMemcpyAsync(O.final_rows.data(), arr.p, nf * sizeof(ga_arrival), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        float ms = 0.0f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ev.e0, ev.e1));
        kernel_ms += ms;
        return GA_OK;
    }
};

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void ga_prop_params_default(ga_prop_params *p, int geometry)
{
    std::memset(p, 0, sizeof(*p));
    p->geometry = geometry;
    p->calc_amp = 1;
    p->bounces = 10;
    p->freq = 0.1;
    p->abs_coeff = 1.0;
    p->turn_ht_min = 0.2;
    p->z_grnd = NAN;
    p->s_max = 1000.0;
    p->ds_min = 0.001;
    p->ds_max = 0.05;
    p->alt_max = NAN;
    p->time_max = 10.0;
    p->r_earth = 6370.0;
    if (geometry == GA_GEOM_SPH) {
        p->src[0] = 30.0; p->src[1] = 0.0; p->src[2] = 0.0;
        p->rng_max = 1000.0;
        p->bounds[0] = -90.0; p->bounds[1] = 90.0; p->bounds[2] = -180.0; p->bounds[3] = 180.0;
    } else {
        p->src[0] = 0.0; p->src[1] = 0.0; p->src[2] = 0.0;
        p->rng_max = 2000.0;   // atmo_io.3d.strat.cpp:30
        p->bounds[0] = -2000.0; p->bounds[1] = 2000.0; p->bounds[2] = -2000.0; p->bounds[3] = 2000.0;
    }
}

void ga_eig_params_default(ga_eig_params *e)
{
    e->incl_min = 0.5;
    e->incl_max = 45.0;
    e->bnc_min = 0;
    e->bnc_max = 0;
    e->iterations = 25;
    e->damping = 1.0e-3;
    e->tolerance = 0.1;
    e->az_dev_lim = 2.0;
    e->incl_step_min = 0.001;
    e->incl_step_max = 0.1;
}

int ga_ctx_create(ga_ctx **out, int dev)
{
    if (!out) return GA_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= dev || dev < 0) return GA_E_NODEV;
    ga_ctx *ctx = new ga_ctx();
    ctx->device = dev;
    if (hipSetDevice(dev) != hipSuccess) { delete ctx; return GA_E_NODEV; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) { delete ctx; return GA_E_NODEV; }
    ctx->n_cu = prop.multiProcessorCount;
    ctx->max_lds = prop.sharedMemPerBlock;
    if (hipStreamCreate(&ctx->stream) != hipSuccess) { delete ctx; return GA_E_HIP; }
    if (hipMalloc(&ctx->d_queue, sizeof(unsigned int)) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return GA_E_NOMEM;
    }
    *out = ctx;
    return GA_OK;
}

void ga_ctx_destroy(ga_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->atmos) (void)hipFree(kv.second.d_tab);
    (void)hipFree(ctx->d_queue);
    (void)hipFree(ctx->d_raw);
    (void)hipFree(ctx->d_yp);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *ga_last_error(const ga_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ga_atmo_load_strat(ga_ctx *ctx, const double *z, const double *c, const double *u, const double *v,
                       const double *rho, int n, int *atmo_id)
{
    if (!ctx || !z || !c || !u || !v || !rho || !atmo_id) return fail(ctx, GA_E_ARG, "null argument");
    if (n < 3) return fail(ctx, GA_E_ARG, "profile needs at least 3 rows (got %d)", n);
    for (int i = 0; i + 1 < n; ++i)
        if (!(z[i + 1] > z[i])) return fail(ctx, GA_E_ARG, "altitudes must increase strictly (row %d)", i + 1);
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(c[i]) || !(c[i] > 0.0) || !std::isfinite(u[i]) || !std::isfinite(v[i]) || !std::isfinite(rho[i]))
            return fail(ctx, GA_E_ARG, "non-finite or non-positive profile value at row %d", i);
    DevAtmo a;
    a.n = n;
    a.z.assign(z, z + n);
    a.c.assign(c, c + n);
    a.u.assign(u, u + n);
    a.v.assign(v, v + n);
    a.rho.assign(rho, rho + n);
    std::vector<double> sc, su, sv, sr;
    spline_slopes(a.z, a.c, sc);
    spline_slopes(a.z, a.u, su);
    spline_slopes(a.z, a.v, sv);
    spline_slopes(a.z, a.rho, sr);
    std::vector<double> tab((size_t)n * NODE, 0.0);
    for (int i = 0; i < n; ++i) {
        double *t = &tab[(size_t)i * NODE];
        t[0] = z[i];
        t[1] = c[i]; t[2] = sc[i];
        t[3] = u[i]; t[4] = su[i];
        t[5] = v[i]; t[6] = sv[i];
        t[7] = rho[i]; t[8] = sr[i];
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (hipMalloc(&a.d_tab, tab.size() * sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "hipMalloc atmosphere");
    if (hipMemcpy(a.d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(a.d_tab);
        return fail(ctx, GA_E_HIP, "atmosphere upload failed");
    }
    int id = ctx->next_id++;
    ctx->atmos[id] = std::move(a);
    *atmo_id = id;
    return GA_OK;
}

int ga_atmo_load_file(ga_ctx *ctx, const char *path, const char *format, int reverse_winds, int *atmo_id)
{
    if (!ctx || !path || !atmo_id) return fail(ctx, GA_E_ARG, "null argument");
    std::string fmt = format ? format : "zTuvdp";
    std::ifstream in(path);
    if (!in.is_open()) return fail(ctx, GA_E_IO, "cannot open atmosphere file %s", path);
    std::vector<double> z, c, u, v, rho;
    std::string line;
    while (std::getline(in, line)) {
        if (line.find('#') == 0) continue;
        std::stringstream ss(line);
        double vals[7];
        int need = (fmt.compare(0, 6, "zTuvdp") == 0) ? 6 : (fmt.compare(0, 7, "zuvwTdp") == 0) ? 7 : (fmt.compare(0, 5, "zcuvd") == 0) ? 5 : -1;
        if (need < 0) return fail(ctx, GA_E_ARG, "unrecognized profile format %s (zTuvdp, zuvwTdp, zcuvd)", fmt.c_str());
        int got = 0;
        while (got < need && (ss >> vals[got])) ++got;
        if (got == 0) continue;   // blank line
        if (got < need) return fail(ctx, GA_E_IO, "short row in %s: '%s'", path, line.c_str());
        double zz, cc, uu, vv, rr;
        if (need == 6) { zz = vals[0]; uu = vals[2]; vv = vals[3]; rr = vals[4]; cc = std::sqrt(0.1 * 1.4 * vals[5] / rr) / 1000.0; }
        else if (need == 7) { zz = vals[0]; uu = vals[1]; vv = vals[2]; rr = vals[5]; cc = std::sqrt(0.1 * 1.4 * vals[6] / rr) / 1000.0; }
        else { zz = vals[0]; cc = vals[1] / 1000.0; uu = vals[2]; vv = vals[3]; rr = vals[4]; }
        if (reverse_winds) { uu /= -1000.0; vv /= -1000.0; }
        else { uu /= 1000.0; vv /= 1000.0; }
        z.push_back(zz); c.push_back(cc); u.push_back(uu); v.push_back(vv); rho.push_back(rr);
    }
    return ga_atmo_load_strat(ctx, z.data(), c.data(), u.data(), v.data(), rho.data(), (int)z.size(), atmo_id);
}

int ga_atmo_release(ga_ctx *ctx, int atmo_id)
{
    if (!ctx) return GA_E_ARG;
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(it->second.d_tab));
    ctx->atmos.erase(it);
    return GA_OK;
}

int ga_fan_run_device(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *d_launches, int n,
                      ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, void *stream)
{
    if (!ctx || !p || (n > 0 && (!d_launches || !d_out || !d_counts || !d_steps))) return fail(ctx, GA_E_ARG, "null argument");
    if (n < 0) return fail(ctx, GA_E_ARG, "negative ray count");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    return fan_enqueue(ctx, atmo_id, p, d_launches, n, d_out, d_counts, d_steps, st);
}

int ga_fan_run(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *launches, int n,
               ga_arrival *out, int out_cap, int *n_out, ga_run_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !p || !n_out || (n > 0 && !launches)) return fail(ctx, GA_E_ARG, "null argument");
    if (n < 0) return fail(ctx, GA_E_ARG, "negative ray count");
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int slots = p->bounces + 1;
    const size_t nl = (size_t)std::max(n, 1);
    DevBuf<double> d_l;
    DevBuf<ga_arrival> d_o;
    DevBuf<int32_t> d_c, d_s;
    if (!d_l.alloc(nl * 2) || !d_o.alloc(nl * slots) || !d_c.alloc(nl) || !d_s.alloc(nl))
        return fail(ctx, GA_E_NOMEM, "device buffers for %d rays", n);
    Events ev;
    if (!ev.ok) return fail(ctx, GA_E_HIP, "hipEventCreate failed");
    HIPCHK(ctx, hipMemcpyAsync(d_l.p, launches, (size_t)n * 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ev.e0, ctx->stream));
    int rc = fan_enqueue(ctx, atmo_id, p, d_l.p, n, d_o.p, d_c.p, d_s.p, ctx->stream);
    if (rc != GA_OK) return rc;
    HIPCHK(ctx, hipEventRecord(ev.e1, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ev.e0, ev.e1));
    std::vector<int32_t> cnt(nl, 0), stp(nl, 0);
    std::vector<ga_arrival> all;
    if (n > 0) {
        HIPCHK(ctx, hipMemcpy(cnt.data(), d_c.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(stp.data(), d_s.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
        all.resize((size_t)n * slots);
        HIPCHK(ctx, hipMemcpy(all.data(), d_o.p, all.size() * sizeof(ga_arrival), hipMemcpyDeviceToHost));
    }
    long long total = 0, steps = 0;
    for (int r = 0; r < n; ++r) { total += cnt[r]; steps += stp[r]; }
    *n_out = (int)total;
    if (stats) {
        stats->ray_steps = steps;
        stats->kernel_ms = ms;
        stats->n_rays = n;
        stats->n_launches = 3;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (total > out_cap || (!out && total > 0)) return fail(ctx, GA_E_CAP, "need %lld arrival slots, have %d", total, out_cap);
    size_t w = 0;
    for (int r = 0; r < n; ++r)
        for (int b = 0; b < cnt[r]; ++b) out[w++] = all[(size_t)r * slots + b];
    return GA_OK;
}

int ga_eig_search(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                  const double *rcvr_xy, int n_rcvr, ga_eigenray *out, int out_cap, int *n_out,
                  double *rcvr_latency_ms, ga_run_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !p || !e || !n_out || (n_rcvr > 0 && !rcvr_xy)) return fail(ctx, GA_E_ARG, "null argument");
    *n_out = 0;
    if (n_rcvr < 0) return fail(ctx, GA_E_ARG, "negative receiver count");
    if (p->geometry != GA_GEOM_3D && p->geometry != GA_GEOM_SPH) return fail(ctx, GA_E_ARG, "bad geometry %d", p->geometry);
    if (e->bnc_min < 0 || e->bnc_max < e->bnc_min) return fail(ctx, GA_E_ARG, "bad bounce range %d..%d", e->bnc_min, e->bnc_max);
    if (e->iterations < 0) return fail(ctx, GA_E_ARG, "negative iteration limit");
    if (!(e->incl_step_max > 0.0) || !(e->incl_step_min > 0.0))
        return fail(ctx, GA_E_ARG, "inclination steps must be positive");
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const DevAtmo &a = it->second;

    const bool sph = p->geometry == GA_GEOM_SPH;
    EigGpu gpu;
    gpu.ctx = ctx;
    gpu.geom = p->geometry;
    FanArgs &F = gpu.F;
    int rc = GA_OK;
    if (!build_medium(ctx, a, p, F.M, F.L, rc)) return rc;
    EigSetup S;
    S.geom = p->geometry;
    S.r0 = F.M.r0;
    S.src_ground = F.M.z0;
    std::vector<double> rcvr(rcvr_xy, rcvr_xy + 2 * (size_t)std::max(n_rcvr, 0));
    double c_arr;   // divisor of the arrival inclination
    if (sph) {
        // main.infraga.sph.cpp:906-909: degrees to radians, source raised to the ground
        const double zsrc = std::max(p->src[2], F.M.z0);
        S.src[0] = zsrc;
        S.src[1] = p->src[0] * (PI / 180.0);
        S.src[2] = p->src[1] * (PI / 180.0);
        for (double &x : rcvr) x *= (PI / 180.0);
        F.src0 = F.M.r0 + zsrc;
        F.src1 = S.src[1];
        F.src2 = S.src[2];
        S.c_src = host_c_at(a, zsrc);
        // atmo::c(src[0], ...) reads the source altitude as a radius, which
        // clamps to the profile's bottom value (geoac.eigenray.sph.cpp:273)
        c_arr = host_c_at(a, a.z.front());
        // theta_grnd = 0.1 deg for a source on the ground (sph.cpp:966-979)
        const double th_grnd = (zsrc - F.M.z0 < 0.01) ? 0.1 * (PI / 180.0) : -89.9 * (PI / 180.0);
        S.th_start = std::max(e->incl_min * (PI / 180.0), th_grnd);
    } else {
        // src[2] = max(topo::z(src), src[2]) (main.infraga.3d.cpp:857)
        F.src0 = p->src[0];
        F.src1 = p->src[1];
        F.src2 = std::max(p->src[2], F.M.zg);
        S.src[0] = F.src0;
        S.src[1] = F.src1;
        S.src[2] = F.src2;
        S.c_src = host_c_at(a, F.src2);
        c_arr = S.c_src;
        S.th_start = e->incl_min * (PI / 180.0);
    }
    F.c_src = c_arr;
    F.freq = p->freq;
    long long rows = (long long)(p->s_max / p->ds_min);
    F.max_steps = (int)std::min<long long>(rows - 1, 2000000000LL);
    F.lds = (size_t)a.n * NODE * sizeof(double) <= ctx->max_lds ? 1 : 0;

    S.th_min = e->incl_min * (PI / 180.0);
    S.th_max = e->incl_max * (PI / 180.0);
    S.bnc_min = e->bnc_min;
    S.bnc_max = e->bnc_max;
    S.iterations = e->iterations;
    S.damping = e->damping;
    S.tolerance = e->tolerance;
    S.az_err_lim = e->az_dev_lim;
    S.dth_big = e->incl_step_max * (PI / 180.0);
    S.dth_sml = e->incl_step_min * (PI / 180.0);
    S.rng_max = p->rng_max;
    S.c_grnd = host_c_at(a, F.M.z0);
    S.spec_chunk = 64;

    rc = gpu.init();
    if (rc != GA_OK) return rc;
    std::vector<ga_arrival> found;
    std::vector<double> lat;
    EigSearchStats es;
    {
        EigSearch search(S, rcvr.data(), n_rcvr);
        rc = search.run([&](const EigBatch &B, EigBatchOut &O) { return gpu(B, O); }, found, lat, es);
    }
    if (rc != GA_OK) return rc;
    *n_out = (int)found.size();
    if (rcvr_latency_ms)
        for (int i = 0; i < n_rcvr; ++i) rcvr_latency_ms[i] = lat[i];
    if (stats) {
        stats->ray_steps = es.ray_steps;
        stats->kernel_ms = gpu.kernel_ms;
        stats->n_rays = (int32_t)std::min<long long>(es.traces, 2147483647LL);
        stats->n_launches = gpu.launches;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if ((long long)found.size() > out_cap || (!out && !found.empty()))
        return fail(ctx, GA_E_CAP, "need %zu eigenray slots, have %d", found.size(), out_cap);
    for (size_t i = 0; i < found.size(); ++i) out[i] = found[i];
    return GA_OK;
}

}  // extern "C"
