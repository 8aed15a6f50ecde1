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
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/infraga_gpu.h"
#include "ga_2d.hpp"
#include "ga_eig_search.hpp"
#include "ga_rdeig.hpp"
#include "ga_rdtrace.hpp"
#include "ga_rngdep.hpp"
#include "ga_trace.hpp"
#include "ga_wvfrm.hpp"
#include "ga_abs_host.hpp"

using namespace ga;

namespace {

// ONE: one ray per lane (grid covers all rays, for launches that share the
// device); otherwise persistent lanes pull rays from the queue counter.
// PATHS: write_rays / write_caustics bookkeeping with path records to A.path.
template <int G, int NEQ, bool LDS, bool ONE, bool PATHS = false>
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
    if constexpr (ONE) {
        bool first = true;
        lane_loop<G, NEQ>(A, M, lane, [&]() {
            const int r = first ? lane : A.n_rays;
            first = false;
            return r;
        });
    } else {
        lane_loop<G, NEQ, PATHS>(A, M, lane, [&]() { return (int)atomicAdd(A.queue, 1u); });
    }
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

// Ring control words: head and tail of the pending records, and the append
// counter on a cache line of its own (the slice kernel's atomics).
constexpr int CTL_HEAD = 0, CTL_TAIL = 1, CTL_APPEND = 32, CTL_WORDS = 64;

// Every ray of the fan as a record that has not started, head 0, tail n.
__global__ void ring_init(RayRing R, int n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        R.iv[i] = (int)i;
        R.iv[R.cap + i] = -1;
    }
    if (i == 0) {
        R.ctl[CTL_HEAD] = 0;
        R.ctl[CTL_TAIL] = (unsigned long long)n;
        R.ctl[CTL_APPEND] = (unsigned long long)n;
    }
}

// One slice: up to R.lanes records from the ring head, spread evenly over
// the workgroups (workgroup b takes records [b * per, (b + 1) * per) of
// the slice), each ray advanced by at most R.slice_steps RK4 steps; the
// records of unfinished rays are appended at the tail, packed per wave.
template <int G, int NEQ, bool LDS>
__global__ void __launch_bounds__((trace_wg<G, NEQ>())) slice_kernel(FanArgs A, RayRing R)
{
    extern __shared__ double lds_tab[];
    const unsigned long long head = R.ctl[CTL_HEAD], tail = R.ctl[CTL_TAIL];
    const long long m = (long long)min(tail - head, (unsigned long long)R.lanes);
    const long long per = (m + gridDim.x - 1) / gridDim.x;
    if ((long long)blockIdx.x * per >= m) return;        // whole workgroup idle this slice
    Medium M = A.M;
    if constexpr (LDS) {
        const int nd = M.n * NODE;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) lds_tab[i] = A.M.tab[i];
        __syncthreads();
        M.tab = lds_tab;
    }
    const long long j = (long long)blockIdx.x * per + threadIdx.x;
    if ((long long)threadIdx.x >= per || j >= m) return;
    RayRec<NEQ> r;
    ring_load(R, (head + (unsigned long long)j) % R.cap, r);
    const bool alive = slice_trace<G, NEQ>(A, M, r, R.slice_steps);
    const unsigned long long act = __ballot(1), live = __ballot(alive);
    const int lane = (int)__lane_id();
    const int leader = __ffsll((long long)act) - 1;
    unsigned long long base = 0;
    if (lane == leader && live) base = atomicAdd(R.ctl + CTL_APPEND, (unsigned long long)__popcll(live));
    base = __shfl(base, leader);
    if (alive) {
        const unsigned long long below = live & ((1ull << lane) - 1ull);
        ring_store(R, (base + (unsigned long long)__popcll(below)) % R.cap, r);
    }
}

// After a slice: the records it took leave the head, the appended ones
// extend the tail.
__global__ void slice_ctl(RayRing R)
{
    const unsigned long long head = R.ctl[CTL_HEAD], tail = R.ctl[CTL_TAIL];
    const unsigned long long m = min(tail - head, (unsigned long long)R.lanes);
    R.ctl[CTL_HEAD] = head + m;
    R.ctl[CTL_TAIL] = R.ctl[CTL_APPEND];
}

// Absorption table records (ga_device.hpp), one per thread.
template <bool SPH>
__global__ void __launch_bounds__(64) abs_fill(Medium M, AbsBasis B, const double *geo, int nrec, double freq,
                                               double *rec, double *err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrec) return;
    const double *g = geo + (size_t)k * ABS_GEO;
    abs_fill_record<SPH>(M, B, g[0], g[1], g[2], g[3], g[4], (int)g[5], freq, rec + (size_t)k * ABS_REC, err + k);
}

// 2D fans (infraga-2d): persistent lanes, the stratified table in LDS when it fits.
#ifndef GA_WG_2D
#define GA_WG_2D 512
#endif
template <int NEQ, bool LDS>
__global__ void __launch_bounds__(GA_WG_2D) fan2d_kernel(Fan2DArgs A)
{
    extern __shared__ double lds_tab[];
    Medium M = A.M;
    if constexpr (LDS) {
        const int nd = M.n * NODE;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) lds_tab[i] = A.M.tab[i];
        __syncthreads();
        M.tab = lds_tab;
    }
    lane_loop_2d<NEQ>(A, M, [&]() { return (int)atomicAdd(A.queue, 1u); });
}

// Range-dependent atmosphere and topography at n points (lat, lon [rad], z
// [km]): per point 4 fields x {f, fx, fy, fz, fxx, fyy, fzz, fxy, fxz, fyz}
// and the topography {z, zx, zy, zxx, zxy, zyy, zxxx, zxxy, zxyy, zyyy}.
__global__ void __launch_bounds__(256) rd_eval_kernel(RdGrid g, TopoGrid t, int has_topo, const double *pts, int n,
                                                      double *aout, double *tout)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    for (int fld = 0; fld < 4; ++fld) {
        double f, df[3], ddf[6];
        rd_field_all(g, fld, x, y, z, f, df, ddf);
        double *o = aout + ((size_t)i * 4 + fld) * 10;
        o[0] = f;
        for (int k = 0; k < 3; ++k) o[1 + k] = df[k];
        for (int k = 0; k < 6; ++k) o[4 + k] = ddf[k];
    }
    double tz[10];
    if (has_topo) topo_all(t, x, y, tz);
    else for (int k = 0; k < 10; ++k) tz[k] = 0.0;
    for (int k = 0; k < 10; ++k) tout[(size_t)i * 10 + k] = tz[k];
}

// Range-dependent fans: the hybrid-spline tables stay in global memory
// (29 MB for an 11 x 11 x 1501 grid) and are read through L1/L2; lanes pull
// rays from the queue counter (ONE: one ray per lane).
#ifndef GA_WG_RD
#define GA_WG_RD 64
#endif
template <int NEQ, bool ONE, bool BLW, bool PATHS = false, bool CART = false>
__global__ void __launch_bounds__(GA_WG_RD) rd_fan_kernel(RdFanArgs A)
{
    // three RK4 rows per lane (state, accumulator, stage point), lane-interleaved
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double rows[3 * NEQ * GA_WG_RD];
    const LaneVec<GA_WG_RD> ya{rows + threadIdx.x}, yb{rows + NEQ * GA_WG_RD + threadIdx.x},
        ys{rows + 2 * NEQ * GA_WG_RD + threadIdx.x};
    if constexpr (ONE) {
        bool first = true;
        rd_lane_loop<NEQ, BLW, false, CART>(A, lane, [&]() {
            const int r = first ? lane : A.n_rays;
            first = false;
            return r;
        }, ya, yb, ys);
    } else {
        rd_lane_loop<NEQ, BLW, PATHS, CART>(A, lane, [&]() { return (int)atomicAdd(A.queue, 1u); }, ya, yb, ys);
    }
}

template <int NEQ, bool CART = false>
__global__ void __launch_bounds__(256) rd_arrival_kernel(RdFanArgs A)
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
    if (CART)
        A.out[slot] = rd3_arrival_row<NEQ>(A, y, ray, bnc, A.launches[2 * ray], A.launches[2 * ray + 1], r[18], r[19],
                                           r[20], r[23], r[21], r[22]);
    else
        A.out[slot] = rd_arrival_row<NEQ>(A, y, ray, bnc, A.launches[2 * ray], A.launches[2 * ray + 1], r[18], r[19],
                                          r[20], r[23], r[24], r[21], r[22]);
}

// Terrain height above sea level at one point (the source of a fan).
// Eigenray traces over the range-dependent medium (same job set as eig_kernel).
template <int NEQ, bool BLW, bool PATHS, bool CART>
__global__ void __launch_bounds__(GA_WG_RD) rd_eig_kernel(RdEigArgs E)
{
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double rows[3 * NEQ * GA_WG_RD];
    const LaneVec<GA_WG_RD> ya{rows + threadIdx.x}, yb{rows + NEQ * GA_WG_RD + threadIdx.x},
        ys{rows + 2 * NEQ * GA_WG_RD + threadIdx.x};
    rd_eig_lane_loop<NEQ, BLW, PATHS, CART>(E, lane, [&]() { return (int)atomicAdd(E.F.queue, 1u); }, ya, yb, ys);
}

// Arrival rows of the range-dependent EIG_FINAL jobs [first, first + n).
template <bool CART>
__global__ void __launch_bounds__(256) rd_eig_arrival_kernel(RdEigArgs E, int first, int n, ga_arrival *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const EigJob J = E.jobs[first + i];
    const double *r = E.res + (size_t)(first + i) * EIG_FULL;
    out[i] = CART ? rd3_eig_arrival_row(E, J, r) : rd_eig_arrival_row(E, J, r);
}

__global__ void rd_ground_kernel(RdMed M, double lat, double lon, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = rd_ground(M, lat, lon) - M.r0;
}

// Eigenray-search traces: same persistent layout as fan_kernel, jobs instead
// of launch angles.  WGS lanes per workgroup share one LDS copy of the table;
// the 6-equation sweep kernel needs <= 110 registers and runs four waves per
// SIMD.
#ifndef GA_EIG_WG6
#define GA_EIG_WG6 1024
#endif
template <int G, int NEQ, bool LDS, int WGS, bool ONE, bool PATHS = false>
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
    if constexpr (ONE) {
        bool first = true;
        eig_lane_loop<G, NEQ>(E, M, lane, [&]() {
            const int r = first ? lane : E.n_jobs;
            first = false;
            return r;
        });
    } else {
        eig_lane_loop<G, NEQ, PATHS>(E, M, lane, [&]() { return (int)atomicAdd(E.F.queue, 1u); });
    }
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
// Absorption table of an atmosphere at one frequency and geometry (see
// ga_device.hpp); ok = false records that no table is used (layout or
// accuracy check failed).
struct AbsTable {
    double freq = 0.0;
    bool sph = false;
    bool ok = false;
    int per = 0;
    double *d = nullptr;
};

struct DevAtmo {
    double *d_tab = nullptr;
    int n = 0;
    std::vector<double> z, c, u, v, rho;
    double *d_topo2d = nullptr;        // 2D terrain spline [n][3] (r, z, slope)
    std::vector<double> topo_r, topo_z;
    std::vector<AbsTable> abs;
    void release_abs()
    {
        for (AbsTable &t : abs)
            if (t.d) (void)hipFree(t.d);
        abs.clear();
    }
};

// Range-dependent atmosphere (hybrid-spline tables) and optional topography.
struct DevRd {
    double *d_axes = nullptr;      // lat [nx], lon [ny], z [nz]
    double *d_tab = nullptr;       // [4][nx][ny][nz][RD_Q]
    int nx = 0, ny = 0, nz = 0;
    double *d_taxes = nullptr;     // topography lat [tnx], lon [tny]
    double *d_ttab = nullptr;      // [tnx][tny][TP_Q]
    int tnx = 0, tny = 0;
    std::vector<double> hx, hy, hz;    // host copies of the node axes [rad, rad, km]
    std::vector<double> thx, thy;      // terrain axes [rad]
    double t_max = 0.0;                // terrain maximum, floored at 0 (set_bndlyr)
    bool cart = false;                 // 3D Cartesian grid: node axes in km
    bool strat = false;                // a stratified profile over terrain (ga_atmo_load_strat_topo)
    void release()
    {
        for (double *p : {d_axes, d_tab, d_taxes, d_ttab})
            if (p) (void)hipFree(p);
        d_axes = d_tab = d_taxes = d_ttab = nullptr;
    }
};

}  // namespace

// Device scratch of one fan launch: raw arrival records, the per-lane copy
// of row k-1 and the work-queue counter.  Grown on demand; a launch in
// flight owns its set, so launches on different streams use different sets.
struct FanBufs {
    unsigned int *queue = nullptr;
    double *scalar = nullptr;      // one-value readback (terrain height at a source)
    double *raw = nullptr;
    size_t raw_cap = 0;            // doubles
    double *yp = nullptr;
    size_t yp_cap = 0;             // doubles
    // fixed: the buffers are views of preallocated memory that must never be
    // grown or freed (a launch that needs more fails instead); max_lanes
    // (> 0) caps the persistent grid so that the lane buffer fits.
    bool fixed = false;
    size_t max_lanes = 0;
    // rotating slice scheduler: ring records, control words, pinned readback
    // of head / tail, a capture stream and the cached graph of one batch of
    // slices (re-captured when the launch or its arguments change)
    double *ring_d = nullptr;      // [(18 + RING_SC)][cap]
    int32_t *ring_i = nullptr;     // [RING_IV][cap]
    size_t ring_cap = 0;
    unsigned long long *ctl = nullptr;
    unsigned long long *h_ctl = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    std::vector<unsigned char> graph_key;
    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_key.clear();
    }
    void release()
    {
        if (fixed) return;
        drop_graph();
        if (ring_d) (void)hipFree(ring_d);
        if (ring_i) (void)hipFree(ring_i);
        if (ctl) (void)hipFree(ctl);
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        ring_d = nullptr; ring_i = nullptr; ring_cap = 0; ctl = nullptr; h_ctl = nullptr; cap_stream = nullptr;
        if (queue) (void)hipFree(queue);
        if (scalar) (void)hipFree(scalar);
        if (raw) (void)hipFree(raw);
        if (yp) (void)hipFree(yp);
        queue = nullptr; scalar = nullptr; raw = nullptr; yp = nullptr; raw_cap = 0; yp_cap = 0;
    }
};

// Device scratch of ga_fan_run_device_noalloc: the queue counter, the
// per-lane row k-1 buffer of NOALLOC_LANES lanes and the raw arrival
// records.  Each device holds its own copy of this array for the life of
// the process; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (scratch_addr), never through
// the array's name.
constexpr size_t NOALLOC_BYTES = size_t(48) << 20;
constexpr size_t NOALLOC_LANES = 131072;                  // 256 CUs x 512 lanes
constexpr size_t NOALLOC_YP_OFF = 256;                    // after the queue counter
constexpr size_t NOALLOC_RAW_OFF = NOALLOC_YP_OFF + NOALLOC_LANES * 18 * sizeof(double);
constexpr size_t NOALLOC_RAW_DOUBLES = (NOALLOC_BYTES - NOALLOC_RAW_OFF) / sizeof(double);
static_assert(NOALLOC_RAW_OFF % 256 == 0, "raw records stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

struct ga_ctx {
    int device = 0;
    unsigned char *scratch_ptr = nullptr;   // this device's address of scratch
    // slice scheduler settings per [geometry sph?][18 equations?] chosen by
    // ga_fan_autotune (0: the built-in defaults)
    int tuned_slice[2][2] = {{0, 0}, {0, 0}};
    int tuned_graph[2][2] = {{0, 0}, {0, 0}};
    hipStream_t stream = nullptr;
    FanBufs fb;                    // scratch of ga_fan_run / ga_fan_run_device
    int n_cu = 0;
    size_t max_lds = 0;
    std::map<int, DevAtmo> atmos;
    std::map<int, DevRd> rds;
    int next_id = 1;
    int64_t wv_recs_per_ray = 0;   // most waveform-row records one ray needed so far
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

// Address of scratch on device dev, looked up once per device with that
// device current (the caller has made it current); nullptr if the lookup
// fails.
unsigned char *scratch_addr(int dev)
{
    static std::mutex mu;
    static std::map<int, unsigned char *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    addr[dev] = (unsigned char *)ptr;
    return (unsigned char *)ptr;
}

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

// The absorption table of atmosphere a at (freq, geometry): built on first
// use (one kernel on a stream of its own, waited for) and kept with the
// atmosphere; a table whose accuracy check fails, or an atmosphere whose
// node spacing is too coarse, runs with direct evaluation.  may_alloc =
// false (callers that allocate nothing) uses a table only if one exists.
// GA_ABS_TABLE=0 turns tables off.
void abs_attach(ga_ctx *ctx, DevAtmo &a, bool sph, double freq, bool may_alloc, Medium &M)
{
    M.abs = nullptr;
    M.abs_per = 0;
    const char *ev = std::getenv("GA_ABS_TABLE");
    if (ev && *ev && std::atoi(ev) == 0) return;
    for (const AbsTable &t : a.abs) {
        if (t.freq == freq && t.sph == sph) {
            if (t.ok) {
                M.abs = t.d;
                M.abs_per = t.per;
            }
            return;
        }
    }
    if (!may_alloc) return;
    AbsTable t;
    t.freq = freq;
    t.sph = sph;
    std::vector<double> geo;
    if (freq > 0.0 && abs_layout(a.z, geo, t.per)) {
        const int nrec = (int)(geo.size() / ABS_GEO);
        const AbsBasis B = abs_basis();
        double *d_geo = nullptr, *d_err = nullptr;
        hipStream_t st = nullptr;
        std::vector<double> err(nrec, 1.0);
        bool good = hipMalloc(&t.d, (size_t)nrec * ABS_REC * sizeof(double)) == hipSuccess &&
                    hipMalloc(&d_geo, geo.size() * sizeof(double)) == hipSuccess &&
                    hipMalloc(&d_err, (size_t)nrec * sizeof(double)) == hipSuccess &&
                    hipStreamCreate(&st) == hipSuccess &&
                    hipMemcpyAsync(d_geo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess;
        if (good) {
            Medium Mf = M;
            auto kern = sph ? abs_fill<true> : abs_fill<false>;
            hipLaunchKernelGGL(kern, dim3((nrec + 63) / 64), dim3(64), 0, st, Mf, B, d_geo, nrec, freq, t.d, d_err);
            good = hipGetLastError() == hipSuccess &&
                   hipMemcpyAsync(err.data(), d_err, (size_t)nrec * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess &&
                   hipStreamSynchronize(st) == hipSuccess;
        }
        if (st) (void)hipStreamDestroy(st);
        if (d_geo) (void)hipFree(d_geo);
        if (d_err) (void)hipFree(d_err);
        double worst = 0.0;
        for (double e : err) worst = std::max(worst, e);
        t.ok = good && worst <= ABS_TOL;
        if (!t.ok && t.d) {
            (void)hipFree(t.d);
            t.d = nullptr;
        }
        (void)hipGetLastError();
    }
    a.abs.push_back(t);
    if (t.ok) {
        M.abs = t.d;
        M.abs_per = t.per;
    }
}

bool build_medium(ga_ctx *ctx, DevAtmo &a, const ga_prop_params *p, Medium &M, Limits &L, int &rc,
                  bool may_alloc = true)
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
    abs_attach(ctx, a, sph, p->freq, may_alloc, M);
    return true;
}

template <int G, int NEQ>
int launch_fan(ga_ctx *ctx, FanBufs &fb, const FanArgs &A, hipStream_t st, bool paths)
{
    size_t lds = A.lds ? (size_t)A.M.n * NODE * sizeof(double) : 0;
    auto kern = paths ? (A.lds ? fan_kernel<G, NEQ, true, false, true> : fan_kernel<G, NEQ, false, false, true>)
              : A.one_shot ? (A.lds ? fan_kernel<G, NEQ, true, true> : fan_kernel<G, NEQ, false, true>)
                           : (A.lds ? fan_kernel<G, NEQ, true, false> : fan_kernel<G, NEQ, false, false>);
    if (lds > 65536) HIPCHK(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    constexpr int W = trace_wg<G, NEQ>();
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, W, lds));
    if (per_cu < 1) per_cu = 1;
    int grid = per_cu * ctx->n_cu;
    int need = (A.n_rays + W - 1) / W;
    grid = (A.one_shot && !paths) ? std::max(1, need) : std::max(1, std::min(grid, need));
    if (fb.max_lanes > 0 && !A.one_shot) grid = std::max(1, std::min(grid, (int)(fb.max_lanes / W)));
    FanArgs B = A;
    B.nlanes = grid * W;
    size_t ypn = (size_t)B.nlanes * 18;
    if (ypn > fb.yp_cap) {
        if (fb.fixed) return fail(ctx, GA_E_CAP, "lane buffer of %zu doubles exceeds the fixed %zu", ypn, fb.yp_cap);
        HIPCHK(ctx, hipStreamSynchronize(st));
        HIPCHK(ctx, hipFree(fb.yp));
        fb.yp = nullptr;
        fb.yp_cap = 0;
        if (hipMalloc(&fb.yp, ypn * sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "lane buffer");
        fb.yp_cap = ypn;
    }
    B.ypbuf = fb.yp;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, A.queue);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(W), lds, st, B);
    HIPCHK(ctx, hipGetLastError());
    long long slots = (long long)A.n_rays * (A.bounces + 1);
    int agrid = (int)((slots + 255) / 256);
    hipLaunchKernelGGL((arrival_kernel<G, NEQ>), dim3(agrid), dim3(256), 0, st, A);
    HIPCHK(ctx, hipGetLastError());
    return GA_OK;
}

// Steps per slice and slices per captured graph of the rotating scheduler;
// GA_FAN_SLICE (steps, 0 = one persistent launch with a ray queue),
// GA_FAN_GRAPH (slices per graph) and GA_FAN_GRID (at most that many
// workgroups) override them.
int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}
constexpr int SLICE_STEPS = 64;
constexpr int SLICES_PER_GRAPH = 32;

// The fan through the rotating slice scheduler (write_rays=false sums):
// ring of ray records, batches of slices captured once into a hipGraph and
// replayed until the ring is empty (one readback of head and tail per
// batch), then the arrival kernel.  The per-ray arithmetic is lane_loop's.
template <int G, int NEQ>
int launch_fan_sliced(ga_ctx *ctx, FanBufs &fb, const FanArgs &A, hipStream_t st, int slice_steps, int graph_k)
{
    size_t lds = A.lds ? (size_t)A.M.n * NODE * sizeof(double) : 0;
    auto kern = A.lds ? slice_kernel<G, NEQ, true> : slice_kernel<G, NEQ, false>;
    if (lds > 65536) HIPCHK(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    constexpr int W = trace_wg<G, NEQ>();
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, W, lds));
    if (per_cu < 1) per_cu = 1;
    int grid = per_cu * ctx->n_cu;
    const int grid_cap = env_int("GA_FAN_GRID", 0);       // fewer workgroups (tests, tuning)
    if (grid_cap > 0) grid = std::min(grid, grid_cap);
    const int lanes = grid * W;
    const int n = A.n_rays;
    const size_t cap = (size_t)n + (size_t)lanes;
    if (cap > fb.ring_cap) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        fb.drop_graph();
        if (fb.ring_d) HIPCHK(ctx, hipFree(fb.ring_d));
        if (fb.ring_i) HIPCHK(ctx, hipFree(fb.ring_i));
        fb.ring_d = nullptr;
        fb.ring_i = nullptr;
        fb.ring_cap = 0;
        if (hipMalloc(&fb.ring_d, cap * (18 + RING_SC) * sizeof(double)) != hipSuccess ||
            hipMalloc(&fb.ring_i, cap * RING_IV * sizeof(int32_t)) != hipSuccess)
            return fail(ctx, GA_E_NOMEM, "ray ring of %zu records", cap);
        fb.ring_cap = cap;
    }
    if (!fb.ctl && hipMalloc(&fb.ctl, CTL_WORDS * sizeof(unsigned long long)) != hipSuccess)
        return fail(ctx, GA_E_NOMEM, "ring control words");
    if (!fb.h_ctl && hipHostMalloc((void **)&fb.h_ctl, 2 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
        fb.h_ctl = nullptr;
        return fail(ctx, GA_E_NOMEM, "pinned readback");
    }
    if (!fb.cap_stream) HIPCHK(ctx, hipStreamCreateWithFlags(&fb.cap_stream, hipStreamNonBlocking));
    const size_t ypn = (size_t)std::max(n, 1) * 18;   // row k-1 per ray
    if (ypn > fb.yp_cap) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        fb.drop_graph();
        HIPCHK(ctx, hipFree(fb.yp));
        fb.yp = nullptr;
        fb.yp_cap = 0;
        if (hipMalloc(&fb.yp, ypn * sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "row buffer");
        fb.yp_cap = ypn;
    }
    FanArgs B = A;
    B.ypbuf = fb.yp;
    B.nlanes = lanes;
    RayRing R;
    R.cap = cap;
    R.y = fb.ring_d;
    R.sc = fb.ring_d + 18 * cap;
    R.iv = fb.ring_i;
    R.ctl = fb.ctl;
    R.lanes = lanes;
    R.slice_steps = std::max(1, slice_steps);
    const int K = std::max(1, graph_k);

    // one batch of K slices as a graph, re-captured when anything it bakes in changes
    std::vector<unsigned char> key(sizeof(void *) + sizeof(int) * 3 + sizeof(size_t) + sizeof(FanArgs) + sizeof(RayRing));
    {
        unsigned char *k = key.data();
        const void *kp = (const void *)kern;
        std::memcpy(k, &kp, sizeof(void *)); k += sizeof(void *);
        std::memcpy(k, &grid, sizeof(int)); k += sizeof(int);
        const int w = W;
        std::memcpy(k, &w, sizeof(int)); k += sizeof(int);
        std::memcpy(k, &K, sizeof(int)); k += sizeof(int);
        std::memcpy(k, &lds, sizeof(size_t)); k += sizeof(size_t);
        std::memcpy(k, &B, sizeof(FanArgs)); k += sizeof(FanArgs);
        std::memcpy(k, &R, sizeof(RayRing));
    }
    if (!fb.graph || fb.graph_key != key) {
        fb.drop_graph();
        hipGraph_t g = nullptr;
        HIPCHK(ctx, hipStreamBeginCapture(fb.cap_stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < K; ++i) {
            hipLaunchKernelGGL(kern, dim3(grid), dim3(W), lds, fb.cap_stream, B, R);
            hipLaunchKernelGGL(slice_ctl, dim3(1), dim3(1), 0, fb.cap_stream, R);
        }
        hipError_t ce = hipStreamEndCapture(fb.cap_stream, &g);
        if (ce != hipSuccess) return fail(ctx, GA_E_HIP, "graph capture of the slice batch: %s", hipGetErrorString(ce));
        hipGraphExec_t ex = nullptr;
        ce = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ce != hipSuccess) return fail(ctx, GA_E_HIP, "graph instantiation: %s", hipGetErrorString(ce));
        fb.graph = ex;
        fb.graph_key = key;
    }
    hipLaunchKernelGGL(ring_init, dim3((n + 255) / 256), dim3(256), 0, st, R, n);
    HIPCHK(ctx, hipGetLastError());
    // every batch advances each live ray by at least one step, so the loop
    // ends; the bound guards against a ring that never drains
    const long long max_batches = 2 + ((long long)A.max_steps * (A.bounces + 1) + (long long)K * R.slice_steps - 1) /
                                          ((long long)K * R.slice_steps) * (((long long)n + lanes - 1) / lanes);
    for (long long b = 0;; ++b) {
        if (b >= max_batches) return fail(ctx, GA_E_HIP, "slice scheduler did not drain after %lld batches", b);
        HIPCHK(ctx, hipGraphLaunch(fb.graph, st));
        HIPCHK(ctx, hipMemcpyAsync(fb.h_ctl, fb.ctl, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        if (fb.h_ctl[0] == fb.h_ctl[1]) break;
    }
    long long slots = (long long)A.n_rays * (A.bounces + 1);
    int agrid = (int)((slots + 255) / 256);
    hipLaunchKernelGGL((arrival_kernel<G, NEQ>), dim3(agrid), dim3(256), 0, st, A);
    HIPCHK(ctx, hipGetLastError());
    return GA_OK;
}

// Set by the request server around its fan launches (one ray per lane).
thread_local int serve_one_shot = 0;

// Inverse spacing of an evenly spaced axis (nodes within 1e-6 of a spacing
// of the uniform positions), else 0; the segment search uses it only as a
// first guess, so any value gives the same segment.
double axis_inv_spacing(const std::vector<double> &v)
{
    const int n = (int)v.size();
    if (n < 2) return 0.0;
    const double d = (v[n - 1] - v[0]) / (n - 1);
    if (!(d > 0.0)) return 0.0;
    for (int i = 0; i < n; ++i)
        if (std::fabs(v[i] - (v[0] + i * d)) > 1e-6 * d) return 0.0;
    return 1.0 / d;
}

// Device views of a loaded hybrid spline and of its terrain.
RdGrid rd_grid_of(const DevRd &R)
{
    RdGrid g;
    g.x = R.d_axes;
    g.y = R.d_axes + R.nx;
    g.z = R.d_axes + R.nx + R.ny;
    g.nx = R.nx;
    g.ny = R.ny;
    g.nz = R.nz;
    g.tab = R.d_tab;
    const std::vector<double> *ax[3] = {&R.hx, &R.hy, &R.hz};
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = ax[d]->front();
        g.hi[d] = ax[d]->back();
        g.iv[d] = axis_inv_spacing(*ax[d]);
    }
    return g;
}

TopoGrid topo_grid_of(const DevRd &R)
{
    TopoGrid t;
    std::memset(&t, 0, sizeof(t));
    if (!R.d_ttab) return t;
    t.x = R.d_taxes;
    t.y = R.d_taxes + R.tnx;
    t.nx = R.tnx;
    t.ny = R.tny;
    t.tab = R.d_ttab;
    const std::vector<double> *ax[2] = {&R.thx, &R.thy};
    for (int d = 0; d < 2; ++d) {
        t.lo[d] = ax[d]->front();
        t.hi[d] = ax[d]->back();
        t.iv[d] = axis_inv_spacing(*ax[d]);
    }
    return t;
}

// Region and ground of a range-dependent atmosphere (set_limits,
// atmo_io.sph.rngdep.cpp:27-43, and set_bndlyr, atmo_state.sph.rngdep.cpp:
// 42-55); z_grnd (NaN: profile bottom) applies without terrain only
// (main.infraga.sph.rngdep.cpp:294-300).
struct RdRegion {
    double lat_min, lat_max, lon_min, lon_max, alt_max, z0, z_max, z_bndlyr;
};

bool rd_region(const DevRd &R, double z_grnd, RdRegion &G)
{
    const double buffer = R.cart ? 1.0e-3 : 1.0e-3 * (PI / 180.0);   // atmo_io.3d.rngdep.cpp:25
    G.lat_min = R.hx.front() + buffer;
    G.lat_max = R.hx.back() - buffer;
    G.lon_min = R.hy.front() + buffer;
    G.lon_max = R.hy.back() - buffer;
    G.alt_max = R.hz.back();
    G.z0 = R.hz.front();
    if (R.d_ttab) {
        G.lat_min = std::max(G.lat_min, R.thx.front() + buffer);
        G.lat_max = std::min(G.lat_max, R.thx.back() - buffer);
        G.lon_min = std::max(G.lon_min, R.thy.front() + buffer);
        G.lon_max = std::min(G.lon_max, R.thy.back() - buffer);
        G.z_max = R.t_max;
        if (R.strat) {
            // the stratified set_bndlyr raises z_max before adding 2 km
            // (atmo_state.sph.strat.cpp:42-54; atmo_state.3d.strat.cpp:36-48: +0.1)
            G.z_max += R.cart ? 0.1 : 0.01;
            G.z_bndlyr = G.z_max + 2.0;
        } else {
            G.z_bndlyr = G.z_max + 2.0;
            G.z_max += 0.01;
        }
        return !(G.lat_max < G.lat_min || G.lon_max < G.lon_min);
    }
    if (!std::isnan(z_grnd)) G.z0 = z_grnd;
    G.z_max = G.z0;
    G.z_bndlyr = G.z0 + 2.0;
    return true;
}

// Range-dependent fan: medium and region as set_limits / set_bndlyr derive
// them (atmo_io.sph.rngdep.cpp:27-43, atmo_state.sph.rngdep.cpp:42-55), the
// source and the reference's z_grnd handling from run_prop
// (main.infraga.sph.rngdep.cpp:229-315).
// Medium, limits and source of a range-dependent run (run_prop /
// run_eig_search set-up, main.infraga.sph.rngdep.cpp:229-315, 877-880).
int rd_setup(ga_ctx *ctx, FanBufs &fb, int rd_id, const ga_prop_params *p, RdFanArgs &A, hipStream_t st)
{
    auto it = ctx->rds.find(rd_id);
    if (it == ctx->rds.end()) return fail(ctx, GA_E_ATMO, "unknown range-dependent atmosphere id %d", rd_id);
    const DevRd &R = it->second;
    const bool cart = p->geometry == GA_GEOM_3D_RNGDEP;
    if (cart != R.cart)
        return fail(ctx, GA_E_ARG, cart ? "GA_GEOM_3D_RNGDEP needs a ga_rngdep_load_xy atmosphere"
                                        : "a ga_rngdep_load_xy atmosphere needs GA_GEOM_3D_RNGDEP");
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    if (!(p->ds_min > 0.0) || !(p->ds_max >= p->ds_min) || !(p->s_max > 0.0))
        return fail(ctx, GA_E_ARG, "invalid step parameters ds_min=%g ds_max=%g s_max=%g", p->ds_min, p->ds_max, p->s_max);
    std::memset(&A, 0, sizeof(A));
    RdMed &M = A.M;
    M.g = rd_grid_of(R);
    M.t = topo_grid_of(R);
    M.is_topo = R.d_ttab ? 1 : 0;
    M.use_blw = p->topo_use_blw ? 1 : 0;
    M.cart = cart ? 1 : 0;
    // topo::bndlyr_param / topo::vert_wind_grad (atmo_state.sph.rngdep.cpp:95-96,
    // atmo_state.3d.rngdep.cpp:81-82, atmo_state.3d.strat.cpp:85-86)
    M.bl_sg = cart ? 0.1 / 1.0 : 0.1 / 0.5;
    M.strat = R.strat ? 1 : 0;
    M.r0 = cart ? 0.0 : p->r_earth;
    M.tweak = std::max(0.0, p->abs_coeff);
    M.sbc[0] = std::pow(10.0, -0.67887);
    M.sbc[1] = std::pow(10.0, -0.10744);
    M.sbc[2] = std::pow(10.0, -3.3979);
    RdRegion G;
    if (!rd_region(R, p->z_grnd, G)) return fail(ctx, GA_E_ARG, "atmosphere grid and topography grid have no overlap");
    const double lat_min = G.lat_min, lat_max = G.lat_max, lon_min = G.lon_min, lon_max = G.lon_max;
    M.z0 = G.z0;
    M.zflat = cart ? 0.0 : G.z0;   // the Cartesian topo::z ignores z0 (atmo_state.3d.rngdep.cpp:45-52)
    M.z_max = G.z_max;
    M.z_bndlyr = G.z_bndlyr;
    Limits &L = A.L;
    L.s_max = p->s_max;
    L.ds_min = p->ds_min;
    L.ds_max = p->ds_max;
    L.alt_max = std::isnan(p->alt_max) ? R.hz.back() : p->alt_max;
    L.rng_max = p->rng_max;
    L.time_max = p->time_max;
    L.z_keep = M.z_max + 4.0 * p->ds_max;
    const double deg = cart ? 1.0 : PI / 180.0;   // bounds and source in degrees (sph) or km (3D)
    L.b0 = std::isnan(p->bounds[0]) ? lat_min : p->bounds[0] * deg;
    L.b1 = std::isnan(p->bounds[1]) ? lat_max : p->bounds[1] * deg;
    L.b2 = std::isnan(p->bounds[2]) ? lon_min : p->bounds[2] * deg;
    L.b3 = std::isnan(p->bounds[3]) ? lon_max : p->bounds[3] * deg;
    if (!cart) {
        double q = p->rng_max / (2.0 * M.r0);
        double sq = q >= PI / 2.0 ? 2.0 : std::sin(q);
        L.gc_thresh = sq >= 1.0 ? 2.0 : sq * sq;
    }
    // source: region centre by default, at max(z_src, terrain) (:242-243, :311-315)
    A.src1 = std::isnan(p->src[0]) ? (lat_max + lat_min) / 2.0 : p->src[0] * deg;
    A.src2 = std::isnan(p->src[1]) ? (lon_max + lon_min) / 2.0 : p->src[1] * deg;
    double zsrc = p->src[2];
    if (M.is_topo) {
        if (!fb.scalar && hipMalloc(&fb.scalar, sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "scalar");
        hipLaunchKernelGGL(rd_ground_kernel, dim3(1), dim3(64), 0, st, M, A.src1, A.src2, fb.scalar);
        HIPCHK(ctx, hipGetLastError());
        double zg = 0.0;
        HIPCHK(ctx, hipMemcpyAsync(&zg, fb.scalar, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        zsrc = std::max(zsrc, zg);
    } else {
        zsrc = std::max(zsrc, M.zflat);   // z_src = max(z_src, topo::z) (main.infraga.3d.rngdep.cpp:308)
    }
    A.src0 = M.r0 + zsrc;
    A.freq = p->freq;
    A.turn_ht_min = p->turn_ht_min;
    A.bounces = p->bounces;
    long long rows = (long long)(p->s_max / p->ds_min);
    A.max_steps = (int)std::min<long long>(rows - 1, 2000000000LL);
    return GA_OK;
}

int rd_fan_enqueue(ga_ctx *ctx, FanBufs &fb, int rd_id, const ga_prop_params *p, const double *d_launches, int n,
                   ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, hipStream_t st,
                   const PathOut *path = nullptr)
{
    RdFanArgs A;
    int rc = rd_setup(ctx, fb, rd_id, p, A, st);
    if (rc != GA_OK) return rc;
    A.freq = p->freq;
    A.turn_ht_min = p->turn_ht_min;
    A.bounces = p->bounces;
    A.n_rays = n;
    long long rows = (long long)(p->s_max / p->ds_min);
    A.max_steps = (int)std::min<long long>(rows - 1, 2000000000LL);
    A.launches = d_launches;
    {
        size_t need = (size_t)std::max(n, 1) * (size_t)(p->bounces + 1) * RAW;
        if (need > fb.raw_cap) {
            HIPCHK(ctx, hipStreamSynchronize(st));
            HIPCHK(ctx, hipFree(fb.raw));
            fb.raw = nullptr;
            fb.raw_cap = 0;
            if (hipMalloc(&fb.raw, need * sizeof(double)) != hipSuccess)
                return fail(ctx, GA_E_NOMEM, "raw arrival buffer (%zu doubles)", need);
            fb.raw_cap = need;
        }
        A.raw = fb.raw;
    }
    A.out = d_out;
    A.counts = d_counts;
    A.steps = d_steps;
    if (!fb.queue && hipMalloc(&fb.queue, sizeof(unsigned int)) != hipSuccess)
        return fail(ctx, GA_E_NOMEM, "queue counter");
    A.queue = fb.queue;
    A.one_shot = path ? 0 : serve_one_shot;
    if (path) A.path = *path;
    if (n == 0) return GA_OK;

    const bool amp = p->calc_amp != 0;
    const bool blw = A.M.is_topo && A.M.use_blw;
    using RdKern = void (*)(RdFanArgs);
    // [amp][one_shot][blw]
    const RdKern kerns[2][2][2] = {
        {{rd_fan_kernel<6, false, false>, rd_fan_kernel<6, false, true>},
         {rd_fan_kernel<6, true, false>, rd_fan_kernel<6, true, true>}},
        {{rd_fan_kernel<18, false, false>, rd_fan_kernel<18, false, true>},
         {rd_fan_kernel<18, true, false>, rd_fan_kernel<18, true, true>}}};
    // [amp][blw], write mode
    const RdKern pkerns[2][2] = {{rd_fan_kernel<6, false, false, true>, rd_fan_kernel<6, false, true, true>},
                                 {rd_fan_kernel<18, false, false, true>, rd_fan_kernel<18, false, true, true>}};
    // 3D Cartesian: [amp][one_shot][blw], write mode [amp][blw]
    const RdKern ckerns[2][2][2] = {
        {{rd_fan_kernel<6, false, false, false, true>, rd_fan_kernel<6, false, true, false, true>},
         {rd_fan_kernel<6, true, false, false, true>, rd_fan_kernel<6, true, true, false, true>}},
        {{rd_fan_kernel<18, false, false, false, true>, rd_fan_kernel<18, false, true, false, true>},
         {rd_fan_kernel<18, true, false, false, true>, rd_fan_kernel<18, true, true, false, true>}}};
    const RdKern cpkerns[2][2] = {{rd_fan_kernel<6, false, false, true, true>, rd_fan_kernel<6, false, true, true, true>},
                                  {rd_fan_kernel<18, false, false, true, true>, rd_fan_kernel<18, false, true, true, true>}};
    const bool cart = p->geometry == GA_GEOM_3D_RNGDEP;
    const RdKern kern = cart ? (path ? cpkerns[amp ? 1 : 0][blw ? 1 : 0] : ckerns[amp ? 1 : 0][A.one_shot ? 1 : 0][blw ? 1 : 0])
                             : (path ? pkerns[amp ? 1 : 0][blw ? 1 : 0] : kerns[amp ? 1 : 0][A.one_shot ? 1 : 0][blw ? 1 : 0]);
    int per_cu = 0;
    constexpr int W = GA_WG_RD;
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, W, 0));
    if (per_cu < 1) per_cu = 1;
    int grid = per_cu * ctx->n_cu;
    int need = (n + W - 1) / W;
    grid = A.one_shot ? std::max(1, need) : std::max(1, std::min(grid, need));
    A.nlanes = grid * W;
    size_t ypn = (size_t)A.nlanes * 18;
    if (ypn > fb.yp_cap) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        HIPCHK(ctx, hipFree(fb.yp));
        fb.yp = nullptr;
        fb.yp_cap = 0;
        if (hipMalloc(&fb.yp, ypn * sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "lane buffer");
        fb.yp_cap = ypn;
    }
    A.ypbuf = fb.yp;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, A.queue);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(W), 0, st, A);
    HIPCHK(ctx, hipGetLastError());
    long long slots = (long long)n * (p->bounces + 1);
    int agrid = (int)((slots + 255) / 256);
    if (cart) {
        if (amp) hipLaunchKernelGGL((rd_arrival_kernel<18, true>), dim3(agrid), dim3(256), 0, st, A);
        else hipLaunchKernelGGL((rd_arrival_kernel<6, true>), dim3(agrid), dim3(256), 0, st, A);
    } else if (amp) hipLaunchKernelGGL(rd_arrival_kernel<18>, dim3(agrid), dim3(256), 0, st, A);
    else hipLaunchKernelGGL(rd_arrival_kernel<6>, dim3(agrid), dim3(256), 0, st, A);
    HIPCHK(ctx, hipGetLastError());
    return GA_OK;
}

// 2D fan (infraga-2d run_prop): always the write_rays bookkeeping; path
// records only for the flags of `path` (nullptr: none).
int fan2d_enqueue(ga_ctx *ctx, FanBufs &fb, int atmo_id, const ga_prop_params *p, const double *d_launches, int n,
                  ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, hipStream_t st, const PathOut *path)
{
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    if (!(p->ds_min > 0.0) || !(p->ds_max >= p->ds_min) || !(p->s_max > 0.0))
        return fail(ctx, GA_E_ARG, "invalid step parameters ds_min=%g ds_max=%g s_max=%g", p->ds_min, p->ds_max, p->s_max);
    const DevAtmo &a = it->second;
    const bool topo = a.d_topo2d != nullptr;
    Fan2DArgs A;
    std::memset(&A, 0, sizeof(A));
    Medium &M = A.M;
    M.tab = a.d_tab;
    M.n = a.n;
    M.zlo = a.z.front();
    M.zhi = a.z.back();
    M.r0 = 0.0;
    // set_limits / set_bndlyr (atmo_io.2d.cpp:23-33, atmo_state.2d.cpp:34-45);
    // z_grnd= only without terrain (main.infraga.2d.cpp:192-198)
    M.z0 = (!topo && !std::isnan(p->z_grnd)) ? p->z_grnd : a.z.front();
    if (topo) {
        double zm = 0.0;
        for (double zz : a.topo_z) zm = std::max(zm, zz);
        M.z_max = zm + 0.1;
    } else {
        M.z_max = M.z0;
    }
    M.zg = M.z0;
    M.z_bndlyr = M.z_max + 2.0;
    M.tweak = std::max(0.0, p->abs_coeff);
    M.sbc[0] = std::pow(10.0, -0.67887);
    M.sbc[1] = std::pow(10.0, -0.10744);
    M.sbc[2] = std::pow(10.0, -3.3979);
    A.T.tab = a.d_topo2d;
    A.T.n = topo ? (int)a.topo_r.size() : 0;
    A.s_max = p->s_max;
    A.ds_min = p->ds_min;
    A.ds_max = p->ds_max;
    A.alt_max = std::isnan(p->alt_max) ? a.z.back() : p->alt_max;
    A.rng_max = !std::isnan(p->rng_max) ? p->rng_max : (topo ? a.topo_r.back() : 2000.0);
    A.z_reflect = std::isnan(p->refl_alt) ? a.z.back() + 1.0 : p->refl_alt;
    // source at r = 0, raised to the ground there (:222)
    double zg0 = M.z0;
    if (topo) {
        // natural spline of the terrain at r = 0 on the host (same Hermite form)
        const std::vector<double> &r = a.topo_r, &zt = a.topo_z;
        std::vector<double> sl;
        spline_slopes(r, zt, sl);
        double rr = std::min(std::max(0.0, r.front()), r.back());
        size_t k = 0;
        while (k + 2 < r.size() && rr > r[k + 1]) ++k;
        const double dx = r[k + 1] - r[k], X = (rr - r[k]) / dx, d = zt[k + 1] - zt[k];
        const double Aa = sl[k] * dx - d, B = -sl[k + 1] * dx + d;
        zg0 = zt[k] + X * (d + (1.0 - X) * (Aa * (1.0 - X) + B * X));
    }
    A.z_src = std::max(p->src[2], zg0);
    A.c_src = host_c_at(a, A.z_src);
    {
        // rho at the source: the same spline through the host copies
        std::vector<double> sl;
        spline_slopes(a.z, a.rho, sl);
        const double zz = std::min(std::max(A.z_src, a.z.front()), a.z.back());
        size_t k = 0;
        while (k + 2 < a.z.size() && zz > a.z[k + 1]) ++k;
        const double dx = a.z[k + 1] - a.z[k], X = (zz - a.z[k]) / dx, d = a.rho[k + 1] - a.rho[k];
        const double Aa = sl[k] * dx - d, B = -sl[k + 1] * dx + d;
        A.rho_src = a.rho[k] + X * (d + (1.0 - X) * (Aa * (1.0 - X) + B * X));
    }
    A.freq = p->freq;
    A.bounces = p->bounces;
    A.n_rays = n;
    long long rows = (long long)(p->s_max / p->ds_min);
    A.max_steps = (int)std::min<long long>(rows - 1, 2000000000LL);
    A.launches = d_launches;
    A.out = d_out;
    A.counts = d_counts;
    A.steps = d_steps;
    if (path) A.path = *path;
    if (!fb.queue && hipMalloc(&fb.queue, sizeof(unsigned int)) != hipSuccess)
        return fail(ctx, GA_E_NOMEM, "queue counter");
    A.queue = fb.queue;
    if (n == 0) return GA_OK;
    const size_t lds = (size_t)a.n * NODE * sizeof(double);
    const bool use_lds = lds <= ctx->max_lds;
    auto kern = p->calc_amp ? (use_lds ? fan2d_kernel<8, true> : fan2d_kernel<8, false>)
                            : (use_lds ? fan2d_kernel<4, true> : fan2d_kernel<4, false>);
    const size_t shm = use_lds ? lds : 0;
    if (shm > 65536) HIPCHK(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    int per_cu = 0;
    HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, GA_WG_2D, shm));
    per_cu = std::max(per_cu, 1);
    const int need = (n + GA_WG_2D - 1) / GA_WG_2D;
    const int grid = std::max(1, std::min(per_cu * ctx->n_cu, need));
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, A.queue);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(GA_WG_2D), shm, st, A);
    HIPCHK(ctx, hipGetLastError());
    return GA_OK;
}

// path: write-mode bookkeeping with path records (nullptr: write_rays=false sums).
int fan_enqueue(ga_ctx *ctx, FanBufs &fb, int atmo_id, const ga_prop_params *p, const double *d_launches, int n,
                ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, hipStream_t st,
                const PathOut *path = nullptr)
{
    if (p->geometry == GA_GEOM_SPH_RNGDEP || p->geometry == GA_GEOM_3D_RNGDEP)
        return rd_fan_enqueue(ctx, fb, atmo_id, p, d_launches, n, d_out, d_counts, d_steps, st, path);
    if (p->geometry == GA_GEOM_2D)
        return fan2d_enqueue(ctx, fb, atmo_id, p, d_launches, n, d_out, d_counts, d_steps, st, path);
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    if (p->geometry != GA_GEOM_SPH && p->geometry != GA_GEOM_3D) return fail(ctx, GA_E_ARG, "bad geometry %d", p->geometry);
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    FanArgs A;
    std::memset(&A, 0, sizeof(A));
    int rc = GA_OK;
    if (!build_medium(ctx, it->second, p, A.M, A.L, rc, !fb.fixed)) return rc;
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
        if (need > fb.raw_cap) {
            if (fb.fixed) return fail(ctx, GA_E_CAP, "raw arrival records of %zu doubles exceed the fixed %zu", need, fb.raw_cap);
            HIPCHK(ctx, hipStreamSynchronize(st));
            HIPCHK(ctx, hipFree(fb.raw));
            fb.raw = nullptr;
            fb.raw_cap = 0;
            if (hipMalloc(&fb.raw, need * sizeof(double)) != hipSuccess)
                return fail(ctx, GA_E_NOMEM, "raw arrival buffer (%zu doubles)", need);
            fb.raw_cap = need;
        }
        A.raw = fb.raw;
    }
    A.out = d_out;
    A.counts = d_counts;
    A.steps = d_steps;
    if (!fb.queue && (fb.fixed || hipMalloc(&fb.queue, sizeof(unsigned int)) != hipSuccess))
        return fail(ctx, GA_E_NOMEM, "queue counter");
    A.queue = fb.queue;
    A.one_shot = fb.fixed ? 0 : serve_one_shot;
    if (path) A.path = *path;
    const bool pw = path != nullptr;
    size_t lds = (size_t)a.n * NODE * sizeof(double);
    A.lds = lds <= ctx->max_lds ? 1 : 0;
    if (n == 0) return GA_OK;
    // the rotating slice scheduler for plain fans; the persistent ray queue
    // for write-mode fans, one-ray-per-lane server launches and the fixed
    // buffers of the no-allocation entry point
    const int ts = ctx->tuned_slice[sph][p->calc_amp ? 1 : 0], tk = ctx->tuned_graph[sph][p->calc_amp ? 1 : 0];
    const int slice = env_int("GA_FAN_SLICE", ts > 0 ? ts : SLICE_STEPS);
    const int gk = env_int("GA_FAN_GRAPH", tk > 0 ? tk : SLICES_PER_GRAPH);
    if (!pw && !A.one_shot && !fb.fixed && slice > 0) {
        if (sph) return p->calc_amp ? launch_fan_sliced<GEOM_SPH, 18>(ctx, fb, A, st, slice, gk)
                                    : launch_fan_sliced<GEOM_SPH, 6>(ctx, fb, A, st, slice, gk);
        return p->calc_amp ? launch_fan_sliced<GEOM_3D, 18>(ctx, fb, A, st, slice, gk)
                           : launch_fan_sliced<GEOM_3D, 6>(ctx, fb, A, st, slice, gk);
    }
    if (sph) return p->calc_amp ? launch_fan<GEOM_SPH, 18>(ctx, fb, A, st, pw) : launch_fan<GEOM_SPH, 6>(ctx, fb, A, st, pw);
    return p->calc_amp ? launch_fan<GEOM_3D, 18>(ctx, fb, A, st, pw) : launch_fan<GEOM_3D, 6>(ctx, fb, A, st, pw);
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

// Pinned host staging buffer (grow-only), so that copies queued on a
// stream run asynchronously to the host.
template <class T> struct PinBuf {
    T *p = nullptr;
    size_t cap = 0;
    bool ensure(size_t n)
    {
        if (n <= cap) return true;
        if (p && hipHostFree(p) != hipSuccess) return false;
        p = nullptr;
        cap = 0;
        if (hipHostMalloc((void **)&p, std::max(n, (size_t)1) * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = std::max(n, (size_t)1);
        return true;
    }
    ~PinBuf() { if (p) (void)hipHostFree(p); }
};

// Device side of a path-record output (PathOut): chunk storage for `cap`
// records plus one partly filled chunk per item, the chunk tags and the
// per-item counts.  gather() puts the records of the used chunks in item
// order on the host.
struct PathDev {
    DevBuf<ga_path_pt> pts;
    DevBuf<int32_t> item, seq, fill, item_pts;
    DevBuf<unsigned int> ctr;
    unsigned int max_chunks = 0;
    int n_items = 0;

    int setup(ga_ctx *ctx, int64_t cap, int items, int flags, hipStream_t st, PathOut &P)
    {
        const long long chunks = (std::max<int64_t>(cap, 0) + PCHUNK - 1) / PCHUNK + (long long)items;
        if (chunks >= 0x7fffffffLL) return fail(ctx, GA_E_ARG, "path capacity too large (%lld chunks)", chunks);
        max_chunks = (unsigned int)std::max(chunks, 1LL);
        n_items = items;
        if (!pts.ensure((size_t)max_chunks * PCHUNK) || !item.ensure(max_chunks) || !seq.ensure(max_chunks) ||
            !fill.ensure(max_chunks) || !item_pts.ensure((size_t)std::max(items, 1)) || !ctr.ensure(1))
            return fail(ctx, GA_E_NOMEM, "path buffers (%u chunks of %d records)", max_chunks, PCHUNK);
        hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, ctr.p);
        HIPCHK(ctx, hipGetLastError());
        P.pts = pts.p;
        P.chunk_item = item.p;
        P.chunk_seq = seq.p;
        P.chunk_fill = fill.p;
        P.item_pts = item_pts.p;
        P.ctr = ctr.p;
        P.max_chunks = max_chunks;
        P.flags = flags;
        return GA_OK;
    }

    // After the stream has finished: *total = records produced; when they
    // fit in cap (and dst is given) they are written to dst in item order.
    int gather(ga_ctx *ctx, ga_path_pt *dst, int64_t cap, int64_t *total)
    {
        unsigned int used = 0;
        HIPCHK(ctx, hipMemcpy(&used, ctr.p, sizeof(unsigned int), hipMemcpyDeviceToHost));
        std::vector<int32_t> cnt((size_t)std::max(n_items, 1), 0);
        if (n_items > 0)
            HIPCHK(ctx, hipMemcpy(cnt.data(), item_pts.p, (size_t)n_items * sizeof(int32_t), hipMemcpyDeviceToHost));
        std::vector<int64_t> off((size_t)n_items + 1, 0);
        for (int i = 0; i < n_items; ++i) off[i + 1] = off[i] + cnt[i];
        *total = off[n_items];
        if (*total > cap || used > max_chunks)
            return fail(ctx, GA_E_CAP, "need %lld path records, have %lld", (long long)*total, (long long)cap);
        if (*total == 0) return GA_OK;
        if (!dst) return fail(ctx, GA_E_ARG, "null path buffer");
        std::vector<int32_t> ci(used), cs(used), cf(used);
        std::vector<ga_path_pt> buf((size_t)used * PCHUNK);
        HIPCHK(ctx, hipMemcpy(ci.data(), item.p, used * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(cs.data(), seq.p, used * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(cf.data(), fill.p, used * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(buf.data(), pts.p, buf.size() * sizeof(ga_path_pt), hipMemcpyDeviceToHost));
        for (unsigned int c = 0; c < used; ++c) {
            const int it = ci[c];
            if (it < 0 || it >= n_items || cf[c] < 0 || cf[c] > PCHUNK)
                return fail(ctx, GA_E_HIP, "corrupt path chunk %u (item %d, fill %d)", c, it, cf[c]);
            const int64_t at = off[it] + (int64_t)cs[c] * PCHUNK;
            if (at + cf[c] > off[it + 1]) return fail(ctx, GA_E_HIP, "path chunk %u overruns item %d", c, it);
            std::memcpy(dst + at, buf.data() + (size_t)c * PCHUNK, (size_t)cf[c] * sizeof(ga_path_pt));
        }
        return GA_OK;
    }
};

// GPU side of the eigenray search: one round of traces per enqueue().
// EIG_SWEEP jobs run on stream st in the 6-equation kernel, EIG_LM/EIG_FINAL
// jobs on stream s2 in the 18-equation kernel, followed there by the arrival
// rows of the EIG_FINAL jobs; results come back through pinned buffers and
// the event `done` marks their arrival.  One round may be in flight at a
// time per EigGpu.
struct EigGpu {
    ga_ctx *ctx = nullptr;
    int geom = GA_GEOM_3D;
    FanArgs F{};                 // medium, limits, source, c_src, freq, max_steps, lds
    bool rd = false;             // range-dependent medium: RF instead of F
    bool cart = false;           // the 3D Cartesian range-dependent medium
    RdFanArgs RF{};
    hipStream_t st = nullptr, s2 = nullptr;
    bool own_st = false;
    DevBuf<EigJob> jobs6, jobs18;
    DevBuf<double> res6, res18, yp6, yp18;
    DevBuf<int32_t> brk6, brk18, st6, st18;
    DevBuf<ga_arrival> arr;
    DevBuf<unsigned int> q;
    PinBuf<EigJob> h_jobs6, h_jobs18;
    PinBuf<double> h_res6, h_res18;
    PinBuf<int32_t> h_brk6, h_brk18, h_st6, h_st18;
    PinBuf<ga_arrival> h_arr;
    hipEvent_t e0 = nullptr, e1 = nullptr, mid = nullptr;
    double kernel_ms = 0.0;
    int launches = 0;
    size_t n6 = 0, n18 = 0, nf = 0;
    int stride = 0;

    ~EigGpu()
    {
        if (s2) { (void)hipStreamSynchronize(s2); (void)hipStreamDestroy(s2); }
        if (own_st && st) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (mid) (void)hipEventDestroy(mid);
    }

    // stream: the stream of the 6-equation kernels (nullptr: a new one)
    int init(hipStream_t stream)
    {
        if (stream) st = stream;
        else { HIPCHK(ctx, hipStreamCreate(&st)); own_st = true; }
        HIPCHK(ctx, hipStreamCreate(&s2));
        HIPCHK(ctx, hipEventCreate(&e0));
        HIPCHK(ctx, hipEventCreate(&e1));
        HIPCHK(ctx, hipEventCreate(&mid));
        if (!q.alloc(2)) return fail(ctx, GA_E_NOMEM, "eigenray queue counters");
        return GA_OK;
    }

    template <int G, int NEQ, int WGS, bool PATHS = false>
    int launch(const EigArgs &E0, DevBuf<double> &yp, unsigned int *queue, hipStream_t s)
    {
        size_t lds = F.lds ? (size_t)F.M.n * NODE * sizeof(double) : 0;
        auto kern = PATHS ? (F.lds ? eig_kernel<G, NEQ, true, WGS, false, PATHS> : eig_kernel<G, NEQ, false, WGS, false, PATHS>)
                  : E0.F.one_shot ? (F.lds ? eig_kernel<G, NEQ, true, WGS, true> : eig_kernel<G, NEQ, false, WGS, true>)
                                  : (F.lds ? eig_kernel<G, NEQ, true, WGS, false> : eig_kernel<G, NEQ, false, WGS, false>);
        if (lds > 65536) HIPCHK(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int per_cu = 0;
        HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WGS, lds));
        per_cu = std::max(per_cu, 1);
        const int need = (E0.n_jobs + WGS - 1) / WGS;
        int grid = (E0.F.one_shot && !PATHS) ? std::max(1, need) : std::max(1, std::min(per_cu * ctx->n_cu, need));
        if (!yp.ensure((size_t)grid * WGS * 18)) return fail(ctx, GA_E_NOMEM, "eigenray lane buffer");
        EigArgs E = E0;
        E.F.ypbuf = yp.p;
        E.F.nlanes = grid * WGS;
        E.F.queue = queue;
        hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, queue);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(WGS), lds, s, E);
        HIPCHK(ctx, hipGetLastError());
        launches += 2;
        return GA_OK;
    }

    // Range-dependent traces: persistent 64-lane workgroups (RK4 rows in LDS).
    template <int NEQ, bool PATHS = false>
    int launch_rd(const EigJob *jobs, int n_jobs, int stride, double *res, int32_t *brk, int32_t *stp,
                  DevBuf<double> &yp, unsigned int *queue, hipStream_t s, const PathOut *path = nullptr)
    {
        const bool blw = RF.M.is_topo && RF.M.use_blw;
        auto kern = cart ? (blw ? rd_eig_kernel<NEQ, true, PATHS, true> : rd_eig_kernel<NEQ, false, PATHS, true>)
                         : (blw ? rd_eig_kernel<NEQ, true, PATHS, false> : rd_eig_kernel<NEQ, false, PATHS, false>);
        int per_cu = 0;
        HIPCHK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, GA_WG_RD, 0));
        per_cu = std::max(per_cu, 1);
        const int need = (n_jobs + GA_WG_RD - 1) / GA_WG_RD;
        const int grid = std::max(1, std::min(per_cu * ctx->n_cu, need));
        if (!yp.ensure((size_t)grid * GA_WG_RD * 18)) return fail(ctx, GA_E_NOMEM, "eigenray lane buffer");
        RdEigArgs E{};
        E.F = RF;
        E.F.ypbuf = yp.p;
        E.F.nlanes = grid * GA_WG_RD;
        E.F.queue = queue;
        if (path) E.F.path = *path;
        E.jobs = jobs;
        E.n_jobs = n_jobs;
        E.stride = stride;
        E.res = res;
        E.brk_leg = brk;
        E.steps = stp;
        hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, s, queue);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(GA_WG_RD), 0, s, E);
        HIPCHK(ctx, hipGetLastError());
        launches += 2;
        return GA_OK;
    }

    int enqueue(const EigBatch &B)
    {
        n6 = B.sweep.size();
        n18 = B.full.size();
        nf = n18 - (size_t)B.n_lm;
        stride = B.sweep_stride;
        // buffers are reused: the previous round has been collected
        if (!jobs6.ensure(n6) || !res6.ensure(n6 * stride) || !brk6.ensure(n6) || !st6.ensure(n6) ||
            !jobs18.ensure(n18) || !res18.ensure(n18 * EIG_FULL) || !brk18.ensure(n18) || !st18.ensure(n18) ||
            !arr.ensure(nf) || !h_jobs6.ensure(n6) || !h_jobs18.ensure(n18) || !h_res6.ensure(n6 * stride) ||
            !h_brk6.ensure(n6) || !h_st6.ensure(n6) || !h_res18.ensure(n18 * EIG_FULL) || !h_brk18.ensure(n18) ||
            !h_st18.ensure(n18) || !h_arr.ensure(nf))
            return fail(ctx, GA_E_NOMEM, "eigenray batch buffers (%zu + %zu traces)", n6, n18);
        if (n6) std::memcpy(h_jobs6.p, B.sweep.data(), n6 * sizeof(EigJob));
        if (n18) std::memcpy(h_jobs18.p, B.full.data(), n18 * sizeof(EigJob));
        if (n6) HIPCHK(ctx, hipMemcpyAsync(jobs6.p, h_jobs6.p, n6 * sizeof(EigJob), hipMemcpyHostToDevice, st));
        if (n18) HIPCHK(ctx, hipMemcpyAsync(jobs18.p, h_jobs18.p, n18 * sizeof(EigJob), hipMemcpyHostToDevice, st));
        HIPCHK(ctx, hipEventRecord(e0, st));
        HIPCHK(ctx, hipStreamWaitEvent(s2, e0, 0));
        if (n6 && rd) {
            int rc = launch_rd<6>(jobs6.p, (int)n6, stride, res6.p, brk6.p, st6.p, yp6, q.p, st);
            if (rc != GA_OK) return rc;
            HIPCHK(ctx, hipMemcpyAsync(h_res6.p, res6.p, n6 * stride * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(h_brk6.p, brk6.p, n6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(h_st6.p, st6.p, n6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        } else if (n6) {
            EigArgs E{};
            E.F = F;
            E.jobs = jobs6.p;
            E.n_jobs = (int)n6;
            E.stride = stride;
            E.res = res6.p;
            E.brk_leg = brk6.p;
            E.steps = st6.p;
            int rc = geom == GA_GEOM_SPH ? launch<GEOM_SPH, 6, GA_EIG_WG6>(E, yp6, q.p, st)
                                         : launch<GEOM_3D, 6, GA_EIG_WG6>(E, yp6, q.p, st);
            if (rc != GA_OK) return rc;
            HIPCHK(ctx, hipMemcpyAsync(h_res6.p, res6.p, n6 * stride * sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(h_brk6.p, brk6.p, n6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(ctx, hipMemcpyAsync(h_st6.p, st6.p, n6 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        }
        if (n18 && rd) {
            int rc = launch_rd<18>(jobs18.p, (int)n18, EIG_FULL, res18.p, brk18.p, st18.p, yp18, q.p + 1, s2);
            if (rc != GA_OK) return rc;
            if (nf) {
                RdEigArgs E{};
                E.F = RF;
                E.jobs = jobs18.p;
                E.n_jobs = (int)n18;
                E.stride = EIG_FULL;
                E.res = res18.p;
                hipLaunchKernelGGL(cart ? rd_eig_arrival_kernel<true> : rd_eig_arrival_kernel<false>,
                                   dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s2, E, B.n_lm, (int)nf, arr.p);
                HIPCHK(ctx, hipGetLastError());
                launches += 1;
                HIPCHK(ctx, hipMemcpyAsync(h_arr.p, arr.p, nf * sizeof(ga_arrival), hipMemcpyDeviceToHost, s2));
            }
            HIPCHK(ctx, hipMemcpyAsync(h_res18.p, res18.p, n18 * EIG_FULL * sizeof(double), hipMemcpyDeviceToHost, s2));
            HIPCHK(ctx, hipMemcpyAsync(h_brk18.p, brk18.p, n18 * sizeof(int32_t), hipMemcpyDeviceToHost, s2));
            HIPCHK(ctx, hipMemcpyAsync(h_st18.p, st18.p, n18 * sizeof(int32_t), hipMemcpyDeviceToHost, s2));
        } else if (n18) {
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
                HIPCHK(ctx, hipMemcpyAsync(h_arr.p, arr.p, nf * sizeof(ga_arrival), hipMemcpyDeviceToHost, s2));
            }
            HIPCHK(ctx, hipMemcpyAsync(h_res18.p, res18.p, n18 * EIG_FULL * sizeof(double), hipMemcpyDeviceToHost, s2));
            HIPCHK(ctx, hipMemcpyAsync(h_brk18.p, brk18.p, n18 * sizeof(int32_t), hipMemcpyDeviceToHost, s2));
            HIPCHK(ctx, hipMemcpyAsync(h_st18.p, st18.p, n18 * sizeof(int32_t), hipMemcpyDeviceToHost, s2));
        }
        HIPCHK(ctx, hipEventRecord(mid, s2));
        HIPCHK(ctx, hipStreamWaitEvent(st, mid, 0));
        HIPCHK(ctx, hipEventRecord(e1, st));
        return GA_OK;
    }

    bool ready() const { return hipEventQuery(e1) != hipErrorNotReady; }

    // Re-traces accepted eigenrays (EIG_FINAL jobs, item = index in jobs)
    // with their path rows into pd, on stream s2; synchronises.
    int trace_paths(const std::vector<EigJob> &jobs, PathDev &pd, int64_t cap)
    {
        const size_t n = jobs.size();
        if (!jobs18.ensure(n) || !res18.ensure(n * EIG_FULL) || !brk18.ensure(n) || !st18.ensure(n) ||
            !h_jobs18.ensure(n))
            return fail(ctx, GA_E_NOMEM, "eigenray path buffers (%zu traces)", n);
        PathOut P{};
        int rc = pd.setup(ctx, cap, (int)n, GA_PATH_RAYS, s2, P);
        if (rc != GA_OK) return rc;
        if (n && rd) {
            std::memcpy(h_jobs18.p, jobs.data(), n * sizeof(EigJob));
            HIPCHK(ctx, hipMemcpyAsync(jobs18.p, h_jobs18.p, n * sizeof(EigJob), hipMemcpyHostToDevice, s2));
            rc = launch_rd<18, true>(jobs18.p, (int)n, EIG_FULL, res18.p, brk18.p, st18.p, yp18, q.p + 1, s2, &P);
            if (rc != GA_OK) return rc;
        } else if (n) {
            std::memcpy(h_jobs18.p, jobs.data(), n * sizeof(EigJob));
            HIPCHK(ctx, hipMemcpyAsync(jobs18.p, h_jobs18.p, n * sizeof(EigJob), hipMemcpyHostToDevice, s2));
            EigArgs E{};
            E.F = F;
            E.F.path = P;
            E.jobs = jobs18.p;
            E.n_jobs = (int)n;
            E.stride = EIG_FULL;
            E.res = res18.p;
            E.brk_leg = brk18.p;
            E.steps = st18.p;
            rc = geom == GA_GEOM_SPH ? launch<GEOM_SPH, 18, trace_wg<GEOM_SPH, 18>(), true>(E, yp18, q.p + 1, s2)
                                     : launch<GEOM_3D, 18, trace_wg<GEOM_3D, 18>(), true>(E, yp18, q.p + 1, s2);
            if (rc != GA_OK) return rc;
        }
        HIPCHK(ctx, hipStreamSynchronize(s2));
        return GA_OK;
    }

    int collect(EigBatchOut &O)
    {
        HIPCHK(ctx, hipEventSynchronize(e1));
        float ms = 0.0f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
        kernel_ms += ms;
        O.sweep_res.assign(h_res6.p, h_res6.p + n6 * stride);
        O.sweep_brk.assign(h_brk6.p, h_brk6.p + n6);
        O.sweep_steps.assign(h_st6.p, h_st6.p + n6);
        O.full_res.assign(h_res18.p, h_res18.p + n18 * EIG_FULL);
        O.full_brk.assign(h_brk18.p, h_brk18.p + n18);
        O.full_steps.assign(h_st18.p, h_st18.p + n18);
        O.final_rows.assign(h_arr.p, h_arr.p + nf);
        return GA_OK;
    }

    int operator()(const EigBatch &B, EigBatchOut &O)
    {
        int rc = enqueue(B);
        if (rc != GA_OK) return rc;
        return collect(O);
    }
};

// Medium, limits and search set-up of an eigenray search (shared by
// ga_eig_search and the request server).  rcvr: receiver pairs in the
// caller's units, converted in place to the search's (sph: radians).
int eig_prepare(ga_ctx *ctx, DevAtmo &a, const ga_prop_params *p, const ga_eig_params *e, FanArgs &F,
                EigSetup &S, std::vector<double> &rcvr)
{
    const bool sph = p->geometry == GA_GEOM_SPH;
    int rc = GA_OK;
    if (!build_medium(ctx, a, p, F.M, F.L, rc)) return rc;
    S.geom = p->geometry;
    S.r0 = F.M.r0;
    S.src_ground = F.M.z0;
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
    S.rngdep = 0;
    S.is_topo = 0;
    return GA_OK;
}

// Search set-up over a range-dependent medium with terrain (run_eig_search
// of main.infraga.sph.rngdep.cpp:877-965): the spherical search with the
// medium's c at the source and the terrain's lowest inclination.
int eig_prepare_rd(ga_ctx *ctx, int rd_id, const ga_prop_params *p, const ga_eig_params *e, RdFanArgs &RF,
                   EigSetup &S, std::vector<double> &rcvr)
{
    int rc = rd_setup(ctx, ctx->fb, rd_id, p, RF, ctx->stream);
    if (rc != GA_OK) return rc;
    const RdMed &M = RF.M;
    S.is_topo = M.is_topo;
    S.dth_big = e->incl_step_max * (PI / 180.0);
    S.dth_sml = e->incl_step_min * (PI / 180.0);
    S.th_min = e->incl_min * (PI / 180.0);
    S.th_max = e->incl_max * (PI / 180.0);
    S.bnc_min = e->bnc_min;
    S.bnc_max = e->bnc_max;
    S.iterations = e->iterations;
    S.damping = e->damping;
    S.tolerance = e->tolerance;
    S.az_err_lim = e->az_dev_lim;
    S.rng_max = p->rng_max;
    S.spec_chunk = 64;
    S.rngdep = 1;
    if (p->geometry == GA_GEOM_3D_RNGDEP) {
        // run_eig_search / run_eig_direct of main.infraga.3d.rngdep.cpp:754-1052:
        // the 3D search of geoac.eigenray.3d.cpp from theta_min, the source
        // raised to the terrain, receivers in km
        S.geom = GA_GEOM_3D;
        S.r0 = 0.0;
        S.src_ground = M.zflat;
        S.src[0] = RF.src1;
        S.src[1] = RF.src2;
        S.src[2] = RF.src0;
        double pt[3] = {RF.src1, RF.src2, RF.src0}, aout[40], tout[10];
        rc = ga_rngdep_eval(ctx, rd_id, pt, 1, aout, tout);
        if (rc != GA_OK) return rc;
        S.c_src = aout[0];             // atmo::c(src[0], src[1], src[2]) (geoac.eigenray.3d.cpp:334)
        S.c_grnd = 0.0;                // read per arrival from the trace result
        S.th_start = S.th_min;
        return GA_OK;
    }
    S.geom = GA_GEOM_SPH;
    S.r0 = M.r0;
    S.src_ground = M.z0;
    const double zsrc = RF.src0 - M.r0;
    S.src[0] = zsrc;
    S.src[1] = RF.src1;
    S.src[2] = RF.src2;
    for (double &x : rcvr) x *= (PI / 180.0);
    double pt[3] = {RF.src1, RF.src2, zsrc}, aout[40], tout[10];
    rc = ga_rngdep_eval(ctx, rd_id, pt, 1, aout, tout);
    if (rc != GA_OK) return rc;
    S.c_src = aout[0];                 // atmo::c(src[0] + r0, src[1], src[2]) (geoac.eigenray.sph.cpp:338)
    S.c_grnd = 0.0;                    // read per arrival from the trace result
    const double zg = M.is_topo ? tout[0] : M.z0;
    double th_grnd;
    if (zsrc - zg < 0.01) {
        if (M.is_topo) {
            // :940-946: the steepest terrain slope over azimuths 0, 3, ..., 357
            // in degrees plus 0.1, then compared with the radian bounds as it stands
            th_grnd = 0.0;
            const double dzgdx = 1.0 / (M.r0 * std::cos(RF.src1)) * tout[2];
            const double dzgdy = 1.0 / M.r0 * tout[1];
            for (double ph = 0.0; ph < 360.0; ph += 3.0)
                th_grnd = std::max(th_grnd, std::atan(dzgdx * std::sin(ph * PI / 180.0) + dzgdy * std::cos(ph * PI / 180.0)) *
                                                (180.0 / PI) + 0.1);
        } else {
            th_grnd = 0.1 * (PI / 180.0);
        }
    } else {
        th_grnd = -89.9 * (PI / 180.0);
    }
    S.th_min = e->incl_min * (PI / 180.0);
    S.th_max = e->incl_max * (PI / 180.0);
    S.th_start = std::max(S.th_min, th_grnd);
    S.bnc_min = e->bnc_min;
    S.bnc_max = e->bnc_max;
    S.iterations = e->iterations;
    S.damping = e->damping;
    S.tolerance = e->tolerance;
    S.az_err_lim = e->az_dev_lim;
    S.dth_big = e->incl_step_max * (PI / 180.0);
    S.dth_sml = e->incl_step_min * (PI / 180.0);
    S.rng_max = p->rng_max;
    S.spec_chunk = 64;
    S.rngdep = 1;
    return GA_OK;
}

int eig_check(ga_ctx *ctx, const ga_prop_params *p, const ga_eig_params *e)
{
    if (p->geometry != GA_GEOM_3D && p->geometry != GA_GEOM_SPH && p->geometry != GA_GEOM_SPH_RNGDEP &&
        p->geometry != GA_GEOM_3D_RNGDEP)
        return fail(ctx, GA_E_ARG, "bad geometry %d", p->geometry);
    if (e->bnc_min < 0 || e->bnc_max < e->bnc_min) return fail(ctx, GA_E_ARG, "bad bounce range %d..%d", e->bnc_min, e->bnc_max);
    if (e->iterations < 0) return fail(ctx, GA_E_ARG, "negative iteration limit");
    if (!(e->incl_step_max > 0.0) || !(e->incl_step_min > 0.0))
        return fail(ctx, GA_E_ARG, "inclination steps must be positive");
    return GA_OK;
}

}  // namespace


// ===========================================================================
// Request server (continuous batching)
// ===========================================================================
namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct EigGroup;

struct SrvReq {
    uint64_t ticket = 0;
    ga_request req{};
    std::vector<double> data;
    bool done = false;
    int status = GA_OK;
    std::vector<ga_arrival> rows;
    long long ray_steps = 0;
    Clock::time_point t_submit, t_admit, t_done;
    double footprint = 0.0;
    EigGroup *group = nullptr;   // eigenray requests: the search they joined
    int r0 = 0;                  // first receiver index inside that search
};

// One running eigenray search shared by the requests that joined it.
struct EigGroup {
    int atmo = 0;
    ga_prop_params p{};
    ga_eig_params e{};
    FanArgs F{};
    EigSetup S{};
    std::unique_ptr<EigSearch> search;
    std::vector<SrvReq *> owner;        // receiver index -> request
    std::vector<int> left;              // per request in `members`: receivers not finished
    std::vector<SrvReq *> members;
    EigBatch batch;
    EigSearchStats st;
    bool inflight = false;
    bool closed = false;                // takes no new receivers
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr, start = nullptr;
    FanBufs fb;
    DevBuf<double> d_l;
    DevBuf<ga_arrival> d_o;
    DevBuf<int32_t> d_c, d_s;
    PinBuf<ga_arrival> h_o;
    PinBuf<int32_t> h_c, h_s;
    std::unique_ptr<EigGpu> eg;
    int kind = -1;                      // -1 idle, GA_REQ_FAN, GA_REQ_EIG
    SrvReq *fan = nullptr;
    EigGroup *group = nullptr;
};

bool same_bytes(const void *a, const void *b, size_t n) { return std::memcmp(a, b, n) == 0; }

}  // namespace

struct ga_server {
    ga_ctx *ctx = nullptr;
    double budget = 0.0;
    int flags = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_done, cv_work;
    std::deque<SrvReq *> submitted;                      // guarded by mu
    std::unordered_map<uint64_t, SrvReq *> reqs;         // guarded by mu
    uint64_t next_ticket = 1;
    bool stop = false;
    // scheduler thread only
    std::deque<SrvReq *> pending, fans;
    std::vector<std::unique_ptr<EigGroup>> groups;
    int active = 0;
    double inflight = 0.0;
    size_t rr = 0;
    // totals (guarded by mu)
    long long steps = 0;
    int launches = 0;
    double kernel_ms = 0.0;

    // Device bytes a request holds while admitted: a fan's launch list, its
    // output slots (arrival rows and raw records), per-ray counts and the
    // row k-1 buffer of its one-ray-per-lane launch (lanes rounded up to a
    // 512-lane workgroup), plus 1 MB of small buffers; an eigenray query
    // about 1 MB per receiver (sweep and Levenberg-Marquardt trace jobs and
    // their results) plus 16 MB.  The atmosphere tables are resident and
    // counted when loaded, not per request.
    double footprint(const SrvReq &r) const
    {
        if (r.req.kind == GA_REQ_FAN) {
            const double n = (double)r.req.n;
            const double slots_ = n * (r.req.prop.bounces + 1);
            const double lanes = std::ceil(n / 512.0) * 512.0;
            return n * 2.0 * sizeof(double) + slots_ * (sizeof(ga_arrival) + RAW * sizeof(double)) +
                   n * 2.0 * sizeof(int32_t) + lanes * 18.0 * sizeof(double) + 1.0e6;
        }
        return r.req.n * 1.0e6 + 16.0e6;
    }

    void finish(SrvReq *r, int status)
    {
        r->t_done = Clock::now();
        inflight -= r->footprint;
        active--;
        std::lock_guard<std::mutex> lk(mu);
        r->status = status;
        r->done = true;
        cv_done.notify_all();
    }

    void admit(SrvReq *r)
    {
        r->t_admit = Clock::now();
        r->footprint = footprint(*r);
        inflight += r->footprint;
        active++;
        if (r->req.kind == GA_REQ_FAN) { fans.push_back(r); return; }
        auto it = ctx->atmos.find(r->req.atmo_id);
        if (it == ctx->atmos.end()) { finish(r, fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", r->req.atmo_id)); return; }
        int rc = eig_check(ctx, &r->req.prop, &r->req.eig);
        if (rc != GA_OK) { finish(r, rc); return; }
        EigGroup *g = nullptr;
        if (flags & GA_SERVE_MERGE)
            for (auto &q : groups)
                if (!q->closed && q->atmo == r->req.atmo_id && same_bytes(&q->p, &r->req.prop, sizeof(ga_prop_params)) &&
                    same_bytes(&q->e, &r->req.eig, sizeof(ga_eig_params))) { g = q.get(); break; }
        std::vector<double> rc_xy = r->data;
        if (!g) {
            auto ng = std::make_unique<EigGroup>();
            ng->atmo = r->req.atmo_id;
            ng->p = r->req.prop;
            ng->e = r->req.eig;
            std::vector<double> none;
            rc = eig_prepare(ctx, it->second, &ng->p, &ng->e, ng->F, ng->S, none);
            if (rc != GA_OK) { finish(r, rc); return; }
            ng->F.one_shot = 1;
            ng->search = std::make_unique<EigSearch>(ng->S);
            ng->search->set_clock(Clock::now());
            g = ng.get();
            groups.push_back(std::move(ng));
        }
        if (r->req.prop.geometry == GA_GEOM_SPH)
            for (double &x : rc_xy) x *= (PI / 180.0);
        r->group = g;
        g->members.push_back(r);
        g->left.push_back(r->req.n);
        r->r0 = g->search->add_receivers(rc_xy.data(), r->req.n);
        g->owner.resize(g->search->n_receivers(), r);
        if (!(flags & GA_SERVE_MERGE)) g->closed = true;
        if (r->req.n == 0) close_member(g, r);
    }

    void close_member(EigGroup *g, SrvReq *r)
    {
        g->search->results_for(r->r0, r->req.n, r->rows);
        for (size_t i = 0; i < g->members.size(); ++i)
            if (g->members[i] == r) {
                g->members.erase(g->members.begin() + i);
                g->left.erase(g->left.begin() + i);
                break;
            }
        finish(r, GA_OK);
    }

    void receivers_done(EigGroup *g)
    {
        for (int rcv : g->search->take_done()) {
            SrvReq *r = g->owner[rcv];
            for (size_t i = 0; i < g->members.size(); ++i)
                if (g->members[i] == r && --g->left[i] == 0) { close_member(g, r); break; }
        }
    }

    Slot *free_slot()
    {
        for (auto &sl : slots)
            if (sl->kind < 0) return sl.get();
        return nullptr;
    }

    // Launches one unit of work on a free slot; false when nothing is ready.
    bool dispatch()
    {
        Slot *sl = free_slot();
        if (!sl) return false;
        // eigenray searches first (their rounds are short and latency-bound)
        for (size_t k = 0; k < groups.size(); ++k) {
            EigGroup *g = groups[(rr + k) % groups.size()].get();
            if (g->inflight) continue;
            bool more = g->search->next_batch(g->batch);
            receivers_done(g);
            if (!more) continue;
            sl->eg->F = g->F;
            sl->eg->geom = g->p.geometry;
            int rc = sl->eg->enqueue(g->batch);
            if (rc != GA_OK) { fail_group(g, rc); return true; }
            g->inflight = true;
            sl->kind = GA_REQ_EIG;
            sl->group = g;
            rr = (rr + k + 1) % std::max<size_t>(groups.size(), 1);
            return true;
        }
        if (fans.empty()) return false;
        SrvReq *r = fans.front();
        fans.pop_front();
        int rc = launch_fan(*sl, r);
        if (rc != GA_OK) { finish(r, rc); return true; }
        sl->kind = GA_REQ_FAN;
        sl->fan = r;
        return true;
    }

    void fail_group(EigGroup *g, int rc)
    {
        std::vector<SrvReq *> m = g->members;
        for (SrvReq *r : m) finish(r, rc);
        g->members.clear();
        g->left.clear();
        g->closed = true;
    }

    int launch_fan(Slot &sl, SrvReq *r)
    {
        const int n = r->req.n, nb = r->req.prop.bounces + 1;
        const size_t nl = (size_t)std::max(n, 1);
        if (!sl.d_l.ensure(nl * 2) || !sl.d_o.ensure(nl * nb) || !sl.d_c.ensure(nl) || !sl.d_s.ensure(nl) ||
            !sl.h_o.ensure(nl * nb) || !sl.h_c.ensure(nl) || !sl.h_s.ensure(nl))
            return fail(ctx, GA_E_NOMEM, "fan buffers for %d rays", n);
        HIPCHK(ctx, hipEventRecord(sl.start, sl.stream));
        if (n > 0) {
            HIPCHK(ctx, hipMemcpyAsync(sl.d_l.p, r->data.data(), (size_t)n * 2 * sizeof(double), hipMemcpyHostToDevice, sl.stream));
            ga_prop_params p = r->req.prop;
            // one ray per lane: the launch's workgroups interleave with other requests'
            serve_one_shot = 1;
            int rc = fan_enqueue(ctx, sl.fb, r->req.atmo_id, &p, sl.d_l.p, n, sl.d_o.p, sl.d_c.p, sl.d_s.p, sl.stream);
            serve_one_shot = 0;
            if (rc != GA_OK) return rc;
            HIPCHK(ctx, hipMemcpyAsync(sl.h_o.p, sl.d_o.p, (size_t)n * nb * sizeof(ga_arrival), hipMemcpyDeviceToHost, sl.stream));
            HIPCHK(ctx, hipMemcpyAsync(sl.h_c.p, sl.d_c.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, sl.stream));
            HIPCHK(ctx, hipMemcpyAsync(sl.h_s.p, sl.d_s.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, sl.stream));
        }
        HIPCHK(ctx, hipEventRecord(sl.done, sl.stream));
        std::lock_guard<std::mutex> lk(mu);
        launches += 3;
        return GA_OK;
    }

    // Retires a finished launch; false when the slot is still busy.
    bool retire(Slot &sl)
    {
        if (sl.kind == GA_REQ_FAN) {
            if (hipEventQuery(sl.done) == hipErrorNotReady) return false;
            SrvReq *r = sl.fan;
            int rc = hipEventSynchronize(sl.done) == hipSuccess ? GA_OK : fail(ctx, GA_E_HIP, "fan launch failed");
            float ms = 0.0f;
            if (rc == GA_OK) (void)hipEventElapsedTime(&ms, sl.start, sl.done);
            const int n = r->req.n, nb = r->req.prop.bounces + 1;
            long long st = 0;
            if (rc == GA_OK)
                for (int i = 0; i < n; ++i) {
                    st += sl.h_s.p[i];
                    for (int b = 0; b < sl.h_c.p[i]; ++b) r->rows.push_back(sl.h_o.p[(size_t)i * nb + b]);
                }
            r->ray_steps = st;
            {
                std::lock_guard<std::mutex> lk(mu);
                steps += st;
                kernel_ms += ms;
            }
            sl.kind = -1;
            sl.fan = nullptr;
            finish(r, rc);
            return true;
        }
        EigGpu &eg = *sl.eg;
        if (!eg.ready()) return false;
        EigGroup *g = sl.group;
        EigBatchOut O;
        const double k0 = eg.kernel_ms;
        const int l0 = eg.launches;
        int rc = eg.collect(O);
        sl.kind = -1;
        sl.group = nullptr;
        g->inflight = false;
        if (rc != GA_OK) { fail_group(g, rc); return true; }
        const long long s0 = g->st.ray_steps;
        g->search->feed(g->batch, O, g->st);
        std::lock_guard<std::mutex> lk(mu);
        steps += g->st.ray_steps - s0;
        kernel_ms += eg.kernel_ms - k0;
        launches += eg.launches - l0;
        return true;
    }

    void loop()
    {
        (void)hipSetDevice(ctx->device);
        Clock::time_point last = Clock::now();
        for (;;) {
            bool progress = false;
            {
                // after destroy() was called, give up on work that makes no
                // progress for a minute (no launch in flight can be waited for)
                std::lock_guard<std::mutex> lk(mu);
                bool inflight_launch = false;
                for (auto &sl : slots) inflight_launch = inflight_launch || sl->kind >= 0;
                if (stop && !inflight_launch && ms_since(last, Clock::now()) > 60000.0) break;
            }
            {
                std::unique_lock<std::mutex> lk(mu);
                while (!submitted.empty()) {
                    pending.push_back(submitted.front());
                    submitted.pop_front();
                }
                bool busy = active > 0 || !pending.empty();
                if (!busy) {
                    if (stop) break;
                    cv_work.wait_for(lk, std::chrono::milliseconds(50));
                    continue;
                }
            }
            // admission in arrival order while the footprint fits
            while (!pending.empty()) {
                if ((flags & GA_SERVE_SERIAL) && active > 0) break;
                SrvReq *r = pending.front();
                if (active > 0 && inflight + footprint(*r) > budget) break;
                pending.pop_front();
                admit(r);
                progress = true;
            }
            for (auto &sl : slots)
                if (sl->kind >= 0 && retire(*sl)) progress = true;
            while (dispatch()) progress = true;
            // searches with no receivers left
            for (size_t i = 0; i < groups.size();) {
                EigGroup *g = groups[i].get();
                if (!g->inflight && g->members.empty() && (g->closed || g->search->idle())) groups.erase(groups.begin() + i);
                else ++i;
            }
            if (progress) last = Clock::now();
            else std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }
};

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
    p->refl_alt = NAN;
    if (geometry == GA_GEOM_2D) {
        // run_prop of infraga-2d (main.infraga.2d.cpp:129-141, geoac.params.2d.cpp,
        // atmo_io.2d.cpp:23-33): source at r = 0, range limit 2000 km (the
        // terrain's end with terrain: NaN selects that default)
        p->src[0] = 0.0; p->src[1] = 0.0; p->src[2] = 0.0;
        p->rng_max = NAN;
        for (double &b : p->bounds) b = NAN;
        return;
    }
    if (geometry == GA_GEOM_SPH) {
        p->src[0] = 30.0; p->src[1] = 0.0; p->src[2] = 0.0;
        p->rng_max = 1000.0;
        p->bounds[0] = -90.0; p->bounds[1] = 90.0; p->bounds[2] = -180.0; p->bounds[3] = 180.0;
    } else if (geometry == GA_GEOM_SPH_RNGDEP) {
        // run_prop (main.infraga.sph.rngdep.cpp:212-227): source at the region
        // centre, two bounces, region from the grid
        p->src[0] = NAN; p->src[1] = NAN; p->src[2] = 0.0;
        p->bounces = 2;
        p->rng_max = 1000.0;
        for (double &b : p->bounds) b = NAN;
    } else if (geometry == GA_GEOM_3D_RNGDEP) {
        // run_prop (main.infraga.3d.rngdep.cpp:205-243): source at the
        // region centre, ten bounces, rng_max of geoac.params.3d.cpp:36
        p->src[0] = NAN; p->src[1] = NAN; p->src[2] = 0.0;
        p->bounces = 10;
        p->rng_max = 1000.0;
        for (double &b : p->bounds) b = NAN;
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
    if (hipMalloc(&ctx->fb.queue, sizeof(unsigned int)) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return GA_E_NOMEM;
    }
    ctx->scratch_ptr = scratch_addr(dev);
    if (!ctx->scratch_ptr) {
        ctx->fb.release();
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return GA_E_HIP;
    }
    *out = ctx;
    return GA_OK;
}

void ga_ctx_destroy(ga_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->atmos) {
        (void)hipFree(kv.second.d_tab);
        if (kv.second.d_topo2d) (void)hipFree(kv.second.d_topo2d);
        kv.second.release_abs();
    }
    for (auto &kv : ctx->rds) kv.second.release();
    ctx->fb.release();
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

// A .met profile in the reference loader's formats and conversions
// (atmo_io.sph.strat.cpp:61-145): z [km], c [km/s], u, v [km/s], rho.
static int read_met(ga_ctx *ctx, const char *path, const char *format, int reverse_winds, std::vector<double> &z,
                    std::vector<double> &c, std::vector<double> &u, std::vector<double> &v, std::vector<double> &rho)
{
    std::string fmt = format ? format : "zTuvdp";
    std::ifstream in(path);
    if (!in.is_open()) return fail(ctx, GA_E_IO, "cannot open atmosphere file %s", path);
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
    return GA_OK;
}

int ga_atmo_load_file(ga_ctx *ctx, const char *path, const char *format, int reverse_winds, int *atmo_id)
{
    if (!ctx || !path || !atmo_id) return fail(ctx, GA_E_ARG, "null argument");
    std::vector<double> z, c, u, v, rho;
    int rc = read_met(ctx, path, format, reverse_winds, z, c, u, v, rho);
    if (rc != GA_OK) return rc;
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
    if (it->second.d_topo2d) HIPCHK(ctx, hipFree(it->second.d_topo2d));
    it->second.release_abs();
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
    return fan_enqueue(ctx, ctx->fb, atmo_id, p, d_launches, n, d_out, d_counts, d_steps, st);
}

int ga_fan_run_device_noalloc(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *d_launches, int n,
                              ga_arrival *d_out, int32_t *d_counts, int32_t *d_steps, void *stream)
{
    if (!ctx || !p || (n > 0 && (!d_launches || !d_out || !d_counts || !d_steps))) return fail(ctx, GA_E_ARG, "null argument");
    if (n < 0) return fail(ctx, GA_E_ARG, "negative ray count");
    if (p->geometry != GA_GEOM_SPH && p->geometry != GA_GEOM_3D)
        return fail(ctx, GA_E_ARG, "the no-allocation fan takes stratified geometries only (got %d)", p->geometry);
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    const size_t slots = (size_t)std::max(n, 1) * (size_t)(p->bounces + 1);
    if (slots > NOALLOC_RAW_DOUBLES / RAW)
        return fail(ctx, GA_E_ARG, "rays x (bounces + 1) = %zu exceeds the no-allocation limit of %zu arrival slots",
                    slots, NOALLOC_RAW_DOUBLES / RAW);
    if (!ctx->scratch_ptr) return fail(ctx, GA_E_HIP, "device scratch address unavailable");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    FanBufs fb;
    fb.fixed = true;
    fb.queue = (unsigned int *)ctx->scratch_ptr;
    fb.yp = (double *)(ctx->scratch_ptr + NOALLOC_YP_OFF);
    fb.yp_cap = NOALLOC_LANES * 18;
    fb.max_lanes = NOALLOC_LANES;
    fb.raw = (double *)(ctx->scratch_ptr + NOALLOC_RAW_OFF);
    fb.raw_cap = NOALLOC_RAW_DOUBLES;
    return fan_enqueue(ctx, fb, atmo_id, p, d_launches, n, d_out, d_counts, d_steps, st);
}

int ga_fan_autotune(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *d_launches, int n,
                    void *stream, ga_fan_tuning *best)
{
    if (!ctx || !p || !best || (n > 0 && !d_launches)) return fail(ctx, GA_E_ARG, "null argument");
    if (n <= 0) return fail(ctx, GA_E_ARG, "autotuning needs a launch list");
    if (p->geometry != GA_GEOM_SPH && p->geometry != GA_GEOM_3D)
        return fail(ctx, GA_E_ARG, "autotuning covers the stratified geometries (got %d)", p->geometry);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const int slots = p->bounces + 1;
    DevBuf<ga_arrival> d_o;
    DevBuf<int32_t> d_c, d_s;
    if (!d_o.alloc((size_t)n * slots) || !d_c.alloc((size_t)n) || !d_s.alloc((size_t)n))
        return fail(ctx, GA_E_NOMEM, "autotune buffers for %d rays", n);
    Events ev;
    if (!ev.ok) return fail(ctx, GA_E_HIP, "hipEventCreate failed");
    const bool sph = p->geometry == GA_GEOM_SPH;
    const int amp = p->calc_amp ? 1 : 0;
    const int save_s = ctx->tuned_slice[sph][amp], save_g = ctx->tuned_graph[sph][amp];
    static const int cand_slice[] = {32, 64, 128, 256, 512};
    static const int cand_graph[] = {8, 32};
    double best_ms = 1e300;
    int bs = SLICE_STEPS, bg = SLICES_PER_GRAPH;
    int rc = GA_OK;
    // one untimed pass with the defaults: buffers, graph capture, code load
    ctx->tuned_slice[sph][amp] = SLICE_STEPS;
    ctx->tuned_graph[sph][amp] = SLICES_PER_GRAPH;
    rc = fan_enqueue(ctx, ctx->fb, atmo_id, p, d_launches, n, d_o.p, d_c.p, d_s.p, st);
    for (int cg : cand_graph)
        for (int cs : cand_slice) {
            if (rc != GA_OK) break;
            ctx->tuned_slice[sph][amp] = cs;
            ctx->tuned_graph[sph][amp] = cg;
            hipError_t e0 = hipEventRecord(ev.e0, st);
            rc = fan_enqueue(ctx, ctx->fb, atmo_id, p, d_launches, n, d_o.p, d_c.p, d_s.p, st);
            if (rc != GA_OK) break;
            hipError_t e1 = hipEventRecord(ev.e1, st);
            if (e0 != hipSuccess || e1 != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
                rc = fail(ctx, GA_E_HIP, "autotune timing failed");
                break;
            }
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev.e0, ev.e1) != hipSuccess) { rc = fail(ctx, GA_E_HIP, "event time"); break; }
            if (ms < best_ms) { best_ms = ms; bs = cs; bg = cg; }
        }
    if (rc != GA_OK) {
        ctx->tuned_slice[sph][amp] = save_s;
        ctx->tuned_graph[sph][amp] = save_g;
        return rc;
    }
    ctx->tuned_slice[sph][amp] = bs;
    ctx->tuned_graph[sph][amp] = bg;
    best->slice_steps = bs;
    best->slices_per_graph = bg;
    best->pass_ms = best_ms;
    return GA_OK;
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
    int rc = fan_enqueue(ctx, ctx->fb, atmo_id, p, d_l.p, n, d_o.p, d_c.p, d_s.p, ctx->stream);
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

int ga_fan_run_paths(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const double *launches, int n,
                     int path_flags, ga_arrival *out, int out_cap, int *n_out, ga_path_pt *pts, int64_t pts_cap,
                     int64_t *n_pts, ga_run_stats *stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !p || !n_out || !n_pts || (n > 0 && !launches)) return fail(ctx, GA_E_ARG, "null argument");
    *n_out = 0;
    *n_pts = 0;
    if (n < 0) return fail(ctx, GA_E_ARG, "negative ray count");
    if (pts_cap < 0) return fail(ctx, GA_E_ARG, "negative path capacity");
    if (p->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    if (path_flags & ~(GA_PATH_RAYS | GA_PATH_CAUSTICS)) return fail(ctx, GA_E_ARG, "unknown path flags %d", path_flags);
    if ((path_flags & GA_PATH_CAUSTICS) && !p->calc_amp)
        return fail(ctx, GA_E_ARG, "caustics need the auxiliary equations (calc_amp)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int slots = p->bounces + 1;
    const size_t nl = (size_t)std::max(n, 1);
    DevBuf<double> d_l;
    DevBuf<ga_arrival> d_o;
    DevBuf<int32_t> d_c, d_s;
    if (!d_l.alloc(nl * 2) || !d_o.alloc(nl * slots) || !d_c.alloc(nl) || !d_s.alloc(nl))
        return fail(ctx, GA_E_NOMEM, "device buffers for %d rays", n);
    PathDev pd;
    PathOut P{};
    int rc = pd.setup(ctx, pts_cap, n, path_flags, ctx->stream, P);
    if (rc != GA_OK) return rc;
    Events ev;
    if (!ev.ok) return fail(ctx, GA_E_HIP, "hipEventCreate failed");
    HIPCHK(ctx, hipMemcpyAsync(d_l.p, launches, (size_t)n * 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ev.e0, ctx->stream));
    rc = fan_enqueue(ctx, ctx->fb, atmo_id, p, d_l.p, n, d_o.p, d_c.p, d_s.p, ctx->stream, &P);
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
    const int prc = pd.gather(ctx, pts, pts_cap, n_pts);
    if (stats) {
        stats->ray_steps = steps;
        stats->kernel_ms = ms;
        stats->n_rays = n;
        stats->n_launches = 4;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (total > out_cap || (!out && total > 0)) return fail(ctx, GA_E_CAP, "need %lld arrival slots, have %d", total, out_cap);
    size_t w = 0;
    for (int r = 0; r < n; ++r)
        for (int b = 0; b < cnt[r]; ++b) out[w++] = all[(size_t)r * slots + b];
    return prc;
}

namespace {

// ---------------------------------------------------------------------------
// -wnl_wvfrm (run_wnl_wvfrm, main.infraga.sph.cpp:1142-1479): ray traces with
// every row recorded (PATH_WVROWS), the reference point, the Burgers
// coefficients and the output scaling on the host from those rows, the
// evolution on the GPU (wv_evolve_kernel, one workgroup per waveform).
// ---------------------------------------------------------------------------
// Records of the waveform trace from their chunks into ray order
// (one workgroup of PCHUNK threads per chunk; off: first record of each ray).
__global__ void wv_place_kernel(const ga_path_pt *pts, const int32_t *citem, const int32_t *cseq,
                                const int32_t *cfill, unsigned int used, const int64_t *off, int n, ga_path_pt *dst)
{
    const unsigned int c = blockIdx.x;
    const int r = threadIdx.x;
    if (c >= used || r >= cfill[c]) return;
    const int it = citem[c];
    if (it < 0 || it >= n) return;
    const int64_t at = off[it] + (int64_t)cseq[c] * PCHUNK + r;
    if (at < off[it + 1]) dst[at] = pts[(size_t)c * PCHUNK + r];
}

// The source waveform's tables: the Tukey window (parameter 0.8), the two
// filters (1 + (m / (g (len/2 + 1)))^n1)^-n2 and the transform's twiddles
// cos, sin of -2 pi j / len (geoac.eqset.sph.strat.cpp:943-973).
void wv_tables(const ga_wvfrm_params *w, std::vector<double> &tukey, std::vector<double> &filt, std::vector<double> &tw)
{
    const int len = w->len, fft_len = len / 2 + 1;
    const double tp = 0.8;
    tukey.assign((size_t)len, 1.0);
    for (int m = 0; m < len; ++m) {
        if (m < tp * (len - 1) / 2.0) tukey[m] = 1.0 / 2.0 * (1.0 + std::cos(PI * (2.0 * m / (tp * (len - 1)) - 1.0)));
        else if (m > (len - 1) * (1.0 - tp / 2.0))
            tukey[m] = 1.0 / 2.0 * (1.0 + std::cos(PI * (2.0 * m / (tp * (len - 1)) - 2.0 / tp + 1.0)));
    }
    filt.assign((size_t)fft_len * 2, 0.0);
    for (int m = 0; m < fft_len; ++m) {
        filt[2 * m] = std::pow(1.0 + std::pow(m / (w->filt1_g * fft_len), w->filt1_n1), -w->filt1_n2);
        filt[2 * m + 1] = std::pow(1.0 + std::pow(m / (w->filt2_g * fft_len), w->filt2_n1), -w->filt2_n2);
    }
    tw.assign((size_t)len, 0.0);
    for (int j = 0; j < len / 2; ++j) {
        const double ang = -2.0 * PI * (double)j / (double)len;
        tw[2 * j] = std::cos(ang);
        tw[2 * j + 1] = std::sin(ang);
    }
}

}  // namespace

static int wnl_wvfrm_impl(ga_ctx *ctx, int atmo_id, const ga_prop_params *p_in, const ga_wvfrm_params *w,
                          const double *launches, int n, const double *t_src, const double *p_src,
                          double *t_init, double *p_init, double *t_out, double *p_out,
                          ga_arrival *arrivals, int32_t *status, int64_t *evolve_steps, ga_wvfrm_stats *stats,
                          double out_step, int snap_cap, double *snaps, int32_t *snap_n)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !p_in || !w) return fail(ctx, GA_E_ARG, "null argument");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n < 0) return fail(ctx, GA_E_ARG, "negative ray count");
    if (n > 0 && (!launches || !t_src || !p_src || !t_init || !p_init || !t_out || !p_out || !arrivals || !status))
        return fail(ctx, GA_E_ARG, "null argument");
    const int geom = p_in->geometry;
    if (geom != GA_GEOM_SPH && geom != GA_GEOM_3D && geom != GA_GEOM_SPH_RNGDEP && geom != GA_GEOM_3D_RNGDEP &&
        geom != GA_GEOM_2D)
        return fail(ctx, GA_E_ARG, "bad geometry %d", geom);
    const bool two_d = geom == GA_GEOM_2D;
    const bool rdm = geom == GA_GEOM_SPH_RNGDEP || geom == GA_GEOM_3D_RNGDEP;
    if (p_in->bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    const int len = w->len;
    int lg = 0;
    while (lg < 30 && (1 << lg) < len) ++lg;
    if (len < WV_MIN_LEN || len > WV_MAX_LEN || (1 << lg) != len)
        return fail(ctx, GA_E_ARG, "waveform length %d: a power of two from %d to %d is needed", len, WV_MIN_LEN,
                    WV_MAX_LEN);
    if (!(w->ds > 0.0) || !std::isfinite(w->ds)) return fail(ctx, GA_E_ARG, "wvfrm_ds must be positive");
    if (!std::isfinite(w->ref)) return fail(ctx, GA_E_ARG, "wvfrm_ref must be finite");
    auto it = ctx->atmos.find(atmo_id);
    if (rdm ? ctx->rds.find(atmo_id) == ctx->rds.end() : it == ctx->atmos.end())
        return fail(ctx, GA_E_ATMO, "unknown %s atmosphere id %d", rdm ? "range-dependent" : "stratified", atmo_id);
    if (n == 0) return GA_OK;
    if (!(t_src[1] > t_src[0])) return fail(ctx, GA_E_ARG, "source waveform times must increase");
    double p0 = 0.0;
    for (int i = 0; i < len; ++i) p0 = std::max(p0, std::fabs(p_src[i]));
    if (!(p0 > 0.0) || !std::isfinite(p0)) return fail(ctx, GA_E_ARG, "source waveform peak must be positive and finite");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const bool sph = geom == GA_GEOM_SPH || geom == GA_GEOM_SPH_RNGDEP;
    // run_wnl_wvfrm traces with the auxiliary equations and stops a ray
    // only at a break or after the last bounce (no turning-height or time
    // limits)
    ga_prop_params p = *p_in;
    p.calc_amp = 1;
    p.turn_ht_min = -1.0e300;
    p.time_max = 1.0e300;
    const int slots = p.bounces + 1;
    const size_t nl = (size_t)n;
    DevBuf<double> d_l;
    DevBuf<ga_arrival> d_o;
    DevBuf<int32_t> d_c, d_s;
    if (!d_l.alloc(nl * 2) || !d_o.alloc(nl * slots) || !d_c.alloc(nl) || !d_s.alloc(nl))
        return fail(ctx, GA_E_NOMEM, "device buffers for %d rays", n);
    HIPCHK(ctx, hipMemcpyAsync(d_l.p, launches, nl * 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    // the 2D trace fills only ground-return slots: an empty last slot marks
    // a ray whose last leg did not end on the ground
    HIPCHK(ctx, hipMemsetAsync(d_o.p, 0, nl * slots * sizeof(ga_arrival), ctx->stream));
    Events ev;
    if (!ev.ok) return fail(ctx, GA_E_HIP, "hipEventCreate failed");

    // 1. the traces, every row of every leg kept (records stay on the GPU);
    // the capacity is sized from the largest ray of the previous call and
    // the trace re-run with the exact count when a ray needs more
    const int64_t per_ray = std::max<int64_t>((int64_t)slots * 2 * 2048, ctx->wv_recs_per_ray + ctx->wv_recs_per_ray / 8);
    int64_t cap = (int64_t)n * per_ray, total = 0;
    double trace_ms = 0.0;
    PathDev pd;
    std::vector<int32_t> item_pts(nl);
    unsigned int used = 0;
    for (int pass = 0;; ++pass) {
        PathOut P{};
        int rc = pd.setup(ctx, cap, n, PATH_WVROWS, ctx->stream, P);
        if (rc != GA_OK) return rc;
        HIPCHK(ctx, hipEventRecord(ev.e0, ctx->stream));
        rc = fan_enqueue(ctx, ctx->fb, atmo_id, &p, d_l.p, n, d_o.p, d_c.p, d_s.p, ctx->stream, &P);
        if (rc != GA_OK) return rc;
        HIPCHK(ctx, hipEventRecord(ev.e1, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0.0f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ev.e0, ev.e1));
        trace_ms += ms;
        HIPCHK(ctx, hipMemcpy(&used, pd.ctr.p, sizeof(unsigned int), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(item_pts.data(), pd.item_pts.p, nl * sizeof(int32_t), hipMemcpyDeviceToHost));
        total = 0;
        int64_t most = 0;
        for (int i = 0; i < n; ++i) { total += item_pts[i]; most = std::max<int64_t>(most, item_pts[i]); }
        ctx->wv_recs_per_ray = std::max(ctx->wv_recs_per_ray, most);
        if (total <= cap && used <= pd.max_chunks) break;
        if (pass > 0) return fail(ctx, GA_E_HIP, "waveform rows overflow (%lld of %lld)", (long long)total, (long long)cap);
        cap = total;
    }
    std::vector<int32_t> cnt(nl), stp(nl);
    std::vector<ga_arrival> all(nl * slots);
    HIPCHK(ctx, hipMemcpy(cnt.data(), d_c.p, nl * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(stp.data(), d_s.p, nl * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(all.data(), d_o.p, all.size() * sizeof(ga_arrival), hipMemcpyDeviceToHost));
    std::vector<int64_t> off(nl + 1, 0);
    for (int i = 0; i < n; ++i) {
        if (item_pts[i] < 2 || (item_pts[i] & 1)) return fail(ctx, GA_E_HIP, "waveform rows of ray %d incomplete", i);
        off[i + 1] = off[i] + item_pts[i];
    }

    // 2. on the GPU: records in ray order, then per ray the reference point
    // and the spline rows of the legs it evolves over
    hipStream_t st = ctx->stream;
    const size_t n_rows = (size_t)(total / 2);
    DevBuf<ga_path_pt> d_rec;
    DevBuf<int64_t> d_off, d_lf;
    DevBuf<int32_t> d_ln;
    DevBuf<double> d_rows, d_tw, d_t, d_tk, d_f, d_u0, d_u;
    DevBuf<WvRay> d_ray;
    DevBuf<long long> d_st;
    std::vector<double> tukey, filt, tw;
    wv_tables(w, tukey, filt, tw);
    std::vector<double> u0((size_t)len);
    for (int m = 0; m < len; ++m) u0[m] = p_src[m] / p0;
    if (!d_rec.alloc((size_t)total) || !d_off.alloc(nl + 1) || !d_lf.alloc(nl * slots) || !d_ln.alloc(nl * slots) ||
        !d_rows.alloc(std::max<size_t>(n_rows, 1) * WV_ROW) || !d_ray.alloc(nl) || !d_tw.alloc(tw.size()) ||
        !d_t.alloc(len) || !d_tk.alloc(len) || !d_f.alloc(filt.size()) || !d_u0.alloc(len) ||
        !d_u.alloc(nl * (size_t)len) || !d_st.alloc(nl))
        return fail(ctx, GA_E_NOMEM, "device buffers for %d waveforms (%zu rows)", n, n_rows);
    if (slots > WV_MAX_SLOTS) return fail(ctx, GA_E_ARG, "waveforms take at most %d bounces", WV_MAX_SLOTS - 1);
    HIPCHK(ctx, hipMemcpyAsync(d_off.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(d_tw.p, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(d_t.p, t_src, (size_t)len * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(d_tk.p, tukey.data(), (size_t)len * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(d_f.p, filt.data(), filt.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(d_u0.p, u0.data(), (size_t)len * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipEventRecord(ev.e0, st));
    if (used > 0)
        hipLaunchKernelGGL(wv_place_kernel, dim3(used), dim3(PCHUNK), 0, st, (const ga_path_pt *)pd.pts.p,
                           (const int32_t *)pd.item.p, (const int32_t *)pd.seq.p, (const int32_t *)pd.fill.p, used,
                           (const int64_t *)d_off.p, n, d_rec.p);
    if (sph)
        hipLaunchKernelGGL((wv_prep_kernel<true>), dim3(n), dim3(WV_PREP_T), 0, st, (const ga_path_pt *)d_rec.p,
                           (const int64_t *)d_off.p, n, slots, w->ref, p0, d_rows.p, d_lf.p, d_ln.p, d_ray.p);
    else if (two_d)
        hipLaunchKernelGGL((wv_prep_kernel<false, true>), dim3(n), dim3(WV_PREP_T), 0, st, (const ga_path_pt *)d_rec.p,
                           (const int64_t *)d_off.p, n, slots, w->ref, p0, d_rows.p, d_lf.p, d_ln.p, d_ray.p);
    else
        hipLaunchKernelGGL((wv_prep_kernel<false>), dim3(n), dim3(WV_PREP_T), 0, st, (const ga_path_pt *)d_rec.p,
                           (const int64_t *)d_off.p, n, slots, w->ref, p0, d_rows.p, d_lf.p, d_ln.p, d_ray.p);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ev.e1, st));

    // 3. the evolution, one workgroup per waveform (rays without a
    // reference point have no legs and leave their waveform as it was)
    Events ev2;
    if (!ev2.ok) return fail(ctx, GA_E_HIP, "hipEventCreate failed");
    WvArgs A;
    std::memset(&A, 0, sizeof(A));
    int med = WV_STRAT;
    if (rdm) {
        RdFanArgs RA;
        int rc = rd_setup(ctx, ctx->fb, atmo_id, &p, RA, st);
        if (rc != GA_OK) return rc;
        A.RM = RA.M;
        med = sph ? WV_RD_SPH : WV_RD_CART;
        // the range-dependent sets floor the step at 1e-6 km; a stratified
        // profile over terrain runs the stratified sets' wnl_wvfrm
        A.ds_floor = RA.M.strat ? (sph ? 1.0e-7 : 1.0e-5) : 1.0e-6;
    } else {
        Limits L;
        int rc = GA_OK;
        if (!build_medium(ctx, it->second, &p, A.M, L, rc)) return rc;
        A.M.abs = nullptr;                        // per-bin absorption is evaluated directly
        A.ds_floor = (sph || two_d) ? 1.0e-7 : 1.0e-5;
        if (two_d) med = WV_2D;
    }
    A.strict = (sph || two_d) ? 1 : 0;
    A.len = len;
    A.lg = lg;
    A.n_wv = n;
    A.ds_wv = w->ds;
    A.tw = d_tw.p;
    A.t = d_t.p;
    A.tukey = d_tk.p;
    A.filt = d_f.p;
    A.u0 = d_u0.p;
    A.u = d_u.p;
    A.slots = slots;
    A.leg_first = d_lf.p;
    A.leg_n = d_ln.p;
    A.rows = d_rows.p;
    A.ray = d_ray.p;
    A.steps = d_st.p;
    const bool dump = snaps && snap_n && snap_cap > 0 && out_step < 1.0e4;
#define WV_EVOLVE(S, MD)                                                                              \
    do {                                                                                              \
        if (dump) hipLaunchKernelGGL((wv_evolve_kernel<S, MD, true>), dim3(n), dim3(WV_T), 0, st, A); \
        else hipLaunchKernelGGL((wv_evolve_kernel<S, MD>), dim3(n), dim3(WV_T), 0, st, A);            \
    } while (0)
    const size_t snap_w = 1 + (size_t)len + 2 * ((size_t)len / 2 + 1);
    DevBuf<double> d_snap;
    DevBuf<int32_t> d_sn;
    A.out_step = out_step;
    A.ref = w->ref;
    A.snap = nullptr;
    A.snap_cap = 0;
    A.snap_n = nullptr;
    if (dump && n > 0) {
        if (!d_snap.ensure((size_t)n * snap_cap * snap_w) || !d_sn.ensure((size_t)n))
            return fail(ctx, GA_E_NOMEM, "waveform dump buffer (%d x %d snapshots)", n, snap_cap);
        A.snap = d_snap.p;
        A.snap_cap = snap_cap;
        A.snap_n = d_sn.p;
    }
    HIPCHK(ctx, hipEventRecord(ev2.e0, st));
    if (med == WV_2D) {
        WV_EVOLVE(false, WV_2D);
    } else if (med == WV_STRAT) {
        if (sph) WV_EVOLVE(true, WV_STRAT);
        else WV_EVOLVE(false, WV_STRAT);
    } else if (med == WV_RD_SPH) {
        WV_EVOLVE(true, WV_RD_SPH);
    } else {
        WV_EVOLVE(false, WV_RD_CART);
    }
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ev2.e1, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    float prep_ms = 0.0f, evo_ms = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&prep_ms, ev.e0, ev.e1));
    HIPCHK(ctx, hipEventElapsedTime(&evo_ms, ev2.e0, ev2.e1));
    const double evolve_ms = evo_ms;
    std::vector<WvRay> rays(nl);
    std::vector<double> uq(nl * (size_t)len);
    std::vector<long long> sq(nl, 0);
    HIPCHK(ctx, hipMemcpy(rays.data(), d_ray.p, nl * sizeof(WvRay), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(uq.data(), d_u.p, uq.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(sq.data(), d_st.p, sq.size() * sizeof(long long), hipMemcpyDeviceToHost));
#undef WV_EVOLVE
    if (dump && n > 0) {
        HIPCHK(ctx, hipMemcpy(snap_n, d_sn.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(snaps, d_snap.p, (size_t)n * snap_cap * snap_w * sizeof(double), hipMemcpyDeviceToHost));
    } else if (snap_n) {
        for (int i = 0; i < n; ++i) snap_n[i] = 0;
    }

    // 4. outputs: wvfrm_init.dat and wvfrm_out.dat rows (main.infraga.sph.cpp:
    // 1380-1386, 1440-1461), the arrival summary
    long long ray_steps = 0, wv_steps = 0;
    for (int i = 0; i < n; ++i) {
        ray_steps += stp[(size_t)i];
        const WvRay &o = rays[(size_t)i];
        const bool ok = two_d ? all[(size_t)i * slots + (slots - 1)].time > 0.0 : cnt[(size_t)i] == slots;
        status[i] = !o.started ? -1 : (ok ? 1 : 0);
        if (ok) arrivals[i] = all[(size_t)i * slots + (slots - 1)];
        else {
            std::memset(&arrivals[i], 0, sizeof(ga_arrival));
            arrivals[i].launch_idx = i;
            arrivals[i].incl = launches[2 * i];
            arrivals[i].az = launches[2 * i + 1];
            arrivals[i].time = o.tt;
        }
        if (evolve_steps) evolve_steps[i] = o.started ? sq[(size_t)i] : 0;
        double *ti = t_init + (size_t)i * len, *pi_ = p_init + (size_t)i * len;
        double *to = t_out + (size_t)i * len, *po = p_out + (size_t)i * len;
        if (!o.started) {
            for (int m = 0; m < len; ++m) ti[m] = pi_[m] = to[m] = po[m] = std::nan("");
            continue;
        }
        wv_steps += sq[(size_t)i];
        const double *u = uq.data() + (size_t)i * len;
        for (int m = 0; m < len; ++m) {
            ti[m] = o.tref + t_src[m];
            pi_[m] = p_src[m];
            to[m] = o.tt + t_src[m];
            po[m] = u[m] * p0 * o.scale;
        }
    }
    trace_ms += prep_ms;
    if (stats) {
        stats->ray_steps = ray_steps;
        stats->wv_steps = wv_steps;
        stats->trace_ms = trace_ms;
        stats->evolve_ms = evolve_ms;
        stats->n_waveforms = n;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GA_OK;
}

// One -eig_direct request: launch estimate [deg] (az NaN: toward the
// receiver) and reflections.
struct EigDirect {
    double incl_deg, az_deg;
    int bounces;
};

static int eig_search_impl(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                           const double *rcvr_xy, int n_rcvr, ga_eigenray *out, int out_cap, int *n_out,
                           double *rcvr_latency_ms, ga_run_stats *stats, bool paths, ga_path_pt *pts,
                           int64_t pts_cap, int64_t *n_pts, const EigDirect *direct = nullptr)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!ctx || !p || !e || !n_out || (n_rcvr > 0 && !rcvr_xy)) return fail(ctx, GA_E_ARG, "null argument");
    *n_out = 0;
    if (n_rcvr < 0) return fail(ctx, GA_E_ARG, "negative receiver count");
    int rc = eig_check(ctx, p, e);
    if (rc != GA_OK) return rc;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    EigGpu gpu;
    gpu.ctx = ctx;
    gpu.geom = p->geometry;
    EigSetup S;
    std::memset(&S, 0, sizeof(S));
    std::vector<double> rcvr(rcvr_xy, rcvr_xy + 2 * (size_t)n_rcvr);
    if (p->geometry == GA_GEOM_SPH_RNGDEP || p->geometry == GA_GEOM_3D_RNGDEP) {
        gpu.rd = true;
        gpu.cart = p->geometry == GA_GEOM_3D_RNGDEP;
        gpu.geom = gpu.cart ? GA_GEOM_3D : GA_GEOM_SPH;
        rc = eig_prepare_rd(ctx, atmo_id, p, e, gpu.RF, S, rcvr);
    } else {
        auto it = ctx->atmos.find(atmo_id);
        if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
        rc = eig_prepare(ctx, it->second, p, e, gpu.F, S, rcvr);
    }
    if (rc != GA_OK) return rc;
    rc = gpu.init(ctx->stream);
    if (rc != GA_OK) return rc;
    std::vector<ga_arrival> found;
    std::vector<EigJob> jobs;
    std::vector<double> lat;
    EigSearchStats es;
    {
        EigSearch search(S, direct ? nullptr : rcvr.data(), direct ? 0 : n_rcvr);
        if (direct) {
            // run_eig_direct: theta_est from incl_est; phi_est = pi/2 - az_est,
            // by default the direction of the receiver (sph: pi/2 - bearing,
            // main.infraga.sph.cpp:1112; 3D: atan2(dy, dx), main.infraga.3d.cpp:1050)
            const double th = direct->incl_deg * (PI / 180.0);
            double ph;
            if (!std::isnan(direct->az_deg)) ph = PI / 2.0 - direct->az_deg * (PI / 180.0);
            else if (S.geom == GA_GEOM_SPH) {
                const double t1 = std::sin(rcvr[1] - S.src[2]);
                const double t2 = std::cos(S.src[1]) * std::tan(rcvr[0]) - std::sin(S.src[1]) * std::cos(rcvr[1] - S.src[2]);
                ph = PI / 2.0 - std::atan2(t1, t2);
            } else {
                ph = std::atan2(rcvr[1] - S.src[1], rcvr[0] - S.src[0]);
            }
            search.add_direct(rcvr.data(), th, ph, direct->bounces);
        }
        rc = search.run([&](const EigBatch &B, EigBatchOut &O) { return gpu(B, O); }, found, lat, es);
        search.result_jobs(jobs);
    }
    if (rc != GA_OK) return rc;
    *n_out = (int)found.size();
    if (rcvr_latency_ms)
        for (int i = 0; i < n_rcvr; ++i) rcvr_latency_ms[i] = lat[i];
    int prc = GA_OK;
    if (paths) {
        // the accepted traces once more, writing their path rows
        PathDev pd;
        prc = gpu.trace_paths(jobs, pd, pts_cap);
        if (prc == GA_OK) prc = pd.gather(ctx, pts, pts_cap, n_pts);
        else *n_pts = 0;
    }
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
    return prc;
}

int ga_eig_search(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                  const double *rcvr_xy, int n_rcvr, ga_eigenray *out, int out_cap, int *n_out,
                  double *rcvr_latency_ms, ga_run_stats *stats)
{
    return eig_search_impl(ctx, atmo_id, p, e, rcvr_xy, n_rcvr, out, out_cap, n_out, rcvr_latency_ms, stats, false,
                           nullptr, 0, nullptr);
}

int ga_eig_direct(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e, const double *rcvr_xy,
                  double incl_est_deg, double az_est_deg, int bounces, ga_eigenray *out, int *found,
                  ga_path_pt *pts, int64_t pts_cap, int64_t *n_pts, ga_run_stats *stats)
{
    if (!ctx || !p || !e || !rcvr_xy || !out || !found || !n_pts) return fail(ctx, GA_E_ARG, "null argument");
    *found = 0;
    *n_pts = 0;
    if (bounces < 0) return fail(ctx, GA_E_ARG, "bounces must be >= 0");
    if (pts_cap < 0) return fail(ctx, GA_E_ARG, "negative path capacity");
    if (!std::isfinite(incl_est_deg)) return fail(ctx, GA_E_ARG, "inclination estimate must be finite");
    ga_eig_params e2 = *e;
    e2.bnc_min = e2.bnc_max = bounces;
    const EigDirect d{incl_est_deg, az_est_deg, bounces};
    return eig_search_impl(ctx, atmo_id, p, &e2, rcvr_xy, 1, out, 1, found, nullptr, stats, true, pts, pts_cap,
                           n_pts, &d);
}

int ga_eig_search_paths(ga_ctx *ctx, int atmo_id, const ga_prop_params *p, const ga_eig_params *e,
                        const double *rcvr_xy, int n_rcvr, ga_eigenray *out, int out_cap, int *n_out,
                        double *rcvr_latency_ms, ga_path_pt *pts, int64_t pts_cap, int64_t *n_pts,
                        ga_run_stats *stats)
{
    if (!n_pts) return fail(ctx, GA_E_ARG, "null argument");
    *n_pts = 0;
    if (pts_cap < 0) return fail(ctx, GA_E_ARG, "negative path capacity");
    return eig_search_impl(ctx, atmo_id, p, e, rcvr_xy, n_rcvr, out, out_cap, n_out, rcvr_latency_ms, stats, true,
                           pts, pts_cap, n_pts);
}



// ---- range-dependent atmosphere and topography ---------------------------

static bool increasing(const double *v, int n)
{
    for (int i = 0; i + 1 < n; ++i)
        if (!(v[i + 1] > v[i])) return false;
    return true;
}

static int rngdep_load(ga_ctx *ctx, const double *lat_deg, int nlat, const double *lon_deg, int nlon, const double *z_km,
                       int nz, const double *c, const double *u, const double *v, const double *rho, int *rd_id, bool cart)
{
    if (!ctx || !lat_deg || !lon_deg || !z_km || !c || !u || !v || !rho || !rd_id) return fail(ctx, GA_E_ARG, "null argument");
    if (nlat < 2 || nlon < 2 || nz < 3) return fail(ctx, GA_E_ARG, "grid needs >= 2 x 2 nodes and >= 3 levels");
    if (!increasing(lat_deg, nlat) || !increasing(lon_deg, nlon) || !increasing(z_km, nz))
        return fail(ctx, GA_E_ARG, "node coordinates must increase strictly");
    const size_t N = (size_t)nlat * nlon * nz;
    std::vector<double> x(nlat), y(nlon), z(z_km, z_km + nz);
    // node files hold degrees; the splines run in radians (atmo_io.sph.rngdep.cpp)
    // node files hold degrees and the splines run in radians
    // (atmo_io.sph.rngdep.cpp); the 3D grid stays in km
    const double deg = cart ? 1.0 : PI / 180.0;
    for (int i = 0; i < nlat; ++i) x[i] = lat_deg[i] * deg;
    for (int j = 0; j < nlon; ++j) y[j] = lon_deg[j] * deg;
    std::vector<double> tab(4 * N * RD_Q);
    const double *fields[4] = {c, u, v, rho};
    std::vector<double> col(nz), sl, dcol(nz), dsl;
    for (int fl = 0; fl < 4; ++fl) {
        const double *F = fields[fl];
        for (size_t q = 0; q < N; ++q)
            if (!std::isfinite(F[q])) return fail(ctx, GA_E_ARG, "non-finite value in field %d", fl);
        auto at = [&](int i, int j, int k) { return F[((size_t)i * nlon + j) * nz + k]; };
        for (int i = 0; i < nlat; ++i)
            for (int j = 0; j < nlon; ++j) {
                const int iu = std::min(i + 1, nlat - 1), id = std::max(i - 1, 0);
                const int ju = std::min(j + 1, nlon - 1), jd = std::max(j - 1, 0);
                for (int k = 0; k < nz; ++k) {
                    col[k] = at(i, j, k);
                    dcol[k] = (at(iu, j, k) - at(id, j, k)) / (x[iu] - x[id]);
                }
                spline_slopes(z, col, sl);
                spline_slopes(z, dcol, dsl);
                for (int k = 0; k < nz; ++k) {
                    double *t = &tab[((((size_t)fl * nlat + i) * nlon + j) * nz + k) * RD_Q];
                    t[0] = col[k];
                    t[1] = sl[k];
                    t[2] = dcol[k];
                    t[3] = dsl[k];
                    t[4] = (at(i, ju, k) - at(i, jd, k)) / (y[ju] - y[jd]);
                }
            }
    }
    std::vector<double> axes;
    axes.insert(axes.end(), x.begin(), x.end());
    axes.insert(axes.end(), y.begin(), y.end());
    axes.insert(axes.end(), z.begin(), z.end());
    HIPCHK(ctx, hipSetDevice(ctx->device));
    DevRd r;
    r.nx = nlat; r.ny = nlon; r.nz = nz;
    r.hx = x; r.hy = y; r.hz = z;
    r.cart = cart;
    if (hipMalloc(&r.d_axes, axes.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&r.d_tab, tab.size() * sizeof(double)) != hipSuccess) {
        r.release();
        return fail(ctx, GA_E_NOMEM, "range-dependent tables (%zu MB)", tab.size() * 8 / 1000000);
    }
    if (hipMemcpy(r.d_axes, axes.data(), axes.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r.d_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        r.release();
        return fail(ctx, GA_E_HIP, "range-dependent table upload failed");
    }
    int id = ctx->next_id++;
    ctx->rds[id] = r;
    *rd_id = id;
    return GA_OK;
}

int ga_rngdep_load(ga_ctx *ctx, const double *lat_deg, int nlat, const double *lon_deg, int nlon, const double *z_km,
                   int nz, const double *c, const double *u, const double *v, const double *rho, int *rd_id)
{
    return rngdep_load(ctx, lat_deg, nlat, lon_deg, nlon, z_km, nz, c, u, v, rho, rd_id, false);
}

int ga_rngdep_load_xy(ga_ctx *ctx, const double *x_km, int nx, const double *y_km, int ny, const double *z_km, int nz,
                      const double *c, const double *u, const double *v, const double *rho, int *rd_id)
{
    return rngdep_load(ctx, x_km, nx, y_km, ny, z_km, nz, c, u, v, rho, rd_id, true);
}

int ga_topo_load(ga_ctx *ctx, int rd_id, const double *lat_deg, int nlat, const double *lon_deg, int nlon,
                 const double *zg_km)
{
    if (!ctx || !lat_deg || !lon_deg || !zg_km) return fail(ctx, GA_E_ARG, "null argument");
    auto it = ctx->rds.find(rd_id);
    if (it == ctx->rds.end()) return fail(ctx, GA_E_ATMO, "unknown range-dependent atmosphere id %d", rd_id);
    if (nlat < 2 || nlon < 2) return fail(ctx, GA_E_ARG, "topography needs >= 2 x 2 nodes");
    if (!increasing(lat_deg, nlat) || !increasing(lon_deg, nlon)) return fail(ctx, GA_E_ARG, "topography nodes must increase");
    std::vector<double> x(nlat), y(nlon);
    const double deg = it->second.cart ? 1.0 : PI / 180.0;
    for (int i = 0; i < nlat; ++i) x[i] = lat_deg[i] * deg;
    for (int j = 0; j < nlon; ++j) y[j] = lon_deg[j] * deg;
    // per node: z, y slope of its row, x slope of its column of z, x slope of
    // its column of y slopes
    std::vector<double> fy((size_t)nlat * nlon), tab((size_t)nlat * nlon * TP_Q), row(nlon), sl, colv(nlat), csl;
    for (int i = 0; i < nlat; ++i) {
        for (int j = 0; j < nlon; ++j) {
            row[j] = zg_km[(size_t)i * nlon + j];
            if (!std::isfinite(row[j])) return fail(ctx, GA_E_ARG, "non-finite topography at (%d, %d)", i, j);
        }
        spline_slopes(y, row, sl);
        for (int j = 0; j < nlon; ++j) fy[(size_t)i * nlon + j] = sl[j];
    }
    for (int j = 0; j < nlon; ++j) {
        for (int i = 0; i < nlat; ++i) colv[i] = zg_km[(size_t)i * nlon + j];
        spline_slopes(x, colv, csl);
        for (int i = 0; i < nlat; ++i) {
            double *t = &tab[((size_t)i * nlon + j) * TP_Q];
            t[0] = colv[i];
            t[1] = fy[(size_t)i * nlon + j];
            t[2] = csl[i];
        }
        for (int i = 0; i < nlat; ++i) colv[i] = fy[(size_t)i * nlon + j];
        spline_slopes(x, colv, csl);
        for (int i = 0; i < nlat; ++i) tab[((size_t)i * nlon + j) * TP_Q + 3] = csl[i];
    }
    std::vector<double> axes(x);
    axes.insert(axes.end(), y.begin(), y.end());
    DevRd &r = it->second;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (r.d_taxes) { HIPCHK(ctx, hipFree(r.d_taxes)); r.d_taxes = nullptr; }
    if (r.d_ttab) { HIPCHK(ctx, hipFree(r.d_ttab)); r.d_ttab = nullptr; }
    if (hipMalloc(&r.d_taxes, axes.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&r.d_ttab, tab.size() * sizeof(double)) != hipSuccess)
        return fail(ctx, GA_E_NOMEM, "topography tables");
    HIPCHK(ctx, hipMemcpy(r.d_taxes, axes.data(), axes.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx, hipMemcpy(r.d_ttab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    r.tnx = nlat;
    r.tny = nlon;
    r.thx = x;
    r.thy = y;
    r.t_max = 0.0;                                   // set_bndlyr starts from 0 (atmo_state.sph.rngdep.cpp:44-48)
    for (size_t q = 0; q < (size_t)nlat * nlon; ++q) r.t_max = std::max(r.t_max, zg_km[q]);
    return GA_OK;
}

int ga_topo_load_2d(ga_ctx *ctx, int atmo_id, const double *r_km, const double *z_km, int n)
{
    if (!ctx || !r_km || !z_km) return fail(ctx, GA_E_ARG, "null argument");
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    if (n < 3) return fail(ctx, GA_E_ARG, "terrain needs at least 3 nodes");
    if (!increasing(r_km, n)) return fail(ctx, GA_E_ARG, "terrain ranges must increase strictly");
    DevAtmo &a = it->second;
    a.topo_r.assign(r_km, r_km + n);
    a.topo_z.assign(z_km, z_km + n);
    std::vector<double> sl;
    spline_slopes(a.topo_r, a.topo_z, sl);
    std::vector<double> tab((size_t)n * 3);
    for (int i = 0; i < n; ++i) {
        tab[3 * i] = r_km[i];
        tab[3 * i + 1] = z_km[i];
        tab[3 * i + 2] = sl[i];
    }
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (a.d_topo2d) HIPCHK(ctx, hipFree(a.d_topo2d));
    a.d_topo2d = nullptr;
    if (hipMalloc(&a.d_topo2d, tab.size() * sizeof(double)) != hipSuccess) return fail(ctx, GA_E_NOMEM, "terrain table");
    HIPCHK(ctx, hipMemcpy(a.d_topo2d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    return GA_OK;
}

int ga_topo_eval_2d(ga_ctx *ctx, int atmo_id, double r, double out[3])
{
    if (!ctx || !out) return fail(ctx, GA_E_ARG, "null argument");
    auto it = ctx->atmos.find(atmo_id);
    if (it == ctx->atmos.end()) return fail(ctx, GA_E_ATMO, "unknown atmosphere id %d", atmo_id);
    const DevAtmo &a = it->second;
    if (a.topo_r.empty()) {
        out[0] = a.z.front();
        out[1] = 0.0;
        out[2] = 0.0;
        return GA_OK;
    }
    const std::vector<double> &x = a.topo_r, &f = a.topo_z;
    std::vector<double> sl;
    spline_slopes(x, f, sl);
    const double rr = std::min(std::max(r, x.front()), x.back());
    size_t k = 0;
    while (k + 2 < x.size() && rr > x[k + 1]) ++k;
    const double dx = x[k + 1] - x[k], X = (rr - x[k]) / dx, d = f[k + 1] - f[k];
    const double A = sl[k] * dx - d, B = -sl[k + 1] * dx + d;
    out[0] = f[k] + X * (d + (1.0 - X) * (A * (1.0 - X) + B * X));
    out[1] = d / dx + (1.0 - 2.0 * X) * (A * (1.0 - X) + B * X) / dx + X * (1.0 - X) * (B - A) / dx;
    out[2] = 2.0 * (B - 2.0 * A + (A - B) * 3.0 * X) / (dx * dx);
    return GA_OK;
}

int ga_rngdep_info(ga_ctx *ctx, int rd_id, double out[8])
{
    if (!ctx || !out) return fail(ctx, GA_E_ARG, "null argument");
    auto it = ctx->rds.find(rd_id);
    if (it == ctx->rds.end()) return fail(ctx, GA_E_ATMO, "unknown range-dependent atmosphere id %d", rd_id);
    RdRegion G;
    if (!rd_region(it->second, NAN, G)) return fail(ctx, GA_E_ARG, "atmosphere grid and topography grid have no overlap");
    out[0] = G.lat_min; out[1] = G.lat_max; out[2] = G.lon_min; out[3] = G.lon_max;
    out[4] = G.alt_max; out[5] = G.z0; out[6] = G.z_max; out[7] = G.z_bndlyr;
    return GA_OK;
}

// fileIO.cpp:18-35: lines not starting with '#', minus one (a file written
// with a trailing newline reads one empty line at the end).
static int file_length(const std::string &path)
{
    std::ifstream in(path);
    if (!in.is_open()) return 0;
    int count = 0;
    std::string line;
    while (!in.eof()) {
        std::getline(in, line);
        if (line.find('#') != 0) ++count;
    }
    return count - 1;
}

// fileIO.cpp:64-85: rows "x y f" with y fastest; n2 = length of the first y run.
static bool file_2d_dims(const std::string &path, int &n1, int &n2)
{
    const int N = file_length(path);
    std::ifstream in(path);
    if (!in.is_open() || N < 2) return false;
    double xj, yj, fj, yprev;
    in >> xj >> yprev >> fj;
    n2 = 0;
    for (int n = 1; n < N; ++n) {
        in >> xj >> yj >> fj;
        if (yj < yprev) { n2 = n; break; }
        yprev = yj;
    }
    if (n2 < 1) return false;
    n1 = N / n2;
    return true;
}

// A terrain file of "x y z" rows, y fastest (file_2d_dims and the loops of
// set_region, atmo_io.sph.rngdep.cpp:203-230, atmo_io.sph.strat.cpp:147-175).
static int read_terrain(ga_ctx *ctx, const char *path, std::vector<double> &tl, std::vector<double> &to,
                        std::vector<double> &tz)
{
    int n1 = 0, n2 = 0;
    if (!file_2d_dims(path, n1, n2) || n1 < 2 || n2 < 2) return fail(ctx, GA_E_IO, "invalid terrain file (%s)", path);
    tl.assign(n1, 0.0);
    to.assign(n2, 0.0);
    tz.assign((size_t)n1 * n2, 0.0);
    std::ifstream in(path);
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            if (!(in >> tl[i] >> to[j] >> tz[(size_t)i * n2 + j]))
                return fail(ctx, GA_E_IO, "cannot read terrain row %d in %s", i * n2 + j, path);
    return GA_OK;
}

static int rngdep_load_files(ga_ctx *ctx, const char *prefix, const char *lat_path, const char *lon_path,
                             const char *topo_path, const char *format, int reverse_winds, int *rd_id, bool cart)
{
    if (!ctx || !prefix || !lat_path || !lon_path || !rd_id) return fail(ctx, GA_E_ARG, "null argument");
    const std::string fmt = format ? format : "zTuvdp";
    const int need = (fmt.compare(0, 6, "zTuvdp") == 0) ? 6 : (fmt.compare(0, 7, "zuvwTdp") == 0) ? 7
                     : (fmt.compare(0, 5, "zcuvd") == 0) ? 5 : -1;
    if (need < 0) return fail(ctx, GA_E_ARG, "unrecognized profile format %s (zTuvdp, zuvwTdp, zcuvd)", fmt.c_str());
    const int nlat = file_length(lat_path), nlon = file_length(lon_path);
    if (nlat < 2 || nlon < 2) return fail(ctx, GA_E_IO, "invalid grid specification (check the node files %s, %s)", lat_path, lon_path);
    std::vector<double> lat(nlat), lon(nlon);
    {
        std::ifstream a(lat_path), b(lon_path);
        for (double &x : lat) if (!(a >> x)) return fail(ctx, GA_E_IO, "cannot read %s", lat_path);
        for (double &x : lon) if (!(b >> x)) return fail(ctx, GA_E_IO, "cannot read %s", lon_path);
    }
    const std::string p0 = std::string(prefix) + "0.met";
    const int nz = file_length(p0);
    if (nz < 3) return fail(ctx, GA_E_IO, "invalid atmospheric specification (%s)", p0.c_str());
    const size_t N = (size_t)nlat * nlon * nz;
    std::vector<double> z(nz), c(N), u(N), v(N), rho(N);
    for (int i = 0; i < nlat; ++i)
        for (int j = 0; j < nlon; ++j) {
            const std::string path = std::string(prefix) + std::to_string(i * nlon + j) + ".met";
            std::ifstream in(path);
            if (!in.is_open()) return fail(ctx, GA_E_IO, "invalid atmospheric specification (%s)", path.c_str());
            std::string line;
            int k = 0;
            while (!in.eof() && k < nz) {
                std::getline(in, line);
                if (line.find('#') == 0) continue;
                std::stringstream ss(line);
                double vals[7];
                int got = 0;
                while (got < need && (ss >> vals[got])) ++got;
                if (got < need) return fail(ctx, GA_E_IO, "short row %d in %s", k, path.c_str());
                const size_t q = ((size_t)i * nlon + j) * nz + k;
                double rr, cc, uu, vv;
                if (need == 6) { uu = vals[2]; vv = vals[3]; rr = vals[4]; cc = std::sqrt(0.1 * 1.4 * vals[5] / rr) / 1000.0; }
                else if (need == 7) { uu = vals[1]; vv = vals[2]; rr = vals[5]; cc = std::sqrt(0.1 * 1.4 * vals[6] / rr) / 1000.0; }
                else { cc = vals[1] / 1000.0; uu = vals[2]; vv = vals[3]; rr = vals[4]; }
                if (reverse_winds) { uu /= -1000.0; vv /= -1000.0; }
                else { uu /= 1000.0; vv /= 1000.0; }
                z[k] = vals[0];   // every profile writes the shared heights, as the reference's loader does
                c[q] = cc; u[q] = uu; v[q] = vv; rho[q] = rr;
                ++k;
            }
            if (k < nz) return fail(ctx, GA_E_IO, "%s has %d rows, expected %d", path.c_str(), k, nz);
        }
    int rc = rngdep_load(ctx, lat.data(), nlat, lon.data(), nlon, z.data(), nz, c.data(), u.data(), v.data(),
                            rho.data(), rd_id, cart);
    if (rc != GA_OK || !topo_path || !topo_path[0]) return rc;
    std::vector<double> tl, to, tz;
    rc = read_terrain(ctx, topo_path, tl, to, tz);
    if (rc == GA_OK) rc = ga_topo_load(ctx, *rd_id, tl.data(), (int)tl.size(), to.data(), (int)to.size(), tz.data());
    if (rc != GA_OK) ga_rngdep_release(ctx, *rd_id);
    return rc;
}

int ga_rngdep_load_files(ga_ctx *ctx, const char *prefix, const char *lat_path, const char *lon_path,
                         const char *topo_path, const char *format, int reverse_winds, int *rd_id)
{
    return rngdep_load_files(ctx, prefix, lat_path, lon_path, topo_path, format, reverse_winds, rd_id, false);
}

// The 3D Cartesian node files (set_region, atmo_io.3d.rngdep.cpp:56-225): the
// same layout with x / y node coordinates and terrain rows in km.
int ga_rngdep_load_files_xy(ga_ctx *ctx, const char *prefix, const char *x_path, const char *y_path,
                            const char *topo_path, const char *format, int reverse_winds, int *rd_id)
{
    return rngdep_load_files(ctx, prefix, x_path, y_path, topo_path, format, reverse_winds, rd_id, true);
}

// A stratified profile over terrain: 4 x 4 identical columns whose grid spans
// the terrain grid exactly (node spacing as numpy's linspace), the terrain,
// and the stratified binaries' semantics (DevRd::strat).  Over identical
// columns the hybrid splines give the 1D splines' values, so the
// range-dependent kernels run the stratified medium with terrain.
static int strat_topo_build(ga_ctx *ctx, const double *z, const double *c, const double *u, const double *v,
                            const double *rho, int nz, const double *tl, int n1, const double *to, int n2,
                            const double *tz, int *atmo_id, bool cart = false)
{
    if (!ctx || !z || !c || !u || !v || !rho || !tl || !to || !tz || !atmo_id) return fail(ctx, GA_E_ARG, "null argument");
    if (nz < 3) return fail(ctx, GA_E_ARG, "profile needs >= 3 levels");
    if (n1 < 2 || n2 < 2) return fail(ctx, GA_E_ARG, "topography needs >= 2 x 2 nodes");
    constexpr int NC = 4;
    double la[NC], lo[NC];
    const double sa = (tl[n1 - 1] - tl[0]) / (NC - 1), so = (to[n2 - 1] - to[0]) / (NC - 1);
    for (int i = 0; i < NC; ++i) {
        la[i] = (i == NC - 1) ? tl[n1 - 1] : i * sa + tl[0];
        lo[i] = (i == NC - 1) ? to[n2 - 1] : i * so + to[0];
    }
    const size_t N = (size_t)NC * NC * nz;
    std::vector<double> cc(N), uu(N), vv(N), rr(N);
    for (size_t q = 0; q < (size_t)NC * NC; ++q)
        for (int k = 0; k < nz; ++k) {
            cc[q * nz + k] = c[k];
            uu[q * nz + k] = u[k];
            vv[q * nz + k] = v[k];
            rr[q * nz + k] = rho[k];
        }
    int id = 0;
    int rc = rngdep_load(ctx, la, NC, lo, NC, z, nz, cc.data(), uu.data(), vv.data(), rr.data(), &id, cart);
    if (rc != GA_OK) return rc;
    rc = ga_topo_load(ctx, id, tl, n1, to, n2, tz);
    if (rc != GA_OK) {
        ga_rngdep_release(ctx, id);
        return rc;
    }
    ctx->rds[id].strat = true;
    *atmo_id = id;
    return GA_OK;
}

int ga_atmo_load_strat_topo(ga_ctx *ctx, const double *z_km, const double *c_kms, const double *u_kms,
                            const double *v_kms, const double *rho_gcc, int n, const double *lat_deg, int nlat,
                            const double *lon_deg, int nlon, const double *zg_km, int *atmo_id)
{
    return strat_topo_build(ctx, z_km, c_kms, u_kms, v_kms, rho_gcc, n, lat_deg, nlat, lon_deg, nlon, zg_km, atmo_id);
}

int ga_atmo_load_file_topo(ga_ctx *ctx, const char *path, const char *topo_path, const char *format,
                           int reverse_winds, int *atmo_id)
{
    if (!ctx || !path || !topo_path || !atmo_id) return fail(ctx, GA_E_ARG, "null argument");
    std::vector<double> z, c, u, v, rho, tl, to, tz;
    int rc = read_met(ctx, path, format, reverse_winds, z, c, u, v, rho);
    if (rc == GA_OK) rc = read_terrain(ctx, topo_path, tl, to, tz);
    if (rc != GA_OK) return rc;
    return strat_topo_build(ctx, z.data(), c.data(), u.data(), v.data(), rho.data(), (int)z.size(), tl.data(),
                            (int)tl.size(), to.data(), (int)to.size(), tz.data(), atmo_id);
}

// The 3D stratified profile over terrain (infraga-3d with topo_file,
// set_region of atmo_io.3d.strat.cpp:146-257): node coordinates and terrain
// rows in km, the Cartesian range-dependent kernels with the stratified 3D
// semantics (boundary-layer top, upper-triangle contraction, turning height
// from the profile bottom, floored caustic rows).
int ga_atmo_load_strat_topo_xy(ga_ctx *ctx, const double *z_km, const double *c_kms, const double *u_kms,
                               const double *v_kms, const double *rho_gcc, int n, const double *x_km, int nx,
                               const double *y_km, int ny, const double *zg_km, int *atmo_id)
{
    return strat_topo_build(ctx, z_km, c_kms, u_kms, v_kms, rho_gcc, n, x_km, nx, y_km, ny, zg_km, atmo_id, true);
}

int ga_atmo_load_file_topo_xy(ga_ctx *ctx, const char *path, const char *topo_path, const char *format,
                              int reverse_winds, int *atmo_id)
{
    if (!ctx || !path || !topo_path || !atmo_id) return fail(ctx, GA_E_ARG, "null argument");
    std::vector<double> z, c, u, v, rho, tx, ty, tz;
    int rc = read_met(ctx, path, format, reverse_winds, z, c, u, v, rho);
    if (rc == GA_OK) rc = read_terrain(ctx, topo_path, tx, ty, tz);
    if (rc != GA_OK) return rc;
    return strat_topo_build(ctx, z.data(), c.data(), u.data(), v.data(), rho.data(), (int)z.size(), tx.data(),
                            (int)tx.size(), ty.data(), (int)ty.size(), tz.data(), atmo_id, true);
}

int ga_rngdep_eval(ga_ctx *ctx, int rd_id, const double *pts, int n, double *atmo_out, double *topo_out)
{
    if (!ctx || (n > 0 && (!pts || !atmo_out || !topo_out))) return fail(ctx, GA_E_ARG, "null argument");
    if (n < 0) return fail(ctx, GA_E_ARG, "negative point count");
    auto it = ctx->rds.find(rd_id);
    if (it == ctx->rds.end()) return fail(ctx, GA_E_ATMO, "unknown range-dependent atmosphere id %d", rd_id);
    if (n == 0) return GA_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const DevRd &r = it->second;
    const RdGrid g = rd_grid_of(r);
    const TopoGrid t = topo_grid_of(r);
    DevBuf<double> d_p, d_a, d_t;
    if (!d_p.alloc((size_t)n * 3) || !d_a.alloc((size_t)n * 40) || !d_t.alloc((size_t)n * 10))
        return fail(ctx, GA_E_NOMEM, "evaluation buffers");
    HIPCHK(ctx, hipMemcpyAsync(d_p.p, pts, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(rd_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, g, t, r.d_ttab ? 1 : 0, d_p.p,
                       n, d_a.p, d_t.p);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(atmo_out, d_a.p, (size_t)n * 40 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(topo_out, d_t.p, (size_t)n * 10 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return GA_OK;
}

int ga_rngdep_release(ga_ctx *ctx, int rd_id)
{
    if (!ctx) return GA_E_ARG;
    auto it = ctx->rds.find(rd_id);
    if (it == ctx->rds.end()) return fail(ctx, GA_E_ATMO, "unknown range-dependent atmosphere id %d", rd_id);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    it->second.release();
    ctx->rds.erase(it);
    return GA_OK;
}

int ga_server_create(ga_ctx *ctx, int n_slots, double hbm_budget_bytes, int flags, ga_server **out)
{
    if (!ctx || !out) return fail(ctx, GA_E_ARG, "null argument");
    *out = nullptr;
    if (n_slots < 1 || n_slots > 16) return fail(ctx, GA_E_ARG, "n_slots must be 1..16 (got %d)", n_slots);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    auto srv = std::make_unique<ga_server>();
    srv->ctx = ctx;
    srv->budget = hbm_budget_bytes > 0.0 ? hbm_budget_bytes : 16.0e9;
    srv->flags = flags;
    for (int i = 0; i < n_slots; ++i) {
        auto sl = std::make_unique<Slot>();
        HIPCHK(ctx, hipStreamCreate(&sl->stream));
        HIPCHK(ctx, hipEventCreate(&sl->done));
        HIPCHK(ctx, hipEventCreate(&sl->start));
        sl->eg = std::make_unique<EigGpu>();
        sl->eg->ctx = ctx;
        int rc = sl->eg->init(sl->stream);
        if (rc != GA_OK) return rc;
        srv->slots.push_back(std::move(sl));
    }
    ga_server *s = srv.release();
    s->th = std::thread([s]() { s->loop(); });
    *out = s;
    return GA_OK;
}

int ga_server_submit(ga_server *srv, const ga_request *req, uint64_t *ticket)
{
    if (!srv || !req || !ticket) return GA_E_ARG;
    if (req->n < 0 || (req->n > 0 && !req->data)) return GA_E_ARG;
    if (req->kind != GA_REQ_FAN && req->kind != GA_REQ_EIG) return GA_E_ARG;
    auto *r = new SrvReq();
    r->req = *req;
    r->data.assign(req->data, req->data + 2 * (size_t)req->n);
    r->req.data = nullptr;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) { delete r; return GA_E_ARG; }
    r->ticket = srv->next_ticket++;
    srv->reqs[r->ticket] = r;
    srv->submitted.push_back(r);
    *ticket = r->ticket;
    srv->cv_work.notify_all();
    return GA_OK;
}

int ga_server_poll(ga_server *srv, uint64_t ticket, int timeout_ms, ga_result *res, ga_arrival *rows, int rows_cap)
{
    if (!srv) return GA_E_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return GA_E_ARG;
    SrvReq *r = it->second;
    if (timeout_ms < 0) srv->cv_done.wait(lk, [&]() { return r->done; });
    else srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&]() { return r->done; });
    if (!r->done) return GA_PENDING;
    if (res) {
        res->status = r->status;
        res->n_rows = (int32_t)r->rows.size();
        res->queue_ms = ms_since(r->t_submit, r->t_admit);
        res->total_ms = ms_since(r->t_submit, r->t_done);
        res->ray_steps = r->ray_steps;
    }
    if ((long long)r->rows.size() > rows_cap || (!rows && !r->rows.empty())) return GA_E_CAP;
    for (size_t i = 0; i < r->rows.size(); ++i) rows[i] = r->rows[i];
    const int st = r->status;
    srv->reqs.erase(it);
    delete r;
    return st;
}

void ga_server_stats(ga_server *srv, int64_t *ray_steps, int32_t *launches, double *kernel_ms)
{
    if (!srv) return;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (ray_steps) *ray_steps = srv->steps;
    if (launches) *launches = srv->launches;
    if (kernel_ms) *kernel_ms = srv->kernel_ms;
}

void ga_server_destroy(ga_server *srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_work.notify_all();
    }
    if (srv->th.joinable()) srv->th.join();
    for (auto &sl : srv->slots) {
        (void)hipStreamSynchronize(sl->stream);
        sl->eg.reset();
        sl->fb.release();
        (void)hipEventDestroy(sl->done);
        (void)hipEventDestroy(sl->start);
        (void)hipStreamDestroy(sl->stream);
    }
    for (auto &kv : srv->reqs) delete kv.second;
    delete srv;
}

}  // extern "C"

int ga_wnl_wvfrm(ga_ctx *ctx, int atmo_id, const ga_prop_params *p_in, const ga_wvfrm_params *w,
                 const double *launches, int n, const double *t_src, const double *p_src,
                 double *t_init, double *p_init, double *t_out, double *p_out,
                 ga_arrival *arrivals, int32_t *status, int64_t *evolve_steps, ga_wvfrm_stats *stats)
{
    return wnl_wvfrm_impl(ctx, atmo_id, p_in, w, launches, n, t_src, p_src, t_init, p_init, t_out, p_out, arrivals,
                          status, evolve_steps, stats, 1.0e10, 0, nullptr, nullptr);
}

int ga_wnl_wvfrm_dump(ga_ctx *ctx, int atmo_id, const ga_prop_params *p_in, const ga_wvfrm_params *w,
                      const double *launches, int n, const double *t_src, const double *p_src,
                      double *t_init, double *p_init, double *t_out, double *p_out,
                      ga_arrival *arrivals, int32_t *status, int64_t *evolve_steps, ga_wvfrm_stats *stats,
                      double out_step, int snap_cap, double *snaps, int32_t *snap_n)
{
    if (!(out_step > 0.0) || snap_cap < 1 || !snaps || !snap_n)
        return fail(ctx, GA_E_ARG, "waveform dumps need out_step > 0, snap_cap >= 1 and output buffers");
    return wnl_wvfrm_impl(ctx, atmo_id, p_in, w, launches, n, t_src, p_src, t_init, p_init, t_out, p_out, arrivals,
                          status, evolve_steps, stats, out_step, snap_cap, snaps, snap_n);
}
