// ga_eig_search.hpp -- host side of the batched eigenray search.
//
// Every (receiver, bounce count) search of run_eig_search
// (src/main.infraga.3d.cpp:940-951, src/main.infraga.sph.cpp:962-989) runs
// as a C++20 coroutine that follows est_eigenray and find_eigenray
// (src/geoac/geoac.eigenray.3d.cpp:34-366, geoac.eigenray.sph.cpp:41-380)
// statement by statement, except that each ray trace is an awaited request.
// Requests from all searches are collected into one batch per round and
// traced on the GPU together; the coroutines resume when their traces are
// back.  Independent work is exposed as follows:
//   - the inclination sweep at the receiver azimuth (every est_eigenray pass
//     before its first azimuth correction) steps by the constant dth_big, so
//     the whole sweep up to theta_max is known in advance and is requested
//     at once; one sweep trace through bnc_max reflections serves every
//     bounce count of that receiver;
//   - after an azimuth correction dth stays constant (3D: one correction,
//     sph: two): a chunk of the following inclinations is requested with
//     each miss;
//   - each accepted estimate starts an independent find_eigenray coroutine,
//     so the next est_eigenray pass of the same search does not wait for it.
// Once mod_dth takes over (3D: second correction, sph: third), dth depends
// on the previous trace and the sweep advances one trace per round.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <coroutine>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <unordered_map>
#include <vector>

#include "ga_eig.hpp"

namespace ga {

struct EigSetup {
    int geom;                      // GA_GEOM_3D or GA_GEOM_SPH
    double src[3];                 // 3D: x, y, z [km]; sph: z [km], lat, lon [rad]; z raised to the ground
    double r0;                     // sph: globe radius [km]
    double src_ground;             // ground height z0 [km]
    double th_start;               // first theta_start (sph: max(theta_min, theta_grnd))
    double th_min, th_max;         // rad
    int bnc_min, bnc_max;
    int iterations;                // LM iteration limit
    double damping, tolerance;     // -, km
    double az_err_lim;             // deg
    double dth_big, dth_sml;       // rad
    double rng_max;                // km
    double c_src, c_grnd;          // c at the source and at the ground [km/s] (LM update)
    int spec_chunk;                // sweep traces requested per miss after one azimuth correction
    int rngdep;                    // a range-dependent medium: the LM update reads the ground
                                   // under row K (topo::z, topo::dz, atmo::c there) from the
                                   // trace result
    int is_topo;                   // geoac::is_topo (the 3D update's terrain-slope form)
};

// One round of traces.  sweep: EIG_SWEEP jobs; full: EIG_LM jobs followed by
// EIG_FINAL jobs (n_lm of the former).
struct EigBatch {
    std::vector<EigJob> sweep;
    std::vector<EigJob> full;
    int n_lm = 0;
    int sweep_stride = 0;          // doubles per sweep result
};

struct EigBatchOut {
    std::vector<double> sweep_res;
    std::vector<int32_t> sweep_brk, sweep_steps;
    std::vector<double> full_res;
    std::vector<int32_t> full_brk, full_steps;
    std::vector<ga_arrival> final_rows;   // one per EIG_FINAL job, in order
};

using EigTracer = std::function<int(const EigBatch &, EigBatchOut &)>;

struct EigSearchStats {
    long long ray_steps = 0;
    long long traces = 0;
    int rounds = 0;
};

class EigSearch {
public:
    EigSearch(const EigSetup &S, const double *rcvr_xy, int n_rcvr) : S_(S), n_rcvr_(0)
    {
        stride_ = 2 * (S.bnc_max + 1);
        init_xy_.assign(rcvr_xy, rcvr_xy + 2 * (size_t)std::max(n_rcvr, 0));
    }
    explicit EigSearch(const EigSetup &S) : S_(S), n_rcvr_(0) { stride_ = 2 * (S.bnc_max + 1); }
    ~EigSearch() { drop_all(); }

    // Runs every search to completion.  out: eigenrays grouped by receiver,
    // then bounce count, then the order the reference finds them in;
    // latency_ms[i]: time from the start of the call until receiver i's last
    // search finished.
    int run(const EigTracer &trace, std::vector<ga_arrival> &out, std::vector<double> &latency_ms,
            EigSearchStats &st)
    {
        start(latency_ms);
        EigBatch B;
        while (next_batch(B)) {
            EigBatchOut O;
            int rc = trace(B, O);
            if (rc != 0) { drop_all(); return rc; }
            feed(B, O, st);
        }
        results(out);
        return 0;
    }

    // Stepwise use (several searches interleaved by a scheduler): start(),
    // then next_batch() / feed() until next_batch() returns false, then
    // results().
    void start(std::vector<double> &latency_ms)
    {
        t0_ = std::chrono::steady_clock::now();
        lat_ = &latency_ms;
        add_receivers(init_xy_.data(), (int)(init_xy_.size() / 2));
        latency_ms.assign(n_rcvr_, 0.0);
    }

    // Adds receivers to a running search (their searches join the next
    // batch); returns the index of the first one.  Before the first call,
    // set_clock() fixes the time origin of the completion times.
    int add_receivers(const double *xy, int n)
    {
        const int base = n_rcvr_;
        rcvr_.insert(rcvr_.end(), xy, xy + 2 * (size_t)n);
        n_rcvr_ += n;
        live_.resize(n_rcvr_, 0);
        done_ms_.resize(n_rcvr_, 0.0);
        for (int r = base; r < n_rcvr_; ++r) {
            const Direct *d = nullptr;
            for (const Direct &q : direct_)
                if (q.rcvr == r) d = &q;
            if (d) {
                spawn(S_.geom == GA_GEOM_SPH ? find_sph(r, d->nb, 0, d->th, d->ph) : find(r, d->nb, 0, d->th, d->ph), r);
                continue;
            }
            for (int nb = S_.bnc_min; nb <= S_.bnc_max; ++nb)
                spawn(S_.geom == GA_GEOM_SPH ? chain_sph(r, nb) : chain(r, nb), r);
            if (S_.bnc_max < S_.bnc_min) receiver_done(r);
        }
        return base;
    }
    void set_clock(std::chrono::steady_clock::time_point t0) { t0_ = t0; }

    // -eig_direct (run_eig_direct, src/main.infraga.3d.cpp:955-1086,
    // src/main.infraga.sph.cpp:998-1139): one find_eigenray from the launch
    // estimate (th_est, ph_est) [rad, geoac convention] with nb reflections
    // toward the receiver xy (search units); no inclination sweep.  Returns
    // the receiver index; call before start().
    int add_direct(const double *xy, double th_est, double ph_est, int nb)
    {
        const int r = (int)(init_xy_.size() / 2);
        direct_.push_back(Direct{r, nb, th_est, ph_est});
        init_xy_.insert(init_xy_.end(), xy, xy + 2);
        return r;
    }

    // Receivers whose searches all finished since the last call.
    std::vector<int> take_done()
    {
        std::vector<int> d;
        d.swap(newly_done_);
        return d;
    }
    bool idle() const { return ready_.empty() && waiting_.empty(); }
    int n_receivers() const { return n_rcvr_; }

    // Eigenrays of receivers [r0, r0 + n), launch_idx renumbered from 0.
    void results_for(int r0, int n, std::vector<ga_arrival> &out)
    {
        std::vector<Found> sel;
        for (const Found &f : found_)
            if (f.rcvr >= r0 && f.rcvr < r0 + n) sel.push_back(f);
        std::sort(sel.begin(), sel.end(), found_order);
        out.clear();
        for (Found &f : sel) {
            f.row.launch_idx = f.rcvr - r0;
            out.push_back(f.row);
        }
    }

    // Advances every search as far as the traces at hand allow; fills B with
    // the traces they wait for.  False: all searches are finished.
    bool next_batch(EigBatch &B)
    {
        while (!ready_.empty()) {
            Handle h = ready_.back();
            ready_.pop_back();
            h.resume();
            if (h.done()) finish(h);
        }
        B.sweep.clear();
        B.full.clear();
        if (waiting_.empty()) return false;
        B.sweep_stride = stride_;
        for (const Key &k : req_sweep_) B.sweep.push_back(job_of(k));
        for (const Key &k : req_full_)
            if (k.mode == EIG_LM) B.full.push_back(job_of(k));
        B.n_lm = (int)B.full.size();
        for (const Key &k : req_full_)
            if (k.mode == EIG_FINAL) B.full.push_back(job_of(k));
        return true;
    }

    void feed(const EigBatch &B, const EigBatchOut &O, EigSearchStats &st)
    {
        st.rounds++;
        st.traces += (long long)B.sweep.size() + (long long)B.full.size();
        for (size_t i = 0; i < B.sweep.size(); ++i) {
            st.ray_steps += O.sweep_steps[i];
            uint32_t id = (uint32_t)sweep_brk_.size();
            sweep_brk_.push_back(O.sweep_brk[i]);
            sweep_store_.insert(sweep_store_.end(), O.sweep_res.begin() + i * stride_,
                                O.sweep_res.begin() + (i + 1) * stride_);
            sweep_idx_[key_of(B.sweep[i])] = id;
        }
        for (size_t i = 0; i < B.full.size(); ++i) {
            st.ray_steps += O.full_steps[i];
            uint32_t id = (uint32_t)full_brk_.size();
            full_brk_.push_back(O.full_brk[i]);
            full_store_.insert(full_store_.end(), O.full_res.begin() + i * EIG_FULL,
                               O.full_res.begin() + (i + 1) * EIG_FULL);
            final_rows_.push_back((int)i >= B.n_lm ? O.final_rows[i - B.n_lm] : ga_arrival{});
            full_idx_[key_of(B.full[i])] = id;
        }
        req_sweep_.clear();
        req_full_.clear();
        ready_.swap(waiting_);
    }

    void results(std::vector<ga_arrival> &out)
    {
        std::sort(found_.begin(), found_.end(), found_order);
        out.clear();
        for (const Found &f : found_) out.push_back(f.row);
        if (lat_) *lat_ = done_ms_;
    }

    // The accepted traces of the eigenrays results() returns, in the same
    // order (call after results()).
    void result_jobs(std::vector<EigJob> &jobs) const
    {
        jobs.clear();
        for (const Found &f : found_) jobs.push_back(f.job);
    }

private:
    static constexpr uint32_t PENDING = 0xffffffffu;

    // ---- coroutine plumbing ------------------------------------------------
    struct Task {
        struct promise_type {
            int rcvr = 0;
            Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
            std::suspend_always initial_suspend() noexcept { return {}; }
            std::suspend_always final_suspend() noexcept { return {}; }
            void return_void() {}
            void unhandled_exception() { std::terminate(); }
        };
        std::coroutine_handle<promise_type> h;
    };
    using Handle = std::coroutine_handle<Task::promise_type>;

    void spawn(Task t, int rcvr)
    {
        t.h.promise().rcvr = rcvr;
        live_[rcvr]++;
        ready_.push_back(t.h);
    }

    void finish(Handle h)
    {
        int r = h.promise().rcvr;
        h.destroy();
        if (--live_[r] == 0) receiver_done(r);
    }
    void receiver_done(int r)
    {
        done_ms_[r] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
        newly_done_.push_back(r);
    }

    void drop_all()
    {
        for (Handle h : ready_) h.destroy();
        for (Handle h : waiting_) h.destroy();
        ready_.clear();
        waiting_.clear();
    }

    // ---- trace cache -------------------------------------------------------
    struct Key {
        uint64_t th, ph;
        int32_t legs, mode;
        bool operator==(const Key &o) const { return th == o.th && ph == o.ph && legs == o.legs && mode == o.mode; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const
        {
            uint64_t h = k.th * 0x9E3779B97F4A7C15ull;
            h ^= k.ph + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
            h ^= (uint64_t)(uint32_t)k.legs * 0x85EBCA77C2B2AE63ull + ((uint64_t)(uint32_t)k.mode << 40);
            return (size_t)(h ^ (h >> 29));
        }
    };
    static uint64_t bits(double x)
    {
        uint64_t b;
        std::memcpy(&b, &x, 8);
        return b;
    }
    static double dbl(uint64_t b)
    {
        double x;
        std::memcpy(&x, &b, 8);
        return x;
    }
    static Key key(double th, double ph, int legs, int mode) { return Key{bits(th), bits(ph), legs, mode}; }
    static EigJob job_of(const Key &k) { return EigJob{dbl(k.th), dbl(k.ph), k.legs, k.mode}; }
    static Key key_of(const EigJob &j) { return key(j.th, j.ph, j.legs, j.mode); }

    void want_sweep(const Key &k)
    {
        auto ins = sweep_idx_.emplace(k, PENDING);
        if (ins.second) req_sweep_.push_back(k);
    }
    void want_full(const Key &k)
    {
        auto ins = full_idx_.emplace(k, PENDING);
        if (ins.second) req_full_.push_back(k);
    }

    // Outcome of trace_legs(nb) for est_eigenray (geoac.eigenray.3d.cpp:66-76).
    struct SweepHit {
        bool brk;
        double x, y;
    };
    struct SweepAwait {
        EigSearch *S;
        Key k;
        int nb;
        double dth, th_max;
        int spec;
        bool strict;     // sweep bound th < th_max (sph) instead of th <= th_max
        bool await_ready()
        {
            auto it = S->sweep_idx_.find(k);
            if (it != S->sweep_idx_.end() && it->second != PENDING) return true;
            S->want_sweep(k);
            // speculative continuation of the sweep: th + dth, th + 2 dth, ...
            double t = dbl(k.th);
            for (int j = 0; j < spec; ++j) {
                t += dth;
                if (strict ? !(t < th_max) : !(t <= th_max)) break;
                S->want_sweep(key(t, dbl(k.ph), k.legs, EIG_SWEEP));
            }
            return false;
        }
        void await_suspend(std::coroutine_handle<> h) { S->waiting_.push_back(Handle::from_address(h.address())); }
        SweepHit await_resume()
        {
            uint32_t id = S->sweep_idx_[k];
            int bl = S->sweep_brk_[id];
            bool brk = bl <= nb;
            int leg = brk ? bl : nb;
            const double *r = &S->sweep_store_[(size_t)id * S->stride_];
            return SweepHit{brk, r[2 * leg], r[2 * leg + 1]};
        }
    };

    struct FullHit {
        bool brk;
        double y[EIG_FULL];
        ga_arrival row;
    };
    struct FullAwait {
        EigSearch *S;
        Key k;
        bool await_ready()
        {
            auto it = S->full_idx_.find(k);
            if (it != S->full_idx_.end() && it->second != PENDING) return true;
            S->want_full(k);
            return false;
        }
        void await_suspend(std::coroutine_handle<> h) { S->waiting_.push_back(Handle::from_address(h.address())); }
        FullHit await_resume()
        {
            uint32_t id = S->full_idx_[k];
            FullHit R;
            R.brk = S->full_brk_[id] <= k.legs;
            std::memcpy(R.y, &S->full_store_[(size_t)id * EIG_FULL], sizeof(R.y));
            R.row = S->final_rows_[id];
            return R;
        }
    };

    SweepAwait sweep(double th, double ph, int legs, int nb, double dth, double th_max, int spec)
    {
        return SweepAwait{this, key(th, ph, legs, EIG_SWEEP), nb, dth, th_max, spec, S_.geom == GA_GEOM_SPH};
    }
    FullAwait full(double th, double ph, int legs, int mode) { return FullAwait{this, key(th, ph, legs, mode)}; }

    // ---- the searches ------------------------------------------------------
    static constexpr double RPD = PI / 180.0;
    static constexpr double DPR = 180.0 / PI;
    static double sq(double x) { return x * x; }

    // globe.cpp:16-28
    double gc_dist(double lat1, double lon1, double lat2, double lon2) const
    {
        double term1 = sq(std::sin((lat2 - lat1) / 2.0));
        double term2 = std::cos(lat1) * std::cos(lat2) * sq(std::sin((lon2 - lon1) / 2.0));
        return 2.0 * S_.r0 * std::asin(std::sqrt(term1 + term2));
    }
    static double bearing(double lat1, double lon1, double lat2, double lon2)
    {
        double term1 = std::sin(lon2 - lon1);
        double term2 = std::cos(lat1) * std::tan(lat2) - std::sin(lat1) * std::cos(lon2 - lon1);
        return std::atan2(term1, term2);
    }

    // Decision log of the 3D search for diagnosis (GA_EIG_LOG=<path>,
    // appended): "E ri nb seq th_est ph_est" per accepted estimate and
    // "L ri nb seq n theta phi dr act" per Levenberg-Marquardt iteration
    // (act: b break, a accept, h halve, s step), angles in radians.
    static FILE *decision_log()
    {
        static FILE *f = []() -> FILE * {
            const char *path = std::getenv("GA_EIG_LOG");
            return (path && *path) ? std::fopen(path, "a") : nullptr;
        }();
        return f;
    }

    double mod_dth(double dr, double dr_dtheta) const   // geoac.eigenray.3d.cpp:32
    {
        return S_.dth_big - (S_.dth_big - S_.dth_sml) * std::exp(-1.0 / 2.0 * sq(dr / dr_dtheta));
    }

    // run_eig_search's loop for one bounce count (main.infraga.3d.cpp:944-951)
    // with est_eigenray (geoac.eigenray.3d.cpp:34-144) inlined.
    Task chain(int ri, int nb)
    {
        const double *src = S_.src;
        const double rcvr[2] = {rcvr_[2 * ri], rcvr_[2 * ri + 1]};
        const double rcvr_rng = std::sqrt(sq(rcvr[0] - src[0]) + sq(rcvr[1] - src[1]));
        const double rcvr_az = std::atan2(rcvr[1] - src[1], rcvr[0] - src[0]);
        int seq = 0;
        double theta_start = S_.th_start;
        while (theta_start < S_.th_max) {
            double th_min = theta_start, th_max = S_.th_max;
            double th_est = 0.0, ph_est = 0.0, th_next = theta_start;
            int iterations = 0;
            double r, r_prev, dth = S_.dth_big, dph = 100.0;
            bool th_max_reached = false, success = false;
            double phi = rcvr_az;
            while (std::fabs(dph * DPR) > S_.az_err_lim && iterations < 5) {
                r = rcvr_rng;
                r_prev = rcvr_rng;
                for (double th = th_min; th <= th_max; th += dth) {
                    if (th + dth >= th_max) th_max_reached = true;
                    // the first pass at the receiver azimuth is shared by every
                    // bounce count and requested whole; later passes per search
                    const int legs = (iterations == 0) ? S_.bnc_max : nb;
                    const int spec = (iterations == 0) ? (1 << 30) : (iterations == 1 ? S_.spec_chunk : 0);
                    SweepHit h = co_await sweep(th, phi, legs, nb, dth, th_max, spec);
                    r = std::sqrt(sq(h.x - src[0]) + sq(h.y - src[1]));
                    if (h.brk) r = S_.rng_max;
                    if (((r - rcvr_rng) * (r_prev - rcvr_rng) <= 0.0) && (r < S_.rng_max) && (r_prev < S_.rng_max) &&
                        (th > th_min)) {
                        if (iterations == 0) th_next = th;
                        dph = std::atan2(h.y - src[1], h.x - src[0]) * DPR;
                        dph -= std::atan2(rcvr[1] - src[1], rcvr[0] - src[0]) * DPR;
                        while (dph > 180.0) dph -= 360.0;
                        while (dph < -180.0) dph += 360.0;
                        if (std::fabs(dph) < S_.az_err_lim) {
                            th_est = th - dth;
                            ph_est = phi;
                            success = true;
                        } else {
                            phi -= dph * RPD * 0.9;
                            th_min = std::max(th - 5.0 * RPD, th_min);
                        }
                        iterations++;
                        break;
                    }
                    if (iterations >= 2) dth = mod_dth(r - rcvr_rng, (r - r_prev) / (2.0 * dth));
                    r_prev = r;
                }
                if (success) break;
                if (th_max_reached) {
                    th_next = th_max;
                    break;
                }
            }
            if (success) spawn(find(ri, nb, seq++, th_est, ph_est), ri);
            theta_start = th_next;
        }
    }

    // find_eigenray (geoac.eigenray.3d.cpp:146-366) with the arithmetic types
    // of the reference (long double for the update, double for theta/phi).
    Task find(int ri, int nb, int seq, double th_est, double ph_est)
    {
        const double *src = S_.src;
        const double rcvr[2] = {rcvr_[2 * ri], rcvr_[2 * ri + 1]};
        double theta = th_est, phi = ph_est;
        FILE *lg = decision_log();
        if (lg) std::fprintf(lg, "E %d %d %d %.17g %.17g\n", ri, nb, seq, th_est, ph_est);
        double dr, dr_prev = 10000.0, step_sc = 1.0;
        long double x, y, c_src, c_grnd, dzg_dx, dzg_dy, ds_dth, ds_dph, dx, dy, dx_dth, dy_dth, dx_dph, dy_dph;
        long double det, dth = 0.0L, dph = 0.0L;   // uninitialised in the reference; zero here
        // the iteration trace reflects at most once (the unconditional break
        // at geoac.eigenray.3d.cpp:206-208)
        const int lm_legs = std::min(nb, 1);
        for (int n = 0; n <= S_.iterations; n++) {
            if (n == S_.iterations) break;
            FullHit h = co_await full(theta, phi, lm_legs, EIG_LM);
            const bool brk = h.brk;
            if (brk && lg) std::fprintf(lg, "L %d %d %d %d %.17g %.17g nan b\n", ri, nb, seq, n, theta, phi);
            if (brk) {
                theta -= dth * step_sc;
                phi -= dph * step_sc;
                step_sc /= 2.0;
            }
            if (!brk) {
                const double *yk = h.y;
                x = yk[0];
                dx = rcvr[0] - x;
                y = yk[1];
                dy = rcvr[1] - y;
                dr = (double)std::sqrt(std::pow(dx, 2) + std::pow(dy, 2));
                if (lg)
                    std::fprintf(lg, "L %d %d %d %d %.17g %.17g %.17g %c\n", ri, nb, seq, n, theta, phi, dr,
                                 dr < S_.tolerance ? 'a' : (n > 0 && dr > dr_prev) ? 'h' : 's');
                if (dr < S_.tolerance) {
                    FullHit f = co_await full(theta, phi, nb, EIG_FINAL);
                    ga_arrival row = f.row;
                    row.launch_idx = ri;
                    row.n_bounce = nb;
                    found_.push_back(Found{ri, nb, seq, row, EigJob{theta, phi, nb, EIG_FINAL}});
                    break;
                } else if (n > 0 && dr > dr_prev) {
                    theta -= dth * step_sc;
                    phi -= dph * step_sc;
                    step_sc /= 2.0;
                } else {
                    step_sc = std::min(1.0, step_sc * 1.25);
                    c_src = S_.c_src;
                    // atmo::c at topo::z under row K: the trace result over a
                    // range-dependent medium, the flat ground's otherwise
                    c_grnd = S_.rngdep ? yk[24] : S_.c_grnd;
                    if (S_.is_topo) {   // geoac.eigenray.3d.cpp:337-342
                        dzg_dx = yk[22];
                        dzg_dy = yk[23];
                        ds_dth = -c_src / c_grnd * (yk[8] - dzg_dx * yk[6] - dzg_dy * yk[7]) /
                                 (yk[5] - dzg_dx * yk[3] - dzg_dy * yk[4]);
                        ds_dph = -c_src / c_grnd * (yk[14] - dzg_dx * yk[12] - dzg_dy * yk[13]) /
                                 (yk[5] - dzg_dx * yk[3] - dzg_dy * yk[4]);
                    } else {
                        ds_dth = -c_src / c_grnd * yk[8] / yk[5];
                        ds_dph = -c_src / c_grnd * yk[14] / yk[5];
                    }
                    dx_dth = yk[6] + yk[3] * ds_dth;
                    dy_dth = yk[7] + yk[4] * ds_dth;
                    dx_dph = yk[12] + yk[3] * ds_dph;
                    dy_dph = yk[13] + yk[4] * ds_dph;
                    det = std::pow(1.0 + S_.damping, 2) * (dx_dth * dy_dph - dx_dph * dy_dth);
                    dth = ((1.0 + S_.damping) * dy_dph * dx - dx_dph * dy) / det;
                    dph = ((1.0 + S_.damping) * dx_dth * dy - dy_dth * dx) / det;
                    theta += dth * step_sc;
                    phi += dph * step_sc;
                    dr_prev = dr;
                }
            }
            if (std::sqrt(dth * dth + dph * dph) * step_sc < 1.0e-12) break;
        }
    }

    // Spherical geometry: est_eigenray (geoac.eigenray.sph.cpp:41-154) inside
    // run_eig_search's loop (main.infraga.sph.cpp:962-989).  src = (z, lat,
    // lon), receivers (lat, lon), all in radians.
    Task chain_sph(int ri, int nb)
    {
        const double *src = S_.src;
        const double rcvr[2] = {rcvr_[2 * ri], rcvr_[2 * ri + 1]};
        const double rcvr_rng = gc_dist(src[1], src[2], rcvr[0], rcvr[1]);
        const double rcvr_az = bearing(src[1], src[2], rcvr[0], rcvr[1]);
        int seq = 0;
        double theta_start = S_.th_start;
        while (theta_start < S_.th_max) {
            double th_min = theta_start, th_max = S_.th_max;
            double th_est = 0.0, ph_est = 0.0, th_next = theta_start;
            int iterations = 0;
            double r, r_prev, dth = S_.dth_big, dph = 100.0;
            bool th_max_reached = false, success = false;
            double phi = PI / 2.0 - rcvr_az;
            while (std::fabs(dph * DPR) > S_.az_err_lim && iterations < 5) {
                r = rcvr_rng;
                r_prev = rcvr_rng;
                for (double th = th_min; th < th_max; th += dth) {
                    if (th + dth >= th_max) th_max_reached = true;
                    const int legs = (iterations == 0) ? S_.bnc_max : nb;
                    const int spec = (iterations == 0) ? (1 << 30) : (iterations <= 2 ? S_.spec_chunk : 0);
                    SweepHit h = co_await sweep(th, phi, legs, nb, dth, th_max, spec);   // h.x, h.y: lat, lon
                    r = gc_dist(src[1], src[2], h.x, h.y);
                    if (h.brk) r = S_.rng_max;
                    if (((r - rcvr_rng) * (r_prev - rcvr_rng) <= 0.0) && (r < S_.rng_max) && (r_prev < S_.rng_max) &&
                        (th > th_min)) {
                        if (iterations == 0) th_next = th;
                        dph = bearing(src[1], src[2], rcvr[0], rcvr[1]) * DPR;
                        dph -= bearing(src[1], src[2], h.x, h.y) * DPR;
                        while (dph > 180.0) dph -= 360.0;
                        while (dph < -180.0) dph += 360.0;
                        if (std::fabs(dph) < S_.az_err_lim) {
                            th_est = th - dth;
                            ph_est = phi;
                            success = true;
                        } else {
                            phi -= dph * RPD * 0.9;
                            th_min = std::max(th - 5.0 * RPD, th_min);
                        }
                        iterations++;
                        break;
                    }
                    if (iterations >= 3) dth = mod_dth(r - rcvr_rng, (r - r_prev) / (2.0 * dth));
                    r_prev = r;
                }
                if (success) break;
                if (th_max_reached) {
                    th_next = th_max;
                    break;
                }
            }
            if (success) spawn(find_sph(ri, nb, seq++, th_est, ph_est), ri);
            theta_start = th_next;
        }
    }

    // find_eigenray, spherical (geoac.eigenray.sph.cpp:157-380): every
    // reflection is traced, the update works in (lat, lon).
    Task find_sph(int ri, int nb, int seq, double th_est, double ph_est)
    {
        const double rcvr[2] = {rcvr_[2 * ri], rcvr_[2 * ri + 1]};
        double theta = th_est, phi = ph_est;
        double dr, dr_prev = 10000.0, step_sc = 1.0;
        long double r_grnd, c_src, c_grnd, dzg_dlat, dzg_dlon, ds_norm, ds_dth, ds_dph, dlat, dlon;
        long double dlat_dth, dlon_dth, dlat_dph, dlon_dph;
        long double det, dth = 0.0L, dph = 0.0L;   // uninitialised in the reference; zero here
        for (int n = 0; n <= S_.iterations; n++) {
            if (n == S_.iterations) break;
            FullHit h = co_await full(theta, phi, nb, EIG_LM);
            const bool brk = h.brk;
            if (brk) {
                theta -= dth * step_sc;
                phi -= dph * step_sc;
                step_sc /= 2.0;
            }
            if (!brk) {
                const double *yk = h.y;
                dr = gc_dist(yk[1], yk[2], rcvr[0], rcvr[1]);
                if (dr < S_.tolerance) {
                    FullHit f = co_await full(theta, phi, nb, EIG_FINAL);
                    ga_arrival row = f.row;
                    row.launch_idx = ri;
                    row.n_bounce = nb;
                    found_.push_back(Found{ri, nb, seq, row, EigJob{theta, phi, nb, EIG_FINAL}});
                    break;
                } else if (n > 0 && dr > dr_prev) {
                    theta -= dth * step_sc;
                    phi -= dph * step_sc;
                    step_sc /= 2.0;
                } else {
                    step_sc = std::min(1.0, step_sc * 1.25);
                    dlat = rcvr[0] - yk[1];
                    dlon = rcvr[1] - yk[2];
                    c_src = S_.c_src;
                    if (S_.rngdep) {
                        // topo::z, topo::dz and atmo::c at the ground under row K
                        r_grnd = yk[21];
                        c_grnd = yk[24];
                        dzg_dlat = (long double)yk[22] / r_grnd;
                        dzg_dlon = (long double)yk[23] / (r_grnd * std::cos(yk[1]));
                    } else {
                        r_grnd = S_.r0 + S_.src_ground;     // topo::z: globe radius + ground height
                        c_grnd = S_.c_grnd;
                        dzg_dlat = 0.0 / r_grnd;             // flat ground: topo::dz = 0
                        dzg_dlon = 0.0 / (r_grnd * std::cos(yk[1]));
                    }
                    ds_norm = yk[3] - dzg_dlat * yk[4] - dzg_dlon * yk[5];
                    ds_dth = -c_src / c_grnd * (yk[6] - dzg_dlat * r_grnd * yk[7] - dzg_dlon * (r_grnd * std::cos(yk[1])) * yk[8]) / ds_norm;
                    ds_dph = -c_src / c_grnd * (yk[12] - dzg_dlat * r_grnd * yk[13] - dzg_dlon * (r_grnd * std::cos(yk[1])) * yk[14]) / ds_norm;
                    dlat_dth = yk[7] + yk[4] / r_grnd * ds_dth;
                    dlon_dth = yk[8] + yk[5] / (r_grnd * std::cos(yk[1])) * ds_dth;
                    dlat_dph = yk[13] + yk[4] / r_grnd * ds_dph;
                    dlon_dph = yk[14] + yk[5] / (r_grnd * std::cos(yk[1])) * ds_dph;
                    // (1 + damping)^2 multiplies the first product only (sph.cpp:351)
                    det = std::pow(1.0 + S_.damping, 2) * dlat_dth * dlon_dph - dlat_dph * dlon_dth;
                    dth = ((1.0 + S_.damping) * dlon_dph * dlat - dlat_dph * dlon) / det;
                    dph = ((1.0 + S_.damping) * dlat_dth * dlon - dlon_dth * dlat) / det;
                    theta += dth * step_sc;
                    phi += dph * step_sc;
                    dr_prev = dr;
                }
            }
            if (std::sqrt(dth * dth + dph * dph) * step_sc < 1.0e-12) break;
        }
    }

    struct Found {
        int rcvr, nb, seq;
        ga_arrival row;
        EigJob job;                // the accepted trace (exact launch angles)
    };
    static bool found_order(const Found &a, const Found &b)
    {
        if (a.rcvr != b.rcvr) return a.rcvr < b.rcvr;
        if (a.nb != b.nb) return a.nb < b.nb;
        return a.seq < b.seq;
    }

    struct Direct {
        int rcvr, nb;
        double th, ph;
    };
    std::vector<Direct> direct_;
    EigSetup S_;
    int n_rcvr_;
    int stride_;
    std::vector<double> rcvr_;
    std::unordered_map<Key, uint32_t, KeyHash> sweep_idx_, full_idx_;
    std::vector<double> sweep_store_, full_store_;
    std::vector<int32_t> sweep_brk_, full_brk_;
    std::vector<ga_arrival> final_rows_;
    std::vector<Key> req_sweep_, req_full_;
    std::vector<Handle> ready_, waiting_;
    std::vector<int> live_;
    std::vector<double> done_ms_;
    std::vector<int> newly_done_;
    std::vector<double> init_xy_;
    std::vector<double> *lat_ = nullptr;
    std::vector<Found> found_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace ga
