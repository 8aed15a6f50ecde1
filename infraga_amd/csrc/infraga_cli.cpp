// infraga_cli.cpp -- drop-in front ends infraga-sph / infraga-3d /
// infraga-sph-rngdep / infraga-3d-rngdep / infraga-2d on the GPU engine.
//
// Same command line as the reference drivers (src/main.infraga.sph.cpp,
// src/main.infraga.3d.cpp, src/main.infraga.sph.rngdep.cpp,
// src/main.infraga.3d.rngdep.cpp, src/main.infraga.2d.cpp): an option
// (-prop, -eig_search, -eig_direct, -wnl_wvfrm; the range-dependent binaries take the
// profile prefix and the two node files), the profile path, then
// key=value parameters with the same names, aliases and defaults.  The
// ray tracing runs on the MI355X through the C ABI of libinfraga_gpu.so; this
// file only parses arguments, builds the launch list with the reference's
// accumulated loops and writes the output files in the reference's format.
//
// Differences from the reference drivers, by design:
//  - an unknown key is an error (exit 2) instead of an interactive prompt;
//  - write_topo is rejected;
//  - verbose=true prints no trace-by-trace log.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "infraga_gpu.h"

#ifndef GA_CLI_GEOM
#define GA_CLI_GEOM GA_GEOM_SPH
#endif
#ifndef GA_CLI_NAME
#define GA_CLI_NAME "infraga-sph"
#endif

namespace {

constexpr double Pi = 3.141592653589793238462643;

bool str2bool(const char *v)
{
    return (std::strcmp(v, "true") == 0 || std::strcmp(v, "True") == 0 || std::strcmp(v, "TRUE") == 0 ||
            std::atoi(v) != 0);
}

bool key(const char *arg, const char *k, const char *&val)
{
    size_t n = std::strlen(k);
    if (std::strncmp(arg, k, n) == 0) { val = arg + n; return true; }
    return false;
}

bool key2(const char *arg, const char *k1, const char *k2, const char *&val)
{
    return key(arg, k1, val) || key(arg, k2, val);
}

int die(ga_ctx *ctx, const char *what)
{
    std::cerr << GA_CLI_NAME << ": " << what;
    if (ctx) std::cerr << ": " << ga_last_error(ctx);
    std::cerr << '\n';
    if (ctx) ga_ctx_destroy(ctx);
    return 1;
}

void usage()
{
    std::cout << "Usage: " << GA_CLI_NAME << " [option] profile.met [parameters]\n"
              << "  -prop        ray paths for a fan of inclinations and azimuths\n"
              << "  -eig_search  eigenrays between the source and a receiver\n"
              << "  -eig_direct  one eigenray from a launch-angle estimate (incl_est, az_est)\n"
              << "  -wnl_wvfrm   weakly non-linear waveform along one ray\n"
              << "Parameters use the key=value names of the reference infraga drivers.\n";
}

// Strip the extension in place, as the reference does (main.infraga.sph.cpp:339-347):
// the header line "profile:" then shows the stripped path.
std::string strip_ext(char *path)
{
    for (int m = (int)std::strlen(path); m >= 0; m--) {
        if (path[m] == '.') { path[m] = '\0'; break; }
    }
    return std::string(path);
}

double profile_bottom(const char *path);

#ifndef GA_CLI_ACCEL
#define GA_CLI_ACCEL 0
#endif

// Rank and size of an mpirun launch (infraga-accel-*: MPICH/Hydra PMI_*,
// Open MPI OMPI_COMM_WORLD_*, PMIx PMIX_RANK); 0 and 1 otherwise.
int env_int(const char *const *names, int dflt)
{
    for (; *names; ++names)
        if (const char *v = std::getenv(*names)) return std::atoi(v);
    return dflt;
}
int mpi_rank()
{
    static const char *const n[] = {"PMI_RANK", "OMPI_COMM_WORLD_RANK", "PMIX_RANK", "MV2_COMM_WORLD_RANK", nullptr};
    return GA_CLI_ACCEL ? env_int(n, 0) : 0;
}
int mpi_size()
{
    static const char *const n[] = {"PMI_SIZE", "OMPI_COMM_WORLD_SIZE", "MV2_COMM_WORLD_SIZE", nullptr};
    return GA_CLI_ACCEL ? std::max(1, env_int(n, 1)) : 1;
}

// The inclinations launched at one azimuth.  Serial drivers accumulate
// theta += step up to max(theta_max, theta_grnd) (main.infraga.sph.cpp:434,
// main.infraga.3d.cpp:421); the MPI drivers step by step x ranks and launch
// theta + step x j on rank j while that stays <= theta_max, with the outer
// bound max(theta_max, theta_grnd) (main.infraga.sph.mpi.cpp:388-399,
// main.infraga.sph.rngdep.mpi.cpp:381-392) or theta_max in 3D
// (main.infraga.3d.mpi.cpp:369-380).
void incl_loop(double theta_min, double theta_max, double theta_grnd, double step, double phi,
               std::vector<double> &launches)
{
    const int world = mpi_size();
    const double top = (world > 1 && GA_CLI_GEOM == GA_GEOM_3D) ? theta_max : std::max(theta_max, theta_grnd);
    if (world <= 1) {
        for (double theta = std::max(theta_min, theta_grnd); theta <= top; theta += step) {
            launches.push_back(theta);
            launches.push_back(phi);
        }
        return;
    }
    for (double theta = std::max(theta_min, theta_grnd); theta <= top; theta += step * world) {
        for (int j = 0; j < world; ++j) {
            if (theta + step * j > theta_max) continue;
            launches.push_back(theta + step * j);
            launches.push_back(phi);
        }
    }
}

// The fan of run_prop with write_rays / write_caustics
// (main.infraga.sph.cpp:388-413, 456-496, 549; main.infraga.3d.cpp:378-400,
// 442-487; main.infraga.sph.rngdep.cpp:382-406, 451-490, 544): rays go to the
// engine in batches; <id>.raypaths.dat gets each ray's sampled rows and a
// blank line, <id>.caustics-<leg>.dat the caustic points of each leg.  The
// reference's sph / range-dependent writers print at precision 8 (the
// setprecision of the first column sticks), the 3D ones at the stream
// default 6.  arr receives the arrivals in launch order.
int fan_with_paths(ga_ctx *ctx, int aid, const ga_prop_params &p, const std::vector<double> &launches,
                   bool write_rays, bool write_caustics, const std::string &id, bool sph,
                   std::vector<ga_arrival> &arr, int &n_out, ga_run_stats &st, bool rd3 = false)
{
    const int n = (int)(launches.size() / 2);
    const int slots = p.bounces + 1;
    std::ofstream raypath;
    std::vector<std::ofstream> caus;
    if (write_rays) {
        raypath.open(id + ".raypaths.dat");
        raypath << (sph ? "# lat [deg]" : "# x [km]") << '\t' << (sph ? "lon [deg]" : "y [km]") << '\t' << "z [km]"
                << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t' << "time [s]" << '\n';
        if (sph) raypath << std::setprecision(8);
    }
    if (write_caustics) {
        for (int b = 0; b <= p.bounces; ++b) {
            caus.emplace_back(id + ".caustics-" + std::to_string(b) + ".dat");
            caus.back() << (sph ? "# lat [deg]" : "# x [km]") << '\t' << (sph ? "lon [deg]" : "y [km]") << '\t'
                        << "z [km]" << '\t' << "time [s]" << '\n';
            if (sph) caus.back() << std::setprecision(8);
        }
    }
    const int flags = (write_rays ? GA_PATH_RAYS : 0) | (write_caustics ? GA_PATH_CAUSTICS : 0);
    const int batch = 4096;
    int64_t cap = (int64_t)1 << 21;
    std::vector<ga_path_pt> pts;
    arr.assign((size_t)std::max(1, n) * slots, ga_arrival{});
    n_out = 0;
    st = ga_run_stats{};
    for (int r0 = 0; r0 < n; r0 += batch) {
        const int nb = std::min(batch, n - r0);
        int nb_out = 0;
        int64_t n_pts = 0;
        ga_run_stats bs;
        int rc;
        for (;;) {
            pts.resize((size_t)cap);
            rc = ga_fan_run_paths(ctx, aid, &p, launches.data() + 2 * (size_t)r0, nb, flags, arr.data() + n_out,
                                  nb * slots, &nb_out, pts.data(), cap, &n_pts, &bs);
            if (rc == GA_E_CAP && n_pts > cap) { cap = n_pts + n_pts / 4; continue; }
            break;
        }
        if (rc != GA_OK) return rc;
        for (int i = 0; i < nb_out; ++i) arr[n_out + i].launch_idx += r0;
        n_out += nb_out;
        st.ray_steps += bs.ray_steps;
        st.kernel_ms += bs.kernel_ms;
        st.n_rays += bs.n_rays;
        st.n_launches += bs.n_launches;
        int64_t w = 0;
        for (int r = 0; r < nb; ++r) {
            for (; w < n_pts && pts[w].launch_idx == r; ++w) {
                const ga_path_pt &P = pts[w];
                if (P.kind == GA_PT_RAY && write_rays) {
                    raypath << P.pos0 << '\t' << P.pos1 << '\t' << P.z << '\t' << P.tl << '\t' << P.absorption
                            << '\t' << P.time << '\n';
                } else if (P.kind == GA_PT_CAUSTIC && write_caustics && P.leg >= 0 && P.leg < (int)caus.size()) {
                    // the 3D range-dependent writer puts a 0.0 column before the time
                    // (main.infraga.3d.rngdep.cpp:467-471)
                    caus[P.leg] << P.pos0 << '\t' << P.pos1 << '\t' << P.z << '\t';
                    if (rd3) caus[P.leg] << 0.0 << '\t';
                    caus[P.leg] << P.time << '\n';
                }
            }
            if (write_rays) raypath << '\n';
        }
    }
    return GA_OK;
}

// .arrivals.dat of the stratified drivers (main.infraga.sph.cpp:354-384 and
// 468-556, main.infraga.3d.cpp): header, then one row per arrival and a blank
// line after each azimuth; z0 for flat ground, topo_file with terrain.
void write_arrivals(const std::string &id, const char *profile, double theta_min, double theta_max, double theta_step,
                    double phi_min, double phi_max, double phi_step, int bounces, double s0, double s1, double z_src,
                    const double *z0, const char *topo_file, double turn_ht_min, double freq, double abs_coeff,
                    bool sph, const std::vector<ga_arrival> &arr, int n_out, int n, const std::vector<int> &az_end)
{
    std::ofstream results(id + ".arrivals.dat");
    results << "# " << GA_CLI_NAME << " -prop summary:" << '\n';
    results << "#" << '\t' << "profile: " << profile << '\n';
    results << "#" << '\t' << "inclination: " << theta_min << ", " << theta_max << ", " << theta_step << '\n';
    results << "#" << '\t' << "azimuth: " << phi_min << ", " << phi_max << ", " << phi_step << '\n';
    results << "#" << '\t' << "bounces: " << bounces << '\n';
    if (sph) results << "#" << '\t' << "source location (lat, lon, alt): " << s0 << ", " << s1 << ", " << z_src << '\n';
    else results << "#" << '\t' << "source location: " << s0 << ", " << s1 << ", " << z_src << '\n';
    if (topo_file) results << "#" << '\t' << "topo file:" << topo_file << '\n';
    else results << "#" << '\t' << "ground elevation: " << *z0 << '\n';
    if (sph) results << "#" << '\t' << "turn_ht_min: " << turn_ht_min << '\n';
    results << "#" << '\t' << "frequency: " << freq << '\n';
    results << "#" << '\t' << "S&B atten coeff: " << abs_coeff << '\n' << '\n';
    results << "# incl [deg]" << '\t' << "az [deg]" << '\t' << "n_b";
    if (sph) results << '\t' << "lat_0 [deg]" << '\t' << "lon_0 [deg]";
    else results << '\t' << "x_0 [km]" << '\t' << "y_0 [km]";
    results << '\t' << "time [s]" << '\t' << "cel [km/s]" << '\t' << "turning ht [km]";
    if (sph) results << '\t' << "inclination [deg]" << '\t' << "back azimuth [deg]";
    else results << '\t' << "arrival incl [deg]" << '\t' << "back az [deg]";
    results << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\n';

    int w = 0, az_i = 0;
    for (int r = 0; r < n; ++r) {
        while (w < n_out && arr[w].launch_idx == r) {
            const ga_arrival &A = arr[w++];
            results << A.incl;
            results << '\t' << A.az;
            results << '\t' << A.n_bounce;
            if (sph) {
                results << '\t' << std::setprecision(8) << A.pos0;
                results << '\t' << std::setprecision(8) << A.pos1;
            } else {
                results << '\t' << A.pos0;
                results << '\t' << A.pos1;
            }
            results << '\t' << A.time << '\t' << A.cel << '\t' << A.turn_ht << '\t' << A.incl_arr << '\t' << A.back_az;
            results << '\t' << A.tl << '\t' << A.absorption << '\n';
        }
        if (r + 1 == az_end[az_i]) { results << '\n'; ++az_i; }
    }
}

int run_prop(int argc, char **argv)
{
    const bool sph = GA_CLI_GEOM == GA_GEOM_SPH;
    double theta_min = 0.5, theta_max = 45.0, theta_step = 0.5;
    double phi_min = -90.0, phi_max = -90.0, phi_step = 1.0;
    int bounces = 10;
    double s0 = sph ? 30.0 : 0.0, s1 = 0.0, z_src = 0.0;
    bool write_rays = true, write_caustics = false, reverse_winds = false, custom_id = false;
    double freq = 0.1, turn_ht_min = 0.2;
    const char *prof_format = "zTuvdp";
    const char *output_id = nullptr;
    ga_prop_params p;
    ga_prop_params_default(&p, GA_CLI_GEOM);
    double z_grnd = NAN;
    const char *topo_file = nullptr;
    bool src_set[2] = {false, false}, bounds_set[4] = {false, false, false, false}, rng_set = false;

    for (int i = 3; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key2(a, "incl_min=", "min_incl=", v)) theta_min = std::atof(v);
        else if (key2(a, "incl_max=", "max_incl=", v)) theta_max = std::atof(v);
        else if (key2(a, "az_min=", "min_az=", v)) phi_min = std::atof(v);
        else if (key2(a, "az_max=", "max_az=", v)) phi_max = std::atof(v);
        else if (key(a, "incl_step=", v)) theta_step = std::atof(v);
        else if (key(a, "az_step=", v)) phi_step = std::atof(v);
        else if (key(a, "inclination=", v)) { theta_min = theta_max = std::atof(v); theta_step = 1.0; }
        else if (key(a, "azimuth=", v)) { phi_min = phi_max = std::atof(v); phi_step = 1.0; }
        else if (key(a, "bounces=", v)) bounces = std::atoi(v);
        else if (key2(a, "src_alt=", "alt_src=", v)) z_src = std::atof(v);
        else if (sph && key2(a, "src_lat=", "lat_src=", v)) { s0 = std::atof(v); src_set[0] = true; }
        else if (sph && key2(a, "src_lon=", "lon_src=", v)) { s1 = std::atof(v); src_set[1] = true; }
        else if (!sph && key2(a, "src_x=", "x_src=", v)) { s0 = std::atof(v); src_set[0] = true; }
        else if (!sph && key2(a, "src_y=", "y_src=", v)) { s1 = std::atof(v); src_set[1] = true; }
        else if (key(a, "turn_ht_min=", v)) turn_ht_min = std::atof(v);
        else if (key(a, "write_rays=", v)) write_rays = str2bool(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key(a, "calc_amp=", v)) p.calc_amp = str2bool(v) ? 1 : 0;
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) { p.rng_max = std::atof(v); rng_set = true; }
        else if (sph && key2(a, "max_time=", "time_max=", v)) p.time_max = std::atof(v);
        else if (sph && key2(a, "min_lat=", "lat_min=", v)) { p.bounds[0] = std::atof(v); bounds_set[0] = true; }
        else if (sph && key2(a, "max_lat=", "lat_max=", v)) { p.bounds[1] = std::atof(v); bounds_set[1] = true; }
        else if (sph && key2(a, "min_lon=", "lon_min=", v)) { p.bounds[2] = std::atof(v); bounds_set[2] = true; }
        else if (sph && key2(a, "max_lon=", "lon_max=", v)) { p.bounds[3] = std::atof(v); bounds_set[3] = true; }
        else if (!sph && key2(a, "min_x=", "x_min=", v)) { p.bounds[0] = std::atof(v); bounds_set[0] = true; }
        else if (!sph && key2(a, "max_x=", "x_max=", v)) { p.bounds[1] = std::atof(v); bounds_set[1] = true; }
        else if (!sph && key2(a, "min_y=", "y_min=", v)) { p.bounds[2] = std::atof(v); bounds_set[2] = true; }
        else if (!sph && key2(a, "max_y=", "y_max=", v)) { p.bounds[3] = std::atof(v); bounds_set[3] = true; }
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "print_resid=", v)) { /* diagnostics only */ }
        else if (key(a, "write_caustics=", v)) write_caustics = str2bool(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else if (key(a, "topo_use_BLw=", v)) p.topo_use_blw = str2bool(v) ? 1 : 0;
        else if (key(a, "write_topo=", v)) {
            if (str2bool(v)) {
                std::cerr << GA_CLI_NAME << ": " << a << " is not supported by this engine\n";
                return 2;
            }
        } else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }

    if (write_caustics) p.calc_amp = 1;   // caustics need the Jacobian (main.infraga.sph.cpp:315, 3d.cpp:308)
    if (!topo_file) p.topo_use_blw = 0;   // the vertical wind needs terrain (atmo::w)
    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int aid = 0;
    if (topo_file) {
        // the stratified profile over terrain (set_region with topo_file,
        // main.infraga.sph.cpp:231-240, main.infraga.3d.cpp:229-238): the
        // range-dependent kernels over identical columns with the
        // stratified semantics (terrain nodes in km for the 3D binary)
        if ((sph ? ga_atmo_load_file_topo : ga_atmo_load_file_topo_xy)(ctx, argv[2], topo_file, prof_format,
                                                                       reverse_winds ? 1 : 0, &aid) != GA_OK)
            return die(ctx, "cannot load the atmosphere and terrain");
        ga_prop_params q;
        ga_prop_params_default(&q, sph ? GA_GEOM_SPH_RNGDEP : GA_GEOM_3D_RNGDEP);
        q.calc_amp = p.calc_amp;
        q.abs_coeff = p.abs_coeff;
        q.alt_max = p.alt_max;
        // the 3D set_limits leaves rng_max at 1000 with terrain (atmo_io.3d.strat.cpp:26-33)
        q.rng_max = (sph || rng_set) ? p.rng_max : 1000.0;
        q.time_max = p.time_max;
        q.ds_min = p.ds_min;
        q.ds_max = p.ds_max;
        q.s_max = p.s_max;
        q.topo_use_blw = p.topo_use_blw;
        for (int k = 0; k < 4; ++k) q.bounds[k] = bounds_set[k] ? p.bounds[k] : NAN;
        q.bounces = bounces;
        q.freq = freq;
        q.turn_ht_min = turn_ht_min;
        double info[8];
        if (ga_rngdep_info(ctx, aid, info) != GA_OK) return die(ctx, "region");
        // source at the terrain's centre unless given (sph :236-239, 3D :235-238)
        const double deg = sph ? Pi / 180.0 : 1.0;
        if (!src_set[0]) s0 = (info[1] + info[0]) / 2.0 / deg;
        if (!src_set[1]) s1 = (info[3] + info[2]) / 2.0 / deg;
        const double la = s0 * deg, lo = s1 * deg;
        double pt[3] = {la, lo, 0.0}, atmo_out[40], topo_out[10];
        if (ga_rngdep_eval(ctx, aid, pt, 1, atmo_out, topo_out) != GA_OK) return die(ctx, "terrain at the source");
        const double zg = topo_out[0];
        z_src = std::max(z_src, zg);                                          // :313
        q.src[0] = s0;
        q.src[1] = s1;
        q.src[2] = z_src;
        std::vector<double> launches;
        std::vector<int> az_end;
        const double r0 = q.r_earth;
        for (double phi = phi_min; phi <= phi_max; phi += phi_step) {
            double theta_grnd = -89.9;   // sph :411-423, 3D :408-417
            if (z_src - zg < 0.01) {
                const double dzgdx = sph ? 1.0 / (r0 * std::cos(la)) * topo_out[2] : topo_out[1];
                const double dzgdy = sph ? 1.0 / r0 * topo_out[1] : topo_out[2];
                theta_grnd = std::atan(dzgdx * std::sin(phi * Pi / 180.0) + dzgdy * std::cos(phi * Pi / 180.0)) * (180.0 / Pi) + 0.1;
            }
            incl_loop(theta_min, theta_max, theta_grnd, theta_step, phi, launches);
            az_end.push_back((int)(launches.size() / 2));
        }
        const int n = (int)(launches.size() / 2);
        std::vector<ga_arrival> arr((size_t)std::max(1, n) * (bounces + 1));
        int n_out = 0;
        ga_run_stats st;
        std::string id = custom_id ? std::string(output_id) : strip_ext(argv[2]);
        if (write_rays || write_caustics) {
            if (fan_with_paths(ctx, aid, q, launches, write_rays, write_caustics, id, sph, arr, n_out, st) != GA_OK)
                return die(ctx, "fan failed");
        } else if (ga_fan_run(ctx, aid, &q, launches.data(), n, arr.data(), (int)arr.size(), &n_out, &st) != GA_OK) {
            return die(ctx, "fan failed");
        }
        write_arrivals(id, argv[2], theta_min, theta_max, theta_step, phi_min, phi_max, phi_step, bounces, s0, s1,
                       z_src, nullptr, topo_file, turn_ht_min, freq, q.abs_coeff, sph, arr, n_out, n, az_end);
        std::cout << GA_CLI_NAME << ": " << n << " rays, " << n_out << " arrivals, " << st.ray_steps << " ray-steps, "
                  << st.kernel_ms << " ms on the GPU\n";
        ga_ctx_destroy(ctx);
        return 0;
    }
    if (ga_atmo_load_file(ctx, argv[2], prof_format, reverse_winds ? 1 : 0, &aid) != GA_OK)
        return die(ctx, "cannot load the atmosphere");

    // ground: the profile bottom unless z_grnd is given (atmo_io set_limits + driver)
    double z0 = std::isnan(z_grnd) ? profile_bottom(argv[2]) : z_grnd;
    p.z_grnd = z0;
    p.bounces = bounces;
    p.freq = freq;
    p.turn_ht_min = turn_ht_min;
    double zg_alt = sph ? ((p.r_earth + z0) - p.r_earth) : z0;
    z_src = std::max(z_src, zg_alt);
    p.src[0] = s0;
    p.src[1] = s1;
    p.src[2] = z_src;

    // launch list: the reference's accumulated loops (main.infraga.sph.cpp:418-434)
    std::vector<double> launches;
    std::vector<int> az_end;      // launch count after each azimuth (blank lines)
    for (double phi = phi_min; phi <= phi_max; phi += phi_step) {
        double theta_grnd = (z_src - zg_alt < 0.01) ? 0.1 : -89.9;
        incl_loop(theta_min, theta_max, theta_grnd, theta_step, phi, launches);
        az_end.push_back((int)(launches.size() / 2));
    }
    const int n = (int)(launches.size() / 2);
    std::vector<ga_arrival> arr((size_t)std::max(1, n) * (bounces + 1));
    int n_out = 0;
    ga_run_stats st;
    std::string id = custom_id ? std::string(output_id) : strip_ext(argv[2]);
    if (write_rays || write_caustics) {
        if (fan_with_paths(ctx, aid, p, launches, write_rays, write_caustics, id, sph, arr, n_out, st) != GA_OK)
            return die(ctx, "fan failed");
    } else if (ga_fan_run(ctx, aid, &p, launches.data(), n, arr.data(), (int)arr.size(), &n_out, &st) != GA_OK) {
        return die(ctx, "fan failed");
    }

    write_arrivals(id, argv[2], theta_min, theta_max, theta_step, phi_min, phi_max, phi_step, bounces, s0, s1, z_src,
                   &z0, nullptr, turn_ht_min, freq, p.abs_coeff, sph, arr, n_out, n, az_end);
    std::cout << GA_CLI_NAME << ": " << n << " rays, " << n_out << " arrivals, " << st.ray_steps << " ray-steps, "
              << st.kernel_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// infraga-sph-rngdep -prop prefix lats.dat lons.dat [key=value ...]
// (run_prop, src/main.infraga.sph.rngdep.cpp:205-562).
int run_prop_rd(int argc, char **argv)
{
    // infraga-3d-rngdep (main.infraga.3d.rngdep.cpp:205-540): the same run
    // over x / y node files in km
    constexpr bool cart = GA_CLI_GEOM == GA_GEOM_3D_RNGDEP;
    if (argc < 5) {
        std::cerr << GA_CLI_NAME << (cart ? ": usage: -prop <profile prefix> <x nodes> <y nodes> [parameters]\n"
                                          : ": usage: -prop <profile prefix> <lat nodes> <lon nodes> [parameters]\n");
        return 2;
    }
    double theta_min = 0.5, theta_max = 45.0, theta_step = 0.5;
    double phi_min = -90.0, phi_max = -90.0, phi_step = 1.0;
    int bounces = cart ? 10 : 2;
    double lat_src = NAN, lon_src = NAN, z_src = 0.0;
    bool write_rays = true, write_caustics = false, reverse_winds = false, custom_id = false;
    double freq = 0.1, turn_ht_min = 0.2;
    const char *prof_format = "zTuvdp";
    const char *topo_file = nullptr;
    const char *output_id = nullptr;
    ga_prop_params p;
    ga_prop_params_default(&p, cart ? GA_GEOM_3D_RNGDEP : GA_GEOM_SPH_RNGDEP);
    p.calc_amp = 1;                                    // run_prop :226
    double z_grnd = NAN;

    for (int i = 5; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key2(a, "incl_min=", "min_incl=", v)) theta_min = std::atof(v);
        else if (key2(a, "incl_max=", "max_incl=", v)) theta_max = std::atof(v);
        else if (key2(a, "az_min=", "min_az=", v)) phi_min = std::atof(v);
        else if (key2(a, "az_max=", "max_az=", v)) phi_max = std::atof(v);
        else if (key(a, "incl_step=", v)) theta_step = std::atof(v);
        else if (key(a, "az_step=", v)) phi_step = std::atof(v);
        else if (key(a, "inclination=", v)) { theta_min = theta_max = std::atof(v); theta_step = 1.0; }
        else if (key(a, "azimuth=", v)) { phi_min = phi_max = std::atof(v); phi_step = 1.0; }
        else if (key(a, "bounces=", v)) bounces = std::atoi(v);
        else if (key2(a, "src_alt=", "alt_src=", v)) z_src = std::atof(v);
        else if (!cart && key2(a, "src_lat=", "lat_src=", v)) lat_src = std::atof(v);
        else if (!cart && key2(a, "src_lon=", "lon_src=", v)) lon_src = std::atof(v);
        else if (cart && key2(a, "src_x=", "x_src=", v)) lat_src = std::atof(v);
        else if (cart && key2(a, "src_y=", "y_src=", v)) lon_src = std::atof(v);
        else if (key(a, "turn_ht_min=", v)) turn_ht_min = std::atof(v);
        else if (key(a, "write_rays=", v)) write_rays = str2bool(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key(a, "calc_amp=", v)) {
            // the 3D binary compares 8 characters and parses "=<value>",
            // which string2bool reads as false (main.infraga.3d.rngdep.cpp:269)
            p.calc_amp = (!cart && str2bool(v)) ? 1 : 0;
        }
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) p.rng_max = std::atof(v);
        else if (!cart && key2(a, "max_time=", "time_max=", v)) p.time_max = std::atof(v);   // no time limit in 3D (:498)
        else if (key2(a, cart ? "min_x=" : "min_lat=", cart ? "x_min=" : "lat_min=", v)) p.bounds[0] = std::atof(v);
        else if (key2(a, cart ? "max_x=" : "max_lat=", cart ? "x_max=" : "lat_max=", v)) p.bounds[1] = std::atof(v);
        else if (key2(a, cart ? "min_y=" : "min_lon=", cart ? "y_min=" : "lon_min=", v)) p.bounds[2] = std::atof(v);
        else if (key2(a, cart ? "max_y=" : "max_lon=", cart ? "y_max=" : "lon_max=", v)) p.bounds[3] = std::atof(v);
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else if (key(a, "topo_use_BLw=", v)) p.topo_use_blw = str2bool(v) ? 1 : 0;
        else if (key(a, "print_resid=", v)) { /* diagnostics only */ }
        else if (key(a, "write_caustics=", v)) write_caustics = str2bool(v);
        else if (key(a, "write_topo=", v)) {
            if (str2bool(v)) {
                std::cerr << GA_CLI_NAME << ": " << a << " is not supported by this engine\n";
                return 2;
            }
        } else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }

    if (write_caustics) p.calc_amp = 1;   // caustics need the Jacobian (main.infraga.sph.cpp:315, 3d.cpp:308)
    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int rid = 0;
    if ((cart ? ga_rngdep_load_files_xy : ga_rngdep_load_files)(ctx, argv[2], argv[3], argv[4], topo_file, prof_format,
                                                               reverse_winds ? 1 : 0, &rid) != GA_OK)
        return die(ctx, "cannot load the range-dependent atmosphere");
    double info[8];
    if (ga_rngdep_info(ctx, rid, info) != GA_OK) return die(ctx, "region");
    const bool is_topo = topo_file != nullptr;
    const double deg = cart ? 1.0 : Pi / 180.0;   // region in radians (sph) or km (3D)
    if (std::isnan(lat_src)) lat_src = (info[1] + info[0]) / 2.0 / deg;   // :242-243
    if (std::isnan(lon_src)) lon_src = (info[3] + info[2]) / 2.0 / deg;
    const double z0 = (!is_topo && !std::isnan(z_grnd)) ? z_grnd : info[5];
    if (!is_topo) p.z_grnd = z0;

    // ground under the source and its slope (topo::z, topo::dz)
    const double la = lat_src * deg, lo = lon_src * deg;
    double pt[3] = {la, lo, 0.0}, atmo_out[40], topo_out[10];
    if (ga_rngdep_eval(ctx, rid, pt, 1, atmo_out, topo_out) != GA_OK) return die(ctx, "terrain at the source");
    // without terrain the spherical ground is z0, the Cartesian topo::z is 0
    // (atmo_state.3d.rngdep.cpp:45-52)
    const double zg = is_topo ? topo_out[0] : (cart ? 0.0 : z0);
    z_src = std::max(z_src, zg);                                                 // :315, 3D :308

    p.bounces = bounces;
    p.freq = freq;
    p.turn_ht_min = turn_ht_min;
    p.src[0] = lat_src;
    p.src[1] = lon_src;
    p.src[2] = z_src;

    // launch list: the reference's accumulated loops with the terrain-slope
    // floor on the inclination (:412-428)
    std::vector<double> launches;
    std::vector<int> az_end;
    const double r0 = p.r_earth;
    for (double phi = phi_min; phi <= phi_max; phi += phi_step) {
        double theta_grnd;
        if (z_src - zg < 0.01) {
            if (is_topo) {
                const double dzgdx = cart ? topo_out[1] : 1.0 / (r0 * std::cos(la)) * topo_out[2];
                const double dzgdy = cart ? topo_out[2] : 1.0 / r0 * topo_out[1];
                theta_grnd = std::atan(dzgdx * std::sin(phi * Pi / 180.0) + dzgdy * std::cos(phi * Pi / 180.0)) * (180.0 / Pi) + 0.1;
            } else {
                theta_grnd = 0.1;
            }
        } else {
            theta_grnd = -89.9;
        }
        incl_loop(theta_min, theta_max, theta_grnd, theta_step, phi, launches);
        az_end.push_back((int)(launches.size() / 2));
    }
    const int n = (int)(launches.size() / 2);
    std::vector<ga_arrival> arr((size_t)std::max(1, n) * (bounces + 1));
    int n_out = 0;
    ga_run_stats st;
    std::string id = custom_id ? std::string(output_id) : std::string(argv[2]);
    if (write_rays || write_caustics) {
        if (fan_with_paths(ctx, rid, p, launches, write_rays, write_caustics, id, !cart, arr, n_out, st, cart) != GA_OK)
            return die(ctx, "fan failed");
    } else if (ga_fan_run(ctx, rid, &p, launches.data(), n, arr.data(), (int)arr.size(), &n_out, &st) != GA_OK) {
        return die(ctx, "fan failed");
    }

    std::ofstream results(id + ".arrivals.dat");
    results << "# " << GA_CLI_NAME << " -prop summary:" << '\n';
    if (cart) {   // main.infraga.3d.rngdep.cpp:346-354
        results << "#" << '\t' << "profile prefix: " << argv[2] << '\n';
        results << "#" << '\t' << "profile grid files: " << argv[3] << ", " << argv[4] << '\n';
    } else {
        results << "#" << '\t' << "profile: " << argv[2] << '\n';
    }
    results << "#" << '\t' << "inclination: " << theta_min << ", " << theta_max << ", " << theta_step << '\n';
    results << "#" << '\t' << "azimuth: " << phi_min << ", " << phi_max << ", " << phi_step << '\n';
    results << "#" << '\t' << "bounces: " << bounces << '\n';
    results << "#" << '\t' << (cart ? "source location: " : "source location (lat, lon, alt): ") << lat_src << ", " << lon_src << ", " << z_src << '\n';
    if (!is_topo) results << "#" << '\t' << "ground elevation: " << z0 << '\n';
    else results << "#" << '\t' << "topo file:" << topo_file << '\n';
    results << "#" << '\t' << "frequency: " << freq << '\n';
    results << "#" << '\t' << "S&B atten coeff: " << p.abs_coeff << '\n' << '\n';
    results << "# incl [deg]" << '\t' << "az [deg]" << '\t' << "n_b" << '\t' << (cart ? "x_0 [km]" : "lat_0 [deg]") << '\t'
            << (cart ? "y_0 [km]" : "lon_0 [deg]")
            << '\t' << "time [s]" << '\t' << "cel [km/s]" << '\t' << "turning ht [km]" << '\t'
            << (cart ? "arrival incl [deg]" : "inclination [deg]") << '\t' << (cart ? "back az [deg]" : "back azimuth [deg]")
            << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\n';
    int w = 0, az_i = 0;
    for (int r = 0; r < n; ++r) {
        while (w < n_out && arr[w].launch_idx == r) {
            const ga_arrival &A = arr[w++];
            results << A.incl << '\t' << A.az << '\t' << A.n_bounce;
            // the spherical writer's setprecision(8) sticks; the 3D one prints
            // at the stream default 6 (main.infraga.3d.rngdep.cpp:506-520)
            if (!cart) results << std::setprecision(8);
            results << '\t' << A.pos0 << '\t' << A.pos1;
            results << '\t' << A.time << '\t' << A.cel << '\t' << A.turn_ht << '\t' << A.incl_arr << '\t' << A.back_az;
            results << '\t' << A.tl << '\t' << A.absorption << '\n';
        }
        if (r + 1 == az_end[az_i]) { results << '\n'; ++az_i; }
    }
    results.close();
    std::cout << GA_CLI_NAME << ": " << n << " rays, " << n_out << " arrivals, " << st.ray_steps << " ray-steps, "
              << st.kernel_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// The profile's lowest altitude (the default ground height).
double profile_bottom(const char *path)
{
    std::ifstream probe(path);
    double z_bottom = 0.0;
    std::string line;
    while (std::getline(probe, line)) {
        if (line.find('#') == 0) continue;
        if (std::sscanf(line.c_str(), "%lf", &z_bottom) == 1) break;
    }
    return z_bottom;
}

// -eig_search (main.infraga.3d.cpp:773-953, main.infraga.sph.cpp:786-996).
int run_eig_search(int argc, char **argv)
{
    const bool sph = GA_CLI_GEOM == GA_GEOM_SPH;
    // defaults: 3D src (0, 0, 0), rcvr (-250, 0); sph src 30 N 0 E, rcvr 30 N 2.5 W
    double s0 = sph ? 30.0 : 0.0, s1 = 0.0, z_src = 0.0;
    double r0 = sph ? 30.0 : -250.0, r1 = sph ? -2.5 : 0.0;
    bool reverse_winds = false, custom_id = false;
    double freq = 0.1;
    const char *prof_format = "zTuvdp";
    const char *output_id = nullptr;
    ga_prop_params p;
    ga_prop_params_default(&p, GA_CLI_GEOM);
    ga_eig_params e;
    ga_eig_params_default(&e);
    double z_grnd = NAN;
    const char *topo_file = nullptr;
    bool set[4] = {false, false, false, false};   // src lat/x, src lon/y, rcvr lat/x, rcvr lon/y given
    bool bounds_set[4] = {false, false, false, false}, rng_set = false;

    for (int i = 3; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key2(a, "incl_min=", "min_incl=", v)) e.incl_min = std::atof(v);
        else if (key2(a, "incl_max=", "max_incl=", v)) e.incl_max = std::atof(v);
        else if (key2(a, "bnc_min=", "min_bnc=", v)) e.bnc_min = std::atoi(v);
        else if (key2(a, "bnc_max=", "max_bnc=", v)) e.bnc_max = std::atoi(v);
        else if (key(a, "bounces=", v)) e.bnc_min = e.bnc_max = std::atoi(v);
        else if (key2(a, "src_alt=", "alt_src=", v)) z_src = std::atof(v);
        else if (sph && key2(a, "src_lat=", "lat_src=", v)) { s0 = std::atof(v); set[0] = true; }
        else if (sph && key2(a, "src_lon=", "lon_src=", v)) { s1 = std::atof(v); set[1] = true; }
        else if (sph && key2(a, "rcvr_lat=", "lat_rcvr=", v)) { r0 = std::atof(v); set[2] = true; }
        else if (sph && key2(a, "rcvr_lon=", "lon_rcvr=", v)) { r1 = std::atof(v); set[3] = true; }
        else if (!sph && key2(a, "src_x=", "x_src=", v)) { s0 = std::atof(v); set[0] = true; }
        else if (!sph && key2(a, "src_y=", "y_src=", v)) { s1 = std::atof(v); set[1] = true; }
        else if (!sph && key2(a, "rcvr_x=", "x_rcvr=", v)) { r0 = std::atof(v); set[2] = true; }
        else if (!sph && key2(a, "rcvr_y=", "y_rcvr=", v)) { r1 = std::atof(v); set[3] = true; }
        else if (key(a, "verbose=", v)) { /* trace-by-trace log not produced */ }
        else if (key(a, "iterations=", v)) e.iterations = (int)std::atof(v);
        else if (key(a, "damping=", v)) e.damping = std::atof(v);
        else if (key(a, "tolerance=", v)) e.tolerance = std::atof(v);
        else if (key(a, "az_dev_lim=", v)) e.az_dev_lim = std::atof(v);
        else if (key2(a, "incl_step_min=", "min_incl_step=", v)) e.incl_step_min = std::atof(v);
        else if (key2(a, "incl_step_max=", "max_incl_step=", v)) e.incl_step_max = std::atof(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) { p.rng_max = std::atof(v); rng_set = true; }
        else if (sph && key2(a, "max_time=", "time_max=", v)) p.time_max = std::atof(v);
        else if (sph && key2(a, "min_lat=", "lat_min=", v)) { p.bounds[0] = std::atof(v); bounds_set[0] = true; }
        else if (sph && key2(a, "max_lat=", "lat_max=", v)) { p.bounds[1] = std::atof(v); bounds_set[1] = true; }
        else if (sph && key2(a, "min_lon=", "lon_min=", v)) { p.bounds[2] = std::atof(v); bounds_set[2] = true; }
        else if (sph && key2(a, "max_lon=", "lon_max=", v)) { p.bounds[3] = std::atof(v); bounds_set[3] = true; }
        else if (!sph && key2(a, "min_x=", "x_min=", v)) { p.bounds[0] = std::atof(v); bounds_set[0] = true; }
        else if (!sph && key2(a, "max_x=", "x_max=", v)) { p.bounds[1] = std::atof(v); bounds_set[1] = true; }
        else if (!sph && key2(a, "min_y=", "y_min=", v)) { p.bounds[2] = std::atof(v); bounds_set[2] = true; }
        else if (!sph && key2(a, "max_y=", "y_max=", v)) { p.bounds[3] = std::atof(v); bounds_set[3] = true; }
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else if (key(a, "topo_use_BLw=", v)) p.topo_use_blw = str2bool(v) ? 1 : 0;
        else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }

    if (!topo_file) p.topo_use_blw = 0;   // the vertical wind needs terrain (atmo::w)
    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int aid = 0;
    double z0 = 0.0, z_rcvr = 0.0;
    if (topo_file) {
        // the stratified profile over terrain (main.infraga.sph.cpp:819-827,
        // 891-899; main.infraga.3d.cpp:812-815): source at the terrain's
        // centre, receiver 2.5 deg east of it (3D: 250 km west), unless
        // given; the search runs over identical columns
        if ((sph ? ga_atmo_load_file_topo : ga_atmo_load_file_topo_xy)(ctx, argv[2], topo_file, prof_format,
                                                                       reverse_winds ? 1 : 0, &aid) != GA_OK)
            return die(ctx, "cannot load the atmosphere and terrain");
        double info[8];
        if (ga_rngdep_info(ctx, aid, info) != GA_OK) return die(ctx, "region");
        const double deg = sph ? Pi / 180.0 : 1.0;
        const double clat = (info[1] + info[0]) / 2.0 / deg, clon = (info[3] + info[2]) / 2.0 / deg;
        if (!set[0]) s0 = clat;
        if (!set[1]) s1 = clon;
        if (!set[2]) r0 = sph ? clat + 0.0 : clat - 250.0;
        if (!set[3]) r1 = sph ? clon + 2.5 : clon;
        double pt[6] = {s0 * deg, s1 * deg, 0.0, r0 * deg, r1 * deg, 0.0};
        double aout[80], tout[20];
        if (ga_rngdep_eval(ctx, aid, pt, 2, aout, tout) != GA_OK) return die(ctx, "terrain");
        z_src = std::max(z_src, tout[0]);
        z_rcvr = tout[10];
        ga_prop_params q;
        ga_prop_params_default(&q, sph ? GA_GEOM_SPH_RNGDEP : GA_GEOM_3D_RNGDEP);
        q.abs_coeff = p.abs_coeff;
        q.alt_max = p.alt_max;
        q.rng_max = (sph || rng_set) ? p.rng_max : 1000.0;
        q.time_max = p.time_max;
        q.ds_min = p.ds_min;
        q.ds_max = p.ds_max;
        q.s_max = p.s_max;
        q.topo_use_blw = p.topo_use_blw;
        for (int k = 0; k < 4; ++k) q.bounds[k] = bounds_set[k] ? p.bounds[k] : NAN;
        p = q;
    } else {
        if (ga_atmo_load_file(ctx, argv[2], prof_format, reverse_winds ? 1 : 0, &aid) != GA_OK)
            return die(ctx, "cannot load the atmosphere");
        z0 = std::isnan(z_grnd) ? profile_bottom(argv[2]) : z_grnd;
        p.z_grnd = z0;
        z_src = std::max(z_src, z0);
        z_rcvr = sph ? (p.r_earth + z0) - p.r_earth : z0;
    }
    p.freq = freq;
    p.src[0] = s0;
    p.src[1] = s1;
    p.src[2] = z_src;
    const double rcvr[2] = {r0, r1};
    std::vector<ga_eigenray> eig(64);
    std::vector<ga_path_pt> pts((size_t)1 << 16);
    int n_out = 0;
    int64_t n_pts = 0;
    double lat_ms = 0.0;
    ga_run_stats st;
    int rc = ga_eig_search_paths(ctx, aid, &p, &e, rcvr, 1, eig.data(), (int)eig.size(), &n_out, &lat_ms,
                                 pts.data(), (int64_t)pts.size(), &n_pts, &st);
    if (rc == GA_E_CAP) {
        eig.resize(std::max<size_t>(eig.size(), (size_t)n_out));
        pts.resize(std::max<size_t>(pts.size(), (size_t)n_pts));
        rc = ga_eig_search_paths(ctx, aid, &p, &e, rcvr, 1, eig.data(), (int)eig.size(), &n_out, &lat_ms,
                                 pts.data(), (int64_t)pts.size(), &n_pts, &st);
    }
    if (rc != GA_OK) return die(ctx, "eigenray search failed");

    std::string id = custom_id ? std::string(output_id) : strip_ext(argv[2]);
    // <id>.eigenray-N.dat, N counting the eigenrays in output order
    // (geoac.eigenray.3d.cpp:219-276, geoac.eigenray.sph.cpp:227-271): sph
    // rows at precision 8; 3D rows at the stream default, the absorption
    // column after the first reflection printed as the bare integral
    // (3d.cpp:263, without the -2 factor of the first leg).
    {
        int64_t w = 0;
        for (int i = 0; i < n_out; ++i) {
            std::ofstream ray(id + ".eigenray-" + std::to_string(i) + ".dat");
            ray << (sph ? "# lat [deg]" : "# x [km]") << '\t' << (sph ? "lon [deg]" : "y [km]") << '\t' << "z [km]"
                << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t' << "time [s]" << '\n';
            if (sph) ray << std::setprecision(8);
            for (; w < n_pts && pts[w].launch_idx == i; ++w) {
                const ga_path_pt &P = pts[w];
                const double ab = (!sph && P.leg > 0) ? P.absorption * -0.5 : P.absorption;
                ray << P.pos0 << '\t' << P.pos1 << '\t' << P.z << '\t' << P.tl << '\t' << ab << '\t' << P.time << '\n';
            }
        }
    }
    std::ofstream out(id + ".arrivals.dat");
    // header (main.infraga.3d.cpp:897-930, main.infraga.sph.cpp:931-960)
    out << "# " << GA_CLI_NAME << " -eig_search summary:" << '\n';
    out << "# " << '\t' << "profile: " << argv[2] << '\n';
    if (sph) {
        out << "# " << '\t' << "source location (lat, lon, alt): " << (s0 * (Pi / 180.0)) * (180.0 / Pi) << ", "
            << (s1 * (Pi / 180.0)) * (180.0 / Pi) << ", " << z_src << '\n';
        out << "# " << '\t' << "receiver location (lat, lon, alt): " << (r0 * (Pi / 180.0)) * (180.0 / Pi) << ", "
            << (r1 * (Pi / 180.0)) * (180.0 / Pi) << ", " << z_rcvr << '\n';
    } else {
        out << "# " << '\t' << "source location: " << s0 << ", " << s1 << ", " << z_src << '\n';
        out << "# " << '\t' << "receiver location: " << r0 << ", " << r1 << ", " << z_rcvr << '\n';
    }
    out << "# " << '\t' << "inclination range: " << (e.incl_min * (Pi / 180.0)) * (180.0 / Pi) << ", "
        << (e.incl_max * (Pi / 180.0)) * (180.0 / Pi) << '\n';
    out << "# " << '\t' << "bounces: " << e.bnc_min << ", " << e.bnc_max << '\n';
    if (topo_file) out << "#" << '\t' << "topo file:" << topo_file << '\n';
    else out << "#" << '\t' << "ground elevation: " << z0 << '\n';
    out << "# " << '\t' << "damping: " << e.damping << '\n';
    out << "# " << '\t' << "frequency: " << freq << '\n';
    out << "# " << '\t' << "S&B atten coeff: " << p.abs_coeff << '\n';
    if (!sph) out << '\n';
    out << "# incl [deg]" << '\t' << "az [deg]" << '\t' << "n_b";
    if (sph) out << '\t' << "lat_0 [deg]" << '\t' << "lon_0 [deg]";
    else out << '\t' << "x_0 [km]" << '\t' << "y_0 [km]";
    out << '\t' << "time [s]" << '\t' << "cel [km/s]" << '\t' << "turning ht [km]";
    if (sph) out << '\t' << "inclination [deg]" << '\t' << "back azimuth [deg]";
    else out << '\t' << "arrival incl [deg]" << '\t' << "back az [deg]";
    out << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\n';
    // rows (geoac.eigenray.3d.cpp:302-315): setprecision(8) stays in effect
    for (int i = 0; i < n_out; ++i) {
        const ga_eigenray &A = eig[i];
        out << std::setprecision(8) << A.incl;
        out << '\t' << std::setprecision(8) << A.az;
        out << '\t' << A.n_bounce;
        out << '\t' << A.pos0 << '\t' << A.pos1 << '\t' << A.time << '\t' << A.cel << '\t' << A.turn_ht;
        out << '\t' << A.incl_arr << '\t' << A.back_az << '\t' << A.tl << '\t' << A.absorption << '\n';
        std::cout << '\t' << "Eigenray identified:" << '\t' << "theta, phi = " << std::setprecision(8) << A.incl << ", "
                  << A.az << " degrees." << '\n';
    }
    out.close();
    std::cout << GA_CLI_NAME << ": " << n_out << " eigenray(s), " << st.n_rays << " ray traces, " << st.ray_steps
              << " ray-steps, " << st.kernel_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// infraga-{sph,3d}-rngdep -eig_search / -eig_direct prefix nodes nodes
// [key=value ...]: run_eig_search and run_eig_direct of
// main.infraga.sph.rngdep.cpp:777-1107 (the spherical search of
// geoac.eigenray.sph.cpp) and main.infraga.3d.rngdep.cpp:755-1050 (the 3D
// search of geoac.eigenray.3d.cpp, node coordinates and receivers in km)
// over the range-dependent medium and terrain.  The search writes
// <id>.arrivals.dat and <id>.eigenray-N.dat, -eig_direct the eigenray file
// of an accepted iteration.
int run_eig_rd(int argc, char **argv, bool direct)
{
    constexpr bool cart = GA_CLI_GEOM == GA_GEOM_3D_RNGDEP;
    if (argc < 5) {
        std::cerr << GA_CLI_NAME << ": usage: " << argv[1]
                  << (cart ? " <profile prefix> <x nodes> <y nodes> [parameters]\n"
                           : " <profile prefix> <lat nodes> <lon nodes> [parameters]\n");
        return 2;
    }
    // sph: lat / lon [deg]; 3D: x / y [km], the defaults taken from the region
    double s0 = 30.0, s1 = 0.0, z_src = 0.0, r0 = 30.0, r1 = -2.5;
    bool s0_set = false, s1_set = false, r0_set = false, r1_set = false;
    // -eig_direct: inclination estimate (sph 10 deg; the 3D binary starts
    // from 0.5 taken as radians, main.infraga.3d.rngdep.cpp:937), azimuth
    // estimate (NaN: toward the receiver)
    double incl_est = cart ? 0.5 * (180.0 / Pi) : 10.0, az_est = NAN;
    int bounces = 0;
    bool reverse_winds = false, custom_id = false;
    double freq = 0.1;
    const char *prof_format = "zTuvdp";
    const char *topo_file = nullptr;
    const char *output_id = nullptr;
    ga_prop_params p;
    ga_prop_params_default(&p, cart ? GA_GEOM_3D_RNGDEP : GA_GEOM_SPH_RNGDEP);
    ga_eig_params e;
    ga_eig_params_default(&e);
    double z_grnd = NAN;
    for (int i = 5; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (!direct && key2(a, "incl_min=", "min_incl=", v)) e.incl_min = std::atof(v);
        else if (!direct && key2(a, "incl_max=", "max_incl=", v)) e.incl_max = std::atof(v);
        else if (!direct && key2(a, "bnc_min=", "min_bnc=", v)) e.bnc_min = std::atoi(v);
        else if (!direct && key2(a, "bnc_max=", "max_bnc=", v)) e.bnc_max = std::atoi(v);
        else if (!direct && key(a, "bounces=", v)) e.bnc_min = e.bnc_max = std::atoi(v);
        else if (direct && key2(a, "incl_est=", "est_incl=", v)) incl_est = std::atof(v);
        else if (direct && key2(a, "az_est=", "est_az=", v)) az_est = std::atof(v);
        else if (direct && key(a, "bounces=", v)) bounces = std::atoi(v);
        else if (key2(a, "src_alt=", "alt_src=", v)) z_src = std::atof(v);
        else if (!cart && key2(a, "src_lat=", "lat_src=", v)) { s0 = std::atof(v); s0_set = true; }
        else if (!cart && key2(a, "src_lon=", "lon_src=", v)) { s1 = std::atof(v); s1_set = true; }
        else if (!cart && key2(a, "rcvr_lat=", "lat_rcvr=", v)) { r0 = std::atof(v); r0_set = true; }
        else if (!cart && key2(a, "rcvr_lon=", "lon_rcvr=", v)) { r1 = std::atof(v); r1_set = true; }
        else if (cart && key2(a, "src_x=", "x_src=", v)) { s0 = std::atof(v); s0_set = true; }
        else if (cart && key2(a, "src_y=", "y_src_y=", v)) { s1 = std::atof(v); s1_set = true; }
        else if (cart && key2(a, "rcvr_x=", "x_rcvr=", v)) { r0 = std::atof(v); r0_set = true; }
        else if (cart && key2(a, "rcvr_y=", "y_rcvr=", v)) { r1 = std::atof(v); r1_set = true; }
        else if (key(a, "verbose=", v)) { /* trace-by-trace log not produced */ }
        else if (key(a, "iterations=", v)) e.iterations = (int)std::atof(v);
        else if (key(a, "damping=", v)) e.damping = std::atof(v);
        else if (key(a, "tolerance=", v)) e.tolerance = std::atof(v);
        else if (!direct && key(a, "az_dev_lim=", v)) e.az_dev_lim = std::atof(v);
        else if (!direct && key2(a, "incl_step_min=", "min_incl_step=", v)) e.incl_step_min = std::atof(v);
        else if (!direct && key2(a, "incl_step_max=", "max_incl_step=", v)) e.incl_step_max = std::atof(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) p.rng_max = std::atof(v);
        else if (!cart && key2(a, "max_time=", "time_max=", v)) p.time_max = std::atof(v);
        else if (key2(a, cart ? "min_x=" : "min_lat=", cart ? "x_min=" : "lat_min=", v)) p.bounds[0] = std::atof(v);
        else if (key2(a, cart ? "max_x=" : "max_lat=", cart ? "x_max=" : "lat_max=", v)) p.bounds[1] = std::atof(v);
        else if (key2(a, cart ? "min_y=" : "min_lon=", cart ? "y_min=" : "lon_min=", v)) p.bounds[2] = std::atof(v);
        else if (key2(a, cart ? "max_y=" : "max_lon=", cart ? "y_max=" : "lon_max=", v)) p.bounds[3] = std::atof(v);
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else if (key(a, "topo_use_BLw=", v)) p.topo_use_blw = str2bool(v) ? 1 : 0;
        else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }
    if (direct && bounces < 0) {
        std::cerr << GA_CLI_NAME << ": bounces must be >= 0\n";
        return 2;
    }
    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int rid = 0;
    if ((cart ? ga_rngdep_load_files_xy : ga_rngdep_load_files)(ctx, argv[2], argv[3], argv[4], topo_file, prof_format,
                                                               reverse_winds ? 1 : 0, &rid) != GA_OK)
        return die(ctx, "cannot load the range-dependent atmosphere");
    double info[8];
    if (ga_rngdep_info(ctx, rid, info) != GA_OK) return die(ctx, "region");
    const bool is_topo = topo_file != nullptr;
    const double deg = cart ? 1.0 : Pi / 180.0;   // region in radians (sph) or km (3D)
    if (cart) {
        // source at the region centre, receiver 250 km east of it (3d.rngdep.cpp:797-798)
        const double cx = (info[1] + info[0]) / 2.0, cy = (info[3] + info[2]) / 2.0;
        if (!s0_set) s0 = cx;
        if (!s1_set) s1 = cy;
        if (!r0_set) r0 = cx + 250.0;
        if (!r1_set) r1 = cy;
    } else if (is_topo) {
        // source at the region centre, receiver 5 deg east of it (sph.rngdep.cpp:812-815)
        const double clat = (info[1] + info[0]) / 2.0 * (180.0 / Pi), clon = (info[3] + info[2]) / 2.0 * (180.0 / Pi);
        if (!s0_set) s0 = clat;
        if (!s1_set) s1 = clon;
        if (!r0_set) r0 = clat + 0.0;
        if (!r1_set) r1 = clon + 5.0;
    }
    const double z0 = (!is_topo && !std::isnan(z_grnd)) ? z_grnd : info[5];
    if (!is_topo) p.z_grnd = z0;
    double pts[6] = {s0 * deg, s1 * deg, 0.0, r0 * deg, r1 * deg, 0.0};
    double aout[80], tout[20];
    if (ga_rngdep_eval(ctx, rid, pts, 2, aout, tout) != GA_OK) return die(ctx, "terrain");
    // without terrain the spherical ground is z0, the Cartesian topo::z is 0
    const double zflat = cart ? 0.0 : z0;
    const double zg_src = is_topo ? tout[0] : zflat, zg_rcvr = is_topo ? tout[10] : zflat;
    z_src = std::max(z_src, zg_src);                                     // sph :880, 3D :845
    p.freq = freq;
    p.src[0] = s0;
    p.src[1] = s1;
    p.src[2] = z_src;
    const double rcvr[2] = {r0, r1};
    std::vector<ga_eigenray> eig(64);
    std::vector<ga_path_pt> pts_out((size_t)1 << 16);
    int n_out = 0;
    int64_t n_pts = 0;
    double lat_ms = 0.0;
    ga_run_stats st;
    auto call = [&]() {
        if (direct)
            return ga_eig_direct(ctx, rid, &p, &e, rcvr, incl_est, az_est, bounces, eig.data(), &n_out, pts_out.data(),
                                 (int64_t)pts_out.size(), &n_pts, &st);
        return ga_eig_search_paths(ctx, rid, &p, &e, rcvr, 1, eig.data(), (int)eig.size(), &n_out, &lat_ms,
                                   pts_out.data(), (int64_t)pts_out.size(), &n_pts, &st);
    };
    int rc = call();
    if (rc == GA_E_CAP) {
        eig.resize(std::max<size_t>(eig.size(), (size_t)n_out));
        pts_out.resize(std::max<size_t>(pts_out.size(), (size_t)n_pts));
        rc = call();
    }
    if (rc != GA_OK) return die(ctx, direct ? "eigenray iteration failed" : "eigenray search failed");
    std::string id = custom_id ? std::string(output_id) : std::string(argv[2]);
    {
        // <id>.eigenray-N.dat: sph rows at precision 8; 3D rows at the
        // stream default, the absorption column after the first reflection
        // the bare integral (geoac.eigenray.3d.cpp:263)
        int64_t w = 0;
        for (int i = 0; i < n_out; ++i) {
            std::ofstream ray(id + ".eigenray-" + std::to_string(i) + ".dat");
            ray << (cart ? "# x [km]" : "# lat [deg]") << '\t' << (cart ? "y [km]" : "lon [deg]") << '\t' << "z [km]"
                << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t' << "time [s]" << '\n';
            if (!cart) ray << std::setprecision(8);
            for (; w < n_pts && pts_out[w].launch_idx == i; ++w) {
                const ga_path_pt &P = pts_out[w];
                const double ab = (cart && P.leg > 0) ? P.absorption * -0.5 : P.absorption;
                ray << P.pos0 << '\t' << P.pos1 << '\t' << P.z << '\t' << P.tl << '\t' << ab << '\t' << P.time << '\n';
            }
        }
    }
    if (direct) {
        if (n_out)
            std::cout << '\t' << "Eigenray identified:" << '\t' << "theta, phi = " << std::setprecision(8) << eig[0].incl
                      << ", " << eig[0].az << " degrees." << '\n';
        else
            std::cout << GA_CLI_NAME << ": no eigenray identified\n";
        std::cout << GA_CLI_NAME << ": " << st.n_rays << " ray traces, " << st.ray_steps << " ray-steps, "
                  << st.kernel_ms << " ms on the GPU\n";
        ga_ctx_destroy(ctx);
        return 0;
    }
    std::ofstream out(id + ".arrivals.dat");
    out << "# " << GA_CLI_NAME << " -eig_search summary:" << '\n';
    out << "# " << '\t' << "profile prefix: " << argv[2] << '\n';
    if (cart) {
        // main.infraga.3d.rngdep.cpp:850-866
        out << "# " << '\t' << "profile grid files: " << argv[3] << ", " << argv[4] << '\n';
        out << "# " << '\t' << "source location: " << s0 << ", " << s1 << ", " << z_src << '\n';
        out << "# " << '\t' << "receiver location: " << r0 << ", " << r1 << ", " << zg_rcvr << '\n';
    } else {
        out << "# " << '\t' << "source location (lat, lon, alt): " << (s0 * (Pi / 180.0)) * (180.0 / Pi) << ", "
            << (s1 * (Pi / 180.0)) * (180.0 / Pi) << ", " << z_src << '\n';
        out << "# " << '\t' << "receiver location (lat, lon, alt): " << (r0 * (Pi / 180.0)) * (180.0 / Pi) << ", "
            << (r1 * (Pi / 180.0)) * (180.0 / Pi) << ", " << zg_rcvr << '\n';
    }
    out << "# " << '\t' << "inclination range: " << (e.incl_min * (Pi / 180.0)) * (180.0 / Pi) << ", "
        << (e.incl_max * (Pi / 180.0)) * (180.0 / Pi) << '\n';
    out << "# " << '\t' << "bounces: " << e.bnc_min << ", " << e.bnc_max << '\n';
    if (!is_topo) out << "#" << '\t' << "ground elevation: " << z0 << '\n';
    else out << "#" << '\t' << "topo file:" << topo_file << '\n';
    out << "# " << '\t' << "damping: " << e.damping << '\n';
    out << "# " << '\t' << "frequency: " << freq << '\n';
    out << "# " << '\t' << "S&B atten coeff: " << p.abs_coeff << '\n';
    if (cart) out << '\n';
    out << "# incl [deg]" << '\t' << "az [deg]" << '\t' << "n_b";
    if (cart) out << '\t' << "x_0 [km]" << '\t' << "y_0 [km]";
    else out << '\t' << "lat_0 [deg]" << '\t' << "lon_0 [deg]";
    out << '\t' << "time [s]" << '\t' << "cel [km/s]" << '\t' << "turning ht [km]";
    if (cart) out << '\t' << "arrival incl [deg]" << '\t' << "back az [deg]";
    else out << '\t' << "inclination [deg]" << '\t' << "back azimuth [deg]";
    out << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\n';
    for (int i = 0; i < n_out; ++i) {
        const ga_eigenray &A = eig[i];
        out << std::setprecision(8) << A.incl << '\t' << std::setprecision(8) << A.az << '\t' << A.n_bounce;
        out << '\t' << A.pos0 << '\t' << A.pos1 << '\t' << A.time << '\t' << A.cel << '\t' << A.turn_ht;
        out << '\t' << A.incl_arr << '\t' << A.back_az << '\t' << A.tl << '\t' << A.absorption << '\n';
        std::cout << '\t' << "Eigenray identified:" << '\t' << "theta, phi = " << std::setprecision(8) << A.incl << ", "
                  << A.az << " degrees." << '\n';
    }
    out.close();
    std::cout << GA_CLI_NAME << ": " << n_out << " eigenray(s), " << st.n_rays << " ray traces, " << st.ray_steps
              << " ray-steps, " << st.kernel_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// infraga-2d -prop (run_prop, main.infraga.2d.cpp:127-400): one azimuth,
// the effective sound speed in its vertical plane.  Always writes
// <id>.raypaths.dat; caustics-N.dat with write_caustics; all files at the
// stream's default precision.
int run_prop_2d(int argc, char **argv)
{
    double theta_min = 0.5, theta_max = 45.0, theta_step = 0.5, azimuth = -90.0;
    double z_src = 0.0, freq = 0.1;
    int bounces = 10;
    bool write_caustics = false, reverse_winds = false, custom_id = false;
    const char *prof_format = "zTuvdp";
    const char *topo_file = nullptr;
    const char *output_id = nullptr;
    ga_prop_params p;
    ga_prop_params_default(&p, GA_GEOM_2D);
    double z_grnd = NAN;

    for (int i = 3; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key2(a, "incl_min=", "min_incl=", v)) theta_min = std::atof(v);
        else if (key2(a, "incl_max=", "max_incl=", v)) theta_max = std::atof(v);
        else if (key(a, "incl_step=", v)) theta_step = std::atof(v);
        else if (key(a, "inclination=", v)) { theta_min = theta_max = std::atof(v); theta_step = 1.0; }
        else if (key(a, "azimuth=", v)) azimuth = std::atof(v);
        else if (key(a, "bounces=", v)) bounces = std::atoi(v);
        else if (key(a, "src_alt=", v)) z_src = std::atof(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key(a, "write_caustics=", v)) write_caustics = str2bool(v);
        else if (key(a, "calc_amp=", v)) p.calc_amp = str2bool(v) ? 1 : 0;
        else if (key2(a, "max_rng=", "rng_max=", v)) p.rng_max = std::atof(v);
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key2(a, "refl_alt=", "alt_refl=", v)) p.refl_alt = std::atof(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else if (key(a, "print_resid=", v)) { /* diagnostics only */ }
        else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }
    if (write_caustics) p.calc_amp = 1;                     // :215

    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int aid = 0;
    if (ga_atmo_load_file(ctx, argv[2], prof_format, reverse_winds ? 1 : 0, &aid) != GA_OK)
        return die(ctx, "cannot load the atmosphere");
    if (topo_file) {
        std::ifstream tf(topo_file);
        if (!tf) return die(ctx, "cannot read the terrain file");
        std::vector<double> r, zt;
        std::string line;
        while (std::getline(tf, line)) {
            if (line.find('#') == 0) continue;
            double a = 0.0, b = 0.0;
            if (std::sscanf(line.c_str(), "%lf %lf", &a, &b) == 2) { r.push_back(a); zt.push_back(b); }
        }
        if (ga_topo_load_2d(ctx, aid, r.data(), zt.data(), (int)r.size()) != GA_OK) return die(ctx, "terrain");
    }
    const double zb = profile_bottom(argv[2]);
    const double z0 = (!topo_file && !std::isnan(z_grnd)) ? z_grnd : zb;
    p.z_grnd = topo_file ? NAN : z0;
    double t0[3];
    if (ga_topo_eval_2d(ctx, aid, 0.0, t0) != GA_OK) return die(ctx, "terrain");
    const double zg_src = topo_file ? t0[0] : z0;
    z_src = std::max(z_src, zg_src);                        // :222
    p.src[2] = z_src;
    p.bounces = bounces;
    p.freq = freq;
    // profile top: the altitude limit and, by default, the reflection altitude above it
    double z_top = zb;
    {
        std::ifstream probe(argv[2]);
        std::string line;
        while (std::getline(probe, line)) {
            if (line.find('#') == 0) continue;
            double zz;
            if (std::sscanf(line.c_str(), "%lf", &zz) == 1) z_top = zz;
        }
    }
    const double alt_max = std::isnan(p.alt_max) ? z_top : p.alt_max;
    const double z_reflect = std::isnan(p.refl_alt) ? z_top + 1.0 : p.refl_alt;

    // launch list (:310-318)
    double theta_grnd;
    if (z_src - t0[0] < 0.01) theta_grnd = topo_file ? std::atan(t0[1]) * (180.0 / Pi) + 0.1 : 0.1;
    else theta_grnd = -89.9;
    std::vector<double> launches;
    for (double theta = std::max(theta_min, theta_grnd); theta <= std::max(theta_max, theta_grnd); theta += theta_step) {
        launches.push_back(theta);
        launches.push_back(azimuth);
    }
    const int n = (int)(launches.size() / 2);
    std::string id = custom_id ? std::string(output_id) : strip_ext(argv[2]);

    std::ofstream results(id + ".arrivals.dat");
    results << "# infraga-2d -prop summary:" << '\n';
    results << "#" << '\t' << "profile: " << argv[2] << '\n';
    results << "#" << '\t' << "inclination: " << theta_min << ", " << theta_max << ", " << theta_step << '\n';
    results << "#" << '\t' << "azimuth: " << azimuth << '\n';
    results << "#" << '\t' << "bounces: " << bounces << '\n';
    results << "#" << '\t' << "source location: " << 0.0 << ", " << z_src << '\n';
    if (!topo_file) results << "#" << '\t' << "ground elevation: " << z0 << '\n';
    else results << "#" << '\t' << "topo file:" << topo_file << '\n';
    if (z_reflect < alt_max) results << '#' << '\t' << "reflection altitude: " << z_reflect << '\n';
    results << "#" << '\t' << "frequency: " << freq << '\n';
    results << "#" << '\t' << "S&B atten coeff: " << p.abs_coeff << '\n' << '\n';
    results << "# incl [deg]" << '\t' << "az [deg]" << '\t' << "n_b" << '\t' << "r_0 [km]" << '\t' << "time [s]" << '\t'
            << "celerity [km/s]" << '\t' << "turning ht [km]" << '\t' << "arrival incl. [deg]" << '\t'
            << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t' << "perp. dist. [km]" << '\n';
    std::ofstream raypath(id + ".raypaths.dat");
    raypath << "# r [km]" << '\t' << "z [km]" << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t'
            << "time [s]" << '\n';
    std::vector<std::ofstream> caus;
    if (write_caustics) {
        for (int b = 0; b <= bounces; ++b) {
            caus.emplace_back(id + ".caustics-" + std::to_string(b) + ".dat");
            caus.back() << "# r [km]" << '\t' << "z [km]" << '\t' << "time [s]" << '\n';
        }
    }
    const int slots = bounces + 1;
    const int flags = GA_PATH_RAYS | (write_caustics ? GA_PATH_CAUSTICS : 0);
    std::vector<ga_arrival> arr;
    std::vector<ga_path_pt> pts;
    int64_t cap = (int64_t)1 << 21;
    ga_run_stats tot{};
    int n_arr = 0;
    for (int r0 = 0; r0 < n; r0 += 4096) {
        const int nb = std::min(4096, n - r0);
        arr.assign((size_t)nb * slots, ga_arrival{});
        int nb_out = 0;
        int64_t n_pts = 0;
        ga_run_stats bs;
        int rc;
        for (;;) {
            pts.resize((size_t)cap);
            rc = ga_fan_run_paths(ctx, aid, &p, launches.data() + 2 * (size_t)r0, nb, flags, arr.data(), nb * slots,
                                  &nb_out, pts.data(), cap, &n_pts, &bs);
            if (rc == GA_E_CAP && n_pts > cap) { cap = n_pts + n_pts / 4; continue; }
            break;
        }
        if (rc != GA_OK) return die(ctx, "fan failed");
        tot.ray_steps += bs.ray_steps;
        tot.kernel_ms += bs.kernel_ms;
        n_arr += nb_out;
        int w = 0;
        int64_t q = 0;
        for (int r = 0; r < nb; ++r) {
            for (; q < n_pts && pts[q].launch_idx == r; ++q) {
                const ga_path_pt &P = pts[q];
                if (P.kind == GA_PT_RAY) {
                    raypath << P.pos0 << '\t' << P.z << '\t' << P.tl << '\t' << P.absorption << '\t' << P.time << '\n';
                } else if (write_caustics && P.leg >= 0 && P.leg < (int)caus.size()) {
                    caus[P.leg] << P.pos0 << '\t' << P.z << '\t' << P.time << '\n' << '\n';
                }
            }
            raypath << '\n';
            for (; w < nb_out && arr[w].launch_idx == r; ++w) {
                const ga_arrival &A = arr[w];
                results << A.incl << '\t' << A.az << '\t' << A.n_bounce << '\t' << A.pos0 << '\t' << A.time << '\t'
                        << A.cel << '\t' << A.turn_ht << '\t' << A.incl_arr << '\t' << A.tl << '\t' << A.absorption
                        << '\t' << A.pos1 << '\n';
            }
        }
    }
    std::cout << GA_CLI_NAME << ": " << n << " rays, " << n_arr << " arrivals, " << tot.ray_steps << " ray-steps, "
              << tot.kernel_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// -eig_direct (run_eig_direct, main.infraga.3d.cpp:955-1086,
// main.infraga.sph.cpp:998-1139): find_eigenray from a launch estimate;
// writes <id>.eigenray-0.dat when it converges (the reference opens no
// arrivals file for this option).
int run_eig_direct(int argc, char **argv)
{
    const bool sph = GA_CLI_GEOM == GA_GEOM_SPH;
    double s0 = sph ? 30.0 : 0.0, s1 = 0.0, z_src = 0.0;
    double r0 = sph ? 30.0 : -250.0, r1 = sph ? -2.5 : 0.0;
    double incl_est = 10.0, az_est = NAN;
    int bounces = 0;
    bool reverse_winds = false, custom_id = false;
    double freq = 0.1;
    const char *prof_format = "zTuvdp";
    const char *output_id = nullptr;
    ga_prop_params p;
    ga_prop_params_default(&p, GA_CLI_GEOM);
    ga_eig_params e;
    ga_eig_params_default(&e);
    double z_grnd = NAN;
    const char *topo_file = nullptr;
    bool set[4] = {false, false, false, false};   // src lat/x, src lon/y, rcvr lat/x, rcvr lon/y given
    bool bounds_set[4] = {false, false, false, false}, rng_set = false;

    for (int i = 3; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key2(a, "incl_est=", "est_incl=", v)) incl_est = std::atof(v);
        else if (key2(a, "az_est=", "est_az=", v)) az_est = std::atof(v);
        else if (key(a, "bounces=", v)) bounces = std::atoi(v);
        else if (key2(a, "src_alt=", "alt_src=", v)) z_src = std::atof(v);
        else if (sph && key2(a, "src_lat=", "lat_src=", v)) { s0 = std::atof(v); set[0] = true; }
        else if (sph && key2(a, "src_lon=", "lon_src=", v)) { s1 = std::atof(v); set[1] = true; }
        else if (sph && key2(a, "rcvr_lat=", "lat_rcvr=", v)) { r0 = std::atof(v); set[2] = true; }
        else if (sph && key2(a, "rcvr_lon=", "lon_rcvr=", v)) { r1 = std::atof(v); set[3] = true; }
        else if (!sph && key2(a, "src_x=", "x_src=", v)) { s0 = std::atof(v); set[0] = true; }
        else if (!sph && key2(a, "src_y=", "y_src_y=", v)) { s1 = std::atof(v); set[1] = true; }
        else if (!sph && key2(a, "rcvr_x=", "x_rcvr=", v)) { r0 = std::atof(v); set[2] = true; }
        else if (!sph && key2(a, "rcvr_y=", "y_rcvr=", v)) { r1 = std::atof(v); set[3] = true; }
        else if (key(a, "verbose=", v)) { /* trace-by-trace log not produced */ }
        else if (key(a, "iterations=", v)) e.iterations = (int)std::atof(v);
        else if (key(a, "damping=", v)) e.damping = std::atof(v);
        else if (key(a, "tolerance=", v)) e.tolerance = std::atof(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) { p.rng_max = std::atof(v); rng_set = true; }
        else if (sph && key2(a, "max_time=", "time_max=", v)) p.time_max = std::atof(v);
        else if (key2(a, sph ? "min_lat=" : "min_x=", sph ? "lat_min=" : "x_min=", v)) { p.bounds[0] = std::atof(v); bounds_set[0] = true; }
        else if (key2(a, sph ? "max_lat=" : "max_x=", sph ? "lat_max=" : "x_max=", v)) { p.bounds[1] = std::atof(v); bounds_set[1] = true; }
        else if (key2(a, sph ? "min_lon=" : "min_y=", sph ? "lon_min=" : "y_min=", v)) { p.bounds[2] = std::atof(v); bounds_set[2] = true; }
        else if (key2(a, sph ? "max_lon=" : "max_y=", sph ? "lon_max=" : "y_max=", v)) { p.bounds[3] = std::atof(v); bounds_set[3] = true; }
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else if (key(a, "topo_use_BLw=", v)) p.topo_use_blw = str2bool(v) ? 1 : 0;
        else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }
    if (bounces < 0) {
        std::cerr << GA_CLI_NAME << ": bounces must be >= 0\n";
        return 2;
    }

    if (!topo_file) p.topo_use_blw = 0;   // the vertical wind needs terrain (atmo::w)
    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int aid = 0;
    if (topo_file) {
        // the stratified profile over terrain (main.infraga.sph.cpp:1030-1040,
        // main.infraga.3d.cpp:992-995): source at the terrain's centre,
        // receiver 2.5 deg east (3D: 250 km west) of it unless given
        if ((sph ? ga_atmo_load_file_topo : ga_atmo_load_file_topo_xy)(ctx, argv[2], topo_file, prof_format,
                                                                       reverse_winds ? 1 : 0, &aid) != GA_OK)
            return die(ctx, "cannot load the atmosphere and terrain");
        double info[8];
        if (ga_rngdep_info(ctx, aid, info) != GA_OK) return die(ctx, "region");
        const double deg = sph ? Pi / 180.0 : 1.0;
        const double clat = (info[1] + info[0]) / 2.0 / deg, clon = (info[3] + info[2]) / 2.0 / deg;
        if (!set[0]) s0 = clat;
        if (!set[1]) s1 = clon;
        if (!set[2]) r0 = sph ? clat + 0.0 : clat - 250.0;
        if (!set[3]) r1 = sph ? clon + 2.5 : clon;
        double pt[3] = {s0 * deg, s1 * deg, 0.0}, aout[40], tout[10];
        if (ga_rngdep_eval(ctx, aid, pt, 1, aout, tout) != GA_OK) return die(ctx, "terrain");
        z_src = std::max(z_src, tout[0]);
        ga_prop_params q;
        ga_prop_params_default(&q, sph ? GA_GEOM_SPH_RNGDEP : GA_GEOM_3D_RNGDEP);
        q.abs_coeff = p.abs_coeff;
        q.alt_max = p.alt_max;
        q.rng_max = (sph || rng_set) ? p.rng_max : 1000.0;
        q.time_max = p.time_max;
        q.ds_min = p.ds_min;
        q.ds_max = p.ds_max;
        q.s_max = p.s_max;
        q.topo_use_blw = p.topo_use_blw;
        for (int k = 0; k < 4; ++k) q.bounds[k] = bounds_set[k] ? p.bounds[k] : NAN;
        p = q;
    } else {
        if (ga_atmo_load_file(ctx, argv[2], prof_format, reverse_winds ? 1 : 0, &aid) != GA_OK)
            return die(ctx, "cannot load the atmosphere");
        const double z0 = std::isnan(z_grnd) ? profile_bottom(argv[2]) : z_grnd;
        p.z_grnd = z0;
        z_src = std::max(z_src, z0);
    }
    p.freq = freq;
    p.src[0] = s0;
    p.src[1] = s1;
    p.src[2] = z_src;
    const double rcvr[2] = {r0, r1};
    ga_eigenray eig;
    int found = 0;
    std::vector<ga_path_pt> pts((size_t)1 << 16);
    int64_t n_pts = 0;
    ga_run_stats st;
    int rc = ga_eig_direct(ctx, aid, &p, &e, rcvr, incl_est, az_est, bounces, &eig, &found, pts.data(),
                           (int64_t)pts.size(), &n_pts, &st);
    if (rc == GA_E_CAP) {
        pts.resize((size_t)n_pts);
        rc = ga_eig_direct(ctx, aid, &p, &e, rcvr, incl_est, az_est, bounces, &eig, &found, pts.data(),
                           (int64_t)pts.size(), &n_pts, &st);
    }
    if (rc != GA_OK) return die(ctx, "eigenray iteration failed");
    std::string id = custom_id ? std::string(output_id) : strip_ext(argv[2]);
    if (found) {
        std::ofstream ray(id + ".eigenray-0.dat");
        ray << (sph ? "# lat [deg]" : "# x [km]") << '\t' << (sph ? "lon [deg]" : "y [km]") << '\t' << "z [km]"
            << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t' << "time [s]" << '\n';
        if (sph) ray << std::setprecision(8);
        for (int64_t w = 0; w < n_pts; ++w) {
            const ga_path_pt &P = pts[w];
            const double ab = (!sph && P.leg > 0) ? P.absorption * -0.5 : P.absorption;
            ray << P.pos0 << '\t' << P.pos1 << '\t' << P.z << '\t' << P.tl << '\t' << ab << '\t' << P.time << '\n';
        }
        std::cout << '\t' << "Eigenray identified:" << '\t' << "theta, phi = " << std::setprecision(8) << eig.incl << ", "
                  << eig.az << " degrees." << '\n';
    } else {
        std::cout << GA_CLI_NAME << ": no eigenray identified\n";
    }
    std::cout << GA_CLI_NAME << ": " << st.n_rays << " ray traces, " << st.ray_steps << " ray-steps, " << st.kernel_ms
              << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// ---- -wnl_wvfrm ------------------------------------------------------------
// Built-in source waveforms (src/util/waveforms.cpp:37-145): a shaped
// impulse p0 x^a (1 - x / (1 + a)) e^-x normalised at its peak x_pk =
// (1 + a) - sqrt(1 + a), x = t / t0, zero before t = 0; the U-wave is its
// central difference (eps = 0.01 s) scaled by 1.126 times the slope at 0;
// the N-wave the ramp -p0 (t - t0/2) / t0 on |t - t0/2| <= t0.  Record
// lengths: t0 len max(x_pk / N_rise, 2 pi' sqrt(1 + a) / N_period) with the
// reference's pi' = 3.14159 (min for an unrecognised option), the N-wave
// t0 / N_period len.  alpha is raised to 0.01 (2 for the U-wave).
double wv_impulse(double t, double t0, double a)
{
    if (t < 0.0) return 0.0;
    const double x = t / t0, x_pk = (a + 1.0) - std::sqrt(a + 1.0);
    const double f = std::pow(x, a) * (1.0 - x / (1.0 + a)) * std::exp(-x);
    return f / (std::pow(x_pk, a) * (1.0 - x_pk / (1.0 + a)) * std::exp(-x_pk));
}

void wv_build(const char *opt, double p0, double t0, double &alpha, int len, int n_rise, int n_period,
              std::vector<double> &t, std::vector<double> &p)
{
    t.assign((size_t)len, 0.0);
    p.assign((size_t)len, 0.0);
    const bool imp = std::strncmp(opt, "impulse", 7) == 0;
    const bool uw = std::strncmp(opt, "Uwave", 5) == 0 || std::strncmp(opt, "u-wave", 6) == 0 ||
                    std::strncmp(opt, "U-wave", 6) == 0;
    const bool nw = std::strncmp(opt, "Nwave", 5) == 0 || std::strncmp(opt, "n-wave", 6) == 0 ||
                    std::strncmp(opt, "N-wave", 6) == 0;
    double duration;
    if (nw) {
        duration = (t0 / n_period) * len;
    } else {
        alpha = std::max(alpha, uw ? 2.0 : 0.01);
        const double a = alpha;
        const double r = ((1.0 + a) - std::sqrt(1.0 + a)) / n_rise, q = 2.0 * 3.14159 * std::sqrt(1 + a) / n_period;
        duration = t0 * len * ((imp || uw) ? std::max(r, q) : std::min(r, q));
    }
    const double eps = 1.0e-2;
    for (int n = 0; n < len; ++n) {
        const double tn = -duration / 2.0 + duration * n / len;
        t[n] = tn;
        if (nw) p[n] = std::fabs(tn - t0 / 2.0) <= t0 ? -p0 * ((tn - t0 / 2.0) / t0) : -p0 * 0.0;
        else if (uw) {
            p[n] = p0 * (wv_impulse(tn + eps, t0, alpha) - wv_impulse(tn - eps, t0, alpha)) / (2.0 * eps);
            p[n] /= 1.126 * (wv_impulse(eps, t0, alpha) - wv_impulse(-eps, t0, alpha)) / (2.0 * eps);
        } else {
            p[n] = p0 * wv_impulse(tn, t0, alpha);
        }
    }
}

// wvfrm_file= (waveforms.cpp:64-90, fileIO.cpp:18-35): the record length is
// the count of lines not starting with '#' less one (the reader's final
// empty read), then the first that many such lines as "t p".
bool wv_load(const char *path, std::vector<double> &t, std::vector<double> &p)
{
    std::ifstream in(path);
    if (!in.is_open()) return false;
    std::vector<std::string> lines;
    std::string line;
    int count = 0;
    while (std::getline(in, line)) {
        if (line.find('#') != 0) { ++count; lines.push_back(line); }
    }
    ++count;   // the reference's eof loop reads one more (empty) line
    const int len = count - 1;
    t.assign((size_t)len, 0.0);
    p.assign((size_t)len, 0.0);
    for (int n = 0; n < len && n < (int)lines.size(); ++n) {
        std::istringstream ss(lines[(size_t)n]);
        ss >> t[n] >> p[n];
    }
    return len > 0;
}

// One waveform; with wvfrm_out_step < 1e4 also the intermediate dumps
// (geoac.eqset.sph.strat.cpp:989-1002, every equation set alike):
// sc_spectrum-<L>.dat (f = n / (len dt), Re U, Im U) and sc_waveform-<L>.dat
// (source time axis, scaled pressure) in the working directory, at the
// stream's default precision, a later dump of the same label replacing the
// file.
static int wv_run_dumped(ga_ctx *ctx, int aid, const ga_prop_params &p, const ga_wvfrm_params &w,
                         const double *launch, const std::vector<double> &t, const std::vector<double> &ps,
                         double *ti, double *pi_, double *to, double *po, ga_arrival *A, int32_t *status,
                         int64_t *steps, ga_wvfrm_stats *st, double out_step, int bounces)
{
    if (!(out_step < 1.0e4))
        return ga_wnl_wvfrm(ctx, aid, &p, &w, launch, 1, t.data(), ps.data(), ti, pi_, to, po, A, status, steps, st);
    const int len = w.len;
    const size_t width = 1 + (size_t)len + 2 * ((size_t)len / 2 + 1);
    int cap = (int)std::min(4096.0, std::ceil(p.s_max / out_step) + 2.0 * (bounces + 2));
    std::vector<double> snaps;
    int32_t cnt = 0;
    for (;;) {
        snaps.assign((size_t)cap * width, 0.0);
        const int rc = ga_wnl_wvfrm_dump(ctx, aid, &p, &w, launch, 1, t.data(), ps.data(), ti, pi_, to, po, A, status,
                                         steps, st, out_step, cap, snaps.data(), &cnt);
        if (rc != GA_OK) return rc;
        if (cnt <= cap) break;
        cap = cnt;   // more crossings than the first guess: once more with room for all
    }
    const double dt = t[1] - t[0];
    for (int j = 0; j < cnt; ++j) {
        const double *S = snaps.data() + (size_t)j * width;
        char name[64];
        std::snprintf(name, sizeof name, "sc_spectrum-%04g.dat", S[0]);
        std::ofstream fs(name);
        for (int b = 0; b <= len / 2; ++b)
            fs << 1.0 / dt * b / len << '\t' << S[1 + len + 2 * b] << '\t' << S[2 + len + 2 * b] << '\n';
        std::snprintf(name, sizeof name, "sc_waveform-%04g.dat", S[0]);
        std::ofstream fw(name);
        for (int m = 0; m < len; ++m) fw << t[(size_t)m] << '\t' << S[1 + m] << '\n';
    }
    return GA_OK;
}

// -wnl_wvfrm (run_wnl_wvfrm: main.infraga.sph.cpp:1142-1479,
// main.infraga.3d.cpp:1087-1430, main.infraga.sph.rngdep.cpp:1108-1410,
// main.infraga.3d.rngdep.cpp:1051-1350): one ray, its waveform set at
// wvfrm_ref km along the ray and evolved to the last ground return;
// <id>.wvfrm_init.dat, <id>.wvfrm_out.dat and, with write_ray=true,
// <id>.raypaths.dat.  The stratified binaries take topo_file= (the source
// then defaults to the terrain's centre), the range-dependent ones their
// node files.  wvfrm_out_step < 1e4 writes the intermediate dumps
// (wv_run_dumped); the waveform length must be a power of two (the engine's
// transform).
int run_wnl_wvfrm(int argc, char **argv)
{
    constexpr bool rd = GA_CLI_GEOM == GA_GEOM_SPH_RNGDEP || GA_CLI_GEOM == GA_GEOM_3D_RNGDEP;
    constexpr bool cart = GA_CLI_GEOM == GA_GEOM_3D || GA_CLI_GEOM == GA_GEOM_3D_RNGDEP;
    const bool sph = !cart;
    const int first = rd ? 5 : 3;
    if (rd && argc < 5) {
        std::cerr << GA_CLI_NAME << ": usage: -wnl_wvfrm <profile prefix> <node file> <node file> [parameters]\n";
        return 2;
    }
    double incl = 15.0, az = 90.0;
    int bounces = 0;
    double s0 = rd ? NAN : (sph ? 30.0 : 0.0), s1 = rd ? NAN : 0.0, z_src = 0.0;
    bool src_set[2] = {false, false};
    const char *topo_file = nullptr;
    bool write_rays = false, reverse_winds = false, custom_id = false;
    double freq = 0.1, wvfrm_ref = 1.0, wvfrm_out_step = 1.0e10;
    double wp0 = 10.0, wt0 = 0.1, walpha = 4.0;
    int n_rise = 10, n_period = 40;
    long len = 8192;
    const char *wvfrm_file = "none", *wvfrm_opt = "impulse";
    const char *prof_format = "zTuvdp", *output_id = nullptr;
    ga_wvfrm_params w{};
    w.ds = 0.5;
    w.filt1_g = 0.5; w.filt1_n1 = 4.0; w.filt1_n2 = 0.5;
    w.filt2_g = 0.75; w.filt2_n1 = 2.0; w.filt2_n2 = 1.0;
    ga_prop_params p;
    ga_prop_params_default(&p, GA_CLI_GEOM);
    double z_grnd = NAN;
    for (int i = first; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key(a, "inclination=", v)) incl = std::atof(v);
        else if (key(a, "azimuth=", v)) az = std::atof(v);
        else if (key(a, "bounces=", v)) bounces = std::max(0, std::atoi(v));
        else if (key2(a, "src_alt=", "alt_src=", v)) z_src = std::atof(v);
        else if (sph && key2(a, "src_lat=", "lat_src=", v)) { s0 = std::atof(v); src_set[0] = true; }
        else if (sph && key2(a, "src_lon=", "lon_src=", v)) { s1 = std::atof(v); src_set[1] = true; }
        else if (!sph && key2(a, "src_x=", "x_src=", v)) { s0 = std::atof(v); src_set[0] = true; }
        // main.infraga.3d.cpp compares "y_src_y=" over 6 characters, so
        // only src_y= sets y there
        else if (!sph && (rd ? key2(a, "src_y=", "y_src=", v) : key(a, "src_y=", v))) { s1 = std::atof(v); src_set[1] = true; }
        else if (rd && key(a, "calc_amp=", v)) { /* the waveform needs the auxiliary equations */ }
        else if (key(a, "write_ray=", v)) write_rays = str2bool(v);
        else if (key(a, "wvfrm_file=", v)) wvfrm_file = v;
        else if (key(a, "wvfrm_opt=", v)) wvfrm_opt = v;
        else if (key(a, "wvfrm_p0=", v)) wp0 = std::atof(v);
        else if (key(a, "wvfrm_t0=", v)) wt0 = std::atof(v);
        else if (key(a, "wvfrm_alpha=", v)) walpha = std::atof(v);
        else if (key(a, "wvfrm_ref=", v)) wvfrm_ref = std::atof(v);
        else if (key(a, "wvfrm_out_step=", v)) wvfrm_out_step = std::atof(v);
        else if (key(a, "wvfrm_ds=", v)) w.ds = std::atof(v);
        // the reference reads N_rise from 15 characters in (two past the '=')
        else if (key(a, "wvfrm_N_rise=", v)) n_rise = (int)std::atof(std::strlen(a) >= 15 ? a + 15 : "");
        else if (key(a, "wvfrm_N_period=", v)) n_period = (int)std::atof(v);
        else if (key(a, "wvfrm_len=", v)) len = (long)std::atof(v);
        else if (key(a, "wvfrm_filt1_g=", v)) w.filt1_g = std::atof(v);
        else if (key(a, "wvfrm_filt1_n1=", v)) w.filt1_n1 = std::atof(v);
        else if (key(a, "wvfrm_filt1_n2=", v)) w.filt1_n2 = std::atof(v);
        else if (key(a, "wvfrm_filt2_g=", v)) w.filt2_g = std::atof(v);
        else if (key(a, "wvfrm_filt2_n1=", v)) w.filt2_n1 = std::atof(v);
        else if (key(a, "wvfrm_filt2_n2=", v)) w.filt2_n2 = std::atof(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) p.rng_max = std::atof(v);
        else if (sph && key2(a, "max_time=", "time_max=", v)) { /* no time limit in -wnl_wvfrm */ }
        else if (sph && key2(a, "min_lat=", "lat_min=", v)) p.bounds[0] = std::atof(v);
        else if (sph && key2(a, "max_lat=", "lat_max=", v)) p.bounds[1] = std::atof(v);
        else if (sph && key2(a, "min_lon=", "lon_min=", v)) p.bounds[2] = std::atof(v);
        else if (sph && key2(a, "max_lon=", "lon_max=", v)) p.bounds[3] = std::atof(v);
        else if (!sph && key2(a, "min_x=", "x_min=", v)) p.bounds[0] = std::atof(v);
        else if (!sph && key2(a, "max_x=", "x_max=", v)) p.bounds[1] = std::atof(v);
        else if (!sph && key2(a, "min_y=", "y_min=", v)) p.bounds[2] = std::atof(v);
        else if (!sph && key2(a, "max_y=", "y_max=", v)) p.bounds[3] = std::atof(v);
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_use_BLw=", v)) p.topo_use_blw = str2bool(v) ? 1 : 0;
        else if (key(a, "topo_file=", v)) topo_file = v;
        else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }
    std::vector<double> t, ps;
    const bool from_file = std::strncmp(wvfrm_file, "none", 4) != 0;
    if (from_file) {
        if (!wv_load(wvfrm_file, t, ps)) {
            std::cerr << GA_CLI_NAME << ": cannot read waveform file " << wvfrm_file << '\n';
            return 1;
        }
        len = (long)t.size();
    } else {
        if (len < 1 || len > (1L << 24)) {
            std::cerr << GA_CLI_NAME << ": bad wvfrm_len\n";
            return 2;
        }
        wv_build(wvfrm_opt, wp0, wt0, walpha, (int)len, n_rise, n_period, t, ps);
    }
    w.len = (int32_t)len;
    w.ref = wvfrm_ref;

    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int aid = 0;
    double z0 = 0.0;
    const bool is_topo = topo_file != nullptr;
    if (rd || is_topo) {
        // range-dependent node files, or the stratified profile over terrain
        // as identical columns (set_region with topo_file); the source at the
        // region's centre unless given (main.infraga.sph.cpp:1187-1191,
        // main.infraga.3d.cpp:1131-1134, main.infraga.sph.rngdep.cpp:1140-1142)
        int rc;
        if (rd) rc = (cart ? ga_rngdep_load_files_xy : ga_rngdep_load_files)(ctx, argv[2], argv[3], argv[4], topo_file,
                                                                          prof_format, reverse_winds ? 1 : 0, &aid);
        else rc = (cart ? ga_atmo_load_file_topo_xy : ga_atmo_load_file_topo)(ctx, argv[2], topo_file, prof_format,
                                                                            reverse_winds ? 1 : 0, &aid);
        if (rc != GA_OK) return die(ctx, "cannot load the atmosphere");
        ga_prop_params q;
        ga_prop_params_default(&q, cart ? GA_GEOM_3D_RNGDEP : GA_GEOM_SPH_RNGDEP);
        q.abs_coeff = p.abs_coeff;
        q.alt_max = p.alt_max;
        q.rng_max = p.rng_max;
        q.ds_min = p.ds_min;
        q.ds_max = p.ds_max;
        q.s_max = p.s_max;
        q.topo_use_blw = is_topo ? p.topo_use_blw : 0;
        for (int k = 0; k < 4; ++k) q.bounds[k] = NAN;
        double info[8];
        if (ga_rngdep_info(ctx, aid, info) != GA_OK) return die(ctx, "region");
        const double deg = cart ? 1.0 : Pi / 180.0;
        if (!src_set[0]) s0 = (info[1] + info[0]) / 2.0 / deg;
        if (!src_set[1]) s1 = (info[3] + info[2]) / 2.0 / deg;
        z0 = (!is_topo && !std::isnan(z_grnd)) ? z_grnd : info[5];
        if (!is_topo) q.z_grnd = z0;
        double pt[3] = {s0 * deg, s1 * deg, 0.0}, atmo_out[40], topo_out[10];
        if (ga_rngdep_eval(ctx, aid, pt, 1, atmo_out, topo_out) != GA_OK) return die(ctx, "terrain at the source");
        const double zg = is_topo ? topo_out[0] : (cart ? 0.0 : z0);
        z_src = std::max(z_src, zg);
        p = q;
    } else {
        if (ga_atmo_load_file(ctx, argv[2], prof_format, reverse_winds ? 1 : 0, &aid) != GA_OK)
            return die(ctx, "cannot load the atmosphere");
        z0 = std::isnan(z_grnd) ? profile_bottom(argv[2]) : z_grnd;
        p.z_grnd = z0;
        z_src = std::max(z_src, z0);
    }
    p.bounces = bounces;
    p.freq = freq;
    p.calc_amp = 1;
    p.src[0] = s0;
    p.src[1] = s1;
    p.src[2] = z_src;
    const std::string id = custom_id ? std::string(output_id) : (rd ? std::string(argv[2]) : strip_ext(argv[2]));
    const double launch[2] = {incl, az};
    std::vector<double> ti((size_t)len), pi_((size_t)len), to((size_t)len), po((size_t)len);
    ga_arrival A{};
    int32_t status = 0;
    int64_t steps = 0;
    ga_wvfrm_stats st{};
    if (wv_run_dumped(ctx, aid, p, w, launch, t, ps, ti.data(), pi_.data(), to.data(), po.data(), &A, &status, &steps,
                      &st, wvfrm_out_step, bounces) != GA_OK)
        return die(ctx, "waveform failed");
    if (write_rays) {
        // the same ray's write_rays rows (m == 1 or m % 15 == 0 of every leg)
        ga_prop_params q = p;
        q.turn_ht_min = -1.0e300;
        q.time_max = 1.0e300;
        std::vector<ga_arrival> arr((size_t)bounces + 1);
        int n_out = 0;
        std::vector<ga_path_pt> pts((size_t)1 << 16);
        int64_t n_pts = 0;
        ga_run_stats rs;
        int rc = ga_fan_run_paths(ctx, aid, &q, launch, 1, GA_PATH_RAYS, arr.data(), bounces + 1, &n_out, pts.data(),
                                  (int64_t)pts.size(), &n_pts, &rs);
        if (rc == GA_E_CAP && n_pts > (int64_t)pts.size()) {
            pts.resize((size_t)n_pts);
            rc = ga_fan_run_paths(ctx, aid, &q, launch, 1, GA_PATH_RAYS, arr.data(), bounces + 1, &n_out, pts.data(),
                                  (int64_t)pts.size(), &n_pts, &rs);
        }
        if (rc != GA_OK) return die(ctx, "ray path failed");
        std::ofstream raypath(id + ".raypaths.dat");
        // the sph writer's header names z first (main.infraga.sph.cpp:1313-1319)
        raypath << (sph ? "# z [km]" : "# x [km]") << '\t' << (sph ? "lat [deg]" : "y [km]") << '\t'
                << (sph ? "lon [deg]" : "z [km]") << '\t' << "trans. coeff. [dB]" << '\t' << "absorption [dB]" << '\t'
                << "time [s]" << '\n';
        if (sph) raypath << std::setprecision(8);
        for (int64_t k = 0; k < n_pts; ++k) {
            const ga_path_pt &P = pts[(size_t)k];
            raypath << P.pos0 << '\t' << P.pos1 << '\t' << P.z << '\t' << P.tl << '\t' << P.absorption << '\t' << P.time
                    << '\n';
        }
    }
    if (status < 0) {
        std::cout << "\n\t\tRay path does not reach the waveform reference location (" << wvfrm_ref << " km).\n\n";
        ga_ctx_destroy(ctx);
        return 1;
    }
    {
        std::ofstream f(id + ".wvfrm_init.dat");
        f << "# t [sec]" << '\t' << "p(t) [Pa]" << '\n';
        for (long n = 0; n < len; ++n) f << std::setprecision(8) << ti[(size_t)n] << '\t' << pi_[(size_t)n] << '\n';
    }
    if (status == 1) {
        std::cout << '\n' << '\t' << "Arrival summary:" << '\n';
        if (sph) {
            std::cout << "\t\tlatitude [deg] = " << std::setprecision(8) << A.pos0 << '\n';
            std::cout << "\t\tlongitude [deg] = " << std::setprecision(8) << A.pos1 << '\n';
        } else {
            std::cout << "\t\tx [km E-W] = " << A.pos0 << '\n';
            std::cout << "\t\ty [km N-S] = " << A.pos1 << '\n';
        }
        std::cout << "\t\ttime [s] = " << A.time << '\n';
        std::cout << "\t\tcelerity [km/s] = " << A.cel << '\n';
        std::cout << "\t\tturning height [km] = " << A.turn_ht << '\n';
        std::cout << "\t\tarrival inclination [deg] = " << A.incl_arr << '\n';
        std::cout << "\t\tback azimuth" << (sph ? " [deg]" : "") << " = " << A.back_az << '\n';
        std::cout << (sph ? "\t\ttrans. coeff. [dB] = " : "\t\tattenuation (geometric) [dB] = ") << A.tl << '\n';
        std::cout << "\t\tabsorption [dB] = " << A.absorption << '\n' << '\n';
    } else {
        std::cout << "\n\t\tRay path does not return to the ground.\n\n";
    }
    std::ofstream f(id + ".wvfrm_out.dat");
    f << "# " << GA_CLI_NAME << " -wnl_wvfrm summary:" << '\n';
    if (rd && cart) {
        f << "#" << '\t' << "profile prefix: " << argv[2] << '\n';
        f << "#" << '\t' << "profile grid files: " << argv[3] << ", " << argv[4] << '\n';
    } else {
        f << "#" << '\t' << "profile: " << argv[2] << '\n';
    }
    f << "#" << '\t' << "inclination: " << (incl * (Pi / 180.0)) * (180.0 / Pi) << '\n';
    // the reference prints geoac::phi (90 - azimuth) here
    f << "#" << '\t' << "azimuth: " << (Pi / 2.0 - az * (Pi / 180.0)) * (180.0 / Pi) << '\n';
    f << "#" << '\t' << "bounces: " << bounces << '\n';
    if (sph)
        f << "#" << '\t' << "source (lat, lon, alt): " << (s0 * (Pi / 180.0)) * (180.0 / Pi) << ", "
          << (s1 * (Pi / 180.0)) * (180.0 / Pi) << ", " << z_src << '\n';
    else f << "#" << '\t' << "source location: " << s0 << ", " << s1 << ", " << z_src << '\n';
    if (is_topo) f << "#" << '\t' << "topo file:" << topo_file << '\n';
    else f << "#" << '\t' << "ground elevation: " << z0 << '\n';
    // the spherical range-dependent writer adds this line, spelling included
    // (main.infraga.sph.rngdep.cpp:1388)
    if (rd && !cart) f << "#" << '\t' << "ground evevation: " << z0 << '\n';
    f << "#" << '\t' << "S&B atten coeff: " << p.abs_coeff << '\n';
    if (from_file) f << "#" << '\t' << "waveform source file: " << wvfrm_file << '\n';
    else {
        f << "#" << '\t' << "waveform option: " << wvfrm_opt << '\n';
        f << "#" << '\t' << "waveform p0: " << wp0 << '\n';
        f << "#" << '\t' << "waveform t0: " << wt0 << '\n';
        f << "#" << '\t' << "waveform alpha: " << walpha << '\n';
    }
    f << "#" << '\t' << "waveform reference location: " << wvfrm_ref << '\n' << '\n';
    f << "# t [sec]" << '\t' << "p(t) [Pa]" << '\n';
    for (long n = 0; n < len; ++n) f << std::setprecision(8) << to[(size_t)n] << '\t' << po[(size_t)n] << '\n';
    std::cout << GA_CLI_NAME << ": waveform of " << len << " samples, " << steps << " evolution steps, "
              << st.ray_steps << " ray-steps, " << st.trace_ms + st.evolve_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

// infraga-2d -wnl_wvfrm (run_wnl_wvfrm, main.infraga.2d.cpp): the waveform
// along one ray of the 2D geometry, every leg (ground or elevated
// reflection) counted toward bounces; <id>.wvfrm_init.dat and
// <id>.wvfrm_out.dat.  write_ray=true is refused (the path file would follow
// the -prop leg count), as are intermediate dumps.
int run_wnl_wvfrm_2d(int argc, char **argv)
{
    double incl = 15.0, az = 90.0, z_src = 0.0, freq = 0.1, wvfrm_ref = 1.0, wvfrm_out_step = 1.0e10;
    double wp0 = 10.0, wt0 = 0.1, walpha = 4.0, z_grnd = NAN;
    int bounces = 0, n_rise = 10, n_period = 40;
    long len = 8192;
    bool reverse_winds = false, custom_id = false;
    const char *wvfrm_file = "none", *wvfrm_opt = "impulse", *prof_format = "zTuvdp", *output_id = nullptr;
    const char *topo_file = nullptr;
    ga_wvfrm_params w{};
    w.ds = 0.5;
    w.filt1_g = 0.5; w.filt1_n1 = 4.0; w.filt1_n2 = 0.5;
    w.filt2_g = 0.75; w.filt2_n1 = 2.0; w.filt2_n2 = 1.0;
    ga_prop_params p;
    ga_prop_params_default(&p, GA_GEOM_2D);
    for (int i = 3; i < argc; i++) {
        const char *a = argv[i], *v = nullptr;
        if (key(a, "inclination=", v)) incl = std::atof(v);
        else if (key(a, "azimuth=", v)) az = std::atof(v);
        else if (key(a, "bounces=", v)) bounces = std::max(0, std::atoi(v));
        else if (key(a, "src_alt=", v)) z_src = std::atof(v);
        else if (key(a, "write_ray=", v)) {
            if (str2bool(v)) {
                std::cerr << GA_CLI_NAME << ": write_ray with -wnl_wvfrm is not supported by this engine\n";
                return 2;
            }
        }
        else if (key(a, "wvfrm_file=", v)) wvfrm_file = v;
        else if (key(a, "wvfrm_opt=", v)) wvfrm_opt = v;
        else if (key(a, "wvfrm_p0=", v)) wp0 = std::atof(v);
        else if (key(a, "wvfrm_t0=", v)) wt0 = std::atof(v);
        else if (key(a, "wvfrm_alpha=", v)) walpha = std::atof(v);
        else if (key(a, "wvfrm_ref=", v)) wvfrm_ref = std::atof(v);
        else if (key(a, "wvfrm_out_step=", v)) wvfrm_out_step = std::atof(v);
        else if (key(a, "wvfrm_ds=", v)) w.ds = std::atof(v);
        else if (key(a, "wvfrm_N_rise=", v)) n_rise = (int)std::atof(std::strlen(a) >= 15 ? a + 15 : "");
        else if (key(a, "wvfrm_N_period=", v)) n_period = (int)std::atof(v);
        else if (key(a, "wvfrm_len=", v)) len = (long)std::atof(v);
        else if (key(a, "wvfrm_filt1_g=", v)) w.filt1_g = std::atof(v);
        else if (key(a, "wvfrm_filt1_n1=", v)) w.filt1_n1 = std::atof(v);
        else if (key(a, "wvfrm_filt1_n2=", v)) w.filt1_n2 = std::atof(v);
        else if (key(a, "wvfrm_filt2_g=", v)) w.filt2_g = std::atof(v);
        else if (key(a, "wvfrm_filt2_n1=", v)) w.filt2_n1 = std::atof(v);
        else if (key(a, "wvfrm_filt2_n2=", v)) w.filt2_n2 = std::atof(v);
        else if (key(a, "freq=", v)) freq = std::atof(v);
        else if (key(a, "abs_coeff=", v)) p.abs_coeff = std::max(0.0, std::atof(v));
        else if (key2(a, "max_alt=", "alt_max=", v)) p.alt_max = std::atof(v);
        else if (key2(a, "max_rng=", "rng_max=", v)) p.rng_max = std::atof(v);
        else if (key2(a, "min_ds=", "ds_min=", v)) p.ds_min = std::atof(v);
        else if (key2(a, "max_ds=", "ds_max=", v)) p.ds_max = std::atof(v);
        else if (key2(a, "max_s=", "s_max=", v)) p.s_max = std::atof(v);
        else if (key2(a, "refl_alt=", "alt_refl=", v)) p.refl_alt = std::atof(v);
        else if (key(a, "write_atmo=", v)) { /* profile dump: not produced */ }
        else if (key(a, "prof_format=", v)) prof_format = v;
        else if (key(a, "reverse_winds=", v)) reverse_winds = str2bool(v);
        else if (key(a, "output_id=", v)) { custom_id = true; output_id = v; }
        else if (key(a, "z_grnd=", v)) z_grnd = std::atof(v);
        else if (key(a, "topo_file=", v)) topo_file = v;
        else {
            std::cerr << GA_CLI_NAME << ": unrecognized parameter entry: " << a << '\n';
            return 2;
        }
    }
    std::vector<double> t, ps;
    const bool from_file = std::strncmp(wvfrm_file, "none", 4) != 0;
    if (from_file) {
        if (!wv_load(wvfrm_file, t, ps)) {
            std::cerr << GA_CLI_NAME << ": cannot read waveform file " << wvfrm_file << '\n';
            return 1;
        }
        len = (long)t.size();
    } else {
        if (len < 1 || len > (1L << 24)) {
            std::cerr << GA_CLI_NAME << ": bad wvfrm_len\n";
            return 2;
        }
        wv_build(wvfrm_opt, wp0, wt0, walpha, (int)len, n_rise, n_period, t, ps);
    }
    w.len = (int32_t)len;
    w.ref = wvfrm_ref;

    ga_ctx *ctx = nullptr;
    if (ga_ctx_create(&ctx, 0) != GA_OK) return die(nullptr, "no usable GPU");
    int aid = 0;
    if (ga_atmo_load_file(ctx, argv[2], prof_format, reverse_winds ? 1 : 0, &aid) != GA_OK)
        return die(ctx, "cannot load the atmosphere");
    if (topo_file) {
        std::ifstream tf(topo_file);
        if (!tf) return die(ctx, "cannot read the terrain file");
        std::vector<double> r, zt;
        std::string line;
        while (std::getline(tf, line)) {
            if (line.find('#') == 0) continue;
            double a = 0.0, b = 0.0;
            if (std::sscanf(line.c_str(), "%lf %lf", &a, &b) == 2) { r.push_back(a); zt.push_back(b); }
        }
        if (ga_topo_load_2d(ctx, aid, r.data(), zt.data(), (int)r.size()) != GA_OK) return die(ctx, "terrain");
    }
    const double zb = profile_bottom(argv[2]);
    const double z0 = (!topo_file && !std::isnan(z_grnd)) ? z_grnd : zb;
    p.z_grnd = topo_file ? NAN : z0;
    double t0[3];
    if (ga_topo_eval_2d(ctx, aid, 0.0, t0) != GA_OK) return die(ctx, "terrain");
    z_src = std::max(z_src, topo_file ? t0[0] : z0);          // z_src = max(z_src, topo::z(r_src))
    p.src[2] = z_src;
    p.bounces = bounces;
    p.freq = freq;
    p.calc_amp = 1;
    const std::string id = custom_id ? std::string(output_id) : strip_ext(argv[2]);
    const double launch[2] = {incl, az};
    std::vector<double> ti((size_t)len), pi_((size_t)len), to((size_t)len), po((size_t)len);
    ga_arrival A{};
    int32_t status = 0;
    int64_t steps = 0;
    ga_wvfrm_stats st{};
    if (wv_run_dumped(ctx, aid, p, w, launch, t, ps, ti.data(), pi_.data(), to.data(), po.data(), &A, &status, &steps,
                      &st, wvfrm_out_step, bounces) != GA_OK)
        return die(ctx, "waveform failed");
    if (status < 0) {
        std::cout << "\n\t\tRay path does not reach the waveform reference location (" << wvfrm_ref << " km).\n\n";
        ga_ctx_destroy(ctx);
        return 1;
    }
    {
        std::ofstream f(id + ".wvfrm_init.dat");
        f << "# t [sec]" << '\t' << "p(t) [Pa]" << '\n';
        for (long n = 0; n < len; ++n) f << std::setprecision(8) << ti[(size_t)n] << '\t' << pi_[(size_t)n] << '\n';
    }
    if (status == 1) {
        std::cout << '\n' << '\t' << "Arrival summary:" << '\n';
        std::cout << "\t\trange [km] = " << A.pos0 << '\n';
        std::cout << "\t\ttime [s] = " << A.time << '\n';
        std::cout << "\t\tcelerity [km/s] = " << A.cel << '\n';
        std::cout << "\t\tturning height [km] = " << A.turn_ht << '\n';
        std::cout << "\t\tarrival inclination [deg] = " << A.incl_arr << '\n';
        std::cout << "\t\ttrans. coeff. [dB] = " << A.tl << '\n';
        std::cout << "\t\tabsorption [dB] = " << A.absorption << '\n' << '\n';
    } else {
        std::cout << "\n\t\tRay path does not return to the ground.\n\n";
    }
    std::ofstream f(id + ".wvfrm_out.dat");
    f << "# " << GA_CLI_NAME << " -wnl_wvfrm summary:" << '\n';
    f << "#" << '\t' << "profile: " << argv[2] << '\n';
    f << "#" << '\t' << "inclination: " << (incl * (Pi / 180.0)) * (180.0 / Pi) << '\n';
    f << "#" << '\t' << "azimuth: " << (Pi / 2.0 - az * (Pi / 180.0)) * (180.0 / Pi) << '\n';
    f << "#" << '\t' << "bounces: " << bounces << '\n';
    f << "#" << '\t' << "source location: " << 0.0 << ", " << z_src << '\n';
    if (topo_file) f << "#" << '\t' << "topo file:" << topo_file << '\n';
    else f << "#" << '\t' << "ground elevation: " << z0 << '\n';
    f << "#" << '\t' << "S&B atten coeff: " << p.abs_coeff << '\n';
    if (from_file) f << "#" << '\t' << "waveform source file: " << wvfrm_file << '\n';
    else {
        f << "#" << '\t' << "waveform option: " << wvfrm_opt << '\n';
        f << "#" << '\t' << "waveform p0: " << wp0 << '\n';
        f << "#" << '\t' << "waveform t0: " << wt0 << '\n';
        f << "#" << '\t' << "waveform alpha: " << walpha << '\n';
    }
    f << "#" << '\t' << "waveform reference location: " << wvfrm_ref << '\n' << '\n';
    f << "# t [sec]" << '\t' << "p(t) [Pa]" << '\n';
    for (long n = 0; n < len; ++n) f << std::setprecision(8) << to[(size_t)n] << '\t' << po[(size_t)n] << '\n';
    std::cout << GA_CLI_NAME << ": waveform of " << len << " samples, " << steps << " evolution steps, "
              << st.ray_steps << " ray-steps, " << st.trace_ms + st.evolve_ms << " ms on the GPU\n";
    ga_ctx_destroy(ctx);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    // under mpirun every rank starts a copy: rank 0 does the whole job on the GPU
    if (mpi_rank() > 0) return 0;
    if (argc < 3) { usage(); return argc < 2 ? 0 : 2; }
    if (GA_CLI_GEOM == GA_GEOM_2D) {
        if (std::strncmp(argv[1], "-prop", 5) == 0) return run_prop_2d(argc, argv);
        if (std::strncmp(argv[1], "-wnl_wvfrm", 10) == 0 || std::strncmp(argv[1], "-w", 2) == 0)
            return run_wnl_wvfrm_2d(argc, argv);
        std::cerr << GA_CLI_NAME << ": option " << argv[1] << " is not supported by this engine\n";
        return 2;
    }
    if (GA_CLI_GEOM == GA_GEOM_3D_RNGDEP || GA_CLI_GEOM == GA_GEOM_SPH_RNGDEP) {
        if (std::strncmp(argv[1], "-prop", 5) == 0) return run_prop_rd(argc, argv);
        if (std::strncmp(argv[1], "-eig_search", 11) == 0) return run_eig_rd(argc, argv, false);
        if (std::strncmp(argv[1], "-eig_direct", 11) == 0) return run_eig_rd(argc, argv, true);
        if (std::strncmp(argv[1], "-wnl_wvfrm", 10) == 0 || std::strncmp(argv[1], "-w", 2) == 0)
            return run_wnl_wvfrm(argc, argv);
        std::cerr << GA_CLI_NAME << ": option " << argv[1] << " is not supported by this engine\n";
        return 2;
    }
    if (std::strncmp(argv[1], "-prop", 5) == 0) return run_prop(argc, argv);
    if (std::strncmp(argv[1], "-eig_search", 11) == 0) return run_eig_search(argc, argv);
    if (std::strncmp(argv[1], "-eig_direct", 11) == 0) return run_eig_direct(argc, argv);
    if (std::strncmp(argv[1], "-wnl_wvfrm", 10) == 0 || std::strncmp(argv[1], "-w", 2) == 0)
        return run_wnl_wvfrm(argc, argv);
    std::cerr << GA_CLI_NAME << ": option " << argv[1] << " is not supported by this engine\n";
    usage();
    return 2;
}
