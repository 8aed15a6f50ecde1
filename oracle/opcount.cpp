// opcount.cpp -- algorithmic FP64 operation count of the propagation path.
//
// TEST / MEASUREMENT INFRASTRUCTURE.  Compiles the oracle restatement
// (ga_oracle.c) with every binary64 value replaced by a counting scalar, runs
// a seeded sample of the 162k-ray spherical fan and reports, per ray-step
// (RK4 iteration), the number of add/sub, mul, div and sqrt operations and of
// transcendental calls.  F_step = add + mul + div + sqrt (each counted once;
// the restatement contains no fused multiply-add).
//
//   g++ -O1 -std=c++17 -o opcount oracle/opcount.cpp && ./opcount profile.txt
//   (profile.txt: n, then n rows z c u v rho in engine units)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static long long g_add, g_mul, g_div, g_sqrt, g_trans, g_cmp;

struct ga_real {
    double v;
    ga_real() = default;
    constexpr ga_real(double x) : v(x) {}
    explicit operator int() const { return (int)v; }
    explicit operator double() const { return v; }
};
static inline ga_real operator+(ga_real a, ga_real b) { ++g_add; return a.v + b.v; }
static inline ga_real operator-(ga_real a, ga_real b) { ++g_add; return a.v - b.v; }
static inline ga_real operator*(ga_real a, ga_real b) { ++g_mul; return a.v * b.v; }
static inline ga_real operator/(ga_real a, ga_real b) { ++g_div; return a.v / b.v; }
static inline ga_real operator-(ga_real a) { return -a.v; }
static inline ga_real operator+(ga_real a) { return a; }
static inline ga_real &operator+=(ga_real &a, ga_real b) { a = a + b; return a; }
static inline ga_real &operator-=(ga_real &a, ga_real b) { a = a - b; return a; }
static inline ga_real &operator*=(ga_real &a, ga_real b) { a = a * b; return a; }
static inline ga_real &operator/=(ga_real &a, ga_real b) { a = a / b; return a; }
#define GA_CMP(op) \
    static inline bool operator op(ga_real a, ga_real b) { ++g_cmp; return a.v op b.v; } \
    static inline bool operator op(ga_real a, double b) { ++g_cmp; return a.v op b; } \
    static inline bool operator op(double a, ga_real b) { ++g_cmp; return a op b.v; }
GA_CMP(<) GA_CMP(>) GA_CMP(<=) GA_CMP(>=) GA_CMP(==) GA_CMP(!=)
static inline ga_real sqrt(ga_real a) { ++g_sqrt; return std::sqrt(a.v); }
static inline ga_real fabs(ga_real a) { return std::fabs(a.v); }
#define GA_T1(f) static inline ga_real f(ga_real a) { ++g_trans; return std::f(a.v); }
GA_T1(exp) GA_T1(sin) GA_T1(cos) GA_T1(tan) GA_T1(asin) GA_T1(log10)
static inline ga_real atan2(ga_real a, ga_real b) { ++g_trans; return std::atan2(a.v, b.v); }
// pow(x, n) for n = 2..5 counts as n - 1 products (pow(x, 2) is exactly
// x*x); other exponents are transcendental calls.
static inline ga_real pow(ga_real a, ga_real b)
{
    if (b.v == 2.0) { ++g_mul; return a.v * a.v; }
    if (b.v == 3.0 || b.v == 4.0 || b.v == 5.0) { g_mul += (long long)b.v - 1; return std::pow(a.v, b.v); }
    ++g_trans;
    return std::pow(a.v, b.v);
}

#define GO_OPCOUNT 1
#define double ga_real
extern "C" {
#include "ga_oracle.c"
}
#undef double

int main(int argc, char **argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: opcount profile.txt [n_rays]\n"); return 2; }
    FILE *f = std::fopen(argv[1], "r");
    if (!f) return 2;
    int n = 0;
    if (std::fscanf(f, "%d", &n) != 1) return 2;
    std::vector<ga_real> z(n), c(n), u(n), v(n), rho(n);
    for (int i = 0; i < n; ++i) {
        double a, b, cc, d, e;
        if (std::fscanf(f, "%lf %lf %lf %lf %lf", &a, &b, &cc, &d, &e) != 5) return 2;
        z[i] = a; c[i] = b; u[i] = cc; v[i] = d; rho[i] = e;
    }
    std::fclose(f);
    int nrays = argc > 2 ? std::atoi(argv[2]) : 96;
    go_atmo A;
    go_atmo_init(&A, GA_GEOM_SPH, z.data(), c.data(), u.data(), v.data(), rho.data(), n);
    go_params P;
    go_params_default(&P, GA_GEOM_SPH, &A, 1);
    go_prop pr;
    pr.bounces = 5; pr.freq = 0.1; pr.turn_ht_min = 0.2; pr.src0 = 30.0; pr.src1 = 0.0; pr.src_alt = 0.0;
    // seeded sample of the 450 x 360 launch grid: every stride-th ray
    std::vector<ga_real> L;
    long long total = 450LL * 360LL, stride = total / nrays;
    for (long long i = 0; i < total; i += stride) {
        long long ia = i / 450, ii = i % 450;
        L.push_back(0.1 + 0.1 * (double)ii);
        L.push_back((double)ia);
    }
    int nl = (int)(L.size() / 2);
    std::vector<go_arrival> out(nl * 6);
    int n_out = 0;
    long long steps = 0;
    g_add = g_mul = g_div = g_sqrt = g_trans = g_cmp = 0;
    go_fan(&A, &P, &pr, L.data(), nl, out.data(), (int)out.size(), &n_out, &steps);
    double s = (double)steps;
    std::printf("{\"rays\": %d, \"ray_steps\": %lld, \"add\": %.2f, \"mul\": %.2f, \"div\": %.2f, \"sqrt\": %.2f, "
                "\"transcendental\": %.2f, \"compare\": %.2f, \"F_step\": %.1f}\n",
                nl, steps, g_add / s, g_mul / s, g_div / s, g_sqrt / s, g_trans / s, g_cmp / s,
                (g_add + g_mul + g_div + g_sqrt) / s);
    return 0;
}
