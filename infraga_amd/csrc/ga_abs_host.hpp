// ga_abs_host.hpp -- host-side layout of the absorption table (the device
// part and the record format are in ga_device.hpp): the pieces of every
// node interval and the interpolation basis.  Plain C++, shared by the
// engine and the host build that tests the table.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace ga {

// Switch heights of the composition fits in sb_smooth (z > h selects the
// upper formula).
constexpr double ABS_SWITCH[5] = {30.0, 76.0, 80.0, 90.0, 95.0};
constexpr int ABS_GEO = 6;   // lo, hi, split_lo, split_hi, other record's index, node interval

// Pieces of the table over the node heights z: per pieces of equal width
// in every interval (per = the widest interval / ABS_PIECE, rounded up),
// then one upper-side record for each piece that contains a switch height;
// a piece that starts at a switch height sends z equal to it to the piece
// below.
// geo receives ABS_GEO doubles per record; false when an interval is wider
// than ABS_MAX_PER pieces allow (no table).
inline bool abs_layout(const std::vector<double> &z, std::vector<double> &geo, int &per)
{
    const int n = (int)z.size();
    geo.clear();
    per = 1;
    if (n < 2) return false;
    for (int i = 0; i + 1 < n; ++i) {
        const double dz = z[i + 1] - z[i];
        if (!(dz > 0.0)) return false;
        const double k = std::ceil(dz / ABS_PIECE - 1.0e-9);
        if (k > ABS_MAX_PER) return false;
        per = std::max(per, (int)k);
    }
    const double inf = std::numeric_limits<double>::infinity();
    const size_t n_main = (size_t)(n - 1) * per;
    std::vector<double> upper;
    geo.resize(n_main * ABS_GEO);
    for (int i = 0; i + 1 < n; ++i) {
        const double dz = z[i + 1] - z[i];
        for (int j = 0; j < per; ++j) {
            const size_t k = (size_t)i * per + j;
            const double lo = (j == 0) ? z[i] : z[i] + dz * j / per;
            const double hi = (j == per - 1) ? z[i + 1] : z[i] + dz * (j + 1) / per;
            double *g = &geo[k * ABS_GEO];
            g[0] = lo; g[1] = hi; g[2] = -inf; g[3] = inf; g[4] = 0.0; g[5] = i;
            for (double b : ABS_SWITCH) {
                if (lo < b && b < hi) {
                    // lower side here, upper side in a record of its own
                    g[1] = b;
                    g[3] = b;
                    g[4] = (double)(n_main + upper.size() / ABS_GEO);
                    upper.insert(upper.end(), {b, hi, -inf, inf, 0.0, (double)i});
                    break;
                }
                if (b == lo && k > 0) {
                    // z == b belongs to the lower formula: the piece below
                    g[2] = b;
                    g[4] = (double)(k - 1);
                    break;
                }
            }
        }
    }
    geo.insert(geo.end(), upper.begin(), upper.end());
    return true;
}

// Chebyshev points and the inverse of the Vandermonde matrix there
// (Gauss-Jordan elimination in long double).
inline AbsBasis abs_basis()
{
    constexpr int N = ABS_DEG + 1;
    AbsBasis B;
    long double a[N][2 * N];
    for (int k = 0; k < N; ++k) {
        const long double t = std::cos(3.14159265358979323846264338327950288L * (k + 0.5L) / N);
        B.t[k] = (double)t;
        long double p = 1.0L;
        for (int j = 0; j < N; ++j) {
            a[k][j] = p;
            p *= (long double)B.t[k];
            a[k][N + j] = (j == k) ? 1.0L : 0.0L;
        }
    }
    for (int c = 0; c < N; ++c) {
        int piv = c;
        for (int r = c + 1; r < N; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        for (int j = 0; j < 2 * N; ++j) std::swap(a[c][j], a[piv][j]);
        const long double d = a[c][c];
        for (int j = 0; j < 2 * N; ++j) a[c][j] /= d;
        for (int r = 0; r < N; ++r) {
            if (r == c) continue;
            const long double f = a[r][c];
            for (int j = 0; j < 2 * N; ++j) a[r][j] -= f * a[c][j];
        }
    }
    // coefficients = inv * values: inv[j][k] = (V^-1)[j][k]
    for (int j = 0; j < N; ++j)
        for (int k = 0; k < N; ++k) B.inv[j][k] = (double)a[j][N + k];
    return B;
}

}  // namespace ga
