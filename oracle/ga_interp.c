/* ga_interp.c -- CPU restatement of the reference's multi-dimensional
 * interpolants (TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of
 * the device spline code; never part of the product path).
 *
 *   go_s2d_*  two-dimensional multi-stage natural cubic spline
 *             (src/util/interpolation.cpp:262-680): natural splines along y
 *             at every x node, then, per evaluation, a natural spline along
 *             x through the y-spline values (solved again on every call).
 *   go_h3d_*  three-dimensional hybrid spline (interpolation.cpp:886-1793):
 *             natural splines along z at every (x, y) node of f and of its
 *             centred x / y differences, and per evaluation a bicubic patch
 *             in (x, y) through corner values and finite-difference slopes
 *             (16 x 16 conversion matrix, interpolation.cpp:686-703).
 *
 * Operation order follows the reference source so that, compiled the same
 * way (gcc -O2, no FMA contraction), results are bit-identical to the
 * reference's own compiled interpolation.cpp.  Reference behaviours kept:
 * the node y-derivative spline reads the x-derivative slopes
 * (interpolation.cpp:1119-1120, 1172-1173); the d2f/dz2 patch of the
 * second-derivative eval_all uses the unclamped z (:1763-1771).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ga_interp.h"

int go_find_segment(double x, const double *xv, int length, int *prev);

static double in_iv(double x, double lo, double hi)   /* interpolation.cpp:52-54 */
{
    double m = x > lo ? x : lo;
    return m < hi ? m : hi;
}

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* Thomas solve of the natural-spline slopes of f over x (stride fs in f,
 * ss in s), as interp::set / set_node_slopes do it. */
static void natural_slopes(const double *x, const double *f, int fs, double *s, int ss, int n)
{
    double *nc = (double *)malloc(sizeof(double) * (size_t)n);
    double *nd = (double *)malloc(sizeof(double) * (size_t)n);
    double ai, bi, ci, di, d1, d2;
    bi = 2.0 / (x[1] - x[0]);
    ci = 1.0 / (x[1] - x[0]);
    di = 3.0 * (f[fs] - f[0]) / pow(x[1] - x[0], 2);
    nc[0] = ci / bi;
    nd[0] = di / bi;
    for (int i = 1; i < n - 1; i++) {
        d1 = x[i] - x[i - 1];
        d2 = x[i + 1] - x[i];
        ai = 1.0 / d1; bi = 2.0 * (1.0 / d1 + 1.0 / d2); ci = 1.0 / d2;
        di = 3.0 * ((f[i * fs] - f[(i - 1) * fs]) / pow(d1, 2) + (f[(i + 1) * fs] - f[i * fs]) / pow(d2, 2));
        nc[i] = ci / (bi - nc[i - 1] * ai);
        nd[i] = (di - nd[i - 1] * ai) / (bi - nc[i - 1] * ai);
    }
    ai = 1.0 / (x[n - 1] - x[n - 2]);
    bi = 2.0 / (x[n - 1] - x[n - 2]);
    di = 3.0 * (f[(n - 1) * fs] - f[(n - 2) * fs]) / pow(x[n - 1] - x[n - 2], 2);
    nd[n - 1] = (di - nd[n - 2] * ai) / (bi - nc[n - 2] * ai);
    s[(n - 1) * ss] = nd[n - 1];
    for (int i = n - 2; i >= 0; i--) s[i * ss] = nd[i] - nc[i] * s[(i + 1) * ss];
    free(nc);
    free(nd);
}

/* Hermite cubic segment (f0, f1, slopes s0, s1, width d) at X = (x - x0)/d:
 * value, first, second and third derivatives in the reference's forms. */
static double herm_v(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return f0 + X * (df + (1.0 - X) * (A * (1.0 - X) + B * X));
}
static double herm_d1(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return df / d + (1.0 - 2.0 * X) * (A * (1.0 - X) + B * X) / d + X * (1.0 - X) * (B - A) / d;
}
static double herm_d2(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    return 2.0 * (B - 2.0 * A + (A - B) * 3.0 * X) / pow(d, 2);
}
static double herm_d3(double f0, double f1, double s0, double s1, double d, double X)
{
    double df = f1 - f0, A = s0 * d - df, B = -s1 * d + df;
    (void)X;
    return 6.0 * (A - B) / pow(d, 3);
}

/* ====================================================================
 * 2D multi-stage spline
 * ================================================================== */
int go_s2d_init(go_s2d *s, const double *x, int nx, const double *y, int ny, const double *f)
{
    if (nx < 2 || ny < 2) return -1;
    s->nx = nx; s->ny = ny; s->accel[0] = 0; s->accel[1] = 0;
    s->x = (double *)malloc(sizeof(double) * nx);
    s->y = (double *)malloc(sizeof(double) * ny);
    s->f = (double *)malloc(sizeof(double) * nx * ny);
    s->fs = (double *)malloc(sizeof(double) * nx * ny);
    memcpy(s->x, x, sizeof(double) * nx);
    memcpy(s->y, y, sizeof(double) * ny);
    memcpy(s->f, f, sizeof(double) * nx * ny);
    for (int i = 0; i < nx; i++)                                   /* interpolation.cpp:279-317 */
        natural_slopes(s->y, s->f + (size_t)i * ny, 1, s->fs + (size_t)i * ny, 1, ny);
    return 0;
}

void go_s2d_free(go_s2d *s)
{
    free(s->x); free(s->y); free(s->f); free(s->fs);
    memset(s, 0, sizeof(*s));
}

/* node value of the y-spline at x node i, derivative order m (0..3) */
static double s2d_node(const go_s2d *s, int i, int ky, double y, int m)
{
    const double *f = s->f + (size_t)i * s->ny, *sl = s->fs + (size_t)i * s->ny;
    double d = s->y[ky + 1] - s->y[ky], X = (y - s->y[ky]) / d;
    switch (m) {
    case 0: return herm_v(f[ky], f[ky + 1], sl[ky], sl[ky + 1], d, X);
    case 1: return herm_d1(f[ky], f[ky + 1], sl[ky], sl[ky + 1], d, X);
    case 2: return herm_d2(f[ky], f[ky + 1], sl[ky], sl[ky + 1], d, X);
    default: return herm_d3(f[ky], f[ky + 1], sl[ky], sl[ky + 1], d, X);
    }
}

/* x-spline through the y-derivative-m node values, its derivative of order q
 * at x (interpolation.cpp:421-546) */
static double s2d_eval(go_s2d *s, double x, double y, int m, int q)
{
    int kx = go_find_segment(x, s->x, s->nx, &s->accel[0]);
    int ky = go_find_segment(y, s->y, s->ny, &s->accel[1]);
    double *nv = (double *)malloc(sizeof(double) * s->nx), *ns = (double *)malloc(sizeof(double) * s->nx);
    for (int i = 0; i < s->nx; i++) nv[i] = s2d_node(s, i, ky, y, m);
    natural_slopes(s->x, nv, 1, ns, 1, s->nx);
    double d = s->x[kx + 1] - s->x[kx], X = (x - s->x[kx]) / d, r;
    switch (q) {
    case 0: r = herm_v(nv[kx], nv[kx + 1], ns[kx], ns[kx + 1], d, X); break;
    case 1: r = herm_d1(nv[kx], nv[kx + 1], ns[kx], ns[kx + 1], d, X); break;
    case 2: r = herm_d2(nv[kx], nv[kx + 1], ns[kx], ns[kx + 1], d, X); break;
    default: r = herm_d3(nv[kx], nv[kx + 1], ns[kx], ns[kx + 1], d, X); break;
    }
    free(nv);
    free(ns);
    return r;
}

double go_s2d_f(go_s2d *s, double x, double y) { return s2d_eval(s, x, y, 0, 0); }
/* n = 0: d/dx, else d/dy (interpolation.cpp:444-473) */
double go_s2d_df(go_s2d *s, double x, double y, int n) { return n == 0 ? s2d_eval(s, x, y, 0, 1) : s2d_eval(s, x, y, 1, 0); }
/* n1 + n2 counts y derivatives (interpolation.cpp:475-508) */
double go_s2d_ddf(go_s2d *s, double x, double y, int n1, int n2)
{
    int m = n1 + n2;
    return s2d_eval(s, x, y, m, 2 - m);
}
/* n1 + n2 + n3 counts y derivatives (interpolation.cpp:510-546) */
double go_s2d_dddf(go_s2d *s, double x, double y, int n1, int n2, int n3)
{
    int m = n1 + n2 + n3;
    return s2d_eval(s, x, y, m, 3 - m);
}

/* ====================================================================
 * 3D hybrid spline
 * ================================================================== */
static const double BICUBIC[16][16] = {                            /* interpolation.cpp:686-703 */
    { 1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0},
    { 0,  0,  0,  0,  1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0},
    {-3,  3,  0,  0, -2, -1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0},
    { 2, -2,  0,  0,  1,  1,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0},
    { 0,  0,  0,  0,  0,  0,  0,  0,  1,  0,  0,  0,  0,  0,  0,  0},
    { 0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  1,  0,  0,  0},
    { 0,  0,  0,  0,  0,  0,  0,  0, -3,  3,  0,  0, -2, -1,  0,  0},
    { 0,  0,  0,  0,  0,  0,  0,  0,  2, -2,  0,  0,  1,  1,  0,  0},
    {-3,  0,  3,  0,  0,  0,  0,  0, -2,  0, -1,  0,  0,  0,  0,  0},
    { 0,  0,  0,  0, -3,  0,  3,  0,  0,  0,  0,  0, -2,  0, -1,  0},
    { 9, -9, -9,  9,  6,  3, -6, -3,  6, -6,  3, -3,  4,  2,  2,  1},
    {-6,  6,  6, -6, -3, -3,  3,  3, -4,  4, -2,  2, -2, -2, -1, -1},
    { 2,  0, -2,  0,  0,  0,  0,  0,  1,  0,  1,  0,  0,  0,  0,  0},
    { 0,  0,  0,  0,  2,  0, -2,  0,  0,  0,  0,  0,  1,  0,  1,  0},
    {-6,  6,  6, -6, -4, -2,  4,  2, -3,  3, -3,  3, -2, -1, -2, -1},
    { 4, -4, -4,  4,  2,  2, -2, -2,  2, -2,  2, -2,  1,  1,  1,  1}};

#define H3(s, a, i, j, k) ((s)->a[(((size_t)(i) * (s)->ny + (j)) * (s)->nz) + (k)])

int go_h3d_init(go_h3d *s, const double *x, int nx, const double *y, int ny, const double *z, int nz,
                const double *f)
{
    if (nx < 2 || ny < 2 || nz < 3) return -1;
    size_t N = (size_t)nx * ny * nz;
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->accel[0] = s->accel[1] = s->accel[2] = 0;
    s->x = (double *)malloc(sizeof(double) * nx);
    s->y = (double *)malloc(sizeof(double) * ny);
    s->z = (double *)malloc(sizeof(double) * nz);
    s->f = (double *)malloc(sizeof(double) * N);
    s->fs = (double *)malloc(sizeof(double) * N);
    s->dxs = (double *)malloc(sizeof(double) * N);
    s->dys = (double *)malloc(sizeof(double) * N);
    memcpy(s->x, x, sizeof(double) * nx);
    memcpy(s->y, y, sizeof(double) * ny);
    memcpy(s->z, z, sizeof(double) * nz);
    memcpy(s->f, f, sizeof(double) * N);
    double *dfdx = (double *)malloc(sizeof(double) * nz), *dfdy = (double *)malloc(sizeof(double) * nz);
    for (int i = 0; i < nx; i++)                                   /* interpolation.cpp:911-957 */
        for (int j = 0; j < ny; j++)
            natural_slopes(s->z, &H3(s, f, i, j, 0), 1, &H3(s, fs, i, j, 0), 1, nz);
    for (int pass = 0; pass < 2; pass++)                           /* :959-1045 */
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++) {
                int iu = imin(i + 1, nx - 1), id = imax(i - 1, 0), ju = imin(j + 1, ny - 1), jd = imax(j - 1, 0);
                for (int k = 0; k < nz; k++) {
                    dfdx[k] = (H3(s, f, iu, j, k) - H3(s, f, id, j, k)) / (s->x[iu] - s->x[id]);
                    dfdy[k] = (H3(s, f, i, ju, k) - H3(s, f, i, jd, k)) / (s->y[ju] - s->y[jd]);
                }
                if (pass == 0) natural_slopes(s->z, dfdx, 1, &H3(s, dxs, i, j, 0), 1, nz);
                else natural_slopes(s->z, dfdy, 1, &H3(s, dys, i, j, 0), 1, nz);
            }
    free(dfdx);
    free(dfdy);
    return 0;
}

void go_h3d_free(go_h3d *s)
{
    free(s->x); free(s->y); free(s->z); free(s->f); free(s->fs); free(s->dxs); free(s->dys);
    memset(s, 0, sizeof(*s));
}

/* vertical node splines (interpolation.cpp:1071-1189) */
static double node_f(const go_h3d *s, double z, int i, int j, int k, int m)
{
    double d = s->z[k + 1] - s->z[k], X = (z - s->z[k]) / d;
    double f0 = H3(s, f, i, j, k), f1 = H3(s, f, i, j, k + 1), s0 = H3(s, fs, i, j, k), s1 = H3(s, fs, i, j, k + 1);
    if (m == 0) return herm_v(f0, f1, s0, s1, d, X);
    if (m == 1) return herm_d1(f0, f1, s0, s1, d, X);
    return herm_d2(f0, f1, s0, s1, d, X);
}
/* centred horizontal difference splines; axis 0: x, 1: y.  Both read the
 * x-difference slopes (the reference's eval_node_dfdy does the same). */
static double node_dh(const go_h3d *s, double z, int i, int j, int k, int axis, int m)
{
    double g0, g1;
    if (axis == 0) {
        int iu = imin(i + 1, s->nx - 1), id = imax(i - 1, 0);
        g0 = (H3(s, f, iu, j, k) - H3(s, f, id, j, k)) / (s->x[iu] - s->x[id]);
        g1 = (H3(s, f, iu, j, k + 1) - H3(s, f, id, j, k + 1)) / (s->x[iu] - s->x[id]);
    } else {
        int ju = imin(j + 1, s->ny - 1), jd = imax(j - 1, 0);
        g0 = (H3(s, f, i, ju, k) - H3(s, f, i, jd, k)) / (s->y[ju] - s->y[jd]);
        g1 = (H3(s, f, i, ju, k + 1) - H3(s, f, i, jd, k + 1)) / (s->y[ju] - s->y[jd]);
    }
    double d = s->z[k + 1] - s->z[k], X = (z - s->z[k]) / d;
    double s0 = H3(s, dxs, i, j, k), s1 = H3(s, dxs, i, j, k + 1);
    return m == 0 ? herm_v(g0, g1, s0, s1, d, X) : herm_d1(g0, g1, s0, s1, d, X);
}

/* kinds of node quantity that the finite differences act on */
enum { Q_F, Q_FZ, Q_FZZ, Q_FX, Q_FY };
static double node_q(const go_h3d *s, double z, int i, int j, int k, int q)
{
    switch (q) {
    case Q_F: return node_f(s, z, i, j, k, 0);
    case Q_FZ: return node_f(s, z, i, j, k, 1);
    case Q_FZZ: return node_f(s, z, i, j, k, 2);
    case Q_FX: return node_dh(s, z, i, j, k, 0, 0);
    default: return node_dh(s, z, i, j, k, 1, 0);
    }
}
/* d/dx (axis 0) or d/dy (axis 1) of a node quantity by centred differences */
static double fd1(const go_h3d *s, double z, int i, int j, int k, int q, int axis)
{
    if (axis == 0) {
        int iu = imin(i + 1, s->nx - 1), id = imax(i - 1, 0);
        return (node_q(s, z, iu, j, k, q) - node_q(s, z, id, j, k, q)) / (s->x[iu] - s->x[id]);
    }
    int ju = imin(j + 1, s->ny - 1), jd = imax(j - 1, 0);
    return (node_q(s, z, i, ju, k, q) - node_q(s, z, i, jd, k, q)) / (s->y[ju] - s->y[jd]);
}
/* d2/dxdy of a node quantity (interpolation.cpp:1227-1238 form) */
static double fdxy(const go_h3d *s, double z, int i, int j, int k, int q)
{
    int iu = imin(i + 1, s->nx - 1), ju = imin(j + 1, s->ny - 1), id = imax(i - 1, 0), jd = imax(j - 1, 0);
    return (node_q(s, z, iu, ju, k, q) - node_q(s, z, iu, jd, k, q) - node_q(s, z, id, ju, k, q) +
            node_q(s, z, id, jd, k, q)) / ((s->x[iu] - s->x[id]) * (s->y[ju] - s->y[jd]));
}

/* Bicubic patch of node quantity q with x / y slopes from the given
 * difference operators: X_vec of interpolation.cpp:1343-1351.  sx, sy, sxy
 * select how the corner slopes are formed:
 *   sx : 0 = fd1(q, x),  1 = node_q(Q_FX) derivative family (see callers)
 * To keep the reference's exact call pattern the callers pass the corner
 * quantities explicitly. */
typedef double (*corner_fn)(const go_h3d *, double, int, int, int);

static double c_f(const go_h3d *s, double z, int i, int j, int k) { return node_f(s, z, i, j, k, 0); }
static double c_fz(const go_h3d *s, double z, int i, int j, int k) { return node_f(s, z, i, j, k, 1); }
static double c_fzz(const go_h3d *s, double z, int i, int j, int k) { return node_f(s, z, i, j, k, 2); }
static double c_fx(const go_h3d *s, double z, int i, int j, int k) { return node_dh(s, z, i, j, k, 0, 0); }
static double c_fy(const go_h3d *s, double z, int i, int j, int k) { return node_dh(s, z, i, j, k, 1, 0); }
static double c_fd_x(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_F, 0); }
static double c_fd_y(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_F, 1); }
static double c_fd_xy(const go_h3d *s, double z, int i, int j, int k) { return fdxy(s, z, i, j, k, Q_F); }
static double c_fzd_x(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_FZ, 0); }
static double c_fzd_y(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_FZ, 1); }
static double c_fzd_xy(const go_h3d *s, double z, int i, int j, int k) { return fdxy(s, z, i, j, k, Q_FZ); }
static double c_fzzd_x(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_FZZ, 0); }
static double c_fzzd_y(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_FZZ, 1); }
static double c_fzzd_xy(const go_h3d *s, double z, int i, int j, int k) { return fdxy(s, z, i, j, k, Q_FZZ); }
static double c_fxd_x(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_FX, 0); }
static double c_fxd_xy(const go_h3d *s, double z, int i, int j, int k) { return fdxy(s, z, i, j, k, Q_FX); }
static double c_fyd_y(const go_h3d *s, double z, int i, int j, int k) { return fd1(s, z, i, j, k, Q_FY, 1); }
static double c_fyd_xy(const go_h3d *s, double z, int i, int j, int k) { return fdxy(s, z, i, j, k, Q_FY); }

/* A = M * X for the corner set (v, vx*dx, vy*dy, vxy*dx*dy) */
static void patch(const go_h3d *s, double z, int kx, int ky, int kz, double dx, double dy, corner_fn v, corner_fn vx,
                  corner_fn vy, corner_fn vxy, double A[16])
{
    double X[16];
    X[0] = v(s, z, kx, ky, kz);          X[8] = vy(s, z, kx, ky, kz) * dy;
    X[1] = v(s, z, kx + 1, ky, kz);      X[9] = vy(s, z, kx + 1, ky, kz) * dy;
    X[2] = v(s, z, kx, ky + 1, kz);      X[10] = vy(s, z, kx, ky + 1, kz) * dy;
    X[3] = v(s, z, kx + 1, ky + 1, kz);  X[11] = vy(s, z, kx + 1, ky + 1, kz) * dy;
    X[4] = vx(s, z, kx, ky, kz) * dx;          X[12] = vxy(s, z, kx, ky, kz) * dx * dy;
    X[5] = vx(s, z, kx + 1, ky, kz) * dx;      X[13] = vxy(s, z, kx + 1, ky, kz) * dx * dy;
    X[6] = vx(s, z, kx, ky + 1, kz) * dx;      X[14] = vxy(s, z, kx, ky + 1, kz) * dx * dy;
    X[7] = vx(s, z, kx + 1, ky + 1, kz) * dx;  X[15] = vxy(s, z, kx + 1, ky + 1, kz) * dx * dy;
    for (int j = 0; j < 16; j++) {
        A[j] = 0.0;
        for (int k = 0; k < 16; k++) A[j] += BICUBIC[j][k] * X[k];
    }
}

/* polynomial sums: which 0 = value, 1 = d/dx, 2 = d/dy, 3 = d2/dxdy */
static double poly(const double A[16], double xs, double ys, double dx, double dy, int which)
{
    double r = 0.0;
    if (which == 0) {
        for (int k1 = 0; k1 < 4; k1++)
            for (int k2 = 0; k2 < 4; k2++) r += A[k2 * 4 + k1] * pow(xs, k1) * pow(ys, k2);
    } else if (which == 1) {
        for (int k1 = 1; k1 < 4; k1++)
            for (int k2 = 0; k2 < 4; k2++) r += A[k2 * 4 + k1] * pow(xs, k1 - 1) * pow(ys, k2) * (float)(k1) / dx;
    } else if (which == 2) {
        for (int k1 = 0; k1 < 4; k1++)
            for (int k2 = 1; k2 < 4; k2++) r += A[k2 * 4 + k1] * pow(xs, k1) * pow(ys, k2 - 1) * (float)(k2) / dy;
    } else {
        for (int k1 = 1; k1 < 4; k1++)
            for (int k2 = 1; k2 < 4; k2++)
                r += A[k2 * 4 + k1] * pow(xs, k1 - 1) * pow(ys, k2 - 1) * (float)(k1 * k2) / (dx * dy);
    }
    return r;
}

typedef struct { int kx, ky, kz; double dx, dy, xs, ys; } cell;

static cell locate(go_h3d *s, double x, double y, double z)
{
    cell c;
    c.kx = go_find_segment(x, s->x, s->nx, &s->accel[0]);
    c.ky = go_find_segment(y, s->y, s->ny, &s->accel[1]);
    c.kz = go_find_segment(z, s->z, s->nz, &s->accel[2]);
    c.dx = s->x[c.kx + 1] - s->x[c.kx]; c.xs = (x - s->x[c.kx]) / c.dx;
    c.dy = s->y[c.ky + 1] - s->y[c.ky]; c.ys = (y - s->y[c.ky]) / c.dy;
    return c;
}

double go_h3d_f(go_h3d *s, double x, double y, double z)        /* :1334-1369 */
{
    double A[16];
    cell c = locate(s, x, y, z);
    patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_f, c_fd_x, c_fd_y, c_fd_xy, A);
    return poly(A, c.xs, c.ys, c.dx, c.dy, 0);
}

double go_h3d_df(go_h3d *s, double x, double y, double z, int n)   /* :1371-1429 */
{
    double A[16];
    cell c = locate(s, x, y, z);
    if (n == 0 || n == 1) patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_f, c_fd_x, c_fd_y, c_fd_xy, A);
    else patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_fz, c_fzd_x, c_fzd_y, c_fzd_xy, A);
    return poly(A, c.xs, c.ys, c.dx, c.dy, n == 0 ? 1 : (n == 1 ? 2 : 0));
}

double go_h3d_ddf(go_h3d *s, double x, double y, double z, int n1, int n2)   /* :1431-1516 */
{
    double A[16];
    cell c = locate(s, x, y, z);
    int which;
    if (n1 == 0 && n2 == 0) {
        patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_fx, c_fxd_x, c_fd_xy, c_fxd_xy, A);
    } else if (n1 == 1 && n2 == 1) {
        patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_fy, c_fd_xy, c_fyd_y, c_fyd_xy, A);
    } else if (n1 == 2 && n2 == 2) {
        patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_fzz, c_fzzd_x, c_fzzd_y, c_fzzd_xy, A);
    } else if ((n1 == 0 && n2 == 1) || (n1 == 1 && n2 == 0)) {
        patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_f, c_fd_x, c_fd_y, c_fd_xy, A);
    } else {
        patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_fz, c_fzd_x, c_fzd_y, c_fzd_xy, A);
    }
    if (n1 == 2 && n2 == 2) which = 0;
    else if ((n1 == 0 && n2 == 0) || (n1 == 0 && n2 == 2) || (n1 == 2 && n2 == 0)) which = 1;
    else if ((n1 == 1 && n2 == 1) || (n1 == 1 && n2 == 2) || (n1 == 2 && n2 == 1)) which = 2;
    else which = 3;
    return poly(A, c.xs, c.ys, c.dx, c.dy, which);
}

/* eval_all with f, df[3], ddf[6] (:1618-1788).  ddf = {xx, yy, zz, xy, xz, yz}. */
void go_h3d_all(go_h3d *s, double x, double y, double z, double *f, double df[3], double ddf[6])
{
    double A[16];
    double xe = in_iv(x, s->x[0], s->x[s->nx - 1]);
    double ye = in_iv(y, s->y[0], s->y[s->ny - 1]);
    double ze = in_iv(z, s->z[0], s->z[s->nz - 1]);
    cell c;
    c.kx = go_find_segment(xe, s->x, s->nx, &s->accel[0]);
    c.ky = go_find_segment(ye, s->y, s->ny, &s->accel[1]);
    c.kz = go_find_segment(ze, s->z, s->nz, &s->accel[2]);
    c.dx = s->x[c.kx + 1] - s->x[c.kx]; c.xs = (xe - s->x[c.kx]) / c.dx;
    c.dy = s->y[c.ky + 1] - s->y[c.ky]; c.ys = (ye - s->y[c.ky]) / c.dy;
    patch(s, ze, c.kx, c.ky, c.kz, c.dx, c.dy, c_f, c_fd_x, c_fd_y, c_fd_xy, A);
    *f = poly(A, c.xs, c.ys, c.dx, c.dy, 0);
    df[0] = poly(A, c.xs, c.ys, c.dx, c.dy, 1);
    df[1] = poly(A, c.xs, c.ys, c.dx, c.dy, 2);
    ddf[3] = poly(A, c.xs, c.ys, c.dx, c.dy, 3);
    patch(s, ze, c.kx, c.ky, c.kz, c.dx, c.dy, c_fz, c_fzd_x, c_fzd_y, c_fzd_xy, A);
    df[2] = poly(A, c.xs, c.ys, c.dx, c.dy, 0);
    ddf[4] = poly(A, c.xs, c.ys, c.dx, c.dy, 1);
    ddf[5] = poly(A, c.xs, c.ys, c.dx, c.dy, 2);
    patch(s, ze, c.kx, c.ky, c.kz, c.dx, c.dy, c_fx, c_fxd_x, c_fd_xy, c_fxd_xy, A);
    ddf[0] = poly(A, c.xs, c.ys, c.dx, c.dy, 1);
    patch(s, ze, c.kx, c.ky, c.kz, c.dx, c.dy, c_fy, c_fd_xy, c_fyd_y, c_fyd_xy, A);
    ddf[1] = poly(A, c.xs, c.ys, c.dx, c.dy, 2);
    /* the reference evaluates this patch at the unclamped z (:1763-1771) */
    patch(s, z, c.kx, c.ky, c.kz, c.dx, c.dy, c_fzz, c_fzzd_x, c_fzzd_y, c_fzzd_xy, A);
    ddf[2] = poly(A, c.xs, c.ys, c.dx, c.dy, 0);
}

void *go_s2d_new(const double *x, int nx, const double *y, int ny, const double *f)
{
    go_s2d *s = (go_s2d *)calloc(1, sizeof(go_s2d));
    if (go_s2d_init(s, x, nx, y, ny, f) != 0) { free(s); return NULL; }
    return s;
}
void *go_h3d_new(const double *x, int nx, const double *y, int ny, const double *z, int nz, const double *f)
{
    go_h3d *s = (go_h3d *)calloc(1, sizeof(go_h3d));
    if (go_h3d_init(s, x, nx, y, ny, z, nz, f) != 0) { free(s); return NULL; }
    return s;
}
void go_s2d_delete(void *h) { if (h) { go_s2d_free((go_s2d *)h); free(h); } }
void go_h3d_delete(void *h) { if (h) { go_h3d_free((go_h3d *)h); free(h); } }
