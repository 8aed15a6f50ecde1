// ref_interp_harness.cpp -- C entry points into the reference's own
// interpolation.cpp (2D multi-stage spline, 3D hybrid spline), built from
// /root/reference by oracle/Makefile into oracle/_ref/libref_interp.so.
// TEST INFRASTRUCTURE ONLY: pins the oracle's restatement (ga_interp.c).
#include <cstring>

#include "util/interpolation.h"

extern "C" {

void *ref_s2d_new(const double *x, int nx, const double *y, int ny, const double *f)
{
    auto *s = new interp::natural_cubic_spline_2D();
    interp::prep(*s, nx, ny);
    for (int i = 0; i < nx; i++) s->x_vals[i] = x[i];
    for (int j = 0; j < ny; j++) s->y_vals[j] = y[j];
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) s->f_vals[i][j] = f[i * ny + j];
    interp::set(*s);
    return s;
}
double ref_s2d_f(void *h, double x, double y) { return interp::eval_f(x, y, *(interp::natural_cubic_spline_2D *)h); }
double ref_s2d_df(void *h, double x, double y, int n) { return interp::eval_df(x, y, n, *(interp::natural_cubic_spline_2D *)h); }
double ref_s2d_ddf(void *h, double x, double y, int n1, int n2)
{
    return interp::eval_ddf(x, y, n1, n2, *(interp::natural_cubic_spline_2D *)h);
}
double ref_s2d_dddf(void *h, double x, double y, int n1, int n2, int n3)
{
    return interp::eval_dddf(x, y, n1, n2, n3, *(interp::natural_cubic_spline_2D *)h);
}
void ref_s2d_delete(void *h)
{
    auto *s = (interp::natural_cubic_spline_2D *)h;
    interp::clear(*s);
    delete s;
}

void *ref_h3d_new(const double *x, int nx, const double *y, int ny, const double *z, int nz, const double *f)
{
    auto *s = new interp::hybrid_spline_3D();
    interp::prep(*s, nx, ny, nz);
    for (int i = 0; i < nx; i++) s->x_vals[i] = x[i];
    for (int j = 0; j < ny; j++) s->y_vals[j] = y[j];
    for (int k = 0; k < nz; k++) s->z_vals[k] = z[k];
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) s->f_vals[i][j][k] = f[((size_t)i * ny + j) * nz + k];
    interp::set(*s);
    return s;
}
double ref_h3d_f(void *h, double x, double y, double z) { return interp::eval_f(x, y, z, *(interp::hybrid_spline_3D *)h); }
double ref_h3d_df(void *h, double x, double y, double z, int n) { return interp::eval_df(x, y, z, n, *(interp::hybrid_spline_3D *)h); }
double ref_h3d_ddf(void *h, double x, double y, double z, int n1, int n2)
{
    return interp::eval_ddf(x, y, z, n1, n2, *(interp::hybrid_spline_3D *)h);
}
void ref_h3d_all(void *h, double x, double y, double z, double out[10])
{
    interp::eval_all(x, y, z, *(interp::hybrid_spline_3D *)h, out[0], out + 1, out + 4);
}
void ref_h3d_delete(void *h)
{
    auto *s = (interp::hybrid_spline_3D *)h;
    interp::clear(*s);
    delete s;
}

}  // extern "C"
