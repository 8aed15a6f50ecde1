/* ga_interp.h -- the oracle's restatement of the reference's 2D
 * multi-stage spline and 3D hybrid spline (src/util/interpolation.cpp).
 * TEST INFRASTRUCTURE ONLY. */
#ifndef GA_INTERP_H_
#define GA_INTERP_H_

#ifdef __cplusplus
extern "C" {
#endif

/* f[i * ny + j] = f(x[i], y[j]); fs: slopes of the y-splines */
typedef struct {
    int nx, ny, accel[2];
    double *x, *y, *f, *fs;
} go_s2d;

int    go_s2d_init(go_s2d *s, const double *x, int nx, const double *y, int ny, const double *f);
void   go_s2d_free(go_s2d *s);
double go_s2d_f(go_s2d *s, double x, double y);
double go_s2d_df(go_s2d *s, double x, double y, int n);
double go_s2d_ddf(go_s2d *s, double x, double y, int n1, int n2);
double go_s2d_dddf(go_s2d *s, double x, double y, int n1, int n2, int n3);

/* f[(i * ny + j) * nz + k] = f(x[i], y[j], z[k]); fs, dxs, dys: slopes of
 * the z-splines of f and of its centred x / y differences */
typedef struct {
    int nx, ny, nz, accel[3];
    double *x, *y, *z, *f, *fs, *dxs, *dys;
} go_h3d;

int    go_h3d_init(go_h3d *s, const double *x, int nx, const double *y, int ny, const double *z, int nz,
                   const double *f);
void   go_h3d_free(go_h3d *s);
double go_h3d_f(go_h3d *s, double x, double y, double z);
double go_h3d_df(go_h3d *s, double x, double y, double z, int n);
double go_h3d_ddf(go_h3d *s, double x, double y, double z, int n1, int n2);
/* f, df = {fx, fy, fz}, ddf = {fxx, fyy, fzz, fxy, fxz, fyz} */
void   go_h3d_all(go_h3d *s, double x, double y, double z, double *f, double df[3], double ddf[6]);

/* opaque handles for ctypes */
void  *go_s2d_new(const double *x, int nx, const double *y, int ny, const double *f);
void  *go_h3d_new(const double *x, int nx, const double *y, int ny, const double *z, int nz, const double *f);
void   go_s2d_delete(void *h);
void   go_h3d_delete(void *h);

#ifdef __cplusplus
}
#endif
#endif
