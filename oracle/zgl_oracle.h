/* CPU restatement of zest's Gauss-Legendre-quadrature transforms, FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library in zest_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product never links or calls it.
 *
 * Conventions follow zest's enums: zernike_norm 0 = normed, 1 = unnormed;
 * sh_norm 0 = geo, 1 = qm; sh_phase -1 = none, +1 = cs; SH layout 0 = LonLat,
 * 1 = LatLon. Coefficients are interleaved {cos, sin} pairs in zest's layouts.
 */
#ifndef ZGL_ORACLE_H
#define ZGL_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

size_t zo_zt_lon_size(size_t order);
size_t zo_zt_lat_size(size_t order);
size_t zo_zt_rad_size(size_t order);
size_t zo_zt_num_coeffs(size_t order);
size_t zo_st_lon_size(int layout, size_t order);
size_t zo_st_lat_size(int layout, size_t order);

void zo_gl_nodes_weights(size_t num_nodes, double *nodes, double *weights);
void zo_gl_packed(size_t num_nodes, double *nodes, double *weights);
void zo_plm_table(int sh_norm, int sh_phase, size_t order, const double *z,
                  size_t nz, double *out);
void zo_radial_table(int zernike_norm, size_t order, const double *r, size_t nr,
                     double *out);
void zo_rdft_forward(size_t n, size_t cols, const double *in, size_t nfreq,
                     double scale, double *out);
void zo_rdft_backward(size_t n, size_t cols, const double *in, size_t nfreq,
                      double *out);

int zo_zt_forward(int zernike_norm, int sh_norm, int sh_phase, size_t grid_order,
                  const double *grid, size_t exp_order, double *coeffs);
int zo_zt_backward(int zernike_norm, int sh_norm, int sh_phase, size_t exp_order,
                   const double *coeffs, size_t grid_order, double *grid);
int zo_st_forward(int sh_norm, int sh_phase, int layout, size_t grid_order,
                  const double *grid, size_t exp_order, double *coeffs);
int zo_st_backward(int sh_norm, int sh_phase, int layout, size_t exp_order,
                   const double *coeffs, size_t grid_order, double *grid);

void zo_zt_evaluate(int zernike_norm, int sh_norm, int sh_phase, size_t order,
                    const double *coeffs, const double *lon, size_t nlon,
                    const double *colat, size_t nlat, const double *r, size_t nrad,
                    double *values);
void zo_st_evaluate(int sh_norm, int sh_phase, size_t order, const double *coeffs,
                    const double *lon, size_t nlon, const double *colat, size_t nlat,
                    double *values);
void zo_st_power_spectrum(int sh_norm, size_t order, const double *a, const double *b,
                          double *out);
void zo_zt_power_spectrum(int sh_norm, size_t order, const double *a, const double *b,
                          double *out);
void zo_wignerd_table(size_t order, double *d);
void zo_rotate(int kind, size_t order, double *coeffs, double alpha, double beta, double gamma,
               int type);
void zo_polar_rotate(int kind, size_t order, double *coeffs, double angle, int type);

#ifdef __cplusplus
}
#endif
#endif
