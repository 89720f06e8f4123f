/* smx_setup.h -- host setup helpers of libsmx_fdtd shared between its C and HIP
 * sources (not part of the C ABI). */
#ifndef SMX_SETUP_H
#define SMX_SETUP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
void smx_cell_coeffs(char cls, const double *mu3, const double *eps3, const double sg[6], double iR,
                     double iI, double out[24]);
void smx_cell_sigma(const int L[4], const double *sigma, int64_t x, int64_t y, int64_t z, double sg[6]);
#ifdef __cplusplus
}
#endif
#endif
