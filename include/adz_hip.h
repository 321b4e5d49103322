/* adz_hip.h -- C ABI of libadz_hip.so, the MI355X (gfx950) implementation of GEM's
 * semi-Lagrangian advection step (gemdyn adz_* layer).
 *
 * Every entry point is extern "C" with plain pointers and sizes; device arrays are
 * hipMalloc'd by the caller unless stated otherwise.  Every function returns 0 on success
 * or an ADZ_E* code; adz_last_error() gives the message of the last failure on the calling
 * thread.  Nothing here throws across the boundary.
 *
 * Array conventions follow the reference: Fortran column-major (i fastest), positions in
 * GLOBAL index space with f(1,1,1) at (1.0,1.0,1.0), AoS (px,py,pz) triplets, and source
 * pointers that address f(1,1,1) of each (possibly halo-extended) array.
 */
#ifndef ADZ_HIP_H
#define ADZ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ADZ_OK 0
#define ADZ_EINVAL 1 /* bad argument (null pointer, size, dtype)            */
#define ADZ_EHIP 2   /* a HIP runtime call failed                           */
#define ADZ_ENOMEM 3 /* device allocation failed                            */
#define ADZ_ESTATE 4 /* call out of order (e.g. tracers before trajectory)  */

/* storage types of fields and positions */
#define ADZ_F64 0
#define ADZ_F32 1
#define ADZ_BF16 2

/* arithmetic of the interpolation */
#define ADZ_MATH_REF 0 /* FP64 weights and accumulation in the reference's x86 AVX2+FMA
                          operation order (turbo_tricublin.c:553-615); bit-exact with the
                          reference build for FP32 fields                              */
#define ADZ_MATH_F32 1 /* index/floor step in FP64, weights and accumulation in FP32    */

typedef struct adz_ztab adz_ztab; /* device-resident vertical descriptor (opaque) */
typedef void* adz_stream;         /* a hipStream_t; NULL = the null stream         */

const char* adz_last_error(void);
int adz_version(void);

/* ------------------------------------------------------------------------------------
 * Level 1: the interpolation kernel (replaces turbo_tricublin.c / tricublin_f90.inc).
 * ----------------------------------------------------------------------------------*/

/* Vertical descriptor.  Replaces Vsearch_setup_plus (turbo_tricublin.c:368-380, which
 * wraps Vsearch_setup, :320-363); `levels` is a HOST array of nk doubles.  Unlike the
 * reference (which never frees its malloc'd ztab) the descriptor is released by
 * adz_vsearch_free.  nk >= 4 is required (the reference silently fills 1.0 below that). */
int adz_vsearch_setup_plus(const double* levels, int nk, int ni, int nj, int offseti, int offsetj,
                           adz_ztab** out);
/* Allocated LOCAL index range of the source arrays addressed through this descriptor,
 * e.g. (1-halo, ni_local+halo) when f(1,1,1) sits inside a halo-extended array.  Default:
 * [1, ni] x [1, nj].  Targets whose 4x4 stencil leaves this range (the reference reads
 * out of bounds there, :744-747) produce NaN instead of a memory fault.  No reference
 * counterpart. */
int adz_vsearch_set_bounds(adz_ztab* lv, int ilo, int ihi, int jlo, int jhi);
int adz_vsearch_free(adz_ztab* lv);

/* Index and weights of n targets: Vcoef_pxyz4_inline (turbo_tricublin.c:382-512).
 * Device outputs: ijk[n] (offset of the stencil corner from f(1,1,1)), zlin[n] (the
 * zlinear flag), coef[20*n] (FP64 weights; may be NULL).  For parity checks. */
int adz_vcoef(int* ijk, int* zlin, double* coef, const void* pxyz, int pdtype, const adz_ztab* lv, int n,
              adz_stream stream);

/* Tricublin_zyx1_m_n (turbo_tricublin.c:793-815): d[f*n + p] = field f at target p.
 * fs: HOST array of m device pointers to f(1,1,1) of each source field.  d, fs[] have
 * storage type fdtype; pxyz has pdtype (ADZ_F32 = the reference's pxpypz). */
int adz_tricublin_zyx1_m_n(void* d, const void* const* fs, const void* pxyz, const adz_ztab* lv, int n, int m,
                           int fdtype, int pdtype, int math, adz_stream stream);

/* Tricublin_mono_zyx_m_n (turbo_tricublin.c:1080-1109): cubic result d, linear result l
 * and min / max of the inner cells (ADZ_MATH_REF follows the reference's AVX2 footprint,
 * x 2 columns * y 4 rows * z 2 planes). */
int adz_tricublin_mono_zyx_m_n(void* d, void* l, void* mi, void* ma, const void* const* fs, const void* pxyz,
                               const adz_ztab* lv, int n, int m, int fdtype, int pdtype, int math,
                               adz_stream stream);

/* Tricublin_zyx3_n (turbo_tricublin.c:1298-1312): d(3,n) from the interleaved f3(3,...). */
int adz_tricublin_zyx3_n(void* d, const void* f3, const void* pxyz, const adz_ztab* lv, int n, int fdtype,
                         int pdtype, int math, adz_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* ADZ_HIP_H */
