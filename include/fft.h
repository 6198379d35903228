/*
 * fft.h -- drop-in C API of the MI355X afft backend.
 *
 * Same enums, types, functions and argument order as the reference's fft.h
 * (fft.h:1-88), so code written against afft compiles and links unchanged.
 * What differs, and why:
 *   - every src/dst pointer is a DEVICE pointer (hipMalloc'd or equivalent);
 *     floats must be 4-byte aligned (doubles 8-byte), any other alignment is
 *     accepted;
 *   - `mem` (>= memsize bytes, any alignment, host memory) holds the whole
 *     plan, as in afft (fft_internal.hpp:24-106): the caller frees it when
 *     done and nothing else needs releasing.  The twiddle tables and the
 *     scratch a plan uses on the GPU are shared by all plans of the process
 *     (a per-device cache and a per-device scratch pool that only grows), so
 *     dropping `mem` leaks nothing.  The afft_XX_Y_destroy calls declared at
 *     the end of this header only invalidate a plan;
 *   - `impl` accepts afft_auto and afft_gpu (see afft_gpu.h); the CPU ISA ids
 *     afft_scalar..afft_neon are not provided by this library, so memsize
 *     returns 0 and create returns NULL for them (afft's convention for an
 *     unsupported implementation, fft.cpp:70-72);
 *   - transform runs on the plan's stream (the null stream unless changed with
 *     afft_gpu_set_stream) and returns after the stream has finished, so it is
 *     synchronous like afft's;
 *   - an invalid format gives memsize (size_t)-1 and create NULL
 *     (impl.hpp:38-39,59-60); a dimension that is not a power of two, or a
 *     transform too large for the device, gives memsize 0 and create NULL
 *     (afft does not check, common.hpp:56-66).
 *
 * Semantics kept from afft: row-major dims with the last one contiguous; unit
 * dims are ignored; forward is exp(-2 pi i jk/N) and inverse exp(+...),
 * unnormalised; real transforms of N-D data halve dim 0 (output
 * (N0/2+1) x N1 x ...); interleaved format ignores the *_im pointers.
 */
#ifndef AFFT_MI355X_FFT_H
#define AFFT_MI355X_FFT_H

#include <stddef.h>

enum
{
  afft_auto = 0,
  afft_scalar = 1,
  afft_sse2 = 2,
  afft_avx2 = 3,
  afft_avx512f = 4,
  afft_neon = 5,
};

enum { afft_split = 0, afft_interleaved = 1 };

#ifdef __cplusplus
extern "C"
{
#endif

/* complex forward, FP32: replaces afft_32_c_* (fft.h:22-28, fft.cpp:139-157) */
typedef struct afft_32_c_struct afft_32_c_type;
size_t afft_32_c_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_32_c_type* afft_32_c_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_32_c_transform(afft_32_c_type* state,
  const float* src_re, const float* src_im, float* dst_re, float* dst_im);

/* complex inverse, FP32: replaces afft_32_ci_* (fft.h:30-35, fft.cpp:159-177) */
typedef struct afft_32_ci_struct afft_32_ci_type;
size_t afft_32_ci_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_32_ci_type* afft_32_ci_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_32_ci_transform(afft_32_ci_type* state,
  const float* src_re, const float* src_im, float* dst_re, float* dst_im);

/* real forward, FP32: replaces afft_32_r_* (fft.h:37-42, fft.cpp:179-197) */
typedef struct afft_32_r_struct afft_32_r_type;
size_t afft_32_r_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_32_r_type* afft_32_r_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_32_r_transform(afft_32_r_type* state,
  const float* src, float* dst_re, float* dst_im);

/* real inverse, FP32: replaces afft_32_ri_* (fft.h:44-49, fft.cpp:199-217) */
typedef struct afft_32_ri_struct afft_32_ri_type;
size_t afft_32_ri_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_32_ri_type* afft_32_ri_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_32_ri_transform(afft_32_ri_type* state,
  const float* src_re, const float* src_im, float* dst);

/* replaces afft_32_align_size (fft.h:50, fft.cpp:219): rounds up to 64 bytes */
size_t afft_32_align_size(size_t sz);

/* FP64 counterparts: replace fft.h:52-80 / fft.cpp:221-301 */
typedef struct afft_64_c_struct afft_64_c_type;
size_t afft_64_c_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_64_c_type* afft_64_c_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_64_c_transform(afft_64_c_type* state,
  const double* src_re, const double* src_im, double* dst_re, double* dst_im);

typedef struct afft_64_ci_struct afft_64_ci_type;
size_t afft_64_ci_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_64_ci_type* afft_64_ci_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_64_ci_transform(afft_64_ci_type* state,
  const double* src_re, const double* src_im, double* dst_re, double* dst_im);

typedef struct afft_64_r_struct afft_64_r_type;
size_t afft_64_r_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_64_r_type* afft_64_r_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_64_r_transform(afft_64_r_type* state,
  const double* src, double* dst_re, double* dst_im);

typedef struct afft_64_ri_struct afft_64_ri_type;
size_t afft_64_ri_memsize(size_t ndim, const size_t* dim, int format, int impl);
afft_64_ri_type* afft_64_ri_create(
  size_t ndim, const size_t* dim, int format, void* mem, int impl);
void afft_64_ri_transform(afft_64_ri_type* state,
  const double* src_re, const double* src_im, double* dst);

size_t afft_64_align_size(size_t sz);

/* replaces afft_supported (fft.h:82, fft.cpp:303-306): 1 for afft_auto and
 * afft_gpu when a GPU is visible, 0 otherwise */
int afft_supported(int impl);

/* replaces afft_alignment (fft.h:84, fft.cpp:308): 64 */
extern const size_t afft_alignment;

/* Additions: invalidate a plan (later transform calls on it fail).  Plans
 * own no device memory, so calling these is optional; `mem` stays owned by
 * the caller, as in afft.  Passing NULL is a no-op. */
void afft_32_c_destroy(afft_32_c_type* state);
void afft_32_ci_destroy(afft_32_ci_type* state);
void afft_32_r_destroy(afft_32_r_type* state);
void afft_32_ri_destroy(afft_32_ri_type* state);
void afft_64_c_destroy(afft_64_c_type* state);
void afft_64_ci_destroy(afft_64_ci_type* state);
void afft_64_r_destroy(afft_64_r_type* state);
void afft_64_ri_destroy(afft_64_ri_type* state);

#ifdef __cplusplus
}
#endif

#endif
