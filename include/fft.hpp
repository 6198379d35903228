// fft.hpp -- C++ RAII wrappers over include/fft.h for the MI355X backend.
//
// Same class templates as the reference's fft.hpp:47-234
// (afft::complex_transform, inverse_complex_transform, real_transform,
// inverse_real_transform, each over float / double and a format), so code
// written against afft's C++ API compiles unchanged.  Differences:
//   - the buffers passed to operator() are device pointers;
//   - the destructor also calls the matching afft_XX_Y_destroy, which only
//     invalidates the plan (plans own no device memory; their twiddle tables
//     and scratch are shared by the process);
//   - a plan can be moved to another HIP stream with set_stream().
#ifndef AFFT_MI355X_FFT_HPP
#define AFFT_MI355X_FFT_HPP

#include <stddef.h>

#include "afft_gpu.h"
#include "fft.h"

namespace afft
{

namespace detail
{
template <typename AfftType>
class owner
{
public:
  typedef size_t (*memsize_fn)(size_t, const size_t*, int, int);
  typedef AfftType* (*create_fn)(size_t, const size_t*, int, void*, int);
  typedef void (*destroy_fn)(AfftType*);

  owner() = default;
  owner(const owner&) = delete;
  owner& operator=(const owner&) = delete;

  owner(memsize_fn memsize, create_fn create, destroy_fn destroy, size_t ndim, const size_t* dim, int format,
        void* mem, int impl)
  : destroy_(destroy)
  {
    if(mem == nullptr)
    {
      size_t size = memsize(ndim, dim, format, impl);
      if(size == 0 || size == (size_t) -1) return;
      allocated_ = new char[size];
      mem = allocated_;
    }
    state_ = create(ndim, dim, format, mem, impl);
  }

  owner(owner&& o) noexcept : allocated_(o.allocated_), state_(o.state_), destroy_(o.destroy_)
  {
    o.allocated_ = nullptr;
    o.state_ = nullptr;
  }

  ~owner()
  {
    if(state_ && destroy_) destroy_(state_);
    delete[] allocated_;
  }

  AfftType* state() const { return state_; }

private:
  char* allocated_ = nullptr;
  AfftType* state_ = nullptr;
  destroy_fn destroy_ = nullptr;
};
}  // namespace detail

#define AFFT_MI355X_WRAPPER(NAME, T, P, K, CALL_ARGS, CALL_PARAMS)                                         \
  template <int format>                                                                                    \
  class NAME<T, format>                                                                                    \
  {                                                                                                        \
  public:                                                                                                  \
    NAME() = default;                                                                                      \
    NAME(size_t ndim, const size_t* dim, void* mem = nullptr, int impl = afft_auto)                        \
    : m_(afft_##P##_##K##_memsize, afft_##P##_##K##_create, afft_##P##_##K##_destroy, ndim, dim, format,   \
         mem, impl)                                                                                        \
    {                                                                                                      \
    }                                                                                                      \
    void operator() CALL_PARAMS { afft_##P##_##K##_transform CALL_ARGS; }                                 \
    explicit operator bool() const { return m_.state() != nullptr; }                                      \
    int set_stream(void* stream) { return afft_gpu_set_stream(m_.state(), stream); }                      \
                                                                                                           \
  private:                                                                                                 \
    detail::owner<afft_##P##_##K##_type> m_;                                                               \
  };

template <typename T, int format = afft_split>
class complex_transform;
template <typename T, int format = afft_split>
class inverse_complex_transform;
template <typename T, int format = afft_split>
class real_transform;
template <typename T, int format = afft_split>
class inverse_real_transform;

AFFT_MI355X_WRAPPER(complex_transform, float, 32, c, (m_.state(), src_re, src_im, dst_re, dst_im),
                    (const float* src_re, const float* src_im, float* dst_re, float* dst_im))
AFFT_MI355X_WRAPPER(complex_transform, double, 64, c, (m_.state(), src_re, src_im, dst_re, dst_im),
                    (const double* src_re, const double* src_im, double* dst_re, double* dst_im))
AFFT_MI355X_WRAPPER(inverse_complex_transform, float, 32, ci, (m_.state(), src_re, src_im, dst_re, dst_im),
                    (const float* src_re, const float* src_im, float* dst_re, float* dst_im))
AFFT_MI355X_WRAPPER(inverse_complex_transform, double, 64, ci, (m_.state(), src_re, src_im, dst_re, dst_im),
                    (const double* src_re, const double* src_im, double* dst_re, double* dst_im))
AFFT_MI355X_WRAPPER(real_transform, float, 32, r, (m_.state(), src, dst_re, dst_im),
                    (const float* src, float* dst_re, float* dst_im))
AFFT_MI355X_WRAPPER(real_transform, double, 64, r, (m_.state(), src, dst_re, dst_im),
                    (const double* src, double* dst_re, double* dst_im))
AFFT_MI355X_WRAPPER(inverse_real_transform, float, 32, ri, (m_.state(), src_re, src_im, dst),
                    (const float* src_re, const float* src_im, float* dst))
AFFT_MI355X_WRAPPER(inverse_real_transform, double, 64, ri, (m_.state(), src_re, src_im, dst),
                    (const double* src_re, const double* src_im, double* dst))

#undef AFFT_MI355X_WRAPPER

}  // namespace afft

#endif
