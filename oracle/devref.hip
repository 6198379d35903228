// devref.hip -- FP64 reference DFT on the device.  TEST INFRASTRUCTURE ONLY:
// the parity tests' checker for transforms too large for the CPU oracle; the
// library never links, loads or calls it.
//
// It shares no code with the library.  Each axis is transformed by a plain
// radix-2 Stockham, one launch per stage, with the data flow of afft's own
// checker ReferenceFft (testing.cpp:870-1002) as the CPU oracle restates it
// (afft_oracle.c stockham_1d): in stage `len`, butterfly s = base + q (q < len)
// takes entries s and s + n/2, multiplies the second by tw[q * (n/2)/len]
// (tw[k] = exp(-+2 pi i k / n), computed on the host in long double and
// rounded once), and writes 2*base + q and 2*base + q + len.  Real transforms
// follow the oracle's data flow (afft_oracle.c orc_rfft / orc_irfft): r2c is
// the c2c of the real input restricted to dim-0 rows 0..n0/2; c2r runs the
// inverse c2c of the inner axes on rows 0..n0/2, keeps the real parts of rows
// 0 and n0/2, extends dim 0 Hermitian-symmetrically and runs the inverse c2c
// of dim 0, whose real part is the result.
//
// Error sums (sum |y - ref|^2, sum |ref|^2) are FP64 block partials that the
// caller adds on the host; they do not use the library's error kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

constexpr int kThreads = 256;
char g_err[256];

bool fail(const char* what, hipError_t e) {
  snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
  return false;
}

unsigned grid_for(size_t n) {
  size_t g = (n + kThreads - 1) / kThreads;
  return (unsigned)(g < 65536 * 8 ? g : 65536 * 8);
}

// one radix-2 Stockham stage along an axis of length n (stride `inner`) for
// `outer` independent lines; cur -> nxt
__global__ void stage_kernel(const double2* __restrict__ cur, double2* __restrict__ nxt, const double2* __restrict__ tw,
                             size_t outer, size_t n, size_t inner, size_t len) {
  const size_t half = n / 2;
  const size_t total = outer * half * inner;
  const size_t tstride = half / len;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
    const size_t i = g % inner;
    const size_t r = g / inner;
    const size_t s = r % half;
    const size_t o = r / half;
    const size_t q = s % len, base = s - q;
    const size_t line = o * n * inner + i;
    const double2 a = cur[line + s * inner];
    const double2 b = cur[line + (s + half) * inner];
    const double2 w = tw[q * tstride];
    const double mr = w.x * b.x - w.y * b.y;
    const double mi = w.x * b.y + w.y * b.x;
    nxt[line + (2 * base + q) * inner] = make_double2(a.x + mr, a.y + mi);
    nxt[line + (2 * base + q + len) * inner] = make_double2(a.x - mr, a.y - mi);
  }
}

// float input (interleaved pairs, split arrays, or reals) -> double2
__global__ void load_kernel(const float* __restrict__ re, const float* __restrict__ im, int mode, size_t count,
                           double2* __restrict__ out) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    double x, y;
    if (mode == 0) {
      x = re[2 * e];
      y = re[2 * e + 1];
    } else if (mode == 1) {
      x = re[e];
      y = im[e];
    } else {
      x = re[e];
      y = 0.0;
    }
    out[e] = make_double2(x, y);
  }
}

// half spectrum [b][h+1][inner] (float, interleaved or split) into rows
// 0..h of a [b][n0][inner] double2 array
__global__ void load_half_kernel(const float* __restrict__ re, const float* __restrict__ im, int il, size_t batch,
                                 size_t n0, size_t inner, double2* __restrict__ out) {
  const size_t rows = n0 / 2 + 1;
  const size_t count = batch * rows * inner;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    const size_t i = e % inner;
    const size_t k = (e / inner) % rows;
    const size_t b = e / (inner * rows);
    const double x = il ? re[2 * e] : re[e];
    const double y = il ? re[2 * e + 1] : im[e];
    out[(b * n0 + k) * inner + i] = make_double2(x, y);
  }
}

// rows 0 and n0/2 real; rows k > n0/2 = conj(row n0 - k)
__global__ void extend_kernel(double2* __restrict__ z, size_t batch, size_t n0, size_t inner) {
  const size_t h = n0 / 2;
  const size_t count = batch * n0 * inner;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    const size_t i = e % inner;
    const size_t k = (e / inner) % n0;
    const size_t b = e / (inner * n0);
    if (k == 0 || k == h) {
      z[e].y = 0.0;
    } else if (k > h) {
      const double2 v = z[(b * n0 + (n0 - k)) * inner + i];
      z[e] = make_double2(v.x, -v.y);
    }
  }
}

// sums of |y - ref|^2 and |ref|^2 over y = [b][rows][inner] against
// ref = [b][n0][inner] (rows <= n0); y complex (il pairs / split) or real
// (mode 2: compared with Re ref).  One partial pair per block.
__global__ void err_kernel(const double2* __restrict__ ref, const float* __restrict__ re, const float* __restrict__ im,
                           int mode, size_t batch, size_t n0, size_t rows, size_t inner, double* __restrict__ part) {
  __shared__ double sd[kThreads], sr[kThreads];
  double d = 0.0, r = 0.0;
  const size_t count = batch * rows * inner;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
    const size_t i = e % inner;
    const size_t k = (e / inner) % rows;
    const size_t b = e / (inner * rows);
    const double2 w = ref[(b * n0 + k) * inner + i];
    double x, y;
    if (mode == 0) {
      x = re[2 * e];
      y = re[2 * e + 1];
    } else if (mode == 1) {
      x = re[e];
      y = im[e];
    } else {
      x = re[e];
      y = 0.0;
    }
    const double wy = mode == 2 ? 0.0 : w.y;
    d += (x - w.x) * (x - w.x) + (y - wy) * (y - wy);
    r += w.x * w.x + wy * wy;
  }
  sd[threadIdx.x] = d;
  sr[threadIdx.x] = r;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s /= 2) {
    if ((int)threadIdx.x < s) {
      sd[threadIdx.x] += sd[threadIdx.x + s];
      sr[threadIdx.x] += sr[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sd[0];
    part[2 * blockIdx.x + 1] = sr[0];
  }
}

bool is_pow2(size_t n) { return n && !(n & (n - 1)); }

// the c2c of one axis (length n, `outer` lines of stride `inner`), in place
// in `data` (work: same size)
bool fft_axis(double2* data, double2* work, size_t outer, size_t n, size_t inner, int inverse, hipStream_t st) {
  if (n <= 1) return true;
  const size_t half = n / 2;
  std::vector<double2> h(half);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (size_t k = 0; k < half; k++) {
    const long double ang = two_pi * (long double)k / (long double)n;
    h[k] = make_double2((double)cosl(ang), (double)(inverse ? sinl(ang) : -sinl(ang)));
  }
  double2* tw = nullptr;
  hipError_t e = hipMalloc(&tw, half * sizeof(double2));
  if (e != hipSuccess) return fail("hipMalloc", e);
  e = hipMemcpyAsync(tw, h.data(), half * sizeof(double2), hipMemcpyHostToDevice, st);
  double2 *cur = data, *nxt = work;
  const size_t total = outer * half * inner;
  for (size_t len = 1; len < n && e == hipSuccess; len *= 2) {
    hipLaunchKernelGGL(stage_kernel, dim3(grid_for(total)), dim3(kThreads), 0, st, cur, nxt, tw, outer, n, inner,
                       len);
    e = hipGetLastError();
    double2* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (e == hipSuccess && cur != data)
    e = hipMemcpyAsync(data, cur, outer * n * inner * sizeof(double2), hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(tw);
  return e == hipSuccess ? true : fail("stage", e);
}

bool dims_ok(int ndim, const size_t* dims) {
  if (ndim < 1 || ndim > 16) {
    snprintf(g_err, sizeof g_err, "ndim %d out of range", ndim);
    return false;
  }
  for (int i = 0; i < ndim; i++)
    if (!is_pow2(dims[i])) {
      snprintf(g_err, sizeof g_err, "dim %d (%zu) is not a power of two", i, dims[i]);
      return false;
    }
  return true;
}

}  // namespace

extern "C" {

const char* devref_last_error() { return g_err; }

// data: [batch][dims...] double2 (device), transformed in place over every
// axis of dims; work: the same size.  Returns 0 on success.
int devref_c2c(int ndim, const size_t* dims, size_t batch, void* data, void* work, int inverse, void* stream) {
  if (!dims_ok(ndim, dims)) return -1;
  size_t total = batch;
  for (int i = 0; i < ndim; i++) total *= dims[i];
  hipStream_t st = (hipStream_t)stream;
  for (int ax = 0; ax < ndim; ax++) {
    size_t inner = 1, outer = batch;
    for (int i = ax + 1; i < ndim; i++) inner *= dims[i];
    for (int i = 0; i < ax; i++) outer *= dims[i];
    if (!fft_axis((double2*)data, (double2*)work, outer, dims[ax], inner, inverse, st)) return -1;
  }
  (void)total;
  return 0;
}

// float -> double2: mode 0 interleaved pairs, 1 split (re, im), 2 reals
int devref_load(const float* re, const float* im, int mode, size_t count, void* out, void* stream) {
  hipLaunchKernelGGL(load_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream, re, im, mode, count,
                     (double2*)out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (fail("load", e), -1);
}

// c2r reference: half spectrum [batch][n0/2+1][inner...] (float; il = 1
// interleaved, 0 split) -> out [batch][n0][inner...] double2 whose real parts
// are the c2r result; work: the same size as out.
int devref_c2r(int ndim, const size_t* dims, size_t batch, const float* re, const float* im, int il, void* out,
               void* work, void* stream) {
  if (!dims_ok(ndim, dims)) return -1;
  hipStream_t st = (hipStream_t)stream;
  const size_t n0 = dims[0];
  size_t inner = 1;
  for (int i = 1; i < ndim; i++) inner *= dims[i];
  hipError_t e = hipMemsetAsync(out, 0, batch * n0 * inner * sizeof(double2), st);
  if (e != hipSuccess) return fail("memset", e), -1;
  const size_t cnt = batch * (n0 / 2 + 1) * inner;
  hipLaunchKernelGGL(load_half_kernel, dim3(grid_for(cnt)), dim3(kThreads), 0, st, re, im, il, batch, n0, inner,
                     (double2*)out);
  if ((e = hipGetLastError()) != hipSuccess) return fail("load_half", e), -1;
  // inner axes on every row (rows above n0/2 are zero and overwritten below)
  for (int ax = 1; ax < ndim; ax++) {
    size_t in2 = 1, outer = batch * n0;
    for (int i = ax + 1; i < ndim; i++) in2 *= dims[i];
    for (int i = 1; i < ax; i++) outer *= dims[i];
    if (!fft_axis((double2*)out, (double2*)work, outer, dims[ax], in2, 1, st)) return -1;
  }
  hipLaunchKernelGGL(extend_kernel, dim3(grid_for(batch * n0 * inner)), dim3(kThreads), 0, st, (double2*)out, batch,
                     n0, inner);
  if ((e = hipGetLastError()) != hipSuccess) return fail("extend", e), -1;
  return fft_axis((double2*)out, (double2*)work, batch, n0, inner, 1, st) ? 0 : -1;
}

// Error sums of y against ref: y = [batch][rows][inner] (mode 0 interleaved
// pairs, 1 split, 2 reals against Re ref), ref = [batch][n0][inner] double2.
// sums[0] = sum |y - ref|^2, sums[1] = sum |ref|^2 (host memory; the call
// waits for the stream).
int devref_err(const void* ref, const float* re, const float* im, int mode, size_t batch, size_t n0, size_t rows,
               size_t inner, double* sums, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const size_t count = batch * rows * inner;
  const unsigned g = grid_for(count);
  double* part = nullptr;
  hipError_t e = hipMalloc(&part, 2 * (size_t)g * sizeof(double));
  if (e != hipSuccess) return fail("hipMalloc", e), -1;
  hipLaunchKernelGGL(err_kernel, dim3(g), dim3(kThreads), 0, st, (const double2*)ref, re, im, mode, batch, n0, rows,
                     inner, part);
  e = hipGetLastError();
  std::vector<double> h(2 * (size_t)g);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(part);
  if (e != hipSuccess) return fail("err", e), -1;
  long double d = 0, r = 0;
  for (unsigned b = 0; b < g; b++) {
    d += h[2 * b];
    r += h[2 * b + 1];
  }
  sums[0] = (double)d;
  sums[1] = (double)r;
  return 0;
}

}  // extern "C"
