// afft_run -- command-line test / benchmark harness over the C ABI, with the
// option set of the reference's `run` and `run_all` (testing.cpp:1190-1350,
// run_all.cpp:149-268):
//
//   afft_run <fft|fft-interleaved> <log2 size or a-b range>... [-d] [-r] [-i]
//            [-b] [-p precision] [-n ops] [--misaligned-by bytes]
//   afft_run --all [fft|fft-interleaved] [-t elements] [-s seed]
//
// Test mode compares the transform with the FP64 plan of the same shape on
// the device and prints the reference harness's error measure
// sqrt(sum (d1-d0)^2 / sum (d0^2 + d1^2)) (testing.cpp:1113-1127); with -p the
// run fails when the error exceeds p*sqrt(log2 N) (testing.cpp:1128,1321-1331).
// Inputs follow testing.cpp:1064-1096: std::mt19937 with the default seed,
// uniform_real_distribution<float>(-1, 1), a re block then an im block; real
// forward transforms zero the im block; real inverse inputs are made
// Hermitian by mirroring.  Bench mode runs the transform on all-zero input
// until about -n flops are done (5 N log2 N complex, 2.5 N log2 N real,
// testing.cpp:1014-1048) and prints GFLOPS and ns per element.  --all runs
// every 1-D size 2^1..2^22, square 2-D and cubic 3-D shapes and random N-D
// shapes through c / ci / r / ri at p = 1.2e-7 (run_all.cpp:58-68,181-255).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "afft_gpu.h"
#include "fft.h"

namespace {

struct Range {
  int begin, end;
};

struct Options {
  bool is_double = false, is_real = false, is_inverse = false, is_bench = false;
  double precision = 0.0;
  bool has_precision = false;
  double num_ops = 1e11;
  long misaligned_by = 0;
  int format = afft_split;
  std::vector<Range> size;
};

bool parse_range(const char* s, Range& r) {
  char* end = nullptr;
  long a = strtol(s, &end, 0);
  if (end == s) return false;
  if (*end == 0) {
    r = {(int)a, (int)a + 1};
    return true;
  }
  if (*end != '-') return false;
  const char* p = end + 1;
  long b = strtol(p, &end, 0);
  if (end == p || *end != 0) return false;
  r = {(int)a, (int)b};
  return true;
}

void usage() {
  fprintf(stderr,
          "usage: afft_run <fft|fft-interleaved> <log2 size|a-b>... [-d] [-r] [-i] [-b] [-p precision] "
          "[-n ops] [--misaligned-by bytes]\n"
          "       afft_run --all [fft|fft-interleaved] [-t elements] [-s seed]\n");
}

#define CHECK_HIP(x)                                                          \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      exit(2);                                                                \
    }                                                                         \
  } while (0)

size_t product(const std::vector<size_t>& v) {
  size_t r = 1;
  for (size_t e : v) r *= e;
  return r;
}

// Row-major complex array of `shape` as re block + im block (host).
struct HostComplex {
  std::vector<double> re, im;
};

size_t mirror_idx(size_t size, size_t i) { return (size - 1) & (size - i); }

// Hermitian mirroring of testing.cpp:1069-1094.
void make_hermitian(const std::vector<size_t>& shape, std::vector<double>& re, std::vector<double>& im) {
  std::vector<double> tr = re, ti = im;
  size_t n = product(shape);
  std::vector<size_t> idx(shape.size()), mid(shape.size());
  for (size_t lin = 0; lin < n; lin++) {
    size_t rem = lin;
    for (size_t d = shape.size(); d-- > 0;) {
      idx[d] = rem % shape[d];
      rem /= shape[d];
    }
    size_t ml = 0;
    for (size_t d = 0; d < shape.size(); d++) ml = ml * shape[d] + mirror_idx(shape[d], idx[d]);
    re[lin] = tr[lin] + tr[ml];
    im[lin] = ti[ml] - ti[lin];
  }
}

// One GPU run of kind (0 c, 1 ci, 2 r, 3 ri) in precision P on host data in
// the canonical layout (complex: full shape; real side: the halved shape for
// r output / ri input), returned in the same canonical layout.
template <typename T>
void run_gpu(int kind, const std::vector<size_t>& shape, int format, long misalign, const std::vector<double>& in_re,
             const std::vector<double>& in_im, std::vector<double>& out_re, std::vector<double>& out_im,
             void* plan_mem, int iters_for_bench, double* seconds) {
  const int prec = sizeof(T) == 4 ? 32 : 64;
  const size_t n = product(shape);
  std::vector<size_t> half = shape;
  // the halved dim is the first non-unit one (HalvedDimFirst, testing.cpp:530-538)
  size_t hd = 0;
  while (hd < half.size() && half[hd] == 1) hd++;
  if (hd == half.size()) hd = 0;
  half[hd] = half[hd] / 2 + 1;
  const size_t nh = product(half);
  const bool real = kind >= 2;
  const size_t nin = kind == 3 ? nh : n, nout = kind == 2 ? nh : n;
  const bool cin = kind != 2, cout = kind != 3;
  const size_t pad = (size_t)misalign;
  // split real-side arrays are placed at afft_XX_align_size offsets (testing.cpp:548-559)
  auto align = [&](size_t k) { return prec == 32 ? afft_32_align_size(k) : afft_64_align_size(k); };
  const size_t in_im_off = (kind == 3) ? align(nh) : nin;
  const size_t out_im_off = (kind == 2) ? align(nh) : nout;
  const size_t in_elems = cin ? (format == afft_interleaved ? 2 * nin : nin + in_im_off) : nin;
  const size_t out_elems = cout ? (format == afft_interleaved ? 2 * nout : nout + out_im_off) : nout;
  std::vector<T> hin(in_elems, T(0));
  if (cin) {
    for (size_t i = 0; i < nin; i++) {
      if (format == afft_interleaved) {
        hin[2 * i] = (T)in_re[i];
        hin[2 * i + 1] = (T)in_im[i];
      } else {
        hin[i] = (T)in_re[i];
        hin[in_im_off + i] = (T)in_im[i];
      }
    }
  } else {
    for (size_t i = 0; i < nin; i++) hin[i] = (T)in_re[i];
  }
  char *din = nullptr, *dout = nullptr;
  CHECK_HIP(hipMalloc(&din, in_elems * sizeof(T) + pad));
  CHECK_HIP(hipMalloc(&dout, out_elems * sizeof(T) + pad));
  T* src = (T*)(din + pad);
  T* dst = (T*)(dout + pad);
  CHECK_HIP(hipMemcpy(src, hin.data(), in_elems * sizeof(T), hipMemcpyHostToDevice));
  std::vector<size_t> dims = shape;
  afft_gpu_plan* p = afft_gpu_plan_create(prec, kind, dims.size(), dims.data(), format, 1, plan_mem);
  if (!p) {
    fprintf(stderr, "plan creation failed: %s\n", afft_gpu_last_error());
    exit(2);
  }
  const T* s_re = src;
  const T* s_im = cin && format == afft_split ? src + in_im_off : nullptr;
  T* d_re = dst;
  T* d_im = cout && format == afft_split ? dst + out_im_off : nullptr;
  if (afft_gpu_execute(p, s_re, s_im, d_re, d_im, nullptr) != 0) {
    fprintf(stderr, "execute failed: %s\n", afft_gpu_last_error());
    exit(2);
  }
  CHECK_HIP(hipDeviceSynchronize());
  if (seconds) {
    auto t0 = std::chrono::high_resolution_clock::now();
    for (int i = 0; i < iters_for_bench; i++) afft_gpu_execute(p, s_re, s_im, d_re, d_im, nullptr);
    CHECK_HIP(hipDeviceSynchronize());
    auto t1 = std::chrono::high_resolution_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  std::vector<T> hout(out_elems);
  CHECK_HIP(hipMemcpy(hout.data(), dst, out_elems * sizeof(T), hipMemcpyDeviceToHost));
  out_re.assign(nout, 0.0);
  out_im.assign(nout, 0.0);
  for (size_t i = 0; i < nout; i++) {
    if (!cout) {
      out_re[i] = hout[i];
    } else if (format == afft_interleaved) {
      out_re[i] = hout[2 * i];
      out_im[i] = hout[2 * i + 1];
    } else {
      out_re[i] = hout[i];
      out_im[i] = hout[out_im_off + i];
    }
  }
  afft_gpu_plan_destroy(p);
  CHECK_HIP(hipFree(din));
  CHECK_HIP(hipFree(dout));
  (void)real;
}

struct Result {
  double error = 0.0, error_factor = 0.0, flops = 0.0, ns_per_element = 0.0;
};

Result test_one(const Options& o, const std::vector<int>& lsz, void* mem) {
  std::vector<size_t> shape;
  for (int e : lsz) shape.push_back((size_t)1 << e);
  const size_t n = product(shape);
  const int kind = (o.is_real ? 2 : 0) + (o.is_inverse ? 1 : 0);
  Result r;
  if (o.is_bench) {
    std::vector<double> zr(n, 0.0), zi(n, 0.0), yr, yi;
    const double per = (o.is_real ? 2.5 : 5.0) * (double)n * std::log2((double)std::max<size_t>(n, 2));
    const int iters = (int)std::max(1.0, o.num_ops / per);
    double sec = 0;
    if (o.is_double)
      run_gpu<double>(kind, shape, o.format, o.misaligned_by, zr, zi, yr, yi, mem, iters, &sec);
    else
      run_gpu<float>(kind, shape, o.format, o.misaligned_by, zr, zi, yr, yi, mem, iters, &sec);
    r.flops = per * iters / sec;
    r.ns_per_element = sec / iters / (double)n * 1e9;
    return r;
  }
  // inputs: mt19937 default seed, uniform(-1, 1) floats, re block then im block
  std::mt19937 mt;
  std::uniform_real_distribution<float> dist(-1.0f, 1.0f);
  std::vector<double> re(n), im(n);
  for (size_t i = 0; i < n; i++) re[i] = dist(mt);
  for (size_t i = 0; i < n; i++) im[i] = dist(mt);
  if (kind == 2) std::fill(im.begin(), im.end(), 0.0);
  if (kind == 3) make_hermitian(shape, re, im);
  // the real-side complex arrays hold the halved shape (first non-unit dim)
  std::vector<double> in_re = re, in_im = im;
  if (kind == 3) {
    std::vector<size_t> half = shape;
    size_t hd = 0;
    while (hd < half.size() && half[hd] == 1) hd++;
    if (hd == half.size()) hd = 0;
    half[hd] = half[hd] / 2 + 1;
    const size_t nh = product(half);
    in_re.assign(nh, 0.0);
    in_im.assign(nh, 0.0);
    std::vector<size_t> idx(shape.size());
    for (size_t lin = 0; lin < nh; lin++) {
      size_t rem = lin;
      for (size_t d = half.size(); d-- > 0;) {
        idx[d] = rem % half[d];
        rem /= half[d];
      }
      size_t full = 0;
      for (size_t d = 0; d < shape.size(); d++) full = full * shape[d] + idx[d];
      in_re[lin] = re[full];
      in_im[lin] = im[full];
    }
  }
  std::vector<double> y32r, y32i, y64r, y64i;
  if (o.is_double) {
    run_gpu<double>(kind, shape, o.format, o.misaligned_by, in_re, in_im, y32r, y32i, mem, 0, nullptr);
  } else {
    run_gpu<float>(kind, shape, o.format, o.misaligned_by, in_re, in_im, y32r, y32i, mem, 0, nullptr);
  }
  run_gpu<double>(kind, shape, afft_interleaved, 0, in_re, in_im, y64r, y64i, mem, 0, nullptr);
  double sum = 0, diff = 0;
  for (size_t i = 0; i < y64r.size(); i++) {
    sum += y64r[i] * y64r[i] + y32r[i] * y32r[i] + y64i[i] * y64i[i] + y32i[i] * y32i[i];
    diff += (y32r[i] - y64r[i]) * (y32r[i] - y64r[i]) + (y32i[i] - y64i[i]) * (y32i[i] - y64i[i]);
  }
  r.error = sum > 0 ? std::sqrt(diff / sum) : 0.0;
  r.error_factor = std::sqrt(std::log2((double)std::max<size_t>(n, 1)));
  return r;
}

bool run_test(const Options& o, void* mem) {
  std::vector<int> sz;
  for (const Range& r : o.size) sz.push_back(r.begin);
  for (;;) {
    for (int e : sz) printf("%2d ", e);
    fflush(stdout);
    Result res = test_one(o, sz, mem);
    if (o.is_bench) {
      printf("%f GFLOPS  %f ns\n", res.flops * 1e-9, res.ns_per_element);
    } else {
      printf("%g\n", res.error);
      if (o.has_precision && res.error > o.precision * res.error_factor) {
        printf("Error exceeds maximal allowed value, which is %g.\n", o.precision * res.error_factor);
        return false;
      }
    }
    bool done = true;
    for (size_t i = sz.size(); i-- > 0;) {
      if (sz[i] + 1 < o.size[i].end) {
        sz[i]++;
        done = false;
        break;
      }
      sz[i] = o.size[i].begin;
    }
    if (done) break;
  }
  return true;
}

// run_all: every 1-D size, square 2-D / cubic 3-D, then random N-D shapes
// (each (ndim, total) class equally likely) until `elements` are processed.
int run_all(int format, double elements, unsigned seed, void* mem) {
  std::vector<std::vector<int>> shapes;
  for (int lg = 1; lg <= 22; lg++) shapes.push_back({lg});
  for (int lg = 1; lg <= 11; lg++) shapes.push_back({lg, lg});
  for (int lg = 1; lg <= 7; lg++) shapes.push_back({lg, lg, lg});
  std::mt19937 rng(seed);
  double done = 0;
  for (const auto& s : shapes) done += std::pow(2.0, [&] {
    int t = 0;
    for (int e : s) t += e;
    return t;
  }());
  while (done < elements) {
    int ndim = std::uniform_int_distribution<int>(1, 5)(rng);
    int total = std::uniform_int_distribution<int>(ndim, 22)(rng);
    std::vector<int> s(ndim, 1);
    for (int i = ndim; i < total; i++) s[std::uniform_int_distribution<int>(0, ndim - 1)(rng)]++;
    shapes.push_back(s);
    done += std::pow(2.0, total);
  }
  int failures = 0;
  for (const auto& s : shapes) {
    for (int kind = 3; kind >= 0; kind--) {
      Options o;
      o.format = format;
      o.is_real = kind >= 2;
      o.is_inverse = kind & 1;
      o.has_precision = true;
      o.precision = 1.2e-7;
      for (int e : s) o.size.push_back({e, e + 1});
      Result res = test_one(o, s, mem);
      const bool ok = res.error <= o.precision * res.error_factor;
      if (!ok) failures++;
      fprintf(ok ? stdout : stderr, "%s %s", ok ? "ok  " : "FAIL", kind == 0 ? "c " : kind == 1 ? "ci" : kind == 2 ? "r " : "ri");
      for (int e : s) fprintf(ok ? stdout : stderr, " %d", e);
      fprintf(ok ? stdout : stderr, "  %g\n", res.error);
    }
  }
  printf("%zu shapes, %d failures\n", shapes.size(), failures);
  return failures ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<char> mem(afft_gpu_plan_memsize());
  if (argc >= 2 && !strcmp(argv[1], "--all")) {
    int format = afft_split;
    double elements = 1 << 26;
    unsigned seed = 5489u;
    for (int i = 2; i < argc; i++) {
      if (!strcmp(argv[i], "fft-interleaved")) format = afft_interleaved;
      else if (!strcmp(argv[i], "fft")) format = afft_split;
      else if (!strcmp(argv[i], "-t") && i + 1 < argc) elements = atof(argv[++i]);
      else if (!strcmp(argv[i], "-s") && i + 1 < argc) seed = (unsigned)atol(argv[++i]);
      else {
        usage();
        return 2;
      }
    }
    return run_all(format, elements, seed, mem.data());
  }
  if (argc < 3) {
    usage();
    return 2;
  }
  Options o;
  if (!strcmp(argv[1], "fft")) o.format = afft_split;
  else if (!strcmp(argv[1], "fft-interleaved")) o.format = afft_interleaved;
  else {
    usage();
    return 2;
  }
  for (int i = 2; i < argc; i++) {
    const char* a = argv[i];
    if (!strcmp(a, "-d")) o.is_double = true;
    else if (!strcmp(a, "-r")) o.is_real = true;
    else if (!strcmp(a, "-i")) o.is_inverse = true;
    else if (!strcmp(a, "-b")) o.is_bench = true;
    else if (!strcmp(a, "-p") && i + 1 < argc) {
      o.precision = atof(argv[++i]);
      o.has_precision = true;
    } else if (!strcmp(a, "-n") && i + 1 < argc) o.num_ops = atof(argv[++i]);
    else if (!strcmp(a, "--misaligned-by") && i + 1 < argc) {
      o.misaligned_by = atol(argv[++i]);
    } else {
      Range r;
      if (!parse_range(a, r) || r.begin < 0 || r.end <= r.begin) {
        usage();
        return 2;
      }
      o.size.push_back(r);
    }
  }
  if (o.size.empty()) {
    usage();
    return 2;
  }
  // buffers must keep element alignment (afft accepts any byte offset; device
  // loads of a float need 4-byte, of a double 8-byte alignment)
  const long el = o.is_double ? 8 : 4;
  if (o.misaligned_by < 0 || o.misaligned_by % el != 0) {
    fprintf(stderr, "--misaligned-by must be a non-negative multiple of %ld\n", el);
    return 2;
  }
  return run_test(o, mem.data()) ? 0 : 1;
}
