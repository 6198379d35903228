// bench2_afft.cpp -- libbench2 adapter: FFTW's benchmark harness driving the
// MI355X afft backend through its C++ wrappers (include/fft.hpp), as the
// reference's misc/libbench2-test/doit.cpp:17-51 drives afft.
//
// What it accepts (can_do): complex FP32 problems of rank >= 1 with power-of-
// two dims, contiguous row-major (libbench2's default strides), no vector
// (batch) dimension, in place or out of place -- the same set the reference
// adapter takes (it asks only for rank >= 1 and a complex problem), minus
// shapes afft itself would not compute (non-power-of-two dims).
//
// libbench2 hands over HOST buffers.  setup() creates the plan (forward for
// sign -1, inverse for +1, interleaved: afft_interleaved) and device buffers;
// doit(iter) uploads p->in, runs the transform iter times on the device and
// downloads the result into p->out, so libbench2's verification reads the
// transform of the input it wrote and its timing covers one upload and one
// download per doit call (amortised over iter); done() releases everything.
//
// Build with FFTW's libbench2 (it provides main(), the timing loop and the
// accuracy check):
//   hipcc -O2 -DAFFT_WITH_LIBBENCH2 -I <fftw-src>/libbench2 -I <fftw-build> -I include \
//         fft_amd/harness/bench2_afft.cpp <fftw-build>/libbench2/libbench2.a -L fft_amd -lafft_mi355x
// Without it (the default build, fft_amd/build.py), bench2_min.h declares the
// interface and the driver at the end of this file runs libbench2's sequence
// (can_do, setup, doit, done) for problems given in libbench2's syntax:
//   afft_bench2 [-n iters] [-w file] <problem>...
//   problem = [i|o]c[f|b]<n>[x<n>...]   e.g. icf1024, ocb64x64
// printing time per transform and MFLOPS (5 N log2 N, libbench2's convention
// for complex transforms); -w writes the output of the last problem (FP32
// interleaved) for an external check.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#ifdef AFFT_WITH_LIBBENCH2
#include "bench-user.h"
#else
#include "bench2_min.h"
#endif
#include "fft.hpp"

BEGIN_BENCH_DOC
BENCH_DOC("name", "afft-mi355x")
BENCH_DOC("version", "0.0")
BENCH_DOC("language", "C++ / HIP")
BENCH_DOC("notes", "afft's C API on one MI355X; host buffers staged through device memory per doit call")
END_BENCH_DOC

namespace {

using Fwd = afft::complex_transform<float, afft_interleaved>;
using Inv = afft::inverse_complex_transform<float, afft_interleaved>;

struct State {
  std::unique_ptr<Fwd> fwd;
  std::unique_ptr<Inv> inv;
  float* d_in = nullptr;
  float* d_out = nullptr;
  size_t n = 0;  // complex elements
};
State g;

bool pow2(int v) { return v >= 1 && (v & (v - 1)) == 0; }

void release() {
  g.fwd.reset();
  g.inv.reset();
  if (g.d_in) (void)hipFree(g.d_in);
  if (g.d_out && g.d_out != g.d_in) (void)hipFree(g.d_out);
  g.d_in = g.d_out = nullptr;
  g.n = 0;
}

}  // namespace

extern "C" {

int can_do(bench_problem* p) {
  if (p->kind != PROBLEM_COMPLEX || !p->sz || p->sz->rnk < 1) return 0;
  if (p->vecsz && p->vecsz->rnk > 0) return 0;
  // contiguous row-major: the last dim has unit stride, each outer dim the
  // product of the inner ones
  int stride = 1;
  for (int i = p->sz->rnk - 1; i >= 0; i--) {
    const bench_iodim& d = p->sz->dims[i];
    if (!pow2(d.n) || d.is != stride || d.os != stride) return 0;
    stride *= d.n;
  }
  return 1;
}

void setup(bench_problem* p) {
  release();
  std::vector<size_t> dim;
  size_t n = 1;
  for (int i = 0; i < p->sz->rnk; i++) {
    dim.push_back((size_t)p->sz->dims[i].n);
    n *= (size_t)p->sz->dims[i].n;
  }
  g.n = n;
  if (p->sign == -1)
    g.fwd = std::make_unique<Fwd>(dim.size(), dim.data());
  else
    g.inv = std::make_unique<Inv>(dim.size(), dim.data());
  (void)hipMalloc(&g.d_in, 2 * n * sizeof(float));
  if (p->in_place)
    g.d_out = g.d_in;
  else
    (void)hipMalloc(&g.d_out, 2 * n * sizeof(float));
}

void doit(int iter, bench_problem* p) {
  const size_t bytes = 2 * g.n * sizeof(float);
  (void)hipMemcpy(g.d_in, p->in, bytes, hipMemcpyHostToDevice);
  for (int i = 0; i < iter; i++) {
    // in place, later iterations transform the previous output, as afft does
    // with libbench2's in-place problems
    if (p->sign == -1)
      (*g.fwd)(g.d_in, nullptr, g.d_out, nullptr);
    else
      (*g.inv)(g.d_in, nullptr, g.d_out, nullptr);
  }
  (void)hipMemcpy(p->out, g.d_out, bytes, hipMemcpyDeviceToHost);
}

void done(bench_problem*) { release(); }

void cleanup(void) { release(); }

void main_init(int*, char***) {}

}  // extern "C"

#ifndef AFFT_WITH_LIBBENCH2
// ---- stand-alone driver (libbench2's problem syntax, timing by wall clock) ----
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

bool parse_problem(const char* s, bench_problem& p, std::vector<bench_iodim>& dims) {
  // [i|o] c [f|b] n[xn...]
  p = bench_problem{};
  p.kind = PROBLEM_COMPLEX;
  p.in_place = 0;
  if (*s == 'i' || *s == 'o') p.in_place = *s++ == 'i';
  if (*s++ != 'c') return false;
  if (*s != 'f' && *s != 'b') return false;
  p.sign = *s++ == 'f' ? -1 : 1;
  dims.clear();
  while (*s) {
    char* end = nullptr;
    long v = strtol(s, &end, 10);
    if (end == s || v < 1) return false;
    dims.push_back({(int)v, 0, 0});
    s = end;
    if (*s == 'x') s++;
    else if (*s) return false;
  }
  int stride = 1;
  for (int i = (int)dims.size() - 1; i >= 0; i--) {
    dims[i].is = dims[i].os = stride;
    stride *= dims[i].n;
  }
  return !dims.empty();
}

}  // namespace

int main(int argc, char** argv) {
  int iters = 10;
  const char* wfile = nullptr;
  main_init(&argc, &argv);
  int rc = 0;
  for (int a = 1; a < argc; a++) {
    if (!strcmp(argv[a], "-n") && a + 1 < argc) {
      iters = atoi(argv[++a]);
      continue;
    }
    if (!strcmp(argv[a], "-w") && a + 1 < argc) {
      wfile = argv[++a];
      continue;
    }
    bench_problem p;
    std::vector<bench_iodim> dims;
    if (!parse_problem(argv[a], p, dims)) {
      fprintf(stderr, "bad problem '%s' (expected [i|o]c[f|b]<n>[x<n>...])\n", argv[a]);
      return 2;
    }
    bench_tensor sz{(int)dims.size(), dims.data()};
    bench_tensor vec{0, nullptr};
    p.sz = &sz;
    p.vecsz = &vec;
    p.pstring = argv[a];
    if (!can_do(&p)) {
      printf("%s: cannot do\n", argv[a]);
      rc = 1;
      continue;
    }
    size_t n = 1;
    for (auto& d : dims) n *= (size_t)d.n;
    std::vector<float> in(2 * n), out(2 * n);
    // the input: x[j] = (u(2j), u(2j+1)) with u a fixed xorshift sequence in [-1, 1)
    unsigned long long st = 88172645463325252ULL;
    for (size_t j = 0; j < 2 * n; j++) {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      in[j] = (float)((double)(st >> 11) / 9007199254740992.0 * 2.0 - 1.0);
    }
    p.in = in.data();
    p.out = p.in_place ? in.data() : out.data();
    setup(&p);
    std::vector<float> saved = in;
    doit(1, &p);  // warm-up (and the output that -w writes)
    std::vector<float> result((float*)p.out, (float*)p.out + 2 * n);
    if (p.in_place) in = saved;
    auto t0 = std::chrono::steady_clock::now();
    doit(iters, &p);
    auto t1 = std::chrono::steady_clock::now();
    done(&p);
    const double s = std::chrono::duration<double>(t1 - t0).count() / iters;
    printf("%s: %.3f us per transform (incl. staging), %.1f MFLOPS\n", argv[a], s * 1e6,
           5.0 * n * std::log2((double)n) / s / 1e6);
    if (wfile) {
      FILE* f = fopen(wfile, "wb");
      if (!f || fwrite(result.data(), sizeof(float), result.size(), f) != result.size()) rc = 1;
      if (f) fclose(f);
    }
  }
  cleanup();
  return rc;
}
#endif
