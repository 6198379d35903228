/*
 * afft_oracle.c -- CPU oracle for the afft transform hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (fft_amd/) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as a checker.
 *
 * What it restates (reference = jkrempus/fft, "afft"):
 *   - orc_fft:    FP64 radix-2 Stockham c2c applied dimension by dimension,
 *                 following ReferenceFft (testing.cpp:870-1002): twiddles
 *                 cos/sin(-+ i*pi/(n/2)) (testing.cpp:912-920), Stockham stage
 *                 loop (testing.cpp:923-951), per-dim application
 *                 (testing.cpp:965-1001).  Unit dims are dropped first, as
 *                 remove_redundant_dimensions does (common.hpp:68-84).
 *                 Forward is exp(-2 pi i jk/N); inverse is exp(+...) and
 *                 unnormalised (fft_internal.hpp:166-198, common.hpp:355-378).
 *   - orc_rfft:   real forward transform with afft's data flow:
 *                 1-D: N reals read as N/2 interleaved complex
 *                 (onedim.hpp:1175-1189), N/2-point c2c, then the forward
 *                 real pass (onedim.hpp:1031-1108, inverse=false) giving bins
 *                 0..N/2.
 *                 N-D: dim 0 is halved (fft_internal.hpp:200-346): rows 2j and
 *                 2j+1 of dim 0 form complex row j (multi.hpp:261-287), an
 *                 N0/2-point c2c runs along dim 0, the row real pass
 *                 (multi.hpp:343-414, inverse=false) gives N0/2+1 rows, then
 *                 each row gets the (ndim-1)-D c2c.
 *   - orc_irfft:  real inverse transform, afft's data flow:
 *                 1-D: inverse real pre-pass (onedim.hpp:1031-1108,
 *                 inverse=true; DC bin uses only the real parts of bins 0 and
 *                 N/2, bin N/4 is 2*conj), then N/2-point inverse c2c whose
 *                 complex output is read as N interleaved reals
 *                 (onedim.hpp:1191-1254).
 *                 N-D: the (ndim-1)-D inverse c2c on each of the N0/2+1 input
 *                 rows, then the inverse row real pass (multi.hpp:343-414,
 *                 inverse=true; rows 0 and N0/2 contribute only their real
 *                 parts), then the N0/2-point inverse c2c along dim 0 whose
 *                 complex row j becomes real rows 2j and 2j+1
 *                 (fft_internal.hpp:472-493).  ri(r(x)) = N*x.
 *   - orc_dft_bruteforce: O(N^2) long-double DFT, the textbook definition, used
 *                 only to pin orc_fft.
 *   - orc_rel_l2 / orc_harness_error: the error metrics.  rel-L2 is
 *                 ||y-ref|| / ||ref||; the harness metric restates
 *                 testing.cpp:1113-1127, sqrt(sum(d1-d0)^2 / sum(d0^2+d1^2)).
 *   - orc_uniform / orc_normal: portable counter-based input generators (the
 *                 reference uses std::mt19937, testing.cpp:1064-1067, whose
 *                 distribution output is not portable across C++ libraries).
 *
 * All complex arrays here are interleaved (re, im) doubles, row-major with the
 * last dim contiguous.  Layout packing (split vs interleaved, the split
 * imaginary offset) is done by the callers.
 *
 * Pinning (see tests/test_oracle.py): orc_fft against orc_dft_bruteforce,
 * closed forms (impulse, constant, tone), linearity and Parseval; orc_rfft and
 * orc_irfft against the full c2c of the (Hermitian-extended) data; all three
 * against the real afft build in oracle/_ref (golden fixtures in tests/golden)
 * within 2*log2(N)*2^-24 relative L2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXDIM 64

static size_t squeeze_dims(size_t ndim, const size_t* dim, size_t* out)
{
  /* remove_redundant_dimensions (common.hpp:68-84) */
  size_t n = 0;
  for(size_t i = 0; i < ndim; i++)
    if(dim[i] != 1) out[n++] = dim[i];
  if(n == 0) out[n++] = 1;
  return n;
}

static size_t prod(const size_t* b, const size_t* e)
{
  size_t r = 1;
  for(; b < e; b++) r *= *b;
  return r;
}

/* One 1-D FP64 Stockham radix-2 transform of n points, in place on
 * interleaved data with element stride `es` (in complex elements).
 * Follows ReferenceFft::Onedim::transform (testing.cpp:906-951). */
static void stockham_1d(size_t n, double* data, size_t es, int inverse,
                        double* work, const double* tw)
{
  if(n <= 1) return;
  double* cur = work;          /* 2n doubles */
  double* nxt = work + 2 * n;  /* 2n doubles */
  for(size_t i = 0; i < n; i++)
  {
    cur[2 * i] = data[2 * i * es];
    cur[2 * i + 1] = data[2 * i * es + 1];
  }
  size_t half = n / 2;
  for(size_t len = 1; len < n; len *= 2)
  {
    size_t tstride = half / len;
    for(size_t base = 0; base < half; base += len)
    {
      for(size_t q = 0; q < len; q++)
      {
        size_t s = base + q;
        const double* a = cur + 2 * s;
        const double* b = cur + 2 * (s + half);
        double wr = tw[2 * (q * tstride)];
        double wi = tw[2 * (q * tstride) + 1];
        double mr = wr * b[0] - wi * b[1];
        double mi = wr * b[1] + wi * b[0];
        double* d0 = nxt + 2 * (2 * base + q);
        double* d1 = nxt + 2 * (2 * base + q + len);
        d0[0] = a[0] + mr; d0[1] = a[1] + mi;
        d1[0] = a[0] - mr; d1[1] = a[1] - mi;
      }
    }
    double* t = cur; cur = nxt; nxt = t;
  }
  for(size_t i = 0; i < n; i++)
  {
    data[2 * i * es] = cur[2 * i];
    data[2 * i * es + 1] = cur[2 * i + 1];
  }
}

static double* make_twiddles(size_t n, int inverse)
{
  /* twiddle[i] = exp(-+ i*pi*i/(n/2)), i < n/2 (testing.cpp:912-920) */
  size_t h = n / 2 ? n / 2 : 1;
  double* tw = (double*) malloc(sizeof(double) * 2 * h);
  const double pi = acos(-1.0);
  for(size_t i = 0; i < h; i++)
  {
    double phi = (inverse ? 1.0 : -1.0) * (double) i * pi / (double) h;
    tw[2 * i] = cos(phi);
    tw[2 * i + 1] = sin(phi);
  }
  return tw;
}

/* c2c along axis `ax` of an interleaved array of shape dim[0..ndim). */
static void fft_axis(size_t ndim, const size_t* dim, size_t ax, double* data,
                     int inverse)
{
  size_t n = dim[ax];
  if(n <= 1) return;
  size_t inner = prod(dim + ax + 1, dim + ndim);
  size_t outer = prod(dim, dim + ax);
  double* work = (double*) malloc(sizeof(double) * 4 * n);
  double* tw = make_twiddles(n, inverse);
  for(size_t o = 0; o < outer; o++)
    for(size_t i = 0; i < inner; i++)
      stockham_1d(n, data + 2 * (o * n * inner + i), inner, inverse, work, tw);
  free(work);
  free(tw);
}

/* Full c2c.  src and dst may alias. */
void orc_fft(size_t ndim_in, const size_t* dim_in, const double* src,
             double* dst, int inverse)
{
  size_t dim[ORC_MAXDIM];
  size_t ndim = squeeze_dims(ndim_in, dim_in, dim);
  size_t total = prod(dim, dim + ndim);
  if(dst != src) memmove(dst, src, sizeof(double) * 2 * total);
  for(size_t ax = 0; ax < ndim; ax++) fft_axis(ndim, dim, ax, dst, inverse);
}

/* Forward real pass over rows: pairs row k with row h-k (h = n/2 rows of m
 * complex elements each), w = W_n^k.  onedim.hpp:1031-1108 (1-D, m = 1) and
 * multi.hpp:343-414 (rows).  `z` holds h+1 rows; row h is written. */
static void real_pass_fwd(size_t n, size_t m, double* z)
{
  size_t h = n / 2;
  const double pi = acos(-1.0);
  for(size_t k = 1; k <= n / 4; k++)
  {
    double ang = -2.0 * pi * (double) k / (double) n;
    double wr = cos(ang), wi = sin(ang);
    double* r0 = z + 2 * k * m;
    double* r1 = z + 2 * (h - k) * m;
    for(size_t c = 0; c < m; c++)
    {
      double s0r = r0[2 * c], s0i = r0[2 * c + 1];
      double s1r = r1[2 * c], s1i = -r1[2 * c + 1]; /* conj(s1) */
      double ar = 0.5 * (s0r + s1r), ai = 0.5 * (s0i + s1i);
      double dr = s0r - s1r, di = s0i - s1i;
      double br = 0.5 * (dr * wr - di * wi), bi = 0.5 * (dr * wi + di * wr);
      /* d0 = a - i b ; d1 = conj(a) - i conj(b) */
      double d0r = ar + bi, d0i = ai - br;
      double d1r = ar - bi, d1i = -ai - br;
      r0[2 * c] = d0r; r0[2 * c + 1] = d0i;
      r1[2 * c] = d1r; r1[2 * c + 1] = d1i;
    }
  }
  double* r0 = z;
  double* rh = z + 2 * h * m;
  for(size_t c = 0; c < m; c++)
  {
    double re = r0[2 * c], im = r0[2 * c + 1];
    r0[2 * c] = re + im; r0[2 * c + 1] = 0.0;
    rh[2 * c] = re - im; rh[2 * c + 1] = 0.0;
  }
}

/* Inverse real pass over rows (inverse=true branch of the same functions). */
static void real_pass_inv(size_t n, size_t m, double* z)
{
  size_t h = n / 2;
  const double pi = acos(-1.0);
  for(size_t k = 1; k <= n / 4; k++)
  {
    double ang = -2.0 * pi * (double) k / (double) n;
    double wr = cos(ang), wi = -sin(ang); /* conj(w) */
    double* r0 = z + 2 * k * m;
    double* r1 = z + 2 * (h - k) * m;
    for(size_t c = 0; c < m; c++)
    {
      double s0r = r0[2 * c], s0i = r0[2 * c + 1];
      double s1r = r1[2 * c], s1i = -r1[2 * c + 1]; /* conj(s1) */
      double ar = s0r + s1r, ai = s0i + s1i;
      double dr = s1r - s0r, di = s1i - s0i;
      double br = dr * wr - di * wi, bi = dr * wi + di * wr;
      double d0r = ar + bi, d0i = ai - br;
      double d1r = ar - bi, d1i = -ai - br;
      r0[2 * c] = d0r; r0[2 * c + 1] = d0i;
      r1[2 * c] = d1r; r1[2 * c + 1] = d1i;
    }
  }
  double* r0 = z;
  double* rh = z + 2 * h * m;
  for(size_t c = 0; c < m; c++)
  {
    double a = r0[2 * c], b = rh[2 * c];
    r0[2 * c] = a + b; r0[2 * c + 1] = a - b;
  }
}

/* Real forward transform.  src: N real values (shape dim), dst: interleaved
 * complex of shape (dim0/2+1, dim1, ...) for N-D or N/2+1 for 1-D. */
void orc_rfft(size_t ndim_in, const size_t* dim_in, const double* src,
              double* dst)
{
  size_t dim[ORC_MAXDIM];
  size_t ndim = squeeze_dims(ndim_in, dim_in, dim);
  size_t n0 = dim[0];
  size_t inner = prod(dim + 1, dim + ndim);
  size_t h = n0 / 2;
  if(n0 == 1)
  {
    /* a single real value (total size 1): DFT is the value itself */
    dst[0] = src[0]; dst[1] = 0.0;
    return;
  }
  /* pack rows 2j / 2j+1 of dim 0 as complex row j */
  for(size_t j = 0; j < h; j++)
    for(size_t c = 0; c < inner; c++)
    {
      dst[2 * (j * inner + c)] = src[(2 * j) * inner + c];
      dst[2 * (j * inner + c) + 1] = src[(2 * j + 1) * inner + c];
    }
  /* N0/2-point c2c along dim 0 (rows of length inner) */
  size_t d2[2] = {h, inner};
  fft_axis(2, d2, 0, dst, 0);
  real_pass_fwd(n0, inner, dst);
  /* (ndim-1)-D c2c on each of the h+1 rows */
  if(ndim > 1)
    for(size_t r = 0; r <= h; r++)
      for(size_t ax = 1; ax < ndim; ax++)
        fft_axis(ndim - 1, dim + 1, ax - 1, dst + 2 * r * inner, 0);
}

/* Real inverse transform.  src: interleaved complex (h+1 rows), dst: reals. */
void orc_irfft(size_t ndim_in, const size_t* dim_in, const double* src,
               double* dst)
{
  size_t dim[ORC_MAXDIM];
  size_t ndim = squeeze_dims(ndim_in, dim_in, dim);
  size_t n0 = dim[0];
  size_t inner = prod(dim + 1, dim + ndim);
  size_t h = n0 / 2;
  if(n0 == 1)
  {
    dst[0] = src[0];
    return;
  }
  double* z = (double*) malloc(sizeof(double) * 2 * (h + 1) * inner);
  memcpy(z, src, sizeof(double) * 2 * (h + 1) * inner);
  if(ndim > 1)
    for(size_t r = 0; r <= h; r++)
      for(size_t ax = 1; ax < ndim; ax++)
        fft_axis(ndim - 1, dim + 1, ax - 1, z + 2 * r * inner, 1);
  real_pass_inv(n0, inner, z);
  size_t d2[2] = {h, inner};
  fft_axis(2, d2, 0, z, 1);
  for(size_t j = 0; j < h; j++)
    for(size_t c = 0; c < inner; c++)
    {
      dst[(2 * j) * inner + c] = z[2 * (j * inner + c)];
      dst[(2 * j + 1) * inner + c] = z[2 * (j * inner + c) + 1];
    }
  free(z);
}

/* Textbook DFT in long double, O(N^2).  Pins orc_fft. */
void orc_dft_bruteforce(size_t n, const double* src, double* dst, int inverse)
{
  const long double pi = acosl(-1.0L);
  long double sgn = inverse ? 1.0L : -1.0L;
  for(size_t k = 0; k < n; k++)
  {
    long double sr = 0, si = 0;
    for(size_t j = 0; j < n; j++)
    {
      size_t m = (j * k) % n;
      long double ang = sgn * 2.0L * pi * (long double) m / (long double) n;
      long double c = cosl(ang), s = sinl(ang);
      sr += src[2 * j] * c - src[2 * j + 1] * s;
      si += src[2 * j] * s + src[2 * j + 1] * c;
    }
    dst[2 * k] = (double) sr;
    dst[2 * k + 1] = (double) si;
  }
}

/* ||y - ref|| / ||ref|| over n values, y in float, ref in double. */
double orc_rel_l2(size_t n, const float* y, const double* ref)
{
  long double num = 0, den = 0;
  for(size_t i = 0; i < n; i++)
  {
    long double d = (long double) y[i] - (long double) ref[i];
    num += d * d;
    den += (long double) ref[i] * (long double) ref[i];
  }
  if(den == 0) return num == 0 ? 0.0 : INFINITY;
  return (double) sqrtl(num / den);
}

/* Same metric with a double-precision candidate. */
double orc_rel_l2_d(size_t n, const double* y, const double* ref)
{
  long double num = 0, den = 0;
  for(size_t i = 0; i < n; i++)
  {
    long double d = (long double) y[i] - (long double) ref[i];
    num += d * d;
    den += (long double) ref[i] * (long double) ref[i];
  }
  if(den == 0) return num == 0 ? 0.0 : INFINITY;
  return (double) sqrtl(num / den);
}

/* testing.cpp:1113-1127: sqrt(sum (d1-d0)^2 / sum (d0^2 + d1^2)). */
double orc_harness_error(size_t n, const float* d1, const double* d0)
{
  long double diff = 0, sum = 0;
  for(size_t i = 0; i < n; i++)
  {
    long double a = d0[i], b = d1[i];
    sum += a * a + b * b;
    diff += (b - a) * (b - a);
  }
  if(sum == 0) return 0.0;
  return (double) sqrtl(diff / sum);
}

/* ---- portable counter-based generators ---------------------------------- */

static uint64_t mix64(uint64_t z)
{
  /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double unit_double(uint64_t seed, uint64_t i)
{
  uint64_t h = mix64(mix64(seed) ^ (i * 0xD1B54A32D192ED03ULL));
  return (double) (h >> 11) * (1.0 / 9007199254740992.0); /* [0, 1) */
}

/* x[i] = float(2u - 1), u in [0,1): uniform on [-1, 1). */
void orc_uniform(uint64_t seed, size_t n, float* out)
{
  for(size_t i = 0; i < n; i++)
    out[i] = (float) (2.0 * unit_double(seed, i) - 1.0);
}

/* Standard normal by Box-Muller in FP64 (u1 in (0,1]), rounded to float. */
void orc_normal(uint64_t seed, size_t n, float* out)
{
  const double two_pi = 2.0 * acos(-1.0);
  for(size_t i = 0; i < n; i++)
  {
    uint64_t pair = i / 2;
    double u1 = 1.0 - unit_double(seed, 2 * pair);
    double u2 = unit_double(seed, 2 * pair + 1);
    double r = sqrt(-2.0 * log(u1));
    double v = (i & 1) ? r * sin(two_pi * u2) : r * cos(two_pi * u2);
    out[i] = (float) v;
  }
}
