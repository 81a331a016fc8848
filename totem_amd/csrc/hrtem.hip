// Batch HRTEM imaging with drawn aberrations: the computation behind the
// reference's Python tool (cuda_stem64.py:164-704).  For every batch element b:
//
//   para_b   parameter draw around the nominal aberrations      (createNewPP2, :176-240)
//   W, E_s   aberration function and spatial-coherence envelope  (:258-533)
//   tl       exp(-2 pi i W / lambda) * E_s^T * focal phase of member 0   (:535-579)
//   I_b      gfsf * |IFFT2(ifftshift(wave * tl))|^2               (:681-700)
//
// k_hrtem_transfer evaluates the transfer function in FP64 and writes the
// product with the exit wave, already quadrant-swapped, as complex64; the
// batched inverse 2-D FFT is the engine's Stockham FFT (1/N^2 folded into its
// column pass); k_hrtem_intensity forms gfsf |x|^2 in FP64.  The envelope uses
// single-precision exp like the reference (getESCoh, :522-533), and is taken at
// the transposed pixel (the reference transposes it, :142).  Only focal member 0
// reaches the output, as in the reference (its quadrant swap and accumulation
// cover the first `batches` products only, :682-700).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "host_common.h"
#include "launch.h"

namespace totem {

namespace {

constexpr int kPara = 49;
constexpr int kDraw = 21;

// createNewPP2 (cuda_stem64.py:176-240): one thread per batch element.
__global__ void k_hrtem_draw(const double* __restrict__ pp2, const double* __restrict__ mrr,
                             const double* __restrict__ ai, int B, double* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double* r = out + (size_t)b * kPara;
  r[0] = pp2[0];
  for (int i = 1; i <= kDraw; ++i) r[i] = pp2[i] + (ai[(size_t)b * kDraw + i - 1] - 0.5) * 2 * mrr[i - 1];
  for (int i = 22; i <= 25; ++i) r[i] = pp2[i];
  // x/y components (destination, magnitude, angle in degrees)
  const int xy[11][3] = {{26, 2, 3},   {28, 4, 5},   {30, 6, 7},   {32, 9, 10},  {34, 11, 12}, {37, 13, 14},
                         {39, 15, 16}, {41, 17, 18}, {43, 19, 20}, {45, 22, 23}, {47, 24, 25}};
  for (int k = 0; k < 11; ++k) {
    const double m = r[xy[k][1]], a = r[xy[k][2]] / 180 * M_PI;
    r[xy[k][0]] = m * cos(a);
    r[xy[k][0] + 1] = m * sin(a);
  }
  r[36] = r[8];
}

struct C2 {
  double x, y;
};
__device__ __forceinline__ C2 cm(C2 a, C2 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

// The aberration function W of getW (:314-336) with getKai's residual chi (:258-311)
// as a sum of terms Re(C s^m w^n), w = lambda (u + i v), s = conj(w); C = (p[ix] + i p[iy]) / div
// (iy < 0: a real coefficient).
struct AbTerm {
  int ix, iy;
  double div;
  int m, n;
};
__constant__ AbTerm kAbTerms[14] = {
    {26, 27, 2, 2, 0},  // A1   Re(A1 s^2) / 2
    {1, -1, 2, 1, 1},   // C1   C1 |w|^2 / 2
    {28, 29, 3, 3, 0},  // A2   Re(A2 s^3) / 3
    {30, 31, 1, 1, 2},  // B2   Re(B2 w^2 s)
    {32, 33, 4, 4, 0},  // A3   Re(A3 s^4) / 4
    {34, 35, 1, 1, 3},  // S3   Re(S3 s w^3)
    {8, -1, 4, 2, 2},   // C3   C3 |w|^4 / 4
    {37, 38, 5, 5, 0},  // A4   Re(A4 s^5) / 5
    {39, 40, 1, 1, 4},  // D4   Re(D4 w^4 s)
    {41, 42, 1, 2, 3},  // B4   Re(B4 w^3 s^2)
    {43, 44, 6, 6, 0},  // A5   Re(A5 s^6) / 6
    {45, 46, 1, 2, 4},  // S5   Re(S5 w^4 s^2)
    {21, -1, 6, 3, 3},  // C5   C5 |w|^6 / 6
    {47, 48, 1, 1, 5},  // D5   Re(D5 w^5 s)
};
constexpr int kTermA2 = 2, kTermA3 = 4, kTermA5 = 10;

struct AbPowers {
  C2 sp[7], wp[7];  // s^k, w^k, k = 0..6
};
__device__ __forceinline__ AbPowers ab_powers(double u, double v, double la) {
  AbPowers P;
  const C2 w = {la * u, la * v}, s = {la * u, -la * v};
  P.sp[0] = P.wp[0] = {1.0, 0.0};
#pragma unroll
  for (int k = 1; k < 7; ++k) {
    P.sp[k] = cm(P.sp[k - 1], s);
    P.wp[k] = cm(P.wp[k - 1], w);
  }
  return P;
}
__device__ __forceinline__ C2 ab_coef(const double* p, const AbTerm& t) {
  return {p[t.ix] / t.div, t.iy >= 0 ? p[t.iy] / t.div : 0.0};
}

__device__ double hrtem_w(const double* p, double u, double v, double la) {
  const AbPowers P = ab_powers(u, v, la);
  double W = 0.0;
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    const AbTerm t = kAbTerms[k];
    const C2 c = ab_coef(p, t), z = cm(c, cm(P.sp[t.m], P.wp[t.n]));
    W += z.x;
  }
  return W;
}

// (1/lambda) grad W of one term, by the Wirtinger derivatives of F = C s^m w^n:
// d/du = lambda (dF/dw + dF/ds), d/dv = i lambda (dF/dw - dF/ds), so
// (1/lambda) dRe(F)/du = Re(C (n s^m w^(n-1) + m s^(m-1) w^n)) and
// (1/lambda) dRe(F)/dv = Re(i C (n s^m w^(n-1) - m s^(m-1) w^n)).
__device__ __forceinline__ void term_grad(C2 c, int m, int n, const AbPowers& P, double& gu, double& gv) {
  C2 dw = {0.0, 0.0}, ds = {0.0, 0.0};
  if (n > 0) {
    dw = cm(P.sp[m], P.wp[n - 1]);
    dw = {n * dw.x, n * dw.y};
  }
  if (m > 0) {
    ds = cm(P.sp[m - 1], P.wp[n]);
    ds = {m * ds.x, m * ds.y};
  }
  const C2 a = cm(c, {dw.x + ds.x, dw.y + ds.y});
  const C2 b = cm(c, {dw.x - ds.x, dw.y - ds.y});
  gu += a.x;
  gv -= b.y;  // Re(i z) = -Im(z)
}

// The gradient the reference's kernels compute (tempgradient_u1/u2 :338-420,
// tempgradient_v1/v2 :431-512): (1/lambda) grad W, except for two terms written
// differently there and kept here as written:
//  * the v-gradient's A2 term is d/dv of Re(A2 s^3)/3 evaluated with A3's
//    coefficients (:461-462), so A2 adds nothing to g_v;
//  * the A5 terms are the gradient of
//      P5 = lambda^5/6 [A5x (u^4 + v^4 - 14uv)(u^2 - v^2) - A5y 2uv (3u^4 + 3v^4 - 10u^2 v^2)]
//    (:411, :504-505) instead of that of Re(A5 s^6)/6.
__device__ void hrtem_grad(const double* p, double u, double v, double la, double& gu, double& gv) {
  const AbPowers P = ab_powers(u, v, la);
  gu = gv = 0.0;
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    if (k == kTermA5) continue;
    const AbTerm t = kAbTerms[k];
    double tu = 0.0, tv = 0.0;
    term_grad(ab_coef(p, t), t.m, t.n, P, tu, tv);
    gu += tu;
    if (k != kTermA2) gv += tv;
  }
  {  // v-gradient's A2 slot with A3's coefficients
    const AbTerm t2 = kAbTerms[kTermA2], t3 = kAbTerms[kTermA3];
    const C2 c = {p[t3.ix] / t2.div, p[t3.iy] / t2.div};
    double tu = 0.0, tv = 0.0;
    term_grad(c, t2.m, t2.n, P, tu, tv);
    gv += tv;
  }
  {  // grad P5
    const double A5x = p[43], A5y = p[44];
    const double u2 = u * u, v2 = v * v, u4 = u2 * u2, v4 = v2 * v2;
    const double f = u4 + v4 - 14 * u * v, g = u2 - v2, q = 3 * u4 + 3 * v4 - 10 * u2 * v2;
    const double k5 = la * la * la * la * la / 6;
    gu += k5 * (A5x * ((4 * u2 * u - 14 * v) * g + f * 2 * u) -
                A5y * (2 * v * q + 2 * u * v * (12 * u2 * u - 20 * u * v2)));
    gv += k5 * (A5x * ((4 * v2 * v - 14 * u) * g - f * 2 * v) -
                A5y * (2 * u * q + 2 * u * v * (12 * v2 * v - 20 * u2 * v)));
  }
}

// One thread per (batch element, pixel): out[b][swap(r, c)] = wave[r][c] * tl_0(r, c)
__global__ void __launch_bounds__(256) k_hrtem_transfer(const double2* __restrict__ wave, const double* __restrict__ U,
                                                        const double* __restrict__ V, const double* __restrict__ paras,
                                                        int N, int B, double la, double alafa, int mtotal,
                                                        double dyita, float2* __restrict__ out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long NN = (long)N * N;
  if (idx >= NN * B) return;
  const int b = (int)(idx / NN);
  const int pix = (int)(idx - (long)b * NN);
  const int r = pix / N, c = pix - r * N;
  const double* p = paras + (size_t)b * kPara;
  const double u = U[pix], v = V[pix];
  const double W = hrtem_w(p, u, v, la);
  // envelope at the transposed pixel
  double gu, gv;
  hrtem_grad(p, U[(size_t)c * N + r], V[(size_t)c * N + r], la, gu, gv);
  const float e = expf((float)(-1 * alafa * alafa * M_PI * M_PI / la / la * (gu * gu + gv * gv)));
  const double img = 2 * M_PI * W / la;
  const double t1 = -1 * M_PI * (1 - (mtotal + 1) * 1.0 / 2) * dyita * (u * u + v * v) * la;
  // exp(i t1) * exp(-i img) * e
  const double ph = t1 - img;
  const double tr = cos(ph) * (double)e, ti = sin(ph) * (double)e;
  const double2 wv = wave[pix];
  const int dr = r + N / 2 >= N ? r - N / 2 : r + N / 2, dc = c + N / 2 >= N ? c - N / 2 : c + N / 2;
  out[(size_t)b * NN + (size_t)dr * N + dc] = make_float2((float)(wv.x * tr - wv.y * ti), (float)(wv.x * ti + wv.y * tr));
}

__global__ void __launch_bounds__(256) k_hrtem_intensity(const float2* __restrict__ x, long n, double gfsf,
                                                         double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[i].x, b = x[i].y;
  out[i] = gfsf * (a * a + b * b);
}

// Weighted sum of the intensities of M fields [M][N][N] (complex64), optionally
// quadrant-swapped on output (fftshift): out[r][c] = sum_m w[m] |x_m(src)|^2.
__global__ void __launch_bounds__(256) k_weighted_intensity(const float2* __restrict__ x, int N, int M,
                                                            const double* __restrict__ w, int shift,
                                                            double* __restrict__ out) {
  const long NN = (long)N * N;
  const long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= NN) return;
  long src = pix;
  if (shift) {
    const int r = (int)(pix / N), c = (int)(pix - (long)r * N);
    const int sr = r + N / 2 >= N ? r - N / 2 : r + N / 2, sc = c + N / 2 >= N ? c - N / 2 : c + N / 2;
    src = (long)sr * N + sc;
  }
  double s = 0.0;
  for (int m = 0; m < M; ++m) {
    const float2 v = x[m * NN + src];
    s += w[m] * ((double)v.x * v.x + (double)v.y * v.y);
  }
  out[pix] = s;
}

// out[m][swap(r, c)] = wave[r][c] * tcc[m][r][c] * ap[r][c] (FP64 product, complex64 out)
__global__ void __launch_bounds__(256) k_tcc_field(const double2* __restrict__ wave, const double2* __restrict__ tcc,
                                                   const double* __restrict__ ap, int N, int M,
                                                   float2* __restrict__ out) {
  const long NN = (long)N * N;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NN * M) return;
  const int m = (int)(idx / NN);
  const long pix = idx - (long)m * NN;
  const int r = (int)(pix / N), c = (int)(pix - (long)r * N);
  const double2 a = wave[pix], b = tcc[idx];
  const double g = ap[pix];
  const double re = (a.x * b.x - a.y * b.y) * g, im = (a.x * b.y + a.y * b.x) * g;
  const int dr = r + N / 2 >= N ? r - N / 2 : r + N / 2, dc = c + N / 2 >= N ? c - N / 2 : c + N / 2;
  out[(long)m * NN + (long)dr * N + dc] = make_float2((float)re, (float)im);
}

struct HrtemScratch {
  std::mutex mu;
  DevBuf wave, u, v, pp2, mrr, ai, paras, field, img;
};

void reserve_buf(DevBuf& b, size_t bytes) {
  if (b.bytes < bytes) b.alloc(bytes);
}

HrtemScratch& scratch() {
  static HrtemScratch sc;
  return sc;
}

}  // namespace

void tcc_image_impl(int N, int M, const double* wave, const double* tcc, const double* ap, const double* gfsf,
                    double* image, hipStream_t st) {
  require(N > 0 && N % 2 == 0 && M > 0, "n must be positive and even, m positive");
  require(size_supported(N), "grid " + std::to_string(N) + " has no compiled FFT plan", TOTEM_ERR_UNSUPPORTED);
  require(wave && tcc && ap && gfsf && image, "wave, tcc, ap, gfsf and image are required");
  HrtemScratch& sc = scratch();
  std::lock_guard<std::mutex> lk(sc.mu);
  const size_t NN = (size_t)N * N;
  reserve_buf(sc.wave, NN * sizeof(double2));
  reserve_buf(sc.u, NN * sizeof(double));              // aperture
  reserve_buf(sc.paras, (size_t)M * NN * sizeof(double2));  // tcc stack
  reserve_buf(sc.pp2, (size_t)M * sizeof(double));     // weights
  reserve_buf(sc.field, (size_t)M * NN * sizeof(float2));
  reserve_buf(sc.img, NN * sizeof(double));
  HIP_OK(hipMemcpyAsync(sc.wave.p, wave, NN * sizeof(double2), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.paras.p, tcc, (size_t)M * NN * sizeof(double2), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.u.p, ap, NN * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.pp2.p, gfsf, (size_t)M * sizeof(double), hipMemcpyHostToDevice, st));
  const long total = (long)M * (long)NN;
  hipLaunchKernelGGL(k_tcc_field, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, sc.wave.as<double2>(),
                     sc.paras.as<double2>(), sc.u.as<double>(), N, M, sc.field.as<float2>());
  fft2d(N, st, sc.field.as<float2>(), sc.field.as<float2>(), M, +1, 2, 0.f);
  hipLaunchKernelGGL(k_weighted_intensity, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, st,
                     sc.field.as<float2>(), N, M, sc.pp2.as<double>(), 0, sc.img.as<double>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(image, sc.img.p, NN * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
}

void weighted_pattern_impl(int N, int M, const void* fields, const double* weights, int shift, double* out,
                           hipStream_t st) {
  require(N > 0 && N % 2 == 0 && M > 0, "n must be positive and even, m positive");
  require(fields && weights && out, "fields, weights and out are required");
  HrtemScratch& sc = scratch();
  std::lock_guard<std::mutex> lk(sc.mu);
  const size_t NN = (size_t)N * N;
  reserve_buf(sc.pp2, (size_t)M * sizeof(double));
  reserve_buf(sc.img, NN * sizeof(double));
  HIP_OK(hipMemcpyAsync(sc.pp2.p, weights, (size_t)M * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_weighted_intensity, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, st,
                     static_cast<const float2*>(fields), N, M, sc.pp2.as<double>(), shift ? 1 : 0,
                     sc.img.as<double>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, sc.img.p, NN * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
}

void hrtem_images_impl(const totem_hrtem_desc& d, double* paras_out, double* images, hipStream_t st) {
  const int N = d.n, B = d.batches;
  require(N > 0 && B > 0, "n and batches must be positive");
  require(N % 2 == 0, "grid side must be even");
  require(size_supported(N), "grid " + std::to_string(N) + " has no compiled FFT plan", TOTEM_ERR_UNSUPPORTED);
  require(d.mtotal >= 1, "mtotal must be >= 1");
  require(d.u && d.v && d.wave && d.pp2 && d.mrr && d.ai && images, "u, v, wave, pp2, mrr, ai and images are required");
  HrtemScratch& sc = scratch();
  std::lock_guard<std::mutex> lk(sc.mu);
  const size_t NN = (size_t)N * N;
  reserve_buf(sc.wave, NN * sizeof(double2));
  reserve_buf(sc.u, NN * sizeof(double));
  reserve_buf(sc.v, NN * sizeof(double));
  reserve_buf(sc.pp2, kPara * sizeof(double));
  reserve_buf(sc.mrr, kDraw * sizeof(double));
  reserve_buf(sc.ai, (size_t)B * kDraw * sizeof(double));
  reserve_buf(sc.paras, (size_t)B * kPara * sizeof(double));
  reserve_buf(sc.field, (size_t)B * NN * sizeof(float2));
  reserve_buf(sc.img, (size_t)B * NN * sizeof(double));
  HIP_OK(hipMemcpyAsync(sc.wave.p, d.wave, NN * sizeof(double2), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.u.p, d.u, NN * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.v.p, d.v, NN * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.pp2.p, d.pp2, kPara * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.mrr.p, d.mrr, kDraw * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(sc.ai.p, d.ai, (size_t)B * kDraw * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_hrtem_draw, dim3((B + 63) / 64), dim3(64), 0, st, sc.pp2.as<double>(), sc.mrr.as<double>(),
                     sc.ai.as<double>(), B, sc.paras.as<double>());
  const long total = (long)B * (long)NN;
  hipLaunchKernelGGL(k_hrtem_transfer, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     sc.wave.as<double2>(), sc.u.as<double>(), sc.v.as<double>(), sc.paras.as<double>(), N, B,
                     d.lambda_nm, d.alafa, d.mtotal, d.delta_yita, sc.field.as<float2>());
  fft2d(N, st, sc.field.as<float2>(), sc.field.as<float2>(), B, +1, 2, 0.f);
  hipLaunchKernelGGL(k_hrtem_intensity, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                     sc.field.as<float2>(), total, d.gfsf, sc.img.as<double>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(images, sc.img.p, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, st));
  if (paras_out)
    HIP_OK(hipMemcpyAsync(paras_out, sc.paras.p, (size_t)B * kPara * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
}

}  // namespace totem
