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
__device__ __forceinline__ double re_mul(double px, double py, C2 a) { return px * a.x - py * a.y; }

// W = Re(A1 w*^2)/2 + C1 |w|^2/2 + chi  (getKai :258-311, getW :314-336)
__device__ double hrtem_w(const double* p, double u, double v, double la) {
  const C2 w = {la * u, la * v}, s = {la * u, -la * v};
  const C2 s2 = cm(s, s), s3 = cm(s2, s), s4 = cm(s2, s2), s5 = cm(s4, s), s6 = cm(s3, s3);
  const C2 w2 = cm(w, w), w3 = cm(w2, w), w4 = cm(w2, w2), w5 = cm(w4, w);
  const double r2 = w.x * w.x + w.y * w.y;  // |w|^2
  double chi = re_mul(p[28] / 3, p[29] / 3, s3) + re_mul(p[30], p[31], cm(w2, s)) +
               re_mul(p[32] / 4, p[33] / 4, s4) + re_mul(p[34], p[35], cm(s, w3)) + p[8] * (r2 * r2) / 4 +
               re_mul(p[37] / 5, p[38] / 5, s5) + re_mul(p[39], p[40], cm(w4, s)) +
               re_mul(p[41], p[42], cm(w3, s2)) + re_mul(p[43] / 6, p[44] / 6, s6) +
               re_mul(p[45], p[46], cm(w4, s2)) + p[21] * (r2 * r2 * r2) / 6 + re_mul(p[47], p[48], cm(w5, s));
  return re_mul(p[26], p[27], s2) / 2 + p[1] * r2 / 2 + chi;
}

// gradient terms exactly as the reference writes them (tempgradient_u1/u2 :338-420,
// tempgradient_v1/v2 :431-512), including its A3 slot in the v-gradient and its A5 terms
__device__ void hrtem_grad(const double* p, double u, double v, double la, double& gu, double& gv) {
  const double A1x = p[26], A1y = p[27], A2x = p[28], A2y = p[29], B2x = p[30], B2y = p[31];
  const double A3x = p[32], A3y = p[33], S3x = p[34], S3y = p[35], C3 = p[8];
  const double A4x = p[37], A4y = p[38], C1 = p[1];
  const double D4x = p[39], D4y = p[40], B4x = p[41], B4y = p[42], A5x = p[43], A5y = p[44];
  const double S5x = p[45], S5y = p[46], C5 = p[21], D5x = p[47], D5y = p[48];
  const double u2 = u * u, v2 = v * v, u3 = u2 * u, v3 = v2 * v, u4 = u2 * u2, v4 = v2 * v2;
  const double r2 = u2 + v2, la2 = la * la, la3 = la2 * la, la4 = la2 * la2, la5 = la4 * la;
  gu = la * (A1x * u + A1y * v) + A2x * la2 / 3 * (3 * u2 - 3 * v2) + A2y * la2 / 3 * 6 * u * v +
       B2x * (3 * u2 + v2) * la2 - B2y * la2 * 2 * u * v +
       1.0 / 4 * la3 * (A3x * (4 * u3 - 12 * u * v2) + A3y * (12 * u2 * v - 4 * v3)) +
       la3 * (2 * u * (S3x * (u2 - v2) - S3y * 2 * u * v) + r2 * (2 * S3x * u - 2 * S3y * v)) + C3 * la3 * r2 * u +
       1.0 / 5 * la4 * (A4x * (5 * u4 + 5 * v4 - 30 * u2 * v2) + A4y * (20 * u3 * v - 20 * u * v3));
  gu += la4 * ((2 * u) * (D4x * (u3 - 3 * u * v2) - D4y * (3 * u2 * v - v3)) +
               r2 * (D4x * (3 * u2 - 3 * v2) - D4y * 6 * u * v)) +
        la4 * (2 * r2 * 2 * u * (B4x * u - B4y * v) + r2 * r2 * B4x) +
        1.0 / 6 * la5 *
            (A5x * (4 * u3 - 14 * v) * (u2 - v2) + A5x * (u4 + v4 - 14 * u * v) * 2 * u -
             A5y * 2 * v * (3 * u4 + 3 * v4 - 10 * u2 * v2) - A5y * 2 * v * u * (12 * u3 - 20 * u * v2)) +
        la5 * (2 * r2 * 2 * u * (S5x * (u2 - v2) - 2 * u * v * S5y) + r2 * r2 * (2 * S5x * u - 2 * v * S5y)) +
        1.0 / 6 * la5 * C5 * 3 * r2 * r2 * 2 * u +
        la5 * (2 * u * (D5x * (u4 + v4 - 6 * u2 * v2) - D5y * 4 * u * v * (u2 - v2)) +
               r2 * (D5x * (4 * u3 - 12 * u * v2) - D5y * (12 * u2 * v - 4 * v3))) +
        C1 * la * u;
  gv = la * (A1y * u - A1x * v) - A3x * la2 / 3 * 6 * u * v + A3y * la2 / 3 * (-3 * v2 + 3 * u2) +
       B2x * la2 * 2 * u * v - B2y * la2 * (u2 + 3 * v2) +
       1.0 / 4 * la3 * (A3x * (4 * v3 - 12 * u2 * v) + A3y * (4 * u3 - 12 * u * v2)) +
       la3 * (2 * v * (S3x * (u2 - v2) - S3y * 2 * u * v) + r2 * (-2 * S3x * v - 2 * S3y * u)) +
       C3 * la3 * r2 * v + 1.0 / 5 * la4 * (A4x * (20 * u * v3 - 20 * u3 * v) + A4y * (5 * v4 + 5 * u4 - 30 * u2 * v2));
  gv += la4 * ((2 * v) * (D4x * (u3 - 3 * u * v2) - D4y * (3 * u2 * v - v3)) +
               r2 * (-D4x * 6 * u * v - D4y * (3 * u2 - 3 * v2))) +
        la4 * (2 * r2 * 2 * v * (B4x * u - B4y * v) - r2 * r2 * B4y) +
        1.0 / 6 * la5 *
            (A5x * (4 * v3 - 14 * u) * (u2 - v2) + A5x * (u4 + v4 - 14 * u * v) * (-2 * v) -
             A5y * 2 * u * (3 * u4 + 3 * v4 - 10 * u2 * v2) - A5y * 2 * v * u * (12 * v3 - 20 * u2 * v)) +
        la5 * (2 * r2 * 2 * v * (S5x * (u2 - v2) - 2 * u * v * S5y) + r2 * r2 * (-2 * S5x * v - 2 * u * S5y)) +
        1.0 / 6 * la5 * C5 * 3 * r2 * r2 * 2 * v +
        la5 * (2 * v * (D5x * (u4 + v4 - 6 * u2 * v2) - D5y * 4 * u * v * (u2 - v2)) +
               r2 * (D5x * (4 * v3 - 12 * u2 * v) - D5y * (4 * u3 - 12 * u * v2))) +
        C1 * la * v;
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
