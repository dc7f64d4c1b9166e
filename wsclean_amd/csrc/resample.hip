// Image resampling by Fourier interpolation (the small-inversion resampler
// of wgridder/wgriddingmsgridder.cpp:290-308, :326-336): 2-D FFT of the
// input, spectrum cropped or zero-padded to the output size, inverse FFT,
// real part / (input pixels), so output pixel m samples the input's
// band-limited function at m * N_in / N_out.  At even sizes the input's
// Nyquist bin is split over +-N/2 when upsampling and the output's Nyquist
// bin collects both input bins +-M/2 when downsampling.
//
// Line FFTs use the LDS Stockham kernel; lengths that are not 2^a 3^b 5^c
// (image sizes are arbitrary) go through Bluestein's algorithm: with the
// chirp w_k = exp(-i pi k^2 / n), X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),
// a circular convolution of length M >= 2n - 1 done with two LDS FFTs of
// length M and the precomputed transform of the chirp.  The chirp angle
// uses k^2 mod 2n in integers, so it is exact before the FP64 sincos.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

constexpr int kBlock = 256;

unsigned Blocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  return unsigned(b < 65536 ? (b ? b : 1) : 65536);
}

bool Smooth235(int n) {
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

__global__ void RealToComplexKernel(const float* __restrict__ in, size_t n,
                                    float2* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock)
    out[i] = make_float2(in[i], 0.f);
}

__global__ void RealPartKernel(const float2* __restrict__ in, size_t n, float scale,
                               float* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock)
    out[i] = in[i].x * scale;
}

// a[l][k] = conj?(x[l][k]) w[k] for k < n, 0 for n <= k < m
__global__ void BluesteinPreKernel(const float2* __restrict__ x, const float2* __restrict__ w,
                                   int n, int m, size_t lines, int inverse,
                                   float2* __restrict__ a) {
  const size_t total = lines * size_t(m);
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < total;
       i += size_t(gridDim.x) * kBlock) {
    const size_t l = i / m;
    const int k = int(i - l * m);
    float2 v = make_float2(0.f, 0.f);
    if (k < n) {
      float2 xv = x[l * n + k];
      if (inverse) xv.y = -xv.y;
      const float2 c = w[k];
      v = make_float2(xv.x * c.x - xv.y * c.y, xv.x * c.y + xv.y * c.x);
    }
    a[i] = v;
  }
}

__global__ void BluesteinMulKernel(float2* __restrict__ a, const float2* __restrict__ B, int m,
                                   size_t lines) {
  const size_t total = lines * size_t(m);
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < total;
       i += size_t(gridDim.x) * kBlock) {
    const float2 u = a[i], v = B[i % m];
    a[i] = make_float2(u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x);
  }
}

// x[l][k] = conj?(w[k] c[l][k])
__global__ void BluesteinPostKernel(const float2* __restrict__ c, const float2* __restrict__ w,
                                    int n, int m, size_t lines, int inverse,
                                    float2* __restrict__ x) {
  const size_t total = lines * size_t(n);
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < total;
       i += size_t(gridDim.x) * kBlock) {
    const size_t l = i / n;
    const int k = int(i - l * n);
    const float2 u = c[l * m + k], v = w[k];
    float2 r = make_float2(u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x);
    if (inverse) r.y = -r.y;
    x[i] = r;
  }
}

// b[j] = conj(w_j) / m for |j| < n (circularly), 0 elsewhere
__global__ void ChirpKernel(const float2* __restrict__ w, int n, int m,
                            float2* __restrict__ b) {
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < m; j += gridDim.x * kBlock) {
    int k = -1;
    if (j < n) k = j;
    else if (j > m - n) k = m - j;
    const float inv_m = 1.0f / float(m);
    b[j] = k >= 0 ? make_float2(w[k].x * inv_m, -w[k].y * inv_m) : make_float2(0.f, 0.f);
  }
}

// out[kx][ky] (wo x ho, x-major) from in[kx][ky] (wi x hi) through the
// per-axis maps: map[k] = {index0, index1} with weights {w0, w1} (index -1:
// no term)
struct AxisTerm {
  int i0, i1;
  float w0, w1;
};

__global__ void SpectrumResizeKernel(const float2* __restrict__ in, int hi,
                                     const AxisTerm* __restrict__ mx, int wo,
                                     const AxisTerm* __restrict__ my, int ho,
                                     float2* __restrict__ out) {
  const size_t total = size_t(wo) * ho;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < total;
       i += size_t(gridDim.x) * kBlock) {
    const int kx = int(i / ho), ky = int(i - size_t(kx) * ho);
    const AxisTerm tx = mx[kx], ty = my[ky];
    float2 s = make_float2(0.f, 0.f);
    const int ix[2] = {tx.i0, tx.i1}, iy[2] = {ty.i0, ty.i1};
    const float wx[2] = {tx.w0, tx.w1}, wy[2] = {ty.w0, ty.w1};
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        if (ix[a] >= 0 && iy[b] >= 0) {
          const float2 v = in[size_t(ix[a]) * hi + iy[b]];
          const float f = wx[a] * wy[b];
          s.x += f * v.x;
          s.y += f * v.y;
        }
    out[i] = s;
  }
}

// terms of output frequency bin ko (see the file comment)
AxisTerm AxisMap(int n_in, int n_out, int ko) {
  const int f = ko < (n_out + 1) / 2 ? ko : ko - n_out;
  const int lim = std::min(n_in, n_out);
  AxisTerm t{-1, -1, 0.f, 0.f};
  if (n_out > n_in && n_in % 2 == 0 && std::abs(f) == n_in / 2) {
    t.i0 = n_in / 2;
    t.w0 = 0.5f;
  } else if (n_out < n_in && n_out % 2 == 0 && f == -(n_out / 2)) {
    t.i0 = n_out / 2;
    t.w0 = 1.f;
    t.i1 = n_in - n_out / 2;
    t.w1 = 1.f;
  } else if ((f >= 0 && f < (lim + 1) / 2) || (f < 0 && -f <= (lim - 1) / 2) || n_out == n_in) {
    t.i0 = ((f % n_in) + n_in) % n_in;
    t.w0 = 1.f;
  }
  return t;
}

}  // namespace

void Engine::LineFFTAny(float2* data, size_t nlines, int n, bool inverse) {
  if (n == 1 || nlines == 0) return;
  if (Smooth235(n) && n <= 16384) {
    LaunchLineFFT(data, int(nlines), MakeFFTDesc(n, Twiddles(n)), inverse, stream_);
    return;
  }
  WG_CHECK(2 * n - 1 <= 16384, "line length above 8192 that is not 2^a 3^b 5^c");
  const int m = GoodFFTSize(2 * n - 1);
  const std::string tag = "bluestein." + std::to_string(n);
  float2* w = static_cast<float2*>(arena_.Get(tag + ".w", size_t(n) * sizeof(float2)));
  float2* B = static_cast<float2*>(arena_.Get(tag + ".B", size_t(m) * sizeof(float2)));
  if (!bluestein_ready_.count(n)) {
    std::vector<float2> hw(n);
    for (int k = 0; k < n; ++k) {
      const long long q = (long long)k * k % (2LL * n);  // k^2 mod 2n, exact
      const double ang = -M_PI * double(q) / double(n);
      hw[k] = make_float2(float(std::cos(ang)), float(std::sin(ang)));
    }
    WG_HIP(hipMemcpyAsync(w, hw.data(), size_t(n) * sizeof(float2), hipMemcpyHostToDevice,
                          stream_));
    ChirpKernel<<<Blocks(m), kBlock, 0, stream_>>>(w, n, m, B);
    LaunchLineFFT(B, 1, MakeFFTDesc(m, Twiddles(m)), false, stream_);
    WG_HIP(hipStreamSynchronize(stream_));  // hw must outlive the copy
    bluestein_ready_.insert(n);
  }
  float2* a = static_cast<float2*>(arena_.Get("bluestein.work", nlines * size_t(m) *
                                                                     sizeof(float2)));
  BluesteinPreKernel<<<Blocks(nlines * m), kBlock, 0, stream_>>>(data, w, n, m, nlines,
                                                                  inverse ? 1 : 0, a);
  const FFTDesc dm = MakeFFTDesc(m, Twiddles(m));
  LaunchLineFFT(a, int(nlines), dm, false, stream_);
  BluesteinMulKernel<<<Blocks(nlines * m), kBlock, 0, stream_>>>(a, B, m, nlines);
  LaunchLineFFT(a, int(nlines), dm, true, stream_);
  BluesteinPostKernel<<<Blocks(nlines * n), kBlock, 0, stream_>>>(a, w, n, m, nlines,
                                                                   inverse ? 1 : 0, data);
  WG_HIP(hipGetLastError());
}

void Engine::Resample(const float* in, int wi, int hi, float* out, int wo, int ho) {
  WG_CHECK(wi > 0 && hi > 0 && wo > 0 && ho > 0, "empty image");
  const size_t nin = size_t(wi) * hi, nout = size_t(wo) * ho;
  const size_t nbig = std::max(nin, nout);
  float2* A = static_cast<float2*>(arena_.Get("resample.a", nbig * sizeof(float2)));
  float2* Bf = static_cast<float2*>(arena_.Get("resample.b", nbig * sizeof(float2)));
  RealToComplexKernel<<<Blocks(nin), kBlock, 0, stream_>>>(in, nin, A);
  LineFFTAny(A, size_t(hi), wi, false);          // rows: [hi][wi]
  LaunchTranspose(A, Bf, hi, wi, 1, stream_);     // -> [wi][hi]
  LineFFTAny(Bf, size_t(wi), hi, false);         // columns
  std::vector<AxisTerm> tx(wo), ty(ho);
  for (int k = 0; k < wo; ++k) tx[k] = AxisMap(wi, wo, k);
  for (int k = 0; k < ho; ++k) ty[k] = AxisMap(hi, ho, k);
  AxisTerm* dmap =
      static_cast<AxisTerm*>(arena_.Get("resample.map", size_t(wo + ho) * sizeof(AxisTerm)));
  WG_HIP(hipMemcpyAsync(dmap, tx.data(), size_t(wo) * sizeof(AxisTerm), hipMemcpyHostToDevice,
                        stream_));
  WG_HIP(hipMemcpyAsync(dmap + wo, ty.data(), size_t(ho) * sizeof(AxisTerm),
                        hipMemcpyHostToDevice, stream_));
  SpectrumResizeKernel<<<Blocks(nout), kBlock, 0, stream_>>>(Bf, hi, dmap, wo, dmap + wo, ho, A);
  LineFFTAny(A, size_t(wo), ho, true);           // [wo][ho]: inverse columns
  LaunchTranspose(A, Bf, wo, ho, 1, stream_);     // -> [ho][wo]
  LineFFTAny(Bf, size_t(ho), wo, true);          // inverse rows
  RealPartKernel<<<Blocks(nout), kBlock, 0, stream_>>>(Bf, nout, float(1.0 / double(nin)), out);
  WG_HIP(hipGetLastError());
  WG_HIP(hipStreamSynchronize(stream_));  // the host maps must outlive their copies
}

}  // namespace wg
