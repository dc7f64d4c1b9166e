// Hand-written LDS Stockham FFT (radix 2/3/4/5/8/9/18, FP32) and the per-plane
// line kernels of the w-gridder:
//
//   invert : box rows --FFT(u)--> crop columns --transpose--> image columns
//            --FFT(v), w-screen exp(-2 pi i w_p (n-1-zc)), accumulate-->
//            dirty[x][y]
//   predict: model[x][y] * correction * exp(+2 pi i w_p (n-1-zc)) --FFT(v)-->
//            box rows of the column spectrum --transpose--> --FFT(u)--> box
//
// Only the occupied uv box (rows/columns holding visibilities) and the
// cropped image are moved through HBM; the zero-padded remainder of each
// nu x nv plane only ever exists in LDS.  One line (<= 16384 complex = 128
// KiB) is transformed per workgroup, entirely inside LDS.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

// Two-level twiddle table: kTwLo entries exp(-2 pi i r / n), r < kTwLo,
// then ceil(n / kTwLo) entries exp(-2 pi i kTwLo m / n); kTwSlots bounds its
// LDS copy for n <= 16384.
constexpr int kTwLo = 128;
constexpr int kTwSlots = kTwLo + 16384 / kTwLo;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
  return make_float2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) {
  return make_float2(a.x * s, a.y * s);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
__device__ __forceinline__ void Dft(float2* v);

template <>
__device__ __forceinline__ void Dft<2, false>(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void Dft<2, true>(float2* v) {
  Dft<2, false>(v);
}

template <bool INV>
__device__ __forceinline__ void Dft4(float2* v) {
  float2 t0 = cadd(v[0], v[2]);
  float2 t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]);
  float2 t3 = rot90<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}
template <>
__device__ __forceinline__ void Dft<4, false>(float2* v) { Dft4<false>(v); }
template <>
__device__ __forceinline__ void Dft<4, true>(float2* v) { Dft4<true>(v); }

template <bool INV>
__device__ __forceinline__ void Dft3(float2* v) {
  constexpr float s3 = 0.86602540378443864676f;
  float2 t1 = cadd(v[1], v[2]);
  float2 t2 = csub(v[1], v[2]);
  float2 y0 = cadd(v[0], t1);
  float2 m = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  float2 r = rot90<INV>(cscale(t2, s3));  // -i s3 t2 forward
  v[0] = y0;
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}
template <>
__device__ __forceinline__ void Dft<3, false>(float2* v) { Dft3<false>(v); }
template <>
__device__ __forceinline__ void Dft<3, true>(float2* v) { Dft3<true>(v); }

template <bool INV>
__device__ __forceinline__ void Dft5(float2* v) {
  constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
  constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
  constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
  constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
  float2 t1 = cadd(v[1], v[4]);
  float2 t2 = cadd(v[2], v[3]);
  float2 t3 = csub(v[1], v[4]);
  float2 t4 = csub(v[2], v[3]);
  float2 y0 = cadd(v[0], cadd(t1, t2));
  float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  float2 b1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
  float2 b2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
  float2 r1 = rot90<INV>(b1);  // -i b1 forward
  float2 r2 = rot90<INV>(b2);
  v[0] = y0;
  v[1] = cadd(a1, r1);
  v[4] = csub(a1, r1);
  v[2] = cadd(a2, r2);
  v[3] = csub(a2, r2);
}
template <>
__device__ __forceinline__ void Dft<5, false>(float2* v) { Dft5<false>(v); }
template <>
__device__ __forceinline__ void Dft<5, true>(float2* v) { Dft5<true>(v); }

// Radix 8 as two radix-2 halves and two radix-4 DFTs:
// X[2m] = DFT4(v_n + v_{n+4})[m],  X[2m+1] = DFT4((v_n - v_{n+4}) w8^n)[m].
template <bool INV>
__device__ __forceinline__ void Dft8(float2* v) {
  constexpr float r = 0.70710678118654752440f;
  float2 a[4], b[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a[i] = cadd(v[i], v[i + 4]);
    b[i] = csub(v[i], v[i + 4]);
  }
  // b1 *= w8, b2 *= w8^2, b3 *= w8^3 (w8 = exp(-+ 2 pi i / 8))
  b[1] = INV ? make_float2(r * (b[1].x - b[1].y), r * (b[1].x + b[1].y))
             : make_float2(r * (b[1].x + b[1].y), r * (b[1].y - b[1].x));
  b[2] = rot90<INV>(b[2]);
  b[3] = INV ? make_float2(-r * (b[3].x + b[3].y), r * (b[3].x - b[3].y))
             : make_float2(r * (b[3].y - b[3].x), -r * (b[3].x + b[3].y));
  Dft4<INV>(a);
  Dft4<INV>(b);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    v[2 * m] = a[m];
    v[2 * m + 1] = b[m];
  }
}
template <>
__device__ __forceinline__ void Dft<8, false>(float2* v) { Dft8<false>(v); }
template <>
__device__ __forceinline__ void Dft<8, true>(float2* v) { Dft8<true>(v); }

// Radix 9 as 3 x 3 (n = 3 n1 + n2, k = k1 + 3 k2): radix-3 DFTs over n1 for
// each n2, twiddles w9^(n2 k1), radix-3 DFTs over n2 for each k1.
template <bool INV>
__device__ __forceinline__ void Dft9(float2* v) {
  // w9^1, w9^2, w9^4 with w9 = exp(-+ 2 pi i / 9)
  constexpr float c1 = 0.76604444311897803520f, s1 = 0.64278760968653932632f;
  constexpr float c2 = 0.17364817766693034885f, s2 = 0.98480775301220805936f;
  constexpr float c4 = -0.93969262078590838405f, s4 = 0.34202014332566873304f;
  float2 y[3][3];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    y[n2][0] = v[n2];
    y[n2][1] = v[n2 + 3];
    y[n2][2] = v[n2 + 6];
    Dft3<INV>(y[n2]);
  }
  const float sg = INV ? 1.f : -1.f;
  y[1][1] = cmul(y[1][1], make_float2(c1, sg * s1));
  y[1][2] = cmul(y[1][2], make_float2(c2, sg * s2));
  y[2][1] = cmul(y[2][1], make_float2(c2, sg * s2));
  y[2][2] = cmul(y[2][2], make_float2(c4, sg * s4));
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    float2 z[3] = {y[0][k1], y[1][k1], y[2][k1]};
    Dft3<INV>(z);
    v[k1] = z[0];
    v[k1 + 3] = z[1];
    v[k1 + 6] = z[2];
  }
}
// Radix 18 as 2 x 9 (n = 9 n1 + n2, k = k1 + 2 k2): radix-2 DFTs over n1
// for each n2, twiddles w18^(n2 k1), radix-9 DFTs over n2 for each k1.
template <bool INV>
__device__ __forceinline__ void Dft18(float2* v) {
  // cos, sin of 2 pi n2 / 18 for n2 = 0 .. 8
  constexpr float c[9] = {1.0f, 0.93969262078590838405f, 0.76604444311897803520f, 0.5f,
                          0.17364817766693034885f, -0.17364817766693034885f, -0.5f,
                          -0.76604444311897803520f, -0.93969262078590838405f};
  constexpr float s[9] = {0.0f, 0.34202014332566873304f, 0.64278760968653932632f,
                          0.86602540378443864676f, 0.98480775301220805936f,
                          0.98480775301220805936f, 0.86602540378443864676f,
                          0.64278760968653932632f, 0.34202014332566873304f};
  const float sg = INV ? 1.f : -1.f;
  float2 a[9], b[9];
  a[0] = cadd(v[0], v[9]);
  b[0] = csub(v[0], v[9]);
#pragma unroll
  for (int n2 = 1; n2 < 9; ++n2) {
    a[n2] = cadd(v[n2], v[9 + n2]);
    b[n2] = cmul(csub(v[n2], v[9 + n2]), make_float2(c[n2], sg * s[n2]));
  }
  Dft9<INV>(a);
  Dft9<INV>(b);
#pragma unroll
  for (int k2 = 0; k2 < 9; ++k2) {
    v[2 * k2] = a[k2];
    v[2 * k2 + 1] = b[k2];
  }
}

template <>
__device__ __forceinline__ void Dft<9, false>(float2* v) { Dft9<false>(v); }
template <>
__device__ __forceinline__ void Dft<9, true>(float2* v) { Dft9<true>(v); }
template <>
__device__ __forceinline__ void Dft<18, false>(float2* v) { Dft18<false>(v); }
template <>
__device__ __forceinline__ void Dft<18, true>(float2* v) { Dft18<true>(v); }

// Twiddle exp(-2 pi i k / n) from the two-level LDS table (lo[k % 128] *
// hi[k / 128]; FP32 product of correctly rounded factors).
__device__ __forceinline__ float2 Twiddle(const float2* twl, int k) {
  const float2 lo = twl[k & (kTwLo - 1)];
  const float2 hi = twl[kTwLo + (k >> 7)];
  return cmul(lo, hi);
}

// One Stockham autosort stage of radix R on an LDS-resident line of length
// n.  Ns is the product of the radices of the previous stages and `magic`
// = floor(2^32 / Ns) + 1, so j / Ns = umulhi(j, magic) exactly for the
// j < n / R <= 8192 used here.  Every thread first reads all its
// butterflies into registers, then (after a barrier) writes them, so a
// single LDS buffer suffices.
// FIRST: the first stage (Ns = 1), whose twiddles are all 1 and skipped.
template <int R, bool INV, int BLOCK, int MAXN, bool FIRST>
__device__ __forceinline__ void StockhamStage(float2* buf, int n, int Ns, uint32_t magic,
                                              const float2* twl) {
  constexpr int NB = (MAXN / R + BLOCK - 1) / BLOCK;
  const int nb = n / R;
  const int tws = n / (Ns * R);
  float2 v[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = threadIdx.x + b * BLOCK;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[j + r * nb];
      if constexpr (!FIRST) {
        // (magic wraps to 1 for Ns = 1, hence the select)
        const int k = Ns > 1 ? j - int(__umulhi(uint32_t(j), magic)) * Ns : 0;
        // w^r for r < R from one table twiddle w = exp(-+2 pi i k tws / n):
        // products of at most four factors (squarings and one extra factor),
        // so the rounding stays a few ulp while the LDS reads drop from
        // 2 (R - 1) to 2 per butterfly
        constexpr int RW = R == 18 ? 9 : R;  // radix 18: w^(9 + r) = w^9 w^r
        float2 w[RW];
        w[1] = Twiddle(twl, k * tws);
        if (INV) w[1].y = -w[1].y;
#pragma unroll
        for (int r = 2; r < RW; ++r) w[r] = (r & 1) ? cmul(w[r - 1], w[1]) : cmul(w[r / 2], w[r / 2]);
#pragma unroll
        for (int r = 1; r < RW; ++r) v[b][r] = cmul(v[b][r], w[r]);
        if constexpr (R == 18) {
          const float2 w9 = cmul(w[8], w[1]);
          v[b][9] = cmul(v[b][9], w9);
#pragma unroll
          for (int r = 1; r < 9; ++r) v[b][9 + r] = cmul(cmul(v[b][9 + r], w9), w[r]);
        }
      }
      Dft<R, INV>(v[b]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = threadIdx.x + b * BLOCK;
    if (j < nb) {
      const int k = FIRST ? 0 : (Ns > 1 ? j - int(__umulhi(uint32_t(j), magic)) * Ns : 0);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[base + r * (FIRST ? 1 : Ns)] = v[b][r];
    }
  }
  __syncthreads();
}

// Full in-LDS FFT; unnormalised, forward = exp(-2 pi i jk/n).  Caller must
// have synchronised after filling buf and the twiddle table twl.
template <bool INV, int BLOCK, int MAXN, bool FIRST>
__device__ __forceinline__ void StageOfRadix(int R, float2* buf, int n, int Ns, uint32_t mg,
                                             const float2* twl) {
  if constexpr (FIRST) {
    // twiddle-free variants for the usual first radices only (less code);
    // the others take the general stage with Ns = 1
    switch (R) {
      case 8: StockhamStage<8, INV, BLOCK, MAXN, true>(buf, n, Ns, mg, twl); return;
      case 9: StockhamStage<9, INV, BLOCK, MAXN, true>(buf, n, Ns, mg, twl); return;
      default: StageOfRadix<INV, BLOCK, MAXN, false>(R, buf, n, Ns, mg, twl); return;
    }
  }
  switch (R) {
    case 8: StockhamStage<8, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
    case 9: StockhamStage<9, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
    case 18: StockhamStage<18, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
    case 4: StockhamStage<4, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
    case 2: StockhamStage<2, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
    case 3: StockhamStage<3, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
    default: StockhamStage<5, INV, BLOCK, MAXN, FIRST>(buf, n, Ns, mg, twl); break;
  }
}

template <bool INV, int BLOCK, int MAXN>
__device__ void LdsFFT(float2* buf, const FFTDesc& d, const float2* twl) {
  int Ns = 1;
  for (int s = 0; s < d.nst; ++s) {
    // the line length is re-read per stage through an empty asm statement,
    // so the compiler cannot hoist the LDS addresses of every radix case
    // (all loop-invariant) out of the stage loop and hold them in registers
    int n = d.n;
    asm volatile("" : "+s"(n));
    if (s == 0)  // the first stage without twiddles
      StageOfRadix<INV, BLOCK, MAXN, true>(d.rad[s], buf, n, Ns, d.magic[s], twl);
    else
      StageOfRadix<INV, BLOCK, MAXN, false>(d.rad[s], buf, n, Ns, d.magic[s], twl);
    Ns *= d.rad[s];
  }
}

// Copies the line's twiddle table to LDS (kTwLo + ceil(n / 128) entries).
template <int BLOCK>
__device__ __forceinline__ void LoadTwiddles(float2* twl, const FFTDesc& d) {
  const int cnt = kTwLo + (d.n + kTwLo - 1) / kTwLo;
  for (int i = threadIdx.x; i < cnt; i += BLOCK) twl[i] = d.tw[i];
}

template <int BLOCK>
__device__ __forceinline__ void ZeroLds(float2* buf, int n) {
  for (int i = threadIdx.x; i < n; i += BLOCK) buf[i] = make_float2(0.f, 0.f);
}

// buf[(origin + i) mod n] += src[i] for i in [0, len), folding periodically.
template <int BLOCK>
__device__ __forceinline__ void FoldLoad(float2* buf, int n, const float2* __restrict__ src,
                                         int len, int origin) {
  for (int k = 0; k * n < origin + len; ++k) {
    const int i0 = max(0, k * n - origin);
    const int i1 = min(len, (k + 1) * n - origin);
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
      float2& d = buf[origin + i - k * n];
      const float2 a = src[i];
      d.x += a.x;
      d.y += a.y;
    }
    // successive folds add to the same slots
    __syncthreads();
    __syncthreads();
  }
}

__device__ __forceinline__ int Wrap(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// n - 1 - zc at image pixel (x, y), FP64 (stable form, 0 beyond horizon).
__device__ __forceinline__ double PixelZ(const DevPlan& dp, int x, int y) {
  const double l = double(dp.nx / 2 - x) * dp.dx_rad + dp.l_shift;
  const double m = double(y - dp.ny / 2) * dp.dy_rad + dp.m_shift;
  const double r2 = l * l + m * m;
  const double z = r2 < 1.0 ? -r2 / (sqrt(1.0 - r2) + 1.0) : 0.0;
  return z - dp.zc;
}

__device__ __forceinline__ double PlaneW(const DevPlan& dp, int p) {
  return dp.w_lo + double(p - (dp.W / 2 - 1)) * dp.dw;
}

// sin(pi x), cos(pi x) for |x| <= 1: x = q / 2 + t with q = rint(2 x) and
// |t| <= 1/4, the Taylor series of sin(pi t) (to t^9) and cos(pi t) (to
// t^10) -- truncation below 2e-9, FP32 Horner evaluation within 1e-7 --
// and the quadrant q mod 4 rotating (sin, cos).  The general sincospif
// spends about twice the instructions on range handling this argument does
// not need.
__device__ __forceinline__ void SinCosPiSmall(float x, float& s, float& c) {
  const float q = rintf(2.0f * x);
  const float t = fmaf(q, -0.5f, x);
  const float u = t * t;
  float sp = 0.082145886611128191f;
  sp = fmaf(sp, u, -0.59926452932079188f);
  sp = fmaf(sp, u, 2.5501640398773451f);
  sp = fmaf(sp, u, -5.1677127800499694f);
  sp = fmaf(sp, u, 3.1415926535897931f);
  const float st = sp * t;
  float cp = -0.025806891390014051f;
  cp = fmaf(cp, u, 0.23533063035889312f);
  cp = fmaf(cp, u, -1.3352627688545893f);
  cp = fmaf(cp, u, 4.0587121264167676f);
  cp = fmaf(cp, u, -4.934802200544679f);
  cp = fmaf(cp, u, 1.0f);
  // q = 0: (st, cp); 1: (cp, -st); 2: (-st, -cp); 3: (-cp, st)
  const int qi = int(q) & 3;
  const float ss = (qi & 1) ? cp : st;
  const float cc = (qi & 1) ? st : cp;
  s = (qi & 2) ? -ss : ss;
  c = ((qi + 1) & 2) ? -cc : cc;
}

// exp(sign * 2 pi i * cycles) with the integer part removed in FP64.
__device__ __forceinline__ float2 Phasor(double cycles, float sign) {
  const double f = cycles - rint(cycles);
  float s, c;
  SinCosPiSmall(2.0f * float(f), s, c);
  return make_float2(c, sign * s);
}

// ---------------------------------------------------------------- kernels

// Register staging of one line's input for the line kernels: element i =
// threadIdx.x + k BLOCK (k < PF), zero past len.  A kernel loads the next
// line into it right after storing the current one to LDS, so those global
// loads are in flight while the current line is transformed.  Capacity:
// PF BLOCK elements (half a maximal line by default).
template <int BLOCK, int MAXN, int PF_ = MAXN / 2 / BLOCK>
struct LineStage {
  static constexpr int PF = PF_;
  float2 v[PF];
  static __device__ __forceinline__ bool Fits(int len, int n) {
    return len <= n && len <= PF * BLOCK;
  }
  __device__ __forceinline__ void Load(const float2* __restrict__ src, int len) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int i = threadIdx.x + k * BLOCK;
      v[k] = i < len ? src[i] : make_float2(0.f, 0.f);
    }
  }
  // Writes the whole line: slot (origin + DIR i) mod n holds element i for
  // i < len and zero for len <= i < n (0 <= origin < n, len <= n), so every
  // slot is written exactly once and no zeroing pass is needed.
  template <int DIR>
  __device__ __forceinline__ void Store(float2* buf, int n, int origin) const {
    auto slot = [&](int i) {
      int s = origin + DIR * i;
      if (DIR > 0)
        s -= s >= n ? n : 0;
      else
        s += s < 0 ? n : 0;
      return s;
    };
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int i = threadIdx.x + k * BLOCK;
      if (i < n) buf[slot(i)] = v[k];
    }
    for (int i = threadIdx.x + PF * BLOCK; i < n; i += BLOCK)
      buf[slot(i)] = make_float2(0.f, 0.f);
  }
};

// (a mod n) for -n < a < 2n
__device__ __forceinline__ int WrapNear(int a, int n) {
  return a < 0 ? a + n : (a >= n ? a - n : a);
}

// Persistent row transforms of the invert pass: line L = (plane L / BV, box
// row L % BV), u-FFT of the box row folded into the grid period, cropped to
// the nx image columns.
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) RowFFTInvertKernel(
    DevPlan dp, FFTDesc fu, const float2* __restrict__ box, float2* __restrict__ rbuf,
    int nlines) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  using Stage = LineStage<BLOCK, MAXN>;
  const bool staged = Stage::Fits(dp.BU, fu.n);
  auto src_of = [&](int L) {
    const int j = L % dp.BV, pl = L / dp.BV;
    return box + size_t(pl) * dp.box_plane + size_t(j) * dp.box_pitch;
  };
  LoadTwiddles<BLOCK>(twl, fu);
  Stage st;
  int L = blockIdx.x;
  if (staged && L < nlines) st.Load(src_of(L), dp.BU);
  for (; L < nlines; L += gridDim.x) {
    if (staged) {
      st.template Store<1>(buf, fu.n, dp.box_u0);
      if (L + int(gridDim.x) < nlines) st.Load(src_of(L + gridDim.x), dp.BU);
      __syncthreads();
    } else {
      ZeroLds<BLOCK>(buf, fu.n);
      __syncthreads();
      FoldLoad<BLOCK>(buf, fu.n, src_of(L), dp.BU, dp.box_u0);
    }
    LdsFFT<false, BLOCK, MAXN>(buf, fu, twl);
    float2* out = rbuf + size_t(L) * dp.nx;
    for (int x = threadIdx.x; x < dp.nx; x += BLOCK)
      out[x] = buf[WrapNear(dp.nx / 2 - x, fu.n)];
    // the next line overwrites buf
    __syncthreads();
  }
}

// One image column x through all planes of the slab: v-FFT of the column
// spectrum, w-screen, real part accumulated in registers.  NY = image rows
// per thread (ny <= NY BLOCK).
template <int BLOCK, int MAXN, int NY>
__global__ void __launch_bounds__(BLOCK) ColFFTInvertKernel(
    DevPlan dp, FFTDesc fv, const float2* __restrict__ cbuf, float* __restrict__ dirty,
    int plane_lo, int nplanes) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  // a quarter line of staging: the per-row registers share the budget
  using Stage = LineStage<BLOCK, MAXN, MAXN / 4 / BLOCK>;
  const int x = blockIdx.x;
  float acc[NY];
  double zz[NY];
#pragma unroll
  for (int k = 0; k < NY; ++k) {
    const int y = threadIdx.x + k * BLOCK;
    acc[k] = 0.f;
    zz[k] = y < dp.ny ? PixelZ(dp, x, y) : 0.0;
  }
  const bool staged = Stage::Fits(dp.BV, fv.n);
  auto src_of = [&](int pl) { return cbuf + (size_t(pl) * dp.nx + x) * dp.BV; };
  LoadTwiddles<BLOCK>(twl, fv);
  Stage st;
  if (staged && nplanes > 0) st.Load(src_of(0), dp.BV);
  for (int pl = 0; pl < nplanes; ++pl) {
    if (staged) {
      st.template Store<1>(buf, fv.n, dp.box_v0);
      if (pl + 1 < nplanes) st.Load(src_of(pl + 1), dp.BV);
      __syncthreads();
    } else {
      ZeroLds<BLOCK>(buf, fv.n);
      __syncthreads();
      FoldLoad<BLOCK>(buf, fv.n, src_of(pl), dp.BV, dp.box_v0);
    }
    LdsFFT<false, BLOCK, MAXN>(buf, fv, twl);
    const double wp = PlaneW(dp, plane_lo + pl);
#pragma unroll
    for (int k = 0; k < NY; ++k) {
      const int y = threadIdx.x + k * BLOCK;
      if (y < dp.ny) {
        const float2 val = buf[WrapNear(y - dp.ny / 2, fv.n)];
        const float2 ph = Phasor(wp * zz[k], -1.f);
        acc[k] += val.x * ph.x - val.y * ph.y;
      }
    }
    __syncthreads();
  }
  float* out = dirty + size_t(x) * dp.ny;
#pragma unroll
  for (int k = 0; k < NY; ++k) {
    const int y = threadIdx.x + k * BLOCK;
    if (y < dp.ny) out[y] += acc[k];
  }
}

// One image column x through all planes of the slab, predict direction:
// model * correction * w-screen, inverse v-FFT, the box rows of the column
// spectrum stored.  NY as in ColFFTInvertKernel.
template <int BLOCK, int MAXN, int NY>
__global__ void __launch_bounds__(BLOCK) ColFFTPredictKernel(
    DevPlan dp, FFTDesc fv, const double* __restrict__ img, const float* __restrict__ corr,
    float2* __restrict__ cbuf, int plane_lo, int nplanes) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  const int x = blockIdx.x;
  float mv[NY];
  double zz[NY];
#pragma unroll
  for (int k = 0; k < NY; ++k) {
    const int y = threadIdx.x + k * BLOCK;
    if (y < dp.ny) {
      const size_t idx = size_t(x) * dp.ny + y;
      mv[k] = float(img[idx]) * corr[idx];
      zz[k] = PixelZ(dp, x, y);
    } else {
      mv[k] = 0.f;
      zz[k] = 0.0;
    }
  }
  LoadTwiddles<BLOCK>(twl, fv);
  for (int pl = 0; pl < nplanes; ++pl) {
    const double wp = PlaneW(dp, plane_lo + pl);
    // slot (y - ny / 2) mod n for y in [0, n) covers the line once: the
    // image rows carry the screened model, the rest is zero
#pragma unroll
    for (int k = 0; k < NY; ++k) {
      const int y = threadIdx.x + k * BLOCK;
      if (y < fv.n)
        buf[WrapNear(y - dp.ny / 2, fv.n)] =
            y < dp.ny ? cscale(Phasor(wp * zz[k], 1.f), mv[k]) : make_float2(0.f, 0.f);
    }
    for (int y = threadIdx.x + NY * BLOCK; y < fv.n; y += BLOCK)
      buf[y - dp.ny / 2] = make_float2(0.f, 0.f);
    __syncthreads();
    LdsFFT<true, BLOCK, MAXN>(buf, fv, twl);
    float2* out = cbuf + (size_t(pl) * dp.nx + x) * dp.BV;
    for (int j = threadIdx.x; j < dp.BV; j += BLOCK)
      out[j] = buf[Wrap(dp.box_v0 + j, fv.n)];
    __syncthreads();
  }
}

// Persistent row transforms of the predict pass: line L = (plane L / BV, box
// row L % BV), the nx image columns placed in the grid period, inverse
// u-FFT, the box row stored.
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) RowFFTPredictKernel(
    DevPlan dp, FFTDesc fu, const float2* __restrict__ rbuf, float2* __restrict__ box,
    int nlines) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  using Stage = LineStage<BLOCK, MAXN>;
  const bool staged = Stage::Fits(dp.nx, fu.n);
  LoadTwiddles<BLOCK>(twl, fu);
  Stage st;
  int L = blockIdx.x;
  if (staged && L < nlines) st.Load(rbuf + size_t(L) * dp.nx, dp.nx);
  for (; L < nlines; L += gridDim.x) {
    if (staged) {
      // image column x goes to slot (nx / 2 - x) mod n
      st.template Store<-1>(buf, fu.n, dp.nx / 2);
      if (L + int(gridDim.x) < nlines) st.Load(rbuf + size_t(L + gridDim.x) * dp.nx, dp.nx);
    } else {
      ZeroLds<BLOCK>(buf, fu.n);
      __syncthreads();
      const float2* src = rbuf + size_t(L) * dp.nx;
      for (int x = threadIdx.x; x < dp.nx; x += BLOCK)
        buf[WrapNear(dp.nx / 2 - x, fu.n)] = src[x];
    }
    __syncthreads();
    LdsFFT<true, BLOCK, MAXN>(buf, fu, twl);
    const int j = L % dp.BV, pl = L / dp.BV;
    float2* out = box + size_t(pl) * dp.box_plane + size_t(j) * dp.box_pitch;
    for (int i = threadIdx.x; i < dp.BU; i += BLOCK) out[i] = buf[Wrap(dp.box_u0 + i, fu.n)];
    __syncthreads();
  }
}

template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) LineFFTKernel(float2* __restrict__ data, FFTDesc d,
                                                       int inverse) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  float2* line = data + size_t(blockIdx.x) * d.n;
  LoadTwiddles<BLOCK>(twl, d);
  for (int i = threadIdx.x; i < d.n; i += BLOCK) buf[i] = line[i];
  __syncthreads();
  if (inverse)
    LdsFFT<true, BLOCK, MAXN>(buf, d, twl);
  else
    LdsFFT<false, BLOCK, MAXN>(buf, d, twl);
  for (int i = threadIdx.x; i < d.n; i += BLOCK) line[i] = buf[i];
}

// ------------------------------------------------------------ long lines
// A line of length N = R M (M <= 16384, R in 2..5) in two passes through HBM,
// by one decimation-in-time step: pass 1 transforms the R interleaved
// subsequences x[R m + r] with the LDS FFT of length M (one workgroup each,
// loading from and folding into the period as the one-pass kernels do);
// pass 2 forms only the outputs it needs, X[k] = sum_r w^(r k) F_r[k mod M]
// with w = exp(-+2 pi i / N), fused with the line's consumer (crop, w-screen
// and accumulate, or box store).  tmp holds F as [line][r][M].

// exp(sign 2 pi i q / N)
__device__ __forceinline__ float2 TwiddleN(long long q, int N, float sign) {
  return Phasor(double(q % N) / double(N), sign);
}

// buf[(s - r) / R] += src[i] for the elements i < len whose slot s =
// (origin + i) mod N (folding, one period at a time) has s mod R = r.
template <int BLOCK>
__device__ __forceinline__ void SubsequenceFoldLoad(float2* buf, const float2* __restrict__ src,
                                                    int len, int origin, int N, int R, int r) {
  const int i0 = Wrap(r - origin, R);
  for (int k = 0; k * N < origin + len; ++k) {
    const int lo = max(0, k * N - origin), hi = min(len, (k + 1) * N - origin);
    const int first = lo + Wrap(i0 - lo, R);
    for (int i = first + R * int(threadIdx.x); i < hi; i += R * BLOCK) {
      float2& d = buf[(origin + i - k * N) / R];
      const float2 a = src[i];
      d.x += a.x;
      d.y += a.y;
    }
    __syncthreads();
  }
}

// Pass 1 of the invert row transform: subsequence r of box row L.
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) RowFFTInvertLongKernel(
    DevPlan dp, FFTDesc fm, int N, int R, const float2* __restrict__ box,
    float2* __restrict__ tmp) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  const int M = fm.n;
  const int L = int(blockIdx.x) / R, r = int(blockIdx.x) % R;
  const int j = L % dp.BV, pl = L / dp.BV;
  LoadTwiddles<BLOCK>(twl, fm);
  ZeroLds<BLOCK>(buf, M);
  __syncthreads();
  SubsequenceFoldLoad<BLOCK>(buf, box + size_t(pl) * dp.box_plane + size_t(j) * dp.box_pitch,
                             dp.BU, dp.box_u0, N, R, r);
  LdsFFT<false, BLOCK, MAXN>(buf, fm, twl);
  float2* out = tmp + (size_t(L) * R + r) * M;
  for (int m = threadIdx.x; m < M; m += BLOCK) out[m] = buf[m];
}

// Pass 2 of the invert row transform: image column x of line L.
__global__ void __launch_bounds__(256) RowCombineInvertKernel(
    DevPlan dp, int N, int R, int M, const float2* __restrict__ tmp, float2* __restrict__ rbuf,
    int nlines) {
  const size_t n = size_t(nlines) * dp.nx;
  for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < n;
       e += size_t(gridDim.x) * 256) {
    const int L = int(e / size_t(dp.nx)), x = int(e - size_t(L) * dp.nx);
    const int k = Wrap(dp.nx / 2 - x, N);
    const float2* f = tmp + size_t(L) * R * M + k % M;
    float2 acc = f[0];
    for (int r = 1; r < R; ++r)
      acc = cadd(acc, cmul(f[size_t(r) * M], TwiddleN((long long)r * k, N, -1.f)));
    rbuf[e] = acc;
  }
}

// Pass 1 of the invert column transform: subsequence r of column line
// (plane, x) of the column spectrum.
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) ColFFTInvertLongKernel(
    DevPlan dp, FFTDesc fm, int N, int R, const float2* __restrict__ cbuf,
    float2* __restrict__ tmp) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  const int M = fm.n;
  const int line = int(blockIdx.x) / R, r = int(blockIdx.x) % R;
  LoadTwiddles<BLOCK>(twl, fm);
  ZeroLds<BLOCK>(buf, M);
  __syncthreads();
  SubsequenceFoldLoad<BLOCK>(buf, cbuf + size_t(line) * dp.BV, dp.BV, dp.box_v0, N, R, r);
  LdsFFT<false, BLOCK, MAXN>(buf, fm, twl);
  float2* out = tmp + (size_t(line) * R + r) * M;
  for (int m = threadIdx.x; m < M; m += BLOCK) out[m] = buf[m];
}

// Pass 2 of the invert column transform: pixel (x, y) through the slab's
// planes, w-screen, real part accumulated into the dirty image.
__global__ void __launch_bounds__(256) ColCombineInvertKernel(
    DevPlan dp, int N, int R, int M, const float2* __restrict__ tmp, float* __restrict__ dirty,
    int plane_lo, int nplanes) {
  const size_t n = size_t(dp.nx) * dp.ny;
  for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < n;
       e += size_t(gridDim.x) * 256) {
    const int x = int(e / size_t(dp.ny)), y = int(e - size_t(x) * dp.ny);
    const int k = Wrap(y - dp.ny / 2, N);
    // twiddles of the subsequences r < R (R <= 5; static indexes)
    float2 tw[5];
#pragma unroll
    for (int r = 0; r < 5; ++r)
      tw[r] = r < R ? TwiddleN((long long)r * k, N, -1.f) : make_float2(0.f, 0.f);
    const double zz = PixelZ(dp, x, y);
    float acc = 0.f;
    for (int pl = 0; pl < nplanes; ++pl) {
      const float2* f = tmp + (size_t(pl) * dp.nx + x) * R * M + k % M;
      float2 val = f[0];
#pragma unroll
      for (int r = 1; r < 5; ++r)
        if (r < R) val = cadd(val, cmul(f[size_t(r) * M], tw[r]));
      const float2 ph = Phasor(PlaneW(dp, plane_lo + pl) * zz, -1.f);
      acc += val.x * ph.x - val.y * ph.y;
    }
    dirty[e] += acc;
  }
}

// Pass 1 of the predict column transform: subsequence r of the screened
// model column x of plane pl (image row y sits at slot (y - ny / 2) mod N).
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) ColFFTPredictLongKernel(
    DevPlan dp, FFTDesc fm, int N, int R, const double* __restrict__ img,
    const float* __restrict__ corr, float2* __restrict__ tmp, int plane_lo) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  const int M = fm.n;
  const int line = int(blockIdx.x) / R, r = int(blockIdx.x) % R;
  const int pl = line / dp.nx, x = line % dp.nx;
  const double wp = PlaneW(dp, plane_lo + pl);
  LoadTwiddles<BLOCK>(twl, fm);
  for (int m = threadIdx.x; m < M; m += BLOCK) {
    const int y = Wrap(R * m + r + dp.ny / 2, N);
    float2 v = make_float2(0.f, 0.f);
    if (y < dp.ny) {
      const size_t idx = size_t(x) * dp.ny + y;
      v = cscale(Phasor(wp * PixelZ(dp, x, y), 1.f), float(img[idx]) * corr[idx]);
    }
    buf[m] = v;
  }
  __syncthreads();
  LdsFFT<true, BLOCK, MAXN>(buf, fm, twl);
  float2* out = tmp + (size_t(line) * R + r) * M;
  for (int m = threadIdx.x; m < M; m += BLOCK) out[m] = buf[m];
}

// Pass 2 of the predict column transform: box row j of column line
// (plane, x).
__global__ void __launch_bounds__(256) ColCombinePredictKernel(
    DevPlan dp, int N, int R, int M, const float2* __restrict__ tmp, float2* __restrict__ cbuf,
    int nplanes) {
  const size_t n = size_t(nplanes) * dp.nx * dp.BV;
  for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < n;
       e += size_t(gridDim.x) * 256) {
    const size_t line = e / size_t(dp.BV);
    const int j = int(e - line * dp.BV);
    const int k = Wrap(dp.box_v0 + j, N);
    const float2* f = tmp + line * R * M + k % M;
    float2 acc = f[0];
    for (int r = 1; r < R; ++r)
      acc = cadd(acc, cmul(f[size_t(r) * M], TwiddleN((long long)r * k, N, 1.f)));
    cbuf[e] = acc;
  }
}

// Pass 1 of the predict row transform: subsequence r of line L of the
// row spectrum (image column x sits at slot (nx / 2 - x) mod N).
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) RowFFTPredictLongKernel(
    DevPlan dp, FFTDesc fm, int N, int R, const float2* __restrict__ rbuf,
    float2* __restrict__ tmp) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  const int M = fm.n;
  const int L = int(blockIdx.x) / R, r = int(blockIdx.x) % R;
  const float2* src = rbuf + size_t(L) * dp.nx;
  LoadTwiddles<BLOCK>(twl, fm);
  ZeroLds<BLOCK>(buf, M);
  __syncthreads();
  const int x0 = Wrap(dp.nx / 2 - r, R);
  for (int x = x0 + R * int(threadIdx.x); x < dp.nx; x += R * BLOCK)
    buf[Wrap(dp.nx / 2 - x, N) / R] = src[x];
  __syncthreads();
  LdsFFT<true, BLOCK, MAXN>(buf, fm, twl);
  float2* out = tmp + (size_t(L) * R + r) * M;
  for (int m = threadIdx.x; m < M; m += BLOCK) out[m] = buf[m];
}

// Pass 2 of the predict row transform: box column i of line L.
__global__ void __launch_bounds__(256) RowCombinePredictKernel(
    DevPlan dp, int N, int R, int M, const float2* __restrict__ tmp, float2* __restrict__ box,
    int nlines) {
  const size_t n = size_t(nlines) * dp.BU;
  for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < n;
       e += size_t(gridDim.x) * 256) {
    const int L = int(e / size_t(dp.BU)), i = int(e - size_t(L) * dp.BU);
    const int k = Wrap(dp.box_u0 + i, N);
    const float2* f = tmp + size_t(L) * R * M + k % M;
    float2 acc = f[0];
    for (int r = 1; r < R; ++r)
      acc = cadd(acc, cmul(f[size_t(r) * M], TwiddleN((long long)r * k, N, 1.f)));
    const int row = L % dp.BV, pl = L / dp.BV;
    box[size_t(pl) * dp.box_plane + size_t(row) * dp.box_pitch + i] = acc;
  }
}

// Plain long-line transform of contiguous lines (test hook): pass 1.
template <int BLOCK, int MAXN>
__global__ void __launch_bounds__(BLOCK) LineFFTLongKernel(const float2* __restrict__ data,
                                                           FFTDesc fm, int N, int R,
                                                           float2* __restrict__ tmp,
                                                           int inverse) {
  extern __shared__ float2 lds[];
  float2* twl = lds;
  float2* buf = lds + kTwSlots;
  const int M = fm.n;
  const int line = int(blockIdx.x) / R, r = int(blockIdx.x) % R;
  LoadTwiddles<BLOCK>(twl, fm);
  for (int m = threadIdx.x; m < M; m += BLOCK) buf[m] = data[size_t(line) * N + R * m + r];
  __syncthreads();
  if (inverse)
    LdsFFT<true, BLOCK, MAXN>(buf, fm, twl);
  else
    LdsFFT<false, BLOCK, MAXN>(buf, fm, twl);
  float2* out = tmp + (size_t(line) * R + r) * M;
  for (int m = threadIdx.x; m < M; m += BLOCK) out[m] = buf[m];
}

// Pass 2 of the plain long-line transform: every output.
__global__ void __launch_bounds__(256) LineCombineKernel(const float2* __restrict__ tmp, int N,
                                                         int R, int M, float2* __restrict__ data,
                                                         int nlines, int inverse) {
  const size_t n = size_t(nlines) * N;
  const float sign = inverse ? 1.f : -1.f;
  for (size_t e = size_t(blockIdx.x) * 256 + threadIdx.x; e < n;
       e += size_t(gridDim.x) * 256) {
    const size_t line = e / size_t(N);
    const int k = int(e - line * N);
    const float2* f = tmp + line * R * M + k % M;
    float2 acc = f[0];
    for (int r = 1; r < R; ++r)
      acc = cadd(acc, cmul(f[size_t(r) * M], TwiddleN((long long)r * k, N, sign)));
    data[e] = acc;
  }
}

// out[p][c][r] = in[p][r][c]; 32x32 tiles through LDS.
__global__ void __launch_bounds__(256) TransposeKernel(const float2* __restrict__ in,
                                                       float2* __restrict__ out, int rows,
                                                       int cols) {
  __shared__ float2 tile[32][33];
  const size_t plane = size_t(blockIdx.z) * rows * cols;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    if (r < rows && c < cols) tile[k][tx] = in[plane + size_t(r) * cols + c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + tx;
    if (r < rows && c < cols) out[plane + size_t(c) * rows + r] = tile[tx][k];
  }
}

// LDS of one line kernel: twiddle table + line of length n.
size_t LineLds(int n) { return size_t(kTwSlots + n) * sizeof(float2); }

// Kernels that use more than the default dynamic LDS must opt in once.
template <typename K>
void AllowLds(K kernel, size_t bytes) {
  AllowDynamicLds(reinterpret_cast<const void*>(kernel), bytes);
}

template <typename F>
void Dispatch(int n, F&& f) {
  if (n <= 2048)
    f(std::integral_constant<int, 256>(), std::integral_constant<int, 2048>());
  else if (n <= 8192)
    f(std::integral_constant<int, 512>(), std::integral_constant<int, 8192>());
  else
    f(std::integral_constant<int, 1024>(), std::integral_constant<int, 16384>());
}

}  // namespace

FFTDesc MakeFFTDesc(int n, const float2* tw) {
  WG_CHECK(n >= 1 && n <= 16384, "FFT length out of range");
  FFTDesc d{};
  d.n = n;
  std::vector<int> r = FFTRadices(n);
  WG_CHECK(r.size() <= 16, "too many FFT stages");
  d.nst = int(r.size());
  uint64_t ns = 1;
  for (size_t i = 0; i < r.size(); ++i) {
    d.rad[i] = r[i];
    // j / Ns = umulhi(j, magic) for j < 2^18 (see StockhamStage)
    d.magic[i] = uint32_t((uint64_t(1) << 32) / ns + 1);
    ns *= uint64_t(r[i]);
  }
  d.tw = tw;
  return d;
}

std::vector<float2> FFTTwiddleTable(int n) {
  const int nhi = (n + kTwLo - 1) / kTwLo;
  std::vector<float2> t(kTwLo + nhi);
  for (int i = 0; i < kTwLo + nhi; ++i) {
    const long m = i < kTwLo ? long(i) : long(i - kTwLo) * kTwLo;
    // exact reduction: angle = 2 pi (m mod n) / n in double
    const double a = 2.0 * M_PI * double(m % n) / double(n);
    t[i] = make_float2(float(std::cos(a)), float(-std::sin(a)));
  }
  return t;
}

namespace {

// Grid of a persistent line kernel: the resident workgroups, at most one
// per line.
template <typename K>
int LineGrid(K kernel, int block, size_t lds, int nlines) {
  const int per_cu = std::max(1, ResidentBlocksPerCU(reinterpret_cast<const void*>(kernel),
                                                     block, lds));
  return std::max(1, std::min(nlines, per_cu * DeviceCUs()));
}

// NY (image rows per thread) of the column kernels: half the line capacity
// when ny fits, so the per-row registers stay small.
template <int BLOCK, int MAXN, typename F>
void DispatchNY(int ny, F&& f) {
  constexpr int NYF = MAXN / BLOCK;
  if (ny <= (NYF / 2) * BLOCK)
    f(std::integral_constant<int, NYF / 2>());
  else
    f(std::integral_constant<int, NYF>());
}

}  // namespace

void LaunchRowFFTInvert(const DevPlan& dp, const FFTDesc& fu, const float2* box,
                        float2* rbuf, int nplanes, hipStream_t s) {
  const int nlines = dp.BV * nplanes;
  if (nlines <= 0) return;
  Dispatch(fu.n, [&](auto B, auto M) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(M)::value;
    auto k = RowFFTInvertKernel<BLOCK, MAXN>;
    AllowLds(k, LineLds(MAXN));
    k<<<LineGrid(k, BLOCK, LineLds(fu.n), nlines), BLOCK, LineLds(fu.n), s>>>(dp, fu, box,
                                                                            rbuf, nlines);
  });
  WG_HIP(hipGetLastError());
}

void LaunchColFFTInvert(const DevPlan& dp, const FFTDesc& fv, const float2* cbuf,
                        float* dirty, int plane_lo, int nplanes, hipStream_t s) {
  Dispatch(fv.n, [&](auto B, auto M) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(M)::value;
    DispatchNY<BLOCK, MAXN>(dp.ny, [&](auto Y) {
      constexpr int NY = decltype(Y)::value;
      AllowLds(ColFFTInvertKernel<BLOCK, MAXN, NY>, LineLds(MAXN));
      ColFFTInvertKernel<BLOCK, MAXN, NY><<<dim3(dp.nx), BLOCK, LineLds(fv.n), s>>>(
          dp, fv, cbuf, dirty, plane_lo, nplanes);
    });
  });
  WG_HIP(hipGetLastError());
}

void LaunchColFFTPredict(const DevPlan& dp, const FFTDesc& fv, const double* img,
                         const float* corr, float2* cbuf, int plane_lo, int nplanes,
                         hipStream_t s) {
  Dispatch(fv.n, [&](auto B, auto M) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(M)::value;
    DispatchNY<BLOCK, MAXN>(dp.ny, [&](auto Y) {
      constexpr int NY = decltype(Y)::value;
      AllowLds(ColFFTPredictKernel<BLOCK, MAXN, NY>, LineLds(MAXN));
      ColFFTPredictKernel<BLOCK, MAXN, NY><<<dim3(dp.nx), BLOCK, LineLds(fv.n), s>>>(
          dp, fv, img, corr, cbuf, plane_lo, nplanes);
    });
  });
  WG_HIP(hipGetLastError());
}

void LaunchRowFFTPredict(const DevPlan& dp, const FFTDesc& fu, const float2* rbuf,
                         float2* box, int nplanes, hipStream_t s) {
  const int nlines = dp.BV * nplanes;
  if (nlines <= 0) return;
  Dispatch(fu.n, [&](auto B, auto M) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(M)::value;
    auto k = RowFFTPredictKernel<BLOCK, MAXN>;
    AllowLds(k, LineLds(MAXN));
    k<<<LineGrid(k, BLOCK, LineLds(fu.n), nlines), BLOCK, LineLds(fu.n), s>>>(dp, fu, rbuf,
                                                                            box, nlines);
  });
  WG_HIP(hipGetLastError());
}

void LaunchTranspose(const float2* in, float2* out, int rows, int cols, int nplanes,
                     hipStream_t s) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32, nplanes);
  TransposeKernel<<<grid, 256, 0, s>>>(in, out, rows, cols);
  WG_HIP(hipGetLastError());
}

void LaunchLineFFT(float2* data, int nlines, const FFTDesc& d, bool inverse, hipStream_t s) {
  Dispatch(d.n, [&](auto B, auto M) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(M)::value;
    AllowLds(LineFFTKernel<BLOCK, MAXN>, LineLds(MAXN));
    LineFFTKernel<BLOCK, MAXN><<<dim3(nlines), BLOCK, LineLds(d.n), s>>>(
        data, d, inverse ? 1 : 0);
  });
  WG_HIP(hipGetLastError());
}

// ---- long lines (N = R M > 16384) ----

int LongFFTRadix(int n) {
  if (n <= kMaxLdsFFT) return 1;
  for (int R = 2; R <= 5; ++R)
    if (n % R == 0 && n / R <= kMaxLdsFFT) return R;
  return 0;
}

namespace {
unsigned CombineBlocks(size_t n) {
  return unsigned(std::min<size_t>((n + 255) / 256, size_t(1) << 20));
}
}  // namespace

void LaunchRowFFTInvertLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                            const float2* box, float2* tmp, float2* rbuf, int nplanes,
                            hipStream_t s) {
  const int nlines = dp.BV * nplanes;
  if (nlines <= 0) return;
  WG_CHECK(size_t(fm.n) * R == size_t(N), "long FFT split does not match the line");
  Dispatch(fm.n, [&](auto B, auto Mx) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(Mx)::value;
    AllowLds(RowFFTInvertLongKernel<BLOCK, MAXN>, LineLds(MAXN));
    RowFFTInvertLongKernel<BLOCK, MAXN><<<dim3(unsigned(nlines) * R), BLOCK, LineLds(fm.n), s>>>(
        dp, fm, N, R, box, tmp);
  });
  const size_t n = size_t(nlines) * dp.nx;
  RowCombineInvertKernel<<<CombineBlocks(n), 256, 0, s>>>(dp, N, R, fm.n, tmp, rbuf, nlines);
  WG_HIP(hipGetLastError());
}

void LaunchColFFTInvertLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                            const float2* cbuf, float2* tmp, float* dirty, int plane_lo,
                            int nplanes, hipStream_t s) {
  if (nplanes <= 0) return;
  WG_CHECK(size_t(fm.n) * R == size_t(N), "long FFT split does not match the line");
  const unsigned lines = unsigned(nplanes) * unsigned(dp.nx);
  Dispatch(fm.n, [&](auto B, auto Mx) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(Mx)::value;
    AllowLds(ColFFTInvertLongKernel<BLOCK, MAXN>, LineLds(MAXN));
    ColFFTInvertLongKernel<BLOCK, MAXN><<<dim3(lines * R), BLOCK, LineLds(fm.n), s>>>(
        dp, fm, N, R, cbuf, tmp);
  });
  const size_t n = size_t(dp.nx) * dp.ny;
  ColCombineInvertKernel<<<CombineBlocks(n), 256, 0, s>>>(dp, N, R, fm.n, tmp, dirty, plane_lo,
                                                          nplanes);
  WG_HIP(hipGetLastError());
}

void LaunchColFFTPredictLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                             const double* img, const float* corr, float2* tmp, float2* cbuf,
                             int plane_lo, int nplanes, hipStream_t s) {
  if (nplanes <= 0) return;
  WG_CHECK(size_t(fm.n) * R == size_t(N), "long FFT split does not match the line");
  const unsigned lines = unsigned(nplanes) * unsigned(dp.nx);
  Dispatch(fm.n, [&](auto B, auto Mx) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(Mx)::value;
    AllowLds(ColFFTPredictLongKernel<BLOCK, MAXN>, LineLds(MAXN));
    ColFFTPredictLongKernel<BLOCK, MAXN><<<dim3(lines * R), BLOCK, LineLds(fm.n), s>>>(
        dp, fm, N, R, img, corr, tmp, plane_lo);
  });
  const size_t n = size_t(nplanes) * dp.nx * dp.BV;
  ColCombinePredictKernel<<<CombineBlocks(n), 256, 0, s>>>(dp, N, R, fm.n, tmp, cbuf, nplanes);
  WG_HIP(hipGetLastError());
}

void LaunchRowFFTPredictLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                             const float2* rbuf, float2* tmp, float2* box, int nplanes,
                             hipStream_t s) {
  const int nlines = dp.BV * nplanes;
  if (nlines <= 0) return;
  WG_CHECK(size_t(fm.n) * R == size_t(N), "long FFT split does not match the line");
  Dispatch(fm.n, [&](auto B, auto Mx) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(Mx)::value;
    AllowLds(RowFFTPredictLongKernel<BLOCK, MAXN>, LineLds(MAXN));
    RowFFTPredictLongKernel<BLOCK, MAXN><<<dim3(unsigned(nlines) * R), BLOCK, LineLds(fm.n),
                                           s>>>(dp, fm, N, R, rbuf, tmp);
  });
  const size_t n = size_t(nlines) * dp.BU;
  RowCombinePredictKernel<<<CombineBlocks(n), 256, 0, s>>>(dp, N, R, fm.n, tmp, box, nlines);
  WG_HIP(hipGetLastError());
}

void LaunchLineFFTLong(float2* data, int nlines, int N, int R, const FFTDesc& fm, float2* tmp,
                       bool inverse, hipStream_t s) {
  WG_CHECK(size_t(fm.n) * R == size_t(N), "long FFT split does not match the line");
  Dispatch(fm.n, [&](auto B, auto Mx) {
    constexpr int BLOCK = decltype(B)::value, MAXN = decltype(Mx)::value;
    AllowLds(LineFFTLongKernel<BLOCK, MAXN>, LineLds(MAXN));
    LineFFTLongKernel<BLOCK, MAXN><<<dim3(unsigned(nlines) * R), BLOCK, LineLds(fm.n), s>>>(
        data, fm, N, R, tmp, inverse ? 1 : 0);
  });
  const size_t n = size_t(nlines) * N;
  LineCombineKernel<<<CombineBlocks(n), 256, 0, s>>>(tmp, N, R, fm.n, data, nlines,
                                                     inverse ? 1 : 0);
  WG_HIP(hipGetLastError());
}

}  // namespace wg
