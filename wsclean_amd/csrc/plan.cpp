// Host-side planning: kernel shape, padded grid, w axis, occupied uv box and
// its tiling.  The arithmetic that decides which grid cell or w plane a
// visibility lands in is repeated bit-for-bit on the device (LocateVis),
// so everything here is plain IEEE double arithmetic without contraction.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <sstream>

#include "common.hpp"

#pragma STDC FP_CONTRACT OFF

namespace wg {

namespace {
const KernelShape kTable[] = {
#include "es_table.inc"
};

// Gauss-Legendre nodes/weights on [-1, 1] (Newton iteration on P_n).
void GaussLegendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
}
}  // namespace

const std::vector<KernelShape>& KernelTable() {
  static const std::vector<KernelShape> t(std::begin(kTable), std::end(kTable));
  return t;
}

void EsPolyCoefficients(int W, double beta, float (*out)[kEsPolyCoefs]) {
  WG_CHECK(W >= 2 && W <= kEsPolyTaps, "kernel support out of range");
  const int D = EsPolyDegree(W), N = D + 1;
  for (int k = 0; k < kEsPolyTaps; ++k)
    for (int d = 0; d < kEsPolyCoefs; ++d) out[k][d] = 0.f;
  for (int k = 0; k < W; ++k) {
    // values at the N Chebyshev points of the first kind in g
    std::vector<double> y(N), a(N, 0.0);
    for (int i = 0; i < N; ++i) {
      const double g = std::cos(M_PI * (i + 0.5) / N);
      const double f = 0.5 * (g + 1.0);
      const double x = (double(k - W / 2 + 1) - f) * 2.0 / double(W);
      y[i] = std::fabs(x) <= 1.0 ? std::exp(beta * (std::sqrt(1.0 - x * x) - 1.0)) : 0.0;
    }
    // Chebyshev coefficients of the interpolant
    for (int j = 0; j < N; ++j) {
      double sum = 0;
      for (int i = 0; i < N; ++i) sum += y[i] * std::cos(M_PI * j * (i + 0.5) / N);
      a[j] = (j == 0 ? 1.0 : 2.0) * sum / N;
    }
    // power basis in g: T_0 = 1, T_1 = g, T_{j+1} = 2 g T_j - T_{j-1}
    std::vector<std::vector<double>> T(N, std::vector<double>(N, 0.0));
    T[0][0] = 1.0;
    if (N > 1) T[1][1] = 1.0;
    for (int j = 2; j < N; ++j)
      for (int d = 0; d < N; ++d)
        T[j][d] = (d > 0 ? 2.0 * T[j - 1][d - 1] : 0.0) - T[j - 2][d];
    std::vector<double> pw(N, 0.0);
    for (int j = 0; j < N; ++j)
      for (int d = 0; d < N; ++d) pw[d] += a[j] * T[j][d];
    for (int d = 0; d < N; ++d) out[k][d] = float(pw[d]);
  }
}

double Fp32EpsilonFloor() {
  double m = 1e300;
  for (const KernelShape& k : KernelTable()) m = std::min(m, ShapeErrorEstimate(k));
  return m;
}

double EsKernelFT(int W, double beta, double t) {
  // Phi(t) = (W/2) * int_{-1}^{1} phi(x) cos(pi W x t) dx with x = sin(th):
  // the integrand exp(beta (cos th - 1)) cos(pi W t sin th) cos th is analytic.
  static std::once_flag once;
  static std::vector<double> gx, gw;
  std::call_once(once, [] { GaussLegendre(256, gx, gw); });
  double s = 0;
  for (size_t i = 0; i < gx.size(); ++i) {
    double th = 0.5 * M_PI * gx[i];
    double c = std::cos(th);
    s += gw[i] * std::exp(beta * (c - 1.0)) * c *
         std::cos(M_PI * W * t * std::sin(th));
  }
  return 0.5 * W * 0.5 * M_PI * s;
}

namespace {
bool Smooth235(int m) {
  for (int p : {2, 3, 5})
    while (m % p == 0) m /= p;
  return m == 1;
}
}  // namespace

namespace {
// Subsequences of a long line (as LongFFTRadix of the FFT code): 1 up to
// kMaxLdsFFT, else the smallest R in 2..5 with N / R <= kMaxLdsFFT, 0 if none.
int SplitRadix(int n) {
  if (n <= kMaxLdsFFT) return 1;
  for (int R = 2; R <= 5; ++R)
    if (n % R == 0 && n / R <= kMaxLdsFFT) return R;
  return 0;
}
}  // namespace

int GoodFFTSize(int n) {
  if (n < 2) n = 2;
  int first = n + (n & 1);
  while (!Smooth235(first) || SplitRadix(first) == 0) first += 2;
  // Among the even 2^a 3^b 5^c sizes up to 25 % above n, the one of least
  // LDS-FFT work: the plane FFTs touch about m lines of m points, each in
  // one LDS pass per Stockham stage plus the load, screen and store, so the
  // cost goes as m^2 (stages + 2); a line longer than kMaxLdsFFT adds a
  // pass through HBM for its R subsequences (+ 2 + R).
  int best = first;
  double best_cost = 1e300;
  const int last = std::max(first, std::min(kMaxFFT, int(1.25 * n)));
  for (int m = first; m <= last; m += 2) {
    const int R = Smooth235(m) ? SplitRadix(m) : 0;
    if (R == 0) continue;
    const double stages = double(FFTRadices(m / R).size() + 2 + (R > 1 ? 2 + R : 0));
    const double cost = double(m) * double(m) * stages;
    if (cost < best_cost) {
      best_cost = cost;
      best = m;
    }
  }
  return best;
}

std::vector<int> FFTRadices(int n) {
  std::vector<int> r;
  int m = n;
  while (m % 8 == 0) { r.push_back(8); m /= 8; }
  while (m % 9 == 0) { r.push_back(9); m /= 9; }
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  while (m % 2 == 0) { r.push_back(2); m /= 2; }
  while (m % 3 == 0) { r.push_back(3); m /= 3; }
  while (m % 5 == 0) { r.push_back(5); m /= 5; }
  WG_CHECK(m == 1, "FFT length " + std::to_string(n) + " is not 2^a 3^b 5^c");
  // a lone factor 2 and a 9 become one radix-18 stage (placed last): one
  // LDS pass less
  auto two = std::find(r.begin(), r.end(), 2);
  auto nine = std::find(r.begin(), r.end(), 9);
  if (two != r.end() && nine != r.end()) {
    r.erase(two);
    r.erase(std::find(r.begin(), r.end(), 9));
    r.push_back(18);
  }
  return r;
}

namespace {

// n - 1 = sqrt(1 - r2) - 1 written without cancellation; 0 beyond the
// horizon as in DirectMSGridder (gridding/directmsgridder.cpp:221-224).
double NMinusOne(double r2) {
  if (r2 >= 1.0) return 0.0;
  return -r2 / (std::sqrt(1.0 - r2) + 1.0);
}

void ZRange(const ImageGeometry& g, double& zmin, double& zmax) {
  // l(x) = (nx/2 - x) dx + l_shift, x in [0, nx);  m(y) = (y - ny/2) dy + m_shift
  double l_a = (g.nx / 2) * g.dx + g.l_shift;
  double l_b = (g.nx / 2 - (g.nx - 1)) * g.dx + g.l_shift;
  double m_a = (0 - g.ny / 2) * g.dy + g.m_shift;
  double m_b = ((g.ny - 1) - g.ny / 2) * g.dy + g.m_shift;
  double lmin = std::min(l_a, l_b), lmax = std::max(l_a, l_b);
  double mmin = std::min(m_a, m_b), mmax = std::max(m_a, m_b);
  double l2max = std::max(lmin * lmin, lmax * lmax);
  double m2max = std::max(mmin * mmin, mmax * mmax);
  double l2min = (lmin <= 0 && lmax >= 0) ? 0.0 : std::min(lmin * lmin, lmax * lmax);
  double m2min = (mmin <= 0 && mmax >= 0) ? 0.0 : std::min(mmin * mmin, mmax * mmax);
  double r2max = l2max + m2max, r2min = l2min + m2min;
  zmin = NMinusOne(std::min(r2max, 0.999999999));
  zmax = NMinusOne(r2min);
  if (r2max >= 1.0) zmax = std::max(zmax, 0.0);
}

int A0(double t, int W) { return int(std::floor(t)) - W / 2 + 1; }

}  // namespace

Plan MakePlan(const ImageGeometry& g, const VisExtent& e, const PlanOptions& opt) {
  WG_CHECK(g.nx >= 8 && g.ny >= 8 && (g.nx % 2) == 0 && (g.ny % 2) == 0,
           "image sizes must be even and >= 8");
  WG_CHECK(g.dx > 0 && g.dy > 0, "pixel sizes must be positive");
  WG_CHECK(g.epsilon > 0, "epsilon must be positive");
  Plan p;
  p.img = g;
  ZRange(g, p.zmin, p.zmax);
  p.zc = 0.5 * (p.zmin + p.zmax);
  const double dz = p.zmax - p.zmin;
  const double wspan = std::max(0.0, e.wmax - e.wmin);

  // ---- choose (W, sigma) by a cost model among shapes meeting epsilon ----
  const double nvis = opt.nvis_hint > 0 ? double(opt.nvis_hint) : 1e6;
  double best = 1e300;
  KernelShape chosen{0, 0, 0, 0};
  for (const KernelShape& k : KernelTable()) {
    if (opt.force_W && k.W != opt.force_W) continue;
    if (opt.force_sigma > 0 && std::fabs(k.sigma - opt.force_sigma) > 1e-9) continue;
    if (k.sigma < opt.sigma_min - 1e-9 || k.sigma > opt.sigma_max + 1e-9) continue;
    if (!opt.force_W && ShapeErrorEstimate(k) > g.epsilon) continue;
    int nu = GoodFFTSize(int(std::ceil(k.sigma * g.nx)));
    int nv = GoodFFTSize(int(std::ceil(k.sigma * g.ny)));
    double dw = dz > 0 ? 1.0 / (k.sigma * dz) : 1e300;
    double planes = std::floor(wspan / dw) + k.W;
    double band_u = std::min<double>(nu, (e.umax - e.umin) * nu * g.dx + k.W);
    double band_v = std::min<double>(nv, (e.vmax - e.vmin) * nv * g.dy + k.W);
    // per W^3 visibility pass, relative to the FFT units below (measured
    // on the 8192^2 workload: matrix-core grid + degrid for W <= 8 take
    // 0.46 FFT units, the VALU kernels for W >= 10 about 2.2)
    const double grid_weight = k.W <= 8 ? 0.46 : 2.2;
    double grid_cost = grid_weight * nvis * k.W * k.W * k.W;
    double fft_cost = planes * (band_v * nu * std::log2(nu) +
                                g.nx * nv * std::log2(nv) + 4.0 * g.nx * g.ny +
                                band_u * band_v) * 4.0;
    double cost = grid_cost + fft_cost;
    if (cost < best) {
      best = cost;
      chosen = k;
    }
  }
  if (chosen.W == 0) {
    char msg[320];
    std::snprintf(msg, sizeof(msg),
                  "epsilon %.3g is below what the FP32 grids and FFT of this library reach "
                  "(%.2g with any kernel shape); the reference switches to FP64 gridding "
                  "there, which this library does not provide",
                  g.epsilon, Fp32EpsilonFloor());
    WG_CHECK(g.epsilon >= Fp32EpsilonFloor(), msg);
    WG_CHECK(false, "no kernel shape reaches epsilon " + std::to_string(g.epsilon) +
                        " within the oversampling limits");
  }
  p.W = chosen.W;
  p.sigma = chosen.sigma;
  p.beta = chosen.beta;
  p.err1d = chosen.err1d;
  p.err_est = ShapeErrorEstimate(chosen);
  // split-precision matrix gridding for W <= 8 where its error still meets
  // epsilon (SplitErrorEstimate), unless the options force either form
  p.split = p.W <= 8 && (opt.split == 1 ||
                         (opt.split < 0 && SplitErrorEstimate(chosen) <= g.epsilon));
  if (p.split) p.err_est = std::max(p.err_est, SplitErrorEstimate(chosen));
  p.nu = GoodFFTSize(int(std::ceil(p.sigma * g.nx)));
  p.nv = GoodFFTSize(int(std::ceil(p.sigma * g.ny)));
  WG_CHECK(p.nu <= kMaxFFT && p.nv <= kMaxFFT && SplitRadix(p.nu) && SplitRadix(p.nv),
           "padded grid larger than " + std::to_string(kMaxFFT) +
               " is not supported by the FFT");
  p.ufac = double(p.nu) * g.dx;
  p.vfac = double(p.nv) * g.dy;

  // ---- w axis ----
  p.dw = dz > 0 ? 1.0 / (p.sigma * dz) : std::max(1.0, 2.0 * wspan + 1.0);
  p.inv_dw = 1.0 / p.dw;
  p.w_lo = e.wmin;
  {
    double t = (e.wmax - p.w_lo) * p.inv_dw + double(p.W / 2 - 1);
    int p0max = A0(t, p.W);
    p.P0 = std::max(1, p0max + 1);
    p.P = p.P0 + p.W - 1;
  }

  // ---- occupied uv box ----
  p.T = opt.T > 0 ? opt.T : 8;
  WG_CHECK(p.T >= 8 && p.T <= 128, "tile size must be in [8, 128]");
  auto box = [&](double cmin, double cmax, double fac, int n, int& origin,
                 bool& wrap, int& ntiles, int& ext) {
    double tmin = cmin * fac + 0.5 * n;
    double tmax = cmax * fac + 0.5 * n;
    int lo = A0(tmin, p.W);
    int hi = A0(tmax, p.W) + p.W - 1;
    int len;
    if (lo >= 0 && hi < n) {
      wrap = false;
      origin = lo;
      len = hi - lo + 1;
    } else {
      wrap = true;
      origin = 0;
      len = n + p.W - 1;
    }
    ntiles = (len - p.W + 1 + p.T - 1) / p.T;
    if (ntiles < 1) ntiles = 1;
    ext = ntiles * p.T + p.W - 1;
  };
  box(e.umin, e.umax, p.ufac, p.nu, p.box_u0, p.wrap_u, p.ntu, p.BU);
  box(e.vmin, e.vmax, p.vfac, p.nv, p.box_v0, p.wrap_v, p.ntv, p.BV);
  WG_CHECK(p.BU <= 2 * p.nu && p.BV <= 2 * p.nv, "uv box larger than twice the grid");

  PlanSortKeys(p);

  // ---- plane batching ----
  // default: as many planes as keep a gridding workgroup (subgrid + staging)
  // within half the LDS of a CU, so two are resident; deeper sub-slabs
  // re-read each visibility fewer times
  {
    const size_t E = size_t(p.T + p.W - 1);
    const int fit = int((kGridLdsTarget - kGridLdsReserve) / (E * E * 8));
    p.Sg = opt.Sg > 0 ? opt.Sg : std::max(1, fit);
  }
  p.Sg = std::min(p.Sg, kMaxSubSlab);
  {
    // the gridding subgrid (Sg planes of (T + W - 1)^2 complex) and the
    // staging reserve must fit the 160 KiB of LDS of one CU
    const size_t E = size_t(p.T + p.W - 1);
    const int max_sg = int((160 * 1024 - kGridLdsReserve) / (E * E * 8));
    WG_CHECK(max_sg >= 1, "tile too large for LDS");
    p.Sg = std::max(1, std::min(p.Sg, max_sg));
    // when every plane fits one sub-slab of the full LDS, take one sub-slab
    // at one workgroup per CU: the slab gather and scatter then fuse into
    // the FFT passes, which saves more than the second workgroup gains
    if (opt.Sg <= 0 && p.P > p.Sg && p.P <= max_sg && p.P <= kMaxSubSlab) p.Sg = p.P;
  }
  // box + row and column line buffers, + the subsequence spectra of a line
  // longer than the LDS FFT
  double per_plane = 8.0 * p.BV * (double(p.BU) + 2.0 * g.nx);
  if (SplitRadix(p.nu) > 1 || SplitRadix(p.nv) > 1)
    per_plane += 8.0 * std::max(double(p.BV) * p.nu, double(g.nx) * p.nv);
  int sf = int(std::max(1.0, std::floor(double(opt.slab_bytes) / per_plane)));
  sf = std::max(p.Sg, (sf / p.Sg) * p.Sg);
  p.Sf = std::min(sf, ((p.P + p.Sg - 1) / p.Sg) * p.Sg);
  return p;
}

void PlanSortKeys(Plan& p) {
  // sort key = ((tile * P0 + p0) * T + dv) * T + du: buckets of one tile and
  // first plane, ordered by the exact first cell inside the bucket (long
  // runs of one cell for the gridding kernels).  When those keys exceed 32
  // bits, the cell order inside a bucket is coarsened to (dv >> s, du >> s)
  // cells: still one bucket per (tile, first plane), only the runs of one
  // exact cell get shorter.
  const uint64_t nb = uint64_t(p.ntu) * uint64_t(p.ntv) * uint64_t(p.P0);
  WG_CHECK(nb < 0xFFFFFFF0ull,
           "more than 2^32 (uv tile, w plane) buckets: use a larger tile (wg_set_tuning)");
  int s = 0;
  uint64_t kt = uint64_t(p.T);
  while (nb * kt * kt >= 0xFFFFFFF0ull) {
    ++s;
    kt = uint64_t((p.T - 1) >> s) + 1;
  }
  p.key_shift = s;
  p.key_T = int(kt);
  p.nbuckets = uint32_t(nb);
  p.nkeys = uint32_t(nb * kt * kt);
  p.key_bits = 1;
  while ((uint64_t(1) << p.key_bits) <= uint64_t(p.nkeys)) ++p.key_bits;
}

std::string Plan::describe() const {
  std::ostringstream s;
  s.precision(17);
  s << "{\"W\":" << W << ",\"sigma\":" << sigma << ",\"beta\":" << beta
    << ",\"err1d\":" << err1d << ",\"err_est\":" << err_est << ",\"nu\":" << nu
    << ",\"nv\":" << nv
    << ",\"P\":" << P << ",\"P0\":" << P0 << ",\"dw\":" << dw
    << ",\"w_lo\":" << w_lo << ",\"zmin\":" << zmin << ",\"zmax\":" << zmax
    << ",\"T\":" << T << ",\"ntu\":" << ntu << ",\"ntv\":" << ntv
    << ",\"BU\":" << BU << ",\"BV\":" << BV << ",\"box_u0\":" << box_u0
    << ",\"box_v0\":" << box_v0 << ",\"wrap_u\":" << (wrap_u ? 1 : 0)
    << ",\"wrap_v\":" << (wrap_v ? 1 : 0) << ",\"key_bits\":" << key_bits
    << ",\"Sg\":" << Sg << ",\"Sf\":" << Sf << ",\"dx\":" << img.dx
    << ",\"dy\":" << img.dy << ",\"zc\":" << zc << ",\"nkeys\":" << nkeys
    << ",\"nbuckets\":" << nbuckets << ",\"key_shift\":" << key_shift
    << ",\"key_T\":" << key_T << ",\"split\":" << (split ? 1 : 0) << "}";
  return s.str();
}

}  // namespace wg
