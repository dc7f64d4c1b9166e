// Shared definitions for the MI355X w-gridder: plan description, error
// handling and small device helpers.  Everything here is plain C++ / HIP for
// gfx950; no CUDA shims and no alternative backends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace wg {

constexpr double kSpeedOfLight = 299792458.0;

// Thrown by host code; the C-ABI converts it into a status code + message.
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define WG_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      throw ::wg::Error(std::string("HIP error ") + hipGetErrorString(e_) + \
                        " at " + __FILE__ + ":" + std::to_string(__LINE__) + \
                        " (" #call ")");                                    \
  } while (0)

#define WG_CHECK(cond, msg)                      \
  do {                                           \
    if (!(cond)) throw ::wg::Error(std::string(msg)); \
  } while (0)

// Piecewise-polynomial form of the ES kernel used by the gridding kernels:
// tap k of W is phi(x) at x = (k - W/2 + 1 - f) 2 / W as a polynomial of
// degree EsPolyDegree(W) in g = 2 f - 1 (f in [0, 1), the fractional grid
// coordinate).  Degrees are the smallest whose interpolation error is below
// 3e-7 of phi's peak (W <= 8: the sqrt singularity at the support edge
// limits it to ~e^-beta) or 1e-8 (W >= 10).
constexpr int kEsPolyTaps = 16;
constexpr int kEsPolyCoefs = 12;
constexpr int EsPolyDegree(int W) { return W <= 8 ? 7 : (W <= 10 ? 8 : (W <= 12 ? 9 : 10)); }
// out[k][d]: coefficient of g^d of tap k (rows k >= W and d > degree are 0),
// from Chebyshev interpolation of phi in FP64, rounded to FP32.
void EsPolyCoefficients(int W, double beta, float (*out)[kEsPolyCoefs]);

// One row of the ES kernel parameter table (gen_es_table.py).
struct KernelShape {
  int W;          // support in cells (even)
  double sigma;   // oversampling factor
  double beta;    // ES shape parameter
  double err1d;   // max 1-D RMS aliasing error over the image
  double amp_rms; // RMS over the image of the correction Phi(0)/Phi(t)
  double amp_max; // its maximum (image edge)
};

// Expected relative RMS error of one shape: aliasing in 3 dimensions with a
// factor-3 margin, plus the FP32 rounding of grid, FFT and accumulation
// amplified by the grid correction in u, v and w, plus a floor of FP32
// rounding that no shape removes.  The aliasing margin covers weighted
// data: uniformly weighted visibilities (dirty beams with strong far
// sidelobes) were measured at up to 4.4 x err1d against the direct FT,
// naturally weighted ones at up to 0.5 x err1d.  The FP32 terms are fitted
// to the worst of invert and predict against the FP64 direct FT on a 96^2
// field, a 4 degree 512^2 field and the 4096^2 precision study, with the
// kernel evaluated as its per-tap polynomial: W 10 .. 16 at amp_rms^3 =
// 100 .. 140 measured 4e-7 .. 3.2e-6 (2.3e-8 amp^3 at most), W 10 at sigma
// 1.8 / 1.9 measured 1.8e-7 .. 3.6e-7, W 8 at sigma 1.9 measured 2.8e-7.
constexpr double kFp32AmpTerm = 2.5e-8;
constexpr double kFp32Floor = 1e-7;
inline double ShapeErrorEstimate(const KernelShape& k) {
  return 3.0 * k.err1d + kFp32AmpTerm * k.amp_rms * k.amp_rms * k.amp_rms + kFp32Floor;
}

// Expected relative RMS error of one shape (W <= 8) gridded and degridded
// with split-precision f16 matrix operations: the FP32 estimate plus the
// products' rounding (about 2^-21 of each product, amplified by the grid
// correction like the FP32 rounding).
constexpr double kSplitAmpTerm = 2.5e-8;
constexpr double kSplitFloor = 1e-7;
inline double SplitErrorEstimate(const KernelShape& k) {
  return ShapeErrorEstimate(k) + kSplitAmpTerm * k.amp_rms * k.amp_rms * k.amp_rms + kSplitFloor;
}

// Smallest epsilon any shape of the table reaches with FP32 grids.
double Fp32EpsilonFloor();

const std::vector<KernelShape>& KernelTable();

// Geometry of the image the gridder produces (the trimmed image of
// WGriddingGridder_Simple) and the accuracy request.
struct ImageGeometry {
  int nx = 0, ny = 0;        // width_t, height_t
  double dx = 0, dy = 0;     // pixel size (radians)
  double l_shift = 0, m_shift = 0;
  double epsilon = 1e-4;
};

// Coordinate extents of one visibility set after the Hermitian flip
// (w >= 0), in wavelengths.
struct VisExtent {
  double umin, umax, vmin, vmax, wmin, wmax;
};

// Everything the kernels need to know about one gridding problem.
struct Plan {
  ImageGeometry img;
  // kernel
  int W = 0;
  double sigma = 0, beta = 0, err1d = 0, err_est = 0;
  int nu = 0, nv = 0;          // padded grid (even, 2^a 3^b 5^c)
  double ufac = 0, vfac = 0;   // grid cells per wavelength: nu*dx, nv*dy
  // n - 1 over the image and the w axis
  double zmin = 0, zmax = 0, zc = 0;
  double w_lo = 0, dw = 1, inv_dw = 1;
  int P0 = 1;                  // number of distinct first planes p0
  int P = 1;                   // number of w planes (P0 + W - 1)
  // occupied uv box (grid storage coordinates) and its tiling
  int T = 32;                  // tile edge (cells)
  int box_u0 = 0, box_v0 = 0;  // grid coordinate of box column/row 0
  bool wrap_u = false, wrap_v = false;
  int ntu = 0, ntv = 0;        // tiles along u / v
  int BU = 0, BV = 0;          // box storage extent (ntu*T + W - 1, ...)
  // sort: key = (bucket * key_T + (dv >> key_shift)) * key_T + (du >> key_shift);
  // key_shift > 0 coarsens the first-cell order inside a bucket when the
  // (bucket, cell) keys would not fit 32 bits
  uint32_t nbuckets = 0;       // (tile, first plane) buckets
  uint32_t nkeys = 0;          // valid keys are [0, nkeys); nkeys = invalid
  int key_bits = 0;
  int key_shift = 0;
  int key_T = 0;               // ((T - 1) >> key_shift) + 1
  // plane batching
  int Sg = 4;                  // planes per gridding sub-slab (LDS bound)
  int Sf = 0;                  // planes per FFT slab (HBM bound)
  // W <= 8: grid and degrid with split-precision f16 matrix operations
  // (each FP32 operand as a pair of halves) instead of FP32 ones
  bool split = false;

  std::string describe() const;
};

// Tunables that select a plan (the autotuner fills these in).
struct PlanOptions {
  int force_W = 0;             // 0: choose from epsilon
  double force_sigma = 0;      // 0: choose
  int T = 0;                   // 0: default
  int Sg = 0;                  // 0: default
  size_t slab_bytes = size_t(8) << 30;  // HBM budget for one FFT slab
  double sigma_min = 1.1, sigma_max = 2.0;
  long long nvis_hint = 0;     // for the cost model
  int split = -1;              // -1: split precision where epsilon allows;
                               // 0: FP32 matrix operations; 1: split for W <= 8
};

// Largest number of w planes of one gridding sub-slab (its LDS subgrid).
constexpr int kMaxSubSlab = 32;
// LDS of a gridding workgroup beyond its subgrid (per-wave coordinate
// staging of the matrix-core kernel); the subgrid gets the rest of 160 KiB.
constexpr int kGridLdsReserve = 33 * 1024;
// Default LDS per gridding workgroup, so that two are resident per CU.
constexpr int kGridLdsTarget = 80 * 1024;

// Sets nbuckets, nkeys, key_bits, key_shift and key_T of a plan whose tile
// grid (ntu, ntv, T) and P0 are set.
void PlanSortKeys(Plan& p);

// Longest line the LDS FFT transforms in one pass; longer lines (up to
// 5 x this) are split into R interleaved subsequences of at most this.
constexpr int kMaxLdsFFT = 16384;
constexpr int kMaxFFT = 5 * kMaxLdsFFT;
int GoodFFTSize(int n);        // even 2^a 3^b 5^c >= n of least LDS-FFT work
std::vector<int> FFTRadices(int n);  // Stockham stage radices for n

// Compute the n-1 range of the image and the full plan.
Plan MakePlan(const ImageGeometry& g, const VisExtent& e,
              const PlanOptions& opt);

// Fourier transform of the ES kernel with support W evaluated at t cycles
// per grid cell (unnormalised: integral of phi over its support at t=0).
double EsKernelFT(int W, double beta, double t);

}  // namespace wg
