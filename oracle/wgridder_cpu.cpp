// oracle/wgridder_cpu.cpp -- FP64 CPU w-gridder.  TEST INFRASTRUCTURE
// ONLY: used by tests/ as a mid-size checker and by bench.py's cpu_baseline
// leg as the timed CPU baseline ("CPU restatement, not ducc0").
//
// Restates the w-gridding algorithm that WSClean reaches through ducc0's
// ms2dirty / dirty2ms (called at wgridder/gridder_simple.h:217-225 and
// wgridder/gridder_simple_implementation.h:186-194; ducc0 itself is absent,
// so this follows the published method of Arras et al. 2021 / Ye et al.
// 2022: ES-kernel gridding in u, v and w, one 2-D FFT per w plane, the
// w screen exp(-2 pi i w_p (n - 1)), and division by the kernel's Fourier
// transform), with the image conventions of DirectMSGridder
// (gridding/directmsgridder.cpp:92-132).  Plane loop, per plane:
//   invert : grid the plane's visibilities into the occupied uv box, FFT the
//            box rows (zero-padded to nu), FFT the image columns (nv), apply
//            the screen and accumulate the cropped image;
//   predict: the adjoint in reverse order.
// Precision is FP64 throughout (WGriddingMSGridder::MakeGridder selects
// double for epsilon <= 1.01e-5, wgridder/wgriddingmsgridder.cpp:43-48).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

// Wall time of the last invert / predict split into the work that scales
// with the visibilities (set-up, shifts, gridding / degridding) and the
// work per w plane (FFTs, screens, crops): the CPU baseline extrapolates a
// row sample to the full visibility count with the plane sweep unchanged.
double g_t_vis = 0, g_t_plane = 0;
int g_planes = 0;
using Clock = std::chrono::steady_clock;
double Since(Clock::time_point t) {
  return std::chrono::duration<double>(Clock::now() - t).count();
}

using cd = std::complex<double>;
constexpr double kC = 299792458.0;

void ParallelFor(int nthreads, size_t n, const std::function<void(size_t, size_t, int)>& f) {
  if (nthreads <= 1 || n < 2) {
    f(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    size_t a = t * per, b = std::min(n, a + per);
    if (a >= b) break;
    th.emplace_back(f, a, b, t);
  }
  for (auto& x : th) x.join();
}

// Iterative Stockham FFT, radix 2/3/4/5, FP64.  sign = -1 forward.
struct Fft {
  int n;
  std::vector<int> rad;
  std::vector<cd> tw;
  explicit Fft(int n_) : n(n_) {
    int m = n;
    while (m % 4 == 0) { rad.push_back(4); m /= 4; }
    while (m % 2 == 0) { rad.push_back(2); m /= 2; }
    while (m % 3 == 0) { rad.push_back(3); m /= 3; }
    while (m % 5 == 0) { rad.push_back(5); m /= 5; }
    tw.resize(n);
    for (int k = 0; k < n; ++k) tw[k] = std::polar(1.0, -2.0 * M_PI * k / n);
  }
  void run(cd* x, cd* scratch, int sign) const {
    cd* a = x;
    cd* b = scratch;
    int Ns = 1;
    const double sg = double(sign);
    for (int R : rad) {
      const int nb = n / R, tws = n / (Ns * R);
      cd wr[5];
      for (int e = 0; e < R; ++e) wr[e] = std::polar(1.0, sign * 2.0 * M_PI * e / R);
      for (int j = 0; j < nb; ++j) {
        const int k = j % Ns;
        cd v[5];
        v[0] = a[j];
        for (int r = 1; r < R; ++r) {
          const cd w = tw[k * r * tws];
          v[r] = a[j + r * nb] * (sign > 0 ? std::conj(w) : w);
        }
        cd y[5];
        if (R == 4) {
          const cd t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
          const cd d = v[1] - v[3];
          const cd t3(-sg * d.imag(), sg * d.real());  // sign * i * d
          y[0] = t0 + t2;
          y[2] = t0 - t2;
          y[1] = t1 + t3;
          y[3] = t1 - t3;
        } else if (R == 2) {
          y[0] = v[0] + v[1];
          y[1] = v[0] - v[1];
        } else {
          for (int q = 0; q < R; ++q) {
            cd s = v[0];
            for (int r = 1; r < R; ++r) s += v[r] * wr[(q * r) % R];
            y[q] = s;
          }
        }
        const int base = (j - k) * R + k;
        for (int r = 0; r < R; ++r) b[base + r * Ns] = y[r];
      }
      std::swap(a, b);
      Ns *= R;
    }
    if (a != x) std::memcpy(x, a, sizeof(cd) * n);
  }
};

double EsKernel(double x, double beta) {
  const double a = 1.0 - x * x;
  return a > 0 ? std::exp(beta * (std::sqrt(a) - 1.0)) : 0.0;
}

// Phi(t) by Gauss-Legendre in theta (x = sin theta).
struct PhiHat {
  std::vector<double> a, b;
  PhiHat(int W, double beta, int nq = 128) {
    std::vector<double> x(nq), w(nq);
    for (int i = 0; i < (nq + 1) / 2; ++i) {
      double z = std::cos(M_PI * (i + 0.75) / (nq + 0.5)), pp = 0;
      for (int it = 0; it < 100; ++it) {
        double p1 = 1, p2 = 0;
        for (int j = 1; j <= nq; ++j) {
          double p3 = p2;
          p2 = p1;
          p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = nq * (z * p1 - p2) / (z * z - 1.0);
        double z1 = z;
        z = z1 - p1 / pp;
        if (std::fabs(z - z1) < 1e-15) break;
      }
      x[i] = -z;
      x[nq - 1 - i] = z;
      w[i] = w[nq - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    for (int q = 0; q < nq; ++q) {
      const double th = 0.5 * M_PI * x[q], c = std::cos(th);
      a.push_back(0.25 * M_PI * W * w[q] * std::exp(beta * (c - 1.0)) * c);
      b.push_back(M_PI * W * std::sin(th));
    }
  }
  double operator()(double t) const {
    double s = 0;
    for (size_t q = 0; q < a.size(); ++q) s += a[q] * std::cos(b[q] * t);
    return s;
  }
};

struct Problem {
  int nx, ny, W, nu, nv, nthreads;
  double dx, dy, ls, ms, beta, sigma;
  size_t n_rows, n_chan;
  const double* uvw;
  const double* freq;
  // derived
  double zmin, zmax, zc, dw, w_lo;
  int P;
  struct Vis {
    int a0u, a0v, p0;
    double fu, fv, fw;
    bool flip;
    double u, v, w;
    size_t idx;
  };
  std::vector<Vis> vis;   // sorted by p0
  std::vector<size_t> p0_off;
  int u_lo, v_lo, BU, BV;
  std::vector<double> corr;  // [x * ny + y]

  static double NM1(double l, double m) {
    const double r2 = l * l + m * m;
    return r2 < 1.0 ? -r2 / (std::sqrt(1.0 - r2) + 1.0) : 0.0;
  }
  double L(int x) const { return double(nx / 2 - x) * dx + ls; }
  double M(int y) const { return double(y - ny / 2) * dy + ms; }

  void Setup(const float* vis_in) {
    zmin = 0;
    zmax = -1;
    for (int x = 0; x < nx; x += std::max(1, nx - 1))
      for (int y = 0; y < ny; y += std::max(1, ny - 1)) {
        double z = NM1(L(x), M(y));
        zmin = std::min(zmin, z);
      }
    // the smallest r^2 on the image (closest point to the phase centre)
    {
      double lmin = std::min(L(0), L(nx - 1)), lmax = std::max(L(0), L(nx - 1));
      double mmin = std::min(M(0), M(ny - 1)), mmax = std::max(M(0), M(ny - 1));
      double lc = std::min(std::max(0.0, lmin), lmax), mc = std::min(std::max(0.0, mmin), mmax);
      zmax = NM1(lc, mc);
    }
    zc = 0.5 * (zmin + zmax);
    const double dz = zmax - zmin;
    std::vector<Vis> all;
    all.reserve(n_rows * n_chan);
    double wmin = 1e300, wmax = -1e300;
    for (size_t r = 0; r < n_rows; ++r)
      for (size_t c = 0; c < n_chan; ++c) {
        const size_t i = r * n_chan + c;
        if (vis_in && vis_in[2 * i] == 0.f && vis_in[2 * i + 1] == 0.f) continue;
        Vis v;
        const double s = freq[c] / kC;
        double um = uvw[3 * r], vm = uvw[3 * r + 1], wm = uvw[3 * r + 2];
        v.flip = wm < 0;
        if (v.flip) { um = -um; vm = -vm; wm = -wm; }
        v.u = um * s;
        v.v = vm * s;
        v.w = wm * s;
        v.idx = i;
        wmin = std::min(wmin, v.w);
        wmax = std::max(wmax, v.w);
        all.push_back(v);
      }
    dw = dz > 0 ? 1.0 / (sigma * dz) : std::max(1.0, 2.0 * (wmax - wmin) + 1.0);
    w_lo = all.empty() ? 0.0 : wmin;
    int p0max = 0, umin = 1 << 30, umax = -(1 << 30), vmin = 1 << 30, vmax = -(1 << 30);
    for (Vis& v : all) {
      const double tu = v.u * nu * dx + 0.5 * nu, tv = v.v * nv * dy + 0.5 * nv;
      const double tw = (v.w - w_lo) / dw + (W / 2 - 1);
      v.a0u = int(std::floor(tu)) - W / 2 + 1;
      v.a0v = int(std::floor(tv)) - W / 2 + 1;
      v.p0 = int(std::floor(tw)) - W / 2 + 1;
      v.fu = tu - std::floor(tu);
      v.fv = tv - std::floor(tv);
      v.fw = tw - std::floor(tw);
      p0max = std::max(p0max, v.p0);
      umin = std::min(umin, v.a0u);
      umax = std::max(umax, v.a0u);
      vmin = std::min(vmin, v.a0v);
      vmax = std::max(vmax, v.a0v);
    }
    P = p0max + W;
    // occupied box; visibilities outside [0, nu) wrap through the fold
    u_lo = all.empty() ? 0 : umin;
    v_lo = all.empty() ? 0 : vmin;
    BU = all.empty() ? 1 : umax - umin + W;
    BV = all.empty() ? 1 : vmax - vmin + W;
    // counting sort by p0
    p0_off.assign(P + 2, 0);
    for (const Vis& v : all) p0_off[v.p0 + 1]++;
    for (int p = 0; p <= P; ++p) p0_off[p + 1] += p0_off[p];
    vis.resize(all.size());
    std::vector<size_t> pos(p0_off.begin(), p0_off.end());
    for (const Vis& v : all) vis[pos[v.p0]++] = v;
    // correction 1 / (Phi_u Phi_v Phi_w) including the (-1)^(q_l + q_m) sign
    PhiHat ph(W, beta);
    std::vector<double> pu(nx), pv(ny);
    for (int x = 0; x < nx; ++x) pu[x] = ph(double(nx / 2 - x) / nu);
    for (int y = 0; y < ny; ++y) pv[y] = ph(double(y - ny / 2) / nv);
    // Phi_w on a fine table over the image's t = dw (n - 1 - zc) range,
    // linearly interpolated (relative error ~1e-9 with 2^16 intervals)
    const double tmax = std::max(1e-12, dw * 0.5 * (zmax - zmin) * 1.000001);
    const int ntab = 1 << 16;
    std::vector<double> tab(ntab + 1);
    ParallelFor(nthreads, ntab + 1, [&](size_t a, size_t b, int) {
      for (size_t k = a; k < b; ++k) tab[k] = ph(-tmax + 2.0 * tmax * double(k) / ntab);
    });
    zpix.resize(size_t(nx) * ny);
    corr.resize(size_t(nx) * ny);
    ParallelFor(nthreads, nx, [&](size_t a, size_t b, int) {
      for (size_t x = a; x < b; ++x)
        for (int y = 0; y < ny; ++y) {
          const double z = NM1(L(int(x)), M(y)) - zc;
          zpix[x * ny + y] = z;
          double f = (dw * z + tmax) / (2.0 * tmax) * ntab;
          f = std::min(std::max(f, 0.0), double(ntab) - 1e-9);
          const int k = int(f);
          const double pw = tab[k] + (f - k) * (tab[k + 1] - tab[k]);
          const int q = (nx / 2 - int(x)) + (y - ny / 2);
          corr[x * ny + y] = ((q & 1) ? -1.0 : 1.0) / (pu[x] * pv[y] * pw);
        }
    });
  }
  std::vector<double> zpix;  // [x * ny + y] n - 1 - zc
  double PlaneW(int p) const { return w_lo + (p - (W / 2 - 1)) * dw; }
  cd Shift(const Vis& v, double sign) const {
    const double cyc = v.u * ls + v.v * ms + v.w * zc;
    return std::polar(1.0, sign * 2.0 * M_PI * (cyc - std::round(cyc)));
  }
  void Taps(double f, double* out) const {
    for (int k = 0; k < W; ++k) out[k] = EsKernel((k - W / 2 + 1 - f) * 2.0 / W, beta);
  }
  int Wrap(int a, int n) const {
    int r = a % n;
    return r < 0 ? r + n : r;
  }
};

void Invert(Problem& pb, const float* vis_in, double* img_out /* [y][x] */) {
  g_t_vis = g_t_plane = 0;
  auto t0 = Clock::now();
  pb.Setup(vis_in);
  const int nx = pb.nx, ny = pb.ny, nu = pb.nu, nv = pb.nv, W = pb.W;
  const int T = pb.nthreads;
  std::vector<double> dirty(size_t(nx) * ny, 0.0);  // [x][y]
  std::vector<std::vector<cd>> boxes(T, std::vector<cd>(size_t(pb.BU) * pb.BV));
  std::vector<cd> rows(size_t(pb.BV) * nx);  // row FFT output: [box row][x]
  std::vector<cd> shifted(pb.vis.size());
  for (size_t i = 0; i < pb.vis.size(); ++i) {
    const auto& v = pb.vis[i];
    cd V(vis_in[2 * v.idx], vis_in[2 * v.idx + 1]);
    if (v.flip) V = std::conj(V);
    shifted[i] = V * pb.Shift(v, -1.0);
  }
  g_t_vis += Since(t0);
  Fft fu(nu), fv(nv);
  for (int p = 0; p < pb.P; ++p) {
    const size_t b0 = pb.p0_off[std::max(0, p - W + 1)], b1 = pb.p0_off[std::min(p, pb.P - W) + 1];
    t0 = Clock::now();
    ParallelFor(T, b1 - b0, [&](size_t a, size_t b, int t) {
      std::vector<cd>& box = boxes[t];
      std::fill(box.begin(), box.end(), cd(0));
      double ku[16], kv[16];
      for (size_t i = b0 + a; i < b0 + b; ++i) {
        const auto& v = pb.vis[i];
        pb.Taps(v.fu, ku);
        pb.Taps(v.fv, kv);
        const double kw = EsKernel((p - v.p0 - W / 2 + 1 - v.fw) * 2.0 / W, pb.beta);
        const cd vw = shifted[i] * kw;
        for (int l = 0; l < W; ++l) {
          cd* row = &box[size_t(v.a0v - pb.v_lo + l) * pb.BU + (v.a0u - pb.u_lo)];
          const cd vl = vw * kv[l];
          for (int k = 0; k < W; ++k) row[k] += vl * ku[k];
        }
      }
    });
    g_t_vis += Since(t0);
    t0 = Clock::now();
    // sum the thread boxes, FFT box rows and keep the cropped columns
    ParallelFor(T, pb.BV, [&](size_t a, size_t b, int) {
      std::vector<cd> line(nu), scr(nu);
      for (size_t j = a; j < b; ++j) {
        std::fill(line.begin(), line.end(), cd(0));
        for (int i = 0; i < pb.BU; ++i) {
          cd s = 0;
          for (int t = 0; t < T; ++t) s += boxes[t][j * pb.BU + i];
          line[pb.Wrap(pb.u_lo + i, nu)] += s;
        }
        fu.run(line.data(), scr.data(), -1);
        for (int x = 0; x < nx; ++x) rows[j * nx + x] = line[pb.Wrap(nx / 2 - x, nu)];
      }
    });
    const double wp = pb.PlaneW(p);
    ParallelFor(T, nx, [&](size_t a, size_t b, int) {
      std::vector<cd> line(nv), scr(nv);
      for (size_t x = a; x < b; ++x) {
        std::fill(line.begin(), line.end(), cd(0));
        for (int j = 0; j < pb.BV; ++j) line[pb.Wrap(pb.v_lo + j, nv)] += rows[j * nx + x];
        fv.run(line.data(), scr.data(), -1);
        for (int y = 0; y < ny; ++y) {
          const double cyc = wp * pb.zpix[x * ny + y];
          double sn, cs;
          sincos(2.0 * M_PI * (cyc - std::round(cyc)), &sn, &cs);
          const cd v = line[pb.Wrap(y - ny / 2, nv)];
          dirty[x * ny + y] += v.real() * cs + v.imag() * sn;
        }
      }
    });
    g_t_plane += Since(t0);
  }
  g_planes = pb.P;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      img_out[size_t(y) * nx + x] = dirty[size_t(x) * ny + y] * pb.corr[size_t(x) * ny + y];
}

void Predict(Problem& pb, const double* img /* [y][x] */, double* vis_out) {
  g_t_vis = g_t_plane = 0;
  auto t0 = Clock::now();
  pb.Setup(nullptr);
  g_t_vis += Since(t0);
  const int nx = pb.nx, ny = pb.ny, nu = pb.nu, nv = pb.nv, W = pb.W;
  const int T = pb.nthreads;
  std::vector<double> model(size_t(nx) * ny);
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      model[size_t(x) * ny + y] = img[size_t(y) * nx + x] * pb.corr[size_t(x) * ny + y];
  std::vector<cd> cols(size_t(nx) * pb.BV), box(size_t(pb.BU) * pb.BV);
  std::vector<cd> acc(pb.vis.size(), cd(0));
  Fft fu(nu), fv(nv);
  for (int p = 0; p < pb.P; ++p) {
    const double wp = pb.PlaneW(p);
    t0 = Clock::now();
    ParallelFor(T, nx, [&](size_t a, size_t b, int) {
      std::vector<cd> line(nv), scr(nv);
      for (size_t x = a; x < b; ++x) {
        std::fill(line.begin(), line.end(), cd(0));
        for (int y = 0; y < ny; ++y) {
          const double cyc = wp * pb.zpix[x * ny + y];
          double sn, cs;
          sincos(2.0 * M_PI * (cyc - std::round(cyc)), &sn, &cs);
          line[pb.Wrap(y - ny / 2, nv)] = model[x * ny + y] * cd(cs, sn);
        }
        fv.run(line.data(), scr.data(), +1);
        for (int j = 0; j < pb.BV; ++j) cols[x * pb.BV + j] = line[pb.Wrap(pb.v_lo + j, nv)];
      }
    });
    ParallelFor(T, pb.BV, [&](size_t a, size_t b, int) {
      std::vector<cd> line(nu), scr(nu);
      for (size_t j = a; j < b; ++j) {
        std::fill(line.begin(), line.end(), cd(0));
        for (int x = 0; x < nx; ++x) line[pb.Wrap(nx / 2 - x, nu)] = cols[size_t(x) * pb.BV + j];
        fu.run(line.data(), scr.data(), +1);
        for (int i = 0; i < pb.BU; ++i) box[j * pb.BU + i] = line[pb.Wrap(pb.u_lo + i, nu)];
      }
    });
    g_t_plane += Since(t0);
    t0 = Clock::now();
    const size_t b0 = pb.p0_off[std::max(0, p - W + 1)], b1 = pb.p0_off[std::min(p, pb.P - W) + 1];
    ParallelFor(T, b1 - b0, [&](size_t a, size_t b, int) {
      double ku[16], kv[16];
      for (size_t i = b0 + a; i < b0 + b; ++i) {
        const auto& v = pb.vis[i];
        pb.Taps(v.fu, ku);
        pb.Taps(v.fv, kv);
        const double kw = EsKernel((p - v.p0 - W / 2 + 1 - v.fw) * 2.0 / W, pb.beta);
        cd s = 0;
        for (int l = 0; l < W; ++l) {
          const cd* row = &box[size_t(v.a0v - pb.v_lo + l) * pb.BU + (v.a0u - pb.u_lo)];
          cd r = 0;
          for (int k = 0; k < W; ++k) r += row[k] * ku[k];
          s += r * kv[l];
        }
        acc[i] += s * kw;
      }
    });
    g_t_vis += Since(t0);
  }
  g_planes = pb.P;
  for (size_t i = 0; i < pb.vis.size(); ++i) {
    const auto& v = pb.vis[i];
    cd V = acc[i] * pb.Shift(v, +1.0);
    if (v.flip) V = std::conj(V);
    vis_out[2 * v.idx] = V.real();
    vis_out[2 * v.idx + 1] = V.imag();
  }
}

Problem MakeProblem(int nx, int ny, double dx, double dy, double ls, double ms, int W,
                    double beta, double sigma, int nu, int nv, size_t n_rows, size_t n_chan,
                    const double* uvw, const double* freq, int nthreads) {
  Problem pb{};
  pb.nx = nx;
  pb.ny = ny;
  pb.dx = dx;
  pb.dy = dy;
  pb.ls = ls;
  pb.ms = ms;
  pb.W = W;
  pb.beta = beta;
  pb.sigma = sigma;
  pb.nu = nu;
  pb.nv = nv;
  pb.n_rows = n_rows;
  pb.n_chan = n_chan;
  pb.uvw = uvw;
  pb.freq = freq;
  pb.nthreads = std::max(1, nthreads);
  return pb;
}

}  // namespace

extern "C" {

// Dirty image out[y * nx + x] (not divided by the image weight).  Zero
// visibilities are skipped.  Returns the number of w planes used.
int oracle_wgrid_invert(int nx, int ny, double dx, double dy, double ls, double ms, int W,
                        double beta, double sigma, int nu, int nv, size_t n_rows, size_t n_chan,
                        const double* uvw, const double* freq, const float* vis, double* out,
                        int nthreads) {
  Problem pb = MakeProblem(nx, ny, dx, dy, ls, ms, W, beta, sigma, nu, nv, n_rows, n_chan, uvw,
                           freq, nthreads);
  Invert(pb, vis, out);
  return pb.P;
}

// Wall seconds of the last invert or predict: out3 = {visibility work,
// plane work, planes}.
void oracle_wgrid_last_timing(double* out3) {
  out3[0] = g_t_vis;
  out3[1] = g_t_plane;
  out3[2] = double(g_planes);
}

// Visibilities out[2 * (row * n_chan + chan) + {0, 1}] from img[y * nx + x].
int oracle_wgrid_predict(int nx, int ny, double dx, double dy, double ls, double ms, int W,
                         double beta, double sigma, int nu, int nv, size_t n_rows, size_t n_chan,
                         const double* uvw, const double* freq, const double* img, double* out,
                         int nthreads) {
  Problem pb = MakeProblem(nx, ny, dx, dy, ls, ms, W, beta, sigma, nu, nv, n_rows, n_chan, uvw,
                           freq, nthreads);
  Predict(pb, img, out);
  return pb.P;
}

}  // extern "C"
