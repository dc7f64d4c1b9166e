/*
 * oracle/dft.c -- FP64 direct Fourier transform oracle.  TEST
 * INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never part of the product.
 *
 * Restates WSClean's DirectMSGridder (gridding/directmsgridder.cpp):
 *   - per-sample contribution, gridSample():            :92-132
 *       angle = -2 pi (u l + v m + w (sqrt(1 - l^2 - m^2) - 1))
 *       I += Re(V) cos(angle) - Im(V) sin(angle)
 *   - the n - 1 table, GetSqrtLMLookupTable():            :205-229
 *       sqrt(1 - l^2 - m^2) - 1 inside the unit circle, 0 outside
 *   - the final image wrap, FinishInversion():             :56-89
 *       after the wrap, pixel (x, y) of the image (index x + y * width)
 *       has l = (width/2 - x) dx + l_shift, m = (y - height/2) dy + m_shift
 *       (even width / height).
 * u, v, w are in wavelengths: uvw [m] * freq / c.  The division by
 * ImageWeight() is left to the caller.
 *
 * The predict direction (DirectMSGridder has none, :184-202) is the adjoint:
 *   V(u, v, w) = sum_{x,y} I(x, y) exp(+2 pi i (u l + v m + w (n - 1)))
 * whose sign is fixed by WSClean's point-source generator, which writes
 * V = exp(+2 pi i (u dl + v dm + w dn)) (gridding/msgridderdata.h:820-843,
 * RotateVisibilities :929-941).
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdlib.h>

#define C_LIGHT 299792458.0

static double pixel_l(int x, int nx, double dx, double ls) {
  return (double)(nx / 2 - x) * dx + ls;
}
static double pixel_m(int y, int ny, double dy, double ms) {
  return (double)(y - ny / 2) * dy + ms;
}
static double n_minus_1(double l, double m) {
  double r2 = l * l + m * m;
  return r2 < 1.0 ? sqrt(1.0 - r2) - 1.0 : 0.0;
}

typedef struct {
  int nx, ny;
  double dx, dy, ls, ms;
  size_t n_rows, n_chan;
  const double* uvw;
  const double* freq;
  const float* vis;      /* complex64 as float pairs */
  const double* img;     /* predict input, x + y * nx */
  double* out;
  size_t npix;
  const int* px;
  const int* py;
  size_t nsel;
  const size_t* sel;     /* flat visibility indices */
  int tid, nthreads;
} job_t;

static void* invert_worker(void* arg) {
  job_t* j = (job_t*)arg;
  for (size_t p = (size_t)j->tid; p < j->npix; p += (size_t)j->nthreads) {
    int x, y;
    if (j->px) {
      x = j->px[p];
      y = j->py[p];
    } else {
      x = (int)(p % (size_t)j->nx);
      y = (int)(p / (size_t)j->nx);
    }
    const double l = pixel_l(x, j->nx, j->dx, j->ls);
    const double m = pixel_m(y, j->ny, j->dy, j->ms);
    const double nm1 = n_minus_1(l, m);
    double acc = 0.0;
    for (size_t r = 0; r < j->n_rows; ++r) {
      const double* uvw = j->uvw + 3 * r;
      for (size_t c = 0; c < j->n_chan; ++c) {
        const float* v = j->vis + 2 * (r * j->n_chan + c);
        if (v[0] == 0.0f && v[1] == 0.0f) continue;
        const double s = j->freq[c] / C_LIGHT;
        const double angle =
            -2.0 * M_PI * (uvw[0] * s * l + uvw[1] * s * m + uvw[2] * s * nm1);
        acc += (double)v[0] * cos(angle) - (double)v[1] * sin(angle);
      }
    }
    j->out[p] = acc;
  }
  return NULL;
}

static void* predict_worker(void* arg) {
  job_t* j = (job_t*)arg;
  const size_t nvis = j->sel ? j->nsel : j->n_rows * j->n_chan;
  /* collect non-zero pixels once per thread */
  size_t np = (size_t)j->nx * (size_t)j->ny, nnz = 0;
  for (size_t p = 0; p < np; ++p)
    if (j->img[p] != 0.0) ++nnz;
  double* pl = (double*)malloc(sizeof(double) * (nnz ? nnz : 1) * 4);
  size_t k = 0;
  for (size_t p = 0; p < np; ++p) {
    if (j->img[p] == 0.0) continue;
    int x = (int)(p % (size_t)j->nx), y = (int)(p / (size_t)j->nx);
    double l = pixel_l(x, j->nx, j->dx, j->ls), m = pixel_m(y, j->ny, j->dy, j->ms);
    pl[4 * k] = l;
    pl[4 * k + 1] = m;
    pl[4 * k + 2] = n_minus_1(l, m);
    pl[4 * k + 3] = j->img[p];
    ++k;
  }
  for (size_t q = (size_t)j->tid; q < nvis; q += (size_t)j->nthreads) {
    const size_t idx = j->sel ? j->sel[q] : q;
    const size_t r = idx / j->n_chan, c = idx % j->n_chan;
    const double s = j->freq[c] / C_LIGHT;
    const double u = j->uvw[3 * r] * s, v = j->uvw[3 * r + 1] * s, w = j->uvw[3 * r + 2] * s;
    double re = 0.0, im = 0.0;
    for (size_t t = 0; t < nnz; ++t) {
      const double angle = 2.0 * M_PI * (u * pl[4 * t] + v * pl[4 * t + 1] + w * pl[4 * t + 2]);
      re += pl[4 * t + 3] * cos(angle);
      im += pl[4 * t + 3] * sin(angle);
    }
    j->out[2 * q] = re;
    j->out[2 * q + 1] = im;
  }
  free(pl);
  return NULL;
}

static void run(job_t* base, void* (*fn)(void*), int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  job_t jobs[256];
  for (int t = 0; t < nthreads; ++t) {
    jobs[t] = *base;
    jobs[t].tid = t;
    jobs[t].nthreads = nthreads;
    pthread_create(&th[t], NULL, fn, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* Full dirty image, out[x + y * nx] (not divided by ImageWeight). */
void oracle_dft_invert(int nx, int ny, double dx, double dy, double l_shift, double m_shift,
                       size_t n_rows, size_t n_chan, const double* uvw, const double* freq,
                       const float* vis, double* out, int nthreads) {
  job_t j = {0};
  j.nx = nx; j.ny = ny; j.dx = dx; j.dy = dy; j.ls = l_shift; j.ms = m_shift;
  j.n_rows = n_rows; j.n_chan = n_chan; j.uvw = uvw; j.freq = freq; j.vis = vis;
  j.out = out; j.npix = (size_t)nx * (size_t)ny;
  run(&j, invert_worker, nthreads);
}

/* Dirty image at npix selected pixels (x[i], y[i]). */
void oracle_dft_invert_pixels(int nx, int ny, double dx, double dy, double l_shift,
                              double m_shift, size_t n_rows, size_t n_chan, const double* uvw,
                              const double* freq, const float* vis, size_t npix, const int* xs,
                              const int* ys, double* out, int nthreads) {
  job_t j = {0};
  j.nx = nx; j.ny = ny; j.dx = dx; j.dy = dy; j.ls = l_shift; j.ms = m_shift;
  j.n_rows = n_rows; j.n_chan = n_chan; j.uvw = uvw; j.freq = freq; j.vis = vis;
  j.out = out; j.npix = npix; j.px = xs; j.py = ys;
  run(&j, invert_worker, nthreads);
}

/* Predicted visibilities out[2 * (row * n_chan + chan) + {0, 1}] from the
 * model image img[x + y * nx]; only non-zero pixels are summed.  With sel
 * non-NULL only the nsel flat indices sel[] are predicted, into out[2 * q]. */
void oracle_dft_predict(int nx, int ny, double dx, double dy, double l_shift, double m_shift,
                        const double* img, size_t n_rows, size_t n_chan, const double* uvw,
                        const double* freq, size_t nsel, const size_t* sel, double* out,
                        int nthreads) {
  job_t j = {0};
  j.nx = nx; j.ny = ny; j.dx = dx; j.dy = dy; j.ls = l_shift; j.ms = m_shift;
  j.n_rows = n_rows; j.n_chan = n_chan; j.uvw = uvw; j.freq = freq; j.img = img;
  j.out = out; j.nsel = nsel; j.sel = sel;
  run(&j, predict_worker, nthreads);
}

/*
 * Dirty image at selected pixels, as oracle_dft_invert_pixels, organised
 * for large visibility sets: threads split the rows, every thread keeps a
 * partial sum per pixel, and the partials are added in thread order.  For
 * one row and one pixel the angle is s_c * theta with s_c = freq[c] / c and
 * theta = -2 pi (u l + v m + w (n - 1)) (u, v, w in metres); when the
 * channels are evenly spaced (every s_c within 4 ulp of s_0 + c ds) the
 * phasors exp(i s_c theta) are formed by the recurrence
 * z_{c+1} = z_c exp(i ds theta) (relative rounding growth ~c * 1e-16),
 * otherwise by cos / sin per channel.  Same formula as invert_worker
 * (gridding/directmsgridder.cpp:92-132); zero visibilities contribute 0.
 */
typedef struct {
  const double* uvw;
  const float* vis;
  size_t n_chan, r0, r1, npix;
  const double* lmn;   /* npix x {l, m, n - 1} */
  double s0, ds;
  const double* s;     /* per-channel scale (fallback) */
  int uniform;
  int per_chan;        /* acc holds npix x n_chan sums (else npix) */
  double* acc;         /* partial sums */
} fast_job_t;

static void* invert_fast_worker(void* arg) {
  fast_job_t* j = (fast_job_t*)arg;
  for (size_t r = j->r0; r < j->r1; ++r) {
    const double u = j->uvw[3 * r], v = j->uvw[3 * r + 1], w = j->uvw[3 * r + 2];
    const float* vr = j->vis + 2 * r * j->n_chan;
    for (size_t p = 0; p < j->npix; ++p) {
      const double theta =
          -2.0 * M_PI * (u * j->lmn[3 * p] + v * j->lmn[3 * p + 1] + w * j->lmn[3 * p + 2]);
      double a = 0.0;
      double* ac = j->per_chan ? j->acc + p * j->n_chan : NULL;
      if (j->uniform) {
        double zr = cos(theta * j->s0), zi = sin(theta * j->s0);
        const double dr = cos(theta * j->ds), di = sin(theta * j->ds);
        for (size_t c = 0; c < j->n_chan; ++c) {
          const double term = (double)vr[2 * c] * zr - (double)vr[2 * c + 1] * zi;
          if (ac)
            ac[c] += term;
          else
            a += term;
          const double t = zr * dr - zi * di;
          zi = zr * di + zi * dr;
          zr = t;
        }
      } else {
        for (size_t c = 0; c < j->n_chan; ++c) {
          const double ang = theta * j->s[c];
          const double term = (double)vr[2 * c] * cos(ang) - (double)vr[2 * c + 1] * sin(ang);
          if (ac)
            ac[c] += term;
          else
            a += term;
        }
      }
      if (!ac) j->acc[p] += a;
    }
  }
  return NULL;
}

static void dft_invert_pixels_fast_impl(int nx, int ny, double dx, double dy, double l_shift,
                                        double m_shift, size_t n_rows, size_t n_chan,
                                        const double* uvw, const double* freq, const float* vis,
                                        size_t npix, const int* xs, const int* ys, double* out,
                                        int nthreads, int* used_recurrence, int per_chan) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  /* sums per thread: npix, or npix x n_chan per channel */
  const size_t nsum = (npix ? npix : 1) * (per_chan && n_chan ? n_chan : 1);
  double* lmn = (double*)malloc(sizeof(double) * 3 * (npix ? npix : 1));
  double* s = (double*)malloc(sizeof(double) * (n_chan ? n_chan : 1));
  double* acc = (double*)calloc((size_t)nthreads * nsum, sizeof(double));
  for (size_t p = 0; p < npix; ++p) {
    const double l = pixel_l(xs[p], nx, dx, l_shift), m = pixel_m(ys[p], ny, dy, m_shift);
    lmn[3 * p] = l;
    lmn[3 * p + 1] = m;
    lmn[3 * p + 2] = n_minus_1(l, m);
  }
  for (size_t c = 0; c < n_chan; ++c) s[c] = freq[c] / C_LIGHT;
  int uniform = n_chan > 1;
  const double ds = n_chan > 1 ? (s[n_chan - 1] - s[0]) / (double)(n_chan - 1) : 0.0;
  for (size_t c = 0; c < n_chan && uniform; ++c) {
    const double e = s[c] - (s[0] + (double)c * ds);
    if (fabs(e) > 4.0 * 2.220446049250313e-16 * fabs(s[c])) uniform = 0;
  }
  pthread_t th[256];
  fast_job_t jobs[256];
  for (int t = 0; t < nthreads; ++t) {
    fast_job_t* j = &jobs[t];
    j->uvw = uvw;
    j->vis = vis;
    j->n_chan = n_chan;
    j->r0 = n_rows * (size_t)t / (size_t)nthreads;
    j->r1 = n_rows * (size_t)(t + 1) / (size_t)nthreads;
    j->npix = npix;
    j->lmn = lmn;
    j->s0 = n_chan ? s[0] : 0.0;
    j->ds = ds;
    j->s = s;
    j->uniform = uniform;
    j->per_chan = per_chan;
    j->acc = acc + (size_t)t * nsum;
    pthread_create(&th[t], NULL, invert_fast_worker, j);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  const size_t nout = per_chan ? npix * n_chan : npix;
  for (size_t p = 0; p < nout; ++p) {
    double a = 0.0;
    for (int t = 0; t < nthreads; ++t) a += acc[(size_t)t * nsum + p];
    out[p] = a;
  }
  if (used_recurrence) *used_recurrence = uniform;
  free(acc);
  free(s);
  free(lmn);
}

void oracle_dft_invert_pixels_fast(int nx, int ny, double dx, double dy, double l_shift,
                                   double m_shift, size_t n_rows, size_t n_chan,
                                   const double* uvw, const double* freq, const float* vis,
                                   size_t npix, const int* xs, const int* ys, double* out,
                                   int nthreads, int* used_recurrence) {
  dft_invert_pixels_fast_impl(nx, ny, dx, dy, l_shift, m_shift, n_rows, n_chan, uvw, freq, vis,
                              npix, xs, ys, out, nthreads, used_recurrence, 0);
}

/* As oracle_dft_invert_pixels_fast, with one sum per channel:
 * out[p * n_chan + c] = the dirty image of channel c alone at pixel p. */
void oracle_dft_invert_pixels_per_channel(int nx, int ny, double dx, double dy, double l_shift,
                                          double m_shift, size_t n_rows, size_t n_chan,
                                          const double* uvw, const double* freq,
                                          const float* vis, size_t npix, const int* xs,
                                          const int* ys, double* out, int nthreads,
                                          int* used_recurrence) {
  dft_invert_pixels_fast_impl(nx, ny, dx, dy, l_shift, m_shift, n_rows, n_chan, uvw, freq, vis,
                              npix, xs, ys, out, nthreads, used_recurrence, 1);
}
