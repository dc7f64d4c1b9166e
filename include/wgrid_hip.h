/*
 * wgrid_hip.h -- C-ABI of the MI355X w-gridder / degridder.
 *
 * Each entry point replaces one member of WSClean's w-gridder boundary,
 * class WGriddingGridderBase (wgridder/gridder_simple.h:20-42), whose
 * concrete implementation WGriddingGridder_Simple<NumT> calls ducc0's
 * ms2dirty / dirty2ms (wgridder/gridder_simple.h:209-227,
 * wgridder/gridder_simple_implementation.h:170-196).  Plain pointers and
 * sizes only; std::complex<float> buffers are passed as float[2 * n].
 *
 * Every function returns 0 on success and a non-zero status on failure;
 * wg_last_error() then returns a thread-local message.  A failed call
 * leaves the handle usable.  Calls on one handle must not overlap; distinct
 * handles may be used from different threads concurrently (each handle owns
 * its own HIP stream on the current device).
 */
#ifndef WGRID_HIP_H_
#define WGRID_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wg_handle wg_handle;

/* Constructor arguments of WGriddingGridder_Simple
 * (wgridder/gridder_simple.h:84-88), in the same order. */
typedef struct wg_params {
  size_t width;          /* untrimmed image width (pixels) */
  size_t height;         /* untrimmed image height */
  size_t width_t;        /* trimmed width gridded (<= width, even) */
  size_t height_t;       /* trimmed height (<= height, even) */
  double pixel_size_x;   /* radians */
  double pixel_size_y;   /* radians */
  double l_shift;        /* phase-centre shift of the image centre */
  double m_shift;
  size_t nthreads;       /* accepted for interface parity; unused on GPU */
  double epsilon;        /* requested accuracy (RMS vs. direct FT) */
  size_t verbosity;      /* 0, 1 or 2; larger values are an error */
  int tuning;            /* non-zero: tuned variant (ms2dirty_tuning) */
} wg_params;

/* new WGriddingGridder_Simple(...)   (gridder_simple_implementation.h:33-50)
 * Fails with a status (the reference throws) if verbosity > 2. */
int wg_create(const wg_params* params, wg_handle** out);
/* ~WGriddingGridder_Simple */
int wg_destroy(wg_handle* h);

/* WGriddingGridderBase::ConstantMemoryUsage (gridder_simple.h:23,
 * gridder_simple_implementation.h:52-60): the device bytes that do not grow
 * with the visibility count (images, the FFT-slab budget, tables). */
int wg_constant_memory_usage(const wg_handle* h, size_t* bytes);
/* WGriddingGridderBase::PerVisibilityMemoryUsage (gridder_simple.h:24,
 * gridder_simple_implementation.h:62-68): device bytes per visibility of
 * one call.  Calls larger than 2^31 - 1 visibilities or than the free
 * device memory holds are split into row chunks by the library, so a
 * caller's chunk of any size works. */
int wg_per_visibility_memory_usage(const wg_handle* h, size_t* bytes);

/* WGriddingGridderBase::InitializeInversion (gridder_simple.h:25,
 * gridder_simple_implementation.h:70-73) */
int wg_initialize_inversion(wg_handle* h);

/* WGriddingGridderBase::AddInversionData (gridder_simple.h:26-28,
 * gridder_simple_implementation.h:75-89).  uvw: n_rows*3 doubles (metres),
 * freq: n_chan doubles (Hz, any order), vis: n_rows*n_chan complex<float>
 * (row-major, weighted).  Zero visibilities are skipped; calls accumulate. */
int wg_add_inversion_data(wg_handle* h, size_t n_rows, size_t n_chan,
                          const double* uvw, const double* freq,
                          const float* vis);

/* WGriddingGridderBase::FinalizeImage (gridder_simple.h:36,
 * gridder_simple_implementation.h:141-144) */
int wg_finalize_image(wg_handle* h, double multiplication_factor);

/* WGriddingGridderBase::RealImage (gridder_simple.h:37,
 * gridder_simple_implementation.h:146-156).  out: width*height floats,
 * index x + y*width; trimmed image centred, NaN border. */
int wg_real_image(wg_handle* h, float* out);

/* WGriddingGridderBase::InitializePrediction (gridder_simple.h:38,
 * gridder_simple_implementation.h:158-167).  image: width*height floats,
 * index x + y*width; the centred trimmed region is used. */
int wg_initialize_prediction(wg_handle* h, const float* image);

/* WGriddingGridderBase::PredictVisibilities (gridder_simple.h:39-41,
 * gridder_simple_implementation.h:169-195).  vis_out: n_rows*n_chan
 * complex<float>, overwritten. */
int wg_predict_visibilities(wg_handle* h, size_t n_rows, size_t n_chan,
                            const double* uvw, const double* freq,
                            float* vis_out);

/* Thread-local message of the last failed call on this thread. */
const char* wg_last_error(void);

/* Polarization codes (the gridder's polarization, WSClean's
 * aocommon::PolarizationEnum; values of the FITS STOKES axis). */
enum {
  WG_POL_I = 1, WG_POL_Q = 2, WG_POL_U = 3, WG_POL_V = 4,
  WG_POL_RR = -1, WG_POL_LL = -2, WG_POL_RL = -3, WG_POL_LR = -4,
  WG_POL_XX = -5, WG_POL_YY = -6, WG_POL_XY = -7, WG_POL_YX = -8
};

/* WGriddingGridderBase::AddInversionDataWithCorrectionCallback
 * (wgridder/gridder_simple.h:29-35; VisibilityCallbackBuffer,
 * wgridder/callback_implementation.h:34-136).  The reference reaches the
 * facet solutions through an MsGridder callback; across the C-ABI they are
 * passed explicitly: `gains` is the cached solution table
 * [n_times][n_chan][n_antennas][nparms] complex<float> (nparms 2: diagonal
 * X, Y; 4: full Jones), indexed per row by time_offsets[row] and the antenna
 * pair antennas[2 row], antennas[2 row + 1] (gridding/visibilitymodifier.h:
 * 539-594).  vis is [n_rows][n_chan][n_pol] complex<float>.  gain_mode in
 * GainMode order: 0 xx, 1 yy, 2 trace (n_pol 1), 3 2-vis diagonal (n_pol 2),
 * 4 full (n_pol 4).  Each visibility gets the forward gain (apply_forward,
 * direction-dependent PSF mode) and then the conjugate gain, the
 * polarizations are collapsed as CollapseData (gridding/msgridderdata.h:
 * 37-52; wgridder/callback_implementation.h:92-95): 2 are summed, 4 linear
 * correlations (XX, XY, YX, YY) are converted to `polarization` (a WG_POL_
 * code: Stokes I = (XX + YY) / 2, Q, U, V, the linear products, or the
 * circular RR, LL, RL, LR), and the result is gridded as by
 * wg_add_inversion_data. */
int wg_add_inversion_data_corrected(wg_handle* h, int gain_mode, size_t n_pol, size_t n_rows,
                                    const double* uvw, const double* freq, size_t n_chan,
                                    const uint32_t* antennas, const float* vis,
                                    const uint64_t* time_offsets, const float* gains,
                                    size_t n_times, size_t n_antennas, size_t nparms,
                                    int apply_forward, int polarization);

/* ------------------------------------------------------------------------
 * Device-resident variants.  Same semantics as the calls above, but uvw,
 * vis and vis_out are HIP device pointers (freq stays on the host).  They
 * correspond to the same reference members (AddInversionData,
 * PredictVisibilities); the caller keeps the data in HBM across calls.
 * Device buffers must hold their data when a call is made: the library
 * works on a HIP stream of its own, which does not wait for work the
 * caller queued on other streams.  Every call has finished with its
 * buffers when it returns.
 */
int wg_add_inversion_data_device(wg_handle* h, size_t n_rows, size_t n_chan,
                                 const double* uvw_dev, const double* freq,
                                 const float* vis_dev);
int wg_predict_visibilities_device(wg_handle* h, size_t n_rows, size_t n_chan,
                                   const double* uvw_dev, const double* freq,
                                   float* vis_out_dev);
/* RealImage / InitializePrediction (gridder_simple.h:37-38) with the
 * width*height float image in device memory: a major cycle's dirty image
 * and model stay in HBM. */
int wg_real_image_device(wg_handle* h, float* image_dev);
int wg_initialize_prediction_device(wg_handle* h, const float* image_dev);

/* ------------------------------------------------------------------------
 * Major-cycle variants.  A predict followed by an invert on the same uvw
 * (WSClean's major cycle, main/wsclean.cpp:1519-1632, with
 * WGriddingMSGridder::PredictMeasurementSet / GridMeasurementSet,
 * wgridder/wgriddingmsgridder.cpp:117-241) shares one plan and one sort.
 * wg_prepare_visibilities builds the device-resident sorted set once;
 * invert/predict then reuse it.  Zero visibilities are gridded as zeros. */
int wg_prepare_visibilities(wg_handle* h, size_t n_rows, size_t n_chan,
                            const double* uvw_dev, const double* freq);
int wg_add_inversion_prepared(wg_handle* h, const float* vis_dev);
int wg_predict_prepared(wg_handle* h, float* vis_out_dev);

/* A major cycle on the prepared set, captured once as a HIP graph and
 * replayed without host launch work (WSClean's major cycle,
 * main/wsclean.cpp:1519-1632, repeats the same invert / predict on the same
 * data).  what & 1: clear the image, invert vis_dev, multiply by factor
 * (InitializeInversion + AddInversionData + FinalizeImage); what & 2:
 * predict vis_out_dev from the handle's image (PredictVisibilities).  The
 * buffers are bound at capture; a replay reads their current contents.
 * The capture runs the cycle once eagerly.  wg_prepare_visibilities drops
 * the graph.  No reference counterpart (launch mechanism). */
int wg_capture_major_cycle(wg_handle* h, const float* vis_dev, float* vis_out_dev,
                           double factor, int what);
int wg_replay_major_cycle(wg_handle* h, int count);

/* Autotuner (the `tuning` flag of wg_params selects ducc0's tuned
 * variant in the reference, gridder_simple.h:217-225 via
 * wgridder/wgriddingmsgridder.cpp:41-54; here it runs this autotuner).
 * Times plan + sort + invert + predict of every candidate -- per kernel
 * support W the smallest oversampling meeting epsilon and two larger ones,
 * 8-cell tiles, and 16-cell tiles for the best two -- on a strided row
 * sample of at most max_sample_vis visibilities of uvw_dev, extrapolates
 * to all rows (per-visibility stages scale with the count, plane FFTs do
 * not), and sets the handle's tuning to the fastest.  The choice is also
 * kept for the process: a handle created with tuning != 0 runs the
 * autotuner in its first invert / predict / prepare call for a geometry and
 * visibility count (rounded to a power of two) and reuses the choice
 * afterwards.  report (optional): JSON list of the trials,
 * fastest first. */
int wg_autotune(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw_dev,
                const double* freq, size_t max_sample_vis, char* report, size_t len);

/* Plan of the last call as JSON (W, sigma, nu, nv, P, box, tiles...). */
int wg_plan_json(const wg_handle* h, char* buf, size_t len);
/* Per-stage device times (ms, HIP events on the handle's stream) accumulated
 * since the last reset while timing is enabled.  out14 = plan, sort, grid,
 * fft, degrid, total; then for the gridding kernel alone: time (ms),
 * launches, algorithmic bytes; the same three for the degridding kernel;
 * then the visibility passes of the gridding and of the degridding kernel
 * (a visibility counts once per plane sub-slab it touches).  No reference
 * counterpart (measurement hook). */
int wg_set_timing(wg_handle* h, int enabled);
int wg_stage_times(wg_handle* h, double* out14, int reset);
/* Tuning overrides (0 = automatic): kernel support, oversampling, tile
 * edge, planes per gridding sub-slab, FFT-slab byte budget. */
int wg_set_tuning(wg_handle* h, int W, double sigma, int tile, int sub_slab,
                  size_t slab_bytes);
/* Smallest epsilon this library accepts: its grids and FFT are FP32 (with
 * FP64 coordinates and phases), and below this no kernel shape meets the
 * error estimate, so wg_add_inversion_data / wg_predict_visibilities with a
 * smaller epsilon fail with a message.  No counterpart in the reference,
 * which switches to FP64 grids at epsilon <= 1.01e-5
 * (wgridder/wgriddingmsgridder.cpp:43-48): a caller keeps the reference's
 * FP64 gridder below this value. */
int wg_min_epsilon(double* epsilon);
/* Arithmetic of the matrix-core gridding (W <= 8): -1 (default) split
 * precision (each FP32 operand as a pair of f16 halves on the f16 matrix
 * instructions) where its error estimate meets epsilon, FP32 otherwise;
 * 0 always FP32; 1 split whenever W <= 8.  No counterpart in the
 * reference (a tuning override like wg_set_tuning). */
int wg_set_precision(wg_handle* h, int split);

/* ------------------------------------------------------------------------
 * Facet batches.  WSClean images a faceted field with one gridder per facet
 * (gridding/msgriddermanager.cpp:41-77) and, in BatchInvert
 * (gridding/msgriddermanager.cpp:227-396 with
 * WGriddingMSGridder::GridSharedMeasurementSetChunk,
 * wgridder/wgriddingmsgridder.cpp:97-115), feeds one chunk read to all of
 * them.  A facet batch replaces that set of WGriddingGridder_Simple objects:
 * facet f has the constructor arguments facets[f] (its own image size,
 * pixel size and l/m shift; the smallest epsilon of the batch applies to
 * all), and its visibilities are the rows [row_offset[f], row_offset[f] +
 * n_rows[f]) of the shared device buffers uvw_dev ([rows][3] f64) and
 * vis_dev ([rows][n_chan] complex64); facets may share rows.  All facets
 * share one sort and one gridding / degridding launch per plane sub-slab.
 * Per-facet images follow RealImage / InitializePrediction
 * (gridder_simple.h:37-38): width*height floats, trimmed image centred,
 * NaN border.  wg_facet_batch_predict overwrites vis_out_dev in the facet
 * layout: facet f's [n_rows[f]][n_chan] block follows the blocks of the
 * facets before it (wg_facet_batch_virtual_count gives the total). */
typedef struct wg_facet_batch wg_facet_batch;
int wg_facet_batch_create(const wg_params* facets, size_t n_facets, wg_facet_batch** out);
int wg_facet_batch_destroy(wg_facet_batch* b);
int wg_facet_batch_prepare(wg_facet_batch* b, size_t n_chan, const double* freq,
                           const double* uvw_dev, const size_t* row_offset,
                           const size_t* n_rows);
/* InitializeInversion / AddInversionData / FinalizeImage for every facet */
int wg_facet_batch_initialize_inversion(wg_facet_batch* b);
int wg_facet_batch_add_inversion(wg_facet_batch* b, const float* vis_dev);
int wg_facet_batch_finalize_image(wg_facet_batch* b, size_t facet, double factor);
int wg_facet_batch_real_image(wg_facet_batch* b, size_t facet, float* out);
int wg_facet_batch_real_image_device(wg_facet_batch* b, size_t facet, float* image_dev);
int wg_facet_batch_initialize_prediction(wg_facet_batch* b, size_t facet, const float* image);
int wg_facet_batch_initialize_prediction_device(wg_facet_batch* b, size_t facet,
                                                const float* image_dev);
/* PredictVisibilities of every facet from its model */
int wg_facet_batch_predict(wg_facet_batch* b, float* vis_out_dev);
int wg_facet_batch_virtual_count(wg_facet_batch* b, size_t* n);
/* The batch's major cycle as a HIP graph, as wg_capture_major_cycle;
 * factors: one FinalizeImage factor per facet. */
int wg_facet_batch_capture_major_cycle(wg_facet_batch* b, const float* vis_dev,
                                       float* vis_out_dev, const double* factors, int what);
int wg_facet_batch_replay_major_cycle(wg_facet_batch* b, int count);
/* Stage timing of the batch, as wg_set_timing / wg_stage_times. */
int wg_facet_batch_set_timing(wg_facet_batch* b, int enabled);
int wg_facet_batch_stage_times(wg_facet_batch* b, double* out14, int reset);
/* Atlas and per-facet plans of the last prepare (JSON). */
int wg_facet_batch_plan_json(wg_facet_batch* b, char* buf, size_t len);

/* ------------------------------------------------------------------------
 * Imaging weights: WSClean's ImageWeights (structures/imageweights.h:22-132,
 * structures/imageweights.cpp:33-390) on the device.  The half-plane grid
 * of weight sums, [height / 2][width] FP64, lives in HBM.
 * mode: 0 natural, 1 uniform, 2 Briggs (robustness briggs_robustness);
 * the other arguments as the constructor (imageweights.cpp:33-48). */
typedef struct wg_image_weights wg_image_weights;
int wg_image_weights_create(int mode, double briggs_robustness, size_t width, size_t height,
                            double pixel_scale_x, double pixel_scale_y, int weights_as_taper,
                            double super_weight, wg_image_weights** out);
int wg_image_weights_destroy(wg_image_weights* w);
/* Grid size after the super-weight division and rounding to even. */
int wg_image_weights_size(const wg_image_weights* w, size_t* width, size_t* height);
/* ImageWeights::Grid(MSProvider&, BandData) (imageweights.cpp:76-114) for
 * one chunk of rows: uvw_dev [rows][3] f64 (metres), weights_dev
 * [rows][chan][pol] f32 (device), freq [chan] (host). */
int wg_image_weights_grid(wg_image_weights* w, size_t n_rows, size_t n_chan, size_t n_pol,
                          const double* uvw_dev, const double* freq, const float* weights_dev);
/* FinishGridding (imageweights.cpp:116-150); a second call fails. */
int wg_image_weights_finish(wg_image_weights* w);
/* SetMinUVRange / SetMaxUVRange / SetTukeyTaper / SetTukeyInnerTaper /
 * SetEdgeTaper / SetEdgeTukeyTaper / SetGaussianTaper / RankFilter
 * (imageweights.cpp:152-356) */
int wg_image_weights_set_min_uv_range(wg_image_weights* w, double min_uv);
int wg_image_weights_set_max_uv_range(wg_image_weights* w, double max_uv);
int wg_image_weights_set_tukey_taper(wg_image_weights* w, double transition, double max_uv);
int wg_image_weights_set_tukey_inner_taper(wg_image_weights* w, double transition,
                                           double min_uv);
int wg_image_weights_set_edge_taper(wg_image_weights* w, double size);
int wg_image_weights_set_edge_tukey_taper(wg_image_weights* w, double transition, double edge);
int wg_image_weights_set_gaussian_taper(wg_image_weights* w, double beam_size);
int wg_image_weights_rank_filter(wg_image_weights* w, double rank_limit, size_t window);
/* GetGrid (imageweights.cpp:270-287): image [height][width] FP64 (host). */
int wg_image_weights_get_grid(wg_image_weights* w, double* image);
/* GetWeight(u, v) (imageweights.h:46) for n (u, v) pairs in wavelengths
 * (host arrays). */
int wg_image_weights_sample(wg_image_weights* w, size_t n, const double* u, const double* v,
                            double* out);

/* Caller-side row preparation on the device: MsGridderData::CalculateWeights
 * + ApplyWeights (gridding/msgridderdata.h:761-789, :815-886) with
 * RotateVisibilities (:929-941) and CollapseData (:37-52), without gain
 * corrections (those are wg_add_inversion_data_corrected).  Per (row,
 * channel, polarization): the point-source visibility exp(+2 pi i (u dl +
 * v dm + w dn) / lambda) when psf (else vis_dev), minus model_dev when
 * given; weight zero for unselected channels (selected: host bytes, NULL =
 * all); weights squared (weighting 1) or set to one (2); times the imaging
 * weight GetWeight(u / lambda, v / lambda) (image_weights NULL: 1); the
 * visibility times that weight; then one visibility per channel (2
 * polarizations: sum; 4 linear correlations XX, XY, YX, YY: converted to
 * `polarization`, a WG_POL_ code, as CollapseData does) into out_dev
 * [rows][chan] complex64, ready for wg_add_inversion_data_device.
 * stats4 (host): total weight (all polarizations), visibility weight sum,
 * max gridded weight, gridded visibility count (first polarization), as
 * MsGridderData accumulates them.  vis_dev, weights_dev, model_dev:
 * [rows][chan][pol] on the device; uvw_dev [rows][3] f64.  Runs on the
 * handle's stream. */
int wg_prepare_rows(wg_handle* h, size_t n_rows, size_t n_chan, size_t n_pol,
                    const double* uvw_dev, const double* freq, const float* vis_dev,
                    const float* weights_dev, const float* model_dev, const uint8_t* selected,
                    int psf, double psf_dl, double psf_dm, int weighting, int polarization,
                    wg_image_weights* image_weights, float* out_dev, double* stats4);

/* wg_prepare_rows without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as wg_prepare_rows; its scratch (statistics partials, channel
 * wavelengths and selection flags) is a fixed ~620 KB device array of the
 * library, one per GPU, held until the process ends, so n_chan is limited to
 * 65536 (larger counts fail with an error).  Calls on one GPU share that
 * array; the library serialises them per device.  The current device must
 * be the handle's. */
int wg_prepare_rows_noalloc(wg_handle* h, size_t n_rows, size_t n_chan, size_t n_pol,
                            const double* uvw_dev, const double* freq, const float* vis_dev,
                            const float* weights_dev, const float* model_dev,
                            const uint8_t* selected, int psf, double psf_dl, double psf_dm,
                            int weighting, int polarization, wg_image_weights* image_weights,
                            float* out_dev, double* stats4);

/* ------------------------------------------------------------------------
 * Reordered measurement-set parts: the temporary files WSClean's
 * ReorderedMsProvider reads (msproviders/reordering.h:22-60,
 * msproviders/reorderedmsprovider.cpp:58-97, msproviders/msreaders/
 * reorderedmsreader.cpp:6-150).  meta_path: the spectral window's
 * "<prefix>-spw<d>-parted-meta.tmp"; data_path: the part's "<part
 * prefix>.tmp"; weight_path ("<part prefix>-w.tmp") and model_path ("<part
 * prefix>-m.tmp") may be NULL (weights 1, no model); n_pol: visibilities
 * per row and channel in the files (1 for a single polarization). */
typedef struct wg_part_reader wg_part_reader;
int wg_part_open(const char* meta_path, const char* data_path, const char* weight_path,
                 const char* model_path, size_t n_pol, wg_part_reader** out);
int wg_part_close(wg_part_reader* p);
/* info6 = rows, channel count, first channel, data description id, has
 * model, polarizations per file; start time (s) and the measurement-set path
 * of the meta header (optional). */
int wg_part_info(const wg_part_reader* p, uint64_t* info6, double* start_time, char* ms_path,
                 size_t ms_path_len);
/* Rows [row0, row0 + n) into host buffers (each may be NULL): uvw [n][3]
 * (metres), antennas [n][2], times [n], field ids [n] (ReadMeta), vis and
 * model [n][chan][pol] complex<float> (ReadData, ReadModel), weights
 * [n][chan][pol] float (ReadWeights). */
int wg_part_read(const wg_part_reader* p, size_t row0, size_t n, double* uvw, uint16_t* antennas,
                 double* times, uint16_t* fields, float* vis, float* weights, float* model);
/* GridMeasurementSet over a whole part (wgridder/wgriddingmsgridder.cpp:
 * 117-183 on a ReorderedMsProvider): chunks of chunk_rows rows (0: automatic)
 * are read from the files into page-locked host buffers (the next chunk on a
 * worker thread while the device grids the current one), weighted and
 * collapsed on the device as wg_prepare_rows does (weighting, polarization,
 * image_weights as there) and added to the handle's image as
 * AddInversionData.  freq: the part's channel frequencies.  stats4: the
 * weight sums of wg_prepare_rows over the part. */
int wg_part_grid(wg_part_reader* p, wg_handle* h, const double* freq, size_t chunk_rows,
                 int weighting, int polarization, wg_image_weights* image_weights,
                 double* stats4);

/* ------------------------------------------------------------------------
 * Small-inversion resampling (wgridder/wgriddingmsgridder.cpp:290-308 after
 * inversion, :326-336 before prediction, schaapcommon::math::Resampler):
 * Fourier interpolation of a float image, on the handle's stream.  Any side
 * up to 16384 that is 2^a 3^b 5^c, any other side up to 8192.
 * wg_resample: device [in_h][in_w] -> device [out_h][out_w].
 * wg_real_image_resampled: RealImage resampled to out_w x out_h (host),
 * the FinishInversion step.  wg_initialize_prediction_resampled: a host
 * model of in_w x in_h resampled to the gridder's width x height, then
 * InitializePrediction, the StartPredict step. */
int wg_resample(wg_handle* h, const float* in_dev, size_t in_w, size_t in_h, float* out_dev,
                size_t out_w, size_t out_h);
int wg_real_image_resampled(wg_handle* h, float* out, size_t out_w, size_t out_h);
int wg_initialize_prediction_resampled(wg_handle* h, const float* image, size_t in_w,
                                       size_t in_h);

/* ------------------------------------------------------------------------
 * Served request stream (no counterpart in the reference, which runs one
 * imaging job per process; it is the request loop a service built on this
 * library runs).  A request is one imaging job of `cycles` major cycles
 * (main/wsclean.cpp:1519-1632) on device-resident data: per cycle the
 * residual visibilities are inverted (wg_add_inversion_prepared on a
 * prepared set), a model is taken from the dirty image (0.1 x the pixels
 * at or above half its peak, standing in for the deconvolution), predicted
 * (wg_predict_prepared) and subtracted from the data.
 *
 * wg_server_create starts `slots` worker threads on the current device.
 * wg_server_submit queues a request and returns at once; requests start in
 * submission order as soon as a worker is free and the request's device
 * memory (an estimate, less what the pooled handle it will run on holds),
 * plus what the running requests have still to allocate, plus
 * reserve_bytes, fits in the device memory free at that moment
 * (hipMemGetInfo); short of memory, idle pooled handles are closed first,
 * and a request that does not fit even alone runs when nothing else does.
 * Requests of one geometry (npix, pixel_size, epsilon) reuse idle gridder
 * handles, at most max_idle_handles of which are kept.  wg_server_wait
 * blocks until request `id` has ended and gives its submission, start and
 * end times in seconds since wg_server_create (times3, may be NULL); a
 * failed request returns non-zero with its message.  The request's buffers
 * must stay valid, and its input buffers unchanged, until then.
 * wg_server_stats: completed requests, peak concurrent requests, pool hits,
 * pool misses, handles closed, the least free device memory seen at an
 * admission (bytes), admission attempts that had to wait, idle handles.
 * wg_server_destroy waits for every submitted request, then closes the
 * idle handles. */
typedef struct wg_server wg_server;
typedef struct wg_request {
  size_t npix;           /* image width = height (even, >= 8) */
  double pixel_size;     /* radians */
  double epsilon;        /* requested accuracy */
  int cycles;            /* major cycles, >= 1 */
  size_t n_rows;
  size_t n_chan;
  const double* uvw_dev; /* [n_rows][3] metres, device */
  const double* freq;    /* [n_chan] Hz, host (copied at submission) */
  const float* vis_dev;  /* [n_rows][n_chan] complex64 data, device */
  float* vis_out_dev;    /* [n_rows][n_chan] complex64, device: the last
                            cycle's predicted visibilities */
  float* image_host;     /* npix * npix float, host: the last dirty image
                            (NULL: not copied) */
  float* residual_dev;   /* [n_rows][n_chan] complex64, device: the residual
                            the last image was made from (NULL: not copied) */
} wg_request;
int wg_server_create(int slots, size_t reserve_bytes, int max_idle_handles, wg_server** out);
int wg_server_submit(wg_server* s, const wg_request* request, uint64_t* id);
int wg_server_wait(wg_server* s, uint64_t id, double* times3);
int wg_server_stats(wg_server* s, uint64_t* out8);
int wg_server_destroy(wg_server* s);

/* ------------------------------------------------------------------------
 * Test entry points (no reference counterpart): expose internals so that
 * the bucket sort and the LDS FFT can be checked bit-for-bit / against a
 * library FFT. */
/* Bucket keys (unsorted, one per visibility) and the stable sorted
 * permutation of the valid ones.  vis may be NULL (no zero skipping). */
int wgx_sort_keys(wg_handle* h, size_t n_rows, size_t n_chan,
                  const double* uvw, const double* freq, const float* vis,
                  uint32_t* keys_out, uint32_t* perm_out, size_t* n_valid,
                  char* plan_json, size_t plan_len);
/* In-place batched complex FFT of nlines lines of length n (host data). */
int wgx_fft(int n, int nlines, int inverse, float* data);
/* The device radix sort on host data: keys_out / perm_out = the n keys and
 * their element indices in the stable order of the low key_bits of the
 * keys (the sort behind the bucket index). */
int wgx_radix_sort(size_t n, int key_bits, const uint32_t* keys_in, uint32_t* keys_out,
                   uint32_t* perm_out);
/* Host-only: the plan (JSON) the library makes for an image geometry and
 * flipped-coordinate extents {umin, umax, vmin, vmax, wmin, wmax} (in
 * wavelengths), with the tuning overrides of wg_set_tuning.  Needs no GPU;
 * used to check kernel index ranges on the CPU. */
int wgx_make_plan(const wg_params* params, const double* extents6, long long nvis,
                  int W, double sigma, int tile, int sub_slab, char* plan_json,
                  size_t plan_len);
/* Largest number of visibilities the drop-in calls (AddInversionData,
 * PredictVisibilities and their device / corrected variants) hand to one
 * engine pass; larger calls are split into row chunks.  0 (the default):
 * 2^31 - 1 or what 80 % of the free device memory holds. */
int wgx_set_max_chunk_vis(wg_handle* h, size_t max_vis);
/* Device bytes the handle holds (its buffer arena). */
int wgx_device_bytes(wg_handle* h, size_t* bytes);
/* Host-only: the gridding kernels' polynomial form of the ES kernel for
 * support W and shape beta: coef[16][12] (row k: coefficients of g^0 .. g^d
 * of tap k, g = 2 f - 1; rows k >= W are zero) and the degree d.  Needs no
 * GPU. */
int wgx_es_poly(int W, double beta, float* coef, int* degree);
/* Host-only: the facet-batch layout (JSON: atlas plan, per-facet tile
 * origin and plan) for facets with flipped-coordinate extents extents6
 * (6 per facet) and visibility counts nvis.  Needs no GPU. */
int wgx_facet_atlas(const wg_params* facets, size_t n_facets, const double* extents6,
                    const long long* nvis, char* json, size_t len);
/* Synthetic visibilities on the device (benchmark inputs):
 * vis(row, chan) = sum_s S_s exp(-2 pi^2 sigma_s^2 (u^2 + v^2))
 *                  * exp(+2 pi i (u l_s + v m_s + w (n_s - 1))) + noise,
 * src: n_src x {l, m, n - 1, S, sigma} (host, FP64), noise: Gaussian sigma
 * per real/imaginary part from a counter-based generator seeded by seed. */
int wgx_synth_visibilities(const double* uvw_dev, size_t n_rows, size_t n_chan,
                           const double* freq, int n_src, const double* src,
                           double noise_sigma, uint64_t seed, float* vis_dev);

#ifdef __cplusplus
}
#endif

#endif /* WGRID_HIP_H_ */
