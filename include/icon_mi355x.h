/* icon_mi355x.h -- C ABI of the MI355X ICON reconstruction core.
 *
 * Drop-in boundary for ICON-GPU's per-slice compressed-sensing NUFFT solve.
 * Plain C types only: device buffers are `void*` obtained from icon_malloc()
 * (or any hipMalloc'd pointer on the plan's device), streams are `void*`
 * (a hipStream_t, NULL = the default stream).  Every call returns an
 * icon_status; icon_last_error() describes the last failure on the calling
 * thread.  Nothing here calls exit().
 *
 * Reference interfaces replaced (paths relative to
 * tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/):
 *   Lib_2D2                 MainCode.cuh:10, MainCode.cu:3-177
 *   Iteration3 / Alpha2     Iteration2.cuh:3, alpha_new2.cu:3-119
 *   cunfft_*Plan_2d2/_1d    nufft_gpu_v8.cuh:153-223, nufft_gpu_v8.cu:281-556,1298-1408
 *   cunfft_trafo_2d_R2C2    nufft_gpu_v8.cu:984-1076
 *   cunfft_trafo_2d2        nufft_gpu_v8.cu:844-949
 *   cunfft_adjoint_2d2      nufft_gpu_v8.cu:1167-1238
 *   cunfft_trafo_1d         nufft_gpu_v8.cu:1479-1516 (readSlicePart.cu:10-22)
 *   cunfft_adjoint_1d       nufft_gpu_v8.cu:1518-1536 (reProject_NFFT.cu:81-98)
 *   initGICONPara2          nufft_gpu_v8.cu:1739-1771
 */
#ifndef ICON_MI355X_H
#define ICON_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  ICON_OK = 0,
  ICON_ERR_ARG = 1,      /* invalid argument (sizes, pointers, flags) */
  ICON_ERR_DEVICE = 2,   /* HIP runtime failure */
  ICON_ERR_IO = 3,       /* file could not be opened / read / written */
  ICON_ERR_CAPACITY = 4, /* batch larger than the plan was created for */
  ICON_ERR_STATE = 5     /* call out of order (e.g. plan not prepared) */
} icon_status;

/* NUFFT window of a plan.  ICON_WINDOW_GAUSS is the reference's Gaussian,
 * m = 6, 14 taps per dimension (nufft_gpu_v8.cu:221-235); icon_plan_create
 * and icon_plan_create_cv always use it.  ICON_WINDOW_KB6 is a Kaiser-Bessel
 * window of width 6 (6 taps per dimension, beta from the oversampling
 * factor): about 5e-6 relative error per transform against the exact
 * non-uniform DFT, for about a third of the gridding work. */
typedef enum {
  ICON_WINDOW_GAUSS = 0,
  ICON_WINDOW_KB6 = 1
} icon_window;

typedef struct icon_plan icon_plan;

/* Geometry of a plan, as derived by the reference's setup
 * (MainCode.cu:63-71, nufft_gpu_v8.cu:471-476). */
typedef struct {
  int32_t N0, N1;        /* width (nx) and thickness */
  int32_t nz;            /* number of tilts */
  int32_t M;             /* nodes = N0 * nz */
  int32_t n0, n1;        /* oversampled grid, n = 2 * nextpow2(N) */
  int32_t m;             /* window half-width: 6 for the Gaussian (2m+2 = 14 taps),
                            3 for Kaiser-Bessel (2m = 6 taps) */
  int32_t iomit;         /* first tilt of minimal |theta| (MainCode.cu:46-49) */
  float sigma0, sigma1;  /* oversampling factors, float as the reference */
  float b0, b1;          /* window shape: Gaussian b, or Kaiser-Bessel beta */
  int32_t max_slices;    /* batch capacity */
  int32_t col_block;     /* grid columns per spread/interp tile */
  int32_t window;        /* icon_window */
  int64_t device_bytes;  /* device memory held by the plan */
} icon_plan_info;

/* Solver parameters: the reference's before[] / betain[] (ICONGPU.cu:380-387). */
typedef struct {
  int32_t pre;        /* INFR iterations before ICON            (-iter a,_,_) */
  int32_t icon;       /* ICON (thresholded) iterations          (-iter _,b,_) */
  int32_t post;       /* INFR iterations after ICON             (-iter _,_,c) */
  int32_t data_type;  /* 1: clamp m <= thr, 2: clamp m >= thr   (-d)        */
  float threshold;    /*                                        (-thr)      */
  int32_t use_graph;  /* 1: replay the iteration loop as a HIP graph       */
  int32_t reuse_data; /* 1: keep b and A^H W b from the previous call on this
                         plan, which must have had the same d_sino and S (the
                         caller keeps the sinograms unchanged); a warm-started
                         continuation then skips the data set-up */
} icon_solve_opts;

/* ---- errors / devices / memory ---------------------------------------- */
const char* icon_last_error(void);
const char* icon_version(void);
int icon_device_count(int* count);
/* The calling thread's current device (thin HIP wrappers). */
int icon_set_device(int device);
int icon_get_device(int* device);
/* Streams and events on the current device (thin HIP wrappers; streams at
 * the default priority): for callers that overlap independent solves. */
int icon_stream_create(void** stream);
int icon_stream_destroy(void* stream);
int icon_stream_synchronize(void* stream);
int icon_event_create(void** event);
int icon_event_destroy(void* event);
int icon_event_record(void* event, void* stream);
int icon_stream_wait_event(void* stream, void* event);
/* Device memory helpers on the current device (thin HIP wrappers). */
int icon_malloc(void** dptr, size_t bytes);
int icon_free(void* dptr);
int icon_memcpy_h2d(void* dst, const void* src, size_t bytes);
int icon_memcpy_d2h(void* dst, const void* src, size_t bytes);
int icon_memset_d(void* dst, int value, size_t bytes);
/* Device-to-device copy, asynchronous on `stream` (NULL = default stream). */
int icon_memcpy_d2d_async(void* dst, const void* src, size_t bytes, void* stream);
int icon_device_synchronize(void);

/* ---- plans ------------------------------------------------------------- */
/* Create a plan on `device` for a tilt series of width nx, reconstruction
 * thickness (0 = nx), nz tilts at the angles of the tilt file (degrees, as
 * read: the plan negates them as MainCode.cu:42 does) and up to max_slices
 * slices per batch.  Replaces cunfft_mallocPlan/initPlan_2d2 + _1d +
 * initGICONPara2. */
int icon_plan_create(icon_plan** plan, int device, int nx, int thickness, int nz,
                     const float* tilt_deg, int max_slices);
/* Cross-validation plan (Lib_2D_crossV2, MainCode_crossV.cu:3-189): the
 * first tilt of least |angle| is left out of the data (its row of the
 * sinograms is skipped; the other nz-1 tilts get their own density weights)
 * and the reconstruction is reprojected at that angle.  Takes the same
 * [S][nz][N0] sinograms as the full plan; icon_reconstruct's d_reproj then
 * holds the cross-validation reprojections. */
int icon_plan_create_cv(icon_plan** plan, int device, int nx, int thickness, int nz,
                        const float* tilt_deg, int max_slices);
/* icon_plan_create / icon_plan_create_cv (cross_validation != 0) with the
 * NUFFT window chosen (icon_window).  No counterpart in the reference, which
 * has only the Gaussian. */
int icon_plan_create_window(icon_plan** plan, int device, int nx, int thickness, int nz,
                            const float* tilt_deg, int max_slices, int cross_validation, int window);
int icon_plan_destroy(icon_plan* plan);
int icon_plan_get_info(const icon_plan* plan, icon_plan_info* info);
/* Tune the spread/interp column tile (8 or 16 grid columns); rebuilds the
 * bin lists.  Optional. */
int icon_plan_set_col_block(icon_plan* plan, int col_block);
/* Autotuner of the column tile (the spread/interpolation tile and bin width):
 * times one iteration's column kernels (2 interpolations + 1 spread, each
 * after its predecessor) over S slices for every tile width the plan's grid
 * and window support, keeps the fastest and reports it in *col_block and, if
 * ms_per_width is not NULL, the measured ms in ms_per_width[0] (C = 8) and
 * [1] (C = 16, or 0 when unsupported).  Overwrites the solve buffers, like
 * icon_time_kernel.  No counterpart in the reference. */
int icon_plan_autotune(icon_plan* plan, int S, int* col_block, double* ms_per_width);

/* ---- operators on S slices (device pointers, batched) ------------------ */
/* Forward NUFFT of S real images [S][N0][N1] to S node vectors [S][M]
 * complex (interleaved float2).  cunfft_trafo_2d_R2C2. */
int icon_nufft_forward_real(icon_plan* plan, int S, const void* d_img, void* d_f, void* stream);
/* Forward NUFFT of complex images [S][N0][N1] (float2).  cunfft_trafo_2d2. */
int icon_nufft_forward_complex(icon_plan* plan, int S, const void* d_img, void* d_f, void* stream);
/* Adjoint NUFFT of node vectors [S][M] (float2) to images [S][N0][N1]
 * (float2).  cunfft_adjoint_2d2. */
int icon_nufft_adjoint(icon_plan* plan, int S, const void* d_f, void* d_img, void* stream);
/* Fourier-slice data of S sinograms [S][nz][N0] (float): the 1D NUFFT of
 * every projection row, [S][M] float2.  readSlicePart + cunfft_trafo_1d. */
int icon_sino_to_b(icon_plan* plan, int S, const void* d_sino, void* d_b, void* stream);
/* Batched 1D FFT of `rows` contiguous rows of length n (256..2048, power of
 * two), in place, sign -1 (forward) or +1 (inverse), unnormalised. */
int icon_fft_rows(int n, int rows, int sign, void* d_data, void* stream);
/* Batched 2D FFT of `count` n0 x n1 grids (row-major, index x*n1 + y). */
int icon_fft_2d(int n0, int n1, int count, int sign, void* d_data, void* stream);

/* ---- the solve ---------------------------------------------------------- */
/* ICON/INFR solve of S slices (MainCode.cu:113-162, Iteration2.cu:3-150):
 *   d_sino    [S][nz][N0] float  projection rows of each slice (input)
 *   d_m       [S][N0][N1] float  reconstruction, index X*N1+Y (output)
 *   d_m_init  [S][N0][N1] float  warm start, or NULL for zeros (reference)
 *   d_reproj  [S][N0]     float  reprojection at the min-|theta| tilt, or NULL
 *   d_resid   [S][iters+1] double data residual R(m_i), i = 0..iters, or NULL:
 *             R(m) = sqrt(sum_j w_j |(A m)_j - b_j|^2 / sum_j w_j |b_j|^2)
 *   d_alpha   [S][iters]  double step sizes, or NULL
 * All pointers are device pointers on the plan's device; the call is
 * asynchronous on `stream`. */
int icon_reconstruct(icon_plan* plan, int S, const void* d_sino, void* d_m, const void* d_m_init,
                     void* d_reproj, void* d_resid, void* d_alpha, const icon_solve_opts* opts,
                     void* stream);

/* The same solve entirely in double precision on the device (the reference's
 * double build, DATATYPE = double, nufft_gpu_v8.cuh:25-33): FP64 window taps
 * and deapodisation (sigma = n/N), FP64 FFTs, spread and interpolation, FP64
 * reductions and update; the Gaussian window.  A precision reference for the
 * FP32 solve, not a fast path.  Host buffers, synchronous, device memory
 * allocated per call: sino [S][nz][nx] float in; m [S][nx][thickness],
 * reproj [S][nx] (or NULL), resid [S][iters+1] (or NULL) and alpha
 * [S][iters] (or NULL) double out. */
int icon_reconstruct_fp64(int device, int nx, int thickness, int nz, const float* tilt_deg, int S, const float* sino,
                          double* m, double* reproj, double* resid, double* alpha, const icon_solve_opts* opts);

/* Host-buffer convenience wrapper of icon_reconstruct (synchronous). */
int icon_reconstruct_host(icon_plan* plan, int S, const float* sino, float* m, const float* m_init,
                          float* reproj, double* resid, double* alpha, const icon_solve_opts* opts);

/* ---- packed solve of several series ------------------------------------ */
/* A batch packs tilt series of uneven width, thickness, tilt range and slice
 * count (one plan each, all on one device) into the same launches: every
 * kernel of the iteration runs once per size class (row length, column grid
 * and tile width) over all series of that class, and the whole solve is
 * replayed as one HIP graph.  Replaces the reference's per-series set-up and
 * slice loop (MainCode.cu:63-71,113-162) for a whole set of series.
 * slices[i] (<= the plan's max_slices) slices of series i are solved by
 * every icon_batch_reconstruct.  The plans must outlive the batch; a batch
 * overwrites their solve buffers. */
typedef struct icon_batch icon_batch;
int icon_batch_create(icon_batch** batch, int nseries, icon_plan* const* plans, const int* slices);
int icon_batch_destroy(icon_batch* batch);
/* icon_reconstruct for every series at once.  Each argument is an array of
 * nseries device pointers laid out as for icon_reconstruct on that series'
 * plan; d_m_init, d_reproj, d_resid and d_alpha may be NULL, or hold NULL
 * entries.  The options (iteration split, data type, threshold) apply to all
 * series; reuse_data is not supported.  use_graph is ignored (always a graph).
 * Asynchronous on `stream`. */
int icon_batch_reconstruct(icon_batch* batch, const void* const* d_sino, void* const* d_m,
                           const void* const* d_m_init, void* const* d_reproj, void* const* d_resid,
                           void* const* d_alpha, const icon_solve_opts* opts, void* stream);
/* Kernel launches of one iteration of the packed solve. */
int icon_batch_launches_per_iteration(const icon_batch* batch, int* launches);
/* Autotuner of the column tile (the spread/interpolation tile and bin width)
 * per column size class: times the class's packed interpolation and spread
 * launches for every tile width it supports and keeps the fastest for all
 * its plans; col_blocks (nseries entries, or NULL) receives the choice per
 * series.  Overwrites the plans' solve buffers.  No counterpart in the
 * reference. */
int icon_batch_autotune(icon_batch* batch, int* col_blocks);

/* ---- host-only helpers (no device needed) -------------------------------- */
/* The plan geometry computed on the host exactly as icon_plan_create does
 * (MainCode.cu:38-78, weight.cu:8-41, px2D.cu:8-47, nufft_gpu_v8.cu:22-42,
 * 221-235, 1242-1271).  Any output pointer may be NULL.  Sizes: nodes 2*M,
 * weights M, phix/phiy M*14, ix0/iy0 M (first wrapped tap index), ck N0*N1,
 * ck1 N0, phi1 N0*14, i10 N0; info as icon_plan_info (device fields 0). */
int icon_geometry_tables(int nx, int thickness, int nz, const float* tilt_deg, icon_plan_info* info,
                         float* nodes, float* weights, float* phix, float* phiy, int* ix0, int* iy0,
                         float* ck, float* ck1, float* phi1, int* i10);

/* MRC I/O of the command line (mrcfile3.c in the reference).
 * Reads slices [y0, y0+count) of a tilt series into out[count][nz][nx]
 * (mrc_read_slice 'y', mrcfile3.c:902-951) and reports nx, ny, nz, mode. */
int icon_mrc_read_y_slices(const char* path, int y0, int count, float* out, int* dims4);
/* Writes one reconstructed slice as saveMRC_real_inv + mrc_update_head do
 * (saveMRC_real.cu:33-56, mrcfile3.c:113-514): header copied from the tilt
 * series `template_path`, image m[width][thickness] transposed to x fastest. */
int icon_mrc_write_slice(const char* path, const char* template_path, const float* m, int width,
                         int thickness);

/* ---- measurement -------------------------------------------------------- */
/* Kernels of one ICON iteration, for icon_time_kernel. */
typedef enum {
  ICON_K_ROW_FWD = 0,   /* deapodise + pad + y-FFT of the image (fused update) */
  ICON_K_COL_FWD = 1,   /* x-FFT of the grid tile + interpolation to the nodes */
  ICON_K_COMBINE = 2,   /* node sums, weights, residual (FP64)                 */
  ICON_K_COL_ADJ = 3,   /* spread of the nodes into the grid tile + x-IFFT     */
  ICON_K_ROW_ADJ = 4,   /* y-IFFT + crop + deapodise + gradient + |r|^2 (FP64)
                           + deapodise, pad and y-FFT of the gradient (the row
                           pass of A rhat, fused)                              */
  ICON_K_SINO = 5,      /* 1D transform of the projection rows                 */
  ICON_K_COMBINE_ALPHA = 6  /* node sums of A rhat, sum w|A rhat|^2 (FP64), alpha */
} icon_kernel_id;

/* Times `reps` launches of one kernel of the iteration over S slices, each
 * right after the kernel that precedes it in an iteration (untimed), all
 * enqueued back to back on the plan's stream after 3 untimed pairs, on the
 * buffers of the plan's last solve.  Each timed launch carries HIP events
 * that take the kernel's own start and end (hipExtLaunchKernel);
 * *avg_ms = median launch duration.  Not part of the reference; measurement
 * only. */
int icon_time_kernel(icon_plan* plan, int kernel, int S, int reps, double* avg_ms);

/* Device-to-device copy of `bytes` repeated `reps` times with a plain vector
 * copy kernel; *avg_ms per copy.  Calibrates HBM bandwidth and counters. */
int icon_copy_probe(void* d_dst, const void* d_src, size_t bytes, int reps, double* avg_ms);

/* ---- file-level entry point -------------------------------------------- */
/* Lib_2D2 (MainCode.cu:3-177): reconstruct slices [slice_begin, slice_end)
 * of the MRC tilt series `mrc_path` with the tilt file `angle_path`, writing
 * <result_dir>/midYYYYY.mrc per slice (f32, nx x thickness, transposed as
 * saveMRC_real_inv) and the reprojection rows into `reproj`
 * ((slice_end - slice_begin) * nx floats, may be NULL).  `before` =
 * {pre, post, threadnum, dataType, skipCV}, `betain` = {threshold,
 * thickness}; iternum = pre + icon + post.  Slices are solved in batches of
 * up to `batch` (<= 0: automatic) on `device`. */
int icon_lib_2d2(int device, int slice_begin, int slice_end, int iternum, const char* mrc_path,
                 float* reproj, const char* angle_path, const char* result_dir, const int* before,
                 const float* betain, int batch);
/* Lib_2D_crossV2 (MainCode_crossV.cu:3-189): the same slices solved without
 * the least-|angle| tilt (icon_plan_create_cv), reprojected at that angle
 * into `reproj`; no slice files are written. */
int icon_lib_2d_crossv(int device, int slice_begin, int slice_end, int iternum, const char* mrc_path,
                       float* reproj, const char* angle_path, const char* result_dir, const int* before,
                       const float* betain, int batch);
/* calFRC_nonsqure (calFRC.cu:93-174): Fourier ring correlation of two
 * nx (fast) x ny images, rings floor(sqrt(di^2 + (dj nx/ny)^2)) < nx/2;
 * frc[0..nx/2-1], frc[0] = 1, 0 for an empty ring.  Host computation. */
int icon_frc(const float* image1, const float* image2, int nx, int ny, float* frc);
/* frcSmooth (ICONGPU.cu:59-91): in-place running mean over `overlap`
 * entries, the window shifted inwards at the ends. */
void icon_frc_smooth(float* frc, int size, int overlap);
/* ICONPreProcess (ICONPreProcess.c:31-86,256-302) on the device: for each
 * image k of the stack d_stack [nz][ny][nx] (in place), subtract the mode of
 * a 256-bin histogram of its central region (10 % margins left out; the bin
 * of the maximum itself is not counted, as in the reference), and, when
 * tilt_deg != NULL and thickness > 0, scale the image to variance
 * 0.33 * thickness / cos(-tilt_deg[k] degrees).  modes (host, may be NULL)
 * receives the subtracted modes.  Synchronous. */
int icon_preprocess(int device, float* d_stack, int nx, int ny, int nz, const float* tilt_deg, int thickness,
                    double* modes, void* stream);
/* icon_preprocess without any device allocation, for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments, results and
 * synchronisation as icon_preprocess; the tilt angles and modes are staged in
 * a fixed device array of the library (48 KiB per device, held until the
 * process ends) instead of buffers allocated per call, so nz is limited to
 * 4096 images per call (ICON_ERR_CAPACITY beyond).  Calls that overlap on one
 * device share that array and race; calls on different devices do not. */
int icon_preprocess_noalloc(int device, float* d_stack, int nx, int ny, int nz, const float* tilt_deg, int thickness,
                            double* modes, void* stream);
/* The same on MRC files: in_path -> out_path (float, header copied from the
 * input, statistics recomputed); tilt_path NULL or thickness <= 0: mode
 * subtraction only. */
int icon_preprocess_file(int device, const char* in_path, const char* out_path, const char* tilt_path,
                         int thickness);
const char* icon_preprocess_error(void);
/* ICONMask3 (ICONCrossVMaskIn2D_weightOnlyOnSampled.cu:369-598,602-1236).
 * icon_mask3_weights: the half-spectrum multiplier [N][N/2+1] (row ky,
 *   column kx <= N/2) = realMask * weighting of genMask: the angular sampling
 *   mask of the tilts (degrees as in the tilt file), 1 inside the FRC-0.5
 *   radius radius05_rate * N/2, a Gaussian edge out to the FRC-0.3 radius
 *   (or gauss_width >= 0 pixels; radius_given selects the '-radius' rules),
 *   times |k|/r_full on sampled frequencies beyond r_full.  Host only.
 * icon_mask3_filter: in place on `count` real N x N slices on the device
 *   (N 8..1024, or a power of two up to 2048): the reference's R2C ->
 *   multiply -> C2R / N^2.  Synchronous.
 * icon_mask3: the program on <input_dir>/midYYYYY.mrc, slices slice_begin ..
 *   slice_end (inclusive), cropped to `thickness` rows around the centre
 *   (+ zshift), written to out_path.  radius > 0 gives the mask radius;
 *   otherwise it comes from the crossV / fullRec FRC files (falling back to
 *   the tilt range's radius when they are not valid); no_mask != 0 only
 *   crops and stacks (simpleCatch). */
int icon_mask3_weights(int N, const float* tilt_deg, int ang_num, float radius05_rate, float radius03_rate,
                       int radius_given, int gauss_width, float mean_ang_step, float* half_mask);
int icon_mask3_filter(int device, float* d_slices, int N, int count, const float* half_mask, void* stream);
int icon_mask3(int device, const char* input_dir, int slice_begin, int slice_end, int thickness,
               const char* tilt_path, const char* out_path, int radius, const char* crossv_frc,
               const char* fullrec_frc, int zshift, int gauss_width, int no_mask);
const char* icon_mask3_error(void);
/* Last failure of icon_lib_2d2 / icon_lib_2d_crossv on any thread. */
const char* icon_lib_2d2_error(void);

/* Tilt file reader (MainCode.cu:38-45): whitespace-separated degrees, as
 * written (not negated).  Stores at most `cap`, reports the count found;
 * ICON_ERR_IO if the file cannot be opened, ICON_ERR_CAPACITY if it holds
 * more than `cap` values. */
int icon_read_tilts(const char* path, float* tilt_deg, int cap, int* count);

/* Slice partition across devices (ICONGPU.cu:393-431): shares proportional
 * to the compute-unit counts `cus`, leftovers round-robin from the device of
 * most units, the last device taking the remainder.  Fills first[d]/last[d]
 * (inclusive) and returns the number of devices used. */
int icon_partition_slices(int ndev, const int* cus, int y_start, int y_end, int* first, int* last);

/* Compute units of a device (the partition weight). */
int icon_device_compute_units(int device, int* cus);

#ifdef __cplusplus
}
#endif
#endif /* ICON_MI355X_H */
