/*
 * quantem_ssb.h -- C ABI of the MI355X single-sideband (SSB) ptychography
 * engine (libquantem_ssb_hip.so, built for gfx950).
 *
 * Every entry point returns an int status (0 = success, negative = error;
 * see QSSB_ERR_*); ssb_last_error() returns the message of the calling
 * thread's last failure.  Array arguments are device pointers unless the
 * name ends in _host.  `stream` is a hipStream_t (NULL = legacy default
 * stream); work is enqueued asynchronously unless a function says it
 * synchronises.  A plan is not thread-safe: use one plan per stream.
 *
 * The entry points replace the CUDA backend of bobleesj/quantem.gpu behind
 * its SSBProtocol (src/quantem/gpu/ssb/compute/protocol.py:61-125); each
 * declaration cites the reference code whose role it takes.  Paths are
 * relative to that repository; cuda/ = src/quantem/gpu/ssb/compute/cuda/.
 */
#ifndef QUANTEM_SSB_H
#define QUANTEM_SSB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSSB_OK 0
#define QSSB_ERR_ARG (-1)         /* bad shape, size or pointer -> ValueError            */
#define QSSB_ERR_OOM (-2)         /* device allocation failed   -> MemoryError           */
#define QSSB_ERR_UNSUPPORTED (-3) /* unsupported scan size etc. -> NotImplementedError   */
#define QSSB_ERR_LAUNCH (-4)      /* HIP launch / runtime error -> RuntimeError          */
#define QSSB_ERR_STATE (-5)       /* call order, empty BF disk  -> ValueError            */

typedef struct ssb_plan ssb_plan;

/*
 * Element types of raw detector frames for the *_typed entry points.  The
 * reference keeps the caller's dtype (cuda/backend.py:195-201) and sums
 * integer frames in uint64 (signed values sign-extended, modulo 2^64) and
 * float frames in float64 (cuda/backend.py:430-441); here float frames are
 * summed in double-float (pairs of float32, error-free transformations).
 * The Fourier stack converts each sample to float32 (astype(complex64),
 * cuda/backend.py:570-571).
 */
#define QSSB_DTYPE_U8 1
#define QSSB_DTYPE_U16 2
#define QSSB_DTYPE_U32 3
#define QSSB_DTYPE_I8 4
#define QSSB_DTYPE_I16 5
#define QSSB_DTYPE_I32 6
#define QSSB_DTYPE_F32 7

/* Bytes per element of a QSSB_DTYPE_* code; 0 for an unknown code. */
int ssb_dtype_size(int dtype);

/*
 * Probe disk centre and radius (pixels) from the plan's last mean DP
 * (ssb_mean_dp_u16): detect_bf_radius of qg/detector/compute/cuda/probe.py:12-123
 * (threshold_ratio 0.1 there).  The SSB backend derives
 * det_sampling = 2*semiangle/radius from it when none is given
 * (cuda/backend.py:271-284).  Synchronises.
 */
int ssb_detect_bf_radius(ssb_plan* plan, double threshold_ratio, int* center_row, int* center_col, int* radius,
                         void* stream);

/*
 * Maximum of `count` uint16 values (device array) into *max_out_host.
 * Synchronises.  Decides whether the BF companion narrows to uint8
 * (cuda/engine.py:602-604).
 */
int ssb_max_u16(const uint16_t* data, long long count, unsigned int* max_out_host, void* stream);

/*
 * Exact detector-major BF columns of the plan's selection:
 * out[b][s] = data[s][bf_rows[bf_offset+b]][bf_cols[bf_offset+b]] for
 * b < count and the N*N scan positions s, as uint16 (narrow_u8 = 0) or
 * uint8 (narrow_u8 = 1; only lossless when every count is <= 255).  The
 * gather of SSBEngine.write_exact_bf_source (cuda/engine.py:577-701); the
 * caller writes the chunks to the companion file.
 */
int ssb_bf_columns_u16(ssb_plan* plan, const uint16_t* data, int bf_offset, int count, int narrow_u8, void* out,
                       void* stream);

/*
 * ssb_bf_columns_u16 for uint8 frames (out is uint8); the reference's
 * companion export accepts uint8 or uint16 counts only (cuda/engine.py:598-602).
 */
int ssb_bf_columns_u8(ssb_plan* plan, const uint8_t* data, int bf_offset, int count, void* out, void* stream);

/*
 * Learned-initializer features of the resident Hermitian stack: for the 8
 * scan-frequency grid points (qrow_host[i], qcol_host[i]), the 32 x 32
 * detector window of G around the rounded BF centre as |G|/max|G| and the
 * phase of G conj(G_centre) / pi; out is [32][32][16] float32 (device).  No
 * reference counterpart (the network only seeds the aberration search).
 * Synchronises.
 */
int ssb_learn_features(ssb_plan* plan, const int* qrow_host, const int* qcol_host, float* out, void* stream);

/*
 * Supported square scan size for an H x W scan: 128, 256, 512 or 1024, the
 * next one at or above the longest side (cuda/backend.py:238-246); 0 for an
 * empty scan.
 */
int ssb_scan_target(int H, int W);

/*
 * Scan normalisation (cuda/backend.py:229-269): centre-crop axes longer than
 * T (= ssb_scan_target(H, W)) and pad shorter ones with the floor of the
 * cropped scan's mean frame, at pad_top = (T-h)/2, pad_left = (T-w)/2.
 * src is [H][W][det_rows][det_cols], dst [T][T][det_rows][det_cols] uint16.
 * Integer-exact.  Synchronises the stream (it allocates and frees a small
 * workspace).
 */
int ssb_normalize_scan_u16(const uint16_t* src, int H, int W, int det_rows, int det_cols, uint16_t* dst, int T,
                           void* stream);

/*
 * ssb_normalize_scan_u16 for frames of any QSSB_DTYPE_* element type: the
 * pad frame is the cropped scan's int64 sum / count truncated toward zero
 * (integers) or its float64 sum / count rounded to float32 (float32 frames,
 * summed here in double-float), as cuda/backend.py:253-266.
 */
int ssb_normalize_scan_typed(const void* src, int dtype, int H, int W, int det_rows, int det_cols, void* dst,
                             int T, void* stream);

/*
 * ssb_normalize_scan_u16 without any device allocation or synchronisation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its partial sums and pad frame live in a library-owned
 * device array of QSSB_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends); inputs whose workspace exceeds it (more than 65,536
 * detector pixels at the largest scans) are rejected with QSSB_ERR_ARG
 * before any launch.  Results are bit-identical to ssb_normalize_scan_u16.
 * All calls on one GPU share the array: calls that overlap in time on one
 * GPU (other streams or threads) race.
 */
#define QSSB_NOALLOC_SCRATCH_BYTES (4456448) /* 4 MiB + 256 KiB */
int ssb_normalize_scan_u16_noalloc(const uint16_t* src, int H, int W, int det_rows, int det_cols, uint16_t* dst,
                                   int T, void* stream);

/* Library version (major*10000 + minor*100 + patch). */
int ssb_version(void);

/* Message of this thread's last error ("" if none). */
const char* ssb_last_error(void);

/*
 * Create a plan for square scans of scan_n x scan_n (128, 256 or 512) and a
 * det_rows x det_cols detector; max_bf bounds the bright-field pixel count.
 * Detector and optics follow CudaSSBBackend.__init__ (cuda/backend.py:168-365)
 * and SSBEngine.__init__ (cuda/engine.py:414-495): wavelength in Angstrom,
 * semiangle and per-pixel angular sampling in radians, scan sampling in
 * Angstrom (scan frequencies are fftfreq(scan_n, sampling) as float32,
 * cuda/backend.py:44-82).
 */
int ssb_plan_create(int scan_n, int det_rows, int det_cols, int max_bf, double wavelength_A,
                    double semiangle_rad, double ang_row_rad, double ang_col_rad,
                    double scan_sampling_row_A, double scan_sampling_col_A, ssb_plan** plan);

/*
 * Override the detector reciprocal sampling (1/Angstrom per pixel) used for
 * the BF k coordinates; by default it is ang/lambda.  SSBEngine derives it as
 * 1/(sampling*gpts) independently of the angular sampling used by the
 * aperture (cuda/engine.py:791-793, 819-830).
 */
int ssb_set_detector_sampling(ssb_plan* plan, double recip_row, double recip_col);

/*
 * Replace the scan frequencies (float32 [scan_n] each, host) -- the
 * q_row[:, 0] / q_col[0, :] of SSBEngine (cuda/engine.py:815-816).
 */
int ssb_set_scan_frequencies(ssb_plan* plan, const float* qrow_host, const float* qcol_host);

/* Release all device memory of the plan (CudaSSBBackend.free, cuda/backend.py:395-418). */
int ssb_plan_destroy(ssb_plan* plan);

/*
 * Exact mean diffraction pattern of `frames` uint16 frames (uint64 sums,
 * float32 mean = f32(sum)/f32(frames)); cuda/backend.py:420-446 and
 * qg/detector/compute/cuda/probe.py:126-147.  sums_out / mean_out may be NULL;
 * the result is always kept in the plan for ssb_select_bf.
 */
int ssb_mean_dp_u16(ssb_plan* plan, const uint16_t* data, int64_t frames, uint64_t* sums_out,
                    float* mean_out, void* stream);

/*
 * ssb_mean_dp_u16 for frames of any QSSB_DTYPE_* element type: exact
 * uint64 sums for integer types (sums_out, may be NULL), double-float sums
 * for float32 (sums_out must be NULL), mean = f32(sum)/f32(frames) times
 * the detector gain when one is set (cuda/backend.py:430-443).
 */
int ssb_mean_dp_typed(ssb_plan* plan, const void* data, int dtype, int64_t frames, uint64_t* sums_out,
                      float* mean_out, void* stream);

/*
 * Per-pixel detector gain (float32 [det_rows][det_cols], device; NULL
 * clears it), copied into the plan.  Applied as the reference applies
 * detector_gain: to the mean DP before the bright-field mask and the probe
 * radius (cuda/backend.py:220-226, 281-282, 442-443) and to every BF sample
 * before the scan FFT (cuda/backend.py:574-576).  Takes effect from the next
 * mean DP / Fourier stack / reconstruction.
 */
int ssb_set_detector_gain(ssb_plan* plan, const float* gain, void* stream);

/*
 * Bright-field selection on the device from the plan's mean DP
 * (cuda/backend.py:448-518).  bf_radius <= 0 selects the automatic
 * mean+std radius.  No host synchronisation; see ssb_get_bf.
 */
int ssb_select_bf(ssb_plan* plan, float threshold, float bf_radius, void* stream);

/*
 * Set the bright-field pixels from the host (raster order), with the disk
 * centre and dc = mean_b G_b(0,0) (SSBEngine(..., bf_inds_row, bf_inds_col,
 * bf_center, dc_value), cuda/engine.py:414-495).
 */
int ssb_set_bf(ssb_plan* plan, const int32_t* rows_host, const int32_t* cols_host, int nbf,
               float center_row, float center_col, float dc_re, float dc_im, void* stream);

/*
 * Copy the selection to the host (synchronises the stream).  info6_host
 * receives {center_row, center_col, radius^2, dc_re, dc_im, status}; status
 * 0 = ok, 1 = no candidates, 2 = empty disk, 3 = more than max_bf pixels.
 * rows_host / cols_host may be NULL; otherwise they hold `capacity` ints.
 */
int ssb_get_bf(ssb_plan* plan, int32_t* rows_host, int32_t* cols_host, int capacity, int* nbf_host,
               float* info6_host, void* stream);

/* Per-BF geometry for a scan rotation (SSBEngine.cache_rotation, cuda/engine.py:784-848). */
int ssb_cache_rotation(ssb_plan* plan, double rotation_rad, void* stream);

/* Device copies of the first `count` cached per-BF kx, ky (1/Angstrom) and aperture A(k). */
int ssb_get_geometry(ssb_plan* plan, float* kx_out, float* ky_out, float* aperture_out, int count,
                     void* stream);

/*
 * Resident Fourier stack from raw uint16 data (CudaSSBBackend._extract_gqk,
 * cuda/backend.py:520-589): BF gather, u16 -> f32, forward 2D FFT over the
 * scan axes, Hermitian half-plane kept in the plan.
 */
int ssb_prepare_g_u16(ssb_plan* plan, const uint16_t* data, void* stream);

/* ssb_prepare_g_u16 for frames of any QSSB_DTYPE_* element type (cuda/backend.py:520-589). */
int ssb_prepare_g_typed(ssb_plan* plan, const void* data, int dtype, void* stream);

/*
 * Export the resident stack in the reference layout G[b][q_row][q_col <= N/2]
 * (complex64); nbf must equal the resident BF count.
 */
int ssb_export_g(ssb_plan* plan, void* g_out, int nbf, void* stream);

/*
 * Import a stack G[b][q_row][q_col] (complex64, gcols = N/2+1 Hermitian or N
 * full columns) for nbf pixels; the BF selection must already be set.
 * Mirrors constructing SSBEngine from G_qk (cuda/engine.py:414-495).
 */
int ssb_import_g(ssb_plan* plan, const void* g_in, int nbf, int gcols, void* stream);

/*
 * Complex object and phase from the resident stack for C10, C12 (Angstrom)
 * and phi12 (rad) at the cached rotation: the Fourier-sum object path
 * (SSBEngine._reconstruct_object_fourier_sum, cuda/engine.py:1047-1107,
 * trotter kernel cuda/kernels/fft256.py:967-1030) and
 * phase = angle(object) (qg/ssb/results.py:117-122).  obj_out
 * ([N][N] complex64) and phase_out ([N][N] float32) may be NULL.
 */
int ssb_object(ssb_plan* plan, double c10, double c12, double phi12, void* obj_out, float* phase_out,
               void* stream);

/*
 * Exact per-BF phase path from the resident Fourier stack (Hermitian half
 * planes, or full planes after ssb_import_g with gcols = N) at the cached
 * rotation: for every BF pixel O_b = IFFT2(G_b conj(Gamma^_b)) with the DC bin
 * set to dc, phi_b = angle(O_b); phase_out = mean_b phi_b ([N][N] float32)
 * and loss_out = mean_R(mean_b phi_b^2 - phase^2) (one float32, device).
 * SSBEngine.reconstruct / reconstruct_with_loss (cuda/engine.py:1153-1191,
 * 1464-1514; kernels cuda/kernels/common.py:768-816, cuda/engine.py:21-73).
 * Either output may be NULL.  Deterministic: fixed reduction order.
 */
int ssb_phase_loss(ssb_plan* plan, double c10, double c12, double phi12, float* phase_out, float* loss_out,
                   void* stream);

/*
 * The exact per-BF phase path for the full 14-coefficient aberration
 * vector (host arrays; order C10, C12, C21, C23, C30, C32, C34, C41, C43,
 * C45, C50, C52, C54, C56; magnitudes in the reference's numeric convention,
 * angles in radians): chi = (2 pi/lambda) sum alpha^(n+1)/(n+1) C_nm
 * cos m(phi - phi_nm).  SSBEngine.reconstruct_full(_with_loss)
 * (cuda/engine.py:1208-1458; chi_full cuda/kernels/common.py:410-532;
 * packing cuda/backend.py:730-769).  Phase = mean per-BF phase.
 */
int ssb_phase_loss_full(ssb_plan* plan, const double* mags_host, const double* angles_host, float* phase_out,
                        float* loss_out, void* stream);

/*
 * Exact losses of n_trials aberration sets in one call, all enqueued on
 * `stream` with no host synchronisation: trial t uses mags_host[14t..],
 * angles_host[14t..] (layout of ssb_phase_loss_full) and, if rotation_host
 * is not NULL, the scan rotation rotation_host[t] (rad).  losses_out is a
 * device array of n_trials floats.  Trials without higher-order terms use
 * the quadratic-chi kernels.  Replaces the per-trial loop of the
 * reference's exact objective (SSBEngine._variance_loss_batch_reconstruct_fallback,
 * cuda/engine.py:1813-1836, driven from cuda/optimizer.py:59-230).
 */
int ssb_batch_phase_loss(ssb_plan* plan, int n_trials, const double* mags_host, const double* angles_host,
                         const double* rotation_host, float* losses_out, void* stream);

/*
 * ssb_batch_phase_loss as a captured hipGraph: the per-trial kernels read
 * their rotation and aberration coefficients from a plan-owned device table,
 * so the sequence is captured once (again when the count, the pattern of
 * higher-order trials or losses_out change) and replayed for each new batch
 * after the table upload, on the plan's own stream ordered after and before
 * `stream`.  Losses are bit-identical to ssb_batch_phase_loss (every trial
 * recomputes its geometry).  Synchronises `stream` once for the upload.
 * Replaces the per-trial loop of cuda/engine.py:1813-1836.
 */
int ssb_batch_phase_loss_graph(ssb_plan* plan, int n_trials, const double* mags_host, const double* angles_host,
                               const double* rotation_host, float* losses_out, void* stream);

/*
 * Batched aberration search over one or more datasets: a handle over
 * n_plans plans (at most 16; each with a resident Hermitian Fourier stack
 * from ssb_prepare_g_*, isotropic detector sampling, one wavelength) and up
 * to max_trials trial aberration sets.  Every trial is evaluated on every
 * dataset in one persistent launch whose work queue spans all datasets
 * (scan sizes may differ), replacing the reference's per-trial loop
 * (cuda/engine.py:1813-1836 driven by cuda/optimizer.py:59-191).  The
 * objective is the phase-image variance var_R angle(O_t(R)) of the
 * Fourier-sum object O_t = IFFT2(S_t), S_t = mean_b G_b conj(Gamma^_b(theta_t))
 * with the DC bin set to dc (the object path of cuda/engine.py:1047-1107 per
 * trial): a contrast measure that the correct aberrations maximise.  It is
 * not the reference's exact objective (ssb_batch_phase_loss).
 */
typedef struct ssb_trials ssb_trials;
int ssb_trials_create(ssb_plan* const* plans, int n_plans, int max_trials, ssb_trials** out);
int ssb_trials_destroy(ssb_trials* trials);

/*
 * Set n_trials aberration sets from host arrays in the layout of
 * ssb_batch_phase_loss (mags_host[14 t ..], angles_host[14 t ..]; odd-order
 * terms C21, C23, C41, C43, C45 must be zero) with per-trial scan rotations
 * rotation_host[t] (rad; NULL = 0).  Uploads them to the device and
 * synchronises `stream`.
 */
int ssb_trials_set(ssb_trials* trials, int n_trials, const double* mags_host, const double* angles_host,
                   const double* rotation_host, void* stream);

/*
 * Phase-image variance of every (plan, trial) into out[p * max_trials + t]
 * (device float32).  use_graph = 0 enqueues the kernels on `stream`;
 * use_graph = 1 captures them once into a hipGraph (again when n_trials or
 * `out` change) and replays it on the handle's own stream, ordered after
 * and before `stream`: new trials need only ssb_trials_set.  Both give the
 * same bits.  The launch takes its work items from a counter owned by the
 * handle (zero again when the launch ends), so calls on one handle must not
 * overlap in time; separate handles may run concurrently.
 */
int ssb_trials_phase_variance(ssb_trials* trials, float* out, int use_graph, void* stream);

/*
 * End-to-end plain SSB from resident raw data without host round trips:
 * mean DP, bright-field selection, geometry, gather + scan FFT fused with the
 * trotter reduction, inverse FFT and angle.  Equivalent to
 * CudaSSBBackend(data, ...) followed by result() (cuda/backend.py:168-851)
 * for C10/C12/phi12 aberrations.
 */
int ssb_recon_u16(ssb_plan* plan, const uint16_t* data, double c10, double c12, double phi12,
                  double rotation_rad, float threshold, float bf_radius, void* obj_out, float* phase_out,
                  void* stream);

/* ssb_recon_u16 for frames of any QSSB_DTYPE_* element type (cuda/backend.py:168-851). */
int ssb_recon_typed(ssb_plan* plan, const void* data, int dtype, double c10, double c12, double phi12,
                    double rotation_rad, float threshold, float bf_radius, void* obj_out, float* phase_out,
                    void* stream);

/*
 * ssb_recon_u16 with its raw-data stream pass ordered after the stream pass
 * of `after`'s most recent ssb_recon_u16_after call (a different plan,
 * usually on another stream): two plans alternating on two streams keep
 * two reconstructions in flight, so that one's VALU-bound trotter runs
 * beside the next one's HBM-bound stream pass while the two stream passes
 * never overlap.  Each plan records its own event after its stream pass;
 * the first call of a pair (nothing recorded on `after` yet) waits on
 * nothing.  Results equal ssb_recon_u16's.  No reference counterpart: the
 * reference reconstructs one dataset at a time (cuda/backend.py:168-851);
 * this serves series of scans.
 */
int ssb_recon_u16_after(ssb_plan* plan, const ssb_plan* after, const uint16_t* data, double c10, double c12,
                        double phi12, double rotation_rad, float threshold, float bf_radius, void* obj_out,
                        float* phase_out, void* stream);

/*
 * Decode HDF5 bitshuffle-LZ4 chunks (one frame or several per chunk) into a
 * uint16 array on the device: chunk c is comp[chunk_offsets[c] ..
 * + chunk_sizes[c]) (device arrays of int64; comp_bytes bytes in all) and
 * decodes to chunk_elems elements at out + c * chunk_elems.  block_bytes is
 * the filter's block size (8192 for the default uint16 blocks; a multiple of
 * 16 up to 8192); every chunk header must agree with it and with
 * chunk_elems.  workspace (device, 16-byte aligned, at least
 * ssb_bslz4_workspace_bytes(...) bytes) holds the status word, the
 * block table and a table of LZ4 sequences (8 bytes per entry,
 * block_bytes / 32 entries per block, at least 64 and at most what a block
 * can hold: 2 KiB per 8 KiB block, a quarter of the decoded size; blocks
 * with more sequences are still decoded).  With status_host non-NULL the call waits and reports a
 * malformed chunk as QSSB_ERR_ARG; with NULL it stays asynchronous and
 * the status word (int at workspace[0], 0 = ok) can be read later.
 * Replaces the CUDA bslz4 decoder (qg/io/backends/cuda/decoder.py:258-1146;
 * format as qg/io/backends/webgpu/bslz4.ts:1-30).
 */
long long ssb_bslz4_workspace_bytes(int n_chunks, long long chunk_elems, int block_bytes);

/*
 * Host-side writer of one bitshuffle-LZ4 chunk of n_elems uint16 values
 * (host memory) in the layout above; returns the chunk's byte count, or a
 * negative status.  out_capacity must be at least
 * ssb_bslz4_encode_bound(n_elems, block_bytes).  The reference compresses
 * its BF companion the same way on the GPU (qg/io/backends/cuda/decoder.py:466-1000).
 */
long long ssb_bslz4_encode_bound(long long n_elems, int block_bytes);
long long ssb_bslz4_encode_u16(const uint16_t* values_host, long long n_elems, int block_bytes, uint8_t* out_host,
                               long long out_capacity);
int ssb_bslz4_decode_u16(const uint8_t* comp, long long comp_bytes, const long long* chunk_offsets,
                         const long long* chunk_sizes, int n_chunks, long long chunk_elems, int block_bytes,
                         uint16_t* out, void* workspace, long long workspace_bytes, int* status_host, void* stream);

/*
 * Integer detector binning of resident uint16 frames src[frames][rows][cols]
 * into out[frames][rows/factor][cols/factor]: each output pixel is the
 * exact sum of its factor x factor block, dead pixels (bad_mask[rows][cols]
 * nonzero; NULL = none) counted as 0, trailing rows/columns that do not
 * fill a block trimmed.  out_bits 32: uint32 sums (the widened default of
 * the reference's binned loads); 16: the low 16 bits, as the reference's
 * uint64 sum cast back to uint16.  Asynchronous on `stream`.
 * Replaces _bin_sum (qg/io/backends/cpu/reference.py:27-42, dead pixels
 * zeroed first at :96-111) and the streaming bin of the CUDA loader
 * (qg/io/load.py:3009-3026, 3398-3415).
 */
int ssb_bin_u16(const uint16_t* src, long long frames, int rows, int cols, int factor, const uint8_t* bad_mask,
                int out_bits, void* out, void* stream);

/*
 * Tuning knobs of the reconstruction path (no reference counterpart: the
 * reference fixes its launch shapes per scan size, cuda/kernels/fft256.py:1051-1058).
 * QSSB_TUNE_STREAM_WG: workgroups of the raw-data stream pass;
 * QSSB_TUNE_GATHER_PAIRS: complex FFTs (two BF pixels each) per gather
 * workgroup; QSSB_TUNE_TROTTER: tile-shape variant of the fused column FFT +
 * trotter; QSSB_TUNE_SUPERSET: 2 (default) the gather reads the frames, 1
 * the stream pass of uint16 frames also copies a sampled bright-field box of
 * every frame compactly for the gather (bit-identical; its extra writes cost
 * the stream pass more than the gather saves at 256x256x128x128);
 * QSSB_TUNE_SPLIT: 2 (default) one pass, 1 the gather and the wave-local
 * trotter of a 256x256 scan in two halves of the BF list, the second half's
 * gather on a second stream beside the first half's trotter (the two group
 * the BF sums differently, so they agree to float rounding).  ssb_tune_candidates writes the
 * allowed values for a scan size
 * (the first is the default) and returns their count; ssb_tune_set accepts
 * one of them or 0 (default); ssb_tune_get reports the value in use.  The
 * knobs change launch shapes only: the mean DP and BF selection are
 * bit-identical for every value; gather and trotter shapes are separate
 * instantiations of the same arithmetic and agree to float rounding.
 */
#define QSSB_TUNE_STREAM_WG 0
#define QSSB_TUNE_GATHER_PAIRS 1
#define QSSB_TUNE_TROTTER 2
#define QSSB_TUNE_SUPERSET 3
#define QSSB_TUNE_SPLIT 4
int ssb_tune_candidates(int scan_n, int knob, int* values_host, int capacity);
int ssb_tune_set(ssb_plan* plan, int knob, int value);
int ssb_tune_get(ssb_plan* plan, int knob, int* value_host);

/*
 * Stage timing of ssb_recon_u16 with hipEvents recorded on the call's own
 * stream (the reference keeps wall-clock stage timings in SSBResult.timings,
 * qg/ssb/results.py:58-80).  ssb_timing_begin reserves events for up to
 * max_calls recon calls; ssb_timing_end synchronises, writes the summed
 * milliseconds of each stage (ssb_stage_count() entries) and the number of
 * recorded calls, and stops timing.  boundary_mask selects the stage
 * boundaries recorded (bit i = boundary before stage i, bit NSTAGES = end;
 * 0 = all); a stage is reported only when both of its boundaries are, so
 * a caller can time one kernel without events between the others.
 */
int ssb_stage_count(void);
const char* ssb_stage_name(int stage);
int ssb_timing_begin(ssb_plan* plan, int max_calls, int boundary_mask);
int ssb_timing_end(ssb_plan* plan, float* stage_ms_host, int n_stages, int* n_calls_host);

#ifdef __cplusplus
}
#endif

#endif /* QUANTEM_SSB_H */
