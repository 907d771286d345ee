// Internal declarations shared by the SSB HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "ssb_device.h"
#include "quantem_ssb.h"

// Device-side status of the bright-field selection.
#define QSSB_DEV_NO_CANDIDATES 1
#define QSSB_DEV_EMPTY_DISK 2
#define QSSB_DEV_CAPACITY 3

namespace qssb {

// Record the calling thread's last error (ssb_last_error) and return `code`
// (or the status of a HIP error); clear it at the start of an entry point.
int report_error(int code, const std::string& msg);
int report_hip_error(hipError_t e, const char* where);
void clear_error();

// Bright-field selection result, resident in device memory so that the
// pipeline can run without a host round trip.
struct BfInfo {
  int nbf;
  int status;
  float center_r;
  float center_c;
  float radius2;
  float dc_re;
  float dc_im;
  int sup_ok;   // every BF pixel lies in the compact superset box (SupBox) of this recon
};

// Superset of the bright field for one reconstruction: a detector box
// [r0, r0 + bh) x [c0, c0 + bw) estimated from sampled frames before the
// stream pass, which then also writes every frame's box compactly
// (C[frame][row - r0][col - c0], bh * bw uint16 per frame).  bw == 0: none.
struct SupBox {
  int r0, c0, bh, bw;
};
// Largest box (pixels) the compact copy holds.
constexpr int kSupMaxPixels = 2048;

// Aberration set in the reference's numeric convention (Angstrom, radians).
struct Theta {
  float c10, c12, phi12, rot;
};

// Device arrays of the per-BF detector coordinates (double-float hi/lo),
// the BF aperture A(k), and the scan frequencies (hi/lo).
struct GeomPtrs {
  const float* kx;
  const float* kxl;
  const float* ky;
  const float* kyl;
  const float* ak;
  const float* qx;
  const float* qxl;
  const float* qy;
  const float* qyl;
};

// Krivanek aberration phase as a polynomial in u = (ux, uy): for each of the
// 14 coefficients C_nm (order C10, C12, C21, C23, C30, C32, C34, C41, C43,
// C45, C50, C52, C54, C56), chi += K * |u|^(n+1-m) * (ca Re(z^m) + sa Im(z^m))
// with z = ux + i uy, K = (2 pi / lambda) lambda^(n+1) C_nm / (n+1),
// (ca, sa) = (cos m phi_nm, sin m phi_nm).  Odd m are the odd part of chi.
struct ChiPoly {
  float K[14];
  float ca[14];
  float sa[14];
  int odd;   // any odd-m term (C21, C23, C41, C43, C45) non-zero: chi(-u) != chi(u)
};

// Matrix of the quadratic aberration phase chi(u) = u^T M u,
// M = pi*lambda*[[C10 + C12 cos2p, C12 sin2p], [C12 sin2p, C10 - C12 cos2p]].
struct QuadChi {
  float mxx, mxy, myy;
};

// Phasor tables of the factorised overlap e^{i theta} = e^{i qx a_b} e^{i qy b_b},
// (a_b, b_b) = 2 M k_b: rowph[b][q] for all rows, colph[b][c] for columns
// c = 0..N/2 (the other columns are their conjugates); and the plane
// echi[c][q] = e^{i chi(Q)} (column-major, like the partial sums).  Written
// by k_phasors, or by the BF gather of a reconstruction on the side (the
// same function: the same values).
struct PhasorJob {
  const BfInfo* info;
  GeomPtrs gp;
  QuadChi m;
  int N;
  float2* rowph;
  float2* colph;
  float2* echi;
};

// The items [gid, gid + stride, ...) of a phasor job.
__device__ __forceinline__ void phasor_work(const PhasorJob& j, long long gid, long long stride) {
  const QuadChi m = j.m;
  const int N = j.N;
  const int nbf = j.info->nbf;
  const int ncol = N / 2 + 1;
  const long long per_b = (long long)N + ncol;
  const long long total = (long long)nbf * per_b;
  const long long NN = (long long)N * N;
  for (long long idx = gid; idx < NN; idx += stride) {
    const int c = (int)(idx / N), q = (int)(idx - (long long)c * N);
    const float x = j.gp.qx[q], y = j.gp.qy[c];
    float sn, cs;
    sincosf(m.mxx * x * x + 2.0f * m.mxy * x * y + m.myy * y * y, &sn, &cs);
    j.echi[idx] = make_float2(cs, sn);
  }
  for (long long idx = gid; idx < total; idx += stride) {
    const int b = (int)(idx / per_b);
    const int r = (int)(idx - (long long)b * per_b);
    const float kx = j.gp.kx[b], ky = j.gp.ky[b];
    float sn, cs;
    if (r < N) {
      const float ta = 2.0f * (m.mxx * kx + m.mxy * ky);
      sincosf(j.gp.qx[r] * ta, &sn, &cs);
      j.rowph[(long long)b * N + r] = make_float2(cs, sn);
    } else {
      const int c = r - N;
      const float tb = 2.0f * (m.mxy * kx + m.myy * ky);
      sincosf(j.gp.qy[c] * tb, &sn, &cs);
      j.colph[(long long)b * ncol + c] = make_float2(cs, sn);
    }
  }
}

// Bytes per sample of a QSSB_DTYPE_* code (0 = unknown); whether its
// stream-pass partials are 64-bit (uint32, int32, float32 samples).
int dtype_size(int dtype);
bool dtype_wide_partials(int dtype);

// Mean DP of `frames` frames of D samples of type `dtype`.  partial holds
// max_ranges * D partial sums (32-bit, or 64-bit when dtype_wide_partials);
// target_wg: workgroups of the stream pass (0 = default 1536).  gain (D
// floats) may be NULL.  sums: exact uint64 sums (integer samples; 0 for
// float32), mean: f32(sum)/f32(frames) * gain, dcpix: f32(sum) * gain
// (signed samples: the signed sum).
hipError_t launch_mean_dp(int dtype, const void* data, long long frames, int D, void* partial, int max_ranges,
                          int target_wg, const float* gain, unsigned long long* sums, float* mean, float* dcpix,
                          hipStream_t s);
// the two passes of launch_mean_dp, separately (stage timing)
hipError_t launch_mean_dp_stream(int dtype, const void* data, long long frames, int D, void* partial, int max_ranges,
                                 int target_wg, hipStream_t s);
hipError_t launch_mean_dp_reduce(int dtype, const void* data, long long frames, int D, const void* partial,
                                 int max_ranges, int target_wg, const float* gain, unsigned long long* sums,
                                 float* mean, float* dcpix, hipStream_t s);
// exact_dc: dc from the exact unsigned integer sums (no gain); otherwise
// from a compensated sum of dcpix over the selected pixels.  box (may be
// NULL): info->sup_ok = every selected pixel inside it.
hipError_t launch_select_bf(const float* mean, const unsigned long long* sums, const float* dcpix, bool exact_dc,
                            int det_rows, int det_cols, float threshold, float bf_radius, int max_bf, int* rows,
                            int* cols, BfInfo* info, hipStream_t s, const SupBox* box = nullptr);
// Superset box from kSupSamples frames spread over the scan (uint16
// frames; part holds kSupSplits * D uint32).
hipError_t launch_superset(const uint16_t* data, long long frames, int det_rows, int det_cols, uint32_t* part,
                           SupBox* box, hipStream_t s);
int superset_part_count();
// The uint16 stream pass that also writes each frame's superset box into C
// (frames * kSupMaxPixels uint16); needs det_cols % 8 == 0 and 16-byte
// aligned data.
hipError_t launch_mean_dp_stream_sup(const uint16_t* data, long long frames, int D, int det_cols, uint32_t* partial,
                                     int max_ranges, int target_wg, const SupBox* box, uint16_t* C, hipStream_t s);
// Probe disk centre and radius of a float32 mean DP; out[4] = {status, row, col, radius}.
hipError_t launch_detect_bf_radius(const float* mean, int R, int C, df ratio, int* out, hipStream_t s);
// rot_dev (device, may be NULL): {cos_a, sin_a} read in place of the arguments.
hipError_t launch_bf_geometry(const int* rows, const int* cols, const BfInfo* info, int max_bf, df recip_y,
                              df recip_x, df cos_a, df sin_a, const Optics& o, float* kx, float* kxl, float* ky,
                              float* kyl, float* ak, hipStream_t s, const df* rot_dev = nullptr);
// pairs: complex FFTs (2 BF pixels each) per workgroup, one of
// gather_pairs_candidates(N) (0 = gather_pairs_default(N)).
int gather_pairs_default(int N);
int gather_pairs_candidates(int N, int* out);   // out[2]; returns the count
// Samples of type dtype become float32 times gain[pixel] (gain may be NULL).
// With box and C (uint16 frames): when info->sup_ok the samples come from
// the compact superset copy C instead of the frames (the same values).
// With job non-NULL the gather's threads also write that phasor job's
// tables (in place of a k_phasors launch after the gather).
hipError_t launch_gather_rowfft(int N, int dtype, const void* data, int D, int det_cols, const int* rows,
                                const int* cols, const BfInfo* info, int nbf_hint, const float2* tw, int pairs,
                                const float* gain, float2* H, hipStream_t s, const SupBox* box = nullptr,
                                const uint16_t* C = nullptr, int part = 0, int parts = 1,
                                const PhasorJob* job = nullptr);
// Exact BF columns out[b][s] for BF pixels [bf_offset, bf_offset + count)
// of the list rows/cols: uint8 frames -> uint8 (in_u8), uint16 frames ->
// uint16, or uint8 when narrow.
hipError_t launch_bf_columns(const void* data, bool in_u8, long long frames, int D, int det_cols, const int* rows,
                             const int* cols, int bf_offset, int count, bool narrow, void* out, hipStream_t s);
// max of n uint16 values into out[0] (partial holds max_u16_parts() uints).
hipError_t launch_max_u16(const uint16_t* data, long long n, unsigned int* partial, unsigned int* out,
                          hipStream_t s);
int max_u16_parts();

// Scan normalisation: crop/pad the H x W scan to T x T (partial holds
// normalize_partial_groups(QSSB_DTYPE_U16, H, T, D) * D uint32, mean D uint16).
hipError_t launch_normalize_scan(const uint16_t* src, int H, int W, int D, uint16_t* dst, int T,
                                 uint32_t* partial, uint16_t* mean, hipStream_t s);
int normalize_partial_groups(int dtype, int H, int T, int D);
// The same for any QSSB_DTYPE_* sample type; partial and mean as laid out by
// normalize_workspace_bytes (mean at *mean_offset of the workspace).
hipError_t launch_normalize_scan_typed(int dtype, const void* src, int H, int W, int D, void* dst, int T,
                                       void* partial, void* mean, hipStream_t s);
size_t normalize_workspace_bytes(int dtype, int H, int T, int D, size_t* mean_offset);

// Constants of the isotropic aperture evaluated from d = Rout^2 - |u|^2.
struct ApIso {
  float rout2;      // Rout^2 (FP32 copy)
  float d_in;       // d above which |u| <= inner radius: A == 1
  float d_out;      // d below which A == 0 (negative margin)
  float rout;       // Rout
  float scale;      // lambda / ang
};

// Number of BF groups (partial planes) of the object-path trotter's tile
// kernels, and the most any of its kernels uses (the partial-plane buffer).
int object_groups(int N);
int trotter_groups_max(int N);
bool optics_isotropic(const Optics& o);
ApIso iso_aperture(const Optics& o, df& rout2_df);

// Phasor tables rowph[b][q] = e^{i qx[q] a_b}, colph[b][c] = e^{i qy[c] b_b}
// (c = 0..N/2) and the plane echi[c][q] = e^{i chi(Q)} for the current
// aberrations.
hipError_t launch_phasors(int N, const BfInfo* info, int nbf_hint, const GeomPtrs& gp, QuadChi m, float2* rowph,
                          float2* colph, float2* echi, hipStream_t s,
                          const QuadChi* m_dev = nullptr);
// Trotter over Hermitian slot columns: from the row-transformed stack H
// (fused scan-column FFT; from_h) or from the resident stack Gt[b][c][q],
// c = 0..N/2.  Writes partial planes S_part[g][c][q] (do_trotter) and, if Gt
// is given with from_h, the resident stack itself.
// variant: kernel of the reconstruction path's trotter, 0 (default) to
// trotter_variants(N) - 1 (at N = 256, 0 is the wave-local kernel and 1..3
// the tile shapes); other paths always use variant 0.  groups_used receives
// the number of partial planes written (launch_finish_object's `groups`).
int trotter_variants(int N);
hipError_t launch_trotter_slots(int N, bool from_h, const float2* src, const BfInfo* info, const GeomPtrs& gp,
                                const float2* rowph, const float2* colph, const float2* echi, const Optics& o,
                                const float2* tw, float2* S_part, float2* Gt, bool do_trotter, int variant,
                                hipStream_t s, int* groups_used = nullptr, int part = 0, int parts = 1);
// Whether launch_trotter_slots takes the wave-local kernel (which can run
// in BF parts: part p of `parts` sums its share of the BF pixels into its
// share of the partial planes).
bool trotter_wave_path(int N, bool from_h, const Optics& o, bool do_trotter, bool has_gt, int variant);
// Trotter from an imported full-plane (non-Hermitian) stack Gt[b][c][q], c < N.
hipError_t launch_trotter_full(int N, const float2* Gt, const BfInfo* info, const GeomPtrs& gp, const float2* rowph,
                               QuadChi m, const Optics& o, float2* S_part, hipStream_t s);
// Sum the partial planes (Kahan, fixed order), divide by N_bf, set the DC
// bin, inverse 2D FFT (1/N^2), write object (row-major c64) and phase.
hipError_t launch_finish_object(int N, const float2* S_part, int groups, const BfInfo* info, const float2* tw,
                                float2* S, float2* obj, float* phase, hipStream_t s);

// Exact per-BF phase path: BF pixels per chunk of column-transformed planes,
// BF splits of the moment planes.
int loss_chunk_bf(int N);
int loss_splits(int N);
// Mean of per-BF phases and the phase-variance loss from the resident
// stack Gt[b][c][q] (gcols = N/2+1 Hermitian, or N full-plane), with the
// phasor tables of a quadratic chi, or (poly != NULL) the full chi; T holds loss_chunk_bf(N) planes, M1/M2
// loss_splits(N) planes, var_part N*N/256 floats.  phase may be NULL; loss
// is a device scalar.
hipError_t launch_phase_moments(int N, const float2* Gt, int gcols, const BfInfo* info, int nbf_hint, const GeomPtrs& gp,
                                const float2* rowph, const float2* colph, const float2* echi, const ChiPoly* poly,
                                const Optics& o,
                                const float2* tw, float2* T, float2* M1, float2* M2, float* var_part,
                                float* phase, float* loss, hipStream_t s,
                                const ChiPoly* poly_dev = nullptr);

// Learned-initializer features of the resident Hermitian stack at the 8
// scan-frequency grid points (qrow[i], qcol[i]) (device arrays): out is
// [32][32][16] (NHWC).
hipError_t launch_features_from_g(const float2* Gt, int N, const int* bf_rows, const int* bf_cols,
                                  const BfInfo* info, const int* qrow, const int* qcol, float* out, hipStream_t s);

// bitshuffle-LZ4 chunk decode (ssb_bslz4.hip)
long long bslz4_slots(long long chunk_elems, int block_elems);
long long bslz4_workspace_bytes(int n_chunks, long long chunk_elems, int block_elems);
long long bslz4_encode_bound(long long n_elems, int block_bytes);
long long bslz4_encode_u16(const uint16_t* v, long long n, int block_bytes, uint8_t* dst, long long cap);
// integer detector binning (ssb_bin.hip): wide = uint32 output, else uint16
// (low 16 bits); bad = per-pixel dead mask (nonzero = dead) or nullptr
hipError_t launch_bin_u16(const uint16_t* src, long long frames, int rows, int cols, int factor,
                          const uint8_t* bad, bool wide, void* out, hipStream_t s);
hipError_t launch_bslz4_decode_u16(const uint8_t* comp, long long comp_bytes, const long long* chunk_off,
                                   const long long* chunk_len, int n_chunks, long long chunk_elems, int block_elems,
                                   uint16_t* out, void* workspace, hipStream_t s);

}  // namespace qssb
