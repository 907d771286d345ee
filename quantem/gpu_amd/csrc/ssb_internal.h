// Internal declarations shared by the SSB HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ssb_device.h"

// Device-side status of the bright-field selection.
#define QSSB_DEV_NO_CANDIDATES 1
#define QSSB_DEV_EMPTY_DISK 2
#define QSSB_DEV_CAPACITY 3

namespace qssb {

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
  int pad;
};

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

// Matrix of the quadratic aberration phase chi(u) = u^T M u,
// M = pi*lambda*[[C10 + C12 cos2p, C12 sin2p], [C12 sin2p, C10 - C12 cos2p]].
struct QuadChi {
  float mxx, mxy, myy;
};

hipError_t launch_mean_dp(const uint16_t* data, long long frames, int D, uint32_t* partial, int max_ranges,
                          unsigned long long* sums, float* mean, hipStream_t s);
hipError_t launch_select_bf(const float* mean, const unsigned long long* sums, int det_rows, int det_cols,
                            float threshold, float bf_radius, int max_bf, int* rows, int* cols, BfInfo* info,
                            hipStream_t s);
hipError_t launch_bf_geometry(const int* rows, const int* cols, const BfInfo* info, int max_bf, df recip_y,
                              df recip_x, df cos_a, df sin_a, const Optics& o, float* kx, float* kxl, float* ky,
                              float* kyl, float* ak, hipStream_t s);
hipError_t launch_gather_rowfft(int N, const uint16_t* data, int D, int det_cols, const int* rows,
                                const int* cols, const BfInfo* info, int nbf_hint, const float2* tw, float2* H,
                                hipStream_t s);

// Column FFT of H fused with the trotter reduction (object path from raw
// data).  Writes per-BF-group partial sums S_part[g][c][q]; optionally the
// transposed Hermitian stack Gt[b][c][q], c = 0..N/2.
hipError_t launch_colfft_trotter(int N, const float2* H, const BfInfo* info, int nbf_hint, const GeomPtrs& gp,
                                 QuadChi m, const Optics& o, const float2* tw, float2* S_part, int max_groups,
                                 float2* Gt, bool do_trotter, hipStream_t s);
// Trotter reduction from a resident stack Gt[b][c][q] with `gcols` stored
// columns (N/2+1 Hermitian, or N full).
hipError_t launch_trotter_from_g(int N, int gcols, const float2* Gt, const BfInfo* info, int nbf_hint,
                                 const GeomPtrs& gp, QuadChi m, const Optics& o, float2* S_part, int max_groups,
                                 hipStream_t s);
// Sum the partial planes (Kahan, fixed order), divide by N_bf, set the DC
// bin, inverse 2D FFT (1/N^2), write object (row-major c64) and phase.
hipError_t launch_finish_object(int N, const float2* S_part, const BfInfo* info, int nbf_hint, int max_groups,
                                const float2* tw, float2* S, float2* obj, float* phase, hipStream_t s);

int group_count(int N, int nbf);
int colfft_group_size(int N);

}  // namespace qssb
