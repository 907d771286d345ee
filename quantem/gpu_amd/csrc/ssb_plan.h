// The plan object behind the opaque ssb_plan handle of quantem_ssb.h: one
// scan/detector geometry's device buffers and settings.  Shared by the
// translation units that implement entry points on plans.
#pragma once

#include <vector>

#include "ssb_internal.h"

using qssb::BfInfo;
using qssb::Optics;

// One trial of a captured exact-objective batch, read by the captured
// kernels from device memory: {cos, sin} of -rotation, the quadratic chi
// matrix and the full chi polynomial.
struct ExactTrial {
  qssb::df cs[2];
  qssb::QuadChi m;
  int higher;
  qssb::ChiPoly cp;
};

struct ssb_plan {
  int n = 0;
  int det_rows = 0, det_cols = 0, D = 0;
  int max_bf = 0;
  int max_groups = 0;
  int max_ranges = 256;
  int nbf_hint = 0;
  int gcols = 0;            // columns of the resident stack (0 = none)
  int g_nbf = 0;
  double wavelength = 0, semiangle = 0, ang_y = 0, ang_x = 0;
  double sampling_row = 0, sampling_col = 0;
  double recip_y = 0, recip_x = 0;
  bool geometry_valid = false;
  double rotation = 0;
  Optics optics{};
  // device buffers
  uint32_t* dp_partial = nullptr;
  void* dp_partial_wide = nullptr;   // 64-bit partials of 32-bit samples (allocated on first use)
  unsigned long long* dp_sums = nullptr;
  float* dp_mean = nullptr;
  float* dp_dcpix = nullptr;         // per-pixel G_b(0,0) contribution, f32(sum) * gain
  float* gain = nullptr;             // detector gain (NULL = none)
  int dp_dtype = QSSB_DTYPE_U16;     // sample type of the last mean DP
  int* bf_rows = nullptr;
  int* bf_cols = nullptr;
  // tuning knobs (ssb_tune_set): stream-pass workgroups, gather FFT pairs,
  // trotter tile shape; 0 = built-in default
  int tune_stream_wg = 0;
  int tune_gather_pairs = 0;
  int tune_trotter = 0;
  int tune_superset = 0;     // 1: compact BF superset copy; 0/2: the gather reads the frames
  int tune_split = 0;        // 0/1: gather and wave-local trotter in two BF halves on two streams; 2: one pass
  // bright-field superset of the reconstruction path (allocated on first use)
  uint32_t* sup_part = nullptr;
  qssb::SupBox* sup_box = nullptr;
  uint16_t* sup_C = nullptr;
  // recorded after the raw-data stream pass of each recon (ssb_recon_u16_after)
  hipEvent_t ev_stream = nullptr;
  bool ev_stream_recorded = false;
  BfInfo* info = nullptr;
  float* kx = nullptr;
  float* kxl = nullptr;
  float* ky = nullptr;
  float* kyl = nullptr;
  float* ak = nullptr;
  float* qx = nullptr;
  float* qxl = nullptr;
  float* qy = nullptr;
  float* qyl = nullptr;
  float2* tw = nullptr;
  float2* H = nullptr;
  float2* S_part = nullptr;
  float2* S = nullptr;
  float2* rowph = nullptr;
  float2* colph = nullptr;
  float2* echi = nullptr;
  float2* Gt = nullptr;
  size_t Gt_bytes = 0;
  int H_cap = 0;              // BF planes the row-transformed stack H holds
  // exact per-BF phase path (allocated on first use)
  float2* T = nullptr;        // column-transformed planes of one BF chunk
  float2* M1 = nullptr;       // phase moments per BF split, (sum, compensation)
  float2* M2 = nullptr;
  float* var_part = nullptr;  // per-block variance sums
  float* loss = nullptr;      // scalar loss when the caller passes none
  // captured exact-objective batch (ssb_batch_phase_loss_graph)
  ExactTrial* xt = nullptr;            // device table [xt_cap]
  int xt_cap = 0;
  std::vector<ExactTrial> xt_host;
  std::vector<int> xt_pattern;         // per trial: higher-order chi (topology of the graph)
  hipStream_t gstream = nullptr;       // the graph's own stream
  // second stream of the reconstruction: the gather of the second half of
  // the BF pixels runs on it while the first half's trotter runs on the
  // caller's stream
  hipStream_t pstream = nullptr;
  hipEvent_t pev_a = nullptr, pev_b = nullptr;
  hipEvent_t gev_in = nullptr, gev_out = nullptr;
  hipGraph_t xgraph = nullptr;
  hipGraphExec_t xexec = nullptr;
  float* xgraph_out = nullptr;
  // stage timing
  std::vector<hipEvent_t> events;   // (NSTAGES + 1) per recorded call
  int timing_calls = 0;
  int timing_max = 0;
  int timing_mask = 0;        // stage boundaries recorded (bit i = boundary i)
};
