// C ABI of the MI355X SSB engine: plan management and pipeline launches.
// Declarations and reference citations: include/quantem_ssb.h.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/quantem_ssb.h"
#include "ssb_internal.h"

using namespace qssb;

#include "ssb_plan.h"

namespace {
// Stages of ssb_recon_u16, in launch order.
const char* const kStageNames[] = {"superset", "mean_dp_stream", "mean_dp_reduce", "select_bf", "geometry",
                                   "gather_rowfft", "colfft_trotter", "finish_object"};
constexpr int NSTAGES = 8;
}

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  std::string msg = std::string(where) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return fail(QSSB_ERR_OOM, msg);
  return fail(QSSB_ERR_LAUNCH, msg);
}

#define HIP_TRY(expr, where)              \
  do {                                    \
    hipError_t _e = (expr);               \
    if (_e != hipSuccess) return hip_fail(_e, where); \
  } while (0)

template <typename T>
hipError_t dmalloc(T** p, size_t count) {
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T) > 0 ? count * sizeof(T) : 1);
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

df to_df(double v) {
  const float hi = (float)v;
  return df{hi, (float)(v - (double)hi)};
}

GeomPtrs geom_ptrs(const ssb_plan* p) {
  return GeomPtrs{p->kx, p->kxl, p->ky, p->kyl, p->ak, p->qx, p->qxl, p->qy, p->qyl};
}

QuadChi quad_chi(const ssb_plan* p, double c10, double c12, double phi12) {
  const double f = M_PI * p->wavelength;
  const double c = std::cos(2.0 * phi12), s = std::sin(2.0 * phi12);
  QuadChi m;
  m.mxx = (float)(f * (c10 + c12 * c));
  m.mxy = (float)(f * (c12 * s));
  m.myy = (float)(f * (c10 - c12 * c));
  return m;
}

int check_plan(const ssb_plan* p) {
  if (!p) return fail(QSSB_ERR_ARG, "plan is NULL");
  g_last_error.clear();
  return QSSB_OK;
}

int ensure_geometry(ssb_plan* p, hipStream_t s) {
  const df ca = to_df(std::cos(-p->rotation)), sa = to_df(std::sin(-p->rotation));
  hipError_t e = launch_bf_geometry(p->bf_rows, p->bf_cols, p->info, p->max_bf, to_df(p->recip_y),
                                    to_df(p->recip_x), ca, sa, p->optics, p->kx, p->kxl, p->ky, p->kyl, p->ak, s);
  if (e != hipSuccess) return hip_fail(e, "bf_geometry");
  p->geometry_valid = true;
  return QSSB_OK;
}

// Gt layout [b][c][q] (q contiguous) <-> reference layout [b][q][c].
__global__ void k_export_g(const float2* __restrict__ Gt, int gcols, int n, int nbf, float2* __restrict__ out,
                           int out_cols) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long total = (long long)nbf * n * out_cols;
  if (idx >= total) return;
  const int c = (int)(idx % out_cols);
  const long long t = idx / out_cols;
  const int q = (int)(t % n);
  const long long b = t / n;
  out[idx] = Gt[(b * gcols + c) * n + q];
}

__global__ void k_import_g(const float2* __restrict__ in, int gcols, int n, int nbf, float2* __restrict__ Gt) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long total = (long long)nbf * n * gcols;
  if (idx >= total) return;
  const int c = (int)(idx % gcols);
  const long long t = idx / gcols;
  const int q = (int)(t % n);
  const long long b = t / n;
  Gt[(b * gcols + c) * n + q] = in[idx];
}

}  // namespace

namespace qssb {
// The thread's last-error record, for entry points in other translation units.
int report_error(int code, const std::string& msg) { return fail(code, msg); }
int report_hip_error(hipError_t e, const char* where) { return hip_fail(e, where); }
void clear_error() { g_last_error.clear(); }
}  // namespace qssb

static void release_events(ssb_plan* p);

extern "C" {

int ssb_version(void) { return 100; }

const char* ssb_last_error(void) { return g_last_error.c_str(); }

int ssb_plan_create(int scan_n, int det_rows, int det_cols, int max_bf, double wavelength_A,
                    double semiangle_rad, double ang_row_rad, double ang_col_rad, double scan_sampling_row_A,
                    double scan_sampling_col_A, ssb_plan** out) {
  g_last_error.clear();
  if (!out) return fail(QSSB_ERR_ARG, "plan output pointer is NULL");
  *out = nullptr;
  if (scan_n != 128 && scan_n != 256 && scan_n != 512 && scan_n != 1024)
    return fail(QSSB_ERR_UNSUPPORTED, "scan size " + std::to_string(scan_n) +
                                          " is not supported; use 128, 256, 512 or 1024");
  if (det_rows < 1 || det_cols < 1 || max_bf < 1)
    return fail(QSSB_ERR_ARG, "detector shape and max_bf must be positive");
  if ((long long)det_rows * det_cols > 65536)
    return fail(QSSB_ERR_UNSUPPORTED, "detectors above 65536 pixels are not supported; bin the detector first");
  if (!(wavelength_A > 0) || !(semiangle_rad > 0) || !(ang_row_rad > 0) || !(ang_col_rad > 0) ||
      !(scan_sampling_row_A > 0) || !(scan_sampling_col_A > 0))
    return fail(QSSB_ERR_ARG, "optics and sampling must be positive");
  ssb_plan* p = new ssb_plan();
  p->n = scan_n;
  p->det_rows = det_rows;
  p->det_cols = det_cols;
  p->D = det_rows * det_cols;
  p->max_bf = max_bf;
  p->nbf_hint = max_bf;
  p->max_groups = trotter_groups_max(scan_n);
  p->wavelength = wavelength_A;
  p->semiangle = semiangle_rad;
  p->ang_y = ang_row_rad;
  p->ang_x = ang_col_rad;
  p->sampling_row = scan_sampling_row_A;
  p->sampling_col = scan_sampling_col_A;
  // reciprocal sampling: ang/lambda, re-derived through the real-space
  // sampling 1/(recip*det) exactly as the engine does (cuda/backend.py:316-327)
  const double ry = ang_row_rad / wavelength_A, rx = ang_col_rad / wavelength_A;
  const double sy = 1.0 / (ry * det_rows), sx = 1.0 / (rx * det_cols);
  p->recip_y = 1.0 / (sy * det_rows);
  p->recip_x = 1.0 / (sx * det_cols);
  Optics& o = p->optics;
  o.wavelength = (float)wavelength_A;
  o.semiangle = (float)semiangle_rad;
  o.ang_y = (float)ang_row_rad;
  o.ang_x = (float)ang_col_rad;
  {
    const float max_ang = std::fmax(o.ang_y, o.ang_x);
    const float inner = (o.semiangle - 0.5f * max_ang) / o.wavelength;
    o.inner2 = (inner > 0.0f) ? inner * inner : 0.0f;
    const float outer = (o.semiangle + 0.5f * max_ang) / o.wavelength;
    o.outer2 = outer * outer * 1.0001f;
    o.lam = to_df(wavelength_A);
    o.semi = to_df(semiangle_rad);
    o.angy = to_df(ang_row_rad);
    o.angx = to_df(ang_col_rad);
  }
  const int n = scan_n;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = dmalloc(&p->dp_partial, (size_t)p->max_ranges * p->D);
  if (e == hipSuccess) e = dmalloc(&p->dp_sums, (size_t)p->D);
  if (e == hipSuccess) e = dmalloc(&p->dp_mean, (size_t)p->D);
  if (e == hipSuccess) e = dmalloc(&p->dp_dcpix, (size_t)p->D);
  if (e == hipSuccess) e = dmalloc(&p->bf_rows, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->bf_cols, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->info, 1);
  if (e == hipSuccess) e = dmalloc(&p->kx, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->kxl, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->ky, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->kyl, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->ak, (size_t)max_bf);
  if (e == hipSuccess) e = dmalloc(&p->qx, (size_t)n);
  if (e == hipSuccess) e = dmalloc(&p->qxl, (size_t)n);
  if (e == hipSuccess) e = dmalloc(&p->qy, (size_t)n);
  if (e == hipSuccess) e = dmalloc(&p->qyl, (size_t)n);
  if (e == hipSuccess) e = dmalloc(&p->tw, (size_t)n);
  if (e == hipSuccess) e = dmalloc(&p->S_part, (size_t)p->max_groups * n * n);
  if (e == hipSuccess) e = dmalloc(&p->S, (size_t)n * n);
  if (e == hipSuccess) e = dmalloc(&p->rowph, (size_t)max_bf * n);
  if (e == hipSuccess) e = dmalloc(&p->colph, (size_t)max_bf * (n / 2 + 1));
  if (e == hipSuccess) e = dmalloc(&p->echi, (size_t)n * n);
  if (e != hipSuccess) {
    ssb_plan_destroy(p);
    return hip_fail(e, "ssb_plan_create allocation");
  }
  std::vector<float> qx(n), qy(n), qxl(n), qyl(n);
  std::vector<float2> tw(n);
  for (int i = 0; i < n; ++i) {
    const int k = (i < (n + 1) / 2) ? i : i - n;   // numpy fftfreq ordering
    const df fx = to_df((double)k / (n * scan_sampling_row_A));
    const df fy = to_df((double)k / (n * scan_sampling_col_A));
    qx[i] = fx.hi;
    qxl[i] = fx.lo;
    qy[i] = fy.hi;
    qyl[i] = fy.lo;
    const double a = -2.0 * M_PI * (double)i / (double)n;
    tw[i] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  BfInfo zero{};
  if (e == hipSuccess) e = hipMemcpy(p->qx, qx.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->qy, qy.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->qxl, qxl.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->qyl, qyl.data(), n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->tw, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->info, &zero, sizeof(BfInfo), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    ssb_plan_destroy(p);
    return hip_fail(e, "ssb_plan_create upload");
  }
  *out = p;
  return QSSB_OK;
}

int ssb_set_detector_sampling(ssb_plan* p, double recip_row, double recip_col) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!(recip_row > 0) || !(recip_col > 0)) return fail(QSSB_ERR_ARG, "reciprocal sampling must be positive");
  p->recip_y = recip_row;
  p->recip_x = recip_col;
  p->geometry_valid = false;
  return QSSB_OK;
}

int ssb_set_scan_frequencies(ssb_plan* p, const float* qrow_host, const float* qcol_host) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!qrow_host || !qcol_host) return fail(QSSB_ERR_ARG, "scan frequencies are NULL");
  HIP_TRY(hipMemcpy(p->qx, qrow_host, (size_t)p->n * 4, hipMemcpyHostToDevice), "set_scan_frequencies");
  HIP_TRY(hipMemcpy(p->qy, qcol_host, (size_t)p->n * 4, hipMemcpyHostToDevice), "set_scan_frequencies");
  HIP_TRY(hipMemset(p->qxl, 0, (size_t)p->n * 4), "set_scan_frequencies");
  HIP_TRY(hipMemset(p->qyl, 0, (size_t)p->n * 4), "set_scan_frequencies");
  return QSSB_OK;
}

int ssb_plan_destroy(ssb_plan* p) {
  if (!p) return QSSB_OK;
  dfree(p->dp_partial);
  dfree(p->dp_partial_wide);
  dfree(p->dp_sums);
  dfree(p->dp_mean);
  dfree(p->dp_dcpix);
  dfree(p->gain);
  dfree(p->bf_rows);
  dfree(p->bf_cols);
  if (p->ev_stream) (void)hipEventDestroy(p->ev_stream);
  dfree(p->info);
  dfree(p->kx);
  dfree(p->kxl);
  dfree(p->ky);
  dfree(p->kyl);
  dfree(p->ak);
  dfree(p->qx);
  dfree(p->qxl);
  dfree(p->qy);
  dfree(p->qyl);
  dfree(p->tw);
  dfree(p->H);
  dfree(p->S_part);
  dfree(p->S);
  dfree(p->rowph);
  dfree(p->colph);
  dfree(p->echi);
  dfree(p->Gt);
  dfree(p->T);
  dfree(p->M1);
  dfree(p->M2);
  dfree(p->var_part);
  dfree(p->loss);
  dfree(p->xt);
  dfree(p->sup_part);
  dfree(p->sup_box);
  dfree(p->sup_C);
  if (p->xexec) (void)hipGraphExecDestroy(p->xexec);
  if (p->xgraph) (void)hipGraphDestroy(p->xgraph);
  if (p->gev_in) (void)hipEventDestroy(p->gev_in);
  if (p->gev_out) (void)hipEventDestroy(p->gev_out);
  if (p->gstream) (void)hipStreamDestroy(p->gstream);
  if (p->pev_a) (void)hipEventDestroy(p->pev_a);
  if (p->pev_b) (void)hipEventDestroy(p->pev_b);
  if (p->pstream) (void)hipStreamDestroy(p->pstream);
  release_events(p);
  delete p;
  return QSSB_OK;
}

static bool dtype_unsigned_int(int dtype) {
  return dtype == QSSB_DTYPE_U8 || dtype == QSSB_DTYPE_U16 || dtype == QSSB_DTYPE_U32;
}

// partial-sum buffer of the stream pass for this sample type
static int dp_partial_for(ssb_plan* p, int dtype, void** out) {
  if (!dtype_wide_partials(dtype)) {
    *out = p->dp_partial;
    return QSSB_OK;
  }
  if (!p->dp_partial_wide) {
    hipError_t e = hipMalloc(&p->dp_partial_wide, (size_t)p->max_ranges * p->D * 8);
    if (e != hipSuccess) {
      p->dp_partial_wide = nullptr;
      return hip_fail(e, "mean DP partial allocation");
    }
  }
  *out = p->dp_partial_wide;
  return QSSB_OK;
}

static int mean_dp_typed(ssb_plan* p, const void* data, int dtype, int64_t frames, uint64_t* sums_out,
                         float* mean_out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!data || frames < 1) return fail(QSSB_ERR_ARG, "data is NULL or has no frames");
  if (!dtype_size(dtype)) return fail(QSSB_ERR_ARG, "unknown sample type " + std::to_string(dtype));
  if (sums_out && dtype == QSSB_DTYPE_F32)
    return fail(QSSB_ERR_ARG, "float32 frames have no integer sums; pass sums_out = NULL");
  void* partial = nullptr;
  rc = dp_partial_for(p, dtype, &partial);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  HIP_TRY(launch_mean_dp(dtype, data, frames, p->D, partial, p->max_ranges, p->tune_stream_wg, p->gain, p->dp_sums,
                         p->dp_mean, p->dp_dcpix, s),
          "mean_dp");
  p->dp_dtype = dtype;
  if (sums_out)
    HIP_TRY(hipMemcpyAsync(sums_out, p->dp_sums, (size_t)p->D * 8, hipMemcpyDeviceToDevice, s), "mean_dp copy");
  if (mean_out)
    HIP_TRY(hipMemcpyAsync(mean_out, p->dp_mean, (size_t)p->D * 4, hipMemcpyDeviceToDevice, s), "mean_dp copy");
  return QSSB_OK;
}

int ssb_dtype_size(int dtype) { return dtype_size(dtype); }

int ssb_mean_dp_u16(ssb_plan* p, const uint16_t* data, int64_t frames, uint64_t* sums_out, float* mean_out,
                    void* stream) {
  return mean_dp_typed(p, data, QSSB_DTYPE_U16, frames, sums_out, mean_out, stream);
}

int ssb_mean_dp_typed(ssb_plan* p, const void* data, int dtype, int64_t frames, uint64_t* sums_out,
                      float* mean_out, void* stream) {
  return mean_dp_typed(p, data, dtype, frames, sums_out, mean_out, stream);
}

int ssb_set_detector_gain(ssb_plan* p, const float* gain, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  if (!gain) {
    if (p->gain) {
      HIP_TRY(hipStreamSynchronize(s), "set_detector_gain sync");
      dfree(p->gain);
      p->gain = nullptr;
    }
    return QSSB_OK;
  }
  if (!p->gain) HIP_TRY(dmalloc(&p->gain, (size_t)p->D), "detector gain allocation");
  HIP_TRY(hipMemcpyAsync(p->gain, gain, (size_t)p->D * 4, hipMemcpyDeviceToDevice, s), "set_detector_gain");
  return QSSB_OK;
}

static bool exact_dc(const ssb_plan* p) { return dtype_unsigned_int(p->dp_dtype) && !p->gain; }

int ssb_select_bf(ssb_plan* p, float threshold, float bf_radius, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  HIP_TRY(launch_select_bf(p->dp_mean, p->dp_sums, p->dp_dcpix, exact_dc(p), p->det_rows, p->det_cols, threshold,
                           bf_radius, p->max_bf, p->bf_rows, p->bf_cols, p->info, s),
          "select_bf");
  p->geometry_valid = false;
  p->gcols = 0;
  return QSSB_OK;
}

int ssb_set_bf(ssb_plan* p, const int32_t* rows_host, const int32_t* cols_host, int nbf, float center_row,
               float center_col, float dc_re, float dc_im, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!rows_host || !cols_host || nbf < 1) return fail(QSSB_ERR_ARG, "BF rows/cols must be non-empty");
  if (nbf > p->max_bf)
    return fail(QSSB_ERR_ARG, "nbf " + std::to_string(nbf) + " exceeds plan capacity " + std::to_string(p->max_bf));
  for (int i = 0; i < nbf; ++i) {
    if (rows_host[i] < 0 || rows_host[i] >= p->det_rows || cols_host[i] < 0 || cols_host[i] >= p->det_cols)
      return fail(QSSB_ERR_ARG, "BF coordinate outside the detector");
  }
  hipStream_t s = as_stream(stream);
  BfInfo info{};
  info.nbf = nbf;
  info.status = 0;
  info.center_r = center_row;
  info.center_c = center_col;
  info.radius2 = 0.0f;
  info.dc_re = dc_re;
  info.dc_im = dc_im;
  HIP_TRY(hipMemcpyAsync(p->bf_rows, rows_host, (size_t)nbf * 4, hipMemcpyHostToDevice, s), "set_bf");
  HIP_TRY(hipMemcpyAsync(p->bf_cols, cols_host, (size_t)nbf * 4, hipMemcpyHostToDevice, s), "set_bf");
  HIP_TRY(hipMemcpyAsync(p->info, &info, sizeof(BfInfo), hipMemcpyHostToDevice, s), "set_bf");
  HIP_TRY(hipStreamSynchronize(s), "set_bf sync");
  p->nbf_hint = nbf;
  p->geometry_valid = false;
  p->gcols = 0;
  return QSSB_OK;
}

int ssb_get_bf(ssb_plan* p, int32_t* rows_host, int32_t* cols_host, int capacity, int* nbf_host,
               float* info6_host, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  BfInfo info{};
  HIP_TRY(hipMemcpyAsync(&info, p->info, sizeof(BfInfo), hipMemcpyDeviceToHost, s), "get_bf");
  HIP_TRY(hipStreamSynchronize(s), "get_bf sync");
  if (nbf_host) *nbf_host = info.nbf;
  if (info6_host) {
    info6_host[0] = info.center_r;
    info6_host[1] = info.center_c;
    info6_host[2] = info.radius2;
    info6_host[3] = info.dc_re;
    info6_host[4] = info.dc_im;
    info6_host[5] = (float)info.status;
  }
  if (info.nbf > 0) p->nbf_hint = info.nbf;
  const int n = info.nbf < capacity ? info.nbf : capacity;
  if (rows_host && n > 0)
    HIP_TRY(hipMemcpy(rows_host, p->bf_rows, (size_t)n * 4, hipMemcpyDeviceToHost), "get_bf rows");
  if (cols_host && n > 0)
    HIP_TRY(hipMemcpy(cols_host, p->bf_cols, (size_t)n * 4, hipMemcpyDeviceToHost), "get_bf cols");
  return QSSB_OK;
}

int ssb_cache_rotation(ssb_plan* p, double rotation_rad, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  p->rotation = rotation_rad;
  return ensure_geometry(p, as_stream(stream));
}

int ssb_get_geometry(ssb_plan* p, float* kx_out, float* ky_out, float* ak_out, int count, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (count < 0 || count > p->max_bf) return fail(QSSB_ERR_ARG, "count exceeds the plan capacity");
  hipStream_t s = as_stream(stream);
  if (!p->geometry_valid) {
    rc = ensure_geometry(p, s);
    if (rc) return rc;
  }
  const size_t bytes = (size_t)count * 4;
  if (kx_out) HIP_TRY(hipMemcpyAsync(kx_out, p->kx, bytes, hipMemcpyDeviceToDevice, s), "get_geometry");
  if (ky_out) HIP_TRY(hipMemcpyAsync(ky_out, p->ky, bytes, hipMemcpyDeviceToDevice, s), "get_geometry");
  if (ak_out) HIP_TRY(hipMemcpyAsync(ak_out, p->ak, bytes, hipMemcpyDeviceToDevice, s), "get_geometry");
  return QSSB_OK;
}

// Row-transformed stack H[b][R_r][slot] for `cap` BF pixels (grown on demand).
static int alloc_h(ssb_plan* p, int cap) {
  if (p->H_cap >= cap) return QSSB_OK;
  dfree(p->H);
  p->H = nullptr;
  p->H_cap = 0;
  hipError_t e = dmalloc(&p->H, (size_t)cap * p->n * (p->n / 2));
  if (e != hipSuccess) return hip_fail(e, "row-transformed stack allocation");
  p->H_cap = cap;
  return QSSB_OK;
}

static int alloc_gt(ssb_plan* p, int gcols, int nbf) {
  const size_t need = (size_t)nbf * gcols * p->n * sizeof(float2);
  if (p->Gt_bytes < need) {
    dfree(p->Gt);
    p->Gt = nullptr;
    p->Gt_bytes = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p->Gt), need);
    if (e != hipSuccess) return hip_fail(e, "resident G allocation");
    p->Gt_bytes = need;
  }
  return QSSB_OK;
}

static int prepare_g_typed(ssb_plan* p, const void* data, int dtype, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!data) return fail(QSSB_ERR_ARG, "data is NULL");
  if (!dtype_size(dtype)) return fail(QSSB_ERR_ARG, "unknown sample type " + std::to_string(dtype));
  hipStream_t s = as_stream(stream);
  {
    // the resident stack is sized by the current selection: read it back
    BfInfo info{};
    HIP_TRY(hipMemcpyAsync(&info, p->info, sizeof(BfInfo), hipMemcpyDeviceToHost, s), "prepare_g info");
    HIP_TRY(hipStreamSynchronize(s), "prepare_g sync");
    if (info.status != 0 || info.nbf < 1) return fail(QSSB_ERR_STATE, "no valid bright-field selection");
    p->nbf_hint = info.nbf;
  }
  rc = alloc_gt(p, p->n / 2 + 1, p->nbf_hint);
  if (rc) return rc;
  rc = alloc_h(p, p->nbf_hint);
  if (rc) return rc;
  HIP_TRY(launch_gather_rowfft(p->n, dtype, data, p->D, p->det_cols, p->bf_rows, p->bf_cols, p->info, p->nbf_hint,
                               p->tw, p->tune_gather_pairs, p->gain, p->H, s),
          "gather_rowfft");
  HIP_TRY(launch_trotter_slots(p->n, true, p->H, p->info, geom_ptrs(p), p->rowph, p->colph, p->echi, p->optics, p->tw,
                               p->S_part, p->Gt, false, p->tune_trotter, s),
          "colfft");
  p->gcols = p->n / 2 + 1;
  p->g_nbf = p->nbf_hint;
  return QSSB_OK;
}

int ssb_prepare_g_u16(ssb_plan* p, const uint16_t* data, void* stream) {
  return prepare_g_typed(p, data, QSSB_DTYPE_U16, stream);
}

int ssb_prepare_g_typed(ssb_plan* p, const void* data, int dtype, void* stream) {
  return prepare_g_typed(p, data, dtype, stream);
}

int ssb_export_g(ssb_plan* p, void* g_out, int nbf, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->gcols) return fail(QSSB_ERR_STATE, "no resident Fourier stack; call ssb_prepare_g_u16 first");
  if (!g_out) return fail(QSSB_ERR_ARG, "output is NULL");
  if (nbf != p->g_nbf)
    return fail(QSSB_ERR_ARG, "output holds " + std::to_string(nbf) + " BF planes; the resident stack has " +
                                  std::to_string(p->g_nbf));
  const int oc = p->n / 2 + 1;
  const long long total = (long long)p->g_nbf * p->n * oc;
  hipLaunchKernelGGL(k_export_g, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, as_stream(stream), p->Gt,
                     p->gcols, p->n, p->g_nbf, reinterpret_cast<float2*>(g_out), oc);
  HIP_TRY(hipGetLastError(), "export_g");
  return QSSB_OK;
}

int ssb_import_g(ssb_plan* p, const void* g_in, int nbf, int gcols, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!g_in) return fail(QSSB_ERR_ARG, "input is NULL");
  if (gcols != p->n && gcols != p->n / 2 + 1)
    return fail(QSSB_ERR_ARG, "G must have N or N/2+1 columns");
  if (nbf != p->nbf_hint) return fail(QSSB_ERR_ARG, "G first dimension must match the BF selection");
  rc = alloc_gt(p, gcols, nbf);
  if (rc) return rc;
  const long long total = (long long)nbf * p->n * gcols;
  hipLaunchKernelGGL(k_import_g, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, as_stream(stream),
                     reinterpret_cast<const float2*>(g_in), gcols, p->n, nbf, p->Gt);
  HIP_TRY(hipGetLastError(), "import_g");
  p->gcols = gcols;
  p->g_nbf = nbf;
  return QSSB_OK;
}

int ssb_object(ssb_plan* p, double c10, double c12, double phi12, void* obj_out, float* phase_out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->gcols) return fail(QSSB_ERR_STATE, "no resident Fourier stack");
  hipStream_t s = as_stream(stream);
  if (!p->geometry_valid) {
    rc = ensure_geometry(p, s);
    if (rc) return rc;
  }
  const QuadChi m = quad_chi(p, c10, c12, phi12);
  HIP_TRY(launch_phasors(p->n, p->info, p->nbf_hint, geom_ptrs(p), m, p->rowph, p->colph, p->echi, s), "phasors");
  int groups = object_groups(p->n);
  if (p->gcols == p->n / 2 + 1) {
    HIP_TRY(launch_trotter_slots(p->n, false, p->Gt, p->info, geom_ptrs(p), p->rowph, p->colph, p->echi, p->optics, p->tw,
                                 p->S_part, nullptr, true, 0, s, &groups),
            "trotter");
  } else {
    HIP_TRY(launch_trotter_full(p->n, p->Gt, p->info, geom_ptrs(p), p->rowph, m, p->optics, p->S_part, s),
            "trotter_full");
  }
  HIP_TRY(launch_finish_object(p->n, p->S_part, groups, p->info, p->tw, p->S,
                               reinterpret_cast<float2*>(obj_out), phase_out, s),
          "finish_object");
  return QSSB_OK;
}

static int alloc_loss(ssb_plan* p) {
  if (p->T) return QSSB_OK;
  const size_t nn = (size_t)p->n * p->n;
  hipError_t e = dmalloc(&p->T, (size_t)loss_chunk_bf(p->n) * nn);
  if (e == hipSuccess) e = dmalloc(&p->M1, (size_t)loss_splits(p->n) * nn);
  if (e == hipSuccess) e = dmalloc(&p->M2, (size_t)loss_splits(p->n) * nn);
  if (e == hipSuccess) e = dmalloc(&p->var_part, (nn + 255) / 256);
  if (e == hipSuccess) e = dmalloc(&p->loss, 1);
  if (e != hipSuccess) {
    dfree(p->T);
    dfree(p->M1);
    dfree(p->M2);
    dfree(p->var_part);
    dfree(p->loss);
    p->T = p->M1 = p->M2 = nullptr;
    p->var_part = p->loss = nullptr;
    return hip_fail(e, "exact phase workspace allocation");
  }
  return QSSB_OK;
}

int ssb_phase_loss(ssb_plan* p, double c10, double c12, double phi12, float* phase_out, float* loss_out,
                   void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!p->gcols) return fail(QSSB_ERR_STATE, "no resident Fourier stack");
  hipStream_t s = as_stream(stream);
  if (!p->geometry_valid) {
    rc = ensure_geometry(p, s);
    if (rc) return rc;
  }
  rc = alloc_loss(p);
  if (rc) return rc;
  const QuadChi m = quad_chi(p, c10, c12, phi12);
  HIP_TRY(launch_phasors(p->n, p->info, p->nbf_hint, geom_ptrs(p), m, p->rowph, p->colph, p->echi, s), "phasors");
  HIP_TRY(launch_phase_moments(p->n, p->Gt, p->gcols, p->info, p->g_nbf, geom_ptrs(p), p->rowph, p->colph, p->echi,
                               nullptr, p->optics, p->tw, p->T, p->M1, p->M2, p->var_part, phase_out,
                               loss_out ? loss_out : p->loss, s),
          "phase_moments");
  return QSSB_OK;
}

int ssb_detect_bf_radius(ssb_plan* p, double threshold_ratio, int* center_row, int* center_col, int* radius,
                         void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!center_row || !center_col || !radius) return fail(QSSB_ERR_ARG, "output pointers must not be NULL");
  if (p->det_rows < 1 || p->det_cols < 1 || p->det_rows > 256 || p->det_cols > 256)
    return fail(QSSB_ERR_UNSUPPORTED, "probe detection supports detectors up to 256 x 256");
  hipStream_t s = as_stream(stream);
  int* dout = nullptr;
  HIP_TRY(dmalloc(&dout, 4), "detect_bf_radius allocation");
  int host[4] = {0, 0, 0, 0};
  hipError_t e = launch_detect_bf_radius(p->dp_mean, p->det_rows, p->det_cols, to_df(threshold_ratio), dout, s);
  if (e == hipSuccess) e = hipMemcpyAsync(host, dout, sizeof(host), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(dout);
  if (e != hipSuccess) return hip_fail(e, "detect_bf_radius");
  if (host[0] == 1)
    return fail(QSSB_ERR_ARG, "Diffraction pattern has no positive finite values - cannot detect BF disk. "
                              "Check that your data is loaded correctly.");
  if (host[0] == 2)
    return fail(QSSB_ERR_ARG, "No pixels above threshold. The diffraction pattern may be too noisy or empty.");
  *center_row = host[1];
  *center_col = host[2];
  *radius = host[3];
  return QSSB_OK;
}

int ssb_max_u16(const uint16_t* data, long long count, unsigned int* max_out_host, void* stream) {
  g_last_error.clear();
  if (!data || !max_out_host) return fail(QSSB_ERR_ARG, "pointers must not be NULL");
  if (count < 1) return fail(QSSB_ERR_ARG, "count must be positive");
  hipStream_t s = as_stream(stream);
  unsigned int* buf = nullptr;
  HIP_TRY(dmalloc(&buf, (size_t)max_u16_parts() + 1), "max allocation");
  hipError_t e = launch_max_u16(data, count, buf, buf + max_u16_parts(), s);
  if (e == hipSuccess) e = hipMemcpyAsync(max_out_host, buf + max_u16_parts(), sizeof(unsigned int),
                                          hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(buf);
  if (e != hipSuccess) return hip_fail(e, "max_u16");
  return QSSB_OK;
}

static int bf_columns(ssb_plan* p, const void* data, bool in_u8, int bf_offset, int count, int narrow_u8, void* out,
                      void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!data || !out) return fail(QSSB_ERR_ARG, "data and output must not be NULL");
  if (p->nbf_hint < 1) return fail(QSSB_ERR_STATE, "no bright-field selection");
  if (bf_offset < 0 || count < 1 || bf_offset + count > p->nbf_hint)
    return fail(QSSB_ERR_ARG, "BF range [" + std::to_string(bf_offset) + ", " + std::to_string(bf_offset + count) +
                                  ") is outside the selection of " + std::to_string(p->nbf_hint));
  HIP_TRY(launch_bf_columns(data, in_u8, (long long)p->n * p->n, p->D, p->det_cols, p->bf_rows, p->bf_cols,
                            bf_offset, count, narrow_u8 != 0, out, as_stream(stream)),
          "bf_columns");
  return QSSB_OK;
}

int ssb_bf_columns_u16(ssb_plan* p, const uint16_t* data, int bf_offset, int count, int narrow_u8, void* out,
                       void* stream) {
  return bf_columns(p, data, false, bf_offset, count, narrow_u8, out, stream);
}

int ssb_bf_columns_u8(ssb_plan* p, const uint8_t* data, int bf_offset, int count, void* out, void* stream) {
  return bf_columns(p, data, true, bf_offset, count, 1, out, stream);
}

int ssb_learn_features(ssb_plan* p, const int* qrow_host, const int* qcol_host, float* out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!qrow_host || !qcol_host || !out) return fail(QSSB_ERR_ARG, "pointers must not be NULL");
  if (p->gcols != p->n / 2 + 1) return fail(QSSB_ERR_STATE, "features read the resident Hermitian stack");
  for (int i = 0; i < 8; ++i)
    if (qrow_host[i] < 0 || qrow_host[i] >= p->n || qcol_host[i] < 0 || qcol_host[i] >= p->n)
      return fail(QSSB_ERR_ARG, "scan-frequency indices out of range");
  hipStream_t s = as_stream(stream);
  int* dq = nullptr;
  HIP_TRY(dmalloc(&dq, 16), "features allocation");
  hipError_t e = hipMemcpyAsync(dq, qrow_host, 8 * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dq + 8, qcol_host, 8 * sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_features_from_g(p->Gt, p->n, p->bf_rows, p->bf_cols, p->info, dq, dq + 8, out, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(dq);
  if (e != hipSuccess) return hip_fail(e, "learn_features");
  return QSSB_OK;
}

int ssb_scan_target(int H, int W) {
  if (H < 1 || W < 1) return 0;
  const int longest = H > W ? H : W;
  return longest <= 128 ? 128 : (longest <= 256 ? 256 : (longest <= 512 ? 512 : 1024));
}

static int normalize_scan_typed(const void* src, int dtype, int H, int W, int det_rows, int det_cols, void* dst,
                                int T, void* stream) {
  g_last_error.clear();
  if (!src || !dst) return fail(QSSB_ERR_ARG, "source and destination must not be NULL");
  if (!dtype_size(dtype)) return fail(QSSB_ERR_ARG, "unknown sample type " + std::to_string(dtype));
  if (H < 1 || W < 1 || det_rows < 1 || det_cols < 1) return fail(QSSB_ERR_ARG, "shapes must be positive");
  if (T != ssb_scan_target(H, W))
    return fail(QSSB_ERR_ARG, "target scan size " + std::to_string(T) + " is not the supported size for a " +
                                  std::to_string(H) + " x " + std::to_string(W) + " scan (" +
                                  std::to_string(ssb_scan_target(H, W)) + ")");
  const long long D = (long long)det_rows * det_cols;
  if (D > (1 << 24)) return fail(QSSB_ERR_UNSUPPORTED, "detector too large");
  hipStream_t s = as_stream(stream);
  size_t mean_off = 0;
  const size_t bytes = normalize_workspace_bytes(dtype, H, T, (int)D, &mean_off);
  uint8_t* ws = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&ws), bytes);
  if (e == hipSuccess) e = launch_normalize_scan_typed(dtype, src, H, W, (int)D, dst, T, ws, ws + mean_off, s);
  hipError_t e2 = hipStreamSynchronize(s);
  dfree(ws);
  if (e != hipSuccess) return hip_fail(e, "normalize_scan");
  if (e2 != hipSuccess) return hip_fail(e2, "normalize_scan sync");
  return QSSB_OK;
}

int ssb_normalize_scan_u16(const uint16_t* src, int H, int W, int det_rows, int det_cols, uint16_t* dst, int T,
                           void* stream) {
  return normalize_scan_typed(src, QSSB_DTYPE_U16, H, W, det_rows, det_cols, dst, T, stream);
}

int ssb_normalize_scan_typed(const void* src, int dtype, int H, int W, int det_rows, int det_cols, void* dst, int T,
                             void* stream) {
  return normalize_scan_typed(src, dtype, H, W, det_rows, det_cols, dst, T, stream);
}


static ChiPoly chi_poly(const ssb_plan* p, const double* mags, const double* angles) {
  static const int kn[14] = {1, 1, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5};
  static const int km[14] = {0, 2, 1, 3, 0, 2, 4, 1, 3, 5, 0, 2, 4, 6};
  ChiPoly cp{};
  const double lam = p->wavelength;
  for (int i = 0; i < 14; ++i) {
    const int n = kn[i], m = km[i];
    cp.K[i] = (float)(2.0 * M_PI / lam * std::pow(lam, n + 1) / (n + 1) * mags[i]);
    cp.ca[i] = (float)std::cos(m * angles[i]);
    cp.sa[i] = (float)std::sin(m * angles[i]);
    if ((m & 1) && cp.K[i] != 0.0f) cp.odd = 1;
  }
  return cp;
}

int ssb_phase_loss_full(ssb_plan* p, const double* mags_host, const double* angles_host, float* phase_out,
                        float* loss_out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!mags_host || !angles_host) return fail(QSSB_ERR_ARG, "aberration vectors must not be NULL");
  if (!p->gcols) return fail(QSSB_ERR_STATE, "no resident Fourier stack");
  hipStream_t s = as_stream(stream);
  if (!p->geometry_valid) {
    rc = ensure_geometry(p, s);
    if (rc) return rc;
  }
  rc = alloc_loss(p);
  if (rc) return rc;
  const ChiPoly cp = chi_poly(p, mags_host, angles_host);
  HIP_TRY(launch_phase_moments(p->n, p->Gt, p->gcols, p->info, p->g_nbf, geom_ptrs(p), p->rowph, p->colph, p->echi,
                               &cp, p->optics, p->tw, p->T, p->M1, p->M2, p->var_part, phase_out,
                               loss_out ? loss_out : p->loss, s),
          "phase_moments_full");
  return QSSB_OK;
}

int ssb_batch_phase_loss(ssb_plan* p, int n_trials, const double* mags_host, const double* angles_host,
                         const double* rotation_host, float* losses_out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (n_trials < 0) return fail(QSSB_ERR_ARG, "n_trials must be non-negative");
  if (n_trials == 0) return QSSB_OK;
  if (!mags_host || !angles_host || !losses_out) return fail(QSSB_ERR_ARG, "trial arrays must not be NULL");
  if (!p->gcols) return fail(QSSB_ERR_STATE, "no resident Fourier stack");
  hipStream_t s = as_stream(stream);
  rc = alloc_loss(p);
  if (rc) return rc;
  const double rot0 = p->rotation;
  bool have_geometry = p->geometry_valid;
  for (int t = 0; t < n_trials; ++t) {
    const double* m = mags_host + (size_t)t * 14;
    const double* a = angles_host + (size_t)t * 14;
    if (rotation_host && (!have_geometry || rotation_host[t] != p->rotation)) {
      p->rotation = rotation_host[t];
      rc = ensure_geometry(p, s);
      if (rc) return rc;
      have_geometry = true;
    } else if (!have_geometry) {
      rc = ensure_geometry(p, s);
      if (rc) return rc;
      have_geometry = true;
    }
    bool higher = false;
    for (int i = 2; i < 14; ++i) higher = higher || (m[i] != 0.0);
    float* lt = losses_out + t;
    if (higher) {
      const ChiPoly cp = chi_poly(p, m, a);
      HIP_TRY(launch_phase_moments(p->n, p->Gt, p->gcols, p->info, p->g_nbf, geom_ptrs(p), p->rowph, p->colph,
                                   p->echi, &cp, p->optics, p->tw, p->T, p->M1, p->M2, p->var_part, nullptr, lt, s),
              "batch phase_moments_full");
    } else {
      const QuadChi qc = quad_chi(p, m[0], m[1], a[1]);
      HIP_TRY(launch_phasors(p->n, p->info, p->nbf_hint, geom_ptrs(p), qc, p->rowph, p->colph, p->echi, s),
              "batch phasors");
      HIP_TRY(launch_phase_moments(p->n, p->Gt, p->gcols, p->info, p->g_nbf, geom_ptrs(p), p->rowph, p->colph,
                                   p->echi, nullptr, p->optics, p->tw, p->T, p->M1, p->M2, p->var_part, nullptr, lt,
                                   s),
              "batch phase_moments");
    }
  }
  // leave the cached rotation as the caller set it
  if (rotation_host && p->rotation != rot0) {
    p->rotation = rot0;
    rc = ensure_geometry(p, s);
    if (rc) return rc;
  }
  return QSSB_OK;
}

// Enqueue the captured form of a batch: per trial the geometry (every
// trial, from the device table), then the phasors and the exact loss of
// ssb_batch_phase_loss reading their coefficients from the table.
static hipError_t enqueue_exact_batch(ssb_plan* p, int n, float* losses_out, hipStream_t s) {
  hipError_t e = hipSuccess;
  for (int t = 0; t < n && e == hipSuccess; ++t) {
    const ExactTrial& h = p->xt_host[t];
    const ExactTrial* d = p->xt + t;
    e = launch_bf_geometry(p->bf_rows, p->bf_cols, p->info, p->max_bf, to_df(p->recip_y), to_df(p->recip_x),
                           h.cs[0], h.cs[1], p->optics, p->kx, p->kxl, p->ky, p->kyl, p->ak, s, d->cs);
    if (e != hipSuccess) break;
    if (h.higher) {
      e = launch_phase_moments(p->n, p->Gt, p->gcols, p->info, p->g_nbf, geom_ptrs(p), p->rowph, p->colph, p->echi,
                               &h.cp, p->optics, p->tw, p->T, p->M1, p->M2, p->var_part, nullptr, losses_out + t, s,
                               &d->cp);
    } else {
      e = launch_phasors(p->n, p->info, p->nbf_hint, geom_ptrs(p), h.m, p->rowph, p->colph, p->echi, s, &d->m);
      if (e == hipSuccess)
        e = launch_phase_moments(p->n, p->Gt, p->gcols, p->info, p->g_nbf, geom_ptrs(p), p->rowph, p->colph,
                                 p->echi, nullptr, p->optics, p->tw, p->T, p->M1, p->M2, p->var_part, nullptr,
                                 losses_out + t, s);
    }
  }
  return e;
}

int ssb_batch_phase_loss_graph(ssb_plan* p, int n_trials, const double* mags_host, const double* angles_host,
                               const double* rotation_host, float* losses_out, void* stream) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (n_trials < 1) return fail(QSSB_ERR_ARG, "n_trials must be positive");
  if (!mags_host || !angles_host || !losses_out) return fail(QSSB_ERR_ARG, "trial arrays must not be NULL");
  if (!p->gcols) return fail(QSSB_ERR_STATE, "no resident Fourier stack");
  hipStream_t s = as_stream(stream);
  rc = alloc_loss(p);
  if (rc) return rc;
  if (p->xt_cap < n_trials) {
    HIP_TRY(hipStreamSynchronize(s), "batch graph sync");
    if (p->gstream) HIP_TRY(hipStreamSynchronize(p->gstream), "batch graph sync");
    dfree(p->xt);
    p->xt = nullptr;
    p->xt_cap = 0;
    HIP_TRY(dmalloc(&p->xt, (size_t)n_trials), "batch table allocation");
    p->xt_cap = n_trials;
    if (p->xexec) (void)hipGraphExecDestroy(p->xexec);
    if (p->xgraph) (void)hipGraphDestroy(p->xgraph);
    p->xexec = nullptr;
    p->xgraph = nullptr;
  }
  if (!p->gstream) {
    HIP_TRY(hipStreamCreateWithFlags(&p->gstream, hipStreamNonBlocking), "batch graph stream");
    HIP_TRY(hipEventCreateWithFlags(&p->gev_in, hipEventDisableTiming), "batch graph event");
    HIP_TRY(hipEventCreateWithFlags(&p->gev_out, hipEventDisableTiming), "batch graph event");
  }
  // the previous replay may still read the device table: order the upload after it
  HIP_TRY(hipStreamWaitEvent(s, p->gev_out, 0), "batch graph order");
  std::vector<int> pattern(n_trials);
  p->xt_host.assign(n_trials, ExactTrial{});
  for (int t = 0; t < n_trials; ++t) {
    const double* m = mags_host + (size_t)t * 14;
    const double* a = angles_host + (size_t)t * 14;
    ExactTrial& x = p->xt_host[t];
    const double rot = rotation_host ? rotation_host[t] : p->rotation;
    x.cs[0] = to_df(std::cos(-rot));
    x.cs[1] = to_df(std::sin(-rot));
    bool higher = false;
    for (int i = 2; i < 14; ++i) higher = higher || (m[i] != 0.0);
    x.higher = higher ? 1 : 0;
    x.m = quad_chi(p, m[0], m[1], a[1]);
    x.cp = chi_poly(p, m, a);
    pattern[t] = x.higher;
  }
  HIP_TRY(hipMemcpyAsync(p->xt, p->xt_host.data(), sizeof(ExactTrial) * n_trials, hipMemcpyHostToDevice, s),
          "batch table upload");
  HIP_TRY(hipStreamSynchronize(s), "batch table upload");   // the host table may change after return
  if (!p->xexec || pattern != p->xt_pattern || p->xgraph_out != losses_out) {
    if (p->xexec) (void)hipGraphExecDestroy(p->xexec);
    if (p->xgraph) (void)hipGraphDestroy(p->xgraph);
    p->xexec = nullptr;
    p->xgraph = nullptr;
    HIP_TRY(hipStreamBeginCapture(p->gstream, hipStreamCaptureModeThreadLocal), "batch capture");
    const hipError_t ee = enqueue_exact_batch(p, n_trials, losses_out, p->gstream);
    const hipError_t ec = hipStreamEndCapture(p->gstream, &p->xgraph);
    if (ee != hipSuccess) return hip_fail(ee, "batch capture enqueue");
    if (ec != hipSuccess) return hip_fail(ec, "batch end capture");
    HIP_TRY(hipGraphInstantiate(&p->xexec, p->xgraph, nullptr, nullptr, 0), "batch graph instantiate");
    p->xt_pattern = pattern;
    p->xgraph_out = losses_out;
  }
  HIP_TRY(hipEventRecord(p->gev_in, s), "batch graph launch");
  HIP_TRY(hipStreamWaitEvent(p->gstream, p->gev_in, 0), "batch graph launch");
  HIP_TRY(hipGraphLaunch(p->xexec, p->gstream), "batch graph launch");
  HIP_TRY(hipEventRecord(p->gev_out, p->gstream), "batch graph launch");
  HIP_TRY(hipStreamWaitEvent(s, p->gev_out, 0), "batch graph launch");
  // the batch leaves the last trial's geometry: restore the cached rotation's
  return ensure_geometry(p, s);
}

// Superset sample sums, box and the compact per-frame copy (allocated on first use).
static int alloc_superset(ssb_plan* p) {
  if (p->sup_C) return QSSB_OK;
  const size_t frames = (size_t)p->n * p->n;
  hipError_t e = dmalloc(&p->sup_part, (size_t)superset_part_count() * p->D);
  if (e == hipSuccess) e = dmalloc(&p->sup_box, 1);
  if (e == hipSuccess) e = dmalloc(&p->sup_C, frames * kSupMaxPixels);
  if (e != hipSuccess) {
    dfree(p->sup_part);
    dfree(p->sup_box);
    dfree(p->sup_C);
    p->sup_part = nullptr;
    p->sup_box = nullptr;
    p->sup_C = nullptr;
    return hip_fail(e, "superset allocation");
  }
  return QSSB_OK;
}

static int recon_typed(ssb_plan* p, const ssb_plan* after, const void* data, int dtype, double c10, double c12,
                       double phi12, double rotation_rad, float threshold, float bf_radius, void* obj_out,
                       float* phase_out, hipStream_t s) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!data) return fail(QSSB_ERR_ARG, "data is NULL");
  if (!dtype_size(dtype)) return fail(QSSB_ERR_ARG, "unknown sample type " + std::to_string(dtype));
  void* partial = nullptr;
  rc = dp_partial_for(p, dtype, &partial);
  if (rc) return rc;
  if (after == p) return fail(QSSB_ERR_ARG, "a plan cannot be ordered after itself");
  if (after) {
    rc = check_plan(after);
    if (rc) return rc;
    if (!p->ev_stream) HIP_TRY(hipEventCreateWithFlags(&p->ev_stream, hipEventDisableTiming), "stream event");
    if (after->ev_stream_recorded) HIP_TRY(hipStreamWaitEvent(s, after->ev_stream, 0), "wait stream pass");
  }
  hipEvent_t* ev = nullptr;
  if (p->timing_max > 0 && p->timing_calls < p->timing_max) {
    ev = &p->events[(size_t)p->timing_calls * (NSTAGES + 1)];
    p->timing_calls += 1;
  }
  auto mark = [&](int i) {
    if (ev && ((p->timing_mask >> i) & 1)) (void)hipEventRecord(ev[i], s);
  };
  const long long frames = (long long)p->n * p->n;
  // bright-field superset (uint16 frames of whole 16-byte detector-row groups):
  // sampled box, copied compactly by the stream pass for the gather
  const bool sup = dtype == QSSB_DTYPE_U16 && p->tune_superset == 1 && p->det_cols % 8 == 0 &&
                   (reinterpret_cast<uintptr_t>(data) & 15u) == 0;
  if (sup) {
    rc = alloc_superset(p);
    if (rc) return rc;
  }
  mark(0);
  if (sup)
    HIP_TRY(launch_superset(reinterpret_cast<const uint16_t*>(data), frames, p->det_rows, p->det_cols,
                            p->sup_part, p->sup_box, s),
            "superset");
  mark(1);
  if (sup)
    HIP_TRY(launch_mean_dp_stream_sup(reinterpret_cast<const uint16_t*>(data), frames, p->D, p->det_cols,
                                      reinterpret_cast<uint32_t*>(partial), p->max_ranges, p->tune_stream_wg,
                                      p->sup_box, p->sup_C, s),
            "mean_dp");
  else
    HIP_TRY(launch_mean_dp_stream(dtype, data, frames, p->D, partial, p->max_ranges, p->tune_stream_wg, s),
            "mean_dp");
  if (p->ev_stream) {
    HIP_TRY(hipEventRecord(p->ev_stream, s), "stream pass event");
    p->ev_stream_recorded = true;
  }
  mark(2);
  HIP_TRY(launch_mean_dp_reduce(dtype, data, frames, p->D, partial, p->max_ranges, p->tune_stream_wg, p->gain,
                                p->dp_sums, p->dp_mean, p->dp_dcpix, s),
          "mean_dp");
  p->dp_dtype = dtype;
  mark(3);
  HIP_TRY(launch_select_bf(p->dp_mean, p->dp_sums, p->dp_dcpix, exact_dc(p), p->det_rows, p->det_cols, threshold,
                           bf_radius, p->max_bf, p->bf_rows, p->bf_cols, p->info, s, sup ? p->sup_box : nullptr),
          "select_bf");
  mark(4);
  p->rotation = rotation_rad;
  rc = ensure_geometry(p, s);
  if (rc) return rc;
  p->gcols = 0;
  rc = alloc_h(p, p->max_bf);
  if (rc) return rc;
  mark(5);
  const QuadChi m = quad_chi(p, c10, c12, phi12);
  int groups = object_groups(p->n);
  // With QSSB_TUNE_SPLIT 1: two halves of the BF list (raster order, so
  // disjoint detector rows), the second half's gather (HBM-bound) on the
  // plan's second stream beside the first half's trotter (VALU-bound).
  // Measured at 256x256x128x128 the two compete for the CUs (the trotter
  // holds one 8-wave workgroup per CU at 225 VGPRs) and the pair is slower
  // than one pass, so one pass is the default.  Stage marks with the split:
  // 5 -> 6 the first half's gather and the phasors, 6 -> 7 both trotters
  // and the second gather.
  const bool two = p->tune_split == 1 &&
                   trotter_wave_path(p->n, true, p->optics, true, false, p->tune_trotter);
  if (two) {
    if (!p->pstream) HIP_TRY(hipStreamCreateWithFlags(&p->pstream, hipStreamNonBlocking), "second stream");
    if (!p->pev_a) HIP_TRY(hipEventCreateWithFlags(&p->pev_a, hipEventDisableTiming), "split event");
    if (!p->pev_b) HIP_TRY(hipEventCreateWithFlags(&p->pev_b, hipEventDisableTiming), "split event");
    HIP_TRY(launch_gather_rowfft(p->n, dtype, data, p->D, p->det_cols, p->bf_rows, p->bf_cols, p->info,
                                 p->nbf_hint, p->tw, p->tune_gather_pairs, p->gain, p->H, s,
                                 sup ? p->sup_box : nullptr, sup ? p->sup_C : nullptr, 0, 2),
            "gather_rowfft");
    HIP_TRY(hipEventRecord(p->pev_a, s), "split event");
    HIP_TRY(hipStreamWaitEvent(p->pstream, p->pev_a, 0), "split wait");
    HIP_TRY(launch_gather_rowfft(p->n, dtype, data, p->D, p->det_cols, p->bf_rows, p->bf_cols, p->info,
                                 p->nbf_hint, p->tw, p->tune_gather_pairs, p->gain, p->H, p->pstream,
                                 sup ? p->sup_box : nullptr, sup ? p->sup_C : nullptr, 1, 2),
            "gather_rowfft");
    HIP_TRY(hipEventRecord(p->pev_b, p->pstream), "split event");
    HIP_TRY(launch_phasors(p->n, p->info, p->nbf_hint, geom_ptrs(p), m, p->rowph, p->colph, p->echi, s),
            "phasors");
    mark(6);
    HIP_TRY(launch_trotter_slots(p->n, true, p->H, p->info, geom_ptrs(p), p->rowph, p->colph, p->echi, p->optics,
                                 p->tw, p->S_part, nullptr, true, p->tune_trotter, s, &groups, 0, 2),
            "colfft_trotter");
    HIP_TRY(hipStreamWaitEvent(s, p->pev_b, 0), "split wait");
    HIP_TRY(launch_trotter_slots(p->n, true, p->H, p->info, geom_ptrs(p), p->rowph, p->colph, p->echi, p->optics,
                                 p->tw, p->S_part, nullptr, true, p->tune_trotter, s, &groups, 1, 2),
            "colfft_trotter");
  } else {
    // The gather's threads also write the phasor tables (one launch fewer;
    // geometry and phasors on a second stream beside the gather measured
    // slower: the cross-stream event pair costs more than they would hide).
    const PhasorJob job{p->info, geom_ptrs(p), m, p->n, p->rowph, p->colph, p->echi};
    HIP_TRY(launch_gather_rowfft(p->n, dtype, data, p->D, p->det_cols, p->bf_rows, p->bf_cols, p->info,
                                 p->nbf_hint, p->tw, p->tune_gather_pairs, p->gain, p->H, s,
                                 sup ? p->sup_box : nullptr, sup ? p->sup_C : nullptr, 0, 1, &job),
            "gather_rowfft");
    mark(6);
    HIP_TRY(launch_trotter_slots(p->n, true, p->H, p->info, geom_ptrs(p), p->rowph, p->colph, p->echi, p->optics,
                                 p->tw, p->S_part, nullptr, true, p->tune_trotter, s, &groups),
            "colfft_trotter");
  }
  mark(7);
  HIP_TRY(launch_finish_object(p->n, p->S_part, groups, p->info, p->tw, p->S,
                               reinterpret_cast<float2*>(obj_out), phase_out, s),
          "finish_object");
  mark(8);
  return QSSB_OK;
}

int ssb_recon_u16(ssb_plan* p, const uint16_t* data, double c10, double c12, double phi12, double rotation_rad,
                  float threshold, float bf_radius, void* obj_out, float* phase_out, void* stream) {
  return recon_typed(p, nullptr, data, QSSB_DTYPE_U16, c10, c12, phi12, rotation_rad, threshold, bf_radius, obj_out,
                     phase_out, as_stream(stream));
}

int ssb_recon_typed(ssb_plan* p, const void* data, int dtype, double c10, double c12, double phi12,
                    double rotation_rad, float threshold, float bf_radius, void* obj_out, float* phase_out,
                    void* stream) {
  return recon_typed(p, nullptr, data, dtype, c10, c12, phi12, rotation_rad, threshold, bf_radius, obj_out,
                     phase_out, as_stream(stream));
}

int ssb_recon_u16_after(ssb_plan* p, const ssb_plan* after, const uint16_t* data, double c10, double c12,
                        double phi12, double rotation_rad, float threshold, float bf_radius, void* obj_out,
                        float* phase_out, void* stream) {
  if (!after) return fail(QSSB_ERR_ARG, "after is NULL");
  return recon_typed(p, after, data, QSSB_DTYPE_U16, c10, c12, phi12, rotation_rad, threshold, bf_radius, obj_out,
                     phase_out, as_stream(stream));
}

int ssb_stage_count(void) { return NSTAGES; }

const char* ssb_stage_name(int stage) {
  if (stage < 0 || stage >= NSTAGES) return "";
  return kStageNames[stage];
}

static void release_events(ssb_plan* p) {
  for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
  p->events.clear();
  p->timing_max = 0;
  p->timing_calls = 0;
}

int ssb_timing_begin(ssb_plan* p, int max_calls, int boundary_mask) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (max_calls < 1 || max_calls > 100000) return fail(QSSB_ERR_ARG, "max_calls must be in [1, 100000]");
  p->timing_mask = boundary_mask ? boundary_mask : ((1 << (NSTAGES + 1)) - 1);
  release_events(p);
  p->events.resize((size_t)max_calls * (NSTAGES + 1));
  for (auto& e : p->events) {
    hipError_t err = hipEventCreate(&e);
    if (err != hipSuccess) {
      e = nullptr;
      release_events(p);
      return hip_fail(err, "timing events");
    }
  }
  p->timing_max = max_calls;
  p->timing_calls = 0;
  return QSSB_OK;
}

int ssb_timing_end(ssb_plan* p, float* stage_ms_host, int n_stages, int* n_calls_host) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!stage_ms_host || n_stages < NSTAGES) return fail(QSSB_ERR_ARG, "stage output too small");
  for (int i = 0; i < NSTAGES; ++i) stage_ms_host[i] = 0.0f;
  const int calls = p->timing_calls;
  for (int c = 0; c < calls; ++c) {
    hipEvent_t* ev = &p->events[(size_t)c * (NSTAGES + 1)];
    int last = NSTAGES;
    while (last > 0 && !((p->timing_mask >> last) & 1)) --last;
    HIP_TRY(hipEventSynchronize(ev[last]), "timing sync");
    for (int i = 0; i < NSTAGES; ++i) {
      if (!((p->timing_mask >> i) & 1) || !((p->timing_mask >> (i + 1)) & 1)) continue;
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "timing elapsed");
      stage_ms_host[i] += ms;
    }
  }
  if (n_calls_host) *n_calls_host = calls;
  release_events(p);
  return QSSB_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// Tuning knobs
// ------------------------------------------------------------------------
static const int kStreamWg[] = {1536, 1024, 2048, 768};

int ssb_tune_candidates(int scan_n, int knob, int* values_host, int capacity) {
  if (scan_n != 128 && scan_n != 256 && scan_n != 512 && scan_n != 1024)
    return fail(QSSB_ERR_UNSUPPORTED, "scan size " + std::to_string(scan_n) + " is not supported");
  int vals[8];
  int n = 0;
  switch (knob) {
    case QSSB_TUNE_STREAM_WG:
      for (int v : kStreamWg) vals[n++] = v;
      break;
    case QSSB_TUNE_GATHER_PAIRS:
      n = gather_pairs_candidates(scan_n, vals);
      break;
    case QSSB_TUNE_TROTTER:
      n = trotter_variants(scan_n);
      for (int i = 0; i < n; ++i) vals[i] = i;
      break;
    case QSSB_TUNE_SUPERSET:
      vals[0] = 2;
      vals[1] = 1;
      n = 2;
      break;
    case QSSB_TUNE_SPLIT:
      vals[0] = 2;
      vals[1] = 1;
      n = 2;
      break;
    default:
      return fail(QSSB_ERR_ARG, "unknown tuning knob " + std::to_string(knob));
  }
  if (values_host)
    for (int i = 0; i < n && i < capacity; ++i) values_host[i] = vals[i];
  return n;
}

int ssb_tune_set(ssb_plan* p, int knob, int value) {
  int rc = check_plan(p);
  if (rc) return rc;
  int vals[8];
  const int n = ssb_tune_candidates(p->n, knob, vals, 8);
  if (n < 0) return n;
  bool ok = (value == 0);
  for (int i = 0; i < n; ++i) ok = ok || (vals[i] == value);
  if (!ok) return fail(QSSB_ERR_ARG, "value " + std::to_string(value) + " is not a candidate of knob " +
                                         std::to_string(knob));
  switch (knob) {
    case QSSB_TUNE_STREAM_WG: p->tune_stream_wg = value; break;
    case QSSB_TUNE_GATHER_PAIRS: p->tune_gather_pairs = value; break;
    case QSSB_TUNE_SUPERSET: p->tune_superset = value; break;
    case QSSB_TUNE_SPLIT: p->tune_split = value; break;
    default: p->tune_trotter = value; break;
  }
  return QSSB_OK;
}

int ssb_tune_get(ssb_plan* p, int knob, int* value_host) {
  int rc = check_plan(p);
  if (rc) return rc;
  if (!value_host) return fail(QSSB_ERR_ARG, "value_host is NULL");
  switch (knob) {
    case QSSB_TUNE_STREAM_WG: *value_host = p->tune_stream_wg ? p->tune_stream_wg : kStreamWg[0]; break;
    case QSSB_TUNE_GATHER_PAIRS:
      *value_host = p->tune_gather_pairs ? p->tune_gather_pairs : gather_pairs_default(p->n);
      break;
    case QSSB_TUNE_TROTTER: *value_host = p->tune_trotter; break;
    case QSSB_TUNE_SUPERSET: *value_host = p->tune_superset ? p->tune_superset : 2; break;
    case QSSB_TUNE_SPLIT: *value_host = p->tune_split ? p->tune_split : 2; break;
    default: return fail(QSSB_ERR_ARG, "unknown tuning knob " + std::to_string(knob));
  }
  return QSSB_OK;
}

// ------------------------------------------------------------------------
// bitshuffle-LZ4 chunk decode
// ------------------------------------------------------------------------
long long ssb_bslz4_workspace_bytes(int n_chunks, long long chunk_elems, int block_bytes) {
  if (n_chunks < 0 || chunk_elems < 1 || block_bytes < 16 || block_bytes > 8192 || block_bytes % 16)
    return fail(QSSB_ERR_ARG, "block_bytes must be a multiple of 16 in [16, 8192] and chunks non-empty");
  return bslz4_workspace_bytes(n_chunks, chunk_elems, block_bytes / 2);
}

int ssb_bslz4_decode_u16(const uint8_t* comp, long long comp_bytes, const long long* chunk_offsets,
                         const long long* chunk_sizes, int n_chunks, long long chunk_elems, int block_bytes,
                         uint16_t* out, void* workspace, long long workspace_bytes, int* status_host, void* stream) {
  const long long need = ssb_bslz4_workspace_bytes(n_chunks, chunk_elems, block_bytes);
  if (need < 0) return (int)need;
  if (n_chunks == 0) return QSSB_OK;
  if (!comp || !chunk_offsets || !chunk_sizes || !out || !workspace)
    return fail(QSSB_ERR_ARG, "bslz4 buffers must not be NULL");
  if (workspace_bytes < need) return fail(QSSB_ERR_ARG, "bslz4 workspace too small");
  if ((reinterpret_cast<uintptr_t>(out) & 15u) || (reinterpret_cast<uintptr_t>(workspace) & 15u))
    return fail(QSSB_ERR_ARG, "bslz4 output and workspace must be 16-byte aligned");
  hipStream_t s = as_stream(stream);
  HIP_TRY(launch_bslz4_decode_u16(comp, comp_bytes, chunk_offsets, chunk_sizes, n_chunks, chunk_elems,
                                  block_bytes / 2, out, workspace, s),
          "bslz4_decode");
  if (status_host) {
    HIP_TRY(hipMemcpyAsync(status_host, workspace, sizeof(int), hipMemcpyDeviceToHost, s), "bslz4 status");
    HIP_TRY(hipStreamSynchronize(s), "bslz4 status");
    if (*status_host)
      return fail(QSSB_ERR_ARG, "malformed bitshuffle-LZ4 chunk (status " + std::to_string(*status_host) + ")");
  }
  return QSSB_OK;
}

int ssb_bin_u16(const uint16_t* src, long long frames, int rows, int cols, int factor, const uint8_t* bad_mask,
                int out_bits, void* out, void* stream) {
  if (frames < 0 || rows < 1 || cols < 1) return fail(QSSB_ERR_ARG, "binning needs frames >= 0 and a non-empty detector");
  if (factor < 1 || factor > 256) return fail(QSSB_ERR_ARG, "det_bin must be in [1, 256]");
  if (factor > rows || factor > cols) return fail(QSSB_ERR_ARG, "det_bin larger than the detector");
  if (out_bits != 16 && out_bits != 32) return fail(QSSB_ERR_ARG, "binned output must be 16 or 32 bits");
  if (frames == 0) return QSSB_OK;
  if (!src || !out) return fail(QSSB_ERR_ARG, "binning buffers must not be NULL");
  HIP_TRY(launch_bin_u16(src, frames, rows, cols, factor, bad_mask, out_bits == 32, out, as_stream(stream)),
          "bin_u16");
  return QSSB_OK;
}

long long ssb_bslz4_encode_bound(long long n_elems, int block_bytes) {
  if (n_elems < 1 || block_bytes < 16 || block_bytes > 8192 || block_bytes % 16)
    return fail(QSSB_ERR_ARG, "block_bytes must be a multiple of 16 in [16, 8192] and the chunk non-empty");
  return bslz4_encode_bound(n_elems, block_bytes);
}

long long ssb_bslz4_encode_u16(const uint16_t* values_host, long long n_elems, int block_bytes, uint8_t* out_host,
                               long long out_capacity) {
  const long long bound = ssb_bslz4_encode_bound(n_elems, block_bytes);
  if (bound < 0) return bound;
  if (!values_host || !out_host) return fail(QSSB_ERR_ARG, "bslz4 encode buffers must not be NULL");
  if (out_capacity < bound) return fail(QSSB_ERR_ARG, "bslz4 encode output too small (see ssb_bslz4_encode_bound)");
  return bslz4_encode_u16(values_host, n_elems, block_bytes, out_host, out_capacity);
}
