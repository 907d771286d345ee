// Allocation-free scan normalisation (ssb_normalize_scan_u16_noalloc).
//
// Its partial sums and pad frame live in `scratch`, a file-scope device array
// of QSSB_NOALLOC_SCRATCH_BYTES; every device holds its own copy, whose
// address hipGetSymbolAddress gives once per device (on the first call with
// that device current) and which is kept in g_scratch_ptr.  Host code uses
// only those addresses, never the symbol.  The kernels write every partial
// and pad element before reading it, so nothing needs clearing between
// calls; calls that overlap on one device share the array and race.
#include <mutex>
#include <string>

#include "ssb_internal.h"

using namespace qssb;

namespace {

constexpr int kMaxDevices = 64;
__device__ __align__(256) uint8_t scratch[QSSB_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(uint8_t** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return report_hip_error(e, "scratch device");
  if (dev < 0 || dev >= kMaxDevices) return report_error(QSSB_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return report_hip_error(e, "scratch symbol address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = reinterpret_cast<uint8_t*>(g_scratch_ptr[dev]);
  return QSSB_OK;
}

}  // namespace

extern "C" int ssb_normalize_scan_u16_noalloc(const uint16_t* src, int H, int W, int det_rows, int det_cols,
                                              uint16_t* dst, int T, void* stream) {
  clear_error();
  if (!src || !dst) return report_error(QSSB_ERR_ARG, "source and destination must not be NULL");
  if (H < 1 || W < 1 || det_rows < 1 || det_cols < 1) return report_error(QSSB_ERR_ARG, "shapes must be positive");
  if (T != ssb_scan_target(H, W))
    return report_error(QSSB_ERR_ARG, "target scan size " + std::to_string(T) +
                                          " is not the supported size for a " + std::to_string(H) + " x " +
                                          std::to_string(W) + " scan (" + std::to_string(ssb_scan_target(H, W)) +
                                          ")");
  const long long D = (long long)det_rows * det_cols;
  if (D > (1 << 24)) return report_error(QSSB_ERR_UNSUPPORTED, "detector too large");
  size_t mean_off = 0;
  const size_t need = normalize_workspace_bytes(QSSB_DTYPE_U16, H, T, (int)D, &mean_off);
  if (need > (size_t)QSSB_NOALLOC_SCRATCH_BYTES)
    return report_error(QSSB_ERR_ARG, "scan " + std::to_string(H) + " x " + std::to_string(W) + " with a " +
                                          std::to_string(det_rows) + " x " + std::to_string(det_cols) +
                                          " detector needs " + std::to_string(need) +
                                          " bytes of scratch; the allocation-free normalisation holds " +
                                          std::to_string((long long)QSSB_NOALLOC_SCRATCH_BYTES) +
                                          " (use ssb_normalize_scan_u16)");
  uint8_t* ws = nullptr;
  const int rc = scratch_of_current_device(&ws);
  if (rc) return rc;
  const hipError_t e = launch_normalize_scan_typed(QSSB_DTYPE_U16, src, H, W, (int)D, dst, T, ws, ws + mean_off,
                                                   reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return report_hip_error(e, "normalize_scan_noalloc");
  return QSSB_OK;
}
