// noalloc.hip -- the one entry point of libhexopt that allocates no device memory.
//
// hexopt_scaled_jacobian_noalloc evaluates the same scaled Jacobian as hexopt_scaled_jacobian
// (the same kernels, scaled_jacobian_kernel then the min and sum reductions) on device arrays
// the caller owns. The kernels' block partials live in `scratch`, a file-scope device array of
// fixed size: each GPU the library runs on holds one copy until the process ends. Its device
// address is looked up once per device (hipGetSymbolAddress with that device current) and
// kept; host code passes that address, never the array's name, to the kernels.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind; calls that
// overlap on one GPU (two host threads, or two streams in flight) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/hexopt.h"
#include "kernels.h"

namespace {

constexpr int kBlock = 256;                    // hexes per block of scaled_jacobian_kernel
constexpr int kMaxHex = 1 << 24;               // 16,777,216 hexes
constexpr int kMaxBlocks = kMaxHex / kBlock;   // 65,536 block partials of each kind
constexpr int kMaxDevices = 64;

// [0, kMaxBlocks): per-block minimum; [kMaxBlocks, 2 kMaxBlocks): per-block sum
static __device__ __align__(256) double scratch[2 * kMaxBlocks];

double* g_ptr[kMaxDevices] = {};
std::mutex g_mu;

thread_local const char* g_err = "";

// Device address of this device's scratch (the current device must be `device`).
hipError_t scratch_ptr(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = (double*)p;
  }
  *out = g_ptr[device];
  return hipSuccess;
}

}  // namespace

extern "C" {

const char* hexopt_noalloc_error(void) { return g_err; }

int hexopt_scaled_jacobian_noalloc(int device, int nV, const double* X, int nH, const int32_t* conn, double* per_hex,
                                   double* min_sum, void* stream) {
  g_err = "";
  if (device < 0 || device >= kMaxDevices) return (g_err = "device index out of range", HEXOPT_EINVAL);
  if (nV < 0 || nH < 0 || !X || !conn || !min_sum) return (g_err = "null array or negative size", HEXOPT_EINVAL);
  if (nH > kMaxHex) return (g_err = "more than 16,777,216 hexes (the scratch array's limit)", HEXOPT_EINVAL);
  if (nH == 0) return (g_err = "mesh without hexes", HEXOPT_EINVAL);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (g_err = hipGetErrorString(e), HEXOPT_EHIP);
  double* s = nullptr;
  if ((e = scratch_ptr(device, &s)) != hipSuccess) return (g_err = hipGetErrorString(e), HEXOPT_EHIP);
  hexopt::launch_scaled_jacobian(X, nV, conn, nH, per_hex, s, s + kMaxBlocks, min_sum, (hipStream_t)stream);
  if ((e = hipGetLastError()) != hipSuccess) return (g_err = hipGetErrorString(e), HEXOPT_EHIP);
  return HEXOPT_OK;
}

}  // extern "C"
