// Allocation-free resampling entry point.
//
// bec_resample allocates a small device buffer for the per-block partial sums
// of |psi|^2 on every call.  Callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) call bec_resample_noalloc instead: the same kernels and the same
// host arithmetic, with the partial rows in the file-scope device array
// `scratch` below.
//
// `scratch` is one array per GPU, held until the process ends.  Host code
// never passes its name to a kernel or a copy: the name is a host-side
// stand-in, not a device address.  Its device address on each GPU is looked
// up once (hipGetSymbolAddress, with that GPU current) and kept in
// g_scratch[device].  Every call on a GPU uses that GPU's array, which keeps
// the last call's rows: calls that overlap on one GPU race.
#include <math.h>
#include <mutex>

#include "../../include/becgpp_hip.h"
#include "bec_ops.h"

#define NOALLOC_ROWS 1024          // per-block rows of the resampling launch (its grid is at most 1024 blocks)
#define NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[NOALLOC_ROWS];

static std::mutex g_scratch_mu;
static double* g_scratch[NOALLOC_MAX_DEVICES];

// Device address of this GPU's scratch array, looked up on the first call
// made with that GPU current.
static double* scratch_ptr() {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw BecError(2, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw BecError(2, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = (double*)ptr;
  }
  return g_scratch[dev];
}

extern "C" int bec_resample_noalloc(int ndim, int n0, double dx0, const void* psi_old, int n1, double dx1,
                                    void* psi_new, double* norm_out, void* stream) {
  try {
    if (!psi_old || !psi_new || !norm_out) throw BecError(1, "null argument");
    if (ndim != 2 && ndim != 3) throw BecError(1, "ndim must be 2 or 3");
    if (n0 < 1 || n1 < 1 || !(dx0 > 0.0) || !(dx1 > 0.0)) throw BecError(1, "grid sizes and spacings must be positive");
    *norm_out = 0.0;
    const double x0 = -(double)(n0 / 2) * dx0;              // first and last node of the old grid
    const double x1 = x0 + (double)(n0 - 1) * dx0;
    if (!(x1 > x0)) return 0;                                 // degenerate: no interpolation
    const long long n = ndim == 2 ? (long long)n1 * n1 : (long long)n1 * n1 * n1;
    const int g = elem_blocks(n);
    if (g > NOALLOC_ROWS)
      throw BecError(1, "bec_resample_noalloc: the launch needs " + std::to_string(g) +
                            " partial rows, more than the scratch array's " + std::to_string(NOALLOC_ROWS));
    double* rows = scratch_ptr();
    hipStream_t s = (hipStream_t)stream;
    double r[NOALLOC_ROWS];
    const int gl = launch_resample((const double2*)psi_old, ndim, n0, x0, x1 - x0, n1, dx1, (double2*)psi_new, rows, s);
    if (gl != g) throw BecError(2, "internal: resampling grid differs from its row count");
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(r, rows, (size_t)g * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // the norm exactly as bec_resample forms it: rows summed in order on the host
    double sum = 0.0;
    for (int b = 0; b < g; ++b) sum += r[b];
    const double nn = sqrt(sum * (ndim == 2 ? dx1 * dx1 : dx1 * dx1 * dx1));
    *norm_out = nn;
    if (std::isfinite(nn) && nn > 1e-14) {
      launch_scale((const double2*)psi_new, (double2*)psi_new, &nn, n, 1, s);
      HIPCHK(hipGetLastError());
    }
    return 0;
  } catch (const BecError& e) {
    bec_set_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bec_set_error(e.what());
    return 2;
  }
}
