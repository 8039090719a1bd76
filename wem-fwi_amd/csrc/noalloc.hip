// Down / Up sweeps that allocate no device memory: the sweep kernel's
// scratch rows and row descriptors live in a file-scope device array that
// every call on a GPU shares.
#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <mutex>
#include <string>

#include "../../include/wemfwi.h"
#include "host.h"
#include "kernels.h"

namespace wemfwi {

// Scratch of the non-allocating sweeps, one copy per GPU for the life of the
// process: [nrows][N] complex scratch rows, then the nrows row descriptors
// (256-byte aligned).  In host code the array's name is not a device
// address: host code uses the per-device address from hipGetSymbolAddress.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static std::mutex& noalloc_mutex() {
  static std::mutex m;
  return m;
}

// device address of `scratch` on the current device, looked up once per
// device
static int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(noalloc_mutex());
  static std::map<int, unsigned char*> addr;
  auto it = addr.find(dev);
  if (it != addr.end()) {
    *out = it->second;
    return WEMFWI_OK;
  }
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(WEMFWI_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
  addr[dev] = (unsigned char*)p;
  *out = (unsigned char*)p;
  return WEMFWI_OK;
}

using cf = std::complex<float>;

}  // namespace wemfwi

using namespace wemfwi;

extern "C" {

long long wemfwi_oneway_noalloc_max_rows(const wemfwi_model* model) {
  if (!model || model->m.N <= 0) return 0;
  const size_t per = (size_t)model->m.N * sizeof(float2) + sizeof(SweepRow);
  // the descriptors start at a 256-byte boundary: keep that slack
  return (long long)((kNoAllocBytes - 256) / per);
}

int wemfwi_oneway_noalloc(wemfwi_model* model, int up, int adjoint, int nrows, const float* freq_hz,
                          const wemfwi_complex* in, wemfwi_complex* out, int add, void* stream) {
  if (!model) return fail(WEMFWI_ERR_ARG, "oneway_noalloc: null model");
  if (nrows < 0 || !freq_hz || !in || !out) return fail(WEMFWI_ERR_ARG, "oneway_noalloc: null argument");
  Model& m = model->m;
  if (int rc = m.require_fft()) return rc;
  if (nrows > wemfwi_oneway_noalloc_max_rows(model))
    return fail(WEMFWI_ERR_ARG, "oneway_noalloc: " + std::to_string(nrows) + " rows of " + std::to_string(m.N) +
                                    " lanes exceed the " + std::to_string(kNoAllocBytes >> 20) +
                                    " MiB scratch (at most " +
                                    std::to_string(wemfwi_oneway_noalloc_max_rows(model)) + " rows)");
  unsigned char* base = nullptr;
  if (int rc = noalloc_scratch(&base)) return rc;
  const size_t rows_bytes = ((size_t)nrows * m.N * sizeof(float2) + 255) & ~(size_t)255;
  return m.oneway_on(up, adjoint, nrows, freq_hz, (const cf*)in, (cf*)out, add, in == out, (float2*)base,
                     (SweepRow*)(base + rows_bytes), true, (hipStream_t)stream);
}

}  // extern "C"
