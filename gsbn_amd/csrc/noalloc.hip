/*
 * noalloc.hip -- the population update for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).
 *
 * gsbn_pop_update takes a workspace argument, scratch_dsup0: the dsup of the
 * population's first unit, computed by a pre-kernel when the WTA maximum is
 * seeded the reference CPU's way (wta_pop_max, Pop.cpp:270) and read by every
 * workgroup of the population kernel.  gsbn_pop_update_noalloc takes that
 * workspace from a file-scope device array instead, so the caller passes
 * inputs and outputs only.  Each GPU holds its own copy of the array (until
 * the process ends); its device address is looked up once per device with
 * hipGetSymbolAddress, and only that address is ever handed to a kernel.
 */
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdint.h>

#include "../../include/gsbn_amd.h"

/* floats of workspace per call: one (the first unit's dsup) */
#define GSBN_NOALLOC_SCRATCH_FLOATS 64
#define GSBN_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) float scratch[GSBN_NOALLOC_SCRATCH_FLOATS];

namespace {

std::mutex g_scratch_mu;
float* g_scratch_ptr[GSBN_NOALLOC_MAX_DEVICES] = {nullptr};

/* the device address of `scratch` on the current device, looked up on the
 * first call on that device */
int scratch_on_current_device(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= GSBN_NOALLOC_MAX_DEVICES) return GSBN_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return GSBN_ERR_STATE;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int gsbn_pop_update_noalloc(const gsbn_pop_args* a, void* stream) {
  if (!a) return GSBN_ERR_ARG;
  /* the workspace this call needs, against what the array holds */
  const int need = a->wta_pop_max ? 1 : 0;
  if (need > GSBN_NOALLOC_SCRATCH_FLOATS) return GSBN_ERR_UNSUPPORTED;
  float* ptr = nullptr;
  const int rc = scratch_on_current_device(&ptr);
  if (rc != 0) return rc;
  gsbn_pop_args b = *a;
  b.scratch_dsup0 = ptr;
  return gsbn_pop_update(&b, stream);
}
