/*
 * wave.h -- wave-level (64-lane) reductions and scans for the gfx950 kernels,
 * on DPP moves (cross-lane operand modifiers of the vector ALU) instead of
 * ds_bpermute, which costs an LDS round trip per step; and the population
 * kernels' ordered one-lane sum of LDS values.
 */
#ifndef GSBN_WAVE_H
#define GSBN_WAVE_H

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

/* Wave-wide inclusive prefix sum with DPP moves (gfx9 DPP: row shifts by
 * 1, 2, 4, 8 within each 16-lane row, out-of-row sources reading 0, then
 * row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3), a
 * few cycles per step instead of an LDS round trip per ds_bpermute. */
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true); /* row_shr:1 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true); /* row_shr:2 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true); /* row_shr:4 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true); /* row_shr:8 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); /* row_bcast:15 */
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); /* row_bcast:31 */
  return x;
}

/* The maximum over the wave's 64 lanes under the reference's sequential
 * comparison (`if (d > max) max = d`, Pop.cpp:270-279): op(l, r) = r > l ?
 * r : l keeps the leftmost of equal values, is associative for NaN-free
 * values (+-0 ties included), and is evaluated as a DPP prefix (lower
 * lanes on the left) whose last lane holds the whole wave's result. */
__device__ __forceinline__ float wave_max_left(float v) {
  const int ninf = __float_as_int(-INFINITY);
  float o;
  o = __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x111, 0xf, 0xf, false));
  v = v > o ? v : o;
  o = __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x112, 0xf, 0xf, false));
  v = v > o ? v : o;
  o = __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x114, 0xf, 0xf, false));
  v = v > o ? v : o;
  o = __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x118, 0xf, 0xf, false));
  v = v > o ? v : o;
  o = __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x142, 0xa, 0xf, false));
  v = v > o ? v : o;
  o = __int_as_float(__builtin_amdgcn_update_dpp(ninf, __float_as_int(v), 0x143, 0xc, 0xf, false));
  v = v > o ? v : o;
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

/* ((p[0] + p[1]) + p[2]) + ... + p[n-1] from 0, in that order (the
 * reference's sequential soft-max normaliser, Pop.cpp:281-291), by the
 * calling lane from LDS (p 16-byte aligned).  32 values per LDS round trip:
 * the eight 16-byte reads of a block are issued before its additions, so a
 * block costs one LDS latency instead of one per read or two. */
__device__ __forceinline__ float ordered_sum_lds(const float* p, int n) {
  const float4* v4 = reinterpret_cast<const float4*>(p);
  float s = 0.0f;
  int m = 0;
  for (; m + 32 <= n; m += 32) {
    float4 x[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = v4[(m >> 2) + i];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      s += x[i].x;
      s += x[i].y;
      s += x[i].z;
      s += x[i].w;
    }
  }
  for (; m + 4 <= n; m += 4) {
    const float4 x = v4[m >> 2];
    s += x.x;
    s += x.y;
    s += x.z;
    s += x.w;
  }
  for (; m < n; ++m) s += p[m];
  return s;
}

#endif /* GSBN_WAVE_H */
