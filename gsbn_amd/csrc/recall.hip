/*
 * recall.hip -- frozen-weight batched recall.
 *
 * A trained population with one recurrent projection recalls B cues in
 * lock-step.  Weights Wij and biases bj are frozen (the reference's recall
 * mode, prn = 0: Proj.cpp:310-363 leaves bj alone and Proj.cpp:365-479 adds
 * the stored Wij to the support without recomputing it), so per cue b and
 * step t only the population dynamics run:
 *
 *   Pop(t)    support from E_{t-1}, dsup, WTA soft-max, spikes
 *             (Pop.cpp:229-323, as the solver's population kernel),
 *   E_t     = E_{t-1} (1 - kepsc) + sum of W over the rows active at t
 *             (jxx decay then row additions, Proj.cpp:310-363, 365-479),
 *
 * a row being active at t when its source unit spiked at t - di - 2.  Each
 * cue starts from the population's initial state (dsup = log(1/M),
 * Pop.cpp:59), an empty delay queue and E = 0, and draws with Philox key
 * seed + b, population 0.
 *
 * Two support paths:
 *   exact  float weights, the additions in the reference's ascending row
 *          order: identical to a single-cue run of the solver (and the
 *          oracle) with the same weights;
 *   mfma   the support of all cues as one product S_{t-d-2} W of the 0/1
 *          spike matrix (B x Nsrc, bf16) and W (Nsrc x N, bf16) on the
 *          gfx950 matrix cores with float accumulation, fused with the
 *          decay: E = E (1 - kepsc) + S W.  Needs one delay for all rows.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gsbn_amd.h"
#include "detmath.h"
#include "gsbn_core.h"
#include "philox.h"
#include "wave.h"

#define RE_THREADS_MAX 1024
#define RE_RING 64 /* spike-history slots; latency di + 2 <= 33 */
#define RE_ROWS_MAX 2048 /* active rows compacted into LDS per (HCU, cue) and step */

/* GEMM tiling: 256 x 256 output tile per workgroup of 16 waves (64 x 64
 * each, 2 x 2 blocks of 32 x 32), K in steps of 64, LDS rows padded by 8 bf16 */
#define GM_BM 256
#define GM_BN 256
#define GM_BK 64
#define GM_LDK (GM_BK + 8)
#define GM_THREADS 1024

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 gsbn_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float gsbn_f32x16;

namespace {

struct ReArgs {
  int B, Bp, H, M, C, N, Kp, W, dmin, dmax, count_from, mode;
  uint64_t seed;
  float taumdt, wtagain, maxfqdt, igain, wgain, lgbias, snoise, kepsc, lg_on, lg_off;
  float* dsup;            /* [B][N] */
  float* epsc;            /* [B][N] */
  int32_t* count;         /* [B][N] */
  const float* bj;        /* [N] */
  const float* wij;       /* [H*C*M] (exact) */
  const int32_t* di;      /* [H*C] */
  const int32_t* inv_off; /* [H][N+1] */
  const int32_t* inv_slot;
  int32_t* spk_cnt;       /* [RE_RING][B][H] */
  uint16_t* spk_list;     /* [RE_RING][B][H][M] */
  uint16_t* S;            /* [RE_RING][B][Kp] bf16 0/1 (mfma) */
  const int16_t* clamp;   /* [B][H] */
  const float* wmask;     /* [B][H] */
  int8_t* raster;         /* [B][N] spikes of the step, or NULL */
};

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

/* exclusive prefix sum across the workgroup (blockDim a multiple of 64,
 * sh: blockDim / 64 ints), with the total */
__device__ __forceinline__ int block_scan_fast_re(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int x = (int)wave_incl_scan((uint32_t)v);
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < nw; ++w) {
    const int t = sh[w];
    if (w < wid) base += t;
    tot += t;
  }
  *total = tot;
  __syncthreads();
  return base + x - v;
}

/* one workgroup per (HCU, cue): the population step and, in the exact mode,
 * the support update of that HCU's units */
__global__ void __launch_bounds__(RE_THREADS_MAX) recall_step_kernel(ReArgs a, int t) {
  extern __shared__ uint32_t sh_bits[];
  __shared__ __attribute__((aligned(16))) float sh_f[RE_THREADS_MAX];
  __shared__ int sh_scan[RE_THREADS_MAX / 64 + 1];
  __shared__ float sh_vsum;
  const int h = blockIdx.x, b = blockIdx.y, j = threadIdx.x;
  const int M = a.M, N = a.N;
  const bool live = j < M;
  const int n = h * M + j;
  const int64_t bn = (int64_t)b * N + n;

  /* support, dsup (Pop.cpp:229-261) */
  float u = 0.0f, dsup = 0.0f, E = 0.0f;
  if (live) {
    float z;
    gsbn_draw_unit(a.seed + (uint64_t)b, 0u, (uint32_t)t, (uint32_t)n, &u, &z);
    const float xi = a.snoise * z;
    E = a.epsc[bn];
    float wsup = 0.0f;
    wsup += a.bj[n] + E;
    const float lginp = (a.clamp[(int64_t)b * a.H + h] == j) ? a.lg_on : a.lg_off;
    float sup = a.lgbias + a.igain * lginp + xi;
    sup += (a.wgain * a.wmask[(int64_t)b * a.H + h]) * wsup;
    const float d0 = a.dsup[bn];
    dsup = d0 + (sup - d0) * a.taumdt;
    a.dsup[bn] = dsup;
  }
  /* WTA maximum: first of equal maxima, as the sequential reference */
  {
    const float mv = wave_max_left(live ? dsup : -INFINITY);
    if ((j & 63) == 0) sh_f[j >> 6] = mv;
  }
  __syncthreads();
  float maxdsup = sh_f[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
    const float o = sh_f[w];
    if (o > maxdsup) maxdsup = o;
  }
  __syncthreads();
  /* soft-max (Pop.cpp:280-302) */
  const float maxact = gsbn_expf(a.wtagain * maxdsup);
  float act = 0.0f;
  if (live) {
    act = gsbn_expf(a.wtagain * (dsup - maxdsup));
    if (maxact < 1.0f) act *= maxact;
    sh_f[j] = act;
  }
  __syncthreads();
  if (j == 0) sh_vsum = ordered_sum_lds(sh_f, M);
  __syncthreads();
  if (live && sh_vsum > 1.0f) act /= sh_vsum;
  /* spikes (Pop.cpp:305-323) */
  const int spk = (live && u < act * a.maxfqdt) ? 1 : 0;
  if (live) {
    if (t > a.count_from) a.count[bn] += spk;
    if (a.raster) a.raster[bn] = (int8_t)spk;
  }
  const int slot = t % RE_RING;
  if (a.mode == GSBN_RECALL_MFMA) {
    if (live) a.S[((int64_t)slot * a.Bp + b) * a.Kp + n] = spk ? (uint16_t)0x3F80u : (uint16_t)0u;
    return;
  }
  /* spike list of this HCU (ascending) for the queue of later steps */
  {
    const int lane = j & 63, wid = j >> 6, nw = blockDim.x >> 6;
    const uint64_t bal = __ballot(spk);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) sh_scan[wid] = __popcll(bal);
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
      if (w < wid) base += sh_scan[w];
      tot += sh_scan[w];
    }
    const int64_t lb = ((int64_t)slot * a.B + b) * a.H + h;
    if (spk) a.spk_list[lb * M + base + before] = (uint16_t)j;
    if (j == 0) a.spk_cnt[lb] = tot;
  }
  /* exact support update: E (1 - kepsc), then + Wij of every row active at
   * t in ascending row order (Proj.cpp:310-363 then 365-479) */
  const int C = a.C, Wd = a.W;
  for (int w = j; w < Wd; w += blockDim.x) sh_bits[w] = 0u;
  __syncthreads();
  const int32_t* off = a.inv_off + (int64_t)h * (N + 1);
  for (int d = a.dmin; d <= a.dmax; ++d) {
    const int s = t - d - 2;
    if (s < 1) continue;
    const int ss = s % RE_RING;
    const int32_t* cnts = a.spk_cnt + ((int64_t)ss * a.B + b) * a.H;
    const uint16_t* lsts = a.spk_list + ((int64_t)ss * a.B + b) * a.H * M;
    for (int hs = j; hs < a.H; hs += blockDim.x) {
      const int c = cnts[hs];
      for (int i = 0; i < c; ++i) {
        const int src = hs * M + lsts[(int64_t)hs * M + i];
        for (int q = off[src]; q < off[src + 1]; ++q) {
          const int k = a.inv_slot[q];
          if (a.di[(int64_t)h * C + k] == d) atomicOr(&sh_bits[k >> 5], 1u << (k & 31));
        }
      }
    }
  }
  __syncthreads();
  /* the active rows in ascending order, compacted into LDS (after the bitmap
   * in the dynamic LDS), so each thread can have eight weight loads in
   * flight instead of one per row; more than RE_ROWS_MAX rows (not reached
   * by recall, whose active rows are the few spiking sources) fall back to
   * walking the bitmap */
  int* const sh_rows = reinterpret_cast<int*>(sh_bits + Wd);
  int nrows;
  {
    const int per = (Wd + (int)blockDim.x - 1) / (int)blockDim.x;
    const int w0 = j * per;
    const int w1 = w0 + per < Wd ? w0 + per : Wd;
    int c = 0;
    for (int w = w0; w < w1; ++w) c += __popc(sh_bits[w]);
    int pos = block_scan_fast_re(c, sh_scan, &nrows);
    if (nrows <= RE_ROWS_MAX) {
      for (int w = w0; w < w1; ++w) {
        uint32_t bits = sh_bits[w];
        while (bits) {
          sh_rows[pos++] = (w << 5) + __ffs(bits) - 1;
          bits &= bits - 1;
        }
      }
    }
    __syncthreads();
  }
  if (live) {
    float e = E * (1 - a.kepsc);
    const float* wrow = a.wij + (int64_t)h * C * M + j;
    if (nrows <= RE_ROWS_MAX) {
      for (int r = 0; r < nrows; r += 8) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = r + q < nrows ? wrow[(int64_t)sh_rows[r + q] * M] : 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (r + q < nrows) e = e + v[q];
      }
    } else {
      for (int w = 0; w < Wd; ++w) {
        uint32_t bits = sh_bits[w];
        while (bits) {
          const int k = (w << 5) + __ffs(bits) - 1;
          bits &= bits - 1;
          e = e + wrow[(int64_t)k * M];
        }
      }
    }
    a.epsc[bn] = e;
  }
}

/* The population step of the mfma mode with one wave per (HCU, cue) and
 * four pairs per workgroup: lane l takes units l, l + 64, ... (UPL per
 * lane), so no workgroup barrier is needed and a hypercolumn of M <= 64 * UPL
 * units occupies one wave instead of ceil(M / 64).  Same operations per unit
 * as recall_step_kernel: the WTA maximum combines the 64-unit ranges left to
 * right with the sequential comparison, and the normaliser is summed in
 * ascending unit order by lane 0 from the wave's LDS row. */
template <int UPL>
__global__ void __launch_bounds__(256) recall_pop_wave_kernel(ReArgs a, int t) {
  __shared__ __attribute__((aligned(16))) float sh_act[4][UPL * 64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t pair = (int64_t)blockIdx.x * 4 + wave; /* b * H + h */
  if (pair >= (int64_t)a.B * a.H) return;              /* uniform per wave */
  const int b = (int)(pair / a.H), h = (int)(pair - (int64_t)b * a.H);
  const int M = a.M, N = a.N;
  const int cl = a.clamp[pair];
  const float wm = a.wmask[pair];
  float* const row = sh_act[wave];
  float dsup[UPL], u[UPL];
  int cnt[UPL];
  /* support, dsup (Pop.cpp:229-261) */
#pragma unroll
  for (int i = 0; i < UPL; ++i) {
    const int j = lane + 64 * i;
    dsup[i] = 0.0f;
    u[i] = 0.0f;
    cnt[i] = 0;
    if (j < M) {
      const int n = h * M + j;
      const int64_t bn = (int64_t)b * N + n;
      const float E = a.epsc[bn];
      const float bjv = a.bj[n];
      const float d0 = a.dsup[bn];
      if (t > a.count_from) cnt[i] = a.count[bn];
      float z;
      gsbn_draw_unit(a.seed + (uint64_t)b, 0u, (uint32_t)t, (uint32_t)n, &u[i], &z);
      const float xi = a.snoise * z;
      float wsup = 0.0f;
      wsup += bjv + E;
      const float lginp = cl == j ? a.lg_on : a.lg_off;
      float sup = a.lgbias + a.igain * lginp + xi;
      sup += (a.wgain * wm) * wsup;
      dsup[i] = d0 + (sup - d0) * a.taumdt;
      a.dsup[bn] = dsup[i];
    }
  }
  /* WTA maximum: each 64-unit range by shuffles, the ranges left to right */
  float maxdsup = 0.0f;
#pragma unroll
  for (int i = 0; i < UPL; ++i) {
    const float mv = wave_max_left(lane + 64 * i < M ? dsup[i] : -INFINITY);
    if (i == 0) {
      maxdsup = mv;
    } else if (64 * i < M && mv > maxdsup) {
      maxdsup = mv;
    }
  }
  /* soft-max (Pop.cpp:280-302) */
  const float maxact = gsbn_expf(a.wtagain * maxdsup);
  float act[UPL];
#pragma unroll
  for (int i = 0; i < UPL; ++i) {
    act[i] = 0.0f;
    if (lane + 64 * i < M) {
      act[i] = gsbn_expf(a.wtagain * (dsup[i] - maxdsup));
      if (maxact < 1.0f) act[i] *= maxact;
      row[lane + 64 * i] = act[i];
    }
  }
  float vsum = 0.0f;
  if (lane == 0) vsum = ordered_sum_lds(row, M);
  vsum = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vsum)));
  /* spikes (Pop.cpp:305-323) */
  const int slot = t % RE_RING;
#pragma unroll
  for (int i = 0; i < UPL; ++i) {
    const int j = lane + 64 * i;
    if (j >= M) continue;
    float ac = act[i];
    if (vsum > 1.0f) ac /= vsum;
    const int spk = u[i] < ac * a.maxfqdt ? 1 : 0;
    const int n = h * M + j;
    const int64_t bn = (int64_t)b * N + n;
    if (t > a.count_from) a.count[bn] = cnt[i] + spk;
    if (a.raster) a.raster[bn] = (int8_t)spk;
    a.S[((int64_t)slot * a.Bp + b) * a.Kp + n] = spk ? (uint16_t)0x3F80u : (uint16_t)0u;
  }
}

/* E[b][n] = beta E[b][n] + sum_k S[b][k] WT[n][k]: bf16 inputs, float
 * accumulation on v_mfma_f32_32x32x16_bf16.  S and WT have Kp columns (Kp a
 * multiple of GM_BK, zero padded); S has ceil(B / GM_BM) * GM_BM rows and WT
 * ceil(N / GM_BN) * GM_BN rows, so tiles never read past them.
 * A 256 x 256 output tile per workgroup of 16 waves (4 x 4, 64 x 64 each:
 * 2 x 2 MFMA blocks), K in steps of 64 through one LDS buffer (rows padded
 * by 8 bf16 against bank conflicts) with the next K tile staged in
 * registers during the current one's MFMAs.  (Measured at this problem's
 * shape: 128 x 128 tiles of 4 waves reach 0.79 PFLOP/s, double-buffered LDS
 * 0.75, 256 x 256 tiles of 16 waves 0.96; larger wave tiles, 8-wave
 * workgroups and K steps of 32 do worse.)
 * Workgroup w is remapped so that consecutive tiles of one XCD (hardware
 * workgroup w runs on XCD w % 8) share their WT tile in that XCD's L2. */
__global__ void __launch_bounds__(GM_THREADS)
    spike_gemm_kernel(const uint16_t* __restrict__ S, const uint16_t* __restrict__ WT,
                      float* __restrict__ E, int Bn, int N, int Kp, float beta, int mtiles,
                      int ntiles) {
  __shared__ __attribute__((aligned(16))) uint16_t sA[GM_BM * GM_LDK];
  __shared__ __attribute__((aligned(16))) uint16_t sB[GM_BN * GM_LDK];
  const int nwg = mtiles * ntiles;
  int wg = blockIdx.x;
  if ((nwg & 7) == 0) wg = (wg & 7) * (nwg >> 3) + (wg >> 3);
  const int nt = wg / mtiles, mt = wg - nt * mtiles;
  const int m0 = mt * GM_BM, n0 = nt * GM_BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;

  gsbn_f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][jj][r] = 0.0f;

  /* global -> register staging: 2 chunks of 16 bytes per thread and matrix;
   * chunk c = tid + 1024 i covers row c / 8, columns (c % 8) * 8 .. +7 of
   * the K tile.  Named registers (no arrays) keep the staging out of
   * scratch. */
  const int crow = tid >> 3, ccol = (tid & 7) * 8; /* chunk i: row crow + 128 i */
  const uint16_t* pa = S + (int64_t)(m0 + crow) * Kp + ccol;
  const uint16_t* pb = WT + (int64_t)(n0 + crow) * Kp + ccol;
  const int64_t rs = (int64_t)128 * Kp; /* 128 rows further */
  uint4 a0, a1, b0, b1;
#define GM_LOAD(k0)                                     \
  a0 = *reinterpret_cast<const uint4*>(pa + (k0));      \
  a1 = *reinterpret_cast<const uint4*>(pa + rs + (k0)); \
  b0 = *reinterpret_cast<const uint4*>(pb + (k0));      \
  b1 = *reinterpret_cast<const uint4*>(pb + rs + (k0));
#define GM_STORE()                                                           \
  *reinterpret_cast<uint4*>(&sA[(crow)*GM_LDK + ccol]) = a0;                 \
  *reinterpret_cast<uint4*>(&sA[(crow + 128) * GM_LDK + ccol]) = a1;         \
  *reinterpret_cast<uint4*>(&sB[(crow)*GM_LDK + ccol]) = b0;                 \
  *reinterpret_cast<uint4*>(&sB[(crow + 128) * GM_LDK + ccol]) = b1;

  const int nk = Kp / GM_BK;
  GM_LOAD(0)
  GM_STORE()
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) {
      GM_LOAD((kt + 1) * GM_BK)
    }
#pragma unroll
    for (int ks = 0; ks < GM_BK / 16; ++ks) {
      const int kk = ks * 16 + (lane >> 5) * 8;
      gsbn_bf16x8 af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
        af[i] = *reinterpret_cast<const gsbn_bf16x8*>(
            &sA[(wm * 64 + i * 32 + (lane & 31)) * GM_LDK + kk]);
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
        bf[jj] = *reinterpret_cast<const gsbn_bf16x8*>(
            &sB[(wn * 64 + jj * 32 + (lane & 31)) * GM_LDK + kk]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
          acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bf[jj], acc[i][jj], 0, 0, 0);
    }
    if (kt + 1 < nk) {
      __syncthreads(); /* every wave is done with this K tile */
      GM_STORE()
      __syncthreads();
    }
  }
#undef GM_LOAD
#undef GM_STORE
  /* accumulator r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4 and
   * column l % 32 of its 32 x 32 block */
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int col = n0 + wn * 64 + jj * 32 + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 64 + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (row < Bn && col < N) {
          float* p = E + (int64_t)row * N + col;
          *p = beta * (*p) + acc[i][jj][r];
        }
      }
    }
}

/* first strict maximum of each (cue, HCU)'s counts (ProcCheck/Pop.cpp:35-50) */
__global__ void recall_argmax_kernel(int B, int H, int M, const int32_t* count, int32_t* idx,
                                     int32_t* cnt) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= (int64_t)B * H) return;
  int mi = -1, mc = 0;
  const int32_t* c = count + g * M;
  for (int j = 0; j < M; ++j) {
    if (c[j] > mc) {
      mi = j;
      mc = c[j];
    }
  }
  idx[g] = mi;
  cnt[g] = mc;
}

__global__ void re_fill_f32(float* p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

int grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 16384);
}

}  // namespace

namespace gsbn {
void set_last_error(const std::string& m);
}

namespace {

int fail(int code, const std::string& m) {
  gsbn::set_last_error(m);
  return code;
}

/* E = beta E + 0: the epilogue of the product when no row is active */
__global__ void re_decay(float* E, int64_t n, float beta) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    E[i] = beta * E[i] + 0.0f;
}

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t count) {
    n = count;
    return count ? hipMalloc(&p, count * sizeof(T)) : hipSuccess;
  }
  hipError_t upload(const T* h, size_t count) {
    hipError_t e = alloc(count);
    if (e == hipSuccess && count) e = hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice);
    return e;
  }
};

}  // namespace

struct gsbn_recall {
  gsbn_recall_desc d;
  int N = 0, Kp = 0, Np = 0, Bp = 0, W = 0, dmin = 1, dmax = 0, delay = 0;
  DBuf<float> dsup, epsc, bj, wij;
  DBuf<int32_t> count, di, inv_off, inv_slot, spk_cnt, out_idx, out_cnt;
  DBuf<uint16_t> spk_list, S, WT;
  DBuf<int16_t> clamp;
  DBuf<float> wmask;
  DBuf<int8_t> raster;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ~gsbn_recall() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" {

int gsbn_spike_gemm(int32_t B, int32_t N, int32_t Kp, const uint16_t* S, const uint16_t* WT,
                    float* E, float beta, void* stream) {
  if (B <= 0 || N <= 0 || Kp <= 0 || Kp % GM_BK || !S || !WT || !E) return GSBN_ERR_ARG;
  const int mtiles = (B + GM_BM - 1) / GM_BM, ntiles = (N + GM_BN - 1) / GM_BN;
  const int64_t nwg = (int64_t)mtiles * ntiles;
  if (nwg > 0x7fffffff) return GSBN_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(spike_gemm_kernel, dim3((unsigned)nwg), dim3(GM_THREADS), 0,
                     (hipStream_t)stream, S, WT, E, B, N, Kp, beta, mtiles, ntiles);
  return ret(hipGetLastError());
}

gsbn_recall* gsbn_recall_new(const gsbn_recall_desc* d) {
  if (!d || d->H <= 0 || d->M <= 0 || d->M > RE_THREADS_MAX || d->C <= 0 || d->batch <= 0 ||
      d->batch > 65535 ||
      !d->ii || !d->di || !d->wij || !d->bj ||
      (d->mode != GSBN_RECALL_EXACT && d->mode != GSBN_RECALL_MFMA)) {
    fail(GSBN_ERR_ARG, "gsbn_recall_new: invalid description");
    return nullptr;
  }
  gsbn_recall* r = new gsbn_recall();
  r->d = *d;
  r->d.ii = nullptr;
  r->d.di = nullptr;
  r->d.wij = nullptr;
  r->d.bj = nullptr;
  const int H = d->H, M = d->M, C = d->C, B = d->batch;
  const int N = H * M;
  const int64_t R = (int64_t)H * C;
  r->N = N;
  r->W = (C + 31) / 32;
  auto bad = [&](int code, const std::string& m) -> gsbn_recall* {
    fail(code, m);
    delete r;
    return nullptr;
  };
  if ((size_t)r->W * 4 > 48 * 1024) return bad(GSBN_ERR_UNSUPPORTED, "gsbn_recall_new: C too large");
  /* delays and the inverse connectivity per (HCU, source unit) */
  std::vector<int32_t> off((size_t)H * (N + 1), 0), slots;
  int dmin = 32, dmax = 0;
  for (int h = 0; h < H; ++h) {
    std::vector<int32_t> cnt(N, 0);
    for (int k = 0; k < C; ++k) {
      const int s = d->ii[(int64_t)h * C + k];
      if (s < 0) continue;
      if (s >= N) return bad(GSBN_ERR_ARG, "gsbn_recall_new: ii out of range");
      const int dl = d->di[(int64_t)h * C + k];
      if (dl < 1 || dl > 31) return bad(GSBN_ERR_ARG, "gsbn_recall_new: delay outside [1, 31]");
      cnt[s]++;
      dmin = std::min(dmin, dl);
      dmax = std::max(dmax, dl);
    }
    int32_t* o = off.data() + (size_t)h * (N + 1);
    o[0] = (int32_t)slots.size();
    for (int s = 0; s < N; ++s) o[s + 1] = o[s] + cnt[s];
    slots.resize((size_t)o[N]);
    std::vector<int32_t> fill(o, o + N);
    for (int k = 0; k < C; ++k) {
      const int s = d->ii[(int64_t)h * C + k];
      if (s >= 0) slots[(size_t)fill[s]++] = k;
    }
  }
  if (dmax == 0) {
    dmin = 1;
    dmax = 0;
  }
  r->dmin = dmin;
  r->dmax = dmax;
  if (d->mode == GSBN_RECALL_MFMA && dmax > 0 && dmin != dmax)
    return bad(GSBN_ERR_UNSUPPORTED, "gsbn_recall_new: the mfma support needs one delay for all rows");
  r->delay = dmax;
  if (slots.empty()) slots.push_back(0);
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t x) {
    if (e == hipSuccess) e = x;
  };
  const int64_t BN = (int64_t)B * N;
  ok(r->dsup.alloc((size_t)BN));
  ok(r->epsc.alloc((size_t)BN));
  ok(r->count.alloc((size_t)BN));
  ok(r->bj.upload(d->bj, (size_t)N));
  ok(r->out_idx.alloc((size_t)B * H));
  ok(r->out_cnt.alloc((size_t)B * H));
  ok(r->clamp.alloc((size_t)B * H));
  ok(r->wmask.alloc((size_t)B * H));
  if (d->mode == GSBN_RECALL_EXACT) {
    ok(r->wij.upload(d->wij, (size_t)(R * M)));
    ok(r->di.upload(d->di, (size_t)R));
    ok(r->inv_off.upload(off.data(), off.size()));
    ok(r->inv_slot.upload(slots.data(), slots.size()));
    ok(r->spk_cnt.alloc((size_t)RE_RING * B * H));
    ok(r->spk_list.alloc((size_t)RE_RING * B * N));
  } else {
    /* dense W^T [n][src] in bf16: the rows of HCU h(n) whose source is src */
    r->Kp = (N + GM_BK - 1) / GM_BK * GM_BK;
    r->Np = (N + GM_BN - 1) / GM_BN * GM_BN;
    std::vector<double> acc((size_t)r->Np * r->Kp, 0.0);
    for (int h = 0; h < H; ++h)
      for (int k = 0; k < C; ++k) {
        const int s = d->ii[(int64_t)h * C + k];
        if (s < 0) continue;
        for (int j = 0; j < M; ++j)
          acc[(size_t)(h * M + j) * r->Kp + s] += (double)d->wij[((int64_t)h * C + k) * M + j];
      }
    std::vector<uint16_t> wt(acc.size());
    for (size_t i = 0; i < acc.size(); ++i) wt[i] = gsbn_f32_to_bf16((float)acc[i]);
    ok(r->WT.upload(wt.data(), wt.size()));
    r->Bp = (B + GM_BM - 1) / GM_BM * GM_BM;
    ok(r->S.alloc((size_t)RE_RING * r->Bp * r->Kp));
  }
  ok(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
  ok(hipEventCreate(&r->ev0));
  ok(hipEventCreate(&r->ev1));
  if (e != hipSuccess) return bad((int)e, std::string("gsbn_recall_new: ") + hipGetErrorString(e));
  return r;
}

void gsbn_recall_free(gsbn_recall* r) { delete r; }

int gsbn_recall_run(gsbn_recall* r, const int16_t* clamp, const float* wmask, int32_t steps,
                    int32_t count_from, int32_t* out_idx, int32_t* out_cnt, int8_t* raster,
                    double* support_ms) {
  if (!r || !clamp || !wmask || steps < 0 || !out_idx || !out_cnt)
    return fail(GSBN_ERR_ARG, "gsbn_recall_run: invalid arguments");
  const gsbn_recall_desc& d = r->d;
  const int B = d.batch, H = d.H, M = d.M, N = r->N;
  const int64_t BN = (int64_t)B * N;
  hipStream_t s = r->stream;
  hipError_t e = hipMemcpyAsync(r->clamp.p, clamp, (size_t)B * H * 2, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(r->wmask.p, wmask, (size_t)B * H * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return fail((int)e, "gsbn_recall_run: upload");
  /* initial state: dsup = log(1/M) (Pop.cpp:59), E = 0, counts 0, empty queue */
  const float dsup0 = (float)log(1.0 / M);
  hipLaunchKernelGGL(re_fill_f32, dim3(grid_of(BN)), dim3(256), 0, s, r->dsup.p, BN, dsup0);
  hipLaunchKernelGGL(re_fill_f32, dim3(grid_of(BN)), dim3(256), 0, s, r->epsc.p, BN, 0.0f);
  hipLaunchKernelGGL(re_fill_f32, dim3(grid_of(BN)), dim3(256), 0, s,
                     reinterpret_cast<float*>(r->count.p), BN, 0.0f);
  if (d.mode == GSBN_RECALL_EXACT) {
    const int64_t nc = (int64_t)RE_RING * B * H;
    hipLaunchKernelGGL(re_fill_f32, dim3(grid_of(nc)), dim3(256), 0, s,
                       reinterpret_cast<float*>(r->spk_cnt.p), nc, 0.0f);
  } else {
    const int64_t ns = (int64_t)RE_RING * r->Bp * r->Kp / 2;
    hipLaunchKernelGGL(re_fill_f32, dim3(grid_of(ns)), dim3(256), 0, s,
                       reinterpret_cast<float*>(r->S.p), ns, 0.0f);
  }
  if (raster && !r->raster.p) {
    e = r->raster.alloc((size_t)BN);
    if (e != hipSuccess) return fail((int)e, "gsbn_recall_run: raster buffer");
  }
  ReArgs a;
  memset(&a, 0, sizeof(a));
  a.B = B;
  a.Bp = r->Bp;
  a.H = H;
  a.M = M;
  a.C = d.C;
  a.N = N;
  a.Kp = r->Kp;
  a.W = r->W;
  a.dmin = r->dmin;
  a.dmax = r->dmax;
  a.count_from = count_from;
  a.mode = d.mode;
  a.seed = d.seed;
  a.taumdt = d.taumdt;
  a.wtagain = d.wtagain;
  a.maxfqdt = d.maxfqdt;
  a.igain = d.igain;
  a.wgain = d.wgain;
  a.lgbias = d.lgbias;
  a.snoise = d.snoise;
  a.kepsc = d.kepsc;
  a.lg_on = d.lg_on;
  a.lg_off = d.lg_off;
  a.dsup = r->dsup.p;
  a.epsc = r->epsc.p;
  a.count = r->count.p;
  a.bj = r->bj.p;
  a.wij = r->wij.p;
  a.di = r->di.p;
  a.inv_off = r->inv_off.p;
  a.inv_slot = r->inv_slot.p;
  a.spk_cnt = r->spk_cnt.p;
  a.spk_list = r->spk_list.p;
  a.S = r->S.p;
  a.clamp = r->clamp.p;
  a.wmask = r->wmask.p;
  a.raster = raster ? r->raster.p : nullptr;
  const int threads = ((M + 63) / 64) * 64;
  const size_t lds = d.mode == GSBN_RECALL_EXACT ? ((size_t)r->W + RE_ROWS_MAX) * 4 : 0;
  const float beta = 1 - d.kepsc;
  double gemm_ms = 0.0;
  /* the mfma mode's population step: one wave per (HCU, cue) up to 256 units */
  const int upl = (M + 63) / 64;
  const bool wave_step = d.mode == GSBN_RECALL_MFMA && upl <= 4;
  const unsigned wave_grid = (unsigned)(((int64_t)B * H + 3) / 4);
  for (int t = 1; t <= steps; ++t) {
    if (wave_step) {
      if (upl == 1)
        hipLaunchKernelGGL(recall_pop_wave_kernel<1>, dim3(wave_grid), dim3(256), 0, s, a, t);
      else if (upl == 2)
        hipLaunchKernelGGL(recall_pop_wave_kernel<2>, dim3(wave_grid), dim3(256), 0, s, a, t);
      else
        hipLaunchKernelGGL(recall_pop_wave_kernel<4>, dim3(wave_grid), dim3(256), 0, s, a, t);
    } else {
      hipLaunchKernelGGL(recall_step_kernel, dim3(H, B), dim3(threads), lds, s, a, t);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail((int)e, "recall_step_kernel");
    if (d.mode == GSBN_RECALL_MFMA) {
      const int src = t - r->delay - 2;
      if (support_ms) (void)hipEventRecord(r->ev0, s);
      if (src >= 1 && r->delay > 0) {
        const uint16_t* Sp = r->S.p + (int64_t)(src % RE_RING) * r->Bp * r->Kp;
        int rc = gsbn_spike_gemm(B, N, r->Kp, Sp, r->WT.p, r->epsc.p, beta, s);
        if (rc) return fail(rc, "gsbn_spike_gemm");
      } else {
        /* no row can be active yet: the decay alone */
        hipLaunchKernelGGL(re_decay, dim3(grid_of(BN)), dim3(256), 0, s, r->epsc.p, BN, beta);
        if ((e = hipGetLastError()) != hipSuccess) return fail((int)e, "re_decay");
      }
      if (support_ms) {
        (void)hipEventRecord(r->ev1, s);
        (void)hipEventSynchronize(r->ev1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, r->ev0, r->ev1);
        gemm_ms += ms;
      }
    }
    if (raster) {
      e = hipMemcpyAsync(raster + (int64_t)(t - 1) * BN, r->raster.p, (size_t)BN,
                         hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return fail((int)e, "gsbn_recall_run: raster");
    }
  }
  const int64_t BH = (int64_t)B * H;
  hipLaunchKernelGGL(recall_argmax_kernel, dim3(grid_of(BH)), dim3(256), 0, s, B, H, M, r->count.p,
                     r->out_idx.p, r->out_cnt.p);
  e = hipMemcpyAsync(out_idx, r->out_idx.p, (size_t)BH * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_cnt, r->out_cnt.p, (size_t)BH * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail((int)e, std::string("gsbn_recall_run: ") + hipGetErrorString(e));
  if (support_ms) *support_ms = gemm_ms;
  return 0;
}

}  // extern "C"
