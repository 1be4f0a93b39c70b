/*
 * kernels.hip -- CDNA4 (gfx950) kernels of the lazy-update BCPNN step and the
 * kernel-level C ABI (include/gsbn_amd.h).
 *
 * Per executed step the library launches, per population, one fused
 * population kernel (one workgroup per hypercolumn, which also pushes the
 * step's spikes into the delay queues) and, per projection, one row/column
 * kernel over a fixed grid that walks the step's active rows and active
 * columns; or one launch per step instead: the fused step kernel (a
 * hypercolumn's population and row/column update in one workgroup) or the
 * chained step kernel (the previous step's row and column updates with this
 * step's population update).  Batches of networks use the same kernels over
 * workgroup maps.  No kernel allocates memory or needs a host round trip:
 * every list length stays on the device.
 *
 * Determinism: every floating-point result is produced by the same sequence
 * of IEEE operations as the reference's sequential CPU loops
 * (Pop.cpp:229-323, Proj.cpp:209-536).  In particular the support sum epsc is
 * accumulated in ascending row order (Proj.cpp:688-690) by the owning
 * hypercolumn's workgroup at the start of the next population update, the
 * WTA normaliser is summed in ascending unit order (Pop.cpp:282-292), and
 * exp/log come from detmath.h.  The file must be compiled with
 * -ffp-contract=off.
 */
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <limits.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/gsbn_amd.h"
#include "detmath.h"
#include "gsbn_core.h"
#include "philox.h"
#include "wave.h"

static_assert(sizeof(gsbn_syn_rec) == 32, "synapse record must be 32 bytes");
static_assert(sizeof(gsbn_syn_t) == sizeof(gsbn_syn_rec), "record layouts differ");
static_assert(sizeof(gsbn_row_rec) == 16, "row record must be 16 bytes");
static_assert(sizeof(gsbn_syn16_rec) == 16 && sizeof(gsbn_syn16_t) == 16,
              "16-bit synapse record must be 16 bytes");

#define POP_MAX_THREADS 1024
#define RC_THREADS 256
#define RC_COLS 4 /* column slots per thread and pass */
#define GSBN_MAX_DEFER 64
#define GSBN_OVF_HELPERS 4096 /* helper waves of a row/column launch with a row overflow queue */
#define GSBN_OVF_QUEUES 64    /* its sub-queues, one counter each (64-byte stride) */

/* the overflow sub-queue of hypercolumn h of batch projection p, and the
 * word of its counter for a step's parity */
__device__ __forceinline__ int ovf_queue(int p, int h) {
  uint32_t x = ((uint32_t)p * 0x9E3779B1u) ^ ((uint32_t)h * 0x85EBCA6Bu);
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  return (int)(x >> 26);
}
__device__ __forceinline__ int ovf_counter(int exec, int q) {
  return ((exec & 1) * GSBN_OVF_QUEUES + q) * 16;
} /* largest ring of deferred column kicks per unit */
#define ALL_THREADS 256
#define ALL_ROWS 16

namespace {
/* One scalar load from every 64-byte line of the first BYTES bytes of the
 * kernel's argument segment, waited for together.  The argument structs are
 * larger than the scalar registers, so the compiler reads them in several
 * dependent rounds of scalar loads; each round that touches a line not yet in
 * the scalar cache costs a trip to L2 or memory.  Touching every line first
 * makes those trips overlap in one round; the later rounds hit. */
template <int BYTES>
__device__ __forceinline__ void touch_kernarg_lines() {
  const uint32_t* k = (const uint32_t*)__builtin_amdgcn_kernarg_segment_ptr();
  uint32_t acc = 0;
#pragma unroll
  for (int o = 0; o < BYTES / 4; o += 16) acc ^= __builtin_nontemporal_load(k + o);
  asm volatile("" ::"s"(acc));
}

/* The same for an argument struct in device memory (a batched launch's
 * per-network arguments; the address must be uniform across the wave). */
template <int BYTES>
__device__ __forceinline__ void touch_lines(const void* p) {
  const uint32_t* k = static_cast<const uint32_t*>(p);
  uint32_t acc = 0;
#pragma unroll
  for (int o = 0; o < BYTES / 4; o += 16) acc ^= __builtin_nontemporal_load(k + o);
  asm volatile("" ::"s"(acc));
}
}  // namespace

namespace gsbn {
/* events for the next timed launch: recorded by the kernel's own dispatch
 * (hipExtLaunchKernelGGL), so that their difference is the kernel's time
 * without the gaps that separate event packets add */
thread_local hipEvent_t tl_ev_start = nullptr, tl_ev_stop = nullptr;
void set_launch_events(hipEvent_t a, hipEvent_t b) {
  tl_ev_start = a;
  tl_ev_stop = b;
}
bool launch_events_pending() { return tl_ev_start != nullptr; }
}  // namespace gsbn

namespace {

inline int ret(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

/* launch that takes the pending timing events, if any */
template <typename K, typename... Args>
inline void launch_timed(K kernel, dim3 g, dim3 b, size_t sh, hipStream_t s, Args... args) {
  if (gsbn::tl_ev_start) {
    hipEvent_t e0 = gsbn::tl_ev_start, e1 = gsbn::tl_ev_stop;
    gsbn::tl_ev_start = gsbn::tl_ev_stop = nullptr;
    hipExtLaunchKernelGGL(kernel, g, b, sh, s, e0, e1, 0, args...);
  } else {
    hipLaunchKernelGGL(kernel, g, b, sh, s, args...);
  }
}

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return ret(e);
}

/* grid of a looping batch launch over n items.  One workgroup per item:
 * capping the grid at one dispatch round (each workgroup then loops over
 * several items) was measured slower on the 512 networks (population and
 * row/column launches 1.5x longer), so the loops run once. */
template <typename K>
unsigned loop_grid(K kernel, int threads, size_t lds, int64_t n) {
  (void)kernel;
  (void)threads;
  (void)lds;
  return (unsigned)n;
}

/* exclusive prefix sum across the workgroup (blockDim multiple of 64);
 * sh needs blockDim/64 + 1 ints */
__device__ int block_excl_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int x = (int)wave_incl_scan((uint32_t)v);
  (void)lane;
  if ((threadIdx.x & 63) == 63) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < nw; ++w) {
      int t = sh[w];
      sh[w] = acc;
      acc += t;
    }
    sh[nw] = acc;
  }
  __syncthreads();
  int res = sh[wid] + x - v;
  *total = sh[nw];
  __syncthreads();
  return res;
}

/* the same with a wave shuffle scan and two barriers (no serial section) */
__device__ __forceinline__ int block_scan_fast(int v, int* sh, int* total) {
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

/* Exclusive workgroup prefix sums of three 32-bit values at once (each may
 * pack two 16-bit counts whose workgroup totals stay below 2^16, so the
 * fields add without carries); the totals go to tot.  sh: 3 words per wave,
 * used by this one scan of the launch (no barrier after it). */
__device__ __forceinline__ void block_scan3(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t* sh,
                                            uint32_t* tot) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const uint32_t x = wave_incl_scan(a), y = wave_incl_scan(b), z = wave_incl_scan(c);
  if (lane == 63) {
    sh[3 * wid] = x;
    sh[3 * wid + 1] = y;
    sh[3 * wid + 2] = z;
  }
  __syncthreads();
  uint32_t bx = 0, by = 0, bz = 0, tx = 0, ty = 0, tz = 0;
  for (int w = 0; w < nw; ++w) {
    const uint32_t u = sh[3 * w], v = sh[3 * w + 1], q = sh[3 * w + 2];
    if (w < wid) {
      bx += u;
      by += v;
      bz += q;
    }
    tx += u;
    ty += v;
    tz += q;
  }
  a = bx + x - a;
  b = by + y - b;
  c = bz + z - c;
  tot[0] = tx;
  tot[1] = ty;
  tot[2] = tz;
}

/* epsc of unit (h,j) for incoming projection m after adding the pending row
 * contributions of the previous step, in ascending row order */
__device__ float finalized_epsc(const gsbn_pop_args& a, const gsbn_pop_proj& p, int m, int h,
                                int j, int cnt) {
  const int64_t N = (int64_t)a.H * a.M;
  float e = a.epsc[m * N + (int64_t)h * a.M + j];
  const float* c = p.contrib + ((int64_t)h * p.C) * a.M + j;
  for (int k = 0; k < cnt; ++k) {
    float v = c[(int64_t)k * a.M];
    if (gsbn_fbits(v) != GSBN_NO_CONTRIB_BITS) e = e + v;
  }
  return e;
}

/* queue push of the rows fwd_row[q0 .. q1) of a unit spiking at exec step
 * s: they become active at s + di + 2; threads tid, tid+nt, ... take them */
__device__ __forceinline__ void push_rows(const gsbn_fanout& f, int q0, int q1, int s, int tid,
                                          int nt) {
  for (int q = q0 + tid; q < q1; q += nt) {
    const int r = f.fwd_row[q];
    const int d = f.fwd_del[q];
    const int hh = r / f.C;
    const int k = r - hh * f.C;
    const int slot = (s + d + 2) % GSBN_FUT;
    atomicOr(f.fut + ((int64_t)slot * f.H + hh) * f.W + (k >> 5), 1u << (k & 31));
  }
}

/* the same for source unit n */
__device__ __forceinline__ void push_unit(const gsbn_fanout& f, int n, int s, int tid, int nt) {
  const int q1 = f.fwd_off[n + 1];
  for (int q = f.fwd_off[n] + tid; q < q1; q += nt) {
    const int r = f.fwd_row[q];
    const int d = f.fwd_del[q];
    const int hh = r / f.C;
    const int k = r - hh * f.C;
    const int slot = (s + d + 2) % GSBN_FUT;
    atomicOr(f.fut + ((int64_t)slot * f.H + hh) * f.W + (k >> 5), 1u << (k & 31));
  }
}

/* support integration of one unit (update_sup_kernel_1_cpu, Pop.cpp:229-261);
 * returns the new dsup and the draws */
__device__ float unit_dsup(const gsbn_pop_args& a, int h, int j, const float* efin, float* u,
                           float* xi) {
  const int64_t N = (int64_t)a.H * a.M;
  const int idx = h * a.M + j;
  float z;
  gsbn_draw_unit(a.seed, (uint32_t)a.pop_id, (uint32_t)a.simstep, (uint32_t)idx, u, &z);
  *xi = a.snoise * z;
  float wsup = 0.0f;
  for (int m = 0; m < a.nproj_in; ++m) wsup += a.bj[m * N + idx] + efin[m];
  float sup = a.lgbias + a.igain * a.lginp[idx] + *xi;
  sup += (a.wgain * a.wmask[h]) * wsup;
  float dsup = a.dsup[idx];
  return dsup + (sup - dsup) * a.taumdt;
}

/* dsup of the population's first unit for the CPU-quirk WTA maximum */
__global__ void pop_dsup0_kernel(gsbn_pop_args a) {
  if (threadIdx.x != 0) return;
  float efin[GSBN_MAX_PROJ_IN];
  for (int m = 0; m < a.nproj_in; ++m) efin[m] = finalized_epsc(a, a.projs[m], m, 0, 0, a.projs[m].pend_cnt[0]);
  float u, xi;
  a.scratch_dsup0[0] = unit_dsup(a, 0, 0, efin, &u, &xi);
}

/* the seven synapse coefficients of gap D: two 16-byte loads (entries are
 * 48 bytes, 16-byte aligned) */
__device__ __forceinline__ gsbn_syn_coef_t load_coef(const float* table, int len, int D) {
  const float* c = gsbn_coef(table, len, D);
  const float4 x = *reinterpret_cast<const float4*>(c);
  const float4 y = *reinterpret_cast<const float4*>(c + 4);
  gsbn_syn_coef_t k;
  k.ezi = x.x;
  k.ezj = x.y;
  k.ee = x.z;
  k.ep = x.w;
  k.cez = y.x;
  k.cpe = y.y;
  k.cpz = y.z;
  return k;
}

/* ---- synapse state access for the FP32 record and the 16-bit forms ---- */
struct SynVals {
  float pij, eij, zi2, zj2, wij;
  int tij;
};

template <int PREC>
struct SynIO {
  gsbn_syn_rec* syn;
  gsbn_syn16_rec* syn16;
  float* wij;
  float ps, es, pinv, einv;

  template <class A>
  __device__ __forceinline__ explicit SynIO(const A& a)
      : syn(a.syn), syn16(a.syn16), wij(a.wij), ps(a.p_scale), es(a.e_scale) {
    pinv = PREC == GSBN_PREC_F16 ? 1.0f / ps : 1.0f;
    einv = PREC == GSBN_PREC_F16 ? 1.0f / es : 1.0f;
  }

  __device__ __forceinline__ void unpack16(const uint4 v, SynVals& s) const {
    s.pij = gsbn_unpack_trace(PREC, (uint16_t)(v.x & 0xffffu), pinv);
    s.eij = gsbn_unpack_trace(PREC, (uint16_t)(v.x >> 16), einv);
    s.zi2 = __uint_as_float(v.y);
    s.zj2 = __uint_as_float(v.z);
    s.tij = (int)v.w;
  }
  __device__ __forceinline__ uint4 pack16(const SynVals& s) const {
    const uint32_t p = gsbn_pack_trace(PREC, s.pij, ps);
    const uint32_t e = gsbn_pack_trace(PREC, s.eij, es);
    return make_uint4(p | (e << 16), __float_as_uint(s.zi2), __float_as_uint(s.zj2), (uint32_t)s.tij);
  }

  /* whole state; the weight is read only when need_w (16-bit forms) */
  __device__ __forceinline__ SynVals load(int64_t i, bool need_w) const {
    SynVals s;
    if constexpr (PREC == GSBN_PREC_F32) {
      const gsbn_syn_rec r = syn[i];
      s.pij = r.pij; s.eij = r.eij; s.zi2 = r.zi2; s.zj2 = r.zj2; s.wij = r.wij; s.tij = r.tij;
    } else {
      unpack16(*reinterpret_cast<const uint4*>(syn16 + i), s);
      s.wij = need_w ? wij[i] : 0.0f;
    }
    return s;
  }
  /* whole state; the weight is written only when w_new (16-bit forms) */
  __device__ __forceinline__ void store(int64_t i, const SynVals& s, bool w_new) const {
    if constexpr (PREC == GSBN_PREC_F32) {
      gsbn_syn_rec r;
      r.pij = s.pij; r.eij = s.eij; r.zi2 = s.zi2; r.zj2 = s.zj2; r.wij = s.wij; r.tij = s.tij;
      r.pad0 = 0; r.pad1 = 0;
      syn[i] = r;
    } else {
      *reinterpret_cast<uint4*>(syn16 + i) = pack16(s);
      if (w_new) wij[i] = s.wij;
    }
  }
  /* the traces and tij only (column updates never touch the weight) */
  __device__ __forceinline__ SynVals load_traces(int64_t i) const {
    SynVals s;
    if constexpr (PREC == GSBN_PREC_F32) {
      const float4 v = *reinterpret_cast<const float4*>(syn + i);
      s.pij = v.x; s.eij = v.y; s.zi2 = v.z; s.zj2 = v.w; s.wij = 0.0f;
      s.tij = syn[i].tij;
    } else {
      unpack16(*reinterpret_cast<const uint4*>(syn16 + i), s);
      s.wij = 0.0f;
    }
    return s;
  }
  __device__ __forceinline__ void store_traces(int64_t i, const SynVals& s) const {
    if constexpr (PREC == GSBN_PREC_F32) {
      *reinterpret_cast<float4*>(syn + i) = make_float4(s.pij, s.eij, s.zi2, s.zj2);
      syn[i].tij = s.tij;
    } else {
      *reinterpret_cast<uint4*>(syn16 + i) = pack16(s);
    }
  }
  /* the values a store and a load would leave (16-bit forms round Pij and
   * Eij to their storage precision) */
  __device__ __forceinline__ void round_traces(SynVals& s) const {
    if constexpr (PREC != GSBN_PREC_F32) {
      s.pij = gsbn_unpack_trace(PREC, (uint16_t)gsbn_pack_trace(PREC, s.pij, ps), pinv);
      s.eij = gsbn_unpack_trace(PREC, (uint16_t)gsbn_pack_trace(PREC, s.eij, es), einv);
    }
  }
  __device__ __forceinline__ void store_tij(int64_t i, int t) const {
    if constexpr (PREC == GSBN_PREC_F32)
      syn[i].tij = t;
    else
      syn16[i].tij = t;
  }
};

/* Deferred column updates: the pending column kicks of one synapse.  ring:
 * unit (h, j)'s ring of K spike steps (K a power of two), T: the unit's
 * spike count the ring is read at.  Applies, oldest first, the steps s with
 * v.tij < s <= upto among the last K - 1 entries, each as the reference's
 * column update of step s (Proj.cpp:481-536): catch up over s - tij, zj2 +=
 * kftj, tij = s, and (16-bit forms) Pij and Eij rounded as the store
 * between two updates rounds them.  Returns whether a kick was applied. */
template <int PREC>
__device__ __forceinline__ bool replay_kicks(const SynIO<PREC>& io, SynVals& v,
                                             const int32_t* ring, int T, int K, int upto,
                                             const float* table, int table_len, float kftj) {
  const int lo = T - K + 1 > 0 ? T - K + 1 : 0;
  int n1 = T - 1;
  while (n1 >= lo && ring[n1 & (K - 1)] > upto) --n1;
  int n0 = n1 + 1;
  while (n0 > lo && ring[(n0 - 1) & (K - 1)] > v.tij) --n0;
  for (int n = n0; n <= n1; ++n) {
    const int sn = ring[n & (K - 1)];
    gsbn_syn_catchup_k(load_coef(table, table_len, sn - v.tij), &v.pij, &v.eij, &v.zi2, &v.zj2);
    v.zj2 = v.zj2 + kftj;
    v.tij = sn;
    io.round_traces(v);
  }
  return n0 <= n1;
}

/* The same for a ring of K = 8, held in registers: the ring is read with two
 * 16-byte loads that need no earlier result (issued beside the synapse
 * record), the entries are put in spike order by register selects, and the
 * coefficients of each kick are loaded while the previous kick is applied,
 * instead of one dependent load per ring entry and per kick.
 * s[q] is the step of spike number T - 7 + q (q = 0..6: the last K - 1
 * spikes); numbers below 0 hold INT_MIN and are never applied. */
struct Ring8 {
  int s[7];
};

__device__ __forceinline__ Ring8 ring8(const int4 lo, const int4 hi, const int T) {
  const int e[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  Ring8 r;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int n = T - 7 + q;
    const int idx = n & 7;
    int v = e[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) v = idx == i ? e[i] : v;
    r.s[q] = n >= 0 ? v : INT_MIN;
  }
  return r;
}

__device__ __forceinline__ Ring8 ring8_load(const int32_t* ring, const int T) {
  const int4* p = reinterpret_cast<const int4*>(ring);
  return ring8(p[0], p[1], T);
}

/* replay_kicks for K = 8 from registers: applies the entries with
 * v.tij < s <= upto in spike order, each as the reference's column update
 * of step s.  With own != nullptr, also loads (first, beside the first
 * kick's) the coefficients of the gap from the last applied kick (or tij)
 * to own_t, when that gap is positive. */
template <int PREC>
__device__ __forceinline__ bool replay_ring8(const SynIO<PREC>& io, SynVals& v, const Ring8& r,
                                             int upto, const float* table, int table_len,
                                             float kftj, int own_t = 0,
                                             gsbn_syn_coef_t* own = nullptr) {
  bool sel[7];
  int tend = v.tij;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    sel[q] = r.s[q] > v.tij && r.s[q] <= upto;
    if (sel[q]) tend = r.s[q];
  }
  if (own && own_t - tend > 0) *own = load_coef(table, table_len, own_t - tend);
  /* the selected entries are consecutive (the ring's steps ascend) */
  gsbn_syn_coef_t cur;
  if (sel[0]) cur = load_coef(table, table_len, r.s[0] - v.tij);
  bool any = false;
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const bool next = q + 1 < 7 && sel[q + 1 < 7 ? q + 1 : 6];
    if (sel[q]) {
      gsbn_syn_coef_t nxt;
      if (next) nxt = load_coef(table, table_len, r.s[q + 1 < 7 ? q + 1 : 6] - r.s[q]);
      gsbn_syn_catchup_k(cur, &v.pij, &v.eij, &v.zi2, &v.zj2);
      v.zj2 = v.zj2 + kftj;
      v.tij = r.s[q];
      io.round_traces(v);
      any = true;
      if (next) cur = nxt;
    } else if (next) {
      cur = load_coef(table, table_len, r.s[q + 1 < 7 ? q + 1 : 6] - v.tij);
    }
  }
  return any;
}

/* per-launch values of a row/column update (taken from the argument struct
 * in the single form, from the step arguments in the batched form) */
struct RcStep {
  int t, exec, P, table_len;
  float kp;
  const float* table;
};

/* What the fused step's row/column phase takes from its population phase
 * (all in LDS): the spike flags, the post-jxx pj and the spike list of the
 * hypercolumn, and the active-row bitmap and list of its one incoming
 * projection. */
struct FusedLds {
  const int* sj;       /* [M] */
  const float* pj;     /* [M] */
  const int* spk;      /* [nspk] */
  int nspk;
  const uint32_t* bitmap;
  const int* rows;     /* [nrows], ascending */
  int nrows;
};

/* Chained step: the previous executed step's row updates of hypercolumn h
 * (the row pass of rowcol_body with the same arithmetic), done by the
 * workgroup of this step's population update.  Thread j (< M) updates
 * synapse j of every active row in ascending row order and adds the row's
 * contribution to e, the support sum of unit j, so the sum is the one the
 * population update would form from memory (ascending row order).  kpre:
 * the first four entries of the row list, read ahead; sjp: unit j's spike
 * of the previous step; pjv: its pj after that step's update.  Thread 0
 * writes each batch's row records once every thread has read them.  Called
 * by every thread of the workgroup (nrows is uniform). */
template <int PREC>
__device__ __forceinline__ float chained_rows(const gsbn_rowcol_args& b, const RcStep& q, int h,
                                              int j, bool live, int nrows, const int* kpre,
                                              int sjp, float pjv, float e) {
  const SynIO<PREC> io(b);
  const int M = b.M, C = b.C;
  const bool learn = q.kp != 0.0f;
  const int t = q.t;
  const int32_t* rl = b.rows_list + (int64_t)h * C;
  for (int base = 0; base < nrows; base += 4) {
    int k[4];
    gsbn_row_rec st[4];
    SynVals sv[4];
    float cr[4][GSBN_COEF_STRIDE];
    gsbn_syn_coef_t cs[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) k[u] = base == 0 ? kpre[u] : (base + u < nrows ? rl[base + u] : 0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (base + u < nrows) {
        const int64_t r = (int64_t)h * C + k[u];
        st[u] = b.rowst[r];
        if (live) sv[u] = io.load(r * M + j, !learn);
      }
    }
    if (b.dK && live) {
      /* deferred column updates: this synapse's pending kicks first */
      const int uj = h * M + j;
      const int32_t* ring = b.dhist + (int64_t)uj * b.dK;
      const int T = b.dcnt[uj];
      for (int u = 0; u < 4; ++u)
        if (base + u < nrows)
          replay_kicks<PREC>(io, sv[u], ring, T, b.dK, t - 1, q.table, q.table_len, b.kftj);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (base + u < nrows) {
        const int Dr = t - st[u].ti;
        if (Dr > 0) {
          const float4* c4 =
              reinterpret_cast<const float4*>(gsbn_coef(q.table, q.table_len, Dr));
#pragma unroll
          for (int w = 0; w < GSBN_COEF_STRIDE / 4; ++w) {
            const float4 x = c4[w];
            cr[u][4 * w] = x.x;
            cr[u][4 * w + 1] = x.y;
            cr[u][4 * w + 2] = x.z;
            cr[u][4 * w + 3] = x.w;
          }
        }
        if (live) {
          const int D = t - sv[u].tij;
          if (D > 0) cs[u] = load_coef(q.table, q.table_len, D);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (base + u >= nrows) break;
      const int64_t r = (int64_t)h * C + k[u];
      const int Dr = t - st[u].ti;
      if (Dr <= 0) {
        st[u].zi += b.kfti;
        st[u].ti = t;
      } else {
        gsbn_row_catchup(cr[u], &st[u].pi, &st[u].ei, &st[u].zi);
        st[u].zi = st[u].zi + b.kfti;
        st[u].ti = t;
      }
      if (live) {
        SynVals& s = sv[u];
        float cval;
        bool w_new = false;
        const int D = t - s.tij;
        if (D <= 0) {
          s.zi2 += b.kfti;
          s.tij = t;
          cval = gsbn_bitsf(GSBN_NO_CONTRIB_BITS);
        } else {
          gsbn_syn_catchup_k(cs[u], &s.pij, &s.eij, &s.zi2, &s.zj2);
          s.zi2 = s.zi2 + b.kfti;
          s.tij = t;
          if (learn) {
            s.wij = gsbn_wij(s.pij, st[u].pi, pjv, b.wgain, b.eps, b.eps2);
            w_new = true;
          }
          cval = s.wij;
        }
        if (sjp > 0) s.zj2 += b.kftj;
        io.store(r * M + j, s, w_new);
        if (gsbn_fbits(cval) != GSBN_NO_CONTRIB_BITS) e = e + cval;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (base + u < nrows) b.rowst[(int64_t)h * C + k[u]] = st[u];
    }
  }
  return e;
}

/* The per-step values of a population update: the argument struct's own
 * fields for a single launch, the batch's shared step fields (with the
 * stimulus and mask rows offset) for a batched one, so that a batched
 * workgroup can read its population's arguments in place, from device
 * memory, instead of copying them into registers. */
struct PopStep {
  int simstep, exec, do_jxx;
  float prn;
  const float* lginp;
  const float* wmask;
  int ovf_keep, ovf_cap; /* the row overflow queue (gsbn_step_args) */
  int32_t* ovf_cnt;
  int32_t* ovf_task;
};

__device__ __forceinline__ PopStep pop_step(const gsbn_pop_args& a) {
  return {a.simstep, a.exec, a.do_jxx, a.prn, a.lginp, a.wmask, 0, 0, nullptr, nullptr};
}

__device__ __forceinline__ PopStep pop_step(const gsbn_pop_args& a, const gsbn_step_args& st) {
  return {st.simstep, st.exec, st.do_jxx, st.prn, a.lginp + st.lginp_offset,
          a.wmask + st.wmask_offset, st.ovf_keep, st.ovf_cap, st.ovf_cnt, st.ovf_task};
}

/* One hypercolumn's population update.  FUSED: `f` also receives what the
 * fused step's row/column phase needs (one incoming projection).
 *
 * Memory order: every load whose address does not depend on this step's
 * results is issued in one batch at the start (the first projection's
 * pointers come from the arguments, and its first four pending support
 * contributions are read before their count is known), and every global
 * store of the update is issued after the last load: loads and stores share
 * one in-order counter, so a load issued behind stores would wait for them. */
template <bool FUSED, int CPREC = -1, int NPRE = 4, bool SPEC = true>
__device__ __forceinline__ void pop_body(const gsbn_pop_args& a, const PopStep& ps, const int h,
                                         FusedLds* f = nullptr,
                                         const gsbn_rowcol_args* rb = nullptr,
                                         const RcStep* rq = nullptr, const int has_prev = 0) {
  /* dynamic LDS: three arrays of blockDim entries, then the active-row
   * bitmap (and, FUSED, the active-row list); see pop_lds() */
  extern __shared__ __attribute__((aligned(16))) uint32_t sh_dyn[];
  float* const sh_f = reinterpret_cast<float*>(sh_dyn);
  int* const sh_ext = reinterpret_cast<int*>(sh_dyn + blockDim.x);
  int* const sh_spk = reinterpret_cast<int*>(sh_dyn + 2 * blockDim.x);
  uint32_t* const sh_bitmap = sh_dyn + 3 * blockDim.x;
  __shared__ int sh_scan[POP_MAX_THREADS / 64 + 1];
  __shared__ uint32_t sh_scan3[3 * (POP_MAX_THREADS / 64)];
  __shared__ int sh_pend[GSBN_MAX_PROJ_IN];
  __shared__ float sh_wmax[POP_MAX_THREADS / 64]; /* per-wave WTA maxima */
  __shared__ float sh_vsum;
  __shared__ int sh_sj[FUSED ? RC_THREADS : 1];
  __shared__ float sh_pj[FUSED ? RC_THREADS : 1];

  const int j = threadIdx.x;
  const int M = a.M;
  const bool live = j < M;
  const int idx = h * M + j;
  const int64_t N = (int64_t)a.H * M;
  const int nin = a.nproj_in;
  const int bd = blockDim.x;

  /* 0. loads that do not depend on this step's results */
  float dsup_old = 0.0f, lg = 0.0f, e0 = 0.0f, bj0 = 0.0f, pj0 = 0.0f, ej0 = 0.0f, zj0 = 0.0f;
  /* pending contributions read before their count is known (NPRE of them:
   * four for a single network, whose busy steps have several rows pending
   * per hypercolumn; one for batches, whose hypercolumns mostly have none,
   * where more speculative reads cost more than the round trip they save;
   * the rest is read eight rows per round trip below) */
  float cpre[NPRE];
#pragma unroll
  for (int q = 0; q < NPRE; ++q) cpre[q] = 0.0f;
  int dcnt0 = 0, dfl0 = 0; /* deferred column updates: this unit's ring state */
  int32_t counter_old = 0;
  uint32_t fut0 = 0u, fut1 = 0u, fut2 = 0u, fut3 = 0u; /* words j + i * blockDim of the ring slot */
  int pend0 = 0;
  const float wmv = ps.wmask[h];
  if (live) {
    dsup_old = a.dsup[idx];
    lg = ps.lginp[idx];
    counter_old = a.counter[idx];
  }
  /* fan-out range of this unit in the first outgoing projection (the queue
   * push of step 5b).  SPEC: read now, for every unit; otherwise read by the
   * spiking units once their spike is known, as is their deferred ring
   * state (measured faster for the single and the batched launches: few
   * units spike in a step) */
  int fo0 = 0, fo1 = 0;
  if (SPEC && a.nproj_out > 0 && live) {
    fo0 = a.out0.fwd_off[idx];
    fo1 = a.out0.fwd_off[idx + 1];
  }
  if (nin > 0) {
    const gsbn_pop_proj& p0 = a.proj0;
    pend0 = p0.pend_cnt[h];
    if (live) {
      e0 = a.epsc[idx];
      bj0 = a.bj[idx];
      pj0 = p0.pj[idx];
      ej0 = p0.ej[idx];
      zj0 = p0.zj[idx];
      /* rows 0..NPRE-1 of the hypercolumn's contribution block (inside the
       * block whatever the count; entries at or beyond the count are not
       * used) */
      const float* c = p0.contrib + ((int64_t)h * p0.C) * M + j;
#pragma unroll
      for (int q = 0; q < NPRE; ++q)
        if (p0.C > q) cpre[q] = c[(int64_t)q * M];
      if (SPEC && p0.dK) {
        dcnt0 = p0.dcnt[idx];
        dfl0 = p0.dfl[idx];
      }
    }
    const uint32_t* wslot = p0.fut + ((int64_t)(ps.exec % GSBN_FUT) * a.H + h) * p0.W;
    if (j < p0.W) fut0 = wslot[j];
    if (j + bd < p0.W) fut1 = wslot[j + bd];
    if (j + 2 * bd < p0.W) fut2 = wslot[j + 2 * bd];
    if (j + 3 * bd < p0.W) fut3 = wslot[j + 3 * bd];
  }
  /* chained step: the previous step's active rows (count, first four
   * entries) and this unit's spike of that step */
  const bool chain = CPREC >= 0 && has_prev;
  int nrp = 0, sjp = 0;
  int kpre[4] = {0, 0, 0, 0};
  if (CPREC >= 0 && has_prev) {
    const int Cp = rb->C;
    nrp = rb->rows_cnt[h];
    const int32_t* rl = rb->rows_list + (int64_t)h * Cp;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (u < Cp) kpre[u] = rl[u];
    if (live) sjp = rb->sj[idx];
  }
  const float d0 = a.wta_pop_max ? a.scratch_dsup0[0] : 0.0f;
  const bool ext_on = a.ext_off && ps.simstep >= 0 && ps.simstep < a.ext_nsteps;
  int ext_b = 0, ext_e = 0;
  if (ext_on) {
    ext_b = a.ext_off[ps.simstep];
    ext_e = a.ext_off[ps.simstep + 1];
  }
  /* the draws and the first projection's postsynaptic traces do not depend
   * on this step's results either: computed while the loads are in flight */
  float u = 0.0f, xi = 0.0f;
  if (live) {
    float z;
    gsbn_draw_unit(a.seed, (uint32_t)a.pop_id, (uint32_t)ps.simstep, (uint32_t)idx, &u, &z);
    xi = a.snoise * z;
  }
  float pj_n = 0.0f, ej_n = 0.0f, zj_n = 0.0f, bj_n = 0.0f;
  bool bj_w = false;
  if (ps.do_jxx && live && nin > 0) {
    /* update_jxx_kernel_cpu (Proj.cpp:310-363) for projection 0; the spike
     * increment of zj follows in step 6 */
    const gsbn_pop_proj& p0 = a.proj0;
    const float kp = p0.taupdt * ps.prn;
    pj_n = pj0 + (ej0 - pj0) * kp;
    ej_n = ej0 + (zj0 - ej0) * p0.ke;
    zj_n = zj0 * (1 - p0.kzj);
    bj_w = kp != 0.0f;
    if (bj_w) bj_n = p0.bgain * gsbn_logf(pj_n + p0.eps);
  }
  if (j > 0 && j < nin) sh_pend[j] = a.projs[j].pend_cnt[h];
  __syncthreads();
  /* the first external spike of each thread (its address needs ext_b) */
  int ext_u0 = -1;
  if (ext_on && ext_b + j < ext_e) ext_u0 = a.ext_units[ext_b + j];

  /* 1. pending support contributions of the previous step's row updates,
   * added in ascending row order (chained: computed here from the rows) */
  float efin[GSBN_MAX_PROJ_IN];
  if constexpr (CPREC >= 0) {
    if (chain) efin[0] = chained_rows<CPREC>(*rb, *rq, h, j, live, nrp, kpre, sjp, pj0, e0);
  }
  for (int m = chain ? 1 : 0; m < nin; ++m) {
    if (!live) {
      efin[m] = 0.0f;
      continue;
    }
    const gsbn_pop_proj& p = m == 0 ? a.proj0 : a.projs[m];
    const int cnt = m == 0 ? pend0 : sh_pend[m];
    float e = m == 0 ? e0 : a.epsc[m * N + idx];
    const float* c = p.contrib + ((int64_t)h * p.C) * M + j;
    int k = 0;
    if (m == 0) {
#pragma unroll
      for (int q = 0; q < NPRE; ++q)
        if (q < cnt && gsbn_fbits(cpre[q]) != GSBN_NO_CONTRIB_BITS) e = e + cpre[q];
      k = NPRE;
    }
    for (; k < cnt; k += 8) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = k + q < cnt ? c[(int64_t)(k + q) * M] : 0.0f;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (k + q < cnt && gsbn_fbits(v[q]) != GSBN_NO_CONTRIB_BITS) e = e + v[q];
    }
    efin[m] = e;
  }

  /* 2. support, dsup (update_sup_kernel_1_cpu, Pop.cpp:229-261) */
  float dsup = 0.0f;
  if (live) {
    float wsup = 0.0f;
    for (int m = 0; m < nin; ++m) wsup += (m == 0 ? bj0 : a.bj[m * N + idx]) + efin[m];
    float sup = a.lgbias + a.igain * lg + xi;
    sup += (a.wgain * wmv) * wsup;
    dsup = dsup_old + (sup - dsup_old) * a.taumdt;
  }

  /* 3. WTA maximum (update_sup_kernel_2_cpu, Pop.cpp:270-279): the
   * reference's sequential `if (d > max) max = d` keeps the first of equal
   * maxima; combining contiguous index ranges left to right with the same
   * comparison gives the same value (+-0 ties included, for NaN-free
   * dsup), here within each wave (wave_max_left), then across waves in
   * order */
  {
    const float mv = wave_max_left(live ? dsup : -INFINITY);
    if ((j & 63) == 0) sh_wmax[j >> 6] = mv;
  }
  __syncthreads();
  float maxdsup = sh_wmax[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
    const float o = sh_wmax[w];
    if (o > maxdsup) maxdsup = o;
  }
  if (a.wta_pop_max) {
    float mx = d0;
    if (maxdsup > mx) mx = maxdsup;
    maxdsup = mx;
  }

  /* soft-max (Pop.cpp:280-302); the normaliser is summed in ascending unit
   * order by one thread, four LDS words per read */
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
  const float vsum = sh_vsum;
  if (live && vsum > 1.0f) act /= vsum;

  /* 4. spikes and counters (Pop.cpp:305-323); external override (Pop.cpp:389-409) */
  int spk = 0;
  if (live) spk = (u < act * a.maxfqdt) ? 1 : 0;
  const int32_t counter_new = counter_old + spk;
  if (a.ext_off) {
    sh_ext[j] = 0;
    __syncthreads();
    for (int t = ext_b + j; t < ext_e; t += bd) {
      const int unit = t == ext_b + j ? ext_u0 : a.ext_units[t];
      if (unit >= h * M && unit < h * M + M) sh_ext[unit - h * M] = 1;
    }
    __syncthreads();
    if (live) spk = sh_ext[j];
  }
  if (FUSED && live) sh_sj[j] = spk;

  if (!SPEC && live && spk) {
    if (a.nproj_out > 0) {
      fo0 = a.out0.fwd_off[idx];
      fo1 = a.out0.fwd_off[idx + 1];
    }
    if (nin > 0 && a.proj0.dK) {
      dcnt0 = a.proj0.dcnt[idx];
      dfl0 = a.proj0.dfl[idx];
    }
  }

  /* 5. spiking units of this hypercolumn, ascending (.ssj, Proj.cpp:632-643).
   * Without the fused caller, and when the first projection's ring slot is
   * in the words read at the start, one scan also places this step's
   * deferred column walks (step 6b) and its active rows (step 7): packed
   * counts of spikes | walks, and of set bits in ring words j + i*bd. */
  int spk_tot;
  int pos_s = 0;
  const bool fast7 = !FUSED && nin > 0 && a.proj0.W <= 4 * bd;
  int wflag = 0, pos_w = 0, walk_tot = 0;
  uint32_t band_pre[4] = {0u, 0u, 0u, 0u}, band_tot[4] = {0u, 0u, 0u, 0u};
  if (fast7) {
    if (a.proj0.dK && live && spk && dcnt0 + 1 - dfl0 >= a.proj0.dK - 1) wflag = 1;
    uint32_t sa = (uint32_t)(live ? spk : 0) | ((uint32_t)wflag << 16);
    uint32_t sb = (uint32_t)__popc(fut0) | ((uint32_t)__popc(fut1) << 16);
    uint32_t sc = (uint32_t)__popc(fut2) | ((uint32_t)__popc(fut3) << 16);
    uint32_t tot[3];
    block_scan3(sa, sb, sc, sh_scan3, tot);
    pos_s = (int)(sa & 0xffffu);
    pos_w = (int)(sa >> 16);
    spk_tot = (int)(tot[0] & 0xffffu);
    walk_tot = (int)(tot[0] >> 16);
    band_pre[0] = sb & 0xffffu;
    band_pre[1] = sb >> 16;
    band_pre[2] = sc & 0xffffu;
    band_pre[3] = sc >> 16;
    band_tot[0] = tot[1] & 0xffffu;
    band_tot[1] = tot[1] >> 16;
    band_tot[2] = tot[2] & 0xffffu;
    band_tot[3] = tot[2] >> 16;
  } else {
    pos_s = block_scan_fast(live ? spk : 0, sh_scan, &spk_tot);
  }
  {
    if (live && spk) {
      sh_spk[pos_s] = j;
      sh_ext[pos_s] = fo0; /* the fan-out range of the first outgoing projection */
      sh_f[pos_s] = __int_as_float(fo1);
    }
  }
  if (FUSED && live) sh_pj[j] = ps.do_jxx ? pj_n : pj0;

  /* 5b. queue push of this step's spikes (what the queue update of the next
   * step would push from si, Proj.cpp:645-664) into the outgoing rings; its
   * fan-out reads come before the stores below */
  if (a.nproj_out > 0 && spk_tot > 0) {
    __syncthreads();
    for (int i = 0; i < spk_tot; ++i)
      push_rows(a.out0, sh_ext[i], __float_as_int(sh_f[i]), ps.exec, threadIdx.x, blockDim.x);
    for (int o = 1; o < a.nproj_out; ++o)
      for (int i = 0; i < spk_tot; ++i)
        push_unit(a.outs[o], h * M + sh_spk[i], ps.exec, threadIdx.x, blockDim.x);
  }

  /* 7a. the first projection's ring slot of this exec step into LDS (from
   * the words read at the start; a slot of more than four words per thread
   * reads the rest here, before the stores) */
  if (nin > 0 && !fast7) {
    const int W = a.proj0.W;
    const uint32_t* words = a.proj0.fut + ((int64_t)(ps.exec % GSBN_FUT) * a.H + h) * W;
    for (int w = j, i = 0; w < W; w += bd, ++i)
      sh_bitmap[w] = i == 0 ? fut0 : i == 1 ? fut1 : i == 2 ? fut2 : i == 3 ? fut3 : words[w];
  }

  /* --- stores: the population's state (steps 1-4) --- */
  if (live) {
    a.dsup[idx] = dsup;
    a.rnd_u[idx] = u;
    a.rnd_n[idx] = xi;
    a.act[idx] = act;
    a.counter[idx] = counter_new;
    a.spike[idx] = (int8_t)spk;
    if (spk) a.spk_list[h * M + pos_s] = j;
  }
  if (j == 0) a.spk_cnt[h] = spk_tot;

  /* 6. postsynaptic traces and bias for every incoming projection
   * (update_jxx_kernel_cpu, Proj.cpp:310-363); epsc keeps the finalised sum
   * when the traces are not updated */
  /* (the first projection's fields are read from the kernel arguments, not
   * through a pointer into device memory: a load there would wait for every
   * store issued so far) */
  if (live && nin > 0) {
    const gsbn_pop_proj& p = a.proj0;
    if (!ps.do_jxx) {
      a.epsc[idx] = efin[0];
    } else {
      a.epsc[idx] = efin[0] * (1 - p.kepsc);
      if (spk > 0) zj_n += p.kftj;
      if (bj_w) a.bj[idx] = bj_n;
      p.pj[idx] = pj_n;
      p.ej[idx] = ej_n;
      p.zj[idx] = zj_n;
    }
  }
  if (live) {
    for (int m = 1; m < nin; ++m) {
      const gsbn_pop_proj& p = a.projs[m];
      if (!ps.do_jxx) {
        a.epsc[m * N + idx] = efin[m];
        continue;
      }
      a.epsc[m * N + idx] = efin[m] * (1 - p.kepsc);
      const float kp = p.taupdt * ps.prn;
      float pj = p.pj[idx], ej = p.ej[idx], zj = p.zj[idx];
      pj += (ej - pj) * kp;
      ej += (zj - ej) * p.ke;
      zj *= (1 - p.kzj);
      if (spk > 0) zj += p.kftj;
      if (kp != 0.0f) a.bj[m * N + idx] = p.bgain * gsbn_logf(pj + p.eps);
      p.pj[idx] = pj;
      p.ej[idx] = ej;
      p.zj[idx] = zj;
    }
  }

  /* 6b. deferred column updates (first projection): this step's spikes go
   * into the postsynaptic units' rings; a unit whose spikes pending since
   * its last column walk reach dK - 1 is listed for a walk of this step */
  if (nin > 0 && a.proj0.dK) {
    const gsbn_pop_proj& p0 = a.proj0;
    const int K = p0.dK;
    int wflag6 = 0, wT = 0;
    if (live && spk) {
      const int n = dcnt0;
      p0.dhist[(int64_t)idx * K + (n & (K - 1))] = ps.simstep;
      p0.dcnt[idx] = n + 1;
      if (n + 1 - dfl0 >= K - 1) {
        p0.dfl[idx] = n + 1;
        wflag6 = 1;
        wT = n + 1;
      }
    }
    if (fast7) {
      /* positions from the scan of step 5 */
      if (wflag6) {
        p0.walk_list[h * M + pos_w] = j;
        p0.walk_T[h * M + pos_w] = wT;
      }
      if (j == 0) p0.walk_cnt[h] = walk_tot;
    } else if (spk_tot > 0) {
      const int wflag = wflag6;
      int wtot;
      const int pos = block_scan_fast(wflag, sh_scan, &wtot);
      if (wflag) {
        p0.walk_list[h * M + pos] = j;
        p0.walk_T[h * M + pos] = wT;
      }
      if (j == 0) p0.walk_cnt[h] = wtot;
    } else if (j == 0) {
      p0.walk_cnt[h] = 0;
    }
  }

  /* 7. active rows of this step (Proj.cpp:621-630): the ring slot of this
   * exec step, listed in ascending order and cleared for its next use */
  /* a generic lambda: the first projection's call (fields in the kernel
   * arguments) and the others' (fields in device memory) are separate code */
  auto active_rows = [&](const gsbn_pop_proj& p, const int m, auto first_tag) {
    (void)first_tag;
    const int C = p.C;
    const int W = p.W;
    uint32_t* words = p.fut + ((int64_t)(ps.exec % GSBN_FUT) * a.H + h) * W;
    if (m > 0) {
      for (int w = j; w < W; w += bd) sh_bitmap[w] = words[w];
    }
    __syncthreads();
    for (int w = j; w < W; w += bd)
      if (sh_bitmap[w]) words[w] = 0u;
    const int per = (W + blockDim.x - 1) / blockDim.x;
    const int w0 = j * per;
    const int w1 = (w0 + per < W) ? w0 + per : W;
    int c = 0;
    for (int w = w0; w < w1; ++w) c += __popc(sh_bitmap[w]);
    int tot;
    int pos = block_scan_fast(c, sh_scan, &tot);
    /* FUSED: the list also goes to LDS after the bitmap */
    int* rows = reinterpret_cast<int*>(sh_bitmap + W);
    for (int w = w0; w < w1; ++w) {
      uint32_t bits = sh_bitmap[w];
      while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        const int k = (w << 5) + b;
        p.rows_list[(int64_t)h * C + pos] = k;
        p.act_stamp[(int64_t)h * C + k] = ps.exec;
        if (FUSED) rows[pos] = k;
        ++pos;
      }
    }
    if (FUSED) {
      f->sj = sh_sj;
      f->pj = sh_pj;
      f->spk = sh_spk;
      f->nspk = spk_tot;
      f->bitmap = sh_bitmap;
      f->rows = rows;
      f->nrows = tot;
    }
    if (j == 0) {
      p.rows_cnt[h] = tot;
      p.pend_cnt[h] = tot;
      if (a.work) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.work[2 * m]),
                  (unsigned long long)tot * (unsigned long long)M);
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.work[2 * m + 1]),
                  (unsigned long long)spk_tot * (unsigned long long)C);
      }
    }
    /* the next projection reuses the bitmap; after the last one the kernel
     * ends, or the fused caller synchronises before reading the lists */
    if (m + 1 < a.nproj_in) __syncthreads();
  };
  if (fast7) {
    /* the first projection's active rows from the ring words in registers
     * (word j + i*bd of the slot in fut_i), placed by the scan of step 5:
     * band i (words i*bd .. i*bd+bd-1) starts after the bands before it */
    const gsbn_pop_proj& p = a.proj0;
    const int C = p.C;
    const int W = p.W;
    uint32_t* words = p.fut + ((int64_t)(ps.exec % GSBN_FUT) * a.H + h) * W;
    const uint32_t fw[4] = {fut0, fut1, fut2, fut3};
    uint32_t base = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int w = j + i * bd;
      uint32_t bits = fw[i];
      if (bits) {
        words[w] = 0u;
        int pos = (int)(base + band_pre[i]);
        while (bits) {
          const int b = __ffs(bits) - 1;
          bits &= bits - 1;
          const int k = (w << 5) + b;
          p.rows_list[(int64_t)h * C + pos] = k;
          p.act_stamp[(int64_t)h * C + k] = ps.exec;
          ++pos;
        }
      }
      base += band_tot[i];
    }
    if (j == 0) {
      const int tot = (int)base;
      p.rows_cnt[h] = tot;
      p.pend_cnt[h] = tot;
      if (a.work) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.work[0]),
                  (unsigned long long)tot * (unsigned long long)M);
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.work[1]),
                  (unsigned long long)spk_tot * (unsigned long long)C);
      }
    }
    if (nin > 1) __syncthreads();
  } else if (nin > 0) {
    active_rows(a.proj0, 0, std::true_type{});
  }
  for (int m = 1; m < nin; ++m) active_rows(a.projs[m], m, std::false_type{});
}

__global__ void __launch_bounds__(POP_MAX_THREADS) pop_kernel(gsbn_pop_args a) {
  touch_kernarg_lines<sizeof(gsbn_pop_args)>();
  pop_body<false, -1, 4, false>(a, pop_step(a), blockIdx.x);
}

/* per-step fields shared by every population / projection of a batch */
__device__ __forceinline__ void apply_step(gsbn_pop_args& a, const gsbn_step_args& st) {
  a.simstep = st.simstep;
  a.exec = st.exec;
  a.prn = st.prn;
  a.do_jxx = st.do_jxx;
  a.lginp += st.lginp_offset;
  a.wmask += st.wmask_offset;
}

/* Workgroup b updates hypercolumns b, b + gridDim.x, ... of the map (the
 * launch gives one hypercolumn to each workgroup, see loop_grid).  The
 * launch is throughput-bound over many short latency chains: six waves per
 * SIMD (80 registers, no spills) keep more of them in flight than the five
 * the kernel's natural register count allows. */
__global__ void __launch_bounds__(POP_MAX_THREADS, 6)
    pop_kernel_batch(const gsbn_pop_args* __restrict__ arr, const int32_t* __restrict__ wg_pop,
                     const int32_t* __restrict__ wg_hcu, int nwg, gsbn_step_args st) {
  for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads(); /* LDS of the previous hypercolumn */
    const gsbn_pop_args& a = arr[wg_pop[w]];
    touch_lines<sizeof(gsbn_pop_args)>(&a);
    pop_body<false, -1, 1, false>(a, pop_step(a, st), wg_hcu[w]);
  }
}

/* One hypercolumn's population update by one wave, UPL units per lane
 * (unit j = lane + 64 u): pop_body's arithmetic, in the same order per unit,
 * for the batched populations it covers — at most one incoming projection
 * (its ring slot in at most four words per lane, C <= 8192) and one
 * outgoing, no external spikes, no population-wide WTA maximum.  A
 * 128-thread workgroup of pop_body spends two waves, kept together by every
 * barrier, on an HCU of up to 128 units and leaves one idle for M <= 64;
 * here each hypercolumn is one wave, and its lists are placed by wave scans
 * taken unit block by unit block (units u = 0 first), which keeps them
 * ascending.  LDS: three arrays of 64 UPL words. */
template <int UPL>
__device__ __forceinline__ void pop_body_w(const gsbn_pop_args& a, const PopStep& ps, const int h) {
  extern __shared__ __attribute__((aligned(16))) uint32_t sh_dyn[];
  float* const sh_f = reinterpret_cast<float*>(sh_dyn);
  int* const sh_ext = reinterpret_cast<int*>(sh_dyn + 64 * UPL);
  int* const sh_spk = reinterpret_cast<int*>(sh_dyn + 128 * UPL);
  const int lane = threadIdx.x;
  const int M = a.M;
  const int nin = a.nproj_in;
  const gsbn_pop_proj& p0 = a.proj0;

  /* 0. loads that do not depend on this step's results */
  float dsup_old[UPL], lg[UPL], e0[UPL], bj0[UPL], pj0[UPL], ej0[UPL], zj0[UPL], cpre[UPL];
  int32_t counter_old[UPL];
  const float wmv = ps.wmask[h];
  int pend0 = 0;
  uint32_t fut[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const int j = lane + 64 * u;
    const int idx = h * M + j;
    dsup_old[u] = lg[u] = e0[u] = bj0[u] = pj0[u] = ej0[u] = zj0[u] = cpre[u] = 0.0f;
    counter_old[u] = 0;
    if (j < M) {
      dsup_old[u] = a.dsup[idx];
      lg[u] = ps.lginp[idx];
      counter_old[u] = a.counter[idx];
      if (nin > 0) {
        e0[u] = a.epsc[idx];
        bj0[u] = a.bj[idx];
        pj0[u] = p0.pj[idx];
        ej0[u] = p0.ej[idx];
        zj0[u] = p0.zj[idx];
        if (p0.C > 0) cpre[u] = p0.contrib[((int64_t)h * p0.C) * M + j];
      }
    }
  }
  if (nin > 0) {
    pend0 = p0.pend_cnt[h];
    const uint32_t* wslot = p0.fut + ((int64_t)(ps.exec % GSBN_FUT) * a.H + h) * p0.W;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (lane + 64 * i < p0.W) fut[i] = wslot[lane + 64 * i];
  }
  /* the draws and the first projection's postsynaptic traces do not depend
   * on this step's results either */
  float ud[UPL], xi[UPL], pj_n[UPL], ej_n[UPL], zj_n[UPL], bj_n[UPL];
  bool bj_w[UPL];
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const int j = lane + 64 * u;
    ud[u] = xi[u] = pj_n[u] = ej_n[u] = zj_n[u] = bj_n[u] = 0.0f;
    bj_w[u] = false;
    if (j < M) {
      float z;
      gsbn_draw_unit(a.seed, (uint32_t)a.pop_id, (uint32_t)ps.simstep, (uint32_t)(h * M + j), &ud[u],
                     &z);
      xi[u] = a.snoise * z;
      if (ps.do_jxx && nin > 0) {
        const float kp = p0.taupdt * ps.prn;
        pj_n[u] = pj0[u] + (ej0[u] - pj0[u]) * kp;
        ej_n[u] = ej0[u] + (zj0[u] - ej0[u]) * p0.ke;
        zj_n[u] = zj0[u] * (1 - p0.kzj);
        bj_w[u] = kp != 0.0f;
        if (bj_w[u]) bj_n[u] = p0.bgain * gsbn_logf(pj_n[u] + p0.eps);
      }
    }
  }

  /* 1. pending support contributions, ascending row order; 2. dsup */
  float efin[UPL], dsup[UPL];
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const int j = lane + 64 * u;
    efin[u] = 0.0f;
    dsup[u] = 0.0f;
    if (j >= M) continue;
    if (nin > 0) {
      float e = e0[u];
      if (0 < pend0 && gsbn_fbits(cpre[u]) != GSBN_NO_CONTRIB_BITS) e = e + cpre[u];
      const float* c = p0.contrib + ((int64_t)h * p0.C) * M + j;
      for (int k = 1; k < pend0; k += 8) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = k + q < pend0 ? c[(int64_t)(k + q) * M] : 0.0f;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (k + q < pend0 && gsbn_fbits(v[q]) != GSBN_NO_CONTRIB_BITS) e = e + v[q];
      }
      efin[u] = e;
    }
    float wsup = 0.0f;
    if (nin > 0) wsup += bj0[u] + efin[u];
    float sup = a.lgbias + a.igain * lg[u] + xi[u];
    sup += (a.wgain * wmv) * wsup;
    dsup[u] = dsup_old[u] + (sup - dsup_old[u]) * a.taumdt;
  }

  /* 3. WTA maximum (Pop.cpp:270-279), unit blocks in order */
  float maxdsup = -INFINITY;
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const float o = wave_max_left(lane + 64 * u < M ? dsup[u] : -INFINITY);
    if (u == 0 || o > maxdsup) maxdsup = o;
  }

  /* soft-max (Pop.cpp:280-302), the normaliser in ascending unit order */
  const float maxact = gsbn_expf(a.wtagain * maxdsup);
  float act[UPL];
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const int j = lane + 64 * u;
    act[u] = 0.0f;
    if (j < M) {
      act[u] = gsbn_expf(a.wtagain * (dsup[u] - maxdsup));
      if (maxact < 1.0f) act[u] *= maxact;
      sh_f[j] = act[u];
    }
  }
  __syncthreads();
  float vsum = 0.0f;
  if (lane == 0) vsum = ordered_sum_lds(sh_f, M);
  vsum = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vsum)));

  /* 4. spikes and counters (Pop.cpp:305-323); the fan-out range and ring
   * state of the spiking units */
  int spk[UPL], fo0[UPL], fo1[UPL], dcnt0[UPL], dfl0[UPL];
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const int j = lane + 64 * u;
    const int idx = h * M + j;
    if (j < M && vsum > 1.0f) act[u] /= vsum;
    spk[u] = (j < M && ud[u] < act[u] * a.maxfqdt) ? 1 : 0;
    fo0[u] = fo1[u] = dcnt0[u] = dfl0[u] = 0;
    if (spk[u]) {
      if (a.nproj_out > 0) {
        fo0[u] = a.out0.fwd_off[idx];
        fo1[u] = a.out0.fwd_off[idx + 1];
      }
      if (nin > 0 && p0.dK) {
        dcnt0[u] = p0.dcnt[idx];
        dfl0[u] = p0.dfl[idx];
      }
    }
  }

  /* 5. positions: spikes | walks per unit block, then the set bits of the
   * ring words (word lane + 64 i in band i) */
  int pos_s[UPL], pos_w[UPL], wflag[UPL];
  int spk_tot = 0, walk_tot = 0;
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    wflag[u] = (nin > 0 && p0.dK && spk[u] && dcnt0[u] + 1 - dfl0[u] >= p0.dK - 1) ? 1 : 0;
    const uint32_t v = (uint32_t)spk[u] | ((uint32_t)wflag[u] << 16);
    const uint32_t incl = wave_incl_scan(v);
    const uint32_t ex = incl - v;
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    pos_s[u] = spk_tot + (int)(ex & 0xffffu);
    pos_w[u] = walk_tot + (int)(ex >> 16);
    spk_tot += (int)(tot & 0xffffu);
    walk_tot += (int)(tot >> 16);
  }
  uint32_t band_pre[4], band_tot[4];
  {
    const uint32_t sb = (uint32_t)__popc(fut[0]) | ((uint32_t)__popc(fut[1]) << 16);
    const uint32_t sc = (uint32_t)__popc(fut[2]) | ((uint32_t)__popc(fut[3]) << 16);
    const uint32_t ib = wave_incl_scan(sb), ic = wave_incl_scan(sc);
    const uint32_t eb = ib - sb, ec = ic - sc;
    const uint32_t tb = (uint32_t)__builtin_amdgcn_readlane((int)ib, 63);
    const uint32_t tc = (uint32_t)__builtin_amdgcn_readlane((int)ic, 63);
    band_pre[0] = eb & 0xffffu;
    band_pre[1] = eb >> 16;
    band_pre[2] = ec & 0xffffu;
    band_pre[3] = ec >> 16;
    band_tot[0] = tb & 0xffffu;
    band_tot[1] = tb >> 16;
    band_tot[2] = tc & 0xffffu;
    band_tot[3] = tc >> 16;
  }
  __syncthreads(); /* the normaliser's reads of sh_f are done */
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    if (!spk[u]) continue;
    sh_spk[pos_s[u]] = lane + 64 * u;
    sh_ext[pos_s[u]] = fo0[u];
    sh_f[pos_s[u]] = __int_as_float(fo1[u]);
  }

  /* 5b. queue push of this step's spikes into the outgoing ring */
  if (a.nproj_out > 0 && spk_tot > 0) {
    __syncthreads();
    for (int i = 0; i < spk_tot; ++i)
      push_rows(a.out0, sh_ext[i], __float_as_int(sh_f[i]), ps.exec, lane, 64);
  }

  /* --- stores: the population's state (steps 1-4) --- */
#pragma unroll
  for (int u = 0; u < UPL; ++u) {
    const int j = lane + 64 * u;
    if (j >= M) continue;
    const int idx = h * M + j;
    a.dsup[idx] = dsup[u];
    a.rnd_u[idx] = ud[u];
    a.rnd_n[idx] = xi[u];
    a.act[idx] = act[u];
    a.counter[idx] = counter_old[u] + spk[u];
    a.spike[idx] = (int8_t)spk[u];
    if (spk[u]) a.spk_list[h * M + pos_s[u]] = j;
  }
  if (lane == 0) a.spk_cnt[h] = spk_tot;

  if (nin > 0) {
    /* 6. postsynaptic traces and bias (Proj.cpp:310-363) */
#pragma unroll
    for (int u = 0; u < UPL; ++u) {
      const int j = lane + 64 * u;
      if (j >= M) continue;
      const int idx = h * M + j;
      if (!ps.do_jxx) {
        a.epsc[idx] = efin[u];
      } else {
        a.epsc[idx] = efin[u] * (1 - p0.kepsc);
        float zj = zj_n[u];
        if (spk[u] > 0) zj += p0.kftj;
        if (bj_w[u]) a.bj[idx] = bj_n[u];
        p0.pj[idx] = pj_n[u];
        p0.ej[idx] = ej_n[u];
        p0.zj[idx] = zj;
      }
    }
    /* 6b. deferred column updates: ring pushes and this step's walks */
    if (p0.dK) {
      const int K = p0.dK;
#pragma unroll
      for (int u = 0; u < UPL; ++u) {
        if (!spk[u]) continue;
        const int j = lane + 64 * u;
        const int idx = h * M + j;
        const int n = dcnt0[u];
        p0.dhist[(int64_t)idx * K + (n & (K - 1))] = ps.simstep;
        p0.dcnt[idx] = n + 1;
        if (wflag[u]) {
          p0.dfl[idx] = n + 1;
          p0.walk_list[h * M + pos_w[u]] = j;
          p0.walk_T[h * M + pos_w[u]] = n + 1;
        }
      }
      if (lane == 0) p0.walk_cnt[h] = walk_tot;
    }
    /* 7. active rows of this step (Proj.cpp:621-630) from the ring words */
    const int C = p0.C;
    uint32_t* words = p0.fut + ((int64_t)(ps.exec % GSBN_FUT) * a.H + h) * p0.W;
    uint32_t base = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int w = lane + 64 * i;
      uint32_t bits = fut[i];
      if (bits) {
        words[w] = 0u;
        int pos = (int)(base + band_pre[i]);
        while (bits) {
          const int b = __ffs(bits) - 1;
          bits &= bits - 1;
          const int k = (w << 5) + b;
          p0.rows_list[(int64_t)h * C + pos] = k;
          p0.act_stamp[(int64_t)h * C + k] = ps.exec;
          ++pos;
        }
      }
      base += band_tot[i];
    }
    if (lane == 0) {
      const int tot = (int)base;
      p0.rows_cnt[h] = tot;
      p0.pend_cnt[h] = tot;
      if (a.work) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.work[0]),
                  (unsigned long long)tot * (unsigned long long)M);
        atomicAdd(reinterpret_cast<unsigned long long*>(&a.work[1]),
                  (unsigned long long)spk_tot * (unsigned long long)C);
      }
    }
    /* rows past the first ovf_keep go to this hypercolumn's overflow
     * sub-queue, one task each (tasks past the sub-queue's share of ovf_cap
     * are not written: the row/column launch then runs every row of that
     * sub-queue's hypercolumns in their own waves) */
    const int keep = ps.ovf_keep;
    if (keep > 0 && (int)base > keep) {
      const int n = (int)base - keep;
      const int qq = ovf_queue(p0.bidx, h);
      const int capq = ps.ovf_cap / GSBN_OVF_QUEUES;
      int b0 = 0;
      if (lane == 0) b0 = atomicAdd(&ps.ovf_cnt[ovf_counter(ps.exec, qq)], n);
      b0 = __builtin_amdgcn_readfirstlane(b0);
      for (int i = lane; i < n; i += 64) {
        if (b0 + i < capq) {
          const int64_t t = (int64_t)qq * capq + b0 + i;
          ps.ovf_task[3 * t] = p0.bidx;
          ps.ovf_task[3 * t + 1] = h;
          ps.ovf_task[3 * t + 2] = keep + i;
        }
      }
    }
  }
}

/* pop_kernel_batch with one 64-thread workgroup (one wave) per hypercolumn
 * of the map, through pop_body_w; UPL = 2 covers M <= 128.  Six waves per
 * SIMD (80 registers, no spills): 24 hypercolumns in flight per CU against
 * the block kernel's 12. */
template <int UPL>
__global__ void __launch_bounds__(64, 6)
    pop_kernel_batch_wave(const gsbn_pop_args* __restrict__ arr, const int32_t* __restrict__ wg_pop,
                          const int32_t* __restrict__ wg_hcu, int nwg, gsbn_step_args st) {
  /* the overflow counter of the next step (the row/column launch of the
   * previous step, which read it, has finished) */
  if (st.ovf_cnt && blockIdx.x == 0 && threadIdx.x < GSBN_OVF_QUEUES)
    st.ovf_cnt[ovf_counter(st.exec + 1, threadIdx.x)] = 0;
  for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
    if (w != (int)blockIdx.x) __syncthreads();
    const gsbn_pop_args& a = arr[wg_pop[w]];
    touch_lines<sizeof(gsbn_pop_args)>(&a);
    pop_body_w<UPL>(a, pop_step(a, st), wg_hcu[w]);
  }
}

__global__ void pop_dsup0_kernel_batch(const gsbn_pop_args* __restrict__ arr, int npop,
                                       gsbn_step_args st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npop) return;
  gsbn_pop_args a = arr[i];
  if (!a.wta_pop_max) return;
  apply_step(a, st);
  float efin[GSBN_MAX_PROJ_IN];
  for (int m = 0; m < a.nproj_in; ++m) efin[m] = finalized_epsc(a, a.projs[m], m, 0, 0, a.projs[m].pend_cnt[0]);
  float u, xi;
  a.scratch_dsup0[0] = unit_dsup(a, 0, 0, efin, &u, &xi);
}

__global__ void __launch_bounds__(POP_MAX_THREADS) pop_flush_kernel(gsbn_pop_args a) {
  __shared__ int sh_pend[GSBN_MAX_PROJ_IN];
  const int h = blockIdx.x, j = threadIdx.x;
  const int64_t N = (int64_t)a.H * a.M;
  if (j < a.nproj_in) sh_pend[j] = a.projs[j].pend_cnt[h];
  __syncthreads();
  if (j < a.M) {
    for (int m = 0; m < a.nproj_in; ++m)
      a.epsc[m * N + (int64_t)h * a.M + j] = finalized_epsc(a, a.projs[m], m, h, j, sh_pend[m]);
  }
  __syncthreads();
  if (j < a.nproj_in) a.projs[j].pend_cnt[h] = 0;
}

/* Part `part` of the row/column update of hypercolumn h.  FUSED: the
 * hypercolumn's active rows, spikes and pj come from the population phase
 * of the same workgroup (fl), not from memory. */
template <int PREC, bool FUSED = false, bool ROWS = true, int NT = RC_THREADS, bool RING8 = false>
__device__ __forceinline__ void rowcol_body(const gsbn_rowcol_args& a, const RcStep& q,
                                            const int h, const int part,
                                            const FusedLds* fl = nullptr, const int keep = INT_MAX,
                                            const int row_first = 0, const bool one_row = false) {
  const SynIO<PREC> io(a);
  const int P = q.P;
  const int M = a.M, C = a.C;
  const int tid = threadIdx.x;
  const int t = q.t;
  const bool learn = q.kp != 0.0f;

  /* rows part, part+P, ... of this HCU's ascending active-row list.  The
   * counts, this thread's first row and the first spiking unit are read in
   * one batch, before either is known to be needed (all inside their
   * arrays). */
  /* keep: the rows this call takes at most (the rest are queued tasks);
   * one_row: only row position row_first, no column work (a queued task) */
  const int rpt = M >= NT ? 1 : NT / M;
  int nrows, nspk, k_pre = 0, j_pre = 0;
  if (FUSED) {
    nrows = fl->nrows;
    nspk = fl->nspk;
  } else {
    nrows = one_row ? 1 : ROWS ? a.rows_cnt[h] : 0;
    if (nrows > keep) nrows = keep;
    nspk = one_row ? 0 : a.spk_cnt[h];
    const int aa0 = row_first + part + P * (tid / M);
    if (ROWS && tid / M < rpt && aa0 < C) k_pre = a.rows_list[(int64_t)h * C + aa0];
    if (!one_row) j_pre = a.spk_list[h * M];
  }
  const int mine = nrows > part ? (nrows - part + P - 1) / P : 0;
  for (int base = 0; base < mine; base += rpt) {
    const int nr = (mine - base < rpt) ? mine - base : rpt;
    /* every thread of a row loads the row record with its synapse record and
     * catches the row's i-traces up itself (all threads of the row compute
     * the same values); the j == 0 thread writes the record back once every
     * read of this pass is done */
    gsbn_row_rec st0;
    int64_t r0 = -1;
    for (int f = tid; f < nr * M; f += NT) {
      const int rr = f / M;
      const int j = f - rr * M;
      const int aa = row_first + part + P * (base + rr);
      const int k = FUSED ? fl->rows[aa] : (base == 0 && f == tid) ? k_pre
                                                                   : a.rows_list[(int64_t)h * C + aa];
      const int64_t r = (int64_t)h * C + k;
      const int64_t sidx = r * M + j;
      gsbn_row_rec st = a.rowst[r];
      SynVals s = io.load(sidx, !learn);
      /* the synapse's own coefficients (gap to this step from tij after
       * any replayed kicks) */
      gsbn_syn_coef_t cown;
      if (RING8 && a.dK == 8) {
        /* deferred column updates: this synapse's pending kicks first, from
         * the unit's ring in registers; the ring is read beside the record */
        const int uj = h * M + j;
        const Ring8 rg = ring8_load(a.dhist + (int64_t)uj * 8, a.dcnt[uj]);
        replay_ring8<PREC>(io, s, rg, t - 1, q.table, q.table_len, a.kftj, t, &cown);
      } else {
        if (a.dK) {
          const int uj = h * M + j;
          replay_kicks<PREC>(io, s, a.dhist + (int64_t)uj * a.dK, a.dcnt[uj], a.dK, t - 1,
                             q.table, q.table_len, a.kftj);
        }
        if (t - s.tij > 0) cown = load_coef(q.table, q.table_len, t - s.tij);
      }
      const int Dr = t - st.ti;
      if (Dr <= 0) {
        st.zi += a.kfti;
        st.ti = t;
      } else {
        gsbn_row_catchup(gsbn_coef(q.table, q.table_len, Dr), &st.pi, &st.ei, &st.zi);
        st.zi = st.zi + a.kfti;
        st.ti = t;
      }
      if (j == 0) {
        st0 = st;
        r0 = r;
      }
      float cval;
      bool w_new = false;
      const int D = t - s.tij;
      if (D <= 0) {
        s.zi2 += a.kfti;
        s.tij = t;
        cval = gsbn_bitsf(GSBN_NO_CONTRIB_BITS);
      } else {
        gsbn_syn_catchup_k(cown, &s.pij, &s.eij, &s.zi2, &s.zj2);
        s.zi2 = s.zi2 + a.kfti;
        s.tij = t;
        if (learn) {
          s.wij = gsbn_wij(s.pij, st.pi, FUSED ? fl->pj[j] : a.pj[h * M + j], a.wgain, a.eps,
                           a.eps2);
          w_new = true;
        }
        cval = s.wij;
      }
      /* column increment of a spiking postsynaptic unit after the row
       * update (the reference's col pass sees D == 0 here) */
      if ((FUSED ? fl->sj[j] : a.sj[h * M + j]) > 0) s.zj2 += a.kftj;
      io.store(sidx, s, w_new);
      a.contrib[((int64_t)h * C + aa) * M + j] = cval;
    }
    __syncthreads();
    if (r0 >= 0) a.rowst[r0] = st0;
  }

  /* slice `part` of the slots of every spiking unit of this HCU */
  if (nspk == 0) return;
  const int slice = (C + P - 1) / P;
  const int k0 = part * slice;
  const int k1 = (k0 + slice < C) ? k0 + slice : C;
  /* RC_COLS slots per thread and pass, in phases so that the loads of
   * different slots are in flight together: slot checks, records,
   * coefficients, then the updates */
  constexpr int U = RC_COLS;
  if (!FUSED && a.dK) {
    /* deferred column updates: walk the listed units' columns, applying
     * their pending kicks (up to this step's) to every slot with ii >= 0
     * that this step's row pass does not own.  The unit's ring entries (in
     * spike order) and the coefficients of the gaps between consecutive
     * entries are the same for every slot of the column: they are read once
     * per walk into LDS, so a slot loads only the coefficients of its first
     * pending kick (the gap from its own tij). */
    __shared__ int sh_wring[GSBN_MAX_DEFER];
    __shared__ float sh_wgc[GSBN_MAX_DEFER * 8];
    for (int s_i = 0; s_i < nspk; ++s_i) {
      const int j = s_i == 0 ? j_pre : a.spk_list[h * M + s_i];
      const int uj = h * M + j;
      const int32_t* ring = a.dhist + (int64_t)uj * a.dK;
      const int T = a.walk_T[h * M + s_i];
      const int lo = T - a.dK + 1 > 0 ? T - a.dK + 1 : 0;
      const int cnt = T - lo;
      __syncthreads(); /* the previous walk's LDS */
      if (tid < cnt) sh_wring[tid] = ring[(lo + tid) & (a.dK - 1)];
      __syncthreads();
      if (tid + 1 < cnt) {
        const gsbn_syn_coef_t c = load_coef(q.table, q.table_len, sh_wring[tid + 1] - sh_wring[tid]);
        float* g = sh_wgc + tid * 8;
        g[0] = c.ezi;
        g[1] = c.ezj;
        g[2] = c.ee;
        g[3] = c.ep;
        g[4] = c.cez;
        g[5] = c.cpe;
        g[6] = c.cpz;
      }
      __syncthreads();
      for (int kb = k0 + tid; kb < k1; kb += U * NT) {
        bool ok[U];
        int64_t sidx[U];
        SynVals v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = kb + u * NT;
          const int r = h * C + (k < k1 ? k : k0);
          sidx[u] = (int64_t)r * M + j;
          ok[u] = (k < k1) & (a.ii[r] >= 0) & (a.act_stamp[r] != q.exec);
          v[u] = io.load_traces(sidx[u]);
        }
        /* first pending kick of each slot (entries above its tij; all
         * entries are at most this step) and its coefficients */
        int f[U];
        gsbn_syn_coef_t c0[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          f[u] = cnt;
          for (int n = cnt - 1; n >= 0 && sh_wring[n] > v[u].tij; --n) f[u] = n;
          if (ok[u] && f[u] < cnt) c0[u] = load_coef(q.table, q.table_len, sh_wring[f[u]] - v[u].tij);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!ok[u] || f[u] >= cnt) continue;
          for (int n = f[u]; n < cnt; ++n) {
            gsbn_syn_coef_t c;
            if (n == f[u]) {
              c = c0[u];
            } else {
              const float* g = sh_wgc + (n - 1) * 8;
              c.ezi = g[0];
              c.ezj = g[1];
              c.ee = g[2];
              c.ep = g[3];
              c.cez = g[4];
              c.cpe = g[5];
              c.cpz = g[6];
            }
            gsbn_syn_catchup_k(c, &v[u].pij, &v[u].eij, &v[u].zi2, &v[u].zj2);
            v[u].zj2 = v[u].zj2 + a.kftj;
            v[u].tij = sh_wring[n];
            io.round_traces(v[u]);
          }
          io.store_traces(sidx[u], v[u]);
        }
      }
    }
    return;
  }
  for (int s_i = 0; s_i < nspk; ++s_i) {
    const int j = FUSED ? fl->spk[s_i] : s_i == 0 ? j_pre : a.spk_list[h * M + s_i];
    for (int kb = k0 + tid; kb < k1; kb += U * NT) {
      bool ok[U];
      int64_t sidx[U];
      int iiv[U], stamp[U];
      SynVals v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = kb + u * NT;
        const int r = h * C + (k < k1 ? k : k0);
        sidx[u] = (int64_t)r * M + j;
        iiv[u] = a.ii[r];
        /* rows updated by this step's row pass are skipped */
        if (FUSED) {
          const int kk = r - h * C;
          stamp[u] = ((fl->bitmap[kk >> 5] >> (kk & 31)) & 1u) ? q.exec : q.exec - 1;
        } else {
          stamp[u] = a.act_stamp[r];
        }
        ok[u] = k < k1;
        /* the record is loaded with the slot checks (a skipped slot only
         * costs an unused load) */
        v[u] = io.load_traces(sidx[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok[u] = ok[u] & (iiv[u] >= 0) & (stamp[u] != q.exec);
        if (!ok[u]) v[u].tij = t;
      }
      gsbn_syn_coef_t cf[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int D = t - v[u].tij;
        if (ok[u] && D > 0) cf[u] = load_coef(q.table, q.table_len, D);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const int D = t - v[u].tij;
        if (D <= 0) {
          v[u].zj2 += a.kftj;
        } else {
          gsbn_syn_catchup_k(cf[u], &v[u].pij, &v[u].eij, &v[u].zi2, &v[u].zj2);
          v[u].zj2 = v[u].zj2 + a.kftj;
        }
        v[u].tij = t;
        io.store_traces(sidx[u], v[u]);
      }
    }
  }
}

/* six waves per SIMD fit the batched FP32 and BF16 forms without spills
 * (one wave more than their natural register count allows); the FP16 form
 * would spill and keeps five */
#define RC_WAVES(PREC) ((PREC) == GSBN_PREC_F16 ? 5 : 6)

/* The single-network launch replays the deferred kicks of its row updates
 * from registers (Ring8), which needs more registers: five waves per SIMD,
 * more than its grid (a few hundred workgroups) fills.  The batched launches
 * fill the device and keep the register-lean replay and six waves. */
template <int PREC>
__global__ void __launch_bounds__(RC_THREADS, 5) rowcol_kernel(gsbn_rowcol_args a) {
  touch_kernarg_lines<sizeof(gsbn_rowcol_args)>();
  const int h = blockIdx.x / a.parts;
  const RcStep q = {a.simstep, a.exec, a.parts, a.table_len, a.kp, a.table};
  rowcol_body<PREC, false, true, RC_THREADS, true>(a, q, h, blockIdx.x - h * a.parts);
}

/* work item i = (map entry i / parts, part i % parts); workgroup b takes
 * items b, b + gridDim.x, ... (see loop_grid) */
template <int PREC>
__global__ void __launch_bounds__(RC_THREADS, RC_WAVES(PREC))
    rowcol_kernel_batch(const gsbn_rowcol_args* __restrict__ arr, const int32_t* __restrict__ wg_proj,
                        const int32_t* __restrict__ wg_hcu, int parts, int nitems,
                        gsbn_step_args st) {
  for (int i = blockIdx.x; i < nitems; i += gridDim.x) {
    const int w = i / parts;
    const gsbn_rowcol_args& a = arr[wg_proj[w]];
    const RcStep q = {st.simstep, st.exec, parts, a.table_lens[st.kp_slot], a.taupdt * st.prn,
                      a.tables[st.kp_slot]};
    rowcol_body<PREC>(a, q, wg_hcu[w], i - w * parts);
  }
}

/* The same with one wave per hypercolumn (64-thread workgroups): a batch's
 * hypercolumns have few active rows and rarely a spiking unit, so most of a
 * 256-thread workgroup would idle while it holds its place on the CU; one
 * wave each keeps four times as many hypercolumns in flight.  Six waves per
 * SIMD as the other batched forms (FP16 five): 2.5 % faster than the five
 * its register count gives on 512 networks, seven slower. */
template <int PREC>
__global__ void __launch_bounds__(64, RC_WAVES(PREC))
    rowcol_kernel_batch_wave(const gsbn_rowcol_args* __restrict__ arr,
                             const int32_t* __restrict__ wg_proj,
                             const int32_t* __restrict__ wg_hcu, int nhcu, int nhelp,
                             gsbn_step_args st) {
  /* workgroups below nhelp (a multiple of the sub-queue count) run the
   * queued rows of this step, helper i those of sub-queue i % 64; the others
   * one hypercolumn each, keeping ovf_keep rows unless their sub-queue
   * overflowed, in which case they run every row themselves */
  const int capq = st.ovf_cap / GSBN_OVF_QUEUES;
  if ((int)blockIdx.x < nhelp) {
    const int qq = (int)blockIdx.x & (GSBN_OVF_QUEUES - 1);
    const int n = st.ovf_cnt[ovf_counter(st.exec, qq)];
    if (n > capq) return;
    const int stride = nhelp / GSBN_OVF_QUEUES;
    for (int i = (int)blockIdx.x / GSBN_OVF_QUEUES; i < n; i += stride) {
      const int32_t* tk = st.ovf_task + 3 * ((int64_t)qq * capq + i);
      const gsbn_rowcol_args& a = arr[tk[0]];
      const RcStep q = {st.simstep, st.exec, 1, a.table_lens[st.kp_slot], a.taupdt * st.prn,
                        a.tables[st.kp_slot]};
      rowcol_body<PREC, false, true, 64>(a, q, tk[1], 0, nullptr, INT_MAX, tk[2], true);
    }
    return;
  }
  for (int w = (int)blockIdx.x - nhelp; w < nhcu; w += (int)gridDim.x - nhelp) {
    const int pj = wg_proj[w], h = wg_hcu[w];
    const gsbn_rowcol_args& a = arr[pj];
    const RcStep q = {st.simstep, st.exec, 1, a.table_lens[st.kp_slot], a.taupdt * st.prn,
                      a.tables[st.kp_slot]};
    int keep = INT_MAX;
    if (nhelp > 0 && st.ovf_cnt[ovf_counter(st.exec, ovf_queue(pj, h))] <= capq) keep = st.ovf_keep;
    rowcol_body<PREC, false, true, 64>(a, q, h, 0, nullptr, keep);
  }
}

/* The one-wave form with a work split per hypercolumn (the graded maps
 * below): the hypercolumns with the most slots get several waves, each
 * taking every P-th active row and a 1/P slice of the column walks, so that
 * the few with many active rows in a step no longer end the launch alone. */
template <int PREC>
__global__ void __launch_bounds__(64, RC_WAVES(PREC))
    rowcol_kernel_graded_wave(const gsbn_rowcol_args* __restrict__ arr,
                              const int32_t* __restrict__ wg_proj,
                              const int32_t* __restrict__ wg_hcu,
                              const int32_t* __restrict__ wg_part, int nwg, gsbn_step_args st) {
  for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
    const gsbn_rowcol_args& a = arr[wg_proj[w]];
    const int pp = wg_part[w];
    const RcStep q = {st.simstep, st.exec, pp >> 16, a.table_lens[st.kp_slot], a.taupdt * st.prn,
                      a.tables[st.kp_slot]};
    rowcol_body<PREC, false, true, 64>(a, q, wg_hcu[w], pp & 0xffff);
  }
}

/* The same with a work split per hypercolumn: workgroup w runs part
 * wg_part[w] & 0xffff of wg_part[w] >> 16 of hypercolumn wg_hcu[w] of
 * projection wg_proj[w] (larger hypercolumns split into more parts). */
template <int PREC>
__global__ void __launch_bounds__(RC_THREADS, RC_WAVES(PREC))
    rowcol_kernel_graded(const gsbn_rowcol_args* __restrict__ arr,
                         const int32_t* __restrict__ wg_proj, const int32_t* __restrict__ wg_hcu,
                         const int32_t* __restrict__ wg_part, int nwg, gsbn_step_args st) {
  for (int w = blockIdx.x; w < nwg; w += gridDim.x) {
    const gsbn_rowcol_args& a = arr[wg_proj[w]];
    const int pp = wg_part[w];
    const RcStep q = {st.simstep, st.exec, pp >> 16, a.table_lens[st.kp_slot], a.taupdt * st.prn,
                      a.tables[st.kp_slot]};
    rowcol_body<PREC>(a, q, wg_hcu[w], pp & 0xffff);
  }
}

/* One launch per step: workgroup h updates hypercolumn h's population and
 * then, from LDS, its rows and columns.  Everything the row/column update
 * reads in a step comes from the same workgroup and the queue pushes go to
 * slots at least three steps ahead, so no workgroup waits for another. */
template <int PREC>
__global__ void __launch_bounds__(RC_THREADS) step_kernel(gsbn_pop_args a, gsbn_rowcol_args b) {
  const int h = blockIdx.x;
  FusedLds f;
  pop_body<true>(a, pop_step(a), h, &f);
  __syncthreads();
  const RcStep q = {b.simstep, b.exec, 1, b.table_len, b.kp, b.table};
  rowcol_body<PREC, true>(b, q, h, 0, &f);
}

template <int PREC>
__global__ void __launch_bounds__(RC_THREADS)
    step_kernel_batch(const gsbn_pop_args* __restrict__ pops, const gsbn_rowcol_args* __restrict__ projs,
                      const int32_t* __restrict__ wg_pop, const int32_t* __restrict__ wg_hcu,
                      const int32_t* __restrict__ in_proj, gsbn_step_args st) {
  const int p = wg_pop[blockIdx.x];
  const int h = wg_hcu[blockIdx.x];
  const gsbn_pop_args& a = pops[p];
  touch_lines<sizeof(gsbn_pop_args)>(&a);
  FusedLds f;
  pop_body<true>(a, pop_step(a, st), h, &f);
  __syncthreads();
  const gsbn_rowcol_args& b = projs[in_proj[p]];
  const RcStep q = {st.simstep, st.exec, 1, b.table_lens[st.kp_slot], b.taupdt * st.prn,
                    b.tables[st.kp_slot]};
  rowcol_body<PREC, true>(b, q, h, 0, &f);
}

/* Chained step, one launch per executed step: workgroup h < H runs the
 * previous step's row updates of hypercolumn h (when has_prev) and this
 * step's population update; the H * b.parts further workgroups run the
 * previous step's column updates.  The two halves touch disjoint state: the
 * population update reads no synapse record, column updates skip the rows
 * the row updates own, and the previous step's spike lists and row stamps
 * (b) are not the ones this step's population update writes (a). */
template <int PREC>
__global__ void __launch_bounds__(RC_THREADS) chained_kernel(gsbn_pop_args a, gsbn_rowcol_args b,
                                                             int has_prev) {
  const RcStep q = {b.simstep, b.exec, b.parts, b.table_len, b.kp, b.table};
  if ((int)blockIdx.x < a.H) {
    /* the population half is the longer path: its waves take issue
     * priority over the column workgroups that share their CUs */
    __builtin_amdgcn_s_setprio(2);
    pop_body<false, PREC>(a, pop_step(a), blockIdx.x, nullptr, &b, &q, has_prev);
    return;
  }
  const int w = blockIdx.x - a.H;
  const int h = w / b.parts;
  rowcol_body<PREC, false, false>(b, q, h, w - h * b.parts);
}

template <int PREC>
__global__ void __launch_bounds__(ALL_THREADS) all_kernel(gsbn_all_args a) {
  __shared__ float sh_pi[ALL_ROWS];
  __shared__ int sh_ii[ALL_ROWS];
  const SynIO<PREC> io(a);
  const int64_t R = (int64_t)a.H * a.C;
  const int M = a.M;
  const int tid = threadIdx.x;
  const int t = a.target;
  const bool learn = a.kp != 0.0f;
  const int64_t nchunks = (R + ALL_ROWS - 1) / ALL_ROWS;
  for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const int64_t r0 = ch * ALL_ROWS;
    const int nr = (R - r0 < ALL_ROWS) ? (int)(R - r0) : ALL_ROWS;
    if (tid < nr) {
      const int64_t r = r0 + tid;
      gsbn_row_rec st = a.rowst[r];
      const int D = t - st.ti;
      if (D > 0) {
        const float* c = gsbn_coef(a.table, a.table_len, D);
        gsbn_row_catchup(c, &st.pi, &st.ei, &st.zi);
        st.ti = t;
        a.rowst[r] = st;
      }
      sh_pi[tid] = st.pi;
      sh_ii[tid] = a.dK ? a.ii[r] : -1;
    }
    __syncthreads();
    for (int f = tid; f < nr * M; f += ALL_THREADS) {
      const int rr = f / M;
      const int j = f - rr * M;
      const int64_t r = r0 + rr;
      const int64_t sidx = r * M + j;
      SynVals s = io.load(sidx, false);
      bool kicked = false;
      if (sh_ii[rr] >= 0) {
        /* deferred column updates: the pending kicks up to the target step,
         * with the old learning rate, as the eager column updates had them */
        const int uj = (int)(r / a.C) * M + j;
        kicked = replay_kicks<PREC>(io, s, a.dhist + (int64_t)uj * a.dK, a.dcnt[uj], a.dK, t,
                                    a.table, a.table_len, a.kftj);
      }
      const int D = t - s.tij;
      if (D <= 0) {
        if (kicked) {
          s.tij = t;
          io.store(sidx, s, false);
        } else {
          io.store_tij(sidx, t);
        }
      } else {
        gsbn_syn_catchup_k(load_coef(a.table, a.table_len, D), &s.pij, &s.eij, &s.zi2, &s.zj2);
        s.tij = t;
        if (learn) {
          const int h = (int)(r / a.C);
          s.wij = gsbn_wij(s.pij, sh_pi[rr], a.pj[h * M + j], a.wgain, a.eps, a.eps2);
        }
        io.store(sidx, s, learn);
      }
    }
    __syncthreads();
  }
}

/* gsbn_proj_materialize: every pending column kick of every slot with
 * ii >= 0, applied in place (grid-stride over chunks of rows, a row's M
 * records contiguous) */
template <int PREC>
__global__ void __launch_bounds__(ALL_THREADS) materialize_kernel(gsbn_all_args a) {
  const SynIO<PREC> io(a);
  const int64_t R = (int64_t)a.H * a.C;
  const int M = a.M;
  const int64_t n = R * M;
  for (int64_t f = (int64_t)blockIdx.x * ALL_THREADS + threadIdx.x; f < n;
       f += (int64_t)gridDim.x * ALL_THREADS) {
    const int64_t r = f / M;
    const int j = (int)(f - r * M);
    if (a.ii[r] < 0) continue;
    const int uj = (int)(r / a.C) * M + j;
    const int T = a.dcnt[uj];
    if (T == 0) continue;
    SynVals s = io.load_traces(f);
    if (replay_kicks<PREC>(io, s, a.dhist + (int64_t)uj * a.dK, T, a.dK, a.target, a.table,
                           a.table_len, a.kftj))
      io.store_traces(f, s);
  }
}

__global__ void jxx_kernel(gsbn_jxx_args a) {
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < a.N; idx += gridDim.x * blockDim.x) {
    float pj = a.pj[idx], ej = a.ej[idx], zj = a.zj[idx];
    const int sj = a.sj[idx];
    a.epsc[idx] *= (1 - a.kepsc);
    pj += (ej - pj) * a.kp;
    ej += (zj - ej) * a.ke;
    zj *= (1 - a.kzj);
    if (sj > 0) zj += a.kftj;
    if (a.kp != 0.0f) a.bj[idx] = a.bgain * gsbn_logf(pj + a.eps);
    a.pj[idx] = pj;
    a.ej[idx] = ej;
    a.zj[idx] = zj;
  }
}

/* one workgroup per source HCU; ascending MCU order within the HCU */
__global__ void __launch_bounds__(1024) que_kernel(gsbn_que_args a) {
  const int hs = blockIdx.x;
  for (int j = 0; j < a.Msrc; ++j) {
    const int n = hs * a.Msrc + j;
    if (a.si[n] > 0) push_unit(a.out, n, a.exec - 1, threadIdx.x, blockDim.x);
  }
}

__global__ void check_count_kernel(int n, const int8_t* spike, int32_t* count) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    count[i] += (spike[i] > 0) ? 1 : 0;
}

__global__ void check_argmax_kernel(int H, int M, int32_t* count, int32_t* out_idx,
                                    int32_t* out_cnt) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= H) return;
  int mi = -1, mc = 0;
  for (int j = 0; j < M; ++j) {
    const int v = count[h * M + j];
    if (v > mc) {
      mi = j;
      mc = v;
    }
    count[h * M + j] = 0;
  }
  out_idx[h] = mi;
  out_cnt[h] = mc;
}

/* one workgroup per population */
__global__ void check_count_multi_kernel(const gsbn_check_item* items) {
  const gsbn_check_item it = items[blockIdx.x];
  const int n = it.H * it.M;
  for (int i = threadIdx.x; i < n; i += blockDim.x) it.count[i] += (it.spike[i] > 0) ? 1 : 0;
}

/* one thread per HCU of one population (blockIdx.y) */
__global__ void check_argmax_multi_kernel(const gsbn_check_item* items, int32_t* out_idx,
                                          int32_t* out_cnt) {
  const gsbn_check_item it = items[blockIdx.y];
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= it.H) return;
  int mi = -1, mc = 0;
  for (int j = 0; j < it.M; ++j) {
    const int v = it.count[h * it.M + j];
    if (v > mc) {
      mi = j;
      mc = v;
    }
    it.count[h * it.M + j] = 0;
  }
  out_idx[it.out_offset + h] = mi;
  out_cnt[it.out_offset + h] = mc;
}

__global__ void syn_field_kernel(gsbn_syn_rec* syn, int64_t n, int field, uint32_t* buf,
                                 int to_records) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t* rec = reinterpret_cast<uint32_t*>(syn + i);
    const int off = field == 5 ? 5 : field;
    if (to_records)
      rec[off] = buf[i];
    else
      buf[i] = rec[off];
  }
}

__global__ void row_field_kernel(gsbn_row_rec* row, int64_t n, int field, uint32_t* buf,
                                 int to_records) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t* rec = reinterpret_cast<uint32_t*>(row + i);
    if (to_records)
      rec[field] = buf[i];
    else
      buf[i] = rec[field];
  }
}

__global__ void syn_fill_kernel(gsbn_syn_rec* syn, int64_t n, gsbn_syn_rec v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    syn[i] = v;
}

__global__ void row_fill_kernel(gsbn_row_rec* row, int64_t n, gsbn_row_rec v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    row[i] = v;
}

/* ---- structural plasticity (Prj::prune, Prj::remove_conn) ---- */

/* one wave per row: the weight test is a wave vote over the row's M weights,
 * read only for connected rows that pass the age test */
__global__ void prune_flags_kernel(gsbn_prune_args a) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= a.R) return; /* uniform per wave */
  uint8_t f = 0;
  if (a.ii[row] >= 0) {
    if (a.rowst[row].ti < a.t_min) {
      f = 1;
    } else {
      bool strong = false;
      const int64_t base = row * a.M;
      if (a.prec == GSBN_PREC_F32) {
        for (int j = lane; j < a.M; j += 64) {
          const float w = a.syn[base + j].wij;
          strong = strong || w >= a.wp || w <= a.wn;
        }
      } else {
        for (int j = lane; j < a.M; j += 64) {
          const float w = a.wij[base + j];
          strong = strong || w >= a.wp || w <= a.wn;
        }
      }
      f = __ballot(strong) ? 0 : 1;
    }
  }
  if (lane == 0) a.flags[row] = f;
}

static_assert(GSBN_FUT == 64, "clear_rows_kernel gives one ring slot to each lane");

/* one wave per listed row; lane s clears the row's bit in ring slot s */
__global__ void clear_rows_kernel(gsbn_clear_rows_args a) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= a.n) return;
  const int64_t r = a.rows[i];
  if (r < 0 || r >= (int64_t)a.H * a.C) return;
  if (lane == 0) {
    a.rowst[r].pi = 0.0f;
    a.rowst[r].ei = 0.0f;
    a.rowst[r].ti = 0;
  }
  const int64_t base = r * a.M;
  for (int j = lane; j < a.M; j += 64) {
    if (a.prec == GSBN_PREC_F32) {
      a.syn[base + j].pij = 0.0f;
      a.syn[base + j].eij = 0.0f;
      a.syn[base + j].zj2 = 0.0f;
    } else {
      a.syn16[base + j].pij = 0;
      a.syn16[base + j].eij = 0;
      a.syn16[base + j].zj2 = 0.0f;
    }
  }
  const int h = (int)(r / a.C), k = (int)(r - (int64_t)h * a.C);
  atomicAnd(a.fut + ((int64_t)lane * a.H + h) * a.W + (k >> 5), ~(1u << (k & 31)));
}

/* plain streaming copy, 16 bytes per lane and iteration (the reference
 * point for the memory-traffic counters) */
__global__ void copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, int64_t n16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void i32_fill_kernel(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void syn16_field_kernel(gsbn_syn16_rec* syn, float* wij, int64_t n, int field,
                                   uint32_t* buf, int to_records, int prec, float ps, float es) {
  const float pinv = prec == GSBN_PREC_F16 ? 1.0f / ps : 1.0f;
  const float einv = prec == GSBN_PREC_F16 ? 1.0f / es : 1.0f;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    gsbn_syn16_rec* r = syn + i;
    if (to_records) {
      const uint32_t v = buf[i];
      switch (field) {
        case 0: r->pij = gsbn_pack_trace(prec, gsbn_bitsf(v), ps); break;
        case 1: r->eij = gsbn_pack_trace(prec, gsbn_bitsf(v), es); break;
        case 2: r->zi2 = gsbn_bitsf(v); break;
        case 3: r->zj2 = gsbn_bitsf(v); break;
        case 4: wij[i] = gsbn_bitsf(v); break;
        default: r->tij = (int32_t)v; break;
      }
    } else {
      uint32_t v;
      switch (field) {
        case 0: v = gsbn_fbits(gsbn_unpack_trace(prec, r->pij, pinv)); break;
        case 1: v = gsbn_fbits(gsbn_unpack_trace(prec, r->eij, einv)); break;
        case 2: v = gsbn_fbits(r->zi2); break;
        case 3: v = gsbn_fbits(r->zj2); break;
        case 4: v = gsbn_fbits(wij[i]); break;
        default: v = (uint32_t)r->tij; break;
      }
      buf[i] = v;
    }
  }
}

__global__ void syn16_fill_kernel(gsbn_syn16_rec* syn, int64_t n, gsbn_syn16_rec v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    syn[i] = v;
}

__global__ void device_math_kernel(int op, int64_t n, const uint32_t* in, uint32_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (op == 0) {
      out[i] = gsbn_fbits(gsbn_expf(gsbn_bitsf(in[i])));
    } else if (op == 1) {
      out[i] = gsbn_fbits(gsbn_logf(gsbn_bitsf(in[i])));
    } else if (op == 2) {
      float u, z;
      const uint64_t seed = (uint64_t)in[4 * i] | ((uint64_t)in[4 * i + 1] << 32);
      gsbn_draw_unit(seed, 0u, in[4 * i + 2], in[4 * i + 3], &u, &z);
      out[2 * i] = gsbn_fbits(u);
      out[2 * i + 1] = gsbn_fbits(z);
    } else if (op == 3) {
      out[i] = gsbn_fbits(gsbn_bitsf(in[2 * i]) / gsbn_bitsf(in[2 * i + 1]));
    } else if (op == 4) {
      out[i] = gsbn_fbits(gsbn_wij(gsbn_bitsf(in[3 * i]), gsbn_bitsf(in[3 * i + 1]),
                                   gsbn_bitsf(in[3 * i + 2]), 1.0f, 1e-3f, 1e-6f));
    } else if (op == 5) {
      out[i] = gsbn_f32_to_bf16(gsbn_bitsf(in[i]));
    } else if (op == 6) {
      out[i] = gsbn_f32_to_f16(gsbn_bitsf(in[i]));
    } else if (op == 7) {
      out[i] = gsbn_fbits(gsbn_f16_to_f32((uint16_t)in[i]));
    }
  }
}

int grid_for(int64_t n, int threads, int cap) {
  int64_t g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace

/* power of two, normal, positive: multiplying by it and by its inverse is exact */
static bool pow2_scale(float x) {
  const uint32_t b = gsbn_fbits(x);
  return (b & 0x807fffffu) == 0 && (b >> 23) != 0 && (b >> 23) != 255;
}

template <class A>
static bool trace_form_ok(const A& a) {
  switch (a.prec) {
    case GSBN_PREC_F32: return a.syn != nullptr || (int64_t)a.H * a.C == 0;
    case GSBN_PREC_BF16: return a.syn16 != nullptr && a.wij != nullptr;
    case GSBN_PREC_F16:
      return a.syn16 != nullptr && a.wij != nullptr && pow2_scale(a.p_scale) && pow2_scale(a.e_scale);
    default: return false;
  }
}

/* dynamic LDS of the population update (pop_body): three arrays of
 * `threads` entries and the active-row bitmap; the fused step adds the
 * active-row list */
/* deferred column updates: dK 0 (off) or a power of two in [2, 64] with its
 * arrays */
static bool deferred_ok(int32_t dK, const void* ii, const void* dhist, const void* dcnt) {
  return dK == 0 || (dK >= 2 && dK <= GSBN_MAX_DEFER && (dK & (dK - 1)) == 0 && ii && dhist && dcnt);
}

static size_t pop_lds(int threads, int64_t max_slots, bool fused) {
  return (size_t)(3 * (int64_t)threads + (max_slots + 31) / 32 + (fused ? max_slots : 0)) *
         sizeof(uint32_t);
}

extern "C" {

int gsbn_pop_update(const gsbn_pop_args* a, void* stream) {
  if (!a || a->H <= 0 || a->M <= 0 || a->M > POP_MAX_THREADS || a->nproj_in < 0 ||
      a->nproj_in > GSBN_MAX_PROJ_IN || (a->nproj_in > 0 && !a->projs) || a->max_slots < 0 ||
      a->nproj_out < 0 || (a->nproj_out > 0 && !a->outs))
    return GSBN_ERR_ARG;
  if (a->wta_pop_max && !a->scratch_dsup0) return GSBN_ERR_ARG;
  if (a->nproj_in > 0 && a->proj0.dK &&
      (!deferred_ok(a->proj0.dK, a->proj0.rows_list, a->proj0.dhist, a->proj0.dcnt) ||
       !a->proj0.dfl || !a->proj0.walk_list || !a->proj0.walk_cnt || !a->proj0.walk_T))
    return GSBN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int threads = ((a->M + 63) / 64) * 64;
  const size_t lds = pop_lds(threads, a->max_slots, false);
  if (lds > 48 * 1024) return GSBN_ERR_UNSUPPORTED;
  if (a->wta_pop_max) {
    hipLaunchKernelGGL(pop_dsup0_kernel, dim3(1), dim3(64), 0, s, *a);
    int rc = launch_status();
    if (rc) return rc;
  }
  launch_timed(pop_kernel, dim3(a->H), dim3(threads), lds, s, *a);
  return launch_status();
}

int gsbn_pop_flush(const gsbn_pop_args* a, void* stream) {
  if (!a || a->H <= 0 || a->M <= 0 || a->M > POP_MAX_THREADS || a->nproj_in < 0 ||
      a->nproj_in > GSBN_MAX_PROJ_IN || (a->nproj_in > 0 && !a->projs))
    return GSBN_ERR_ARG;
  const int threads = ((a->M + 63) / 64) * 64;
  hipLaunchKernelGGL(pop_flush_kernel, dim3(a->H), dim3(threads), 0, (hipStream_t)stream, *a);
  return launch_status();
}

int gsbn_proj_materialize(const gsbn_all_args* a, void* stream) {
  if (!a || a->H <= 0 || a->M <= 0 || a->C < 0 || !a->table || a->table_len < 2 ||
      !trace_form_ok(*a) || a->dK <= 0 || !deferred_ok(a->dK, a->ii, a->dhist, a->dcnt))
    return GSBN_ERR_ARG;
  const int64_t n = (int64_t)a->H * a->C * a->M;
  if (n == 0) return 0;
  const int grid = grid_for(n, ALL_THREADS, 16384);
  hipStream_t s = (hipStream_t)stream;
  switch (a->prec) {
    case GSBN_PREC_F32:
      hipLaunchKernelGGL(materialize_kernel<GSBN_PREC_F32>, dim3(grid), dim3(ALL_THREADS), 0, s, *a);
      break;
    case GSBN_PREC_BF16:
      hipLaunchKernelGGL(materialize_kernel<GSBN_PREC_BF16>, dim3(grid), dim3(ALL_THREADS), 0, s, *a);
      break;
    default:
      hipLaunchKernelGGL(materialize_kernel<GSBN_PREC_F16>, dim3(grid), dim3(ALL_THREADS), 0, s, *a);
      break;
  }
  return launch_status();
}

int gsbn_proj_all(const gsbn_all_args* a, void* stream) {
  if (!a || a->H <= 0 || a->M <= 0 || a->C < 0 || !a->table || a->table_len < 2 ||
      !trace_form_ok(*a) || !deferred_ok(a->dK, a->ii, a->dhist, a->dcnt))
    return GSBN_ERR_ARG;
  const int64_t R = (int64_t)a->H * a->C;
  if (R == 0) return 0;
  const int grid = grid_for((R + ALL_ROWS - 1) / ALL_ROWS, 1, 8192);
  hipStream_t s = (hipStream_t)stream;
  switch (a->prec) {
    case GSBN_PREC_F32:
      launch_timed(all_kernel<GSBN_PREC_F32>, dim3(grid), dim3(ALL_THREADS), 0, s, *a);
      break;
    case GSBN_PREC_BF16:
      launch_timed(all_kernel<GSBN_PREC_BF16>, dim3(grid), dim3(ALL_THREADS), 0, s, *a);
      break;
    default:
      launch_timed(all_kernel<GSBN_PREC_F16>, dim3(grid), dim3(ALL_THREADS), 0, s, *a);
      break;
  }
  return launch_status();
}

int gsbn_proj_jxx(const gsbn_jxx_args* a, void* stream) {
  if (!a || a->N < 0) return GSBN_ERR_ARG;
  if (a->N == 0) return 0;
  launch_timed(jxx_kernel, dim3(grid_for(a->N, 256, 4096)), dim3(256), 0,
                     (hipStream_t)stream, *a);
  return launch_status();
}

int gsbn_rowcol_resident(int32_t prec, int32_t* out) {
  if (!out) return GSBN_ERR_ARG;
  int dev = 0, cus = 0, per = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) {
    switch (prec) {
      case GSBN_PREC_F32:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, rowcol_kernel<GSBN_PREC_F32>, RC_THREADS, 0);
        break;
      case GSBN_PREC_BF16:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, rowcol_kernel<GSBN_PREC_BF16>, RC_THREADS, 0);
        break;
      default:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, rowcol_kernel<GSBN_PREC_F16>, RC_THREADS, 0);
        break;
    }
  }
  if (e != hipSuccess) return ret(e);
  *out = per * cus;
  return 0;
}

int gsbn_proj_rowcol(const gsbn_rowcol_args* a, void* stream) {
  if (!a || a->H <= 0 || a->M <= 0 || a->C < 0 || !a->table || a->table_len < 2 || a->parts < 0 ||
      !trace_form_ok(*a) || !deferred_ok(a->dK, a->ii, a->dhist, a->dcnt) ||
      (a->dK && !a->walk_T))
    return GSBN_ERR_ARG;
  gsbn_rowcol_args b = *a;
  if (b.parts == 0) {
    int32_t res = 0;
    const int rc = gsbn_rowcol_resident(b.prec, &res);
    if (rc) return rc;
    const int p = res / b.H;
    b.parts = p < 1 ? 1 : (p > 32 ? 32 : p);
  }
  const int64_t grid = (int64_t)b.H * b.parts;
  if (grid > 0x7fffffff) return GSBN_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  switch (b.prec) {
    case GSBN_PREC_F32:
      launch_timed(rowcol_kernel<GSBN_PREC_F32>, dim3((unsigned)grid), dim3(RC_THREADS), 0, s, b);
      break;
    case GSBN_PREC_BF16:
      launch_timed(rowcol_kernel<GSBN_PREC_BF16>, dim3((unsigned)grid), dim3(RC_THREADS), 0, s, b);
      break;
    default:
      launch_timed(rowcol_kernel<GSBN_PREC_F16>, dim3((unsigned)grid), dim3(RC_THREADS), 0, s, b);
      break;
  }
  return launch_status();
}

int gsbn_pop_update_batch(const gsbn_pop_args* pops_dev, int32_t npop, const int32_t* wg_pop,
                          const int32_t* wg_hcu, int32_t nwg, int32_t threads, int32_t max_slots,
                          int32_t any_wta_pop_max, const gsbn_step_args* step, void* stream) {
  if (!pops_dev || !wg_pop || !wg_hcu || !step || nwg <= 0 || npop <= 0 || threads <= 0 ||
      threads > POP_MAX_THREADS || threads % 64 || max_slots < 0)
    return GSBN_ERR_ARG;
  const size_t lds = pop_lds(threads, max_slots, false);
  if (lds > 48 * 1024) return GSBN_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  if (any_wta_pop_max) {
    hipLaunchKernelGGL(pop_dsup0_kernel_batch, dim3((npop + 63) / 64), dim3(64), 0, s, pops_dev,
                       npop, *step);
    int rc = launch_status();
    if (rc) return rc;
  }
  launch_timed(pop_kernel_batch, dim3(loop_grid(pop_kernel_batch, threads, lds, nwg)),
               dim3(threads), lds, s, pops_dev, wg_pop, wg_hcu, nwg, *step);
  return launch_status();
}

int gsbn_pop_update_batch_wave(const gsbn_pop_args* pops_dev, const int32_t* wg_pop,
                               const int32_t* wg_hcu, int32_t nwg, int32_t units_per_lane,
                               const gsbn_step_args* step, void* stream) {
  if (!pops_dev || !wg_pop || !wg_hcu || !step || nwg <= 0 ||
      (units_per_lane != 1 && units_per_lane != 2))
    return GSBN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const size_t lds = (size_t)3 * 64 * units_per_lane * sizeof(uint32_t);
  if (units_per_lane == 1)
    launch_timed(pop_kernel_batch_wave<1>, dim3(loop_grid(pop_kernel_batch_wave<1>, 64, lds, nwg)),
                 dim3(64), lds, s, pops_dev, wg_pop, wg_hcu, nwg, *step);
  else
    launch_timed(pop_kernel_batch_wave<2>, dim3(loop_grid(pop_kernel_batch_wave<2>, 64, lds, nwg)),
                 dim3(64), lds, s, pops_dev, wg_pop, wg_hcu, nwg, *step);
  return launch_status();
}

int gsbn_proj_rowcol_batch(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                           const int32_t* wg_hcu, int32_t nhcu, int32_t parts,
                           const gsbn_step_args* step, void* stream) {
  if (!projs_dev || !wg_proj || !wg_hcu || !step || nhcu <= 0 || parts <= 0 ||
      step->kp_slot < 0 || step->kp_slot >= GSBN_KP_SLOTS || step->prec < GSBN_PREC_F32 ||
      step->prec > GSBN_PREC_F16)
    return GSBN_ERR_ARG;
  const int64_t items = (int64_t)nhcu * parts;
  if (items > 0x7fffffff) return GSBN_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  const dim3 b(RC_THREADS);
  const int n = (int)items;
  switch (step->prec) {
    case GSBN_PREC_F32:
      launch_timed(rowcol_kernel_batch<GSBN_PREC_F32>,
                   dim3(loop_grid(rowcol_kernel_batch<GSBN_PREC_F32>, RC_THREADS, 0, n)), b, 0, s,
                   projs_dev, wg_proj, wg_hcu, parts, n, *step);
      break;
    case GSBN_PREC_BF16:
      launch_timed(rowcol_kernel_batch<GSBN_PREC_BF16>,
                   dim3(loop_grid(rowcol_kernel_batch<GSBN_PREC_BF16>, RC_THREADS, 0, n)), b, 0,
                   s, projs_dev, wg_proj, wg_hcu, parts, n, *step);
      break;
    default:
      launch_timed(rowcol_kernel_batch<GSBN_PREC_F16>,
                   dim3(loop_grid(rowcol_kernel_batch<GSBN_PREC_F16>, RC_THREADS, 0, n)), b, 0,
                   s, projs_dev, wg_proj, wg_hcu, parts, n, *step);
      break;
  }
  return launch_status();
}



int gsbn_proj_rowcol_batch_wave(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                                const int32_t* wg_hcu, int32_t nhcu, const gsbn_step_args* step,
                                void* stream) {
  if (!projs_dev || !wg_proj || !wg_hcu || !step || nhcu <= 0 || step->kp_slot < 0 ||
      step->kp_slot >= GSBN_KP_SLOTS || step->prec < GSBN_PREC_F32 || step->prec > GSBN_PREC_F16 ||
      (step->ovf_keep > 0 && (!step->ovf_cnt || !step->ovf_task || step->ovf_cap < 0)))
    return GSBN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const dim3 b(64);
  /* helper waves for the queued rows: enough to run a busy step's queue in
   * about one pass of the device */
  const int nhelp = step->ovf_keep > 0 ? GSBN_OVF_HELPERS : 0;
  const unsigned g = (unsigned)(nhcu + nhelp);
  switch (step->prec) {
    case GSBN_PREC_F32:
      launch_timed(rowcol_kernel_batch_wave<GSBN_PREC_F32>, dim3(g), b, 0, s, projs_dev, wg_proj,
                   wg_hcu, nhcu, nhelp, *step);
      break;
    case GSBN_PREC_BF16:
      launch_timed(rowcol_kernel_batch_wave<GSBN_PREC_BF16>, dim3(g), b, 0, s, projs_dev, wg_proj,
                   wg_hcu, nhcu, nhelp, *step);
      break;
    default:
      launch_timed(rowcol_kernel_batch_wave<GSBN_PREC_F16>, dim3(g), b, 0, s, projs_dev, wg_proj,
                   wg_hcu, nhcu, nhelp, *step);
      break;
  }
  return launch_status();
}

int gsbn_proj_rowcol_graded(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                            const int32_t* wg_hcu, const int32_t* wg_part, int32_t nwg,
                            const gsbn_step_args* step, void* stream) {
  if (!projs_dev || !wg_proj || !wg_hcu || !wg_part || !step || nwg <= 0 || step->kp_slot < 0 ||
      step->kp_slot >= GSBN_KP_SLOTS || step->prec < GSBN_PREC_F32 || step->prec > GSBN_PREC_F16)
    return GSBN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const dim3 b(RC_THREADS);
  switch (step->prec) {
    case GSBN_PREC_F32:
      launch_timed(rowcol_kernel_graded<GSBN_PREC_F32>,
                   dim3(loop_grid(rowcol_kernel_graded<GSBN_PREC_F32>, RC_THREADS, 0, nwg)), b, 0,
                   s, projs_dev, wg_proj, wg_hcu, wg_part, nwg, *step);
      break;
    case GSBN_PREC_BF16:
      launch_timed(rowcol_kernel_graded<GSBN_PREC_BF16>,
                   dim3(loop_grid(rowcol_kernel_graded<GSBN_PREC_BF16>, RC_THREADS, 0, nwg)), b,
                   0, s, projs_dev, wg_proj, wg_hcu, wg_part, nwg, *step);
      break;
    default:
      launch_timed(rowcol_kernel_graded<GSBN_PREC_F16>,
                   dim3(loop_grid(rowcol_kernel_graded<GSBN_PREC_F16>, RC_THREADS, 0, nwg)), b,
                   0, s, projs_dev, wg_proj, wg_hcu, wg_part, nwg, *step);
      break;
  }
  return launch_status();
}

int gsbn_proj_rowcol_graded_wave(const gsbn_rowcol_args* projs_dev, const int32_t* wg_proj,
                                  const int32_t* wg_hcu, const int32_t* wg_part, int32_t nwg,
                                  const gsbn_step_args* step, void* stream) {
  if (!projs_dev || !wg_proj || !wg_hcu || !wg_part || !step || nwg <= 0 || step->kp_slot < 0 ||
      step->kp_slot >= GSBN_KP_SLOTS || step->prec < GSBN_PREC_F32 || step->prec > GSBN_PREC_F16)
    return GSBN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const dim3 b(64);
  switch (step->prec) {
    case GSBN_PREC_F32:
      launch_timed(rowcol_kernel_graded_wave<GSBN_PREC_F32>,
                   dim3(loop_grid(rowcol_kernel_graded_wave<GSBN_PREC_F32>, 64, 0, nwg)), b, 0, s,
                   projs_dev, wg_proj, wg_hcu, wg_part, nwg, *step);
      break;
    case GSBN_PREC_BF16:
      launch_timed(rowcol_kernel_graded_wave<GSBN_PREC_BF16>,
                   dim3(loop_grid(rowcol_kernel_graded_wave<GSBN_PREC_BF16>, 64, 0, nwg)), b, 0,
                   s, projs_dev, wg_proj, wg_hcu, wg_part, nwg, *step);
      break;
    default:
      launch_timed(rowcol_kernel_graded_wave<GSBN_PREC_F16>,
                   dim3(loop_grid(rowcol_kernel_graded_wave<GSBN_PREC_F16>, 64, 0, nwg)), b, 0,
                   s, projs_dev, wg_proj, wg_hcu, wg_part, nwg, *step);
      break;
  }
  return launch_status();
}

int gsbn_step_fused(const gsbn_pop_args* a, const gsbn_rowcol_args* b, void* stream) {
  if (!a || !b || a->H <= 0 || a->M <= 0 || a->M > RC_THREADS || a->nproj_in != 1 || !a->projs ||
      a->max_slots < b->C || a->nproj_out < 0 || (a->nproj_out > 0 && !a->outs) || b->H != a->H ||
      b->M != a->M || b->C < 0 || !b->table || b->table_len < 2 || !b->contrib ||
      !trace_form_ok(*b))
    return GSBN_ERR_ARG;
  if (a->wta_pop_max && !a->scratch_dsup0) return GSBN_ERR_ARG;
  /* the fused row/column phase does eager column updates only */
  if (b->dK || a->proj0.dK) return GSBN_ERR_UNSUPPORTED;
  const size_t lds = pop_lds(RC_THREADS, a->max_slots, true);
  if (lds > 48 * 1024) return GSBN_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  if (a->wta_pop_max) {
    hipLaunchKernelGGL(pop_dsup0_kernel, dim3(1), dim3(64), 0, s, *a);
    int rc = launch_status();
    if (rc) return rc;
  }
  const dim3 g((unsigned)a->H), t(RC_THREADS);
  switch (b->prec) {
    case GSBN_PREC_F32:
      launch_timed(step_kernel<GSBN_PREC_F32>, g, t, lds, s, *a, *b);
      break;
    case GSBN_PREC_BF16:
      launch_timed(step_kernel<GSBN_PREC_BF16>, g, t, lds, s, *a, *b);
      break;
    default:
      launch_timed(step_kernel<GSBN_PREC_F16>, g, t, lds, s, *a, *b);
      break;
  }
  return launch_status();
}

int gsbn_step_chained(const gsbn_pop_args* a, const gsbn_rowcol_args* b, int32_t has_prev,
                      void* stream) {
  if (!a || !b || a->H <= 0 || a->M <= 0 || a->M > RC_THREADS || a->nproj_in != 1 ||
      a->wta_pop_max || a->nproj_out < 0 || (a->nproj_out > 0 && !a->outs))
    return GSBN_ERR_ARG;
  if (has_prev && (b->H != a->H || b->M != a->M || b->C != a->proj0.C || b->C < 0 || !b->table ||
                   b->table_len < 2 || !trace_form_ok(*b) || b->parts < 1))
    return GSBN_ERR_ARG;
  /* the previous step's lists and stamps must not be this step's */
  if (has_prev && (b->spk_list == a->spk_list || b->spk_cnt == a->spk_cnt ||
                   b->act_stamp == a->proj0.act_stamp || !b->rows_list || !b->rows_cnt ||
                   !b->rowst || !b->sj))
    return GSBN_ERR_ARG;
  const size_t lds = pop_lds(RC_THREADS, a->max_slots, false);
  if (lds > 48 * 1024) return GSBN_ERR_UNSUPPORTED;
  const int64_t grid = (int64_t)a->H * (1 + (has_prev ? b->parts : 0));
  if (grid > 0x7fffffff) return GSBN_ERR_UNSUPPORTED;
  const dim3 g((unsigned)grid), t(RC_THREADS);
  hipStream_t s = (hipStream_t)stream;
  const int hp = has_prev ? 1 : 0;
  /* without a previous step the form only selects the instance */
  switch (has_prev ? b->prec : GSBN_PREC_F32) {
    case GSBN_PREC_F32:
      launch_timed(chained_kernel<GSBN_PREC_F32>, g, t, lds, s, *a, *b, hp);
      break;
    case GSBN_PREC_BF16:
      launch_timed(chained_kernel<GSBN_PREC_BF16>, g, t, lds, s, *a, *b, hp);
      break;
    default:
      launch_timed(chained_kernel<GSBN_PREC_F16>, g, t, lds, s, *a, *b, hp);
      break;
  }
  return launch_status();
}

int gsbn_chained_resident(int32_t prec, int32_t* out) {
  if (!out) return GSBN_ERR_ARG;
  int dev = 0, cus = 0, per = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) {
    switch (prec) {
      case GSBN_PREC_F32:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, chained_kernel<GSBN_PREC_F32>, RC_THREADS, 0);
        break;
      case GSBN_PREC_BF16:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, chained_kernel<GSBN_PREC_BF16>, RC_THREADS, 0);
        break;
      default:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, chained_kernel<GSBN_PREC_F16>, RC_THREADS, 0);
        break;
    }
  }
  if (e != hipSuccess) return ret(e);
  *out = per * cus;
  return 0;
}

int gsbn_step_fused_batch(const gsbn_pop_args* pops_dev, int32_t npop,
                          const gsbn_rowcol_args* projs_dev, const int32_t* wg_pop,
                          const int32_t* wg_hcu, int32_t nwg, const int32_t* in_proj,
                          int32_t max_m, int32_t max_slots, int32_t any_wta_pop_max,
                          const gsbn_step_args* step, void* stream) {
  if (!pops_dev || !projs_dev || !wg_pop || !wg_hcu || !in_proj || !step || nwg <= 0 ||
      npop <= 0 || max_m <= 0 || max_m > RC_THREADS || max_slots < 0 ||
      step->kp_slot < 0 || step->kp_slot >= GSBN_KP_SLOTS || step->prec < GSBN_PREC_F32 ||
      step->prec > GSBN_PREC_F16)
    return GSBN_ERR_ARG;
  const size_t lds = pop_lds(RC_THREADS, max_slots, true);
  if (lds > 48 * 1024) return GSBN_ERR_UNSUPPORTED;
  hipStream_t s = (hipStream_t)stream;
  if (any_wta_pop_max) {
    hipLaunchKernelGGL(pop_dsup0_kernel_batch, dim3((npop + 63) / 64), dim3(64), 0, s, pops_dev,
                       npop, *step);
    int rc = launch_status();
    if (rc) return rc;
  }
  const dim3 g((unsigned)nwg), b(RC_THREADS);
  switch (step->prec) {
    case GSBN_PREC_F32:
      launch_timed(step_kernel_batch<GSBN_PREC_F32>, g, b, lds, s, pops_dev, projs_dev, wg_pop,
                   wg_hcu, in_proj, *step);
      break;
    case GSBN_PREC_BF16:
      launch_timed(step_kernel_batch<GSBN_PREC_BF16>, g, b, lds, s, pops_dev, projs_dev, wg_pop,
                   wg_hcu, in_proj, *step);
      break;
    default:
      launch_timed(step_kernel_batch<GSBN_PREC_F16>, g, b, lds, s, pops_dev, projs_dev, wg_pop,
                   wg_hcu, in_proj, *step);
      break;
  }
  return launch_status();
}

int gsbn_proj_que(const gsbn_que_args* a, void* stream) {
  if (!a || a->Hsrc <= 0 || a->Msrc <= 0 || a->exec < 0 || !a->si || !a->out.fwd_off ||
      !a->out.fwd_row || !a->out.fwd_del || !a->out.fut || a->out.H <= 0 || a->out.C <= 0 ||
      a->out.W != (a->out.C + 31) / 32)
    return GSBN_ERR_ARG;
  const int threads = a->Msrc >= 1024 ? 1024 : ((a->Msrc + 63) / 64) * 64;
  launch_timed(que_kernel, dim3(a->Hsrc), dim3(threads), 0, (hipStream_t)stream, *a);
  return launch_status();
}

int gsbn_check_count(int32_t n, const int8_t* spike, int32_t* count, void* stream) {
  if (n < 0 || (n && (!spike || !count))) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(check_count_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0,
                     (hipStream_t)stream, n, spike, count);
  return launch_status();
}

int gsbn_check_argmax(int32_t H, int32_t M, int32_t* count, int32_t* out_idx, int32_t* out_cnt,
                      void* stream) {
  if (H <= 0 || M <= 0 || !count || !out_idx || !out_cnt) return GSBN_ERR_ARG;
  hipLaunchKernelGGL(check_argmax_kernel, dim3((H + 63) / 64), dim3(64), 0, (hipStream_t)stream,
                     H, M, count, out_idx, out_cnt);
  return launch_status();
}

int gsbn_check_count_multi(const gsbn_check_item* items_dev, int32_t nitems, void* stream) {
  if (!items_dev || nitems <= 0) return GSBN_ERR_ARG;
  hipLaunchKernelGGL(check_count_multi_kernel, dim3(nitems), dim3(256), 0, (hipStream_t)stream,
                     items_dev);
  return launch_status();
}

int gsbn_check_argmax_multi(const gsbn_check_item* items_dev, int32_t nitems, int32_t* out_idx,
                            int32_t* out_cnt, void* stream) {
  if (!items_dev || nitems <= 0 || !out_idx || !out_cnt || nitems > 65535) return GSBN_ERR_ARG;
  /* up to 64*16 = 1024 HCUs per population */
  hipLaunchKernelGGL(check_argmax_multi_kernel, dim3(16, nitems), dim3(64), 0,
                     (hipStream_t)stream, items_dev, out_idx, out_cnt);
  return launch_status();
}

int gsbn_syn_field(gsbn_syn_rec* syn, int64_t n, int field, void* buf, int to_records,
                   void* stream) {
  if (n < 0 || field < 0 || field > 5 || (n && (!syn || !buf))) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(syn_field_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, syn, n, field, (uint32_t*)buf, to_records);
  return launch_status();
}

int gsbn_row_field(gsbn_row_rec* row, int64_t n, int field, void* buf, int to_records,
                   void* stream) {
  if (n < 0 || field < 0 || field > 3 || (n && (!row || !buf))) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(row_field_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, row, n, field, (uint32_t*)buf, to_records);
  return launch_status();
}

int gsbn_proj_prune_flags(const gsbn_prune_args* a, void* stream) {
  if (!a || a->R < 0 || a->M <= 0 || a->M > 1024 || a->prec < 0 || a->prec > 2) return GSBN_ERR_ARG;
  if (a->R == 0) return 0;
  if (!a->ii || !a->rowst || !a->flags || (a->prec == GSBN_PREC_F32 ? !a->syn : !a->wij))
    return GSBN_ERR_ARG;
  const int64_t blocks = (a->R + 3) / 4;
  if (blocks > 0x7fffffff) return GSBN_ERR_ARG;
  hipLaunchKernelGGL(prune_flags_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, *a);
  return launch_status();
}

int gsbn_proj_clear_rows(const gsbn_clear_rows_args* a, void* stream) {
  if (!a || a->n < 0 || a->M <= 0 || a->M > 1024 || a->H <= 0 || a->C <= 0 ||
      a->W != (a->C + 31) / 32 || a->prec < 0 || a->prec > 2)
    return GSBN_ERR_ARG;
  if (a->n == 0) return 0;
  if (!a->rows || !a->rowst || !a->fut || (a->prec == GSBN_PREC_F32 ? !a->syn : !a->syn16))
    return GSBN_ERR_ARG;
  const int64_t blocks = (a->n + 3) / 4;
  if (blocks > 0x7fffffff) return GSBN_ERR_ARG;
  hipLaunchKernelGGL(clear_rows_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, *a);
  return launch_status();
}

int gsbn_syn_fill(gsbn_syn_rec* syn, int64_t n, gsbn_syn_rec v, void* stream) {
  if (n < 0 || (n && !syn)) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(syn_fill_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, syn, n, v);
  return launch_status();
}

int gsbn_row_fill(gsbn_row_rec* row, int64_t n, gsbn_row_rec v, void* stream) {
  if (n < 0 || (n && !row)) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(row_fill_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, row, n, v);
  return launch_status();
}

int gsbn_syn16_field(gsbn_syn16_rec* syn, float* wij, int64_t n, int field, void* buf,
                     int to_records, int prec, float p_scale, float e_scale, void* stream) {
  if (n < 0 || field < 0 || field > 5 || prec < GSBN_PREC_BF16 || prec > GSBN_PREC_F16 ||
      (n && (!syn || !buf || (field == 4 && !wij))) ||
      (prec == GSBN_PREC_F16 && (!pow2_scale(p_scale) || !pow2_scale(e_scale))))
    return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(syn16_field_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, syn, wij, n, field, (uint32_t*)buf, to_records, prec,
                     p_scale, e_scale);
  return launch_status();
}

int gsbn_syn16_fill(gsbn_syn16_rec* syn, int64_t n, gsbn_syn16_rec v, void* stream) {
  if (n < 0 || (n && !syn)) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(syn16_fill_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, syn, n, v);
  return launch_status();
}

int gsbn_device_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes < 0 || bytes % 16 || (bytes && (!dst || !src)) ||
      ((uintptr_t)dst | (uintptr_t)src) % 16)
    return GSBN_ERR_ARG;
  if (bytes == 0) return 0;
  /* up to 65536 workgroups: 5.1 TB/s on 2 GiB against 4.6 TB/s with 8192 */
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(bytes / 16, 256, 65536)), dim3(256), 0,
                     (hipStream_t)stream, (uint4*)dst, (const uint4*)src, bytes / 16);
  return launch_status();
}

int gsbn_i32_fill(int32_t* p, int64_t n, int32_t v, void* stream) {
  if (n < 0 || (n && !p)) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(i32_fill_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0,
                     (hipStream_t)stream, p, n, v);
  return launch_status();
}

int gsbn_device_math(int op, int64_t n, const void* in_host, void* out_host) {
  if (op < 0 || op > 7 || n < 0) return GSBN_ERR_ARG;
  if (n == 0) return 0;
  const int in_w = op == 2 ? 4 : op == 3 ? 2 : op == 4 ? 3 : 1;
  const int out_w = op == 2 ? 2 : 1;
  uint32_t *din = nullptr, *dout = nullptr;
  hipError_t e = hipMalloc(&din, (size_t)n * in_w * 4);
  if (e != hipSuccess) return ret(e);
  e = hipMalloc(&dout, (size_t)n * out_w * 4);
  if (e != hipSuccess) {
    (void)hipFree(din);
    return ret(e);
  }
  e = hipMemcpy(din, in_host, (size_t)n * in_w * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(device_math_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, 0, op, n,
                       din, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out_host, dout, (size_t)n * out_w * 4, hipMemcpyDeviceToHost);
  (void)hipFree(din);
  (void)hipFree(dout);
  return ret(e);
}

}  // extern "C"
