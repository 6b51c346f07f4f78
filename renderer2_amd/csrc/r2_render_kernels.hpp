// Integrator kernels, included once per record precision: the including file
// defines R2_PREC_MODE (0: FP32 records, 1: packed records) and R2K_NS, the
// namespace of that instantiation.
// Volume integrator kernels for gfx950.
//
// Work item = (sample chunk, pixel), items numbered so that 64 consecutive items
// are one 8x8 pixel tile of one chunk (coherent primary rays).  A persistent
// grid of wavefronts runs the path state machine of r2_pathsm.hpp; each lane
// runs the K samples of its item in order and writes the chunk sum; a second
// kernel adds the chunk sums in chunk order.  Results are independent of
// scheduling: every path's random stream is keyed by (seed, stream, pixel,
// sample) and the reduction order is fixed.
#pragma once
#include "r2_internal.hpp"
#include "r2_pathsm.hpp"
#include <algorithm>
#include <cstdio>

namespace R2K_NS {
using namespace r2;

__device__ __forceinline__ void item_pixel(uint32_t q, int W, int H, int& x, int& y) {
  if ((W & 7) == 0 && (H & 7) == 0) {
    const uint32_t tile = q >> 6, in = q & 63u;
    const uint32_t tx = tile % (uint32_t)(W >> 3), ty = tile / (uint32_t)(W >> 3);
    x = (int)(tx * 8u + (in & 7u));
    y = (int)(ty * 8u + (in >> 3));
  } else {
    x = (int)(q % (uint32_t)W);
    y = (int)(q / (uint32_t)W);
  }
}

// Persistent path state machine (see r2_pathsm.hpp).  Render mode: item =
// (sample chunk, pixel), output = chunk sums.  Probe mode (probe_pix != null):
// item = one (pixel, sample) path, output = its radiance and lookup counters.
#ifndef R2_EV_ITERS
#define R2_EV_ITERS 12    // event passes per scheduling round
#endif
#ifndef R2_TR_MIN_LANES
#define R2_TR_MIN_LANES 16  // leave the traversal round when fewer lanes are still traversing
#endif
#ifndef R2_RNG_TRAV_TOPUP
#define R2_RNG_TRAV_TOPUP 1  // top up the random words of traversing lanes every traversal round
#endif
#ifndef R2_TR_MIN_ROUNDS
#define R2_TR_MIN_ROUNDS 8
#endif
#ifdef R2_SM_STATS
// debug-build scheduling statistics, kept per wavefront in LDS and added to the
// global table once when the wavefront exits: [0] outer iterations, [1] event
// iterations, [2] traversal rounds, [3] lanes active in traversal rounds,
// [4] / [5] wave cycles in event passes / traversal rounds, [16+2s] executions
// of state s, [17+2s] lanes in those executions, [48+s] wave cycles in state s
__device__ unsigned long long g_sm_stats[64];
__device__ __forceinline__ unsigned long long* wave_stats() {
  __shared__ unsigned long long st[4][64];
  return st[(threadIdx.x >> 6) & 3];
}
__device__ __forceinline__ void stats_begin() { wave_stats()[threadIdx.x & 63] = 0ull; }
__device__ __forceinline__ void stats_end() {
  atomicAdd(&g_sm_stats[threadIdx.x & 63], wave_stats()[threadIdx.x & 63]);
}
// leader lane of the active lanes adds v to counter i
__device__ __forceinline__ void stat_add(int i, uint64_t v) {
  const uint64_t m = __ballot(true);
  if ((threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) wave_stats()[i] += v;
}
#define SM_STAT(i, v) stat_add((i), (uint64_t)(v))
__device__ __forceinline__ void stat_cycles(int i, uint64_t c) { stat_add(i, c); }
__device__ __forceinline__ void stat_state(int st) {
  stat_add(16 + 2 * st, 1);
  stat_add(17 + 2 * st, __popcll(__ballot(true)));
}
#else
#define SM_STAT(i, v) do {} while (0)
__device__ __forceinline__ void stat_state(int) {}
__device__ __forceinline__ void stats_begin() {}
__device__ __forceinline__ void stats_end() {}
#endif
// end of one path: radiance, NaN guard, accumulation into the item or probe output
template <typename Acc>
__device__ __forceinline__ void finish_path(const DScene& s, PathLane& p, Cnt& cnt, uint32_t npix, int chunk,
                                            Acc* __restrict__ partial, const int32_t* __restrict__ probe_pix,
                                            float* __restrict__ probe_out, uint32_t* __restrict__ probe_cnt) {
  v3 L = p.L;
  if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z))
    L = mk(0.f, 0.f, 0.f);
  else
    L = L / (float)s.spp;
  if (probe_pix) {
    probe_out[3 * p.item] = L.x;
    probe_out[3 * p.item + 1] = L.y;
    probe_out[3 * p.item + 2] = L.z;
    if (probe_cnt)
      for (int c = 0; c < kNumCnt; ++c) probe_cnt[p.item * kNumCnt + c] = cnt.c[c];
#pragma unroll
    for (int c = 0; c < kNumCnt; ++c) cnt.c[c] = 0;
    p.stage = ST_FETCH;
    return;
  }
  if constexpr (sizeof(Acc) == 8) {
    p.dax += (double)L.x;
    p.day += (double)L.y;
    p.daz += (double)L.z;
  } else {
    p.ax += L.x;
    p.ay += L.y;
    p.az += L.z;
  }
  ++p.s_cur;
  if (p.s_cur < p.s_end) {
    p.stage = ST_START;
  } else {
    const uint32_t c = (uint32_t)(p.item / npix);
    Acc* o = partial + ((size_t)c * npix + p.pix) * 3;
    if constexpr (sizeof(Acc) == 8) {
      o[0] = p.dax;
      o[1] = p.day;
      o[2] = p.daz;
    } else {
      o[0] = p.ax;
      o[1] = p.ay;
      o[2] = p.az;
    }
    p.stage = ST_FETCH;
  }
}

// Run state kSt for the lanes that are in it, or in kSt2 when the code of kSt
// also serves kSt2 (a divergent branch the wave skips when no lane is).
template <bool kSingle, bool kCount, int kSt, int kSt2 = -1>
__device__ __forceinline__ void event_if(const DScene& s, PathLane& p, Cnt& cnt) {
  if (p.stage == kSt || p.stage == kSt2) {
#ifdef R2_SM_STATS
    stat_state(kSt);
    const uint64_t t0 = clock64();
    event_step<kSingle, kCount>(s, p, cnt, kSt);
    stat_cycles(48 + kSt, clock64() - t0);
#else
    event_step<kSingle, kCount>(s, p, cnt, kSt);
#endif
  }
}

// One scheduling round of a wavefront: shading events, batched across the
// wavefront, then traversal rounds in which every traversing lane advances one
// DDA / tracking step.
template <bool kSingle, bool kCount, typename Acc>
__device__ __forceinline__ void run_round(const DScene& s, PathLane& p, Cnt& cnt, int lane, uint32_t npix, int chunk,
                                          Acc* __restrict__ partial, const int32_t* __restrict__ probe_pix,
                                          float* __restrict__ probe_out, uint32_t* __restrict__ probe_cnt) {
  // ---- shading events, batched across the wavefront
#ifdef R2_SM_STATS
  const uint64_t te0 = clock64();
#endif
  // pass-through: the states run in the forward order of the path's state graph,
  // so within one pass a lane advances through every state that follows its own
  // in that order (e.g. SH_NEXT -> BRDF_DONE -> DIRECT_DONE -> MOVE -> BOUNCE ->
  // a new traversal), and each state's code runs at most once per pass for the
  // wave; states whose code shares its expensive part run as one block (BOUNCE /
  // SH_SEG: scene intersection + traversal set-up; NOHIT / BRDF_DONE: environment
  // lookup; BRDF_SAMPLE / DIRECT_DONE: phase sample; ST_MOVE: the voxel-border
  // move of every ray spawned at a scatter point)
  for (int k = 0; k < R2_EV_ITERS; ++k) {
    if (!__any(p.stage >= ST_START)) break;
    SM_STAT(1, 1);
    // random words for this pass (a lane in ST_START starts a new stream)
    if (p.stage > ST_START && p.stage != ST_FINISH) p.rng.top_up();
    if (p.stage == ST_FINISH) {
      stat_state(ST_FINISH);
#ifdef R2_SM_STATS
      const uint64_t t0 = clock64();
#endif
      finish_path<Acc>(s, p, cnt, npix, chunk, partial, probe_pix, probe_out, probe_cnt);
#ifdef R2_SM_STATS
      stat_cycles(48 + ST_FINISH, clock64() - t0);
#endif
    }
    event_if<kSingle, kCount, ST_START>(s, p, cnt);
    event_if<kSingle, kCount, ST_SH_NEXT>(s, p, cnt);
    event_if<kSingle, kCount, ST_AFTER_SAMPLE>(s, p, cnt);
    event_if<kSingle, kCount, ST_NOHIT, ST_BRDF_DONE>(s, p, cnt);
    event_if<kSingle, kCount, ST_LIGHT_DONE>(s, p, cnt);
    event_if<kSingle, kCount, ST_SCATTER>(s, p, cnt);
    event_if<kSingle, kCount, ST_BRDF_SAMPLE, ST_DIRECT_DONE>(s, p, cnt);
    event_if<kSingle, kCount, ST_MOVE>(s, p, cnt);
    event_if<kSingle, kCount, ST_BOUNCE, ST_SH_SEG>(s, p, cnt);
  }
#ifdef R2_SM_STATS
  SM_STAT(4, clock64() - te0);
#endif
  // ---- traversal rounds: every traversing lane advances one DDA / tracking step
#ifdef R2_SM_STATS
  const uint64_t tt0 = clock64();
#endif
  for (int k = 0; k < 256; ++k) {
    const uint64_t tm = __ballot(p.stage == ST_TRAV);
    if (tm == 0) break;
    if (k >= R2_TR_MIN_ROUNDS && __popcll(tm) < R2_TR_MIN_LANES) break;
    SM_STAT(2, 1);
    SM_STAT(3, __popcll(tm));
#if R2_RNG_TRAV_TOPUP
    if (p.stage == ST_TRAV) p.rng.top_up();
#endif
    if (p.stage == ST_TRAV) {
      if (traversal_step<kSingle, kCount>(s, p, cnt)) end_traversal(p);
    }
  }
#ifdef R2_SM_STATS
  SM_STAT(5, clock64() - tt0);
#endif
}

#ifndef R2_SM_WAVES
#define R2_SM_WAVES 3  // wavefronts per SIMD the register budget is sized for (measured best)
#endif
template <bool kSingle, bool kCount, typename Acc>
__global__ void __launch_bounds__(256, R2_SM_WAVES) k_paths_sm(const DScene* __restrict__ sp, uint64_t* __restrict__ counter,
                                                  uint64_t total, int chunk, Acc* __restrict__ partial,
                                                  const int32_t* __restrict__ probe_pix,
                                                  const int32_t* __restrict__ probe_smp, float* __restrict__ probe_out,
                                                  uint32_t* __restrict__ probe_cnt) {
  const DScene& s = kconst(sp);
  const int lane = threadIdx.x & 63;
  const uint32_t npix = (uint32_t)s.W * (uint32_t)s.H;
  __shared__ ColdLane cold[256];
  PathLane p(cold[threadIdx.x], s);
  p.stage = ST_FETCH;
  p.item = 0;
  Cnt cnt;
#pragma unroll
  for (int k = 0; k < kNumCnt; ++k) cnt.c[k] = 0;
  stats_begin();
  for (int guard = 0; guard < (1 << 24); ++guard) {
    // ---- work fetch: one atomic per wavefront, items handed out by prefix rank
    const uint64_t need = __ballot(p.stage == ST_FETCH);
    if (need) {
      const int n = __popcll(need);
      const int leader = __ffsll((unsigned long long)need) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)n);
      base = __shfl(base, leader);
      if (p.stage == ST_FETCH) {
        const uint64_t item = base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
        if (item < total) {
          p.item = item;
          p.ax = p.ay = p.az = 0.f;
          p.dax = p.day = p.daz = 0.0;
          if (probe_pix) {
            p.pix = (uint32_t)probe_pix[item];
            p.s_cur = probe_smp[item];
            p.s_end = p.s_cur + 1;
          } else {
            // items of a launch fit 32 bits (the lane keeps them as 32-bit values)
            const uint32_t it = (uint32_t)item, c = it / npix, q = it - c * npix;
            int x, y;
            item_pixel(q, s.W, s.H, x, y);
            p.pix = (uint32_t)y * (uint32_t)s.W + (uint32_t)x;
            p.s_cur = (int)c * chunk;
            p.s_end = min(s.spp, p.s_cur + chunk);
          }
          p.stage = ST_START;
        } else {
          p.stage = ST_IDLE;
        }
      }
    }
    if (__ballot(p.stage != ST_IDLE) == 0) break;
    SM_STAT(0, 1);
    run_round<kSingle, kCount, Acc>(s, p, cnt, lane, npix, chunk, partial, probe_pix, probe_out, probe_cnt);
  }
  stats_end();
}

// Batched form for the render server: one launch serves several requests, each
// an entry with its own scene, item range and chunk-sum buffer.  A wavefront
// serves one entry at a time (the scene stays wave-uniform); when the entry is
// exhausted and all its lanes are idle it moves to the next entry.  Per item the
// computation is the one of k_paths_sm, so results do not depend on the mix.

template <bool kSingle, typename Acc>
__global__ void __launch_bounds__(256, R2_SM_WAVES) k_paths_batch(const BatchEntry* __restrict__ ents, int n_ent,
                                                                  uint64_t* __restrict__ counters) {
  const int lane = threadIdx.x & 63;
  const int e0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * 4u + (threadIdx.x >> 6)) % (unsigned)n_ent));
  __shared__ ColdLane cold[256];
  Cnt cnt;
#pragma unroll
  for (int k = 0; k < kNumCnt; ++k) cnt.c[k] = 0;
  stats_begin();
  // one pass over the entries: an entry is left only when none of its items is
  // left to hand out, so it never needs a second visit
#pragma unroll 1
  for (int v = 0; v < n_ent; ++v) {
    const int e = __builtin_amdgcn_readfirstlane(e0 + v < n_ent ? e0 + v : e0 + v - n_ent);
    const BatchEntry& E = (&kconst(ents))[e];
    const DScene& s = kconst(E.scene);
    const uint32_t npix = (uint32_t)s.W * (uint32_t)s.H;
    const int chunk = E.chunk;
    Acc* partial = (Acc*)E.partial;
    PathLane p(cold[threadIdx.x], s);
    p.stage = ST_FETCH;
    p.item = 0;
    for (int guard = 0; guard < (1 << 24); ++guard) {
      const uint64_t need = __ballot(p.stage == ST_FETCH);
      if (need) {
        const int n = __popcll(need);
        const int leader = __ffsll((unsigned long long)need) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd((unsigned long long*)&counters[e], (unsigned long long)n);
        base = __shfl(base, leader);
        if (p.stage == ST_FETCH) {
          const uint64_t local = base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
          if (local < E.count) {
            const uint64_t item = E.first + local;
            p.item = item;
            p.ax = p.ay = p.az = 0.f;
            p.dax = p.day = p.daz = 0.0;
            // items of a launch fit 32 bits (the lane keeps them as 32-bit values)
            const uint32_t it = (uint32_t)item, c = it / npix, q = it - c * npix;
            int x, y;
            item_pixel(q, s.W, s.H, x, y);
            p.pix = (uint32_t)y * (uint32_t)s.W + (uint32_t)x;
            p.s_cur = (int)c * chunk;
            p.s_end = min(s.spp, p.s_cur + chunk);
            p.stage = ST_START;
          } else {
            p.stage = ST_IDLE;
          }
        }
      }
      if (__ballot(p.stage != ST_IDLE) == 0) break;
      run_round<kSingle, false, Acc>(s, p, cnt, lane, npix, chunk, partial, nullptr, nullptr, nullptr);
    }
  }
  stats_end();
}

// out[p] = sum over chunks in chunk order (FP32), or a Neumaier-compensated
// FP64 sum of FP64 chunk sums, rounded once to FP32.
template <typename Acc>
__global__ void k_reduce(const Acc* __restrict__ partial, int nchunks, uint32_t n3, float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  if constexpr (sizeof(Acc) == 4) {
    float acc = 0.f;
    for (int c = 0; c < nchunks; ++c) acc += partial[(size_t)c * n3 + i];
    out[i] = acc;
  } else {
    double sum = 0.0, comp = 0.0;
    for (int c = 0; c < nchunks; ++c) {
      const double v = partial[(size_t)c * n3 + i];
      const double t = sum + v;
      comp += (fabs(sum) >= fabs(v)) ? (sum - t) + v : (v - t) + sum;
      sum = t;
    }
    out[i] = (float)(sum + comp);
  }
}



// host-side launchers of this instantiation
void launch_sm(hipStream_t st, int blocks, const DScene* d_scene, uint64_t* counter, uint64_t total, int chunk,
               void* partial, bool fp64, bool single, bool count, const int32_t* pp, const int32_t* ps, float* po,
               uint32_t* pc) {
#define R2K_LAUNCH(S, C, A)                                                                                        \
  hipLaunchKernelGGL((k_paths_sm<S, C, A>), dim3(blocks), dim3(256), 0, st, d_scene, counter, total, chunk, (A*)partial, \
                     pp, ps, po, pc)
  if (count) {
    if (single) R2K_LAUNCH(true, true, float);
    else R2K_LAUNCH(false, true, float);
  } else if (fp64) {
    if (single) R2K_LAUNCH(true, false, double);
    else R2K_LAUNCH(false, false, double);
  } else {
    if (single) R2K_LAUNCH(true, false, float);
    else R2K_LAUNCH(false, false, float);
  }
#undef R2K_LAUNCH
}

void launch_batch_group(hipStream_t st, int blocks, const void* ents, int n, uint64_t* counters, bool single,
                        bool fp64) {
  const BatchEntry* de = (const BatchEntry*)ents;
  if (single && fp64)
    hipLaunchKernelGGL((k_paths_batch<true, double>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else if (single)
    hipLaunchKernelGGL((k_paths_batch<true, float>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else if (fp64)
    hipLaunchKernelGGL((k_paths_batch<false, double>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else
    hipLaunchKernelGGL((k_paths_batch<false, float>), dim3(blocks), dim3(256), 0, st, de, n, counters);
}

void launch_reduce_k(hipStream_t st, const void* partial, bool fp64, int nchunks, uint32_t n3, float* out) {
  if (fp64)
    hipLaunchKernelGGL(k_reduce<double>, dim3((n3 + 255) / 256), dim3(256), 0, st, (const double*)partial, nchunks, n3,
                       out);
  else
    hipLaunchKernelGGL(k_reduce<float>, dim3((n3 + 255) / 256), dim3(256), 0, st, (const float*)partial, nchunks, n3,
                       out);
}

}  // namespace R2K_NS
