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

// Persistent two-phase path program (see r2_pathsm.hpp).  Render mode: item =
// (sample chunk, pixel), output = chunk sums.  Probe mode (probe_pix != null):
// item = one (pixel, sample) path, output = its radiance and lookup counters.
#ifdef R2_SM_STATS
// debug-build scheduling statistics, kept per wavefront in LDS and added to the
// global table once when the wavefront exits: [0] scheduling steps, [1] / [2]
// shading passes and their lanes, [3] / [4] segment passes and their lanes,
// [7] / [8] traversal steps and their lanes, [5] / [6] / [9] wave cycles in
// shading / segment / traversal steps
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
#define SM_CLOCK() clock64()
#else
#define SM_STAT(i, v) do {} while (0)
#define SM_CLOCK() 0ull
__device__ __forceinline__ void stats_begin() {}
__device__ __forceinline__ void stats_end() {}
#endif

// end of one path: NaN guard, radiance / spp, accumulation into the item (or the
// probe output); the lane starts the item's next sample, or fetches a new item
template <typename Acc>
__device__ __forceinline__ void finish_path(const DScene& s, PathLane& p, Cnt& cnt, uint32_t npix, int chunk,
                                            Acc* __restrict__ partial, const int32_t* __restrict__ probe_pix,
                                            float* __restrict__ probe_out, uint32_t* __restrict__ probe_cnt) {
  v3 L = p.L;
  if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z))
    L = mk(0.f, 0.f, 0.f);
  else
    L = L / (float)s.spp;
  if (probe_pix) {  // one path per item
    const uint32_t item = __float_as_uint(p.ax);
    probe_out[3 * item] = L.x;
    probe_out[3 * item + 1] = L.y;
    probe_out[3 * item + 2] = L.z;
    if (probe_cnt)
      for (int c = 0; c < kNumCnt; ++c) probe_cnt[item * kNumCnt + c] = cnt.c[c];
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
  // the item is sample chunk c of the pixel: samples [c chunk, min(spp, (c + 1) chunk))
  const uint32_t smp = p.rng.c1 - (uint32_t)s.sample_begin, c = smp / (uint32_t)chunk;
  if (smp + 1u < min((uint32_t)s.spp, (c + 1u) * (uint32_t)chunk)) {
    ++p.rng.c1;
    p.cres = CR_START;
  } else {
    Acc* o = partial + ((size_t)c * npix + p.rng.c0) * 3;
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

// One shading pass of a lane at a vertex (ST_SHADE): the direct light of the
// previous vertex, then the end of the ray that arrived (a collision becomes the
// next scatter vertex, a surface hit the next surface vertex; an escaped camera
// ray sees the environment; anything else ends the path), and for a path that
// ends the start of the item's next sample (vertex 0: the camera ray).
template <bool kSingle, bool kSurf, bool kCount, typename Acc>
__device__ __forceinline__ void shade(const DScene& s, PathLane& p, Cnt& cnt, uint32_t npix, int chunk,
                                      Acc* __restrict__ partial, const int32_t* __restrict__ probe_pix,
                                      float* __restrict__ probe_out, uint32_t* __restrict__ probe_cnt) {
  bool scatter = false;
  Scatter sc;
  if (p.cres == CR_COLLIDED) {
    p.thr = p.thr * flight_collision<kSingle, kCount>(s, p, sc, cnt);
    scatter = !any_zero(p.thr);
  } else if (p.cres == CR_ESCAPED) {
    float lpdf = 0.f;
    const v3 Le = env_le<kCount>(s.env, p.ray.d, lpdf, cnt);
    p.L = p.L + Le * p.thr;
  }
  // every path out of a shading pass starts its walks with a segment
  dead_traversal(p);
  dead(p.sh_seg);
  dead(p.sh_Tr);
  dead(p.si_hit);
  if (scatter) {
    shade_scatter<kSingle, kCount>(s, p, sc, cnt);
    return;
  }
  if (kSurf && p.cres == CR_SURFACE) {
    shade_surface<kCount>(s, p, cnt);
    return;
  }
  if (p.cres != CR_START) {
    finish_path<Acc>(s, p, cnt, npix, chunk, partial, probe_pix, probe_out, probe_cnt);
    if (p.stage == ST_FETCH) return;
  }
  // vertex 0 of the item's next sample: the camera ray
  dead_walk(p);
  p.rng.init(p.rng.c0, p.rng.c1, 0u);
  const Ray r = camera_ray(s, (int)(p.rng.c0 % (uint32_t)s.W), (int)(p.rng.c0 / (uint32_t)s.W), p.rng);
  p.c.c_o = r.o;
  p.c.c_d = r.d;
  p.c.c_tmax = r.tmax;
  p.c.c_med = r.medium;
  p.L = mk(0.f, 0.f, 0.f);
  p.thr = mk(1.f, 1.f, 1.f);
  p.b = 0;
  p.iters = 0;
  p.k = 0;
  p.c.flags = FL_CONT;
  p.cres = CR_NONE;
  walk_begin(p, RL_CONT);
}

// Work fetch of the standalone kernel: one atomic per wavefront, items handed
// out by prefix rank; a lane that gets an item starts its first path at the next
// shading pass.
struct FrameFetch {
  uint64_t* counter;
  uint64_t total;
  int chunk;
  const int32_t* probe_pix;
  const int32_t* probe_smp;
  __device__ __forceinline__ int chunk_len() const { return chunk; }
  __device__ __forceinline__ void operator()(const DScene& s, PathLane& p, int lane, uint32_t npix) const {
    const uint64_t need = __ballot(p.stage == ST_FETCH);
    if (!need) return;
    const int n = __popcll(need);
    const int leader = __ffsll((unsigned long long)need) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)n);
    base = __shfl(base, leader);
    if (p.stage != ST_FETCH) return;
    const uint64_t item = base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
    if (item >= total) {
      p.stage = ST_IDLE;
      return;
    }
    p.ax = p.ay = p.az = 0.f;
    p.dax = p.day = p.daz = 0.0;
    if (probe_pix) {
      p.rng.c0 = (uint32_t)probe_pix[item];
      p.rng.c1 = (uint32_t)(s.sample_begin + probe_smp[item]);
      p.ax = __uint_as_float((uint32_t)item);
    } else {
      // items of a launch fit 32 bits (the lane keeps them as 32-bit values)
      const uint32_t it = (uint32_t)item, c = it / npix, q = it - c * npix;
      int x, y;
      item_pixel(q, s.W, s.H, x, y);
      p.rng.c0 = (uint32_t)y * (uint32_t)s.W + (uint32_t)x;
      p.rng.c1 = (uint32_t)s.sample_begin + c * (uint32_t)chunk;
    }
    p.c.flags = 0u;
    p.cres = CR_START;
    p.stage = ST_SHADE;
  }
};

// One scheduling step of a wavefront.  Its lanes wait at one of three blocks of
// code: shading (a vertex, or a new item to fetch), the start of a walk segment,
// or a traversal step.  The wavefront runs one block per step for the lanes
// waiting at it, chosen from ballots so that the costly blocks run with many
// lanes: shading once at least R2_SHADE_MIN lanes wait for it (or nothing else
// is left), followed by segment starts for the lanes at them; else segment
// starts once at least R2_SEG_MIN lanes, or R2_SEG_RATIO times as many lanes as
// are in traversals, wait for them; else a traversal step.  Returns false when
// every lane is idle.
#ifndef R2_SHADE_MIN
#define R2_SHADE_MIN 48
#endif
// Issue priority of a wavefront (s_setprio) while it runs a block: the short,
// load-bound traversal steps above segment starts above shading, so that a
// SIMD's shading wavefronts fill the cycles in which the others wait on memory
// (+1.7 % at C2 over equal priorities; shading first measured 1 % slower).
constexpr short kPrioTrav = 2, kPrioSeg = 1;
#ifndef R2_SEG_MIN
#define R2_SEG_MIN 32
#endif
#ifndef R2_SEG_RATIO
#define R2_SEG_RATIO 2
#endif
template <bool kSingle, bool kSurf, bool kCount, typename Acc, typename Fetch>
__device__ __forceinline__ bool sched_step(const DScene& s, PathLane& p, Cnt& cnt, int lane, uint32_t npix,
                                           Acc* __restrict__ partial, const int32_t* __restrict__ probe_pix,
                                           float* __restrict__ probe_out, uint32_t* __restrict__ probe_cnt,
                                           const Fetch& fetch) {
  const bool at_shade = p.stage == ST_SHADE || p.stage == ST_FETCH;
  const bool at_seg = p.stage == ST_SEG;
  const bool at_trav = p.stage == ST_TRAV;
  const int nh = __popcll(__ballot(at_shade)), ns = __popcll(__ballot(at_seg)), nt = __popcll(__ballot(at_trav));
  if (nh + ns + nt == 0) return false;
#ifdef R2_SM_STATS
  const uint64_t t0 = SM_CLOCK();
#endif
  if (nh >= R2_SHADE_MIN || ns + nt == 0) {
    fetch(s, p, lane, npix);
    SM_STAT(1, 1);
    SM_STAT(2, __popcll(__ballot(p.stage == ST_SHADE)));
    if (p.stage == ST_SHADE)
      shade<kSingle, kSurf, kCount, Acc>(s, p, cnt, npix, fetch.chunk_len(), partial, probe_pix, probe_out, probe_cnt);
#ifdef R2_SM_STATS
    SM_STAT(5, SM_CLOCK() - t0);
#endif
    // most shaded lanes now start a walk: their first segment runs in the same
    // step, together with the lanes that were already waiting at a segment
    SM_STAT(3, 1);
    SM_STAT(4, __popcll(__ballot(p.stage == ST_SEG)));
    __builtin_amdgcn_s_setprio(kPrioSeg);
    if (p.stage == ST_SEG) {
      p.rng.top_up();
      seg_step<kSingle, kSurf>(s, p);
    }
    __builtin_amdgcn_s_setprio(0);
  } else if (ns >= R2_SEG_MIN || ns >= R2_SEG_RATIO * nt) {
    SM_STAT(3, 1);
    SM_STAT(4, ns);
    __builtin_amdgcn_s_setprio(kPrioSeg);
    if (at_seg) {
      p.rng.top_up();
      seg_step<kSingle, kSurf>(s, p);
    }
    __builtin_amdgcn_s_setprio(0);
#ifdef R2_SM_STATS
    SM_STAT(6, SM_CLOCK() - t0);
#endif
  } else {
    SM_STAT(7, 1);
    SM_STAT(8, nt);
    __builtin_amdgcn_s_setprio(kPrioTrav);
    if (at_trav) {
      p.rng.top_up();
      trav_step<kSingle, kCount>(s, p, cnt);
    }
    __builtin_amdgcn_s_setprio(0);
#ifdef R2_SM_STATS
    SM_STAT(9, SM_CLOCK() - t0);
#endif
  }
  return true;
}

#ifndef R2_SM_WAVES
#define R2_SM_WAVES 4  // wavefronts per SIMD the register budget is sized for
#endif
template <bool kSingle, bool kSurf, bool kCount, typename Acc>
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
  p.cres = CR_START;
  Cnt cnt;
#pragma unroll
  for (int k = 0; k < kNumCnt; ++k) cnt.c[k] = 0;
  const FrameFetch fetch{counter, total, chunk, probe_pix, probe_smp};
  stats_begin();
  for (uint32_t guard = 0; guard < (1u << 30); ++guard) {
    SM_STAT(0, 1);
    if (!sched_step<kSingle, kSurf, kCount, Acc>(s, p, cnt, lane, npix, partial, probe_pix, probe_out, probe_cnt, fetch))
      break;
  }
  stats_end();
}

// Batched form for the render server: one launch serves several requests, each
// an entry with its own scene, item range and chunk-sum buffer.  A wavefront
// serves one entry at a time (the scene stays wave-uniform); when the entry is
// exhausted and all its lanes are idle it moves to the next entry.  Per item the
// computation is the one of k_paths_sm, so results do not depend on the mix.
struct BatchFetch {
  const BatchEntry* e;
  uint64_t* counter;
  __device__ __forceinline__ int chunk_len() const { return e->chunk; }
  __device__ __forceinline__ void operator()(const DScene& s, PathLane& p, int lane, uint32_t npix) const {
    const uint64_t need = __ballot(p.stage == ST_FETCH);
    if (!need) return;
    const int n = __popcll(need);
    const int leader = __ffsll((unsigned long long)need) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd((unsigned long long*)counter, (unsigned long long)n);
    base = __shfl(base, leader);
    if (p.stage != ST_FETCH) return;
    const uint64_t local = base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
    if (local >= e->count) {
      p.stage = ST_IDLE;
      return;
    }
    const uint64_t item = e->first + local;
    p.ax = p.ay = p.az = 0.f;
    p.dax = p.day = p.daz = 0.0;
    // items of a launch fit 32 bits (the lane keeps them as 32-bit values)
    const uint32_t it = (uint32_t)item, c = it / npix, q = it - c * npix;
    int x, y;
    item_pixel(q, s.W, s.H, x, y);
    p.rng.c0 = (uint32_t)y * (uint32_t)s.W + (uint32_t)x;
    p.rng.c1 = (uint32_t)s.sample_begin + c * (uint32_t)e->chunk;
    p.c.flags = 0u;
    p.cres = CR_START;
    p.stage = ST_SHADE;
  }
};

template <bool kSingle, bool kSurf, typename Acc>
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
    Acc* partial = (Acc*)E.partial;
    PathLane p(cold[threadIdx.x], s);
    p.stage = ST_FETCH;
    p.cres = CR_START;
    const BatchFetch fetch{&E, &counters[e]};
    for (uint32_t guard = 0; guard < (1u << 30); ++guard) {
      if (!sched_step<kSingle, kSurf, false, Acc>(s, p, cnt, lane, npix, partial, nullptr, nullptr, nullptr, fetch))
        break;
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



// host-side launchers of this instantiation.  Scenes with analytic surfaces run
// the general-instance form with the surface code; the others the forms without.
void launch_sm(hipStream_t st, int blocks, const DScene* d_scene, uint64_t* counter, uint64_t total, int chunk,
               void* partial, bool fp64, bool single, bool surf, bool count, const int32_t* pp, const int32_t* ps,
               float* po, uint32_t* pc) {
#define R2K_LAUNCH(S, F, C, A)                                                                                  \
  hipLaunchKernelGGL((k_paths_sm<S, F, C, A>), dim3(blocks), dim3(256), 0, st, d_scene, counter, total, chunk, \
                     (A*)partial, pp, ps, po, pc)
  if (surf) {
    if (count) R2K_LAUNCH(false, true, true, float);
    else if (fp64) R2K_LAUNCH(false, true, false, double);
    else R2K_LAUNCH(false, true, false, float);
  } else if (count) {
    if (single) R2K_LAUNCH(true, false, true, float);
    else R2K_LAUNCH(false, false, true, float);
  } else if (fp64) {
    if (single) R2K_LAUNCH(true, false, false, double);
    else R2K_LAUNCH(false, false, false, double);
  } else {
    if (single) R2K_LAUNCH(true, false, false, float);
    else R2K_LAUNCH(false, false, false, float);
  }
#undef R2K_LAUNCH
}

void launch_batch_group(hipStream_t st, int blocks, const void* ents, int n, uint64_t* counters, bool single,
                        bool surf, bool fp64) {
  const BatchEntry* de = (const BatchEntry*)ents;
  if (surf && fp64)
    hipLaunchKernelGGL((k_paths_batch<false, true, double>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else if (surf)
    hipLaunchKernelGGL((k_paths_batch<false, true, float>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else if (single && fp64)
    hipLaunchKernelGGL((k_paths_batch<true, false, double>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else if (single)
    hipLaunchKernelGGL((k_paths_batch<true, false, float>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else if (fp64)
    hipLaunchKernelGGL((k_paths_batch<false, false, double>), dim3(blocks), dim3(256), 0, st, de, n, counters);
  else
    hipLaunchKernelGGL((k_paths_batch<false, false, float>), dim3(blocks), dim3(256), 0, st, de, n, counters);
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
