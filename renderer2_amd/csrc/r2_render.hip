// Volume integrator kernels for gfx950.
//
// Work item = (sample chunk, pixel).  A persistent grid of wavefronts pulls
// 64-item batches from one global counter (one atomic per wavefront); the 64
// lanes of a batch are one 8x8 pixel tile of one chunk, so primary rays of a
// wavefront are coherent.  Each lane runs the K samples of its chunk in order
// (CPathIntegrator::Li per sample, path_integrator_impl.hpp:120-247) and
// writes the chunk sum; a second kernel adds the chunk sums in chunk order.
// Results are independent of scheduling: every path's random stream is keyed by
// (seed, stream, pixel, sample) and the reduction order is fixed.
#include "r2_internal.hpp"
#include "r2_volume.hpp"
#include <algorithm>

namespace {
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

template <typename Acc>
__global__ void __launch_bounds__(256) k_render(const DScene* __restrict__ sp, uint32_t* __restrict__ counter,
                                                uint64_t total_items, int chunk, Acc* __restrict__ partial) {
  const DScene& s = *sp;
  const uint32_t npix = (uint32_t)s.W * (uint32_t)s.H;
  const int lane = threadIdx.x & 63;
  Cnt cnt;
  while (true) {
    uint64_t base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long*)counter, 64ull);
    base = __shfl(base, 0);
    if (base >= total_items) break;
    const uint64_t item = base + (uint64_t)lane;
    if (item < total_items) {
      const uint32_t c = (uint32_t)(item / npix), q = (uint32_t)(item % npix);
      int x, y;
      item_pixel(q, s.W, s.H, x, y);
      const uint32_t pix = (uint32_t)y * (uint32_t)s.W + (uint32_t)x;
      Acc ax = 0, ay = 0, az = 0;
      const int s0 = (int)c * chunk, s1 = min(s.spp, s0 + chunk);
      for (int smp = s0; smp < s1; ++smp) {
        Rng rng;
        rng.init(s.seed, s.stream, pix, (uint32_t)(s.sample_begin + smp), 0u);
        const v3 L = Li<false>(s, x, y, rng, cnt);
        ax += (Acc)L.x;
        ay += (Acc)L.y;
        az += (Acc)L.z;
      }
      Acc* o = partial + ((size_t)c * npix + pix) * 3;
      o[0] = ax;
      o[1] = ay;
      o[2] = az;
    }
  }
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

__global__ void __launch_bounds__(256) k_paths(const DScene* __restrict__ sp, int n, const int32_t* __restrict__ pix,
                                               const int32_t* __restrict__ smp, float* __restrict__ out,
                                               uint32_t* __restrict__ cnt_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DScene& s = *sp;
  const int p = pix[i];
  const int x = p % s.W, y = p / s.W;
  Rng rng;
  rng.init(s.seed, s.stream, (uint32_t)p, (uint32_t)smp[i], 0u);
  Cnt cnt;
  for (int k = 0; k < kNumCnt; ++k) cnt.c[k] = 0;
  const v3 L = Li<true>(s, x, y, rng, cnt);
  out[3 * i] = L.x;
  out[3 * i + 1] = L.y;
  out[3 * i + 2] = L.z;
  if (cnt_out)
    for (int k = 0; k < kNumCnt; ++k) cnt_out[(size_t)i * kNumCnt + k] = cnt.c[k];
}

int g_num_cus = 0;

}  // namespace

namespace r2host {

int launch_render(r2_ctx* ctx, const r2::DScene& scene, float* d_out, bool fp64) {
  using namespace r2;
  const uint64_t npix = (uint64_t)scene.W * scene.H;
  const uint64_t target = 1ull << 21;
  int chunk = (int)std::min<uint64_t>((uint64_t)scene.spp, std::max<uint64_t>(1, (npix * scene.spp + target - 1) / target));
  const int nchunks = (scene.spp + chunk - 1) / chunk;
  const uint64_t total = (uint64_t)nchunks * npix;
  const size_t acc_bytes = fp64 ? 8 : 4;
  const size_t part_bytes = (size_t)nchunks * npix * 3 * acc_bytes;
  int rc = ensure_scratch(ctx, part_bytes);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  R2_HIP(ctx, hipMemcpyAsync(ctx->d_scene, &scene, sizeof(DScene), hipMemcpyHostToDevice, st));
  R2_HIP(ctx, hipMemsetAsync(ctx->d_counter, 0, 8, st));
  if (!g_num_cus) {
    int v = 0;
    R2_HIP(ctx, hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device));
    g_num_cus = v > 0 ? v : 256;
  }
  const int blocks = (int)std::min<uint64_t>((uint64_t)g_num_cus * 16, (total + 255) / 256);
  R2_HIP(ctx, hipEventRecord(ctx->ev0, st));
  if (fp64)
    hipLaunchKernelGGL(k_render<double>, dim3(blocks), dim3(256), 0, st, ctx->d_scene, ctx->d_counter, total, chunk,
                       (double*)ctx->scratch);
  else
    hipLaunchKernelGGL(k_render<float>, dim3(blocks), dim3(256), 0, st, ctx->d_scene, ctx->d_counter, total, chunk,
                       (float*)ctx->scratch);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipEventRecord(ctx->ev1, st));
  const uint32_t n3 = (uint32_t)(npix * 3);
  if (fp64)
    hipLaunchKernelGGL(k_reduce<double>, dim3((n3 + 255) / 256), dim3(256), 0, st, (const double*)ctx->scratch, nchunks,
                       n3, d_out);
  else
    hipLaunchKernelGGL(k_reduce<float>, dim3((n3 + 255) / 256), dim3(256), 0, st, (const float*)ctx->scratch, nchunks,
                       n3, d_out);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(st));
  float ms = 0.f;
  R2_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->stats.last_kernel_ms = ms;
  ctx->stats.last_paths = (int64_t)(npix * scene.spp);
  ctx->stats.last_launches = 2;
  return R2_OK;
}

int launch_paths(r2_ctx* ctx, const r2::DScene& scene, int n, const int32_t* d_pix, const int32_t* d_smp, float* d_out,
                 uint32_t* d_cnt) {
  hipStream_t st = ctx->stream;
  R2_HIP(ctx, hipMemcpyAsync(ctx->d_scene, &scene, sizeof(r2::DScene), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_paths, dim3((n + 255) / 256), dim3(256), 0, st, ctx->d_scene, n, d_pix, d_smp, d_out, d_cnt);
  R2_HIP(ctx, hipGetLastError());
  R2_HIP(ctx, hipStreamSynchronize(st));
  return R2_OK;
}

}  // namespace r2host
