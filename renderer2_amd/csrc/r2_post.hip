// Post-process of a rendered frame (rt::Raytracer::renderFrame after the spp
// loop, raytracer/raytracer.cu:53-127,192-206): global log-average tonemapping
// c / (c + k), gamma and bytes.  The column and row sums run in the same order
// as the reference kernels (one thread per column over the rows, then one
// thread over the columns); log / exp / pow are the double-evaluated forms of
// r2_detmath.hpp so that the result does not depend on the device library.
#include "r2_internal.hpp"
#include "r2_detmath.hpp"
#include <cfloat>

namespace {

using namespace r2;

// avg[x] = sum_y log(r + g + b + FLT_MIN) / (W H 3)  (computeGlobalTonemapping1)
__global__ void k_tm_columns(const float* __restrict__ hdr, int w, int h, float* __restrict__ avg) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= w) return;
  const float divisor = (float)(h * w * 3);
  float a = 0.f;
  for (int y = 0; y < h; ++y) {
    const float* p = hdr + 3 * ((size_t)y * w + x);
    a += det_logf(p[0] + p[1] + p[2] + FLT_MIN) / divisor;
  }
  avg[x] = a;
}

// k = exp(sum_x avg[x]) - W H FLT_MIN  (computeGlobalTonemapping2)
__global__ void k_tm_factor(const float* __restrict__ avg, int w, int h, float* __restrict__ factor) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float r = 0.f;
  for (int x = 0; x < w; ++x) r += avg[x];
  *factor = det_expf(r) - (float)(w * h) * FLT_MIN;
}

// applyTonemapping, correctGamma, fillByteFrame
__global__ void k_tm_apply(const float* __restrict__ hdr, size_t n3, const float* __restrict__ factor, float inv_gamma,
                           float* __restrict__ ldr, uint8_t* __restrict__ bytes) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const float k = *factor;
  const float c = hdr[i];
  const float t = det_powf(c / (c + k), inv_gamma);
  if (ldr) ldr[i] = t;
  if (bytes) bytes[i] = (uint8_t)roundf(t * 255.0f);
}

}  // namespace

extern "C" int r2_postprocess(r2_ctx* ctx, const float* hdr, int32_t w, int32_t h, float gamma, float* ldr,
                              uint8_t* bytes, int32_t on_device) {
  if (!ctx || !hdr || w <= 0 || h <= 0 || !(gamma > 0.f)) return R2_ERR_INVALID;
  if ((int64_t)w * h > (1ll << 28)) return r2host::fail(ctx, R2_ERR_UNSUPPORTED, "postprocess: frame too large");
  hipStream_t st = ctx->stream;
  const size_t n3 = (size_t)w * h * 3;
  void* buf = nullptr;
  // scratch: column sums, factor, and (host mode) the frame in and out
  const size_t need = sizeof(float) * ((size_t)w + 64) + (on_device ? 0 : n3 * (4 + 4 + 1) + 512);
  if (r2host::pool_get(ctx, 23, need, &buf) != hipSuccess)
    return r2host::fail(ctx, R2_ERR_OOM, "postprocess: scratch");
  float* avg = (float*)buf;
  float* factor = avg + w;
  const float* d_hdr = hdr;
  float* d_ldr = ldr;
  uint8_t* d_bytes = bytes;
  if (!on_device) {
    float* base = factor + 64;
    float* in = base;
    float* outf = base + n3;
    uint8_t* outb = (uint8_t*)(outf + n3);
    R2_HIP(ctx, hipMemcpyAsync(in, hdr, n3 * 4, hipMemcpyHostToDevice, st));
    d_hdr = in;
    d_ldr = ldr ? outf : nullptr;
    d_bytes = bytes ? outb : nullptr;
  }
  hipLaunchKernelGGL(k_tm_columns, dim3((w + 127) / 128), dim3(128), 0, st, d_hdr, w, h, avg);
  hipLaunchKernelGGL(k_tm_factor, dim3(1), dim3(64), 0, st, avg, w, h, factor);
  hipLaunchKernelGGL(k_tm_apply, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, d_hdr, n3, factor,
                     1.f / gamma, d_ldr, d_bytes);
  R2_HIP(ctx, hipGetLastError());
  if (!on_device) {
    if (ldr) R2_HIP(ctx, hipMemcpyAsync(ldr, d_ldr, n3 * 4, hipMemcpyDeviceToHost, st));
    if (bytes) R2_HIP(ctx, hipMemcpyAsync(bytes, d_bytes, n3, hipMemcpyDeviceToHost, st));
  }
  R2_HIP(ctx, hipStreamSynchronize(st));
  return R2_OK;
}
