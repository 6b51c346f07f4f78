/*
 * Oracle (test infrastructure only): CPU restatement of the frame post-process of
 * FrederikBoehm/Renderer2 -- global log-average tonemapping, gamma and bytes
 * (raytracer/raytracer.cu:53-127: computeGlobalTonemapping1/2, applyTonemapping,
 * correctGamma, fillByteFrame; called at raytracer.cu:192-206).  The column sums
 * run over rows in order and the row of column sums in order, as the reference's
 * kernels do; log/exp/pow are the double-evaluated r2o_det_* functions.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include "r2o.h"

void r2o_postprocess(const float* hdr, int w, int h, float gamma, float* ldr, uint8_t* bytes) {
  const float divisor = (float)(h * w * 3);
  float* avg = (float*)malloc(sizeof(float) * (size_t)w);
  for (int x = 0; x < w; ++x) {
    float a = 0.f;
    for (int y = 0; y < h; ++y) {
      const float* p = hdr + 3 * ((size_t)y * w + x);
      a += r2o_det_logf(p[0] + p[1] + p[2] + FLT_MIN) / divisor;
    }
    avg[x] = a;
  }
  float r = 0.f;
  for (int x = 0; x < w; ++x) r += avg[x];
  const float k = r2o_det_expf(r) - (float)(w * h) * FLT_MIN;
  free(avg);
  const float inv_gamma = 1.f / gamma;
  for (size_t i = 0; i < (size_t)w * h * 3; ++i) {
    const float c = hdr[i];
    const float t = r2o_det_powf(c / (c + k), inv_gamma);
    if (ldr) ldr[i] = t;
    if (bytes) bytes[i] = (uint8_t)roundf(t * 255.0f);
  }
}
