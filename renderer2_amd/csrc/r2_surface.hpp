// Analytic surfaces of a render scene and their material, for the surface
// vertices of the path integrator.
//
//  * shapes: CCircle::intersect (circle.hpp:88-113) and Sphere::intersect
//    (sphere.hpp:33-90), reported to the closest-hit program, which normalises
//    the normal (shaders.optix.cu:115-130, 170-189);
//  * material: CMaterial (material.hpp:67-100), an Oren-Nayar diffuse lobe
//    (oren_nayar_brdf.hpp:42-84) and a Trowbridge-Reitz microfacet lobe
//    (microfacet_brdf.hpp:57-100, microfacet_distribution.hpp:40-140) mixed by the
//    dielectric Fresnel weight of |cos theta_o| (fresnel.hpp:28-54);
//  * the tangent frame CCoordinateFrame::fromNormal (objects.hpp:28-40).
// The reference evaluates M_PI, M_1_PI and a few literals in double inside float
// code; dmul() keeps those products in double and rounds once, as it does.
// Included by r2_volume.hpp ahead of the scene intersection: it uses that
// file's Ray, frame_basic and cosine_hemisphere.
#pragma once

namespace r2 {

// ------------------------------------------------------------------ shapes --
__device__ __forceinline__ bool circle_hit(const DSurface& sf, const Ray& ray, float tmax, float& t_out, v3& n_out) {
  const v3 n = mk(sf.normal[0], sf.normal[1], sf.normal[2]);
  const v3 c = mk(sf.pos[0], sf.pos[1], sf.pos[2]);
  const float den = dot(ray.d, n);
  if (den == 0.f) return false;
  const float t = dot(c - ray.o, n) / den;
  if (!(t > 0.f && t <= tmax)) return false;
  const v3 ip = ray.o + t * ray.d;
  if (!(length(ip - c) < sf.radius)) return false;
  t_out = t;
  n_out = n;
  return true;
}
// the quadratic in model space (the ray through CRay::transform), t compared
// against the world ray's tmax, the far root when the origin is inside
__device__ __forceinline__ bool sphere_hit(const DSurface& sf, const Ray& ray, float tmax, float& t_out, v3& n_out) {
  const v3 om = mul_point(sf.w2m, ray.o);
  const v3 dm = normalize(mul_dir(sf.w2m, ray.d));
  const float a = dot(dm, dm);
  const float b = 2.f * dot(dm, om);
  const float c = dot(om, om) - sf.radius * sf.radius;
  const float disc = b * b - 4.f * a * c;
  float t;
  if (disc == 0.f) {
    t = (-b + sqrtf(disc)) / (2.f * a);
    if (!(t > 0.f && t <= tmax)) return false;
  } else if (disc > 0.f) {
    const float sq = sqrtf(disc);
    const float den = 1.f / (2.f * a);
    const float t1 = (-b + sq) * den, t2 = (-b - sq) * den;
    const float mn = gmin(t1, t2), mx = gmax(t1, t2);
    if (!(mx > 0.f)) return false;
    if (mn > 0.f && mn <= tmax) t = mn;
    else if (mn <= 0.f && mx <= tmax) t = mx;
    else return false;
  } else {
    return false;
  }
  t_out = t;
  n_out = normalize(mul_dir(sf.m2w, om + t * dm));
  return true;
}
__device__ __forceinline__ bool surface_hit(const DSurface& sf, const Ray& ray, float tmax, float& t, v3& n) {
  return sf.shape == 0 ? circle_hit(sf, ray, tmax, t, n) : sphere_hit(sf, ray, tmax, t, n);
}

// ------------------------------------------------------------------ frame --
// tangent-to-world columns (T, B, N) and their glm::inverse (cofactors over the
// determinant, in glm's order)
struct SFrame {
  float t2w[3][3], w2t[3][3];  // [column][row]
};
__device__ __forceinline__ SFrame surface_frame(v3 n) {
  v3 T, B;
  frame_basic(n, T, B);
  SFrame f;
  f.t2w[0][0] = T.x; f.t2w[0][1] = T.y; f.t2w[0][2] = T.z;
  f.t2w[1][0] = B.x; f.t2w[1][1] = B.y; f.t2w[1][2] = B.z;
  f.t2w[2][0] = n.x; f.t2w[2][1] = n.y; f.t2w[2][2] = n.z;
  const float(*m)[3] = f.t2w;
  const float det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                    m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2]) +
                    m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
  const float o = 1.f / det;
  f.w2t[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * o;
  f.w2t[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * o;
  f.w2t[2][0] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * o;
  f.w2t[0][1] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) * o;
  f.w2t[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) * o;
  f.w2t[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * o;
  f.w2t[0][2] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) * o;
  f.w2t[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * o;
  f.w2t[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) * o;
  return f;
}
__device__ __forceinline__ v3 mat3_mul(const float (&m)[3][3], v3 v) {
  return mk(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
            m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}

// --------------------------------------------------------------- material --
constexpr double kDPi = 3.14159265358979323846;
constexpr double kD1Pi = 0.318309886183790671538;
__device__ __forceinline__ float dmul(double c, float x) { return (float)(c * (double)x); }

// CFresnel::evaluate with both indices
__device__ __forceinline__ float fresnel_ie(float cos_i, float eta_i_in, float eta_t_in) {
  cos_i = cos_i < -1.f ? -1.f : (cos_i > 1.f ? 1.f : cos_i);
  float eta_i = eta_i_in, eta_t = eta_t_in;
  if (!(cos_i > 0.f)) {
    eta_i = eta_t_in;
    eta_t = eta_i_in;
    cos_i = fabsf(cos_i);
  }
  const float sin_i = sqrtf(gmax(0.f, 1.f - cos_i * cos_i));
  const float sin_t = eta_i / eta_t * sin_i;
  if (sin_t >= 1.f) return 1.f;
  const float cos_t = sqrtf(gmax(0.f, 1.f - sin_t * sin_t));
  const float rpar = ((eta_t * cos_i) - (eta_i * cos_t)) / ((eta_t * cos_i) + (eta_i * cos_t));
  const float rper = ((eta_i * cos_i) - (eta_t * cos_t)) / ((eta_i * cos_i) + (eta_t * cos_t));
  return (rpar * rpar + rper * rper) / 2.f;
}

// brdf_functions.hpp:7-55, sameHemisphere (functions.hpp:58-60)
__device__ __forceinline__ float sin_theta(v3 w) { return sqrtf(gmax(0.f, 1.f - w.z * w.z)); }
__device__ __forceinline__ float clamp11(float x) { return x < -1.f ? -1.f : (x > 1.f ? 1.f : x); }
__device__ __forceinline__ float cos_phi(v3 w) {
  const float st = sin_theta(w);
  return st == 0.f ? 1.f : clamp11(w.x / st);
}
__device__ __forceinline__ float sin_phi(v3 w) {
  const float st = sin_theta(w);
  return st == 0.f ? 0.f : clamp11(w.y / st);
}
__device__ __forceinline__ bool same_hemisphere(v3 a, v3 b) { return a.z * b.z > 0.f; }

__device__ __forceinline__ float on_f(const DSurface& m, v3 wo, v3 wi) {
  const float st_i = sin_theta(wi), st_o = sin_theta(wo);
  float max_cos = 0.f;
  if ((double)st_i > 1.0e-4 && (double)st_o > 1.0e-4)
    max_cos = gmax(0.f, cos_phi(wi) * cos_phi(wo) + sin_phi(wi) * sin_phi(wo));
  float sin_a, tan_b;
  if (fabsf(wi.z) > fabsf(wo.z)) {
    sin_a = st_o;
    tan_b = st_i / fabsf(wi.z);
  } else {
    sin_a = st_i;
    tan_b = st_o / fabsf(wo.z);
  }
  return dmul(kD1Pi, m.on_a + m.on_b * max_cos * sin_a * tan_b);
}
__device__ __forceinline__ float on_pdf(v3 wo, v3 wi) { return same_hemisphere(wo, wi) ? dmul(kD1Pi, fabsf(wi.z)) : 0.f; }

__device__ __forceinline__ float ggx_D(const DSurface& m, v3 h) {
  const float c2 = h.z * h.z;
  const float t2 = gmax(0.f, 1.f - c2) / c2;
  if (isinf(t2)) return 0.f;
  const float c4 = c2 * c2;
  const float cp = cos_phi(h), sp = sin_phi(h);
  const float e = (cp * cp / (m.alpha_x * m.alpha_x) + sp * sp / (m.alpha_y * m.alpha_y)) * t2;
  const float e1 = 1.f + e;
  return (float)(1.0 / ((((kDPi * (double)m.alpha_x) * (double)m.alpha_y) * (double)c4) * (double)e1 * (double)e1));
}
__device__ __forceinline__ float ggx_lambda(const DSurface& m, v3 w) {
  const float at = fabsf(sin_theta(w) / w.z);
  if (isinf(at)) return 0.f;
  const float cp = cos_phi(w), sp = sin_phi(w);
  const float alpha = sqrtf(cp * cp * m.alpha_x * m.alpha_x + sp * sp * m.alpha_y * m.alpha_y);
  const float a2t2 = (alpha * at) * (alpha * at);
  return (-1.f + sqrtf(1.f + a2t2)) / 2.f;
}
__device__ __forceinline__ float ggx_pdf(const DSurface& m, v3 wo, v3 wh) {
  const float g1 = 1.f / (1.f + ggx_lambda(m, wo));
  return ggx_D(m, wh) * g1 * fabsf(dot(wo, wh)) / fabsf(wo.z);
}
// visible-slope sampling of the unit-roughness distribution (Heitz and d'Eon
// 2014, with pbrt-v3's rational fit of the y slope), stretched by the alphas
__device__ __forceinline__ v3 ggx_sample_h(const DSurface& m, v3 wo, float u1, float u2) {
  const bool flip = wo.z < 0.f;
  const v3 wi = flip ? -wo : wo;
  const v3 ws = normalize(mk(m.alpha_x * wi.x, m.alpha_y * wi.y, wi.z));
  const float ct = ws.z;
  float sx, sy;
  if ((double)ct > 0.9999) {
    const float r = sqrtf(u1 / (1.f - u1));
    const float phi = dmul(6.28318530718, u2);
    sx = r * cosf(phi);
    sy = r * sinf(phi);
  } else {
    const float st = sqrtf(gmax(0.f, 1.f - ct * ct));
    const float tt = st / ct;
    const float a = 1.f / tt;
    const float g1 = 2.f / (1.f + sqrtf(1.f + 1.f / (a * a)));
    const float A = 2.f * u1 / g1 - 1.f;
    float tmp = 1.f / (A * A - 1.f);
    if ((double)tmp > 1e10) tmp = 1e10f;
    const float B = tt;
    const float D = sqrtf(gmax(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.f));
    const float x1 = B * tmp - D, x2 = B * tmp + D;
    sx = (A < 0.f || x2 > 1.f / tt) ? x1 : x2;
    float S;
    if (u2 > 0.5f) {
      S = 1.f;
      u2 = 2.f * (u2 - 0.5f);
    } else {
      S = -1.f;
      u2 = 2.f * (0.5f - u2);
    }
    const float z = (u2 * (u2 * (u2 * 0.27385f - 0.73369f) + 0.46341f)) /
                    (u2 * (u2 * (u2 * 0.093073f + 0.309420f) - 1.f) + 0.597999f);
    sy = S * z * sqrtf(1.f + sx * sx);
  }
  const float cp = cos_phi(ws), sp = sin_phi(ws);
  const float tmp = cp * sx - sp * sy;
  sy = sp * sx + cp * sy;
  sx = tmp;
  sx = m.alpha_x * sx;
  sy = m.alpha_y * sy;
  const v3 h = normalize(mk(-sx, -sy, 1.f));
  return flip ? -h : h;
}
__device__ __forceinline__ float mf_f(const DSurface& m, v3 wo, v3 wi) {
  const float co = fabsf(wo.z), ci = fabsf(wi.z);
  v3 h = wi + wo;
  if (ci == 0.f || co == 0.f) return 0.f;
  if (h.x == 0.f && h.y == 0.f && h.z == 0.f) return 0.f;
  h = normalize(h);
  const v3 hf = h.z < 0.f ? -h : h;
  const float F = fresnel_ie(dot(wi, hf), m.eta_i, m.eta_t);
  const float G = 1.f / (1.f + ggx_lambda(m, wo) + ggx_lambda(m, wi));
  return ggx_D(m, h) * G * F / (4.f * ci * co);
}
__device__ __forceinline__ float mf_pdf(const DSurface& m, v3 wo, v3 wi) {
  if (!same_hemisphere(wo, wi)) return 0.f;
  const v3 h = normalize(wo + wi);
  return ggx_pdf(m, wo, h) / (4.f * dot(wo, h));
}
__device__ __forceinline__ v3 mat_f(const DSurface& m, v3 wo, v3 wi) {
  const float fd = on_f(m, wo, wi), fg = mf_f(m, wo, wi);
  const float w = fresnel_ie(fabsf(wo.z), m.eta_i, m.eta_t);
  return (mk(m.diffuse[0], m.diffuse[1], m.diffuse[2]) * fd) * (1.f - w) +
         (mk(m.specular[0], m.specular[1], m.specular[2]) * fg) * w;
}
__device__ __forceinline__ float mat_pdf(const DSurface& m, v3 wo, v3 wi) {
  const float w = fresnel_ie(fabsf(wo.z), m.eta_i, m.eta_t);
  return on_pdf(wo, wi) * (1.f - w) + mf_pdf(m, wo, wi) * w;
}
// CMaterial::sampleF: one lobe chosen by the Fresnel weight, its f and its own
// pdf; *wi and *pdf stay untouched when the microfacet lobe rejects the draw
template <typename Rng>
__device__ __forceinline__ v3 mat_sample(const DSurface& m, v3 wo, v3& wi, Rng& rng, float& pdf) {
  const float p = fresnel_ie(fabsf(wo.z), m.eta_i, m.eta_t);
  if (rng.uni() < p) {
    const v3 spec = mk(m.specular[0], m.specular[1], m.specular[2]);
    if (wo.z == 0.f) return spec * 0.f;
    const float u1 = rng.uni(), u2 = rng.uni();
    const v3 h = ggx_sample_h(m, wo, u1, u2);
    const float dh = dot(wo, h);
    if (dh < 0.f) return spec * 0.f;
    const v3 I = -wo;
    const float d = dot(h, I);
    const v3 r = mk(I.x - h.x * d * 2.f, I.y - h.y * d * 2.f, I.z - h.z * d * 2.f);
    wi = r;
    if (!same_hemisphere(wo, r)) return spec * 0.f;
    pdf = ggx_pdf(m, wo, h) / (4.f * dh);
    return spec * mf_f(m, wo, r);
  }
  v3 w = cosine_hemisphere(rng);
  if (wo.z < 0.f) w.z = -w.z;
  wi = w;
  pdf = on_pdf(wo, w);
  return mk(m.diffuse[0], m.diffuse[1], m.diffuse[2]) * on_f(m, wo, w);
}

}  // namespace r2
