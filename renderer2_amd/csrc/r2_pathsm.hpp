// Persistent path state machine for the volume integrator (gfx950).
//
// CPathIntegrator::Li (path_integrator_impl.hpp:120-247) with direct()
// (:23-118) and CDeviceScene::intersectTr (device_scene_impl.hpp:93-114) is
// re-expressed as explicit per-lane states.  Every free-flight (sampleDDA /
// sampleInternal) and transmittance (trDDA / trInternal) traversal of every
// path runs through ONE shared step loop, so the lanes of a wavefront advance
// their traversals together whatever shading stage each path is in; shading
// events run in batches between traversal rounds.  A lane whose path ends takes
// the next sample of its work item, or a new item: the wavefront ballots the
// lanes that need work, takes that many items with one atomic, and hands them
// out by prefix rank.  Random numbers are drawn in exactly the order of the
// recursive formulation, so each path's result is independent of scheduling.
#pragma once
#include "r2_volume.hpp"

namespace r2 {

enum : int {
  ST_IDLE = 0,    // no more work
  ST_FETCH,       // needs a work item (handled wave-wide)
  ST_TRAV,        // inside a traversal
  ST_START,       // new sample of the current item
  ST_BOUNCE,      // top of the bounce loop
  ST_AFTER_SAMPLE,
  ST_NOHIT,
  ST_SCATTER,
  ST_LIGHT_DONE,
  ST_BRDF_SAMPLE,
  ST_BRDF_DONE,
  ST_DIRECT_DONE,
  ST_SH_SEG,
  ST_SH_NEXT,
  ST_MOVE,        // spawn a ray from the scatter point and move it to the voxel border
  ST_FINISH,
};

enum : int { TK_FLIGHT_DDA = 1, TK_TR_DDA = 2, TK_FLIGHT_DELTA = 3, TK_TR_RATIO = 4 };

// Path state touched only at the start and end of a path and in shading events
// lives in LDS, one slot per lane (51 dwords: an odd stride, so the 64 lanes of
// a wavefront fall on distinct banks; 3 blocks of 256 lanes fit a CU's 160 KB);
// the state the traversal loop touches stays in registers.  This keeps the
// kernel within the register budget of three wavefronts per SIMD without
// spilling to scratch.
struct ColdLane {
  uint32_t pix;
  int32_t sh_next, sh_seg;  // shadow ray walk: medium after the segment, segments walked
  int32_t s_cur, s_end;
  uint32_t item;
  float ax, ay, az;     // FP32 item sum (same order as the reference's framebuffer +=)
  v3 Ld, pdir, pLe;     // pending NEE sample (light direction / phase direction, radiance)
  float ppdf;
  v3 mi_pos, mi_n;      // scatter event
  float S[6];
  float mi_ior;
  int32_t mi_inst;
  v3 sh_p0, sh_p1, si_pos;
  uint32_t pad;
  uint32_t ring[8];     // random words of the path (PathRng)
};
static_assert(sizeof(ColdLane) == 51 * 4, "ColdLane must keep an odd dword stride");

// The random stream of one path: Philox4x32-10 keyed (seed, stream), counter
// (pixel, sample, block, 0); draw k is word k & 3 of block k >> 2, the same
// stream as Rng.  Words are read from an 8-word ring in LDS that holds blocks
// b and b + 1 (block b at words 4 (b & 1) ..).  top_up() computes the next block
// ahead of use at points that the whole wavefront passes (the top of an event
// pass, of a traversal round), so that after it every lane has at least five
// words in hand; a draw computes a block itself only when its lane has run out.
// A wavefront therefore runs about one block computation per pass or round,
// where a lazily refilled buffer runs one at every draw site at which any of
// its lanes happens to cross into a new block.
struct PathRng {
  const uint32_t k0, k1;  // launch-uniform: kept in scalar registers
  uint32_t c0, c1;
  uint32_t n;   // words drawn
  uint32_t hi;  // blocks 0 .. hi - 1 computed; the ring holds the last two of them (block 0 alone after init)
  uint32_t (&ring)[8];
  __device__ PathRng(uint32_t (&r)[8], uint32_t seed, uint32_t stream) : k0(seed), k1(stream), ring(r) {}
  __device__ __forceinline__ void block(uint32_t b) {
    uint32_t c[4] = {c0, c1, b, 0u};
    philox4x32_10(c, k0, k1);
    const uint32_t o = (b & 1u) * 4u;
    ring[o] = c[0];
    ring[o + 1] = c[1];
    ring[o + 2] = c[2];
    ring[o + 3] = c[3];
  }
  __device__ __forceinline__ void init(uint32_t ctr0, uint32_t ctr1) {
    c0 = ctr0;
    c1 = ctr1;
    n = 0;
    block(0);
    hi = 1;
  }
  // make words n .. 4 (n / 4 + 2) - 1 available
  __device__ __forceinline__ void top_up() {
    const uint32_t want = (n >> 2) + 2u;
    if (hi < want) {
      if (hi + 1u < want) block(want - 2u);
      block(want - 1u);
      hi = want;
    }
  }
  __device__ __forceinline__ uint32_t next_u32() {
    if ((n >> 2) >= hi) {  // ran past the words in hand: compute the block now
      block(n >> 2);
      hi = (n >> 2) + 1u;
    }
    return ring[n++ & 7u];
  }
  __device__ __forceinline__ float uni() { return (float)(next_u32() >> 8) * 5.9604644775390625e-08f; }
};

struct PathLane {
  // work item (LDS)
  uint32_t& pix;
  int32_t &s_cur, &s_end;
  uint32_t& item;
  float &ax, &ay, &az;
  double dax, day, daz;  // FP64 item sum (compensated framebuffer form)
  // path
  PathRng rng;
  int stage, b, iters;
  v3 L, thr;
  v3& Ld;
  Ray ray;
  v3& si_pos;
  int si_medium, si_hit;
  // scatter event (LDS)
  v3 &mi_pos, &mi_n;
  float (&S)[6];
  float& mi_ior;
  int32_t& mi_inst;
  // pending NEE sample (LDS)
  v3 &pdir, &pLe;
  float& ppdf;
  // shadow ray walk (intersectTr); sh_ret is also the state that follows ST_MOVE
  v3 &sh_p0, &sh_p1;
  int32_t &sh_next, &sh_seg;
  int sh_cur, sh_hit, sh_ret;
  float sh_Tr;
  // traversal
  int tk, tinst, tsteps, tres;
  v3 to, td, tri;
  float tt, ttau, tmip, ttmax, ttr;

  __device__ PathLane(ColdLane& c, const DScene& s)
      : pix(c.pix), s_cur(c.s_cur), s_end(c.s_end), item(c.item), ax(c.ax), ay(c.ay), az(c.az),
        rng(c.ring, s.seed, s.stream),
        Ld(c.Ld), si_pos(c.si_pos), mi_pos(c.mi_pos), mi_n(c.mi_n), S(c.S), mi_ior(c.mi_ior), mi_inst(c.mi_inst),
        pdir(c.pdir), pLe(c.pLe), ppdf(c.ppdf), sh_p0(c.sh_p0), sh_p1(c.sh_p1), sh_next(c.sh_next),
        sh_seg(c.sh_seg) {}
};

template <bool kSingle>
__device__ __forceinline__ int lane_grid(const DScene& s, int inst) {
  return kSingle ? s.inst[0].grid : s.inst[inst].grid;
}

// --------------------------------------------------------------- traversal --
// Prologue of sampleDDA / sampleInternal / trDDA / trInternal on instance k.
template <bool kSingle>
__device__ __forceinline__ void start_traversal(const DScene& s, PathLane& p, int k, const Ray& rw, int kind) {
  const int ki = kSingle ? 0 : k;
  const DInstance& in = s.inst[ki];
  const DGrid& g = s.grid[in.grid];
  const Ray rm = ray_transform2(rw, in.w2m);
  p.tinst = k;
  p.tk = kind;
  p.tsteps = 0;
  p.tres = 0;
  p.tt = 0.f;
  p.ttr = 1.f;
  if ((s.flags & 1u) != 0u) {  // TK_FLIGHT_DDA / TK_TR_DDA (brick grid on: same for every lane)
    const Ray ir = ray_transform2(rm, g.m2s);
    p.to = ir.o;
    p.td = ir.d;
    p.ttmax = ir.tmax;
    p.tri = mk(1.f / ir.d.x, 1.f / ir.d.y, 1.f / ir.d.z);
    p.ttau = -logf(1.f - p.rng.uni());
    p.tmip = 3.f;
  } else {
    const Ray ir = ray_transform2(rm, g.m2i);
    p.to = ir.o;
    p.td = ir.d;
    p.ttmax = ir.tmax;
  }
  p.stage = ST_TRAV;
}

// One step of the traversal loop body.  Returns true when the traversal ended;
// p.tres = 1 for a real collision of a free flight.
template <bool kSingle, bool kCount>
__device__ __forceinline__ bool traversal_step(const DScene& s, PathLane& p, Cnt& cnt) {
  const int ki = kSingle ? 0 : p.tinst;
  const DInstance& in = s.inst[ki];
  const DGrid& g = s.grid[in.grid];
  const float ratio = in.ratio;
  if (p.tk == TK_FLIGHT_DDA || p.tk == TK_TR_DDA) {
    if (!(p.tsteps < kLoopCap && p.tt < p.ttmax)) return true;
    ++p.tsteps;
    count<kCount>(cnt, kCntDda);
    const v3 cur = p.to + p.tt * p.td;
    const int imip = (int)roundf(p.tmip);
    const float maj = ratio * lookup_majorant<kCount>(g, cur, imip, cnt);
    const float dt = step_dda(cur, p.tri, imip);
    const float prev = p.tt;
    p.tt += div_sigma_t(g, dt);
    if (p.tt == prev) return true;
    p.ttau -= maj * dt;
    p.tmip = gmin(p.tmip + 0.25f, 3.f);
    if (p.ttau > 0.f) return false;
    p.tt += p.ttau / (maj * g.sigma_t);
    if (p.tt >= p.ttmax) return true;
    const float u0 = p.rng.uni(), u1 = p.rng.uni(), u2 = p.rng.uni();
    const float d = lookup_density<kCount>(g, p.to + p.tt * p.td, mk(u0, u1, u2), cnt);
    if (p.rng.uni() * maj < d * ratio) {
      if (p.tk == TK_FLIGHT_DDA) {
        p.tres = 1;
        return true;
      }
      const float inv_max = 1.f / (ratio / g.inv_max_density);
      const float vol_maj = 1.f / inv_max;
      p.ttr *= gmax(0.f, 1.f - vol_maj / maj);
      if (p.ttr < 0.1f) {
        const float prob = 1.f - p.ttr;
        if (p.rng.uni() < prob) {
          p.ttr = 0.f;
          return true;
        }
        p.ttr /= 1.f - prob;
      }
    }
    p.ttau = -logf(1.f - p.rng.uni());
    p.tmip = gmax(0.f, p.tmip - 2.f);
    return false;
  }
  // delta / ratio tracking against the global majorant (sampleInternal / trInternal)
  if (!(p.tsteps < kLoopCap)) return true;
  ++p.tsteps;
  count<kCount>(cnt, kCntFlight);
  const float inv_max = 1.f / (ratio / g.inv_max_density);
  p.tt -= div_sigma_t(g, logf(1.f - p.rng.uni()) * inv_max);
  if (p.tt >= p.ttmax) return true;
  const Rec fd = filtered_data<kCount>(g, p.to + p.tt * p.td, cnt);
  if (p.tk == TK_FLIGHT_DELTA) {
    if (ratio * fd.density * inv_max > p.rng.uni()) {
      p.tres = 1;
      return true;
    }
    return false;
  }
  p.ttr *= 1.f - gmax(0.f, ratio * fd.density * inv_max);
  if (p.ttr < 1.f) {
    const float pr = 1.f - p.ttr;
    if (p.rng.uni() < pr) {
      p.ttr = 0.f;
      return true;
    }
    p.ttr /= 1.f - pr;
  }
  return false;
}

// Collision tail of sampleDDA / sampleInternal + CMediumInstance::sample's
// transform back to world space; returns the albedo, fills the scatter record.
template <bool kSingle, bool kCount>
__device__ __forceinline__ v3 flight_collision(const DScene& s, PathLane& p, Cnt& cnt) {
  const int ki = kSingle ? 0 : p.tinst;
  const DInstance& in = s.inst[ki];
  const DGrid& g = s.grid[in.grid];
  Rec fd;
  Ray rw;
  const Ray ir = mkray(p.to, p.td, p.tt, -1);
  if (p.tk == TK_FLIGHT_DDA) {
    const v3 sp = p.to + p.tt * p.td;
    const v3 pos = mk(div_size(g, sp.x - (float)g.ibb_min[0], 0), div_size(g, sp.y - (float)g.ibb_min[1], 1),
                      div_size(g, sp.z - (float)g.ibb_min[2], 2));
    fd = filtered_data<kCount>(g, pos, cnt);
    rw = ray_transform2(ir, g.s2m);
  } else {
    Cnt dummy;
    fd = filtered_data<false>(g, p.to + p.tt * p.td, dummy);  // same lookup as the accepted step
    rw = ray_transform2(ir, g.i2m);
  }
  MHit h;
  const v3 alb = collision_albedo(g, fd, rw, h);
  p.mi_inst = p.tinst;
  p.mi_pos = mul_point(in.m2w, h.pos);
  p.mi_n = normalize(mul_dir(in.m2w, h.normal));
#pragma unroll
  for (int i = 0; i < 6; ++i) p.S[i] = h.S[i];
  p.mi_ior = h.ior;
  count<kCount>(cnt, kCntScatter);
  return alb;
}

template <bool kSingle>
__device__ __forceinline__ Ray lane_move_to_voxel_border(const DScene& s, int k, const Ray& r) {
  return inst_move_to_voxel_border(s, kSingle ? 0 : k, r);
}

__device__ __forceinline__ void start_shadow(PathLane& p, const Ray& r, int ret) {
  p.sh_Tr = 1.f;
  p.sh_p0 = r.o;
  p.sh_p1 = r.o + r.tmax * r.d;
  p.sh_cur = r.medium;
  p.sh_seg = 0;
  p.sh_ret = ret;
  p.stage = ST_SH_SEG;
}

// One transition of a lane that is not traversing, not fetching and not idle.
// `st` is the state to run (the lane's stage); passing a wave-uniform value lets
// the switch compile to one uniform branch.
template <bool kSingle, bool kCount>
__device__ __forceinline__ void event_step(const DScene& s, PathLane& p, Cnt& cnt, int st) {
  const bool dda = (s.flags & 1u) != 0u;
  switch (st) {
    case ST_START: {
      p.rng.init(p.pix, (uint32_t)(s.sample_begin + p.s_cur));
      p.ray = camera_ray(s, (int)(p.pix % (uint32_t)s.W), (int)(p.pix / (uint32_t)s.W), p.rng);
      p.L = mk(0.f, 0.f, 0.f);
      p.thr = mk(1.f, 1.f, 1.f);
      p.b = 0;
      p.iters = 0;
      p.stage = ST_BOUNCE;
      break;
    }
    case ST_BOUNCE: {
      // also runs the lanes in ST_SH_SEG: both intersect the scene and may start a
      // traversal (Li's free flight, or the transmittance of one intersectTr
      // segment), so a wavefront runs that code once for both
      const bool sh = p.stage == ST_SH_SEG;
      if (sh ? ++p.sh_seg > 64 : (!(p.b < s.max_depth) || ++p.iters > 4 * s.max_depth + 64)) {
        if (sh) {
          p.sh_hit = 1;
          p.stage = p.sh_ret;
        } else {
          p.stage = ST_FINISH;
        }
        break;
      }
      Ray r = sh ? spawn_ray(p.sh_p0, p.sh_p1, p.sh_cur) : p.ray;
      const Hit h = scene_intersect(s, r);  // shortens r.tmax to the hit, as the reference's intersect does
      Ray fr = r;
      int next;
      if (sh) {
        p.sh_hit = h.hit;
        p.sh_next = r.medium < 0 ? h.medium : -1;
        p.sh_p0 = h.pos;
        next = (h.hit && r.medium >= 0 && (h.medium < 0 || h.medium == p.sh_cur)) ? ST_TRAV : ST_SH_NEXT;
      } else {
        p.ray = r;
        p.si_pos = h.pos;
        p.si_medium = h.medium;
        p.si_hit = h.hit;
        p.tres = 0;
        p.tk = 0;
        // the free flight starts on the current ray (inside a medium) or on the ray
        // entering the medium that was hit
        next = ST_NOHIT;
        if (h.medium >= 0) {
          if (r.medium >= 0) {
            next = h.medium != r.medium ? ST_FINISH : ST_TRAV;
          } else {
            Ray mr = mkray(h.pos, normalize(r.d), kDefaultTmax, h.medium);
            ray_offset(mr, r.d);
            const Hit se = scene_intersect(s, mr);
            next = !se.hit ? ST_AFTER_SAMPLE : ((se.medium >= 0 && h.medium != se.medium) ? ST_FINISH : ST_TRAV);
            fr = mr;
          }
        }
      }
      if (next == ST_TRAV)
        start_traversal<kSingle>(s, p, fr.medium, fr,
                                 sh ? (dda ? TK_TR_DDA : TK_TR_RATIO) : (dda ? TK_FLIGHT_DDA : TK_FLIGHT_DELTA));
      else
        p.stage = next;
      break;
    }
    case ST_AFTER_SAMPLE: {
      const bool collided = p.tres != 0;
      if (collided) p.thr = p.thr * flight_collision<kSingle, kCount>(s, p, cnt);
      p.tres = 0;
      if (p.thr.x == 0.f || p.thr.y == 0.f || p.thr.z == 0.f) {
        p.stage = ST_FINISH;
        break;
      }
      p.stage = collided ? ST_SCATTER : ST_NOHIT;
      break;
    }
    case ST_NOHIT: {
      // also runs the lanes in ST_BRDF_DONE: both look up the environment (the
      // radiance of an escaped camera ray, or the light along direct()'s phase
      // sample), so a wavefront runs that lookup once for both
      const bool brdf = p.stage == ST_BRDF_DONE;
      const bool escaped = !brdf && p.b == 0 && !p.si_hit;
      float lpdf = 0.f;
      v3 Le = mk(0.f, 0.f, 0.f);
      if (brdf || escaped) Le = env_le<kCount>(s.env, brdf ? p.pdir : p.ray.d, lpdf, cnt);
      if (brdf) {
        const float ph = p.ppdf;
        const float mis = (1.f * ph) / (1.f * ph + 1.f * lpdf);
        const float tr = p.sh_hit ? 0.f : p.sh_Tr;
        p.Ld = p.Ld + ((Le * (mis * ph)) * tr) / ph;
        p.stage = ST_DIRECT_DONE;
        break;
      }
      if (escaped) p.L = p.L + Le * p.thr;
      if (!p.si_hit) {
        p.stage = ST_FINISH;
        break;
      }
      Ray nr = mkray(p.si_pos, normalize(p.ray.d), kDefaultTmax, p.ray.medium < 0 ? p.si_medium : -1);
      ray_offset(nr, p.ray.d);
      p.ray = nr;
      p.stage = ST_BOUNCE;
      break;
    }
    case ST_SCATTER: {
      p.Ld = mk(0.f, 0.f, 0.f);
      v3 ldir;
      float lpdf = 0.f;
      const v3 Le = env_sample<kCount>(s.env, p.rng, ldir, lpdf, cnt);
      if (lpdf > 0.f) {
        p.pdir = normalize(ldir);
        p.pLe = Le;
        p.ppdf = lpdf;
        p.sh_ret = ST_LIGHT_DONE;  // shadow ray toward the light sample (ST_MOVE)
        p.stage = ST_MOVE;
      } else {
        p.stage = ST_BRDF_SAMPLE;
      }
      break;
    }
    case ST_LIGHT_DONE: {
      const v3 wo = -p.ray.d;
      const float ph = phase_p(p.S, p.mi_n, p.mi_ior, wo, p.pdir, p.rng);
      const float mis = (1.f * p.ppdf) / (1.f * p.ppdf + 1.f * ph);
      const float tr = p.sh_hit ? 0.f : p.sh_Tr;
      p.Ld = p.Ld + ((p.pLe * (mis * ph)) * tr) / p.ppdf;
      p.stage = ST_BRDF_SAMPLE;
      break;
    }
    case ST_BRDF_SAMPLE: {
      // also runs the lanes in ST_DIRECT_DONE: both start with a phase sample at
      // the scatter point (direct()'s phase-sampled light direction, then Li's
      // continuation direction), so a wavefront runs that code once for both
      const bool cont = p.stage == ST_DIRECT_DONE;
      if (cont) p.L = p.L + p.thr * p.Ld;
      const v3 wo = -p.ray.d;
      v3 wi;
      const float ph = phase_sample(p.S, p.mi_n, p.mi_ior, wo, wi, p.rng);
      if (!cont) {
        if (ph > 0.f) {
          p.pdir = normalize(wi);
          p.ppdf = ph;
          p.sh_ret = ST_BRDF_DONE;  // shadow ray along the phase sample (ST_MOVE)
          p.stage = ST_MOVE;
        } else {
          p.stage = ST_DIRECT_DONE;
        }
        break;
      }
      p.pdir = normalize(wi);
      p.Ld = wi;  // the continuation ray is offset along the unnormalised sample
      if (p.b > 3) {
        const float q = 1.f - (p.thr.x + p.thr.y + p.thr.z) / 3.f;
        if (p.rng.uni() <= q) {
          p.stage = ST_FINISH;
          break;
        }
        p.thr = p.thr / (1.f - q);
      }
      ++p.b;
      p.sh_ret = ST_BOUNCE;
      p.stage = ST_MOVE;
      break;
    }
    case ST_MOVE: {
      // the ray from the scatter point along pdir, offset along the normal (shadow
      // rays) or along the unnormalised direction sample (continuation), moved to
      // the voxel border: the tails of direct() and of Li's bounce
      const bool cont = p.sh_ret == ST_BOUNCE;
      Ray r = mkray(p.mi_pos, p.pdir, kDefaultTmax, p.mi_inst);
      ray_offset(r, cont ? p.Ld : p.mi_n);
      r = lane_move_to_voxel_border<kSingle>(s, p.mi_inst, r);
      if (cont) {
        p.ray = r;
        p.stage = ST_BOUNCE;
      } else {
        start_shadow(p, r, p.sh_ret);
      }
      break;
    }
    case ST_SH_NEXT: {
      if (!p.sh_hit) {
        p.stage = p.sh_ret;
        break;
      }
      p.sh_cur = p.sh_next;
      p.stage = ST_SH_SEG;
      break;
    }
    default:
      break;
  }
}

// Called when a traversal ended: hand the result to the owning stage.
__device__ __forceinline__ void end_traversal(PathLane& p) {
  if (p.tk == TK_TR_DDA || p.tk == TK_TR_RATIO) {
    p.sh_Tr *= p.ttr;
    p.stage = ST_SH_NEXT;
  } else {
    p.stage = ST_AFTER_SAMPLE;
  }
}


}  // namespace r2
