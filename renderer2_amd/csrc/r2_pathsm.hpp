// Path integrator as a two-phase lane program for gfx950 wavefronts.
//
// CPathIntegrator::Li (path_integrator_impl.hpp:120-247) with direct() (:23-118)
// and CDeviceScene::intersectTr (device_scene_impl.hpp:93-114), re-expressed so
// that a path alternates between two kinds of work:
//   * SHADE at a vertex (the camera, or a scatter event): everything that does
//     not walk through the volume.  The direct light of the previous vertex once
//     its shadow walks are done; the collision record of the ray that arrived
//     (trilinear voxel fetch, Fresnel-mixed albedo); direct()'s light sample and
//     phase evaluation and its phase sample; Li's continuation sample and Russian
//     roulette; the three rays spawned at the scatter point and moved to the
//     voxel border.  A lane whose path ends accumulates it and starts its next
//     sample in the same pass.
//   * WALK of the vertex's rays, one after another: the light-sample shadow ray
//     and the phase-sample shadow ray (intersectTr segment walks with DDA or
//     ratio-tracking transmittance), then the continuation (Li's bounce loop up
//     to the next collision: scene intersection, the entry-segment double
//     tracking, bounding-box pass-throughs, free flight).
// Each vertex and each of its rays draws from a Philox stream of its own
// (counter word 3 = 4 k + role, the oracle's path streams), so the rays need not
// run in the reference's interleaved order, and a path on the GPU follows the
// FP32 oracle's path draw for draw until a decision falls within rounding of its
// threshold.  A wavefront shades all its lanes that are at a vertex in one pass
// and advances all its walking lanes one walk step per round, so it runs one of
// two code blocks at a time instead of a mix of shading states.
#pragma once
#include "r2_volume.hpp"

namespace r2 {

// Where a lane waits.  A walking lane either starts a segment (scene
// intersection, traversal set-up) or advances a traversal.  One variable holds
// it: stores of different states in the arms of a branch then merge into one
// value select, where stores to different fields could merge into a store
// through a selected address and keep the lane's state out of registers.
enum : int {
  ST_IDLE = 0,  // no more work
  ST_FETCH,     // needs a work item (handled wave-wide)
  ST_SHADE,     // at a vertex
  ST_SEG,       // walking: at the start of a segment
  ST_TRAV,      // walking: inside a traversal
};
// roles of a vertex's streams; the rays are walked in this order
enum : int { RL_VERTEX = 0, RL_LIGHT = 1, RL_PHASE = 2, RL_CONT = 3 };
// how the ray that arrived at a vertex ended (set by the continuation walk)
enum : int { CR_START = 0, CR_NONE, CR_END, CR_ESCAPED, CR_COLLIDED, CR_SURFACE };
enum : int { TK_FLIGHT_DDA = 1, TK_TR_DDA = 2, TK_FLIGHT_DELTA = 3, TK_TR_RATIO = 4 };
// ColdLane::flags: the rays of the vertex that wait for their walks
enum : uint32_t { FL_LIGHT = 1u, FL_PHASE = 2u, FL_CONT = 4u };
// caps on the segments of one intersectTr walk and on the bounce loop's
// iterations, which the reference loops until a miss and does not count box
// pass-throughs: 64 (beyond 4 max_depth iterations) plus 4 per instance beyond
// the fourth; a segment or pass-through enters or leaves an instance box, so a
// walk that crosses every box once stays below them (the oracle uses the same caps)
__device__ __forceinline__ int inst_extra(const DScene& s) { return s.n_inst > 4 ? 4 * (s.n_inst - 4) : 0; }

// Path state that is touched only in shading passes and at walk boundaries
// lives in LDS, one slot per lane; the state the walk steps touch stays in
// registers.  39 dwords: an odd stride, so the 64 lanes of a wavefront fall on
// distinct banks, and four blocks of 256 lanes (156 KB) fit a CU's 160 KB of
// LDS.  The lane's pixel and sample are the counter words 0 and 1 of its random
// stream (registers); the item's chunk follows from the sample.
struct ColdLane {
  float ax, ay, az;  // FP32 item sum (the reference's framebuffer += order); a probe keeps its item in ax
  uint32_t ring[8];  // random words of the current stream (PathRng)
  // rays of the current vertex waiting for their walk (the light-sample shadow
  // ray, walked first, goes straight to the walk registers): the phase-sample
  // shadow ray as segment end points and medium (intersectTr's p0, p1,
  // currentMedium), the continuation as a ray
  v3 lb0, lb1;
  int32_t lb_med;
  v3 c_o, c_d;
  float c_tmax;
  int32_t c_med;
  // direct light of the vertex per strategy, throughput * (Le * (mis * p)) / pdf,
  // added to L times the shadow walk's transmittance when that walk ends
  v3 nee_a, nee_b;
  uint32_t flags;  // FL_*: rays of the vertex waiting for their walks
  v3 L, thr;       // radiance and throughput of the path
};
static_assert(sizeof(ColdLane) == 39 * 4, "ColdLane must keep an odd dword stride");

// The random stream of one (path, vertex, role): Philox4x32-10 keyed (seed,
// stream), counter (pixel, sample, block, 4 k + role); draw j is word j & 3 of
// block j >> 2.  Words are read from an 8-word ring in LDS holding blocks b and
// b + 1 (block b at words 4 (b & 1) ..).  top_up() computes blocks ahead of use at
// points that all lanes of a pass or step reach together, so that after it every
// lane has at least R2_RNG_AHEAD words in hand; a draw computes a block itself
// only when its lane has run out.  init() computes nothing: the first top_up()
// or draw does.
#ifndef R2_RNG_AHEAD
#define R2_RNG_AHEAD 4
#endif
static_assert(R2_RNG_AHEAD >= 1 && R2_RNG_AHEAD <= 5, "the ring holds two blocks");
struct PathRng {
  const uint32_t k0, k1;  // launch-uniform: kept in scalar registers
  uint32_t c0, c1, c3;
  uint32_t n;   // words drawn
  uint32_t hi;  // blocks 0 .. hi - 1 computed; the ring holds the last two of them
  uint32_t (&ring)[8];
  __device__ PathRng(uint32_t (&r)[8], uint32_t seed, uint32_t stream) : k0(seed), k1(stream), ring(r) {}
  __device__ __forceinline__ void block(uint32_t b) {
    uint32_t c[4] = {c0, c1, b, c3};
    philox4x32_10(c, k0, k1);
    const uint32_t o = (b & 1u) * 4u;
    ring[o] = c[0];
    ring[o + 1] = c[1];
    ring[o + 2] = c[2];
    ring[o + 3] = c[3];
  }
  __device__ __forceinline__ void init(uint32_t ctr0, uint32_t ctr1, uint32_t sub) {
    c0 = ctr0;
    c1 = ctr1;
    c3 = sub;
    n = 0;
    hi = 0;
  }
  // make at least the next R2_RNG_AHEAD words (at most 5) available
  __device__ __forceinline__ void top_up() {
    const uint32_t want = (n + (uint32_t)R2_RNG_AHEAD + 3u) >> 2;
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
  ColdLane& c;
  // work item: the pixel is rng.c0, the sample rng.c1 (+ the launch's first);
  // its FP32 sum in LDS
  float &ax, &ay, &az;
  double dax, day, daz;  // FP64 item sum (compensated framebuffer form)
  PathRng rng;
  int stage, role, cres;
  int b, k, iters;
  v3 &L, &thr;
  // continuation walk: the ray of the bounce loop, and its last scene
  // intersection (a surface hit keeps the surface in si_medium and its normal in
  // ray.o until it is shaded).  Shadow walks keep intersectTr's p0 in ray.o, p1
  // in ray.d and currentMedium in ray.medium, the segment's hit flag in si_hit.
  // During a shadow traversal si_hit = 2 says that the segment to follow it is
  // known to meet nothing.
  Ray ray;
  v3 si_pos;
  int si_medium, si_hit;
  int sh_seg;
  float sh_Tr;
  // traversal
  int tk, tinst, tsteps, tres;
  v3 to, td, tri;
  float tt, ttau, tmip, ttmax, ttr;

  __device__ PathLane(ColdLane& cl, const DScene& s)
      : c(cl), ax(cl.ax), ay(cl.ay), az(cl.az), rng(cl.ring, s.seed, s.stream), L(cl.L), thr(cl.thr) {}
};

// scatter record of a collision (CMediumInstance::sample's SMediumInteraction)
struct Scatter {
  v3 pos, n;
  float S[6];
  float ior;
  int inst;
};

__device__ __forceinline__ bool any_zero(v3 a) { return a.x == 0.f || a.y == 0.f || a.z == 0.f; }

// The lane's fields are carried around the scheduling loop, and the compiler
// cannot see which of them a block's successors read before writing, so it
// keeps every field in a register through every block.  A block marks the fields
// that are always written again before their next read as dead (an undefined
// value: no instruction), which frees their registers for its temporaries.
template <typename T>
__device__ __forceinline__ void dead(T& x) { x = __builtin_nondeterministic_value(x); }
__device__ __forceinline__ void dead(v3& a) { dead(a.x); dead(a.y); dead(a.z); }
// the traversal fields: start_traversal writes all that a traversal reads
__device__ __forceinline__ void dead_traversal(PathLane& p) {
  dead(p.tk); dead(p.tinst); dead(p.tsteps); dead(p.tres);
  dead(p.to); dead(p.td); dead(p.tri);
  dead(p.tt); dead(p.ttau); dead(p.tmip); dead(p.ttmax); dead(p.ttr);
}
// the walk fields: a segment writes the intersection fields before reading them,
// walk_begin the shadow walk's counters
__device__ __forceinline__ void dead_walk(PathLane& p) {
  dead(p.si_pos); dead(p.si_medium); dead(p.si_hit); dead(p.sh_seg); dead(p.sh_Tr);
}

// --------------------------------------------------------------- traversal --
// Prologue of sampleDDA / sampleInternal / trDDA / trInternal on instance k.
template <bool kSingle>
__device__ __forceinline__ void start_traversal(const DScene& s, PathLane& p, int k, const Ray& rw, int kind) {
  const int ki = kSingle ? 0 : k;
  const DInstance& in = inst_at(s, ki);
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
  const DInstance& in = inst_at(s, ki);
  const DGrid& g = s.grid[in.grid];
  const float ratio = in.ratio;
  if (p.tk == TK_FLIGHT_DDA || p.tk == TK_TR_DDA) {
    if (!(p.tsteps < kLoopCap && p.tt < p.ttmax)) return true;
    ++p.tsteps;
    count<kCount>(cnt, kCntDda);
    const v3 cur = p.to + p.tt * p.td;
    const int imip = (int)(p.tmip + 0.5f);  // roundf: tmip is a multiple of 0.25 in [0, 3]
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
// transform back to world space: returns the albedo, fills the scatter record.
template <bool kSingle, bool kCount>
__device__ __forceinline__ v3 flight_collision(const DScene& s, const PathLane& p, Scatter& sc, Cnt& cnt) {
  const int ki = kSingle ? 0 : p.tinst;
  const DInstance& in = inst_at(s, ki);
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
  sc.inst = p.tinst;
  sc.pos = mul_point(in.m2w, h.pos);
  sc.n = normalize(mul_dir(in.m2w, h.normal));
#pragma unroll
  for (int i = 0; i < 6; ++i) sc.S[i] = h.S[i];
  sc.ior = h.ior;
  count<kCount>(cnt, kCntScatter);
  return alb;
}

template <bool kSingle>
__device__ __forceinline__ Ray lane_move_to_voxel_border(const DScene& s, int k, const Ray& r) {
  return inst_move_to_voxel_border(s, kSingle ? 0 : k, r);
}

// a ray spawned at a scatter point along dir, its origin offset along `off` (the
// voxel normal for shadow rays, the unnormalised sample for the continuation),
// moved to the voxel border (the tails of direct() and of Li's bounce)
template <bool kSingle>
__device__ __forceinline__ Ray spawn_at_scatter(const DScene& s, const Scatter& sc, v3 dir, v3 off) {
  Ray r = mkray(sc.pos, dir, kDefaultTmax, sc.inst);
  ray_offset(r, off);
  return lane_move_to_voxel_border<kSingle>(s, sc.inst, r);
}

// -------------------------------------------------------------------- walk --
// Begin the walk of the vertex's ray `role`: the light-sample shadow ray is in
// the walk registers already, the others wait in the lane's LDS slots.
__device__ __forceinline__ void walk_begin(PathLane& p, int role) {
  p.role = role;
  p.rng.init(p.rng.c0, p.rng.c1, 4u * (uint32_t)p.k + (uint32_t)role);
  p.stage = ST_SEG;
  if (role == RL_CONT) {
    p.ray = mkray(p.c.c_o, p.c.c_d, p.c.c_tmax, p.c.c_med);
    return;
  }
  if (role == RL_PHASE) {
    p.ray.o = p.c.lb0;
    p.ray.d = p.c.lb1;
    p.ray.medium = p.c.lb_med;
  }
  p.sh_seg = 0;
  p.sh_Tr = 1.f;
}
// walk the vertex's next ray after `after`, or go back to shading
__device__ __forceinline__ void walk_next(PathLane& p, int after) {
  const uint32_t f = p.c.flags;
  if (after < RL_LIGHT && (f & FL_LIGHT)) walk_begin(p, RL_LIGHT);
  else if (after < RL_PHASE && (f & FL_PHASE)) walk_begin(p, RL_PHASE);
  else if (after < RL_CONT && (f & FL_CONT)) walk_begin(p, RL_CONT);
  else p.stage = ST_SHADE;
}

// end of a shadow walk: the strategy's share of direct() times the throughput
// is added to the path's radiance, with Tr = 0 when the walk hit a surface
__device__ __forceinline__ void shadow_done(PathLane& p) {
  const float tr = p.si_hit ? 0.f : p.sh_Tr;
  p.L = p.L + (p.role == RL_LIGHT ? p.c.nee_a : p.c.nee_b) * tr;
  walk_next(p, p.role);
}

__device__ __forceinline__ void cont_done(PathLane& p, int cr) {
  p.cres = cr;
  p.stage = ST_SHADE;  // the continuation is the vertex's last ray
}
// the bounce loop's no-scatter branch: the environment for an escaped camera
// ray (looked up by the next shading pass), the end of the path on a miss,
// else a pass-through: the ray continues from the box it hit without a bounce
__device__ __forceinline__ void cont_nohit(PathLane& p) {
  if (!p.si_hit) {
    cont_done(p, p.b == 0 ? CR_ESCAPED : CR_END);
    return;
  }
  Ray nr = mkray(p.si_pos, normalize(p.ray.d), kDefaultTmax, p.ray.medium < 0 ? p.si_medium : -1);
  ray_offset(nr, p.ray.d);
  p.ray = nr;
  p.stage = ST_SEG;
}
// A lane at the start of a segment (ST_SEG).  For the continuation this is the
// top of Li's bounce loop: depth and iteration caps, scene intersection, and the
// start of the free flight on the current ray (inside a medium) or on the ray
// entering the medium that was hit (the entry segment is sampled again from
// inside on the pass-through that follows a miss: the reference's double
// tracking, path_integrator_impl.hpp:133-160,188-191).  For a shadow ray it is
// one intersectTr segment: spawnRay(p0, p1, currentMedium), scene intersection,
// then the transmittance traversal of the segment if it lies inside a medium.
// Both roles share one scene intersection and one traversal set-up, so a pass
// whose lanes hold both kinds of ray runs each once.
template <bool kSingle, bool kSurf>
__device__ __forceinline__ void seg_step(const DScene& s, PathLane& p) {
  dead_traversal(p);  // read by the step that ended the last traversal
  const bool dda = (s.flags & 1u) != 0u;
  const bool cont = p.role == RL_CONT;
  if (cont ? (!(p.b < s.max_depth) || ++p.iters > 4 * s.max_depth + 64 + inst_extra(s))
           : ++p.sh_seg > 64 + inst_extra(s)) {
    if (cont) {
      cont_done(p, CR_END);
    } else {
      p.si_hit = 1;
      shadow_done(p);
    }
    return;
  }
  const int cur = p.ray.medium;
  Ray r = cont ? p.ray : spawn_ray(p.ray.o, p.ray.d, cur);
  const Hit h = scene_intersect<kSurf, kSingle>(s, r);  // shortens r.tmax to the hit, as the reference's intersect does
  Ray tr = r;  // the ray of the traversal to start
  int kind;
  if (cont) {
    p.ray = r;
    p.si_pos = h.pos;
    p.si_medium = h.medium;
    p.si_hit = h.hit;
    if (h.medium < 0) {
      if (kSurf && h.surf >= 0) {  // a surface vertex (path_integrator_impl.hpp:194-223)
        p.si_medium = h.surf;
        p.ray.o = h.normal;
        cont_done(p, CR_SURFACE);
        return;
      }
      cont_nohit(p);
      return;
    }
    if (r.medium >= 0) {
      if (h.medium != r.medium) {
        cont_done(p, CR_END);
        return;
      }
    } else {
      Ray mr = mkray(h.pos, normalize(r.d), kDefaultTmax, h.medium);
      ray_offset(mr, r.d);
      const Hit se = scene_intersect<kSurf, kSingle>(s, mr);
      if (!se.hit) {
        cont_nohit(p);
        return;
      }
      if (se.medium >= 0 && h.medium != se.medium) {
        cont_done(p, CR_END);
        return;
      }
      tr = mr;
    }
    kind = dda ? TK_FLIGHT_DDA : TK_FLIGHT_DELTA;
  } else {
    p.si_hit = h.hit;
    if (kSurf && h.surf >= 0) {  // intersectTr stops at a surface, whose zero emission zeroes the term
      shadow_done(p);
      return;
    }
    p.ray.o = h.pos;
    p.ray.medium = r.medium < 0 ? h.medium : -1;
    if (!(h.hit && r.medium >= 0 && (h.medium < 0 || h.medium == cur))) {
      if (!h.hit) shadow_done(p);
      return;
    }
    kind = dda ? TK_TR_DDA : TK_TR_RATIO;
    // The segment after this one starts at the box face, outside every medium,
    // toward p1.  Its scene intersection is made now (the same one, so the
    // same result); when it meets nothing (si_hit = 2 through the traversal),
    // the traversal's last step ends the walk without another segment pass.
    if (p.sh_seg < 64 + inst_extra(s)) {
      Ray er = spawn_ray(p.ray.o, p.ray.d, p.ray.medium);
      if (!scene_intersect<kSurf, kSingle>(s, er).hit) p.si_hit = 2;
    }
  }
  start_traversal<kSingle>(s, p, tr.medium, tr, kind);
}

// A lane inside a traversal (ST_TRAV): one DDA / tracking step; at the end
// of the traversal the continuation has collided (next vertex) or leaves the
// medium box, and a shadow ray's segment ends at a box face.
template <bool kSingle, bool kCount>
__device__ __forceinline__ void trav_step(const DScene& s, PathLane& p, Cnt& cnt) {
  if (!traversal_step<kSingle, kCount>(s, p, cnt)) return;
  if (p.role == RL_CONT) {
    if (p.tres) cont_done(p, CR_COLLIDED);
    else cont_nohit(p);
  } else {
    p.sh_Tr *= p.ttr;  // intersectTr continues from the box face
    if (p.si_hit == 2) {  // the next segment, intersected already, meets nothing
      ++p.sh_seg;
      p.si_hit = 0;
      shadow_done(p);
    } else {
      p.stage = ST_SEG;
    }
  }
}

// ------------------------------------------------------------------- shade --
// The scatter vertex k (k >= 1) of a path at a collision: direct()'s two
// strategies up to their shadow walks, Li's continuation and Russian roulette,
// in the reference's order of draws from the vertex's own stream.
template <bool kSingle, bool kCount>
__device__ __forceinline__ void shade_scatter(const DScene& s, PathLane& p, const Scatter& sc, Cnt& cnt) {
  ++p.k;
  p.rng.init(p.rng.c0, p.rng.c1, 4u * (uint32_t)p.k);
  p.rng.top_up();
  uint32_t flags = 0u;
  const v3 wo = -p.ray.d;
  dead_walk(p);
  dead(p.ray.o); dead(p.ray.d); dead(p.ray.tmax); dead(p.ray.medium);
  // direct(): light sample, its shadow ray and the phase function toward it
  v3 ldir;
  float lpdf = 0.f;
  const v3 Le = env_sample<kCount>(s.env, p.rng, ldir, lpdf, cnt);
  if (lpdf > 0.f) {
    const v3 dl = normalize(ldir);
    const Ray r = spawn_at_scatter<kSingle>(s, sc, dl, sc.n);
    p.ray.o = r.o;  // walked first: straight to the walk registers (wo is taken)
    p.ray.d = r.o + r.tmax * r.d;
    p.ray.medium = r.medium;
    const float ph = phase_p(sc.S, sc.n, sc.ior, wo, dl, p.rng);
    const float mis = (1.f * lpdf) / (1.f * lpdf + 1.f * ph);
    p.c.nee_a = p.thr * ((Le * (mis * ph)) / lpdf);
    flags |= FL_LIGHT;
  }
  p.rng.top_up();
  // direct(): phase sample, its shadow ray and the environment along it
  {
    v3 wi;
    const float ph = phase_sample(sc.S, sc.n, sc.ior, wo, wi, p.rng);
    if (ph > 0.f) {
      const v3 dw = normalize(wi);
      const Ray r = spawn_at_scatter<kSingle>(s, sc, dw, sc.n);
      p.c.lb0 = r.o;
      p.c.lb1 = r.o + r.tmax * r.d;
      p.c.lb_med = r.medium;
      float epdf = 0.f;
      const v3 Le2 = env_le<kCount>(s.env, dw, epdf, cnt);
      const float mis = (1.f * ph) / (1.f * ph + 1.f * epdf);
      p.c.nee_b = p.thr * ((Le2 * (mis * ph)) / ph);
      flags |= FL_PHASE;
    }
  }
  p.rng.top_up();
  // Li: continuation direction; its ray is offset along the unnormalised sample
  {
    v3 wi;
    (void)phase_sample(sc.S, sc.n, sc.ior, wo, wi, p.rng);
    const Ray r = spawn_at_scatter<kSingle>(s, sc, normalize(wi), wi);
    p.c.c_o = r.o;
    p.c.c_d = r.d;
    p.c.c_tmax = r.tmax;
    p.c.c_med = r.medium;
    flags |= FL_CONT;
  }
  // Russian roulette (the vertex's direct light above took the throughput
  // before it)
  if (p.b > 3) {
    const float q = 1.f - (p.thr.x + p.thr.y + p.thr.z) / 3.f;
    if (p.rng.uni() <= q) flags &= ~FL_CONT;
    else p.thr = p.thr / (1.f - q);
  }
  ++p.b;
  // the bounce loop's depth test: a continuation past max_depth is never walked
  if (!(p.b < s.max_depth)) flags &= ~FL_CONT;
  p.c.flags = flags;
  p.cres = CR_NONE;
  walk_next(p, RL_VERTEX);
}

// The surface vertex k (k >= 1) of a path whose continuation hit surface
// p.si_medium at p.si_pos with normal p.ray.o: direct()'s light sample and BSDF
// sample in the tangent frame of the normal, their shadow rays starting in the
// medium the incoming ray travelled in, Li's BSDF continuation, its throughput
// and Russian roulette, in the reference's order of draws from the vertex's own
// stream.  As at a scatter vertex, the direct light of the vertex is stored
// multiplied by the throughput it arrived with, before the throughput update.
template <bool kCount>
__device__ __forceinline__ void shade_surface(const DScene& s, PathLane& p, Cnt& cnt) {
  const DSurface& sf = s.surf[p.si_medium];
  const v3 n = p.ray.o, pos = p.si_pos, rd = p.ray.d;
  const int cur = p.ray.medium;
  dead_walk(p);
  dead(p.ray.o); dead(p.ray.d); dead(p.ray.tmax); dead(p.ray.medium);
  const v3 thr0 = p.thr;
  ++p.k;
  p.rng.init(p.rng.c0, p.rng.c1, 4u * (uint32_t)p.k);
  p.rng.top_up();
  uint32_t flags = 0u;
  const SFrame fr = surface_frame(n);
  const v3 wot = mat3_mul(fr.w2t, -rd);
  {
    v3 ldir;
    float lpdf = 0.f;
    const v3 Le = env_sample<kCount>(s.env, p.rng, ldir, lpdf, cnt);
    if (lpdf > 0.f) {
      const v3 dl = normalize(ldir);
      const v3 lt = normalize(mat3_mul(fr.w2t, dl));
      Ray r = mkray(pos, dl, kDefaultTmax, cur);
      ray_offset(r, n);
      p.ray.o = r.o;  // walked first: straight to the walk registers
      p.ray.d = r.o + r.tmax * r.d;
      p.ray.medium = r.medium;
      const v3 f = mat_f(sf, wot, lt) * fabsf(dot(n, r.d));
      const float sp = mat_pdf(sf, wot, lt);
      const float mis = (1.f * lpdf) / (1.f * lpdf + 1.f * sp);
      p.c.nee_a = (thr0 * ((f * mis) * Le)) / lpdf;
      flags |= FL_LIGHT;
    }
  }
  p.rng.top_up();
  {
    float sp = 0.f;
    v3 wit = mk(0.f, 0.f, 0.f);
    v3 f = mat_sample(sf, wot, wit, p.rng, sp);
    v3 wi = normalize(mat3_mul(fr.t2w, wit));
    f = f * fabsf(dot(n, wi));
    if (sp > 0.f) {
      wi = normalize(wi);
      Ray r = mkray(pos, wi, kDefaultTmax, cur);
      ray_offset(r, n);
      p.c.lb0 = r.o;
      p.c.lb1 = r.o + r.tmax * r.d;
      p.c.lb_med = r.medium;
      float epdf = 0.f;
      const v3 Le2 = env_le<kCount>(s.env, wi, epdf, cnt);
      const float mis = (1.f * sp) / (1.f * sp + 1.f * epdf);
      p.c.nee_b = (thr0 * ((f * mis) * Le2)) / sp;
      flags |= FL_PHASE;
    }
  }
  p.rng.top_up();
  // Li: the BSDF continuation (ray.transform(worldToTangent) of the incoming ray)
  {
    const v3 wt = normalize(mat3_mul(fr.w2t, rd));
    v3 wi = mk(0.f, 0.f, 0.f);
    float pdf = 0.f;
    const v3 f = mat_sample(sf, -wt, wi, p.rng, pdf);
    if (pdf != 0.f) {
      Ray r = mkray(pos, normalize(mat3_mul(fr.t2w, wi)), kDefaultTmax, cur);
      ray_offset(r, n);
      const float cosine = fabsf(dot(n, r.d));
      p.thr = p.thr * ((f * cosine) / pdf);
      p.c.c_o = r.o;
      p.c.c_d = r.d;
      p.c.c_tmax = r.tmax;
      p.c.c_med = r.medium;
      flags |= FL_CONT;
      if (p.b > 3) {
        const float q = 1.f - (p.thr.x + p.thr.y + p.thr.z) / 3.f;
        if (p.rng.uni() <= q) flags &= ~FL_CONT;
        else p.thr = p.thr / (1.f - q);
      }
    }
  }
  ++p.b;
  if (!(p.b < s.max_depth)) flags &= ~FL_CONT;
  p.c.flags = flags;
  p.cres = CR_NONE;
  walk_next(p, RL_VERTEX);
}

}  // namespace r2
