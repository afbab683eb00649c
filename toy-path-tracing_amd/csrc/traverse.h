// Traversal device code shared by the trace kernels and the path tracer:
// watertight triangle test (R/math/ray.rs:107-209), node slab test
// (R/qbvh.rs:578-607) over FP32 or 8-bit quantised nodes, and the per-lane
// two-level closest-hit / shadow walk in the reference's visit order
// (R/qbvh.rs:609-675, R/scene/mod.rs:340-395, 686-723).
//
// Bit-exactness with the reference CPU code rests on the same float
// expression trees (-ffp-contract=off; fma only where the reference calls
// mul_add; IEEE division and square root), the slab test's compare-select
// min/max (x86 minps/maxps semantics of wide::f32x4::fast_min/fast_max), and
// the reference visit order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tpt_internal.h"

namespace tpt {

constexpr int kBlock = 256;
constexpr int kLdsDepth = 16;     // LDS stack entries per lane of the one-thread-per-pixel kernels
constexpr int kMinLdsDepth = 4;   // the range of the trace kernels' tunable LDS depth
constexpr int kMaxLdsDepth = 32;
constexpr int kStackDepth = 128;
constexpr uint32_t kEmptyChild = kChildLeafBit;

// ---------------------------------------------------------------------------
// device math

struct F3 {
  float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }  // minps
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }  // maxps

__device__ __forceinline__ float difference_of_products(float a, float b, float c, float d) {
  float cd = c * d;
  float dop = __builtin_fmaf(a, b, -cd);
  float err = __builtin_fmaf(-c, d, cd);
  return dop + err;
}

// gamma(n) = n*eps/2 / (1 - n*eps/2) in f32, R/math/mod.rs:143-146
__device__ __forceinline__ constexpr float gamma_n(int n) {
  return ((float)n * (__FLT_EPSILON__ * 0.5f)) / (1.0f - (float)n * (__FLT_EPSILON__ * 0.5f));
}

__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }  // f32::max
__device__ __forceinline__ float max3abs(float a, float b, float c) {
  return rmax(fabsf(a), rmax(fabsf(b), fabsf(c)));
}

struct TriResult {
  float t, b0, b1, b2;
};

// A ray in the space of the mesh being traversed: origin, direction, the
// slab test's reciprocal direction, and the ray-only part of the watertight
// triangle test (R/math/ray.rs:120-144): the axis kz of the largest |d|
// component and the shear constants sx = -d.x/d.z, sy = -d.y/d.z, sz = 1/d.z
// of the permuted direction.  They are the same values the reference
// recomputes for every triangle, evaluated once per ray (kz = 3: d.z == 0,
// no triangle is hit).
struct Ray3 {
  F3 o, d, inv;
  int kz;
  float sx, sy, sz;
  // Constants of the conservative slab test (slab_cons): per axis the near
  // plane's t is fma(plane, ns, no) and the far plane's fma(plane, fs, fo),
  // rows selects the node rows that hold each axis's near plane (bit a: the
  // direction's axis a is negative, so its near plane is the child's max),
  // cons is false where the test's error bounds do not hold (a zero or
  // out-of-range direction component): such a ray takes the exact test.
  float ns[3], no[3], fs[3], fo[3];
  uint32_t rows;
  bool cons;
};

__device__ __forceinline__ void tri_setup(Ray3* r) {
  const float ax = fabsf(r->d.x), ay = fabsf(r->d.y), az = fabsf(r->d.z);
  const int kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  const float dx = kz == 0 ? r->d.y : (kz == 1 ? r->d.z : r->d.x);
  const float dy = kz == 0 ? r->d.z : (kz == 1 ? r->d.x : r->d.y);
  const float dz = kz == 0 ? r->d.x : (kz == 1 ? r->d.y : r->d.z);
  if (dz == 0.0f) {
    r->kz = 3;
    r->sx = r->sy = r->sz = 0.0f;
    return;
  }
  r->kz = kz;
  r->sx = -dx / dz;
  r->sy = -dy / dz;
  r->sz = 1.0f / dz;
}

// intersect_triangle, R/math/ray.rs:107-209.  Same operation order as the
// reference, including the fp64 recomputation of the edge functions.  The
// zero-area test ((v2 - v0) x (v1 - v0) has length 0, :114-116) depends on
// the triangle only and is precomputed into its record (bit 31 of the
// triangle id); the caller skips such triangles.
__device__ __forceinline__ bool intersect_triangle(const Ray3& r, float t_max, F3 v0, F3 v1, F3 v2, TriResult* out) {
  if (r.kz == 3) return false;
  const F3 ro = r.o;
  const int kz = r.kz;
  F3 p0 = v0 - ro, p1 = v1 - ro, p2 = v2 - ro;
  float p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
  if (kz == 0) {  // (kx, ky, kz) = (1, 2, 0)
    p0x = p0.y; p0y = p0.z; p0z = p0.x;
    p1x = p1.y; p1y = p1.z; p1z = p1.x;
    p2x = p2.y; p2y = p2.z; p2z = p2.x;
  } else if (kz == 1) {  // (2, 0, 1)
    p0x = p0.z; p0y = p0.x; p0z = p0.y;
    p1x = p1.z; p1y = p1.x; p1z = p1.y;
    p2x = p2.z; p2y = p2.x; p2z = p2.y;
  } else {  // (0, 1, 2)
    p0x = p0.x; p0y = p0.y; p0z = p0.z;
    p1x = p1.x; p1y = p1.y; p1z = p1.z;
    p2x = p2.x; p2y = p2.y; p2z = p2.z;
  }
  const float sx = r.sx, sy = r.sy, sz = r.sz;
  p0x += sx * p0z;
  p0y += sy * p0z;
  p1x += sx * p1z;
  p1y += sy * p1z;
  p2x += sx * p2z;
  p2y += sy * p2z;
  float e0 = difference_of_products(p1x, p2y, p1y, p2x);
  float e1 = difference_of_products(p2x, p0y, p2y, p0x);
  float e2 = difference_of_products(p0x, p1y, p0y, p1x);
  if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
    e0 = (float)((double)p2y * (double)p1x - (double)p2x * (double)p1y);
    e1 = (float)((double)p0y * (double)p2x - (double)p0x * (double)p2y);
    e2 = (float)((double)p1y * (double)p0x - (double)p1x * (double)p0y);
  }
  if ((e0 < 0.0f || e1 < 0.0f || e2 < 0.0f) && (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f)) return false;
  const float det = e0 + e1 + e2;
  if (det == 0.0f) return false;
  p0z *= sz;
  p1z *= sz;
  p2z *= sz;
  const float t_scaled = e0 * p0z + e1 * p1z + e2 * p2z;
  if ((det < 0.0f && (t_scaled >= 0.0f || t_scaled < t_max * det)) ||
      (det > 0.0f && (t_scaled <= 0.0f || t_scaled > t_max * det)))
    return false;
  const float inv_det = 1.0f / det;
  const float t = t_scaled * inv_det;
  const float max_zt = max3abs(p0z, p1z, p2z);
  const float delta_z = gamma_n(3) * max_zt;
  const float max_xt = max3abs(p0x, p1x, p2x);
  const float max_yt = max3abs(p0y, p1y, p2y);
  const float delta_x = gamma_n(5) * (max_xt + max_zt);
  const float delta_y = gamma_n(5) * (max_yt + max_zt);
  const float delta_e = 2.0f * (gamma_n(2) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
  const float max_e = max3abs(e0, e1, e2);
  const float delta_t =
      3.0f * (gamma_n(3) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * fabsf(inv_det);
  if (t <= delta_t) return false;
  out->t = t;
  out->b0 = e0 * inv_det;
  out->b1 = e1 * inv_det;
  out->b2 = e2 * inv_det;
  return true;
}

// ---------------------------------------------------------------------------
// scene view passed to kernels by value

struct SceneView {
  const GpuNode* nodes;      // all BLAS nodes
  const GpuNodeQ8* qnodes;   // quantised copy, same indices
  const GpuTri* tris;        // leaf-ordered triangles of all meshes
  const uint32_t* tri_rank;  // per mesh triangle index -> leaf-ordered slot
  const GpuInstance* inst;
  const GpuNode* tlas;
  const GpuNodeQ8* qtlas;
  const uint32_t* tlas_prim;  // instance ids in TLAS leaf order
  uint32_t instance_count;
  uint32_t tlas_node_count;
};


__device__ __forceinline__ float safe_inv(float d) { return d == 0.0f ? __builtin_inff() : 1.0f / d; }

// Conservative slab test: per-ray constants.
//
// The reference's slab test (R/qbvh.rs:578-607) computes, per plane p of axis
// a, r = fl(fl(p - o) * inv) and accepts a child when
// max(near_x, near_y, near_z, 0) <= min(far_x, far_y, far_z, t_max).  The
// conservative test computes each near plane as f_n = fma(p, ns, no) with
// ns = fl(inv (1 - k)), no = fl(-o ns - m) and each far plane as
// f_f = fma(p, fs, fo) with fs = fl(inv (1 + k)), fo = fl(-o fs + m), where
// k = 2^-20 and m = |o inv| 2^-21 + 2^-100.  With u = 2^-24 and T the exact t
// of the plane, |r - T| <= (2u + u^2)|T| and
//   f_n <= (1 - k)(1 + u)^2 T + u(1 + u)(|o inv|(1 + u) + m) - m(1 + u)
// so for T >= 0, f_n <= r as soon as k >= 4u (the T terms) and
// m >= 2u |o inv| (the constant terms); for T < 0 the same bound makes f_n
// negative, where the reference's near value is negative too and both are
// clamped to 0.  Symmetrically f_f >= r whenever T >= 0, and a far plane
// with T < 0 makes the reference reject the child.  Hence every child the
// reference accepts is accepted (the test answers a superset), with one
// fused multiply-add per plane, IEEE min3 / max3 and no compare-selects.
// The bounds need finite products: rays with a direction component of zero
// or of reciprocal magnitude outside [2^-60, 2^60], or |o inv| >= 2^100, keep
// the reference's test (cons = false).
__device__ __forceinline__ void cons_setup(Ray3* r) {
  constexpr float kLo = 1.0f - 0x1p-20f, kHi = 1.0f + 0x1p-20f;
  const float o[3] = {r->o.x, r->o.y, r->o.z};
  const float inv[3] = {r->inv.x, r->inv.y, r->inv.z};
  bool ok = true;
  uint32_t rows = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float ai = fabsf(inv[a]);
    const float ao = fabsf(o[a] * inv[a]);
    ok = ok && ai >= 0x1p-60f && ai <= 0x1p60f && ao < 0x1p100f;
    const float m = __builtin_fmaf(ao, 0x1p-21f, 0x1p-100f);
    const float sn = inv[a] * kLo, sf = inv[a] * kHi;
    r->ns[a] = sn;
    r->fs[a] = sf;
    r->no[a] = __builtin_fmaf(-o[a], sn, -m);
    r->fo[a] = __builtin_fmaf(-o[a], sf, m);
    rows |= (inv[a] < 0.0f ? 1u : 0u) << a;
  }
  r->rows = rows;
  r->cons = ok;
}

// glam Mat4::transform_point3 / transform_vector3 (x*px, then y*py + r, z*pz + r, w + r)
__device__ __forceinline__ void to_local(const GpuInstance& in, F3 o, F3 d, Ray3* r) {
  const float* m = in.m;
  float rx = m[0] * o.x, ry = m[1] * o.x, rz = m[2] * o.x;
  rx = m[3] * o.y + rx; ry = m[4] * o.y + ry; rz = m[5] * o.y + rz;
  rx = m[6] * o.z + rx; ry = m[7] * o.z + ry; rz = m[8] * o.z + rz;
  rx = m[9] + rx; ry = m[10] + ry; rz = m[11] + rz;
  r->o = f3(rx, ry, rz);
  rx = m[0] * d.x; ry = m[1] * d.x; rz = m[2] * d.x;
  rx = m[3] * d.y + rx; ry = m[4] * d.y + ry; rz = m[5] * d.y + rz;
  rx = m[6] * d.z + rx; ry = m[7] * d.z + ry; rz = m[8] * d.z + rz;
  r->d = f3(rx, ry, rz);
  r->inv = f3(safe_inv(rx), safe_inv(ry), safe_inv(rz));
  tri_setup(r);
  cons_setup(r);
}

__device__ __forceinline__ float q8_scale(int e) { return __int_as_float((e + 127) << 23); }

// Conservative slab test of one node (see cons_setup): the hit mask of a
// superset of the children the reference's test accepts, and per child the
// conservative entry distance lo (<= the reference's tmin; used only to order
// children).  8-bit nodes are tested against their decoded boxes (which
// contain the FP32 boxes, and the reference's test is monotone in the box),
// ray-relative: per node and axis S = 2^e ns and B = fma(origin, ns, no), so
// a plane origin + q 2^e costs one byte conversion and one fma(q, S, B).
// That form rounds S (exactly: a power-of-two scale) and B once, and drops the
// rounding of origin + q 2^e; the extra error is below
// 3u(|B| + 255|S|) + u 255 |S|, so B is moved outward by
// 2^-20 (|B| + 255|S|) (16u of the same quantity) for the near and far planes.
template <bool kQ8>
__device__ __forceinline__ uint32_t slab4_cons(const GpuNode* __restrict__ nodes, const GpuNodeQ8* __restrict__ qnodes,
                                               uint32_t ni, const Ray3& r, float t_max, float lo[4],
                                               uint32_t child[4]) {
  float tn[3][4], tf[3][4];
  bool fp32 = !kQ8;
  if (kQ8) {
    const uint4* qp = reinterpret_cast<const uint4*>(qnodes + ni);
    const uint4 a = qp[0], b = qp[1], c = qp[2], ch = qp[3];
    if ((a.w >> 24) & 1u) {
      fp32 = true;
    } else {
      const float org[3] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)};
      // q words: b.x = lox[0..3], b.y = hix, b.z = loy, b.w = hiy, c.x = loz, c.y = hiz
      const uint32_t wlo[3] = {b.x, b.z, c.x}, whi[3] = {b.y, b.w, c.y};
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const int e = (int)(int8_t)((a.w >> (8 * ax)) & 0xFFu);
        const bool neg = (r.rows >> ax) & 1u;
        const uint32_t wn = neg ? whi[ax] : wlo[ax], wf = neg ? wlo[ax] : whi[ax];
        const float sn = __builtin_ldexpf(r.ns[ax], e), sf = __builtin_ldexpf(r.fs[ax], e);
        float bn = __builtin_fmaf(org[ax], r.ns[ax], r.no[ax]);
        float bf = __builtin_fmaf(org[ax], r.fs[ax], r.fo[ax]);
        bn = __builtin_fmaf(-0x1p-20f, __builtin_fmaf(255.0f, fabsf(sn), fabsf(bn)), bn);
        bf = __builtin_fmaf(0x1p-20f, __builtin_fmaf(255.0f, fabsf(sf), fabsf(bf)), bf);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          tn[ax][i] = __builtin_fmaf((float)((wn >> (8 * i)) & 0xFFu), sn, bn);
          tf[ax][i] = __builtin_fmaf((float)((wf >> (8 * i)) & 0xFFu), sf, bf);
        }
      }
      child[0] = ch.x;
      child[1] = ch.y;
      child[2] = ch.z;
      child[3] = ch.w;
    }
  }
  if (fp32) {
    const float4* p = reinterpret_cast<const float4*>(nodes + ni);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const uint32_t neg = (r.rows >> ax) & 1u;
      const float4 nr = p[2 * ax + neg], fr = p[2 * ax + 1 - neg];
      const float nv[4] = {nr.x, nr.y, nr.z, nr.w}, fv[4] = {fr.x, fr.y, fr.z, fr.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        tn[ax][i] = __builtin_fmaf(nv[i], r.ns[ax], r.no[ax]);
        tf[ax][i] = __builtin_fmaf(fv[i], r.fs[ax], r.fo[ax]);
      }
    }
    const uint4 ch = *reinterpret_cast<const uint4*>(p + 6);
    child[0] = ch.x;
    child[1] = ch.y;
    child[2] = ch.z;
    child[3] = ch.w;
  }
  uint32_t bits = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float l = fmaxf(fmaxf(tn[0][i], tn[1][i]), fmaxf(tn[2][i], 0.0f));
    const float h = fminf(fminf(tf[0][i], tf[1][i]), fminf(tf[2][i], t_max));
    lo[i] = l;
    if (l <= h && child[i] != kEmptyChild) bits |= 1u << i;
  }
  return bits;
}

// Node slab test, intersect_4_aabbs (R/qbvh.rs:578-607).  Writes per-slot
// tmin and child ref; returns the 4-bit hit mask restricted to non-empty slots.
template <bool kQ8>
__device__ __forceinline__ uint32_t slab4(const GpuNode* __restrict__ nodes, const GpuNodeQ8* __restrict__ qnodes,
                                          uint32_t ni, const Ray3& r, float t_max, float tmin[4],
                                          uint32_t child[4]) {
  float lx[4], hx[4], ly[4], hy[4], lz[4], hz[4];
  bool fp32 = !kQ8;
  if (kQ8) {
    const uint4* qp = reinterpret_cast<const uint4*>(qnodes + ni);
    const uint4 a = qp[0], b = qp[1], c = qp[2], ch = qp[3];
    const uint32_t flags = (a.w >> 24) & 0xFFu;
    if (flags & 1u) {
      fp32 = true;
    } else {
      const float ox = __uint_as_float(a.x), oy = __uint_as_float(a.y), oz = __uint_as_float(a.z);
      const float sxs = q8_scale((int)(int8_t)(a.w & 0xFFu));
      const float sys = q8_scale((int)(int8_t)((a.w >> 8) & 0xFFu));
      const float szs = q8_scale((int)(int8_t)((a.w >> 16) & 0xFFu));
      // q bytes: b.x = lox[0..3], b.y = hix, b.z = loy, b.w = hiy, c.x = loz, c.y = hiz.
      // A plane is origin + q * 2^e: the product of a byte and a power of two
      // is exact, so one fused multiply-add gives the bits of the rounded sum
      // origin + (q * 2^e) that the builder's containment check used.
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int sh = 8 * i;
        lx[i] = __builtin_fmaf((float)((b.x >> sh) & 0xFFu), sxs, ox);
        hx[i] = __builtin_fmaf((float)((b.y >> sh) & 0xFFu), sxs, ox);
        ly[i] = __builtin_fmaf((float)((b.z >> sh) & 0xFFu), sys, oy);
        hy[i] = __builtin_fmaf((float)((b.w >> sh) & 0xFFu), sys, oy);
        lz[i] = __builtin_fmaf((float)((c.x >> sh) & 0xFFu), szs, oz);
        hz[i] = __builtin_fmaf((float)((c.y >> sh) & 0xFFu), szs, oz);
      }
      (void)c;
      child[0] = ch.x;
      child[1] = ch.y;
      child[2] = ch.z;
      child[3] = ch.w;
    }
  }
  if (fp32) {
    const float4* p = reinterpret_cast<const float4*>(nodes + ni);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4], f = p[5];
    const uint4 ch = *reinterpret_cast<const uint4*>(p + 6);
    lx[0] = a.x; lx[1] = a.y; lx[2] = a.z; lx[3] = a.w;
    hx[0] = b.x; hx[1] = b.y; hx[2] = b.z; hx[3] = b.w;
    ly[0] = c.x; ly[1] = c.y; ly[2] = c.z; ly[3] = c.w;
    hy[0] = d.x; hy[1] = d.y; hy[2] = d.z; hy[3] = d.w;
    lz[0] = e.x; lz[1] = e.y; lz[2] = e.z; lz[3] = e.w;
    hz[0] = f.x; hz[1] = f.y; hz[2] = f.z; hz[3] = f.w;
    child[0] = ch.x;
    child[1] = ch.y;
    child[2] = ch.z;
    child[3] = ch.w;
  }
  uint32_t bits = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float t1x = (lx[i] - r.o.x) * r.inv.x, t2x = (hx[i] - r.o.x) * r.inv.x;
    const float t1y = (ly[i] - r.o.y) * r.inv.y, t2y = (hy[i] - r.o.y) * r.inv.y;
    const float t1z = (lz[i] - r.o.z) * r.inv.z, t2z = (hz[i] - r.o.z) * r.inv.z;
    const float tminx = sel_min(t1x, t2x), tmaxx = sel_max(t1x, t2x);
    const float tminy = sel_min(t1y, t2y), tmaxy = sel_max(t1y, t2y);
    const float tminz = sel_min(t1z, t2z), tmaxz = sel_max(t1z, t2z);
    const float tn = sel_max(sel_max(sel_max(tminx, tminy), tminz), 0.0f);
    const float tx = sel_min(sel_min(sel_min(tmaxx, tmaxy), tmaxz), t_max);
    tmin[i] = tn;
    if (tn <= tx && child[i] != kEmptyChild) bits |= 1u << i;
  }
  return bits;
}

// Per-lane traversal stack: depth() entries in LDS laid out [depth][thread]
// (the bank is the thread, so lanes at different depths never conflict),
// continued in a private array up to the reference's 128 entries
// (R/qbvh.rs:609).  kDepth > 0 fixes the LDS depth at compile time (the
// path kernels and the trace kernels' default); kDepth == 0 reads it from
// lds_depth (the trace kernels at a tuned depth, LDS sized at launch).  A
// runtime depth costs the shadow kernel two VGPRs (82 against 80: five
// waves per SIMD instead of six), so the default stays a constant.  The
// stack pointer is a register of its own: it must not share an aggregate
// with the dynamically indexed spill array, or it would be kept in scratch.
template <int kDepth>
struct LaneStackT {
  uint32_t* lds;    // lds[d * kBlock] is this thread's entry d
  uint32_t* spill;  // private array of at least kStackDepth - depth() entries
  int sp;
  int lds_depth;    // read only when kDepth == 0
  bool overflow;
  __device__ __forceinline__ int depth() const {
    if constexpr (kDepth > 0) return kDepth;
    else return lds_depth;
  }
  __device__ __forceinline__ void push(uint32_t v) {
    if (sp < depth()) {
      lds[sp * kBlock] = v;
    } else if (sp < kStackDepth) {
      spill[sp - depth()] = v;
    } else {
      overflow = true;  // the reference panics here; the entry is dropped
      return;
    }
    ++sp;
  }
  __device__ __forceinline__ uint32_t pop() {
    --sp;
    return sp < depth() ? lds[sp * kBlock] : spill[sp - depth()];
  }
};
using LaneStack = LaneStackT<kLdsDepth>;

// Leaves of the last visited node, in processing order (q0 first).
struct LeafQueue {
  uint32_t q0, q1, q2, q3;
  int n;
  __device__ __forceinline__ void clear() {
    q0 = q1 = q2 = q3 = 0;
    n = 0;
  }
  __device__ __forceinline__ uint32_t pop() {
    const uint32_t v = q0;
    q0 = q1;
    q1 = q2;
    q2 = q3;
    n--;
    return v;
  }
};

__device__ __forceinline__ void cswap_desc(float& ta, uint32_t& ca, int& sa, float& tb, uint32_t& cb, int& sb) {
  // descending by (t, slot): swap when (ta, sa) < (tb, sb)
  const bool sw = (ta < tb) || (ta == tb && sa < sb);
  if (sw) {
    float t = ta; ta = tb; tb = t;
    uint32_t c = ca; ca = cb; cb = c;
    int s = sa; sa = sb; sb = s;
  }
}

// One node visit.  The reference sorts the hit children by tmin with a stable
// insertion sort (ties keep slot order; a hit tmin is never NaN) and walks
// them far->near: leaves are handled at once, interiors pushed so that the
// nearest pops first.  The same order results from a 4-input sorting network
// on (tmin, slot) descending; pushes and leaf handling are independent, so
// interiors are pushed here and leaves queued to run before the next pop.
//
// kOrder: kVisitReference is that order.  The early-exit any-hit query's
// answer does not depend on the order; for it kVisitSlot keeps slot order
// (measured fastest) and kVisitNearLast only moves the nearest hit child to
// the end of the walk, so that it pops first.
enum VisitOrder { kVisitReference = 0, kVisitNearLast = 1, kVisitSlot = 2 };

//
// kCons: test the children with the conservative slab test (slab4_cons) on
// rays that allow it; it accepts every child the reference's test accepts,
// so the walk only tests more triangles, never fewer, and orders children by
// its conservative entry distances (an order that can differ from the
// reference's only between children whose entry distances tie within
// rounding).
template <bool kQ8, int kOrder = kVisitReference, bool kCons = false, class Stack = LaneStack>
__device__ __forceinline__ void visit_node(const GpuNode* nodes, const GpuNodeQ8* qnodes, uint32_t ni,
                                           const Ray3& r, float t_max, Stack& st, LeafQueue& lq) {
  float tmin[4];
  uint32_t child[4];
  uint32_t bits;
  if (kCons && r.cons)
    bits = slab4_cons<kQ8>(nodes, qnodes, ni, r, t_max, tmin, child);
  else
    bits = slab4<kQ8>(nodes, qnodes, ni, r, t_max, tmin, child);
  float t[4];
  uint32_t c[4];
  int s[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool h = (bits >> i) & 1u;
    t[i] = h ? tmin[i] : -1.0f;  // misses sort last (a hit tmin is >= 0)
    c[i] = child[i];
    s[i] = h ? i : -1;
  }
  if (kOrder == kVisitReference) {
    cswap_desc(t[0], c[0], s[0], t[1], c[1], s[1]);
    cswap_desc(t[2], c[2], s[2], t[3], c[3], s[3]);
    cswap_desc(t[0], c[0], s[0], t[2], c[2], s[2]);
    cswap_desc(t[1], c[1], s[1], t[3], c[3], s[3]);
    cswap_desc(t[1], c[1], s[1], t[2], c[2], s[2]);
  } else if (kOrder == kVisitNearLast) {
    // the nearest hit child into slot 3 (a miss has t = -1, any hit is larger)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const bool sw = s[i] >= 0 && (s[3] < 0 || t[i] < t[3]);
      if (sw) {
        float tt = t[i]; t[i] = t[3]; t[3] = tt;
        uint32_t cc = c[i]; c[i] = c[3]; c[3] = cc;
        int ss = s[i]; s[i] = s[3]; s[3] = ss;
      }
    }
  }
  lq.clear();
  uint32_t q[4] = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (s[k] < 0) continue;
    const uint32_t ck = c[k];
    if (ck & kChildLeafBit) {
      if (((ck >> kChildCountShift) & kChildCountMask) == 0) continue;
      if (lq.n == 0) q[0] = ck;
      else if (lq.n == 1) q[1] = ck;
      else if (lq.n == 2) q[2] = ck;
      else q[3] = ck;
      lq.n++;
    } else {
      st.push(ck);
    }
  }
  lq.q0 = q[0];
  lq.q1 = q[1];
  lq.q2 = q[2];
  lq.q3 = q[3];
}

enum QueryMode { kClosest = 0, kAny = 1 };

// Child order of the early-exit any-hit walk (the exact rule runs the
// closest-hit walk and keeps the reference order): slot order, the fastest
// of the three on a 2.1M-triangle field and the Cornell path tracer.
constexpr int kAnyVisitOrder = kVisitSlot;

struct Counters {
  uint32_t nodes, tris, insts;
};

// Traversal state of one lane.  Single-level scenes (one instance, no TLAS
// nodes) keep only the BLAS part; two-level scenes interleave TLAS visits and
// BLAS walks on the same stack, BLAS entries above `blas_base`.
struct Lane {
  Ray3 r;         // ray in the current instance's space
  F3 wo, wd, winv;// world ray (two-level)
  float closest;  // current t_max
  float strict;   // kAny: hits at t >= strict are ignored
  uint32_t skip_inst, skip_tri;
  // result
  uint32_t inst, tri;
  float t, b0, b1, b2;
  bool hit;
  // BLAS walk
  uint32_t tri_cur, tri_end, tri_base;
  LeafQueue bq;
  // two-level
  int level, blas_base;
  uint32_t cur_inst, inst_cur, inst_end;
  LeafQueue tq;
};

template <class Stack>
__device__ __forceinline__ void enter_instance(const SceneView& sc, Lane& L, Stack& st, uint32_t ii,
                                               const F3& wo, const F3& wd) {
  const GpuInstance in = sc.inst[ii];
  to_local(in, wo, wd, &L.r);
  L.cur_inst = ii;
  L.tri_base = in.tri_base;
  L.blas_base = st.sp;
  L.bq.clear();
  if (in.node_count == 0) {  // root leaf, R/qbvh.rs:615-620
    L.tri_cur = in.tri_base;
    L.tri_end = in.tri_base + in.tri_count;
  } else {
    L.tri_cur = L.tri_end = 0;
    st.push(in.node_base);
  }
  L.level = 1;
}

template <bool kTwoLevel, class Stack>
__device__ __forceinline__ void lane_begin(const SceneView& sc, Lane& L, Stack& st, F3 o, F3 d, float t_max,
                                           float strict, uint32_t skip_inst, uint32_t skip_tri) {
  L.closest = t_max;
  L.strict = strict;
  L.skip_inst = skip_inst;
  L.skip_tri = skip_tri;
  L.hit = false;
  L.inst = L.tri = TPT_INVALID_INDEX;
  L.t = __builtin_inff();
  L.b0 = L.b1 = L.b2 = 0.0f;
  L.tri_cur = L.tri_end = 0;
  L.bq.clear();
  L.tq.clear();
  st.sp = 0;
  st.overflow = false;
  if (kTwoLevel) {
    L.wo = o;
    L.wd = d;
    L.winv = f3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
    L.level = 0;
    L.inst_cur = L.inst_end = 0;
    if (sc.tlas_node_count == 0)
      L.inst_end = sc.instance_count;  // TLAS root leaf over all instances
    else
      st.push(0);
  } else {
    enter_instance(sc, L, st, 0, o, d);
  }
}

// One iteration: at most one node visit and one triangle test (the "if-if"
// form: a lane that pops a node with leaf children tests a triangle in the
// same iteration).  Returns true when the lane's query is finished.
template <bool kQ8, int kMode, bool kCount, bool kTwoLevel, class Stack>
__device__ __forceinline__ bool lane_step(const SceneView& sc, Lane& L, Stack& st, Counters& cnt) {
  if (kTwoLevel && L.level == 0) {
    if (L.inst_cur >= L.inst_end) {
      if (L.tq.n == 0) {
        if (st.sp == 0) return true;
        const uint32_t ni = st.pop();
        if (kCount) cnt.nodes++;
        Ray3 w;
        w.o = L.wo;
        w.d = L.wd;
        w.inv = L.winv;
        visit_node<kQ8, kMode == kAny ? kAnyVisitOrder : kVisitReference>(sc.tlas, sc.qtlas, ni, w, L.closest, st,
                                                                        L.tq);
      }
      if (L.tq.n > 0) {
        const uint32_t c = L.tq.pop();
        L.inst_cur = c & kChildOffsetMask;
        L.inst_end = L.inst_cur + ((c >> kChildCountShift) & kChildCountMask);
      }
    }
    if (L.inst_cur >= L.inst_end) return false;
    if (kCount) cnt.insts++;
    enter_instance(sc, L, st, sc.tlas_prim[L.inst_cur++], L.wo, L.wd);
  }
  // BLAS part
  if (L.tri_cur >= L.tri_end) {
    if (L.bq.n == 0) {
      if (st.sp == (kTwoLevel ? L.blas_base : 0)) {
        if (kTwoLevel) {
          L.level = 0;
          return false;
        }
        return true;
      }
      const uint32_t ni = st.pop();
      if (kCount) cnt.nodes++;
      // The conservative slab test serves the early-exit any-hit walk (its
      // answer depends only on which triangles are tested) and 8-bit nodes
      // (whose boxes already only grow the candidate set); FP32 closest-hit
      // walks keep the reference's test and visit order.
      visit_node<kQ8, kMode == kAny ? kAnyVisitOrder : kVisitReference, (kMode == kAny || kQ8)>(
          sc.nodes, sc.qnodes, ni, L.r, L.closest, st, L.bq);
    }
    if (L.bq.n > 0) {
      const uint32_t c = L.bq.pop();
      L.tri_cur = L.tri_base + (c & kChildOffsetMask);
      L.tri_end = L.tri_cur + ((c >> kChildCountShift) & kChildCountMask);
    }
  }
  if (L.tri_cur < L.tri_end) {
    const GpuTri* tp = sc.tris + L.tri_cur;
    const float4 a = reinterpret_cast<const float4*>(tp)[0];
    const float4 b = reinterpret_cast<const float4*>(tp)[1];
    const float2 c = reinterpret_cast<const float2*>(tp)[4];
    const uint32_t prim = __float_as_uint(c.y) & kTriIdMask;
    const bool degenerate = (__float_as_uint(c.y) & kTriDegenerate) != 0;
    ++L.tri_cur;
    const bool skip = (kMode == kAny) && L.cur_inst == L.skip_inst && prim == L.skip_tri;
    if (!skip) {
      if (kCount) cnt.tris++;
      TriResult h;
      if (!degenerate &&
          intersect_triangle(L.r, L.closest, f3(a.x, a.y, a.z), f3(a.w, b.x, b.y), f3(b.z, b.w, c.x), &h) &&
          (kMode != kAny || h.t < L.strict)) {
        L.closest = h.t;
        L.hit = true;
        L.inst = L.cur_inst;
        L.tri = prim;
        L.t = h.t;
        L.b0 = h.b0;
        L.b1 = h.b1;
        L.b2 = h.b2;
        if (kMode == kAny) return true;
      }
    }
  }
  return false;
}

__device__ __forceinline__ void load_ray(const TptRay* rays, uint32_t i, F3* o, F3* d, float* tmax) {
  const float4* p = reinterpret_cast<const float4*>(rays + i);
  const float4 a = p[0], b = p[1];
  *o = f3(a.x, a.y, a.z);
  *tmax = a.w;
  *d = f3(b.x, b.y, b.z);
}


}  // namespace tpt
