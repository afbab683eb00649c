// Ray generation and sampler kernels: PinholeCamera::generate_ray_differential
// (R/scene/camera.rs:48-96) driven by PathSampler::camera_sample
// (R/sampler/path_sampler.rs:89-91), and the per-vertex random numbers of
// PathSampler::path_vertex_randoms (R/sampler/path_sampler.rs:97-131).
#include <hip/hip_runtime.h>

#include <string>

#include "sampler.h"
#include "tpt_device.h"

namespace tpt {

struct CamF3 {
  float x, y, z;
};
__device__ __forceinline__ CamF3 cadd(CamF3 a, CamF3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ CamF3 csub(CamF3 a, CamF3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ CamF3 cmul(float s, CamF3 a) { return {s * a.x, s * a.y, s * a.z}; }
// glam Vec3::normalize: v * (1 / sqrt(dot(v, v)))
__device__ __forceinline__ CamF3 cnormalize(CamF3 a) {
  const float r = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x * r, a.y * r, a.z * r};
}
__device__ __forceinline__ CamF3 cnormalize_or_zero(CamF3 a) {
  const float r = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  if (__builtin_isfinite(r) && r > 0.0f) return {a.x * r, a.y * r, a.z * r};
  return {0.0f, 0.0f, 0.0f};
}

struct CameraView {
  CamF3 eye, fwd, right, up;
  float tan_half, spread;
};

// direction_for_sample, R/scene/camera.rs:80-87
__device__ __forceinline__ CamF3 direction_for(const CameraView& c, float rx, float ry, float aspect, float px,
                                               float py) {
  const float sx = ((2.0f * px / rx) - 1.0f) * aspect * c.tan_half;
  const float sy = (1.0f - (2.0f * py / ry)) * c.tan_half;
  return cnormalize(cadd(cadd(c.fwd, cmul(sx, c.right)), cmul(sy, c.up)));
}

__global__ __launch_bounds__(256) void camera_kernel(CameraView cam, uint32_t w, uint32_t h, uint32_t spp,
                                                     uint32_t si, TptRay* __restrict__ rays,
                                                     float* __restrict__ diff) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= w * h) return;
  const uint32_t x = i % w, y = i / w;
  smp::PathSampler ps{smp::zsobol_frame(spp, w, h), x, y, si};
  float u[2];
  ps.s2(smp::kCameraDim, u);
  const float rx = (float)w, ry = (float)h, aspect = rx / ry;
  const float px = ((float)x + 0.5f) + 0.5f * smp::tent(u[0]);
  const float py = ((float)y + 0.5f) + 0.5f * smp::tent(u[1]);
  const CamF3 d = direction_for(cam, rx, ry, aspect, px, py);
  TptRay r;
  r.o[0] = cam.eye.x;
  r.o[1] = cam.eye.y;
  r.o[2] = cam.eye.z;
  r.tmax = __builtin_inff();
  r.d[0] = d.x;
  r.d[1] = d.y;
  r.d[2] = d.z;
  r.pad = 0.0f;
  rays[i] = r;
  if (diff) {
    // differential_scale, R/scene/camera.rs:99-101
    const float sc = fmaxf(1.0f / sqrtf((float)(spp < 1 ? 1 : spp)), 0.125f);
    const CamF3 dx = direction_for(cam, rx, ry, aspect, px + 1.0f, py);
    const CamF3 dy = direction_for(cam, rx, ry, aspect, px, py + 1.0f);
    const CamF3 rxd = cnormalize_or_zero(cadd(d, cmul(sc, csub(dx, d))));
    const CamF3 ryd = cnormalize_or_zero(cadd(d, cmul(sc, csub(dy, d))));
    float* o = diff + 7 * (size_t)i;
    o[0] = rxd.x;
    o[1] = rxd.y;
    o[2] = rxd.z;
    o[3] = ryd.x;
    o[4] = ryd.y;
    o[5] = ryd.z;
    o[6] = fmaxf(cam.spread * sc, 0.0f);  // RayCone::new(0, spread * scale)
  }
}

__global__ __launch_bounds__(256) void randoms_kernel(uint32_t w, uint32_t h, uint32_t spp, uint32_t si,
                                                      uint32_t depths, float* __restrict__ values,
                                                      uint32_t* __restrict__ seeds) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= w * h) return;
  smp::PathSampler ps{smp::zsobol_frame(spp, w, h), i % w, i / w, si};
  float* o = values + (size_t)i * (2 + 14 * depths);
  uint32_t* sd = seeds + (size_t)i * (1 + depths);
  ps.s2(smp::kCameraDim, o);
  sd[0] = ps.aux_seed(smp::kInitialAuxSeedDim);
  for (uint32_t k = 0; k < depths; ++k) {
    const uint32_t b = smp::kVertexBase + k * smp::kVertexStride;
    float* v = o + 2 + 14 * k;
    v[0] = ps.s1(b + 0);   // u_category
    v[1] = ps.s1(b + 1);   // u_tree
    v[2] = ps.s1(b + 2);   // u_light_aux
    ps.s2(b + 3, v + 3);   // u_surface
    v[5] = ps.s1(b + 5);   // u_lobe
    v[6] = ps.s1(b + 6);   // u_layer
    ps.s2(b + 7, v + 7);   // u_dir
    v[9] = ps.s1(b + 9);   // u_extra0
    v[10] = ps.s1(b + 10);
    v[11] = ps.s1(b + 11);
    v[12] = ps.s1(b + 12);  // u_extra3
    v[13] = ps.s1(b + 13);  // u_rr
    sd[1 + k] = ps.aux_seed(b + 14);
  }
}

}  // namespace tpt

using namespace tpt;

extern "C" {

int32_t tpt_camera_rays(TptDevice* d, const TptCamera* c, uint32_t w, uint32_t h, uint32_t spp, uint32_t si,
                        TptRay* rays, float* diff, uint32_t flags) {
  try {
    if (!d || !c || !rays) {
      set_error("tpt_camera_rays: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (w == 0 || h == 0) return TPT_OK;
    if ((uint64_t)w * h > 0x7FFFFFFFull || spp == 0 || si >= smp::pow2_ceil(spp)) {
      set_error("tpt_camera_rays: frame too large, spp == 0 or sample index beyond the pow2-rounded spp");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    CameraView v;
    v.eye = {c->eye[0], c->eye[1], c->eye[2]};
    v.fwd = {c->forward[0], c->forward[1], c->forward[2]};
    v.right = {c->right[0], c->right[1], c->right[2]};
    v.up = {c->image_up[0], c->image_up[1], c->image_up[2]};
    v.tan_half = c->tan_half_fov_y;
    v.spread = c->pixel_spread;
    const size_t n = (size_t)w * h;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    TptRay* drays = rays;
    float* ddiff = diff;
    DeviceBuffer tmp;
    if (!dev_ptr) {
      if (!tmp.ensure(n * (sizeof(TptRay) + (diff ? 28 : 0)))) return TPT_ERROR_DEVICE;
      drays = (TptRay*)tmp.ptr;
      ddiff = diff ? (float*)((char*)tmp.ptr + n * sizeof(TptRay)) : nullptr;
    }
    hipLaunchKernelGGL(camera_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, v, w, h, spp, si,
                       drays, ddiff);
    if (!hip_ok(hipGetLastError(), "camera_kernel launch")) return TPT_ERROR_DEVICE;
    if (!dev_ptr) {
      if (!hip_ok(hipMemcpyAsync(rays, drays, n * sizeof(TptRay), hipMemcpyDeviceToHost, d->stream), "copy") ||
          (diff && !hip_ok(hipMemcpyAsync(diff, ddiff, n * 28, hipMemcpyDeviceToHost, d->stream), "copy")) ||
          !hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize"))
        return TPT_ERROR_DEVICE;
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}

int32_t tpt_path_randoms(TptDevice* d, uint32_t w, uint32_t h, uint32_t spp, uint32_t si, uint32_t depths,
                         float* values, uint32_t* seeds, uint32_t flags) {
  try {
    if (!d || !values || !seeds) {
      set_error("tpt_path_randoms: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (w == 0 || h == 0) return TPT_OK;
    if ((uint64_t)w * h > 0x7FFFFFFFull || spp == 0 || si >= smp::pow2_ceil(spp) || depths > 4096) {
      set_error("tpt_path_randoms: bad frame, spp, sample index or depth count");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    const size_t n = (size_t)w * h;
    const size_t nv = n * (2 + 14 * (size_t)depths) * 4, ns = n * (1 + (size_t)depths) * 4;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    float* dv = values;
    uint32_t* dsd = seeds;
    DeviceBuffer tmp;
    if (!dev_ptr) {
      if (!tmp.ensure(nv + ns)) return TPT_ERROR_DEVICE;
      dv = (float*)tmp.ptr;
      dsd = (uint32_t*)((char*)tmp.ptr + nv);
    }
    hipLaunchKernelGGL(randoms_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, w, h, spp, si,
                       depths, dv, dsd);
    if (!hip_ok(hipGetLastError(), "randoms_kernel launch")) return TPT_ERROR_DEVICE;
    if (!dev_ptr) {
      if (!hip_ok(hipMemcpyAsync(values, dv, nv, hipMemcpyDeviceToHost, d->stream), "copy") ||
          !hip_ok(hipMemcpyAsync(seeds, dsd, ns, hipMemcpyDeviceToHost, d->stream), "copy") ||
          !hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize"))
        return TPT_ERROR_DEVICE;
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}

}  // extern "C"
