// Spatial adaptivity around the MGPBD solve: the active/passive tet zone of
// cons/adaptive_zone.py and the compliance mask of cons/mgpbd_adaptive.py. The zone is an
// int32 flag per tet on the device; the mask inflates alpha~ of passive tets (their
// constraints drop out of the solve) and keeps the original values for the restore.
//
// Float semantics follow the reference kernels: FP32 differences and squares summed left to
// right, correctly rounded square root and division, no contraction into FMAs, so the flags
// match a numpy FP32 restatement bit for bit.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>

#include "mgpbd_internal.h"

namespace mgpbd_zone {

constexpr int kBlock = 256;

inline int blocks(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// |p - c|^2 = (dx dx + dy dy) + dz dz in FP32 without contraction
__device__ __forceinline__ float dist2(float px, float py, float pz, float cx, float cy, float cz) {
  const float dx = __fsub_rn(px, cx), dy = __fsub_rn(py, cy), dz = __fsub_rn(pz, cz);
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// active[t] = 1 when some vertex of t lies strictly inside the ball (cons/adaptive_zone.py:35-48)
__global__ void __launch_bounds__(kBlock)
k_proximity(int nt, const int* __restrict__ t_i, const float* __restrict__ x, float cx, float cy,
            float cz, float r2, int* __restrict__ active) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  const int4 v = reinterpret_cast<const int4*>(t_i)[t];
  const int ids[4] = {v.x, v.y, v.z, v.w};
  int a = 0;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const float* p = x + 3 * (size_t)ids[l];
    if (dist2(p[0], p[1], p[2], cx, cy, cz) < r2) a = 1;
  }
  active[t] = a;
}

// active[t] = 1 when the fastest vertex of t moves faster than threshold:
// max_l |x_l - x_old_l| / dt > threshold (cons/adaptive_zone.py:50-64)
__global__ void __launch_bounds__(kBlock)
k_strain(int nt, const int* __restrict__ t_i, const float* __restrict__ x,
         const float* __restrict__ x_old, float threshold, float dt, int* __restrict__ active) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  const int4 v = reinterpret_cast<const int4*>(t_i)[t];
  const int ids[4] = {v.x, v.y, v.z, v.w};
  float vmax = 0.0f;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const float* p = x + 3 * (size_t)ids[l];
    const float* q = x_old + 3 * (size_t)ids[l];
    const float vel = __fdiv_rn(__fsqrt_rn(dist2(p[0], p[1], p[2], q[0], q[1], q[2])), dt);
    if (vel > vmax) vmax = vel;
  }
  active[t] = vmax > threshold ? 1 : 0;
}

// One ring of face neighbours: out[t] = in[t] or some face neighbour active (the face graph
// is symmetric, so this is the reference's scatter to the neighbours of active tets,
// cons/adaptive_zone.py:66-80, written as a gather)
__global__ void __launch_bounds__(kBlock)
k_expand(int nt, const int* __restrict__ nbr, const int* __restrict__ in, int* __restrict__ out) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  int a = in[t] != 0;
  const int4 nb = reinterpret_cast<const int4*>(nbr)[t];
  if (nb.x >= 0 && in[nb.x]) a = 1;
  if (nb.y >= 0 && in[nb.y]) a = 1;
  if (nb.z >= 0 && in[nb.z]) a = 1;
  if (nb.w >= 0 && in[nb.w]) a = 1;
  out[t] = a;
}

// saved = alpha; alpha[t] = passive_value where active[t] == 0 (cons/mgpbd_adaptive.py:64-79)
__global__ void __launch_bounds__(kBlock)
k_mask(int nt, const int* __restrict__ active, float* __restrict__ alpha, float* __restrict__ saved,
       float passive_value) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  const float a = alpha[t];
  saved[t] = a;
  if (active[t] == 0) alpha[t] = passive_value;
}

}  // namespace mgpbd_zone

using namespace mgpbd_zone;

#define ZONE_CHECK(ctx, name)                                                     \
  do {                                                                            \
    if (!(ctx)) return mgpbd_fail(MGPBD_EINVAL, std::string(name) + ": null context"); \
  } while (0)

extern "C" {

int mgpbd_zone_proximity(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_p,
                         float cx, float cy, float cz, float radius, int32_t* active) {
  ZONE_CHECK(ctx, "mgpbd_zone_proximity");
  if (n_tets < 0 || (n_tets > 0 && (!t_i || !v_p || !active)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_zone_proximity: invalid arguments");
  if (n_tets == 0) return 0;
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  const float r2 = radius * radius;   // one FP32 product (the reference's r2 = radius * radius)
  k_proximity<<<blocks(n_tets), kBlock, 0, s>>>(n_tets, t_i, v_p, cx, cy, cz, r2, active);
  MGPBD_LAUNCH_CHECK("mgpbd_zone_proximity");
  return 0;
}

int mgpbd_zone_strain(mgpbd_ctx* ctx, int n_tets, const int32_t* t_i, const float* v_p,
                      const float* v_p_old, float threshold, float dt, int32_t* active) {
  ZONE_CHECK(ctx, "mgpbd_zone_strain");
  if (n_tets < 0 || !(dt > 0.0f) || (n_tets > 0 && (!t_i || !v_p || !v_p_old || !active)))
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_zone_strain: invalid arguments");
  if (n_tets == 0) return 0;
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  k_strain<<<blocks(n_tets), kBlock, 0, s>>>(n_tets, t_i, v_p, v_p_old, threshold, dt, active);
  MGPBD_LAUNCH_CHECK("mgpbd_zone_strain");
  return 0;
}

int mgpbd_zone_expand(mgpbd_ctx* ctx, int n_tets, const int32_t* nbr, int32_t* active,
                      int32_t* scratch, int n_rings) {
  ZONE_CHECK(ctx, "mgpbd_zone_expand");
  if (n_tets < 0 || n_rings < 0 || (n_tets > 0 && (!nbr || !active || !scratch)) ||
      active == scratch)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_zone_expand: invalid arguments");
  if (n_tets == 0 || n_rings == 0) return 0;
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  int32_t* in = active;
  int32_t* out = scratch;
  for (int r = 0; r < n_rings; ++r) {
    k_expand<<<blocks(n_tets), kBlock, 0, s>>>(n_tets, nbr, in, out);
    std::swap(in, out);
  }
  if (in != active)
    MGPBD_HIP_TRY(hipMemcpyAsync(active, in, sizeof(int32_t) * (size_t)n_tets,
                                 hipMemcpyDeviceToDevice, s));
  MGPBD_LAUNCH_CHECK("mgpbd_zone_expand");
  return 0;
}

int mgpbd_zone_mask_alpha(mgpbd_ctx* ctx, int n_tets, const int32_t* active, float* alpha,
                          float* saved, float passive_value) {
  ZONE_CHECK(ctx, "mgpbd_zone_mask_alpha");
  if (n_tets < 0 || (n_tets > 0 && (!active || !alpha || !saved)) || alpha == saved)
    return mgpbd_fail(MGPBD_EINVAL, "mgpbd_zone_mask_alpha: invalid arguments");
  if (n_tets == 0) return 0;
  hipStream_t s = (hipStream_t)mgpbd_internal_stream(ctx);
  k_mask<<<blocks(n_tets), kBlock, 0, s>>>(n_tets, active, alpha, saved, passive_value);
  MGPBD_LAUNCH_CHECK("mgpbd_zone_mask_alpha");
  return 0;
}

}  // extern "C"
