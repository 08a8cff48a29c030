// FP32 device math for the path tracer: vectors, orthonormal frames, PCG32,
// sample transforms and the spectral tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "device_scene.hpp"

// Device helpers are also callable from host code, so the same functions can be
// exercised on the CPU (tests and host-side checks).
#define TT_HD __host__ __device__ __forceinline__

namespace ttgpu {

TT_HD int as_int(float x) { return __builtin_bit_cast(int, x); }
TT_HD float as_float(int x) { return __builtin_bit_cast(float, x); }
TT_HD float as_float(uint32_t x) { return __builtin_bit_cast(float, x); }
TT_HD uint32_t as_uint(float x) { return __builtin_bit_cast(uint32_t, x); }
// Fast transcendental instructions on the GPU; libm on the host.
TT_HD void fast_sincos(float x, float *s, float *c) {
#if defined(__HIP_DEVICE_COMPILE__)
  __sincosf(x, s, c);
#else
  *s = sinf(x);
  *c = cosf(x);
#endif
}
TT_HD float fast_exp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(x);
#else
  return expf(x);
#endif
}
TT_HD float fast_log(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __logf(x);
#else
  return logf(x);
#endif
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kFltEps = 1.1920928955078125e-07f;

TT_HD float3 mk(float x, float y, float z) { return make_float3(x, y, z); }
// float3/float4 arithmetic (+, -, *, unary -) comes from HIP's vector types.
TT_HD float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TT_HD float3 cross(float3 a, float3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
TT_HD float length(float3 a) { return sqrtf(dot(a, a)); }
TT_HD float3 normalize(float3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
TT_HD float3 xyz(float4 a) { return mk(a.x, a.y, a.z); }
TT_HD float sqr(float x) { return x * x; }
TT_HD float clampdot(float3 a, float3 b) { return fmaxf(0.f, dot(a, b)); }

TT_HD float4 f4(float v) { return make_float4(v, v, v, v); }
TT_HD float at(float4 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : (i == 2 ? a.z : a.w)); }
TT_HD bool all_zero(float4 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f && a.w == 0.f; }

// Orthonormal frame with z = n, Duff et al. (2017) "Building an Orthonormal Basis, Revisited".
struct Frame {
  float3 x, y, z;
  TT_HD float3 to_world(float3 v) const { return x * v.x + y * v.y + z * v.z; }
  TT_HD float3 to_local(float3 v) const { return mk(dot(x, v), dot(y, v), dot(z, v)); }
};
TT_HD Frame make_frame(float3 n) {
  float sign = copysignf(1.0f, n.z);
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  Frame f;
  f.x = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  f.y = mk(b, sign + n.y * n.y * a, -n.y);
  f.z = n;
  return f;
}

// ---- PCG32 (O'Neill 2014), XSH-RR output of a 64-bit LCG -------------------
struct Pcg {
  uint64_t state, inc;
  TT_HD uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
  }
  TT_HD void seed(uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  // Uniform float in [0,1) from the top 24 bits.
  TT_HD float uniform() { return (float)(next() >> 8) * 5.9604644775390625e-08f; }
  TT_HD uint32_t bounded(uint32_t bound) {
    uint32_t threshold = (~bound + 1u) % bound;
    for (;;) {
      uint32_t r = next();
      if (r >= threshold) return r % bound;
    }
  }
};

TT_HD Pcg load_rng(const PathQueues &q, uint32_t i) {
  Pcg r;
  uint2 s = q.rng_state[i], c = q.rng_inc[i];
  r.state = ((uint64_t)s.y << 32) | s.x;
  r.inc = ((uint64_t)c.y << 32) | c.x;
  return r;
}
TT_HD void store_rng(const PathQueues &q, uint32_t i, const Pcg &r) {
  q.rng_state[i] = make_uint2((uint32_t)r.state, (uint32_t)(r.state >> 32));
}

// ---- sample transforms (src/sampler.cxx:6-125) ------------------------------
TT_HD float3 to_uniform_sphere(float r0, float r1) {
  float z = 1.f - 2.f * r1;
  float s = sqrtf(fmaxf(0.f, r1 * (1.f - r1)));
  float sn, cs;
  fast_sincos(2.f * kPi * r0, &sn, &cs);
  return mk(2.f * cs * s, 2.f * sn * s, z);
}
TT_HD float3 to_uniform_hemisphere(float r0, float r1) {
  float3 v = to_uniform_sphere(r0, r1);
  v.z = fabsf(v.z);
  return v;
}
TT_HD float3 to_cos_hemisphere(float r0, float r1) {
  float rho = sqrtf(1.f - r0), z = sqrtf(r0);
  float sn, cs;
  fast_sincos(2.f * kPi * r1, &sn, &cs);
  return mk(cs * rho, sn * rho, z);
}
TT_HD float3 to_sphere_section(float cos_open, float r0, float r1) {
  float z = 1.f - r1 * (1.f - cos_open);
  float s = sqrtf(fmaxf(0.f, 1.f - z * z));
  float sn, cs;
  fast_sincos(2.f * kPi * r0, &sn, &cs);
  return mk(cs * s, sn * s, z);
}
TT_HD float sphere_section_pdf(float cos_open) { return 1.f / (2.f * kPi * (1.f - cos_open)); }
TT_HD float3 to_sphere_section(float r0, float r1, float phi0, float z0, float phi1, float z1) {
  float z = z0 + r1 * (z1 - z0);
  float s = sqrtf(fmaxf(0.f, 1.f - z * z));
  float sn, cs;
  fast_sincos(phi0 + (phi1 - phi0) * r0, &sn, &cs);
  return mk(cs * s, sn * s, z);
}

// Area -> solid angle (src/sampler.hxx:523-528).
TT_HD float area_to_solid_angle(float len, float3 dir, float3 n) {
  return len * len / fabsf(dot(dir, n) + 2.220446049250313e-16f);
}

// ---- spectral tables in constant memory ------------------------------------
__constant__ float c_spec_to_rgb[3][kNBins];
__constant__ float c_smits[7][kNBins];

TT_HD int bin_of(uint32_t packed, int i) { return (packed >> (8 * i)) & 0xFF; }

TT_HD float4 take(const float *spec, uint32_t bins) {
  return make_float4(spec[bin_of(bins, 0)], spec[bin_of(bins, 1)], spec[bin_of(bins, 2)], spec[bin_of(bins, 3)]);
}

// SpectralSelectionToRGB (src/spectral.cxx:130-140).
TT_HD float3 spectral_to_rgb(float4 v, uint32_t bins) {
  float3 r = mk(0.f, 0.f, 0.f);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int b = bin_of(bins, j);
    float x = at(v, j);
    r.x += x * c_spec_to_rgb[0][b];
    r.y += x * c_spec_to_rgb[1][b];
    r.z += x * c_spec_to_rgb[2][b];
  }
  return r;
}

// RGBToSpectralSelection (src/spectral.cxx:143-177): Smits basis on the selected bins.
TT_HD float4 rgb_to_spectral(float3 c, uint32_t bins) {
  const float red = c.x, green = c.y, blue = c.z;
  float a, b, cc;
  int ia, ib, ic;
  if (red <= green && red <= blue) {
    if (green <= blue) { a = red; ia = 0; b = green - red; ib = 5; cc = blue - green; ic = 3; }
    else { a = red; ia = 0; b = blue - red; ib = 5; cc = green - blue; ic = 2; }
  } else if (green <= red && green <= blue) {
    if (red <= blue) { a = green; ia = 0; b = red - green; ib = 6; cc = blue - red; ic = 3; }
    else { a = green; ia = 0; b = blue - green; ib = 6; cc = red - blue; ic = 1; }
  } else {
    if (red <= green) { a = blue; ia = 0; b = red - blue; ib = 4; cc = green - red; ic = 2; }
    else { a = blue; ia = 0; b = green - blue; ib = 4; cc = red - green; ic = 1; }
  }
  float r[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int k = bin_of(bins, j);
    r[j] = a * c_smits[ia][k] + b * c_smits[ib][k] + cc * c_smits[ic][k];
  }
  return make_float4(r[0], r[1], r[2], r[3]);
}

}  // namespace ttgpu
