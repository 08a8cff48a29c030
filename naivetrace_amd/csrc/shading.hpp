// FP32 surface shaders and phase functions on the GPU.
//
// Same sampling procedures, densities and delta-lobe conventions as ToyTrace's
// shaders (src/shader.cxx:251-1081, src/shader_physics.hxx,
// src/normaldistributionfunction.hxx) and phase functions
// (src/phasefunctions.cxx); GGX metal is new (Heitz 2018 visible normals).
#pragma once
#include "device_math.hpp"

namespace ttgpu {

constexpr float kDblEps = 2.220446049250313e-16f;  // the reference's Epsilon, as a float

struct SurfHit {
  float3 pos;
  float3 ng;     // unoriented geometric normal
  float3 n;      // geometric normal oriented towards the incoming ray
  float3 ns;     // shading normal, oriented like n
  float offset;  // anti-self-intersection distance along ng
  int material;
  int prim;      // leaf-order primitive
};

struct Scatter {
  float3 dir;
  float4 value;
  float pdf;     // > 0
  bool delta;
};

TT_HD float3 offset_origin(const SurfHit &h, float3 dir) {
  return h.pos + h.ng * (dot(dir, h.ng) > 0.f ? h.offset : -h.offset);
}

TT_HD float3 reflected(float3 wi, float3 n) { return n * (2.f * dot(wi, n)) - wi; }
TT_HD bool refracted(float3 wi, float3 m, float eta, float3 &out) {
  const float c = dot(m, wi);
  const float t2 = 1.f - eta * eta * (1.f - c * c);
  if (t2 < 0.f) return false;
  const float s = c >= 0.f ? 1.f : -1.f;
  out = m * (eta * c - s * sqrtf(t2)) - wi * eta;
  return true;
}
TT_HD float3 half_vector(float3 wi, float3 wo) {
  float3 r = wo + wi;
  float n = length(r);
  if (n > 0.f) return r * (1.f / n);
  return make_frame(wi).x;
}
TT_HD bool half_vector_refracted(float3 wi, float3 wo, float eta, float3 &out) {
  float3 r = -(wi * eta + wo);
  float n = length(r);
  r = n > 0.f ? r * (1.f / n) : wi;
  if (dot(r, wi) * dot(r, wo) < 0.f) { out = r; return true; }
  return false;
}
TT_HD float schlick(float ks, float c) {
  float m = 1.f - c;
  float m2 = m * m;
  return ks + (1.f - ks) * (m2 * m2 * m);
}
TT_HD float fresnel(float cs_i, float cs_t, float eta) {
  float rs = (eta * cs_i - cs_t) / (eta * cs_i + cs_t);
  float rp = (eta * cs_t - cs_i) / (eta * cs_t + cs_i);
  return 0.5f * (rs * rs + rp * rp);
}
TT_HD float fresnel_walter(float cos_n_wi, float eta) {
  float c = fabsf(cos_n_wi);
  float tmp = 1.f / (eta * eta) - 1.f + c * c;
  if (tmp < 0.f) return 1.f;
  float g = sqrtf(tmp);
  if (g == 0.f && c == 0.f) return 1.f;
  float nom = c * (g + c) - 1.f, den = c * (g - c) + 1.f;
  return 0.5f * sqr((g - c) / (g + c)) * (1.f + sqr(nom / den));
}

// Beckmann NDF (normaldistributionfunction.hxx:8-43).
TT_HD float beckmann_d(float alpha, float c) {
  if (c <= 0.f) return 0.f;
  float c2 = c * c, a2 = alpha * alpha;
  return fast_exp((c2 - 1.f) / (c2 * a2)) / (a2 * c2 * c2 * kPi);
}
TT_HD float3 beckmann_sample(float alpha, float r0, float r1) {
  float t1 = -alpha * alpha * fast_log(r0);
  float t = 1.f / (t1 + 1.f);
  float z = sqrtf(t), rho = sqrtf(fmaxf(0.f, 1.f - t));
  float sn, cs;
  fast_sincos(2.f * kPi * r1, &sn, &cs);
  return mk(cs * rho, sn * rho, z);
}
TT_HD float g1_vcavity(float cvm, float cvn, float cnm) {
  if (cvm * cvn < 0.f) return 0.f;
  return fminf(1.f, 2.f * fabsf(cnm * cvn) / (fabsf(cvm) + kDblEps));
}
constexpr float kWiBelowHorizonFix = 1.e-6f;

// GGX NDF, Smith Lambda, visible-normal sampling (Heitz 2018).
TT_HD float ggx_d(float a, float3 h) {
  if (h.z <= 0.f) return 0.f;
  float a2 = a * a;
  float t = (h.x * h.x + h.y * h.y) / a2 + h.z * h.z;
  return 1.f / (kPi * a2 * t * t);
}
TT_HD float ggx_lambda(float a, float3 w) {
  float z2 = w.z * w.z;
  if (z2 <= 0.f) return 1e30f;
  return 0.5f * (-1.f + sqrtf(1.f + a * a * (1.f - z2) / z2));
}
TT_HD float3 ggx_sample_visible(float a, float3 wi, float r0, float r1) {
  float3 vh = normalize(mk(a * wi.x, a * wi.y, wi.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  float3 T1 = lensq > 0.f ? mk(-vh.y, vh.x, 0.f) * (1.f / sqrtf(lensq)) : mk(1.f, 0.f, 0.f);
  float3 T2 = cross(vh, T1);
  float rr = sqrtf(r0);
  float sn, cs;
  fast_sincos(2.f * kPi * r1, &sn, &cs);
  float t1 = rr * cs, t2 = rr * sn;
  float s = 0.5f * (1.f + vh.z);
  t2 = (1.f - s) * sqrtf(fmaxf(0.f, 1.f - t1 * t1)) + s * t2;
  float3 nh = T1 * t1 + T2 * t2 + vh * sqrtf(fmaxf(0.f, 1.f - t1 * t1 - t2 * t2));
  return normalize(mk(a * nh.x, a * nh.y, fmaxf(0.f, nh.z)));
}

// ---- per-shader sampling / evaluation ---------------------------------------
TT_HD float my_roughness(const DevShader &s) {
  return (s.kind == TTGPU_SHADER_BECKMANN_METAL || s.kind == TTGPU_SHADER_GGX_METAL || s.kind == TTGPU_SHADER_GLOSSY_DIELECTRIC)
             ? s.alpha : 0.f;
}

TT_HD float4 diffuse_eval(const DevShader &s, const SurfHit &h, uint32_t bins, float3 wo, float *pdf) {
  *pdf = fmaxf(0.f, dot(h.ns, wo)) * kInvPi;
  return dot(h.n, wo) > 0.f ? kInvPi * take(s.reflectance, bins) : f4(0.f);
}
TT_HD Scatter diffuse_sample(const DevShader &s, const SurfHit &h, uint32_t bins, Pcg &rng) {
  float r0 = rng.uniform(), r1 = rng.uniform();
  float3 v = to_cos_hemisphere(r0, r1);
  Scatter out;
  out.dir = make_frame(h.ns).to_world(v);
  out.pdf = v.z * kInvPi;
  out.delta = false;
  out.value = dot(h.n, out.dir) > 0.f ? kInvPi * take(s.reflectance, bins) : f4(0.f);
  return out;
}

// Microfacet metal: Beckmann with V-cavity VNDF (shader.cxx:67-123, 634-707), or GGX.
TT_HD float4 metal_core(const DevShader &s, float alpha, float3 ng_l, float3 wi, float3 wh, float3 wo,
                                             uint32_t bins, float *pdf) {
  float4 col = take(s.reflectance, bins);
  if (s.kind == TTGPU_SHADER_BECKMANN_METAL) {
    const float ns_out = wo.z, ns_in = wi.z, ns_wh = wh.z;
    const float wh_out = dot(wo, wh), wh_in = dot(wi, wh);
    float D = beckmann_d(alpha, fabsf(ns_wh));
    float G = fminf(g1_vcavity(wh_in, ns_in, ns_wh), g1_vcavity(wh_out, ns_out, ns_wh));
    float3 whf = wh.z < 0.f ? -wh : wh;
    float3 whp = mk(-whf.x, -whf.y, whf.z);
    float a = clampdot(wi, whp), b = clampdot(whf, wi);
    float prob = (2.f * b + kWiBelowHorizonFix) / (b + a + kWiBelowHorizonFix);
    float sp = D * fabsf(whf.z) * prob;
    *pdf = sp * 0.25f / (fabsf(dot(whf, wi)) + kDblEps);
    if (dot(ng_l, wo) <= 0.f) return f4(0.f);
    float f = G * D * 0.25f / (fabsf(ns_in * ns_out) + kDblEps);
    float fr = fabsf(wh_in);
    return make_float4(f * schlick(col.x, fr), f * schlick(col.y, fr), f * schlick(col.z, fr), f * schlick(col.w, fr));
  }
  // GGX
  float3 h = wh.z < 0.f ? -wh : wh;
  float lw = ggx_lambda(alpha, wi);
  // Density of the visible normal for the direction the sampler used: wi, or wi lifted just above
  // the shading hemisphere when it lies below (then the value is zero and the path ends).
  const float3 wv = wi.z > 0.f ? wi : normalize(mk(wi.x, wi.y, 1e-7f));
  const float lv = wi.z > 0.f ? lw : ggx_lambda(alpha, wv);
  float pv = (1.f / (1.f + lv)) * clampdot(wv, h) * ggx_d(alpha, h) / wv.z;
  *pdf = pv * 0.25f / (fabsf(dot(wi, h)) + kDblEps);
  if (dot(ng_l, wo) <= 0.f || wi.z <= 0.f || wo.z <= 0.f) return f4(0.f);
  float D = ggx_d(alpha, h);
  float G = 1.f / (1.f + lw + ggx_lambda(alpha, wo));
  float f = D * G * 0.25f / (fabsf(wi.z * wo.z) + kDblEps);
  float fr = fabsf(dot(h, wi));
  return make_float4(f * schlick(col.x, fr), f * schlick(col.y, fr), f * schlick(col.z, fr), f * schlick(col.w, fr));
}
TT_HD float4 metal_eval(const DevShader &s, const SurfHit &h, float3 wi_w, uint32_t bins, float rough, float3 wo_w,
                                             float *pdf) {
  Frame f = make_frame(h.ns);
  float3 wi = f.to_local(wi_w), wo = f.to_local(wo_w);
  float alpha = fmaxf(s.alpha, rough);
  float3 wh = s.kind == TTGPU_SHADER_BECKMANN_METAL ? normalize(wi + wo) : half_vector(wi, wo);
  return metal_core(s, alpha, f.to_local(h.n), wi, wh, wo, bins, pdf);
}
TT_HD Scatter metal_sample(const DevShader &s, const SurfHit &h, float3 wi_w, uint32_t bins, float rough, Pcg &rng) {
  Frame f = make_frame(h.ns);
  float3 wi = f.to_local(wi_w);
  float alpha = fmaxf(s.alpha, rough);
  float r0 = rng.uniform(), r1 = rng.uniform();
  float3 wh;
  if (s.kind == TTGPU_SHADER_BECKMANN_METAL) {
    wh = beckmann_sample(alpha, r0, r1);
    float3 whp = mk(-wh.x, -wh.y, wh.z);
    float prob = clampdot(wi, whp) / (clampdot(wi, wh) + clampdot(wi, whp) + kWiBelowHorizonFix);
    if (rng.uniform() < prob) wh = whp;
  } else {
    float3 wip = wi.z >= 0.f ? wi : mk(wi.x, wi.y, 1e-7f);
    wh = ggx_sample_visible(alpha, normalize(wip), r0, r1);
  }
  float3 wo = reflected(wi, wh);
  Scatter out;
  float pdf;
  float3 whe = s.kind == TTGPU_SHADER_BECKMANN_METAL ? wh : half_vector(wi, wo);
  out.value = metal_core(s, alpha, f.to_local(h.n), wi, whe, wo, bins, &pdf);
  out.dir = f.to_world(wo);
  out.pdf = pdf;
  out.delta = false;
  return out;
}

// Specular dielectric, regular and mollified (shader.cxx:400-604).
TT_HD float eta_i_over_t(const DevShader &s, const SurfHit &h, float3 wi, float lambda0) {
  bool entering = dot(h.ng, wi) > 0.f;
  float ior = s.ior_lambda_coeff != 0.f ? s.ior_ratio + s.ior_lambda_coeff * lambda0 : s.ior_ratio;
  return entering ? 1.f / ior : ior;
}
// importance: transport of importance (light subpaths), without the 1 / eta^2 radiance scaling
// on refraction (shader.cxx:425, 498, 566).
TT_HD Scatter dielectric_sample(const DevShader &s, const SurfHit &h, float3 wi, float rough, float lambda0, Pcg &rng,
                                bool importance = false) {
  float abs_ni = fabsf(dot(h.ns, wi));
  float eta = eta_i_over_t(s, h, wi, lambda0);
  float fr = 1.f;
  float3 wt;
  bool has_wt = refracted(wi, h.ns, eta, wt);
  if (has_wt) fr = fresnel(abs_ni, fabsf(dot(wt, h.ns)), eta);
  const float prob_refl = has_wt ? fmaxf(0.1f, fminf(0.9f, fr)) : 1.f;
  bool do_refl = has_wt ? rng.uniform() < prob_refl : true;
  Scatter out;
  float val;
  float3 centre = do_refl ? reflected(wi, h.ns) : wt;
  if (rough > 0.f) {
    const float open = 1.f - rough;
    const float sp = sphere_section_pdf(open);
    float r0 = rng.uniform(), r1 = rng.uniform();
    out.dir = make_frame(centre).to_world(to_sphere_section(open, r0, r1));
    out.pdf = (do_refl ? prob_refl : 1.f - prob_refl) * sp;
    out.delta = false;
    val = (do_refl ? fr : 1.f - fr) * sp;
  } else {
    out.dir = centre;
    out.pdf = do_refl ? prob_refl : 1.f - prob_refl;
    out.delta = true;
    val = do_refl ? fr : 1.f - fr;
  }
  val /= fabsf(dot(out.dir, h.ns));
  if (dot(out.dir, h.n) < 0.f && !importance) val *= eta * eta;
  out.value = f4(val);
  if (s.ior_lambda_coeff != 0.f) out.value.y = out.value.z = out.value.w = 0.f;
  return out;
}
TT_HD float4 dielectric_eval(const DevShader &s, const SurfHit &h, float3 wi, float rough, float lambda0, float3 wo,
                                                  float *pdf, bool importance = false) {
  if (!(rough > 0.f)) { *pdf = 0.f; return f4(0.f); }
  const float open = 1.f - rough;
  float eta = eta_i_over_t(s, h, wi, lambda0);
  float abs_ni = fabsf(dot(h.ns, wi));
  float fr = 1.f;
  float3 wt;
  bool has_wt = refracted(wi, h.ns, eta, wt);
  if (has_wt) fr = fresnel(abs_ni, fabsf(dot(wt, h.ns)), eta);
  const float prob_refl = has_wt ? fmaxf(0.1f, fminf(0.9f, fr)) : 1.f;
  const float sp = sphere_section_pdf(open);
  const bool maybe_refl = dot(reflected(wi, h.ns), wo) > open;
  float tp = maybe_refl ? sp * prob_refl : 0.f;
  float tv = maybe_refl ? prob_refl * fr : 0.f;
  if (has_wt && dot(wt, wo) > open) {
    tv += (1.f - fr) * sp;
    tp += (1.f - prob_refl) * sp;
  }
  if (dot(wo, h.n) < 0.f && !importance) tv *= eta * eta;
  tv /= fabsf(dot(wo, h.ns));
  *pdf = tp;
  return f4(tv);
}

// Glossy transmissive dielectric (shader.cxx:125-195, 710-787).
TT_HD float glossy_diel_core(float3 wi, float3 ng, float alpha, float balpha, float eta, float3 wo, float *pdf) {
  const float ns_out = wo.z, ns_in = wi.z;
  float pwot = 0.f;
  float3 wht;
  if (half_vector_refracted(wi, wo, eta, wht)) {
    float frt = fresnel_walter(fabsf(dot(wht, wi)), eta);
    float nt = beckmann_d(balpha, fabsf(wht.z)) * fabsf(wht.z);
    float d = eta * dot(wht, wi) + dot(wht, wo);
    float pt = d == 0.f ? nt : nt * fabsf(dot(wht, wo)) / (d * d);
    pwot = pt * (1.f - frt);
  }
  float3 whr = half_vector(wi, wo);
  float frr = fresnel_walter(fabsf(dot(whr, wi)), eta);
  float nr = beckmann_d(balpha, fabsf(whr.z)) * fabsf(whr.z);
  *pdf = nr * 0.25f / (fabsf(dot(whr, wi)) + kDblEps) * frr + pwot;
  if (dot(ng, wo) >= 0.f) {
    float fr = fresnel_walter(fabsf(dot(whr, wi)), eta);
    float D = beckmann_d(alpha, fabsf(whr.z));
    float G = fminf(g1_vcavity(dot(whr, wi), ns_in, whr.z), g1_vcavity(dot(wo, whr), ns_out, whr.z));
    return fr * G * D * 0.25f / (fabsf(ns_in * ns_out) + kDblEps);
  }
  float3 w;
  if (!half_vector_refracted(wi, wo, eta, w)) return 0.f;
  float3 wt = w.z < 0.f ? -w : w;
  const float whi = dot(wi, wt), who = dot(wo, wt);
  float fr = fresnel_walter(fabsf(whi), eta);
  float D = beckmann_d(alpha, fabsf(wt.z));
  float den = eta * whi + who;
  float J = den == 0.f ? 1.f : fabsf(who) / (den * den);
  float G = fmaxf(g1_vcavity(whi, ns_in, wt.z) + g1_vcavity(who, ns_out, wt.z) - 1.f, 0.f);
  return (1.f - fr) * fabsf(whi) * G * D * J / (fabsf(ns_in * ns_out) + kDblEps);
}
TT_HD float4 glossy_diel_eval(const DevShader &s, const SurfHit &h, float3 wi_w, float rough, float3 wo_w, float *pdf) {
  Frame f = make_frame(h.ns);
  float3 wi = f.to_local(wi_w), wo = f.to_local(wo_w);
  float eta = dot(h.ng, wi_w) < 0.f ? s.ior_ratio : 1.f / s.ior_ratio;
  float alpha = fmaxf(rough, s.alpha);
  return f4(glossy_diel_core(wi, f.to_local(h.n), alpha, alpha * (1.2f - 0.2f * fabsf(wi.z)), eta, wo, pdf));
}
TT_HD Scatter glossy_diel_sample(const DevShader &s, const SurfHit &h, float3 wi_w, float rough, Pcg &rng) {
  Frame f = make_frame(h.ns);
  float3 wi = f.to_local(wi_w);
  float eta = dot(h.ng, wi_w) < 0.f ? s.ior_ratio : 1.f / s.ior_ratio;
  float alpha = fmaxf(rough, s.alpha), balpha = alpha * (1.2f - 0.2f * fabsf(wi.z));
  float r0 = rng.uniform(), r1 = rng.uniform(), r2 = rng.uniform();
  float3 wh = beckmann_sample(balpha, r0, r1);
  float fr = fresnel_walter(fabsf(dot(wh, wi)), eta);
  float3 wt;
  if (!refracted(wi, wh, eta, wt)) fr = 1.f;
  float3 wo = r2 < fr ? reflected(wi, wh) : wt;
  Scatter out;
  float pdf;
  out.value = f4(glossy_diel_core(wi, f.to_local(h.n), alpha, balpha, eta, wo, &pdf));
  out.dir = f.to_world(wo);
  out.pdf = pdf;
  out.delta = false;
  return out;
}

TT_HD float average_projected_schlick(float ks) { return ks + (1.f - ks) * 2.f / 42.f; }

TT_HD Scatter sample_bsdf(const DevShader &s, const SurfHit &h, float3 wi, uint32_t bins, float rough, float lambda0, Pcg &rng,
                          bool importance = false) {
  switch (s.kind) {
    case TTGPU_SHADER_DIFFUSE: return diffuse_sample(s, h, bins, rng);
    case TTGPU_SHADER_MIRROR: {
      Scatter o;
      o.dir = reflected(wi, h.ns);
      o.pdf = 1.f;
      o.delta = true;
      o.value = dot(h.n, o.dir) < 0.f ? f4(0.f) : (1.f / dot(h.ns, o.dir)) * take(s.reflectance, bins);
      return o;
    }
    case TTGPU_SHADER_DIELECTRIC: return dielectric_sample(s, h, wi, rough, lambda0, rng, importance);
    case TTGPU_SHADER_BECKMANN_METAL:
    case TTGPU_SHADER_GGX_METAL: return metal_sample(s, h, wi, bins, rough, rng);
    case TTGPU_SHADER_INVISIBLE: {
      Scatter o;
      o.dir = -wi;
      o.pdf = 1.f;
      o.delta = true;
      o.value = f4(-1.f / fabsf(dot(wi, h.ns)));
      return o;
    }
    case TTGPU_SHADER_GLOSSY_DIELECTRIC: return glossy_diel_sample(s, h, wi, rough, rng);
    case TTGPU_SHADER_DENSE_DIELECTRIC: {
      float ci = fmaxf(0.f, dot(h.ns, wi));
      float rf = schlick(s.specular, ci);
      if (rng.uniform() < rf) {
        Scatter o;
        o.dir = reflected(wi, h.ns);
        o.pdf = rf;
        o.delta = true;
        o.value = dot(h.n, o.dir) >= 0.f ? f4(rf / (ci + kDblEps)) : f4(0.f);
        return o;
      }
      Scatter o = diffuse_sample(s, h, bins, rng);
      float ce = fmaxf(0.f, dot(h.ns, o.dir));
      float k = (1.f - rf) * (1.f - schlick(s.specular, ce)) / (1.f - average_projected_schlick(s.specular));
      o.value = k * o.value;
      o.pdf *= (1.f - rf);
      return o;
    }
  }
  Scatter o;
  o.dir = wi; o.pdf = 1.f; o.delta = true; o.value = f4(0.f);
  return o;
}

TT_HD float4 eval_bsdf(const DevShader &s, const SurfHit &h, float3 wi, uint32_t bins, float rough, float lambda0, float3 wo, float *pdf,
                       bool importance = false) {
  switch (s.kind) {
    case TTGPU_SHADER_DIFFUSE: return diffuse_eval(s, h, bins, wo, pdf);
    case TTGPU_SHADER_DIELECTRIC: return dielectric_eval(s, h, wi, rough, lambda0, wo, pdf, importance);
    case TTGPU_SHADER_BECKMANN_METAL:
    case TTGPU_SHADER_GGX_METAL: return metal_eval(s, h, wi, bins, rough, wo, pdf);
    case TTGPU_SHADER_GLOSSY_DIELECTRIC: return glossy_diel_eval(s, h, wi, rough, wo, pdf);
    case TTGPU_SHADER_DENSE_DIELECTRIC: {
      if (dot(h.n, wo) > 0.f) {
        float ce = fmaxf(0.f, dot(h.ns, wo)), ci = fmaxf(0.f, dot(h.ns, wi));
        float rf = schlick(s.specular, ci);
        float4 v = diffuse_eval(s, h, bins, wo, pdf);
        float k = (1.f - rf) * (1.f - schlick(s.specular, ce)) / (1.f - average_projected_schlick(s.specular));
        *pdf *= (1.f - rf);
        return k * v;
      }
      *pdf = 0.f;
      return f4(0.f);
    }
    default: *pdf = 0.f; return f4(0.f);
  }
}

// ---- phase functions (phasefunctions.cxx) -------------------------------------
TT_HD float3 mu_to_direction(float mu, float r2) {
  float sn, cs;
  fast_sincos(2.f * kPi * r2, &sn, &cs);
  float rho = sqrtf(fmaxf(0.f, 1.f - mu * mu));
  return mk(sn * rho, cs * rho, mu);
}
TT_HD float hg_sample_mu(float r, float g) {
  float s = 2.f * r - 1.f, g2 = g * g;
  if (fabsf(g) > 1.0e-3f) {
    float t1 = (1.f - g2) / (1.f + g * s);
    return (1.f + g2 - t1 * t1) / (2.f * g);
  }
  float s2 = s * s;
  return s + 1.5f * g * (1.f - s2) - 2.f * g2 * s * (1.f - s2);
}
TT_HD float hg_value(float mu, float g) {
  float g2 = g * g, term = 1.f + g2 - 2.f * g * mu;
  return (1.f / (4.f * kPi)) * (1.f - g2) / (sqrtf(term) * term);
}
TT_HD float rayleigh_sample_mu(float r) {
  float a = cbrtf(-2.f + 4.f * r + sqrtf(5.f - 16.f * r + 16.f * r * r));
  return a - 1.f / a;
}
TT_HD float rayleigh_value(float mu) { return (3.f / (16.f * kPi)) * (1.f + mu * mu); }

TT_HD void phase_sample(int kind, float g, float3 wi, Pcg &rng, float3 &dir, float &val) {
  float r0 = rng.uniform(), r1 = rng.uniform();
  if (kind == TTGPU_PHASE_UNIFORM) {
    dir = to_uniform_sphere(r0, r1);
    val = 1.f / (4.f * kPi);
    return;
  }
  Frame f = make_frame(-wi);
  float mu = kind == TTGPU_PHASE_RAYLEIGH ? rayleigh_sample_mu(r0) : hg_sample_mu(r0, g);
  mu = fminf(1.f, fmaxf(-1.f, mu));
  dir = f.to_world(mu_to_direction(mu, r1));
  val = kind == TTGPU_PHASE_RAYLEIGH ? rayleigh_value(mu) : hg_value(mu, g);
}
TT_HD float phase_eval(int kind, float g, float3 wi, float3 wo) {
  if (kind == TTGPU_PHASE_UNIFORM) return 1.f / (4.f * kPi);
  float mu = dot(-wi, wo);
  return kind == TTGPU_PHASE_RAYLEIGH ? rayleigh_value(mu) : hg_value(mu, g);
}

}  // namespace ttgpu
