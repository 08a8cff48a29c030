// Guided sampling on the GPU: cell lookup in the guiding kd-tree and the baked
// 8-lobe von Mises-Fisher mixture (pdf and sampling), plus the one-sample
// mixture with the BSDF of the guided integrator.
//
// vMF density k / (2 pi (1 - e^{-2k})) exp(k (mu.x - 1)) and inverse-CDF sampling in
// w = cos(theta) follow src/distribution_mixture_models.cxx:88-150 with exact exp/log;
// the BSDF/guide mixture follows SampleWithBinaryMixtureDensity
// (src/renderingalgorithms_pathtracing_guided.cxx:473-544).
#pragma once
#include "device_math.hpp"
#include "shading.hpp"

namespace ttgpu {

TT_HD int guide_lookup(const DeviceScene &sc, float3 p) {
  int h = sc.guide_root;
  while (h >= 0) {
    const GuideNode n = sc.guide_nodes[h];
    const float c = n.axis == 0 ? p.x : (n.axis == 1 ? p.y : p.z);
    h = c < n.split ? n.left : n.right;
  }
  return ~h;
}

// Mixtures: GuideCell (8 lobes, kd-tree cells) or the learned guide's per-path 16-lobe mixture;
// both hold lobe[N] = (mean, concentration), wn[N] = weight x normalisation, cdf[N].
template <class M>
TT_HD constexpr int lobes_of() {
  return (int)(sizeof(M::wn) / sizeof(float));
}

template <class M>
TT_HD float guide_pdf(const M &c, float3 x) {
  float p = 0.f;
#pragma unroll
  for (int k = 0; k < lobes_of<M>(); ++k) {
    const float4 l = c.lobe[k];
    p += c.wn[k] * expf(l.w * (l.x * x.x + l.y * x.y + l.z * x.z - 1.f));
  }
  return p;
}

template <class M>
TT_HD float3 guide_sample(const M &c, float u0, float u1, float u2) {
  int idx = lobes_of<M>() - 1;
#pragma unroll
  for (int k = lobes_of<M>() - 1; k >= 0; --k)
    if (u0 < c.cdf[k]) idx = k;
  const float4 l = c.lobe[idx];
  const float k = l.w;
  // w = 1 + log(u2 + (1 - u2) e^{-2k}) / k, written with log1p/expm1 for small k.
  const float w = fminf(1.f, fmaxf(-1.f, 1.f + log1pf((1.f - u2) * expm1f(-2.f * k)) / k));
  const float rho = sqrtf(fmaxf(0.f, 1.f - w * w));
  float sn, cs;
  fast_sincos(2.f * kPi * u1, &sn, &cs);
  const Frame f = make_frame(mk(l.x, l.y, l.z));
  return normalize(f.to_world(mk(rho * cs, rho * sn, w)));
}

// GuidingProbMixShaderAmount (src/shader.cxx:224-227, 937-941): probability of keeping the BSDF sample.
TT_HD float guide_prob_bsdf(const DevShader &s) {
  if (s.kind == TTGPU_SHADER_GLOSSY_DIELECTRIC) return s.alpha > 0.05f ? 0.1f : 0.5f;
  return 0.5f;
}

// One-sample mixture of BSDF and guide: the BSDF is sampled first and a delta lobe is returned as
// it is; otherwise the BSDF sample is kept with probability alpha or replaced by a guide sample.
// The returned pdf is (1 - alpha) p_guide + alpha p_bsdf.
template <class M>
TT_HD Scatter sample_guided(const DevShader &shd, const M &cell, const SurfHit &h, float3 wi, uint32_t bins, float rough,
                            float lambda0, Pcg &rng) {
  Scatter s = sample_bsdf(shd, h, wi, bins, rough, lambda0, rng);
  if (s.delta) return s;
  const float alpha = guide_prob_bsdf(shd);
  if (rng.uniform() < alpha) {
    s.pdf = (1.f - alpha) * guide_pdf(cell, s.dir) + alpha * s.pdf;
    return s;
  }
  const float u0 = rng.uniform(), u1 = rng.uniform(), u2 = rng.uniform();
  s.dir = guide_sample(cell, u0, u1, u2);
  float pb;
  s.value = eval_bsdf(shd, h, wi, bins, rough, lambda0, s.dir, &pb);
  s.pdf = (1.f - alpha) * guide_pdf(cell, s.dir) + alpha * pb;
  return s;
}

// Mixture pdf of a direction (EvaluateWithBinaryMixtureDensity), given the BSDF's own pdf.
template <class M>
TT_HD float guided_pdf(const DevShader &shd, const M &cell, float3 dir, float bsdf_pdf) {
  const float alpha = guide_prob_bsdf(shd);
  return (1.f - alpha) * guide_pdf(cell, dir) + alpha * bsdf_pdf;
}

TT_HD float power_heuristic(float a, float b) {
  const float a2 = a * a, b2 = b * b;
  return a2 > 0.f ? a2 / (a2 + b2) : 0.f;
}

}  // namespace ttgpu
