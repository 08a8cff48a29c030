// ORACLE — test infrastructure only.  A CPU (FP64) restatement of the algorithms
// on ToyTrace's pt2 path, used as the parity checker for the GPU library.
// Nothing in the product links or calls this code.
//
// Math foundations: vectors, frames, spectral tables, PCG32, sample transforms.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <array>
#include <algorithm>
#include <limits>
#include <vector>

namespace ttref {

constexpr double Pi = 3.14159265358979323846264338327950288;
constexpr double Epsilon = std::numeric_limits<double>::epsilon();      // vec3f.hxx:15
constexpr double Infinity = std::numeric_limits<double>::infinity();    // vec3f.hxx:20
constexpr double LargeNumber = std::numeric_limits<double>::max() / 16; // vec3f.hxx:22
constexpr double UnitSphereSurfaceArea = 4. * Pi;                        // vec3f.hxx:25
constexpr double UnitHalfSphereSurfaceArea = 2. * Pi;                    // vec3f.hxx:27
constexpr int NBINS = 36;                                                // spectral.hxx:274
constexpr double LambdaMin = 370., LambdaMax = 730.;                     // spectral.hxx:271-272

inline double Sqr(double x) { return x * x; }

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  constexpr V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 Cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double Length(V3 a) { return std::sqrt(Dot(a, a)); }
inline V3 Normalized(V3 a) { return a / Length(a); }
inline double ClampDot(V3 a, V3 b) { return std::max(0., Dot(a, b)); }   // vec3f.hxx:214-218
inline double AbsDot(V3 a, V3 b) { return std::abs(Dot(a, b)); }

// 3x3 matrix stored by columns.
struct M3 {
  V3 c0, c1, c2;
  V3 operator*(V3 v) const { return c0 * v.x + c1 * v.y + c2 * v.z; }
  V3 TMul(V3 v) const { return {Dot(c0, v), Dot(c1, v), Dot(c2, v)}; }  // transpose * v
};

// Duff et al. (2017) "Building an Orthonormal Basis, Revisited", listing 3
// (same construction as OrthogonalSystemZAligned, vec3f.hxx:242-258).
inline M3 OrthogonalSystemZAligned(V3 Z) {
  double sign = std::copysign(1.0, Z.z);
  const double a = -1.0 / (sign + Z.z);
  const double b = Z.x * Z.y * a;
  V3 X{1.0 + sign * Z.x * Z.x * a, sign * b, -sign * Z.x};
  V3 Y{b, sign + Z.y * Z.y * a, -Z.y};
  return {X, Y, Z};
}

// ---- spectral tables (spectral.hxx:287-289, spectral.cxx:13-45) --------
// Bin centres 375..725 nm; Smits (2000) basis spectra; spectrum->RGB rows.
extern const double kBinCenters[NBINS];
extern const double kSmits[7][NBINS];
extern const double kSpecToRGB[3][NBINS];
extern const double kSunSpectrum[NBINS];   // light.hxx:260

void RGBToSpectrum(const double rgb[3], double out[NBINS]);          // spectral.cxx:81-116
void SpectrumToRGB(const double spec[NBINS], double out_rgb[3]);     // spectral.cxx:118-128

// ---- PCG32 (O'Neill 2014, XSH-RR 64/32), the generator of src/sampler.cxx:371-393 --
struct Pcg32 {
  uint64_t state = 0x853c49e6748fea9bULL, inc = 0xda3e39cb94b95bdbULL;
  Pcg32() = default;
  Pcg32(uint64_t initstate, uint64_t initseq) { Seed(initstate, initseq); }
  void Seed(uint64_t initstate, uint64_t initseq) {
    state = 0u;
    inc = (initseq << 1u) | 1u;
    NextUInt();
    state += initstate;
    NextUInt();
  }
  uint32_t NextUInt() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
  }
  // Unbiased bounded integer by rejection.
  uint32_t NextUInt(uint32_t bound) {
    uint32_t threshold = (~bound + 1u) % bound;
    for (;;) {
      uint32_t r = NextUInt();
      if (r >= threshold) return r % bound;
    }
  }
  // 32 random mantissa bits in [0,1).
  double NextDouble() {
    uint64_t bits = ((uint64_t)NextUInt() << 20) | 0x3ff0000000000000ULL;
    double d;
    std::memcpy(&d, &bits, 8);
    return d - 1.0;
  }
};

struct Sampler {
  Pcg32 rng;
  double Uniform01() { return rng.NextDouble(); }
  void UniformUnitSquare(double r[2]) { r[0] = rng.NextDouble(); r[1] = rng.NextDouble(); }
  int UniformInt(int a, int b_inclusive) { return a + (int)rng.NextUInt((uint32_t)(b_inclusive + 1 - a)); }
};

// ---- sample transforms (src/sampler.cxx:6-125) --------------------------
inline V3 ToUniformDisc(const double r[2]) {
  double s = std::sqrt(r[1]), om = 2. * Pi * r[0];
  return {s * std::cos(om), s * std::sin(om), 0.};
}
inline V3 ToUniformSphere(const double r[2]) {
  double z = 1. - 2. * r[1];
  double s = std::sqrt(r[1] * (1. - r[1]));
  double om = 2. * Pi * r[0];
  return {2. * std::cos(om) * s, 2. * std::sin(om) * s, z};
}
inline V3 ToUniformHemisphere(const double r[2]) {
  V3 v = ToUniformSphere(r);
  v.z = v.z >= 0. ? v.z : -v.z;
  return v;
}
inline V3 ToUniformSphereSection(double cos_open, const double r[2]) {
  double z = 1. - r[1] * (1. - cos_open);
  double s = std::sqrt(1. - z * z);
  double om = 2. * Pi * r[0];
  return {std::cos(om) * s, std::sin(om) * s, z};
}
inline double UniformSphereSectionPdf(double cos_open) { return 1. / (2. * Pi * (1. - cos_open)); }
inline V3 ToUniformSphereSection(const double r[2], double phi0, double z0, double phi1, double z1) {
  const double z = z0 + r[1] * (z1 - z0);
  double s = std::sqrt(1. - z * z);
  s = std::isnan(s) ? 0. : s;
  double om = phi0 + (phi1 - phi0) * r[0];
  return {std::cos(om) * s, std::sin(om) * s, z};
}
inline V3 ToCosHemisphere(const double r[2]) {
  double rho = std::sqrt(1. - r[0]), z = std::sqrt(r[0]), om = 2. * Pi * r[1];
  return {std::cos(om) * rho, std::sin(om) * rho, z};
}
inline V3 ToTriangleBarycentricCoords(const double r[2]) {
  double s = std::sqrt(r[0]);
  return {1. - s, (1. - r[1]) * s, r[1] * s};
}

// Pdf value with the delta flag in the sign bit (src/sampler.hxx:552-590).
struct Pdf {
  double v = std::numeric_limits<double>::quiet_NaN();
  Pdf() = default;
  Pdf(double x) : v(x) {}
  static Pdf Delta(double x) { return Pdf(std::copysign(x, -1.)); }
  double val() const { return std::abs(v); }
  bool IsDelta() const { return std::signbit(v); }
};

// PdfConversion::AreaToSolidAngle (src/sampler.hxx:523-528)
inline double AreaToSolidAngle(double len, V3 dir, V3 n) { return Sqr(len) / std::abs(Dot(dir, n) + Epsilon); }

// TowerSampling: linear search from the rear (src/sampler.hxx:148-166).
template <int n_> inline int TowerSampling(const double *probs, double r) {
  int n = n_ - 1;
  while (r >= probs[n] && n > 0) { r -= probs[n]; --n; }
  return n;
}
// BisectionSearch / TowerSamplingBisection (src/sampler.hxx:170-207).
inline long BisectionSearch(const double *vals, long size, double r) {
  if (r < vals[0]) return 0;
  if (r >= vals[size - 1]) return size;
  long first = 0, last = size - 1;
  while (last > first + 1) {
    long c = (first + last) / 2;
    if (r < vals[c]) last = c; else first = c;
  }
  return last;
}
inline long TowerSamplingBisection(const double *cmf, long size, double r) {
  long u = BisectionSearch(cmf, size, r);
  return u >= size ? size - 1 : u;
}
inline void ComputeNormalizedCumSum(double *w, long n) {
  for (long i = 1; i < n; ++i) w[i] += w[i - 1];
  double inv = 1. / w[n - 1];
  for (long i = 0; i < n; ++i) w[i] *= inv;
}
inline double ProbabilityFromCmf(const double *cmf, long idx) { return cmf[idx] - (idx > 0 ? cmf[idx - 1] : 0.); }

// Projections (src/vec3f.hxx:273-344).  KartesianToSpherical treats component 2
// as the polar axis ("y is up" in the reference's own wording).
inline void KartesianToSpherical(V3 xyz, double &phi, double &theta) {
  const double x = xyz.x, z = xyz.y, y = xyz.z;
  const double ax = std::abs(x), az = std::abs(z);
  const double r = Length(xyz);
  theta = std::acos(y / r);
  if (ax > az) {
    phi = std::atan2(z, ax);
    phi = (x > 0) ? phi : Pi - phi;
  } else {
    phi = std::atan2(x, az);
    phi = (z > 0) ? Pi / 2. - phi : 3. / 2. * Pi + phi;
  }
}

// PBRT's gamma(n) error bound (src/util.hxx:124-129), float flavour.
inline float GammaF(int n) {
  constexpr float e = std::numeric_limits<float>::epsilon();
  return (n * e) / (1 - n * e);
}

// PBRT quadratic with error bounds (src/util.hxx:132-199).
template <class T> inline bool Quadratic(T a, T b, T c, T &t0, T &t1) {
  const T d = b * b - T(4) * a * c;
  if (d < T(0)) return false;
  const T sd = std::sqrt(d);
  const T q = b < 0 ? -b + sd : -b - sd;
  t0 = q / T(2) / a;
  t1 = T(2) * c / q;
  if (t0 > t1) std::swap(t0, t1);
  return true;
}
template <class T>
inline bool QuadraticErr(T A, T B, T C, T eA, T eB, T eC, T &t0, T &t1, T &err0, T &err1) {
  constexpr T eps = std::numeric_limits<T>::epsilon();
  const T D = B * B - T(4) * A * C;
  if (D < T(0)) return false;
  const T sD = std::sqrt(D);
  {
    const T xi = B < T(0) ? T(1) : T(-1);
    const T G = B - xi * sD;
    const T Ainv = T(1) / A, sDinv = T(1) / sD;
    const T E1 = std::abs(G * Ainv) + T(3) / T(4) * std::abs(sD * Ainv) + std::abs(C * sDinv) +
                 std::abs(B * B * Ainv * sDinv) / T(4);
    const T E2 = eA * std::abs((C * xi * Ainv * sDinv - G * Ainv * Ainv / T(2))) +
                 eB / T(2) * std::abs(Ainv * (B * xi * sDinv - T(1))) + eC * std::abs(sDinv);
    err0 = eps * E1 + E2;
  }
  {
    const T xi = B < T(0) ? T(1) : T(-1);
    const T G = B - xi * sD;
    const T GGsDinv = T(1) / (G * G * sD);
    const T E1 = std::abs(GGsDinv) * (T(4) * std::abs(C * G * sD) + T(3) * std::abs(C * D) +
                                      T(4) * std::abs(A * C * C) + std::abs(B * B * C));
    const T E2 = std::abs(GGsDinv) * (eA * T(4) * std::abs(C * C) + T(2) * eB * std::abs(C * (B * xi - sD)) +
                                      eC * std::abs((T(4) * A * C * xi - T(2) * G * sD)));
    err1 = eps * E1 + E2;
  }
  const T q = B < 0 ? -B + sD : -B - sD;
  t0 = q / T(2) / A;
  t1 = T(2) * C / q;
  if (t0 > t1) { std::swap(t0, t1); std::swap(err0, err1); }
  return true;
}

}  // namespace ttref
