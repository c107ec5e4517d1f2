// Pd-H EAM functions (potential.cpp:16-205) with analytic derivatives by
// forward (value, slope) accumulation through the product polynomials; the
// reference's central-FD derivatives are formed by the callers.  Templated on
// the floating type: double on the device (the FP64 kernels), long double on
// the host (building the FP32 tables of the mixed path).  Constants are those
// of potential.cpp, in its order, as the doubles the reference parses.  The
// radial polynomials' scale t = r / 5 is applied as r * 0.2 (within an ulp
// of the quotient): an FP64 divide is a ~10-instruction sequence, and the
// Pd-Pd terms of one quadrature point would take four of them.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace a2c {
namespace eam {

template <typename T>
struct VDT { T v, d; };
using VD = VDT<double>;
#define A2C_HD __host__ __device__ __forceinline__

template <typename T>
struct same { using type = T; };  // keeps an argument out of template deduction
template <typename T>
A2C_HD VDT<T> vmul(VDT<T> a, typename same<VDT<T>>::type b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <typename T>
A2C_HD VDT<T> quadf(T t, typename same<T>::type b, typename same<T>::type c) {
  return {t * (t - b) + c, T(2) * t - b};
}

// ---- embedding F_Pd(rho), potential.cpp:16-38 --------------------------------
template <typename T = double>
A2C_HD T F_Pd(T rho) {
  T r = rho / 50.0;
  T y = 295878.9003038662 * (r - 0.20581955357385892) * (r - 0.081228755904399) * r *
             (r + 0.05298811034615951) * (r * (r - 2.4242616904962846) + 1.4791899886249564) *
             (r * (r - 2.1376274623740064) + 1.2169215689822592) *
             (r * (r - 1.6486007989726832) + 0.8159825255339774) *
             (r * (r - 1.0749204110338482) + 0.42007491336688396) *
             (r * (r - 0.5128056047933808) + 0.12468685331167456);
  return y - 50.0 * 0.251546 * r;
}
template <typename T = double>
A2C_HD VDT<T> F_Pd_vd(T rho) {
  T r = rho / 50.0;
  VDT<T> p = {295878.9003038662 * (r - 0.20581955357385892), 295878.9003038662};
  p = vmul(p, {r - 0.081228755904399, 1.0});
  p = vmul(p, {r, 1.0});
  p = vmul(p, {r + 0.05298811034615951, 1.0});
  p = vmul(p, quadf(r, 2.4242616904962846, 1.4791899886249564));
  p = vmul(p, quadf(r, 2.1376274623740064, 1.2169215689822592));
  p = vmul(p, quadf(r, 1.6486007989726832, 0.8159825255339774));
  p = vmul(p, quadf(r, 1.0749204110338482, 0.42007491336688396));
  p = vmul(p, quadf(r, 0.5128056047933808, 0.12468685331167456));
  return {p.v - 50.0 * 0.251546 * r, p.d / 50.0 - 0.251546};
}

// ---- embedding F_H(rho), potential.cpp:45-57 ---------------------------------
template <typename T = double>
A2C_HD T F_H(T rho) {
  const T c = 0.000197047, d = 1.18860, a = 9.99780, b = 60.0155, e1 = 0.0540638;
  T u = rho + e1;
  T y = -c * (1.0 / (2.0 + d) * std::pow(u, 2.0 + d) - (a + b) / (1.0 + d) * std::pow(u, 1.0 + d) +
                   a * b / d * std::pow(u, d));
  return y + 0.0296604 * rho;
}
template <typename T = double>
A2C_HD VDT<T> F_H_vd(T rho) {
  const T c = 0.000197047, d = 1.18860, a = 9.99780, b = 60.0155, e1 = 0.0540638;
  T u = rho + e1;
  T pd = std::pow(u, d);            // u^d
  T p1 = pd * u, p2 = p1 * u;  // u^(1+d), u^(2+d)
  T v = -c * (1.0 / (2.0 + d) * p2 - (a + b) / (1.0 + d) * p1 + a * b / d * pd) + 0.0296604 * rho;
  T dv = -c * (p1 - (a + b) * pd + a * b * pd / u) + 0.0296604;
  return {v, dv};
}

// ---- density f_Pd(r), potential.cpp:64-98 -------------------------------------
template <typename T = double>
A2C_HD T f_Pd(T r) {
  T t = r * 0.2, rr = t * 5.0;
  if (rr <= 0.0) return 0.0;
  if (rr > 5.0) return 0.001510569378297 * std::exp(-0.077976566126292 / 0.001510569378297 * (rr - 5.0));
  T ra = t * (t * (t - 2.7267629107325706) + 1.8716766113599643) *
              (t * (t - 2.50290548851635) + 1.668549182690922) *
              (t * (t - 2.0924467509943674) + 1.3150372774478005) *
              (t * (t - 1.564328475106985) + 0.8987511149780485) *
              (t * (t - 1.009780903403673) + 0.5124363774128722) *
              (t * (t - 0.5304054524800665) + 0.2169886022464641) *
              (t * (t - 0.1356566408715063) + 0.035852347523891395);
  return -0.02972698211669922 +
         t * (0.6676807403564453 + t * (-255.8965835571289 + t * (14673.409149169922 + -2.597301181336601e7 * ra)));
}
template <typename T = double>
A2C_HD VDT<T> f_Pd_vd(T r) {
  T t = r * 0.2, rr = t * 5.0;
  if (rr <= 0.0) return {0.0, 0.0};
  if (rr > 5.0) {
    const T k = -0.077976566126292 / 0.001510569378297;
    T y = 0.001510569378297 * std::exp(k * (rr - 5.0));
    return {y, k * y};
  }
  VDT<T> p = {t, 1.0};
  p = vmul(p, quadf(t, 2.7267629107325706, 1.8716766113599643));
  p = vmul(p, quadf(t, 2.50290548851635, 1.668549182690922));
  p = vmul(p, quadf(t, 2.0924467509943674, 1.3150372774478005));
  p = vmul(p, quadf(t, 1.564328475106985, 0.8987511149780485));
  p = vmul(p, quadf(t, 1.009780903403673, 0.5124363774128722));
  p = vmul(p, quadf(t, 0.5304054524800665, 0.2169886022464641));
  p = vmul(p, quadf(t, 0.1356566408715063, 0.035852347523891395));
  VDT<T> h = {14673.409149169922 + -2.597301181336601e7 * p.v, -2.597301181336601e7 * p.d};
  h = {-255.8965835571289 + t * h.v, h.v + t * h.d};
  h = {0.6676807403564453 + t * h.v, h.v + t * h.d};
  h = {-0.02972698211669922 + t * h.v, h.v + t * h.d};
  return {h.v, h.d * 0.2};
}

// ---- density f_H(r), potential.cpp:105-119 ------------------------------------
template <typename T = double>
A2C_HD T f_H(T r) {
  if (r <= 0.0) return 0.0;
  if (r > 5.35) return 0.009986990526578 * std::exp(-0.013075667086408 / 0.009986990526578 * (r - 5.35));
  return 11.0025 * std::exp(-1.30927 * r);
}
template <typename T = double>
A2C_HD VDT<T> f_H_vd(T r) {
  if (r <= 0.0) return {0.0, 0.0};
  if (r > 5.35) {
    const T k = -0.013075667086408 / 0.009986990526578;
    T y = 0.009986990526578 * std::exp(k * (r - 5.35));
    return {y, k * y};
  }
  T y = 11.0025 * std::exp(-1.30927 * r);
  return {y, -1.30927 * y};
}

// ---- pair phi_Pd(r), potential.cpp:126-155 ------------------------------------
template <typename T = double>
A2C_HD T phi_Pd(T r) {
  T t = r * 0.2, rr = t * 5.0;
  if (rr <= 0.0) return 0.0;
  if (rr > 5.05) return -0.016958780164980 * std::exp(0.089019679316328 / -0.016958780164980 * (rr - 5.05));
  T y = -79415.24035137112 * (t - 1.0699996145674568) * (t - 1.06015072612581) *
             (t - 0.42433991011376526) * (t + 0.06169160085238687) *
             (t * (t - 2.0586473420376348) + 1.0683922574015199) *
             (t * (t - 1.6696359816422877) + 0.7337878627470482) *
             (t * (t - 1.1690370066230809) + 0.3909805777737639) *
             (t * (t - 0.2635598721249787) + 0.033551116514910245);
  return y + 2 * 0.251546 * f_Pd(rr);
}
// phi_Pd with derivative; fpd is f_Pd(r) with derivative (shared with the
// density/embedding terms of the same point).
template <typename T = double>
A2C_HD VDT<T> phi_Pd_vd(T r, VDT<T> fpd) {
  T t = r * 0.2, rr = t * 5.0;
  if (rr <= 0.0) return {0.0, 0.0};
  if (rr > 5.05) {
    const T k = 0.089019679316328 / -0.016958780164980;
    T y = -0.016958780164980 * std::exp(k * (rr - 5.05));
    return {y, k * y};
  }
  VDT<T> p = {-79415.24035137112 * (t - 1.0699996145674568), -79415.24035137112};
  p = vmul(p, {t - 1.06015072612581, 1.0});
  p = vmul(p, {t - 0.42433991011376526, 1.0});
  p = vmul(p, {t + 0.06169160085238687, 1.0});
  p = vmul(p, quadf(t, 2.0586473420376348, 1.0683922574015199));
  p = vmul(p, quadf(t, 1.6696359816422877, 0.7337878627470482));
  p = vmul(p, quadf(t, 1.1690370066230809, 0.3909805777737639));
  p = vmul(p, quadf(t, 0.2635598721249787, 0.033551116514910245));
  return {p.v + 2 * 0.251546 * fpd.v, p.d * 0.2 + 2 * 0.251546 * fpd.d};
}

// ---- pair phi_PdH / phi_H, potential.cpp:162-200 ------------------------------
template <int HH, typename T = double>
A2C_HD T phi_morse(T r) {
  const T fr = HH ? -0.003702722457147 : -3.353001113951400e-04;
  const T fd = HH ? 0.005465945471933 : 7.147062196813336e-04;
  const T D = HH ? 0.0661496 : 0.2494540, al = HH ? 3.67263 : 4.82613;
  const T be = HH ? 1.47797 : 2.13158, r0 = HH ? 2.51980 : 1.50964;
  if (r <= 0.0) return 0.0;
  if (r > 5.35) return fr * std::exp(fd / fr * (r - 5.35));
  return D * (be * std::exp(-al * (r - r0)) - al * std::exp(-be * (r - r0)));
}
template <int HH, typename T = double>
A2C_HD VDT<T> phi_morse_vd(T r) {
  const T fr = HH ? -0.003702722457147 : -3.353001113951400e-04;
  const T fd = HH ? 0.005465945471933 : 7.147062196813336e-04;
  const T D = HH ? 0.0661496 : 0.2494540, al = HH ? 3.67263 : 4.82613;
  const T be = HH ? 1.47797 : 2.13158, r0 = HH ? 2.51980 : 1.50964;
  if (r <= 0.0) return {0.0, 0.0};
  if (r > 5.35) {
    const T k = fd / fr;
    T y = fr * std::exp(k * (r - 5.35));
    return {y, k * y};
  }
  T ea = std::exp(-al * (r - r0)), eb = std::exp(-be * (r - r0));
  return {D * (be * ea - al * eb), D * al * be * (eb - ea)};
}

}  // namespace eam
}  // namespace a2c
