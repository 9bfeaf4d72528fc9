// Fermi-Dirac integrals I_p(alpha), p = -1/2, 1/2, 3/2, and alpha-derivatives of
// I_{-1/2}, with the inverse of I_{1/2} (fermi_dirac.f90).  A 6-column table on
// alpha in [-30, 300], h = 0.05, built once on the host by 20-point
// Gauss-Legendre panels, evaluated by cubic Hermite interpolation using the
// exact derivative relations; Sommerfeld / classical expansions outside.
// The same __host__ __device__ code serves the host (units, kernels) and the
// per-point electron functional on the GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace hy {
namespace fd {

constexpr double AMIN = -30.0, AMAX = 300.0, H = 0.05;
constexpr double INVH = 20.0;  // 1/H, exact: the node search multiplies instead of dividing
constexpr int NTAB = 6601;  // nint((AMAX - AMIN)/H) + 1
constexpr int NC = 6;       // columns: I-1/2, I1/2, I3/2, I'-1/2, I''-1/2, I'''-1/2
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double SQRTPI = 1.77245385090551602729816748334114518;

// Host-built table, tab[i*NC + c] (hy_model.hip).
const double* host_table();

__host__ __device__ inline int node(double a) {
  int i = (int)((a - AMIN) * INVH);
  return i < 0 ? 0 : (i > NTAB - 2 ? NTAB - 2 : i);
}

// cubic Hermite of column `col` with derivative values d0, d1 at the nodes i, i+1
__host__ __device__ inline double hermite(const double* tab, int col, double a, int i, double d0, double d1) {
  double t = (a - (AMIN + i * H)) * INVH;
  double t2 = t * t, t3 = t2 * t;
  double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
  double y0 = tab[i * NC + col], y1 = tab[(i + 1) * NC + col];
  return h00 * y0 + h10 * H * d0 + h01 * y1 + h11 * H * d1;
}

// m-th alpha-derivative of I_p from the Sommerfeld expansion (alpha >> 1).
// The tails are out of line: they are rarely taken, and inlined they would
// dominate the register allocation of the field kernels.
__host__ __device__ __attribute__((noinline)) double sommerfeld(double p, double a, int m) {
  double c[3] = {1.0, (PI * PI / 6.0) * p * (p + 1.0),
                 (7.0 * PI * PI * PI * PI / 360.0) * (p + 1.0) * p * (p - 1.0) * (p - 2.0)};
  double v = 0.0;
  for (int j = 0; j < 3; ++j) {
    double ex = p + 1.0 - 2 * j;
    double coef = c[j] / (p + 1.0);
    for (int l = 0; l < m; ++l) coef *= (ex - l);
    v += coef * pow(a, ex - m);
  }
  return v;
}

// m-th derivative of I_p from the classical expansion (alpha << -1); gp1 = Gamma(p+1)
__host__ __device__ __attribute__((noinline)) double classical(double p, double a, int m, double gp1) {
  double v = 0.0, sgn = 1.0;
  for (int j = 1; j <= 12; ++j) {
    v += sgn * exp(j * a) * pow((double)j, m - p - 1.0);
    sgn = -sgn;
  }
  return v * gp1;
}

__host__ __device__ inline double im12(const double* tab, double a) {
  if (a > AMAX) return sommerfeld(-0.5, a, 0);
  if (a < AMIN) return classical(-0.5, a, 0, SQRTPI);
  int i = node(a);
  return hermite(tab, 0, a, i, tab[i * NC + 3], tab[(i + 1) * NC + 3]);
}
__host__ __device__ inline double ip12(const double* tab, double a) {
  if (a > AMAX) return sommerfeld(0.5, a, 0);
  if (a < AMIN) return classical(0.5, a, 0, 0.5 * SQRTPI);
  int i = node(a);
  return hermite(tab, 1, a, i, 0.5 * tab[i * NC + 0], 0.5 * tab[(i + 1) * NC + 0]);
}
__host__ __device__ inline double ip32(const double* tab, double a) {
  if (a > AMAX) return sommerfeld(1.5, a, 0);
  if (a < AMIN) return classical(1.5, a, 0, 0.75 * SQRTPI);
  int i = node(a);
  return hermite(tab, 2, a, i, 1.5 * tab[i * NC + 1], 1.5 * tab[(i + 1) * NC + 1]);
}
__host__ __device__ inline double dim12(const double* tab, double a) {
  if (a > AMAX) return sommerfeld(-0.5, a, 1);
  if (a < AMIN) return classical(-0.5, a, 1, SQRTPI);
  int i = node(a);
  return hermite(tab, 3, a, i, tab[i * NC + 4], tab[(i + 1) * NC + 4]);
}
__host__ __device__ inline double d2im12(const double* tab, double a) {
  if (a > AMAX) return sommerfeld(-0.5, a, 2);
  if (a < AMIN) return classical(-0.5, a, 2, SQRTPI);
  int i = node(a);
  return hermite(tab, 4, a, i, tab[i * NC + 5], tab[(i + 1) * NC + 5]);
}
// xi = I'_{-1/2}/I_{-1/2}^2 and its alpha-derivative (fermi_dirac.f90:225-238)
__host__ __device__ inline double xi(const double* tab, double a) {
  double im = im12(tab, a);
  return dim12(tab, a) / (im * im);
}
__host__ __device__ inline double dxi(const double* tab, double a) {
  double im = im12(tab, a), d1 = dim12(tab, a), d2 = d2im12(tab, a);
  return (d2 * im - 2.0 * d1 * d1) / (im * im * im);
}

// alpha with I_{1/2}(alpha) = y > 0: safeguarded Newton (fermi_dirac.f90:241-267)
__host__ __device__ inline double alpha_from_ip12(const double* tab, double y) {
  const double huge = 1.7976931348623157e308;
  double alpha = fmax(log(2.0 * y / SQRTPI), pow(1.5 * y, 2.0 / 3.0) - 1.0);
  if (y < 1.0) alpha = log(2.0 * y / SQRTPI);
  double lo = -huge, hi = huge;
  for (int it = 0; it < 100; ++it) {
    double f = ip12(tab, alpha) - y;
    if (f > 0.0) hi = fmin(hi, alpha);
    else lo = fmax(lo, alpha);
    double df = 0.5 * im12(tab, alpha);
    double step = f / df;
    if (fabs(step) < 1.0e-14 * fmax(1.0, fabs(alpha))) break;
    alpha = alpha - step;
    if (alpha <= lo || alpha >= hi) {
      if (lo > -huge && hi < huge) alpha = 0.5 * (lo + hi);
    }
  }
  return alpha;
}

// I_{1/2}(a) and I_{-1/2}(a) from one node lookup (the pair every Newton
// iteration needs; same interpolants as ip12 and im12).
__host__ __device__ inline void ip12_im12(const double* tab, double a, double& ip, double& im) {
  if (a > AMAX || a < AMIN) {
    ip = ip12(tab, a);
    im = im12(tab, a);
    return;
  }
  const int i = node(a);
  const double* r0 = tab + i * NC;
  const double* r1 = r0 + NC;
  const double t = (a - (AMIN + i * H)) * INVH;
  const double t2 = t * t, t3 = t2 * t;
  const double h00 = 2 * t3 - 3 * t2 + 1, h10 = (t3 - 2 * t2 + t) * H, h01 = -2 * t3 + 3 * t2,
               h11 = (t3 - t2) * H;
  const double m0 = r0[0], m1 = r1[0];
  ip = h00 * r0[1] + h10 * (0.5 * m0) + h01 * r1[1] + h11 * (0.5 * m1);
  im = h00 * m0 + h10 * r0[3] + h01 * m1 + h11 * r1[3];
}

// I_{-1/2}, I'_{-1/2}, I''_{-1/2} at a from one node lookup (the gradient term
// needs all three: xi, d xi/d alpha and d n/d alpha).
__host__ __device__ inline void im12_d1_d2(const double* tab, double a, double& im, double& d1, double& d2) {
  if (a > AMAX || a < AMIN) {
    im = im12(tab, a);
    d1 = dim12(tab, a);
    d2 = d2im12(tab, a);
    return;
  }
  const int i = node(a);
  const double* r0 = tab + i * NC;
  const double* r1 = r0 + NC;
  const double t = (a - (AMIN + i * H)) * INVH;
  const double t2 = t * t, t3 = t2 * t;
  const double h00 = 2 * t3 - 3 * t2 + 1, h10 = (t3 - 2 * t2 + t) * H, h01 = -2 * t3 + 3 * t2,
               h11 = (t3 - t2) * H;
  im = h00 * r0[0] + h10 * r0[3] + h01 * r1[0] + h11 * r1[3];
  d1 = h00 * r0[3] + h10 * r0[4] + h01 * r1[3] + h11 * r1[4];
  d2 = h00 * r0[4] + h10 * r0[5] + h01 * r1[4] + h11 * r1[5];
}

// alpha_from_ip12 started at `a_start` instead of the closed-form guess: the
// same safeguarded Newton iteration and stopping rule, so it converges to the
// same root to the iteration's 1e-14 relative step tolerance.  Started from
// the expansion about the uniform state, it needs two to three iterations
// for the small density excursions of a run.
__host__ __device__ inline double alpha_from_ip12_from(const double* tab, double y, double a_start) {
  const double huge = 1.7976931348623157e308;
  double alpha = a_start;
  double lo = -huge, hi = huge;
  for (int it = 0; it < 100; ++it) {
    double ip, im;
    ip12_im12(tab, alpha, ip, im);
    const double f = ip - y;
    if (f > 0.0) hi = fmin(hi, alpha);
    else lo = fmax(lo, alpha);
    const double step = f / (0.5 * im);
    if (fabs(step) < 1.0e-14 * fmax(1.0, fabs(alpha))) break;
    alpha = alpha - step;
    if (alpha <= lo || alpha >= hi) {
      if (lo > -huge && hi < huge) alpha = 0.5 * (lo + hi);
    }
  }
  return alpha;
}

// ---------------------------------------------------------------- field kernels
// The forms the field kernels evaluate (also compiled for the host, where
// hy_fd_eval exposes them for checks against the functions above).
// Device-side table access: buffer loads through a wave-uniform descriptor
// over the whole table with 32-bit byte offsets.  (Plain loads from the table
// pointer become generic-address loads, whose waits also drain LDS traffic.)
// node() keeps every offset inside the table; the descriptor's record count
// bounds it as well (out-of-range buffer loads return 0).
__host__ __device__ inline double tabd(const double* tab, int idx) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(tab), (short)0, NTAB * NC * 8, 0x00020000);
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (unsigned)idx * 8u, 0, 0));
#else
  return tab[idx];
#endif
}

// ---- the expansions outside the table, inline for the field kernels.
// The same series as sommerfeld() and classical(), with the half-integer
// powers formed from one square root and integer powers instead of pow():
// they agree to a few ulp.  Kept inline because a call from a kernel makes
// the compiler reserve registers for the calling convention across the
// whole kernel (the electron x-line kernel needed about 30 more VGPRs).
// Sommerfeld (a > AMAX): I_p^(m) = sum_j c_j/(p+1) prod_{l<m}(e_j - l) a^(e_j - m),
// e_j = p + 1 - 2 j, c = {1, pi^2/6 p(p+1), 7 pi^4/360 (p+1)p(p-1)(p-2)}.
constexpr double SOM_C1M = (PI * PI / 6.0) * (-0.5) * 0.5;                          // p = -1/2
constexpr double SOM_C2M = (7.0 * PI * PI * PI * PI / 360.0) * 0.5 * (-0.5) * (-1.5) * (-2.5);
constexpr double SOM_C1P = (PI * PI / 6.0) * 0.5 * 1.5;                             // p = +1/2
constexpr double SOM_C2P = (7.0 * PI * PI * PI * PI / 360.0) * 1.5 * 0.5 * (-0.5) * (-1.5);
// classical (a < AMIN): I_p^(m) = Gamma(p+1) sum_{j=1..12} (-1)^(j+1) e^(j a) j^(m-p-1)
__host__ __device__ __attribute__((always_inline)) inline void tail_ip12_im12(double a, double& ip, double& im) {
  if (a > AMAX) {
    const double r = sqrt(a), ia = 1.0 / a, ia2 = ia * ia;
    // p = 1/2: e = 3/2, -1/2, -5/2; p = -1/2: e = 1/2, -3/2, -7/2 (coefficient c_j / (p+1))
    ip = (r * a) / 1.5 + (SOM_C1P / 1.5) * (r * ia) + (SOM_C2P / 1.5) * (r * ia2 * ia);
    im = r / 0.5 + (SOM_C1M / 0.5) * (r * ia2) + (SOM_C2M / 0.5) * (r * ia2 * ia2);
    return;
  }
  double sp = 0.0, sm = 0.0, sgn = 1.0;
#pragma unroll 1
  for (int j = 1; j <= 12; ++j) {
    const double e = exp(j * a), rj = sqrt((double)j);
    sp += sgn * e / (j * rj);  // j^(-3/2)
    sm += sgn * e / rj;        // j^(-1/2)
    sgn = -sgn;
  }
  ip = sp * (0.5 * SQRTPI);
  im = sm * SQRTPI;
}
__host__ __device__ __attribute__((always_inline)) inline void tail_im12_d1_d2(double a, double& im, double& d1, double& d2) {
  if (a > AMAX) {
    const double r = sqrt(a), ia = 1.0 / a, ia2 = ia * ia;
    // p = -1/2, e = 1/2, -3/2, -7/2; m-th derivative: prod_{l<m} (e - l) a^(e - m)
    const double c0 = 1.0 / 0.5, c1 = SOM_C1M / 0.5, c2 = SOM_C2M / 0.5;
    im = c0 * r + c1 * (r * ia2) + c2 * (r * ia2 * ia2);
    d1 = c0 * 0.5 * (r * ia) + c1 * (-1.5) * (r * ia2 * ia) + c2 * (-3.5) * (r * ia2 * ia2 * ia);
    d2 = c0 * (0.5 * -0.5) * (r * ia2) + c1 * (-1.5 * -2.5) * (r * ia2 * ia2) +
         c2 * (-3.5 * -4.5) * (r * ia2 * ia2 * ia2);
    return;
  }
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, sgn = 1.0;
#pragma unroll 1
  for (int j = 1; j <= 12; ++j) {
    const double e = sgn * exp(j * a), rj = sqrt((double)j);
    s0 += e / rj;      // j^(-1/2)
    s1 += e * rj;      // j^(1/2)
    s2 += e * (j * rj);  // j^(3/2)
    sgn = -sgn;
  }
  im = s0 * SQRTPI;
  d1 = s1 * SQRTPI;
  d2 = s2 * SQRTPI;
}

// ip12_im12 with device gathers
__host__ __device__ inline void ip12_im12_dev(const double* tab, double a, double& ip, double& im) {
  if (a > AMAX || a < AMIN) {
    tail_ip12_im12(a, ip, im);
    return;
  }
  const int i = node(a);
  const int o = i * NC;
  const double t = (a - (AMIN + i * H)) * INVH;
  const double t2 = t * t, t3 = t2 * t;
  const double h00 = 2 * t3 - 3 * t2 + 1, h10 = (t3 - 2 * t2 + t) * H, h01 = -2 * t3 + 3 * t2,
               h11 = (t3 - t2) * H;
  const double m0 = tabd(tab, o), p0 = tabd(tab, o + 1), d0 = tabd(tab, o + 3);
  const double m1 = tabd(tab, o + NC), p1 = tabd(tab, o + NC + 1), d1 = tabd(tab, o + NC + 3);
  ip = h00 * p0 + h10 * (0.5 * m0) + h01 * p1 + h11 * (0.5 * m1);
  im = h00 * m0 + h10 * d0 + h01 * m1 + h11 * d1;
}

// im12_d1_d2 with device gathers
__host__ __device__ inline void im12_d1_d2_dev(const double* tab, double a, double& im, double& d1, double& d2) {
  if (a > AMAX || a < AMIN) {
    tail_im12_d1_d2(a, im, d1, d2);
    return;
  }
  const int i = node(a);
  const int o = i * NC;
  const double t = (a - (AMIN + i * H)) * INVH;
  const double t2 = t * t, t3 = t2 * t;
  const double h00 = 2 * t3 - 3 * t2 + 1, h10 = (t3 - 2 * t2 + t) * H, h01 = -2 * t3 + 3 * t2,
               h11 = (t3 - t2) * H;
  const double a0 = tabd(tab, o), a3 = tabd(tab, o + 3), a4 = tabd(tab, o + 4), a5 = tabd(tab, o + 5);
  const double b0 = tabd(tab, o + NC), b3 = tabd(tab, o + NC + 3), b4 = tabd(tab, o + NC + 4),
               b5 = tabd(tab, o + NC + 5);
  im = h00 * a0 + h10 * a3 + h01 * b0 + h11 * b3;
  d1 = h00 * a3 + h10 * a4 + h01 * b3 + h11 * b4;
  d2 = h00 * a4 + h10 * a5 + h01 * b4 + h11 * b5;
}

// 1/d from the hardware reciprocal and one Newton-Raphson correction
__host__ __device__ inline double rcp_nr(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(d);
  return fma(fma(-d, r, 1.0), r, r);
#else
  return 1.0 / d;
#endif
}

// The I_{1/2} interpolant at a inside the table and its first two
// alpha-derivatives: the exact derivatives of the cubic piece of node i
// (the piece the Newton iteration solves on).
__host__ __device__ inline void ip12_d2_dev(const double* tab, double a, int i, double& ip, double& dp, double& d2p) {
  const int o = i * NC;
  const double t = (a - (AMIN + i * H)) * INVH;
  const double t2 = t * t, t3 = t2 * t;
  const double h00 = 2 * t3 - 3 * t2 + 1, h10 = (t3 - 2 * t2 + t) * H, h01 = -2 * t3 + 3 * t2,
               h11 = (t3 - t2) * H;
  const double g00 = 6 * t2 - 6 * t, g10 = (3 * t2 - 4 * t + 1) * H, g01 = -6 * t2 + 6 * t, g11 = (3 * t2 - 2 * t) * H;
  const double k00 = 12 * t - 6, k10 = (6 * t - 4) * H, k01 = -12 * t + 6, k11 = (6 * t - 2) * H;
  const double q0 = 0.5 * tabd(tab, o), p0 = tabd(tab, o + 1);
  const double q1 = 0.5 * tabd(tab, o + NC), p1 = tabd(tab, o + NC + 1);
  ip = h00 * p0 + h10 * q0 + h01 * p1 + h11 * q1;
  dp = (g00 * p0 + g10 * q0 + g01 * p1 + g11 * q1) * INVH;
  d2p = (k00 * p0 + k10 * q0 + k01 * p1 + k11 * q1) * (INVH * INVH);
}

// alpha_from_ip12_from with device gathers: the same iteration and safeguard;
// the step's division is the hardware reciprocal refined once.  It first
// takes one Halley step on the cubic piece that holds the start, with the
// piece's own first and second derivatives; Halley's error on a cubic is
// at most about (p''/(2p'))^2 + |p'''/(6p')| <= 0.5 times step^3, so a step
// below 1e-5 max(1, |alpha|) that stays in the piece ends the solve (error
// <= 5e-16 max(1, |alpha|)^3, below the 1e-14 of the reference's stopping
// point).  Otherwise the reference's Newton iteration continues from the
// better of the two points; besides its rule (stop before a step below
// 1e-14 max(1, |alpha|)), a Newton step below 2e-8 max(1, |alpha|) that
// needed no safeguard ends it: the error after it is at most
// |I''/(2 I')| step^2 <= 2e-16 max(1, |alpha|)^2 (|I''/(2 I')| <= 1/2).
// From the field kernels' second-order start this is one table evaluation
// per point at the density excursions of a run (|dn/n0| of a few 1e-3).
__host__ __device__ inline double alpha_from_ip12_dev(const double* tab, double y, double a_start) {
  const double huge = 1.7976931348623157e308;
  double alpha = a_start;
  if (!(alpha > AMAX || alpha < AMIN)) {
    const int i = node(alpha);
    double ip, dp, d2p;
    ip12_d2_dev(tab, alpha, i, ip, dp, d2p);
    const double f = ip - y;
    const double step = 2.0 * f * dp * rcp_nr(2.0 * dp * dp - f * d2p);
    if (fabs(step) < 1.0e-14 * fmax(1.0, fabs(alpha))) return alpha;
    const double an = alpha - step;
    if (!(an > AMAX || an < AMIN) && node(an) == i && fabs(step) < 1.0e-5 * fmax(1.0, fabs(an))) return an;
    if (fabs(step) < 0.5) alpha = an;  // a sound start for the safeguarded iteration
  }
  double lo = -huge, hi = huge;
  for (int it = 0; it < 100; ++it) {
    double ip, im;
    ip12_im12_dev(tab, alpha, ip, im);
    const double f = ip - y;
    if (f > 0.0) hi = fmin(hi, alpha);
    else lo = fmax(lo, alpha);
    const double step = f * rcp_nr(0.5 * im);
    if (fabs(step) < 1.0e-14 * fmax(1.0, fabs(alpha))) break;
    alpha = alpha - step;
    if (alpha <= lo || alpha >= hi) {
      if (lo > -huge && hi < huge) alpha = 0.5 * (lo + hi);
    } else if (fabs(step) < 2.0e-8 * fmax(1.0, fabs(alpha))) {
      break;
    }
  }
  return alpha;
}

}  // namespace fd
}  // namespace hy
