// Register-resident DFTs and the two-level (L = L1 x L2) column FFT used by
// the batched y/z pencil passes.
//
// A column of length L is split as n = L2*n1 + n2, k = k1 + L1*k2:
//   X[k1 + L1 k2] = sum_n2 W_L2^(n2 k2) [ W_L^(n2 k1) sum_n1 W_L1^(n1 k1) x[L2 n1 + n2] ]
// Step 1 (registers): DFTs of length L1 over n1, one per n2.  Step 2: the
// twiddle W_L^(n2 k1).  Step 3: one transpose through LDS.  Step 4
// (registers): DFTs of length L2 over n2, one per k1.  A group of TPC threads
// shares a column (TPC divides L1 and L2); columns with L <= 32 are done by
// one thread entirely in registers.  Small DFTs (2, 3, 4 and their products up
// to 16, plus 24 and 32) are unrolled at compile time with constant twiddles.
#pragma once
#include "hy_internal.h"

namespace hy {

// ------------------------------------------------ compile-time cos/sin(2 pi k/n)
namespace ct {
constexpr double PI = 3.14159265358979323846264338327950288;
// Taylor series on |x| <= pi/4 (terms to x^25: error below 1e-26)
constexpr double sin_small(double x) {
  double t = x, s = x;
  for (int i = 1; i <= 12; ++i) { t *= -x * x / ((2.0 * i) * (2.0 * i + 1.0)); s += t; }
  return s;
}
constexpr double cos_small(double x) {
  double t = 1.0, s = 1.0;
  for (int i = 1; i <= 12; ++i) { t *= -x * x / ((2.0 * i - 1.0) * (2.0 * i)); s += t; }
  return s;
}
// angle 2 pi k / n reduced exactly in units of pi/(4n) to the first octant
constexpr void sincos2pi(long k, long n, double& s, double& c) {
  long a = ((8 * k) % (8 * n) + 8 * n) % (8 * n);  // angle = pi a / (4 n), a in [0, 8n)
  double sg_s = 1.0, sg_c = 1.0;
  if (a >= 4 * n) { a -= 4 * n; sg_s = -1.0; sg_c = -1.0; }  // +pi
  bool swap = false;
  if (a >= 2 * n) {                                             // +pi/2: (s,c) -> (c,-s)
    a -= 2 * n;
    swap = true;
  }
  // now angle in [0, pi/2); fold about pi/4
  double ss, cc;
  if (a > n) {
    const double x = PI * (double)(2 * n - a) / (4.0 * (double)n);
    ss = cos_small(x); cc = sin_small(x);
  } else {
    const double x = PI * (double)a / (4.0 * (double)n);
    ss = sin_small(x); cc = cos_small(x);
  }
  if (swap) { s = sg_s * cc; c = -sg_c * ss; }
  else { s = sg_s * ss; c = sg_c * cc; }
}
constexpr double cos2pi(long k, long n) { double s = 0, c = 0; sincos2pi(k, n, s, c); return c; }
constexpr double sin2pi(long k, long n) { double s = 0, c = 0; sincos2pi(k, n, s, c); return s; }
}  // namespace ct

// v * W_N^e with W_N = exp(-2 pi i / N) (forward) or its conjugate (INV)
template <class T, int N, int E, bool INV>
__device__ inline cpx<T> twc(cpx<T> v) {
  constexpr int e = ((E % N) + N) % N;
  if constexpr (e == 0) {
    return v;
  } else if constexpr (4 * e == N) {  // -i (forward) / +i (inverse)
    return INV ? mk<T>(-v.y, v.x) : mk<T>(v.y, -v.x);
  } else if constexpr (2 * e == N) {
    return mk<T>(-v.x, -v.y);
  } else if constexpr (4 * e == 3 * N) {
    return INV ? mk<T>(v.y, -v.x) : mk<T>(-v.y, v.x);
  } else {
    constexpr double c = ct::cos2pi(e, N);
    constexpr double s = INV ? ct::sin2pi(e, N) : -ct::sin2pi(e, N);
    const T cr = (T)c, sr = (T)s;
    return mk<T>(v.x * cr - v.y * sr, v.x * sr + v.y * cr);
  }
}

template <class T, int N, int R, int M, bool INV, int K1>
__device__ inline void rdft_combine(cpx<T> (&s)[R][M], cpx<T>* v);

// In-register DFT of length N (natural order in and out), mixed radix 4/2/3,
// decimation in time; all twiddles are compile-time constants.
template <class T, int N, bool INV>
__device__ inline void rdft(cpx<T>* v) {
  if constexpr (N == 1) {
    return;
  } else if constexpr (N == 2 || N == 3 || N == 4) {
    dftR<T, INV, N>(v);
  } else {
    constexpr int R = (N % 4 == 0) ? 4 : ((N % 2 == 0) ? 2 : 3);
    constexpr int M = N / R;
    cpx<T> s[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int m = 0; m < M; ++m) s[r][m] = v[R * m + r];
#pragma unroll
    for (int r = 0; r < R; ++r) rdft<T, M, INV>(s[r]);
    rdft_combine<T, N, R, M, INV, 0>(s, v);
  }
}

// X[k1 + M k2] = sum_r W_R^(r k2) (W_N^(r k1) S_r[k1]), k1 unrolled by recursion
template <class T, int N, int R, int M, bool INV, int K1>
__device__ inline void rdft_combine(cpx<T> (&s)[R][M], cpx<T>* v) {
  if constexpr (K1 < M) {
    cpx<T> t[R];
    t[0] = s[0][K1];
    if constexpr (R > 1) t[1] = twc<T, N, K1, INV>(s[1][K1]);
    if constexpr (R > 2) t[2] = twc<T, N, 2 * K1, INV>(s[2][K1]);
    if constexpr (R > 3) t[3] = twc<T, N, 3 * K1, INV>(s[3][K1]);
    dftR<T, INV, R>(t);
#pragma unroll
    for (int k2 = 0; k2 < R; ++k2) v[K1 + M * k2] = t[k2];
    rdft_combine<T, N, R, M, INV, K1 + 1>(s, v);
  }
}

// ------------------------------------------------------ two-level column plans
// (L1, L2, TPC) per length; L1 * L2 = L, TPC divides both.  TPC = 1 keeps a
// whole column in one thread's registers (columns of at most 16 points);
// longer columns are split over TPC threads with one LDS transpose, which
// keeps a thread well under 128 VGPRs (a 32-point FP32 column in one thread
// needs about 140 and runs at 3 waves per SIMD).
struct ColPlan { int L1, L2, TPC; };
__host__ __device__ constexpr ColPlan col_plan(int L, int /*elem_bytes*/) {
  return L == 256 ? ColPlan{16, 16, 16}
       : L == 128 ? ColPlan{16, 8, 8}
       : L == 96  ? ColPlan{12, 8, 4}
       : L == 64  ? ColPlan{8, 8, 8}
       : L == 48  ? ColPlan{12, 4, 4}
       : L == 32  ? ColPlan{8, 4, 4}
       : L == 24  ? ColPlan{6, 4, 2}
                  : ColPlan{L, 1, 1};
}

}  // namespace hy
