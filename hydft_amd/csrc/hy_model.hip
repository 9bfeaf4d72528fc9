// The time-dependent model (model.f90) on the GPU with a spectral state.
//
// State: the masked half spectra of n, mom(3) and, for the Maxwell closure,
// Pi(6), stored only on the kept 2/3-rule box (kx index < Kx, |my| <= My,
// |mz| <= Mz) in FP64.  The reference dealiases the state after every stage,
// so its real-space state is always band limited and equals the backward
// transform of this spectral state.
//
// One RK3 stage:
//   A  state -> x-lines (z, y inverse passes) -> x kernel A: n, u, mu_loc, flux
//      -> y, z forward passes
//   K1 mu_hat = F(mu_loc) + K(k) n_hat'   (+ pending Maxwell relaxation)
//   B  mu_hat, Pi_hat -> x-lines -> x kernel B: n (f_ext - grad mu), -u.grad Pi
//      -> y, z forward passes
//   K2 R_hat assembly (continuity, flux divergence, stress divergence) and the
//      low-storage RK3 update S += dt (a R + b R_old), R_old = R.
// Real fields never leave LDS/registers: each x kernel inverse-transforms its
// inputs along x, does the pointwise physics and forward-transforms the results.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <vector>

#include "hy_fd.h"
#include "hy_internal.h"

namespace hy {

// ------------------------------------------------------------ FD table (host)
static std::vector<double> g_fdtab;
static std::once_flag g_fdonce;

static void gauss_legendre(int n, double* x, double* w) {  // quadrature.f90:17-38
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double z = std::cos(fd::PI * (i - 0.25) / (n + 0.5)), pp = 0, z1;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0, p3;
      for (int j = 1; j <= n; ++j) {
        p3 = p2; p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1.0e-15) break;
    }
    x[i - 1] = -z; x[n - i] = z;
    w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp); w[n - i] = w[i - 1];
  }
}
static void fermi_derivs(double y, double& f, double& p1, double& p2, double& p3) {
  if (y > 0.0) { double g = std::exp(-y); f = g / (1.0 + g); }
  else { double g = std::exp(y); f = 1.0 / (1.0 + g); }
  p1 = f * (1.0 - f);
  p2 = p1 * (1.0 - 2.0 * f);
  p3 = p1 * (1.0 - 6.0 * f + 6.0 * f * f);
}
static void fd_build() {  // fermi_dirac.f90:30-74
  const int ngl = 20;
  double xg[ngl], wg[ngl];
  gauss_legendre(ngl, xg, wg);
  g_fdtab.assign((size_t)fd::NTAB * fd::NC, 0.0);
  for (int i = 0; i < fd::NTAB; ++i) {
    double alpha = fd::AMIN + i * fd::H;
    double xmax = std::max(alpha, 0.0) + 45.0;
    double s[6] = {0, 0, 0, 0, 0, 0}, f, p1, p2, p3;
    for (int q = 0; q < ngl; ++q) {
      double t = 0.5 * (xg[q] + 1.0), x = t * t;
      fermi_derivs(x - alpha, f, p1, p2, p3);
      double e = wg[q] * 0.5 * 2.0;
      s[0] += e * f; s[1] += e * t * t * f; s[2] += e * std::pow(t, 4) * f;
      s[3] += e * p1; s[4] += e * p2; s[5] += e * p3;
    }
    int npan = (int)std::ceil(xmax - 1.0);
    for (int ip = 1; ip <= npan; ++ip) {
      double a = 1.0 + (ip - 1), b = std::min(a + 1.0, xmax);
      double xm = 0.5 * (a + b), xr = 0.5 * (b - a);
      for (int q = 0; q < ngl; ++q) {
        double x = xm + xr * xg[q];
        fermi_derivs(x - alpha, f, p1, p2, p3);
        double e = wg[q] * xr, t = 1.0 / std::sqrt(x), sx = std::sqrt(x);
        s[0] += e * t * f; s[1] += e * sx * f; s[2] += e * x * sx * f;
        s[3] += e * t * p1; s[4] += e * t * p2; s[5] += e * t * p3;
      }
    }
    for (int c = 0; c < 6; ++c) g_fdtab[(size_t)i * fd::NC + c] = s[c];
  }
}
const double* fd::host_table() {
  std::call_once(g_fdonce, fd_build);
  return g_fdtab.data();
}

// ------------------------------------------------------------ device control
struct Ctl {
  double t, dt, ts;          // model time, step size, current stage time
  double relax_fac, relax_phi;
  double umax;
  double t_vext;             // time of the last v_ext evaluation (stage start)
  int step;
  int nonpos;                // set when a non-positive density is seen
  int nodt;                  // compute_dt could not find a time step
  // step limit of a run in a batch of runs of different lengths: at the start
  // of each recorded step sequence, halt = (stop > 0 && step >= stop), and
  // every kernel of the sequence returns at once while halt is set
  int stop;
  int halt;
};

__constant__ double c_rk_a[3] = {32.0 / 60.0, 25.0 / 60.0, 45.0 / 60.0};  // model.f90:51-52
__constant__ double c_rk_b[3] = {0.0, -17.0 / 60.0, -25.0 / 60.0};
static const double h_rk_a[3] = {32.0 / 60.0, 25.0 / 60.0, 45.0 / 60.0};
static const double h_rk_b[3] = {0.0, -17.0 / 60.0, -25.0 / 60.0};

enum CtlOp { CTL_STEP_BEGIN = 2, CTL_SEQ_BEGIN = 3, CTL_OPEN = 4 };  // the ends of stages and steps: K2
struct CtlK {
  Ctl* c;
  int op;
  double tend;
};
// The ends of a stage and of a step (time, stage time, Maxwell relaxation
// factors of the next stage, step count).  K2 applies them from its first
// thread: nothing in K2 reads the fields they change, and the next stage's
// kernels follow K2 on the stream.
__device__ inline void ctl_stage_end(Ctl* c, int s, double tau, int evolve) {
  c->t_vext = c->ts;
  double dts = (c_rk_a[s] + c_rk_b[s]) * c->dt;
  c->ts = c->ts + dts;
  if (evolve) {
    double fac = exp(-dts / tau);
    c->relax_fac = fac;
    c->relax_phi = (tau / dts) * (1.0 - fac);
  }
}
__device__ inline void ctl_step_end(Ctl* c) {
  c->t = c->t + c->dt;
  c->step += 1;
  c->ts = c->t;
}
__device__ __attribute__((always_inline)) inline void ctl_body(const CtlK& k) {
  Ctl* c = k.c;
  const int op = k.op;
  const double tend = k.tend;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (op == CTL_SEQ_BEGIN) {
    c->halt = (c->stop > 0 && c->step >= c->stop) ? 1 : 0;
    return;
  }
  if (op == CTL_OPEN) {  // steps taken one model at a time ignore a batch's step limit
    c->halt = 0;
    return;
  }
  if (c->halt) return;
  if (op == CTL_STEP_BEGIN) {
    if (tend > 0.0 && c->t + c->dt > tend) c->dt = tend - c->t;  // src/hydft.f90:44
    c->ts = c->t;
  }
}
__global__ void k_ctl(CtlK k) { ctl_body(k); }
__global__ void k_ctl_b(const CtlK* __restrict__ ks) { ctl_body(ks[blockIdx.z]); }
__global__ void k_set_stop(Ctl* c, int stop) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->stop = stop;
    c->halt = 0;
  }
}

// ------------------------------------------------------------ global memory
// (Device compilation; a host-side build of this file, such as a CPU
// emulation, uses plain pointers.)
// Loads and stores through the global address space with a 32-bit byte
// offset from a uniform base.  The x-line kernels take their buffer pointers
// from an argument block in memory, and the batched K1/K2 kernels from a
// device array; the compiler cannot place such pointers in an address space,
// and this form gives scalar-base global accesses instead of 64-bit flat
// addresses per lane.  Every buffer addressed this way is below
// 4 GiB.
template <class V>
__device__ inline V gld(const void* base, size_t elem) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) const char gchar;
  typedef __attribute__((address_space(1))) const V gV;
  return *(const gV*)((const gchar*)(uintptr_t)base + (unsigned)(elem * sizeof(V)));
#else
  return reinterpret_cast<const V*>(base)[elem];
#endif
}
template <class V>
__device__ inline void gst(void* base, size_t elem, V v) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(1))) char gchar;
  typedef __attribute__((address_space(1))) V gV;
  *(gV*)((gchar*)(uintptr_t)base + (unsigned)(elem * sizeof(V))) = v;
#else
  reinterpret_cast<V*>(base)[elem] = v;
#endif
}
// complex elements move as two-component vectors (one 8- or 16-byte access)
template <class T>
__device__ inline cpx<T> gldc(const void* base, size_t elem) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = gld<v2>(base, elem);
  return mk<T>(v.x, v.y);
#else
  return reinterpret_cast<const cpx<T>*>(base)[elem];
#endif
}
template <class T>
__device__ inline void gstc(void* base, size_t elem, cpx<T> c) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 v;
  v.x = c.x;
  v.y = c.y;
  gst<v2>(base, elem, v);
#else
  reinterpret_cast<cpx<T>*>(base)[elem] = c;
#endif
}

// ------------------------------------------------------------ x-line inputs
struct XIn {
  const void* p;   // (Kx, nlines) spectral x-lines
  int is64;        // element type c64 (1) or compute type (0)
  int mul;         // 0 none, 1 i kx, 2 -kx^2
  double k0off;    // added to the (kx=0, line 0) entry (1-D first pass)
};
enum XMul { XM_NONE = 0, XM_IK = 1, XM_NEGK2 = 2 };

// The lane's inputs of a first transform stage, gathered before any is used
// (kx of column kk is kk * dkx, computed: a table load would be a
// generic-address load whose wait also drains the wave's pending loads):
// one uniform branch per field on the element type, then every load of the
// field issued back to back (a wait between them would expose each load's
// full latency); the kx multipliers and the k = 0 offset follow once all
// loads are in flight.  ok[i] false: element i is not read (zero).
//
// M >= 0: the input is a compute-type x-line slot with no k = 0 offset and
// the multiplier M (an XMul), known at compile time (every input of a 3-D
// grid); M = -1: element type, offset and multiplier are read from X.
template <class T, int NI, int M = -1>
__device__ inline void xin_gather(const XIn* X, size_t nlines, size_t line, int Kx, const int* kk, const bool* ok,
                                  cpx<T>* v) {
#pragma unroll
  for (int i = 0; i < NI; ++i) v[i] = mk<T>(0, 0);
  if (!X) return;
  const unsigned base = (unsigned)(line * Kx);
  if constexpr (M >= 0) {
    const unsigned nb = (unsigned)(nlines * Kx * sizeof(cpx<T>));
#pragma unroll
    for (int i = 0; i < NI; ++i) v[i] = bload_c<T>(X->p, nb, base + (unsigned)kk[i], ok[i]);
  } else if (X->is64) {
    const unsigned nb = (unsigned)(nlines * Kx * sizeof(c64));
#pragma unroll
    for (int i = 0; i < NI; ++i) v[i] = cvt<T>(bload_c<double>(X->p, nb, base + (unsigned)kk[i], ok[i]));
  } else {
    const unsigned nb = (unsigned)(nlines * Kx * sizeof(cpx<T>));
#pragma unroll
    for (int i = 0; i < NI; ++i) v[i] = bload_c<T>(X->p, nb, base + (unsigned)kk[i], ok[i]);
  }
}
template <class T, int NI, int M = -1>
__device__ inline void xin_finish(const XIn* X, size_t line, double dkx, const int* kk, const bool* ok, cpx<T>* v) {
  if (!X) return;
  if constexpr (M >= 0) {
    if constexpr (M == XM_IK) {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const T k = (T)((double)kk[i] * dkx);
        v[i] = mk<T>(-k * v[i].y, k * v[i].x);  // skipped elements are zero
      }
    } else if constexpr (M == XM_NEGK2) {
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const double kd = (double)kk[i] * dkx;
        const T k2 = (T)(kd * kd);
        v[i] = mk<T>(-k2 * v[i].x, -k2 * v[i].y);
      }
    }
    return;
  }
  if (X->k0off != 0.0 && line == 0) {
#pragma unroll
    for (int i = 0; i < NI; ++i)
      if (ok[i] && kk[i] == 0) v[i].x = (T)((double)v[i].x + X->k0off);
  }
  if (X->mul == XM_IK) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const T k = (T)((double)kk[i] * dkx);
      if (ok[i]) v[i] = mk<T>(-k * v[i].y, k * v[i].x);
    }
  } else if (X->mul == XM_NEGK2) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const double kd = (double)kk[i] * dkx;
      const T k2 = (T)(kd * kd);
      if (ok[i]) v[i] = mk<T>(-k2 * v[i].x, -k2 * v[i].y);
    }
  }
}

// ---- x-line transforms on TPC lanes of a wave
// An NX-point line is transformed by TPC lanes (64 / TPC lines per wave, see
// XWave) with a Stockham sequence whose first stage
// reads its operands from the caller (global loads or the lane's registers)
// and whose last stage hands its results back in registers; only the middle
// stages exchange data through two LDS buffers (one barrier per stage).  The
// inverse transforms use the radix order radix_of(NX, s) and the forward ones
// the reversed order, so the points a lane holds after an inverse transform
// are exactly the operands of its first forward stage: lane `lane` (of the
// line's TPC lanes) owns the XQ<NX, TPC>::Q points x = xown<NX, TPC>(lane, q).
template <int L, bool REV>
__host__ __device__ constexpr int prad(int s) { return REV ? radix_of(L, nstages_of(L) - 1 - s) : radix_of(L, s); }
template <int L, bool REV>
__host__ __device__ constexpr int pns(int s) { int p = 1; for (int i = 0; i < s; ++i) p *= prad<L, REV>(i); return p; }

// LDS slot of element e of a line buffer or twiddle table: bits 4-5 of e are
// XORed into bits 0-1 and 2-3.  The strided Stockham stores (stride 4 in the
// first stage, groups of 4 in the second) and the twiddle reads then hit 16
// different bank pairs per 16 lanes.  The map permutes every aligned block
// of 16 elements (and [16, 24)), so it is a bijection on [0, NX) for every
// supported NX.
__device__ constexpr int swz(int e) { return e ^ (((e >> 4) & 3) * 5); }

template <int NX, int TPC = 64>
struct XQ {
  static constexpr int RL = radix_of(NX, nstages_of(NX) - 1);  // last inverse radix
  static constexpr int NBL = NX / RL;
  static constexpr int PERL = (NBL + TPC - 1) / TPC;
  static constexpr int Q = PERL * RL;  // points per lane
};
// point x of the lane's q-th value, or -1 (lane idle for that slot)
template <int NX, int TPC = 64>
__device__ inline int xown(int lane, int q) {
  const int j = lane + TPC * (q / XQ<NX, TPC>::RL);
  return j < XQ<NX, TPC>::NBL ? j + (q % XQ<NX, TPC>::RL) * XQ<NX, TPC>::NBL : -1;
}

template <class T, int L, bool INV, bool REV, int TPC, int S, class In, class Out>
__device__ inline void xstage(const cpx<T>* src, cpx<T>* dst, const cpx<T>* tw, int lane, In& in, Out& out) {
  constexpr int NST = nstages_of(L);
  constexpr int R = prad<L, REV>(S);
  constexpr int NS = pns<L, REV>(S);
  constexpr int NB = L / R;
  constexpr int PER = (NB + TPC - 1) / TPC;
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int j = lane + TPC * b;
    if (NB % TPC == 0 || j < NB) {
      cpx<T> v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (S == 0) v[r] = in(b, r, j + r * NB);
        else v[r] = src[swz(j + r * NB)];
      }
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cpx<T> w = tw[swz(k * r * (L / (NS * R)))];
          v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
        }
      }
      dftR<T, INV, R>(v);
      if constexpr (S == NST - 1) {  // NS == NB here: outputs sit where the inputs were
#pragma unroll
        for (int r = 0; r < R; ++r) out(b, r, j + r * NB, v[r]);
      } else {
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[swz(base + r * NS)] = v[r];
      }
    }
  }
}

// stage S reads buf[(S+1)%2] (S > 0) and writes buf[S%2]
template <class T, int L, bool INV, bool REV, int TPC, int S = 0, class In, class Out>
__device__ inline void xfft(cpx<T>* b0, cpx<T>* b1, const cpx<T>* tw, int lane, In& in, Out& out) {
  if constexpr (S < nstages_of(L)) {
    xstage<T, L, INV, REV, TPC, S>(S % 2 ? b0 : b1, S % 2 ? b1 : b0, tw, lane, in, out);
    if constexpr (S + 1 < nstages_of(L)) __syncthreads();
    xfft<T, L, INV, REV, TPC, S + 1>(b0, b1, tw, lane, in, out);
  }
}

// Inverse x transform of two spectral x-lines packed as a + i b (a c2r pair:
// DC and Nyquist imaginary parts dropped); the lane's points of a(x) and
// b(x) are returned in qa[q], qb[q] (qb may be null).  Lanes of a line
// beyond the last (lok false) read nothing.  MA, MB: the inputs' compile-time
// multipliers (see xin_gather), or -1.  The transform is split in two:
// x_inv_load gathers the first stage's operands into registers, x_inv_run
// transforms them, so a caller can issue the next pair's loads before it
// transforms the current pair and overlap their latency with that work.
//
// The first stage's operands: element e = j + r NB0 of the line for
// j = lane + TPC b, held as i = b R0 + r.
template <class T, int NX, int TPC>
struct XInvIn {
  static constexpr int R0 = prad<NX, false>(0), NB0 = NX / R0, NI = ((NB0 + TPC - 1) / TPC) * R0;
  cpx<T> va[NI], vb[NI];
  __device__ static int kk(int lane, int i) {
    const int j = lane + TPC * (i / R0), e = j + (i % R0) * NB0;
    return 2 * e > NX ? NX - e : e;
  }
};
template <class T, int NX, int TPC = 64, int MA = -1, int MB = -1>
__device__ inline void x_inv_load(const XIn* A, const XIn* B, size_t nlines, size_t line, bool lok, int Kx,
                                  double dkx, int lane, XInvIn<T, NX, TPC>& P) {
  using P_t = XInvIn<T, NX, TPC>;
  constexpr int NI = P_t::NI;
  int kk[NI];
  bool ok[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int j = lane + TPC * (i / P_t::R0);
    kk[i] = P_t::kk(lane, i);
    ok[i] = lok && (P_t::NB0 % TPC == 0 || j < P_t::NB0) && kk[i] < Kx;
  }
  xin_gather<T, NI, MA>(A, nlines, line, Kx, kk, ok, P.va);
  xin_gather<T, NI, MB>(B, nlines, line, Kx, kk, ok, P.vb);
  xin_finish<T, NI, MA>(A, line, dkx, kk, ok, P.va);
  xin_finish<T, NI, MB>(B, line, dkx, kk, ok, P.vb);
}
template <class T, int NX, int TPC = 64>
__device__ inline void x_inv_run(cpx<T>* b0, cpx<T>* b1, const cpx<T>* tw, const XInvIn<T, NX, TPC>& P, int lane,
                                 T* qa, T* qb) {
  using P_t = XInvIn<T, NX, TPC>;
#pragma unroll
  for (int q = 0; q < XQ<NX, TPC>::Q; ++q) {  // slots of idle lanes stay defined
    qa[q] = 0;
    if (qb) qb[q] = 0;
  }
  __syncthreads();  // the previous transform's readers of b0/b1 are done
  auto in = [&](int bb, int r, int e) -> cpx<T> {
    const int i = bb * P_t::R0 + r, k = P_t::kk(lane, i);
    cpx<T> a = P.va[i], b = P.vb[i];
    if (k == 0 || 2 * k == NX) { a.y = 0; b.y = 0; }
    return 2 * e > NX ? mk<T>(a.x + b.y, -a.y + b.x)   // conj(a) + i conj(b)
                      : mk<T>(a.x - b.y, a.y + b.x);   // a + i b
  };
  auto out = [&](int b, int r, int, cpx<T> v) {
    const int q = b * XQ<NX, TPC>::RL + r;
    qa[q] = v.x;
    if (qb) qb[q] = v.y;
  };
  xfft<T, NX, true, false, TPC>(b0, b1, tw, lane, in, out);
}
template <class T, int NX, int TPC = 64, int MA = -1, int MB = -1>
__device__ inline void x_inv_pair(cpx<T>* b0, cpx<T>* b1, const cpx<T>* tw, const XIn* A, const XIn* B,
                                  size_t nlines, size_t line, bool lok, int Kx, double dkx, int lane, T* qa, T* qb) {
  XInvIn<T, NX, TPC> P;
  x_inv_load<T, NX, TPC, MA, MB>(A, B, nlines, line, lok, Kx, dkx, lane, P);
  x_inv_run<T, NX, TPC>(b0, b1, tw, P, lane, qa, qb);
}

// Forward x transform of two real lines packed as va + i vb, where
// val(q) = (va, vb) at the lane's q-th point; stores the first Kx
// coefficients of each, scaled by 1/NX, to outA / outB (compute type).
// Lanes of a line beyond the last (lok false) store nothing.
template <class T, int NX, int TPC, class F>
__device__ inline void x_fwd_pair(cpx<T>* b0, cpx<T>* b1, const cpx<T>* tw, F&& val, void* outA, void* outB,
                                  size_t line, bool lok, int Kx, int lane) {
  constexpr int NST = nstages_of(NX);
  cpx<T>* res = (NST - 1) % 2 ? b1 : b0;  // where the last stage stores
  __syncthreads();
  auto in = [&](int b, int r, int) -> cpx<T> { return val(b * XQ<NX, TPC>::RL + r); };
  auto out = [&](int, int, int e, cpx<T> v) { res[swz(e)] = v; };
  xfft<T, NX, false, true, TPC>(b0, b1, tw, lane, in, out);
  __syncthreads();
  const T sc = (T)(0.5 / NX);
  for (int k = lane; lok && k < Kx; k += TPC) {
    cpx<T> Z = res[swz(k)], Zc = conjc(res[swz((NX - k) % NX)]);
    gstc<T>(outA, line * Kx + k, scal(Z + Zc, sc));
    if (outB) {
      cpx<T> D = Z - Zc;
      gstc<T>(outB, line * Kx + k, mk<T>(D.y * sc, -D.x * sc));
    }
  }
}

// ---- x-line layout of the x-line kernels: LPW lines per 64-lane wave, TPC =
// 64 / LPW lanes per line, with the Stockham transforms above.  TPC is the
// smallest power of two that keeps a lane at no more than 4 points of its
// line (the 256-point case): every lane then works in every stage, where one
// line per wave would leave 3 of 4 lanes idle at NX = 64.  The kernels take
// the layout as a policy: the lane's line (line_in_wave) and position g in
// it, its points x(g, q) for q < PTS, the pair transforms and per-line
// reductions.
__host__ __device__ constexpr int x_tpc(int NX) {
  int t = 1;
  while (t < 64 && ((NX / radix_of(NX, nstages_of(NX) - 1)) + t - 1) / t * radix_of(NX, nstages_of(NX) - 1) > 4) t *= 2;
  return t;
}
// every Stockham stage of an NX-point line on TPC lanes has at most one
// butterfly per lane
template <int NX, int TPC>
__host__ __device__ constexpr bool x_one_butterfly_per_lane() {
  for (int s = 0; s < nstages_of(NX); ++s)
    if (NX / radix_of(NX, s) > TPC) return false;
  return true;
}
template <class T, int NX>
struct XWave {
  static constexpr int TPC = x_tpc(NX), LPW = 64 / TPC, PTS = XQ<NX, TPC>::Q;
  // One LDS buffer serves both sides of the Stockham exchange when each
  // stage has one butterfly per lane: the workgroup is one wave, a stage's
  // writes depend on all of its reads (they wait for the read data), and the
  // LDS executes a wave's accesses in order, so a stage may overwrite what it
  // read.  With two butterflies per lane the first one's writes would precede
  // the second one's reads, and the stages ping-pong between two buffers.
  // (A 256-point line: 4 KiB of LDS per wave instead of 6, which lets the
  // 66-VGPR phase B hold 7 waves per SIMD.)
  static constexpr bool ONEBUF = x_one_butterfly_per_lane<NX, TPC>();
  struct Smem {
    cpx<T> b0[LPW * NX], b1[ONEBUF ? 1 : LPW * NX], tw[NX];
  };
  __device__ static cpx<T>* buf1(Smem& s, int lw) { return ONEBUF ? s.b0 + lw * NX : s.b1 + lw * NX; }
  __device__ static void init(Smem& s, const cpx<T>* twg, int lane) {
    for (int j = lane; j < NX; j += 64) s.tw[swz(j)] = twg[j];
  }
  __device__ static int xpos(int g, int q) { return xown<NX, TPC>(g, q); }
  template <int MA = -1, int MB = -1>
  __device__ static void inv_pair(Smem& s, const XIn* A, const XIn* B, size_t nlines, size_t line, bool lok,
                                  int Kx, double dkx, int g, int lw, T* qa, T* qb) {
    x_inv_pair<T, NX, TPC, MA, MB>(s.b0 + lw * NX, buf1(s, lw), s.tw, A, B, nlines, line, lok, Kx, dkx, g, qa, qb);
  }
  using InvIn = XInvIn<T, NX, TPC>;
  template <int MA = -1, int MB = -1>
  __device__ static void inv_load(const XIn* A, const XIn* B, size_t nlines, size_t line, bool lok, int Kx,
                                  double dkx, int g, InvIn& P) {
    x_inv_load<T, NX, TPC, MA, MB>(A, B, nlines, line, lok, Kx, dkx, g, P);
  }
  __device__ static void inv_run(Smem& s, const InvIn& P, int g, int lw, T* qa, T* qb) {
    x_inv_run<T, NX, TPC>(s.b0 + lw * NX, buf1(s, lw), s.tw, P, g, qa, qb);
  }
  template <class F>
  __device__ static void fwd_pair(Smem& s, F&& val, void* oA, void* oB, size_t line, bool lok, int Kx, int g, int lw) {
    x_fwd_pair<T, NX, TPC>(s.b0 + lw * NX, buf1(s, lw), s.tw, val, oA, oB, line, lok, Kx, g);
  }
  // sums / maxima over the TPC lanes of a line (aligned groups of the wave)
  __device__ static double line_sum(double v) {
    for (int o = TPC / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
  }
  __device__ static double line_max(double v) {
    for (int o = TPC / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
  }
};

// ---------------------------------------------------------- physics args
enum OutA { OA_MU = 0, OA_U1, OA_U2, OA_U3, OA_F11, OA_F12, OA_F13, OA_F22, OA_F23, OA_F33 };
enum OutB { OB_H1 = 0, OB_H2, OB_H3, OB_ADV1 };  // ADV1..ADV6 = 3..8

// Arguments of the x-line kernels.  They are written once per configuration
// into device memory and the kernels read them through a pointer: a by-value
// kernel argument whose arrays are indexed at run time, or whose members have
// their address taken, is copied whole into per-lane scratch memory in every
// workgroup.
// Density table of the mixed electron functional: RT_N intervals over
// |n - n0| <= RT_R n0.  4-point Lagrange interpolation on it reproduces the
// per-point Fermi-Dirac evaluation (table, Halley and Newton) to about
// 1e-13 relative for alpha0 in [-8, 60]; points outside take that evaluation.
constexpr int RT_N = 2048;
constexpr double RT_R = 0.2;

struct PhysArgs {
  int nd, Kx, ny, nz;
  double dx, dy, dz;
  double dkx;  // 2 pi / lx: kx of column kk is kk * dkx
  double n0, mass, beta, cn, alpha0, cg;
  double inv_n0, inv_beta, inv_cn;  // reciprocals: the per-point functional multiplies
  double dadn, d2adn2;  // d alpha / d n, d2 alpha / d n2 at n0: start of the per-point Newton iteration
  // density table (mixed precision only, null otherwise): at x_i = -rt_r + i h
  // (x = (n - n0) / n0, i <= RT_N) the node holds {alpha - alpha0, xi,
  // d xi / d n, 0}; rt_in bounds |x| where the 4-point stencil is inside
  const double* rtab;
  double rt_r, rt_invh, rt_in;
  int ideal, tf, grad;
  const double* fdtab;
  Ctl* ctl;
  // phase A inputs
  XIn n, m[3], gx, gy, gz, lap;
  // phase B inputs
  XIn mux, muy, muz;
  XIn px[6], py[6], pz[6];
  int evolve;
  // external force
  int vext_kind, vext_td;
  double amp, ramp, omega, kvec[3];
  const double* fext;  // gaussian: 3 real fields
  const double* vsc;   // mode / driven: sin(kx x_i), then cos(kx x_i), i < nx
  double* red;         // per-line partial sums
  // outputs: phase A in the order of code[]; phase B: H_1..H_nd, then ADV_1..6
  int nout;
  int code[10];
  void* dst[10];
};

// Compile-time configuration of the x-line kernels.  XCfg<XF_RUNTIME> reads
// every choice from the argument block.  The other forms are for 3-D grids,
// whose x-line inputs are all compute-type slots with no k = 0 offset and a
// multiplier fixed by their role (XMul): XF_ANY3 fixes that layout (phase B
// of any model; phase A of the other functionals, FP32), XF_NOF3 also the
// absence of an external force (phase B without the force code: 102 against
// 73 VGPRs at 256 points in FP32); XF_IDEAL also fixes phase A of a classical
// ion model (the ideal term as the only pointwise functional; mu, u and the
// six fluxes emitted in the order the host lists them).  Fixing these at compile time
// removes the uniform branches on the inputs' types and on the output codes,
// with the register moves at their joins.  (A fixed form of the electron
// functional and output list was tried: its straight-line code needed more
// than the 128 VGPRs of 4 waves per SIMD and spilled; the electrons take
// XF_ANY3, which keeps both run-time.)
enum XFun { XF_RUNTIME = 0, XF_IDEAL = 1, XF_ANY3 = 2, XF_NOF3 = 3 };
template <int F>
struct XCfg {
  static constexpr bool D3 = F != XF_RUNTIME;
  static constexpr bool IDEAL = F == XF_IDEAL;
  static constexpr bool NOFORCE = F == XF_NOF3;
  static constexpr int MN = D3 ? XM_NONE : -1, MK = D3 ? XM_IK : -1;
};

// Common preamble of the x-line kernels: the lane's line and position in it.
#define HY_XLINE_PREAMBLE(ND3)                                                         \
  constexpr int PTS = X::PTS;                                                          \
  __shared__ typename X::Smem sm;                                                      \
  const PhysArgs& a = *pa;                                                             \
  const int lane = threadIdx.x;                                                        \
  const int lw = lane / X::TPC, g = lane % X::TPC;                                     \
  const size_t line = (size_t)xcd_work_index(blockIdx.x, gridDim.x) * X::LPW + lw;     \
  const size_t nlines = (size_t)a.ny * a.nz;                                           \
  const bool lok = line < nlines;                                                      \
  const int nd = (ND3) ? 3 : a.nd;                                                     \
  X::init(sm, twg, lane);

template <class T, int NX, class X, class C>
__device__ __attribute__((always_inline)) inline void phaseA_body(const PhysArgs* __restrict__ pa, const cpx<T>* __restrict__ twg) {
  constexpr bool FIX = C::IDEAL;
  HY_XLINE_PREAMBLE(C::D3)
  constexpr int MN = C::MN, MK = C::MK;
  const bool ideal = FIX || a.ideal != 0;
  const bool tf = !FIX && a.tf != 0;
  const bool grad = !FIX && a.grad != 0;
  // per-point values are held in the transform type T: in mixed precision
  // they come out of (and go back into) FP32 transforms; the pointwise
  // functional below is evaluated in FP64
  // Transform pairs ordered so that few fields are live at once: the
  // functional needs n - n0, |grad n|^2 and lap n; the outputs need mom.
  // The gradient path transforms three pairs before the functional; in the
  // 3-D forms each of the second and third pairs' loads is issued before the
  // previous pair is transformed, so their latency overlaps that work (the
  // run-time form issues the second pair's after, within the 96 VGPRs of 5
  // waves per SIMD in FP32; issuing the (m1, m2) loads before the functional
  // as well spills).
  T dn[PTS], m0[PTS], g2[PTS], lap[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) { m0[p] = 0; g2[p] = 0; lap[p] = 0; }
  if (grad) {
    typename X::InvIn P0, P1;
    X::template inv_load<MN, MN>(&a.n, &a.lap, nlines, line, lok, a.Kx, a.dkx, g, P0);
    if constexpr (C::D3)
      X::template inv_load<MK, MN>(&a.gx, nd >= 2 ? &a.gy : nullptr, nlines, line, lok, a.Kx, a.dkx, g, P1);
    X::inv_run(sm, P0, g, lw, dn, lap);
    if constexpr (!C::D3)
      X::template inv_load<MK, MN>(&a.gx, nd >= 2 ? &a.gy : nullptr, nlines, line, lok, a.Kx, a.dkx, g, P1);
    X::template inv_load<MN, MN>(nd >= 3 ? &a.gz : nullptr, &a.m[0], nlines, line, lok, a.Kx, a.dkx, g, P0);
    T gx[PTS], gy[PTS];
    X::inv_run(sm, P1, g, lw, gx, gy);
#pragma unroll
    for (int p = 0; p < PTS; ++p) g2[p] = gx[p] * gx[p] + gy[p] * gy[p];
    X::inv_run(sm, P0, g, lw, gx, m0);
#pragma unroll
    for (int p = 0; p < PTS; ++p) g2[p] += gx[p] * gx[p];
  } else {
    X::template inv_pair<MN, MN>(sm, &a.n, &a.m[0], nlines, line, lok, a.Kx, a.dkx, g, lw, dn, m0);
  }
  // pointwise functional in FP64
  T mu[PTS], inv[PTS];  // mu_loc, 1 / (m n)
  bool bad = false;
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    const bool live = lok && X::xpos(g, p) >= 0;
    const double n = a.n0 + (double)dn[p];
    if (live && !(n > 0.0)) bad = true;
    // divisions are reciprocal multiplies (1 ulp apart from the quotients)
    inv[p] = (T)fd::rcp_nr(a.mass * n);
    // mu_loc with its k = 0 constants (ln n0 / beta, alpha0 / beta) dropped:
    // only grad(mu) enters the momentum equation (model.f90:99).
    double v = ideal ? log1p((double)dn[p] * a.inv_n0) * a.inv_beta : 0.0;  // fe_ideal_classical.f90:40
    if (tf || grad) {
      const double x = (double)dn[p] * a.inv_n0;
      if (a.rtab && fabs(x) < a.rt_in) {
        // mixed precision: the node values around x, 4-point Lagrange weights
        const double u = (x + a.rt_r) * a.rt_invh;
        const int i = (int)u - 1;  // u >= 2 here, so the stencil is i..i+3, inside the table
        const double t = u - (double)(i + 1);
        const double tm1 = t - 1.0, tm2 = t - 2.0, tp1 = t + 1.0;
        const double w0 = -t * tm1 * tm2 * (1.0 / 6.0), w1 = tp1 * tm1 * tm2 * 0.5;
        const double w2 = -tp1 * t * tm2 * 0.5, w3 = tp1 * t * tm1 * (1.0 / 6.0);
        double r[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double w = k == 0 ? w0 : (k == 1 ? w1 : (k == 2 ? w2 : w3));
          const double* q = a.rtab + 4 * (i + k);
          r[0] += w * gld<double>(q, 0);
          r[1] += w * gld<double>(q, 1);
          r[2] += w * gld<double>(q, 2);
        }
        if (tf) v += r[0] * a.inv_beta;
        if (grad) v -= a.cg * (r[2] * (double)g2[p] + 2.0 * r[1] * (double)lap[p]);
      } else {
        const double al = live && n > 0.0
                              ? fd::alpha_from_ip12_dev(a.fdtab, n * a.inv_cn,
                                                        a.alpha0 + (double)dn[p] * (a.dadn + 0.5 * a.d2adn2 * (double)dn[p]))
                              : a.alpha0;
        if (tf) v += (al - a.alpha0) * a.inv_beta;  // fe_kinetic_tfk.f90:80-81
        if (grad) {                                 // fe_kinetic_tfk.f90:130-137
          double im, d1, d2;
          fd::im12_d1_d2_dev(a.fdtab, al, im, d1, d2);
          const double rim = fd::rcp_nr(im);
          const double xi = d1 * rim * rim;
          const double dxi = (d2 * im - 2.0 * d1 * d1) * (rim * rim * rim);
          const double dxidn = dxi * rim * (2.0 * a.inv_cn);
          v -= a.cg * (dxidn * (double)g2[p] + 2.0 * xi * (double)lap[p]);
        }
      }
    }
    mu[p] = (T)v;
    }
  if (bad) a.ctl->nonpos = 1;
  T m1[PTS], m2[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) { m1[p] = 0; m2[p] = 0; }
  if (nd >= 2)
    X::template inv_pair<MN, MN>(sm, &a.m[1], nd >= 3 ? &a.m[2] : nullptr, nlines, line, lok, a.Kx, a.dkx, g, lw,
                                 m1, m2);
  // outputs, one uniform switch per output (not per point)
  auto fill = [&](int c, T* v) {
#pragma unroll
    for (int p = 0; p < PTS; ++p) v[p] = 0;
    // every case names its arrays directly: selecting an array through a
    // run-time pointer would move the arrays to scratch memory
#define HY_FILL(EXPR) for (int p = 0; p < PTS; ++p) v[p] = (T)(EXPR)
    switch (c) {
      case OA_MU: HY_FILL(mu[p]); break;
      case OA_U1: HY_FILL(m0[p] * inv[p]); break;
      case OA_U2: HY_FILL(m1[p] * inv[p]); break;
      case OA_U3: HY_FILL(m2[p] * inv[p]); break;
      case OA_F11: HY_FILL(m0[p] * (m0[p] * inv[p])); break;
      case OA_F12: HY_FILL(m0[p] * (m1[p] * inv[p])); break;
      case OA_F13: HY_FILL(m0[p] * (m2[p] * inv[p])); break;
      case OA_F22: HY_FILL(m1[p] * (m1[p] * inv[p])); break;
      case OA_F23: HY_FILL(m1[p] * (m2[p] * inv[p])); break;
      case OA_F33: HY_FILL(m2[p] * (m2[p] * inv[p])); break;
#undef HY_FILL
      default: break;
    }
  };
  if constexpr (FIX) {
    // the host lists mu, u1, u2, u3, F11, F12, F13, F22, F23, F33
    auto pair = [&](int o, int ca, int cb) {
      T va[PTS], vb[PTS];
      fill(ca, va);
      fill(cb, vb);
      X::fwd_pair(sm, [&](int q) { return mk<T>(va[q], vb[q]); }, a.dst[o], a.dst[o + 1], line, lok, a.Kx, g, lw);
    };
    pair(0, OA_MU, OA_U1);
    pair(2, OA_U2, OA_U3);
    pair(4, OA_F11, OA_F12);
    pair(6, OA_F13, OA_F22);
    pair(8, OA_F23, OA_F33);
  } else {
    for (int o = 0; o < a.nout; o += 2) {
      const bool two = o + 1 < a.nout;
      T va[PTS], vb[PTS];
      fill(a.code[o], va);
      fill(two ? a.code[o + 1] : -1, vb);
      X::fwd_pair(sm, [&](int q) { return mk<T>(va[q], vb[q]); }, a.dst[o], two ? a.dst[o + 1] : nullptr, line,
                  lok, a.Kx, g, lw);
    }
  }
}

template <class T, int NX, class X, bool MX, class C>
__device__ __attribute__((always_inline)) inline void phaseB_body(const PhysArgs* __restrict__ pa, const cpx<T>* __restrict__ twg) {
  HY_XLINE_PREAMBLE(C::D3)
  constexpr int MN = C::MN, MK = C::MK;
  const int iy = lok ? (int)(line % a.ny) : 0, iz = lok ? (int)(line / a.ny) : 0;
  T dn[PTS], gmx[PTS], gmy[PTS], gmz[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) { gmy[p] = 0; gmz[p] = 0; }
  {  // the second pair's loads are issued before the first pair is transformed
    typename X::InvIn P0, P1;
    X::template inv_load<MN, MK>(&a.n, &a.mux, nlines, line, lok, a.Kx, a.dkx, g, P0);
    if (nd >= 2)
      X::template inv_load<MN, MN>(&a.muy, nd >= 3 ? &a.muz : nullptr, nlines, line, lok, a.Kx, a.dkx, g, P1);
    X::inv_run(sm, P0, g, lw, dn, gmx);
    if (nd >= 2) X::inv_run(sm, P1, g, lw, gmy, gmz);
  }
  // external force (external_potential.f90:65-93) at the stage-start time,
  // evaluated point by point.  The wave force amp k sin(kx x + ph), with the
  // line's phase ph = ky y + kz z - w t, is formed as
  // amp k (sin(kx x) cos(ph) + cos(kx x) sin(ph)) from the model's table of
  // sin(kx x), cos(kx x): one sincos per lane instead of one sin per point.
  const bool wave_f = !C::NOFORCE && (a.vext_kind == HY_VEXT_MODE || a.vext_kind == HY_VEXT_DRIVEN);
  const bool gauss_f = !C::NOFORCE && a.vext_kind == HY_VEXT_GAUSSIAN && lok;
  double amp = 0.0, sph = 0.0, cph = 0.0;
  if (wave_f) {
    const double t = a.vext_td ? a.ctl->ts : 0.0;
    amp = a.amp;
    if (a.vext_td && a.ramp > 0.0) amp = amp * fmin(1.0, t / a.ramp);
    sincos(a.kvec[1] * (iy * a.dy) + a.kvec[2] * (iz * a.dz) - a.omega * t, &sph, &cph);
  }
  const size_t nr = (size_t)NX * a.ny * a.nz;
  const double* fe = a.fext + (lok ? line : 0) * NX;
  T h0[PTS], h1[PTS], h2[PTS];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    const int x = X::xpos(g, p);
    const bool live = lok && x >= 0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if (wave_f) {
      const int xi = x >= 0 ? x : 0;
      const double sv = amp * (gld<double>(a.vsc, xi) * cph + gld<double>(a.vsc, NX + xi) * sph);
      f0 = sv * a.kvec[0]; f1 = sv * a.kvec[1]; f2 = sv * a.kvec[2];
    } else if (gauss_f && x >= 0) {
      f0 = fe[x]; f1 = fe[nr + x]; f2 = fe[2 * nr + x];
    }
    const double n = a.n0 + dn[p];
    h0[p] = (T)(n * (f0 - gmx[p]));
    h1[p] = (T)(n * (f1 - gmy[p]));
    h2[p] = (T)(n * (f2 - gmz[p]));
    if (live) { s0 += n * f0; s1 += n * f1; s2 += n * f2; }
  }
  if (!C::NOFORCE && a.vext_kind != HY_VEXT_NONE) {
    s0 = X::line_sum(s0); s1 = X::line_sum(s1); s2 = X::line_sum(s2);
    if (g == 0 && lok) { a.red[line * 3] = s0; a.red[line * 3 + 1] = s1; a.red[line * 3 + 2] = s2; }
  }
  // H_1, H_2 share a transform (H_2 = 0 in 1-D and is not stored); H_3 goes alone
  X::fwd_pair(sm, [&](int q) { return mk<T>(h0[q], h1[q]); }, a.dst[0], nd >= 2 ? a.dst[1] : nullptr, line, lok,
              a.Kx, g, lw);
  if (nd >= 3)
    X::fwd_pair(sm, [&](int q) { return mk<T>(h2[q], T(0)); }, a.dst[2], nullptr, line, lok, a.Kx, g, lw);
  if constexpr (MX) {
  if (!a.evolve) return;
  // Maxwell stress advection dPi_c = -sum_d u_d d_d Pi_c (model.f90:126-134),
  // two components per pass.  PF: each pair's loads are issued before the
  // previous pair is transformed (FP32; the FP64 form, at 2 waves per SIMD
  // either way, runs slower so)
  constexpr bool PF = sizeof(T) == 4;
  T u0[PTS], u1[PTS], u2[PTS];
  {
    T m0[PTS], m1[PTS], m2[PTS];
#pragma unroll
    for (int p = 0; p < PTS; ++p) { m1[p] = 0; m2[p] = 0; }
    typename X::InvIn Q0, Q1;
    X::template inv_load<MN, MN>(&a.m[0], nd >= 2 ? &a.m[1] : nullptr, nlines, line, lok, a.Kx, a.dkx, g, Q0);
    if (PF && nd >= 3) X::template inv_load<MN, MN>(&a.m[2], nullptr, nlines, line, lok, a.Kx, a.dkx, g, Q1);
    X::inv_run(sm, Q0, g, lw, m0, m1);
    if (!PF && nd >= 3) X::template inv_load<MN, MN>(&a.m[2], nullptr, nlines, line, lok, a.Kx, a.dkx, g, Q1);
    if (nd >= 3) X::inv_run(sm, Q1, g, lw, m2, nullptr);
#pragma unroll
    for (int p = 0; p < PTS; ++p) {
      const double inv = fd::rcp_nr(a.mass * (a.n0 + dn[p]));
      u0[p] = m0[p] * inv; u1[p] = nd >= 2 ? m1[p] * inv : 0.0; u2[p] = m2[p] * inv;
    }
  }
  for (int c = 0; c < 6; c += 2) {
    T xa[PTS], ya[PTS], za[PTS], xb[PTS], yb[PTS], zb[PTS];
#pragma unroll
    for (int p = 0; p < PTS; ++p) { ya[p] = 0; za[p] = 0; yb[p] = 0; zb[p] = 0; }
    if (nd == 1) {
      X::inv_pair(sm, &a.px[c], &a.px[c + 1], nlines, line, lok, a.Kx, a.dkx, g, lw, xa, xb);
    } else {
      typename X::InvIn Q0, Q1;
      X::template inv_load<MK, MN>(&a.px[c], &a.py[c], nlines, line, lok, a.Kx, a.dkx, g, Q0);
      if (PF) X::template inv_load<MK, MN>(&a.px[c + 1], &a.py[c + 1], nlines, line, lok, a.Kx, a.dkx, g, Q1);
      X::inv_run(sm, Q0, g, lw, xa, ya);
      if (!PF) X::template inv_load<MK, MN>(&a.px[c + 1], &a.py[c + 1], nlines, line, lok, a.Kx, a.dkx, g, Q1);
      if (PF && nd >= 3)
        X::template inv_load<MN, MN>(&a.pz[c], &a.pz[c + 1], nlines, line, lok, a.Kx, a.dkx, g, Q0);
      X::inv_run(sm, Q1, g, lw, xb, yb);
      if (!PF && nd >= 3)
        X::template inv_load<MN, MN>(&a.pz[c], &a.pz[c + 1], nlines, line, lok, a.Kx, a.dkx, g, Q0);
      if (nd >= 3) X::inv_run(sm, Q0, g, lw, za, zb);
    }
    X::fwd_pair(
        sm,
        [&](int p) {
          const double va = -(u0[p] * xa[p]) - u1[p] * ya[p] - u2[p] * za[p];
          const double vb = -(u0[p] * xb[p]) - u1[p] * yb[p] - u2[p] * zb[p];
          return mk<T>((T)va, (T)vb);
        },
        a.dst[nd + c], a.dst[nd + c + 1], line, lok, a.Kx, g, lw);
  }
  }
}

// max |u_d| over a line (model.f90:171-172)
template <class T, int NX, class X>
__device__ __attribute__((always_inline)) inline void umax_body(const PhysArgs* __restrict__ pa, const cpx<T>* __restrict__ twg) {
  HY_XLINE_PREAMBLE(false)
  T dn[PTS], m0[PTS], m1[PTS], m2[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) { m1[p] = 0; m2[p] = 0; }
  X::inv_pair(sm, &a.n, &a.m[0], nlines, line, lok, a.Kx, a.dkx, g, lw, dn, m0);
  if (nd >= 2) X::inv_pair(sm, &a.m[1], nd >= 3 ? &a.m[2] : nullptr, nlines, line, lok, a.Kx, a.dkx, g, lw, m1, m2);
  double um = 0.0;
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    if (!lok || X::xpos(g, p) < 0) continue;
    const double inv = fd::rcp_nr(a.mass * (a.n0 + dn[p]));
    um = fmax(um, fabs(m0[p] * inv));
    if (nd >= 2) um = fmax(um, fabs(m1[p] * inv));
    if (nd >= 3) um = fmax(um, fabs(m2[p] * inv));
  }
  um = X::line_max(um);
  if (g == 0 && lok) a.red[line] = um;
}
#undef HY_XLINE_PREAMBLE

// The x kernels take the argument block pointer as a kernel parameter (one
// model) or the base of a device array of blocks (a batch: model blockIdx.z).
// A pointer that arrives inside a struct or is read from memory would lose
// what the compiler knows about kernel parameters, and every field of the
// block would then be loaded per lane instead of once per wave.
template <class T, int NX, class X, class C>
__global__ __launch_bounds__(64, sizeof(T) == 4 ? 5 : 1) void k_phaseA(const PhysArgs* __restrict__ pa, const cpx<T>* __restrict__ tw) {
  phaseA_body<T, NX, X, C>(pa, tw);
}
template <class T, int NX, class X, class C>
__global__ __launch_bounds__(64, sizeof(T) == 4 ? 5 : 1) void k_phaseA_b(const PhysArgs* __restrict__ pas, const cpx<T>* __restrict__ tw) {
  if (gate_closed(&pas[blockIdx.z].ctl->halt)) return;
  phaseA_body<T, NX, X, C>(pas + blockIdx.z, tw);
}
template <class T, int NX, class X, bool MX, class C>
__global__ __launch_bounds__(64) void k_phaseB(const PhysArgs* __restrict__ pa, const cpx<T>* __restrict__ tw) {
  phaseB_body<T, NX, X, MX, C>(pa, tw);
}
template <class T, int NX, class X, bool MX, class C>
__global__ __launch_bounds__(64) void k_phaseB_b(const PhysArgs* __restrict__ pas, const cpx<T>* __restrict__ tw) {
  if (gate_closed(&pas[blockIdx.z].ctl->halt)) return;
  phaseB_body<T, NX, X, MX, C>(pas + blockIdx.z, tw);
}
template <class T, int NX, class X>
__global__ __launch_bounds__(64) void k_umax(const PhysArgs* __restrict__ pa, const cpx<T>* __restrict__ tw) {
  umax_body<T, NX, X>(pa, tw);
}
template <class T, int NX, class X>
__global__ __launch_bounds__(64) void k_umax_b(const PhysArgs* __restrict__ pas, const cpx<T>* __restrict__ tw) {
  if (gate_closed(&pas[blockIdx.z].ctl->halt)) return;
  umax_body<T, NX, X>(pas + blockIdx.z, tw);
}

// compute_dt from the per-line maxima (model.f90:169-181)
struct DtK {
  const double* red;
  int nlines;
  Ctl* c;
  double wmax, numax, dxmin, cfl;
};
__device__ __attribute__((always_inline)) inline void dt_body(const DtK& k) {
  const double* __restrict__ red = k.red;
  const int nlines = k.nlines;
  Ctl* c = k.c;
  const double wmax = k.wmax, numax = k.numax, dxmin = k.dxmin, cfl = k.cfl;
  if (gate_closed(&c->halt)) return;  // uniform over the block
  __shared__ double sh[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nlines; i += blockDim.x) m = fmax(m, red[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double umax = sh[0];
    const double huge = 1.7976931348623157e308;
    double dtm = huge;
    if (wmax > 0.0) dtm = fmin(dtm, 1.7 / wmax);
    if (numax > 0.0) dtm = fmin(dtm, 2.5 / numax);
    if (umax > 0.0) dtm = fmin(dtm, dxmin / umax);
    c->umax = umax;
    if (dtm == huge) c->nodt = 1;
    else c->dt = cfl * dtm;
  }
}
__global__ void k_dt(DtK k) { dt_body(k); }
__global__ void k_dt_b(const DtK* __restrict__ ks) { dt_body(ks[blockIdx.z]); }

// sum over lines of `ncomp` partials, scaled: out[c] = scale * sum_l red[l*ncomp + c]

// ------------------------------------------------------------ spectral kernels
struct SpecArgs {
  int Kx, Ky, Kz, My, Mz, ny, nz, nd;
  double dkx, dky, dkz;  // 2 pi / L_d (0 for an absent axis): k_d = signed mode index * dk_d
  size_t nt;   // Kx*Ky*Kz (< 2^31 for every supported grid)
};
// Wavenumbers of kept-box entry e (grid.f90:77-88: k = 2 pi wrap(m) / L).
// 32-bit index arithmetic and no table loads: a 64-bit division by a run-time
// value costs about a hundred instructions, a generic-address table load a
// full wait.
__device__ inline void spec_k(const SpecArgs& a, size_t e, double k[3]) {
  const unsigned u = (unsigned)e, kxn = (unsigned)a.Kx, kyn = (unsigned)a.Ky;
  const unsigned q = u / kxn;
  const int ix = (int)(u - q * kxn);
  const unsigned tz = q / kyn;
  const int ty = (int)(q - tz * kyn);
  const int sy = ty <= a.My ? ty : ty - a.Ky;
  const int sz = (int)tz <= a.Mz ? (int)tz : (int)tz - a.Kz;
  k[0] = ix * a.dkx;
  k[1] = sy * a.dky;
  k[2] = sz * a.dkz;
}
// K1/K2 field accesses go through gld/gst (global, not flat, in the batched
// kernels, whose pointers come from a device array)
template <class T>
__device__ inline c64 ld(const void* p, size_t i) {
  return cvt<double>(gldc<T>(p, i));
}

// Newtonian stress Pi0 from u_hat (stress.f90:76-82) evaluated spectrally
__device__ inline void pi0_hat(const c64* u, const double* k, double eta, double xi, c64* out) {
  auto g = [&](int i, int j) { return mk<double>(-k[j] * u[i].y, k[j] * u[i].x); };
  c64 divu = g(0, 0) + g(1, 1) + g(2, 2);
  for (int c = 0; c < 3; ++c)
    out[c] = scal(scal(g(c, c), 2.0) - scal(divu, 2.0 / 3.0), eta) + scal(divu, xi);
  out[3] = scal(g(0, 1) + g(1, 0), eta);
  out[4] = scal(g(0, 2) + g(2, 0), eta);
  out[5] = scal(g(1, 2) + g(2, 1), eta);
}

struct K1Args {
  SpecArgs s;
  const void* fmu;        // F(mu_loc) (compute type)
  const c64* N;           // state n_hat
  const double* kcomb;    // hartree v(k) - ry c_cor(k)/beta on the kept box
  double n0;
  void* muhat;            // out (compute type)
  // Maxwell
  int evolve, relax;
  c64* P;                 // state Pi_hat (6)
  const void* fu[3];      // F(u) of this stage (compute type)
  void* uold;             // u_hat at the previous stage start (3, transform type: F(u) is stored in it)
  double eta, xi;
  const Ctl* ctl;
};

template <class T>
__device__ __attribute__((always_inline)) inline void spec1_body(const K1Args& a) {
  const size_t nt = a.s.nt;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nt; e += (size_t)gridDim.x * blockDim.x) {
    c64 nh = gldc<double>(a.N, e);
    if (e == 0) nh.x -= a.n0;  // fe_correlation_ry.f90:78
    c64 mu = ld<T>(a.fmu, e) + scal(nh, gld<double>(a.kcomb, e));
    if (e == 0) mu = mk<double>(0, 0);  // only grad(mu) is used
    gstc<T>(a.muhat, e, cvt<T>(mu));
    if (a.evolve) {
      double k[3];
      spec_k(a.s, e, k);
      c64 un[3];
      for (int d = 0; d < 3; ++d) un[d] = d < a.s.nd ? ld<T>(a.fu[d], e) : mk<double>(0, 0);
      if (a.relax) {  // model.f90:209-214, applied with this stage's u = u after the update
        c64 uo[3] = {cvt<double>(gldc<T>(a.uold, e)), cvt<double>(gldc<T>(a.uold, nt + e)),
                     cvt<double>(gldc<T>(a.uold, 2 * nt + e))};
        c64 p0n[6], p0o[6];
        pi0_hat(un, k, a.eta, a.xi, p0n);
        pi0_hat(uo, k, a.eta, a.xi, p0o);
        const double fac = a.ctl->relax_fac, phi = a.ctl->relax_phi;
        for (int c = 0; c < 6; ++c) {
          c64 pc = gldc<double>(a.P, c * nt + e);
          gstc<double>(a.P, c * nt + e, p0n[c] - scal(p0n[c] - p0o[c], phi) + scal(pc - p0o[c], fac));
        }
      }
      for (int d = 0; d < 3; ++d) gstc<T>(a.uold, d * nt + e, cvt<T>(un[d]));  // exact: un is F(u) in T
    }
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_spec1(K1Args a) { spec1_body<T>(a); }
template <class T>
__global__ __launch_bounds__(256) void k_spec1_b(const K1Args* __restrict__ as) {
  if (gate_closed(&as[blockIdx.z].ctl->halt)) return;  // a run past its step limit in a batch
  spec1_body<T>(as[blockIdx.z]);
}

struct K2Args {
  SpecArgs s;
  c64* S;                 // state: n, mom(3), Pi(6) (FP64)
  void* R;                // R_old (same layout, transform type T), or the FP64 output in rhs mode
  const void* fl[6];      // F(flux_ij): 11 12 13 22 23 33
  const void* fh[3];      // F(H_i)
  const void* fu[3];      // F(u) (Newtonian stress)
  const void* fadv[6];    // F(adv_c) (Maxwell)
  double mass, inv_mass, eta, xi;
  int stress_newt, evolve, ext_active;
  const double* red;      // external force: per-line partial sums of n f_i (3 per line, phase B)
  int nlines;             // their line count
  double red_scale;       // 1 / (number of grid points)
  int stage;              // 0..2
  int rhs_only;
  int end_ops;            // 1: end the stage; 2: also end the step (first thread of the first workgroup)
  double tau;             // Maxwell relaxation time (the stage end's factors)
  Ctl* ctl;
};

// ND is a template parameter so that every loop below unrolls and every
// array index is a constant: a run-time index into the by-value K2Args (or
// into a local array) would put them in per-lane scratch memory.
template <class T, int ND>
__device__ __attribute__((always_inline)) inline void spec2_body(const K2Args& a) {
  const size_t nt = a.s.nt;
  constexpr int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  constexpr int idx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};  // stress.f90:92
  const double dt = a.ctl->dt;
  const double ra = c_rk_a[a.stage], rb = c_rk_b[a.stage];
  if (a.end_ops && blockIdx.x == 0 && threadIdx.x == 0 && !a.ctl->halt) {
    ctl_stage_end(a.ctl, a.stage, a.tau, a.evolve);
    if (a.end_ops == 2) ctl_step_end(a.ctl);
  }
  // mean(n f_ext) of this stage (the k = 0 momentum RHS, model.f90:119-122):
  // the first workgroup sums phase B's per-line partial sums itself before it
  // handles mode 0, which only its first thread does
  __shared__ double nf[3];
  if (a.ext_active && blockIdx.x == 0) {
    __shared__ double sh[256];
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int i = threadIdx.x; i < a.nlines; i += blockDim.x) v += a.red[(size_t)i * 3 + c];
      sh[threadIdx.x] = v;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) nf[c] = sh[0] * a.red_scale;
      __syncthreads();
    }
  }
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nt; e += (size_t)gridDim.x * blockDim.x) {
    double k[3];
    spec_k(a.s, e, k);
    // the state of this mode, loaded once (the update below stores S[0]
    // before it reads S[1..3]; reloading them after that store re-reads HBM)
    c64 sv0[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) sv0[f] = gldc<double>(a.S, f * nt + e);
    // continuity: dn = -div(mom)/m (model.f90:95-96)
    c64 km = mk<double>(0, 0);
#pragma unroll
    for (int d = 0; d < ND; ++d) km = km + scal(sv0[1 + d], k[d]);
    c64 r[10];
    r[0] = mk<double>(km.y * a.inv_mass, -km.x * a.inv_mass);
    // stress divergence f_i = d_j Pi_ij
    c64 pih[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) pih[c] = mk<double>(0, 0);
    const bool have_pi = a.evolve || a.stress_newt;
    if (a.evolve) {
#pragma unroll
      for (int c = 0; c < 6; ++c) pih[c] = gldc<double>(a.S, (4 + c) * nt + e);
    } else if (a.stress_newt) {
      c64 u[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) u[d] = d < ND ? ld<T>(a.fu[d], e) : mk<double>(0, 0);
      pi0_hat(u, k, a.eta, a.xi, pih);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      c64 acc = mk<double>(0, 0);
      if (i < ND) {
        acc = ld<T>(a.fh[i], e);  // F(n (f_ext - d_i mu))
#pragma unroll
        for (int j = 0; j < ND; ++j) {  // - i k_j F(m_i u_j)
          const c64 fv = ld<T>(a.fl[sym[i][j]], e);
          acc.x += k[j] * fv.y;
          acc.y -= k[j] * fv.x;
        }
        if (have_pi) {
#pragma unroll
          for (int j = 0; j < ND; ++j) {
            const c64 pv = pih[idx[i][j]];
            acc.x -= k[j] * pv.y;
            acc.y += k[j] * pv.x;
          }
        }
        // k = 0: only the external force survives (the internal force has its mean
        // removed, model.f90:119-122)
        if (e == 0) acc = mk<double>(a.ext_active ? nf[i] : 0.0, 0.0);
      }
      r[1 + i] = acc;
    }
    if (a.evolve) {
#pragma unroll
      for (int c = 0; c < 6; ++c) r[4 + c] = ld<T>(a.fadv[c], e);
    }
    // R_old is kept in the transform type: in mixed precision the RHS is
    // formed from FP32 spectra (the continuity term excepted) and is
    // FP32-accurate already; the update itself is FP64.
    auto put = [&](int f, c64 rf) {
      if (a.rhs_only) {
        gstc<double>(a.R, f * nt + e, rf);
        return;
      }
      c64 sv = f < 4 ? sv0[f] : gldc<double>(a.S, f * nt + e);
      if (a.stage == 0) {
        sv = sv + scal(rf, dt * ra);
      } else {
        const c64 ro = cvt<double>(gldc<T>(a.R, f * nt + e));
        sv = sv + scal(scal(rf, ra) + scal(ro, rb), dt);
      }
      gstc<double>(a.S, f * nt + e, sv);
      gstc<T>(a.R, f * nt + e, cvt<T>(rf));
    };
#pragma unroll
    for (int f = 0; f < 4; ++f) put(f, r[f]);
    if (a.evolve) {
#pragma unroll
      for (int f = 4; f < 10; ++f) put(f, r[f]);
    }
  }
}
template <class T, int ND>
__global__ __launch_bounds__(256) void k_spec2(K2Args a) { spec2_body<T, ND>(a); }
template <class T, int ND>
__global__ __launch_bounds__(256) void k_spec2_b(const K2Args* __restrict__ as) {
  if (gate_closed(&as[blockIdx.z].ctl->halt)) return;
  spec2_body<T, ND>(as[blockIdx.z]);
}

// probe: spec_mode of n/n0 and mom_x (diagnostics.f90:70-77) after a step
__global__ void k_probe(const Ctl* c, const c64* S, size_t nt, long long idx, double n0, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  c64 a = idx >= 0 ? S[idx] : mk<double>(0, 0);
  c64 b = idx >= 0 ? S[nt + idx] : mk<double>(0, 0);
  out[0] = c->t;
  out[1] = a.x / n0; out[2] = a.y / n0;
  out[3] = b.x; out[4] = b.y;
  out[5] = c->dt;
}

// the same row into a ring of `ring` rows at slot (step - 1) % ring, where
// step is the device step count after the step (captured steps: every replay
// appends its rows)
struct ProbeK {
  const Ctl* c;
  const c64* S;
  size_t nt;
  long long idx;
  double n0;
  double* out;
  int ring;
};
__device__ __attribute__((always_inline)) inline void probe_ring_body(const ProbeK& k) {
  const Ctl* c = k.c;
  const c64* S = k.S;
  const size_t nt = k.nt;
  const long long idx = k.idx;
  const double n0 = k.n0;
  double* out = k.out;
  const int ring = k.ring;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->halt) return;
  const int slot = (int)(((long long)c->step - 1) % ring + ring) % ring;
  c64 a = idx >= 0 ? S[idx] : mk<double>(0, 0);
  c64 b = idx >= 0 ? S[nt + idx] : mk<double>(0, 0);
  double* o = out + 6 * (size_t)slot;
  o[0] = c->t;
  o[1] = a.x / n0; o[2] = a.y / n0;
  o[3] = b.x; o[4] = b.y;
  o[5] = c->dt;
}
__global__ void k_probe_ring(ProbeK k) { probe_ring_body(k); }
__global__ void k_probe_ring_b(const ProbeK* __restrict__ ks) { probe_ring_body(ks[blockIdx.z]); }

// full half spectrum <-> kept box
__global__ void k_full_to_trunc(const c64* __restrict__ full, c64* __restrict__ tr, SpecArgs a, int nkx) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.nt; e += (size_t)gridDim.x * blockDim.x) {
    int ix = (int)(e % a.Kx), ty = (int)((e / a.Kx) % a.Ky), tz = (int)(e / ((size_t)a.Kx * a.Ky));
    int jy = ty <= a.My ? ty : ty - a.Ky + a.ny;
    int jz = tz <= a.Mz ? tz : tz - a.Kz + a.nz;
    tr[e] = full[((size_t)jz * a.ny + jy) * nkx + ix];
  }
}
__global__ void k_trunc_to_full(const c64* __restrict__ tr, c64* __restrict__ full, SpecArgs a, int nkx) {
  size_t n = (size_t)nkx * a.ny * a.nz;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(e % nkx), j = (int)((e / nkx) % a.ny), k = (int)(e / ((size_t)nkx * a.ny));
    int my = j <= a.ny / 2 ? j : j - a.ny, mz = k <= a.nz / 2 ? k : k - a.nz;
    c64 v = mk<double>(0, 0);
    if (i < a.Kx && abs(my) <= a.My && abs(mz) <= a.Mz) {
      int ty = my >= 0 ? my : my + a.Ky, tz = mz >= 0 ? mz : mz + a.Kz;
      v = tr[((size_t)tz * a.Ky + ty) * a.Kx + i];
    }
    full[e] = v;
  }
}

// ------------------------------------------------------------ the model
struct Model {
  Grid g;
  int prec = HY_PREC_FP64;
  double mass = 1, n0 = 0;
  int ideal = 0, tf = 0, grad = 0, hartree = 0, ry = 0;
  double beta = 3, cn = 0, alpha0 = 0, cg = 0, e2 = 1.0 / 3.0, kappa = 0;
  double dadn = 0;  // 1 / (dn/dalpha) at alpha0: 2 / (cn I_{-1/2}(alpha0))
  double d2adn2 = 0;  // d2 alpha / dn2 at alpha0: -dadn^2 I'_{-1/2} / I_{-1/2}
  std::vector<hy_term_desc> terms;
  std::vector<std::vector<double>> ktabs, ctabs;
  int closure = 0, evolve = 0, stress_newt = 0;
  double eta = 0, xi = 0, tau = 0, eta_l = 0;
  int vext_kind = 0, vext_td = 0;
  double vamp = 0, vomega = 0, vwidth = 1, vramp = 0, kvec[3] = {0, 0, 0};
  int vmode[3] = {1, 0, 0};
  double dt_fixed = 0, cfl = 0.4;
  double wmax = 0, numax = 0, dxmin = 0;
  // device
  c64* S = nullptr;     // 10 * nt
  c64* R = nullptr;     // 10 * nt
  c64* uold = nullptr;  // 3 * nt
  double* kcomb = nullptr;
  double* fext = nullptr;
  double* vsc = nullptr;  // wave force: sin, cos of kx x at the nx grid points
  double* rtab = nullptr;  // mixed electrons: the functional's factors on a density grid
  double* fdtab = nullptr;
  void* xb = nullptr;   // x-line slots (compute type), each Kx*ny*nz
  void* zb = nullptr;   // z-pass outputs, each Kx*Ky*nz
  void* fb = nullptr;   // truncated spectra, each nt
  double* red = nullptr;
  Ctl* ctl = nullptr;
  PhysArgs* d_xargs = nullptr;  // XA_N device copies of the x-kernel arguments
  PhysArgs* p_xargs = nullptr;  // pinned host staging of the same (source of the uploads)
  // transfer buffers of set_state / get_state (a real field and its full
  // half spectrum), allocated on first use and kept: freeing device memory
  // waits for all work on the device, which would stall other models' streams
  double* d_real = nullptr;
  c64* d_full = nullptr;
  bool xargs_ok[4] = {false, false, false, false};
  // captured RK3 steps (hy_model_graph_*)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_nsteps = 0, g_probe = 0, g_ring = 0;
  // probe ring of captured and batched steps: its address is baked into the
  // graph and into every batch built on the model, so it only grows while
  // neither holds it (a grown ring drops the graph; a batch refuses)
  double* probe = nullptr;
  int probe_cap = 0;
  int nbatch = 0;       // live batches that hold this model's arguments
  // probe rows of hy_model_advance: a buffer of its own, so that an advance
  // never moves the ring a graph or a batch writes into
  double* aprobe = nullptr;
  int aprobe_cap = 0;
  // work buffers of the real-space diagnostics (mu, budgets), kept
  struct Work { void* p; size_t bytes; };
  std::vector<Work> work;
  double* full_kh = nullptr;  // hartree v(k) on the full half spectrum
  double* full_kr = nullptr;  // ry -c_cor(k)/beta on the full half spectrum
  int nxslots = 0, nzslots = 0, nfslots = 0;
  size_t xsz = 0, zsz = 0, nt = 0;
  size_t esz = 16;      // compute element size
  hipStream_t st = nullptr;
  int relax_pending = 0;
  int irecompute = 0;
  // host mirror of the time
  double h_t = 0, h_dt = 0;
  int h_step = 0;
  // per-kernel-class profiling (HIP events on the model stream): bit c of
  // prof set records the launches of class c; prof_open marks a begin event
  // whose end is still to come
  unsigned prof = 0;
  bool prof_open = false;
  struct Ev { int cls; hipEvent_t a, b; double bytes; };
  std::vector<Ev> evs;
  std::vector<hipEvent_t> pool;
};

enum ProfCls { PC_ZINV = 0, PC_YINV, PC_XA, PC_XB, PC_YFWD, PC_ZFWD, PC_K1, PC_K2, PC_UMAX, PC_MISC, PC_N };
static const char* prof_names[PC_N] = {"pencil_z_inverse", "pencil_y_inverse", "x_phaseA", "x_phaseB",
                                       "pencil_y_forward", "pencil_z_forward", "spec_mu", "spec_update",
                                       "x_umax", "misc"};
static hipEvent_t prof_event(Model& m) {
  if (!m.pool.empty()) { hipEvent_t e = m.pool.back(); m.pool.pop_back(); return e; }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}
static void prof_begin(Model& m, int cls, double bytes) {
  if (!((m.prof >> cls) & 1u)) return;
  Model::Ev ev{cls, prof_event(m), prof_event(m), bytes};
  (void)hipEventRecord(ev.a, m.st);
  m.evs.push_back(ev);
  m.prof_open = true;
}
static void prof_end(Model& m) {
  if (!m.prof_open) return;
  (void)hipEventRecord(m.evs.back().b, m.st);
  m.prof_open = false;
}

static bool mixed(const Model& m) { return m.prec == HY_PREC_MIXED; }
static void* xslot(Model& m, int i) { return (char*)m.xb + m.xsz * m.esz * i; }
static void* zslot(Model& m, int i) { return (char*)m.zb + m.zsz * m.esz * i; }
static void* fslot(Model& m, int i) { return (char*)m.fb + m.nt * m.esz * i; }
static c64* sfield(Model& m, int f) { return m.S + m.nt * f; }

// PencilField::in_f64 of a buffer holding the compute type
static int ct64(const Model& m) { return m.esz == sizeof(c64) ? 1 : 0; }

static SpecArgs spec_args(const Model& m) {
  SpecArgs a;
  const Grid& g = m.g;
  a.Kx = g.Kx; a.Ky = g.Ky; a.Kz = g.Kz; a.My = g.My; a.Mz = g.Mz; a.ny = g.ny; a.nz = g.nz;
  a.nd = g.ndim; a.nt = m.nt;
  a.dkx = 2.0 * fd::PI / g.lx;
  a.dky = g.ny > 1 ? 2.0 * fd::PI / g.ly : 0.0;
  a.dkz = g.nz > 1 ? 2.0 * fd::PI / g.lz : 0.0;
  return a;
}
static int nblocks(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 16384); }

// x-line slots
enum XS { XS_N = 0, XS_M1, XS_M2, XS_M3, XS_GY, XS_GZ, XS_LAP, XS_A0 = 7 /*10 outputs*/ };
// phase B inputs reuse XS_A0.., then outputs
static int XS_B_MU() { return 7; }
static int XS_B_MUY() { return 8; }
static int XS_B_MUZ() { return 9; }
static int XS_B_P(int c) { return 10 + c; }
static int XS_B_PY(int c) { return 16 + c; }
static int XS_B_PZ(int c) { return 22 + c; }
static int XS_B_OUT(const Model& m) { return m.evolve ? 28 : 10; }
// truncated-spectrum slots
static int FS_A(int o) { return o; }          // phase A outputs, 0..9
static int FS_B(int o) { return 10 + o; }     // phase B outputs, 10..18
static const int FS_MU = 19;
static const int FS_RHS = 20;                 // rhs-only output: 10 slots of c64 (20..)

// Compulsory bytes of a pencil pass: every distinct input read once (fields
// that share an input, a field and its derivatives, count it once), every
// output written once.
static double pencil_bytes(const Model& m, const std::vector<PencilField>& f, size_t A, size_t B, size_t L,
                           size_t K) {
  double b = 0.0;
  for (size_t i = 0; i < f.size(); ++i) {
    bool seen = false;
    for (size_t j = 0; j < i && !seen; ++j) seen = f[j].in == f[i].in;
    if (!seen) b += (double)A * B * K * (f[i].in_f64 ? 16.0 : (double)m.esz);
    b += (double)A * B * L * (double)m.esz;
  }
  return b;
}

// The step-limit gate of a pass: only passes recorded for a batch take it
// (a run's own launches and graphs have no step limit, and a gate read at
// the start of every workgroup costs latency).
static const int* batch_gate(const Model& m) { return recorder() ? &m.ctl->halt : nullptr; }

// ---- inverse chains: truncated spectrum -> x-line slot (returns the XIn)
struct InvSpec {
  const void* src;
  int src64;
  double k0off;
  int mx, my, mz;  // XMul / MulKind per axis
  int xslot;       // destination x-line slot (3-D/2-D)
};

static int inv_chain(Model& m, const std::vector<InvSpec>& fs, std::vector<XIn>& xin) {
  const Grid& g = m.g;
  xin.clear();
  const int nd = g.ndim;
  if (g.nz > 1) {
    // Fields that differ only in their y or x multiplier share one z pass
    // (n and its y derivative; mu for its x and y derivatives; each stress
    // component for its x and y derivatives).  A k = 0 offset only reaches
    // the ky = 0 row, which an i ky multiplier zeroes, so such a field may
    // also share the z pass of one with another offset.
    std::vector<PencilField> zf, yf;
    std::vector<int> zid(fs.size());
    for (size_t i = 0; i < fs.size(); ++i) {
      const auto& f = fs[i];
      const int in64 = f.src64 ? 1 : ct64(m);
      int u = -1;
      for (size_t j = 0; g.ny > 1 && j < zf.size() && u < 0; ++j)
        if (zf[j].in == f.src && zf[j].in_f64 == in64 && zf[j].mul == f.mz &&
            (zf[j].k0_offset == f.k0off || f.my == MUL_IK))
          u = (int)j;
      if (u < 0) {
        u = (int)zf.size();
        zf.push_back({f.src, zslot(m, u), f.mz, in64, f.k0off});
      }
      zid[i] = u;
      yf.push_back({zslot(m, u), xslot(m, f.xslot), f.my, ct64(m), 0.0});
    }
    if ((int)fs.size() > m.nzslots) return set_error(HY_ERR_INVALID, "internal: z slots");
    prof_begin(m, PC_ZINV, pencil_bytes(m, zf, (size_t)g.Kx * g.Ky, 1, g.nz, g.Kz));
    HY_TRY(pencil_pass(g, 2, true, true, g.Kx * g.Ky, 1, zf.data(), (int)zf.size(), m.st, batch_gate(m)));
    prof_end(m);
    if (g.ny > 1) {
      prof_begin(m, PC_YINV, pencil_bytes(m, yf, g.Kx, g.nz, g.ny, g.Ky));
      HY_TRY(pencil_pass(g, 1, true, true, g.Kx, g.nz, yf.data(), (int)yf.size(), m.st, batch_gate(m)));
      prof_end(m);
    }
    for (auto& f : fs) xin.push_back({xslot(m, f.xslot), 0, f.mx, 0.0});
    // ny == 1 with nz > 1: the z output is already (Kx, 1, nz) x-lines
    if (g.ny == 1)
      for (size_t i = 0; i < fs.size(); ++i) xin[i].p = zslot(m, zid[i]);
  } else if (g.ny > 1) {
    std::vector<PencilField> yf;
    for (auto& f : fs) yf.push_back({f.src, xslot(m, f.xslot), f.my, f.src64 ? 1 : ct64(m), f.k0off});
    prof_begin(m, PC_YINV, pencil_bytes(m, yf, g.Kx, 1, g.ny, g.Ky));
    HY_TRY(pencil_pass(g, 1, true, true, g.Kx, 1, yf.data(), (int)yf.size(), m.st, batch_gate(m)));
    prof_end(m);
    for (auto& f : fs) xin.push_back({xslot(m, f.xslot), 0, f.mx, 0.0});
  } else {
    for (auto& f : fs) xin.push_back({f.src, f.src64, f.mx, f.k0off});
  }
  (void)nd;
  return HY_OK;
}

// ---- forward chains: x-line slot -> truncated spectrum (returns the x kernel destinations)
static void* fwd_dst(Model& m, int xs, int fs) {
  const Grid& g = m.g;
  return (g.ny > 1 || g.nz > 1) ? xslot(m, xs) : fslot(m, fs);
}
static int fwd_chain(Model& m, const std::vector<std::pair<int, int>>& f /* (xslot, fslot) */) {
  const Grid& g = m.g;
  if (f.empty()) return HY_OK;
  if (g.nz > 1) {
    std::vector<PencilField> yf, zf;
    for (size_t i = 0; i < f.size(); ++i) {
      if (g.ny > 1) yf.push_back({xslot(m, f[i].first), zslot(m, (int)i), MUL_NONE, ct64(m), 0.0});
      const void* zin = g.ny > 1 ? zslot(m, (int)i) : xslot(m, f[i].first);
      zf.push_back({zin, fslot(m, f[i].second), MUL_NONE, ct64(m), 0.0});
    }
    if ((int)f.size() > m.nzslots) return set_error(HY_ERR_INVALID, "internal: z slots");
    if (g.ny > 1) {
      prof_begin(m, PC_YFWD, pencil_bytes(m, yf, g.Kx, g.nz, g.ny, g.Ky));
      HY_TRY(pencil_pass(g, 1, false, true, g.Kx, g.nz, yf.data(), (int)yf.size(), m.st, batch_gate(m)));
      prof_end(m);
    }
    prof_begin(m, PC_ZFWD, pencil_bytes(m, zf, (size_t)g.Kx * g.Ky, 1, g.nz, g.Kz));
    HY_TRY(pencil_pass(g, 2, false, true, g.Kx * g.Ky, 1, zf.data(), (int)zf.size(), m.st, batch_gate(m)));
    prof_end(m);
  } else if (g.ny > 1) {
    std::vector<PencilField> yf;
    for (auto& p : f) yf.push_back({xslot(m, p.first), fslot(m, p.second), MUL_NONE, ct64(m), 0.0});
    prof_begin(m, PC_YFWD, pencil_bytes(m, yf, g.Kx, 1, g.ny, g.Ky));
    HY_TRY(pencil_pass(g, 1, false, true, g.Kx, 1, yf.data(), (int)yf.size(), m.st, batch_gate(m)));
    prof_end(m);
  }
  return HY_OK;
}

// ---- x kernel dispatch
enum XK { XK_A = 0, XK_B = 1, XK_UMAX = 2 };
// device argument slots: one per distinct argument set
enum XSlot { XA_STAGE = 0, XA_RELAX = 1, XB_STAGE = 2, XU_DT = 3, XA_N = 4 };
#define HY_FN(K) reinterpret_cast<const void*>(&K)
// cfg: the XFun form chosen by xk_cfg; the fixed forms are built for lines
// of 32 points and more (the 3-D grids of the timed runs), shorter lines
// always take the run-time form.  The batched twins follow the same choice.
template <class T, int NX>
static int launch_xk(int which, bool maxwell, int cfg, const PhysArgs* a, int nlines, const cpx<T>* tw,
                     hipStream_t s) {
  using X = XWave<T, NX>;
  using CR = XCfg<XF_RUNTIME>;
  const dim3 grid((unsigned)((nlines + X::LPW - 1) / X::LPW));
  if (which == XK_A) {
    if constexpr (NX >= 32) {
      if (cfg == XF_IDEAL)
        return launch_x(HY_FN((k_phaseA<T, NX, X, XCfg<XF_IDEAL>>)), HY_FN((k_phaseA_b<T, NX, X, XCfg<XF_IDEAL>>)),
                        grid, s, a, tw);
      if constexpr (sizeof(T) == 4) {
        if (cfg == XF_ANY3)
          return launch_x(HY_FN((k_phaseA<T, NX, X, XCfg<XF_ANY3>>)), HY_FN((k_phaseA_b<T, NX, X, XCfg<XF_ANY3>>)),
                          grid, s, a, tw);
      }
    }
    return launch_x(HY_FN((k_phaseA<T, NX, X, CR>)), HY_FN((k_phaseA_b<T, NX, X, CR>)), grid, s, a, tw);
  }
  if (which == XK_B) {
    // phase B without the Maxwell advection code where the closure has no
    // evolving stress (fewer registers).  The fixed 3-D form is taken in FP32
    // only: in FP64 it needs more registers than the run-time form (138 against
    // 126 VGPRs at 256 points, which drops a wave per SIMD).
    if constexpr (NX >= 32 && sizeof(T) == 4) {
      if (cfg == XF_ANY3) {
        using C3 = XCfg<XF_ANY3>;
        if (maxwell)
          return launch_x(HY_FN((k_phaseB<T, NX, X, true, C3>)), HY_FN((k_phaseB_b<T, NX, X, true, C3>)), grid, s,
                          a, tw);
        return launch_x(HY_FN((k_phaseB<T, NX, X, false, C3>)), HY_FN((k_phaseB_b<T, NX, X, false, C3>)), grid, s,
                        a, tw);
      }
      if (cfg == XF_NOF3) {
        using C3 = XCfg<XF_NOF3>;
        if (maxwell)
          return launch_x(HY_FN((k_phaseB<T, NX, X, true, C3>)), HY_FN((k_phaseB_b<T, NX, X, true, C3>)), grid, s,
                          a, tw);
        return launch_x(HY_FN((k_phaseB<T, NX, X, false, C3>)), HY_FN((k_phaseB_b<T, NX, X, false, C3>)), grid, s,
                        a, tw);
      }
    }
    if (maxwell)
      return launch_x(HY_FN((k_phaseB<T, NX, X, true, CR>)), HY_FN((k_phaseB_b<T, NX, X, true, CR>)), grid, s, a,
                      tw);
    return launch_x(HY_FN((k_phaseB<T, NX, X, false, CR>)), HY_FN((k_phaseB_b<T, NX, X, false, CR>)), grid, s, a,
                    tw);
  }
  return launch_x(HY_FN((k_umax<T, NX, X>)), HY_FN((k_umax_b<T, NX, X>)), grid, s, a, tw);
}
template <class T>
static int dispatch_xk(int which, int cfg, const Model& m, const PhysArgs* a, const cpx<T>* tw) {
  const int nlines = m.g.ny * m.g.nz;
  switch (m.g.nx) {
#define HY_XK(LL) case LL: HY_TRY((launch_xk<T, LL>(which, m.evolve != 0, cfg, a, nlines, tw, m.st))); break;
    HY_XK(4) HY_XK(8) HY_XK(12) HY_XK(16) HY_XK(24) HY_XK(32) HY_XK(48) HY_XK(64) HY_XK(96)
    HY_XK(128) HY_XK(256)
#undef HY_XK
    default: return set_error(HY_ERR_UNSUPPORTED, "x kernel: unsupported nx");
  }
  HY_HIP(hipGetLastError());
  HY_TRACE(which == XK_A ? "x_phaseA" : (which == XK_B ? "x_phaseB" : "x_umax"), m.st);
  return HY_OK;
}
// The kernel form for an argument set (XFun): a fixed form only where every
// input the kernel reads has the layout and multiplier that form assumes and,
// for phase A, the functional terms and output list are the form's own.
static int xk_cfg(const Model& m, int which, const PhysArgs& a) {
  const Grid& g = m.g;
  if (which == XK_UMAX || g.ndim != 3 || g.ny < 2 || g.nz < 2 || a.nd != 3) return XF_RUNTIME;
  bool ok = true;
  auto in = [&](const XIn& x, int mul) { ok = ok && x.p && !x.is64 && x.k0off == 0.0 && x.mul == mul; };
  in(a.n, XM_NONE);
  if (which == XK_A) {
    for (int d = 0; d < 3; ++d) in(a.m[d], XM_NONE);
    if (a.grad) { in(a.gx, XM_IK); in(a.gy, XM_NONE); in(a.gz, XM_NONE); in(a.lap, XM_NONE); }
    if (!ok) return XF_RUNTIME;
    if (!(a.ideal && !a.tf && !a.grad) || a.nout != 10) return XF_ANY3;
    static const int order[10] = {OA_MU, OA_U1, OA_U2, OA_U3, OA_F11, OA_F12, OA_F13, OA_F22, OA_F23, OA_F33};
    for (int o = 0; o < 10; ++o) ok = ok && a.code[o] == order[o];
    return ok ? XF_IDEAL : XF_ANY3;
  }
  in(a.mux, XM_IK); in(a.muy, XM_NONE); in(a.muz, XM_NONE);
  if (a.evolve) {
    for (int d = 0; d < 3; ++d) in(a.m[d], XM_NONE);
    for (int c = 0; c < 6; ++c) { in(a.px[c], XM_IK); in(a.py[c], XM_NONE); in(a.pz[c], XM_NONE); }
  }
  if (!ok) return XF_RUNTIME;
  return a.vext_kind == HY_VEXT_NONE ? XF_NOF3 : XF_ANY3;
}
// Compulsory bytes of one x-kernel launch: every distinct input x-line the
// kernel reads once, every output written once.
static double xk_bytes(const Model& m, int which, const PhysArgs& a) {
  std::vector<const void*> seen;
  double b = 0.0;
  const double line = (double)m.g.Kx * m.g.ny * m.g.nz;
  auto in = [&](const XIn& x) {
    if (!x.p) return;
    for (auto q : seen) if (q == x.p) return;
    seen.push_back(x.p);
    b += line * (x.is64 ? 16.0 : (double)m.esz);
  };
  const int nd = m.g.ndim;
  in(a.n);
  if (which == XK_A || which == XK_UMAX || a.evolve)
    for (int d = 0; d < nd; ++d) in(a.m[d]);
  if (which == XK_A && a.grad) { in(a.gx); in(a.gy); in(a.gz); in(a.lap); }
  if (which == XK_B) {
    in(a.mux); in(a.muy); in(a.muz);
    if (a.evolve)
      for (int c = 0; c < 6; ++c) { in(a.px[c]); in(a.py[c]); in(a.pz[c]); }
  }
  b += line * a.nout * (double)m.esz;
  return b;
}
// Device copy of an argument set.  The sets depend only on the model's
// configuration, so each slot is uploaded on its first use, from pinned host
// staging and in stream order.  Inside a graph capture the upload becomes a
// copy node of the graph (its source stays valid and unchanged); a set that
// changes while a capture or a recording is active is an error.
static int xargs_dev(Model& m, int slot, const PhysArgs& a, const PhysArgs** out) {
  const bool same = m.xargs_ok[slot] && std::memcmp(&m.p_xargs[slot], &a, sizeof(PhysArgs)) == 0;
  if (!same) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HY_HIP(hipStreamIsCapturing(m.st, &cs));
    if (cs == hipStreamCaptureStatusActive || recorder()) {
      // a captured or recorded launch reads the slot as it is at replay
      if (m.xargs_ok[slot]) return set_error(HY_ERR_INVALID, "x-kernel arguments changed during graph capture");
    } else {
      HY_HIP(hipStreamSynchronize(m.st));  // no queued upload still reads the staging slot
    }
    std::memcpy(&m.p_xargs[slot], &a, sizeof(PhysArgs));
    HY_HIP(hipMemcpyAsync(m.d_xargs + slot, &m.p_xargs[slot], sizeof(PhysArgs), hipMemcpyHostToDevice, m.st));
    m.xargs_ok[slot] = true;
  }
  *out = m.d_xargs + slot;
  return HY_OK;
}
// Every input an x-line kernel will read must be set: a missing one would be
// read through a null pointer.
static int check_xk_inputs(const Model& m, int which, const PhysArgs& a) {
  const int nd = m.g.ndim;
  bool ok = a.n.p != nullptr && a.ctl != nullptr;
  for (int d = 0; d < nd; ++d) ok = ok && a.m[d].p != nullptr;
  if (which == XK_A && a.grad) {
    ok = ok && a.gx.p && a.lap.p && (nd < 2 || a.gy.p) && (nd < 3 || a.gz.p) && a.fdtab;
  }
  if (which == XK_A && a.tf) ok = ok && a.fdtab;
  if (which == XK_B) {
    ok = ok && a.mux.p && (nd < 2 || a.muy.p) && (nd < 3 || a.muz.p);
    if (a.evolve)
      for (int c = 0; c < 6; ++c) ok = ok && a.px[c].p && (nd < 2 || a.py[c].p) && (nd < 3 || a.pz[c].p);
  }
  if (which != XK_UMAX)
    for (int o = 0; o < a.nout; ++o) ok = ok && a.dst[o] != nullptr;
  if (!ok) return set_error(HY_ERR_INVALID, "internal: an x-line kernel input or output is not set");
  // the kernels address their inputs with 32-bit byte offsets
  if ((double)a.Kx * a.ny * a.nz * sizeof(c64) >= 4294967296.0)
    return set_error(HY_ERR_INVALID, "grid too large: an x-line field must stay below 4 GiB");
  return HY_OK;
}

static int run_xk(Model& m, int which, int slot, const PhysArgs& a) {
  HY_TRY(check_xk_inputs(m, which, a));
  const TwiddleSet& t = m.g.tw(m.g.nx);
  const PhysArgs* da = nullptr;
  HY_TRY(xargs_dev(m, slot, a, &da));
  prof_begin(m, which == XK_A ? PC_XA : (which == XK_B ? PC_XB : PC_UMAX), xk_bytes(m, which, a));
  const int cfg = xk_cfg(m, which, a);
  int rc = mixed(m) ? dispatch_xk<float>(which, cfg, m, da, t.d32) : dispatch_xk<double>(which, cfg, m, da, t.d64);
  prof_end(m);
  return rc;
}

static PhysArgs phys_base(Model& m) {
  PhysArgs a;
  std::memset(&a, 0, sizeof(a));
  const Grid& g = m.g;
  a.nd = g.ndim; a.Kx = g.Kx; a.ny = g.ny; a.nz = g.nz;
  a.dx = g.dx; a.dy = g.dy; a.dz = g.dz; a.dkx = 2.0 * fd::PI / g.lx;
  a.n0 = m.n0; a.mass = m.mass; a.beta = m.beta; a.cn = m.cn; a.alpha0 = m.alpha0; a.cg = m.cg;
  a.inv_n0 = 1.0 / m.n0; a.inv_beta = m.beta > 0.0 ? 1.0 / m.beta : 0.0; a.inv_cn = m.cn > 0.0 ? 1.0 / m.cn : 0.0;
  a.dadn = m.dadn;
  a.d2adn2 = m.d2adn2;
  a.ideal = m.ideal; a.tf = m.tf; a.grad = m.grad;
  a.fdtab = m.fdtab; a.ctl = m.ctl; a.evolve = m.evolve;
  a.rtab = m.rtab;
  a.rt_r = RT_R;
  a.rt_invh = RT_N / (2.0 * RT_R);
  a.rt_in = RT_R - 2.0 * (2.0 * RT_R / RT_N);
  a.vext_kind = m.vext_kind; a.vext_td = m.vext_td; a.amp = m.vamp; a.ramp = m.vramp;
  a.omega = m.vomega; a.kvec[0] = m.kvec[0]; a.kvec[1] = m.kvec[1]; a.kvec[2] = m.kvec[2];
  a.fext = m.fext; a.vsc = m.vsc; a.red = m.red;
  return a;
}

// state (n, mom) -> x-lines of n - n0 and mom, plus the electron gradient inputs
static int stage_inputs(Model& m, PhysArgs& a, bool with_grad) {
  const int nd = m.g.ndim;
  std::vector<InvSpec> fs;
  fs.push_back({sfield(m, 0), 1, -m.n0, XM_NONE, MUL_NONE, MUL_NONE, XS_N});
  for (int d = 0; d < nd; ++d) fs.push_back({sfield(m, 1 + d), 1, 0.0, XM_NONE, MUL_NONE, MUL_NONE, XS_M1 + d});
  const int nbase = (int)fs.size();
  if (with_grad) {
    if (nd >= 2) fs.push_back({sfield(m, 0), 1, 0.0, XM_NONE, MUL_IK, MUL_NONE, XS_GY});
    if (nd >= 3) fs.push_back({sfield(m, 0), 1, 0.0, XM_NONE, MUL_NONE, MUL_IK, XS_GZ});
    // Laplacian: -k^2 at the first pass of the chain
    if (nd == 1) fs.push_back({sfield(m, 0), 1, 0.0, XM_NEGK2, MUL_NONE, MUL_NONE, XS_LAP});
    else if (nd == 2) fs.push_back({sfield(m, 0), 1, 0.0, XM_NONE, MUL_NEGK2, MUL_NONE, XS_LAP});
    else fs.push_back({sfield(m, 0), 1, 0.0, XM_NONE, MUL_NONE, MUL_NEGK2, XS_LAP});
  }
  std::vector<XIn> xin;
  HY_TRY(inv_chain(m, fs, xin));
  a.n = xin[0];
  for (int d = 0; d < nd; ++d) a.m[d] = xin[1 + d];
  if (with_grad) {
    int q = nbase;
    a.gx = xin[0];
    a.gx.k0off = 0.0;
    a.gx.mul = XM_IK;
    if (nd >= 2) a.gy = xin[q++];
    if (nd >= 3) a.gz = xin[q++];
    a.lap = xin[q++];
  }
  return HY_OK;
}

static int launch_k1(Model& m, int relax);
static int launch_k2(Model& m, int stage, int rhs_only, int end_ops);

// One RHS evaluation up to K2; the stage time is ctl->ts.  end_ops: the
// control updates K2 applies after the stage (see K2Args).
static int stage(Model& m, int s, int rhs_only, int end_ops = 0) {
  const Grid& g = m.g;
  const int nd = g.ndim;
  // ---- phase A
  PhysArgs a = phys_base(m);
  HY_TRY(stage_inputs(m, a, m.grad));
  const bool need_u = m.stress_newt || m.evolve;
  std::vector<std::pair<int, int>> fwd;
  auto addA = [&](int code) {
    int o = a.nout++;
    a.code[o] = code;
    a.dst[o] = fwd_dst(m, XS_A0 + o, FS_A(code));
    fwd.push_back({XS_A0 + o, FS_A(code)});
  };
  addA(OA_MU);
  if (need_u) for (int d = 0; d < nd; ++d) addA(OA_U1 + d);
  const int fcode[3][3] = {{OA_F11, OA_F12, OA_F13}, {OA_F12, OA_F22, OA_F23}, {OA_F13, OA_F23, OA_F33}};
  for (int i = 0; i < nd; ++i)
    for (int j = i; j < nd; ++j) addA(fcode[i][j]);
  HY_TRY(run_xk(m, XK_A, XA_STAGE, a));
  HY_TRY(fwd_chain(m, fwd));
  // ---- K1: mu_hat (+ pending Maxwell relaxation)
  HY_TRY(launch_k1(m, m.relax_pending && !rhs_only));
  if (m.evolve && !rhs_only) m.relax_pending = 0;
  // ---- phase B
  PhysArgs b = phys_base(m);
  b.n = a.n;
  for (int d = 0; d < nd; ++d) b.m[d] = a.m[d];
  std::vector<InvSpec> fs;
  fs.push_back({fslot(m, FS_MU), 0, 0.0, XM_IK, MUL_NONE, MUL_NONE, XS_B_MU()});
  if (nd >= 2) fs.push_back({fslot(m, FS_MU), 0, 0.0, XM_NONE, MUL_IK, MUL_NONE, XS_B_MUY()});
  if (nd >= 3) fs.push_back({fslot(m, FS_MU), 0, 0.0, XM_NONE, MUL_NONE, MUL_IK, XS_B_MUZ()});
  if (m.evolve) {
    for (int c = 0; c < 6; ++c) {
      fs.push_back({sfield(m, 4 + c), 1, 0.0, XM_IK, MUL_NONE, MUL_NONE, XS_B_P(c)});
      if (nd >= 2) fs.push_back({sfield(m, 4 + c), 1, 0.0, XM_NONE, MUL_IK, MUL_NONE, XS_B_PY(c)});
      if (nd >= 3) fs.push_back({sfield(m, 4 + c), 1, 0.0, XM_NONE, MUL_NONE, MUL_IK, XS_B_PZ(c)});
    }
  }
  std::vector<XIn> xin;
  HY_TRY(inv_chain(m, fs, xin));
  int q = 0;
  b.mux = xin[q++];
  if (nd >= 2) b.muy = xin[q++];
  if (nd >= 3) b.muz = xin[q++];
  if (m.evolve)
    for (int c = 0; c < 6; ++c) {
      b.px[c] = xin[q++];
      if (nd >= 2) b.py[c] = xin[q++];
      if (nd >= 3) b.pz[c] = xin[q++];
    }
  std::vector<std::pair<int, int>> fwdb;
  const int xo = XS_B_OUT(m);
  auto addB = [&](int code) {
    int o = b.nout++;
    b.code[o] = code;
    b.dst[o] = fwd_dst(m, xo + o, FS_B(code));
    fwdb.push_back({xo + o, FS_B(code)});
  };
  for (int i = 0; i < nd; ++i) addB(OB_H1 + i);
  if (m.evolve) for (int c = 0; c < 6; ++c) addB(OB_ADV1 + c);
  HY_TRY(run_xk(m, XK_B, XB_STAGE, b));
  HY_TRY(fwd_chain(m, fwdb));
  // ---- K2: RHS assembly and the RK update
  HY_TRY(launch_k2(m, s, rhs_only, end_ops));
  return HY_OK;
}

static int launch_k1(Model& m, int relax) {
  K1Args a;
  std::memset(&a, 0, sizeof(a));
  a.s = spec_args(m);
  a.fmu = fslot(m, FS_A(OA_MU));
  a.N = sfield(m, 0);
  a.kcomb = m.kcomb;
  a.n0 = m.n0;
  a.muhat = fslot(m, FS_MU);
  a.evolve = m.evolve;
  a.relax = relax;
  a.P = sfield(m, 4);
  for (int d = 0; d < 3; ++d) a.fu[d] = fslot(m, FS_A(OA_U1 + d));
  a.uold = m.uold;
  a.eta = m.eta; a.xi = m.xi;
  a.ctl = m.ctl;
  {
    const double nt = (double)m.nt, es = (double)m.esz;
    double b = nt * (es + 16.0 + 8.0 + es);
    // Maxwell: F(u) read and stored as the stage-start u; with a relaxation,
    // the previous stage-start u read and Pi read and written
    if (m.evolve) b += nt * (3 * es + 3 * es + (relax ? 3 * es + 6 * 16.0 * 2 : 0.0));
    prof_begin(m, PC_K1, b);
  }
  if (mixed(m)) HY_TRY(launch(HY_FN(k_spec1<float>), HY_FN(k_spec1_b<float>), dim3(nblocks(m.nt)), dim3(256), m.st, a));
  else HY_TRY(launch(HY_FN(k_spec1<double>), HY_FN(k_spec1_b<double>), dim3(nblocks(m.nt)), dim3(256), m.st, a));
  prof_end(m);
  HY_HIP(hipGetLastError());
  HY_TRACE("spec_mu", m.st);
  return HY_OK;
}

static int launch_k2(Model& m, int stage_idx, int rhs_only, int end_ops) {
  K2Args a;
  std::memset(&a, 0, sizeof(a));
  a.s = spec_args(m);
  a.S = m.S;
  a.R = rhs_only ? fslot(m, FS_RHS) : (void*)m.R;
  const int fo[6] = {OA_F11, OA_F12, OA_F13, OA_F22, OA_F23, OA_F33};
  for (int i = 0; i < 6; ++i) a.fl[i] = fslot(m, FS_A(fo[i]));
  for (int i = 0; i < 3; ++i) a.fh[i] = fslot(m, FS_B(OB_H1 + i));
  for (int i = 0; i < 3; ++i) a.fu[i] = fslot(m, FS_A(OA_U1 + i));
  for (int c = 0; c < 6; ++c) a.fadv[c] = fslot(m, FS_B(OB_ADV1 + c));
  a.mass = m.mass; a.inv_mass = 1.0 / m.mass; a.eta = m.eta; a.xi = m.xi;
  a.stress_newt = m.stress_newt; a.evolve = m.evolve;
  a.ext_active = m.vext_kind != HY_VEXT_NONE;
  a.red = m.red;
  a.nlines = m.g.ny * m.g.nz;
  a.red_scale = 1.0 / (double)m.g.nreal();
  a.stage = stage_idx;
  a.rhs_only = rhs_only;
  a.end_ops = rhs_only ? 0 : end_ops;
  a.tau = m.tau;
  a.ctl = m.ctl;
  {
    const double nt = (double)m.nt, es = (double)m.esz;
    const int nd = m.g.ndim;
    const int nfl = nd * (nd + 1) / 2;
    const int ns = m.evolve ? 10 : 4;
    double b = nt * es * (nfl + nd + (m.stress_newt ? nd : 0) + (m.evolve ? 6 : 0));
    // S read and written; R written (FP64 in rhs mode, else the transform
    // type) and, after the first stage, read
    b += nt * 16.0 * ns * 2 + nt * ns * (rhs_only ? 16.0 : es * (stage_idx > 0 ? 2.0 : 1.0));
    if (m.vext_kind != HY_VEXT_NONE) b += 8.0 * 3.0 * m.g.ny * m.g.nz;  // phase B's per-line force sums
    prof_begin(m, PC_K2, b);
  }
  const dim3 gr(nblocks(m.nt)), bl(256);
  switch (m.g.ndim * 2 + (mixed(m) ? 1 : 0)) {
#define HY_K2(T, ND) HY_TRY(launch(HY_FN((k_spec2<T, ND>)), HY_FN((k_spec2_b<T, ND>)), gr, bl, m.st, a))
    case 2: HY_K2(double, 1); break;
    case 3: HY_K2(float, 1); break;
    case 4: HY_K2(double, 2); break;
    case 5: HY_K2(float, 2); break;
    case 6: HY_K2(double, 3); break;
    default: HY_K2(float, 3); break;
#undef HY_K2
  }
  prof_end(m);
  HY_HIP(hipGetLastError());
  HY_TRACE("spec_update", m.st);
  return HY_OK;
}

// Pending Maxwell relaxation after the last stage: needs u_hat of the new state.
static int finish_relax(Model& m) {
  if (!m.evolve || !m.relax_pending) return HY_OK;
  const int nd = m.g.ndim;
  PhysArgs a = phys_base(m);
  HY_TRY(stage_inputs(m, a, false));
  // only u = mom / (m n) is formed here: no functional, so none of its inputs
  // (the gradient term's grad n and lap n are not loaded on this path)
  a.ideal = 0;
  a.tf = 0;
  a.grad = 0;
  std::vector<std::pair<int, int>> fwd;
  a.nout = 0;
  for (int d = 0; d < nd; ++d) {
    int o = a.nout++;
    a.code[o] = OA_U1 + d;
    a.dst[o] = fwd_dst(m, XS_A0 + o, FS_A(OA_U1 + d));
    fwd.push_back({XS_A0 + o, FS_A(OA_U1 + d)});
  }
  HY_TRY(run_xk(m, XK_A, XA_RELAX, a));
  HY_TRY(fwd_chain(m, fwd));
  // mu is not recomputed here: K1 rewrites mu_hat from the stale F(mu_loc) but
  // it is never read before the next stage recomputes it.
  HY_TRY(launch_k1(m, 1));
  m.relax_pending = 0;
  return HY_OK;
}

struct SetDtK {
  Ctl* c;
  double dt;
};
__device__ __attribute__((always_inline)) inline void set_dt_body(const SetDtK& k) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && !k.c->halt) k.c->dt = k.dt;
}
__global__ void k_set_dt(SetDtK k) { set_dt_body(k); }
__global__ void k_set_dt_b(const SetDtK* __restrict__ ks) { set_dt_body(ks[blockIdx.z]); }

static int compute_dt_dev(Model& m) {
  if (m.dt_fixed > 0.0) {  // model.f90:155-158
    HY_TRY(launch(HY_FN(k_set_dt), HY_FN(k_set_dt_b), dim3(1), dim3(1), m.st, SetDtK{m.ctl, m.dt_fixed}));
    HY_HIP(hipGetLastError());
    return HY_OK;
  }
  PhysArgs a = phys_base(m);
  HY_TRY(stage_inputs(m, a, false));
  HY_TRY(run_xk(m, XK_UMAX, XU_DT, a));
  const DtK dk{m.red, m.g.ny * m.g.nz, m.ctl, m.wmax, m.numax, m.dxmin, m.cfl};
  return launch(HY_FN(k_dt), HY_FN(k_dt_b), dim3(1), dim3(256), m.st, dk);
}

static int ctl_op(Model& m, int op, double tend) {
  const CtlK k{m.ctl, op, tend};
  return launch(HY_FN(k_ctl), HY_FN(k_ctl_b), dim3(1), dim3(1), m.st, k);
}

// One RK3 step: the step's start on the control block, then three stages,
// whose K2 launches end each stage and, after the third, the step.
static int step_once(Model& m, double tend) {
  HY_TRY(ctl_op(m, CTL_STEP_BEGIN, tend));
  for (int s = 0; s < 3; ++s) {
    HY_TRY(stage(m, s, 0, s < 2 ? 1 : 2));
    if (m.evolve) m.relax_pending = 1;
  }
  return HY_OK;
}

static int read_ctl(Model& m, Ctl& c) {
  HY_HIP(hipMemcpyAsync(&c, m.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, m.st));
  HY_HIP(hipStreamSynchronize(m.st));
  m.h_t = c.t; m.h_dt = c.dt; m.h_step = c.step;
  return HY_OK;
}

// ---------------------------------------------------------- host helpers
static double interp_linear(const std::vector<double>& x, const std::vector<double>& y, double xq) {
  size_t n = x.size();  // utils.f90:30-55
  if (n == 1) return y[0];
  size_t lo;
  if (xq <= x[0]) lo = 0;
  else if (xq >= x[n - 1]) lo = n - 2;
  else {
    size_t l = 0, h = n - 1;
    while (h - l > 1) { size_t mid = (l + h) / 2; if (x[mid] <= xq) l = mid; else h = mid; }
    lo = l;
  }
  double w = (xq - x[lo]) / (x[lo + 1] - x[lo]);
  return (1.0 - w) * y[lo] + w * y[lo + 1];
}

// Per-term linear kernels K_i(k) (free_energy.f90:22-59 and each term's kernel);
// kt, ct: the RY term's radial table
static double term_kernel_desc(const hy_term_desc& d, const std::vector<double>& kt, const std::vector<double>& ct,
                               double k) {
  const double* tab = fd::host_table();
  const double pi = 3.14159265358979323846264338327950288;
  switch (d.kind) {
    case HY_TERM_IDEAL: return 1.0 / (d.beta * d.n0);
    case HY_TERM_TF: return 2.0 / (d.beta * d.cn * fd::im12(tab, d.alpha0));
    case HY_TERM_GRADIENT: return 2.0 * d.cg * fd::xi(tab, d.alpha0) * k * k;
    case HY_TERM_HARTREE: {
      double den = k * k + d.kappa * d.kappa;
      return den > 0.0 ? 4.0 * pi * d.e2 / den : 0.0;
    }
    case HY_TERM_RY: {
      double k1 = std::max(k, kt[0]);
      double c = interp_linear(kt, ct, std::min(k1, kt.back()));
      if (d.subtract_mean_field) c = c + d.beta * 4.0 * pi * d.e2 / (k1 * k1 + d.kappa * d.kappa);
      return -c / d.beta;
    }
  }
  return 0.0;
}
static double term_kernel(const Model& m, size_t t, double k) {
  return term_kernel_desc(m.terms[t], m.ktabs[t], m.ctabs[t], k);
}
static double fun_kernel(const Model& m, double k) {
  double kk = 0.0;
  for (size_t t = 0; t < m.terms.size(); ++t) kk += term_kernel(m, t, k);
  return kk;
}

static void model_free(Model& m) {
  for (void* p : {(void*)m.S, (void*)m.R, (void*)m.uold, (void*)m.kcomb, (void*)m.fext, (void*)m.vsc, (void*)m.rtab,
                  (void*)m.fdtab,
                  m.xb, m.zb, m.fb, (void*)m.red, (void*)m.ctl, (void*)m.probe, (void*)m.aprobe, (void*)m.d_xargs,
                  (void*)m.d_real, (void*)m.d_full})
    if (p) (void)hipFree(p);
  for (auto& w : m.work)
    if (w.p) (void)hipFree(w.p);
  m.work.clear();
  if (m.full_kh) (void)hipFree(m.full_kh);
  if (m.full_kr) (void)hipFree(m.full_kr);
  m.full_kh = m.full_kr = nullptr;
  if (m.gexec) (void)hipGraphExecDestroy(m.gexec);
  if (m.graph) (void)hipGraphDestroy(m.graph);
  m.gexec = nullptr;
  m.graph = nullptr;
  if (m.p_xargs) (void)hipHostFree(m.p_xargs);
  m.p_xargs = nullptr;
  for (auto& e : m.evs) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  for (auto e : m.pool) (void)hipEventDestroy(e);
  m.evs.clear();
  m.pool.clear();
  if (m.st) (void)hipStreamDestroy(m.st);
  grid_free(m.g);
}

static int dmalloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return set_error(HY_ERR_OOM, "model: device allocation failed");
  }
  return HY_OK;
}

int model_build(Model& m, const hy_model_desc* d);
int fft_r2c_full_api(Grid& g, const double* u, c64* uhat, hipStream_t s);
int fft_c2r_full_api(Grid& g, const c64* uhat, double* u, hipStream_t s);

// ------------------------------------------------- functional mu and budgets
// Real-space evaluation in FP64 on the full half spectrum, following
// functional_mu (functional.f90:56-67) and diag_budgets (diagnostics.f90:41-60).
struct FullArgs {
  int nkx, ny, nz, Kx, My, Mz;
  const double *kx, *ky, *kz;
};
__device__ inline bool fkept(const FullArgs& a, int i, int j, int k) {
  int my = j <= a.ny / 2 ? j : j - a.ny, mz = k <= a.nz / 2 ? k : k - a.nz;
  return i < a.Kx && abs(my) <= a.My && abs(mz) <= a.Mz;
}
// out = mask (kh N + kr (N - n0 delta_k0))  | op 1: out = mask i k_dir N | op 2: out = -mask k^2 N
__global__ void k_full_conv(FullArgs a, int op, const double* __restrict__ kdir, size_t kstride, int kn,
                            const c64* __restrict__ N, const double* __restrict__ kh,
                            const double* __restrict__ kr, double n0, c64* __restrict__ out) {
  size_t n = (size_t)a.nkx * a.ny * a.nz;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(e % a.nkx), j = (int)((e / a.nkx) % a.ny), k = (int)(e / ((size_t)a.nkx * a.ny));
    double m = fkept(a, i, j, k) ? 1.0 : 0.0;
    c64 v = N[e];
    c64 r;
    if (op == 0) {
      c64 vp = v;
      if (e == 0) vp.x -= n0;
      r = scal(v, kh[e]) + scal(vp, kr[e]);
    } else if (op == 1) {
      double kd = kdir[(e / kstride) % (size_t)kn];
      r = mk<double>(-kd * v.y, kd * v.x);
    } else {
      double k2 = a.kx[i] * a.kx[i] + a.ky[j] * a.ky[j] + a.kz[k] * a.kz[k];
      r = scal(v, -k2);
    }
    out[e] = scal(r, m);
  }
}

struct PointArgs {
  size_t nr;
  int nx, ny, nz, nd;
  double dx, dy, dz, lx, ly, lz;
  const double* n;
  const double* mom;     // 3 fields (budgets)
  const double* conv;    // mu: convolution sum | budgets: hartree w
  const double* wry;     // budgets: ry w
  const double* gn;      // 3 gradient fields (gradient term)
  const double* lap;
  double n0, beta, cn, alpha0, cg;
  int ideal, tf, grad, hartree, ry;
  const double* fdtab;
  // external potential for the budgets
  int vext_kind;
  double vamp, vt, vomega, vramp, vwidth, kvec[3];
  int vext_td;
  double* out;           // mu output | budgets: per-block partials (8 per block)
  int* flag;             // set on non-positive density
};

__global__ void k_point_mu(PointArgs a) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.nr; e += (size_t)gridDim.x * blockDim.x) {
    double n = a.n[e];
    if (!(n > 0.0)) { *a.flag = 1; a.out[e] = 0.0; continue; }
    double mu = a.conv[e];
    if (a.ideal) mu += log(n) / a.beta;
    if (a.tf || a.grad) {
      double al = fd::alpha_from_ip12(a.fdtab, n / a.cn);
      if (a.tf) mu += al / a.beta;
      if (a.grad) {
        double xi = fd::xi(a.fdtab, al);
        double dxidn = fd::dxi(a.fdtab, al) / (0.5 * a.cn * fd::im12(a.fdtab, al));
        double g2 = a.gn[e] * a.gn[e] + a.gn[a.nr + e] * a.gn[a.nr + e] + a.gn[2 * a.nr + e] * a.gn[2 * a.nr + e];
        mu -= a.cg * (dxidn * g2 + 2.0 * xi * a.lap[e]);
      }
    }
    a.out[e] = mu;
  }
}

__device__ inline double vext_at(const PointArgs& a, size_t e) {
  int i = (int)(e % a.nx), j = (int)((e / a.nx) % a.ny), k = (int)(e / ((size_t)a.nx * a.ny));
  double x = i * a.dx, y = j * a.dy, z = k * a.dz;
  if (a.vext_kind == HY_VEXT_MODE || a.vext_kind == HY_VEXT_DRIVEN) {
    double t = a.vext_td ? a.vt : 0.0;
    double amp = a.vamp;
    if (a.vext_td && a.vramp > 0.0) amp = amp * fmin(1.0, t / a.vramp);
    double ph = a.kvec[0] * x + a.kvec[1] * y + a.kvec[2] * z - a.vomega * t;
    return amp * cos(ph);
  }
  if (a.vext_kind == HY_VEXT_GAUSSIAN) {
    double xx = x - 0.5 * a.lx, yy = a.ny > 1 ? y - 0.5 * a.ly : 0.0, zz = a.nz > 1 ? z - 0.5 * a.lz : 0.0;
    double r2 = xx * xx + yy * yy + zz * zz;
    return -a.vamp * exp(-0.5 * r2 / (a.vwidth * a.vwidth));
  }
  return 0.0;
}

// partial sums: 0 sum mom^2/n, 1 n(ln n - 1), 2 n w_h, 3 (n - n0) w_ry,
// 4 n alpha - 2/3 cn I_3/2(alpha), 5 xi |grad n|^2, 6 n v_ext
__global__ void k_point_budget(PointArgs a) {
  __shared__ double sh[7][256];
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.nr; e += (size_t)gridDim.x * blockDim.x) {
    double n = a.n[e];
    if (!(n > 0.0)) { *a.flag = 1; continue; }
    for (int d = 0; d < a.nd; ++d) acc[0] += a.mom[d * a.nr + e] * a.mom[d * a.nr + e] / n;
    if (a.ideal) acc[1] += n * (log(n) - 1.0);
    if (a.hartree) acc[2] += n * a.conv[e];
    if (a.ry) acc[3] += (n - a.n0) * a.wry[e];
    if (a.tf || a.grad) {
      double al = fd::alpha_from_ip12(a.fdtab, n / a.cn);
      if (a.tf) acc[4] += n * al - (2.0 / 3.0) * a.cn * fd::ip32(a.fdtab, al);
      if (a.grad) {
        double g2 = a.gn[e] * a.gn[e] + a.gn[a.nr + e] * a.gn[a.nr + e] + a.gn[2 * a.nr + e] * a.gn[2 * a.nr + e];
        acc[5] += fd::xi(a.fdtab, al) * g2;
      }
    }
    if (a.vext_kind != HY_VEXT_NONE) acc[6] += n * vext_at(a, e);
  }
  for (int q = 0; q < 7; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 7; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 7; ++q) a.out[blockIdx.x * 8 + q] = sh[q][0];
}

static FullArgs full_args_m(const Model& m) {
  FullArgs a;
  const Grid& g = m.g;
  a.nkx = g.nkx; a.ny = g.ny; a.nz = g.nz; a.Kx = g.Kx; a.My = g.My; a.Mz = g.Mz;
  a.kx = g.d_kx; a.ky = g.d_ky; a.kz = g.d_kz;
  return a;
}

// hartree v(k) and ry -c_cor(k)/beta on the full half spectrum
static int full_kernels(Model& m, double** kh, double** kr) {
  const Grid& g = m.g;
  std::vector<double> h(g.nfull(), 0.0), r(g.nfull(), 0.0);
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nkx; ++i) {
        size_t e = ((size_t)k * g.ny + j) * g.nkx + i;
        double kabs = std::sqrt(g.kx[i] * g.kx[i] + g.ky[j] * g.ky[j] + g.kz[k] * g.kz[k]);
        for (size_t t = 0; t < m.terms.size(); ++t) {
          if (m.terms[t].kind == HY_TERM_HARTREE) h[e] += term_kernel(m, t, kabs);
          if (m.terms[t].kind == HY_TERM_RY) r[e] += term_kernel(m, t, kabs);
        }
      }
  HY_TRY(dmalloc((void**)kh, sizeof(double) * g.nfull()));
  HY_TRY(dmalloc((void**)kr, sizeof(double) * g.nfull()));
  HY_HIP(hipMemcpy(*kh, h.data(), sizeof(double) * g.nfull(), hipMemcpyHostToDevice));
  HY_HIP(hipMemcpy(*kr, r.data(), sizeof(double) * g.nfull(), hipMemcpyHostToDevice));
  return HY_OK;
}

// the same tables, built on first use and kept with the model
static int full_kernels_kept(Model& m, double** kh, double** kr) {
  if (!m.full_kh) {
    double *a = nullptr, *b = nullptr;
    const int rc = full_kernels(m, &a, &b);
    if (rc != HY_OK) {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      return rc;
    }
    m.full_kh = a;
    m.full_kr = b;
  }
  *kh = m.full_kh;
  *kr = m.full_kr;
  return HY_OK;
}

// Work buffers of hy_model_mu / hy_model_budgets: the i-th request of a call
// gets the model's i-th kept buffer, allocated on first use (or when a call
// needs it larger) and kept until the model is destroyed.  Freeing device
// memory waits for all queued work on the device, so per-call buffers would
// stall every other stream each time a run takes its diagnostics.
struct DevBufs {
  Model& m;
  size_t next = 0;
  explicit DevBufs(Model& mm) : m(mm) {}
  template <class T> int get(T** out, size_t n) {
    const size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
    if (next == m.work.size()) m.work.push_back({nullptr, 0});
    auto& w = m.work[next++];
    if (w.bytes < bytes) {
      if (w.p) (void)hipFree(w.p);
      w.p = nullptr;
      w.bytes = 0;
      HY_TRY(dmalloc(&w.p, bytes));
      w.bytes = bytes;
    }
    *out = (T*)w.p;
    return HY_OK;
  }
};

// real fields needed by mu and budgets, from a real density on the device:
// Nhat, conv (hartree+ry or hartree only), wry, grad n, lap n
static int real_terms(Model& m, DevBufs& B, const double* dn, bool split, c64** Nh, double** conv,
                      double** wry, double** gn, double** lap) {
  Grid& g = m.g;
  const size_t nr = g.nreal(), nf = g.nfull();
  c64 *N, *W;
  HY_TRY(B.get(&N, nf));
  HY_TRY(B.get(&W, nf));
  double *kh, *kr;
  HY_TRY(full_kernels_kept(m, &kh, &kr));
  HY_TRY(fft_r2c_full_api(g, dn, N, m.st));
  FullArgs fa = full_args_m(m);
  const int nb = nblocks(nf);
  HY_TRY(B.get(conv, nr));
  double* zero;
  HY_TRY(B.get(&zero, nf));
  HY_HIP(hipMemsetAsync(zero, 0, sizeof(double) * nf, m.st));
  hipLaunchKernelGGL(k_full_conv, dim3(nb), dim3(256), 0, m.st, fa, 0, g.d_kx, (size_t)1, g.nkx, N, kh,
                     split ? zero : kr, m.n0, W);
  HY_TRY(fft_c2r_full_api(g, W, *conv, m.st));
  *wry = nullptr;
  if (split) {
    HY_TRY(B.get(wry, nr));
    hipLaunchKernelGGL(k_full_conv, dim3(nb), dim3(256), 0, m.st, fa, 0, g.d_kx, (size_t)1, g.nkx, N, zero, kr,
                       m.n0, W);
    HY_TRY(fft_c2r_full_api(g, W, *wry, m.st));
  }
  *gn = nullptr; *lap = nullptr;
  if (m.grad) {
    HY_TRY(B.get(gn, nr * 3));
    HY_TRY(B.get(lap, nr));
    HY_HIP(hipMemsetAsync(*gn, 0, sizeof(double) * nr * 3, m.st));
    for (int d = 0; d < g.ndim; ++d) {
      const double* kdir = d == 0 ? g.d_kx : (d == 1 ? g.d_ky : g.d_kz);
      size_t kstride = d == 0 ? 1 : (d == 1 ? (size_t)g.nkx : (size_t)g.nkx * g.ny);
      int kn = d == 0 ? g.nkx : (d == 1 ? g.ny : g.nz);
      hipLaunchKernelGGL(k_full_conv, dim3(nb), dim3(256), 0, m.st, fa, 1, kdir, kstride, kn, N, kh, kr, m.n0, W);
      HY_TRY(fft_c2r_full_api(g, W, *gn + nr * d, m.st));
    }
    hipLaunchKernelGGL(k_full_conv, dim3(nb), dim3(256), 0, m.st, fa, 2, g.d_kx, (size_t)1, g.nkx, N, kh, kr,
                       m.n0, W);
    HY_TRY(fft_c2r_full_api(g, W, *lap, m.st));
  }
  HY_HIP(hipGetLastError());
  *Nh = N;
  return HY_OK;
}

static PointArgs point_args(Model& m) {
  PointArgs a;
  std::memset(&a, 0, sizeof(a));
  const Grid& g = m.g;
  a.nr = g.nreal(); a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nd = g.ndim;
  a.dx = g.dx; a.dy = g.dy; a.dz = g.dz; a.lx = g.lx; a.ly = g.ly; a.lz = g.lz;
  a.n0 = m.n0; a.beta = m.beta; a.cn = m.cn; a.alpha0 = m.alpha0; a.cg = m.cg;
  a.ideal = m.ideal; a.tf = m.tf; a.grad = m.grad; a.hartree = m.hartree; a.ry = m.ry;
  a.fdtab = m.fdtab;
  a.vext_kind = m.vext_kind; a.vamp = m.vamp; a.vomega = m.vomega; a.vramp = m.vramp; a.vwidth = m.vwidth;
  a.kvec[0] = m.kvec[0]; a.kvec[1] = m.kvec[1]; a.kvec[2] = m.kvec[2]; a.vext_td = m.vext_td;
  return a;
}

}  // namespace hy

// =========================================================================== ABI
using namespace hy;

struct hy_model_s {
  Model m;
};

namespace hy {
int fft_r2c_full_api(Grid& g, const double* u, c64* uhat, hipStream_t s);
int fft_c2r_full_api(Grid& g, const c64* uhat, double* u, hipStream_t s);
int spec_grad_api(Grid& g, const double* u, double* du3, hipStream_t s);

int model_build(Model& m, const hy_model_desc* d) {
  HY_TRY(grid_build(m.g, d->nx, d->ny, d->nz, d->lx, d->ly, d->lz, d->dealias, d->precision));
  const Grid& g = m.g;
  m.prec = d->precision;
  m.esz = mixed(m) ? sizeof(c32) : sizeof(c64);
  m.mass = d->mass;
  m.n0 = d->n0;
  if (!(m.mass > 0.0) || !(m.n0 > 0.0)) return set_error(HY_ERR_INVALID, "model: mass and n0 must be > 0");
  if (d->nterms < 0 || (d->nterms > 0 && !d->terms)) return set_error(HY_ERR_INVALID, "model: terms");
  m.terms.assign(d->terms, d->terms + d->nterms);
  m.ktabs.resize(d->nterms);
  m.ctabs.resize(d->nterms);
  for (int t = 0; t < d->nterms; ++t) {
    const hy_term_desc& td = d->terms[t];
    switch (td.kind) {
      case HY_TERM_IDEAL: m.ideal = 1; m.beta = td.beta; break;
      case HY_TERM_TF: m.tf = 1; m.beta = td.beta; m.cn = td.cn; m.alpha0 = td.alpha0; break;
      case HY_TERM_GRADIENT: m.grad = 1; m.beta = td.beta; m.cn = td.cn; m.alpha0 = td.alpha0; m.cg = td.cg; break;
      case HY_TERM_HARTREE: m.hartree = 1; break;
      case HY_TERM_RY:
        m.ry = 1;
        if (td.ntab < 1 || !td.k_tab || !td.c_tab) return set_error(HY_ERR_INVALID, "ry term needs a structure table");
        m.ktabs[t].assign(td.k_tab, td.k_tab + td.ntab);
        m.ctabs[t].assign(td.c_tab, td.c_tab + td.ntab);
        break;
      default: return set_error(HY_ERR_UNKNOWN_KIND, "unknown free-energy term kind");
    }
    if ((td.kind == HY_TERM_TF || td.kind == HY_TERM_GRADIENT) && !(td.cn > 0.0))
      return set_error(HY_ERR_INVALID, "tf/gradient term needs cn > 0 (kind = electron)");
  }
  for (auto& td : m.terms)
    if (td.kind != HY_TERM_HARTREE && td.kind != HY_TERM_RY) m.beta = td.beta;
  if (d->closure == HY_CLOSURE_NEWTONIAN) { m.evolve = 0; m.tau = 0.0; }
  else if (d->closure == HY_CLOSURE_MAXWELL) { m.evolve = d->tau > 0.0; m.tau = d->tau; }
  else return set_error(HY_ERR_UNKNOWN_KIND, "unknown stress closure");
  m.closure = d->closure;
  m.eta = d->eta; m.xi = d->xi; m.eta_l = d->eta_l;
  m.stress_newt = !m.evolve && (m.eta != 0.0 || m.xi != 0.0);
  m.vext_kind = d->vext_kind;
  if (m.vext_kind < 0 || m.vext_kind > 3) return set_error(HY_ERR_UNKNOWN_KIND, "unknown external potential");
  m.vext_td = m.vext_kind == HY_VEXT_DRIVEN;
  m.vamp = d->vext_amp; m.vomega = d->vext_omega; m.vwidth = d->vext_width; m.vramp = d->vext_ramp;
  const double twopi = 2.0 * 3.14159265358979323846264338327950288;
  for (int i = 0; i < 3; ++i) m.vmode[i] = d->vext_mode[i];
  m.kvec[0] = twopi * m.vmode[0] / g.lx;  // external_potential.f90:46-48
  m.kvec[1] = g.ny == 1 ? 0.0 : twopi * m.vmode[1] / g.ly;
  m.kvec[2] = g.nz == 1 ? 0.0 : twopi * m.vmode[2] / g.lz;
  m.dt_fixed = d->dt;
  m.cfl = d->cfl;
  HY_HIP(hipStreamCreateWithFlags(&m.st, hipStreamNonBlocking));
  // sizes and buffers
  m.nt = g.ntrunc();
  m.xsz = (size_t)g.Kx * g.ny * g.nz;
  m.zsz = (size_t)g.Kx * g.Ky * g.nz;
  m.nxslots = m.evolve ? 37 : 17;
  m.nzslots = 24;
  m.nfslots = 20;
  HY_TRY(dmalloc((void**)&m.S, sizeof(c64) * m.nt * 10));
  HY_TRY(dmalloc((void**)&m.R, sizeof(c64) * m.nt * 10));
  HY_TRY(dmalloc((void**)&m.uold, sizeof(c64) * m.nt * 3));
  HY_TRY(dmalloc(&m.xb, m.esz * m.xsz * m.nxslots));
  HY_TRY(dmalloc(&m.zb, m.esz * m.zsz * m.nzslots));
  // FS slots in the compute type plus 10 FP64 rhs-output slots
  HY_TRY(dmalloc(&m.fb, m.esz * m.nt * m.nfslots + sizeof(c64) * m.nt * 10));
  HY_TRY(dmalloc((void**)&m.red, sizeof(double) * std::max<size_t>((size_t)g.ny * g.nz * 4, 4096)));
  HY_TRY(dmalloc((void**)&m.ctl, sizeof(Ctl)));
  HY_HIP(hipMemsetAsync(m.S, 0, sizeof(c64) * m.nt * 10, m.st));
  HY_HIP(hipMemsetAsync(m.R, 0, sizeof(c64) * m.nt * 10, m.st));
  HY_HIP(hipMemsetAsync(m.uold, 0, sizeof(c64) * m.nt * 3, m.st));
  HY_HIP(hipMemsetAsync(m.xb, 0, m.esz * m.xsz * m.nxslots, m.st));
  HY_HIP(hipMemsetAsync(m.fb, 0, m.esz * m.nt * m.nfslots + sizeof(c64) * m.nt * 10, m.st));
  Ctl c;
  std::memset(&c, 0, sizeof(c));
  c.dt = m.dt_fixed;
  HY_HIP(hipMemcpyAsync(m.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, m.st));
  m.h_dt = m.dt_fixed;
  // state n = n0, mom = 0 (species.f90:162-163)
  double n0v = m.n0;
  HY_HIP(hipMemcpyAsync(m.S, &n0v, sizeof(double), hipMemcpyHostToDevice, m.st));
  // FD table
  if (m.tf || m.grad) {
    const double* tab = fd::host_table();
    HY_TRY(dmalloc((void**)&m.fdtab, sizeof(double) * fd::NTAB * fd::NC));
    HY_HIP(hipMemcpy(m.fdtab, tab, sizeof(double) * fd::NTAB * fd::NC, hipMemcpyHostToDevice));
    m.dadn = 2.0 / (m.cn * fd::im12(tab, m.alpha0));
    m.d2adn2 = -m.dadn * m.dadn * fd::dim12(tab, m.alpha0) / fd::im12(tab, m.alpha0);
    if (mixed(m)) {
      // the reference's evaluation at each node (fermi_dirac.f90:241-267, fe_kinetic_tfk.f90:130-137)
      std::vector<double> rt(4 * (size_t)(RT_N + 1), 0.0);
      const double h = 2.0 * RT_R / RT_N;
      for (int i = 0; i <= RT_N; ++i) {
        const double n = m.n0 * (1.0 + (-RT_R + i * h));
        const double al = fd::alpha_from_ip12(tab, n / m.cn);
        double im, d1, d2;
        fd::im12_d1_d2(tab, al, im, d1, d2);
        rt[4 * i + 0] = al - m.alpha0;
        rt[4 * i + 1] = d1 / (im * im);
        rt[4 * i + 2] = (d2 * im - 2.0 * d1 * d1) / (im * im * im) / im * (2.0 / m.cn);
      }
      HY_TRY(dmalloc((void**)&m.rtab, sizeof(double) * rt.size()));
      HY_HIP(hipMemcpy(m.rtab, rt.data(), sizeof(double) * rt.size(), hipMemcpyHostToDevice));
    }
  }
  HY_TRY(dmalloc((void**)&m.d_xargs, sizeof(PhysArgs) * XA_N));
  if (hipHostMalloc((void**)&m.p_xargs, sizeof(PhysArgs) * XA_N, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    m.p_xargs = nullptr;
    return set_error(HY_ERR_OOM, "model: pinned host allocation failed");
  }
  std::memset(m.p_xargs, 0, sizeof(PhysArgs) * XA_N);
  // combined convolution kernel on the kept box, and compute_dt constants
  std::vector<double> kc(m.nt, 0.0);
  double wmax = 0.0;
  for (size_t e = 0; e < m.nt; ++e) {
    int ix = (int)(e % g.Kx), ty = (int)((e / g.Kx) % g.Ky), tz = (int)(e / ((size_t)g.Kx * g.Ky));
    int jy = ty <= g.My ? ty : ty - g.Ky + g.ny;
    int jz = tz <= g.Mz ? tz : tz - g.Kz + g.nz;
    double k2 = g.kx[ix] * g.kx[ix] + g.ky[jy] * g.ky[jy] + g.kz[jz] * g.kz[jz];
    double kabs = std::sqrt(k2);
    double v = 0.0;
    for (size_t t = 0; t < m.terms.size(); ++t) {
      int kind = m.terms[t].kind;
      if (kind == HY_TERM_HARTREE || kind == HY_TERM_RY) v += term_kernel(m, t, kabs);
    }
    kc[e] = v;
    if (kabs != 0.0) {  // model.f90:159-168
      double c2 = m.n0 * fun_kernel(m, kabs) / m.mass;
      if (m.evolve) c2 = c2 + m.eta_l / (m.mass * m.tau);
      double kk = kabs * kabs * std::max(c2, 0.0);
      wmax = std::max(wmax, std::sqrt(kk));
    }
  }
  m.wmax = wmax;
  m.numax = m.evolve ? 0.0 : m.eta_l / m.mass * g.kmax_dealiased * g.kmax_dealiased;
  m.dxmin = g.dx;
  if (g.ny > 1) m.dxmin = std::min(m.dxmin, g.dy);
  if (g.nz > 1) m.dxmin = std::min(m.dxmin, g.dz);
  HY_TRY(dmalloc((void**)&m.kcomb, sizeof(double) * m.nt));
  HY_HIP(hipMemcpy(m.kcomb, kc.data(), sizeof(double) * m.nt, hipMemcpyHostToDevice));
  // wave force: the x factor of sin(kx x + ky y + kz z - w t) per grid point,
  // as sin(kx x) and cos(kx x) (the x-line kernels add the line's phase)
  if (m.vext_kind == HY_VEXT_MODE || m.vext_kind == HY_VEXT_DRIVEN) {
    std::vector<double> sc(2 * (size_t)g.nx);
    for (int i = 0; i < g.nx; ++i) {
      const double ax = m.kvec[0] * (i * g.dx);
      sc[i] = std::sin(ax);
      sc[g.nx + i] = std::cos(ax);
    }
    HY_TRY(dmalloc((void**)&m.vsc, sizeof(double) * sc.size()));
    HY_HIP(hipMemcpy(m.vsc, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice));
  }
  // static gaussian force -grad(v) (external_potential.f90:84-91)
  if (m.vext_kind == HY_VEXT_GAUSSIAN) {
    std::vector<double> v(g.nreal());
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
          double x = i * g.dx - 0.5 * g.lx;
          double y = g.ny > 1 ? j * g.dy - 0.5 * g.ly : 0.0;
          double z = g.nz > 1 ? k * g.dz - 0.5 * g.lz : 0.0;
          double r2 = x * x + y * y + z * z;
          v[((size_t)k * g.ny + j) * g.nx + i] = -m.vamp * std::exp(-0.5 * r2 / (m.vwidth * m.vwidth));
        }
    double* dv;
    HY_TRY(dmalloc((void**)&m.fext, sizeof(double) * g.nreal() * 3));
    HY_TRY(dmalloc((void**)&dv, sizeof(double) * g.nreal()));
    HY_HIP(hipMemcpy(dv, v.data(), sizeof(double) * g.nreal(), hipMemcpyHostToDevice));
    int rc = spec_grad_api(m.g, dv, m.fext, m.st);
    HY_HIP(hipStreamSynchronize(m.st));
    (void)hipFree(dv);
    HY_TRY(rc);
    std::vector<double> f(g.nreal() * 3);
    HY_HIP(hipMemcpy(f.data(), m.fext, sizeof(double) * f.size(), hipMemcpyDeviceToHost));
    for (auto& x : f) x = -x;
    HY_HIP(hipMemcpy(m.fext, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice));
  }
  HY_HIP(hipStreamSynchronize(m.st));
  return HY_OK;
}
}  // namespace hy

extern "C" {

int hy_model_create(const hy_model_desc* d, hy_model* out) {
  if (!d || !out) return set_error(HY_ERR_INVALID, "hy_model_create: null argument");
  auto* h = new hy_model_s();
  int rc = model_build(h->m, d);
  if (rc != HY_OK) {
    model_free(h->m);
    delete h;
    return rc;
  }
  *out = h;
  return HY_OK;
}

int hy_model_destroy(hy_model h) {
  if (!h) return HY_OK;
  (void)hipStreamSynchronize(h->m.st);
  model_free(h->m);
  delete h;
  return HY_OK;
}

static int transfer_bufs(Model& m, double** dr, c64** df) {
  if (!m.d_real) HY_TRY(dmalloc((void**)&m.d_real, sizeof(double) * m.g.nreal()));
  if (!m.d_full) HY_TRY(dmalloc((void**)&m.d_full, sizeof(c64) * m.g.nfull()));
  *dr = m.d_real;
  *df = m.d_full;
  return HY_OK;
}

// Real fields -> kept-box spectra, in FP64 (the state is FP64 in every mode).
int hy_model_set_state(hy_model h, const double* h_n, const double* h_mom3, const double* h_pi6) {
  if (!h || !h_n) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  Grid& g = m.g;
  const size_t nr = g.nreal();
  double* dr;
  c64* df;
  int rc = HY_OK;
  HY_TRY(transfer_bufs(m, &dr, &df));
  SpecArgs sa = spec_args(m);
  auto put = [&](const double* src, int f) -> int {
    HY_HIP(hipMemcpyAsync(dr, src, sizeof(double) * nr, hipMemcpyHostToDevice, m.st));
    HY_TRY(fft_r2c_full_api(g, dr, df, m.st));
    hipLaunchKernelGGL(k_full_to_trunc, dim3(nblocks(m.nt)), dim3(256), 0, m.st, df, sfield(m, f), sa, g.nkx);
    HY_HIP(hipGetLastError());
    return HY_OK;
  };
  rc = put(h_n, 0);
  for (int d = 0; d < 3 && rc == HY_OK; ++d) {
    if (h_mom3) rc = put(h_mom3 + nr * d, 1 + d);
    else rc = hipMemsetAsync(sfield(m, 1 + d), 0, sizeof(c64) * m.nt, m.st) == hipSuccess ? HY_OK : HY_ERR_HIP;
  }
  if (m.evolve)
    for (int c = 0; c < 6 && rc == HY_OK; ++c) {
      if (h_pi6) rc = put(h_pi6 + nr * c, 4 + c);
      else rc = hipMemsetAsync(sfield(m, 4 + c), 0, sizeof(c64) * m.nt, m.st) == hipSuccess ? HY_OK : HY_ERR_HIP;
    }
  // a new state starts clean: the non-positive-density and no-time-step
  // conditions seen on the previous state are cleared
  const int clear2[2] = {0, 0};
  static_assert(offsetof(Ctl, nodt) == offsetof(Ctl, nonpos) + sizeof(int), "Ctl: nonpos, nodt adjacent");
  if (rc == HY_OK && hipMemcpyAsync(&m.ctl->nonpos, clear2, sizeof(clear2), hipMemcpyHostToDevice, m.st) != hipSuccess)
    rc = set_error(HY_ERR_HIP, "hy_model_set_state: control block");
  (void)hipStreamSynchronize(m.st);
  m.relax_pending = 0;
  return rc;
}

static int trunc_to_real(Model& m, const c64* tr, double* h_out) {
  Grid& g = m.g;
  const size_t nr = g.nreal();
  double* dr;
  c64* df;
  int rc = HY_OK;
  HY_TRY(transfer_bufs(m, &dr, &df));
  SpecArgs sa = spec_args(m);
  hipLaunchKernelGGL(k_trunc_to_full, dim3(nblocks(g.nfull())), dim3(256), 0, m.st, tr, df, sa, g.nkx);
  rc = fft_c2r_full_api(g, df, dr, m.st);
  if (rc == HY_OK && hipMemcpyAsync(h_out, dr, sizeof(double) * nr, hipMemcpyDeviceToHost, m.st) != hipSuccess)
    rc = set_error(HY_ERR_HIP, "copy");
  if (hipStreamSynchronize(m.st) != hipSuccess) rc = set_error(HY_ERR_HIP, "sync");
  return rc;
}

int hy_model_get_state(hy_model h, double* h_n, double* h_mom3, double* h_pi6) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  const size_t nr = m.g.nreal();
  HY_TRY(finish_relax(m));
  if (h_n) HY_TRY(trunc_to_real(m, sfield(m, 0), h_n));
  if (h_mom3)
    for (int d = 0; d < 3; ++d) HY_TRY(trunc_to_real(m, sfield(m, 1 + d), h_mom3 + nr * d));
  if (h_pi6) {
    if (!m.evolve) return set_error(HY_ERR_INVALID, "get_state: no stress state (newtonian closure)");
    for (int c = 0; c < 6; ++c) HY_TRY(trunc_to_real(m, sfield(m, 4 + c), h_pi6 + nr * c));
  }
  return HY_OK;
}

int hy_model_set_time(hy_model h, double t, int step) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  Ctl c;
  HY_TRY(read_ctl(m, c));
  c.t = t; c.ts = t; c.step = step;
  HY_HIP(hipMemcpyAsync(m.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, m.st));
  HY_HIP(hipStreamSynchronize(m.st));
  m.h_t = t; m.h_step = step;
  return HY_OK;
}

int hy_model_get_time(hy_model h, double* t, double* dt, int* step) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Ctl c;
  HY_TRY(read_ctl(h->m, c));
  if (t) *t = c.t;
  if (dt) *dt = c.dt;
  if (step) *step = c.step;
  return HY_OK;
}

int hy_model_set_dt(hy_model h, double dt) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  const int clear = 0;  // a given step clears an earlier no-time-step condition
  HY_HIP(hipMemcpyAsync(&m.ctl->dt, &dt, sizeof(double), hipMemcpyHostToDevice, m.st));
  HY_HIP(hipMemcpyAsync(&m.ctl->nodt, &clear, sizeof(int), hipMemcpyHostToDevice, m.st));
  HY_HIP(hipStreamSynchronize(m.st));
  m.h_dt = dt;
  return HY_OK;
}

int hy_model_compute_dt(hy_model h, double* h_dt) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  HY_TRY(ctl_op(m, CTL_OPEN, 0.0));
  HY_TRY(finish_relax(m));
  HY_TRY(compute_dt_dev(m));
  Ctl c;
  HY_TRY(read_ctl(m, c));
  if (c.nodt) return set_error(HY_ERR_NO_DT, "cannot determine a time step; set dt in &time");
  if (h_dt) *h_dt = c.dt;
  return HY_OK;
}

int hy_model_advance(hy_model h, int nsteps, int dt_every, double tend, const int* probe3,
                     double* h_probe, int* h_taken) {
  if (!h || nsteps < 0) return set_error(HY_ERR_INVALID, "hy_model_advance: bad argument");
  Model& m = h->m;
  HY_TRY(ctl_op(m, CTL_OPEN, 0.0));
  const Grid& g = m.g;
  long long pidx = -1;
  if (probe3) {
    if (!h_probe) return set_error(HY_ERR_INVALID, "probe buffer missing");
    int mx = probe3[0], my = probe3[1], mz = probe3[2];
    if (mx < 0 || mx >= g.nkx) return set_error(HY_ERR_INVALID, "probe: mx out of range");
    int wy = ((my % g.ny) + g.ny) % g.ny, wz = ((mz % g.nz) + g.nz) % g.nz;
    int sy = wy <= g.ny / 2 ? wy : wy - g.ny, sz = wz <= g.nz / 2 ? wz : wz - g.nz;
    if (mx < g.Kx && std::abs(sy) <= g.My && std::abs(sz) <= g.Mz) {
      int ty = sy >= 0 ? sy : sy + g.Ky, tz = sz >= 0 ? sz : sz + g.Kz;
      pidx = ((long long)tz * g.Ky + ty) * g.Kx + mx;
    }
    if (m.aprobe_cap < nsteps) {
      if (m.aprobe) (void)hipFree(m.aprobe);
      m.aprobe = nullptr;
      m.aprobe_cap = 0;
      HY_TRY(dmalloc((void**)&m.aprobe, sizeof(double) * 6 * std::max(nsteps, 1)));
      m.aprobe_cap = nsteps;
    }
  }
  // tend > 0: stop at tend and clamp the last step (src/hydft.f90:41,44);
  // tend < 0: stop once t >= |tend| without clamping (the test loops).
  const double tstop = std::fabs(tend);
  const double tclamp = tend > 0.0 ? tend : 0.0;
  int taken = 0;
  for (int i = 0; i < nsteps; ++i) {
    if (tend != 0.0) {
      Ctl c;
      HY_TRY(read_ctl(m, c));
      if (c.t >= tstop) break;
    }
    if (dt_every > 0) {
      if (m.irecompute == 0) {
        HY_TRY(finish_relax(m));
        HY_TRY(compute_dt_dev(m));
        if (tend != 0.0 || i == 0) {
          Ctl c;
          HY_TRY(read_ctl(m, c));
          if (c.nodt) return set_error(HY_ERR_NO_DT, "cannot determine a time step; set dt in &time");
        }
      }
      m.irecompute = (m.irecompute + 1) % dt_every;
    }
    HY_TRY(step_once(m, tclamp));
    if (probe3) hipLaunchKernelGGL(k_probe, dim3(1), dim3(1), 0, m.st, m.ctl, m.S, m.nt, pidx, m.n0, m.aprobe + 6 * (size_t)taken);
    ++taken;
  }
  HY_TRY(finish_relax(m));
  Ctl c;
  HY_TRY(read_ctl(m, c));
  if (probe3 && taken > 0) {
    HY_HIP(hipMemcpy(h_probe, m.aprobe, sizeof(double) * 6 * taken, hipMemcpyDeviceToHost));
  }
  if (h_taken) *h_taken = taken;
  if (c.nonpos) return set_error(HY_ERR_NONPOSITIVE, "density became non-positive; reduce the amplitude or the time step");
  if (c.nodt) return set_error(HY_ERR_NO_DT, "cannot determine a time step; set dt in &time");
  return HY_OK;
}

// ---- captured steps
// Probe row index of mode (mx, my, mz) in the kept box, or -1 (not kept).
static int probe_index(const Model& m, const int* probe3, long long* pidx) {
  const Grid& g = m.g;
  *pidx = -1;
  if (!probe3) return HY_OK;
  int mx = probe3[0], my = probe3[1], mz = probe3[2];
  if (mx < 0 || mx >= g.nkx) return set_error(HY_ERR_INVALID, "probe: mx out of range");
  int wy = ((my % g.ny) + g.ny) % g.ny, wz = ((mz % g.nz) + g.nz) % g.nz;
  int sy = wy <= g.ny / 2 ? wy : wy - g.ny, sz = wz <= g.nz / 2 ? wz : wz - g.nz;
  if (mx < g.Kx && std::abs(sy) <= g.My && std::abs(sz) <= g.Mz) {
    int ty = sy >= 0 ? sy : sy + g.Ky, tz = sz >= 0 ? sz : sz + g.Kz;
    *pidx = ((long long)tz * g.Ky + ty) * g.Kx + mx;
  }
  return HY_OK;
}

// Make the model ready to emit a replayable sequence of nsteps steps: probe
// ring allocated, no pending relaxation, start of the dt cadence, stream idle.
static int prepare_steps(Model& m, int nsteps, int dt_every, const int* probe3, int ring) {
  if (nsteps < 1 || dt_every < 0 || (probe3 && ring < nsteps)) return set_error(HY_ERR_INVALID, "bad step sequence");
  if (dt_every > 0 && nsteps % dt_every != 0)
    return set_error(HY_ERR_INVALID, "nsteps must be a multiple of dt_every");
  if (probe3 && m.probe_cap < ring) {
    if (m.nbatch > 0)
      return set_error(HY_ERR_INVALID, "the model's probe ring is held by a batch; destroy the batch or ask for "
                                       "a ring no longer than " + std::to_string(m.probe_cap));
    HY_HIP(hipStreamSynchronize(m.st));
    // the captured graph writes into the old ring: it goes with it
    if (m.gexec) { (void)hipGraphExecDestroy(m.gexec); m.gexec = nullptr; }
    if (m.graph) { (void)hipGraphDestroy(m.graph); m.graph = nullptr; }
    m.g_probe = 0;
    m.g_ring = 0;
    if (m.probe) (void)hipFree(m.probe);
    m.probe = nullptr;
    m.probe_cap = 0;
    HY_TRY(dmalloc((void**)&m.probe, sizeof(double) * 6 * ring));
    m.probe_cap = ring;
  }
  HY_TRY(finish_relax(m));
  if (dt_every > 0) m.irecompute = 0;
  HY_HIP(hipStreamSynchronize(m.st));
  return HY_OK;
}

// The launches of nsteps steps as hy_model_advance(nsteps, dt_every, tend=0)
// takes them, each step followed by its probe row into the ring; ends with
// no pending relaxation and at the start of the dt cadence again.
static int emit_steps(Model& m, int nsteps, int dt_every, long long pidx, bool probe, int ring) {
  HY_TRY(ctl_op(m, CTL_SEQ_BEGIN, 0.0));
  for (int i = 0; i < nsteps; ++i) {
    if (dt_every > 0) {
      if (m.irecompute == 0) {
        HY_TRY(finish_relax(m));
        HY_TRY(compute_dt_dev(m));
      }
      m.irecompute = (m.irecompute + 1) % dt_every;
    }
    HY_TRY(step_once(m, 0.0));
    if (probe)
      HY_TRY(launch(HY_FN(k_probe_ring), HY_FN(k_probe_ring_b), dim3(1), dim3(1), m.st,
                    ProbeK{m.ctl, m.S, m.nt, pidx, m.n0, m.probe, ring}));
  }
  return finish_relax(m);
}

int hy_model_graph_build(hy_model h, int nsteps, int dt_every, const int* probe3, int ring) {
  if (!h) return set_error(HY_ERR_INVALID, "hy_model_graph_build: null argument");
  Model& m = h->m;
  long long pidx = -1;
  HY_TRY(probe_index(m, probe3, &pidx));
  HY_TRY(prepare_steps(m, nsteps, dt_every, probe3, ring));
  if (m.gexec) { (void)hipGraphExecDestroy(m.gexec); m.gexec = nullptr; }
  if (m.graph) { (void)hipGraphDestroy(m.graph); m.graph = nullptr; }
  const unsigned prof = m.prof;
  m.prof = 0;
  HY_HIP(hipStreamBeginCapture(m.st, hipStreamCaptureModeThreadLocal));
  const int rc = emit_steps(m, nsteps, dt_every, pidx, probe3 != nullptr, ring);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(m.st, &graph);
  m.prof = prof;
  // The argument-block uploads made during the capture are copy nodes of the
  // graph: the device slots hold them only once the graph has run.  Upload
  // the staging now, so that eager steps before the first replay read the
  // same blocks (the staging is what the graph copies as well).
  if (ec == hipSuccess) {
    HY_HIP(hipMemcpyAsync(m.d_xargs, m.p_xargs, sizeof(PhysArgs) * XA_N, hipMemcpyHostToDevice, m.st));
    HY_HIP(hipStreamSynchronize(m.st));
  }
  if (rc != HY_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess) return set_error(HY_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  m.graph = graph;
  HY_HIP(hipGraphInstantiate(&m.gexec, m.graph, nullptr, nullptr, 0));
  m.g_nsteps = nsteps;
  m.g_probe = probe3 ? 1 : 0;
  m.g_ring = probe3 ? ring : 0;
  return HY_OK;
}

int hy_model_graph_launch(hy_model h, int nlaunch) {
  if (!h || nlaunch < 0) return set_error(HY_ERR_INVALID, "hy_model_graph_launch: bad argument");
  Model& m = h->m;
  if (!m.gexec) return set_error(HY_ERR_INVALID, "hy_model_graph_launch: no graph (hy_model_graph_build)");
  for (int i = 0; i < nlaunch; ++i) HY_HIP(hipGraphLaunch(m.gexec, m.st));
  return HY_OK;
}

int hy_model_graph_probe(hy_model h, double* h_probe) {
  if (!h || !h_probe) return set_error(HY_ERR_INVALID, "hy_model_graph_probe: null argument");
  Model& m = h->m;
  if (!m.gexec || !m.g_probe) return set_error(HY_ERR_INVALID, "hy_model_graph_probe: graph has no probe");
  HY_HIP(hipStreamSynchronize(m.st));
  HY_HIP(hipMemcpy(h_probe, m.probe, sizeof(double) * 6 * m.g_ring, hipMemcpyDeviceToHost));
  return HY_OK;
}

int hy_model_idle(hy_model h, int* idle) {
  if (!h || !idle) return set_error(HY_ERR_INVALID, "hy_model_idle: null argument");
  const hipError_t e = hipStreamQuery(h->m.st);
  if (e == hipSuccess) { *idle = 1; return HY_OK; }
  if (e == hipErrorNotReady) { (void)hipGetLastError(); *idle = 0; return HY_OK; }
  return set_error(HY_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int hy_model_check(hy_model h) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Ctl c;
  HY_TRY(read_ctl(h->m, c));
  if (c.nonpos) return set_error(HY_ERR_NONPOSITIVE, "density became non-positive; reduce the amplitude or the time step");
  if (c.nodt) return set_error(HY_ERR_NO_DT, "cannot determine a time step; set dt in &time");
  return HY_OK;
}

int hy_model_device_bytes(hy_model h, unsigned long long* bytes) {
  if (!h || !bytes) return set_error(HY_ERR_INVALID, "null argument");
  const Model& m = h->m;
  const Grid& g = m.g;
  double b = sizeof(c64) * (double)m.nt * (10 + 10 + 3 + 10);             // S, R, u_old, rhs slots
  b += (double)m.esz * ((double)m.xsz * m.nxslots + (double)m.zsz * m.nzslots + (double)m.nt * m.nfslots);
  b += sizeof(double) * ((double)m.nt + std::max<double>((double)g.ny * g.nz * 4, 4096));  // kcomb, reductions
  if (m.fext) b += sizeof(double) * 3.0 * g.nreal();
  if (m.vsc) b += sizeof(double) * 2.0 * g.nx;
  if (m.rtab) b += sizeof(double) * 4.0 * (RT_N + 1);
  if (m.fdtab) b += sizeof(double) * (double)fd::NTAB * fd::NC;
  b += sizeof(double) * 6.0 * (m.probe_cap + m.aprobe_cap) + sizeof(PhysArgs) * XA_N + sizeof(Ctl);
  b += (double)g.scratch_bytes;
  if (m.d_real) b += sizeof(double) * (double)g.nreal();
  if (m.d_full) b += sizeof(c64) * (double)g.nfull();
  for (const auto& w : m.work) b += (double)w.bytes;
  if (m.full_kh) b += 2.0 * sizeof(double) * (double)g.nfull();
  *bytes = (unsigned long long)b;
  return HY_OK;
}

// ---- batches: one launch per kernel for many models of the same shape
struct hy_batch_s {
  std::vector<hy_model> models;
  hipStream_t st = nullptr;
  struct Item {
    const void* fn;
    dim3 grid, block;
    void* d_args;      // B argument structs, or B argument blocks (x-line kernels)
    bool xline;
    const void* tw;
  };
  std::vector<Item> items;
  void* d_all = nullptr;
  bool holds = false;  // the models count this batch in their nbatch
  bool stops_set = false;
  int nsteps = 0;      // steps per launch
};

static void batch_free(hy_batch_s* b) {
  if (!b) return;
  if (b->holds) {
    // the models run on alone afterwards: no step limit
    if (b->stops_set)
      for (hy_model h : b->models) hipLaunchKernelGGL(k_set_stop, dim3(1), dim3(1), 0, b->st, h->m.ctl, 0);
    for (hy_model h : b->models) --h->m.nbatch;
  }
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->d_all) (void)hipFree(b->d_all);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

int hy_batch_create(const hy_model* ms, int n, int nsteps, int dt_every, const int* probe3, int ring,
                    hy_batch* out) {
  if (!ms || !out || n < 1) return set_error(HY_ERR_INVALID, "hy_batch_create: bad argument");
  int maxz = 0;
  HY_HIP(hipDeviceGetAttribute(&maxz, hipDeviceAttributeMaxGridDimZ, 0));
  if (n > maxz) return set_error(HY_ERR_INVALID, "hy_batch_create: too many models for one launch");
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (ms[i] == ms[j]) return set_error(HY_ERR_INVALID, "hy_batch_create: a model appears twice");
  std::vector<Recorder> recs(n);
  for (int i = 0; i < n; ++i) {
    if (!ms[i]) return set_error(HY_ERR_INVALID, "hy_batch_create: null model");
    Model& m = ms[i]->m;
    long long pidx = -1;
    HY_TRY(probe_index(m, probe3, &pidx));
    HY_TRY(prepare_steps(m, nsteps, dt_every, probe3, ring));
    const unsigned prof = m.prof;
    m.prof = 0;
    recorder() = &recs[i];
    const int rc = emit_steps(m, nsteps, dt_every, pidx, probe3 != nullptr, ring);
    recorder() = nullptr;
    m.prof = prof;
    HY_TRY(rc);
    HY_HIP(hipStreamSynchronize(m.st));  // argument-block uploads done during the recording
  }
  // every model must have produced the same sequence of launches
  const size_t L = recs[0].items.size();
  for (int i = 1; i < n; ++i) {
    if (recs[i].items.size() != L)
      return set_error(HY_ERR_INVALID, "hy_batch_create: models differ in shape or options (model " +
                                           std::to_string(i) + ": " + std::to_string(recs[i].items.size()) +
                                           " launches, model 0: " + std::to_string(L) + ")");
    for (size_t p = 0; p < L; ++p) {
      const LaunchRec& a = recs[0].items[p];
      const LaunchRec& c = recs[i].items[p];
      if (a.fn_batched != c.fn_batched || a.grid.x != c.grid.x || a.grid.y != c.grid.y || a.block.x != c.block.x ||
          a.xline != c.xline || a.args.size() != c.args.size())
        return set_error(HY_ERR_INVALID, "hy_batch_create: models differ in shape or options (model " +
                                             std::to_string(i) + ", launch " + std::to_string(p) + ": " +
                                             (a.fn_batched != c.fn_batched ? "kernel" : "grid or arguments") +
                                             ", grid " + std::to_string(a.grid.x) + "x" + std::to_string(a.grid.y) +
                                             " vs " + std::to_string(c.grid.x) + "x" + std::to_string(c.grid.y) +
                                             ", block " + std::to_string(a.block.x) + " vs " +
                                             std::to_string(c.block.x) + ", args " + std::to_string(a.args.size()) +
                                             " vs " + std::to_string(c.args.size()) + ")");
    }
  }
  // one device allocation: per launch position, n argument structs (or blocks)
  auto* b = new hy_batch_s();
  b->models.assign(ms, ms + n);
  std::vector<size_t> off(L);
  size_t total = 0;
  for (size_t p = 0; p < L; ++p) {
    off[p] = total;
    const size_t each = recs[0].items[p].xline ? sizeof(PhysArgs) : recs[0].items[p].args.size();
    total += (each * n + 255) & ~(size_t)255;
  }
  std::vector<unsigned char> host(total, 0);
  for (size_t p = 0; p < L; ++p) {
    const bool xl = recs[0].items[p].xline;
    for (int i = 0; i < n; ++i) {
      const LaunchRec& it = recs[i].items[p];
      if (xl) {
        const Model& m = ms[i]->m;
        const PhysArgs* blk = reinterpret_cast<const PhysArgs*>(it.xblock);
        const long slot = (long)(blk - m.d_xargs);
        if (slot < 0 || slot >= XA_N) {
          batch_free(b);
          return set_error(HY_ERR_INVALID, "hy_batch_create: unknown x-kernel argument block");
        }
        std::memcpy(&host[off[p] + sizeof(PhysArgs) * i], &m.p_xargs[slot], sizeof(PhysArgs));
      } else {
        std::memcpy(&host[off[p] + it.args.size() * i], it.args.data(), it.args.size());
      }
    }
  }
  if (hipMalloc(&b->d_all, std::max<size_t>(total, 256)) != hipSuccess) {
    (void)hipGetLastError();
    b->d_all = nullptr;
    batch_free(b);
    return set_error(HY_ERR_OOM, "hy_batch_create: device allocation failed");
  }
  if (hipMemcpy(b->d_all, host.data(), total, hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
    batch_free(b);
    return set_error(HY_ERR_HIP, "hy_batch_create: upload failed");
  }
  for (size_t p = 0; p < L; ++p) {
    const LaunchRec& it = recs[0].items[p];
    // the twiddle table of a line length is the same in every model of the batch
    b->items.push_back({it.fn_batched, dim3(it.grid.x, it.grid.y, (unsigned)n), it.block,
                        (char*)b->d_all + off[p], it.xline, it.xtw});
    if (it.pencil[0]) {  // pencil passes: the plan for this batch size
      auto& bi = b->items.back();
      const int rc = pencil_batch_plan(it.pencil, n, bi.d_args, b->st, &bi.fn, &bi.grid, &bi.block);
      if (rc != HY_OK) {
        batch_free(b);
        return rc;
      }
    }
  }
  for (hy_model h : b->models) ++h->m.nbatch;
  b->holds = true;
  b->nsteps = nsteps;
  *out = b;
  return HY_OK;
}

// One pass of a batch's recorded launches on stream st.
static int batch_enqueue(hy_batch b, hipStream_t st) {
  for (auto& it : b->items) {
    hipError_t e;
    if (it.xline) {
      void* args[] = {&it.d_args, const_cast<void**>(&it.tw)};
      e = hipLaunchKernel(it.fn, it.grid, it.block, args, 0, st);
    } else {
      void* args[] = {&it.d_args};
      e = hipLaunchKernel(it.fn, it.grid, it.block, args, 0, st);
    }
    if (e != hipSuccess) return set_error(HY_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
  }
  return HY_OK;
}

int hy_batch_launch(hy_batch b, int nlaunch) {
  if (!b || nlaunch < 0) return set_error(HY_ERR_INVALID, "hy_batch_launch: bad argument");
  for (hy_model h : b->models) HY_HIP(hipStreamSynchronize(h->m.st));  // the models' own queued work first
  for (int l = 0; l < nlaunch; ++l) HY_TRY(batch_enqueue(b, b->st));
  return HY_OK;
}

// ---- groups of batches: runs on uneven grids advanced by one graph launch
struct hy_group_s {
  std::vector<hy_batch> batches;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<hipEvent_t> events;
  hipEvent_t done = nullptr;  // recorded after each hy_group_launch
};

static void group_free(hy_group_s* g) {
  if (!g) return;
  if (g->st) (void)hipStreamSynchronize(g->st);
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (auto e : g->events) (void)hipEventDestroy(e);
  if (g->done) (void)hipEventDestroy(g->done);
  if (g->st) (void)hipStreamDestroy(g->st);
  delete g;
}

int hy_group_create(const hy_batch* batches, int n, hy_group* out) {
  if (!batches || n < 1 || !out) return set_error(HY_ERR_INVALID, "hy_group_create: bad argument");
  for (int i = 0; i < n; ++i) {
    if (!batches[i]) return set_error(HY_ERR_INVALID, "hy_group_create: null batch");
    for (int j = 0; j < i; ++j) {
      if (batches[i] == batches[j]) return set_error(HY_ERR_INVALID, "hy_group_create: a batch appears twice");
      // the branches run concurrently: a model in two of them would race
      for (hy_model a : batches[i]->models)
        for (hy_model c : batches[j]->models)
          if (a == c) return set_error(HY_ERR_INVALID, "hy_group_create: a model is in two batches");
    }
  }
  auto* g = new hy_group_s();
  g->batches.assign(batches, batches + n);
  auto fail = [&](int rc) {
    group_free(g);
    return rc;
  };
  if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(HY_ERR_HIP, "hy_group_create: stream"));
  for (hy_batch b : g->batches)
    if (hipStreamSynchronize(b->st) != hipSuccess) return fail(set_error(HY_ERR_HIP, "hy_group_create: sync"));
  if (hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess)
    return fail(set_error(HY_ERR_HIP, "hy_group_create: event"));
  g->events.resize(1 + n, nullptr);
  for (auto& e : g->events)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(set_error(HY_ERR_HIP, "hy_group_create: event"));
  // capture: fork one branch per batch (its own stream joins the capture
  // through an event), each branch one pass of the batch's launches, join
  if (hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(set_error(HY_ERR_HIP, "hy_group_create: begin capture"));
  int rc = HY_OK;
  if (hipEventRecord(g->events[0], g->st) != hipSuccess) rc = set_error(HY_ERR_HIP, "hy_group_create: fork");
  for (int i = 0; i < n && rc == HY_OK; ++i) {
    hy_batch b = g->batches[i];
    if (hipStreamWaitEvent(b->st, g->events[0], 0) != hipSuccess) { rc = set_error(HY_ERR_HIP, "hy_group_create: branch"); break; }
    rc = batch_enqueue(b, b->st);
    if (rc == HY_OK && (hipEventRecord(g->events[1 + i], b->st) != hipSuccess ||
                        hipStreamWaitEvent(g->st, g->events[1 + i], 0) != hipSuccess))
      rc = set_error(HY_ERR_HIP, "hy_group_create: join");
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(g->st, &graph);
  if (rc != HY_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(rc);
  }
  if (ec != hipSuccess) return fail(set_error(HY_ERR_HIP, std::string("hy_group_create: end capture: ") + hipGetErrorString(ec)));
  g->graph = graph;
  if (hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(set_error(HY_ERR_HIP, "hy_group_create: instantiate"));
  *out = g;
  return HY_OK;
}

int hy_group_launch(hy_group g, int nlaunch) {
  if (!g || nlaunch < 0) return set_error(HY_ERR_INVALID, "hy_group_launch: bad argument");
  for (hy_batch b : g->batches) {
    HY_HIP(hipStreamSynchronize(b->st));
    for (hy_model h : b->models) HY_HIP(hipStreamSynchronize(h->m.st));
  }
  for (int l = 0; l < nlaunch; ++l) HY_HIP(hipGraphLaunch(g->gexec, g->st));
  HY_HIP(hipEventRecord(g->done, g->st));
  return HY_OK;
}

int hy_group_idle(hy_group g, int* idle) {
  if (!g || !idle) return set_error(HY_ERR_INVALID, "hy_group_idle: null argument");
  // the event after the last replays, not the stream: a stream that has run
  // graph launches can read as busy after it drained
  const hipError_t e = hipEventQuery(g->done);
  if (e == hipSuccess) { *idle = 1; return HY_OK; }
  if (e == hipErrorNotReady) { (void)hipGetLastError(); *idle = 0; return HY_OK; }
  return set_error(HY_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int hy_group_sync(hy_group g) {
  if (!g) return set_error(HY_ERR_INVALID, "null argument");
  HY_HIP(hipStreamSynchronize(g->st));
  return HY_OK;
}

int hy_group_destroy(hy_group g) {
  group_free(g);
  return HY_OK;
}

int hy_batch_set_stops(hy_batch b, const int* stops) {
  if (!b || !stops) return set_error(HY_ERR_INVALID, "hy_batch_set_stops: null argument");
  for (size_t i = 0; i < b->models.size(); ++i)
    if (stops[i] < 0 || stops[i] % b->nsteps != 0)
      return set_error(HY_ERR_INVALID, "hy_batch_set_stops: stop " + std::to_string(stops[i]) + " of model " +
                                           std::to_string(i) + " is not a non-negative multiple of the batch's " +
                                           std::to_string(b->nsteps) + " steps per launch");
  // stream ordered after the launches already queued
  for (size_t i = 0; i < b->models.size(); ++i)
    hipLaunchKernelGGL(k_set_stop, dim3(1), dim3(1), 0, b->st, b->models[i]->m.ctl, stops[i]);
  HY_HIP(hipGetLastError());
  b->stops_set = true;
  return HY_OK;
}

int hy_batch_idle(hy_batch b, int* idle) {
  if (!b || !idle) return set_error(HY_ERR_INVALID, "hy_batch_idle: null argument");
  const hipError_t e = hipStreamQuery(b->st);
  if (e == hipSuccess) { *idle = 1; return HY_OK; }
  if (e == hipErrorNotReady) { (void)hipGetLastError(); *idle = 0; return HY_OK; }
  return set_error(HY_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int hy_batch_sync(hy_batch b) {
  if (!b) return set_error(HY_ERR_INVALID, "null argument");
  HY_HIP(hipStreamSynchronize(b->st));
  return HY_OK;
}

int hy_batch_destroy(hy_batch b) {
  batch_free(b);
  return HY_OK;
}

int hy_model_kept(hy_model h, int* kept3) {
  if (!h || !kept3) return set_error(HY_ERR_INVALID, "hy_model_kept: null argument");
  kept3[0] = h->m.g.Kx;
  kept3[1] = h->m.g.Ky;
  kept3[2] = h->m.g.Kz;
  return HY_OK;
}

int hy_model_probe_ring(hy_model h, int ring, double* h_probe) {
  if (!h || !h_probe || ring < 1) return set_error(HY_ERR_INVALID, "hy_model_probe_ring: bad argument");
  Model& m = h->m;
  if (ring > m.probe_cap) return set_error(HY_ERR_INVALID, "hy_model_probe_ring: ring larger than the model's");
  HY_HIP(hipStreamSynchronize(m.st));
  HY_HIP(hipMemcpy(h_probe, m.probe, sizeof(double) * 6 * ring, hipMemcpyDeviceToHost));
  return HY_OK;
}

int hy_model_rhs(hy_model h, double t, double* h_dn, double* h_dmom3, double* h_dpi6) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  HY_TRY(ctl_op(m, CTL_OPEN, 0.0));
  HY_TRY(finish_relax(m));
  Ctl c;
  HY_TRY(read_ctl(m, c));
  c.ts = t;
  HY_HIP(hipMemcpyAsync(m.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, m.st));
  HY_TRY(stage(m, 0, 1));
  const c64* R = reinterpret_cast<const c64*>(fslot(m, FS_RHS));
  const size_t nr = m.g.nreal();
  if (h_dn) HY_TRY(trunc_to_real(m, R, h_dn));
  if (h_dmom3)
    for (int d = 0; d < 3; ++d) HY_TRY(trunc_to_real(m, R + m.nt * (1 + d), h_dmom3 + nr * d));
  if (h_dpi6 && m.evolve)
    for (int cc = 0; cc < 6; ++cc) HY_TRY(trunc_to_real(m, R + m.nt * (4 + cc), h_dpi6 + nr * cc));
  Ctl c2;
  HY_TRY(read_ctl(m, c2));
  if (c2.nonpos) return set_error(HY_ERR_NONPOSITIVE, "density became non-positive");
  return HY_OK;
}

int hy_model_probe(hy_model h, int mx, int my, int mz, double* out4) {
  if (!h || !out4) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  const Grid& g = m.g;
  HY_TRY(finish_relax(m));
  if (mx < 0 || mx >= g.nkx) return set_error(HY_ERR_INVALID, "probe: mx out of range");
  int wy = ((my % g.ny) + g.ny) % g.ny, wz = ((mz % g.nz) + g.nz) % g.nz;
  int sy = wy <= g.ny / 2 ? wy : wy - g.ny, sz = wz <= g.nz / 2 ? wz : wz - g.nz;
  c64 a = mk<double>(0, 0), b = mk<double>(0, 0);
  if (mx < g.Kx && std::abs(sy) <= g.My && std::abs(sz) <= g.Mz) {
    int ty = sy >= 0 ? sy : sy + g.Ky, tz = sz >= 0 ? sz : sz + g.Kz;
    size_t idx = ((size_t)tz * g.Ky + ty) * g.Kx + mx;
    HY_HIP(hipMemcpy(&a, m.S + idx, sizeof(c64), hipMemcpyDeviceToHost));
    HY_HIP(hipMemcpy(&b, m.S + m.nt + idx, sizeof(c64), hipMemcpyDeviceToHost));
  }
  out4[0] = a.x / m.n0; out4[1] = a.y / m.n0; out4[2] = b.x; out4[3] = b.y;
  return HY_OK;
}


int hy_model_mu(hy_model h, const double* h_n, double* h_mu) {
  if (!h || !h_n || !h_mu) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  const size_t nr = m.g.nreal();
  DevBufs B(m);
  double *dn, *out;
  int* flag;
  HY_TRY(B.get(&dn, nr));
  HY_TRY(B.get(&out, nr));
  HY_TRY(B.get(&flag, 1));
  HY_HIP(hipMemsetAsync(flag, 0, sizeof(int), m.st));
  HY_HIP(hipMemcpyAsync(dn, h_n, sizeof(double) * nr, hipMemcpyHostToDevice, m.st));
  c64* N;
  double *conv, *wry, *gn, *lap;
  HY_TRY(real_terms(m, B, dn, false, &N, &conv, &wry, &gn, &lap));
  PointArgs a = point_args(m);
  a.n = dn; a.conv = conv; a.gn = gn; a.lap = lap; a.out = out; a.flag = flag;
  hipLaunchKernelGGL(k_point_mu, dim3(nblocks(nr)), dim3(256), 0, m.st, a);
  HY_HIP(hipGetLastError());
  int bad = 0;
  HY_HIP(hipMemcpyAsync(h_mu, out, sizeof(double) * nr, hipMemcpyDeviceToHost, m.st));
  HY_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, m.st));
  HY_HIP(hipStreamSynchronize(m.st));
  if (bad) return set_error(HY_ERR_NONPOSITIVE, "density is not positive");
  return HY_OK;
}

int hy_model_budgets(hy_model h, double* out6) {
  if (!h || !out6) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  const Grid& g = m.g;
  HY_TRY(finish_relax(m));
  const size_t nr = g.nreal();
  DevBufs B(m);
  double *dn, *dmom, *dn0;
  c64* F;
  int* flag;
  HY_TRY(B.get(&dn, nr));
  HY_TRY(B.get(&dmom, nr * 3));
  HY_TRY(B.get(&F, g.nfull()));
  HY_TRY(B.get(&flag, 1));
  HY_HIP(hipMemsetAsync(flag, 0, sizeof(int), m.st));
  SpecArgs sa = spec_args(m);
  for (int f = 0; f < 4; ++f) {
    hipLaunchKernelGGL(k_trunc_to_full, dim3(nblocks(g.nfull())), dim3(256), 0, m.st, sfield(m, f), F, sa, g.nkx);
    HY_TRY(fft_c2r_full_api(m.g, F, f == 0 ? dn : dmom + nr * (f - 1), m.st));
  }
  (void)dn0;
  c64* N;
  double *conv, *wry, *gn, *lap;
  HY_TRY(real_terms(m, B, dn, true, &N, &conv, &wry, &gn, &lap));
  Ctl c;
  HY_TRY(read_ctl(m, c));
  PointArgs a = point_args(m);
  a.n = dn; a.mom = dmom; a.conv = conv; a.wry = wry; a.gn = gn; a.lap = lap; a.flag = flag;
  a.vt = c.step > 0 ? c.t_vext : 0.0;
  const int nbk = 256;
  double* part;
  HY_TRY(B.get(&part, nbk * 8));
  a.out = part;
  hipLaunchKernelGGL(k_point_budget, dim3(nbk), dim3(256), 0, m.st, a);
  HY_HIP(hipGetLastError());
  std::vector<double> hp(nbk * 8);
  c64 k0[4];
  int bad = 0;
  HY_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * nbk * 8, hipMemcpyDeviceToHost, m.st));
  for (int f = 0; f < 4; ++f) HY_HIP(hipMemcpyAsync(&k0[f], sfield(m, f), sizeof(c64), hipMemcpyDeviceToHost, m.st));
  HY_HIP(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, m.st));
  HY_HIP(hipStreamSynchronize(m.st));
  if (bad) return set_error(HY_ERR_NONPOSITIVE, "density is not positive");
  double sum[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nbk; ++b)
    for (int q = 0; q < 7; ++q) sum[q] += hp[b * 8 + q];
  const double V = g.volume, Nr = (double)nr;
  // mass and momentum are the k = 0 modes of the spectral state (mean(u) = u_hat(0))
  out6[0] = k0[0].x * V;
  for (int d = 0; d < 3; ++d) out6[1 + d] = k0[1 + d].x * V;
  out6[4] = 0.5 * (sum[0] / Nr) * V / m.mass;
  double fe = 0.0;
  if (m.ideal) fe += (sum[1] / Nr) * V / m.beta;
  if (m.hartree) fe += 0.5 * (sum[2] / Nr) * V;
  // w_ry holds B(K_ry N') with K_ry = -c_cor/beta, so -1/2 <(n-n0) c_cor*N'> V/beta
  // (fe_correlation_ry.f90:96-97) is +1/2 <(n-n0) w_ry> V
  if (m.ry) fe += 0.5 * (sum[3] / Nr) * V;
  if (m.tf) fe += (sum[4] / Nr) * V / m.beta;
  if (m.grad) fe += m.cg * (sum[5] / Nr) * V;
  if (m.vext_kind != HY_VEXT_NONE) fe += (sum[6] / Nr) * V;
  out6[5] = fe;
  return HY_OK;
}

int hy_model_profile_classes(hy_model h, unsigned mask) {
  if (!h) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  HY_HIP(hipStreamSynchronize(m.st));
  for (auto& e : m.evs) { m.pool.push_back(e.a); m.pool.push_back(e.b); }
  m.evs.clear();
  m.prof = mask & ((1u << PC_N) - 1u);
  m.prof_open = false;
  return HY_OK;
}

int hy_model_profile(hy_model h, int enable) { return hy_model_profile_classes(h, enable ? ~0u : 0u); }

int hy_model_profile_read(hy_model h, int maxcls, int* launches, double* total_ms, double* bytes) {
  if (!h || !launches || !total_ms || !bytes) return set_error(HY_ERR_INVALID, "null argument");
  Model& m = h->m;
  HY_HIP(hipStreamSynchronize(m.st));
  for (int c = 0; c < maxcls; ++c) { launches[c] = 0; total_ms[c] = 0.0; bytes[c] = 0.0; }
  for (auto& e : m.evs) {
    if (e.cls >= maxcls) continue;
    float ms = 0.0f;
    HY_HIP(hipEventElapsedTime(&ms, e.a, e.b));
    launches[e.cls] += 1;
    total_ms[e.cls] += ms;
    bytes[e.cls] += e.bytes;
  }
  return PC_N;
}

const char* hy_profile_class_name(int cls) { return (cls >= 0 && cls < PC_N) ? prof_names[cls] : ""; }

int hy_model_kernel(hy_model h, double k, double* out) {
  if (!h || !out) return set_error(HY_ERR_INVALID, "null argument");
  *out = fun_kernel(h->m, k);
  return HY_OK;
}

int hy_fd_eval(int which, int n, const double* x, double* y) {
  if (n < 0 || (n > 0 && (!x || !y))) return set_error(HY_ERR_INVALID, "null argument");
  const double* tab = fd::host_table();
  for (int i = 0; i < n; ++i) {
    double a = x[i];
    switch (which) {
      case 0: y[i] = fd::im12(tab, a); break;
      case 1: y[i] = fd::ip12(tab, a); break;
      case 2: y[i] = fd::ip32(tab, a); break;
      case 3: y[i] = fd::dim12(tab, a); break;
      case 4: y[i] = fd::d2im12(tab, a); break;
      case 5: y[i] = fd::xi(tab, a); break;
      case 6: y[i] = fd::dxi(tab, a); break;
      case 7:
        if (!(a > 0.0)) return set_error(HY_ERR_INVALID, "fermi_dirac: I_{1/2} inverse needs y > 0");
        y[i] = fd::alpha_from_ip12(tab, a);
        break;
      // the field kernels' forms (table gathers, inline expansions outside it)
      case 8: case 9: case 10: {
        double im, d1, d2;
        fd::im12_d1_d2_dev(tab, a, im, d1, d2);
        y[i] = which == 8 ? im : (which == 9 ? d1 : d2);
        break;
      }
      case 11: {
        double ip, im;
        fd::ip12_im12_dev(tab, a, ip, im);
        y[i] = ip;
        break;
      }
      case 12: {
        if (!(a > 0.0)) return set_error(HY_ERR_INVALID, "fermi_dirac: I_{1/2} inverse needs y > 0");
        // started from the reference's closed-form guess (fermi_dirac.f90:246-248)
        double a0 = std::fmax(std::log(2.0 * a / fd::SQRTPI), std::pow(1.5 * a, 2.0 / 3.0) - 1.0);
        if (a < 1.0) a0 = std::log(2.0 * a / fd::SQRTPI);
        y[i] = fd::alpha_from_ip12_dev(tab, a, a0);
        break;
      }
      default: return set_error(HY_ERR_INVALID, "hy_fd_eval: unknown function");
    }
  }
  return HY_OK;
}

}  // extern "C"

// ============================================ functional terms on device buffers
// functional_t / fe_term_t (functional.f90:17-90, free_energy.f90:22-59) on a
// drop-in grid: per-term add_mu, kernel and energy of the reference's terms on
// FP64 device fields in the reference layouts.
namespace hy {

// W = mask K (N - n0sub delta_k0) on the full half spectrum
__global__ void k_fun_conv(FullArgs a, const c64* __restrict__ N, const double* __restrict__ kern, double n0sub,
                           c64* __restrict__ W) {
  const size_t n = (size_t)a.nkx * a.ny * a.nz;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % a.nkx), j = (int)((e / a.nkx) % a.ny), k = (int)(e / ((size_t)a.nkx * a.ny));
    c64 v = N[e];
    if (e == 0) v.x -= n0sub;
    W[e] = scal(v, fkept(a, i, j, k) ? kern[e] : 0.0);
  }
}

__global__ void k_fun_fill(double* __restrict__ u, size_t n, double v) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) u[e] = v;
}
__global__ void k_fun_clear(int* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

struct FunPointArgs {
  size_t nr;
  int kind;              // HY_TERM_*
  const double* n;
  const double* w;       // hartree / ry: the convolution field, added as is
  const double* gn;      // gradient: 3 fields
  const double* lap;
  double beta, cn, cg;
  const double* fdtab;
  double* mu;
  int* flag;
};

// mu += the term's pointwise part (fe_ideal_classical.f90:33-41,
// fe_kinetic_tfk.f90:63-82 and 114-138, and the convolution terms)
__global__ void k_fun_point(FunPointArgs a) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < a.nr; e += (size_t)gridDim.x * blockDim.x) {
    if (a.kind == HY_TERM_HARTREE || a.kind == HY_TERM_RY) {
      a.mu[e] += a.w[e];
      continue;
    }
    const double n = a.n[e];
    if (!(n > 0.0)) { *a.flag = 1; continue; }
    if (a.kind == HY_TERM_IDEAL) {
      a.mu[e] += log(n) / a.beta;
    } else if (a.kind == HY_TERM_TF) {
      a.mu[e] += fd::alpha_from_ip12(a.fdtab, n / a.cn) / a.beta;
    } else if (a.kind == HY_TERM_GRADIENT) {
      const double al = fd::alpha_from_ip12(a.fdtab, n / a.cn);
      const double xi = fd::xi(a.fdtab, al);
      const double dxidn = fd::dxi(a.fdtab, al) / (0.5 * a.cn * fd::im12(a.fdtab, al));
      const double g2 = a.gn[e] * a.gn[e] + a.gn[a.nr + e] * a.gn[a.nr + e] + a.gn[2 * a.nr + e] * a.gn[2 * a.nr + e];
      a.mu[e] -= a.cg * (dxidn * g2 + 2.0 * xi * a.lap[e]);
    }
  }
}

}  // namespace hy

struct hy_fun_s {
  hy::Grid* g = nullptr;                    // the grid's (borrowed: the grid outlives the functional)
  std::vector<hy_term_desc> terms;
  std::vector<std::vector<double>> kt, ct;  // RY radial tables
  std::vector<double*> ktab;                // hartree / ry: K_i(|k|) on the full half spectrum (device)
  double* fdtab = nullptr;
  c64* W = nullptr;      // work spectrum (destroyed by each backward transform)
  double* w = nullptr;   // convolution field
  double* gn = nullptr;  // gradient term: grad n (3 fields) and lap n
  double* lap = nullptr;
  double* part = nullptr;  // energy partial sums: 256 blocks x 8
  int* flag = nullptr;
};

namespace {
constexpr int FUN_NB = 256;

void fun_free(hy_fun_s* f) {
  if (!f) return;
  for (double* p : f->ktab) if (p) (void)hipFree(p);
  for (void* p : {(void*)f->fdtab, (void*)f->W, (void*)f->w, (void*)f->gn, (void*)f->lap, (void*)f->part,
                  (void*)f->flag})
    if (p) (void)hipFree(p);
  delete f;
}

int fun_term_index(const hy_fun_s* f, int term) {
  if (!f || term < 0 || term >= (int)f->terms.size()) return set_error(HY_ERR_INVALID, "hy_fun: no such term");
  return HY_OK;
}

// mu += dF_t/dn for term t, on stream st (no synchronisation)
int fun_add_mu_async(hy_fun_s* f, int t, const double* n, const c64* N, double* mu, hipStream_t st) {
  Grid& g = *f->g;
  const hy_term_desc& d = f->terms[t];
  const size_t nr = g.nreal(), nf = g.nfull();
  FullArgs fa;
  fa.nkx = g.nkx; fa.ny = g.ny; fa.nz = g.nz; fa.Kx = g.Kx; fa.My = g.My; fa.Mz = g.Mz;
  fa.kx = g.d_kx; fa.ky = g.d_ky; fa.kz = g.d_kz;
  FunPointArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nr = nr; a.kind = d.kind; a.n = n; a.mu = mu; a.flag = f->flag;
  a.beta = d.beta; a.cn = d.cn; a.cg = d.cg; a.fdtab = f->fdtab;
  if (d.kind == HY_TERM_HARTREE || d.kind == HY_TERM_RY) {
    // hartree: mu += B(v N mask); ry: mu -= B(c_cor (N - n0 delta_k0) mask) / beta
    hipLaunchKernelGGL(k_fun_conv, dim3(nblocks(nf)), dim3(256), 0, st, fa, N, f->ktab[t],
                       d.kind == HY_TERM_RY ? d.n0 : 0.0, f->W);
    HY_TRY(fft_c2r_full_api(g, f->W, f->w, st));
    a.w = f->w;
  } else if (d.kind == HY_TERM_GRADIENT) {
    for (int dd = 0; dd < 3; ++dd) {
      double* dst = f->gn + nr * dd;
      if (dd >= g.ndim) {
        hipLaunchKernelGGL(k_fun_fill, dim3(nblocks(nr)), dim3(256), 0, st, dst, nr, 0.0);
        continue;
      }
      const double* kdir = dd == 0 ? g.d_kx : (dd == 1 ? g.d_ky : g.d_kz);
      const size_t kstride = dd == 0 ? 1 : (dd == 1 ? (size_t)g.nkx : (size_t)g.nkx * g.ny);
      const int kn = dd == 0 ? g.nkx : (dd == 1 ? g.ny : g.nz);
      hipLaunchKernelGGL(k_full_conv, dim3(nblocks(nf)), dim3(256), 0, st, fa, 1, kdir, kstride, kn, N,
                         (const double*)nullptr, (const double*)nullptr, 0.0, f->W);
      HY_TRY(fft_c2r_full_api(g, f->W, dst, st));
    }
    hipLaunchKernelGGL(k_full_conv, dim3(nblocks(nf)), dim3(256), 0, st, fa, 2, g.d_kx, (size_t)1, g.nkx, N,
                       (const double*)nullptr, (const double*)nullptr, 0.0, f->W);
    HY_TRY(fft_c2r_full_api(g, f->W, f->lap, st));
    a.gn = f->gn;
    a.lap = f->lap;
  }
  hipLaunchKernelGGL(k_fun_point, dim3(nblocks(nr)), dim3(256), 0, st, a);
  HY_HIP(hipGetLastError());
  return HY_OK;
}

int fun_check(hy_fun_s* f, hipStream_t st) {
  int bad = 0;
  HY_HIP(hipMemcpyAsync(&bad, f->flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HY_HIP(hipStreamSynchronize(st));
  if (bad) return set_error(HY_ERR_NONPOSITIVE, "functional: density is not positive");
  return HY_OK;
}
}  // namespace

extern "C" {

int hy_fun_create(hy_grid gh, int nterms, const hy_term_desc* terms, hy_fun* out) {
  if (!gh || !out || nterms < 0 || (nterms > 0 && !terms)) return set_error(HY_ERR_INVALID, "hy_fun_create: bad argument");
  auto* f = new hy_fun_s();
  f->g = &gh->g;
  Grid& g = gh->g;
  f->terms.assign(terms, terms + nterms);
  f->kt.resize(nterms);
  f->ct.resize(nterms);
  f->ktab.assign(nterms, nullptr);
  bool grad = false, fd_needed = false;
  auto fail = [&](int rc) { fun_free(f); return rc; };
  for (int t = 0; t < nterms; ++t) {
    const hy_term_desc& d = terms[t];
    if (d.kind < HY_TERM_IDEAL || d.kind > HY_TERM_RY) return fail(set_error(HY_ERR_UNKNOWN_KIND, "unknown free-energy term kind"));
    if (!(d.beta > 0.0) || !(d.n0 > 0.0)) return fail(set_error(HY_ERR_INVALID, "term: beta and n0 must be > 0"));
    if ((d.kind == HY_TERM_TF || d.kind == HY_TERM_GRADIENT) && !(d.cn > 0.0))
      return fail(set_error(HY_ERR_INVALID, "tf/gradient term needs cn > 0 (kind = electron)"));
    if (d.kind == HY_TERM_RY) {
      if (d.ntab < 1 || !d.k_tab || !d.c_tab) return fail(set_error(HY_ERR_INVALID, "ry term needs a structure table"));
      f->kt[t].assign(d.k_tab, d.k_tab + d.ntab);
      f->ct[t].assign(d.c_tab, d.c_tab + d.ntab);
    }
    grad |= d.kind == HY_TERM_GRADIENT;
    fd_needed |= d.kind == HY_TERM_TF || d.kind == HY_TERM_GRADIENT;
  }
  const size_t nr = g.nreal(), nf = g.nfull();
  // convolution tables: the terms' own kernels K_i(|k|) (fe_hartree.f90:41-43,
  // fe_correlation_ry.f90:59-61 with the 1/beta of ry_add_mu folded in)
  for (int t = 0; t < nterms; ++t) {
    const hy_term_desc& d = f->terms[t];
    if (d.kind != HY_TERM_HARTREE && d.kind != HY_TERM_RY) continue;
    std::vector<double> h(nf);
    for (int k = 0; k < g.nz; ++k)
      for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nkx; ++i) {
          const double kabs = std::sqrt(g.kx[i] * g.kx[i] + g.ky[j] * g.ky[j] + g.kz[k] * g.kz[k]);
          h[((size_t)k * g.ny + j) * g.nkx + i] = term_kernel_desc(d, f->kt[t], f->ct[t], kabs);
        }
    if (dmalloc((void**)&f->ktab[t], sizeof(double) * nf) != HY_OK) return fail(HY_ERR_OOM);
    if (hipMemcpy(f->ktab[t], h.data(), sizeof(double) * nf, hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_error(HY_ERR_HIP, "hy_fun_create: upload"));
  }
  if (fd_needed) {
    if (dmalloc((void**)&f->fdtab, sizeof(double) * fd::NTAB * fd::NC) != HY_OK) return fail(HY_ERR_OOM);
    if (hipMemcpy(f->fdtab, fd::host_table(), sizeof(double) * fd::NTAB * fd::NC, hipMemcpyHostToDevice) != hipSuccess)
      return fail(set_error(HY_ERR_HIP, "hy_fun_create: upload"));
  }
  if (dmalloc((void**)&f->W, sizeof(c64) * nf) != HY_OK || dmalloc((void**)&f->w, sizeof(double) * nr) != HY_OK ||
      dmalloc((void**)&f->part, sizeof(double) * FUN_NB * 8) != HY_OK || dmalloc((void**)&f->flag, sizeof(int)) != HY_OK)
    return fail(HY_ERR_OOM);
  if (grad && (dmalloc((void**)&f->gn, sizeof(double) * nr * 3) != HY_OK ||
               dmalloc((void**)&f->lap, sizeof(double) * nr) != HY_OK))
    return fail(HY_ERR_OOM);
  *out = f;
  return HY_OK;
}

int hy_fun_destroy(hy_fun f) {
  fun_free(f);
  return HY_OK;
}

int hy_fun_nterms(hy_fun f, int* n) {
  if (!f || !n) return set_error(HY_ERR_INVALID, "null argument");
  *n = (int)f->terms.size();
  return HY_OK;
}

int hy_fun_add_mu(hy_fun f, int term, const double* d_n, const double* d_nhat, double* d_mu, void* stream) {
  HY_TRY(fun_term_index(f, term));
  if (!d_n || !d_nhat || !d_mu) return set_error(HY_ERR_INVALID, "null argument");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fun_clear, dim3(1), dim3(1), 0, st, f->flag);
  HY_TRY(fun_add_mu_async(f, term, d_n, reinterpret_cast<const c64*>(d_nhat), d_mu, st));
  return fun_check(f, st);
}

int hy_fun_mu(hy_fun f, const double* d_n, const double* d_nhat, double* d_mu, void* stream) {
  if (!f || !d_n || !d_nhat || !d_mu) return set_error(HY_ERR_INVALID, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const size_t nr = f->g->nreal();
  hipLaunchKernelGGL(k_fun_clear, dim3(1), dim3(1), 0, st, f->flag);
  hipLaunchKernelGGL(k_fun_fill, dim3(nblocks(nr)), dim3(256), 0, st, d_mu, nr, 0.0);  // functional.f90:61
  for (int t = 0; t < (int)f->terms.size(); ++t)
    HY_TRY(fun_add_mu_async(f, t, d_n, reinterpret_cast<const c64*>(d_nhat), d_mu, st));
  return fun_check(f, st);
}

int hy_fun_kernel(hy_fun f, int term, double k, double* h_K) {
  if (!f || !h_K) return set_error(HY_ERR_INVALID, "null argument");
  if (term >= 0) {
    HY_TRY(fun_term_index(f, term));
    *h_K = term_kernel_desc(f->terms[term], f->kt[term], f->ct[term], k);
    return HY_OK;
  }
  double kk = 0.0;  // functional.f90:70-78
  for (size_t t = 0; t < f->terms.size(); ++t) kk += term_kernel_desc(f->terms[t], f->kt[t], f->ct[t], k);
  *h_K = kk;
  return HY_OK;
}

int hy_fun_energy(hy_fun f, int term, const double* d_n, const double* d_nhat, double* h_F, void* stream) {
  if (!f || !d_n || !d_nhat || !h_F) return set_error(HY_ERR_INVALID, "null argument");
  if (term >= 0) HY_TRY(fun_term_index(f, term));
  hipStream_t st = (hipStream_t)stream;
  Grid& g = *f->g;
  const size_t nr = g.nreal(), nf = g.nfull();
  const c64* N = reinterpret_cast<const c64*>(d_nhat);
  FullArgs fa;
  fa.nkx = g.nkx; fa.ny = g.ny; fa.nz = g.nz; fa.Kx = g.Kx; fa.My = g.My; fa.Mz = g.Mz;
  fa.kx = g.d_kx; fa.ky = g.d_ky; fa.kz = g.d_kz;
  double total = 0.0;
  const int t0 = term >= 0 ? term : 0, t1 = term >= 0 ? term + 1 : (int)f->terms.size();
  for (int t = t0; t < t1; ++t) {
    const hy_term_desc& d = f->terms[t];
    PointArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nr = nr; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nd = 0;  // no momentum sums
    a.n = d_n; a.n0 = d.n0; a.beta = d.beta; a.cn = d.cn; a.alpha0 = d.alpha0; a.cg = d.cg;
    a.fdtab = f->fdtab; a.vext_kind = HY_VEXT_NONE;
    a.out = f->part; a.flag = f->flag;
    if (d.kind == HY_TERM_HARTREE || d.kind == HY_TERM_RY) {
      hipLaunchKernelGGL(k_fun_conv, dim3(nblocks(nf)), dim3(256), 0, st, fa, N, f->ktab[t],
                         d.kind == HY_TERM_RY ? d.n0 : 0.0, f->W);
      HY_TRY(fft_c2r_full_api(g, f->W, f->w, st));
      if (d.kind == HY_TERM_HARTREE) { a.hartree = 1; a.conv = f->w; }
      else { a.ry = 1; a.wry = f->w; }
    } else if (d.kind == HY_TERM_GRADIENT) {
      for (int dd = 0; dd < 3; ++dd) {
        double* dst = f->gn + nr * dd;
        if (dd >= g.ndim) {
          hipLaunchKernelGGL(k_fun_fill, dim3(nblocks(nr)), dim3(256), 0, st, dst, nr, 0.0);
          continue;
        }
        const double* kdir = dd == 0 ? g.d_kx : (dd == 1 ? g.d_ky : g.d_kz);
        const size_t kstride = dd == 0 ? 1 : (dd == 1 ? (size_t)g.nkx : (size_t)g.nkx * g.ny);
        const int kn = dd == 0 ? g.nkx : (dd == 1 ? g.ny : g.nz);
        hipLaunchKernelGGL(k_full_conv, dim3(nblocks(nf)), dim3(256), 0, st, fa, 1, kdir, kstride, kn, N,
                           (const double*)nullptr, (const double*)nullptr, 0.0, f->W);
        HY_TRY(fft_c2r_full_api(g, f->W, dst, st));
      }
      a.grad = 1; a.gn = f->gn;
    } else if (d.kind == HY_TERM_IDEAL) {
      a.ideal = 1;
    } else {
      a.tf = 1;
    }
    hipLaunchKernelGGL(k_fun_clear, dim3(1), dim3(1), 0, st, f->flag);
    hipLaunchKernelGGL(k_point_budget, dim3(FUN_NB), dim3(256), 0, st, a);
    HY_HIP(hipGetLastError());
    std::vector<double> hp(FUN_NB * 8);
    HY_HIP(hipMemcpyAsync(hp.data(), f->part, sizeof(double) * FUN_NB * 8, hipMemcpyDeviceToHost, st));
    HY_TRY(fun_check(f, st));
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < FUN_NB; ++b)
      for (int q = 0; q < 7; ++q) s[q] += hp[b * 8 + q];
    const double V = g.volume, Nr = (double)nr;
    switch (d.kind) {  // the terms' energy functions (fe_*.f90 energy)
      case HY_TERM_IDEAL: total += (s[1] / Nr) * V / d.beta; break;
      case HY_TERM_HARTREE: total += 0.5 * (s[2] / Nr) * V; break;
      case HY_TERM_RY: total += 0.5 * (s[3] / Nr) * V; break;  // w holds B(-c_cor N' / beta)
      case HY_TERM_TF: total += (s[4] / Nr) * V / d.beta; break;
      case HY_TERM_GRADIENT: total += d.cg * (s[5] / Nr) * V; break;
    }
  }
  *h_F = total;
  return HY_OK;
}

}  // extern "C"
