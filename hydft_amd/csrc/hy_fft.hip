// Grid, batched y/z pencil FFT passes, x-line transforms and the drop-in
// spectral operators (fft_t, spec_t, stress_t of the reference) on FP64
// device buffers in the reference layouts.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "hy_internal.h"
#include "hy_regfft.h"

namespace hy {

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;
int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

static int trace_level() {
  static int lv = -1;
  if (lv < 0) {
    const char* e = std::getenv("HYDFT_TRACE");
    lv = (e && e[0] == '1') ? 1 : 0;
  }
  return lv;
}
int trace_launch(const char* what, hipStream_t s) {
  if (!trace_level()) return HY_OK;
  std::fprintf(stderr, "[hydft] launched %s\n", what);
  std::fflush(stderr);
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipGetLastError();
  std::fprintf(stderr, "[hydft] done %s: %s\n", what, hipGetErrorString(e));
  std::fflush(stderr);
  if (e != hipSuccess) return set_error(HY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return HY_OK;
}

// ------------------------------------------------------------------- grid
const TwiddleSet& Grid::tw(int L) const {
  for (const auto& t : tws)
    if (t.L == L) return t;
  return tws.front();
}

static int wrapi(int m, int n) { return m <= n / 2 ? m : m - n; }  // grid.f90:116-123

int grid_build(Grid& g, int nx, int ny, int nz, double lx, double ly, double lz, int dealias,
               int precision) {
  if (nx < 2) return set_error(HY_ERR_INVALID, "grid: nx must be >= 2");
  if (nx % 2) return set_error(HY_ERR_INVALID, "grid: nx must be even");
  if (ny > 1 && ny % 2) return set_error(HY_ERR_INVALID, "grid: ny must be even (or 1)");
  if (nz > 1 && nz % 2) return set_error(HY_ERR_INVALID, "grid: nz must be even (or 1)");
  if (ny < 1 || nz < 1) return set_error(HY_ERR_INVALID, "grid: ny, nz must be >= 1");
  if (!supported_len(nx) || (ny > 1 && !supported_len(ny)) || (nz > 1 && !supported_len(nz)))
    return set_error(HY_ERR_UNSUPPORTED,
                     "grid: each of nx, ny, nz must be 1 (ny, nz) or one of 4 8 12 16 24 32 48 64 96 128 256");
  if (precision != HY_PREC_FP64 && precision != HY_PREC_MIXED)
    return set_error(HY_ERR_INVALID, "grid: unknown precision");
  const double twopi = 2.0 * 3.14159265358979323846264338327950288;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.nkx = nx / 2 + 1;
  g.ndim = 1;
  if (ny > 1) g.ndim = 2;
  if (nz > 1) g.ndim = 3;
  g.lx = lx; g.ly = ly; g.lz = lz;
  g.dx = lx / nx; g.dy = ly / ny; g.dz = lz / nz;
  g.volume = lx;
  if (ny > 1) g.volume *= ly;
  if (nz > 1) g.volume *= lz;
  g.dealias = dealias ? 1 : 0;
  g.precision = precision;
  g.kx.assign(g.nkx, 0.0); g.ky.assign(ny, 0.0); g.kz.assign(nz, 0.0);
  for (int i = 0; i < g.nkx; ++i) g.kx[i] = twopi * i / lx;
  if (ny > 1) for (int j = 0; j < ny; ++j) g.ky[j] = twopi * wrapi(j, ny) / ly;
  if (nz > 1) for (int k = 0; k < nz; ++k) g.kz[k] = twopi * wrapi(k, nz) / lz;
  // The reference mask (grid.f90:90-108) is a per-axis test, so the kept set is
  // a box: kx index < Kx, |my| <= My, |mz| <= Mz.  Derive the box with the same
  // floating-point comparisons and verify it point by point.
  const double huge = 1.7976931348623157e308;
  double kxc = twopi * (nx / 2) / lx * (2.0 / 3.0), kyc = huge, kzc = huge;
  if (ny > 1) kyc = twopi * (ny / 2) / ly * (2.0 / 3.0);
  if (nz > 1) kzc = twopi * (nz / 2) / lz * (2.0 / 3.0);
  auto keepx = [&](int i) { return dealias ? !(std::fabs(g.kx[i]) >= kxc) : (i != g.nkx - 1); };
  auto keepy = [&](int j) { return dealias ? !(std::fabs(g.ky[j]) >= kyc) : !(ny > 1 && j == ny / 2); };
  auto keepz = [&](int k) { return dealias ? !(std::fabs(g.kz[k]) >= kzc) : !(nz > 1 && k == nz / 2); };
  g.Kx = 0;
  while (g.Kx < g.nkx && keepx(g.Kx)) ++g.Kx;
  for (int i = g.Kx; i < g.nkx; ++i)
    if (keepx(i)) return set_error(HY_ERR_INVALID, "grid: kept kx set is not contiguous");
  g.My = 0;
  if (ny > 1) { while (g.My + 1 <= ny / 2 && keepy(g.My + 1)) ++g.My; }
  g.Mz = 0;
  if (nz > 1) { while (g.Mz + 1 <= nz / 2 && keepz(g.Mz + 1)) ++g.Mz; }
  for (int j = 0; j < ny; ++j) {
    bool box = std::abs(wrapi(j, ny)) <= g.My;
    if (keepy(j) != box) return set_error(HY_ERR_INVALID, "grid: kept ky set is not a box");
  }
  for (int k = 0; k < nz; ++k) {
    bool box = std::abs(wrapi(k, nz)) <= g.Mz;
    if (keepz(k) != box) return set_error(HY_ERR_INVALID, "grid: kept kz set is not a box");
  }
  if (g.Kx < 1 || !keepy(0) || !keepz(0)) return set_error(HY_ERR_INVALID, "grid: empty kept set");
  g.Ky = ny > 1 ? 2 * g.My + 1 : 1;
  g.Kz = nz > 1 ? 2 * g.Mz + 1 : 1;
  g.kmax_dealiased = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < g.nkx; ++i)
        if (keepx(i) && keepy(j) && keepz(k)) {
          double k2 = g.kx[i] * g.kx[i] + g.ky[j] * g.ky[j] + g.kz[k] * g.kz[k];
          g.kmax_dealiased = std::max(g.kmax_dealiased, std::sqrt(k2));
        }
  // device wavenumbers
  HY_HIP(hipMalloc(&g.d_kx, sizeof(double) * g.nkx));
  HY_HIP(hipMalloc(&g.d_ky, sizeof(double) * ny));
  HY_HIP(hipMalloc(&g.d_kz, sizeof(double) * nz));
  HY_HIP(hipMemcpy(g.d_kx, g.kx.data(), sizeof(double) * g.nkx, hipMemcpyHostToDevice));
  HY_HIP(hipMemcpy(g.d_ky, g.ky.data(), sizeof(double) * ny, hipMemcpyHostToDevice));
  HY_HIP(hipMemcpy(g.d_kz, g.kz.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  // twiddles exp(-2 pi i j/L), computed in extended precision and rounded once
  int lens[3] = {nx, ny, nz};
  for (int L : lens) {
    if (L < 2) continue;
    bool have = false;
    for (auto& t : g.tws) have |= (t.L == L);
    if (have) continue;
    TwiddleSet t;
    t.L = L;
    std::vector<c64> h64(L);
    std::vector<c32> h32(L);
    const long double tp = 6.283185307179586476925286766559005768L;
    for (int j = 0; j < L; ++j) {
      long double a = tp * (long double)j / (long double)L;
      double c = (double)cosl(a), s = (double)-sinl(a);
      h64[j] = mk<double>(c, s);
      h32[j] = mk<float>((float)cosl(a), (float)-sinl(a));
    }
    HY_HIP(hipMalloc(&t.d64, sizeof(c64) * L));
    HY_HIP(hipMalloc(&t.d32, sizeof(c32) * L));
    HY_HIP(hipMemcpy(t.d64, h64.data(), sizeof(c64) * L, hipMemcpyHostToDevice));
    HY_HIP(hipMemcpy(t.d32, h32.data(), sizeof(c32) * L, hipMemcpyHostToDevice));
    g.tws.push_back(t);
  }
  return HY_OK;
}

void grid_free(Grid& g) {
  (void)hipFree(g.d_kx); (void)hipFree(g.d_ky); (void)hipFree(g.d_kz);
  for (auto& t : g.tws) { (void)hipFree(t.d64); (void)hipFree(t.d32); }
  g.tws.clear();
  if (g.d_scratch) (void)hipFree(g.d_scratch);
  g.d_scratch = nullptr;
  g.d_kx = g.d_ky = g.d_kz = nullptr;
}

// ---------------------------------------------------- batched pencil passes
struct PencilArgs {
  PencilField f[MAXPF];
  int A, B, L, K, h;        // contiguous count, outer count, FFT length, stored count, #nonneg stored
  double scale;             // forward: 1/L
  // Wavenumbers are computed as (signed index) x spacing: a table load here
  // would be a dependent load in front of each multiplied element.
  double dkax;              // 2 pi / (box length) along the pass axis
  // decoding of the contiguous index for MUL_NEGK2 (first inverse pass only):
  double dkx; int ax_x;     // kx spacing; x count in the contiguous index
  double dky; int ay_n;     // ky spacing; y stored count in the contiguous index (0 = none)
  int ay_h, ay_L;                // y stored-index mapping (h, L) when ay_n > 0
  int y_from_pencil;             // 1: ky comes from the pencil position (y pass)
  int field_fastest;             // 1: workgroup order with the field index fastest (inverse passes)
  const int* halt;               // every workgroup returns when *halt != 0 (null: never)
};

// signed index of FFT position p on an axis of length n (grid.f90:116-123)
__device__ inline int wrap_index(int p, int n) { return p <= n / 2 ? p : p - n; }
// Written without FMA contraction: the single and batched forms of a pencil
// kernel then round every product and sum alike, so batched runs equal
// single runs bit for bit.
template <class T>
__device__ inline cpx<T> apply_mul(cpx<T> v, int mul, int p, int ga, const PencilArgs& a) {
#pragma clang fp contract(off)
  if (mul == MUL_IK) {
    double k = (double)wrap_index(p, a.L) * a.dkax;
    return mk<T>((T)(-k * (double)v.y), (T)(k * (double)v.x));
  }
  if (mul == MUL_NEGK2) {
    int ix = ga % a.ax_x;
    double kx = (double)ix * a.dkx;
    double ky = 0.0, kz = 0.0;
    if (a.y_from_pencil) {
      ky = (double)wrap_index(p, a.L) * a.dkax;
    } else {
      kz = (double)wrap_index(p, a.L) * a.dkax;
      if (a.ay_n > 0) {
        int iy = ga / a.ax_x;
        int jy = iy < a.ay_h ? iy : iy + (a.ay_L - a.ay_n);
        ky = (double)wrap_index(jy, a.ay_L) * a.dky;
      }
    }
    double k2 = kx * kx + ky * ky + kz * kz;
    return mk<T>((T)(-k2 * (double)v.x), (T)(-k2 * (double)v.y));
  }
  return v;
}

// Field and twiddle pointers are device (global) memory.  The batched kernels
// read them from a device array, where the compiler cannot infer that and
// would emit flat accesses; these helpers make every pencil access a global
// one.  (A host-side build of this file uses plain pointers.)
template <class T>
__device__ inline cpx<T> gload_c(const void* p, size_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = ((const __attribute__((address_space(1))) v2*)p)[i];
  return mk<T>(v.x, v.y);
#else
  return reinterpret_cast<const cpx<T>*>(p)[i];
#endif
}
template <class T>
__device__ inline void gstore_c(void* p, size_t i, cpx<T> c) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 v;
  v.x = c.x;
  v.y = c.y;
  ((__attribute__((address_space(1))) v2*)p)[i] = v;
#else
  reinterpret_cast<cpx<T>*>(p)[i] = c;
#endif
}

// NP elements idx[i] of the nel-element buffer p (ok[i] false: zero), all
// loads issued before any is used: one uniform branch on the element type
// around the group, and branch-free skipped loads (bload_c).
template <class T, int NP>
__device__ inline void gather_any(const void* p, size_t nel, const unsigned* idx, const bool* ok, int is64,
                                  cpx<T>* v) {
  if (is64) {
#pragma unroll
    for (int i = 0; i < NP; ++i) v[i] = cvt<T>(bload_c<double>(p, (unsigned)(nel * sizeof(c64)), idx[i], ok[i]));
  } else {
#pragma unroll
    for (int i = 0; i < NP; ++i) v[i] = cvt<T>(bload_c<float>(p, (unsigned)(nel * sizeof(c32)), idx[i], ok[i]));
  }
}

// Block shape of the pencil kernels: NA contiguous columns times TPC threads
// per column (the column plan of hy_regfft.h: columns of at most 16 points take
// one thread each).  NA is a template parameter: the plan autotuner
// (pencil_plan below) times the candidates of each pass shape once and keeps
// the fastest.  The default (the choice without timing) takes 64 columns per
// block for one-thread columns; FP32 takes 16 for columns of at most 48 points,
// 24 at 64 and 128 points and 32 otherwise; FP64 takes 8, or as many as fill
// one wave where a column has few threads (32 at 24 points, 16 at 32, 48 and
// 96 points).
template <class T, int L>
__host__ __device__ constexpr int pencil_tpc() { return col_plan(L, (int)sizeof(cpx<T>)).TPC; }
template <class T, int L>
__host__ __device__ constexpr int pencil_na_default() {
  return pencil_tpc<T, L>() == 1 ? 64
         : (sizeof(T) == 4 ? (L <= 48 ? 16 : (L == 64 || L == 128 ? 24 : 32))
                           : (64 / pencil_tpc<T, L>() > 8 ? 64 / pencil_tpc<T, L>() : 8));
}
// a candidate: at most 1024 threads and 64 KiB of LDS per workgroup
template <class T, int L, int NA>
__host__ __device__ constexpr bool pencil_na_ok() {
  return (pencil_tpc<T, L>() == 1 ? (NA == 32 || NA == 64 || NA == 128)
                                  : (NA == 8 || NA == 16 || NA == 24 || NA == 32)) &&
         NA * pencil_tpc<T, L>() <= 1024 &&
         (pencil_tpc<T, L>() == 1 || (size_t)L * NA * sizeof(cpx<T>) <= 65536);
}
template <class T, int L, int NA_>
struct PencilShape {
  static constexpr ColPlan P = col_plan(L, (int)sizeof(cpx<T>));
  static constexpr int L1 = P.L1, L2 = P.L2, TPC = P.TPC;
  static constexpr int NA = NA_;
  static constexpr int NT = NA * TPC;
  static constexpr int J1 = L2 / TPC;  // step-1 DFTs (length L1) per thread
  static constexpr int J2 = L1 / TPC;  // step-2 DFTs (length L2) per thread
};

// stored index of position p along the axis, or -1 outside the kept range
__device__ inline int kept_index(int p, int h, int tailoff) {
  return p < h ? p : (p >= h + tailoff ? p - tailoff : -1);
}

// Inverse pass: in (A, K, B) stored spectrum (kept entries only) -> out (A, L, B).
template <class T>
struct PencilK {
  PencilArgs a;
  const cpx<T>* tw;
};

template <class T, int L, int NA_>
__device__ __attribute__((always_inline)) inline void pencil_inv_body(const PencilArgs& a, const cpx<T>* __restrict__ twg) {
  if (gate_closed(a.halt)) return;  // uniform over the workgroup
  using S = PencilShape<T, L, NA_>;
  constexpr int L1 = S::L1, L2 = S::L2, TPC = S::TPC, NA = S::NA;
  const int al = threadIdx.x % NA, g = threadIdx.x / NA;
  // 1-D grid of ntile * B * nf workgroups in XCD-aware order.  With
  // field_fastest (z inverse passes, whose fields often share an input: a
  // field and its derivatives), the field index varies fastest, so fields
  // of one input load the same tile back to back on one XCD, the later ones
  // from its L2; otherwise the tile index does (neighbouring tiles share
  // cache lines of a row).
  const int ntile = (a.A + NA - 1) / NA;
  const unsigned nf = gridDim.x / ((unsigned)ntile * (unsigned)a.B);
  const unsigned w = xcd_work_index(blockIdx.x, gridDim.x);
  int fi, tile, b;
  if (a.field_fastest) {
    fi = (int)(w % nf);
    tile = (int)((w / nf) % (unsigned)ntile);
    b = (int)(w / (nf * (unsigned)ntile));
  } else {
    tile = (int)(w % (unsigned)ntile);
    b = (int)((w / (unsigned)ntile) % (unsigned)a.B);
    fi = (int)(w / ((unsigned)ntile * (unsigned)a.B));
  }
  const int ga = tile * NA + al;
  const PencilField& f = a.f[fi];
  const bool live = ga < a.A;
  const int tailoff = L - a.K;
  // the NP elements at positions pos[i] along the pass axis (zero outside the
  // kept range), with the k = 0 offset and the wavenumber multiplier
  auto gather = [&](auto np, const int* pos, cpx<T>* v) {
    constexpr int NP = decltype(np)::value;
    int t[NP];
    bool ok[NP];
    unsigned idx[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      t[i] = kept_index(pos[i], a.h, tailoff);
      ok[i] = t[i] >= 0 && live;
      idx[i] = ((unsigned)b * (unsigned)a.K + (unsigned)t[i]) * (unsigned)a.A + (unsigned)ga;
    }
    gather_any<T, NP>(f.in, (size_t)a.B * a.K * a.A, idx, ok, f.in_f64, v);
    if (f.k0_offset != 0.0 && ga == 0 && b == 0) {
#pragma unroll
      for (int i = 0; i < NP; ++i)
        if (ok[i] && t[i] == 0) v[i].x = (T)((double)v[i].x + f.k0_offset);
    }
    if (f.mul != MUL_NONE) {
#pragma unroll
      for (int i = 0; i < NP; ++i)
        if (ok[i]) v[i] = apply_mul<T>(v[i], f.mul, pos[i], ga, a);
    }
  };
  void* out = f.out;
  if constexpr (TPC == 1) {
    cpx<T> v[L];
    int pos[L];
#pragma unroll
    for (int p = 0; p < L; ++p) pos[p] = p;
    gather(std::integral_constant<int, L>{}, pos, v);
    rdft<T, L, true>(v);
    if (live) {
#pragma unroll
      for (int p = 0; p < L; ++p) gstore_c<T>(out, ((size_t)b * L + p) * a.A + ga, v[p]);
    }
  } else {
    __shared__ cpx<T> lds[L * NA];
    // the cross twiddles W_L^(n2 k1) from an LDS copy of the table: one
    // global load per thread (issued with the column loads) instead of
    // L1 - 1 dependent loads after the first DFT
    // (issued after the column loads: the first DFT waits for those anyway)
    __shared__ cpx<T> stw[L];
    cpx<T> v[S::J1][L1];
#pragma unroll
    for (int j = 0; j < S::J1; ++j) {
      const int n2 = g + TPC * j;
      int pos[L1];
#pragma unroll
      for (int n1 = 0; n1 < L1; ++n1) pos[n1] = L2 * n1 + n2;
      gather(std::integral_constant<int, L1>{}, pos, v[j]);
    }
#pragma unroll
    for (int r = 0; r < (L + S::NT - 1) / S::NT; ++r) {
      const int e = (int)threadIdx.x + S::NT * r;
      if (e < L) stw[e] = gload_c<T>(twg, e);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < S::J1; ++j) {
      const int n2 = g + TPC * j;
      rdft<T, L1, true>(v[j]);
#pragma unroll
      for (int k1 = 0; k1 < L1; ++k1) {
        cpx<T> w = v[j][k1];
        if (k1 > 0) w = cmulc(w, stw[(n2 * k1) % L]);
        lds[(k1 * L2 + n2) * NA + al] = w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < S::J2; ++j) {
      const int k1 = g + TPC * j;
      cpx<T> w[L2];
#pragma unroll
      for (int n2 = 0; n2 < L2; ++n2) w[n2] = lds[(k1 * L2 + n2) * NA + al];
      rdft<T, L2, true>(w);
      if (live) {
#pragma unroll
        for (int k2 = 0; k2 < L2; ++k2) gstore_c<T>(out, ((size_t)b * L + k1 + L1 * k2) * a.A + ga, w[k2]);
      }
    }
  }
}

// Forward pass: in (A, L, B) -> out (A, K, B) kept entries, scaled by 1/L.
template <class T, int L, int NA_>
__device__ __attribute__((always_inline)) inline void pencil_fwd_body(const PencilArgs& a, const cpx<T>* __restrict__ twg) {
  if (gate_closed(a.halt)) return;  // uniform over the workgroup
  using S = PencilShape<T, L, NA_>;
  constexpr int L1 = S::L1, L2 = S::L2, TPC = S::TPC, NA = S::NA;
  const int al = threadIdx.x % NA, g = threadIdx.x / NA;
  // 1-D grid of ntile * B * nf workgroups in XCD-aware order
  const int ntile = (a.A + NA - 1) / NA;
  const unsigned w = xcd_work_index(blockIdx.x, gridDim.x);
  const int tile = (int)(w % (unsigned)ntile);
  const int b = (int)((w / (unsigned)ntile) % (unsigned)a.B);
  const int fi = (int)(w / ((unsigned)ntile * (unsigned)a.B));
  const int ga = tile * NA + al;
  const PencilField& f = a.f[fi];
  const bool live = ga < a.A;
  const int tailoff = L - a.K;
  const T sc = (T)a.scale;
  auto gather = [&](auto np, const int* pos, cpx<T>* v) {
    constexpr int NP = decltype(np)::value;
    bool ok[NP];
    unsigned idx[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      ok[i] = live;
      idx[i] = ((unsigned)b * (unsigned)L + (unsigned)pos[i]) * (unsigned)a.A + (unsigned)ga;
    }
    gather_any<T, NP>(f.in, (size_t)a.B * L * a.A, idx, ok, f.in_f64, v);
  };
  void* out = f.out;
  auto store = [&](int p, cpx<T> v) {
    const int t = kept_index(p, a.h, tailoff);
    if (t >= 0 && live) gstore_c<T>(out, ((size_t)b * a.K + t) * a.A + ga, apply_mul<T>(scal(v, sc), f.mul, p, ga, a));
  };
  if constexpr (TPC == 1) {
    cpx<T> v[L];
    int pos[L];
#pragma unroll
    for (int p = 0; p < L; ++p) pos[p] = p;
    gather(std::integral_constant<int, L>{}, pos, v);
    rdft<T, L, false>(v);
#pragma unroll
    for (int p = 0; p < L; ++p) store(p, v[p]);
  } else {
    __shared__ cpx<T> lds[L * NA];
    // the cross twiddles W_L^(n2 k1) from an LDS copy of the table: one
    // global load per thread (issued with the column loads) instead of
    // L1 - 1 dependent loads after the first DFT
    // (issued after the column loads: the first DFT waits for those anyway)
    __shared__ cpx<T> stw[L];
    cpx<T> v[S::J1][L1];
#pragma unroll
    for (int j = 0; j < S::J1; ++j) {
      const int n2 = g + TPC * j;
      int pos[L1];
#pragma unroll
      for (int n1 = 0; n1 < L1; ++n1) pos[n1] = L2 * n1 + n2;
      gather(std::integral_constant<int, L1>{}, pos, v[j]);
    }
#pragma unroll
    for (int r = 0; r < (L + S::NT - 1) / S::NT; ++r) {
      const int e = (int)threadIdx.x + S::NT * r;
      if (e < L) stw[e] = gload_c<T>(twg, e);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < S::J1; ++j) {
      const int n2 = g + TPC * j;
      rdft<T, L1, false>(v[j]);
#pragma unroll
      for (int k1 = 0; k1 < L1; ++k1) {
        cpx<T> w = v[j][k1];
        if (k1 > 0) w = cmul(w, stw[(n2 * k1) % L]);
        lds[(k1 * L2 + n2) * NA + al] = w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < S::J2; ++j) {
      const int k1 = g + TPC * j;
      cpx<T> w[L2];
#pragma unroll
      for (int n2 = 0; n2 < L2; ++n2) w[n2] = lds[(k1 * L2 + n2) * NA + al];
      rdft<T, L2, false>(w);
#pragma unroll
      for (int k2 = 0; k2 < L2; ++k2) store(k1 + L1 * k2, w[k2]);
    }
  }
}

// Axis tags of the single-run pencil kernels: the y and z passes run kernels
// of their own names, so kernel traces and counters tell the two axes apart
// (the batched kernels serve both).
struct AxisY {};
struct AxisZ {};

template <class T, int L, int NA, class AX>
__global__ __launch_bounds__((PencilShape<T, L, NA>::NT)) void k_pencil_inv(PencilK<T> k) {
  pencil_inv_body<T, L, NA>(k.a, k.tw);
}
template <class T, int L, int NA>
__global__ __launch_bounds__((PencilShape<T, L, NA>::NT)) void k_pencil_inv_b(const PencilK<T>* __restrict__ ks) {
  pencil_inv_body<T, L, NA>(ks[blockIdx.z].a, ks[blockIdx.z].tw);
}
template <class T, int L, int NA, class AX>
__global__ __launch_bounds__((PencilShape<T, L, NA>::NT)) void k_pencil_fwd(PencilK<T> k) {
  pencil_fwd_body<T, L, NA>(k.a, k.tw);
}
template <class T, int L, int NA>
__global__ __launch_bounds__((PencilShape<T, L, NA>::NT)) void k_pencil_fwd_b(const PencilK<T>* __restrict__ ks) {
  pencil_fwd_body<T, L, NA>(ks[blockIdx.z].a, ks[blockIdx.z].tw);
}

// The single and batched kernels of a pencil plan, and its launch shape.
struct PencilPlanFns {
  const void* single = nullptr;
  const void* batched = nullptr;
  unsigned nt = 0;
};
template <class T, int L, int NA>
static PencilPlanFns pencil_fns_na(bool inverse, bool y) {
  PencilPlanFns p;
  if constexpr (pencil_na_ok<T, L, NA>()) {
    if (inverse)
      p.single = y ? reinterpret_cast<const void*>(&k_pencil_inv<T, L, NA, AxisY>)
                   : reinterpret_cast<const void*>(&k_pencil_inv<T, L, NA, AxisZ>);
    else
      p.single = y ? reinterpret_cast<const void*>(&k_pencil_fwd<T, L, NA, AxisY>)
                   : reinterpret_cast<const void*>(&k_pencil_fwd<T, L, NA, AxisZ>);
    p.batched = inverse ? reinterpret_cast<const void*>(&k_pencil_inv_b<T, L, NA>)
                        : reinterpret_cast<const void*>(&k_pencil_fwd_b<T, L, NA>);
    p.nt = (unsigned)PencilShape<T, L, NA>::NT;
  }
  return p;
}
template <class T, int L>
static PencilPlanFns pencil_fns_l(bool inverse, int na, bool y) {
  switch (na) {
    case 8: return pencil_fns_na<T, L, 8>(inverse, y);
    case 16: return pencil_fns_na<T, L, 16>(inverse, y);
    case 24: return pencil_fns_na<T, L, 24>(inverse, y);
    case 32: return pencil_fns_na<T, L, 32>(inverse, y);
    case 64: return pencil_fns_na<T, L, 64>(inverse, y);
    case 128: return pencil_fns_na<T, L, 128>(inverse, y);
  }
  return PencilPlanFns{};
}
// y: the single-run kernel of the y pass (else the z pass)
template <class T>
static PencilPlanFns pencil_fns(int L, bool inverse, int na, bool y) {
  switch (L) {
#define HY_PL(LL) case LL: return pencil_fns_l<T, LL>(inverse, na, y);
    HY_PL(4) HY_PL(8) HY_PL(12) HY_PL(16) HY_PL(24) HY_PL(32) HY_PL(48) HY_PL(64) HY_PL(96)
    HY_PL(128) HY_PL(256)
#undef HY_PL
  }
  return PencilPlanFns{};
}
template <class T>
static int pencil_default_na(int L) {
  switch (L) {
#define HY_PL(LL) case LL: return pencil_na_default<T, LL>();
    HY_PL(4) HY_PL(8) HY_PL(12) HY_PL(16) HY_PL(24) HY_PL(32) HY_PL(48) HY_PL(64) HY_PL(96)
    HY_PL(128) HY_PL(256)
#undef HY_PL
  }
  return 0;
}
static const int kNaCandidates[] = {8, 16, 24, 32, 64, 128};

// ------------------------------------------------------------ plan autotuner
// The column count per workgroup of each pencil pass shape is chosen by
// timing the candidates once, on the pass's own buffers, the first time the
// shape runs outside a graph capture or a batch recording (FFTW_MEASURE's role
// in the reference, fft.f90:46-47), and kept for the process.  Only
// out-of-place passes are timed (a candidate's run writes the same output as
// the pass itself); in-place passes and passes inside a capture take the
// shape's known choice or the default.  HYDFT_TUNE=0 keeps the defaults.  The
// arithmetic of a column does not depend on the choice, so results are
// bitwise the same under every plan.
struct PlanKey {
  int f64, L, inv, A, B, nf, z;
  bool operator<(const PlanKey& o) const {
    return std::tie(f64, L, inv, A, B, nf, z) < std::tie(o.f64, o.L, o.inv, o.A, o.B, o.nf, o.z);
  }
};
static std::mutex g_plan_mutex;
static std::map<PlanKey, int> g_plans;
static int g_tune = -1;  // -1: from HYDFT_TUNE (default on)
static bool tuning_enabled() {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  if (g_tune < 0) {
    const char* e = std::getenv("HYDFT_TUNE");
    g_tune = (e && e[0] == '0') ? 0 : 1;
  }
  return g_tune == 1;
}
static int plan_lookup(const PlanKey& k) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  auto it = g_plans.find(k);
  return it == g_plans.end() ? 0 : it->second;
}
static void plan_store(const PlanKey& k, int na) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  g_plans[k] = na;
}

// Times candidate launches round-robin (one run of each per round, after a
// warm-up round), so that clock and cache changes during the measurement
// fall on every candidate alike; best[i] is candidate i's fastest run (ms).
struct Cand {
  const void* fn;
  dim3 grid, block;
};
// A pass in the step pipeline reads inputs that an earlier pass left in HBM,
// while a pass repeated back to back finds them in the 256 MiB last-level
// cache.  Each timed run is therefore preceded by a write of this buffer,
// larger than that cache (allocated on the first measurement, kept).
constexpr size_t FLUSH_BYTES = (size_t)512 << 20;
static void* g_flush = nullptr;
static int flush_cache(hipStream_t s) {
  {
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    if (!g_flush && hipMalloc(&g_flush, FLUSH_BYTES) != hipSuccess) {
      (void)hipGetLastError();
      g_flush = nullptr;
      return set_error(HY_ERR_OOM, "pencil plan: cannot allocate the cache-flush buffer");
    }
  }
  HY_HIP(hipMemsetAsync(g_flush, 0, FLUSH_BYTES, s));
  return HY_OK;
}

static int time_candidates(const std::vector<Cand>& c, void** args, hipStream_t s, int rounds, std::vector<float>& best) {
  best.assign(c.size(), 1e30f);
  std::vector<hipEvent_t> ev(2 * c.size());
  for (auto& e : ev) HY_HIP(hipEventCreate(&e));
  int rc = HY_OK;
  for (int r = 0; r <= rounds && rc == HY_OK; ++r) {
    for (size_t i = 0; i < c.size(); ++i) {
      if (flush_cache(s) != HY_OK) {
        rc = HY_ERR_OOM;
        break;
      }
      if (hipEventRecord(ev[2 * i], s) != hipSuccess ||
          hipLaunchKernel(c[i].fn, c[i].grid, c[i].block, args, 0, s) != hipSuccess ||
          hipEventRecord(ev[2 * i + 1], s) != hipSuccess) {
        (void)hipGetLastError();
        rc = set_error(HY_ERR_HIP, "pencil plan timing failed");
        break;
      }
    }
    if (rc != HY_OK || hipStreamSynchronize(s) != hipSuccess) {
      rc = rc != HY_OK ? rc : set_error(HY_ERR_HIP, "pencil plan timing failed");
      break;
    }
    if (r == 0) continue;  // the warm-up round
    for (size_t i = 0; i < c.size(); ++i) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
      best[i] = std::min(best[i], ms);
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
}
// The fastest candidate; the default keeps its place unless another is 3 % faster.
static int pick(const std::vector<int>& na, const std::vector<float>& t, int na0) {
  int bi = -1;
  for (size_t i = 0; i < na.size(); ++i) if (na[i] == na0) bi = (int)i;
  for (size_t i = 0; i < na.size(); ++i)
    if (bi < 0 || t[i] < 0.97f * t[bi]) bi = (int)i;
  return na[bi];
}

template <class T>
static unsigned pencil_grid_x(const PencilArgs& a, int nf, int na) {
  return (unsigned)((a.A + na - 1) / na) * (unsigned)a.B * (unsigned)nf;
}

// The plan of a single (not batched) pass: the known one, or one measured now.
template <class T>
static int choose_na(const PencilArgs& a, bool inverse, int nf, const cpx<T>* tw, hipStream_t s) {
  const PlanKey key{sizeof(T) == 8 ? 1 : 0, a.L, inverse ? 1 : 0, a.A, a.B, nf, 1};
  if (int na = plan_lookup(key)) return na;
  const int na0 = pencil_default_na<T>(a.L);
  bool can = tuning_enabled() && !recorder();
  if (can) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) can = false;
    for (int i = 0; i < nf && can; ++i) can = a.f[i].in != a.f[i].out;  // in-place passes are not repeated
  }
  if (!can) return na0;
  const PencilK<T> k{a, tw};
  void* args[] = {const_cast<PencilK<T>*>(&k)};
  std::vector<Cand> cands;
  std::vector<int> nas;
  for (int na : kNaCandidates) {
    const PencilPlanFns p = pencil_fns<T>(a.L, inverse, na, a.y_from_pencil != 0);
    if (!p.single) continue;
    cands.push_back({p.single, dim3(pencil_grid_x<T>(a, nf, na)), dim3(p.nt)});
    nas.push_back(na);
  }
  std::vector<float> t;
  HY_TRY(time_candidates(cands, args, s, 4, t));
  const int best_na = pick(nas, t, na0);
  plan_store(key, best_na);
  return best_na;
}

// The plan of a batched pass (a batch of `nz` runs, grid z = run): the known
// one for this batch size, or one measured now with the batch's own packed
// arguments (the candidates rewrite the runs' transform slots, which the next
// launch recomputes from the state).
template <class T>
static int batch_plan_t(const std::array<int, 8>& pk, int nz, void* d_args, hipStream_t s, const void** fn,
                        dim3* grid, dim3* block) {
  const int L = pk[2], inv = pk[3], A = pk[4], B = pk[5], nf = pk[6];
  const PlanKey key{pk[1], L, inv, A, B, nf, nz};
  int na = plan_lookup(key);
  PencilArgs a;
  std::memset(&a, 0, sizeof(a));
  a.A = A;
  a.B = B;
  if (!na) {
    na = pk[7];
    if (tuning_enabled()) {
      void* args[] = {&d_args};
      std::vector<Cand> cands;
      std::vector<int> nas;
      for (int c : kNaCandidates) {
        const PencilPlanFns p = pencil_fns<T>(L, inv != 0, c, false);
        if (!p.batched) continue;
        cands.push_back({p.batched, dim3(pencil_grid_x<T>(a, nf, c), 1, (unsigned)nz), dim3(p.nt)});
        nas.push_back(c);
      }
      std::vector<float> t;
      HY_TRY(time_candidates(cands, args, s, 3, t));
      na = pick(nas, t, pk[7]);
      plan_store(key, na);
    }
  }
  const PencilPlanFns p = pencil_fns<T>(L, inv != 0, na, false);
  if (!p.batched) return set_error(HY_ERR_INVALID, "pencil plan: no kernel for the chosen shape");
  *fn = p.batched;
  *grid = dim3(pencil_grid_x<T>(a, nf, na), 1, (unsigned)nz);
  *block = dim3(p.nt);
  return HY_OK;
}
int pencil_batch_plan(const std::array<int, 8>& pk, int nz, void* d_args, hipStream_t s, const void** fn,
                      dim3* grid, dim3* block) {
  return pk[1] ? batch_plan_t<double>(pk, nz, d_args, s, fn, grid, block)
               : batch_plan_t<float>(pk, nz, d_args, s, fn, grid, block);
}

template <class T>
static int dispatch_pencil(const PencilArgs& a, bool inverse, int nf, const cpx<T>* tw, hipStream_t s) {
  const int na = choose_na<T>(a, inverse, nf, tw, s);
  const PencilPlanFns p = pencil_fns<T>(a.L, inverse, na, a.y_from_pencil != 0);
  if (!p.single) return set_error(HY_ERR_UNSUPPORTED, "pencil: unsupported length");
  const PencilK<T> k{a, tw};
  if (Recorder* r = recorder()) {
    LaunchRec it{p.batched, dim3(pencil_grid_x<T>(a, nf, na)), dim3(p.nt), {}};
    it.args.resize(sizeof(k));
    std::memcpy(it.args.data(), &k, sizeof(k));
    it.pencil = {1, sizeof(T) == 8 ? 1 : 0, a.L, inverse ? 1 : 0, a.A, a.B, nf, na};
    r->items.push_back(std::move(it));
    return HY_OK;
  }
  HY_TRY(launch(p.single, p.batched, dim3(pencil_grid_x<T>(a, nf, na)), dim3(p.nt), s, k));
  HY_HIP(hipGetLastError());
  HY_TRACE(inverse ? "pencil_inverse" : "pencil_forward", s);
  return HY_OK;
}

Recorder*& recorder() {
  static thread_local Recorder* r = nullptr;
  return r;
}

int pencil_pass(const Grid& g, int axis, bool inverse, bool trunc, int A, int B,
                const PencilField* f, int nf, hipStream_t s, const int* halt) {
  if (nf <= 0) return HY_OK;
  PencilArgs a;
  std::memset(&a, 0, sizeof(a));
  a.halt = halt;
  a.A = A;
  a.B = B;
  a.L = axis == 1 ? g.ny : g.nz;
  a.K = trunc ? (axis == 1 ? g.Ky : g.Kz) : a.L;
  a.h = trunc ? (a.K + 1) / 2 : a.L;
  a.scale = 1.0 / a.L;
  const double twopi = 2.0 * 3.14159265358979323846264338327950288;
  a.dkax = twopi / (axis == 1 ? g.ly : g.lz);
  a.dkx = twopi / g.lx;
  a.ax_x = trunc ? g.Kx : g.nkx;
  a.dky = twopi / g.ly;
  a.y_from_pencil = axis == 1 ? 1 : 0;
  a.field_fastest = inverse && axis == 2 ? 1 : 0;
  if (axis == 2 && g.ny > 1) {
    a.ay_n = trunc ? g.Ky : g.ny;
    a.ay_h = trunc ? (g.Ky + 1) / 2 : g.ny;
    a.ay_L = g.ny;
  }
  // the kernels address their inputs with 32-bit byte offsets
  if ((double)A * B * std::max(a.K, a.L) * sizeof(c64) >= 4294967296.0)
    return set_error(HY_ERR_INVALID, "grid too large: a pencil field must stay below 4 GiB");
  const bool f32 = (g.precision == HY_PREC_MIXED) && trunc;
  for (int i0 = 0; i0 < nf; i0 += MAXPF) {
    int n = std::min(MAXPF, nf - i0);
    for (int i = 0; i < n; ++i) a.f[i] = f[i0 + i];
    const TwiddleSet& t = g.tw(a.L);
    if (f32) HY_TRY(dispatch_pencil<float>(a, inverse, n, t.d32, s));
    else HY_TRY(dispatch_pencil<double>(a, inverse, n, t.d64, s));
  }
  return HY_OK;
}

// ------------------------------------------------------ plain x-line passes
// Two real lines per block packed as one complex FFT of length NX.
template <int NX>
__global__ __launch_bounds__(64) void k_xplain_r2c(const double* __restrict__ u, c64* __restrict__ out,
                                                   int nlines, int Kout, double scale,
                                                   const c64* __restrict__ twg) {
  __shared__ c64 buf[NX];
  __shared__ c64 tw[NX];
  const int tid = threadIdx.x;
  const int l0 = blockIdx.x * 2, l1 = l0 + 1;
  const bool has1 = l1 < nlines;
  for (int j = tid; j < NX; j += 64) tw[j] = twg[j];
  for (int x = tid; x < NX; x += 64)
    buf[x] = mk<double>(u[(size_t)l0 * NX + x], has1 ? u[(size_t)l1 * NX + x] : 0.0);
  __syncthreads();
  fft_lds<double, NX, 1, 64, false>(buf, tw, tid);
  for (int k = tid; k < Kout; k += 64) {
    c64 Z = buf[k], Zc = conjc(buf[(NX - k) % NX]);
    c64 A = scal(Z + Zc, 0.5 * scale);
    c64 D = Z - Zc;
    out[(size_t)l0 * Kout + k] = A;
    if (has1) out[(size_t)l1 * Kout + k] = mk<double>(0.5 * scale * D.y, -0.5 * scale * D.x);
  }
}

template <int NX>
__global__ __launch_bounds__(64) void k_xplain_c2r(const c64* __restrict__ in, double* __restrict__ u,
                                                   int nlines, int Kin, const c64* __restrict__ twg) {
  __shared__ c64 buf[NX];
  __shared__ c64 tw[NX];
  const int tid = threadIdx.x;
  const int l0 = blockIdx.x * 2, l1 = l0 + 1;
  const bool has1 = l1 < nlines;
  for (int j = tid; j < NX; j += 64) tw[j] = twg[j];
  for (int k = tid; k < NX; k += 64) {
    const int kk = k <= NX / 2 ? k : NX - k;
    c64 A = kk < Kin ? in[(size_t)l0 * Kin + kk] : mk<double>(0, 0);
    c64 B = (has1 && kk < Kin) ? in[(size_t)l1 * Kin + kk] : mk<double>(0, 0);
    if (kk == 0 || 2 * kk == NX) { A.y = 0; B.y = 0; }
    if (k > NX / 2) { A = conjc(A); B = conjc(B); }
    buf[k] = mk<double>(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  fft_lds<double, NX, 1, 64, true>(buf, tw, tid);
  for (int x = tid; x < NX; x += 64) {
    u[(size_t)l0 * NX + x] = buf[x].x;
    if (has1) u[(size_t)l1 * NX + x] = buf[x].y;
  }
}

static int xplain(const Grid& g, bool fwd, const void* in, void* out, int K, hipStream_t s) {
  const int nlines = g.ny * g.nz;
  dim3 grid((nlines + 1) / 2);
  const c64* tw = g.tw(g.nx).d64;
  switch (g.nx) {
#define HY_XP(LL)                                                                             \
  case LL:                                                                                    \
    if (fwd) hipLaunchKernelGGL((k_xplain_r2c<LL>), grid, dim3(64), 0, s, (const double*)in,  \
                                (c64*)out, nlines, K, 1.0 / LL, tw);                          \
    else hipLaunchKernelGGL((k_xplain_c2r<LL>), grid, dim3(64), 0, s, (const c64*)in,         \
                            (double*)out, nlines, K, tw);                                     \
    break;
    HY_XP(4) HY_XP(8) HY_XP(12) HY_XP(16) HY_XP(24) HY_XP(32) HY_XP(48) HY_XP(64) HY_XP(96)
    HY_XP(128) HY_XP(256)
#undef HY_XP
    default: return set_error(HY_ERR_UNSUPPORTED, "x pass: unsupported length");
  }
  HY_HIP(hipGetLastError());
  HY_TRACE(fwd ? "xplain_r2c" : "xplain_c2r", s);
  return HY_OK;
}

// Full-layout (nkx, ny, nz) transforms used by the drop-in operators.
// r2c: u (real) -> uhat;  c2r: work (spectrum, overwritten) -> u.
static int fft_r2c_full(const Grid& g, const double* u, c64* uhat, hipStream_t s) {
  HY_TRY(xplain(g, true, u, uhat, g.nkx, s));
  if (g.ny > 1) {
    PencilField f{uhat, uhat, MUL_NONE, 1, 0.0};
    HY_TRY(pencil_pass(g, 1, false, false, g.nkx, g.nz, &f, 1, s));
  }
  if (g.nz > 1) {
    PencilField f{uhat, uhat, MUL_NONE, 1, 0.0};
    HY_TRY(pencil_pass(g, 2, false, false, g.nkx * g.ny, 1, &f, 1, s));
  }
  return HY_OK;
}
static int fft_c2r_full_inplace(const Grid& g, c64* work, double* u, hipStream_t s) {
  if (g.nz > 1) {
    PencilField f{work, work, MUL_NONE, 1, 0.0};
    HY_TRY(pencil_pass(g, 2, true, false, g.nkx * g.ny, 1, &f, 1, s));
  }
  if (g.ny > 1) {
    PencilField f{work, work, MUL_NONE, 1, 0.0};
    HY_TRY(pencil_pass(g, 1, true, false, g.nkx, g.nz, &f, 1, s));
  }
  return xplain(g, false, work, u, g.nkx, s);
}

// --------------------------------------------- full-spectrum multiply kernels
enum SpecOp { SOP_COPY = 0, SOP_MASK, SOP_IK_MASK, SOP_NEGK2_MASK, SOP_KERNEL_MASK, SOP_ACC_IK };

struct FullSpecArgs {
  int nkx, ny, nz, Kx, My, Mz;
  const double *kx, *ky, *kz;
};

__device__ inline bool full_kept(const FullSpecArgs& a, int i, int j, int k) {
  int my = j <= a.ny / 2 ? j : j - a.ny;
  int mz = k <= a.nz / 2 ? k : k - a.nz;
  return i < a.Kx && (my < 0 ? -my : my) <= a.My && (mz < 0 ? -mz : mz) <= a.Mz;
}

// kdir: wavenumbers of the operator's axis, indexed by (e / kstride) % kn; the
// host selects the array (no pointer selection inside the kernel).
__global__ void k_specop(FullSpecArgs a, int op, const double* __restrict__ kdir, size_t kstride, int kn,
                         const c64* __restrict__ in, const double* __restrict__ kern, c64* __restrict__ out) {
  size_t n = (size_t)a.nkx * a.ny * a.nz;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int i = e % a.nkx;
    int j = (e / a.nkx) % a.ny;
    int k = e / ((size_t)a.nkx * a.ny);
    double m = full_kept(a, i, j, k) ? 1.0 : 0.0;
    c64 v = in[e];
    double kd = kdir[(e / kstride) % (size_t)kn];
    switch (op) {
      case SOP_COPY: out[e] = v; break;
      case SOP_MASK: out[e] = scal(v, m); break;
      case SOP_IK_MASK: out[e] = scal(mk<double>(-kd * v.y, kd * v.x), m); break;
      case SOP_NEGK2_MASK: {
        double k2 = a.kx[i] * a.kx[i] + a.ky[j] * a.ky[j] + a.kz[k] * a.kz[k];
        out[e] = scal(v, -k2 * m);
        break;
      }
      case SOP_KERNEL_MASK: out[e] = scal(v, kern[e] * m); break;
      case SOP_ACC_IK: out[e] = out[e] + mk<double>(-kd * v.y, kd * v.x); break;
    }
  }
}

static FullSpecArgs full_args(const Grid& g) {
  FullSpecArgs a;
  a.nkx = g.nkx; a.ny = g.ny; a.nz = g.nz; a.Kx = g.Kx; a.My = g.My; a.Mz = g.Mz;
  a.kx = g.d_kx; a.ky = g.d_ky; a.kz = g.d_kz;
  return a;
}
static int specop(const Grid& g, int op, int dir, const c64* in, const double* kern, c64* out,
                  hipStream_t s) {
  size_t n = g.nfull();
  int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
  const double* kdir = dir == 0 ? g.d_kx : (dir == 1 ? g.d_ky : g.d_kz);
  size_t kstride = dir == 0 ? 1 : (dir == 1 ? (size_t)g.nkx : (size_t)g.nkx * g.ny);
  int kn = dir == 0 ? g.nkx : (dir == 1 ? g.ny : g.nz);
  hipLaunchKernelGGL(k_specop, dim3(blocks), dim3(256), 0, s, full_args(g), op, kdir, kstride, kn, in, kern, out);
  HY_HIP(hipGetLastError());
  HY_TRACE("specop", s);
  return HY_OK;
}

// Deterministic two-level sum of a real array.
__global__ void k_sum_partial(const double* __restrict__ u, size_t n, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    acc += u[e];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_sum_final(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double sh[256];
  double acc = 0.0;
  for (int e = threadIdx.x; e < np; e += blockDim.x) acc += part[e];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// -------------------------------------------------------- drop-in scratch
// Eight full half-spectra, allocated on first use of a drop-in operator.
static int grid_scratch(Grid& g, c64** s) {
  const size_t need = sizeof(c64) * g.nfull() * 8 + sizeof(double) * 1024;
  if (g.scratch_bytes < need) {
    if (g.d_scratch) HY_HIP(hipFree(g.d_scratch));
    g.d_scratch = nullptr;
    if (hipMalloc(&g.d_scratch, need) != hipSuccess) {
      (void)hipGetLastError();
      g.scratch_bytes = 0;
      return set_error(HY_ERR_OOM, "grid: cannot allocate drop-in scratch");
    }
    g.scratch_bytes = need;
  }
  for (int i = 0; i < 8; ++i) s[i] = reinterpret_cast<c64*>(g.d_scratch) + g.nfull() * i;
  return HY_OK;
}
static double* grid_red(Grid& g) {
  return reinterpret_cast<double*>(reinterpret_cast<c64*>(g.d_scratch) + g.nfull() * 8);
}

static int spec_grad_impl(Grid& g, const double* d_u, double* d_du3, hipStream_t st) {
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  HY_TRY(fft_r2c_full(g, d_u, s[1], st));
  for (int d = 0; d < 3; ++d) {
    double* dst = d_du3 + g.nreal() * d;
    if (d >= g.ndim) {
      HY_HIP(hipMemsetAsync(dst, 0, sizeof(double) * g.nreal(), st));
      continue;
    }
    HY_TRY(specop(g, SOP_IK_MASK, d, s[1], nullptr, s[0], st));
    HY_TRY(fft_c2r_full_inplace(g, s[0], dst, st));
  }
  return HY_OK;
}

// entry points for the model translation unit
int fft_r2c_full_api(Grid& g, const double* u, c64* uhat, hipStream_t s) { return fft_r2c_full(g, u, uhat, s); }
// destroys its input spectrum
int fft_c2r_full_api(Grid& g, const c64* uhat, double* u, hipStream_t s) {
  return fft_c2r_full_inplace(g, const_cast<c64*>(uhat), u, s);
}
int spec_grad_api(Grid& g, const double* u, double* du3, hipStream_t s) { return spec_grad_impl(g, u, du3, s); }

}  // namespace hy

// ======================================================================= ABI
using namespace hy;


extern "C" {

int hy_init(int device) {
  int n = 0;
  HY_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_error(HY_ERR_INVALID, "hy_init: no such device");
  HY_HIP(hipSetDevice(device));
  return HY_OK;
}
const char* hy_last_error(void) { return g_last_error.c_str(); }
int hy_version(void) { return 100; }
int hy_device_sync(void) {
  HY_HIP(hipDeviceSynchronize());
  return HY_OK;
}

int hy_device_alloc(unsigned long long bytes, void** d_out) {
  if (!d_out) return set_error(HY_ERR_INVALID, "null argument");
  if (hipMalloc(d_out, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    *d_out = nullptr;
    return set_error(HY_ERR_OOM, "hy_device_alloc: out of device memory");
  }
  return HY_OK;
}
int hy_device_free(void* d) {
  if (d) HY_HIP(hipFree(d));
  return HY_OK;
}
int hy_copy_to_device(void* d_dst, const void* h_src, unsigned long long bytes) {
  HY_HIP(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return HY_OK;
}
int hy_copy_to_host(void* h_dst, const void* d_src, unsigned long long bytes) {
  HY_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return HY_OK;
}

int hy_grid_create(int nx, int ny, int nz, double lx, double ly, double lz, int dealias,
                   int precision, hy_grid* out) {
  if (!out) return set_error(HY_ERR_INVALID, "hy_grid_create: null output");
  auto* h = new hy_grid_s();
  int rc = grid_build(h->g, nx, ny, nz, lx, ly, lz, dealias, precision);
  if (rc != HY_OK) {
    grid_free(h->g);
    delete h;
    return rc;
  }
  *out = h;
  return HY_OK;
}
int hy_grid_destroy(hy_grid g) {
  if (!g) return HY_OK;
  grid_free(g->g);
  delete g;
  return HY_OK;
}
int hy_grid_query(hy_grid g, int* nkx, int* ndim, int* kept3, double* volume, double* kmax) {
  if (!g) return set_error(HY_ERR_INVALID, "null grid");
  if (nkx) *nkx = g->g.nkx;
  if (ndim) *ndim = g->g.ndim;
  if (kept3) { kept3[0] = g->g.Kx; kept3[1] = g->g.Ky; kept3[2] = g->g.Kz; }
  if (volume) *volume = g->g.volume;
  if (kmax) *kmax = g->g.kmax_dealiased;
  return HY_OK;
}
int hy_grid_mask(hy_grid gh, double* h_mask) {
  if (!gh || !h_mask) return set_error(HY_ERR_INVALID, "null argument");
  const Grid& g = gh->g;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nkx; ++i) {
        int my = j <= g.ny / 2 ? j : j - g.ny, mz = k <= g.nz / 2 ? k : k - g.nz;
        bool kept = i < g.Kx && std::abs(my) <= g.My && std::abs(mz) <= g.Mz;
        h_mask[((size_t)k * g.ny + j) * g.nkx + i] = kept ? 1.0 : 0.0;
      }
  return HY_OK;
}

int hy_fft_r2c(hy_grid gh, const double* d_u, double* d_uhat, void* stream) {
  if (!gh || !d_u || !d_uhat) return set_error(HY_ERR_INVALID, "null argument");
  return fft_r2c_full(gh->g, d_u, reinterpret_cast<c64*>(d_uhat), (hipStream_t)stream);
}

int hy_fft_c2r(hy_grid gh, const double* d_uhat, double* d_u, void* stream) {
  if (!gh || !d_u || !d_uhat) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_HIP(hipMemcpyAsync(s[0], d_uhat, sizeof(c64) * g.nfull(), hipMemcpyDeviceToDevice, st));
  return fft_c2r_full_inplace(g, s[0], d_u, st);
}

int hy_spec_grad_hat(hy_grid gh, const double* d_uhat, int idir, double* d_du, void* stream) {
  if (!gh || !d_uhat || !d_du) return set_error(HY_ERR_INVALID, "null argument");
  if (idir < 1 || idir > 3) return set_error(HY_ERR_INVALID, "spectral: bad direction");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_TRY(specop(g, SOP_IK_MASK, idir - 1, reinterpret_cast<const c64*>(d_uhat), nullptr, s[0], st));
  return fft_c2r_full_inplace(g, s[0], d_du, st);
}

int hy_spec_grad(hy_grid gh, const double* d_u, double* d_du3, void* stream) {
  if (!gh || !d_u || !d_du3) return set_error(HY_ERR_INVALID, "null argument");
  return spec_grad_impl(gh->g, d_u, d_du3, (hipStream_t)stream);
}

int hy_spec_lap(hy_grid gh, const double* d_u, double* d_lu, void* stream) {
  if (!gh || !d_u || !d_lu) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_TRY(fft_r2c_full(g, d_u, s[1], st));
  HY_TRY(specop(g, SOP_NEGK2_MASK, 0, s[1], nullptr, s[0], st));
  return fft_c2r_full_inplace(g, s[0], d_lu, st);
}

int hy_spec_div(hy_grid gh, const double* d_v3, double* d_dv, void* stream) {
  if (!gh || !d_v3 || !d_dv) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_HIP(hipMemsetAsync(s[0], 0, sizeof(c64) * g.nfull(), st));
  for (int d = 0; d < g.ndim; ++d) {
    HY_TRY(fft_r2c_full(g, d_v3 + g.nreal() * d, s[1], st));
    HY_TRY(specop(g, SOP_ACC_IK, d, s[1], nullptr, s[0], st));
  }
  HY_TRY(specop(g, SOP_MASK, 0, s[0], nullptr, s[0], st));
  return fft_c2r_full_inplace(g, s[0], d_dv, st);
}

int hy_spec_convolve(hy_grid gh, const double* d_kernel, const double* d_uhat, double* d_cu,
                     void* stream) {
  if (!gh || !d_kernel || !d_uhat || !d_cu) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_TRY(specop(g, SOP_KERNEL_MASK, 0, reinterpret_cast<const c64*>(d_uhat), d_kernel, s[0], st));
  return fft_c2r_full_inplace(g, s[0], d_cu, st);
}

int hy_spec_dealias(hy_grid gh, double* d_u, void* stream) {
  if (!gh || !d_u) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_TRY(fft_r2c_full(g, d_u, s[0], st));
  HY_TRY(specop(g, SOP_MASK, 0, s[0], nullptr, s[0], st));
  return fft_c2r_full_inplace(g, s[0], d_u, st);
}

int hy_spec_mean(hy_grid gh, const double* d_u, double* h_mean, void* stream) {
  if (!gh || !d_u || !h_mean) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  double* red = grid_red(g);
  const int np = 512;
  hipLaunchKernelGGL(k_sum_partial, dim3(np), dim3(256), 0, st, d_u, g.nreal(), red + 1);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, red + 1, np, red);
  HY_HIP(hipGetLastError());
  double sum = 0.0;
  HY_HIP(hipMemcpyAsync(&sum, red, sizeof(double), hipMemcpyDeviceToHost, st));
  HY_HIP(hipStreamSynchronize(st));
  *h_mean = sum / (double)g.nreal();
  return HY_OK;
}

}  // extern "C"

// ---- spec_mean without device allocation
// The partial sums of hy_spec_mean_noalloc live in this file-scope device
// array (one copy per GPU, held until the process ends): MEAN_PARTS partials
// and the total.  Host code never passes the array's name to a kernel (in
// host code it is a stand-in, not a device address); it passes the address
// that hipGetSymbolAddress returns for the current device.
namespace hy {
constexpr int MEAN_PARTS = 512;
constexpr int MEAN_SCRATCH = 1024;  // doubles (8 KiB)
static_assert(MEAN_PARTS + 1 <= MEAN_SCRATCH, "mean scratch too small");
}  // namespace hy
static __device__ __align__(256) double g_mean_scratch[hy::MEAN_SCRATCH];

namespace hy {
constexpr int MAX_DEVICES = 64;
static std::mutex g_mean_mutex;
static double* g_mean_ptr[MAX_DEVICES] = {nullptr};

// Device address of g_mean_scratch on the current device, looked up on the
// first call on that device.
static int mean_scratch_ptr(double** out) {
  int dev = -1;
  HY_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(HY_ERR_INVALID, "spec_mean_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mean_mutex);
  if (!g_mean_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_mean_scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(HY_ERR_HIP, std::string("spec_mean_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    }
    g_mean_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_mean_ptr[dev];
  return HY_OK;
}
}  // namespace hy

extern "C" {

int hy_spec_mean_noalloc(hy_grid gh, const double* d_u, double* h_mean, void* stream) {
  if (!gh || !d_u || !h_mean) return set_error(HY_ERR_INVALID, "null argument");
  const Grid& g = gh->g;
  const size_t n = g.nreal();
  // the partials do not grow with the field: any grid the library accepts
  // fits; the check guards the fixed layout of the array
  if (n == 0 || MEAN_PARTS + 1 > MEAN_SCRATCH)
    return set_error(HY_ERR_INVALID, "spec_mean_noalloc: field size outside the scratch layout (" +
                                         std::to_string(MEAN_SCRATCH) + " doubles)");
  double* red = nullptr;
  HY_TRY(mean_scratch_ptr(&red));
  hipStream_t st = (hipStream_t)stream;
  // the same kernels and partial count as hy_spec_mean: the same sum, bit for bit
  hipLaunchKernelGGL(k_sum_partial, dim3(MEAN_PARTS), dim3(256), 0, st, d_u, n, red + 1);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, red + 1, MEAN_PARTS, red);
  HY_HIP(hipGetLastError());
  double sum = 0.0;
  HY_HIP(hipMemcpyAsync(&sum, red, sizeof(double), hipMemcpyDeviceToHost, st));
  HY_HIP(hipStreamSynchronize(st));
  *h_mean = sum / (double)n;
  return HY_OK;
}

int hy_spec_mode(hy_grid gh, const double* d_u, int mx, int my, int mz, double* h_re_im,
                 void* stream) {
  if (!gh || !d_u || !h_re_im) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  if (mx < 0 || mx >= g.nkx) return set_error(HY_ERR_INVALID, "spec_mode: mx out of range");
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  HY_TRY(fft_r2c_full(g, d_u, s[0], st));
  int jy = ((my % g.ny) + g.ny) % g.ny, kz = ((mz % g.nz) + g.nz) % g.nz;
  size_t idx = ((size_t)kz * g.ny + jy) * g.nkx + mx;
  c64 v;
  HY_HIP(hipMemcpyAsync(&v, s[0] + idx, sizeof(c64), hipMemcpyDeviceToHost, st));
  HY_HIP(hipStreamSynchronize(st));
  h_re_im[0] = v.x;
  h_re_im[1] = v.y;
  return HY_OK;
}

// stress_pi0 (stress.f90:60-83) evaluated spectrally: gu_ij = B(i k_j F(u_i) mask),
// so each Pi component is the backward transform of the same linear combination.
__global__ void k_pi0_combo(FullSpecArgs a, int comp, double eta, double xi,
                            const c64* __restrict__ u0, const c64* __restrict__ u1,
                            const c64* __restrict__ u2, int ndim, c64* __restrict__ out) {
  size_t n = (size_t)a.nkx * a.ny * a.nz;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int i = e % a.nkx, j = (e / a.nkx) % a.ny, k = e / ((size_t)a.nkx * a.ny);
    double m = full_kept(a, i, j, k) ? 1.0 : 0.0;
    const double k0 = a.kx[i], k1 = a.ky[j], k2 = a.kz[k];
    const c64 z = mk<double>(0, 0);
    const c64 v0 = ndim > 0 ? u0[e] : z, v1 = ndim > 1 ? u1[e] : z, v2 = ndim > 2 ? u2[e] : z;
    // spectrum of du_i/dx_j is i k_j uhat_i
    const c64 g00 = mk<double>(-k0 * v0.y, k0 * v0.x), g11 = mk<double>(-k1 * v1.y, k1 * v1.x),
              g22 = mk<double>(-k2 * v2.y, k2 * v2.x);
    const c64 divu = g00 + g11 + g22;
    c64 r;
    if (comp < 3) {
      const c64 gii = comp == 0 ? g00 : (comp == 1 ? g11 : g22);
      r = scal(scal(gii, 2.0) - scal(divu, 2.0 / 3.0), eta) + scal(divu, xi);
    } else if (comp == 3) {  // g01 + g10
      r = scal(mk<double>(-k1 * v0.y, k1 * v0.x) + mk<double>(-k0 * v1.y, k0 * v1.x), eta);
    } else if (comp == 4) {  // g02 + g20
      r = scal(mk<double>(-k2 * v0.y, k2 * v0.x) + mk<double>(-k0 * v2.y, k0 * v2.x), eta);
    } else {                 // g12 + g21
      r = scal(mk<double>(-k2 * v1.y, k2 * v1.x) + mk<double>(-k1 * v2.y, k1 * v2.x), eta);
    }
    out[e] = scal(r, m);
  }
}

int hy_stress_pi0(hy_grid gh, double eta, double xi, const double* d_u3, double* d_pi6, void* stream) {
  if (!gh || !d_u3 || !d_pi6) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  for (int d = 0; d < g.ndim; ++d) HY_TRY(fft_r2c_full(g, d_u3 + g.nreal() * d, s[1 + d], st));
  size_t n = g.nfull();
  int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
  for (int c = 0; c < 6; ++c) {
    hipLaunchKernelGGL(k_pi0_combo, dim3(blocks), dim3(256), 0, st, full_args(g), c, eta, xi, s[1],
                       s[2], s[3], g.ndim, s[0]);
    HY_HIP(hipGetLastError());
    HY_TRY(fft_c2r_full_inplace(g, s[0], d_pi6 + g.nreal() * c, st));
  }
  return HY_OK;
}

int hy_stress_div(hy_grid gh, const double* d_pi6, double* d_f3, void* stream) {
  if (!gh || !d_pi6 || !d_f3) return set_error(HY_ERR_INVALID, "null argument");
  Grid& g = gh->g;
  c64* s[8];
  HY_TRY(grid_scratch(g, s));
  hipStream_t st = (hipStream_t)stream;
  static const int idx[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};  // stress.f90:92
  for (int c = 0; c < 6; ++c) HY_TRY(fft_r2c_full(g, d_pi6 + g.nreal() * c, s[1 + c], st));
  for (int i = 0; i < 3; ++i) {
    double* dst = d_f3 + g.nreal() * i;
    if (i >= g.ndim) {
      HY_HIP(hipMemsetAsync(dst, 0, sizeof(double) * g.nreal(), st));
      continue;
    }
    HY_HIP(hipMemsetAsync(s[0], 0, sizeof(c64) * g.nfull(), st));
    for (int j = 0; j < g.ndim; ++j) HY_TRY(specop(g, SOP_ACC_IK, j, s[1 + idx[i][j]], nullptr, s[0], st));
    HY_TRY(specop(g, SOP_MASK, 0, s[0], nullptr, s[0], st));
    HY_TRY(fft_c2r_full_inplace(g, s[0], dst, st));
  }
  return HY_OK;
}

// ---- pencil plans (the autotuner's choices)
int hy_plan_tuning(int on) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  g_tune = on ? 1 : 0;
  return HY_OK;
}

int hy_plan_reset(void) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  g_plans.clear();
  return HY_OK;
}

int hy_plan_list(int max, int* out8) {
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  int n = 0;
  for (const auto& kv : g_plans) {
    if (out8 && n < max) {
      const PlanKey& k = kv.first;
      const int v[8] = {k.f64, k.L, k.inv, k.A, k.B, k.nf, k.z, kv.second};
      std::memcpy(out8 + 8 * n, v, sizeof(v));
    }
    ++n;
  }
  return n;
}

}  // extern "C"
