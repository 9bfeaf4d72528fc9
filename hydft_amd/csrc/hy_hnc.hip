// Batched OZ-HNC iteration on the GPU: the direct solve of oz_hnc.f90
// (Picard + Ng iteration on gamma_s = h - c_s, restarts with halved mixing)
// for many state points at once.  The radial Fourier-Bessel pair is a DST-I
// of length nr (N + 1 = 8193 = 3 * 2731 at the default nr: no short FFT), so
// each transform of all active solves is one FP64 product with the DST-I
// matrix: passes that rebuild the matrix entries from a table in LDS (or
// stream the matrix from HBM for large nr), and rocBLAS dgemm on the matrix
// cores for more than DST_STREAM_MAX active solves.  The
// pointwise closure, the OZ relation, the residual and the Ng dot products
// are HIP kernels with one workgroup per solve.  Every solve keeps its own
// control state on the host (iteration count, Ng history length, cooldown,
// mixing, restarts), with the rules of the reference's loop.
#include <rocblas/rocblas.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "hy_internal.h"

namespace hy {
namespace {

constexpr double PI_ = 3.14159265358979323846264338327950288;
constexpr int HW = 256;  // threads per solve workgroup

// S[j][m] = 2 sin(pi (j+1)(m+1) / (N+1)): the unnormalised DST-I (RODFT00);
// the product is reduced mod 2(N+1) in integers before the sine.
__global__ void k_dst_matrix(double* __restrict__ S, int n) {
  const size_t total = (size_t)n * n;
  const long long two = 2LL * (n + 1);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const long long j = (long long)(e % n), m = (long long)(e / n);
    const long long q = ((j + 1) * (m + 1)) % two;
    S[e] = 2.0 * sin(PI_ * (double)q / (double)(n + 1));
  }
}

// The matrix entry (j, m) is T[q], q = (j+1)(m+1) mod 2(n+1): the same
// value from a table of 2(n+1) entries.
__global__ void k_dst_table(double* __restrict__ T, int n) {
  const int two = 2 * (n + 1);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < two; q += gridDim.x * blockDim.x)
    T[q] = 2.0 * sin(PI_ * (double)q / (double)(n + 1));
}

struct Dims {
  int nr;
  double dr, dk, cap;
};

// Y[:, p] = S X[:, p] for p < na <= NA_MAX: one streaming pass over the
// matrix (used when its table does not fit in LDS, see k_dst_tab).  Up to
// DST_STREAM_MAX active solves take ceil(na / NA_MAX) such passes; a library GEMM with so few columns keeps only M/128 workgroups busy
// (65 at nr = 8193) and is 2-7x slower.  Block (row tile, chunk): thread = row
// j, the chunk's X rows staged in LDS; per-chunk partial sums, added in a
// fixed order by k_dst_sum (deterministic).
constexpr int NA_MAX = 32, DCH = 32, DTILE = 256, DSTAGE = 256;
constexpr int DST_STREAM_MAX = 192;
template <int NA>
__global__ __launch_bounds__(DTILE) void k_dst_few(const double* __restrict__ S, const double* __restrict__ X,
                                                   double* __restrict__ part, int nr, int na) {
  const int j = blockIdx.x * DTILE + threadIdx.x;
  const int c = blockIdx.y;
  const int clen = (nr + DCH - 1) / DCH;
  const int m0 = c * clen, m1 = min(nr, m0 + clen);
  constexpr int ST = NA <= 16 ? DSTAGE : DSTAGE * 16 / NA;  // X rows staged: <= 32 KB of LDS
  __shared__ double xs[ST][NA];
  double acc[NA];
#pragma unroll
  for (int p = 0; p < NA; ++p) acc[p] = 0.0;
  for (int mb = m0; mb < m1; mb += ST) {
    const int mn = min(ST, m1 - mb);
    __syncthreads();
    for (int e = threadIdx.x; e < ST * NA; e += DTILE) {
      const int mm = e / NA, p = e % NA;
      xs[mm][p] = (mm < mn && p < na) ? X[(size_t)p * nr + mb + mm] : 0.0;
    }
    __syncthreads();
    if (j < nr) {
      const double* sp = S + (size_t)mb * nr + j;
      int mm = 0;
      for (; mm + 4 <= mn; mm += 4) {  // four independent loads in flight
        const double s0 = sp[(size_t)mm * nr], s1 = sp[(size_t)(mm + 1) * nr];
        const double s2 = sp[(size_t)(mm + 2) * nr], s3 = sp[(size_t)(mm + 3) * nr];
#pragma unroll
        for (int p = 0; p < NA; ++p)
          acc[p] = fma(s3, xs[mm + 3][p], fma(s2, xs[mm + 2][p], fma(s1, xs[mm + 1][p], fma(s0, xs[mm][p], acc[p]))));
      }
      for (; mm < mn; ++mm) {
        const double sv = sp[(size_t)mm * nr];
#pragma unroll
        for (int p = 0; p < NA; ++p) acc[p] = fma(sv, xs[mm][p], acc[p]);
      }
    }
  }
  if (j < nr) {
#pragma unroll
    for (int p = 0; p < NA; ++p)
      if (p < na) part[((size_t)c * NA_MAX + p) * nr + j] = acc[p];
  }
}
// The same product without reading the matrix: the workgroup holds the
// 2(n+1)-entry table in LDS and walks row j's entries q = (j+1)(m+1) mod
// 2(n+1) by adding j+1 per column.  A workgroup covers TTILE rows, R per
// thread (R = 2 for wide passes: each staged X value then feeds two FMAs);
// grid (row tiles, chunks), per-chunk partial sums as in k_dst_few.  Needs
// 2(n+1) <= TAB_MAX.
constexpr int TAB_MAX = 2 * (8192 + 1), TTILE = 1024;
template <int NA>
constexpr int tab_rows() { return NA >= 32 ? 2 : 1; }
template <int NA>
__global__ __launch_bounds__(TTILE / tab_rows<NA>()) void k_dst_tab(const double* __restrict__ T,
                                                                  const double* __restrict__ X,
                                                                  double* __restrict__ part, int nr, int na, int nch) {
  constexpr int R = tab_rows<NA>(), NT = TTILE / R;
  constexpr int ST = NA <= 8 ? 256 : 2048 / NA;  // X rows staged: <= 16 KB of LDS beside the table
  __shared__ double tab[TAB_MAX];
  __shared__ double xs[ST][NA];
  const int two = 2 * (nr + 1);
  for (int e = threadIdx.x; e < two; e += NT) tab[e] = T[e];
  const int c = blockIdx.y;
  const int clen = (nr + nch - 1) / nch;
  const int m0 = c * clen, m1 = min(nr, m0 + clen);
  int j[R], step[R], q[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    j[r] = blockIdx.x * TTILE + r * NT + threadIdx.x;
    // rows past the last (their sums are discarded) stay on entry 0; for a
    // live row step = j + 1 <= nr < two, so one conditional subtraction keeps
    // q in [0, two)
    const bool live = j[r] < nr;
    step[r] = live ? j[r] + 1 : 0;
    q[r] = live ? (int)(((long long)(j[r] + 1) * (long long)(m0 + 1)) % two) : 0;
  }
  double acc[R][NA];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int p = 0; p < NA; ++p) acc[r][p] = 0.0;
  for (int mb = m0; mb < m1; mb += ST) {
    const int mn = min(ST, m1 - mb);
    __syncthreads();  // the table (first pass) and the previous stage
    for (int e = threadIdx.x; e < ST * NA; e += NT) {
      const int mm = e / NA, p = e % NA;
      xs[mm][p] = (mm < mn && p < na) ? X[(size_t)p * nr + mb + mm] : 0.0;
    }
    __syncthreads();
    for (int mm = 0; mm < mn; ++mm) {
      double sv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        sv[r] = tab[q[r]];
        q[r] += step[r];
        if (q[r] >= two) q[r] -= two;
      }
#pragma unroll
      for (int p = 0; p < NA; ++p) {
        const double x = xs[mm][p];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r][p] = fma(sv[r], x, acc[r][p]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (j[r] < nr) {
#pragma unroll
      for (int p = 0; p < NA; ++p)
        if (p < na) part[((size_t)c * NA_MAX + p) * nr + j[r]] = acc[r][p];
    }
  }
}

// Y[:, p] = sum over the nch chunks of the partial sums, in chunk order
__global__ void k_dst_sum(const double* __restrict__ part, double* __restrict__ Y, int nr, int na, int nch) {
  const size_t n = (size_t)na * nr;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(e / nr), j = (int)(e % nr);
    double v = 0.0;
    for (int c = 0; c < nch; ++c) v += part[((size_t)c * NA_MAX + p) * nr + j];
    Y[e] = v;
  }
}

// X[:, p] = r * (exp(min(-bvs + gs, cap)) - 1 - gs) for active solve act[p]
__global__ __launch_bounds__(HW) void k_closure(const int* __restrict__ act, Dims d, const double* __restrict__ gs,
                                                const double* __restrict__ bvs, double* __restrict__ X) {
  const int p = blockIdx.x, i = act[p];
  const size_t o = (size_t)i * d.nr, q = (size_t)p * d.nr;
  for (int j = threadIdx.x; j < d.nr; j += HW) {
    const double g = gs[o + j];
    const double cs = exp(fmin(-bvs[o + j] + g, d.cap)) - 1.0 - g;
    X[q + j] = (j + 1) * d.dr * cs;
  }
}

// ck = 2 pi dr Y / k - bvl_k; unstable if any n0 ck >= 1; X = k gk with
// gk = n0 ck^2 / (1 - n0 ck) - bvl_k
__global__ __launch_bounds__(HW) void k_oz(const int* __restrict__ act, Dims d, const double* __restrict__ Y,
                                           const double* __restrict__ bvlk, const double* __restrict__ n0,
                                           double* __restrict__ X, double* __restrict__ stats) {
  const int p = blockIdx.x, i = act[p];
  const size_t o = (size_t)i * d.nr, q = (size_t)p * d.nr;
  const double nn = n0[i];
  int bad = 0;
  for (int j = threadIdx.x; j < d.nr; j += HW) {
    const double k = (j + 1) * d.dk;
    const double ck = 2.0 * PI_ * d.dr * Y[q + j] / k - bvlk[o + j];
    if (nn * ck >= 1.0) bad = 1;
    const double gk = nn * ck * ck / (1.0 - nn * ck) - bvlk[o + j];
    X[q + j] = k * gk;
  }
  __shared__ int sb[HW];
  sb[threadIdx.x] = bad;
  __syncthreads();
  for (int s = HW / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sb[threadIdx.x] |= sb[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) stats[(size_t)i * 8 + 1] = sb[0] ? 1.0 : 0.0;
}

// gnew = dk Y / (4 pi^2 r); the history shift (hin[s0] = gs, hout[s0] = gnew)
// and, per solve: resid = max|gnew - gs| and the Ng sums with the two older
// history entries s1, s2: d0 = gnew - gs, d01 = d0 - (hout1 - hin1),
// d02 = d0 - (hout2 - hin2); a11, a12, a22, b1, b2.
__global__ __launch_bounds__(HW) void k_resid(const int* __restrict__ act, Dims d, const double* __restrict__ Y,
                                              const double* __restrict__ gs, double* __restrict__ gnew,
                                              double* __restrict__ hin, double* __restrict__ hout, size_t hstride,
                                              int s0, int s1, int s2, double* __restrict__ stats) {
  const int p = blockIdx.x, i = act[p];
  const size_t o = (size_t)i * d.nr, q = (size_t)p * d.nr;
  double mx = 0.0, a11 = 0.0, a12 = 0.0, a22 = 0.0, b1 = 0.0, b2 = 0.0;
  const double sc = d.dk / (4.0 * PI_ * PI_);
  for (int j = threadIdx.x; j < d.nr; j += HW) {
    const double gn = sc * Y[q + j] / ((j + 1) * d.dr);
    const double g = gs[o + j];
    gnew[o + j] = gn;
    const double d0 = gn - g;
    mx = fmax(mx, fabs(d0));
    const double d01 = d0 - (hout[s1 * hstride + o + j] - hin[s1 * hstride + o + j]);
    const double d02 = d0 - (hout[s2 * hstride + o + j] - hin[s2 * hstride + o + j]);
    a11 += d01 * d01; a12 += d01 * d02; a22 += d02 * d02;
    b1 += d0 * d01; b2 += d0 * d02;
    hin[s0 * hstride + o + j] = g;
    hout[s0 * hstride + o + j] = gn;
  }
  __shared__ double sm[6][HW];
  sm[0][threadIdx.x] = mx; sm[1][threadIdx.x] = a11; sm[2][threadIdx.x] = a12;
  sm[3][threadIdx.x] = a22; sm[4][threadIdx.x] = b1; sm[5][threadIdx.x] = b2;
  __syncthreads();
  for (int s = HW / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sm[0][threadIdx.x] = fmax(sm[0][threadIdx.x], sm[0][threadIdx.x + s]);
      for (int c = 1; c < 6; ++c) sm[c][threadIdx.x] += sm[c][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* st = stats + (size_t)i * 8;
    st[0] = sm[0][0];
    for (int c = 1; c < 6; ++c) st[1 + c] = sm[c][0];
  }
}

// per-solve actions
enum HncAct { A_NONE = 0, A_TAKE_NEW = 1, A_MIX = 2, A_RESET = 4, A_NG_TRY = 5 };

// Update of gs per solve.  A_NG_TRY: gtry = (1 - c1 - c2) hout0 + c1 hout1 +
// c2 hout2 is taken if max(gtry - bvs) < cap, else the mixing step (the Ng
// acceptance rule); coef[4 i ..] = mix or (c1, c2, mix).
__global__ __launch_bounds__(HW) void k_apply(const int* __restrict__ all, Dims d, const int* __restrict__ action,
                                              const double* __restrict__ coef, double* __restrict__ gs,
                                              const double* __restrict__ gnew, const double* __restrict__ hout,
                                              size_t hstride, int s0, int s1, int s2, const double* __restrict__ gs0,
                                              const double* __restrict__ bvs) {
  const int i = all[blockIdx.x];
  int a = action[i];
  if (a == A_NONE) return;
  const size_t o = (size_t)i * d.nr;
  const double c1 = coef[4 * i], c2 = coef[4 * i + 1];
  double mix = c1;
  if (a == A_NG_TRY) {
    double mx = -1.0e300;
    for (int j = threadIdx.x; j < d.nr; j += HW) {
      const double gt = (1.0 - c1 - c2) * hout[s0 * hstride + o + j] + c1 * hout[s1 * hstride + o + j] +
                        c2 * hout[s2 * hstride + o + j];
      mx = fmax(mx, gt - bvs[o + j]);
    }
    __shared__ double sm[HW];
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int s = HW / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
      __syncthreads();
    }
    if (!(sm[0] < d.cap)) {
      a = A_MIX;
      mix = coef[4 * i + 2];
    }
  }
  for (int j = threadIdx.x; j < d.nr; j += HW) {
    switch (a) {
      case A_TAKE_NEW: gs[o + j] = gnew[o + j]; break;
      case A_MIX: gs[o + j] = (1.0 - mix) * gs[o + j] + mix * gnew[o + j]; break;
      case A_RESET: gs[o + j] = gs0[o + j]; break;
      case A_NG_TRY:
        gs[o + j] = (1.0 - c1 - c2) * hout[s0 * hstride + o + j] + c1 * hout[s1 * hstride + o + j] +
                    c2 * hout[s2 * hstride + o + j];
        break;
      default: break;
    }
  }
}

// DST-I table and (when needed) matrix of the last nr, kept for the life of
// the process
struct DstCache {
  std::mutex mu;
  int nr = 0;
  double* S = nullptr;
  double* T = nullptr;  // its 2(nr+1) distinct entries
};
DstCache& dst_cache() {
  static DstCache c;
  return c;
}

int halloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return set_error(HY_ERR_OOM, "hnc: device allocation failed");
  }
  return HY_OK;
}

struct Solve {
  int state = 0;  // 0 active, 1 converged, 2 stopped at maxiter, 3 unstable (all restarts)
  int attempt = 0, it = 0, nhist = 0, cooldown = 0, total_it = 0;
  double mix = 0.0, resid_prev = 1.0e100, resid = 0.0;
};

}  // namespace
}  // namespace hy

using namespace hy;

int hy_hnc_iterate(int nsolve, int nr, double dr, double dk, const double* h_n0, const double* h_bvs,
                   const double* h_bvlk, const double* h_gs0, double tol, double mix, int maxiter, double* h_gs,
                   int* h_status, int* h_niter, double* h_resid) {
  if (nsolve < 1 || nr < 2 || !h_n0 || !h_bvs || !h_bvlk || !h_gs0 || !h_gs || !h_status || !h_niter || !h_resid ||
      !(dr > 0.0) || !(dk > 0.0) || maxiter < 1 || !(mix > 0.0))
    return set_error(HY_ERR_INVALID, "hy_hnc_iterate: bad argument");
  if ((size_t)nr * nr > (size_t)1 << 31) return set_error(HY_ERR_INVALID, "hy_hnc_iterate: nr too large");
  // one call at a time: a call with another nr replaces the cached DST-I table
  // and matrix, which a concurrent call may still be using
  static std::mutex call_mutex;
  std::lock_guard<std::mutex> call_lock(call_mutex);
  const int expcap_i = 40, npicard = 5, nattempt = 4;
  const Dims d{nr, dr, dk, (double)expcap_i};
  const size_t vec = (size_t)nsolve * nr;
  // device buffers: gs, gnew, gs0, bvs, bvlk, X, Y, hin[3], hout[3]
  std::vector<void*> allocs;
  auto dm = [&](void** p, size_t bytes) -> int {
    const int rc = halloc(p, bytes);
    if (rc == HY_OK) allocs.push_back(*p);
    return rc;
  };
  auto free_all = [&]() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
  };
  double *gs, *gnew, *gs0, *bvs, *bvlk, *X, *Y, *hin, *hout, *n0, *stats, *coef, *part;
  int *act, *all, *action;
  int rc = HY_OK;
  if ((rc = dm((void**)&gs, vec * 8)) || (rc = dm((void**)&gnew, vec * 8)) || (rc = dm((void**)&gs0, vec * 8)) ||
      (rc = dm((void**)&bvs, vec * 8)) || (rc = dm((void**)&bvlk, vec * 8)) || (rc = dm((void**)&X, vec * 8)) ||
      (rc = dm((void**)&Y, vec * 8)) || (rc = dm((void**)&hin, 3 * vec * 8)) || (rc = dm((void**)&hout, 3 * vec * 8)) ||
      (rc = dm((void**)&n0, nsolve * 8)) || (rc = dm((void**)&stats, (size_t)nsolve * 8 * 8)) ||
      (rc = dm((void**)&part, (size_t)DCH * NA_MAX * nr * 8)) ||
      (rc = dm((void**)&coef, (size_t)nsolve * 4 * 8)) || (rc = dm((void**)&act, nsolve * 4)) ||
      (rc = dm((void**)&all, nsolve * 4)) || (rc = dm((void**)&action, nsolve * 4))) {
    free_all();
    return rc;
  }
  hipStream_t st = nullptr;
  rocblas_handle blas = nullptr;
  auto fail = [&](int code, const std::string& msg) {
    if (st) (void)hipStreamSynchronize(st);
    if (blas) rocblas_destroy_handle(blas);
    if (st) (void)hipStreamDestroy(st);
    free_all();
    return set_error(code, msg);
  };
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return fail(HY_ERR_HIP, "hnc: stream");
  if (rocblas_create_handle(&blas) != rocblas_status_success || rocblas_set_stream(blas, st) != rocblas_status_success)
    return fail(HY_ERR_HIP, "hnc: rocblas handle");
  // the DST-I matrix (and its table) for this nr
  double* S = nullptr;
  double* T = nullptr;
  {
    DstCache& c = dst_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.nr != nr) {
      if (c.S || c.T) {
        (void)hipDeviceSynchronize();
        if (c.S) (void)hipFree(c.S);
        if (c.T) (void)hipFree(c.T);
        c.S = c.T = nullptr;
      }
      c.nr = nr;
    }
    // the table serves every pass when it fits in LDS; the 8 nr^2-byte matrix
    // is built only for the GEMM path (more than DST_STREAM_MAX solves) or
    // when the table does not fit
    const bool tab_fits = 2 * (nr + 1) <= TAB_MAX;
    const bool need_s = !tab_fits || nsolve > DST_STREAM_MAX;
    bool built = false;
    if (tab_fits && !c.T) {
      if (halloc((void**)&c.T, (size_t)2 * (nr + 1) * 8) != HY_OK) return fail(HY_ERR_OOM, "hnc: DST table");
      hipLaunchKernelGGL(k_dst_table, dim3(64), dim3(256), 0, st, c.T, nr);
      built = true;
    }
    if (need_s && !c.S) {
      if (halloc((void**)&c.S, (size_t)nr * nr * 8) != HY_OK) return fail(HY_ERR_OOM, "hnc: DST matrix");
      hipLaunchKernelGGL(k_dst_matrix, dim3(4096), dim3(256), 0, st, c.S, nr);
      built = true;
    }
    if (built && hipStreamSynchronize(st) != hipSuccess) return fail(HY_ERR_HIP, "hnc: DST matrix kernel");
    S = c.S;
    T = c.T;
  }
  if (hipMemcpy(bvs, h_bvs, vec * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(bvlk, h_bvlk, vec * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(gs0, h_gs0, vec * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(gs, h_gs0, vec * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n0, h_n0, nsolve * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(hin, 0, 3 * vec * 8) != hipSuccess || hipMemset(hout, 0, 3 * vec * 8) != hipSuccess)
    return fail(HY_ERR_HIP, "hnc: upload");

  std::vector<Solve> sv(nsolve);
  for (auto& s : sv) s.mix = mix;
  std::vector<int> h_act, h_all(nsolve), h_action(nsolve);
  std::vector<double> h_stats((size_t)nsolve * 8), h_coef((size_t)nsolve * 4, 0.0);
  for (int i = 0; i < nsolve; ++i) h_all[i] = i;
  if (hipMemcpy(all, h_all.data(), nsolve * 4, hipMemcpyHostToDevice) != hipSuccess) return fail(HY_ERR_HIP, "hnc");
  const double one = 1.0, zero = 0.0;
  // Y = S X over the na active columns: streaming passes over S of up to
  // NA_MAX columns each for up to DST_STREAM_MAX columns, matrix cores beyond
  auto dst = [&](int na) -> bool {
    if (na <= DST_STREAM_MAX) {
      // matrix-free from the LDS table when it fits, else streaming S
      const bool tab = 2 * (nr + 1) <= TAB_MAX;
      const int tiles = tab ? (nr + TTILE - 1) / TTILE : (nr + DTILE - 1) / DTILE;
      const int nch = tab ? std::max(1, std::min(DCH, 256 / tiles)) : DCH;  // about one workgroup per CU
      const dim3 g(tiles, nch);
      for (int p0 = 0; p0 < na; p0 += NA_MAX) {
        const int nap = std::min(NA_MAX, na - p0);
        const double* Xp = X + (size_t)p0 * nr;
        double* Yp = Y + (size_t)p0 * nr;
        if (tab) {
          if (nap <= 1) hipLaunchKernelGGL(k_dst_tab<1>, g, dim3(TTILE), 0, st, T, Xp, part, nr, nap, nch);
          else if (nap <= 2) hipLaunchKernelGGL(k_dst_tab<2>, g, dim3(TTILE), 0, st, T, Xp, part, nr, nap, nch);
          else if (nap <= 4) hipLaunchKernelGGL(k_dst_tab<4>, g, dim3(TTILE), 0, st, T, Xp, part, nr, nap, nch);
          else if (nap <= 8) hipLaunchKernelGGL(k_dst_tab<8>, g, dim3(TTILE), 0, st, T, Xp, part, nr, nap, nch);
          else if (nap <= 16) hipLaunchKernelGGL(k_dst_tab<16>, g, dim3(TTILE), 0, st, T, Xp, part, nr, nap, nch);
          else hipLaunchKernelGGL(k_dst_tab<32>, g, dim3(TTILE / tab_rows<32>()), 0, st, T, Xp, part, nr, nap, nch);
        } else {
          if (nap <= 1) hipLaunchKernelGGL(k_dst_few<1>, g, dim3(DTILE), 0, st, S, Xp, part, nr, nap);
          else if (nap <= 2) hipLaunchKernelGGL(k_dst_few<2>, g, dim3(DTILE), 0, st, S, Xp, part, nr, nap);
          else if (nap <= 4) hipLaunchKernelGGL(k_dst_few<4>, g, dim3(DTILE), 0, st, S, Xp, part, nr, nap);
          else if (nap <= 8) hipLaunchKernelGGL(k_dst_few<8>, g, dim3(DTILE), 0, st, S, Xp, part, nr, nap);
          else if (nap <= 16) hipLaunchKernelGGL(k_dst_few<16>, g, dim3(DTILE), 0, st, S, Xp, part, nr, nap);
          else hipLaunchKernelGGL(k_dst_few<32>, g, dim3(DTILE), 0, st, S, Xp, part, nr, nap);
        }
        hipLaunchKernelGGL(k_dst_sum, dim3(std::min<size_t>(((size_t)nap * nr + 255) / 256, 4096)), dim3(256), 0,
                           st, part, Yp, nr, nap, nch);
      }
      return hipGetLastError() == hipSuccess;
    }
    return rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, nr, na, nr, &one, S, nr, X, nr, &zero,
                         Y, nr) == rocblas_status_success;
  };
  long long step = 0;  // global iteration counter: history ring position
  for (;;) {
    h_act.clear();
    for (int i = 0; i < nsolve; ++i)
      if (sv[i].state == 0) h_act.push_back(i);
    const int na = (int)h_act.size();
    if (na == 0) break;
    const int s0 = (int)(step % 3), s1 = (int)((step + 2) % 3), s2 = (int)((step + 1) % 3);
    ++step;
    if (hipMemcpyAsync(act, h_act.data(), na * 4, hipMemcpyHostToDevice, st) != hipSuccess) return fail(HY_ERR_HIP, "hnc");
    hipLaunchKernelGGL(k_closure, dim3(na), dim3(HW), 0, st, act, d, gs, bvs, X);
    if (!dst(na)) return fail(HY_ERR_HIP, "hnc: DST product");
    hipLaunchKernelGGL(k_oz, dim3(na), dim3(HW), 0, st, act, d, Y, bvlk, n0, X, stats);
    if (!dst(na)) return fail(HY_ERR_HIP, "hnc: DST product");
    hipLaunchKernelGGL(k_resid, dim3(na), dim3(HW), 0, st, act, d, Y, gs, gnew, hin, hout, vec, s0, s1, s2, stats);
    if (hipMemcpyAsync(h_stats.data(), stats, (size_t)nsolve * 8 * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(HY_ERR_HIP, "hnc: iteration");
    // decisions (oz_hnc.f90 direct solve loop; the Python host restatement
    // hnc._solve_direct follows the same rules)
    for (int i = 0; i < nsolve; ++i) h_action[i] = A_NONE;
    for (int i : h_act) {
      Solve& s = sv[i];
      const double* stt = &h_stats[(size_t)i * 8];
      s.it += 1;
      s.total_it = s.it;
      bool unstable = stt[1] != 0.0;
      const double resid = stt[0];
      if (!unstable) {
        s.resid = resid;
        if (resid < tol) {
          h_action[i] = A_TAKE_NEW;
          s.state = 1;
          continue;
        }
        if (resid > 1.0e6) unstable = true;
      }
      if (unstable) {
        if (s.attempt + 1 < nattempt) {  // restart with halved mixing from the initial guess
          s.attempt += 1;
          s.mix *= 0.5;
          s.it = 0;
          s.nhist = 0;
          s.cooldown = 0;
          s.resid_prev = 1.0e100;
          h_action[i] = A_RESET;
        } else {
          s.state = 3;
        }
        continue;
      }
      s.nhist = std::min(s.nhist + 1, 3);
      if (resid > 10.0 * s.resid_prev) s.cooldown = 10;
      h_coef[4 * i] = s.mix;
      h_coef[4 * i + 1] = 0.0;
      h_action[i] = A_MIX;
      if (s.nhist == 3 && s.it > npicard && s.cooldown == 0) {
        const double a11 = stt[2], a12 = stt[3], a22 = stt[4], b1 = stt[5], b2 = stt[6];
        const double det = a11 * a22 - a12 * a12;
        if (std::fabs(det) > 1.0e-14 * a11 * a22 && a11 > 0.0 && a22 > 0.0) {
          const double c1 = (b1 * a22 - b2 * a12) / det;
          const double c2 = (a11 * b2 - a12 * b1) / det;
          if (std::fabs(c1) < 20.0 && std::fabs(c2) < 20.0) {
            h_coef[4 * i] = c1;
            h_coef[4 * i + 1] = c2;
            h_coef[4 * i + 2] = s.mix;
            h_action[i] = A_NG_TRY;
          }
        }
      } else if (s.cooldown > 0) {
        s.cooldown -= 1;
      }
      s.resid_prev = resid;
      if (s.it >= maxiter) s.state = 2;
    }
    if (hipMemcpyAsync(coef, h_coef.data(), (size_t)nsolve * 4 * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(action, h_action.data(), nsolve * 4, hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(HY_ERR_HIP, "hnc");
    hipLaunchKernelGGL(k_apply, dim3(nsolve), dim3(HW), 0, st, all, d, action, coef, gs, gnew, hout, vec, s0, s1, s2,
                       gs0, bvs);
  }
  if (hipMemcpyAsync(h_gs, gs, vec * 8, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return fail(HY_ERR_HIP, "hnc: download");
  for (int i = 0; i < nsolve; ++i) {
    h_status[i] = sv[i].state == 1 ? 1 : (sv[i].state == 2 ? 0 : -1);
    h_niter[i] = sv[i].total_it;
    h_resid[i] = sv[i].resid;
  }
  rocblas_destroy_handle(blas);
  (void)hipStreamDestroy(st);
  free_all();
  return HY_OK;
}
