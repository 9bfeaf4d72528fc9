// Internal declarations of libhydft: error state, complex arithmetic, the
// grid (wavenumbers, 2/3-rule kept box, twiddle tables) and the LDS
// Stockham FFT building blocks shared by every kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <array>
#include <vector>

#include "../../include/hydft.h"

namespace hy {

// ------------------------------------------------------------------ errors
int set_error(int code, const std::string& msg);
// HYDFT_TRACE=1: synchronise after every launch and print its name to stderr
// (debugging aid; off by default).
int trace_launch(const char* what, hipStream_t s);
#define HY_TRACE(what, s) HY_TRY(::hy::trace_launch(what, s))
#define HY_HIP(call)                                                               \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return ::hy::set_error(HY_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define HY_TRY(call)                \
  do {                              \
    int r_ = (call);                \
    if (r_ != HY_OK) return r_;     \
  } while (0)

// ------------------------------------------------------------ launches
// Every kernel launch of the model's step path goes through hy::launch.  A
// kernel K with argument struct A comes in two forms: K(A) (one model) and
// K_b(const A*) (a batch: model blockIdx.z takes args[blockIdx.z]).  While a
// recorder is active on the calling host thread, launch() appends the batched
// form, the launch shape and the arguments to it instead of launching; a batch
// of models with identical step sequences then replays each recorded launch
// once for all of them.
struct LaunchRec {
  const void* fn_batched;
  dim3 grid, block;
  std::vector<unsigned char> args;
  // x-line kernels: (argument block pointer, twiddle pointer) parameters; the
  // batch gathers the blocks themselves into one device array
  bool xline = false;
  const void* xblock = nullptr;
  const void* xtw = nullptr;
  // pencil passes: {1, f64, L, inverse, A, B, nf, columns per workgroup}
  // (the batch retunes the plan for its size); pencil[0] == 0 otherwise
  std::array<int, 8> pencil{};
};
struct Recorder {
  std::vector<LaunchRec> items;
};
Recorder*& recorder();  // the calling thread's active recorder (null: launch)

template <class A>
inline int launch(const void* fn_single, const void* fn_batched, dim3 grid, dim3 block, hipStream_t s,
                  const A& a) {
  if (Recorder* r = recorder()) {
    LaunchRec it{fn_batched, grid, block, {}};
    it.args.resize(sizeof(A));
    std::memcpy(it.args.data(), &a, sizeof(A));
    r->items.push_back(std::move(it));
    return HY_OK;
  }
  void* args[] = {const_cast<A*>(&a)};
  const hipError_t e = hipLaunchKernel(fn_single, grid, block, args, 0, s);
  if (e != hipSuccess) return set_error(HY_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return HY_OK;
}

// x-line kernels: K(const Block* pa, const Tw* tw) / K_b(const Block* blocks, const Tw* tw)
inline int launch_x(const void* fn_single, const void* fn_batched, dim3 grid, hipStream_t s, const void* pa,
                    const void* tw) {
  if (Recorder* r = recorder()) {
    LaunchRec it{fn_batched, grid, dim3(64), {}};
    it.xline = true;
    it.xblock = pa;
    it.xtw = tw;
    r->items.push_back(std::move(it));
    return HY_OK;
  }
  void* args[] = {const_cast<void**>(&pa), const_cast<void**>(&tw)};
  const hipError_t e = hipLaunchKernel(fn_single, grid, dim3(64), args, 0, s);
  if (e != hipSuccess) return set_error(HY_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return HY_OK;
}

// ---------------------------------------------------------------- complex
template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};
using c64 = cpx<double>;
using c32 = cpx<float>;

template <class T>
__host__ __device__ inline cpx<T> mk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <class T>
__host__ __device__ inline cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <class T>
__host__ __device__ inline cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <class T>
__host__ __device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T>
__host__ __device__ inline cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <class T>
__host__ __device__ inline cpx<T> conjc(cpx<T> a) { return mk<T>(a.x, -a.y); }
template <class T>
__host__ __device__ inline cpx<T> scal(cpx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <class T>
__host__ __device__ inline cpx<T> mul_i(cpx<T> a) { return mk<T>(-a.y, a.x); }  // i*a
template <class TO, class TI>
__host__ __device__ inline cpx<TO> cvt(cpx<TI> a) { return mk<TO>((TO)a.x, (TO)a.y); }

// Element `elem` of a device buffer of `nbytes` bytes (below 4 GiB) holding
// cpx<E>, or zero when !ok, without a branch: the offset of a skipped
// element is the buffer's end, and a raw buffer load past the end returns
// zero without touching memory.  A group of such loads issues back to back;
// a branch around each load would make the compiler wait for every load
// before issuing the next.
template <class E>
__device__ inline cpx<E> bload_c(const void* base, unsigned nbytes, unsigned elem, bool ok) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes, 0x00020000);
  const unsigned off = ok ? elem * (unsigned)sizeof(cpx<E>) : nbytes;
  if constexpr (sizeof(E) == 4) return __builtin_bit_cast(cpx<E>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else return __builtin_bit_cast(cpx<E>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
#else
  return ok ? reinterpret_cast<const cpx<E>*>(base)[elem] : mk<E>(0, 0);
#endif
}

// ------------------------------------------------------ XCD-aware workgroup order
// Workgroups of a launch are dealt round-robin over the chip's 8 XCDs
// (workgroup i runs on XCD i % 8), each with its own L2.  Kernels whose
// neighbouring work items share cache lines (column tiles of one row,
// adjacent x-lines) take their work index from this map, so that
// consecutive work indices run on one XCD at about the same time.
constexpr unsigned NXCD = 8;
__device__ inline unsigned xcd_work_index(unsigned i, unsigned n) {
  const unsigned q = n / NXCD;
  return i < q * NXCD ? (i % NXCD) * q + i / NXCD : i;
}

// ------------------------------------------------------ radix plans (2^a 3^b)
// Stage radices: as many 4s as possible, then one 2, then 3s.
__host__ __device__ constexpr int n4_of(int L) { int n = 0; while (L % 4 == 0) { L /= 4; ++n; } return n; }
__host__ __device__ constexpr int rem4(int L) { while (L % 4 == 0) L /= 4; return L; }
__host__ __device__ constexpr int n2_of(int L) { return rem4(L) % 2 == 0 ? 1 : 0; }
__host__ __device__ constexpr int n3_of(int L) {
  int t = rem4(L); if (t % 2 == 0) t /= 2; int n = 0; while (t % 3 == 0) { t /= 3; ++n; } return n;
}
__host__ __device__ constexpr int nstages_of(int L) { return n4_of(L) + n2_of(L) + n3_of(L); }
__host__ __device__ constexpr int radix_of(int L, int s) {
  return s < n4_of(L) ? 4 : (s < n4_of(L) + n2_of(L) ? 2 : 3);
}
__host__ __device__ constexpr int ns_before(int L, int s) {
  int p = 1; for (int i = 0; i < s; ++i) p *= radix_of(L, i); return p;
}
__host__ __device__ constexpr bool supported_len(int L) {
  return L == 4 || L == 8 || L == 12 || L == 16 || L == 24 || L == 32 || L == 48 || L == 64 ||
         L == 96 || L == 128 || L == 256;
}

// ----------------------------------------------------------- DFT codelets
// Forward transforms use exp(-2 pi i jk/R); INV flips the sign.
template <class T, bool INV>
__device__ inline void dft2(cpx<T>* v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = a + b; v[1] = a - b;
}
template <class T, bool INV>
__device__ inline void dft3(cpx<T>* v) {
  const T c = T(-0.5);
  const T s = INV ? T(0.86602540378443864676) : T(-0.86602540378443864676);
  cpx<T> a = v[0], t1 = v[1] + v[2], t2 = v[1] - v[2];
  cpx<T> m = mk<T>(a.x + c * t1.x, a.y + c * t1.y);
  cpx<T> is = mk<T>(-s * t2.y, s * t2.x);  // i*s*t2
  v[0] = a + t1;
  v[1] = m + is;
  v[2] = m - is;
}
template <class T, bool INV>
__device__ inline void dft4(cpx<T>* v) {
  cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], t3 = v[1] - v[3];
  cpx<T> it3 = mul_i(t3);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  if (INV) { v[1] = t1 + it3; v[3] = t1 - it3; }
  else     { v[1] = t1 - it3; v[3] = t1 + it3; }
}
template <class T, bool INV, int R>
__device__ inline void dftR(cpx<T>* v) {
  if constexpr (R == 2) dft2<T, INV>(v);
  else if constexpr (R == 3) dft3<T, INV>(v);
  else dft4<T, INV>(v);
}

// One Stockham autosort stage (decimation in time) over NA pencils of length L
// stored in LDS as buf[p*NA + a].  In place: every thread loads its
// butterflies, the block synchronises, then stores.  tw[j] = exp(-2 pi i j/L).
template <class T, int L, int NA, int NT, bool INV, int S>
__device__ inline void fft_stage(cpx<T>* buf, const cpx<T>* tw, int tid) {
  constexpr int R = radix_of(L, S);
  constexpr int NS = ns_before(L, S);
  constexpr int NB = L / R;
  constexpr int TOT = NB * NA;
  constexpr int PER = (TOT + NT - 1) / NT;
  cpx<T> v[PER][R];
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int idx = tid + b * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int a = idx % NA, j = idx / NA;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = buf[(j + r * NB) * NA + a];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int idx = tid + b * NT;
    if (TOT % NT == 0 || idx < TOT) {
      const int a = idx % NA, j = idx / NA;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[k * r * (L / (NS * R))];
          v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
        }
      }
      dftR<T, INV, R>(v[b]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(base + r * NS) * NA + a] = v[b][r];
    }
  }
  __syncthreads();
}

template <class T, int L, int NA, int NT, bool INV, int S = 0>
__device__ inline void fft_lds(cpx<T>* buf, const cpx<T>* tw, int tid) {
  if constexpr (S < nstages_of(L)) {
    fft_stage<T, L, NA, NT, INV, S>(buf, tw, tid);
    fft_lds<T, L, NA, NT, INV, S + 1>(buf, tw, tid);
  }
}

// ------------------------------------------------------------------- grid
// Wavenumbers, the kept box of the 2/3 rule and twiddle tables on the device.
struct TwiddleSet {
  int L = 0;
  c64* d64 = nullptr;
  c32* d32 = nullptr;
};

struct Grid {
  int nx = 0, ny = 1, nz = 1, nkx = 0, ndim = 1;
  double lx = 0, ly = 0, lz = 0, dx = 0, dy = 0, dz = 0, volume = 0, kmax_dealiased = 0;
  int dealias = 1, precision = HY_PREC_FP64;
  // kept box: kx index in [0,Kx); signed ky in [-My,My] (Ky = 2My+1 or 1); same for z
  int Kx = 0, Ky = 1, Kz = 1, My = 0, Mz = 0;
  std::vector<double> kx, ky, kz;  // full wavenumber arrays (nkx, ny, nz)
  double* d_kx = nullptr;          // device copies
  double* d_ky = nullptr;
  double* d_kz = nullptr;
  std::vector<TwiddleSet> tws;
  const TwiddleSet& tw(int L) const;
  // scratch owned by the grid for the drop-in spectral operators
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  size_t nreal() const { return (size_t)nx * ny * nz; }
  size_t nfull() const { return (size_t)nkx * ny * nz; }
  size_t ntrunc() const { return (size_t)Kx * Ky * Kz; }
};

int grid_build(Grid& g, int nx, int ny, int nz, double lx, double ly, double lz, int dealias,
               int precision);
void grid_free(Grid& g);

// Descriptor of one pencil field for the batched y/z passes.
enum MulKind { MUL_NONE = 0, MUL_IK = 1, MUL_NEGK2 = 2 };
struct PencilField {
  const void* in;
  void* out;
  int mul;          // MulKind, applied on the spectral side
  int in_f64;       // input element type: 1 = c64, 0 = c32 (outputs are the pass's compute type)
  double k0_offset; // added to the k = 0 entry on load (inverse first pass only)
};
constexpr int MAXPF = 24;

// Launch helpers (hy_fft.hip).  `axis` is 1 (y) or 2 (z).
//   inverse: in (A, Kp, B) truncated-or-full spectrum -> out (A, L, B)
//   forward: in (A, L, B) -> out (A, Kp, B), scaled by 1/L
// A is the contiguous count (elements), B the outer count.
// `halt`: a device flag that makes every workgroup of the pass return at once
// when set (a run that reached its step limit in a batch); null: never.
int pencil_pass(const Grid& g, int axis, bool inverse, bool trunc, int A, int B,
                const PencilField* f, int nf, hipStream_t s, const int* halt = nullptr);

// The plan of a recorded pencil pass (LaunchRec::pencil) in a batch of nz
// runs whose packed arguments are at d_args: the batched kernel, grid and block
// to launch (measured once per shape and batch size, then kept).
int pencil_batch_plan(const std::array<int, 8>& pencil, int nz, void* d_args, hipStream_t s, const void** fn,
                      dim3* grid, dim3* block);

// A device gate flag, read through the global address space.
__device__ inline bool gate_closed(const int* h) {
#if defined(__HIP_DEVICE_COMPILE__)
  return h && *((const __attribute__((address_space(1))) int*)h) != 0;
#else
  return h && *h != 0;
#endif
}

}  // namespace hy

// The grid handle of the C ABI (hy_grid): a grid and its transform plans.
struct hy_grid_s {
  hy::Grid g;
};
