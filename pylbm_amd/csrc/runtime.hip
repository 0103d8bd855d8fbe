// Core runtime of the MI355X lattice Boltzmann time-stepper (libpylbm_hip.so):
// device memory, copies, streams, events, and the scheme-independent kernels
// (periodic ghost fill, boundary link lists, FP64 compensated population sums).
#include <stdarg.h>

#include <mutex>

#include "lbm_device.h"

namespace lbm {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ------------------------------------------------------------- ghost fill
// One thread per ghost cell of the halo shell, partitioned into
//   A: outer-axis ghost slabs (full mid/inner extent)
//   B: mid-axis ghost slabs (interior outer, full inner)
//   C: inner-axis ghost slabs (interior outer and mid)
// For every population that enters the interior through all the ghost axes of
// the cell (component >= depth on the low side, <= -depth on the high side) the
// value at the fully wrapped interior position is copied.  all_pops copies every
// population instead (the full axis-by-axis wrap).
struct ShellDesc {
  int64_t nA, nB, nC;
  int sidesA, sidesB, sidesC;  // bit 0: low face slab, bit 1: high face slab
};

// coordinate of ghost layer g (0 <= g < h * number of sides) along an axis
__device__ __forceinline__ int slab_coord(int g, int n, int h, int sides) {
  if (sides == 3) return g < h ? g : n - 2 * h + g;
  if (sides == 1) return g;
  return n - h + g;
}

__host__ __device__ __forceinline__ int nsides(int sides) { return (sides & 1) + ((sides >> 1) & 1); }

__device__ __forceinline__ bool ghost_axis(int64_t c, int64_t n, int h, int64_t* wrapped,
                                           int* depth) {
  const int64_t nin = n - 2 * h;
  if (c < h) {
    *wrapped = c + nin;
    *depth = (int)(h - c);  // > 0: low side
    return true;
  }
  if (c >= n - h) {
    *wrapped = c - nin;
    *depth = -(int)(c - (n - h) + 1);  // < 0: high side
    return true;
  }
  *wrapped = c;
  *depth = 0;
  return false;
}

// src and dst are the same array: every read is an interior slot and every write a
// ghost slot, so they never touch the same element and may be declared restrict,
// which lets all loads of a cell be issued before its stores.
template <typename T, int QMAX>
__global__ void __launch_bounds__(256) ghost_fill_kernel(const T* __restrict__ src_f,
                                                         T* __restrict__ dst_f, lbm_layout L,
                                                         lbm_velocities V, ShellDesc S,
                                                         int all_pops) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= S.nA + S.nB + S.nC) return;
  const int64_t n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
  const int h0 = L.halo[0], h1 = L.halo[1], h2 = L.halo[2];
  int64_t o, m, i;
  if (t < S.nA) {  // outer ghost slabs: layers x n1 x n2
    int64_t r = t;
    i = r % n2;
    r /= n2;
    m = r % n1;
    r /= n1;
    o = slab_coord((int)r, (int)n0, h0, S.sidesA);
  } else if (t < S.nA + S.nB) {  // mid ghost slabs: (n0-2h0) x layers x n2
    const int nl = h1 * nsides(S.sidesB);
    int64_t r = t - S.nA;
    i = r % n2;
    r /= n2;
    m = slab_coord((int)(r % nl), (int)n1, h1, S.sidesB);
    o = h0 + r / nl;
  } else {  // inner ghost slabs: (n0-2h0) x (n1-2h1) x layers
    const int nl = h2 * nsides(S.sidesC);
    int64_t r = t - (S.nA + S.nB);
    i = slab_coord((int)(r % nl), (int)n2, h2, S.sidesC);
    r /= nl;
    m = h1 + r % (n1 - 2 * h1);
    o = h0 + r / (n1 - 2 * h1);
  }
  int64_t wo, wm, wi;
  int d0, d1, d2;
  ghost_axis(o, n0, h0, &wo, &d0);
  ghost_axis(m, n1, h1, &wm, &d1);
  ghost_axis(i, n2, h2, &wi, &d2);
  const int64_t dst = cell_offset(L, o, m, i);
  const int64_t src = cell_offset(L, wo, wm, wi);
  const int64_t ps = L.pop_stride;
  T val[QMAX];
  bool take[QMAX];
#pragma unroll
  for (int k = 0; k < QMAX; ++k) {
    take[k] = k < L.ns;
    if (!all_pops) {
      const int c0 = V.v[k][0], c1 = V.v[k][1], c2 = V.v[k][2];
      take[k] = take[k] && (d0 == 0 || (d0 > 0 ? c0 >= d0 : c0 <= d0)) &&
                (d1 == 0 || (d1 > 0 ? c1 >= d1 : c1 <= d1)) &&
                (d2 == 0 || (d2 > 0 ? c2 >= d2 : c2 <= d2));
    }
    if (take[k]) val[k] = src_f[k * ps + src];
  }
#pragma unroll
  for (int k = 0; k < QMAX; ++k)
    if (take[k]) dst_f[k * ps + dst] = val[k];
}

template <typename T>
static void ghost_launch(void* f, const lbm_layout& L, const lbm_velocities& V, const ShellDesc& S,
                         int all_pops, unsigned blocks, hipStream_t s) {
  const T* src = (const T*)f;
  T* dst = (T*)f;
  if (L.ns <= 4) ghost_fill_kernel<T, 4><<<blocks, 256, 0, s>>>(src, dst, L, V, S, all_pops);
  else if (L.ns <= 9) ghost_fill_kernel<T, 9><<<blocks, 256, 0, s>>>(src, dst, L, V, S, all_pops);
  else if (L.ns <= 19) ghost_fill_kernel<T, 19><<<blocks, 256, 0, s>>>(src, dst, L, V, S, all_pops);
  else ghost_fill_kernel<T, LBM_MAX_Q><<<blocks, 256, 0, s>>>(src, dst, L, V, S, all_pops);
}

// ------------------------------------------------------------ link lists
template <typename S>
__global__ void __launch_bounds__(256) links_kernel(S* __restrict__ f,
                                                    const uint32_t* __restrict__ store,
                                                    const uint32_t* __restrict__ load,
                                                    const typename Codec<S>::C* __restrict__ rhs,
                                                    const int8_t* __restrict__ sgn, float sign,
                                                    int64_t n) {
  using C = typename Codec<S>::C;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  // deviation formats: rhs already carries the rest-population shift, so the
  // stored deviations are combined directly; load_dev / store with w = 0 keep
  // every bit (a sign-1 link without rhs copies the stored value, -0 included)
  C v = Codec<S>::load_dev(f, load[t], (C)0);
  const C a = sgn ? (C)sgn[t] : (C)sign;
  C r = a * v;
  if (rhs) r += rhs[t];
  Codec<S>::store(f, store[t], r, (C)0);
}

template <typename S>
__global__ void __launch_bounds__(256) links_gather_kernel(const S* __restrict__ f,
                                                           const uint32_t* __restrict__ load,
                                                           const typename Codec<S>::C* __restrict__ rhs,
                                                           const int8_t* __restrict__ sgn, float sign,
                                                           int64_t n,
                                                           typename Codec<S>::C* __restrict__ work) {
  using C = typename Codec<S>::C;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const C a = sgn ? (C)sgn[t] : (C)sign;
  C r = a * Codec<S>::load_dev(f, load[t], (C)0);
  if (rhs) r += rhs[t];
  work[t] = r;
}

template <typename S>
__global__ void __launch_bounds__(256) links_scatter_kernel(S* __restrict__ f,
                                                            const uint32_t* __restrict__ store,
                                                            int64_t n,
                                                            const typename Codec<S>::C* __restrict__ work) {
  using C = typename Codec<S>::C;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  Codec<S>::store(f, store[t], work[t], (C)0);
}

template <typename S>
static int links_launch(void* d_f, const uint32_t* st, const uint32_t* ld, const void* rhs,
                        const int8_t* sgn, double sign, int64_t n, int two_phase, void* work,
                        hipStream_t s) {
  using C = typename Codec<S>::C;
  const int64_t blocks = (n + 255) / 256;
  if (!two_phase) {
    links_kernel<S><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        (S*)d_f, st, ld, (const C*)rhs, sgn, (float)sign, n);
  } else {
    LBM_REQUIRE(work != nullptr, "two-phase link list needs a workspace");
    links_gather_kernel<S><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
        (const S*)d_f, ld, (const C*)rhs, sgn, (float)sign, n, (C*)work);
    links_scatter_kernel<S><<<dim3((unsigned)blocks), dim3(256), 0, s>>>((S*)d_f, st, n,
                                                                        (const C*)work);
  }
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

// Interpolated (Bouzidi) links, pylbm/boundary.py:608-623 and 745-760:
// f[store] = ca f[load0] + cb f[load1] + rhs with one coefficient pair per link
// (ca = s or -s, cb = 1 - s; the host picks the loads and s from the wall
// distance).  work == nullptr: the value is stored at once (the list loads no slot
// it stores); otherwise it is gathered into work and links_scatter_kernel stores it.
template <typename S>
__global__ void __launch_bounds__(256) bouzidi_kernel(S* __restrict__ f,
                                                      const uint32_t* __restrict__ store,
                                                      const uint32_t* __restrict__ load0,
                                                      const uint32_t* __restrict__ load1,
                                                      const typename Codec<S>::C* __restrict__ ca,
                                                      const typename Codec<S>::C* __restrict__ cb,
                                                      const typename Codec<S>::C* __restrict__ rhs,
                                                      int64_t n, typename Codec<S>::C* __restrict__ work) {
  using C = typename Codec<S>::C;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const C v0 = Codec<S>::load_dev(f, load0[t], (C)0);
  const C v1 = Codec<S>::load_dev(f, load1[t], (C)0);
  C r;
  {
    // every product and sum rounds on its own (no FMA): the arithmetic the generated
    // step uses for the same links folded into its pulls, and numpy's
#pragma clang fp contract(off)
    r = ca[t] * v0 + cb[t] * v1;
    if (rhs) r += rhs[t];
  }
  if (work)
    work[t] = r;
  else
    Codec<S>::store(f, store[t], r, (C)0);
}

template <typename S>
static int bouzidi_launch(void* d_f, const uint32_t* st, const uint32_t* l0, const uint32_t* l1,
                          const void* ca, const void* cb, const void* rhs, int64_t n, int two_phase,
                          void* work, hipStream_t s) {
  using C = typename Codec<S>::C;
  const int64_t blocks = (n + 255) / 256;
  LBM_REQUIRE(!two_phase || work != nullptr, "two-phase link list needs a workspace");
  bouzidi_kernel<S><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(
      (S*)d_f, st, l0, l1, (const C*)ca, (const C*)cb, (const C*)rhs, n, two_phase ? (C*)work : nullptr);
  if (two_phase)
    links_scatter_kernel<S><<<dim3((unsigned)blocks), dim3(256), 0, s>>>((S*)d_f, st, n, (const C*)work);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

// ca = sign * s and cb = 1 - s from the FP64 per-link s, each rounded once to C
template <typename C>
__global__ void __launch_bounds__(256) bouzidi_coefficients_kernel(const double* __restrict__ dist, double sign,
                                                                   C* __restrict__ ca, C* __restrict__ cb,
                                                                   int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double s = dist[t];
  ca[t] = (C)(sign * s);
  cb[t] = (C)(1.0 - s);
}

// pylbm's FP64 rhs of one link list in the form the link kernels take it: the
// compute type, and for the deviation formats (which combine stored deviations)
// the rest-population term a0 w(load0) [+ a1 w(load1)] - w(store) added, in the
// association the host lists use (contraction off, so every product and sum
// rounds once as numpy's do).  Plain lists: a0 = sign; Bouzidi lists (load1 set):
// a0 = sign dist, a1 = 1 - dist.
struct RestTable {
  double w[LBM_MAX_Q];
};

__device__ __forceinline__ int population_of(const lbm_layout& L, uint32_t idx) {
  if (L.pop_stride > L.row_stride) return (int)((int64_t)idx / L.pop_stride);
  return (int)(((int64_t)idx / L.pop_stride) % L.ns);
}

template <typename C>
__global__ void __launch_bounds__(256) rhs_prepare_kernel(const lbm_layout L, const RestTable W, int shift,
                                                          const uint32_t* __restrict__ store,
                                                          const uint32_t* __restrict__ load0,
                                                          const uint32_t* __restrict__ load1,
                                                          const double* __restrict__ dist, double sign,
                                                          const double* __restrict__ rhs, C* __restrict__ out,
                                                          int64_t n) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double r = rhs ? rhs[t] : 0.0;
  if (shift) {
    const double ws = W.w[population_of(L, store[t])];
    const double w0 = W.w[population_of(L, load0[t])];
    if (load1) {
      const double s = dist[t];
      const double w1 = W.w[population_of(L, load1[t])];
      r = r + (((sign * s) * w0 + (1.0 - s) * w1) - ws);
    } else {
      r = (r + sign * w0) - ws;
    }
  }
  out[t] = (C)r;
}

// Several independent link lists (no list reads or writes a slot another one writes)
// in one launch: thread t belongs to the list j with end[j - 1] <= t < end[j] and
// applies that list's link t - end[j - 1] exactly as links_kernel does (or, for a
// one-phase interpolated list, load1[j] set, exactly as bouzidi_kernel does).
constexpr int LBM_MAX_LINK_GROUP = 8;

template <typename C>
struct LinkGroup {
  const uint32_t* store[LBM_MAX_LINK_GROUP];
  const uint32_t* load[LBM_MAX_LINK_GROUP];
  const C* rhs[LBM_MAX_LINK_GROUP];
  const int8_t* sgn[LBM_MAX_LINK_GROUP];
  const uint32_t* load1[LBM_MAX_LINK_GROUP];
  const C* ca[LBM_MAX_LINK_GROUP];
  const C* cb[LBM_MAX_LINK_GROUP];
  float sign[LBM_MAX_LINK_GROUP];
  int64_t end[LBM_MAX_LINK_GROUP];
  int count;
};

template <typename S>
__global__ void __launch_bounds__(256) links_group_kernel(S* __restrict__ f,
                                                          const LinkGroup<typename Codec<S>::C> G) {
  using C = typename Codec<S>::C;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.end[G.count - 1]) return;
  int j = 0;
  while (t >= G.end[j]) ++j;
  const int64_t i = t - (j ? G.end[j - 1] : 0);
  C r;
  {
    // no FMA, as in bouzidi_kernel: the step's folded pulls round every product
    // and sum on its own
#pragma clang fp contract(off)
    if (G.load1[j]) {
      const C v0 = Codec<S>::load_dev(f, G.load[j][i], (C)0);
      const C v1 = Codec<S>::load_dev(f, G.load1[j][i], (C)0);
      r = G.ca[j][i] * v0 + G.cb[j][i] * v1;
    } else {
      C v = Codec<S>::load_dev(f, G.load[j][i], (C)0);
      const C a = G.sgn[j] ? (C)G.sgn[j][i] : (C)G.sign[j];
      r = a * v;
    }
    if (G.rhs[j]) r += G.rhs[j][i];
  }
  Codec<S>::store(f, G.store[j][i], r, (C)0);
}

template <typename S>
static int links_group_launch(void* d_f, const lbm_link_set* sets, int nsets, hipStream_t s) {
  using C = typename Codec<S>::C;
  LinkGroup<C> G{};
  int64_t total = 0;
  G.count = nsets;
  for (int j = 0; j < nsets; ++j) {
    G.store[j] = sets[j].d_store;
    G.load[j] = sets[j].d_load;
    G.rhs[j] = (const C*)sets[j].d_rhs;
    G.sgn[j] = sets[j].d_sign;
    G.load1[j] = sets[j].d_load1;
    G.ca[j] = (const C*)sets[j].d_ca;
    G.cb[j] = (const C*)sets[j].d_cb;
    G.sign[j] = (float)sets[j].sign;
    total += sets[j].n;
    G.end[j] = total;
  }
  if (total == 0) return 0;
  links_group_kernel<S><<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>((S*)d_f, G);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

static int links_group_dispatch(void* d_f, const lbm_layout* layout, const lbm_link_set* sets, int nsets,
                                hipStream_t s) {
  LBM_DISPATCH_STORAGE(layout->storage, links_group_launch, d_f, sets, nsets, s);
}

// ---------------------------------------------------- compensated sums
// Neumaier accumulation per thread, TwoSum-based pairwise combination of
// (sum, compensation) pairs across the block and then across blocks.
struct Acc {
  double s, c;
};

__device__ __forceinline__ void neumaier_add(Acc& a, double x) {
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x))
    a.c += (a.s - t) + x;
  else
    a.c += (x - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ Acc acc_merge(Acc a, Acc b) {
  Acc r = a;
  neumaier_add(r, b.s);
  r.c += b.c;
  return r;
}

constexpr int kRedThreads = 256;
constexpr int64_t kMaxRedBlocks = 1024;   // blocks per population of the first reduction pass

template <typename S>
__global__ void __launch_bounds__(kRedThreads) reduce_pop_kernel(const S* __restrict__ f,
                                                                 lbm_layout L,
                                                                 const double* __restrict__ wshift,
                                                                 double* __restrict__ partial) {
  // grid: (nblocks, ns); rows (o, m) of the interior are distributed over blocks
  const int k = blockIdx.y;
  const int64_t nrows = interior(L, 0) * interior(L, 1);
  const int64_t nin = interior(L, 2);
  const double w = wshift ? wshift[k] : 0.0;
  Acc a{0.0, 0.0};
  const S* fk = f + k * L.pop_stride;
  for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    const int64_t o = L.halo[0] + r / interior(L, 1);
    const int64_t m = L.halo[1] + r % interior(L, 1);
    const int64_t base = cell_offset(L, o, m, L.halo[2]);
    for (int64_t i = threadIdx.x; i < nin; i += blockDim.x)
      neumaier_add(a, Codec<S>::widen(fk[base + i], w));
  }
  __shared__ Acc sh[kRedThreads];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = acc_merge(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * ((int64_t)k * gridDim.x + blockIdx.x)] = sh[0].s;
    partial[2 * ((int64_t)k * gridDim.x + blockIdx.x) + 1] = sh[0].c;
  }
}

__global__ void __launch_bounds__(kRedThreads) reduce_final_kernel(double* __restrict__ partial,
                                                                   int64_t nblocks,
                                                                   double* __restrict__ out) {
  const int k = blockIdx.x;
  Acc a{0.0, 0.0};
  for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
    Acc p{partial[2 * (k * nblocks + b)], partial[2 * (k * nblocks + b) + 1]};
    a = acc_merge(a, p);
  }
  __shared__ Acc sh[kRedThreads];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = acc_merge(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * k] = sh[0].s;
    out[2 * k + 1] = sh[0].c;
  }
}

// ------------------------------------------------------------ plain copy
// 16-byte vector copy, grid-stride; the bandwidth reference of the chip and the
// known-size kernel that calibrates the HBM byte counters.
__global__ void __launch_bounds__(256) copy_kernel(uint4* __restrict__ dst,
                                                   const uint4* __restrict__ src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

template <typename S>
static int reduce_launch(const void* d_f, const lbm_layout* L, const double* wshift, double* partial,
                         dim3 grid, hipStream_t s) {
  reduce_pop_kernel<S><<<grid, kRedThreads, 0, s>>>((const S*)d_f, *L, wshift, partial);
  return 0;
}

static int reduce_launch_any(int storage, const void* d_f, const lbm_layout* L, const double* wshift,
                             double* partial, dim3 grid, hipStream_t s) {
  LBM_DISPATCH_STORAGE(storage, reduce_launch, d_f, L, wshift, partial, grid, s);
}

// ---------------------------------------- scratch of the allocation-free reduction
// The per-block (sum, compensation) pairs of every population plus the final pairs,
// for the largest grid lbm_reduce_blocks gives (kMaxRedBlocks blocks) and LBM_MAX_Q
// populations: 2 * 32 * 1025 doubles = 512.5 KiB, held by every GPU the library runs
// on until the process ends.  In host code `scratch` names a host-side stand-in;
// the device address comes from hipGetSymbolAddress, per device.
constexpr int64_t kScratchDoubles = 2 * (int64_t)LBM_MAX_Q * (kMaxRedBlocks + 1);
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
static int scratch_for_current_device(double** out) {
  int dev = 0;
  LBM_CHECK_HIP(hipGetDevice(&dev));
  LBM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d is beyond the scratch table (%d devices)", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    LBM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// both reduction entry points: partials of nb blocks per population into `partial`
// (2 ns (nb + 1) doubles), the final pairs at its tail, copied back to the host
static int reduce_into(const void* d_f, const lbm_layout* layout, const double* d_wshift, double* partial,
                       int64_t nb, double* h_sum, double* h_comp, hipStream_t s) {
  dim3 grid((unsigned)nb, (unsigned)layout->ns);
  int rc = reduce_launch_any(layout->storage, d_f, layout, d_wshift, partial, grid, s);
  if (rc) return rc;
  LBM_CHECK_HIP(hipGetLastError());
  double* out = partial + 2 * (int64_t)layout->ns * nb;
  reduce_final_kernel<<<dim3((unsigned)layout->ns), kRedThreads, 0, s>>>(partial, nb, out);
  LBM_CHECK_HIP(hipGetLastError());
  double tmp[2 * LBM_MAX_Q];
  LBM_CHECK_HIP(hipMemcpyAsync(tmp, out, sizeof(double) * 2 * layout->ns, hipMemcpyDeviceToHost, s));
  LBM_CHECK_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < layout->ns; ++k) {
    h_sum[k] = tmp[2 * k];
    h_comp[k] = tmp[2 * k + 1];
  }
  return 0;
}

}  // namespace lbm

using namespace lbm;

// ======================================================================= C ABI
extern "C" {

const char* lbm_last_error(void) { return g_err; }

int lbm_device_count(int* count) {
  LBM_REQUIRE(count, "count is NULL");
  LBM_CHECK_HIP(hipGetDeviceCount(count));
  return 0;
}

int lbm_set_device(int device) {
  LBM_CHECK_HIP(hipSetDevice(device));
  // set-up of the allocation-free entry points on this device
  double* p = nullptr;
  return scratch_for_current_device(&p);
}

int lbm_device_name(int device, char* buf, int len) {
  hipDeviceProp_t p;
  LBM_CHECK_HIP(hipGetDeviceProperties(&p, device));
  snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return 0;
}

int lbm_malloc(void** d_ptr, uint64_t bytes) {
  LBM_REQUIRE(d_ptr, "d_ptr is NULL");
  *d_ptr = nullptr;
  LBM_CHECK_HIP(hipMalloc(d_ptr, bytes ? bytes : 256));
  return 0;
}

int lbm_free(void* d_ptr) {
  if (d_ptr) LBM_CHECK_HIP(hipFree(d_ptr));
  return 0;
}

int lbm_memcpy_h2d(void* d_dst, const void* h_src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  LBM_CHECK_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  LBM_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int lbm_memcpy_d2h(void* h_dst, const void* d_src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  LBM_CHECK_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  LBM_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int lbm_memcpy_d2d(void* d_dst, const void* d_src, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  LBM_CHECK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return 0;
}

int lbm_memset(void* d_ptr, int value, uint64_t bytes, void* stream) {
  if (!bytes) return 0;
  LBM_CHECK_HIP(hipMemsetAsync(d_ptr, value, bytes, (hipStream_t)stream));
  return 0;
}

int lbm_stream_create(void** stream) {
  LBM_REQUIRE(stream, "stream is NULL");
  hipStream_t s;
  LBM_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return 0;
}

int lbm_stream_destroy(void* stream) {
  if (stream) LBM_CHECK_HIP(hipStreamDestroy((hipStream_t)stream));
  return 0;
}

int lbm_stream_sync(void* stream) {
  LBM_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int lbm_device_sync(void) {
  LBM_CHECK_HIP(hipDeviceSynchronize());
  return 0;
}

int lbm_event_create(void** event) {
  LBM_REQUIRE(event, "event is NULL");
  hipEvent_t e;
  LBM_CHECK_HIP(hipEventCreate(&e));
  *event = (void*)e;
  return 0;
}

int lbm_event_destroy(void* event) {
  if (event) LBM_CHECK_HIP(hipEventDestroy((hipEvent_t)event));
  return 0;
}

int lbm_event_record(void* event, void* stream) {
  LBM_CHECK_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return 0;
}

int lbm_event_elapsed_ms(void* start, void* stop, float* ms) {
  LBM_CHECK_HIP(hipEventSynchronize((hipEvent_t)stop));
  LBM_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return 0;
}

int lbm_event_sync(void* event) {
  LBM_CHECK_HIP(hipEventSynchronize((hipEvent_t)event));
  return 0;
}

int lbm_stream_wait_event(void* stream, void* event) {
  LBM_CHECK_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
  return 0;
}

int lbm_host_malloc(void** h_ptr, uint64_t bytes) {
  LBM_REQUIRE(h_ptr, "h_ptr is NULL");
  *h_ptr = nullptr;
  LBM_CHECK_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 64, hipHostMallocDefault));
  return 0;
}

int lbm_host_free(void* h_ptr) {
  if (h_ptr) LBM_CHECK_HIP(hipHostFree(h_ptr));
  return 0;
}

static int ghost_fill_impl(void* d_f, const lbm_layout* L, const lbm_velocities* V, int all_pops,
                           void* stream) {
  int rc = check_layout(L);
  if (rc) return rc;
  LBM_REQUIRE(V && V->q == L->ns, "velocity table does not match the layout");
  const int64_t n0 = L->n[0], n1 = L->n[1], n2 = L->n[2];
  const int h0 = L->halo[0], h1 = L->halo[1], h2 = L->halo[2];
  // directional mode: only the slabs of periodic faces (their first ghost axis
  // is the slab's axis); full mode: every slab
  const int faces = all_pops ? 0x3F : V->periodic_faces;
  ShellDesc S;
  S.sidesA = faces & 3;
  S.sidesB = (faces >> 2) & 3;
  S.sidesC = (faces >> 4) & 3;
  S.nA = (int64_t)h0 * nsides(S.sidesA) * n1 * n2;
  S.nB = (n0 - 2 * h0) * (int64_t)h1 * nsides(S.sidesB) * n2;
  S.nC = (n0 - 2 * h0) * (n1 - 2 * h1) * (int64_t)h2 * nsides(S.sidesC);
  const int64_t total = S.nA + S.nB + S.nC;
  if (total == 0) return 0;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  switch (L->elem_size) {
    case 8: ghost_launch<uint64_t>(d_f, *L, *V, S, all_pops, blocks, s); break;
    case 4: ghost_launch<uint32_t>(d_f, *L, *V, S, all_pops, blocks, s); break;
    case 2: ghost_launch<uint16_t>(d_f, *L, *V, S, all_pops, blocks, s); break;
    default: LBM_REQUIRE(false, "bad element size %d", L->elem_size);
  }
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_ghost_fill(void* d_f, const lbm_layout* layout, const lbm_velocities* vel, void* stream) {
  return ghost_fill_impl(d_f, layout, vel, 0, stream);
}

int lbm_ghost_fill_all(void* d_f, const lbm_layout* layout, const lbm_velocities* vel,
                       void* stream) {
  return ghost_fill_impl(d_f, layout, vel, 1, stream);
}

int lbm_links_apply(void* d_f, const lbm_layout* layout, const uint32_t* d_store,
                    const uint32_t* d_load, const void* d_rhs, const int8_t* d_sign, double sign,
                    int64_t n, int two_phase, void* d_work, void* stream) {
  int rc = check_layout(layout);
  if (rc) return rc;
  if (n <= 0) return 0;
  LBM_REQUIRE(d_store && d_load, "store/load index arrays are NULL");
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(layout->storage, links_launch, d_f, d_store, d_load, d_rhs, d_sign, sign, n,
                       two_phase, d_work, s);
}

int lbm_links_apply_bouzidi(void* d_f, const lbm_layout* layout, const uint32_t* d_store,
                            const uint32_t* d_load0, const uint32_t* d_load1, const void* d_ca,
                            const void* d_cb, const void* d_rhs, int64_t n, int two_phase, void* d_work,
                            void* stream) {
  int rc = check_layout(layout);
  if (rc) return rc;
  if (n <= 0) return 0;
  LBM_REQUIRE(n < ((int64_t)1 << 31) * 256, "link list too long");
  LBM_REQUIRE(d_store && d_load0 && d_load1 && d_ca && d_cb, "Bouzidi index or coefficient arrays are NULL");
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(layout->storage, bouzidi_launch, d_f, d_store, d_load0, d_load1, d_ca, d_cb, d_rhs,
                       n, two_phase, d_work, s);
}

int lbm_links_bouzidi_coefficients(const double* d_dist, double sign, void* d_ca, void* d_cb, int64_t n,
                                   int storage, void* stream) {
  if (n <= 0) return 0;
  LBM_REQUIRE(d_dist && d_ca && d_cb, "Bouzidi coefficient arrays are NULL");
  LBM_REQUIRE(n < ((int64_t)1 << 31) * 256, "link list too long");
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (storage == LBM_F64)
    bouzidi_coefficients_kernel<double><<<blocks, 256, 0, s>>>(d_dist, sign, (double*)d_ca, (double*)d_cb, n);
  else
    bouzidi_coefficients_kernel<float><<<blocks, 256, 0, s>>>(d_dist, sign, (float*)d_ca, (float*)d_cb, n);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_links_prepare_rhs(const lbm_layout* layout, const double* h_wshift, const uint32_t* d_store,
                          const uint32_t* d_load0, const uint32_t* d_load1, const double* d_dist, double sign,
                          const double* d_rhs, void* d_out, int64_t n, void* stream) {
  int rc = check_layout(layout);
  if (rc) return rc;
  if (n <= 0) return 0;
  LBM_REQUIRE(n < ((int64_t)1 << 31) * 256, "link list too long");
  LBM_REQUIRE(d_out != nullptr, "rhs output array is NULL");
  const int shift = is_deviation_storage(layout->storage) ? 1 : 0;
  RestTable W{};
  if (shift) {
    LBM_REQUIRE(h_wshift != nullptr, "deviation storage needs the rest populations");
    LBM_REQUIRE(d_store && d_load0, "store/load index arrays are NULL");
    LBM_REQUIRE(!d_load1 || d_dist, "a Bouzidi list needs its dist array");
    for (int k = 0; k < layout->ns; ++k) W.w[k] = h_wshift[k];
  }
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  if (layout->storage == LBM_F64)
    rhs_prepare_kernel<double><<<blocks, 256, 0, s>>>(*layout, W, shift, d_store, d_load0, d_load1, d_dist, sign,
                                                      d_rhs, (double*)d_out, n);
  else
    rhs_prepare_kernel<float><<<blocks, 256, 0, s>>>(*layout, W, shift, d_store, d_load0, d_load1, d_dist, sign,
                                                     d_rhs, (float*)d_out, n);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_links_apply_sets(void* d_f, const lbm_layout* layout, const lbm_link_set* sets, int nsets,
                         void* stream) {
  int rc = check_layout(layout);
  if (rc) return rc;
  LBM_REQUIRE(nsets >= 0 && (nsets == 0 || sets != nullptr), "link sets: NULL array");
  hipStream_t s = (hipStream_t)stream;
  int j = 0;
  while (j < nsets) {
    // a group: set j and the following sets marked merge_prev (at most 8, none two-phase)
    int k = j + 1;
    if (!sets[j].two_phase) {
      while (k < nsets && k - j < LBM_MAX_LINK_GROUP && sets[k].merge_prev && !sets[k].two_phase) ++k;
    }
    if (k - j == 1 && sets[j].d_load1) {
      const lbm_link_set& L = sets[j];
      rc = lbm_links_apply_bouzidi(d_f, layout, L.d_store, L.d_load, L.d_load1, L.d_ca, L.d_cb, L.d_rhs, L.n,
                                   L.two_phase, L.d_work, stream);
    } else if (k - j == 1) {
      const lbm_link_set& L = sets[j];
      rc = lbm_links_apply(d_f, layout, L.d_store, L.d_load, L.d_rhs, L.d_sign, L.sign, L.n, L.two_phase,
                           L.d_work, stream);
    } else {
      for (int q = j; q < k; ++q) {
        LBM_REQUIRE(sets[q].n == 0 || (sets[q].d_store && sets[q].d_load), "store/load index arrays are NULL");
        LBM_REQUIRE(sets[q].n == 0 || !sets[q].d_load1 || (sets[q].d_ca && sets[q].d_cb),
                    "Bouzidi coefficient arrays are NULL");
      }
      rc = links_group_dispatch(d_f, layout, sets + j, k - j, s);
    }
    if (rc) return rc;
    j = k;
  }
  return 0;
}

int lbm_copy(void* d_dst, const void* d_src, uint64_t bytes, void* stream) {
  LBM_REQUIRE(bytes % 16 == 0, "lbm_copy needs a multiple of 16 bytes");
  LBM_REQUIRE(((uintptr_t)d_dst | (uintptr_t)d_src) % 16 == 0, "lbm_copy needs 16-byte alignment");
  const int64_t n = (int64_t)(bytes / 16);
  if (n == 0) return 0;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  copy_kernel<<<dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream>>>((uint4*)d_dst,
                                                                           (const uint4*)d_src, n);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_reduce_blocks(const lbm_layout* layout, int64_t* nblocks) {
  int rc = check_layout(layout);
  if (rc) return rc;
  const int64_t rows = interior(*layout, 0) * interior(*layout, 1);
  *nblocks = rows < kMaxRedBlocks ? rows : kMaxRedBlocks;
  return 0;
}

int lbm_reduce_populations(const void* d_f, const lbm_layout* layout, const double* d_wshift,
                           double* d_partial, double* h_sum, double* h_comp, void* stream) {
  int64_t nb = 0;
  int rc = lbm_reduce_blocks(layout, &nb);
  if (rc) return rc;
  LBM_REQUIRE(d_partial && h_sum && h_comp, "NULL output");
  return reduce_into(d_f, layout, d_wshift, d_partial, nb, h_sum, h_comp, (hipStream_t)stream);
}

int lbm_reduce_populations_noalloc(const void* d_f, const lbm_layout* layout, const double* d_wshift,
                                   double* h_sum, double* h_comp, void* stream) {
  int64_t nb = 0;
  int rc = lbm_reduce_blocks(layout, &nb);
  if (rc) return rc;
  LBM_REQUIRE(h_sum && h_comp, "NULL output");
  const int64_t need = 2 * (int64_t)layout->ns * (nb + 1);
  LBM_REQUIRE(need <= kScratchDoubles, "the reduction needs %lld doubles of scratch, the maximum is %lld",
              (long long)need, (long long)kScratchDoubles);
  double* part = nullptr;
  rc = scratch_for_current_device(&part);
  if (rc) return rc;
  return reduce_into(d_f, layout, d_wshift, part, nb, h_sum, h_comp, (hipStream_t)stream);
}

}  // extern "C"
