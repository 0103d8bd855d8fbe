// Device-side helpers shared by the core runtime and the generated scheme
// kernels: population storage codecs, layout arithmetic, error plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "pylbm_hip.h"

namespace lbm {

// ------------------------------------------------------------------ errors
// One message per host thread; set by every failing entry point.
void set_error(const char* fmt, ...);

#define LBM_CHECK_HIP(expr)                                                  \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::lbm::set_error("%s failed: %s", #expr, hipGetErrorString(e_));       \
      return (int)e_;                                                        \
    }                                                                        \
  } while (0)

#define LBM_REQUIRE(cond, ...)                                               \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::lbm::set_error(__VA_ARGS__);                                         \
      return -1;                                                             \
    }                                                                        \
  } while (0)

// ------------------------------------------------------------ bfloat16 bits
struct bf16 {
  uint16_t bits;
};

__device__ __forceinline__ float bf16_to_float(bf16 h) {
  return __uint_as_float(((uint32_t)h.bits) << 16);
}

// round to nearest even (no NaN payload handling: populations are finite)
__device__ __forceinline__ bf16 float_to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  bf16 r;
  r.bits = (uint16_t)(u >> 16);
  return r;
}

// Population stores; LBM_NT_STORES=1 makes them non-temporal (streamed past the caches).
#ifndef LBM_NT_STORES
#define LBM_NT_STORES 0
#endif
template <typename T>
__device__ __forceinline__ void put(T* p, T v) {
#if LBM_NT_STORES
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// FP32 deviation storage (distinct type so it gets its own codec)
struct f32dev {
  float x;
};

// ------------------------------------------------------------ storage codec
// Codec<S>::C is the compute type.  load/store convert a stored population
// to/from its full value; load_dev/store_dev to/from its deviation from the
// rest population w (= w_k rho0): the deviation formats store it directly, the
// FP32 full-value format subtracts / adds w (exact for w/2 <= f <= 2w).
template <typename S>
struct Codec;

template <>
struct Codec<double> {
  using C = double;
  static constexpr int kind = LBM_F64;
  __device__ __forceinline__ static C decode(double r, C) { return r; }
  __device__ __forceinline__ static C decode_dev(double r, C w) { return r - w; }
  __device__ __forceinline__ static double encode(C v, C) { return v; }
  __device__ __forceinline__ static double encode_dev(C d, C w) { return w + d; }
  __device__ __forceinline__ static C load(const double* p, int64_t i, C) { return p[i]; }
  __device__ __forceinline__ static void store(double* p, int64_t i, C v, C) { put(p + i, v); }
  __device__ __forceinline__ static C load_dev(const double* p, int64_t i, C w) { return p[i] - w; }
  __device__ __forceinline__ static void store_dev(double* p, int64_t i, C d, C w) { put(p + i, w + d); }
  __device__ __forceinline__ static double widen(double s, double) { return s; }
};

template <>
struct Codec<float> {
  using C = float;
  static constexpr int kind = LBM_F32;
  __device__ __forceinline__ static C decode(float r, C) { return r; }
  __device__ __forceinline__ static C decode_dev(float r, C w) { return r - w; }
  __device__ __forceinline__ static float encode(C v, C) { return v; }
  __device__ __forceinline__ static float encode_dev(C d, C w) { return w + d; }
  __device__ __forceinline__ static C load(const float* p, int64_t i, C) { return p[i]; }
  __device__ __forceinline__ static void store(float* p, int64_t i, C v, C) { put(p + i, v); }
  __device__ __forceinline__ static C load_dev(const float* p, int64_t i, C w) { return p[i] - w; }
  __device__ __forceinline__ static void store_dev(float* p, int64_t i, C d, C w) { put(p + i, w + d); }
  __device__ __forceinline__ static double widen(float s, double) { return (double)s; }
};

template <>
struct Codec<f32dev> {
  using C = float;
  static constexpr int kind = LBM_F32D;
  __device__ __forceinline__ static C decode(f32dev r, C w) { return w + r.x; }
  __device__ __forceinline__ static C decode_dev(f32dev r, C) { return r.x; }
  __device__ __forceinline__ static f32dev encode(C v, C w) { return f32dev{v - w}; }
  __device__ __forceinline__ static f32dev encode_dev(C d, C) { return f32dev{d}; }
  __device__ __forceinline__ static C load(const f32dev* p, int64_t i, C w) { return w + p[i].x; }
  __device__ __forceinline__ static void store(f32dev* p, int64_t i, C v, C w) { put(&p[i].x, v - w); }
  __device__ __forceinline__ static C load_dev(const f32dev* p, int64_t i, C) { return p[i].x; }
  __device__ __forceinline__ static void store_dev(f32dev* p, int64_t i, C d, C) { put(&p[i].x, d); }
  __device__ __forceinline__ static double widen(f32dev s, double w) { return w + (double)s.x; }
};

template <>
struct Codec<_Float16> {
  using C = float;
  static constexpr int kind = LBM_F16D;
  __device__ __forceinline__ static C decode(_Float16 r, C w) { return w + (float)r; }
  __device__ __forceinline__ static C decode_dev(_Float16 r, C) { return (float)r; }
  __device__ __forceinline__ static _Float16 encode(C v, C w) { return (_Float16)(v - w); }
  __device__ __forceinline__ static _Float16 encode_dev(C d, C) { return (_Float16)d; }
  __device__ __forceinline__ static C load(const _Float16* p, int64_t i, C w) {
    return w + (float)p[i];
  }
  __device__ __forceinline__ static void store(_Float16* p, int64_t i, C v, C w) {
    put(p + i, (_Float16)(v - w));
  }
  __device__ __forceinline__ static C load_dev(const _Float16* p, int64_t i, C) { return (float)p[i]; }
  __device__ __forceinline__ static void store_dev(_Float16* p, int64_t i, C d, C) { put(p + i, (_Float16)d); }
  __device__ __forceinline__ static double widen(_Float16 s, double w) { return w + (double)(float)s; }
};

template <>
struct Codec<bf16> {
  using C = float;
  static constexpr int kind = LBM_BF16D;
  __device__ __forceinline__ static C decode(bf16 r, C w) { return w + bf16_to_float(r); }
  __device__ __forceinline__ static C decode_dev(bf16 r, C) { return bf16_to_float(r); }
  __device__ __forceinline__ static bf16 encode(C v, C w) { return float_to_bf16(v - w); }
  __device__ __forceinline__ static bf16 encode_dev(C d, C) { return float_to_bf16(d); }
  __device__ __forceinline__ static C load(const bf16* p, int64_t i, C w) {
    return w + bf16_to_float(p[i]);
  }
  __device__ __forceinline__ static void store(bf16* p, int64_t i, C v, C w) {
    put(&p[i].bits, float_to_bf16(v - w).bits);
  }
  __device__ __forceinline__ static C load_dev(const bf16* p, int64_t i, C) { return bf16_to_float(p[i]); }
  __device__ __forceinline__ static void store_dev(bf16* p, int64_t i, C d, C) { put(&p[i].bits, float_to_bf16(d).bits); }
  __device__ __forceinline__ static double widen(bf16 s, double w) {
    return w + (double)bf16_to_float(s);
  }
};

// ------------------------------------------------- vector load / store
// VEC consecutive stored populations per thread.  Loads may be aligned to one
// element only (pulled rows are shifted by the velocity); stores are aligned to
// the whole vector (interior rows start 128-byte aligned, threads own VEC-aligned
// runs).
template <typename S> struct Bits;
template <> struct Bits<double> { using T = uint64_t; };
template <> struct Bits<float> { using T = uint32_t; };
template <> struct Bits<f32dev> { using T = uint32_t; };
template <> struct Bits<_Float16> { using T = uint16_t; };
template <> struct Bits<bf16> { using T = uint16_t; };

// Population loads; LBM_NT_LOADS=1 makes the vector loads non-temporal (measured 9 %
// slower for the D3Q19 FP32 step on MI355X, hence off).
#ifndef LBM_NT_LOADS
#define LBM_NT_LOADS 0
#endif

template <typename S, int V>
__device__ __forceinline__ void vload(S (&out)[V], const S* p) {
#if LBM_NT_LOADS
  if constexpr (V > 1) {
    using B = typename Bits<S>::T;
    typedef B vec_t __attribute__((ext_vector_type(V), aligned(sizeof(B))));
    const vec_t v = __builtin_nontemporal_load((const vec_t*)p);
    __builtin_memcpy(out, &v, sizeof(S) * V);
    return;
  }
#endif
  __builtin_memcpy(out, p, sizeof(S) * V);
}

// Pulled vector of V populations at p[D], p[D + 1], ..., p[D + V - 1], where p is the
// vector-aligned element of the cell (element index a multiple of V) and D the
// compile-time pull offset along the contiguous axis (minus the velocity's inner
// component).  4- and 8-byte formats load p + D directly (dword-aligned for every
// D).  For the 2-byte formats with V = 2 or 4 the element p + D is only 2-byte
// aligned when D is odd, and such a load would be split into 16-bit loads: the
// vector is then assembled from aligned dword loads around it, joined with
// v_alignbyte_b32 (bytes 2-5 of the 8 bytes {hi, lo}).  Every element read lies
// within one element of [p + D, p + D + V - 1].  That stays inside the row:
// |D| <= halo, and where the vector reaches the outer end of a halo (|D| = halo,
// odd) the row's lead before it and its padding after it are not empty (the lead
// aligns the interior start, lead + halo = 0 mod 64 elements, and the pitch rounds
// lead + 2 halo + interior, interior = 0 mod V, up to a multiple of 64).
template <typename S, int V, int D>
__device__ __forceinline__ void vload_pull(S (&out)[V], const S* p) {
  if constexpr (sizeof(S) != 2 || (V != 2 && V != 4) || ((D % V) + V) % V == 0) {
    vload<S, V>(out, p + D);
  } else if constexpr (V == 2) {
    const uint32_t* w = (const uint32_t*)(p + (D - 1));
    const uint32_t o = __builtin_amdgcn_alignbyte(w[1], w[0], 2);
    __builtin_memcpy(&out[0], &o, 4);
  } else {
    constexpr int R = ((D % 4) + 4) % 4;
    uint32_t o0, o1;
    if constexpr (R == 2) {
      const uint32_t* w = (const uint32_t*)(p + D);
      o0 = w[0];
      o1 = w[1];
    } else if constexpr (R == 1) {
      const uint32_t* w = (const uint32_t*)(p + (D - 1));
      const uint2 A = *(const uint2*)w;
      const uint32_t b0 = w[2];
      o0 = __builtin_amdgcn_alignbyte(A.y, A.x, 2);
      o1 = __builtin_amdgcn_alignbyte(b0, A.y, 2);
    } else {
      const uint32_t* w = (const uint32_t*)(p + (D + 1));
      const uint2 A = *(const uint2*)w;
      const uint32_t pm = w[-1];
      o0 = __builtin_amdgcn_alignbyte(A.x, pm, 2);
      o1 = __builtin_amdgcn_alignbyte(A.y, A.x, 2);
    }
    __builtin_memcpy(&out[0], &o0, 4);
    __builtin_memcpy(&out[2], &o1, 4);
  }
}

template <typename S, int V>
__device__ __forceinline__ void vstore(S* p, const S (&in)[V]) {
  using B = typename Bits<S>::T;
  if constexpr (V == 1) {
    B b;
    __builtin_memcpy(&b, in, sizeof(S));
    put((B*)p, b);
  } else {
    typedef B vec_t __attribute__((ext_vector_type(V)));
    vec_t v;
    __builtin_memcpy(&v, in, sizeof(S) * V);
    put((vec_t*)p, v);
  }
}

inline int storage_elem_size(int storage) {
  switch (storage) {
    case LBM_F64: return 8;
    case LBM_F32: return 4;
    case LBM_F16D: return 2;
    case LBM_BF16D: return 2;
    case LBM_F32D: return 4;
    default: return 0;
  }
}

inline bool is_deviation_storage(int storage) {
  return storage == LBM_F16D || storage == LBM_BF16D || storage == LBM_F32D;
}

// Calls FN<S>(args...) with S the storage type of `kind`; returns -1 on an unknown kind.
#define LBM_DISPATCH_STORAGE(kind, FN, ...)                                   \
  do {                                                                        \
    switch (kind) {                                                           \
      case LBM_F64: return FN<double>(__VA_ARGS__);                           \
      case LBM_F32: return FN<float>(__VA_ARGS__);                            \
      case LBM_F16D: return FN<_Float16>(__VA_ARGS__);                        \
      case LBM_BF16D: return FN<::lbm::bf16>(__VA_ARGS__);                    \
      case LBM_F32D: return FN<::lbm::f32dev>(__VA_ARGS__);                   \
    }                                                                         \
    ::lbm::set_error("unknown storage kind %d", (int)(kind));                 \
    return -1;                                                                \
  } while (0)

// interior extents of the canonical axes
__host__ __device__ __forceinline__ int64_t interior(const lbm_layout& L, int a) {
  return L.n[a] - 2 * L.halo[a];
}

// element offset of (o, m, i) inside one population (halo-inclusive indices)
__host__ __device__ __forceinline__ int64_t cell_offset(const lbm_layout& L, int64_t o, int64_t m,
                                                        int64_t i) {
  return (o * L.n[1] + m) * L.row_stride + L.lead + i;
}

// one past the largest element index of the array (any arrangement)
__host__ __device__ __forceinline__ int64_t layout_elements(const lbm_layout& L) {
  const int64_t rows = L.n[0] * L.n[1];
  return (L.ns - 1) * L.pop_stride + (rows - 1) * L.row_stride + L.pitch;
}

// true when every element index and every offset between two elements fits int32
__host__ __device__ __forceinline__ bool layout_fits_int32(const lbm_layout& L) {
  return layout_elements(L) < ((int64_t)1 << 31);
}

inline int check_layout(const lbm_layout* L) {
  LBM_REQUIRE(L != nullptr, "layout is NULL");
  LBM_REQUIRE(L->ns > 0 && L->ns <= LBM_MAX_Q, "ns=%d out of range [1, %d]", L->ns, LBM_MAX_Q);
  LBM_REQUIRE(L->dim >= 1 && L->dim <= 3, "dim=%d out of range", L->dim);
  LBM_REQUIRE(storage_elem_size(L->storage) == L->elem_size, "storage %d / elem_size %d mismatch",
              L->storage, L->elem_size);
  for (int a = 0; a < 3; ++a) {
    LBM_REQUIRE(L->n[a] >= 1 && L->halo[a] >= 0 && L->n[a] > 2 * L->halo[a],
                "axis %d: extent %lld halo %d invalid", a, (long long)L->n[a], L->halo[a]);
  }
  LBM_REQUIRE(L->lead >= 0 && L->lead + L->n[2] <= L->pitch, "row lead/pitch too small");
  const int64_t rows = L->n[0] * L->n[1];
  const bool planes = L->row_stride >= L->pitch && L->pop_stride >= rows * L->row_stride;
  const bool interleaved = L->pop_stride >= L->pitch && L->row_stride >= L->ns * L->pop_stride;
  LBM_REQUIRE(planes || interleaved, "pop_stride / row_stride overlap");
  // element indices are 64-bit: arrays of 2^31 elements and more take the kernels'
  // 64-bit index variants (layout_elements), link lists their own 32-bit check
  const int64_t last = layout_elements(*L);
  LBM_REQUIRE(last > 0 && last < ((int64_t)1 << 62), "array of %lld elements out of range", (long long)last);
  return 0;
}

}  // namespace lbm
