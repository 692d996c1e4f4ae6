// Raw buffer loads and stores for the step kernels (gfx9 buffer instructions).
//
// Each access is base (a 128-bit resource in SGPRs) + a 32-bit per-lane byte offset (VGPR) + a
// 32-bit wave-uniform byte offset (SGPR). A line of PTS values per thread then costs one offset
// VGPR in total: the per-value part of the address (q * T rows) is uniform and goes in the SGPR
// operand. Flat 64-bit addressing would keep one VGPR pair per value live across a whole line
// transform. Accesses at or beyond `bytes` of the resource read zero and drop the write.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace s2t {

// word 3 of a raw (stride 0, untyped) buffer resource on gfx9
constexpr int BUF_WORD3 = 0x00020000;

// A per-lane offset beyond every resource's range (and beyond it plus any SGPR offset used here):
// the load returns zero and the store is dropped, with no memory access. Predicating an access
// by its offset keeps it out of a branch, so the wait counts of the loads around it stay exact.
constexpr uint32_t BUF_OOB = 0x80000000u;

struct Buf {
  __amdgpu_buffer_rsrc_t r;
};

__device__ __forceinline__ Buf make_buf(const void* base, uint32_t bytes) {
  return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                               BUF_WORD3)};
}

__device__ __forceinline__ float2 ld2(Buf b, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)voff, (int)soff, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}

__device__ __forceinline__ void st2(Buf b, float2 x, uint32_t voff, uint32_t soff) {
  typedef unsigned u32x2 __attribute__((vector_size(8)));
  u32x2 v;
  v[0] = __float_as_uint(x.x);
  v[1] = __float_as_uint(x.y);
  __builtin_amdgcn_raw_buffer_store_b64(v, b.r, (int)voff, (int)soff, 0);
}

// 16-byte access: two adjacent complex values
__device__ __forceinline__ void ld4(Buf b, uint32_t voff, uint32_t soff, float2& lo, float2& hi) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(b.r, (int)voff, (int)soff, 0);
  lo = make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
  hi = make_float2(__uint_as_float(v[2]), __uint_as_float(v[3]));
}

// The 16-byte element is written as two 8-byte stores. A store of more than 8 bytes reads its
// data VGPRs after it issues, and for gfx950 the compiler does not keep a VALU write of those
// VGPRs away from it: with stores issued back to back (no branch between them) a write landed
// one instruction after the store and corrupted the stored values now and then (results that
// differed run to run). 8-byte stores are not affected.
__device__ __forceinline__ void st4(Buf b, float2 lo, float2 hi, uint32_t voff, uint32_t soff) {
  st2(b, lo, voff, soff);
  st2(b, hi, voff, soff + 8u);
}

// Opaque copy of a value: loop-invariant code that depends on it cannot be hoisted out of the
// loop it is laundered in (keeps per-line address and factor arrays from being held live across
// line transforms).
template <class T>
__device__ __forceinline__ T launder(T x) {
  asm volatile("" : "+v"(x));
  return x;
}

}  // namespace s2t
