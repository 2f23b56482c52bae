// Raw buffer access helpers shared by the strided derivative kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

namespace cssx {

// element at a wave-uniform row base plus a 32-bit per-lane byte offset
template <typename T>
__device__ __forceinline__ T &at_row(const T *row, unsigned byte_ofs) {
    return *reinterpret_cast<T *>(const_cast<char *>(reinterpret_cast<const char *>(row)) + byte_ofs);
}

// Raw buffer access to one field: the lane's byte offset stays in one VGPR for every row and
// the row's byte offset is a scalar register (buffer_load ... v_lane, s[rsrc], s_row offen),
// so a sweep row costs one memory instruction and no address arithmetic per lane.  The
// resource spans 2^32-1 bytes; the host checks that a launch's offsets fit.
typedef unsigned int cssx_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int cssx_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void *base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, -1, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned vofs, unsigned sofs);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned vofs, unsigned sofs) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vofs, sofs, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned vofs, unsigned sofs) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vofs, sofs, 0));
}
__device__ __forceinline__ void bstore(float v, __amdgpu_buffer_rsrc_t r, unsigned vofs, unsigned sofs) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, vofs, sofs, 0);
}
__device__ __forceinline__ void bstore(double v, __amdgpu_buffer_rsrc_t r, unsigned vofs, unsigned sofs) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(cssx_u32x2, v), r, vofs, sofs, 0);
}

// the value of lane-1 (lane 0 gets 0) and of lane+1 (lane 63 gets 0) by DPP wave shifts
// (wave_shr:1 / wave_shl:1) instead of an LDS permute.  A source lane that is disabled reads as
// zero, so a shift must not feed a select on the receiving side (the compiler can turn that
// select into an exec-masked branch around the shift): mask on the sending side instead.
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ T from_prev(T v) { return dpp_shift<0x138>(v); }
template <typename T>
__device__ __forceinline__ T from_next(T v) { return dpp_shift<0x130>(v); }

// v from lane l of the wavefront as a wave-uniform value (v_readlane)
__device__ __forceinline__ float lane_val(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_val(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

}  // namespace cssx
