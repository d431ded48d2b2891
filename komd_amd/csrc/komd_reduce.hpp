// Deterministic block reductions of the energy / virial passes.
#pragma once
#include <hip/hip_runtime.h>

namespace komd {

// Components of a symmetric 3x3 tensor in the order XX XY YY XZ YZ ZZ
// (PressureSensor.h:18).
constexpr int kSym = 6;

// w += sym(a (x) b) = (a_p b_q + a_q b_p) / 2 per component.
template <typename T>
__device__ __forceinline__ void sym_outer(T* w, T ax, T ay, T az, T bx, T by, T bz) {
    w[0] += ax * bx;
    w[1] += T(0.5) * (ax * by + ay * bx);
    w[2] += ay * by;
    w[3] += T(0.5) * (ax * bz + az * bx);
    w[4] += T(0.5) * (ay * bz + az * by);
    w[5] += az * bz;
}

// v from the lane that DPP control CTRL selects within its 16-lane row
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// the same for a double (both halves moved)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over aligned groups of `lps` lanes (2, 4, 8 or 16; uniform), the group total in every
// lane of the group: quad_perm [1,0,3,2] and [2,3,0,1], then row_half_mirror (lane i and
// 7 - i of each 8) and row_mirror (i and 15 - i of each 16).  No LDS; a + b = b + a, so the
// lanes of a group agree bit for bit.
__device__ __forceinline__ float group_sum(float v, int lps) {
    v += dpp_f32<0xB1>(v);
    if (lps > 2) v += dpp_f32<0x4E>(v);
    if (lps > 4) v += dpp_f32<0x141>(v);
    if (lps > 8) v += dpp_f32<0x140>(v);
    return v;
}

// Sum of v over the workgroup in a fixed order (blockDim.x a multiple of 64, at most
// 1024); every thread returns the total.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[16];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int k = 0; k < nw; k++) s += sh[k];
    return s;
}

// part[k * nparts + slot] = block total of w[k], k < kSym.
template <typename T>
__device__ __forceinline__ void block_store_sym(double* __restrict__ part, int nparts, int slot, const T* w) {
#pragma unroll
    for (int k = 0; k < kSym; k++) {
        const double s = block_sum((double)w[k]);
        if (threadIdx.x == 0) part[(size_t)k * nparts + slot] = s;
    }
}

}  // namespace komd
