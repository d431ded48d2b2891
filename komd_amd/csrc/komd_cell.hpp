// Capacities of the cell-tile force kernel (k_cell.hip).
#pragma once
#include "komd_internal.hpp"
#include "komd_reduce.hpp"

namespace komd {
constexpr int kCellThreads = 256;  // one workgroup (4 waves) per home cell
// pair-pass workgroups resident per CU: 8 waves per SIMD (at most 64 VGPRs) = 8 workgroups of 4 waves
constexpr int kCellWgPerCu = 8;
constexpr int kMaxM = 768;         // staged neighbourhood capacity when the three-body pass runs
constexpr int kChunkLJ = 768;      // staging chunk of the pair-only pass (any neighbourhood size)
constexpr int kMaxOwn = 64;        // home sites per cell when the three-body pass runs
constexpr int kFullCap = 96;       // full neighbour list capacity per home site (three-body pass)
constexpr int kSkinCap = 128;      // Verlet list capacity per site (within cutoff + skin)
// Verlet-list entry: cell-sorted slot of the neighbour (25 bits, so at most 2^25 - 1 sites)
// | the offset class of its cell (0..124) << 25
constexpr uint32_t kSkinIdxBits = 25;
constexpr uint32_t kSkinIdxMask = (1u << kSkinIdxBits) - 1u;
constexpr int kRows = 64;          // sites per cell held by a per-cell partial force row
constexpr int kSurM = 896;         // staged neighbourhood capacity of the surrogate force pass

// Buffers of the Verlet-list passes (one copy per context in device memory, read where they
// are used: as kernel arguments they would stay in scalar registers through a whole kernel).
struct SkinArgs {
    const uint32_t* vl;        // [N][kSkinCap] entries: cell-sorted slot | offset class << 25
    const uint32_t* vlc;       // [N] lengths
    const uint16_t* vld;       // [N][kSkinCap] cell direction | rank in that cell << 5 (or null)
    const float4* spos;        // [N] FP32 positions relative to the build-time cell centres
    const float4* spb;         // [N] the same at the last rebuild
    const float4* skin_off;    // [125] offset of each offset class, image code in .w
    uint32_t* req;             // [2] rebuild words of the two step parities
    unsigned long long* pair_counters;  // in-range unique pairs per workgroup slot
    uint32_t* flags;
    double dt, lim2;           // time step; squared displacement that triggers a rebuild
};

// Would the next Integrator::Step0 (Integrator.cpp:28-32) move site i more than sqrt(lim2)
// from its build position?  The step (v + f dt/(2m)) dt of the drift, after this step's
// second half-kick has been written to v, plus the displacement since the rebuild from the
// FP32 cell-relative positions (both relative to the same build-time cell centre; their
// rounding, below 3e-7, lies inside the margin of lim2).
__device__ __forceinline__ bool skin_exceeded(const double* __restrict__ vx, const double* __restrict__ vy,
                                              const double* __restrict__ vz, const double* __restrict__ mass,
                                              uint32_t i, double dt_half, double dt, double fx, double fy, double fz,
                                              float4 q, float4 qb, double lim2) {
    const double k = dt_half / mass[i];
    const double x = (double)(q.x - qb.x) + (vx[i] + fx * k) * dt;
    const double y = (double)(q.y - qb.y) + (vy[i] + fy * k) * dt;
    const double z = (double)(q.z - qb.z) + (vz[i] + fz * k) * dt;
    return !(x * x + y * y + z * z <= lim2);
}

// ---- helpers shared by the list-driven kernels (k_cell.hip, k_table.hip)

__device__ __forceinline__ double image_shift(uint32_t code, double L) {
    return code == 1u ? L : (code == 2u ? -L : 0.0);
}

// |(r_a + s_a) - (r_b + s_b)|^2 in FP64, reference operation order, no contraction
// (the pragma covers the operators written here; __dmul_rn / __dadd_rn would not be
// covered, their bodies are plain * and + compiled with contraction on).
__device__ __forceinline__ double exact_r2(const double* __restrict__ rx, const double* __restrict__ ry,
                                           const double* __restrict__ rz, const Grid& g, uint32_t wa, uint32_t wb) {
#pragma clang fp contract(off)
    const uint32_t i = wa & kIdxMask, j = wb & kIdxMask, ca = wa >> kIdxBits, cb = wb >> kIdxBits;
    const double xa = rx[i] + image_shift(ca & 3u, g.size[0]);
    const double ya = ry[i] + image_shift((ca >> 2) & 3u, g.size[1]);
    const double za = rz[i] + image_shift((ca >> 4) & 3u, g.size[2]);
    const double xb = rx[j] + image_shift(cb & 3u, g.size[0]);
    const double yb = ry[j] + image_shift((cb >> 2) & 3u, g.size[1]);
    const double zb = rz[j] + image_shift((cb >> 4) & 3u, g.size[2]);
    const double dx = xa - xb, dy = ya - yb, dz = za - zb;
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

// Sum over the wave in a fixed order, the total in every lane: quads by quad_perm,
// 16-lane rows by row_ror (DPP, no LDS), then the four row sums read out.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x124>(v);  // row_ror:4
    v += dpp_f64<0x128>(v);  // row_ror:8
    auto row = [&](int l) {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                                __builtin_amdgcn_readlane(__double2loint(v), l));
    };
    return (row(0) + row(16)) + (row(32) + row(48));
}

// the lanes of the wave whose predicate holds (the builtin takes the predicate as a
// lane mask; the int-taking __ballot rebuilds it from a VGPR)
__device__ __forceinline__ unsigned long long ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// LDS writes of this wave visible to this wave's later reads.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

}  // namespace komd
