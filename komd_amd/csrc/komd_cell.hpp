// Capacities of the cell-tile force kernel (k_cell.hip).
#pragma once
#include "komd_internal.hpp"

namespace komd {
constexpr int kCellThreads = 256;  // one workgroup (4 waves) per home cell
// pair-pass workgroups resident per CU: 8 waves per SIMD (at most 64 VGPRs) = 8 workgroups of 4 waves
constexpr int kCellWgPerCu = 8;
constexpr int kMaxM = 768;         // staged neighbourhood capacity when the three-body pass runs
constexpr int kChunkLJ = 768;      // staging chunk of the pair-only pass (any neighbourhood size)
constexpr int kMaxOwn = 64;        // home sites per cell when the three-body pass runs
constexpr int kFullCap = 96;       // full neighbour list capacity per home site (three-body pass)
constexpr int kRows = 64;          // sites per cell held by a per-cell partial force row
}  // namespace komd
