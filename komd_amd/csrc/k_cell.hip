// Cell-tile pair kernel (LJ 12-6, force clamp, full neighbour lists) and the
// per-site Axilrod-Teller-Muto kernel -- gfx950.
//
// k_cell: one 256-thread workgroup per domain cell ("home cell"); on grids below
// 1024 cells, `split` workgroups share a cell's home sites (about 5 each).
//  * The 27 neighbour cells' sites are staged in LDS (in chunks) as FP32
//    positions relative to the home cell centre, periodic images folded into
//    the staging offset.
//  * Every home site is served by 4-32 consecutive lanes striding over the
//    staged candidates.  FP32 distance test; a candidate within a few ulps of
//    rc^2 is decided in FP64 with the reference's operation order, so the
//    interaction set is the reference's.  LJ force summed per lane in FP64 and
//    combined with shuffles, then clamped (Limit).
//  * When a three-body term follows, each site's full neighbour list
//    (r_j - r_i in FP32, site index, image code) is compacted with a ballot and
//    written to HBM in the site's cell-sorted slot.
// LJ equals the reference's C08 cell-pair walk (LinkedCells.cpp:164-224), in
// which every unshifted member receives the pair force once (LJ12_6.cpp:15-43);
// the clamp is Limit_Force (Limit.cpp:19-25), applied after the two-body terms
// and before the three-body term as main.cpp:57-66 registers them.
//
// k_atm: one wave per site, four independent waves per workgroup.  The site's
// list is loaded into registers (lane a holds x_a = r_a - r_i) and into LDS.
// Pair filter in registers: y is rotated around the 64-lane circle with DPP
// (wave_rol), so shifts k = 1..31 for every lane plus k = 32 for lanes < 32
// visit every unordered neighbour pair once (r_jk^2 from FP16 (x, y) pairs and
// FP32 z, a 1 % candidate margin); the candidates are compacted with a ballot
// into a 128-entry LDS ring as (lane, k) and evaluated 64 at a time without
// divergence, the exact cutoff decision taken there.  Lists longer than 64
// use a circulant walk over LDS in FP32.  Every site accumulates only its
// own three-body force, so there is no scatter and no atomic; each triplet is
// visited by its three members and a visit carries V/3 of the energy.
//
// ATM: V = nu (1 + 3 cos g1 cos g2 cos g3)/(r_ij r_ik r_jk)^3.  With
// x = r_j - r_i, y = r_k - r_i, z = y - x, X,Y,Z = |.|^2, p1 = x.y, p2 = x.z,
// p3 = y.z, D = XYZ, P = -p1 p2 p3:   V = nu D^{-5/2} (D + 3P) and
//   F_i = -dV/dr_i = dV/dx + dV/dy
//       = 3 nu D^{-5/2} [ -(D + 5P)(x/X + y/Y) - p2 p3 (x + y) - (p1 p3 + p1 p2) z ].
// The reference mode applies +grad V (comp_force of ATM.h:36-69 is
// +(1/r) dV/dr, used with that sign in ATM_NOLIST.cpp:148-158) and keeps only
// triplets whose cells are all equal or all distinct (44 C08 cell triples,
// LinkedCells.h:126-152, plus same-cell triples, ATM_NOLIST.h:112-157).
// Hybrid: the triplet energy is weighted by w = (l_i + l_j + l_k)/3, so site i
// also receives -(sum V/3) dl_i/dx_axis.
#include "komd_cell.hpp"
#include "komd_reduce.hpp"

namespace komd {
namespace {

struct CellArgs {
    const float4* __restrict__ spos;
    const float2* __restrict__ sprm;
    const uint32_t* __restrict__ sid;
    const uint32_t* __restrict__ start;
    const double* __restrict__ rx;
    const double* __restrict__ ry;
    const double* __restrict__ rz;
    const double* __restrict__ eps;
    const double* __restrict__ sig;
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    float4* __restrict__ nl;       // [N][kFullCap] neighbour list of every site (cell-sorted order)
    uint32_t* __restrict__ nlc;    // [N] its length
    uint16_t* __restrict__ nld;    // [N][kFullCap] neighbour cell direction | rank in that cell << 5
    uint32_t* __restrict__ vl;     // NL 3: [N][kSkinCap] Verlet list (cell-sorted slot | direction << 26)
    uint32_t* __restrict__ vlc;    // NL 3: [N] its length
    uint16_t* __restrict__ vld;    // NL 4: [N][kSkinCap] neighbour cell direction | rank in that cell << 5
    const uint32_t* __restrict__ req;  // NL 3: rebuild word of the step (0: the lists stay)
    double* __restrict__ part;
    double* __restrict__ vpart;    // [kSym][nblocks] virial tensor partials (energy mode)
    uint32_t* __restrict__ flags;
    unsigned long long* __restrict__ counters;
    double* __restrict__ vx;       // second half-kick fused into the force write (kick > 0)
    double* __restrict__ vy;
    double* __restrict__ vz;
    const double* __restrict__ mass;
    Grid g;
    int chunk, nblocks, split;
    // cx > 0: XCD x owns cells [x cx, (x+1) cx); when its n cells fill one round of `slots`
    // resident workgroups and part of a second, those past 2 slots - n are served by two
    // workgroups each (placed last), so the second round holds half-cells only
    int cx, slots;
    float rc2f, band, sig2, eps24, eps4;
    float rl2f;     // NL 3: (cutoff + skin)^2
    float c24, c4;  // 24 and 4 times the pair-force scale (mixed sites: e24 = c24 eps_ij)
    double rc2, limit, kick;
};

// Per axis, the offset of a neighbour cell falls in one of five classes: 0 same cell
// column, 1 / 2 the cell below / above, 3 / 4 the periodic image below / above (the home
// cell is the first / last of the axis).  A Verlet-list entry carries the triple as one
// number cx + 5 (cy + 5 cz) < 125.
__device__ __forceinline__ float axis_offset(int cls, int n, double cs, double L) {
    const int d = cls == 0 ? 0 : (cls == 1 ? -1 : (cls == 2 ? 1 : (cls == 3 ? n - 1 : -(n - 1))));
    const double s = cls == 3 ? -L : (cls == 4 ? L : 0.0);
    return (float)((double)d * cs + s);  // the FP64 expression of nb_cell
}

// Neighbour cell `dir` (0..26, (dz, dy, dx) order, dx fastest) of `cell`: its flat index,
// the periodic image offset relative to the home cell centre (FP32 of the FP64 value),
// its image code in the site word and (cls) its offset class.
__device__ __forceinline__ int nb_cell(const Grid& g, int cell, int dir, float* off, uint32_t* code,
                                       uint32_t* cls = nullptr) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int hx = cell % nx, hy = (cell / nx) % ny, hz = cell / (nx * ny);
    const int dx = dir % 3 - 1, dy = (dir / 3) % 3 - 1, dz = dir / 9 - 1;
    int x = hx + dx, y = hy + dy, z = hz + dz;
    uint32_t cx = 0, cy = 0, cz = 0;
    double sx = 0, sy = 0, sz = 0;
    if (x < 0) { x += nx; cx = 2; sx = -g.size[0]; } else if (x >= nx) { x -= nx; cx = 1; sx = g.size[0]; }
    if (y < 0) { y += ny; cy = 2; sy = -g.size[1]; } else if (y >= ny) { y -= ny; cy = 1; sy = g.size[1]; }
    if (z < 0) { z += nz; cz = 2; sz = -g.size[2]; } else if (z >= nz) { z -= nz; cz = 1; sz = g.size[2]; }
    off[0] = (float)((double)(x - hx) * g.cs + sx);
    off[1] = (float)((double)(y - hy) * g.cs + sy);
    off[2] = (float)((double)(z - hz) * g.cs + sz);
    *code = (cx | (cy << 2) | (cz << 4)) << kIdxBits;
    if (cls) {
        auto c1 = [](int d, uint32_t img) { return img == 2u ? 3u : (img == 1u ? 4u : (d < 0 ? 1u : (d > 0 ? 2u : 0u))); };
        *cls = c1(dx, cx) + 5u * (c1(dy, cy) + 5u * c1(dz, cz));
    }
    return x + nx * (y + ny * z);
}

// Wave 0 of a workgroup (lanes 0..26 carry a neighbour cell each): the 27 cells around
// `cell` -- start of their sites, the prefix of their counts (nb_pref[27] = sites of the
// neighbourhood), the periodic image offset of each relative to the home cell centre
// (FP32) and its image code in the site word.
__device__ __forceinline__ void nb_setup(const Grid& g, const uint32_t* __restrict__ start, int cell, int tid,
                                         int* nb_start, int* nb_pref, float (*nb_off)[3], uint32_t* nb_code,
                                         uint32_t* nb_cls) {
    int cnt = 0;
    if (tid < 27) {
        const int nc = nb_cell(g, cell, tid, nb_off[tid], &nb_code[tid], &nb_cls[tid]);
        nb_start[tid] = (int)start[nc];
        cnt = (int)(start[nc + 1] - start[nc]);
    }
    int incl = cnt;  // inclusive scan across the wave
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (tid >= off) incl += o;
    }
    if (tid < 27) nb_pref[tid + 1] = incl;
    if (tid == 0) nb_pref[0] = 0;
}

// the neighbour cell holding the gidx-th site of the neighbourhood (binary search of the prefix)
__device__ __forceinline__ int nb_of(const int* nb_pref, int gidx) {
    int k = 0;
#pragma unroll
    for (int step = 16; step > 0; step >>= 1)
        if (k + step < 27 && nb_pref[k + step] <= gidx) k += step;
    return k;
}

// a neighbour's packed position relative to the home cell centre: image offset folded
// into the FP32 coordinates, image code into the site word
__device__ __forceinline__ float4 staged(float4 v, const float* off, uint32_t code) {
    return make_float4(v.x + off[0], v.y + off[1], v.z + off[2], __uint_as_float(__float_as_uint(v.w) | code));
}

// NL: 0 none, 1 neighbour list, 2 neighbour list + cell directions, 3 Verlet list
// (every site within cutoff + skin; the step's rebuild word gates the launch), 4 Verlet
// list + cell directions.
template <bool UNIFORM, bool MOLS, bool LJ, int NL, bool ENERGY, bool LIMIT>
__global__ __launch_bounds__(kCellThreads) void k_cell(CellArgs a) {
    constexpr bool VL = NL >= 3;
    if (VL && a.req[0] == 0u) return;
    constexpr int kStage = kChunkLJ;
    constexpr int kWaves = kCellThreads / 64;
    __shared__ float4 pos[kStage + 32];  // + 32: a group's lanes past the chunk read in bounds
    __shared__ uint16_t sdir[(NL == 2 || NL == 4) ? kStage : 1];
    __shared__ uint32_t sref[VL ? kStage : 1];  // cell-sorted slot | offset class << 25
    __shared__ float2 cprm[UNIFORM ? 1 : kStage];
    __shared__ uint32_t cmol[MOLS ? kStage : 1];
    __shared__ int nb_start[27];
    __shared__ int nb_pref[28];
    __shared__ float nb_off[27][3];
    __shared__ uint32_t nb_code[27];
    __shared__ uint32_t nb_cls[27];
    __shared__ double red[2][kWaves];
    __shared__ unsigned long long red_cnt[kWaves];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int b = blockIdx.x;
    // XCD-aware: consecutive blocks go round-robin to the 8 XCDs; each XCD gets a contiguous slab
    int cell, part, nparts;  // part of nparts: this workgroup's share of the cell's home sites
    if (a.cx == 0) {
        const int lb = (b & 7) * a.chunk + (b >> 3);
        cell = lb / a.split;
        part = lb - cell * a.split;
        nparts = a.split;
    } else {
        const int x = b & 7, j = b >> 3;
        const int n = max(0, min(a.cx, a.g.ncell - x * a.cx));
        const int w = (n <= a.slots || n >= 2 * a.slots) ? n : 2 * a.slots - n;  // whole cells of this XCD
        if (j < w) {
            cell = x * a.cx + j;
            part = 0;
            nparts = 1;
        } else {
            const int i = j - w;
            cell = w + (i >> 1) < n ? x * a.cx + w + (i >> 1) : a.g.ncell;  // past the slab: no cell
            part = i & 1;
            nparts = 2;
        }
    }
    double u_acc = 0.0, w_acc = 0.0;
    double vir[kSym] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // pair virial r_ij (x) F_ij (energy mode)
    unsigned long long npairs = 0;
    const bool live = cell < a.g.ncell;
    int m = 0, hs = 0, nh = 0;
    if (live) {
        if (tid < 64) nb_setup(a.g, a.start, cell, tid, nb_start, nb_pref, nb_off, nb_code, nb_cls);
        __syncthreads();
        // workgroup-uniform values read from LDS: keep them in scalar registers
        m = __builtin_amdgcn_readfirstlane(nb_pref[27]);
        hs = __builtin_amdgcn_readfirstlane(nb_start[13]);
        nh = __builtin_amdgcn_readfirstlane(nb_pref[14] - nb_pref[13]);
        // this workgroup's share of the home sites
        const int nhp = (nh + nparts - 1) / nparts;
        const int hb = min(nh, part * nhp);
        hs += hb;
        nh = min(nh - hb, nhp);
    }
    const int lps = nh <= 8 ? 32 : (nh <= 16 ? 16 : (nh <= 32 ? 8 : 4));
    const int per = kCellThreads / lps;
    const int h = tid / lps, sub = tid % lps;
    // the lps lanes of one home site are consecutive lanes of one wave
    const int gshift = (lane / lps) * lps;
    const unsigned long long gmask = ((1ull << lps) - 1ull) << gshift;  // the lanes of my home site

    for (int base = 0; live && base < nh; base += per) {
        const bool active = base + h < nh;
        const int p = hs + base + h;
        float4 qi = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 pi = make_float2(0.f, 0.f);
        uint32_t mi = 0;
        if (active) {
            qi = a.spos[p];
            if (!UNIFORM) pi = a.sprm[p];
            if (MOLS) mi = a.sid[p];
        }
        const uint32_t wi = __float_as_uint(qi.w);
        const uint32_t oi = wi & kIdxMask;
        double ax = 0.0, ay = 0.0, az = 0.0;
        uint32_t ncnt = 0;
        for (int c0 = 0; c0 < m; c0 += kStage) {
            const int mc = min(kStage, m - c0);
            // the first staging needs no barrier (the stage is untouched since the
            // neighbour-cell barrier), so the home-site loads above overlap its loads
            if (base > 0 || c0 > 0) __syncthreads();
            for (int t = tid; t < mc; t += kCellThreads) {
                const int gidx = c0 + t;
                const int k = nb_of(nb_pref, gidx);
                const int j = nb_start[k] + (gidx - nb_pref[k]);
                pos[t] = staged(a.spos[j], nb_off[k], nb_code[k]);
                if (!UNIFORM) cprm[t] = a.sprm[j];
                if (MOLS) cmol[t] = a.sid[j];
                if (NL == 2 || NL == 4) sdir[t] = (uint16_t)(k | ((gidx - nb_pref[k]) << 5));
                if (VL) sref[t] = (uint32_t)j | (nb_cls[k] << kSkinIdxBits);
            }
            __syncthreads();
            // every lane of the workgroup runs the same trip count (ballots need the whole wave)
            for (int t0 = 0; t0 < mc; t0 += lps) {
                const int t = t0 + sub;
                const bool cand = active && t < mc;
                const float4 c = pos[t];  // t < kStage + 32; lanes without a candidate ignore it
                const uint32_t w = __float_as_uint(c.w);
                const uint32_t oj = w & kIdxMask;
                const float dx = c.x - qi.x, dy = c.y - qi.y, dz = c.z - qi.z;
                // lanes without a candidate test +inf: every predicate below is then one
                // compare, which the ballots take as a lane mask directly
                const float r2 = cand ? dx * dx + dy * dy + dz * dz : __builtin_huge_valf();
                bool in;
                if (VL) {
                    // Verlet list: a superset bound, no exact decision needed
                    in = r2 <= a.rl2f;
                } else {
                    in = r2 <= a.rc2f;
                    // a distance within a few ulps of rc is decided in FP64 (rare: wave-uniform branch)
                    const bool amb = fabsf(r2 - a.rc2f) <= a.band;
                    if (ballot(amb) != 0ull && amb) in = !(exact_r2(a.rx, a.ry, a.rz, a.g, wi, w) > a.rc2);
                }
                const uint32_t oh = in ? oj : oi;  // hit: in range and not i itself
                // Verlet lists of multi-site molecules hold sites of other molecules only
                const bool hit = oh != oi && !(VL && MOLS && cmol[t] == mi);
                if (!VL) npairs += (unsigned long long)__popcll(ballot(oh > oi));  // wave-uniform
                if (VL) {
                    const unsigned long long bal = ballot(hit) & gmask;
                    if (hit) {
                        const uint32_t s = ncnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                        if (s < (uint32_t)kSkinCap) {
                            a.vl[(size_t)p * kSkinCap + s] = sref[t];
                            if (NL == 4) a.vld[(size_t)p * kSkinCap + s] = sdir[t];
                        }
                    }
                    ncnt += (uint32_t)__popcll(bal);
                } else if (NL) {
                    const unsigned long long bal = ballot(hit) & gmask;
                    if (hit) {
                        const uint32_t s = ncnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                        if (s < (uint32_t)kFullCap) {
                            const size_t slot = (size_t)p * kFullCap + s;
                            a.nl[slot] = make_float4(dx, dy, dz, __uint_as_float(w));
                            if (NL == 2) a.nld[slot] = sdir[t];
                        }
                    }
                    ncnt += (uint32_t)__popcll(bal);
                }
                if (LJ && hit) {
                    if (MOLS && cmol[t] == mi) continue;
                    float sig2 = a.sig2, e24 = a.eps24, e4 = a.eps4;
                    if (!UNIFORM) {
                        const float2 pj = cprm[t];
                        const float s = 0.5f * (pi.x + pj.x);
                        const float e = pi.y * pj.y;
                        sig2 = s * s;
                        e24 = a.c24 * e;
                        e4 = a.c4 * e;
                    }
                    const float inv = __builtin_amdgcn_rcpf(r2);
                    const float s2 = sig2 * inv;
                    const float s6 = s2 * s2 * s2;
                    const float fac = e24 * (2.f * s6 * s6 - s6) * inv;
                    ax -= (double)(dx * fac);  // force on i = (r_i - r_j) fac
                    ay -= (double)(dy * fac);
                    az -= (double)(dz * fac);
                    if (ENERGY) {
                        // each pair is visited from both members: half of U and of the virial
                        u_acc += 0.5 * (double)(e4 * (s6 * s6 - s6));
                        w_acc += 0.5 * (double)(r2 * fac);
                        const float hf = 0.5f * fac;
                        sym_outer<double>(vir, dx, dy, dz, dx * hf, dy * hf, dz * hf);
                    }
                }
            }
        }
        if (VL && active && sub == 0) {
            a.vlc[p] = min(ncnt, (uint32_t)kSkinCap);
            if (ncnt > (uint32_t)kSkinCap) atomicOr(a.flags, FLAG_VL_OVERFLOW);
        } else if (NL && active && sub == 0) {
            a.nlc[p] = min(ncnt, (uint32_t)kFullCap);
            if (ncnt > (uint32_t)kFullCap) atomicOr(a.flags, FLAG_NL_OVERFLOW);
        }
        if (LJ || LIMIT) {
            for (int o = 1; o < lps; o <<= 1) {
                ax += __shfl_xor(ax, o);
                ay += __shfl_xor(ay, o);
                az += __shfl_xor(az, o);
            }
            if (active && sub == 0) {
                double x = a.fx[oi] + ax, y = a.fy[oi] + ay, z = a.fz[oi] + az;
                if (LIMIT) {
                    const double fm = a.limit * a.eps[oi] / a.sig[oi];
                    x = x < fm ? x : fm; x = x > -fm ? x : -fm;
                    y = y < fm ? y : fm; y = y > -fm ? y : -fm;
                    z = z < fm ? z : fm; z = z > -fm ? z : -fm;
                }
                a.fx[oi] = x;
                a.fy[oi] = y;
                a.fz[oi] = z;
                if (a.kick > 0.0) kick_site(a.vx, a.vy, a.vz, a.mass, oi, a.kick, x, y, z);
            }
        }
    }

    // in-range pair count (the same in every lane of a wave): one slot per workgroup,
    // accumulated across launches
    // (a single global counter would serialise thousands of same-address atomics; an
    // atomic add without return leaves the workgroup at once, where += waits for a load)
    if (lane == 0) red_cnt[wv] = npairs;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kWaves; k++) t += red_cnt[k];
        atomicAdd(&a.counters[b], t);
    }
    if (ENERGY) {
        u_acc = wave_sum(u_acc);
        w_acc = wave_sum(w_acc);
        if (lane == 0) {
            red[0][wv] = u_acc;
            red[1][wv] = w_acc;
        }
        __syncthreads();
        if (tid == 0) {
            double s0 = 0, s1 = 0;
            for (int k = 0; k < kWaves; k++) { s0 += red[0][k]; s1 += red[1][k]; }
            a.part[b] = s0;
            a.part[a.nblocks + b] = s1;
        }
        block_store_sym(a.vpart, a.nblocks, b, vir);
    }
}

// ---------------------------------------------------------------------------
struct AtmArgs {
    const float4* __restrict__ nl;
    const uint32_t* __restrict__ nlc;
    const uint16_t* __restrict__ nld;
    const uint32_t* __restrict__ sorted;
    const double* __restrict__ rx;
    const double* __restrict__ ry;
    const double* __restrict__ rz;
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    double* __restrict__ esite;        // per-site three-body energy share (optional)
    const float* __restrict__ lam;     // hybrid: exact-ATM weight per site
    const float* __restrict__ dlam;    // hybrid: d lambda / d axis coordinate per site
    double* __restrict__ part;
    double* __restrict__ vpart;        // [kSym][nblocks] virial tensor partials (energy mode)
    unsigned long long* __restrict__ counters;
    double* __restrict__ vx;           // second half-kick fused into the force write (kick > 0)
    double* __restrict__ vy;
    double* __restrict__ vz;
    const double* __restrict__ mass;
    // Verlet-list steps (SKIN): their buffers, read from device memory where they are used
    // (as kernel arguments they would stay in scalar registers through the whole kernel)
    const SkinArgs* __restrict__ sk;
    int par;                           // step parity: the next step's rebuild word is sk->req[par ^ 1]
    Grid g;
    int n, chunk, nblocks, hyb_axis;
    int h16;                           // 1e-2 <= rc <= 1e4: the FP16 rotation filter is exact enough
    int lj;                            // also the LJ force of the site's list (uniform sites, no clamp)
    float sig2, eps24;
    float rc2f, band, band3, nu;
    double rc2, kick;
};

constexpr int kAtmWaves = 4;
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// ATM: 1 physical, 2 reference semantics.  SKIN (physical, force pass, no hybrid): the
// site's in-range neighbours come from its Verlet list and the current positions; the
// kernel then writes the complete force (it is the step's only force kernel) and raises
// the next step's rebuild word when the next drift would leave the skin.
template <int ATM, bool ENERGY, bool HYB, bool SKIN = false>
__global__ __launch_bounds__(64 * kAtmWaves) __attribute__((amdgpu_waves_per_eu(8))) void k_atm(AtmArgs a) {
    static_assert(!SKIN || (ATM == 1 && !ENERGY && !HYB), "Verlet lists: physical ATM force pass only");
    __shared__ float4 wbuf[kAtmWaves][kFullCap];
    __shared__ float4 woff[SKIN ? 125 : 1];  // offset of each Verlet-list offset class, image code in .w
    __shared__ float2 wsq[kAtmWaves][kFullCap];  // |x_a|^2 and its reciprocal per neighbour
    __shared__ uint32_t qbuf[kAtmWaves][128];  // ring of 128 candidate pairs
    __shared__ uint8_t wdir[ATM == 2 ? kAtmWaves : 1][ATM == 2 ? kFullCap : 1];
    __shared__ float wlam[HYB ? kAtmWaves : 1][HYB ? kFullCap : 1];
    __shared__ double red[kAtmWaves];
    __shared__ unsigned long long red_cnt[kAtmWaves];
    __shared__ unsigned long long red_np[SKIN ? kAtmWaves : 1];

    const int lane = threadIdx.x & 63;
    const int wvu = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    const int lb = (b & 7) * a.chunk + (b >> 3);  // XCD-aware: contiguous site slabs per XCD
    const int p = lb * kAtmWaves + wvu;
    if (SKIN) {
        // the offsets of the 125 classes of the Verlet-list entries (k_skin_offsets)
        if (threadIdx.x < 125) woff[threadIdx.x] = a.sk->skin_off[threadIdx.x];
        __syncthreads();
    }
    double e_acc = 0.0;
    double ljx = 0.0, ljy = 0.0, ljz = 0.0;  // this lane's LJ pairs (a.lj)
    float vir[kSym] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // triplet virial share of this lane's visits
    unsigned long long nvisit = 0;  // accepted visits of this wave (the same in every lane)
    unsigned long long npair = 0;   // SKIN: in-range pairs j > i of this wave (the same in every lane)
    bool p_skip = false;
    if (p < a.n) {
        const uint32_t oi = a.sorted[p];
        int nn;
        if (SKIN) {
            // the in-range neighbours of the Verlet list at the current positions:
            // r_j - r_i = (q_j + o) - q_i with q the FP32 positions relative to the build-time
            // cell centres and o the entry's cell offset (the list pass's staging arithmetic,
            // so a fresh list gives the same vectors); within a few ulps of rc the FP64
            // distance decides, as in the list pass.  Both row halves are loaded at once, each
            // lane only within the list's length; slots past the length read site p.
            const SkinArgs* sk = a.sk;
            const float4* spos = sk->spos;
            const float4 qi = spos[p];
            const int nv = __builtin_amdgcn_readfirstlane((int)sk->vlc[p]);
            const uint32_t* vrow = sk->vl + (size_t)p * kSkinCap;
            const uint32_t e0 = lane < nv ? vrow[lane] : (uint32_t)p;
            const uint32_t e1 = 64 + lane < nv ? vrow[64 + lane] : (uint32_t)p;
            int cnt = 0;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (h == 1 && nv <= 64) break;  // wave-uniform: a list of at most 64 entries
                const int s = h * 64 + lane;
                const bool live = s < nv;
                const uint32_t e = live ? (h ? e1 : e0) : (uint32_t)p;
                const float4 q = spos[e & kSkinIdxMask];
                const float4 o = woff[e >> kSkinIdxBits];
                float4 d;
                d.x = (q.x + o.x) - qi.x;
                d.y = (q.y + o.y) - qi.y;
                d.z = (q.z + o.z) - qi.z;
                const uint32_t wj = __float_as_uint(q.w) | __float_as_uint(o.w);
                d.w = __uint_as_float(wj);
                const float r2 = live ? d.x * d.x + d.y * d.y + d.z * d.z : __builtin_huge_valf();
                bool in = r2 <= a.rc2f;
                const bool amb = fabsf(r2 - a.rc2f) <= a.band;
                if (ballot(amb) != 0ull && amb) in = !(exact_r2(a.rx, a.ry, a.rz, a.g, oi, wj) > a.rc2);
                const unsigned long long bal = ballot(in);
                npair += (unsigned long long)__popcll(ballot(in && (wj & kIdxMask) > oi));
                if (in) {
                    const int slot = cnt + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    if (slot < kFullCap) wbuf[wvu][slot] = d;
                }
                cnt += (int)__popcll(bal);
            }
            if (cnt > kFullCap && lane == 0) atomicOr(sk->flags, FLAG_NL_OVERFLOW);
            nn = min(cnt, kFullCap);
            wave_sync();
        } else {
            nn = __builtin_amdgcn_readfirstlane((int)a.nlc[p]);
        }
        float li = 1.f;
        if (HYB) li = a.lam[oi];
        const float4* row = a.nl + (size_t)p * kFullCap;
        float4 xa = make_float4(1.0e6f * (float)(lane + 1), 0.f, 0.f, 0.f);  // distinct far-away filler
        bool weighted = HYB ? li != 0.f : true;
        for (int s = lane; s < nn; s += 64) {
            const float4 v = SKIN ? wbuf[wvu][s] : row[s];
            if (!SKIN) wbuf[wvu][s] = v;
            const float X = v.x * v.x + v.y * v.y + v.z * v.z;
            const float iX = __builtin_amdgcn_rcpf(X);
            wsq[wvu][s] = make_float2(X, iX);
            if (!ENERGY && !HYB && a.lj) {
                // LJ 12-6 of the pair (the pair pass's expression): force on i = -x fac,
                // per-pair FP32 terms summed in FP64
                const float s2 = a.sig2 * iX;
                const float s6 = s2 * s2 * s2;
                const float fac = a.eps24 * (2.f * s6 * s6 - s6) * iX;
                ljx -= (double)(v.x * fac);
                ljy -= (double)(v.y * fac);
                ljz -= (double)(v.z * fac);
            }
            if (s == lane) xa = v;
            if (ATM == 2) wdir[wvu][s] = (uint8_t)(a.nld[(size_t)p * kFullCap + s] & 31u);
            if (HYB) {
                const float lj = a.lam[__float_as_uint(v.w) & kIdxMask];
                wlam[wvu][s] = lj;
                weighted = weighted || lj != 0.f;
            }
        }
        wave_sync();
        // hybrid: a site whose neighbourhood lies wholly outside the exact region has no weighted triplet
        if (HYB && ballot(weighted) == 0ull) p_skip = true;
        if (!p_skip) {
            uint32_t* ring = qbuf[wvu];
            // the visit energy is needed by the energy passes, the hybrid weight gradient
            // and the per-site energy output only
            const bool want_e = ENERGY || HYB || a.esite != nullptr;
            float fx = 0.f, fy = 0.f, fz = 0.f, ev = 0.f;
            // ring entries: a | b << 16 from the circulant filter, a | k << 16 (b = a + k mod 64)
            // from the rotation filter (wave-uniform: lists of at most 64 and a cutoff in the
            // FP16 filter's range)
            const bool rot = nn <= 64 && a.h16;
            // dense evaluation of 64 candidate pairs (x = r_j - r_i, y = r_k - r_i), `live`
            // lanes only.  The filters pass every pair with r_jk^2 <= rc^2 + band; the
            // cutoff is decided here: FP32 r_jk^2 against rc^2, and within the band the
            // FP64 distance with the reference's operation order (rare: wave-uniform
            // branch), so the triplet set is the reference's.
            auto force = [&](uint32_t e, bool live) {
                const int ia = (int)(e & 0xFFFFu);
                const int ib = rot ? (int)((e + (e >> 16)) & 63u) : (int)(e >> 16);
                const float4 xv = wbuf[wvu][ia], yv = wbuf[wvu][ib];
                const float zx = yv.x - xv.x, zy = yv.y - xv.y, zz = yv.z - xv.z;
                const float Z = zx * zx + zy * zy + zz * zz;
                bool in = live && Z <= a.rc2f;
                const float p2 = xv.x * zx + xv.y * zy + xv.z * zz;
                const bool amb = live && fabsf(Z - a.rc2f) <= a.band3;
                if (ballot(amb) != 0ull && amb)
                    in = !(exact_r2(a.rx, a.ry, a.rz, a.g, __float_as_uint(xv.w), __float_as_uint(yv.w)) > a.rc2);
                nvisit += (unsigned long long)__popcll(ballot(in));
                float wt = 1.f;
                if (HYB) wt = (li + wlam[wvu][ia] + wlam[wvu][ib]) * (1.f / 3.f);
                const float2 sx = wsq[wvu][ia], sy = wsq[wvu][ib];
                const float X = sx.x, Y = sy.x;
                const float p1 = X + p2, p3 = p2 + Z;  // y = x + z
                const float D = X * Y * Z;
                const float Pp = -(p1 * p2 * p3);
                const float rD = __builtin_amdgcn_rsqf(D);
                const float rD2 = rD * rD;
                const float D52 = rD2 * rD2 * rD;
                // rejected and idle lanes select zeros (their D may be 0 or garbage)
                float K = in ? 3.f * a.nu * D52 * wt : 0.f;
                if (ATM == 2) K = -K;  // reference: +grad V
                const float A = -(D + 5.f * Pp);
                const float t1 = p2 * p3, t23 = p1 * (p3 + p2);
                // with z = y - x the bracket is (A/X - t1 + t23) x + (A/Y - t1 - t23) y
                const float cx = K * (A * sx.y - t1 + t23), cy = K * (A * sy.y - t1 - t23);
                fx = fmaf(cy, yv.x, fmaf(cx, xv.x, fx));
                fy = fmaf(cy, yv.y, fmaf(cx, xv.y, fy));
                fz = fmaf(cy, yv.z, fmaf(cx, xv.z, fz));
                if (want_e) ev += in ? a.nu * D52 * (D + 3.f * Pp) : 0.f;  // unweighted V of this visit
                if (ENERGY) {
                    // (r_i - centroid) (x) F_i, r_i - centroid = -(x + y)/3; the three member
                    // visits sum to sum_m r_m (x) F_m.  Physical sign in both modes.
                    const float gx = cx * xv.x + cy * yv.x, gy = cx * xv.y + cy * yv.y, gz = cx * xv.z + cy * yv.z;
                    const float sc = ATM == 2 ? (1.f / 3.f) : (-1.f / 3.f);
                    sym_outer<float>(vir, sc * (xv.x + yv.x), sc * (xv.y + yv.y), sc * (xv.z + yv.z), gx, gy, gz);
                }
            };
            int head = 0, tail = 0;
            // ballot-compact this instruction's candidates into the ring; dense flush at 64
            auto push = [&](bool hit, uint32_t e) {
                const unsigned long long bal = ballot(hit);
                if (hit) {
                    const int pos = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    ring[(tail + pos) & 127] = e;
                }
                tail = __builtin_amdgcn_readfirstlane(tail + (int)__popcll(bal));
                if (tail - head >= 64) {
                    wave_sync();
                    force(ring[(head + lane) & 127], true);
                    head += 64;
                    wave_sync();
                }
            };
            const float lim = a.rc2f + a.band3;  // candidate bound of the circulant filter
            auto accept = [&](bool in, int ia, int ib) {
                if (ATM == 2 && in) {
                    const int da = wdir[wvu][ia], db = wdir[wvu][ib];
                    const bool same = (da == 13) && (db == 13);
                    const bool distinct = (da != 13) && (db != 13) && (da != db);
                    in = same || distinct;
                }
                if (HYB && in) in = (li + wlam[wvu][ia] + wlam[wvu][ib]) != 0.f;
                return in;
            };
            if (rot) {
                // filter in registers: lane a holds x_a; wave_rol moves lane a+1's value to
                // lane a, so after k rotations lane a holds y = x_{a+k mod 64}.
                // (x, y) as an FP16 pair, z in FP32; r_jk^2 from v_dot2_f32_f16 (exact FP16
                // products, FP32 sum).  Rounding x and y to FP16 (unit roundoff u = 2^-11)
                // moves r_jk^2 by at most 2|z|u(|x| + |y| + |z|) <= 6u rc^2 = 0.003 rc^2 near
                // the cutoff, inside the 1 % margin, so every pair within rc is a candidate.
                // Filler lanes convert to FP16 infinity: their r_jk^2 is inf or NaN, never a
                // candidate.  wave_rol has a source lane for every lane: plain DPP moves.
                const half2_t oxy = {(_Float16)xa.x, (_Float16)xa.y};
                uint32_t mxy = __builtin_bit_cast(uint32_t, oxy);
                float yz = xa.z;
                const float limh = a.rc2f * 1.01f + a.band3;
                auto step = [&](int k, bool valid) {
                    mxy = (uint32_t)__builtin_amdgcn_mov_dpp((int)mxy, 0x134, 0xF, 0xF, false);
                    yz = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(yz), 0x134, 0xF, 0xF, false));
                    const half2_t d = __builtin_bit_cast(half2_t, mxy) - oxy;
                    const float dz = yz - xa.z;
                    const float Z = __builtin_amdgcn_fdot2(d, d, dz * dz, false);
                    bool in = valid && Z <= limh;
                    if (ATM == 2 || HYB) in = accept(in, lane, (lane + k) & 63);
                    push(in, (uint32_t)lane | ((uint32_t)k << 16));
                };
#pragma clang loop unroll(disable)
                for (int k = 1; k < 31; k += 2) {
                    step(k, true);
                    step(k + 1, true);
                }
                step(31, true);
                step(32, lane < 32);  // the pairs (a, a + 32) once
            } else {
                // long lists (and cutoffs outside the FP16 filter's range): circulant pairs
                // (a, a+s+1 mod n) for s < H = (n-1)/2, plus s = H for a < n/2 when n is
                // even, read from LDS, r_jk^2 in FP32
                const int H = (nn - 1) >> 1;
                const int halfn = (nn & 1) ? 0 : (nn >> 1);
                for (int sft = 0; sft <= H; sft++) {
                    const int lims = sft < H ? nn : halfn;
                    for (int a0 = 0; a0 < lims; a0 += 64) {
                        const int ia = a0 + lane;
                        bool in = false;
                        int ib = 0;
                        if (ia < lims) {
                            const float4 xv = wbuf[wvu][ia];
                            ib = ia + sft + 1;
                            if (ib >= nn) ib -= nn;
                            const float4 yv = wbuf[wvu][ib];
                            const float zx = yv.x - xv.x, zy = yv.y - xv.y, zz = yv.z - xv.z;
                            const float Z = zx * zx + zy * zy + zz * zz;
                            in = accept(Z <= lim, ia, ib);
                        }
                        push(in, (uint32_t)ia | ((uint32_t)ib << 16));
                    }
                }
            }
            wave_sync();
            if (tail > head) {
                const bool live = lane < tail - head;
                force(live ? ring[(head + lane) & 127] : 0u, live);
            }
            double dx = wave_sum((double)fx + ljx), dy = wave_sum((double)fy + ljy), dz = wave_sum((double)fz + ljz);
            const double de = want_e ? wave_sum((double)ev) * (1.0 / 3.0) : 0.0;  // this site's third of its triplets
            if (lane == 0) {
                if (HYB) {  // -V grad(w): site i receives -(sum V/3) dl_i/dx_axis
                    const double t = -de * (double)a.dlam[oi];
                    if (a.hyb_axis == 0) dx += t; else if (a.hyb_axis == 1) dy += t; else dz += t;
                }
                // SKIN: the only force kernel of the step writes the force outright
                const double nfx = (SKIN ? 0.0 : a.fx[oi]) + dx, nfy = (SKIN ? 0.0 : a.fy[oi]) + dy,
                             nfz = (SKIN ? 0.0 : a.fz[oi]) + dz;
                a.fx[oi] = nfx;
                a.fy[oi] = nfy;
                a.fz[oi] = nfz;
                if (a.kick > 0.0) kick_site(a.vx, a.vy, a.vz, a.mass, oi, a.kick, nfx, nfy, nfz);
                if (SKIN) {
                    const SkinArgs* sk = a.sk;
                    if (skin_exceeded(a.vx, a.vy, a.vz, a.mass, oi, a.kick, sk->dt, nfx, nfy, nfz, sk->spos[p],
                                      sk->spb[p], sk->lim2))
                        atomicOr(sk->req + (a.par ^ 1), 1u);
                }
                if (a.esite) a.esite[oi] = HYB ? de * (double)li : de;
            }
            e_acc = HYB ? de * (double)li : de;
        }
    }
    // triplet visits: one slot per workgroup, accumulated across launches
    if (lane == 0) {
        red[wvu] = e_acc;
        red_cnt[wvu] = nvisit;
        if (SKIN) red_np[wvu] = npair;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kAtmWaves; k++) t += red_cnt[k];
        atomicAdd(&a.counters[b], t);  // no return: the workgroup does not wait on it
        if (SKIN) {
            unsigned long long tp = 0;
            for (int k = 0; k < kAtmWaves; k++) tp += red_np[k];
            atomicAdd(&a.sk->pair_counters[b], tp);
        }
        if (ENERGY) {
            double s = 0;
            for (int k = 0; k < kAtmWaves; k++) s += red[k];
            a.part[b] = s;
        }
    }
    if (ENERGY) block_store_sym(a.vpart, a.nblocks, b, vir);
}

// The FP32 offset and image code of each of the 125 Verlet-list offset classes: the list
// pass's nb_cell arithmetic per axis, computed once per run.
__global__ void k_skin_offsets(Grid g, float4* __restrict__ out) {
    const int t = threadIdx.x;
    if (t >= 125) return;
    const int cx = t % 5, cy = (t / 5) % 5, cz = t / 25;
    auto img = [](int c) { return c == 3 ? 2u : (c == 4 ? 1u : 0u); };
    out[t] = make_float4(axis_offset(cx, g.nx, g.cs, g.size[0]), axis_offset(cy, g.ny, g.cs, g.size[1]),
                         axis_offset(cz, g.nz, g.cs, g.size[2]),
                         __uint_as_float((img(cx) | (img(cy) << 2) | (img(cz) << 4)) << kIdxBits));
}

// ---------------------------------------------------------------------------
// Verlet-list step of the learned and hybrid three-body kinds: each site's in-range
// neighbours at the current positions, filtered from its Verlet list as the Verlet-list ATM
// pass does (the same FP32 vectors and the FP64 decision near the cutoff), compacted into
// the neighbour-list rows the three-body kernels read (r_j - r_i with the site word, the cell
// direction and rank), and the LJ 12-6 force of those pairs written as the site's force:
// the first force term of the step (uniform sites, no clamp).  One 16-lane DPP row per site,
// four per wave.
struct RefreshArgs {
    const SkinArgs* __restrict__ sk;
    const uint32_t* __restrict__ sorted;
    const double* __restrict__ rx;
    const double* __restrict__ ry;
    const double* __restrict__ rz;
    float4* __restrict__ nl;
    uint32_t* __restrict__ nlc;
    uint16_t* __restrict__ nld;
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    Grid g;
    int n, chunk;
    float rc2f, band, sig2, eps24;
    double rc2;
};

constexpr int kLjWaves = 4;

// The sum over the L-lane group (L = 4, 8 or 16) in every lane of the group, a fixed order:
// quads by quad_perm, then the two quads of a half-row by row_half_mirror (L = 8) or the
// four quads of a row by row_ror (L = 16).
template <int L>
__device__ __forceinline__ double group_sum(double v) {
    v += dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
    if (L == 4) {
    } else if (L == 8) {
        v += dpp_f64<0x141>(v);  // row_half_mirror
    } else {
        v += dpp_f64<0x124>(v);  // row_ror:4
        v += dpp_f64<0x128>(v);  // row_ror:8
    }
    return v;
}

constexpr int kRefWaves = 4;
constexpr int kRefSites = 4 * kRefWaves;  // sites per workgroup

// 16 lanes per site: narrower groups (the k_skin_lj rule) break the compacted row writes into
// pieces of 64 B or less across many rows (256,000 atoms: 298 us at 4 lanes against 159).
__global__ __launch_bounds__(64 * kRefWaves) void k_skin_refresh(RefreshArgs a) {
    constexpr int L = 16, R = 64 / L;  // lanes per site, sites per wave
    __shared__ float4 woff[125];
    const int lane = threadIdx.x & 63, sub = lane & (L - 1);
    const int wvu = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const SkinArgs* sk = a.sk;
    if (threadIdx.x < 125) woff[threadIdx.x] = sk->skin_off[threadIdx.x];
    __syncthreads();
    const int b = blockIdx.x;
    // XCD-aware site slabs; the R lane groups of a wave take consecutive sites
    const int p = (((b & 7) * a.chunk + (b >> 3)) * kRefWaves + wvu) * R + lane / L;
    const bool act = p < a.n;
    const uint32_t self = act ? (uint32_t)p : 0u;  // dead groups read slot 0, class 0
    const unsigned long long rmask = ((1ull << L) - 1ull) << (lane & (64 - L));  // my group's lanes
    const float4* spos = sk->spos;
    const uint32_t oi = act ? a.sorted[p] : 0u;
    const float4 qi = spos[self];
    const int nv = act ? (int)sk->vlc[p] : 0;
    const uint32_t* vrow = sk->vl + (size_t)self * kSkinCap;
    const uint16_t* drow = sk->vld + (size_t)self * kSkinCap;
    float4* out = a.nl + (size_t)self * kFullCap;
    uint16_t* dout = a.nld + (size_t)self * kFullCap;
    double ljx = 0.0, ljy = 0.0, ljz = 0.0;
    unsigned long long npair = 0;
    int cnt = 0;  // in-range entries of my group so far (the same in the group's lanes)
    for (int t0 = 0; ballot(t0 < nv) != 0ull; t0 += L) {  // wave-uniform trip count
        const int s = t0 + sub;
        const bool live = s < nv;
        const uint32_t e = live ? vrow[s] : self;
        const float4 q = spos[e & kSkinIdxMask];
        const float4 o = woff[e >> kSkinIdxBits];
        float4 d;
        d.x = (q.x + o.x) - qi.x;
        d.y = (q.y + o.y) - qi.y;
        d.z = (q.z + o.z) - qi.z;
        const uint32_t wj = __float_as_uint(q.w) | __float_as_uint(o.w);
        d.w = __uint_as_float(wj);
        const float r2 = live ? d.x * d.x + d.y * d.y + d.z * d.z : __builtin_huge_valf();
        bool in = r2 <= a.rc2f;
        const bool amb = fabsf(r2 - a.rc2f) <= a.band;
        if (ballot(amb) != 0ull && amb) in = !(exact_r2(a.rx, a.ry, a.rz, a.g, oi, wj) > a.rc2);
        const unsigned long long bal = ballot(in) & rmask;
        npair += (unsigned long long)__popcll(ballot(in && (wj & kIdxMask) > oi));
        if (in) {
            // compacted in entry order, as the list pass orders its rows
            const int slot = cnt + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (slot < kFullCap) {
                out[slot] = d;
                dout[slot] = drow[s];
            }
            // LJ 12-6 of the pair (the pair pass's expression): force on i = -x fac
            const float iX = __builtin_amdgcn_rcpf(r2);
            const float s2 = a.sig2 * iX;
            const float s6 = s2 * s2 * s2;
            const float fac = a.eps24 * (2.f * s6 * s6 - s6) * iX;
            ljx -= (double)(d.x * fac);
            ljy -= (double)(d.y * fac);
            ljz -= (double)(d.z * fac);
        }
        cnt += (int)__popcll(bal);
    }
    ljx = group_sum<L>(ljx);
    ljy = group_sum<L>(ljy);
    ljz = group_sum<L>(ljz);
    if (act && sub == 0) {
        a.nlc[p] = (uint32_t)min(cnt, kFullCap);
        if (cnt > kFullCap) atomicOr(sk->flags, FLAG_NL_OVERFLOW);
        a.fx[oi] = ljx;
        a.fy[oi] = ljy;
        a.fz[oi] = ljz;
    }
    if (lane == 0 && npair) atomicAdd(&sk->pair_counters[b], npair);
}

// ---------------------------------------------------------------------------
// Verlet-list step without a three-body term (LJ 12-6 of uniform sites, no clamp): each
// site's in-range partners filtered from its Verlet list at the current positions (the FP32
// vectors and the FP64 decision near the cutoff of the Verlet-list ATM pass), their LJ force
// added to the site's force (zero, or the step's bonded terms, which ran first: the order of
// the per-step run), then Step1 and the next step's rebuild decision.  Lists of multi-site
// molecules hold sites of other molecules only (the exclusion of LJ12_6.cpp:19 taken at the
// build).  A group of L lanes per site (skin_lanes), 64 / L sites per wave; the group sums
// its FP64 partial forces by DPP (group_sum), in a fixed order.
struct SkinLjArgs {
    const SkinArgs* __restrict__ sk;
    const uint32_t* __restrict__ sorted;
    const double* __restrict__ rx;
    const double* __restrict__ ry;
    const double* __restrict__ rz;
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    double* __restrict__ vx;
    double* __restrict__ vy;
    double* __restrict__ vz;
    const double* __restrict__ mass;
    Grid g;
    int n, chunk, add, par;
    float rc2f, band, sig2, eps24;
    double rc2, kick;
};
template <int L>
__global__ __launch_bounds__(64 * kLjWaves) void k_skin_lj(SkinLjArgs a) {
    static_assert(L == 4 || L == 8 || L == 16, "lanes per site");
    constexpr int R = 64 / L;  // sites per wave
    __shared__ float4 woff[125];
    __shared__ unsigned long long red_np[kLjWaves];
    const int lane = threadIdx.x & 63, sub = lane & (L - 1);
    const int wvu = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const SkinArgs* sk = a.sk;
    if (threadIdx.x < 125) woff[threadIdx.x] = sk->skin_off[threadIdx.x];
    __syncthreads();
    const int b = blockIdx.x;
    // XCD-aware site slabs; the R lane groups of a wave take consecutive sites
    const int p = (((b & 7) * a.chunk + (b >> 3)) * kLjWaves + wvu) * R + lane / L;
    const bool act = p < a.n;
    const uint32_t self = act ? (uint32_t)p : 0u;  // dead lanes read slot 0, class 0
    const float4* spos = sk->spos;
    const uint32_t oi = act ? a.sorted[p] : 0u;
    const float4 qi = spos[self];
    const int nv = act ? (int)sk->vlc[p] : 0;
    const uint32_t* vrow = sk->vl + (size_t)self * kSkinCap;
    double ljx = 0.0, ljy = 0.0, ljz = 0.0;
    unsigned long long npair = 0;
    for (int t0 = 0; ballot(t0 < nv) != 0ull; t0 += L) {  // wave-uniform trip count
        const int s = t0 + sub;
        const bool live = s < nv;
        const uint32_t e = live ? vrow[s] : self;
        const float4 q = spos[e & kSkinIdxMask];
        const float4 o = woff[e >> kSkinIdxBits];
        const float dx = (q.x + o.x) - qi.x, dy = (q.y + o.y) - qi.y, dz = (q.z + o.z) - qi.z;
        const uint32_t wj = __float_as_uint(q.w) | __float_as_uint(o.w);
        const float r2 = live ? dx * dx + dy * dy + dz * dz : __builtin_huge_valf();
        bool in = r2 <= a.rc2f;
        const bool amb = fabsf(r2 - a.rc2f) <= a.band;
        if (ballot(amb) != 0ull && amb) in = !(exact_r2(a.rx, a.ry, a.rz, a.g, oi, wj) > a.rc2);
        npair += (unsigned long long)__popcll(ballot(in && (wj & kIdxMask) > oi));
        if (in) {
            // the pair pass's LJ 12-6 expression: force on i = -(r_j - r_i) fac
            const float iX = __builtin_amdgcn_rcpf(r2);
            const float s2 = a.sig2 * iX;
            const float s6 = s2 * s2 * s2;
            const float fac = a.eps24 * (2.f * s6 * s6 - s6) * iX;
            ljx -= (double)(dx * fac);
            ljy -= (double)(dy * fac);
            ljz -= (double)(dz * fac);
        }
    }
    ljx = group_sum<L>(ljx);
    ljy = group_sum<L>(ljy);
    ljz = group_sum<L>(ljz);
    if (act && sub == 0) {
        const double nfx = (a.add ? a.fx[oi] : 0.0) + ljx, nfy = (a.add ? a.fy[oi] : 0.0) + ljy,
                     nfz = (a.add ? a.fz[oi] : 0.0) + ljz;
        a.fx[oi] = nfx;
        a.fy[oi] = nfy;
        a.fz[oi] = nfz;
        kick_site(a.vx, a.vy, a.vz, a.mass, oi, a.kick, nfx, nfy, nfz);
        if (skin_exceeded(a.vx, a.vy, a.vz, a.mass, oi, a.kick, sk->dt, nfx, nfy, nfz, qi, sk->spb[p], sk->lim2))
            atomicOr(sk->req + (a.par ^ 1), 1u);
    }
    if (lane == 0) red_np[wvu] = npair;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tp = 0;
        for (int k = 0; k < kLjWaves; k++) tp += red_np[k];
        atomicAdd(&sk->pair_counters[b], tp);
    }
}

template <bool U, bool M, bool LJ, int NL, bool E, bool LIM>
void go(komd_ctx* c, const CellArgs& a) {
    k_cell<U, M, LJ, NL, E, LIM><<<a.nblocks, kCellThreads, 0, c->stream>>>(a);
}

template <bool U, bool M, bool LIM>
void dispatch_lim(komd_ctx* c, const CellArgs& a, bool lj, int nl, bool energy) {
    if (nl == 0) {
        if (energy) go<U, M, true, 0, true, LIM>(c, a);
        else go<U, M, true, 0, false, LIM>(c, a);
    } else if (nl == 1) {
        if (lj && energy) go<U, M, true, 1, true, LIM>(c, a);
        else if (lj) go<U, M, true, 1, false, LIM>(c, a);
        else go<U, M, false, 1, false, false>(c, a);
    } else {
        if (lj && energy) go<U, M, true, 2, true, LIM>(c, a);
        else if (lj) go<U, M, true, 2, false, LIM>(c, a);
        else go<U, M, false, 2, false, false>(c, a);
    }
}

template <bool U, bool M>
hipError_t dispatch(komd_ctx* c, const CellArgs& a, bool lj, int nl, bool energy, bool lim) {
    if (lim) dispatch_lim<U, M, true>(c, a, lj, nl, energy);
    else dispatch_lim<U, M, false>(c, a, lj, nl, energy);
    return hipGetLastError();
}

}  // namespace

// Workgroups of the pair pass per XCD (blocks = 8 x that).  cx = 0: the plain mapping, with
// `cell_split` workgroups per cell on small grids.  Otherwise XCD x owns cx consecutive cells;
// with S < n < 2S cells for S resident slots, 2S - n of them get one workgroup and the rest
// two (a partial second round of half-cells instead of whole ones); otherwise one each.
// (Measured at 32k: the last n - S cells in S / (n - S) short pieces each, after S whole
// cells, took 24.9 us against 23.1: the pieces repeat the neighbourhood staging.)
static void cell_jobs(const komd_ctx* c, int* cx, int* chunk) {
    const int C = c->grid.ncell;
    if (c->cell_split > 1 || c->xcd_slots <= 0) {
        *cx = 0;
        *chunk = (C * c->cell_split + 7) / 8;
        return;
    }
    *cx = (C + 7) / 8;
    int jmax = 0;
    for (int x = 0; x < 8; x++) {
        const int n = std::max(0, std::min(*cx, C - x * *cx));
        const int w = (n <= c->xcd_slots || n >= 2 * c->xcd_slots) ? n : 2 * c->xcd_slots - n;
        jmax = std::max(jmax, w + 2 * (n - w));
    }
    *chunk = jmax;
}

int cell_blocks(komd_ctx* c) {
    int cx, chunk;
    cell_jobs(c, &cx, &chunk);
    return 8 * chunk;
}
int atm_blocks(komd_ctx* c) { return 8 * (((int)((c->n + kAtmWaves - 1) / kAtmWaves) + 7) / 8); }

hipError_t launch_cell(komd_ctx* c, bool lj, int nl, bool energy, bool limit, double scale) {
    auto& d = c->d;
    CellArgs a{};
    a.spos = d.spos; a.sprm = d.sprm; a.sid = d.sid; a.start = d.start;
    a.rx = d.rx; a.ry = d.ry; a.rz = d.rz; a.eps = d.eps; a.sig = d.sig;
    a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.nl = d.nl; a.nlc = d.nlc; a.nld = d.nld;
    a.part = d.red; a.vpart = d.vpart; a.flags = d.flags; a.counters = d.cnt_pair;
    a.g = c->grid;
    a.split = c->cell_split;
    cell_jobs(c, &a.cx, &a.chunk);
    a.slots = c->xcd_slots;
    a.nblocks = 8 * a.chunk;
    a.rc2 = c->prm.cutoff * c->prm.cutoff;
    a.rc2f = (float)a.rc2;
    a.band = 4e-6f * a.rc2f;
    a.sig2 = c->sig_u * c->sig_u;
    // LJ's 24 eps and 4 eps with the pair-force scale folded in (1 for LJ 12-6, the
    // mean-field factor for ATM2B): no extra operation per pair
    a.eps24 = (float)(24.0 * (double)c->eps_u * scale);
    a.eps4 = (float)(4.0 * (double)c->eps_u * scale);
    a.c24 = (float)(24.0 * scale);
    a.c4 = (float)(4.0 * scale);
    a.limit = c->prm.limit_factor;
    a.vx = d.vx; a.vy = d.vy; a.vz = d.vz; a.mass = d.mass;
    a.kick = (lj && !energy && c->kick_in == KICK_IN_CELL) ? 0.5 * c->prm.delta_t : 0.0;
    if (nl && (!d.nl || !d.nlc || (nl == 2 && !d.nld))) return hipErrorInvalidValue;
    const bool lim = limit && c->prm.limit_factor > 0;
    ScopedTimer tm(c, nl ? (lj ? "cell_lj_nl" : "cell_nl") : "cell_lj");
    const bool U = c->uniform, M = c->multisite;
    if (U && !M) return dispatch<true, false>(c, a, lj, nl, energy, lim);
    if (U && M) return dispatch<true, true>(c, a, lj, nl, energy, lim);
    if (!U && !M) return dispatch<false, false>(c, a, lj, nl, energy, lim);
    return dispatch<false, true>(c, a, lj, nl, energy, lim);
}

hipError_t launch_skin_offsets(komd_ctx* c) {
    k_skin_offsets<<<1, 128, 0, c->stream>>>(c->grid, c->d.skin_off);
    return hipGetLastError();
}

size_t skin_args_bytes() { return sizeof(SkinArgs); }

// The Verlet-list force pass's buffer table and constants, written before a skinned run.
hipError_t upload_skin_args(komd_ctx* c) {
    auto& d = c->d;
    if (!d.vl || !d.vlc || !d.spb || !d.skin_off || !d.skin_req || !d.skin_args || c->n > (int64_t)kSkinIdxMask)
        return hipErrorInvalidValue;
    SkinArgs h{};
    h.vl = d.vl; h.vlc = d.vlc; h.vld = d.vld; h.spos = d.spos; h.spb = d.spb; h.skin_off = d.skin_off;
    h.req = d.skin_req;
    h.pair_counters = d.cnt_trip + atm_blocks(c);
    h.flags = d.flags;
    h.dt = c->prm.delta_t;
    // rebuild once a site would move further than half the skin, less a margin for the
    // FP32 distance tests of the list pass
    const double lim = 0.5 * c->skin - 1e-5 * (c->prm.cutoff + c->skin);
    h.lim2 = lim > 0 ? lim * lim : 0.0;
    return h2d(c, d.skin_args, &h, sizeof(h));
}

// The Verlet-list pass of a skinned step (parity par): every site within cutoff + skin,
// run only when the step's rebuild word is set.
hipError_t launch_skin_list(komd_ctx* c, int par, bool dirs) {
    auto& d = c->d;
    CellArgs a{};
    a.vld = d.vld;
    a.spos = d.spos; a.sprm = d.sprm; a.sid = d.sid; a.start = d.start;
    a.rx = d.rx; a.ry = d.ry; a.rz = d.rz; a.eps = d.eps; a.sig = d.sig;
    a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.vl = d.vl; a.vlc = d.vlc; a.req = d.skin_req + par;
    a.part = d.red; a.vpart = d.vpart; a.flags = d.flags; a.counters = d.cnt_pair;
    a.g = c->grid;
    a.split = c->cell_split;
    cell_jobs(c, &a.cx, &a.chunk);
    a.slots = c->xcd_slots;
    a.nblocks = 8 * a.chunk;
    const double rl = c->prm.cutoff + c->skin;
    a.rl2f = (float)(rl * rl);
    a.rc2 = c->prm.cutoff * c->prm.cutoff;
    a.rc2f = (float)a.rc2;
    if (!d.vl || !d.vlc || !d.skin_req || !c->uniform || c->n > (int64_t)kSkinIdxMask ||
        (c->multisite && (dirs || !d.sid)))
        return hipErrorInvalidValue;
    ScopedTimer tm(c, "cell_nl");
    if (c->multisite) {
        k_cell<true, true, false, 3, false, false><<<a.nblocks, kCellThreads, 0, c->stream>>>(a);
    } else if (dirs) {
        if (!d.vld) return hipErrorInvalidValue;
        k_cell<true, false, false, 4, false, false><<<a.nblocks, kCellThreads, 0, c->stream>>>(a);
    } else {
        k_cell<true, false, false, 3, false, false><<<a.nblocks, kCellThreads, 0, c->stream>>>(a);
    }
    return hipGetLastError();
}

// Lanes per site of the Verlet-list LJ walk (k_skin_lj): the narrowest of 4, 8
// and 16 that still gives about eight waves per SIMD (n L / 64 >= 8 x 1,024).  Narrow groups
// share a wave's trip count and its per-site tail among more sites (262,144 chain sites:
// 52.0 us at 4 lanes, 53.4 at 8, 68.8 at 16); small systems need the waves (4,000 atoms:
// 0.0222 ms per step at 16 lanes, 0.0245 at 8).
int skin_lanes(int64_t n) {
    for (int L = 4; L < 16; L *= 2)
        if (n * L >= 8 * 1024 * 64) return L;
    return 16;
}

// LJ from the Verlet lists for a skinned step of parity par without a three-body term; add:
// the step's bonded forces are already in f (else f is written outright).
hipError_t launch_skin_lj(komd_ctx* c, int par, bool add) {
    auto& d = c->d;
    if (!d.skin_args || !c->uniform || c->n > (int64_t)kSkinIdxMask) return hipErrorInvalidValue;
    SkinLjArgs a{};
    a.sk = d.skin_args;
    a.sorted = d.sorted; a.rx = d.rx; a.ry = d.ry; a.rz = d.rz;
    a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.vx = d.vx; a.vy = d.vy; a.vz = d.vz; a.mass = d.mass;
    a.g = c->grid;
    a.n = (int)c->n;
    a.add = add ? 1 : 0;
    a.par = par;
    // at most atm_blocks workgroups: the pair-count slots of the Verlet-list passes
    const int lanes = skin_lanes(c->n);
    const int sites = kLjWaves * 64 / lanes;  // per workgroup
    const int blocks = 8 * ((((int)c->n + sites - 1) / sites + 7) / 8);
    a.chunk = blocks / 8;
    a.rc2 = c->prm.cutoff * c->prm.cutoff;
    a.rc2f = (float)a.rc2;
    a.band = 4e-6f * a.rc2f;
    a.sig2 = c->sig_u * c->sig_u;
    a.eps24 = 24.f * c->eps_u;
    a.kick = 0.5 * c->prm.delta_t;
    ScopedTimer tm(c, "skin_lj");
    if (lanes == 4)
        k_skin_lj<4><<<blocks, 64 * kLjWaves, 0, c->stream>>>(a);
    else if (lanes == 8)
        k_skin_lj<8><<<blocks, 64 * kLjWaves, 0, c->stream>>>(a);
    else
        k_skin_lj<16><<<blocks, 64 * kLjWaves, 0, c->stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_skin_refresh(komd_ctx* c) {
    auto& d = c->d;
    if (!d.skin_args || !d.vld || !d.nl || !d.nlc || !d.nld || !c->uniform || c->multisite) return hipErrorInvalidValue;
    RefreshArgs a{};
    a.sk = d.skin_args;
    a.sorted = d.sorted; a.rx = d.rx; a.ry = d.ry; a.rz = d.rz;
    a.nl = d.nl; a.nlc = d.nlc; a.nld = d.nld;
    a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.g = c->grid;
    a.n = (int)c->n;
    // at most atm_blocks workgroups: the pair-count slots of the Verlet-list passes
    const int blocks = 8 * ((((int)c->n + kRefSites - 1) / kRefSites + 7) / 8);
    a.chunk = blocks / 8;
    a.rc2 = c->prm.cutoff * c->prm.cutoff;
    a.rc2f = (float)a.rc2;
    a.band = 4e-6f * a.rc2f;
    a.sig2 = c->sig_u * c->sig_u;
    a.eps24 = 24.f * c->eps_u;
    ScopedTimer tm(c, "skin_refresh");
    k_skin_refresh<<<blocks, 64 * kRefWaves, 0, c->stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_atm(komd_ctx* c, int atm, bool energy, const float* lam, const float* dlam, double* esite, bool lj,
                      int skin_par) {
    auto& d = c->d;
    AtmArgs a{};
    a.nl = d.nl; a.nlc = d.nlc; a.nld = d.nld; a.sorted = d.sorted;
    a.rx = d.rx; a.ry = d.ry; a.rz = d.rz; a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.esite = esite; a.lam = lam; a.dlam = dlam; a.part = d.red; a.vpart = d.vpart; a.counters = d.cnt_trip;
    a.g = c->grid;
    a.n = (int)c->n;
    a.nblocks = atm_blocks(c);
    a.chunk = a.nblocks / 8;
    a.hyb_axis = c->hyb_axis;
    a.rc2 = c->prm.cutoff * c->prm.cutoff;
    a.rc2f = (float)a.rc2;
    a.band = 4e-6f * a.rc2f;
    a.band3 = 1e-5f * a.rc2f;
    a.h16 = c->prm.cutoff >= 1e-2 && c->prm.cutoff <= 1e4;
    a.lj = lj && !energy && !lam;
    a.sig2 = c->sig_u * c->sig_u;
    a.eps24 = 24.f * c->eps_u;
    a.nu = (float)c->prm.nu;
    a.vx = d.vx; a.vy = d.vy; a.vz = d.vz; a.mass = d.mass;
    a.kick = (!energy && c->kick_in == KICK_IN_ATM) ? 0.5 * c->prm.delta_t : 0.0;
    ScopedTimer tm(c, "atm");
    const bool hyb = lam != nullptr;
    if (skin_par >= 0) {
        // Verlet-list step: physical ATM with LJ, NVE (fused half-kick), no hybrid; the
        // buffers in d.skin_args (upload_skin_args)
        if (atm != 1 || energy || hyb || !a.lj || !(a.kick > 0.0) || !d.skin_args) return hipErrorInvalidValue;
        a.sk = d.skin_args;
        a.par = skin_par;
        k_atm<1, false, false, true><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
    } else if (hyb) {
        if (energy) k_atm<1, true, true><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
        else k_atm<1, false, true><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
    } else if (atm == 2) {
        if (energy) k_atm<2, true, false><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
        else k_atm<2, false, false><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
    } else {
        if (energy) k_atm<1, true, false><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
        else k_atm<1, false, false><<<a.nblocks, 64 * kAtmWaves, 0, c->stream>>>(a);
    }
    return hipGetLastError();
}

}  // namespace komd
