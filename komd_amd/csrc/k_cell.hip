// Cell-tile force kernel: LJ 12-6, force clamp and Axilrod-Teller-Muto in one
// pass -- gfx950.
//
// One 256-thread workgroup per domain cell ("home cell"):
//  1. the 27 neighbour cells' sites are staged in LDS as FP32 positions
//     relative to the home cell centre (periodic images folded into the
//     staging offset);
//  2. every home site is served by 4-16 lanes striding over the staged
//     candidates: FP32 distance test (a candidate within a few ulps of rc^2
//     is decided in FP64 with the reference's operation order), LJ force
//     summed per lane in FP64 and combined with shuffles, then clamped
//     (Limit) -- and, when the three-body term is on, the site's full
//     neighbour list (staging indices) is appended in LDS;
//  3. three-body pass: each wave takes home sites in turn, copies the site's
//     neighbour vectors into a per-wave LDS buffer and deals the n(n-1)/2
//     neighbour pairs to its 64 lanes (circulant order, shift s and pair
//     (a, a+s mod n)).  Every site accumulates only its own three-body force,
//     so the triplet loop has no atomics and no scatter; the per-lane FP32
//     sums are reduced across the wave in FP64 and added to the site's force.
// Each triplet is visited by its three members; the energy share of a visit
// is V/3.
//
// LJ equals the reference's C08 cell-pair walk (LinkedCells.cpp:164-224), in
// which every unshifted member receives the pair force once (LJ12_6.cpp:15-43);
// the clamp is Limit_Force (Limit.cpp:19-25), applied after the two-body terms
// and before the three-body term as main.cpp:57-66 registers them.
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
    double* __restrict__ esite;        // per-site three-body energy share (optional)
    const float* __restrict__ lam;     // hybrid: exact-ATM weight per site
    const float* __restrict__ dlam;    // hybrid: d lambda / d axis coordinate per site
    double* __restrict__ part;
    uint32_t* __restrict__ flags;
    unsigned long long* __restrict__ counters;
    Grid g;
    int chunk, nblocks, hyb_axis;
    float rc2f, band, band3, sig2, eps24, eps4, nu;
    double rc2, limit;
};

__device__ __forceinline__ double image_shift(uint32_t code, double L) {
    return code == 1u ? L : (code == 2u ? -L : 0.0);
}

// |(r_a + s_a) - (r_b + s_b)|^2 in FP64, reference operation order, no contraction.
__device__ __forceinline__ double exact_r2(const CellArgs& a, uint32_t wa, uint32_t wb) {
    const uint32_t i = wa & kIdxMask, j = wb & kIdxMask, ca = wa >> kIdxBits, cb = wb >> kIdxBits;
    const double xa = __dadd_rn(a.rx[i], image_shift(ca & 3u, a.g.size[0]));
    const double ya = __dadd_rn(a.ry[i], image_shift((ca >> 2) & 3u, a.g.size[1]));
    const double za = __dadd_rn(a.rz[i], image_shift((ca >> 4) & 3u, a.g.size[2]));
    const double xb = __dadd_rn(a.rx[j], image_shift(cb & 3u, a.g.size[0]));
    const double yb = __dadd_rn(a.ry[j], image_shift((cb >> 2) & 3u, a.g.size[1]));
    const double zb = __dadd_rn(a.rz[j], image_shift((cb >> 4) & 3u, a.g.size[2]));
    const double dx = __dsub_rn(xa, xb), dy = __dsub_rn(ya, yb), dz = __dsub_rn(za, zb);
    double s = __dmul_rn(dx, dx);
    s = __dadd_rn(s, __dmul_rn(dy, dy));
    s = __dadd_rn(s, __dmul_rn(dz, dz));
    return s;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// LDS writes of this wave visible to this wave's later reads.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ATM: 0 off, 1 physical, 2 reference semantics.
template <bool UNIFORM, bool MOLS, bool LJ, int ATM, bool ENERGY, bool HYB, bool LIMIT>
__global__ __launch_bounds__(kCellThreads) void k_cell(CellArgs a) {
    constexpr int kStage = ATM ? kMaxM : kChunkLJ;
    constexpr int kWaves = kCellThreads / 64;
    __shared__ float4 pos[kStage];
    __shared__ uint8_t sdir[ATM == 2 ? kMaxM : 1];
    __shared__ float slam[HYB ? kMaxM : 1];
    __shared__ float2 cprm[UNIFORM ? 1 : kStage];
    __shared__ uint32_t cmol[MOLS ? kStage : 1];
    __shared__ uint16_t nbl[ATM ? kMaxOwn : 1][ATM ? kFullCap : 1];
    __shared__ float4 wbuf[ATM ? kWaves : 1][ATM ? kFullCap : 1];
    __shared__ uint8_t wdir[ATM == 2 ? kWaves : 1][ATM == 2 ? kFullCap : 1];
    __shared__ float wlam[HYB ? kWaves : 1][HYB ? kFullCap : 1];
    __shared__ uint32_t hcnt[kMaxOwn];
    __shared__ uint32_t pcnt[kMaxOwn];
    __shared__ uint8_t own_of[kMaxOwn];    // three-body pass: home sites grouped by wave
    __shared__ uint8_t own_cnt[kCellThreads / 64];
    __shared__ int nb_start[27];
    __shared__ int nb_pref[28];
    __shared__ float nb_off[27][3];
    __shared__ uint32_t nb_code[27];
    __shared__ double red[3][kWaves];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int b = blockIdx.x;
    const int cell = (b & 7) * a.chunk + (b >> 3);  // XCD-aware: each XCD gets a contiguous slab
    double u_acc = 0.0, w_acc = 0.0, e_acc = 0.0;
    unsigned long long npairs = 0, nvisit = 0;
    const bool live = cell < a.g.ncell;
    int m = 0, hs = 0, nh = 0;
    if (live) {
        const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
        const int hx = cell % nx, hy = (cell / nx) % ny, hz = cell / (nx * ny);
        if (tid < 27) {
            const int dx = tid % 3 - 1, dy = (tid / 3) % 3 - 1, dz = tid / 9 - 1;
            int x = hx + dx, y = hy + dy, z = hz + dz;
            uint32_t cx = 0, cy = 0, cz = 0;
            double sx = 0, sy = 0, sz = 0;
            if (x < 0) { x += nx; cx = 2; sx = -a.g.size[0]; } else if (x >= nx) { x -= nx; cx = 1; sx = a.g.size[0]; }
            if (y < 0) { y += ny; cy = 2; sy = -a.g.size[1]; } else if (y >= ny) { y -= ny; cy = 1; sy = a.g.size[1]; }
            if (z < 0) { z += nz; cz = 2; sz = -a.g.size[2]; } else if (z >= nz) { z -= nz; cz = 1; sz = a.g.size[2]; }
            const int nc = x + nx * (y + ny * z);
            nb_start[tid] = (int)a.start[nc];
            nb_pref[tid + 1] = (int)(a.start[nc + 1] - a.start[nc]);
            nb_off[tid][0] = (float)((double)(x - hx) * a.g.cs + sx);
            nb_off[tid][1] = (float)((double)(y - hy) * a.g.cs + sy);
            nb_off[tid][2] = (float)((double)(z - hz) * a.g.cs + sz);
            nb_code[tid] = (cx | (cy << 2) | (cz << 4)) << kIdxBits;
        }
        __syncthreads();
        if (tid == 0) {
            nb_pref[0] = 0;
            for (int k = 0; k < 27; k++) nb_pref[k + 1] += nb_pref[k];
        }
        __syncthreads();
        m = nb_pref[27];
        hs = nb_start[13];
        nh = nb_pref[14] - nb_pref[13];
    }
    bool ok = live;
    if (ATM && live && (m > kMaxM || nh > kMaxOwn)) {
        if (tid == 0) atomicOr(a.flags, FLAG_CELL_OVERFLOW);
        ok = false;
    }
    const int lps = nh <= 16 ? 16 : (nh <= 32 ? 8 : 4);
    const int per = kCellThreads / lps;
    const int h = tid / lps, sub = tid % lps;

    // ---- pair pass (LJ, clamp, neighbour lists)
    for (int base = 0; ok && base < nh; base += per) {
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
        if (tid < per) { hcnt[tid] = 0u; pcnt[tid] = 0u; }
        for (int c0 = 0; c0 < m; c0 += kStage) {
            const int mc = min(kStage, m - c0);
            __syncthreads();
            if (!ATM || base == 0) {  // three-body mode: whole neighbourhood staged once (one pass)
                for (int t = tid; t < mc; t += kCellThreads) {
                    const int gidx = c0 + t;
                    int k = 0;
#pragma unroll
                    for (int step = 16; step > 0; step >>= 1)
                        if (k + step < 27 && nb_pref[k + step] <= gidx) k += step;
                    const int j = nb_start[k] + (gidx - nb_pref[k]);
                    const float4 v = a.spos[j];
                    pos[t] = make_float4(v.x + nb_off[k][0], v.y + nb_off[k][1], v.z + nb_off[k][2],
                                         __uint_as_float(__float_as_uint(v.w) | nb_code[k]));
                    if (!UNIFORM) cprm[t] = a.sprm[j];
                    if (MOLS) cmol[t] = a.sid[j];
                    if (ATM == 2) sdir[t] = (uint8_t)k;
                    if (HYB) slam[t] = a.lam[__float_as_uint(v.w) & kIdxMask];
                }
            }
            __syncthreads();
            if (active) {
                for (int t = sub; t < mc; t += lps) {
                    const float4 c = pos[t];
                    const uint32_t w = __float_as_uint(c.w);
                    const uint32_t oj = w & kIdxMask;
                    if (oj == oi) continue;
                    const float dx = c.x - qi.x, dy = c.y - qi.y, dz = c.z - qi.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    bool in;
                    if (fabsf(r2 - a.rc2f) <= a.band)
                        in = !(exact_r2(a, wi, w) > a.rc2);
                    else
                        in = r2 <= a.rc2f;
                    if (!in) continue;
                    if (oj > oi) atomicAdd(&pcnt[h], 1u);
                    if (ATM) {
                        const uint32_t s = atomicAdd(&hcnt[h], 1u);
                        if (s < (uint32_t)kFullCap) nbl[h][s] = (uint16_t)(c0 + t);
                    }
                    if (LJ) {
                        if (MOLS && cmol[t] == mi) continue;
                        float sig2 = a.sig2, e24 = a.eps24, e4 = a.eps4;
                        if (!UNIFORM) {
                            const float2 pj = cprm[t];
                            const float s = 0.5f * (pi.x + pj.x);
                            const float e = pi.y * pj.y;
                            sig2 = s * s;
                            e24 = 24.f * e;
                            e4 = 4.f * e;
                        }
                        const float inv = __builtin_amdgcn_rcpf(r2);
                        const float s2 = sig2 * inv;
                        const float s6 = s2 * s2 * s2;
                        const float fac = e24 * (2.f * s6 * s6 - s6) * inv;
                        ax -= (double)(dx * fac);  // force on i = (r_i - r_j) fac
                        ay -= (double)(dy * fac);
                        az -= (double)(dz * fac);
                        if (ENERGY) {
                            u_acc += 0.5 * (double)(e4 * (s6 * s6 - s6));
                            w_acc += 0.5 * (double)(r2 * fac);
                        }
                    }
                }
            }
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
            }
        }
        __syncthreads();
        if (tid < per && base + tid < nh) {
            npairs += pcnt[tid];
            if (ATM && hcnt[tid] > (uint32_t)kFullCap) atomicOr(a.flags, FLAG_NL_OVERFLOW);
        }
    }

    if (ATM && ok) {
        // ---- three-body pass.  Home sites are dealt to the 4 waves by cost (pairs
        // n(n-1)/2, longest first to the least loaded wave) so the waves finish together.
        if (tid == 0) {
            uint32_t load[kWaves] = {0, 0, 0, 0};
            uint8_t cntw[kWaves] = {0, 0, 0, 0};
            uint8_t tmpw[kMaxOwn];
            uint64_t done = 0;
            for (int r = 0; r < nh; r++) {
                int best = -1;
                uint32_t bc = 0;
                for (int o = 0; o < nh; o++) {
                    if ((done >> o) & 1ull) continue;
                    const uint32_t n = min(hcnt[o], (uint32_t)kFullCap);
                    const uint32_t cst = n * (n - 1) / 2 + 64;
                    if (best < 0 || cst > bc) { best = o; bc = cst; }
                }
                done |= 1ull << best;
                int wmin = 0;
                for (int k = 1; k < kWaves; k++)
                    if (load[k] < load[wmin]) wmin = k;
                load[wmin] += bc;
                tmpw[best] = (uint8_t)wmin;
                cntw[wmin]++;
            }
            uint8_t off[kWaves];
            off[0] = 0;
            for (int k = 1; k < kWaves; k++) off[k] = off[k - 1] + cntw[k - 1];
            for (int k = 0; k < kWaves; k++) own_cnt[k] = cntw[k];
            for (int o = 0; o < nh; o++) own_of[off[tmpw[o]]++] = (uint8_t)o;
        }
        __syncthreads();
        const int home0 = nb_pref[13];
        int obase = 0;
        for (int k = 0; k < wv; k++) obase += own_cnt[k];
        for (int oo = 0; oo < own_cnt[wv]; oo++) {
            const int o = own_of[obase + oo];
            const int nn = (int)min(hcnt[o], (uint32_t)kFullCap);
            const float4 pi4 = pos[home0 + o];
            const uint32_t oi = __float_as_uint(pi4.w) & kIdxMask;
            float li = 1.f;
            if (HYB) li = slam[home0 + o];
            for (int s = lane; s < nn; s += 64) {
                const int t = nbl[o][s];
                const float4 v = pos[t];
                wbuf[wv][s] = make_float4(v.x - pi4.x, v.y - pi4.y, v.z - pi4.z, v.w);
                if (ATM == 2) wdir[wv][s] = sdir[t];
                if (HYB) wlam[wv][s] = slam[t];
            }
            wave_sync();
            float fx = 0.f, fy = 0.f, fz = 0.f, ev = 0.f;
            unsigned int visits = 0;
            // circulant enumeration of the n(n-1)/2 pairs: for every shift s < H = (n-1)/2
            // the pairs (a, a+s+1 mod n), a = 0..n-1, plus for even n the shift H with
            // a < n/2.  The shift loop is wave-uniform and unrolled by 4 so that a lane
            // loads its x_a once and four y_b together; lanes take a.
            const int H = (nn - 1) >> 1;
            const int halfn = (nn & 1) ? 0 : (nn >> 1);
            auto pair = [&](const float4 xa, const float4 yb, int ib, int ia) {
                const float zx = yb.x - xa.x, zy = yb.y - xa.y, zz = yb.z - xa.z;
                const float Z = zx * zx + zy * zy + zz * zz;
                bool in;
                if (fabsf(Z - a.rc2f) <= a.band3)
                    in = !(exact_r2(a, __float_as_uint(xa.w), __float_as_uint(yb.w)) > a.rc2);
                else
                    in = Z <= a.rc2f;
                if (!in) return;
                if (ATM == 2) {
                    const int da = wdir[wv][ia], db = wdir[wv][ib];
                    const bool same = (da == 13) && (db == 13);
                    const bool distinct = (da != 13) && (db != 13) && (da != db);
                    if (!same && !distinct) return;
                }
                float wt = 1.f;
                if (HYB) {
                    wt = (li + wlam[wv][ia] + wlam[wv][ib]) * (1.f / 3.f);
                    if (wt == 0.f) return;
                }
                visits++;
                const float X = xa.x * xa.x + xa.y * xa.y + xa.z * xa.z;
                const float Y = yb.x * yb.x + yb.y * yb.y + yb.z * yb.z;
                const float p1 = xa.x * yb.x + xa.y * yb.y + xa.z * yb.z;
                const float p2 = xa.x * zx + xa.y * zy + xa.z * zz;
                const float p3 = yb.x * zx + yb.y * zy + yb.z * zz;
                const float D = X * Y * Z;
                const float Pp = -(p1 * p2 * p3);
                const float rD = __builtin_amdgcn_rsqf(D);
                const float rD2 = rD * rD;
                const float D52 = rD2 * rD2 * rD;
                float K = 3.f * a.nu * D52 * wt;
                if (ATM == 2) K = -K;  // reference: +grad V
                const float A = -(D + 5.f * Pp);
                const float iX = __builtin_amdgcn_rcpf(X), iY = __builtin_amdgcn_rcpf(Y);
                const float t1 = p2 * p3, t23 = p1 * (p3 + p2);
                fx += K * (A * (xa.x * iX + yb.x * iY) - t1 * (xa.x + yb.x) - t23 * zx);
                fy += K * (A * (xa.y * iX + yb.y * iY) - t1 * (xa.y + yb.y) - t23 * zy);
                fz += K * (A * (xa.z * iX + yb.z * iY) - t1 * (xa.z + yb.z) - t23 * zz);
                ev += a.nu * D52 * (D + 3.f * Pp);  // unweighted V of this visit
            };
            for (int a0 = 0; nn > 1 && a0 < nn; a0 += 64) {
                const int ia = a0 + lane;
                if (ia < nn) {
                    const float4 xa = wbuf[wv][ia];
                    for (int sft = 0; sft < H; sft += 4) {
                        // issue the four loads before any data-dependent branch
                        int ib[4];
                        float4 yb[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            int t = ia + sft + u + 1;
                            t -= (t >= nn) ? nn : 0;
                            t -= (t >= nn) ? nn : 0;
                            ib[u] = t;
                            yb[u] = wbuf[wv][t];
                        }
#pragma unroll
                        for (int u = 0; u < 4; u++)
                            if (sft + u < H) pair(xa, yb[u], ib[u], ia);
                    }
                    if (ia < halfn) pair(xa, wbuf[wv][ia + halfn], ia + halfn, ia);  // even n: shift n/2
                }
            }
            double dx = wave_sum((double)fx), dy = wave_sum((double)fy), dz = wave_sum((double)fz);
            const double de = wave_sum((double)ev) * (1.0 / 3.0);  // this site's third of its triplets
            nvisit += visits;
            if (lane == 0) {
                if (HYB) {  // -V grad(w): site i receives -(sum V/3) dl_i/dx_axis
                    const double t = -de * (double)a.dlam[oi];
                    if (a.hyb_axis == 0) dx += t; else if (a.hyb_axis == 1) dy += t; else dz += t;
                }
                a.fx[oi] += dx;
                a.fy[oi] += dy;
                a.fz[oi] += dz;
                if (a.esite) a.esite[oi] = HYB ? de * (double)li : de;
            }
            if (ENERGY && lane == 0) e_acc += HYB ? de * (double)li : de;
            wave_sync();
        }
    }

    // ---- counters and energy partials
    npairs = (unsigned long long)wave_sum((double)npairs);
    nvisit = (unsigned long long)wave_sum((double)nvisit);
    if (lane == 0) {
        if (npairs) atomicAdd(&a.counters[0], npairs);
        if (nvisit) atomicAdd(&a.counters[1], nvisit);
    }
    if (ENERGY) {
        u_acc = wave_sum(u_acc);
        w_acc = wave_sum(w_acc);
        e_acc = wave_sum(e_acc);
        if (lane == 0) {
            red[0][wv] = u_acc;
            red[1][wv] = w_acc;
            red[2][wv] = e_acc;
        }
        __syncthreads();
        if (tid == 0) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int k = 0; k < kWaves; k++) { s0 += red[0][k]; s1 += red[1][k]; s2 += red[2][k]; }
            a.part[b] = s0;
            a.part[a.nblocks + b] = s1;
            a.part[2 * a.nblocks + b] = s2;
        }
    }
}

template <bool U, bool M, bool LJ, int ATM, bool E, bool HYB, bool LIM>
void go(komd_ctx* c, const CellArgs& a) {
    k_cell<U, M, LJ, ATM, E, HYB, LIM><<<a.nblocks, kCellThreads, 0, c->stream>>>(a);
}

template <bool U, bool M, bool LIM>
void dispatch_lim(komd_ctx* c, const CellArgs& a, bool lj, int atm, bool energy, bool hyb) {
    if (hyb) {
        if (energy) go<U, M, true, 1, true, true, LIM>(c, a);
        else go<U, M, true, 1, false, true, LIM>(c, a);
    } else if (atm == 0) {
        if (energy) go<U, M, true, 0, true, false, LIM>(c, a);
        else go<U, M, true, 0, false, false, LIM>(c, a);
    } else if (atm == 1) {
        if (lj && energy) go<U, M, true, 1, true, false, LIM>(c, a);
        else if (lj) go<U, M, true, 1, false, false, LIM>(c, a);
        else if (energy) go<U, M, false, 1, true, false, false>(c, a);
        else go<U, M, false, 1, false, false, false>(c, a);
    } else {
        if (lj && energy) go<U, M, true, 2, true, false, LIM>(c, a);
        else if (lj) go<U, M, true, 2, false, false, LIM>(c, a);
        else if (energy) go<U, M, false, 2, true, false, false>(c, a);
        else go<U, M, false, 2, false, false, false>(c, a);
    }
}

template <bool U, bool M>
hipError_t dispatch(komd_ctx* c, const CellArgs& a, bool lj, int atm, bool energy, bool hyb, bool lim) {
    if (lim) dispatch_lim<U, M, true>(c, a, lj, atm, energy, hyb);
    else dispatch_lim<U, M, false>(c, a, lj, atm, energy, hyb);
    return hipGetLastError();
}

}  // namespace

int cell_blocks(komd_ctx* c) { return 8 * ((c->grid.ncell + 7) / 8); }

hipError_t launch_cell(komd_ctx* c, bool lj, int atm, bool energy, bool limit, const float* lam, const float* dlam,
                       double* esite) {
    auto& d = c->d;
    CellArgs a{};
    a.spos = d.spos; a.sprm = d.sprm; a.sid = d.sid; a.start = d.start;
    a.rx = d.rx; a.ry = d.ry; a.rz = d.rz; a.eps = d.eps; a.sig = d.sig;
    a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.esite = esite; a.lam = lam; a.dlam = dlam;
    a.part = d.red; a.flags = d.flags; a.counters = d.counters;
    a.g = c->grid;
    a.chunk = (c->grid.ncell + 7) / 8;
    a.nblocks = 8 * a.chunk;
    a.hyb_axis = c->hyb_axis;
    a.rc2 = c->prm.cutoff * c->prm.cutoff;
    a.rc2f = (float)a.rc2;
    a.band = 4e-6f * a.rc2f;
    a.band3 = 1e-5f * a.rc2f;
    a.sig2 = c->sig_u * c->sig_u;
    a.eps24 = 24.f * c->eps_u;
    a.eps4 = 4.f * c->eps_u;
    a.nu = (float)c->prm.nu;
    a.limit = c->prm.limit_factor;
    const bool lim = limit && c->prm.limit_factor > 0;
    const bool hyb = lam != nullptr;
    ScopedTimer tm(c, atm ? (lj ? "cell_lj_atm" : "cell_atm") : "cell_lj");
    const bool U = c->uniform, M = c->multisite;
    if (U && !M) return dispatch<true, false>(c, a, lj, atm, energy, hyb, lim);
    if (U && M) return dispatch<true, true>(c, a, lj, atm, energy, hyb, lim);
    if (!U && !M) return dispatch<false, false>(c, a, lj, atm, energy, hyb, lim);
    return dispatch<false, true>(c, a, lj, atm, energy, hyb, lim);
}

}  // namespace komd
