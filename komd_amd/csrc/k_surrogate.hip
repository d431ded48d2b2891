// Learned three-body surrogate -- gfx950.
//
// The surrogate replaces the ATM energy by sum_i E_i, E_i = MLP(x_i), over
// per-site descriptors of the full neighbour list (cell-sorted order p):
//
//   k_sur_desc   four lanes per site: 32 radial Gaussians and the moment tensors
//                T_a^L = sum_j psi_a(r_ij) u_ij^(x L) (4 channels, L = 1..3, 19
//                unique symmetric components per channel); the 16 angular
//                invariants are <T_a^L, T_b^L>.  Writes x = (D - shift) * scale
//                and T.
//   k_sur_mlp    one wave per 16 sites, BF16 MFMA (v_mfma_f32_16x16x32_bf16)
//                with FP32 accumulation: forward 48 -> 64 -> 64 -> 1 (SiLU) and
//                the input-gradient backward g = dE/dx.  The hidden units are
//                stored in a lane-matched order (kappa below) so that a layer's
//                accumulator registers are already the next layer's B operand:
//                no LDS round trip between layers.  Writes the per-site gradient
//                S = dE/dD (48, scaled; the force pass forms dE/dT_a^L from it and T).
//   k_sur_force  one workgroup per cell: F = -dE/dr as a sum over directed pairs of
//                c_i(d_ij) = dE_i/dd_ij from site i's own state, +c to i and -c to j,
//                accumulated in LDS over the cell's 27-cell neighbourhood and written
//                as partial rows; k_sur_apply sums the 27 rows covering each site.
//
// Training (loss L = wE/N sum (E_i - E*_i)^2 + wF/(3N) sum |F_i - F*_i|^2):
//   k_sur_apply<TRAIN>  delta_i = dL/dF_i
//   k_sur_jvp           w_a = dL/dg_a = -scale * (dD_a along the field delta)
//   k_sur_train         forward + tangent + backward signals per 16-site tile
//                       (MFMA), weight gradients as MFMA products over 32-site
//                       chunks through LDS, per-workgroup partial sums
//   k_sur_adam          fixed-order reduction of the partials and Adam.
// The same formulas are restated in FP64 in oracle/surrogate.py.
#include <math.h>

#include <cmath>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "komd_cell.hpp"
#include "komd_reduce.hpp"

namespace komd {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NR = 32;             // radial descriptors
constexpr int NA = 4;              // angular channels
constexpr int NC = 19;             // unique components of T^1, T^2, T^3
constexpr int NT = NA * NC;        // 76 moment values per site
constexpr int ND = 48;             // descriptors
constexpr int NH = 64;             // hidden width
constexpr int NS = 112;            // staged state row of the force pass: 32 radial + 4 x 20 angular
constexpr int O_W1 = 0, O_B1 = O_W1 + NH * ND, O_W2 = O_B1 + NH, O_B2 = O_W2 + NH * NH, O_W3 = O_B2 + NH,
              O_B3 = O_W3 + NH, O_SHIFT = O_B3 + 1, O_SCALE = O_SHIFT + ND, NPARAM = O_SCALE + ND;
static_assert(NPARAM == KOMD_SUR_NPARAM, "parameter layout");
constexpr int NOPS = 30 * 64;      // packed BF16 operand tables (bf8 entries)
constexpr int kOp1 = 0, kOp2 = 512, kOp3 = 1024, kOp4 = 1536;

constexpr float kRMu0 = 0.85f;
constexpr float kRInvW = 31.f / 1.6f;   // 1 / spacing = 1 / width
constexpr float kAInvW = 1.f / 0.35f;
__device__ __forceinline__ float a_mu(int ch) {
    return ch == 0 ? 1.0f : (ch == 1 ? 1.35f : (ch == 2 ? 1.75f : 2.15f));
}
// angular invariants (a, b, L)
__device__ __forceinline__ void inv_of(int m, int& ia, int& ib, int& L) {
    const int k = m < 12 ? m % 6 : 0;
    const int pa[6] = {0, 0, 0, 0, 1, 1}, pb[6] = {0, 1, 2, 3, 1, 2};
    if (m < 12) {
        ia = pa[k];
        ib = pb[k];
        L = m < 6 ? 1 : 2;
    } else {
        ia = ib = m - 12;
        L = 3;
    }
}
__device__ __forceinline__ int l_off(int L) { return L == 1 ? 0 : (L == 2 ? 3 : 9); }
__device__ __forceinline__ int l_cnt(int L) { return L == 1 ? 3 : (L == 2 ? 6 : 10); }
// multiplicity of a unique component in the full contraction
__device__ __forceinline__ float mult(int c) {
    return c < 6 ? 1.f : (c < 9 ? 2.f : (c < 12 ? 1.f : (c < 18 ? 3.f : 6.f)));
}

// unique components of u, u u, u u u: x y z | xx yy zz xy xz yz |
// xxx yyy zzz xxy xxz xyy yyz xzz yzz xyz
__device__ __forceinline__ void comps(float x, float y, float z, float t[NC]) {
    t[0] = x; t[1] = y; t[2] = z;
    t[3] = x * x; t[4] = y * y; t[5] = z * z; t[6] = x * y; t[7] = x * z; t[8] = y * z;
    t[9] = t[3] * x; t[10] = t[4] * y; t[11] = t[5] * z; t[12] = t[3] * y; t[13] = t[3] * z;
    t[14] = x * t[4]; t[15] = t[4] * z; t[16] = x * t[5]; t[17] = y * t[5]; t[18] = t[6] * z;
}

// d(u^(x L)) along the transverse displacement w (times r): symmetrised slots
__device__ __forceinline__ void sym_comps(float x, float y, float z, float wx, float wy, float wz, float t[NC]) {
    t[0] = wx; t[1] = wy; t[2] = wz;
    t[3] = 2.f * wx * x; t[4] = 2.f * wy * y; t[5] = 2.f * wz * z;
    t[6] = wx * y + x * wy; t[7] = wx * z + x * wz; t[8] = wy * z + y * wz;
    t[9] = 3.f * wx * x * x; t[10] = 3.f * wy * y * y; t[11] = 3.f * wz * z * z;
    t[12] = 2.f * wx * x * y + x * x * wy; t[13] = 2.f * wx * x * z + x * x * wz;
    t[14] = wx * y * y + 2.f * x * y * wy; t[15] = 2.f * wy * y * z + y * y * wz;
    t[16] = wx * z * z + 2.f * x * z * wz; t[17] = wy * z * z + 2.f * y * z * wz;
    t[18] = wx * y * z + x * wy * z + x * y * wz;
}

// <A, d(psi u^L)/dd> summed over L = 1..3 for one channel: psi' S u + psi/r (V - s u),
// V = V1 + 2 V2 + 3 V3, s = s1 + 2 s2 + 3 s3, S = s1 + s2 + s3 (A given by its 19
// unique symmetric components).
__device__ __forceinline__ void chan_grad(const float* A, float ux, float uy, float uz, float dpsi, float psir,
                                          float& cx, float& cy, float& cz) {
    const float v1x = A[0], v1y = A[1], v1z = A[2];
    const float v2x = A[3] * ux + A[6] * uy + A[7] * uz;
    const float v2y = A[6] * ux + A[4] * uy + A[8] * uz;
    const float v2z = A[7] * ux + A[8] * uy + A[5] * uz;
    const float xx = ux * ux, yy = uy * uy, zz = uz * uz, xy = 2.f * ux * uy, xz = 2.f * ux * uz,
                yz = 2.f * uy * uz;
    const float v3x = A[9] * xx + A[14] * yy + A[16] * zz + A[12] * xy + A[13] * xz + A[18] * yz;
    const float v3y = A[12] * xx + A[10] * yy + A[17] * zz + A[14] * xy + A[18] * xz + A[15] * yz;
    const float v3z = A[13] * xx + A[15] * yy + A[11] * zz + A[18] * xy + A[16] * xz + A[17] * yz;
    const float s1 = v1x * ux + v1y * uy + v1z * uz;
    const float s2 = v2x * ux + v2y * uy + v2z * uz;
    const float s3 = v3x * ux + v3y * uy + v3z * uz;
    const float S = s1 + s2 + s3;
    const float sw = s1 + 2.f * s2 + 3.f * s3;
    const float vx = v1x + 2.f * v2x + 3.f * v3x, vy = v1y + 2.f * v2y + 3.f * v3y, vz = v1z + 2.f * v2z + 3.f * v3z;
    const float a = dpsi * S - psir * sw;
    cx += a * ux + psir * vx;
    cy += a * uy + psir * vy;
    cz += a * uz + psir * vz;
}

struct Pair {
    float r, ir, ux, uy, uz, fc, dfc;
};
__device__ __forceinline__ Pair pair_of(float4 v, float pirc) {
    Pair q;
    const float r2 = v.x * v.x + v.y * v.y + v.z * v.z;
    q.ir = __builtin_amdgcn_rsqf(r2);
    q.r = r2 * q.ir;
    q.ux = v.x * q.ir;
    q.uy = v.y * q.ir;
    q.uz = v.z * q.ir;
    float sn, cs;
    __sincosf(q.r * pirc, &sn, &cs);
    q.fc = 0.5f * (cs + 1.f);
    q.dfc = -0.5f * pirc * sn;
    return q;
}
// psi_a and dpsi_a/dr
__device__ __forceinline__ void ang_basis(const Pair& q, int ch, float& psi, float& dpsi) {
    const float x = (q.r - a_mu(ch)) * kAInvW;
    const float g = __expf(-x * x);
    const float ir3 = q.ir * q.ir * q.ir;
    psi = g * q.fc * ir3;
    dpsi = g * ir3 * (q.dfc - 2.f * x * kAInvW * q.fc - 3.f * q.ir * q.fc);
}

// exp(-(y - k)^2), k = 0..31, in blocks of 4: one exponential for the block's first
// value and one for the ratio, g_{a+t+1} = g_{a+t} exp(2(y-a-t) - 1), the ratio itself
// falling by e^-2 per step.  A block whose first value underflows (|y - a| > 9.3) has
// every value below exp(-39).
__device__ __forceinline__ void radial_gauss(float y, float g[NR]) {
#pragma unroll
    for (int b = 0; b < NR / 4; b++) {
        const float e = y - (float)(4 * b);
        float v = __expf(-e * e);
        float R = __expf(2.f * e - 1.f);
        g[4 * b] = v;
#pragma unroll
        for (int t = 1; t < 4; t++) {
            v *= R;
            R *= 0.1353352832366127f;  // e^-2
            g[4 * b + t] = v;
        }
    }
}

// lanes per site of the per-site kernels (neighbours strided over the lanes, sums
// combined with shuffles inside the lane group)
constexpr int kLps = 4;
// Sites of the 27-cell neighbourhood held by the surrogate force pass (kSurM: on average
// about 394 at density 0.8 with cells of edge 2.63, 500 with the 2.85 cells of a 0.3 skin;
// a larger neighbourhood is reported as a cell overflow).

// ---------------------------------------------------------------------------
// descriptors
struct DescArgs {
    const float4* __restrict__ nl;
    const uint32_t* __restrict__ nlc;
    const float* __restrict__ params;  // shift / scale (nullptr: raw descriptors)
    float* __restrict__ X;             // [n][48]
    float* __restrict__ T;             // [n][76]
    const float* __restrict__ lam;     // hybrid: sites with lambda = 1 carry no surrogate energy (skipped)
    const uint32_t* __restrict__ sorted;
    int n;
    float pirc;
    // LJ 12-6 of every listed pair added to the site's force by the radial role (the
    // step's pair pass then only builds lists): uniform sites, no clamp, no hybrid
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    int lj;
    float sig2, eps24;
};

// Radial window: the kWin Gaussians around y.  Every Gaussian outside
// k0 .. k0 + kWin - 1, k0 = clamp(floor(y) - 4, 0, NR - kWin), lies at |y - k| >= 4,
// i.e. below exp(-16) = 1.1e-7, under the FP32 resolution of the sums; the descriptor
// and force passes skip them (the training kernels keep all 32).
constexpr int kWin = 9;
__device__ __forceinline__ int win_start(float y) {
    const int k = (int)floorf(y) - 4;
    return k < 0 ? 0 : (k > NR - kWin ? NR - kWin : k);
}
// g[t] = exp(-(e - t)^2), t < kWin, e = y - k0: one exponential for g[0], one for the
// ratio g[t+1] / g[t] = exp(2(e - t) - 1), which falls by e^-2 per step.
__device__ __forceinline__ void win_gauss(float e, float g[kWin]) {
    float v = __expf(-e * e);
    float R = __expf(2.f * e - 1.f);
    g[0] = v;
#pragma unroll
    for (int t = 1; t < kWin; t++) {
        v *= R;
        R *= 0.1353352832366127f;  // e^-2
        g[t] = v;
    }
}

// One workgroup per 32 sites.  Waves 0-1 (4 lanes per site) accumulate the 76 moment
// sums and write the angular invariants and T; waves 2-3 (4 lanes per site) accumulate
// the 32 radial sums.  Each lane then holds one of the two accumulator sets only, which
// halves the registers of the single-role form and doubles its occupancy.
constexpr int kDescSites = 32;
// LDS writes of this wave visible to its later reads
__device__ __forceinline__ void desc_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
template <bool ANG>
__device__ __forceinline__ void desc_role(const DescArgs& a, int p, int sub, float* __restrict__ lds) {
    constexpr int NACC = ANG ? NT : NR;
    float acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) acc[k] = 0.f;
    int nn = 0;
    if (p < a.n) {
        nn = (int)a.nlc[p];
        if (a.lam && a.lam[a.sorted[p]] == 1.f) nn = 0;  // weight 1 - lambda = 0: descriptors unused
    }
    const float4* row = a.nl + (size_t)(p < a.n ? p : 0) * kFullCap;
    // neighbour entries are loaded one iteration ahead of their use
    float4 vnext = sub < nn ? row[sub] : make_float4(1.f, 0.f, 0.f, 0.f);
    for (int s = sub; s < nn; s += kLps) {
        const float4 v = vnext;
        if (s + kLps < nn) vnext = row[s + kLps];
        const Pair q = pair_of(v, a.pirc);
        if (ANG) {
            float t[NC];
            comps(q.ux, q.uy, q.uz, t);
#pragma unroll
            for (int ch = 0; ch < NA; ch++) {
                const float x = (q.r - a_mu(ch)) * kAInvW;
                const float psi = __expf(-x * x) * q.fc * (q.ir * q.ir * q.ir);
#pragma unroll
                for (int c = 0; c < NC; c++) acc[ch * NC + c] += psi * t[c];
            }
        } else {
            float g[NR];
            radial_gauss((q.r - kRMu0) * kRInvW, g);
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] += g[k] * q.fc;
        }
    }
    if (ANG) {
        // the 4 lanes' moment sums combined through LDS (a register reduction of 76 values
        // keeps a temporary per value): every lane stores its sums, each lane adds a
        // quarter of the components over the 4 lanes (fixed order), every lane reads the
        // totals back
        const int l = threadIdx.x & 127;
        float* mine = lds + l * (NT + 1);
        float* site = lds + (l & ~3) * (NT + 1);
#pragma unroll
        for (int k = 0; k < NT; k++) mine[k] = acc[k];
        desc_wave_sync();
        float tot[(NT + kLps - 1) / kLps];
#pragma unroll
        for (int q = 0; q < (NT + kLps - 1) / kLps; q++) {
            const int k = q * kLps + sub;
            tot[q] = k < NT ? ((site[k] + site[(NT + 1) + k]) + site[2 * (NT + 1) + k]) + site[3 * (NT + 1) + k]
                            : 0.f;
        }
        desc_wave_sync();
#pragma unroll
        for (int q = 0; q < (NT + kLps - 1) / kLps; q++) {
            const int k = q * kLps + sub;
            if (k < NT) site[k] = tot[q];
        }
        desc_wave_sync();
    } else {
        static_assert(kLps == 2 || kLps == 4 || kLps == 8 || kLps == 16, "DPP group sum");
#pragma unroll
        for (int k = 0; k < NACC; k++) acc[k] = group_sum(acc[k], kLps);
    }
    if (p >= a.n) return;
    float4* xo = reinterpret_cast<float4*>(a.X + (size_t)p * ND);
    if (ANG) {
        // invariants and T from the site's summed moments in LDS (same address in its 4 lanes)
        const float* M = lds + ((threadIdx.x & 127) & ~3) * (NT + 1);
        float D[16];
#pragma unroll
        for (int m = 0; m < 16; m++) {
            int ia, ib, L;
            inv_of(m, ia, ib, L);
            float sm = 0.f;
#pragma unroll
            for (int c = 0; c < 10; c++)
                if (c < l_cnt(L)) {
                    const int cc = l_off(L) + c;
                    sm += mult(cc) * M[ia * NC + cc] * M[ib * NC + cc];
                }
            D[m] = sm;
        }
        if (a.params) {
#pragma unroll
            for (int m = 0; m < 16; m++) D[m] = (D[m] - a.params[O_SHIFT + NR + m]) * a.params[O_SCALE + NR + m];
        }
        // the lanes of the site share the stores: invariants are float4 chunks 8..11 of the row
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k == sub) xo[8 + k] = make_float4(D[4 * k], D[4 * k + 1], D[4 * k + 2], D[4 * k + 3]);
        float4* to = reinterpret_cast<float4*>(a.T + (size_t)p * NT);
#pragma unroll
        for (int k = 0; k < NT / 4; k++)
            if (k % kLps == sub) to[k] = make_float4(M[4 * k], M[4 * k + 1], M[4 * k + 2], M[4 * k + 3]);
    } else {
        if (a.params) {
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] = (acc[k] - a.params[O_SHIFT + k]) * a.params[O_SCALE + k];
        }
        // radial descriptors are float4 chunks 0..7 of the row
#pragma unroll
        for (int k = 0; k < NR / 4; k++)
            if (k % kLps == sub) xo[k] = make_float4(acc[4 * k], acc[4 * k + 1], acc[4 * k + 2], acc[4 * k + 3]);
        if (a.lj) {
            // LJ 12-6 of the site's pairs, a second walk over its list (L2-resident) once the
            // radial sums are stored: the pair pass's expression, FP32 per pair, force on
            // i = -x fac, FP64 sums, combined over the quad in a fixed order
            double ljx = 0.0, ljy = 0.0, ljz = 0.0;
            for (int s = sub; s < nn; s += kLps) {
                const float4 v = row[s];
                const float inv = __builtin_amdgcn_rcpf(v.x * v.x + v.y * v.y + v.z * v.z);
                const float s2 = a.sig2 * inv;
                const float s6 = s2 * s2 * s2;
                const float fac = a.eps24 * (2.f * s6 * s6 - s6) * inv;
                ljx -= (double)(v.x * fac);
                ljy -= (double)(v.y * fac);
                ljz -= (double)(v.z * fac);
            }
            static_assert(kLps == 4, "LJ sum over a quad");
            ljx += dpp_f64<0xB1>(ljx); ljx += dpp_f64<0x4E>(ljx);
            ljy += dpp_f64<0xB1>(ljy); ljy += dpp_f64<0x4E>(ljy);
            ljz += dpp_f64<0xB1>(ljz); ljz += dpp_f64<0x4E>(ljz);
            if (sub == 0) {
                const uint32_t oi = a.sorted[p];
                a.fx[oi] = a.fx[oi] + ljx;
                a.fy[oi] = a.fy[oi] + ljy;
                a.fz[oi] = a.fz[oi] + ljz;
            }
        }
    }
}

// 4 waves per SIMD (the LDS of the angular reduction allows 4 workgroups per CU)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_sur_desc(DescArgs a) {
    __shared__ float red[128 * (NT + 1)];  // moment sums of the angular waves' lanes
    const int l = threadIdx.x & 127;
    const int p = blockIdx.x * kDescSites + l / kLps, sub = l % kLps;
    if (threadIdx.x < 128)
        desc_role<true>(a, p, sub, red);
    else
        desc_role<false>(a, p, sub, red);
}

// ---------------------------------------------------------------------------
// network on the MFMA cores
//
// v_mfma_f32_16x16x32_bf16, lane l (i = l % 16, q = l / 16):
//   A operand  A[i][8q + t], t = 0..7;  B operand  B[8q + t][i];  D  D[4q + r][i], r = 0..3.
// Layers are computed transposed, z^T = W x^T, so the 16 columns are 16 sites and
// lane l serves site l % 16.  Output row rho = 16 m + 4 q + r of block m holds hidden
// unit kappa(rho) = 32 (m/2) + 8 q + 4 (m%2) + r: then the registers
// {block 2kk, block 2kk+1} of a lane are exactly the units 32 kk + 8 q + t it must
// supply as the next product's B operand for K-step kk.
__device__ __forceinline__ int kappa(int rho) {
    const int m = rho >> 4, w = rho & 15;
    return 32 * (m >> 1) + 8 * (w >> 2) + 4 * (m & 1) + (w & 3);
}
__device__ __forceinline__ f4 mfma(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ bf8 pack2(const f4& lo, const f4& hi) {
    bf8 r;
    r[0] = (__bf16)lo[0]; r[1] = (__bf16)lo[1]; r[2] = (__bf16)lo[2]; r[3] = (__bf16)lo[3];
    r[4] = (__bf16)hi[0]; r[5] = (__bf16)hi[1]; r[6] = (__bf16)hi[2]; r[7] = (__bf16)hi[3];
    return r;
}
__device__ __forceinline__ bf8 load8(const float* p, bool ok) {
    bf8 r;
    if (ok) {
        const float4 u = reinterpret_cast<const float4*>(p)[0], v = reinterpret_cast<const float4*>(p)[1];
        r[0] = (__bf16)u.x; r[1] = (__bf16)u.y; r[2] = (__bf16)u.z; r[3] = (__bf16)u.w;
        r[4] = (__bf16)v.x; r[5] = (__bf16)v.y; r[6] = (__bf16)v.z; r[7] = (__bf16)v.w;
    } else {
#pragma unroll
        for (int t = 0; t < 8; t++) r[t] = (__bf16)0.f;
    }
    return r;
}
// SiLU and its first two derivatives
__device__ __forceinline__ void silu(float z, float& h, float& d, float& dd) {
    const float s = __builtin_amdgcn_rcpf(1.f + __expf(-z));
    h = z * s;
    d = s * (1.f + z * (1.f - s));
    dd = s * (1.f - s) * (2.f + z * (1.f - 2.f * s));
}

// Packs the four operand tables (BF16, operand order) from the FP32 parameters:
//   op1  W1 rows in kappa order (layer 1 A)    op2  W2 rows in kappa order (layer 2 A)
//   op3  W2^T rows in kappa order (backward)   op4  W1^T, descriptor rows (backward)
__global__ __launch_bounds__(256) void k_sur_pack(const float* __restrict__ prm, bf8* __restrict__ ops) {
    for (int e = threadIdx.x; e < NOPS; e += 256) {
        const int tbl = e >> 9, idx = e & 511, blk = idx >> 6, lane = idx & 63;
        const int m = blk >> 1, kk = blk & 1, i = lane & 15, q = lane >> 4;
        bf8 v;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int k = 32 * kk + 8 * q + t;
            float x = 0.f;
            if (tbl == 0) x = k < ND ? prm[O_W1 + kappa(16 * m + i) * ND + k] : 0.f;
            else if (tbl == 1) x = prm[O_W2 + kappa(16 * m + i) * NH + k];
            else if (tbl == 2) x = prm[O_W2 + k * NH + kappa(16 * m + i)];
            else x = (m < 3) ? prm[O_W1 + k * ND + 16 * m + i] : 0.f;
            v[t] = (__bf16)x;
        }
        ops[e] = v;
    }
}

// Per-lane copies of the unit vectors in register order (m, r) -> kappa(16 m + 4 q + r).
struct LaneVecs {
    float b1[16], b2[16], w3[16];
};
__device__ __forceinline__ void lane_vecs(const float* __restrict__ prm, int q, LaneVecs& v) {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int u = kappa(16 * m + 4 * q + r);
            v.b1[4 * m + r] = prm[O_B1 + u];
            v.b2[4 * m + r] = prm[O_B2 + u];
            v.w3[4 * m + r] = prm[O_W3 + u];
        }
}

// Forward and input-gradient backward of one 16-site tile.  On return z1/h1/d1/dd1,
// z2/h2/d2/dd2 hold the activations (register order), g[3] the gradient dE/dx
// (row 16 m + 4 q + r = descriptor), e the site energy (same in the 4 lanes of a site).
struct TileFwd {
    f4 h1[4], d1[4], dd1[4], h2[4], d2[4], dd2[4], t2[4], q1[4], g[3];
    float e;
};
__device__ __forceinline__ void tile_forward(const bf8* __restrict__ ops, int lane, const bf8 xb[2],
                                             const LaneVecs& v, float b3, TileFwd& o) {
    f4 z1[4], z2[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(ops[kOp1 + (2 * m) * 64 + lane], xb[0], acc);
        acc = mfma(ops[kOp1 + (2 * m + 1) * 64 + lane], xb[1], acc);
        z1[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float h, d, dd;
            silu(z1[m][r] + v.b1[4 * m + r], h, d, dd);
            o.h1[m][r] = h;
            o.d1[m][r] = d;
            o.dd1[m][r] = dd;
        }
    const bf8 hb0 = pack2(o.h1[0], o.h1[1]), hb1 = pack2(o.h1[2], o.h1[3]);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(ops[kOp2 + (2 * m) * 64 + lane], hb0, acc);
        acc = mfma(ops[kOp2 + (2 * m + 1) * 64 + lane], hb1, acc);
        z2[m] = acc;
    }
    float e = 0.f;
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float h, d, dd;
            silu(z2[m][r] + v.b2[4 * m + r], h, d, dd);
            o.h2[m][r] = h;
            o.d2[m][r] = d;
            o.dd2[m][r] = dd;
            e += v.w3[4 * m + r] * o.h2[m][r];
            o.t2[m][r] = o.d2[m][r] * v.w3[4 * m + r];
        }
    e += __shfl_xor(e, 16);
    e += __shfl_xor(e, 32);
    o.e = e + b3;
    // q1 = W2^T t2 (kappa rows), g = W1^T (d1 * q1)
    const bf8 tb0 = pack2(o.t2[0], o.t2[1]), tb1 = pack2(o.t2[2], o.t2[3]);
    f4 d1q[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(ops[kOp3 + (2 * m) * 64 + lane], tb0, acc);
        acc = mfma(ops[kOp3 + (2 * m + 1) * 64 + lane], tb1, acc);
        o.q1[m] = acc;
        d1q[m] = o.d1[m] * acc;
    }
    const bf8 db0 = pack2(d1q[0], d1q[1]), db1 = pack2(d1q[2], d1q[3]);
#pragma unroll
    for (int m = 0; m < 3; m++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(ops[kOp4 + (2 * m) * 64 + lane], db0, acc);
        acc = mfma(ops[kOp4 + (2 * m + 1) * 64 + lane], db1, acc);
        o.g[m] = acc;
    }
}

struct MlpArgs {
    const bf8* __restrict__ ops;
    const float* __restrict__ params;
    const float* __restrict__ X;
    const float* __restrict__ T;
    const uint32_t* __restrict__ sorted;
    const float* __restrict__ lam;   // hybrid: exact-ATM weight per site (site order), or nullptr
    float* __restrict__ S;           // [n][48] dE/dD (scaled, weighted by the surrogate share)
    float* __restrict__ E;           // [n] per-site energy (cell-sorted), optional
    double* __restrict__ part;       // per-workgroup energy partials, optional
    int n;
};

// 2 waves per SIMD (a 512-workgroup persistent grid fills both)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_sur_mlp(MlpArgs a) {
    __shared__ bf8 ops[NOPS];
    __shared__ double red[4];
    for (int e = threadIdx.x; e < NOPS; e += 256) ops[e] = a.ops[e];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, q = lane >> 4, si = lane & 15;
    LaneVecs v;
    lane_vecs(a.params, q, v);
    const float b3 = a.params[O_B3];
    float sc[12];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) sc[4 * m + r] = a.params[O_SCALE + 16 * m + 4 * q + r];
    const int ntiles = (a.n + 15) >> 4;
    double eacc = 0.0;
    for (int tile = blockIdx.x * 4 + wv; tile < ntiles; tile += gridDim.x * 4) {
        const int p = tile * 16 + si;
        const bool ok = p < a.n;
        const float* xr = a.X + (size_t)(ok ? p : 0) * ND;
        bf8 xb[2];
        xb[0] = load8(xr + 8 * q, ok);
        xb[1] = load8(xr + 32 + 8 * q, ok && q < 2);
        TileFwd o;
        tile_forward(ops, lane, xb, v, b3, o);
        float wgt = 1.f;
        if (a.lam && ok) wgt = 1.f - a.lam[a.sorted[p]];
        if (ok && q == 0) {
            if (a.E) a.E[p] = o.e;
            eacc += (double)(o.e * wgt);
        }
        // G = dE/dD = g * scale (weighted by the surrogate share in hybrid runs)
        float G[12];
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) G[4 * m + r] = o.g[m][r] * sc[4 * m + r] * wgt;
        // lane q holds descriptors 16 m + 4 q + r: float4 chunks q, 4 + q, 8 + q of the row
        if (ok) {
            float4* grow = reinterpret_cast<float4*>(a.S + (size_t)p * ND);
#pragma unroll
            for (int m = 0; m < 3; m++)
                grow[4 * m + q] = make_float4(G[4 * m], G[4 * m + 1], G[4 * m + 2], G[4 * m + 3]);
        }
    }
    if (a.part) {
        for (int o = 32; o > 0; o >>= 1) eacc += __shfl_xor(eacc, o);
        if (lane == 0) red[wv] = eacc;
        __syncthreads();
        if (threadIdx.x == 0) a.part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

// ---------------------------------------------------------------------------
// forces from the states
//
// F = -dE/dr is a sum over directed pairs of c_i(d_ij) = dE_i/dd_ij, which needs only
// site i's own state: +c goes to i, -c to j.  k_sur_force runs one workgroup per home
// cell (XCD-aware order): the home sites' states are staged in LDS, every pair of a
// home site adds -c to an LDS accumulator slot of its neighbour (27 neighbour cells,
// slot = cell direction and rank in the cell) and its own sum to its home slot; the
// accumulator is written as one partial row per (home cell, direction, rank).
// k_sur_apply sums, for each site, the 27 rows that cover it (its own cell and the 26
// cells that see it) and hands the force to its consumer.  No state is gathered, no
// global atomic is used, and the result does not depend on scheduling.
enum { FORCE_APPLY = 0, FORCE_TRAIN = 1, FORCE_OUT = 2 };

struct ForceArgs {
    const float4* __restrict__ nl;
    const uint32_t* __restrict__ nlc;
    const uint16_t* __restrict__ nld;
    const uint32_t* __restrict__ inv;
    const uint32_t* __restrict__ sorted;
    const uint32_t* __restrict__ key;
    const uint32_t* __restrict__ start;
    const float* __restrict__ S;      // [n][48] dE/dD of every site (cell-sorted)
    const float* __restrict__ T;      // [n][76] moment tensors of every site (cell-sorted)
    const float* __restrict__ E;      // hybrid: per-site surrogate energies (cell-sorted)
    const float* __restrict__ dlam;   // hybrid: d lambda / d axis coordinate (site order)
    const float* __restrict__ lam;    // hybrid: lambda (site order)
    float4* __restrict__ P;           // partial rows [ncell][27][kRows]
    uint32_t* __restrict__ flags;
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    const float* __restrict__ fref;   // training: exact forces [n][3], site order
    float4* __restrict__ delta;       // training: dL/dF per site (cell-sorted)
    float* __restrict__ fout;         // FORCE_OUT: forces [n][3], site order
    double* __restrict__ part;        // training: per-block loss partials
    double* __restrict__ vpart;       // energy: [kSym][vparts] virial partials
    double* __restrict__ vx;          // second half-kick fused into the force write (kick > 0)
    double* __restrict__ vy;
    double* __restrict__ vz;
    const double* __restrict__ mass;
    const SkinArgs* __restrict__ sk;  // Verlet-list step: the next step's rebuild decision (or null)
    int par;                          // its parity
    Grid g;
    int n, hyb_axis, chunk, vparts;
    float pirc, cF, lF;
    double kick;
};

__device__ __forceinline__ int wrap_cell(const Grid& g, int x, int y, int z) {
    x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    y = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
    z = z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z);
    return x + g.nx * (y + g.ny * z);
}

// lps lanes per home site (16, 8 or 4 by the cell's occupancy) stride over the site's
// list; each lane holds the site's angular state in registers and reads the radial
// window of a pair from the staged state in LDS.  c = dE_i/dd_ij: +c is summed in
// registers and combined inside the lane group once per site, -c goes to j's slot of
// the LDS accumulator (cell direction and rank carried by the list entry).
// The accumulator is fixed point (kFixScale per unit, 32-bit LDS integer adds, which
// run far faster than LDS float adds and make the sums independent of their order).
// Every add is below kFixMax in magnitude and a slot takes at most kRows + 1 adds (one
// per home site listing it, plus its own sum when it is a home site), so no sum can
// wrap.  A block that meets a larger or non-finite value (each lane folds the range test
// into a flag, the adds themselves run unconditionally) sets `big` and runs its pairs
// again with float adds.
// VIR: also the virial sum over directed pairs of -(r_j - r_i) (x) c (c is the force on i
// from E_i through the pair; -c acts on j), per-block partials in a.vpart.
constexpr int kSurForceThreads = 256;
constexpr float kFixScale = 1048576.f;  // 2^20
constexpr float kFixMax = 16.f;         // kFixMax * kFixScale = 2^24
static_assert((kRows + 1) * (double)kFixMax * kFixScale < 2147483648.0, "fixed-point slot sum can wrap");
__device__ __forceinline__ bool fix_ok(float x, float y, float z) {
    return fabsf(x) < kFixMax && fabsf(y) < kFixMax && fabsf(z) < kFixMax;  // false for NaN
}
template <bool VIR>
__global__ __launch_bounds__(kSurForceThreads) void k_sur_force(ForceArgs a) {
    __shared__ int nb_start[27], nb_cnt[27], nb_base[28];
    __shared__ union {
        int i[3][kSurM];
        float f[3][kSurM];
    } acc;
    __shared__ float4 hst4[kRows * NS / 4];
    __shared__ int bad, big;
    const int tid = threadIdx.x, b = blockIdx.x;
    const int cell = (b & 7) * a.chunk + (b >> 3);  // XCD-aware: contiguous slab of cells per XCD
    if (cell >= a.g.ncell) {
        if (VIR && tid < kSym) a.vpart[(size_t)tid * a.vparts + b] = 0.0;
        return;
    }
    const int hx = cell % a.g.nx, hy = (cell / a.g.nx) % a.g.ny, hz = cell / (a.g.nx * a.g.ny);
    if (tid < 64) {  // wave 0: the 27 neighbour cells, the prefix of their counts, capacity check
        int cnt = 0;
        if (tid < 27) {
            const int nc = wrap_cell(a.g, hx + tid % 3 - 1, hy + (tid / 3) % 3 - 1, hz + tid / 9 - 1);
            cnt = (int)(a.start[nc + 1] - a.start[nc]);
            nb_start[tid] = (int)a.start[nc];
            nb_cnt[tid] = cnt;
        }
        int incl = cnt;  // inclusive scan across the wave
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            const int o = __shfl_up(incl, off);
            if (tid >= off) incl += o;
        }
        if (tid < 27) nb_base[tid + 1] = incl;
        const bool over = __ballot(tid < 27 && cnt > kRows) != 0ull || __shfl(incl, 26) > kSurM;
        if (tid == 0) {
            nb_base[0] = 0;
            bad = over;
            big = 0;
            if (over) atomicOr(a.flags, FLAG_CELL_OVERFLOW);
        }
    }
    __syncthreads();
    if (bad) {
        if (VIR && tid < kSym) a.vpart[(size_t)tid * a.vparts + b] = 0.0;
        return;
    }
    const int m = nb_base[27], hs = nb_start[13], nh = nb_cnt[13], self = nb_base[13];
    float wv[kSym];
    for (int e = tid; e < m; e += kSurForceThreads) acc.i[0][e] = acc.i[1][e] = acc.i[2][e] = 0;
    // home-site state: dE/dD_radial (32) and, per angular channel q, the 20 slots of
    // dE/dT_q = sum over the invariants <T_a^L, T_b^L> containing q of dE/dinv * T_partner
    for (int e = tid; e < nh * (NR / 4); e += kSurForceThreads) {
        const int h = e / (NR / 4), k4 = e % (NR / 4);
        hst4[h * (NS / 4) + k4] = reinterpret_cast<const float4*>(a.S + (size_t)(hs + h) * ND)[k4];
    }
    for (int e = tid; e < nh * NA; e += kSurForceThreads) {
        const int h = e / NA, q = e % NA;
        const float* Tm = a.T + (size_t)(hs + h) * NT;
        const float* ga = a.S + (size_t)(hs + h) * ND + NR;
        float A[20];
#pragma unroll
        for (int c = 0; c < 20; c++) A[c] = 0.f;
#pragma unroll
        for (int m = 0; m < 16; m++) {
            int ia, ib, L;
            inv_of(m, ia, ib, L);
            const float g = ga[m];
            const float w = ia == q ? g : 0.f, w2 = ib == q ? g : 0.f;
#pragma unroll
            for (int c = 0; c < 10; c++)
                if (c < l_cnt(L)) {
                    const int cc = l_off(L) + c;
                    A[cc] += w * Tm[ib * NC + cc] + w2 * Tm[ia * NC + cc];
                }
        }
        float4* ar = hst4 + h * (NS / 4) + (NR + 20 * q) / 4;
#pragma unroll
        for (int k = 0; k < 5; k++) ar[k] = make_float4(A[4 * k], A[4 * k + 1], A[4 * k + 2], A[4 * k + 3]);
    }
    __syncthreads();
    const float* hst = reinterpret_cast<const float*>(hst4);
    const int lps = nh <= 16 ? 16 : (nh <= 32 ? 8 : 4);
    const int per = kSurForceThreads / lps;
    // one pass over the home sites' pairs; FIXED: fixed-point adds, else float adds
    auto pairs = [&](auto fixed_tag) {
        constexpr bool fixed = decltype(fixed_tag)::value;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < kSym; k++) wv[k] = 0.f;
        for (int h0 = 0; h0 < nh; h0 += per) {
            const int h = h0 + tid / lps, sub = tid % lps;
            const bool live = h < nh;
            float fx = 0.f, fy = 0.f, fz = 0.f;
            if (live) {
                const float* si = hst + h * NS;
                float A[80];
                {
                    const float4* ar = reinterpret_cast<const float4*>(si + NR);
#pragma unroll
                    for (int k = 0; k < 20; k++) {
                        const float4 t4 = ar[k];
                        A[4 * k] = t4.x; A[4 * k + 1] = t4.y; A[4 * k + 2] = t4.z; A[4 * k + 3] = t4.w;
                    }
                }
                const int p = hs + h;
                int nn = (int)a.nlc[p];
                if (a.dlam && a.lam[a.sorted[p]] == 1.f) nn = 0;  // hybrid: no surrogate energy on this site
                const float4* row = a.nl + (size_t)p * kFullCap;
                const uint16_t* drow = a.nld + (size_t)p * kFullCap;
                // the next entry is loaded one iteration ahead of its use
                float4 vnext = sub < nn ? row[sub] : make_float4(1.f, 0.f, 0.f, 0.f);
                uint32_t dnext = sub < nn ? drow[sub] : 0u;
                for (int s = sub; s < nn; s += lps) {
                    const float4 v = vnext;
                    const uint32_t dr = dnext;
                    if (s + lps < nn) {
                        vnext = row[s + lps];
                        dnext = drow[s + lps];
                    }
                    const Pair q = pair_of(v, a.pirc);
                    // radial: sum_k S_k d(G_k fc)/dr = dfc sa - 2 fc / w ((y - k0) sa - sc),
                    // sa = sum_t S_{k0+t} g_t, sc = sum_t t S_{k0+t} g_t
                    const float y = (q.r - kRMu0) * kRInvW;
                    const int k0 = win_start(y);
                    const float ew = y - (float)k0;
                    float g[kWin];
                    win_gauss(ew, g);
                    float sa = 0.f, sc = 0.f;
#pragma unroll
                    for (int t = 0; t < kWin; t++) {
                        const float w = si[k0 + t] * g[t];
                        sa += w;
                        sc += (float)t * w;
                    }
                    const float coef = q.dfc * sa - 2.f * kRInvW * q.fc * (ew * sa - sc);
                    float cx = coef * q.ux, cy = coef * q.uy, cz = coef * q.uz;
#pragma unroll
                    for (int ch = 0; ch < NA; ch++) {
                        float psi, dpsi;
                        ang_basis(q, ch, psi, dpsi);
                        chan_grad(A + 20 * ch, q.ux, q.uy, q.uz, dpsi, psi * q.ir, cx, cy, cz);
                    }
                    fx += cx;
                    fy += cy;
                    fz += cz;
                    if (VIR) sym_outer<float>(wv, -v.x, -v.y, -v.z, cx, cy, cz);
                    const int d = (int)(dr & 31u), rk = (int)(dr >> 5);
                    if (rk < nb_cnt[d]) {  // lists and cells come from one binning
                        const int slot = nb_base[d] + rk;
                        if constexpr (!fixed) {
                            atomicAdd(&acc.f[0][slot], -cx);
                            atomicAdd(&acc.f[1][slot], -cy);
                            atomicAdd(&acc.f[2][slot], -cz);
                        } else {  // out of range: the sums of this pass are discarded
                            ok = ok & fix_ok(cx, cy, cz);
                            atomicAdd(&acc.i[0][slot], __float2int_rn(-cx * kFixScale));
                            atomicAdd(&acc.i[1][slot], __float2int_rn(-cy * kFixScale));
                            atomicAdd(&acc.i[2][slot], __float2int_rn(-cz * kFixScale));
                        }
                    }
                }
            }
            fx = group_sum(fx, lps);  // lps = 16, 8 or 4
            fy = group_sum(fy, lps);
            fz = group_sum(fz, lps);
            if (live && sub == 0) {
                if constexpr (!fixed) {
                    atomicAdd(&acc.f[0][self + h], fx);
                    atomicAdd(&acc.f[1][self + h], fy);
                    atomicAdd(&acc.f[2][self + h], fz);
                } else {
                    ok = ok & fix_ok(fx, fy, fz);
                    atomicAdd(&acc.i[0][self + h], __float2int_rn(fx * kFixScale));
                    atomicAdd(&acc.i[1][self + h], __float2int_rn(fy * kFixScale));
                    atomicAdd(&acc.i[2][self + h], __float2int_rn(fz * kFixScale));
                }
            }
        }
        if (!ok) big = 1;
    };
    pairs(std::true_type{});
    __syncthreads();
    if (big) {  // block-uniform: read after the barrier, written only in the fixed pass
        for (int e = tid; e < m; e += kSurForceThreads) acc.f[0][e] = acc.f[1][e] = acc.f[2][e] = 0.f;
        __syncthreads();
        pairs(std::false_type{});
        __syncthreads();
    }
    // the rows out in one flat pass over (direction, rank); a wave covers one direction
    static_assert(kRows == 64 && kSurForceThreads % kRows == 0, "one wave per direction row");
    float4* out = a.P + (size_t)cell * 27 * kRows;
    const bool as_int = !big;
    for (int idx = tid; idx < 27 * kRows; idx += kSurForceThreads) {
        const int d = idx / kRows, r = idx % kRows;
        if (r < nb_cnt[d]) {
            const int e = nb_base[d] + r;
            out[idx] = as_int ? make_float4((float)acc.i[0][e] * (1.f / kFixScale), (float)acc.i[1][e] * (1.f / kFixScale),
                                            (float)acc.i[2][e] * (1.f / kFixScale), 0.f)
                              : make_float4(acc.f[0][e], acc.f[1][e], acc.f[2][e], 0.f);
        }
    }
    if (VIR) block_store_sym(a.vpart, a.vparts, b, wv);
}

// dst[k * stride] = src ? src[k] : 0 for k < kSym (the extra virial slot of a pass).
__global__ void k_set_sym_slot(double* __restrict__ dst, int stride, const double* __restrict__ src) {
    const int k = threadIdx.x;
    if (k < kSym) dst[(size_t)k * stride] = src ? src[k] : 0.0;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_sur_apply(ForceArgs a) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    double lsum = 0.0;
    if (p < a.n) {
        const uint32_t oi = a.sorted[p];
        const int cb = (int)a.key[oi];
        const int r = p - (int)a.start[cb];
        const int bx = cb % a.g.nx, by = (cb / a.g.nx) % a.g.ny, bz = cb / (a.g.nx * a.g.ny);
        float4 F = make_float4(0.f, 0.f, 0.f, 0.f);
        if (r < kRows) {  // larger cells are reported by the force pass
#pragma unroll
            for (int d = 0; d < 27; d++) {
                // the cell that sees this one in direction d lies at offset -d
                const int ca = wrap_cell(a.g, bx - (d % 3 - 1), by - ((d / 3) % 3 - 1), bz - (d / 9 - 1));
                const float4 t = a.P[((size_t)ca * 27 + d) * kRows + r];
                F.x += t.x;
                F.y += t.y;
                F.z += t.z;
            }
        }
        if (MODE == FORCE_APPLY) {
            double dx = F.x, dy = F.y, dz = F.z;
            if (a.dlam) {  // hybrid: E_i grad(lambda_i) from the weight 1 - lambda_i of E_i
                const double t = (double)a.E[p] * (double)a.dlam[oi];
                if (a.hyb_axis == 0) dx += t; else if (a.hyb_axis == 1) dy += t; else dz += t;
            }
            const double nfx = a.fx[oi] + dx, nfy = a.fy[oi] + dy, nfz = a.fz[oi] + dz;
            a.fx[oi] = nfx;
            a.fy[oi] = nfy;
            a.fz[oi] = nfz;
            if (a.kick > 0.0) {
                kick_site(a.vx, a.vy, a.vz, a.mass, oi, a.kick, nfx, nfy, nfz);
                const SkinArgs* sk = a.sk;
                if (sk && skin_exceeded(a.vx, a.vy, a.vz, a.mass, oi, a.kick, sk->dt, nfx, nfy, nfz, sk->spos[p],
                                        sk->spb[p], sk->lim2))
                    atomicOr(sk->req + (a.par ^ 1), 1u);
            }
        } else if (MODE == FORCE_OUT) {
            a.fout[3 * (size_t)oi] = F.x;
            a.fout[3 * (size_t)oi + 1] = F.y;
            a.fout[3 * (size_t)oi + 2] = F.z;
        } else {
            const float ex = F.x - a.fref[3 * (size_t)oi], ey = F.y - a.fref[3 * (size_t)oi + 1],
                        ez = F.z - a.fref[3 * (size_t)oi + 2];
            a.delta[p] = make_float4(a.cF * ex, a.cF * ey, a.cF * ez, 0.f);
            lsum = (double)a.lF * (double)(ex * ex + ey * ey + ez * ez);
        }
    }
    if (MODE == FORCE_TRAIN) {
        __shared__ double red[4];
        for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsum;
        __syncthreads();
        if (threadIdx.x == 0) a.part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
}

// ---------------------------------------------------------------------------
// training: descriptor tangents along delta, w_a = -scale * dD_a
struct JvpArgs {
    const float4* __restrict__ nl;
    const uint32_t* __restrict__ nlc;
    const uint32_t* __restrict__ inv;
    const float* __restrict__ T;
    const float4* __restrict__ delta;
    const float* __restrict__ params;
    float* __restrict__ W;            // [n][48]
    int n;
    float pirc;
};

__global__ __launch_bounds__(256) void k_sur_jvp(JvpArgs a) {
    const int tid = blockIdx.x * 256 + threadIdx.x;
    const int p = tid / kLps, sub = tid % kLps;
    if (p >= a.n) return;
    float dR[NR], dM[NT];
#pragma unroll
    for (int k = 0; k < NR; k++) dR[k] = 0.f;
#pragma unroll
    for (int k = 0; k < NT; k++) dM[k] = 0.f;
    const float4 dp = a.delta[p];
    const int nn = (int)a.nlc[p];
    const float4* row = a.nl + (size_t)p * kFullCap;
    float4 vnext = sub < nn ? row[sub] : make_float4(1.f, 0.f, 0.f, 0.f);
    for (int s = sub; s < nn; s += kLps) {
        const float4 v = vnext;
        if (s + kLps < nn) vnext = row[s + kLps];
        const uint32_t pj = a.inv[__float_as_uint(v.w) & kIdxMask];
        const float4 dj = a.delta[pj];
        const Pair q = pair_of(v, a.pirc);
        const float dvx = dj.x - dp.x, dvy = dj.y - dp.y, dvz = dj.z - dp.z;
        const float ud = q.ux * dvx + q.uy * dvy + q.uz * dvz;
        const float y = (q.r - kRMu0) * kRInvW;
        float g[NR];
        radial_gauss(y, g);
#pragma unroll
        for (int k = 0; k < NR; k++) dR[k] += g[k] * (q.dfc - 2.f * (y - (float)k) * kRInvW * q.fc) * ud;
        float t[NC], w[NC];
        comps(q.ux, q.uy, q.uz, t);
        sym_comps(q.ux, q.uy, q.uz, dvx - ud * q.ux, dvy - ud * q.uy, dvz - ud * q.uz, w);
#pragma unroll
        for (int ch = 0; ch < NA; ch++) {
            float psi, dpsi;
            ang_basis(q, ch, psi, dpsi);
            const float c1 = dpsi * ud, c2 = psi * q.ir;
#pragma unroll
            for (int c = 0; c < NC; c++) dM[ch * NC + c] += c1 * t[c] + c2 * w[c];
        }
    }
#pragma unroll
    for (int o = 1; o < kLps; o <<= 1) {
#pragma unroll
        for (int k = 0; k < NR; k++) dR[k] += __shfl_xor(dR[k], o);
#pragma unroll
        for (int k = 0; k < NT; k++) dM[k] += __shfl_xor(dM[k], o);
    }
    const float4* tr = reinterpret_cast<const float4*>(a.T + (size_t)p * NT);
    float Tm[NT];
#pragma unroll
    for (int k = 0; k < NT / 4; k++) {
        const float4 t = tr[k];
        Tm[4 * k] = t.x; Tm[4 * k + 1] = t.y; Tm[4 * k + 2] = t.z; Tm[4 * k + 3] = t.w;
    }
    float dD[ND];
#pragma unroll
    for (int k = 0; k < NR; k++) dD[k] = dR[k];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        int ia, ib, L;
        inv_of(m, ia, ib, L);
        float sm = 0.f;
#pragma unroll
        for (int c = 0; c < 10; c++)
            if (c < l_cnt(L)) {
                const int cc = l_off(L) + c;
                sm += mult(cc) * (dM[ia * NC + cc] * Tm[ib * NC + cc] + Tm[ia * NC + cc] * dM[ib * NC + cc]);
            }
        dD[NR + m] = sm;
    }
    float4* wo = reinterpret_cast<float4*>(a.W + (size_t)p * ND);
#pragma unroll
    for (int k = 0; k < ND / 4; k++) {
        if (k % kLps != sub) continue;
        float o[4];
#pragma unroll
        for (int t = 0; t < 4; t++) o[t] = -a.params[O_SCALE + 4 * k + t] * dD[4 * k + t];
        wo[k] = make_float4(o[0], o[1], o[2], o[3]);
    }
}

// ---------------------------------------------------------------------------
// training: weight gradients
//
// Per 16-site tile (lane = site, kappa register order):
//   forward z1 h1 z2 h2 E;  eps = dL/dE = cE (E - E*)
//   tangent  u = W1 w, dh1 = d1 u, dz2 = W2 dh1            (MFMA)
//   signals  t2 = d2 w3, e2 = w3 dd2 dz2, s2 = eps w3 d2 + e2,
//            q1 = W2^T t2, back = W2^T s2 (MFMA), s1 = d1 back + dd1 u q1, delta1 = d1 q1
//   grads    dW2 += s2 h1^T + t2 dh1^T,  dW1 += s1 x^T + delta1 w^T   (MFMA, K = 32 sites)
//            db2 += s2, db1 += s1, dw3 += eps h2 + d2 dz2, db3 += eps  (VALU)
// The unit-major BF16 tables of a 32-site chunk live in LDS; after the loop the same
// LDS holds the workgroup's gradient sum, written as one partial row.
struct TrainArgs {
    const bf8* __restrict__ ops;
    const float* __restrict__ params;
    const float* __restrict__ X;
    const float* __restrict__ W;
    const float* __restrict__ eref;    // exact per-site energies, site order
    const uint32_t* __restrict__ sorted;
    float* __restrict__ gpart;         // [gridDim][NPARAM]
    double* __restrict__ lpart;        // [gridDim] energy-loss partials
    int n;
    float cE, lE;
};

constexpr int kTabU = 64 * 32;  // one unit-major table: 64 units x 32 sites (BF16)
constexpr int kTrainLds = NOPS * 16 + 4 * 8 * kTabU * 2;  // ops + 4 waves x 8 tables

__global__ __launch_bounds__(256) void k_sur_train(TrainArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    bf8* ops = reinterpret_cast<bf8*>(lds);
    __bf16* tabs = reinterpret_cast<__bf16*>(lds + NOPS * 16);
    for (int e = threadIdx.x; e < NOPS; e += 256) ops[e] = a.ops[e];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, q = lane >> 4, si = lane & 15;
    // this wave's tables: 0 s2, 1 t2, 2 s1, 3 delta1 (A side), 4 h1, 5 dh1, 6 x, 7 w (B side)
    __bf16* tb = tabs + (size_t)wv * 8 * kTabU;
    LaneVecs v;
    lane_vecs(a.params, q, v);
    const float b3 = a.params[O_B3];
    f4 aW2[16], aW1[12];
#pragma unroll
    for (int k = 0; k < 16; k++) aW2[k] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 12; k++) aW1[k] = f4{0.f, 0.f, 0.f, 0.f};
    float gb1[16], gb2[16], gw3[16], gb3 = 0.f;
#pragma unroll
    for (int k = 0; k < 16; k++) gb1[k] = gb2[k] = gw3[k] = 0.f;
    double lacc = 0.0;
    const int nchunks = (a.n + 31) >> 5;
    for (int ch = blockIdx.x * 4 + wv; ch < nchunks; ch += gridDim.x * 4) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int p = ch * 32 + h * 16 + si;
            const bool ok = p < a.n;
            const size_t pr = (size_t)(ok ? p : 0);
            bf8 xb[2], wb[2];
            xb[0] = load8(a.X + pr * ND + 8 * q, ok);
            xb[1] = load8(a.X + pr * ND + 32 + 8 * q, ok && q < 2);
            wb[0] = load8(a.W + pr * ND + 8 * q, ok);
            wb[1] = load8(a.W + pr * ND + 32 + 8 * q, ok && q < 2);
            TileFwd o;
            tile_forward(ops, lane, xb, v, b3, o);
            float eps = 0.f;
            if (ok) {
                const float de = o.e - a.eref[a.sorted[p]];
                eps = a.cE * de;
                if (q == 0) lacc += (double)a.lE * (double)(de * de);
            }
            // tangent
            f4 u[4], dh1[4], dz2[4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(ops[kOp1 + (2 * m) * 64 + lane], wb[0], acc);
                acc = mfma(ops[kOp1 + (2 * m + 1) * 64 + lane], wb[1], acc);
                u[m] = acc;
                dh1[m] = o.d1[m] * acc;
            }
            const bf8 hb0 = pack2(dh1[0], dh1[1]), hb1 = pack2(dh1[2], dh1[3]);
            f4 s2[4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(ops[kOp2 + (2 * m) * 64 + lane], hb0, acc);
                acc = mfma(ops[kOp2 + (2 * m + 1) * 64 + lane], hb1, acc);
                dz2[m] = acc;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float w3 = v.w3[4 * m + r];
                    s2[m][r] = eps * w3 * o.d2[m][r] + w3 * o.dd2[m][r] * acc[r];
                    gw3[4 * m + r] += eps * o.h2[m][r] + o.d2[m][r] * acc[r];
                    gb2[4 * m + r] += s2[m][r];
                }
            }
            if (q == 0) gb3 += eps;
            const bf8 sb0 = pack2(s2[0], s2[1]), sb1 = pack2(s2[2], s2[3]);
            f4 s1[4], dl1[4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                f4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma(ops[kOp3 + (2 * m) * 64 + lane], sb0, acc);
                acc = mfma(ops[kOp3 + (2 * m + 1) * 64 + lane], sb1, acc);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    s1[m][r] = o.d1[m][r] * acc[r] + o.dd1[m][r] * u[m][r] * o.q1[m][r];
                    dl1[m][r] = o.d1[m][r] * o.q1[m][r];
                    gb1[4 * m + r] += s1[m][r];
                }
            }
            // unit-major tables, column = site of the chunk
            const int col = h * 16 + si;
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int uu = kappa(16 * m + 4 * q + r) * 32 + col;
                    tb[0 * kTabU + uu] = (__bf16)s2[m][r];
                    tb[1 * kTabU + uu] = (__bf16)o.t2[m][r];
                    tb[2 * kTabU + uu] = (__bf16)s1[m][r];
                    tb[3 * kTabU + uu] = (__bf16)dl1[m][r];
                    tb[4 * kTabU + uu] = (__bf16)o.h1[m][r];
                    tb[5 * kTabU + uu] = (__bf16)dh1[m][r];
                }
#pragma unroll
            for (int kk = 0; kk < 2; kk++)
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const int k = 32 * kk + 8 * q + t;
                    if (k < ND) {
                        tb[6 * kTabU + k * 32 + col] = xb[kk][t];
                        tb[7 * kTabU + k * 32 + col] = wb[kk][t];
                    }
                }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // weight gradients over the 32 sites: A[unit][site], B[site][unit] from the tables
        const int i = lane & 15;
        const bf8* t8 = reinterpret_cast<const bf8*>(tb);  // a row of 32 sites = 4 bf8
#pragma unroll
        for (int M = 0; M < 4; M++) {
            const bf8 As2 = t8[(0 * kTabU + (16 * M + i) * 32) / 8 + q];
            const bf8 At2 = t8[(1 * kTabU + (16 * M + i) * 32) / 8 + q];
            const bf8 As1 = t8[(2 * kTabU + (16 * M + i) * 32) / 8 + q];
            const bf8 Ad1 = t8[(3 * kTabU + (16 * M + i) * 32) / 8 + q];
#pragma unroll
            for (int N = 0; N < 4; N++) {
                const bf8 Bh1 = t8[(4 * kTabU + (16 * N + i) * 32) / 8 + q];
                const bf8 Bdh = t8[(5 * kTabU + (16 * N + i) * 32) / 8 + q];
                aW2[4 * M + N] = mfma(As2, Bh1, aW2[4 * M + N]);
                aW2[4 * M + N] = mfma(At2, Bdh, aW2[4 * M + N]);
            }
#pragma unroll
            for (int N = 0; N < 3; N++) {
                const bf8 Bx = t8[(6 * kTabU + (16 * N + i) * 32) / 8 + q];
                const bf8 Bw = t8[(7 * kTabU + (16 * N + i) * 32) / 8 + q];
                aW1[3 * M + N] = mfma(As1, Bx, aW1[3 * M + N]);
                aW1[3 * M + N] = mfma(Ad1, Bw, aW1[3 * M + N]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // vector gradients: sum over the 16 sites of a lane group
#pragma unroll
    for (int k = 0; k < 16; k++) {
        for (int o = 1; o < 16; o <<= 1) {
            gb1[k] += __shfl_xor(gb1[k], o);
            gb2[k] += __shfl_xor(gb2[k], o);
            gw3[k] += __shfl_xor(gw3[k], o);
        }
    }
    for (int o = 1; o < 64; o <<= 1) gb3 += __shfl_xor(gb3, o);
    for (int o = 1; o < 64; o <<= 1) lacc += __shfl_xor(lacc, o);
    // workgroup sum in LDS (reusing the tables), then one partial row
    __syncthreads();
    float* acc = reinterpret_cast<float*>(lds + NOPS * 16);
    for (int e = threadIdx.x; e < NPARAM; e += 256) acc[e] = 0.f;
    __shared__ double lred[4];
    __syncthreads();
#pragma unroll
    for (int M = 0; M < 4; M++) {
#pragma unroll
        for (int N = 0; N < 4; N++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                atomicAdd(&acc[O_W2 + (16 * M + 4 * q + r) * NH + 16 * N + (lane & 15)], aW2[4 * M + N][r]);
#pragma unroll
        for (int N = 0; N < 3; N++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                atomicAdd(&acc[O_W1 + (16 * M + 4 * q + r) * ND + 16 * N + (lane & 15)], aW1[3 * M + N][r]);
    }
    if (si == 0) {
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int u = kappa(16 * m + 4 * q + r);
                atomicAdd(&acc[O_B1 + u], gb1[4 * m + r]);
                atomicAdd(&acc[O_B2 + u], gb2[4 * m + r]);
                atomicAdd(&acc[O_W3 + u], gw3[4 * m + r]);
            }
    }
    if (lane == 0) {
        atomicAdd(&acc[O_B3], gb3);
        lred[wv] = lacc;
    }
    __syncthreads();
    float* dst = a.gpart + (size_t)blockIdx.x * NPARAM;
    for (int e = threadIdx.x; e < NPARAM; e += 256) dst[e] = e < O_SHIFT ? acc[e] : 0.f;
    if (threadIdx.x == 0) a.lpart[blockIdx.x] = lred[0] + lred[1] + lred[2] + lred[3];
}

// Fixed-order sum of the partial rows, optional gradient copy, Adam on the weights;
// block 0 also folds this step's loss partials into the running loss sum.
__global__ __launch_bounds__(256) void k_sur_adam(float* __restrict__ prm, float* __restrict__ mom,
                                                  float* __restrict__ vel, const float* __restrict__ gpart,
                                                  int nrows, float* __restrict__ grad_out, float lr, float bc1,
                                                  float bc2, int update, const double* __restrict__ lpart_e,
                                                  int ne, const double* __restrict__ lpart_f, int nf,
                                                  double* __restrict__ lacc) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < O_SHIFT) {
        float g = 0.f;
        for (int b = 0; b < nrows; b++) g += gpart[(size_t)b * NPARAM + i];
        if (grad_out) grad_out[i] = g;
        if (update) {
            const float m = 0.9f * mom[i] + 0.1f * g;
            const float v = 0.999f * vel[i] + 0.001f * g * g;
            mom[i] = m;
            vel[i] = v;
            prm[i] -= lr * (m * bc1) / (sqrtf(v * bc2) + 1e-8f);
        }
    } else if (i < NPARAM && grad_out) {
        grad_out[i] = 0.f;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < ne; b++) s += lpart_e[b];
        for (int b = 0; b < nf; b++) s += lpart_f[b];
        lacc[0] += s;
        lacc[1] = s;
    }
}

// three-body force error of the surrogate against the exact forces in f (judge)
__global__ __launch_bounds__(256) void k_sur_judge(int n, const float* __restrict__ fs, const double* __restrict__ fx,
                                                   const double* __restrict__ fy, const double* __restrict__ fz,
                                                   double* __restrict__ part) {
    double se = 0, sr = 0, mx = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double ex = fs[3 * (size_t)i] - fx[i], ey = fs[3 * (size_t)i + 1] - fy[i],
                     ez = fs[3 * (size_t)i + 2] - fz[i];
        const double e2 = ex * ex + ey * ey + ez * ez;
        se += e2;
        sr += fx[i] * fx[i] + fy[i] * fy[i] + fz[i] * fz[i];
        mx = fmax(mx, e2);
    }
    __shared__ double r[3][256];
    r[0][threadIdx.x] = se;
    r[1][threadIdx.x] = sr;
    r[2][threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r[0][threadIdx.x] += r[0][threadIdx.x + o];
            r[1][threadIdx.x] += r[1][threadIdx.x + o];
            r[2][threadIdx.x] = fmax(r[2][threadIdx.x], r[2][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = r[0][0];
        part[3 * blockIdx.x + 1] = r[1][0];
        part[3 * blockIdx.x + 2] = r[2][0];
    }
}

// FP64 site-order forces -> FP32 AoS (training targets)
__global__ __launch_bounds__(256) void k_f64_to_aos(int n, const double* __restrict__ fx, const double* __restrict__ fy,
                                                    const double* __restrict__ fz, float* __restrict__ out,
                                                    const double* __restrict__ e, float* __restrict__ eout) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i] = (float)fx[i];
    out[3 * (size_t)i + 1] = (float)fy[i];
    out[3 * (size_t)i + 2] = (float)fz[i];
    eout[i] = (float)e[i];
}

// per-block FP64 sums of x and x^2 of every raw descriptor
__global__ __launch_bounds__(256) void k_sur_colstats(int n, const float* __restrict__ X, double* __restrict__ part) {
    const int k = threadIdx.x;  // one column per thread (k < 96: sum, sum of squares)
    if (k >= 2 * ND) return;
    const int col = k % ND;
    double s = 0.0;
    for (int p = blockIdx.x; p < n; p += gridDim.x) {
        const double x = X[(size_t)p * ND + col];
        s += k < ND ? x : x * x;
    }
    part[(size_t)blockIdx.x * 2 * ND + k] = s;
}

// hybrid weights: lambda = 1 inside |x - c| < h (minimum image), cos^2 descent to 0 over b
__global__ __launch_bounds__(256) void k_hyb_weights(int n, const double* __restrict__ r, double c, double h,
                                                     double b, double L, float* __restrict__ lam,
                                                     float* __restrict__ dlam) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double d = r[i] - c;
    d -= L * rint(d / L);
    const double ad = fabs(d);
    double l = 0.0, dl = 0.0;
    if (ad <= h) {
        l = 1.0;
    } else if (ad < h + b) {
        const double t = M_PI * 0.5 * (ad - h) / b;
        const double cs = cos(t), sn = sin(t);
        l = cs * cs;
        dl = -2.0 * cs * sn * (M_PI * 0.5 / b) * (d > 0 ? 1.0 : -1.0);
    }
    lam[i] = (float)l;
    dlam[i] = (float)dl;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
struct Surrogate {
    float* params = nullptr;
    bf8* ops = nullptr;
    bool have_params = false;
    int64_t n = 0;
    float *X = nullptr, *T = nullptr, *S = nullptr, *E = nullptr;
    // training
    float *mom = nullptr, *vel = nullptr, *gpart = nullptr, *grad = nullptr, *W = nullptr, *fj = nullptr;
    float4* delta = nullptr;
    double *lpart = nullptr, *lacc = nullptr, *save = nullptr, *stats = nullptr;
    int64_t step = 0;
    std::vector<double*> snap_r;
    std::vector<float*> snap_f, snap_e;
    // hybrid
    float *lam = nullptr, *dlam = nullptr;
};

namespace {

constexpr int kTrainBlocks = 64;
constexpr int kJudgeBlocks = 256;

template <typename T>
hipError_t salloc(T** p, size_t count) {
    if (*p) return hipSuccess;
    return hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
}

Surrogate* sur_of(komd_ctx* c) {
    if (!c->sur) c->sur = new Surrogate();
    return c->sur;
}

// Frees every buffer whose size follows the site count, and the training snapshots.
void free_site_buffers(Surrogate* s) {
    void* ptrs[] = {s->X, s->T, s->S, s->E, s->W, s->delta, s->fj, s->save, s->lam, s->dlam, s->lpart};
    for (void* p : ptrs)
        if (p) hipFree(p);
    s->X = s->T = s->S = s->E = s->W = s->fj = s->lam = s->dlam = nullptr;
    s->delta = nullptr;
    s->save = s->lpart = nullptr;
    for (auto p : s->snap_r) hipFree(p);
    for (auto p : s->snap_f) hipFree(p);
    for (auto p : s->snap_e) hipFree(p);
    s->snap_r.clear();
    s->snap_f.clear();
    s->snap_e.clear();
}

hipError_t ensure_site_buffers(komd_ctx* c) {
    Surrogate* s = sur_of(c);
    if (s->n != c->n) {
        free_site_buffers(s);
        s->n = c->n;
    }
    const size_t n = (size_t)c->n;
    KOMD_CHECK_HIP(salloc(&s->X, n * ND));
    KOMD_CHECK_HIP(salloc(&s->T, n * NT));
    KOMD_CHECK_HIP(salloc(&s->S, n * ND));
    KOMD_CHECK_HIP(salloc(&s->E, n));
    KOMD_CHECK_HIP(salloc(&s->params, NPARAM));
    KOMD_CHECK_HIP(salloc(&s->ops, NOPS));
    return hipSuccess;
}

int mlp_blocks(komd_ctx* c) {
    const int64_t tiles = (c->n + 15) / 16;
    return (int)std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 512));
}
int site_blocks(komd_ctx* c) { return (int)((c->n + 255) / 256); }
int group_blocks(komd_ctx* c) { return (int)((c->n * kLps + 255) / 256); }  // kLps lanes per site
float pirc_of(komd_ctx* c) { return (float)(M_PI / c->prm.cutoff); }

hipError_t launch_desc(komd_ctx* c, bool raw, const float* lam = nullptr) {
    ScopedTimer tm(c, "sur_desc");
    Surrogate* s = c->sur;
    DescArgs a{c->d.nl, c->d.nlc, raw ? nullptr : s->params, s->X, s->T, lam, c->d.sorted, (int)c->n, pirc_of(c)};
    a.fx = c->d.fx; a.fy = c->d.fy; a.fz = c->d.fz;
    a.lj = c->lj_in_desc && !raw && !lam;
    a.sig2 = c->sig_u * c->sig_u;
    a.eps24 = 24.f * c->eps_u;
    k_sur_desc<<<(int)((c->n + kDescSites - 1) / kDescSites), 256, 0, c->stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_mlp(komd_ctx* c, bool energy, const float* lam) {
    ScopedTimer tm(c, "sur_mlp");
    Surrogate* s = c->sur;
    MlpArgs a{s->ops, s->params, s->X, s->T, c->d.sorted, lam, s->S, s->E, energy ? c->d.red : nullptr, (int)c->n};
    k_sur_mlp<<<mlp_blocks(c), 256, 0, c->stream>>>(a);
    return hipGetLastError();
}

ForceArgs force_args(komd_ctx* c) {
    Surrogate* s = c->sur;
    ForceArgs a{};
    a.nl = c->d.nl; a.nlc = c->d.nlc; a.nld = c->d.nld; a.inv = c->d.inv; a.sorted = c->d.sorted;
    a.key = c->d.key; a.start = c->d.start; a.S = s->S; a.T = s->T; a.E = s->E; a.P = c->d.rows;
    a.flags = c->d.flags;
    a.g = c->grid;
    a.chunk = (c->grid.ncell + 7) / 8;
    a.fx = c->d.fx; a.fy = c->d.fy; a.fz = c->d.fz;
    a.n = (int)c->n;
    a.hyb_axis = c->hyb_axis;
    a.pirc = pirc_of(c);
    return a;
}

// descriptors + network + forces on an up-to-date neighbour list
hipError_t surrogate_pass(komd_ctx* c, bool energy, bool hybrid, int mode) {
    Surrogate* s = c->sur;
    if (!s || !s->have_params) return hipErrorNotSupported;
    KOMD_CHECK_HIP(ensure_site_buffers(c));
    KOMD_CHECK_HIP(launch_desc(c, false, hybrid ? s->lam : nullptr));
    KOMD_CHECK_HIP(launch_mlp(c, energy, hybrid ? s->lam : nullptr));
    ForceArgs a = force_args(c);
    if (hybrid) {
        a.dlam = s->dlam;
        a.lam = s->lam;
    }
    {
        ScopedTimer tm(c, "sur_force");
        if (energy) {
            a.vpart = c->d.vpart;
            a.vparts = surrogate_virial_blocks(c);
            k_sur_force<true><<<8 * a.chunk, kSurForceThreads, 0, c->stream>>>(a);
        } else {
            k_sur_force<false><<<8 * a.chunk, kSurForceThreads, 0, c->stream>>>(a);
        }
        KOMD_CHECK_HIP(hipGetLastError());
    }
    ScopedTimer tm(c, "sur_apply");
    if (mode == FORCE_OUT) {
        a.fout = s->fj;
        k_sur_apply<FORCE_OUT><<<site_blocks(c), 256, 0, c->stream>>>(a);
    } else {
        a.vx = c->d.vx; a.vy = c->d.vy; a.vz = c->d.vz; a.mass = c->d.mass;
        a.kick = (!energy && c->kick_in == KICK_IN_APPLY) ? 0.5 * c->prm.delta_t : 0.0;
        if (a.kick > 0.0 && c->skin_step_par >= 0) {  // Verlet-list step: decide the next rebuild
            a.sk = c->d.skin_args;
            a.par = c->skin_step_par;
        }
        k_sur_apply<FORCE_APPLY><<<site_blocks(c), 256, 0, c->stream>>>(a);
    }
    return hipGetLastError();
}

hipError_t hybrid_weights(komd_ctx* c) {
    Surrogate* s = sur_of(c);
    KOMD_CHECK_HIP(ensure_site_buffers(c));
    KOMD_CHECK_HIP(salloc(&s->lam, (size_t)c->n));
    KOMD_CHECK_HIP(salloc(&s->dlam, (size_t)c->n));
    const double* r = c->hyb_axis == 0 ? c->d.rx : (c->hyb_axis == 1 ? c->d.ry : c->d.rz);
    k_hyb_weights<<<site_blocks(c), 256, 0, c->stream>>>((int)c->n, r, c->hyb_center, c->hyb_half, c->hyb_blend,
                                                         c->grid.size[c->hyb_axis], s->lam, s->dlam);
    return hipGetLastError();
}

}  // namespace

// forces_3b / enqueue_forces entry: the neighbour lists are current.  With energy the
// partials are d.red[0, mlp_blocks) plus one slot for an exact-ATM share (0 here).
hipError_t launch_surrogate_forces(komd_ctx* c, bool energy, int hybrid) {
    (void)hybrid;
    KOMD_CHECK_HIP(surrogate_pass(c, energy, false, FORCE_APPLY));
    if (energy) {
        KOMD_CHECK_HIP(hipMemsetAsync(c->d.red + mlp_blocks(c), 0, sizeof(double), c->stream));
        const int vb = surrogate_virial_blocks(c);
        k_set_sym_slot<<<1, 64, 0, c->stream>>>(c->d.vpart + (vb - 1), vb, nullptr);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    return hipSuccess;
}

// Hybrid three-body pass: exact ATM weighted by the members' mean lambda, surrogate
// site energies weighted by 1 - lambda.  With energy, the ATM share is reduced first
// and placed after the surrogate's partials.
hipError_t hybrid_forces(komd_ctx* c, bool energy) {
    Surrogate* s = c->sur;
    if (!s || !s->have_params || c->hyb_blend <= 0) return hipErrorNotSupported;
    KOMD_CHECK_HIP(hybrid_weights(c));
    KOMD_CHECK_HIP(launch_atm(c, 1, energy, s->lam, s->dlam, nullptr));
    if (energy) {
        KOMD_CHECK_HIP(launch_reduce_parts(c, c->d.red, atm_blocks(c), 1, c->d.scal + 12));
        KOMD_CHECK_HIP(launch_reduce_parts(c, c->d.vpart, atm_blocks(c), kSym, c->d.scal + 16));
    }
    KOMD_CHECK_HIP(surrogate_pass(c, energy, true, FORCE_APPLY));
    if (energy) {
        KOMD_CHECK_HIP(hipMemcpyAsync(c->d.red + mlp_blocks(c), c->d.scal + 12, sizeof(double),
                                      hipMemcpyDeviceToDevice, c->stream));
        const int vb = surrogate_virial_blocks(c);
        k_set_sym_slot<<<1, 64, 0, c->stream>>>(c->d.vpart + (vb - 1), vb, c->d.scal + 16);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    return hipSuccess;
}

int surrogate_energy_blocks(komd_ctx* c) { return mlp_blocks(c) + 1; }
int surrogate_virial_blocks(komd_ctx* c) { return 8 * ((c->grid.ncell + 7) / 8) + 1; }  // k_sur_force blocks + 1

hipError_t surrogate_prepare(komd_ctx* c, bool hybrid) {
    Surrogate* s = c->sur;
    if (!s || !s->have_params || (hybrid && c->hyb_blend <= 0)) return hipErrorNotSupported;
    KOMD_CHECK_HIP(ensure_site_buffers(c));
    if (hybrid) {
        KOMD_CHECK_HIP(salloc(&s->lam, (size_t)c->n));
        KOMD_CHECK_HIP(salloc(&s->dlam, (size_t)c->n));
    }
    return hipSuccess;
}

void surrogate_sites_changed(komd_ctx* c) {
    Surrogate* s = c->sur;
    if (!s || s->n == c->n) return;
    hipStreamSynchronize(c->stream);
    free_site_buffers(s);
    s->n = c->n;
}

void surrogate_free(komd_ctx* c) {
    Surrogate* s = c->sur;
    if (!s) return;
    void* ptrs[] = {s->params, s->ops, s->X, s->T, s->S, s->E, s->mom, s->vel, s->gpart, s->grad, s->W, s->fj,
                    s->delta, s->lpart, s->lacc, s->save, s->stats, s->lam, s->dlam};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (auto p : s->snap_r) hipFree(p);
    for (auto p : s->snap_f) hipFree(p);
    for (auto p : s->snap_e) hipFree(p);
    delete s;
    c->sur = nullptr;
}

// ---------------------------------------------------------------------------
// C ABI helpers shared with komd_api.hip
komd_status api_fail(komd_ctx* c, komd_status s, const std::string& msg);
komd_status api_sync(komd_ctx* c, const char* where);
hipError_t api_prepare_lists(komd_ctx* c);  // container update + neighbour lists

namespace {

#define SCALL(c, expr, where)                                                                            \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e == hipErrorNotSupported) return api_fail(c, KOMD_ERR_STATE, "surrogate parameters not set"); \
        if (_e != hipSuccess) return api_fail(c, KOMD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(_e)); \
    } while (0)

komd_status need(komd_ctx* c) {
    if (!c) return KOMD_ERR_ARG;
    if (!c->have_sites) return api_fail(c, KOMD_ERR_STATE, "no sites: call komd_set_sites first");
    return KOMD_OK;
}

// save / restore positions and forces around passes that rebuild the container
hipError_t save_state(komd_ctx* c) {
    Surrogate* s = sur_of(c);
    KOMD_CHECK_HIP(ensure_site_buffers(c));  // buffers of the current site count
    KOMD_CHECK_HIP(salloc(&s->save, 6 * (size_t)c->n));
    const size_t nb = (size_t)c->n * 8;
    double* src[6] = {c->d.rx, c->d.ry, c->d.rz, c->d.fx, c->d.fy, c->d.fz};
    for (int k = 0; k < 6; k++)
        KOMD_CHECK_HIP(hipMemcpyAsync(s->save + k * c->n, src[k], nb, hipMemcpyDeviceToDevice, c->stream));
    return hipSuccess;
}
hipError_t restore_state(komd_ctx* c) {
    Surrogate* s = c->sur;
    const size_t nb = (size_t)c->n * 8;
    double* dst[6] = {c->d.rx, c->d.ry, c->d.rz, c->d.fx, c->d.fy, c->d.fz};
    for (int k = 0; k < 6; k++)
        KOMD_CHECK_HIP(hipMemcpyAsync(dst[k], s->save + k * c->n, nb, hipMemcpyDeviceToDevice, c->stream));
    c->nl_valid = false;
    return hipSuccess;
}

template <typename T>
komd_status to_site_order(komd_ctx* c, const T* dev_sorted, int width, T* out, const char* where) {
    const size_t n = (size_t)c->n;
    std::vector<T> h(n * width);
    std::vector<uint32_t> srt(n);
    if (hipMemcpy(h.data(), dev_sorted, h.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(srt.data(), c->d.sorted, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return api_fail(c, KOMD_ERR_HIP, where);
    for (size_t p = 0; p < n; p++)
        for (int k = 0; k < width; k++) out[(size_t)srt[p] * width + k] = h[p * width + k];
    return KOMD_OK;
}

// one training step on stored snapshot k: the container is rebuilt from its positions
hipError_t train_step(komd_ctx* c, int64_t k, float wE, float wF, float lr, bool update, float* grad_out) {
    Surrogate* s = c->sur;
    const size_t n = (size_t)c->n;
    const size_t nb = n * 8;
    KOMD_CHECK_HIP(hipMemcpyAsync(c->d.rx, s->snap_r[(size_t)k], nb, hipMemcpyDeviceToDevice, c->stream));
    KOMD_CHECK_HIP(hipMemcpyAsync(c->d.ry, s->snap_r[(size_t)k] + n, nb, hipMemcpyDeviceToDevice, c->stream));
    KOMD_CHECK_HIP(hipMemcpyAsync(c->d.rz, s->snap_r[(size_t)k] + 2 * n, nb, hipMemcpyDeviceToDevice, c->stream));
    KOMD_CHECK_HIP(api_prepare_lists(c));
    KOMD_CHECK_HIP(launch_desc(c, false));
    KOMD_CHECK_HIP(launch_mlp(c, false, nullptr));
    const int fb = site_blocks(c);
    {
        ScopedTimer tm(c, "sur_force");
        ForceArgs a = force_args(c);
        a.fref = s->snap_f[(size_t)k];
        a.delta = s->delta;
        a.part = s->lpart + kTrainBlocks;
        a.cF = 2.f * wF / (3.f * (float)n);
        a.lF = wF / (3.f * (float)n);
        k_sur_force<false><<<8 * a.chunk, kSurForceThreads, 0, c->stream>>>(a);
        KOMD_CHECK_HIP(hipGetLastError());
        k_sur_apply<FORCE_TRAIN><<<fb, 256, 0, c->stream>>>(a);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    {
        ScopedTimer tm(c, "sur_jvp");
        JvpArgs a{c->d.nl, c->d.nlc, c->d.inv, s->T, s->delta, s->params, s->W, (int)n, pirc_of(c)};
        k_sur_jvp<<<group_blocks(c), 256, 0, c->stream>>>(a);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    {
        ScopedTimer tm(c, "sur_train");
        TrainArgs a{s->ops, s->params, s->X, s->W, s->snap_e[(size_t)k], c->d.sorted, s->gpart, s->lpart, (int)n,
                    2.f * wE / (float)n, wE / (float)n};
        k_sur_train<<<kTrainBlocks, 256, kTrainLds, c->stream>>>(a);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    {
        ScopedTimer tm(c, "sur_adam");
        if (update) s->step++;
        const float bc1 = 1.f / (1.f - powf(0.9f, (float)std::max<int64_t>(1, s->step)));
        const float bc2 = 1.f / (1.f - powf(0.999f, (float)std::max<int64_t>(1, s->step)));
        k_sur_adam<<<(NPARAM + 255) / 256, 256, 0, c->stream>>>(s->params, s->mom, s->vel, s->gpart, kTrainBlocks,
                                                               grad_out, lr, bc1, bc2, update ? 1 : 0, s->lpart,
                                                               kTrainBlocks, s->lpart + kTrainBlocks, fb, s->lacc);
        KOMD_CHECK_HIP(hipGetLastError());
        if (update) {
            k_sur_pack<<<1, 256, 0, c->stream>>>(s->params, s->ops);
            KOMD_CHECK_HIP(hipGetLastError());
        }
    }
    return hipSuccess;
}

hipError_t ensure_train_buffers(komd_ctx* c) {
    Surrogate* s = sur_of(c);
    KOMD_CHECK_HIP(ensure_site_buffers(c));
    const size_t n = (size_t)c->n;
    KOMD_CHECK_HIP(salloc(&s->W, n * ND));
    KOMD_CHECK_HIP(salloc(&s->delta, n));
    KOMD_CHECK_HIP(salloc(&s->gpart, (size_t)kTrainBlocks * NPARAM));
    KOMD_CHECK_HIP(salloc(&s->grad, NPARAM));
    KOMD_CHECK_HIP(salloc(&s->lpart, (size_t)kTrainBlocks + (size_t)site_blocks(c)));
    KOMD_CHECK_HIP(salloc(&s->lacc, 2));
    KOMD_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sur_train),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kTrainLds));
    if (!s->mom) {
        KOMD_CHECK_HIP(salloc(&s->mom, NPARAM));
        KOMD_CHECK_HIP(salloc(&s->vel, NPARAM));
        KOMD_CHECK_HIP(dzero(c, s->mom, NPARAM * 4));
        KOMD_CHECK_HIP(dzero(c, s->vel, NPARAM * 4));
        s->step = 0;
    }
    return hipSuccess;
}

}  // namespace
}  // namespace komd

using namespace komd;

extern "C" {

komd_status komd_surrogate_set_params(komd_ctx* c, const float* params, int64_t n) {
    komd_status st = need(c);
    if (st) return st;
    if (!params || n != NPARAM) return api_fail(c, KOMD_ERR_ARG, "expected KOMD_SUR_NPARAM parameters");
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite((double)params[i])) return api_fail(c, KOMD_ERR_ARG, "non-finite surrogate parameter");
    Surrogate* s = sur_of(c);
    SCALL(c, ensure_site_buffers(c), "komd_surrogate_set_params");
    SCALL(c, h2d(c, s->params, params, NPARAM * 4), "komd_surrogate_set_params");
    k_sur_pack<<<1, 256, 0, c->stream>>>(s->params, s->ops);
    SCALL(c, hipGetLastError(), "komd_surrogate_set_params");
    s->have_params = true;
    return api_sync(c, "komd_surrogate_set_params");
}

komd_status komd_surrogate_get_params(komd_ctx* c, float* params, int64_t n) {
    if (!c || !params || n != NPARAM) return KOMD_ERR_ARG;
    if (!c->sur || !c->sur->params) return api_fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    SCALL(c, hipStreamSynchronize(c->stream), "komd_surrogate_get_params");
    SCALL(c, hipMemcpy(params, c->sur->params, NPARAM * 4, hipMemcpyDeviceToHost), "komd_surrogate_get_params");
    return KOMD_OK;
}

komd_status komd_surrogate_descriptors(komd_ctx* c, float* out) {
    komd_status st = need(c);
    if (st) return st;
    if (!out) return KOMD_ERR_ARG;
    SCALL(c, ensure_site_buffers(c), "komd_surrogate_descriptors");
    SCALL(c, api_prepare_lists(c), "komd_surrogate_descriptors");
    SCALL(c, launch_desc(c, true), "komd_surrogate_descriptors");
    st = api_sync(c, "komd_surrogate_descriptors");
    if (st) return st;
    return to_site_order(c, c->sur->X, ND, out, "komd_surrogate_descriptors");
}

komd_status komd_surrogate_energies(komd_ctx* c, float* out) {
    komd_status st = need(c);
    if (st) return st;
    if (!out) return KOMD_ERR_ARG;
    if (!c->sur || !c->sur->have_params) return api_fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    SCALL(c, api_prepare_lists(c), "komd_surrogate_energies");
    SCALL(c, launch_desc(c, false), "komd_surrogate_energies");
    SCALL(c, launch_mlp(c, false, nullptr), "komd_surrogate_energies");
    st = api_sync(c, "komd_surrogate_energies");
    if (st) return st;
    return to_site_order(c, c->sur->E, 1, out, "komd_surrogate_energies");
}

komd_status komd_atm_site_energies(komd_ctx* c, double* out) {
    komd_status st = need(c);
    if (st) return st;
    if (!out) return KOMD_ERR_ARG;
    SCALL(c, save_state(c), "komd_atm_site_energies");
    SCALL(c, api_prepare_lists(c), "komd_atm_site_energies");
    SCALL(c, launch_atm(c, 1, false, nullptr, nullptr, c->d.esite), "komd_atm_site_energies");
    SCALL(c, restore_state(c), "komd_atm_site_energies");
    st = api_sync(c, "komd_atm_site_energies");
    if (st) return st;
    SCALL(c, hipMemcpy(out, c->d.esite, (size_t)c->n * 8, hipMemcpyDeviceToHost), "komd_atm_site_energies");
    return KOMD_OK;
}

komd_status komd_train_add_snapshot(komd_ctx* c) {
    komd_status st = need(c);
    if (st) return st;
    Surrogate* s = sur_of(c);
    // buffers of the current site count first: a change of count drops the stored set
    SCALL(c, ensure_site_buffers(c), "komd_train_add_snapshot");
    const size_t n = (size_t)c->n;
    double* r = nullptr;
    float *f = nullptr, *e = nullptr;
    if (hipMalloc((void**)&r, 3 * n * 8) != hipSuccess || hipMalloc((void**)&f, 3 * n * 4) != hipSuccess ||
        hipMalloc((void**)&e, n * 4) != hipSuccess) {
        if (r) hipFree(r);
        if (f) hipFree(f);
        return api_fail(c, KOMD_ERR_NOMEM, "komd_train_add_snapshot: device allocation");
    }
    s->snap_r.push_back(r);
    s->snap_f.push_back(f);
    s->snap_e.push_back(e);
    SCALL(c, save_state(c), "komd_train_add_snapshot");
    SCALL(c, api_prepare_lists(c), "komd_train_add_snapshot");  // wraps positions, zeroes forces
    SCALL(c, launch_atm(c, 1, false, nullptr, nullptr, c->d.esite), "komd_train_add_snapshot");
    const size_t nb = n * 8;
    SCALL(c, hipMemcpyAsync(r, c->d.rx, nb, hipMemcpyDeviceToDevice, c->stream), "komd_train_add_snapshot");
    SCALL(c, hipMemcpyAsync(r + n, c->d.ry, nb, hipMemcpyDeviceToDevice, c->stream), "komd_train_add_snapshot");
    SCALL(c, hipMemcpyAsync(r + 2 * n, c->d.rz, nb, hipMemcpyDeviceToDevice, c->stream), "komd_train_add_snapshot");
    k_f64_to_aos<<<site_blocks(c), 256, 0, c->stream>>>((int)n, c->d.fx, c->d.fy, c->d.fz, f, c->d.esite, e);
    SCALL(c, hipGetLastError(), "komd_train_add_snapshot");
    SCALL(c, restore_state(c), "komd_train_add_snapshot");
    return api_sync(c, "komd_train_add_snapshot");
}

komd_status komd_train_clear(komd_ctx* c) {
    if (!c) return KOMD_ERR_ARG;
    if (c->sur) {
        hipStreamSynchronize(c->stream);
        for (auto p : c->sur->snap_r) hipFree(p);
        for (auto p : c->sur->snap_f) hipFree(p);
        for (auto p : c->sur->snap_e) hipFree(p);
        c->sur->snap_r.clear();
        c->sur->snap_f.clear();
        c->sur->snap_e.clear();
    }
    return KOMD_OK;
}

komd_status komd_train_size(komd_ctx* c, int64_t* out) {
    if (!c || !out) return KOMD_ERR_ARG;
    *out = c->sur ? (int64_t)c->sur->snap_r.size() : 0;
    return KOMD_OK;
}

komd_status komd_train_stats(komd_ctx* c, double out[3]) {
    if (!c || !out) return KOMD_ERR_ARG;
    if (!c->sur || c->sur->snap_e.empty()) return api_fail(c, KOMD_ERR_STATE, "no training snapshots");
    const size_t n = (size_t)c->n;
    std::vector<float> e(n), f(3 * n);
    double se = 0, se2 = 0, sf2 = 0;
    for (size_t k = 0; k < c->sur->snap_e.size(); k++) {
        SCALL(c, hipMemcpy(e.data(), c->sur->snap_e[k], n * 4, hipMemcpyDeviceToHost), "komd_train_stats");
        SCALL(c, hipMemcpy(f.data(), c->sur->snap_f[k], 3 * n * 4, hipMemcpyDeviceToHost), "komd_train_stats");
        for (float x : e) {
            se += x;
            se2 += (double)x * x;
        }
        for (float x : f) sf2 += (double)x * x;
    }
    const double cnt = (double)n * (double)c->sur->snap_e.size();
    out[0] = se / cnt;
    out[1] = sqrt(std::max(0.0, se2 / cnt - out[0] * out[0]));
    out[2] = sqrt(sf2 / (3.0 * cnt));
    return KOMD_OK;
}

komd_status komd_train_standardize(komd_ctx* c) {
    komd_status st = need(c);
    if (st) return st;
    Surrogate* s = sur_of(c);
    if (s->snap_r.empty()) return api_fail(c, KOMD_ERR_STATE, "no training snapshots");
    SCALL(c, ensure_site_buffers(c), "komd_train_standardize");
    if (!s->have_params) {
        SCALL(c, dzero(c, s->params, NPARAM * 4), "komd_train_standardize");
        s->have_params = true;
    }
    SCALL(c, salloc(&s->stats, (size_t)kJudgeBlocks * 2 * ND), "komd_train_standardize");
    SCALL(c, save_state(c), "komd_train_standardize");
    const size_t n = (size_t)c->n, nb = n * 8;
    std::vector<double> tot(2 * ND, 0.0), part((size_t)kJudgeBlocks * 2 * ND);
    for (size_t k = 0; k < s->snap_r.size(); k++) {
        SCALL(c, hipMemcpyAsync(c->d.rx, s->snap_r[k], nb, hipMemcpyDeviceToDevice, c->stream), "standardize");
        SCALL(c, hipMemcpyAsync(c->d.ry, s->snap_r[k] + n, nb, hipMemcpyDeviceToDevice, c->stream), "standardize");
        SCALL(c, hipMemcpyAsync(c->d.rz, s->snap_r[k] + 2 * n, nb, hipMemcpyDeviceToDevice, c->stream), "standardize");
        SCALL(c, api_prepare_lists(c), "komd_train_standardize");
        SCALL(c, launch_desc(c, true), "komd_train_standardize");
        k_sur_colstats<<<kJudgeBlocks, 128, 0, c->stream>>>((int)n, s->X, s->stats);
        SCALL(c, hipGetLastError(), "komd_train_standardize");
        // host buffers are pageable: synchronous copies after the stream has drained
        SCALL(c, hipStreamSynchronize(c->stream), "komd_train_standardize");
        SCALL(c, hipMemcpy(part.data(), s->stats, part.size() * 8, hipMemcpyDeviceToHost), "komd_train_standardize");
        for (int b = 0; b < kJudgeBlocks; b++)
            for (int j = 0; j < 2 * ND; j++) tot[(size_t)j] += part[(size_t)b * 2 * ND + j];
    }
    SCALL(c, restore_state(c), "komd_train_standardize");
    std::vector<float> prm(NPARAM);
    SCALL(c, hipMemcpy(prm.data(), s->params, NPARAM * 4, hipMemcpyDeviceToHost), "komd_train_standardize");
    const double cnt = (double)n * (double)s->snap_r.size();
    for (int j = 0; j < ND; j++) {
        const double mean = tot[(size_t)j] / cnt;
        const double var = std::max(0.0, tot[(size_t)ND + j] / cnt - mean * mean);
        prm[O_SHIFT + j] = (float)mean;
        prm[O_SCALE + j] = (float)(1.0 / (sqrt(var) + 1e-6 * (fabs(mean) + 1e-12)));
    }
    SCALL(c, h2d(c, s->params, prm.data(), NPARAM * 4), "komd_train_standardize");
    k_sur_pack<<<1, 256, 0, c->stream>>>(s->params, s->ops);
    SCALL(c, hipGetLastError(), "komd_train_standardize");
    return api_sync(c, "komd_train_standardize");
}

komd_status komd_train_gradient(komd_ctx* c, int64_t k, float wE, float wF, float* grad, double* loss) {
    komd_status st = need(c);
    if (st) return st;
    Surrogate* s = c->sur;
    if (!s || !s->have_params) return api_fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    if (k < 0 || k >= (int64_t)s->snap_r.size()) return api_fail(c, KOMD_ERR_ARG, "snapshot index out of range");
    SCALL(c, ensure_train_buffers(c), "komd_train_gradient");
    SCALL(c, save_state(c), "komd_train_gradient");
    SCALL(c, hipMemsetAsync(s->lacc, 0, 16, c->stream), "komd_train_gradient");
    SCALL(c, train_step(c, k, wE, wF, 0.f, false, s->grad), "komd_train_gradient");
    SCALL(c, restore_state(c), "komd_train_gradient");
    st = api_sync(c, "komd_train_gradient");
    if (st) return st;
    if (grad) SCALL(c, hipMemcpy(grad, s->grad, NPARAM * 4, hipMemcpyDeviceToHost), "komd_train_gradient");
    if (loss) SCALL(c, hipMemcpy(loss, s->lacc, 8, hipMemcpyDeviceToHost), "komd_train_gradient");
    return KOMD_OK;
}

komd_status komd_train_epoch(komd_ctx* c, float lr, float wE, float wF, uint64_t seed, double* loss) {
    komd_status st = need(c);
    if (st) return st;
    Surrogate* s = c->sur;
    if (!s || !s->have_params) return api_fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    if (s->snap_r.empty()) return api_fail(c, KOMD_ERR_STATE, "no training snapshots");
    SCALL(c, ensure_train_buffers(c), "komd_train_epoch");
    SCALL(c, save_state(c), "komd_train_epoch");
    SCALL(c, hipMemsetAsync(s->lacc, 0, 16, c->stream), "komd_train_epoch");
    // order of the snapshots: Fisher-Yates with a SplitMix64 stream
    std::vector<int64_t> order(s->snap_r.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int64_t)i;
    uint64_t z = seed;
    for (size_t i = order.size(); i > 1; i--) {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        std::swap(order[i - 1], order[(size_t)(x % i)]);
    }
    for (size_t i = 0; i < order.size(); i++) {
        SCALL(c, train_step(c, order[i], wE, wF, lr, true, nullptr), "komd_train_epoch");
        if ((i & 63) == 63) {
            st = api_sync(c, "komd_train_epoch");
            if (st) return st;
        }
    }
    SCALL(c, restore_state(c), "komd_train_epoch");
    st = api_sync(c, "komd_train_epoch");
    if (st) return st;
    double acc[2];
    SCALL(c, hipMemcpy(acc, s->lacc, 16, hipMemcpyDeviceToHost), "komd_train_epoch");
    if (loss) *loss = acc[0] / (double)order.size();
    return KOMD_OK;
}

komd_status komd_surrogate_judge(komd_ctx* c, double out[2]) {
    komd_status st = need(c);
    if (st) return st;
    if (!out) return KOMD_ERR_ARG;
    Surrogate* s = c->sur;
    if (!s || !s->have_params) return api_fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    SCALL(c, ensure_site_buffers(c), "komd_surrogate_judge");
    SCALL(c, salloc(&s->fj, 3 * (size_t)c->n), "komd_surrogate_judge");
    SCALL(c, salloc(&s->stats, (size_t)kJudgeBlocks * 2 * ND), "komd_surrogate_judge");
    SCALL(c, save_state(c), "komd_surrogate_judge");
    SCALL(c, api_prepare_lists(c), "komd_surrogate_judge");  // forces zeroed
    SCALL(c, surrogate_pass(c, false, false, FORCE_OUT), "komd_surrogate_judge");
    SCALL(c, launch_atm(c, 1, false, nullptr, nullptr, nullptr), "komd_surrogate_judge");
    k_sur_judge<<<kJudgeBlocks, 256, 0, c->stream>>>((int)c->n, s->fj, c->d.fx, c->d.fy, c->d.fz, s->stats);
    SCALL(c, hipGetLastError(), "komd_surrogate_judge");
    SCALL(c, restore_state(c), "komd_surrogate_judge");
    st = api_sync(c, "komd_surrogate_judge");
    if (st) return st;
    std::vector<double> part(3 * kJudgeBlocks);  // pageable: a synchronous copy after the drain
    SCALL(c, hipMemcpy(part.data(), s->stats, part.size() * 8, hipMemcpyDeviceToHost), "komd_surrogate_judge");
    double se = 0, sr = 0, mx = 0;
    for (int b = 0; b < kJudgeBlocks; b++) {
        se += part[3 * b];
        sr += part[3 * b + 1];
        mx = std::max(mx, part[3 * b + 2]);
    }
    const double rms_ref = sqrt(sr / (double)c->n);
    out[0] = rms_ref > 0 ? sqrt(se / (double)c->n) / rms_ref : 0.0;
    out[1] = rms_ref > 0 ? sqrt(mx) / rms_ref : 0.0;
    return KOMD_OK;
}

komd_status komd_set_hybrid(komd_ctx* c, int32_t axis, double center, double half_width, double blend) {
    if (!c) return KOMD_ERR_ARG;
    if (axis < 0 || axis > 2 || !(half_width >= 0) || !(blend > 0))
        return api_fail(c, KOMD_ERR_ARG, "hybrid region: axis 0..2, half_width >= 0, blend > 0");
    c->hyb_axis = axis;
    c->hyb_center = center;
    c->hyb_half = half_width;
    c->hyb_blend = blend;
    c->graph_key = -1;
    return KOMD_OK;
}

}  // extern "C"
