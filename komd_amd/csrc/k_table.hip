// Tabulated pair force (the IBI force functor) over the full neighbour lists -- gfx950.
//
// IBI_Force (plugins/IBI/IBI_ForceFunctor.cpp:24-42): for every pair within the cutoff,
// F = evaluateAt(|dr|) of a piecewise-linear table (FunctionPL, IBI_Math.h:30-72) and the
// force (dr / |dr|) F on the first member, its negative on the second, dr = r_i - r_j.
// No molecule exclusion.
//
// One wave per site (cell-sorted order), four independent waves per workgroup, site slabs
// per XCD.  Lane a takes neighbour a of the site's list (x = r_j - r_i in FP32; k_cell wrote
// the list with the reference's FP64 cutoff decision), so the force on i is -x F(r)/r and
// each pair is visited from both members, each adding its own force: no scatter, no atomic.
// The table sits in global memory (a few KB, cache-resident); a uniform table finds its
// segment from (r - x0)/dx with a one-node correction, any other by binary search.  The
// defaults below the first and above the last node make the table discontinuous there, so
// a distance within a few ulps of either end is decided in FP64 from the absolute
// positions, in the reference's operation order.
#include <algorithm>
#include <vector>

#include "komd_cell.hpp"

namespace komd {

struct PairTable {
    int nodes = 0;
    bool uniform = false;
    bool have_u = false;
    double x0 = 0, xl = 0;
    float y_low = 0, y_high = 0, u_low = 0, u_high = 0;
    float* tx = nullptr;    // [nodes] x nodes (FP32)
    float4* seg = nullptr;  // [nodes] (x[k-1], y[k-1], slope of y on [x[k-1], x[k]], 0); [0] = (x0, y0, 0, 0)
    float2* useg = nullptr; // [nodes] (u[k-1], slope of u) for the potential, same layout
};

namespace {

constexpr int kTabWaves = 4;

struct TableArgs {
    const float4* __restrict__ nl;
    const uint32_t* __restrict__ nlc;
    const uint32_t* __restrict__ sorted;
    const double* __restrict__ rx;
    const double* __restrict__ ry;
    const double* __restrict__ rz;
    double* __restrict__ fx;
    double* __restrict__ fy;
    double* __restrict__ fz;
    double* __restrict__ vx;  // second half-kick fused into the force write (kick > 0)
    double* __restrict__ vy;
    double* __restrict__ vz;
    const double* __restrict__ mass;
    const float* __restrict__ tx;
    const float4* __restrict__ seg;
    const float2* __restrict__ useg;
    double* __restrict__ part;   // [nblocks] energy partials (energy mode)
    double* __restrict__ vpart;  // [kSym][nblocks] virial partials (energy mode)
    Grid g;
    int n, chunk, nblocks, nodes, uniform;
    float x0f, inv_dx, band;
    float y_low, y_high, u_low, u_high;
    double x0, xl, kick;
};

// Index k of the first node with r <= x[k] (0 <= k < nodes), for x0 <= r <= x_last up to
// rounding: a uniform table from (r - x0)/dx, moved to the exact node; otherwise binary search.
__device__ __forceinline__ int find_segment(const TableArgs& a, float r) {
    if (a.uniform) {
        int k = (int)((r - a.x0f) * a.inv_dx) + 1;
        k = max(1, min(k, a.nodes - 1));
        while (k > 1 && r <= a.tx[k - 1]) k--;
        while (k < a.nodes - 1 && r > a.tx[k]) k++;
        return r <= a.tx[0] ? 0 : k;
    }
    int lo = 0, hi = a.nodes - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (r <= a.tx[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

template <bool ENERGY>
__global__ __launch_bounds__(64 * kTabWaves) void k_pair_table(TableArgs a) {
    __shared__ double red[kTabWaves];
    const int lane = threadIdx.x & 63;
    const int wvu = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    const int lb = (b & 7) * a.chunk + (b >> 3);  // XCD-aware: contiguous site slabs per XCD
    const int p = lb * kTabWaves + wvu;
    double e_site = 0.0;
    double vir[kSym] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (p < a.n) {
        const int nn = __builtin_amdgcn_readfirstlane((int)a.nlc[p]);
        const uint32_t oi = a.sorted[p];
        const float4* row = a.nl + (size_t)p * kFullCap;
        double ax = 0.0, ay = 0.0, az = 0.0, eu = 0.0;
        for (int s = lane; s < nn; s += 64) {
            const float4 v = row[s];
            const float r2 = v.x * v.x + v.y * v.y + v.z * v.z;
            const float r = __builtin_sqrtf(r2);
            bool below = r < a.x0f, above = r > (float)a.xl;
            // within a few ulps of either end: the reference's FP64 distance decides
            const bool amb = fabsf(r - a.x0f) <= a.band || fabsf(r - (float)a.xl) <= a.band;
            if (ballot(amb) != 0ull && amb) {
                const double r64 = sqrt(exact_r2(a.rx, a.ry, a.rz, a.g, oi, __float_as_uint(v.w)));
                below = r64 < a.x0;
                above = r64 > a.xl;
            }
            float F, U = 0.f;
            if (above) {
                F = a.y_high;
                U = a.u_high;
            } else if (below) {
                F = a.y_low;
                U = a.u_low;
            } else {
                const int k = find_segment(a, r);
                const float4 sg = a.seg[k];  // ya + (yb - ya)/(xb - xa) (r - xa), IBI_Math.h:49-56
                const float t = r - sg.x;
                F = sg.y + sg.z * t;
                if (ENERGY && a.useg) {
                    const float2 us = a.useg[k];
                    U = us.x + us.y * t;
                }
            }
            const float fac = F / r;  // force on i = -x F / r
            ax -= (double)(v.x * fac);
            ay -= (double)(v.y * fac);
            az -= (double)(v.z * fac);
            if (ENERGY) {
                // each pair is visited from both members: half of U and of the virial
                eu += 0.5 * (double)U;
                const float hf = 0.5f * fac;
                sym_outer<double>(vir, v.x, v.y, v.z, v.x * hf, v.y * hf, v.z * hf);
            }
        }
        const double dx = wave_sum(ax), dy = wave_sum(ay), dz = wave_sum(az);
        if (ENERGY) e_site = wave_sum(eu);
        if (lane == 0) {
            const double nfx = a.fx[oi] + dx, nfy = a.fy[oi] + dy, nfz = a.fz[oi] + dz;
            a.fx[oi] = nfx;
            a.fy[oi] = nfy;
            a.fz[oi] = nfz;
            if (a.kick > 0.0) kick_site(a.vx, a.vy, a.vz, a.mass, oi, a.kick, nfx, nfy, nfz);
        }
    }
    if (!ENERGY) return;
    if (lane == 0) red[wvu] = e_site;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < kTabWaves; k++) s += red[k];
        a.part[b] = s;
    }
    block_store_sym(a.vpart, a.nblocks, b, vir);
}

}  // namespace

int table_blocks(komd_ctx* c) { return 8 * (((int)((c->n + kTabWaves - 1) / kTabWaves) + 7) / 8); }

hipError_t launch_table(komd_ctx* c, bool energy) {
    const PairTable* t = c->tab;
    if (!t || !c->d.nl || !c->d.nlc) return hipErrorInvalidValue;
    auto& d = c->d;
    TableArgs a{};
    a.nl = d.nl; a.nlc = d.nlc; a.sorted = d.sorted;
    a.rx = d.rx; a.ry = d.ry; a.rz = d.rz;
    a.fx = d.fx; a.fy = d.fy; a.fz = d.fz;
    a.vx = d.vx; a.vy = d.vy; a.vz = d.vz; a.mass = d.mass;
    a.tx = t->tx; a.seg = t->seg; a.useg = t->have_u ? t->useg : nullptr;
    a.part = d.red; a.vpart = d.vpart;
    a.g = c->grid;
    a.n = (int)c->n;
    a.nblocks = table_blocks(c);
    a.chunk = a.nblocks / 8;
    a.nodes = t->nodes;
    a.uniform = t->uniform ? 1 : 0;
    a.x0 = t->x0;
    a.xl = t->xl;
    a.x0f = (float)t->x0;
    a.inv_dx = (float)((double)(t->nodes - 1) / (t->xl - t->x0));
    // FP32 list vectors carry a few 1e-7 of absolute error per component (cell-relative
    // coordinates up to two cells long): the band is wider by an order of magnitude
    a.band = (float)(1e-5 * std::max({c->grid.cs, fabs(t->x0), fabs(t->xl)}));
    a.y_low = t->y_low; a.y_high = t->y_high; a.u_low = t->u_low; a.u_high = t->u_high;
    a.kick = (!energy && c->kick_in == KICK_IN_TABLE) ? 0.5 * c->prm.delta_t : 0.0;
    ScopedTimer tm(c, "table");
    if (energy) k_pair_table<true><<<a.nblocks, 64 * kTabWaves, 0, c->stream>>>(a);
    else k_pair_table<false><<<a.nblocks, 64 * kTabWaves, 0, c->stream>>>(a);
    return hipGetLastError();
}

void table_free(komd_ctx* c) {
    PairTable* t = c->tab;
    if (!t) return;
    if (t->tx) hipFree(t->tx);
    if (t->seg) hipFree(t->seg);
    if (t->useg) hipFree(t->useg);
    delete t;
    c->tab = nullptr;
}

// Host copies of the table in the kernel's layout; the slopes are formed in FP64 as
// linearInterpolation does (IBI_Math.h:49-56) and rounded once.
hipError_t table_upload(komd_ctx* c, int64_t nodes, const double* x, const double* y, double y_low, double y_high,
                        const double* u, double u_low, double u_high) {
    PairTable* t = new PairTable();
    t->nodes = (int)nodes;
    t->x0 = x[0];
    t->xl = x[nodes - 1];
    t->y_low = (float)y_low;
    t->y_high = (float)y_high;
    t->u_low = (float)u_low;
    t->u_high = (float)u_high;
    t->have_u = u != nullptr;
    const double dx = (x[nodes - 1] - x[0]) / (double)(nodes - 1);
    t->uniform = true;
    for (int64_t k = 0; k < nodes && t->uniform; k++)
        t->uniform = fabs(x[k] - (x[0] + (double)k * dx)) <= 1e-9 * fabs(dx);
    std::vector<float> tx((size_t)nodes);
    std::vector<float4> seg((size_t)nodes);
    std::vector<float2> useg((size_t)nodes);
    for (int64_t k = 0; k < nodes; k++) {
        tx[(size_t)k] = (float)x[k];
        if (k == 0) {
            seg[0] = make_float4((float)x[0], (float)y[0], 0.f, 0.f);
            useg[0] = make_float2(u ? (float)u[0] : 0.f, 0.f);
            continue;
        }
        const double h = x[k] - x[k - 1];
        seg[(size_t)k] = make_float4((float)x[k - 1], (float)y[k - 1], (float)((y[k] - y[k - 1]) / h), 0.f);
        if (u) useg[(size_t)k] = make_float2((float)u[k - 1], (float)((u[k] - u[k - 1]) / h));
    }
    hipError_t e = hipMalloc((void**)&t->tx, tx.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&t->seg, seg.size() * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc((void**)&t->useg, useg.size() * sizeof(float2));
    if (e == hipSuccess) e = h2d(c, t->tx, tx.data(), tx.size() * sizeof(float));
    if (e == hipSuccess) e = h2d(c, t->seg, seg.data(), seg.size() * sizeof(float4));
    if (e == hipSuccess) e = h2d(c, t->useg, useg.data(), useg.size() * sizeof(float2));
    if (e != hipSuccess) {
        PairTable* old = c->tab;
        c->tab = t;
        table_free(c);
        c->tab = old;
        return e;
    }
    table_free(c);
    c->tab = t;
    return hipSuccess;
}

}  // namespace komd
