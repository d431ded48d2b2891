// Container update (COM, periodic wrap, force reset, counting-sort binning),
// split velocity Verlet, temperature and velocity rescaling -- gfx950.
//
// All FP64 arithmetic here reproduces the reference's operation order with
// no FMA contraction, so binning and integration are bit-identical to the
// FP64 restatement (LinkedCells.cpp:87-103, Integrator.cpp:28-37).
#pragma clang fp contract(off)

#include "komd_internal.hpp"
#include "komd_reduce.hpp"

namespace komd {
namespace {

constexpr int kBlock = 256;
// k_com_wrap_key: one molecule per thread and no shared memory; 64-thread workgroups spread a
// small system's few waves over all CUs (32k sites: 500 workgroups instead of 125)
constexpr int kKeyBlock = 64;

__device__ __forceinline__ int cell_coord(double com, double low, double cs, int nd) {
    // LinkedCells.cpp:93-94: fBin = (com - low)/cell_size; clamp((int64)fBin, 0, n-1)
    const double fbin = (com - low) / cs;
    if (!(fbin > 0.0)) return 0;                 // (int64) truncation of (-1,0] and clamp of <= -1 (and NaN)
    if (fbin >= (double)nd) return nd - 1;
    return (int)fbin;
}

// Integrator::Step0 of the sites s0..s1-1 (Integrator.cpp:28-32), optionally after the
// velocity rescale v *= beta of the previous step's thermostat.
__device__ __forceinline__ void drift_unit(int s0, int s1, double* __restrict__ rx, double* __restrict__ ry,
                                           double* __restrict__ rz, double* __restrict__ vx,
                                           double* __restrict__ vy, double* __restrict__ vz,
                                           const double* __restrict__ fx, const double* __restrict__ fy,
                                           const double* __restrict__ fz, const double* __restrict__ mass,
                                           double dt, double dt_half, const double* __restrict__ beta_src) {
    const double beta = beta_src ? beta_src[0] : 1.0;
    for (int s = s0; s < s1; s++) {
        double x = vx[s], y = vy[s], z = vz[s];
        if (beta_src) { x = x * beta; y = y * beta; z = z * beta; }
        const double k = dt_half / mass[s];
        x = x + fx[s] * k;
        y = y + fy[s] * k;
        z = z + fz[s] * k;
        vx[s] = x; vy[s] = y; vz[s] = z;
        rx[s] = rx[s] + x * dt;
        ry[s] = ry[s] + y * dt;
        rz[s] = rz[s] + z * dt;
    }
}

// Container update of one molecule (sites s0..s1-1):
//  com = sum(r m)/sum(m); wrap r by the COM; com again   (MoleculeContainer.cpp:55-75, LinkedCells.cpp:390-400)
//  f = 0 when reset_f                                    (FReset_Kernel LinkedCells.cpp:386-388)
//  key = flat domain cell of the COM, slot = atomic arrival index
__device__ __forceinline__ void bin_unit(int s0, int s1, double* __restrict__ rx, double* __restrict__ ry,
                                         double* __restrict__ rz, double* __restrict__ fx, double* __restrict__ fy,
                                         double* __restrict__ fz, const double* __restrict__ mass,
                                         uint32_t* __restrict__ key, uint32_t* __restrict__ slot,
                                         uint32_t* __restrict__ count, const Grid& g, int bin_site, int reset_f,
                                         uint32_t* __restrict__ flags) {
    double cx, cy, cz;
    if (bin_site) {
        // every site on its own: wrap and bin by its own position (s1 == s0 + 1)
        cx = rx[s0]; cy = ry[s0]; cz = rz[s0];
    } else {
        double ax = 0, ay = 0, az = 0, tm = 0;
        for (int s = s0; s < s1; s++) {
            const double ms = mass[s];
            ax = ax + rx[s] * ms; ay = ay + ry[s] * ms; az = az + rz[s] * ms;
            tm = tm + ms;
        }
        cx = ax / tm; cy = ay / tm; cz = az / tm;
    }
    double ox = 0, oy = 0, oz = 0;
    if (cx < g.low[0]) ox = g.size[0];
    if (cx > g.high[0]) ox = -g.size[0];
    if (cy < g.low[1]) oy = g.size[1];
    if (cy > g.high[1]) oy = -g.size[1];
    if (cz < g.low[2]) oz = g.size[2];
    if (cz > g.high[2]) oz = -g.size[2];
    const bool wrap = (ox != 0) | (oy != 0) | (oz != 0);
    if (wrap) {
        for (int s = s0; s < s1; s++) {
            rx[s] = rx[s] + ox; ry[s] = ry[s] + oy; rz[s] = rz[s] + oz;
        }
        if (bin_site) {
            cx = rx[s0]; cy = ry[s0]; cz = rz[s0];
        } else {  // second updateCOM (LinkedCells.cpp:76)
            double ax = 0, ay = 0, az = 0, tm = 0;
            for (int s = s0; s < s1; s++) {
                const double ms = mass[s];
                ax = ax + rx[s] * ms; ay = ay + ry[s] * ms; az = az + rz[s] * ms;
                tm = tm + ms;
            }
            cx = ax / tm; cy = ay / tm; cz = az / tm;
        }
    }
    if (!(cx == cx) || !(cy == cy) || !(cz == cz) || cx - cx != 0.0 || cy - cy != 0.0 || cz - cz != 0.0)
        atomicOr(flags, FLAG_NONFINITE);
    const int ix = cell_coord(cx, g.low[0], g.cs, g.nx);
    const int iy = cell_coord(cy, g.low[1], g.cs, g.ny);
    const int iz = cell_coord(cz, g.low[2], g.cs, g.nz);
    const uint32_t k = (uint32_t)(ix + g.nx * (iy + g.ny * iz));
    const uint32_t base = atomicAdd(&count[k], (uint32_t)(s1 - s0));
    for (int s = s0; s < s1; s++) {
        key[s] = k;
        slot[s] = base + (uint32_t)(s - s0);
        if (reset_f) { fx[s] = 0.0; fy[s] = 0.0; fz[s] = 0.0; }
    }
}

// One thread per molecule (a single-site molecule is one site).
//  [optional] v *= beta; v += f dt/(2m); r += v dt       (rescale + Integrator Step0)
//  then the container update of bin_unit
__global__ __launch_bounds__(kBlock) void k_com_wrap_key(
    int nmol, const uint32_t* __restrict__ mol_start, double* __restrict__ rx, double* __restrict__ ry,
    double* __restrict__ rz, double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz, const double* __restrict__ mass,
    uint32_t* __restrict__ key, uint32_t* __restrict__ slot, uint32_t* __restrict__ count, Grid g, int bin_site,
    int drift, int reset_f, double dt, double dt_half, const double* __restrict__ beta_src,
    uint32_t* __restrict__ flags) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const int s0 = mol_start ? (int)mol_start[m] : m;
    const int s1 = mol_start ? (int)mol_start[m + 1] : m + 1;
    if (drift) drift_unit(s0, s1, rx, ry, rz, vx, vy, vz, fx, fy, fz, mass, dt, dt_half, beta_src);
    bin_unit(s0, s1, rx, ry, rz, fx, fy, fz, mass, key, slot, count, g, bin_site, reset_f, flags);
}

// Step0 of a Verlet-list run (single-site molecules, or sites binned on their own
// positions).  req[0] != 0 marks a rebuild step:
// then the container update follows as in k_com_wrap_key.  Otherwise the site keeps its
// build-time cell and
// no wrap happens (the lists carry the image offsets of the build); only its FP32
// position relative to that cell's centre is refreshed, as k_sort_repack computes it.
// zero_f: the force is zeroed on every step (the step's first force kernel adds to it).
// Thread 0 clears the request word of the other step parity (the step's last force kernel
// sets it when some site would move beyond half the skin at the next drift) and counts
// the rebuilds.  words: [0], [1] the rebuild words of the two parities, [2] rebuilds.
__global__ __launch_bounds__(kBlock) void k_skin_drift(
    int n, double* __restrict__ rx, double* __restrict__ ry, double* __restrict__ rz, double* __restrict__ vx,
    double* __restrict__ vy, double* __restrict__ vz, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, const double* __restrict__ mass, uint32_t* __restrict__ key,
    uint32_t* __restrict__ slot, uint32_t* __restrict__ count, Grid g, int bin_site, double dt, double dt_half,
    uint32_t* __restrict__ words, int par, float4* __restrict__ spos,
    const uint32_t* __restrict__ inv, int zero_f, uint32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool rebuild = words[par] != 0u;
    if (i == 0) {
        words[par ^ 1] = 0u;
        if (rebuild) words[2] += 1u;
    }
    if (i >= n) return;
    drift_unit(i, i + 1, rx, ry, rz, vx, vy, vz, fx, fy, fz, mass, dt, dt_half, nullptr);
    if (rebuild) {
        bin_unit(i, i + 1, rx, ry, rz, fx, fy, fz, mass, key, slot, count, g, bin_site, 1, flags);
        return;
    }
    if (zero_f) { fx[i] = 0.0; fy[i] = 0.0; fz[i] = 0.0; }  // the step's forces add onto zero
    const double x = rx[i], y = ry[i], z = rz[i];
    if (!(x - x == 0.0) || !(y - y == 0.0) || !(z - z == 0.0)) atomicOr(flags, FLAG_NONFINITE);
    const uint32_t k = key[i];
    const int ix = (int)(k % (uint32_t)g.nx);
    const int iy = (int)((k / (uint32_t)g.nx) % (uint32_t)g.ny);
    const int iz = (int)(k / (uint32_t)(g.nx * g.ny));
    const double ccx = g.low[0] + ((double)ix + 0.5) * g.cs;
    const double ccy = g.low[1] + ((double)iy + 0.5) * g.cs;
    const double ccz = g.low[2] + ((double)iz + 0.5) * g.cs;
    spos[inv[i]] = make_float4((float)(x - ccx), (float)(y - ccy), (float)(z - ccz), __uint_as_float((uint32_t)i));
}

__global__ void k_skin_arm(uint32_t* __restrict__ req) {
    if (threadIdx.x == 0) req[0] = 1u;
}

// Grids above kFuseCells cells: tiles of 1024 cells, coalesced.  k_scan_tiles sums each tile;
// k_scan_fin offsets a tile by the sum of the tiles before it, scans it in LDS,
// writes start[] and zeroes the counts.
constexpr int kScanTile = 1024;
// req (Verlet-list runs): the step's rebuild word; the binning kernels return at once
// on steps that keep their lists.
__global__ __launch_bounds__(kScanTile) void k_scan_tiles(int ncell, const uint32_t* __restrict__ count,
                                                          uint32_t* __restrict__ tsum,
                                                          const uint32_t* __restrict__ req) {
    if (req && req[0] == 0u) return;
    __shared__ uint32_t ws[kScanTile / 64];
    const int i = blockIdx.x * kScanTile + threadIdx.x;
    uint32_t v = i < ncell ? count[i] : 0u;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < kScanTile / 64; k++) s += ws[k];
        tsum[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kScanTile) void k_scan_fin(int ncell, uint32_t* __restrict__ count,
                                                        uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ tsum,
                                                        const uint32_t* __restrict__ req) {
    if (req && req[0] == 0u) return;
    __shared__ uint32_t part[kScanTile];
    __shared__ uint32_t base;
    const int t = threadIdx.x;
    if (t < 64) {
        uint32_t s = 0;
        for (int k = t; k < (int)blockIdx.x; k += 64) s += tsum[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (t == 0) base = s;
    }
    const int i = blockIdx.x * kScanTile + t;
    const uint32_t c = i < ncell ? count[i] : 0u;
    part[t] = c;
    __syncthreads();
    for (int off = 1; off < kScanTile; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    if (i < ncell) {
        start[i] = base + part[t] - c;
        count[i] = 0;
        if (i == ncell - 1) start[ncell] = base + part[t];
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter(int n, const uint32_t* __restrict__ key,
                                                    const uint32_t* __restrict__ slot,
                                                    const uint32_t* __restrict__ start, uint32_t* __restrict__ tmp,
                                                    const uint32_t* __restrict__ req) {
    if (req && req[0] == 0u) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmp[start[key[i]] + slot[i]] = (uint32_t)i;
}

// Grids of at most kFuseCells cells: every workgroup scans the counts itself in LDS
// (block 0 also writes start[]) and scatters its sites, so no separate scan launch
// is needed; k_sort_repack zeroes the counts afterwards.
constexpr int kFuseCells = 4096;
__global__ __launch_bounds__(kBlock) void k_scan_scatter(int n, int ncell, const uint32_t* __restrict__ count,
                                                         const uint32_t* __restrict__ key,
                                                         const uint32_t* __restrict__ slot,
                                                         uint32_t* __restrict__ start, uint32_t* __restrict__ tmp,
                                                         const uint32_t* __restrict__ req) {
    if (req && req[0] == 0u) return;
    __shared__ uint32_t pre[kFuseCells + 1];
    __shared__ uint32_t wsum[kBlock / 64];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int chunk = (ncell + kBlock - 1) / kBlock;
    const int b0 = min(ncell, t * chunk), b1 = min(ncell, b0 + chunk);
    uint32_t s = 0;
    for (int i = b0; i < b1; i++) {
        const uint32_t c = count[i];
        pre[i] = c;
        s += c;
    }
    uint32_t v = s;  // inclusive scan of the chunk sums across the wave
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint32_t run = v - s, total = 0;
    for (int k = 0; k < kBlock / 64; k++) {
        if (k < w) run += wsum[k];
        total += wsum[k];
    }
    for (int i = b0; i < b1; i++) {
        const uint32_t c = pre[i];
        pre[i] = run;
        run += c;
    }
    if (t == 0) pre[ncell] = total;
    __syncthreads();
    if (blockIdx.x == 0)
        for (int i = t; i <= ncell; i += kBlock) start[i] = pre[i];
    const int i = blockIdx.x * blockDim.x + t;
    if (i < n) tmp[pre[key[i]] + slot[i]] = (uint32_t)i;
}

// Stable order inside a cell (ascending site index, as Cell::addIndex appends in
// site order, Cell.cpp:32-34) and the FP32 repack of the cell-sorted positions
// relative to the cell centre.
__global__ __launch_bounds__(kBlock) void k_sort_repack(
    int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ start, const uint32_t* __restrict__ tmp,
    uint32_t* __restrict__ sorted, const double* __restrict__ rx, const double* __restrict__ ry,
    const double* __restrict__ rz, float4* __restrict__ spos, const double* __restrict__ eps,
    const double* __restrict__ sig, float2* __restrict__ sprm, const uint32_t* __restrict__ idcls,
    uint32_t* __restrict__ sid, uint32_t* __restrict__ inv, Grid g, uint32_t* __restrict__ zero_count,
    const uint32_t* __restrict__ req, float4* __restrict__ spb) {
    if (req && req[0] == 0u) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (zero_count && i < g.ncell) zero_count[i] = 0;  // histogram of the next binning
    if (i >= n) return;
    const uint32_t k = key[i];
    const uint32_t s = start[k], e = start[k + 1];
    uint32_t rank = 0;
    for (uint32_t t = s; t < e; t++) rank += (tmp[t] < (uint32_t)i) ? 1u : 0u;
    const uint32_t p = s + rank;
    sorted[p] = (uint32_t)i;
    inv[i] = p;
    const int ix = (int)(k % (uint32_t)g.nx);
    const int iy = (int)((k / (uint32_t)g.nx) % (uint32_t)g.ny);
    const int iz = (int)(k / (uint32_t)(g.nx * g.ny));
    const double ccx = g.low[0] + ((double)ix + 0.5) * g.cs;
    const double ccy = g.low[1] + ((double)iy + 0.5) * g.cs;
    const double ccz = g.low[2] + ((double)iz + 0.5) * g.cs;
    const float4 q = make_float4((float)(rx[i] - ccx), (float)(ry[i] - ccy), (float)(rz[i] - ccz),
                                 __uint_as_float((uint32_t)i));
    spos[p] = q;
    if (spb) spb[p] = q;  // Verlet-list runs: the position of the rebuild
    if (sprm) sprm[p] = make_float2((float)sig[i], (float)sqrt(eps[i]));
    if (sid) sid[p] = idcls[i];
}

// Integrator::Step1 (Integrator.cpp:34-37) plus the kinetic sums of
// TemperatureSensor::measure (TemperatureSensor.cpp:32-36) as per-block partials.
__global__ __launch_bounds__(kBlock) void k_kick(int n, double* __restrict__ vx, double* __restrict__ vy,
                                                 double* __restrict__ vz, const double* __restrict__ fx,
                                                 const double* __restrict__ fy, const double* __restrict__ fz,
                                                 const double* __restrict__ mass, const double* __restrict__ rx,
                                                 const double* __restrict__ ry, const double* __restrict__ rz,
                                                 double dt_half, int with_temp, Grid g, double* __restrict__ part,
                                                 uint32_t* __restrict__ flags) {
    double mv2 = 0, cnt = 0, ke = 0;
    bool bad = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double k = dt_half / mass[i];
        double x = vx[i] + fx[i] * k;
        double y = vy[i] + fy[i] * k;
        double z = vz[i] + fz[i] * k;
        vx[i] = x; vy[i] = y; vz[i] = z;
        if (with_temp) {
            const double px = rx[i], py = ry[i], pz = rz[i];
            double vv = 0;
            vv = vv + x * x; vv = vv + y * y; vv = vv + z * z;
            bad |= !(vv - vv == 0.0);
            ke = ke + mass[i] * vv;
            if (px >= g.low[0] && py >= g.low[1] && pz >= g.low[2] && px <= g.high[0] && py <= g.high[1] &&
                pz <= g.high[2]) {
                mv2 = mv2 + mass[i] * vv;
                cnt = cnt + 1.0;
            }
        }
    }
    if (bad) atomicOr(flags, FLAG_NONFINITE);
    if (!with_temp) return;
    __shared__ double sh[3][kBlock];
    sh[0][threadIdx.x] = mv2; sh[1][threadIdx.x] = cnt; sh[2][threadIdx.x] = ke;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[0 * gridDim.x + blockIdx.x] = sh[0][0];
        part[1 * gridDim.x + blockIdx.x] = sh[1][0];
        part[2 * gridDim.x + blockIdx.x] = sh[2][0];
    }
}

// Integrator::Step0 without the container update (Integrator.cpp:28-32).
__global__ __launch_bounds__(kBlock) void k_kick_drift(int n, double* __restrict__ rx, double* __restrict__ ry,
                                                       double* __restrict__ rz, double* __restrict__ vx,
                                                       double* __restrict__ vy, double* __restrict__ vz,
                                                       const double* __restrict__ fx, const double* __restrict__ fy,
                                                       const double* __restrict__ fz, const double* __restrict__ mass,
                                                       double dt, double dt_half) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double k = dt_half / mass[i];
    const double x = vx[i] + fx[i] * k, y = vy[i] + fy[i] * k, z = vz[i] + fz[i] * k;
    vx[i] = x; vy[i] = y; vz[i] = z;
    rx[i] = rx[i] + x * dt;
    ry[i] = ry[i] + y * dt;
    rz[i] = rz[i] + z * dt;
}

// PressureSensor::MomentumKernel (PressureSensor.cpp:142-150): sum_i m v_a v_b over
// all sites, components XX XY YY XZ YZ ZZ, per-block partials.
__global__ __launch_bounds__(kBlock) void k_kin_tensor(int n, const double* __restrict__ vx,
                                                       const double* __restrict__ vy,
                                                       const double* __restrict__ vz,
                                                       const double* __restrict__ mass, double* __restrict__ part) {
    double w[kSym] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double x = vx[i], y = vy[i], z = vz[i], m = mass[i];
        w[0] += x * x * m;
        w[1] += x * y * m;
        w[2] += y * y * m;
        w[3] += x * z * m;
        w[4] += y * z * m;
        w[5] += z * z * m;
    }
    block_store_sym(part, gridDim.x, blockIdx.x, w);
}

// Deterministic sum of nparts partials (fixed order, one wave per slot).
__device__ double sum_partials(const double* __restrict__ p, int nparts) {
    __shared__ double sh[kBlock];
    double s = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// T = temp_factor * mv2/(3 count); beta = sqrt(t0/T); v *= beta
// (TemperatureSensor.cpp:41-42, VelocityScaling.cpp:12-25).  Every block sums
// the kinetic partials itself in the same order, so beta is identical everywhere.
// t0 comes from scal[kScalT0] when t0_dev is set (the captured step of komd_run, so
// that a replay uses the target of the current run, not the one of the capture).
__global__ __launch_bounds__(kBlock) void k_rescale(int n, double* __restrict__ vx, double* __restrict__ vy,
                                                    double* __restrict__ vz, const double* __restrict__ part,
                                                    int nparts, double temp_factor, double t0, int t0_dev,
                                                    double* __restrict__ scal, int apply) {
    const double mv2 = sum_partials(part, nparts);
    const double cnt = sum_partials(part + nparts, nparts);
    const double ke = sum_partials(part + 2 * nparts, nparts);
    const double T = temp_factor * (mv2 / (cnt * 3.0));
    const double beta = sqrt((t0_dev ? scal[kScalT0] : t0) / T);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[kScalMv2] = mv2; scal[kScalCount] = cnt; scal[kScalBeta] = beta; scal[kScalTemp] = T;
        scal[kScalKin] = 0.5 * ke;
    }
    if (!apply) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        vx[i] = vx[i] * beta; vy[i] = vy[i] * beta; vz[i] = vz[i] * beta;
    }
}

__global__ void k_run_scalars(double* __restrict__ scal, double t0) {
    if (threadIdx.x == 0) {
        scal[kScalBeta] = 1.0;
        scal[kScalT0] = t0;
    }
}

__global__ void k_flag_snapshot(const uint32_t* __restrict__ flags, uint32_t* __restrict__ host_slot) {
    if (threadIdx.x == 0) host_slot[0] = flags[0];
}

__global__ __launch_bounds__(kBlock) void k_reduce_final(const double* __restrict__ part, int nparts, int nslots,
                                                         double* __restrict__ out) {
    for (int s = 0; s < nslots; s++) {
        const double v = sum_partials(part + (size_t)s * nparts, nparts);
        if (threadIdx.x == 0) out[s] = v;
    }
}

}  // namespace

hipError_t launch_com_wrap_key(komd_ctx* c, bool do_drift, bool reset_f) {
    ScopedTimer tm(c, "bin_key");
    auto& d = c->d;
    const bool site = c->prm.bin_mode == KOMD_BIN_SITE;
    const int units = site ? (int)c->n : (int)c->nmol;
    const uint32_t* ms = (site || !c->multisite) ? nullptr : d.mol_start;
    k_com_wrap_key<<<(units + kKeyBlock - 1) / kKeyBlock, kKeyBlock, 0, c->stream>>>(
        units, ms, d.rx, d.ry, d.rz, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.key, d.slot, d.count, c->grid,
        site ? 1 : 0, do_drift ? 1 : 0, reset_f ? 1 : 0, c->prm.delta_t, c->prm.delta_t * 0.5, nullptr, d.flags);
    return hipGetLastError();
}

// rescale-fused variant used by komd_run: v *= beta (from scal[kScalBeta]) before the kick.
hipError_t launch_com_wrap_key_beta(komd_ctx* c, const double* beta_dev) {
    ScopedTimer tm(c, "bin_key");
    auto& d = c->d;
    const bool site = c->prm.bin_mode == KOMD_BIN_SITE;
    const int units = site ? (int)c->n : (int)c->nmol;
    const uint32_t* ms = (site || !c->multisite) ? nullptr : d.mol_start;
    k_com_wrap_key<<<(units + kKeyBlock - 1) / kKeyBlock, kKeyBlock, 0, c->stream>>>(
        units, ms, d.rx, d.ry, d.rz, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.key, d.slot, d.count, c->grid,
        site ? 1 : 0, 1, 1, c->prm.delta_t, c->prm.delta_t * 0.5, beta_dev, d.flags);
    return hipGetLastError();
}

hipError_t launch_skin_drift(komd_ctx* c, int par, bool zero_f) {
    ScopedTimer tm(c, "bin_key");
    auto& d = c->d;
    const int n = (int)c->n;
    if (c->multisite && c->prm.bin_mode != KOMD_BIN_SITE) return hipErrorInvalidValue;
    k_skin_drift<<<(n + kKeyBlock - 1) / kKeyBlock, kKeyBlock, 0, c->stream>>>(
        n, d.rx, d.ry, d.rz, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.key, d.slot, d.count, c->grid,
        c->prm.bin_mode == KOMD_BIN_SITE ? 1 : 0, c->prm.delta_t, c->prm.delta_t * 0.5, d.skin_req, par, d.spos,
        d.inv, zero_f ? 1 : 0, d.flags);
    return hipGetLastError();
}

hipError_t launch_skin_arm(komd_ctx* c, int par) {
    k_skin_arm<<<1, 64, 0, c->stream>>>(c->d.skin_req + par);
    return hipGetLastError();
}

hipError_t launch_bin(komd_ctx* c, const uint32_t* req, float4* spb) {
    auto& d = c->d;
    const int n = (int)c->n;
    const int C = c->grid.ncell;
    const bool fused = C <= kFuseCells;
    if (fused) {
        ScopedTimer tm(c, "bin_scatter");
        k_scan_scatter<<<std::max(1, (n + kBlock - 1) / kBlock), kBlock, 0, c->stream>>>(n, C, d.count, d.key, d.slot,
                                                                                       d.start, d.tmp, req);
        KOMD_CHECK_HIP(hipGetLastError());
    } else {
        {
            ScopedTimer tm(c, "bin_scan");
            const int tiles = (C + kScanTile - 1) / kScanTile;
            k_scan_tiles<<<tiles, kScanTile, 0, c->stream>>>(C, d.count, d.scanp, req);
            KOMD_CHECK_HIP(hipGetLastError());
            k_scan_fin<<<tiles, kScanTile, 0, c->stream>>>(C, d.count, d.start, d.scanp, req);
            KOMD_CHECK_HIP(hipGetLastError());
        }
        ScopedTimer tm(c, "bin_scatter");
        k_scatter<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(n, d.key, d.slot, d.start, d.tmp, req);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    {
        ScopedTimer tm(c, "bin_sort");
        const int units = std::max(n, fused ? C : 0);
        k_sort_repack<<<(units + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(
            n, d.key, d.start, d.tmp, d.sorted, d.rx, d.ry, d.rz, d.spos, d.eps, d.sig, c->uniform ? nullptr : d.sprm,
            d.idcls, c->multisite ? d.sid : nullptr, d.inv, c->grid, fused ? d.count : nullptr, req, spb);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t launch_kick_drift(komd_ctx* c) {
    ScopedTimer tm(c, "kick_drift");
    auto& d = c->d;
    const int n = (int)c->n;
    k_kick_drift<<<(n + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(n, d.rx, d.ry, d.rz, d.vx, d.vy, d.vz, d.fx,
                                                                        d.fy, d.fz, d.mass, c->prm.delta_t,
                                                                        c->prm.delta_t * 0.5);
    return hipGetLastError();
}

hipError_t launch_kick(komd_ctx* c, bool with_temp) {
    ScopedTimer tm(c, "kick");
    auto& d = c->d;
    const int n = (int)c->n;
    k_kick<<<c->red_blocks, kBlock, 0, c->stream>>>(n, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.rx, d.ry, d.rz,
                                                     c->prm.delta_t * 0.5, with_temp ? 1 : 0, c->grid, d.red, d.flags);
    return hipGetLastError();
}

hipError_t launch_temperature(komd_ctx* c) {
    // kinetic partials without changing v (dt_half = 0 makes the kick an identity: v + f*0 = v)
    ScopedTimer tm(c, "temperature");
    auto& d = c->d;
    const int n = (int)c->n;
    k_kick<<<c->red_blocks, kBlock, 0, c->stream>>>(n, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.rx, d.ry, d.rz,
                                                     0.0, 1, c->grid, d.red, d.flags);
    KOMD_CHECK_HIP(hipGetLastError());
    k_rescale<<<1, kBlock, 0, c->stream>>>(n, d.vx, d.vy, d.vz, d.red, c->red_blocks, c->prm.temp_factor, 1.0, 0,
                                           d.scal, 0);
    return hipGetLastError();
}

hipError_t launch_rescale(komd_ctx* c, double t0) {
    ScopedTimer tm(c, "rescale");
    auto& d = c->d;
    const int n = (int)c->n;
    const int blocks = std::min(1024, (n + kBlock - 1) / kBlock);
    k_rescale<<<blocks, kBlock, 0, c->stream>>>(n, d.vx, d.vy, d.vz, d.red, c->red_blocks, c->prm.temp_factor, t0,
                                                0, d.scal, 1);
    return hipGetLastError();
}

hipError_t launch_beta_only(komd_ctx* c) {
    auto& d = c->d;
    k_rescale<<<1, kBlock, 0, c->stream>>>((int)c->n, d.vx, d.vy, d.vz, d.red, c->red_blocks, c->prm.temp_factor,
                                           0.0, 1, d.scal, 0);
    return hipGetLastError();
}

hipError_t launch_run_scalars(komd_ctx* c, double t0) {
    k_run_scalars<<<1, 64, 0, c->stream>>>(c->d.scal, t0);
    return hipGetLastError();
}

hipError_t launch_flag_snapshot(komd_ctx* c, uint32_t* host_slot) {
    k_flag_snapshot<<<1, 64, 0, c->stream>>>(c->d.flags, host_slot);
    return hipGetLastError();
}

hipError_t launch_kin_tensor(komd_ctx* c) {
    k_kin_tensor<<<kKinBlocks, kBlock, 0, c->stream>>>((int)c->n, c->d.vx, c->d.vy, c->d.vz, c->d.mass, c->d.vpart);
    return hipGetLastError();
}

hipError_t launch_reduce_parts(komd_ctx* c, const double* part, int nparts, int nslots, double* out_dev) {
    k_reduce_final<<<1, kBlock, 0, c->stream>>>(part, nparts, nslots, out_dev);
    return hipGetLastError();
}

}  // namespace komd
