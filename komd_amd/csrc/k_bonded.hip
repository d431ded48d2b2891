// Harmonic bond and angle forces, force clamp -- gfx950.
#include "komd_internal.hpp"
#include "komd_reduce.hpp"

namespace komd {
namespace {

// ---- harmonic bonds and angles ---------------------------------------
// FP64 minimum-image displacements rounded to FP32, FP32 force math, FP64
// atomic accumulation.  U_bond = kb/2 (r - r0)^2, U_angle = ka/2 (theta - theta0)^2.
__device__ __forceinline__ float min_image(double d, double L) { return (float)(d - L * rint(d / L)); }

__global__ __launch_bounds__(256) void k_bonds(int nb, const uint32_t* __restrict__ bonds, float kb, float r0,
                                               const double* __restrict__ rx, const double* __restrict__ ry,
                                               const double* __restrict__ rz, double* __restrict__ fx,
                                               double* __restrict__ fy, double* __restrict__ fz, Grid g,
                                               double* __restrict__ part, double* __restrict__ vpart, int vparts,
                                               int energy) {
    double u = 0;
    float w[kSym] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nb; t += gridDim.x * blockDim.x) {
        const uint32_t i = bonds[2 * t], j = bonds[2 * t + 1];
        const float dx = min_image(rx[j] - rx[i], g.size[0]);
        const float dy = min_image(ry[j] - ry[i], g.size[1]);
        const float dz = min_image(rz[j] - rz[i], g.size[2]);
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        const float dl = len - r0;
        const float gfac = -kb * dl / len;  // F_j = g d, F_i = -g d
        unsafeAtomicAdd(&fx[j], (double)(gfac * dx));
        unsafeAtomicAdd(&fy[j], (double)(gfac * dy));
        unsafeAtomicAdd(&fz[j], (double)(gfac * dz));
        unsafeAtomicAdd(&fx[i], (double)(-gfac * dx));
        unsafeAtomicAdd(&fy[i], (double)(-gfac * dy));
        unsafeAtomicAdd(&fz[i], (double)(-gfac * dz));
        if (energy) {
            u += 0.5 * (double)(kb * dl * dl);
            sym_outer<float>(w, dx, dy, dz, gfac * dx, gfac * dy, gfac * dz);  // (r_j - r_i) (x) F_j
        }
    }
    if (energy) {
        block_store_sym(vpart, vparts, blockIdx.x, w);
        __shared__ double sh[256];
        sh[threadIdx.x] = u;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
    }
}

__global__ __launch_bounds__(256) void k_angles(int na, const uint32_t* __restrict__ ang, float ka, float th0,
                                                const double* __restrict__ rx, const double* __restrict__ ry,
                                                const double* __restrict__ rz, double* __restrict__ fx,
                                                double* __restrict__ fy, double* __restrict__ fz, Grid g,
                                                double* __restrict__ part, double* __restrict__ vpart, int vparts,
                                                int vslot0, int energy) {
    double u = 0;
    float w[kSym] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < na; t += gridDim.x * blockDim.x) {
        const uint32_t i = ang[3 * t], j = ang[3 * t + 1], k = ang[3 * t + 2];
        const float ax = min_image(rx[i] - rx[j], g.size[0]);
        const float ay = min_image(ry[i] - ry[j], g.size[1]);
        const float az = min_image(rz[i] - rz[j], g.size[2]);
        const float bx = min_image(rx[k] - rx[j], g.size[0]);
        const float by = min_image(ry[k] - ry[j], g.size[1]);
        const float bz = min_image(rz[k] - rz[j], g.size[2]);
        const float la2 = ax * ax + ay * ay + az * az, lb2 = bx * bx + by * by + bz * bz;
        const float inv_ab = rsqrtf(la2 * lb2);
        float c = (ax * bx + ay * by + az * bz) * inv_ab;
        c = fminf(1.f, fmaxf(-1.f, c));
        const float th = acosf(c);
        const float dth = th - th0;
        const float s = fmaxf(sqrtf(fmaxf(0.f, 1.f - c * c)), 1e-6f);
        const float pre = ka * dth / s;
        const float ia2 = 1.f / la2, ib2 = 1.f / lb2;
        const float fix = pre * (bx * inv_ab - c * ax * ia2), fiy = pre * (by * inv_ab - c * ay * ia2),
                    fiz = pre * (bz * inv_ab - c * az * ia2);
        const float fkx = pre * (ax * inv_ab - c * bx * ib2), fky = pre * (ay * inv_ab - c * by * ib2),
                    fkz = pre * (az * inv_ab - c * bz * ib2);
        unsafeAtomicAdd(&fx[i], (double)fix); unsafeAtomicAdd(&fy[i], (double)fiy); unsafeAtomicAdd(&fz[i], (double)fiz);
        unsafeAtomicAdd(&fx[k], (double)fkx); unsafeAtomicAdd(&fy[k], (double)fky); unsafeAtomicAdd(&fz[k], (double)fkz);
        unsafeAtomicAdd(&fx[j], -(double)(fix + fkx));
        unsafeAtomicAdd(&fy[j], -(double)(fiy + fky));
        unsafeAtomicAdd(&fz[j], -(double)(fiz + fkz));
        if (energy) {
            u += 0.5 * (double)(ka * dth * dth);
            sym_outer<float>(w, ax, ay, az, fix, fiy, fiz);  // levers from the centre site j
            sym_outer<float>(w, bx, by, bz, fkx, fky, fkz);
        }
    }
    if (energy) {
        block_store_sym(vpart, vparts, vslot0 + blockIdx.x, w);
        __shared__ double sh[256];
        sh[threadIdx.x] = u;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
    }
}

// Bonds and angles grouped by molecule (CSR): one thread per molecule updates the
// forces of its own sites with plain read-modify-writes (no two threads share a
// site), in a fixed order.  Same terms and arithmetic as k_bonds / k_angles.
__global__ __launch_bounds__(256) void k_bonded_mol(int nmol, const uint32_t* __restrict__ bstart,
                                                    const uint32_t* __restrict__ astart,
                                                    const uint32_t* __restrict__ bonds,
                                                    const uint32_t* __restrict__ ang, float kb, float r0, float ka,
                                                    float th0, const double* __restrict__ rx,
                                                    const double* __restrict__ ry, const double* __restrict__ rz,
                                                    double* __restrict__ fx, double* __restrict__ fy,
                                                    double* __restrict__ fz, Grid g, double* __restrict__ part,
                                                    double* __restrict__ vpart, int vparts, int energy) {
    double ub = 0, ua = 0;
    float w[kSym] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < nmol; m += gridDim.x * blockDim.x) {
        for (uint32_t t = bstart[m]; t < bstart[m + 1]; t++) {
            const uint32_t i = bonds[2 * t], j = bonds[2 * t + 1];
            const float dx = min_image(rx[j] - rx[i], g.size[0]);
            const float dy = min_image(ry[j] - ry[i], g.size[1]);
            const float dz = min_image(rz[j] - rz[i], g.size[2]);
            const float len = sqrtf(dx * dx + dy * dy + dz * dz);
            const float dl = len - r0;
            const float gfac = -kb * dl / len;
            fx[j] += (double)(gfac * dx);
            fy[j] += (double)(gfac * dy);
            fz[j] += (double)(gfac * dz);
            fx[i] -= (double)(gfac * dx);
            fy[i] -= (double)(gfac * dy);
            fz[i] -= (double)(gfac * dz);
            if (energy) {
                ub += 0.5 * (double)(kb * dl * dl);
                sym_outer<float>(w, dx, dy, dz, gfac * dx, gfac * dy, gfac * dz);
            }
        }
        for (uint32_t t = astart[m]; t < astart[m + 1]; t++) {
            const uint32_t i = ang[3 * t], j = ang[3 * t + 1], k = ang[3 * t + 2];
            const float ax = min_image(rx[i] - rx[j], g.size[0]);
            const float ay = min_image(ry[i] - ry[j], g.size[1]);
            const float az = min_image(rz[i] - rz[j], g.size[2]);
            const float bx = min_image(rx[k] - rx[j], g.size[0]);
            const float by = min_image(ry[k] - ry[j], g.size[1]);
            const float bz = min_image(rz[k] - rz[j], g.size[2]);
            const float la2 = ax * ax + ay * ay + az * az, lb2 = bx * bx + by * by + bz * bz;
            const float inv_ab = rsqrtf(la2 * lb2);
            float c = (ax * bx + ay * by + az * bz) * inv_ab;
            c = fminf(1.f, fmaxf(-1.f, c));
            const float th = acosf(c);
            const float dth = th - th0;
            const float s = fmaxf(sqrtf(fmaxf(0.f, 1.f - c * c)), 1e-6f);
            const float pre = ka * dth / s;
            const float ia2 = 1.f / la2, ib2 = 1.f / lb2;
            const float fix = pre * (bx * inv_ab - c * ax * ia2), fiy = pre * (by * inv_ab - c * ay * ia2),
                        fiz = pre * (bz * inv_ab - c * az * ia2);
            const float fkx = pre * (ax * inv_ab - c * bx * ib2), fky = pre * (ay * inv_ab - c * by * ib2),
                        fkz = pre * (az * inv_ab - c * bz * ib2);
            fx[i] += (double)fix; fy[i] += (double)fiy; fz[i] += (double)fiz;
            fx[k] += (double)fkx; fy[k] += (double)fky; fz[k] += (double)fkz;
            fx[j] -= (double)(fix + fkx);
            fy[j] -= (double)(fiy + fky);
            fz[j] -= (double)(fiz + fkz);
            if (energy) {
                ua += 0.5 * (double)(ka * dth * dth);
                sym_outer<float>(w, ax, ay, az, fix, fiy, fiz);
                sym_outer<float>(w, bx, by, bz, fkx, fky, fkz);
            }
        }
    }
    if (energy) {
        block_store_sym(vpart, vparts, blockIdx.x, w);
        __shared__ double sh[2][256];
        sh[0][threadIdx.x] = ub;
        sh[1][threadIdx.x] = ua;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) {
                sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
                sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[blockIdx.x] = sh[0][0];
            part[gridDim.x + blockIdx.x] = sh[1][0];
        }
    }
}

// Limit_Force (Limit.cpp:19-25): f = max(min(f, fmax), -fmax), fmax = limit eps / sig.
__global__ __launch_bounds__(256) void k_limit(int n, double lim, const double* __restrict__ eps,
                                               const double* __restrict__ sig, double* __restrict__ fx,
                                               double* __restrict__ fy, double* __restrict__ fz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fm = lim * eps[i] / sig[i];
    double x = fx[i], y = fy[i], z = fz[i];
    x = x < fm ? x : fm; x = x > -fm ? x : -fm;
    y = y < fm ? y : fm; y = y > -fm ? y : -fm;
    z = z < fm ? z : fm; z = z > -fm ? z : -fm;
    fx[i] = x; fy[i] = y; fz[i] = z;
}

}  // namespace

// Energy mode: U partials in d.red ([0, 256) bonds, [256, 512) angles) and the virial
// tensor in d.vpart as kSym slots of 512 partials; the caller zeroes both first.
hipError_t launch_bonded(komd_ctx* c, bool energy) {
    ScopedTimer tm(c, "bonded");
    auto& d = c->d;
    const int blocks = 256;
    if (c->bonded_by_mol) {
        const int nmol = (int)c->nmol;
        k_bonded_mol<<<blocks, 256, 0, c->stream>>>(nmol, d.bond_mol, d.angle_mol, d.bonds, d.angles, (float)c->kb,
                                                     (float)c->r0, (float)c->ka, (float)c->theta0, d.rx, d.ry, d.rz,
                                                     d.fx, d.fy, d.fz, c->grid, d.red, d.vpart, 2 * blocks,
                                                     energy ? 1 : 0);
        return hipGetLastError();
    }
    if (c->nbonds > 0) {
        k_bonds<<<blocks, 256, 0, c->stream>>>((int)c->nbonds, d.bonds, (float)c->kb, (float)c->r0, d.rx, d.ry, d.rz,
                                               d.fx, d.fy, d.fz, c->grid, d.red, d.vpart, 2 * blocks,
                                               energy ? 1 : 0);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    if (c->nangles > 0) {
        k_angles<<<blocks, 256, 0, c->stream>>>((int)c->nangles, d.angles, (float)c->ka, (float)c->theta0, d.rx, d.ry,
                                                d.rz, d.fx, d.fy, d.fz, c->grid, d.red + blocks, d.vpart, 2 * blocks,
                                                blocks, energy ? 1 : 0);
        KOMD_CHECK_HIP(hipGetLastError());
    }
    return hipSuccess;
}

hipError_t launch_limit(komd_ctx* c) {
    if (!(c->prm.limit_factor > 0)) return hipSuccess;
    ScopedTimer tm(c, "limit");
    auto& d = c->d;
    const int n = (int)c->n;
    k_limit<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->prm.limit_factor, d.eps, d.sig, d.fx, d.fy, d.fz);
    return hipGetLastError();
}

}  // namespace komd
