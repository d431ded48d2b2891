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

// Is the same-molecule pair (i, j), i < j, within the cutoff, decided as the pair pass
// decides it from the lower-index site: FP64 |(r_i + 0) - (r_j + s_j)|^2 in that order,
// s_j the periodic image of j nearest i (the image of j's cell seen from i's cell)?
__device__ __forceinline__ uint32_t intra_in_range(double xi, double yi, double zi, double xj, double yj,
                                                   double zj, const Grid& g, double rc2) {
#pragma clang fp contract(off)
    const double Lx = g.size[0], Ly = g.size[1], Lz = g.size[2];
    const double ex = xj - xi, ey = yj - yi, ez = zj - zi;
    const double sx = ex > 0.5 * Lx ? -Lx : (ex < -0.5 * Lx ? Lx : 0.0);
    const double sy = ey > 0.5 * Ly ? -Ly : (ey < -0.5 * Ly ? Ly : 0.0);
    const double sz = ez > 0.5 * Lz ? -Lz : (ez < -0.5 * Lz ? Lz : 0.0);
    const double dx = (xi + 0.0) - (xj + sx), dy = (yi + 0.0) - (yj + sy), dz = (zi + 0.0) - (zj + sz);
    double d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    return d2 > rc2 ? 0u : 1u;
}
constexpr int kMolRegs = 4;

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
                                                    double* __restrict__ vpart, int vparts, int energy,
                                                    const uint32_t* __restrict__ mol_start, double rc2,
                                                    unsigned long long* __restrict__ intra, int nslots) {
    double ub = 0, ua = 0;
    float w[kSym] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    uint32_t nintra = 0;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < nmol; m += gridDim.x * blockDim.x) {
        if (intra) {
            // Verlet-list runs: the lists hold no same-molecule pairs, so the in-range ones are
            // counted here (molecules of up to kMolRegs sites from registers)
            const uint32_t s0 = mol_start[m], s1 = mol_start[m + 1];
            const int ns = (int)(s1 - s0);
            if (ns <= kMolRegs) {
                double x[kMolRegs], y[kMolRegs], z[kMolRegs];
#pragma unroll
                for (int u = 0; u < kMolRegs; u++) {
                    const uint32_t su = u < ns ? s0 + u : s0;
                    x[u] = rx[su]; y[u] = ry[su]; z[u] = rz[su];
                }
#pragma unroll
                for (int u = 0; u < kMolRegs; u++)
#pragma unroll
                    for (int w2 = u + 1; w2 < kMolRegs; w2++)
                        if (w2 < ns) nintra += intra_in_range(x[u], y[u], z[u], x[w2], y[w2], z[w2], g, rc2);
            } else {
                for (uint32_t i = s0; i < s1; i++)
                    for (uint32_t j = i + 1; j < s1; j++)
                        nintra += intra_in_range(rx[i], ry[i], rz[i], rx[j], ry[j], rz[j], g, rc2);
            }
        }
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
    if (intra) {
        // one add per workgroup, each to its own slot (same-address atomics serialise)
        __shared__ uint32_t wsum[4];
        for (int o = 32; o > 0; o >>= 1) nintra += __shfl_xor(nintra, o);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = nintra;
        __syncthreads();
        if (threadIdx.x == 0)
            atomicAdd(&intra[blockIdx.x % nslots],
                      (unsigned long long)wsum[0] + wsum[1] + wsum[2] + wsum[3]);
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

// Bonded forces per site (force passes of topologies whose terms cross molecules, in place of
// the atomics of k_bonds / k_angles): each site sums the bond and angle terms it takes part
// in, with their FP32 expressions per term, in FP64 in the order of its term list, and adds
// the total to its force once (a fixed order).  Term reference: term << 3 | angle << 2 | role (bond: 0 first site, 1 second;
// angle: 0, 1 middle, 2).  intra (Verlet-list runs of multi-site molecules, whose lists leave
// same-molecule pairs out): also count the in-range pairs (i, j > i) of the site's molecule,
// decided as the pair pass decides them, into the Verlet-list pair counters (without terms:
// the count alone, for runs without bonded forces).
constexpr int kSiteTerms = 6;
constexpr uint32_t kNoTerm = 0xffffffffu;

__device__ __forceinline__ void site_term(uint32_t e, const uint32_t* __restrict__ bonds,
                                          const uint32_t* __restrict__ ang, float kb, float r0, float ka, float th0,
                                          const double* __restrict__ rx, const double* __restrict__ ry,
                                          const double* __restrict__ rz, const Grid& g, double& sx, double& sy,
                                          double& sz) {
    const uint32_t t = e >> 3, role = e & 3u;
    if (e & 4u) {
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
        if (role == 0) {
            sx += (double)fix; sy += (double)fiy; sz += (double)fiz;
        } else if (role == 2) {
            sx += (double)fkx; sy += (double)fky; sz += (double)fkz;
        } else {
            sx -= (double)(fix + fkx); sy -= (double)(fiy + fky); sz -= (double)(fiz + fkz);
        }
    } else {
        const uint32_t i = bonds[2 * t], j = bonds[2 * t + 1];
        const float dx = min_image(rx[j] - rx[i], g.size[0]);
        const float dy = min_image(ry[j] - ry[i], g.size[1]);
        const float dz = min_image(rz[j] - rz[i], g.size[2]);
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        const float dl = len - r0;
        const float gfac = -kb * dl / len;
        if (role == 1) {
            sx += (double)(gfac * dx); sy += (double)(gfac * dy); sz += (double)(gfac * dz);
        } else {
            sx -= (double)(gfac * dx); sy -= (double)(gfac * dy); sz -= (double)(gfac * dz);
        }
    }
}

__global__ __launch_bounds__(256) void k_bonded_site(int n, const uint32_t* __restrict__ tstart,
                                                     const uint32_t* __restrict__ terms,
                                                     const uint32_t* __restrict__ bonds,
                                                     const uint32_t* __restrict__ ang, float kb, float r0, float ka,
                                                     float th0, const double* __restrict__ rx,
                                                     const double* __restrict__ ry, const double* __restrict__ rz,
                                                     double* __restrict__ fx, double* __restrict__ fy,
                                                     double* __restrict__ fz, Grid g,
                                                     const uint32_t* __restrict__ mol_of,
                                                     const uint32_t* __restrict__ mol_start, double rc2,
                                                     unsigned long long* __restrict__ intra, int nslots) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nintra = 0;
    if (i < n) {
        if (tstart) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            const uint32_t t0 = tstart[i], t1 = tstart[i + 1];
            for (uint32_t tb = t0; tb < t1; tb += kSiteTerms) {
                uint32_t e[kSiteTerms];
#pragma unroll
                for (int k = 0; k < kSiteTerms; k++) e[k] = tb + k < t1 ? terms[tb + k] : kNoTerm;
#pragma unroll
                for (int k = 0; k < kSiteTerms; k++)
                    if (e[k] != kNoTerm) site_term(e[k], bonds, ang, kb, r0, ka, th0, rx, ry, rz, g, sx, sy, sz);
            }
            fx[i] = fx[i] + sx;
            fy[i] = fy[i] + sy;
            fz[i] = fz[i] + sz;
        }
        if (intra) {
            const uint32_t s1 = mol_start[mol_of[i] + 1];
            const double xi = rx[i], yi = ry[i], zi = rz[i];
            for (uint32_t jb = (uint32_t)i + 1; jb < s1; jb += kMolRegs) {
                double x[kMolRegs], y[kMolRegs], z[kMolRegs];
#pragma unroll
                for (int u = 0; u < kMolRegs; u++) {
                    const uint32_t j = jb + u < s1 ? jb + u : (uint32_t)i;
                    x[u] = rx[j]; y[u] = ry[j]; z[u] = rz[j];
                }
#pragma unroll
                for (int u = 0; u < kMolRegs; u++)
                    if (jb + u < s1) nintra += intra_in_range(xi, yi, zi, x[u], y[u], z[u], g, rc2);
            }
        }
    }
    if (intra) {
        // one add per wave, each workgroup to its own slot (same-address atomics serialise)
        for (int o = 32; o > 0; o >>= 1) nintra += __shfl_xor(nintra, o);
        if ((threadIdx.x & 63) == 0 && nintra) atomicAdd(&intra[blockIdx.x % nslots], (unsigned long long)nintra);
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
hipError_t launch_bonded(komd_ctx* c, bool energy, bool intra) {
    ScopedTimer tm(c, "bonded");
    auto& d = c->d;
    const int blocks = 256;
    if (intra && (!d.mol_start || !d.cnt_trip || !c->ids_runs)) return hipErrorInvalidValue;
    if (!energy && !c->bonded_by_mol) return launch_bonded_sites(c, true, intra);
    if (c->bonded_by_mol) {
        const int nmol = (int)c->nmol;
        // the Verlet-list passes' first pair-count slot
        unsigned long long* ic = intra ? d.cnt_trip + atm_blocks(c) : nullptr;
        k_bonded_mol<<<blocks, 256, 0, c->stream>>>(nmol, d.bond_mol, d.angle_mol, d.bonds, d.angles, (float)c->kb,
                                                     (float)c->r0, (float)c->ka, (float)c->theta0, d.rx, d.ry, d.rz,
                                                     d.fx, d.fy, d.fz, c->grid, d.red, d.vpart, 2 * blocks,
                                                     energy ? 1 : 0, d.mol_start,
                                                     c->prm.cutoff * c->prm.cutoff, ic, atm_blocks(c));
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

hipError_t launch_bonded_sites(komd_ctx* c, bool terms, bool intra) {
    ScopedTimer tm(c, "bonded");
    auto& d = c->d;
    if ((terms && (!d.site_tstart || !d.site_terms)) ||
        (intra && (!d.mol_start || !d.idcls || !d.cnt_trip || !c->ids_runs)))
        return hipErrorInvalidValue;
    if (!terms && !intra) return hipSuccess;
    const int n = (int)c->n;
    k_bonded_site<<<(n + 255) / 256, 256, 0, c->stream>>>(
        n, terms ? d.site_tstart : nullptr, d.site_terms, d.bonds, d.angles, (float)c->kb, (float)c->r0,
        (float)c->ka, (float)c->theta0, d.rx, d.ry, d.rz, d.fx, d.fy, d.fz, c->grid, d.idcls, d.mol_start,
        c->prm.cutoff * c->prm.cutoff, intra ? d.cnt_trip + atm_blocks(c) : nullptr, atm_blocks(c));
    return hipGetLastError();
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
