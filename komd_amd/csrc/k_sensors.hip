// Radial distribution sensor -- gfx950.
//
// RDFSensor::measure (RDFSensor.cpp:23-35) histograms, for every ordered pair of
// distinct molecules, the minimum-image distance of their centres of mass
// (kernel RDFSensor.cpp:59-72).  Here:
//  * k_mol_com restates MoleculeContainer::updateCOM (MoleculeContainer.cpp:55-75)
//    per molecule, in FP64 with the reference's operation order.
//  * k_rdf walks unordered molecule pairs in 256 x 256 tiles (the column tile's
//    centres staged in LDS, one row molecule per thread) and adds 2 to the bin of
//    each pair within max_r: the minimum-image vector of (j, i) is exactly the
//    negation of that of (i, j), since round() is odd, so both orders fall in one
//    bin.  Each workgroup keeps a u32 histogram in LDS and adds it to the u64
//    totals once, after all its tiles.
// The file compiles without FMA contraction, so distances and bins are the FP64
// values the reference's expressions give.
#pragma clang fp contract(off)

#include <mutex>

#include "komd_internal.hpp"

namespace komd {
namespace {

constexpr int kRdfThreads = 256;  // one row molecule per thread
constexpr int kRdfTile = 256;     // molecules per tile side

// Scratch of komd_rdf_noalloc: the molecule centres (x, y, z blocks of kRdfScratchMol
// doubles) and then the kRdfMaxBins u64 counts, 6.1 MiB in all.  Each GPU the library runs
// on holds its own copy until the process ends; host code reaches it only through the
// per-device address from hipGetSymbolAddress (g_scratch below), never by this name.
constexpr size_t kRdfScratchWords = 3 * (size_t)kRdfScratchMol + kRdfMaxBins;
static __device__ __align__(256) double scratch[kRdfScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device ordinal

// com = (sum_s m_s r_s) / (sum_s m_s) over the molecule's sites in order.
__global__ __launch_bounds__(256) void k_mol_com(int nmol, const uint32_t* __restrict__ ms,
                                                 const double* __restrict__ rx, const double* __restrict__ ry,
                                                 const double* __restrict__ rz, const double* __restrict__ mass,
                                                 double* __restrict__ cx, double* __restrict__ cy,
                                                 double* __restrict__ cz) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const uint32_t s0 = ms ? ms[m] : (uint32_t)m, s1 = ms ? ms[m + 1] : (uint32_t)m + 1u;
    double x = 0.0, y = 0.0, z = 0.0, tm = 0.0;
    for (uint32_t s = s0; s < s1; s++) {
        const double w = mass[s];
        x += rx[s] * w;
        y += ry[s] * w;
        z += rz[s] * w;
        tm += w;
    }
    cx[m] = x / tm;
    cy[m] = y / tm;
    cz[m] = z / tm;
}

struct RdfArgs {
    const double* __restrict__ cx;
    const double* __restrict__ cy;
    const double* __restrict__ cz;
    unsigned long long* __restrict__ out;  // [nbins]
    int nmol, ntile, nbins;
    double Lx, Ly, Lz, max_r, max_r2_hi, delta_r;
};

// delta - L round(delta / L) (RDFSensor.cpp:65), round half away from zero.
__device__ __forceinline__ double min_image(double d, double L) { return d - L * round(d / L); }

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* __restrict__ p, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0ull;
}

__global__ __launch_bounds__(kRdfThreads) void k_rdf(RdfArgs a) {
    extern __shared__ unsigned int hist[];  // [nbins]
    __shared__ double tx[kRdfTile], ty[kRdfTile], tz[kRdfTile];
    const int tid = threadIdx.x;
    for (int k = tid; k < a.nbins; k += kRdfThreads) hist[k] = 0u;
    const long long npair_tiles = (long long)a.ntile * a.ntile;
    for (long long t = blockIdx.x; t < npair_tiles; t += gridDim.x) {
        const int I = (int)(t / a.ntile), J = (int)(t % a.ntile);
        if (J < I) continue;  // block-uniform: unordered tile pairs only
        __syncthreads();      // previous tile's reads are done
        const int j0 = J * kRdfTile;
        if (j0 + tid < a.nmol) {
            tx[tid] = a.cx[j0 + tid];
            ty[tid] = a.cy[j0 + tid];
            tz[tid] = a.cz[j0 + tid];
        }
        __syncthreads();
        const int i = I * kRdfTile + tid;
        if (i >= a.nmol) continue;
        const double xi = a.cx[i], yi = a.cy[i], zi = a.cz[i];
        const int jn = min(kRdfTile, a.nmol - j0);
        const int jb = (I == J) ? tid + 1 : 0;  // j > i inside the diagonal tile
        for (int jj = jb; jj < jn; jj++) {
            const double dx = min_image(xi - tx[jj], a.Lx);
            const double dy = min_image(yi - ty[jj], a.Ly);
            const double dz = min_image(zi - tz[jj], a.Lz);
            const double r2 = dx * dx + dy * dy + dz * dz;  // (x^2 + y^2) + z^2, as Array::L2
            if (r2 > a.max_r2_hi) continue;                 // far outside max_r: no sqrt needed
            const double r = sqrt(r2);
            if (r > a.max_r) continue;
            unsigned long long bin = (unsigned long long)(r / a.delta_r);
            if (bin > (unsigned long long)(a.nbins - 1)) bin = (unsigned long long)(a.nbins - 1);
            atomicAdd(&hist[bin], 2u);
        }
    }
    __syncthreads();
    for (int k = tid; k < a.nbins; k += kRdfThreads)
        if (hist[k]) atomicAdd(&a.out[k], (unsigned long long)hist[k]);
}

// One sample into the given buffers: centres [nmol] x 3, counts [nbins] (cleared here by a
// fill kernel, so the sequence holds no memset node when captured).
hipError_t rdf_sample(komd_ctx* c, double max_r, double delta_r, int nbins, double* comx, double* comy, double* comz,
                      unsigned long long* counts) {
    auto& d = c->d;
    const int nmol = (int)c->nmol;
    k_zero_u64<<<(nbins + 255) / 256, 256, 0, c->stream>>>(counts, nbins);
    KOMD_CHECK_HIP(hipGetLastError());
    if (nmol < 2) return hipSuccess;
    const uint32_t* ms = c->multisite ? d.mol_start : nullptr;
    k_mol_com<<<(nmol + 255) / 256, 256, 0, c->stream>>>(nmol, ms, d.rx, d.ry, d.rz, d.mass, comx, comy, comz);
    KOMD_CHECK_HIP(hipGetLastError());
    RdfArgs a{};
    a.cx = comx; a.cy = comy; a.cz = comz; a.out = counts;
    a.nmol = nmol;
    a.ntile = (nmol + kRdfTile - 1) / kRdfTile;
    a.nbins = nbins;
    a.Lx = c->grid.size[0]; a.Ly = c->grid.size[1]; a.Lz = c->grid.size[2];
    a.max_r = max_r;
    a.max_r2_hi = max_r * max_r * (1.0 + 1e-9);
    a.delta_r = delta_r;
    // at most 4096 tile slots per workgroup keeps every u32 bin below 2^31
    const long long slots = (long long)a.ntile * a.ntile;
    const long long grid = std::max<long long>(std::min<long long>(slots, 4096), (slots + 4095) / 4096);
    k_rdf<<<(unsigned)grid, kRdfThreads, (size_t)nbins * sizeof(unsigned int), c->stream>>>(a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_rdf(komd_ctx* c, double max_r, double delta_r, int nbins) {
    auto& d = c->d;
    if (!d.comx) {
        const size_t bytes = (size_t)std::max<int64_t>(c->nmol, 1) * sizeof(double);
        KOMD_CHECK_HIP(hipMalloc((void**)&d.comx, bytes));
        KOMD_CHECK_HIP(hipMalloc((void**)&d.comy, bytes));
        KOMD_CHECK_HIP(hipMalloc((void**)&d.comz, bytes));
        KOMD_CHECK_HIP(hipMalloc((void**)&d.rdf, (size_t)kRdfMaxBins * sizeof(unsigned long long)));
    }
    ScopedTimer tm(c, "rdf");
    return rdf_sample(c, max_r, delta_r, nbins, d.comx, d.comy, d.comz, d.rdf);
}

// The same sample in the static scratch (nmol <= kRdfScratchMol, checked by the caller):
// no allocation, counts left at rdf_scratch_counts(c).
hipError_t launch_rdf_noalloc(komd_ctx* c, double max_r, double delta_r, int nbins) {
    double* base = c->rdf_scratch;
    if (!base) return hipErrorInvalidSymbol;
    return rdf_sample(c, max_r, delta_r, nbins, base, base + kRdfScratchMol, base + 2 * (size_t)kRdfScratchMol,
                      rdf_scratch_counts(c));
}

unsigned long long* rdf_scratch_counts(komd_ctx* c) {
    return c->rdf_scratch ? reinterpret_cast<unsigned long long*>(c->rdf_scratch + 3 * (size_t)kRdfScratchMol)
                          : nullptr;
}

// Device address of `scratch` on the current device, looked up once per device and kept
// (komd_create calls this with the context's device current).
hipError_t rdf_scratch_init(komd_ctx* c) {
    int dev = 0;
    KOMD_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<double*>(p);
    }
    c->rdf_scratch = g_scratch[dev];
    return hipSuccess;
}

}  // namespace komd
