// Internal state and launch wrappers of libkomd_hip.so (gfx950).
//
// Device layout (HBM), N sites, C = nx*ny*nz domain cells:
//   original site order, FP64 SoA : rx ry rz vx vy vz fx fy fz (9 x 8N B), mass, eps, sig
//   idcls[N] u32   dense molecule-id class (LJ same-molecule exclusion, LJ12_6.cpp:19)
//   mol_of[N] u32, mol_start[M+1] u32   molecule grouping for the COM (MoleculeContainer.cpp:55-75)
//   key[N] u32     domain flat cell of every site (x + nx (y + ny z), 0-based)
//   slot[N] u32    arrival slot inside its cell (atomic histogram)
//   count[C], start[C+1] u32, tmp[N], sorted[N] u32   counting sort, stable per cell
//   inv[N] u32     cell-sorted slot of every site (inverse of sorted)
//   spos[N] float4 cell-sorted FP32 position relative to the cell centre, .w = site index bits
//   sprm[N] float2 cell-sorted (sigma, sqrt(eps)) when the sites are not uniform
//   sid[N]  u32    cell-sorted id class when molecules have several sites
//   nl[N][96] float4  full neighbour list of every site, cell-sorted order:
//                  (r_j - r_i) in FP32, .w = j | image code << 26; nlc[N] its length,
//                  nld[N][96] u8 neighbour cell direction (reference-mode ATM only)
//   esite[N] f64   per-site three-body energy share (training targets)
//   vl[N][128] u32 Verlet list of skinned runs: cell-sorted slot of every site within
//                  cutoff + skin at the last rebuild | the offset class of its cell << 25;
//                  vlc[N] its length, spb[N] float4 the FP32 positions of that rebuild
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/komd.h"

namespace komd {

constexpr uint32_t kIdxBits = 26;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;

// error flags raised by kernels (bitmask in dev->flags)
enum : uint32_t { FLAG_NL_OVERFLOW = 1u, FLAG_NONFINITE = 2u, FLAG_CELL_OVERFLOW = 4u, FLAG_VL_OVERFLOW = 8u };

// Slots of the device scalars (DevState::scal, 32 doubles).
enum : int {
    kScalMv2 = 0,    // sum m v.v of the sites inside the box (last temperature pass)
    kScalCount = 1,  // sites inside the box
    kScalBeta = 2,   // rescale factor applied at the next kick-drift of a thermostatted run
    kScalTemp = 3,   // last measured temperature
    kScalKin = 4,    // kinetic energy
    kScalT0 = 5,     // thermostat target of the running komd_run (read by the captured step)
    kScalHybU = 12,  // hybrid pass: exact-ATM energy share
    kScalHybW = 16,  // hybrid pass: exact-ATM virial share [16, 22)
    kScalOut = 24    // results of one reduction [24, 32)
};

struct Grid {
    int nx, ny, nz;         // domain cells
    int ncell;              // nx*ny*nz
    double low[3], high[3], size[3], cs;
};

struct Partials {
    double* buf = nullptr;  // per-block FP64 partial sums, [nslots][nblocks]
    int nblocks = 0;
};

struct Surrogate;  // k_surrogate.hip
struct SkinArgs;   // k_cell.hip
struct PairTable;  // k_table.hip

struct DevState {
    // sites
    double *rx = nullptr, *ry = nullptr, *rz = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double *mass = nullptr, *eps = nullptr, *sig = nullptr;
    uint32_t* idcls = nullptr;
    uint32_t* mol_start = nullptr;
    // binning
    uint32_t *key = nullptr, *slot = nullptr, *count = nullptr, *start = nullptr, *tmp = nullptr,
             *sorted = nullptr, *inv = nullptr, *scanp = nullptr;
    float4* spos = nullptr;
    float2* sprm = nullptr;
    uint32_t* sid = nullptr;
    // full neighbour lists (three-body and surrogate passes)
    float4* nl = nullptr;
    uint32_t* nlc = nullptr;
    uint16_t* nld = nullptr;      // [N][kFullCap] neighbour cell direction | rank in that cell << 5
    // Verlet lists of skinned runs (komd_set_skin): every site within cutoff + skin at the
    // last rebuild, entry = cell-sorted slot of the neighbour | offset class of its cell << 25
    uint32_t* vl = nullptr;       // [N][kSkinCap]
    uint32_t* vlc = nullptr;      // [N] lengths
    uint16_t* vld = nullptr;      // [N][kSkinCap] cell direction | rank << 5 (learned and hybrid kinds)
    float4* spb = nullptr;        // [N] cell-sorted FP32 positions of the last rebuild
    float4* skin_off = nullptr;   // [125] FP32 offset (and image code) of each list offset class
    SkinArgs* skin_args = nullptr;  // the Verlet-list force pass's buffer table (device copy)
    uint32_t* skin_req = nullptr; // [0], [1] rebuild words, one per step parity; [2] rebuilds so far
    // bonded
    uint32_t *bonds = nullptr, *angles = nullptr;
    uint32_t *bond_mol = nullptr, *angle_mol = nullptr;  // CSR of bonds / angles per molecule
    // per site: the bond and angle terms it takes part in (CSR: term << 3 | angle << 2 | role)
    uint32_t *site_tstart = nullptr, *site_terms = nullptr;
    // reductions and flags
    double* red = nullptr;        // partials (see k_container.hip)
    double* vpart = nullptr;      // [6][blocks] virial / kinetic tensor partials (energy passes)
    double* scal = nullptr;       // device scalars [32]: [0] mv2, [1] count, [2] beta, ...
    uint32_t* flags = nullptr;
    unsigned long long* counters = nullptr;  // unused legacy slot
    unsigned long long* cnt_pair = nullptr;  // [cell blocks] in-range pairs per workgroup slot
    unsigned long long* cnt_trip = nullptr;  // [ATM blocks] triplet visits per workgroup slot
    double* esite = nullptr;      // per-site energy scratch [N]
    float4* rows = nullptr;       // per-cell partial force rows [C][27][kRows] (surrogate force pass)
    // radial distribution sensor (allocated on first use)
    double *comx = nullptr, *comy = nullptr, *comz = nullptr;  // [M] molecule centres of mass
    unsigned long long* rdf = nullptr;                         // [kRdfMaxBins] ordered-pair counts
};

constexpr int kRdfMaxBins = 8192;  // u32 LDS histogram of 32 KiB per workgroup
constexpr int kRdfScratchMol = 262144;  // molecules komd_rdf_noalloc's static scratch holds

struct Timer {
    double total_ms = 0;
    int64_t launches = 0;
};

}  // namespace komd

struct komd_ctx {
    komd::Grid grid{};
    komd_params prm{};
    int64_t n = 0;
    int64_t nmol = 0;
    bool uniform = true;       // all sites share sigma/epsilon
    bool multisite = false;    // some molecule has more than one site
    float sig_u = 1.f, eps_u = 1.f;
    double sig0 = 1.0, eps0 = 1.0;  // the first site's sigma and epsilon (ATM2B's one component)
    int64_t nbonds = 0, nangles = 0;
    std::vector<uint32_t> mol_start_h;  // host copy of the molecule ranges
    bool bonded_by_mol = false;         // every bond/angle inside one molecule: per-molecule kernel
    bool ids_runs = true;               // every molecule id forms one run of consecutive sites
    double kb = 0, r0 = 0, ka = 0, theta0 = 0;
    hipStream_t stream = nullptr;
    komd::DevState d;
    std::string err;
    bool have_sites = false;
    bool nl_valid = false;
    int red_blocks = 0;
    int cell_split = 1;        // workgroups sharing one home cell in the pair pass (small grids)
    int xcd_slots = 0;         // pair pass: workgroups resident at once per XCD (0: not queried yet)
    // timing
    bool timing = false;
    std::unordered_map<std::string, komd::Timer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // graph of one step (skinned runs: graph_exec for even steps, graph_exec_odd for odd ones)
    hipGraphExec_t graph_exec = nullptr;
    hipGraphExec_t graph_exec_odd = nullptr;
    // kStepsPerGraph steps in one graph (skinned runs: starting at parity 0); the gap between
    // two graph launches is several times that between two kernels of one graph
    hipGraphExec_t graph_exec_multi = nullptr;
    int graph_key = -1;
    // Verlet lists (komd_set_skin): skin > 0 enables them for the runs that can use them;
    // skin_par is the parity of the next queued step
    double skin = 0.0;
    int skin_par = 0;
    // surrogate
    komd::Surrogate* sur = nullptr;
    // pair force of komd_run / komd_compute_forces / komd_measure_observables
    int pair_kind = KOMD_PAIR_LJ12_6;
    komd::PairTable* tab = nullptr;  // KOMD_PAIR_TABLE (komd_set_pair_table)
    // pinned host slots the flag snapshots of komd_run land in, and their events
    uint32_t* flag_host = nullptr;
    hipEvent_t flag_ev[2] = {nullptr, nullptr};
    // this device's static RDF scratch (komd_rdf_noalloc); null when its lookup failed
    double* rdf_scratch = nullptr;
    std::string rdf_scratch_err;
    // Integrator::Step1 fused into the last force kernel of an NVE step (no kinetic sums
    // are needed there): 0 none (k_kick runs), else the kernel that applies it
    int kick_in = 0;
    bool lj_in_desc = false;  // set by the NVE step's force pass: the descriptor pass also sums LJ
    int skin_step_par = -1;   // parity of the Verlet-list step being queued (-1: none)
    int hyb_axis = 0;
    double hyb_center = 0, hyb_half = 0, hyb_blend = 0;
    // exact-ATM judge of surrogate runs
    int64_t judge_every = 0;
    double judge_max = 0;
    int64_t judge_checks = 0;
    double judge_last = 0, judge_worst = 0;
    int judge_fell = 0;
};

namespace komd {

// kernel timing helpers (komd_api.hip)
void timer_begin(komd_ctx* c, const char* name, hipEvent_t* a, hipEvent_t* b);
void timer_end(komd_ctx* c, const char* name, hipEvent_t a, hipEvent_t b);

struct ScopedTimer {
    komd_ctx* c;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    ScopedTimer(komd_ctx* c_, const char* n_) : c(c_), name(n_) {
        if (c->timing) timer_begin(c, name, &a, &b);
    }
    ~ScopedTimer() {
        if (c->timing) timer_end(c, name, a, b);
    }
};

// ---- host <-> device transfers on the context stream ---------------------
// The context stream is non-blocking, so nothing orders null-stream memsets and copies
// with the kernels queued on it: every transfer into device memory goes through the
// context stream.  h2d returns once the copy is complete (the source may be reused).
inline hipError_t h2d(komd_ctx* c, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}
inline hipError_t dzero(komd_ctx* c, void* dst, size_t bytes) { return hipMemsetAsync(dst, 0, bytes, c->stream); }

// ---- launch wrappers --------------------------------------------------
// k_container.hip
hipError_t launch_com_wrap_key(komd_ctx* c, bool do_drift, bool reset_f = true);
// scan + scatter + stable sort + FP32 repack; req (Verlet-list runs): the step's rebuild
// word, the kernels return at once when it is zero; spb: a copy of the FP32 positions
hipError_t launch_bin(komd_ctx* c, const uint32_t* req = nullptr, float4* spb = nullptr);
// Verlet-list runs: Step0 + (rebuild steps) container update, else the FP32 position
// refresh; par is the step parity whose rebuild word is read
hipError_t launch_skin_drift(komd_ctx* c, int par, bool zero_f = false);
hipError_t launch_skin_arm(komd_ctx* c, int par);  // rebuild word of parity par = 1
hipError_t launch_kick_drift(komd_ctx* c);      // integrate0
hipError_t launch_kick(komd_ctx* c, bool with_temp);  // integrate1 (+ kinetic partials)
// v += f (dt/2)/m in the reference's FP64 operation order (Integrator.cpp:34-37).
// Contraction into an FMA is switched off for the operators written in this body
// (the pragma does not reach into other functions, e.g. __dmul_rn / __dadd_rn, whose
// bodies are plain * and + compiled with contraction on).
__device__ __forceinline__ void kick_site(double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
                                          const double* __restrict__ mass, uint32_t i, double dt_half, double fx,
                                          double fy, double fz) {
#pragma clang fp contract(off)
    const double k = dt_half / mass[i];
    vx[i] = vx[i] + fx * k;
    vy[i] = vy[i] + fy * k;
    vz[i] = vz[i] + fz * k;
}
hipError_t launch_temperature(komd_ctx* c);     // kinetic partials -> scal[0..1]
hipError_t launch_rescale(komd_ctx* c, double t0);    // beta from scal, v *= beta
hipError_t launch_beta_only(komd_ctx* c);             // beta = sqrt(scal[kScalT0]/T) into scal[kScalBeta]
// scal[kScalBeta] = 1 and scal[kScalT0] = t0 before a thermostatted komd_run (outside any capture)
hipError_t launch_run_scalars(komd_ctx* c, double t0);
// copies the error flags to a pinned host slot (flag checks that do not drain the stream)
hipError_t launch_flag_snapshot(komd_ctx* c, uint32_t* host_slot);
hipError_t launch_reduce_parts(komd_ctx* c, const double* part, int nparts, int nslots, double* out_dev);
hipError_t launch_kin_tensor(komd_ctx* c);      // sum m v_a v_b partials -> d.vpart ([6][kKinBlocks])
constexpr int kKinBlocks = 256;
// k_sensors.hip: one RDF sample of the molecules' COM distances into d.rdf (nbins counts)
hipError_t launch_rdf(komd_ctx* c, double max_r, double delta_r, int nbins);
// the same sample in the device's static scratch (no allocation); counts at rdf_scratch_counts
hipError_t launch_rdf_noalloc(komd_ctx* c, double max_r, double delta_r, int nbins);
unsigned long long* rdf_scratch_counts(komd_ctx* c);
// looks up the static scratch's address on the current device (once per device)
hipError_t rdf_scratch_init(komd_ctx* c);
// k_cell.hip.  launch_cell: LJ (+ clamp) on cell tiles; nl: 0 no neighbour list, 1 write
// the full lists, 2 lists + cell directions.  launch_atm: per-site ATM from the lists;
// atm: 1 physical, 2 reference; lam/dlam: hybrid weights (nullptr otherwise); esite:
// per-site energy share out (optional).
enum { KICK_NONE = 0, KICK_IN_CELL = 1, KICK_IN_ATM = 2, KICK_IN_APPLY = 3, KICK_IN_TABLE = 4 };
// scale multiplies the LJ force and energy (1: LJ 12-6; the ATM2B factor for KOMD_PAIR_ATM2B)
hipError_t launch_cell(komd_ctx* c, bool lj, int nl, bool energy, bool limit, double scale = 1.0);
// lj: the kernel also adds the LJ force of every listed pair (uniform single-site
// systems without the force clamp; the pair pass then skips LJ)
// skin_par >= 0: a Verlet-list step of that parity (launch_skin_list built or kept the lists)
hipError_t launch_atm(komd_ctx* c, int atm, bool energy, const float* lam, const float* dlam, double* esite,
                      bool lj = false, int skin_par = -1);
// dirs: also the cell direction and rank of every entry (d.vld), for launch_skin_refresh
hipError_t launch_skin_list(komd_ctx* c, int par, bool dirs);
// LJ from the Verlet lists + Step1 + rebuild decision (runs without a three-body term)
hipError_t launch_skin_lj(komd_ctx* c, int par, bool add);
// the in-range neighbour-list rows (d.nl, d.nlc, d.nld) of every site from its Verlet list
// at the current positions, and the LJ 12-6 force written as the site's force
hipError_t launch_skin_refresh(komd_ctx* c);
hipError_t launch_skin_offsets(komd_ctx* c);  // d.skin_off from the grid
hipError_t upload_skin_args(komd_ctx* c);     // d.skin_args (before a skinned run)
size_t skin_args_bytes();
int cell_blocks(komd_ctx* c);
int atm_blocks(komd_ctx* c);
// k_table.hip: tabulated pair force over current neighbour lists; with energy the pair
// energy and virial partials are left in d.red / d.vpart, table_blocks of them
hipError_t launch_table(komd_ctx* c, bool energy);
int table_blocks(komd_ctx* c);
hipError_t table_upload(komd_ctx* c, int64_t nodes, const double* x, const double* y, double y_low, double y_high,
                        const double* u, double u_low, double u_high);
void table_free(komd_ctx* c);
// k_bonded.hip
// intra: also count the in-range same-molecule pairs into the Verlet-list pair counters
// (per-molecule grouping only; Verlet lists exclude those pairs)
// intra: also count the in-range same-molecule pairs into the Verlet-list pair counters
// (Verlet lists exclude those pairs)
hipError_t launch_bonded(komd_ctx* c, bool energy, bool intra = false);
// the per-site bonded kernel (terms: add the bond and angle forces; intra: the count)
hipError_t launch_bonded_sites(komd_ctx* c, bool terms, bool intra);
hipError_t launch_limit(komd_ctx* c);
// k_surrogate.hip.  Both expect current neighbour lists.  With energy, the three-body
// energy partials are left in d.red, `three_body_energy_blocks` of them.
hipError_t launch_surrogate_forces(komd_ctx* c, bool energy, int hybrid);
hipError_t hybrid_forces(komd_ctx* c, bool energy);
int surrogate_energy_blocks(komd_ctx* c);
// With energy, the surrogate pass leaves its virial partials in d.vpart as kSym slots of
// surrogate_virial_blocks partials (the last one holds the hybrid's exact-ATM share).
int surrogate_virial_blocks(komd_ctx* c);
hipError_t surrogate_prepare(komd_ctx* c, bool hybrid);  // allocates the pass buffers (not capturable)
void surrogate_free(komd_ctx* c);
// komd_set_sites hook: when the site count changes, frees the surrogate's per-site buffers
// and the training snapshots (both sized by the old count); parameters are kept
void surrogate_sites_changed(komd_ctx* c);

}  // namespace komd

#define KOMD_CHECK_HIP(expr)                                                          \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) return _e;                                             \
    } while (0)
