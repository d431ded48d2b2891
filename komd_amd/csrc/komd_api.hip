// C ABI of libkomd_hip.so: context, state transfer, step orchestration.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <vector>

#include "komd_cell.hpp"
#include "komd_reduce.hpp"

namespace komd {
hipError_t launch_com_wrap_key_beta(komd_ctx* c, const double* beta_dev);

void timer_begin(komd_ctx* c, const char* name, hipEvent_t* a, hipEvent_t* b) {
    (void)name;
    auto grab = [&]() {
        hipEvent_t e;
        if (!c->event_pool.empty()) {
            e = c->event_pool.back();
            c->event_pool.pop_back();
        } else {
            hipEventCreate(&e);
        }
        return e;
    };
    *a = grab();
    *b = grab();
    hipEventRecord(*a, c->stream);
}

void timer_end(komd_ctx* c, const char* name, hipEvent_t a, hipEvent_t b) {
    hipEventRecord(b, c->stream);
    c->pending.push_back({std::string(name), {a, b}});
}
}  // namespace komd

using namespace komd;

namespace {

komd_status fail(komd_ctx* c, komd_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

komd_status hip_fail(komd_ctx* c, hipError_t e, const char* where) {
    return fail(c, KOMD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCALL(c, expr, where)                         \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return hip_fail(c, _e, where); \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

void free_sites(komd_ctx* c) {
    auto& d = c->d;
    void* ptrs[] = {d.rx, d.ry, d.rz, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz, d.mass, d.eps, d.sig, d.idcls, d.mol_start,
                    d.key, d.slot, d.count, d.start, d.tmp, d.sorted, d.inv, d.scanp, d.spos, d.sprm, d.sid, d.nl, d.nlc, d.nld, d.rows, d.red, d.cnt_pair, d.cnt_trip,
                    d.esite, d.vpart, d.comx, d.comy, d.comz, d.rdf, d.vl, d.vlc, d.vld, d.spb, d.skin_off, d.skin_args, d.skin_req};
    for (void* p : ptrs)
        if (p) hipFree(p);
    DevState keep;
    keep.scal = d.scal;
    keep.flags = d.flags;
    keep.counters = d.counters;
    keep.bonds = d.bonds;
    keep.bond_mol = d.bond_mol;
    keep.angle_mol = d.angle_mol;
    keep.angles = d.angles;
    keep.site_tstart = d.site_tstart;
    keep.site_terms = d.site_terms;
    d = keep;
    c->have_sites = false;
    c->nl_valid = false;
}

// Frees the bonded topology (terms, per-molecule and per-site tables): no bonded terms.
void drop_topology(komd_ctx* c) {
    auto& d = c->d;
    void* old[] = {d.bonds, d.angles, d.bond_mol, d.angle_mol, d.site_tstart, d.site_terms};
    for (void* p : old)
        if (p) hipFree(p);
    d.bonds = d.angles = d.bond_mol = d.angle_mol = d.site_tstart = d.site_terms = nullptr;
    c->nbonds = c->nangles = 0;
    c->bonded_by_mol = false;
}

void drop_graph(komd_ctx* c) {
    if (c->graph_exec) hipGraphExecDestroy(c->graph_exec);
    if (c->graph_exec_odd) hipGraphExecDestroy(c->graph_exec_odd);
    if (c->graph_exec_multi) hipGraphExecDestroy(c->graph_exec_multi);
    c->graph_exec = nullptr;
    c->graph_exec_odd = nullptr;
    c->graph_exec_multi = nullptr;
    c->graph_key = -1;
}

// Steps per multi-step graph of komd_run: even (a skinned launch ends at the step parity it
// started with) and a divisor of kFlagEvery.
constexpr int kStepsPerGraph = 8;

// Captures what fn queues on the context stream into an executable graph.
template <class F>
hipError_t capture_graph(komd_ctx* c, hipGraphExec_t* out, F fn) {
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t e1 = fn();
    const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (e1 == hipSuccess) e1 = e2;
    if (e1 == hipSuccess) e1 = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    if (graph) hipGraphDestroy(graph);
    return e1;
}

// Collect finished event pairs into the per-kernel totals.
void drain_timers(komd_ctx* c) {
    for (auto& pe : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            auto& t = c->timers[pe.first];
            t.total_ms += ms;
            t.launches += 1;
        }
        c->event_pool.push_back(pe.second.first);
        c->event_pool.push_back(pe.second.second);
    }
    c->pending.clear();
}

// Synchronises and turns kernel-raised flags into a status.
komd_status sync_check(komd_ctx* c, const char* where) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, where);
    if (c->timing) drain_timers(c);
    uint32_t flags = 0;
    e = hipMemcpy(&flags, c->d.flags, sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(c, e, where);
    if (flags) {
        dzero(c, c->d.flags, sizeof(uint32_t));
        if (flags & FLAG_NL_OVERFLOW)
            return fail(c, KOMD_ERR_OVERFLOW,
                        std::string(where) + ": a site has more than " + std::to_string(kFullCap) +
                            " in-range neighbours (the neighbour-list capacity)");
        if (flags & FLAG_VL_OVERFLOW)
            return fail(c, KOMD_ERR_OVERFLOW,
                        std::string(where) + ": a site has more than " + std::to_string(kSkinCap) +
                            " neighbours within cutoff + skin (the Verlet-list capacity; use a smaller skin)");
        if (flags & FLAG_CELL_OVERFLOW)
            return fail(c, KOMD_ERR_OVERFLOW,
                        std::string(where) + ": surrogate force pass: a cell neighbourhood holds more than " +
                            std::to_string(kSurM) + " sites or a cell more than " + std::to_string(kRows) +
                            " (use a smaller cell_size)");
        if (flags & FLAG_NONFINITE)
            return fail(c, KOMD_ERR_NONFINITE, std::string(where) + ": non-finite position or velocity");
    }
    return KOMD_OK;
}

// Sums the per-workgroup count slots: [0] unique in-range pairs (the pair pass's slots plus
// those of the Verlet-list ATM pass, cnt_trip[atm_blocks ..]), [1] triplet visits.
komd_status read_counters(komd_ctx* c, int64_t out[2], bool reset) {
    HIPCALL(c, hipStreamSynchronize(c->stream), "komd_counters");
    const int na = atm_blocks(c);
    std::vector<unsigned long long> hp((size_t)cell_blocks(c)), ht((size_t)2 * na);
    HIPCALL(c, hipMemcpy(hp.data(), c->d.cnt_pair, hp.size() * 8, hipMemcpyDeviceToHost), "komd_counters");
    HIPCALL(c, hipMemcpy(ht.data(), c->d.cnt_trip, ht.size() * 8, hipMemcpyDeviceToHost), "komd_counters");
    unsigned long long pairs = 0, trip = 0;
    for (auto x : hp) pairs += x;
    for (int k = 0; k < na; k++) {
        trip += ht[(size_t)k];
        pairs += ht[(size_t)(na + k)];
    }
    out[0] = (int64_t)pairs;
    out[1] = (int64_t)trip;
    if (reset) {
        HIPCALL(c, dzero(c, c->d.cnt_pair, hp.size() * 8), "komd_counters");
        HIPCALL(c, dzero(c, c->d.cnt_trip, ht.size() * 8), "komd_counters");
    }
    return KOMD_OK;
}

komd_status need_sites(komd_ctx* c) {
    if (!c) return KOMD_ERR_ARG;
    if (!c->have_sites) return fail(c, KOMD_ERR_STATE, "no sites: call komd_set_sites first");
    return KOMD_OK;
}

bool is_atm(int kind) { return kind == KOMD_3B_ATM_EXACT || kind == KOMD_3B_ATM_REF; }

// Neighbour-list buffers, allocated on first use by a three-body or surrogate pass.
hipError_t ensure_nl(komd_ctx* c, bool dirs) {
    auto& d = c->d;
    if (!d.nl) KOMD_CHECK_HIP(dalloc(&d.nl, (size_t)c->n * kFullCap));
    if (!d.nlc) KOMD_CHECK_HIP(dalloc(&d.nlc, (size_t)c->n));
    if (dirs && !d.nld) KOMD_CHECK_HIP(dalloc(&d.nld, (size_t)c->n * kFullCap));
    if (dirs && !d.rows) KOMD_CHECK_HIP(dalloc(&d.rows, (size_t)c->grid.ncell * 27 * kRows));
    return hipSuccess;
}

// Verlet-list buffers, allocated on the first skinned run (dirs: the entries' cell
// directions and ranks, for the learned and hybrid kinds).
hipError_t ensure_skin(komd_ctx* c, bool dirs) {
    auto& d = c->d;
    if (!d.vl) KOMD_CHECK_HIP(dalloc(&d.vl, (size_t)c->n * kSkinCap));
    if (dirs && !d.vld) KOMD_CHECK_HIP(dalloc(&d.vld, (size_t)c->n * kSkinCap));
    if (!d.vlc) KOMD_CHECK_HIP(dalloc(&d.vlc, (size_t)c->n));
    if (!d.spb) KOMD_CHECK_HIP(dalloc(&d.spb, (size_t)c->n));
    if (!d.skin_off) KOMD_CHECK_HIP(dalloc(&d.skin_off, 125));
    if (!d.skin_args) KOMD_CHECK_HIP(hipMalloc((void**)&d.skin_args, skin_args_bytes()));
    if (!d.skin_req) {
        KOMD_CHECK_HIP(dalloc(&d.skin_req, 4));
        KOMD_CHECK_HIP(dzero(c, d.skin_req, 4 * sizeof(uint32_t)));
    }
    return hipSuccess;
}
}  // namespace

namespace komd {
komd_status api_fail(komd_ctx* c, komd_status s, const std::string& msg) { return fail(c, s, msg); }
komd_status api_sync(komd_ctx* c, const char* where) { return sync_check(c, where); }
// COM, wrap, force reset, binning, then the full neighbour lists (no pair force)
hipError_t api_prepare_lists(komd_ctx* c) {
    KOMD_CHECK_HIP(launch_com_wrap_key(c, false));
    KOMD_CHECK_HIP(launch_bin(c));
    KOMD_CHECK_HIP(ensure_nl(c, true));
    return launch_cell(c, false, 2, false, false);
}
}  // namespace komd

namespace {

int atm_code(int kind) { return kind == KOMD_3B_ATM_EXACT ? 1 : (kind == KOMD_3B_ATM_REF ? 2 : 0); }
// neighbour-list mode of a three-body kind: 0 none, 1 lists, 2 lists + cell directions
// (the reference-mode ATM filter and the surrogate's per-cell force rows read them)
int nl_mode(int kind) { return kind == KOMD_3B_ATM_EXACT ? 1 : (kind ? 2 : 0); }
bool table_pair(const komd_ctx* c) { return c->pair_kind == KOMD_PAIR_TABLE; }
// ATM2B's pair-force factor 1 - 2 nu rho / (3 eps sig^6) (potentials/ATM2B.cpp:9-20): rho is
// the molecule number density, eps and sigma those of the first site (one component)
double atm2b_factor(const komd_ctx* c) {
    const Grid& g = c->grid;
    const double rho = (double)c->nmol / (g.size[0] * g.size[1] * g.size[2]);
    return 1.0 - (2.0 * c->prm.nu * rho) / (3.0 * c->eps0 * std::pow(c->sig0, 6));
}
// the LJ pass's force scale of a pair kind
double pair_scale(const komd_ctx* c, int kind) { return kind == KOMD_PAIR_ATM2B ? atm2b_factor(c) : 1.0; }
// lists of a step: the three-body kind's, and at least plain lists for the tabulated pair force
int lists_of(const komd_ctx* c, int kind3b) { return table_pair(c) ? std::max(1, nl_mode(kind3b)) : nl_mode(kind3b); }
// The kernel that applies the second half-kick of an NVE step: the last force kernel that
// completes a site's force (KICK_NONE: a separate k_kick after the forces).
int kick_kernel(const komd_ctx* c, int kind3b) {
    if (is_atm(kind3b)) return KICK_IN_ATM;
    if (kind3b == KOMD_3B_SURROGATE || kind3b == KOMD_3B_HYBRID) return KICK_IN_APPLY;
    if (!table_pair(c)) return KICK_IN_CELL;
    return c->prm.limit_factor > 0 ? KICK_NONE : KICK_IN_TABLE;  // the clamp follows the table pass
}

// Three-body forces on their own (ForceFunctor3B::operator(), ForceFunctor3B.cpp:10-14).
hipError_t forces_3b(komd_ctx* c, int kind, bool energy) {
    if (is_atm(kind)) {
        KOMD_CHECK_HIP(ensure_nl(c, kind == KOMD_3B_ATM_REF));
        KOMD_CHECK_HIP(launch_cell(c, false, atm_code(kind), false, false));
        return launch_atm(c, atm_code(kind), energy, nullptr, nullptr, nullptr);
    }
    if (kind == KOMD_3B_SURROGATE || kind == KOMD_3B_HYBRID) {
        KOMD_CHECK_HIP(ensure_nl(c, true));
        KOMD_CHECK_HIP(launch_cell(c, false, 2, false, false));
        if (kind == KOMD_3B_SURROGATE) return launch_surrogate_forces(c, energy, 0);
        return hybrid_forces(c, energy);
    }
    return hipErrorInvalidValue;
}

// The force part of a step in the reference order (Simulation.cpp:55-56, main.cpp:57-66):
// LJ, bonded (FENE's slot), Limit, then three-body.  The bonded kernel runs first so
// that the cell-tile pass sees the complete two-body force when it clamps it; LJ,
// clamp and ATM share that one pass.
hipError_t enqueue_forces(komd_ctx* c, int kind3b, bool bonded) {
    if (bonded) KOMD_CHECK_HIP(launch_bonded(c, false));
    const int atm = atm_code(kind3b);
    const int nl = lists_of(c, kind3b);
    if (nl) KOMD_CHECK_HIP(ensure_nl(c, nl == 2));
    // uniform single-site LJ without the clamp: the three-body kernel sums the LJ force
    // from the site's compacted list (one pair per lane) instead of the pair pass's
    // candidate loop, where a lane's rare hits left most lanes idle
    // (the surrogate's descriptor pass likewise walks every site's list)
    const bool lj_mv = c->pair_kind == KOMD_PAIR_LJ12_6 && c->uniform && !c->multisite && !(c->prm.limit_factor > 0);
    const bool lj_atm = atm && lj_mv, lj_sur = kind3b == KOMD_3B_SURROGATE && lj_mv;
    if (table_pair(c)) {
        // tabulated pair force (the IBI functor in LJ's registry slot, IBI.cpp:51-55) from
        // the lists, then the clamp over the complete two-body force
        KOMD_CHECK_HIP(launch_cell(c, false, nl, false, false));
        KOMD_CHECK_HIP(launch_table(c, false));
        KOMD_CHECK_HIP(launch_limit(c));
    } else {
        KOMD_CHECK_HIP(launch_cell(c, !(lj_atm || lj_sur), nl, false, true, pair_scale(c, c->pair_kind)));
    }
    if (atm) return launch_atm(c, atm, false, nullptr, nullptr, nullptr, lj_atm);
    if (kind3b == KOMD_3B_SURROGATE) {
        c->lj_in_desc = lj_sur;
        const hipError_t e = launch_surrogate_forces(c, false, 0);
        c->lj_in_desc = false;
        return e;
    }
    if (kind3b == KOMD_3B_HYBRID) return hybrid_forces(c, false);
    return hipSuccess;
}

// Runs that use Verlet lists (komd_set_skin): NVE with LJ 12-6 of uniform sites without the
// clamp, and either the exact ATM, the surrogate or the hybrid slab (single-site, no bonded
// terms), or no three-body term: single-site sites, or multi-site molecules binned by site
// whose ids each form one run of sites (the per-site bonded kernel counts the same-molecule
// pairs the lists leave out, from the molecule ranges).
bool skin_run(const komd_ctx* c, int kind3b, bool bonded, bool thermo) {
    const bool three = kind3b == KOMD_3B_ATM_EXACT || kind3b == KOMD_3B_SURROGATE || kind3b == KOMD_3B_HYBRID;
    const bool two = kind3b == 0 && (!c->multisite || (c->prm.bin_mode == KOMD_BIN_SITE && c->ids_runs));
    const bool sites = three ? (!bonded && !c->multisite) : two;
    return c->skin > 0 && c->n <= (int64_t)kSkinIdxMask && sites && !thermo &&
           c->pair_kind == KOMD_PAIR_LJ12_6 && c->uniform && !(c->prm.limit_factor > 0);
}

// One step with Verlet lists, parity par (Simulation.cpp:49-74 with the pair list of
// LinkedCells::updatePairList rebuilt only when needed, PairList.h:22-24): Step0 and the
// refreshed FP32 positions; on a rebuild step the container update, binning and the list
// pass (the other kernels return at once).  Exact ATM: LJ + ATM from the lists with Step1
// fused.  Learned and hybrid kinds: the in-range rows and LJ from the lists, then their own
// three-body passes, the last of which applies Step1 and decides the next rebuild.  No
// three-body term: the bonded terms (FENE's slot, Simulation.cpp:55-56), then LJ from the
// lists added to them with Step1 fused.
hipError_t enqueue_skin_step(komd_ctx* c, int kind3b, int par, bool bonded) {
    const bool two = kind3b == 0;
    // multi-site molecules: a bonded kernel runs (if only to count the in-range same-molecule
    // pairs) and the forces add onto the zero the drift writes
    const bool sites = two && (bonded || c->multisite);
    KOMD_CHECK_HIP(launch_skin_drift(c, par, sites));
    KOMD_CHECK_HIP(launch_bin(c, c->d.skin_req + par, c->d.spb));
    const bool exact = kind3b == KOMD_3B_ATM_EXACT;
    KOMD_CHECK_HIP(launch_skin_list(c, par, !exact && !two));
    if (two) {
        if (bonded) KOMD_CHECK_HIP(launch_bonded(c, false, c->multisite));
        else if (sites) KOMD_CHECK_HIP(launch_bonded_sites(c, false, true));  // the count alone
        return launch_skin_lj(c, par, sites);
    }
    hipError_t e;
    if (exact) {
        c->kick_in = KICK_IN_ATM;
        e = launch_atm(c, 1, false, nullptr, nullptr, nullptr, true, par);
    } else {
        KOMD_CHECK_HIP(launch_skin_refresh(c));
        c->kick_in = KICK_IN_APPLY;
        c->skin_step_par = par;
        e = kind3b == KOMD_3B_SURROGATE ? launch_surrogate_forces(c, false, 0) : hybrid_forces(c, false);
    }
    c->kick_in = KICK_NONE;
    c->skin_step_par = -1;
    return e;
}

// One reference-order step (Simulation.cpp:49-74) queued on the stream.
hipError_t enqueue_step(komd_ctx* c, int kind3b, bool bonded, bool thermo) {
    if (thermo)
        KOMD_CHECK_HIP(launch_com_wrap_key_beta(c, c->d.scal + 2));
    else
        KOMD_CHECK_HIP(launch_com_wrap_key(c, true));
    KOMD_CHECK_HIP(launch_bin(c));
    // without the thermostat no kinetic sums are needed: the last force kernel of the
    // step applies the second half-kick to the sites whose force it completes
    const int kick = thermo ? KICK_NONE : kick_kernel(c, kind3b);
    c->kick_in = kick;
    const hipError_t e = enqueue_forces(c, kind3b, bonded);
    c->kick_in = KICK_NONE;
    KOMD_CHECK_HIP(e);
    if (kick == KICK_NONE) KOMD_CHECK_HIP(launch_kick(c, thermo));
    if (thermo) KOMD_CHECK_HIP(launch_beta_only(c));  // beta for the next step's rescale
    return hipSuccess;
}

}  // namespace

extern "C" {

komd_status komd_device_name(char* buf, int32_t len) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return KOMD_ERR_HIP;
    snprintf(buf, (size_t)len, "%s%s%s, %d CUs, %.0f GB", p.name, p.name[0] ? " " : "", p.gcnArchName,
             p.multiProcessorCount, (double)p.totalGlobalMem / 1e9);
    return KOMD_OK;
}

komd_status komd_create(const komd_box* box, const komd_params* params, komd_ctx** out) {
    if (!box || !params || !out) return KOMD_ERR_ARG;
    *out = nullptr;
    komd_ctx* c = new komd_ctx();
    c->prm = *params;
    Grid& g = c->grid;
    for (int k = 0; k < 3; k++) {
        g.low[k] = box->low[k];
        g.high[k] = box->high[k];
        g.size[k] = box->high[k] - box->low[k];
    }
    g.cs = box->cell_size;
    if (!(params->cutoff > 0) || !(box->cell_size >= params->cutoff)) {
        // FileInput.cpp:117-120
        c->err = "cell size must be larger or equal to cutoff radius";
        *out = c;
        return KOMD_ERR_ARG;
    }
    // LinkedCells.cpp:19: num_cells_domain = floor(dom_size / cell_size)
    g.nx = (int)floor(g.size[0] / g.cs);
    g.ny = (int)floor(g.size[1] / g.cs);
    g.nz = (int)floor(g.size[2] / g.cs);
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) {
        c->err = "need at least 3 domain cells per dimension";
        *out = c;
        return KOMD_ERR_ARG;
    }
    g.ncell = g.nx * g.ny * g.nz;
    if (params->max_neighbors != 0 && params->max_neighbors != kFullCap) {
        c->err = "max_neighbors: the neighbour-list capacity is fixed at " + std::to_string(kFullCap) +
                 " sites (pass 0 or " + std::to_string(kFullCap) + ")";
        *out = c;
        return KOMD_ERR_ARG;
    }
    c->prm.max_neighbors = kFullCap;
    if (c->prm.temp_factor <= 0) c->prm.temp_factor = 1.0;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = dalloc(&c->d.scal, 32);
    if (e == hipSuccess) e = dalloc(&c->d.flags, 1);
    if (e == hipSuccess) e = dalloc(&c->d.counters, 2);
    if (e == hipSuccess) e = dzero(c, c->d.flags, sizeof(uint32_t));
    if (e == hipSuccess) e = dzero(c, c->d.counters, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) e = dzero(c, c->d.scal, 32 * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->flag_host, 2 * sizeof(uint32_t));
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipEventCreateWithFlags(&c->flag_ev[k], hipEventDisableTiming);
    if (e == hipSuccess) {
        // this device's static RDF scratch; a failed lookup only disables komd_rdf_noalloc
        const hipError_t es = rdf_scratch_init(c);
        if (es != hipSuccess) {
            c->rdf_scratch_err = hipGetErrorString(es);
            (void)hipGetLastError();  // not left for the next launch check to report
        }
    }
    *out = c;
    if (e != hipSuccess) return hip_fail(c, e, "komd_create");
    return KOMD_OK;
}

void komd_destroy(komd_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    drop_graph(c);
    surrogate_free(c);
    table_free(c);
    free_sites(c);
    if (c->d.scal) hipFree(c->d.scal);
    if (c->d.flags) hipFree(c->d.flags);
    if (c->d.counters) hipFree(c->d.counters);
    if (c->flag_host) hipHostFree(c->flag_host);
    for (auto e : c->flag_ev)
        if (e) hipEventDestroy(e);
    drop_topology(c);
    for (auto& pe : c->pending) {
        hipEventDestroy(pe.second.first);
        hipEventDestroy(pe.second.second);
    }
    for (auto e : c->event_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* komd_last_error(const komd_ctx* c) { return c ? c->err.c_str() : "null context"; }

komd_status komd_grid_dims(const komd_ctx* c, int64_t dims[3]) {
    if (!c || !dims) return KOMD_ERR_ARG;
    dims[0] = c->grid.nx;
    dims[1] = c->grid.ny;
    dims[2] = c->grid.nz;
    return KOMD_OK;
}

komd_status komd_set_sites(komd_ctx* c, int64_t n, const double* r, const double* v, const double* mass,
                           const double* eps, const double* sig, const uint64_t* mol_id) {
    if (!c || n <= 0 || !r || !v || !mass || !eps || !sig || !mol_id) return KOMD_ERR_ARG;
    if (n > (int64_t)kIdxMask) return fail(c, KOMD_ERR_ARG, "at most 2^26-1 sites");
    drop_graph(c);
    free_sites(c);
    const int64_t n_old = c->n;
    c->n = n;
    surrogate_sites_changed(c);
    auto& d = c->d;
    // molecules: runs of equal consecutive ids (CheckpointIO.cpp:63-72); id classes for the exclusion
    std::vector<uint32_t> mstart;
    mstart.push_back(0);
    for (int64_t i = 1; i < n; i++)
        if (mol_id[i] != mol_id[i - 1]) mstart.push_back((uint32_t)i);
    mstart.push_back((uint32_t)n);
    // the bonded terms index sites and are grouped by molecule: they hold only for the same
    // sites and molecules
    if (n != n_old || mstart != c->mol_start_h) drop_topology(c);
    c->nmol = (int64_t)mstart.size() - 1;
    c->mol_start_h = mstart;
    c->multisite = c->nmol != n;
    std::map<uint64_t, uint32_t> cls;
    std::vector<uint32_t> idc((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        auto it = cls.find(mol_id[i]);
        if (it == cls.end()) it = cls.emplace(mol_id[i], (uint32_t)cls.size()).first;
        idc[(size_t)i] = it->second;
    }
    if (!c->multisite && (int64_t)cls.size() != n) c->multisite = true;  // repeated ids: exclusion matters
    c->ids_runs = (int64_t)cls.size() == c->nmol;  // no id reappears after another
    c->uniform = true;
    for (int64_t i = 1; i < n; i++)
        if (eps[i] != eps[0] || sig[i] != sig[0]) { c->uniform = false; break; }
    c->eps_u = (float)eps[0];
    c->sig_u = (float)sig[0];
    c->eps0 = eps[0];
    c->sig0 = sig[0];
    // SoA split of the reference's interleaved d3 layout
    std::vector<double> buf((size_t)n);
    const size_t nb = (size_t)n * sizeof(double);
    const size_t C = (size_t)c->grid.ncell;
    // Small grids cannot fill the GPU with one workgroup per cell: the pair pass then
    // splits each cell's home sites over several workgroups of about 5 sites.
    {
        const int64_t C0 = c->grid.ncell;
        const int64_t mean = (n + C0 - 1) / std::max<int64_t>(C0, 1);
        c->cell_split = C0 >= 1024 ? 1 : (int)std::min<int64_t>(8, std::max<int64_t>(1, (mean + 4) / 5));
        if (c->xcd_slots == 0) {  // 8 XCDs; kCellWgPerCu workgroups of the pair pass per CU
            int dev = 0;
            hipDeviceProp_t p;
            c->xcd_slots = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
                               ? std::max(1, p.multiProcessorCount / 8) * kCellWgPerCu
                               : -1;
        }
    }
    const int pair_blocks = cell_blocks(c);
    c->red_blocks = (int)std::min<int64_t>(1024, (n + 255) / 256);
    const size_t red_n = std::max<size_t>({(size_t)2 * pair_blocks, (size_t)atm_blocks(c), (size_t)3 * c->red_blocks,
                                           (size_t)1024}) + 64;
    hipError_t e = hipSuccess;
    double** f64[] = {&d.rx, &d.ry, &d.rz, &d.vx, &d.vy, &d.vz, &d.fx, &d.fy, &d.fz, &d.mass, &d.eps, &d.sig, &d.esite};
    for (auto pp : f64)
        if (e == hipSuccess) e = dalloc(pp, (size_t)n);
    uint32_t** u32n[] = {&d.idcls, &d.key, &d.slot, &d.tmp, &d.sorted, &d.inv, &d.sid};
    for (auto pp : u32n)
        if (e == hipSuccess) e = dalloc(pp, (size_t)n);
    if (e == hipSuccess) e = dalloc(&d.mol_start, mstart.size());
    if (e == hipSuccess) e = dalloc(&d.count, C);
    if (e == hipSuccess) e = dalloc(&d.start, C + 1);
    if (e == hipSuccess) e = dalloc(&d.scanp, (C + 1023) / 1024);
    if (e == hipSuccess) e = dalloc(&d.spos, (size_t)n);
    if (e == hipSuccess) e = dalloc(&d.sprm, (size_t)n);
    if (e == hipSuccess) e = dalloc(&d.red, red_n);
    if (e == hipSuccess)
        e = dalloc(&d.vpart, (size_t)kSym * std::max<size_t>({(size_t)cell_blocks(c) + 1, (size_t)atm_blocks(c),
                                                               (size_t)512, (size_t)kKinBlocks}));
    if (e == hipSuccess) e = dalloc(&d.cnt_pair, (size_t)cell_blocks(c));
    // triplet visits, then the in-range pairs of Verlet-list ATM passes
    if (e == hipSuccess) e = dalloc(&d.cnt_trip, (size_t)2 * atm_blocks(c));
    if (e == hipSuccess) e = dzero(c, d.cnt_pair, (size_t)cell_blocks(c) * 8);
    if (e == hipSuccess) e = dzero(c, d.cnt_trip, (size_t)2 * atm_blocks(c) * 8);
    if (e != hipSuccess) {
        free_sites(c);
        return fail(c, KOMD_ERR_NOMEM, std::string("device allocation: ") + hipGetErrorString(e));
    }
    for (int k = 0; k < 3 && e == hipSuccess; k++) {
        for (int64_t i = 0; i < n; i++) buf[(size_t)i] = r[3 * i + k];
        e = h2d(c, k == 0 ? d.rx : (k == 1 ? d.ry : d.rz), buf.data(), nb);
        for (int64_t i = 0; i < n && e == hipSuccess; i++) buf[(size_t)i] = v[3 * i + k];
        if (e == hipSuccess)
            e = h2d(c, k == 0 ? d.vx : (k == 1 ? d.vy : d.vz), buf.data(), nb);
    }
    if (e == hipSuccess) e = h2d(c, d.mass, mass, nb);
    if (e == hipSuccess) e = h2d(c, d.eps, eps, nb);
    if (e == hipSuccess) e = h2d(c, d.sig, sig, nb);
    if (e == hipSuccess) e = h2d(c, d.idcls, idc.data(), (size_t)n * 4);
    if (e == hipSuccess) e = h2d(c, d.mol_start, mstart.data(), mstart.size() * 4);
    if (e == hipSuccess) e = dzero(c, d.fx, nb);
    if (e == hipSuccess) e = dzero(c, d.fy, nb);
    if (e == hipSuccess) e = dzero(c, d.fz, nb);
    if (e == hipSuccess) e = dzero(c, d.count, C * 4);
    if (e != hipSuccess) {
        free_sites(c);
        return hip_fail(c, e, "komd_set_sites");
    }
    c->have_sites = true;
    return KOMD_OK;
}

komd_status komd_set_state(komd_ctx* c, const double* r, const double* v, const double* f) {
    komd_status s = need_sites(c);
    if (s) return s;
    auto& d = c->d;
    const int64_t n = c->n;
    std::vector<double> buf((size_t)n);
    const double* src[3] = {r, v, f};
    double* dst[3][3] = {{d.rx, d.ry, d.rz}, {d.vx, d.vy, d.vz}, {d.fx, d.fy, d.fz}};
    for (int q = 0; q < 3; q++) {
        if (!src[q]) continue;
        for (int k = 0; k < 3; k++) {
            for (int64_t i = 0; i < n; i++) buf[(size_t)i] = src[q][3 * i + k];
            HIPCALL(c, h2d(c, dst[q][k], buf.data(), (size_t)n * 8), "komd_set_state");
        }
    }
    c->nl_valid = false;
    return KOMD_OK;
}

komd_status komd_get_sites(komd_ctx* c, double* r, double* v, double* f) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, hipStreamSynchronize(c->stream), "komd_get_sites");
    auto& d = c->d;
    const int64_t n = c->n;
    std::vector<double> buf((size_t)n);
    double* dst[3] = {r, v, f};
    double* src[3][3] = {{d.rx, d.ry, d.rz}, {d.vx, d.vy, d.vz}, {d.fx, d.fy, d.fz}};
    for (int q = 0; q < 3; q++) {
        if (!dst[q]) continue;
        for (int k = 0; k < 3; k++) {
            HIPCALL(c, hipMemcpy(buf.data(), src[q][k], (size_t)n * 8, hipMemcpyDeviceToHost), "komd_get_sites");
            for (int64_t i = 0; i < n; i++) dst[q][3 * i + k] = buf[(size_t)i];
        }
    }
    return KOMD_OK;
}

komd_status komd_set_topology(komd_ctx* c, int64_t nb, const uint32_t* bonds, double kb, double r0, int64_t na,
                              const uint32_t* angles, double ka, double theta0) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (nb < 0 || na < 0 || (nb && !bonds) || (na && !angles)) return fail(c, KOMD_ERR_ARG, "bad topology");
    for (int64_t i = 0; i < 2 * nb; i++)
        if ((int64_t)bonds[i] >= c->n) return fail(c, KOMD_ERR_ARG, "bond index out of range");
    for (int64_t i = 0; i < 3 * na; i++)
        if ((int64_t)angles[i] >= c->n) return fail(c, KOMD_ERR_ARG, "angle index out of range");
    if (nb >= (int64_t)1 << 29 || na >= (int64_t)1 << 29) return fail(c, KOMD_ERR_ARG, "at most 2^29-1 bonds and angles");
    drop_graph(c);
    drop_topology(c);
    // Group the terms by molecule when every bond and angle lies inside one molecule: one
    // thread per molecule then updates its own sites without atomics.
    const std::vector<uint32_t>& ms = c->mol_start_h;
    const int64_t nmol = c->nmol;
    std::vector<uint32_t> mol_of((size_t)c->n);
    for (int64_t m = 0; m < nmol; m++)
        for (uint32_t i = ms[(size_t)m]; i < ms[(size_t)m + 1]; i++) mol_of[i] = (uint32_t)m;
    bool by_mol = true;
    for (int64_t t = 0; t < nb && by_mol; t++) by_mol = mol_of[bonds[2 * t]] == mol_of[bonds[2 * t + 1]];
    for (int64_t t = 0; t < na && by_mol; t++)
        by_mol = mol_of[angles[3 * t]] == mol_of[angles[3 * t + 1]] && mol_of[angles[3 * t + 1]] == mol_of[angles[3 * t + 2]];
    std::vector<uint32_t> b_sorted(bonds, bonds + 2 * nb), a_sorted(angles, angles + 3 * na);
    std::vector<uint32_t> b_start((size_t)nmol + 1, 0), a_start((size_t)nmol + 1, 0);
    if (by_mol) {
        // stable counting sort by molecule
        for (int64_t t = 0; t < nb; t++) b_start[mol_of[bonds[2 * t]] + 1]++;
        for (int64_t t = 0; t < na; t++) a_start[mol_of[angles[3 * t]] + 1]++;
        for (int64_t m = 0; m < nmol; m++) {
            b_start[(size_t)m + 1] += b_start[(size_t)m];
            a_start[(size_t)m + 1] += a_start[(size_t)m];
        }
        std::vector<uint32_t> bp(b_start.begin(), b_start.end() - 1), ap(a_start.begin(), a_start.end() - 1);
        for (int64_t t = 0; t < nb; t++) {
            const uint32_t q = bp[mol_of[bonds[2 * t]]]++;
            b_sorted[2 * q] = bonds[2 * t];
            b_sorted[2 * q + 1] = bonds[2 * t + 1];
        }
        for (int64_t t = 0; t < na; t++) {
            const uint32_t q = ap[mol_of[angles[3 * t]]]++;
            for (int k = 0; k < 3; k++) a_sorted[3 * q + k] = angles[3 * t + k];
        }
        HIPCALL(c, dalloc(&c->d.bond_mol, b_start.size()), "komd_set_topology");
        HIPCALL(c, dalloc(&c->d.angle_mol, a_start.size()), "komd_set_topology");
        HIPCALL(c, h2d(c, c->d.bond_mol, b_start.data(), b_start.size() * 4), "komd_set_topology");
        HIPCALL(c, h2d(c, c->d.angle_mol, a_start.data(), a_start.size() * 4), "komd_set_topology");
    }
    c->bonded_by_mol = by_mol;
    {
        // per site: the terms it takes part in (term << 3 | angle << 2 | role), bonds first,
        // each in the order of the uploaded term arrays
        const size_t n = (size_t)c->n;
        std::vector<uint32_t> ts(n + 1, 0), tl((size_t)(2 * nb + 3 * na));
        for (int64_t t = 0; t < 2 * nb; t++) ts[b_sorted[(size_t)t] + 1]++;
        for (int64_t t = 0; t < 3 * na; t++) ts[a_sorted[(size_t)t] + 1]++;
        for (size_t i = 0; i < n; i++) ts[i + 1] += ts[i];
        std::vector<uint32_t> pos(ts.begin(), ts.end() - 1);
        for (int64_t t = 0; t < nb; t++)
            for (uint32_t r = 0; r < 2; r++) tl[pos[b_sorted[(size_t)(2 * t + r)]]++] = ((uint32_t)t << 3) | r;
        for (int64_t t = 0; t < na; t++)
            for (uint32_t r = 0; r < 3; r++)
                tl[pos[a_sorted[(size_t)(3 * t + r)]]++] = ((uint32_t)t << 3) | 4u | r;
        HIPCALL(c, dalloc(&c->d.site_tstart, n + 1), "komd_set_topology");
        HIPCALL(c, dalloc(&c->d.site_terms, std::max<size_t>(1, tl.size())), "komd_set_topology");
        HIPCALL(c, h2d(c, c->d.site_tstart, ts.data(), (n + 1) * 4), "komd_set_topology");
        if (!tl.empty()) HIPCALL(c, h2d(c, c->d.site_terms, tl.data(), tl.size() * 4), "komd_set_topology");
    }
    HIPCALL(c, dalloc(&c->d.bonds, (size_t)std::max<int64_t>(1, 2 * nb)), "komd_set_topology");
    HIPCALL(c, dalloc(&c->d.angles, (size_t)std::max<int64_t>(1, 3 * na)), "komd_set_topology");
    if (nb) HIPCALL(c, h2d(c, c->d.bonds, b_sorted.data(), (size_t)(2 * nb) * 4), "komd_set_topology");
    if (na) HIPCALL(c, h2d(c, c->d.angles, a_sorted.data(), (size_t)(3 * na) * 4), "komd_set_topology");
    c->nbonds = nb;
    c->nangles = na;
    c->kb = kb;
    c->r0 = r0;
    c->ka = ka;
    c->theta0 = theta0;
    return KOMD_OK;
}

komd_status komd_update_container(komd_ctx* c) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_com_wrap_key(c, false), "komd_update_container");
    HIPCALL(c, launch_bin(c), "komd_update_container");
    c->nl_valid = false;
    return sync_check(c, "komd_update_container");
}

komd_status komd_get_cells(komd_ctx* c, int64_t* site_cell, uint32_t* cell_count, uint32_t* cell_start,
                           uint32_t* sorted_idx) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, hipStreamSynchronize(c->stream), "komd_get_cells");
    const Grid& g = c->grid;
    const int64_t n = c->n;
    std::vector<uint32_t> start((size_t)g.ncell + 1);
    HIPCALL(c, hipMemcpy(start.data(), c->d.start, start.size() * 4, hipMemcpyDeviceToHost), "komd_get_cells");
    if (cell_start) memcpy(cell_start, start.data(), start.size() * 4);
    if (cell_count)
        for (int k = 0; k < g.ncell; k++) cell_count[k] = start[(size_t)k + 1] - start[(size_t)k];
    if (sorted_idx)
        HIPCALL(c, hipMemcpy(sorted_idx, c->d.sorted, (size_t)n * 4, hipMemcpyDeviceToHost), "komd_get_cells");
    if (site_cell) {
        std::vector<uint32_t> key((size_t)n);
        HIPCALL(c, hipMemcpy(key.data(), c->d.key, (size_t)n * 4, hipMemcpyDeviceToHost), "komd_get_cells");
        const int64_t ax = g.nx + 2, ay = g.ny + 2;
        for (int64_t i = 0; i < n; i++) {
            const int64_t k = key[(size_t)i];
            const int64_t x = k % g.nx, y = (k / g.nx) % g.ny, z = k / ((int64_t)g.nx * g.ny);
            site_cell[i] = (x + 1) + ax * ((y + 1) + ay * (z + 1));
        }
    }
    return KOMD_OK;
}

komd_status komd_forces_pair(komd_ctx* c, int32_t kind) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (kind == KOMD_PAIR_TABLE) {
        if (!c->tab) return fail(c, KOMD_ERR_STATE, "no pair table: call komd_set_pair_table first");
        HIPCALL(c, ensure_nl(c, false), "komd_forces_pair");
        HIPCALL(c, launch_cell(c, false, 1, false, false), "komd_forces_pair");
        HIPCALL(c, launch_table(c, false), "komd_forces_pair");
        return sync_check(c, "komd_forces_pair");
    }
    if (kind != KOMD_PAIR_LJ12_6 && kind != KOMD_PAIR_ATM2B) return fail(c, KOMD_ERR_ARG, "unknown pair kind");
    HIPCALL(c, launch_cell(c, true, 0, false, false, pair_scale(c, kind)), "komd_forces_pair");
    return sync_check(c, "komd_forces_pair");
}

komd_status komd_set_pair_table(komd_ctx* c, int64_t nodes, const double* x, const double* force, double force_low,
                                double force_high, const double* potential, double potential_low,
                                double potential_high) {
    if (!c) return KOMD_ERR_ARG;
    if (nodes < 2 || nodes > (1 << 20) || !x || !force)
        return fail(c, KOMD_ERR_ARG, "pair table: 2 to 2^20 nodes, x and force required");
    for (int64_t k = 0; k < nodes; k++) {
        if (!std::isfinite(x[k]) || !std::isfinite(force[k]) || (potential && !std::isfinite(potential[k])))
            return fail(c, KOMD_ERR_ARG, "pair table: non-finite node or value");
        if (k && !(x[k] > x[k - 1])) return fail(c, KOMD_ERR_ARG, "pair table: nodes must increase strictly");
    }
    if (!std::isfinite(force_low) || !std::isfinite(force_high) || !std::isfinite(potential_low) ||
        !std::isfinite(potential_high))
        return fail(c, KOMD_ERR_ARG, "pair table: non-finite default value");
    if (c->stream) HIPCALL(c, hipStreamSynchronize(c->stream), "komd_set_pair_table");
    drop_graph(c);  // a captured step holds the old table's addresses
    HIPCALL(c, table_upload(c, nodes, x, force, force_low, force_high, potential, potential_low, potential_high),
            "komd_set_pair_table");
    return KOMD_OK;
}

komd_status komd_select_pair(komd_ctx* c, int32_t kind) {
    if (!c) return KOMD_ERR_ARG;
    if (kind != KOMD_PAIR_LJ12_6 && kind != KOMD_PAIR_TABLE && kind != KOMD_PAIR_ATM2B)
        return fail(c, KOMD_ERR_ARG, "unknown pair kind");
    if (kind == KOMD_PAIR_TABLE && !c->tab)
        return fail(c, KOMD_ERR_STATE, "no pair table: call komd_set_pair_table first");
    if (kind != c->pair_kind) drop_graph(c);
    c->pair_kind = kind;
    return KOMD_OK;
}

komd_status komd_forces_bonded(komd_ctx* c) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (c->nbonds + c->nangles == 0) return KOMD_OK;
    HIPCALL(c, launch_bonded(c, false), "komd_forces_bonded");
    return sync_check(c, "komd_forces_bonded");
}

komd_status komd_forces_limit(komd_ctx* c) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_limit(c), "komd_forces_limit");
    return sync_check(c, "komd_forces_limit");
}

komd_status komd_forces_3b(komd_ctx* c, int32_t kind) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (kind < KOMD_3B_ATM_EXACT || kind > KOMD_3B_HYBRID) return fail(c, KOMD_ERR_ARG, "unknown 3-body kind");
    hipError_t e = forces_3b(c, kind, false);
    if (e == hipErrorNotSupported) return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    HIPCALL(c, e, "komd_forces_3b");
    return sync_check(c, "komd_forces_3b");
}

komd_status komd_integrate0(komd_ctx* c) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_kick_drift(c), "komd_integrate0");
    c->nl_valid = false;
    return sync_check(c, "komd_integrate0");
}

komd_status komd_integrate1(komd_ctx* c) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_kick(c, false), "komd_integrate1");
    return sync_check(c, "komd_integrate1");
}

komd_status komd_measure_temperature(komd_ctx* c, double* t) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_temperature(c), "komd_measure_temperature");
    s = sync_check(c, "komd_measure_temperature");
    if (s) return s;
    double sc[5];
    HIPCALL(c, hipMemcpy(sc, c->d.scal, sizeof(sc), hipMemcpyDeviceToHost), "komd_measure_temperature");
    if (t) *t = sc[3];
    return KOMD_OK;
}

komd_status komd_rescale(komd_ctx* c, double t0) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_rescale(c, t0), "komd_rescale");
    return sync_check(c, "komd_rescale");
}

komd_status komd_compute_forces(komd_ctx* c, int32_t kind3b, int32_t use_bonded) {
    komd_status s = need_sites(c);
    if (s) return s;
    HIPCALL(c, launch_com_wrap_key(c, false), "komd_compute_forces");
    HIPCALL(c, launch_bin(c), "komd_compute_forces");
    hipError_t e = enqueue_forces(c, kind3b, use_bonded && (c->nbonds + c->nangles) > 0);
    if (e == hipErrorNotSupported) return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
    HIPCALL(c, e, "komd_compute_forces");
    return sync_check(c, "komd_compute_forces");
}

}  // extern "C"

namespace {
// Error flags copied to a pinned host slot every kFlagEvery steps; the host reads the previous
// copy (kFlagEvery steps older) when it queues the next one, so an overflow or a non-finite
// state stops the run within 2 kFlagEvery steps without draining the queue at every check.
constexpr int64_t kFlagEvery = 32;
static_assert(kFlagEvery % kStepsPerGraph == 0 && kStepsPerGraph % 2 == 0, "multi-step graphs");

struct FlagWatch {
    bool pending[2] = {false, false};
    // after step k (0-based) of nsteps has been queued
    komd_status after(komd_ctx* c, int64_t k, int64_t nsteps) {
        if ((k + 1) % kFlagEvery != 0 || k + 1 >= nsteps) return KOMD_OK;
        const int slot = (int)(((k + 1) / kFlagEvery) & 1), prev = slot ^ 1;
        if (pending[prev]) {
            HIPCALL(c, hipEventSynchronize(c->flag_ev[prev]), "komd_run");
            pending[prev] = false;
            if (c->flag_host[prev]) {
                komd_status s = sync_check(c, "komd_run");
                if (s) return s;
            }
        }
        HIPCALL(c, launch_flag_snapshot(c, c->flag_host + slot), "komd_run");
        HIPCALL(c, hipEventRecord(c->flag_ev[slot], c->stream), "komd_run");
        pending[slot] = true;
        return KOMD_OK;
    }
};

// komd_run with Verlet lists (skin_run): the first step always rebuilds; a step's last
// kernel arms the next step's rebuild when some site would leave half the skin.  The step
// parities alternate (each step reads its own rebuild word and clears the other's): two
// single-step graphs, one per parity, and one graph of kStepsPerGraph steps from parity 0.  The run ends with a container update without force reset, so the
// positions are wrapped into the box and the grid matches them, as after a plain run.
komd_status run_skin(komd_ctx* c, int64_t nsteps, int kind3b, bool bonded) {
    HIPCALL(c, ensure_skin(c, kind3b != KOMD_3B_ATM_EXACT && kind3b != 0), "komd_run");
    HIPCALL(c, launch_skin_offsets(c), "komd_run");
    HIPCALL(c, upload_skin_args(c), "komd_run");
    HIPCALL(c, launch_skin_arm(c, c->skin_par), "komd_run");
    if (c->timing || nsteps < 3) {
        for (int64_t k = 0; k < nsteps; k++) {
            hipError_t e = enqueue_skin_step(c, kind3b, c->skin_par, bonded);
            if (e == hipErrorNotSupported) return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
            HIPCALL(c, e, "komd_run");
            c->skin_par ^= 1;
            if (c->timing && (k % 64) == 63) {
                komd_status s = sync_check(c, "komd_run");
                if (s) return s;
            }
        }
    } else {
        const int key = (1 << 20) + (bonded ? 1 << 16 : 0) + kind3b;
        if (!c->graph_exec || !c->graph_exec_odd || !c->graph_exec_multi || c->graph_key != key) {
            drop_graph(c);
            hipError_t e = hipSuccess;
            for (int par = 0; par < 2 && e == hipSuccess; par++)
                e = capture_graph(c, par ? &c->graph_exec_odd : &c->graph_exec,
                                  [&] { return enqueue_skin_step(c, kind3b, par, bonded); });
            if (e == hipSuccess)
                e = capture_graph(c, &c->graph_exec_multi, [&] {
                    hipError_t es = hipSuccess;
                    for (int k = 0; k < kStepsPerGraph && es == hipSuccess; k++)
                        es = enqueue_skin_step(c, kind3b, k & 1, bonded);
                    return es;
                });
            if (e == hipErrorNotSupported) {
                drop_graph(c);
                return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
            }
            if (e != hipSuccess) drop_graph(c);
            HIPCALL(c, e, "komd_run capture");
            c->graph_key = key;
        }
        FlagWatch fw;
        for (int64_t k = 0; k < nsteps;) {
            // whole multi-step graphs from parity 0, single steps to align and for the rest
            const bool multi = c->skin_par == 0 && nsteps - k >= kStepsPerGraph;
            const int steps = multi ? kStepsPerGraph : 1;
            hipGraphExec_t ge = multi ? c->graph_exec_multi : (c->skin_par ? c->graph_exec_odd : c->graph_exec);
            HIPCALL(c, hipGraphLaunch(ge, c->stream), "komd_run launch");
            if (!multi) c->skin_par ^= 1;
            for (int j = 0; j < steps; j++, k++) {
                komd_status s = fw.after(c, k, nsteps);
                if (s) return s;
            }
        }
    }
    if (nsteps > 0) {
        HIPCALL(c, launch_com_wrap_key(c, false, false), "komd_run");
        HIPCALL(c, launch_bin(c), "komd_run");
    }
    return sync_check(c, "komd_run");
}

komd_status run_plain(komd_ctx* c, int64_t nsteps, int32_t kind3b, int32_t use_bonded, double t0) {
    komd_status s = KOMD_OK;
    const bool bonded = use_bonded && (c->nbonds + c->nangles) > 0;
    const bool thermo = t0 > 0;
    // beta = 1 for the first kick-drift and the target the captured step reads, stream-ordered
    // ahead of the steps (the graph holds no host value of t0, so one graph serves every target)
    if (thermo) HIPCALL(c, launch_run_scalars(c, t0), "komd_run");
    {
        // device buffers of the chosen passes exist before any graph capture (no allocation while capturing)
        const int nl = lists_of(c, kind3b);
        if (nl) HIPCALL(c, ensure_nl(c, nl == 2), "komd_run");
        if (kind3b == KOMD_3B_SURROGATE || kind3b == KOMD_3B_HYBRID) {
            hipError_t e = surrogate_prepare(c, kind3b == KOMD_3B_HYBRID);
            if (e == hipErrorNotSupported) return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
            HIPCALL(c, e, "komd_run");
        }
    }
    if (skin_run(c, kind3b, bonded, thermo)) return run_skin(c, nsteps, kind3b, bonded);
    if (c->timing || nsteps < 3) {
        for (int64_t k = 0; k < nsteps; k++) {
            hipError_t e = enqueue_step(c, kind3b, bonded, thermo);
            if (e == hipErrorNotSupported) return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
            HIPCALL(c, e, "komd_run");
            if (c->timing && (k % 64) == 63) {
                s = sync_check(c, "komd_run");
                if (s) return s;
            }
        }
    } else {
        const int key = (c->pair_kind * 8 + kind3b) * 4 + (bonded ? 2 : 0) + (thermo ? 1 : 0);
        if (!c->graph_exec || !c->graph_exec_multi || c->graph_key != key) {
            drop_graph(c);
            hipError_t e = capture_graph(c, &c->graph_exec, [&] { return enqueue_step(c, kind3b, bonded, thermo); });
            if (e == hipSuccess)
                e = capture_graph(c, &c->graph_exec_multi, [&] {
                    hipError_t es = hipSuccess;
                    for (int k = 0; k < kStepsPerGraph && es == hipSuccess; k++)
                        es = enqueue_step(c, kind3b, bonded, thermo);
                    return es;
                });
            if (e == hipErrorNotSupported) {
                drop_graph(c);
                return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
            }
            if (e != hipSuccess) drop_graph(c);
            HIPCALL(c, e, "komd_run capture");
            c->graph_key = key;
        }
        FlagWatch fw;
        for (int64_t k = 0; k < nsteps;) {
            const bool multi = nsteps - k >= kStepsPerGraph;
            HIPCALL(c, hipGraphLaunch(multi ? c->graph_exec_multi : c->graph_exec, c->stream), "komd_run launch");
            for (int j = 0; j < (multi ? kStepsPerGraph : 1); j++, k++) {
                s = fw.after(c, k, nsteps);
                if (s) return s;
            }
        }
    }
    if (thermo && nsteps > 0) HIPCALL(c, launch_rescale(c, t0), "komd_run");
    return sync_check(c, "komd_run");
}

}  // namespace

extern "C" {

komd_status komd_run(komd_ctx* c, int64_t nsteps, int32_t kind3b, int32_t use_bonded, double t0) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (nsteps < 0) return fail(c, KOMD_ERR_ARG, "negative step count");
    if (kind3b < 0 || kind3b > KOMD_3B_HYBRID) return fail(c, KOMD_ERR_ARG, "unknown 3-body kind");
    const bool learned = kind3b == KOMD_3B_SURROGATE || kind3b == KOMD_3B_HYBRID;
    if (!learned || c->judge_every <= 0 || c->judge_fell) {
        if (learned && c->judge_fell) {
            // a judged run fell back: the exact ATM replaces the surrogate from here on
            kind3b = KOMD_3B_ATM_EXACT;
        }
        return run_plain(c, nsteps, kind3b, use_bonded, t0);
    }
    // exact ATM as judge every judge_every steps; fallback when the error is too large
    int64_t done = 0;
    while (done < nsteps) {
        const int64_t seg = std::min<int64_t>(c->judge_every, nsteps - done);
        s = run_plain(c, seg, kind3b, use_bonded, t0);
        if (s) return s;
        done += seg;
        double err[2];
        s = komd_surrogate_judge(c, err);
        if (s) return s;
        c->judge_checks++;
        c->judge_last = err[0];
        c->judge_worst = std::max(c->judge_worst, err[0]);
        if (err[0] > c->judge_max) {
            c->judge_fell = 1;
            kind3b = KOMD_3B_ATM_EXACT;
            s = komd_compute_forces(c, kind3b, use_bonded);  // current forces from the exact ATM
            if (s) return s;
            if (done < nsteps) return run_plain(c, nsteps - done, kind3b, use_bonded, t0);
            return KOMD_OK;
        }
    }
    return KOMD_OK;
}

komd_status komd_set_skin(komd_ctx* c, double skin) {
    if (!c) return KOMD_ERR_ARG;
    if (!(skin >= 0) || !std::isfinite(skin)) return fail(c, KOMD_ERR_ARG, "skin must be finite and >= 0");
    if (skin > 0 && c->have_sites && c->n > (int64_t)kSkinIdxMask)
        return fail(c, KOMD_ERR_ARG, "skin: Verlet lists hold at most 2^25 - 1 sites");
    if (skin > 0 && c->grid.cs < c->prm.cutoff + skin)
        return fail(c, KOMD_ERR_ARG,
                    "skin: the cell edge (" + std::to_string(c->grid.cs) + ") must be at least cutoff + skin (" +
                        std::to_string(c->prm.cutoff + skin) + ")");
    c->skin = skin;
    drop_graph(c);
    if (c->d.skin_req) HIPCALL(c, dzero(c, c->d.skin_req + 2, sizeof(uint32_t)), "komd_set_skin");
    return KOMD_OK;
}

komd_status komd_skin_rebuilds(komd_ctx* c, int64_t* rebuilds) {
    if (!c || !rebuilds) return KOMD_ERR_ARG;
    *rebuilds = 0;
    if (!c->d.skin_req) return KOMD_OK;
    uint32_t h = 0;
    HIPCALL(c, hipStreamSynchronize(c->stream), "komd_skin_rebuilds");
    HIPCALL(c, hipMemcpy(&h, c->d.skin_req + 2, sizeof(uint32_t), hipMemcpyDeviceToHost), "komd_skin_rebuilds");
    *rebuilds = (int64_t)h;
    return KOMD_OK;
}

komd_status komd_set_judge(komd_ctx* c, int64_t every, double max_rel_rms) {
    if (!c) return KOMD_ERR_ARG;
    if (every < 0 || !(max_rel_rms > 0)) return fail(c, KOMD_ERR_ARG, "judge: every >= 0, max_rel_rms > 0");
    c->judge_every = every;
    c->judge_max = max_rel_rms;
    c->judge_checks = 0;
    c->judge_last = c->judge_worst = 0;
    c->judge_fell = 0;
    return KOMD_OK;
}

komd_status komd_judge_report(komd_ctx* c, int64_t* checks, double* last_rel_rms, double* worst_rel_rms,
                              int32_t* fell_back) {
    if (!c) return KOMD_ERR_ARG;
    if (checks) *checks = c->judge_checks;
    if (last_rel_rms) *last_rel_rms = c->judge_last;
    if (worst_rel_rms) *worst_rel_rms = c->judge_worst;
    if (fell_back) *fell_back = c->judge_fell;
    return KOMD_OK;
}

}  // extern "C"

namespace {
// The energy passes of komd_measure_observables; they overwrite the forces, which the
// caller restores on every exit.
komd_status measure_observables(komd_ctx* c, int32_t kind3b, komd_observables* out, const std::vector<double>& fsave) {
    komd_status s = KOMD_OK;
    auto& d = c->d;
    int64_t cnt0[2];
    s = read_counters(c, cnt0, false);
    if (s) return s;
    HIPCALL(c, launch_com_wrap_key(c, false), "komd_measure_observables");
    HIPCALL(c, launch_bin(c), "komd_measure_observables");
    double* out_dev = d.scal + kScalOut;
    double vals[8];
    // reduce `nslots` slots of `nparts` partials and copy them to vals
    auto reduce = [&](const double* part, int nparts, int nslots) -> komd_status {
        HIPCALL(c, launch_reduce_parts(c, part, nparts, nslots, out_dev), "komd_measure_observables");
        komd_status st = sync_check(c, "komd_measure_observables");
        if (st) return st;
        HIPCALL(c, hipMemcpy(vals, out_dev, (size_t)nslots * sizeof(double), hipMemcpyDeviceToHost),
                "komd_measure_observables");
        return KOMD_OK;
    };
    // virial tensor partials of the pass just run, added to out->virial_tensor
    auto add_virial = [&](int nparts) -> komd_status {
        komd_status st = reduce(d.vpart, nparts, kSym);
        if (st) return st;
        for (int k = 0; k < kSym; k++) out->virial_tensor[k] += vals[k];
        return KOMD_OK;
    };
    const int atm = atm_code(kind3b);
    const int nl = lists_of(c, kind3b);
    if (nl) HIPCALL(c, ensure_nl(c, nl == 2), "komd_measure_observables");
    if (table_pair(c)) {
        HIPCALL(c, launch_cell(c, false, nl, false, false), "komd_measure_observables");
        HIPCALL(c, launch_table(c, true), "komd_measure_observables");
        if ((s = reduce(d.red, table_blocks(c), 1))) return s;
        out->u_lj = vals[0];
        if ((s = add_virial(table_blocks(c)))) return s;
    } else {
        HIPCALL(c, launch_cell(c, true, nl, true, false, pair_scale(c, c->pair_kind)), "komd_measure_observables");
        if ((s = reduce(d.red, cell_blocks(c), 1))) return s;
        out->u_lj = vals[0];
        if ((s = add_virial(cell_blocks(c)))) return s;
    }
    if (atm) {
        HIPCALL(c, launch_atm(c, atm, true, nullptr, nullptr, nullptr), "komd_measure_observables");
        if ((s = reduce(d.red, atm_blocks(c), 1))) return s;
        out->u_3b = vals[0];
        if ((s = add_virial(atm_blocks(c)))) return s;
    }
    if (c->nbonds + c->nangles > 0) {
        HIPCALL(c, dzero(c, d.red, 512 * sizeof(double)), "komd_measure_observables");
        HIPCALL(c, dzero(c, d.vpart, (size_t)kSym * 512 * sizeof(double)), "komd_measure_observables");
        HIPCALL(c, launch_bonded(c, true), "komd_measure_observables");
        if ((s = reduce(d.red, 256, 2))) return s;
        out->u_bond = vals[0];
        out->u_angle = vals[1];
        if ((s = add_virial(512))) return s;
    }
    if (kind3b && !atm) {
        hipError_t e = forces_3b(c, kind3b, true);
        if (e == hipErrorNotSupported) return fail(c, KOMD_ERR_STATE, "surrogate parameters not set");
        HIPCALL(c, e, "komd_measure_observables");
        if ((s = reduce(d.red, surrogate_energy_blocks(c), 1))) return s;
        out->u_3b = vals[0];
        if ((s = add_virial(surrogate_virial_blocks(c)))) return s;
    }
    out->virial = out->virial_tensor[0] + out->virial_tensor[2] + out->virial_tensor[5];
    int64_t cnt1[2];
    s = read_counters(c, cnt1, false);
    if (s) return s;
    out->pairs = cnt1[0] - cnt0[0];
    out->triplets = (cnt1[1] - cnt0[1]) / 3;  // every triplet is visited by its 3 members
    // kinetic tensor, kinetic energy and temperature
    HIPCALL(c, launch_kin_tensor(c), "komd_measure_observables");
    if ((s = reduce(d.vpart, kKinBlocks, kSym))) return s;
    for (int k = 0; k < kSym; k++) out->kinetic_tensor[k] = vals[k];
    HIPCALL(c, launch_temperature(c), "komd_measure_observables");
    s = sync_check(c, "komd_measure_observables");
    if (s) return s;
    double sc[5];
    HIPCALL(c, hipMemcpy(sc, d.scal, sizeof(sc), hipMemcpyDeviceToHost), "komd_measure_observables");
    out->temperature = sc[3];
    out->kinetic = sc[4];
    out->n_in_box = (int64_t)sc[1];
    double fm = 0;  // from the forces as they were
    for (double x : fsave) fm = std::max(fm, fabs(x));
    out->fmax = fm;
    return KOMD_OK;
}
}  // namespace

extern "C" {

komd_status komd_measure_observables(komd_ctx* c, int32_t kind3b, komd_observables* out) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (!out) return KOMD_ERR_ARG;
    memset(out, 0, sizeof(*out));
    // keep the current forces: the energy passes rebuild them
    std::vector<double> fsave((size_t)(3 * c->n));
    s = komd_get_sites(c, nullptr, nullptr, fsave.data());
    if (s) return s;
    s = measure_observables(c, kind3b, out, fsave);
    const std::string first_err = c->err;
    // forces back as they were on every exit, also after a failed pass
    const komd_status s2 = komd_set_state(c, nullptr, nullptr, fsave.data());
    if (s) {
        c->err = first_err;
        return s;
    }
    return s2;
}

}  // extern "C"

namespace {
// Arguments of one RDF sample (RDFSensor.cpp:19: nbins = (uint64)(max_r / delta_r)).
komd_status rdf_args(komd_ctx* c, double max_r, double delta_r, int64_t nbins, const uint64_t* counts) {
    komd_status s = need_sites(c);
    if (s) return s;
    if (!counts) return fail(c, KOMD_ERR_ARG, "rdf: counts is NULL");
    if (!(max_r > 0) || !(delta_r > 0)) return fail(c, KOMD_ERR_ARG, "rdf: max_r and delta_r must be positive");
    const double q = max_r / delta_r;
    if (!(q >= 1.0) || !(q < (double)kRdfMaxBins + 1.0) || (int64_t)(uint64_t)q != nbins)
        return fail(c, KOMD_ERR_ARG,
                    "rdf: nbins must equal (uint64)(max_r / delta_r), between 1 and " + std::to_string(kRdfMaxBins));
    return KOMD_OK;
}
}  // namespace

extern "C" {

komd_status komd_rdf_noalloc(komd_ctx* c, double max_r, double delta_r, int64_t nbins, uint64_t* counts) {
    komd_status s = rdf_args(c, max_r, delta_r, nbins, counts);
    if (s) return s;
    if (c->nmol > kRdfScratchMol)
        return fail(c, KOMD_ERR_ARG, "rdf_noalloc: more than " + std::to_string(kRdfScratchMol) +
                                         " molecules, the capacity of the static scratch");
    if (!c->rdf_scratch) return fail(c, KOMD_ERR_HIP, "rdf_noalloc: no scratch address: " + c->rdf_scratch_err);
    HIPCALL(c, launch_rdf_noalloc(c, max_r, delta_r, (int)nbins), "komd_rdf_noalloc");
    s = sync_check(c, "komd_rdf_noalloc");
    if (s) return s;
    HIPCALL(c, hipMemcpy(counts, rdf_scratch_counts(c), (size_t)nbins * sizeof(uint64_t), hipMemcpyDeviceToHost),
            "komd_rdf_noalloc");
    return KOMD_OK;
}

komd_status komd_rdf(komd_ctx* c, double max_r, double delta_r, int64_t nbins, uint64_t* counts) {
    komd_status s = rdf_args(c, max_r, delta_r, nbins, counts);
    if (s) return s;
    HIPCALL(c, launch_rdf(c, max_r, delta_r, (int)nbins), "komd_rdf");
    s = sync_check(c, "komd_rdf");
    if (s) return s;
    HIPCALL(c, hipMemcpy(counts, c->d.rdf, (size_t)nbins * sizeof(uint64_t), hipMemcpyDeviceToHost), "komd_rdf");
    return KOMD_OK;
}

komd_status komd_timing_enable(komd_ctx* c, int32_t enable) {
    if (!c) return KOMD_ERR_ARG;
    if (c->stream) hipStreamSynchronize(c->stream);
    drain_timers(c);
    c->timing = enable != 0;
    c->timers.clear();
    return KOMD_OK;
}

komd_status komd_timing_get(komd_ctx* c, const char* name, double* total_ms, int64_t* launches) {
    if (!c || !name) return KOMD_ERR_ARG;
    if (c->stream) hipStreamSynchronize(c->stream);
    drain_timers(c);
    auto it = c->timers.find(name);
    if (total_ms) *total_ms = it == c->timers.end() ? 0.0 : it->second.total_ms;
    if (launches) *launches = it == c->timers.end() ? 0 : it->second.launches;
    return KOMD_OK;
}

komd_status komd_counters(komd_ctx* c, int64_t* pairs, int64_t* triplets, int32_t reset) {
    if (!c) return KOMD_ERR_ARG;
    int64_t v[2] = {0, 0};
    if (c->have_sites) {
        komd_status s = read_counters(c, v, reset != 0);
        if (s) return s;
    }
    if (pairs) *pairs = v[0];
    if (triplets) *triplets = v[1] / 3;  // every triplet is visited by its 3 members
    return KOMD_OK;
}

}  // extern "C"
