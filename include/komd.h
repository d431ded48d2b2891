/*
 * komd.h -- C ABI of the MI355X-native KoMD force hot path (libkomd_hip.so).
 *
 * The library owns device-resident state (positions, velocities, forces, the
 * linked-cell grid, neighbour lists, surrogate weights) on one HIP stream.
 * Host buffers passed in or out are owned by the caller and copied.  Every
 * call completes before it returns, mirroring the Kokkos::fence that ends
 * each reference functor (potentials/ForceFunctor.cpp:10-14), except
 * komd_run, which queues many steps and synchronises once at the end.
 *
 * Errors are status codes; komd_last_error() returns the message.  Nothing
 * throws across this boundary.  One context per host thread; a context is
 * not thread-safe (the reference orchestrates from one thread too).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * KoMD's src/).  Units are the caller's: the library applies no unit
 * conversion (the reference's conversions, util/constants.h:12-20, live in
 * the host layer, komd_amd/host/).
 */
#ifndef KOMD_H
#define KOMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    KOMD_OK = 0,
    KOMD_ERR_ARG = 1,        /* invalid argument (sizes, pointers, modes)          */
    KOMD_ERR_HIP = 2,        /* HIP runtime error (no device, launch failure, ...)  */
    KOMD_ERR_STATE = 3,      /* call out of order (e.g. forces before set_sites)    */
    KOMD_ERR_OVERFLOW = 4,   /* a fixed-capacity buffer overflowed (neighbour list) */
    KOMD_ERR_NONFINITE = 5,  /* a position/velocity/force became NaN or Inf          */
    KOMD_ERR_NOMEM = 6       /* device allocation failed                            */
} komd_status;

/* Binning rule of the container. */
enum {
    KOMD_BIN_COM = 0,   /* every site goes to its molecule's COM cell and wraps with
                           it (LinkedCells.cpp:87-103, :390-400) -- reference        */
    KOMD_BIN_SITE = 1   /* every site binned and wrapped on its own position       */
};

/* Three-body force kinds for komd_forces_3b. */
enum {
    KOMD_3B_ATM_EXACT = 1,   /* physical ATM, F = -grad V, every triplet within cutoff   */
    KOMD_3B_ATM_REF = 2,     /* reference ATM_NOLIST semantics: F = +grad V and only
                                triplets whose cells are all equal or all distinct
                                (ATM_NOLIST.cpp:20-159, LinkedCells.h:126-152)         */
    KOMD_3B_SURROGATE = 3,   /* learned descriptor-MLP surrogate of the ATM force         */
    KOMD_3B_HYBRID = 4       /* exact ATM in a slab, surrogate elsewhere, cos^2 blend    */
};

/* Pair force kinds for komd_forces_pair and komd_select_pair. */
enum {
    KOMD_PAIR_LJ12_6 = 1,   /* LJ 12-6, Lorentz-Berthelot mixing, same-molecule exclusion  */
    KOMD_PAIR_TABLE = 2,    /* tabulated radial force (the IBI force functor), no exclusion */
    KOMD_PAIR_ATM2B = 3     /* LJ 12-6 times 1 - 2 nu rho/(3 eps sig^6): the mean-field pair
                               form of ATM (potentials/ATM2B.cpp:9-56; rho = molecules per
                               volume, eps and sigma of the first site); energy scaled alike */
};

typedef struct {
    double low[3];       /* domain low corner  (config->domainLow)            */
    double high[3];      /* domain high corner (config->domainHigh)           */
    double cell_size;    /* linked-cell edge  (config->cell_size), >= cutoff  */
} komd_box;

typedef struct {
    double cutoff;        /* LJ and ATM cutoff radius (config->cutoff)                     */
    double delta_t;       /* time step (config->delta_t)                                   */
    double limit_factor;  /* Limit clamp fmax = limit*eps/sig (Limit.cpp:19-25); <= 0: off  */
    double nu;            /* ATM strength, caller units (ATM.cpp:13 converts before use)   */
    double temp_factor;   /* T = temp_factor * sum(m v.v)/(3 N_in_box) (TemperatureSensor)  */
    int32_t bin_mode;     /* KOMD_BIN_COM or KOMD_BIN_SITE                                 */
    int32_t max_neighbors;/* full neighbour-list capacity per site: fixed at 96; 0 or 96,     *
                           * any other value fails komd_create with KOMD_ERR_ARG              */
    int32_t reserved[4];
} komd_params;

typedef struct {
    double kinetic;        /* sum 0.5 m v.v over all sites                      */
    double temperature;    /* temp_factor * sum(m v.v)/(3 N_in_box)             */
    double u_lj;           /* pair energy: LJ 12-6 unshifted (LJ12_6_Sensor.cpp:76), or the
                              tabulated potential with KOMD_PAIR_TABLE selected (0 without one) */
    double u_3b;           /* three-body energy (exact ATM or surrogate)        */
    double u_bond;         /* harmonic bond energy                              */
    double u_angle;        /* harmonic angle energy                             */
    double virial;         /* trace of virial_tensor                            */
    double fmax;           /* largest |F| component                             */
    int64_t n_in_box;      /* sites inside [low, high] (temperature count)      */
    int64_t pairs;         /* unique LJ pairs within cutoff at the last pass    */
    int64_t triplets;      /* unique ATM triplets within cutoff at the last pass */
    /* Stress tensor parts, components XX XY YY XZ YZ ZZ (PressureSensor.h:18).
     * kinetic_tensor: sum_i m v_a v_b (PressureSensor.cpp:142-150).
     * virial_tensor: sum over interactions of sum_members r_a F_b, from relative
     * vectors (periodic-safe): LJ pairs dr_a F_b (PressureSensor.cpp:152-174);
     * ATM triplets with the physical sign for both ATM kinds, zz included
     * (PressureSensor.cpp:223-287 omits zz); harmonic bonds and angles; the
     * surrogate -sum_ij r_ij,a dE_i/dr_ij,b.  The hybrid blend's -V grad(w) and
     * E grad(lambda) terms act as external fields and are not included.
     * PressureSensor::write divides (kinetic + virial) by the volume. */
    double kinetic_tensor[6];
    double virial_tensor[6];
} komd_observables;

typedef struct komd_ctx komd_ctx;

/* Context owning device buffers and the HIP stream.  Replaces the Registry-owned
 * LinkedCells container (main.cpp:48, LinkedCells.cpp:11-45) and force functors
 * (main.cpp:57-66).  Validates cell_size >= cutoff (FileInput.cpp:117-120) and
 * at least 3 domain cells per dimension. */
komd_status komd_create(const komd_box* box, const komd_params* params, komd_ctx** out);
void komd_destroy(komd_ctx* ctx);
const char* komd_last_error(const komd_ctx* ctx);
komd_status komd_device_name(char* buf, int32_t len);

/* Domain cells per dimension, floor(L / cell_size) (LinkedCells.cpp:19). */
komd_status komd_grid_dims(const komd_ctx* ctx, int64_t dims[3]);

/* Copies sites in the reference's SoA/AoS layout (container/SOA.h:11-38):
 * r, v interleaved d3 (3n doubles), per-site mass/eps/sig, molecule id per site.
 * Consecutive sites with equal ids form one molecule (CheckpointIO.cpp:63-72).
 * Replaces MoleculeContainer::constructSOAs (MoleculeContainer.cpp:37-53).
 * Forces are zeroed.  A call with a different site count than the last one drops the
 * surrogate training set (its snapshots hold the old count); surrogate parameters stay.
 * A call that changes the site count or the molecules drops the bonded topology
 * (komd_set_topology indexes sites and groups its terms by molecule). */
komd_status komd_set_sites(komd_ctx* ctx, int64_t n, const double* r, const double* v, const double* mass,
                           const double* eps, const double* sig, const uint64_t* mol_id);
/* Overwrites positions / velocities / forces (any pointer may be NULL). */
komd_status komd_set_state(komd_ctx* ctx, const double* r, const double* v, const double* f);
/* Copies state out (any pointer may be NULL); replaces writeSOA2AOS (MoleculeContainer.cpp:24-35). */
komd_status komd_get_sites(komd_ctx* ctx, double* r, double* v, double* f);

/* Harmonic bonds (U = kb/2 (r-r0)^2) and angles (U = ka/2 (theta-theta0)^2, middle site
 * second).  New physics: the reference's only bonded force is FENE (potentials/FENE.cpp:15-43). */
komd_status komd_set_topology(komd_ctx* ctx, int64_t nbonds, const uint32_t* bonds, double kb, double r0,
                              int64_t nangles, const uint32_t* angles, double ka, double theta0);

/* COM -> periodic wrap -> force reset -> counting-sort binning.  Replaces
 * LinkedCells::updateContainer (LinkedCells.cpp:65-81). */
komd_status komd_update_container(komd_ctx* ctx);

/* Bit-exact view of the grid after komd_update_container:
 *   site_cell[n]      augmented flat cell index x + (nx+2)(y + (ny+2) z) per site
 *                     (LinkedCells.cpp:92-100, Vec3D.h:73-78)
 *   cell_count[nc]    sites per domain cell, nc = nx*ny*nz, domain flat order
 *   cell_start[nc+1]  CSR offsets into sorted_idx
 *   sorted_idx[n]     site indices grouped by cell, ascending inside a cell
 *                     (the reference appends in site order, Cell.cpp:32-34)
 * Any pointer may be NULL. */
komd_status komd_get_cells(komd_ctx* ctx, int64_t* site_cell, uint32_t* cell_count, uint32_t* cell_start,
                           uint32_t* sorted_idx);

/* LJ 12-6 over all site pairs of different molecules within cutoff whose cells are
 * neighbours (ForceFunctor::operator() -> LJ12_6::handlePairList, LJ12_6.cpp:10-43).
 * Adds to the forces.  Builds no neighbour list: komd_forces_3b builds the full lists
 * (every in-range neighbour of every site, capacity 96) that its kernels read. */
komd_status komd_forces_pair(komd_ctx* ctx, int32_t kind);
/* Tabulated pair force of KOMD_PAIR_TABLE: for every site pair within the cutoff
 * (no molecule exclusion), F(r) from a piecewise-linear table and the force (r_i - r_j)/r F on
 * i, its negative on j (IBI_Force::operator(), plugins/IBI/IBI_ForceFunctor.cpp:24-42).
 * F(r) is FunctionPL::evaluateAt (plugins/IBI/IBI_Math.h:30-72): force_high above the last
 * node, force_low below the first, otherwise linear interpolation on the segment whose upper
 * node is the first node >= r (at r equal to the first node: the first value).  x: nodes,
 * 2 to 2^20 of them, strictly increasing; force: values at the nodes.  potential (may be NULL)
 * is the pair potential on the same nodes with its own defaults, used only for energies.
 * Replaces FunctionPL::setXValues / setYValues / setLowerDefault / setUpperDefault on
 * IBI_ForceFunctor::getForceFunction() and getPotentialFunction() (IBI_Math.h:15-32,
 * IBI.cpp:61-69).  The table is copied. */
komd_status komd_set_pair_table(komd_ctx* ctx, int64_t nodes, const double* x, const double* force,
                                double force_low, double force_high, const double* potential, double potential_low,
                                double potential_high);
/* The pair force that komd_run, komd_compute_forces and komd_measure_observables apply:
 * KOMD_PAIR_LJ12_6 (default), KOMD_PAIR_TABLE (needs komd_set_pair_table) or
 * KOMD_PAIR_ATM2B, as the IBI plugin replaces the LJ functor in the registry
 * (plugins/IBI/IBI.cpp:51-55) and main.cpp registers ATM2B in LJ's place. */
komd_status komd_select_pair(komd_ctx* ctx, int32_t kind);
/* Harmonic bond and angle forces (new; replaces the reference's FENE functor slot). */
komd_status komd_forces_bonded(komd_ctx* ctx);
/* Per-component clamp (Limit::handlePairList, Limit.cpp:19-25); no-op when limit_factor <= 0. */
komd_status komd_forces_limit(komd_ctx* ctx);
/* Three-body forces (ForceFunctor3B::operator() -> ATM_NOLIST / ATM, ForceFunctor3B.cpp:10-14). */
komd_status komd_forces_3b(komd_ctx* ctx, int32_t kind);

/* Split velocity Verlet (Integrator.cpp:10-37). */
komd_status komd_integrate0(komd_ctx* ctx);
komd_status komd_integrate1(komd_ctx* ctx);
/* TemperatureSensor::measure (TemperatureSensor.cpp:22-43). */
komd_status komd_measure_temperature(komd_ctx* ctx, double* temperature);
/* VelocityScaling::apply (VelocityScaling.cpp:12-25): v *= sqrt(T0 / T_last_measured). */
komd_status komd_rescale(komd_ctx* ctx, double t0);

/* Energies, virial and counts of the current configuration (FP64 reductions).
 * Recomputes the force passes in energy mode; forces are left as they were.
 * kind3b: 0 none, else a KOMD_3B_* kind.  Replaces LJ12_6_Sensor / PressureSensor. */
komd_status komd_measure_observables(komd_ctx* ctx, int32_t kind3b, komd_observables* out);

/* One sample of RDFSensor::measure (RDFSensor.cpp:23-35, kernel :59-72): for every
 * ordered pair of distinct molecules, the minimum-image distance r of their
 * centres of mass (delta - L round(delta / L) per axis, MoleculeContainer.cpp:
 * 55-91 for the COM); when r <= max_r, one count in bin clamp((uint64)(r/delta_r),
 * 0, nbins - 1).  nbins must equal (uint64)(max_r / delta_r) (RDFSensor.cpp:19),
 * at most 8192.  counts receives this sample's nbins totals (overwritten);
 * RDFSensor::write's averaging and g(r) normalisation stay with the caller. */
komd_status komd_rdf(komd_ctx* ctx, double max_r, double delta_r, int64_t nbins, uint64_t* counts);

/* komd_rdf without any allocation, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The molecule centres and the counts live in a static
 * device array of 6.1 MiB that each GPU the library runs on holds until the process ends:
 * at most 262,144 molecules (more: KOMD_ERR_ARG) and 8,192 bins.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (contexts on several host
 * threads or streams) race; calls in sequence do not.  Results equal komd_rdf's. */
komd_status komd_rdf_noalloc(komd_ctx* ctx, double max_r, double delta_r, int64_t nbins, uint64_t* counts);

/* Simulation::run main loop (Simulation.cpp:49-74) on the device: nsteps x
 * {integrate0, update_container, LJ, [bonded], [limit], [3-body], integrate1,
 *  [temperature + rescale to t0 when t0 > 0]}.  kind3b: 0 none, else KOMD_3B_*.
 * use_bonded: run the bonded kernel.  Forces must be current on entry
 * (komd_compute_forces).  Synchronises once at the end.
 * Kernel-raised errors (KOMD_ERR_OVERFLOW, KOMD_ERR_NONFINITE) are checked every 32
 * steps without draining the queue, so up to 64 further steps may have run on capped
 * lists or non-finite values before the call returns the error: after such an error the
 * state is undefined (restore it with komd_set_state). */
komd_status komd_run(komd_ctx* ctx, int64_t nsteps, int32_t kind3b, int32_t use_bonded, double t0);
/* The force part of one step: update_container + LJ + [bonded] + [limit] + [3-body]
 * (the pre-loop pass of Simulation.cpp:15-24). */
komd_status komd_compute_forces(komd_ctx* ctx, int32_t kind3b, int32_t use_bonded);

/* Verlet lists for komd_run, the reference's reuse of its pair list over several steps
 * (LinkedCells::updatePairList every pair_steps steps, PairList.h:22-24, LinkedCells.cpp:130-131)
 * made exact: the list holds every site within cutoff + skin at its last rebuild, the force
 * kernel tests the cutoff at the current positions each step (the same FP32 test and FP64
 * decision near the cutoff as a fresh list), and the device rebuilds the container and the
 * lists before any site has moved more than skin/2 since the last rebuild.  Every step's
 * interaction set therefore equals the one of a container update at that step.  Applies to
 * komd_run without thermostat (t0 <= 0) with LJ 12-6 of uniform sites without the clamp and
 * either KOMD_3B_ATM_EXACT, KOMD_3B_SURROGATE or KOMD_3B_HYBRID (single-site, no bonded
 * terms), or no three-body term: single-site molecules, or multi-site molecules with
 * KOMD_BIN_SITE whose ids each form one run of consecutive sites (their lists leave
 * same-molecule pairs out, LJ12_6.cpp:19); bonded terms allowed; other runs ignore it.  The run's last
 * container update wraps the positions and leaves the grid consistent with them.
 * skin = 0 (default) turns it off; skin > 0 needs cell_size >= cutoff + skin. */
komd_status komd_set_skin(komd_ctx* ctx, double skin);
/* Container updates with list rebuilds of skinned runs since the last komd_set_skin (the
 * first step of every skinned komd_run is one). */
komd_status komd_skin_rebuilds(komd_ctx* ctx, int64_t* rebuilds);

/* Per-kernel device time measured with HIP events on the context's stream.
 * enable = 1 starts accumulating (komd_run then launches without a graph). */
komd_status komd_timing_enable(komd_ctx* ctx, int32_t enable);
/* name: "pair", "atm", "bin_key", "bin_scan", "bin_scatter", "bin_sort", "kick_drift",
 * "kick", "bonded", "limit", "desc", "mlp", "desc_bwd", ... Total ms and launches. */
komd_status komd_timing_get(komd_ctx* ctx, const char* name, double* total_ms, int64_t* launches);
/* Cumulative in-range counts since the last reset (for roofline accounting). */
komd_status komd_counters(komd_ctx* ctx, int64_t* pairs, int64_t* triplets, int32_t reset);

/* ---- learned three-body surrogate ------------------------------------- */
/* Descriptor-MLP surrogate of the ATM energy (no reference counterpart; the nearest
 * reference method is the mean-field ATM2B scaling, potentials/ATM2B.cpp:9-20).
 * Per site i, within the cutoff:
 *   32 radial descriptors  sum_j exp(-((r_ij - mu_k)/w)^2) fc(r_ij),  mu_k = 0.85 + k*1.6/31, w = 1.6/31
 *   16 angular invariants  <T_a^L, T_b^L>, T_a^L = sum_j psi_a(r_ij) u_ij^(x L), L = 1, 2, 3,
 *                          psi_a(r) = exp(-((r - nu_a)/0.35)^2) fc(r) / r^3, nu = 1.0, 1.35, 1.75, 2.15
 *   fc(r) = (cos(pi r / rc) + 1) / 2
 * x = (D - shift) * scale -> 64 -> 64 -> 1 with SiLU; BF16 MFMA operands, FP32 accumulation.
 * The three-body energy is sum_i E_i; forces are -dE/dr by a hand-written backward. */
#define KOMD_SUR_NDESC 48
#define KOMD_SUR_H1 64
#define KOMD_SUR_H2 64
/* Number of float parameters: W1[64][48], b1[64], W2[64][64], b2[64], w3[64], b3[1],
 * then the 48 descriptor shifts and the 48 descriptor scales. */
#define KOMD_SUR_NPARAM (64 * 48 + 64 + 64 * 64 + 64 + 64 + 1 + 48 + 48)
komd_status komd_surrogate_set_params(komd_ctx* ctx, const float* params, int64_t n);
komd_status komd_surrogate_get_params(komd_ctx* ctx, float* params, int64_t n);
/* Raw (unstandardised) descriptors of every site of the current configuration,
 * [n][48] FP32 in site order.  Rebuilds the container and neighbour lists. */
komd_status komd_surrogate_descriptors(komd_ctx* ctx, float* desc_out);
/* Per-site surrogate energies E_i ([n], FP32, site order) of the current configuration. */
komd_status komd_surrogate_energies(komd_ctx* ctx, float* e_out);
/* Exact per-site ATM energy shares V/3 ([n], FP64, site order) of the current configuration. */
komd_status komd_atm_site_energies(komd_ctx* ctx, double* e_out);

/* Training set on the device: komd_train_add_snapshot stores the current positions with
 * their exact (physical) ATM forces and per-site energy shares; positions, velocities and
 * forces of the context are left as they were. */
komd_status komd_train_add_snapshot(komd_ctx* ctx);
komd_status komd_train_clear(komd_ctx* ctx);
komd_status komd_train_size(komd_ctx* ctx, int64_t* n_snapshots);
/* Target statistics of the stored set: [0] mean E_i, [1] std E_i, [2] RMS force component. */
komd_status komd_train_stats(komd_ctx* ctx, double out[3]);
/* Sets shift = mean and scale = 1/std of every descriptor over the stored snapshots
 * (allocates zero parameters first when none are set). */
komd_status komd_train_standardize(komd_ctx* ctx);
/* Loss L = wE/N sum_i (E_i - E*_i)^2 + wF/(3N) sum_i |F_i - F*_i|^2 of one stored snapshot
 * and its gradient with respect to the parameters (shift/scale entries 0). */
komd_status komd_train_gradient(komd_ctx* ctx, int64_t snapshot, float w_energy, float w_force, float* grad,
                                double* loss);
/* One epoch: every stored snapshot once, in an order drawn from seed, one Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8) per snapshot with learning rate lr.  loss: mean
 * of the per-snapshot losses before their steps.  Positions and forces of the context
 * are left as they were. */
komd_status komd_train_epoch(komd_ctx* ctx, float lr, float w_energy, float w_force, uint64_t seed,
                             double* loss);
/* Exact ATM as judge: three-body force error of the surrogate at the current
 * configuration, out[0] = RMS |F_sur - F_atm| / RMS |F_atm|, out[1] = max |F_sur - F_atm| /
 * RMS |F_atm|.  Positions and forces are left as they were. */
komd_status komd_surrogate_judge(komd_ctx* ctx, double out[2]);
/* komd_run with KOMD_3B_SURROGATE or KOMD_3B_HYBRID judges every `every` steps (0: never);
 * when the relative RMS error exceeds max_rel_rms the run continues with the exact ATM. */
komd_status komd_set_judge(komd_ctx* ctx, int64_t every, double max_rel_rms);
/* Judge history since the last komd_set_judge: checks, last and worst relative RMS error,
 * and whether the run fell back to the exact ATM. */
komd_status komd_judge_report(komd_ctx* ctx, int64_t* checks, double* last_rel_rms, double* worst_rel_rms,
                              int32_t* fell_back);
/* Hybrid region for KOMD_3B_HYBRID: weight l = 1 (exact ATM) for |x - center| < half_width
 * along axis (minimum image), cos^2 descent to 0 over `blend`, surrogate weight 1 - l.
 * A triplet's ATM energy is weighted by the mean l of its members, a site's surrogate
 * energy by 1 - l_i (the AdResS-style weighting of plugins/ADR/AdResS.cpp:60-74, applied
 * to energies so that the blended force is conservative). */
komd_status komd_set_hybrid(komd_ctx* ctx, int32_t axis, double center, double half_width, double blend);

#ifdef __cplusplus
}
#endif
#endif /* KOMD_H */
