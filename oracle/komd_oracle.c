/*
 * komd_oracle.c -- FP64 CPU restatement of KoMD's force hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
 * tests under tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py, never by the product library (komd_amd/).  It restates the
 * reference algorithm step by step, in plain C with OpenMP, so that GPU
 * results can be compared with it on identical inputs.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the KoMD source tree, src/).  Pins (what checks each function against
 * something other than itself) are listed in tests/test_oracle_*.py:
 *   - or_grid_dims / or_update_com / or_periodic / or_bin: integer closed
 *     forms, invariants (sum of counts, stable ascending order, halo lists
 *     equal source lists), the 3-site..4-site example checkpoint.
 *   - or_lj_c08: brute-force minimum-image O(N^2) (or_lj_brute), textbook LJ
 *     closed form F(2^(1/6)) = 0, F(1) = 24, finite differences of the energy.
 *   - or_atm_c08 / or_atm_complete: textbook ATM energy (equilateral value
 *     11 nu / (8 r^9)), finite differences of or_atm_energy_triplet (which
 *     pins the sign of the reference's comp_force), brute-force O(N^3)
 *     minimum-image triplet sets (or_atm_brute) with the reference coverage
 *     rule stated independently (cells all equal or all distinct).
 *   - or_vv_step0/1: harmonic-oscillator closed form, time reversibility.
 *   - or_bonded: finite differences, sum of forces = 0.
 *   - or_virial: minus the strain derivative of the total energy of an
 *     isolated cluster, W_ab = -dU/de_ab under r -> (I + e E_ab) r, from the
 *     brute-force energies (tests/test_oracle_sensors.py).
 *   - or_rdf: exact coordination shells of a perfect FCC lattice, the count of
 *     ordered pairs within max_r from a numpy distance matrix, and a numpy
 *     histogram of the same distances.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off), so that the
 * FP64 operation order is exactly the one written here (no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Unit constants: util/constants.h:12-20 and math/SciValue.{h,cpp}.         */
/* SciValue keeps (mantissa, exponent) normalised to 0.1 < |m| < 10          */
/* (SciValue.h normalize) and converts with m * pow(10, e) (SciValue.cpp:8). */
/* ------------------------------------------------------------------------ */
typedef struct { double m; int64_t e; } sci_t;

static sci_t sci_norm(double m, int64_t e) {
    double a = fabs(m);
    while (a >= 10.0) { m /= 10.0; e += 1; a = fabs(m); }
    while (a <= 0.1) { m *= 10.0; e -= 1; a = fabs(m); }
    sci_t s = {m, e};
    return s;
}
static sci_t sci_mul(sci_t a, sci_t b) { return sci_norm(a.m * b.m, a.e + b.e); }   /* SciValue.cpp:38-40 */
static sci_t sci_div(sci_t a, sci_t b) { return sci_norm(a.m / b.m, a.e - b.e); }   /* SciValue.cpp:42-44 */
static sci_t sci_from(double v) { sci_t s = {v, 0}; if (v == 0.0) return s; return sci_norm(v, 0); }
static double sci_val(sci_t s) { return s.m * pow(10.0, (double)s.e); }

/* Temperature factor of TemperatureSensor.cpp:39-42:
 *   factor = conv_Aps_ms * conv_Aps_ms * (conv_Da_kg / kB), T = factor * (mv2 / (3 n)). */
EXPORT double or_temperature_of(double mv2, double n_sites) {
    sci_t kB = sci_norm(1.380649, -23);
    sci_t da = sci_norm(1.66053906892, -27);
    sci_t aps = sci_norm(1.0, 2);
    sci_t conv = sci_div(da, kB);
    sci_t factor = sci_mul(sci_mul(aps, aps), conv);
    double x = mv2 / (n_sites * 3.0);
    return sci_val(sci_mul(factor, sci_from(x)));
}

/* ATM strength in internal units, ATM.cpp:13 / ATM_NOLIST.cpp:17:
 * nu = energy_3b * 1e9 * double(conv_J_Ei). */
EXPORT double or_nu_internal(double energy_3b) {
    sci_t conv = sci_norm(6.022140754, 22);
    return energy_3b * 1e+9 * sci_val(conv);
}

/* epsilon conversion of the component deck: eps_internal = eps_J * conv_J_Ei. */
EXPORT double or_conv_J_Ei(void) { return sci_val(sci_norm(6.022140754, 22)); }

/* ------------------------------------------------------------------------ */
/* Grid, COM, periodic wrap, binning: container/LinkedCells.cpp             */
/* ------------------------------------------------------------------------ */

/* LinkedCells.cpp:16-20: num_cells_domain = ufloor(dom_size / cell_size). */
EXPORT void or_grid_dims(const double* low, const double* high, double cell_size, int64_t* nd) {
    for (int d = 0; d < 3; d++) {
        double size = high[d] - low[d];
        nd[d] = (int64_t)floor(size / cell_size);
    }
}

/* Reference cell bounds (LinkedCells.cpp:24-44) and halo classification
 * (LinkedCells.cpp:107-113): is_halo = (low < m_low) - (high > m_high).
 * Writes is_halo per augmented cell and dim (values -1, 0, 1). */
EXPORT void or_cell_halo(const double* low, const double* high, double cell_size, const int64_t* nd, int8_t* is_halo) {
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    /* bounds per dim are separable: compute the 1-d sequences */
    for (int d = 0; d < 3; d++) {
        double lo = low[d] - cell_size;
        for (int64_t c = 0; c < na[d]; c++) {
            double hi = (c == na[d] - 2) ? high[d] : (lo + cell_size);
            int8_t h = (int8_t)((lo < low[d]) - (hi > high[d]));
            /* store into all cells with coordinate c in dim d */
            for (int64_t z = 0; z < na[2]; z++)
                for (int64_t y = 0; y < na[1]; y++)
                    for (int64_t x = 0; x < na[0]; x++) {
                        int64_t cc = d == 0 ? x : (d == 1 ? y : z);
                        if (cc == c) is_halo[3 * (x + na[0] * (y + na[1] * z)) + d] = h;
                    }
            lo = hi;
        }
    }
}

/* MoleculeContainer.cpp:55-75: com = sum(r*m)/sum(m) per molecule, in site
 * order, broadcast to every site of the molecule. r and com are interleaved
 * d3 arrays (3*n doubles) like the reference's View<d3*>. */
EXPORT void or_update_com(int64_t nmol, const int64_t* mol_start, const double* r, const double* m, double* com) {
    for (int64_t mi = 0; mi < nmol; mi++) {
        int64_t s0 = mol_start[mi], s1 = mol_start[mi + 1];
        double c[3] = {0, 0, 0};
        double total = 0;
        for (int64_t s = s0; s < s1; s++) {
            const double mass = m[s];
            for (int d = 0; d < 3; d++) c[d] = c[d] + r[3 * s + d] * mass;
            total += mass;
        }
        for (int d = 0; d < 3; d++) c[d] = c[d] / total;
        for (int64_t s = s0; s < s1; s++)
            for (int d = 0; d < 3; d++) com[3 * s + d] = c[d];
    }
}

/* LinkedCells.cpp:390-400: strict single-image wrap by the COM. */
EXPORT void or_periodic(int64_t n, const double* com, double* r, const double* low, const double* high) {
    double size[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    for (int64_t i = 0; i < n; i++) {
        double off[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) {
            if (com[3 * i + d] < low[d]) off[d] = size[d];
            if (com[3 * i + d] > high[d]) off[d] = -size[d];
        }
        for (int d = 0; d < 3; d++) r[3 * i + d] = r[3 * i + d] + off[d];
    }
}

/* Augmented cell coordinate of one key (LinkedCells.cpp:92-97):
 * fBin = (com - low)/cell_size, clamp((int64)fBin, 0, n-1), +1 halo. */
static int64_t bin1(double com, double low, double cell_size, int64_t nd) {
    double fbin = (com - low) / cell_size;
    int64_t c = (int64_t)fbin;
    if (c < 0) c = 0;
    if (c > nd - 1) c = nd - 1;
    return c + 1;
}

/* writeIndices (LinkedCells.cpp:83-121).
 *  site_cell[i]   augmented flat index x + na_x*(y + na_y*z) (Vec3D.h:73-78)
 *  cell_count[c]  per augmented cell, halo cells hold copies of their source
 *  cell_start[c]  CSR offsets (size n_aug + 1) into cell_list
 *  cell_list      site indices, per cell in insertion order
 * Returns the total number of entries (N + halo copies).  Pass cell_list =
 * NULL to query the size. */
EXPORT int64_t or_bin(int64_t nmol, const int64_t* mol_start, const double* com, const double* low, const double* high,
                      double cell_size, const int64_t* nd, int64_t* site_cell, int64_t* cell_count, int64_t* cell_start,
                      int64_t* cell_list) {
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    int64_t ncell = na[0] * na[1] * na[2];
    int8_t* halo = (int8_t*)malloc((size_t)ncell * 3);
    or_cell_halo(low, high, cell_size, nd, halo);
    memset(cell_count, 0, sizeof(int64_t) * (size_t)ncell);
    int64_t nsites = mol_start[nmol];
    /* domain cells: every site of a molecule goes to the COM's cell */
    for (int64_t mi = 0; mi < nmol; mi++) {
        int64_t s0 = mol_start[mi];
        int64_t cx = bin1(com[3 * s0 + 0], low[0], cell_size, nd[0]);
        int64_t cy = bin1(com[3 * s0 + 1], low[1], cell_size, nd[1]);
        int64_t cz = bin1(com[3 * s0 + 2], low[2], cell_size, nd[2]);
        int64_t c = cx + na[0] * (cy + na[1] * cz);
        for (int64_t s = s0; s < mol_start[mi + 1]; s++) { site_cell[s] = c; cell_count[c]++; }
    }
    /* halo cells copy the source cell coord + is_halo * nd (LinkedCells.cpp:107-120) */
    int64_t* src = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    for (int64_t z = 0; z < na[2]; z++)
        for (int64_t y = 0; y < na[1]; y++)
            for (int64_t x = 0; x < na[0]; x++) {
                int64_t c = x + na[0] * (y + na[1] * z);
                const int8_t* h = halo + 3 * c;
                if (h[0] == 0 && h[1] == 0 && h[2] == 0) { src[c] = -1; continue; }
                int64_t sx = x + h[0] * nd[0], sy = y + h[1] * nd[1], sz = z + h[2] * nd[2];
                src[c] = sx + na[0] * (sy + na[1] * sz);
                cell_count[c] = cell_count[src[c]];
            }
    int64_t total = 0;
    for (int64_t c = 0; c < ncell; c++) { cell_start[c] = total; total += cell_count[c]; }
    cell_start[ncell] = total;
    if (cell_list) {
        int64_t* fill = (int64_t*)calloc((size_t)ncell, sizeof(int64_t));
        for (int64_t s = 0; s < nsites; s++) {
            int64_t c = site_cell[s];
            cell_list[cell_start[c] + fill[c]++] = s;
        }
        for (int64_t c = 0; c < ncell; c++) {
            if (src[c] < 0) continue;
            int64_t sc = src[c];
            for (int64_t k = 0; k < cell_count[sc]; k++) cell_list[cell_start[c] + k] = cell_list[cell_start[sc] + k];
        }
        free(fill);
    }
    free(src);
    free(halo);
    return total;
}

/* ------------------------------------------------------------------------ */
/* C08 cell-pair and cell-triplet enumeration: LinkedCells.h:64-152,        */
/* LinkedCells.cpp:443-572.                                                  */
/* ------------------------------------------------------------------------ */
static const int PAIR_OFF[13][2][3] = {
    /* LinkedCells.h:74-86 (o, x, y, z, xy, yz, xz, xyz as in :64-71) */
    {{0, 0, 0}, {0, 1, 0}}, {{0, 1, 0}, {0, 0, 1}}, {{0, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {0, 1, 1}},
    {{1, 0, 0}, {0, 1, 1}}, {{1, 0, 0}, {0, 1, 0}}, {{1, 0, 0}, {0, 0, 1}}, {{0, 0, 0}, {1, 0, 0}},
    {{0, 0, 0}, {1, 1, 0}}, {{1, 1, 0}, {0, 0, 1}}, {{0, 1, 0}, {1, 0, 1}}, {{0, 0, 0}, {1, 0, 1}},
    {{0, 0, 0}, {1, 1, 1}}};

#define O_ {0, 0, 0}
#define X_ {1, 0, 0}
#define Y_ {0, 1, 0}
#define Z_ {0, 0, 1}
#define XY_ {1, 1, 0}
#define YZ_ {0, 1, 1}
#define XZ_ {1, 0, 1}
#define XYZ_ {1, 1, 1}
static const int TRIPLE_OFF[44][3][3] = {
    /* LinkedCells.h:126-152, same order */
    {O_, X_, Y_}, {O_, X_, XY_}, {O_, Y_, XY_}, {O_, X_, Z_}, {O_, X_, XZ_}, {O_, X_, YZ_}, {O_, X_, XYZ_},
    {O_, Y_, Z_}, {O_, Y_, XZ_}, {O_, Y_, YZ_}, {O_, Y_, XYZ_}, {O_, XY_, Z_}, {O_, XY_, XZ_}, {O_, XY_, YZ_},
    {O_, XY_, XYZ_}, {O_, Z_, XZ_}, {O_, Z_, YZ_}, {O_, Z_, XYZ_}, {O_, XZ_, YZ_}, {O_, XZ_, XYZ_}, {O_, YZ_, XYZ_},
    {X_, Y_, XY_}, {X_, Y_, Z_}, {X_, Y_, XZ_}, {X_, Y_, YZ_}, {X_, Y_, XYZ_}, {X_, XY_, Z_}, {X_, XY_, YZ_},
    {Y_, XY_, Z_}, {Y_, XY_, XZ_}, {X_, Z_, XZ_}, {Y_, Z_, XZ_}, {XY_, Z_, XZ_}, {X_, Z_, YZ_}, {Y_, Z_, YZ_},
    {XY_, Z_, YZ_}, {X_, Z_, XYZ_}, {Y_, Z_, XYZ_}, {XY_, Z_, XYZ_}, {X_, XZ_, YZ_}, {Y_, XZ_, YZ_},
    {XY_, XZ_, YZ_}, {Y_, XZ_, XYZ_}, {X_, YZ_, XYZ_}};

/* getCell?Shift (LinkedCells.cpp:494-502): dom * (coord >= max) - dom * (coord <= 0), max = na - 1 */
static void cell_shift(const int64_t* c, const int64_t* na, const double* dom, double* s) {
    for (int d = 0; d < 3; d++) s[d] = dom[d] * (double)(c[d] >= na[d] - 1) - dom[d] * (double)(c[d] <= 0);
}
static int nonzero3(const double* s) { return s[0] != 0 || s[1] != 0 || s[2] != 0; }

/* The C08 base-cell walk of LCC08Iterator::operator++ (LinkedCells.cpp:448-468):
 * for colour 0..7 starting at begin = (c&1, c>>1&1, c>>2&1), step 2 per dim
 * while coord < max - 2 (max = na - 1), x fastest.  Fills base cells in the
 * iterator's order; returns their number. */
static int64_t c08_bases(const int64_t* na, int64_t* out) {
    int64_t nb = 0;
    int64_t max[3] = {na[0] - 1, na[1] - 1, na[2] - 1};
    for (int color = 0; color < 8; color++) {
        int64_t b[3] = {color & 1, (color >> 1) & 1, (color >> 2) & 1};
        for (int64_t z = b[2];; z += 2) {
            for (int64_t y = b[1];; y += 2) {
                for (int64_t x = b[0];; x += 2) {
                    if (out) { out[3 * nb] = x; out[3 * nb + 1] = y; out[3 * nb + 2] = z; }
                    nb++;
                    if (!(x < max[0] - 2)) break;
                }
                if (!(y < max[1] - 2)) break;
            }
            if (!(z < max[2] - 2)) break;
        }
    }
    return nb;
}

/* LJ12_6_Force (potentials/LJ12_6.cpp:15-43) on one candidate pair.  The
 * total potential follows the sensor (sensors/LJ12_6_Sensor.cpp:215-251):
 * counted when shift1 == 0 and the ids differ. */
static inline void lj_pair(int64_t i0, int64_t i1, const double* s0, const double* s1, const double* r,
                           const uint64_t* id, const double* sig, const double* eps, double cutoff2, double* f,
                           double* u, double* vir, int atomic) {
    if (id[i0] == id[i1]) return;
    double dr[3];
    for (int d = 0; d < 3; d++) dr[d] = (r[3 * i0 + d] + s0[d]) - (r[3 * i1 + d] + s1[d]);
    double dr2 = 0;
    for (int d = 0; d < 3; d++) dr2 += dr[d] * dr[d];
    if (dr2 > cutoff2) return;
    const double invdr2 = 1. / dr2;
    const double sigma = (sig[i0] + sig[i1]) / 2.0;
    const double epsilon = sqrt(eps[i0] * eps[i1]);
    const double sig2 = sigma * sigma;
    const double lj6 = pow(sig2 * invdr2, 3.0);
    const double lj12 = pow(lj6, 2.0);
    const double fac = 24.0 * epsilon * (2.0 * lj12 - lj6) * invdr2;
    double force[3] = {dr[0] * fac, dr[1] * fac, dr[2] * fac};
    int u0 = !nonzero3(s0), u1 = !nonzero3(s1);
    if (u0) {
        for (int d = 0; d < 3; d++) {
            if (atomic) {
#pragma omp atomic
                f[3 * i0 + d] += force[d];
            } else f[3 * i0 + d] += force[d];
        }
    }
    if (u1) {
        for (int d = 0; d < 3; d++) {
            if (atomic) {
#pragma omp atomic
                f[3 * i1 + d] -= force[d];
            } else f[3 * i1 + d] -= force[d];
        }
    }
    if (u && u1) {
        double uu = 4.0 * epsilon * (lj12 - lj6);
        if (atomic) {
#pragma omp atomic
            *u += uu;
        } else *u += uu;
    }
    if (vir && u1) { /* scalar virial sum r_ij . F_ij, each physical pair once */
        double w = (dr[0] * force[0] + dr[1] * force[1] + dr[2] * force[2]);
        if (atomic) {
#pragma omp atomic
            *vir += w;
        } else *vir += w;
    }
}

/* LJ forces over the reference's pair set: C08 cell pairs (LinkedCells.cpp:
 * 164-191, kernel LinkedCells.h:199-222: i = local % n0, j = local / n0)
 * followed by intra-cell pairs of domain cells (LinkedCells.cpp:199-223,
 * LinkedCells.h:234-258: triangular decode, pair (idx[i+1], idx[j])).
 * f is accumulated (the caller zeroes it, FReset_Kernel LinkedCells.cpp:386).
 * nthreads > 1 runs the cell pairs in parallel with atomic adds, as the
 * reference's Kokkos OpenMP backend does (Kokkos_Wrapper.h:24-28). */
EXPORT void or_lj_c08(int64_t n, const double* r, const uint64_t* id, const double* sig, const double* eps,
                      double cutoff, const double* low, const double* high, const int64_t* nd,
                      const int64_t* cell_start, const int64_t* cell_list, double* f, double* u_total,
                      double* virial, int nthreads) {
    (void)n;
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    double dom[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    double cutoff2 = pow(cutoff, 2);
    int64_t nb = c08_bases(na, NULL);
    int64_t* bases = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)nb);
    c08_bases(na, bases);
    int atomic = nthreads > 1;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t b = 0; b < nb; b++) {
        for (int o = 0; o < 13; o++) {
            int64_t c0[3], c1[3];
            for (int d = 0; d < 3; d++) { c0[d] = bases[3 * b + d] + PAIR_OFF[o][0][d]; c1[d] = bases[3 * b + d] + PAIR_OFF[o][1][d]; }
            double s0[3], s1[3];
            cell_shift(c0, na, dom, s0);
            cell_shift(c1, na, dom, s1);
            if (nonzero3(s0) && nonzero3(s1)) continue; /* checkState LinkedCells.cpp:471-474 */
            int64_t f0 = c0[0] + na[0] * (c0[1] + na[1] * c0[2]);
            int64_t f1 = c1[0] + na[0] * (c1[1] + na[1] * c1[2]);
            int64_t n0 = cell_start[f0 + 1] - cell_start[f0], n1 = cell_start[f1 + 1] - cell_start[f1];
            for (int64_t local = 0; local < n0 * n1; local++) {
                int64_t i = local % n0, j = local / n0;
                lj_pair(cell_list[cell_start[f0] + i], cell_list[cell_start[f1] + j], s0, s1, r, id, sig, eps, cutoff2,
                        f, u_total, virial, atomic);
            }
        }
    }
    /* intra-cell pairs, domain cells only */
    int64_t ncell = na[0] * na[1] * na[2];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t c = 0; c < ncell; c++) {
        int64_t x = c % na[0], y = (c / na[0]) % na[1], z = c / (na[0] * na[1]);
        if (x == 0 || y == 0 || z == 0 || x == na[0] - 1 || y == na[1] - 1 || z == na[2] - 1) continue;
        int64_t nc = cell_start[c + 1] - cell_start[c];
        const double zero[3] = {0, 0, 0};
        for (int64_t local = 0; local < nc * (nc - 1) / 2; local++) {
            int64_t i = (int64_t)floor((-1 + sqrt(1 + 8 * (double)local)) / 2.0);
            int64_t j = local - i * (i + 1) / 2;
            i += 1;
            lj_pair(cell_list[cell_start[c] + i], cell_list[cell_start[c] + j], zero, zero, r, id, sig, eps, cutoff2, f,
                    u_total, virial, atomic);
        }
    }
    free(bases);
}

/* Minimum-image brute force LJ (all pairs, distinct molecule ids), FP64.
 * Independent pin of the C08 enumeration on small systems. */
EXPORT void or_lj_brute(int64_t n, const double* r, const uint64_t* id, const double* sig, const double* eps,
                        double cutoff, const double* box, double* f, double* u_total) {
    double rc2 = cutoff * cutoff;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            if (id[i] == id[j]) continue;
            double d[3], d2 = 0;
            for (int k = 0; k < 3; k++) {
                d[k] = r[3 * i + k] - r[3 * j + k];
                d[k] -= box[k] * nearbyint(d[k] / box[k]);
                d2 += d[k] * d[k];
            }
            if (d2 > rc2) continue;
            double s = 0.5 * (sig[i] + sig[j]), e = sqrt(eps[i] * eps[j]);
            double sr2 = s * s / d2, sr6 = sr2 * sr2 * sr2;
            double fac = 24.0 * e * (2.0 * sr6 * sr6 - sr6) / d2;
            for (int k = 0; k < 3; k++) { f[3 * i + k] += d[k] * fac; f[3 * j + k] -= d[k] * fac; }
            if (u_total) *u_total += 4.0 * e * (sr6 * sr6 - sr6);
        }
}

/* ------------------------------------------------------------------------ */
/* Axilrod-Teller-Muto                                                      */
/* ------------------------------------------------------------------------ */

/* comp_force (potentials/ATM.h:36-69), restated term by term. */
EXPORT double or_atm_comp_force(double nu, double r_ab, double r_ac, double r_bc) {
    const double r_ab2 = r_ab * r_ab;
    const double r_ab3 = r_ab2 * r_ab;
    const double r_ab4 = r_ab2 * r_ab2;
    const double r_ab6 = r_ab3 * r_ab3;
    const double r_ac2 = r_ac * r_ac;
    const double r_ac3 = r_ac2 * r_ac;
    const double r_ac5 = r_ac3 * r_ac2;
    const double r_bc2 = r_bc * r_bc;
    const double r_bc3 = r_bc2 * r_bc;
    const double r_bc5 = r_bc3 * r_bc2;
    return ((3. * nu) / (8. * r_ab)) *
           (-(8.) / (r_ab4 * r_ac3 * r_bc3) - (1.) / (r_ac5 * r_bc5) + (5. * r_ac) / (r_ab6 * r_bc5) +
            (5. * r_bc) / (r_ab6 * r_ac5) - (1.) / (r_ab2 * r_ac3 * r_bc5) - (1.) / (r_ab2 * r_ac5 * r_bc3) -
            (3.) / (r_ab4 * r_ac * r_bc5) - (3.) / (r_ab4 * r_ac5 * r_bc) - (5.) / (r_ab6 * r_ac * r_bc3) -
            (5.) / (r_ab6 * r_ac3 * r_bc) + (6.) / (r_ab4 * r_ac3 * r_bc3));
}

/* Textbook triple-dipole energy V = nu (1 + 3 cos g1 cos g2 cos g3) / (r12 r13 r23)^3,
 * written with the law of cosines (no reference counterpart: the reference computes no
 * ATM energy; ATM.cpp and ATM_NOLIST.cpp only add forces). */
EXPORT double or_atm_energy_triplet(double nu, const double* ri, const double* rj, const double* rk) {
    double a2 = 0, b2 = 0, c2 = 0;
    for (int d = 0; d < 3; d++) {
        double x = rj[d] - ri[d], y = rk[d] - ri[d], z = rk[d] - rj[d];
        a2 += x * x; b2 += y * y; c2 += z * z;   /* a = |ij|, b = |ik|, c = |jk| */
    }
    double a = sqrt(a2), b = sqrt(b2), c = sqrt(c2);
    double ci = (a2 + b2 - c2) / (2 * a * b);
    double cj = (a2 + c2 - b2) / (2 * a * c);
    double ck = (b2 + c2 - a2) / (2 * b * c);
    return nu * (1.0 + 3.0 * ci * cj * ck) / (a2 * a * b2 * b * c2 * c);
}

/* ATM_NOLIST_Force (potentials/ATM_NOLIST.cpp:124-159) with a sign switch:
 * sign = +1 reproduces the reference (which applies +grad V), sign = -1 is
 * the physical force -grad V.  Forces go to unshifted members only. */
static inline void atm_triplet(int64_t i0, int64_t i1, int64_t i2, const double* s0, const double* s1, const double* s2,
                               const double* r, double nu, double cutoff2, double sign, double* f, int atomic,
                               int64_t* hits) {
    double r0[3], r1[3], r2[3], d01[3], d02[3], d12[3];
    for (int d = 0; d < 3; d++) {
        r0[d] = r[3 * i0 + d] + s0[d];
        r1[d] = r[3 * i1 + d] + s1[d];
        r2[d] = r[3 * i2 + d] + s2[d];
        d01[d] = r0[d] - r1[d];
        d02[d] = r0[d] - r2[d];
        d12[d] = r1[d] - r2[d];
    }
    double q01 = 0, q02 = 0, q12 = 0;
    for (int d = 0; d < 3; d++) q01 += d01[d] * d01[d];
    for (int d = 0; d < 3; d++) q02 += d02[d] * d02[d];
    for (int d = 0; d < 3; d++) q12 += d12[d] * d12[d];
    if (q01 > cutoff2 || q02 > cutoff2 || q12 > cutoff2) return;
    if (hits) {
        if (atomic) {
#pragma omp atomic
            (*hits)++;
        } else (*hits)++;
    }
    const double a = sqrt(q01), b = sqrt(q02), c = sqrt(q12);
    const double v01 = sign * or_atm_comp_force(nu, a, b, c);
    const double v02 = sign * or_atm_comp_force(nu, b, a, c);
    const double v12 = sign * or_atm_comp_force(nu, c, a, b);
    double f0[3], f1[3], f2[3];
    for (int d = 0; d < 3; d++) {
        f0[d] = d01[d] * v01 + d02[d] * v02;
        f1[d] = d01[d] * (-v01) + d12[d] * v12;
        f2[d] = d02[d] * (-v02) + d12[d] * (-v12);
    }
    const double* sh[3] = {s0, s1, s2};
    const int64_t idx[3] = {i0, i1, i2};
    const double* fv[3] = {f0, f1, f2};
    for (int m = 0; m < 3; m++) {
        if (nonzero3(sh[m])) continue;
        for (int d = 0; d < 3; d++) {
            if (atomic) {
#pragma omp atomic
                f[3 * idx[m] + d] += fv[m][d];
            } else f[3 * idx[m] + d] += fv[m][d];
        }
    }
}

/* Direct ATM over the reference's triplet set (potentials/ATM_NOLIST.cpp:20-122):
 * the 44 C08 cell triples per base cell (LinkedCells.h:126-152; walk
 * LinkedCells.cpp:509-534, skipped when all three cells are halo), candidates
 * decoded i = local % n0, j = (local / n0) % n1, k = local / (n0 n1)
 * (ATM_NOLIST.h:76-110), then same-cell triples of domain cells (ATM_NOLIST.h:
 * 112-157: tetrahedral/triangular decode).  sign: +1 reference, -1 physical.
 * Returns the number of in-range triplet evaluations. */
EXPORT int64_t or_atm_c08(int64_t n, const double* r, double nu, double cutoff, const double* low, const double* high,
                          const int64_t* nd, const int64_t* cell_start, const int64_t* cell_list, double sign, double* f,
                          int nthreads) {
    (void)n;
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    double dom[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    double cutoff2 = pow(cutoff, 2);
    int64_t nb = c08_bases(na, NULL);
    int64_t* bases = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)nb);
    c08_bases(na, bases);
    int atomic = nthreads > 1;
    int64_t hits = 0;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
    for (int64_t b = 0; b < nb; b++) {
        for (int o = 0; o < 44; o++) {
            int64_t c[3][3];
            double s[3][3];
            int64_t fc[3], ncnt[3];
            for (int m = 0; m < 3; m++) {
                for (int d = 0; d < 3; d++) c[m][d] = bases[3 * b + d] + TRIPLE_OFF[o][m][d];
                cell_shift(c[m], na, dom, s[m]);
                fc[m] = c[m][0] + na[0] * (c[m][1] + na[1] * c[m][2]);
                ncnt[m] = cell_start[fc[m] + 1] - cell_start[fc[m]];
            }
            if (nonzero3(s[0]) && nonzero3(s[1]) && nonzero3(s[2])) continue; /* LinkedCells.cpp:531-534 */
            int64_t tot = ncnt[0] * ncnt[1] * ncnt[2];
            for (int64_t local = 0; local < tot; local++) {
                int64_t i = local % ncnt[0];
                int64_t j = (local / ncnt[0]) % ncnt[1];
                int64_t k = local / (ncnt[0] * ncnt[1]);
                atm_triplet(cell_list[cell_start[fc[0]] + i], cell_list[cell_start[fc[1]] + j],
                            cell_list[cell_start[fc[2]] + k], s[0], s[1], s[2], r, nu, cutoff2, sign, f, atomic, &hits);
            }
        }
    }
    int64_t ncell = na[0] * na[1] * na[2];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t cc = 0; cc < ncell; cc++) {
        int64_t x = cc % na[0], y = (cc / na[0]) % na[1], z = cc / (na[0] * na[1]);
        if (x == 0 || y == 0 || z == 0 || x == na[0] - 1 || y == na[1] - 1 || z == na[2] - 1) continue;
        int64_t nc = cell_start[cc + 1] - cell_start[cc];
        const double zero[3] = {0, 0, 0};
        /* tri_sums[p] = (p+1)(p+2)(p+3)/6 (ATM_NOLIST.cpp:106-109) */
        for (int64_t local = 0; local < nc * (nc - 1) * (nc - 2) / 6; local++) {
            int64_t k_idx = -1;
            for (int64_t p = 0; p < nc; p++) {
                if ((p + 1) * (p + 2) * (p + 3) / 6 > local) { k_idx = p; break; }
            }
            int64_t k = k_idx + 2;
            int64_t tmp = local;
            if (k_idx > 0) tmp -= k_idx * (k_idx + 1) * (k_idx + 2) / 6;
            int64_t j = (int64_t)floor((-1 + sqrt(1 + 8 * (double)tmp)) / 2.0);
            int64_t i = tmp - j * (j + 1) / 2;
            j += 1;
            atm_triplet(cell_list[cell_start[cc] + i], cell_list[cell_start[cc] + j], cell_list[cell_start[cc] + k],
                        zero, zero, zero, r, nu, cutoff2, sign, f, atomic, &hits);
        }
    }
    free(bases);
    return hits;
}

/* Minimum-image displacement b - a in a periodic box. */
static inline void mi_disp(const double* a, const double* b, const double* box, double* d) {
    for (int k = 0; k < 3; k++) {
        d[k] = b[k] - a[k];
        d[k] -= box[k] * nearbyint(d[k] / box[k]);
    }
}

/* Brute-force O(N^3) ATM over minimum-image triplets with all three pair
 * distances <= cutoff.  coverage: 0 = every triplet (physics),
 * 1 = reference coverage restated independently: the three sites' cells
 * (site_cell) are all equal or pairwise distinct.  sign as above.
 * Forces per member, energy (textbook) summed over counted triplets. */
EXPORT int64_t or_atm_brute(int64_t n, const double* r, double nu, double cutoff, const double* box,
                            const int64_t* site_cell, int coverage, double sign, double* f, double* energy) {
    double rc2 = cutoff * cutoff;
    int64_t hits = 0;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            double dij[3];
            mi_disp(r + 3 * i, r + 3 * j, box, dij);
            double qij = dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2];
            if (qij > rc2) continue;
            for (int64_t k = j + 1; k < n; k++) {
                double dik[3];
                mi_disp(r + 3 * i, r + 3 * k, box, dik);
                double qik = dik[0] * dik[0] + dik[1] * dik[1] + dik[2] * dik[2];
                if (qik > rc2) continue;
                double djk[3];
                for (int d = 0; d < 3; d++) djk[d] = dik[d] - dij[d];
                double qjk = djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2];
                if (qjk > rc2) continue;
                if (coverage == 1) {
                    int64_t ci = site_cell[i], cj = site_cell[j], ck = site_cell[k];
                    int same = (ci == cj) && (cj == ck);
                    int distinct = (ci != cj) && (cj != ck) && (ci != ck);
                    if (!same && !distinct) continue;
                }
                hits++;
                double pi[3] = {0, 0, 0}, pj[3], pk[3];
                for (int d = 0; d < 3; d++) { pj[d] = dij[d]; pk[d] = dik[d]; }
                double a = sqrt(qij), b = sqrt(qik), c = sqrt(qjk);
                /* reference member order (0,1,2) = (i,j,k): dr0_1 = ri - rj, etc. */
                double v01 = sign * or_atm_comp_force(nu, a, b, c);
                double v02 = sign * or_atm_comp_force(nu, b, a, c);
                double v12 = sign * or_atm_comp_force(nu, c, a, b);
                for (int d = 0; d < 3; d++) {
                    double d01 = -dij[d], d02 = -dik[d], d12 = -djk[d];
                    f[3 * i + d] += d01 * v01 + d02 * v02;
                    f[3 * j + d] += -d01 * v01 + d12 * v12;
                    f[3 * k + d] += -d02 * v02 - d12 * v12;
                }
                if (energy) *energy += or_atm_energy_triplet(nu, pi, pj, pk);
            }
        }
    return hits;
}

/* Complete ATM over cell lists (every minimum-image triplet once, owner = the
 * lowest site index), O(N * neighbours^2).  coverage/sign as or_atm_brute.
 * Optional per-site energy share e_site[i] += V/3 for each member and total
 * energy.  Used as the FP64 reference of the GPU's physical ATM on large
 * snapshots. */
EXPORT int64_t or_atm_complete(int64_t n, const double* r, double nu, double cutoff, const double* low,
                               const double* high, const int64_t* nd, const int64_t* site_cell,
                               const int64_t* cell_start, const int64_t* cell_list, int coverage, double sign,
                               double* f, double* energy, double* e_site, int nthreads) {
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    double box[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    double rc2 = cutoff * cutoff;
    int64_t hits = 0;
    double etot = 0;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads) reduction(+ : hits, etot)
#endif
    {
        int64_t cap = 4096;
        int64_t* nb = (int64_t*)malloc(sizeof(int64_t) * cap);
        double* dv = (double*)malloc(sizeof(double) * 3 * cap);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t i = 0; i < n; i++) {
            int64_t ci = site_cell[i];
            int64_t cx = ci % na[0], cy = (ci / na[0]) % na[1], cz = ci / (na[0] * na[1]);
            int64_t m = 0;
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        int64_t x = cx + dx, y = cy + dy, z = cz + dz;
                        /* wrap inside the domain range 1..nd */
                        if (x < 1) x += nd[0]; if (x > nd[0]) x -= nd[0];
                        if (y < 1) y += nd[1]; if (y > nd[1]) y -= nd[1];
                        if (z < 1) z += nd[2]; if (z > nd[2]) z -= nd[2];
                        int64_t c = x + na[0] * (y + na[1] * z);
                        for (int64_t q = cell_start[c]; q < cell_start[c + 1]; q++) {
                            int64_t j = cell_list[q];
                            if (j <= i) continue;
                            double d[3];
                            mi_disp(r + 3 * i, r + 3 * j, box, d);
                            double q2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                            if (q2 > rc2) continue;
                            if (m == cap) {
                                cap *= 2;
                                nb = (int64_t*)realloc(nb, sizeof(int64_t) * cap);
                                dv = (double*)realloc(dv, sizeof(double) * 3 * cap);
                            }
                            nb[m] = j;
                            dv[3 * m] = d[0]; dv[3 * m + 1] = d[1]; dv[3 * m + 2] = d[2];
                            m++;
                        }
                    }
            for (int64_t a = 0; a < m; a++)
                for (int64_t b = a + 1; b < m; b++) {
                    double djk[3];
                    for (int d = 0; d < 3; d++) djk[d] = dv[3 * b + d] - dv[3 * a + d];
                    double qjk = djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2];
                    if (qjk > rc2) continue;
                    int64_t j = nb[a], k = nb[b];
                    if (coverage == 1) {
                        int64_t cj = site_cell[j], ck = site_cell[k];
                        int same = (ci == cj) && (cj == ck);
                        int distinct = (ci != cj) && (cj != ck) && (ci != ck);
                        if (!same && !distinct) continue;
                    }
                    hits++;
                    const double* dij = dv + 3 * a;
                    const double* dik = dv + 3 * b;
                    double qij = dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2];
                    double qik = dik[0] * dik[0] + dik[1] * dik[1] + dik[2] * dik[2];
                    double ra = sqrt(qij), rb = sqrt(qik), rcn = sqrt(qjk);
                    double v01 = sign * or_atm_comp_force(nu, ra, rb, rcn);
                    double v02 = sign * or_atm_comp_force(nu, rb, ra, rcn);
                    double v12 = sign * or_atm_comp_force(nu, rcn, ra, rb);
                    double fi[3], fj[3], fk[3];
                    for (int d = 0; d < 3; d++) {
                        double d01 = -dij[d], d02 = -dik[d], d12 = -djk[d];
                        fi[d] = d01 * v01 + d02 * v02;
                        fj[d] = -d01 * v01 + d12 * v12;
                        fk[d] = -d02 * v02 - d12 * v12;
                    }
                    for (int d = 0; d < 3; d++) {
#ifdef _OPENMP
#pragma omp atomic
#endif
                        f[3 * i + d] += fi[d];
#ifdef _OPENMP
#pragma omp atomic
#endif
                        f[3 * j + d] += fj[d];
#ifdef _OPENMP
#pragma omp atomic
#endif
                        f[3 * k + d] += fk[d];
                    }
                    if (energy || e_site) {
                        double z0[3] = {0, 0, 0};
                        double v = or_atm_energy_triplet(nu, z0, dij, dik);
                        etot += v;
                        if (e_site) {
                            double v3 = v / 3.0;
#ifdef _OPENMP
#pragma omp atomic
#endif
                            e_site[i] += v3;
#ifdef _OPENMP
#pragma omp atomic
#endif
                            e_site[j] += v3;
#ifdef _OPENMP
#pragma omp atomic
#endif
                            e_site[k] += v3;
                        }
                    }
                }
        }
        free(nb);
        free(dv);
    }
    if (energy) *energy += etot;
    return hits;
}

/* ------------------------------------------------------------------------ */
/* Limit (potentials/Limit.cpp:19-25): f = max(min(f, fmax), -fmax) per      */
/* component, fmax = limit * eps / sig.                                     */
/* ------------------------------------------------------------------------ */
EXPORT void or_limit(int64_t n, double limit_factor, const double* eps, const double* sig, double* f) {
    for (int64_t i = 0; i < n; i++) {
        const double fmax = limit_factor * eps[i] / sig[i];
        for (int d = 0; d < 3; d++) {
            double v = f[3 * i + d];
            v = v < fmax ? v : fmax;     /* Kokkos::min(a, b) */
            v = v > -fmax ? v : -fmax;   /* Kokkos::max(a, b) */
            f[3 * i + d] = v;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Velocity Verlet (integrators/Integrator.cpp:10-37), temperature          */
/* (sensors/TemperatureSensor.cpp:22-43), rescale (thermostats/              */
/* VelocityScaling.cpp:12-25).                                              */
/* ------------------------------------------------------------------------ */
EXPORT void or_vv_step0(int64_t n, double dt, double* r, double* v, const double* f, const double* m) {
    const double dt_halve = dt * 0.5;
    for (int64_t i = 0; i < n; i++) {
        const double dtInv2m = dt_halve / m[i];
        for (int d = 0; d < 3; d++) v[3 * i + d] = v[3 * i + d] + f[3 * i + d] * dtInv2m;
        for (int d = 0; d < 3; d++) r[3 * i + d] = r[3 * i + d] + v[3 * i + d] * dt;
    }
}

EXPORT void or_vv_step1(int64_t n, double dt, double* v, const double* f, const double* m) {
    const double dt_halve = dt * 0.5;
    for (int64_t i = 0; i < n; i++) {
        const double dtInv2m = dt_halve / m[i];
        for (int d = 0; d < 3; d++) v[3 * i + d] = v[3 * i + d] + f[3 * i + d] * dtInv2m;
    }
}

/* Sum of m v.v over sites inside [low, high] (pointInBox, math/Geometry.h:14-17)
 * and their count; temperature = factor * (mv2 / (3 n)). */
EXPORT void or_kinetic_sum(int64_t n, const double* r, const double* v, const double* m, const double* low,
                           const double* high, double* mv2_out, double* count_out) {
    double mv2 = 0, cnt = 0;
    for (int64_t i = 0; i < n; i++) {
        const double* p = r + 3 * i;
        if (!(p[0] >= low[0] && p[1] >= low[1] && p[2] >= low[2] && p[0] <= high[0] && p[1] <= high[1] &&
              p[2] <= high[2]))
            continue;
        const double* vel = v + 3 * i;
        double vv = 0;
        for (int d = 0; d < 3; d++) vv += vel[d] * vel[d];
        mv2 += m[i] * vv;
        cnt += 1;
    }
    *mv2_out = mv2;
    *count_out = cnt;
}

EXPORT void or_rescale(int64_t n, double beta, double* v) {
    for (int64_t i = 0; i < 3 * n; i++) v[i] = v[i] * beta;
}

/* ------------------------------------------------------------------------ */
/* Harmonic bonds and angles (no reference counterpart; the reference's only */
/* intramolecular force is FENE, potentials/FENE.cpp:15-43).                 */
/*   bond : U = kb/2 (r - r0)^2                                             */
/*   angle: U = ka/2 (theta - theta0)^2, theta at the middle site           */
/* Displacements use the minimum image. Returns the bonded energy.          */
/* ------------------------------------------------------------------------ */
EXPORT double or_bonded(int64_t nb, const uint32_t* bonds, double kb, double r0, int64_t nang, const uint32_t* angles,
                        double ka, double theta0, const double* r, const double* box, double* f) {
    double u = 0;
    for (int64_t b = 0; b < nb; b++) {
        uint32_t i = bonds[2 * b], j = bonds[2 * b + 1];
        double d[3];
        mi_disp(r + 3 * i, r + 3 * j, box, d); /* d = rj - ri */
        double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double dl = len - r0;
        u += 0.5 * kb * dl * dl;
        double g = -kb * dl / len; /* F_j = g d, F_i = -g d */
        for (int k = 0; k < 3; k++) { f[3 * j + k] += g * d[k]; f[3 * i + k] -= g * d[k]; }
    }
    for (int64_t a = 0; a < nang; a++) {
        uint32_t i = angles[3 * a], j = angles[3 * a + 1], k = angles[3 * a + 2];
        double e1[3], e2[3];
        mi_disp(r + 3 * j, r + 3 * i, box, e1); /* ri - rj */
        mi_disp(r + 3 * j, r + 3 * k, box, e2); /* rk - rj */
        double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        double c = (e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]) / (l1 * l2);
        if (c > 1) c = 1;
        if (c < -1) c = -1;
        double th = acos(c);
        double dth = th - theta0;
        u += 0.5 * ka * dth * dth;
        double s = sqrt(1 - c * c);
        if (s < 1e-12) s = 1e-12;
        double pre = ka * dth / s; /* dU/dc = -ka dth / sin(theta) ; F = -dU/dr = pre * dc/dr */
        double fi[3], fk[3];
        for (int q = 0; q < 3; q++) {
            fi[q] = pre * (e2[q] / (l1 * l2) - c * e1[q] / (l1 * l1));
            fk[q] = pre * (e1[q] / (l1 * l2) - c * e2[q] / (l2 * l2));
            f[3 * i + q] += fi[q];
            f[3 * k + q] += fk[q];
            f[3 * j + q] -= fi[q] + fk[q];
        }
    }
    return u;
}

/* ------------------------------------------------------------------------ */
/* One reference-order MD step (Simulation.cpp:49-74) for the CPU baseline.  */
/* Single-site molecules.  atm: 0 off, 1 reference triplet set (direct,     */
/* ATM_NOLIST), sign as above.  thermostat: T0 > 0 rescales every step.      */
/* ------------------------------------------------------------------------ */
EXPORT int or_run(int64_t n, double* r, double* v, double* f, const double* m, const uint64_t* id, const double* sig,
                  const double* eps, const double* low, const double* high, double cell_size, double cutoff,
                  double dt, double nu, int atm, double atm_sign, double limit_factor, double T0, double temp_factor,
                  int64_t nsteps, int nthreads) {
    int64_t nd[3];
    or_grid_dims(low, high, cell_size, nd);
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    int64_t ncell = na[0] * na[1] * na[2];
    int64_t* mol_start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i <= n; i++) mol_start[i] = i;
    double* com = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    int64_t* site_cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* cell_count = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    int64_t* cell_start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
    int64_t cap = 2 * n + 1;
    int64_t* cell_list = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    for (int64_t s = 0; s < nsteps; s++) {
        or_vv_step0(n, dt, r, v, f, m);
        or_update_com(n, mol_start, r, m, com);
        or_periodic(n, com, r, low, high);
        memset(f, 0, sizeof(double) * 3 * (size_t)n);
        or_update_com(n, mol_start, r, m, com);
        int64_t tot = or_bin(n, mol_start, com, low, high, cell_size, nd, site_cell, cell_count, cell_start, NULL);
        if (tot > cap) { cap = tot; cell_list = (int64_t*)realloc(cell_list, sizeof(int64_t) * (size_t)cap); }
        or_bin(n, mol_start, com, low, high, cell_size, nd, site_cell, cell_count, cell_start, cell_list);
        or_lj_c08(n, r, id, sig, eps, cutoff, low, high, nd, cell_start, cell_list, f, NULL, NULL, nthreads);
        if (limit_factor > 0) or_limit(n, limit_factor, eps, sig, f);
        if (atm) or_atm_c08(n, r, nu, cutoff, low, high, nd, cell_start, cell_list, atm_sign, f, nthreads);
        or_vv_step1(n, dt, v, f, m);
        if (T0 > 0) {
            double mv2, cnt;
            or_kinetic_sum(n, r, v, m, low, high, &mv2, &cnt);
            double T = temp_factor > 0 ? temp_factor * (mv2 / (cnt * 3.0)) : or_temperature_of(mv2, cnt);
            or_rescale(n, sqrt(T0 / T), v);
        }
    }
    free(mol_start); free(com); free(site_cell); free(cell_count); free(cell_start); free(cell_list);
    return 0;
}

EXPORT int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Sensors: virial tensor and radial distribution.                          */
/* ------------------------------------------------------------------------ */

/* w6 += sym(a (x) b), components XX XY YY XZ YZ ZZ (sensors/PressureSensor.h:18). */
static inline void sym_add(double* w6, const double* a, const double* b) {
    w6[0] += a[0] * b[0];
    w6[1] += 0.5 * (a[0] * b[1] + a[1] * b[0]);
    w6[2] += a[1] * b[1];
    w6[3] += 0.5 * (a[0] * b[2] + a[2] * b[0]);
    w6[4] += 0.5 * (a[1] * b[2] + a[2] * b[1]);
    w6[5] += a[2] * b[2];
}

/* Virial tensor W = sum over interactions of sum_members r_m (x) F_m, written as
 * pair terms of minimum-image relative vectors (periodic-safe):
 *   LJ pairs of different molecules within cutoff: d (x) F_i, d = r_i - r_j
 *     (PressureSensor.cpp:152-174);
 *   ATM triplets within cutoff, physical force (F = -grad V), coverage as
 *     or_atm_brute: sum over the three pairs d_ab (x) d_ab v_ab with
 *     v_ab = -comp_force (PressureSensor.cpp:223-287, zz included);
 *   harmonic bonds d (x) F_j, d = r_j - r_i; angles (r_i - r_j) (x) F_i +
 *     (r_k - r_j) (x) F_k.
 * nu = 0 skips the ATM; nb / nang = 0 skip the bonded terms.  Brute-force
 * neighbour search, O(N^2 + N m^2). */
EXPORT void or_virial(int64_t n, const double* r, const uint64_t* id, const double* sig, const double* eps,
                      double cutoff, const double* box, double nu, const int64_t* site_cell, int coverage,
                      int64_t nb, const uint32_t* bonds, double kb, double r0, int64_t nang, const uint32_t* angles,
                      double ka, double theta0, double* w6) {
    const double rc2 = cutoff * cutoff;
    for (int k = 0; k < 6; k++) w6[k] = 0.0;
    int64_t cap = 256;
    int64_t* nbr = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    double* dv = (double*)malloc(sizeof(double) * 3 * (size_t)cap);
    for (int64_t i = 0; i < n; i++) {
        int64_t m = 0;
        for (int64_t j = i + 1; j < n; j++) {
            double d[3];
            mi_disp(r + 3 * i, r + 3 * j, box, d); /* r_j - r_i */
            double q = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (q > rc2) continue;
            if (id[i] != id[j]) {
                double dr[3] = {-d[0], -d[1], -d[2]}; /* r_i - r_j */
                double invdr2 = 1. / q;
                double sigma = (sig[i] + sig[j]) / 2.0;
                double epsilon = sqrt(eps[i] * eps[j]);
                double lj6 = pow(sigma * sigma * invdr2, 3.0);
                double lj12 = pow(lj6, 2.0);
                double fac = 24.0 * epsilon * (2.0 * lj12 - lj6) * invdr2;
                double fi[3] = {dr[0] * fac, dr[1] * fac, dr[2] * fac};
                sym_add(w6, dr, fi);
            }
            if (m == cap) {
                cap *= 2;
                nbr = (int64_t*)realloc(nbr, sizeof(int64_t) * (size_t)cap);
                dv = (double*)realloc(dv, sizeof(double) * 3 * (size_t)cap);
            }
            nbr[m] = j;
            dv[3 * m] = d[0]; dv[3 * m + 1] = d[1]; dv[3 * m + 2] = d[2];
            m++;
        }
        if (nu == 0.0) continue;
        for (int64_t a = 0; a < m; a++)
            for (int64_t b = a + 1; b < m; b++) {
                const double* dij = dv + 3 * a;
                const double* dik = dv + 3 * b;
                double djk[3] = {dik[0] - dij[0], dik[1] - dij[1], dik[2] - dij[2]};
                double qjk = djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2];
                if (qjk > rc2) continue;
                if (coverage == 1) {
                    int64_t ci = site_cell[i], cj = site_cell[nbr[a]], ck = site_cell[nbr[b]];
                    int same = (ci == cj) && (cj == ck);
                    int distinct = (ci != cj) && (cj != ck) && (ci != ck);
                    if (!same && !distinct) continue;
                }
                double qij = dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2];
                double qik = dik[0] * dik[0] + dik[1] * dik[1] + dik[2] * dik[2];
                double ra = sqrt(qij), rb = sqrt(qik), rcn = sqrt(qjk);
                double v01 = -or_atm_comp_force(nu, ra, rb, rcn);
                double v02 = -or_atm_comp_force(nu, rb, ra, rcn);
                double v12 = -or_atm_comp_force(nu, rcn, ra, rb);
                double d01[3], d02[3], d12[3], f01[3], f02[3], f12[3];
                for (int q = 0; q < 3; q++) {
                    d01[q] = -dij[q]; d02[q] = -dik[q]; d12[q] = -djk[q];
                    f01[q] = d01[q] * v01; f02[q] = d02[q] * v02; f12[q] = d12[q] * v12;
                }
                sym_add(w6, d01, f01);
                sym_add(w6, d02, f02);
                sym_add(w6, d12, f12);
            }
    }
    free(nbr);
    free(dv);
    for (int64_t t = 0; t < nb; t++) {
        uint32_t i = bonds[2 * t], j = bonds[2 * t + 1];
        double d[3];
        mi_disp(r + 3 * i, r + 3 * j, box, d);
        double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double g = -kb * (len - r0) / len;
        double fj[3] = {g * d[0], g * d[1], g * d[2]};
        sym_add(w6, d, fj);
    }
    for (int64_t t = 0; t < nang; t++) {
        uint32_t i = angles[3 * t], j = angles[3 * t + 1], k = angles[3 * t + 2];
        double e1[3], e2[3];
        mi_disp(r + 3 * j, r + 3 * i, box, e1);
        mi_disp(r + 3 * j, r + 3 * k, box, e2);
        double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        double l2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        double c = (e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]) / (l1 * l2);
        if (c > 1) c = 1;
        if (c < -1) c = -1;
        double dth = acos(c) - theta0;
        double s = sqrt(1 - c * c);
        if (s < 1e-12) s = 1e-12;
        double pre = ka * dth / s;
        double fi[3], fk[3];
        for (int q = 0; q < 3; q++) {
            fi[q] = pre * (e2[q] / (l1 * l2) - c * e1[q] / (l1 * l1));
            fk[q] = pre * (e1[q] / (l1 * l2) - c * e2[q] / (l2 * l2));
        }
        sym_add(w6, e1, fi);
        sym_add(w6, e2, fk);
    }
}

/* RDFSensor::RDF_Kernel over every ordered pair of distinct molecules
 * (RDFSensor.cpp:59-72): delta = c0 - c1, dr = delta - L round(delta / L) per
 * axis, r = sqrt(x^2 + y^2 + z^2) summed in that order (Array::L2), skipped
 * when r > max_r, else one count in clamp((uint64)(r / dr), 0, nbins - 1)
 * (getBin, RDFSensor.h:43-48).  com: molecule centres, 3 per molecule. */
EXPORT void or_rdf(int64_t m, const double* com, const double* box, double max_r, double delta_r, int64_t nbins,
                   uint64_t* counts, int nthreads) {
    for (int64_t b = 0; b < nbins; b++) counts[b] = 0;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        uint64_t* loc = (uint64_t*)calloc((size_t)nbins, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = 0; i < m; i++)
            for (int64_t j = 0; j < m; j++) {
                if (i == j) continue;
                double d[3];
                for (int q = 0; q < 3; q++) {
                    double delta = com[3 * i + q] - com[3 * j + q];
                    d[q] = delta - box[q] * round(delta / box[q]);
                }
                double s = 0;
                for (int q = 0; q < 3; q++) s += d[q] * d[q];
                double r = sqrt(s);
                if (r > max_r) continue;
                uint64_t bin = (uint64_t)(r / delta_r);
                if (bin > (uint64_t)(nbins - 1)) bin = (uint64_t)(nbins - 1);
                loc[bin]++;
            }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int64_t b = 0; b < nbins; b++) counts[b] += loc[b];
        free(loc);
    }
}

/* ------------------------------------------------------------------ tabulated pair force (IBI) */

/* FunctionPL::evaluateAt (plugins/IBI/IBI_Math.h:30-41) with findUpperNode (:59-67) and
 * linearInterpolation (:49-56): def_high above the last node, def_low below the first,
 * otherwise ya + (yb - ya) / (xb - xa) * (x - xa) on the segment whose upper node is the
 * first node >= x.  At x equal to the first node the reference's lower index is -1 (a
 * read before the table); this restatement returns the first node's value there, the
 * limit of the first segment. */
EXPORT double or_table_eval(double x, int64_t n, const double* xv, const double* yv, double def_low, double def_high) {
    if (x > xv[n - 1]) return def_high;
    if (x < xv[0]) return def_low;
    int64_t i = 0;
    while (x > xv[i]) {
        i++;
        if (i == n) break;
    }
    if (i == 0) return yv[0];
    const double ya = yv[i - 1], yb = yv[i], xa = xv[i - 1], xb = xv[i];
    return ya + (yb - ya) / (xb - xa) * (x - xa);
}

typedef void (*pair_fn)(int64_t i0, int64_t i1, const double* s0, const double* s1, void* ctx);

/* The reference's pair set (the walk of or_lj_c08: 13 C08 cell pairs per base cell, then
 * the intra-cell pairs of domain cells), one callback per candidate. */
static void c08_pair_walk(const int64_t* nd, const double* low, const double* high, const int64_t* cell_start,
                          const int64_t* cell_list, pair_fn fn, void* ctx, int nthreads) {
    int64_t na[3] = {nd[0] + 2, nd[1] + 2, nd[2] + 2};
    double dom[3] = {high[0] - low[0], high[1] - low[1], high[2] - low[2]};
    int64_t nb = c08_bases(na, NULL);
    int64_t* bases = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)nb);
    c08_bases(na, bases);
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t b = 0; b < nb; b++) {
        for (int o = 0; o < 13; o++) {
            int64_t c0[3], c1[3];
            for (int d = 0; d < 3; d++) {
                c0[d] = bases[3 * b + d] + PAIR_OFF[o][0][d];
                c1[d] = bases[3 * b + d] + PAIR_OFF[o][1][d];
            }
            double s0[3], s1[3];
            cell_shift(c0, na, dom, s0);
            cell_shift(c1, na, dom, s1);
            if (nonzero3(s0) && nonzero3(s1)) continue; /* checkState LinkedCells.cpp:471-474 */
            int64_t f0 = c0[0] + na[0] * (c0[1] + na[1] * c0[2]);
            int64_t f1 = c1[0] + na[0] * (c1[1] + na[1] * c1[2]);
            int64_t n0 = cell_start[f0 + 1] - cell_start[f0], n1 = cell_start[f1 + 1] - cell_start[f1];
            for (int64_t local = 0; local < n0 * n1; local++) {
                int64_t i = local % n0, j = local / n0;
                fn(cell_list[cell_start[f0] + i], cell_list[cell_start[f1] + j], s0, s1, ctx);
            }
        }
    }
    int64_t ncell = na[0] * na[1] * na[2];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
    for (int64_t c = 0; c < ncell; c++) {
        int64_t x = c % na[0], y = (c / na[0]) % na[1], z = c / (na[0] * na[1]);
        if (x == 0 || y == 0 || z == 0 || x == na[0] - 1 || y == na[1] - 1 || z == na[2] - 1) continue;
        int64_t nc = cell_start[c + 1] - cell_start[c];
        const double zero[3] = {0, 0, 0};
        for (int64_t local = 0; local < nc * (nc - 1) / 2; local++) {
            int64_t i = (int64_t)floor((-1 + sqrt(1 + 8 * (double)local)) / 2.0);
            int64_t j = local - i * (i + 1) / 2;
            i += 1;
            fn(cell_list[cell_start[c] + i], cell_list[cell_start[c] + j], zero, zero, ctx);
        }
    }
    free(bases);
}

typedef struct {
    const double* r;
    double cutoff2;
    int64_t nodes;
    const double *x, *y, *u; /* force table, optional potential table on the same nodes */
    double y_low, y_high, u_low, u_high;
    double* f;
    double* u_total; /* optional */
    double* vir;     /* optional, [6] XX XY YY XZ YZ ZZ */
    int atomic;
} table_ctx;

static inline void add_to(double* p, double v, int atomic) {
    if (atomic) {
#pragma omp atomic
        *p += v;
    } else *p += v;
}

/* IBI_Force::operator() (plugins/IBI/IBI_ForceFunctor.cpp:24-42) on one candidate pair:
 * dr = (r0 + s0) - (r1 + s1), skipped when |dr|^2 > cutoff^2, F = evaluateAt(|dr|),
 * force = (dr / |dr|) F added to the unshifted first member and subtracted from the
 * unshifted second.  No molecule exclusion (the functor has none).  Energy and virial
 * (not in the reference functor) count each physical pair once, as lj_pair does. */
static void table_pair(int64_t i0, int64_t i1, const double* s0, const double* s1, void* p) {
    table_ctx* t = (table_ctx*)p;
    double dr[3];
    for (int d = 0; d < 3; d++) dr[d] = (t->r[3 * i0 + d] + s0[d]) - (t->r[3 * i1 + d] + s1[d]);
    double dr2 = 0;
    for (int d = 0; d < 3; d++) dr2 += dr[d] * dr[d];
    if (dr2 > t->cutoff2) return;
    const double rr = sqrt(dr2);
    const double F = or_table_eval(rr, t->nodes, t->x, t->y, t->y_low, t->y_high);
    double force[3];
    for (int d = 0; d < 3; d++) force[d] = (dr[d] / rr) * F;
    const int u0 = !nonzero3(s0), u1 = !nonzero3(s1);
    for (int d = 0; d < 3; d++) {
        if (u0) add_to(&t->f[3 * i0 + d], force[d], t->atomic);
        if (u1) add_to(&t->f[3 * i1 + d], -force[d], t->atomic);
    }
    if (!u1) return;
    if (t->u_total && t->u)
        add_to(t->u_total, or_table_eval(rr, t->nodes, t->x, t->u, t->u_low, t->u_high), t->atomic);
    if (t->vir) {
        const double w[6] = {dr[0] * force[0], dr[0] * force[1], dr[1] * force[1],
                             dr[0] * force[2], dr[1] * force[2], dr[2] * force[2]};
        for (int k = 0; k < 6; k++) add_to(&t->vir[k], w[k], t->atomic);
    }
}

/* The tabulated pair force over the reference's pair set; f accumulated. */
EXPORT void or_table_c08(int64_t n, const double* r, double cutoff, const double* low, const double* high,
                         const int64_t* nd, const int64_t* cell_start, const int64_t* cell_list, int64_t nodes,
                         const double* x, const double* y, double y_low, double y_high, const double* u, double u_low,
                         double u_high, double* f, double* u_total, double* vir, int nthreads) {
    (void)n;
    table_ctx t = {r, pow(cutoff, 2), nodes, x, y, u, y_low, y_high, u_low, u_high, f, u_total, vir, nthreads > 1};
    c08_pair_walk(nd, low, high, cell_start, cell_list, table_pair, &t, nthreads);
}

/* Minimum-image brute force of the same force (every pair within the cutoff, no
 * exclusion), FP64: the independent pin of the walk on small systems. */
EXPORT void or_table_brute(int64_t n, const double* r, double cutoff, const double* box, int64_t nodes,
                           const double* x, const double* y, double y_low, double y_high, double* f) {
    const double rc2 = cutoff * cutoff;
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = i + 1; j < n; j++) {
            double d[3], d2 = 0;
            for (int k = 0; k < 3; k++) {
                d[k] = r[3 * i + k] - r[3 * j + k];
                d[k] -= box[k] * nearbyint(d[k] / box[k]);
                d2 += d[k] * d[k];
            }
            if (d2 > rc2) continue;
            const double rr = sqrt(d2);
            const double F = or_table_eval(rr, nodes, x, y, y_low, y_high);
            for (int k = 0; k < 3; k++) {
                f[3 * i + k] += d[k] / rr * F;
                f[3 * j + k] -= d[k] / rr * F;
            }
        }
}
