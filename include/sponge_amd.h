/*
 * sponge_amd.h -- C ABI of the MI355X (gfx950) nonbonded hot path:
 * cell-grid neighbor list, LJ + PME direct-space pair forces with exclusions,
 * and the PME reciprocal part (order-4 B-spline spread, 3D FFT convolution,
 * force gather).
 *
 * Every entry point mirrors one method of the reference's modules and keeps
 * its argument meaning:
 *   NEIGHBOR_LIST              SPONGE/neighbor_list/neighbor_list.h:7-127
 *   LENNARD_JONES_INFORMATION  SPONGE/Lennard_Jones_force/Lennard_Jones_force.h:87-173
 *   Particle_Mesh              SPONGE/PM_force/PM_force.h:152-209
 *
 * Conventions (those of the reference):
 *   - VECTOR arrays are float[3*n] AoS (x, y, z); LTMatrix3 is float[6] =
 *     {a11, a21, a22, a31, a32, a33}; the box is passed as the two LTMatrix3 of
 *     the reference's Boundary (cell and rcell), in HOST memory.
 *   - Per-atom arrays (crd, frc, charge, energies, virials, exclusion CSR,
 *     atom_local) are DEVICE pointers owned by the caller.
 *   - Force, energy and virial outputs are ACCUMULATED (+=) into caller-zeroed
 *     arrays; energies only when need_energy, virials only when need_virial.
 *   - Every call is asynchronous on the stream given at *_initial (NULL =
 *     the default stream) unless it says it synchronises.
 *   - Return value: 0 on success; a HIP runtime error code (< 1000) when a
 *     runtime call fails; otherwise a SPONGE error code (1001..1010,
 *     SPONGE/utils/control/error.hpp:5-29).  Nothing calls exit().
 */
#ifndef SPONGE_AMD_H
#define SPONGE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum sponge_amd_status {
    SPONGE_AMD_SUCCESS = 0,
    SPONGE_AMD_ERROR_NOT_IMPLEMENTED = 1001,
    SPONGE_AMD_ERROR_BAD_FILE_FORMAT = 1002,
    SPONGE_AMD_ERROR_CONFLICTING_COMMAND = 1003,
    SPONGE_AMD_ERROR_MISSING_COMMAND = 1004,
    SPONGE_AMD_ERROR_TYPE_ERROR_COMMAND = 1005,
    SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND = 1006,
    SPONGE_AMD_ERROR_SIMULATION_BREAKDOWN = 1007,
    SPONGE_AMD_ERROR_MALLOC_FAILED = 1008,
    SPONGE_AMD_ERROR_OVERFLOW = 1009,
    SPONGE_AMD_ERROR_OPEN_FILE_FAILED = 1010
};

/* Update modes of NEIGHBOR_LIST::Update (neighbor_list.h:112-116). */
enum sponge_amd_nl_update { SPONGE_AMD_NL_CONDITIONAL = 0, SPONGE_AMD_NL_FORCED = 1 };

/* Overflow flags reported by sponge_nl_check_overflow. */
enum sponge_amd_nl_flag {
    SPONGE_AMD_NL_GRID_OVERFLOW = 1,  /* a cell held more than max_atom_in_grid_numbers */
    SPONGE_AMD_NL_LIST_OVERFLOW = 2,  /* an atom had more than 2 * max_neighbor_numbers
                                         partners in both directions: its row was cut */
    SPONGE_AMD_NL_STREAM_OVERFLOW = 4 /* a cell's stencil held more than 65535 atoms: the
                                         rows of that cell were left empty (no limit of the
                                         options can lift this; an error in every mode) */
};

typedef struct sponge_nl sponge_nl;
typedef struct sponge_lj sponge_lj;
typedef struct sponge_pme sponge_pme;

/* ------------------------------------------------------------------------- */
/* Library                                                                   */
/* ------------------------------------------------------------------------- */

/* Version string of the library. */
const char* sponge_amd_version(void);

/* Ewald splitting parameter: bisection on erfc(beta*cutoff)/cutoff = tolerance
 * (Get_Beta, SPONGE/PM_force/PM_force.cpp:246-274). Host only. */
float sponge_amd_get_beta(float cutoff, float tolerance);

/* Automatic FFT dimension for a box length in grid units
 * (Get_Fft_Patameter, SPONGE/common.cpp:639-659). Host only. */
int sponge_amd_get_fft_parameter(float length);

/* Per-kernel event timers (no counterpart call in the reference, whose
 * TIME_RECORDER, SPONGE/utils/control/time_recorder.hpp:10-42, times host
 * regions): when enabled, HIP events are recorded on the launching stream
 * around the hot kernels ("lj_pairs", "nl_search", "pme_spread",
 * "pme_gather", "pme_fft_z_r2c", "pme_fft_y", "pme_fft_x_conv",
 * "pme_ifft_y", "pme_fft_z_c2r").  sponge_amd_kernel_timing synchronises
 * with the recorded events and returns the accumulated milliseconds and
 * launch count of one timer.  Timers also run inside the NVE driver's HIP
 * graph: each replay records fresh events.
 * sponge_amd_set_kernel_timing_filter limits the timers to a comma-separated
 * list of names (NULL or "" = all), so that only the kernel being measured
 * carries events. */
int sponge_amd_set_kernel_timing(int enable);
int sponge_amd_set_kernel_timing_filter(const char* names);
int sponge_amd_kernel_timing(const char* name, double* total_ms, long long* launches);
int sponge_amd_reset_kernel_timing(void);

/* ------------------------------------------------------------------------- */
/* Neighbor list: NEIGHBOR_LIST (SPONGE/neighbor_list/neighbor_list.h:7-127)  */
/* ------------------------------------------------------------------------- */

/* The [neighbor_list] mdin keys read by NEIGHBOR_LIST::Initial
 * (neighbor_list.cpp:271-300, 871-972). */
typedef struct sponge_nl_options {
    int refresh_interval;         /* 0: rebuild when any atom moved > skin*skin_permit */
    float skin_permit;            /* 0.5 */
    int max_neighbor_numbers;     /* 1200 */
    int max_atom_in_grid_numbers; /* 150 (reported as a flag; never drops atoms) */
    int throw_error_when_overflow;/* 0 */
    int check_overflow_interval;  /* 150 */
} sponge_nl_options;

void sponge_nl_default_options(sponge_nl_options* opt);

/* NEIGHBOR_LIST::Initial(controller, atom_numbers, cutoff, skin, boundary)
 * (neighbor_list.cpp:871-972).  Fails with VALUE_ERROR_COMMAND when the cell
 * grid has 2 or fewer cells along an axis (the reference's "box is too
 * small", :949-954, raised there as MallocFailed). */
int sponge_nl_initial(sponge_nl** out, int atom_numbers, float cutoff, float skin,
                      const float* cell, const float* rcell, const sponge_nl_options* opt,
                      void* stream);

/* NEIGHBOR_LIST::Update(atom_local, local_atom_numbers, ghost_numbers, crd,
 * boundary, step, update, excluded_list_start, excluded_list,
 * excluded_numbers) (neighbor_list.cpp:974-1025).  atom_local maps local to
 * global ids (NULL = identity); the exclusion CSR is indexed by global id and
 * holds, per atom, ascending global ids j > i.  One GPU holds no ghost atoms:
 * ghost_numbers must be 0. */
int sponge_nl_update(sponge_nl* nl, const int* atom_local, int local_atom_numbers,
                     int ghost_numbers, const float* crd, const float* cell, const float* rcell,
                     int step, int update, const int* excluded_list_start,
                     const int* excluded_list, const int* excluded_numbers);

/* NEIGHBOR_LIST::Check_Overflow (neighbor_list.cpp:1027-1103), reporting
 * instead of re-initialising: synchronises the stream, returns the sticky
 * flags (sponge_amd_nl_flag bits) and clears them.  Returns OVERFLOW when a
 * flag was set and throw_error_when_overflow is on, and always on
 * SPONGE_AMD_NL_STREAM_OVERFLOW.  The reference's recovery (re-initialising
 * with 1.1x the limit that overflowed) is the caller's: the NVE driver and the
 * C++ facade do it. */
int sponge_nl_check_overflow(sponge_nl* nl, int* flags);

/* The options the list was initialised with (the public fields
 * max_neighbor_numbers, max_atom_in_grid_numbers, ... of NEIGHBOR_LIST,
 * neighbor_list.h:7-127). */
int sponge_nl_get_options(sponge_nl* nl, sponge_nl_options* opt);

/* The half list in the reference's representation: for local atom i,
 * counts[i] neighbours in serial[i*max_numbers ...] (local ids, j > i by
 * global id, exclusions removed), as ATOM_GROUP d_nl holds it
 * (SPONGE/common.h:177-182).  Device arrays; entry order within an atom is
 * unspecified, as in the reference. */
int sponge_nl_export_half_list(sponge_nl* nl, int* counts, int* serial, int max_numbers);

/* Fills and returns (in *d_nl) a device array of ATOM_GROUP {int atom_numbers;
 * int ghost_numbers; int* atom_serial;} for modules that read
 * NEIGHBOR_LIST::d_nl directly.  The storage is owned by the list. */
int sponge_nl_atom_group(sponge_nl* nl, void** d_nl);

/* The address of that ATOM_GROUP array without filling it (allocated on
 * first use, N x 16 bytes, stable until sponge_nl_clear): the value of
 * NEIGHBOR_LIST::d_nl (neighbor_list.h:42) that callers pass to the force
 * modules as the list's name, as SPONGE/main.cpp:1452-1470 does. */
int sponge_nl_atom_group_address(sponge_nl* nl, void** d_nl);

/* Statistics of the current list: entries and the cell grid dims. */
int sponge_nl_info(sponge_nl* nl, long long* list_entries, int* grid_dims);

/* Number of list builds since Initial (synchronises). */
int sponge_nl_rebuild_count(sponge_nl* nl, long long* count);

/* NEIGHBOR_LIST::Clear (neighbor_list.cpp:1105-1119). */
int sponge_nl_clear(sponge_nl* nl);

/* ------------------------------------------------------------------------- */
/* LJ + PME direct: LENNARD_JONES_INFORMATION (Lennard_Jones_force.h:87-173)  */
/* ------------------------------------------------------------------------- */

/* LENNARD_JONES_INFORMATION::Initial (Lennard_Jones_force.cpp:147-250).
 * LJ_A / LJ_B are the in-memory tables (file values x12 and x6, lj.hpp:47,57)
 * of size T(T+1)/2; atom_LJ_type has atom_numbers entries.  Host arrays.
 * At most 64 LJ types (the pair kernel holds the T x T table in LDS);
 * more return SPONGE_AMD_ERROR_NOT_IMPLEMENTED. */
int sponge_lj_initial(sponge_lj** out, int atom_numbers, int atom_type_numbers,
                      const float* LJ_A, const float* LJ_B, const int* atom_LJ_type, float cutoff,
                      void* stream);

/* long_range_factor of Initial (Lennard_Jones_force.cpp:208-235). */
int sponge_lj_long_range_factor(sponge_lj* lj, float* factor);

/* LJ_PME_Direct_Force_With_Atom_Energy_And_Virial (Lennard_Jones_force.cpp:
 * 326-384) over the list held by nl.  frc, atom_energy, atom_virial are
 * accumulated; atom_direct_pme_energy (PM's d_direct_atom_energy) and the
 * module's own LJ energy array are overwritten when need_atom_energy.
 * With solvent_numbers > 0 the last solvent_numbers atoms are the solvent
 * tail (Lennard_Jones_force.cpp:55-58): only the pairs with a member below it
 * are computed here, the tail's own pairs by sponge_solvent_lj_direct_force. */
int sponge_lj_pme_direct_force(sponge_lj* lj, sponge_nl* nl, int atom_numbers,
                               int local_atom_numbers, int solvent_numbers, int ghost_numbers,
                               const float* crd, const float* charge, float* frc,
                               const float* cell, const float* rcell, float pme_beta,
                               int need_atom_energy, float* atom_energy, int need_virial,
                               float* atom_virial, float* atom_direct_pme_energy);

/* The same force call over a caller-provided half list in the reference's
 * representation (counts[N], serial[N*max_numbers] of local ids), so the pair
 * kernel can be checked on exactly the reference's pairs. */
int sponge_lj_pme_direct_force_list(sponge_lj* lj, const int* counts, const int* serial,
                                    int max_numbers, int atom_numbers, const float* crd,
                                    const float* charge, float* frc, const float* cell,
                                    const float* rcell, float pme_beta, int need_atom_energy,
                                    float* atom_energy, int need_virial, float* atom_virial,
                                    float* atom_direct_pme_energy);

/* Allocation-free form of sponge_lj_pme_direct_force_list.  It has no
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No module handle: the LJ tables are the caller's device
 * arrays (d_LJ_A, d_LJ_B of T(T+1)/2 entries in the in-memory convention,
 * d_atom_LJ_type of atom_numbers entries), and the pair kernel's working
 * arrays live in a fixed device array of the library (one per GPU, ~17 MB,
 * held until the process ends).  The call allocates nothing and frees nothing.
 * Outputs as sponge_lj_pme_direct_force_list: frc, atom_energy, atom_virial
 * accumulated; atom_LJ_energy and atom_direct_pme_energy (either may be NULL)
 * overwritten when need_atom_energy.  Limits: atom_numbers <=
 * SPONGE_AMD_NOALLOC_MAX_ATOMS and atom_type_numbers <=
 * SPONGE_AMD_NOALLOC_MAX_TYPES, else SPONGE_AMD_ERROR_OVERFLOW.  Every call
 * on one GPU shares that GPU's working array: calls that overlap on one GPU
 * (several streams or host threads) race and must be serialised by the
 * caller.  Asynchronous on `stream` (NULL = the default stream). */
#define SPONGE_AMD_NOALLOC_MAX_ATOMS 262144
#define SPONGE_AMD_NOALLOC_MAX_TYPES 64
int sponge_lj_pme_direct_force_list_noalloc(
    int atom_numbers, int atom_type_numbers, const float* d_LJ_A, const float* d_LJ_B,
    const int* d_atom_LJ_type, float cutoff, const int* counts, const int* serial,
    int max_numbers, const float* crd, const float* charge, float* frc, const float* cell,
    const float* rcell, float pme_beta, int need_atom_energy, float* atom_energy, int need_virial,
    float* atom_virial, float* atom_LJ_energy, float* atom_direct_pme_energy, void* stream);

/* Long_Range_Correction (Lennard_Jones_force.cpp:287-308): d_potential[0] +=
 * factor/volume; diagonal of d_virial[0] += 2*factor/volume. */
int sponge_lj_long_range_correction(sponge_lj* lj, int need_pressure, float* d_virial,
                                    int need_potential, float* d_potential, float volume);

/* Step_Print (Lennard_Jones_force.cpp:386-400): synchronises; LJ_short is the
 * sum of the per-atom LJ energies (double accumulation), LJ_long the last
 * long-range term added. */
int sponge_lj_step_print(sponge_lj* lj, float* LJ_short, float* LJ_long);

/* Device pointer of the per-atom LJ energy array (d_LJ_energy_atom). */
float* sponge_lj_atom_energy(sponge_lj* lj);

int sponge_lj_clear(sponge_lj* lj);

/* ------------------------------------------------------------------------- */
/* solvent_LJ: SOLVENT_LENNARD_JONES                                          */
/* (SPONGE/Lennard_Jones_force/solvent_LJ.h:9-50)                             */
/* ------------------------------------------------------------------------- */

typedef struct sponge_solvent_lj sponge_solvent_lj;

/* SOLVENT_LENNARD_JONES::Initial(controller, lj, lj_soft, md_info,
 * default_enable) (solvent_LJ.cpp:149-248) over the LJ module `lj`.
 * res_start (residue_numbers entries) and mass (atom_numbers) are host
 * arrays; residues are contiguous atom ranges (the reference's update groups
 * of a water / ion system).  solvent_LJ is the mdin key: -1 when absent (then
 * on when default_enable -- the reference passes mode >= NVT -- and there are
 * at least 10 residues), 0 false, 1 true.  The water size is the last
 * residue's (3 or 4 atoms, else off); the solvent is the run of trailing
 * residues of that size whose first three atoms have water masses (O in
 * (15.9, 16.1), H in (1.007, 1.009)).  A module that is off computes nothing,
 * as the reference's. */
int sponge_solvent_lj_initial(sponge_solvent_lj** out, sponge_lj* lj, int atom_numbers,
                              int residue_numbers, const int* res_start, const float* mass,
                              int solvent_LJ, int default_enable, void* stream);

/* The public fields is_initialized, solvent_numbers, solvent_start and
 * water_points (solvent_LJ.h:13-17). */
int sponge_solvent_lj_info(sponge_solvent_lj* h, int* enabled, int* solvent_numbers,
                           int* solvent_start_residue, int* water_points);

/* SOLVENT_LENNARD_JONES::LJ_PME_Direct_Force_With_Atom_Energy_And_Virial(
 * atom_numbers, residue_numbers, d_res_start, crd, charge, frc, boundary, nl,
 * pme_beta, need_atom_energy, atom_energy, need_virial, atom_lj_virial,
 * atom_direct_pme_energy) (solvent_LJ.cpp:24-147, 256-432) over the list
 * held by nl: for every water w of the tail, the pairs of each of its atoms
 * with every partner of its OXYGEN's half list (the hydrogens' own lists are
 * not used), strict r < cutoff; frc, atom_energy, atom_virial, the LJ
 * module's per-atom LJ energy and atom_direct_pme_energy accumulated, pair
 * energies and virials on the oxygen.  Call after sponge_lj_pme_direct_force
 * with solvent_numbers = this module's solvent_numbers, which computes every
 * pair with a non-solvent member. */
int sponge_solvent_lj_direct_force(sponge_solvent_lj* h, sponge_nl* nl, int atom_numbers,
                                   int residue_numbers, const float* crd, const float* charge,
                                   float* frc, const float* cell, const float* rcell,
                                   float pme_beta, int need_atom_energy, float* atom_energy,
                                   int need_virial, float* atom_virial,
                                   float* atom_direct_pme_energy);

int sponge_solvent_lj_clear(sponge_solvent_lj* h);

/* ------------------------------------------------------------------------- */
/* PME: Particle_Mesh (SPONGE/PM_force/PM_force.h:152-209)                    */
/* ------------------------------------------------------------------------- */

/* The [PM] / [PME] mdin keys read by Particle_Mesh::Initial
 * (PM_force.cpp:315-502). fft* <= 0 selects the automatic size. */
typedef struct sponge_pme_options {
    int fftx, ffty, fftz;          /* -1 */
    float grid_spacing;            /* 1.0 Angstrom */
    float direct_tolerance;        /* 1e-5 */
    int update_interval;           /* 1 */
    int calculate_reciprocal_part; /* 1 */
    int calculate_excluded_part;   /* 1 */
} sponge_pme_options;

void sponge_pme_default_options(sponge_pme_options* opt);

/* Particle_Mesh::Initial(controller, atom_numbers, boundary, box_length,
 * cutoff, ...) (PM_force.cpp:295-654), any lower-triangular cell (box_length
 * for the grid sizing = the lengths of the cell vectors). FFT sizes must be
 * 2,3,5,7-smooth and at most 512. */
int sponge_pme_initial(sponge_pme** out, int atom_numbers, const float* cell, const float* rcell,
                       float cutoff, const sponge_pme_options* opt, void* stream);

/* beta and the grid: out_beta, out_fft[3]. */
int sponge_pme_info(sponge_pme* pme, float* beta, int* fft);

/* Device pointer of d_direct_atom_energy (written by the LJ direct call). */
float* sponge_pme_direct_atom_energy(sponge_pme* pme);

/* PME_Reciprocal_Force_With_Energy_And_Virial(crd, boundary, charge, force,
 * need_virial, need_energy, d_virial, d_potential, step)
 * (PM_force.cpp:1070-1168): on steps with step % update_interval == 0, frc +=
 * update_interval * F_reciprocal; when need_energy, d_potential[0] +=
 * E_reciprocal + E_self + E_neutralizing; when need_virial, d_virial[0] +=
 * the reciprocal virial. */
int sponge_pme_reciprocal_force(sponge_pme* pme, const float* crd, const float* cell,
                                const float* rcell, const float* charge, float* frc,
                                int need_virial, int need_energy, float* d_virial,
                                float* d_potential, int step);

/* MPI_PME_Excluded_Force_With_Atom_Energy(local_atom_numbers, atom_local,
 * atom_local_id, crd, boundary, charge, excluded_list_start, excluded_list,
 * excluded_numbers, frc, need_energy, atom_energy, need_virial, atom_virial)
 * (PM_force.cpp:1745-1842); the exclusion CSR is indexed by local id. */
int sponge_pme_excluded_force(sponge_pme* pme, int local_atom_numbers, const int* atom_local,
                              const int* atom_local_id, const float* crd, const float* cell,
                              const float* rcell, const float* charge,
                              const int* excluded_list_start, const int* excluded_list,
                              const int* excluded_numbers, float* frc, int need_energy,
                              float* atom_energy, int need_virial, float* atom_virial);

/* Step_Print (PM_force.cpp:1844-1869): synchronises; PM = direct + reciprocal
 * + self + correction of the last energy step. */
int sponge_pme_step_print(sponge_pme* pme, float* PM, float* direct, float* reciprocal,
                          float* self_energy, float* correction);

/* The real-space grids of the last reciprocal step, [fftx][ffty][fftz]
 * (z fastest): Q (spread charges) and FBCFQ (convolved potential).  Device
 * arrays of fftx*ffty*fftz floats; either may be NULL. */
int sponge_pme_grids(sponge_pme* pme, float* Q, float* FBCFQ);

/* Particle_Mesh::Update_Box (PM_force.cpp:1240-1256, up_box_bc :1171-1238):
 * the BC / virial_BC tables and the neutralizing factor for a new
 * box on the same grid, as a barostat step needs (host
 * LTMatrix3 pair).  Synchronises.  The neighbor list is re-initialized on a
 * box change, as the reference does (SPONGE/main.cpp:2104-2110). */
int sponge_pme_update_box(sponge_pme* pme, const float* cell, const float* rcell);

/* Autotuner of the FFT passes (no counterpart in the reference, which calls
 * cuFFT / FFTW plans): lines per workgroup of the z, y and x passes, each the
 * fastest of {2, 4, 8, 16} by event timing of `reps` reciprocal-force calls on
 * crd / charge (device arrays) on the module's stream; chosen[3] (host, may
 * be NULL) gets the choice.  Forces do not depend on it.  Synchronises. */
int sponge_pme_autotune(sponge_pme* pme, const float* crd, const float* charge, int reps,
                        int* chosen);

/* Sets the FFT tiling (host int[3]: z, y, x lines per workgroup; even, 2..16). */
int sponge_pme_set_fft_tiling(sponge_pme* pme, const int* tiles);

int sponge_pme_clear(sponge_pme* pme);

/* ------------------------------------------------------------------------- */
/* LJ soft core: LJ_SOFT_CORE (SPONGE/Lennard_Jones_force/LJ_soft_core.h)      */
/* ------------------------------------------------------------------------- */

typedef struct sponge_ljsc sponge_ljsc;

/* The mdin keys read by LJ_SOFT_CORE::Initial (LJ_soft_core.cpp:490-560):
 * lambda_lj, soft_core_alpha (0.5), soft_core_powfer (1.0), soft_core_sigma
 * (3.0 A), soft_core_sigma_min (0.0). */
typedef struct sponge_ljsc_options {
    float lambda;
    float alpha;
    float p;
    float sigma;
    float sigma_min;
} sponge_ljsc_options;

void sponge_ljsc_default_options(sponge_ljsc_options* opt);

/* LJ_SOFT_CORE::Initial: the A- and B-state tables (lower triangles, already
 * scaled by 12 and 6 as Native_Load_LJ_Soft_Core leaves them) and per-atom
 * types, host arrays; also the long-range factors of the lambda-mixed and the
 * B - A C6 sums (LJ_soft_core.cpp:640-668).  At most 64 types per state. */
int sponge_ljsc_initial(sponge_ljsc** out, int atom_numbers, int type_numbers_A,
                        int type_numbers_B, const float* LJ_AA, const float* LJ_AB,
                        const float* LJ_BA, const float* LJ_BB, const int* atom_type_A,
                        const int* atom_type_B, float cutoff, const sponge_ljsc_options* opt,
                        void* stream);

int sponge_ljsc_long_range_factors(sponge_ljsc* h, float* factor, float* factor_TI);

/* LJ_Soft_Core_PME_Direct_Force_With_Atom_Energy_And_Virial
 * (LJ_soft_core.cpp:757-826): frc += soft-core LJ + direct Coulomb forces over
 * the list; with need_atom_energy, atom_energy += LJ + Coulomb per atom and
 * atom_direct_pme_energy (may be NULL) = the direct Coulomb part; with
 * need_virial, atom_virial += the pair virial.  Device arrays. */
int sponge_ljsc_direct_force(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
                             float* frc, const float* cell, const float* rcell, float pme_beta,
                             int need_atom_energy, float* atom_energy, int need_virial,
                             float* atom_virial, float* atom_direct_pme_energy);

/* Get_Partial_H_Partial_Lambda_With_Columb_Direct (LJ_soft_core.cpp:828-880):
 * dU/dlambda of the LJ part (with the long-range TI term) and of the direct
 * Coulomb part (charge_BA = charge_B - charge_A, device, may be NULL);
 * host outputs; synchronises. */
int sponge_ljsc_du_dlambda(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
                           const float* charge_BA, const float* cell, const float* rcell,
                           float pme_beta, float* du_lj, float* du_direct);

int sponge_ljsc_long_range_correction(sponge_ljsc* h, int need_pressure, float* d_virial,
                                      int need_potential, float* d_potential, float volume);

/* LJ_soft_short of the last energy call (sum of the per-atom LJ energies);
 * synchronises. */
int sponge_ljsc_step_print(sponge_ljsc* h, float* LJ_soft_short);

int sponge_ljsc_clear(sponge_ljsc* h);

/* ------------------------------------------------------------------------- */
/* Learned reciprocal-force surrogate                                         */
/* ------------------------------------------------------------------------- */

/* A per-atom MLP (48 -> 64 -> 64 -> 3, SiLU, BF16 MFMA with FP32
 * accumulation) over charge-environment features of the neighbor-list
 * partners within feature_cutoff predicts the PME reciprocal force as
 * q_i (c_0 U_0 + c_1 U_1 + c_2 U_2): U_0 the reciprocal pair field of the
 * list partners (smooth cutoff), U_1 that of the excluded partners, U_2 the
 * charge dipole of the environment.  No counterpart in the reference: it
 * stands beside Particle_Mesh::PME_Reciprocal_Force_With_Energy_And_Virial
 * (PM_force.cpp:1070-1169), which stays the fallback; its impulse
 * multiple-time-step switch (PME update_interval, PM_force.cpp:1082,
 * 1135-1140) is the method the surrogate is compared with.  Training frames
 * (features, basis, target force) stay resident in device memory. */
typedef struct sponge_rsur sponge_rsur;

typedef struct sponge_rsur_options {
    float feature_cutoff; /* Rf in A, at most the neighbor list's cutoff (default 10) */
    int max_frames;       /* training frames kept on the device (default 0) */
    int seed;             /* initial weights and frame shuffling (default 1) */
} sponge_rsur_options;

void sponge_rsur_default_options(sponge_rsur_options* opt);

/* Floats in a model blob: parameters (W1[64][48], b1[64], W2[64][64], b2[64],
 * W3[3][64], b3[3]) followed by the feature mean[48] and inverse std[48]. */
int sponge_rsur_model_size(void);

/* beta is the Ewald splitting parameter of the PME module it stands in for. */
int sponge_rsur_initial(sponge_rsur** out, int atom_numbers, float beta,
                        const sponge_rsur_options* opt, void* stream);

/* Features and basis of the current coordinates over the neighbor list
 * (which must have been built; its partners within the cutoff are complete)
 * and the exclusion CSR (device arrays, may be NULL). */
int sponge_rsur_features(sponge_rsur* h, sponge_nl* nl, const float* crd, const float* cell,
                         const float* rcell, const float* charge, const int* excluded_list_start,
                         const int* excluded_list, const int* excluded_numbers);

/* The last features [N][48] and basis [N][3][3] in atom order (device arrays;
 * either may be NULL). */
int sponge_rsur_export_features(sponge_rsur* h, float* feat, float* basis);

/* frc += predicted reciprocal force of the last features (device VECTOR[N]). */
int sponge_rsur_predict(sponge_rsur* h, float* frc);

/* Model blob of sponge_rsur_model_size() floats, host memory; synchronise. */
int sponge_rsur_get_model(sponge_rsur* h, float* model);
int sponge_rsur_set_model(sponge_rsur* h, const float* model);

/* Stores the last features and basis with the target rec_frc - mean(rec_frc)
 * (device VECTOR[N], the PME reciprocal force of the same coordinates). */
int sponge_rsur_add_frame(sponge_rsur* h, const float* rec_frc);
int sponge_rsur_frame_count(sponge_rsur* h, int* n);

/* Feature mean and inverse standard deviation from the first n stored frames. */
int sponge_rsur_fit_normalization(sponge_rsur* h, int n_frames);

/* Loss (mean squared force error per component) and its gradient on one
 * stored frame (host arrays of model-parameter length / 1 double; either may
 * be NULL); synchronises. */
int sponge_rsur_gradient(sponge_rsur* h, int frame, float* grad, double* loss);

/* Adam over the first n_train_frames stored frames, one step per frame,
 * `epochs` passes in shuffled order; epoch_loss[epochs] (host, may be NULL)
 * gets the mean loss of each pass.  Synchronises. */
int sponge_rsur_train(sponge_rsur* h, int n_train_frames, int epochs, float lr, double* epoch_loss);

/* Relative RMS force errors over stored frames [begin, end) (frames are
 * kept in atom order): out[0] the model, out[1] the physics baseline
 * c = (1, 1, 0); with lag > 0 each frame f is also predicted anchored on
 * frame f - lag (prediction + that frame's target - its prediction):
 * out[2] model anchored, out[3] baseline anchored, out[4] the stale target
 * of frame f - lag itself.  Synchronises. */
int sponge_rsur_evaluate(sponge_rsur* h, int frame_begin, int frame_end, int lag, double* out);

int sponge_rsur_clear(sponge_rsur* h);

/* ------------------------------------------------------------------------- */
/* NVE step driver (the caller side of the path)                              */
/* ------------------------------------------------------------------------- */

/* One periodic orthorhombic system of rigid 3-site waters (+ ions) owned on
 * the device, stepped as SPONGE's main loop steps it in NVE mode with SETTLE:
 * Main_Calculate_Force (SPONGE/main.cpp:1355-1626: reset, neighbor_list
 * CONDITIONAL update, PME excluded correction, LJ + PME direct, long-range
 * correction on energy steps, PME reciprocal) followed by Main_Iteration
 * (:1690-1777: remember, Leap_Frog, Do_SETTLE).  All arrays are HOST arrays
 * copied at set-up; water_atoms holds (O, H1, H2) per water, whose rigid
 * geometry is taken from the first water's coordinates.  dt_ps is in ps. */
typedef struct sponge_md sponge_md;

int sponge_md_initial(sponge_md** out, int atom_numbers, const float* crd, const float* vel,
                      const float* mass, const float* charge, int atom_type_numbers,
                      const float* LJ_A, const float* LJ_B, const int* atom_LJ_type,
                      const int* excluded_list_start, const int* excluded_numbers,
                      const int* excluded_list, int n_water, const int* water_atoms,
                      const float* box_length, float cutoff, float skin, float dt_ps,
                      const sponge_nl_options* nl_opt, const sponge_pme_options* pme_opt,
                      void* stream);

/* Advances `steps` steps asynchronously.  Energies are evaluated on steps
 * with step % energy_interval == 0 (0 = never).  use_graph != 0 captures the
 * force-only step into one HIP graph on first use and replays it (default
 * schedule only: automatic list refresh, reciprocal part every step); while
 * at least 8 force-only steps remain, 8 steps per graph launch.
 * After every step with (step + 1) % check_overflow_interval == 0 (the
 * neighbor-list option; <= 0 = never) the list's overflow flags are read
 * (a host synchronisation), as NEIGHBOR_LIST::Check_Overflow does
 * (neighbor_list.cpp:1027-1103): on an overflow the list is re-initialised
 * with 1.1x the limit that overflowed and rebuilt on the next step, or, with
 * throw_error_when_overflow (or a stream overflow, which no limit lifts), the
 * call returns SPONGE_AMD_ERROR_OVERFLOW at that step. */
int sponge_md_run(sponge_md* md, int steps, int energy_interval, int use_graph);

/* Synchronises and reports out[10] = {potential, kinetic, LJ_short, LJ_long,
 * PM, PM_direct, PM_reciprocal, PM_self, PM_correction, sum of per-atom
 * energies} of the last energy step (kinetic from the current velocities). */
int sponge_md_energies(sponge_md* md, double* out);

/* Copies the current coordinates, velocities and last forces to host arrays
 * of 3*atom_numbers floats (any may be NULL); synchronises. */
int sponge_md_get_state(sponge_md* md, float* crd, float* vel, float* frc);

/* Replaces the coordinates and / or velocities from host arrays (either may be
 * NULL); synchronises.  A coordinate change is picked up by the next
 * CONDITIONAL list update through the displacement check. */
int sponge_md_set_state(sponge_md* md, const float* crd, const float* vel);

/* The module handles the driver owns (for timing and list export). */
int sponge_md_handles(sponge_md* md, sponge_nl** nl, sponge_lj** lj, sponge_pme** pme);

/* Inner rows of the driver's pair kernel (no counterpart in the reference):
 * the list's entries within cutoff + delta (Angstrom) at the last prune,
 * pruned again once an atom moved more than delta / 2 since then or the list
 * was rebuilt, so the pair kernel skips most pairs beyond the cutoff.  The
 * pair set of every force is unchanged (every pair within the cutoff is in
 * the inner rows); the list itself (sponge_nl_export_half_list) is the
 * reference's.  delta = 0 reads the full rows.  Default
 * SPONGE_AMD_MD_LIST_PRUNE_DELTA; at most the skin. */
#define SPONGE_AMD_MD_LIST_PRUNE_DELTA 0.0f
int sponge_md_set_list_prune(sponge_md* md, float delta);

/* The driver's list is rebuilt from the coordinates of the next step, alone or
 * batched (NEIGHBOR_LIST::Update with update = FORCED, neighbor_list.cpp:
 * 983-998, for the list the driver holds). */
int sponge_md_request_list_update(sponge_md* md);

/* sponge_pme_autotune on the driver's PME module and current coordinates;
 * tiles[3] (host, may be NULL) gets the choice.  Synchronises. */
int sponge_md_autotune(sponge_md* md, int reps, int* tiles);

/* Whether a plain PME step runs its reciprocal half (bin, spread, FFT
 * convolution) on a second stream beside the list and pair kernels, joined
 * before the gather (no counterpart in the reference, whose Main_Calculate_
 * Force, SPONGE/main.cpp, evaluates the terms one after another): mode 1 on,
 * 0 off, -1 automatic (on from SPONGE_AMD_MD_OVERLAP_MIN_ATOMS atoms, where
 * the fork and join cost less than the overlap saves).  The captured step is
 * rebuilt.  The forces are the same either way up to the order of the float
 * additions into the force array. */
#define SPONGE_AMD_MD_OVERLAP_MIN_ATOMS 12000
int sponge_md_set_reciprocal_overlap(sponge_md* md, int mode);

/* Soft-core mode of a lambda window: the step's LJ + PME direct term is the
 * soft-core module sc (LJ_Soft_Core_PME_Direct_Force_With_Atom_Energy_And_
 * Virial, SPONGE/main.cpp:1460-1465, over the driver's list) instead of the
 * plain LJ module, and energy steps take its Long_Range_Correction
 * (:1479-1482) and LJ_soft_short; the reference initialises one of the two
 * modules (main.cpp:1179-1180, the one whose input exists).  The charges of
 * the system are the window's (lambda-mixed) charges.  sc must have the
 * system's atom count; it stays owned by the caller; NULL returns to the plain
 * module.  The batched driver runs soft-core systems too. */
int sponge_md_set_soft_core(sponge_md* md, sponge_ljsc* sc);

/* Learned multiple-time-step schedule (no counterpart in the reference; the
 * reference's own way to skip reciprocal steps is PME update_interval,
 * PM_force.cpp:1082, 1135-1140): with rsur set, the steps with step % period
 * != 0 that evaluate no energy take the surrogate's predicted reciprocal force
 * instead of PME's.  The first surrogate step after the PME step of every
 * check_interval-th period (0 = never) is also evaluated with PME; when the
 * prediction's relative RMS error exceeds bound, that step uses PME's force
 * (the fallback) and the surrogate is switched off: every later step runs
 * plain PME until this call installs a model again (the driver reads the
 * fallback counter once per schedule cycle that holds a checked step).  The
 * surrogate runs on the driver's stream from then on; its feature cutoff must
 * not exceed the driver's cutoff.  rsur = NULL restores plain PME.  Graph
 * replay captures whole cycles of the schedule.
 * The surrogate's force is anchored on the last PME step and is not the
 * gradient of an energy: this schedule does not conserve energy the way
 * plain PME (or the reference's impulse schedule) does. */
int sponge_md_set_surrogate(sponge_md* md, sponge_rsur* rsur, int period, int check_interval,
                            float bound);

/* Force of the surrogate steps (no counterpart in the reference):
 * mode 1 (default), impulse: the surrogate's force on every step and, on each
 * period-th (PME) step, the kick period x (F_PME - F_surrogate) -- the
 * reference's update_interval splitting (PM_force.cpp:1082, 1135-1140) with the
 * surrogate as the part evaluated every step, which is time-reversible and
 * conserves energy as far as the surrogate force is a gradient; mode 0,
 * anchored: F_surrogate(t) + F_PME(t0) - F_surrogate(t0) on every step, the
 * smaller instantaneous error but no conserved energy.  The checked steps
 * measure the anchored force's error in both modes. */
int sponge_md_set_surrogate_mode(sponge_md* md, int mode);

/* out[4] = {surrogate steps, checked steps, fallbacks, switched off (0/1)};
 * *last_error = the relative RMS error of the last checked step (either may
 * be NULL). */
int sponge_md_surrogate_stats(sponge_md* md, long long* out, double* last_error);

/* Runs n_frames * stride plain steps (eager); the last step of every stride
 * stores a training frame (features and PME reciprocal force of that step's
 * coordinates) into rsur.  Synchronises. */
int sponge_md_collect_frames(sponge_md* md, sponge_rsur* rsur, int n_frames, int stride);

int sponge_md_clear(sponge_md* md);

/* ------------------------------------------------------------------------- */
/* Batched driver                                                              */
/* ------------------------------------------------------------------------- */

/* Many independent systems (each a sponge_md with its own neighbor list, PME
 * grid and FFT tiling; plain or soft-core) advanced together: one step of
 * every system is recorded through the same code that launches it alone, and
 * each kernel of the step then runs as ONE launch over all systems (each
 * workgroup finds its system from the prefix of the systems' workgroup
 * counts); steps_per_launch steps of these launches form one HIP graph.
 * The reciprocal chain of each step (PME bin, spread, FFT convolution) runs
 * on a second stream of the batch, forked at the step's start and joined
 * before the gather, as in sponge_md_set_reciprocal_overlap.
 * No counterpart in the reference (one process runs one system there).  The
 * systems stay owned by the caller. */
typedef struct sponge_batch sponge_batch;

int sponge_batch_initial(sponge_batch** out, int n_systems, sponge_md* const* systems);

/* sponge_pme_autotune for every distinct grid shape of the batch (the same
 * choice for systems that share a shape), then, per FFT axis, those choices
 * against one common value of {2, 4, 8, 16} timed on batched steps (one warm
 * and `reps` timed launches of 10 steps per candidate: the call advances
 * every system by 150 + 150 reps steps); tiles[3 * n_systems] (host, may be
 * NULL) gets each system's final choice.  Synchronises. */
int sponge_batch_autotune(sponge_batch* b, int reps, int* tiles);

/* Advances every system by `steps` steps: whole launches of the graph of
 * steps_per_launch steps, the remainder step by step per system.  Systems on
 * a schedule a graph cannot hold (list not built yet, fixed refresh or PME
 * intervals, a surrogate) all run eagerly instead.  Kernel timers must be off
 * (CONFLICTING_COMMAND otherwise).  Asynchronous. */
int sponge_batch_run(sponge_batch* b, int steps, int steps_per_launch);
int sponge_batch_synchronize(sponge_batch* b);
int sponge_batch_clear(sponge_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* SPONGE_AMD_H */
