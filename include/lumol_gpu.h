/*
 * lumol_gpu.h -- C ABI of the MI355X (gfx950) force-and-energy backend for
 * Lumol.
 *
 * The library keeps one system resident in HBM per handle: positions,
 * velocities, masses, charges, kinds, molecule topology, the potential
 * tables, the cell list and the Ewald k-space caches.  Host buffers cross
 * the boundary only in lg_system_set_* / lg_system_get_* / lg_forces /
 * lg_potential_energy / lg_virial; MD steps and MC moves run on the device.
 *
 * Every entry point returns an int status (LG_OK = 0).  No exception, panic
 * or abort crosses the boundary; lg_last_error() returns a thread-local
 * message for the last failure on the calling thread.  Calls on one handle
 * are ordered on the handle's stream and must not overlap from several host
 * threads (the reference serialises Ewald behind a write lock,
 * lumol-core/src/energy/global/ewald.rs:847-874; the handle does the same by
 * contract).
 *
 * Units are Lumol's internal units (A, fs, u, K; energy in u*A^2/fs^2), and
 * vectors are AoS triples of doubles, as Lumol's Vector3D
 * (lumol-core/src/types/vectors.rs:59).
 *
 * Citations are to the reference checkout (lumol-org/lumol).
 */
#ifndef LUMOL_GPU_H
#define LUMOL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    LG_OK = 0,
    LG_E_INVALID = 1,     /* bad argument or inconsistent system            */
    LG_E_UNSUPPORTED = 2, /* valid for Lumol but outside the GPU kernel set */
    LG_E_DEVICE = 3,      /* HIP runtime error                              */
    LG_E_STATE = 4,       /* call order (e.g. commit without a cost)        */
    LG_E_NONFINITE = 5    /* energy or coordinates are not finite           */
};

/* ---- closed set of GPU potential kinds ----------------------------------
 * Replaces the open `Potential`/`PairPotential` trait objects
 * (lumol-core/src/energy/mod.rs:98-163); kernels cannot call host closures.
 * `params` of each kind, in the reference's argument order, and where the
 * reference implements it (P pair, B bond, A angle, D dihedral):
 *   NULL {}                                functions.rs:30-50     P B A D
 *   LJ {sigma, epsilon}                    functions.rs:52-107    P
 *   HARMONIC {k, x0}                       functions.rs:109-160   P B A D
 *   BUCKINGHAM {A, C, rho}                 functions.rs:260-327   P
 *   BORN (Born-Mayer-Huggins)
 *        {A, C, D, sigma, rho}             functions.rs:330-395   P
 *   MORSE {A, x0, depth}                   functions.rs:397-437   P B A D
 *   GAUSSIAN {a, b}, b > 0                 functions.rs:439-499   P
 *   MIE {sigma, epsilon, n, m}, m < n      functions.rs:501-592   P
 *   COSINE_HARMONIC {k, x0}                functions.rs:162-210     A D
 *   TORSION {k, delta, n}, n integer >= 0  functions.rs:212-258       D
 * A kind used where the reference has no such trait implementation returns
 * LG_E_UNSUPPORTED; parameters the reference's constructors reject
 * (Gaussian::new, Mie::new assert) return LG_E_INVALID.
 */
enum {
    LG_POT_NULL = 0,
    LG_POT_LJ = 1,
    LG_POT_HARMONIC = 2,
    LG_POT_BUCKINGHAM = 3,
    LG_POT_BORN = 4,
    LG_POT_MORSE = 5,
    LG_POT_GAUSSIAN = 6,
    LG_POT_MIE = 7,
    LG_POT_COSINE_HARMONIC = 8,
    LG_POT_TORSION = 9
};

/* PairRestriction (lumol-core/src/energy/restrictions.rs:13-33). */
enum {
    LG_RESTRICT_NONE = 0,
    LG_RESTRICT_INTRA = 1,
    LG_RESTRICT_INTER = 2,
    LG_RESTRICT_EXCLUDE12 = 3,
    LG_RESTRICT_EXCLUDE13 = 4,
    LG_RESTRICT_EXCLUDE14 = 5,
    LG_RESTRICT_SCALE14 = 6
};

/* Arithmetic form of the evaluation. */
enum {
    LG_PRECISION_FP64 = 0,  /* all FP64: matches the reference to 1e-10 rel   */
    LG_PRECISION_MIXED = 1  /* FP32 pair/erfc/trig, FP64 positions/sums        */
};

enum { LG_INTEGRATOR_VELOCITY_VERLET = 0, LG_INTEGRATOR_LEAP_FROG = 1 };

/* Energy terms of EnergyEvaluator (lumol-core/src/sys/energy.rs:47-158). */
typedef struct lg_energy_terms {
    double pairs;      /* EnergyEvaluator::pairs      energy.rs:47   */
    double pairs_tail; /* EnergyEvaluator::pairs_tail energy.rs:63   */
    double bonds;      /* EnergyEvaluator::bonds      energy.rs:90   */
    double angles;     /* EnergyEvaluator::angles     energy.rs:110  */
    double dihedrals;  /* EnergyEvaluator::dihedrals  energy.rs:131  */
    double coulomb;    /* EnergyEvaluator::coulomb    energy.rs:145  */
    double global;     /* EnergyEvaluator::global     energy.rs:152  */
    double total;      /* PotentialEnergy::compute    compute.rs:114 */
    double coulomb_real; /* Ewald real space   ewald.rs:431 (Wolf: pairs+self) */
    double coulomb_self; /* Ewald self energy  ewald.rs:619              */
    double coulomb_kspace; /* Ewald k-space    ewald.rs:677              */
} lg_energy_terms;

typedef struct lg_context lg_context;
typedef struct lg_system lg_system;

/* ---- errors --------------------------------------------------------------- */
/* Message of the last failed call on this thread (never NULL). */
const char* lg_last_error(void);
/* Library version string. */
const char* lg_version(void);

/* ---- context: one per GPU ---------------------------------------------------- */
/* Opens device `device` and creates the stream the context's systems use.
 * No counterpart in the reference (rayon's global pool, lumol-core/Cargo.toml). */
int lg_context_create(int device, lg_context** out);
int lg_context_destroy(lg_context* ctx);

/* ---- system: replaces System / Configuration ----------------------------------
 * System::new / with_cell (lumol-core/src/sys/system.rs:41-80). */
int lg_system_create(lg_context* ctx, size_t natoms, lg_system** out);
int lg_system_destroy(lg_system* sys);
int lg_system_size(const lg_system* sys, size_t* natoms);

/* UnitCell::ortho / cubic (lumol-core/src/sys/config/cells.rs:60-100).
 * Orthorhombic only; lengths all zero = UnitCell::infinite (cells.rs:50).
 * Triclinic cells return LG_E_UNSUPPORTED. */
int lg_system_set_cell(lg_system* sys, double a, double b, double c);

/* Particle SoA (lumol-core/src/sys/config/particles.rs:31-46). */
int lg_system_set_positions(lg_system* sys, const double* xyz /* 3N */);
int lg_system_get_positions(lg_system* sys, double* xyz /* 3N */);
int lg_system_set_velocities(lg_system* sys, const double* v /* 3N */);
int lg_system_get_velocities(lg_system* sys, double* v /* 3N */);
int lg_system_set_masses(lg_system* sys, const double* mass /* N */);
int lg_system_set_charges(lg_system* sys, const double* q /* N */);
/* ParticleKind per atom (particles.rs:36, interactions.rs:94-103). */
int lg_system_set_kinds(lg_system* sys, const uint32_t* kinds /* N */, uint32_t nkinds);

/* System::add_bond (lumol-core/src/sys/config/configuration.rs:150-200).
 * Molecules are the connected components of the bond graph and must be
 * contiguous atom ranges; angles, dihedrals and bond paths follow
 * Bonding::rebuild (lumol-core/src/sys/config/bonding.rs:78-155). */
int lg_system_set_bonds(lg_system* sys, size_t nbonds, const uint32_t* pairs /* 2*nbonds */);
int lg_system_molecule_count(const lg_system* sys, size_t* nmol);

/* Evaluation form; see LG_PRECISION_*. */
int lg_system_set_precision(lg_system* sys, int precision);

/* ---- potentials ------------------------------------------------------------------
 * System::set_pair_potential (system.rs:122-131) with a PairInteraction
 * (lumol-core/src/energy/pairs.rs:57-297): cutoff `r >= cutoff -> 0`,
 * `shifted` subtracts V(cutoff) from the energy only, `tail` enables
 * tail corrections, `restriction`/`scale14` as PairRestriction.
 * Returns LG_E_INVALID when the cutoff is above half the smallest cell
 * length (system.rs:123-129 panics there). */
int lg_pair_set(lg_system* sys, uint32_t kind_i, uint32_t kind_j, int pot_kind, const double* params,
                double cutoff, int shifted, int tail, int restriction, double scale14);
/* lg_pair_set with the potential evaluated through
 * TableComputation::new(potential, table_size, table_max)
 * (lumol-core/src/energy/computations.rs:96-117): energy and force are
 * tabulated at r = i * table_max / table_size and interpolated linearly,
 * zero from the last bin on (:122-146); tails stay those of the potential
 * (:148-175) and a shift is the table's energy at the cutoff. */
int lg_pair_set_table(lg_system* sys, uint32_t kind_i, uint32_t kind_j, int pot_kind, const double* params,
                      uint64_t table_size, double table_max, double cutoff, int shifted, int tail,
                      int restriction, double scale14);
/* System::set_bond_potential / set_angle_potential / set_dihedral_potential
 * (system.rs:133-170). */
int lg_bond_set(lg_system* sys, uint32_t kind_i, uint32_t kind_j, int pot_kind, const double* params);
int lg_angle_set(lg_system* sys, uint32_t kind_i, uint32_t kind_j, uint32_t kind_k, int pot_kind,
                 const double* params);
int lg_dihedral_set(lg_system* sys, uint32_t kind_i, uint32_t kind_j, uint32_t kind_k, uint32_t kind_m,
                    int pot_kind, const double* params);

/* System::set_coulomb_potential with SharedEwald::new(Ewald::new(cutoff,
 * kmax, alpha)) (ewald.rs:278-305, 847-956); alpha NaN selects pi/cutoff
 * (ewald.rs:279).  `restriction` as CoulombicPotential::set_restriction
 * (global/mod.rs:197-202); Scale14 is rejected as ewald.rs:388 asserts. */
int lg_coulomb_ewald(lg_system* sys, double cutoff, int kmax, double alpha, int restriction);
/* Wolf::new(cutoff) (wolf.rs:68-84).
 * Both return LG_E_INVALID when the cutoff is above half the smallest
 * length of a periodic cell (system.rs:157-167 panics there); evaluations
 * re-check it, since the cell may change after the potential is set. */
int lg_coulomb_wolf(lg_system* sys, double cutoff, int restriction, double scale14);
int lg_coulomb_clear(lg_system* sys);

/* ---- evaluation --------------------------------------------------------------------
 * System::forces (system.rs:316 -> compute.rs:30-108).  Writes 3N doubles;
 * with accumulate != 0 adds into `forces` instead (GlobalPotential::forces
 * accumulates, global/mod.rs:95). */
int lg_forces(lg_system* sys, double* forces /* 3N */, int accumulate);
/* System::potential_energy and the EnergyEvaluator terms
 * (system.rs:263, compute.rs:111-128, energy.rs:47-158). */
int lg_potential_energy(lg_system* sys, lg_energy_terms* out);
/* lg_potential_energy that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-block energy and virial partial sums go to a static device array of
 * 2^20 doubles (8 MiB per GPU, held until the process ends) instead of
 * buffers the library allocates; the results are bit-identical to
 * lg_potential_energy's.  The system must have been evaluated since its last
 * change (lg_potential_energy or lg_forces), else LG_E_STATE; partial sums
 * beyond the array return LG_E_INVALID; a neighbour list that would have to
 * grow returns LG_E_STATE.  Calls that overlap on one GPU (from several host
 * threads or contexts) race on the shared array. */
int lg_potential_energy_noalloc(lg_system* sys, lg_energy_terms* out);
/* System::virial with the atomic definition (system.rs:286,
 * compute.rs:195-255), row-major 3x3. */
int lg_virial(lg_system* sys, double* virial /* 9 */);
/* MolecularVirial (compute.rs:278-364), row-major 3x3: the pair virials
 * between molecules weighted by (r_ab . r_ij) / r_ab^2 (r_ij between the
 * centres of mass), the tail corrections, no bond/angle/dihedral term, and
 * for Ewald the real-space pairs between molecules plus the k-space atomic
 * virial minus sum_i F_i (x) (r_i - com) (ewald.rs:533-570, 736-753; Wolf
 * wolf.rs:285-324).  The virial of System::virial when the simulated degrees
 * of freedom are molecules (compute.rs:372-381). */
int lg_molecular_virial(lg_system* sys, double* virial /* 9 */);
/* KineticEnergy / Temperature (compute.rs:133-172; dof = 3N, system.rs:249). */
int lg_kinetic_energy(lg_system* sys, double* out);
int lg_temperature(lg_system* sys, double* out);
/* Pair set of the device neighbour search: every inter-molecular (i < j)
 * with r < cutoff, r computed by the kernels' own minimum image and norm.
 * Replaces the brute-force `for j in i+1..n` candidate loop
 * (compute.rs:40) as a checkable set; `cutoff` must not exceed the largest
 * interaction cutoff.  Writes min(count, capacity) pairs (2 uint32 each,
 * unordered) and the total in *count. */
int lg_pairs_within(lg_system* sys, double cutoff, uint32_t* pairs, uint64_t capacity, uint64_t* count);
/* Number of k-vectors of the Ewald half-space set (ewald.rs:130-183). */
int lg_ewald_kvector_count(lg_system* sys, int* out);

/* ---- molecular dynamics --------------------------------------------------------------
 * MolecularDynamics::new(integrator) + Simulation::run (lumol-sim/src/md/
 * molecular_dynamics.rs:11-83, simulations.rs:69-97).  `integrator` is
 * VelocityVerlet (integrators.rs:22-70) or LeapFrog (:128-169); the first
 * step starts from zero accelerations as both setups do (:40-42, :146-148).
 * berendsen_temperature >= 0 enables BerendsenThermostat::new(T, tau)
 * (thermostats.rs:91-121; T = 0 is valid there, :103), applied after each
 * step; a negative or NaN temperature means no thermostat. */
int lg_md_setup(lg_system* sys, int integrator, double timestep, double berendsen_temperature,
                double berendsen_tau);
/* Runs `nsteps` steps on the device (positions/velocities stay in HBM).
 * Returns LG_E_NONFINITE when a position, velocity or acceleration became
 * non-finite (checked on the device every step, reported after each chunk of
 * at most 200 steps; the state of the end of that chunk stays on the
 * device).  The reference's Simulation::run only warns, every 10 000 steps,
 * about coordinates above 1e6 A or velocities above 1000 A/fs
 * (lumol-sim/src/simulations.rs:94-96, 129-155). */
int lg_md_run(lg_system* sys, uint64_t nsteps);

/* ---- Monte Carlo energy cache ---------------------------------------------------------
 * EnergyCache::init (lumol-core/src/sys/cache.rs:75-98).  Computes the
 * cached energy terms and the Ewald rho(k) cache on the device; `energy`
 * (may be NULL) receives EnergyCache::energy (cache.rs:101-113).  Setting
 * positions, charges, cell or potentials, or running MD, drops the cache; the
 * calls below re-initialise it when needed. */
int lg_cache_init(lg_system* sys, double* energy);
/* EnergyCache::energy (cache.rs:101-113) of the cached configuration. */
int lg_cache_energy(lg_system* sys, double* energy);
/* EnergyCache::move_molecule_cost (cache.rs:145-220): energy change of
 * moving molecule `molid` to `newpos` (3 * molecule size doubles), pairs +
 * coulomb (SharedEwald/Wolf GlobalCache::move_molecule_cost, ewald.rs:933-945,
 * wolf.rs:121-165).  Stages the move.  At most 32 atoms per molecule. */
int lg_move_cost(lg_system* sys, uint32_t molid, const double* newpos, double* cost);
/* EnergyCache::update (cache.rs:118-127) + Translate::apply
 * (lumol-sim/src/mc/moves/translate.rs:115-127, COM wrap when wrap != 0).
 * LG_E_STATE when no move is staged. */
int lg_move_commit(lg_system* sys, int wrap);
/* Rejected move: drop the staged update (monte_carlo.rs:318-321). */
int lg_move_discard(lg_system* sys);
/* MonteCarlo::propagate (lumol-sim/src/mc/monte_carlo.rs:289-330) with one
 * Translate move (translate.rs:41-148) repeated for attempts attempt0 ..
 * attempt0 + nattempts - 1, entirely on the device: uniform molecule choice,
 * displacement U(+-delta/sqrt(3)) per axis, Metropolis on beta * cost, COM
 * wrap on accept.  Random numbers: Philox4x32-10 with key = seed and counter
 * (attempt, w, 0, 0), w = 0 (molecule, x), 1 (y, z), 2 (acceptance); 53-bit
 * uniforms from word pairs.  Replaces the reference's xorshift stream (same
 * distribution, different numbers).  `accepted` may be NULL. */
int lg_mc_run(lg_system* sys, uint64_t seed, uint64_t attempt0, uint64_t nattempts, double beta, double delta,
              uint64_t* accepted);
/* lg_mc_run for `count` independent chains (distinct systems of one
 * context) in one launch, one workgroup per chain; per-chain arguments are
 * arrays.  No counterpart in the reference, which runs one chain per
 * process; this is the batched form of the same propagate loop. */
int lg_mc_run_batch(lg_system* const* systems, size_t count, const uint64_t* seeds, const uint64_t* attempt0,
                    const uint64_t* nattempts, const double* beta, const double* delta, uint64_t* accepted);

/* ---- launch tuning --------------------------------------------------------------------
 * No counterpart in the reference (a CPU code has no launch shapes).  Knobs of
 * the device kernels; 0 leaves the built-in size-based choice:
 *   pair_lanes    lanes per atom in the pair and neighbour-list kernels (4..64;
 *                 16 for systems with a pair potential other than LJ,
 *                 Harmonic and Null, or a table)
 *   kspace_split  row ranges of the k-space force kernel (1..32)
 *   rho_width     atoms per block of the row form of the rho(k) kernel (a
 *                 multiple of 64 up to 512)
 *   nlist_build   neighbour-list build: 1 lanes per atom, 2 one workgroup per
 *                 cell with candidates staged in LDS, 3 direct: workgroups
 *                 per cell over the atoms' cell tags, no cell list (up to
 *                 32768 atoms; a box of fewer than 256 cells deals each
 *                 cell's home atoms to up to 8 workgroups).  Default: 3 when
 *                 there are >= 128 cells and >= 8 atoms per cell and it
 *                 applies, else 2 there, else 1
 *   md_graph      lg_md_run: 1 (default) launches every kernel; 2 captures a
 *                 chunk of steps in a HIP graph once and replays it (eager
 *                 launches while a kernel profile is active).  On ROCm 7.0
 *                 the replay measured 4-18% slower than eager launches, so
 *                 it is off by default.
 *   rho_kind      rho(k) kernel: 1 row form (one lane per run of five
 *                 k-vectors, a loop over atoms, the phase rows of a few
 *                 atoms at a time staged in LDS up to kmax 12; the default),
 *                 2 matrix-core form (rho as a complex product over atoms on
 *                 v_mfma 16x16x4; kmax <= 16, else the row form), 3 the row
 *                 form reading the phase rows from the table (no staging;
 *                 the same sums, bit for bit).
 * lumol_amd.tuning picks them per size bucket by timing the candidates.
 * Results do not depend on them beyond summation order. */
typedef struct lg_tuning {
    int pair_lanes;
    int kspace_split;
    int rho_width;
    int nlist_build;
    int md_graph;
    int rho_kind;
} lg_tuning;
int lg_system_set_tuning(lg_system* sys, const lg_tuning* tuning);
/* The values in effect (after the automatic choice); kspace_split and
 * rho_width are 0 for systems without Ewald k-space. */
int lg_system_get_tuning(lg_system* sys, lg_tuning* tuning);

/* ---- kernel timing ---------------------------------------------------------------
 * No counterpart in the reference (its benches time whole calls with
 * criterion, benches/argon.rs etc.).  Between begin and end, every launch of a kernel
 * class in `class_mask` (bit = LG_KERNEL_*) is timed on the system's stream:
 * the pair, intra-molecular, rho(k) and k-space force launches by their
 * kernels' own start and end timestamps (the events of
 * hipExtLaunchKernelGGL, which agree with a kernel trace), the others by a
 * HIP event pair around the launch; end synchronises and returns the summed
 * device time and the launch count per class. */
enum {
    LG_KERNEL_NLIST = 0,        /* displacement check + cell list + Verlet list */
    LG_KERNEL_PAIRS = 1,        /* k_pairs: LJ + Ewald real space / Wolf, and, in
                                   the FP64 form, the intra-molecular terms (extra
                                   blocks of the same launch)                     */
    LG_KERNEL_MOLECULES = 2,    /* k_intra_terms: the intra-molecular terms in a
                                   launch of their own (mixed form)               */
    LG_KERNEL_KSPACE_RHO = 3,   /* k_rho_partial + k_rho_reduce                  */
    LG_KERNEL_KSPACE_FORCE = 4, /* k_kspace_force                                */
    LG_KERNEL_INTEGRATE = 5,    /* integrator halves + thermostat                */
    LG_KERNEL_CLASSES = 6
};
typedef struct lg_profile {
    double ms[6];
    uint64_t launches[6];
} lg_profile;
int lg_profile_begin(lg_system* sys, uint32_t class_mask);
int lg_profile_end(lg_system* sys, lg_profile* out);

#ifdef __cplusplus
}
#endif

#endif /* LUMOL_GPU_H */
