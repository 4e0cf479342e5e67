/* mjg: MI355X-native force-and-integrate engine for Mjolnir-style coarse-grained MD.
 *
 * C ABI of libmjg.so.  Plain pointers and sizes only; every buffer passed in is owned by the
 * caller and copied, every device buffer is owned by the engine.  Each entry point replaces a
 * piece of the reference's class surface (paths relative to the Mjolnir source tree):
 *
 *   System<traits>                       mjolnir/core/System.hpp:18-207
 *   ForceFieldBase<traits>               mjolnir/core/ForceFieldBase.hpp:13-52
 *   LocalInteractionBase / Global...     mjolnir/core/LocalInteractionBase.hpp:10-41,
 *                                        mjolnir/core/GlobalInteractionBase.hpp:10-38
 *   SpatialPartition (neighbour lists)   mjolnir/core/SpatialPartitionBase.hpp:18-145
 *   BAOABLangevinIntegrator              mjolnir/core/BAOABLangevinIntegrator.hpp:17-240
 *   GJFNVTLangevinIntegrator             mjolnir/core/GJFNVTLangevinIntegrator.hpp:17-244
 *
 * Error convention: the reference throws std::runtime_error / std::out_of_range at set-up and
 * keeps the hot path noexcept (ForceFieldBase.hpp:33-37).  Here every call returns an
 * mjg_status; set-up errors carry a message available from mjg_last_error().  Device faults
 * and non-finite state found during stepping are reported as MJG_DEVICE_ERROR /
 * MJG_NAN_DETECTED by the call that observes them.
 *
 * Threading: one engine is driven by one host thread at a time (calls on the same engine are
 * not re-entrant); engines are independent and own one HIP stream each.
 */
#ifndef MJG_H
#define MJG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mjg_engine mjg_engine;
typedef int mjg_status;

enum {
    MJG_OK              = 0,
    MJG_INVALID_ARG     = 1,
    MJG_OOM             = 2,
    MJG_DEVICE_ERROR    = 3,
    MJG_NAN_DETECTED    = 4,
    MJG_NOT_INITIALIZED = 5,
    MJG_UNSUPPORTED     = 6
};

/* arithmetic of the force kernels */
enum {
    MJG_FP64       = 0, /* FP64 throughout */
    MJG_FP32_FIXED = 1  /* FP32 pair/local arithmetic, int64 fixed-point force sums,
                           FP64 integrator state.  On an engine (mjg_engine_create) the
                           LJ / excluded-volume / Debye-Hueckel / 3SPN2 excluded-volume pair
                           kernel runs this way; local terms, 3SPN2 base pairing and cross
                           stacking stay FP64 there.  Batches run every term this way. */
};
/* boundary (read_system.hpp:100-299: boundary_type "Unlimited" / "PeriodicCuboid") */
enum { MJG_UNLIMITED = 0, MJG_PERIODIC = 1 };

/* local interactions: interaction:potential (read_local_interaction.hpp:31-715) */
enum {
    MJG_BOND_HARMONIC     = 1, /* BondLength:Harmonic  {k, v0}        BondLengthInteraction.hpp:126-148 */
    MJG_BOND_GOCONTACT    = 2, /* BondLength:GoContact {k, v0}        BondLengthGoContactInteraction.hpp:126-167 */
    MJG_CONTACT_GOCONTACT = 3, /* Contact:GoContact    {k, v0}        GoContactInteraction.hpp:159-228 */
    MJG_ANGLE_HARMONIC    = 4, /* BondAngle:Harmonic   {k, v0}        BondAngleInteraction.hpp:129-189 */
    MJG_ANGLE_GAUSSIAN    = 5, /* BondAngle:Gaussian   {k, sigma, v0} BondAngleInteraction.hpp:129-189 */
    MJG_DIHEDRAL_CLEMENTI = 6, /* DihedralAngle:ClementiDihedral {k1, k3, v0} DihedralAngleInteraction.hpp:157-227 */
    MJG_BOND_3SPN2        = 7, /* BondLength:3SPN2Bond {k, v0}        ThreeSPN2BondPotential.hpp:22-35 */
    MJG_DIHEDRAL_GAUSSIAN = 8, /* DihedralAngle:Gaussian {k, sigma, v0}, periodic  PeriodicGaussianPotential.hpp:21-44 */
    MJG_BASE_STACKING_3SPN2 = 9 /* 3SPN2BaseStacking, indices {S5, B5, B3, P5, S3, P3} (P5/P3 may be
                                   0xFFFFFFFF: absent), {stack kind 4 b5 + b3}
                                   ThreeSPN2BaseStackingInteraction.hpp:281-428 */
};
/* global pair potentials (read_global_potential.hpp:262-705) */
enum {
    MJG_PAIR_LJ = 1, /* LennardJones {sigma, epsilon} per participant */
    MJG_PAIR_EV = 2, /* ExcludedVolume {radius} per participant, common epsilon */
    MJG_PAIR_DH = 3, /* DebyeHuckel {charge} per participant */
    MJG_PAIR_3SPN2_EV = 4,       /* 3SPN2ExcludedVolume {bead kind P 0, S 1, A 2, T 3, G 4, C 5}
                                    ThreeSPN2ExcludedVolumePotential.hpp:47-64 */
    MJG_3SPN2_BASEPAIR = 5,      /* 3SPN2BasePair {base A 0 T 1 G 2 C 3, sugar index} per base
                                    ThreeSPN2BasePairInteraction.hpp:289-514 */
    MJG_3SPN2_CROSSSTACKING = 6  /* 3SPN2CrossStacking {base, sugar, 5' base, 3' base; -1 absent}
                                    ThreeSPN2CrossStackingInteraction.hpp:340-640 */
};
/* spatial_partition.type (read_spatial_partition.hpp:62-116) */
enum { MJG_PART_CELLLIST = 1, MJG_PART_VERLET = 2, MJG_PART_NAIVE = 3 };
/* ignore.molecule (IgnoreMolecule.hpp:101-121) */
enum { MJG_IGNORE_NOTHING = 0, MJG_IGNORE_SELF = 1, MJG_IGNORE_OTHERS = 2 };
/* integrator.type (read_integrator.hpp) */
enum { MJG_INTEG_BAOAB = 1, MJG_INTEG_GJF = 2, MJG_INTEG_GBAOAB = 3, MJG_INTEG_UNDERDAMPED = 4 };

/* Library version string. */
const char* mjg_version(void);

/* Create an engine for n_particles particles on HIP device `device`.
 * Replaces System(num_particles, boundary) (System.hpp:43-51) and the choice of
 * SimulatorTraits<precision, boundary> (read_input_file.hpp:22-145).
 * lower/upper are read only for MJG_PERIODIC (CuboidalPeriodicBoundary, BoundaryCondition.hpp:75-78). */
mjg_status mjg_engine_create(int precision, int boundary, const double* lower, const double* upper,
                             size_t n_particles, int device, mjg_engine** out);
void mjg_engine_destroy(mjg_engine* e);
/* Text of the last error on this engine ("" if none). */
const char* mjg_last_error(const mjg_engine* e);

/* System state: masses (rmass = 1/mass), positions and velocities, each 3*N doubles for
 * vectors (System.hpp:139-150).  Positions are wrapped into the box as System::initialize
 * does (System.hpp:63-67). */
mjg_status mjg_system_set_state(mjg_engine* e, const double* mass, const double* pos, const double* vel);
/* Positions only (3*N doubles), wrapped into the box as above: what a host-side integrator that
 * moved them hands to the force field before ForceFieldBase::calc_force(System&)
 * (BAOABLangevinIntegrator.hpp:195-220).  Masses, velocities, lists and their margins stay. */
mjg_status mjg_system_set_positions(mjg_engine* e, const double* pos);
/* Download positions, velocities and the current forces (any pointer may be NULL). */
mjg_status mjg_system_get_state(mjg_engine* e, double* pos, double* vel, double* force);
/* System attributes "temperature" [K] and "ionic_strength" [M] (System.hpp:164-167). */
mjg_status mjg_system_set_attribute(mjg_engine* e, const char* key, double value);

/* Add a local interaction ([[forcefields.local]]): n_terms terms of `arity` indices and
 * the potential's parameters (see the enum).  `topology` is the connection kind the
 * interaction writes into the topology ("bond", "contact", or "none"; write_topology,
 * BondLengthInteraction.hpp:232-245).  `margin` is used by MJG_CONTACT_GOCONTACT only.
 * The new interaction's id (its position in calc_energy / format_energy order) goes to *id. */
mjg_status mjg_add_local(mjg_engine* e, int kind, const char* topology, size_t n_terms,
                         const uint32_t* indices, const double* params, double margin, int* id);

/* Add a global pair interaction ([[forcefields.global]] interaction = "Pair"):
 * cutoff_ratio is the relative cutoff, epsilon the ExcludedVolume energy scale,
 * ignore.molecule, ignore.particles_within as (kind, distance) pairs, and the participants
 * with their per-particle parameters (parameters = [{index = ..., ...}]). */
mjg_status mjg_add_global(mjg_engine* e, int potential, double cutoff_ratio, double epsilon,
                          int partition, double margin, int ignore_molecule,
                          size_t n_within, const char* const* within_kinds, const uint32_t* within_dist,
                          size_t n_participants, const uint32_t* participants, const double* params,
                          int* id);

/* ConstraintForceField ([[forcefields.constraint]], read_constraint.hpp:12-46,
 * ConstraintForceField.hpp:15-81): n bond constraints, indices[2k], indices[2k+1] held at
 * v0[k], written into the topology with connection kind `topology` ("none": not written), with
 * the Gauss-Seidel sweep cap and tolerance the gBAOABLangevin integrator uses.  At most one set,
 * added before mjg_forcefield_initialize. */
mjg_status mjg_add_constraint(mjg_engine* e, const char* topology, size_t n, const uint32_t* indices,
                              const double* v0, size_t max_iteration, double tolerance);

/* ForceField::initialize (ForceField.hpp:47-63): topology and molecules from the local
 * interactions, exclusion lists, potential parameters, first neighbour lists. */
mjg_status mjg_forcefield_initialize(mjg_engine* e);
size_t mjg_num_interactions(const mjg_engine* e);

/* ForceFieldBase::calc_force (ForceField.hpp:101-109) on the device state; forces are
 * recomputed from zero. */
mjg_status mjg_calc_force(mjg_engine* e);
/* ForceFieldBase::calc_force_and_energy (ForceField.hpp:126-135); total energy to *energy. */
mjg_status mjg_calc_force_and_energy(mjg_engine* e, double* energy);
/* Per-interaction energies in interaction-id order (what format_energy prints,
 * ForceField.hpp:154-161) and their sum; does not change the state. */
mjg_status mjg_calc_energies(mjg_engine* e, double* per_interaction, double* total);
/* ForceFieldBase::calc_force_and_virial / calc_force_virial_energy (ForceFieldBase.hpp:34,37):
 * not provided.  No kernel accumulates the virial tensor, since no integrator on this path reads it
 * (only the reference's constant-pressure integrators do).  Returns MJG_UNSUPPORTED and fills the 9
 * entries of `virial` (if not NULL) with NaN, so that nothing downstream mistakes it for zero. */
mjg_status mjg_calc_force_and_virial(mjg_engine* e, double* virial);
/* Name of interaction `id` as the reference's name() prints it in the .ene header
 * (LocalInteractionBase::name, GlobalInteractionBase::name; e.g. "BondLength:Harmonic",
 * "GlobalPairLennardJones"); written NUL-terminated into buf (at most cap bytes), its length to *len. */
mjg_status mjg_interaction_name(const mjg_engine* e, int id, char* buf, size_t cap, size_t* len);
/* Force of interaction `id` alone, downloaded into force_out (3*N doubles). */
mjg_status mjg_calc_force_of(mjg_engine* e, int id, double* force_out);
/* ForceFieldBase::reduce_margin (ForceField.hpp:75-92): rebuild lists whose skin is spent. */
mjg_status mjg_reduce_margin(mjg_engine* e, double dmargin);
/* Current neighbour pairs (i < j) of the global interaction with id `id`, as the
 * reference's NeighborList holds them (NeighborList.hpp:151-266).  Writes min(n, capacity)
 * pairs (2 uint32 each) and the total count to *n_pairs. */
mjg_status mjg_neighbor_pairs(mjg_engine* e, int id, uint32_t* pairs, size_t capacity, size_t* n_pairs);

/* Integrator: type, delta_t, per-particle gammas (BAOAB, g-BAOAB, UnderdampedLangevin) or alphas
 * (G-JF), and the seed of the counter-based noise stream (read_integrator.hpp:18-307).  g-BAOAB
 * (gBAOABLangevinIntegrator.hpp:348-433) applies the constraints of mjg_add_constraint;
 * UnderdampedLangevin (UnderdampedLangevinIntegrator.hpp:459-542) keeps the acceleration
 * f/m + noise of the last force call on the device. */
mjg_status mjg_integrator_set(mjg_engine* e, int type, double dt, const double* coefs, uint64_t seed);
/* SystemMotionRemover(translation, rotation, rescale) of the integrator (SystemMotionRemover.hpp:
 * 21-148; TOML integrator.remove.{translation, rotation, rescale}, read_integrator.hpp:18-35):
 * after every step, remove the centre-of-mass velocity and the rigid rotation about the centre of
 * mass, then optionally rescale to the kinetic energy before the removal.  Default: all off. */
mjg_status mjg_integrator_set_remover(mjg_engine* e, int translation, int rotation, int rescale);
/* Integrator::initialize: coefficients from the temperature, then the first calc_force
 * (BAOABLangevinIntegrator.hpp:135-168, GJFNVTLangevinIntegrator.hpp:44-75). */
mjg_status mjg_integrator_initialize(mjg_engine* e);
/* Position of the counter-based noise stream: the step count the next step draws its noise at.
 * Restoring a checkpoint sets it (the reference restores its generator from the state string of
 * <prefix>_rng.msg, MsgPackSaver.hpp:70-88).  UnderdampedLangevin then redraws the stored
 * acceleration's noise at that count, as its initialize does after a restart (:404-457). */
mjg_status mjg_set_step_count(mjg_engine* e, uint64_t step);
/* n_steps calls of Integrator::step (BAOABLangevinIntegrator.hpp:170-240) on the device,
 * including the skin check and list rebuilds; returns after the work is complete. */
mjg_status mjg_step(mjg_engine* e, size_t n_steps);
/* Steps taken and neighbour-list rebuilds performed so far. */
mjg_status mjg_counters(mjg_engine* e, uint64_t* steps, uint64_t* rebuilds);
/* Kinetic energy sum_i m_i v_i^2 / 2 (EnergyObserver.hpp:89-150 prints it beside the terms). */
mjg_status mjg_kinetic_energy(mjg_engine* e, double* ke);
/* mjg_kinetic_energy without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The per-particle terms and the sum go to a static device array of the library (8 MiB
 * per GPU, held until the process ends) instead of the engine's scratch buffer; the kernels and
 * the sum order are mjg_kinetic_energy's, so the result is the same bit for bit.  At most 1048576
 * particles (MJG_INVALID_ARG above that).  Calls on engines of one GPU share the array: calls that
 * overlap in time on one GPU race and must be serialized by the caller. */
mjg_status mjg_kinetic_energy_static(mjg_engine* e, double* ke);

/* EnergyObserver (EnergyObserver.hpp:27-150): create prefix.ene (truncating it) and write the
 * unit line and the column header: "# timestep", one column per interaction named as the
 * reference names it, "kinetic_energy" and one "attribute:<key>" column per system attribute. */
mjg_status mjg_observer_energy_open(mjg_engine* e, const char* prefix);
/* EnergyObserver::output: append the line of step `step` (energies by ForceField::format_energy,
 * the kinetic energy, the attributes) in the reference's fixed-width format.  A non-finite
 * energy gives MJG_NAN_DETECTED after the line is written. */
mjg_status mjg_observer_energy_output(mjg_engine* e, uint64_t step);
/* XYZObserver (XYZObserver.hpp:24-107; output.format = "xyz", read_observer.hpp:28-34): create
 * (truncate) prefix_position.xyz and prefix_velocity.xyz; mjg_simulate appends one frame per
 * output ("N", "step = S", then "name x y z" per particle, fixed with 8 decimals).  names: one
 * string per particle, or NULL for "X" (read_system.hpp:277). */
mjg_status mjg_observer_xyz_open(mjg_engine* e, const char* prefix, const char* const* names);
/* DCDObserver (DCDObserver.hpp:27-341; output.format = "dcd", read_observer.hpp:35-41): create
 * (truncate) prefix_position.dcd and prefix_velocity.dcd; mjg_simulate writes the header
 * (save interval, total steps, number of molecules, dt, unit-cell flag), one float32 frame per
 * output (with the unit-cell record in the position file of a periodic box), and the frame
 * count at the end.  The 80-byte title record carries this library's own text. */
mjg_status mjg_observer_dcd_open(mjg_engine* e, const char* prefix);
/* MolecularDynamicsSimulator::initialize/run/finalize (MolecularDynamicsSimulator.hpp:72-123):
 * take total_step integrator steps; the open observers (XYZ, DCD, then energy, the order of
 * read_observer.hpp:84-98) output at every multiple of save_step and after the last step. */
mjg_status mjg_simulate(mjg_engine* e, uint64_t total_step, uint64_t save_step);

/* Timing hook with no counterpart in the reference (it has no GPU kernels to time): launch
 * the force kernel of interaction `id` `reps` times back to back on the engine's stream into a
 * scratch force buffer, bracketed by HIP events, and return the average kernel time in ms.
 * Also returns the number of half-list pairs the interaction currently holds (0 for locals). */
mjg_status mjg_profile_interaction(mjg_engine* e, int id, int reps, double* avg_ms, uint64_t* half_pairs);

/* ------------------------------------------------------------------------------------
 * Batched replicas.  The reference runs one System per process
 * (MolecularDynamicsSimulator holds a single system_, MolecularDynamicsSimulator.hpp:153-165);
 * a batch holds many independent ones -- each with the same description as an engine above --
 * and advances all resident replicas in one persistent launch per segment of steps, one
 * workgroup per replica (at most 2048 particles each).  Replicas can be added and retired
 * between calls; their state stays in device memory.  MJG_FP32_FIXED runs pair and local
 * arithmetic in FP32 with int64 fixed-point force sums and the FP64 integrator; MJG_FP64 runs
 * the same kernel in FP64.
 * ------------------------------------------------------------------------------------ */
typedef struct mjg_batch mjg_batch;

mjg_status mjg_batch_create(int precision, int device, mjg_batch** out);
void mjg_batch_destroy(mjg_batch* b);
const char* mjg_batch_last_error(const mjg_batch* b);
/* A new replica: System(num_particles, boundary) with masses, positions, velocities
 * (System.hpp:43-51,139-150); its id goes to *replica. */
mjg_status mjg_batch_add_replica(mjg_batch* b, int boundary, const double* lower, const double* upper, size_t n,
                                 const double* mass, const double* pos, const double* vel, int* replica);
/* System attributes of one replica (System.hpp:164-167). */
mjg_status mjg_batch_set_attribute(mjg_batch* b, int replica, const char* key, double value);
/* Local / global interactions of one replica; same meaning as mjg_add_local / mjg_add_global. */
mjg_status mjg_batch_add_local(mjg_batch* b, int replica, int kind, const char* topology, size_t n_terms,
                               const uint32_t* indices, const double* params);
mjg_status mjg_batch_add_global(mjg_batch* b, int replica, int potential, double cutoff_ratio, double epsilon,
                                int partition, double margin, int ignore_molecule, size_t n_within,
                                const char* const* within_kinds, const uint32_t* within_dist, size_t n_participants,
                                const uint32_t* participants, const double* params);
/* Integrator of one replica (as mjg_integrator_set). */
mjg_status mjg_batch_set_integrator(mjg_batch* b, int replica, int type, double dt, const double* coefs, uint64_t seed);
/* ForceField::initialize + Integrator::initialize for every replica added since the last call
 * (ForceField.hpp:47-63, BAOABLangevinIntegrator.hpp:135-168); they become resident. */
mjg_status mjg_batch_initialize(mjg_batch* b);
/* Every resident replica takes n_steps Integrator::step calls (BAOABLangevinIntegrator.hpp:170-240).
 * A replica whose positions or force contributions stop being finite is stopped at that step and
 * takes no further steps; the other replicas complete the call, which then returns
 * MJG_NAN_DETECTED (mjg_batch_last_error names the replicas; mjg_batch_replica_status tells each
 * one's).  The reference has one system per process and only logs a NaN energy
 * (GlobalForceField.hpp:181-186); here one failing replica never stops the others.  The same holds
 * for mjg_batch_initialize, mjg_batch_advance and mjg_batch_profile. */
mjg_status mjg_batch_step(mjg_batch* b, size_t n_steps);
/* Steps one replica still has to take under mjg_batch_advance. */
mjg_status mjg_batch_set_steps(mjg_batch* b, int replica, uint64_t steps);
/* Position of one replica's noise stream, as mjg_set_step_count for an engine: a replica restored
 * from a checkpoint continues the stream (MsgPackSaver.hpp:70-88 saves the generator's state). Its
 * lists, and an UnderdampedLangevin replica's acceleration, are rebuilt at its next step. */
mjg_status mjg_batch_set_step_count(mjg_batch* b, int replica, uint64_t step);
/* One launch of up to `segment` steps for every resident replica with steps remaining. */
mjg_status mjg_batch_advance(mjg_batch* b, int segment);
/* Free a replica's device memory. */
mjg_status mjg_batch_retire(mjg_batch* b, int replica);
mjg_status mjg_batch_get_state(mjg_batch* b, int replica, double* pos, double* vel, double* force);
/* Launch parameters of a replica that change no result (forces are exact fixed-point sums over the
 * pairs inside the cutoff, whatever list holds them), with no counterpart in the reference (which
 * fixes its cell size and skin from the input, PeriodicGridCellList.hpp:133-169,
 * read_spatial_partition.hpp:62-116): the replica kernel's own list skin (relative margin, at most
 * the input's) for LJ / excluded-volume pairs and for Debye-Hueckel, the cell-grid resolution of its
 * list builds (grid_bits 4: 16 cells per dimension, 3: 8) and the lanes sharing one row of such a
 * build (1, 2 or 4).  By default mjg_batch_initialize picks them with a measured search: replicas of
 * one size bucket (half an octave), size class, box kind and interaction make-up share the result,
 * found once per batch by timing a short segment of a clone of the first such replica for each
 * candidate.  mjg_batch_set_autotune(b, 0) (or MJG_AUTOTUNE=0) uses fixed defaults instead;
 * mjg_batch_set_tuning fixes one replica's (before mjg_batch_initialize). */
mjg_status mjg_batch_set_autotune(mjg_batch* b, int on);
mjg_status mjg_batch_set_tuning(mjg_batch* b, int replica, double list_skin_short, double list_skin_dh, int grid_bits,
                                int grid_lanes);
mjg_status mjg_batch_replica_tuning(mjg_batch* b, int replica, double* list_skin_short, double* list_skin_dh,
                                    int* grid_bits, int* grid_lanes);
/* MJG_OK for a replica that steps normally, MJG_NAN_DETECTED for one stopped with non-finite state
 * (it keeps its last state until mjg_batch_retire). */
mjg_status mjg_batch_replica_status(mjg_batch* b, int replica, int* status);
/* Steps taken, steps remaining, list rebuilds and device bytes of one replica. */
mjg_status mjg_batch_replica_info(mjg_batch* b, int replica, uint64_t* step, uint64_t* remaining, uint64_t* rebuilds,
                                  size_t* footprint_bytes);
/* Current neighbour-list entries of each global interaction of one replica, in the order they
 * were added (the size of SpatialPartitionBase::neighbors(), SpatialPartitionBase.hpp:138-139);
 * writes min(max_globals, number of globals) counts. */
mjg_status mjg_batch_list_sizes(mjg_batch* b, int replica, uint64_t* entries, int max_globals);
/* Forces and per-interaction energies of one replica at its current positions
 * (calc_force_and_energy / format_energy, ForceField.hpp:126-161). */
mjg_status mjg_batch_evaluate(mjg_batch* b, int replica, double* per_interaction, double* total, double* force);
/* EnergyObserver::output (EnergyObserver.hpp:89-150) for every resident replica in one launch:
 * per-interaction energies and the kinetic energy at the current state.  The forces are recomputed
 * from the same lists, so the trajectories are unchanged. */
mjg_status mjg_batch_observe(mjg_batch* b);
/* The last observation of one replica: per-interaction energies (locals, then globals, in the
 * order they were added) and the kinetic energy sum of m |v|^2 / 2. */
mjg_status mjg_batch_energies(mjg_batch* b, int replica, double* per_interaction, double* kinetic);
/* Sum over replicas of particles x steps taken, and list rebuilds. */
mjg_status mjg_batch_counters(mjg_batch* b, uint64_t* particle_steps, uint64_t* rebuilds);
/* Timing hook with no counterpart in the reference: `reps` launches of a `segment_steps` segment
 * over every resident replica, bracketed by HIP events; average launch time in ms and the
 * particle-steps of one launch. */
mjg_status mjg_batch_profile(mjg_batch* b, int segment_steps, int reps, double* avg_ms, uint64_t* particle_steps);

#ifdef __cplusplus
}
#endif
#endif
