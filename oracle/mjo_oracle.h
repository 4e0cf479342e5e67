/* CPU restatement of Mjolnir's force-and-step path, in plain C99.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke entry point and
 * the CPU-baseline leg of the benchmark call it to compare against the HIP engine.  The
 * product (libmjg.so) never links, loads or calls it.
 *
 * Every routine restates one piece of the reference and cites the file:line it follows
 * (paths relative to the Mjolnir source tree).  Arithmetic is written in the reference's
 * operation order and the library is built with -ffp-contract=off, so on the same inputs its
 * forces and energies agree bit for bit with the reference's own classes built the same way
 * (oracle/ref_driver.cpp); the tests check exactly that against committed golden vectors.
 */
#ifndef MJO_ORACLE_H
#define MJO_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* local interaction kinds (interaction:potential) */
enum {
    ORC_BOND_HARMONIC      = 1, /* BondLength:Harmonic      BondLengthInteraction.hpp:126-148 */
    ORC_BOND_GOCONTACT     = 2, /* BondLengthGoContact      BondLengthGoContactInteraction.hpp:126-167 */
    ORC_CONTACT_GOCONTACT  = 3, /* Contact:GoContact        GoContactInteraction.hpp:159-228 */
    ORC_ANGLE_HARMONIC     = 4, /* BondAngle:Harmonic       BondAngleInteraction.hpp:129-189 */
    ORC_ANGLE_GAUSSIAN     = 5, /* BondAngle:Gaussian       BondAngleInteraction.hpp:129-189 */
    ORC_DIHEDRAL_CLEMENTI  = 6, /* DihedralAngle:ClementiDihedral DihedralAngleInteraction.hpp:157-227 */
    ORC_BOND_3SPN2         = 7, /* BondLength:3SPN2Bond {k, v0}  ThreeSPN2BondPotential.hpp:22-35 */
    ORC_DIHEDRAL_PGAUSSIAN = 8, /* DihedralAngle:Gaussian {k, sigma, v0}, PeriodicGaussianPotential.hpp:21-44 */
    ORC_3SPN2_STACKING     = 9  /* 3SPN2BaseStacking {stack kind}, indices (S5, B5, B3, P5, S3, P3), P may be
                                   ORC_NIL; ThreeSPN2BaseStackingInteraction.hpp:127-428 */
};
/* global pair potentials */
enum {
    ORC_PAIR_LJ = 1,  /* GlobalPairLennardJonesInteraction.hpp:148-198 */
    ORC_PAIR_EV = 2,  /* GlobalPairExcludedVolumeInteraction.hpp:148-195 */
    ORC_PAIR_DH = 3,  /* GlobalPairInteraction.hpp:148-213 + DebyeHuckelPotential.hpp:47-65 */
    ORC_PAIR_3SPN2_EV = 4,   /* Pair:3SPN2ExcludedVolume {bead kind P,S,A,T,G,C = 0..5}
                                ThreeSPN2ExcludedVolumePotential.hpp:27-446 */
    ORC_3SPN2_BASEPAIR = 5,  /* 3SPN2BasePair {base A,T,G,C = 0..3, sugar index}
                                ThreeSPN2BasePairInteraction.hpp:130-514 */
    ORC_3SPN2_CROSSSTACK = 6 /* 3SPN2CrossStacking {base, sugar, 5' base, 3' base (ORC_NIL if none)}
                                ThreeSPN2CrossStackingInteraction.hpp:133-640 */
};
enum { ORC_PART_CELLLIST = 1, ORC_PART_VERLET = 2, ORC_PART_NAIVE = 3 };
#define ORC_NIL 0xffffffffu   /* absent index (a nucleotide without phosphate, a strand end) */
enum { ORC_IGNORE_NOTHING = 0, ORC_IGNORE_SELF = 1, ORC_IGNORE_OTHERS = 2 };
enum { ORC_INTEG_BAOAB = 1, ORC_INTEG_GJF = 2, ORC_INTEG_GBAOAB = 3, ORC_INTEG_UNDERDAMPED = 4 };

typedef struct orc_engine orc_engine;

/* physical constants after read_units("kcal/mol", "angstrom") (read_units.hpp:32-92) */
double orc_kB(void);
double orc_eps0(void);
double orc_L_to_volume(void);

/* Philox4x32-10 (Salmon et al., SC'11) and the engine's noise stream built on it */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_noise3(uint64_t seed, uint64_t step, uint32_t particle, double out[3]);

orc_engine* orc_create(int periodic, const double* lower, const double* upper, size_t n);
void orc_destroy(orc_engine* e);
int  orc_set_state(orc_engine* e, const double* mass, const double* pos, const double* vel);
int  orc_get_state(const orc_engine* e, double* pos, double* vel, double* force);
int  orc_set_attribute(orc_engine* e, const char* key, double value);

/* returns the interaction id (>= 0) or -1 */
int orc_add_local(orc_engine* e, int kind, const char* topology, size_t n_terms,
                  const uint32_t* indices, const double* params, double margin);
int orc_add_global(orc_engine* e, int potential, double cutoff_ratio, double epsilon,
                   int partition, double margin, int ignore_molecule,
                   size_t n_within, const char* const* within_kinds, const uint32_t* within_dist,
                   size_t n_participants, const uint32_t* participants, const double* params);

/* ConstraintForceField (ConstraintForceField.hpp:15-81, read_constraint.hpp:12-46): bond
 * constraints (i, j, v0) for gBAOABLangevin, with the connection kind they write into the
 * topology ("none": none), the iteration cap and the tolerance; one set per engine */
int orc_add_constraint(orc_engine* e, const char* topology, size_t n, const uint32_t* indices, const double* v0,
                       size_t max_iteration, double tolerance);

/* ForceField::initialize (ForceField.hpp:47-63): topology, molecules, exclusions, lists */
int orc_initialize(orc_engine* e);

size_t orc_num_interactions(const orc_engine* e);
/* calc_energy of one interaction (locals first, then globals, in insertion order) */
double orc_calc_energy_of(orc_engine* e, int id);
/* calc_force of one interaction, added into force_out (3N) */
int orc_calc_force_of(orc_engine* e, int id, double* force_out);
/* ForceField::calc_force_and_energy on the engine's own force array; returns the energy */
double orc_calc_force_and_energy(orc_engine* e);
/* neighbour list of global interaction `gid` (0-based among globals): pairs (i, j) */
size_t orc_pairs_of(const orc_engine* e, int gid, uint32_t* out, size_t cap);
/* reduce_margin on all interactions (ForceField.hpp:75-92); returns #rebuilt lists */
int orc_reduce_margin(orc_engine* e, double dmargin);

int orc_set_integrator(orc_engine* e, int type, double dt, const double* coefs, uint64_t seed);
/* SystemMotionRemover(translation, rotation, rescale), applied after every step */
int orc_set_remover(orc_engine* e, int translation, int rotation, int rescale);
int orc_integrator_initialize(orc_engine* e);
int orc_step(orc_engine* e, size_t nsteps);
uint64_t orc_step_count(const orc_engine* e);
size_t orc_rebuild_count(const orc_engine* e);

#ifdef __cplusplus
}
#endif
#endif
