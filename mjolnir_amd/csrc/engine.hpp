// Host side of the engine: device-resident state for one system, the interaction tables
// uploaded from the reference-style descriptors, and the launch sequences for the force call
// and the integrator step.
#pragma once
#include "device_common.hpp"
#include "force_kernels.hpp"
#include "integrate_kernels.hpp"
#include "constraint_kernels.hpp"
#include "nlist_kernels.hpp"
#include "tile_kernels.hpp"

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace mjg
{

struct DeviceError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};
struct NonFinite : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

struct LocalIA
{
    int kind = 0, arity = 0, npar = 0, nbody = 0;
    int topo_kind = -1;           // interned connection kind, -1 for "none"
    std::vector<uint32_t> idx;    // n_terms * arity
    std::vector<double> par;      // n_terms * npar
    double margin = 0.5;
    uint32_t slot_off = 0, term_off = 0;
    size_t n_terms() const {return arity ? idx.size() / arity : 0;}
};

struct GlobalIA
{
    int pot = 0, partition = 0, ignore_mol = 0;
    double cutoff_ratio = 0, eps = 0, margin = 0;
    std::vector<std::pair<int, uint32_t>> within;  // (interned kind, distance)
    std::vector<uint32_t> part;                    // participants in input order
    std::vector<double> par;                       // per participant, npar columns
    int npar = 0;

    // derived at initialize
    PairParams pp{};
    double max_cutoff = 0, r_list = 0;
    ListRule rule{};
    Grid grid{};          // periodic grid (host copy)

    // device
    uint32_t *d_part = nullptr, *d_pslot = nullptr, *d_mol = nullptr, *d_ex_start = nullptr, *d_ex = nullptr;
    double* d_ppar = nullptr;
    Grid* d_grid = nullptr;
    CellSort sb;
    uint32_t *d_cell_start = nullptr, *d_cell_end = nullptr;
    uint32_t *d_nl = nullptr, *d_nn = nullptr, *d_order = nullptr, *d_spart = nullptr;
    double4* d_spos = nullptr;    // position of each sorted slot at the last cell-list build
    uint32_t cap = 0;             // row length of the list: nl[t * cap + k]
    uint32_t tpp = 8;             // lanes per participant in k_verlet / k_pairs
    uint32_t mb_tpp = 16;         // lanes per participant in k_tsp_pairs (one entry per lane)
    bool small = false;           // one-workgroup rebuild (np <= kSmallMaxPart)
    double* d_epair = nullptr;
    double4* d_fg = nullptr;      // this interaction's force on every particle
    unsigned long long* d_acc = nullptr;   // 3SPN2 base pairing / cross stacking: int64 fixed-point sums
    // cell-tile pair kernel (periodic cell lists): staging slot of every list entry and the
    // staging capacity in particles
    bool tiled = false;
    uint16_t* d_nl16 = nullptr;
    uint32_t tile_cap = 0;
    bool tile_fixed = false;
    bool multibody() const {return pot == kBasePair || pot == kCrossStack;}
};

// ConstraintForceField (ConstraintForceField.hpp:15-81)
struct ConstraintSet
{
    bool set = false;
    int topo_kind = -1;
    std::vector<uint32_t> idx;    // 2 per constraint
    std::vector<double> v0;
    unsigned long long max_it = 0;
    double tol = 0;
};

class Engine
{
  public:
    Engine(int precision, int boundary, const double* lower, const double* upper, size_t n, int device);
    ~Engine();

    void set_state(const double* mass, const double* pos, const double* vel);
    void set_positions(const double* pos);
    void get_state(double* pos, double* vel, double* force);
    void set_attribute(const std::string& key, double value);
    int add_local(int kind, const char* topology, size_t n_terms, const uint32_t* idx, const double* par, double margin);
    int add_global(int pot, double cutoff_ratio, double eps, int partition, double margin, int ignore_mol,
                   size_t n_within, const char* const* kinds, const uint32_t* dists,
                   size_t np, const uint32_t* part, const double* par);
    void add_constraint(const char* topology, size_t n, const uint32_t* idx, const double* v0, size_t max_it,
                        double tol);
    void initialize();
    size_t num_interactions() const {return locals_.size() + globals_.size();}
    size_t num_locals() const {return locals_.size();}

    void calc_force();
    double calc_force_and_energy();
    void calc_energies(double* per, double* total);
    void calc_force_of(int id, double* out);
    void reduce_margin(double dmargin);
    size_t neighbor_pairs(int id, uint32_t* pairs, size_t cap);

    void set_integrator(int type, double dt, const double* coefs, uint64_t seed);
    void set_remover(int translation, int rotation, int rescale);
    void integrator_initialize();
    void step(size_t n);
    void counters(uint64_t* steps, uint64_t* rebuilds);
    void set_step_count(uint64_t step);
    double kinetic_energy();
    double kinetic_energy_static();   // same sum, in the library's static device scratch
    // observers and the simulator loop
    std::string interaction_name(size_t id) const;
    void energy_observer_open(const std::string& prefix);
    bool energy_observer_output(uint64_t step);
    void xyz_observer_open(const std::string& prefix, const char* const* names);
    void dcd_observer_open(const std::string& prefix);
    void simulate(uint64_t total_step, uint64_t save_step);
    double profile_interaction(int id, int reps, uint64_t* half_pairs);

    std::string last_error;

  private:
    int kind_id(const std::string& name);
    void write_topology();
    void make_exclusions(GlobalIA& g);
    void setup_global(GlobalIA& g);
    void alloc_list(GlobalIA& g, uint32_t cap);
    void launch_rebuilds(int force);
    void build_lists_now(bool all);
    ListSeg list_seg(GlobalIA& g, size_t gi);
    ListSegTable list_table();
    PairSegTable pair_table(int only);
    void launch_multibody(const GlobalIA& g, bool energy);
    void launch_tiled(GlobalIA& g, int gi, bool energy);
    bool grow_tiles();
    uint32_t tile_lanes(const GlobalIA& g) const;
    void forces(double4* f, bool energy, int only, bool kick, bool advance = false);
    void launch_ul_accel(bool kick);
    void launch_step();
    void setup_constraints();
    void run_segment(size_t n);
    void sync();
    void check(hipError_t e, const char* what);
    template<typename T> T* dalloc(size_t n);
    template<typename T> void upload(T* d, const T* h, size_t n);
    template<typename T> void download(T* h, const T* d, size_t n);

    int precision_, device_;
    size_t n_;
    Box box_{};
    hipStream_t stream_ = nullptr;
    std::vector<void*> allocations_;

    std::vector<double> mass_;
    bool has_T_ = false, has_I_ = false, initialized_ = false, integ_ready_ = false, state_set_ = false;
    double T_ = 0, I_ = 0;
    std::string ene_path_;          // open energy observer ("" if none)
    std::string xyz_prefix_, dcd_prefix_;   // open trajectory observers ("" if none)
    std::vector<std::string> names_;        // particle names of the XYZ frames
    uint64_t dcd_frames_ = 0;
    uint32_t n_molecules_ = 0;              // Topology::number_of_molecules
    void trajectory_output(uint64_t step);
    void dcd_begin(uint64_t total_step, uint64_t save_step);
    void dcd_end();

    double4 *d_pos_ = nullptr, *d_vel_ = nullptr, *d_force_ = nullptr, *d_force_tmp_ = nullptr;
    double4 *d_bak_pos_ = nullptr, *d_bak_vel_ = nullptr, *d_bak_force_ = nullptr;
    double *d_mass_ = nullptr, *d_rmass_ = nullptr;

    std::vector<LocalIA> locals_;
    std::vector<GlobalIA> globals_;
    std::vector<std::string> kinds_;
    std::vector<std::vector<std::pair<uint32_t, int>>> adj_;
    std::vector<uint32_t> mol_;

    // local tables
    uint32_t* d_lidx_ = nullptr;
    double* d_lpar_ = nullptr;
    double* d_tf_ = nullptr;
    double* d_te_ = nullptr;
    uint32_t *d_gstart_ = nullptr, *d_gslot_ = nullptr;
    LocalSegTable segtab_{};
    uint32_t local_blocks_ = 0, n_slots_ = 0, n_terms_ = 0;

    // integrator
    int integ_ = 0;
    int rm_translation_ = 0, rm_rotation_ = 0, rm_rescale_ = 0;   // SystemMotionRemover
    double dt_ = 0;
    uint64_t seed_ = 0;
    std::vector<double> coefs_;
    double *d_c1_ = nullptr, *d_c2_ = nullptr, *d_c3_ = nullptr;
    double4 *d_accel_ = nullptr, *d_bak_accel_ = nullptr;   // UnderdampedLangevin acceleration

    // g-BAOAB constraints
    ConstraintSet con_;
    ConTable ctab_{};
    double4 *d_old_pos_ = nullptr, *d_old_rattle_ = nullptr;
    double dt_corr_ = 0;

    // control
    StepCtl* d_sctl_ = nullptr;
    StepCtl* d_bak_sctl_ = nullptr;
    ListCtl* d_lctl_ = nullptr;
    ListCtl* d_bak_lctl_ = nullptr;
    double* d_abs_margin_ = nullptr;
    int* d_never_ = nullptr;
    double* d_energy_ = nullptr;    // per interaction + total
    double* d_scratch_ = nullptr;   // N doubles
    size_t scratch_n_ = 0;

    // graph of one step segment
    hipGraphExec_t graph_exec_ = nullptr;
    size_t graph_steps_ = 0;
    bool use_graph_ = true;
};

} // mjg
