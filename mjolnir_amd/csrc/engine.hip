// Host side of the engine (see engine.hpp) and the C ABI declared in include/mjg.h.
#include "engine.hpp"
#include "observers.hpp"
#include "../../include/mjg.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <mutex>
#include <new>
#include <sstream>

namespace mjg
{

// ------------------------------------------------------------------------------------
// physical constants after read_units("kcal/mol", "angstrom") (read_units.hpp:32-92,
// Unit.hpp:23-44,150-159), formed in the same operation order
// ------------------------------------------------------------------------------------
static double kB_kcal()
{
    const double J_to_cal = 1.0 / 4.184;
    return 1.38064852e-23 * (J_to_cal / 1000.0) * 6.022140857e23;
}
static double eps0_kcal_angstrom()
{
    const double J_to_cal = 1.0 / 4.184;
    double eps0 = (8.854187817e-12 / 1.6021766208e-19) / 1.6021766208e-19;
    eps0 = eps0 * (1000.0 / J_to_cal) / 6.022140857e23;
    return eps0 / 1e+10;
}
static double L_to_volume_angstrom() {return 1e-3 * std::pow(1e+10, 3);}

static uint32_t blocks_for(size_t n, uint32_t bs) {return (uint32_t)std::max<size_t>(1, (n + bs - 1) / bs);}

void Engine::check(hipError_t e, const char* what)
{
    if(e != hipSuccess)
    {
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
    }
}

template<typename T> T* Engine::dalloc(size_t n)
{
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(1, n) * sizeof(T);
    const hipError_t e = hipMalloc(&p, bytes);
    if(e != hipSuccess)
    {
        (void)hipGetLastError();
        throw std::bad_alloc();
    }
    allocations_.push_back(p);
    return static_cast<T*>(p);
}
template<typename T> void Engine::upload(T* d, const T* h, size_t n)
{
    if(n == 0) {return;}
    check(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, stream_), "upload");
    check(hipStreamSynchronize(stream_), "upload sync");
}
template<typename T> void Engine::download(T* h, const T* d, size_t n)
{
    if(n == 0) {return;}
    check(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, stream_), "download");
    check(hipStreamSynchronize(stream_), "download sync");
}
void Engine::sync() {check(hipStreamSynchronize(stream_), "synchronize");}

Engine::Engine(int precision, int boundary, const double* lower, const double* upper, size_t n, int device)
    : precision_(precision), device_(device), n_(n)
{
    if(precision != MJG_FP64 && precision != MJG_FP32_FIXED) {throw std::invalid_argument("unknown precision");}
    if(boundary != MJG_UNLIMITED && boundary != MJG_PERIODIC) {throw std::invalid_argument("unknown boundary");}
    if(n == 0 || n >= 0xffffffffull) {throw std::invalid_argument("particle count out of range");}
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    box_.periodic = (boundary == MJG_PERIODIC);
    if(box_.periodic)
    {
        if(lower == nullptr || upper == nullptr) {throw std::invalid_argument("periodic boundary needs lower/upper");}
        // CuboidalPeriodicBoundary::set_boundary (BoundaryCondition.hpp:115-123)
        for(int d = 0; d < 3; ++d)
        {
            if(!(upper[d] > lower[d])) {throw std::invalid_argument("upper must exceed lower");}
            box_.lower[d] = lower[d]; box_.upper[d] = upper[d];
            box_.width[d] = upper[d] - lower[d];
            box_.halfw[d] = box_.width[d] * 0.5;
        }
    }
    d_pos_ = dalloc<double4>(n); d_vel_ = dalloc<double4>(n);
    d_force_ = dalloc<double4>(n); d_force_tmp_ = dalloc<double4>(n);
    d_bak_pos_ = dalloc<double4>(n); d_bak_vel_ = dalloc<double4>(n); d_bak_force_ = dalloc<double4>(n);
    d_mass_ = dalloc<double>(n); d_rmass_ = dalloc<double>(n);
    d_sctl_ = dalloc<StepCtl>(1); d_bak_sctl_ = dalloc<StepCtl>(1);
    StepCtl z{};
    upload(d_sctl_, &z, 1);
    scratch_n_ = n;
    d_scratch_ = dalloc<double>(n);
    // Steps are captured into a hipGraph per segment, except with MJG_NO_GRAPH=1 and under a
    // preloaded profiler library (rocprofv3 preloads librocprofiler-sdk-tool): a kernel-trace run
    // of the graph path once ended in a segmentation fault inside the runtime and the profiler,
    // whose cause was not found, so profiled runs launch the same kernels on the stream instead.
    const char* ng = std::getenv("MJG_NO_GRAPH");
    const char* pre = std::getenv("LD_PRELOAD");
    const bool profiled = pre && std::strstr(pre, "rocprof") != nullptr;
    use_graph_ = !(ng && ng[0] == '1') && !profiled;
}

Engine::~Engine()
{
    if(graph_exec_) {(void)hipGraphExecDestroy(graph_exec_);}
    if(stream_) {(void)hipStreamSynchronize(stream_);}
    for(void* p : allocations_) {(void)hipFree(p);}
    if(stream_) {(void)hipStreamDestroy(stream_);}
}

void Engine::set_state(const double* mass, const double* pos, const double* vel)
{
    if(!mass || !pos || !vel) {throw std::invalid_argument("null state pointer");}
    mass_.assign(mass, mass + n_);
    std::vector<double> rm(n_);
    std::vector<double4> p(n_), v(n_);
    for(size_t i = 0; i < n_; ++i)
    {
        if(!(mass[i] > 0.0)) {throw std::invalid_argument("masses must be positive");}
        rm[i] = 1.0 / mass[i];
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        if(box_.periodic)
        {
            // System::initialize wraps positions (System.hpp:63-67, BoundaryCondition.hpp:93-102)
            if(x < box_.lower[0]) {x += box_.width[0];} else if(x >= box_.upper[0]) {x -= box_.width[0];}
            if(y < box_.lower[1]) {y += box_.width[1];} else if(y >= box_.upper[1]) {y -= box_.width[1];}
            if(z < box_.lower[2]) {z += box_.width[2];} else if(z >= box_.upper[2]) {z -= box_.width[2];}
        }
        p[i] = make_double4(x, y, z, 0.0);
        v[i] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0);
    }
    upload(d_mass_, mass_.data(), n_);
    upload(d_rmass_, rm.data(), n_);
    upload(d_pos_, p.data(), n_);
    upload(d_vel_, v.data(), n_);
    if(d_old_pos_) {upload(d_old_pos_, p.data(), n_);}   // g-BAOAB: the last step end is now
    std::vector<double4> zf(n_, make_double4(0, 0, 0, 0));
    upload(d_force_, zf.data(), n_);
    state_set_ = true;
}

// Positions only (a host integrator moved them): wrapped into the box as set_state does; masses,
// velocities, forces, neighbour lists and their margins are left as they are.
void Engine::set_positions(const double* pos)
{
    if(!pos) {throw std::invalid_argument("null position pointer");}
    if(!state_set_) {throw std::logic_error("set the state first");}
    std::vector<double4> p(n_);
    for(size_t i = 0; i < n_; ++i)
    {
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        if(box_.periodic)
        {
            if(x < box_.lower[0]) {x += box_.width[0];} else if(x >= box_.upper[0]) {x -= box_.width[0];}
            if(y < box_.lower[1]) {y += box_.width[1];} else if(y >= box_.upper[1]) {y -= box_.width[1];}
            if(z < box_.lower[2]) {z += box_.width[2];} else if(z >= box_.upper[2]) {z -= box_.width[2];}
        }
        p[i] = make_double4(x, y, z, 0.0);
    }
    upload(d_pos_, p.data(), n_);
}

void Engine::get_state(double* pos, double* vel, double* force)
{
    std::vector<double4> h(n_);
    double4* src[3] = {d_pos_, d_vel_, d_force_};
    double* dst[3] = {pos, vel, force};
    for(int k = 0; k < 3; ++k)
    {
        if(!dst[k]) {continue;}
        download(h.data(), src[k], n_);
        for(size_t i = 0; i < n_; ++i) {dst[k][3 * i] = h[i].x; dst[k][3 * i + 1] = h[i].y; dst[k][3 * i + 2] = h[i].z;}
    }
}

void Engine::set_attribute(const std::string& key, double value)
{
    if(key == "temperature") {T_ = value; has_T_ = true;}
    else if(key == "ionic_strength") {I_ = value; has_I_ = true;}
    else {throw std::invalid_argument("unknown attribute: " + key);}
}

int Engine::kind_id(const std::string& name)
{
    for(size_t i = 0; i < kinds_.size(); ++i) {if(kinds_[i] == name) {return (int)i;}}
    kinds_.push_back(name);
    return (int)kinds_.size() - 1;
}

int Engine::add_local(int kind, const char* topology, size_t n_terms, const uint32_t* idx, const double* par, double margin)
{
    if(initialized_) {throw std::invalid_argument("interactions must be added before initialize");}
    if(!globals_.empty())
    {
        // interaction ids follow calc_energy order: every local, then every global
        throw std::invalid_argument("add local interactions before global ones");
    }
    LocalIA L;
    L.kind = kind;
    switch(kind)
    {
        case kBondHarmonic: case kBondGo: case kContactGo: L.arity = 2; L.npar = 2; break;
        case kAngleHarmonic: L.arity = 3; L.npar = 2; break;
        case kAngleGaussian: L.arity = 3; L.npar = 3; break;
        case kDihedralClementi: L.arity = 4; L.npar = 3; break;
        case kBond3SPN2: L.arity = 2; L.npar = 2; break;
        case kDihedralPGauss: L.arity = 4; L.npar = 3; break;
        case kBaseStacking: L.arity = 6; L.npar = 1; break;
        default: throw std::invalid_argument("unknown local interaction kind");
    }
    // base stacking reads {Si, Bi, Bj} and keeps the phosphates for its topology only
    L.nbody = (kind == kBaseStacking) ? 3 : L.arity;
    if(n_terms > 0 && (!idx || !par)) {throw std::invalid_argument("null term arrays");}
    for(size_t t = 0; t < n_terms * L.arity; ++t)
    {
        const bool optional = (kind == kBaseStacking) && (t % 6 == 3 || t % 6 == 5);
        if(idx[t] >= n_ && !(optional && idx[t] == kNil))
        {
            throw std::out_of_range("local interaction index out of range");
        }
    }
    if(kind == kBaseStacking)
    {
        for(size_t t = 0; t < n_terms; ++t)
        {
            if(!(par[t] >= 0 && par[t] < 16)) {throw std::invalid_argument("3SPN2 stacking kind out of range");}
        }
    }
    if(locals_.size() + 1 >= (size_t)kMaxLocalSegs) {throw std::invalid_argument("too many local interactions");}
    const std::string topo = topology ? topology : "";
    L.topo_kind = (topo.empty() || topo == "none") ? -1 : kind_id(topo);
    L.idx.assign(idx, idx + n_terms * L.arity);
    L.par.assign(par, par + n_terms * L.npar);
    L.margin = margin;
    locals_.push_back(std::move(L));
    return (int)locals_.size() - 1;
}

int Engine::add_global(int pot, double cutoff_ratio, double eps, int partition, double margin, int ignore_mol,
                       size_t n_within, const char* const* kinds, const uint32_t* dists,
                       size_t np, const uint32_t* part, const double* par)
{
    if(initialized_) {throw std::invalid_argument("interactions must be added before initialize");}
    if(globals_.size() >= (size_t)kMaxGlobals) {throw std::invalid_argument("too many global interactions");}
    GlobalIA G;
    G.pot = pot;
    switch(pot)
    {
        case kPairLJ: G.npar = 2; break;
        case kPairEV: case kPairDH: case kPair3SPN2EV: G.npar = 1; break;
        case kBasePair: G.npar = 2; break;
        case kCrossStack: G.npar = 4; break;
        default: throw std::invalid_argument("unknown pair potential");
    }
    if(partition != MJG_PART_CELLLIST && partition != MJG_PART_VERLET && partition != MJG_PART_NAIVE)
    {
        throw std::invalid_argument("unknown spatial partition");
    }
    if(ignore_mol < 0 || ignore_mol > 2) {throw std::invalid_argument("unknown ignore.molecule");}
    if(np < 1 || !part || !par) {throw std::invalid_argument("a pair interaction needs participants");}
    if(!(cutoff_ratio > 0.0) || !(margin >= 0.0)) {throw std::invalid_argument("cutoff and margin must be positive");}
    for(size_t k = 0; k < np; ++k) {if(part[k] >= n_) {throw std::out_of_range("participant index out of range");}}
    G.cutoff_ratio = cutoff_ratio; G.eps = eps; G.partition = partition; G.margin = margin; G.ignore_mol = ignore_mol;
    for(size_t k = 0; k < n_within; ++k)
    {
        if(!kinds || !kinds[k] || !dists) {throw std::invalid_argument("null ignore.particles_within entry");}
        G.within.emplace_back(kind_id(kinds[k]), dists[k]);
    }
    G.part.assign(part, part + np);
    G.par.assign(par, par + np * G.npar);
    for(size_t k = 0; k < np; ++k)
    {
        const double* p = G.par.data() + k * G.npar;
        if(pot == kPair3SPN2EV && !(p[0] >= 0 && p[0] < 6)) {throw std::invalid_argument("3SPN2 bead kind out of range");}
        if(pot == kBasePair || pot == kCrossStack)
        {
            if(!(p[0] >= 0 && p[0] < 4)) {throw std::invalid_argument("3SPN2 base kind out of range");}
            if(!(p[1] >= 0 && p[1] < (double)n_)) {throw std::out_of_range("3SPN2 sugar index out of range");}
        }
        if(pot == kCrossStack)
        {
            for(int c = 2; c < 4; ++c)
            {
                if(!(p[c] == -1.0 || (p[c] >= 0 && p[c] < (double)n_)))
                {
                    throw std::out_of_range("3SPN2 neighbouring base index out of range");
                }
            }
        }
    }
    globals_.push_back(std::move(G));
    return (int)globals_.size() - 1;
}

// ConstraintForceField (read_constraint.hpp:12-46): one set of bond constraints per engine
void Engine::add_constraint(const char* topology, size_t n, const uint32_t* idx, const double* v0, size_t max_it,
                            double tol)
{
    if(initialized_) {throw std::invalid_argument("constraints must be added before initialize");}
    if(con_.set) {throw std::invalid_argument("only one constraint set is supported");}
    if(n > 0 && (!idx || !v0)) {throw std::invalid_argument("null constraint arrays");}
    for(size_t k = 0; k < n; ++k)
    {
        if(idx[2 * k] >= n_ || idx[2 * k + 1] >= n_) {throw std::out_of_range("constraint index out of range");}
        if(idx[2 * k] == idx[2 * k + 1]) {throw std::invalid_argument("a constraint needs two particles");}
        if(!(v0[k] > 0)) {throw std::invalid_argument("constraint length must be positive");}
    }
    if(!(tol >= 0)) {throw std::invalid_argument("constraint tolerance must be non-negative");}
    const std::string topo = topology ? topology : "";
    con_.set = true;
    con_.topo_kind = (topo.empty() || topo == "none") ? -1 : kind_id(topo);
    con_.idx.assign(idx, idx + 2 * n);
    con_.v0.assign(v0, v0 + n);
    con_.max_it = max_it;
    con_.tol = tol;
}

// Topology::add_connection (Topology.hpp:161-191) for every local interaction with a
// connection kind, in the order the force field writes them (ForceField.hpp:53-56), then
// construct_molecules (Topology.hpp:280-317): only "bond" edges join molecules.
void Engine::write_topology()
{
    adj_.assign(n_, {});
    auto add = [&](uint32_t i, uint32_t j, int kind) {
        for(int s = 0; s < 2; ++s)
        {
            const uint32_t a = s ? j : i, b = s ? i : j;
            auto& v = adj_[a];
            if(std::find(v.begin(), v.end(), std::make_pair(b, kind)) == v.end()) {v.emplace_back(b, kind);}
        }
    };
    for(const auto& L : locals_)
    {
        if(L.topo_kind < 0) {continue;}
        if(L.kind == kBaseStacking)
        {
            // ThreeSPN2BaseStackingInteraction::write_topology (:125-179): every bead of the 5'
            // nucleotide to every bead of the 3' one, phosphates when present
            for(size_t t = 0; t < L.n_terms(); ++t)
            {
                const uint32_t* ix = L.idx.data() + t * 6;
                const uint32_t S5 = ix[0], B5 = ix[1], B3 = ix[2], P5 = ix[3], S3 = ix[4], P3 = ix[5];
                add(S5, S3, L.topo_kind);
                add(S5, B3, L.topo_kind);
                add(B5, S3, L.topo_kind);
                add(B5, B3, L.topo_kind);
                if(P5 != kNil) {add(P5, S3, L.topo_kind); add(P5, B3, L.topo_kind);}
                if(P3 != kNil) {add(S5, P3, L.topo_kind); add(B5, P3, L.topo_kind);}
                if(P5 != kNil && P3 != kNil) {add(P5, P3, L.topo_kind);}
            }
            continue;
        }
        for(size_t t = 0; t < L.n_terms(); ++t)
        {
            const uint32_t* ix = L.idx.data() + t * L.arity;
            for(int a = 0; a < L.arity; ++a)
                for(int b = a + 1; b < L.arity; ++b) {add(ix[a], ix[b], L.topo_kind);}
        }
    }
    // ConstraintForceField::write_topology (ConstraintForceField.hpp:49-60), after the locals
    // (ForceField.hpp:53-56)
    if(con_.topo_kind >= 0)
    {
        for(size_t k = 0; k < con_.v0.size(); ++k) {add(con_.idx[2 * k], con_.idx[2 * k + 1], con_.topo_kind);}
    }
    const int bond = kind_id("bond");
    const uint32_t UNINIT = 0xffffffffu;
    mol_.assign(n_, UNINIT);
    uint32_t next = 0;
    for(size_t i = 0; i < n_; ++i)
    {
        for(const auto& e : adj_[i])
        {
            if(e.second != bond) {continue;}
            if(mol_[e.first] != UNINIT) {mol_[i] = mol_[e.first]; break;}
        }
        if(mol_[i] == UNINIT) {mol_[i] = next++;}
    }
    n_molecules_ = next;
}

// ExclusionList::make (ExclusionList.hpp:86-234): per particle, itself plus every particle
// reachable within `dist` edges of each ignored connection kind
// (Topology::list_adjacent_within, Topology.hpp:133-153,243-264), sorted and unique.
void Engine::make_exclusions(GlobalIA& g)
{
    std::vector<uint32_t> start(n_ + 1), all;
    std::vector<uint32_t> tmp;
    std::vector<std::pair<uint32_t, uint32_t>> stack;
    for(size_t i = 0; i < n_; ++i)
    {
        start[i] = (uint32_t)all.size();
        tmp.clear();
        tmp.push_back((uint32_t)i);
        for(const auto& w : g.within)
        {
            // depth-first walk over `kind` edges up to `dist` steps
            stack.clear();
            stack.emplace_back((uint32_t)i, w.second);
            while(!stack.empty())
            {
                const auto top = stack.back();
                stack.pop_back();
                if(top.second == 0) {continue;}
                for(const auto& e : adj_[top.first])
                {
                    if(e.second != w.first) {continue;}
                    tmp.push_back(e.first);
                    stack.emplace_back(e.first, top.second - 1);
                }
            }
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        all.insert(all.end(), tmp.begin(), tmp.end());
    }
    start[n_] = (uint32_t)all.size();
    // a cell list with no exclusion (every particle's list holds only itself, which never matches a
    // partner), molecule or base rule decides a pair by its lower index alone
    g.rule.last_part = g.part.empty() ? kNoSlot : g.part.back();
    g.rule.simple = all.size() == n_ && !g.rule.pairwise && g.rule.ignore_mol == 0 && g.rule.base_rule == 0;
    g.d_ex_start = dalloc<uint32_t>(n_ + 1);
    g.d_ex = dalloc<uint32_t>(all.size());
    upload(g.d_ex_start, start.data(), n_ + 1);
    upload(g.d_ex, all.data(), all.size());
    g.d_mol = dalloc<uint32_t>(n_);
    upload(g.d_mol, mol_.data(), n_);
}

void Engine::alloc_list(GlobalIA& g, uint32_t cap)
{
    const uint32_t np = (uint32_t)g.part.size();
    g.cap = cap;
    g.d_nl = dalloc<uint32_t>((size_t)np * cap);
    if(g.tiled) {g.d_nl16 = dalloc<uint16_t>((size_t)np * cap);}
}

// The cell-tile pair kernel (tile_kernels.hpp) serves a pair interaction on a periodic cell list
// with at least kTileMinNp participants (MJG_TILE_MIN_NP overrides it; MJG_NO_TILE=1 turns the
// kernel off): it needs a launch of one workgroup per cell, which fills the GPU only for large
// grids.  Its staging capacity starts at twice the mean 27-cell neighbourhood and grows with the
// largest one a build reports (ListCtl::maxnbr), up to what the LDS of one workgroup holds after
// the kernel's static LDS (896 bytes in the current build; 4 KiB are set aside for it).
constexpr uint32_t kTileMinNp = 16384;
constexpr uint32_t kTileMaxCap = (uint32_t)((160 * 1024 - 4096) / kTileBytesPerParticle);
static uint32_t tile_min_np()
{
    if(const char* v = std::getenv("MJG_TILE_MIN_NP")) {return (uint32_t)std::max(0L, std::atol(v));}
    return kTileMinNp;
}

// lanes per participant: about six partners per lane, a power of two in [1, 64]
static uint32_t lanes_for(uint32_t maxcount)
{
    uint32_t t = 1;
    while(t < 64 && t * 6 < maxcount) {t <<= 1;}
    return t;
}

void Engine::setup_global(GlobalIA& g)
{
    const uint32_t np = (uint32_t)g.part.size();
    // per-particle parameter table (zero for non-participants)
    std::vector<double> ppar(n_ * g.npar, 0.0);
    for(uint32_t k = 0; k < np; ++k)
        for(int c = 0; c < g.npar; ++c) {ppar[(size_t)g.part[k] * g.npar + c] = g.par[(size_t)k * g.npar + c];}

    PairParams& pp = g.pp;
    pp = PairParams{};
    pp.kind = g.pot;
    pp.npar = g.npar;
    if(g.pot == kPairLJ)
    {
        // LennardJonesPotential.hpp:34-38 and max_cutoff :82-98 over the parameter table
        pp.cutoff_ratio2 = g.cutoff_ratio * g.cutoff_ratio;
        pp.coef_at_cutoff = std::pow(1.0 / g.cutoff_ratio, 12) - std::pow(1.0 / g.cutoff_ratio, 6);
        double ms = 0;
        for(size_t i = 0; i < n_; ++i) {ms = std::max(ms, ppar[2 * i]);}
        g.max_cutoff = ms * g.cutoff_ratio;
    }
    else if(g.pot == kPairEV)
    {
        // ExcludedVolumePotential.hpp:44-48, max_cutoff :90-106
        pp.cutoff_ratio2 = g.cutoff_ratio * g.cutoff_ratio;
        pp.coef_at_cutoff = std::pow(1.0 / g.cutoff_ratio, 12);
        pp.eps = g.eps;
        double mr = 0;
        for(size_t i = 0; i < n_; ++i) {mr = std::max(mr, ppar[i]);}
        g.max_cutoff = mr * 2 * g.cutoff_ratio;
    }
    else if(g.pot == kPair3SPN2EV)
    {
        // ThreeSPN2ExcludedVolumePotential.hpp:120-134: epsilon 1 kJ/mol in kcal/mol; the list
        // cutoff is the largest sigma of the table (:176-178)
        pp.eps = 1.0 * (1.0 / 4.184);
        g.max_cutoff = 7.1;
    }
    else if(g.multibody())
    {
        g.max_cutoff = 18.0;   // ThreeSPN2BasePairGlobalPotentialParameter::cutoff
    }
    else
    {
        // DebyeHuckelPotential.hpp:126-174
        if(!has_T_ || !has_I_)
        {
            throw std::invalid_argument("DebyeHuckel requires `temperature` and `ionic_strength` attributes");
        }
        const double pi = 3.14159265358979323846264338;
        const double kB = kB_kcal(), NA = 6.022140857e23, eps0 = eps0_kcal_angstrom();
        const double T = T_, c = I_;
        const double epsk = (249.4 - 0.788 * T + 7.2e-4 * T * T) *
                            (1. - 2.551e-1 * c + 5.151e-2 * c * c - 6.889e-3 * c * c * c);
        pp.inv4pieps = 1.0 / (4 * pi * eps0 * epsk);
        const double I = c / L_to_volume_angstrom();
        pp.debye = std::sqrt((eps0 * epsk * kB * T) / (2 * NA * I));
        pp.inv_debye = 1. / pp.debye;
        pp.abs_cutoff = g.cutoff_ratio * pp.debye;
        pp.coef_at_cutoff = std::exp(-g.cutoff_ratio) / (pp.debye * g.cutoff_ratio);
        g.max_cutoff = pp.debye * g.cutoff_ratio;
        if(!std::isfinite(g.max_cutoff) || !(g.max_cutoff > 0)) {throw std::invalid_argument("invalid Debye length");}
    }
    g.r_list = g.max_cutoff * (1. + g.margin);

    // participant slot per particle (kNoSlot for non-participants)
    std::vector<uint32_t> pslot(n_, kNoSlot);
    for(uint32_t k = 0; k < np; ++k) {pslot[g.part[k]] = k;}
    g.rule.last_slot = np - 1;
    g.rule.pairwise = (g.partition != MJG_PART_CELLLIST);
    g.rule.use_distance = (g.partition != MJG_PART_NAIVE);
    g.rule.ignore_mol = g.ignore_mol;
    g.rule.r_list2 = g.r_list * g.r_list;

    g.d_part = dalloc<uint32_t>(np);
    upload(g.d_part, g.part.data(), np);
    g.d_pslot = dalloc<uint32_t>(n_);
    upload(g.d_pslot, pslot.data(), n_);
    g.d_ppar = dalloc<double>(n_ * g.npar);
    upload(g.d_ppar, ppar.data(), n_ * g.npar);
    g.rule.base_rule = (g.pot == kPair3SPN2EV) ? 1 : g.multibody() ? 2 : 0;
    g.rule.base_stride = g.npar;
    g.rule.base_par = g.d_ppar;
    make_exclusions(g);

    g.d_grid = dalloc<Grid>(1);
    if(g.partition == MJG_PART_CELLLIST)
    {
        if(box_.periodic)
        {
            // PeriodicGridCellList::set_cutoff / initialize (:67-96, :152-160)
            const double me = 1e-6;
            const double r = 1 / (g.max_cutoff * (1 + g.margin) + me);
            Grid gr{};
            for(int d = 0; d < 3; ++d)
            {
                const double f = std::floor(box_.width[d] * r);
                gr.dim[d] = (int)std::max<double>(3.0, f);
                gr.inv_w[d] = 1.0 / (box_.width[d] / gr.dim[d]);
                gr.lo[d] = box_.lower[d];
            }
            gr.periodic = 1;
            const uint64_t nc = (uint64_t)gr.dim[0] * gr.dim[1] * gr.dim[2];
            if(nc > kMaxCells) {throw std::invalid_argument("periodic box needs more than 2^18 cells");}
            gr.ncell = (uint32_t)nc;
            g.grid = gr;
            upload(g.d_grid, &gr, 1);
        }
        g.sb.keys = dalloc<uint32_t>(np); g.sb.vals = dalloc<uint32_t>(np);
        g.sb.ukeys = dalloc<uint32_t>(np);
        g.sb.count = dalloc<uint32_t>(kMaxCells);
        check(hipMemsetAsync(g.sb.count, 0, kMaxCells * sizeof(uint32_t), stream_), "memset");
        g.d_cell_start = dalloc<uint32_t>(kMaxCells);
        g.d_cell_end = dalloc<uint32_t>(kMaxCells);
    }
    // one-workgroup rebuild: the participants and the cells must fit its LDS tables
    g.small = (np <= kSmallMaxPart) && (!box_.periodic || g.partition != MJG_PART_CELLLIST ||
                                        g.grid.ncell <= kSmallMaxCells);
    if(g.partition != MJG_PART_CELLLIST && g.sb.keys == nullptr)
    {
        g.sb.keys = dalloc<uint32_t>(np); g.sb.vals = dalloc<uint32_t>(np);
    }
    const char* no_tile = std::getenv("MJG_NO_TILE");
    g.tiled = box_.periodic && g.partition == MJG_PART_CELLLIST && !g.multibody() && np >= tile_min_np() &&
              !(no_tile && no_tile[0] == '1');
    if(g.tiled)
    {
        const double mean_nbr = 27.0 * (double)np / (double)g.grid.ncell;
        g.tile_cap = (uint32_t)std::min<double>(kTileMaxCap, std::max(256.0, 1.25 * mean_nbr + 64.0));
        // MJG_TILE_CAP fixes the staging capacity (no growth): a capacity below the largest
        // neighbourhood serves every build from global memory
        if(const char* v = std::getenv("MJG_TILE_CAP"))
        {
            g.tile_cap = (uint32_t)std::min<long>(kTileMaxCap, std::max(1L, std::atol(v)));
            g.tile_fixed = true;
        }
    }
    g.d_nn = dalloc<uint32_t>(np);
    g.d_order = dalloc<uint32_t>(np);
    g.d_spart = dalloc<uint32_t>(np);
    if(g.partition == MJG_PART_CELLLIST) {g.d_spos = dalloc<double4>(np);}
    g.d_epair = dalloc<double>(np);
    g.d_fg = dalloc<double4>(n_);
    check(hipMemsetAsync(g.d_fg, 0, n_ * sizeof(double4), stream_), "memset");
    if(g.multibody()) {g.d_acc = dalloc<unsigned long long>(3 * n_ + 1);}   // sums and the flag word
    sync();
    g.tpp = 8;
    alloc_list(g, 64);
}

ListSeg Engine::list_seg(GlobalIA& g, size_t gi)
{
    ListSeg s{};
    s.np = (uint32_t)g.part.size();
    s.tpp = g.tpp;
    s.cap = g.cap;
    s.allpairs = (g.partition != MJG_PART_CELLLIST);
    s.part = g.d_part;
    s.skeys = g.sb.keys;
    s.svals = g.sb.vals;
    s.cell_start = g.d_cell_start;
    s.cell_end = g.d_cell_end;
    s.grid = g.d_grid;
    s.pslot = g.d_pslot;
    s.mol = g.d_mol;
    s.ex_start = g.d_ex_start;
    s.ex = g.d_ex;
    s.nl = g.d_nl;
    s.nl16 = g.tiled ? g.d_nl16 : nullptr;
    s.nn = g.d_nn;
    s.order = g.d_order;
    s.spart = g.d_spart;
    s.spos = g.d_spos;
    s.ctl = d_lctl_ + gi;
    s.rule = g.rule;
    return s;
}

// k_verlet's segments: every cell-list interaction the cell-tile kernel does not serve (those get
// their own build, k_verlet_tiled)
ListSegTable Engine::list_table()
{
    ListSegTable t{};
    uint32_t block = 0;
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        GlobalIA& g = globals_[gi];
        if(g.tiled) {continue;}
        ListSeg& s = t.seg[t.n++];
        s = list_seg(g, gi);
        s.block_begin = block;
        s.nblocks = std::min<uint32_t>(blocks_for((size_t)s.np * s.tpp, 256), kVerletMaxBlocks);
        block += s.nblocks;
    }
    return t;
}

PairSegTable Engine::pair_table(int only)
{
    PairSegTable t{};
    uint32_t block = 0;
    const int nl = (int)locals_.size();
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        if(only >= 0 && only != nl + (int)gi) {continue;}
        GlobalIA& g = globals_[gi];
        if(g.multibody() || g.tiled) {continue;}
        PairSeg& s = t.seg[t.n++];
        s.kind = g.pot;
        s.tpp = g.tpp;
        s.np = (uint32_t)g.part.size();
        s.block_begin = block;
        s.cap = g.cap;
        s.part = g.d_part;
        s.order = g.d_order;
        s.nl = g.d_nl;
        s.nn = g.d_nn;
        s.ppar = g.d_ppar;
        s.fout = g.d_fg;
        s.eout = g.d_epair;
        s.pp = g.pp;
        block += blocks_for((size_t)s.np * s.tpp, 256);
    }
    return t;
}

// Launch the skin check and the rebuild pipeline of every global interaction; the kernels
// check ListCtl::rebuild themselves, so a captured step always launches the same sequence.
// force != 0 rebuilds every list now (initialize, explicit reduce_margin, capacity growth).
void Engine::launch_rebuilds(int force)
{
    PrepSegTable small{}, large{};
    std::vector<int> large_ids;
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        GlobalIA& g = globals_[gi];
        PrepSeg ps{};
        ps.np = (uint32_t)g.part.size();
        ps.periodic_grid = box_.periodic;
        ps.never_rebuild = (g.partition == MJG_PART_NAIVE);
        ps.allpairs = (g.partition != MJG_PART_CELLLIST);
        ps.abs_margin = g.max_cutoff * g.margin;
        ps.r_list = g.r_list;
        ps.part = g.d_part;
        ps.skeys = g.sb.keys;
        ps.svals = g.sb.vals;
        ps.cell_start = g.d_cell_start;
        ps.cell_end = g.d_cell_end;
        ps.grid = g.d_grid;
        ps.ctl = d_lctl_ + gi;
        if(g.small || g.partition != MJG_PART_CELLLIST) {small.seg[small.n++] = ps;}
        else {large.seg[large.n++] = ps; large_ids.push_back((int)gi);}
    }
    if(small.n > 0)
    {
        hipLaunchKernelGGL(k_prep_small, dim3(small.n), dim3(1024), 0, stream_, d_pos_, small, d_sctl_, force);
    }
    if(large.n > 0)
    {
        hipLaunchKernelGGL(k_margin_large, dim3(1), dim3(64), 0, stream_, large, d_sctl_, force);
        for(int gi : large_ids)
        {
            GlobalIA& g = globals_[gi];
            const uint32_t np = (uint32_t)g.part.size();
            ListCtl* ctl = d_lctl_ + gi;
            if(!box_.periodic)
            {
                hipLaunchKernelGGL(k_bbox, dim3(std::min<uint32_t>(blocks_for(np, 256), 1024)), dim3(256), 0, stream_,
                                   d_pos_, g.d_part, np, ctl);
                hipLaunchKernelGGL(k_grid_open, dim3(1), dim3(1), 0, stream_, ctl, g.d_grid, g.r_list);
            }
            // a periodic grid's cell count is known here; an open grid's is set on the device
            const uint32_t cells = box_.periodic ? g.grid.ncell : kMaxCells;
            hipLaunchKernelGGL(k_cellkey, dim3(blocks_for(np, 256)), dim3(256), 0, stream_, d_pos_, g.d_part, np,
                               g.d_grid, g.sb.ukeys, g.sb.count, ctl);
            hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, stream_, g.d_grid, g.sb.count, g.d_cell_start,
                               g.d_cell_end, ctl);
            hipLaunchKernelGGL(k_cell_place, dim3(blocks_for(np, 256)), dim3(256), 0, stream_, g.sb.ukeys, np,
                               g.d_cell_end, g.sb.keys, g.sb.vals, ctl);
            hipLaunchKernelGGL(k_cell_order, dim3(std::min<uint32_t>(blocks_for(cells, 256), 1024)), dim3(256), 0,
                               stream_, g.d_grid, g.d_cell_start, g.d_cell_end, g.sb.count, g.sb.vals, d_pos_,
                               g.d_part, g.d_spart, g.d_spos, ctl);
        }
    }
    // the slot copies of the interactions that k_prep_small prepares (k_cell_order made the others')
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        GlobalIA& g = globals_[gi];
        if(g.partition != MJG_PART_CELLLIST || !g.small) {continue;}
        const uint32_t np = (uint32_t)g.part.size();
        hipLaunchKernelGGL(k_sorted_slots, dim3(std::min<uint32_t>(blocks_for(np, 256), 1024)), dim3(256), 0, stream_,
                           d_pos_, g.d_part, g.sb.vals, np, g.d_spart, g.d_spos, d_lctl_ + gi);
    }
    const ListSegTable lt = list_table();
    uint32_t blocks = 0;
    for(int k = 0; k < lt.n; ++k) {blocks += lt.seg[k].nblocks;}
    if(blocks > 0) {hipLaunchKernelGGL(k_verlet, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, lt);}
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        GlobalIA& g = globals_[gi];
        if(!g.tiled) {continue;}
        const ListSeg s = list_seg(g, gi);
        const size_t lds = (size_t)g.tile_cap * kBuildBytesPerParticle;
        hipLaunchKernelGGL(k_verlet_tiled, dim3(std::min<uint32_t>(g.grid.ncell, kVerletMaxBlocks)), dim3(kTileThreads),
                           lds, stream_, d_pos_, box_, s, g.tile_cap);
    }
}

// Build the lists of every global interaction now (initialize / explicit reduce_margin),
// growing the list capacity until every partner list fits, then size the lanes per
// participant from the longest list.
void Engine::build_lists_now(bool)
{
    std::vector<ListCtl> ctl(globals_.size());
    for(int pass = 0; pass < 8; ++pass)
    {
        launch_rebuilds(1);
        check(hipGetLastError(), "rebuild launch");
        download(ctl.data(), d_lctl_, globals_.size());
        bool again = false;
        for(size_t g = 0; g < globals_.size(); ++g)
        {
            GlobalIA& G = globals_[g];
            if(ctl[g].maxcount > G.cap)
            {
                alloc_list(G, (uint32_t)(ctl[g].maxcount + ctl[g].maxcount / 4 + 16));
                again = true;
            }
            const uint32_t t = lanes_for(ctl[g].maxcount);
            if(t != G.tpp) {G.tpp = t; again = again || false;}
            uint32_t mb = 1;
            while(mb < 64 && mb < ctl[g].maxcount) {mb <<= 1;}
            G.mb_tpp = mb;
        }
        if(graph_exec_) {(void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr;}
        if(!again) {break;}
    }
    // the decision flags were consumed by this build
    for(auto& c : ctl) {c.rebuild = 0;}
    download(ctl.data(), d_lctl_, globals_.size());
    for(auto& c : ctl) {c.rebuild = 0;}
    upload(d_lctl_, ctl.data(), globals_.size());
    grow_tiles();
}

void Engine::initialize()
{
    if(!state_set_) {throw std::invalid_argument("system state must be set before initialize");}
    if(initialized_) {throw std::invalid_argument("force field already initialized");}
    // the cell-tile kernel stages up to kTileMaxCap particles in the LDS of one workgroup
    const int tile_lds = (int)((size_t)kTileMaxCap * kTileBytesPerParticle);
    check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_tiled<true, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds), "tile attribute");
    check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_tiled<true, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds), "tile attribute");
    check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_tiled<false, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds), "tile attribute");
    check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pairs_tiled<false, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, tile_lds), "tile attribute");
    check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_verlet_tiled), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)((size_t)kTileMaxCap * kBuildBytesPerParticle)), "tile attribute");
    write_topology();

    // local tables: one segment per interaction, body slots in interaction/term/body order
    std::vector<uint32_t> lidx;
    std::vector<double> lpar;
    std::vector<std::vector<uint32_t>> per_particle(n_);
    segtab_.n = 0;
    uint32_t block = 0, slot = 0, term = 0;
    for(size_t l = 0; l < locals_.size(); ++l)
    {
        LocalIA& L = locals_[l];
        LocalSeg s{};
        s.kind = L.kind; s.arity = L.arity; s.npar = L.npar;
        s.n_terms = (uint32_t)L.n_terms();
        s.block_begin = block;
        s.idx_off = (uint32_t)lidx.size();
        s.par_off = (uint32_t)lpar.size();
        s.slot_off = slot;
        s.term_off = term;
        L.slot_off = slot; L.term_off = term;
        s.nbody = L.nbody;
        lidx.insert(lidx.end(), L.idx.begin(), L.idx.end());
        lpar.insert(lpar.end(), L.par.begin(), L.par.end());
        for(uint32_t t = 0; t < s.n_terms; ++t)
            for(int b = 0; b < L.nbody; ++b) {per_particle[L.idx[(size_t)t * L.arity + b]].push_back(slot + t * L.nbody + b);}
        segtab_.seg[segtab_.n++] = s;
        block += blocks_for(s.n_terms, 256) * (s.n_terms > 0 ? 1 : 0);
        slot += s.n_terms * L.nbody;
        term += s.n_terms;
    }
    local_blocks_ = block;
    n_slots_ = slot;
    n_terms_ = term;
    if(!locals_.empty())
    {
        std::vector<uint32_t> gstart(n_ + 1), gslot;
        for(size_t i = 0; i < n_; ++i)
        {
            gstart[i] = (uint32_t)gslot.size();
            gslot.insert(gslot.end(), per_particle[i].begin(), per_particle[i].end());
        }
        gstart[n_] = (uint32_t)gslot.size();
        d_lidx_ = dalloc<uint32_t>(lidx.size()); upload(d_lidx_, lidx.data(), lidx.size());
        d_lpar_ = dalloc<double>(lpar.size()); upload(d_lpar_, lpar.data(), lpar.size());
        d_tf_ = dalloc<double>(3 * (size_t)n_slots_);
        d_te_ = dalloc<double>(n_terms_);
        d_gstart_ = dalloc<uint32_t>(n_ + 1); upload(d_gstart_, gstart.data(), n_ + 1);
        d_gslot_ = dalloc<uint32_t>(gslot.size()); upload(d_gslot_, gslot.data(), gslot.size());
    }

    d_lctl_ = dalloc<ListCtl>(std::max<size_t>(1, globals_.size()));
    d_bak_lctl_ = dalloc<ListCtl>(std::max<size_t>(1, globals_.size()));
    d_abs_margin_ = dalloc<double>(std::max<size_t>(1, globals_.size()));
    d_never_ = dalloc<int>(std::max<size_t>(1, globals_.size()));
    d_energy_ = dalloc<double>(num_interactions() + 1);
    std::vector<double> absm;
    std::vector<int> never;
    for(auto& g : globals_)
    {
        setup_global(g);
        absm.push_back(g.max_cutoff * g.margin);
        never.push_back(g.partition == MJG_PART_NAIVE);
    }
    if(!globals_.empty())
    {
        upload(d_abs_margin_, absm.data(), absm.size());
        upload(d_never_, never.data(), never.size());
        std::vector<ListCtl> ctl(globals_.size());
        upload(d_lctl_, ctl.data(), ctl.size());
        build_lists_now(true);
    }
    initialized_ = true;
}

// ForceField::calc_force into f: locals (term kernel), every pair interaction (fused kernel),
// then k_finish sums them per particle (and with `kick` applies the closing half-kick of the
// integrator step).  With `energy`, per-interaction energies land in d_energy_.  only >= 0
// restricts the call to one interaction id.
void Engine::forces(double4* f, bool energy, int only, bool kick, bool advance)
{
    const int nl = (int)locals_.size();
    const bool run_locals = local_blocks_ > 0 && (only < 0 || only < nl);
    if(run_locals)
    {
        if(only >= 0)
        {
            // other segments' slots must read as zero
            hipLaunchKernelGGL(k_zero_doubles, dim3(blocks_for(3 * (size_t)n_slots_, 256)), dim3(256), 0, stream_,
                               d_tf_, 3 * n_slots_);
        }
        if(energy)
        {
            hipLaunchKernelGGL(k_local_terms<true>, dim3(local_blocks_), dim3(256), 0, stream_, d_pos_, box_, segtab_,
                               d_lidx_, d_lpar_, d_tf_, d_te_, only);
        }
        else
        {
            hipLaunchKernelGGL(k_local_terms<false>, dim3(local_blocks_), dim3(256), 0, stream_, d_pos_, box_, segtab_,
                               d_lidx_, d_lpar_, d_tf_, d_te_, only);
        }
    }
    const PairSegTable pt = pair_table(only);
    if(pt.n > 0)
    {
        uint32_t blocks = 0;
        for(int k = 0; k < pt.n; ++k) {blocks += blocks_for((size_t)pt.seg[k].np * pt.seg[k].tpp, 256);}
        if(precision_ == MJG_FP32_FIXED)
        {
            if(energy) {hipLaunchKernelGGL(k_pairs_mixed<true>, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, pt);}
            else {hipLaunchKernelGGL(k_pairs_mixed<false>, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, pt);}
        }
        else
        {
            if(energy) {hipLaunchKernelGGL(k_pairs<true>, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, pt);}
            else {hipLaunchKernelGGL(k_pairs<false>, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, pt);}
        }
    }
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        if(only >= 0 && only != nl + (int)gi) {continue;}
        if(globals_[gi].multibody()) {launch_multibody(globals_[gi], energy);}
        else if(globals_[gi].tiled) {launch_tiled(globals_[gi], (int)gi, energy);}
    }
    if(energy)
    {
        for(int l = 0; l < nl; ++l)
        {
            if(only >= 0 && only != l) {continue;}
            const auto& L = locals_[l];
            hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, stream_, d_te_ + L.term_off, (uint32_t)L.n_terms(),
                               d_energy_ + l);
        }
        for(size_t gi = 0; gi < globals_.size(); ++gi)
        {
            const int id = nl + (int)gi;
            if(only >= 0 && only != id) {continue;}
            hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, stream_, globals_[gi].d_epair,
                               (uint32_t)globals_[gi].part.size(), d_energy_ + id);
        }
    }
    FinishArgs fa{};
    fa.gstart = run_locals ? d_gstart_ : nullptr;
    fa.gslot = d_gslot_;
    fa.tf = d_tf_;
    fa.ng = 0;
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        if(only >= 0 && only != nl + (int)gi) {continue;}
        fa.fg[fa.ng++] = globals_[gi].d_fg;
    }
    fa.halfdt = dt_ / 2;
    fa.kick = kick ? 1 : 0;
    fa.advance = (kick || advance) ? 1 : 0;
    hipLaunchKernelGGL(k_finish, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, (uint32_t)n_, fa, f, d_vel_,
                       d_rmass_, &d_sctl_->step);
}

// 3SPN2 base pairing / cross stacking: clear the fixed-point sums, one thread per participant
// row, then convert the sums into the interaction's force array.
void Engine::launch_multibody(const GlobalIA& g, bool energy)
{
    const uint32_t np = (uint32_t)g.part.size();
    hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(3 * n_ + 1, 256)), dim3(256), 0, stream_, g.d_acc,
                       (uint32_t)(3 * n_ + 1));
    TspSeg s{};
    s.kind = g.pot; s.np = np; s.cap = g.cap; s.tpp = g.mb_tpp;
    s.part = g.d_part; s.order = g.d_order; s.nl = g.d_nl; s.nn = g.d_nn;
    s.ppar = g.d_ppar; s.acc = g.d_acc; s.n = (uint32_t)n_; s.eout = g.d_epair;
    const uint32_t blocks = blocks_for((size_t)np * s.tpp, 256);
    if(energy) {hipLaunchKernelGGL(k_tsp_pairs<true>, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, s);}
    else {hipLaunchKernelGGL(k_tsp_pairs<false>, dim3(blocks), dim3(256), 0, stream_, d_pos_, box_, s);}
    hipLaunchKernelGGL(k_fix_to_force, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, g.d_acc, g.d_fg, (uint32_t)n_);
}

// Cell-tile pair kernel of one interaction: one workgroup per cell, its LDS the staging of the
// 27-cell neighbourhood, the fused kernel's lanes per row (the same sums bit for bit).
void Engine::launch_tiled(GlobalIA& g, int gi, bool energy)
{
    TileSeg s{};
    PairSeg& ps = s.ps;
    ps.kind = g.pot; ps.tpp = g.tpp; ps.np = (uint32_t)g.part.size(); ps.block_begin = 0; ps.cap = g.cap;
    ps.part = g.d_part; ps.order = g.d_order; ps.nl = g.d_nl; ps.nn = g.d_nn; ps.ppar = g.d_ppar;
    ps.fout = g.d_fg; ps.eout = g.d_epair; ps.pp = g.pp;
    s.nl16 = g.d_nl16; s.spart = g.d_spart; s.cell_start = g.d_cell_start; s.cell_end = g.d_cell_end;
    s.grid = g.d_grid; s.ctl = d_lctl_ + gi; s.cap_nbr = g.tile_cap;
    const bool mixed = (precision_ == MJG_FP32_FIXED);
    // FP64 keeps the gather kernel's lanes per row (its sums in the same order); the mixed sums are
    // exact, so its lanes per row are chosen for the tile alone
    s.tpp = mixed ? tile_lanes(g) : g.tpp;
    const size_t lds = (size_t)g.tile_cap * (mixed ? kTileBytesPerParticleF32 : kTileBytesPerParticle);
    const dim3 grid(g.grid.ncell), block(kTileThreads);
    if(mixed)
    {
        if(energy) {hipLaunchKernelGGL((k_pairs_tiled<true, true>), grid, block, lds, stream_, d_pos_, box_, s);}
        else {hipLaunchKernelGGL((k_pairs_tiled<true, false>), grid, block, lds, stream_, d_pos_, box_, s);}
    }
    else
    {
        if(energy) {hipLaunchKernelGGL((k_pairs_tiled<false, true>), grid, block, lds, stream_, d_pos_, box_, s);}
        else {hipLaunchKernelGGL((k_pairs_tiled<false, false>), grid, block, lds, stream_, d_pos_, box_, s);}
    }
}

// Lanes per list row of the mixed cell-tile kernel: MJG_TILE_TPP (a power of two up to 64) or half
// the gather kernel's (about twelve entries a lane for the longest row; measured on a 262,144-particle
// LJ fluid, 8 and 16 lanes per row ran 10 % faster than 4 or 32).
uint32_t Engine::tile_lanes(const GlobalIA& g) const
{
    if(const char* v = std::getenv("MJG_TILE_TPP"))
    {
        const long t = std::atol(v);
        if(t >= 1 && t <= 64 && (t & (t - 1)) == 0) {return (uint32_t)t;}
    }
    return std::max<uint32_t>(1, g.tpp / 2);
}

// After lists were built: a tiled interaction whose largest neighbourhood outgrew its staging
// capacity gets a larger one (its launches served that build from global memory meanwhile).
// True if a capacity changed, which invalidates a captured step graph.
bool Engine::grow_tiles()
{
    bool any = false;
    std::vector<ListCtl> ctl(globals_.size());
    bool have = false;
    for(size_t gi = 0; gi < globals_.size(); ++gi)
    {
        GlobalIA& g = globals_[gi];
        if(!g.tiled) {continue;}
        if(!have) {download(ctl.data(), d_lctl_, globals_.size()); have = true;}
        const uint32_t need = ctl[gi].maxnbr;
        if(need > g.tile_cap && g.tile_cap < kTileMaxCap && !g.tile_fixed)
        {
            g.tile_cap = std::min<uint32_t>(kTileMaxCap, need + need / 4 + 32);
            any = true;
        }
    }
    if(any && graph_exec_) {(void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr;}
    return any;
}

void Engine::calc_force()
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    forces(d_force_, false, -1, false);
    check(hipGetLastError(), "calc_force launch");
    sync();
}

double Engine::calc_force_and_energy()
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    forces(d_force_, true, -1, false);
    check(hipGetLastError(), "calc_force_and_energy launch");
    std::vector<double> e(num_interactions());
    download(e.data(), d_energy_, e.size());
    double tot = 0;
    for(double x : e) {tot += x;}
    return tot;
}

void Engine::calc_energies(double* per, double* total)
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    forces(d_force_tmp_, true, -1, false);
    check(hipGetLastError(), "calc_energies launch");
    std::vector<double> e(num_interactions());
    download(e.data(), d_energy_, e.size());
    double tot = 0;
    for(size_t k = 0; k < e.size(); ++k) {if(per) {per[k] = e[k];} tot += e[k];}
    if(total) {*total = tot;}
}

void Engine::calc_force_of(int id, double* out)
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    if(id < 0 || (size_t)id >= num_interactions()) {throw std::out_of_range("interaction id out of range");}
    forces(d_force_tmp_, false, id, false);
    check(hipGetLastError(), "calc_force_of launch");
    std::vector<double4> h(n_);
    download(h.data(), d_force_tmp_, n_);
    for(size_t i = 0; i < n_; ++i) {out[3 * i] = h[i].x; out[3 * i + 1] = h[i].y; out[3 * i + 2] = h[i].z;}
}

void Engine::reduce_margin(double dmargin)
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    if(globals_.empty()) {return;}
    std::vector<ListCtl> ctl(globals_.size());
    download(ctl.data(), d_lctl_, globals_.size());
    bool any = false;
    for(size_t g = 0; g < globals_.size(); ++g)
    {
        if(globals_[g].partition == MJG_PART_NAIVE) {continue;}
        ctl[g].margin -= dmargin;
        any = any || ctl[g].margin < 0;
    }
    upload(d_lctl_, ctl.data(), globals_.size());
    // a rebuild resets the margin of every list (rebuilding a list early is always valid)
    if(any) {build_lists_now(true);}
}

size_t Engine::neighbor_pairs(int id, uint32_t* pairs, size_t cap)
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    const int gi = id - (int)locals_.size();
    if(gi < 0 || (size_t)gi >= globals_.size()) {throw std::out_of_range("not a global interaction id");}
    GlobalIA& g = globals_[gi];
    const uint32_t np = (uint32_t)g.part.size();
    std::vector<uint32_t> nn(np), order(np), nl((size_t)np * g.cap);
    download(nn.data(), g.d_nn, np);
    download(order.data(), g.d_order, np);
    download(nl.data(), g.d_nl, nl.size());
    std::vector<std::pair<uint32_t, uint32_t>> out;
    for(uint32_t t = 0; t < np; ++t)
    {
        const uint32_t i = g.part[order[t]];
        for(uint32_t k = 0; k < nn[t]; ++k)
        {
            const uint32_t j = nl[(size_t)t * g.cap + k];
            if(i < j) {out.emplace_back(i, j);}
        }
    }
    std::sort(out.begin(), out.end());
    if(pairs)
    {
        for(size_t k = 0; k < out.size() && k < cap; ++k) {pairs[2 * k] = out[k].first; pairs[2 * k + 1] = out[k].second;}
    }
    return out.size();
}

void Engine::set_integrator(int type, double dt, const double* coefs, uint64_t seed)
{
    if(type != MJG_INTEG_BAOAB && type != MJG_INTEG_GJF && type != MJG_INTEG_GBAOAB && type != MJG_INTEG_UNDERDAMPED)
    {
        throw std::invalid_argument("unknown integrator");
    }
    if(!(dt > 0) || !coefs) {throw std::invalid_argument("invalid integrator parameters");}
    integ_ = type; dt_ = dt; seed_ = seed;
    coefs_.assign(coefs, coefs + n_);
    integ_ready_ = false;
    if(graph_exec_) {(void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr;}
}

// SystemMotionRemover(translation, rotation, rescale) (SystemMotionRemover.hpp:21-31), applied
// after every step as the integrators do (BAOABLangevinIntegrator.hpp:236-237).
void Engine::set_remover(int translation, int rotation, int rescale)
{
    rm_translation_ = translation != 0;
    rm_rotation_ = rotation != 0;
    rm_rescale_ = rescale != 0;
    if(graph_exec_) {(void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr;}
}

// reset_parameters (BAOABLangevinIntegrator.hpp:73-99; GJFNVTLangevinIntegrator.hpp:106-134;
// UnderdampedLangevinIntegrator.hpp:320-350) and the first force call (:151-166 / :58-73 /
// :404-457).
void Engine::integrator_initialize()
{
    if(!initialized_) {throw std::logic_error("force field not initialized");}
    if(integ_ == 0) {throw std::logic_error("integrator not set");}
    if(!has_T_) {throw std::out_of_range("Langevin integrator requires the `temperature` attribute");}
    const double kBT = kB_kcal() * T_;
    const double noise_coef = std::sqrt(2 * kB_kcal() * T_ / dt_);
    std::vector<double> c1(n_), c2(n_), c3(n_);
    for(size_t i = 0; i < n_; ++i)
    {
        if(integ_ == MJG_INTEG_UNDERDAMPED)
        {
            c1[i] = coefs_[i];
            c2[i] = noise_coef * std::sqrt(coefs_[i] * (1.0 / mass_[i]));
        }
        else if(integ_ == MJG_INTEG_BAOAB || integ_ == MJG_INTEG_GBAOAB)
        {
            // gBAOABLangevinIntegrator::reset_parameter (:82-107) is BAOAB's
            const double gamma_dt = -1 * coefs_[i] * dt_;
            c1[i] = std::exp(gamma_dt);
            c2[i] = std::sqrt(kBT * (1 - std::exp(2 * gamma_dt)) * (1.0 / mass_[i]));
        }
        else
        {
            const double alpha = coefs_[i], m = mass_[i];
            c1[i] = 1.0 / (1.0 + alpha * dt_ * 0.5 / m);
            c2[i] = std::sqrt(2 * alpha * kBT * dt_);
            c3[i] = 1.0 - alpha * c1[i] * dt_ / m;
        }
    }
    if(!d_c1_) {d_c1_ = dalloc<double>(n_); d_c2_ = dalloc<double>(n_); d_c3_ = dalloc<double>(n_);}
    upload(d_c1_, c1.data(), n_);
    upload(d_c2_, c2.data(), n_);
    upload(d_c3_, c3.data(), n_);
    if(integ_ == MJG_INTEG_GBAOAB) {setup_constraints();}
    forces(d_force_, false, -1, false);
    if(integ_ == MJG_INTEG_GBAOAB)
    {
        hipLaunchKernelGGL(k_copy_double4, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, (uint32_t)n_, d_pos_,
                           d_old_pos_);
    }
    if(integ_ == MJG_INTEG_UNDERDAMPED)
    {
        if(!d_accel_) {d_accel_ = dalloc<double4>(n_); d_bak_accel_ = dalloc<double4>(n_);}
        launch_ul_accel(false);
    }
    check(hipGetLastError(), "integrator_initialize launch");
    sync();
    integ_ready_ = true;
}

// gBAOABLangevinIntegrator::initialize (:290-345): v0^2, 1/v0^2 and the reduced masses of the
// constraints, the correction time step (one correction per A step), and the connected
// components of the constraint graph (union-find), each listing its constraints in input order.
void Engine::setup_constraints()
{
    const size_t m = con_.v0.size();
    dt_corr_ = dt_ * 0.5 / 1;
    if(!d_old_pos_) {d_old_pos_ = dalloc<double4>(n_); d_old_rattle_ = dalloc<double4>(n_);}
    ctab_ = ConTable{};
    ctab_.max_it = con_.max_it;
    ctab_.tol = con_.tol;
    ctab_.r_dt_corr = 1. / dt_corr_;
    if(m == 0) {return;}
    std::vector<double> sq(m), rsq(m), redm(m);
    for(size_t k = 0; k < m; ++k)
    {
        sq[k] = std::pow(con_.v0[k], 2);
        rsq[k] = 1.0 / sq[k];
        redm[k] = 1.0 / ((1.0 / mass_[con_.idx[2 * k]]) + (1.0 / mass_[con_.idx[2 * k + 1]]));
    }
    std::vector<uint32_t> root(n_);
    for(size_t i = 0; i < n_; ++i) {root[i] = (uint32_t)i;}
    auto find = [&](uint32_t a) {
        while(root[a] != a) {root[a] = root[root[a]]; a = root[a];}
        return a;
    };
    for(size_t k = 0; k < m; ++k)
    {
        const uint32_t a = find(con_.idx[2 * k]), b = find(con_.idx[2 * k + 1]);
        if(a != b) {root[std::max(a, b)] = std::min(a, b);}
    }
    std::map<uint32_t, std::vector<uint32_t>> comps;
    for(size_t k = 0; k < m; ++k) {comps[find(con_.idx[2 * k])].push_back((uint32_t)k);}
    std::vector<uint32_t> start, list;
    for(const auto& kv : comps)
    {
        start.push_back((uint32_t)list.size());
        list.insert(list.end(), kv.second.begin(), kv.second.end());
    }
    start.push_back((uint32_t)list.size());
    uint32_t* d_start = dalloc<uint32_t>(start.size());
    uint32_t* d_list = dalloc<uint32_t>(list.size());
    uint32_t* d_pair = dalloc<uint32_t>(2 * m);
    double* d_sq = dalloc<double>(m);
    double* d_rsq = dalloc<double>(m);
    double* d_redm = dalloc<double>(m);
    upload(d_start, start.data(), start.size());
    upload(d_list, list.data(), list.size());
    upload(d_pair, con_.idx.data(), 2 * m);
    upload(d_sq, sq.data(), m);
    upload(d_rsq, rsq.data(), m);
    upload(d_redm, redm.data(), m);
    ctab_.ncomp = (uint32_t)comps.size();
    ctab_.comp_start = d_start;
    ctab_.comp_con = d_list;
    ctab_.pair = d_pair;
    ctab_.sq_v0 = d_sq;
    ctab_.r_sq_v0 = d_rsq;
    ctab_.redm = d_redm;
}

void Engine::launch_ul_accel(bool kick)
{
    hipLaunchKernelGGL(k_ul_accel, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, (uint32_t)n_, d_vel_, d_accel_,
                       d_force_, d_rmass_, d_c1_, d_c2_, dt_ / 2, seed_, d_sctl_, kick ? 1 : 0);
}

void Engine::launch_step()
{
    const uint32_t nb = blocks_for(n_, 256);
    if(integ_ == MJG_INTEG_UNDERDAMPED)
    {
        // UnderdampedLangevinIntegrator::step (:459-542)
        hipLaunchKernelGGL(k_ul_a, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_pos_, d_vel_, d_accel_, d_c1_, dt_,
                           dt_ / 2, dt_ * dt_ / 2, box_, d_sctl_);
        if(!globals_.empty()) {launch_rebuilds(0);}
        forces(d_force_, false, -1, false, true);
        launch_ul_accel(true);
        if(rm_translation_ || rm_rotation_)
        {
            hipLaunchKernelGGL(k_remove_motion, dim3(1), dim3(kRemoverThreads), 0, stream_, (uint32_t)n_, d_pos_,
                               d_vel_, d_mass_, rm_translation_, rm_rotation_, rm_rescale_);
        }
        return;
    }
    if(integ_ == MJG_INTEG_GBAOAB)
    {
        // gBAOABLangevinIntegrator::step (:348-433)
        const uint32_t nc = blocks_for(ctab_.ncomp, 64);
        const bool con = ctab_.ncomp > 0 && ctab_.max_it > 0;
        auto cvel = [&]() {
            if(con) {hipLaunchKernelGGL(k_con_velocity, dim3(nc), dim3(64), 0, stream_, ctab_, d_pos_, d_vel_, d_rmass_, box_);}
        };
        auto drift = [&]() {
            hipLaunchKernelGGL(k_gb_drift, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_pos_, d_vel_, d_old_rattle_,
                               dt_corr_);
            if(con)
            {
                hipLaunchKernelGGL(k_con_coordinate, dim3(nc), dim3(64), 0, stream_, ctab_, d_pos_, d_vel_,
                                   d_old_rattle_, d_rmass_, box_);
            }
            cvel();
        };
        hipLaunchKernelGGL(k_gb_kick, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_vel_, d_force_, d_rmass_, dt_ / 2);
        cvel();
        drift();
        hipLaunchKernelGGL(k_gb_o, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_vel_, d_c1_, d_c2_, seed_, d_sctl_);
        cvel();
        drift();
        hipLaunchKernelGGL(k_gb_wrap, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_pos_, d_old_pos_, box_, d_sctl_);
        if(!globals_.empty()) {launch_rebuilds(0);}
        forces(d_force_, false, -1, true);
        cvel();
        if(rm_translation_ || rm_rotation_)
        {
            hipLaunchKernelGGL(k_remove_motion, dim3(1), dim3(kRemoverThreads), 0, stream_, (uint32_t)n_, d_pos_,
                               d_vel_, d_mass_, rm_translation_, rm_rotation_, rm_rescale_);
        }
        hipLaunchKernelGGL(k_copy_double4, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_pos_, d_old_pos_);
        return;
    }
    if(integ_ == MJG_INTEG_BAOAB)
    {
        hipLaunchKernelGGL(k_baoab_a, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_pos_, d_vel_, d_force_,
                           d_rmass_, d_c1_, d_c2_, dt_ / 2, box_, seed_, d_sctl_);
    }
    else
    {
        hipLaunchKernelGGL(k_gjf_a, dim3(nb), dim3(256), 0, stream_, (uint32_t)n_, d_pos_, d_vel_, d_force_,
                           d_rmass_, d_c1_, d_c2_, d_c3_, dt_, box_, seed_, d_sctl_);
    }
    if(!globals_.empty()) {launch_rebuilds(0);}
    forces(d_force_, false, -1, true);
    if(rm_translation_ || rm_rotation_)
    {
        hipLaunchKernelGGL(k_remove_motion, dim3(1), dim3(kRemoverThreads), 0, stream_, (uint32_t)n_, d_pos_, d_vel_,
                           d_mass_, rm_translation_, rm_rotation_, rm_rescale_);
    }
}

void Engine::run_segment(size_t n)
{
    const size_t ng = std::max<size_t>(1, globals_.size());
    for(int attempt = 0; attempt < 6; ++attempt)
    {
        check(hipMemcpyAsync(d_bak_pos_, d_pos_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_), "backup");
        check(hipMemcpyAsync(d_bak_vel_, d_vel_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_), "backup");
        check(hipMemcpyAsync(d_bak_force_, d_force_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_), "backup");
        check(hipMemcpyAsync(d_bak_sctl_, d_sctl_, sizeof(StepCtl), hipMemcpyDeviceToDevice, stream_), "backup");
        if(integ_ == MJG_INTEG_UNDERDAMPED)
        {
            check(hipMemcpyAsync(d_bak_accel_, d_accel_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_),
                  "backup");
        }
        if(!globals_.empty())
        {
            check(hipMemcpyAsync(d_bak_lctl_, d_lctl_, ng * sizeof(ListCtl), hipMemcpyDeviceToDevice, stream_), "backup");
        }
        if(use_graph_)
        {
            if(graph_exec_ == nullptr || graph_steps_ != n)
            {
                if(graph_exec_) {(void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr;}
                hipGraph_t graph = nullptr;
                check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
                for(size_t s = 0; s < n; ++s) {launch_step();}
                check(hipStreamEndCapture(stream_, &graph), "end capture");
                check(hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0), "graph instantiate");
                (void)hipGraphDestroy(graph);
                graph_steps_ = n;
            }
            check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
        }
        else
        {
            for(size_t s = 0; s < n; ++s) {launch_step();}
            check(hipGetLastError(), "step launch");
        }
        StepCtl sc{};
        download(&sc, d_sctl_, 1);
        bool overflow = false;
        if(!globals_.empty())
        {
            std::vector<ListCtl> ctl(globals_.size());
            download(ctl.data(), d_lctl_, globals_.size());
            for(size_t g = 0; g < globals_.size(); ++g)
            {
                if(ctl[g].maxcount > globals_[g].cap)
                {
                    alloc_list(globals_[g], (uint32_t)(ctl[g].maxcount + ctl[g].maxcount / 4 + 16));
                    overflow = true;
                }
            }
        }
        if(!overflow)
        {
            if(sc.nonfinite) {throw NonFinite("non-finite particle position");}
            grow_tiles();
            return;
        }
        // a partner list did not fit: restore the segment's start, force a rebuild with the
        // larger capacity and run the segment again
        if(graph_exec_) {(void)hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr;}
        check(hipMemcpyAsync(d_pos_, d_bak_pos_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_), "restore");
        check(hipMemcpyAsync(d_vel_, d_bak_vel_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_), "restore");
        check(hipMemcpyAsync(d_force_, d_bak_force_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_), "restore");
        check(hipMemcpyAsync(d_sctl_, d_bak_sctl_, sizeof(StepCtl), hipMemcpyDeviceToDevice, stream_), "restore");
        if(integ_ == MJG_INTEG_UNDERDAMPED)
        {
            check(hipMemcpyAsync(d_accel_, d_bak_accel_, n_ * sizeof(double4), hipMemcpyDeviceToDevice, stream_),
                  "restore");
        }
        check(hipMemcpyAsync(d_lctl_, d_bak_lctl_, ng * sizeof(ListCtl), hipMemcpyDeviceToDevice, stream_), "restore");
        if(integ_ == MJG_INTEG_GBAOAB)
        {
            hipLaunchKernelGGL(k_copy_double4, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, (uint32_t)n_, d_pos_,
                               d_old_pos_);
        }
        sync();
        build_lists_now(true);
        forces(d_force_, false, -1, false);
        sync();
    }
    throw DeviceError("neighbour list capacity did not converge");
}

void Engine::step(size_t n)
{
    if(!integ_ready_) {throw std::logic_error("integrator not initialized");}
    const size_t seg = 100;
    while(n > 0)
    {
        const size_t k = std::min(n, seg);
        run_segment(k);
        n -= k;
    }
}

// the position of the noise stream (a restored checkpoint continues the stream where it stopped)
void Engine::set_step_count(uint64_t step)
{
    StepCtl sc{};
    download(&sc, d_sctl_, 1);
    sc.step = step;
    upload(d_sctl_, &sc, 1);
    if(integ_ready_ && integ_ == MJG_INTEG_UNDERDAMPED)
    {
        // the stored acceleration carries the noise of the count it was drawn at
        launch_ul_accel(false);
        check(hipGetLastError(), "set_step_count launch");
        sync();
    }
}

void Engine::counters(uint64_t* steps, uint64_t* rebuilds)
{
    StepCtl sc{};
    download(&sc, d_sctl_, 1);
    if(steps) {*steps = sc.step;}
    if(rebuilds) {*rebuilds = sc.rebuilds;}
}

// ------------------------------------------------------------------------------------
// observers and the simulator loop
// ------------------------------------------------------------------------------------

// the energy-column name each interaction prints (LocalInteractionBase::name / GlobalInteraction
// name(): BondLengthInteraction.hpp:88, BondLengthGoContactInteraction.hpp:102,
// GoContactInteraction.hpp:134, BondAngleInteraction.hpp:100, DihedralAngleInteraction.hpp:95,
// GlobalPairLennardJonesInteraction.hpp:134, GlobalPairExcludedVolumeInteraction.hpp:134,
// GlobalPairInteraction.hpp:112)
std::string Engine::interaction_name(size_t id) const
{
    if(id < locals_.size())
    {
        switch(locals_[id].kind)
        {
            case MJG_BOND_HARMONIC: return "BondLength:Harmonic";
            case MJG_BOND_GOCONTACT: return "BondLengthGoContact";
            case MJG_CONTACT_GOCONTACT: return "Contact:GoContact";
            case MJG_ANGLE_HARMONIC: return "BondAngle:Harmonic";
            case MJG_ANGLE_GAUSSIAN: return "BondAngle:Gaussian";
            case MJG_DIHEDRAL_CLEMENTI: return "DihedralAngle:ClementiDihedral";
            case MJG_BOND_3SPN2: return "BondLength:ThreeSPN2BondPotential";
            case MJG_DIHEDRAL_GAUSSIAN: return "DihedralAngle:PeriodicGaussian";
            case MJG_BASE_STACKING_3SPN2: return "3SPN2BaseStacking";
            default: return "Unknown";
        }
    }
    switch(globals_.at(id - locals_.size()).pot)
    {
        case MJG_PAIR_LJ: return "GlobalPairLennardJones";
        case MJG_PAIR_EV: return "GlobalPairExcludedVolume";
        case MJG_PAIR_DH: return "Pair:DebyeHuckel";
        case MJG_PAIR_3SPN2_EV: return "Pair:3SPN2ExcludedVolume";
        case MJG_3SPN2_BASEPAIR: return "3SPN2BasePair";
        case MJG_3SPN2_CROSSSTACKING: return "3SPN2CrossStacking";
        default: return "Unknown";
    }
}

// EnergyObserver constructor + initialize (EnergyObserver.hpp:27-63): truncate prefix.ene and
// write the unit line and the column header.  System attributes print in key order, as the
// reference's std::map iterates them.
void Engine::energy_observer_open(const std::string& prefix)
{
    if(!initialized_) {throw std::logic_error("force field not initialized");}
    const std::string path = prefix + ".ene";
    std::ofstream ofs(path);
    if(!ofs.good()) {throw std::runtime_error("[error] EnergyObserver: file open error: " + path);}
    ofs << "# unit of length : angstrom, unit of energy : kcal/mol\n";
    ofs << "# timestep  ";
    for(size_t k = 0; k < num_interactions(); ++k) {ofs << interaction_name(k) << ' ';}
    ofs << " kinetic_energy";
    if(has_I_) {ofs << " attribute:ionic_strength";}
    if(has_T_) {ofs << " attribute:temperature";}
    ofs << '\n';
    ene_path_ = path;
}

// EnergyObserver::output (EnergyObserver.hpp:87-150) with ForceField::format_energy
// (LocalForceField.hpp:173-200, GlobalForceField.hpp:170-191): step, one fixed-format column per
// interaction (width max(name length, 10)), the kinetic energy, the attributes.  Returns false
// when a value is not finite (the reference writes the line, then throws).
bool Engine::energy_observer_output(uint64_t step)
{
    if(ene_path_.empty()) {throw std::logic_error("no energy observer is open");}
    const size_t ni = num_interactions();
    std::vector<double> e(std::max<size_t>(1, ni));
    double total = 0;
    calc_energies(e.data(), &total);
    const double ek = kinetic_energy();
    std::ostringstream line;
    line << std::setw(11) << std::left << std::to_string(step) << ' ';
    bool ok = std::isfinite(total) && std::isfinite(ek);
    for(size_t k = 0; k < ni; ++k)
    {
        const size_t w = std::max<size_t>(interaction_name(k).size(), 10);
        line << std::setw((int)w) << std::fixed << std::right << e[k] << ' ';
    }
    line << std::setw(14) << std::right << std::fixed << ek;
    if(has_I_) {line << ' ' << std::setw(10 + 14) << std::right << std::fixed << I_;}
    if(has_T_) {line << ' ' << std::setw(10 + 11) << std::right << std::fixed << T_;}
    std::ofstream ofs(ene_path_, std::ios::app);
    ofs << line.str() << std::endl;
    return ok;
}

// MolecularDynamicsSimulator::initialize/run/finalize (MolecularDynamicsSimulator.hpp:72-123):
// the observers' initialize, output at every step that is a multiple of save_step and after the
// last step, then finalize; the steps between outputs run on the device in one call.  The
// trajectory observers write before the energy observer, which read_observer appends last
// (read_observer.hpp:84-98), so a step whose energy is not finite still gets its frame.
void Engine::simulate(uint64_t total_step, uint64_t save_step)
{
    if(!integ_ready_) {throw std::logic_error("integrator not initialized");}
    if(save_step == 0) {throw std::invalid_argument("save_step must be positive");}
    if(!dcd_prefix_.empty()) {dcd_begin(total_step, save_step);}
    auto output = [&](uint64_t step) {
        trajectory_output(step);
        if(!ene_path_.empty() && !energy_observer_output(step)) {throw NonFinite("Energy value becomes NaN");}
    };
    uint64_t done = 0;
    while(done < total_step)
    {
        if(done % save_step == 0) {output(done);}
        const uint64_t next = std::min<uint64_t>(total_step, (done / save_step + 1) * save_step);
        step((size_t)(next - done));
        done = next;
    }
    output(total_step);
    if(!dcd_prefix_.empty()) {dcd_end();}
}

double Engine::kinetic_energy()
{
    hipLaunchKernelGGL(k_kinetic, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, (uint32_t)n_, d_vel_, d_mass_,
                       d_scratch_);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, stream_, d_scratch_, (uint32_t)n_, d_energy_ + num_interactions());
    check(hipGetLastError(), "kinetic launch");
    double ke = 0;
    download(&ke, d_energy_ + num_interactions(), 1);
    return ke;
}

// ------------------------------------------------------------------------------------
// Allocation-free kinetic energy: the per-particle terms and their sum live in a file-scope device
// array, `scratch`, instead of the engine's scratch buffer.  Every GPU the library runs on holds
// its own copy (8 MiB) until the process ends; its address on a device is looked up once, with
// that device current, and kept.  Calls on one GPU share the array: overlapping calls race.
// Nothing in it has to start at zero (k_kinetic writes every term it sums, k_sum the result).
// ------------------------------------------------------------------------------------
constexpr size_t kStaticMaxParticles = size_t(1) << 20;
static __device__ __align__(256) double scratch[kStaticMaxParticles + 1];

namespace
{
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}

static double* static_scratch(int device)
{
    if(device < 0 || device >= kMaxDevices) {throw std::out_of_range("device ordinal out of range");}
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if(g_scratch_ptr[device] == nullptr)
    {
        auto fail = [](hipError_t e, const char* what) {
            if(e != hipSuccess) {throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));}
        };
        int prev = 0;
        fail(hipGetDevice(&prev), "hipGetDevice");
        fail(hipSetDevice(device), "hipSetDevice");
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        fail(e, "hipGetSymbolAddress(scratch)");
        if(p == nullptr) {throw DeviceError("hipGetSymbolAddress(scratch) returned no address");}
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    return g_scratch_ptr[device];
}

double Engine::kinetic_energy_static()
{
    if(n_ > kStaticMaxParticles)
    {
        throw std::invalid_argument("mjg_kinetic_energy_static: " + std::to_string(n_) +
                                    " particles exceed its static scratch of 1048576; use mjg_kinetic_energy");
    }
    double* s = static_scratch(device_);
    hipLaunchKernelGGL(k_kinetic, dim3(blocks_for(n_, 256)), dim3(256), 0, stream_, (uint32_t)n_, d_vel_, d_mass_, s);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, stream_, s, (uint32_t)n_, s + kStaticMaxParticles);
    check(hipGetLastError(), "kinetic launch");
    double ke = 0;
    download(&ke, s + kStaticMaxParticles, 1);
    return ke;
}

double Engine::profile_interaction(int id, int reps, uint64_t* half_pairs)
{
    if(!initialized_) {throw std::logic_error("not initialized");}
    if(id < 0 || (size_t)id >= num_interactions() || reps < 1) {throw std::out_of_range("bad profile request");}
    hipEvent_t a, b;
    check(hipEventCreate(&a), "event");
    check(hipEventCreate(&b), "event");
    const int nl = (int)locals_.size();
    uint64_t hp = 0;
    check(hipEventRecord(a, stream_), "event record");
    const PairSegTable pt = pair_table(id);
    uint32_t pblocks = 0;
    for(int k = 0; k < pt.n; ++k) {pblocks += blocks_for((size_t)pt.seg[k].np * pt.seg[k].tpp, 256);}
    for(int r = 0; r < reps; ++r)
    {
        if(id < nl)
        {
            hipLaunchKernelGGL(k_local_terms<false>, dim3(local_blocks_), dim3(256), 0, stream_, d_pos_, box_, segtab_,
                               d_lidx_, d_lpar_, d_tf_, d_te_, id);
        }
        else if(globals_[id - nl].multibody()) {launch_multibody(globals_[id - nl], false);}
        else if(globals_[id - nl].tiled) {launch_tiled(globals_[id - nl], id - nl, false);}
        else if(precision_ == MJG_FP32_FIXED)
        {
            hipLaunchKernelGGL(k_pairs_mixed<false>, dim3(pblocks), dim3(256), 0, stream_, d_pos_, box_, pt);
        }
        else
        {
            hipLaunchKernelGGL(k_pairs<false>, dim3(pblocks), dim3(256), 0, stream_, d_pos_, box_, pt);
        }
    }
    check(hipEventRecord(b, stream_), "event record");
    check(hipEventSynchronize(b), "event sync");
    check(hipGetLastError(), "profile launch");
    float ms = 0;
    check(hipEventElapsedTime(&ms, a, b), "event elapsed");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if(id >= nl)
    {
        GlobalIA& g = globals_[id - nl];
        const uint32_t np = (uint32_t)g.part.size();
        std::vector<uint32_t> nn(np);
        download(nn.data(), g.d_nn, np);
        for(uint32_t c : nn) {hp += c;}
        hp /= 2;
    }
    if(half_pairs) {*half_pairs = hp;}
    return (double)ms / reps;
}

} // mjg

// ======================================================================================
// C ABI
// ======================================================================================
struct mjg_engine
{
    mjg::Engine* impl = nullptr;
    std::string err;
};

namespace
{
template<typename F>
mjg_status guarded(mjg_engine* e, F&& f)
{
    if(e == nullptr || e->impl == nullptr) {return MJG_INVALID_ARG;}
    try
    {
        f(*e->impl);
        e->err.clear();
        return MJG_OK;
    }
    catch(const mjg::NonFinite& x) {e->err = x.what(); return MJG_NAN_DETECTED;}
    catch(const mjg::DeviceError& x) {e->err = x.what(); return MJG_DEVICE_ERROR;}
    catch(const std::bad_alloc&) {e->err = "device or host allocation failed"; return MJG_OOM;}
    catch(const std::logic_error& x)
    {
        e->err = x.what();
        if(dynamic_cast<const std::invalid_argument*>(&x) || dynamic_cast<const std::out_of_range*>(&x))
        {
            return MJG_INVALID_ARG;
        }
        return MJG_NOT_INITIALIZED;
    }
    catch(const std::exception& x) {e->err = x.what(); return MJG_INVALID_ARG;}
}
} // namespace

extern "C" {

const char* mjg_version(void) {return "mjg 0.1 (gfx950)";}

mjg_status mjg_engine_create(int precision, int boundary, const double* lower, const double* upper,
                             size_t n_particles, int device, mjg_engine** out)
{
    if(out == nullptr) {return MJG_INVALID_ARG;}
    *out = nullptr;
    mjg_engine* e = new(std::nothrow) mjg_engine;
    if(e == nullptr) {return MJG_OOM;}
    try
    {
        e->impl = new mjg::Engine(precision, boundary, lower, upper, n_particles, device);
    }
    catch(const std::bad_alloc&) {delete e; return MJG_OOM;}
    catch(const mjg::DeviceError&) {delete e; return MJG_DEVICE_ERROR;}
    catch(const std::exception&) {delete e; return MJG_INVALID_ARG;}
    *out = e;
    return MJG_OK;
}

void mjg_engine_destroy(mjg_engine* e)
{
    if(e == nullptr) {return;}
    delete e->impl;
    delete e;
}

const char* mjg_last_error(const mjg_engine* e) {return e ? e->err.c_str() : "null engine";}

mjg_status mjg_system_set_state(mjg_engine* e, const double* mass, const double* pos, const double* vel)
{
    return guarded(e, [&](mjg::Engine& x) {x.set_state(mass, pos, vel);});
}
mjg_status mjg_system_set_positions(mjg_engine* e, const double* pos)
{
    return guarded(e, [&](mjg::Engine& x) {x.set_positions(pos);});
}
mjg_status mjg_system_get_state(mjg_engine* e, double* pos, double* vel, double* force)
{
    return guarded(e, [&](mjg::Engine& x) {x.get_state(pos, vel, force);});
}
mjg_status mjg_system_set_attribute(mjg_engine* e, const char* key, double value)
{
    if(key == nullptr) {return MJG_INVALID_ARG;}
    return guarded(e, [&](mjg::Engine& x) {x.set_attribute(key, value);});
}
mjg_status mjg_add_local(mjg_engine* e, int kind, const char* topology, size_t n_terms, const uint32_t* indices,
                         const double* params, double margin, int* id)
{
    return guarded(e, [&](mjg::Engine& x) {
        const int k = x.add_local(kind, topology, n_terms, indices, params, margin);
        if(id) {*id = k;}
    });
}
mjg_status mjg_add_global(mjg_engine* e, int potential, double cutoff_ratio, double epsilon, int partition,
                          double margin, int ignore_molecule, size_t n_within, const char* const* within_kinds,
                          const uint32_t* within_dist, size_t n_participants, const uint32_t* participants,
                          const double* params, int* id)
{
    return guarded(e, [&](mjg::Engine& x) {
        const int k = x.add_global(potential, cutoff_ratio, epsilon, partition, margin, ignore_molecule, n_within,
                                   within_kinds, within_dist, n_participants, participants, params);
        // interaction ids follow calc_energy order: locals first, then globals
        if(id) {*id = (int)x.num_locals() + k;}
    });
}
mjg_status mjg_set_step_count(mjg_engine* e, uint64_t step)
{
    return guarded(e, [&](mjg::Engine& x) {x.set_step_count(step);});
}
mjg_status mjg_add_constraint(mjg_engine* e, const char* topology, size_t n, const uint32_t* indices,
                              const double* v0, size_t max_iteration, double tolerance)
{
    return guarded(e, [&](mjg::Engine& x) {x.add_constraint(topology, n, indices, v0, max_iteration, tolerance);});
}
mjg_status mjg_forcefield_initialize(mjg_engine* e)
{
    return guarded(e, [&](mjg::Engine& x) {x.initialize();});
}
size_t mjg_num_interactions(const mjg_engine* e) {return (e && e->impl) ? e->impl->num_interactions() : 0;}
mjg_status mjg_calc_force(mjg_engine* e) {return guarded(e, [&](mjg::Engine& x) {x.calc_force();});}
mjg_status mjg_calc_force_and_energy(mjg_engine* e, double* energy)
{
    return guarded(e, [&](mjg::Engine& x) {const double v = x.calc_force_and_energy(); if(energy) {*energy = v;}});
}
mjg_status mjg_calc_energies(mjg_engine* e, double* per_interaction, double* total)
{
    return guarded(e, [&](mjg::Engine& x) {x.calc_energies(per_interaction, total);});
}
mjg_status mjg_calc_force_and_virial(mjg_engine* e, double* virial)
{
    // no kernel accumulates the virial tensor (no integrator on this path reads it)
    if(e == nullptr || e->impl == nullptr) {return MJG_INVALID_ARG;}
    if(virial) {for(int k = 0; k < 9; ++k) {virial[k] = __builtin_nan("");}}
    e->err = "the virial is not computed by this library (constant-volume integrators only)";
    return MJG_UNSUPPORTED;
}
mjg_status mjg_interaction_name(const mjg_engine* e, int id, char* buf, size_t cap, size_t* len)
{
    if(e == nullptr || e->impl == nullptr || id < 0 || (size_t)id >= e->impl->num_interactions()) {return MJG_INVALID_ARG;}
    const std::string nm = e->impl->interaction_name((size_t)id);
    if(len) {*len = nm.size();}
    if(buf && cap > 0)
    {
        const size_t k = std::min(cap - 1, nm.size());
        std::memcpy(buf, nm.data(), k);
        buf[k] = '\0';
    }
    return MJG_OK;
}
mjg_status mjg_calc_force_of(mjg_engine* e, int id, double* force_out)
{
    if(force_out == nullptr) {return MJG_INVALID_ARG;}
    return guarded(e, [&](mjg::Engine& x) {x.calc_force_of(id, force_out);});
}
mjg_status mjg_reduce_margin(mjg_engine* e, double dmargin)
{
    return guarded(e, [&](mjg::Engine& x) {x.reduce_margin(dmargin);});
}
mjg_status mjg_neighbor_pairs(mjg_engine* e, int id, uint32_t* pairs, size_t capacity, size_t* n_pairs)
{
    return guarded(e, [&](mjg::Engine& x) {
        const size_t n = x.neighbor_pairs(id, pairs, capacity);
        if(n_pairs) {*n_pairs = n;}
    });
}
mjg_status mjg_integrator_set(mjg_engine* e, int type, double dt, const double* coefs, uint64_t seed)
{
    return guarded(e, [&](mjg::Engine& x) {x.set_integrator(type, dt, coefs, seed);});
}
mjg_status mjg_observer_energy_open(mjg_engine* e, const char* prefix)
{
    if(!prefix) {return MJG_INVALID_ARG;}
    return guarded(e, [&](mjg::Engine& x) {x.energy_observer_open(prefix);});
}
mjg_status mjg_observer_energy_output(mjg_engine* e, uint64_t step)
{
    return guarded(e, [&](mjg::Engine& x) {
        if(!x.energy_observer_output(step)) {throw mjg::NonFinite("Energy value becomes NaN");}
    });
}
mjg_status mjg_observer_xyz_open(mjg_engine* e, const char* prefix, const char* const* names)
{
    if(!prefix) {return MJG_INVALID_ARG;}
    return guarded(e, [&](mjg::Engine& x) {x.xyz_observer_open(prefix, names);});
}
mjg_status mjg_observer_dcd_open(mjg_engine* e, const char* prefix)
{
    if(!prefix) {return MJG_INVALID_ARG;}
    return guarded(e, [&](mjg::Engine& x) {x.dcd_observer_open(prefix);});
}
mjg_status mjg_simulate(mjg_engine* e, uint64_t total_step, uint64_t save_step)
{
    return guarded(e, [&](mjg::Engine& x) {x.simulate(total_step, save_step);});
}
mjg_status mjg_integrator_set_remover(mjg_engine* e, int translation, int rotation, int rescale)
{
    return guarded(e, [&](mjg::Engine& x) {x.set_remover(translation, rotation, rescale);});
}
mjg_status mjg_integrator_initialize(mjg_engine* e)
{
    return guarded(e, [&](mjg::Engine& x) {x.integrator_initialize();});
}
mjg_status mjg_step(mjg_engine* e, size_t n_steps)
{
    return guarded(e, [&](mjg::Engine& x) {x.step(n_steps);});
}
mjg_status mjg_counters(mjg_engine* e, uint64_t* steps, uint64_t* rebuilds)
{
    return guarded(e, [&](mjg::Engine& x) {x.counters(steps, rebuilds);});
}
mjg_status mjg_kinetic_energy(mjg_engine* e, double* ke)
{
    return guarded(e, [&](mjg::Engine& x) {const double v = x.kinetic_energy(); if(ke) {*ke = v;}});
}
mjg_status mjg_kinetic_energy_static(mjg_engine* e, double* ke)
{
    return guarded(e, [&](mjg::Engine& x) {const double v = x.kinetic_energy_static(); if(ke) {*ke = v;}});
}

mjg_status mjg_profile_interaction(mjg_engine* e, int id, int reps, double* avg_ms, uint64_t* half_pairs)
{
    return guarded(e, [&](mjg::Engine& x) {
        const double v = x.profile_interaction(id, reps, half_pairs);
        if(avg_ms) {*avg_ms = v;}
    });
}

} // extern "C"
