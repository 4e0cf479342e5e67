// Host side of the batched replica engine (kernels in batch_kernels.hpp) and its C ABI.
//
// A batch holds independent replicas (each one a System + ForceField + Integrator of its own,
// as the reference would hold in separate processes).  Replicas can be added while others are
// resident: each one owns a device arena sized by its footprint, and every launch runs the
// replicas that are active, largest first, one workgroup each.
#include "batch_kernels.hpp"
#include "topology.hpp"
#include "../../include/mjg.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace mjg
{
// the segment kernels are compiled in batch_seg.hip
MJG_REP_ALL_INSTANCES(extern)

struct BatchDeviceError : std::runtime_error {using std::runtime_error::runtime_error;};
struct BatchNonFinite : std::runtime_error {using std::runtime_error::runtime_error;};

struct RepLocal
{
    int kind = 0, arity = 0, npar = 0, topo = -1;
    std::vector<uint32_t> idx;
    std::vector<double> par;
};
struct RepGlobalIn
{
    int pot = 0, partition = 0, ignore_mol = 0, npar = 0;
    double cutoff_ratio = 0, eps = 0, margin = 0;
    std::vector<std::pair<int, uint32_t>> within;
    std::vector<uint32_t> part;
    std::vector<double> par;
    uint32_t cap = 0;
};

// Launch parameters of one replica that change no result (forces are exact fixed-point sums over
// the pairs inside the cutoff, whatever list holds them): the replica kernel's own list skin for
// LJ / excluded-volume pairs and for Debye-Hueckel (never above the input's margin), the cell-grid
// resolution of its list builds (16 or 8 cells per dimension: the cell size) and the lanes that
// share one row of a grid build.  Chosen per replica by Batch::tuned (or set explicitly).
struct RepTune
{
    double m_short = 0.4;
    double m_dh = 0.1;
    int grid_bits = 4;
    int grid_lanes = 1;
};

struct Replica
{
    // host description
    Box box{};
    size_t n = 0;
    std::vector<double> mass, pos, vel;
    double T = 0, I = 0;
    bool has_T = false, has_I = false;
    Topo topo;
    std::vector<RepLocal> locals;
    std::vector<RepGlobalIn> globals;
    int integ = 0;
    double dt = 0;
    uint64_t seed = 0;
    std::vector<double> coefs;
    // device
    void* arena = nullptr;
    size_t footprint = 0;
    RepDesc desc{};
    int slot = -1;                 // index into the batch's state array
    size_t lds = 0;                // dynamic LDS of its workgroup
    int cls = 0;                   // size class (block shape)
    bool ready = false, active = false;
    bool observed = false;          // an energy observation is held for this replica
    bool failed = false;            // stopped with a non-finite position or force; never stepped again
    bool tsp = false;               // has 3SPN2 terms: runs in the kernel variant that holds them
    RepTune tune;                   // launch parameters (see RepTune)
    bool tune_set = false;          // set explicitly or tuned already
};

static void hip_check(hipError_t e, const char* what)
{
    if(e != hipSuccess) {throw BatchDeviceError(std::string(what) + ": " + hipGetErrorString(e));}
}
static size_t align256(size_t x) {return (x + 255) & ~(size_t)255;}

// Skin of a replica's neighbour list.  The replica kernel's forces are exact fixed-point sums
// over the pairs within the cutoff, so the skin changes no result, only the balance between
// the pair loop (list entries grow as (1 + m)^3) and the rebuilds (their rate grows as 1 / m):
// the kernel uses its own margin, at most the input's (RepTune; untuned defaults 0.4 for the
// short-ranged LJ and excluded-volume pairs, 0.1 for Debye-Hueckel, whose cutoff spans many
// particle spacings; MJG_LIST_MARGIN_SHORT / MJG_LIST_MARGIN_DH override those defaults).
static RepTune default_tune()
{
    RepTune t;
    if(const char* c = std::getenv("MJG_LIST_MARGIN_SHORT")) {t.m_short = std::atof(c);}
    if(const char* c = std::getenv("MJG_LIST_MARGIN_DH")) {t.m_dh = std::atof(c);}
    return t;
}
static double list_margin(int pot, double input_margin, const RepTune& t)
{
    double m = input_margin;
    if(pot == kPairLJ || pot == kPairEV) {m = std::min(m, t.m_short);}
    else if(pot == kPairDH) {m = std::min(m, t.m_dh);}
    return std::max(m, 0.02);
}

class Batch
{
  public:
    Batch(int precision, int device) : precision_(precision)
    {
        if(precision != MJG_FP32_FIXED && precision != MJG_FP64) {throw std::invalid_argument("unknown precision");}
        hip_check(hipSetDevice(device), "hipSetDevice");
        for(int c = 0; c < kRepClasses; ++c)
        {
            hip_check(hipStreamCreateWithFlags(&cls_stream_[c], hipStreamNonBlocking), "hipStreamCreate");
        }
        stream_ = cls_stream_[0];
        hip_check(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming), "hipEventCreate");
        for(int c = 0; c < kRepClasses; ++c) {hip_check(hipEventCreateWithFlags(&ev_join_[c], hipEventDisableTiming), "hipEventCreate");}
        const char* seg = std::getenv("MJG_SEGMENT");
        if(seg) {segment_ = std::max(1, std::atoi(seg));}
        trace_ = std::getenv("MJG_BATCH_TRACE") != nullptr;
        if(const char* a = std::getenv("MJG_AUTOTUNE")) {autotune_ = a[0] != '0';}
        set_lds_limits<float>();
        set_lds_limits<double>();
    }
    ~Batch()
    {
        for(int c = 0; c < kRepClasses; ++c) {if(cls_stream_[c]) {(void)hipStreamSynchronize(cls_stream_[c]);}}
        for(auto& r : reps_) {if(r && r->arena) {(void)hipFree(r->arena);}}
        if(d_descs_) {(void)hipFree(d_descs_);}
        if(d_states_) {(void)hipFree(d_states_);}
        if(ev_fork_) {(void)hipEventDestroy(ev_fork_);}
        for(int c = 0; c < kRepClasses; ++c) {if(ev_join_[c]) {(void)hipEventDestroy(ev_join_[c]);}}
        for(int c = 0; c < kRepClasses; ++c) {if(cls_stream_[c]) {(void)hipStreamDestroy(cls_stream_[c]);}}
    }

    int begin_replica(int boundary, const double* lower, const double* upper, size_t n, const double* mass,
                      const double* pos, const double* vel)
    {
        if(n == 0 || n > kRepMaxN) {throw std::invalid_argument("replica size must be 1.." + std::to_string(kRepMaxN));}
        if(!mass || !pos || !vel) {throw std::invalid_argument("null state pointer");}
        auto r = std::make_unique<Replica>();
        r->n = n;
        r->box.periodic = (boundary == MJG_PERIODIC);
        if(boundary != MJG_PERIODIC && boundary != MJG_UNLIMITED) {throw std::invalid_argument("unknown boundary");}
        if(r->box.periodic)
        {
            if(!lower || !upper) {throw std::invalid_argument("periodic boundary needs lower/upper");}
            for(int d = 0; d < 3; ++d)
            {
                if(!(upper[d] > lower[d])) {throw std::invalid_argument("upper must exceed lower");}
                r->box.lower[d] = lower[d]; r->box.upper[d] = upper[d];
                r->box.width[d] = upper[d] - lower[d]; r->box.halfw[d] = r->box.width[d] * 0.5;
            }
        }
        r->mass.assign(mass, mass + n);
        r->pos.assign(pos, pos + 3 * n);
        r->vel.assign(vel, vel + 3 * n);
        for(size_t i = 0; i < n; ++i) {if(!(mass[i] > 0)) {throw std::invalid_argument("masses must be positive");}}
        r->topo.reset(n);
        reps_.push_back(std::move(r));
        return (int)reps_.size() - 1;
    }

    Replica& rep(int id)
    {
        if(id < 0 || (size_t)id >= reps_.size() || !reps_[id]) {throw std::out_of_range("replica id out of range");}
        return *reps_[id];
    }

    void set_attribute(int id, const std::string& key, double v)
    {
        Replica& r = rep(id);
        if(key == "temperature") {r.T = v; r.has_T = true;}
        else if(key == "ionic_strength") {r.I = v; r.has_I = true;}
        else {throw std::invalid_argument("unknown attribute: " + key);}
    }

    void add_local(int id, int kind, const char* topology, size_t nt, const uint32_t* idx, const double* par)
    {
        Replica& r = rep(id);
        if(r.ready) {throw std::invalid_argument("replica already initialized");}
        if(!r.globals.empty()) {throw std::invalid_argument("add local interactions before global ones");}
        if(r.locals.size() >= (size_t)kRepMaxLocals) {throw std::invalid_argument("too many local interactions");}
        RepLocal L;
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
        if(nt > 0 && (!idx || !par)) {throw std::invalid_argument("null term arrays");}
        for(size_t t = 0; t < nt * L.arity; ++t)
        {
            // a base-stacking term may lack its phosphates (positions 3 and 5)
            const bool optional = (kind == kBaseStacking) && (t % 6 == 3 || t % 6 == 5);
            if(idx[t] >= r.n && !(optional && idx[t] == kNil)) {throw std::out_of_range("term index out of range");}
        }
        if(kind == kBaseStacking)
        {
            for(size_t t = 0; t < nt; ++t)
            {
                if(!(par[t] >= 0 && par[t] < 16)) {throw std::invalid_argument("3SPN2 stacking kind out of range");}
            }
        }
        const std::string topo = topology ? topology : "";
        L.topo = (topo.empty() || topo == "none") ? -1 : r.topo.kind_id(topo);
        L.idx.assign(idx, idx + nt * L.arity);
        L.par.assign(par, par + nt * L.npar);
        r.locals.push_back(std::move(L));
    }

    void add_global(int id, int pot, double cutoff_ratio, double eps, int partition, double margin, int ignore_mol,
                    size_t n_within, const char* const* kinds, const uint32_t* dists, size_t np, const uint32_t* part,
                    const double* par)
    {
        Replica& r = rep(id);
        if(r.ready) {throw std::invalid_argument("replica already initialized");}
        if(r.globals.size() >= (size_t)kRepMaxGlobals) {throw std::invalid_argument("too many global interactions");}
        RepGlobalIn G;
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
        if(np < 1 || !part || !par) {throw std::invalid_argument("a pair interaction needs participants");}
        if(ignore_mol < 0 || ignore_mol > 2) {throw std::invalid_argument("unknown ignore.molecule");}
        for(size_t k = 0; k < np; ++k) {if(part[k] >= r.n) {throw std::out_of_range("participant index out of range");}}
        G.cutoff_ratio = cutoff_ratio; G.eps = eps; G.partition = partition; G.margin = margin; G.ignore_mol = ignore_mol;
        for(size_t k = 0; k < n_within; ++k)
        {
            if(!kinds || !kinds[k] || !dists) {throw std::invalid_argument("null ignore.particles_within entry");}
            G.within.emplace_back(r.topo.kind_id(kinds[k]), dists[k]);
        }
        G.part.assign(part, part + np);
        G.par.assign(par, par + np * G.npar);
        for(size_t k = 0; k < np; ++k)
        {
            const double* q = G.par.data() + k * G.npar;
            if(pot == kPair3SPN2EV && !(q[0] >= 0 && q[0] < 6)) {throw std::invalid_argument("3SPN2 bead kind out of range");}
            if(pot == kBasePair || pot == kCrossStack)
            {
                if(!(q[0] >= 0 && q[0] < 4)) {throw std::invalid_argument("3SPN2 base kind out of range");}
                if(!(q[1] >= 0 && q[1] < (double)r.n)) {throw std::out_of_range("3SPN2 sugar index out of range");}
            }
            if(pot == kCrossStack)
            {
                for(int c = 2; c < 4; ++c)
                {
                    if(!(q[c] == -1.0 || (q[c] >= 0 && q[c] < (double)r.n)))
                    {
                        throw std::out_of_range("3SPN2 neighbouring base index out of range");
                    }
                }
            }
        }
        r.globals.push_back(std::move(G));
    }

    void set_integrator(int id, int type, double dt, const double* coefs, uint64_t seed)
    {
        Replica& r = rep(id);
        if(type != MJG_INTEG_BAOAB && type != MJG_INTEG_GJF && type != MJG_INTEG_UNDERDAMPED)
        {
            throw std::invalid_argument("unknown integrator");
        }
        if(!(dt > 0) || !coefs) {throw std::invalid_argument("invalid integrator parameters");}
        r.integ = type; r.dt = dt; r.seed = seed;
        r.coefs.assign(coefs, coefs + r.n);
    }

    // Prepare every replica that is not resident yet, make it active, build its lists and
    // its first forces.
    void initialize()
    {
        std::vector<int> fresh;
        for(size_t id = 0; id < reps_.size(); ++id)
        {
            if(reps_[id] && !reps_[id]->ready)
            {
                Replica& r = *reps_[id];
                if(!r.tune_set) {r.tune = autotune_ ? tuned(r) : default_tune(); r.tune_set = true;}
                prepare((int)id);
                fresh.push_back((int)id);
            }
        }
        if(fresh.empty()) {return;}
        settle(fresh);
        download_states();
        for(int id : fresh)
        {
            Replica& r = *reps_[id];
            r.active = true;
            // non-finite input (positions or first forces): the replica never steps
            if(h_states_[r.slot].nonfinite) {r.failed = true; r.active = false; failed_now_.push_back(id);}
        }
        descs_dirty_ = true;
    }

    // launch parameters of a replica (set before initialize; overrides the tuner)
    void set_tuning(int id, double m_short, double m_dh, int grid_bits, int grid_lanes)
    {
        Replica& r = rep(id);
        if(r.ready) {throw std::invalid_argument("replica already initialized");}
        if(!(m_short > 0) || !(m_dh > 0) || (grid_bits != 3 && grid_bits != 4) ||
           (grid_lanes != 1 && grid_lanes != 2 && grid_lanes != 4))
        {
            throw std::invalid_argument("tuning: margins > 0, grid_bits 3 or 4, grid_lanes 1, 2 or 4");
        }
        r.tune.m_short = m_short; r.tune.m_dh = m_dh; r.tune.grid_bits = grid_bits; r.tune.grid_lanes = grid_lanes;
        r.tune_set = true;
    }
    RepTune tuning(int id) {return rep(id).tune;}
    void set_autotune(bool on) {autotune_ = on;}

  private:
    // First evaluation of fresh replicas: lists, first forces; list capacities grown until they fit.
    void settle(const std::vector<int>& fresh)
    {
        std::vector<int> todo = fresh;
        for(int pass = 0; pass < 8 && !todo.empty(); ++pass)
        {
            launch(todo, 0, 1);
            download_states();
            std::vector<int> again;
            for(int id : todo)
            {
                Replica& r = *reps_[id];
                const RepState& st = h_states_[r.slot];
                bool over = false;
                for(size_t g = 0; g < r.globals.size(); ++g)
                {
                    if(st.maxpairs[g] > r.globals[g].cap) {r.globals[g].cap = grown_cap(st.maxpairs[g]); over = true;}
                }
                if(over) {r.topo.reset(r.n); r.ready = false; prepare(id); again.push_back(id);}
            }
            todo.swap(again);
        }
        if(!todo.empty()) {throw BatchDeviceError("replica list capacity did not converge");}
    }

    // ---- the autotuner ------------------------------------------------------------------
    // Replicas of one size bucket (half an octave), class, box kind, precision and interaction
    // make-up share their tuning, found once by a measured coordinate search: a clone of the first
    // such replica runs a short segment alone (one workgroup, the replica's own per-step latency)
    // for each candidate, and the fastest one is kept.  The candidates change no result.
    struct TuneKey
    {
        int cls, tsp, periodic, bucket;
        uint32_t kinds;
        bool operator<(const TuneKey& o) const
        {
            return std::tie(cls, tsp, periodic, bucket, kinds) < std::tie(o.cls, o.tsp, o.periodic, o.bucket, o.kinds);
        }
    };
    TuneKey tune_key(const Replica& r) const
    {
        TuneKey k{};
        k.cls = 0;
        while(r.n > kRepClassMaxN[k.cls]) {++k.cls;}
        k.periodic = r.box.periodic;
        k.bucket = (int)std::floor(2.0 * std::log2((double)r.n));
        for(const auto& G : r.globals) {k.kinds |= 1u << G.pot;}
        for(const auto& L : r.locals) {k.tsp |= (L.kind == kBond3SPN2 || L.kind == kDihedralPGauss || L.kind == kBaseStacking);}
        return k;
    }
    RepTune tuned(const Replica& r)
    {
        const TuneKey key = tune_key(r);
        const auto it = tune_cache_.find(key);
        if(it != tune_cache_.end()) {return it->second;}
        RepTune best = default_tune();
        bool has_short = false, has_dh = false, has_grid = false;
        for(const auto& G : r.globals)
        {
            has_short = has_short || G.pot == kPairLJ || G.pot == kPairEV;
            has_dh = has_dh || G.pot == kPairDH;
            has_grid = has_grid || (G.partition != MJG_PART_NAIVE && G.part.size() > kGridMinNp);
        }
        double best_t = trial(r, best);
        auto consider = [&](RepTune t) {
            const double x = trial(r, t);
            if(x < best_t) {best_t = x; best = t;}
        };
        if(has_short) {for(double m : {0.2, 0.3, 0.5, 0.6}) {RepTune t = best; t.m_short = m; consider(t);}}
        if(has_dh) {for(double m : {0.05, 0.2, 0.3}) {RepTune t = best; t.m_dh = m; consider(t);}}
        if(has_grid)
        {
            const RepTune base = best;
            for(int b : {3, 4})
                for(int l : {1, 2, 4})
                {
                    if(b == base.grid_bits && l == base.grid_lanes) {continue;}
                    RepTune t = base; t.grid_bits = b; t.grid_lanes = l; consider(t);
                }
        }
        failed_now_.clear();
        descs_dirty_ = true;
        tune_cache_[key] = best;
        if(trace_)
        {
            std::fprintf(stderr, "[mjg batch] tuned n=%zu: skin %.2f / DH %.2f, grid %d bits, %d lanes (%.3f ms)\n", r.n,
                         best.m_short, best.m_dh, best.grid_bits, best.grid_lanes, best_t);
        }
        return best;
    }
    // time of kTuneSteps steps of a clone of `src` alone, after kTuneWarm steps (ms; infinite
    // if the clone fails)
    static constexpr int kTuneWarm = 100, kTuneSteps = 300;
    double trial(const Replica& src, const RepTune& t)
    {
        auto c = std::make_unique<Replica>();
        c->box = src.box; c->n = src.n; c->mass = src.mass; c->pos = src.pos; c->vel = src.vel;
        c->T = src.T; c->I = src.I; c->has_T = src.has_T; c->has_I = src.has_I;
        c->locals = src.locals; c->globals = src.globals;
        for(auto& G : c->globals) {G.cap = 0;}
        c->integ = src.integ; c->dt = src.dt; c->seed = src.seed; c->coefs = src.coefs;
        c->topo.reset(c->n);
        c->tune = t;
        c->tune_set = true;
        reps_.push_back(std::move(c));
        const int id = (int)reps_.size() - 1;
        double ms = INFINITY;
        try
        {
            prepare(id);
            settle({id});
            reps_[id]->active = true;
            set_remaining(*reps_[id], kTuneWarm + kTuneSteps);
            run_segment({id}, kTuneWarm);
            const auto t0 = std::chrono::steady_clock::now();
            run_segment({id}, kTuneSteps);
            const auto t1 = std::chrono::steady_clock::now();
            if(!reps_[id]->failed) {ms = std::chrono::duration<double, std::milli>(t1 - t0).count();}
        }
        catch(const BatchDeviceError&) {throw;}
        catch(const std::exception&) {ms = INFINITY;}
        retire(id);
        reps_.pop_back();
        return ms;
    }
  public:

    // Advance every active replica by up to n steps (fewer if it has fewer remaining).  A replica
    // that stops with non-finite state is deactivated; the others go on (failures_of_call()).
    void step(size_t n)
    {
        std::vector<int> act = active_ids();
        if(act.empty()) {return;}
        download_states();
        for(int id : act) {h_states_[reps_[id]->slot].remaining = n;}
        upload_states();
        size_t done = 0;
        while(done < n)
        {
            const int seg = (int)std::min<size_t>(segment_, n - done);
            act = active_ids();
            if(act.empty()) {break;}
            run_segment(act, seg);
            done += seg;
        }
    }

    // replicas that failed (non-finite state) during the calls since the last clear
    const std::vector<int>& failures_of_call() const {return failed_now_;}
    void clear_failures() {failed_now_.clear();}
    bool failed(int id) {return rep(id).failed;}

    void set_steps(int id, uint64_t steps) {set_remaining(rep(id), steps);}

    // position of a replica's noise stream (a restored checkpoint continues it); its lists, and
    // an UnderdampedLangevin replica's stored acceleration, are rebuilt at its next step
    void set_step_count(int id, uint64_t step)
    {
        Replica& r = rep(id);
        if(!r.ready) {throw std::logic_error("replica not initialized");}
        RepState st = state_of(r);
        st.step = step;
        st.force_rebuild = 1;
        hip_check(hipMemcpyAsync(r.desc.st, &st, sizeof(RepState), hipMemcpyHostToDevice, stream_), "upload");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }

    // run `seg` steps of the replicas whose remaining count is positive (serving)
    void advance(int seg)
    {
        std::vector<int> act = active_ids();
        if(!act.empty()) {run_segment(act, seg);}
    }

    void retire(int id)
    {
        Replica& r = rep(id);
        hip_check(hipStreamSynchronize(stream_), "sync");
        if(r.arena) {hip_check(hipFree(r.arena), "hipFree");}
        free_slots_.push_back(r.slot);
        reps_[id].reset();
        descs_dirty_ = true;
    }

    RepState state(int id)
    {
        Replica& r = rep(id);
        if(!r.ready) {throw std::logic_error("replica not initialized");}
        return state_of(r);
    }

    void get_state(int id, double* pos, double* vel, double* force)
    {
        Replica& r = rep(id);
        if(!r.ready) {throw std::logic_error("replica not initialized");}
        std::vector<double4> h(r.n);
        double4* src[3] = {r.desc.pos, r.desc.vel, r.desc.force};
        double* dst[3] = {pos, vel, force};
        for(int k = 0; k < 3; ++k)
        {
            if(!dst[k]) {continue;}
            hip_check(hipMemcpyAsync(h.data(), src[k], r.n * sizeof(double4), hipMemcpyDeviceToHost, stream_), "download");
            hip_check(hipStreamSynchronize(stream_), "sync");
            for(size_t i = 0; i < r.n; ++i) {dst[k][3 * i] = h[i].x; dst[k][3 * i + 1] = h[i].y; dst[k][3 * i + 2] = h[i].z;}
        }
    }

    // forces and per-interaction energies at the current positions (no step)
    void evaluate(int id, double* energies, double* total, double* force)
    {
        Replica& r = rep(id);
        if(!r.ready) {throw std::logic_error("replica not initialized");}
        launch_eval({id});
        RepState s = state(id);
        double tot = 0;
        const int ne = (int)r.locals.size() + (int)r.globals.size();
        for(int k = 0; k < ne; ++k) {if(energies) {energies[k] = s.energy[k];} tot += s.energy[k];}
        if(total) {*total = tot;}
        if(force) {get_state(id, nullptr, nullptr, force);}
    }

    // EnergyObserver::output for every active replica at once: one evaluation launch (forces
    // recomputed from the same lists, so the trajectory is unchanged), then the states
    void observe()
    {
        std::vector<int> act = active_ids();
        if(act.empty()) {return;}
        launch_eval(act);
        for(int id : act) {reps_[id]->observed = true;}
    }

    // One evaluation launch (forces and energies, no step); a replica whose list had to be rebuilt
    // and did not fit (a forced rebuild after a restore) is evaluated again with a larger list.
    void launch_eval(const std::vector<int>& ids)
    {
        launch(ids, 0, 2);
        download_states();
        std::vector<int> again;
        for(int id : ids) {if(h_states_[reps_[id]->slot].overflow) {again.push_back(id);}}
        for(int pass = 0; pass < 6 && !again.empty(); ++pass)
        {
            for(int id : again) {grow_if_overflow(*reps_[id], false);}
            launch(again, 0, 2);
            download_states();
            std::vector<int> still;
            for(int id : again) {if(h_states_[reps_[id]->slot].overflow) {still.push_back(id);}}
            again.swap(still);
        }
        if(!again.empty()) {throw BatchDeviceError("replica list capacity did not converge");}
    }

    // the last observation of one replica: per-interaction energies and the kinetic energy
    void energies(int id, double* per, double* kinetic)
    {
        Replica& r = rep(id);
        if(!r.observed) {throw std::logic_error("replica has no energy observation yet");}
        const RepState& s = h_states_[r.slot];
        const int ne = (int)r.locals.size() + (int)r.globals.size();
        for(int k = 0; k < ne; ++k) {if(per) {per[k] = s.energy[k];}}
        if(kinetic) {*kinetic = s.kinetic;}
    }

    size_t footprint(int id) {return rep(id).footprint;}
    size_t num_interactions(int id) {Replica& r = rep(id); return r.locals.size() + r.globals.size();}
    int num_globals(int id) {return (int)rep(id).globals.size();}

    void counters(uint64_t* steps, uint64_t* rebuilds)
    {
        uint64_t s = 0, b = 0;
        download_states();
        for(size_t id = 0; id < reps_.size(); ++id)
        {
            if(!reps_[id] || !reps_[id]->ready) {continue;}
            const RepState& st = h_states_[reps_[id]->slot];
            s += st.step * reps_[id]->n;
            b += st.rebuilds;
        }
        if(steps) {*steps = s;}
        if(rebuilds) {*rebuilds = b;}
    }

    // time `reps` launches of an nsteps segment over every active replica (HIP events)
    double profile(int nsteps, int reps, uint64_t* particle_steps)
    {
        std::vector<int> act = active_ids();
        if(act.empty() || reps < 1) {throw std::logic_error("nothing to profile");}
        uint64_t ps = 0;
        download_states();
        for(int id : act) {h_states_[reps_[id]->slot].remaining = (uint64_t)nsteps * reps; ps += reps_[id]->n;}
        upload_states();
        upload_descs(act);
        hipEvent_t a, b;
        hip_check(hipEventCreate(&a), "event");
        hip_check(hipEventCreate(&b), "event");
        hip_check(hipEventRecord(a, stream_), "record");
        for(int k = 0; k < reps; ++k) {launch_descs(nsteps, 0);}
        hip_check(hipEventRecord(b, stream_), "record");
        hip_check(hipEventSynchronize(b), "event sync");
        hip_check(hipGetLastError(), "profile launch");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, a, b), "elapsed");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        download_states();
        note_failures(act);
        if(particle_steps) {*particle_steps = ps * (uint64_t)nsteps;}
        return (double)ms / reps;
    }

  private:

    std::vector<int> active_ids()
    {
        std::vector<int> v;
        for(size_t id = 0; id < reps_.size(); ++id) {if(reps_[id] && reps_[id]->active) {v.push_back((int)id);}}
        return v;
    }

    void set_remaining(Replica& r, uint64_t n)
    {
        RepState st = state_of(r);
        st.remaining = n;
        hip_check(hipMemcpyAsync(r.desc.st, &st, sizeof(RepState), hipMemcpyHostToDevice, stream_), "upload");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }

    // state array shared by every replica (one slot each), mirrored on the host
    int take_slot()
    {
        if(!free_slots_.empty()) {const int k = free_slots_.back(); free_slots_.pop_back(); return k;}
        if(n_slots_ == states_cap_)
        {
            const size_t cap = std::max<size_t>(64, 2 * states_cap_);
            RepState* d = nullptr;
            hip_check(hipMalloc(&d, cap * sizeof(RepState)), "hipMalloc");
            if(d_states_)
            {
                hip_check(hipMemcpyAsync(d, d_states_, n_slots_ * sizeof(RepState), hipMemcpyDeviceToDevice, stream_), "copy");
                hip_check(hipStreamSynchronize(stream_), "sync");
                hip_check(hipFree(d_states_), "hipFree");
            }
            d_states_ = d;
            states_cap_ = cap;
            // every replica's state pointer moves with the array
            for(auto& r : reps_) {if(r && r->slot >= 0) {r->desc.st = d_states_ + r->slot;}}
            descs_dirty_ = true;
        }
        h_states_.resize(n_slots_ + 1);
        return (int)n_slots_++;
    }
    void download_states()
    {
        if(n_slots_ == 0) {return;}
        h_states_.resize(n_slots_);
        hip_check(hipMemcpyAsync(h_states_.data(), d_states_, n_slots_ * sizeof(RepState), hipMemcpyDeviceToHost, stream_),
                  "download");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }
    void upload_states()
    {
        if(n_slots_ == 0) {return;}
        hip_check(hipMemcpyAsync(d_states_, h_states_.data(), n_slots_ * sizeof(RepState), hipMemcpyHostToDevice, stream_),
                  "upload");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }

    // Lay out the replica's device arena and upload everything (set-up only).
    void prepare(int id)
    {
        Replica& r = *reps_[id];
        if(r.integ == 0) {throw std::invalid_argument("replica " + std::to_string(id) + " has no integrator");}
        if(!r.has_T) {throw std::invalid_argument("Langevin integrator requires the `temperature` attribute");}
        // topology, molecules
        for(auto& L : r.locals)
        {
            if(L.kind != kBaseStacking) {r.topo.add_terms(L.idx, L.arity, L.topo); continue;}
            if(L.topo < 0) {continue;}
            // ThreeSPN2BaseStackingInteraction::write_topology (:125-179): every bead of the 5'
            // nucleotide to every bead of the 3' one, phosphates when present
            for(size_t t = 0; t + 6 <= L.idx.size(); t += 6)
            {
                const uint32_t S5 = L.idx[t], B5 = L.idx[t + 1], B3 = L.idx[t + 2];
                const uint32_t P5 = L.idx[t + 3], S3 = L.idx[t + 4], P3 = L.idx[t + 5];
                r.topo.add_connection(S5, S3, L.topo);
                r.topo.add_connection(S5, B3, L.topo);
                r.topo.add_connection(B5, S3, L.topo);
                r.topo.add_connection(B5, B3, L.topo);
                if(P5 != kNil) {r.topo.add_connection(P5, S3, L.topo); r.topo.add_connection(P5, B3, L.topo);}
                if(P3 != kNil) {r.topo.add_connection(S5, P3, L.topo); r.topo.add_connection(B5, P3, L.topo);}
                if(P5 != kNil && P3 != kNil) {r.topo.add_connection(P5, P3, L.topo);}
            }
        }
        r.topo.construct_molecules();
        const size_t n = r.n;

        // local terms, grouped by interaction, as structure of arrays
        size_t n_terms = 0;
        for(const RepLocal& L : r.locals) {n_terms += L.idx.size() / L.arity;}
        std::vector<unsigned long long> t_idx(n_terms, 0ull);
        std::vector<uint32_t> t_kind(n_terms, 0u);
        std::vector<double> t_par(3 * n_terms, 0.0);
        {
            size_t t = 0;
            for(size_t l = 0; l < r.locals.size(); ++l)
            {
                const RepLocal& L = r.locals[l];
                const size_t nt = L.idx.size() / L.arity;
                for(size_t q = 0; q < nt; ++q, ++t)
                {
                    // a base-stacking term keeps {Si, Bi, Bj}: its phosphates only enter the topology
                    const int packed = (L.kind == kBaseStacking) ? 3 : L.arity;
                    unsigned long long ix = 0;
                    for(int a = 0; a < packed; ++a) {ix |= (unsigned long long)L.idx[q * L.arity + a] << (16 * a);}
                    t_idx[t] = ix;
                    t_kind[t] = (uint32_t)L.kind | ((uint32_t)l << 8);
                    for(int c = 0; c < L.npar; ++c) {t_par[c * n_terms + t] = L.par[q * L.npar + c];}
                }
            }
        }

        // per global: tables
        struct GHost
        {
            std::vector<uint16_t> part;
            std::vector<uint32_t> far_start, far;
            std::vector<unsigned long long> ex_near;
            std::vector<double> ppar;
        };
        std::vector<GHost> gh(r.globals.size());
        for(size_t g = 0; g < r.globals.size(); ++g)
        {
            RepGlobalIn& G = r.globals[g];
            GHost& h = gh[g];
            h.part.assign(G.part.begin(), G.part.end());
            h.ppar.assign(n * G.npar, 0.0);
            for(size_t k = 0; k < G.part.size(); ++k)
                for(int c = 0; c < G.npar; ++c) {h.ppar[(size_t)G.part[k] * G.npar + c] = G.par[k * G.npar + c];}
            {
                // exclusions of each particle against higher indices: a 64-bit mask for the
                // next 64 indices, a CSR list for the rest
                std::vector<uint32_t> ex_start, ex;
                r.topo.exclusions(G.within, ex_start, ex);
                h.ex_near.assign(n, 0ull);
                h.far_start.assign(n + 1, 0u);
                for(size_t p = 0; p < n; ++p)
                {
                    h.far_start[p] = (uint32_t)h.far.size();
                    for(uint32_t k = ex_start[p]; k < ex_start[p + 1]; ++k)
                    {
                        const uint32_t q = ex[k];
                        if(q <= p) {continue;}
                        if(q - p <= 64) {h.ex_near[p] |= 1ull << (q - p - 1);}
                        else {h.far.push_back(q);}
                    }
                }
                h.far_start[n] = (uint32_t)h.far.size();
            }
            if(G.cap == 0)
            {
                // Every participant pair: at most 2,096,128 entries (8 MiB) for the largest replica,
                // which HBM holds for thousands of replicas, so a list never overflows and no
                // segment runs twice or stops for a host-side re-layout (a chain's Debye-Hueckel
                // list grows several-fold as it compacts).  Above kRepAllPairsMax entries the first
                // guess is a tenth of the pairs, grown on overflow.  MJG_BATCH_INITIAL_CAP
                // overrides the guess (tests use it to exercise the growth path).
                const uint64_t np = G.part.size();
                const uint64_t all = np * (np - 1) / 2 + 1;
                G.cap = (uint32_t)(all <= kRepAllPairsMax ? all : std::max<uint64_t>(1024, np * (np - 1) / 10));
                if(const char* c = std::getenv("MJG_BATCH_INITIAL_CAP")) {G.cap = (uint32_t)std::max(1, std::atoi(c));}
            }
        }

        // arena layout
        size_t off = 0;
        auto take = [&](size_t bytes) {const size_t o = off; off += align256(bytes); return o;};
        const size_t o_pos = take(n * sizeof(double4)), o_vel = take(n * sizeof(double4)), o_force = take(n * sizeof(double4));
        const size_t o_bpos = take(n * sizeof(double4)), o_bvel = take(n * sizeof(double4)), o_bforce = take(n * sizeof(double4));
        const size_t o_acc = take(r.integ == MJG_INTEG_UNDERDAMPED ? n * sizeof(double4) : 0);
        const size_t o_rm = take(n * 8), o_c1 = take(n * 8), o_c2 = take(n * 8), o_c3 = take(n * 8);
        const size_t o_mass = take(n * 8);
        const size_t o_mol = take(n * 4), o_tidx = take(n_terms * 8), o_tkind = take(n_terms * 4);
        const size_t o_tpar = take(n_terms * 24);
        const size_t o_bst = take(sizeof(RepState));
        if(r.slot < 0) {r.slot = take_slot();}
        std::vector<size_t> o_part(gh.size()), o_exnear(gh.size()), o_ppar(gh.size()), o_exs(gh.size()), o_ex(gh.size()),
            o_pairs(gh.size());
        for(size_t g = 0; g < gh.size(); ++g)
        {
            o_part[g] = take(gh[g].part.size() * 2);
            o_exnear[g] = take(n * 8);
            o_ppar[g] = take(gh[g].ppar.size() * 8);
            o_exs[g] = take((n + 1) * 4);
            o_ex[g] = take(std::max<size_t>(1, gh[g].far.size()) * 4);
            o_pairs[g] = take((size_t)r.globals[g].cap * 4);
        }
        if(r.arena) {hip_check(hipFree(r.arena), "hipFree");}
        r.footprint = off;
        if(hipMalloc(&r.arena, off) != hipSuccess) {(void)hipGetLastError(); r.arena = nullptr; throw std::bad_alloc();}
        char* A = static_cast<char*>(r.arena);

        // state and coefficients
        const double kBT = kB_kcal() * r.T;
        std::vector<double4> p(n), v(n), z(n, make_double4(0, 0, 0, 0));
        std::vector<double> rm(n), c1(n), c2(n), c3(n);
        for(size_t i = 0; i < n; ++i)
        {
            double x = r.pos[3 * i], y = r.pos[3 * i + 1], zz = r.pos[3 * i + 2];
            adjust_position_host(r.box, x, y, zz);
            p[i] = make_double4(x, y, zz, 0);
            v[i] = make_double4(r.vel[3 * i], r.vel[3 * i + 1], r.vel[3 * i + 2], 0);
            rm[i] = 1.0 / r.mass[i];
            if(r.integ == MJG_INTEG_UNDERDAMPED)
            {
                // UnderdampedLangevinIntegrator::update (:320-350): gamma, sqrt(2 kB T / dt) sqrt(gamma / m)
                c1[i] = r.coefs[i];
                c2[i] = std::sqrt(2 * kB_kcal() * r.T / r.dt) * std::sqrt(r.coefs[i] * rm[i]);
                c3[i] = 0;
            }
            else if(r.integ == MJG_INTEG_BAOAB)
            {
                const double gamma_dt = -1 * r.coefs[i] * r.dt;
                c1[i] = std::exp(gamma_dt);
                c2[i] = std::sqrt(kBT * (1 - std::exp(2 * gamma_dt)) * rm[i]);
                c3[i] = 0;
            }
            else
            {
                const double alpha = r.coefs[i], m = r.mass[i];
                c1[i] = 1.0 / (1.0 + alpha * r.dt * 0.5 / m);
                c2[i] = std::sqrt(2 * alpha * kBT * r.dt);
                c3[i] = 1.0 - alpha * c1[i] * r.dt / m;
            }
        }
        RepState st{};
        for(size_t g = 0; g < r.globals.size(); ++g) {st.margin[g] = -1.0;}
        st.force_rebuild = 1;
        auto up = [&](size_t o, const void* src, size_t bytes) {
            if(bytes) {hip_check(hipMemcpyAsync(A + o, src, bytes, hipMemcpyHostToDevice, stream_), "upload");}
        };
        up(o_pos, p.data(), n * sizeof(double4));
        up(o_vel, v.data(), n * sizeof(double4));
        up(o_force, z.data(), n * sizeof(double4));
        up(o_mass, r.mass.data(), n * 8);
        up(o_rm, rm.data(), n * 8); up(o_c1, c1.data(), n * 8); up(o_c2, c2.data(), n * 8); up(o_c3, c3.data(), n * 8);
        up(o_mol, r.topo.mol.data(), n * 4);
        up(o_tidx, t_idx.data(), n_terms * 8);
        up(o_tkind, t_kind.data(), n_terms * 4);
        up(o_tpar, t_par.data(), n_terms * 24);
        hip_check(hipMemcpyAsync(d_states_ + r.slot, &st, sizeof(RepState), hipMemcpyHostToDevice, stream_), "upload");

        RepDesc& D = r.desc;
        D = RepDesc{};
        D.n = (uint32_t)n;
        D.n_terms = (uint32_t)n_terms;
        D.n_local = (int)r.locals.size();
        D.ng = (int)r.globals.size();
        D.integ = r.integ;
        D.dt = r.dt;
        D.halfdt = r.dt / 2;
        D.seed = r.seed;
        D.box = r.box;
        D.pos = reinterpret_cast<double4*>(A + o_pos);
        D.vel = reinterpret_cast<double4*>(A + o_vel);
        D.force = reinterpret_cast<double4*>(A + o_force);
        D.bak_pos = reinterpret_cast<double4*>(A + o_bpos);
        D.bak_vel = reinterpret_cast<double4*>(A + o_bvel);
        D.bak_force = reinterpret_cast<double4*>(A + o_bforce);
        D.accel = reinterpret_cast<double4*>(A + o_acc);
        D.rmass = reinterpret_cast<const double*>(A + o_rm);
        D.mass = reinterpret_cast<const double*>(A + o_mass);
        D.c1 = reinterpret_cast<const double*>(A + o_c1);
        D.c2 = reinterpret_cast<const double*>(A + o_c2);
        D.c3 = reinterpret_cast<const double*>(A + o_c3);
        D.mol = reinterpret_cast<const uint32_t*>(A + o_mol);
        D.t_idx = reinterpret_cast<const unsigned long long*>(A + o_tidx);
        D.t_kind = reinterpret_cast<const uint32_t*>(A + o_tkind);
        D.t_par = reinterpret_cast<const double*>(A + o_tpar);
        D.st = d_states_ + r.slot;
        D.bak_st = reinterpret_cast<RepState*>(A + o_bst);
        for(size_t g = 0; g < r.globals.size(); ++g)
        {
            RepGlobalIn& G = r.globals[g];
            GHost& h = gh[g];
            up(o_part[g], h.part.data(), h.part.size() * 2);
            up(o_exnear[g], h.ex_near.data(), n * 8);
            up(o_ppar[g], h.ppar.data(), h.ppar.size() * 8);
            up(o_exs[g], h.far_start.data(), (n + 1) * 4);
            up(o_ex[g], h.far.data(), h.far.size() * 4);
            RepGlobal& RG = D.g[g];
            RG = RepGlobal{};
            RG.kind = G.pot;
            RG.base_rule = (G.pot == kPair3SPN2EV) ? 1 : (G.pot == kBasePair || G.pot == kCrossStack) ? 2 : 0;
            RG.pairwise = (G.partition != MJG_PART_CELLLIST);
            RG.use_distance = (G.partition != MJG_PART_NAIVE);
            RG.never_rebuild = (G.partition == MJG_PART_NAIVE);
            RG.ignore_mol = G.ignore_mol;
            RG.np = (uint32_t)G.part.size();
            RG.last_slot = RG.np - 1;
            RG.sorted = std::is_sorted(G.part.begin(), G.part.end()) ? 1 : 0;
            RG.npar = G.npar;
            RG.par_col = D.par_cols;
            RG.part_off = D.part_total;
            D.par_cols += (uint32_t)G.npar;
            D.part_total += (RG.np + 7u) & ~7u;
            if(RG.use_distance && RG.np > kGridMinNp) {D.grid_np = std::max(D.grid_np, RG.np);}
            RG.cap = G.cap;
            RG.part = reinterpret_cast<const uint16_t*>(A + o_part[g]);
            RG.ex_near = reinterpret_cast<const unsigned long long*>(A + o_exnear[g]);
            RG.ppar = reinterpret_cast<const double*>(A + o_ppar[g]);
            RG.far_start = reinterpret_cast<const uint32_t*>(A + o_exs[g]);
            RG.far = reinterpret_cast<const uint32_t*>(A + o_ex[g]);
            RG.pairs = reinterpret_cast<uint32_t*>(A + o_pairs[g]);
            double max_cutoff = 0;
            if(G.pot == kPairLJ)
            {
                RG.cutoff_ratio2 = G.cutoff_ratio * G.cutoff_ratio;
                RG.coef_at_cutoff = std::pow(1.0 / G.cutoff_ratio, 12) - std::pow(1.0 / G.cutoff_ratio, 6);
                double ms = 0;
                for(size_t i = 0; i < n; ++i) {ms = std::max(ms, h.ppar[2 * i]);}
                max_cutoff = ms * G.cutoff_ratio;
            }
            else if(G.pot == kPairEV)
            {
                RG.cutoff_ratio2 = G.cutoff_ratio * G.cutoff_ratio;
                RG.coef_at_cutoff = std::pow(1.0 / G.cutoff_ratio, 12);
                RG.eps = G.eps;
                double mr = 0;
                for(size_t i = 0; i < n; ++i) {mr = std::max(mr, h.ppar[i]);}
                max_cutoff = mr * 2 * G.cutoff_ratio;
            }
            else if(G.pot == kPair3SPN2EV)
            {
                // ThreeSPN2ExcludedVolumePotential.hpp:120-134 (1 kJ/mol), list cutoff: the largest
                // sigma of the table (:176-178)
                RG.eps = 1.0 * (1.0 / 4.184);
                max_cutoff = 7.1;
            }
            else if(G.pot == kBasePair || G.pot == kCrossStack)
            {
                max_cutoff = 18.0;   // ThreeSPN2BasePairGlobalPotentialParameter::cutoff
            }
            else
            {
                if(!r.has_I) {throw std::invalid_argument("DebyeHuckel requires `temperature` and `ionic_strength`");}
                const DHParams dh = debye_huckel_params(r.T, r.I, G.cutoff_ratio);
                RG.debye = dh.debye; RG.inv_debye = dh.inv_debye; RG.abs_cutoff = dh.abs_cutoff;
                RG.inv4pieps = dh.inv4pieps; RG.coef_at_cutoff = dh.coef_at_cutoff;
                max_cutoff = dh.debye * G.cutoff_ratio;
            }
            const double m = list_margin(G.pot, G.margin, r.tune);
            const double r_list = max_cutoff * (1. + m);
            RG.r_list2 = r_list * r_list;
            RG.abs_margin = max_cutoff * m;
        }
        // the tuned grid (16 or 8 cells per dimension) where it fits in LDS
        D.grid_bits = (uint32_t)r.tune.grid_bits;
        D.grid_lanes = (uint32_t)r.tune.grid_lanes;
        r.lds = rep_lds_bytes(D.n, D.par_cols, D.part_total, real_size(), D.grid_np, D.grid_bits);
        if(r.lds > kRepLdsLimit)
        {
            D.grid_bits = 3;
            r.lds = rep_lds_bytes(D.n, D.par_cols, D.part_total, real_size(), D.grid_np, D.grid_bits);
        }
        if(r.lds > kRepLdsLimit)
        {
            throw std::invalid_argument("replica needs " + std::to_string(r.lds) + " B of LDS (limit " +
                                        std::to_string(kRepLdsLimit) + ")");
        }
        r.cls = 0;
        while(r.n > kRepClassMaxN[r.cls]) {++r.cls;}
        r.tsp = false;
        for(const auto& L : r.locals) {r.tsp = r.tsp || L.kind == kBond3SPN2 || L.kind == kDihedralPGauss || L.kind == kBaseStacking;}
        for(const auto& G : r.globals) {r.tsp = r.tsp || G.pot == kPair3SPN2EV || G.pot == kBasePair || G.pot == kCrossStack;}
        hip_check(hipStreamSynchronize(stream_), "upload sync");
        r.ready = true;
        descs_dirty_ = true;
    }

    static void adjust_position_host(const Box& b, double& x, double& y, double& z)
    {
        if(!b.periodic) {return;}
        if(x < b.lower[0]) {x += b.width[0];} else if(x >= b.upper[0]) {x -= b.width[0];}
        if(y < b.lower[1]) {y += b.width[1];} else if(y >= b.upper[1]) {y -= b.width[1];}
        if(z < b.lower[2]) {z += b.width[2];} else if(z >= b.upper[2]) {z -= b.width[2];}
    }

    // Upload the descriptors of `ids` grouped by size class, largest replicas first within a
    // class (they have the longest critical path).
    void upload_descs(const std::vector<int>& ids)
    {
        if(!descs_dirty_ && ids == last_ids_) {return;}
        last_ids_ = ids;
        descs_dirty_ = false;
        std::vector<int> order(ids);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const Replica& x = *reps_[a];
            const Replica& y = *reps_[b];
            if(x.cls != y.cls) {return x.cls > y.cls;}
            if(x.tsp != y.tsp) {return y.tsp;}
            return x.n > y.n;
        });
        std::vector<RepDesc> h;
        for(int c = 0; c < kRepClasses; ++c) {grp_count_[c] = 0; grp_tsp_[c] = 0; grp_lds_[c] = 0;}
        for(int id : order)
        {
            const Replica& r = *reps_[id];
            h.push_back(r.desc);
            grp_count_[r.cls] += 1;
            grp_tsp_[r.cls] += r.tsp ? 1 : 0;
            grp_lds_[r.cls] = std::max(grp_lds_[r.cls], r.lds);
        }
        // groups are stored class 2, 1, 0
        size_t off = 0;
        for(int c = kRepClasses - 1; c >= 0; --c) {grp_first_[c] = off; off += grp_count_[c];}
        if(h.size() > desc_cap_)
        {
            if(d_descs_) {hip_check(hipFree(d_descs_), "hipFree");}
            desc_cap_ = std::max<size_t>(h.size(), 2 * desc_cap_);
            hip_check(hipMalloc(&d_descs_, desc_cap_ * sizeof(RepDesc)), "hipMalloc");
        }
        hip_check(hipMemcpyAsync(d_descs_, h.data(), h.size() * sizeof(RepDesc), hipMemcpyHostToDevice, stream_), "upload");
        hip_check(hipStreamSynchronize(stream_), "upload sync");
    }

    size_t real_size() const {return precision_ == MJG_FP32_FIXED ? sizeof(float) : sizeof(double);}

    // list capacity after an overflow at `seen` entries
    static uint32_t grown_cap(uint32_t seen) {return seen + seen / 2 + 1024;}
    static constexpr uint64_t kRepAllPairsMax = uint64_t(1) << 21;

    template<typename Real, bool ENERGY, int BS, int PPT>
    void set_lds_limit()
    {
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_replica_segment<Real, ENERGY, BS, PPT, false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRepLdsLimit),
                  "LDS attribute");
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_replica_segment<Real, ENERGY, BS, PPT, true>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)kRepLdsLimit),
                  "LDS attribute");
    }
    template<typename Real, int C> using Shape = RepShape<Real, C>;
    template<typename Real, int C>
    void set_lds_limits_class()
    {
        set_lds_limit<Real, false, Shape<Real, C>::bs, Shape<Real, C>::ppt>();
        set_lds_limit<Real, true, Shape<Real, C>::bs, Shape<Real, C>::ppt>();
    }
    template<typename Real>
    void set_lds_limits()
    {
        set_lds_limits_class<Real, 0>();
        set_lds_limits_class<Real, 1>();
        set_lds_limits_class<Real, 2>();
    }

    template<typename Real, bool ENERGY, int C>
    void launch_shape(int nsteps, int eval_only)
    {
        constexpr int BS = Shape<Real, C>::bs, PPT = Shape<Real, C>::ppt;
        static_assert(BS * PPT == (int)kRepClassMaxN[C], "class shape must cover its size");
        // the class's replicas without 3SPN2 terms, then (same stream) those with them
        const size_t plain = grp_count_[C] - grp_tsp_[C];
        if(plain > 0)
        {
            hipLaunchKernelGGL((k_replica_segment<Real, ENERGY, BS, PPT, false>), dim3((uint32_t)plain), dim3(BS),
                               grp_lds_[C], cls_stream_[C], d_descs_ + grp_first_[C], nsteps, eval_only);
        }
        if(grp_tsp_[C] > 0)
        {
            hipLaunchKernelGGL((k_replica_segment<Real, ENERGY, BS, PPT, true>), dim3((uint32_t)grp_tsp_[C]), dim3(BS),
                               grp_lds_[C], cls_stream_[C], d_descs_ + grp_first_[C] + plain, nsteps, eval_only);
        }
    }
    template<typename Real, bool ENERGY>
    void launch_class(int c, int nsteps, int eval_only)
    {
        if(c == 0) {launch_shape<Real, ENERGY, 0>(nsteps, eval_only);}
        else if(c == 1) {launch_shape<Real, ENERGY, 1>(nsteps, eval_only);}
        else {launch_shape<Real, ENERGY, 2>(nsteps, eval_only);}
    }

    // mode 0: step segment, 1: evaluate after a forced rebuild (set-up), 2: evaluate with energies.
    // Each size class runs on its own stream: the classes fork from stream_ (class 0's stream),
    // all three are launched, and only then does stream_ wait for the other two, so the classes
    // run concurrently and a segment takes as long as its longest class.
    void launch_descs(int nsteps, int mode)
    {
        const bool fp32 = (precision_ == MJG_FP32_FIXED);
        hip_check(hipEventRecord(ev_fork_, stream_), "record");
        for(int c = kRepClasses - 1; c >= 0; --c)
        {
            if(grp_count_[c] == 0) {continue;}
            if(c != 0) {hip_check(hipStreamWaitEvent(cls_stream_[c], ev_fork_, 0), "wait");}
            const int ns = (mode == 0) ? nsteps : 0, ev = (mode == 0) ? 0 : 1;
            if(mode == 2) {if(fp32) {launch_class<float, true>(c, ns, ev);} else {launch_class<double, true>(c, ns, ev);}}
            else {if(fp32) {launch_class<float, false>(c, ns, ev);} else {launch_class<double, false>(c, ns, ev);}}
            if(c != 0) {hip_check(hipEventRecord(ev_join_[c], cls_stream_[c]), "record");}
        }
        for(int c = 1; c < kRepClasses; ++c)
        {
            if(grp_count_[c] != 0) {hip_check(hipStreamWaitEvent(stream_, ev_join_[c], 0), "wait");}
        }
    }

    void launch(const std::vector<int>& ids, int nsteps, int mode)
    {
        upload_descs(ids);
        launch_descs(nsteps, mode);
        hip_check(hipGetLastError(), "replica launch");
        hip_check(hipStreamSynchronize(stream_), "replica sync");
    }

    // after a launch: grow the list capacities that overflowed and reset the replica to run again,
    // from its segment-start copy after a step launch (from_backup) or from its current state after
    // an evaluation launch, which moved nothing (true if the replica must run again)
    bool grow_if_overflow(Replica& r, bool from_backup = true)
    {
        const RepState s = state_of(r);
        bool over = false;
        for(size_t g = 0; g < r.globals.size(); ++g)
        {
            if(s.maxpairs[g] > r.globals[g].cap)
            {
                r.globals[g].cap = grown_cap(s.maxpairs[g]);
                over = true;
            }
        }
        if(!over) {return false;}
        relayout(r, from_backup);
        return true;
    }

    // lay the replica's arena out again with its (new) list capacities and restore its state: the
    // segment-start copy (from_backup) or the current one; the lists are rebuilt at its next step
    void relayout(Replica& r, bool from_backup)
    {
        std::vector<double> pos(3 * r.n), vel(3 * r.n), force(3 * r.n);
        download_xyz(r, from_backup ? r.desc.bak_pos : r.desc.pos, pos);
        download_xyz(r, from_backup ? r.desc.bak_vel : r.desc.vel, vel);
        download_xyz(r, from_backup ? r.desc.bak_force : r.desc.force, force);
        RepState bak{};
        hip_check(hipMemcpyAsync(&bak, from_backup ? r.desc.bak_st : r.desc.st, sizeof(RepState),
                                 hipMemcpyDeviceToHost, stream_), "download");
        hip_check(hipStreamSynchronize(stream_), "sync");
        r.pos = pos;
        r.vel = vel;
        r.topo.reset(r.n);
        r.ready = false;
        int id = -1;
        for(size_t k = 0; k < reps_.size(); ++k) {if(reps_[k].get() == &r) {id = (int)k;}}
        prepare(id);
        // positions are already wrapped; forces and counters are the restored ones
        std::vector<double4> f(r.n);
        for(size_t i = 0; i < r.n; ++i) {f[i] = make_double4(force[3 * i], force[3 * i + 1], force[3 * i + 2], 0.0);}
        hip_check(hipMemcpyAsync(r.desc.force, f.data(), r.n * sizeof(double4), hipMemcpyHostToDevice, stream_), "upload");
        bak.force_rebuild = 1;
        bak.overflow = 0;
        for(size_t g = 0; g < r.globals.size(); ++g) {bak.maxpairs[g] = 0;}
        hip_check(hipMemcpyAsync(r.desc.st, &bak, sizeof(RepState), hipMemcpyHostToDevice, stream_), "upload");
        hip_check(hipStreamSynchronize(stream_), "sync");
    }

    void download_xyz(Replica& r, const double4* d, std::vector<double>& out)
    {
        std::vector<double4> h(r.n);
        hip_check(hipMemcpyAsync(h.data(), d, r.n * sizeof(double4), hipMemcpyDeviceToHost, stream_), "download");
        hip_check(hipStreamSynchronize(stream_), "sync");
        for(size_t i = 0; i < r.n; ++i) {out[3 * i] = h[i].x; out[3 * i + 1] = h[i].y; out[3 * i + 2] = h[i].z;}
    }

    RepState state_of(Replica& r)
    {
        RepState s{};
        hip_check(hipMemcpyAsync(&s, r.desc.st, sizeof(RepState), hipMemcpyDeviceToHost, stream_), "download");
        hip_check(hipStreamSynchronize(stream_), "sync");
        return s;
    }

    // A replica whose state stopped being finite is taken out of the launches (its workgroup
    // already ended its segment there); it stays queryable until it is retired.
    void note_failures(const std::vector<int>& ids)
    {
        for(int id : ids)
        {
            Replica& r = *reps_[id];
            // a replica whose list overflowed runs its segment again (its state is restored)
            if(r.failed || h_states_[r.slot].overflow || !h_states_[r.slot].nonfinite) {continue;}
            r.failed = true;
            r.active = false;
            failed_now_.push_back(id);
            descs_dirty_ = true;
        }
    }

    void run_segment(const std::vector<int>& act, int seg)
    {
        launch(act, seg, 0);
        download_states();
        note_failures(act);
        // a replica whose list overflowed reruns the segment from its segment-start copy
        std::vector<int> again;
        for(int id : act)
        {
            const RepState& s = h_states_[reps_[id]->slot];
            if(s.overflow && !reps_[id]->failed) {again.push_back(id);}
        }
        for(int pass = 0; pass < 6 && !again.empty(); ++pass)
        {
            if(trace_) {std::fprintf(stderr, "[mjg batch] %zu replica(s) overflowed a list; rerunning the segment\n", again.size());}
            for(int id : again) {grow_if_overflow(*reps_[id]);}
            launch(again, seg, 0);
            download_states();
            note_failures(again);
            std::vector<int> still;
            for(int id : again) {if(h_states_[reps_[id]->slot].overflow && !reps_[id]->failed) {still.push_back(id);}}
            again.swap(still);
        }
        if(!again.empty()) {throw BatchDeviceError("replica list capacity did not converge");}
        // A list that came within an eighth of its capacity (chains compact over a run, and their
        // Debye-Hueckel lists grow) is enlarged now, from the replica's current state, so that it
        // does not overflow in a later segment, which would run that segment a second time.
        for(int id : act)
        {
            Replica& r = *reps_[id];
            if(r.failed) {continue;}
            const RepState& st = h_states_[r.slot];
            bool tight = false;
            for(size_t g = 0; g < r.globals.size(); ++g) {tight = tight || list_tight(r, g, st);}
            if(tight) {grow_ahead(r);}
        }
    }

    // A list within an eighth of its capacity, unless it holds every participant pair (which no
    // build can exceed)
    static bool list_tight(const Replica& r, size_t g, const RepState& st)
    {
        const uint64_t np = r.globals[g].part.size(), cap = r.globals[g].cap;
        return cap < np * (np - 1) / 2 + 1 && st.maxpairs[g] > cap - cap / 8;
    }

    // enlarge the replica's nearly full lists as an overflow at their capacity would, keeping its
    // current state (its next step rebuilds the lists)
    void grow_ahead(Replica& r)
    {
        const RepState st = state_of(r);
        for(size_t g = 0; g < r.globals.size(); ++g)
        {
            const uint32_t cap = r.globals[g].cap;
            if(list_tight(r, g, st))
            {
                if(trace_)
                {
                    std::fprintf(stderr, "[mjg batch] replica n=%zu: list %zu at %u of %u entries; enlarged\n", r.n, g,
                                 st.maxpairs[g], cap);
                }
                r.globals[g].cap = grown_cap(std::max(st.maxpairs[g], cap));
            }
        }
        relayout(r, false);
        h_states_[r.slot] = state_of(r);
    }

    int precision_;
    hipStream_t stream_ = nullptr;                 // class 0's stream; set-up copies and joins
    hipStream_t cls_stream_[kRepClasses] = {};
    hipEvent_t ev_fork_ = nullptr, ev_join_[kRepClasses] = {};
    size_t grp_count_[kRepClasses] = {}, grp_first_[kRepClasses] = {}, grp_lds_[kRepClasses] = {};
    size_t grp_tsp_[kRepClasses] = {};             // replicas with 3SPN2 terms, last in their class
    std::vector<std::unique_ptr<Replica>> reps_;
    RepDesc* d_descs_ = nullptr;
    size_t desc_cap_ = 0;
    int segment_ = 100;
    bool autotune_ = true;         // MJG_AUTOTUNE=0 starts a batch with the tuner off
    std::map<TuneKey, RepTune> tune_cache_;
    bool trace_ = false;           // MJG_BATCH_TRACE: report segment reruns on stderr
    bool descs_dirty_ = true;
    std::vector<int> last_ids_;
    RepState* d_states_ = nullptr;
    size_t states_cap_ = 0, n_slots_ = 0;
    std::vector<RepState> h_states_;
    std::vector<int> free_slots_;
    std::vector<int> failed_now_;
};

} // mjg

// ======================================================================================
// C ABI
// ======================================================================================
struct mjg_batch
{
    mjg::Batch* impl = nullptr;
    std::string err;
};

namespace
{
template<typename F>
mjg_status bguard(mjg_batch* b, F&& f)
{
    if(b == nullptr || b->impl == nullptr) {return MJG_INVALID_ARG;}
    try
    {
        b->impl->clear_failures();
        f(*b->impl);
        b->err.clear();
        // replicas that stopped with non-finite state during this call: the others completed it
        const std::vector<int>& bad = b->impl->failures_of_call();
        if(!bad.empty())
        {
            b->err = "non-finite state in replica(s)";
            for(int id : bad) {b->err += " " + std::to_string(id);}
            b->err += "; they were stopped, the other replicas completed the call";
            return MJG_NAN_DETECTED;
        }
        return MJG_OK;
    }
    catch(const mjg::BatchNonFinite& x) {b->err = x.what(); return MJG_NAN_DETECTED;}
    catch(const mjg::BatchDeviceError& x) {b->err = x.what(); return MJG_DEVICE_ERROR;}
    catch(const std::bad_alloc&) {b->err = "device or host allocation failed"; return MJG_OOM;}
    catch(const std::logic_error& x)
    {
        b->err = x.what();
        if(dynamic_cast<const std::invalid_argument*>(&x) || dynamic_cast<const std::out_of_range*>(&x))
        {
            return MJG_INVALID_ARG;
        }
        return MJG_NOT_INITIALIZED;
    }
    catch(const std::exception& x) {b->err = x.what(); return MJG_INVALID_ARG;}
}
} // namespace

extern "C" {

mjg_status mjg_batch_create(int precision, int device, mjg_batch** out)
{
    if(out == nullptr) {return MJG_INVALID_ARG;}
    *out = nullptr;
    mjg_batch* b = new(std::nothrow) mjg_batch;
    if(b == nullptr) {return MJG_OOM;}
    try {b->impl = new mjg::Batch(precision, device);}
    catch(const std::bad_alloc&) {delete b; return MJG_OOM;}
    catch(const mjg::BatchDeviceError&) {delete b; return MJG_DEVICE_ERROR;}
    catch(const std::exception&) {delete b; return MJG_INVALID_ARG;}
    *out = b;
    return MJG_OK;
}
void mjg_batch_destroy(mjg_batch* b)
{
    if(!b) {return;}
    delete b->impl;
    delete b;
}
const char* mjg_batch_last_error(const mjg_batch* b) {return b ? b->err.c_str() : "null batch";}

mjg_status mjg_batch_add_replica(mjg_batch* b, int boundary, const double* lower, const double* upper, size_t n,
                                 const double* mass, const double* pos, const double* vel, int* replica)
{
    return bguard(b, [&](mjg::Batch& x) {
        const int id = x.begin_replica(boundary, lower, upper, n, mass, pos, vel);
        if(replica) {*replica = id;}
    });
}
mjg_status mjg_batch_set_attribute(mjg_batch* b, int replica, const char* key, double value)
{
    if(!key) {return MJG_INVALID_ARG;}
    return bguard(b, [&](mjg::Batch& x) {x.set_attribute(replica, key, value);});
}
mjg_status mjg_batch_add_local(mjg_batch* b, int replica, int kind, const char* topology, size_t n_terms,
                               const uint32_t* indices, const double* params)
{
    return bguard(b, [&](mjg::Batch& x) {x.add_local(replica, kind, topology, n_terms, indices, params);});
}
mjg_status mjg_batch_add_global(mjg_batch* b, int replica, int potential, double cutoff_ratio, double epsilon,
                                int partition, double margin, int ignore_molecule, size_t n_within,
                                const char* const* within_kinds, const uint32_t* within_dist, size_t n_participants,
                                const uint32_t* participants, const double* params)
{
    return bguard(b, [&](mjg::Batch& x) {
        x.add_global(replica, potential, cutoff_ratio, epsilon, partition, margin, ignore_molecule, n_within,
                     within_kinds, within_dist, n_participants, participants, params);
    });
}
mjg_status mjg_batch_set_integrator(mjg_batch* b, int replica, int type, double dt, const double* coefs, uint64_t seed)
{
    return bguard(b, [&](mjg::Batch& x) {x.set_integrator(replica, type, dt, coefs, seed);});
}
mjg_status mjg_batch_initialize(mjg_batch* b) {return bguard(b, [&](mjg::Batch& x) {x.initialize();});}
mjg_status mjg_batch_step(mjg_batch* b, size_t n_steps) {return bguard(b, [&](mjg::Batch& x) {x.step(n_steps);});}
mjg_status mjg_batch_set_steps(mjg_batch* b, int replica, uint64_t steps)
{
    return bguard(b, [&](mjg::Batch& x) {x.set_steps(replica, steps);});
}
mjg_status mjg_batch_set_step_count(mjg_batch* b, int replica, uint64_t step)
{
    return bguard(b, [&](mjg::Batch& x) {x.set_step_count(replica, step);});
}
mjg_status mjg_batch_advance(mjg_batch* b, int segment) {return bguard(b, [&](mjg::Batch& x) {x.advance(segment);});}
mjg_status mjg_batch_retire(mjg_batch* b, int replica) {return bguard(b, [&](mjg::Batch& x) {x.retire(replica);});}
mjg_status mjg_batch_get_state(mjg_batch* b, int replica, double* pos, double* vel, double* force)
{
    return bguard(b, [&](mjg::Batch& x) {x.get_state(replica, pos, vel, force);});
}
mjg_status mjg_batch_set_autotune(mjg_batch* b, int on) {return bguard(b, [&](mjg::Batch& x) {x.set_autotune(on != 0);});}
mjg_status mjg_batch_set_tuning(mjg_batch* b, int replica, double list_skin_short, double list_skin_dh, int grid_bits,
                                int grid_lanes)
{
    return bguard(b, [&](mjg::Batch& x) {x.set_tuning(replica, list_skin_short, list_skin_dh, grid_bits, grid_lanes);});
}
mjg_status mjg_batch_replica_tuning(mjg_batch* b, int replica, double* list_skin_short, double* list_skin_dh,
                                    int* grid_bits, int* grid_lanes)
{
    return bguard(b, [&](mjg::Batch& x) {
        const mjg::RepTune t = x.tuning(replica);
        if(list_skin_short) {*list_skin_short = t.m_short;}
        if(list_skin_dh) {*list_skin_dh = t.m_dh;}
        if(grid_bits) {*grid_bits = t.grid_bits;}
        if(grid_lanes) {*grid_lanes = t.grid_lanes;}
    });
}
mjg_status mjg_batch_replica_status(mjg_batch* b, int replica, int* status)
{
    if(status == nullptr) {return MJG_INVALID_ARG;}
    return bguard(b, [&](mjg::Batch& x) {*status = x.failed(replica) ? MJG_NAN_DETECTED : MJG_OK;});
}
mjg_status mjg_batch_replica_info(mjg_batch* b, int replica, uint64_t* step, uint64_t* remaining, uint64_t* rebuilds,
                                  size_t* footprint_bytes)
{
    return bguard(b, [&](mjg::Batch& x) {
        const mjg::RepState s = x.state(replica);
        if(step) {*step = s.step;}
        if(remaining) {*remaining = s.remaining;}
        if(rebuilds) {*rebuilds = s.rebuilds;}
        if(footprint_bytes) {*footprint_bytes = x.footprint(replica);}
    });
}
mjg_status mjg_batch_list_sizes(mjg_batch* b, int replica, uint64_t* entries, int max_globals)
{
    if(!entries && max_globals > 0) {return MJG_INVALID_ARG;}
    return bguard(b, [&](mjg::Batch& x) {
        const mjg::RepState s = x.state(replica);
        const int ng = std::min(max_globals, x.num_globals(replica));
        for(int g = 0; g < ng; ++g) {entries[g] = s.npairs[g];}
    });
}
mjg_status mjg_batch_observe(mjg_batch* b) {return bguard(b, [&](mjg::Batch& x) {x.observe();});}
mjg_status mjg_batch_energies(mjg_batch* b, int replica, double* per_interaction, double* kinetic)
{
    return bguard(b, [&](mjg::Batch& x) {x.energies(replica, per_interaction, kinetic);});
}
mjg_status mjg_batch_evaluate(mjg_batch* b, int replica, double* per_interaction, double* total, double* force)
{
    return bguard(b, [&](mjg::Batch& x) {x.evaluate(replica, per_interaction, total, force);});
}
mjg_status mjg_batch_counters(mjg_batch* b, uint64_t* particle_steps, uint64_t* rebuilds)
{
    return bguard(b, [&](mjg::Batch& x) {x.counters(particle_steps, rebuilds);});
}
mjg_status mjg_batch_profile(mjg_batch* b, int segment_steps, int reps, double* avg_ms, uint64_t* particle_steps)
{
    return bguard(b, [&](mjg::Batch& x) {
        const double v = x.profile(segment_steps, reps, particle_steps);
        if(avg_ms) {*avg_ms = v;}
    });
}

} // extern "C"
