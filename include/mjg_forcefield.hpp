/* mjg_forcefield.hpp: the reference-side C++ binding of libmjg.so.
 *
 * Header-only adapter that puts the C ABI of mjg.h behind the reference's own class surface, so that
 * a Mjolnir build can hold the GPU engine wherever it holds its CPU objects:
 *
 *   mjg::GpuForceField<traitsT>                 : mjolnir::ForceFieldBase<traitsT>
 *       every virtual of mjolnir/core/ForceFieldBase.hpp:27-51, held as the
 *       std::unique_ptr<ForceFieldBase<traitsT>> the integrators take
 *       (mjolnir/core/BAOABLangevinIntegrator.hpp:27);
 *   mjg::GpuBAOABLangevinIntegrator<traitsT>    the shape of mjolnir/core/BAOABLangevinIntegrator.hpp:34-67
 *   mjg::GpuGJFNVTLangevinIntegrator<traitsT>   the shape of mjolnir/core/GJFNVTLangevinIntegrator.hpp:34-72
 *       constructor (dt, per-particle gammas / alphas, SystemMotionRemover), initialize, step,
 *       update, delta_t, parameters: what MolecularDynamicsSimulator<traitsT, integratorT> calls
 *       (mjolnir/core/MolecularDynamicsSimulator.hpp:72-123).
 *
 * Two ways to use them:
 *   - GpuForceField under the reference's own integrator: every calc_force uploads the positions the
 *     host integrator moved (mjg_system_set_positions), computes on the device and downloads the
 *     forces into the System.  Correct and PCIe-bound: N * 48 bytes cross the bus per call.
 *   - GpuForceField with a Gpu*Integrator: the step runs on the device (mjg_step) and the System is
 *     refreshed from it after each step, so observers see the same state they see on the CPU.
 *
 * Needs the reference headers (include path to the Mjolnir source tree) and libmjg.so.  Device
 * errors in noexcept calls become a sticky status (status()); set-up errors throw mjg::Error.
 */
#ifndef MJG_FORCEFIELD_HPP
#define MJG_FORCEFIELD_HPP

#include <mjolnir/core/ForceFieldBase.hpp>
#include <mjolnir/core/RandomNumberGenerator.hpp>
#include <mjolnir/core/SystemMotionRemover.hpp>
#include <mjolnir/core/Topology.hpp>

#include "mjg.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iomanip>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mjg
{

struct Error : std::runtime_error
{
    Error(mjg_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    mjg_status status;
};

template<typename traitsT>
class GpuForceField final : public mjolnir::ForceFieldBase<traitsT>
{
  public:
    using base_type       = mjolnir::ForceFieldBase<traitsT>;
    using real_type       = typename base_type::real_type;
    using system_type     = typename base_type::system_type;
    using topology_type   = typename base_type::topology_type;
    using constraint_type = typename base_type::constraint_type;

    // Takes ownership of an engine whose interactions were added with mjg_add_local / mjg_add_global
    // (and mjg_add_constraint) but that is not initialized yet; `topology` is the reference's
    // Topology of the same interactions, which observers read (e.g. DCDObserver's molecule count).
    GpuForceField(mjg_engine* engine, topology_type topology)
        : e_(engine), topology_(std::move(topology)), constraint_("none", typename constraint_type::container_type{}, 0, 0)
    {
        if(e_ == nullptr) {throw Error(MJG_INVALID_ARG, "mjg::GpuForceField: null engine");}
    }
    GpuForceField(mjg_engine* engine, topology_type topology, constraint_type constraint)
        : e_(engine), topology_(std::move(topology)), constraint_(std::move(constraint))
    {
        if(e_ == nullptr) {throw Error(MJG_INVALID_ARG, "mjg::GpuForceField: null engine");}
    }
    ~GpuForceField() override {mjg_engine_destroy(e_);}
    GpuForceField(const GpuForceField&) = delete;
    GpuForceField& operator=(const GpuForceField&) = delete;

    // ForceField::initialize (ForceField.hpp:47-63): state and attributes, then topology,
    // exclusions, parameters and the first neighbour lists on the device.
    void initialize(const system_type& sys) override
    {
        upload_state(sys);
        upload_attributes(sys);
        check(mjg_forcefield_initialize(e_), "mjg_forcefield_initialize");
        initialized_ = true;
    }
    // ForceField::update: system parameters changed (temperature, ionic strength): everything that
    // depends on them is rebuilt from the current state.
    void update(const system_type& sys) override
    {
        upload_state(sys);
        upload_attributes(sys);
        check(mjg_forcefield_initialize(e_), "mjg_forcefield_initialize");
    }
    void reduce_margin(const real_type dmargin, const system_type& sys) override
    {
        set_positions(sys);
        check(mjg_reduce_margin(e_, static_cast<double>(dmargin)), "mjg_reduce_margin");
    }
    // scale_margin follows a box change (constant-pressure integrators); the engine's box is fixed.
    void scale_margin(const real_type scale, const system_type&) override
    {
        if(scale != real_type(1))
        {
            throw Error(MJG_UNSUPPORTED, "mjg::GpuForceField: the box is fixed; scale_margin is not supported");
        }
    }

    void calc_force(system_type& sys) const noexcept override
    {
        if(!set_positions_nothrow(sys)) {return;}
        note(mjg_calc_force(e_));
        download_forces(sys);
    }
    // No kernel accumulates the virial (mjg_calc_force_and_virial returns MJG_UNSUPPORTED and NaN):
    // the forces are computed, the virial is set to NaN and status() reports MJG_UNSUPPORTED.
    void calc_force_and_virial(system_type& sys) const noexcept override
    {
        calc_force(sys);
        double v[9];
        note(mjg_calc_force_and_virial(e_, v));
        for(std::size_t k = 0; k < 9; ++k) {sys.virial()[k] = static_cast<real_type>(v[k]);}
    }
    real_type calc_energy(const system_type& sys) const noexcept override
    {
        if(!set_positions_nothrow(sys)) {return std::numeric_limits<real_type>::quiet_NaN();}
        double total = 0;
        note(mjg_calc_energies(e_, nullptr, &total));
        return static_cast<real_type>(total);
    }
    real_type calc_force_and_energy(system_type& sys) const noexcept override
    {
        if(!set_positions_nothrow(sys)) {return std::numeric_limits<real_type>::quiet_NaN();}
        double total = 0;
        note(mjg_calc_force_and_energy(e_, &total));
        download_forces(sys);
        return static_cast<real_type>(total);
    }
    real_type calc_force_virial_energy(system_type& sys) const noexcept override
    {
        const real_type e = calc_force_and_energy(sys);
        double v[9];
        note(mjg_calc_force_and_virial(e_, v));
        for(std::size_t k = 0; k < 9; ++k) {sys.virial()[k] = static_cast<real_type>(v[k]);}
        return e;
    }

    // LocalForceField / GlobalForceField::format_energy_name (LocalForceField.hpp:158-171): every
    // interaction's name followed by a space, locals first.
    void format_energy_name(std::string& fmt) const override
    {
        for(std::size_t k = 0; k < size(); ++k) {fmt += name(k); fmt += ' ';}
    }
    // LocalForceField / GlobalForceField::format_energy (LocalForceField.hpp:173-200): each energy
    // right-aligned in max(name length, 10) columns, fixed notation, then a space; returns the total.
    real_type format_energy(const system_type& sys, std::string& fmt) const override
    {
        std::vector<double> e(std::max<std::size_t>(1, size()), 0.0);
        double total = 0;
        if(set_positions_nothrow(sys)) {note(mjg_calc_energies(e_, e.data(), &total));}
        std::ostringstream oss;
        real_type sum = 0;
        for(std::size_t k = 0; k < size(); ++k)
        {
            const std::size_t w = std::max<std::size_t>(name(k).size(), 10);
            oss << std::setw(static_cast<int>(w)) << std::fixed << std::right << static_cast<real_type>(e[k]) << ' ';
            sum += static_cast<real_type>(e[k]);
        }
        fmt += oss.str();
        return sum;
    }

    topology_type const& topology() const noexcept override {return topology_;}
    constraint_type const& constraint() const noexcept override {return constraint_;}

    // the engine, for the device-side integrators below
    mjg_engine* engine() const noexcept {return e_;}
    // MJG_OK, or the first non-OK status of a noexcept call since the last clear_status()
    mjg_status status() const noexcept {return status_;}
    void clear_status() noexcept {status_ = MJG_OK;}
    std::size_t size() const noexcept {return mjg_num_interactions(e_);}

    std::string name(std::size_t k) const
    {
        std::size_t len = 0;
        if(mjg_interaction_name(e_, static_cast<int>(k), nullptr, 0, &len) != MJG_OK) {return "Unknown";}
        std::string s(len + 1, '\0');
        mjg_interaction_name(e_, static_cast<int>(k), &s[0], s.size(), &len);
        s.resize(len);
        return s;
    }

    // the device state into the System (positions, velocities, forces)
    void download_state(system_type& sys) const noexcept
    {
        const std::size_t n = sys.size();
        pos_.resize(3 * n);
        vel_.resize(3 * n);
        frc_.resize(3 * n);
        if(!note(mjg_system_get_state(e_, pos_.data(), vel_.data(), frc_.data()))) {return;}
        for(std::size_t i = 0; i < n; ++i)
        {
            for(std::size_t d = 0; d < 3; ++d)
            {
                sys.position(i)[d] = static_cast<real_type>(pos_[3 * i + d]);
                sys.velocity(i)[d] = static_cast<real_type>(vel_[3 * i + d]);
                sys.force(i)[d] = static_cast<real_type>(frc_[3 * i + d]);
            }
        }
    }
    void upload_state(const system_type& sys)
    {
        const std::size_t n = sys.size();
        std::vector<double> m(n), x(3 * n), v(3 * n);
        for(std::size_t i = 0; i < n; ++i)
        {
            m[i] = static_cast<double>(sys.mass(i));
            for(std::size_t d = 0; d < 3; ++d)
            {
                x[3 * i + d] = static_cast<double>(sys.position(i)[d]);
                v[3 * i + d] = static_cast<double>(sys.velocity(i)[d]);
            }
        }
        check(mjg_system_set_state(e_, m.data(), x.data(), v.data()), "mjg_system_set_state");
    }
    void check(mjg_status s, const char* what) const
    {
        if(s != MJG_OK) {throw Error(s, std::string(what) + ": " + mjg_last_error(e_));}
    }

  private:
    void upload_attributes(const system_type& sys)
    {
        for(const char* key : {"temperature", "ionic_strength"})
        {
            if(sys.has_attribute(key))
            {
                check(mjg_system_set_attribute(e_, key, static_cast<double>(sys.attribute(key))), "mjg_system_set_attribute");
            }
        }
    }
    void set_positions(const system_type& sys)
    {
        if(!set_positions_nothrow(sys)) {throw Error(status_, std::string("mjg_system_set_positions: ") + mjg_last_error(e_));}
    }
    bool set_positions_nothrow(const system_type& sys) const noexcept
    {
        const std::size_t n = sys.size();
        pos_.resize(3 * n);
        for(std::size_t i = 0; i < n; ++i)
        {
            for(std::size_t d = 0; d < 3; ++d) {pos_[3 * i + d] = static_cast<double>(sys.position(i)[d]);}
        }
        return note(mjg_system_set_positions(e_, pos_.data()));
    }
    void download_forces(system_type& sys) const noexcept
    {
        const std::size_t n = sys.size();
        frc_.resize(3 * n);
        if(!note(mjg_system_get_state(e_, nullptr, nullptr, frc_.data()))) {return;}
        // ForceFieldBase::calc_force adds to the forces the System holds (the integrators zero them)
        for(std::size_t i = 0; i < n; ++i)
        {
            for(std::size_t d = 0; d < 3; ++d) {sys.force(i)[d] += static_cast<real_type>(frc_[3 * i + d]);}
        }
    }
    bool note(mjg_status s) const noexcept
    {
        if(s != MJG_OK && status_ == MJG_OK) {status_ = s;}
        return s == MJG_OK;
    }

    mjg_engine* e_;
    topology_type topology_;
    constraint_type constraint_;
    bool initialized_ = false;
    mutable mjg_status status_ = MJG_OK;
    mutable std::vector<double> pos_, vel_, frc_;
};

namespace detail
{
// SystemMotionRemover keeps its flags protected (SystemMotionRemover.hpp:35-38); a member pointer
// named through a derived class reads them.
template<typename traitsT>
struct RemoverFlags : mjolnir::SystemMotionRemover<traitsT>
{
    using base = mjolnir::SystemMotionRemover<traitsT>;
    static bool translation(const base& r) {return r.*(&RemoverFlags::translation_);}
    static bool rotation(const base& r) {return r.*(&RemoverFlags::rotation_);}
    static bool rescale(const base& r) {return r.*(&RemoverFlags::rescale_);}
};

// The counter-based noise stream's seed: two draws of the reference's generator, so that the run
// is a function of the input seed as the reference's is (its mt19937 stream itself is not
// reproduced; RandomNumberGenerator.hpp:44-62).
template<typename traitsT>
std::uint64_t philox_seed(mjolnir::RandomNumberGenerator<traitsT>& rng)
{
    const auto a = static_cast<std::uint64_t>(rng.uniform_real01() * 4294967296.0);
    const auto b = static_cast<std::uint64_t>(rng.uniform_real01() * 4294967296.0);
    return (a << 32) ^ b;
}

// Integrator::step on the device: the engine's step (skin check, rebuilds, force call, closing
// half-kick and the remover), then the System refreshed from the device state.
template<typename traitsT, int kType>
class GpuLangevin
{
  public:
    using traits_type     = traitsT;
    using real_type       = typename traits_type::real_type;
    using system_type     = mjolnir::System<traits_type>;
    using forcefield_type = std::unique_ptr<mjolnir::ForceFieldBase<traits_type>>;
    using rng_type        = mjolnir::RandomNumberGenerator<traits_type>;
    using remover_type    = mjolnir::SystemMotionRemover<traits_type>;

    GpuLangevin(const real_type dt, std::vector<real_type>&& coefs, remover_type&& remover)
        : dt_(dt), coefs_(std::move(coefs)), translation_(RemoverFlags<traitsT>::translation(remover)),
          rotation_(RemoverFlags<traitsT>::rotation(remover)), rescale_(RemoverFlags<traitsT>::rescale(remover))
    {}

    // BAOABLangevinIntegrator::initialize (:135-168) / GJFNVTLangevinIntegrator::initialize (:44-75):
    // parameters from the temperature, then the first force call, here on the device.
    void initialize(system_type& sys, forcefield_type& ff, rng_type& rng)
    {
        gff_ = dynamic_cast<GpuForceField<traitsT>*>(ff.get());
        if(gff_ == nullptr) {throw Error(MJG_INVALID_ARG, "mjg: a GPU integrator needs an mjg::GpuForceField");}
        if(!sys.has_attribute("temperature"))
        {
            throw std::out_of_range("mjg: Langevin integrator requires the `temperature` attribute");
        }
        if(coefs_.size() != sys.size()) {throw Error(MJG_INVALID_ARG, "mjg: one friction coefficient per particle");}
        seed_ = philox_seed(rng);
        gff_->upload_state(sys);
        configure();
        gff_->download_state(sys);
    }
    real_type step(const real_type time, system_type& sys, forcefield_type&, rng_type&)
    {
        gff_->check(mjg_step(gff_->engine(), 1), "mjg_step");
        gff_->download_state(sys);
        return time + dt_;
    }
    // update(sys[, dt]) (BAOABLangevinIntegrator.hpp:46-65): parameters from the new temperature
    void update(const system_type& sys)
    {
        if(gff_ == nullptr) {return;}
        gff_->check(mjg_system_set_attribute(gff_->engine(), "temperature", static_cast<double>(sys.attribute("temperature"))),
                    "mjg_system_set_attribute");
        configure();
    }
    void update(const system_type& sys, const real_type newdt) {dt_ = newdt; update(sys);}

    real_type delta_t() const noexcept {return dt_;}
    std::vector<real_type> const& parameters() const noexcept {return coefs_;}

  private:
    void configure()
    {
        std::vector<double> c(coefs_.begin(), coefs_.end());
        mjg_engine* e = gff_->engine();
        gff_->check(mjg_integrator_set(e, kType, static_cast<double>(dt_), c.data(), seed_), "mjg_integrator_set");
        gff_->check(mjg_integrator_set_remover(e, translation_, rotation_, rescale_), "mjg_integrator_set_remover");
        gff_->check(mjg_integrator_initialize(e), "mjg_integrator_initialize");
    }

    real_type dt_;
    std::vector<real_type> coefs_;
    int translation_, rotation_, rescale_;
    std::uint64_t seed_ = 0;
    GpuForceField<traitsT>* gff_ = nullptr;
};
} // detail

// BAOABLangevinIntegrator(dt, gammas, remover) (BAOABLangevinIntegrator.hpp:34-39)
template<typename traitsT>
using GpuBAOABLangevinIntegrator = detail::GpuLangevin<traitsT, MJG_INTEG_BAOAB>;
// GJFNVTLangevinIntegrator(dt, alphas, remover) (GJFNVTLangevinIntegrator.hpp:34-41)
template<typename traitsT>
using GpuGJFNVTLangevinIntegrator = detail::GpuLangevin<traitsT, MJG_INTEG_GJF>;

} // mjg
#endif
