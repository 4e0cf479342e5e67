// Checker of the drop-in boundary: the reference's own classes driving libmjg.so through the C++
// adapter include/mjg_forcefield.hpp, against the reference's own ForceField on the same input.
//
// TEST INFRASTRUCTURE ONLY: compiled against the reference sources where they lie (oracle/Makefile)
// into oracle/_ref/, linked with mjolnir_amd/libmjg.so.  It reuses ref_driver.cpp's mjspec reader
// and object construction (the objects read_input_file would build).
//
// Commands (each prints one JSON line):
//   dropin_check forces <spec>          ForceFieldBase calls through both force fields: calc_force,
//                                       calc_energy, calc_force_and_energy, format_energy_name,
//                                       format_energy, calc_force_and_virial
//   dropin_check run <spec> <nsteps>    the reference's integrator (BAOAB / G-JF / UnderdampedLangevin
//                                       as the spec names) stepping the reference ForceField and, as
//                                       a std::unique_ptr<ForceFieldBase>, mjg::GpuForceField
//   dropin_check sim <spec> <total> <save> <prefix>
//                                       the reference's MolecularDynamicsSimulator with
//                                       mjg::GpuForceField and mjg::GpuBAOAB / GpuGJF and the
//                                       reference's EnergyObserver: writes <prefix>.ene
#define REF_DRIVER_NO_MAIN
#include "ref_driver.cpp"

#include <mjolnir/core/MolecularDynamicsSimulator.hpp>
#include <mjolnir/core/ObserverContainer.hpp>

#include "../include/mjg_forcefield.hpp"

#include <cmath>
#include <cstring>

namespace chk
{
using namespace mjolnir;

static int local_code(const std::string& inter, const std::string& pot)
{
    if(inter == "BondLength" && pot == "Harmonic") {return MJG_BOND_HARMONIC;}
    if(inter == "BondLength" && pot == "GoContact") {return MJG_BOND_GOCONTACT;}
    if(inter == "Contact" && pot == "GoContact") {return MJG_CONTACT_GOCONTACT;}
    if(inter == "BondAngle" && pot == "Harmonic") {return MJG_ANGLE_HARMONIC;}
    if(inter == "BondAngle" && pot == "Gaussian") {return MJG_ANGLE_GAUSSIAN;}
    if(inter == "DihedralAngle" && pot == "ClementiDihedral") {return MJG_DIHEDRAL_CLEMENTI;}
    if(inter == "BondLength" && pot == "3SPN2Bond") {return MJG_BOND_3SPN2;}
    if(inter == "DihedralAngle" && pot == "Gaussian") {return MJG_DIHEDRAL_GAUSSIAN;}
    if(inter == "3SPN2BaseStacking") {return MJG_BASE_STACKING_3SPN2;}
    throw std::runtime_error("no GPU local interaction for " + inter + ":" + pot);
}
static int global_code(const std::string& pot)
{
    if(pot == "LennardJones") {return MJG_PAIR_LJ;}
    if(pot == "ExcludedVolume") {return MJG_PAIR_EV;}
    if(pot == "DebyeHuckel") {return MJG_PAIR_DH;}
    if(pot == "3SPN2ExcludedVolume") {return MJG_PAIR_3SPN2_EV;}
    if(pot == "3SPN2BasePair") {return MJG_3SPN2_BASEPAIR;}
    if(pot == "3SPN2CrossStacking") {return MJG_3SPN2_CROSSSTACKING;}
    throw std::runtime_error("no GPU global interaction for " + pot);
}
static int partition_code(const std::string& p)
{
    if(p == "CellList") {return MJG_PART_CELLLIST;}
    if(p == "VerletList") {return MJG_PART_VERLET;}
    return MJG_PART_NAIVE;
}
static int ignore_code(const std::string& m)
{
    if(m == "Self" || m == "Intra") {return MJG_IGNORE_SELF;}
    if(m == "Others" || m == "Inter") {return MJG_IGNORE_OTHERS;}
    return MJG_IGNORE_NOTHING;
}
static void ok(mjg_engine* e, mjg_status s, const char* what)
{
    if(s != MJG_OK) {throw std::runtime_error(std::string(what) + ": " + (e ? mjg_last_error(e) : "create failed"));}
}

// The C-ABI calls a maintainer's reader makes for the same [[forcefields]] tables: one engine with
// every interaction of the spec (interactions only; the state comes from the System at initialize).
static mjg_engine* make_engine(const drv::SpecIn& s)
{
    mjg_engine* e = nullptr;
    ok(nullptr, mjg_engine_create(MJG_FP64, s.periodic ? MJG_PERIODIC : MJG_UNLIMITED, s.lower, s.upper,
                                  s.mass.size(), 0, &e), "mjg_engine_create");
    ok(e, mjg_system_set_state(e, s.mass.data(), s.x.data(), s.v.data()), "mjg_system_set_state");
    for(const auto& kv : s.attributes) {ok(e, mjg_system_set_attribute(e, kv.first.c_str(), kv.second), "attribute");}
    int id = 0;
    for(const auto& l : s.local)
    {
        std::vector<uint32_t> idx;
        std::vector<double> par;
        for(const auto& t : l.idx) {for(auto i : t) {idx.push_back(static_cast<uint32_t>(i));}}
        for(const auto& p : l.par) {par.insert(par.end(), p.begin(), p.end());}
        ok(e, mjg_add_local(e, local_code(l.interaction, l.potential), l.topology.c_str(), l.idx.size(), idx.data(),
                            par.data(), l.margin, &id), "mjg_add_local");
    }
    for(const auto& g : s.global)
    {
        std::vector<const char*> kinds;
        std::vector<uint32_t> dists, part;
        std::vector<double> par;
        for(const auto& kv : g.within) {kinds.push_back(kv.first.c_str()); dists.push_back(static_cast<uint32_t>(kv.second));}
        for(auto p : g.participants) {part.push_back(static_cast<uint32_t>(p));}
        for(const auto& p : g.par) {par.insert(par.end(), p.begin(), p.end());}
        ok(e, mjg_add_global(e, global_code(g.potential), g.cutoff, g.epsilon, partition_code(g.partition), g.margin,
                             ignore_code(g.ignore_mol), kinds.size(), kinds.data(), dists.data(), part.size(),
                             part.data(), par.data(), &id), "mjg_add_global");
    }
    if(s.has_constraint)
    {
        std::vector<uint32_t> idx;
        for(const auto& c : s.con_idx) {idx.push_back((uint32_t)c[0]); idx.push_back((uint32_t)c[1]);}
        ok(e, mjg_add_constraint(e, s.con_topology.c_str(), s.con_v0.size(), idx.data(), s.con_v0.data(),
                                 s.con_max_iteration, s.con_tolerance), "mjg_add_constraint");
    }
    return e;
}

template<typename traitsT>
static std::unique_ptr<ForceFieldBase<traitsT>> make_gpu_ff(const drv::SpecIn& s, const ForceFieldBase<traitsT>& ref)
{
    // the reference's Topology of the same interactions (what the observers read)
    return std::unique_ptr<ForceFieldBase<traitsT>>(new mjg::GpuForceField<traitsT>(make_engine(s), ref.topology()));
}

template<typename traitsT>
static double max_rel_force_err(const System<traitsT>& a, const System<traitsT>& b)
{
    double rms = 0, worst = 0;
    for(std::size_t i = 0; i < a.size(); ++i) {rms += math::length_sq(a.force(i));}
    rms = std::sqrt(rms / std::max<std::size_t>(1, a.size()));
    for(std::size_t i = 0; i < a.size(); ++i)
    {
        const double d = math::length(a.force(i) - b.force(i));
        worst = std::max(worst, d / (math::length(a.force(i)) + rms));
    }
    return worst;
}

template<typename traitsT>
static void zero(System<traitsT>& sys)
{
    for(std::size_t i = 0; i < sys.size(); ++i) {sys.force(i) = math::make_coordinate<typename traitsT::coordinate_type>(0, 0, 0);}
}

// the numbers of two format_energy lines, compared with a relative tolerance
static double max_rel_number_diff(const std::string& a, const std::string& b, std::size_t* count)
{
    std::istringstream sa(a), sb(b);
    double x, y, worst = 0;
    std::size_t n = 0;
    while(sa >> x)
    {
        if(!(sb >> y)) {return INFINITY;}
        worst = std::max(worst, std::abs(x - y) / std::max(1.0, std::abs(x)));
        ++n;
    }
    if(sb >> y) {return INFINITY;}
    *count = n;
    return worst;
}

template<typename traitsT>
static int cmd_forces(const drv::SpecIn& s)
{
    auto a = drv::build<traitsT>(s);
    auto b = drv::build<traitsT>(s);
    auto& sa = *a.sys;
    auto& sb = *b.sys;
    a.ff->initialize(sa);
    auto gff = make_gpu_ff<traitsT>(s, *a.ff);
    gff->initialize(sb);
    ForceFieldBase<traitsT>& ref = *a.ff;
    ForceFieldBase<traitsT>& gpu = *gff;

    zero(sa); zero(sb);
    ref.calc_force(sa);
    gpu.calc_force(sb);
    const double f_err = max_rel_force_err(sa, sb);

    const double e_ref = ref.calc_energy(sa), e_gpu = gpu.calc_energy(sb);
    zero(sa); zero(sb);
    const double fe_ref = ref.calc_force_and_energy(sa), fe_gpu = gpu.calc_force_and_energy(sb);
    const double fe_err = max_rel_force_err(sa, sb);

    std::string n_ref, n_gpu, l_ref, l_gpu;
    ref.format_energy_name(n_ref);
    gpu.format_energy_name(n_gpu);
    const double t_ref = ref.format_energy(sa, l_ref), t_gpu = gpu.format_energy(sb, l_gpu);
    std::size_t nums = 0;
    const double line_err = max_rel_number_diff(l_ref, l_gpu, &nums);

    zero(sb);
    auto* g = dynamic_cast<mjg::GpuForceField<traitsT>*>(gff.get());
    g->clear_status();
    gpu.calc_force_and_virial(sb);
    const bool virial_nan = std::isnan(sb.virial()[0]);
    const int virial_status = g->status();
    g->clear_status();

    std::printf("{\"force_rel_err\": %.3e, \"force_and_energy_rel_err\": %.3e, \"energy_ref\": %.17g, "
                "\"energy_gpu\": %.17g, \"force_and_energy_ref\": %.17g, \"force_and_energy_gpu\": %.17g, "
                "\"names_equal\": %s, \"name_ref\": \"%s\", \"name_gpu\": \"%s\", \"format_line_equal_width\": %s, "
                "\"format_numbers\": %zu, \"format_rel_err\": %.3e, \"format_total_ref\": %.17g, "
                "\"format_total_gpu\": %.17g, \"virial_nan\": %s, \"virial_status\": %d}\n",
                f_err, fe_err, e_ref, e_gpu, fe_ref, fe_gpu, n_ref == n_gpu ? "true" : "false", n_ref.c_str(),
                n_gpu.c_str(), l_ref.size() == l_gpu.size() ? "true" : "false", nums, line_err, t_ref, t_gpu,
                virial_nan ? "true" : "false", virial_status);
    return 0;
}

// the reference's integrator, unchanged, with either force field behind the same base pointer
template<typename traitsT, typename integT>
static int run_impl(const drv::SpecIn& s, std::size_t nsteps)
{
    using real_type = typename traitsT::real_type;
    auto a = drv::build<traitsT>(s);
    auto b = drv::build<traitsT>(s);
    auto& sa = *a.sys;
    auto& sb = *b.sys;
    RandomNumberGenerator<traitsT> ra(s.seed), rb(s.seed);
    sa.initialize(ra);
    sb.initialize(rb);
    a.ff->initialize(sa);
    std::unique_ptr<ForceFieldBase<traitsT>> gff = make_gpu_ff<traitsT>(s, *a.ff);
    gff->initialize(sb);
    std::vector<real_type> ca(s.coefs.begin(), s.coefs.end()), cb(s.coefs.begin(), s.coefs.end());
    integT ia(s.dt, std::move(ca), SystemMotionRemover<traitsT>(s.rm_translation, s.rm_rotation, s.rm_rescale));
    integT ib(s.dt, std::move(cb), SystemMotionRemover<traitsT>(s.rm_translation, s.rm_rotation, s.rm_rescale));
    ia.initialize(sa, a.ff, ra);
    ib.initialize(sb, gff, rb);
    real_type ta = 0, tb = 0;
    double worst = 0;
    std::printf("{\"frames\": [");
    for(std::size_t step = 1; step <= nsteps; ++step)
    {
        ta = ia.step(ta, sa, a.ff, ra);
        tb = ib.step(tb, sb, gff, rb);
        if(step % 100 == 0 || step == nsteps)
        {
            double d = 0;
            for(std::size_t i = 0; i < sa.size(); ++i)
            {
                auto dr = sa.position(i) - sb.position(i);
                if(s.periodic) {dr = sa.adjust_direction(sb.position(i), sa.position(i));}
                d = std::max(d, std::max(std::abs(dr[0]), std::max(std::abs(dr[1]), std::abs(dr[2]))));
            }
            worst = std::max(worst, d);
            std::printf("%s[%zu, %.3e]", step == 100 || step == nsteps && nsteps < 100 ? "" : ", ", step, d);
        }
    }
    const int st = dynamic_cast<mjg::GpuForceField<traitsT>*>(gff.get())->status();
    std::printf("], \"max_abs_dx\": %.3e, \"status\": %d}\n", worst, st);
    return 0;
}

template<typename traitsT>
static int cmd_run(const drv::SpecIn& s, std::size_t nsteps)
{
    if(s.integrator == "BAOABLangevin") {return chk::run_impl<traitsT, BAOABLangevinIntegrator<traitsT>>(s, nsteps);}
    if(s.integrator == "GJFNVTLangevin") {return chk::run_impl<traitsT, GJFNVTLangevinIntegrator<traitsT>>(s, nsteps);}
    if(s.integrator == "UnderdampedLangevin")
    {
        return chk::run_impl<traitsT, UnderdampedLangevinIntegrator<traitsT>>(s, nsteps);
    }
    throw std::runtime_error("dropin_check run: unsupported integrator " + s.integrator);
}

// MolecularDynamicsSimulator<traits, GPU integrator> with the GPU force field and the reference's
// EnergyObserver (MolecularDynamicsSimulator.hpp:72-123, EnergyObserver.hpp:27-150)
template<typename traitsT, typename integT>
static int sim_impl(const drv::SpecIn& s, std::size_t total, std::size_t save, const std::string& prefix)
{
    using real_type = typename traitsT::real_type;
    using sim_type = MolecularDynamicsSimulator<traitsT, integT>;
    auto a = drv::build<traitsT>(s);
    a.ff->initialize(*a.sys);   // the reference's topology, for the GPU force field's observers
    auto gff = make_gpu_ff<traitsT>(s, *a.ff);
    std::vector<real_type> coefs(s.coefs.begin(), s.coefs.end());
    integT integ(s.dt, std::move(coefs), SystemMotionRemover<traitsT>(s.rm_translation, s.rm_rotation, s.rm_rescale));
    ObserverContainer<traitsT> obs(false);
    obs.push_back(std::unique_ptr<ObserverBase<traitsT>>(new EnergyObserver<traitsT>(prefix)));
    sim_type sim(total, save, /*checkpoint*/ total, std::move(*a.sys), std::move(gff), std::move(integ),
                 std::move(obs), RandomNumberGenerator<traitsT>(s.seed));
    sim.initialize();
    while(sim.step()) {}
    sim.finalize();
    std::printf("{\"written\": \"%s.ene\"}\n", prefix.c_str());
    return 0;
}

template<typename traitsT>
static int cmd_sim(const drv::SpecIn& s, std::size_t total, std::size_t save, const std::string& prefix)
{
    if(s.integrator == "BAOABLangevin") {return chk::sim_impl<traitsT, mjg::GpuBAOABLangevinIntegrator<traitsT>>(s, total, save, prefix);}
    if(s.integrator == "GJFNVTLangevin") {return chk::sim_impl<traitsT, mjg::GpuGJFNVTLangevinIntegrator<traitsT>>(s, total, save, prefix);}
    throw std::runtime_error("dropin_check sim: unsupported integrator " + s.integrator);
}

template<template<typename, typename> class B>
static int dispatch(const std::string& cmd, const drv::SpecIn& s, int argc, char** argv)
{
    using traits = SimulatorTraits<double, B>;
    if(cmd == "forces") {return chk::cmd_forces<traits>(s);}
    if(cmd == "run" && argc > 3) {return chk::cmd_run<traits>(s, std::stoul(argv[3]));}
    if(cmd == "sim" && argc > 5) {return chk::cmd_sim<traits>(s, std::stoul(argv[3]), std::stoul(argv[4]), argv[5]);}
    throw std::runtime_error("usage: dropin_check forces|run|sim <spec> ...");
}
} // chk

int main(int argc, char** argv)
{
    if(argc < 3)
    {
        std::cerr << "usage: dropin_check forces|run|sim <spec> ...\n";
        return 2;
    }
    try
    {
        mjolnir::LoggerManager::set_default_logger("/dev/null");
        drv::set_units_kcal_angstrom<double>();
        const auto spec = drv::read_spec(argv[2]);
        if(spec.periodic) {return chk::dispatch<mjolnir::CuboidalPeriodicBoundary>(argv[1], spec, argc, argv);}
        return chk::dispatch<mjolnir::UnlimitedBoundary>(argv[1], spec, argc, argv);
    }
    catch(const std::exception& e)
    {
        std::cerr << "dropin_check: " << e.what() << std::endl;
        return 1;
    }
}
