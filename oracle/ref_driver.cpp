// Checker-side driver that runs Mjolnir's own force-and-step classes on an mjspec file.
//
// TEST INFRASTRUCTURE ONLY: this file is compiled against the reference sources where they
// lie (see oracle/Makefile) into oracle/_ref/.  It is used to write golden vectors and to time
// the reference CPU build as a baseline; nothing in the product links or calls it.
//
// It builds the objects the reference's TOML front end would build, without toml11 (absent
// here): System (mjolnir/core/System.hpp:43), ForceField (mjolnir/core/ForceField.hpp:32),
// GlobalPairInteraction with LorentzBerthelotRule / ExcludedVolumeParameterList /
// DebyeHuckelParameterList and a CellList / VerletList / Naive partition
// (mjolnir/input/read_global_interaction.hpp, read_spatial_partition.hpp:62-116), the local
// BondLength / BondAngle / DihedralAngle / Contact interactions
// (mjolnir/input/read_local_interaction.hpp), a ConstraintForceField
// (mjolnir/input/read_constraint.hpp), and BAOABLangevinIntegrator / gBAOABLangevinIntegrator /
// GJFNVTLangevinIntegrator / UnderdampedLangevinIntegrator (mjolnir/input/read_integrator.hpp).
// Units are set exactly as read_units does for "kcal/mol" and "angstrom"
// (mjolnir/input/read_units.hpp:32-92).
//
// Commands:
//   ref_driver forces <spec> <out>         per-interaction energies, forces and pair lists
//   ref_driver run    <spec> <nsteps> <out> [every]   integrate, dump x/v (every `every` steps)
//   ref_driver bench  <spec> <nsteps>      time the step loop only, print one JSON line
//   ref_driver save   <spec> <prefix>      MsgPack checkpoint of the initialized system
//   ref_driver sim    <spec> <total> <save> <prefix>   simulator loop with the energy, XYZ and
//                                          DCD observers (prefix.ene, prefix_position.xyz, ...)
#include <mjolnir/core/SimulatorTraits.hpp>
#include <mjolnir/core/BoundaryCondition.hpp>
#include <mjolnir/core/System.hpp>
#include <mjolnir/core/Unit.hpp>
#include <mjolnir/core/ForceField.hpp>
#include <mjolnir/core/RandomNumberGenerator.hpp>
#include <mjolnir/core/BAOABLangevinIntegrator.hpp>
#include <mjolnir/core/GJFNVTLangevinIntegrator.hpp>
#include <mjolnir/core/UnderdampedLangevinIntegrator.hpp>
#include <mjolnir/core/gBAOABLangevinIntegrator.hpp>
#include <mjolnir/core/SystemMotionRemover.hpp>
#include <mjolnir/core/EnergyObserver.hpp>
#include <mjolnir/core/XYZObserver.hpp>
#include <mjolnir/core/DCDObserver.hpp>
#include <mjolnir/core/MsgPackSaver.hpp>
#include <mjolnir/core/PeriodicGridCellList.hpp>
#include <mjolnir/core/UnlimitedGridCellList.hpp>
#include <mjolnir/core/VerletList.hpp>
#include <mjolnir/core/NaivePairCalculation.hpp>
#include <mjolnir/forcefield/global/GlobalPairInteraction.hpp>
#include <mjolnir/forcefield/global/GlobalPairLennardJonesInteraction.hpp>
#include <mjolnir/forcefield/global/GlobalPairExcludedVolumeInteraction.hpp>
#include <mjolnir/forcefield/global/LennardJonesPotential.hpp>
#include <mjolnir/forcefield/global/ExcludedVolumePotential.hpp>
#include <mjolnir/forcefield/global/DebyeHuckelPotential.hpp>
#include <mjolnir/forcefield/local/BondLengthInteraction.hpp>
#include <mjolnir/forcefield/local/BondAngleInteraction.hpp>
#include <mjolnir/forcefield/local/DihedralAngleInteraction.hpp>
#include <mjolnir/forcefield/local/ContactInteraction.hpp>
#include <mjolnir/forcefield/local/GoContactInteraction.hpp>
#include <mjolnir/forcefield/local/HarmonicPotential.hpp>
#include <mjolnir/forcefield/local/GaussianPotential.hpp>
#include <mjolnir/forcefield/local/ClementiDihedralPotential.hpp>
#include <mjolnir/forcefield/local/GoContactPotential.hpp>
#include <mjolnir/forcefield/local/PeriodicGaussianPotential.hpp>
#include <mjolnir/forcefield/3SPN2/ThreeSPN2BondPotential.hpp>
#include <mjolnir/forcefield/3SPN2/ThreeSPN2BaseStackingInteraction.hpp>
#include <mjolnir/forcefield/3SPN2/ThreeSPN2BasePairInteraction.hpp>
#include <mjolnir/forcefield/3SPN2/ThreeSPN2CrossStackingInteraction.hpp>
#include <mjolnir/forcefield/3SPN2/ThreeSPN2ExcludedVolumePotential.hpp>
#ifdef MJOLNIR_WITH_OPENMP
#include <mjolnir/omp/omp.hpp>
#endif

#include <chrono>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace drv
{
using namespace mjolnir;

struct LocalIn
{
    std::string interaction, potential, topology;
    double margin;
    std::vector<std::vector<std::size_t>> idx;
    std::vector<std::vector<double>> par;
};
struct GlobalIn
{
    std::string potential, partition, ignore_mol;
    double cutoff, epsilon, margin;
    std::map<std::string, std::size_t> within;
    std::vector<std::size_t> participants;
    std::vector<std::vector<double>> par;
};
struct SpecIn
{
    bool periodic = false;
    double lower[3] = {0, 0, 0}, upper[3] = {0, 0, 0};
    std::map<std::string, double> attributes;
    std::vector<double> mass, x, v;  // x, v: 3N
    std::vector<LocalIn> local;
    std::vector<GlobalIn> global;
    bool has_integrator = false;
    bool rm_translation = false, rm_rotation = false, rm_rescale = false;
    std::string integrator;
    // constraint block (read_constraint.hpp:12-46)
    bool has_constraint = false;
    std::string con_topology;
    std::size_t con_max_iteration = 0;
    double con_tolerance = 0;
    std::vector<std::array<std::size_t, 2>> con_idx;
    std::vector<double> con_v0;
    double dt = 0;
    std::uint32_t seed = 0;
    std::vector<double> coefs;
};

static SpecIn read_spec(const std::string& path)
{
    std::ifstream ifs(path);
    if(!ifs) {throw std::runtime_error("cannot open " + path);}
    SpecIn s;
    std::string tok;
    ifs >> tok; int ver; ifs >> ver;
    if(tok != "mjspec") {throw std::runtime_error("not an mjspec file");}
    while(ifs >> tok)
    {
        if(tok == "end") {break;}
        if(tok == "boundary")
        {
            std::string kind; ifs >> kind;
            s.periodic = (kind == "periodic");
            if(s.periodic)
            {
                ifs >> s.lower[0] >> s.lower[1] >> s.lower[2] >> s.upper[0] >> s.upper[1] >> s.upper[2];
            }
        }
        else if(tok == "attribute")
        {
            std::string k; double val; ifs >> k >> val; s.attributes[k] = val;
        }
        else if(tok == "particles")
        {
            std::size_t n; ifs >> n;
            s.mass.resize(n); s.x.resize(3 * n); s.v.resize(3 * n);
            for(std::size_t i = 0; i < n; ++i)
            {
                ifs >> s.mass[i] >> s.x[3*i] >> s.x[3*i+1] >> s.x[3*i+2]
                    >> s.v[3*i] >> s.v[3*i+1] >> s.v[3*i+2];
            }
        }
        else if(tok == "local")
        {
            LocalIn l; std::size_t n;
            ifs >> l.interaction >> l.potential >> l.topology >> n >> l.margin;
            std::size_t arity = 2, npar = 2;
            if(l.interaction == "BondAngle") {arity = 3;}
            if(l.interaction == "DihedralAngle") {arity = 4;}
            if(l.potential == "Gaussian" || l.potential == "ClementiDihedral") {npar = 3;}
            // 3SPN2 base stacking: S5 B5 B3 P5 S3 P3 (absent phosphate 4294967295), stack kind
            if(l.interaction == "3SPN2BaseStacking") {arity = 6; npar = 1;}
            for(std::size_t t = 0; t < n; ++t)
            {
                std::vector<std::size_t> ix(arity); std::vector<double> p(npar);
                for(auto& i : ix) {ifs >> i;}
                for(auto& q : p)  {ifs >> q;}
                l.idx.push_back(ix); l.par.push_back(p);
            }
            s.local.push_back(l);
        }
        else if(tok == "global")
        {
            GlobalIn g; std::size_t n; std::string within;
            ifs >> g.potential >> n >> g.cutoff >> g.epsilon >> g.partition >> g.margin
                >> g.ignore_mol >> within;
            if(within != "-")
            {
                std::stringstream ss(within); std::string item;
                while(std::getline(ss, item, ','))
                {
                    const auto c = item.find(':');
                    g.within[item.substr(0, c)] = std::stoul(item.substr(c + 1));
                }
            }
            std::size_t npar = (g.potential == "LennardJones") ? 2 : 1;
            if(g.potential == "3SPN2BasePair") {npar = 2;}        // base, sugar
            if(g.potential == "3SPN2CrossStacking") {npar = 4;}   // base, sugar, 5' base, 3' base (-1: none)
            for(std::size_t t = 0; t < n; ++t)
            {
                std::size_t i; std::vector<double> p(npar);
                ifs >> i; for(auto& q : p) {ifs >> q;}
                g.participants.push_back(i); g.par.push_back(p);
            }
            s.global.push_back(g);
        }
        else if(tok == "integrator")
        {
            s.has_integrator = true;
            ifs >> s.integrator >> s.dt >> s.seed;
            s.coefs.resize(s.mass.size());
            for(auto& c : s.coefs) {ifs >> c;}
        }
        else if(tok == "constraint")
        {
            std::size_t n = 0;
            s.has_constraint = true;
            ifs >> s.con_topology >> n >> s.con_max_iteration >> s.con_tolerance;
            for(std::size_t k = 0; k < n; ++k)
            {
                std::array<std::size_t, 2> ij; double v0;
                ifs >> ij[0] >> ij[1] >> v0;
                s.con_idx.push_back(ij); s.con_v0.push_back(v0);
            }
        }
        else if(tok == "remover")
        {
            int t = 0, r = 0, c = 0;
            ifs >> t >> r >> c;
            s.rm_translation = t != 0; s.rm_rotation = r != 0; s.rm_rescale = c != 0;
        }
        else {throw std::runtime_error("unknown token " + tok);}
    }
    return s;
}

// read_units for energy = "kcal/mol", length = "angstrom" (read_units.hpp:32-92).
template<typename realT>
static void set_units_kcal_angstrom()
{
    using phys_type = physics::constants<realT>;
    using unit_type = unit::constants<realT>;
    phys_type::set_kB(phys_type::kB() * (unit_type::J_to_cal() / 1000.0) * unit_type::avogadro_constant());
    phys_type::set_eps0(phys_type::eps0() * (1000.0 / unit_type::J_to_cal()) / unit_type::avogadro_constant());
    phys_type::set_energy_unit("kcal/mol");
    phys_type::set_eps0(phys_type::eps0() / unit_type::m_to_angstrom());
    phys_type::set_m_to_length(unit_type::m_to_angstrom());
    phys_type::set_length_to_m(unit_type::angstrom_to_m());
    phys_type::set_L_to_volume(1e-3 * std::pow(unit_type::m_to_angstrom(), 3));
    phys_type::set_volume_to_L(1e+3 * std::pow(unit_type::angstrom_to_m(), 3));
    phys_type::set_length_unit("angstrom");
}

template<typename traitsT>
struct Built
{
    using real_type = typename traitsT::real_type;
    using coord     = typename traitsT::coordinate_type;
    using system_type = System<traitsT>;
    using ff_type     = ForceField<traitsT>;

    std::unique_ptr<system_type> sys;
    std::unique_ptr<ForceFieldBase<traitsT>> ff;
    // pair-list dumpers, one per global interaction, in input order
    std::vector<std::function<void(std::ostream&)>> dump_pairs;
};

template<typename traitsT, typename potT, typename partitionBase>
static std::unique_ptr<partitionBase> make_partition(const GlobalIn& g)
{
    using boundary_type = typename traitsT::boundary_type;
    if(g.partition == "CellList")
    {
        if(is_cuboidal_periodic_boundary<boundary_type>::value)
        {
            return std::unique_ptr<partitionBase>(new PeriodicGridCellList<traitsT, potT>(g.margin));
        }
        return std::unique_ptr<partitionBase>(new UnlimitedGridCellList<traitsT, potT>(g.margin));
    }
    if(g.partition == "VerletList")
    {
        return std::unique_ptr<partitionBase>(new VerletList<traitsT, potT>(g.margin));
    }
    if(g.partition == "Naive")
    {
        return std::unique_ptr<partitionBase>(new NaivePairCalculation<traitsT, potT>());
    }
    throw std::runtime_error("unknown partition " + g.partition);
}

#ifdef MJOLNIR_WITH_OPENMP
// the OpenMP specialisations keep their partition private; the timing build never dumps lists
template<typename interactionT>
static std::function<void(std::ostream&)> pair_dumper(const interactionT*, std::vector<std::size_t>)
{
    return [](std::ostream& os) {os << "pairs 0\n";};
}
#else
template<typename interactionT>
static std::function<void(std::ostream&)> pair_dumper(const interactionT* inter, std::vector<std::size_t> participants)
{
    return [inter, participants](std::ostream& os) {
        const auto& nl = inter->partition().neighbors();
        std::vector<std::pair<std::size_t, std::size_t>> pairs;
        for(std::size_t k = 0; k + 1 < participants.size(); ++k) // leading participants
        {
            const auto i = participants[k];
            for(const auto& ptnr : inter->partition().partners(i))
            {
                pairs.emplace_back(i, ptnr.index);
            }
        }
        (void)nl;
        os << "pairs " << pairs.size() << "\n";
        for(const auto& p : pairs) {os << p.first << ' ' << p.second << "\n";}
    };
}
#endif

template<typename traitsT>
static Built<traitsT> build(const SpecIn& s)
{
    using real_type   = typename traitsT::real_type;
    using coord       = typename traitsT::coordinate_type;
    using boundary_t  = typename traitsT::boundary_type;
    Built<traitsT> b;

    boundary_t bdry;
    if(s.periodic)
    {
        bdry.set_boundary(coord(s.lower[0], s.lower[1], s.lower[2]), coord(s.upper[0], s.upper[1], s.upper[2]));
    }
    const std::size_t N = s.mass.size();
    b.sys.reset(new System<traitsT>(N, bdry));
    auto& sys = *b.sys;
    for(const auto& kv : s.attributes) {sys.attribute(kv.first) = kv.second;}
    for(std::size_t i = 0; i < N; ++i)
    {
        sys.mass(i)     = s.mass[i];
        sys.rmass(i)    = 1.0 / s.mass[i];
        sys.position(i) = coord(s.x[3*i], s.x[3*i+1], s.x[3*i+2]);
        sys.velocity(i) = coord(s.v[3*i], s.v[3*i+1], s.v[3*i+2]);
        sys.force(i)    = coord(0, 0, 0);
        sys.name(i)     = "X";
        sys.group(i)    = "NONE";
    }
    sys.velocity_initialized() = true;

    LocalForceField<traitsT> local;
    for(const auto& l : s.local)
    {
        if(l.interaction == "BondLength" && l.potential == "Harmonic")
        {
            using pot = HarmonicPotential<real_type>;
            using I = BondLengthInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1]}}, pot(l.par[t][0], l.par[t][1]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "BondLength" && l.potential == "GoContact")
        {
            using pot = GoContactPotential<real_type>;
            using I = BondLengthInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1]}}, pot(l.par[t][0], l.par[t][1]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "Contact" && l.potential == "GoContact")
        {
            using pot = GoContactPotential<real_type>;
            using I = ContactInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1]}}, pot(l.par[t][0], l.par[t][1]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c), l.margin)));
        }
        else if(l.interaction == "BondAngle" && l.potential == "Gaussian")
        {
            using pot = GaussianPotential<real_type>;
            using I = BondAngleInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1], l.idx[t][2]}},
                               pot(l.par[t][0], l.par[t][1], l.par[t][2]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "BondAngle" && l.potential == "Harmonic")
        {
            using pot = HarmonicPotential<real_type>;
            using I = BondAngleInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1], l.idx[t][2]}},
                               pot(l.par[t][0], l.par[t][1]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "DihedralAngle" && l.potential == "ClementiDihedral")
        {
            using pot = ClementiDihedralPotential<real_type>;
            using I = DihedralAngleInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1], l.idx[t][2], l.idx[t][3]}},
                               pot(l.par[t][0], l.par[t][1], l.par[t][2]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "BondLength" && l.potential == "3SPN2Bond")
        {
            using pot = ThreeSPN2BondPotential<real_type>;
            using I = BondLengthInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1]}}, pot(l.par[t][0], l.par[t][1]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "DihedralAngle" && l.potential == "Gaussian")
        {
            // read_local_interaction.hpp:282-290: a Gaussian dihedral is PeriodicGaussianPotential
            using pot = PeriodicGaussianPotential<real_type>;
            using I = DihedralAngleInteraction<traitsT, pot>;
            typename I::container_type c;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                c.emplace_back(typename I::indices_type{{l.idx[t][0], l.idx[t][1], l.idx[t][2], l.idx[t][3]}},
                               pot(l.par[t][0], l.par[t][1], l.par[t][2]));
            }
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(new I(l.topology, std::move(c))));
        }
        else if(l.interaction == "3SPN2BaseStacking")
        {
            using I = ThreeSPN2BaseStackingInteraction<traitsT>;
            using nuc_t = parameter_3SPN2::NucleotideInfo;
            using bsk = parameter_3SPN2::base_stack_kind;
            const std::size_t nil_in = 4294967295u;
            typename I::container_type c;
            std::vector<nuc_t> nucs;
            for(std::size_t t = 0; t < l.idx.size(); ++t)
            {
                const auto& ix = l.idx[t];
                c.emplace_back(typename I::indices_type{{ix[0], ix[1], ix[2]}}, static_cast<bsk>((int)l.par[t][0]));
                nuc_t n5, n3;
                n5.strand = 0; n5.S = ix[0]; n5.B = ix[1]; n5.P = (ix[3] == nil_in) ? nuc_t::nil() : ix[3];
                n3.strand = 0; n3.S = ix[4]; n3.B = ix[2]; n3.P = (ix[5] == nil_in) ? nuc_t::nil() : ix[5];
                nucs.push_back(n5); nucs.push_back(n3);
            }
            ThreeSPN2BaseStackingPotential<real_type> pot{ThreeSPN2BaseStackingPotentialParameter<real_type>{}};
            local.emplace(std::unique_ptr<LocalInteractionBase<traitsT>>(
                new I(l.topology, std::move(c), std::move(pot), std::move(nucs))));
        }
        else
        {
            throw std::runtime_error("unsupported local " + l.interaction + ":" + l.potential);
        }
    }

    GlobalForceField<traitsT> global;
    for(const auto& g : s.global)
    {
        IgnoreMolecule<typename Topology::molecule_id_type> ign_mol(g.ignore_mol);
        IgnoreGroup<typename Topology::group_id_type> ign_grp({});
        if(g.potential == "LennardJones")
        {
            using pot = LennardJonesPotential<real_type>;
            using I = GlobalPairInteraction<traitsT, pot>;
            std::vector<std::pair<std::size_t, typename pot::parameter_type>> ps;
            for(std::size_t t = 0; t < g.participants.size(); ++t)
            {
                ps.emplace_back(g.participants[t], typename pot::parameter_type{g.par[t][0], g.par[t][1]});
            }
            ParameterList<traitsT, pot> pl(make_unique<LorentzBerthelotRule<traitsT, pot>>(ps, g.within, ign_mol, ign_grp));
            SpatialPartition<traitsT, pot> part(make_partition<traitsT, pot, SpatialPartitionBase<traitsT, pot>>(g));
            auto* inter = new I(pot(g.cutoff), std::move(pl), std::move(part));
            b.dump_pairs.push_back(pair_dumper(inter, g.participants));
            global.emplace(std::unique_ptr<GlobalInteractionBase<traitsT>>(inter));
        }
        else if(g.potential == "ExcludedVolume")
        {
            using pot = ExcludedVolumePotential<real_type>;
            using I = GlobalPairInteraction<traitsT, pot>;
            std::vector<std::pair<std::size_t, typename pot::parameter_type>> ps;
            for(std::size_t t = 0; t < g.participants.size(); ++t)
            {
                ps.emplace_back(g.participants[t], typename pot::parameter_type{g.par[t][0]});
            }
            ParameterList<traitsT, pot> pl(make_unique<ExcludedVolumeParameterList<traitsT>>(ps, g.within, ign_mol, ign_grp));
            SpatialPartition<traitsT, pot> part(make_partition<traitsT, pot, SpatialPartitionBase<traitsT, pot>>(g));
            auto* inter = new I(pot(g.cutoff, g.epsilon), std::move(pl), std::move(part));
            b.dump_pairs.push_back(pair_dumper(inter, g.participants));
            global.emplace(std::unique_ptr<GlobalInteractionBase<traitsT>>(inter));
        }
        else if(g.potential == "DebyeHuckel")
        {
            using pot = DebyeHuckelPotential<real_type>;
            using PL  = DebyeHuckelParameterList<traitsT>;
            using I = GlobalPairInteraction<traitsT, pot>;
            std::vector<std::pair<std::size_t, typename PL::parameter_type>> ps;
            for(std::size_t t = 0; t < g.participants.size(); ++t)
            {
                ps.emplace_back(g.participants[t], typename PL::parameter_type{g.par[t][0]});
            }
            ParameterList<traitsT, pot> pl(make_unique<PL>(ps, g.within, ign_mol, ign_grp));
            SpatialPartition<traitsT, pot> part(make_partition<traitsT, pot, SpatialPartitionBase<traitsT, pot>>(g));
            auto* inter = new I(pot(g.cutoff), std::move(pl), std::move(part));
            b.dump_pairs.push_back(pair_dumper(inter, g.participants));
            global.emplace(std::unique_ptr<GlobalInteractionBase<traitsT>>(inter));
        }
        else if(g.potential == "3SPN2ExcludedVolume")
        {
            using pot = ThreeSPN2ExcludedVolumePotential<real_type>;
            using PL  = ThreeSPN2ExcludedVolumeParameterList<traitsT>;
            using I = GlobalPairInteraction<traitsT, pot>;
            std::vector<std::pair<std::size_t, typename PL::parameter_type>> ps;
            for(std::size_t t = 0; t < g.participants.size(); ++t)
            {
                ps.emplace_back(g.participants[t], static_cast<parameter_3SPN2::bead_kind>((int)g.par[t][0]));
            }
            ParameterList<traitsT, pot> pl(make_unique<PL>(ThreeSPN2ExcludedVolumePotentialParameter<real_type>{},
                                                           ps, g.within, ign_mol, ign_grp));
            SpatialPartition<traitsT, pot> part(make_partition<traitsT, pot, SpatialPartitionBase<traitsT, pot>>(g));
            auto* inter = new I(pot(), std::move(pl), std::move(part));
            b.dump_pairs.push_back(pair_dumper(inter, g.participants));
            global.emplace(std::unique_ptr<GlobalInteractionBase<traitsT>>(inter));
        }
        else if(g.potential == "3SPN2BasePair")
        {
            using I = ThreeSPN2BasePairInteraction<traitsT>;
            using pot = typename I::potential_type;
            using PL = typename I::parameter_list_type;
            std::vector<std::pair<std::size_t, typename PL::parameter_type>> ps;
            for(std::size_t t = 0; t < g.participants.size(); ++t)
            {
                typename PL::parameter_type p;
                p.base = static_cast<parameter_3SPN2::base_kind>((int)g.par[t][0]);
                p.S_idx = static_cast<std::size_t>(g.par[t][1]);
                ps.emplace_back(g.participants[t], p);
            }
            PL pl(ThreeSPN2BasePairGlobalPotentialParameter<real_type>{}, ps, g.within, ign_mol, ign_grp);
            typename I::partition_type part(make_partition<traitsT, pot, SpatialPartitionBase<traitsT, pot>>(g));
            auto* inter = new I(pot{}, std::move(pl), std::move(part));
            b.dump_pairs.push_back([](std::ostream&) {});   // no partition accessor on this class
            global.emplace(std::unique_ptr<GlobalInteractionBase<traitsT>>(inter));
        }
        else if(g.potential == "3SPN2CrossStacking")
        {
            using I = ThreeSPN2CrossStackingInteraction<traitsT>;
            using pot = typename I::potential_type;
            using PL = typename I::parameter_list_type;
            std::vector<std::pair<std::size_t, typename PL::parameter_type>> ps;
            for(std::size_t t = 0; t < g.participants.size(); ++t)
            {
                typename PL::parameter_type p;
                p.base = static_cast<parameter_3SPN2::base_kind>((int)g.par[t][0]);
                p.S = static_cast<std::size_t>(g.par[t][1]);
                p.B5 = (g.par[t][2] < 0) ? pot::invalid() : static_cast<std::size_t>(g.par[t][2]);
                p.B3 = (g.par[t][3] < 0) ? pot::invalid() : static_cast<std::size_t>(g.par[t][3]);
                ps.emplace_back(g.participants[t], p);
            }
            PL pl(ThreeSPN2CrossStackingGlobalPotentialParameter<real_type>{}, ps, g.within, ign_mol, ign_grp);
            typename I::partition_type part(make_partition<traitsT, pot, SpatialPartitionBase<traitsT, pot>>(g));
            auto* inter = new I(pot{}, std::move(pl), std::move(part));
            b.dump_pairs.push_back([](std::ostream&) {});
            global.emplace(std::unique_ptr<GlobalInteractionBase<traitsT>>(inter));
        }
        else
        {
            throw std::runtime_error("unsupported global " + g.potential);
        }
    }
    ConstraintForceField<traitsT> constraint;
    if(s.has_constraint)
    {
        typename ConstraintForceField<traitsT>::container_type cs;
        for(std::size_t k = 0; k < s.con_idx.size(); ++k) {cs.emplace_back(s.con_idx[k], s.con_v0[k]);}
        constraint = ConstraintForceField<traitsT>(s.con_topology, std::move(cs), s.con_max_iteration,
                                                   s.con_tolerance);
    }
    b.ff.reset(new ForceField<traitsT>(std::move(local), std::move(global),
                                        ExternalForceField<traitsT>{}, std::move(constraint)));
    return b;
}

static void write_vec3(std::ostream& os, const char* tag, const std::vector<double>& v)
{
    const std::size_t n = v.size() / 3;
    os << tag << ' ' << n << "\n";
    char buf[128];
    for(std::size_t i = 0; i < n; ++i)
    {
        std::snprintf(buf, sizeof(buf), "%.17g %.17g %.17g\n", v[3*i], v[3*i+1], v[3*i+2]);
        os << buf;
    }
}

template<typename traitsT>
static std::vector<double> forces_of(const System<traitsT>& sys)
{
    std::vector<double> f(3 * sys.size());
    for(std::size_t i = 0; i < sys.size(); ++i)
    {
        f[3*i] = math::X(sys.force(i)); f[3*i+1] = math::Y(sys.force(i)); f[3*i+2] = math::Z(sys.force(i));
    }
    return f;
}

template<typename traitsT>
static void zero_forces(System<traitsT>& sys)
{
    using coord = typename traitsT::coordinate_type;
    for(std::size_t i = 0; i < sys.size(); ++i) {sys.force(i) = coord(0, 0, 0);}
}

// forces: initialise the force field, then per interaction (local in order, then global in
// order) the energy from calc_energy (what EnergyObserver prints, EnergyObserver.hpp:89-150)
// and the force from calc_force with zeroed force arrays; then the total force from
// ForceField::calc_force and the total from calc_force_and_energy.
template<typename traitsT>
static int cmd_forces(const SpecIn& s, const std::string& out)
{
    auto b = build<traitsT>(s);
    auto& sys = *b.sys;
    auto& ff  = *b.ff;
    auto& ffc = dynamic_cast<ForceField<traitsT>&>(ff);
    sys.initialize(*(new RandomNumberGenerator<traitsT>(1)));
    ff.initialize(sys);

    std::ofstream os(out);
    char buf[128];
    for(auto& inter : ffc.local())
    {
        zero_forces(sys);
        const double e = inter->calc_energy(sys);
        inter->calc_force(sys);
        std::snprintf(buf, sizeof(buf), "%.17g", e);
        os << "interaction " << inter->name() << " energy " << buf << "\n";
        write_vec3(os, "force", forces_of(sys));
    }
    std::size_t gidx = 0;
    for(auto& inter : ffc.global())
    {
        zero_forces(sys);
        const double e = inter->calc_energy(sys);
        inter->calc_force(sys);
        std::snprintf(buf, sizeof(buf), "%.17g", e);
        os << "interaction " << inter->name() << " energy " << buf << "\n";
        write_vec3(os, "force", forces_of(sys));
        b.dump_pairs.at(gidx++)(os);
    }
    zero_forces(sys);
    const double etot = ff.calc_force_and_energy(sys);
    std::snprintf(buf, sizeof(buf), "%.17g", etot);
    os << "total energy " << buf << "\n";
    write_vec3(os, "force", forces_of(sys));
    os << "end\n";
    return 0;
}

// save: the system after ForceField::initialize and one calc_force, written by MsgPackSaver
// (MsgPackSaver.hpp:98-186) to <prefix>_system.msg
template<typename traitsT>
static int cmd_save(const SpecIn& s, const std::string& prefix)
{
    auto b = build<traitsT>(s);
    auto& sys = *b.sys;
    sys.initialize(*(new RandomNumberGenerator<traitsT>(1)));
    b.ff->initialize(sys);
    zero_forces(sys);
    b.ff->calc_force(sys);
    MsgPackSaver<traitsT> saver(prefix);
    saver.save(sys);
    return 0;
}

template<typename traitsT, typename integT>
static int run_impl(const SpecIn& s, std::size_t nsteps, const std::string& out, std::size_t every, bool bench)
{
    using real_type = typename traitsT::real_type;
    auto b = build<traitsT>(s);
    auto& sys = *b.sys;
    RandomNumberGenerator<traitsT> rng(s.seed);
    sys.initialize(rng);
    b.ff->initialize(sys);
    std::vector<real_type> coefs(s.coefs.begin(), s.coefs.end());
    integT integ(s.dt, std::move(coefs), SystemMotionRemover<traitsT>(s.rm_translation, s.rm_rotation, s.rm_rescale));
    integ.initialize(sys, b.ff, rng);

    std::ofstream os;
    if(!bench) {os.open(out);}
    auto dump = [&](std::size_t step) {
        std::vector<double> x(3 * sys.size()), v(3 * sys.size());
        for(std::size_t i = 0; i < sys.size(); ++i)
        {
            x[3*i] = math::X(sys.position(i)); x[3*i+1] = math::Y(sys.position(i)); x[3*i+2] = math::Z(sys.position(i));
            v[3*i] = math::X(sys.velocity(i)); v[3*i+1] = math::Y(sys.velocity(i)); v[3*i+2] = math::Z(sys.velocity(i));
        }
        os << "step " << step << "\n";
        write_vec3(os, "position", x);
        write_vec3(os, "velocity", v);
    };
    real_type t = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for(std::size_t step = 0; step < nsteps; ++step)
    {
        if(!bench && every != 0 && step % every == 0) {dump(step);}
        t = integ.step(t, sys, b.ff, rng);
    }
    const auto t1 = std::chrono::steady_clock::now();
    if(bench)
    {
        const double sec = std::chrono::duration<double>(t1 - t0).count();
        int threads = 1;
#ifdef MJOLNIR_WITH_OPENMP
        threads = omp_get_max_threads();
#endif
        std::printf("{\"seconds\": %.9g, \"steps\": %zu, \"particles\": %zu, \"threads\": %d, "
                    "\"particle_steps_per_s\": %.9g}\n", sec, nsteps, sys.size(), threads,
                    double(sys.size()) * double(nsteps) / sec);
        return 0;
    }
    dump(nsteps);
    os << "end\n";
    return 0;
}

// MolecularDynamicsSimulator::initialize/step/finalize (MolecularDynamicsSimulator.hpp:72-123)
// with the EnergyObserver, XYZObserver and DCDObserver attached in that order: prefix.ene,
// prefix_{position,velocity}.{xyz,dcd} after `total` steps, output every `save`
template<typename traitsT, typename integT>
static int sim_impl(const SpecIn& s, std::size_t total, std::size_t save, const std::string& prefix)
{
    using real_type = typename traitsT::real_type;
    auto b = build<traitsT>(s);
    auto& sys = *b.sys;
    RandomNumberGenerator<traitsT> rng(s.seed);
    sys.initialize(rng);
    b.ff->initialize(sys);
    std::vector<real_type> coefs(s.coefs.begin(), s.coefs.end());
    integT integ(s.dt, std::move(coefs), SystemMotionRemover<traitsT>(s.rm_translation, s.rm_rotation, s.rm_rescale));
    integ.initialize(sys, b.ff, rng);
    std::vector<std::unique_ptr<ObserverBase<traitsT>>> obs;
    obs.emplace_back(new EnergyObserver<traitsT>(prefix));
    obs.emplace_back(new XYZObserver<traitsT>(prefix));
    obs.emplace_back(new DCDObserver<traitsT>(prefix));
    for(auto& o : obs) {o->initialize(total, save, integ.delta_t(), sys, b.ff);}
    real_type t = 0;
    for(std::size_t step = 0; step < total; ++step)
    {
        if(step % save == 0) {for(auto& o : obs) {o->output(step, integ.delta_t(), sys, b.ff);}}
        t = integ.step(t, sys, b.ff, rng);
    }
    for(auto& o : obs) {o->output(total, integ.delta_t(), sys, b.ff);}
    for(auto& o : obs) {o->finalize(total, integ.delta_t(), sys, b.ff);}
    return 0;
}

template<typename traitsT>
static int cmd_sim(const SpecIn& s, std::size_t total, std::size_t save, const std::string& prefix)
{
    if(s.integrator == "BAOABLangevin") {return sim_impl<traitsT, BAOABLangevinIntegrator<traitsT>>(s, total, save, prefix);}
    if(s.integrator == "GJFNVTLangevin") {return sim_impl<traitsT, GJFNVTLangevinIntegrator<traitsT>>(s, total, save, prefix);}
    if(s.integrator == "gBAOABLangevin") {return sim_impl<traitsT, gBAOABLangevinIntegrator<traitsT>>(s, total, save, prefix);}
    if(s.integrator == "UnderdampedLangevin")
    {
        return sim_impl<traitsT, UnderdampedLangevinIntegrator<traitsT>>(s, total, save, prefix);
    }
    throw std::runtime_error("unsupported integrator " + s.integrator);
}

template<typename traitsT>
static int cmd_run(const SpecIn& s, std::size_t nsteps, const std::string& out, std::size_t every, bool bench)
{
    if(!s.has_integrator) {throw std::runtime_error("spec has no integrator");}
    if(s.integrator == "BAOABLangevin")
    {
        return run_impl<traitsT, BAOABLangevinIntegrator<traitsT>>(s, nsteps, out, every, bench);
    }
    if(s.integrator == "GJFNVTLangevin")
    {
        return run_impl<traitsT, GJFNVTLangevinIntegrator<traitsT>>(s, nsteps, out, every, bench);
    }
    if(s.integrator == "gBAOABLangevin")
    {
        return run_impl<traitsT, gBAOABLangevinIntegrator<traitsT>>(s, nsteps, out, every, bench);
    }
    if(s.integrator == "UnderdampedLangevin")
    {
        return run_impl<traitsT, UnderdampedLangevinIntegrator<traitsT>>(s, nsteps, out, every, bench);
    }
    throw std::runtime_error("unsupported integrator " + s.integrator);
}

#ifdef MJOLNIR_WITH_OPENMP
template<typename realT, template<typename, typename> class B>
using Traits = OpenMPSimulatorTraits<realT, B>;
#else
template<typename realT, template<typename, typename> class B>
using Traits = SimulatorTraits<realT, B>;
#endif

template<template<typename, typename> class B>
static int dispatch(const std::string& cmd, const SpecIn& s, int argc, char** argv)
{
    using traits = Traits<double, B>;
    if(cmd == "forces") {return cmd_forces<traits>(s, argv[3]);}
    if(cmd == "run")
    {
        const std::size_t every = (argc > 5) ? std::stoul(argv[5]) : 0;
        return cmd_run<traits>(s, std::stoul(argv[3]), argv[4], every, false);
    }
    if(cmd == "bench") {return cmd_run<traits>(s, std::stoul(argv[3]), "", 0, true);}
    if(cmd == "sim") {return cmd_sim<traits>(s, std::stoul(argv[3]), std::stoul(argv[4]), argv[5]);}
    if(cmd == "save") {return cmd_save<traits>(s, argv[3]);}
    throw std::runtime_error("unknown command " + cmd);
}
} // drv

#ifndef REF_DRIVER_NO_MAIN
int main(int argc, char** argv)
{
    if(argc < 4)
    {
        std::cerr << "usage: ref_driver forces|run|bench <spec> ...\n";
        return 2;
    }
    try
    {
        mjolnir::LoggerManager::set_default_logger("/dev/null");
        drv::set_units_kcal_angstrom<double>();
        const auto spec = drv::read_spec(argv[2]);
        if(spec.periodic) {return drv::dispatch<mjolnir::CuboidalPeriodicBoundary>(argv[1], spec, argc, argv);}
        return drv::dispatch<mjolnir::UnlimitedBoundary>(argv[1], spec, argc, argv);
    }
    catch(const std::exception& e)
    {
        std::cerr << "ref_driver: " << e.what() << std::endl;
        return 1;
    }
}
#endif // REF_DRIVER_NO_MAIN
