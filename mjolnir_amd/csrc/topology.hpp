// Host-side topology and exclusion lists (set-up only, never on the step path).
//
// Topology::add_connection (Topology.hpp:161-191) for every local interaction that writes a
// connection kind, in the order the force field writes them (ForceField.hpp:53-56);
// construct_molecules (Topology.hpp:280-317), where only "bond" edges join molecules; and
// ExclusionList::make (ExclusionList.hpp:86-234): per particle, itself plus everything
// reachable within `dist` edges of each ignored kind (list_adjacent_within,
// Topology.hpp:133-153,243-264), sorted and unique.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace mjg
{

struct Topo
{
    std::vector<std::string> kinds;
    std::vector<std::vector<std::pair<uint32_t, int>>> adj;
    std::vector<uint32_t> mol;

    int kind_id(const std::string& name)
    {
        for(size_t i = 0; i < kinds.size(); ++i) {if(kinds[i] == name) {return (int)i;}}
        kinds.push_back(name);
        return (int)kinds.size() - 1;
    }

    void reset(size_t n) {adj.assign(n, {}); mol.clear();}

    void add_connection(uint32_t i, uint32_t j, int kind)
    {
        for(int s = 0; s < 2; ++s)
        {
            const uint32_t a = s ? j : i, b = s ? i : j;
            auto& v = adj[a];
            if(std::find(v.begin(), v.end(), std::make_pair(b, kind)) == v.end()) {v.emplace_back(b, kind);}
        }
    }

    // connect every pair of bodies of each term (BondLength: 1 pair, BondAngle: 3, Dihedral: 6)
    void add_terms(const std::vector<uint32_t>& idx, int arity, int kind)
    {
        if(kind < 0) {return;}
        for(size_t t = 0; t + arity <= idx.size(); t += arity)
            for(int a = 0; a < arity; ++a)
                for(int b = a + 1; b < arity; ++b) {add_connection(idx[t + a], idx[t + b], kind);}
    }

    void construct_molecules()
    {
        const int bond = kind_id("bond");
        const uint32_t UNINIT = 0xffffffffu;
        mol.assign(adj.size(), UNINIT);
        uint32_t next = 0;
        for(size_t i = 0; i < adj.size(); ++i)
        {
            for(const auto& e : adj[i])
            {
                if(e.second != bond) {continue;}
                if(mol[e.first] != UNINIT) {mol[i] = mol[e.first]; break;}
            }
            if(mol[i] == UNINIT) {mol[i] = next++;}
        }
    }

    // exclusion CSR for ignore.particles_within = {(kind, dist)}
    void exclusions(const std::vector<std::pair<int, uint32_t>>& within,
                    std::vector<uint32_t>& start, std::vector<uint32_t>& all) const
    {
        const size_t n = adj.size();
        start.assign(n + 1, 0);
        all.clear();
        std::vector<uint32_t> tmp;
        std::vector<std::pair<uint32_t, uint32_t>> stack;
        for(size_t i = 0; i < n; ++i)
        {
            start[i] = (uint32_t)all.size();
            tmp.clear();
            tmp.push_back((uint32_t)i);
            for(const auto& w : within)
            {
                stack.clear();
                stack.emplace_back((uint32_t)i, w.second);
                while(!stack.empty())
                {
                    const auto top = stack.back();
                    stack.pop_back();
                    if(top.second == 0) {continue;}
                    for(const auto& e : adj[top.first])
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
        start[n] = (uint32_t)all.size();
    }
};

// physical constants after read_units("kcal/mol", "angstrom") (read_units.hpp:32-92,
// Unit.hpp:23-44,150-159), formed in the same operation order
inline double kB_kcal()
{
    const double J_to_cal = 1.0 / 4.184;
    return 1.38064852e-23 * (J_to_cal / 1000.0) * 6.022140857e23;
}
inline double eps0_kcal_angstrom()
{
    const double J_to_cal = 1.0 / 4.184;
    double eps0 = (8.854187817e-12 / 1.6021766208e-19) / 1.6021766208e-19;
    eps0 = eps0 * (1000.0 / J_to_cal) / 6.022140857e23;
    return eps0 / 1e+10;
}
inline double L_to_volume_angstrom() {return 1e-3 * std::pow(1e+10, 3);}

// DebyeHuckelPotential::calc_parameters (DebyeHuckelPotential.hpp:126-174)
struct DHParams
{
    double inv4pieps, debye, inv_debye, abs_cutoff, coef_at_cutoff;
};
inline DHParams debye_huckel_params(double T, double c, double cutoff_ratio)
{
    const double pi = 3.14159265358979323846264338;
    const double kB = kB_kcal(), NA = 6.022140857e23, eps0 = eps0_kcal_angstrom();
    const double epsk = (249.4 - 0.788 * T + 7.2e-4 * T * T) *
                        (1. - 2.551e-1 * c + 5.151e-2 * c * c - 6.889e-3 * c * c * c);
    DHParams p{};
    p.inv4pieps = 1.0 / (4 * pi * eps0 * epsk);
    const double I = c / L_to_volume_angstrom();
    p.debye = std::sqrt((eps0 * epsk * kB * T) / (2 * NA * I));
    p.inv_debye = 1. / p.debye;
    p.abs_cutoff = cutoff_ratio * p.debye;
    p.coef_at_cutoff = std::exp(-cutoff_ratio) / (p.debye * cutoff_ratio);
    return p;
}

} // mjg
