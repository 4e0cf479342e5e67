// Trajectory observers of the simulator loop: XYZObserver (XYZObserver.hpp:24-107) and
// DCDObserver (DCDObserver.hpp:27-341).  Host code: the state is downloaded once per output
// frame (positions and velocities, wrapped into the box as the engine keeps them) and written
// in the reference's file layouts, so a trajectory written here reads the same as one written
// by the reference.  Included once, by engine.hip.
#pragma once
#include "engine.hpp"

#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace mjg
{

namespace
{
template<typename T> void put(std::ostream& os, const T& v) {os.write(reinterpret_cast<const char*>(&v), sizeof(T));}

void truncate_or_throw(const std::string& fname, const char* who)
{
    std::ofstream ofs(fname);
    if(!ofs.good()) {throw std::runtime_error(std::string("[error] ") + who + ": file open error: " + fname);}
}

// 80-byte title record of the DCD header (any text; readers show it as a remark)
constexpr char kDcdTitle[80] = "libmjg (MI355X) trajectory, Mjolnir DCDObserver layout";
} // namespace

// XYZObserver constructor (:24-32): truncate prefix_position.xyz and prefix_velocity.xyz.
// Particle names default to "X", the name read_system gives a particle without one
// (read_system.hpp:277).
void Engine::xyz_observer_open(const std::string& prefix, const char* const* names)
{
    truncate_or_throw(prefix + "_position.xyz", "XYZObserver");
    truncate_or_throw(prefix + "_velocity.xyz", "XYZObserver");
    names_.assign(n_, "X");
    if(names)
    {
        for(size_t i = 0; i < n_; ++i)
        {
            if(!names[i]) {throw std::invalid_argument("null particle name");}
            names_[i] = names[i];
        }
    }
    xyz_prefix_ = prefix;
}

// DCDObserver constructor (:27-36): truncate prefix_position.dcd and prefix_velocity.dcd.
void Engine::dcd_observer_open(const std::string& prefix)
{
    truncate_or_throw(prefix + "_position.dcd", "DCDObserver");
    truncate_or_throw(prefix + "_velocity.dcd", "DCDObserver");
    dcd_prefix_ = prefix;
}

// DCDObserver::initialize / write_header (:39-51, :119-199): three Fortran records per file --
// the 84-byte control record (frame count patched at the end, first index 0, save interval,
// total steps, number of molecules, dt as float, the unit-cell flag, version 24), the title
// record, and the particle count.
void Engine::dcd_begin(uint64_t total_step, uint64_t save_step)
{
    dcd_frames_ = 0;
    for(const char* suffix : {"_position.dcd", "_velocity.dcd"})
    {
        const std::string fname = dcd_prefix_ + suffix;
        std::ofstream ofs(fname, std::ios::binary | std::ios::trunc);
        if(!ofs.good()) {throw std::runtime_error("[error] DCDObserver: file open error: " + fname);}
        const int32_t zero = 0;
        put(ofs, int32_t(84));
        ofs.write("CORD", 4);
        put(ofs, zero);                              // frames, patched by dcd_end
        put(ofs, zero);                              // index of the first frame
        put(ofs, int32_t(save_step));
        put(ofs, int32_t(total_step));
        put(ofs, int32_t(n_molecules_));
        for(int k = 0; k < 4; ++k) {put(ofs, zero);}
        put(ofs, float(dt_));
        put(ofs, int32_t(box_.periodic ? 1 : 0));
        for(int k = 0; k < 8; ++k) {put(ofs, zero);}
        put(ofs, int32_t(24));
        put(ofs, int32_t(84));

        put(ofs, int32_t(84));
        put(ofs, int32_t(1));
        ofs.write(kDcdTitle, 80);
        put(ofs, int32_t(84));

        put(ofs, int32_t(4));
        put(ofs, int32_t(n_));
        put(ofs, int32_t(4));
    }
}

// DCDObserver::finalize (:74-102): the frame count at byte 8 of both files.
void Engine::dcd_end()
{
    for(const char* suffix : {"_position.dcd", "_velocity.dcd"})
    {
        std::fstream f(dcd_prefix_ + suffix, std::ios::binary | std::ios::in | std::ios::out);
        if(!f.good()) {throw std::runtime_error("[error] DCDObserver: file open error: " + dcd_prefix_ + suffix);}
        f.seekp(2 * sizeof(int32_t), std::ios::beg);
        put(f, int32_t(dcd_frames_));
    }
}

// XYZObserver::output (:46-79) and DCDObserver::output (:266-341) for one step.
void Engine::trajectory_output(uint64_t step)
{
    if(xyz_prefix_.empty() && dcd_prefix_.empty()) {return;}
    std::vector<double> x(3 * n_), v(3 * n_);
    get_state(x.data(), v.data(), nullptr);
    if(!xyz_prefix_.empty())
    {
        const std::vector<double>* src[2] = {&x, &v};
        const char* suffix[2] = {"_position.xyz", "_velocity.xyz"};
        for(int k = 0; k < 2; ++k)
        {
            std::ostringstream os;
            os << n_ << "\nstep = " << step << '\n';
            const std::vector<double>& a = *src[k];
            for(size_t i = 0; i < n_; ++i)
            {
                os << names_[i] << ' ' << std::fixed << std::setprecision(8) << a[3 * i] << ' ' << a[3 * i + 1] << ' '
                   << a[3 * i + 2] << '\n';
            }
            std::ofstream ofs(xyz_prefix_ + suffix[k], std::ios::app);
            ofs << os.str();
        }
    }
    if(!dcd_prefix_.empty())
    {
        dcd_frames_ += 1;
        std::vector<float> buf(n_);
        const int32_t bytes = int32_t(sizeof(float) * n_);
        const std::vector<double>* src[2] = {&x, &v};
        const char* suffix[2] = {"_position.dcd", "_velocity.dcd"};
        for(int k = 0; k < 2; ++k)
        {
            std::ofstream ofs(dcd_prefix_ + suffix[k], std::ios::app | std::ios::binary);
            if(k == 0 && box_.periodic)
            {
                // unit cell record (write_unitcell_if_needed, :225-253): A, gamma, B, beta, alpha, C
                put(ofs, int32_t(6 * sizeof(double)));
                put(ofs, box_.width[0]);
                put(ofs, 90.0);
                put(ofs, box_.width[1]);
                put(ofs, 90.0);
                put(ofs, 90.0);
                put(ofs, box_.width[2]);
                put(ofs, int32_t(6 * sizeof(double)));
            }
            const std::vector<double>& a = *src[k];
            for(int c = 0; c < 3; ++c)
            {
                for(size_t i = 0; i < n_; ++i) {buf[i] = static_cast<float>(a[3 * i + c]);}
                put(ofs, bytes);
                ofs.write(reinterpret_cast<const char*>(buf.data()), bytes);
                put(ofs, bytes);
            }
        }
    }
}

} // mjg
