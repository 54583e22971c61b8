// Stand-in for the two types of PeTar's src/profile.hpp that its GPU profile
// (src/force_gpu_cuda.hpp:6-93) is made of, with their interface
// (src/profile.hpp:176-265): Tprofile (a named wall-time accumulator) and
// NumCounter (a named count), each printing itself as one column.  Test
// scaffolding for the adapter built with GPU_PROFILE; a PeTar build includes
// its own profile.hpp.
#pragma once
#include <iomanip>
#include <iostream>

#include "particle_simulator.hpp"

#ifndef PROFILE_PRINT_WIDTH
#define PROFILE_PRINT_WIDTH 20
#endif

struct Tprofile {
    PS::F64 time;
    PS::F64 tbar;
    const char *name;
    Tprofile(const char *_name) : time(0.0), tbar(0.0), name(_name) {}
    void start() { time -= PS::GetWtime(); }
    void end() { time += PS::GetWtime(); }
    void print(std::ostream &fout, const PS::S32 divider = 1) { fout << name << ": " << time / divider << std::endl; }
    void dump(std::ostream &fout, const PS::S32 divider = 1, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        fout << std::setw(width) << time / divider;
    }
    void dumpName(std::ostream &fout, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        fout << std::setw(width) << name;
    }
    void reset() { time = tbar = 0.0; }
};

struct NumCounter {
    PS::S64 n;
    const char *name;
    NumCounter(const char *_name) : n(0), name(_name) {}
    NumCounter &operator+=(const PS::S64 _n)
    {
        n += _n;
        return *this;
    }
    NumCounter &operator=(const PS::S64 _n)
    {
        n = _n;
        return *this;
    }
    void dump(std::ostream &fout, const PS::S32 divider = 1, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        if (divider == 1) fout << std::setw(width) << n;
        else fout << std::setw(width) << (PS::F64)n / divider;
    }
    void dumpName(std::ostream &fout, const PS::S32 width = PROFILE_PRINT_WIDTH) const
    {
        fout << std::setw(width) << name;
    }
};
