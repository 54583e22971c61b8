// Minimal stand-in for the parts of FDPS's particle_simulator.hpp that the
// soft-force adapter touches (test scaffolding for tests/cpp/fdps_replay.cpp,
// not used by any build of the reference): scalar typedefs, F64vec (x, y, z),
// the symmetric F64mat (xx, yy, zz, xy, xz, yz), the members of
// SPJQuadrupoleInAndOut and SPJMonopoleInAndOut that PeTar reads
// (src/force_gpu_cuda.cu:596-608) and the wall clock GetWtime.
#pragma once
#include <chrono>
namespace PS {
typedef int S32;
typedef long long S64;
typedef double F64;
typedef float F32;
struct F64vec {
    F64 x, y, z;
    F64vec() : x(0), y(0), z(0) {}
    F64vec(F64 a, F64 b, F64 c) : x(a), y(b), z(c) {}
};
struct F64mat {
    F64 xx, yy, zz, xy, xz, yz;
};
struct SPJQuadrupoleInAndOut {
    F64 mass;
    F64vec pos;
    F64mat quad;
    F64 getCharge() const { return mass; }
    F64vec getPos() const { return pos; }
};
struct SPJMonopoleInAndOut {
    F64 mass;
    F64vec pos;
    F64 getCharge() const { return mass; }
    F64vec getPos() const { return pos; }
};
inline F64 GetWtime()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace PS
