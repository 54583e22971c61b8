// mt19937.hpp -- Mersenne Twister MT19937 (Matsumoto & Nishimura, ACM TOMACS
// 8(1), 1998), with the init_genrand seeding and the genrand_res53 /
// genrand_real2 output transforms of the authors' reference release.  FDPS's
// MTTS class, used by the reference's Plummer generator
// (src/particle_distribution_generator.hpp:190-191), is this generator.
#pragma once
#include <cstdint>

namespace pth {

class MT19937 {
  public:
    explicit MT19937(uint32_t seed = 5489u) { seed_(seed); }

    void seed_(uint32_t s)
    {
        st_[0] = s;
        for (int i = 1; i < N; i++)
            st_[i] = 1812433253u * (st_[i - 1] ^ (st_[i - 1] >> 30)) + (uint32_t)i;
        idx_ = N;
    }

    uint32_t next_u32()
    {
        if (idx_ >= N) twist_();
        uint32_t y = st_[idx_++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= (y >> 18);
        return y;
    }

    // 53-bit resolution uniform on [0,1)
    double res53()
    {
        uint32_t a = next_u32() >> 5, b = next_u32() >> 6;
        return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
    }

    // 32-bit resolution uniform on [0,1)
    double real2() { return next_u32() * (1.0 / 4294967296.0); }

  private:
    static constexpr int N = 624, M = 397;
    uint32_t st_[N];
    int idx_;

    void twist_()
    {
        static const uint32_t mag[2] = {0u, 0x9908b0dfu};
        int k = 0;
        for (; k < N - M; k++) {
            uint32_t y = (st_[k] & 0x80000000u) | (st_[k + 1] & 0x7fffffffu);
            st_[k] = st_[k + M] ^ (y >> 1) ^ mag[y & 1u];
        }
        for (; k < N - 1; k++) {
            uint32_t y = (st_[k] & 0x80000000u) | (st_[k + 1] & 0x7fffffffu);
            st_[k] = st_[k + (M - N)] ^ (y >> 1) ^ mag[y & 1u];
        }
        uint32_t y = (st_[N - 1] & 0x80000000u) | (st_[0] & 0x7fffffffu);
        st_[N - 1] = st_[M - 1] ^ (y >> 1) ^ mag[y & 1u];
        idx_ = 0;
    }
};

}  // namespace pth
