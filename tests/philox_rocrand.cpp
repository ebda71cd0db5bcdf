// Independent Philox4x32-10 for cross-checking the oracle: ROCm's rocRAND
// engine (rocrand/rocrand_philox4x32_10.h), host side only.  Reads lines of
// "c0 c1 c2 c3 k0 k1" (decimal) on stdin and prints the four output words.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

struct RawPhilox : rocrand_device::philox4x32_10_engine {
    RawPhilox() : rocrand_device::philox4x32_10_engine(0ull, 0ull, 0ull) {}
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
    RawPhilox e;
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = e.block(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
