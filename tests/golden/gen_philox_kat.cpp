// Writes Philox4x32-10 known-answer vectors computed by rocRAND's host-callable
// engine (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), an implementation
// independent of this repository's.  Build and run with make_philox_kat.sh.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

struct Probe : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

int main() {
    Probe p;
    const uint32_t cases[][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
        {1u, 2u, 3u, 4u, 5u, 6u},
        {7u, 0x03000005u, 11u, 1023u, 0xdeadbeefu, 0x12345678u},
    };
    std::printf("[\n");
    const int n = sizeof(cases) / sizeof(cases[0]);
    for (int i = 0; i < n; ++i) {
        const uint32_t* c = cases[i];
        uint4 o = p.rounds(uint4{c[0], c[1], c[2], c[3]}, uint2{c[4], c[5]});
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    c[0], c[1], c[2], c[3], c[4], c[5], o.x, o.y, o.z, o.w, i + 1 < n ? "," : "");
    }
    std::printf("]\n");
    return 0;
}
