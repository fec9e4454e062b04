// cvrp_types.h -- host/device descriptors of the CVRP genetic algorithm.
#pragma once
#include <stdint.h>

namespace emtf {

struct CvrpDev {
    int32_t n, cap, pop;
    int32_t woff;      // offset making every split ring value W + woff non-negative (see cvrp.hip)
    int64_t tour_off;  // tours[pop][n] uint8
    int64_t cost_off;  // cost[pop] int32
    int64_t xy_off;    // xy[n+1] uint32: coordinates relative to the depot, (x & 0xFFFF) | (y << 16)
    int64_t dem_off;   // demand[n+1] uint8
};

struct CvrpArgs {
    const CvrpDev* inst;
    int32_t n_inst, g0, g1, init;
    uint8_t* tours;
    int32_t* cost;
    const uint32_t* xy;
    const uint8_t* dem;
    float p_swap, p_2opt;
    int32_t n_islands, migrate_interval, migrate_count;
    uint32_t seed_lo, seed_hi;
};

}  // namespace emtf
