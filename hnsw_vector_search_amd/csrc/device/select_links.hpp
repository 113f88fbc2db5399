#pragma once

// Neighbour choice of inserted nodes on the GPU (select_links.hip).

#include <hip/hip_runtime.h>

#include <cstdint>

#include "beam_args.hpp"
#include "device_common.hpp"

namespace vdx::dev {

struct SelectArgs {
    const float* vectors;  // the graph's rows (ld floats each)
    std::uint32_t ld, l2;
    std::uint32_t nb, n8, tail, first_tail;  // the dimension split of the exact distance order
    const Neighbor* found;         // per node: its layer-0 search results, sorted by (distance, id)
    std::uint64_t found_stride;
    const std::uint32_t* found_count;
    std::uint32_t m;               // keep at most m (M <= 32)
    std::uint64_t count;           // nodes
    Neighbor* chosen;              // per node: m slots
    std::uint32_t* chosen_count;
};

hipError_t launch_select_links(const SelectArgs& a, BeamShape shape, hipStream_t stream);

}  // namespace vdx::dev
