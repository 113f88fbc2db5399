// Instantiations of the register-resident beam (hnsw_beam_reg.hpp) for R capacity 128.

#include "hnsw_beam_reg.hpp"

namespace vdx::dev {

const void* hnsw_beam_reg_ptr_c2(BeamShape shape, int mode) { return hnsw_beam_reg_ptr<2>(shape, mode); }

}  // namespace vdx::dev
