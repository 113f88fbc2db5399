// Instantiations of the register-resident beam (hnsw_beam_reg.hpp) for R capacity 64.

#include "hnsw_beam_reg.hpp"

namespace vdx::dev {

const void* hnsw_beam_reg_ptr_c1(BeamShape shape, int mode) { return hnsw_beam_reg_ptr<1>(shape, mode); }

}  // namespace vdx::dev
