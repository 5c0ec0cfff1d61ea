// Device code of libmjx_warp_amd.so: the stage kernels of one instantiation
// list of stages.hpp, named by MJH_KERNEL_LIST at compile time (the build
// compiles this file once per list).
#include "stages.hpp"

#ifndef MJH_KERNEL_LIST
#error "compile with -DMJH_KERNEL_LIST=<a list of stages.hpp>"
#endif

namespace mjh {
MJH_KERNEL_LIST(MJH_INSTANTIATE_STAGE, MJH_INSTANTIATE_GROUPS)
}  // namespace mjh
