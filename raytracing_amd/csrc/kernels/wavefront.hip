// The wavefront integrator's kernels
#include "kernel_api.hpp"
#include "path_tracer.hpp"
