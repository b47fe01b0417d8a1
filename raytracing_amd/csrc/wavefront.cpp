// The wavefront integrator's host driver: see wavefront.hpp
#include "wavefront.hpp"
