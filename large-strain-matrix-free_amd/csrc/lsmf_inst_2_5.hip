// Kernel instantiation unit for dim = 2, N = p + 1 = n_q_1d = 5.
#include "lsmf_launch.hpp"
LSMF_INSTANTIATE(2, 5)
