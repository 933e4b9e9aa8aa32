// Kernel instantiation unit for dim = 3, N = p + 1 = n_q_1d = 5.
#include "lsmf_launch.hpp"
LSMF_INSTANTIATE(3, 5)
