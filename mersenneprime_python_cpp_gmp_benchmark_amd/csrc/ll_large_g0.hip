// ll_large_g0.hip -- kernels of the large-engine schedules of group 0 (ll_large_scheds.inc).
#define LL_LARGE_GROUP 0
#include "ll_large_group.hpp"
