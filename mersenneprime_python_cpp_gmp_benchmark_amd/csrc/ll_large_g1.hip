// ll_large_g1.hip -- kernels of the large-engine schedules of group 1 (ll_large_scheds.inc).
#define LL_LARGE_GROUP 1
#include "ll_large_group.hpp"
