// ll_large_g3.hip -- kernels of the large-engine schedules of group 3 (ll_large_scheds.inc).
#define LL_LARGE_GROUP 3
#include "ll_large_group.hpp"
