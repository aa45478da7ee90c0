// ll_large_g2.hip -- kernels of the large-engine schedules of group 2 (ll_large_scheds.inc).
#define LL_LARGE_GROUP 2
#include "ll_large_group.hpp"
