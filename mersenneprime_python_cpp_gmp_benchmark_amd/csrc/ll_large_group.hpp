// ll_large_group.hpp -- the body of one ll_large_g<n>.hip unit: instantiates the kernels of the
// schedules of group LL_LARGE_GROUP (ll_large_scheds.inc) and returns them by (M, W, H, CW).
#include "ll_large_impl.hpp"

#define LL_CAT2(a, b) a##b
#define LL_CAT(a, b) LL_CAT2(a, b)

// a row expands only in the unit of its group
#define LLS(G, TR, MV, WV, HV, CWV) LL_CAT(LLS_G, G)(TR, MV, WV, HV, CWV)
#if LL_LARGE_GROUP == 0
#define LLS_G0(TR, MV, WV, HV, CWV) LLS_TR##TR(MV, WV, HV, CWV)
#else
#define LLS_G0(TR, MV, WV, HV, CWV)
#endif
#if LL_LARGE_GROUP == 1
#define LLS_G1(TR, MV, WV, HV, CWV) LLS_TR##TR(MV, WV, HV, CWV)
#else
#define LLS_G1(TR, MV, WV, HV, CWV)
#endif
#if LL_LARGE_GROUP == 2
#define LLS_G2(TR, MV, WV, HV, CWV) LLS_TR##TR(MV, WV, HV, CWV)
#else
#define LLS_G2(TR, MV, WV, HV, CWV)
#endif
#if LL_LARGE_GROUP == 3
#define LLS_G3(TR, MV, WV, HV, CWV) LLS_TR##TR(MV, WV, HV, CWV)
#else
#define LLS_G3(TR, MV, WV, HV, CWV)
#endif
#define LLS_MATCH(MV, WV, HV, CWV) (M == (MV) && W == (WV) && H == (HV) && CW == (CWV))
#define LLS_TR1(MV, WV, HV, CWV) \
    if (!gl && LLS_MATCH(MV, WV, HV, CWV)) { L = make_large_launch<false, WV, HV, CWV>(); return true; }
#define LLS_TR2(MV, WV, HV, CWV) \
    if (gl && LLS_MATCH(MV, WV, HV, CWV)) { L = make_large_launch<true, WV, HV, CWV>(); return true; }
#define LLS_TR3(MV, WV, HV, CWV) LLS_TR1(MV, WV, HV, CWV) LLS_TR2(MV, WV, HV, CWV)

namespace llh {
bool LL_CAT(large_launch_group, LL_LARGE_GROUP)(int transform, u64 M, int W, int H, int CW, LargeLaunch &L) {
    const bool gl = transform == kNTT;
#include "ll_large_scheds.inc"
    return false;
}
}  // namespace llh
