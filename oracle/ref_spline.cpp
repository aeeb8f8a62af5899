// The reference's cubic spline kernel, for pinning the oracle's.
//
// TEST INFRASTRUCTURE ONLY.  spline_ker (nbr_search.cpp:39-53) is an inline
// function of nbr_search.cpp that the reference defines but never calls, so
// no object of the reference exports it.  This translation unit compiles the
// reference's nbr_search.cpp where it lies (oracle/Makefile passes -iquote to
// the reference tree) and exposes the function through a C entry point, so
// that tests compare orc_spline with it bit for bit.  Built into
// oracle/_ref/libasmref_spline.so only where the reference tree exists.
#include "nbr_search.cpp"

extern "C" float ref_spline_ker(const float* x, const float* y, float hinv)
{
    return spline_ker(x, y, hinv);
}
