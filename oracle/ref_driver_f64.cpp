// ORACLE -- test infrastructure only.
//
// The reference's Thomas solver (src/Common/Algorithms.h:21-38) instantiated
// with FTYPE = double, the FP64 configuration the reference selects by
// changing the FTYPE macro (src/Common/Geometry.h:21).  Algorithms.h needs
// only that macro, so it is defined here instead of including Geometry.h.
#define FTYPE double
#include "Algorithms.h"

extern "C" void ref_thomas_f64(double* a, double* b, double* c, double* d, double* x, int n) {
    Common::SolveTridiagonal(a, b, c, d, x, n);
}
