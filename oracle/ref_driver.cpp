// ORACLE -- test infrastructure only.
//
// Thin C entry points around the two pieces of the reference that compile
// from their own sources as they lie (no stand-in headers needed):
//   src/Common/Algorithms.h:21-38  Common::SolveTridiagonal  (FTYPE = float, as shipped)
//   src/Common/Geometry.h:538-562  Common::FluidParams       (both constructors)
// Built by oracle/Makefile (target `ref`) into oracle/_ref/libref.so, only when
// the reference tree is present.  The 3D/2D solver sources include
// <cuda_runtime.h>, which this image lacks, so they are not built.
#include "Geometry.h"
#include "Algorithms.h"

extern "C" {

void ref_thomas_f32(float* a, float* b, float* c, float* d, float* x, int n) {
    Common::SolveTridiagonal(a, b, c, d, x, n);
}

// FluidParams(Re, Pr, lambda): out = {v_T, v_vis, t_vis, t_phi}
void ref_fluid_params_normalized(double Re, double Pr, double lambda, float* out) {
    Common::FluidParams p(Re, Pr, lambda);
    out[0] = p.v_T; out[1] = p.v_vis; out[2] = p.t_vis; out[3] = p.t_phi;
}

// FluidParams(vis, rho, R, k, cv)
void ref_fluid_params_dimensional(double vis, double rho, double R, double k, double cv,
                                  float* out) {
    Common::FluidParams p(vis, rho, R, k, cv);
    out[0] = p.v_T; out[1] = p.v_vis; out[2] = p.t_vis; out[3] = p.t_phi;
}

}
