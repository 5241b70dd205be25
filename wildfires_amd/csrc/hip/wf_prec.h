/*
 * wf_prec.h -- the precision of one compilation of the kernel sources.
 *
 * The stage, pressure, projection and solver sources are compiled twice:
 *   WF_MIXED=0  namespace wf64: every field, table and spectrum in FP64 (the
 *               form that matches the reference to round-off);
 *   WF_MIXED=1  namespace wf32: the mixed form.  u, v, w, T, p, the derived
 *               fields, the grid tables and the pressure spectrum are FP32, and
 *               so is the stencil, FFT and fuel-update arithmetic.  FP64 stays
 *               where single precision would lose the solution: the fuel
 *               fraction Y is stored and accumulated in FP64 (fuel consumption
 *               stays conservative), the tridiagonal solves of the lowest
 *               wavenumbers (singular or near-singular under the Neumann z
 *               conditions) run in FP64, and so do the convergence reductions.
 * wf_real is the type of the first group; `double` in the kernel sources means
 * FP64 in both compilations.  WR(x) writes a floating literal in wf_real, so
 * the FP32 compilation does no hidden FP64 arithmetic (it is built with
 * -Wdouble-promotion -Werror=double-promotion).
 */
#ifndef WF_PREC_H
#define WF_PREC_H

#include <hip/hip_runtime.h>

#ifndef WF_MIXED
#define WF_MIXED 0
#endif

#if WF_MIXED
#define WF_NS wf32
#else
#define WF_NS wf64
#endif

namespace WF_NS {
#if WF_MIXED
typedef float wf_real;
typedef float2 wf_real2;
#else
typedef double wf_real;
typedef double2 wf_real2;
#endif
}  // namespace WF_NS

#define WR(x) ((wf_real)(x))

#endif
