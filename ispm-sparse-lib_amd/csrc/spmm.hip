/* spmm.hip -- multi-vector SlELL product: one pass over the matrix serves k = 2..8
 * right-hand sides (kernel in spmv_kernel.h, one row per thread). */
#include "spmv_kernel.h"

namespace ispm {
using namespace spmvk;

int spmm_dispatch(const ispm_launch *p, const ispm_slell_desc *A, int vt, int k,
                  const void *X, size_t ldx, void *Y, size_t ldy, const layout_hints *hints)
{
  if (use_stream(p, A)) {
    /* streamed mapping (short irregular rows): elements staged in LDS, one row per thread */
    switch (k) {
      case 2: return launch_stream_spmm_types<2>(A, vt, X, ldx, Y, ldy, hints);
      case 3: return launch_stream_spmm_types<3>(A, vt, X, ldx, Y, ldy, hints);
      case 4: return launch_stream_spmm_types<4>(A, vt, X, ldx, Y, ldy, hints);
      case 5: return launch_stream_spmm_types<5>(A, vt, X, ldx, Y, ldy, hints);
      case 6: return launch_stream_spmm_types<6>(A, vt, X, ldx, Y, ldy, hints);
      case 7: return launch_stream_spmm_types<7>(A, vt, X, ldx, Y, ldy, hints);
      case 8: return launch_stream_spmm_types<8>(A, vt, X, ldx, Y, ldy, hints);
      default: return fail(ISPM_ERR_UNSUPPORTED, "spmm: k=%d", k);
    }
  }
  switch (k) {
    case 2: return launch_types_h<2, 1>(p, A, vt, X, ldx, Y, ldy);
    case 3: return launch_types_h<3, 1>(p, A, vt, X, ldx, Y, ldy);
    case 4: return launch_types_h<4, 1>(p, A, vt, X, ldx, Y, ldy);
    case 5: return launch_types_h<5, 1>(p, A, vt, X, ldx, Y, ldy);
    case 6: return launch_types_h<6, 1>(p, A, vt, X, ldx, Y, ldy);
    case 7: return launch_types_h<7, 1>(p, A, vt, X, ldx, Y, ldy);
    case 8: return launch_types_h<8, 1>(p, A, vt, X, ldx, Y, ldy);
    default: return fail(ISPM_ERR_UNSUPPORTED, "spmm: k=%d", k);
  }
}

}  // namespace ispm
