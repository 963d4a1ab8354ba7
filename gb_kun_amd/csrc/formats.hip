// Optimized matrix representations (SELL-C-sigma, segmented CSR, LAV).
#include "gbk_objects.h"

namespace gbk {

int format_build(gbk_matrix A) {
    return fail(GBK_INVALID_VALUE, "matrix storage format %d is not available in this build", A->format);
}

void format_free(gbk_matrix A) { (void)A; }

int format_spmv(gbk_matrix A, int transposed, const void* x, void* y, int add, int mul, int accum, bool* handled) {
    (void)A; (void)transposed; (void)x; (void)y; (void)add; (void)mul; (void)accum;
    *handled = false;
    return GBK_SUCCESS;
}

int64_t format_bytes(gbk_matrix A) {
    (void)A;
    return 0;
}

}  // namespace gbk
