// mxv / vxm: the L3 dispatch of the reference (src/backend/operations/operations.h:151-211)
// on top of the L2 kernels.
//
//   dense u  -> pull SpMV: unmasked in the matrix's format (spmv.h:29-67 /
//               92-131), masked always on CSR (mxv) / CSC (vxm) (spmv.h:68-78,
//               133-145);
//   sparse u -> push SpMSpV over the CSC (mxv) / CSR (vxm) (spmspv.h:126-146);
//   then w is made sparse or dense by the 30% rule (vector.h:117-129).
#include "gbk_objects.h"

namespace gbk {

int accum_into(int type, void* y, const void* t, int64_t n, int accum);
int build_flag_mask(gbk_vector mask, uint8_t* flag);

static int mask_field(const gbk_desc* d) { return d ? d->mask : GBK_DEFAULT; }

static int check_binop(int accum) {
    if (accum < GBK_ACC_NONE || accum > GBK_ACC_NE) return fail(GBK_INVALID_VALUE, "unknown binary operator %d", accum);
    return GBK_SUCCESS;
}

// w = accum(w, A u) (transposed = false) or accum(w, u A) (transposed = true).
int matvec(gbk_vector w, gbk_vector mask, int accum, gbk_semiring op, gbk_matrix A, gbk_vector u,
           const gbk_desc* desc, bool transposed) {
    if (!w || !u || !A) return fail(GBK_UNINITIALIZED_OBJECT, "uninitialized object in %s", transposed ? "vxm" : "mxv");
    if (!A->built) return fail(GBK_UNINITIALIZED_OBJECT, "matrix not built");
    GBK_TRY(ensure_init());
    GBK_TRY(check_binop(accum));
    if (w->n != A->n || u->n != A->n || (mask && mask->n != A->n))
        return fail(GBK_DIMENSION_MISMATCH, "dimension mismatch");
    if (w->type != A->type || u->type != A->type)
        return fail(GBK_DOMAIN_MISMATCH, "vector and matrix element types differ");
    const int mfield = mask_field(desc);

    if (u->storage == GBK_STORAGE_DENSE) {
        // pull: x = u (dense), y = w (made dense first, as getDense() does)
        GBK_TRY(vec_force_dense(w));
        const void* x = u->dense;
        void* y = w->dense;
        if (!mask) {
            if (transposed && A->format == GBK_FMT_LAV)
                return fail(GBK_INVALID_VALUE, "unsupported matrix storage format in VSpM");
            if (x == y) {
                void* t = A->ws.buf[0];
                GBK_TRY(spmv_unmasked(A, transposed, x, t, op.add, op.mul, GBK_ACC_NONE));
                GBK_TRY(accum_into(A->type, y, t, A->n, accum));
            } else {
                GBK_TRY(spmv_unmasked(A, transposed, x, y, op.add, op.mul, accum));
            }
        } else {
            const bool comp = (mfield == GBK_STR_COMP);
            const void* xin = x;
            if (x == y) {  // masked kernels read x while writing y: keep a copy
                void* t = A->ws.buf[3];
                GBK_HIP(hipMemcpyAsync(t, x, type_size(A->type) * A->n, hipMemcpyDeviceToDevice, stream()));
                xin = t;
            }
            if (mask->storage == GBK_STORAGE_DENSE) {
                GBK_TRY(spmv_masked_dense(A, transposed, xin, y, op.add, op.mul, accum, mask->dense, mask->type, comp));
            } else if (!comp) {
                GBK_TRY(spmv_masked_rows(A, transposed, xin, y, op.add, op.mul, accum, mask->ids, mask->nvals));
            } else {
                uint8_t* flag = (uint8_t*)A->ws.buf[2];
                GBK_TRY(build_flag_mask(mask, flag));
                GBK_TRY(spmv_masked_dense(A, transposed, xin, y, op.add, op.mul, accum, flag, GBK_BOOL, true));
            }
        }
    } else {
        // push: x = u (sparse); mxv pushes over the CSC, vxm over the CSR
        GBK_TRY(vec_force_dense(w));
        const int32_t* ids = u->ids;
        const void* vals = u->svals;
        if (u == w) return fail(GBK_INVALID_OBJECT, "aliased input and output of a sparse product");
        GBK_TRY(spmspv(A, transposed, ids, vals, u->nvals, w->dense, op.add, op.mul, accum, mask, mfield));
    }
    return vec_convert_if_required(w);
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_mxv(gbk_vector w, gbk_vector mask, int accum, gbk_semiring op, gbk_matrix A, gbk_vector u,
                       const gbk_desc* desc) {
    return matvec(w, mask, accum, op, A, u, desc, false);
}

extern "C" int gbk_vxm(gbk_vector w, gbk_vector mask, int accum, gbk_semiring op, gbk_vector u, gbk_matrix A,
                       const gbk_desc* desc) {
    return matvec(w, mask, accum, op, A, u, desc, true);
}
