// mfh_fast_posv.hip -- FAST-mode batched SPD solve (Cholesky with the
// right-hand side carried as row n, then the back substitution); the kernel
// templates are in mfh_fast_kernels.h.
#include "mfh_fast_kernels.h"

int mfh_launch_chol_fast_posv(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count)
{
    return dispatch_fast<true>(h, B, q, rb, count);
}
