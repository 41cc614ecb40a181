// mfh_fast.hip -- FAST-mode batched Cholesky (no right-hand side) and the
// per-bucket dispatcher of the FAST kernels; the kernel templates are in
// mfh_fast_kernels.h, the SPD-solve instantiations in mfh_fast_posv.hip.
#include "mfh_fast_kernels.h"

// Wavefronts per matrix by row blocks (matrices of up to 8 * rb rows), as
// measured on MI355X by matflat_amd/autotune.py (8,192 SPD solves per
// bucket): one wavefront wins up to 104 rows (e.g. 89..96 rows: 0.440 ms on
// one against 0.479 on two and 0.720 on four), two from 97 to 128 rows
// (97..104 rows inside the full SPD-solve step: 0.41 against 0.49 ms).
int mfh_fast_default_waves(int rb) { return (rb >= 13 && rb <= 16) ? 2 : 1; }

int mfh_launch_chol_fast_bucket(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count, bool posv)
{
    return posv ? mfh_launch_chol_fast_posv(h, B, q, rb, count) : dispatch_fast<false>(h, B, q, rb, count);
}
