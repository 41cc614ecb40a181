// mfh_fast.hip -- FAST-mode batched Cholesky (no right-hand side) and the
// per-bucket dispatcher of the FAST kernels; the kernel templates are in
// mfh_fast_kernels.h, the SPD-solve instantiations in mfh_fast_posv.hip.
#include "mfh_fast_kernels.h"

// Wavefronts per matrix by row blocks (matrices of up to 8 * rb rows), as
// measured on MI355X by matflat_amd/autotune.py (its table is
// matflat_amd/tuning/gfx950.json): one wavefront wins up to 104 rows (e.g.
// 89..96 rows: 0.42 ms on one against 0.47 on two and 0.72 on four; 97..104
// rows, close alone, 2.04 against 2.26 ms for the whole 100,000-solve step),
// two from 105 to 128 rows.
int mfh_fast_default_waves(int rb) { return (rb >= 14 && rb <= 16) ? 2 : 1; }

int mfh_launch_chol_fast_bucket(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count, bool posv)
{
    return posv ? mfh_launch_chol_fast_posv(h, B, q, rb, count) : dispatch_fast<false>(h, B, q, rb, count);
}
