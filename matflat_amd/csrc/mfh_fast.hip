// mfh_fast.hip -- FAST-mode batched Cholesky (no right-hand side) and the
// per-bucket dispatcher of the FAST kernels; the kernel templates are in
// mfh_fast_kernels.h, the SPD-solve instantiations in mfh_fast_posv.hip.
#include "mfh_fast_kernels.h"

// Measured on MI355X by matflat_amd/autotune.py (8192 SPD solves per bucket):
// with the register caps of fast_min_waves, one wavefront per matrix wins in
// every bucket (89..96 rows: 0.49 ms against 0.67 ms on two wavefronts).
int mfh_fast_default_waves(int) { return 1; }

int mfh_launch_chol_fast_bucket(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count, bool posv)
{
    return posv ? mfh_launch_chol_fast_posv(h, B, q, rb, count) : dispatch_fast<false>(h, B, q, rb, count);
}
