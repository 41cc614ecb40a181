// mfh_internal.h -- host-side declarations shared by the library's sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/matflat_hip.h"
#include "mfh_device.h"

// Size buckets of the exact kernels: bucket b holds keys in (16b, 16(b+1)];
// the key is the number of rows a workgroup handles (n + 1 for the SPD solve).
#define MFH_NBUCKETS 9
#define MFH_BUCKET_STEP 16
// Largest lda * n of one matrix (2 GiB of doubles): kernels address a matrix
// with 32-bit byte offsets from its base pointer.
#define MFH_MAX_MATRIX_ELEMS (1LL << 28)

// Largest order of the LDS-resident exact kernels (mfh_exact.hip); larger
// orders, up to MFH_EXACT_MAX_N, run the kernels of mfh_general.hip.
#define MFH_LDS_MAX_N 128

// The FAST kernels use finer buckets: (8f, 8f + 8] rows.
#define MFH_FINE_STEP 8

// One kernel launch bracketed by events (mfh_set_timing).
struct MfhTimedLaunch {
    hipEvent_t start, stop;
    int kind, rows;
};

struct mfh_handle {
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cus = 0;
    void* ws = nullptr;  // device workspace
    size_t ws_bytes = 0;
    void* aux = nullptr;  // second device buffer: staging of calls that run another entry's kernels on ws
    size_t aux_bytes = 0;
    bool capturing = false;  // plans pre-size the workspace; no allocation while set
    bool no_alloc = false;   // inside mfh_dposv_vbatched_noalloc: ws is the static scratch, nothing is created
    void* scratch = nullptr; // this device's copy of the library's static scratch array (device address)
    size_t scratch_bytes = 0;
    int math_mode = MFH_MATH_EXACT;
    int fast_waves[18] = {};              // FAST Cholesky waves per matrix by row blocks (0 = default)
    int chol_shape[MFH_NBUCKETS][2] = {}; // exact Cholesky: lanes per matrix, rows per lane, by bucket (0 = default)
    bool timing = false;                  // bracket launches with events
    int bucket_streams = 2;               // FAST bucket launches forked over this many streams (1 = in order)
    std::vector<hipStream_t> side;        // side streams of the fork (default priority), created on first use
    std::vector<hipEvent_t> side_ev;      // fork / join events, one per side stream plus one
    std::vector<MfhTimedLaunch> timed;    // recorded launches not yet read
    std::vector<hipEvent_t> spare_events; // reusable events
};

// Thread-local error text behind mfh_last_error().
void mfh_set_error(const std::string& msg);
int mfh_fail(int status, const std::string& msg);

// Grows the handle workspace to at least bytes (device memory).
int mfh_ws_reserve(mfh_handle* h, size_t bytes);
// Side streams for forking independent launches off the handle's stream
// (created once, default priority); fills at least k of them.
int mfh_side_streams(mfh_handle* h, int k);

// Grows the handle's second buffer (aux) to at least bytes.
int mfh_aux_reserve(mfh_handle* h, size_t bytes);

// Per-launch timing: when h->timing is set (and no plan is being captured),
// mfh_timing_begin records a start event of a new record on the handle's
// stream and mfh_timing_end the stop event of the newest record.
int mfh_timing_begin(mfh_handle* h, int kind, int rows);
int mfh_timing_end(mfh_handle* h);

#define MFH_HIP_CHECK(expr)                                                          \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return mfh_fail(MFH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Batch descriptors passed by value to kernels.  Pointer arrays may be null,
// in which case the strided fields describe a uniform batch.
struct MfhGetrfBatch {
    const int* m;
    const int* n;
    double* const* a;
    const int* lda;
    int* const* piv;
    int m0, n0, lda0;
    double* a0;
    int64_t stride_a;
    int* piv0;
    int64_t stride_piv;
    int* sign;
    int* info;
};

struct MfhCholBatch {
    const int* n;
    double* const* a;
    const int* lda;
    double* const* x;  // posv only
    int* info;
};

struct MfhTrsvBatch {
    const int* n;
    const double* const* a;
    const int* lda;
    double* const* x;
    const int* incx;
    int* info;
    int uplo, trans, unit;
};

struct MfhGetrsBatch {
    const int* n;
    const double* const* lu;
    const int* lda;
    const int* const* piv;
    const double* const* b;
    double* const* x;
    int* info;
};

struct MfhGetriBatch {
    const int* n;
    double* const* a;
    const int* lda;
    const int* const* piv;
    int* info;
};

struct MfhGemmBatch {
    const int* transa;
    const int* transb;
    const int* m;
    const int* n;
    const int* k;
    const double* alpha;
    const double* const* a;
    const int* lda;
    const double* const* b;
    const int* ldb;
    const double* beta;
    double* const* c;
    const int* ldc;
    int* info;
};

struct MfhGemvBatch {
    const int* trans;
    const int* m;
    const int* n;
    const double* const* a;
    const int* lda;
    const double* const* x;
    const int* incx;
    double* const* y;
    const int* incy;
    int* info;
};

struct MfhAdjointBatch {
    const int* n;
    const double* const* l;
    const int* lda;
    const double* const* x;
    const double* const* xbar;
    double* const* abar;
    const int* ldabar;
    double* const* bbar;
    int* info;
};

struct MfhCholAdjointBatch {
    const int* n;
    const double* const* l;
    const int* ldl;
    const double* const* lbar;
    const int* ldlbar;
    double* const* abar;
    const int* ldabar;
    int* info;
};

// Device workspace bytes the exact kernels need for a batch of count.
size_t mfh_exact_ws_bytes(int count);

// Workspace bytes of one GEMM call (mfh_gemm.hip); plans reserve this.
size_t mfh_gemm_ws_bytes(int count, int max_m, int max_n, bool fast);

// Shapes (lanes per matrix G, rows per lane S) the exact Cholesky kernel is
// built for; a bucket of up to R rows can take any of them with G * S >= R.
bool mfh_chol_shape_ok(int bucket, int g, int s);

// Built-in choice of wavefronts per matrix (1 or 2) of the FAST Cholesky
// kernel for matrices of up to 8 * rb rows.
int mfh_fast_default_waves(int rb);

// FAST-mode Cholesky / SPD solve for one fine size bucket (mfh_fast.hip):
// every matrix of the queue has at most 8 * rb rows (n, or n + 1 for posv).
int mfh_launch_chol_fast_bucket(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count, bool posv);
// The SPD-solve half of it (mfh_fast_posv.hip).
int mfh_launch_chol_fast_posv(mfh_handle* h, const MfhCholBatch& B, MfhQueue q, int rb, long count);

// Workspace bytes of the blocked LU solvers (mfh_mixed.hip).
size_t mfh_gesv_ws_bytes(int count, int max_n, bool mixed);

// Launchers (mfh_exact.hip, mfh_gemm.hip, mfh_adjoint.hip, mfh_mixed.hip).
int mfh_launch_getrf(mfh_handle* h, int count, const MfhGetrfBatch& B, int max_m, int max_n);
int mfh_launch_potrf(mfh_handle* h, int count, const MfhCholBatch& B, int max_n, bool posv);
int mfh_launch_trsv(mfh_handle* h, int count, const MfhTrsvBatch& B, int max_n);
int mfh_launch_getrs(mfh_handle* h, int count, const MfhGetrsBatch& B, int max_n);
int mfh_launch_gemm(mfh_handle* h, int count, const MfhGemmBatch& B, int max_m, int max_n,
                    int max_k, bool complex);
int mfh_launch_getri(mfh_handle* h, int count, const MfhGetriBatch& B, int max_n);
int mfh_launch_gemv(mfh_handle* h, int count, const MfhGemvBatch& B, int max_m, int max_n);
int mfh_launch_posv_backward(mfh_handle* h, int count, const MfhAdjointBatch& B, int max_n);
int mfh_launch_potrf_backward(mfh_handle* h, int count, const MfhCholAdjointBatch& B, int max_n);
// Factorization.Gevd over a batch (mfh_gevd.hip): argument check, the exact
// Cholesky of B, then one wavefront per matrix.  Uses h->aux (2 count ints).
int mfh_launch_gevd(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, double* const* b,
                    const int* ldb, double* const* w, int* info, int max_n);
int mfh_launch_gesv_mixed(mfh_handle* h, int count, const int* n, const double* const* a,
                          const int* lda, const double* const* b, double* const* x, int max_iter,
                          double tol_per_n, int* iters, double* berr, int* info, int max_n,
                          bool mixed);
// mfh_general.hip: double matrices above MFH_LDS_MAX_N rows (the size sort
// has reported invalid ones), and the complex128 entry points.
size_t mfh_potrf_gen_aux_bytes(const mfh_handle* h, int count, int max_n, bool cplx);
int mfh_launch_getrf_gen_d(mfh_handle* h, int count, const MfhGetrfBatch& B, int max_m, int max_n);
int mfh_launch_potrf_gen_d(mfh_handle* h, int count, const MfhCholBatch& B, int max_n, bool posv);
int mfh_launch_trsv_gen_d(mfh_handle* h, int count, const MfhTrsvBatch& B, int max_n);
int mfh_launch_getrs_gen_d(mfh_handle* h, int count, const MfhGetrsBatch& B, int max_n);
int mfh_launch_zgetrf(mfh_handle* h, int count, const int* m, const int* n, double* const* a, const int* lda,
                      int* const* piv, int* sign, int* info, int max_m, int max_n);
int mfh_launch_zpotrf(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, int* info,
                      int max_n);
int mfh_launch_ztrsv(mfh_handle* h, int uplo, int trans, int count, const int* n, const double* const* a,
                     const int* lda, double* const* x, const int* incx, int* info, int max_n);
int mfh_launch_getrf_blocked(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                             int* const* piv, int* info, int max_n);
