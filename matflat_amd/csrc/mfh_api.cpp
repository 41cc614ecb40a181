// mfh_api.cpp -- the C ABI of libmatflat_hip.so: argument validation,
// handles, workspace and hipGraph plans.  Host argument checks mirror the
// order and meaning of MatFlat's own (ArgumentException /
// ArgumentNullException before any write, e.g. Factorization.Lu.cs:124-147);
// a failed check returns -(argument position) and launches nothing.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>

#include "mfh_internal.h"

namespace {
thread_local std::string g_error;
}

// Static device scratch of mfh_dposv_vbatched_noalloc: the size sort's keys
// and item lists (2 ints per matrix) and its histograms, for up to
// MFH_NOALLOC_MAX_COUNT matrices.  Every GPU the library runs on holds its
// own copy (8 MiB + 4 KiB) until the process ends.  Host code uses only the
// device address that hipGetSymbolAddress returns for the current device,
// never the array's name.
#define MFH_NOALLOC_SCRATCH_BYTES ((size_t)8 * MFH_NOALLOC_MAX_COUNT + 4096)
static __device__ __align__(256) unsigned char scratch[MFH_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64];  // per device: address of this device's scratch, looked up once
}  // namespace

// The scratch address on the current device (looked up on first use per
// device, with that device current).
static int scratch_address(int device, void** ptr)
{
    if (device < 0 || device >= 64) return mfh_fail(MFH_ERR_UNSUPPORTED, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return mfh_fail(MFH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return MFH_OK;
}

void mfh_set_error(const std::string& msg) { g_error = msg; }

int mfh_fail(int status, const std::string& msg)
{
    g_error = msg;
    return status;
}

int mfh_ws_reserve(mfh_handle* h, size_t bytes)
{
    if (h->ws_bytes >= bytes) return MFH_OK;
    if (h->no_alloc) return mfh_fail(MFH_ERR_STATE, "the static scratch is too small for this call");
    if (h->capturing) return mfh_fail(MFH_ERR_STATE, "workspace must be reserved before graph capture");
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    if (h->ws) {
        MFH_HIP_CHECK(hipStreamSynchronize(h->stream));
        MFH_HIP_CHECK(hipFree(h->ws));
        h->ws = nullptr;
        h->ws_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->ws, want);
    if (e != hipSuccess) return mfh_fail(MFH_ERR_ALLOC, std::string("hipMalloc workspace: ") + hipGetErrorString(e));
    h->ws_bytes = want;
    return MFH_OK;
}

int mfh_aux_reserve(mfh_handle* h, size_t bytes)
{
    if (h->aux_bytes >= bytes) return MFH_OK;
    if (h->no_alloc) return mfh_fail(MFH_ERR_STATE, "no second buffer in a call that allocates nothing");
    if (h->capturing) return mfh_fail(MFH_ERR_STATE, "workspace must be reserved before graph capture");
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    if (h->aux) {
        MFH_HIP_CHECK(hipStreamSynchronize(h->stream));
        MFH_HIP_CHECK(hipFree(h->aux));
        h->aux = nullptr;
        h->aux_bytes = 0;
    }
    hipError_t e = hipMalloc(&h->aux, want);
    if (e != hipSuccess) return mfh_fail(MFH_ERR_ALLOC, std::string("hipMalloc workspace: ") + hipGetErrorString(e));
    h->aux_bytes = want;
    return MFH_OK;
}

int mfh_side_streams(mfh_handle* h, int k)
{
    while ((int)h->side.size() < k) {
        hipStream_t s;
        MFH_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        h->side.push_back(s);
    }
    while ((int)h->side_ev.size() < k + 1) {
        hipEvent_t ev;
        MFH_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        h->side_ev.push_back(ev);
    }
    return MFH_OK;
}

namespace {
int take_event(mfh_handle* h, hipEvent_t* ev)
{
    if (!h->spare_events.empty()) {
        *ev = h->spare_events.back();
        h->spare_events.pop_back();
        return MFH_OK;
    }
    MFH_HIP_CHECK(hipEventCreate(ev));
    return MFH_OK;
}
}  // namespace

int mfh_timing_begin(mfh_handle* h, int kind, int rows)
{
    if (!h->timing || h->capturing || h->no_alloc) return MFH_OK;
    MfhTimedLaunch t{nullptr, nullptr, kind, rows};
    int st = take_event(h, &t.start);
    if (st) return st;
    st = take_event(h, &t.stop);
    if (st) {
        h->spare_events.push_back(t.start);
        return st;
    }
    h->timed.push_back(t);
    MFH_HIP_CHECK(hipEventRecord(t.start, h->stream));
    return MFH_OK;
}

int mfh_timing_end(mfh_handle* h)
{
    if (!h->timing || h->capturing || h->no_alloc || h->timed.empty()) return MFH_OK;
    MFH_HIP_CHECK(hipEventRecord(h->timed.back().stop, h->stream));
    return MFH_OK;
}

#define MFH_REQUIRE(cond, pos, msg)                  \
    do {                                             \
        if (!(cond)) return mfh_fail(-(pos), (msg)); \
    } while (0)

extern "C" {

const char* mfh_last_error(void) { return g_error.c_str(); }

int mfh_version(void) { return 100; }

int mfh_create(mfh_handle** handle)
{
    MFH_REQUIRE(handle != nullptr, 1, "handle must not be null");
    mfh_handle* h = new mfh_handle();
    hipError_t e = hipGetDevice(&h->device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&h->num_cus, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e != hipSuccess) {
        delete h;
        return mfh_fail(MFH_ERR_HIP, std::string("no usable HIP device: ") + hipGetErrorString(e));
    }
    // set-up of the static scratch on this device (device current)
    int st = scratch_address(h->device, &h->scratch);
    if (st) {
        delete h;
        return st;
    }
    h->scratch_bytes = MFH_NOALLOC_SCRATCH_BYTES;
    *handle = h;
    return MFH_OK;
}

int mfh_destroy(mfh_handle* h)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    if (h->ws || h->aux) hipStreamSynchronize(h->stream);
    if (h->ws) hipFree(h->ws);
    if (h->aux) hipFree(h->aux);
    for (auto& t : h->timed) {
        hipEventSynchronize(t.stop);
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    for (auto ev : h->spare_events) hipEventDestroy(ev);
    for (auto s : h->side) {
        hipStreamSynchronize(s);
        hipStreamDestroy(s);
    }
    for (auto ev : h->side_ev) hipEventDestroy(ev);
    delete h;
    return MFH_OK;
}

int mfh_set_stream(mfh_handle* h, void* stream)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    h->stream = (hipStream_t)stream;
    return MFH_OK;
}

int mfh_set_math_mode(mfh_handle* h, int mode)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(mode == MFH_MATH_EXACT || mode == MFH_MATH_FAST, 2, "Invalid enum value.");
    h->math_mode = mode;
    return MFH_OK;
}

int mfh_set_fast_waves(mfh_handle* h, int row_blocks, int waves)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(row_blocks >= 1 && row_blocks <= 17, 2, "row_blocks must be in [1, 17]");
    MFH_REQUIRE((waves == 0 || waves == 1 || waves == 2 || waves == 4) && !(waves == 2 && row_blocks < 2) &&
                    !(waves == 4 && row_blocks < 4),
                3, "waves must be 0 (default), 1, 2 or 4 (2 needs row_blocks >= 2, 4 needs row_blocks >= 4)");
    h->fast_waves[row_blocks] = waves;
    return MFH_OK;
}

int mfh_set_chol_shape(mfh_handle* h, int bucket, int lanes, int rows_per_lane)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(bucket >= 0 && bucket < MFH_NBUCKETS, 2, "bucket must be in [0, 8]");
    const bool reset = lanes == 0 && rows_per_lane == 0;
    MFH_REQUIRE(reset || mfh_chol_shape_ok(bucket, lanes, rows_per_lane), 3,
                "(lanes, rows_per_lane) must be a built shape covering the bucket's rows, or (0, 0)");
    h->chol_shape[bucket][0] = lanes;
    h->chol_shape[bucket][1] = rows_per_lane;
    return MFH_OK;
}

int mfh_set_bucket_streams(mfh_handle* h, int streams)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(streams >= 1 && streams <= 4, 2, "streams must be in [1, 4]");
    h->bucket_streams = streams;
    return MFH_OK;
}

int mfh_get_bucket_streams(mfh_handle* h, int* streams)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(streams != nullptr, 2, "streams must not be null");
    *streams = h->bucket_streams;
    return MFH_OK;
}

int mfh_get_fast_waves(mfh_handle* h, int row_blocks, int* waves)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(row_blocks >= 1 && row_blocks <= 17, 2, "row_blocks must be in [1, 17]");
    MFH_REQUIRE(waves != nullptr, 3, "waves must not be null");
    *waves = h->fast_waves[row_blocks] > 0 ? h->fast_waves[row_blocks] : mfh_fast_default_waves(row_blocks);
    return MFH_OK;
}

int mfh_get_math_mode(mfh_handle* h, int* mode)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(mode != nullptr, 2, "mode must not be null");
    *mode = h->math_mode;
    return MFH_OK;
}

int mfh_set_timing(mfh_handle* h, int enable)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(enable == 0 || enable == 1, 2, "enable must be 0 or 1");
    h->timing = enable != 0;
    return MFH_OK;
}

int mfh_get_timing(mfh_handle* h, int capacity, int* n_records, int* kind, int* rows, float* ms)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(capacity >= 0, 2, "capacity must be >= 0");
    MFH_REQUIRE(n_records != nullptr, 3, "n_records must not be null");
    MFH_REQUIRE(capacity == 0 || kind != nullptr, 4, "kind must not be null");
    MFH_REQUIRE(capacity == 0 || rows != nullptr, 5, "rows must not be null");
    MFH_REQUIRE(capacity == 0 || ms != nullptr, 6, "ms must not be null");
    *n_records = (int)h->timed.size();
    int st = MFH_OK;
    for (size_t i = 0; i < h->timed.size(); ++i) {
        MfhTimedLaunch& t = h->timed[i];
        if ((int)i < capacity && st == MFH_OK) {
            float e = 0.0f;
            hipError_t he = hipEventSynchronize(t.stop);
            if (he == hipSuccess) he = hipEventElapsedTime(&e, t.start, t.stop);
            if (he != hipSuccess) st = mfh_fail(MFH_ERR_HIP, std::string("timing events: ") + hipGetErrorString(he));
            kind[i] = t.kind;
            rows[i] = t.rows;
            ms[i] = e;
        }
        h->spare_events.push_back(t.start);
        h->spare_events.push_back(t.stop);
    }
    h->timed.clear();
    return st;
}

int mfh_get_stream(mfh_handle* h, void** stream)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(stream != nullptr, 2, "stream must not be null");
    *stream = (void*)h->stream;
    return MFH_OK;
}

int mfh_synchronize(mfh_handle* h)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_HIP_CHECK(hipStreamSynchronize(h->stream));
    return MFH_OK;
}

int mfh_dgetrf_vbatched(mfh_handle* h, int count, const int* m, const int* n, double* const* a, const int* lda,
                        int* const* piv, int* sign, int* info, int max_m, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(m != nullptr, 3, "m must not be null");
    MFH_REQUIRE(n != nullptr, 4, "n must not be null");
    MFH_REQUIRE(a != nullptr, 5, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 6, "lda must not be null");
    MFH_REQUIRE(piv != nullptr, 7, "piv must not be null");
    MFH_REQUIRE(sign != nullptr, 8, "sign must not be null");
    MFH_REQUIRE(info != nullptr, 9, "info must not be null");
    MFH_REQUIRE(max_m >= 1 && max_m <= MFH_EXACT_MAX_N, 10, "max_m must be in [1, MFH_EXACT_MAX_N]");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 11, "max_n must be in [1, MFH_EXACT_MAX_N]");
    MfhGetrfBatch B{m, n, a, lda, piv, 0, 0, 0, nullptr, 0, nullptr, 0, sign, info};
    return mfh_launch_getrf(h, count, B, max_m, max_n);
}

int mfh_dgetrf_strided_batched(mfh_handle* h, int count, int m, int n, double* a, int lda, int64_t stride_a, int* piv,
                               int64_t stride_piv, int* sign, int* info)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(m >= 1 && m <= MFH_EXACT_MAX_N, 3, "The number of rows must be in [1, MFH_EXACT_MAX_N].");
    MFH_REQUIRE(n >= 1 && n <= MFH_EXACT_MAX_N, 4, "The number of columns must be in [1, MFH_EXACT_MAX_N].");
    MFH_REQUIRE(a != nullptr, 5, "a must not be null");
    MFH_REQUIRE(lda >= m, 6, "The leading dimension must be greater than or equal to the number of rows.");
    MFH_REQUIRE(stride_a >= (int64_t)lda * n, 7, "stride_a must be >= lda*n");
    MFH_REQUIRE(piv != nullptr, 8, "piv must not be null");
    MFH_REQUIRE(stride_piv >= m, 9, "stride_piv must be >= m");
    MFH_REQUIRE(sign != nullptr, 10, "sign must not be null");
    MFH_REQUIRE(info != nullptr, 11, "info must not be null");
    MfhGetrfBatch B{nullptr, nullptr, nullptr, nullptr, nullptr, m, n, lda, a, stride_a, piv, stride_piv, sign, info};
    return mfh_launch_getrf(h, count, B, m, n);
}

int mfh_dpotrf_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, int* info,
                        int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(info != nullptr, 6, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 7, "max_n must be in [1, MFH_EXACT_MAX_N]");
    MfhCholBatch B{n, a, lda, nullptr, info};
    return mfh_launch_potrf(h, count, B, max_n, false);
}

int mfh_dposv_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, double* const* x,
                       int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(x != nullptr, 6, "x must not be null");
    MFH_REQUIRE(info != nullptr, 7, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 8, "max_n must be in [1, MFH_EXACT_MAX_N]");
    MfhCholBatch B{n, a, lda, x, info};
    return mfh_launch_potrf(h, count, B, max_n, true);
}

int mfh_dposv_vbatched_noalloc(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                               double* const* x, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    MFH_REQUIRE(count <= MFH_NOALLOC_MAX_COUNT, 2, "count must be <= MFH_NOALLOC_MAX_COUNT (the static scratch size)");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(h->scratch != nullptr, 1, "the handle has no scratch address");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_SMALL_MAX_N, 8, "max_n must be in [1, MFH_SMALL_MAX_N]");
    // The scratch replaces the handle workspace for this call only; with
    // no_alloc set, nothing below allocates, creates a stream or an event:
    // the size buckets run in order unless the handle's side streams exist.
    void* saved_ws = h->ws;
    const size_t saved_bytes = h->ws_bytes;
    h->ws = h->scratch;
    h->ws_bytes = h->scratch_bytes;
    h->no_alloc = true;
    const int st = mfh_dposv_vbatched(h, count, n, a, lda, x, info, max_n);
    h->no_alloc = false;
    h->ws = saved_ws;
    h->ws_bytes = saved_bytes;
    return st;
}

int mfh_dtrsv_vbatched(mfh_handle* h, int uplo, int trans, int diag, int count, const int* n, const double* const* a,
                       const int* lda, double* const* x, const int* incx, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(uplo == MFH_UPPER || uplo == MFH_LOWER, 2, "Invalid enum value.");
    MFH_REQUIRE(trans == MFH_NOTRANS || trans == MFH_TRANS, 3, "Invalid enum value.");
    MFH_REQUIRE(diag == MFH_NONUNIT || diag == MFH_UNIT, 4, "Invalid enum value.");
    MFH_REQUIRE(count >= 0, 5, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 6, "n must not be null");
    MFH_REQUIRE(a != nullptr, 7, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 8, "lda must not be null");
    MFH_REQUIRE(x != nullptr, 9, "x must not be null");
    MFH_REQUIRE(incx != nullptr, 10, "incx must not be null");
    MFH_REQUIRE(info != nullptr, 11, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 12, "max_n must be in [1, MFH_EXACT_MAX_N]");
    MfhTrsvBatch B{n, a, lda, x, incx, info, uplo, trans, diag == MFH_UNIT ? 1 : 0};
    return mfh_launch_trsv(h, count, B, max_n);
}

int mfh_dgetrs_vbatched(mfh_handle* h, int count, const int* n, const double* const* lu, const int* lda,
                        const int* const* piv, const double* const* b, double* const* x, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(lu != nullptr, 4, "lu must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(piv != nullptr, 6, "piv must not be null");
    MFH_REQUIRE(b != nullptr, 7, "b must not be null");
    MFH_REQUIRE(x != nullptr, 8, "x must not be null");
    MFH_REQUIRE(info != nullptr, 9, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 10, "max_n must be in [1, MFH_EXACT_MAX_N]");
    MfhGetrsBatch B{n, lu, lda, piv, b, x, info};
    return mfh_launch_getrs(h, count, B, max_n);
}

int mfh_dgetri_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                        const int* const* piv, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(piv != nullptr, 6, "piv must not be null");
    MFH_REQUIRE(info != nullptr, 7, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_SMALL_MAX_N, 8, "max_n must be in [1, MFH_SMALL_MAX_N]");
    MFH_REQUIRE((long long)count * max_n <= 0x7fffffffLL, 2, "count * max_n must be < 2^31");
    MfhGetriBatch B{n, a, lda, piv, info};
    return mfh_launch_getri(h, count, B, max_n);
}

int mfh_zgetrf_vbatched(mfh_handle* h, int count, const int* m, const int* n, double* const* a, const int* lda,
                        int* const* piv, int* sign, int* info, int max_m, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(m != nullptr, 3, "m must not be null");
    MFH_REQUIRE(n != nullptr, 4, "n must not be null");
    MFH_REQUIRE(a != nullptr, 5, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 6, "lda must not be null");
    MFH_REQUIRE(piv != nullptr, 7, "piv must not be null");
    MFH_REQUIRE(sign != nullptr, 8, "sign must not be null");
    MFH_REQUIRE(info != nullptr, 9, "info must not be null");
    MFH_REQUIRE(max_m >= 1 && max_m <= MFH_EXACT_MAX_N, 10, "max_m must be in [1, MFH_EXACT_MAX_N]");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 11, "max_n must be in [1, MFH_EXACT_MAX_N]");
    return mfh_launch_zgetrf(h, count, m, n, a, lda, piv, sign, info, max_m, max_n);
}

int mfh_zpotrf_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, int* info,
                        int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(info != nullptr, 6, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 7, "max_n must be in [1, MFH_EXACT_MAX_N]");
    return mfh_launch_zpotrf(h, count, n, a, lda, info, max_n);
}

int mfh_ztrsv_vbatched(mfh_handle* h, int uplo, int trans, int count, const int* n, const double* const* a,
                       const int* lda, double* const* x, const int* incx, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(uplo == MFH_UPPER || uplo == MFH_LOWER, 2, "Invalid enum value.");
    MFH_REQUIRE(trans >= MFH_NOTRANS && trans <= MFH_CONJNOTRANS, 3, "Invalid enum value.");
    MFH_REQUIRE(count >= 0, 4, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 5, "n must not be null");
    MFH_REQUIRE(a != nullptr, 6, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 7, "lda must not be null");
    MFH_REQUIRE(x != nullptr, 8, "x must not be null");
    MFH_REQUIRE(incx != nullptr, 9, "incx must not be null");
    MFH_REQUIRE(info != nullptr, 10, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 11, "max_n must be in [1, MFH_EXACT_MAX_N]");
    return mfh_launch_ztrsv(h, uplo, trans, count, n, a, lda, x, incx, info, max_n);
}

static int gemm_common(mfh_handle* h, int count, const int* transa, const int* transb, const int* m, const int* n,
                       const int* k, const double* alpha, const double* const* a, const int* lda,
                       const double* const* b, const int* ldb, const double* beta, double* const* c,
                       const int* ldc, int* info, int max_m, int max_n, int max_k, bool cplx)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(transa != nullptr, 3, "transa must not be null");
    MFH_REQUIRE(transb != nullptr, 4, "transb must not be null");
    MFH_REQUIRE(m != nullptr, 5, "m must not be null");
    MFH_REQUIRE(n != nullptr, 6, "n must not be null");
    MFH_REQUIRE(k != nullptr, 7, "k must not be null");
    MFH_REQUIRE(alpha != nullptr, 8, "alpha must not be null");
    MFH_REQUIRE(a != nullptr, 9, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 10, "lda must not be null");
    MFH_REQUIRE(b != nullptr, 11, "b must not be null");
    MFH_REQUIRE(ldb != nullptr, 12, "ldb must not be null");
    MFH_REQUIRE(beta != nullptr, 13, "beta must not be null");
    MFH_REQUIRE(c != nullptr, 14, "c must not be null");
    MFH_REQUIRE(ldc != nullptr, 15, "ldc must not be null");
    MFH_REQUIRE(info != nullptr, 16, "info must not be null");
    MFH_REQUIRE(max_m >= 1 && max_m <= 4096, 17, "max_m must be in [1, 4096]");
    MFH_REQUIRE(max_n >= 1 && max_n <= 4096, 18, "max_n must be in [1, 4096]");
    MFH_REQUIRE(max_k >= 1, 19, "max_k must be >= 1");
    MfhGemmBatch B{transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, info};
    return mfh_launch_gemm(h, count, B, max_m, max_n, max_k, cplx);
}

int mfh_dgemm_vbatched(mfh_handle* h, int count, const int* transa, const int* transb, const int* m, const int* n,
                       const int* k, const double* alpha, const double* const* a, const int* lda,
                       const double* const* b, const int* ldb, const double* beta, double* const* c, const int* ldc,
                       int* info, int max_m, int max_n, int max_k)
{
    return gemm_common(h, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, info, max_m, max_n,
                       max_k, false);
}

int mfh_zgemm_vbatched(mfh_handle* h, int count, const int* transa, const int* transb, const int* m, const int* n,
                       const int* k, const double* alpha, const double* const* a, const int* lda,
                       const double* const* b, const int* ldb, const double* beta, double* const* c, const int* ldc,
                       int* info, int max_m, int max_n, int max_k)
{
    return gemm_common(h, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, info, max_m, max_n,
                       max_k, true);
}

int mfh_dgemv_vbatched(mfh_handle* h, int count, const int* trans, const int* m, const int* n,
                       const double* const* a, const int* lda, const double* const* x, const int* incx,
                       double* const* y, const int* incy, int* info, int max_m, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(trans != nullptr, 3, "trans must not be null");
    MFH_REQUIRE(m != nullptr, 4, "m must not be null");
    MFH_REQUIRE(n != nullptr, 5, "n must not be null");
    MFH_REQUIRE(a != nullptr, 6, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 7, "lda must not be null");
    MFH_REQUIRE(x != nullptr, 8, "x must not be null");
    MFH_REQUIRE(incx != nullptr, 9, "incx must not be null");
    MFH_REQUIRE(y != nullptr, 10, "y must not be null");
    MFH_REQUIRE(incy != nullptr, 11, "incy must not be null");
    MFH_REQUIRE(info != nullptr, 12, "info must not be null");
    MFH_REQUIRE(max_m >= 1, 13, "max_m must be >= 1");
    MFH_REQUIRE(max_n >= 1, 14, "max_n must be >= 1");
    MfhGemvBatch B{trans, m, n, a, lda, x, incx, y, incy, info};
    return mfh_launch_gemv(h, count, B, max_m, max_n);
}

int mfh_dgesv_mixed_vbatched(mfh_handle* h, int count, const int* n, const double* const* a, const int* lda,
                             const double* const* b, double* const* x, int max_iter, double tol_per_n, int* iters,
                             double* berr, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(b != nullptr, 6, "b must not be null");
    MFH_REQUIRE(x != nullptr, 7, "x must not be null");
    MFH_REQUIRE(max_iter >= 0 && max_iter <= 100, 8, "max_iter must be in [0, 100]");
    MFH_REQUIRE(tol_per_n > 0.0, 9, "tol_per_n must be > 0");
    MFH_REQUIRE(iters != nullptr, 10, "iters must not be null");
    MFH_REQUIRE(berr != nullptr, 11, "berr must not be null");
    MFH_REQUIRE(info != nullptr, 12, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_BLOCKED_MAX_N, 13, "max_n must be in [1, MFH_BLOCKED_MAX_N]");
    return mfh_launch_gesv_mixed(h, count, n, a, lda, b, x, max_iter, tol_per_n, iters, berr, info, max_n, true);
}

int mfh_dgesv_vbatched(mfh_handle* h, int count, const int* n, const double* const* a, const int* lda,
                       const double* const* b, double* const* x, double* berr, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(b != nullptr, 6, "b must not be null");
    MFH_REQUIRE(x != nullptr, 7, "x must not be null");
    MFH_REQUIRE(berr != nullptr, 8, "berr must not be null");
    MFH_REQUIRE(info != nullptr, 9, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_BLOCKED_MAX_N, 10, "max_n must be in [1, MFH_BLOCKED_MAX_N]");
    return mfh_launch_gesv_mixed(h, count, n, a, lda, b, x, 0, 1.0, nullptr, berr, info, max_n, false);
}

int mfh_dgetrf_blocked_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda,
                                int* const* piv, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(piv != nullptr, 6, "piv must not be null");
    MFH_REQUIRE(info != nullptr, 7, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_BLOCKED_MAX_N, 8, "max_n must be in [1, MFH_BLOCKED_MAX_N]");
    return mfh_launch_getrf_blocked(h, count, n, a, lda, piv, info, max_n);
}

int mfh_dposv_backward_vbatched(mfh_handle* h, int count, const int* n, const double* const* l, const int* lda,
                                const double* const* x, const double* const* xbar, double* const* abar,
                                const int* ldabar, double* const* bbar, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(l != nullptr, 4, "l must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(x != nullptr, 6, "x must not be null");
    MFH_REQUIRE(xbar != nullptr, 7, "xbar must not be null");
    MFH_REQUIRE(abar != nullptr, 8, "abar must not be null");
    MFH_REQUIRE(ldabar != nullptr, 9, "ldabar must not be null");
    MFH_REQUIRE(bbar != nullptr, 10, "bbar must not be null");
    MFH_REQUIRE(info != nullptr, 11, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_SMALL_MAX_N, 12, "max_n must be in [1, MFH_SMALL_MAX_N]");
    MfhAdjointBatch B{n, l, lda, x, xbar, abar, ldabar, bbar, info};
    return mfh_launch_posv_backward(h, count, B, max_n);
}

int mfh_dpotrf_backward_vbatched(mfh_handle* h, int count, const int* n, const double* const* l, const int* ldl,
                                 const double* const* lbar, const int* ldlbar, double* const* abar,
                                 const int* ldabar, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(l != nullptr, 4, "l must not be null");
    MFH_REQUIRE(ldl != nullptr, 5, "ldl must not be null");
    MFH_REQUIRE(lbar != nullptr, 6, "lbar must not be null");
    MFH_REQUIRE(ldlbar != nullptr, 7, "ldlbar must not be null");
    MFH_REQUIRE(abar != nullptr, 8, "abar must not be null");
    MFH_REQUIRE(ldabar != nullptr, 9, "ldabar must not be null");
    MFH_REQUIRE(info != nullptr, 10, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_SMALL_MAX_N, 11, "max_n must be in [1, MFH_SMALL_MAX_N]");
    MfhCholAdjointBatch B{n, l, ldl, lbar, ldlbar, abar, ldabar, info};
    return mfh_launch_potrf_backward(h, count, B, max_n);
}

int mfh_dgevd_vbatched(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, double* const* b,
                       const int* ldb, double* const* w, int* info, int max_n)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 0, 2, "count must be >= 0");
    if (count == 0) return MFH_OK;
    MFH_REQUIRE(n != nullptr, 3, "n must not be null");
    MFH_REQUIRE(a != nullptr, 4, "a must not be null");
    MFH_REQUIRE(lda != nullptr, 5, "lda must not be null");
    MFH_REQUIRE(b != nullptr, 6, "b must not be null");
    MFH_REQUIRE(ldb != nullptr, 7, "ldb must not be null");
    MFH_REQUIRE(w != nullptr, 8, "w must not be null");
    MFH_REQUIRE(info != nullptr, 9, "info must not be null");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_SMALL_MAX_N, 10, "max_n must be in [1, MFH_SMALL_MAX_N]");
    return mfh_launch_gevd(h, count, n, a, lda, b, ldb, w, info, max_n);
}

// ---------------------------------------------------------------- plans
}  // extern "C"

struct mfh_plan {
    mfh_handle* h = nullptr;
    int op = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* ws = nullptr;   // the plan's own workspace and second buffer: later calls on the handle may regrow
    void* aux = nullptr;  // the handle's, and the graph must keep pointing at live memory
};

// Records one call into a hipGraph.  The plan owns a workspace and a second
// buffer sized exactly as the launchers ask for them; the capture runs on a
// private stream (the handle's stream may be the null stream, which cannot be
// captured) with the handle temporarily pointing at the plan's buffers.
template <typename F>
static int capture_plan(mfh_handle* h, int op, size_t ws_bytes, size_t aux_bytes, F call, mfh_plan** plan)
{
    // side streams of the bucket fork exist before the capture starts
    int sst = mfh_side_streams(h, h->bucket_streams);
    if (sst) return sst;
    void* pws = nullptr;
    void* paux = nullptr;
    MFH_HIP_CHECK(hipMalloc(&pws, ws_bytes < 256 ? 256 : ws_bytes));
    if (aux_bytes) {
        hipError_t ea = hipMalloc(&paux, aux_bytes);
        if (ea != hipSuccess) {
            hipFree(pws);
            return mfh_fail(MFH_ERR_ALLOC, std::string("hipMalloc plan buffer: ") + hipGetErrorString(ea));
        }
    }
    hipStream_t cap;
    hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e != hipSuccess) {
        hipFree(pws);
        if (paux) hipFree(paux);
        return mfh_fail(MFH_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipStream_t saved = h->stream;
    void* saved_ws = h->ws;
    size_t saved_bytes = h->ws_bytes;
    void* saved_aux = h->aux;
    size_t saved_aux_bytes = h->aux_bytes;
    h->stream = cap;
    h->ws = pws;
    h->ws_bytes = ws_bytes;
    h->aux = paux;
    h->aux_bytes = aux_bytes;
    h->capturing = true;
    int st = MFH_OK;
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        st = call();
        e = hipStreamEndCapture(cap, &graph);
    }
    h->stream = saved;
    h->ws = saved_ws;
    h->ws_bytes = saved_bytes;
    h->aux = saved_aux;
    h->aux_bytes = saved_aux_bytes;
    h->capturing = false;
    hipStreamDestroy(cap);
    hipGraphExec_t exec = nullptr;
    if (st == MFH_OK && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (st != MFH_OK || e != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        hipFree(pws);
        if (paux) hipFree(paux);
        if (st != MFH_OK) return st;
        return mfh_fail(MFH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    mfh_plan* p = new mfh_plan();
    p->h = h;
    p->op = op;
    p->graph = graph;
    p->exec = exec;
    p->ws = pws;
    p->aux = paux;
    *plan = p;
    return MFH_OK;
}

extern "C" {

int mfh_plan_create_dposv(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, double* const* x,
                          int* info, int max_n, mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(plan != nullptr, 9, "plan must not be null");
    MFH_REQUIRE(count >= 1, 2, "count must be >= 1");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 8, "max_n must be in [1, MFH_EXACT_MAX_N]");
    return capture_plan(
        h, MFH_OP_DPOSV, mfh_exact_ws_bytes(count), mfh_potrf_gen_aux_bytes(h, count, max_n, false),
        [&]() { return mfh_dposv_vbatched(h, count, n, a, lda, x, info, max_n); }, plan);
}

int mfh_plan_create_dgetrf(mfh_handle* h, int count, const int* m, const int* n, double* const* a, const int* lda,
                           int* const* piv, int* sign, int* info, int max_m, int max_n, mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(plan != nullptr, 12, "plan must not be null");
    MFH_REQUIRE(count >= 1, 2, "count must be >= 1");
    return capture_plan(
        h, MFH_OP_DGETRF, mfh_exact_ws_bytes(count), 0,
        [&]() { return mfh_dgetrf_vbatched(h, count, m, n, a, lda, piv, sign, info, max_m, max_n); }, plan);
}

int mfh_plan_create_dpotrf(mfh_handle* h, int count, const int* n, double* const* a, const int* lda, int* info,
                           int max_n, mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(plan != nullptr, 8, "plan must not be null");
    MFH_REQUIRE(count >= 1, 2, "count must be >= 1");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_EXACT_MAX_N, 7, "max_n must be in [1, MFH_EXACT_MAX_N]");
    return capture_plan(
        h, MFH_OP_DPOTRF, mfh_exact_ws_bytes(count), mfh_potrf_gen_aux_bytes(h, count, max_n, false),
        [&]() { return mfh_dpotrf_vbatched(h, count, n, a, lda, info, max_n); }, plan);
}

int mfh_plan_create_dtrsv(mfh_handle* h, int uplo, int trans, int diag, int count, const int* n,
                          const double* const* a, const int* lda, double* const* x, const int* incx, int* info,
                          int max_n, mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(plan != nullptr, 13, "plan must not be null");
    MFH_REQUIRE(count >= 1, 5, "count must be >= 1");
    return capture_plan(
        h, MFH_OP_DTRSV, mfh_exact_ws_bytes(count), 0,
        [&]() { return mfh_dtrsv_vbatched(h, uplo, trans, diag, count, n, a, lda, x, incx, info, max_n); }, plan);
}

int mfh_plan_create_dgesv(mfh_handle* h, int count, const int* n, const double* const* a, const int* lda,
                          const double* const* b, double* const* x, double* berr, int* info, int max_n,
                          mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(plan != nullptr, 11, "plan must not be null");
    MFH_REQUIRE(count >= 1, 2, "count must be >= 1");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_BLOCKED_MAX_N, 10, "max_n must be in [1, MFH_BLOCKED_MAX_N]");
    return capture_plan(
        h, MFH_OP_DGESV, mfh_gesv_ws_bytes(count, max_n, false), 0,
        [&]() { return mfh_dgesv_vbatched(h, count, n, a, lda, b, x, berr, info, max_n); }, plan);
}

int mfh_plan_create_dgesv_mixed(mfh_handle* h, int count, const int* n, const double* const* a, const int* lda,
                                const double* const* b, double* const* x, int max_iter, double tol_per_n,
                                int* iters, double* berr, int* info, int max_n, mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(plan != nullptr, 14, "plan must not be null");
    MFH_REQUIRE(count >= 1, 2, "count must be >= 1");
    MFH_REQUIRE(max_n >= 1 && max_n <= MFH_BLOCKED_MAX_N, 13, "max_n must be in [1, MFH_BLOCKED_MAX_N]");
    return capture_plan(
        h, MFH_OP_DGESV_MIXED, mfh_gesv_ws_bytes(count, max_n, true), 0,
        [&]() {
            return mfh_dgesv_mixed_vbatched(h, count, n, a, lda, b, x, max_iter, tol_per_n, iters, berr, info,
                                            max_n);
        },
        plan);
}

static int plan_gemm(mfh_handle* h, int count, const int* transa, const int* transb, const int* m, const int* n,
                     const int* k, const double* alpha, const double* const* a, const int* lda,
                     const double* const* b, const int* ldb, const double* beta, double* const* c, const int* ldc,
                     int* info, int max_m, int max_n, int max_k, bool cplx, mfh_plan** plan)
{
    MFH_REQUIRE(h != nullptr, 1, "handle must not be null");
    MFH_REQUIRE(count >= 1, 2, "count must be >= 1");
    MFH_REQUIRE(plan != nullptr, 20, "plan must not be null");
    MFH_REQUIRE(max_m >= 1 && max_m <= 4096, 17, "max_m must be in [1, 4096]");
    MFH_REQUIRE(max_n >= 1 && max_n <= 4096, 18, "max_n must be in [1, 4096]");
    const size_t ws = mfh_gemm_ws_bytes(count, max_m, max_n, h->math_mode == MFH_MATH_FAST);
    return capture_plan(
        h, cplx ? MFH_OP_ZGEMM : MFH_OP_DGEMM, ws, 0,
        [&]() {
            return cplx ? mfh_zgemm_vbatched(h, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                                             info, max_m, max_n, max_k)
                        : mfh_dgemm_vbatched(h, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                                             info, max_m, max_n, max_k);
        },
        plan);
}

int mfh_plan_create_dgemm(mfh_handle* h, int count, const int* transa, const int* transb, const int* m, const int* n,
                          const int* k, const double* alpha, const double* const* a, const int* lda,
                          const double* const* b, const int* ldb, const double* beta, double* const* c,
                          const int* ldc, int* info, int max_m, int max_n, int max_k, mfh_plan** plan)
{
    return plan_gemm(h, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, info, max_m, max_n,
                     max_k, false, plan);
}

int mfh_plan_create_zgemm(mfh_handle* h, int count, const int* transa, const int* transb, const int* m, const int* n,
                          const int* k, const double* alpha, const double* const* a, const int* lda,
                          const double* const* b, const int* ldb, const double* beta, double* const* c,
                          const int* ldc, int* info, int max_m, int max_n, int max_k, mfh_plan** plan)
{
    return plan_gemm(h, count, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, info, max_m, max_n,
                     max_k, true, plan);
}

int mfh_plan_get_op(mfh_plan* plan, int* op)
{
    MFH_REQUIRE(plan != nullptr, 1, "plan must not be null");
    MFH_REQUIRE(op != nullptr, 2, "op must not be null");
    *op = plan->op;
    return MFH_OK;
}

int mfh_plan_execute(mfh_plan* plan)
{
    MFH_REQUIRE(plan != nullptr, 1, "plan must not be null");
    MFH_HIP_CHECK(hipGraphLaunch(plan->exec, plan->h->stream));
    return MFH_OK;
}

int mfh_plan_destroy(mfh_plan* plan)
{
    MFH_REQUIRE(plan != nullptr, 1, "plan must not be null");
    if (plan->exec) hipGraphExecDestroy(plan->exec);
    if (plan->graph) hipGraphDestroy(plan->graph);
    if (plan->ws || plan->aux) hipStreamSynchronize(plan->h->stream);
    if (plan->aux) hipFree(plan->aux);
    if (plan->ws) hipFree(plan->ws);
    delete plan;
    return MFH_OK;
}

}  // extern "C"
