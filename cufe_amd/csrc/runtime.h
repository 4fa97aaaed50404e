// runtime.h -- device state shared by the C ABI: uploaded constant tables,
// device buffers with RAII, and the Gaussian-sampler driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "host_tables.h"
#include "kernels.h"

namespace ipfe {

extern const uint32_t *g_aes_te;

// Per-process device state (one GPU, the one current at the library's first
// call): constant tables resident in HBM, in static device arrays (set-up
// allocates nothing).
struct Device {
    int id = 0;                  // the HIP device the state belongs to
    uint32_t *tw_all = nullptr;  // [4][3][N]
    uint32_t *aes_te = nullptr;  // [1024]
    NinvConsts ninv;
    ScaleConsts scale;
    BarrettConsts barrett;
    hipStream_t stream = nullptr;
};
// Initialises on first use; returns nullptr (error set) on failure.
Device *device();

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count)
    {
        reset();
        if (count == 0) return 0;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            ipfe_set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
            return -1;
        }
        n = count;
        return 0;
    }
    T *get() const { return p; }
};

// One Gaussian stream request for the sampler driver.
struct StreamReq {
    uint32_t rk[60];
    uint16_t nonce;
    int which;  // sampler width index (GaussStream::which)
    int ncalls;
    uint32_t *out0;
    uint32_t *out1;
    long long out_stride;
    int coef0_only;
    int slen;          // samples per call
    int nmod;          // output limbs and their moduli / fold shifts
    uint32_t q[4], fold[4];
};
// A stream of the reference parameter set (4096 samples per call, 3 limbs);
// which = 1, 2, 3 for S1, S2, S3.
StreamReq make_stream(const uint8_t key[32], uint16_t nonce, int which, int ncalls, uint32_t *out0, uint32_t *out1,
                      long long out_stride, int coef0_only = 0);
struct PsetSpec;
// A stream of parameter set ps: sampler which = 1, 2, 3 at that set's widths.
StreamReq make_stream_ps(const PsetSpec &ps, const uint8_t key[32], uint16_t nonce, int which, int ncalls,
                         uint32_t *out0, uint32_t *out1, long long out_stride);

// Samples every request (device outputs), bit-exact with the reference
// sequential sampler; splits into groups that fit scratch_budget bytes.
int run_gauss(const std::vector<StreamReq> &reqs, hipStream_t st, size_t scratch_budget = (size_t)8 << 30);

// One group of sampler streams, split into host planning, table upload, the
// kernel launches (no host round trip, so a graph can capture them) and the
// status check that reports a speculative window that was too short.
struct GaussPlan {
    int n = 0, level = 0;
    int full = 0;  // all 24 AES blocks per chunk (set after a chunk needed them)
    long long tiles = 0, chunks = 0, full_chunks = 0, nfirst = 0;
    std::vector<GaussStream> hs;
    std::vector<long long> tile_prefix, chunk_prefix, first_prefix;
    DevBuf<GaussStream> d_streams;
    DevBuf<long long> d_tp, d_cp, d_fp, d_first;
    DevBuf<int> d_status;
    // device metadata the kernels read: the buffers above, or a region bound
    // with bind_meta
    GaussStream *m_streams = nullptr;
    long long *m_tp = nullptr, *m_cp = nullptr, *m_fp = nullptr, *m_first = nullptr;
    int *m_status = nullptr;
    DevBuf<uint8_t> own_cnt;
    DevBuf<uint64_t> own_vals;
    DevBuf<int> own_pos;
    uint8_t *s_cnt = nullptr;
    uint64_t *s_vals = nullptr;
    int *s_pos = nullptr;

    // alloc_meta false: plan only; the caller binds the device metadata
    // (bind_meta) and the chunk scratch (bind_scratch_region) to its memory
    int build(const StreamReq *reqs, int n, int level, bool alloc_meta = true);
    size_t meta_bytes() const;
    int bind_meta(uint8_t *base);
    // the chunk scratch as one region of 256-byte aligned pieces
    size_t bound_scratch_bytes() const;
    int bind_scratch_region(uint8_t *base);
    // chunk scratch: own allocation, or a caller's buffer of scratch_bytes()
    // (chunks counts, then 8 values and 1 position per full-stream chunk)
    size_t scratch_bytes() const;
    int alloc_scratch();
    int bind_scratch(uint8_t *cnt, uint64_t *vals, int *pos);
    void rekey(const StreamReq *reqs);
    int upload(hipStream_t st);
    int launch(hipStream_t st);
    // overflow: a window was short; need_full: a chunk needed AES blocks the
    // default set leaves out (rerun with full = 1)
    int check(hipStream_t st, int *overflow, int *need_full);
};

// Speculative window (chunks) for a stream of ncalls calls of slen samples of
// sampler width `which` at margin level m, and the scratch bytes one stream
// needs at level 0 (the three-argument forms: 4096 samples per call).
long long gauss_window(int which, int slen, int ncalls, int m);
long long gauss_window(int which, int ncalls, int m);
size_t gauss_bytes(int which, int slen, int ncalls, int coef0_only);
size_t gauss_bytes(int which, int ncalls, int coef0_only);

int sync_stream(hipStream_t st);

// [p, p + bytes) lies inside one block from cufe_host_alloc (pinned).
bool host_pinned(const void *p, size_t bytes);

}  // namespace ipfe
