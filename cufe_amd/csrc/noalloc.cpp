// noalloc.cpp -- cufe_gauss_sample_noalloc: the Gaussian sampling of
// cufe_gauss_sample for callers that cannot allocate device memory (code under
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The caller passes the keys, nonces and a device output buffer it owns.
// Every other device buffer the sampler's kernels use -- stream descriptors,
// tile / chunk prefixes, status words and the per-chunk counts, accepted
// values and call labels -- is carved out of `scratch`, a static device array
// of kScratchBytes.  Host code never passes the array's name to a kernel: the
// array's device address comes from hipGetSymbolAddress, resolved once per
// device with that device current, and only that address is used.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; nothing it reads is assumed to start at zero (the status words and
// labels are cleared by fill kernels in each launch, counts are written for
// every chunk before they are read, values only where they are read).  Calls
// from several threads of one process are serialised by a mutex; calls that
// overlap on one GPU from other code paths sharing the array would race.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/cufe_ipfe.h"
#include "runtime.h"

using namespace ipfe;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU, held until the process ends
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;                         // serialises calls (one array per device)
uint8_t *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// Device address of the current device's copy of `scratch` (g_mu held).
uint8_t *scratch_address(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) {
        ipfe_set_error("cufe_gauss_sample_noalloc: device %d out of range", dev);
        return nullptr;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ipfe_set_error("cufe_gauss_sample_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                           hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (uint8_t *)p;
    }
    return g_scratch_ptr[dev];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes a plan of these streams binds at this level (metadata + chunk scratch).
size_t plan_bytes(GaussPlan &p, const StreamReq *reqs, int n, int level)
{
    p.build(reqs, n, level, /*alloc_meta=*/false);
    return align256(p.meta_bytes()) + p.bound_scratch_bytes();
}

}  // namespace

extern "C" int cufe_gauss_sample_noalloc(const unsigned char *keys, const uint16_t *nonces, int nstreams, int which,
                                         int ncalls, uint32_t *d_out)
{
    if (which < 1 || which > 3 || ncalls < 0 || nstreams < 0 || (nstreams > 0 && (!keys || !nonces))) {
        ipfe_set_error("cufe_gauss_sample_noalloc: bad arguments (which=%d ncalls=%d nstreams=%d)", which, ncalls,
                       nstreams);
        return -1;
    }
    if (ncalls == 0 || nstreams == 0) return 0;
    if (!d_out) {
        ipfe_set_error("cufe_gauss_sample_noalloc: null output");
        return -1;
    }
    Device *dev = device();  // set-up allocates nothing (static tables)
    if (!dev) return -1;
    int cur = 0;
    HIP_OK(hipGetDevice(&cur));
    if (cur != dev->id) {
        ipfe_set_error("cufe_gauss_sample_noalloc: the library runs on device %d, current device is %d", dev->id,
                       cur);
        return -1;
    }
    // one stream's window must fit the array (planning only, on the host)
    {
        GaussPlan probe;
        const StreamReq rq = make_stream(keys, nonces[0], which, ncalls, nullptr, nullptr, 0);
        const size_t one = plan_bytes(probe, &rq, 1, 0);
        if (one > kScratchBytes) {
            ipfe_set_error("cufe_gauss_sample_noalloc: one stream of %d calls needs %zu bytes of scratch, more than "
                           "the %zu-byte maximum",
                           ncalls, one, kScratchBytes);
            return -1;
        }
    }
    // the output must be device memory the caller owns
    hipPointerAttribute_t attr;
    memset(&attr, 0, sizeof attr);
    if (hipPointerGetAttributes(&attr, d_out) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        ipfe_set_error("cufe_gauss_sample_noalloc: the output is not device memory");
        return -1;
    }
    const size_t per_stream = (size_t)ncalls * 3 * IPFE_N;
    // ... and hold the whole output: the kernels write [nstreams][ncalls][3][N]
    hipDeviceptr_t alloc_base = nullptr;
    size_t alloc_size = 0;
    if (hipMemGetAddressRange(&alloc_base, &alloc_size, (hipDeviceptr_t)d_out) != hipSuccess) {
        (void)hipGetLastError();
        ipfe_set_error("cufe_gauss_sample_noalloc: the output's allocation is unknown");
        return -1;
    }
    const size_t out_bytes = (size_t)nstreams * per_stream * 4;
    if ((char *)d_out < (char *)alloc_base || (size_t)((char *)d_out - (char *)alloc_base) + out_bytes > alloc_size) {
        ipfe_set_error("cufe_gauss_sample_noalloc: the output needs %zu bytes; its allocation holds %zu from there",
                       out_bytes, alloc_size - (size_t)((char *)d_out - (char *)alloc_base));
        return -1;
    }
    std::vector<StreamReq> reqs;
    reqs.reserve((size_t)nstreams);
    for (int t = 0; t < nstreams; ++t) {
        uint32_t *o = d_out + (size_t)t * per_stream;
        reqs.push_back(make_stream(keys + 32 * (size_t)t, nonces[t], which, ncalls, o, o + 3 * IPFE_N,
                                   (long long)3 * IPFE_N));
    }
    std::lock_guard<std::mutex> lk(g_mu);
    uint8_t *base = scratch_address(cur);
    if (!base) return -1;
    hipStream_t st = dev->stream;
    GaussPlan probe;
    const size_t one = plan_bytes(probe, reqs.data(), 1, 0);
    // streams are sampled in groups whose plan fits the array, one after another
    int i = 0;
    while (i < nstreams) {
        int j = std::min(nstreams, i + std::max(1, (int)(kScratchBytes / one)));
        GaussPlan plan;
        while (j > i + 1 && plan_bytes(plan, reqs.data() + i, j - i, 0) > kScratchBytes) j = i + (j - i) / 2;
        for (int level = 0;;) {
            if (plan_bytes(plan, reqs.data() + i, j - i, level) > kScratchBytes) {
                if (j - i > 1) {  // a widened window: fewer streams per group
                    j = i + (j - i) / 2;
                    continue;
                }
                ipfe_set_error("cufe_gauss_sample_noalloc: the widened window of retry %d exceeds the %zu-byte "
                               "scratch",
                               level, kScratchBytes);
                return -1;
            }
            plan.bind_meta(base);
            plan.bind_scratch_region(base + align256(plan.meta_bytes()));
            int overflow = 0;
            for (;;) {
                int need_full = 0;
                if (plan.upload(st) || plan.launch(st) || plan.check(st, &overflow, &need_full)) return -1;
                if (!need_full || plan.full) break;
                plan.full = 1;  // a chunk needed the AES blocks the default set omits
            }
            if (!overflow) break;
            if (++level > 6) {
                ipfe_set_error("cufe_gauss_sample_noalloc: speculative window exhausted");
                return -1;
            }
        }
        i = j;
    }
    return sync_stream(st);
}
