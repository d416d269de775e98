// saber_hip_xdrop_noalloc: the X-drop pass of saber_hip_xdrop over inputs and
// outputs already in device memory, allocating nothing. Its scratch (claim
// counters, the score table, one 32-byte descriptor per hit) is the
// file-scope array below, one copy per device; its address on a device is
// taken once, when a context is created on that device, and only that
// address is passed to kernels and fills.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/saber_hip.h"
#include "common.h"
#include "ctx.h"

using namespace saber;

namespace {

constexpr uint64_t kCounterBytes = (uint64_t)kXdropGroups * kXdropCounterStride * 4;   // 2 KiB
constexpr uint64_t kTableBytes = (uint64_t)kXdropTable * sizeof(short);               // 16 KiB
constexpr uint64_t kDescOffset = kCounterBytes + kTableBytes;
constexpr uint64_t kScratchBytes = kDescOffset + (uint64_t)SABER_HIP_XDROP_NOALLOC_MAX_HITS * 32;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

struct MatrixArg {
    int32_t m[576];
};

// kXdropTable entries: row a (query code, rows >= 24 as X) x column c
// (subject byte, columns >= 24 as X).
__global__ void xdrop_table_kernel(MatrixArg mat, short* table) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kXdropTable) return;
    const int a = i >> 8, c = i & 255;
    table[i] = (short)mat.m[(a < 24 ? a : kCodeX) * 24 + (c < 24 ? c : kCodeX)];
}

}  // namespace

namespace saber {

// The scratch array's address on the current device (called with the
// context's device current, at context creation).
int xdrop_noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(SABER_HIP_EINVAL, "device ordinal above the scratch table");
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[device]) return SABER_HIP_OK;
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_ptr[device] = static_cast<unsigned char*>(p);
    return SABER_HIP_OK;
}

}  // namespace saber

extern "C" int saber_hip_xdrop_noalloc(saber_hip_ctx* c, const uint8_t* qres, const uint64_t* qoff,
                                       const int8_t* qbias, const uint8_t* sres, const uint64_t* soff,
                                       const uint32_t* hits, uint64_t nhits, const saber_hip_scheme* scheme,
                                       int32_t x, uint32_t* out) {
    if (!c || !scheme) return fail(SABER_HIP_EINVAL, "null argument");
    if (nhits == 0) return SABER_HIP_OK;
    if (!qres || !qoff || !sres || !soff || !hits || !out) return fail(SABER_HIP_EINVAL, "null device pointer");
    if (nhits > SABER_HIP_XDROP_NOALLOC_MAX_HITS)
        return fail(SABER_HIP_EINVAL, "more hits than the allocation-free scratch holds (SABER_HIP_XDROP_NOALLOC_MAX_HITS)");
    bool wide = false;
    MatrixArg mat;
    for (int i = 0; i < 576; ++i) {
        const int32_t v = scheme->matrix[i];
        if (v < -32767 || v > 32767) return fail(SABER_HIP_EINVAL, "substitution score outside the int16 range");
        if (v < -128 || v > 127) wide = true;
        mat.m[i] = v;
    }
    std::lock_guard<std::mutex> g(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> g2(g_mu);
        base = c->device >= 0 && c->device < kMaxDevices ? g_ptr[c->device] : nullptr;
    }
    if (!base) return fail(SABER_HIP_ESTATE, "scratch address not set up on this device");
    hipStream_t st = c->stream;
    uint32_t* counters = reinterpret_cast<uint32_t*>(base);
    short* table = reinterpret_cast<short*>(base + kCounterBytes);
    void* desc = base + kDescOffset;
    HIP_TRY(fill_u32(counters, kXdropGroups * kXdropCounterStride, 0, st));
    SABER_LAUNCH(xdrop_table_kernel, dim3((kXdropTable + 255) / 256), dim3(256), 0, st, mat, table);
    HIP_TRY(hipGetLastError());
    const bool bias = qbias != nullptr;
    XdropArgs a{};
    a.hits = hits;
    a.nhits = (uint32_t)nhits;
    a.qoff = qoff;
    a.qres = qres;
    a.qbias = qbias;
    a.soff = soff;
    a.sres = sres;
    a.table = table;
    a.x = x;
    a.next_chunk = counters;
    a.desc = desc;
    a.res = out;
    const int grid = (int)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)c->cus * xdrop_blocks_per_cu(bias, wide), (nhits + 255) / 256));
    a.chunk = (uint32_t)std::max<uint64_t>(64, nhits / (16ull * grid));
    HIP_TRY(launch_xdrop(a, bias, wide, grid, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SABER_HIP_OK;
}
