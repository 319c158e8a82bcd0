// rrg_noalloc.hip -- the device-resident decode without device allocation.
//
// rrg_gpu_decode_device grows its context's device buffers with hipMalloc as
// inputs need them.  Callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid it) call
// rrg_gpu_decode_device_noalloc instead: the same decode, with every device
// buffer of the call (input copy, finder targets, segment descriptors and
// results, cell regions, per-lane scratch, tail windows, CRC pieces) carved
// from one file-scope device array, and the output written to the caller's
// buffer.
//
// The array is a single region per device, held from the library's first use
// of it on that device until the process ends.  Every call on a device shares
// it and it keeps the last call's contents; calls that overlap on one device
// race.  Everything that must start at zero is cleared by the decode itself.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "rrg_engine.h"

namespace {

using namespace rrg::host;

// Size of the array, the largest compressed input it serves and the inflate
// lanes it plans for (their per-lane scratch comes out of the array too).
constexpr size_t kScratchBytes = 256ull << 20;
constexpr size_t kMaxInput = 8ull << 20;
constexpr uint64_t kLanes = 8192;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (`device`), looked up
// once per device.  In host code the array's name is only a stand-in for the
// symbol; its memory is reached through this address.
hipError_t scratch_on(int device, uint8_t** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = static_cast<uint8_t*>(p);
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" int rrg_gpu_decode_device_noalloc(void* ctx, const void* d_file, size_t len, const uint8_t* h_file,
                                             const rrg_cfg* cfg, uint8_t* d_out, size_t out_cap, uint64_t* out_len,
                                             uint64_t* chunk_ends, size_t chunk_cap, size_t* n_chunks,
                                             rrg_stats* stats, rrg_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (!ctx || !d_file || !d_out) {
        set_err(err, RRG_E_ARG, "null context, input or output");
        return RRG_E_ARG;
    }
    if (len > kMaxInput) {
        set_err(err, RRG_E_ARG, "input of %zu bytes exceeds the %zu-byte limit of the no-allocation decode", len,
                kMaxInput);
        return RRG_E_ARG;
    }
    if (reinterpret_cast<uintptr_t>(d_out) & 255) {
        set_err(err, RRG_E_ARG, "output buffer must be 256-byte aligned");
        return RRG_E_ARG;
    }
    Ctx& c = *static_cast<Ctx*>(ctx);
    HIPCHK(hipSetDevice(c.device));
    uint8_t* base = nullptr;
    HIPCHK(scratch_on(c.device, &base));
    // every buffer of the context is swapped for an empty one carved from the
    // array (the output: the caller's), and restored afterwards
    Arena arena;
    arena.base = base;
    arena.cap = kScratchBytes;
    std::vector<DevBuf*> bufs = c.bufs();
    std::vector<DevBuf> saved;
    saved.reserve(bufs.size());
    for (DevBuf* b : bufs) {
        saved.push_back(*b);
        *b = DevBuf{};
        b->arena = &arena;
    }
    c.out.p = d_out;
    c.out.cap = out_cap;
    c.out.fixed = true;
    const uint64_t lanes_before = c.max_lanes;
    c.max_lanes = kLanes;
    c.no_alloc = true;  // uploads use the pinned staging buffer only if one exists already
    int rc = decode_device_impl(c, d_file, len, h_file, cfg, out_len, chunk_ends, chunk_cap, n_chunks, stats, err);
    (void)hipStreamSynchronize(c.s);  // no kernel of this call still uses the array
    c.no_alloc = false;
    c.max_lanes = lanes_before;
    for (size_t i = 0; i < bufs.size(); ++i) *bufs[i] = saved[i];
    if (rc == RRG_E_DEVICE && arena.fixed_short) {
        set_err(err, RRG_E_ARG, "output capacity %zu bytes is too small: the decode needs %zu", out_cap,
                arena.fixed_short);
        return RRG_E_ARG;
    }
    if (rc == RRG_E_DEVICE && arena.exhausted) {
        set_err(err, RRG_E_ARG, "the decode needs more than the %zu MiB no-allocation workspace", kScratchBytes >> 20);
        return RRG_E_ARG;
    }
    return rc;
}
