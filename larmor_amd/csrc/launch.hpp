// Host-side launch helpers of liblarmor_sd.so, shared by every translation unit and by both
// arithmetic forms (FP64, namespace lsd; plain FP32, namespace lsd32): error reporting, the
// stream-ordered allocator, the opt-in per-launch timing and the launchers of persistent kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/larmor_sd.h"

namespace lsd {

constexpr int LAUNCH_WAVE = 64;
constexpr int WAVES_PER_BLOCK = 4;

extern thread_local std::string g_error;

// The low-resolution solve's LDS factor capacity for the row class rpl (4, 6 or 8 rows per lane),
// 20 columns unless lsd_tune_lr_capacity set another (sd_api.hip).
int lr_capacity(int rpl);

inline int fail(const std::string& message) {
    g_error = message;
    return LSD_ERR_INVALID;
}

inline int hip_fail(hipError_t e, const char* what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return LSD_ERR_HIP;
}

#define LSD_HIP(call)                                                                          \
    do {                                                                                       \
        const hipError_t lsd_e = (call);                                                       \
        if (lsd_e != hipSuccess) {                                                             \
            return hip_fail(lsd_e, #call);                                                     \
        }                                                                                      \
    } while (0)

inline int device_alloc(void** p, size_t bytes, hipStream_t stream) {
    if (bytes == 0) {
        bytes = 8;
    }
    LSD_HIP(hipMallocAsync(p, bytes, stream));
    return 0;
}

// Opt-in per-launch timing: events recorded on the launch stream around each kernel, and for
// the kernels that count their algorithmic work, a device counter slot per launch.
struct ProfiledLaunch {
    const char* name;
    hipEvent_t start;
    hipEvent_t stop;
    unsigned long long* work;
};
extern bool g_profiling;
extern std::vector<ProfiledLaunch> g_launches;
extern unsigned long long* g_work_slots; // one zeroed counter per profiled launch
constexpr long WORK_SLOTS = 16384;

// Gives a launch's arguments its work counter (kernels that count none ignore it).
template <class Args>
inline bool attach_work(Args&, unsigned long long*) {
    return false;
}

// Launches kern over `waves` waves, waves_per_block to a block with lds_bytes of dynamic LDS.
template <class Kern, class Args>
inline int launch(const char* name, Kern kern, const Args& args, int waves, size_t lds_bytes,
           hipStream_t stream, int waves_per_block = WAVES_PER_BLOCK) {
    const int blocks = (waves + waves_per_block - 1) / waves_per_block;
    ProfiledLaunch rec{name, nullptr, nullptr, nullptr};
    Args a = args;
    if (g_profiling) {
        LSD_HIP(hipEventCreate(&rec.start));
        LSD_HIP(hipEventCreate(&rec.stop));
        LSD_HIP(hipEventRecord(rec.start, stream));
        const long at = static_cast<long>(g_launches.size());
        unsigned long long* slot =
            (g_work_slots != nullptr && at < WORK_SLOTS) ? g_work_slots + at : nullptr;
        if (attach_work(a, slot)) {
            rec.work = slot;
        }
    }
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(LAUNCH_WAVE * waves_per_block), lds_bytes, stream, a);
    LSD_HIP(hipGetLastError());
    if (g_profiling) {
        LSD_HIP(hipEventRecord(rec.stop, stream));
        g_launches.push_back(rec);
    }
    return 0;
}


// Launches a persistent kernel (one that strides over its items) on at most `waves` waves and
// no more than the device keeps resident at once: one full-occupancy wave set.
template <class Kern, class Args>
inline int launch_fill(const char* name, Kern kern, const Args& args, int waves, size_t lds_bytes,
                       hipStream_t stream) {
    int blocks_per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, kern, LAUNCH_WAVE * WAVES_PER_BLOCK,
                                                     lds_bytes) == hipSuccess &&
        blocks_per_cu > 0) {
        int dev = 0;
        int cus = 0;
        LSD_HIP(hipGetDevice(&dev));
        LSD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const long resident = static_cast<long>(blocks_per_cu) * cus * WAVES_PER_BLOCK;
        if (resident > 0 && resident < waves) {
            waves = static_cast<int>(resident);
        }
    }
    return launch(name, kern, args, waves, lds_bytes, stream);
}

// Scratch of the global-memory rerun launches of one kind of solve.  Either the first launch
// that needs it allocates it (stream-ordered; the caller frees it) and later launches reuse it,
// or the caller hands in a fixed region (fixed = true) whose size sets the rerun's wave count.
struct RerunScratch {
    void* base = nullptr;
    size_t bytes = 0;
    bool fixed = false;
};

inline std::string shape_tuple(int64_t r, int64_t c) {
    return "(" + std::to_string(r) + ", " + std::to_string(c) + ")";
}

// Which instantiation covers a problem size.
inline int ncpl_for(int64_t n) {
    if (n <= 64) return 1;
    if (n <= 192) return 3;
    if (n <= 320) return 5;
    if (n <= 640) return 10;
    return -1;
}
inline int rpl_for(int64_t m) {
    if (m <= 256) return 4;
    if (m <= 384) return 6;
    if (m <= 512) return 8;
    return -1;
}
// Row slots of the plain FP64 fits, which also take up to 768 and 1,024 measurements (12 and 16
// slots); the other forms stop at 512.
constexpr int64_t MAX_MEASUREMENTS_WIDE = 1024;
inline int rpl_for_wide(int64_t m) {
    const int r = rpl_for(m);
    if (r > 0) return r;
    if (m <= 768) return 12;
    if (m <= MAX_MEASUREMENTS_WIDE) return 16;
    return -1;
}

// Grid of the main launch: enough waves to fill the chip, never more than voxels.
inline int main_waves(int64_t n_vox) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (cus <= 0) {
        cus = 256;
    }
    // Up to 8 waves per SIMD; each persistent launch then clamps to what its kernel keeps
    // resident (launch_fill).
    const int64_t target = static_cast<int64_t>(cus) * 32;
    int64_t waves = std::min<int64_t>(target, std::max<int64_t>(n_vox, 1));
    waves = (waves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK * WAVES_PER_BLOCK;
    return static_cast<int>(waves);
}


// The device memory of one call.  By default each buffer is a stream-ordered allocation, freed
// (stream-ordered) when the call returns.  Given a fixed region instead (the allocation-free
// entry point), buffers are carved from it in order, 256-byte aligned, and a request beyond its
// end fails; what is left after the last buffer serves the rerun scratch.
class Workspace {
public:
    explicit Workspace(hipStream_t stream) : stream_(stream) {}
    Workspace(void* base, size_t bytes, hipStream_t stream)
        : stream_(stream), base_(static_cast<char*>(base)), capacity_(bytes), fixed_(true) {}
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() {
        for (void* p : held_) {
            (void)hipFreeAsync(p, stream_);
        }
    }
    template <class T>
    int get(T** out, long count) {
        const size_t bytes = sizeof(T) * static_cast<size_t>(std::max(1L, count));
        if (fixed_) {
            const size_t at = align(used_);
            if (at > capacity_ || bytes > capacity_ - at) {
                return fail("the fit needs more than the " + std::to_string(capacity_) +
                            " bytes of the allocation-free workspace; pass fewer voxels per call");
            }
            *out = reinterpret_cast<T*>(base_ + at);
            used_ = at + bytes;
            return 0;
        }
        void* p = nullptr;
        const int rc = device_alloc(&p, bytes, stream_);
        if (rc) return rc;
        held_.push_back(p);
        *out = static_cast<T*>(p);
        return 0;
    }
    // Frees p with the call's own buffers (the rerun scratch allocated by a launcher).
    void adopt(void* p) {
        if (p && !fixed_) held_.push_back(p);
    }
    bool fixed() const { return fixed_; }
    // The region after the last buffer, for the rerun scratch (empty, to be allocated by the
    // launcher, when the workspace is not fixed).
    RerunScratch rest() const {
        RerunScratch r;
        if (fixed_) {
            const size_t at = align(used_);
            r.base = base_ + std::min(at, capacity_);
            r.bytes = at < capacity_ ? capacity_ - at : 0;
            r.fixed = true;
        }
        return r;
    }

private:
    static size_t align(size_t at) { return (at + 255) & ~static_cast<size_t>(255); }
    hipStream_t stream_;
    char* base_ = nullptr;
    size_t capacity_ = 0;
    size_t used_ = 0;
    bool fixed_ = false;
    std::vector<void*> held_;
};

} // namespace lsd
