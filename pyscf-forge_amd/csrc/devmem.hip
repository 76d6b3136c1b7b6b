// Large device blocks (sigma plan workspaces, tens of GB) kept across plans.
// Releasing blocks of that size and allocating them again costs up to about
// a second in the driver (measured on MI355X: plan creation of an FP64 and a
// bf16x3 plan at (16o,16e) 0.23 s in a fresh process, 0.92 s once earlier
// plans had been freed), so a block freed through big_free stays allocated
// for the next big_alloc of about its size on the same device.
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace {

constexpr size_t BIG_MIN = (size_t)1 << 28;    // smaller blocks are not kept
constexpr size_t BIG_KEEP = 4;                 // kept blocks per process

struct Blk {
        void *p;
        size_t bytes;
        int dev;
};

std::mutex mu;
std::unordered_map<void *, Blk> live;   // blocks handed out by big_alloc
std::vector<Blk> kept;                  // oldest first

void release_device(int dev)
{
        for (auto it = kept.begin(); it != kept.end();) {
                if (dev < 0 || it->dev == dev) {
                        (void)hipFree(it->p);
                        it = kept.erase(it);
                } else {
                        ++it;
                }
        }
}

}  // namespace

hipError_t fcis::big_alloc(void **p, size_t bytes)
{
        *p = nullptr;
        if (bytes < BIG_MIN) return dev_malloc(p, bytes);
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) {
                (void)hipGetLastError();
                return hipErrorInvalidDevice;
        }
        std::lock_guard<std::mutex> g(mu);
        // the smallest kept block of this device that holds `bytes` and wastes
        // at most a quarter of it
        int best = -1;
        for (size_t i = 0; i < kept.size(); i++) {
                const Blk &b = kept[i];
                if (b.dev == dev && b.bytes >= bytes && b.bytes - bytes <= bytes / 4 &&
                    (best < 0 || b.bytes < kept[best].bytes))
                        best = (int)i;
        }
        if (best >= 0) {
                const Blk b = kept[best];
                kept.erase(kept.begin() + best);
                live[b.p] = b;
                *p = b.p;
                return hipSuccess;
        }
        hipError_t e = dev_malloc(p, bytes);
        if (e != hipSuccess) {
                release_device(dev);
                e = dev_malloc(p, bytes);
                if (e != hipSuccess) return e;
        }
        live[*p] = Blk{*p, bytes, dev};
        return hipSuccess;
}

void fcis::big_free(void *p)
{
        if (p == nullptr) return;
        std::lock_guard<std::mutex> g(mu);
        auto it = live.find(p);
        if (it == live.end()) {
                (void)hipFree(p);
                return;
        }
        const Blk b = it->second;
        live.erase(it);
        // work still in flight on the block finishes before it can be reused
        // (hipFree synchronizes the same way)
        (void)hipDeviceSynchronize();
        kept.push_back(b);
        while (kept.size() > BIG_KEEP) {
                (void)hipFree(kept.front().p);
                kept.erase(kept.begin());
        }
}

size_t fcis::big_kept_bytes(int dev)
{
        std::lock_guard<std::mutex> g(mu);
        size_t s = 0;
        for (const Blk &b : kept)
                if (dev < 0 || b.dev == dev) s += b.bytes;
        return s;
}

void fcis::big_release(int dev)
{
        std::lock_guard<std::mutex> g(mu);
        release_device(dev);
}

extern "C" {

int fcis_kept_memory(int device, uint64_t *bytes)
{
        if (bytes == nullptr) {
                fcis::set_error("fcis_kept_memory: null output");
                return FCIS_ERR_INVALID;
        }
        *bytes = (uint64_t)fcis::big_kept_bytes(device);
        return FCIS_SUCCESS;
}

int fcis_release_kept_memory(int device)
{
        fcis::big_release(device);
        return FCIS_SUCCESS;
}

}  // extern "C"
