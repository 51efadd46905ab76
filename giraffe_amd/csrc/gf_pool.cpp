// Device memory cache for model storage.  hipFree waits for the whole device,
// so a server that creates and destroys models while others run would stall
// every resident request at each finished one.  Freed blocks are kept per
// device and reused for requests of at least their size and at most twice it;
// the memory stays with the process (it is the high-water mark of the models
// alive at once).  Callers free a block only after its last use on the device
// has completed.
#include "gf_internal.h"

#include <map>
#include <mutex>
#include <unordered_map>

namespace {

struct Pool {
    std::mutex mu;
    std::multimap<size_t, void *> free_by_size;  // per device key = (device << 48) | bytes
    std::unordered_map<void *, std::pair<int, size_t>> live;
};

Pool &pool()
{
    static Pool *p = new Pool();  // never destroyed: HIP may be gone at exit
    return *p;
}

size_t round_up(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

int gf_pool_alloc(void **out, size_t bytes)
{
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    size_t b = round_up(bytes ? bytes : 1);
    const size_t key = ((size_t)dev << 48) | b, key_max = ((size_t)dev << 48) | (2 * b);
    Pool &P = pool();
    {
        std::lock_guard<std::mutex> lk(P.mu);
        auto it = P.free_by_size.lower_bound(key);
        if (it != P.free_by_size.end() && it->first <= key_max) {
            *out = it->second;
            P.live[it->second] = {dev, it->first & ((1ull << 48) - 1)};
            P.free_by_size.erase(it);
            return GF_OK;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        // release the cache of this device and retry once
        std::lock_guard<std::mutex> lk(P.mu);
        for (auto it = P.free_by_size.begin(); it != P.free_by_size.end();) {
            if ((int)(it->first >> 48) == dev) {
                (void)hipFree(it->second);
                it = P.free_by_size.erase(it);
            } else {
                ++it;
            }
        }
        (void)hipGetLastError();
        e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            gf_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return GF_E_NOMEM;
        }
    }
    std::lock_guard<std::mutex> lk(P.mu);
    P.live[p] = {dev, b};
    *out = p;
    return GF_OK;
}

void gf_pool_free(void *p)
{
    if (!p) return;
    Pool &P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) {
        (void)hipFree(p);  // not from the cache
        return;
    }
    P.free_by_size.emplace(((size_t)it->second.first << 48) | it->second.second, p);
    P.live.erase(it);
}

// Small pinned host blocks (per-model norm mirrors): kept and reused as well,
// since hipHostFree also waits for the device.
namespace {
std::mutex pinned_mu;
std::multimap<size_t, void *> pinned_free;
std::unordered_map<void *, size_t> pinned_live;
}  // namespace

int gf_pinned_alloc(void **out, size_t bytes)
{
    size_t b = round_up(bytes ? bytes : 1);
    {
        std::lock_guard<std::mutex> lk(pinned_mu);
        auto it = pinned_free.find(b);
        if (it != pinned_free.end()) {
            *out = it->second;
            pinned_live[it->second] = b;
            pinned_free.erase(it);
            return GF_OK;
        }
    }
    void *p = nullptr;
    GF_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
    std::lock_guard<std::mutex> lk(pinned_mu);
    pinned_live[p] = b;
    *out = p;
    return GF_OK;
}

void gf_pinned_free(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(pinned_mu);
    auto it = pinned_live.find(p);
    if (it == pinned_live.end()) {
        (void)hipHostFree(p);
        return;
    }
    pinned_free.emplace(it->second, p);
    pinned_live.erase(it);
}
