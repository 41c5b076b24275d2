// Bookkeeping of the library's caching device allocator (runtime.hip), kept free of HIP calls so
// that its reuse rules can be unit-tested on the host.
//
// A block remembers the stream it was allocated on (its owner).  A block freed on its owner goes
// to the free list with that tag and is handed out again only to a request on the same device
// and stream: stream order then keeps the reuse behind every use queued before the free.  A block
// freed on another stream (the caller switched streams while the block was live) is "fenced":
// runtime.hip records an event on the owner and one on the freeing stream, and the block may be
// handed to a request on any stream of its device once that request's stream waits on both.
#ifndef RRTMGP_HIP_CACHE_POLICY_H
#define RRTMGP_HIP_CACHE_POLICY_H

#include <cstddef>
#include <cstdint>
#include <vector>

namespace rrtmgp_hip
{
struct CacheBlock
{
    void* ptr = nullptr;
    size_t bytes = 0;
    int device = 0;
    void* owner = nullptr;         // stream the block is ordered on
    void* fence[2] = {nullptr, nullptr};   // events to wait on before reuse (fenced blocks)
    bool fenced() const { return fence[0] != nullptr || fence[1] != nullptr; }
};

class CachePolicy
{
    public:
        // Allocation granularity: 512 B below 1 MiB, 2 MiB above.
        static size_t round_block(size_t bytes)
        {
            const size_t q = bytes > (size_t(1) << 20) ? (size_t(2) << 20) : size_t(512);
            return (bytes + q - 1) / q * q;
        }

        // Takes the smallest free block of the device that fits `bytes` (rounded) without wasting
        // more than max(bytes, 4 MiB), and that the stream may use: one owned by the stream, or
        // a fenced one.  The block becomes live, owned by `stream`; its fences (if any) are
        // returned in out.fence for the caller to wait on and release.
        bool take(int device, void* stream, size_t bytes, CacheBlock& out)
        {
            const size_t limit = (2 * bytes > bytes + (size_t(4) << 20)) ? 2 * bytes : bytes + (size_t(4) << 20);
            size_t best = free_.size();
            for (size_t i = 0; i < free_.size(); ++i)
            {
                const CacheBlock& b = free_[i];
                if (b.device != device || b.bytes < bytes || b.bytes > limit) continue;
                if (!b.fenced() && b.owner != stream) continue;
                if (best == free_.size() || b.bytes < free_[best].bytes) best = i;
            }
            if (best == free_.size()) return false;
            out = free_[best];
            free_.erase(free_.begin() + std::ptrdiff_t(best));
            CacheBlock live = out;
            live.owner = stream;
            live.fence[0] = live.fence[1] = nullptr;
            live_.push_back(live);
            return true;
        }

        void add_live(void* ptr, size_t bytes, int device, void* stream)
        {
            CacheBlock b;
            b.ptr = ptr; b.bytes = bytes; b.device = device; b.owner = stream;
            live_.push_back(b);
        }

        // Looks up a live block.  `cross` tells whether the freeing stream differs from the owner
        // (the caller then records the two fence events and calls finish_release with them).
        bool find_live(void* ptr, void* free_stream, CacheBlock& out, bool& cross) const
        {
            for (const auto& b : live_)
                if (b.ptr == ptr)
                {
                    out = b;
                    cross = (b.owner != free_stream);
                    return true;
                }
            return false;
        }

        // Moves a live block to the free list, fenced with the given events (both NULL for a
        // free on the owner stream).
        bool finish_release(void* ptr, void* fence_owner, void* fence_free)
        {
            for (size_t i = 0; i < live_.size(); ++i)
                if (live_[i].ptr == ptr)
                {
                    CacheBlock b = live_[i];
                    live_.erase(live_.begin() + std::ptrdiff_t(i));
                    b.fence[0] = fence_owner;
                    b.fence[1] = fence_free;
                    free_.push_back(b);
                    return true;
                }
            return false;
        }

        // A stream that is about to be destroyed (after it was synchronized): blocks it owns are
        // handed to the default stream (NULL), so no later wait names the destroyed stream.
        void forget_stream(void* stream)
        {
            if (stream == nullptr) return;
            for (auto& b : live_) if (b.owner == stream) b.owner = nullptr;
            for (auto& b : free_) if (b.owner == stream) b.owner = nullptr;
        }

        // Empties the free list (the caller frees the blocks and destroys their fences).
        std::vector<CacheBlock> drain_free()
        {
            std::vector<CacheBlock> out;
            out.swap(free_);
            return out;
        }

        size_t live_bytes() const { size_t n = 0; for (const auto& b : live_) n += b.bytes; return n; }
        size_t cached_bytes() const { size_t n = 0; for (const auto& b : free_) n += b.bytes; return n; }
        size_t n_free() const { return free_.size(); }
        size_t n_live() const { return live_.size(); }

    private:
        std::vector<CacheBlock> free_;
        std::vector<CacheBlock> live_;
};
}  // namespace rrtmgp_hip

#endif
