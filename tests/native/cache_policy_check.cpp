// Host-side checks of the caching allocator's reuse rules (rte-rrtmgp-cpp_amd/csrc/cache_policy.h).
// Streams and events are opaque handles here; the checks print "ok" and return 0 when all hold.
#include <cstdio>
#include <cstdlib>

#include "cache_policy.h"

using rrtmgp_hip::CacheBlock;
using rrtmgp_hip::CachePolicy;

#define CHECK(c)                                                                  \
    do { if (!(c)) { std::printf("FAILED line %d: %s\n", __LINE__, #c); return 1; } } while (0)

int main()
{
    void* A = reinterpret_cast<void*>(0x100);   // two streams
    void* B = reinterpret_cast<void*>(0x200);
    void* evA = reinterpret_cast<void*>(0x1);    // two events
    void* evB = reinterpret_cast<void*>(0x2);
    char mem[4][8];
    const size_t MiB = size_t(1) << 20;

    CHECK(CachePolicy::round_block(1) == 512);
    CHECK(CachePolicy::round_block(MiB) == MiB);
    CHECK(CachePolicy::round_block(MiB + 1) == 2 * MiB);

    // A block allocated and freed on stream A is reused on A only.
    {
        CachePolicy p;
        p.add_live(mem[0], 4096, 0, A);
        CacheBlock b;
        bool cross = true;
        CHECK(p.find_live(mem[0], A, b, cross));
        CHECK(!cross);
        CHECK(p.finish_release(mem[0], nullptr, nullptr));
        CHECK(!p.take(0, B, 4096, b));          // other stream: no
        CHECK(!p.take(1, A, 4096, b));          // other device: no
        CHECK(p.take(0, A, 4096, b));           // owner: yes
        CHECK(b.ptr == mem[0] && !b.fenced());
        CHECK(p.n_live() == 1 && p.n_free() == 0);
    }
    // Allocated on A, freed on B (the caller switched streams): the free is a cross-stream free;
    // the block comes back fenced by both events and any stream of the device may take it, after
    // which it is owned by the taker and carries no fence any more.
    {
        CachePolicy p;
        p.add_live(mem[1], 8192, 0, A);
        CacheBlock b;
        bool cross = false;
        CHECK(p.find_live(mem[1], B, b, cross));
        CHECK(cross && b.owner == A);
        CHECK(p.finish_release(mem[1], evA, evB));
        CHECK(p.take(0, B, 8192, b));
        CHECK(b.fence[0] == evA && b.fence[1] == evB);
        CacheBlock again;
        bool cr = false;
        CHECK(p.find_live(mem[1], B, again, cr));
        CHECK(again.owner == B && !again.fenced() && !cr);
        // freed on its new owner: plain reuse rules
        CHECK(p.finish_release(mem[1], nullptr, nullptr));
        CHECK(!p.take(0, A, 8192, b));
        CHECK(p.take(0, B, 8192, b) && !b.fenced());
    }
    // Size rules: the smallest fitting block, never one more than max(2x, +4 MiB) larger.
    {
        CachePolicy p;
        p.add_live(mem[0], 64 * MiB, 0, A);
        p.add_live(mem[1], 8 * MiB, 0, A);
        p.add_live(mem[2], 6 * MiB, 0, A);
        for (int i = 0; i < 3; ++i) CHECK(p.finish_release(mem[i], nullptr, nullptr));
        CacheBlock b;
        CHECK(p.take(0, A, 5 * MiB, b) && b.ptr == mem[2]);     // 6 MiB: smallest that fits
        CHECK(p.take(0, A, 5 * MiB, b) && b.ptr == mem[1]);     // 8 MiB: within +4 MiB
        CHECK(!p.take(0, A, 5 * MiB, b));                         // 64 MiB: too wasteful
        CHECK(p.take(0, A, 40 * MiB, b) && b.ptr == mem[0]);    // within 2x
    }
    // A destroyed stream's blocks go to the default stream.
    {
        CachePolicy p;
        p.add_live(mem[3], 4096, 0, B);
        CHECK(p.finish_release(mem[3], nullptr, nullptr));
        p.forget_stream(B);
        CacheBlock b;
        CHECK(!p.take(0, B, 4096, b));
        CHECK(p.take(0, nullptr, 4096, b) && b.ptr == mem[3]);
    }
    // Unknown pointers are not found.
    {
        CachePolicy p;
        CacheBlock b;
        bool cross = false;
        CHECK(!p.find_live(mem[0], A, b, cross));
        CHECK(!p.finish_release(mem[0], nullptr, nullptr));
    }
    std::printf("ok\n");
    return 0;
}
