// outring.hip -- host ownership of streaming outputs: per signal, a ring of
// ARGUS_OUT_RING_DEPTH reusable host buffers, refilled from a free pool that
// is shared by every signal ID (streamOutRing / outBufPool,
// internal/demod/gpudemod/streaming_out_ring.go:22-112).
//
// A streaming call hands each signal's outputs to its consumer in the ring's
// next buffer.  The buffer stays untouched for the next DEPTH - 1 calls, so a
// consumer that is up to DEPTH - 1 calls behind still reads its own data, and
// the producer makes no allocation once the buffers have grown to size.
//  * next(id, n): the ring of `id` (created empty on first use) takes its
//    current slot; a slot whose buffer holds fewer than n samples draws one
//    from the pool (the slot's old buffer is released).  The ring then
//    advances (streaming_out_ring.go:32-45).
//  * pool get(n): the first free buffer that holds n, else the largest free
//    one (released if it is still too small), else a new buffer of exactly n
//    samples (56-78).
//  * pool put: released buffers join the pool, up to 4 x 32 of them; the
//    rest are freed (81-89).
//  * prune(active): the rings of signals outside `active` return their
//    buffers to the pool (in ascending ID order) and are dropped (105-112).
//    The batch runner prunes with the job list of every streaming call, as
//    it prunes signal states.
// Buffers are plain host memory (the outputs are copied in from the runner's
// pinned staging).  Not thread-safe: one owner, like the runner.
#include "common.hpp"
#include "outring.hpp"

#include <algorithm>
#include <cstdlib>
#include <unordered_set>

namespace {

constexpr int kMaxPooled = ARGUS_OUT_RING_DEPTH * 32;

void release(gpud_out_rings* o, ArgusOutBuf& b)
{
    if (b.p) {
        std::free(b.p);
        o->released++;
    }
    b = ArgusOutBuf{};
}

ArgusOutBuf pool_get(gpud_out_rings* o, int64_t n)
{
    int best = -1;
    for (int i = 0; i < (int)o->free.size(); i++) {
        if (o->free[(size_t)i].cap >= n) {
            best = i;
            break;
        }
        if (best < 0 || o->free[(size_t)i].cap > o->free[(size_t)best].cap) best = i;
    }
    if (best >= 0) {
        ArgusOutBuf b = o->free[(size_t)best];
        o->free[(size_t)best] = o->free.back();
        o->free.pop_back();
        if (b.cap >= n) return b;
        release(o, b);  // the largest pooled buffer is still too small
    }
    ArgusOutBuf b;
    b.cap = n;
    b.p = (gpud_float2*)std::malloc((size_t)(n > 0 ? n : 1) * sizeof(gpud_float2));
    if (b.p) o->allocated++;
    else b.cap = 0;
    return b;
}

void pool_put(gpud_out_rings* o, ArgusOutBuf b)
{
    if (!b.p) return;
    if ((int)o->free.size() >= kMaxPooled) {
        release(o, b);
        return;
    }
    o->free.push_back(b);
}

}  // namespace

int argus_out_ring_next(gpud_out_rings* o, int64_t signal_id, int64_t n, gpud_float2** buf)
{
    if (!o || n < 0 || !buf) return -1;
    ArgusOutRing& ring = o->rings[signal_id];
    ArgusOutBuf& slot = ring.bufs[ring.idx];
    if (slot.cap < n) {
        // the slot's old buffer is dropped, not pooled (streaming_out_ring.go:33-35)
        release(o, slot);
        slot = pool_get(o, n);
        if (!slot.p) return -3;
    }
    *buf = slot.p;
    ring.idx = (ring.idx + 1) % ARGUS_OUT_RING_DEPTH;
    return 0;
}

void argus_out_rings_prune(gpud_out_rings* o, const int64_t* active, int n_active)
{
    if (!o || o->rings.empty()) return;
    std::unordered_set<int64_t> keep(active, active + (n_active > 0 ? n_active : 0));
    // The reference visits its map in Go's random order; here in ascending
    // ID order, so which buffers the bounded pool keeps is reproducible.
    std::vector<int64_t> drop;
    for (const auto& kv : o->rings)
        if (!keep.count(kv.first)) drop.push_back(kv.first);
    std::sort(drop.begin(), drop.end());
    for (int64_t id : drop) {
        auto it = o->rings.find(id);
        for (ArgusOutBuf& b : it->second.bufs) pool_put(o, b);
        o->rings.erase(it);
    }
}

void argus_out_rings_free(gpud_out_rings* o)
{
    if (!o) return;
    for (auto& kv : o->rings)
        for (ArgusOutBuf& b : kv.second.bufs) release(o, b);
    for (ArgusOutBuf& b : o->free) release(o, b);
    delete o;
}

extern "C" {

GPUD_API int gpud_out_rings_create(gpud_out_rings** out)
{
    if (!out) return -1;
    *out = new gpud_out_rings();
    return 0;
}

GPUD_API int gpud_out_rings_destroy(gpud_out_rings* o)
{
    argus_out_rings_free(o);
    return 0;
}

GPUD_API int gpud_out_rings_next(gpud_out_rings* o, int64_t signal_id, int64_t n, gpud_float2** buf)
{
    return argus_out_ring_next(o, signal_id, n, buf);
}

GPUD_API int gpud_out_rings_prune(gpud_out_rings* o, const int64_t* active, int n_active)
{
    if (!o || n_active < 0 || (n_active > 0 && !active)) return -1;
    argus_out_rings_prune(o, active, n_active);
    return 0;
}

GPUD_API int gpud_out_rings_pool_get(gpud_out_rings* o, int64_t n, gpud_float2** buf, int64_t* cap)
{
    if (!o || n < 0 || !buf) return -1;
    ArgusOutBuf b = pool_get(o, n);
    if (!b.p) return -3;
    *buf = b.p;
    if (cap) *cap = b.cap;
    return 0;
}

GPUD_API int gpud_out_rings_pool_put(gpud_out_rings* o, gpud_float2* buf, int64_t cap)
{
    if (!o || !buf || cap < 0) return -1;
    ArgusOutBuf b;
    b.p = buf;
    b.cap = cap;
    pool_put(o, b);
    return 0;
}

GPUD_API int gpud_out_rings_stats(const gpud_out_rings* o, int* n_rings, int* n_free, int64_t* allocated,
                                  int64_t* released)
{
    if (!o) return -1;
    if (n_rings) *n_rings = (int)o->rings.size();
    if (n_free) *n_free = (int)o->free.size();
    if (allocated) *allocated = o->allocated;
    if (released) *released = o->released;
    return 0;
}

GPUD_API int gpud_out_rings_slot(const gpud_out_rings* o, int64_t signal_id, int* next_idx, int64_t* caps)
{
    if (!o) return -1;
    auto it = o->rings.find(signal_id);
    if (it == o->rings.end()) return -1;
    if (next_idx) *next_idx = it->second.idx;
    if (caps)
        for (int i = 0; i < ARGUS_OUT_RING_DEPTH; i++) caps[i] = it->second.bufs[i].cap;
    return 0;
}

}  // extern "C"
