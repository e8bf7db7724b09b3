// ingest.hip -- the step before the extraction path: SDRIQ001 capture files
// (internal/iqfile/iqfile.go) and a pinned host ring that moves a source's
// blocks to the device ahead of the consumer (gpud_iqfile_*, gpud_ingest_*).
// Host code only; no kernels.
//
// Capture format (iqfile.go:5-11, little-endian): "SDRIQ001", int64 sample
// rate, float64 centre frequency, int64 sample count, then count complex64
// samples.  Reads go straight from the file into the caller's buffer (or a
// pinned ring slot) with pread, so a capture of any size streams without
// being loaded; ReadIQ's loop-at-end replay (iqfile.go:121-142) is kept.
//
// Ingest ring: depth slots, each a pinned host buffer, a device buffer and an
// event.  A producer (a thread replaying a capture, or the caller pushing
// blocks from a live source) fills the next free slot's host buffer, queues
// its H2D copy on the ring's own stream and records the event; the consumer
// acquires slots in order (waiting on the event only if the copy is still
// running) and releases each once the work that reads its device buffer has
// completed.  With depth >= 3 the copy of block k+1 and the read of block
// k+2 overlap the GPU work on block k.
#include "common.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

namespace {

constexpr char kMagic[8] = {'S', 'D', 'R', 'I', 'Q', '0', '0', '1'};
constexpr int64_t kHeader = 32;

void put_u64(unsigned char* p, uint64_t v)
{
    for (int i = 0; i < 8; i++) p[i] = (unsigned char)(v >> (8 * i));
}

uint64_t get_u64(const unsigned char* p)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

bool pread_all(int fd, void* dst, size_t bytes, int64_t off)
{
    char* d = (char*)dst;
    while (bytes > 0) {
        const ssize_t r = pread(fd, d, bytes, (off_t)off);
        if (r <= 0) return false;
        d += r;
        bytes -= (size_t)r;
        off += r;
    }
    return true;
}

}  // namespace

struct gpud_iqfile {
    int fd = -1;
    int64_t sample_rate = 0;
    double center_hz = 0;
    int64_t n = 0;
    int64_t pos = 0;
    std::mutex mu;
};

enum SlotState { kFree = 0, kFilling, kReady, kAcquired };

struct IngestSlot {
    gpud_float2* host = nullptr;  // pinned
    gpud_float2* dev = nullptr;
    hipEvent_t ev = nullptr;
    int n = 0;
    int64_t seq = -1;
    SlotState state = kFree;
};

struct gpud_ingest {
    int block = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<IngestSlot> slots;
    std::mutex mu;
    std::condition_variable cv;
    int64_t produced = 0;  // next sequence number to fill
    int64_t consumed = 0;  // next sequence number to acquire
    int error = 0;         // first producer error (sticky)
    bool stop = false;
    std::thread producer;
    gpud_iqfile* file = nullptr;
};

// Fill and queue the copy of the next block; blocks while the ring is full.
// fill(host, n) writes the block's samples into the pinned buffer.
template <typename Fill>
static int ingest_produce(gpud_ingest* g, int n, Fill fill)
{
    IngestSlot* sl = nullptr;
    int64_t seq = 0;
    {
        std::unique_lock<std::mutex> lk(g->mu);
        g->cv.wait(lk, [&] { return g->stop || g->slots[(size_t)(g->produced % g->slots.size())].state == kFree; });
        if (g->stop) return -2;
        seq = g->produced;
        sl = &g->slots[(size_t)(seq % g->slots.size())];
        sl->state = kFilling;
    }
    int err = fill(sl->host, n) ? 0 : -3;
    if (!err && n > 0)
        err = (int)hipMemcpyAsync(sl->dev, sl->host, (size_t)n * sizeof(gpud_float2), hipMemcpyHostToDevice, g->stream);
    if (!err) err = (int)hipEventRecord(sl->ev, g->stream);
    {
        std::lock_guard<std::mutex> lk(g->mu);
        if (err) {
            sl->state = kFree;
            if (!g->error) g->error = err;
        } else {
            sl->n = n;
            sl->seq = seq;
            sl->state = kReady;
            g->produced = seq + 1;
        }
    }
    g->cv.notify_all();
    return err;
}

extern "C" {

// ------------------------------------------------------------- capture files

GPUD_API int gpud_iqfile_write(const char* path, const gpud_float2* iq, int64_t n, int64_t sample_rate,
                               double center_hz)
{
    if (!path || n < 0 || (n > 0 && !iq)) return -1;
    FILE* f = std::fopen(path, "wb");
    if (!f) return -2;
    unsigned char hdr[kHeader];
    std::memcpy(hdr, kMagic, 8);
    put_u64(hdr + 8, (uint64_t)sample_rate);
    uint64_t cbits;
    std::memcpy(&cbits, &center_hz, 8);
    put_u64(hdr + 16, cbits);
    put_u64(hdr + 24, (uint64_t)n);
    bool ok = std::fwrite(hdr, 1, kHeader, f) == (size_t)kHeader;
    // complex64 little-endian: the in-memory layout on this (little-endian) host
    if (ok && n > 0) ok = std::fwrite(iq, sizeof(gpud_float2), (size_t)n, f) == (size_t)n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : -3;
}

// Open a capture: its header (iqfile.go:66-90).  -2 cannot open, -3 bad
// magic or short header, -4 fewer samples than the header declares.
GPUD_API int gpud_iqfile_open(const char* path, gpud_iqfile** out, int64_t* sample_rate, double* center_hz,
                              int64_t* n_samples)
{
    if (!path || !out) return -1;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return -2;
    unsigned char hdr[kHeader];
    if (!pread_all(fd, hdr, kHeader, 0) || std::memcmp(hdr, kMagic, 8) != 0) {
        ::close(fd);
        return -3;
    }
    const int64_t n = (int64_t)get_u64(hdr + 24);
    const off_t size = lseek(fd, 0, SEEK_END);
    if (n < 0 || (int64_t)size < kHeader + n * 8) {
        ::close(fd);
        return -4;
    }
    auto* f = new gpud_iqfile();
    f->fd = fd;
    f->sample_rate = (int64_t)get_u64(hdr + 8);
    const uint64_t cbits = get_u64(hdr + 16);
    std::memcpy(&f->center_hz, &cbits, 8);
    f->n = n;
    if (sample_rate) *sample_rate = f->sample_rate;
    if (center_hz) *center_hz = f->center_hz;
    if (n_samples) *n_samples = n;
    *out = f;
    return 0;
}

// ReadIQ (iqfile.go:121-142): n samples from the current position, looping
// to the start at the end; an empty capture yields zeros.
GPUD_API int gpud_iqfile_read(gpud_iqfile* f, gpud_float2* dst, int n)
{
    if (!f || n < 0 || (n > 0 && !dst)) return -1;
    std::lock_guard<std::mutex> lk(f->mu);
    if (f->n == 0) {
        std::memset(dst, 0, (size_t)n * sizeof(gpud_float2));
        return 0;
    }
    int done = 0;
    while (done < n) {
        const int64_t run = std::min<int64_t>(n - done, f->n - f->pos);
        if (!pread_all(f->fd, dst + done, (size_t)run * sizeof(gpud_float2), kHeader + f->pos * 8)) return -3;
        done += (int)run;
        f->pos += run;
        if (f->pos >= f->n) f->pos = 0;
    }
    return 0;
}

GPUD_API int gpud_iqfile_seek(gpud_iqfile* f, int64_t pos)
{
    if (!f || pos < 0 || (f->n > 0 && pos >= f->n) || (f->n == 0 && pos != 0)) return -1;
    std::lock_guard<std::mutex> lk(f->mu);
    f->pos = pos;
    return 0;
}

GPUD_API int gpud_iqfile_close(gpud_iqfile* f)
{
    if (!f) return 0;
    if (f->fd >= 0) ::close(f->fd);
    delete f;
    return 0;
}

// ------------------------------------------------------------- ingest ring

GPUD_API int gpud_ingest_create(int block_samples, int depth, gpud_ingest** out)
{
    if (!out || block_samples <= 0 || depth < 2 || depth > 64) return -1;
    auto* g = new gpud_ingest();
    g->block = block_samples;
    int err = (int)hipGetDevice(&g->device);
    if (!err) err = (int)hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    g->slots.resize((size_t)depth);
    for (auto& sl : g->slots) {
        if (err) break;
        err = (int)hipHostMalloc((void**)&sl.host, (size_t)block_samples * sizeof(gpud_float2), hipHostMallocDefault);
        if (!err) err = (int)hipMalloc((void**)&sl.dev, (size_t)block_samples * sizeof(gpud_float2));
        if (!err) err = (int)hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    }
    if (err) {
        for (auto& sl : g->slots) {
            if (sl.host) hipHostFree(sl.host);
            if (sl.dev) hipFree(sl.dev);
            if (sl.ev) hipEventDestroy(sl.ev);
        }
        if (g->stream) hipStreamDestroy(g->stream);
        delete g;
        return err;
    }
    *out = g;
    return 0;
}

GPUD_API int gpud_ingest_destroy(gpud_ingest* g)
{
    if (!g) return 0;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        g->stop = true;
    }
    g->cv.notify_all();
    if (g->producer.joinable()) g->producer.join();
    hipStreamSynchronize(g->stream);
    for (auto& sl : g->slots) {
        hipHostFree(sl.host);
        hipFree(sl.dev);
        hipEventDestroy(sl.ev);
    }
    hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

// Replay a capture: a producer thread reads block after block (looping at
// the end, as the replay Source does) and copies them ahead of the consumer.
// The file must stay open until the ingest is destroyed.  Unpaced: the ring
// is kept full.
GPUD_API int gpud_ingest_start_file(gpud_ingest* g, gpud_iqfile* f)
{
    if (!g || !f) return -1;
    std::lock_guard<std::mutex> lk(g->mu);
    if (g->producer.joinable()) return -1;
    g->file = f;
    g->producer = std::thread([g] {
        (void)hipSetDevice(g->device);
        for (;;) {
            const int err = ingest_produce(g, g->block, [g](gpud_float2* host, int n) {
                return gpud_iqfile_read(g->file, host, n) == 0;
            });
            if (err) return;  // stopped, or the error is kept for the consumer
        }
    });
    return 0;
}

// Push one block (n <= block_samples) from the caller's source; waits while
// the ring is full.  Not with a replaying capture.
GPUD_API int gpud_ingest_push(gpud_ingest* g, const gpud_float2* iq, int n)
{
    if (!g || n < 0 || n > g->block || (n > 0 && !iq)) return -1;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        if (g->producer.joinable()) return -1;
        if (g->error) return g->error;
    }
    return ingest_produce(g, n, [iq](gpud_float2* host, int m) {
        if (m > 0) std::memcpy(host, iq, (size_t)m * sizeof(gpud_float2));
        return true;
    });
}

// The next block in order, resident on the device: *dev stays valid until
// gpud_ingest_release(seq).  Waits for the block to be produced and its copy
// to finish.  timeout_ms < 0 waits without bound; 1 = timed out (nothing
// acquired).
GPUD_API int gpud_ingest_acquire(gpud_ingest* g, const gpud_float2** dev, int* n, int64_t* seq, int timeout_ms)
{
    if (!g || !dev || !n || !seq) return -1;
    IngestSlot* sl = nullptr;
    {
        std::unique_lock<std::mutex> lk(g->mu);
        auto ready = [&] {
            IngestSlot& s = g->slots[(size_t)(g->consumed % g->slots.size())];
            return g->error != 0 || (s.state == kReady && s.seq == g->consumed);
        };
        if (timeout_ms < 0)
            g->cv.wait(lk, ready);
        else if (!g->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
            return 1;
        if (g->error) return g->error;
        sl = &g->slots[(size_t)(g->consumed % g->slots.size())];
        sl->state = kAcquired;
        *seq = g->consumed++;
    }
    ARGUS_CHECK(hipEventSynchronize(sl->ev));
    *dev = sl->dev;
    *n = sl->n;
    return 0;
}

// Hand a block's slot back to the producer.
GPUD_API int gpud_ingest_release(gpud_ingest* g, int64_t seq)
{
    if (!g || seq < 0) return -1;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        IngestSlot& sl = g->slots[(size_t)(seq % (int64_t)g->slots.size())];
        if (sl.seq != seq || sl.state != kAcquired) return -1;
        sl.state = kFree;
    }
    g->cv.notify_all();
    return 0;
}

GPUD_API int gpud_ingest_stats(gpud_ingest* g, int64_t* produced, int64_t* consumed, int* queued)
{
    if (!g) return -1;
    std::lock_guard<std::mutex> lk(g->mu);
    if (produced) *produced = g->produced;
    if (consumed) *consumed = g->consumed;
    if (queued) *queued = (int)(g->produced - g->consumed);
    return 0;
}

}  // extern "C"
