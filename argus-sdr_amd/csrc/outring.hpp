// outring.hpp -- per-signal output rings and their shared free pool (host;
// semantics in outring.hip).
#pragma once
#include "../../include/argus_gpu.h"

#include <stdint.h>
#include <unordered_map>
#include <vector>

#define ARGUS_OUT_RING_DEPTH GPUD_OUT_RING_DEPTH

struct ArgusOutBuf {
    gpud_float2* p = nullptr;
    int64_t cap = 0;  // samples the buffer holds
};

struct ArgusOutRing {
    ArgusOutBuf bufs[ARGUS_OUT_RING_DEPTH];
    int idx = 0;  // the slot the next call takes
};

struct gpud_out_rings {
    std::unordered_map<int64_t, ArgusOutRing> rings;
    std::vector<ArgusOutBuf> free;
    int64_t allocated = 0;  // buffers allocated so far
    int64_t released = 0;   // buffers freed so far
};

// The ring of signal_id takes its next buffer, grown to hold n samples.
int argus_out_ring_next(gpud_out_rings* o, int64_t signal_id, int64_t n, gpud_float2** buf);
// Rings of signals outside active[0..n_active) go back to the pool.
void argus_out_rings_prune(gpud_out_rings* o, const int64_t* active, int n_active);
void argus_out_rings_free(gpud_out_rings* o);
