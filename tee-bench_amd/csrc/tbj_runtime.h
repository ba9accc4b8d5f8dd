// Host runtime: device state, per-stream sessions with cached workspace,
// error plumbing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/tbj.h"

namespace tbj {

struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Device workspace of one session.  Normally buffers are grown on demand
// with hipMalloc and kept between calls, keyed by name; `cap` (bytes, 0 =
// none) bounds the total and turns an over-cap request into TBJ_ERR_OOM.
// In arena mode (set_arena) every request is carved from one fixed region
// instead, front to back, and nothing is allocated: reset() starts a call
// over, and a request past the region's end returns TBJ_ERR_TOO_LARGE.
class Workspace {
public:
    int get(const char* name, size_t bytes, void** out);
    void release_all();
    size_t total() const { return total_; }
    size_t cap = 0;

    void set_arena(void* base, size_t bytes) {
        arena_ = (uint8_t*)base;
        arena_bytes_ = bytes;
        arena_used_ = 0;
        arena_n_ = 0;
    }
    bool arena() const { return arena_ != nullptr; }
    void reset() {
        arena_used_ = 0;
        arena_n_ = 0;
    }

private:
    std::unordered_map<std::string, Buf> bufs_;
    size_t total_ = 0;
    // arena mode: fixed table of the current call's named buffers (no host
    // allocation either)
    static constexpr int kArenaNames = 64;
    uint8_t* arena_ = nullptr;
    size_t arena_bytes_ = 0, arena_used_ = 0;
    int arena_n_ = 0;
    const char* arena_name_[kArenaNames] = {};
    Buf arena_buf_[kArenaNames];
};

// Process-wide state of the device the library runs on.
struct Device {
    bool ready = false;
    int device = -1;
    int num_cus = 0;
    // default priority; used to capture HIP graphs when the caller's stream
    // is the legacy default stream, which cannot be captured
    hipStream_t own_stream = nullptr;
};

// One session per caller stream: its own workspace, events and side stream,
// so that calls on different streams (from different host threads) run
// concurrently; calls on one stream serialise on `mu`.
struct Context {
    std::mutex mu;
    int num_cus = 0;
    // a second stream for work beside the caller's (forked and joined
    // through the two events)
    hipStream_t side_stream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // a few device scalars the host reads in the middle of a call while the
    // stream runs on (pinned, allocated by the first call that needs it);
    // host_ev keeps the system-scope release so that the copy is visible
    uint32_t* pinned = nullptr;
    hipEvent_t host_ev = nullptr;
    Workspace ws;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t pev[24] = {};
    // creates the streams and events (every event, the per-phase ones too,
    // so that a call allocates nothing on the device afterwards)
    int setup(int num_cus);
    void destroy();
};

// Per-phase device timing (tbj_options.profile): mark(phase) records an event
// that closes the time since the previous mark into phase_ms[phase].
class PhaseTimer {
public:
    PhaseTimer(Context& c, hipStream_t st, bool on) : c_(c), st_(st), on_(on) {}
    int begin();
    int mark(int phase);
    int finish(tbj_result* res);

private:
    Context& c_;
    hipStream_t st_;
    bool on_;
    int n_ = 0;
    int phase_[24] = {};
};

Device& dev();

#define TBJ_HIP(call)                                            \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) {                                  \
            tbj::set_last_error(#call, e_, __FILE__, __LINE__);   \
            return TBJ_ERR_HIP;                                  \
        }                                                        \
    } while (0)

#define TBJ_TRY(expr)            \
    do {                         \
        int s_ = (expr);         \
        if (s_ != TBJ_OK) return s_; \
    } while (0)

void set_last_error(const char* what, hipError_t e, const char* file, int line);
void set_last_message(const char* msg);

// Algorithms (implemented in the tbj_*.cpp files), each on one session.
int run_radix_join(Context& C, int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S,
                   uint64_t nS, void* out, uint64_t out_cap, const tbj_options& opt,
                   tbj_result* res, hipStream_t st);
// R alone through the exact radix passes on key bits [shift, shift + bits)
// (1 <= bits <= 22; the bulk build of a no-partitioning join's table):
// *Rb = the partitioned copy (session workspace), (*bounds)[p] = {start, end}
// of partition p = (key bits [shift, shift + b1)) * 2^(bits - b1) + (key bits
// [shift + b1, shift + bits)), p < 2^bits; *b1 = the first pass's bits (bits
// itself for one pass).
int partition_relation(Context& C, const tbj_tuple* R, uint32_t nR, uint32_t bits,
                       uint32_t shift, hipStream_t st, const uint2** Rb, const uint2** bounds,
                       uint32_t* b1);
int run_pht(Context& C, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
            tbj_pair* pairs, uint64_t pairs_cap, const tbj_options& opt, tbj_result* res,
            hipStream_t st);
int run_sm64(Context& C, int algo, const tbj_tuple64* R, uint64_t nR, const tbj_tuple64* S,
             uint64_t nS, tbj_pair64* pairs, uint64_t pairs_cap, const tbj_options& opt,
             tbj_result* res, hipStream_t st);
int run_cht(Context& C, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
            tbj_pair* pairs, uint64_t pairs_cap, const tbj_options& opt, tbj_result* res,
            hipStream_t st);

}  // namespace tbj
