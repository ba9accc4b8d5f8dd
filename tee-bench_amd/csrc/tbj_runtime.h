// Host runtime: device context, cached workspace, error plumbing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <unordered_map>

#include "../../include/tbj.h"

namespace tbj {

struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Device workspace, grown on demand and kept between calls.  `cap` (bytes,
// 0 = none) bounds the total and turns an over-cap request into TBJ_ERR_OOM.
class Workspace {
public:
    int get(const char* name, size_t bytes, void** out);
    void release_all();
    size_t total() const { return total_; }
    size_t cap = 0;

private:
    std::unordered_map<std::string, Buf> bufs_;
    size_t total_ = 0;
};

struct Context {
    bool ready = false;
    int device = -1;
    int num_cus = 0;
    hipStream_t own_stream = nullptr;
    // a second stream for work beside the caller's (forked and joined
    // through the two events)
    hipStream_t side_stream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    Workspace ws;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t pev[24] = {};
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

Context& ctx();

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

// Algorithms (implemented in the tbj_*.cpp files).
int run_radix_join(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
                   tbj_pair* pairs, uint64_t pairs_cap, const tbj_options& opt, tbj_result* res,
                   hipStream_t st);
int run_pht(const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS, tbj_pair* pairs,
            uint64_t pairs_cap, const tbj_options& opt, tbj_result* res, hipStream_t st);
int run_sm64(int algo, const tbj_tuple64* R, uint64_t nR, const tbj_tuple64* S, uint64_t nS,
             tbj_pair64* pairs, uint64_t pairs_cap, const tbj_options& opt, tbj_result* res,
             hipStream_t st);
int run_cht(const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS, tbj_pair* pairs,
            uint64_t pairs_cap, const tbj_options& opt, tbj_result* res, hipStream_t st);

}  // namespace tbj
