// C ABI entry points of libtbj.so (declared in include/tbj.h).
//
// State: one Device (the GPU the library runs on) and one session per caller
// stream (workspace, events, side stream), so that calls on different streams
// from different host threads do not serialise on each other.  A lock guards
// only the session table; each call holds its own session's lock.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>

#include "tbj_runtime.h"

namespace tbj {

static Device g_dev;
static std::mutex g_mu;  // g_dev, the session table, the scratch addresses
static std::unordered_map<void*, std::unique_ptr<Context>> g_sessions;
static thread_local char g_last_error[512];

Device& dev() { return g_dev; }

void set_last_error(const char* what, hipError_t e, const char* file, int line) {
    snprintf(g_last_error, sizeof(g_last_error), "%s failed: %s (%s:%d)", what,
             hipGetErrorString(e), file, line);
}

void set_last_message(const char* msg) { snprintf(g_last_error, sizeof(g_last_error), "%s", msg); }

// ---------------------------------------------------------------------------
// Fixed device scratch of the no-allocation entry point: one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a host-side stand-in, not a device
// address): each device's address comes from hipGetSymbolAddress at
// tbj_init, with that device current, and only that address is used.
constexpr size_t kNoAllocBytes = (size_t)TBJ_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::unique_ptr<Context> g_noalloc[kMaxDevices];

int Workspace::get(const char* name, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    if (arena_) {
        for (int i = 0; i < arena_n_; i++) {
            if (!strcmp(arena_name_[i], name)) {
                if (arena_buf_[i].bytes >= bytes) {
                    *out = arena_buf_[i].ptr;
                    return TBJ_OK;
                }
                break;  // larger than before: carve a new block
            }
        }
        const size_t want = (bytes + 255) & ~(size_t)255;
        if (arena_used_ + want > arena_bytes_ || arena_n_ >= kArenaNames) {
            set_last_message("no-allocation scratch exhausted (workspace above "
                             "TBJ_NOALLOC_SCRATCH_MB)");
            return TBJ_ERR_TOO_LARGE;
        }
        arena_name_[arena_n_] = name;
        arena_buf_[arena_n_].ptr = arena_ + arena_used_;
        arena_buf_[arena_n_].bytes = want;
        arena_n_++;
        *out = arena_ + arena_used_;
        arena_used_ += want;
        return TBJ_OK;
    }
    Buf& b = bufs_[name];
    if (b.bytes >= bytes) {
        *out = b.ptr;
        return TBJ_OK;
    }
    // grow with 12.5% headroom so that slightly larger calls reuse the buffer
    size_t want = bytes + bytes / 8;
    want = (want + 4095) & ~(size_t)4095;
    if (cap && total_ - b.bytes + want > cap) {
        want = bytes;
        if (total_ - b.bytes + want > cap) return TBJ_ERR_OOM;
    }
    if (b.ptr) {
        (void)hipFree(b.ptr);
        total_ -= b.bytes;
        b.ptr = nullptr;
        b.bytes = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        return TBJ_ERR_OOM;
    }
    b.ptr = p;
    b.bytes = want;
    total_ += want;
    *out = p;
    return TBJ_OK;
}

void Workspace::release_all() {
    if (arena_) {
        reset();
        return;
    }
    for (auto& kv : bufs_)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    bufs_.clear();
    total_ = 0;
}

int Context::setup(int cus) {
    num_cus = cus;
    // call and phase timing: read after the call's stream synchronisation,
    // so no system-scope fence (cache writeback) at each record, which
    // stalls the work behind it
    for (auto& e : ev)
        if (!e) TBJ_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    for (auto& e : pev)
        if (!e) TBJ_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    if (!side_stream) TBJ_HIP(hipStreamCreateWithFlags(&side_stream, hipStreamNonBlocking));
    if (!fork_ev) TBJ_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    if (!join_ev) TBJ_HIP(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming));
    if (!host_ev) TBJ_HIP(hipEventCreateWithFlags(&host_ev, hipEventDisableTiming));
    return TBJ_OK;
}

void Context::destroy() {
    ws.release_all();
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    for (auto& e : pev)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (side_stream) (void)hipStreamDestroy(side_stream), side_stream = nullptr;
    if (fork_ev) (void)hipEventDestroy(fork_ev), fork_ev = nullptr;
    if (join_ev) (void)hipEventDestroy(join_ev), join_ev = nullptr;
    if (host_ev) (void)hipEventDestroy(host_ev), host_ev = nullptr;
    if (pinned) (void)hipHostFree(pinned), pinned = nullptr;
}

int PhaseTimer::begin() {
    if (!on_) return TBJ_OK;
    n_ = 0;
    TBJ_HIP(hipEventRecord(c_.pev[n_++], st_));
    return TBJ_OK;
}

int PhaseTimer::mark(int phase) {
    if (!on_ || n_ >= 24) return TBJ_OK;
    phase_[n_] = phase;
    TBJ_HIP(hipEventRecord(c_.pev[n_++], st_));
    return TBJ_OK;
}

int PhaseTimer::finish(tbj_result* res) {
    if (!on_) return TBJ_OK;
    TBJ_HIP(hipEventSynchronize(c_.pev[n_ - 1]));
    for (int i = 1; i < n_; i++) {
        float ms = 0;
        TBJ_HIP(hipEventElapsedTime(&ms, c_.pev[i - 1], c_.pev[i]));
        res->phase_ms[phase_[i]] += ms;
    }
    return TBJ_OK;
}

static void release_sessions_locked() {
    for (auto& kv : g_sessions) {
        std::lock_guard<std::mutex> g(kv.second->mu);
        kv.second->destroy();
    }
    g_sessions.clear();
}

static int init_locked(int device) {
    if (g_dev.ready && g_dev.device == device) return TBJ_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        set_last_message("no HIP device visible");
        return TBJ_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n || device >= kMaxDevices) return TBJ_ERR_INVALID;
    // another device than before: the old device's sessions go
    if (g_dev.ready) release_sessions_locked();
    TBJ_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TBJ_HIP(hipGetDeviceProperties(&prop, device));
    g_dev.device = device;
    g_dev.num_cus = prop.multiProcessorCount;
    if (!g_dev.own_stream) TBJ_HIP(hipStreamCreateWithFlags(&g_dev.own_stream, hipStreamNonBlocking));
    // this device's copy of the no-allocation scratch (device current)
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        TBJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_last_message("hipGetSymbolAddress returned no address for the scratch");
            return TBJ_ERR_HIP;
        }
        g_scratch_ptr[device] = p;
    }
    if (!g_noalloc[device]) {
        std::unique_ptr<Context> c(new Context());
        TBJ_TRY(c->setup(g_dev.num_cus));
        c->ws.set_arena(g_scratch_ptr[device], kNoAllocBytes);
        g_noalloc[device] = std::move(c);
    }
    g_dev.ready = true;
    return TBJ_OK;
}

// The session of `stream` (created on first use); initialises device 0 when
// the library was not initialised.
static int session(void* stream, Context** out) {
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_dev.ready) TBJ_TRY(init_locked(0));
    auto it = g_sessions.find(stream);
    if (it == g_sessions.end()) {
        std::unique_ptr<Context> c(new Context());
        TBJ_TRY(c->setup(g_dev.num_cus));
        it = g_sessions.emplace(stream, std::move(c)).first;
    }
    *out = it->second.get();
    return TBJ_OK;
}

static bool radix_family(int algo) {
    return algo == TBJ_RHO || algo == TBJ_RHT || algo == TBJ_RJ || algo == TBJ_RHO_ATOMIC ||
           algo == TBJ_RSM || algo == TBJ_PSM || algo == TBJ_MWAY;
}

static int dispatch(Context& C, int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S,
                    uint64_t nS, void* out, uint64_t out_cap, tbj_options o, tbj_result* res,
                    hipStream_t st) {
    memset(res, 0, sizeof(*res));
    C.ws.cap = o.workspace_cap;
    if (o.output != TBJ_OUT_PAIRS && o.output != TBJ_OUT_RELATION) return TBJ_ERR_INVALID;
    if (o.output == TBJ_OUT_RELATION && !radix_family(algo)) return TBJ_ERR_UNSUPPORTED;
    switch (algo) {
        case TBJ_RHO:
        case TBJ_RHT:
        case TBJ_RJ:
        case TBJ_RHO_ATOMIC:
        case TBJ_RSM:
        case TBJ_MWAY:
            return run_radix_join(C, algo, R, nR, S, nS, out, out_cap, o, res, st);
        case TBJ_PSM:
            // global sort order: range-partition by the learned key CDF
            if (o.partitioner == TBJ_PART_AUTO) o.partitioner = TBJ_PART_LEARNED;
            return run_radix_join(C, algo, R, nR, S, nS, out, out_cap, o, res, st);
        case TBJ_PHT:
        case TBJ_NPO_ST:
            return run_pht(C, R, nR, S, nS, (tbj_pair*)out, out_cap, o, res, st);
        case TBJ_CHT:
            return run_cht(C, R, nR, S, nS, (tbj_pair*)out, out_cap, o, res, st);
        default:
            return TBJ_ERR_UNSUPPORTED;
    }
}

static int check_args(int algo, const void* R, uint64_t nR, const void* S, uint64_t nS,
                      const void* out, uint64_t out_cap, const tbj_options& o, tbj_result* res) {
    if (!res || algo < 0 || algo >= TBJ_ALGO_COUNT) return TBJ_ERR_INVALID;
    if ((nR && !R) || (nS && !S)) return TBJ_ERR_INVALID;
    if (o.materialize && out_cap && !out) return TBJ_ERR_INVALID;
    return TBJ_OK;
}

}  // namespace tbj

using namespace tbj;

extern "C" {

int tbj_init(int device) {
    std::lock_guard<std::mutex> g(g_mu);
    return init_locked(device);
}

void tbj_shutdown(void) {
    std::lock_guard<std::mutex> g(g_mu);
    release_sessions_locked();
}

int tbj_release_stream(void* stream) {
    std::unique_ptr<Context> c;
    {
        std::lock_guard<std::mutex> g(g_mu);
        auto it = g_sessions.find(stream);
        if (it == g_sessions.end()) return TBJ_OK;
        c = std::move(it->second);
        g_sessions.erase(it);
    }
    std::lock_guard<std::mutex> g(c->mu);  // waits for a call still running on it
    c->destroy();
    return TBJ_OK;
}

void tbj_default_options(tbj_options* opt) {
    if (!opt) return;
    memset(opt, 0, sizeof(*opt));
    opt->partitioner = TBJ_PART_AUTO;
    opt->materialize = 0;
    opt->output = TBJ_OUT_PAIRS;
}

const char* tbj_status_str(int s) {
    switch (s) {
        case TBJ_OK: return "ok";
        case TBJ_ERR_INVALID: return "invalid argument";
        case TBJ_ERR_NO_DEVICE: return g_last_error[0] ? g_last_error : "no HIP device";
        case TBJ_ERR_OOM: return "out of device memory";
        case TBJ_ERR_HIP: return g_last_error[0] ? g_last_error : "HIP runtime error";
        case TBJ_ERR_UNSUPPORTED: return "unsupported algorithm or option";
        case TBJ_ERR_TOO_LARGE: return g_last_error[0] ? g_last_error : "relation too large";
        default: return "unknown status";
    }
}

static const char* kAlgoNames[TBJ_ALGO_COUNT] = {"RHO", "RHT", "RJ",         "PHT",   "CHT",
                                                 "RSM", "PSM", "RHO_atomic", "NPO_st", "MWAY"};

const char* tbj_algo_name(int algo) {
    if (algo < 0 || algo >= TBJ_ALGO_COUNT) return nullptr;
    return kAlgoNames[algo];
}

int tbj_algo_from_name(const char* name) {
    if (!name) return TBJ_ERR_INVALID;
    for (int i = 0; i < TBJ_ALGO_COUNT; i++)
        if (!strcmp(name, kAlgoNames[i])) return i;
    return TBJ_ERR_INVALID;
}

int tbj_join(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
             tbj_pair* pairs, uint64_t pairs_cap, const tbj_options* opt, tbj_result* res,
             void* stream) {
    tbj_options o;
    if (opt) o = *opt;
    else tbj_default_options(&o);
    TBJ_TRY(check_args(algo, R, nR, S, nS, pairs, pairs_cap, o, res));
    Context* C = nullptr;
    TBJ_TRY(session(stream, &C));
    std::lock_guard<std::mutex> g(C->mu);
    return dispatch(*C, algo, R, nR, S, nS, pairs, pairs_cap, o, res, (hipStream_t)stream);
}

int tbj_join_noalloc(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
                     tbj_pair* pairs, uint64_t pairs_cap, const tbj_options* opt,
                     tbj_result* res, void* stream) {
    tbj_options o;
    if (opt) o = *opt;
    else tbj_default_options(&o);
    TBJ_TRY(check_args(algo, R, nR, S, nS, pairs, pairs_cap, o, res));
    if (!radix_family(algo)) return TBJ_ERR_UNSUPPORTED;
    if (nR + nS > (uint64_t)TBJ_NOALLOC_MAX_TUPLES) {
        set_last_message("tbj_join_noalloc: nR + nS above TBJ_NOALLOC_MAX_TUPLES");
        return TBJ_ERR_TOO_LARGE;
    }
    // sample-sized regions need more room than exact partitions
    o.sampled_regions = -1;
    o.workspace_cap = 0;
    Context* C = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (!g_dev.ready) {
            set_last_message("tbj_join_noalloc: call tbj_init first");
            return TBJ_ERR_NO_DEVICE;
        }
        C = g_noalloc[g_dev.device].get();
    }
    std::lock_guard<std::mutex> g(C->mu);
    C->ws.reset();
    return dispatch(*C, algo, R, nR, S, nS, pairs, pairs_cap, o, res, (hipStream_t)stream);
}

int tbj_join3(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
              const tbj_tuple* T, uint64_t nT, tbj_tuple* out, uint64_t out_cap,
              const tbj_options* opt, tbj_result* res, void* stream) {
    tbj_options o;
    if (opt) o = *opt;
    else tbj_default_options(&o);
    o.materialize = 1;
    o.output = TBJ_OUT_RELATION;
    TBJ_TRY(check_args(algo, R, nR, S, nS, out, out_cap, o, res));
    if ((nT && !T) || !radix_family(algo)) return nT && !T ? TBJ_ERR_INVALID : TBJ_ERR_UNSUPPORTED;
    Context* C = nullptr;
    TBJ_TRY(session(stream, &C));
    std::lock_guard<std::mutex> g(C->mu);
    hipStream_t st = (hipStream_t)stream;
    // join 1 writes out1 = (key, R.payload) into the session's workspace,
    // sized for one match per S tuple and redone once with the exact count
    // when R's keys repeat
    uint64_t cap1 = std::max<uint64_t>(1, std::max(nR, nS));
    tbj_tuple* mid = nullptr;
    tbj_result r1;
    for (int attempt = 0; attempt < 2; attempt++) {
        TBJ_TRY(C->ws.get("j3_mid", sizeof(tbj_tuple) * cap1, (void**)&mid));
        TBJ_TRY(dispatch(*C, algo, R, nR, S, nS, mid, cap1, o, &r1, st));
        if (r1.matches <= cap1) break;
        cap1 = r1.matches;
    }
    // join 2: T as the build side, out1 as the probe side (secure_joins.cpp:383)
    TBJ_TRY(dispatch(*C, algo, T, nT, mid, r1.pairs_written, out, out_cap, o, res, st));
    res->ms_total += r1.ms_total;
    return TBJ_OK;
}

int tbj_join_host(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
                  tbj_pair* pairs, uint64_t pairs_cap, const tbj_options* opt,
                  tbj_result* res) {
    if (!res) return TBJ_ERR_INVALID;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (!g_dev.ready) TBJ_TRY(init_locked(0));
    }
    const bool mat = opt && opt->materialize && pairs && pairs_cap;
    void *dR = nullptr, *dS = nullptr, *dP = nullptr;
    int rc = TBJ_OK;
    if (hipMalloc(&dR, (nR ? nR : 1) * sizeof(tbj_tuple)) != hipSuccess ||
        hipMalloc(&dS, (nS ? nS : 1) * sizeof(tbj_tuple)) != hipSuccess ||
        (mat && hipMalloc(&dP, pairs_cap * sizeof(tbj_pair)) != hipSuccess)) {
        (void)hipGetLastError();
        rc = TBJ_ERR_OOM;
    }
    if (rc == TBJ_OK && nR && hipMemcpy(dR, R, nR * sizeof(tbj_tuple), hipMemcpyHostToDevice) != hipSuccess)
        rc = TBJ_ERR_HIP;
    if (rc == TBJ_OK && nS && hipMemcpy(dS, S, nS * sizeof(tbj_tuple), hipMemcpyHostToDevice) != hipSuccess)
        rc = TBJ_ERR_HIP;
    if (rc == TBJ_OK)
        rc = tbj_join(algo, (const tbj_tuple*)dR, nR, (const tbj_tuple*)dS, nS, (tbj_pair*)dP,
                      mat ? pairs_cap : 0, opt, res, nullptr);
    if (rc == TBJ_OK && mat && res->pairs_written &&
        hipMemcpy(pairs, dP, res->pairs_written * sizeof(tbj_pair), hipMemcpyDeviceToHost) != hipSuccess)
        rc = TBJ_ERR_HIP;
    if (dR) (void)hipFree(dR);
    if (dS) (void)hipFree(dS);
    if (dP) (void)hipFree(dP);
    return rc;
}

int tbj_join64(int algo, const tbj_tuple64* R, uint64_t nR, const tbj_tuple64* S, uint64_t nS,
               tbj_pair64* pairs, uint64_t pairs_cap, const tbj_options* opt, tbj_result* res,
               void* stream) {
    tbj_options o;
    if (opt) o = *opt;
    else tbj_default_options(&o);
    TBJ_TRY(check_args(algo, R, nR, S, nS, pairs, pairs_cap, o, res));
    if (algo == TBJ_MWAY) algo = TBJ_RSM;
    if (algo != TBJ_PSM && algo != TBJ_RSM) return TBJ_ERR_UNSUPPORTED;
    if (o.output != TBJ_OUT_PAIRS) return TBJ_ERR_UNSUPPORTED;
    Context* C = nullptr;
    TBJ_TRY(session(stream, &C));
    std::lock_guard<std::mutex> g(C->mu);
    memset(res, 0, sizeof(*res));
    C->ws.cap = o.workspace_cap;
    return run_sm64(*C, algo, R, nR, S, nS, pairs, pairs_cap, o, res, (hipStream_t)stream);
}

}  // extern "C"
