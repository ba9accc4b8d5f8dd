// C ABI entry points of libtbj.so (declared in include/tbj.h).
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "tbj_runtime.h"

namespace tbj {

static Context g_ctx;
static std::mutex g_mu;
static char g_last_error[512];

Context& ctx() { return g_ctx; }

void set_last_error(const char* what, hipError_t e, const char* file, int line) {
    snprintf(g_last_error, sizeof(g_last_error), "%s failed: %s (%s:%d)", what,
             hipGetErrorString(e), file, line);
}

int Workspace::get(const char* name, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
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

int PhaseTimer::begin() {
    if (!on_) return TBJ_OK;
    for (auto& e : c_.pev)
        if (!e) TBJ_HIP(hipEventCreate(&e));
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

void Workspace::release_all() {
    for (auto& kv : bufs_)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    bufs_.clear();
    total_ = 0;
}

static int init_locked(int device) {
    if (g_ctx.ready && g_ctx.device == device) return TBJ_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        snprintf(g_last_error, sizeof(g_last_error), "no HIP device visible");
        return TBJ_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return TBJ_ERR_INVALID;
    TBJ_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TBJ_HIP(hipGetDeviceProperties(&prop, device));
    g_ctx.device = device;
    g_ctx.num_cus = prop.multiProcessorCount;
    for (auto& e : g_ctx.ev)
        if (!e) TBJ_HIP(hipEventCreate(&e));
    // default priority; used to capture HIP graphs when the caller's stream
    // is the legacy default stream, which cannot be captured
    if (!g_ctx.own_stream) TBJ_HIP(hipStreamCreateWithFlags(&g_ctx.own_stream, hipStreamNonBlocking));
    if (!g_ctx.side_stream)
        TBJ_HIP(hipStreamCreateWithFlags(&g_ctx.side_stream, hipStreamNonBlocking));
    if (!g_ctx.fork_ev) TBJ_HIP(hipEventCreateWithFlags(&g_ctx.fork_ev, hipEventDisableTiming));
    if (!g_ctx.join_ev) TBJ_HIP(hipEventCreateWithFlags(&g_ctx.join_ev, hipEventDisableTiming));
    g_ctx.ready = true;
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
    g_ctx.ws.release_all();
}

void tbj_default_options(tbj_options* opt) {
    if (!opt) return;
    memset(opt, 0, sizeof(*opt));
    opt->partitioner = TBJ_PART_AUTO;
    opt->materialize = 0;
}

const char* tbj_status_str(int s) {
    switch (s) {
        case TBJ_OK: return "ok";
        case TBJ_ERR_INVALID: return "invalid argument";
        case TBJ_ERR_NO_DEVICE: return "no HIP device";
        case TBJ_ERR_OOM: return "out of device memory";
        case TBJ_ERR_HIP: return g_last_error[0] ? g_last_error : "HIP runtime error";
        case TBJ_ERR_UNSUPPORTED: return "unsupported algorithm or option";
        case TBJ_ERR_TOO_LARGE: return "relation too large";
        default: return "unknown status";
    }
}

static const char* kAlgoNames[TBJ_ALGO_COUNT] = {"RHO", "RHT", "RJ",         "PHT",   "CHT",
                                                 "RSM", "PSM", "RHO_atomic", "NPO_st"};

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
    if (!res || algo < 0 || algo >= TBJ_ALGO_COUNT) return TBJ_ERR_INVALID;
    if ((nR && !R) || (nS && !S)) return TBJ_ERR_INVALID;
    tbj_options o;
    if (opt) o = *opt;
    else tbj_default_options(&o);
    if (o.materialize && pairs_cap && !pairs) return TBJ_ERR_INVALID;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ctx.ready) {
        int s = init_locked(0);
        if (s != TBJ_OK) return s;
    }
    memset(res, 0, sizeof(*res));
    g_ctx.ws.cap = o.workspace_cap;
    hipStream_t st = (hipStream_t)stream;
    switch (algo) {
        case TBJ_RHO:
        case TBJ_RHT:
        case TBJ_RJ:
        case TBJ_RHO_ATOMIC:
        case TBJ_RSM:
            return run_radix_join(algo, R, nR, S, nS, pairs, pairs_cap, o, res, st);
        case TBJ_PSM:
            // global sort order: range-partition by the learned key CDF
            if (o.partitioner == TBJ_PART_AUTO) o.partitioner = TBJ_PART_LEARNED;
            return run_radix_join(algo, R, nR, S, nS, pairs, pairs_cap, o, res, st);
        case TBJ_PHT:
        case TBJ_NPO_ST:
            return run_pht(R, nR, S, nS, pairs, pairs_cap, o, res, st);
        case TBJ_CHT:
            return run_cht(R, nR, S, nS, pairs, pairs_cap, o, res, st);
        default:
            return TBJ_ERR_UNSUPPORTED;
    }
}

int tbj_join_host(int algo, const tbj_tuple* R, uint64_t nR, const tbj_tuple* S, uint64_t nS,
                  tbj_pair* pairs, uint64_t pairs_cap, const tbj_options* opt,
                  tbj_result* res) {
    if (!res) return TBJ_ERR_INVALID;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (!g_ctx.ready) {
            int s = init_locked(0);
            if (s != TBJ_OK) return s;
        }
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
    if (!res || algo < 0 || algo >= TBJ_ALGO_COUNT) return TBJ_ERR_INVALID;
    if ((nR && !R) || (nS && !S)) return TBJ_ERR_INVALID;
    if (algo != TBJ_PSM && algo != TBJ_RSM) return TBJ_ERR_UNSUPPORTED;
    tbj_options o;
    if (opt) o = *opt;
    else tbj_default_options(&o);
    if (o.materialize && pairs_cap && !pairs) return TBJ_ERR_INVALID;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ctx.ready) {
        int s = init_locked(0);
        if (s != TBJ_OK) return s;
    }
    memset(res, 0, sizeof(*res));
    g_ctx.ws.cap = o.workspace_cap;
    return run_sm64(algo, R, nR, S, nS, pairs, pairs_cap, o, res, (hipStream_t)stream);
}

}  // extern "C"
