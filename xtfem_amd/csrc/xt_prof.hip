// Per-launch kernel timing with HIP events on the launching stream.
//
// When a kind is enabled, its launch sites bracket each kernel launch with an event
// pair and add the launch's algorithmic bytes and flops (each matrix and vector
// element the kernel's format needs, read once and written once).  xt_prof_read
// synchronises on the last event and sums the per-launch durations.  Process-wide
// state: enable, run, read from one host thread.
#include <vector>

#include <atomic>

#include "xt_common.h"

namespace xt {
namespace {
struct Kind {
    std::vector<hipEvent_t> ev;   // pairs: ev[2i] start, ev[2i+1] stop
    size_t used = 0;              // events recorded since enable
    double bytes = 0.0, flops = 0.0;          // as the launches perform them
    double alg_bytes = 0.0, alg_flops = 0.0;  // algorithmic (the unpaired computation)
};
unsigned g_mask = 0;
Kind g_kind[XT_PROF_NKINDS];

hipEvent_t next_event(Kind &k) {
    if (k.used == k.ev.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        k.ev.push_back(e);
    }
    return k.ev[k.used++];
}
}  // namespace

bool prof_on(int kind) { return (g_mask >> kind) & 1u; }

namespace {
bool g_graphs = true;
int g_knob[XT_TUNE_NKNOBS] = {};   // 0 = the built-in choice
}  // namespace
std::atomic<unsigned long long> g_dev_epoch{0};
bool graphs_on() { return g_graphs; }
int tune_knob(int knob) { return (knob >= 0 && knob < XT_TUNE_NKNOBS) ? g_knob[knob] : 0; }
// bumped by every knob change: captured graphs keyed on it never replay another variant
static std::atomic<long long> g_knob_epoch{0};
long long tune_epoch() { return g_knob_epoch.load(std::memory_order_relaxed); }

void prof_begin(int kind, hipStream_t s) {
    if (!prof_on(kind)) return;
    Kind &k = g_kind[kind];
    if (k.used % 2) return;                      // unmatched begin: keep the pair structure
    hipEvent_t e = next_event(k);
    if (e) (void)hipEventRecord(e, s);
}

void prof_end(int kind, hipStream_t s, double bytes, double flops, bool alg) {
    if (!prof_on(kind)) return;
    Kind &k = g_kind[kind];
    if (k.used % 2 == 0) return;
    hipEvent_t e = next_event(k);
    if (!e) { --k.used; return; }
    (void)hipEventRecord(e, s);
    k.bytes += bytes;
    k.flops += flops;
    if (alg) {
        k.alg_bytes += bytes;
        k.alg_flops += flops;
    }
}

void prof_alg(int kind, double bytes, double flops) {
    if (!prof_on(kind)) return;
    g_kind[kind].alg_bytes += bytes;
    g_kind[kind].alg_flops += flops;
}
}  // namespace xt

using namespace xt;

extern "C" {

int xt_tune_set(int knob, int value) {
    XT_ARG(knob >= 0 && knob < XT_TUNE_NKNOBS, "tuning knob");
    static const int allowed[XT_TUNE_NKNOBS][5] = {
        {0, 2, 4, 8, 8},       // XT_TUNE_SPMM_UNR
        {0, 1, 2, 4, 4},       // XT_TUNE_SPMM_WPB
        {0, 4, 8, 4, 4},       // XT_TUNE_SPMM_R1SPLIT
        {0, 64, 128, 64, 64},  // XT_TUNE_BAND_TILE
        {0, 1, 2, 4, 4},       // XT_TUNE_GMRES_CHECK
        {0, 1, 2, 1, 1},       // XT_TUNE_SPMM_LDS
        {0, 1, 2, 3, 3},       // XT_TUNE_BAND_SMALLC
        {0, 1, 2, 4, 8},       // XT_TUNE_IR_FP32_CYCLE
        {0, 2, 8, 32, 128},    // XT_TUNE_IR_MIN_GAIN
        {0, 1, 4, 16, 64},     // XT_TUNE_IR_SKIP32
        {0, 1, 2, 3, 4},       // XT_TUNE_BAND_CHAIN
        {0, 1, 2, 3, 5},       // XT_TUNE_BAND_LEVELS
        {0, 1, 2, 2, 2},       // XT_TUNE_GMRES_REORTH
        {0, 1, 2, 2, 2},       // XT_TUNE_GMRES_COMPACT
        {0, 1, 2, 2, 2},       // XT_TUNE_BAND_CHAIN_FEW
        {0, 1, 2, 2, 2},       // XT_TUNE_BAND_PAIR
    };
    bool ok = false;
    for (int v : allowed[knob]) ok = ok || v == value;
    XT_ARG(ok, "value not allowed for this knob");
    if (g_knob[knob] != value) g_knob_epoch.fetch_add(1, std::memory_order_relaxed);
    g_knob[knob] = value;
    return XT_OK;
}

int xt_tune_get(int knob) { return (knob >= 0 && knob < XT_TUNE_NKNOBS) ? g_knob[knob] : -1; }

int xt_graphs_enable(int on) {
    g_graphs = on != 0;
    return XT_OK;
}

int xt_prof_enable(unsigned mask) {
    g_mask = mask;
    for (auto &k : g_kind) {
        k.used = 0;
        k.bytes = k.flops = k.alg_bytes = k.alg_flops = 0.0;
        if (!mask) {   // disabling releases the event pool
            for (hipEvent_t e : k.ev) (void)hipEventDestroy(e);
            k.ev.clear();
        }
    }
    return XT_OK;
}

int xt_prof_read(int kind, double *ms, long long *launches, double *bytes, double *flops) {
    XT_ARG(kind >= 0 && kind < XT_PROF_NKINDS, "profile kind");
    Kind &k = g_kind[kind];
    const size_t pairs = k.used / 2;
    double total = 0.0;
    if (pairs) XT_HIP(hipEventSynchronize(k.ev[2 * pairs - 1]));
    for (size_t i = 0; i < pairs; ++i) {
        float t = 0.0f;
        XT_HIP(hipEventElapsedTime(&t, k.ev[2 * i], k.ev[2 * i + 1]));
        total += t;
    }
    if (ms) *ms = total;
    if (launches) *launches = (long long)pairs;
    if (bytes) *bytes = k.bytes;
    if (flops) *flops = k.flops;
    return XT_OK;
}

int xt_prof_read_alg(int kind, double *bytes, double *flops) {
    XT_ARG(kind >= 0 && kind < XT_PROF_NKINDS, "profile kind");
    if (bytes) *bytes = g_kind[kind].alg_bytes;
    if (flops) *flops = g_kind[kind].alg_flops;
    return XT_OK;
}

}  // extern "C"
