// tuner.cpp -- launch-configuration choices timed on first use (tuner.hpp).
#include "tuner.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace cir {

namespace {

struct Entry {
    int choice = 0;
    std::vector<float> ms;
};

std::mutex& tune_mu() {
    static std::mutex m;
    return m;
}
std::map<TuneKey, Entry>& tune_map() {
    static std::map<TuneKey, Entry> m;
    return m;
}

}  // namespace

namespace {
thread_local bool t_det = false;
}

bool deterministic_mode() {
    if (t_det) return true;
    const char* e = std::getenv("CIR_DETERMINISTIC");
    return e && std::atoi(e) == 1;
}
DeterministicScope::DeterministicScope(bool on) : prev(t_det) { t_det = prev || on; }
DeterministicScope::~DeterministicScope() { t_det = prev; }

int tune_lookup(const TuneKey& key) {
    if (deterministic_mode()) return -1;
    std::lock_guard<std::mutex> lk(tune_mu());
    auto it = tune_map().find(key);
    return it == tune_map().end() ? -1 : it->second.choice;
}

bool tune_allowed(hipStream_t s) {
    if (deterministic_mode()) return false;
    const char* e = std::getenv("CIR_NO_TUNE");
    if (e && std::atoi(e) == 1) return false;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
    return st == hipStreamCaptureStatusNone;
}

int tune_time(const TuneKey& key, int ncand, hipStream_t s, const std::function<int(int)>& run, int* err) {
    *err = CIR_OK;
    Entry ent;
    ent.ms.assign(ncand, 0.f);
    EventPair ev;
    hipError_t he = ev.create();
    if (he != hipSuccess) {
        *err = hip_fail(he, "hipEventCreate", __FILE__, __LINE__);
        return 0;
    }
    float best = 0.f;
    for (int i = 0; i < ncand; ++i) {
        int rc = run(i);  // warm (first use of the code object, caches)
        if (rc == CIR_OK && (he = hipEventRecord(ev.a, s)) != hipSuccess) rc = hip_fail(he, "hipEventRecord", __FILE__, __LINE__);
        if (rc == CIR_OK) rc = run(i);
        if (rc == CIR_OK && (he = hipEventRecord(ev.b, s)) != hipSuccess) rc = hip_fail(he, "hipEventRecord", __FILE__, __LINE__);
        if (rc == CIR_OK && (he = hipEventSynchronize(ev.b)) != hipSuccess)
            rc = hip_fail(he, "hipEventSynchronize", __FILE__, __LINE__);
        float ms = 0.f;
        if (rc == CIR_OK && (he = hipEventElapsedTime(&ms, ev.a, ev.b)) != hipSuccess)
            rc = hip_fail(he, "hipEventElapsedTime", __FILE__, __LINE__);
        if (rc != CIR_OK) {
            *err = rc;
            return 0;
        }
        ent.ms[i] = ms;
        if (i == 0 || ms < best) {
            best = ms;
            ent.choice = i;
        }
    }
    std::lock_guard<std::mutex> lk(tune_mu());
    auto ins = tune_map().emplace(key, ent);
    return ins.first->second.choice;  // a concurrent first call may have kept its own
}

std::string tune_report() {
    std::lock_guard<std::mutex> lk(tune_mu());
    std::string out;
    char buf[256];
    for (const auto& kv : tune_map()) {
        std::snprintf(buf, sizeof buf, "%s %lld %lld %lld %lld %lld -> %d (", kv.first.family.c_str(),
                      (long long)kv.first.shape[0], (long long)kv.first.shape[1], (long long)kv.first.shape[2],
                      (long long)kv.first.shape[3], (long long)kv.first.shape[4], kv.second.choice);
        out += buf;
        for (size_t i = 0; i < kv.second.ms.size(); ++i) {
            std::snprintf(buf, sizeof buf, "%s%.4f", i ? " " : "", double(kv.second.ms[i]));
            out += buf;
        }
        out += " ms)\n";
    }
    return out;
}

}  // namespace cir
