// tuner.hpp -- launch-configuration choices timed on first use.
//
// A choice is kept per key: a kernel family and a size bucket (the shape
// rounded so that nearby sizes share one measurement), for the life of the
// process.  The first call of a key times every candidate with HIP events on
// the caller's stream and keeps the fastest; later calls of the key, and
// calls made while the stream is being captured into a graph or with
// CIR_NO_TUNE=1 in the environment, take the kept choice or the default.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <string>

namespace cir {

struct TuneKey {
    std::string family;
    std::array<int64_t, 5> shape{};  // size bucket and shape class, family-defined
    bool operator<(const TuneKey& o) const {
        return family != o.family ? family < o.family : shape < o.shape;
    }
};

// Power-of-two bucket of a size (sizes 1 .. 2^k map to 2^k).
inline int64_t size_bucket(int64_t x) {
    int64_t b = 1;
    while (b < x) b <<= 1;
    return b;
}

// The kept choice of key, or -1.
int tune_lookup(const TuneKey& key);
// Whether a choice may be timed now on stream s (not under CIR_NO_TUNE=1,
// not while s is capturing).
bool tune_allowed(hipStream_t s);
// Times run(i) for each candidate i < ncand on stream s (once to warm, once
// timed), keeps and returns the fastest.  run must be safe to repeat: every
// candidate leaves the same valid result.  On a HIP error the default 0 is
// kept and the error returned through *err.
int tune_time(const TuneKey& key, int ncand, hipStream_t s, const std::function<int(int)>& run, int* err);
// Every kept choice as text, one line per key: family, shape, choice and the
// candidates' times in ms.
std::string tune_report();

// Deterministic mode of the calling thread (a run with the deterministic key
// on, or CIR_DETERMINISTIC=1 in the environment): launch shapes are the fixed
// defaults (no timing, no kept choices), the eigensolver takes its two-stage
// form, the propagation one K segment, and the reductions a fixed order.
bool deterministic_mode();
struct DeterministicScope {
    bool prev;
    explicit DeterministicScope(bool on);
    ~DeterministicScope();
};
// Sizes the two-stage eigensolver takes (the deterministic key's range).
bool deterministic_eigh_size(int n);
int deterministic_eigh_max();

}  // namespace cir
