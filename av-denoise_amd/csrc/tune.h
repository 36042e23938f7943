// The launch-site tuner (the reference's KernelSelector over VariantSite candidates,
// tune/selector.rs:109, tune/site.rs:56-157): a site offers candidate launches of one stage (tile
// shapes, references per wave); the tuner runs each on the site's own buffers, drops a candidate
// whose outputs differ from the baseline's (candidate 0) by a single bit, times the rest with HIP
// events and keeps the fastest, once per device, site and shape key.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "common.h"

namespace avd {

struct TuneEntry {
    int device = 0;
    std::string site, key;
    int choice = 0, candidates = 0;
    std::vector<double> ms;  // best time of each candidate (negative: did not match the baseline)
};

class Tuner {
  public:
    static Tuner &instance();
    void set_enabled(bool on);
    bool enabled() const;
    // `run(c)` enqueues candidate c on `s` (writing the site's outputs), `clear()` resets the
    // outputs it accumulates into, `read()` returns them as bytes (after a synchronize).  Returns
    // the chosen candidate (0 when tuning is off or every other candidate mismatched).
    int choose(const std::string &site, const std::string &key, int candidates, hipStream_t s,
               const std::function<hipError_t(int)> &run, const std::function<hipError_t()> &clear,
               const std::function<std::vector<uint8_t>()> &read);
    std::vector<TuneEntry> entries() const;

  private:
    Tuner() = default;
};

}  // namespace avd
