#include "tune.h"

#include <algorithm>
#include <mutex>

#include "engine.h"

namespace avd {

namespace {

std::mutex g_mu;
bool g_enabled = true;
std::vector<TuneEntry> g_entries;

constexpr int kReps = 3;

}  // namespace

Tuner &Tuner::instance()
{
    static Tuner t;
    return t;
}

void Tuner::set_enabled(bool on)
{
    std::lock_guard<std::mutex> lock(g_mu);
    g_enabled = on;
}

bool Tuner::enabled() const
{
    std::lock_guard<std::mutex> lock(g_mu);
    return g_enabled;
}

std::vector<TuneEntry> Tuner::entries() const
{
    std::lock_guard<std::mutex> lock(g_mu);
    return g_entries;
}

int Tuner::choose(const std::string &site, const std::string &key, int candidates, hipStream_t s,
                  const std::function<hipError_t(int)> &run, const std::function<hipError_t()> &clear,
                  const std::function<std::vector<uint8_t>()> &read)
{
    if (candidates <= 1 || !enabled()) return 0;
    int dev = 0;
    check_hip(hipGetDevice(&dev), "tuner device");
    {
        std::lock_guard<std::mutex> lock(g_mu);
        for (const TuneEntry &e : g_entries)
            if (e.device == dev && e.site == site && e.key == key) return e.choice;
    }
    TuneEntry entry;
    entry.device = dev;
    entry.site = site;
    entry.key = key;
    entry.candidates = candidates;
    hipEvent_t start, stop;
    check_hip(hipEventCreate(&start), "tuner event");
    check_hip(hipEventCreate(&stop), "tuner event");
    std::vector<uint8_t> baseline;
    try {
        for (int c = 0; c < candidates; ++c) {
            check_hip(clear(), "tuner clear");
            check_hip(run(c), "tuner run");
            check_hip(hipStreamSynchronize(s), "tuner run");
            const std::vector<uint8_t> out = read();
            if (c == 0) baseline = out;
            if (c > 0 && out != baseline) {  // exact fidelity: any differing bit rules it out
                entry.ms.push_back(-1.0);
                continue;
            }
            float best = 0.0f;
            for (int rep = 0; rep < kReps; ++rep) {
                check_hip(clear(), "tuner clear");
                check_hip(hipEventRecord(start, s), "tuner event");
                check_hip(run(c), "tuner run");
                check_hip(hipEventRecord(stop, s), "tuner event");
                check_hip(hipEventSynchronize(stop), "tuner event");
                float ms = 0.0f;
                check_hip(hipEventElapsedTime(&ms, start, stop), "tuner event");
                best = rep == 0 ? ms : std::min(best, ms);
            }
            entry.ms.push_back(best);
        }
    } catch (...) {
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        throw;
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    int choice = 0;
    for (int c = 1; c < candidates; ++c)
        if (entry.ms[c] >= 0.0 && entry.ms[c] < entry.ms[choice]) choice = c;
    entry.choice = choice;
    std::lock_guard<std::mutex> lock(g_mu);
    g_entries.push_back(entry);
    return choice;
}

}  // namespace avd
