// The host wire API: a denoiser over frames held as host bytes (the reference's HostDenoiser,
// av-denoise/src/host/mod.rs:27-345), driving an Engine on a stream of its own.
#pragma once
#include <deque>
#include <functional>
#include <memory>
#include <vector>

#include "engine.h"

namespace avd {

// Readbacks in flight at once (MAX_PENDING, host/mod.rs:27) and output plane sets rotated through
// (OUTPUT_SETS, host/mod.rs:30): each frame in flight reads its own set.
constexpr size_t kMaxPending = 2;
constexpr size_t kOutputSets = 2;
static_assert(kOutputSets >= kMaxPending, "a set is never rewritten before its readback lands");

class HostDenoiser {
  public:
    explicit HostDenoiser(const avd_engine_spec &spec);
    ~HostDenoiser();
    HostDenoiser(const HostDenoiser &) = delete;
    HostDenoiser &operator=(const HostDenoiser &) = delete;

    // One frame, one byte buffer per plane (Y, then U and V), each of its plane's wire length.
    void push(const void *const *planes, const uint64_t *lengths, uint32_t count);
    void push_priming(const void *const *planes, const uint64_t *lengths, uint32_t count);
    // Waits for the oldest frame in flight and copies it out; false when none is in flight.
    bool recv(void *const *planes, const uint64_t *capacities, uint32_t count);
    // Drains the frames in flight and the temporal tail through `sink`, then starts a new stream.
    void flush(const std::function<void(const void *const *, const uint64_t *, uint32_t)> &sink);
    void reset_stream();
    uint32_t pending() const { return (uint32_t)pending_.size(); }
    uint32_t planes() const { return (uint32_t)dims_.size(); }
    uint64_t input_length(uint32_t i) const { return lengths_.at(i); }
    uint64_t output_length(uint32_t i) const { return out_lengths_.at(i); }
    avd_window_span window_span() const { return engine_->window_span(); }

  private:
    struct Pending {
        size_t set;
        hipEvent_t ready;
    };
    void check_planes(const void *const *planes, const uint64_t *lengths, uint32_t count) const;
    std::vector<avd_device_plane> upload(const void *const *planes);
    Pending emit_into_set(size_t set);
    void wait(const Pending &p);
    // Runs `call`; a failure poisons the stream until reset_stream (a context frame after a push
    // changes no state and does not).
    void guarded(const std::function<void()> &call);
    void check_usable() const;
    void release();

    hipStream_t stream_ = nullptr;
    std::unique_ptr<Engine> engine_;
    std::vector<std::pair<uint32_t, uint32_t>> dims_;
    std::vector<uint64_t> lengths_, out_lengths_;            // wire bytes per input / output plane
    std::vector<uint64_t> device_bytes_, out_device_bytes_;  // the same, rounded up to whole words
    std::vector<DeviceBuffer> input_;     // one device plane per plane
    std::vector<DeviceBuffer> output_[kOutputSets];
    std::vector<void *> host_[kOutputSets];  // pinned readback buffers
    std::deque<Pending> pending_;
    std::vector<hipEvent_t> free_events_;
    size_t next_set_ = 0;
    bool poisoned_ = false;
};

}  // namespace avd
