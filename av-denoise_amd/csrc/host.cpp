// HostDenoiser (av-denoise/src/host/mod.rs:111-345): frames in and out as host wire bytes, one
// buffer per plane.  Each push copies the planes into device planes of whole words and pushes them
// through the engine; every frame the push readies is emitted into one of two output plane sets and
// read back asynchronously into pinned host memory, so at most two readbacks are in flight.  recv
// waits for the oldest and copies it out; flush drains the rest and the temporal tail.
#include "host.h"

#include <cstring>

namespace avd {

namespace {

uint64_t wire_bytes(avd_sample_format f, uint64_t pixels)
{
    return f.kind == AVD_FORMAT_U8 ? pixels : pixels * 2;  // U16 planes: little-endian 16-bit samples
}

}  // namespace

HostDenoiser::HostDenoiser(const avd_engine_spec &spec)
{
    const avd_geometry &g = spec.geometry;
    validate_format(g.input);
    validate_format(g.output);
    if (g.input.kind == AVD_FORMAT_F32 || g.output.kind == AVD_FORMAT_F32)
        throw Failure(AVD_ERR_INVALID_GEOMETRY, "the wire API carries 8- and 16-bit samples");
    dims_.push_back({g.width, g.height});
    const uint32_t cw = (g.width + 1) / 2, ch = (g.height + 1) / 2;
    if (g.chroma == AVD_CHROMA_420) dims_.insert(dims_.end(), 2, {cw, ch});
    else if (g.chroma == AVD_CHROMA_422) dims_.insert(dims_.end(), 2, {cw, g.height});
    else if (g.chroma == AVD_CHROMA_444) dims_.insert(dims_.end(), 2, {g.width, g.height});
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "create host denoiser stream");
    try {
        avd_engine_spec own = spec;
        own.stream = stream_;
        engine_.reset(new Engine(own));
        for (const auto &d : dims_) {
            const uint64_t px = (uint64_t)d.first * d.second;
            lengths_.push_back(wire_bytes(g.input, px));
            out_lengths_.push_back(wire_bytes(g.output, px));
            device_bytes_.push_back(plane_bytes(g.input, px));
            out_device_bytes_.push_back(plane_bytes(g.output, px));
        }
        input_ = std::vector<DeviceBuffer>(dims_.size());
        for (size_t i = 0; i < dims_.size(); ++i) input_[i].allocate(device_bytes_[i]);
        for (size_t s = 0; s < kOutputSets; ++s) {
            output_[s] = std::vector<DeviceBuffer>(dims_.size());
            for (size_t i = 0; i < dims_.size(); ++i) {
                output_[s][i].allocate(out_device_bytes_[i]);
                void *h = nullptr;
                check_hip(hipHostMalloc(&h, out_device_bytes_[i], hipHostMallocDefault), "pinned readback buffer");
                host_[s].push_back(h);
            }
        }
    } catch (...) {
        release();
        throw;
    }
}

HostDenoiser::~HostDenoiser() { release(); }

void HostDenoiser::release()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto &p : pending_) (void)hipEventDestroy(p.ready);
    pending_.clear();
    for (hipEvent_t e : free_events_) (void)hipEventDestroy(e);
    free_events_.clear();
    for (size_t s = 0; s < kOutputSets; ++s) {
        for (void *h : host_[s]) (void)hipHostFree(h);
        host_[s].clear();
    }
    engine_.reset();
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
}

void HostDenoiser::check_planes(const void *const *planes, const uint64_t *lengths, uint32_t count) const
{
    if (count != dims_.size())
        throw Failure(AVD_ERR_PLANE_MISMATCH,
                      "expected " + std::to_string(dims_.size()) + " planes, got " + std::to_string(count));
    for (uint32_t i = 0; i < count; ++i) {
        if (!planes[i]) throw Failure(AVD_ERR_NULL_ARGUMENT, "plane " + std::to_string(i) + " is null");
        if (lengths[i] != lengths_[i])
            throw Failure(AVD_ERR_PLANE_MISMATCH, "plane " + std::to_string(i) + " holds " +
                                                      std::to_string(lengths[i]) + " bytes, expected " +
                                                      std::to_string(lengths_[i]));
    }
}

std::vector<avd_device_plane> HostDenoiser::upload(const void *const *planes)
{
    std::vector<avd_device_plane> out;
    for (size_t i = 0; i < dims_.size(); ++i) {
        // From pageable memory the copy has left the caller's buffer when it returns.
        check_hip(hipMemcpyAsync(input_[i].ptr, planes[i], lengths_[i], hipMemcpyHostToDevice, stream_),
                  "upload plane");
        out.push_back(avd_device_plane{input_[i].ptr, device_bytes_[i], dims_[i].first, dims_[i].second});
    }
    return out;
}

HostDenoiser::Pending HostDenoiser::emit_into_set(size_t set)
{
    std::vector<avd_device_plane> planes;
    for (size_t i = 0; i < dims_.size(); ++i)
        planes.push_back(avd_device_plane{output_[set][i].ptr, out_device_bytes_[i], dims_[i].first, dims_[i].second});
    engine_->emit(planes.data(), (uint32_t)planes.size());
    for (size_t i = 0; i < dims_.size(); ++i)
        check_hip(hipMemcpyAsync(host_[set][i], output_[set][i].ptr, out_lengths_[i], hipMemcpyDeviceToHost, stream_),
                  "read back plane");
    hipEvent_t ev;
    if (!free_events_.empty()) {
        ev = free_events_.back();
        free_events_.pop_back();
    } else {
        check_hip(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "create readback event");
    }
    check_hip(hipEventRecord(ev, stream_), "record readback event");
    return Pending{set, ev};
}

void HostDenoiser::wait(const Pending &p)
{
    check_hip(hipEventSynchronize(p.ready), "wait for readback");
    free_events_.push_back(p.ready);
}

void HostDenoiser::guarded(const std::function<void()> &call)
{
    try {
        call();
    } catch (const Failure &f) {
        // A context frame after a push changes no state (DenoiserError::poisons_stream).
        if (f.status != AVD_ERR_CONTEXT_AFTER_PUSH) poisoned_ = true;
        throw;
    } catch (...) {
        poisoned_ = true;
        throw;
    }
}

void HostDenoiser::check_usable() const
{
    if (poisoned_) throw Failure(AVD_ERR_POISONED, "denoiser failed earlier, reset the stream before using it again");
}

void HostDenoiser::push(const void *const *planes, const uint64_t *lengths, uint32_t count)
{
    check_usable();
    if (pending_.size() >= kMaxPending)
        throw Failure(AVD_ERR_QUEUE_FULL, "denoiser queue is full, collect the pending frame before pushing more");
    check_planes(planes, lengths, count);
    guarded([&] {
        const std::vector<avd_device_plane> dev = upload(planes);
        const uint32_t ready = engine_->push(dev.data(), (uint32_t)dev.size());
        // Emitting past the free output sets would overwrite a set whose readback has not landed.
        if (pending_.size() + ready > kOutputSets)
            throw Failure(AVD_ERR_GPU, "push readied " + std::to_string(ready) + " frames with " +
                                           std::to_string(pending_.size()) + " in flight, past the output sets");
        for (uint32_t k = 0; k < ready; ++k) {
            const size_t set = next_set_;
            next_set_ = (next_set_ + 1) % kOutputSets;
            pending_.push_back(emit_into_set(set));
        }
    });
}

void HostDenoiser::push_priming(const void *const *planes, const uint64_t *lengths, uint32_t count)
{
    check_usable();
    check_planes(planes, lengths, count);
    guarded([&] {
        const std::vector<avd_device_plane> dev = upload(planes);
        engine_->push_context(dev.data(), (uint32_t)dev.size());
    });
}

bool HostDenoiser::recv(void *const *planes, const uint64_t *capacities, uint32_t count)
{
    check_usable();
    if (pending_.empty()) return false;
    if (count != dims_.size()) throw Failure(AVD_ERR_PLANE_MISMATCH, "wrong number of output planes");
    for (uint32_t i = 0; i < count; ++i) {
        if (!planes[i]) throw Failure(AVD_ERR_NULL_ARGUMENT, "output plane is null");
        if (capacities[i] < out_lengths_[i])
            throw Failure(AVD_ERR_PLANE_MISMATCH, "output plane " + std::to_string(i) + " is too small");
    }
    guarded([&] {
        const Pending p = pending_.front();
        pending_.pop_front();
        wait(p);
        for (uint32_t i = 0; i < count; ++i) std::memcpy(planes[i], host_[p.set][i], out_lengths_[i]);
    });
    return true;
}

void HostDenoiser::flush(const std::function<void(const void *const *, const uint64_t *, uint32_t)> &sink)
{
    check_usable();
    guarded([&] {
        auto hand = [&](size_t set) {
            std::vector<const void *> ptrs(host_[set].begin(), host_[set].end());
            sink(ptrs.data(), out_lengths_.data(), (uint32_t)ptrs.size());
        };
        while (!pending_.empty()) {
            const Pending p = pending_.front();
            pending_.pop_front();
            wait(p);
            hand(p.set);
        }
        const uint32_t tail = engine_->finish();
        for (uint32_t k = 0; k < tail; ++k) {
            const Pending p = emit_into_set(0);
            wait(p);
            hand(p.set);
        }
        next_set_ = 0;
    });
}

void HostDenoiser::reset_stream()
{
    check_hip(hipStreamSynchronize(stream_), "reset stream");
    for (auto &p : pending_) free_events_.push_back(p.ready);
    pending_.clear();
    engine_->reset();
    next_set_ = 0;
    poisoned_ = false;
}

}  // namespace avd
