// Recording pipeline of the fused step (engine.h, RecordPipeline).  The reference writes its
// recordings at each record step on the stepping thread (spike_recorder.cpp:40-60,
// voltage_recorder.cpp:54-77), and the simulation waits for them; here the copies of the rings go
// to pinned host buffers on the state's stream, so the next interval's steps follow them
// directly, and a writer thread formats and writes the files.
#include <chrono>

#include "../bn_errors.h"
#include "../bn_step.h"
#include "engine.h"

namespace bn {

RecordPipeline::RecordPipeline(GpuSimulationState& s, std::vector<Monitor*> monitors)
    : s_(s), monitors_(std::move(monitors)) {
    const uint32_t T = s_.options().recording_interval_steps;
    spk_elems_ = s_.d_spike_bits_ring() ? (size_t)s_.spike_words() * T : 0;
    v_elems_ = s_.d_v_ring() ? (size_t)std::max<uint32_t>(s_.n_vrec(), 1) * T : 0;
    BN_HIP_CHECK(hipGetDevice(&device_), "recording pipeline device");
    for (Slot& sl : slot_) {
        if (spk_elems_)
            BN_HIP_CHECK(hipHostMalloc((void**)&sl.spk, spk_elems_ * sizeof(uint32_t), hipHostMallocDefault),
                         "pinned spike ring buffer");
        if (v_elems_)
            BN_HIP_CHECK(hipHostMalloc((void**)&sl.v, v_elems_ * sizeof(float), hipHostMallocDefault),
                         "pinned voltage ring buffer");
        BN_HIP_CHECK(hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming), "recording event");
    }
    thread_ = std::thread([this] { writer(); });
}

RecordPipeline::~RecordPipeline() {
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return jobs_.empty(); });
        stop_ = true;
    }
    cv_.notify_all();
    if (thread_.joinable()) thread_.join();
    for (Slot& sl : slot_) {
        if (sl.spk) (void)hipHostFree(sl.spk);
        if (sl.v) (void)hipHostFree(sl.v);
        if (sl.ready) (void)hipEventDestroy(sl.ready);
    }
}

void RecordPipeline::submit(uint32_t step) {
    const int k = next_;
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !slot_[k].busy; });   // both buffers still being written: wait
        slot_[k].busy = true;
    }
    Slot& sl = slot_[k];
    hipStream_t st = s_.stream();
    // The interval's rows, then the rings cleared for the next interval (stream order: the next
    // steps write them only after the copies).
    if (spk_elems_) {
        BN_HIP_CHECK(hipMemcpyAsync(sl.spk, s_.d_spike_bits_ring(), spk_elems_ * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, st),
                     "D2H spike ring");
        BN_CALL(launch_fill_u32(s_.d_spike_bits_ring(), 0u, spk_elems_, st));
    }
    if (v_elems_) {
        BN_HIP_CHECK(hipMemcpyAsync(sl.v, s_.d_v_ring(), v_elems_ * sizeof(float), hipMemcpyDeviceToHost, st),
                     "D2H voltage ring");
        BN_CALL(launch_fill_u32(reinterpret_cast<uint32_t*>(s_.d_v_ring()), 0u, v_elems_, st));
    }
    BN_HIP_CHECK(hipEventRecord(sl.ready, st), "recording event");
    s_.mark_rings_flushed(step);
    {
        std::lock_guard<std::mutex> lk(mu_);
        jobs_.emplace_back(k, step);
    }
    cv_.notify_all();
    next_ ^= 1;
}

void RecordPipeline::drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return jobs_.empty(); });
    if (error_) {
        std::exception_ptr e = error_;
        error_ = nullptr;
        std::rethrow_exception(e);
    }
}

void RecordPipeline::writer() {
    (void)hipSetDevice(device_);
    for (;;) {
        std::pair<int, uint32_t> job;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
            if (jobs_.empty()) return;
            job = jobs_.front();
        }
        const Slot& sl = slot_[job.first];
        std::exception_ptr err;
        std::chrono::duration<double> dt{0.0};
        try {
            BN_HIP_CHECK(hipEventSynchronize(sl.ready), "recording copies");   // the interval's steps and copies
            const auto t0 = std::chrono::steady_clock::now();
            for (Monitor* m : monitors_) m->record_from_rings(job.second, sl.spk, sl.v);
            dt = std::chrono::steady_clock::now() - t0;
        } catch (...) {
            err = std::current_exception();
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (err && !error_) error_ = err;
            writer_seconds_ += dt.count();
            ++intervals_;
            slot_[job.first].busy = false;
            jobs_.pop_front();
        }
        cv_.notify_all();
    }
}

}  // namespace bn
