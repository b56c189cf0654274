// Frame file writer and FPS log (output.h).
#include "output.h"

#include <chrono>
#include <cstring>

namespace sph {

namespace {
double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

FrameWriter::FrameWriter(SimulationBase& sim, const std::string& path, int fields)
    : sim_(sim), fields_(fields), path_(path) {
    if (fields_ <= 0 || fields_ > 3) throw Error(SPH_ERR_INVALID, "frames: fields must be 1, 2 or 3");
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) throw Error(SPH_ERR_INVALID, "frames: cannot open " + path);
    const uint32_t head[4] = {0x46485053u /* "SPHF" */, 1u, (uint32_t)fields_, 0u};
    if (std::fwrite(head, sizeof head, 1, f_) != 1) {
        std::fclose(f_);
        throw Error(SPH_ERR_INVALID, "frames: cannot write " + path);
    }
    for (Slot& s : slots_) hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate");
    worker_ = std::thread([this] { run(); });
}

FrameWriter::~FrameWriter() {
    try {
        close();
    } catch (...) {
    }
    for (Slot& s : slots_) {
        if (s.host) (void)hipHostFree(s.host);
        if (s.done) (void)hipEventDestroy(s.done);
    }
}

void FrameWriter::write() {
    if (!f_) throw Error(SPH_ERR_STATE, "frames: file is closed");
    const int n = sim_.particle_count();
    const int per = (fields_ & 1) + ((fields_ >> 1) & 1);
    int k = -1;
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] {
            for (int i = 0; i < 3; ++i)
                if (!slots_[i].busy) return true;
            return failed_;
        });
        if (failed_) throw Error(SPH_ERR_INVALID, "frames: writing " + path_ + " failed");
        for (int i = 0; i < 3; ++i)
            if (!slots_[i].busy) {
                k = i;
                break;
            }
        slots_[k].busy = true;
    }
    Slot& s = slots_[k];
    try {
        const size_t need = (size_t)n * per;
        if (s.cap < need) {
            if (s.host) hip_check(hipHostFree(s.host), "hipHostFree");
            s.host = nullptr;
            s.cap = 0;
            hip_check(hipHostMalloc((void**)&s.host, std::max<size_t>(need, 1) * sizeof(float4), hipHostMallocDefault),
                      "hipHostMalloc");
            s.cap = need;
        }
        s.n = n;
        s.frame = (unsigned long long)sim_.frames_done;
        s.time = sim_.sim_time;
        const float4* pos = sim_.d_pos.get();
        const float4* vel = sim_.d_vel.get();
        hipStream_t st = sim_.stream();
        size_t off = 0;
        if (n > 0 && (fields_ & 1)) {
            hip_check(hipMemcpyAsync(s.host, pos, n * sizeof(float4), hipMemcpyDeviceToHost, st), "frames copy");
            off = n;
        }
        if (n > 0 && (fields_ & 2))
            hip_check(hipMemcpyAsync(s.host + off, vel, n * sizeof(float4), hipMemcpyDeviceToHost, st), "frames copy");
        hip_check(hipEventRecord(s.done, st), "hipEventRecord");
    } catch (...) {
        // the slot was never queued: give it back so that later writes do
        // not wait for it
        {
            std::lock_guard<std::mutex> lk(m_);
            s.busy = false;
        }
        cv_.notify_all();
        throw;
    }
    {
        std::lock_guard<std::mutex> lk(m_);
        queue_.push_back(k);
    }
    cv_.notify_all();
}

void FrameWriter::run() {
    for (;;) {
        int k;
        {
            std::unique_lock<std::mutex> lk(m_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) return;
            k = queue_.front();
            queue_.pop_front();
        }
        Slot& s = slots_[k];
        bool ok = hipEventSynchronize(s.done) == hipSuccess;
        const int per = (fields_ & 1) + ((fields_ >> 1) & 1);
        const uint32_t nf[2] = {(uint32_t)s.n, (uint32_t)fields_};
        ok = ok && std::fwrite(&s.frame, sizeof s.frame, 1, f_) == 1 && std::fwrite(&s.time, sizeof s.time, 1, f_) == 1 &&
             std::fwrite(nf, sizeof nf, 1, f_) == 1;
        if (ok && s.n > 0) ok = std::fwrite(s.host, sizeof(float4), (size_t)s.n * per, f_) == (size_t)s.n * per;
        {
            std::lock_guard<std::mutex> lk(m_);
            s.busy = false;
            if (ok) ++written_;
            else failed_ = true;
        }
        cv_.notify_all();
    }
}

long long FrameWriter::close() {
    if (!f_) return written_;
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    std::fclose(f_);
    f_ = nullptr;
    if (failed_) throw Error(SPH_ERR_INVALID, "frames: writing " + path_ + " failed");
    return written_;
}

PerfLog::PerfLog(SimulationBase& sim, const std::string& path) : sim_(sim) {
    f_ = std::fopen(path.c_str(), "w");
    if (!f_) throw Error(SPH_ERR_INVALID, "performance log: cannot open " + path);
}

PerfLog::~PerfLog() {
    if (f_) std::fclose(f_);
}

void PerfLog::frame() {
    sim_.synchronize();
    const double t = now_ms();
    if (!started_) {
        started_ = true;
        t_last_ = t;
        return;
    }
    total_ms_ += t - t_last_;
    t_last_ = t;
    ++count_;
    if (count_ > 5) {  // glwidget.cpp:65-72: every 6 frames
        const float fps = (float)(1000.0 / (total_ms_ / count_));
        std::fprintf(f_, "%g\n", fps);
        std::fflush(f_);
        total_ms_ = 0.0;
        count_ = 0;
        ++lines_;
    }
}

}  // namespace sph
