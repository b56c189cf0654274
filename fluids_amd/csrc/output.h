// Output boundary: the particle buffers leave the device as a frame file
// (in place of the GL vertex buffer the reference renders from,
// opencl/clallocator.h:165-183, 300-374) and frame times as an FPS log (the
// -o option, main.cpp:19, 34-35; gui/glwidget.cpp:48-73;
// gui/mainwindow.cpp:99-101).
//
// Frame file: "SPHF", u32 version (1), u32 fields (bit 0 positions, bit 1
// velocities), u32 0; then per frame: u64 frame index, f64 simulated time,
// u32 n, u32 fields, n float4 positions, n float4 velocities (as present).
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "solver.h"

namespace sph {

class FrameWriter {
public:
    FrameWriter(SimulationBase& sim, const std::string& path, int fields);
    ~FrameWriter();
    FrameWriter(const FrameWriter&) = delete;
    FrameWriter& operator=(const FrameWriter&) = delete;
    // queues the solver's current stored-order state: copies to pinned host
    // memory on the solver's stream, written by a background thread
    void write();
    // writes every queued frame and closes the file; returns frames written
    long long close();
    long long frames() const { return written_; }

private:
    struct Slot {
        float4* host = nullptr;
        size_t cap = 0;  // float4 elements
        hipEvent_t done = nullptr;
        unsigned long long frame = 0;
        double time = 0.0;
        int n = 0;
        bool busy = false;
    };
    void run();
    SimulationBase& sim_;
    std::FILE* f_ = nullptr;
    int fields_ = 1;
    Slot slots_[3];
    std::deque<int> queue_;
    std::mutex m_;
    std::condition_variable cv_;
    std::thread worker_;
    bool stop_ = false, failed_ = false;
    long long written_ = 0;
    std::string path_;
};

// FPS log: frame times measured between calls (after waiting for the
// solver's stream), one line "fps" per 6 frames as the reference's widget
// averages them; times are kept in microseconds instead of whole
// milliseconds so that sub-millisecond frames give a finite rate.
class PerfLog {
public:
    PerfLog(SimulationBase& sim, const std::string& path);
    ~PerfLog();
    void frame();
    long long lines() const { return lines_; }

private:
    SimulationBase& sim_;
    std::FILE* f_ = nullptr;
    bool started_ = false;
    double t_last_ = 0.0, total_ms_ = 0.0;
    int count_ = 0;
    long long lines_ = 0;
};

}  // namespace sph
