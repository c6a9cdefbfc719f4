#include "server.hpp"

namespace vtx {

Server::Server(Engine& e, const vtx_server_config& cfg) : e_(e), cfg_(cfg) {
    if (cfg_.hbm_budget_bytes) {
        limit_ = (size_t)cfg_.hbm_budget_bytes;
    } else {
        size_t free_b = 0, total_b = 0;
        VTX_HIP(hipMemGetInfo(&free_b, &total_b));
        limit_ = free_b > cfg_.hbm_reserve_bytes ? free_b - (size_t)cfg_.hbm_reserve_bytes : 0;
    }
    e_.set_use_graph(cfg_.use_graph != 0);
    e_.reserve(cfg_.reserve_particles);
    worker_ = std::thread([this] { worker_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    if (in_flight_) (void)hipStreamSynchronize(e_.stream());
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
}

Job& Server::get(int32_t id) {
    auto it = jobs_.find(id);
    VTX_REQUIRE(it != jobs_.end(), VTX_EINVAL, "no such job");
    return it->second;
}

const Job& Server::job(int32_t id) {
    Job& j = get(id);
    settle(j);
    return j;
}

// Submitted jobs are packed here, in submission order, while the caller's thread registers
// further jobs or runs frames. The engine is not otherwise touched from this thread: a queued
// job's segment is inactive, so frames, commits and table rebuilds skip it, and its packing
// reads and writes that segment object alone.
void Server::worker_loop() {
    for (;;) {
        Job* j = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu_);
            work_cv_.wait(lk, [this] { return stop_ || !to_pack_.empty(); });
            if (to_pack_.empty()) return;        // stopping, nothing left
            j = to_pack_.front();
            to_pack_.pop_front();
        }
        size_t bytes = 0;
        uint32_t n = 0;
        uint64_t work = 0;
        bool reject = false;
        try {
            e_.prepare_segment(*j->sp);
            bytes = Engine::bytes_estimate(*j->sp);
            n = j->sp->n();
            work = (uint64_t)n * (uint64_t)j->sp->prm.num_substeps;
            reject = bytes > limit_ || (cfg_.max_live_particles && n > cfg_.max_live_particles);
        } catch (const std::exception&) {
            reject = true;                       // e.g. beyond a capacity limit of the engine
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            j->bytes = bytes;
            j->n = n;
            j->work = work;
            j->reject = reject;
            j->prepared = true;
        }
        done_cv_.notify_all();
    }
}

void Server::settle(Job& j) {
    if (j.status != VTX_JOB_QUEUED || j.sp == nullptr) return;
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&j] { return j.prepared; });
    }
    j.sp = nullptr;                              // packed: the job no longer needs the worker
    if (j.reject) {                              // could never be admitted
        j.status = VTX_JOB_REJECTED;
        queue_.erase(std::find(queue_.begin(), queue_.end(), j.id));
        e_.remove_segment(j.seg);
        j.seg = -1;
    }
}

// A job is a pending (inactive) segment of the server's solver; the caller registers its
// cloths and colliders through the normal solver calls on that segment, then submits it.
int32_t Server::new_job(const vtx_params& p, int32_t frames, int32_t* seg) {
    VTX_REQUIRE(frames >= 1, VTX_EINVAL, "a job needs at least one frame");
    Job j;
    j.id = next_id_++;
    j.seg = e_.add_segment(p, /*active=*/false);
    j.frames = frames;
    jobs_[j.id] = j;
    if (seg) *seg = j.seg;
    return j.id;
}

// The job is queued at once; its host packing runs on the worker thread, and the first look
// at the job (admission, a status query) waits for that packing to finish.
void Server::submit(int32_t id) {
    Job& j = get(id);
    VTX_REQUIRE(j.status == VTX_JOB_PENDING, VTX_ESTATE, "job was already submitted");
    j.sp = &e_.seg(j.seg);
    j.status = VTX_JOB_QUEUED;
    queue_.push_back(id);
    {
        std::lock_guard<std::mutex> lk(mu_);
        j.prepared = false;
        to_pack_.push_back(&j);
    }
    work_cv_.notify_one();
}

// Step-time admission: the frame time of the live set is predicted as (particle-substeps) x
// (measured ms per particle-substep). Until a step has been measured, one job is admitted per
// step. A job always fits an empty server (so one that alone exceeds the budget still runs).
// The step-time budget bounds the predicted frame of every live job: live particle-substeps x
// the measured frame time per particle-substep (admission itself is timed apart, admit_ms).
bool Server::fits(const Job& j, uint32_t admitted_now) const {
    if (resident_ + j.bytes > limit_) return false;
    if (cfg_.max_live_particles && live_particles_ + j.n > cfg_.max_live_particles) return false;
    if (cfg_.max_live_jobs && live_.size() + staged_.size() >= cfg_.max_live_jobs) return false;
    if (cfg_.step_budget_ms > 0.f && !(live_.empty() && staged_.empty())) {
        if (ms_per_work_ == 0.0) return admitted_now == 0;
        if ((double)(live_work_ + j.work) * ms_per_work_ > (double)cfg_.step_budget_ms) return false;
    }
    return true;
}

// Admission first come, first served: the queued jobs that fit leave the queue. An empty server
// activates them for this step's frame; a running server stages them (uploads on the engine's
// copy stream while this step's frame runs) and they join the live set at the next step.
uint32_t Server::admit(bool stage, uint32_t before) {
    uint32_t n = 0;
    while (!queue_.empty()) {
        Job& j = get(queue_.front());
        settle(j);
        if (j.status == VTX_JOB_REJECTED) continue;   // settle took it out of the queue
        if (!fits(j, before + n)) break;
        if (stage) {
            // a frame is in flight: an upload that fails (e.g. device memory) leaves the job
            // queued and ends this step's admission, so the step's own accounting still runs
            try {
                e_.stage(j.seg);
            } catch (const Error&) {
                break;
            }
            staged_.push_back(j.id);
        } else {
            e_.set_active(j.seg, true);
            live_.push_back(j.id);
        }
        queue_.pop_front();
        j.status = VTX_JOB_RUNNING;
        resident_ += j.bytes;
        live_particles_ += j.n;
        live_work_ += j.work;
        n++;
    }
    return n;
}

// Wait for the frame in flight, count it for its jobs and the cost model, then retire the jobs
// that reached their horizon (read back and removed) or diverged; the figures go to carry_.
void Server::complete() {
    using clk = std::chrono::steady_clock;
    if (!in_flight_) return;
    e_.synchronize();
    in_flight_ = false;
    float frame_ms = 0.f;
    VTX_HIP(hipEventElapsedTime(&frame_ms, ev0_, ev1_));
    carry_.frame_ms += frame_ms;
    uint64_t work = 0;
    for (int32_t id : flight_) {
        Job& j = get(id);
        j.done++;
        work += j.work;
    }
    flight_.clear();
    carry_.particle_substeps += work;
    if (work) {
        const double r = frame_ms / (double)work;
        ms_per_work_ = ms_per_work_ == 0.0 ? r : 0.8 * ms_per_work_ + 0.2 * r;
    }
    const auto r0 = clk::now();
    std::vector<int32_t> bad;
    if (cfg_.check_every > 0 && steps_ % (uint64_t)cfg_.check_every == 0 && !live_.empty())
        bad = e_.nonfinite_segments();
    std::vector<int32_t> keep;
    for (int32_t id : live_) {
        Job& j = get(id);
        bool failed = false;
        for (int32_t s : bad) failed |= (s == j.seg);
        if (failed || j.done >= j.frames) {
            if (!failed) {
                j.result.resize((size_t)3 * j.n);
                e_.read_particles(j.seg, 0, j.result.data());
            }
            j.status = failed ? VTX_JOB_FAILED : VTX_JOB_DONE;
            e_.remove_segment(j.seg);
            j.seg = -1;
            resident_ -= j.bytes;
            live_particles_ -= j.n;
            live_work_ -= j.work;
            (failed ? carry_.failed : carry_.completed)++;
        } else {
            keep.push_back(id);
        }
    }
    live_.swap(keep);
    carry_.retire_ms += std::chrono::duration<double, std::milli>(clk::now() - r0).count();
}

void Server::drain() { complete(); }

// One step: (pipelined: complete the previous step's frame first) staged jobs join, queued jobs
// are admitted, the frame of every live job is launched, jobs admitted while it runs are
// uploaded on the copy stream; without pipelining the step then waits for its frame and
// retires, with pipelining it returns and the next step does.
void Server::step(float frame_dt, vtx_server_stats* out) {
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point a) { return std::chrono::duration<double, std::milli>(clk::now() - a).count(); };
    const auto t0 = clk::now();
    complete();                                  // pipelined: the previous step's frame
    vtx_server_stats st = carry_;
    carry_ = vtx_server_stats{};
    for (int32_t id : staged_) {                 // staged last step: uploaded, now live
        e_.set_active(get(id).seg, true);
        live_.push_back(id);
    }
    staged_.clear();
    if (live_.empty()) st.admitted += admit(false, 0);
    bool launched = false;
    if (!live_.empty()) {
        const auto c0 = clk::now();
        e_.commit();                             // uploads of jobs activated now, launch tables
        st.admit_ms += ms_since(c0);
        if (!ev0_) {
            VTX_HIP(hipEventCreate(&ev0_));
            VTX_HIP(hipEventCreate(&ev1_));
        }
        const hipStream_t es = e_.stream();
        const auto launch = clk::now();          // pipelined step_ms: launch to launch
        if (cfg_.pipelined && last_launch_ != clk::time_point{})
            st.step_ms = std::chrono::duration<double, std::milli>(launch - last_launch_).count();
        last_launch_ = launch;
        VTX_HIP(hipEventRecord(ev0_, es));
        e_.simulate(frame_dt);
        e_.join();                               // the frame's end is the end of every stream group
        VTX_HIP(hipEventRecord(ev1_, es));
        in_flight_ = true;
        flight_ = live_;
        launched = true;
        const auto s0 = clk::now();
        st.admitted += admit(true, st.admitted);   // overlaps the frame
        st.admit_ms += ms_since(s0);
    }
    steps_++;
    if (!cfg_.pipelined) {
        complete();
        st.frame_ms += carry_.frame_ms;
        st.particle_substeps += carry_.particle_substeps;
        st.completed += carry_.completed;
        st.failed += carry_.failed;
        st.retire_ms += carry_.retire_ms;
        carry_ = vtx_server_stats{};
    }
    st.live_jobs = (uint32_t)live_.size();
    st.queued_jobs = (uint32_t)queue_.size();
    st.live_particles = live_particles_;
    if (cfg_.pipelined && !launched && last_launch_ != clk::time_point{}) {
        // the last frame's period ends here; the next launch starts a new run of periods
        st.step_ms = std::chrono::duration<double, std::milli>(clk::now() - last_launch_).count();
        last_launch_ = clk::time_point{};
    } else if (!(cfg_.pipelined && launched && st.step_ms > 0.0)) {   // first launch, or synchronous
        st.step_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    if (out) *out = st;
}

// A running job leaves the device (its state parked on the host, its device ranges freed) and
// queues again behind the jobs already waiting; when re-admitted it continues from the frame it
// reached, with the same result as an uninterrupted run.
void Server::preempt(int32_t id) {
    complete();
    Job& j = get(id);
    VTX_REQUIRE(j.status == VTX_JOB_RUNNING, VTX_ESTATE, "only a running job can be preempted");
    e_.park(j.seg);
    auto it = std::find(live_.begin(), live_.end(), id);
    if (it != live_.end()) live_.erase(it);
    else staged_.erase(std::find(staged_.begin(), staged_.end(), id));   // running jobs are in one of the two
    resident_ -= j.bytes;
    live_particles_ -= j.n;
    live_work_ -= j.work;
    j.status = VTX_JOB_QUEUED;
    queue_.push_back(id);
}

void Server::release(int32_t id) {
    Job& j = get(id);
    settle(j);
    VTX_REQUIRE(j.status != VTX_JOB_RUNNING && j.status != VTX_JOB_QUEUED, VTX_ESTATE,
                "job is still queued or running");
    if (j.seg >= 0) e_.remove_segment(j.seg);
    jobs_.erase(id);
}

}  // namespace vtx
