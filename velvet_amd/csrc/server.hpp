// Continuous-batching server over the engine: jobs (independent simulations with a frame
// horizon) queue up, are admitted while device memory and the live-particle budget allow, all
// admitted jobs advance one frame per step in the same launches, and finished jobs leave.
#pragma once
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace vtx {

struct Job {
    int32_t id = -1, seg = -1;
    int32_t frames = 0, done = 0;
    int32_t status = VTX_JOB_PENDING;
    size_t bytes = 0;
    uint32_t n = 0;
    uint64_t work = 0;           // particle-substeps per frame
    std::vector<float> result;   // final positions (xyz) once done
    // host packing of a submitted job runs on the server's worker thread; these are written
    // there under the server's mutex
    Segment* sp = nullptr;       // the job's segment (stable while the job is queued)
    bool prepared = false;       // packing finished (bytes, n, work valid)
    bool reject = false;         // could never be admitted, or packing failed
};

class Server {
public:
    Server(Engine& e, const vtx_server_config& cfg);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    int32_t new_job(const vtx_params& p, int32_t frames, int32_t* seg);
    void submit(int32_t job);
    void step(float frame_dt, vtx_server_stats* out);
    const Job& job(int32_t id);  // waits for the job's packing if it is still running
    void release(int32_t id);
    void preempt(int32_t id);
    void drain();                // pipelined: complete the frame in flight

private:
    Job& get(int32_t id);
    bool fits(const Job& j, uint32_t admitted_now) const;
    uint32_t admit(bool stage, uint32_t before);   // before: jobs admitted earlier in this step
    void complete();             // wait for the frame in flight, count it, retire finished jobs
    void worker_loop();          // packs submitted jobs (Engine::prepare_segment) off the caller's thread
    void settle(Job& j);         // wait for a submitted job's packing; reject it if it can never run

    Engine& e_;
    vtx_server_config cfg_;
    std::map<int32_t, Job> jobs_;
    std::deque<int32_t> queue_;
    std::vector<int32_t> live_;
    std::vector<int32_t> staged_;  // admitted while a frame ran: uploaded, live from the next step
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;   // GPU time of the frame
    int32_t next_id_ = 0;
    size_t limit_ = 0;           // device bytes jobs may occupy
    size_t resident_ = 0;        // estimated bytes of the live jobs
    uint64_t live_particles_ = 0;
    uint64_t live_work_ = 0;     // particle-substeps per frame of the live jobs
    double ms_per_work_ = 0.0;   // measured frame time per particle-substep (EMA); 0 = not yet
    uint64_t steps_ = 0;
    bool in_flight_ = false;     // a launched frame not yet completed
    std::vector<int32_t> flight_;   // the jobs of that frame
    vtx_server_stats carry_{};   // completed-frame figures for the next step's statistics
    std::chrono::steady_clock::time_point last_launch_{};
    std::mutex mu_;              // guards the packing queue and the jobs' packing results
    std::condition_variable work_cv_, done_cv_;
    std::deque<Job*> to_pack_;
    bool stop_ = false;
    std::thread worker_;
};

}  // namespace vtx
