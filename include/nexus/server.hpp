// Request loop for encrypted inference over the evaluator programs.
//
// A request is a list of evaluator calls (GELU, softmax, LayerNorm, bootstrap)
// on its own input ciphertexts.  Workers, one per lane of the context (a HIP
// stream with its own memory cache), take requests in arrival order as soon
// as they are free and the request's ciphertext footprint fits in the memory
// budget; requests on different lanes overlap on the device while sharing the
// device-resident keys.  Each request's result depends only on its inputs, so
// it is the same whether it runs alone or among others.
#pragma once

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bootstrapper.hpp"
#include "ckks.hpp"

namespace nexus {

enum class JobKind : int { Gelu = 0, Softmax = 1, LayerNorm = 2, Bootstrap = 3 };

struct Job {
    JobKind kind;
    int len = 0;  // row length of softmax / LayerNorm
    Ciphertext input;
};

class Server {
public:
    struct Options {
        int lanes = 2;               // concurrent requests (HIP streams)
        size_t memory_budget = 0;    // bytes of in-flight request footprint; 0 = no limit
        double footprint_factor = 6; // footprint = factor x input ciphertext bytes
    };
    struct Stats {
        size_t submitted = 0, completed = 0;
        double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
        size_t peak_concurrent = 0;
        size_t peak_footprint = 0;
    };

    // Keys and the context are shared by every worker; each worker holds its
    // own CKKSEvaluator (its constant caches are per worker).  boot may be null.
    Server(const CKKSContext &ctx, const SecretKey *sk, RelinKeys &rk, GaloisKeys &gk, double scale, Options opt,
           const Bootstrapper::Params *boot = nullptr);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Thread-safe.  The inputs must be complete on the caller's stream (submit
    // synchronises it).
    uint64_t submit(std::vector<Job> jobs);
    // Blocks until the request is done; one output per job.
    std::vector<Ciphertext> wait(uint64_t id);
    void drain();
    Stats stats() const;

private:
    struct Request {
        uint64_t id;
        std::vector<Job> jobs;
        std::vector<Ciphertext> outputs;
        size_t footprint = 0;
        double t_submit = 0, t_done = 0;
        bool done = false;
        std::string error;
    };
    void worker(int lane);
    void run_lane(int lane);
    static double now_ms();

    const CKKSContext &ctx_;
    const SecretKey *sk_;
    RelinKeys &rk_;
    GaloisKeys &gk_;
    double scale_;
    Options opt_;
    bool have_boot_ = false;
    Bootstrapper::Params boot_;

    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Request>> queue_;
    std::map<uint64_t, std::shared_ptr<Request>> all_;
    std::vector<double> latencies_;
    size_t in_flight_ = 0, in_flight_bytes_ = 0, submitted_ = 0;
    size_t peak_concurrent_ = 0, peak_footprint_ = 0;
    uint64_t next_id_ = 1;
    bool stop_ = false;
    std::string lane_error_;
    std::vector<std::thread> workers_;
};

}  // namespace nexus
