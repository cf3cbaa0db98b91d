// Request loop for encrypted inference over the evaluator programs.
//
// A request is a list of evaluator calls (GELU, softmax, LayerNorm, bootstrap,
// an encrypted BERT layer) on its own input ciphertexts.  Workers, one per lane of the context (a HIP
// stream with its own memory cache), take requests in arrival order as soon
// as they are free and the request's ciphertext footprint fits in the memory
// budget; requests on different lanes overlap on the device while sharing the
// device-resident keys.  Each request's result depends only on its inputs, so
// it is the same whether it runs alone or among others.
//
// Continuous batching of BERT requests: a request made of one BERT-layer job
// (a sequence encrypted at rows [0, len)) is packed, when a lane takes it,
// with the oldest queued BERT requests whose rows still fit in one set of
// ciphertexts (128 token rows at N = 2^16); each is moved to its rows by a
// token rotation, the set runs one layer call with attention kept inside each
// sequence, and each request's rows are moved back and the others cleared.
// At most bert_concurrency BERT calls run at once (one call keeps the GPU
// busy), so requests arriving meanwhile queue up and are packed together.
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

#include "bert.hpp"
#include "bootstrapper.hpp"
#include "ckks.hpp"

namespace nexus {

enum class JobKind : int { Gelu = 0, Softmax = 1, LayerNorm = 2, Bootstrap = 3, Bert = 4 };

struct Job {
    JobKind kind;
    int len = 0;  // row length of softmax / LayerNorm; sequence length of a BERT job
    Ciphertext input;
    std::vector<Ciphertext> inputs;  // BERT job: hidden / 256 ciphertexts, rows [0, len)
    // BERT job over a set the caller packed: its sequences' rows (len unused);
    // the set runs as one layer call and is never packed with other requests
    std::vector<BertLayer::Segment> segments;
};

class Server {
public:
    struct Options {
        int lanes = 2;               // concurrent requests (HIP streams)
        size_t memory_budget = 0;    // bytes of in-flight request footprint; 0 = no limit
        double footprint_factor = 6; // footprint = factor x input ciphertext bytes
        bool per_op = false;         // the lanes' evaluators run in per-op mode (Evaluator::set_per_op)
        bool pack = true;            // continuous batching of BERT requests
        int bert_concurrency = 1;    // BERT layer calls in flight at once (one call fills the GPU)
        // sets of queued BERT requests run in one lockstep layer call when more
        // are queued than one set holds (BertLayer::forward_sets: bootstrapping
        // steps over all sets, 1.06x the throughput of one call per set)
        int bert_sets_per_call = 2;
        size_t bert_footprint = size_t(8) << 30;  // working set of one BERT layer call
    };
    struct Stats {
        size_t submitted = 0, completed = 0;
        double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
        size_t peak_concurrent = 0;
        size_t peak_footprint = 0;
        size_t packed_calls = 0, packed_requests = 0;  // BERT sets and the requests they carried
    };

    // Keys and the context are shared by every worker; each worker holds its
    // own CKKSEvaluator (its constant caches are per worker).  boot may be null.
    // bert: a layer whose weights every lane shares (needs boot).
    Server(const CKKSContext &ctx, const SecretKey *sk, RelinKeys &rk, GaloisKeys &gk, double scale, Options opt,
           const Bootstrapper::Params *boot = nullptr, const BertLayer *bert = nullptr);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Thread-safe.  The inputs must be complete on the caller's stream (submit
    // synchronises it).
    uint64_t submit(std::vector<Job> jobs);
    // Blocks until the request is done; one output per job (hidden / 256 for a
    // BERT job).
    std::vector<Ciphertext> wait(uint64_t id);
    void drain();
    Stats stats() const;
    // While paused, queued requests are not admitted (requests submitted
    // meanwhile are batched together once resumed).
    void pause(bool paused);
    // BERT layer calls allowed in flight at once (1..lanes)
    void set_bert_concurrency(int n);

private:
    struct Request {
        uint64_t id;
        std::vector<Job> jobs;
        std::vector<Ciphertext> outputs;
        size_t footprint = 0, charged = 0;
        double t_submit = 0, t_done = 0;
        bool done = false;
        std::string error;
    };
    void worker(int lane);
    void run_lane(int lane);
    static bool packable(const Request &r)
    {
        return r.jobs.size() == 1 && r.jobs[0].kind == JobKind::Bert && r.jobs[0].segments.empty();
    }
    static double now_ms();

    const CKKSContext &ctx_;
    const SecretKey *sk_;
    RelinKeys &rk_;
    GaloisKeys &gk_;
    double scale_;
    Options opt_;
    bool have_boot_ = false;
    Bootstrapper::Params boot_;
    const BertLayer *bert_ = nullptr;
    size_t tokens_ = 0;

    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Request>> queue_;
    std::map<uint64_t, std::shared_ptr<Request>> all_;
    std::vector<double> latencies_;
    size_t in_flight_ = 0, in_flight_bytes_ = 0, submitted_ = 0;
    int bert_running_ = 0;
    int dead_lanes_ = 0;  // lanes that failed to start
    size_t peak_concurrent_ = 0, peak_footprint_ = 0, packed_calls_ = 0, packed_requests_ = 0;
    uint64_t next_id_ = 1;
    bool stop_ = false, paused_ = false;
    std::string lane_error_;
    std::vector<std::thread> workers_;
};

}  // namespace nexus
