// Request loop over the evaluator programs (include/nexus/server.hpp).
#include "../../../include/nexus/server.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>

namespace nexus {

double Server::now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

Server::Server(const CKKSContext &ctx, const SecretKey *sk, RelinKeys &rk, GaloisKeys &gk, double scale, Options opt,
               const Bootstrapper::Params *boot, const BertLayer *bert)
    : ctx_(ctx), sk_(sk), rk_(rk), gk_(gk), scale_(scale), opt_(opt), bert_(bert)
{
    if (opt_.lanes < 1 || opt_.lanes > 16) throw std::invalid_argument("Server: lanes must be 1..16");
    if (boot) {
        have_boot_ = true;
        boot_ = *boot;
    }
    if (bert_) {
        if (!boot) throw std::invalid_argument("Server: a BERT layer needs the bootstrapper parameters");
        if (bert_->context()->get() != ctx.get()) throw std::invalid_argument("Server: BERT layer on another context");
        tokens_ = bert_->tokens_per_ct();
    }
    for (int i = 0; i < opt_.lanes; i++) workers_.emplace_back(&Server::worker, this, i + 1);
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : workers_) t.join();
}

uint64_t Server::submit(std::vector<Job> jobs)
{
    if (jobs.empty()) throw std::invalid_argument("Server: empty request");
    {
        std::lock_guard<std::mutex> g(mu_);
        if (!lane_error_.empty()) throw std::runtime_error("Server: a lane failed to start: " + lane_error_);
    }
    auto r = std::make_shared<Request>();
    size_t bytes = 0;
    const size_t n = ctx_.poly_degree();
    size_t extra = 0;
    for (const Job &j : jobs) {
        if (j.kind == JobKind::Bert) {
            if (!bert_) throw std::invalid_argument("Server: no BERT layer configured");
            if (j.inputs.size() != (size_t)bert_->shape().hidden / BertLayer::kGroup)
                throw std::invalid_argument("Server: a BERT job needs hidden / 256 input ciphertexts");
            if (!j.segments.empty()) {
                std::vector<bool> used(tokens_, false);
                for (const BertLayer::Segment &g : j.segments) {
                    if (g.len < 1 || g.start < 0 || (size_t)g.start + (size_t)g.len > tokens_)
                        throw std::invalid_argument("Server: segment out of range");
                    for (int i = g.start; i < g.start + g.len; i++) {
                        if (used[i]) throw std::invalid_argument("Server: overlapping segments");
                        used[i] = true;
                    }
                }
            } else if (j.len < 1 || (size_t)j.len > tokens_) {
                throw std::invalid_argument("Server: sequence length out of range");
            }
            for (const Ciphertext &c : j.inputs)
                if (!c.get()) throw std::invalid_argument("Server: job without input");
            extra += opt_.bert_footprint;
            continue;
        }
        if (!j.input.get()) throw std::invalid_argument("Server: job without input");
        bytes += j.input.size() * j.input.coeff_modulus_size() * n * 8;
    }
    // the inputs were produced on the caller's stream: complete them before a
    // worker's stream reads them
    ctx_.synchronize();
    r->jobs = std::move(jobs);
    r->footprint = (size_t)(opt_.footprint_factor * (double)bytes) + extra;
    {
        std::lock_guard<std::mutex> g(mu_);
        r->id = next_id_++;
        r->t_submit = now_ms();
        queue_.push_back(r);
        all_[r->id] = r;
        submitted_++;
    }
    cv_work_.notify_one();
    return r->id;
}

std::vector<Ciphertext> Server::wait(uint64_t id)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = all_.find(id);
    if (it == all_.end()) throw std::invalid_argument("Server: unknown request id");
    std::shared_ptr<Request> r = it->second;
    cv_done_.wait(lk, [&] { return r->done; });
    all_.erase(id);
    lk.unlock();
    if (!r->error.empty()) throw std::runtime_error("Server: request failed: " + r->error);
    // copy the outputs onto the caller's stream and memory, and complete the
    // copies before the worker lane's buffers go back to its cache
    std::vector<Ciphertext> res;
    for (const Ciphertext &o : r->outputs) res.push_back(Ciphertext(o));
    ctx_.synchronize();
    r->outputs.clear();
    return res;
}

void Server::pause(bool paused)
{
    {
        std::lock_guard<std::mutex> g(mu_);
        paused_ = paused;
    }
    cv_work_.notify_all();
}

void Server::set_bert_concurrency(int n)
{
    if (n < 1 || n > opt_.lanes) throw std::invalid_argument("Server: BERT concurrency must be 1..lanes");
    {
        std::lock_guard<std::mutex> g(mu_);
        opt_.bert_concurrency = n;
    }
    cv_work_.notify_all();
}

void Server::drain()
{
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return queue_.empty() && in_flight_ == 0; });
}

Server::Stats Server::stats() const
{
    std::lock_guard<std::mutex> g(mu_);
    Stats s;
    s.submitted = submitted_;
    s.completed = latencies_.size();
    s.peak_concurrent = peak_concurrent_;
    s.peak_footprint = peak_footprint_;
    s.packed_calls = packed_calls_;
    s.packed_requests = packed_requests_;
    if (!latencies_.empty()) {
        std::vector<double> v = latencies_;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double pos = p * (double)(v.size() - 1);
            const size_t lo = (size_t)pos;
            const size_t hi = std::min(lo + 1, v.size() - 1);
            return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s.mean_ms = sum / (double)v.size();
        s.max_ms = v.back();
    }
    return s;
}

void Server::worker(int lane)
{
    try {
        run_lane(lane);
    } catch (const std::exception &e) {
        // a lane that cannot start takes no requests; once no lane is left, every
        // queued request fails with its error instead of waiting forever
        std::vector<std::shared_ptr<Request>> failed;
        {
            std::lock_guard<std::mutex> g(mu_);
            lane_error_ = e.what();
            if (++dead_lanes_ == opt_.lanes) {
                const double t = now_ms();
                for (auto &r : queue_) {
                    r->error = std::string("no lane left to serve the request: ") + lane_error_;
                    r->t_done = t;
                    r->done = true;
                }
                queue_.clear();
            }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

void Server::run_lane(int lane)
{
    // this thread's work runs on its own lane (stream + memory cache)
    check(nx_context_bind_lane(ctx_.get(), (size_t)lane));
    CKKSEvaluator ck(&ctx_, sk_, &rk_, &gk_, scale_);
    ck.evaluator.set_per_op(opt_.per_op);
    std::unique_ptr<Bootstrapper> boot;
    if (have_boot_) boot.reset(new Bootstrapper(ck, boot_));
    std::unique_ptr<BertLayer> bert;
    if (bert_) bert.reset(new BertLayer(*bert_, ck, *boot));
    GELUEvaluator gelu(ck);
    SoftmaxEvaluator softmax(ck);
    LNEvaluator ln(ck);
    for (;;) {
        // a request, plus (continuous batching) queued BERT requests packed with it
        std::vector<std::shared_ptr<Request>> group;
        std::vector<int> set_of;  // set of each member (sets run in one lockstep call)
        bool bert_call = false;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                if (stop_) {
                    check(nx_context_bind_lane(ctx_.get(), 0));
                    return;
                }
                // admission: the oldest request whose footprint fits; with nothing
                // in flight the oldest is always admitted
                auto pick = queue_.end();
                for (auto it = queue_.begin(); !paused_ && it != queue_.end(); ++it) {
                    const bool has_bert = std::any_of((*it)->jobs.begin(), (*it)->jobs.end(),
                                                      [](const Job &j) { return j.kind == JobKind::Bert; });
                    if (has_bert && bert_running_ >= opt_.bert_concurrency) continue;
                    if (opt_.memory_budget == 0 || in_flight_ == 0 ||
                        in_flight_bytes_ + (*it)->footprint <= opt_.memory_budget) {
                        pick = it;
                        break;
                    }
                }
                if (pick != queue_.end()) {
                    std::shared_ptr<Request> r = *pick;
                    queue_.erase(pick);
                    r->charged = r->footprint;
                    group.push_back(r);
                    set_of.push_back(0);
                    for (const Job &j : r->jobs)
                        if (j.kind == JobKind::Bert) {
                            bert_call = true;
                            break;
                        }
                    if (bert_call) bert_running_++;
                    if (opt_.pack && bert && packable(*r)) {
                        // later BERT requests whose rows still fit ride in the same set;
                        // with more queued, further sets run in the same lockstep call
                        const Ciphertext &c0 = r->jobs[0].inputs[0];
                        auto compatible = [&](const Request &q) {
                            return packable(q) && q.jobs[0].inputs[0].scale() == c0.scale() &&
                                   q.jobs[0].inputs[0].coeff_modulus_size() == c0.coeff_modulus_size();
                        };
                        size_t used = (size_t)r->jobs[0].len;
                        int set = 0;
                        for (;;) {
                            for (auto it = queue_.begin(); it != queue_.end() && used < tokens_;) {
                                const Request &q = **it;
                                if (compatible(q) && used + (size_t)q.jobs[0].len <= tokens_) {
                                    used += (size_t)q.jobs[0].len;
                                    (*it)->charged = 0;  // shares the call's working memory
                                    group.push_back(*it);
                                    set_of.push_back(set);
                                    it = queue_.erase(it);
                                } else {
                                    ++it;
                                }
                            }
                            if (set + 1 >= opt_.bert_sets_per_call) break;
                            // the next set starts with the oldest compatible request left
                            auto nx = std::find_if(queue_.begin(), queue_.end(),
                                                   [&](const std::shared_ptr<Request> &q) { return compatible(*q); });
                            if (nx == queue_.end()) break;
                            set++;
                            used = (size_t)(*nx)->jobs[0].len;
                            (*nx)->charged = 0;
                            group.push_back(*nx);
                            set_of.push_back(set);
                            queue_.erase(nx);
                        }
                    }
                    for (auto &g : group) {
                        in_flight_++;
                        in_flight_bytes_ += g->charged;
                    }
                    peak_concurrent_ = std::max(peak_concurrent_, in_flight_);
                    peak_footprint_ = std::max(peak_footprint_, in_flight_bytes_);
                    if (group.size() > 1) {
                        packed_calls_++;
                        packed_requests_ += group.size();
                    }
                    break;
                }
                cv_work_.wait(lk);
            }
        }
        std::vector<std::vector<Ciphertext>> outs(group.size());
        std::string err;
        try {
            if (group.size() > 1) {
                // pack: member k at rows [off_k, off_k + len_k) of its set, one
                // layer call (sets in lockstep when there are several), unpack:
                // every member of the call gets its rows moved back, the others cleared
                const int nsets = set_of.back() + 1;
                std::vector<std::vector<Ciphertext>> sets(nsets);
                std::vector<std::vector<BertLayer::Segment>> segs(nsets);
                std::vector<int> off(nsets, 0);
                std::vector<BertLayer::Segment> where(group.size());
                for (size_t k = 0; k < group.size(); k++) {
                    const int st = set_of[k];
                    const Job &j = group[k]->jobs[0];
                    std::vector<Ciphertext> sh = off[st] ? bert->shift_rows(j.inputs, off[st]) : j.inputs;
                    if (sets[st].empty())
                        sets[st] = std::move(sh);
                    else
                        for (size_t c = 0; c < sets[st].size(); c++) ck.evaluator.add_inplace(sets[st][c], sh[c]);
                    where[k] = {off[st], j.len};
                    segs[st].push_back(where[k]);
                    off[st] += j.len;
                }
                std::vector<std::vector<Ciphertext>> ys;
                if (nsets > 1)
                    ys = bert->forward_sets(sets, segs);
                else
                    ys.push_back(bert->forward(sets[0], segs[0]));
                for (size_t k = 0; k < group.size(); k++)
                    outs[k] = bert->extract_rows(ys[set_of[k]], where[k].start, where[k].len);
            } else {
                for (Job &j : group[0]->jobs) {
                    if (j.kind == JobKind::Bert) {
                        if (!bert) throw std::invalid_argument("server has no BERT layer");
                        std::vector<Ciphertext> y =
                            j.segments.empty() ? bert->forward(j.inputs, j.len) : bert->forward(j.inputs, j.segments);
                        for (auto &c : y) outs[0].push_back(std::move(c));
                        continue;
                    }
                    Ciphertext out;
                    switch (j.kind) {
                    case JobKind::Gelu: gelu.gelu(j.input, out); break;
                    case JobKind::Softmax: softmax.softmax(j.input, out, j.len); break;
                    case JobKind::LayerNorm: ln.layer_norm(j.input, out, j.len); break;
                    case JobKind::Bootstrap:
                        if (!boot) throw std::invalid_argument("server has no bootstrapper");
                        boot->bootstrap(j.input, out);
                        break;
                    default: throw std::invalid_argument("unknown job kind");
                    }
                    outs[0].push_back(std::move(out));
                }
            }
            // completion on this lane's stream before the outputs are handed over
            ctx_.synchronize();
        } catch (const std::exception &e) {
            err = e.what();
            try {
                ctx_.synchronize();
            } catch (...) {
            }
        }
        const double t = now_ms();
        for (auto &g : group) g->jobs.clear();  // inputs: their last use on this lane is complete
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (bert_call) bert_running_--;
            for (size_t k = 0; k < group.size(); k++) {
                Request &g = *group[k];
                g.outputs = std::move(outs[k]);
                g.error = err;
                g.t_done = t;
                g.done = true;
                if (err.empty()) latencies_.push_back(t - g.t_submit);
                in_flight_--;
                in_flight_bytes_ -= g.charged;
            }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

}  // namespace nexus
