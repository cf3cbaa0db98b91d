// C ABI of the request loop (include/nexus_eval.h, serving section).
#include "../../../include/nexus/server.hpp"
#include "eval_handles.h"

using nexus_capi::need;
using nexus_capi::run;

struct nx_bootstrapper;  // capi_boot.cpp
namespace nexus_capi {
const nexus::Bootstrapper::Params *bootstrapper_params(const nx_bootstrapper *bt);  // capi_boot.cpp
}

struct nx_server {
    std::unique_ptr<nexus::Server> impl;
    nx_evaluator *ev;
};

extern "C" {

int nx_server_create(nx_evaluator *ev, int lanes, size_t memory_budget, const nx_bootstrapper *boot,
                     nx_server **out)
{
    return run([&] {
        need(ev);
        need(out);
        if (!ev->rk || !ev->gk) throw std::invalid_argument("the evaluator needs relinearisation and Galois keys");
        nexus::Server::Options o;
        o.lanes = lanes;
        o.memory_budget = memory_budget;
        o.per_op = ev->ckks->evaluator.per_op();
        auto s = std::make_unique<nx_server>();
        s->ev = ev;
        s->impl = std::make_unique<nexus::Server>(*ev->ctx, ev->sk.get(), *ev->rk, *ev->gk, ev->ckks->scale, o,
                                                  boot ? nexus_capi::bootstrapper_params(boot) : nullptr);
        *out = s.release();
    });
}

int nx_server_create_bert(nx_evaluator *ev, int lanes, size_t memory_budget, const nx_bootstrapper *boot,
                          const nx_bert_layer *layer, int pack, size_t bert_footprint, nx_server **out)
{
    return run([&] {
        need(ev);
        need(boot);
        need(layer);
        need(out);
        if (!ev->rk || !ev->gk) throw std::invalid_argument("the evaluator needs relinearisation and Galois keys");
        nexus::Server::Options o;
        o.lanes = lanes;
        o.memory_budget = memory_budget;
        o.pack = pack != 0;
        o.per_op = ev->ckks->evaluator.per_op();
        if (bert_footprint) o.bert_footprint = bert_footprint;
        auto s = std::make_unique<nx_server>();
        s->ev = ev;
        s->impl = std::make_unique<nexus::Server>(*ev->ctx, ev->sk.get(), *ev->rk, *ev->gk, ev->ckks->scale, o,
                                                  nexus_capi::bootstrapper_params(boot), layer->layer.get());
        *out = s.release();
    });
}

int nx_server_set_bert_concurrency(nx_server *s, int n)
{
    return run([&] {
        need(s);
        s->impl->set_bert_concurrency(n);
    });
}

int nx_server_destroy(nx_server *s)
{
    return run([&] { delete s; });
}

int nx_server_submit(nx_server *s, const int *kinds, const int *lens, const nx_ciphertext *const *inputs,
                     size_t count, uint64_t *request_id)
{
    return run([&] {
        need(s);
        need(kinds);
        need(inputs);
        need(request_id);
        std::vector<nexus::Job> jobs;
        for (size_t k = 0; k < count; k++) {
            need(inputs[k]);
            if (kinds[k] < 0 || kinds[k] > 3) throw std::invalid_argument("unknown job kind");
            nexus::Job j;
            j.kind = (nexus::JobKind)kinds[k];
            j.len = lens ? lens[k] : 0;
            j.input = nexus::Ciphertext::copy_of(*s->ev->ctx, inputs[k]);
            jobs.push_back(std::move(j));
        }
        *request_id = s->impl->submit(std::move(jobs));
    });
}

int nx_server_submit_bert(nx_server *s, const nx_ciphertext *const *x, size_t n, int seq_len, uint64_t *request_id)
{
    return run([&] {
        need(s);
        need(x);
        need(request_id);
        nexus::Job j;
        j.kind = nexus::JobKind::Bert;
        j.len = seq_len;
        for (size_t k = 0; k < n; k++) {
            need(x[k]);
            j.inputs.push_back(nexus::Ciphertext::copy_of(*s->ev->ctx, x[k]));
        }
        std::vector<nexus::Job> jobs;
        jobs.push_back(std::move(j));
        *request_id = s->impl->submit(std::move(jobs));
    });
}

int nx_server_submit_bert_set(nx_server *s, const nx_ciphertext *const *x, size_t n, const int *starts,
                              const int *lens, size_t segments, uint64_t *request_id)
{
    return run([&] {
        need(s);
        need(x);
        need(starts);
        need(lens);
        need(request_id);
        if (!segments) throw std::invalid_argument("a packed set needs at least one segment");
        nexus::Job j;
        j.kind = nexus::JobKind::Bert;
        for (size_t k = 0; k < segments; k++) j.segments.push_back({starts[k], lens[k]});
        for (size_t k = 0; k < n; k++) {
            need(x[k]);
            j.inputs.push_back(nexus::Ciphertext::copy_of(*s->ev->ctx, x[k]));
        }
        std::vector<nexus::Job> jobs;
        jobs.push_back(std::move(j));
        *request_id = s->impl->submit(std::move(jobs));
    });
}

int nx_server_wait(nx_server *s, uint64_t request_id, nx_ciphertext *const *outputs, size_t count)
{
    return run([&] {
        need(s);
        need(outputs);
        std::vector<nexus::Ciphertext> res = s->impl->wait(request_id);
        if (res.size() != count) throw std::invalid_argument("output count does not match the request's jobs");
        for (size_t k = 0; k < count; k++) {
            need(outputs[k]);
            nexus::check(nx_ciphertext_copy(outputs[k], res[k].get()));
        }
    });
}

int nx_server_pause(nx_server *s, int paused)
{
    return run([&] {
        need(s);
        s->impl->pause(paused != 0);
    });
}

int nx_server_drain(nx_server *s)
{
    return run([&] {
        need(s);
        s->impl->drain();
    });
}

int nx_server_stats(const nx_server *s, size_t *completed, double *p50_ms, double *p99_ms, double *mean_ms,
                    size_t *peak_concurrent)
{
    return run([&] {
        need(s);
        const nexus::Server::Stats st = s->impl->stats();
        if (completed) *completed = st.completed;
        if (p50_ms) *p50_ms = st.p50_ms;
        if (p99_ms) *p99_ms = st.p99_ms;
        if (mean_ms) *mean_ms = st.mean_ms;
        if (peak_concurrent) *peak_concurrent = st.peak_concurrent;
    });
}

int nx_server_batching_stats(const nx_server *s, size_t *packed_calls, size_t *packed_requests, size_t *peak_footprint)
{
    return run([&] {
        need(s);
        const nexus::Server::Stats st = s->impl->stats();
        if (packed_calls) *packed_calls = st.packed_calls;
        if (packed_requests) *packed_requests = st.packed_requests;
        if (peak_footprint) *peak_footprint = st.peak_footprint;
    });
}

}  // extern "C"
