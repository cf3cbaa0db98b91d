// C ABI of the encrypted BERT layer (include/nexus_eval.h, BERT section).
#include <cstring>

#include "../../../include/nexus/bert.hpp"
#include "eval_handles.h"

using nexus_capi::need;
using nexus_capi::run;
using nexus_capi::store;

namespace {

std::vector<nexus::Ciphertext> inputs(const nx_bert_layer *l, const nx_ciphertext *const *x, size_t n)
{
    need(x);
    std::vector<nexus::Ciphertext> v;
    v.reserve(n);
    for (size_t i = 0; i < n; i++) {
        need(x[i]);
        v.push_back(nexus::Ciphertext::copy_of(*l->bt->ev->ctx, x[i]));
    }
    return v;
}

void outputs(const std::vector<nexus::Ciphertext> &r, nx_ciphertext *const *out, size_t n)
{
    need(out);
    if (r.size() != n) throw std::invalid_argument("wrong number of output ciphertexts");
    for (size_t i = 0; i < n; i++) {
        need(out[i]);
        store(r[i], out[i]);
    }
}

}  // namespace

extern "C" {

int nx_bert_layer_create(nx_bootstrapper *bt, int hidden, int heads, int ffn, const double *wq, const double *wk,
                         const double *wv, const double *wo, const double *w1, const double *w2, int baby,
                         nx_bert_layer **out)
{
    return run([&] {
        need(bt);
        need(out);
        nexus::BertShape shape{hidden, heads, ffn};
        nexus::BertLayer::Options opt;
        if (baby > 0) opt.baby = baby;
        auto l = std::make_unique<nx_bert_layer>();
        l->bt = bt;
        l->layer = std::make_unique<nexus::BertLayer>(*bt->ev->ckks, *bt->bt, shape,
                                                      std::vector<const double *>{wq, wk, wv, wo, w1, w2}, opt);
        *out = l.release();
    });
}

int nx_bert_layer_destroy(nx_bert_layer *layer)
{
    return run([&] { delete layer; });
}

int nx_bert_layer_rotation_steps(size_t slots, int hidden, int heads, int ffn, int baby, int *steps, size_t capacity,
                                 size_t *count)
{
    return run([&] {
        need(count);
        nexus::BertLayer::Options opt;
        if (baby > 0) opt.baby = baby;
        const std::vector<int> s = nexus::BertLayer::rotation_steps(slots, nexus::BertShape{hidden, heads, ffn}, opt);
        *count = s.size();
        if (steps) std::memcpy(steps, s.data(), std::min(capacity, s.size()) * sizeof(int));
    });
}

int nx_bert_layer_forward(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t n, int seq_len,
                          nx_ciphertext *const *out)
{
    return run([&] {
        need(layer);
        outputs(layer->layer->forward(inputs(layer, x, n), seq_len), out, n);
    });
}

int nx_bert_layer_forward_packed(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t n, const int *starts,
                                 const int *lens, size_t segments, nx_ciphertext *const *out)
{
    return run([&] {
        need(layer);
        need(starts);
        need(lens);
        std::vector<nexus::BertLayer::Segment> seg(segments);
        for (size_t k = 0; k < segments; k++) seg[k] = {starts[k], lens[k]};
        outputs(layer->layer->forward(inputs(layer, x, n), seg), out, n);
    });
}

int nx_bert_layer_forward_sets(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t sets, size_t n,
                               const int *seg_counts, const int *starts, const int *lens, nx_ciphertext *const *out)
{
    return run([&] {
        need(layer);
        need(x);
        need(seg_counts);
        need(starts);
        need(lens);
        need(out);
        std::vector<std::vector<nexus::Ciphertext>> xs(sets);
        std::vector<std::vector<nexus::BertLayer::Segment>> segs(sets);
        size_t k = 0;
        for (size_t s = 0; s < sets; s++) {
            xs[s] = inputs(layer, x + s * n, n);
            if (seg_counts[s] < 1) throw std::invalid_argument("every set needs a segment");
            for (int j = 0; j < seg_counts[s]; j++, k++) segs[s].push_back({starts[k], lens[k]});
        }
        std::vector<std::vector<nexus::Ciphertext>> ys = layer->layer->forward_sets(xs, segs);
        for (size_t s = 0; s < sets; s++) outputs(ys[s], out + s * n, n);
    });
}

int nx_bert_layer_linear(nx_bert_layer *layer, int matrix, const nx_ciphertext *const *x, size_t n_in, double factor,
                         nx_ciphertext *const *out, size_t n_out)
{
    return run([&] {
        need(layer);
        if (matrix < 0 || matrix >= nexus::BertLayer::kMatrices) throw std::invalid_argument("unknown matrix");
        outputs(layer->layer->linear((nexus::BertLayer::Matrix)matrix, inputs(layer, x, n_in), factor), out, n_out);
    });
}

int nx_bert_layer_attention(nx_bert_layer *layer, const nx_ciphertext *const *q, const nx_ciphertext *const *k,
                            const nx_ciphertext *const *v, size_t n, int seq_len, nx_ciphertext *const *out)
{
    return run([&] {
        need(layer);
        outputs(layer->layer->attention(inputs(layer, q, n), inputs(layer, k, n), inputs(layer, v, n), seq_len), out,
                n);
    });
}

int nx_bert_layer_layer_norm(nx_bert_layer *layer, const nx_ciphertext *const *x, size_t n,
                             nx_ciphertext *const *out)
{
    return run([&] {
        need(layer);
        outputs(layer->layer->layer_norm(inputs(layer, x, n)), out, n);
    });
}

int nx_bert_layer_set_stage_timing(nx_bert_layer *layer, int on)
{
    return run([&] {
        need(layer);
        layer->layer->set_stage_timing(on != 0);
    });
}

int nx_bert_layer_stage_ms(const nx_bert_layer *layer, char *names, size_t names_len, double *ms, size_t capacity,
                           size_t *count)
{
    return run([&] {
        need(layer);
        need(count);
        const auto &st = layer->layer->stage_ms();
        *count = st.size();
        std::string joined;
        for (size_t i = 0; i < st.size(); i++) {
            if (i) joined += ",";
            joined += st[i].first;
            if (ms && i < capacity) ms[i] = st[i].second;
        }
        if (names && names_len) {
            std::strncpy(names, joined.c_str(), names_len - 1);
            names[names_len - 1] = 0;
        }
    });
}

}  // extern "C"
