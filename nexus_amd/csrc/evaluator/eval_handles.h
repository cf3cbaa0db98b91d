// Handle of the evaluator-program C ABI (include/nexus_eval.h) and the
// exception-to-status mapping shared by its translation units.
#pragma once
#include <memory>

#include "../../../include/nexus/bert.hpp"
#include "../../../include/nexus/bootstrapper.hpp"
#include "../../../include/nexus/ckks.hpp"
#include "../../../include/nexus_eval.h"

struct nx_evaluator {
    std::unique_ptr<nexus::CKKSContext> ctx;
    std::unique_ptr<nexus::SecretKey> sk;
    std::unique_ptr<nexus::KSwitchKeys> rk, gk;
    std::unique_ptr<nexus::CKKSEvaluator> ckks;
};

struct nx_bootstrapper {
    nx_evaluator *ev;
    std::unique_ptr<nexus::Bootstrapper> bt;
};

struct nx_bert_layer {
    nx_bootstrapper *bt;
    std::unique_ptr<nexus::BertLayer> layer;
};

namespace nx {
void set_last_error(const char *msg);  // capi.cpp
}

namespace nexus_capi {

template <class F>
int run(F &&f)
{
    try {
        f();
        return NX_OK;
    } catch (const std::invalid_argument &e) {
        nx::set_last_error(e.what());
        return NX_ERR_INVALID_ARGUMENT;
    } catch (const std::out_of_range &e) {
        nx::set_last_error(e.what());
        return NX_ERR_OUT_OF_RANGE;
    } catch (const std::logic_error &e) {
        nx::set_last_error(e.what());
        return NX_ERR_LOGIC;
    } catch (const std::bad_alloc &) {
        nx::set_last_error("out of host memory");
        return NX_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        nx::set_last_error(e.what());
        return NX_ERR_UNKNOWN;
    }
}

inline void store(const nexus::Ciphertext &res, nx_ciphertext *out) { nexus::check(nx_ciphertext_copy(out, res.get())); }

inline void need(const void *p)
{
    if (!p) throw std::invalid_argument("null argument");
}

}  // namespace nexus_capi
