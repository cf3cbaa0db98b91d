// Definitions of the opaque C-ABI handle types (shared by the C-ABI sources).
#pragma once
#include "../../include/nexus_ckks.h"
#include "context.h"

struct nx_context {
    nx::Context impl;
    nx_context(int logn, const std::vector<int> &bits, uint32_t hw, size_t sparse, size_t special = 1)
        : impl(logn, bits, hw, sparse, special)
    {
    }
};
struct nx_ciphertext {
    nx::Ciphertext impl;
};
struct nx_plaintext {
    nx::Plaintext impl;
};
struct nx_secret_key {
    nx::SecretKey impl;
};
struct nx_kswitch_keys {
    nx::KSwitchKeys impl;
};
struct nx_ct_batch {
    nx::CtBatch impl;
};
struct nx_periodic_plain {
    nx::PeriodicPlain impl;
};

namespace nx {

// Runs f, converting exceptions into status codes and the thread-local message.
template <class F>
int guard_call(F &&f)
{
    try {
        f();
        return NX_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::invalid_argument &e) {
        set_last_error(e.what());
        return NX_ERR_INVALID_ARGUMENT;
    } catch (const std::out_of_range &e) {
        set_last_error(e.what());
        return NX_ERR_OUT_OF_RANGE;
    } catch (const std::logic_error &e) {
        set_last_error(e.what());
        return NX_ERR_LOGIC;
    } catch (const std::bad_alloc &e) {
        set_last_error(e.what());
        return NX_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return NX_ERR_UNKNOWN;
    }
}

inline void need_ptr(const void *p, const char *what)
{
    if (!p) fail(kInvalidArgument, std::string(what) + " is null");
}

}  // namespace nx
