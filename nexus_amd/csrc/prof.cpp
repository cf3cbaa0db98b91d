#include "prof.h"

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nexus_ckks.h"
#include "../../include/nexus_prof.h"
#include "context.h"
#include "handles.h"

namespace nx {

namespace {

struct Stat {
    size_t launches = 0;
    double ms = 0, bytes = 0, flops = 0;
};

struct Pending {
    std::string cls;
    hipEvent_t a, b;
    double bytes, flops;
};

std::mutex g_mu;
bool g_all = false;
std::vector<std::string> g_sel;
std::map<std::string, Stat> g_stats;
std::map<std::pair<std::string, hipStream_t>, hipEvent_t> g_open;  // open launch per (class, stream)
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;

hipEvent_t take()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    NX_HIP(hipEventCreate(&e));
    return e;
}

void drain()
{
    for (auto &p : g_pending) {
        NX_HIP(hipEventSynchronize(p.b));
        float ms = 0;
        NX_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        Stat &s = g_stats[p.cls];
        s.ms += ms;
        s.bytes += p.bytes;
        s.flops += p.flops;
        s.launches++;
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

bool selected(const char *cls)
{
    if (g_all) return true;
    for (auto &s : g_sel)
        if (s == cls) return true;
    return false;
}

}  // namespace

bool prof_active(const char *cls)
{
    std::lock_guard<std::mutex> g(g_mu);
    return selected(cls);
}

void prof_record(const char *cls, hipStream_t s, bool begin, double bytes, double flops)
{
    std::lock_guard<std::mutex> g(g_mu);
    if (!selected(cls)) return;
    if (begin) {
        hipEvent_t e = take();
        NX_HIP(hipEventRecord(e, s));
        g_open[{cls, s}] = e;
    } else {
        auto it = g_open.find({cls, s});
        if (it == g_open.end()) return;
        hipEvent_t b = take();
        NX_HIP(hipEventRecord(b, s));
        g_pending.push_back({cls, it->second, b, bytes, flops});
        g_open.erase(it);
        if (g_pending.size() > 8192) drain();
    }
}

// ---------------------------------------------------------------- operation tally
namespace {
std::mutex g_tally_mu;
bool g_tally_on = false;
std::map<std::pair<std::string, size_t>, double> g_tally;
}  // namespace

void tally(const char *kind, size_t limbs, double count)
{
    if (!g_tally_on) return;
    std::lock_guard<std::mutex> g(g_tally_mu);
    if (g_tally_on) g_tally[{kind, limbs}] += count;
}

}  // namespace nx

extern "C" {

int nx_profile_select(const char *classes)
{
    try {
        std::lock_guard<std::mutex> g(nx::g_mu);
        nx::drain();
        nx::g_stats.clear();
        nx::g_sel.clear();
        nx::g_all = false;
        std::string s = classes ? classes : "";
        if (s == "*") {
            nx::g_all = true;
        } else {
            size_t pos = 0;
            while (pos <= s.size()) {
                size_t c = s.find(',', pos);
                if (c == std::string::npos) c = s.size();
                if (c > pos) nx::g_sel.push_back(s.substr(pos, c - pos));
                pos = c + 1;
            }
        }
        return NX_OK;
    } catch (const nx::Error &e) {
        nx::set_last_error(e.what());
        return e.code;
    }
}

int nx_profile_read(const char *kernel_class, size_t *launches, double *total_ms, double *bytes)
{
    try {
        std::lock_guard<std::mutex> g(nx::g_mu);
        nx::drain();
        nx::Stat st;
        auto it = nx::g_stats.find(kernel_class ? kernel_class : "");
        if (it != nx::g_stats.end()) st = it->second;
        if (launches) *launches = st.launches;
        if (total_ms) *total_ms = st.ms;
        if (bytes) *bytes = st.bytes;
        return NX_OK;
    } catch (const nx::Error &e) {
        nx::set_last_error(e.what());
        return e.code;
    }
}

int nx_profile_read_flops(const char *kernel_class, double *flops)
{
    try {
        std::lock_guard<std::mutex> g(nx::g_mu);
        nx::drain();
        auto it = nx::g_stats.find(kernel_class ? kernel_class : "");
        if (flops) *flops = it == nx::g_stats.end() ? 0.0 : it->second.flops;
        return NX_OK;
    } catch (const nx::Error &e) {
        nx::set_last_error(e.what());
        return e.code;
    }
}

int nx_profile_classes(char *buf, size_t len)
{
    try {
        std::lock_guard<std::mutex> g(nx::g_mu);
        nx::drain();
        std::string out;
        for (auto &kv : nx::g_stats) {
            if (!out.empty()) out += ",";
            out += kv.first;
        }
        if (buf && len) {
            std::strncpy(buf, out.c_str(), len - 1);
            buf[len - 1] = 0;
        }
        return NX_OK;
    } catch (const nx::Error &e) {
        nx::set_last_error(e.what());
        return e.code;
    }
}

int nx_op_tally(int start, char *buf, size_t len, size_t *needed)
{
    return nx::guard_call([&] {
        std::lock_guard<std::mutex> g(nx::g_tally_mu);
        if (start) {
            nx::g_tally.clear();
            nx::g_tally_on = true;
            if (needed) *needed = 0;
            return;
        }
        nx::g_tally_on = false;
        std::string out = "{";
        std::string kind;
        for (auto &kv : nx::g_tally) {
            if (kv.first.first != kind) {
                if (!kind.empty()) out += "},";
                kind = kv.first.first;
                out += "\"" + kind + "\": {";
            } else {
                out += ", ";
            }
            char num[64];
            std::snprintf(num, sizeof num, "\"%zu\": %.0f", kv.first.second, kv.second);
            out += num;
        }
        if (!kind.empty()) out += "}";
        out += "}";
        if (needed) *needed = out.size() + 1;
        if (buf && len) {
            std::strncpy(buf, out.c_str(), len - 1);
            buf[len - 1] = 0;
        }
    });
}

int nx_profile_valu_rate(int kind, double *ops_per_s)
{
    return nx::guard_call([&] {
        nx::need_ptr(ops_per_s, "ops_per_s");
        *ops_per_s = nx::probe_valu_rate(kind, nullptr);
    });
}

}  // extern "C"
