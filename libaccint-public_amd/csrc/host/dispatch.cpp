// dispatch.cpp -- the per-class dispatch table of the J/K kernels: JSON form,
// the autotuner entry of Engine, and LIBACCINT_DISPATCH_TABLE.
#include <libaccint/engine.hpp>

#include "dispatch_detail.hpp"
#include "json_reader.hpp"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace libaccint {

namespace {
constexpr int kClasses = 55;

// pair class p = La (La + 1) / 2 + Lb, La >= Lb
void pair_am(int p, int& la, int& lb) {
    la = 0;
    while ((la + 1) * (la + 2) / 2 <= p) ++la;
    lb = p - la * (la + 1) / 2;
}

DispatchEntry from_tune(const lai_class_tune& t) {
    DispatchEntry e;
    e.cls = t.cls;
    e.am = quartet_class_am(t.cls);
    e.strategy = static_cast<GpuExecutionStrategy>(t.strategy);
    e.chunk = t.chunk;
    e.blocks_per_cu = t.blocks_per_cu;
    e.ms = t.ms;
    e.ms_default = t.ms_default;
    return e;
}
}  // namespace

int quartet_class_index(const std::array<int, 4>& am) noexcept {
    for (int x : am)
        if (x < 0 || x > 3) return -1;
    if (am[0] < am[1] || am[2] < am[3]) return -1;
    const int pb = am[0] * (am[0] + 1) / 2 + am[1], pk = am[2] * (am[2] + 1) / 2 + am[3];
    if (pb < pk) return -1;
    return pb * (pb + 1) / 2 + pk;
}

std::array<int, 4> quartet_class_am(int cls) {
    if (cls < 0 || cls >= kClasses) throw InvalidArgumentException("quartet class " + std::to_string(cls) + " out of range");
    int pb, pk, la, lb, lc, ld;
    pair_am(cls, pb, pk);
    pair_am(pb, la, lb);
    pair_am(pk, lc, ld);
    return {la, lb, lc, ld};
}

BackendType DispatchPolicy::select_backend(WorkUnitType work_type, Size batch_size, int total_am, Size n_primitives,
                                           BackendHint hint, bool gpu_available) const {
    if (hint == BackendHint::ForceCPU) return BackendType::CPU;
    if (hint == BackendHint::ForceGPU) return gpu_available ? BackendType::HIP : BackendType::CPU;
    if (!gpu_available) return BackendType::CPU;
    const DispatchConfig& c = config_;
    const bool single = work_type == WorkUnitType::SingleShellPair || work_type == WorkUnitType::SingleShellQuartet;
    bool gpu = false;
    switch (work_type) {
        case WorkUnitType::SingleShellPair:
        case WorkUnitType::SingleShellQuartet:
            gpu = total_am >= c.high_am_threshold + 2;
            break;
        case WorkUnitType::ShellSetPair:
            gpu = batch_size >= c.min_gpu_batch_size || n_primitives >= c.min_gpu_primitives ||
                  total_am >= c.high_am_threshold;
            break;
        case WorkUnitType::ShellSetQuartet:
            gpu = batch_size >= c.min_gpu_batch_size || n_primitives >= c.min_gpu_primitives ||
                  total_am >= c.high_am_threshold - 1;
            break;
        case WorkUnitType::FullBasis:
            gpu = batch_size >= c.min_gpu_shells;
            break;
    }
    if (hint == BackendHint::PreferCPU) {
        gpu = !single && gpu && (batch_size >= 2 * c.min_gpu_batch_size || total_am >= c.high_am_threshold + 2);
    } else if (hint == BackendHint::PreferGPU && !single) {
        gpu = gpu || batch_size >= c.min_gpu_batch_size / 4 || total_am >= c.high_am_threshold - 1;
    }
    return gpu ? BackendType::HIP : BackendType::CPU;
}

std::string_view to_string(GpuExecutionStrategy s) noexcept {
    switch (s) {
        case GpuExecutionStrategy::ThreadPerQuartet: return "thread_per_quartet";
        case GpuExecutionStrategy::WarpPerQuartet: return "warp_per_quartet";
        case GpuExecutionStrategy::BlockPerQuartet: return "block_per_quartet";
    }
    return "unknown";
}

GpuExecutionStrategy gpu_execution_strategy_from_string(std::string_view s) {
    if (s == "thread_per_quartet") return GpuExecutionStrategy::ThreadPerQuartet;
    if (s == "warp_per_quartet") return GpuExecutionStrategy::WarpPerQuartet;
    if (s == "block_per_quartet") return GpuExecutionStrategy::BlockPerQuartet;
    throw std::invalid_argument("Unknown GpuExecutionStrategy: " + std::string(s));
}

// {"format": "libaccint-dispatch-table", "version": 1, "entries": [{"class": c,
//  "am": [la, lb, lc, ld], "strategy": "...", "chunk": n, "blocks_per_cu": n,
//  "ms": t, "ms_default": t}, ...]}; an entry names its class by "class",
// "am" or both (then they must agree).
DispatchTable DispatchTable::parse_json(const std::string& text) {
    const detail::JValue root = detail::JParser(text, "dispatch table JSON").parse();
    const detail::JValue* entries = root.get("entries");
    if (root.kind != detail::JValue::Object || !entries || entries->kind != detail::JValue::Array)
        throw InvalidArgumentException("dispatch table JSON: expected an object with an 'entries' array");
    if (const detail::JValue* v = root.get("version"))
        if (v->kind != detail::JValue::Number || v->num != 1.0)
            throw InvalidArgumentException("dispatch table JSON: unsupported version");
    auto num = [](const detail::JValue* v, const char* key) -> double {
        if (!v || v->kind != detail::JValue::Number)
            throw InvalidArgumentException(std::string("dispatch table JSON: '") + key + "' must be a number");
        return v->num;
    };
    DispatchTable t;
    for (const auto& x : entries->arr) {
        if (x.kind != detail::JValue::Object) throw InvalidArgumentException("dispatch table JSON: entry is not an object");
        DispatchEntry e;
        int cls = -1;
        if (const detail::JValue* c = x.get("class")) cls = static_cast<int>(num(c, "class"));
        if (const detail::JValue* a = x.get("am")) {
            if (a->kind != detail::JValue::Array || a->arr.size() != 4)
                throw InvalidArgumentException("dispatch table JSON: 'am' must hold four integers");
            std::array<int, 4> am{};
            for (int i = 0; i < 4; ++i) am[i] = static_cast<int>(num(&a->arr[i], "am"));
            const int ca = quartet_class_index(am);
            if (ca < 0 || ca >= kClasses)
                throw InvalidArgumentException("dispatch table JSON: 'am' is not a canonical class up to f");
            if (cls >= 0 && cls != ca) throw InvalidArgumentException("dispatch table JSON: 'class' and 'am' disagree");
            cls = ca;
        }
        if (cls < 0 || cls >= kClasses) throw InvalidArgumentException("dispatch table JSON: entry without a valid class");
        e.cls = cls;
        e.am = quartet_class_am(cls);
        const detail::JValue* st = x.get("strategy");
        if (!st || st->kind != detail::JValue::String)
            throw InvalidArgumentException("dispatch table JSON: 'strategy' must be a string");
        e.strategy = gpu_execution_strategy_from_string(st->str);
        if (const detail::JValue* v = x.get("chunk")) e.chunk = static_cast<int>(num(v, "chunk"));
        if (const detail::JValue* v = x.get("blocks_per_cu")) e.blocks_per_cu = static_cast<int>(num(v, "blocks_per_cu"));
        if (const detail::JValue* v = x.get("ms")) e.ms = num(v, "ms");
        if (const detail::JValue* v = x.get("ms_default")) e.ms_default = num(v, "ms_default");
        t.entries.push_back(e);
    }
    return t;
}

std::string DispatchTable::dump_json() const {
    std::ostringstream o;
    o << std::setprecision(6);
    o << "{\n  \"format\": \"libaccint-dispatch-table\",\n  \"version\": 1,\n  \"entries\": [";
    for (size_t i = 0; i < entries.size(); ++i) {
        const DispatchEntry& e = entries[i];
        const std::array<int, 4> am = quartet_class_am(e.cls);
        o << (i ? ",\n    " : "\n    ") << "{\"class\": " << e.cls << ", \"am\": [" << am[0] << ", " << am[1]
          << ", " << am[2] << ", " << am[3] << "], \"strategy\": \"" << to_string(e.strategy)
          << "\", \"chunk\": " << e.chunk << ", \"blocks_per_cu\": " << e.blocks_per_cu << ", \"ms\": " << e.ms
          << ", \"ms_default\": " << e.ms_default << "}";
    }
    o << (entries.empty() ? "]\n}\n" : "\n  ]\n}\n");
    return o.str();
}

DispatchTable DispatchTable::from_json(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("dispatch table: cannot read " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_json(ss.str());
}

void DispatchTable::to_json(const std::string& path) const {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("dispatch table: cannot write " + path);
    f << dump_json();
    if (!f) throw std::runtime_error("dispatch table: write failed for " + path);
}

namespace detail {
DispatchTable autotune_handle(lai_basis* b, int flags, const Real* d_D, hipStream_t s, int reps, const char* what) {
    std::vector<lai_class_tune> tab(kClasses);
    int n = 0;
    device::check(lai_dev_autotune(b, d_D, flags, reps, tab.data(), kClasses, &n, s), what);
    DispatchTable t;
    for (int i = 0; i < n && i < kClasses; ++i) t.entries.push_back(from_tune(tab[i]));
    return t;
}

DispatchTable dispatch_table_of(lai_basis* b, const char* what) {
    std::vector<lai_class_tune> tab(kClasses);
    int n = 0;
    device::check(lai_dev_get_tuning(b, tab.data(), kClasses, &n), what);
    DispatchTable t;
    for (int i = 0; i < n && i < kClasses; ++i) t.entries.push_back(from_tune(tab[i]));
    return t;
}

void set_dispatch_table_of(lai_basis* b, const DispatchTable& table, const char* what) {
    std::vector<lai_class_tune> tab;
    for (const auto& e : table.entries) {
        lai_class_tune t{};
        t.cls = e.cls;
        t.strategy = static_cast<int>(e.strategy);
        t.chunk = e.chunk;
        t.blocks_per_cu = e.blocks_per_cu;
        tab.push_back(t);
    }
    device::check(lai_dev_set_tuning(b, tab.empty() ? nullptr : tab.data(), static_cast<int>(tab.size())), what);
}

void apply_dispatch_table_env(lai_basis* b) {
    const char* path = std::getenv("LIBACCINT_DISPATCH_TABLE");
    if (path && path[0] != '\0') set_dispatch_table_of(b, DispatchTable::from_json(path), "LIBACCINT_DISPATCH_TABLE");
}
}  // namespace detail

DispatchTable Engine::autotune(const JKOptions& o, const Real* d_D, hipStream_t s, int reps, const Real* D_host) {
    if (!d_D) throw InvalidArgumentException("autotune: null device density");
    prepare_worklist(o, D_host);
    lai_basis* b = device_basis();
    device::check(lai_dev_set_mixed_prim_cut(b, o.fp32_prim_cut), "autotune");
    const int flags = o.mixed_precision ? LAI_JK_MIXED : LAI_JK_FP64;
    DispatchTable t = detail::autotune_handle(b, flags, d_D, s, reps, "autotune");
    tuned_flags_ = flags;
    return t;
}

DispatchTable Engine::dispatch_table() { return detail::dispatch_table_of(device_basis(), "dispatch_table"); }

void Engine::set_dispatch_table(const DispatchTable& table) {
    detail::set_dispatch_table_of(device_basis(), table, "set_dispatch_table");
}

DispatchTable BatchEngine::autotune(const JKOptions& o, const Real* d_D, hipStream_t s, int reps,
                                    const Real* D_host) {
    if (!d_D) throw InvalidArgumentException("BatchEngine::autotune: null device density");
    prepare_worklist(o, D_host);
    lai_basis* b = device_handle();
    device::check(lai_dev_set_mixed_prim_cut(b, o.fp32_prim_cut), "BatchEngine::autotune");
    return detail::autotune_handle(b, o.mixed_precision ? LAI_JK_MIXED : LAI_JK_FP64, d_D, s, reps,
                                   "BatchEngine::autotune");
}

DispatchTable BatchEngine::dispatch_table() { return detail::dispatch_table_of(device_handle(), "BatchEngine::dispatch_table"); }

void BatchEngine::set_dispatch_table(const DispatchTable& table) {
    detail::set_dispatch_table_of(device_handle(), table, "BatchEngine::set_dispatch_table");
}

}  // namespace libaccint
