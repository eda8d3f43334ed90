// Batched retained-operator evaluation over instances of uneven sizes.
//
// A batch holds size classes.  Class c stacks `instances` copies of one
// workload (Nx, Ny, Nz, F) as a single workload with F * instances fields:
// instance i occupies fields [F*i, F*i + F), i.e. planes z + Nz*(f + F*i), so
// every stage of a class is ONE launch over the planes of all its
// instances.  One evaluation of a class is the retained operator
//     compact retained --pruned inverse--> real --pruned forward--> compact
//     --K^2-grouped vertical GEMM--> modal
// and one evaluation of the batch runs every class back to back on one
// stream.  skb_batch_capture records that sequence once into a hipGraph;
// skb_batch_launch replays it with a single launch call.  The kernels clear
// nothing with hipMemsetAsync, so the graph replays correctly.
//
// skb_batch_autotune times every candidate of every stage of every class
// with HIP events and keeps the fastest: the launch shape (workgroups per
// launch: one per work item, or a persistent grid) of each stage and, for
// the vertical product of FP32 square families, its arithmetic (split-BF16
// or FP32 matrix cores).  Persistent kernels are separate instantiations and
// the two product forms round differently, so results agree across choices
// to the FP32 tolerance, not bit for bit.  skb_batch_tuning writes the
// choices as a spec string ("c<class>.s<stage>=<grid>;c<class>.p=<products>;
// ..."), and skb_batch_set_tuning applies one, so a tuning can be kept and
// reused without timing again.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

struct skb_batch_class {
    skb_workload w{};        // the stacked workload (fields = F * instances)
    uint64_t instances = 0, fields = 0, modes = 0, nj = 0;
    skb_pruned_plan* pruned = nullptr;
    skb_vertical_plan* vertical = nullptr;
    const void* x = nullptr;  // caller-owned input (skb_batch_bind)
    void* modal = nullptr;    // caller-owned output
    skb::DeviceBuffer real, back;
    int grid[5] = {0, 0, 0, 0, 0};  // pruned stages 0..3, vertical GEMM
    int products = SKB_PRODUCTS_SPLIT_BF16;
    ~skb_batch_class() {
        if (pruned) skb_pruned_plan_destroy(pruned);
        if (vertical) skb_vertical_plan_destroy(vertical);
    }
};

struct skb_batch {
    int precision = SKB_F32;
    std::vector<std::unique_ptr<skb_batch_class>> classes;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~skb_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

using namespace skb;

namespace {

void run_class(skb_batch_class& c, hipStream_t s) {
    check_status(skb_pruned_inverse(c.pruned, c.x, c.real.ptr, s));
    check_status(skb_pruned_forward(c.pruned, c.real.ptr, c.back.ptr, s));
    check_status(skb_vertical_forward(c.vertical, c.back.ptr, c.modal, s));
}

void require_bound(const skb_batch& b) {
    for (const auto& c : b.classes)
        if (!c->x || !c->modal) fail(SKB_LOGIC_ERROR, "every batch class needs bound input and output arrays");
}

void run_all(skb_batch& b, hipStream_t s) {
    for (auto& c : b.classes) run_class(*c, s);
}

void drop_graph(skb_batch& b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

void run_stage(skb_batch_class& c, int stage, hipStream_t s) {
    switch (stage) {
        case 0: check_status(skb_pruned_forward_rows(c.pruned, c.real.ptr, s)); break;
        case 1: check_status(skb_pruned_forward_columns(c.pruned, c.back.ptr, s)); break;
        case 2: check_status(skb_pruned_inverse_columns(c.pruned, c.x, s)); break;
        case 3: check_status(skb_pruned_inverse_rows(c.pruned, c.real.ptr, s)); break;
        default: check_status(skb_vertical_forward(c.vertical, c.back.ptr, c.modal, s)); break;
    }
}

void apply_grid(skb_batch_class& c, int stage, int grid) {
    if (stage < 4) check_status(skb_pruned_plan_set_grid(c.pruned, stage, grid));
    else check_status(skb_vertical_plan_set_grid(c.vertical, grid));
    c.grid[stage] = grid;
}

void apply_products(skb_batch_class& c, int products) {
    check_status(skb_vertical_plan_set_products(c.vertical, products));
    c.products = products;
}

// The product forms a class's vertical GEMM can take: split-BF16 and FP32
// where the family allows the split, FP32 (FP64) alone elsewhere.
std::vector<int> product_candidates(skb_batch_class& c) {
    const int was = c.products;
    apply_products(c, SKB_PRODUCTS_SPLIT_BF16);
    const bool split = skb_vertical_plan_products(c.vertical, 0) == SKB_PRODUCTS_SPLIT_BF16;
    apply_products(c, was);
    return split ? std::vector<int>{SKB_PRODUCTS_SPLIT_BF16, SKB_PRODUCTS_FP32} : std::vector<int>{SKB_PRODUCTS_FP32};
}

std::string tuning_spec(const skb_batch& b) {
    std::ostringstream o;
    for (size_t ci = 0; ci < b.classes.size(); ++ci) {
        const auto& c = *b.classes[ci];
        for (int st = 0; st < 5; ++st) o << "c" << ci << ".s" << st << "=" << c.grid[st] << ";";
        o << "c" << ci << ".p=" << c.products << ";";
    }
    return o.str();
}

}  // namespace

extern "C" {

int skb_batch_create(int precision, skb_batch** batch) {
    return guarded([&] {
        if (!batch) fail(SKB_INVALID_ARGUMENT, "null batch pointer");
        *batch = nullptr;
        check_precision(precision);
        auto b = std::make_unique<skb_batch>();
        b->precision = precision;
        *batch = b.release();
    });
}

int skb_batch_destroy(skb_batch* batch) {
    return guarded([&] { delete batch; });
}

int skb_batch_add_class(skb_batch* batch, const skb_workload* workload, uint64_t instances,
                        const skb_retained_mode* modes, uint64_t mode_count, uint64_t nj,
                        const skb_mode_group* groups, uint64_t group_count, const double* forward,
                        const double* inverse, int* class_index) {
    return guarded([&] {
        if (!batch || !workload || !class_index) fail(SKB_INVALID_ARGUMENT, "null argument");
        validate_workload(workload);
        if (instances == 0) fail(SKB_INVALID_ARGUMENT, "a size class needs at least one instance");
        auto c = std::make_unique<skb_batch_class>();
        c->w = *workload;
        c->w.fields = workload->fields * instances;
        c->instances = instances;
        c->fields = workload->fields;
        c->modes = mode_count;
        c->nj = nj;
        check_status(skb_pruned_plan_create(&c->w, modes, mode_count, batch->precision, &c->pruned));
        check_status(skb_vertical_plan_create(&c->w, nj, groups, group_count, forward, inverse, batch->precision,
                                              &c->vertical));
        const uint64_t cb = complex_bytes(batch->precision), rb = real_bytes(batch->precision);
        const uint64_t planes = c->w.nz * c->w.fields;
        c->back.allocate(mode_count * planes * cb);
        c->real.allocate(c->w.nx * c->w.ny * planes * rb);
        drop_graph(*batch);
        batch->classes.push_back(std::move(c));
        *class_index = static_cast<int>(batch->classes.size() - 1);
    });
}

// Binds the class's caller-owned device arrays in the stacked layout: input
// compact retained (plane + P*mode, P = Nz*F*instances; elements
// modes*P) and output modal (j + Nj*(field + F*instances*mode)).  A captured
// graph records the pointers, so binding drops it.
int skb_batch_bind(skb_batch* batch, int cls, const void* input, void* output) {
    return guarded([&] {
        if (!batch || cls < 0 || cls >= static_cast<int>(batch->classes.size()))
            fail(SKB_INVALID_ARGUMENT, "bad class index");
        if (!input || !output) fail(SKB_INVALID_ARGUMENT, "null device array");
        auto& c = *batch->classes[cls];
        c.x = input;
        c.modal = output;
        drop_graph(*batch);
    });
}

int skb_batch_run(skb_batch* batch, void* stream) {
    return guarded([&] {
        if (!batch) fail(SKB_INVALID_ARGUMENT, "null batch");
        require_bound(*batch);
        run_all(*batch, as_stream(stream));
    });
}

int skb_batch_capture(skb_batch* batch, void* stream) {
    return guarded([&] {
        if (!batch) fail(SKB_INVALID_ARGUMENT, "null batch");
        if (!stream) fail(SKB_INVALID_ARGUMENT, "graph capture needs a non-default stream");
        const hipStream_t s = as_stream(stream);
        require_bound(*batch);
        drop_graph(*batch);
        check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            run_all(*batch, s);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check_hip(hipStreamEndCapture(s, &batch->graph), "hipStreamEndCapture");
        check_hip(hipGraphInstantiate(&batch->exec, batch->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    });
}

int skb_batch_launch(skb_batch* batch, void* stream) {
    return guarded([&] {
        if (!batch) fail(SKB_INVALID_ARGUMENT, "null batch");
        if (!batch->exec) fail(SKB_LOGIC_ERROR, "skb_batch_launch before skb_batch_capture");
        check_hip(hipGraphLaunch(batch->exec, as_stream(stream)), "hipGraphLaunch");
    });
}

int skb_batch_autotune(skb_batch* batch, int reps, void* stream, char* report, size_t capacity) {
    return guarded([&] {
        if (!batch || reps < 1) fail(SKB_INVALID_ARGUMENT, "bad autotune arguments");
        require_bound(*batch);
        const hipStream_t s = as_stream(stream);
        hipEvent_t a = nullptr, e = nullptr;
        check_hip(hipEventCreate(&a), "hipEventCreate");
        check_hip(hipEventCreate(&e), "hipEventCreate");
        std::ostringstream js;
        js << "[";
        try {
            for (size_t ci = 0; ci < batch->classes.size(); ++ci) {
                auto& c = *batch->classes[ci];
                run_all(*batch, s);  // every intermediate holds valid data
                js << (ci ? "," : "") << "{\"class\":" << ci << ",\"nx\":" << c.w.nx << ",\"nz\":" << c.w.nz
                   << ",\"instances\":" << c.instances << ",\"stages\":[";
                for (int stage = 0; stage < 5; ++stage) {
                    const uint64_t items = stage < 4 ? skb_pruned_plan_items(c.pruned, stage)
                                                     : skb_vertical_plan_items(c.vertical);
                    std::vector<int> cand = {0};
                    for (uint64_t g : {256ull, 512ull, 1024ull, 2048ull, 4096ull})
                        if (g < items) cand.push_back(static_cast<int>(g));
                    const std::vector<int> prods = stage == 4 ? product_candidates(c) : std::vector<int>{c.products};
                    int best = 0, best_p = prods[0];
                    float best_ms = 1e30f;
                    for (int pr : prods) {
                        if (stage == 4) apply_products(c, pr);
                        for (int g : cand) {
                            apply_grid(c, stage, g);
                            run_stage(c, stage, s);  // warm
                            check_hip(hipEventRecord(a, s), "hipEventRecord");
                            for (int r = 0; r < reps; ++r) run_stage(c, stage, s);
                            check_hip(hipEventRecord(e, s), "hipEventRecord");
                            check_hip(hipEventSynchronize(e), "hipEventSynchronize");
                            float ms = 0.f;
                            check_hip(hipEventElapsedTime(&ms, a, e), "hipEventElapsedTime");
                            // anything but the default (one workgroup per item, the
                            // first product form) must win by 2% to replace it
                            const bool first = g == 0 && pr == prods[0];
                            if (ms < (first ? best_ms : 0.98f * best_ms)) {
                                best_ms = ms;
                                best = g;
                                best_p = pr;
                            }
                        }
                    }
                    if (stage == 4) apply_products(c, best_p);
                    apply_grid(c, stage, best);
                    js << (stage ? "," : "") << "{\"stage\":" << stage << ",\"items\":" << items
                       << ",\"grid\":" << best << ",\"us\":" << 1e3 * best_ms / reps;
                    if (stage == 4) js << ",\"products\":" << best_p << ",\"product_forms\":" << prods.size();
                    js << "}";
                }
                js << "]}";
            }
        } catch (...) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(e);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(e);
        js << "]";
        drop_graph(*batch);  // launch shapes changed: capture again
        const std::string out = js.str();
        if (report && capacity) {
            const size_t n = std::min(capacity - 1, out.size());
            std::memcpy(report, out.data(), n);
            report[n] = '\0';
        }
    });
}

int skb_batch_tuning(const skb_batch* batch, char* spec, size_t capacity) {
    return guarded([&] {
        if (!batch || !spec || capacity == 0) fail(SKB_INVALID_ARGUMENT, "null argument");
        const std::string out = tuning_spec(*batch);
        if (out.size() + 1 > capacity) fail(SKB_INVALID_ARGUMENT, "tuning spec buffer too small");
        std::memcpy(spec, out.c_str(), out.size() + 1);
    });
}

int skb_batch_set_tuning(skb_batch* batch, const char* spec) {
    return guarded([&] {
        if (!batch || !spec) fail(SKB_INVALID_ARGUMENT, "null argument");
        // parse everything first, so a bad spec changes nothing
        struct Choice {
            int cls, stage, value;  // stage 5 = products
        };
        std::vector<Choice> choices;
        const std::string text(spec);
        size_t at = 0;
        while (at < text.size()) {
            size_t end = text.find(';', at);
            if (end == std::string::npos) end = text.size();
            const std::string tok = text.substr(at, end - at);
            at = end + 1;
            if (tok.empty()) continue;
            int cls = -1, stage = -1, value = 0, used = 0;
            if (std::sscanf(tok.c_str(), "c%d.s%d=%d%n", &cls, &stage, &value, &used) == 3 &&
                used == static_cast<int>(tok.size())) {
                if (stage < 0 || stage > 4 || value < 0) fail(SKB_INVALID_ARGUMENT, "bad tuning entry: " + tok);
            } else if (std::sscanf(tok.c_str(), "c%d.p=%d%n", &cls, &value, &used) == 2 &&
                       used == static_cast<int>(tok.size())) {
                stage = 5;
                if (value != SKB_PRODUCTS_FP32 && value != SKB_PRODUCTS_SPLIT_BF16)
                    fail(SKB_INVALID_ARGUMENT, "bad tuning entry: " + tok);
            } else {
                fail(SKB_INVALID_ARGUMENT, "bad tuning entry: " + tok);
            }
            if (cls < 0 || cls >= static_cast<int>(batch->classes.size()))
                fail(SKB_INVALID_ARGUMENT, "tuning entry for a class the batch does not have: " + tok);
            choices.push_back({cls, stage, value});
        }
        for (const auto& ch : choices) {
            auto& c = *batch->classes[ch.cls];
            if (ch.stage == 5) apply_products(c, ch.value);
            else apply_grid(c, ch.stage, ch.value);
        }
        drop_graph(*batch);
    });
}

}  // extern "C"
