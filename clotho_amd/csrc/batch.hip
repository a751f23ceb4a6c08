// batch.hip — replicate batches: R independent populations advanced together.
//
// Each replicate is an ordinary context (own buffers, GenState and Philox
// streams keyed by its replicate number).  A batched generation launches every
// kernel of the generation once for all replicates (pipeline.hip,
// launch_batch_generation): the kernels read each replicate's Params and
// Buffers from a device array and walk per-kernel block-offset tables, so
// replicates of uneven size share one persistent grid.  All replicates of a
// batch advance in lockstep (same parent/child parity).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "pipeline.hpp"

using namespace clotho;

struct clotho_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<clotho_ctx*> reps;
    std::vector<RepView> host_views;  // last uploaded copy
    RepView* d_views = nullptr;
    uint32_t* d_offs = nullptr;
    uint32_t* d_live_offs = nullptr;  // [kLiveKinds][R + 1], written on the device each generation
    BatchLaunch L;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int cur = 0;
    Timing timing;
    std::string err;
};

namespace {

#define BATCH_TRY(b, expr)                                                   \
    do {                                                                     \
        const hipError_t e_ = (expr);                                        \
        if (e_ != hipSuccess) {                                              \
            (b)->err = std::string(#expr) + ": " + hipGetErrorString(e_);   \
            return CLOTHO_ERR_HIP;                                           \
        }                                                                    \
    } while (0)

void drop_graphs(clotho_batch* b) {
    for (hipGraphExec_t& g : b->exec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Upload each replicate's Params/Buffers and the per-kernel block offsets
// when they changed (a replicate's launch shape may have been retuned).
int sync_views(clotho_batch* b) {
    const uint32_t R = (uint32_t)b->reps.size();
    std::vector<RepView> v(R);
    for (uint32_t r = 0; r < R; ++r) {
        std::memset(&v[r], 0, sizeof(RepView));
        v[r].P = b->reps[r]->p;
        v[r].B = b->reps[r]->b;
        v[r].B.by_parent = nullptr;  // replicates recombine in child order
        // batched recombination blocks take kBatchChunkDiv times the rows of
        // a replicate's own launch: the replicates supply the parallelism,
        // and every block pays a fixed set-up (offset search, tables, partials)
        v[r].P.chunks = std::max<uint32_t>(1, (v[r].P.chunks + kBatchChunkDiv - 1) / kBatchChunkDiv);
    }
    if (b->d_views && b->host_views.size() == R &&
        std::memcmp(v.data(), b->host_views.data(), sizeof(RepView) * R) == 0)
        return CLOTHO_OK;
    drop_graphs(b);
    std::vector<uint32_t> offs((size_t)kOpCount * (R + 1), 0);
    for (uint32_t r = 0; r < R; ++r) {
        uint32_t blocks[kOpCount];
        batch_op_blocks(v[r].P, blocks);
        for (int k = 0; k < kOpCount; ++k)
            offs[(size_t)k * (R + 1) + r + 1] = offs[(size_t)k * (R + 1) + r] + blocks[k];
    }
    if (!b->d_views) {
        BATCH_TRY(b, hipMalloc((void**)&b->d_views, sizeof(RepView) * R));
        BATCH_TRY(b, hipMalloc((void**)&b->d_offs, sizeof(uint32_t) * offs.size()));
        BATCH_TRY(b, hipMalloc((void**)&b->d_live_offs, sizeof(uint32_t) * kLiveKinds * (R + 1)));
    }
    BATCH_TRY(b, hipMemcpy(b->d_views, v.data(), sizeof(RepView) * R, hipMemcpyHostToDevice));
    BATCH_TRY(b, hipMemcpy(b->d_offs, offs.data(), sizeof(uint32_t) * offs.size(),
                           hipMemcpyHostToDevice));
    b->host_views = v;
    BatchLaunch& L = b->L;
    L.views = b->d_views;
    L.offs = b->d_offs;
    L.live_offs = b->d_live_offs;
    L.R = R;
    for (int k = 0; k < kOpCount; ++k) L.total[k] = offs[(size_t)k * (R + 1) + R];
    L.traits = v[0].P.traits;
    L.wide_events = false;
    for (const RepView& x : v) L.wide_events = L.wide_events || x.P.ev_cap > 64;
    L.cus = (uint32_t)std::max(1, b->reps[0]->prop.multiProcessorCount);
    return CLOTHO_OK;
}

int capture_batch(clotho_batch* b) {
    for (int k = 0; k < 2; ++k) {
        if (b->exec[k]) continue;
        hipGraph_t g = nullptr;
        BATCH_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e1 = launch_batch_generation(b->L, (uint32_t)k, b->stream, nullptr);
        const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            b->err = std::string("batch capture: ") +
                     hipGetErrorString(e1 != hipSuccess ? e1 : e2);
            return CLOTHO_ERR_HIP;
        }
        const hipError_t e3 = hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) {
            b->err = std::string("batch instantiate: ") + hipGetErrorString(e3);
            return CLOTHO_ERR_HIP;
        }
    }
    return CLOTHO_OK;
}

}  // namespace

extern "C" {

int clotho_batch_create(const clotho_config* cfgs, uint32_t count, int device,
                        clotho_batch** out) {
    if (!cfgs || !out || count == 0) return CLOTHO_ERR_ARG;
    *out = nullptr;
    for (uint32_t r = 0; r < count; ++r)
        if (cfgs[r].trait_count != cfgs[0].trait_count || cfgs[r].compact_every)
            return CLOTHO_ERR_ARG;
    clotho_batch* b = new (std::nothrow) clotho_batch();
    if (!b) return CLOTHO_ERR_ARG;
    b->device = device;
    auto bail = [&](int rc) {
        clotho_batch_destroy(b);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(CLOTHO_ERR_HIP);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(CLOTHO_ERR_HIP);
    for (uint32_t r = 0; r < count; ++r) {
        clotho_ctx* c = nullptr;
        const int rc = clotho_create(&cfgs[r], device, &c);
        if (rc != CLOTHO_OK) return bail(rc);
        b->reps.push_back(c);
    }
    const int rc = sync_views(b);
    if (rc != CLOTHO_OK) return bail(rc);
    *out = b;
    return CLOTHO_OK;
}

void clotho_batch_destroy(clotho_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_offs) (void)hipFree(b->d_offs);
    if (b->d_live_offs) (void)hipFree(b->d_live_offs);
    for (clotho_ctx* c : b->reps) clotho_destroy(c);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* clotho_batch_last_error(const clotho_batch* b) { return b ? b->err.c_str() : ""; }

uint32_t clotho_batch_size(const clotho_batch* b) { return b ? (uint32_t)b->reps.size() : 0u; }

clotho_ctx* clotho_batch_replicate(clotho_batch* b, uint32_t r) {
    return (b && r < b->reps.size()) ? b->reps[r] : nullptr;
}

int clotho_batch_autotune(clotho_batch* b, uint32_t* percent) {
    if (!b) return CLOTHO_ERR_ARG;
    BATCH_TRY(b, hipSetDevice(b->device));
    for (clotho_ctx* c : b->reps) BATCH_TRY(b, hipStreamSynchronize(c->stream));
    int rc = sync_views(b);
    if (rc != CLOTHO_OK) return rc;
    const uint32_t par = (uint32_t)b->cur;
    const uint32_t nums[3] = {1, 1, 2}, dens[3] = {2, 1, 1};
    EventSet events;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BATCH_TRY(b, events.add(&e0));
    BATCH_TRY(b, events.add(&e1));
    float best[3] = {1e30f, 1e30f, 1e30f};
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < 3; ++i) {
            b->L.grid_scale_num = nums[i];
            b->L.grid_scale_den = dens[i];
            BATCH_TRY(b, launch_batch_replay(b->L, par, b->stream));  // warm-up
            BATCH_TRY(b, hipEventRecord(e0, b->stream));
            for (int r = 0; r < 3; ++r) BATCH_TRY(b, launch_batch_replay(b->L, par, b->stream));
            BATCH_TRY(b, hipEventRecord(e1, b->stream));
            BATCH_TRY(b, hipEventSynchronize(e1));
            float ms = 0.f;
            BATCH_TRY(b, hipEventElapsedTime(&ms, e0, e1));
            best[i] = std::min(best[i], ms);
        }
    int bi = 1;
    for (int i = 0; i < 3; ++i)
        if (best[i] < best[bi]) bi = i;
    b->L.grid_scale_num = nums[bi];
    b->L.grid_scale_den = dens[bi];
    drop_graphs(b);  // captured launches bake the grid in
    if (percent) *percent = 100u * nums[bi] / dens[bi];
    return CLOTHO_OK;
}

void clotho_batch_set_timing(clotho_batch* b, int enabled) {
    if (b) b->timing.enabled = enabled != 0;
}

int clotho_batch_timings(clotho_batch* b, double* ms, int cap) {
    if (!b || !ms) return CLOTHO_ERR_ARG;
    for (int i = 0; i < cap && i < 6; ++i) ms[i] = b->timing.ms[i];
    return b->timing.generations;
}

int clotho_batch_run(clotho_batch* b, uint32_t n, int mode, double* elapsed_ms) {
    if (!b) return CLOTHO_ERR_ARG;
    if (mode < CLOTHO_BATCH_GRAPH || mode > CLOTHO_BATCH_SEQUENTIAL) {
        b->err = "batch_run: unknown mode";
        return CLOTHO_ERR_ARG;
    }
    BATCH_TRY(b, hipSetDevice(b->device));
    for (clotho_ctx* c : b->reps) {
        if (c->cur != b->cur) {
            b->err = "batch_run: a replicate was advanced outside the batch";
            return CLOTHO_ERR_STATE;
        }
        if ((c->next_n && c->next_n != c->p.n) || c->p.n_par != c->p.n) {
            b->err = "batch_run: replicate population sizes are fixed inside a batch";
            return CLOTHO_ERR_ARG;
        }
        BATCH_TRY(b, hipStreamSynchronize(c->stream));  // uploads done
    }
    int rc = sync_views(b);
    if (rc != CLOTHO_OK) return rc;
    if (mode == CLOTHO_BATCH_GRAPH) {
        rc = capture_batch(b);
        if (rc != CLOTHO_OK) return rc;
    } else if (mode == CLOTHO_BATCH_SEQUENTIAL) {
        for (clotho_ctx* c : b->reps) {
            rc = capture_graphs(c);
            if (rc != CLOTHO_OK) {
                b->err = "replicate graph: " + c->err;
                return rc;
            }
        }
    }
    Timing& tm = b->timing;
    for (double& v : tm.ms) v = 0.0;
    tm.generations = 0;
    const bool phase_timing = tm.enabled && mode == CLOTHO_BATCH_EAGER;
    std::vector<std::vector<hipEvent_t>> marks;
    EventSet events;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BATCH_TRY(b, events.add(&e0));
    BATCH_TRY(b, events.add(&e1));
    BATCH_TRY(b, hipEventRecord(e0, b->stream));
    for (uint32_t g = 0; g < n; ++g) {
        const uint32_t par = (uint32_t)b->cur;
        if (mode == CLOTHO_BATCH_GRAPH) {
            BATCH_TRY(b, hipGraphLaunch(b->exec[par], b->stream));
        } else if (mode == CLOTHO_BATCH_EAGER) {
            hipEvent_t* mk = nullptr;
            if (phase_timing) {
                marks.emplace_back(6);
                for (hipEvent_t& e : marks.back()) BATCH_TRY(b, events.add(&e));
                mk = marks.back().data();
            }
            BATCH_TRY(b, launch_batch_generation(b->L, par, b->stream, mk));
        } else {
            for (clotho_ctx* c : b->reps) BATCH_TRY(b, hipGraphLaunch(c->graphs.exec[par], b->stream));
        }
        b->cur = 1 - b->cur;
        for (clotho_ctx* c : b->reps) c->cur = b->cur;
    }
    BATCH_TRY(b, hipEventRecord(e1, b->stream));
    BATCH_TRY(b, hipEventSynchronize(e1));
    float ms = 0.f;
    BATCH_TRY(b, hipEventElapsedTime(&ms, e0, e1));
    if (elapsed_ms) *elapsed_ms = ms;
    if (phase_timing) {
        for (auto& mk : marks) {
            float t = 0.f;
            BATCH_TRY(b, hipEventElapsedTime(&t, mk[0], mk[5]));
            tm.ms[0] += t;
            for (int q = 0; q < 5; ++q) {
                BATCH_TRY(b, hipEventElapsedTime(&t, mk[q], mk[q + 1]));
                tm.ms[1 + q] += t;
            }
        }
        tm.generations = (int)n;
    }
    for (uint32_t r = 0; r < b->reps.size(); ++r) {
        clotho_ctx* c = b->reps[r];
        GenState st;
        rc = read_state(c, &st);
        if (rc != CLOTHO_OK) {
            b->err = c->err;
            return rc;
        }
        c->has_population = true;
        if (st.error) {
            b->err = "replicate " + std::to_string(r) + ": allele capacity exceeded at generation " +
                     std::to_string(st.error_gen);
            return CLOTHO_ERR_CAPACITY;
        }
    }
    return CLOTHO_OK;
}

}  // extern "C"
