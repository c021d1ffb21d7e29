// Host runtime and C ABI of the gfx950 hexahedral Q1 conduction solver.
//
// An operator handle owns the device copies of one prepared operator (the
// reference's ThermalSystem arrays, cpp/src/thermal/thermal_system.cpp:252-363,
// in both precisions) and a stream.  A batch owns the per-right-hand-side
// vectors and scalar states of any number of slots and runs the whole MPIR
// iteration on the device: a captured HIP graph of several steps (outer
// kernel, inner apply, inner update) is replayed until a device counter of
// slots still iterating reaches zero, so the host waits once per graph, not
// once per inner product as the reference's device paths do
// (cpp/cuda/inner_pcg.cu:522-528, runtime.py:156-160).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pcbhip.h"
#include "q1hex_device.h"

using namespace pcbhip;

namespace {

thread_local std::string g_last_error;

struct Failure {
    int code;
    std::string message;
};

[[noreturn]] void fail(int code, const std::string& message) { throw Failure{code, message}; }

void check(hipError_t err, const char* what) {
    if (err != hipSuccess) fail(PCBHIP_RUNTIME, std::string("HIP ") + what + ": " + hipGetErrorString(err));
}

template <class F>
int guarded(F&& body) {
    try {
        body();
        g_last_error.clear();
        return PCBHIP_OK;
    } catch (const Failure& f) {
        g_last_error = f.message;
        return f.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return PCBHIP_NO_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PCBHIP_RUNTIME;
    }
}

void* device_alloc(size_t bytes) {
    void* p = nullptr;
    const hipError_t err = hipMalloc(&p, bytes == 0 ? 256 : bytes);
    if (err == hipErrorOutOfMemory) fail(PCBHIP_NO_MEMORY, "out of device memory");
    check(err, "hipMalloc");
    return p;
}

size_t align256(size_t bytes) { return (bytes + 255) & ~size_t(255); }

int strips_x(int C) { return (C + 1 + kStripCols - 1) / kStripCols; }
int groups_y(int R, int rows) { return (R + 1 + rows * kWavesPerGroup - 1) / (rows * kWavesPerGroup); }

// Planes per z block: the whole depth when the launch already has
// ``target`` workgroups, else split into the blocks needed to reach it
// (each block re-reads one halo plane on each side, so never fewer than 2
// planes).
int planes_per_block(int S, long long groups_xy, long long target) {
    const int planes = S + 1;
    if (groups_xy >= target) return planes;
    const long long split = (target + groups_xy - 1) / groups_xy;
    int zb = (int)((planes + split - 1) / split);
    zb += zb & 1;  // whole plane pairs: z blocks never cut a reduction cell
    return zb < 2 ? 2 : (zb > planes ? planes : zb);
}

}  // namespace

struct pcbhip_operator {
    int device = 0;
    int S = 0, R = 0, C = 0;
    long long N = 0, E = 0;
    hipStream_t stream = nullptr;
    void* block = nullptr;  // one allocation for every array below
    float* coef32 = nullptr;
    double* coef64 = nullptr;
    float* robin32 = nullptr;
    double* robin64 = nullptr;
    float* dinv32 = nullptr;
    double* dinv64 = nullptr;
    double* diag64 = nullptr;
    int rows32 = 4, zb32 = 0, rows64 = 4, zb64 = 0;  // standalone apply tile shapes

    OperatorDev view() const {
        OperatorDev v{};
        v.S = S;
        v.R = R;
        v.C = C;
        v.N = N;
        v.coef32 = coef32;
        v.coef64 = coef64;
        v.robin32 = robin32;
        v.robin64 = robin64;
        v.dinv32 = dinv32;
        v.dinv64 = dinv64;
        return v;
    }
    ~pcbhip_operator() {
        if (block) (void)hipFree(block);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

struct SlotHost {
    const pcbhip_operator* op = nullptr;
    void* block = nullptr;
    double* b = nullptr;
    double* x[2] = {nullptr, nullptr};
    double* x0 = nullptr;
    void* r = nullptr;
    void* p[2] = {nullptr, nullptr};
    void* q = nullptr;
    void* c = nullptr;
    double* part = nullptr;
    double* history = nullptr;
    int history_rows = 0;
    pcbhip_mpir_config config{1e-10, 0.0, 2e-3, 8, 200};
    int tiles_outer = 0, tiles_inner = 0, tiles_flat = 0, cells = 0, part_stride = 0;
};

}  // namespace

struct pcbhip_batch {
    int device = 0;
    int inner64 = 0;
    hipStream_t stream = nullptr;
    std::vector<SlotHost> slots;
    SlotDev* d_slots = nullptr;
    SlotState* d_states = nullptr;
    SlotConfig* d_configs = nullptr;
    Tile* d_outer = nullptr;
    Tile* d_inner = nullptr;
    Tile* d_flat = nullptr;
    int n_outer = 0, n_inner = 0, n_flat = 0;
    int rows_inner = 4;
    int outer_every = 8;
    int tiled_slots = 0;  // slots whose tiles are in the tables
    long long target_groups = kTargetGroups;
    int* d_active = nullptr;
    int* h_active = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
    bool configs_dirty = true;

    ~pcbhip_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (SlotHost& s : slots) {
            if (s.block) (void)hipFree(s.block);
            if (s.history) (void)hipFree(s.history);
        }
        for (void* p : {(void*)d_slots, (void*)d_states, (void*)d_configs, (void*)d_outer,
                        (void*)d_inner, (void*)d_flat, (void*)d_active})
            if (p) (void)hipFree(p);
        if (h_active) (void)hipHostFree(h_active);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void set_device(int device) { check(hipSetDevice(device), "hipSetDevice"); }

void check_config(const pcbhip_mpir_config& c) {
    // MPIRConfig.__post_init__ (src/electrical/matrix_free_mpir_fem/solver.py:63-79).
    for (double v : {c.relative_tolerance, c.absolute_tolerance, c.inner_relative_tolerance})
        if (!std::isfinite(v) || v < 0.0) fail(PCBHIP_INVALID_INPUT, "tolerances must be finite and non-negative");
    if (c.relative_tolerance == 0.0 && c.absolute_tolerance == 0.0)
        fail(PCBHIP_INVALID_INPUT, "at least one outer tolerance must be positive");
    if (!(c.inner_relative_tolerance > 0.0 && c.inner_relative_tolerance < 1.0))
        fail(PCBHIP_INVALID_INPUT, "inner_relative_tolerance must be between zero and one");
    if (c.max_outer_iterations < 1 || c.max_inner_iterations < 1)
        fail(PCBHIP_INVALID_INPUT, "iteration limits must be positive");
}

hipStream_t stream_of(const pcbhip_operator* op, void* stream) {
    return stream ? static_cast<hipStream_t>(stream) : op->stream;
}

template <typename T>
void apply_dev(const pcbhip_operator* op, const T* x, T* y, int mode, hipStream_t stream) {
    set_device(op->device);
    const T* coef;
    const T* robin;
    if constexpr (sizeof(T) == 4) {
        coef = op->coef32;
        robin = op->robin32;
    } else {
        coef = op->coef64;
        robin = op->robin64;
    }
    const int rows = sizeof(T) == 4 ? op->rows32 : op->rows64;
    const int zb = sizeof(T) == 4 ? op->zb32 : op->zb64;
    check(launch::apply<T>(coef, robin, x, y, op->S, op->R, op->C, mode, rows, zb, stream), "apply launch");
}

template <typename T>
void apply_host(const pcbhip_operator* op, const T* x, T* y, int mode) {
    if (!op || !x || !y) fail(PCBHIP_INVALID_INPUT, "null argument");
    set_device(op->device);
    const size_t bytes = sizeof(T) * (size_t)op->N;
    T* dx = static_cast<T*>(device_alloc(2 * align256(bytes)));
    T* dy = reinterpret_cast<T*>(reinterpret_cast<char*>(dx) + align256(bytes));
    try {
        check(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, op->stream), "copy x");
        apply_dev<T>(op, dx, dy, mode, op->stream);
        check(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, op->stream), "copy y");
        check(hipStreamSynchronize(op->stream), "apply");
    } catch (...) {
        (void)hipFree(dx);
        throw;
    }
    check(hipFree(dx), "hipFree");
}

// Tiles of one slot: stencil workgroups in (z block, y group, x strip)
// order, so a contiguous range of tiles is a compact piece of the board.
int add_tiles(std::vector<Tile>& out, int slot, int nsx, int nsy, int S, int zb) {
    const int nsz = (S + 1 + zb - 1) / zb;
    int index = 0;
    for (int sz = 0; sz < nsz; ++sz)
        for (int sy = 0; sy < nsy; ++sy)
            for (int sx = 0; sx < nsx; ++sx)
                out.push_back(Tile{slot, index++, (short)sx, (short)sy, (short)sz, (short)zb});
    return index;
}

void upload_tiles(pcbhip_batch* b, const std::vector<Tile>& t, Tile** d, int* n) {
    if (*d) check(hipFree(*d), "hipFree");
    *n = (int)t.size();
    *d = static_cast<Tile*>(device_alloc(sizeof(Tile) * t.size()));
    check(hipMemcpyAsync(*d, t.data(), sizeof(Tile) * t.size(), hipMemcpyHostToDevice, b->stream), "upload");
}

// (Re)builds the three tile tables for the batch's current shape.  The
// z split is chosen over the whole batch: many slots fill the device without
// splitting any of them.
void build_tiles(pcbhip_batch* b, const std::vector<char>* keep = nullptr) {
    const int rows_inner = b->inner64 ? kRowsInner64 : b->rows_inner;
    long long xy_outer = 0, xy_inner = 0;
    for (size_t i = 0; i < b->slots.size(); ++i) {
        if (keep && !(*keep)[i]) continue;
        const SlotHost& s = b->slots[i];
        xy_outer += (long long)strips_x(s.op->C) * groups_y(s.op->R, kRowsOuter);
        xy_inner += (long long)strips_x(s.op->C) * groups_y(s.op->R, rows_inner);
    }
    std::vector<Tile> outer, inner, flat;
    for (int i = 0; i < (int)b->slots.size(); ++i) {
        SlotHost& s = b->slots[(size_t)i];
        const pcbhip_operator* op = s.op;
        const int nsx = strips_x(op->C);
        const int nyo = groups_y(op->R, kRowsOuter), nyi = groups_y(op->R, rows_inner);
        const long long share_o = xy_outer > 0 ? (long long)nsx * nyo * b->target_groups / xy_outer : 0;
        const long long share_i = xy_inner > 0 ? (long long)nsx * nyi * b->target_groups / xy_inner : 0;
        const int zbo = planes_per_block(op->S, (long long)nsx * nyo, std::max(share_o, (long long)nsx * nyo));
        const int zbi = planes_per_block(op->S, (long long)nsx * nyi, std::max(share_i, (long long)nsx * nyi));
        s.tiles_flat = (int)((op->N + kFlatChunk - 1) / kFlatChunk);
        if (!keep || (*keep)[(size_t)i]) {
            // A finished slot's tiles leave the tables, so its early exits
            // stop costing launch slots; its partial layout does not change.
            s.tiles_outer = add_tiles(outer, i, nsx, nyo, op->S, zbo);
            s.tiles_inner = add_tiles(inner, i, nsx, nyi, op->S, zbi);
            for (int k = 0; k < s.tiles_flat; ++k) flat.push_back(Tile{i, k, 0, 0, 0, 0});
        }
        s.cells = cell_count(op->S, op->R, op->C);
        if (std::max(s.cells, s.tiles_flat) > s.part_stride)
            fail(PCBHIP_RUNTIME, "partial count exceeds the partial-sum area");
    }
    b->tiled_slots = 0;
    for (size_t i = 0; i < b->slots.size(); ++i) b->tiled_slots += (!keep || (*keep)[i]) ? 1 : 0;
    if (outer.empty()) outer.push_back(Tile{0, 0, 0, 0, 0, 1});
    if (inner.empty()) inner.push_back(Tile{0, 0, 0, 0, 0, 1});
    if (flat.empty()) flat.push_back(Tile{0, 0, 0, 0, 0, 0});
    upload_tiles(b, outer, &b->d_outer, &b->n_outer);
    upload_tiles(b, inner, &b->d_inner, &b->n_inner);
    upload_tiles(b, flat, &b->d_flat, &b->n_flat);
    b->configs_dirty = true;
    if (b->exec) {
        check(hipGraphExecDestroy(b->exec), "graph exec destroy");
        b->exec = nullptr;
    }
}

void build_graph(pcbhip_batch* b, int steps) {
    if (b->exec) {
        check(hipGraphExecDestroy(b->exec), "graph exec destroy");
        b->exec = nullptr;
    }
    if (b->graph) {
        check(hipGraphDestroy(b->graph), "graph destroy");
        b->graph = nullptr;
    }
    check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    const int n_slots = (int)b->slots.size();
    for (int s = 0; s < steps; ++s) {
        hipError_t err = hipSuccess;
        // The outer step runs every outer_every inner iterations: a slot whose
        // inner solve ended waits at most that long, and the idle outer
        // kernels stay out of most steps.
        if (s % b->outer_every == 0)
            err = launch::outer_step(b->inner64 != 0, n_slots, b->d_slots, b->d_states, b->d_configs, b->d_outer,
                                     b->n_outer, b->d_active, b->stream);
        if (err == hipSuccess)
            err = launch::inner_step(b->inner64 != 0, b->rows_inner, n_slots, b->d_slots, b->d_states, b->d_configs,
                                     b->d_inner, b->n_inner, b->d_flat, b->n_flat, b->d_active, b->stream);
        if (err != hipSuccess) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(b->stream, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
            check(err, "capture launch");
        }
    }
    check(hipMemcpyAsync(b->h_active, b->d_active, sizeof(int), hipMemcpyDeviceToHost, b->stream), "capture copy");
    check(hipStreamEndCapture(b->stream, &b->graph), "end capture");
    check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
    b->graph_steps = steps;
}

}  // namespace

extern "C" {

const char* pcbhip_last_error(void) { return g_last_error.c_str(); }

int pcbhip_device_count(int* count) {
    return guarded([&] {
        if (!count) fail(PCBHIP_INVALID_INPUT, "null argument");
        int n = 0;
        const hipError_t err = hipGetDeviceCount(&n);
        *count = err == hipSuccess ? n : 0;
    });
}

int pcbhip_operator_create(int device, int slabs, int rows, int cols, const double* coef, const double* robin,
                           const uint8_t* free_nodes, pcbhip_operator** out) {
    return guarded([&] {
        if (!out || !coef || !robin || !free_nodes) fail(PCBHIP_INVALID_INPUT, "null argument");
        *out = nullptr;
        if (slabs < 1 || rows < 1 || cols < 1) fail(PCBHIP_INVALID_INPUT, "element grid must be positive");
        const long long E = (long long)slabs * rows * cols;
        const long long N = (long long)(slabs + 1) * (rows + 1) * (cols + 1);
        if (N > (1LL << 31) - 1) fail(PCBHIP_INVALID_INPUT, "node count exceeds 2^31 - 1");
        std::vector<double> coef64(3 * (size_t)E), robin64((size_t)N);
        std::vector<float> coef32(3 * (size_t)E), robin32((size_t)N);
        for (long long i = 0; i < 3 * E; ++i) {
            if (!std::isfinite(coef[i]) || coef[i] < 0.0)
                fail(PCBHIP_INVALID_INPUT, "element coefficients must be finite and non-negative");
            coef64[i] = coef[i] / 36.0;
            coef32[i] = (float)coef64[i];
        }
        for (long long n = 0; n < N; ++n) {
            if (free_nodes[n]) {
                if (!std::isfinite(robin[n]) || robin[n] < 0.0)
                    fail(PCBHIP_INVALID_INPUT, "robin conductance must be finite and non-negative");
                robin64[n] = robin[n];
            } else {
                robin64[n] = -1.0;
            }
            robin32[n] = (float)robin64[n];
        }
        set_device(device);
        auto op = std::make_unique<pcbhip_operator>();
        op->device = device;
        op->S = slabs;
        op->R = rows;
        op->C = cols;
        op->N = N;
        op->E = E;
        check(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking), "stream create");
        op->zb32 = planes_per_block(slabs, (long long)strips_x(cols) * groups_y(rows, op->rows32), kTargetGroups);
        op->zb64 = planes_per_block(slabs, (long long)strips_x(cols) * groups_y(rows, op->rows64), kTargetGroups);
        const size_t s_c32 = align256(sizeof(float) * 3 * E), s_c64 = align256(sizeof(double) * 3 * E);
        const size_t s_n32 = align256(sizeof(float) * N), s_n64 = align256(sizeof(double) * N);
        char* base = static_cast<char*>(device_alloc(s_c32 + s_c64 + 2 * s_n32 + 3 * s_n64 + 256));
        op->block = base;
        op->coef32 = reinterpret_cast<float*>(base);
        base += s_c32;
        op->coef64 = reinterpret_cast<double*>(base);
        base += s_c64;
        op->robin32 = reinterpret_cast<float*>(base);
        base += s_n32;
        op->dinv32 = reinterpret_cast<float*>(base);
        base += s_n32;
        op->robin64 = reinterpret_cast<double*>(base);
        base += s_n64;
        op->dinv64 = reinterpret_cast<double*>(base);
        base += s_n64;
        op->diag64 = reinterpret_cast<double*>(base);
        base += s_n64;
        int* d_bad = reinterpret_cast<int*>(base);
        check(hipMemcpyAsync(op->coef32, coef32.data(), sizeof(float) * 3 * E, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->coef64, coef64.data(), sizeof(double) * 3 * E, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->robin32, robin32.data(), sizeof(float) * N, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->robin64, robin64.data(), sizeof(double) * N, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemsetAsync(d_bad, 0, sizeof(int), op->stream), "memset");
        check(launch::diagonal(op->coef64, op->robin64, op->diag64, op->dinv64, op->dinv32, slabs, rows, cols, d_bad,
                               op->stream),
              "diagonal launch");
        int bad = 0;
        check(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, op->stream), "copy");
        check(hipStreamSynchronize(op->stream), "operator create");
        if (bad) fail(PCBHIP_INVALID_INPUT, "every free node must have positive conductance");
        *out = op.release();
    });
}

int pcbhip_operator_destroy(pcbhip_operator* op) {
    return guarded([&] {
        if (op) {
            set_device(op->device);
            delete op;
        }
    });
}

int64_t pcbhip_operator_size(const pcbhip_operator* op) { return op ? op->N : 0; }

int pcbhip_operator_diagonal(const pcbhip_operator* op, double* diag) {
    return guarded([&] {
        if (!op || !diag) fail(PCBHIP_INVALID_INPUT, "null argument");
        set_device(op->device);
        check(hipMemcpyAsync(diag, op->diag64, sizeof(double) * op->N, hipMemcpyDeviceToHost, op->stream), "copy");
        check(hipStreamSynchronize(op->stream), "diagonal");
    });
}

int pcbhip_apply_f32(const pcbhip_operator* op, const float* x, float* y) {
    return guarded([&] { apply_host<float>(op, x, y, 0); });
}
int pcbhip_apply_f64(const pcbhip_operator* op, const double* x, double* y) {
    return guarded([&] { apply_host<double>(op, x, y, 0); });
}
int pcbhip_stiffness_f64(const pcbhip_operator* op, const double* x, double* y) {
    return guarded([&] { apply_host<double>(op, x, y, 1); });
}
int pcbhip_apply_f32_dev(const pcbhip_operator* op, const float* x, float* y, void* stream) {
    return guarded([&] {
        if (!op || !x || !y) fail(PCBHIP_INVALID_INPUT, "null argument");
        apply_dev<float>(op, x, y, 0, stream_of(op, stream));
    });
}
int pcbhip_apply_f64_dev(const pcbhip_operator* op, const double* x, double* y, void* stream) {
    return guarded([&] {
        if (!op || !x || !y) fail(PCBHIP_INVALID_INPUT, "null argument");
        apply_dev<double>(op, x, y, 0, stream_of(op, stream));
    });
}
int pcbhip_stiffness_f64_dev(const pcbhip_operator* op, const double* x, double* y, void* stream) {
    return guarded([&] {
        if (!op || !x || !y) fail(PCBHIP_INVALID_INPUT, "null argument");
        apply_dev<double>(op, x, y, 1, stream_of(op, stream));
    });
}

int pcbhip_batch_create(int count, const pcbhip_operator* const* ops, int inner_fp64, pcbhip_batch** out) {
    return guarded([&] {
        if (!out || !ops || count < 1) fail(PCBHIP_INVALID_INPUT, "a batch needs at least one operator");
        *out = nullptr;
        const int device = ops[0]->device;
        for (int i = 0; i < count; ++i) {
            if (!ops[i]) fail(PCBHIP_INVALID_INPUT, "null operator");
            if (ops[i]->device != device) fail(PCBHIP_INVALID_INPUT, "a batch lives on one device");
        }
        set_device(device);
        auto b = std::make_unique<pcbhip_batch>();
        b->device = device;
        b->inner64 = inner_fp64 ? 1 : 0;
        check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream create");
        const size_t ti = b->inner64 ? sizeof(double) : sizeof(float);
        b->slots.resize((size_t)count);
        for (int i = 0; i < count; ++i) {
            SlotHost& s = b->slots[(size_t)i];
            const pcbhip_operator* op = ops[i];
            s.op = op;
            // One partial per reduction cell of the stencils or per chunk of
            // the pointwise kernel, whichever is more.
            const long long flat = (op->N + kFlatChunk - 1) / kFlatChunk;
            s.part_stride = (int)std::max((long long)cell_count(op->S, op->R, op->C), flat);
            s.history_rows = s.config.max_outer_iterations;
            // Vectors of the slot in one allocation.  History room is
            // allocated on demand in pcbhip_batch_set.
            const size_t n64 = align256(sizeof(double) * op->N), nti = align256(ti * op->N);
            const size_t part = align256(sizeof(double) * 4 * (size_t)s.part_stride);
            char* base = static_cast<char*>(device_alloc(4 * n64 + 5 * nti + part));
            s.block = base;
            s.b = reinterpret_cast<double*>(base);
            s.x[0] = reinterpret_cast<double*>(base + n64);
            s.x[1] = reinterpret_cast<double*>(base + 2 * n64);
            s.x0 = reinterpret_cast<double*>(base + 3 * n64);
            base += 4 * n64;
            s.r = base;
            s.p[0] = base + nti;
            s.p[1] = base + 2 * nti;
            s.q = base + 3 * nti;
            s.c = base + 4 * nti;
            base += 5 * nti;
            s.part = reinterpret_cast<double*>(base);
            // Zeroed vectors: the kernels load p, c and x unconditionally and
            // weight them by 0 where the algorithm starts from zero.
            check(hipMemsetAsync(s.b, 0, 4 * n64 + 5 * nti, b->stream), "memset");
        }
        build_tiles(b.get());
        b->d_slots = static_cast<SlotDev*>(device_alloc(sizeof(SlotDev) * count));
        b->d_states = static_cast<SlotState*>(device_alloc(sizeof(SlotState) * count));
        b->d_configs = static_cast<SlotConfig*>(device_alloc(sizeof(SlotConfig) * count));
        b->d_active = static_cast<int*>(device_alloc(sizeof(int)));
        check(hipHostMalloc(reinterpret_cast<void**>(&b->h_active), sizeof(int), hipHostMallocDefault),
              "pinned alloc");
        check(hipStreamSynchronize(b->stream), "batch create");
        *out = b.release();
    });
}

int pcbhip_batch_destroy(pcbhip_batch* batch) {
    return guarded([&] {
        if (batch) {
            set_device(batch->device);
            delete batch;
        }
    });
}

int pcbhip_batch_set(pcbhip_batch* b, int slot, const double* rhs, const double* x0,
                     const pcbhip_mpir_config* config) {
    return guarded([&] {
        if (!b || slot < 0 || slot >= (int)b->slots.size()) fail(PCBHIP_INVALID_INPUT, "slot out of range");
        set_device(b->device);
        SlotHost& s = b->slots[(size_t)slot];
        const size_t bytes = sizeof(double) * (size_t)s.op->N;
        if (config) {
            check_config(*config);
            s.config = *config;
        }
        if (rhs) {
            for (long long i = 0; i < s.op->N; ++i)
                if (!std::isfinite(rhs[i])) fail(PCBHIP_INVALID_INPUT, "rhs must contain only finite values");
            check(hipMemcpyAsync(s.b, rhs, bytes, hipMemcpyHostToDevice, b->stream), "upload rhs");
        }
        if (x0) {
            check(hipMemcpyAsync(s.x0, x0, bytes, hipMemcpyHostToDevice, b->stream), "upload x0");
        } else {
            check(hipMemsetAsync(s.x0, 0, bytes, b->stream), "memset");
        }
        if (s.history_rows < s.config.max_outer_iterations || !s.history) {
            if (s.history) check(hipFree(s.history), "hipFree");
            s.history_rows = s.config.max_outer_iterations;
            s.history = static_cast<double*>(device_alloc(sizeof(double) * 4 * (size_t)s.history_rows));
        }
        check(hipStreamSynchronize(b->stream), "batch set");
        b->configs_dirty = true;
    });
}

int pcbhip_batch_set_shape(pcbhip_batch* b, int rows_inner, int target_groups) {
    return guarded([&] {
        if (!b) fail(PCBHIP_INVALID_INPUT, "null batch");
        if (rows_inner != 4 && rows_inner != 8) fail(PCBHIP_INVALID_INPUT, "rows_inner must be 4 or 8");
        if (target_groups < 1) fail(PCBHIP_INVALID_INPUT, "target_groups must be positive");
        set_device(b->device);
        b->rows_inner = rows_inner;
        b->target_groups = target_groups;
        build_tiles(b);
        check(hipStreamSynchronize(b->stream), "set shape");
    });
}

int pcbhip_operator_set_shape(pcbhip_operator* op, int rows32, int zb32, int rows64, int zb64) {
    return guarded([&] {
        if (!op) fail(PCBHIP_INVALID_INPUT, "null operator");
        if ((rows32 != 4 && rows32 != 8) || (rows64 != 2 && rows64 != 4))
            fail(PCBHIP_INVALID_INPUT, "rows must be 4 or 8 (FP32) and 2 or 4 (FP64)");
        if (zb32 < 1 || zb64 < 1) fail(PCBHIP_INVALID_INPUT, "planes per z block must be positive");
        op->rows32 = rows32;
        op->zb32 = std::min(zb32 + (zb32 & 1), op->S + 1);
        op->rows64 = rows64;
        op->zb64 = std::min(zb64 + (zb64 & 1), op->S + 1);
    });
}

int pcbhip_batch_rhs_dev(pcbhip_batch* b, int slot, double** rhs_dev) {
    return guarded([&] {
        if (!b || !rhs_dev || slot < 0 || slot >= (int)b->slots.size())
            fail(PCBHIP_INVALID_INPUT, "slot out of range");
        *rhs_dev = b->slots[(size_t)slot].b;
    });
}

int pcbhip_batch_run(pcbhip_batch* b) {
    return guarded([&] {
        if (!b) fail(PCBHIP_INVALID_INPUT, "null batch");
        set_device(b->device);
        const int count = (int)b->slots.size();
        if (b->configs_dirty) {
            std::vector<SlotDev> devs((size_t)count);
            std::vector<SlotConfig> cfgs((size_t)count);
            for (int i = 0; i < count; ++i) {
                SlotHost& s = b->slots[(size_t)i];
                if (!s.history) {
                    s.history_rows = s.config.max_outer_iterations;
                    s.history = static_cast<double*>(device_alloc(sizeof(double) * 4 * (size_t)s.history_rows));
                }
                SlotDev& d = devs[(size_t)i];
                d.op = s.op->view();
                d.b64 = s.b;
                d.x64[0] = s.x[0];
                d.x64[1] = s.x[1];
                d.r = s.r;
                d.p[0] = s.p[0];
                d.p[1] = s.p[1];
                d.q = s.q;
                d.c = s.c;
                d.part = s.part;
                d.history = s.history;
                d.cells = s.cells;
                d.tiles_flat = s.tiles_flat;
                d.part_stride = s.part_stride;
                cfgs[(size_t)i] = SlotConfig{s.config.relative_tolerance, s.config.absolute_tolerance,
                                             s.config.inner_relative_tolerance, s.config.max_outer_iterations,
                                             s.config.max_inner_iterations};
            }
            check(hipMemcpyAsync(b->d_slots, devs.data(), sizeof(SlotDev) * count, hipMemcpyHostToDevice,
                                 b->stream),
                  "upload slots");
            check(hipMemcpyAsync(b->d_configs, cfgs.data(), sizeof(SlotConfig) * count, hipMemcpyHostToDevice,
                                 b->stream),
                  "upload configs");
            b->configs_dirty = false;
        }
        // Reset: x = x0 in buffer 0, fresh scalar states, every slot active.
        std::vector<SlotState> states((size_t)count);
        long long cap = 0;
        for (int i = 0; i < count; ++i) {
            SlotHost& s = b->slots[(size_t)i];
            check(hipMemcpyAsync(s.x[0], s.x0, sizeof(double) * s.op->N, hipMemcpyDeviceToDevice, b->stream),
                  "reset x");
            std::memset(&states[(size_t)i], 0, sizeof(SlotState));
            states[(size_t)i].phase = kPhaseOuter;
            cap = std::max(cap, (long long)(s.config.max_outer_iterations + 1) *
                                    (s.config.max_inner_iterations + b->outer_every) + 2 * b->outer_every);
        }
        check(hipMemcpyAsync(b->d_states, states.data(), sizeof(SlotState) * count, hipMemcpyHostToDevice,
                             b->stream),
              "upload states");
        check(hipMemcpyAsync(b->d_active, &count, sizeof(int), hipMemcpyHostToDevice, b->stream), "upload active");
        check(hipStreamSynchronize(b->stream), "reset");
        // Graph length: enough steps that the wait per replay is a small
        // share of the work, few enough that steps launched after the last
        // slot finished cost little.
        const int steps = 32;
        if (!b->exec || b->graph_steps != steps) build_graph(b, steps);
        long long done_steps = 0;
        std::vector<SlotState> host_states;
        for (;;) {
            check(hipGraphLaunch(b->exec, b->stream), "graph launch");
            check(hipStreamSynchronize(b->stream), "mpir steps");
            done_steps += steps;
            const int active = *b->h_active;
            if (active <= 0) break;
            if (done_steps > cap) fail(PCBHIP_RUNTIME, "MPIR iteration did not terminate within its limits");
            if (count > 1 && active * 2 <= b->tiled_slots) {
                // Half the tiled slots have finished: drop their tiles.
                host_states.resize((size_t)count);
                check(hipMemcpy(host_states.data(), b->d_states, sizeof(SlotState) * count, hipMemcpyDeviceToHost),
                      "copy states");
                std::vector<char> keep((size_t)count);
                for (int i = 0; i < count; ++i) keep[(size_t)i] = host_states[(size_t)i].phase != kPhaseDone;
                build_tiles(b, &keep);
                build_graph(b, steps);
            }
        }
        if (b->tiled_slots != count) build_tiles(b);
    });
}

int pcbhip_batch_result(pcbhip_batch* b, int slot, double* x, pcbhip_mpir_result* result, double* history) {
    return guarded([&] {
        if (!b || slot < 0 || slot >= (int)b->slots.size()) fail(PCBHIP_INVALID_INPUT, "slot out of range");
        set_device(b->device);
        SlotHost& s = b->slots[(size_t)slot];
        SlotState st{};
        check(hipMemcpyAsync(&st, b->d_states + slot, sizeof(SlotState), hipMemcpyDeviceToHost, b->stream), "copy");
        check(hipStreamSynchronize(b->stream), "result");
        if (x) {
            check(hipMemcpyAsync(x, s.x[st.xcur], sizeof(double) * s.op->N, hipMemcpyDeviceToHost, b->stream),
                  "copy x");
        }
        if (history && st.outer > 0) {
            check(hipMemcpyAsync(history, s.history, sizeof(double) * 4 * (size_t)st.outer, hipMemcpyDeviceToHost,
                                 b->stream),
                  "copy history");
        }
        check(hipStreamSynchronize(b->stream), "result");
        if (result) {
            result->converged = st.converged;
            result->outer_iterations = st.outer;
            result->inner_iterations = st.total_inner;
            result->high_operator_applications = st.high_apps;
            result->low_operator_applications = st.low_apps;
            result->not_spd = st.status == kStatusNotSpd ? 1 : 0;
            result->relative_residual = st.rel_outer;
        }
        if (st.status == kStatusNotSpd)
            fail(PCBHIP_RUNTIME, "inner PCG requires a finite symmetric positive-definite operator");
    });
}

int pcbhip_solve_mpir(const pcbhip_operator* op, const double* rhs, double* x, const pcbhip_mpir_config* config,
                      int inner_fp64, pcbhip_mpir_result* result, double* history) {
    pcbhip_batch* b = nullptr;
    int rc = pcbhip_batch_create(1, &op, inner_fp64, &b);
    if (rc != PCBHIP_OK) return rc;
    rc = pcbhip_batch_set(b, 0, rhs, x, config);
    if (rc == PCBHIP_OK) rc = pcbhip_batch_run(b);
    if (rc == PCBHIP_OK) rc = pcbhip_batch_result(b, 0, x, result, history);
    const std::string message = g_last_error;
    (void)pcbhip_batch_destroy(b);
    g_last_error = message;
    return rc;
}

int pcbhip_copy_dev(const float* src, float* dst, int64_t n_floats, void* stream) {
    return guarded([&] {
        if (!src || !dst || n_floats % 4 != 0) fail(PCBHIP_INVALID_INPUT, "copy needs a multiple of 4 floats");
        check(launch::copy(src, dst, n_floats, static_cast<hipStream_t>(stream)), "copy launch");
    });
}

}  // extern "C"
