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
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

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

// Graph capture and the calls HIP may synchronise in (device and pinned
// allocation and release, graph teardown, device-wide waits) never overlap
// inside this library: HIP lists them as unsafe while a stream captures, and
// a server's pump thread captures its graph while host threads build
// operators and batches.  Each such call and each capture holds this lock.
std::mutex& device_mutex() {
    static std::mutex m;
    return m;
}
using DeviceLock = std::lock_guard<std::mutex>;

void* device_alloc(size_t bytes) {
    void* p = nullptr;
    hipError_t err;
    {
        DeviceLock lock(device_mutex());
        err = hipMalloc(&p, bytes == 0 ? 256 : bytes);
    }
    if (err == hipErrorOutOfMemory) fail(PCBHIP_NO_MEMORY, "out of device memory");
    check(err, "hipMalloc");
    return p;
}

hipError_t device_free(void* p) {
    DeviceLock lock(device_mutex());
    return hipFree(p);
}

// Operator memory comes from the stream-ordered allocator on the operator's
// own stream, so releasing it waits for that stream only, not for the whole
// device (which a server in another thread keeps busy).  The allocation is
// complete, and usable from every stream, when this returns.
void* stream_alloc(hipStream_t stream, size_t bytes) {
    void* p = nullptr;
    hipError_t err;
    {
        DeviceLock lock(device_mutex());
        err = hipMallocAsync(&p, bytes == 0 ? 256 : bytes, stream);
    }
    if (err == hipErrorOutOfMemory) fail(PCBHIP_NO_MEMORY, "out of device memory");
    check(err, "hipMallocAsync");
    check(hipStreamSynchronize(stream), "allocation");
    return p;
}

hipError_t stream_free(hipStream_t stream, void* p) {
    DeviceLock lock(device_mutex());
    return hipFreeAsync(p, stream);
}

hipError_t pinned_free(void* p) {
    DeviceLock lock(device_mutex());
    return hipHostFree(p);
}

hipError_t stream_create(hipStream_t* stream) {
    DeviceLock lock(device_mutex());
    return hipStreamCreateWithFlags(stream, hipStreamNonBlocking);
}

hipError_t stream_destroy(hipStream_t stream) {
    DeviceLock lock(device_mutex());
    return hipStreamDestroy(stream);
}

hipError_t graph_release(hipGraphExec_t exec, hipGraph_t graph) {
    DeviceLock lock(device_mutex());
    hipError_t err = hipSuccess;
    if (exec) err = hipGraphExecDestroy(exec);
    if (graph) {
        const hipError_t e = hipGraphDestroy(graph);
        if (err == hipSuccess) err = e;
    }
    return err;
}

size_t align256(size_t bytes) { return (bytes + 255) & ~size_t(255); }

int strips_x(int C) { return (C + 1 + kStripCols - 1) / kStripCols; }
int groups_y(int R, int rows) { return (R + 1 + rows * kWavesPerGroup - 1) / (rows * kWavesPerGroup); }

// Planes per z block: the whole depth when the launch already has
// ``target`` workgroups, else split into the blocks needed to reach it
// (each block re-reads one halo plane on each side, so never fewer than 2
// planes).
int planes_per_block(int S, long long groups_xy, long long target) {
    // A tile holds at most kMaxTilePairs plane pairs of reduction cells.
    const int planes = std::min(S + 1, 2 * kMaxTilePairs);
    if (groups_xy >= target) return planes;
    const long long split = (target + groups_xy - 1) / groups_xy;
    int zb = (int)((planes + split - 1) / split);
    zb += zb & 1;  // whole plane pairs: z blocks never cut a reduction cell
    return zb < 2 ? 2 : (zb > planes ? planes : zb);
}

// Bit z set for each of the first 64 node planes (sheet: layers) of
// ``per_plane`` nodes whose robin words are all exactly +0: every node free,
// no Robin term.  The stencils skip the robin load there.
unsigned long long plain_planes(const std::vector<double>& robin64, long long planes, long long per_plane) {
    unsigned long long mask = 0;
    for (long long z = 0; z < planes && z < 64; ++z) {
        bool plain = true;
        for (long long i = 0; i < per_plane && plain; ++i) {
            const double v = robin64[(size_t)(z * per_plane + i)];
            plain = v == 0.0 && !std::signbit(v);
        }
        if (plain) mask |= 1ull << z;
    }
    return mask;
}

// Bit ez set for each of the first kMaxUniformLayers element layers whose
// elements all have the same (a, b, c); their values into ``values``
// (layer-major, 3 per layer, zero for the others).
unsigned long long uniform_layers(const std::vector<double>& coef64, int S, long long per_layer,
                                  std::vector<double>& values) {
    const long long E = (long long)S * per_layer;
    values.assign(3 * kMaxUniformLayers, 0.0);
    unsigned long long mask = 0;
    for (int ez = 0; ez < S && ez < kMaxUniformLayers; ++ez) {
        bool same = true;
        for (int d = 0; d < 3 && same; ++d) {
            const double* c = coef64.data() + d * E + ez * per_layer;
            for (long long i = 1; i < per_layer && same; ++i) same = c[i] == c[0] && !std::signbit(c[i]);
            same = same && !std::signbit(c[0]);
        }
        if (!same) continue;
        mask |= 1ull << ez;
        for (int d = 0; d < 3; ++d) values[(size_t)(3 * ez + d)] = coef64[(size_t)(d * E + ez * per_layer)];
    }
    return mask;
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
    unsigned long long plain = 0;  // bit z: robin is +0 throughout node plane (sheet: layer) z
    unsigned long long uniform = 0;  // hex: bit ez, every element of layer ez has the same coefficients
    float* ucoef32 = nullptr;        // hex: their values, (kMaxUniformLayers, 3), for the standalone apply
    double* ucoef64 = nullptr;
    float* dinv32 = nullptr;
    double* dinv64 = nullptr;
    double* diag64 = nullptr;
    float* nrec32 = nullptr;  // (N, 2) records (1 / diagonal, robin)
    double* nrec64 = nullptr;
    int kind = kHex;
    int* via_ptr = nullptr;  // sheet operators: node-ordered via CSR (null without vias)
    int* via_nbr = nullptr;
    float* via_g32 = nullptr;
    double* via_g64 = nullptr;
    int rows32 = 4, zb32 = 0, rows64 = 4, zb64 = 0;  // standalone apply tile shapes
    int mfma32 = 0;  // standalone FP32 element action: 0 vector, 2 or 3 bf16 terms on the matrix cores
    // Two-level coarse space (nc = 0: none).  A replaced inverse stays
    // allocated until the operator goes: slots made before the change keep
    // the view they were made with.
    float* cinv = nullptr;
    int cblock = 0, crows = 0, ccols = 0, nc = 0;
    std::vector<float*> retired;
    // Device staging of the host-buffer calls (applies, preconditioner),
    // kept for the operator's life: freeing device memory waits for the
    // whole device, which a server in another thread keeps busy.
    mutable std::mutex staging_mutex;
    mutable void* staging = nullptr;
    mutable size_t staging_bytes = 0;

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
        v.plain = plain;
        v.uniform = uniform;
        v.dinv32 = dinv32;
        v.dinv64 = dinv64;
        v.nrec32 = nrec32;
        v.nrec64 = nrec64;
        v.kind = kind;
        v.via_ptr = via_ptr;
        v.via_nbr = via_nbr;
        v.via_g32 = via_g32;
        v.via_g64 = via_g64;
        v.cinv = cinv;
        v.cblock = cblock;
        v.crows = crows;
        v.ccols = ccols;
        v.nc = nc;
        return v;
    }
    // Node layers of the coarse space: the node planes of a hexahedral
    // operator, the copper layers of a sheet operator.
    int layers() const { return kind == kSheet ? S : S + 1; }
    ~pcbhip_operator() {
        for (void* p : {block, (void*)cinv, staging, (void*)via_ptr})
            if (p) (void)stream_free(stream, p);
        for (float* p : retired) (void)stream_free(stream, p);
        if (stream) (void)stream_destroy(stream);
    }
};

namespace {

struct SlotHost {
    const pcbhip_operator* op = nullptr;
    OperatorDev view{};  // the operator as it was when the slot was made
    void* block = nullptr;
    double* b = nullptr;
    double* x[2] = {nullptr, nullptr};
    double* x0 = nullptr;
    void* state[2] = {nullptr, nullptr};
    double* part = nullptr;
    double* cg = nullptr;
    double* ch = nullptr;
    double* cw = nullptr;
    double* history = nullptr;
    int history_rows = 0;
    pcbhip_mpir_config config{1e-10, 0.0, 2e-3, 8, 200};
    int cells = 0, part_stride = 0;
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
    Tile* d_block = nullptr;
    Tile* d_patches = nullptr;  // coarse restriction tiles of the two-level slots
    Tile* d_rows = nullptr;     // coarse row tiles
    SlotGroup* d_groups = nullptr;
    int* d_counts = nullptr;  // live table lengths: outer, inner, (unused), block, patches, coarse rows
    int n_outer = 0, n_inner = 0, n_block = 0, n_patches = 0, n_rows = 0;
    bool plain = true;  // some tiled slot is Jacobi-only
    int block_group = 0;  // right-hand sides per block-apply group (0: no block apply)
    int block_mfma = 3;  // element action of the block apply (0, 2, 3)
    int rows_inner = 4;
    // Longer replays and sparser outer steps cost a slot whose inner solve
    // ended at most outer_every idle steps, and save the launches of the
    // outer pair (about 11 us) in the other steps: 115.5 against 117.0 ms
    // for one 4.5 M-node solve, the 200-board batch unchanged.
    int outer_every = 32;
    int replay_steps = 64;   // steps per captured replay
    int compact_percent = 25;  // drop finished slots' tiles once this share of the tiled slots has ended
    int tiled_slots = 0;  // slots whose tiles are in the tables
    long long target_groups = kPassTargetGroups;
    int* d_active = nullptr;
    int* h_active = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
    bool configs_dirty = true;
    bool ran = false;  // pcbhip_batch_run has completed once
    int h_counts[6] = {0, 0, 0, 0, 0, 0};  // source of the last counts upload (synchronised before reuse)

    launch::CoarseLaunch coarse() const {
        launch::CoarseLaunch c;
        c.patches = d_patches;
        c.rows = d_rows;
        c.grid_patches = n_patches;
        c.grid_rows = n_rows;
        c.plain = plain;
        return c;
    }

    ~pcbhip_batch() {
        (void)graph_release(exec, graph);
        for (SlotHost& s : slots) {
            if (s.block) (void)device_free(s.block);
            if (s.history) (void)device_free(s.history);
        }
        for (void* p : {(void*)d_slots, (void*)d_states, (void*)d_configs, (void*)d_outer, (void*)d_inner,
                        (void*)d_block, (void*)d_patches, (void*)d_rows, (void*)d_groups, (void*)d_active,
                        (void*)d_counts})
            if (p) (void)device_free(p);
        if (h_active) (void)pinned_free(h_active);
        if (stream) (void)stream_destroy(stream);
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
    if (op->kind == kSheet) {
        check(launch::sheet_apply<T>(op->view(), x, y, mode, stream), "sheet apply launch");
        return;
    }
    CoefSource<T> coef;
    const T* robin;
    if constexpr (sizeof(T) == 4) {
        coef = {op->coef32, op->uniform, op->ucoef32};
        robin = op->robin32;
    } else {
        coef = {op->coef64, op->uniform, op->ucoef64};
        robin = op->robin64;
    }
    const int rows = sizeof(T) == 4 ? op->rows32 : op->rows64;
    const int zb = sizeof(T) == 4 ? op->zb32 : op->zb64;
    check(launch::apply<T>(coef, robin, op->plain, x, y, op->S, op->R, op->C, mode, rows, zb, op->mfma32, stream),
          "apply launch");
}

// The operator's staging buffer, at least ``bytes`` long; the caller holds
// op->staging_mutex.  It only grows (to the largest request so far).
char* staging_of(const pcbhip_operator* op, size_t bytes) {
    if (op->staging_bytes < bytes) {
        if (op->staging) {
            check(stream_free(op->stream, op->staging), "hipFreeAsync");
            op->staging = nullptr;
            op->staging_bytes = 0;
        }
        op->staging = stream_alloc(op->stream, bytes);
        op->staging_bytes = bytes;
    }
    return static_cast<char*>(op->staging);
}

template <typename T>
void apply_host(const pcbhip_operator* op, const T* x, T* y, int mode) {
    if (!op || !x || !y) fail(PCBHIP_INVALID_INPUT, "null argument");
    set_device(op->device);
    const size_t bytes = sizeof(T) * (size_t)op->N;
    std::lock_guard<std::mutex> guard(op->staging_mutex);
    // Room for FP64 x and y at once, so FP32 and FP64 calls share one buffer.
    char* base = staging_of(op, 2 * align256(sizeof(double) * (size_t)op->N));
    T* dx = reinterpret_cast<T*>(base);
    T* dy = reinterpret_cast<T*>(base + align256(bytes));
    check(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, op->stream), "copy x");
    apply_dev<T>(op, dx, dy, mode, op->stream);
    check(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, op->stream), "copy y");
    check(hipStreamSynchronize(op->stream), "apply");
}

SlotDev slot_dev(const SlotHost& s) {
    SlotDev d{};
    d.op = s.view;
    d.b64 = s.b;
    d.x64[0] = s.x[0];
    d.x64[1] = s.x[1];
    d.state[0] = s.state[0];
    d.state[1] = s.state[1];
    d.part = s.part;
    d.cg = s.cg;
    d.ch = s.ch;
    d.cw = s.cw;
    d.history = s.history;
    d.cells = s.cells;
    d.part_stride = s.part_stride;
    return d;
}

SlotConfig slot_config(const pcbhip_mpir_config& c) {
    return SlotConfig{c.relative_tolerance, c.absolute_tolerance, c.inner_relative_tolerance, c.max_outer_iterations,
                      c.max_inner_iterations};
}

// Device bytes of one slot's vectors: b and the x pair in FP64, the pair of
// KNode state arrays (r, q, p, c per node) in the inner precision, the
// partial sums and, for a two-level slot, its three coarse vectors.
size_t slot_vector_bytes(const SlotHost& s, int inner64) {
    const size_t ti = inner64 ? sizeof(double) : sizeof(float);
    return 3 * align256(sizeof(double) * s.op->N) + 2 * align256(4 * ti * s.op->N) +
           align256(sizeof(double) * kInnerSums * (size_t)s.part_stride) +
           3 * align256(sizeof(double) * (size_t)s.view.nc);
}

// Points the slot's vectors into one block laid out as slot_vector_bytes
// counts them (the batch adds an x0 buffer behind its own block).
void carve_slot(SlotHost& s, char* base, int inner64) {
    const size_t ti = inner64 ? sizeof(double) : sizeof(float);
    const size_t n64 = align256(sizeof(double) * s.op->N), nst = align256(4 * ti * s.op->N);
    s.block = base;
    s.b = reinterpret_cast<double*>(base);
    s.x[0] = reinterpret_cast<double*>(base + n64);
    s.x[1] = reinterpret_cast<double*>(base + 2 * n64);
    base += 3 * n64;
    s.state[0] = base;
    s.state[1] = base + nst;
    s.part = reinterpret_cast<double*>(base + 2 * nst);
    base += 2 * nst + align256(sizeof(double) * kInnerSums * (size_t)s.part_stride);
    const size_t ncb = align256(sizeof(double) * (size_t)s.view.nc);
    s.cg = s.view.nc ? reinterpret_cast<double*>(base) : nullptr;
    s.ch = s.view.nc ? reinterpret_cast<double*>(base + ncb) : nullptr;
    s.cw = s.view.nc ? reinterpret_cast<double*>(base + 2 * ncb) : nullptr;
}

// Coarse tiles of a two-level slot: one restriction tile per patch and one
// row tile per kCoarseRows rows of its inverse.
void add_coarse_tiles(std::vector<Tile>& patches, std::vector<Tile>& rows, int slot, const OperatorDev& v) {
    if (v.nc <= 0) return;
    const int layers = v.kind == kSheet ? v.S : v.S + 1;
    int index = 0;
    for (int l = 0; l < layers; ++l)
        for (int cy = 0; cy < v.crows; ++cy)
            for (int cx = 0; cx < v.ccols; ++cx)
                patches.push_back(Tile{slot, index++, (short)cx, (short)cy, (short)l, 1});
    for (int row = 0; row < v.nc; row += kCoarseRows) rows.push_back(Tile{slot, row, 0, 0, 0, 1});
}

// Tiles of one slot: stencil workgroups in (y group, x strip, z block)
// order, so a contiguous range of tiles is a compact piece of the board and
// the z blocks of one column, which share two node planes and sweep towards
// or away from each other, run side by side.
int add_tiles(std::vector<Tile>& out, int slot, int nsx, int nsy, int S, int zb) {
    const int nsz = (S + 1 + zb - 1) / zb;
    int index = 0;
    for (int sy = 0; sy < nsy; ++sy)
        for (int sx = 0; sx < nsx; ++sx)
            for (int sz = 0; sz < nsz; ++sz)
                out.push_back(Tile{slot, index++, (short)sx, (short)sy, (short)sz, (short)zb});
    return index;
}

int add_sheet_tiles(std::vector<Tile>& out, int slot, int nsx, int nsy, int layers) {
    int index = 0;
    for (int l = 0; l < layers; ++l)
        for (int sy = 0; sy < nsy; ++sy)
            for (int sx = 0; sx < nsx; ++sx) out.push_back(Tile{slot, index++, (short)sx, (short)sy, (short)l, 1});
    return index;
}

void upload_tiles(pcbhip_batch* b, const std::vector<Tile>& t, Tile** d, int* n) {
    if (*d) check(device_free(*d), "hipFree");
    *n = (int)t.size();
    *d = static_cast<Tile*>(device_alloc(sizeof(Tile) * std::max<size_t>(1, t.size())));
    if (!t.empty())
        check(hipMemcpyAsync(*d, t.data(), sizeof(Tile) * t.size(), hipMemcpyHostToDevice, b->stream), "upload");
}

// Appends the tiles of the slots that ``slots`` holds (nullptr entries are
// skipped; the index in the vector is the slot number).  Slots that belong to
// one of ``groups`` get no inner tiles of their own: each group gets block
// tiles (4-row strips, the tile's slot field holding the group index).  The
// z split is chosen over all listed slots and groups together: many slots
// fill the device without splitting any of them.
void collect_tiles(const std::vector<const SlotHost*>& slots, const std::vector<SlotGroup>& groups,
                   int rows_inner, long long target_groups, std::vector<Tile>& outer, std::vector<Tile>& inner,
                   std::vector<Tile>& block) {
    std::vector<char> grouped(slots.size(), 0);
    for (const SlotGroup& g : groups)
        for (int k = 0; k < g.count; ++k) grouped[(size_t)g.slot[k]] = 1;
    long long xy_outer = 0, xy_inner = 0;
    for (size_t i = 0; i < slots.size(); ++i) {
        const SlotHost* s = slots[i];
        if (!s) continue;
        // Sheet tiles are fixed (one per layer strip); they still count
        // toward filling the device.
        const long long layers = s->op->kind == kSheet ? s->op->S : 1;
        xy_outer += layers * strips_x(s->op->C) * groups_y(s->op->R, kRowsOuter);
        if (!grouped[i]) xy_inner += layers * strips_x(s->op->C) * groups_y(s->op->R, rows_inner);
    }
    for (const SlotGroup& g : groups) {
        const pcbhip_operator* op = slots[(size_t)g.slot[0]]->op;
        xy_inner += (long long)strips_x(op->C) * groups_y(op->R, 4);
    }
    for (int i = 0; i < (int)slots.size(); ++i) {
        const SlotHost* s = slots[(size_t)i];
        if (!s) continue;
        const pcbhip_operator* op = s->op;
        const int nsx = strips_x(op->C);
        if (op->kind == kSheet) {
            // One layer of 4-row wave strips per tile, for both stencils.
            add_sheet_tiles(outer, i, nsx, groups_y(op->R, kCellRows), op->S);
            if (!grouped[(size_t)i]) add_sheet_tiles(inner, i, nsx, groups_y(op->R, kCellRows), op->S);
            continue;
        }
        const int nyo = groups_y(op->R, kRowsOuter), nyi = groups_y(op->R, rows_inner);
        const long long share_o = (long long)nsx * nyo * target_groups / xy_outer;
        const int zbo = planes_per_block(op->S, (long long)nsx * nyo, std::max(share_o, (long long)nsx * nyo));
        add_tiles(outer, i, nsx, nyo, op->S, zbo);
        if (!grouped[(size_t)i]) {
            const long long share_i = (long long)nsx * nyi * target_groups / xy_inner;
            const int zbi = planes_per_block(op->S, (long long)nsx * nyi, std::max(share_i, (long long)nsx * nyi));
            add_tiles(inner, i, nsx, nyi, op->S, zbi);
        }
    }
    // Block tiles: the groups of one operator interleaved at every spatial
    // tile, so the workgroups that read one piece of the operator run
    // together and share it through the XCD's L2.
    std::vector<char> placed(groups.size(), 0);
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        if (placed[gi]) continue;
        const pcbhip_operator* op = slots[(size_t)groups[gi].slot[0]]->op;
        std::vector<int> same;
        for (size_t gj = gi; gj < groups.size(); ++gj)
            if (!placed[gj] && slots[(size_t)groups[gj].slot[0]]->op == op) {
                same.push_back((int)gj);
                placed[gj] = 1;
            }
        const int nsx = strips_x(op->C), ny = groups_y(op->R, 4);
        const long long share = (long long)nsx * ny * target_groups / xy_inner;
        const int zb = planes_per_block(op->S, (long long)nsx * ny, std::max(share, (long long)nsx * ny));
        const int nsz = (op->S + 1 + zb - 1) / zb;
        int index = 0;
        for (int sy = 0; sy < ny; ++sy)
            for (int sx = 0; sx < nsx; ++sx)
                for (int sz = 0; sz < nsz; ++sz, ++index)
                    for (int g : same) block.push_back(Tile{g, index, (short)sx, (short)sy, (short)sz, (short)zb});
    }
}

// Groups of at most ``size`` listed slots that share an operator, in slot
// order, split evenly (none when size is 0).  A lone slot is a group of one,
// so a right-hand side gets the same bits alone as in a group.
std::vector<SlotGroup> group_slots(const std::vector<const SlotHost*>& slots, int size) {
    std::vector<SlotGroup> out;
    if (size <= 0) return out;
    std::vector<const pcbhip_operator*> ops;
    std::vector<std::vector<int>> members;
    for (int i = 0; i < (int)slots.size(); ++i) {
        // Two-level slots keep tiles of their own (the block pass is Jacobi).
        if (!slots[(size_t)i] || slots[(size_t)i]->op->kind != kHex || slots[(size_t)i]->view.nc > 0) continue;
        const pcbhip_operator* op = slots[(size_t)i]->op;
        size_t k = 0;
        while (k < ops.size() && ops[k] != op) ++k;
        if (k == ops.size()) {
            ops.push_back(op);
            members.emplace_back();
        }
        members[k].push_back(i);
    }
    for (const std::vector<int>& m : members) {
        const int n = (int)m.size();
        const int chunks = (n + size - 1) / size;
        int start = 0;
        for (int c = 0; c < chunks; ++c) {
            const int len = n / chunks + (c < n % chunks ? 1 : 0);
            SlotGroup g{};
            g.count = len;
            for (int k = 0; k < len; ++k) g.slot[k] = m[(size_t)(start + k)];
            start += len;
            out.push_back(g);
        }
    }
    return out;
}

// Partial-sum layout of a slot: one partial per reduction cell of the
// stencils for each of the sums a kernel keeps.
void set_partial_layout(SlotHost& s) {
    const pcbhip_operator* op = s.op;
    s.cells = op->kind == kSheet ? sheet_cell_count(op->S, op->R, op->C) : cell_count(op->S, op->R, op->C);
    s.part_stride = s.cells;
}

// (Re)builds the tile tables for the batch's current shape; with
// ``keep``, only the slots it marks (finished slots leave the tables, so their
// early exits stop costing launch slots; their partial layout does not change).
void build_tiles(pcbhip_batch* b, const std::vector<char>* keep = nullptr) {
    const int rows_inner = b->inner64 ? kRowsInner64 : b->rows_inner;
    std::vector<const SlotHost*> listed(b->slots.size(), nullptr);
    b->tiled_slots = 0;
    for (size_t i = 0; i < b->slots.size(); ++i) {
        if (keep && !(*keep)[i]) continue;
        listed[i] = &b->slots[i];
        b->tiled_slots += 1;
    }
    std::vector<Tile> outer, inner, block, patches, rows;
    const std::vector<SlotGroup> groups = group_slots(listed, b->inner64 ? 0 : b->block_group);
    collect_tiles(listed, groups, rows_inner, b->target_groups, outer, inner, block);
    b->plain = false;
    for (size_t i = 0; i < listed.size(); ++i) {
        if (!listed[i]) continue;
        add_coarse_tiles(patches, rows, (int)i, listed[i]->view);
        if (listed[i]->view.nc == 0) b->plain = true;
    }
    if (outer.empty()) outer.push_back(Tile{0, 0, 0, 0, 0, 1});
    if (inner.empty() && block.empty()) inner.push_back(Tile{0, 0, 0, 0, 0, 1});
    upload_tiles(b, outer, &b->d_outer, &b->n_outer);
    upload_tiles(b, inner, &b->d_inner, &b->n_inner);
    upload_tiles(b, block, &b->d_block, &b->n_block);
    upload_tiles(b, patches, &b->d_patches, &b->n_patches);
    upload_tiles(b, rows, &b->d_rows, &b->n_rows);
    if (b->d_groups) check(device_free(b->d_groups), "hipFree");
    b->d_groups = static_cast<SlotGroup*>(device_alloc(sizeof(SlotGroup) * std::max<size_t>(1, groups.size())));
    if (!groups.empty())
        check(hipMemcpyAsync(b->d_groups, groups.data(), sizeof(SlotGroup) * groups.size(), hipMemcpyHostToDevice,
                             b->stream),
              "groups");
    if (!b->d_counts) b->d_counts = static_cast<int*>(device_alloc(6 * sizeof(int)));
    b->h_counts[0] = b->n_outer;
    b->h_counts[1] = b->n_inner;
    b->h_counts[2] = 0;
    b->h_counts[3] = b->n_block;
    b->h_counts[4] = b->n_patches;
    b->h_counts[5] = b->n_rows;
    check(hipMemcpyAsync(b->d_counts, b->h_counts, 6 * sizeof(int), hipMemcpyHostToDevice, b->stream), "counts");
    // Tables and counts are read through stream order; the sources above are
    // host memory that changes at the next build, so wait here.
    check(hipStreamSynchronize(b->stream), "tile upload");
    b->configs_dirty = true;
    if (b->exec) {
        check(graph_release(b->exec, nullptr), "graph exec destroy");
        b->exec = nullptr;
    }
}

void build_graph(pcbhip_batch* b, int steps) {
    check(graph_release(b->exec, b->graph), "graph destroy");
    b->exec = nullptr;
    b->graph = nullptr;
    DeviceLock lock(device_mutex());
    check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    const int n_slots = (int)b->slots.size();
    for (int s = 0; s < steps; ++s) {
        hipError_t err = hipSuccess;
        // The outer step runs every outer_every inner iterations: a slot whose
        // inner solve ended waits at most that long, and the idle outer
        // kernels stay out of most steps.
        if (s % b->outer_every == 0)
            err = launch::outer_step(b->inner64 != 0, n_slots, b->d_slots, b->d_states, b->d_configs, b->d_outer,
                                     b->d_counts, b->n_outer, b->d_active, b->coarse(), b->stream);
        if (err == hipSuccess)
            err = launch::inner_step(b->inner64 != 0, b->rows_inner, n_slots, b->d_slots, b->d_states, b->d_configs,
                                     b->d_inner, b->d_counts, std::max(b->n_inner, 1), false, nullptr, b->d_block,
                                     b->d_groups, b->n_block, b->block_mfma, b->d_active, b->coarse(), b->stream);
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

// ------------------------------------------------------------------ server

namespace {

// A slot whose solve has ended is first "copying": its solution and history
// are copied into pinned host memory on the replay stream, ahead of the next
// replay, and it becomes "finished" (collectable without any device wait)
// once a replay launched after those copies has been retired.
enum SlotStatus : int { kSlotFree = 0, kSlotQueued = 1, kSlotRunning = 2, kSlotFinished = 3, kSlotCopying = 4 };

struct ServerSlot {
    SlotHost h;
    size_t block_bytes = 0;
    int status = kSlotFree;
    int64_t tag = 0;
    long long first_replay = 0;  // sequence number of the first replay that runs it
    long long cap_steps = 0;     // steps after which the slot must have ended
    SlotState final_state{};
    // Pinned result buffer (solution, then history rows) from the server's
    // pinned pool, while copying or finished.
    double* host = nullptr;
    size_t host_bytes = 0;
    bool copy_enqueued = false;
    long long copy_seq = 0;  // the replay launched right after the result copies
};

// Pinned sources of one set of table uploads.  Two sets are used in turn;
// ``used`` is recorded after the uploads that read a set, and the host waits
// on it before writing the set again.
struct Staging {
    int* counts = nullptr;
    Tile* tiles = nullptr;
    size_t tile_cap = 0;
    SlotDev* devs = nullptr;
    SlotConfig* cfgs = nullptr;
    SlotState* fresh = nullptr;
    int* fresh_index = nullptr;
    hipEvent_t used = nullptr;
    bool recorded = false;
};

struct Replay {
    long long seq;
    int parity;
};

}  // namespace

struct pcbhip_server {
    int device = 0;
    int inner64 = 0;
    int rows_inner = 4;
    int outer_every = 8;
    int steps = 16;
    int max_slots = 0;
    long long target_groups = kPassTargetGroups;
    hipStream_t stream = nullptr;       // the MPIR replays and table uploads
    hipStream_t copy_stream = nullptr;  // admissions and results
    std::vector<ServerSlot> slots;
    std::vector<std::pair<void*, size_t>> pool;  // released slot blocks
    std::vector<std::pair<void*, size_t>> pinned_pool;  // released pinned result buffers
    size_t bytes_in_use = 0, bytes_pooled = 0;
    SlotDev* d_slots = nullptr;
    SlotState* d_states = nullptr;
    SlotConfig* d_configs = nullptr;
    SlotState* d_fresh = nullptr;
    int* d_fresh_index = nullptr;
    int* d_active = nullptr;  // decremented by the finish kernels; the server reads states instead
    int* d_counts = nullptr;  // outer, inner, (unused), (unused), patches, coarse rows
    int* d_tickets = nullptr;  // tile tickets of the looping pass (kPassTickets)
    Tile* d_tiles = nullptr;  // outer | inner | patches | coarse rows tables at their capacities
    int cap_outer = 0, cap_inner = 0, cap_patches = 0, cap_rows = 0;
    Staging stage[2];
    int stage_next = 0;
    SlotState* h_states = nullptr;  // pinned, two snapshots of max_slots states
    hipEvent_t replay_done[2] = {nullptr, nullptr};
    hipEvent_t replay_start[2] = {nullptr, nullptr};  // timing pairs of the replays in flight
    std::vector<Replay> inflight;  // oldest first, at most two
    long long seq = 0;             // replays launched
    long long graph_builds = 0;
    double wait_s = 0.0;           // host time blocked on replays in flight (the device was the bottleneck)
    double replay_s = 0.0;         // device time of the retired replays, start to end
    bool tables_dirty = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~pcbhip_server() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (copy_stream) (void)hipStreamSynchronize(copy_stream);
        (void)graph_release(exec, graph);
        for (ServerSlot& s : slots) {
            if (s.h.block) (void)device_free(s.h.block);
            if (s.host) (void)pinned_free(s.host);
        }
        for (auto& b : pool) (void)device_free(b.first);
        for (auto& b : pinned_pool) (void)pinned_free(b.first);
        for (void* p : {(void*)d_slots, (void*)d_states, (void*)d_configs, (void*)d_fresh, (void*)d_fresh_index,
                        (void*)d_active, (void*)d_counts, (void*)d_tickets, (void*)d_tiles})
            if (p) (void)device_free(p);
        for (Staging& st : stage) {
            for (void* p : {(void*)st.counts, (void*)st.tiles, (void*)st.devs, (void*)st.cfgs, (void*)st.fresh,
                            (void*)st.fresh_index})
                if (p) (void)pinned_free(p);
            if (st.used) (void)hipEventDestroy(st.used);
        }
        if (h_states) (void)pinned_free(h_states);
        for (hipEvent_t e : replay_done)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : replay_start)
            if (e) (void)hipEventDestroy(e);
        if (copy_stream) (void)stream_destroy(copy_stream);
        if (stream) (void)stream_destroy(stream);
    }
};

namespace {

template <typename T>
T* pinned_alloc(size_t count) {
    void* p = nullptr;
    hipError_t err;
    {
        DeviceLock lock(device_mutex());
        err = hipHostMalloc(&p, sizeof(T) * (count ? count : 1), hipHostMallocDefault);
    }
    check(err, "pinned alloc");
    return static_cast<T*>(p);
}

// Waits for the oldest replay in flight and takes its snapshot of the slot
// states: slots it shows finished leave the tables at the next upload.
void server_retire_oldest(pcbhip_server* sv) {
    const Replay r = sv->inflight.front();
    sv->inflight.erase(sv->inflight.begin());
    const auto t0 = std::chrono::steady_clock::now();
    check(hipEventSynchronize(sv->replay_done[r.parity]), "replay");
    sv->wait_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, sv->replay_start[r.parity], sv->replay_done[r.parity]) == hipSuccess)
        sv->replay_s += 1e-3 * ms;
    const SlotState* snap = sv->h_states + (size_t)r.parity * sv->max_slots;
    for (int i = 0; i < sv->max_slots; ++i) {
        ServerSlot& s = sv->slots[(size_t)i];
        // Result copies enqueued before this replay are complete.
        if (s.status == kSlotCopying && s.copy_enqueued && s.copy_seq <= r.seq) s.status = kSlotFinished;
        if (s.status != kSlotRunning || s.first_replay > r.seq) continue;
        if (snap[i].phase == kPhaseDone) {
            s.status = kSlotCopying;
            s.copy_enqueued = false;
            s.final_state = snap[i];
            sv->tables_dirty = true;
        } else if ((r.seq - s.first_replay + 1) * sv->steps > s.cap_steps) {
            fail(PCBHIP_RUNTIME, "MPIR iteration did not terminate within its limits");
        }
    }
}

// A pinned host buffer of at least ``bytes`` from the server's pool (best
// fit), else a new one.
double* server_take_pinned(pcbhip_server* sv, size_t bytes, size_t* got) {
    size_t best = sv->pinned_pool.size();
    for (size_t i = 0; i < sv->pinned_pool.size(); ++i)
        if (sv->pinned_pool[i].second >= bytes &&
            (best == sv->pinned_pool.size() || sv->pinned_pool[i].second < sv->pinned_pool[best].second))
            best = i;
    if (best < sv->pinned_pool.size()) {
        void* p = sv->pinned_pool[best].first;
        *got = sv->pinned_pool[best].second;
        sv->pinned_pool.erase(sv->pinned_pool.begin() + (long)best);
        return static_cast<double*>(p);
    }
    *got = bytes;
    return pinned_alloc<double>(bytes / sizeof(double));
}

// Enqueues, on the replay stream, the result copies of the slots that ended
// since the last call; they complete before the next replay starts.
void server_enqueue_copies(pcbhip_server* sv) {
    for (ServerSlot& s : sv->slots) {
        if (s.status != kSlotCopying || s.copy_enqueued) continue;
        const size_t nx = sizeof(double) * (size_t)s.h.op->N;
        const size_t nh = sizeof(double) * 4 * (size_t)std::max(s.h.history_rows, 1);
        if (!s.host) s.host = server_take_pinned(sv, nx + nh, &s.host_bytes);
        check(hipMemcpyAsync(s.host, s.h.x[s.final_state.xcur], nx, hipMemcpyDeviceToHost, sv->stream), "result copy");
        if (s.final_state.outer > 0)
            check(hipMemcpyAsync(s.host + s.h.op->N, s.h.history, sizeof(double) * 4 * (size_t)s.final_state.outer,
                                 hipMemcpyDeviceToHost, sv->stream),
                  "history copy");
        s.copy_enqueued = true;
        s.copy_seq = sv->seq;
    }
}

// Copies enqueued with no replay after them: wait for them here.
void server_settle_copies(pcbhip_server* sv) {
    bool pending = false;
    for (const ServerSlot& s : sv->slots) pending = pending || (s.status == kSlotCopying && s.copy_enqueued);
    if (!pending) return;
    const auto t0 = std::chrono::steady_clock::now();
    check(hipStreamSynchronize(sv->stream), "result copies");
    sv->wait_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (ServerSlot& s : sv->slots)
        if (s.status == kSlotCopying && s.copy_enqueued) s.status = kSlotFinished;
}

void server_drain(pcbhip_server* sv) {
    while (!sv->inflight.empty()) server_retire_oldest(sv);
}

void server_build_graph(pcbhip_server* sv) {
    check(graph_release(sv->exec, sv->graph), "graph destroy");
    sv->exec = nullptr;
    sv->graph = nullptr;
    Tile* outer = sv->d_tiles;
    Tile* inner = outer + sv->cap_outer;
    // Two-level slots: the coarse kernels over their tables (grids of the
    // tables' capacities; workgroups past the live counts return at once).
    launch::CoarseLaunch coarse;
    coarse.patches = inner + sv->cap_inner;
    coarse.rows = coarse.patches + sv->cap_patches;
    coarse.grid_patches = sv->cap_patches;
    coarse.grid_rows = sv->cap_rows;
    DeviceLock lock(device_mutex());
    check(hipStreamBeginCapture(sv->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int s = 0; s < sv->steps; ++s) {
        hipError_t err = hipSuccess;
        if (s % sv->outer_every == 0)
            err = launch::outer_step(sv->inner64 != 0, sv->max_slots, sv->d_slots, sv->d_states, sv->d_configs,
                                     outer, sv->d_counts, sv->cap_outer, sv->d_active, coarse, sv->stream);
        if (err == hipSuccess)
            err = launch::inner_step(sv->inner64 != 0, sv->rows_inner, sv->max_slots, sv->d_slots, sv->d_states,
                                     sv->d_configs, inner, sv->d_counts, std::min(sv->cap_inner, kPassGroups), true,
                                     sv->d_tickets, nullptr, nullptr, 0, 0, sv->d_active, coarse, sv->stream);
        if (err != hipSuccess) {
            hipGraph_t dropped = nullptr;
            (void)hipStreamEndCapture(sv->stream, &dropped);
            if (dropped) (void)hipGraphDestroy(dropped);
            check(err, "capture launch");
        }
    }
    check(hipStreamEndCapture(sv->stream, &sv->graph), "end capture");
    check(hipGraphInstantiate(&sv->exec, sv->graph, nullptr, nullptr, 0), "graph instantiate");
    sv->graph_builds += 1;
}

int round_capacity(size_t n) { return (int)std::max<size_t>(256, ((n + n / 4 + 255) / 256) * 256); }

// Grows the device tables when they are too short and shrinks them when
// they are mostly empty (workgroups past the live count still cost dispatch);
// either way the replays in flight are drained and the graph is captured
// again with the new grids.  Returns whether it drained.
// Capacity of a coarse table: none while no two-level slot runs.
int coarse_capacity(size_t n) { return n ? round_capacity(n) : 0; }

bool server_reserve(pcbhip_server* sv, size_t n_outer, size_t n_inner, size_t n_patches, size_t n_rows) {
    const bool fits = n_outer <= (size_t)sv->cap_outer && n_inner <= (size_t)sv->cap_inner &&
                      n_patches <= (size_t)sv->cap_patches && n_rows <= (size_t)sv->cap_rows;
    const size_t live = n_outer + n_inner + n_patches + n_rows;
    const size_t cap = (size_t)sv->cap_outer + sv->cap_inner + sv->cap_patches + sv->cap_rows;
    const bool sparse = cap > 3 * 1024 && live * 3 < cap;
    // Coarse kernels stay out of the graph once no two-level slot is left.
    const bool idle_coarse = sv->cap_patches > 0 && n_patches == 0;
    if (sv->exec && fits && !sparse && !idle_coarse) return false;
    server_drain(sv);
    check(hipStreamSynchronize(sv->stream), "drain");
    if (!fits || sparse || idle_coarse || !sv->d_tiles) {
        if (sv->d_tiles) check(device_free(sv->d_tiles), "hipFree");
        sv->d_tiles = nullptr;
        sv->cap_outer = round_capacity(n_outer);
        sv->cap_inner = round_capacity(n_inner);
        sv->cap_patches = coarse_capacity(n_patches);
        sv->cap_rows = coarse_capacity(n_rows);
        sv->d_tiles = static_cast<Tile*>(device_alloc(
            sizeof(Tile) * ((size_t)sv->cap_outer + sv->cap_inner + sv->cap_patches + sv->cap_rows)));
    }
    server_build_graph(sv);
    return true;
}

// Uploads the tile tables of the running slots, every slot's descriptor and
// configuration, and fresh states for the slots admitted in this call; all
// of it runs on the replay stream between two replays.
void server_upload(pcbhip_server* sv, const std::vector<int>& admitted) {
    const int rows_inner = sv->inner64 ? kRowsInner64 : sv->rows_inner;
    std::vector<Tile> outer, inner, patches, rows;
    // A drain inside the reserve can finish slots: list again after one.
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<const SlotHost*> listed((size_t)sv->max_slots, nullptr);
        for (int i = 0; i < sv->max_slots; ++i)
            if (sv->slots[(size_t)i].status == kSlotRunning) listed[(size_t)i] = &sv->slots[(size_t)i].h;
        outer.clear();
        inner.clear();
        patches.clear();
        rows.clear();
        std::vector<Tile> block;
        collect_tiles(listed, {}, rows_inner, sv->target_groups, outer, inner, block);
        for (int i = 0; i < sv->max_slots; ++i)
            if (listed[(size_t)i]) add_coarse_tiles(patches, rows, i, listed[(size_t)i]->view);
        if (!server_reserve(sv, outer.size(), inner.size(), patches.size(), rows.size())) break;
    }
    Staging& st = sv->stage[sv->stage_next];
    sv->stage_next ^= 1;
    if (st.recorded) check(hipEventSynchronize(st.used), "staging");
    const size_t n_tiles = outer.size() + inner.size() + patches.size() + rows.size();
    if (n_tiles > st.tile_cap) {
        if (st.tiles) check(pinned_free(st.tiles), "pinned free");
        st.tile_cap = n_tiles + n_tiles / 4;
        st.tiles = pinned_alloc<Tile>(st.tile_cap);
    }
    st.counts[0] = (int)outer.size();
    st.counts[1] = (int)inner.size();
    st.counts[2] = 0;
    st.counts[3] = 0;
    st.counts[4] = (int)patches.size();
    st.counts[5] = (int)rows.size();
    Tile* staged = st.tiles;
    for (const std::vector<Tile>* table : {&outer, &inner, &patches, &rows}) {
        std::copy(table->begin(), table->end(), staged);
        staged += table->size();
    }
    for (int i = 0; i < sv->max_slots; ++i) {
        const ServerSlot& s = sv->slots[(size_t)i];
        const bool bound = s.status == kSlotRunning || s.status == kSlotFinished || s.status == kSlotCopying;
        st.devs[i] = bound ? slot_dev(s.h) : SlotDev{};
        st.cfgs[i] = bound ? slot_config(s.h.config) : SlotConfig{};
    }
    for (size_t k = 0; k < admitted.size(); ++k) {
        std::memset(&st.fresh[k], 0, sizeof(SlotState));
        st.fresh[k].phase = kPhaseOuter;
        st.fresh_index[k] = admitted[k];
    }
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    Tile* dest = sv->d_tiles;
    staged = st.tiles;
    const int caps[4] = {sv->cap_outer, sv->cap_inner, sv->cap_patches, sv->cap_rows};
    int k = 0;
    for (const std::vector<Tile>* table : {&outer, &inner, &patches, &rows}) {
        if (!table->empty())
            check(hipMemcpyAsync(dest, staged, sizeof(Tile) * table->size(), h2d, sv->stream), "tiles");
        staged += table->size();
        dest += caps[k++];
    }
    check(hipMemcpyAsync(sv->d_counts, st.counts, 6 * sizeof(int), h2d, sv->stream), "counts");
    check(hipMemcpyAsync(sv->d_slots, st.devs, sizeof(SlotDev) * sv->max_slots, h2d, sv->stream), "slots");
    check(hipMemcpyAsync(sv->d_configs, st.cfgs, sizeof(SlotConfig) * sv->max_slots, h2d, sv->stream), "configs");
    if (!admitted.empty()) {
        check(hipMemcpyAsync(sv->d_fresh, st.fresh, sizeof(SlotState) * admitted.size(), h2d, sv->stream), "fresh");
        check(hipMemcpyAsync(sv->d_fresh_index, st.fresh_index, sizeof(int) * admitted.size(), h2d, sv->stream),
              "fresh index");
        check(launch::scatter_states(sv->d_fresh_index, sv->d_fresh, sv->d_states, (int)admitted.size(),
                                     sv->stream),
              "scatter states");
    }
    check(hipEventRecord(st.used, sv->stream), "event");
    st.recorded = true;
    sv->tables_dirty = false;
}

void server_launch(pcbhip_server* sv) {
    const int parity = (int)(sv->seq & 1);
    check(hipEventRecord(sv->replay_start[parity], sv->stream), "event");
    check(hipGraphLaunch(sv->exec, sv->stream), "graph launch");
    check(hipMemcpyAsync(sv->h_states + (size_t)parity * sv->max_slots, sv->d_states,
                         sizeof(SlotState) * sv->max_slots, hipMemcpyDeviceToHost, sv->stream),
          "state snapshot");
    check(hipEventRecord(sv->replay_done[parity], sv->stream), "event");
    sv->inflight.push_back(Replay{sv->seq, parity});
    sv->seq += 1;
}

void* server_take_block(pcbhip_server* sv, size_t bytes, size_t* got) {
    // Best fit among the released blocks, else a new allocation.
    size_t best = sv->pool.size();
    for (size_t i = 0; i < sv->pool.size(); ++i)
        if (sv->pool[i].second >= bytes && (best == sv->pool.size() || sv->pool[i].second < sv->pool[best].second))
            best = i;
    if (best < sv->pool.size()) {
        void* p = sv->pool[best].first;
        *got = sv->pool[best].second;
        sv->pool.erase(sv->pool.begin() + (long)best);
        sv->bytes_pooled -= *got;
        return p;
    }
    *got = bytes;
    return device_alloc(bytes);
}

}  // namespace

namespace {

// Scratch of the allocation-free preconditioner entry point
// (pcbhip_operator_precondition_f32_noalloc): the two FP64 coarse vectors
// g = Z^T r and w = cinv g of at most kMaxCoarse values each, 256 KiB, which
// every device the library runs on holds until the process ends.  In host
// code the array's name is not a device address: the address comes from
// hipGetSymbolAddress, once per device with that device current, and is kept
// in the table of noalloc_scratch().
constexpr size_t kNoallocScratchDoubles = 2 * (size_t)kMaxCoarse;
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

// The device address of ``scratch`` on ``device`` (which must be current),
// looked up on the first call on that device and kept.
double* noalloc_scratch(const int device) {
    static std::mutex table_lock;
    static std::vector<double*> table;
    std::lock_guard<std::mutex> guard(table_lock);
    if ((size_t)device >= table.size()) table.resize((size_t)device + 1, nullptr);
    if (!table[(size_t)device]) {
        void* ptr = nullptr;
        check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        if (!ptr) fail(PCBHIP_RUNTIME, "no device address for the preconditioner scratch");
        table[(size_t)device] = static_cast<double*>(ptr);
    }
    return table[(size_t)device];
}

// One rocBLAS handle per device for the coarse factorisations, made on first
// use and kept (its workspace too) for the life of the process; calls take
// the handle's lock and set its stream.
struct SolverHandle {
    std::mutex lock;
    rocblas_handle handle = nullptr;
};

SolverHandle& solver_handle(int device) {
    static std::mutex table_lock;
    static std::vector<std::unique_ptr<SolverHandle>> table;
    std::lock_guard<std::mutex> guard(table_lock);
    if ((size_t)device >= table.size()) table.resize((size_t)device + 1);
    if (!table[(size_t)device]) {
        auto h = std::make_unique<SolverHandle>();
        rocblas_status st;
        {
            DeviceLock lock(device_mutex());
            st = rocblas_create_handle(&h->handle);
        }
        if (st != rocblas_status_success) fail(PCBHIP_RUNTIME, "rocblas_create_handle failed");
        table[(size_t)device] = std::move(h);
    }
    return *table[(size_t)device];
}

// Coarse-size checks shared by the coarse entry points; returns (rows, cols).
std::pair<int, int> coarse_grid(const pcbhip_operator* op, int block, int nc) {
    if (block < 1) fail(PCBHIP_INVALID_INPUT, "block must be a positive node count");
    const long long crows = (op->R + 1 + block - 1) / block, ccols = (op->C + 1 + block - 1) / block;
    const long long expected = (long long)op->layers() * crows * ccols;
    if (nc != expected)
        fail(PCBHIP_INVALID_INPUT, "coarse size " + std::to_string(nc) + " does not match the patches of block " +
                                       std::to_string(block) + " (" + std::to_string(expected) + ")");
    if (nc > kMaxCoarse)
        fail(PCBHIP_INVALID_INPUT, "coarse size " + std::to_string(nc) + " exceeds the limit of " +
                                       std::to_string(kMaxCoarse));
    return {(int)crows, (int)ccols};
}

// Per-device state of the FP32 vector runtime: a stream of its own
// (non-blocking, default priority: no implicit waits on other streams, so it
// never disturbs a server capturing its graph in another thread) and the dot
// product's partial sums and result (kDotGroupsMax + 1 doubles), made on
// first use with the device current and kept.  Every call returns when its
// work is done, so the runtime's calls order with every other call of this
// library; a call holds the lock from its first launch until it has waited.
struct VectorRuntime {
    std::mutex lock;
    hipStream_t stream = nullptr;
    double* scratch = nullptr;
    // The basis projections' partial sums and results (kBasisRowsMax rows),
    // allocated at the first projection.
    double* basis_scratch = nullptr;
};

VectorRuntime& vector_runtime(int device) {
    static std::mutex table_lock;
    static std::vector<std::unique_ptr<VectorRuntime>> table;
    std::lock_guard<std::mutex> guard(table_lock);
    if ((size_t)device >= table.size()) table.resize((size_t)device + 1);
    if (!table[(size_t)device]) {
        auto vr = std::make_unique<VectorRuntime>();
        {
            DeviceLock lock(device_mutex());
            check(hipStreamCreateWithFlags(&vr->stream, hipStreamNonBlocking), "vector stream");
        }
        vr->scratch = static_cast<double*>(device_alloc(sizeof(double) * (kDotGroupsMax + 1)));
        table[(size_t)device] = std::move(vr);
    }
    return *table[(size_t)device];
}

void check_vector_args(int device, const void* a, const void* b, int64_t n) {
    if (device < 0) fail(PCBHIP_INVALID_INPUT, "device must be non-negative");
    if (n < 0) fail(PCBHIP_INVALID_INPUT, "vector length must be non-negative");
    if (n > 0 && (!a || !b)) fail(PCBHIP_INVALID_INPUT, "null vector");
}

// Runs ``body(runtime)`` on the device's vector stream and waits for it.
template <class F>
void on_vector_stream(int device, const char* what, F&& body) {
    set_device(device);
    VectorRuntime& vr = vector_runtime(device);
    std::lock_guard<std::mutex> guard(vr.lock);
    body(vr);
    check(hipStreamSynchronize(vr.stream), what);
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
        // Stencil offsets inside a node plane are 32-bit byte offsets below
        // the buffer range (2^30) of the widest per-node record, the FP64
        // inner state of 32 bytes.
        if ((long long)(rows + 1) * (cols + 1) >= (1LL << 25))
            fail(PCBHIP_INVALID_INPUT, "a node plane must hold fewer than 2^25 nodes");
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
        check(stream_create(&op->stream), "stream create");
        op->zb32 = planes_per_block(slabs, (long long)strips_x(cols) * groups_y(rows, op->rows32), kTargetGroups);
        op->zb64 = planes_per_block(slabs, (long long)strips_x(cols) * groups_y(rows, op->rows64), kTargetGroups);
        op->plain = plain_planes(robin64, slabs + 1, (long long)(rows + 1) * (cols + 1));
        std::vector<double> uvalues64;
        op->uniform = uniform_layers(coef64, slabs, (long long)rows * cols, uvalues64);
        const std::vector<float> uvalues32(uvalues64.begin(), uvalues64.end());
        // On the device an element's coefficients are one record (a, b, c).
        std::vector<double> rec64(3 * (size_t)E);
        std::vector<float> rec32(3 * (size_t)E);
        for (long long e = 0; e < E; ++e)
            for (int d = 0; d < 3; ++d) {
                rec64[(size_t)(3 * e + d)] = coef64[(size_t)(d * E + e)];
                rec32[(size_t)(3 * e + d)] = coef32[(size_t)(d * E + e)];
            }
        const size_t s_c32 = align256(sizeof(float) * 3 * E), s_c64 = align256(sizeof(double) * 3 * E);
        const size_t s_n32 = align256(sizeof(float) * N), s_n64 = align256(sizeof(double) * N);
        const size_t s_u32 = align256(sizeof(float) * uvalues32.size());
        const size_t s_u64 = align256(sizeof(double) * uvalues64.size());
        char* base = static_cast<char*>(stream_alloc(
            op->stream, s_c32 + s_c64 + 2 * s_n32 + 3 * s_n64 + 2 * s_n32 + 2 * s_n64 + s_u32 + s_u64 + 256));
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
        op->nrec32 = reinterpret_cast<float*>(base);
        base += 2 * s_n32;
        op->nrec64 = reinterpret_cast<double*>(base);
        base += 2 * s_n64;
        op->ucoef32 = reinterpret_cast<float*>(base);
        base += s_u32;
        op->ucoef64 = reinterpret_cast<double*>(base);
        base += s_u64;
        int* d_bad = reinterpret_cast<int*>(base);
        check(hipMemcpyAsync(op->ucoef32, uvalues32.data(), sizeof(float) * uvalues32.size(), hipMemcpyHostToDevice,
                             op->stream),
              "upload");
        check(hipMemcpyAsync(op->ucoef64, uvalues64.data(), sizeof(double) * uvalues64.size(), hipMemcpyHostToDevice,
                             op->stream),
              "upload");
        check(hipMemcpyAsync(op->coef32, rec32.data(), sizeof(float) * 3 * E, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->coef64, rec64.data(), sizeof(double) * 3 * E, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->robin32, robin32.data(), sizeof(float) * N, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->robin64, robin64.data(), sizeof(double) * N, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemsetAsync(d_bad, 0, sizeof(int), op->stream), "memset");
        check(launch::diagonal(op->coef64, op->robin64, op->diag64, op->dinv64, op->dinv32, slabs, rows, cols, d_bad,
                               op->stream),
              "diagonal launch");
        check(launch::pack_nodes(op->dinv32, op->robin32, op->dinv64, op->robin64, op->nrec32, op->nrec64, N,
                                 op->stream),
              "node records");
        int bad = 0;
        check(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, op->stream), "copy");
        check(hipStreamSynchronize(op->stream), "operator create");
        if (bad) fail(PCBHIP_INVALID_INPUT, "every free node must have positive conductance");
        *out = op.release();
    });
}

int pcbhip_sheet_operator_create(int device, int layers, int rows, int cols, const double* coef,
                                 const uint8_t* free_nodes, int64_t via_count, const int64_t* via_a,
                                 const int64_t* via_b, const double* via_g, pcbhip_operator** out) {
    return guarded([&] {
        if (!out || !coef || !free_nodes || (via_count > 0 && (!via_a || !via_b || !via_g)))
            fail(PCBHIP_INVALID_INPUT, "null argument");
        *out = nullptr;
        if (layers < 1 || rows < 1 || cols < 1) fail(PCBHIP_INVALID_INPUT, "element grid must be positive");
        if (via_count < 0) fail(PCBHIP_INVALID_INPUT, "via count must be non-negative");
        const long long E = (long long)layers * rows * cols;
        const long long N = (long long)layers * (rows + 1) * (cols + 1);
        if (N > (1LL << 31) - 1 || 2 * via_count > (1LL << 31) - 1)
            fail(PCBHIP_INVALID_INPUT, "node or via count exceeds 2^31 - 1");
        // A via's far end is loaded at its node index as a 32-bit byte offset
        // below the buffer range (2^30) of the widest per-node record, the
        // FP64 inner state of 32 bytes.
        if (N >= (1LL << 25)) fail(PCBHIP_INVALID_INPUT, "a sheet operator must hold fewer than 2^25 nodes");
        std::vector<double> coef64(2 * (size_t)E), robin64((size_t)N);
        std::vector<float> coef32(2 * (size_t)E), robin32((size_t)N);
        for (long long i = 0; i < 2 * E; ++i) {
            if (!std::isfinite(coef[i]) || coef[i] < 0.0)
                fail(PCBHIP_INVALID_INPUT, "element coefficients must be finite and non-negative");
            coef64[i] = coef[i] / 6.0;
            coef32[i] = (float)coef64[i];
        }
        for (long long n = 0; n < N; ++n) {
            robin64[n] = free_nodes[n] ? 0.0 : -1.0;
            robin32[n] = (float)robin64[n];
        }
        // Node-ordered via lists: every via appears at both ends, in via order.
        std::vector<int> ptr;
        std::vector<int> nbr;
        std::vector<double> g64;
        std::vector<float> g32;
        if (via_count > 0) {
            ptr.assign((size_t)N + 1, 0);
            for (long long k = 0; k < via_count; ++k) {
                if (via_a[k] < 0 || via_a[k] >= N || via_b[k] < 0 || via_b[k] >= N || via_a[k] == via_b[k])
                    fail(PCBHIP_INVALID_INPUT, "via endpoints must be two distinct nodes of the mesh");
                if (!std::isfinite(via_g[k]) || via_g[k] <= 0.0)
                    fail(PCBHIP_INVALID_INPUT, "via conductance must be finite and positive");
                ptr[(size_t)via_a[k] + 1] += 1;
                ptr[(size_t)via_b[k] + 1] += 1;
            }
            for (long long n = 0; n < N; ++n) ptr[(size_t)n + 1] += ptr[(size_t)n];
            std::vector<int> fill(ptr.begin(), ptr.end() - 1);
            nbr.resize((size_t)(2 * via_count));
            g64.resize((size_t)(2 * via_count));
            for (long long k = 0; k < via_count; ++k) {
                const int a = (int)via_a[k], b = (int)via_b[k];
                nbr[(size_t)fill[(size_t)a]] = b;
                g64[(size_t)fill[(size_t)a]++] = via_g[k];
                nbr[(size_t)fill[(size_t)b]] = a;
                g64[(size_t)fill[(size_t)b]++] = via_g[k];
            }
            g32.assign(g64.begin(), g64.end());
        }
        set_device(device);
        auto op = std::make_unique<pcbhip_operator>();
        op->device = device;
        op->kind = kSheet;
        op->S = layers;
        op->R = rows;
        op->C = cols;
        op->N = N;
        op->E = E;
        check(stream_create(&op->stream), "stream create");
        op->plain = plain_planes(robin64, layers, (long long)(rows + 1) * (cols + 1));
        const size_t s_c32 = align256(sizeof(float) * 2 * E), s_c64 = align256(sizeof(double) * 2 * E);
        const size_t s_n32 = align256(sizeof(float) * N), s_n64 = align256(sizeof(double) * N);
        char* base = static_cast<char*>(
            stream_alloc(op->stream, s_c32 + s_c64 + 2 * s_n32 + 3 * s_n64 + 2 * s_n32 + 2 * s_n64 + 256));
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
        op->nrec32 = reinterpret_cast<float*>(base);
        base += 2 * s_n32;
        op->nrec64 = reinterpret_cast<double*>(base);
        base += 2 * s_n64;
        int* d_bad = reinterpret_cast<int*>(base);
        if (via_count > 0) {
            const size_t s_ptr = align256(sizeof(int) * ((size_t)N + 1)), s_nbr = align256(sizeof(int) * nbr.size());
            const size_t s_g32 = align256(sizeof(float) * g32.size()), s_g64 = align256(sizeof(double) * g64.size());
            char* vb = static_cast<char*>(stream_alloc(op->stream, s_ptr + s_nbr + s_g32 + s_g64));
            op->via_ptr = reinterpret_cast<int*>(vb);
            op->via_nbr = reinterpret_cast<int*>(vb + s_ptr);
            op->via_g32 = reinterpret_cast<float*>(vb + s_ptr + s_nbr);
            op->via_g64 = reinterpret_cast<double*>(vb + s_ptr + s_nbr + s_g32);
            check(hipMemcpyAsync(op->via_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, op->stream),
                  "upload");
            check(hipMemcpyAsync(op->via_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice, op->stream),
                  "upload");
            check(hipMemcpyAsync(op->via_g32, g32.data(), sizeof(float) * g32.size(), hipMemcpyHostToDevice,
                                 op->stream),
                  "upload");
            check(hipMemcpyAsync(op->via_g64, g64.data(), sizeof(double) * g64.size(), hipMemcpyHostToDevice,
                                 op->stream),
                  "upload");
        }
        check(hipMemcpyAsync(op->coef32, coef32.data(), sizeof(float) * 2 * E, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->coef64, coef64.data(), sizeof(double) * 2 * E, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->robin32, robin32.data(), sizeof(float) * N, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemcpyAsync(op->robin64, robin64.data(), sizeof(double) * N, hipMemcpyHostToDevice, op->stream),
              "upload");
        check(hipMemsetAsync(d_bad, 0, sizeof(int), op->stream), "memset");
        check(launch::sheet_diagonal(op->view(), op->diag64, op->dinv64, op->dinv32, d_bad, op->stream),
              "diagonal launch");
        check(launch::pack_nodes(op->dinv32, op->robin32, op->dinv64, op->robin64, op->nrec32, op->nrec64, N,
                                 op->stream),
              "node records");
        int bad = 0;
        check(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, op->stream), "copy");
        check(hipStreamSynchronize(op->stream), "operator create");
        if (bad) fail(PCBHIP_INVALID_INPUT, "every free node must have positive conductivity coupling");
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
        check(stream_create(&b->stream), "stream create");
        b->slots.resize((size_t)count);
        for (int i = 0; i < count; ++i) {
            SlotHost& s = b->slots[(size_t)i];
            const pcbhip_operator* op = ops[i];
            s.op = op;
            s.view = op->view();
            set_partial_layout(s);
            s.history_rows = s.config.max_outer_iterations;
            // Vectors of the slot in one allocation, x0 behind them.  History
            // room is allocated on demand in pcbhip_batch_set.
            const size_t vec = slot_vector_bytes(s, b->inner64);
            char* base = static_cast<char*>(device_alloc(vec + align256(sizeof(double) * op->N)));
            carve_slot(s, base, b->inner64);
            s.x0 = reinterpret_cast<double*>(base + vec);
            // Zeroed vectors: the kernels load r, p, q, c and x unconditionally and
            // weight them by 0 where the algorithm starts from zero.
            check(hipMemsetAsync(base, 0, vec, b->stream), "memset");
        }
        build_tiles(b.get());
        b->d_slots = static_cast<SlotDev*>(device_alloc(sizeof(SlotDev) * count));
        b->d_states = static_cast<SlotState*>(device_alloc(sizeof(SlotState) * count));
        b->d_configs = static_cast<SlotConfig*>(device_alloc(sizeof(SlotConfig) * count));
        b->d_active = static_cast<int*>(device_alloc(sizeof(int)));
        b->h_active = pinned_alloc<int>(1);
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
            if (s.history) check(device_free(s.history), "hipFree");
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

int pcbhip_batch_set_schedule(pcbhip_batch* b, int replay_steps, int outer_every, int compact_percent) {
    return guarded([&] {
        if (!b) fail(PCBHIP_INVALID_INPUT, "null batch");
        if (replay_steps < 1 || replay_steps > 1024) fail(PCBHIP_INVALID_INPUT, "replay_steps must be in 1..1024");
        if (outer_every < 1 || outer_every > replay_steps)
            fail(PCBHIP_INVALID_INPUT, "outer_every must be in 1..replay_steps");
        if (compact_percent < 1 || compact_percent > 100)
            fail(PCBHIP_INVALID_INPUT, "compact_percent must be in 1..100");
        b->replay_steps = replay_steps;
        b->outer_every = outer_every;
        b->compact_percent = compact_percent;
        b->graph_steps = 0;  // capture again at the next run
    });
}

int pcbhip_batch_set_block(pcbhip_batch* b, int group_size, int mfma) {
    return guarded([&] {
        if (!b) fail(PCBHIP_INVALID_INPUT, "null batch");
        if (group_size < 0 || group_size > kGroupMax)
            fail(PCBHIP_INVALID_INPUT, "group_size must be in 0.." + std::to_string(kGroupMax));
        if (mfma != 0 && mfma != 2 && mfma != 3) fail(PCBHIP_INVALID_INPUT, "engine must be 0, 2 or 3");
        set_device(b->device);
        b->block_group = group_size;
        b->block_mfma = mfma;
        build_tiles(b);
    });
}

int pcbhip_operator_set_engine(pcbhip_operator* op, int mfma) {
    return guarded([&] {
        if (!op) fail(PCBHIP_INVALID_INPUT, "null operator");
        if (mfma != 0 && mfma != 2 && mfma != 3) fail(PCBHIP_INVALID_INPUT, "engine must be 0, 2 or 3");
        op->mfma32 = mfma;
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

int pcbhip_operator_set_coarse(pcbhip_operator* op, int block, int nc, const double* cinv) {
    return guarded([&] {
        if (!op) fail(PCBHIP_INVALID_INPUT, "null operator");
        if (nc == 0) {
            if (op->cinv) op->retired.push_back(op->cinv);
            op->cinv = nullptr;
            op->cblock = op->crows = op->ccols = op->nc = 0;
            return;
        }
        if (!cinv) fail(PCBHIP_INVALID_INPUT, "null coarse inverse");
        const std::pair<int, int> grid = coarse_grid(op, block, nc);
        const size_t count = (size_t)nc * (size_t)nc;
        std::vector<float> values(count);
        for (size_t i = 0; i < count; ++i) {
            if (!std::isfinite(cinv[i])) fail(PCBHIP_INVALID_INPUT, "coarse inverse must be finite");
            values[i] = (float)cinv[i];
        }
        set_device(op->device);
        float* d = static_cast<float*>(stream_alloc(op->stream, sizeof(float) * count));
        hipError_t err = hipMemcpyAsync(d, values.data(), sizeof(float) * count, hipMemcpyHostToDevice, op->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(op->stream);
        if (err != hipSuccess) {
            (void)stream_free(op->stream, d);
            check(err, "coarse upload");
        }
        if (op->cinv) op->retired.push_back(op->cinv);
        op->cinv = d;
        op->cblock = block;
        op->crows = grid.first;
        op->ccols = grid.second;
        op->nc = nc;
    });
}


int pcbhip_operator_set_coarse_matrix(pcbhip_operator* op, int block, int nc, const double* matrix,
                                      double* inverse) {
    return guarded([&] {
        if (!op || !matrix) fail(PCBHIP_INVALID_INPUT, "null argument");
        const std::pair<int, int> grid = coarse_grid(op, block, nc);
        const size_t count = (size_t)nc * (size_t)nc;
        for (size_t i = 0; i < count; ++i)
            if (!std::isfinite(matrix[i])) fail(PCBHIP_INVALID_INPUT, "coarse matrix must be finite");
        set_device(op->device);
        const size_t b64 = align256(sizeof(double) * count);
        // The factor in place, then the FP32 inverse the solves read (kept),
        // and the FP64 inverse when the caller wants it back.
        char* work = static_cast<char*>(stream_alloc(op->stream, 2 * b64 + 256));
        double* packed = reinterpret_cast<double*>(work);
        double* full = inverse ? reinterpret_cast<double*>(work + b64) : nullptr;
        rocblas_int* info = reinterpret_cast<rocblas_int*>(work + 2 * b64);
        float* d = nullptr;
        try {
            d = static_cast<float*>(stream_alloc(op->stream, sizeof(float) * count));
            check(hipMemcpyAsync(packed, matrix, sizeof(double) * count, hipMemcpyHostToDevice, op->stream),
                  "coarse upload");
            rocblas_int h_info[2] = {0, 0};
            {
                SolverHandle& solver = solver_handle(op->device);
                std::lock_guard<std::mutex> guard(solver.lock);
                if (rocblas_set_stream(solver.handle, op->stream) != rocblas_status_success)
                    fail(PCBHIP_RUNTIME, "rocblas_set_stream failed");
                if (rocsolver_dpotrf(solver.handle, rocblas_fill_lower, nc, packed, nc, info) !=
                    rocblas_status_success)
                    fail(PCBHIP_RUNTIME, "rocsolver_dpotrf failed");
                check(hipMemcpyAsync(&h_info[0], info, sizeof(rocblas_int), hipMemcpyDeviceToHost, op->stream),
                      "info");
                check(hipStreamSynchronize(op->stream), "potrf");
                if (h_info[0] != 0)
                    fail(PCBHIP_INVALID_INPUT,
                         "the coarse operator is not positive definite; the fine operator must be SPD for the "
                         "two-level preconditioner");
                if (rocsolver_dpotri(solver.handle, rocblas_fill_lower, nc, packed, nc, info) !=
                    rocblas_status_success)
                    fail(PCBHIP_RUNTIME, "rocsolver_dpotri failed");
                check(hipMemcpyAsync(&h_info[1], info, sizeof(rocblas_int), hipMemcpyDeviceToHost, op->stream),
                      "info");
                check(launch::symmetric_inverse(packed, nc, d, full, op->stream), "symmetric inverse launch");
                if (inverse)
                    check(hipMemcpyAsync(inverse, full, sizeof(double) * count, hipMemcpyDeviceToHost, op->stream),
                          "inverse download");
                check(hipStreamSynchronize(op->stream), "potri");
            }
            if (h_info[1] != 0) fail(PCBHIP_INVALID_INPUT, "the coarse operator is singular");
        } catch (...) {
            if (d) (void)stream_free(op->stream, d);
            (void)stream_free(op->stream, work);
            throw;
        }
        (void)stream_free(op->stream, work);
        if (op->cinv) op->retired.push_back(op->cinv);
        op->cinv = d;
        op->cblock = block;
        op->crows = grid.first;
        op->ccols = grid.second;
        op->nc = nc;
    });
}

int pcbhip_operator_galerkin_f64(const pcbhip_operator* op, int block, int count, const double* v, double* out) {
    return guarded([&] {
        if (!op || !v || !out) fail(PCBHIP_INVALID_INPUT, "null argument");
        if (block < 1) fail(PCBHIP_INVALID_INPUT, "block must be a positive node count");
        if (count < 1) fail(PCBHIP_INVALID_INPUT, "count must be positive");
        OperatorDev view = op->view();
        view.cblock = block;
        view.crows = (op->R + 1 + block - 1) / block;
        view.ccols = (op->C + 1 + block - 1) / block;
        const long long nc = (long long)op->layers() * view.crows * view.ccols;
        if (nc > kMaxCoarse)
            fail(PCBHIP_INVALID_INPUT, "coarse size " + std::to_string(nc) + " exceeds the limit of " +
                                           std::to_string(kMaxCoarse));
        view.nc = (int)nc;
        set_device(op->device);
        const size_t n64 = align256(sizeof(double) * (size_t)op->N);
        const size_t c64 = align256(sizeof(double) * (size_t)nc * (size_t)count);
        std::lock_guard<std::mutex> guard(op->staging_mutex);
        char* base = staging_of(op, std::max(2 * n64 + 2 * c64, 2 * n64));
        double* fine = reinterpret_cast<double*>(base);
        double* y = reinterpret_cast<double*>(base + n64);
        double* dv = reinterpret_cast<double*>(base + 2 * n64);
        double* dout = reinterpret_cast<double*>(base + 2 * n64 + c64);
        const size_t bytes = sizeof(double) * (size_t)nc * (size_t)count;
        check(hipMemcpyAsync(dv, v, bytes, hipMemcpyHostToDevice, op->stream), "copy v");
        for (int k = 0; k < count; ++k) {
            check(launch::coarse_spread(view, dv + (size_t)k * nc, fine, op->stream), "spread launch");
            apply_dev<double>(op, fine, y, 0, op->stream);
            check(launch::coarse_restrict(view, y, dout + (size_t)k * nc, op->stream), "restrict launch");
        }
        check(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, op->stream), "copy out");
        check(hipStreamSynchronize(op->stream), "galerkin");
    });
}

int pcbhip_operator_precondition_f32(const pcbhip_operator* op, const float* r, float* z) {
    return guarded([&] {
        if (!op || !r || !z) fail(PCBHIP_INVALID_INPUT, "null argument");
        set_device(op->device);
        const size_t bytes = sizeof(float) * (size_t)op->N;
        const size_t coarse = align256(sizeof(double) * (size_t)std::max(op->nc, 1));
        std::lock_guard<std::mutex> guard(op->staging_mutex);
        char* base = staging_of(op, std::max(2 * align256(bytes) + 2 * coarse,
                                             2 * align256(sizeof(double) * (size_t)op->N)));
        float* dr = reinterpret_cast<float*>(base);
        float* dz = reinterpret_cast<float*>(base + align256(bytes));
        double* g = reinterpret_cast<double*>(base + 2 * align256(bytes));
        double* w = reinterpret_cast<double*>(base + 2 * align256(bytes) + coarse);
        check(hipMemcpyAsync(dr, r, bytes, hipMemcpyHostToDevice, op->stream), "copy r");
        check(launch::precondition(op->view(), dr, dz, g, w, op->stream), "precondition launch");
        check(hipMemcpyAsync(z, dz, bytes, hipMemcpyDeviceToHost, op->stream), "copy z");
        check(hipStreamSynchronize(op->stream), "precondition");
    });
}

int pcbhip_operator_precondition_f32_noalloc(const pcbhip_operator* op, const float* r, float* z, void* stream) {
    return guarded([&] {
        if (!op || !r || !z) fail(PCBHIP_INVALID_INPUT, "null argument");
        const size_t need = 2 * (size_t)std::max(op->nc, 0);
        if (need > kNoallocScratchDoubles)
            fail(PCBHIP_INVALID_INPUT, "coarse size " + std::to_string(op->nc) +
                                           " exceeds the allocation-free scratch of " +
                                           std::to_string(kNoallocScratchDoubles / 2) + " coarse unknowns");
        set_device(op->device);
        double* g = noalloc_scratch(op->device);
        // Both coarse vectors are written in full before they are read (the
        // restriction writes every patch, the coarse rows every row), so the
        // scratch needs no clearing between calls.
        check(launch::precondition(op->view(), r, z, g, g + kMaxCoarse, stream_of(op, stream)),
              "precondition launch");
    });
}

int pcbhip_batch_rhs_dev(pcbhip_batch* b, int slot, double** rhs_dev) {
    return guarded([&] {
        if (!b || !rhs_dev || slot < 0 || slot >= (int)b->slots.size())
            fail(PCBHIP_INVALID_INPUT, "slot out of range");
        *rhs_dev = b->slots[(size_t)slot].b;
    });
}

namespace {

// Uploads every slot's descriptor and limits when they changed since the
// last upload (history rows are allocated on first use).
void upload_slots(pcbhip_batch* b) {
    if (!b->configs_dirty) return;
    const int count = (int)b->slots.size();
    std::vector<SlotDev> devs((size_t)count);
    std::vector<SlotConfig> cfgs((size_t)count);
    for (int i = 0; i < count; ++i) {
        SlotHost& s = b->slots[(size_t)i];
        if (!s.history) {
            s.history_rows = s.config.max_outer_iterations;
            s.history = static_cast<double*>(device_alloc(sizeof(double) * 4 * (size_t)s.history_rows));
        }
        devs[(size_t)i] = slot_dev(s);
        cfgs[(size_t)i] = slot_config(s.config);
    }
    check(hipMemcpyAsync(b->d_slots, devs.data(), sizeof(SlotDev) * count, hipMemcpyHostToDevice, b->stream),
          "upload slots");
    check(hipMemcpyAsync(b->d_configs, cfgs.data(), sizeof(SlotConfig) * count, hipMemcpyHostToDevice, b->stream),
          "upload configs");
    // The host vectors are the copies' sources: wait before they go.
    check(hipStreamSynchronize(b->stream), "upload");
    b->configs_dirty = false;
}

}  // namespace

int pcbhip_batch_run(pcbhip_batch* b) {
    return guarded([&] {
        if (!b) fail(PCBHIP_INVALID_INPUT, "null batch");
        set_device(b->device);
        const int count = (int)b->slots.size();
        upload_slots(b);
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
        const int steps = b->replay_steps;
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
            if (count > 1 && (long long)(b->tiled_slots - active) * 100 >= (long long)b->compact_percent * b->tiled_slots) {
                // Enough of the tiled slots have finished: drop their tiles.
                host_states.resize((size_t)count);
                check(hipMemcpyAsync(host_states.data(), b->d_states, sizeof(SlotState) * count,
                                     hipMemcpyDeviceToHost, b->stream),
                      "copy states");
                check(hipStreamSynchronize(b->stream), "copy states");
                std::vector<char> keep((size_t)count);
                for (int i = 0; i < count; ++i) keep[(size_t)i] = host_states[(size_t)i].phase != kPhaseDone;
                build_tiles(b, &keep);
                build_graph(b, steps);
            }
        }
        if (b->tiled_slots != count) build_tiles(b);
        b->ran = true;
    });
}

int pcbhip_batch_time_pass(pcbhip_batch* b, int launches, double* avg_us) {
    return guarded([&] {
        if (!b || !avg_us || launches < 1) fail(PCBHIP_INVALID_INPUT, "time_pass needs a batch and launches >= 1");
        if (!b->ran) fail(PCBHIP_INVALID_INPUT, "run the batch before timing its pass");
        set_device(b->device);
        upload_slots(b);
        const int count = (int)b->slots.size();
        std::vector<SlotState> saved((size_t)count);
        check(hipMemcpyAsync(saved.data(), b->d_states, sizeof(SlotState) * count, hipMemcpyDeviceToHost, b->stream),
              "copy states");
        check(hipStreamSynchronize(b->stream), "copy states");
        std::vector<SlotState> timing = saved;
        for (SlotState& st : timing) {
            st.phase = kPhaseInner;
            st.first = 0;
            st.alpha = 0.0;
            st.beta = 0.0;
        }
        check(hipMemcpyAsync(b->d_states, timing.data(), sizeof(SlotState) * count, hipMemcpyHostToDevice, b->stream),
              "states");
        check(hipStreamSynchronize(b->stream), "states");
        hipEvent_t ev[2] = {nullptr, nullptr};
        float ms = 0.0f;
        hipError_t err = hipEventCreate(&ev[0]);
        if (err == hipSuccess) err = hipEventCreate(&ev[1]);
        if (err == hipSuccess) err = hipEventRecord(ev[0], b->stream);
        for (int k = 0; k < launches && err == hipSuccess; ++k)
            err = launch::inner_pass(b->inner64 != 0, b->rows_inner, b->d_slots, b->d_states, b->d_inner, b->d_counts,
                                     std::max(b->n_inner, 1), false, nullptr, b->d_block, b->d_groups, b->n_block,
                                     b->block_mfma, b->coarse(), b->stream);
        if (err == hipSuccess) err = hipEventRecord(ev[1], b->stream);
        if (err == hipSuccess) err = hipEventSynchronize(ev[1]);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, ev[0], ev[1]);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        check(err, "timed passes");
        check(hipMemcpyAsync(b->d_states, saved.data(), sizeof(SlotState) * count, hipMemcpyHostToDevice, b->stream),
              "restore states");
        check(hipStreamSynchronize(b->stream), "restore states");
        *avg_us = 1e3 * (double)ms / launches;
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

int pcbhip_streams_dev(const float* const* in, int reads, float* const* out, int writes, int64_t n_floats,
                       void* stream) {
    return guarded([&] {
        if (!in || reads < 1 || reads > kStreamsMax || writes < 0 || writes > kStreamsMax || (writes > 0 && !out))
            fail(PCBHIP_INVALID_INPUT, "streams needs 1.." + std::to_string(kStreamsMax) + " inputs and 0.." +
                                           std::to_string(kStreamsMax) + " outputs");
        if (n_floats < 4 || n_floats % 4 != 0) fail(PCBHIP_INVALID_INPUT, "streams needs a multiple of 4 floats");
        for (int k = 0; k < reads; ++k)
            if (!in[k]) fail(PCBHIP_INVALID_INPUT, "null input array");
        for (int k = 0; k < writes; ++k)
            if (!out[k]) fail(PCBHIP_INVALID_INPUT, "null output array");
        check(launch::streams(in, reads, out, writes, n_floats, static_cast<hipStream_t>(stream)), "streams launch");
    });
}

int pcbhip_vec_create(int device, int64_t n, float** v) {
    return guarded([&] {
        if (!v) fail(PCBHIP_INVALID_INPUT, "null output");
        check_vector_args(device, v, v, n);
        set_device(device);
        VectorRuntime& vr = vector_runtime(device);
        *v = nullptr;
        const size_t bytes = sizeof(float) * (size_t)std::max<int64_t>(n, 1);
        float* p = static_cast<float*>(device_alloc(bytes));
        std::lock_guard<std::mutex> guard(vr.lock);
        hipError_t err = hipMemsetAsync(p, 0, bytes, vr.stream);
        if (err == hipSuccess) err = hipStreamSynchronize(vr.stream);
        if (err != hipSuccess) {
            (void)device_free(p);
            check(err, "vector clear");
        }
        *v = p;
    });
}

int pcbhip_vec_destroy(int device, float* v) {
    return guarded([&] {
        if (device < 0) fail(PCBHIP_INVALID_INPUT, "device must be non-negative");
        if (!v) return;
        set_device(device);
        check(device_free(v), "hipFree");
    });
}

int pcbhip_vec_write(int device, float* dst, const float* src, int64_t n) {
    return guarded([&] {
        check_vector_args(device, dst, src, n);
        if (n == 0) return;
        on_vector_stream(device, "vector write", [&](VectorRuntime& vr) {
            check(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, vr.stream),
                  "vector write");
        });
    });
}

int pcbhip_vec_read(int device, const float* src, float* dst, int64_t n) {
    return guarded([&] {
        check_vector_args(device, dst, src, n);
        if (n == 0) return;
        on_vector_stream(device, "vector read", [&](VectorRuntime& vr) {
            check(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, vr.stream),
                  "vector read");
        });
    });
}

int pcbhip_vec_copy(int device, const float* src, float* dst, int64_t n) {
    return guarded([&] {
        check_vector_args(device, dst, src, n);
        if (n == 0) return;
        on_vector_stream(device, "vector copy", [&](VectorRuntime& vr) {
            check(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, vr.stream),
                  "vector copy");
        });
    });
}

int pcbhip_vec_dot(int device, const float* x, const float* y, int64_t n, double* out) {
    return guarded([&] {
        if (!out) fail(PCBHIP_INVALID_INPUT, "null output");
        check_vector_args(device, x, y, n);
        if (n == 0) {
            *out = 0.0;
            return;
        }
        on_vector_stream(device, "dot", [&](VectorRuntime& vr) {
            check(launch::dot(x, y, n, vr.scratch, vr.scratch + kDotGroupsMax, vr.stream), "dot launch");
            check(hipMemcpyAsync(out, vr.scratch + kDotGroupsMax, sizeof(double), hipMemcpyDeviceToHost, vr.stream),
                  "dot read");
        });
    });
}

int pcbhip_vec_axpy(int device, float alpha, const float* x, const float* y, float* out, int64_t n) {
    return guarded([&] {
        check_vector_args(device, x, y, n);
        check_vector_args(device, out, out, n);
        if (n == 0) return;
        on_vector_stream(device, "axpy", [&](VectorRuntime& vr) {
            check(launch::axpy(alpha, x, y, out, n, vr.stream), "axpy launch");
        });
    });
}

int pcbhip_vec_divide(int device, const float* num, const float* den, float* out, int64_t n) {
    return guarded([&] {
        check_vector_args(device, num, den, n);
        check_vector_args(device, out, out, n);
        if (n == 0) return;
        on_vector_stream(device, "divide", [&](VectorRuntime& vr) {
            check(launch::divide(num, den, out, n, vr.stream), "divide launch");
        });
    });
}

int pcbhip_vec_project(int device, const float* const* rows, int count, const float* w, int64_t n, double* out) {
    return guarded([&] {
        if (!rows || !out) fail(PCBHIP_INVALID_INPUT, "null argument");
        if (count < 1 || count > kBasisRowsMax)
            fail(PCBHIP_INVALID_INPUT, "a projection takes 1.." + std::to_string(kBasisRowsMax) + " rows");
        for (int k = 0; k < count; ++k) check_vector_args(device, rows[k], w, n);
        if (n == 0) {
            for (int k = 0; k < count; ++k) out[k] = 0.0;
            return;
        }
        on_vector_stream(device, "project", [&](VectorRuntime& vr) {
            if (!vr.basis_scratch)
                vr.basis_scratch = static_cast<double*>(
                    device_alloc(sizeof(double) * (size_t)kBasisRowsMax * (size_t)(kDotGroupsMax + 1)));
            double* results = vr.basis_scratch + (size_t)kBasisRowsMax * kDotGroupsMax;
            check(launch::project(rows, count, w, n, vr.basis_scratch, results, vr.stream), "project launch");
            check(hipMemcpyAsync(out, results, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost, vr.stream),
                  "project read");
        });
    });
}

int pcbhip_vec_combine(int device, const float* const* rows, const double* coef, int count, const float* y,
                       float* out, int64_t n) {
    return guarded([&] {
        if (!rows || !coef) fail(PCBHIP_INVALID_INPUT, "null argument");
        if (count < 1 || count > kBasisRowsMax)
            fail(PCBHIP_INVALID_INPUT, "a combination takes 1.." + std::to_string(kBasisRowsMax) + " rows");
        for (int k = 0; k < count; ++k) {
            check_vector_args(device, rows[k], out, n);
            if (!std::isfinite(coef[k])) fail(PCBHIP_INVALID_INPUT, "combination coefficients must be finite");
        }
        if (n == 0) return;
        on_vector_stream(device, "combine", [&](VectorRuntime& vr) {
            check(launch::combine(rows, coef, count, y, out, n, vr.stream), "combine launch");
        });
    });
}

int pcbhip_operator_precondition_f32_dev(const pcbhip_operator* op, const float* r, float* z) {
    return guarded([&] {
        if (!op || !r || !z) fail(PCBHIP_INVALID_INPUT, "null argument");
        set_device(op->device);
        const size_t coarse = align256(sizeof(double) * (size_t)std::max(op->nc, 1));
        std::lock_guard<std::mutex> guard(op->staging_mutex);
        char* base = staging_of(op, std::max(2 * coarse, 2 * align256(sizeof(double) * (size_t)op->N)));
        double* g = reinterpret_cast<double*>(base);
        double* w = reinterpret_cast<double*>(base + coarse);
        check(launch::precondition(op->view(), r, z, g, w, op->stream), "precondition launch");
        check(hipStreamSynchronize(op->stream), "precondition");
    });
}

int pcbhip_operator_synchronize(const pcbhip_operator* op) {
    return guarded([&] {
        if (!op) fail(PCBHIP_INVALID_INPUT, "null operator");
        set_device(op->device);
        check(hipStreamSynchronize(op->stream), "operator stream");
    });
}

int pcbhip_copy_dev(const float* src, float* dst, int64_t n_floats, void* stream) {
    return guarded([&] {
        if (!src || !dst || n_floats % 4 != 0) fail(PCBHIP_INVALID_INPUT, "copy needs a multiple of 4 floats");
        check(launch::copy(src, dst, n_floats, static_cast<hipStream_t>(stream)), "copy launch");
    });
}

int pcbhip_operator_bytes(const pcbhip_operator* op, int64_t* bytes) {
    return guarded([&] {
        if (!op || !bytes) fail(PCBHIP_INVALID_INPUT, "null argument");
        const size_t e = (size_t)op->E, n = (size_t)op->N;
        const size_t uniform_values = op->kind == kSheet ? 0 : align256(4 * 3 * kMaxUniformLayers) +
                                                                   align256(8 * 3 * kMaxUniformLayers);
        const size_t words = op->kind == kSheet ? 2 : 3;  // coefficients per element
        *bytes = (int64_t)(align256(4 * words * e) + align256(8 * words * e) + align256(4 * n) + 2 * align256(8 * n) +
                           align256(4 * n) + align256(8 * n) + align256(8 * n) + align256(16 * n) + uniform_values);
    });
}

int pcbhip_server_create(int device, int max_slots, int inner_fp64, int steps_per_replay, pcbhip_server** out) {
    return guarded([&] {
        if (!out) fail(PCBHIP_INVALID_INPUT, "null argument");
        *out = nullptr;
        if (max_slots < 1 || max_slots > 65536) fail(PCBHIP_INVALID_INPUT, "max_slots must be in 1..65536");
        if (steps_per_replay < 1 || steps_per_replay > 1024)
            fail(PCBHIP_INVALID_INPUT, "steps_per_replay must be in 1..1024");
        set_device(device);
        auto sv = std::make_unique<pcbhip_server>();
        sv->device = device;
        sv->inner64 = inner_fp64 ? 1 : 0;
        sv->max_slots = max_slots;
        sv->steps = steps_per_replay;
        sv->slots.resize((size_t)max_slots);
        check(stream_create(&sv->stream), "stream create");
        check(stream_create(&sv->copy_stream), "stream create");
        const size_t n = (size_t)max_slots;
        sv->d_slots = static_cast<SlotDev*>(device_alloc(sizeof(SlotDev) * n));
        sv->d_states = static_cast<SlotState*>(device_alloc(sizeof(SlotState) * n));
        sv->d_configs = static_cast<SlotConfig*>(device_alloc(sizeof(SlotConfig) * n));
        sv->d_fresh = static_cast<SlotState*>(device_alloc(sizeof(SlotState) * n));
        sv->d_fresh_index = static_cast<int*>(device_alloc(sizeof(int) * n));
        sv->d_active = static_cast<int*>(device_alloc(sizeof(int)));
        sv->d_counts = static_cast<int*>(device_alloc(6 * sizeof(int)));
        sv->d_tickets = static_cast<int*>(device_alloc(kPassTickets * sizeof(int)));
        for (Staging& st : sv->stage) {
            st.counts = pinned_alloc<int>(6);
            st.devs = pinned_alloc<SlotDev>(n);
            st.cfgs = pinned_alloc<SlotConfig>(n);
            st.fresh = pinned_alloc<SlotState>(n);
            st.fresh_index = pinned_alloc<int>(n);
            check(hipEventCreateWithFlags(&st.used, hipEventDisableTiming), "event create");
        }
        sv->h_states = pinned_alloc<SlotState>(2 * n);
        for (hipEvent_t& e : sv->replay_done) check(hipEventCreate(&e), "event");
        for (hipEvent_t& e : sv->replay_start) check(hipEventCreate(&e), "event");
        // Every slot starts finished, so the finish kernels skip empty slots.
        std::vector<SlotState> done(n);
        for (SlotState& st : done) {
            std::memset(&st, 0, sizeof(SlotState));
            st.phase = kPhaseDone;
        }
        check(hipMemcpyAsync(sv->d_states, done.data(), sizeof(SlotState) * n, hipMemcpyHostToDevice, sv->stream),
              "states");
        check(hipMemsetAsync(sv->d_slots, 0, sizeof(SlotDev) * n, sv->stream), "memset");
        check(hipMemsetAsync(sv->d_configs, 0, sizeof(SlotConfig) * n, sv->stream), "memset");
        check(hipMemsetAsync(sv->d_counts, 0, 6 * sizeof(int), sv->stream), "memset");
        check(hipMemsetAsync(sv->d_tickets, 0, kPassTickets * sizeof(int), sv->stream), "memset");
        check(hipMemsetAsync(sv->d_active, 0, sizeof(int), sv->stream), "memset");
        check(hipStreamSynchronize(sv->stream), "server create");
        *out = sv.release();
    });
}

int pcbhip_server_destroy(pcbhip_server* server) {
    return guarded([&] {
        if (server) {
            set_device(server->device);
            delete server;
        }
    });
}

int pcbhip_server_set_shape(pcbhip_server* sv, int rows_inner, int target_groups) {
    return guarded([&] {
        if (!sv) fail(PCBHIP_INVALID_INPUT, "null server");
        if (rows_inner != 4 && rows_inner != 8) fail(PCBHIP_INVALID_INPUT, "rows_inner must be 4 or 8");
        if (target_groups < 1) fail(PCBHIP_INVALID_INPUT, "target_groups must be positive");
        set_device(sv->device);
        server_drain(sv);
        check(hipStreamSynchronize(sv->stream), "drain");
        sv->rows_inner = rows_inner;
        sv->target_groups = target_groups;
        if (sv->exec) {
            check(graph_release(sv->exec, nullptr), "graph exec destroy");
            sv->exec = nullptr;
        }
        sv->tables_dirty = true;
    });
}

int pcbhip_server_slot_bytes(const pcbhip_server* sv, const pcbhip_operator* op, int max_outer, int64_t* bytes) {
    return guarded([&] {
        if (!sv || !op || !bytes || max_outer < 1) fail(PCBHIP_INVALID_INPUT, "invalid argument");
        SlotHost h;
        h.op = op;
        h.view = op->view();
        set_partial_layout(h);
        *bytes = (int64_t)(slot_vector_bytes(h, sv->inner64) + align256(sizeof(double) * 4 * (size_t)max_outer));
    });
}

int pcbhip_server_submit(pcbhip_server* sv, const pcbhip_operator* op, const double* rhs, const double* x0,
                         const pcbhip_mpir_config* config, int64_t tag, int* slot_out) {
    return guarded([&] {
        if (!sv || !op || !rhs || !config || !slot_out) fail(PCBHIP_INVALID_INPUT, "null argument");
        if (op->device != sv->device) fail(PCBHIP_INVALID_INPUT, "operator and server live on different devices");
        check_config(*config);
        for (long long i = 0; i < op->N; ++i)
            if (!std::isfinite(rhs[i])) fail(PCBHIP_INVALID_INPUT, "rhs must contain only finite values");
        if (x0)
            for (long long i = 0; i < op->N; ++i)
                if (!std::isfinite(x0[i])) fail(PCBHIP_INVALID_INPUT, "initial guess must be finite");
        int index = -1;
        for (int i = 0; i < sv->max_slots && index < 0; ++i)
            if (sv->slots[(size_t)i].status == kSlotFree) index = i;
        if (index < 0) fail(PCBHIP_FULL, "every slot of the server is in use");
        set_device(sv->device);
        ServerSlot& s = sv->slots[(size_t)index];
        SlotHost h;
        h.op = op;
        h.view = op->view();
        h.config = *config;
        set_partial_layout(h);
        h.history_rows = config->max_outer_iterations;
        const size_t vec = slot_vector_bytes(h, sv->inner64);
        const size_t bytes = vec + align256(sizeof(double) * 4 * (size_t)h.history_rows);
        size_t got = 0;
        char* block = static_cast<char*>(server_take_block(sv, bytes, &got));
        carve_slot(h, block, sv->inner64);
        h.history = reinterpret_cast<double*>(block + vec);
        const size_t n_bytes = sizeof(double) * (size_t)op->N;
        // Zeroed vectors (the kernels weight p, c and x by 0 where the
        // algorithm starts from zero, and a reused block holds old values).
        hipError_t err = hipMemsetAsync(block, 0, bytes, sv->copy_stream);
        if (err == hipSuccess) err = hipMemcpyAsync(h.b, rhs, n_bytes, hipMemcpyHostToDevice, sv->copy_stream);
        if (err == hipSuccess && x0) err = hipMemcpyAsync(h.x[0], x0, n_bytes, hipMemcpyHostToDevice, sv->copy_stream);
        if (err == hipSuccess) err = hipStreamSynchronize(sv->copy_stream);
        if (err != hipSuccess) {
            sv->pool.emplace_back(block, got);
            sv->bytes_pooled += got;
            check(err, "slot upload");
        }
        s.h = h;
        s.block_bytes = got;
        s.status = kSlotQueued;
        s.tag = tag;
        s.cap_steps = (long long)(config->max_outer_iterations + 1) *
                          (config->max_inner_iterations + sv->outer_every) + 2LL * sv->outer_every;
        sv->bytes_in_use += got;
        *slot_out = index;
    });
}

int pcbhip_server_step(pcbhip_server* sv, int* running, int* finished) {
    return guarded([&] {
        if (!sv) fail(PCBHIP_INVALID_INPUT, "null server");
        set_device(sv->device);
        if (sv->inflight.size() >= 2) server_retire_oldest(sv);
        std::vector<int> admitted;
        for (int i = 0; i < sv->max_slots; ++i) {
            ServerSlot& s = sv->slots[(size_t)i];
            if (s.status != kSlotQueued) continue;
            s.status = kSlotRunning;
            s.first_replay = sv->seq;
            admitted.push_back(i);
            sv->tables_dirty = true;
        }
        int n_running = 0;
        for (const ServerSlot& s : sv->slots) n_running += s.status == kSlotRunning ? 1 : 0;
        if (n_running == 0) {
            server_drain(sv);
            server_enqueue_copies(sv);
            server_settle_copies(sv);
        } else {
            if (sv->tables_dirty || !sv->exec) {
                server_upload(sv, admitted);
                // server_upload may have drained: slots it finished leave at the next call.
                for (int i : admitted) sv->slots[(size_t)i].first_replay = sv->seq;
            }
            server_enqueue_copies(sv);
            server_launch(sv);
        }
        int r = 0, f = 0;
        for (const ServerSlot& s : sv->slots) {
            r += (s.status == kSlotRunning || s.status == kSlotCopying) ? 1 : 0;
            f += s.status == kSlotFinished ? 1 : 0;
        }
        if (running) *running = r;
        if (finished) *finished = f;
    });
}

int pcbhip_server_finished(const pcbhip_server* sv, int* slots, int64_t* tags, int capacity, int* count) {
    return guarded([&] {
        if (!sv || !count || capacity < 0 || (capacity > 0 && (!slots || !tags)))
            fail(PCBHIP_INVALID_INPUT, "invalid argument");
        int n = 0;
        for (int i = 0; i < sv->max_slots && n < capacity; ++i) {
            if (sv->slots[(size_t)i].status != kSlotFinished) continue;
            slots[n] = i;
            tags[n] = sv->slots[(size_t)i].tag;
            ++n;
        }
        *count = n;
    });
}

int pcbhip_server_collect(pcbhip_server* sv, int slot, double* x, pcbhip_mpir_result* result, double* history) {
    return guarded([&] {
        if (!sv || slot < 0 || slot >= sv->max_slots) fail(PCBHIP_INVALID_INPUT, "slot out of range");
        ServerSlot& s = sv->slots[(size_t)slot];
        if (s.status != kSlotFinished) fail(PCBHIP_INVALID_INPUT, "slot has no finished result");
        set_device(sv->device);
        const SlotState& st = s.final_state;
        // The result copies completed before a replay that has been retired,
        // so the pinned buffer holds the final solution and history.
        if (x) std::memcpy(x, s.host, sizeof(double) * (size_t)s.h.op->N);
        if (history && st.outer > 0) std::memcpy(history, s.host + s.h.op->N, sizeof(double) * 4 * (size_t)st.outer);
        if (result) {
            result->converged = st.converged;
            result->outer_iterations = st.outer;
            result->inner_iterations = st.total_inner;
            result->high_operator_applications = st.high_apps;
            result->low_operator_applications = st.low_apps;
            result->not_spd = st.status == kStatusNotSpd ? 1 : 0;
            result->relative_residual = st.rel_outer;
        }
        // Release: replays still in flight may list the slot, but its device
        // state stays "done" until an admission rewrites it in stream order.
        sv->pool.emplace_back(s.h.block, s.block_bytes);
        sv->bytes_pooled += s.block_bytes;
        sv->bytes_in_use -= s.block_bytes;
        sv->pinned_pool.emplace_back(s.host, s.host_bytes);
        const bool not_spd = st.status == kStatusNotSpd;
        s = ServerSlot{};
        if (not_spd) fail(PCBHIP_RUNTIME, "inner PCG requires a finite symmetric positive-definite operator");
    });
}

int pcbhip_server_stats(const pcbhip_server* sv, pcbhip_server_info* out) {
    return guarded([&] {
        if (!sv || !out) fail(PCBHIP_INVALID_INPUT, "null argument");
        pcbhip_server_info st{};
        for (const ServerSlot& s : sv->slots) {
            st.queued += s.status == kSlotQueued ? 1 : 0;
            st.running += (s.status == kSlotRunning || s.status == kSlotCopying) ? 1 : 0;
            st.finished += s.status == kSlotFinished ? 1 : 0;
            st.free_slots += s.status == kSlotFree ? 1 : 0;
        }
        st.bytes_in_use = (int64_t)sv->bytes_in_use;
        st.bytes_pooled = (int64_t)sv->bytes_pooled;
        st.replays = sv->seq;
        st.graph_builds = sv->graph_builds;
        st.steps_per_replay = sv->steps;
        st.wait_ms = 1e3 * sv->wait_s;
        st.replay_ms = 1e3 * sv->replay_s;
        *out = st;
    });
}

int pcbhip_server_trim(pcbhip_server* sv) {
    return guarded([&] {
        if (!sv) fail(PCBHIP_INVALID_INPUT, "null server");
        set_device(sv->device);
        for (auto& b : sv->pool) check(device_free(b.first), "hipFree");
        sv->pool.clear();
        sv->bytes_pooled = 0;
        for (auto& b : sv->pinned_pool) check(pinned_free(b.first), "pinned free");
        sv->pinned_pool.clear();
    });
}

}  // extern "C"
