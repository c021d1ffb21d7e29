// Device-side data layout shared by the gfx950 kernels and the host runtime
// of the hexahedral Q1 conduction solver.
//
// Nodes of one problem are the flat C-order grid ((z * (R + 1) + y) * (C + 1)
// + x) with x fastest, elements ((ez * R + ey) * C + ex); local corners are
// ordered 4 dz + 2 dy + dx.  This is the reference's own layout
// (cpp/include/pcbcore/thermal/thermal_system.hpp:39-46), so host arrays cross
// the boundary without a permutation.
//
// A "slot" is one right-hand side of one operator: its FP64 solution and
// residual, its inner Krylov vectors and its scalar state all live in HBM, and
// every kernel finds its work through a tile table, so one launch serves one
// solve, a batch of uneven problems or the requests in flight of a server.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcbhip {

// Lanes of a wavefront and the node columns one wave strip writes: lane l
// holds element/node column x0 - 1 + l, the two edge lanes are halo.
constexpr int kWave = 64;
constexpr int kStripCols = kWave - 2;
// Waves per workgroup of the stencil kernels (stacked in y).
constexpr int kWavesPerGroup = 4;
constexpr int kGroupThreads = kWave * kWavesPerGroup;
// Rows per wave strip of the FP64 outer-residual kernel and of the FP64
// inner kernel; the FP32 inner kernel and the standalone apply are built for
// 4 and 8 rows and the tile shape picks one.
constexpr int kRowsOuter = 4;  // multiples of the 4-row reduction cell
constexpr int kRowsInner64 = 4;
// Stencil workgroups wanted per launch before the z sweep is split into
// blocks (256 CUs, several waves each): the standalone apply's default, and
// the fused pass's (three workgroups per CU fit at its 168 VGPRs, so about
// three rounds of them).
constexpr int kTargetGroups = 1024;
constexpr int kPassTargetGroups = 2048;
// Workgroups of the persistent outer-residual kernel.
constexpr int kOuterGroups = 512;
// Workgroups a server launches for the fused PCG pass over its tile table
// (two per CU; the pass takes the live entries by ticket, eight ticket
// counters per pass kernel, one per XCD, for the Jacobi and the two-level
// pass).
constexpr int kPassGroups = 512;
constexpr int kPassTickets = 16;
// Threads of the one-workgroup-per-slot finish kernels.
constexpr int kFinishThreads = 512;

// Reduction cells of the stencil kernels: every inner product is summed per
// (node-plane pair, 16 node rows, one strip of 62 node columns) and stored as
// one partial, whatever tile shape computed it.  A wave sums 4-row groups
// (kCellRows); the workgroup adds the four groups of a cell in order.
constexpr int kCellRows = 4;
constexpr int kCellPlanes = 2;
constexpr int kCellBandRows = kCellRows * kWavesPerGroup;
// Plane pairs one stencil tile may hold (its z block is at most 32 planes).
constexpr int kMaxTilePairs = 16;
// Sums per reduction cell: four of the outer residual kernel and five of
// the fused PCG pass (p.q, r.z, q.z, q.D^-1 q, r.r).
constexpr int kOuterSums = 4;
constexpr int kInnerSums = 5;

__host__ __device__ inline int cell_bands(const int R) { return (R + 1 + kCellBandRows - 1) / kCellBandRows; }

__host__ __device__ inline int cell_count(const int S, const int R, const int C) {
    const int nsx = (C + 1 + kStripCols - 1) / kStripCols;
    const int nzp = (S + 1 + kCellPlanes - 1) / kCellPlanes;
    return nsx * cell_bands(R) * nzp;
}

enum Phase : int { kPhaseOuter = 0, kPhaseInner = 1, kPhaseDone = 2 };
enum Status : int { kStatusOk = 0, kStatusNotSpd = 1 };
// Outer step bookkeeping: no correction pending, a PCG correction pending,
// or a zero correction pending (the inner right-hand side was zero).
enum Pending : int { kPendingNone = 0, kPendingCorrection = 1, kPendingZero = 2 };

// The operator of one problem, read-only on the device.  Coefficients are
// stored pre-divided by 36: with M = W / 6, W = [[2, 1], [1, 2]], the element
// action a U_x u is (a / 36) times integer-weighted sums of edge differences.
// A hexahedral element's three coefficients are one record (a, b, c), so a
// stencil fetches an element row with one load instead of three loads from
// three arrays.  ``robin`` holds the lumped Robin conductance of a free node
// and -1 on a fixed node (the free mask and the Robin term in one array);
// ``nrec`` repeats 1 / diag and robin per node as one record (d, robin), the
// fused pass's one node-operator stream on planes that are not plain (plain
// planes read ``dinv`` alone).
//
// A second operator kind, the layered-sheet DC operator of copper layers
// joined by vias (src/electrical/matrix_free_mpir_fem/pcb.py:575-604): S node
// layers of (R + 1) x (C + 1) nodes, no coupling between layers but the vias,
// coefficients (2, S, R, C) = (c_x / 6, c_y / 6) of the 2-D Q1 element
// c_x U_x + c_y U_y, and per node a CSR list of its vias (other end, g).
enum OperatorKind : int { kHex = 0, kSheet = 1 };

// Whether bit z of a 64-bit node-plane or element-layer mask is set (planes
// and layers from 64 on never are).
__host__ __device__ inline bool mask_bit(const unsigned long long mask, const int z) {
    return z >= 0 && z < 64 && ((mask >> z) & 1ull);
}

// Uniform element layers: every element of layer ez < 64 has the same
// (a, b, c); a dielectric slab of a board is such a layer.  The stencils
// fetch no rows of coefficients there.  With a table of the layers' values
// (the standalone apply, whose source is a kernel argument) the loads go to
// the out-of-range offset and the values come from the table; without one
// (the MPIR kernels, whose operator view sits in memory) every element in
// the grid reads the layer's first element, one cache line per wave.
// Elements outside the grid read zero either way.
constexpr int kMaxUniformLayers = 64;

template <typename T>
struct CoefSource {
    const T* coef;               // (S, R, C, 3) element records (a, b, c)
    unsigned long long uniform;  // bit ez: layer ez is uniform
    const T* values;             // (kMaxUniformLayers, 3) a, b, c of the uniform layers, or null
};

struct OperatorDev {
    int S, R, C, kind;
    long long N;  // nodes
    const float* coef32;    // hex: (S, R, C, 3) records a/36, b/36, c/36; sheet: (2, S, R, C) c_x/6, c_y/6
    const double* coef64;
    const float* robin32;   // (N)
    const double* robin64;
    // Bit z (z < 64): every node of node plane z (sheet: layer z) is free
    // with a Robin term of exactly +0.  The stencils read robin there through
    // an out-of-range offset, which returns that +0 without touching memory.
    unsigned long long plain;
    unsigned long long uniform;  // hex: bit ez, every element of layer ez has the same coefficients
    const float* dinv32;    // (N) 1 / diagonal, 1 on fixed rows
    const double* dinv64;
    const float* nrec32;    // (N, 2) records (1 / diagonal, robin)
    const double* nrec64;
    const int* via_ptr;     // sheet: (N + 1) CSR offsets, or null without vias
    const int* via_nbr;     // sheet: other end of each via entry
    const float* via_g32;   // sheet: via conductance of each entry
    const double* via_g64;
    // Two-level coarse correction (null cinv: Jacobi only).  Patches are
    // cblock x cblock nodes of one node layer; coarse index
    // (layer * crows + y / cblock) * ccols + x / cblock; cinv is the dense
    // FP32 inverse of the coarse Galerkin matrix, nc x nc, row-major.
    const float* cinv;
    int cblock, crows, ccols, nc;
};

// Reduction cells of the sheet kernels: one per (layer, 16 node rows, strip).
__host__ __device__ inline int sheet_cell_count(const int layers, const int R, const int C) {
    const int nsx = (C + 1 + kStripCols - 1) / kStripCols;
    return layers * cell_bands(R) * nsx;
}

// The inner Krylov vectors of one node side by side: the inner-precision
// residual r, operator-times-direction q, direction p and correction c.  A
// pass moves a node's four words with one 16-byte access each way (two for
// FP64), so it streams one state array in and one out instead of four and
// four.
template <typename TI>
struct KNode {
    TI r, q, p, c;
};

// One right-hand side.  ``x64`` and ``state`` are ping-pong pairs: a stencil
// kernel reads one buffer at its halo while neighbouring workgroups write
// the other, so no workgroup ever reads a value another rewrites in the same
// launch.  ``state`` holds (N) KNode<TI> in the inner precision.
struct SlotDev {
    OperatorDev op;
    const double* b64;  // (N) right-hand side
    double* x64[2];     // (N) solution pair
    void* state[2];
    double* part;       // kInnerSums * part_stride partial sums
    double* history;    // 4 * max_outer: (outer, fp64 rel before, inner its, inner rel)
    // Coarse vectors of a two-level slot (nc each): g the patch sums of the
    // vector being restricted, h = cinv g, w = cinv Z^T r for the current r.
    double* cg;
    double* ch;
    double* cw;
    int cells;          // reduction cells of the stencil kernels (see cell_count)
    int part_stride;
};

struct SlotConfig {
    double rtol, atol, inner_tol;
    int max_outer, max_inner;
};

// Mutable scalars of one slot.  Written only by the one-workgroup "finish"
// kernels that follow each compute kernel; read by every later launch.
// ``pcur`` selects the current buffer of the state pair; when an inner solve
// ends, its final correction is the c of the other buffer.
struct SlotState {
    int phase, first, pending, outer;
    int inner_its, total_inner, high_apps, low_apps;
    int status, converged, xcur, pcur;
    int started, pad0, pad1, pad2;
    double rz, rhs_norm, rel_outer, rel_inner;
    double alpha, beta, scale, target;
};

// One workgroup of work: slot, index of the tile within the slot (its
// partial-sum slot), x strip, y group, z block and planes per z block.
struct Tile {
    int slot, index;
    short sx, sy, sz, zb;
};

// Tile tables of the coarse kernels reuse Tile: a restriction tile is one
// patch (sx, sy, sz = coarse column, row, node layer); a coarse row tile is
// kCoarseRows rows of cinv starting at ``index``.
constexpr int kCoarseRows = 4;
// Workgroups of the coarse kernels; each loops over its table's entries.
constexpr int kCoarseGroups = 1024;
// Largest coarse space a slot may carry (its inverse is nc^2 FP32 values).
constexpr int kMaxCoarse = 16384;

// Arrays the multi-stream bandwidth reference takes on each side.
constexpr int kStreamsMax = 16;

// Workgroups of the first stage of the vector runtime's dot product.
constexpr int kDotGroupsMax = 1024;

// Basis rows one Gram-Schmidt projection or combination takes at once.
constexpr int kBasisRowsMax = 64;

// Slots that share one operator, applied together by the block apply
// (at most kGroupMax right-hand sides per group).
constexpr int kGroupMax = 64;
struct SlotGroup {
    int count;
    int slot[kGroupMax];
};

namespace launch {
// Host-side launchers of the kernels (q1hex_kernels.hip).
// mfma selects the element action of FP32 applies: 0 vector FMAs, 2 or 3
// the matrix cores with a two- or three-term bf16 split of the operand.
template <typename T>
hipError_t apply(const CoefSource<T>& coef, const T* robin, unsigned long long plain, const T* x, T* y, int S, int R,
                 int C, int mode, int rows, int zb, int mfma, hipStream_t stream);
hipError_t diagonal(const double* coef, const double* robin, double* diag, double* dinv64, float* dinv32, int S, int R,
                    int C, int* bad, hipStream_t stream);
// The layered-sheet operator: y = A x (mode 0) or y = K x (mode 1), and its
// Jacobi diagonal (vias included).
template <typename T>
hipError_t sheet_apply(const OperatorDev& op, const T* x, T* y, int mode, hipStream_t stream);
hipError_t sheet_diagonal(const OperatorDev& op, double* diag, double* dinv64, float* dinv32, int* bad,
                          hipStream_t stream);
// The node records (1 / diagonal, robin) of either operator kind, in both
// precisions, from the separate arrays.
hipError_t pack_nodes(const float* dinv32, const float* robin32, const double* dinv64, const double* robin64,
                      float* nrec32, double* nrec64, long long n, hipStream_t stream);
hipError_t copy(const float* src, float* dst, long long n_floats, hipStream_t stream);
// Bandwidth reference: 1..kStreamsMax input and 0..kStreamsMax output arrays
// of n_floats (a multiple of 4) streamed together.
hipError_t streams(const float* const* in, int reads, float* const* out, int writes, long long n_floats,
                   hipStream_t stream);
// FP32 vector runtime: dot (FP64, an order fixed by n; ``partial`` holds
// kDotGroupsMax doubles, ``out`` one), out = y + alpha x, out = num / den.
int dot_groups(long long n);
hipError_t dot(const float* x, const float* y, long long n, double* partial, double* out, hipStream_t stream);
hipError_t axpy(float alpha, const float* x, const float* y, float* out, long long n, hipStream_t stream);
hipError_t divide(const float* num, const float* den, float* out, long long n, hipStream_t stream);
// GMRES basis products on 1..kBasisRowsMax rows of n floats: out[k] = rows[k]
// . w in FP64, each with the dot's bits (``partial`` holds count *
// kDotGroupsMax doubles, ``out`` count); out = y + sum_k coef[k] rows[k] in
// FP64, rounded once (y may be null, out may alias y or a row).
hipError_t project(const float* const* rows, int count, const float* w, long long n, double* partial, double* out,
                   hipStream_t stream);
hipError_t combine(const float* const* rows, const double* coef, int count, const float* y, float* out, long long n,
                   hipStream_t stream);
// One outer step (FP64 residual and decision) for the slots in their outer
// phase, and one fused Jacobi-PCG pass for the slots in their inner phase.
// ``counts`` (device) holds the live lengths of the outer, inner and block
// tile tables (entries 0, 1 and 3) and of the coarse restriction and coarse
// row tables (entries 4 and 5).  The grids may be longer than the live
// tables (workgroups past the count return at once), so a captured graph
// survives changes of the tables; with ``loop`` the inner grid may also be
// shorter and its workgroups take the live inner tiles by ticket from
// ``tickets`` (kPassTickets device ints, zero before the first pass; the
// finish kernel of every step zeroes them again).  Block tiles (a
// slot group per tile, ``groups``) are launched when grid_block > 0, with
// the element action block_mfma (0, 2, 3 as in apply).
// Two-level slots: ``coarse`` describes the coarse kernels' tables (counts
// entries 4 and 5) and which pass kernels run.  With plain = false no slot
// is Jacobi-only; with grid_patches = 0 none is two-level.
struct CoarseLaunch {
    const Tile* patches = nullptr;  // restriction tiles
    const Tile* rows = nullptr;     // coarse row tiles
    int grid_patches = 0, grid_rows = 0;
    bool plain = true;
};
hipError_t outer_step(bool inner64, int n_slots, const SlotDev* slots, SlotState* states, const SlotConfig* configs,
                      const Tile* outer_tiles, const int* counts, int grid_outer, int* active,
                      const CoarseLaunch& coarse, hipStream_t stream);
hipError_t inner_step(bool inner64, int rows_inner, int n_slots, const SlotDev* slots, SlotState* states,
                      const SlotConfig* configs, const Tile* inner_tiles, const int* counts, int grid_inner, bool loop,
                      int* tickets, const Tile* block_tiles, const SlotGroup* groups, int grid_block, int block_mfma,
                      int* active, const CoarseLaunch& coarse, hipStream_t stream);
// The fused PCG pass alone (inner_step without its coarse kernels and its
// finish kernel).
hipError_t inner_pass(bool inner64, int rows_inner, const SlotDev* slots, const SlotState* states,
                      const Tile* inner_tiles, const int* counts, int grid_inner, bool loop, int* tickets,
                      const Tile* block_tiles, const SlotGroup* groups, int grid_block, int block_mfma,
                      const CoarseLaunch& coarse, hipStream_t stream);
// z = M^-1 r of one operator in FP32 (Jacobi, plus the coarse correction
// when the operator has one); g and w are nc-long FP64 scratch vectors.
hipError_t precondition(const OperatorDev& op, const float* r, float* z, double* g, double* w, hipStream_t stream);
// The Galerkin assembly's transfers in FP64 for the patches op describes:
// fine = Z v, and g = Z^T y (free nodes only).
hipError_t coarse_spread(const OperatorDev& op, const double* v, double* fine, hipStream_t stream);
hipError_t coarse_restrict(const OperatorDev& op, const double* y, double* g, hipStream_t stream);
// The full symmetric inverse from the lower triangle potri leaves
// (column-major ``packed``): FP32 row-major, and FP64 too when inverse64 is set.
hipError_t symmetric_inverse(const double* packed, int n, float* inverse32, double* inverse64, hipStream_t stream);
// states[index[i]] = fresh[i] for i < count (device arrays).
hipError_t scatter_states(const int* index, const SlotState* fresh, SlotState* states, int count, hipStream_t stream);
}  // namespace launch

}  // namespace pcbhip
