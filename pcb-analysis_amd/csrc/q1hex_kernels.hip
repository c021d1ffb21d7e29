// gfx950 kernels of the matrix-free hexahedral Q1 conduction solve.
//
// The operator is the reference's node-owned Q1 action
// (src/thermal/matrix_free_mpir_fem/native/hex_q1.cpp:90-193,
// src/thermal/matrix_free_mpir_fem/cuda.py:19-94): for every free node the
// sum over its adjacent elements of (a U_x + b U_y + c U_z) restricted to free
// columns, plus the lumped Robin conductance times the node value; fixed rows
// are the identity.  Instead of the reference's 64 weight products per
// element and node, the element action is sum-factorised: with
// U_x = M(z) M(y) S(x), S u is an edge difference and M = W / 6 a 2-point
// average, so each element costs edge differences, integer-weighted sums and
// three scalings (the 1/36 is folded into the stored coefficients).
//
// Layout and mapping (see q1hex_device.h): one wavefront owns a strip of 64
// node columns (two halo lanes) by ROWS node rows and sweeps the z node
// planes, keeping two planes in registers.  x-neighbours move between lanes
// with cross-lane shuffles, y-neighbours stay in the same lane's registers,
// so the stencil needs no LDS and no barrier; loads are 64-wide coalesced
// rows of x.  Every output node is written by exactly one lane; reductions
// are wave butterflies, a fixed-order sum over the workgroup's waves, one
// partial per workgroup and a fixed-order sum over a slot's partials by the
// workgroup that finishes last, so results never depend on scheduling, batch
// composition or launch order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "q1hex_device.h"

namespace pcbhip {

// ------------------------------------------------------- buffer access
//
// Every stencil access is a raw buffer load or store: a wave-uniform base
// (the node plane or element layer, in scalar registers) plus a per-lane
// 32-bit element offset that is fixed for the whole sweep.  Offsets of nodes
// outside the grid are set to kOutside, beyond the descriptor's range, so the
// hardware returns zero for them without a branch or a select.
constexpr unsigned kOutside = 0x10000000u;
constexpr int kRange = 0x3FFFFFFF;

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc rsrc(const void* base, const int bytes = kRange) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
__device__ __forceinline__ float ld(const Rsrc r, const unsigned off, float) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off * 4u, 0, 0));
}
__device__ __forceinline__ double ld(const Rsrc r, const unsigned off, double) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off * 8u, 0, 0));
}
template <typename T>
__device__ __forceinline__ T ldt(const T* base, const unsigned off) {
    return ld(rsrc(base), off, T());
}
__device__ __forceinline__ void store(const float v, const Rsrc r, const unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off * 4u, 0, 0);
}
__device__ __forceinline__ void store(const double v, const Rsrc r, const unsigned off) {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off * 8u, 0, 0);
}
template <typename T>
__device__ __forceinline__ void stt(T* base, const unsigned off, const T v) {
    store(v, rsrc(base), off);
}

// Byte offset of the KNode at node offset ``off``: node offsets at or past
// 2^30 / sizeof(KNode) (kOutside among them) are clamped to that bound, which
// lands at byte 2^30, past kRange, so they read zero as the scalar loads do
// (kOutside times 16 or 32 would wrap to 0 in 32 bits).  Operator creation
// keeps real offsets below the bound.
template <typename TI>
__device__ __forceinline__ unsigned knode_bytes(const unsigned off) {
    constexpr unsigned kSize = sizeof(KNode<TI>);
    constexpr unsigned kBound = (1u << 30) / kSize;
    return (off < kBound ? off : kBound) * kSize;
}
// The loaded vectors are bit-cast whole: this compiler's __builtin_bit_cast
// of a single vector element (v.y) reads element 0.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ KNode<float> ld_knode(const KNode<float>* base, const unsigned off) {
    const f32x4 v = __builtin_bit_cast(
        f32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc(base), knode_bytes<float>(off), 0, 0));
    return {v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ KNode<double> ld_knode(const KNode<double>* base, const unsigned off) {
    const Rsrc r = rsrc(base);
    const unsigned b = knode_bytes<double>(off);
    const f64x2 lo = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, b, 0, 0));
    const f64x2 hi = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, b + 16u, 0, 0));
    return {lo.x, lo.y, hi.x, hi.y};
}
__device__ __forceinline__ void st_knode(KNode<float>* base, const unsigned off, const KNode<float>& v) {
    const f32x4 w{v.r, v.q, v.p, v.c};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, w), rsrc(base), knode_bytes<float>(off), 0, 0);
}
__device__ __forceinline__ void st_knode(KNode<double>* base, const unsigned off, const KNode<double>& v) {
    const f64x2 lo{v.r, v.q}, hi{v.p, v.c};
    const Rsrc d = rsrc(base);
    const unsigned b = knode_bytes<double>(off);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, lo), d, b, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, hi), d, b + 16u, 0, 0);
}
// One word of a KNode (0 r, 1 q, 2 p, 3 c).
template <typename TI>
__device__ __forceinline__ TI ld_knode_word(const KNode<TI>* base, const unsigned off, const int word) {
    return ld(rsrc(reinterpret_cast<const TI*>(base) + word), knode_bytes<TI>(off) / (unsigned)sizeof(TI), TI());
}

// Operator records: an element's (a, b, c) and a node's (1 / diag, robin),
// one access each.  Record offsets are clamped as KNode offsets are, so
// kOutside reads zero.
template <typename T, int WORDS>
__device__ __forceinline__ unsigned rec_bytes(const unsigned off) {
    constexpr unsigned kSize = WORDS * sizeof(T);
    constexpr unsigned kBound = ((1u << 30) + kSize - 1) / kSize;
    return (off < kBound ? off : kBound) * kSize;
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ void ld_rec3(const Rsrc r, const unsigned off, float& a, float& b, float& c) {
    const f32x3 v = __builtin_bit_cast(f32x3, __builtin_amdgcn_raw_buffer_load_b96(r, rec_bytes<float, 3>(off), 0, 0));
    a = v.x;
    b = v.y;
    c = v.z;
}
__device__ __forceinline__ void ld_rec3(const Rsrc r, const unsigned off, double& a, double& b, double& c) {
    const unsigned o = rec_bytes<double, 3>(off);
    const f64x2 ab = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0));
    c = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, o + 16u, 0, 0));
    a = ab.x;
    b = ab.y;
}
// Two consecutive words at byte offset ``bytes`` (at or past kRange: zero).
__device__ __forceinline__ void ld_pair(const float* base, const unsigned bytes, float& a, float& b) {
    const f32x2 v = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(rsrc(base), bytes, 0, 0));
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ void ld_pair(const double* base, const unsigned bytes, double& a, double& b) {
    const f64x2 v = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rsrc(base), bytes, 0, 0));
    a = v.x;
    b = v.y;
}
__device__ __forceinline__ float fmadd(const float a, const float b, const float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(const double a, const double b, const double c) {
    return __builtin_fma(a, b, c);
}

// Neighbour lanes by a full-wave DPP shift (a VALU move, no LDS crossbar):
// lane_next gives lane l the value of lane l + 1 (wave_shl:1), lane_prev that
// of lane l - 1 (wave_shr:1); the lane with no source (63 and 0) keeps its
// own value.
__device__ __forceinline__ int lane_next_bits(const int b) {
    return __builtin_amdgcn_update_dpp(b, b, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ int lane_prev_bits(const int b) {
    return __builtin_amdgcn_update_dpp(b, b, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ float lane_next(const float v) {
    return __builtin_bit_cast(float, lane_next_bits(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float lane_prev(const float v) {
    return __builtin_bit_cast(float, lane_prev_bits(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double lane_next(const double v) {
    typedef int i2 __attribute__((ext_vector_type(2)));
    i2 b = __builtin_bit_cast(i2, v);
    b.x = lane_next_bits(b.x);
    b.y = lane_next_bits(b.y);
    return __builtin_bit_cast(double, b);
}
__device__ __forceinline__ double lane_prev(const double v) {
    typedef int i2 __attribute__((ext_vector_type(2)));
    i2 b = __builtin_bit_cast(i2, v);
    b.x = lane_prev_bits(b.x);
    b.y = lane_prev_bits(b.y);
    return __builtin_bit_cast(double, b);
}

// acc + a b accumulated in FP64.  For FP32 operands the FP64 product is
// exact, so one fused multiply-add rounds exactly as multiply-then-add does;
// FP64 operands keep the separate multiply and add.
__device__ __forceinline__ double dot_add(const double acc, const float a, const float b) {
    return __builtin_fma((double)a, (double)b, acc);
}
__device__ __forceinline__ double dot_add(const double acc, const double a, const double b) { return acc + a * b; }

template <typename T>
struct NodeVal {
    T u;      // operand value at the node (unmasked)
    T robin;  // Robin conductance of a free node, negative on a fixed node
};

// What the fused PCG pass carries from a node's load to its emit: the new
// direction (the operand), the Robin word, the updated residual, the Jacobi
// inverse and the updated correction.
template <typename T>
struct InnerVal {
    T u;
    T robin;
    T r;
    T d;
    T c;  // the updated correction (owned nodes)
};

// The same with the preconditioned residual z = D^-1 r + Z w of a
// two-level slot (the coarse term read where the node is loaded).
template <typename T>
struct InnerValZ {
    T u;
    T robin;
    T r;
    T d;
    T c;
    T z;
};

// A stencil's node loader takes (plane base, lane offset, owned) and may
// take the node's strip row and node layer too (a two-level pass needs its
// patch); row -1 marks a node reached through a via, found from its index.
//
// A loader is either a callable that loads and computes in one step, or a
// two-phase loader with fetch(...) (the raw loads only) and make(raw) (the
// arithmetic only): the stencil then issues every load of a plane step
// before any of them is waited for.
// std::declval for device code (unevaluated contexts only; never defined).
template <class T>
__host__ __device__ T&& any_value() noexcept;
template <class L, class = void>
struct HasFetch : std::false_type {};
template <class L>
struct HasFetch<L, std::void_t<decltype(any_value<L&>().fetch(0LL, 0u, false, 0, 0))>> : std::true_type {};

template <class Load>
__device__ __forceinline__ auto fetch_node(Load& load, const long long pb, const unsigned off, const bool owned,
                                           const int row, const int layer) {
    if constexpr (HasFetch<Load>::value)
        return load.fetch(pb, off, owned, row, layer);
    else if constexpr (std::is_invocable_v<Load&, long long, unsigned, bool, int, int>)
        return load(pb, off, owned, row, layer);
    else
        return load(pb, off, owned);
}
template <class Load, class Raw>
__device__ __forceinline__ auto make_node(Load& load, const Raw& raw) {
    if constexpr (HasFetch<Load>::value)
        return load.make(raw);
    else
        return raw;
}
template <class Load>
__device__ __forceinline__ auto call_load(Load& load, const long long pb, const unsigned off, const bool owned,
                                          const int row, const int layer) {
    return make_node(load, fetch_node(load, pb, off, owned, row, layer));
}

// The operand with fixed columns zeroed, for any node-value type with u and robin.
template <class V>
__device__ __forceinline__ auto masked(const V& v) -> decltype(v.u) {
    using T = decltype(v.u);
    return v.robin >= T(0) ? v.u : T(0);
}

// ------------------------------------------------- MFMA element action
//
// The element action as a matrix product on the gfx950 matrix cores:
// P = [U_x; U_y; U_z] X, with the 24 x 8 stack of 36 U (entries +-1, +-2,
// +-4: exact in bf16) against 8 corner values of 32 elements, then
// y = (a/36) P_x + (b/36) P_y + (c/36) P_z per corner.  The FP32 corner
// values are split exactly into three bf16 terms (hi + mid + lo = v), the
// products are exact and v_mfma_f32_32x32x16_bf16 accumulates in FP32, so
// the action keeps FP32 accuracy.  Operand layout (checked on the device):
// lane l holds A[l % 32][8 (l / 32) + i] and B[8 (l / 32) + i][l % 32],
// i = 0..7, and D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] in register r.
// With K = 16 = [hi | mid] in one product and [lo | 0] in a second, the
// rows of D are (direction r / 4, corner 4 (l / 32) + r % 4) and its
// columns the 32 elements.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// 36 U_d[c][k] for direction d (0: x, 1: y, 2: z), corners c, k = 4dz+2dy+dx.
__device__ __forceinline__ float unit36(const int d, const int c, const int k) {
    const int sz = ((c >> 2) ^ (k >> 2)) & 1, sy = ((c >> 1) ^ (k >> 1)) & 1, sx = (c ^ k) & 1;
    const float wz = sz ? 1.f : 2.f, wy = sy ? 1.f : 2.f, wx = sx ? 1.f : 2.f;
    const float ez = sz ? -1.f : 1.f, ey = sy ? -1.f : 1.f, ex = sx ? -1.f : 1.f;
    return d == 0 ? wz * wy * ex : (d == 1 ? wz * ey * wx : ez * wy * wx);
}

__device__ __forceinline__ bf16x8 unit_operand() {
    const int m = threadIdx.x & 31;
    bf16x8 a;
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = (__bf16)(m < 24 ? unit36(m >> 3, m & 7, i) : 0.f);
    return a;
}

// v_permlane32_swap of a and b: lo = [a lanes 0..31, b lanes 0..31],
// hi = [a lanes 32..63, b lanes 32..63] (checked on the device).
__device__ __forceinline__ void pair_halves(const unsigned a, const unsigned b, unsigned& lo, unsigned& hi) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    lo = r[0];
    hi = r[1];
}

__device__ __forceinline__ unsigned bf16_bits(const __bf16 v) { return (unsigned)__builtin_bit_cast(unsigned short, v); }

// Matrix-core operand dwords of one node row of one plane: the exact
// three-term bf16 split of the lane's node value m and of its x neighbour n
// (corners dx = 0 and dx = 1 of the lane's element), packed as the bf16 pair
// of one operand dword, with the lane halves already arranged for the two
// products: q[0] / q[1] the [hi | mid] dwords of products A / B, q[2] / q[3]
// the [lo | 0] dwords.  Each node row of a plane is prepared once and serves
// the two element rows that touch it.
template <int TERMS>
__device__ __forceinline__ void plane_operands(const float m, const float n, unsigned (&q)[4]) {
    const __bf16 hm = (__bf16)m, hn = (__bf16)n;
    const float rm = m - (float)hm, rn = n - (float)hn;
    const __bf16 mm = (__bf16)rm, mn = (__bf16)rn;
    const unsigned ph = bf16_bits(hm) | (bf16_bits(hn) << 16);
    const unsigned pm = bf16_bits(mm) | (bf16_bits(mn) << 16);
    pair_halves(ph, pm, q[0], q[1]);
    if constexpr (TERMS == 3) {
        const __bf16 lm = (__bf16)(rm - (float)mm), ln = (__bf16)(rn - (float)mn);
        const unsigned pl = bf16_bits(lm) | (bf16_bits(ln) << 16);
        pair_halves(pl, 0u, q[2], q[3]);
    }
}

// Accumulates the action of the elements of strip row j into the corner
// accumulators lower[t], upper[t] (t = 2dy+dx).  lo0/lo1 and hi0/hi1 are
// the operand dwords (plane_operands) of node rows j and j + 1 of the lower
// and upper plane, i.e. corners (0,1), (2,3), (4,5), (6,7).  Product A
// covers the elements of lanes 0..31, product B those of lanes 32..63.
template <int TERMS>
__device__ __forceinline__ void element_row_mfma(const bf16x8 unit, const unsigned (&lo0)[4],
                                                 const unsigned (&lo1)[4], const unsigned (&hi0)[4],
                                                 const unsigned (&hi1)[4], const float a, const float b,
                                                 const float c, float (&lower)[4], float (&upper)[4]) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 b1a = {lo0[0], lo1[0], hi0[0], hi1[0]};
    const u4 b1b = {lo0[1], lo1[1], hi0[1], hi1[1]};
    floatx16 da = {}, db = {};
    da = __builtin_amdgcn_mfma_f32_32x32x16_bf16(unit, __builtin_bit_cast(bf16x8, b1a), da, 0, 0, 0);
    db = __builtin_amdgcn_mfma_f32_32x32x16_bf16(unit, __builtin_bit_cast(bf16x8, b1b), db, 0, 0, 0);
    if constexpr (TERMS == 3) {
        const u4 b2a = {lo0[2], lo1[2], hi0[2], hi1[2]};
        const u4 b2b = {lo0[3], lo1[3], hi0[3], hi1[3]};
        da = __builtin_amdgcn_mfma_f32_32x32x16_bf16(unit, __builtin_bit_cast(bf16x8, b2a), da, 0, 0, 0);
        db = __builtin_amdgcn_mfma_f32_32x32x16_bf16(unit, __builtin_bit_cast(bf16x8, b2b), db, 0, 0, 0);
    }
    // Coefficients of the elements each product covers, in both lane halves.
    unsigned ua, ub, va, vb, wa, wb;
    pair_halves(__float_as_uint(a), __float_as_uint(a), ua, ub);
    pair_halves(__float_as_uint(b), __float_as_uint(b), va, vb);
    pair_halves(__float_as_uint(c), __float_as_uint(c), wa, wb);
    const float aa = __uint_as_float(ua), ab = __uint_as_float(ub);
    const float ba = __uint_as_float(va), bb = __uint_as_float(vb);
    const float ca = __uint_as_float(wa), cb = __uint_as_float(wb);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const float ya = fmadd(aa, da[t], fmadd(ba, da[4 + t], ca * da[8 + t]));
        const float yb = fmadd(ab, db[t], fmadd(bb, db[4 + t], cb * db[8 + t]));
        // A holds lower corners in lanes 0..31 and upper corners in 32..63
        // (elements 0..31), B the same for elements 32..63.
        unsigned lw, uw;
        pair_halves(__float_as_uint(ya), __float_as_uint(yb), lw, uw);
        lower[t] = __uint_as_float(lw);
        upper[t] = __uint_as_float(uw);
    }
}

// The sum-factorised Q1 action over one wave strip.
//
// ``load(base, off, owned)`` returns the operand at flat node base + off
// (``base`` the wave-uniform start of the node plane, ``off`` the lane's
// offset in it, kOutside outside the grid, where loads return zero); it is
// called for the ROWS + 2 rows of every plane.  ``owned`` marks the nodes
// this lane writes.  ``emit(base, off, row, conduction, value, parity)``
// receives, for every owned node (``row`` its row in the strip, ``parity``
// the plane's z & 1 as a std::integral_constant), the conduction sum K u
// restricted to free columns (no Robin term); ``plane_done(z)`` follows every
// output plane, in uniform control flow.
//
// DOWN sweeps the z block from its top plane down.  Neighbouring z blocks of
// a column run side by side in opposite directions, so the planes they share
// are read by both at the same time and the second read is served by the
// XCD's L2.  Each output node's sum is the same in both directions: the
// element layer below and the element layer above are each summed from zero
// (their element rows in order, the corner terms by fused multiply-adds) and
// the two partials then added, a missing layer contributing +0.
template <typename T, int ROWS, int MFMA = 0, bool DOWN = false, bool TABLE = false, bool GROUPED = true, class Load,
          class Emit, class PlaneDone>
__device__ __forceinline__ void strip_apply(const CoefSource<T>& cs, const int S, const int R, const int C,
                                            const int x0, const int y0, const int zlo, const int zhi, Load&& load,
                                            Emit&& emit, PlaneDone&& plane_done) {
    const int lane = threadIdx.x & (kWave - 1);
    const int xc = x0 - 1 + lane;
    const bool col_in = xc >= 0 && xc <= C;
    const bool ecol_in = xc >= 0 && xc < C && lane < kWave - 1;
    const bool out_lane = lane >= 1 && lane <= kWave - 2 && col_in;
    const long long NC = C + 1;
    const long long NR = R + 1;

    // Lane offsets of the strip's node rows inside a plane and of its element
    // rows inside a layer.  Only the first row's offset is kept in a
    // register (kOutside in a column outside the grid); row i adds i rows
    // to it, and a row outside the grid (a wave-uniform test) is kOutside.
    // Offsets of outside columns are then kOutside + i (C + 1), which every
    // load treats as outside (at or past kOutside).
    // A one-phase loader (the standalone apply) keeps every row's offset in
    // registers instead, which its register budget allows.
    constexpr bool kRebuildOffsets = HasFetch<std::decay_t<Load>>::value;
    unsigned nbase = col_in ? (unsigned)((y0 - 1) * (int)NC + xc) : kOutside;
    unsigned ebase = ecol_in ? (unsigned)((y0 - 1) * C + xc) : kOutside;
    auto row_noff = [&](const int i) -> unsigned {
        const int y = y0 - 1 + i;
        return (y >= 0 && y <= R) ? nbase + (unsigned)(i * (int)NC) : kOutside;
    };
    auto row_eoff = [&](const int j) -> unsigned {
        const int ye = y0 - 1 + j;
        return (ye >= 0 && ye < R) ? ebase + (unsigned)(j * C) : kOutside;
    };
    [[maybe_unused]] unsigned noff_kept[ROWS + 2], eoff_kept[ROWS + 1];
    if constexpr (!kRebuildOffsets) {
#pragma unroll
        for (int i = 0; i < ROWS + 2; ++i) noff_kept[i] = row_noff(i);
#pragma unroll
        for (int j = 0; j < ROWS + 1; ++j) eoff_kept[j] = row_eoff(j);
    }
    auto noff = [&](const int i) -> unsigned {
        if constexpr (kRebuildOffsets)
            return row_noff(i);
        else
            return noff_kept[i];
    };
    auto eoff = [&](const int j) -> unsigned {
        if constexpr (kRebuildOffsets)
            return row_eoff(j);
        else
            return eoff_kept[j];
    };

    // Raw loads of a node (what fetch returns) and the node value made from
    // them (NodeVal, InnerVal, InnerValZ).
    using Raw = std::decay_t<decltype(fetch_node(load, 0LL, 0u, false, 0, 0))>;
    using V = std::decay_t<decltype(make_node(load, any_value<const Raw&>()))>;
    T mlo[ROWS + 2], nlo[ROWS + 2], mhi[ROWS + 2], nhi[ROWS + 2];
    V vlo[ROWS], vhi[ROWS];
    // The partial of the layer swept last that belongs to the next output
    // plane: the layer below it going up, the layer above it going down.
    T heldL[ROWS], heldR[ROWS];
    [[maybe_unused]] bf16x8 unit;
    [[maybe_unused]] unsigned qlo[ROWS + 2][4], qhi[ROWS + 2][4];
    if constexpr (MFMA != 0) unit = unit_operand();

    // Planes zlo - 1 and zhi are loaded for the elements between them and the
    // block's output planes but belong to neighbouring z blocks.
    auto fetch_plane = [&](const int z, Raw (&raw)[ROWS + 2]) {
        const long long pb = (long long)z * NR * NC;
        const bool out_plane = z >= zlo && z < zhi;
#pragma unroll
        for (int i = 0; i < ROWS + 2; ++i) {
            const bool owned = out_plane && out_lane && i >= 1 && i <= ROWS && y0 - 1 + i <= R;
            raw[i] = fetch_node(load, pb, noff(i), owned, i, z);
        }
    };
    auto make_plane = [&](const Raw (&raw)[ROWS + 2], T (&m)[ROWS + 2], T (&n)[ROWS + 2], V (&v)[ROWS],
                          unsigned (&q)[ROWS + 2][4]) {
#pragma unroll
        for (int i = 0; i < ROWS + 2; ++i) {
            const V nv = make_node(load, raw[i]);
            m[i] = masked(nv);
            if (i >= 1 && i <= ROWS) v[i - 1] = nv;
        }
        // The x neighbours (lane_next) are kept only for the matrix-core
        // operands; the vector action forms them where it uses them, so they
        // hold no registers across steps.
        if constexpr (MFMA != 0) {
#pragma unroll
            for (int i = 0; i < ROWS + 2; ++i) n[i] = lane_next(m[i]);
#pragma unroll
            for (int i = 0; i < ROWS + 2; ++i) plane_operands<MFMA>(m[i], n[i], q[i]);
        }
    };

    // Output plane z: its two layer partials (L for the corners right of the
    // node's element column, R for those left of it, from the lane below).
    auto finish_plane = [&](const int z, const T (&La)[ROWS], const T (&Ra)[ROWS], const T (&Lb)[ROWS],
                            const T (&Rb)[ROWS], const V (&v)[ROWS]) {
        const long long pb = (long long)z * NR * NC;
        auto rows = [&](auto parity) {
#pragma unroll
            for (int i = 0; i < ROWS; ++i) {
                const T L = La[i] + Lb[i];
                const T Rt = Ra[i] + Rb[i];
                const T total = L + lane_prev(Rt);
                if (out_lane && y0 + i <= R) emit(pb, noff(i + 1), i, total, v[i], parity);
            }
        };
        if (z & 1)
            rows(std::integral_constant<int, 1>{});
        else
            rows(std::integral_constant<int, 0>{});
        plane_done(z);
    };

    // The three coefficient rows of element layer ez (loads only).
    auto fetch_layer = [&](const int ez, T (&ca)[ROWS + 1], T (&cb)[ROWS + 1], T (&cc)[ROWS + 1]) {
        // Uniform layers as described at CoefSource.
        const Rsrc rl = rsrc(cs.coef + 3 * (long long)ez * R * C);
        const bool uniform = mask_bit(cs.uniform, ez);
        if constexpr (TABLE) {
            const T* uv = cs.values + 3 * (ez & (kMaxUniformLayers - 1));
            const T ua = uv[0], ub = uv[1], uc = uv[2];
#pragma unroll
            for (int j = 0; j < ROWS + 1; ++j) {
                const unsigned off = uniform ? kOutside : eoff(j);
                T la, lb, lc;
                ld_rec3(rl, off, la, lb, lc);
                const bool in = eoff(j) < kOutside;
                ca[j] = uniform ? (in ? ua : T(0)) : la;
                cb[j] = uniform ? (in ? ub : T(0)) : lb;
                cc[j] = uniform ? (in ? uc : T(0)) : lc;
            }
        } else {
            // In-grid offsets are below kOutside (a power of two), so masking
            // with it maps them to the first element and leaves kOutside.
            const unsigned keep = uniform ? kOutside : ~0u;
#pragma unroll
            for (int j = 0; j < ROWS + 1; ++j) ld_rec3(rl, eoff(j) & keep, ca[j], cb[j], cc[j]);
        }
    };

    // Element layer ez between the lo (ez) and hi (ez + 1) planes in
    // registers, from its coefficient rows: its part of the lo plane (aL, aR)
    // and of the hi plane (bL, bR), each summed from zero.
    auto layer = [&](const T (&ca)[ROWS + 1], const T (&cb)[ROWS + 1], const T (&cc)[ROWS + 1], T (&aL)[ROWS],
                     T (&aR)[ROWS], T (&bL)[ROWS], T (&bR)[ROWS]) {
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
            aL[i] = T(0);
            aR[i] = T(0);
            bL[i] = T(0);
            bR[i] = T(0);
        }
        if constexpr (MFMA != 0) {
            static_assert(sizeof(T) == 4 && (MFMA == 2 || MFMA == 3), "the matrix-core action is FP32, 2 or 3 terms");
#pragma unroll
            for (int j = 0; j < ROWS + 1; ++j) {
                float lower[4], upper[4];
                element_row_mfma<MFMA>(unit, qlo[j], qlo[j + 1], qhi[j], qhi[j + 1], ca[j], cb[j], cc[j], lower, upper);
                if (j >= 1) {
                    aL[j - 1] += lower[0];
                    aR[j - 1] += lower[1];
                    bL[j - 1] += upper[0];
                    bR[j - 1] += upper[1];
                }
                if (j < ROWS) {
                    aL[j] += lower[2];
                    aR[j] += lower[3];
                    bL[j] += upper[2];
                    bR[j] += upper[3];
                }
            }
        } else {
            // Per node row: x differences in both planes, z differences in both
            // columns and their 2-point x average.  Every rounding is explicit
            // (the file is built without contraction): 2 a + b is one fused op
            // and exact in its doubling, and each corner accumulates its three
            // terms with fused multiply-adds, so the bits do not depend on the
            // tile shape that computes an element.
            T dxl[ROWS + 2], dxh[ROWS + 2], zw0[ROWS + 2], zw1[ROWS + 2];
            T nlo[ROWS + 2], nhi[ROWS + 2];
#pragma unroll
            for (int i = 0; i < ROWS + 2; ++i) {
                nlo[i] = lane_next(mlo[i]);
                nhi[i] = lane_next(mhi[i]);
            }
#pragma unroll
            for (int i = 0; i < ROWS + 2; ++i) {
                dxl[i] = nlo[i] - mlo[i];
                dxh[i] = nhi[i] - mhi[i];
                const T dz = mhi[i] - mlo[i];
                const T dzn = nhi[i] - nlo[i];
                zw0[i] = fmadd(T(2), dz, dzn);
                zw1[i] = fmadd(T(2), dzn, dz);
            }

#pragma unroll
            for (int j = 0; j < ROWS + 1; ++j) {
                // a U_x: y average of the x differences, then z average: tx[z][y].
                const T wl0 = fmadd(T(2), dxl[j], dxl[j + 1]);
                const T wl1 = fmadd(T(2), dxl[j + 1], dxl[j]);
                const T wh0 = fmadd(T(2), dxh[j], dxh[j + 1]);
                const T wh1 = fmadd(T(2), dxh[j + 1], dxh[j]);
                const T tx00 = fmadd(T(2), wl0, wh0);
                const T tx01 = fmadd(T(2), wl1, wh1);
                const T tx10 = fmadd(T(2), wh0, wl0);
                const T tx11 = fmadd(T(2), wh1, wl1);
                // b U_y: y differences in both columns, x average, z average: ty[z][x].
                const T dyl0 = mlo[j + 1] - mlo[j];
                const T dyl1 = nlo[j + 1] - nlo[j];
                const T dyh0 = mhi[j + 1] - mhi[j];
                const T dyh1 = nhi[j + 1] - nhi[j];
                const T vl0 = fmadd(T(2), dyl0, dyl1);
                const T vl1 = fmadd(T(2), dyl1, dyl0);
                const T vh0 = fmadd(T(2), dyh0, dyh1);
                const T vh1 = fmadd(T(2), dyh1, dyh0);
                const T ty00 = fmadd(T(2), vl0, vh0);
                const T ty01 = fmadd(T(2), vl1, vh1);
                const T ty10 = fmadd(T(2), vh0, vl0);
                const T ty11 = fmadd(T(2), vh1, vl1);
                // c U_z: z differences, x average (per node row), y average: tz[y][x].
                const T tz00 = fmadd(T(2), zw0[j], zw0[j + 1]);
                const T tz01 = fmadd(T(2), zw1[j], zw1[j + 1]);
                const T tz10 = fmadd(T(2), zw0[j + 1], zw0[j]);
                const T tz11 = fmadd(T(2), zw1[j + 1], zw1[j]);
                const T a = ca[j], b = cb[j], c = cc[j];
                // Corner (z, y, x) receives s(x) a tx[z][y] + s(y) b ty[z][x] + s(z) c tz[y][x],
                // s(0) = -1, s(1) = +1.  Row j is output row j - 1, row j + 1 is output row j.
                if (j >= 1) {
                    aL[j - 1] = fmadd(-a, tx00, fmadd(-b, ty00, fmadd(-c, tz00, aL[j - 1])));
                    aR[j - 1] = fmadd(a, tx00, fmadd(-b, ty01, fmadd(-c, tz01, aR[j - 1])));
                    bL[j - 1] = fmadd(-a, tx10, fmadd(-b, ty10, fmadd(c, tz00, bL[j - 1])));
                    bR[j - 1] = fmadd(a, tx10, fmadd(-b, ty11, fmadd(c, tz01, bR[j - 1])));
                }
                if (j < ROWS) {
                    aL[j] = fmadd(-a, tx01, fmadd(b, ty00, fmadd(-c, tz10, aL[j])));
                    aR[j] = fmadd(a, tx01, fmadd(b, ty01, fmadd(-c, tz11, aR[j])));
                    bL[j] = fmadd(-a, tx11, fmadd(b, ty10, fmadd(c, tz10, bL[j])));
                    bR[j] = fmadd(a, tx11, fmadd(b, ty11, fmadd(c, tz11, bR[j])));
                }
            }
        }
    };


    // Output planes [zlo, zhi) need element layers zlo - 1 .. zhi - 1 (those
    // that exist); the layer outside the block only completes a plane.  Each
    // step fetches one plane and one layer's coefficients.  GROUPED: every
    // load of a step is issued before the step's arithmetic: the scheduling
    // barrier keeps the compiler from sinking loads towards their uses, which
    // left the two-phase loaders of the MPIR kernels three dependent round
    // trips to memory per step.  The standalone apply, whose loader is one
    // phase, measured about 2 % faster without the barrier.
    const int first_layer = zlo > 0 ? zlo - 1 : 0;
    const int last_layer = zhi - 1 < S - 1 ? zhi - 1 : S - 1;
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
        heldL[i] = T(0);
        heldR[i] = T(0);
    }
    // Step k handles element layer ez(k) and fetches plane pz(k): going up,
    // ez = first_layer + k and pz = ez + 1; going down, ez = last_layer - k
    // and pz = ez.
    const int steps = last_layer - first_layer + 1;
    // The arithmetic of step k: the plane it fetched becomes the new plane
    // in registers, its layer's partials are formed, and the plane the layer
    // completes is emitted.
    auto step_body = [&](const int ez, const Raw (&raw)[ROWS + 2], const T (&ca)[ROWS + 1], const T (&cb)[ROWS + 1],
                         const T (&cc)[ROWS + 1]) {
        T aL[ROWS], aR[ROWS], bL[ROWS], bR[ROWS];
        if constexpr (!DOWN) {
            make_plane(raw, mhi, nhi, vhi, qhi);
            layer(ca, cb, cc, aL, aR, bL, bR);
            if (ez >= zlo) finish_plane(ez, heldL, heldR, aL, aR, vlo);
#pragma unroll
            for (int i = 0; i < ROWS + 2; ++i) {
                mlo[i] = mhi[i];
                nlo[i] = nhi[i];
                if constexpr (MFMA != 0) {
#pragma unroll
                    for (int w = 0; w < 4; ++w) qlo[i][w] = qhi[i][w];
                }
            }
#pragma unroll
            for (int i = 0; i < ROWS; ++i) {
                vlo[i] = vhi[i];
                heldL[i] = bL[i];
                heldR[i] = bR[i];
            }
        } else {
            make_plane(raw, mlo, nlo, vlo, qlo);
            layer(ca, cb, cc, aL, aR, bL, bR);
            if (ez + 1 < zhi) finish_plane(ez + 1, bL, bR, heldL, heldR, vhi);
#pragma unroll
            for (int i = 0; i < ROWS + 2; ++i) {
                mhi[i] = mlo[i];
                nhi[i] = nlo[i];
                if constexpr (MFMA != 0) {
#pragma unroll
                    for (int w = 0; w < 4; ++w) qhi[i][w] = qlo[i][w];
                }
            }
#pragma unroll
            for (int i = 0; i < ROWS; ++i) {
                vhi[i] = vlo[i];
                heldL[i] = aL[i];
                heldR[i] = aR[i];
            }
        }
    };
    {
        Raw r0[ROWS + 2];
        fetch_plane(DOWN ? last_layer + 1 : first_layer, r0);
        if constexpr (DOWN)
            make_plane(r0, mhi, nhi, vhi, qhi);
        else
            make_plane(r0, mlo, nlo, vlo, qlo);
    }
    for (int k = 0; k < steps; ++k) {
        const int ez = DOWN ? last_layer - k : first_layer + k;
        Raw raw[ROWS + 2];
        T ca[ROWS + 1], cb[ROWS + 1], cc[ROWS + 1];
        fetch_layer(ez, ca, cb, cc);
        fetch_plane(DOWN ? ez : ez + 1, raw);
        if constexpr (GROUPED) __builtin_amdgcn_sched_barrier(0);
        step_body(ez, raw, ca, cb, cc);
        // Two-phase loaders (the MPIR kernels): the row offsets are formed
        // again in each step from the two bases rather than held in
        // registers across the sweep, which keeps the fused FP32 pass at
        // three waves per SIMD.
        if constexpr (kRebuildOffsets) asm volatile("" : "+v"(nbase), "+v"(ebase));
    }
    const T none[ROWS] = {};
    if constexpr (!DOWN) {
        if (zhi == S + 1) finish_plane(S, heldL, heldR, none, none, vlo);
    } else {
        if (zlo == 0) finish_plane(0, none, none, heldL, heldR, vhi);
    }
}

// The layered-sheet DC action over one wave strip of one copper layer
// (src/electrical/matrix_free_mpir_fem/pcb.py:575-604,
// native/layered_dc_q1.cpp): 2-D Q1 elements c_x U_x + c_y U_y with
// U_x = M(y) S(x), U_y = S(y) M(x), sum-factorised as in the hexahedral
// stencil (x differences averaged in y with weights 2, 1 and the 1/6 folded
// into the coefficients, and likewise for y), four node rows per wave, plus
// every owned node's vias: sum g (w_n - w_other) over its CSR entries, in
// their fixed order, with w the masked operand.  ``load`` is the hexahedral
// stencil's: it is also called at flat index (0, other) for a via's far end.
template <typename T, class Load, class Emit>
__device__ __forceinline__ void sheet_apply(const T* __restrict__ coef, const int* __restrict__ via_ptr,
                                            const int* __restrict__ via_nbr, const T* __restrict__ via_g,
                                            const int layers, const int R, const int C, const int layer, const int x0,
                                            const int y0, Load&& load, Emit&& emit) {
    constexpr int ROWS = kCellRows;
    const int lane = threadIdx.x & (kWave - 1);
    const int xc = x0 - 1 + lane;
    const bool col_in = xc >= 0 && xc <= C;
    const bool ecol_in = xc >= 0 && xc < C && lane < kWave - 1;
    const bool out_lane = lane >= 1 && lane <= kWave - 2 && col_in;
    const long long NC = C + 1;
    const long long E = (long long)layers * R * C;
    const long long pb = (long long)layer * (R + 1) * NC;
    unsigned noff[ROWS + 2], eoff[ROWS + 1];
#pragma unroll
    for (int i = 0; i < ROWS + 2; ++i) {
        const int y = y0 - 1 + i;
        noff[i] = (col_in && y >= 0 && y <= R) ? (unsigned)(y * (int)NC + xc) : kOutside;
    }
#pragma unroll
    for (int j = 0; j < ROWS + 1; ++j) {
        const int ye = y0 - 1 + j;
        eoff[j] = (ecol_in && ye >= 0 && ye < R) ? (unsigned)(ye * C + xc) : kOutside;
    }
    using V = std::decay_t<decltype(call_load(load, 0LL, 0u, false, 0, 0))>;
    T m[ROWS + 2], n[ROWS + 2];
    V v[ROWS];
#pragma unroll
    for (int i = 0; i < ROWS + 2; ++i) {
        const bool owned = out_lane && i >= 1 && i <= ROWS && y0 - 1 + i <= R;
        const V nv = call_load(load, pb, noff[i], owned, i, layer);
        m[i] = masked(nv);
        if (i >= 1 && i <= ROWS) v[i - 1] = nv;
    }
#pragma unroll
    for (int i = 0; i < ROWS + 2; ++i) n[i] = lane_next(m[i]);
    const T* layer_coef = coef + (long long)layer * R * C;
    const Rsrc rx = rsrc(layer_coef), ry = rsrc(layer_coef + E);
    T Lacc[ROWS], Racc[ROWS];
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
        Lacc[i] = T(0);
        Racc[i] = T(0);
    }
#pragma unroll
    for (int j = 0; j < ROWS + 1; ++j) {
        const T cx = ld(rx, eoff[j], T()), cy = ld(ry, eoff[j], T());
        const T dx0 = n[j] - m[j], dx1 = n[j + 1] - m[j + 1];
        const T dy0 = m[j + 1] - m[j], dy1 = n[j + 1] - n[j];
        const T tx0 = fmadd(T(2), dx0, dx1), tx1 = fmadd(T(2), dx1, dx0);
        const T ty0 = fmadd(T(2), dy0, dy1), ty1 = fmadd(T(2), dy1, dy0);
        // Corner (y, x) receives s(x) c_x tx[y] + s(y) c_y ty[x], s(0) = -1,
        // s(1) = +1; node row j is output row j - 1.
        if (j >= 1) {
            Lacc[j - 1] = fmadd(-cx, tx0, fmadd(-cy, ty0, Lacc[j - 1]));
            Racc[j - 1] = fmadd(cx, tx0, fmadd(-cy, ty1, Racc[j - 1]));
        }
        if (j < ROWS) {
            Lacc[j] = fmadd(-cx, tx1, fmadd(cy, ty0, Lacc[j]));
            Racc[j] = fmadd(cx, tx1, fmadd(cy, ty1, Racc[j]));
        }
    }
#pragma unroll
    for (int i = 0; i < ROWS; ++i) {
        T total = Lacc[i] + lane_prev(Racc[i]);
        if (out_lane && y0 + i <= R) {
            if (via_ptr) {
                const long long node = pb + noff[i + 1];
                const int lo = via_ptr[node], hi = via_ptr[node + 1];
                const T own = masked(v[i]);
                for (int k = lo; k < hi; ++k) {
                    const V far = call_load(load, 0LL, (unsigned)via_nbr[k], false, -1, -1);
                    total = fmadd(via_g[k], own - masked(far), total);
                }
            }
            emit(pb, noff[i + 1], i, total, v[i], std::integral_constant<int, 0>{});
        }
    }
}

// Workgroups are dealt to the eight XCDs round-robin; give each XCD a
// contiguous range of tiles so neighbouring strips (which share halo rows)
// meet in the same L2.
__device__ __forceinline__ int xcd_tile(const int block, const int count) {
    constexpr int kXcds = 8;
    const int q = count / kXcds, r = count % kXcds;
    const int k = block % kXcds, j = block / kXcds;
    return k * q + (k < r ? k : r) + j;
}

// Sweep direction of z block sz of a column of nsz blocks: the even blocks
// go down and the odd ones up, so every pair of neighbouring blocks meets at
// the planes they share at the same time (both at their start or both at
// their end).  A column of one block goes up.
__device__ __forceinline__ bool sweep_down(const int sz, const int nsz) { return nsz > 1 && (sz & 1) == 0; }

// strip_apply in the direction of z block sz of nsz.
template <typename T, int ROWS, int MFMA = 0, bool TABLE = false, bool GROUPED = true, class Load, class Emit,
          class PlaneDone>
__device__ __forceinline__ void strip_sweep(const int sz, const int nsz, const CoefSource<T>& coef, const int S,
                                            const int R, const int C, const int x0, const int y0, const int zlo,
                                            const int zhi, Load&& load, Emit&& emit, PlaneDone&& plane_done) {
    if (sweep_down(sz, nsz))
        strip_apply<T, ROWS, MFMA, true, TABLE, GROUPED>(coef, S, R, C, x0, y0, zlo, zhi, load, emit, plane_done);
    else
        strip_apply<T, ROWS, MFMA, false, TABLE, GROUPED>(coef, S, R, C, x0, y0, zlo, zhi, load, emit, plane_done);
}


// ------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset >>= 1) v += __shfl_xor(v, offset);
    return v;
}

// Reduction cells of the stencil kernels.  Every inner product is summed per
// cell of (node-plane pair, 16 node rows, one 62-column strip): each lane sums
// each plane of the pair from zero (its rows in order, into ``t``) and adds
// the two plane sums (``v``), so the z direction of the sweep does not matter
// (the one addition of the two planes commutes); each wave sums its 4-row
// groups (then a wave butterfly) into a table in LDS, and at the end of the
// tile the workgroup adds the four groups of each cell in row order and
// stores one partial per cell.  The cells do
// not depend on the tile shape (rows per wave, planes per z block) or on
// which other slots share the launch, so a slot's bits depend only on its own
// data.  Partials become visible to the finish kernel at the launch
// boundary; no workgroup needs a device-scope fence.
template <int NV, int ROWS>
struct CellSums {
    static constexpr int kGroups = ROWS / kCellRows;  // 4-row groups per wave
    static constexpr int kEntries = kWavesPerGroup * kGroups;
    double t[NV][kGroups];  // per-lane sums of the current plane
    int planes;             // planes of the current pair folded so far

    __device__ static double (&table())[kMaxTilePairs][NV][kEntries] {
        __shared__ double sums[kMaxTilePairs][NV][kEntries];
        return sums;
    }
    // Per-lane sums of the pair's planes folded so far, kept in LDS (each
    // lane its own words) so they hold no registers during the sweep.
    __device__ static double (&pair())[NV][kGroups][kGroupThreads] {
        __shared__ double held[NV][kGroups][kGroupThreads];
        return held;
    }
    __device__ void clear() {
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int g = 0; g < kGroups; ++g) t[k][g] = 0.0;
        planes = 0;
    }
    // The plane's sums join the pair: the first plane's as they are, the
    // second's by one addition.
    __device__ void fold() {
        const int lane = threadIdx.x;
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int g = 0; g < kGroups; ++g) {
                double& v = pair()[k][g][lane];
                v = planes == 0 ? t[k][g] : v + t[k][g];
                t[k][g] = 0.0;
            }
        ++planes;
    }
    __device__ void store_groups(const int pair_index, const int groups) {
        const int wave = threadIdx.x / kWave;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            if (g >= groups) break;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const double s = wave_sum(pair()[k][g][threadIdx.x]);
                if ((threadIdx.x & (kWave - 1)) == 0) table()[pair_index][k][wave * kGroups + g] = s;
            }
        }
        clear();
    }
    // After output plane z of a tile whose z block starts at plane zlo (even),
    // in either direction: the pair is complete after its second plane (the
    // last plane S alone when S is even).
    __device__ void plane_done(const int z, const int zlo, const int S) {
        fold();
        if (planes == ((z | 1) <= S ? 2 : 1)) store_groups(z / kCellPlanes - zlo / kCellPlanes, kGroups);
    }
    // Sheet kernels: the wave's one 4-row group of its layer.
    __device__ void sheet_done() {
        fold();
        store_groups(0, 1);
    }
    // End of a tile, reached by every wave of the workgroup: the cells of
    // ``pairs`` plane pairs from pair pz0 (sheet: one layer, pz0 = layer) and
    // of the workgroup's rows from y0 on.  Group j of a cell is table entry
    // (c * 4 + j) * step; groups that start past row R hold no nodes.
    __device__ void flush(const int pairs, const int pz0, const int y0, const int sx, const int R, const int C,
                          const int cells, const int step, double* __restrict__ part, const int stride) {
        __syncthreads();
        const int nsx = (C + 1 + kStripCols - 1) / kStripCols;
        const int nb = cell_bands(R);
        for (int i = threadIdx.x; i < pairs * NV * cells; i += kGroupThreads) {
            const int c = i % cells, k = (i / cells) % NV, p = i / (cells * NV);
            const int band = y0 / kCellBandRows + c;
            if (band >= nb) continue;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < kWavesPerGroup; ++j) {
                const int group = c * kWavesPerGroup + j;
                if (y0 + group * kCellRows <= R) s += table()[p][k][group * step];
            }
            part[k * stride + ((pz0 + p) * nb + band) * nsx + sx] = s;
        }
        __syncthreads();
    }
    __device__ void flush_hex(const int zlo, const int zhi, const int y0, const int sx, const int R, const int C,
                              double* __restrict__ part, const int stride) {
        flush((zhi - zlo + 1) / kCellPlanes, zlo / kCellPlanes, y0, sx, R, C, kGroups, 1, part, stride);
    }
    __device__ void flush_sheet(const int layer, const int y0, const int sx, const int R, const int C,
                                double* __restrict__ part, const int stride) {
        flush(1, layer, y0, sx, R, C, 1, kGroups, part, stride);
    }
};

// Fixed-order sum of ``count`` partials of one slot by one workgroup of
// THREADS threads: thread t adds partials t, t + THREADS, t + 2 THREADS, ...
// in that order (kUnroll loads per sum in flight before the adds), then a
// wave butterfly, and a butterfly over the wave sums.  The result is for
// thread 0.
template <int NV, int THREADS>
__device__ void slot_sum(const double* __restrict__ part, const int stride, const int count, double (&total)[NV]) {
    constexpr int kWaves = THREADS / kWave;
    constexpr int kUnroll = 8;
    __shared__ double waves[kWaves][NV];
    const int tid = threadIdx.x;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
#pragma unroll 1
    for (int i = tid; i < count; i += kUnroll * THREADS) {
        double v[NV][kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int idx = i + u * THREADS;
#pragma unroll
            for (int k = 0; k < NV; ++k) v[k][u] = idx < count ? part[k * stride + idx] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
#pragma unroll
            for (int k = 0; k < NV; ++k) s[k] += v[k][u];
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double w = wave_sum(s[k]);
        if ((tid & (kWave - 1)) == 0) waves[tid / kWave][k] = w;
    }
    __syncthreads();
    // The wave sums by a butterfly over the first wave (a fixed order too).
    static_assert(kWaves <= kWave, "one lane per wave sum");
    if (tid < kWave) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const double t = wave_sum(tid < kWaves ? waves[tid][k] : 0.0);
            if (tid == 0) total[k] = t;
        }
    }
}

// --------------------------------------------------- standalone applies

// y = A x (mode 0: fixed rows identity, Robin on the diagonal) or y = K x
// (mode 1: unconstrained stiffness, every node free, no Robin).
template <typename T, int ROWS, int MODE, int MFMA>
__global__ __launch_bounds__(kGroupThreads) void k_apply(const CoefSource<T> coef, const T* __restrict__ robin,
                                                          const unsigned long long plain, const T* __restrict__ x,
                                                          T* __restrict__ y, const int S,
                                                          const int R, const int C, const int nsx, const int nsy,
                                                          const int zb) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // Tiles in (y group, x strip, z block) order: the z blocks of a column
    // are dealt to one XCD side by side.
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int nsz = (S + zb) / zb;
    const int sz = t % nsz;
    const int sx = (t / nsz) % nsx;
    const int sy = t / (nsz * nsx);
    const int x0 = sx * kStripCols;
    const int y0 = (sy * kWavesPerGroup + wave) * ROWS;
    const int zlo = sz * zb;
    const int zhi = zlo + zb < S + 1 ? zlo + zb : S + 1;
    if (y0 > R) return;
    strip_sweep<T, ROWS, MFMA, true, false>(
        sz, nsz, coef, S, R, C, x0, y0, zlo, zhi,
        [&](const long long pb, const unsigned off, const bool, const int, const int z) -> NodeVal<T> {
            if (MODE == 1) return {ldt(x + pb, off), T(0)};
            return {ldt(x + pb, off), ldt(robin + pb, mask_bit(plain, z) ? kOutside : off)};
        },
        [&](const long long pb, const unsigned off, const int, const T conduction, const NodeVal<T>& v, auto) {
            stt(y + pb, off, v.robin >= T(0) ? conduction + v.robin * v.u : v.u);
        },
        [](const int) {});
}

// Standalone sheet apply (mode 0: system with fixed rows identity; mode 1:
// unconstrained K, every node free).  Tile = (strip, 16-row group, layer).
template <typename T, int MODE>
__global__ __launch_bounds__(kGroupThreads) void k_sheet_apply(const OperatorDev op, const T* __restrict__ coef,
                                                               const T* __restrict__ robin,
                                                               const T* __restrict__ via_g,
                                                               const T* __restrict__ x, T* __restrict__ y,
                                                               const int nsx, const int nsy) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int sx = t % nsx;
    const int sy = (t / nsx) % nsy;
    const int layer = t / (nsx * nsy);
    const int x0 = sx * kStripCols;
    const int y0 = (sy * kWavesPerGroup + wave) * kCellRows;
    if (y0 > op.R) return;
    sheet_apply<T>(
        coef, op.via_ptr, op.via_nbr, via_g, op.S, op.R, op.C, layer, x0, y0,
        [&](const long long pb, const unsigned off, const bool) -> NodeVal<T> {
            if (MODE == 1) return {ldt(x + pb, off), T(0)};
            return {ldt(x + pb, off), ldt(robin + pb, off)};
        },
        [&](const long long pb, const unsigned off, const int, const T conduction, const NodeVal<T>& v, auto) {
            stt(y + pb, off, v.robin >= T(0) ? conduction : v.u);
        });
}

// Jacobi diagonal of the sheet operator: 2 (c_x + c_y) / 6 per adjacent
// element corner (U_x[i][i] = U_y[i][i] = 1 / 3) plus the conductance of
// every via at the node; 1 on fixed rows (pcb.py:607-624).
__global__ __launch_bounds__(256) void k_sheet_diagonal(const OperatorDev op, double* __restrict__ diag,
                                                        double* __restrict__ dinv64, float* __restrict__ dinv32,
                                                        int* bad) {
    const long long NC = op.C + 1, NR = op.R + 1;
    const long long E = (long long)op.S * op.R * op.C;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= op.N) return;
    const int x = (int)(n % NC);
    const int y = (int)((n / NC) % NR);
    const int l = (int)(n / (NC * NR));
    double d = 0.0;
    for (int ey = y - 1; ey <= y; ++ey) {
        if (ey < 0 || ey >= op.R) continue;
        for (int ex = x - 1; ex <= x; ++ex) {
            if (ex < 0 || ex >= op.C) continue;
            const long long e = ((long long)l * op.R + ey) * op.C + ex;
            d += 2.0 * (op.coef64[e] + op.coef64[E + e]);
        }
    }
    if (op.via_ptr)
        for (int k = op.via_ptr[n]; k < op.via_ptr[n + 1]; ++k) d += op.via_g64[k];
    double value = 1.0;
    if (op.robin64[n] >= 0.0) {
        value = d;
        if (!(value > 0.0)) atomicAdd(bad, 1);
    }
    diag[n] = value;
    dinv64[n] = 1.0 / value;
    dinv32[n] = (float)(1.0 / value);
}

// The node records (1 / diagonal, robin) from the separate arrays.
__global__ __launch_bounds__(256) void k_pack_nodes(const float* __restrict__ dinv32,
                                                    const float* __restrict__ robin32,
                                                    const double* __restrict__ dinv64,
                                                    const double* __restrict__ robin64, float* __restrict__ nrec32,
                                                    double* __restrict__ nrec64, const long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    nrec32[2 * i] = dinv32[i];
    nrec32[2 * i + 1] = robin32[i];
    nrec64[2 * i] = dinv64[i];
    nrec64[2 * i + 1] = robin64[i];
}

// Jacobi diagonal: sum of the element diagonals ((a + b + c) / 9 for every
// corner of the trilinear unit tensors) over the adjacent elements plus the
// Robin conductance; 1 on fixed rows (operator.py:289-306).  ``bad`` counts
// free nodes whose conductance is not positive.
__global__ __launch_bounds__(256) void k_diagonal(const double* __restrict__ coef, const double* __restrict__ robin,
                                                  double* __restrict__ diag, double* __restrict__ dinv64,
                                                  float* __restrict__ dinv32, const int S, const int R, const int C,
                                                  int* bad) {
    const long long NC = C + 1, NR = R + 1;
    const long long N = (long long)(S + 1) * NR * NC;
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int x = (int)(n % NC);
    const int y = (int)((n / NC) % NR);
    const int z = (int)(n / (NC * NR));
    double d = 0.0;
    for (int ez = z - 1; ez <= z; ++ez) {
        if (ez < 0 || ez >= S) continue;
        for (int ey = y - 1; ey <= y; ++ey) {
            if (ey < 0 || ey >= R) continue;
            for (int ex = x - 1; ex <= x; ++ex) {
                if (ex < 0 || ex >= C) continue;
                const long long e = ((long long)ez * R + ey) * C + ex;
                d += 4.0 * ((coef[3 * e] + coef[3 * e + 1]) + coef[3 * e + 2]);
            }
        }
    }
    const double rb = robin[n];
    double value = 1.0;
    if (rb >= 0.0) {
        value = d + rb;
        if (!(value > 0.0)) atomicAdd(bad, 1);
    }
    diag[n] = value;
    dinv64[n] = 1.0 / value;
    dinv32[n] = (float)(1.0 / value);
}

// Plain copy, the reference point for bandwidth and for the hardware byte
// counters.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                              const long long n4) {
    // One 16-byte element per thread over a grid of the whole array: of the
    // copy shapes timed on gfx950 (more loads in flight per thread,
    // grid-stride loops, non-temporal hints) this one moves the most bytes.
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}

// A bandwidth reference for kernels that stream several arrays at once:
// ``reads`` input arrays and ``writes`` output arrays of one length, one
// 16-byte element of every array per thread (the inputs summed, the sum
// written to every output), so the only cost is the traffic of that many
// concurrent streams.
struct StreamSet {
    const float4* in[kStreamsMax];
    float4* out[kStreamsMax];
};

__global__ __launch_bounds__(256) void k_streams(const StreamSet set, const int reads, const int writes,
                                                 const long long n4) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    float4 sum = set.in[0][i];
    for (int k = 1; k < reads; ++k) {
        const float4 v = set.in[k][i];
        sum.x += v.x;
        sum.y += v.y;
        sum.z += v.z;
        sum.w += v.w;
    }
    for (int k = 0; k < writes; ++k) set.out[k][i] = sum;
}

// ------------------------------------------------------- vector runtime

// The FP32 vector operations of the inner-solve runtime, for callers that
// drive a PCG loop from the host.  The dot product accumulates FP32 products
// (exact in FP64) in FP64 in an order fixed by the length alone: workgroup g
// sums the contiguous range g of kDotGroupsMax at most (threads stride it,
// then a fixed LDS tree), and one workgroup adds the partials in the same
// tree, so a vector gives the same bits on every call.
constexpr int kVecThreads = 256;

__device__ double block_tree_sum(double v, double* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int w = kVecThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    return lds[0];
}

__global__ __launch_bounds__(kVecThreads) void k_dot_partial(const float* __restrict__ x, const float* __restrict__ y,
                                                             const long long n, const long long per,
                                                             double* __restrict__ partial) {
    __shared__ double lds[kVecThreads];
    const long long lo = (long long)blockIdx.x * per;
    const long long hi = lo + per < n ? lo + per : n;
    double acc = 0.0;
    for (long long i = lo + threadIdx.x; i < hi; i += kVecThreads) acc += (double)x[i] * (double)y[i];
    const double total = block_tree_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kVecThreads) void k_dot_final(const double* __restrict__ partial, const int groups,
                                                           double* __restrict__ out) {
    __shared__ double lds[kVecThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < groups; i += kVecThreads) acc += partial[i];
    const double total = block_tree_sum(acc, lds);
    if (threadIdx.x == 0) out[0] = total;
}

// out = y + alpha x with the two roundings of FP32 arithmetic (product, then
// sum); no contraction into an FMA (the library is built without it).
__global__ __launch_bounds__(kVecThreads) void k_axpy(const float alpha, const float* __restrict__ x,
                                                      const float* __restrict__ y, float* __restrict__ out,
                                                      const long long n) {
    const long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x;
    if (i < n) {
        const float scaled = alpha * x[i];
        out[i] = y[i] + scaled;
    }
}

// out = num / den, correctly rounded FP32 division.
__global__ __launch_bounds__(kVecThreads) void k_divide(const float* __restrict__ num, const float* __restrict__ den,
                                                        float* __restrict__ out, const long long n) {
    const long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x;
    if (i < n) out[i] = num[i] / den[i];
}

// Gram-Schmidt products of a restarted GMRES basis: ``count`` stored rows
// of n floats against one vector.  The rows travel as kernel arguments, so a
// launch needs no device table.
struct BasisRows {
    const float* row[kBasisRowsMax];
};
struct BasisCoefs {
    double c[kBasisRowsMax];
};

// partial[k * groups + g] = the FP64 sum of rows[k][i] * w[i] over workgroup
// g's range, in exactly the order of k_dot_partial: one row's projection has
// the bits of the dot product of the same two vectors.  The workgroup's
// range of w is read once from HBM and again from cache for every further
// row.
__global__ __launch_bounds__(kVecThreads) void k_project_partial(const BasisRows rows, const int count,
                                                                 const float* __restrict__ w, const long long n,
                                                                 const long long per, double* __restrict__ partial) {
    __shared__ double lds[kVecThreads];
    const long long lo = (long long)blockIdx.x * per;
    const long long hi = lo + per < n ? lo + per : n;
    for (int k = 0; k < count; ++k) {
        const float* __restrict__ v = rows.row[k];
        double acc = 0.0;
        for (long long i = lo + threadIdx.x; i < hi; i += kVecThreads) acc += (double)v[i] * (double)w[i];
        const double total = block_tree_sum(acc, lds);
        if (threadIdx.x == 0) partial[(long long)k * gridDim.x + blockIdx.x] = total;
        __syncthreads();  // every thread has read lds[0] before the next row writes it
    }
}

// One workgroup per row: out[k] sums that row's partials as k_dot_final does.
__global__ __launch_bounds__(kVecThreads) void k_project_final(const double* __restrict__ partial, const int groups,
                                                               double* __restrict__ out) {
    __shared__ double lds[kVecThreads];
    const double* __restrict__ mine = partial + (long long)blockIdx.x * groups;
    double acc = 0.0;
    for (int i = threadIdx.x; i < groups; i += kVecThreads) acc += mine[i];
    const double total = block_tree_sum(acc, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = total;
}

// out[i] = y[i] + sum_k coef[k] rows[k][i], accumulated in FP64 in row order
// and rounded once to FP32 (y null: zero).  Each element's inputs are read
// before it is written, so out may be y or one of the rows.
__global__ __launch_bounds__(kVecThreads) void k_combine(const BasisRows rows, const BasisCoefs coef, const int count,
                                                         const float* y, float* out, const long long n) {
    const long long i = (long long)blockIdx.x * kVecThreads + threadIdx.x;
    if (i < n) {
        double acc = y ? (double)y[i] : 0.0;
        for (int k = 0; k < count; ++k) acc += coef.c[k] * (double)rows.row[k][i];
        out[i] = (float)acc;
    }
}

// ---------------------------------------------------------- MPIR kernels

template <typename TI>
struct InnerOf;
template <>
struct InnerOf<float> {
    __device__ static const float* dinv(const OperatorDev& op) { return op.dinv32; }
    __device__ static const float* nrec(const OperatorDev& op) { return op.nrec32; }
    __device__ static const float* coef(const OperatorDev& op) { return op.coef32; }
    __device__ static CoefSource<float> coefs(const OperatorDev& op) { return {op.coef32, op.uniform, nullptr}; }
    __device__ static const float* robin(const OperatorDev& op) { return op.robin32; }
    __device__ static const float* via_g(const OperatorDev& op) { return op.via_g32; }
};
template <>
struct InnerOf<double> {
    __device__ static const double* dinv(const OperatorDev& op) { return op.dinv64; }
    __device__ static const double* nrec(const OperatorDev& op) { return op.nrec64; }
    __device__ static const double* coef(const OperatorDev& op) { return op.coef64; }
    __device__ static CoefSource<double> coefs(const OperatorDev& op) { return {op.coef64, op.uniform, nullptr}; }
    __device__ static const double* robin(const OperatorDev& op) { return op.robin64; }
    __device__ static const double* via_g(const OperatorDev& op) { return op.via_g64; }
};

// Outer step (hex_q1.cpp:870-911 and solver.py:638-686): apply a pending
// correction x += (double) c, form the FP64 residual r = b - A x and its norm,
// decide convergence, and start the inner solve on r rounded to the inner
// precision (with the first r.z of the Jacobi PCG, hex_q1.cpp:434-446).
template <typename TI>
__global__ __launch_bounds__(kGroupThreads) void k_outer(const SlotDev* __restrict__ slots,
                                                         const SlotState* __restrict__ states,
                                                         const Tile* __restrict__ tiles,
                                                         const int* __restrict__ counts) {
    // Persistent over the tile table: in most steps no slot is in its outer
    // phase and each workgroup only reads phases.
    const int count = counts[0];
    for (int t = blockIdx.x; t < count; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SlotState& st = states[tile.slot];
        if (st.phase != kPhaseOuter) continue;
        const SlotDev& sd = slots[tile.slot];
        const OperatorDev& op = sd.op;
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        const int x0 = tile.sx * kStripCols;
        const int y0 = (tile.sy * kWavesPerGroup + wave) * kRowsOuter;
        const int zlo = tile.sz * tile.zb;
        const int zhi = zlo + tile.zb < op.S + 1 ? zlo + tile.zb : op.S + 1;
        const int pending = st.pending;
        const int xcur = st.xcur;
        // Weight of the pending correction: loaded always, added with weight
        // 0 or 1 (the correction buffer starts zeroed), so no load is branched.
        const double cw = pending == kPendingCorrection ? 1.0 : 0.0;
        const double* xin = sd.x64[xcur];
        double* xout = sd.x64[1 - xcur];
        // The pending correction is the c of the buffer that the inner
        // solve's last pass wrote; the new inner residual starts the other
        // buffer with zero q, p and c.
        const KNode<TI>* __restrict__ cin = static_cast<const KNode<TI>*>(sd.state[1 - st.pcur]);
        KNode<TI>* __restrict__ sout = static_cast<KNode<TI>*>(sd.state[st.pcur]);
        const TI* __restrict__ dinv = InnerOf<TI>::dinv(op);
        const double* __restrict__ robin = op.robin64;
        const double* __restrict__ b = sd.b64;
        // b.b, r.r (FP64), r.r and r.z (inner precision) per reduction cell.
        CellSums<kOuterSums, kRowsOuter> acc;
        acc.clear();
        auto load = [&](const long long pb, const unsigned off, const bool, const int,
                        const int layer) -> NodeVal<double> {
            const double xv = ldt(xin + pb, off) + cw * (double)ld_knode_word(cin + pb, off, 3);
            return {xv, ldt(robin + pb, mask_bit(op.plain, layer) ? kOutside : off)};
        };
        auto emit = [&](const long long pb, const unsigned off, const int row, const double conduction,
                        const NodeVal<double>& v, auto parity) {
            if (pending != kPendingNone) stt(xout + pb, off, v.u);
            const double ax = v.robin >= 0.0 ? conduction + v.robin * v.u : v.u;
            const double bv = ldt(b + pb, off);
            const double r = bv - ax;
            const TI ri = (TI)r;
            st_knode(sout + pb, off, KNode<TI>{ri, TI(0), TI(0), TI(0)});
            const int g = row / kCellRows;
            acc.t[0][g] += bv * bv;
            acc.t[1][g] += r * r;
            acc.t[2][g] = dot_add(acc.t[2][g], ri, ri);
            acc.t[3][g] = dot_add(acc.t[3][g], ri, ri * ldt(dinv + pb, off));
        };
        const int y_group = tile.sy * kWavesPerGroup * kRowsOuter;
        if (op.kind == kSheet) {
            if (y0 <= op.R) {
                sheet_apply<double>(op.coef64, op.via_ptr, op.via_nbr, op.via_g64, op.S, op.R, op.C, tile.sz, x0, y0,
                                    load, emit);
                acc.sheet_done();
            }
            acc.flush_sheet(tile.sz, y_group, tile.sx, op.R, op.C, sd.part, sd.part_stride);
            continue;
        }
        if (y0 <= op.R) {
            // Stores wait for their plane's emit, so a plane's loads issue back to back.
            strip_sweep<double, kRowsOuter>(tile.sz, (op.S + tile.zb) / tile.zb, InnerOf<double>::coefs(op), op.S,
                                            op.R, op.C, x0, y0, zlo, zhi, load, emit,
                                            [&](const int z) { acc.plane_done(z, zlo, op.S); });
        }
        acc.flush_hex(zlo, zhi, y_group, tile.sx, op.R, op.C, sd.part, sd.part_stride);
    }
}

// Outer step decision, one workgroup per slot: ||b|| (first call), ||r||,
// convergence, and the start of the inner solve (hex_q1.cpp:870-911,
// solver.py:638-686).
template <typename TI>
__global__ __launch_bounds__(kFinishThreads) void k_finish_outer(const SlotDev* __restrict__ slots,
                                                                SlotState* states,
                                                                const SlotConfig* __restrict__ configs,
                                                                int* active) {
    // The slot's descriptor is read with its phase, before the test: a slot
    // in another phase reads nothing more (its vectors may be released).
    const int slot = blockIdx.x;
    SlotState& st = states[slot];
    const SlotDev& sd = slots[slot];
    const double* part = sd.part;
    const int stride = sd.part_stride, cells = sd.cells;
    if (st.phase != kPhaseOuter) return;
    double total[kOuterSums];
    slot_sum<kOuterSums, kFinishThreads>(part, stride, cells, total);
    if (threadIdx.x != 0) return;
    const SlotConfig cfg = configs[slot];
    const int pending = st.pending;
    if (!st.started) {
        const double bn = sqrt(total[0]);
        st.scale = bn > 0.0 ? bn : 1.0;
        st.target = cfg.atol + cfg.rtol * st.scale;
        st.started = 1;
    }
    if (pending != kPendingNone) st.xcur = 1 - st.xcur;
    st.pending = kPendingNone;
    st.high_apps += 1;
    const double rn = sqrt(total[1]);
    st.rel_outer = rn / st.scale;
    if (rn <= st.target) {
        st.converged = 1;
        st.phase = kPhaseDone;
        atomicSub(active, 1);
        return;
    }
    if (st.outer >= cfg.max_outer) {
        st.phase = kPhaseDone;
        atomicSub(active, 1);
        return;
    }
    const double rhs_norm = sqrt(total[2]);
    if (rhs_norm == 0.0) {
        // A zero inner right-hand side: the correction is zero (solver.py:188-189).
        double* h = sd.history + 4 * st.outer;
        h[0] = st.outer + 1;
        h[1] = st.rel_outer;
        h[2] = 0.0;
        h[3] = 0.0;
        st.outer += 1;
        st.pending = kPendingZero;
        return;
    }
    st.rhs_norm = rhs_norm;
    st.rz = total[3];
    st.first = 1;
    st.inner_its = 0;
    st.phase = kPhaseInner;
}

// The fused Jacobi-PCG pass (hex_q1.cpp:448-489 restated so that one pass
// over the vectors does one iteration).  Pass i of an inner solve, on one
// tile of one slot:
//   r_{i-1} = r_{i-2} - alpha_{i-1} q_{i-1}   (every loaded node, halo included)
//   c_{i-1} = c_{i-2} + alpha_{i-1} p_{i-1}   (owned nodes)
//   p_i     = D^-1 r_{i-1} + beta_i p_{i-1}
//   q_i     = A p_i
// and, per reduction cell, p_i.q_i, r_{i-1}.z_{i-1}, q_i.z_{i-1},
// q_i.D^-1 q_i and r_{i-1}.r_{i-1}.  The first pass has alpha = beta = 0
// and no correction (the buffers start zeroed, so the zero weights meet
// finite values).  r, p and q are ping-pong pairs: a pass reads the current
// buffers at its halo and writes the others.  The finish kernel then
// completes iteration i - 1 exactly as the two-kernel form does (tolerance
// on ||r_{i-1}||, the curvature of p_i, alpha_i = r.z / p.q) and forms
// beta_{i+1} from r_i.z_i = r.z - 2 alpha q.z + alpha^2 q.D^-1 q, the one
// value the pass cannot yet have summed.
// The fused pass's node loader, in two phases (see fetch_node): fetch issues
// the node's loads of its state (r, q, p, c in one access), its operator
// record (1/diag and Robin in one access; on a plain plane 1/diag alone, and
// the record's load goes to the out-of-range offset), and for a two-level
// slot the coarse term w at its patch; make forms
//   r = r - alpha q,  z = r D^-1 (+ w),  p = z + beta p,  c = keep c + alpha p
// (c is formed at halo nodes too, and only owned nodes store it).
template <typename TI>
struct PassRaw {
    KNode<TI> s;
    TI d, rb;
};
template <typename TI>
struct PassRawZ {
    KNode<TI> s;
    TI d, rb;
    double w;
    bool in;
};
template <typename TI, bool COARSE>
struct PassLoader {
    const KNode<TI>* sin;
    const TI* dinv;
    const TI* nrec;
    unsigned long long plain;
    TI alpha, beta, keep;
    const double* cw;
    int crows, ccols, cblock, R, C;
    int ccol;
    const int* crow;  // patch row of each strip row (two-level slots)

    __device__ auto fetch(const long long pb, const unsigned off, const bool owned, const int row,
                          const int layer) const {
        const KNode<TI> s = ld_knode(sin + pb, off);
        // One two-word load: the record (1/diag, robin), or on a plain plane
        // (1/diag[n], 1/diag[n + 1]) from the 1/diag array, whose second word
        // is not used (the array is followed by other operator arrays in its
        // allocation, so the last node's load stays inside it).  Both byte
        // offsets of kOutside are at least 2^30, past kRange.
        const bool plain_plane = mask_bit(plain, layer);
        TI d, rb;
        ld_pair(plain_plane ? dinv + pb : nrec + 2 * pb,
                plain_plane ? off * (unsigned)sizeof(TI) : rec_bytes<TI, 2>(off), d, rb);
        rb = plain_plane ? TI(0) : rb;
        if constexpr (COARSE) {
            // The patch index does not depend on loaded data and is always in
            // range, so w is fetched with the node's other loads.
            int patch;
            if (row >= 0) {
                patch = (layer * crows + crow[row]) * ccols + ccol;
            } else {
                // A via's far end (always inside the grid): its layer, row and
                // column from the index.
                const long long plane = (long long)(R + 1) * (C + 1);
                const long long n = pb + off;
                const int l = (int)(n / plane);
                const int rem = (int)(n - l * plane);
                const int y = rem / (C + 1), x = rem - y * (C + 1);
                patch = (l * crows + y / cblock) * ccols + x / cblock;
            }
            return PassRawZ<TI>{s, d, rb, cw[patch], off < kOutside};
        } else {
            return PassRaw<TI>{s, d, rb};
        }
    }
    __device__ InnerVal<TI> make(const PassRaw<TI>& x) const {
        const TI r = x.s.r - alpha * x.s.q;
        const TI p = r * x.d + beta * x.s.p;
        return InnerVal<TI>{p, x.rb, r, x.d, keep * x.s.c + alpha * x.s.p};
    }
    __device__ InnerValZ<TI> make(const PassRawZ<TI>& x) const {
        const TI r = x.s.r - alpha * x.s.q;
        // The free mask is applied after the load.
        const TI wv = (x.in && x.rb >= TI(0)) ? (TI)x.w : TI(0);
        const TI z = r * x.d + wv;
        const TI p = z + beta * x.s.p;
        return InnerValZ<TI>{p, x.rb, r, x.d, keep * x.s.c + alpha * x.s.p, z};
    }
};

template <typename TI, int ROWS, int MFMA, bool COARSE = false>
__device__ __forceinline__ void inner_step_slot(const SlotDev& sd, const SlotState& st, const Tile& tile) {
    const OperatorDev& op = sd.op;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int x0 = tile.sx * kStripCols;
    const int y0 = (tile.sy * kWavesPerGroup + wave) * ROWS;
    const int zlo = tile.sz * tile.zb;
    const int zhi = zlo + tile.zb < op.S + 1 ? zlo + tile.zb : op.S + 1;
    const bool first = st.first != 0;
    const TI alpha = first ? TI(0) : (TI)st.alpha;
    const TI beta = first ? TI(0) : (TI)st.beta;
    const TI keep = first ? TI(0) : TI(1);
    const int cur = st.pcur;
    const KNode<TI>* __restrict__ sin = static_cast<const KNode<TI>*>(sd.state[cur]);
    KNode<TI>* __restrict__ sout = static_cast<KNode<TI>*>(sd.state[1 - cur]);
    const TI* __restrict__ dinv = InnerOf<TI>::dinv(op);
    const TI* __restrict__ nrec = InnerOf<TI>::nrec(op);
    CellSums<kInnerSums, ROWS> acc;
    acc.clear();
    // Two-level slots: the coarse term of z at a node is w at its patch
    // (free nodes; nodes outside the grid read none).  The patch column of
    // the lane and the patch rows of the strip are fixed for the sweep, and
    // every index is clamped into the grid.
    [[maybe_unused]] int ccol = 0;
    [[maybe_unused]] int crow[ROWS + 2];
    if constexpr (COARSE) {
        const int cb = op.cblock;
        const int xc = x0 - 1 + (int)(threadIdx.x & (kWave - 1));
        ccol = min(max(xc, 0), op.C) / cb;
        const int ybase = op.kind == kSheet ? (tile.sy * kWavesPerGroup + wave) * kCellRows : y0;
#pragma unroll
        for (int i = 0; i < ROWS + 2; ++i) crow[i] = min(max(ybase - 1 + i, 0), op.R) / cb;
    }
    const PassLoader<TI, COARSE> load{sin, dinv, nrec, op.plain, alpha, beta, keep, sd.cw, op.crows,
                                      op.ccols, op.cblock, op.R, op.C, ccol, crow};
    // Stores wait for their plane's emit, so a plane's loads issue back to back.
    auto emit = [&](const long long pb, const unsigned off, const int row, const TI conduction, const auto& v,
                    auto parity) {
        const TI ap = v.robin >= TI(0) ? conduction + v.robin * v.u : v.u;
        st_knode(sout + pb, off, KNode<TI>{v.r, ap, v.u, v.c});
        TI z;
        if constexpr (COARSE)
            z = v.z;
        else
            z = v.r * v.d;
        const int g = row / kCellRows;
        acc.t[0][g] = dot_add(acc.t[0][g], v.u, ap);
        acc.t[1][g] = dot_add(acc.t[1][g], v.r, z);
        acc.t[2][g] = dot_add(acc.t[2][g], ap, z);
        acc.t[3][g] = dot_add(acc.t[3][g], ap, ap * v.d);
        acc.t[4][g] = dot_add(acc.t[4][g], v.r, v.r);
    };
    if (op.kind == kSheet) {
        // Sheet tiles are one layer of 4-row wave strips.
        const int ys = (tile.sy * kWavesPerGroup + wave) * kCellRows;
        if (ys <= op.R) {
            sheet_apply<TI>(InnerOf<TI>::coef(op), op.via_ptr, op.via_nbr, InnerOf<TI>::via_g(op), op.S, op.R, op.C,
                            tile.sz, x0, ys, load, emit);
            acc.sheet_done();
        }
        acc.flush_sheet(tile.sz, tile.sy * kCellBandRows, tile.sx, op.R, op.C, sd.part, sd.part_stride);
        return;
    }
    if (y0 <= op.R) {
        strip_sweep<TI, ROWS, MFMA>(
            tile.sz, (op.S + tile.zb) / tile.zb, InnerOf<TI>::coefs(op), op.S, op.R, op.C, x0, y0, zlo, zhi, load,
            emit, [&](const int z) { acc.plane_done(z, zlo, op.S); });
    }
    acc.flush_hex(zlo, zhi, tile.sy * kWavesPerGroup * ROWS, tile.sx, op.R, op.C, sd.part, sd.part_stride);
}

// One fused PCG pass, one tile of one slot per workgroup.  LOOP: the grid
// is shorter than the table and loops over its live entries (a server
// launches a fixed grid over a table that grows and shrinks, and spends no
// dispatch on empty entries); otherwise the grid is the table.  COARSE: the
// pass of two-level slots (the others' tiles are skipped; the plain pass
// skips the two-level slots' tiles), a kernel of its own so the plain pass
// keeps its registers.
template <typename TI, int ROWS, bool LOOP, bool COARSE>
__global__ __launch_bounds__(kGroupThreads) void k_inner_step(const SlotDev* __restrict__ slots,
                                                              const SlotState* __restrict__ states,
                                                              const Tile* __restrict__ tiles,
                                                              const int* __restrict__ counts, int* tickets) {
    const int count = counts[1];
    if constexpr (LOOP) {
        // Workgroups take tiles by ticket, so a workgroup that drew short
        // tiles takes more of them and none waits on a precomputed share.
        // Workgroup b draws from the contiguous range of XCD b % 8 (the
        // dispatcher deals workgroups to the XCDs in turn), so neighbouring
        // tiles still meet in one L2.  The finish kernel that follows every
        // pass resets the tickets.
        constexpr int kXcds = 8;
        const int xcd = blockIdx.x % kXcds;
        const int q = count / kXcds, rem = count % kXcds;
        const int start = xcd * q + (xcd < rem ? xcd : rem);
        const int len = q + (xcd < rem ? 1 : 0);
        int* ticket = tickets + (COARSE ? kXcds : 0) + xcd;
        __shared__ int drawn;
        for (;;) {
            if (threadIdx.x == 0) drawn = atomicAdd(ticket, 1);
            __syncthreads();
            const int j = drawn;
            __syncthreads();
            if (j >= len) break;
            const Tile tile = tiles[start + j];
            const SlotState& st = states[tile.slot];
            if (st.phase != kPhaseInner) continue;
            const SlotDev& sd = slots[tile.slot];
            if ((sd.op.nc > 0) != COARSE) continue;
            inner_step_slot<TI, ROWS, 0, COARSE>(sd, st, tile);
        }
    } else {
        const int t = xcd_tile(blockIdx.x, gridDim.x);
        if (t >= count) return;
        const Tile tile = tiles[t];
        const SlotState& st = states[tile.slot];
        if (st.phase != kPhaseInner) return;
        const SlotDev& sd = slots[tile.slot];
        if ((sd.op.nc > 0) != COARSE) return;
        inner_step_slot<TI, ROWS, 0, COARSE>(sd, st, tile);
    }
}

// The multi-right-hand-side block pass: one tile of a group of slots that
// share an operator per workgroup, the slots one after another, so the
// group's coefficient, Robin and diagonal lines are fetched from HBM once and
// served from the caches to the rest of the group; MFMA selects the
// matrix-core element action.  Each slot keeps its own scalars (per-column
// alpha and beta) and its own reduction cells.
template <int ROWS, int MFMA>
__global__ __launch_bounds__(kGroupThreads) void k_inner_step_block(const SlotDev* __restrict__ slots,
                                                                    const SlotState* __restrict__ states,
                                                                    const Tile* __restrict__ tiles,
                                                                    const int* __restrict__ counts,
                                                                    const SlotGroup* __restrict__ groups) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= counts[3]) return;
    const Tile tile = tiles[t];
    const SlotGroup& group = groups[tile.slot];
    for (int k = 0; k < group.count; ++k) {
        const int slot = group.slot[k];
        const SlotState& st = states[slot];
        if (st.phase != kPhaseInner) continue;
        inner_step_slot<float, ROWS, MFMA>(slots[slot], st, tile);
    }
}

// Sum of a[k] b[k] over k < n by one workgroup of THREADS threads: each
// thread takes k = tid, tid + THREADS, ... in order, then the lanes of each
// wave by the butterfly and the waves in order.  The result is thread 0's.
template <int THREADS>
__device__ double block_dot(const double* __restrict__ a, const double* __restrict__ b, const int n) {
    __shared__ double waves[THREADS / kWave];
    double s = 0.0;
#pragma unroll 4
    for (int k = threadIdx.x; k < n; k += THREADS) s += a[k] * b[k];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) waves[threadIdx.x / kWave] = s;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < THREADS / kWave; ++w) t += waves[w];
    return t;
}

// The decisions of one pass, one workgroup per slot: the end test of
// iteration i - 1 (hex_q1.cpp:469-489: tolerance, iteration limit, a zero or
// non-finite r.z), then the curvature check and alpha_i of iteration i
// (hex_q1.cpp:454-466) and beta_{i+1}.  A pass whose previous iteration ends
// the inner solve was not an iteration: its apply is not counted and its
// outputs (in the buffers that stay unused) are dropped; c already holds the
// final correction.
//
// Two-level slots (nc > 0): M^-1 = D^-1 + Z cinv Z^T.  The coarse kernels
// have left g = Z^T q_i and h = cinv g; q_i.M^-1 q_i gains g.h, and the
// coarse part of the next z follows r: w_i = w_{i-1} - alpha_i h, which is
// cinv Z^T r_i for r_i = r_{i-1} - alpha_i q_i.
template <typename TI>
__global__ __launch_bounds__(kFinishThreads) void k_finish_step(const SlotDev* __restrict__ slots, SlotState* states,
                                                                const SlotConfig* __restrict__ configs, int* active,
                                                                int* tickets) {
    // The tile tickets of the looping pass that ran before this kernel start
    // again from zero for the next pass.
    if (tickets && blockIdx.x == 0 && threadIdx.x < kPassTickets) tickets[threadIdx.x] = 0;
    // As in k_finish_outer, the descriptor is read with the phase, and thread
    // 0's scalars are fetched before the sum, so their latency hides behind it.
    const int slot = blockIdx.x;
    SlotState& st = states[slot];
    const SlotDev& sd = slots[slot];
    const double* part = sd.part;
    const int stride = sd.part_stride, cells = sd.cells;
    const int nc = sd.op.nc;
    if (st.phase != kPhaseInner) return;
    int first = 0, inner_its = 0;
    double rhs_norm = 1.0, rz_prev = 0.0;
    SlotConfig cfg{};
    if (threadIdx.x == 0) {
        first = st.first;
        inner_its = st.inner_its;
        rhs_norm = st.rhs_norm;
        rz_prev = st.rz;
        cfg = configs[slot];
    }
    double total[kInnerSums];
    slot_sum<kInnerSums, kFinishThreads>(part, stride, cells, total);
    double gh = 0.0;
    if (nc > 0) gh = block_dot<kFinishThreads>(sd.cg, sd.ch, nc);
    int go = 0;
    double alpha = 0.0;
    if (threadIdx.x == 0) {
        const double pq = total[0], rz = total[1], qz = total[2], qdq = total[3], rr = total[4];
        bool ended = false;
        if (first == 0) {
            inner_its += 1;
            st.inner_its = inner_its;
            const double rel = sqrt(rr) / rhs_norm;
            st.rel_inner = rel;
            ended = rel <= cfg.inner_tol || inner_its >= cfg.max_inner || !isfinite(rz) || rz_prev == 0.0;
            if (ended) {
                double* h = sd.history + 4 * st.outer;
                h[0] = st.outer + 1;
                h[1] = st.rel_outer;
                h[2] = inner_its;
                h[3] = rel;
                st.total_inner += inner_its;
                st.outer += 1;
                st.pending = kPendingCorrection;
                st.phase = kPhaseOuter;
            }
        }
        if (!ended) {
            st.low_apps += 1;
            if (!isfinite(pq) || pq <= 0.0) {
                st.status = kStatusNotSpd;
                st.phase = kPhaseDone;
                atomicSub(active, 1);
            } else {
                alpha = (double)(TI)(rz / pq);
                const double next_rz = rz - 2.0 * alpha * qz + alpha * alpha * (qdq + gh);
                st.rz = rz;
                st.alpha = alpha;
                st.beta = rz != 0.0 ? (double)(TI)(next_rz / rz) : 0.0;
                st.first = 0;
                st.pcur = 1 - st.pcur;
                go = 1;
            }
        }
    }
    if (nc == 0) return;
    __shared__ double shared_alpha;
    __shared__ int shared_go;
    if (threadIdx.x == 0) {
        shared_alpha = alpha;
        shared_go = go;
    }
    __syncthreads();
    if (!shared_go) return;
    const double a = shared_alpha;
    for (int k = threadIdx.x; k < nc; k += kFinishThreads) sd.cw[k] -= a * sd.ch[k];
}

// ------------------------------------------------- two-level coarse kernels
//
// The coarse space of the reference's AggregationCoarseCorrection
// (src/electrical/matrix_free_mpir_fem/two_level.py:54-229; native
// precondition hex_q1.cpp:388-432): functions constant over cblock x cblock
// node patches of one node layer, zero on fixed nodes.  Z^T v sums v over the
// free nodes of each patch; Z w spreads w[patch] over the patch's free nodes.

// Z^T v at one patch, one workgroup of four waves: wave k sums patch rows
// k, k + 4, ... (each lane its columns down those rows, in row order), then
// the lanes by the butterfly and the four waves in order.  Every thread
// returns the sum.
template <typename TI>
__device__ double patch_sum(const OperatorDev& op, const TI* __restrict__ v, const TI* __restrict__ robin,
                            const int layer, const int cy, const int cx) {
    __shared__ double waves[kWavesPerGroup];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int b = op.cblock;
    const long long NC = op.C + 1, NR = op.R + 1;
    const int y_hi = min(op.R + 1, (cy + 1) * b), x_lo = cx * b, x_hi = min(op.C + 1, (cx + 1) * b);
    double s = 0.0;
#pragma unroll 4
    for (int y = cy * b + wave; y < y_hi; y += kWavesPerGroup) {
        const long long row = ((long long)layer * NR + y) * NC;
        for (int x = x_lo + lane; x < x_hi; x += kWave) {
            const TI value = v[row + x];
            if (robin[row + x] >= TI(0)) s += (double)value;
        }
    }
    s = wave_sum(s);
    if (lane == 0) waves[wave] = s;
    __syncthreads();
    return ((waves[0] + waves[1]) + waves[2]) + waves[3];
}

// One row of cinv g by one wave: lane l takes columns l, l + 64, ... in
// order, then the butterfly.  Every lane returns the sum.
__device__ __forceinline__ double coarse_row(const float* __restrict__ cinv, const int nc,
                                             const double* __restrict__ g, const int row) {
    const int lane = threadIdx.x & (kWave - 1);
    const float* a = cinv + (long long)row * nc;
    double s = 0.0;
#pragma unroll 4
    for (int k = lane; k < nc; k += kWave) s += (double)a[k] * g[k];
    return wave_sum(s);
}

// Z^T v at one patch by one wave: lane l takes the patch's nodes l, l + 64,
// ... in row-major order (free nodes only), then the butterfly.  Every lane
// returns the sum.  Node n's value is v[n * stride] (a word of a KNode array
// has stride 4).
template <typename TI>
__device__ double patch_sum_wave(const OperatorDev& op, const TI* __restrict__ v, const int stride,
                                 const TI* __restrict__ robin, const int layer, const int cy, const int cx) {
    const int lane = threadIdx.x & (kWave - 1);
    const int b = op.cblock;
    const long long NC = op.C + 1, NR = op.R + 1;
    const int y_lo = cy * b, x_lo = cx * b;
    const int h = min(op.R + 1, y_lo + b) - y_lo, w = min(op.C + 1, x_lo + b) - x_lo;
    const int count = h * w;
    double s = 0.0;
#pragma unroll 4
    for (int k = lane; k < count; k += kWave) {
        const int yy = k / w, xx = k - yy * w;
        const long long idx = ((long long)layer * NR + y_lo + yy) * NC + x_lo + xx;
        const TI value = v[idx * stride];
        if (robin[idx] >= TI(0)) s += (double)value;
    }
    return wave_sum(s);
}

// The restriction of the two-level slots, one patch per wave: the waves of
// the grid loop over the patch table on their own (no barrier), so many
// slots in flight cost loop trips, not dispatches.  which 0: slots that have
// just entered an inner solve (r_0, written by the outer step); which 1:
// after a pass, its new q.
template <typename TI>
__global__ __launch_bounds__(kGroupThreads) void k_coarse_restrict(const SlotDev* __restrict__ slots,
                                                                   const SlotState* __restrict__ states,
                                                                   const Tile* __restrict__ tiles,
                                                                   const int* __restrict__ counts, const int which) {
    const int count = counts[4];
    const int waves = (int)gridDim.x * kWavesPerGroup;
    const int first = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kWavesPerGroup + (int)(threadIdx.x / kWave));
    for (int t = first; t < count; t += waves) {
        const Tile tile = tiles[t];
        const SlotState& st = states[tile.slot];
        if (st.phase != kPhaseInner || (which == 0 && st.first == 0)) continue;
        const SlotDev& sd = slots[tile.slot];
        const OperatorDev& op = sd.op;
        // r of the current state, or q of the state the pass just wrote.
        const KNode<TI>* kn = static_cast<const KNode<TI>*>(which == 0 ? sd.state[st.pcur] : sd.state[1 - st.pcur]);
        const TI* v = which == 0 ? &kn->r : &kn->q;
        const double s = patch_sum_wave<TI>(op, v, 4, InnerOf<TI>::robin(op), tile.sz, tile.sy, tile.sx);
        if ((threadIdx.x & (kWave - 1)) == 0)
            sd.cg[((long long)tile.sz * op.crows + tile.sy) * op.ccols + tile.sx] = s;
    }
}

// cinv g of the two-level slots, kCoarseRows rows per table entry, the grid
// looping over the table: into w (which 0, the start of an inner solve) or
// h (which 1, after a pass).
__global__ __launch_bounds__(kWave * kCoarseRows) void k_coarse_rows(const SlotDev* __restrict__ slots,
                                                                    const SlotState* __restrict__ states,
                                                                    const Tile* __restrict__ tiles,
                                                                    const int* __restrict__ counts, const int which) {
    const int count = counts[5];
    for (int t = blockIdx.x; t < count; t += gridDim.x) {
        const Tile tile = tiles[t];
        const SlotState& st = states[tile.slot];
        if (st.phase != kPhaseInner || (which == 0 && st.first == 0)) continue;
        const SlotDev& sd = slots[tile.slot];
        const int row = tile.index + (int)(threadIdx.x / kWave);
        if (row >= sd.op.nc) continue;
        const double s = coarse_row(sd.op.cinv, sd.op.nc, sd.cg, row);
        if ((threadIdx.x & (kWave - 1)) == 0) (which == 0 ? sd.cw : sd.ch)[row] = s;
    }
}

// The standalone preconditioner z = M^-1 r of one operator in FP32
// (operator.py:366-371): patch sums, coarse rows, then z = D^-1 r + Z w.
template <typename T>
__global__ __launch_bounds__(kGroupThreads) void k_op_restrict(const OperatorDev op, const T* __restrict__ r,
                                                               double* __restrict__ g) {
    const int patch = blockIdx.x;
    const int per_layer = op.crows * op.ccols;
    const int layer = patch / per_layer, cy = (patch % per_layer) / op.ccols, cx = patch % op.ccols;
    const double s = patch_sum<T>(op, r, InnerOf<T>::robin(op), layer, cy, cx);
    if (threadIdx.x == 0) g[patch] = s;
}

// fine = Z v in FP64: every free node takes its patch's coarse value, fixed
// nodes 0 (the Galerkin assembly's prolongation).
__global__ __launch_bounds__(256) void k_op_spread(const OperatorDev op, const double* __restrict__ v,
                                                   double* __restrict__ fine) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= op.N) return;
    double value = 0.0;
    if (op.robin64[n] >= 0.0) {
        const long long NC = op.C + 1, plane = (long long)(op.R + 1) * NC;
        const int l = (int)(n / plane);
        const int rem = (int)(n - l * plane);
        const int y = (int)(rem / NC), x = (int)(rem - y * NC);
        value = v[((long long)l * op.crows + y / op.cblock) * op.ccols + x / op.cblock];
    }
    fine[n] = value;
}

__global__ __launch_bounds__(kWave * kCoarseRows) void k_op_rows(const OperatorDev op, const double* __restrict__ g,
                                                                double* __restrict__ w) {
    const int row = blockIdx.x * kCoarseRows + (int)(threadIdx.x / kWave);
    if (row >= op.nc) return;
    const double s = coarse_row(op.cinv, op.nc, g, row);
    if ((threadIdx.x & (kWave - 1)) == 0) w[row] = s;
}

__global__ __launch_bounds__(256) void k_op_prolong(const OperatorDev op, const float* __restrict__ r,
                                                    float* __restrict__ z, const double* __restrict__ w) {
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= op.N) return;
    float value = r[n] * op.dinv32[n];
    if (op.nc > 0 && op.robin32[n] >= 0.0f) {
        const long long NC = op.C + 1, plane = (long long)(op.R + 1) * NC;
        const int l = (int)(n / plane);
        const int rem = (int)(n - l * plane);
        const int y = (int)(rem / NC), x = (int)(rem - y * NC);
        value += (float)w[((long long)l * op.crows + y / op.cblock) * op.ccols + x / op.cblock];
    }
    z[n] = value;
}

// Writes fresh states of newly admitted slots of a server: one thread per
// listed slot.  Slots already iterating keep the states their finish kernels
// wrote.
__global__ void k_scatter_states(const int* __restrict__ index, const SlotState* __restrict__ fresh,
                                 SlotState* states, const int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) states[index[i]] = fresh[i];
}

// The inverse from LAPACK-style potri (its lower triangle, column-major
// n x n in ``packed``) as a full symmetric matrix: FP32 row-major into
// ``inverse32`` and, when ``inverse64`` is not null, FP64.
__global__ __launch_bounds__(256) void k_symmetric_inverse(const double* __restrict__ packed, const int n,
                                                           float* __restrict__ inverse32,
                                                           double* __restrict__ inverse64) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)n * n) return;
    const int r = (int)(k / n), c = (int)(k - (long long)r * n);
    // Element (r, c) of the lower triangle sits at r + c n (column-major).
    const double v = r >= c ? packed[r + (long long)c * n] : packed[c + (long long)r * n];
    inverse32[k] = (float)v;
    if (inverse64) inverse64[k] = v;
}

// ------------------------------------------------------ host launchers

namespace launch {

template <typename T>
hipError_t apply(const CoefSource<T>& coef, const T* robin, const unsigned long long plain, const T* x, T* y, int S,
                 int R, int C, int mode, int rows, int zb, int mfma, hipStream_t stream) {
    const int nsx = (C + 1 + kStripCols - 1) / kStripCols;
    const int nsy = (R + 1 + rows * kWavesPerGroup - 1) / (rows * kWavesPerGroup);
    const int nsz = (S + 1 + zb - 1) / zb;
    const dim3 grid(nsx * nsy * nsz), block(kGroupThreads);
#define PCBHIP_APPLY(ROWS, MODE, MF)                                                                               \
    hipLaunchKernelGGL((k_apply<T, ROWS, MODE, MF>), grid, block, 0, stream, coef, robin, plain, x, y, S, R, C, nsx, \
                       nsy, zb)
#define PCBHIP_APPLY_ROWS(MF)                                                                                      \
    if (rows == 8) {                                                                                               \
        if (mode == 0) PCBHIP_APPLY(8, 0, MF); else PCBHIP_APPLY(8, 1, MF);                                        \
    } else {                                                                                                       \
        if (mode == 0) PCBHIP_APPLY(4, 0, MF); else PCBHIP_APPLY(4, 1, MF);                                        \
    }
    if constexpr (sizeof(T) == 4) {
        if (mfma == 3) {
            PCBHIP_APPLY_ROWS(3)
        } else if (mfma == 2) {
            PCBHIP_APPLY_ROWS(2)
        } else {
            PCBHIP_APPLY_ROWS(0)
        }
    } else {
        if (rows == 2) {
            if (mode == 0) PCBHIP_APPLY(2, 0, 0); else PCBHIP_APPLY(2, 1, 0);
        } else {
            if (mode == 0) PCBHIP_APPLY(4, 0, 0); else PCBHIP_APPLY(4, 1, 0);
        }
    }
#undef PCBHIP_APPLY_ROWS
#undef PCBHIP_APPLY
    return hipGetLastError();
}

template hipError_t apply<float>(const CoefSource<float>&, const float*, unsigned long long, const float*, float*, int,
                                 int, int, int, int, int, int, hipStream_t);
template hipError_t apply<double>(const CoefSource<double>&, const double*, unsigned long long, const double*, double*,
                                  int, int, int, int, int, int, int, hipStream_t);

template <typename T>
hipError_t sheet_apply(const OperatorDev& op, const T* x, T* y, int mode, hipStream_t stream) {
    const int nsx = (op.C + 1 + kStripCols - 1) / kStripCols;
    const int nsy = (op.R + 1 + kCellRows * kWavesPerGroup - 1) / (kCellRows * kWavesPerGroup);
    const dim3 grid(nsx * nsy * op.S), block(kGroupThreads);
    const T* coef;
    const T* robin;
    const T* via_g;
    if constexpr (sizeof(T) == 4) {
        coef = op.coef32;
        robin = op.robin32;
        via_g = op.via_g32;
    } else {
        coef = op.coef64;
        robin = op.robin64;
        via_g = op.via_g64;
    }
    if (mode == 0)
        hipLaunchKernelGGL((k_sheet_apply<T, 0>), grid, block, 0, stream, op, coef, robin, via_g, x, y, nsx, nsy);
    else
        hipLaunchKernelGGL((k_sheet_apply<T, 1>), grid, block, 0, stream, op, coef, robin, via_g, x, y, nsx, nsy);
    return hipGetLastError();
}

template hipError_t sheet_apply<float>(const OperatorDev&, const float*, float*, int, hipStream_t);
template hipError_t sheet_apply<double>(const OperatorDev&, const double*, double*, int, hipStream_t);

hipError_t sheet_diagonal(const OperatorDev& op, double* diag, double* dinv64, float* dinv32, int* bad,
                          hipStream_t stream) {
    const int blocks = (int)((op.N + 255) / 256);
    hipLaunchKernelGGL(k_sheet_diagonal, dim3(blocks), dim3(256), 0, stream, op, diag, dinv64, dinv32, bad);
    return hipGetLastError();
}

hipError_t diagonal(const double* coef, const double* robin, double* diag, double* dinv64, float* dinv32, int S,
                    int R, int C, int* bad, hipStream_t stream) {
    const long long N = (long long)(S + 1) * (R + 1) * (C + 1);
    const int blocks = (int)((N + 255) / 256);
    hipLaunchKernelGGL(k_diagonal, dim3(blocks), dim3(256), 0, stream, coef, robin, diag, dinv64, dinv32, S, R, C,
                       bad);
    return hipGetLastError();
}

hipError_t pack_nodes(const float* dinv32, const float* robin32, const double* dinv64, const double* robin64,
                      float* nrec32, double* nrec64, long long n, hipStream_t stream) {
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_pack_nodes, dim3(blocks), dim3(256), 0, stream, dinv32, robin32, dinv64, robin64, nrec32,
                       nrec64, n);
    return hipGetLastError();
}

int dot_groups(long long n) {
    const long long groups = (n + 4095) / 4096;
    return (int)(groups < 1 ? 1 : (groups > kDotGroupsMax ? kDotGroupsMax : groups));
}

hipError_t dot(const float* x, const float* y, long long n, double* partial, double* out, hipStream_t stream) {
    const int groups = dot_groups(n);
    const long long per = (n + groups - 1) / groups;
    hipLaunchKernelGGL(k_dot_partial, dim3(groups), dim3(kVecThreads), 0, stream, x, y, n, per, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kVecThreads), 0, stream, partial, groups, out);
    return hipGetLastError();
}

hipError_t axpy(float alpha, const float* x, const float* y, float* out, long long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const long long blocks = (n + kVecThreads - 1) / kVecThreads;
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)blocks), dim3(kVecThreads), 0, stream, alpha, x, y, out, n);
    return hipGetLastError();
}

hipError_t divide(const float* num, const float* den, float* out, long long n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const long long blocks = (n + kVecThreads - 1) / kVecThreads;
    hipLaunchKernelGGL(k_divide, dim3((unsigned)blocks), dim3(kVecThreads), 0, stream, num, den, out, n);
    return hipGetLastError();
}

hipError_t project(const float* const* rows, int count, const float* w, long long n, double* partial, double* out,
                   hipStream_t stream) {
    if (count < 1 || count > kBasisRowsMax || n <= 0) return hipErrorInvalidValue;
    BasisRows set{};
    for (int k = 0; k < count; ++k) set.row[k] = rows[k];
    const int groups = dot_groups(n);
    const long long per = (n + groups - 1) / groups;
    hipLaunchKernelGGL(k_project_partial, dim3(groups), dim3(kVecThreads), 0, stream, set, count, w, n, per, partial);
    hipLaunchKernelGGL(k_project_final, dim3(count), dim3(kVecThreads), 0, stream, partial, groups, out);
    return hipGetLastError();
}

hipError_t combine(const float* const* rows, const double* coef, int count, const float* y, float* out, long long n,
                   hipStream_t stream) {
    if (count < 1 || count > kBasisRowsMax) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    BasisRows set{};
    BasisCoefs c{};
    for (int k = 0; k < count; ++k) {
        set.row[k] = rows[k];
        c.c[k] = coef[k];
    }
    const long long blocks = (n + kVecThreads - 1) / kVecThreads;
    hipLaunchKernelGGL(k_combine, dim3((unsigned)blocks), dim3(kVecThreads), 0, stream, set, c, count, y, out, n);
    return hipGetLastError();
}

hipError_t streams(const float* const* in, int reads, float* const* out, int writes, long long n_floats,
                   hipStream_t stream) {
    StreamSet set{};
    for (int k = 0; k < reads; ++k) set.in[k] = reinterpret_cast<const float4*>(in[k]);
    for (int k = 0; k < writes; ++k) set.out[k] = reinterpret_cast<float4*>(out[k]);
    const long long n4 = n_floats / 4;
    const int blocks = (int)((n4 + 255) / 256);
    hipLaunchKernelGGL(k_streams, dim3(blocks), dim3(256), 0, stream, set, reads, writes, n4);
    return hipGetLastError();
}

hipError_t copy(const float* src, float* dst, long long n_floats, hipStream_t stream) {
    const long long n4 = n_floats / 4;
    const int blocks = (int)((n4 + 255) / 256);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<const float4*>(src),
                       reinterpret_cast<float4*>(dst), n4);
    return hipGetLastError();
}

// The restriction and the coarse rows of the two-level slots (which as in
// k_coarse_restrict).
template <typename TI>
void coarse_kernels(const SlotDev* slots, const SlotState* states, const int* counts, const CoarseLaunch& coarse,
                    const int which, hipStream_t stream) {
    if (coarse.grid_patches <= 0) return;
    const int patch_groups = (coarse.grid_patches + kWavesPerGroup - 1) / kWavesPerGroup;
    const int patches = patch_groups < kCoarseGroups ? patch_groups : kCoarseGroups;
    const int rows = coarse.grid_rows < kCoarseGroups ? coarse.grid_rows : kCoarseGroups;
    hipLaunchKernelGGL(k_coarse_restrict<TI>, dim3(patches), dim3(kGroupThreads), 0, stream, slots, states,
                       coarse.patches, counts, which);
    hipLaunchKernelGGL(k_coarse_rows, dim3(rows > 0 ? rows : 1), dim3(kWave * kCoarseRows), 0, stream, slots, states,
                       coarse.rows, counts, which);
}

hipError_t outer_step(bool inner64, int n_slots, const SlotDev* slots, SlotState* states,
                      const SlotConfig* configs, const Tile* outer_tiles, const int* counts, int grid_outer,
                      int* active, const CoarseLaunch& coarse, hipStream_t stream) {
    const dim3 block(kGroupThreads);
    const dim3 grid(grid_outer < kOuterGroups ? grid_outer : kOuterGroups);
    if (inner64) {
        hipLaunchKernelGGL(k_outer<double>, grid, block, 0, stream, slots, states, outer_tiles, counts);
        hipLaunchKernelGGL(k_finish_outer<double>, dim3(n_slots), dim3(kFinishThreads), 0, stream, slots, states,
                           configs, active);
        coarse_kernels<double>(slots, states, counts, coarse, 0, stream);
    } else {
        hipLaunchKernelGGL(k_outer<float>, grid, block, 0, stream, slots, states, outer_tiles, counts);
        hipLaunchKernelGGL(k_finish_outer<float>, dim3(n_slots), dim3(kFinishThreads), 0, stream, slots, states,
                           configs, active);
        coarse_kernels<float>(slots, states, counts, coarse, 0, stream);
    }
    return hipGetLastError();
}

hipError_t inner_pass(bool inner64, int rows_inner, const SlotDev* slots, const SlotState* states,
                      const Tile* inner_tiles, const int* counts, int grid_inner, bool loop, int* tickets,
                      const Tile* block_tiles, const SlotGroup* groups, int grid_block, int block_mfma,
                      const CoarseLaunch& coarse, hipStream_t stream) {
    // A looping grid needs a ticket set and every XCD range drawn by some workgroup.
    if (loop && (!tickets || grid_inner < 8)) return hipErrorInvalidValue;
    const dim3 block(kGroupThreads);
    const dim3 grid(grid_inner);
    const bool two_level = coarse.grid_patches > 0;
#define PCBHIP_PASS(TI, ROWS, COARSE)                                                                              \
    do {                                                                                                           \
        if (loop)                                                                                                  \
            hipLaunchKernelGGL((k_inner_step<TI, ROWS, true, COARSE>), grid, block, 0, stream, slots, states,        \
                               inner_tiles, counts, tickets);                                                      \
        else                                                                                                       \
            hipLaunchKernelGGL((k_inner_step<TI, ROWS, false, COARSE>), grid, block, 0, stream, slots, states,       \
                               inner_tiles, counts, nullptr);                                                      \
    } while (0)
    if (inner64) {
        if (coarse.plain) PCBHIP_PASS(double, kRowsInner64, false);
        if (two_level) PCBHIP_PASS(double, kRowsInner64, true);
        return hipGetLastError();
    }
    if (rows_inner == 8) {
        if (coarse.plain) PCBHIP_PASS(float, 8, false);
        if (two_level) PCBHIP_PASS(float, 8, true);
    } else {
        if (coarse.plain) PCBHIP_PASS(float, 4, false);
        if (two_level) PCBHIP_PASS(float, 4, true);
    }
#undef PCBHIP_PASS
    if (grid_block > 0) {
        if (block_mfma == 3) {
            hipLaunchKernelGGL((k_inner_step_block<4, 3>), dim3(grid_block), block, 0, stream, slots, states,
                               block_tiles, counts, groups);
        } else if (block_mfma == 2) {
            hipLaunchKernelGGL((k_inner_step_block<4, 2>), dim3(grid_block), block, 0, stream, slots, states,
                               block_tiles, counts, groups);
        } else {
            hipLaunchKernelGGL((k_inner_step_block<4, 0>), dim3(grid_block), block, 0, stream, slots, states,
                               block_tiles, counts, groups);
        }
    }
    return hipGetLastError();
}

hipError_t inner_step(bool inner64, int rows_inner, int n_slots, const SlotDev* slots, SlotState* states,
                      const SlotConfig* configs, const Tile* inner_tiles, const int* counts, int grid_inner, bool loop,
                      int* tickets, const Tile* block_tiles, const SlotGroup* groups, int grid_block, int block_mfma,
                      int* active, const CoarseLaunch& coarse, hipStream_t stream) {
    const hipError_t err = inner_pass(inner64, rows_inner, slots, states, inner_tiles, counts, grid_inner, loop,
                                      tickets, block_tiles, groups, grid_block, block_mfma, coarse, stream);
    if (err != hipSuccess) return err;
    const dim3 block(kFinishThreads);
    if (inner64) {
        coarse_kernels<double>(slots, states, counts, coarse, 1, stream);
        hipLaunchKernelGGL(k_finish_step<double>, dim3(n_slots), block, 0, stream, slots, states, configs, active,
                           tickets);
    } else {
        coarse_kernels<float>(slots, states, counts, coarse, 1, stream);
        hipLaunchKernelGGL(k_finish_step<float>, dim3(n_slots), block, 0, stream, slots, states, configs, active,
                           tickets);
    }
    return hipGetLastError();
}

hipError_t coarse_spread(const OperatorDev& op, const double* v, double* fine, hipStream_t stream) {
    hipLaunchKernelGGL(k_op_spread, dim3((unsigned)((op.N + 255) / 256)), dim3(256), 0, stream, op, v, fine);
    return hipGetLastError();
}

hipError_t symmetric_inverse(const double* packed, int n, float* inverse32, double* inverse64, hipStream_t stream) {
    const long long count = (long long)n * n;
    hipLaunchKernelGGL(k_symmetric_inverse, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, packed, n,
                       inverse32, inverse64);
    return hipGetLastError();
}

hipError_t coarse_restrict(const OperatorDev& op, const double* y, double* g, hipStream_t stream) {
    hipLaunchKernelGGL(k_op_restrict<double>, dim3(op.nc), dim3(kGroupThreads), 0, stream, op, y, g);
    return hipGetLastError();
}

hipError_t precondition(const OperatorDev& op, const float* r, float* z, double* g, double* w, hipStream_t stream) {
    if (op.nc > 0) {
        hipLaunchKernelGGL(k_op_restrict<float>, dim3(op.nc), dim3(kGroupThreads), 0, stream, op, r, g);
        hipLaunchKernelGGL(k_op_rows, dim3((op.nc + kCoarseRows - 1) / kCoarseRows), dim3(kWave * kCoarseRows), 0,
                           stream, op, g, w);
    }
    hipLaunchKernelGGL(k_op_prolong, dim3((unsigned)((op.N + 255) / 256)), dim3(256), 0, stream, op, r, z, w);
    return hipGetLastError();
}

hipError_t scatter_states(const int* index, const SlotState* fresh, SlotState* states, int count,
                          hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_states, dim3((count + 63) / 64), dim3(64), 0, stream, index, fresh, states, count);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace pcbhip
