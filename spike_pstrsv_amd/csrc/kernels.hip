// gfx950 kernels of the Spike sparse triangular solve (FP64, and FP32 block
// solves for the iteratively refined mode).
//
// Block solve (pass 1 / pass 2): one 64-lane wavefront per partition.  The
// partition's x segment lives in LDS (k + XS_PAD slots; slot k holds 0.0 and
// is the target of padding entries).  Step values are pre-scaled by their
// row's 1/diag; before the steps run, xs[r] holds the row's right-hand side
// times 1/diag, and a step slot computes
//     xs[row] = xs[row] - sum_e val_e * xs[col_e]
// (the chunks of a long row subtract their partial sums in turn).  Steps execute in order and each step only reads rows written by
// earlier steps; LDS operations of one wavefront complete in program order, so
// no flags or barriers are needed between steps.  Step records are streamed
// from HBM through a register ring D steps deep, so the latency of the loads
// is off the dependency chain.  The G lanes of a row slot are reduced with
// DPP (quad_perm / row_half_mirror / row_mirror), in a fixed order, so results
// are bit-for-bit reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.hpp"

namespace spike {

namespace {

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    int lo = (int)(unsigned)(u & 0xffffffffull);
    int hi = (int)(unsigned)(u >> 32);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Raw buffer loads: a descriptor per array and partition, a constant per-lane
// byte offset (VGPR) and a per-step byte offset (SGPR), so refilling the ring
// needs no vector address arithmetic.  Word 3 0x00020000: raw dword access on
// gfx9-family parts.  Offsets past num_records read 0.
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ uint16_t buf_u16(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void buf_2f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, double &a, double &b)
{
    u32x4_t v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
    u32x2_t lo = {v.x, v.y}, hi = {v.z, v.w};
    a = __builtin_bit_cast(double, lo);
    b = __builtin_bit_cast(double, hi);
}

// Every lane receives the values held by the four 16-lane rows of v (v is
// uniform within each row): x[t] = v of row t.  gfx950's permlane16 swap
// exchanges odd and even rows, the permlane32 swap the two halves; three
// swaps per 32-bit half, all in VGPRs.
__device__ __forceinline__ void rows4_f64(double v, double x[4])
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)(u & 0xffffffffull), hi = (unsigned)(u >> 32);
    const auto alo = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);   // rows {0,0,2,2}, {1,1,3,3}
    const auto ahi = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const auto blo = __builtin_amdgcn_permlane32_swap(alo[0], alo[0], false, false);   // row 0, row 2
    const auto bhi = __builtin_amdgcn_permlane32_swap(ahi[0], ahi[0], false, false);
    const auto clo = __builtin_amdgcn_permlane32_swap(alo[1], alo[1], false, false);   // row 1, row 3
    const auto chi = __builtin_amdgcn_permlane32_swap(ahi[1], ahi[1], false, false);
    auto mk = [](unsigned l, unsigned h) {
        return __longlong_as_double((long long)(((unsigned long long)h << 32) | l));
    };
    x[0] = mk(blo[0], bhi[0]);
    x[1] = mk(clo[0], chi[0]);
    x[2] = mk(blo[1], bhi[1]);
    x[3] = mk(clo[1], chi[1]);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) { return dpp_f64<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) { return dpp_f32<CTRL>(v); }

// Sum over aligned groups of G lanes; every lane of a group gets the same bits.
template <int G, class T>
__device__ __forceinline__ T group_sum(T v)
{
    if constexpr (G >= 2) v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    if constexpr (G >= 4) v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v += dpp_move<0x141>(v);  // row_half_mirror
    if constexpr (G >= 16) v += dpp_move<0x140>(v); // row_mirror
    return v;
}

// Value-type helpers of the block solve: FP64 or FP32 step values.
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <class T> struct VT;
template <> struct VT<double> {
    static __device__ __forceinline__ const double *val(const BlockArgs &A) { return A.st.val; }
    static __device__ __forceinline__ double load(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s)
    {
        return buf_f64(r, v, s);
    }
};
template <> struct VT<float> {
    static __device__ __forceinline__ const float *val(const BlockArgs &A) { return A.st.val32; }
    static __device__ __forceinline__ float load(__amdgpu_buffer_rsrc_t r, unsigned v, unsigned s)
    {
        return buf_f32(r, v, s);
    }
};

// Writes one solution entry: x = v, or x += v for a refinement update.
__device__ __forceinline__ void put_x(double *x, size_t i, double v, int accum)
{
    x[i] = accum ? x[i] + v : v;
}

// Caller-side addressing of one partition's right-hand sides and solutions.
struct RhsMap {
    const double *b;
    double *x;
    int ldb, ldx, nr, row0, mn, mflip;
    // caller index of lower-form row g
    __device__ __forceinline__ int at(int g) const
    {
        const int l = g - row0;
        return mflip ? mn - 1 - l : l;
    }
    __device__ __forceinline__ double bv(int r, int g) const
    {
        return b[(size_t)(r < nr ? r : nr - 1) * ldb + at(g)];
    }
};
__device__ __forceinline__ RhsMap rhs_map(const BlockArgs &A, const DevPart &pt)
{
    RhsMap m;
    m.b = A.b + pt.boff;
    m.x = A.x + pt.boff;
    m.ldb = pt.ld ? pt.ld : A.ldb;
    m.ldx = pt.ld ? pt.ld : A.ldx;
    m.nr = pt.mnr ? pt.mnr : A.nr;
    m.row0 = pt.row0;
    m.mn = pt.mn;
    m.mflip = pt.mflip;
    return m;
}

// Bytes of a block-solve x segment (kmax + XS_PAD slots of NR values), rounded
// up to 8 so that the FP64 coupling scratch can follow it.
template <class T>
__host__ __device__ __forceinline__ size_t xs_bytes(int kmax, int NR)
{
    return ((size_t)(kmax + XS_PAD) * NR * sizeof(T) + 7) & ~(size_t)7;
}

// xs[i*NR + r] = b_r[s+i] (times 1/diag when scaled) for i < k.  Each lane
// takes UB rows per trip at clamped indices, so all loads of a trip are in
// flight together (one memory round trip per trip, not per row).
template <int NR, class T>
__device__ __forceinline__ void load_rhs(const BlockArgs &A, const RhsMap &m, T *xs, int s, int k, int roff,
                                         bool scaled)
{
    constexpr int UB = NR >= 8 ? 1 : 8 / NR;   // 8 values of b per lane and trip
    const int lane = threadIdx.x;
    for (int i0 = lane; i0 < k; i0 += 64 * UB) {
        T sc[UB], bv[UB][NR];
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const int i = min(i0 + 64 * u, k - 1);
            sc[u] = scaled ? (T)A.inv_diag[s + i] : (T)1;
#pragma unroll
            for (int r = 0; r < NR; r++) bv[u][r] = (T)m.bv(roff + r, s + i);
        }
        // every value of the trip is used here, so no load sinks into the
        // conditional stores below (which would wait per row again)
#pragma unroll
        for (int u = 0; u < UB; u++) {
            asm volatile("" ::"v"(sc[u]));
#pragma unroll
            for (int r = 0; r < NR; r++) asm volatile("" ::"v"(bv[u][r]));
        }
#pragma unroll
        for (int u = 0; u < UB; u++)
            if (i0 + 64 * u < k)
#pragma unroll
                for (int r = 0; r < NR; r++) xs[(i0 + 64 * u) * NR + r] = bv[u][r] * sc[u];
    }
}

// Values published by other workgroups during a launch (the sync-free level
// solve, the fused Spike solve) carry their readiness: the vector starts
// filled with kNotReady, a signalling-NaN pattern that arithmetic never
// produces (results of NaN operands are quiet NaNs); each value is published
// by one 64-bit agent-scope store and read by polling with agent-scope loads,
// which are performed at the agent's coherence point (the XCDs' L2 caches are
// not coherent with each other).
constexpr unsigned long long kNotReady = 0x7FF4DEADBEEF0001ull;

__device__ __forceinline__ double load_agent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool not_ready(double v)
{
    return (unsigned long long)__double_as_longlong(v) == kNotReady;
}

// Polls v[u] = src[idx[u]] (u with in[u]) until none is kNotReady; false if
// the wait gave up.
template <int U, class I>
__device__ __forceinline__ bool poll_values(const double *src, const I (&idx)[U], const bool (&in)[U],
                                            double (&v)[U], int spin_limit, int *abort_word, int *fail_word)
{
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = load_agent(&src[idx[u]]);   // all polls in flight together
    for (int spins = 0, nap = 1;; spins++) {
        bool ready = true;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u] && not_ready(v[u])) ready = false;
        if (ready) return true;
        if (spins >= spin_limit ||
            ((spins & 63) == 63 && __hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            // the first lane to give up also raises the caller-visible
            // failure word (host memory, read by the library after the solve)
            if (__hip_atomic_exchange(abort_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 && fail_word)
                __hip_atomic_store(fail_word, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
#pragma unroll
            for (int u = 0; u < U; u++)
                if (not_ready(v[u])) v[u] = __builtin_nan("");   // never pass the pattern on
            return false;
        }
        // back off (up to about 1000 cycles), then poll again only what is
        // still unpublished, so waiting waves do not flood the memory system
        for (int z = 0; z < nap; z++) __builtin_amdgcn_s_sleep(4);
        nap = nap < 4 ? nap * 2 : 4;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u] && not_ready(v[u])) v[u] = load_agent(&src[idx[u]]);
    }
}

// The fused Spike solve's view of its published vectors (one launch, see
// FusedArgs): xcur holds x^(0) .. x^(S) of this solve's parity, mstride values
// apart (nred x nr_slots each), xnext the other parity's, which this solve
// refills with kNotReady for the next one.
struct FusedCtx {
    double *xcur, *xnext;
    size_t mstride;
    int S, spin_limit;
    int *abort_word, *fail_word;
};

// Pass-2 right-hand side of one partition in LDS, NR right-hand sides
// interleaved (xs[i*NR + r]): b for every row, then b - sum_j c_ij xr[col_j]
// on the coupling rows (times 1/diag when scaled).  The coupling update is
// FP64 for FP32 factors too: the products are gathered lane-parallel into the
// FP64 LDS scratch prod[] and each coupling row then sums its own products in
// entry order, exactly as k_coupling does, so the result depends neither on
// the lane count nor on which of the two computes it.  Right-hand sides
// r >= nr reuse column nr-1.
template <int NR, class T, bool FUSED>
__device__ __forceinline__ void pass2_rhs(const BlockArgs &A, const FusedCtx &F, const DevPart &pt, T *xs,
                                          double *prod, bool scaled, int roff, int RS)
{
    const int lane = threadIdx.x, s = pt.s, k = pt.k;
    const RhsMap m = rhs_map(A, pt);
    if (pt.ccnt == 0) {
        load_rhs<NR, T>(A, m, xs, s, k, roff, scaled);
        return;
    }
    // The coupling rows of this lane (up to UR of them when the partition has
    // at most 64 * UR; otherwise a loop after the barrier): their indices,
    // right-hand sides and 1/diag are loaded first, so that these loads, the
    // plain right-hand sides and the coupling products overlap instead of
    // following each other.
    constexpr int UR = 2;
    const bool few = pt.ccnt <= 64 * UR;
    int gr[UR], pa[UR], pb[UR];
#pragma unroll
    for (int u = 0; u < UR; u++) {
        const int cc = pt.cbeg + min(lane + 64 * u, pt.ccnt - 1);
        gr[u] = A.cpl_rows[cc];
        pa[u] = A.cpl_fused ? A.cpl_ptr[cc] : 0;
        pb[u] = A.cpl_fused ? A.cpl_ptr[cc + 1] : 0;
    }
    const int e0 = pt.cpe0, e1 = pt.cpe1;
    load_rhs<NR, T>(A, m, xs, s, k, roff, scaled);
    T scr[UR];
    double bvr[UR][NR];
    if (few) {
#pragma unroll
        for (int u = 0; u < UR; u++) {
            scr[u] = scaled ? (T)A.inv_diag[gr[u]] : (T)1;
#pragma unroll
            for (int r = 0; r < NR; r++)
                bvr[u][r] = A.cpl_fused ? m.bv(roff + r, gr[u])
                                        : A.bc[(size_t)(pt.cbeg + min(lane + 64 * u, pt.ccnt - 1)) * RS + roff + r];
        }
    }
    if (A.cpl_fused) {
        const double *cv = A.cpl_val;
        // UC entries per lane per trip at clamped indices: values and
        // columns, then the gathers, then the products
        constexpr int UC = NR >= 8 ? 1 : 8 / NR;
        for (int e = e0 + lane; e < e1; e += 64 * UC) {
            double w[UC], xv[UC][NR];
            int col[UC];
#pragma unroll
            for (int u = 0; u < UC; u++) {
                const int ee = min(e + 64 * u, e1 - 1);
                w[u] = cv[ee];
                col[u] = A.cpl_col[ee];
            }
            if constexpr (FUSED) {
                // x_K of the last sweep, published by other workgroups
                unsigned idx[UC * NR];
                bool in[UC * NR];
                double pv[UC * NR];
#pragma unroll
                for (int u = 0; u < UC; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        idx[u * NR + r] = (unsigned)col[u] * (unsigned)RS + (unsigned)(roff + r);
                        in[u * NR + r] = e + 64 * u < e1;
                    }
                (void)poll_values<UC * NR>(F.xcur + (size_t)F.S * F.mstride, idx, in, pv, F.spin_limit,
                                           F.abort_word, F.fail_word);
#pragma unroll
                for (int u = 0; u < UC; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) xv[u][r] = pv[u * NR + r];
            } else {
#pragma unroll
                for (int u = 0; u < UC; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) xv[u][r] = A.xr[(size_t)col[u] * RS + roff + r];
            }
#pragma unroll
            for (int u = 0; u < UC; u++) {   // all loads before the conditional stores
                asm volatile("" ::"v"(w[u]));
#pragma unroll
                for (int r = 0; r < NR; r++) asm volatile("" ::"v"(xv[u][r]));
            }
#pragma unroll
            for (int u = 0; u < UC; u++)
                if (e + 64 * u < e1)
#pragma unroll
                    for (int r = 0; r < NR; r++) prod[(e + 64 * u - e0) * NR + r] = w[u] * xv[u][r];
        }
    }
    __syncthreads();
    if (few) {
#pragma unroll
        for (int u = 0; u < UR; u++)
            if (lane + 64 * u < pt.ccnt) {
                double acc[NR];
#pragma unroll
                for (int r = 0; r < NR; r++) acc[r] = 0.0;
                for (int e = pa[u]; e < pb[u]; e++)
#pragma unroll
                    for (int r = 0; r < NR; r++) acc[r] += prod[(e - e0) * NR + r];
#pragma unroll
                for (int r = 0; r < NR; r++)
                    xs[(gr[u] - s) * NR + r] = A.cpl_fused ? (T)(bvr[u][r] - acc[r]) * scr[u] : (T)bvr[u][r] * scr[u];
            }
        return;
    }
    for (int c = lane; c < pt.ccnt; c += 64) {
        const int g = A.cpl_rows[pt.cbeg + c];
        const T sc = scaled ? (T)A.inv_diag[g] : (T)1;
        if (A.cpl_fused) {
            double acc[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] = 0.0;
            for (int e = A.cpl_ptr[pt.cbeg + c]; e < A.cpl_ptr[pt.cbeg + c + 1]; e++)
#pragma unroll
                for (int r = 0; r < NR; r++) acc[r] += prod[(e - e0) * NR + r];
#pragma unroll
            for (int r = 0; r < NR; r++) xs[(g - s) * NR + r] = (T)(m.bv(roff + r, g) - acc[r]) * sc;
        } else {
#pragma unroll
            for (int r = 0; r < NR; r++)
                xs[(g - s) * NR + r] = (T)A.bc[(size_t)(pt.cbeg + c) * RS + roff + r] * sc;
        }
    }
}

// Value type of one block-solve LDS slot with NR interleaved right-hand
// sides (NR = 1 or 2): the NR values of a slot are read and written as one.
template <class T, int NR>
struct SlotVec;
template <class T>
struct SlotVec<T, 1> {
    using type = T;
};
template <class T>
struct SlotVec<T, 2> {
    typedef T type __attribute__((ext_vector_type(2)));
};
template <class T>
struct SlotVec<T, 4> {
    typedef T type __attribute__((ext_vector_type(4)));
};
template <class T>
struct SlotVec<T, 8> {
    typedef T type __attribute__((ext_vector_type(8)));
};

// Block solve with NR right-hand sides sharing each step record; x segments
// interleaved in LDS as xs[i*NR + r] (i < k+2, slot k holds 0).  A launch of
// nr_slots right-hand sides runs nr_slots / NR waves per partition
// (blockIdx.y picks columns roff .. roff+NR-1); the reduced vectors hold
// nr_slots interleaved values per index.
//
// block_pass is the body of one (partition p, RHS tile yb) wave; with FUSED
// it runs as a work item of the single-launch solve (k_spike_fused): pass 1
// publishes y at the dependents into x^(0) of the current parity (and
// refills the other parity's vectors of its dependents for the next solve),
// pass 2 polls x^(S) of the coupling columns instead of reading a reduced
// vector written by an earlier launch.
template <int G, int M, int D, int PASS, int NR, class T, bool FUSED>
__device__ __forceinline__ void block_pass(const BlockArgs &A, const FusedCtx &F, int p, int yb,
                                           unsigned char *smem_blk)
{
    T *xs = reinterpret_cast<T *>(smem_blk);
    constexpr int R = 64 / G;
    const DevPart pt = A.parts[p];
    // the step range is loaded with the partition record (one round trip
    // for both, not one after the other around the early returns)
    const int64_t sb = A.st.part_begin[p], se = A.st.part_begin[p + 1];
    asm volatile("" ::"s"(sb), "s"(se));
    const int lane = threadIdx.x;
    const int slot = lane / G, li = lane % G;
    const int s = pt.s, k = pt.k;
    // Pass 1 is needed when the partition holds dependents or has no coupling
    // (then its pass-1 result is final); pass 2 only when it has coupling.
    if (PASS == 1 && pt.hi1 < 0) return;
    if (PASS == 2 && pt.f >= k) return;
    const RhsMap m = rhs_map(A, pt);
    const int roff = yb * NR, RS = A.nr_slots;

    constexpr unsigned TB = sizeof(T);
    const int nst = (int)(se - sb);
    const unsigned nstu = (unsigned)(nst > 0 ? nst : 1);
    const __amdgpu_buffer_rsrc_t bval = make_rsrc(VT<T>::val(A) + (size_t)sb * M * 64, nstu * M * 64u * TB);
    const __amdgpu_buffer_rsrc_t bcol = make_rsrc(A.st.col + (size_t)sb * M * 64, nstu * M * 128u);
    const __amdgpu_buffer_rsrc_t brow = make_rsrc(A.st.row + (size_t)sb * R, nstu * R * 2u);
    // Register ring of D step records (slot d holds step base + d).  The
    // first D records are loaded before the right-hand sides, so their
    // latency overlaps with that prologue (measured on a 1M-row banded
    // matrix: 4 us per solve at one RHS; 38.6 -> 34.8 ms for a batch of 4096
    // matrices in tiles of up to 8 RHS).  The loop runs whole ring turns (the
    // last < D steps run in the unrolled tail below), each step refilling a
    // slot with the record D steps ahead (clamped to the last step), so
    // loads stay D steps in flight.
    // Columns and rows are held 32-bit: the 16-bit buffer loads zero-extend,
    // so no instruction widens them at their use (16-bit ring fields cost two
    // masking instructions per step on the step chain).
    T rv[D][M];
    unsigned rc[D][M];
    unsigned rrow[D];
    auto load_step = [&](int d, unsigned st) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            rv[d][m] = VT<T>::load(bval, (unsigned)(m * 64 + lane) * TB, st * (unsigned)(M * 64) * TB);
            rc[d][m] = buf_u16(bcol, (unsigned)(m * 64 + lane) * 2u, st * (unsigned)(M * 128));
        }
        rrow[d] = buf_u16(brow, (unsigned)slot * 2u, st * (unsigned)(R * 2));
    };
    if (nst > 0)
#pragma unroll
        for (int d = 0; d < D; d++) load_step(d, (unsigned)(d < nst ? d : nst - 1));

    // Pass 1 with one RHS: the rows of the first 64*UP dependents (written
    // back in the epilogue) are loaded here, with the right-hand
    // sides, so the epilogue issues no loads for them.  (Several RHS load
    // them per trip in the epilogue: holding them across the step loop raised
    // the register count of those variants.)
    constexpr int UP = 8;
    constexpr bool dep_early = PASS == 1 && NR == 1;
    int ir0[UP];
#pragma unroll
    for (int u = 0; u < UP; u++) ir0[u] = s;
    // (global rows: the partition start is subtracted at their use in the
    // epilogue; subtracting it here made the compiler wait for these loads,
    // and with them for every ring load issued before, ahead of the
    // right-hand sides' loads)
    if (dep_early && pt.dcnt > 0)
#pragma unroll
        for (int u = 0; u < UP; u++) ir0[u] = A.dep_rows[pt.dbeg + min(lane + 64 * u, pt.dcnt - 1)];
    // right-hand sides; pass 2 replaces them by b - C x_K on the coupling rows
    if (PASS == 1) {
        load_rhs<NR, T>(A, m, xs, s, k, roff, true);
    } else {
        pass2_rhs<NR, T, FUSED>(A, F, pt, xs, reinterpret_cast<double *>(smem_blk + xs_bytes<T>(A.kmax, NR)), true,
                                roff, RS);
    }
    if (lane == 0)
#pragma unroll
        for (int r = 0; r < NR; r++) xs[k * NR + r] = (T)0;
    __syncthreads();
    if (nst <= 0) goto epilogue;
    {
        // One step from ring slot d: LDS reads first (operands and the row's
        // accumulator), then the dot products.  Empty row slots name slot
        // k+1, so no lane tests for them.
        //
        // Lane 0 of the group starts its dot product from the row's
        // accumulator, the other lanes from the zero slot k, so no lane
        // branches.  With one RHS lane 0 writes the row and every other lane
        // its own spare slot k+2+lane; with several every lane of the slot
        // writes the same sums to the row.  Each RHS gets the same
        // arithmetic, so a column gives the same bits in any tile.  The NR
        // values of an LDS slot move as one vector.  The addresses select
        // between the row's LDS address and an address fixed for the lane
        // (opaque to the compiler, which would otherwise form each address
        // after its select), and the addresses of step d+1 are formed during
        // step d, after its LDS reads are issued, off the path from step d's
        // write to step d+1's reads; the empty asm statement keeps them there
        // (the scheduler otherwise hoisted them, and the ring waits they
        // carry, to the start of a ring turn, where they waited for ring
        // loads issued only a few steps before).  Measured on a 1M-row banded
        // matrix at 4 lanes per row, one RHS: 44.1 / 47.6 us per pass against
        // 44.7 / 48.8.
        static_assert(M == 1, "one step entry per lane");
        using SlotT = typename SlotVec<T, NR>::type;
        using LdsV = __attribute__((address_space(3))) SlotT;
        LdsV *const xl = (LdsV *)xs;
        LdsV *zero_p = xl + k, *spare_p = xl + (k + 2 + lane);
        asm volatile("" : "+v"(zero_p), "+v"(spare_p));
        LdsV *a_col = xl, *a_acc = xl, *a_wr = xl;
        auto addr = [&](int d) __attribute__((always_inline)) {
            unsigned row = rrow[d], col = rc[d][0];
            asm volatile("" : "+v"(row), "+v"(col));
            LdsV *const rp = xl + row;
            a_col = xl + col;
            a_acc = li == 0 ? rp : zero_p;
            a_wr = NR == 1 && li != 0 ? spare_p : rp;
        };
        addr(0);
        // mid(): ring refills placed after the step's LDS reads (the
        // scheduling barriers keep them there, in the reads' latency, rather
        // than between the previous write and these reads)
        auto step = [&](int d, auto &&mid) __attribute__((always_inline)) {
            const SlotT xv = *a_col;
            SlotT t = *a_acc;
            LdsV *const wr = a_wr;
            __builtin_amdgcn_sched_barrier(0);
            mid();
            addr((d + 1) % D);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (NR == 1) {
                t = fma(-rv[d][0], xv, t);
                t = group_sum<G>(t);
            } else {
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    t[r] = fma(-rv[d][0], xv[r], t[r]);
                    t[r] = group_sum<G>(t[r]);
                }
            }
            *wr = t;
        };
        // Refill slot d with step sidx + D (clamped to the last step), after
        // the last use of the slot's registers.
        auto refill = [&](int d, int sidx) __attribute__((always_inline)) {
            load_step(d, (unsigned)(sidx + D < nst ? sidx + D : nst - 1));
        };
        int base = 0;
        for (; base + D <= nst; base += D) {
#pragma unroll
            for (int d = 0; d < D; d++) {
                // refill the previous step's slot (its last use was that
                // step's FMA); slot D-1 of a turn is refilled in the next
                // turn's first step, which the tail never needs
                step(d, [&] {
                    if (d > 0)
                        refill(d - 1, base + d - 1);
                    else if (base > 0)
                        refill(D - 1, base - 1);
                });
            }
        }
#pragma unroll
        for (int d = 0; d < D; d++)
            if (base + d < nst) step(d, [] {});
    }
epilogue:
    __syncthreads();

    // The epilogue loops read UE (UW) items per lane and trip from LDS at
    // clamped indices before their conditional stores.  Global loads are
    // issued before any store: a load behind a store makes its wait cover
    // that store too (one counter for both), i.e. a write round trip per trip.
    constexpr int UE = NR >= 4 ? 1 : 4 / NR;
    constexpr int UW = NR >= 8 ? 1 : 8 / NR;
    if (PASS == 1) {
        int q = 0;   // dependents written so far
        if (dep_early) {
            T v[UP];
#pragma unroll
            for (int u = 0; u < UP; u++) v[u] = xs[ir0[u] - s];
#pragma unroll
            for (int u = 0; u < UP; u++)
                if (lane + 64 * u < pt.dcnt) {
                    const size_t at = (size_t)(pt.dbeg + lane + 64 * u) * RS + roff;
                    if constexpr (FUSED) store_agent(F.xcur + at, (double)v[u]);
                    else A.xred[at] = (double)v[u];
                }
            q = 64 * UP;
        }
        // the rest (several RHS, or more dependents): indices loaded per trip
        for (int b = q; b < pt.dcnt; b += 64 * UE) {
            const int q0 = b + lane;
            int ir[UE];
#pragma unroll
            for (int u = 0; u < UE; u++) ir[u] = A.dep_rows[pt.dbeg + min(q0 + 64 * u, pt.dcnt - 1)] - s;
#pragma unroll
            for (int u = 0; u < UE; u++) asm volatile("" ::"v"(ir[u]));   // all loads before the stores
            T v[UE][NR];
#pragma unroll
            for (int u = 0; u < UE; u++)
#pragma unroll
                for (int r = 0; r < NR; r++) v[u][r] = xs[ir[u] * NR + r];
#pragma unroll
            for (int u = 0; u < UE; u++)
                if (q0 + 64 * u < pt.dcnt)
#pragma unroll
                    for (int r = 0; r < NR; r++) {
                        const size_t at = (size_t)(pt.dbeg + q0 + 64 * u) * RS + roff + r;
                        if constexpr (FUSED) store_agent(F.xcur + at, (double)v[u][r]);
                        else A.xred[at] = (double)v[u][r];
                    }
        }
        if constexpr (FUSED) {
            // the next solve (other parity) starts with every value of these
            // dependents unpublished: nothing reads that parity during this
            // solve, and the launch boundary orders these stores before it
            const int nv = pt.dcnt * NR;
            for (int mm = 0; mm <= F.S; mm++)
                for (int i = lane; i < nv; i += 64)
                    store_agent(F.xnext + (size_t)mm * F.mstride + (size_t)(pt.dbeg + i / NR) * RS + roff + i % NR,
                                __longlong_as_double((long long)kNotReady));
        }
    }
    // pass 2 writes the solution; without coupling, pass 1's result is final
    if (PASS == 2 || pt.f >= k) {
        if (!A.accum) {
            // plain write-back: LDS reads and stores only
            for (int i0 = lane; i0 < k; i0 += 64 * UW) {
                T v[UW][NR];
#pragma unroll
                for (int u = 0; u < UW; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) v[u][r] = xs[min(i0 + 64 * u, k - 1) * NR + r];
#pragma unroll
                for (int u = 0; u < UW; u++)
                    if (i0 + 64 * u < k) {
                        const int gm = m.at(s + i0 + 64 * u);
#pragma unroll
                        for (int r = 0; r < NR; r++)
                            if (roff + r < m.nr) m.x[(size_t)(roff + r) * m.ldx + gm] = (double)v[u][r];
                    }
            }
        } else {
            // refinement update x += d: the old values of a trip are loaded
            // together (columns beyond nr read column nr-1, unused)
            for (int i0 = lane; i0 < k; i0 += 64 * UE) {
                double old[UE][NR];
#pragma unroll
                for (int u = 0; u < UE; u++) {
                    const int gm = m.at(s + min(i0 + 64 * u, k - 1));
#pragma unroll
                    for (int r = 0; r < NR; r++) old[u][r] = m.x[(size_t)min(roff + r, m.nr - 1) * m.ldx + gm];
                }
#pragma unroll
                for (int u = 0; u < UE; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) asm volatile("" ::"v"(old[u][r]));   // all loads first
#pragma unroll
                for (int u = 0; u < UE; u++)
                    if (i0 + 64 * u < k) {
                        const int gm = m.at(s + i0 + 64 * u);
#pragma unroll
                        for (int r = 0; r < NR; r++)
                            if (roff + r < m.nr)
                                m.x[(size_t)(roff + r) * m.ldx + gm] = old[u][r] + (double)xs[(i0 + 64 * u) * NR + r];
                    }
            }
        }
    }
}

template <int G, int M, int D, int PASS, int NR, class T>
__global__ __launch_bounds__(64) void k_block_solve(BlockArgs A)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_blk[];
    block_pass<G, M, D, PASS, NR, T, false>(A, FusedCtx{}, blockIdx.x, blockIdx.y, smem_blk);
}

// One sweep item of the fused solve: x^(m) = y - W x^(m-1) on the reduced
// rows [ch.x, ch.y) (entries [ch.z, ch.w) of the reduced CSR) for the NR
// right-hand sides roff .. roff+NR-1 (nr_slots = RS values per index).  The
// entries are taken 64*UW at a time (UW*NR = 32 values per lane): values and
// columns, then the polled x^(m-1) values; the products go to the LDS
// scratch prod.  Each row's products of the batch are then summed by a group
// of 8 lanes (lane li adds entries li, li+8, ... in order, then a fixed DPP
// tree), and the batch sums are added in batch order to the row's sum in
// rowacc; the result depends neither on NR nor on the lane that runs the row.
// y (the row's own x^(0)) is loaded with the row bounds and polled again only
// if it was not yet published then.  LDS: prod (64*32 doubles), rowacc
// (kFusedRowsPerItem*NR doubles), the row bounds (kFusedRowsPerItem+1 ints).
template <int NR>
__device__ __forceinline__ void sweep_item(const FusedArgs &FA, const FusedCtx &F, int4 ch, int m, int roff, int RS,
                                           unsigned char *lds)
{
    constexpr int UW = 32 / NR >= 1 ? 32 / NR : 1;
    constexpr int VR = kFusedRowsPerItem / 64;
    double *prod = reinterpret_cast<double *>(lds);
    double *rowacc = prod + 64 * 32;
    int *rbs = reinterpret_cast<int *>(rowacc + kFusedRowsPerItem * NR);
    const int lane = threadIdx.x, g = lane >> 3, li = lane & 7;
    const int nrows = ch.y - ch.x;
    const double *xprev = F.xcur + (size_t)(m - 1) * F.mstride;
    double *xout = F.xcur + (size_t)m * F.mstride;
    unsigned yidx[VR * NR];
    bool yin[VR * NR];
    double yv[VR * NR];
#pragma unroll
    for (int v = 0; v < VR; v++) {
        const int j = lane + 64 * v;
        const bool ok = j < nrows;
        const int qc = ch.x + (ok ? j : nrows - 1);
        const int rb = FA.rp[qc];
        if (ok) rbs[j] = rb;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            yidx[v * NR + r] = (unsigned)qc * (unsigned)RS + (unsigned)(roff + r);
            yin[v * NR + r] = ok;
            yv[v * NR + r] = load_agent(F.xcur + yidx[v * NR + r]);
        }
    }
    if (lane == 0) rbs[nrows] = ch.w;
    for (int eb = ch.z; eb < ch.w; eb += 64 * UW) {
        unsigned idx[UW * NR];
        bool in[UW * NR];
        double w[UW], xv[UW * NR];
#pragma unroll
        for (int u = 0; u < UW; u++) {
            const int e = eb + lane + 64 * u;
            const bool ok = e < ch.w;
            const int ee = ok ? e : ch.w - 1;
            const int c = FA.ci[ee];
            const double wv = FA.v[ee];
            w[u] = ok ? wv : 0.0;
#pragma unroll
            for (int r = 0; r < NR; r++) {
                idx[u * NR + r] = (unsigned)c * (unsigned)RS + (unsigned)(roff + r);
                in[u * NR + r] = ok;
            }
        }
        (void)poll_values<UW * NR>(xprev, idx, in, xv, F.spin_limit, F.abort_word, F.fail_word);
        __syncthreads();   // the previous batch's sums have read prod (and rbs is written)
#pragma unroll
        for (int u = 0; u < UW; u++)
#pragma unroll
            for (int r = 0; r < NR; r++) prod[(lane + 64 * u) * NR + r] = w[u] * xv[u * NR + r];
        __syncthreads();
        const int ee = min(ch.w, eb + 64 * UW);
        for (int j = g; j < nrows; j += 8) {
            const int lo = max(rbs[j], eb), hi = min(rbs[j + 1], ee);
            double part[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) part[r] = 0.0;
            for (int e = lo + li; e < hi; e += 32) {
                double t4[4][NR];
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) t4[u][r] = e + 8 * u < hi ? prod[(e + 8 * u - eb) * NR + r] : 0.0;
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) part[r] += t4[u][r];
            }
#pragma unroll
            for (int r = 0; r < NR; r++) {
                part[r] = group_sum<8>(part[r]);
                if (li == 0) rowacc[j * NR + r] = (eb == ch.z ? 0.0 : rowacc[j * NR + r]) + part[r];
            }
        }
    }
    __syncthreads();
    bool yready = true;
#pragma unroll
    for (int q = 0; q < VR * NR; q++)
        if (yin[q] && not_ready(yv[q])) yready = false;
    if (!yready) (void)poll_values<VR * NR>(F.xcur, yidx, yin, yv, F.spin_limit, F.abort_word, F.fail_word);
#pragma unroll
    for (int v = 0; v < VR; v++)
#pragma unroll
        for (int r = 0; r < NR; r++)
            if (yin[v * NR + r]) {
                const int j = lane + 64 * v;
                const double sum = ch.w > ch.z ? rowacc[j * NR + r] : 0.0;
                store_agent(xout + yidx[v * NR + r], yv[v * NR + r] - sum);
            }
}

// Single-launch Spike solve (FusedArgs): a persistent grid whose workgroups
// (one wave each) take tickets from the in-order queue until it is empty.
// Ticket order: pass-1 items (partition p, RHS tile), sweep items of sweep 1
// .. S (chunk, tile), pass-2 items.  An item polls only values published by
// items of smaller tickets, which running workgroups hold, so the lowest
// unfinished item never waits for an unfinished one.  The next ticket is
// taken when an item starts, so its round trip overlaps the item.
template <int G, int M, int D, int NR, class T>
__global__ __launch_bounds__(64) void k_spike_fused(BlockArgs A, FusedArgs FA, int P)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_f[];
    const int lane = threadIdx.x;
    const int RS = A.nr_slots, tiles = RS / NR;
    unsigned *c_par = FA.ctr, *c_queue = FA.ctr + 32, *c_exit = FA.ctr + 64;
    const unsigned par = __hip_atomic_load(c_par, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1u;
    FusedCtx F;
    F.mstride = (size_t)FA.nred * RS;
    const size_t half = (size_t)(FA.S + 1) * F.mstride;
    F.xcur = FA.X + (size_t)par * half;
    F.xnext = FA.X + (size_t)(par ^ 1u) * half;
    F.S = FA.S;
    F.spin_limit = FA.spin_limit;
    F.abort_word = FA.abort_word;
    F.fail_word = FA.fail_word;
    const unsigned np = (unsigned)P * tiles, nc = (unsigned)FA.nchunks * tiles;
    const unsigned p2 = np + (unsigned)FA.S * nc, total = p2 + np;
    unsigned raw = 0;
    if (lane == 0) raw = __hip_atomic_fetch_add(c_queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned t = (unsigned)__builtin_amdgcn_readfirstlane((int)raw);
    while (t < total) {
        if (lane == 0) raw = __hip_atomic_fetch_add(c_queue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t < np) {
            block_pass<G, M, D, 1, NR, T, true>(A, F, (int)(t % (unsigned)P), (int)(t / (unsigned)P), smem_f);
        } else if (t < p2) {
            const unsigned u = t - np, w = u % nc;
            const int4 ch = reinterpret_cast<const int4 *>(FA.chunks)[w % (unsigned)FA.nchunks];
            sweep_item<NR>(FA, F, ch, 1 + (int)(u / nc), (int)(w / (unsigned)FA.nchunks) * NR, RS, smem_f);
        } else {
            const unsigned u = t - p2;
            block_pass<G, M, D, 2, NR, T, true>(A, F, (int)(u % (unsigned)P), (int)(u / (unsigned)P), smem_f);
        }
        __syncthreads();   // the next item may overwrite what this one left in LDS
        t = (unsigned)__builtin_amdgcn_readfirstlane((int)raw);
    }
    // The last workgroup to leave resets the queue and the abort word and
    // flips the parity: every workgroup read the parity before its exit
    // increment, and nothing of this launch touches the counters afterwards.
    if (lane == 0 && __hip_atomic_fetch_add(c_exit, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
        __hip_atomic_store(c_queue, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(c_exit, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(FA.abort_word, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(c_par, par ^ 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// X[0 .. len) = kNotReady; ctr[0 .. 96) = 0; *abort_word = 0.
__global__ __launch_bounds__(256) void k_fused_reset(size_t len, double *X, unsigned *ctr, int *abort_word)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < len) X[i] = __longlong_as_double((long long)kNotReady);
    if (i < 96) ctr[i] = 0u;
    if (i == 0) *abort_word = 0;
}

// Forwarding block solve (16 lanes per row slot, 4 slots per step).  Values
// are pre-scaled by 1/diag; xs[r] starts as b_r/d_r.  Step s of a partition:
//   E_s  = xs[row] - sum_early val * xs[col]   (lanes of the slot, DPP sum)
//   x_s  = E_s - sum_t late[s][slot][t] * x_{s-1}[t]
// where x_{s-1}[t] comes from the previous step's registers (row t of the
// wave, gathered by permlane swaps).  E_{s+1} only depends on steps <= s-1,
// so its LDS reads and its reduction are issued during step s and stay off
// the step-to-step chain, which is two swap levels + the late FMAs.  Rows of
// empty slots are k+1; non-leader lanes write to their own spare slot.
template <int D, int PASS>
__global__ __launch_bounds__(64) void k_block_solve_fwd(BlockArgs A)
{
    extern __shared__ double xs[];
    constexpr int G = 16, R = 4;
    const int p = blockIdx.x;
    const DevPart pt = A.parts[p];
    const int lane = threadIdx.x;
    const int slot = lane / G, li = lane % G;
    const int s = pt.s, k = pt.k;
    if (PASS == 1 && pt.hi1 < 0) return;
    if (PASS == 2 && pt.f >= k) return;
    const double *__restrict__ invd = A.inv_diag;
    const RhsMap m = rhs_map(A, pt);

    if (PASS == 1) {
        for (int r = lane; r < k; r += 64) {
            const int g = s + r;
            xs[r] = m.bv(0, g) * invd[g];
        }
    } else {
        pass2_rhs<1, double, false>(A, FusedCtx{}, pt, xs, xs + A.kmax + XS_PAD, true, 0, 1);
    }
    if (lane == 0) {
        xs[k] = 0.0;
        xs[k + 1] = 0.0;
    }
    __syncthreads();

    const int64_t sb = A.st.part_begin[p];
    const int nst = (int)(A.st.part_begin[p + 1] - sb);
    if (nst > 0) {
        const unsigned nstu = (unsigned)nst;
        const __amdgpu_buffer_rsrc_t bval = make_rsrc(A.st.val + (size_t)sb * 64, nstu * 512u);
        const __amdgpu_buffer_rsrc_t bcol = make_rsrc(A.st.col + (size_t)sb * 64, nstu * 128u);
        const __amdgpu_buffer_rsrc_t brow = make_rsrc(A.st.row + (size_t)sb * R, nstu * R * 2u);
        const __amdgpu_buffer_rsrc_t blate = make_rsrc(A.st.late + (size_t)sb * R * R, nstu * R * R * 8u);
        // register ring (32-bit index fields, as in k_block_solve)
        double rv[D];
        unsigned rc[D];
        unsigned rrow[D];
        double rl[D][R];
#pragma unroll
        for (int d = 0; d < D; d++) {
            rv[d] = 0.0;
            rc[d] = (unsigned)k;
            rrow[d] = (unsigned)(k + 1);
#pragma unroll
            for (int t = 0; t < R; t++) rl[d][t] = 0.0;
        }
        double E = 0.0, xprev = 0.0;
        const unsigned ulane = (unsigned)lane, uslot = (unsigned)slot;
        auto refill = [&](int d, int sidx) __attribute__((always_inline)) {
            const unsigned nxt = (unsigned)(sidx + D < nst ? sidx + D : nst - 1);
            rv[d] = buf_f64(bval, ulane * 8u, nxt * 512u);
            rc[d] = buf_u16(bcol, ulane * 2u, nxt * 128u);
            rrow[d] = buf_u16(brow, uslot * 2u, nxt * (unsigned)(R * 2));
            buf_2f64(blate, uslot * 32u, nxt * (unsigned)(R * R * 8), rl[d][0], rl[d][1]);
            buf_2f64(blate, uslot * 32u + 16u, nxt * (unsigned)(R * R * 8), rl[d][2], rl[d][3]);
        };
        // late part of the step in ring slot d, given its early sum
        auto finish = [&](int d, double e) __attribute__((always_inline)) {
            double xp[R];
            rows4_f64(xprev, xp);
            const double l = fma(rl[d][1], xp[1], rl[d][0] * xp[0]) + fma(rl[d][3], xp[3], rl[d][2] * xp[2]);
            const double x = e - l;
            const unsigned wr = li == 0 ? (unsigned)rrow[d] : (unsigned)(k + 2 + lane);
            xs[wr] = x;
            xprev = x;
        };
        // One step: issue the LDS reads of the next step's early part, run
        // this step's late part (covers the LDS latency), then reduce the
        // next step's early part.  Scheduling barriers keep that order.
        auto pos = [&](int d, bool refill_slot, int sidx) __attribute__((always_inline)) {
            const int dn = (d + 1) % D;
            const double xv = xs[rc[dn]];
            const double bv = xs[li == 0 ? (unsigned)rrow[dn] : (unsigned)k];
            __builtin_amdgcn_sched_barrier(0);
            finish(d, E);
            __builtin_amdgcn_sched_barrier(0);
            E = group_sum<G>(fma(-rv[dn], xv, bv));
            if (refill_slot) refill(d, sidx);
        };
        int base = -D;
        for (; base + D <= nst; base += D) {
#pragma unroll
            for (int d = 0; d < D; d++) pos(d, true, base + d);
        }
#pragma unroll
        for (int d = 0; d < D; d++)
            if (base + d < nst) pos(d, false, 0);
    }
    __syncthreads();

    if (PASS == 1)
        for (int q = lane; q < pt.dcnt; q += 64) A.xred[pt.dbeg + q] = xs[A.dep_rows[pt.dbeg + q] - s];
    if (PASS == 2 || pt.f >= k)
        for (int r = lane; r < k; r += 64) put_x(m.x, m.at(s + r), xs[r], A.accum);
}

template <int G, int D, int PASS, int NR, class T>
hipError_t launch_gdt(int P, const BlockArgs &a, hipStream_t s)
{
    constexpr int M = 1;
    size_t lds = xs_bytes<T>(a.kmax, NR) + (PASS == 2 && a.cpl_fused ? (size_t)a.cmax * NR * sizeof(double) : 0);
    auto kern = k_block_solve<G, M, D, PASS, NR, T>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(P, a.nr_slots / NR), dim3(64), lds, s, a);
    return hipGetLastError();
}

template <int G, int D, int PASS, int NR>
hipError_t launch_gd(int P, const BlockArgs &a, hipStream_t s)
{
    return a.fp32 ? launch_gdt<G, D, PASS, NR, float>(P, a, s) : launch_gdt<G, D, PASS, NR, double>(P, a, s);
}

template <int G, int PASS>
hipError_t launch_g(int P, const BlockArgs &a, hipStream_t s)
{
    if (a.tile < 1 || a.tile > a.nr_slots || a.nr_slots % a.tile) return hipErrorInvalidValue;
    switch (a.tile) {
    case 1: return a.ring >= 16 ? launch_gd<G, 16, PASS, 1>(P, a, s) : launch_gd<G, 8, PASS, 1>(P, a, s);
    case 2: return a.ring >= 16 ? launch_gd<G, 16, PASS, 2>(P, a, s) : launch_gd<G, 8, PASS, 2>(P, a, s);
    case 4: return launch_gd<G, 8, PASS, 4>(P, a, s);
    case 8: return launch_gd<G, 8, PASS, 8>(P, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int D, int PASS>
hipError_t launch_fwd(int P, const BlockArgs &a, hipStream_t s)
{
    size_t lds = (size_t)(a.kmax + XS_PAD + (PASS == 2 && a.cpl_fused ? a.cmax : 0)) * sizeof(double);
    auto kern = k_block_solve_fwd<D, PASS>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(P), dim3(64), lds, s, a);
    return hipGetLastError();
}

template <int PASS>
hipError_t launch_pass(int P, const BlockArgs &a, hipStream_t s)
{
    if (a.st.fwd) {
        if (a.st.G != 16 || a.nr_slots != 1 || a.fp32) return hipErrorInvalidValue;
        return a.ring >= 16 ? launch_fwd<16, PASS>(P, a, s) : launch_fwd<8, PASS>(P, a, s);
    }
    switch (a.st.G) {
    case 1: return launch_g<1, PASS>(P, a, s);
    case 2: return launch_g<2, PASS>(P, a, s);
    case 4: return launch_g<4, PASS>(P, a, s);
    case 8: return launch_g<8, PASS>(P, a, s);
    case 16: return launch_g<16, PASS>(P, a, s);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------- reduced

// Exact forward substitution of the unit lower reduced system, one workgroup,
// level by level; NR right-hand sides interleaved (X[q*NR + r]).
template <bool IN_LDS, int NR>
__global__ __launch_bounds__(1024) void k_reduced_solve(int nred, int nlev, const int *__restrict__ levptr,
                                                        const int *__restrict__ rlist, const int *__restrict__ rp,
                                                        const int *__restrict__ ci, const double *__restrict__ v,
                                                        double *xred)
{
    extern __shared__ double xl[];
    const int tid = threadIdx.x, nt = blockDim.x;
    double *X = xred;
    if (IN_LDS) {
        for (int i = tid; i < nred * NR; i += nt) xl[i] = xred[i];
        __syncthreads();
        X = xl;
    }
    for (int L = 0; L < nlev; L++) {
        const int b = levptr[L], e = levptr[L + 1];
        for (int idx = b + tid; idx < e; idx += nt) {
            const int q = rlist[idx];
            double acc[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] = 0.0;
            for (int j = rp[q]; j < rp[q + 1]; j++) {
                const double w = v[j];
                const int c = ci[j];
#pragma unroll
                for (int r = 0; r < NR; r++) acc[r] = fma(w, X[c * NR + r], acc[r]);
            }
#pragma unroll
            for (int r = 0; r < NR; r++) X[q * NR + r] -= acc[r];
        }
        __syncthreads();
    }
    if (IN_LDS)
        for (int i = tid; i < nred * NR; i += nt) xred[i] = xl[i];
}

// One Jacobi sweep of the unit lower reduced system x = y - W x over a list
// of rows: xout[r] = y[r] - (W xin)[r] for the listed rows, G lanes per row,
// fixed summation order.  Rows of the list flagged with the top bit only copy
// xin[r] to xout[r] (rows that converged in the previous sweep, so that both
// buffers hold their final value).  Each lane issues all of its column and
// value loads, then all of its gathers, before the FMAs; each lane sums its
// entries in ascending order, then the group sum runs in a fixed order.
template <int G, int U, int NR>
__global__ __launch_bounds__(256) void k_jacobi_sweep(int nrows, const int4 *__restrict__ rows,
                                                      const int *__restrict__ ci, const double *__restrict__ v,
                                                      const double *__restrict__ y, const double *__restrict__ xin,
                                                      double *__restrict__ xout)
{
    const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int q = tid / G, li = tid % G;
    int r = -1;
    bool copy = false;
    int4 dsc = make_int4(0, 0, 0, 0);
    if (q < nrows) {
        dsc = rows[q];                      // {row | copy flag, first entry, end}
        copy = dsc.x < 0;
        r = dsc.x & 0x7fffffff;
    }
    // y[r] (xin[r] for a copy row), loaded with the entries
    double base[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) base[t] = 0.0;
    if (li == 0 && r >= 0)
#pragma unroll
        for (int t = 0; t < NR; t++) base[t] = copy ? xin[(size_t)r * NR + t] : y[(size_t)r * NR + t];
    double acc[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) acc[t] = 0.0;
    if (r >= 0 && !copy) {
        // U entries per lane per trip, the last trip predicated, so a row of
        // up to U * G entries issues all of its loads at once
        const int e = dsc.z;
        for (int j = dsc.y + li; j < e; j += U * G) {
            int c[U];
            double w[U], xv[U][NR];
#pragma unroll
            for (int u = 0; u < U; u++) {
                // unconditional loads at a clamped index (no branch between
                // the loads), masked entries then use the row itself
                const bool ok = j + u * G < e;
                const int jj = ok ? j + u * G : e - 1;
                const int cv = ci[jj];
                const double wv = v[jj];
                c[u] = ok ? cv : r;
                w[u] = ok ? wv : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int t = 0; t < NR; t++) xv[u][t] = xin[(size_t)c[u] * NR + t];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int t = 0; t < NR; t++)
                    if (j + u * G < e) acc[t] = fma(w[u], xv[u][t], acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NR; t++) acc[t] = group_sum<G>(acc[t]);
    if (li == 0 && r >= 0)
#pragma unroll
        for (int t = 0; t < NR; t++) xout[(size_t)r * NR + t] = base[t] - acc[t];
}

// One Jacobi sweep m over items (runs of one partition's dependents; see
// JacobiItems).  A workgroup first gathers xin at its item's distinct columns
// into LDS, NR values per column (only when the list code has the top bit
// set: some row of the item computes in this sweep), then walks its rows
// with 32 groups of 8 lanes: rows with need >= m compute xout = y - W xin
// (lane li adds entries li, li+8, ... in order, four per trip, then
// group_sum<8>: the order of k_jacobi_sweep<8, 4, NR>, so both give the same
// bits), rows with need == m-1 (or need == 0 at m == 2) copy xin to xout,
// and the others (already final in both buffers) are left alone.
template <int NR>
__global__ __launch_bounds__(256) void k_jacobi_items(JacobiItems J, int m, const int *__restrict__ list,
                                                      const double *__restrict__ xin, double *__restrict__ xout)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_j[];
    double *xk = reinterpret_cast<double *>(smem_j);
    constexpr int UG = 4;
    const int tid = threadIdx.x, g = tid >> 3, li = tid & 7;
    const int code = list[blockIdx.x];
    const int4 it = reinterpret_cast<const int4 *>(J.items)[code & 0x7fffffff];
    const int q0 = it.x, q1 = it.y, kb = it.z, nkv = (it.w - it.z) * NR;
    if (code < 0) {
        for (int i0 = 0; i0 < nkv; i0 += 256 * UG) {
            double xv[UG];
#pragma unroll
            for (int u = 0; u < UG; u++) {
                const int i = min(i0 + tid + 256 * u, nkv - 1);
                xv[u] = xin[(size_t)J.kcol[kb + i / NR] * NR + (unsigned)(i % NR)];
            }
#pragma unroll
            for (int u = 0; u < UG; u++) {
                const int i = i0 + tid + 256 * u;
                if (i < nkv) xk[i] = xv[u];
            }
        }
        __syncthreads();
    }
    for (int q = q0 + g; q < q1; q += 32) {
        const int nd = J.need[q];
        if (nd >= m) {
            const int e1 = J.rp[q + 1];
            double acc[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] = 0.0;
            for (int e = J.rp[q] + li; e < e1; e += 32) {
                int c[4];
                double w[4], xv[4][NR];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const bool ok = e + 8 * u < e1;
                    const int ee = ok ? e + 8 * u : e1 - 1;
                    const int cv = J.lci[ee];
                    const double wv = J.v[ee];
                    c[u] = ok ? cv : 0;
                    w[u] = ok ? wv : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < NR; r++) xv[u][r] = xk[c[u] * NR + r];
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (e + 8 * u < e1)
#pragma unroll
                        for (int r = 0; r < NR; r++) acc[r] = fma(w[u], xv[u][r], acc[r]);
            }
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] = group_sum<8>(acc[r]);
            if (li == 0)
#pragma unroll
                for (int r = 0; r < NR; r++) xout[(size_t)q * NR + r] = J.y[(size_t)q * NR + r] - acc[r];
        } else if (nd == m - 1 || (nd == 0 && m == 2)) {
            if (li < NR)
                for (int r = li; r < NR; r += 8) xout[(size_t)q * NR + r] = xin[(size_t)q * NR + r];
        }
    }
}

// ---------------------------------------------------------------- coupling

// One workgroup per partition: threads over its coupling rows.
template <int NR>
__global__ __launch_bounds__(256) void k_coupling(const DevPart *__restrict__ parts, const int *__restrict__ rows,
                                                  const int *__restrict__ ptr, const int *__restrict__ col,
                                                  const double *__restrict__ val, const double *__restrict__ xred,
                                                  BlockArgs A, double *__restrict__ bc)
{
#pragma clang fp contract(off)
    const DevPart pt = parts[blockIdx.x];
    const RhsMap m = rhs_map(A, pt);
    // U entries per trip at clamped indices: values and columns, then the
    // gathers, then the sums in entry order (8 per trip, 4 with 8 RHS to
    // keep four waves per SIMD: a coupling row of the banded matrices has up
    // to about 16 entries, and each trip is two dependent round trips)
    constexpr int U = NR >= 8 ? 4 : 8;
    for (int c = pt.cbeg + threadIdx.x; c < pt.cbeg + pt.ccnt; c += blockDim.x) {
        const int g = rows[c];
        const int j0 = ptr[c], j1 = ptr[c + 1];
        // the row's right-hand sides, loaded with its first entries
        double bv[NR];
#pragma unroll
        for (int t = 0; t < NR; t++) bv[t] = m.bv(t, g);
        double acc[NR];
#pragma unroll
        for (int t = 0; t < NR; t++) acc[t] = 0.0;
        for (int j = j0; j < j1; j += U) {
            double w[U], xv[U][NR];
            size_t x0[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int jj = min(j + u, j1 - 1);
                w[u] = val[jj];
                x0[u] = (size_t)col[jj] * NR;
            }
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int t = 0; t < NR; t++) xv[u][t] = xred[x0[u] + t];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j + u < j1)
#pragma unroll
                    for (int t = 0; t < NR; t++) acc[t] += w[u] * xv[u][t];
        }
#pragma unroll
        for (int t = 0; t < NR; t++) bc[(size_t)c * NR + t] = bv[t] - acc[t];
    }
}

// ---------------------------------------------------------------- residual

// r = b - A x in FP64, one workgroup per block of at most 64 lower-form rows
// (the row blocks of a factor), G = 4 lanes per row (strided over its
// entries), NR right-hand sides (nr of the block's matrix active).  r uses
// the same caller addressing as b (ldr replaces ldb when the block has no
// leading dimension of its own).  Each lane loads U entries per trip at
// clamped indices, then their x values, then sums them in entry order; lane
// sums are combined with DPP in a fixed order.
template <int NR>
__global__ __launch_bounds__(256) void k_residual(const DevPart *__restrict__ parts, const int *__restrict__ ia,
                                                  const int *__restrict__ ja, const double *__restrict__ a,
                                                  BlockArgs A, double *__restrict__ r, int ldr, int c0)
{
    constexpr int G = 4, U = NR >= 8 ? 2 : (NR >= 4 ? 4 : 8);
    const DevPart pt = parts[blockIdx.x];
    const RhsMap m = rhs_map(A, pt);
    const int ld_r = pt.ld ? pt.ld : ldr;
    double *rp = r + pt.boff;
    const int li = threadIdx.x % G;
    for (int base = 0; base < pt.k; base += 256 / G) {
        const int lr = base + threadIdx.x / G;
        const int row = pt.s + lr;
        // the row's right-hand sides, loaded with its entries rather than
        // after the gathers (columns beyond nr read column nr-1, unused)
        const int gb = m.at(pt.s + min(lr, pt.k - 1));
        double bv[NR];
#pragma unroll
        for (int t = 0; t < NR; t++) bv[t] = m.b[(size_t)min(c0 + t, m.nr - 1) * m.ldb + gb];
        double acc[NR];
#pragma unroll
        for (int t = 0; t < NR; t++) acc[t] = 0.0;
        if (lr < pt.k) {
            const int e1 = ia[row + 1];
            for (int e = ia[row] + li; e < e1; e += U * G) {
                double v[U], xv[U][NR];
                int cm[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int ee = min(e + u * G, e1 - 1);
                    v[u] = a[ee];
                    cm[u] = m.at(ja[ee]);
                }
#pragma unroll
                for (int u = 0; u < U; u++)
#pragma unroll
                    for (int t = 0; t < NR; t++)
                        xv[u][t] = m.x[(size_t)min(c0 + t, m.nr - 1) * m.ldx + cm[u]];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (e + u * G < e1)
#pragma unroll
                        for (int t = 0; t < NR; t++) acc[t] = fma(v[u], xv[u][t], acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < NR; t++) acc[t] = group_sum<G>(acc[t]);
        if (lr < pt.k && li == 0) {
            const int gm = m.at(row);
#pragma unroll
            for (int t = 0; t < NR; t++)
                if (c0 + t < m.nr) rp[(size_t)(c0 + t) * ld_r + gm] = bv[t] - acc[t];
        }
    }
}

// ---------------------------------------------------------------- levels

// One dependency level of a lower-form triangular solve: its rows depend only
// on rows of earlier levels, so G lanes per row compute x[row] = (b[row] -
// sum_j a_rj x[j]) / a_rr in parallel, reading earlier x from global memory
// (the previous launch has completed).  Values and accumulation in T (FP64,
// or FP32 for the refined mode); b and x are FP64 in the caller's index order
// (flip: n-1-i), NR right-hand sides.  Each lane sums its entries (first +
// lane, stride G) in ascending order, then the group sum runs in a fixed
// order.  Row records and entries are stored in level order, so a lane issues
// its record load, then all of its column and value loads (U per trip, the
// last trip predicated), then all of its gathers.
template <int G, int NR, class T>
__global__ __launch_bounds__(256) void k_level(int r0, int cnt, const int4 *__restrict__ rec,
                                               const int *__restrict__ lja, const T *__restrict__ lval,
                                               const T *__restrict__ invd, const double *__restrict__ b, int ldb,
                                               double *x, int ldx, int nr, int n, int flip)
{
    constexpr int U = 4;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int q = gid / G, li = gid % G;
    const int4 rc = q < cnt ? rec[r0 + q] : make_int4(-1, 0, 0, 0);
    const bool ok = rc.x >= 0;   // padding records have row -1 and no entries
    const int row = rc.x;
    const int gm = flip ? n - 1 - row : row;
    // 1/diag and b of the row (row 0 for padding, unused), loaded with the
    // entries rather than after the gathers and the group sum
    const int rs = ok ? row : 0, gs = flip ? n - 1 - rs : rs;
    const T d = invd[rs];
    double bv[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) bv[t] = b[(size_t)min(t, nr - 1) * ldb + gs];
    T acc[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) acc[t] = (T)0;
    for (int e = rc.y + li; e < rc.z; e += U * G) {
        int c[U];
        T w[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // unconditional loads at a clamped index, masked entries then
            // use the row itself (not used)
            const bool in = e + u * G < rc.z;
            const int ee = in ? e + u * G : rc.z - 1;
            const int cl = lja[ee];
            const T wl = lval[ee];
            const int cc = in ? cl : row;
            c[u] = flip ? n - 1 - cc : cc;
            w[u] = in ? wl : (T)0;
        }
        // (columns t >= nr read column nr-1: unconditional, so the gathers
        // are issued together; their sums are not stored)
        T xv[U][NR];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int t = 0; t < NR; t++) xv[u][t] = (T)x[(size_t)min(t, nr - 1) * ldx + c[u]];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int t = 0; t < NR; t++)
                if (e + u * G < rc.z) acc[t] = fma(w[u], xv[u][t], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < NR; t++) acc[t] = group_sum<G>(acc[t]);
    if (ok && li == 0) {
#pragma unroll
        for (int t = 0; t < NR; t++)
            if (t < nr) x[(size_t)t * ldx + gm] = (double)(((T)bv[t] - acc[t]) * d);
    }
}

// Single-RHS level step of an FP32 level factor with one FP64 refinement
// step, G lanes per row as in k_level<G, 1, float>: x0 = (b - L x0) / d
// (k_level's operations, same bits), r = b - A x0 in FP64 (lane partial sums
// over the row's entries, the group sum, then the diagonal term; rec.w is the
// row's first entry in the lower-form CSR a64, diagonal last) and
// dx = (r - L dx) / d.  x0 goes to x and dx to dl, both at the caller's
// index; x += dl after the last level completes the step.  One launch per
// level instead of two, plus no residual launch.
template <int G>
__global__ __launch_bounds__(256) void k_level_ref1(int r0, int cnt, const int4 *__restrict__ rec,
                                                    const int *__restrict__ lja, const float *__restrict__ lval,
                                                    const float *__restrict__ invd, const double *__restrict__ a64,
                                                    const double *b, double *x, double *dl, int n, int flip)
{
    constexpr int U = 4;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int q = gid / G, li = gid % G;
    const int4 rc = q < cnt ? rec[r0 + q] : make_int4(-1, 0, 0, 0);
    const bool ok = rc.x >= 0;   // padding records have row -1 and no entries
    const int row = rc.x;
    // 1/diag, b and the FP64 diagonal of the row (row 0 for padding, unused),
    // loaded with the entries
    const int rs = ok ? row : 0, gs = flip ? n - 1 - rs : rs;
    const float d = invd[rs];
    const double bv = b[gs];
    const int abase = rc.w - rc.y;   // a64 index of level-ordered entry e: abase + e
    const double ad = a64[ok ? abase + rc.z : 0];
    float acc0 = 0.f, acc1 = 0.f;
    double acc64 = 0.0;
    for (int e = rc.y + li; e < rc.z; e += U * G) {
        int c[U];
        float w[U];
        double w64[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // unconditional loads at a clamped index, masked entries then
            // use the row itself (not used)
            const bool in = e + u * G < rc.z;
            const int ee = in ? e + u * G : rc.z - 1;
            const int cl = lja[ee];
            const float wl = lval[ee];
            const double wl64 = a64[abase + ee];
            const int cc = in ? cl : row;
            c[u] = flip ? n - 1 - cc : cc;
            w[u] = in ? wl : 0.f;
            w64[u] = in ? wl64 : 0.0;
        }
        double xv[U], dv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            xv[u] = x[c[u]];
            dv[u] = dl[c[u]];
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (e + u * G < rc.z) {
                acc0 = fma(w[u], (float)xv[u], acc0);
                acc64 = fma(w64[u], xv[u], acc64);
                acc1 = fma(w[u], (float)dv[u], acc1);
            }
    }
    acc0 = group_sum<G>(acc0);
    acc64 = group_sum<G>(acc64);
    acc1 = group_sum<G>(acc1);
    const float x0 = ((float)bv - acc0) * d;
    const double res = bv - fma(ad, (double)x0, acc64);
    const float dx = ((float)res - acc1) * d;
    if (ok && li == 0) {
        x[gs] = (double)x0;
        dl[gs] = (double)dx;
    }
}

// Multi-RHS level step on interleaved vectors (v[i*NRP + r]): NRP lanes per
// row, one per right-hand side, each summing the row's entries in order
// (the entry's value and column are the same for all lanes of the row; the
// x loads of one entry are one contiguous NRP*8-byte segment).  Record,
// then entries (U per trip, the last trip predicated), then gathers, as in
// k_level.
template <int NRP, class T>
__global__ __launch_bounds__(256) void k_level_il(int r0, int cnt, const int4 *__restrict__ rec,
                                                  const int *__restrict__ lja, const T *__restrict__ lval,
                                                  const T *__restrict__ invd, const double *__restrict__ bi,
                                                  double *xi)
{
    constexpr int U = 4;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int q = gid / NRP, r = gid % NRP;
    if (q >= cnt) return;
    const int4 rc = rec[r0 + q];
    asm volatile("" ::"v"(rc.y), "v"(rc.z));   // one record load, not two around the branch
    if (rc.x < 0) return;   // padding
    const int row = rc.x;
    const double bv = bi[(size_t)row * NRP + r];
    const T d = invd[row];
    T acc = (T)0;
    for (int e = rc.y; e < rc.z; e += U) {
        int c[U];
        T w[U];
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // unconditional loads at a clamped index, masked entries then
            // use the row itself (not used)
            const bool in = e + u < rc.z;
            const int ee = in ? e + u : rc.z - 1;
            const int cl = lja[ee];
            const T wl = lval[ee];
            c[u] = in ? cl : row;
            w[u] = in ? wl : (T)0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) xv[u] = xi[(size_t)c[u] * NRP + r];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (e + u < rc.z) acc = fma(w[u], (T)xv[u], acc);
    }
    xi[(size_t)row * NRP + r] = (double)(((T)bv - acc) * d);
}

// ---------------------------------------------------------------- sync-free

// Sync-free form of the level schedule: one launch solves every row.  Rows
// are taken in level order (position q of the level-ordered records; every
// level starts at a multiple of 64 positions, so a wavefront never holds rows
// of two levels), so a row only waits for rows at smaller positions, whose
// workgroups were dispatched before its own and are running or done.
//
// The solution vector itself carries readiness: before the launch it is
// filled with kNotReady, a signalling-NaN pattern that arithmetic never
// produces (results of NaN operands are quiet NaNs), and each x value is
// published by one 64-bit store.  A row polls the x values it reads until
// none is kNotReady and computes with the values it polled.  Loads and stores
// are agent-scope atomics, performed at the agent's coherence point (the
// XCDs' L2 caches are not coherent with each other).  Waits are bounded:
// after kSpinLimit polls (about a second) a row gives up, raises the shared
// abort word and writes a quiet NaN, so a broken dispatch-order assumption
// shows as NaN results instead of a hang; once the abort word is set every
// waiting row gives up at once.  (The limit is a launch argument so that a
// test can force the give-up path with a tiny one.)

// Interleaved right-hand sides (v[i*NRP + r]), NRP lanes per row as in
// k_level_il and with its arithmetic (same bits); each lane polls its own
// right-hand side's values.
template <int NRP, class T>
__global__ __launch_bounds__(256) void k_sf_il(int cnt, const int4 *__restrict__ rec, const int *__restrict__ lja,
                                               const T *__restrict__ lval, const T *__restrict__ invd,
                                               const double *__restrict__ bi, double *xi, int spin_limit,
                                               int *abort_word, int *fail_word)
{
    constexpr int U = 4;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int q = gid / NRP, r = gid % NRP;
    if (q >= cnt) return;
    const int4 rc = rec[q];
    asm volatile("" ::"v"(rc.y), "v"(rc.z));   // one record load, not two around the branch
    if (rc.x < 0) return;   // padding
    const int row = rc.x;
    const double bv = bi[(size_t)row * NRP + r];
    const T d = invd[row];
    T acc = (T)0;
    bool ok = true;
    for (int e = rc.y; e < rc.z && ok; e += U) {
        size_t idx[U];
        bool in[U];
        T w[U];
        double xv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // unconditional loads at a clamped index; a masked entry polls
            // the last entry again and is not used
            in[u] = e + u < rc.z;
            const int ee = in[u] ? e + u : rc.z - 1;
            idx[u] = (size_t)lja[ee] * NRP + r;
            const T wl = lval[ee];
            w[u] = in[u] ? wl : (T)0;
        }
        ok = poll_values<U>(xi, idx, in, xv, spin_limit, abort_word, fail_word);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u]) acc = fma(w[u], (T)xv[u], acc);
    }
    store_agent(&xi[(size_t)row * NRP + r], ok ? (double)(((T)bv - acc) * d) : __builtin_nan(""));
}

// FP32 level solve with one FP64 refinement step in one sync-free launch, on
// interleaved right-hand sides (NRP lanes per row as in k_sf_il).  Per row
// and right-hand side, in level order:
//   x0 = (b - sum_e l_e x0[c_e]) / d      FP32 values (k_sf_il<NRP, float>)
//   r  = b - sum_e a_e x0[c_e] - a_d x0    FP64, CSR order with the diagonal
//                                          last (k_residual_il)
//   dx = (r - sum_e l_e dx[c_e]) / d       FP32 values (k_sf_il<NRP, float>)
// with the same operations in the same order as those three launches, so
// x = x0 + dx has the same bits as the launch sequence.  Both solves walk the
// same levels, so one wait per operand serves both: x0 and dx are FP32
// numbers and are published together as one 64-bit word, x0 in the high
// half and dx in the low half (a poll sees both or neither).  No published
// word equals kNotReady: its high half is a NaN and its low half a finite
// number, while a NaN x0 makes r, and so dx, NaN.  rec.w is the row's first
// entry in the lower-form CSR (a64, diagonal last), whose strictly lower
// entries are the row's level-ordered entries in the same order.
__device__ __forceinline__ double pack_pair(float hi, float lo)
{
    return __longlong_as_double((long long)(((unsigned long long)__float_as_uint(hi) << 32) | __float_as_uint(lo)));
}
__device__ __forceinline__ float pair_hi(double v)
{
    return __uint_as_float((unsigned)((unsigned long long)__double_as_longlong(v) >> 32));
}
__device__ __forceinline__ float pair_lo(double v)
{
    return __uint_as_float((unsigned)(unsigned long long)__double_as_longlong(v));
}

template <int NRP>
__global__ __launch_bounds__(256) void k_sf_il_ref1(int cnt, const int4 *__restrict__ rec, const int *__restrict__ lja,
                                                    const float *__restrict__ lval, const float *__restrict__ invd,
                                                    const double *__restrict__ a64, const double *__restrict__ bi,
                                                    double *pub, int spin_limit, int *abort_word, int *fail_word)
{
    constexpr int U = 4;
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int q = gid / NRP, r = gid % NRP;
    if (q >= cnt) return;
    const int4 rc = rec[q];
    asm volatile("" ::"v"(rc.y), "v"(rc.z), "v"(rc.w));   // one record load, not two around the branch
    if (rc.x < 0) return;   // padding
    const int row = rc.x;
    const double bv = bi[(size_t)row * NRP + r];
    const float d = invd[row];
    const int abase = rc.w - rc.y;   // a64 index of level-ordered entry e: abase + e
    const double ad = a64[abase + rc.z];   // the diagonal, after the strictly lower entries
    float acc0 = 0.f, acc1 = 0.f;
    double acc64 = 0.0;
    bool ok = true;
    for (int e = rc.y; e < rc.z && ok; e += U) {
        size_t idx[U];
        bool in[U];
        float w[U];
        double w64[U], pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            // unconditional loads at a clamped index; a masked entry polls
            // the last entry again and is not used
            in[u] = e + u < rc.z;
            const int ee = in[u] ? e + u : rc.z - 1;
            idx[u] = (size_t)lja[ee] * NRP + r;
            const float wl = lval[ee];
            const double wl64 = a64[abase + ee];
            w[u] = in[u] ? wl : 0.f;
            w64[u] = in[u] ? wl64 : 0.0;
        }
        ok = poll_values<U>(pub, idx, in, pv, spin_limit, abort_word, fail_word);
#pragma unroll
        for (int u = 0; u < U; u++)
            if (in[u]) {
                const float x0c = pair_hi(pv[u]);
                acc0 = fma(w[u], x0c, acc0);
                acc64 = fma(w64[u], (double)x0c, acc64);
                acc1 = fma(w[u], pair_lo(pv[u]), acc1);
            }
    }
    const float x0 = ((float)bv - acc0) * d;
    const double res = bv - fma(ad, (double)x0, acc64);
    const float dx = ((float)res - acc1) * d;
    const float qnan = __builtin_nanf("");
    store_agent(&pub[(size_t)row * NRP + r], ok ? pack_pair(x0, dx) : pack_pair(qnan, qnan));
}

// v[0 .. len) = kNotReady; clears the abort word.
__global__ __launch_bounds__(256) void k_not_ready(size_t len, double *v, int *abort_word)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < len) v[i] = __longlong_as_double((long long)kNotReady);
    if (i == 0) *abort_word = 0;
}

// Interleave / de-interleave between the caller's column-major layout
// (caller index flip ? n-1-i : i, leading dimension ld, nr columns) and
// v[i*nrp + r] over lower-form rows i (nrp = 1 << lgp <= 32).  A workgroup
// moves one tile of 2048 values, TR = 2048 / nrp rows by nrp columns, through
// LDS, so both the column side (TR contiguous rows per column) and the
// interleaved side (one contiguous stretch) are accessed in whole lines.  The
// LDS tile is column-major with one pad value per column.
constexpr int kTile = 2048;

// Columns r >= nr are filled with 0.
__global__ __launch_bounds__(256) void k_interleave(int n, int nr, int lgp, int flip, const double *__restrict__ src,
                                                    int ld, double *__restrict__ dst)
{
    __shared__ double tile[kTile + 32];
    const int nrp = 1 << lgp, lgr = 11 - lgp, TR = 1 << lgr;
    const int i0 = blockIdx.x << lgr;
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int f = threadIdx.x + 256 * u, r = f >> lgr, i = i0 + (f & (TR - 1));
        v[u] = r < nr && i < n ? src[(size_t)r * ld + (flip ? n - 1 - i : i)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int f = threadIdx.x + 256 * u;
        tile[(f >> lgr) * (TR + 1) + (f & (TR - 1))] = v[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int g = threadIdx.x + 256 * u, row = g >> lgp;
        if (i0 + row < n) dst[((size_t)i0 << lgp) + g] = tile[(g & (nrp - 1)) * (TR + 1) + row];
    }
}
// x (caller layout) = v, or x += v with accum, for the first nr columns;
// with v2, v + v2 takes the place of v; with pairs, v holds k_sf_il_ref1's
// packed (x0, dx) words and x0 + dx (in FP64) takes the place of v.
__global__ __launch_bounds__(256) void k_deinterleave(int n, int nr, int lgp, int flip, const double *__restrict__ v,
                                                      const double *__restrict__ v2, double *x, int ld, int accum,
                                                      int pairs)
{
    __shared__ double tile[kTile + 32];
    const int nrp = 1 << lgp, lgr = 11 - lgp, TR = 1 << lgr;
    const int i0 = blockIdx.x << lgr;
    double w[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int g = threadIdx.x + 256 * u;
        const size_t idx = ((size_t)i0 << lgp) + g;
        const bool ok = i0 + (g >> lgp) < n;
        w[u] = ok ? v[idx] : 0.0;
        if (pairs) w[u] = (double)pair_hi(w[u]) + (double)pair_lo(w[u]);
        if (v2 != nullptr && ok) w[u] += v2[idx];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int g = threadIdx.x + 256 * u;
        tile[(g & (nrp - 1)) * (TR + 1) + (g >> lgp)] = w[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int f = threadIdx.x + 256 * u, r = f >> lgr, row = f & (TR - 1), i = i0 + row;
        if (r < nr && i < n) {
            double *d = x + (size_t)r * ld + (flip ? n - 1 - i : i);
            const double t = tile[r * (TR + 1) + row];
            *d = accum ? *d + t : t;
        }
    }
}

// r[i*nrp + t] = b[i*nrp + t] - sum_e a_e x[ja_e*nrp + t] over row i of the
// lower-form CSR (diagonal included) in FP64, on interleaved vectors: one
// lane per row and right-hand side, so the x values of one entry are one
// contiguous segment.  Products are summed in entry order.
__global__ __launch_bounds__(256) void k_residual_il(int n, int lgp, const int *__restrict__ ia,
                                                     const int *__restrict__ ja, const double *__restrict__ a,
                                                     const double *__restrict__ bi, const double *__restrict__ xi,
                                                     double *__restrict__ r)
{
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int i = (int)(g >> lgp), t = (int)(g & ((1u << lgp) - 1));
    if (i >= n) return;
    const double bv = bi[g];
    const int e0 = ia[i], e1 = ia[i + 1];
    double acc = 0.0;
    for (int e = e0; e < e1; e++) acc = fma(a[e], xi[((size_t)ja[e] << lgp) + t], acc);
    r[g] = bv - acc;
}

// x[:, t] += d[:, t] for t < nr (refinement update of the level method).
__global__ __launch_bounds__(256) void k_axpy(int n, int nr, const double *__restrict__ d, int ldd, double *x, int ldx)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int t = 0; t < nr; t++) x[(size_t)t * ldx + i] += d[(size_t)t * ldd + i];
}

// One column per grid row, 1024 entries per workgroup: all four loads of a
// lane are issued before its stores.
__global__ __launch_bounds__(256) void k_copy_cols(int n, ColCopy cc)
{
    const double *__restrict__ src = cc.src[blockIdx.y];
    double *__restrict__ dst = cc.dst[blockIdx.y];
    const int i0 = blockIdx.x * 1024 + threadIdx.x;
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = i0 + u * 256;
        v[u] = i < n ? src[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = i0 + u * 256;
        if (i < n) dst[i] = v[u];
    }
}

// ---------------------------------------------------------------- spikes
//
// Spike rows r = 0 .. rows-1 of a partition are its rows f .. g (row r is
// matrix row base + r, base = s + f).  Column k of the spike (K_p's k-th
// coupling column) satisfies, row by row,
//     w_r = (c_r - sum_{j in-block} l_rj w_(j-base)) / l_rr,
// c_r the row's entry in that coupling column (or 0), with the row's
// off-diagonal entries taken in entry order.  Every kernel below runs one
// thread per spike column and the rows in order.

// Staged entries of the ring build.  Spike row r of a partition occupies
// [pia[r], pia[r+1]) of (pcode, pval): first its reciprocal diagonal, then
// its rinfo.z coupling entries (pcode = -1 - k for K_p's k-th column), zero
// padding to a multiple of four, then its other off-diagonal entries in
// rinfo.w chunks of four (zero-padded).  Their pcode is the byte offset
// (x * kstride * 8) of the LDS row x that holds the column's value at the
// entry's column:
//   x = c & (R - 1)       spike row c = ja - base with r - c < R (ring);
//   x = R + 1 + fslot[c]  spike row c further back (its far row);
//   x = R                 a column of the partition before f, or padding
//                         (zero row, value 0).
// With gfar, a spike row c further back is read from W instead: pcode =
// -1 - wslot[c].  One wavefront per spike row; K_p is ascending.
__global__ __launch_bounds__(256) void k_spike_codes(const SpikePart *__restrict__ sps, const int *__restrict__ kcols,
                                                     const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const double *__restrict__ a, const int4 *__restrict__ rinfo,
                                                     const int *__restrict__ pia, int R, int kstride, int gfar,
                                                     int *pcode, double *pval)
{
    const SpikePart sp = sps[blockIdx.y];
    const int r = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
    if (r >= sp.rows) return;
    const int base = sp.s + sp.f, rb = kstride * (int)sizeof(double), zcode = R * rb;
    const int4 *ri = rinfo + sp.slot0;
    const int4 inf = ri[r];
    const int e0 = ia[base + r], e1 = ia[base + r + 1] - 1, ncp = inf.z, nin = e1 - e0 - ncp;
    const int q = pia[sp.slot0 + r], hdr = (1 + ncp + 3) / 4 * 4;
    for (int l = lane; l < hdr; l += 64) {
        int c = zcode;
        double v = 0.0;
        if (l == 0) {
            v = 1.0 / a[e1];
        } else if (l <= ncp) {
            const int j = ja[e0 + l - 1];
            const int *kc = kcols + sp.kbeg;
            int lo = 0, hi = sp.kcnt - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (kc[mid] < j) lo = mid + 1;
                else hi = mid;
            }
            c = -1 - lo;
            v = a[e0 + l - 1];
        }
        pcode[q + l] = c;
        pval[q + l] = v;
    }
    for (int m = lane; m < 4 * inf.w; m += 64) {
        int c = zcode;
        double v = 0.0;
        if (m < nin) {
            const int e = e0 + ncp + m, j = ja[e];
            if (j >= base) {
                const int cr = j - base;
                c = r - cr < R ? (cr & (R - 1)) * rb : gfar ? -1 - ri[cr].x : (R + 1 + ri[cr].y) * rb;
            }
            v = a[e];
        }
        pcode[q + hdr + m] = c;
        pval[q + hdr + m] = v;
    }
}

// One workgroup per (partition, tile of blockDim columns), one thread per
// column.  Each thread keeps its column's last R rows (ring), a zero row and
// the far rows (rows that a row R or more further on reads; the host assigns
// their far slots so that live ones never share a slot) in LDS: row x of
// column t at lrow[x * kstride + t].  A thread reads and writes only its own
// column there, so the rows need no barrier.  The rows' staged entries are
// the same for every thread: they are copied sb rows at a time into LDS,
// double-buffered, with the next stage's loads issued before the current
// stage's rows (row offsets two stages ahead), so a row reads only LDS; its
// row info is read one row ahead.  Per chunk of four entries a thread reads
// the codes and values (three 16-byte reads), then per entry spends an add,
// an LDS read and an FMA into one of four partial sums (entry position mod
// 4), which are added at the end of the row.  rinfo per spike row: (W slot or -1, far slot or
// -1, coupling entries, chunks).
constexpr int kRingPF = 4;   // staged entries per thread and stage

// Value of this thread's column at a chunk entry's code: an LDS row, or
// (GF, negative code) a row stored in W (a flat load covers both).
template <bool GF>
__device__ __forceinline__ double ring_value(const char *mine, const double *Wr, int kcnt, int code)
{
    if (!GF) return *reinterpret_cast<const double *>(mine + code);
    const double *p = code >= 0 ? reinterpret_cast<const double *>(mine + code) : Wr + (int64_t)(-1 - code) * kcnt;
    return *p;
}

template <bool GF>
__global__ __launch_bounds__(256) void k_spike_ring(const SpikePart *__restrict__ sps, const int *__restrict__ pia,
                                                    const int *__restrict__ pcode, const double *__restrict__ pval,
                                                    const int4 *__restrict__ rinfo, int R, int nlr, int kstride,
                                                    int sb, int cap, double *W)
{
    extern __shared__ double lds_sr[];
    const int bd = blockDim.x, tid = threadIdx.x;
    int4 *rib = reinterpret_cast<int4 *>(lds_sr);                       // 2 x sb row info
    double *evb = reinterpret_cast<double *>(rib + 2 * sb);             // 2 x cap values
    int *ecb = reinterpret_cast<int *>(evb + 2 * (size_t)cap);          // 2 x cap codes
    double *lrow = reinterpret_cast<double *>(ecb + 2 * (size_t)cap);   // nlr x kstride
    int *rpb = reinterpret_cast<int *>(lrow + (size_t)nlr * kstride);   // 3 x (sb + 1) row offsets
    const SpikePart sp = sps[blockIdx.x];
    const int col = blockIdx.y * bd + tid;
    const bool active = col < sp.kcnt;
    const int lt = min(tid, kstride - 1);
    const int rows = sp.rows;
    const int *iap = pia + sp.slot0;
    const int4 *rip = rinfo + sp.slot0;
    double *Wc = W + sp.woff + col;
    const double *Wr = W + sp.woff + min(col, sp.kcnt - 1);          // far reads (idle lanes: a real column)
    const char *mine = reinterpret_cast<const char *>(lrow + lt);   // + code: this thread's entry of an LDS row

    lrow[(size_t)R * kstride + lt] = 0.0;
    int prow = 0;
    if (tid <= sb) {
        rpb[tid] = iap[min(tid, rows)];
        prow = iap[min(sb + tid, rows)];
    }
    __syncthreads();
    int pc[kRingPF];
    double pv[kRingPF];
    int4 pinf = make_int4(-1, -1, 0, 0);
    {
        const int e0 = rpb[0], ne = rpb[sb] - e0;
#pragma unroll
        for (int u = 0; u < kRingPF; u++) {
            const int i = tid + u * bd;
            pc[u] = i < ne ? pcode[e0 + i] : 0;
            pv[u] = i < ne ? pval[e0 + i] : 0.0;
        }
        if (tid < sb && tid < rows) pinf = rip[tid];
    }
    for (int r0 = 0, k = 0; r0 < rows; r0 += sb, k++) {
        const int *rp = rpb + (k % 3) * (sb + 1);
        int *rpn = rpb + ((k + 1) % 3) * (sb + 1);
        const int *ec = ecb + (k & 1) * cap;
        const double *ev = evb + (k & 1) * cap;
        const int4 *rin = rib + (k & 1) * sb;
        {
            int *ecw = ecb + (k & 1) * cap;
            double *evw = evb + (k & 1) * cap;
            const int E0 = rp[0], ne = rp[sb] - E0;
#pragma unroll
            for (int u = 0; u < kRingPF; u++) {
                const int i = tid + u * bd;
                if (i < ne) {
                    ecw[i] = pc[u];
                    evw[i] = pv[u];
                }
            }
            if (tid < sb) rib[(k & 1) * sb + tid] = pinf;
            if (tid <= sb) rpn[tid] = prow;
        }
        __syncthreads();
        // the next stage's entries and row info, the offsets of the one after
        const int r1n = r0 + sb;
        if (r1n < rows) {
            const int e0n = __builtin_amdgcn_readfirstlane(rpn[0]);
            const int nen = __builtin_amdgcn_readfirstlane(rpn[sb]) - e0n;
#pragma unroll
            for (int u = 0; u < kRingPF; u++) {
                const int i = tid + u * bd;
                if (i < nen) {
                    pc[u] = pcode[e0n + i];
                    pv[u] = pval[e0n + i];
                }
            }
            if (tid < sb && r1n + tid < rows) pinf = rip[r1n + tid];
            if (tid <= sb) prow = iap[min(r1n + sb + tid, rows)];
        }
        const int r1 = min(r0 + sb, rows);
        int q = 0;   // the row's first staged entry
        int4 inf = rin[0];
        for (int r = r0; r < r1; r++) {
            const int4 ninf = rin[min(r + 1 - r0, sb - 1)];   // the next row's (unused after the last)
            const int ncp = __builtin_amdgcn_readfirstlane(inf.z), nch = __builtin_amdgcn_readfirstlane(inf.w);
            const int wsl = __builtin_amdgcn_readfirstlane(inf.x), fsl = __builtin_amdgcn_readfirstlane(inf.y);
            const double rinv = ev[q];
            double w = 0.0;
            const int hdr = (1 + ncp + 3) & ~3;
            if (ncp > 0) {
                // coupling entries: at most one per column, so order is free
                for (int h = 0; h < hdr; h += 4) {
                    const int4 cd = *reinterpret_cast<const int4 *>(ec + q + h);
                    const double2 v01 = *reinterpret_cast<const double2 *>(ev + q + h);
                    const double2 v23 = *reinterpret_cast<const double2 *>(ev + q + h + 2);
                    w = cd.x == -1 - col ? v01.x : w;
                    w = cd.y == -1 - col ? v01.y : w;
                    w = cd.z == -1 - col ? v23.x : w;
                    w = cd.w == -1 - col ? v23.y : w;
                }
            }
            q += hdr;
            // four chunks at a time (sixteen entries): their codes, then their
            // values in the ring, then the products, in four partial sums (one
            // per chunk position) added at the end
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int cb = 0;
            for (; cb + 4 <= nch; cb += 4, q += 16) {
                int4 cd[4];
                double2 va[4], vb[4];
                double x[4][4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    cd[c] = *reinterpret_cast<const int4 *>(ec + q + 4 * c);
                    va[c] = *reinterpret_cast<const double2 *>(ev + q + 4 * c);
                    vb[c] = *reinterpret_cast<const double2 *>(ev + q + 4 * c + 2);
                }
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    x[c][0] = ring_value<GF>(mine, Wr, sp.kcnt, cd[c].x);
                    x[c][1] = ring_value<GF>(mine, Wr, sp.kcnt, cd[c].y);
                    x[c][2] = ring_value<GF>(mine, Wr, sp.kcnt, cd[c].z);
                    x[c][3] = ring_value<GF>(mine, Wr, sp.kcnt, cd[c].w);
                }
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    s0 += va[c].x * x[c][0];
                    s1 += va[c].y * x[c][1];
                    s2 += vb[c].x * x[c][2];
                    s3 += vb[c].y * x[c][3];
                }
            }
            for (; cb < nch; cb++, q += 4) {
                const int4 cd = *reinterpret_cast<const int4 *>(ec + q);
                const double2 va = *reinterpret_cast<const double2 *>(ev + q);
                const double2 vb = *reinterpret_cast<const double2 *>(ev + q + 2);
                s0 += va.x * ring_value<GF>(mine, Wr, sp.kcnt, cd.x);
                s1 += va.y * ring_value<GF>(mine, Wr, sp.kcnt, cd.y);
                s2 += vb.x * ring_value<GF>(mine, Wr, sp.kcnt, cd.z);
                s3 += vb.y * ring_value<GF>(mine, Wr, sp.kcnt, cd.w);
            }
            w -= (s0 + s1) + (s2 + s3);
            w *= rinv;
            if (active) {
                lrow[(size_t)(r & (R - 1)) * kstride + lt] = w;
                if (fsl >= 0) lrow[(size_t)(R + 1 + fsl) * kstride + lt] = w;
                if (wsl >= 0) Wc[(int64_t)wsl * sp.kcnt] = w;
            }
            inf = ninf;
        }
    }
}

// Fallback (partitions whose ring does not fit LDS): the spike rows are all
// stored in W (row r at W[woff + r * kcnt]) and a row reads the earlier rows
// from there.  One thread per column: workgroup (p, y) takes columns
// [256 y, 256 y + 256) of partition p.
__global__ __launch_bounds__(256) void k_spike_direct(const SpikePart *__restrict__ sps, const int *__restrict__ kcols,
                                                      const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const double *__restrict__ a, double *W)
{
    const SpikePart sp = sps[blockIdx.x];
    const int base = sp.s + sp.f;
    const int j = (int)(blockIdx.y * blockDim.x + threadIdx.x);
    if (j >= sp.kcnt) return;
    const int myk = kcols[sp.kbeg + j];
    double *Wc = W + sp.woff + j;
    for (int r = 0; r < sp.rows; r++) {
        const int e0 = ia[base + r], e1 = ia[base + r + 1] - 1;
        double w = 0.0;
        int e = e0;
        for (; e < e1 && ja[e] < sp.s; e++)
            if (ja[e] == myk) w = a[e];
        for (; e < e1; e++) {
            const int c = ja[e] - base;
            if (c >= 0) w -= a[e] * Wc[(int64_t)c * sp.kcnt];
        }
        Wc[(int64_t)r * sp.kcnt] = w / a[e1];
    }
}

// ---------------------------------------------------------------- assembly

__device__ __forceinline__ bool keep_entry(double w, double tol)
{
    return tol < 0.0 ? (w != 0.0) : (fabs(w) > tol);
}

__global__ __launch_bounds__(256) void k_reduced_count(int nred, const int64_t *__restrict__ wrow,
                                                       const int *__restrict__ kcnt, const double *__restrict__ tol,
                                                       const double *__restrict__ W, int *cnt, double *dropped)
{
    const int q = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (q >= nred) return;
    const int64_t w0 = wrow[q];
    int c = 0;
    double d = 0.0;
    if (w0 >= 0) {
        const int m = kcnt[q];
        const double t = tol[q];
        for (int l = lane; l < m; l += 64) {
            const double w = W[w0 + l];
            if (keep_entry(w, t)) c++;
            else d += fabs(w);
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        c += __shfl_xor(c, off);
        d += __shfl_xor(d, off);
    }
    if (lane == 0) {
        cnt[q] = c;
        dropped[q] = d;
    }
}

__global__ __launch_bounds__(256) void k_reduced_fill(int nred, const int64_t *__restrict__ wrow,
                                                      const int *__restrict__ kbeg, const int *__restrict__ kcnt,
                                                      const double *__restrict__ tol, const double *__restrict__ W,
                                                      const int *__restrict__ kridx, const int *__restrict__ rp,
                                                      int *ci, double *v)
{
    const int q = blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (q >= nred) return;
    const int64_t w0 = wrow[q];
    if (w0 < 0) return;
    const int m = kcnt[q], kb = kbeg[q];
    const double t = tol[q];
    int pos = rp[q];
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int l0 = 0; l0 < m; l0 += 64) {
        const int l = l0 + lane;
        double w = 0.0;
        bool keep = false;
        if (l < m) {
            w = W[w0 + l];
            keep = keep_entry(w, t);
        }
        const unsigned long long mask = __ballot(keep);
        if (keep) {
            const int at = pos + __popcll(mask & below);
            ci[at] = kridx[kb + l];
            v[at] = w;
        }
        pos += __popcll(mask);
    }
}

// Dynamic LDS of the fused solve: the block solve's x segment and its pass-2
// coupling products, or a sweep item's scratch (products, row sums, row
// bounds), whichever is larger.
template <int NR, class T>
size_t fused_lds_t(const BlockArgs &a)
{
    const size_t blk = xs_bytes<T>(a.kmax, NR) + (size_t)a.cmax * NR * sizeof(double);
    const size_t swp = (size_t)(64 * 32 + kFusedRowsPerItem * NR) * sizeof(double) + (kFusedRowsPerItem + 1) * sizeof(int);
    return std::max(blk, swp);
}

template <int G, int D, int NR, class T>
hipError_t fused_gdt(int P, const BlockArgs &a, const FusedArgs &f, int ncu, hipStream_t s, int *per_cu_out)
{
    const size_t lds = fused_lds_t<NR, T>(a);
    auto kern = k_spike_fused<G, 1, D, NR, T>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds);
    if (e != hipSuccess) return e;
    if (per_cu_out) {   // query only
        *per_cu_out = per_cu;
        return hipSuccess;
    }
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const long tiles = a.nr_slots / NR;
    const long want = std::max((long)P * tiles, (long)f.nchunks * tiles);
    const long grid = std::max(1L, std::min((long)per_cu * ncu, want));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), lds, s, a, f, P);
    return hipGetLastError();
}

template <int G>
hipError_t fused_g(int P, const BlockArgs &a, const FusedArgs &f, int ncu, hipStream_t s, int *q)
{
    if (a.tile < 1 || a.tile > a.nr_slots || a.nr_slots % a.tile) return hipErrorInvalidValue;
#define SPIKE_FUSED_T(NR, D)                                                                            \
    return a.fp32 ? fused_gdt<G, D, NR, float>(P, a, f, ncu, s, q) : fused_gdt<G, D, NR, double>(P, a, f, ncu, s, q)
    switch (a.tile) {
    case 1: SPIKE_FUSED_T(1, 16);
    case 2: SPIKE_FUSED_T(2, 16);
    case 4: SPIKE_FUSED_T(4, 8);
    case 8: SPIKE_FUSED_T(8, 8);
    default: return hipErrorInvalidValue;
    }
#undef SPIKE_FUSED_T
}

hipError_t fused_dispatch(int P, const BlockArgs &a, const FusedArgs &f, int ncu, hipStream_t s, int *q)
{
    if (a.st.fwd) return hipErrorInvalidValue;
    switch (a.st.G) {
    case 1: return fused_g<1>(P, a, f, ncu, s, q);
    case 2: return fused_g<2>(P, a, f, ncu, s, q);
    case 4: return fused_g<4>(P, a, f, ncu, s, q);
    case 8: return fused_g<8>(P, a, f, ncu, s, q);
    case 16: return fused_g<16>(P, a, f, ncu, s, q);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_block_solve(int pass, int P, const BlockArgs &a, hipStream_t s)
{
    if (P <= 0) return hipSuccess;
    return pass == 1 ? launch_pass<1>(P, a, s) : launch_pass<2>(P, a, s);
}

size_t spike_fused_lds(const BlockArgs &a)
{
    if (a.st.fwd) return 0;
    int per_cu = 0;
    FusedArgs f{};
    if (fused_dispatch(1, a, f, 1, nullptr, &per_cu) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        return 0;
    }
    switch (a.tile) {
    case 1: return a.fp32 ? fused_lds_t<1, float>(a) : fused_lds_t<1, double>(a);
    case 2: return a.fp32 ? fused_lds_t<2, float>(a) : fused_lds_t<2, double>(a);
    case 4: return a.fp32 ? fused_lds_t<4, float>(a) : fused_lds_t<4, double>(a);
    case 8: return a.fp32 ? fused_lds_t<8, float>(a) : fused_lds_t<8, double>(a);
    default: return 0;
    }
}

hipError_t launch_spike_fused(int P, const BlockArgs &a, const FusedArgs &f, int ncu, hipStream_t s)
{
    if (P <= 0) return hipSuccess;
    return fused_dispatch(P, a, f, ncu, s, nullptr);
}

// The no-alloc entry point's scratch.  In host code the name is a stand-in,
// not a device address: only the address from hipGetSymbolAddress (one per
// device, each device holding its own copy) is passed to kernels.
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

hipError_t noalloc_scratch(double **ptr)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = static_cast<double *>(p);
    }
    *ptr = addr[dev];
    return hipSuccess;
}

hipError_t launch_fused_reset(double *X, size_t len, unsigned *ctr, int *abort_word, hipStream_t s)
{
    const size_t n = std::max<size_t>(len, 96);
    hipLaunchKernelGGL(k_fused_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, len, X, ctr, abort_word);
    return hipGetLastError();
}

template <bool IN_LDS, int NR>
hipError_t launch_reduced_nr(int nred, int nlev, const int *levptr, const int *rlist, const int *rp, const int *ci,
                             const double *v, double *xred, hipStream_t s)
{
    size_t lds = IN_LDS ? (size_t)nred * NR * sizeof(double) : 0;
    auto kern = k_reduced_solve<IN_LDS, NR>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(1), dim3(1024), lds, s, nred, nlev, levptr, rlist, rp, ci, v, xred);
    return hipGetLastError();
}

hipError_t launch_reduced_solve(int nred, int nlev, const int *levptr, const int *rlist, const int *rp,
                                const int *ci, const double *v, double *xred, bool in_lds, int nr_slots,
                                hipStream_t s)
{
    if (nlev <= 0) return hipSuccess;
#define SPIKE_RED(NR)                                                                                    \
    return in_lds ? launch_reduced_nr<true, NR>(nred, nlev, levptr, rlist, rp, ci, v, xred, s)           \
                  : launch_reduced_nr<false, NR>(nred, nlev, levptr, rlist, rp, ci, v, xred, s)
    switch (nr_slots) {
    case 1: SPIKE_RED(1);
    case 2: SPIKE_RED(2);
    case 4: SPIKE_RED(4);
    case 8: SPIKE_RED(8);
    default: return hipErrorInvalidValue;
    }
#undef SPIKE_RED
}

template <int G, int U, int NR>
static void jacobi_go(int nrows, const int4 *r4, const int *ci, const double *v, const double *y, const double *xin,
                      double *xout, hipStream_t s)
{
    const long threads = (long)nrows * G;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    hipLaunchKernelGGL((k_jacobi_sweep<G, U, NR>), grid, block, 0, s, nrows, r4, ci, v, y, xin, xout);
}

hipError_t launch_jacobi_sweep(int nrows, const int *rows4, const int *ci, const double *v, const double *y,
                               const double *xin, double *xout, int nr_slots, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    // one lane layout for every RHS count: a column gets the same bits
    // whatever else is solved with it
    const int4 *r4 = (const int4 *)rows4;
    switch (nr_slots) {
    case 1: jacobi_go<8, 4, 1>(nrows, r4, ci, v, y, xin, xout, s); break;
    case 2: jacobi_go<8, 4, 2>(nrows, r4, ci, v, y, xin, xout, s); break;
    case 4: jacobi_go<8, 4, 4>(nrows, r4, ci, v, y, xin, xout, s); break;
    case 8: jacobi_go<8, 4, 8>(nrows, r4, ci, v, y, xin, xout, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t jacobi_items_lds(int kmax, int nr_slots)
{
    return (size_t)std::max(kmax, 1) * (size_t)nr_slots * sizeof(double);
}

hipError_t launch_jacobi_items(const JacobiItems &j, int m, const int *list, int nitems, const double *xin,
                               double *xout, int nr_slots, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
    const size_t lds = jacobi_items_lds(j.kmax, nr_slots);
    if (lds > block_solve_lds_limit()) return hipErrorInvalidValue;
    const void *fn = nullptr;
    switch (nr_slots) {
    case 1: fn = (const void *)k_jacobi_items<1>; break;
    case 2: fn = (const void *)k_jacobi_items<2>; break;
    case 4: fn = (const void *)k_jacobi_items<4>; break;
    case 8: fn = (const void *)k_jacobi_items<8>; break;
    default: return hipErrorInvalidValue;
    }
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const dim3 grid((unsigned)nitems), block(256);
    switch (nr_slots) {
    case 1: hipLaunchKernelGGL(k_jacobi_items<1>, grid, block, lds, s, j, m, list, xin, xout); break;
    case 2: hipLaunchKernelGGL(k_jacobi_items<2>, grid, block, lds, s, j, m, list, xin, xout); break;
    case 4: hipLaunchKernelGGL(k_jacobi_items<4>, grid, block, lds, s, j, m, list, xin, xout); break;
    case 8: hipLaunchKernelGGL(k_jacobi_items<8>, grid, block, lds, s, j, m, list, xin, xout); break;
    }
    return hipGetLastError();
}

hipError_t launch_coupling(int P, const DevPart *parts, const int *cpl_rows, const int *cpl_ptr, const int *cpl_col,
                           const double *cpl_val, const double *xred, const BlockArgs &a, double *bc, hipStream_t s)
{
    if (P <= 0) return hipSuccess;
    const dim3 grid(P), block(256);
    switch (a.nr_slots) {
    case 1: hipLaunchKernelGGL(k_coupling<1>, grid, block, 0, s, parts, cpl_rows, cpl_ptr, cpl_col, cpl_val, xred, a, bc); break;
    case 2: hipLaunchKernelGGL(k_coupling<2>, grid, block, 0, s, parts, cpl_rows, cpl_ptr, cpl_col, cpl_val, xred, a, bc); break;
    case 4: hipLaunchKernelGGL(k_coupling<4>, grid, block, 0, s, parts, cpl_rows, cpl_ptr, cpl_col, cpl_val, xred, a, bc); break;
    case 8: hipLaunchKernelGGL(k_coupling<8>, grid, block, 0, s, parts, cpl_rows, cpl_ptr, cpl_col, cpl_val, xred, a, bc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_residual(int P, const DevPart *parts, const int *ia, const int *ja, const double *a_,
                           const BlockArgs &a, int nr_max, double *r, int ldr, hipStream_t s)
{
    if (P <= 0 || nr_max <= 0) return hipSuccess;
    const dim3 grid(P), block(256);
    for (int c0 = 0; c0 < nr_max; c0 += 8) {
        const int left = nr_max - c0;
        if (left == 1)
            hipLaunchKernelGGL(k_residual<1>, grid, block, 0, s, parts, ia, ja, a_, a, r, ldr, c0);
        else if (left <= 4)
            hipLaunchKernelGGL(k_residual<4>, grid, block, 0, s, parts, ia, ja, a_, a, r, ldr, c0);
        else
            hipLaunchKernelGGL(k_residual<8>, grid, block, 0, s, parts, ia, ja, a_, a, r, ldr, c0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int G, int NR, class T>
hipError_t launch_level_t(int r0, int cnt, const LevelArgs &a, hipStream_t s)
{
    const dim3 grid((unsigned)(((size_t)cnt * G + 255) / 256)), block(256);
    hipLaunchKernelGGL((k_level<G, NR, T>), grid, block, 0, s, r0, cnt, (const int4 *)a.rec, a.lja,
                       (const T *)(a.fp32 ? (const void *)a.lval32 : (const void *)a.lval),
                       (const T *)(a.fp32 ? (const void *)a.invd32 : (const void *)a.invd), a.b, a.ldb, a.x, a.ldx,
                       a.nr, a.n, a.flip);
    return hipGetLastError();
}

template <int G, int NR>
hipError_t launch_level_g(int r0, int cnt, const LevelArgs &a, hipStream_t s)
{
    return a.fp32 ? launch_level_t<G, NR, float>(r0, cnt, a, s) : launch_level_t<G, NR, double>(r0, cnt, a, s);
}

template <int G>
hipError_t launch_level_nr(int r0, int cnt, const LevelArgs &a, hipStream_t s)
{
    switch (a.nr_slots) {
    case 1: return launch_level_g<G, 1>(r0, cnt, a, s);
    case 2: return launch_level_g<G, 2>(r0, cnt, a, s);
    case 4: return launch_level_g<G, 4>(r0, cnt, a, s);
    case 8: return launch_level_g<G, 8>(r0, cnt, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_level(int r0, int cnt, const LevelArgs &a, hipStream_t s)
{
    if (cnt <= 0) return hipSuccess;
    switch (a.G) {
    case 1: return launch_level_nr<1>(r0, cnt, a, s);
    case 2: return launch_level_nr<2>(r0, cnt, a, s);
    case 4: return launch_level_nr<4>(r0, cnt, a, s);
    case 8: return launch_level_nr<8>(r0, cnt, a, s);
    case 16: return launch_level_nr<16>(r0, cnt, a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int G>
static hipError_t level_ref1_g(int r0, int cnt, const LevelArgs &a, double *dl, hipStream_t s)
{
    const dim3 grid((unsigned)(((size_t)cnt * G + 255) / 256)), block(256);
    hipLaunchKernelGGL((k_level_ref1<G>), grid, block, 0, s, r0, cnt, (const int4 *)a.rec, a.lja, a.lval32, a.invd32,
                       a.a64, a.b, a.x, dl, a.n, a.flip);
    return hipGetLastError();
}

hipError_t launch_level_ref1(int r0, int cnt, const LevelArgs &a, double *dl, hipStream_t s)
{
    if (!a.fp32 || a.a64 == nullptr || a.lval32 == nullptr || a.invd32 == nullptr || dl == nullptr)
        return hipErrorInvalidValue;
    if (cnt <= 0) return hipSuccess;
    switch (a.G) {
    case 1: return level_ref1_g<1>(r0, cnt, a, dl, s);
    case 2: return level_ref1_g<2>(r0, cnt, a, dl, s);
    case 4: return level_ref1_g<4>(r0, cnt, a, dl, s);
    case 8: return level_ref1_g<8>(r0, cnt, a, dl, s);
    case 16: return level_ref1_g<16>(r0, cnt, a, dl, s);
    default: return hipErrorInvalidValue;
    }
}

template <int NRP>
hipError_t launch_level_il_n(int r0, int cnt, const LevelArgs &a, const double *bi, double *xi, hipStream_t s)
{
    const dim3 grid((unsigned)(((size_t)cnt * NRP + 255) / 256)), block(256);
    if (a.fp32)
        hipLaunchKernelGGL((k_level_il<NRP, float>), grid, block, 0, s, r0, cnt, (const int4 *)a.rec, a.lja,
                           a.lval32, a.invd32, bi, xi);
    else
        hipLaunchKernelGGL((k_level_il<NRP, double>), grid, block, 0, s, r0, cnt, (const int4 *)a.rec, a.lja, a.lval,
                           a.invd, bi, xi);
    return hipGetLastError();
}

hipError_t launch_level_il(int r0, int cnt, int nrp, const LevelArgs &a, const double *bi, double *xi, hipStream_t s)
{
    if (cnt <= 0) return hipSuccess;
    switch (nrp) {
    case 2: return launch_level_il_n<2>(r0, cnt, a, bi, xi, s);
    case 4: return launch_level_il_n<4>(r0, cnt, a, bi, xi, s);
    case 8: return launch_level_il_n<8>(r0, cnt, a, bi, xi, s);
    case 16: return launch_level_il_n<16>(r0, cnt, a, bi, xi, s);
    case 32: return launch_level_il_n<32>(r0, cnt, a, bi, xi, s);
    default: return hipErrorInvalidValue;
    }
}

template <int NRP>
static hipError_t sf_il_n(int npos, const LevelArgs &a, const double *bi, double *xi, int *abort_word, hipStream_t s)
{
    const dim3 grid((unsigned)(((size_t)npos * NRP + 255) / 256)), block(256);
    if (a.fp32)
        hipLaunchKernelGGL((k_sf_il<NRP, float>), grid, block, 0, s, npos, (const int4 *)a.rec, a.lja, a.lval32,
                           a.invd32, bi, xi, a.spin_limit, abort_word, a.fail_word);
    else
        hipLaunchKernelGGL((k_sf_il<NRP, double>), grid, block, 0, s, npos, (const int4 *)a.rec, a.lja, a.lval,
                           a.invd, bi, xi, a.spin_limit, abort_word, a.fail_word);
    return hipGetLastError();
}

hipError_t launch_sync_free(int npos, int n, int nrp, const LevelArgs &a, const double *bi, double *xi,
                            int *abort_word, hipStream_t s)
{
    if (npos <= 0 || n <= 0) return hipSuccess;
    // the vector the rows publish into starts as kNotReady
    const size_t len = (size_t)n * nrp;
    hipLaunchKernelGGL(k_not_ready, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, xi, abort_word);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (nrp) {
    case 2: return sf_il_n<2>(npos, a, bi, xi, abort_word, s);
    case 4: return sf_il_n<4>(npos, a, bi, xi, abort_word, s);
    case 8: return sf_il_n<8>(npos, a, bi, xi, abort_word, s);
    case 16: return sf_il_n<16>(npos, a, bi, xi, abort_word, s);
    case 32: return sf_il_n<32>(npos, a, bi, xi, abort_word, s);
    default: return hipErrorInvalidValue;
    }
}

template <int NRP>
static hipError_t sf_il_ref1_n(int npos, const LevelArgs &a, const double *bi, double *pub, int *abort_word,
                               hipStream_t s)
{
    const dim3 grid((unsigned)(((size_t)npos * NRP + 255) / 256)), block(256);
    hipLaunchKernelGGL((k_sf_il_ref1<NRP>), grid, block, 0, s, npos, (const int4 *)a.rec, a.lja, a.lval32, a.invd32,
                       a.a64, bi, pub, a.spin_limit, abort_word, a.fail_word);
    return hipGetLastError();
}

hipError_t launch_sync_free_refined1(int npos, int n, int nrp, const LevelArgs &a, const double *bi, double *pub,
                                     int *abort_word, hipStream_t s)
{
    if (!a.fp32 || a.a64 == nullptr || a.lval32 == nullptr || a.invd32 == nullptr) return hipErrorInvalidValue;
    if (npos <= 0 || n <= 0) return hipSuccess;
    const size_t len = (size_t)n * nrp;
    hipLaunchKernelGGL(k_not_ready, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, len, pub, abort_word);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (nrp) {
    case 2: return sf_il_ref1_n<2>(npos, a, bi, pub, abort_word, s);
    case 4: return sf_il_ref1_n<4>(npos, a, bi, pub, abort_word, s);
    case 8: return sf_il_ref1_n<8>(npos, a, bi, pub, abort_word, s);
    case 16: return sf_il_ref1_n<16>(npos, a, bi, pub, abort_word, s);
    case 32: return sf_il_ref1_n<32>(npos, a, bi, pub, abort_word, s);
    default: return hipErrorInvalidValue;
    }
}

namespace {
// log2 of an interleave width (a power of two <= 32), or -1.
int il_log2(int nrp)
{
    for (int l = 0; l <= 5; l++)
        if (nrp == 1 << l) return l;
    return -1;
}
}  // namespace

hipError_t launch_interleave(int n, int nr, int nrp, int flip, const double *src, int ld, double *dst, hipStream_t s)
{
    const int lgp = il_log2(nrp);
    if (lgp < 0 || nr > nrp) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    const int tr = kTile >> lgp;
    hipLaunchKernelGGL(k_interleave, dim3((unsigned)((n + tr - 1) / tr)), dim3(256), 0, s, n, nr, lgp, flip, src, ld,
                       dst);
    return hipGetLastError();
}

hipError_t launch_deinterleave(int n, int nr, int nrp, int flip, const double *v, double *x, int ld, int accum,
                               hipStream_t s)
{
    const int lgp = il_log2(nrp);
    if (lgp < 0 || nr > nrp) return hipErrorInvalidValue;
    if (n <= 0 || nr <= 0) return hipSuccess;
    const int tr = kTile >> lgp;
    hipLaunchKernelGGL(k_deinterleave, dim3((unsigned)((n + tr - 1) / tr)), dim3(256), 0, s, n, nr, lgp, flip, v,
                       (const double *)nullptr, x, ld, accum, 0);
    return hipGetLastError();
}

hipError_t launch_deinterleave_sum(int n, int nr, int nrp, int flip, const double *v, const double *v2, double *x,
                                   int ld, hipStream_t s)
{
    const int lgp = il_log2(nrp);
    if (lgp < 0 || nr > nrp || v2 == nullptr) return hipErrorInvalidValue;
    if (n <= 0 || nr <= 0) return hipSuccess;
    const int tr = kTile >> lgp;
    hipLaunchKernelGGL(k_deinterleave, dim3((unsigned)((n + tr - 1) / tr)), dim3(256), 0, s, n, nr, lgp, flip, v, v2,
                       x, ld, 0, 0);
    return hipGetLastError();
}

hipError_t launch_deinterleave_pairs(int n, int nr, int nrp, int flip, const double *v, double *x, int ld,
                                     hipStream_t s)
{
    const int lgp = il_log2(nrp);
    if (lgp < 0 || nr > nrp) return hipErrorInvalidValue;
    if (n <= 0 || nr <= 0) return hipSuccess;
    const int tr = kTile >> lgp;
    hipLaunchKernelGGL(k_deinterleave, dim3((unsigned)((n + tr - 1) / tr)), dim3(256), 0, s, n, nr, lgp, flip, v,
                       (const double *)nullptr, x, ld, 0, 1);
    return hipGetLastError();
}

hipError_t launch_residual_il(int n, int nrp, const int *ia, const int *ja, const double *a, const double *bi,
                              const double *xi, double *r, hipStream_t s)
{
    const int lgp = il_log2(nrp);
    if (lgp < 0) return hipErrorInvalidValue;
    const size_t tot = (size_t)n << lgp;
    if (tot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_residual_il, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, lgp, ia, ja, a, bi, xi,
                       r);
    return hipGetLastError();
}

hipError_t launch_axpy(int n, int nr, const double *d, int ldd, double *x, int ldx, hipStream_t s)
{
    if (n <= 0 || nr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, s, n, nr, d, ldd, x, ldx);
    return hipGetLastError();
}

hipError_t launch_copy_cols(int n, int ncols, const ColCopy &cc, hipStream_t s)
{
    if (n <= 0 || ncols <= 0) return hipSuccess;
    if (ncols > kMaxCopyCols) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy_cols, dim3((unsigned)((n + 1023) / 1024), (unsigned)ncols), dim3(256), 0, s, n, cc);
    return hipGetLastError();
}

size_t spike_ring_lds(int nlr, int kstride, int sb, int cap)
{
    return 2 * (size_t)sb * 4 * sizeof(int) + 2 * (size_t)cap * (sizeof(double) + sizeof(int)) +
           (size_t)nlr * kstride * sizeof(double) + 3 * (size_t)(sb + 1) * sizeof(int);
}

int spike_ring_stage_entries(int block_threads) { return kRingPF * block_threads; }

hipError_t launch_spike_codes(int nsp, int max_rows, const SpikePart *sp, const int *kcols, const int *ia,
                              const int *ja, const double *a, const int *rinfo, const int *pia, int R, int kstride,
                              bool gfar, int *pcode, double *pval, hipStream_t s)
{
    if (nsp <= 0 || max_rows <= 0) return hipSuccess;
    if (R < 1 || (R & (R - 1)) || kstride < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_spike_codes, dim3((unsigned)((max_rows + 3) / 4), (unsigned)nsp), dim3(256), 0, s, sp, kcols,
                       ia, ja, a, reinterpret_cast<const int4 *>(rinfo), pia, R, kstride, gfar ? 1 : 0, pcode, pval);
    return hipGetLastError();
}

hipError_t launch_spike_ring(int nsp, int max_kcnt, int block_threads, const SpikePart *sp, const int *pia,
                             const int *pcode, const double *pval, const int *rinfo, int R, int nlr, int kstride,
                             int sb, int cap, bool gfar, double *W, hipStream_t s)
{
    if (nsp <= 0 || max_kcnt <= 0) return hipSuccess;
    if (block_threads < 64 || block_threads > 256 || block_threads % 64 || R < 1 || (R & (R - 1)) || nlr <= R ||
        kstride < 1 || kstride > block_threads || (max_kcnt > block_threads && kstride != block_threads) ||
        (max_kcnt <= block_threads && kstride < max_kcnt) || sb < 1 || sb >= block_threads || cap < 4 || cap % 4 ||
        cap > kRingPF * block_threads)
        return hipErrorInvalidValue;
    const size_t lds = spike_ring_lds(nlr, kstride, sb, cap);
    if (lds > block_solve_lds_limit()) return hipErrorInvalidValue;
    const void *fn = gfar ? (const void *)k_spike_ring<true> : (const void *)k_spike_ring<false>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const dim3 grid((unsigned)nsp, (unsigned)((max_kcnt + block_threads - 1) / block_threads));
    if (gfar)
        hipLaunchKernelGGL(k_spike_ring<true>, grid, dim3(block_threads), lds, s, sp, pia, pcode, pval,
                           reinterpret_cast<const int4 *>(rinfo), R, nlr, kstride, sb, cap, W);
    else
        hipLaunchKernelGGL(k_spike_ring<false>, grid, dim3(block_threads), lds, s, sp, pia, pcode, pval,
                           reinterpret_cast<const int4 *>(rinfo), R, nlr, kstride, sb, cap, W);
    return hipGetLastError();
}

hipError_t launch_spike_direct(int nsp, int max_kcnt, const SpikePart *sp, const int *kcols, const int *ia,
                               const int *ja, const double *a, double *W, hipStream_t s)
{
    if (nsp <= 0 || max_kcnt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spike_direct, dim3((unsigned)nsp, (unsigned)((max_kcnt + 255) / 256)), dim3(256), 0, s, sp,
                       kcols, ia, ja, a, W);
    return hipGetLastError();
}

hipError_t launch_reduced_count(int nred, const int64_t *wrow, const int *kcnt, const double *tol, const double *W,
                                int *cnt, double *dropped, hipStream_t s)
{
    if (nred <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduced_count, dim3((nred + 3) / 4), dim3(256), 0, s, nred, wrow, kcnt, tol, W, cnt,
                       dropped);
    return hipGetLastError();
}

hipError_t launch_reduced_fill(int nred, const int64_t *wrow, const int *kbeg, const int *kcnt, const double *tol,
                               const double *W, const int *kridx, const int *rp, int *ci, double *v, hipStream_t s)
{
    if (nred <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduced_fill, dim3((nred + 3) / 4), dim3(256), 0, s, nred, wrow, kbeg, kcnt, tol, W, kridx,
                       rp, ci, v);
    return hipGetLastError();
}

size_t block_solve_lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 65536;
    return (size_t)v;
}

}  // namespace spike
