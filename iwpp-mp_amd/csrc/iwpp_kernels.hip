// HIP kernels for IWPP grayscale morphological reconstruction by dilation on
// CDNA4 (gfx950, wave64).  See iwpp_api.cpp for the host driver.
//
// Update rule everywhere (reference: IWPP MP/src/MorphologicalReconstruction.cu:42-47
// and the sweep recurrence s = min(max(prev, old), mask) of
// IWPP MP DQ/src/MorphologicalReconstruction.cu:861):
//     J(p) <- max(J(p), min(J(q), I(p)))   for a neighbour q of p,
// with pixels outside the image absent.  Along a line this is a scan with the
// associative pair operator  f_(v,c)(x) = max(v, min(x, c)),
//     (v1,c1) then (v2,c2)  =  (max(v2, min(v1, c2)), min(c1, c2)),
// identity (0, ~0).  Absent pixels are (0, 0): they neither rise nor conduct.
#include "iwpp_internal.h"

#include <cstdlib>
#include <type_traits>

namespace iwpp {

#define DEV static __device__ __forceinline__

// Every pointer the kernels dereference outside LDS points to global memory.
// Saying so (address space 1) makes the compiler emit global_* rather than
// flat_* instructions for pointers it reads from descriptors; flat accesses
// also count on the LDS/scalar counter and so serialise the memory waits.
#define GAS __attribute__((address_space(1)))
template <class T>
DEV GAS T* gp(T* p) { return (GAS T*)p; }
template <class T>
DEV const GAS T* gp(const T* p) { return (const GAS T*)p; }
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // 16-byte access

// ---------------------------------------------------------------- atomics ---
// Inside the queue kernels every access to shared mutable data -- marker
// words, queue entries, counters, flags, state words -- is a relaxed
// agent-scope atomic.  On gfx950 these are the device-coherent forms
// (global_load/store ... sc1, and RMWs performed at the device coherence
// point), so they are coherent across the 8 XCDs without L2 maintenance.
// Ordering between a wave's pixel stores and the atomics that hand a
// megapixel to another wave is enforced by waiting for the wave's outstanding
// memory operations (mem_wait: s_waitcnt vmcnt(0), with compiler barriers):
//   visit: take item (RMW) -> mem_wait -> load -> ... -> store -> mem_wait -> notify (RMW).
// An agent-scope release/acquire fence would do the same job but, on
// gfx950, writes back / invalidates the whole XCD L2 each time; measured, that
// was the dominant cost of a visit.  The mask is read-only and uses plain loads.
DEV uint32_t a_load(const uint32_t* p) {
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV void a_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_or(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_and(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_and(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_xchg(uint32_t* p, uint32_t v) {
    return __hip_atomic_exchange(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_cas(uint32_t* p, uint32_t expected, uint32_t desired) {
    __hip_atomic_compare_exchange_strong(gp(p), &expected, desired, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return expected;  // the value seen before the exchange
}
DEV unsigned long long a_load64(const unsigned long long* p) {
    return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV void a_store64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV void mem_wait() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0) expcnt(0) lgkmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Device-coherent 32-bit accesses to the marker and the queues.
DEV uint32_t ld_coh(const void* p) {
    return __hip_atomic_load(gp(reinterpret_cast<const uint32_t*>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
DEV void st_coh(void* p, uint32_t v) {
    __hip_atomic_store(gp(reinterpret_cast<uint32_t*>(p)), v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

DEV uint32_t* ctl_at(uint32_t* ctl, int slot) { return ctl + slot * kCtlStride; }

DEV int lane_id() { return (int)__lane_id(); }
DEV uint32_t rdlane(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
DEV uint32_t first_lane(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
DEV uint64_t ballot(bool p) { return __ballot(p); }
DEV bool wave_any(bool p) { return ballot(p) != 0; }

// ------------------------------------------------------------ DPP moves -----
// update_dpp(old, src): lanes whose source is out of range or whose row is
// masked off keep `old`.
template <int CTRL, int ROWMASK = 0xF>
DEV uint32_t dpp(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, ROWMASK, 0xF, false);
}
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_SHL1 = 0x101, DPP_ROW_SHL2 = 0x102, DPP_ROW_SHL4 = 0x104,
              DPP_ROW_SHL8 = 0x108;
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// The same with 0 for lanes whose source is out of range (bound_ctrl): a move
// the compiler can fold into the VOP2 instruction that consumes it.
template <int CTRL>
DEV uint32_t dpp_z(uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)src, CTRL, 0xF, 0xF, true);
}
// lane i receives v[i-1]; lane 0 receives fill.
DEV uint32_t from_left(uint32_t v, uint32_t fill) { return dpp<DPP_WAVE_SHR1>(fill, v); }
// lane i receives v[i+1]; lane 63 receives fill.
DEV uint32_t from_right(uint32_t v, uint32_t fill) { return dpp<DPP_WAVE_SHL1>(fill, v); }

// One combine step: (e) earlier then own.
DEV void pq_combine(uint32_t& P, uint32_t& Q, uint32_t pe, uint32_t qe) {
    P = max(P, min(pe, Q));
    Q = min(Q, qe);
}
template <int CTRL, int ROWMASK = 0xF>
DEV void scan_step(uint32_t& P, uint32_t& Q) {
    const uint32_t pe = dpp<CTRL, ROWMASK>(0u, P);
    const uint32_t qe = dpp<CTRL, ROWMASK>(0xFFFFFFFFu, Q);
    pq_combine(P, Q, pe, qe);
}
// Inclusive scan over lanes 0..63 in increasing lane order.
DEV void scan_lanes_fwd(uint32_t& P, uint32_t& Q) {
    scan_step<DPP_ROW_SHR1>(P, Q);
    scan_step<DPP_ROW_SHR2>(P, Q);
    scan_step<DPP_ROW_SHR4>(P, Q);
    scan_step<DPP_ROW_SHR8>(P, Q);
    scan_step<DPP_ROW_BCAST15, 0xA>(P, Q);
    scan_step<DPP_ROW_BCAST31, 0xC>(P, Q);
}
// Inclusive scan over lanes 63..0 in decreasing lane order.
DEV void scan_lanes_bwd(uint32_t& P, uint32_t& Q, int lane) {
    scan_step<DPP_ROW_SHL1>(P, Q);
    scan_step<DPP_ROW_SHL2>(P, Q);
    scan_step<DPP_ROW_SHL4>(P, Q);
    scan_step<DPP_ROW_SHL8>(P, Q);
    // Each 16-lane row's first lane now holds that row's full suffix.
    const int row = lane >> 4;
    {
        const uint32_t p16 = rdlane(P, 16), q16 = rdlane(Q, 16);
        const uint32_t p48 = rdlane(P, 48), q48 = rdlane(Q, 48);
        if (row == 0) pq_combine(P, Q, p16, q16);
        if (row == 2) pq_combine(P, Q, p48, q48);
    }
    {
        const uint32_t p32 = rdlane(P, 32), q32 = rdlane(Q, 32);
        if (row < 2) pq_combine(P, Q, p32, q32);
    }
}

// Carry into wave wv of a workgroup from the NWV per-wave summaries of one
// chunk in LDS (sp, sq), in scan direction D: xin enters wave wv, xall
// leaves the whole chunk; carry enters the chunk.  Up to 4 waves fold the
// summaries one after another; more waves scan them across the lanes of one
// 16-lane DPP row (NWV <= 16).
template <int NWV, int D>
DEV void wave_carry(const uint32_t* sp, const uint32_t* sq, int wv, int lane, uint32_t carry,
                    uint32_t& xin, uint32_t& xall) {
    static_assert(NWV <= 16, "one DPP row");
    if constexpr (NWV <= 4) {
        xin = carry;
        xall = carry;
#pragma unroll
        for (int uu = 0; uu < NWV; ++uu) {
            const int u = D > 0 ? uu : NWV - 1 - uu;
            const uint32_t tp = sp[u], tq = sq[u];
            if (D > 0 ? (u < wv) : (u > wv)) xin = max(tp, min(xin, tq));
            xall = max(tp, min(xall, tq));
        }
    } else {
        uint32_t P = 0, Q = 0xFFFFFFFFu;
        if (lane < NWV) {  // lane l: the l-th wave in scan order
            const int src = D > 0 ? lane : NWV - 1 - lane;
            P = sp[src];
            Q = sq[src];
        }
        scan_step<DPP_ROW_SHR1>(P, Q);
        scan_step<DPP_ROW_SHR2>(P, Q);
        scan_step<DPP_ROW_SHR4>(P, Q);
        scan_step<DPP_ROW_SHR8>(P, Q);
        const int w = __builtin_amdgcn_readfirstlane(wv);
        const int pos = D > 0 ? w : NWV - 1 - w;  // this wave's place in scan order
        uint32_t pp = 0, pq = 0xFFFFFFFFu;
        if (pos > 0) {
            pp = rdlane(P, pos - 1);
            pq = rdlane(Q, pos - 1);
        }
        xin = max(pp, min(carry, pq));
        xall = max(rdlane(P, NWV - 1), min(carry, rdlane(Q, NWV - 1)));
    }
}

// ------------------------------------------------------ vector row access ---
template <class T, int VB>
struct VecIO {
    static constexpr int NW = VB / 4;                 // 32-bit words per lane
    static constexpr int NPX = VB / (int)sizeof(T);   // pixels per lane
    static constexpr int PPW = 4 / (int)sizeof(T);    // pixels per word
    static constexpr int BITS = 8 * (int)sizeof(T);
    static constexpr uint32_t PM = (uint32_t)((1u << BITS) - 1u);

    // Every lane loads unconditionally (lanes past the row end re-read the
    // row's first vector) and clears the bytes beyond the row afterwards.
    // With no divergent branch around the load, a wave's loads of successive
    // rows stay in flight together; behind a branch the compiler waits for
    // each one at the join.  A vector never reaches past round_up(rowbytes,
    // VB), which the pitch covers (VB = 16 is used only for 16-byte pitches).
    DEV void load(const uint8_t* row, int o, int rowbytes, uint32_t (&w)[NW]) {
        load_raw(row, o, rowbytes, w);
        uint32_t k[NW];
        keep(o, rowbytes, k);
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] &= k[i];
    }
    // The load without the masking (bytes past the row end are garbage).
    DEV void load_raw(const uint8_t* row, int o, int rowbytes, uint32_t (&w)[NW]) {
        const int oc = o < rowbytes ? o : 0;
        if constexpr (VB == 16) {
            const u32x4 t = *gp(reinterpret_cast<const u32x4*>(row + oc));
            w[0] = t.x; w[1] = t.y; w[2] = t.z; w[3] = t.w;
        } else {
            w[0] = *gp(reinterpret_cast<const uint32_t*>(row + oc));
        }
    }
    // Byte masks of a lane's words: the bytes inside the row.
    DEV void keep(int o, int rowbytes, uint32_t (&k)[NW]) {
        const int valid = rowbytes - o;  // bytes of this vector inside the row
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int vb = valid - 4 * i;
            k[i] = vb >= 4 ? 0xFFFFFFFFu : vb <= 0 ? 0u : (0xFFFFFFFFu >> (8 * (4 - vb)));
        }
    }
    DEV void store(uint8_t* row, int o, int rowbytes, const uint32_t (&w)[NW]) {
        if (o + VB <= rowbytes) {
            if constexpr (VB == 16) {
                *gp(reinterpret_cast<u32x4*>(row + o)) = u32x4{w[0], w[1], w[2], w[3]};
            } else {
                *gp(reinterpret_cast<uint32_t*>(row + o)) = w[0];
            }
        } else if (o < rowbytes) {
#pragma unroll
            for (int k = 0; k < NPX; ++k) {
                const int ob = o + k * (int)sizeof(T);
                if (ob < rowbytes)
                    *gp(reinterpret_cast<T*>(row + ob)) = (T)((w[k / PPW] >> (BITS * (k % PPW))) & PM);
            }
        }
    }
    DEV void unpack(const uint32_t (&w)[NW], uint32_t (&px)[NPX]) {
#pragma unroll
        for (int k = 0; k < NPX; ++k) px[k] = (w[k / PPW] >> (BITS * (k % PPW))) & PM;
    }
    DEV void pack(const uint32_t (&px)[NPX], uint32_t (&w)[NW]) {
#pragma unroll
        for (int i = 0; i < NW; ++i) w[i] = 0;
#pragma unroll
        for (int k = 0; k < NPX; ++k) w[k / PPW] |= px[k] << (BITS * (k % PPW));
    }
};

// ------------------------------------------------- lane pixels for scans ---
// A lane's marker (v) and mask (c) pixels for the row scans.  With 16-byte
// lanes they are held as 16-bit pairs, pixel k in the low half and pixel
// k + NP in the high half, so that one packed min/max (v_pk_min_u16 /
// v_pk_max_u16) advances both halves of the lane's segment; the halves are
// joined through their (P, Q) summaries.  4-byte u8 lanes pair bytes 0 | 2 and
// 1 | 3 the same way; 4-byte u16 lanes hold one pixel per register.  The scan
// step is out = max(v, min(x, c)); a segment's summary (P, Q) maps an incoming
// x to max(P, min(x, Q)).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
DEV uint32_t pk_max(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
DEV uint32_t pk_min(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
// v_perm_b32: byte i of the result is byte sel[i] of {hi:lo} (0x0C gives 0).
DEV uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) { return __builtin_amdgcn_perm(hi, lo, sel); }

template <class T, int VB>
struct LanePx {
    using V = VecIO<T, VB>;
    static constexpr bool PAIRS = VB == 16 || (VB == 4 && sizeof(T) == 1);
    static constexpr int NW = V::NW;
    static constexpr int NP = PAIRS ? V::NPX / 2 : V::NPX;  // registers per plane
    uint32_t v[NP], c[NP];
    uint32_t hp = 0, hq = 0;  // summary of the half scanned first (pairs)

    static __device__ __forceinline__ void split(const uint32_t (&w)[NW], uint32_t (&p)[NP]) {
        if constexpr (!PAIRS) {
            V::unpack(w, p);
        } else if constexpr (VB == 4) {  // u8: p[k] = byte k | byte k+2 << 16
            p[0] = perm(0u, w[0], 0x0C020C00u);
            p[1] = perm(0u, w[0], 0x0C030C01u);
        } else if constexpr (sizeof(T) == 1) {  // p[k] = byte k | byte k+8 << 16
#pragma unroll
            for (int k = 0; k < NP; ++k)
                p[k] = perm(w[2 + k / 4], w[k / 4],
                            (uint32_t)(k % 4) | 0x0C00u | ((uint32_t)(4 + k % 4) << 16) | 0x0C000000u);
        } else {  // p[k] = pixel k | pixel k+4 << 16
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const uint32_t h = 2 * (k % 2);
                p[k] = perm(w[2 + k / 2], w[k / 2], h | (h + 1) << 8 | (h + 4) << 16 | (h + 5) << 24);
            }
        }
    }
    static __device__ __forceinline__ void join(const uint32_t (&p)[NP], uint32_t (&w)[NW]) {
        if constexpr (!PAIRS) {
            V::pack(p, w);
        } else if constexpr (VB == 4) {
            w[0] = perm(p[1], p[0], 0x06020400u);
        } else if constexpr (sizeof(T) == 1) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const uint32_t t01 = perm(p[4 * i + 1], p[4 * i], 0x06020400u);
                const uint32_t t23 = perm(p[4 * i + 3], p[4 * i + 2], 0x06020400u);
                w[i] = perm(t23, t01, 0x05040100u);
                w[2 + i] = perm(t23, t01, 0x07060302u);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                w[i] = perm(p[2 * i + 1], p[2 * i], 0x05040100u);
                w[2 + i] = perm(p[2 * i + 1], p[2 * i], 0x07060302u);
            }
        }
    }
    static __device__ __forceinline__ uint32_t vmax(uint32_t a, uint32_t b) { return PAIRS ? pk_max(a, b) : max(a, b); }
    static __device__ __forceinline__ uint32_t vmin(uint32_t a, uint32_t b) { return PAIRS ? pk_min(a, b) : min(a, b); }

    __device__ __forceinline__ void load(const uint32_t (&jw)[NW], const uint32_t (&iw)[NW]) {
        split(jw, v);
        split(iw, c);
    }
    __device__ __forceinline__ void store(uint32_t (&jw)[NW]) const { join(v, jw); }
    __device__ __forceinline__ bool clip() {  // whether any pixel changed
        bool d = false;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const uint32_t o = vmin(v[k], c[k]);
            d |= o != v[k];
            v[k] = o;
        }
        return d;
    }
    __device__ __forceinline__ void wall() {  // no rise, no conduction
#pragma unroll
        for (int k = 0; k < NP; ++k) c[k] = 0;
    }
    // The lane's mask minimum (per half with pairs): the Q of its summaries in
    // both directions, for reduce_masked.
    __device__ __forceinline__ uint32_t mask_min() const {
        uint32_t q = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < NP; ++k) q = vmin(q, c[k]);
        return q;
    }
    // Summary (P, Q) of the lane's pixels in scan direction D.
    template <int D>
    __device__ __forceinline__ void reduce(uint32_t& P, uint32_t& Q) {
        reduce_masked<D>(P, Q, mask_min());
    }
    // The same with the lane's mask minimum q (mask_min) given.
    template <int D>
    __device__ __forceinline__ void reduce_masked(uint32_t& P, uint32_t& Q, uint32_t q) {
        uint32_t p = 0;
#pragma unroll
        for (int kk = 0; kk < NP; ++kk) {
            const int k = D > 0 ? kk : NP - 1 - kk;
            p = vmax(v[k], vmin(p, c[k]));
        }
        if constexpr (!PAIRS) {
            P = p;
            Q = q;
        } else {
            const uint32_t pa = p & 0xFFFFu, pb = p >> 16, qa = q & 0xFFFFu, qb = q >> 16;
            // forward: low half then high half; backward: high then low
            hp = D > 0 ? pa : pb;
            hq = D > 0 ? qa : qb;
            P = D > 0 ? max(pb, min(pa, qb)) : max(pa, min(pb, qa));
            Q = min(qa, qb);
        }
    }
    // Scan the lane's pixels in direction D from x, the value next to its
    // first pixel in that direction (after reduce<D>).  Returns whether any
    // pixel changed (TRACK; always false otherwise).
    template <int D, bool TRACK = true>
    __device__ __forceinline__ bool apply(uint32_t x) {
        if constexpr (PAIRS) {
            const uint32_t x2 = max(hp, min(x, hq));  // entering the second half
            x = D > 0 ? (x | x2 << 16) : (x2 | x << 16);
        }
        bool d = false;
#pragma unroll
        for (int kk = 0; kk < NP; ++kk) {
            const int k = D > 0 ? kk : NP - 1 - kk;
            x = vmax(v[k], vmin(x, c[k]));
            if constexpr (TRACK) d |= x != v[k];
            v[k] = x;
        }
        return d;
    }
};

DEV bool tile_on(const TileMask& tm, int x, int y) {
    return tm.active == nullptr || tm.active[(size_t)(y / tm.th) * tm.ntx + x / tm.tw] != 0;
}

// Whether any tile under pixels [x0, x1) of row y is on (x0 < x1).  The same
// value in every lane when the arguments are.
DEV bool span_on(const TileMask& tm, int x0, int x1, int y) {
    if (tm.active == nullptr) return true;
    const uint8_t* row = tm.active + (size_t)(y / tm.th) * tm.ntx;
    for (int t = x0 / tm.tw; t <= (x1 - 1) / tm.tw; ++t)
        if (row[t]) return true;
    return false;
}

// ------------------------------------------------------------ descriptors ---
// The descriptor of image i, by value: from the kernel argument for a single
// image (scalar loads of the argument segment), else from the device table
// (the index is wave-uniform, so scalar loads as well).  Returning a
// reference would merge two address spaces into a generic pointer and turn
// every field read into a flat load waited on one at a time.
DEV ImgDesc desc(const Descs& a, int i) {
    if (a.d == nullptr) return a.one;
    return a.d[__builtin_amdgcn_readfirstlane(i)];
}
// Image owning global work item `idx` for a prefix field (binary search).
#define IWPP_FIND(NAME, FIELD)                                           \
    DEV int NAME(const Descs& a, int idx) {                              \
        if (a.d == nullptr) return 0;                                    \
        int lo = 0, hi = a.n - 1;                                        \
        while (lo < hi) {                                                \
            const int mid = (lo + hi + 1) >> 1;                          \
            if (a.d[mid].FIELD <= idx) lo = mid;                         \
            else hi = mid - 1;                                           \
        }                                                                \
        return lo;                                                       \
    }
IWPP_FIND(search_row, row_base)
IWPP_FIND(search_cw, cw_base)
IWPP_FIND(search_sw, sw_base)
IWPP_FIND(search_se, se_base)
#undef IWPP_FIND
// One load from the work-item table when there is one (a dependent chain of
// loads per work item otherwise: measured as the main cost of batched passes).
DEV int find_row(const Descs& a, int r) { return a.img_of_row ? (int)a.img_of_row[r] : search_row(a, r); }
DEV int find_cw(const Descs& a, int w) { return a.img_of_cw ? (int)a.img_of_cw[w] : search_cw(a, w); }
DEV int find_sw(const Descs& a, int w) { return a.img_of_sw ? (int)a.img_of_sw[w] : search_sw(a, w); }
DEV int find_mp(const Descs& a, int g) { return a.img_of_mp ? (int)a.img_of_mp[g] : 0; }
DEV Image image_of(const ImgDesc& d) { return Image{d.J, d.I, d.pj, d.pi, d.W, d.H}; }
DEV uint64_t first_lane64(uint64_t v) {
    return (uint64_t)first_lane((uint32_t)v) | (uint64_t)first_lane((uint32_t)(v >> 32)) << 32;
}
// The descriptor with every field in scalar registers.  It is the same in
// every lane, but read from the device table (several images) it arrives in
// vector registers (the compiler cannot use scalar loads on memory the kernel
// may write), and everything computed from it -- block geometry, row
// addresses, 64-bit offsets -- would then be computed per lane.
DEV ImgDesc scalar_desc(ImgDesc d) {
    d.J = (uint8_t*)first_lane64((uint64_t)d.J);
    d.I = (const uint8_t*)first_lane64((uint64_t)d.I);
    d.P = (uint8_t*)first_lane64((uint64_t)d.P);
    d.pj = (size_t)first_lane64((uint64_t)d.pj);
    d.pi = (size_t)first_lane64((uint64_t)d.pi);
    d.pp = (size_t)first_lane64((uint64_t)d.pp);
    d.W = (int)first_lane((uint32_t)d.W);
    d.H = (int)first_lane((uint32_t)d.H);
    d.nmx = (int)first_lane((uint32_t)d.nmx);
    d.nmy = (int)first_lane((uint32_t)d.nmy);
    d.R = (int)first_lane((uint32_t)d.R);
    d.strips = (int)first_lane((uint32_t)d.strips);
    d.mp_base = (int)first_lane((uint32_t)d.mp_base);
    d.row_base = (int)first_lane((uint32_t)d.row_base);
    d.cw_base = (int)first_lane((uint32_t)d.cw_base);
    d.sw_base = (int)first_lane((uint32_t)d.sw_base);
    d.tm.active = (const uint8_t*)first_lane64((uint64_t)d.tm.active);
    d.tm.tw = (int)first_lane((uint32_t)d.tm.tw);
    d.tm.th = (int)first_lane((uint32_t)d.tm.th);
    d.tm.ntx = (int)first_lane((uint32_t)d.tm.ntx);
    d.rounds = (int)first_lane((uint32_t)d.rounds);
    d.edge_base = (long long)first_lane64((uint64_t)d.edge_base);
    d.se_base = (int)first_lane((uint32_t)d.se_base);
    return d;
}
// The descriptor of the image owning global megapixel m, for the queue
// kernels: the kernel argument for a single image (ONE: scalar registers, no
// load), else the device table, the fields a visit uses taken through
// readfirstlane so that the address arithmetic built on them stays scalar.
template <bool ONE>
DEV ImgDesc mp_desc(const Descs& a, uint32_t m) {
    if constexpr (ONE) {
        return a.one;
    } else {
        ImgDesc d = desc(a, find_mp(a, (int)m));
        d.J = (uint8_t*)first_lane64((uint64_t)d.J);
        d.I = (const uint8_t*)first_lane64((uint64_t)d.I);
        d.pj = (size_t)first_lane64((uint64_t)d.pj);
        d.pi = (size_t)first_lane64((uint64_t)d.pi);
        d.W = (int)first_lane((uint32_t)d.W);
        d.H = (int)first_lane((uint32_t)d.H);
        d.nmx = (int)first_lane((uint32_t)d.nmx);
        d.mp_base = (int)first_lane((uint32_t)d.mp_base);
        return d;
    }
}

// ------------------------------------------------------------------ clip ----
template <class T, int VB>
__global__ __launch_bounds__(kWG) void k_clip(Image im) {
    using V = VecIO<T, VB>;
    const int rowbytes = im.W * (int)sizeof(T);
    const int nvec = (rowbytes + VB - 1) / VB;
    const long total = (long)nvec * im.H;
    for (long i = (long)blockIdx.x * kWG + threadIdx.x; i < total; i += (long)gridDim.x * kWG) {
        const int y = (int)(i / nvec), o = (int)(i % nvec) * VB;
        uint8_t* jr = im.J + (size_t)y * im.pj;
        const uint8_t* ir = im.I + (size_t)y * im.pi;
        uint32_t jw[V::NW], iw[V::NW], jp[V::NPX], ip[V::NPX];
        V::load(jr, o, rowbytes, jw);
        V::load(ir, o, rowbytes, iw);
        V::unpack(jw, jp);
        V::unpack(iw, ip);
        bool ch = false;
#pragma unroll
        for (int k = 0; k < V::NPX; ++k) {
            ch |= jp[k] > ip[k];
            jp[k] = min(jp[k], ip[k]);
        }
        if (ch) {
            V::pack(jp, jw);
            V::store(jr, o, rowbytes, jw);
        }
    }
}

// -------------------------------------------------------------- row pass ----
// One workgroup of NWV waves per row (grid-stride).  A row is split in chunks
// of (NWV*64 lanes x VB bytes), all held in registers; forward then backward
// scans with carries through LDS.  In place.  CLIP folds marker =
// min(marker, mask) into the load.
template <class T, int VB, int NCH, int NWV, bool CLIP>
__global__ __launch_bounds__(NWV * 64) void k_row_pass(Image im, TileMask tm, uint32_t* changed) {
    using V = VecIO<T, VB>;
    using L = LanePx<T, VB>;
    constexpr int NW = V::NW, NT = NWV * 64;
    __shared__ uint32_t sP[2][NCH][NWV], sQ[2][NCH][NWV];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rowbytes = im.W * (int)sizeof(T);
    bool any_change = false;
    for (int y = blockIdx.x; y < im.H; y += gridDim.x) {
        uint8_t* jr = im.J + (size_t)y * im.pj;
        const uint8_t* ir = im.I + (size_t)y * im.pi;
        uint32_t jv[NCH][NW], iv[NCH][NW];
        bool wall[NCH], dch[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int o = (ch * NT + tid) * VB;
            // A chunk whose tiles are all off is neither read nor written:
            // zero marker and mask make it a wall (carry 0, a lower bound).
            // The clipping round reads every chunk.
            const int cx0 = ch * NT * VB / (int)sizeof(T);
            if (!CLIP && tm.active &&
                (cx0 >= im.W || !span_on(tm, cx0, min(im.W, cx0 + NT * VB / (int)sizeof(T)), y))) {
#pragma unroll
                for (int i = 0; i < NW; ++i) jv[ch][i] = iv[ch][i] = 0;
                wall[ch] = true;
                dch[ch] = false;
                continue;
            }
            V::load(jr, o, rowbytes, jv[ch]);
            V::load(ir, o, rowbytes, iv[ch]);
            wall[ch] = o < rowbytes && !tile_on(tm, o / (int)sizeof(T), y);
            dch[ch] = false;
        }
        // forward (left to right)
        uint32_t carry = 0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            L px;
            px.load(jv[ch], iv[ch]);
            if (CLIP) dch[ch] |= px.clip();
            if (wall[ch]) px.wall();
            uint32_t Pi, Qi;
            px.template reduce<1>(Pi, Qi);
            scan_lanes_fwd(Pi, Qi);
            const uint32_t Pe = from_left(Pi, 0u), Qe = from_left(Qi, 0xFFFFFFFFu);
            if (lane == 63) {
                sP[0][ch][wv] = Pi;
                sQ[0][ch][wv] = Qi;
            }
            __syncthreads();
            uint32_t xin, xall;
            wave_carry<NWV, 1>(sP[0][ch], sQ[0][ch], wv, lane, carry, xin, xall);
            carry = xall;
            dch[ch] |= px.template apply<1>(max(Pe, min(xin, Qe)));  // from the value left of this lane
            px.store(jv[ch]);
        }
        // backward (right to left)
        carry = 0;
#pragma unroll
        for (int ch = NCH - 1; ch >= 0; --ch) {
            L px;
            px.load(jv[ch], iv[ch]);
            if (wall[ch]) px.wall();
            uint32_t Pi, Qi;
            px.template reduce<-1>(Pi, Qi);
            scan_lanes_bwd(Pi, Qi, lane);
            const uint32_t Pe = from_right(Pi, 0u), Qe = from_right(Qi, 0xFFFFFFFFu);
            if (lane == 0) {
                sP[1][ch][wv] = Pi;
                sQ[1][ch][wv] = Qi;
            }
            __syncthreads();
            uint32_t xin, xall;
            wave_carry<NWV, -1>(sP[1][ch], sQ[1][ch], wv, lane, carry, xin, xall);
            carry = xall;
            dch[ch] |= px.template apply<-1>(max(Pe, min(xin, Qe)));  // from the value right of this lane
            px.store(jv[ch]);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (dch[ch]) V::store(jr, (ch * NT + tid) * VB, rowbytes, jv[ch]);
            any_change |= dch[ch];
        }
    }
    if (any_change && changed) *changed = 1u;
}

// Row pass with the whole row held in registers by one 256-lane workgroup:
// NCH chunks of 256 lanes x 16 bytes, read once, swept forward and backward,
// written once (only the vectors that changed).  A direction takes one
// barrier: every (chunk, wave) pair publishes its summary (P, Q) in LDS, then
// each wave folds the summaries that precede it in scan order to get the
// value entering each of its chunks; within a wave the lanes' exclusive
// prefixes come from a DPP scan made before the barrier.  All of a row's
// loads are issued together, with no branch around them (a lane with nothing
// to read re-reads the row's first vector, and its data is replaced by
// zeros), so they are in flight at once.  Pixel handling (walls for tiles
// that are off, clipping, packed pairs) is LanePx's, as in k_row_pass.  A
// wave's part of a chunk whose tiles are all off is not read or written
// (outside the clipping round): zero marker and mask make it a wall.
// An empty asm that claims to change x: values the compiler derived from x
// earlier are then recomputed from it rather than kept alive in registers
// until their next use.
DEV void reg_fence(uint32_t& x) { asm volatile("" : "+v"(x)); }
template <int NCH, int NW>
DEV void reg_fence(uint32_t (&a)[NCH][NW], int ch) {
#pragma unroll
    for (int i = 0; i < NW; ++i) reg_fence(a[ch][i]);
}
// Inclusive scan, across the lanes of a wave, of NE (<= 64) summaries
// (P, Q) stored in LDS in row order; lane l holds the l-th in scan
// direction D.  Returns P: with the value 0 entering the row, the value
// leaving the l-th summary.
template <int NE, int D>
DEV uint32_t entry_scan(const uint32_t* sp, const uint32_t* sq, int lane) {
    uint32_t P = 0, Q = 0xFFFFFFFFu;
    if (lane < NE) {
        const int e = D > 0 ? lane : NE - 1 - lane;
        P = sp[e];
        Q = sq[e];
    }
    scan_lanes_fwd(P, Q);
    return P;
}
DEV void ld16(const uint8_t* base, uint32_t off, uint32_t (&w)[4]) {
    const u32x4 t = *gp(reinterpret_cast<const u32x4*>(base + off));
    w[0] = t.x;
    w[1] = t.y;
    w[2] = t.z;
    w[3] = t.w;
}

template <class T, int NT, int NCH, bool CLIP>
__global__ __launch_bounds__(NT) void k_row_pass_held(Image im, TileMask tm, uint32_t* changed) {
    using L = LanePx<T, 16>;
    constexpr int NWV = NT / 64, PPL = 16 / (int)sizeof(T);  // pixels per lane vector
    static_assert(NCH <= 16 && NCH * NWV <= 64, "chunk flags fit a mask, entries a wave");
    __shared__ uint32_t sP[2][NCH][NWV], sQ[2][NCH][NWV];
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    bool any_change = false;
    for (int y = blockIdx.x; y < im.H; y += gridDim.x) {
        // The thread index and the row length pass through opaque copies in
        // every iteration: hoisted out of the loop, the per-lane offsets and
        // conditions derived from them would stay live in registers across
        // the whole kernel.
        int tid = threadIdx.x, rowbytes = im.W * (int)sizeof(T);
        asm volatile("" : "+v"(tid));
        asm volatile("" : "+s"(rowbytes));
        const int lane = tid & 63;
        uint8_t* jr = im.J + (size_t)y * im.pj;
        const uint8_t* ir = im.I + (size_t)y * im.pi;
        // per-chunk lane bits: `read` = the vector is read, `wall` = no
        // conduction through it
        uint32_t read = 0, wall = 0;
        if (tm.active) {
            const uint8_t* trow = tm.active + (size_t)(y / tm.th) * tm.ntx;
            uint32_t on = 0;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                // every lane loads (no branch around the loads)
                const int x0 = (ch * NT + tid) * PPL;
                const uint8_t f = *gp(trow + min(x0, im.W - 1) / tm.tw);
                if (x0 < im.W && f != 0) on |= 1u << ch;
            }
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                const bool inside = (ch * NT + tid) * 16 < rowbytes;
                const bool lane_on = on >> ch & 1u;
                const bool wave_on = wave_any(lane_on);  // (every lane votes)
                if (inside && (CLIP || wave_on)) read |= 1u << ch;
                if (!lane_on) wall |= 1u << ch;
            }
        } else {
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
                if ((ch * NT + tid) * 16 < rowbytes) read |= 1u << ch;
            wall = ~read;
        }
        uint32_t jv[NCH][4], iv[NCH][4];
        uint32_t xp[NCH], xq[NCH];  // lane-exclusive prefix within the wave, then the value in
        uint32_t hpq[NCH];          // summary of the lane's first-scanned half (packed pairs)
        uint32_t dch = 0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const uint32_t o = (read >> ch & 1u) ? (uint32_t)(ch * NT + tid) * 16u : 0u;
            ld16(jr, o, jv[ch]);
            ld16(ir, o, iv[ch]);
        }
        // The chunk flags, re-read opaquely at each use: otherwise the
        // compiler splits them into one register per chunk up front.
        const auto bit = [](uint32_t m, int ch) {
            asm volatile("" : "+v"(m));
            return m >> ch & 1u;
        };
        // ---- A: clip, walls, forward summaries of every chunk
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            reg_fence(jv, ch);
            reg_fence(iv, ch);
            {
                // vectors not read, and the bytes of the last vector past the
                // row's end, become zero (marker and mask), so they are walls
                const int lim = bit(read, ch) ? rowbytes - (ch * NT + tid) * 16 : 0;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int vb = lim - 4 * i;
                    const uint32_t keep = vb >= 4 ? 0xFFFFFFFFu
                                        : vb <= 0 ? 0u : (0xFFFFFFFFu >> (8 * (4 - vb)));
                    jv[ch][i] &= keep;
                    iv[ch][i] &= keep;
                }
            }
            L px;
            px.load(jv[ch], iv[ch]);
            if (CLIP && px.clip()) {  // the clipped marker is kept (and written back)
                px.store(jv[ch]);
                dch |= 1u << ch;
            }
            // a wall's mask is zero from here on: no rise, no conduction
            const uint32_t cm = bit(wall, ch) ? 0u : 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < 4; ++i) iv[ch][i] &= cm;
#pragma unroll
            for (int k = 0; k < L::NP; ++k) px.c[k] &= cm;
            uint32_t Pi, Qi;
            px.template reduce<1>(Pi, Qi);
            hpq[ch] = px.hp | px.hq << 16;
            scan_lanes_fwd(Pi, Qi);
            xp[ch] = from_left(Pi, 0u);
            xq[ch] = from_left(Qi, 0xFFFFFFFFu);
            if (lane == 63) {
                sP[0][ch][wv] = Pi;
                sQ[0][ch][wv] = Qi;
            }
            reg_fence(xp[ch]);  // one chunk at a time: fewer live registers
            reg_fence(xq[ch]);
            reg_fence(hpq[ch]);
        }
        __syncthreads();
        {
            const uint32_t inc = entry_scan<NCH * NWV, 1>(&sP[0][0][0], &sQ[0][0][0], lane);
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                const int pos = ch * NWV + wv;  // this (chunk, wave) in scan order
                const uint32_t xin = pos > 0 ? rdlane(inc, pos - 1) : 0u;
                xp[ch] = max(xp[ch], min(xin, xq[ch]));
            }
        }
        // ---- B: forward scan of every chunk, then its backward summaries
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            reg_fence(jv, ch);
            reg_fence(iv, ch);
            L px;
            px.load(jv[ch], iv[ch]);
            px.hp = hpq[ch] & 0xFFFFu;
            px.hq = hpq[ch] >> 16;
            const bool d = px.template apply<1>(xp[ch]);
            uint32_t Pi, Qi;
            px.template reduce<-1>(Pi, Qi);
            hpq[ch] = px.hp | px.hq << 16;
            if (d) {
                px.store(jv[ch]);
                dch |= 1u << ch;
            }
            scan_lanes_bwd(Pi, Qi, lane);
            xp[ch] = from_right(Pi, 0u);
            xq[ch] = from_right(Qi, 0xFFFFFFFFu);
            if (lane == 0) {
                sP[1][ch][wv] = Pi;
                sQ[1][ch][wv] = Qi;
            }
            reg_fence(jv, ch);
            reg_fence(xp[ch]);
            reg_fence(xq[ch]);
            reg_fence(hpq[ch]);
        }
        __syncthreads();
        {
            const uint32_t inc = entry_scan<NCH * NWV, -1>(&sP[1][0][0], &sQ[1][0][0], lane);
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                const int pos = NCH * NWV - 1 - (ch * NWV + wv);
                const uint32_t xin = pos > 0 ? rdlane(inc, pos - 1) : 0u;
                xp[ch] = max(xp[ch], min(xin, xq[ch]));
            }
        }
        // ---- C: backward scan of every chunk
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            reg_fence(jv, ch);
            reg_fence(iv, ch);
            L px;
            px.load(jv[ch], iv[ch]);
            px.hp = hpq[ch] & 0xFFFFu;
            px.hq = hpq[ch] >> 16;
            if (px.template apply<-1>(xp[ch])) {
                px.store(jv[ch]);
                dch |= 1u << ch;
            }
            reg_fence(jv, ch);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (!(dch >> ch & 1u)) continue;
            const int o = (ch * NT + tid) * 16;
            if (o + 16 <= rowbytes) {
                *gp(reinterpret_cast<u32x4*>(jr + o)) = u32x4{jv[ch][0], jv[ch][1], jv[ch][2], jv[ch][3]};
            } else {  // the row's last, partial vector
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    const int ob = o + k * (int)sizeof(T);
                    if (ob < rowbytes)
                        *gp(reinterpret_cast<T*>(jr + ob)) =
                            (T)(jv[ch][k * (int)sizeof(T) / 4] >> (8 * ((k * (int)sizeof(T)) % 4)));
                }
            }
        }
        any_change |= dch != 0;
    }
    if (any_change && changed) *changed = 1u;
}

// Streaming row pass for wide rows: one chunk (256 lanes x VB bytes) in
// registers at a time, carried along the row.  DIR = +1 / -1 sweeps one
// direction; DIR = 0 sweeps forward then backward in the same launch, the
// backward sweep re-reading the row this workgroup just wrote (each lane
// reads back its own vectors, which the caches still hold), so the memory
// traffic stays at one read of marker and mask and one write of the marker.
// One direction (DIR = +1 left to right, -1 right to left) over one row; the
// direction is a template parameter so that every per-pixel index is a
// constant and the pixels stay in fixed registers.
template <class T, int VB, int dir, bool clip>
DEV void row_stream_dir(const Image& im, int y, const TileMask& tm, uint32_t (&sP)[2][4],
                        uint32_t (&sQ)[2][4], bool& any_change) {
    using V = VecIO<T, VB>;
    constexpr int NW = V::NW;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rowbytes = im.W * (int)sizeof(T);
    const int nch = (rowbytes + kWG * VB - 1) / (kWG * VB);
    uint8_t* jr = im.J + (size_t)y * im.pj;
    const uint8_t* ir = im.I + (size_t)y * im.pi;
    {
        uint32_t carry = 0;
        for (int n = 0; n < nch; ++n) {
            const int ch = dir > 0 ? n : nch - 1 - n;
            const int o = (ch * kWG + tid) * VB;
            if (!clip && tm.active) {
                // all tiles of the chunk off: skip it, restart the carry at 0
                // (a lower bound of the marker), touch no LDS
                const int cx0 = ch * kWG * VB / (int)sizeof(T);
                if (!span_on(tm, cx0, min(im.W, cx0 + kWG * VB / (int)sizeof(T)), y)) {
                    carry = 0;
                    continue;
                }
            }
            uint32_t jw[NW], iw[NW];
            V::load(jr, o, rowbytes, jw);
            V::load(ir, o, rowbytes, iw);
            LanePx<T, VB> L;
            L.load(jw, iw);
            bool d = clip && L.clip();
            if (o < rowbytes && !tile_on(tm, o / (int)sizeof(T), y)) L.wall();
            uint32_t P, Q;
            L.template reduce<dir>(P, Q);
            uint32_t Pi = P, Qi = Q, Pe, Qe;
            if (dir > 0) {
                scan_lanes_fwd(Pi, Qi);
                Pe = from_left(Pi, 0u);
                Qe = from_left(Qi, 0xFFFFFFFFu);
                if (lane == 63) { sP[n & 1][wv] = Pi; sQ[n & 1][wv] = Qi; }
            } else {
                scan_lanes_bwd(Pi, Qi, lane);
                Pe = from_right(Pi, 0u);
                Qe = from_right(Qi, 0xFFFFFFFFu);
                if (lane == 0) { sP[n & 1][wv] = Pi; sQ[n & 1][wv] = Qi; }
            }
            __syncthreads();
            uint32_t xin = carry, xall = carry;
#pragma unroll
            for (int uu = 0; uu < 4; ++uu) {
                const int u = dir > 0 ? uu : 3 - uu;
                const uint32_t tp = sP[n & 1][u], tq = sQ[n & 1][u];
                if (dir > 0 ? (u < wv) : (u > wv)) xin = max(tp, min(xin, tq));
                xall = max(tp, min(xall, tq));
            }
            carry = xall;
            d |= L.template apply<dir>(max(Pe, min(xin, Qe)));
            if (d) {
                L.store(jw);
                V::store(jr, o, rowbytes, jw);
            }
            any_change |= d;
            __syncthreads();  // the LDS slot pair is reused two chunks later
        }
    }
}

// Streaming sweep of one row: DIR = +1 / -1 one direction, DIR = 0 forward
// then backward.  CLIP applies to the first direction swept.
template <class T, int VB, int DIR, bool CLIP>
DEV void row_stream(const Image& im, int y, const TileMask& tm, uint32_t (&sP)[2][4],
                    uint32_t (&sQ)[2][4], bool& any_change) {
    if constexpr (DIR >= 0) row_stream_dir<T, VB, 1, CLIP>(im, y, tm, sP, sQ, any_change);
    if constexpr (DIR == 0) row_stream_dir<T, VB, -1, false>(im, y, tm, sP, sQ, any_change);
    if constexpr (DIR < 0) row_stream_dir<T, VB, -1, CLIP>(im, y, tm, sP, sQ, any_change);
}

// Streaming row pass for wide rows: one chunk (256 lanes x VB bytes) in
// registers at a time, carried along the row.  DIR = +1 / -1 sweeps one
// direction; DIR = 0 sweeps forward then backward in the same launch, the
// backward sweep re-reading the row this workgroup just wrote (each lane
// reads back its own vectors, which the caches still hold), so the memory
// traffic stays at one read of marker and mask and one write of the marker.
template <class T, int VB, int DIR, bool CLIP>
__global__ __launch_bounds__(kWG) void k_row_pass_long(Image im, TileMask tm, uint32_t* changed) {
    __shared__ uint32_t sP[2][4], sQ[2][4];
    bool any_change = false;
    for (int y = blockIdx.x; y < im.H; y += gridDim.x)
        row_stream<T, VB, DIR, CLIP>(im, y, tm, sP, sQ, any_change);
    if (any_change && changed) *changed = 1u;
}

// The same over the rows of several images (global row index -> image).
// Rows of images whose sweep rounds are done are skipped; in the first round
// such rows are only clipped (marker = min(marker, mask)).
template <class T, int VB>
DEV void row_clip(const Image& im, int y, bool& any_change, int t = threadIdx.x, int nt = kWG) {
    using V = VecIO<T, VB>;
    const int rowbytes = im.W * (int)sizeof(T);
    uint8_t* jr = im.J + (size_t)y * im.pj;
    const uint8_t* ir = im.I + (size_t)y * im.pi;
    for (int o = t * VB; o < rowbytes; o += nt * VB) {
        uint32_t jw[V::NW], iw[V::NW], a[V::NPX], b[V::NPX];
        V::load(jr, o, rowbytes, jw);
        V::load(ir, o, rowbytes, iw);
        V::unpack(jw, a);
        V::unpack(iw, b);
        bool ch = false;
#pragma unroll
        for (int k = 0; k < V::NPX; ++k) {
            ch |= a[k] > b[k];
            a[k] = min(a[k], b[k]);
        }
        if (ch) {
            V::pack(a, jw);
            V::store(jr, o, rowbytes, jw);
            any_change = true;
        }
    }
}

template <class T, int VB, bool CLIP>
__global__ __launch_bounds__(kWG) void k_row_pass_batch(Descs ds, uint32_t* changed) {
    __shared__ uint32_t sP[2][4], sQ[2][4];
    bool any_change = false;
    for (int r = blockIdx.x; r < ds.total; r += gridDim.x) {
        const ImgDesc d = desc(ds, find_row(ds, r));
        if (ds.round >= d.rounds) {  // the same for the whole workgroup
            if (CLIP) row_clip<T, VB>(image_of(d), r - d.row_base, any_change);
            continue;
        }
        row_stream<T, VB, 0, CLIP>(image_of(d), r - d.row_base, d.tm, sP, sQ, any_change);
    }
    if (any_change && changed) *changed = 1u;
}

// Row pass with one wave per row, for rows of at most NCH chunks of
// (64 lanes x VB bytes): the row is held in registers, scanned forward then
// backward with the carry between chunks kept in the wave (no LDS, no
// barriers), and written back where it changed.  Rows are global work items
// of one or several images, as in k_row_pass_batch; a row of an image whose
// sweep rounds are done is only clipped (first round) or skipped.  Short rows
// leave most lanes of a 256-lane workgroup idle; here a workgroup sweeps four
// rows at once.
template <class T, int VB, int NCH, bool CLIP>
__global__ __launch_bounds__(kWG) void k_row_pass_wave(Descs ds, uint32_t* changed) {
    using V = VecIO<T, VB>;
    using L = LanePx<T, VB>;
    constexpr int NW = V::NW, CB = 64 * VB, WPB = kWG / 64;
    const int lane = lane_id();
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    bool any_change = false;
    for (int r = blockIdx.x * WPB + wv; r < ds.total; r += gridDim.x * WPB) {
        const ImgDesc d = desc(ds, find_row(ds, r));
        const Image im = image_of(d);
        const int y = r - d.row_base;
        if (ds.round >= d.rounds) {  // the same in the whole wave
            if (CLIP) row_clip<T, VB>(im, y, any_change, lane, 64);
            continue;
        }
        const TileMask& tm = d.tm;
        const int rowbytes = im.W * (int)sizeof(T);
        const int nch = (rowbytes + CB - 1) / CB;  // <= NCH (the launch checks)
        uint8_t* jr = im.J + (size_t)y * im.pj;
        const uint8_t* ir = im.I + (size_t)y * im.pi;
        uint32_t jv[NCH][NW], iv[NCH][NW];
        bool wall[NCH], dch[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            dch[ch] = false;
            wall[ch] = true;
#pragma unroll
            for (int i = 0; i < NW; ++i) jv[ch][i] = iv[ch][i] = 0;
            if (ch >= nch) continue;
            // a chunk whose tiles are all off stays a zero wall (not read,
            // not written); the clipping round reads every chunk
            const int cx0 = ch * CB / (int)sizeof(T);
            if (!CLIP && tm.active && !span_on(tm, cx0, min(im.W, cx0 + CB / (int)sizeof(T)), y))
                continue;
            const int o = ch * CB + lane * VB;
            V::load(jr, o, rowbytes, jv[ch]);
            V::load(ir, o, rowbytes, iv[ch]);
            wall[ch] = o < rowbytes && !tile_on(tm, o / (int)sizeof(T), y);
        }
        uint32_t carry = 0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (ch >= nch) break;
            L px;
            px.load(jv[ch], iv[ch]);
            if (CLIP) dch[ch] |= px.clip();
            if (wall[ch]) px.wall();
            uint32_t Pi, Qi;
            px.template reduce<1>(Pi, Qi);
            scan_lanes_fwd(Pi, Qi);
            const uint32_t Pe = from_left(Pi, 0u), Qe = from_left(Qi, 0xFFFFFFFFu);
            dch[ch] |= px.template apply<1>(max(Pe, min(carry, Qe)));
            carry = max(rdlane(Pi, 63), min(carry, rdlane(Qi, 63)));
            px.store(jv[ch]);
        }
        carry = 0;
#pragma unroll
        for (int ch = NCH - 1; ch >= 0; --ch) {
            if (ch >= nch) continue;
            L px;
            px.load(jv[ch], iv[ch]);
            if (wall[ch]) px.wall();
            uint32_t Pi, Qi;
            px.template reduce<-1>(Pi, Qi);
            scan_lanes_bwd(Pi, Qi, lane);
            const uint32_t Pe = from_right(Pi, 0u), Qe = from_right(Qi, 0xFFFFFFFFu);
            dch[ch] |= px.template apply<-1>(max(Pe, min(carry, Qe)));
            carry = max(rdlane(Pi, 0), min(carry, rdlane(Qi, 0)));
            px.store(jv[ch]);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (ch >= nch) break;
            if (dch[ch]) V::store(jr, ch * CB + lane * VB, rowbytes, jv[ch]);
            any_change |= dch[ch];
        }
    }
    if (any_change && changed) *changed = 1u;
}

// ----------------------------------------------------------- column pass ----
// Each wave sweeps one block of (64 lanes x VB bytes) columns x R rows,
// reading src and writing dst (out of place).  Inside the block each row
// uses the block's previous output row; neighbours outside the block come
// from src, so the pass is deterministic.
//
// SEED (the last pass of the sweep phase, in place, no tile schedule): the
// pass also sets flags[mp_base + my * nmx + mx] = 1 for every megapixel that
// holds a pixel p with a neighbour q such that min(J(q), I(p)) > J(p) at the
// end of the pass (the seeding test of k_unstable_scan), for the pairs with
// p and q both in this block: when a row is computed it is final, and so is
// the row before it, so the pairs inside the row and between the two rows
// are tested (both ways) at once.  Pairs across a strip edge or a block
// boundary are k_seed_edges' (after the pass, on final values).  The flags
// must be zero before the pass.  For k_seed_edges the pass writes the strip's
// border columns as it leaves them: edges[(strip * 2 + side) * H + y] =
// marker | mask << 16 of the strip's first (side 0) or last (side 1) pixel.
//
// RASTER: after its column step, each row is scanned forward and backward
// across the strip (lane scans of (P, Q) summaries, as in the row passes),
// entering from the pixels just left and right of the strip in the same row
// (read from S: lower bounds of the reconstruction, in place or not).  A pass
// then carries a rise down (or up) and sideways within its block, along the
// staircase paths a column pass followed by a row pass would need rounds for:
// Vincent's raster scan, per block.  CLIP: the marker is clipped to the mask
// as it is read, the pixels read from outside the block or strip included.
template <class T, int VB, int CONN, int DIR, bool INPLACE, bool SEED = false,
          bool RASTER = false, bool CLIP = false>
DEV void col_block(const uint8_t* S, uint8_t* D, const uint8_t* __restrict__ I, size_t ps,
                   size_t pd, size_t pi, int W, int H, int y0, int y1, int strip,
                   const TileMask& tm, uint32_t* changed, uint8_t* flags = nullptr,
                   int mp_base = 0, int nmx = 0, uint32_t* edges = nullptr,
                   bool far_end = false) {
    using V = VecIO<T, VB>;
    using L = LanePx<T, VB>;
    constexpr int NW = V::NW, NP = L::NP;
    static_assert(!CLIP || RASTER, "clipping passes are raster passes");
    // in-place raster passes that do not clip follow another pass of the
    // round (or an earlier round), whose scans left their rows closed
    constexpr bool SKIP_CLOSED = RASTER && INPLACE && !CLIP;
    const int lane = lane_id();
    const int rowbytes = W * (int)sizeof(T);
    const int xb0 = strip * 64 * VB;  // first byte of the strip
    if (xb0 >= rowbytes) return;      // whole wave past the row end (no barriers below)
    const int o = xb0 + lane * VB;
    const int xs0 = xb0 / (int)sizeof(T);             // first pixel of the strip
    const int xs1 = (xb0 + 64 * VB) / (int)sizeof(T); // first pixel right of the strip
    // With a tile schedule, the lane's tile is read once per tile row the
    // block spans (a block of at most tm.th rows spans two; taller ones read
    // it again at each row, the slow way): a per-row read of the table in the
    // row loop would hold every row on a dependent load.  A block whose tiles
    // are all off is skipped in both directions of a round (the same test in
    // both), so its marker keeps its value, which neighbouring blocks read as
    // a lower bound.  (Out of place, a skipped block leaves the destination
    // plane as it was: only the in-place mode is used with tile schedules.)
    const bool lane_px = o < rowbytes;
    const int ytb = tm.active ? (y0 / tm.th + 1) * tm.th : 0;  // first row of the next tile row
    const bool two_rows = tm.active == nullptr || y1 - 1 < ytb + tm.th;
    const bool on0 = lane_px && tile_on(tm, o / (int)sizeof(T), y0);
    const bool on1 = lane_px && tile_on(tm, o / (int)sizeof(T), min(y1 - 1, ytb));
    if (tm.active && two_rows && !wave_any(on0 || on1)) return;
    auto row_on = [&](int y) -> bool {
        if (tm.active == nullptr) return lane_px;
        if (two_rows) return y < ytb ? on0 : on1;
        return lane_px && tile_on(tm, o / (int)sizeof(T), y);
    };
    // The previous output row, as the lane's pixel registers (16-bit pairs
    // with 16-byte lanes: pixel k and k + NP in one register).
    uint32_t prev[NP];
    {
        const int yh = DIR > 0 ? y0 - 1 : y1;
        uint32_t w[NW], wm[NW];
        if (yh >= 0 && yh < H) {
            V::load(S + (size_t)yh * ps, o, rowbytes, w);
            if constexpr (CLIP) V::load(I + (size_t)yh * pi, o, rowbytes, wm);
        } else {
#pragma unroll
            for (int i = 0; i < NW; ++i) w[i] = wm[i] = 0;
        }
        L::split(w, prev);
        if constexpr (CLIP) {
            uint32_t pm[NP];
            L::split(wm, pm);
#pragma unroll
            for (int k = 0; k < NP; ++k) prev[k] = L::vmin(prev[k], pm[k]);
        }
    }
    // One pixel of S in every lane (0 outside the image), clipped to its mask
    // under CLIP: loaded from a clamped address with no branch (y and x are
    // the same in every lane).
    auto px_at = [&](int y, int x) -> uint32_t {
        const bool in = y >= 0 && y < H && x >= 0 && x < W;
        const size_t yc = (size_t)min(max(y, 0), H - 1), xc = (size_t)min(max(x, 0), W - 1);
        uint32_t v = *gp(reinterpret_cast<const T*>(S + yc * ps + xc * sizeof(T)));
        if constexpr (CLIP)
            v = min(v, (uint32_t)*gp(reinterpret_cast<const T*>(I + yc * pi + xc * sizeof(T))));
        return in ? v : 0u;
    };
    bool any_change = false;
    // ---- seeding state (SEED): r1 = the row computed before (final), c1 its
    // mask, v1 what r1 offers a vertical neighbour inside the strip
    // (8-connectivity: the pixel and its left and right neighbours; 4: the
    // pixel).
    const int mx_lane = (o / (int)sizeof(T)) >> 6;
    const bool lane_in = o < rowbytes;
    uint32_t r1[NP], c1[NP], v1[NP];
    uint32_t acc = 0;  // this lane's pixels of band `band`: a flag
    int band = -1;
    // neighbour maxima inside the strip (pixels outside count as 0); with
    // `self` the pixel itself too
    auto nbmax = [&](const uint32_t (&Rw)[NP], uint32_t (&Dst)[NP], bool self) {
        uint32_t left, right;
        if constexpr (L::PAIRS) {
            left = from_left(Rw[NP - 1] >> 16, 0u) | (Rw[NP - 1] << 16);
            right = (Rw[0] >> 16) | (from_right(Rw[0] & 0xFFFFu, 0u) << 16);
        } else {
            left = from_left(Rw[NP - 1], 0u);
            right = from_right(Rw[0], 0u);
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const uint32_t a = k > 0 ? Rw[k - 1] : left, b = k < NP - 1 ? Rw[k + 1] : right;
            Dst[k] = self ? L::vmax(Rw[k], L::vmax(a, b)) : L::vmax(a, b);
        }
    };
    auto flag_band = [&](int y, bool u) {  // accumulated per 64-row band
        const int b = y >> 6;
        if (b != band) {
            if (acc && lane_in) flags[mp_base + band * nmx + mx_lane] = 1;
            band = b;
            acc = 0;
        }
        acc |= u ? 1u : 0u;
    };
    // the strip's border columns of one 64-row band (lane = row), stored
    // for the rows of this block when the pass leaves the band
    uint32_t eL = 0, eR = 0;
    int eband = -1;
    auto edge_flush = [&]() {
        const int yr = eband * 64 + lane;
        if (eband >= 0 && yr >= y0 && yr < y1) {
            *gp(edges + (size_t)strip * 2 * H + yr) = eL;
            *gp(edges + ((size_t)strip * 2 + 1) * H + yr) = eR;
        }
    };
    // Rows go in groups of CU, double-buffered: the loads of the next group
    // (marker, mask, and the single pixels beside the strip) are issued
    // before the rows of this group are computed, so that a wave's memory
    // latency overlaps its own arithmetic (a raster row's scans take about as
    // long as the loads of a group).  Every load is unconditional: rows past
    // the block re-read its last row (in bounds, not computed), pixels
    // outside the image come from clamped addresses and are replaced by 0.
    // 8-connectivity takes the strip's outside neighbours in the previous
    // row from memory in a column pass; a raster pass has them already (the
    // previous row's entry pixels).
    constexpr int CU = RASTER ? (VB == 16 ? 2 : 4) : (VB == 16 ? 4 : 8);
    const int n = y1 - y0;
    struct Group {
        uint32_t jw[CU][NW], iw[CU][NW], lc[CU], rc[CU];
    };
    auto load_group = [&](int i0, Group& g) {
#pragma unroll
        for (int u = 0; u < CU; ++u) {
            const int i = min(i0 + u, n - 1);
            const int y = DIR > 0 ? y0 + i : y1 - 1 - i;
            V::load(S + (size_t)y * ps, o, rowbytes, g.jw[u]);
            V::load(I + (size_t)y * pi, o, rowbytes, g.iw[u]);
            g.lc[u] = g.rc[u] = 0;
            if constexpr (RASTER) {  // the pixels the row scans enter from, in this row
                g.lc[u] = px_at(y, xs0 - 1);
                g.rc[u] = px_at(y, xs1);
            } else if constexpr (CONN == 8) {  // the outside neighbours in the previous row
                g.lc[u] = px_at(y - DIR, xs0 - 1);
                g.rc[u] = px_at(y - DIR, xs1);
            }
        }
    };
    // raster: the previous row's entry pixels (for the block's first row, the
    // row before the block)
    uint32_t pl = 0, pr = 0;
    if constexpr (RASTER && CONN == 8) {
        pl = px_at(DIR > 0 ? y0 - 1 : y1, xs0 - 1);
        pr = px_at(DIR > 0 ? y0 - 1 : y1, xs1);
    }
    // WALLS (a compile-time copy of whether a tile schedule is active): the
    // row loop is instantiated with and without walls, chosen once per wave
    // (as a per-row test the compiler turned it into per-register selects and
    // copies: 36 of the raster row's 283 VALU instructions).
    auto row_step = [&](auto wallc, int i, const uint32_t (&jw)[NW], const uint32_t (&iw)[NW],
                        uint32_t lc, uint32_t rc) {
        constexpr bool WALLS = decltype(wallc)::value;
        const int y = DIR > 0 ? y0 + i : y1 - 1 - i;
        const uint32_t le = RASTER ? pl : lc, re = RASTER ? pr : rc;
        uint32_t v[NP], c[NP], m[NP];
        L::split(jw, v);
        L::split(iw, c);
        // A raster row tests for a change once, at the end, against the
        // loaded values (every step only raises pixels, the clip only lowers
        // them); a column row tests its one step.
        uint32_t v0[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) v0[k] = v[k];
        bool d = false;
        if constexpr (CLIP) {
#pragma unroll
            for (int k = 0; k < NP; ++k) v[k] = L::vmin(v[k], c[k]);
        }
        // walls: pixels of tiles that are off (lanes past the row end have a
        // zero mask already)
        if constexpr (WALLS) {
            const uint32_t wm = row_on(y) ? 0xFFFFFFFFu : 0u;
#pragma unroll
            for (int k = 0; k < NP; ++k) c[k] &= wm;
        }
        if constexpr (CONN == 8) {
            // left neighbour of the lane's first register, right neighbour
            // of its last: the adjacent lane's pixel (outside the strip:
            // the loaded one) joined with this lane's own pixel
            uint32_t left, right;
            if constexpr (L::PAIRS) {
                left = from_left(prev[NP - 1] >> 16, le) | (prev[NP - 1] << 16);
                right = (prev[0] >> 16) | (from_right(prev[0] & 0xFFFFu, re) << 16);
            } else {
                left = from_left(prev[NP - 1], le);
                right = from_right(prev[0], re);
            }
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const uint32_t a = k > 0 ? prev[k - 1] : left;
                const uint32_t b = k < NP - 1 ? prev[k + 1] : right;
                m[k] = L::vmax(prev[k], L::vmax(a, b));
            }
        } else {
#pragma unroll
            for (int k = 0; k < NP; ++k) m[k] = prev[k];
        }
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const uint32_t out = L::vmax(v[k], L::vmin(m[k], c[k]));
            if constexpr (!RASTER) d |= out != v[k];
            prev[k] = out;
        }
        if constexpr (RASTER) {
            // SKIP_CLOSED: the scans change nothing in a row where no single
            // step can raise a pixel (every pixel at least min(its mask, each
            // neighbour along the row), the pixels beside the strip counted
            // as neighbours): the scans are such steps, in sequence.  Rows
            // that a previous pass left closed mostly stay closed after the
            // column step (about half of the upward pass's rows at 32768^2
            // after the first downward one: its VALU instructions per wave
            // 146K -> 115K), so the wave tests that first (about a fifth of
            // the scans' instructions) and scans only when some lane finds a
            // pixel that can rise.
            bool open = true;
            if constexpr (SKIP_CLOSED) {
                uint32_t left, right;
                if constexpr (L::PAIRS) {
                    left = from_left(prev[NP - 1] >> 16, lc) | (prev[NP - 1] << 16);
                    right = (prev[0] >> 16) | (from_right(prev[0] & 0xFFFFu, rc) << 16);
                } else {
                    left = from_left(prev[NP - 1], lc);
                    right = from_right(prev[0], rc);
                }
                bool rise = false;
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    const uint32_t a = k > 0 ? prev[k - 1] : left;
                    const uint32_t b = k < NP - 1 ? prev[k + 1] : right;
                    rise |= L::vmax(prev[k], L::vmin(L::vmax(a, b), c[k])) != prev[k];
                }
                open = wave_any(rise);
            }
            if (open) {
                // the row across the strip, forward then backward, entering
                // from the pixels beside the strip
                L px;
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    px.v[k] = prev[k];
                    px.c[k] = c[k];
                }
                const uint32_t qm = px.mask_min();
                uint32_t Pi, Qi;
                px.template reduce_masked<1>(Pi, Qi, qm);
                scan_lanes_fwd(Pi, Qi);
                px.template apply<1, false>(max(from_left(Pi, 0u), min(lc, from_left(Qi, 0xFFFFFFFFu))));
                px.template reduce_masked<-1>(Pi, Qi, qm);
                scan_lanes_bwd(Pi, Qi, lane);
                px.template apply<-1, false>(max(from_right(Pi, 0u), min(rc, from_right(Qi, 0xFFFFFFFFu))));
#pragma unroll
                for (int k = 0; k < NP; ++k) prev[k] = px.v[k];
            }
#pragma unroll
            for (int k = 0; k < NP; ++k) d |= prev[k] != v0[k];
        }
        any_change |= d;
        if (!INPLACE || d) {  // in place: only vectors that rose are written
            uint32_t ow[NW];
            L::join(prev, ow);
            V::store(D + (size_t)y * pd, o, rowbytes, ow);
        }
        if constexpr (SEED) {
            {  // the strip's border pixels (final) into lane y % 64
                const uint32_t f = rdlane(prev[0], 0), cf = rdlane(c[0], 0);
                const uint32_t l = rdlane(prev[NP - 1], 63), cl = rdlane(c[NP - 1], 63);
                if ((y >> 6) != eband) {
                    edge_flush();
                    eband = y >> 6;
                }
                const bool here = lane == (y & 63);
                eL = here ? (f & 0xFFFFu) | (cf << 16) : eL;
                eR = here ? (L::PAIRS ? (l >> 16) | (cl & 0xFFFF0000u) : l | (cl << 16)) : eR;
            }
            // row y against its own neighbours and row r1 = y - DIR, and
            // r1 against row y (the row before y in the block: i > 0).  A
            // raster row passes the first test by construction: its column
            // step took r1's final pixels inside the strip, and its scans
            // (or the closure test that found them idle) leave it stable
            // along the strip.
            uint32_t h[NP];
            if constexpr (CONN == 8 || !RASTER) nbmax(prev, h, false);
            bool uy = false, ur = false;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const uint32_t dn = CONN == 8 ? L::vmax(h[k], prev[k]) : prev[k];
                if constexpr (!RASTER) {
                    const uint32_t q = i > 0 ? L::vmax(h[k], v1[k]) : h[k];
                    uy |= L::vmax(L::vmin(q, c[k]), prev[k]) != prev[k];
                    v1[k] = dn;
                }
                if (i > 0) ur |= L::vmax(L::vmin(dn, c1[k]), r1[k]) != r1[k];
                r1[k] = prev[k];
                c1[k] = c[k];
            }
            if (i > 0) flag_band(y - DIR, ur);
            flag_band(y, uy);
        }
        if constexpr (RASTER) {
            pl = lc;
            pr = rc;
        }
    };
    auto run_group = [&](auto wallc, int i0, const Group& g) {
#pragma unroll
        for (int u = 0; u < CU; ++u) {
            if (i0 + u >= n) break;  // the same in every lane
            row_step(wallc, i0 + u, g.jw[u], g.iw[u], g.lc[u], g.rc[u]);
        }
    };
    auto run_block = [&](auto wallc) {
        Group ga, gb;
        load_group(0, ga);
        for (int i0 = 0; i0 < n; i0 += 2 * CU) {
            load_group(i0 + CU, gb);
            run_group(wallc, i0, ga);
            if (i0 + CU >= n) break;
            load_group(i0 + 2 * CU, ga);
            run_group(wallc, i0 + CU, gb);
        }
    };
    if (tm.active) run_block(std::true_type{});
    else run_block(std::false_type{});
    if constexpr (SEED) {
        const int yt = DIR > 0 ? y1 : y0 - 1;  // the row beyond the last one computed
        if (far_end && n > 0 && yt >= 0 && yt < H) {
            // a fix-up window's far end: row yt (not changed by this pass,
            // final) and the window's last row r1, both ways
            uint32_t w[NW], wm[NW], t[NP], ct[NP], hr[NP], ht[NP];
            V::load(S + (size_t)yt * ps, o, rowbytes, w);
            V::load(I + (size_t)yt * pi, o, rowbytes, wm);
            L::split(w, t);
            L::split(wm, ct);
            if constexpr (CONN == 8) {
                nbmax(r1, hr, true);
                nbmax(t, ht, true);
            }
            bool ut = false, ur = false;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const uint32_t fr = CONN == 8 ? hr[k] : r1[k], ft = CONN == 8 ? ht[k] : t[k];
                ut |= L::vmax(L::vmin(fr, ct[k]), t[k]) != t[k];
                ur |= L::vmax(L::vmin(ft, c1[k]), r1[k]) != r1[k];
            }
            flag_band(DIR > 0 ? y1 - 1 : y0, ur);
            flag_band(yt, ut);
        }
        if (n > 0) {
            flag_band(band >= 0 ? -64 : 0, false);  // flush
            edge_flush();
        }
    }
    if (any_change && changed) *changed = 1u;
}

// Global wave w -> (image, strip, segment).  MODE 0: read J write P (every
// pixel: the result of the blocked pass, neighbours outside a block from J);
// 1: read P write J; 2: in place on J, writing only the vectors that rose.
// In place, the rows just outside a block and the pixels just outside a
// strip belong to other waves and may be read before or after those waves
// raise them: either value is a lower bound of the reconstruction, so the
// reconstruction is unchanged (only the order of propagation is), and every
// unchanged vector saves its write.
template <class T, int VB, int CONN, int DIR, int MODE>
__global__ __launch_bounds__(kWG) void k_col_pass(Descs ds, uint32_t* changed, uint8_t* flags,
                                                  uint32_t* edges) {
    // the wave's index, in a scalar register: every address derived from it
    // (block geometry, row bases) is then computed once per wave, not per lane
    const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (w >= ds.total) return;
    const ImgDesc d = scalar_desc(desc(ds, find_cw(ds, w)));
    if (ds.round >= d.rounds) return;  // this image's sweep rounds are done
    const int local = w - d.cw_base;
    const int strip = local % d.strips, seg = local / d.strips;
    int y0 = seg * d.R, y1 = min(d.H, y0 + d.R);
    if constexpr (MODE == 5 || MODE == 7) {
        if (ds.seam > 0) {  // fix-up window after the block's top boundary (down) or before its bottom one
            const int k = min(ds.seam, d.R / 2);
            if (DIR > 0) {
                if (seg == 0) return;
                y1 = min(d.H, y0 + k);
            } else {
                if (y1 >= d.H) return;
                y0 = y1 - k;
            }
        }
    }
    uint8_t* const J = d.J;
    if constexpr (MODE == 0)
        col_block<T, VB, CONN, DIR, false>(J, d.P, d.I, d.pj, d.pp, d.pi, d.W, d.H, y0, y1, strip, d.tm, changed);
    else if constexpr (MODE == 1)
        col_block<T, VB, CONN, DIR, false>(d.P, J, d.I, d.pp, d.pj, d.pi, d.W, d.H, y0, y1, strip, d.tm, changed);
    else if constexpr (MODE == 2)
        col_block<T, VB, CONN, DIR, true>(J, J, d.I, d.pj, d.pj, d.pi, d.W, d.H, y0, y1, strip, d.tm, changed);
    else if constexpr (MODE == 3)  // in place, with the seeding flags
        col_block<T, VB, CONN, DIR, true, true>(J, J, d.I, d.pj, d.pj, d.pi, d.W, d.H, y0, y1, strip, d.tm,
                                               changed, flags, d.mp_base, d.nmx, edges + d.edge_base);
    else if constexpr (MODE == 4)  // raster, read J write P
        col_block<T, VB, CONN, DIR, false, false, true>(J, d.P, d.I, d.pj, d.pp, d.pi, d.W, d.H, y0, y1,
                                                       strip, d.tm, changed);
    else if constexpr (MODE == 5)  // raster, in place
        col_block<T, VB, CONN, DIR, true, false, true>(J, J, d.I, d.pj, d.pj, d.pi, d.W, d.H, y0, y1, strip,
                                                      d.tm, changed);
    else if constexpr (MODE == 6)  // raster, in place, clipping first
        col_block<T, VB, CONN, DIR, true, false, true, true>(J, J, d.I, d.pj, d.pj, d.pi, d.W, d.H, y0, y1,
                                                            strip, d.tm, changed);
    else  // raster, in place, with the seeding flags
        col_block<T, VB, CONN, DIR, true, true, true>(J, J, d.I, d.pj, d.pj, d.pi, d.W, d.H, y0, y1, strip,
                                                     d.tm, changed, flags, d.mp_base, d.nmx,
                                                     edges + d.edge_base, ds.seam > 0);
}

// ------------------------------------------------- megapixel (64 x 64) tile --
// A wave owns one megapixel at a time.  Its LDS tile holds rows -1..64 of the
// 72 columns -4..67 (the megapixel, its one-pixel halo, and three columns on
// each side that are loaded with the halo columns and never read); absent
// pixels are 0.  In registers every pixel is a word (marker | mask << 16); in
// LDS:
//  * uint8 images: 16-bit pixels (marker | mask << 8), row stride 148 bytes
//    = 37 words (odd: a lane-per-row word access is bank-conflict free; a
//    lane-per-column 16-bit access reads one contiguous row per step).
//    9.8 KiB per wave: four waves per SIMD fit the LDS.
//  * uint16 images: 32-bit pixels (marker | mask << 16), row stride 73 words.
// A fetching lane's 4 pixels land on word boundaries (8 or 16 bytes) and are
// written with no per-pixel test.  Every sweep is a per-lane sequential scan:
// lane = row for the row sweeps, lane = column for the column sweeps.
typedef __attribute__((address_space(3))) uint8_t lds8;
typedef __attribute__((address_space(3))) uint16_t lds16;
typedef __attribute__((address_space(3))) uint32_t lds32;

template <class T>
struct TileFmt;
template <>
struct TileFmt<uint8_t> {
    static constexpr int PXB = 2;    // bytes per pixel
    static constexpr int RB = 148;   // row stride, bytes
};
template <>
struct TileFmt<uint16_t> {
    static constexpr int PXB = 4;
    static constexpr int RB = 73 * 4;
};
template <class T>
constexpr int TILE_BYTES = 66 * TileFmt<T>::RB;
constexpr int ROSE_BYTES = 4 * 64;  // per-border "rose in this visit" flags
template <class T>
constexpr int WAVE_LDS = (TILE_BYTES<T> + ROSE_BYTES + 15) / 16 * 16;

template <class T>
constexpr int toff(int r, int c) {  // r in [-1, 64], c in [-4, 67]
    return (r + 1) * TileFmt<T>::RB + (c + 4) * TileFmt<T>::PXB;
}
// 16-bit tile pixel (marker | mask << 8) -> register word (marker | mask << 16)
DEV uint32_t widen8(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c010c00u); }

// Pixel at byte offset o of the tile as a word (marker | mask << 16).
template <class T>
DEV uint32_t tld_at(const lds8* p) {
    if constexpr (sizeof(T) == 1) return widen8(*reinterpret_cast<const lds16*>(p));
    else return *reinterpret_cast<const lds32*>(p);
}
template <class T>
DEV uint32_t tld(const lds8* s, int r, int c) { return tld_at<T>(s + toff<T>(r, c)); }
// The marker half of the pixel at byte offset o (zero-extended; one LDS load,
// no repacking), and a store of the marker half alone (the mask half of a
// tile pixel never changes).
template <class T>
DEV uint32_t tld_marker_at(const lds8* p) {
    if constexpr (sizeof(T) == 1) return *p;
    else return *reinterpret_cast<const lds16*>(p);
}
template <class T>
DEV void tst_marker_at(lds8* p, uint32_t w) {
    if constexpr (sizeof(T) == 1) *p = (uint8_t)w;
    else *reinterpret_cast<lds16*>(p) = (uint16_t)w;
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
DEV u16x2 as2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
DEV uint32_t as1(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
// w = (j | m << 16); p carries the neighbour's value in its low half.
// Returns (max(j, min(m, p)), m): packed 16-bit min/max with op_sel.
DEV uint32_t upd(uint32_t w, uint32_t p) {
    const u16x2 W = as2(w), P = as2(p);
    const u16x2 t = __builtin_elementwise_min(W.yx, P.xx);  // (min(m,p), min(j,p) <= m)
    return as1(__builtin_elementwise_max(W, t));
}
// The largest low half among p and its left and right lanes' (0 past the
// wave's ends).  Two 16-bit maxima, each with its lane move folded in as a
// DPP operand; the empty asm between them keeps the compiler from merging
// them into one three-operand maximum, which cannot take a DPP operand.
DEV uint32_t max_lr_lo(uint32_t p) {
    uint32_t t = max((uint16_t)dpp_z<DPP_WAVE_SHR1>(p), (uint16_t)p);
    asm volatile("" : "+v"(t));
    return max((uint16_t)dpp_z<DPP_WAVE_SHL1>(p), (uint16_t)t);
}
// min(J(p), I(q)) > J(q) for a pixel value p and a packed neighbour word q.
DEV bool raises(uint32_t p, uint32_t q) { return min(p, q >> 16) > (q & 0xFFFFu); }

DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
DEV uint32_t ldp(const uint8_t* base, size_t pitch, int x, int y) {
    return *gp(reinterpret_cast<const T*>(base + (size_t)y * pitch + (size_t)x * sizeof(T)));
}

// Loads the megapixel and its halo into the wave's LDS tile (absent pixels 0).
// The 66 pixels of a row with its halo lie among the 72 from x0 - 4 (x0 - 4
// is a multiple of 4 pixels, and pitches and bases of 4 bytes): 18 lanes per
// row, 4 pixels each (one 32-bit word for uint8, two for uint16), 3 rows per
// load instruction, so the halo columns come with the rows (one cache line or
// two per row) instead of one line per halo pixel.  All loads are issued
// before any is used, from clamped (always valid) addresses, in one batch
// for both pixel types (uint16 tiles hold the kernel at two waves per SIMD by
// their LDS, so the 88 registers of a uint16 batch cost no occupancy).
// A lane's 4 pixels go to LDS as 2 (uint8) or 4 (uint16) words built with
// byte permutes, masked to 0 where they lie outside the image; rows outside
// the image are written as 0 (a separate path: megapixels on the image's top
// or bottom edge).
template <class T>
DEV void tile_fetch(const Image& im, int mx, int my, lds8* s) {
    constexpr int WPL = (int)sizeof(T);                 // 32-bit words per lane and plane
    constexpr int RPI = 3, NIT = 22, NB = 1, BIT = NIT / NB;
    constexpr int RB = TileFmt<T>::RB;
    const int lane = lane_id();
    const int x0 = mx * kMP, y0 = my * kMP;
    const int rl = lane / 18, d = lane - rl * 18;  // rl == 3: idle lanes
    const int xw = x0 - 4 + 4 * d;                  // first pixel of the lane's 4
    const int lastw = (im.W * (int)sizeof(T) - 1) & ~3;  // the row's last word (bytes)
    size_t off[WPL];
#pragma unroll
    for (int w = 0; w < WPL; ++w) off[w] = (size_t)min(max(xw * (int)sizeof(T) + 4 * w, 0), lastw);
    // byte masks of the lane's tile words: pixels outside the image are 0
    uint32_t keep[2 * WPL];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool in = xw + k >= 0 && xw + k < im.W;
        if constexpr (sizeof(T) == 1) {
            if (k % 2 == 0) keep[k / 2] = in ? 0xFFFFu : 0u;
            else keep[k / 2] |= in ? 0xFFFF0000u : 0u;
        } else {
            keep[k] = in ? 0xFFFFFFFFu : 0u;
        }
    }
    const bool rows_inside = y0 > 0 && y0 + kMP < im.H;  // rows -1..64 all in the image
    lds8* a = s + rl * RB + 4 * d * TileFmt<T>::PXB;      // the lane's 4 pixels of row rl - 1
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        uint32_t gj[BIT][WPL], gi[BIT][WPL];
        if (rows_inside) {
            // a scalar base per row group and a 32-bit lane offset: no 64-bit
            // address arithmetic per load (idle lanes re-read row 2's words)
            const int y = y0 - 1 + b * BIT * RPI;
            const uint8_t* bj = im.J + (size_t)y * im.pj;
            const uint8_t* bi = im.I + (size_t)y * im.pi;
            const int rr = min(rl, RPI - 1);
            const uint32_t sj = RPI * (uint32_t)im.pj, si = RPI * (uint32_t)im.pi;
            uint32_t oj = (uint32_t)rr * (uint32_t)im.pj, oi = (uint32_t)rr * (uint32_t)im.pi;
#pragma unroll
            for (int u = 0; u < BIT; ++u) {
#pragma unroll
                for (int w = 0; w < WPL; ++w) {
                    gj[u][w] = ld_coh(bj + (oj + (uint32_t)off[w]));
                    gi[u][w] = *gp(reinterpret_cast<const uint32_t*>(bi + (oi + (uint32_t)off[w])));
                }
                oj += sj;
                oi += si;
            }
        } else {
#pragma unroll
            for (int u = 0; u < BIT; ++u) {
                const int yc = min(max(y0 - 1 + (b * BIT + u) * RPI + rl, 0), im.H - 1);
#pragma unroll
                for (int w = 0; w < WPL; ++w) {
                    gj[u][w] = ld_coh(im.J + (size_t)yc * im.pj + off[w]);
                    gi[u][w] = *gp(reinterpret_cast<const uint32_t*>(im.I + (size_t)yc * im.pi + off[w]));
                }
            }
        }
        if (rl < RPI) {
#pragma unroll
            for (int u = 0; u < BIT; ++u) {
                const int it = b * BIT + u;
                uint32_t m[2 * WPL];
                if (rows_inside) {
#pragma unroll
                    for (int k = 0; k < 2 * WPL; ++k) m[k] = keep[k];
                } else {
                    const int y = y0 - 1 + it * RPI + rl;
                    const bool rv = y >= 0 && y < im.H;
#pragma unroll
                    for (int k = 0; k < 2 * WPL; ++k) m[k] = rv ? keep[k] : 0u;
                }
                lds32* o = reinterpret_cast<lds32*>(a + it * RPI * RB);
                if constexpr (sizeof(T) == 1) {
                    // [j0 m0 j1 m1], [j2 m2 j3 m3]
                    o[0] = __builtin_amdgcn_perm(gi[u][0], gj[u][0], 0x05010400u) & m[0];
                    o[1] = __builtin_amdgcn_perm(gi[u][0], gj[u][0], 0x07030602u) & m[1];
                } else {
                    // (j | m << 16) per pixel
                    o[0] = __builtin_amdgcn_perm(gi[u][0], gj[u][0], 0x05040100u) & m[0];
                    o[1] = __builtin_amdgcn_perm(gi[u][0], gj[u][0], 0x07060302u) & m[1];
                    o[2] = __builtin_amdgcn_perm(gi[u][1], gj[u][1], 0x05040100u) & m[2];
                    o[3] = __builtin_amdgcn_perm(gi[u][1], gj[u][1], 0x07060302u) & m[3];
                }
            }
        }
    }
    wave_lds_sync();
}

// Row sweeps over the interior, lane = row: left to right, then right to
// left.  The lane's 66 pixels (halo included) are read into registers first,
// so the reads issue back to back, both dependent chains run on registers,
// and the interior goes back to LDS once, if it changed.  A lane owns its row
// in a row sweep (its column in a column sweep): no other lane touches those
// pixels.  ch[0], ch[1]: whether each direction changed a pixel.
template <class T>
DEV void sweep_rows_pair(lds8* s, int lane, bool (&ch)[2], bool last_forward) {
    lds32* row = reinterpret_cast<lds32*>(s + (lane + 1) * TileFmt<T>::RB);
    uint32_t w[66];  // columns -1..64
    if constexpr (sizeof(T) == 1) {
        // word k holds columns 2k - 4 (low) and 2k - 3 (high)
#pragma unroll
        for (int k = 1; k <= 34; ++k) {
            const uint32_t v = row[k];
            if (k > 1) w[2 * k - 3] = widen8(v);
            if (k < 34) w[2 * k - 2] = widen8(v >> 16);
        }
    } else {
#pragma unroll
        for (int c = -1; c <= kMP; ++c) w[c + 1] = row[c + 4];
    }
    ch[0] = ch[1] = false;
    uint32_t p = w[0];
#pragma unroll
    for (int c = 1; c <= kMP; ++c) {
        const uint32_t nw = upd(w[c], p);
        ch[0] |= nw != w[c];
        w[c] = nw;
        p = nw;
    }
    if (last_forward && !wave_any(ch[0])) return;  // converged: nothing to sweep or write
    p = w[65];
#pragma unroll
    for (int c = kMP; c >= 1; --c) {
        const uint32_t nw = upd(w[c], p);
        ch[1] |= nw != w[c];
        w[c] = nw;
        p = nw;
    }
    if (ch[0] || ch[1]) {
        if constexpr (sizeof(T) == 1) {  // marker bytes only
            lds8* rb = reinterpret_cast<lds8*>(row);
#pragma unroll
            for (int c = 0; c < kMP; ++c) rb[(c + 4) * 2] = (uint8_t)w[c + 1];
        } else {
#pragma unroll
            for (int c = 0; c < kMP; ++c) row[c + 4] = w[c + 1];
        }
    }
}

// Column sweeps over the interior, lane = column: top to bottom, then bottom
// to top, the same way.  8-connectivity adds the previous row's left/right
// neighbours (DPP); the halo columns' share of the left and right columns
// has been folded in once per visit (tile_raise_sides), so lanes 0 and 63
// take 0 past the wave's ends.
template <class T, int CONN>
DEV void sweep_cols_pair(lds8* s, int lane, bool (&ch)[2], bool last_forward) {
    constexpr int RB = TileFmt<T>::RB;
    lds8* col = s + (lane + 4) * TileFmt<T>::PXB;  // row -1 of the lane's column
    uint32_t w[66];
#pragma unroll
    for (int r = 0; r < 66; ++r) w[r] = tld_at<T>(col + r * RB);
    ch[0] = ch[1] = false;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const int DIR = d == 0 ? 1 : -1;
        uint32_t p = d == 0 ? w[0] : w[65];
#pragma unroll
        for (int k = 0; k < kMP; ++k) {
            const int r = DIR > 0 ? 1 + k : kMP - k;  // w index of the row updated
            const uint32_t q = CONN == 8 ? max_lr_lo(p) : p;  // upd reads only q's low half
            const uint32_t nw = upd(w[r], q);
            ch[d] |= nw != w[r];
            w[r] = nw;
            p = nw;
        }
        if (d == 0 && last_forward && !wave_any(ch[0])) return;  // converged
    }
    if (ch[0] || ch[1]) {
#pragma unroll
        for (int r = 1; r <= kMP; ++r) tst_marker_at<T>(col + r * RB, w[r]);
    }
}

// 8-connectivity: raises the left and right columns' pixels by their
// neighbours in the halo columns (rows r - 1, r, r + 1), lane = row.  The
// halo does not change during a visit and the marker only rises, so once is
// enough; the column sweeps then leave the halo columns out.  Returns whether
// a pixel rose.
template <class T>
DEV bool tile_raise_sides(lds8* s) {
    const int r = lane_id();
    bool d = false;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int c = side ? kMP - 1 : 0, h = side ? kMP : -1;  // the column, its halo column
        const uint32_t hm = max(max(tld_marker_at<T>(s + toff<T>(r - 1, h)),
                                    tld_marker_at<T>(s + toff<T>(r, h))),
                                tld_marker_at<T>(s + toff<T>(r + 1, h)));
        lds8* pp = s + toff<T>(r, c);
        const uint32_t w = tld_at<T>(pp);
        const uint32_t nw = upd(w, hm);
        if (nw != w) {
            tst_marker_at<T>(pp, nw);
            d = true;
        }
    }
    wave_lds_sync();
    return wave_any(d);
}

// Waves per SIMD the queue kernels are compiled for: what the LDS tile
// allows (uint8 tiles: four; uint16 tiles: two).
template <class T>
constexpr int kQueueWaves = sizeof(T) == 1 ? 4 : 2;

// Sweeps until the four directions in a row change nothing, or max_sweeps:
// rows forward, rows backward, columns forward, columns backward, in turn.
// Each sweep is idempotent, so three quiet sweeps after the last change mean
// every direction leaves the tile as it is.  Two sweeps share one pass over
// the registers; when two quiet sweeps precede a pair and its forward sweep
// is quiet too, the tile has converged and the backward sweep is left out.
// `changed`: some sweep raised a pixel.
template <class T, int CONN>
DEV bool tile_converge(lds8* s, int max_sweeps, bool& changed) {
    const int lane = lane_id();
    int quiet = -1;
    changed = CONN == 8 && tile_raise_sides<T>(s);
    for (int k = 0; k < max_sweeps; k += 2) {
        bool ch[2];
        if ((k & 2) == 0) sweep_rows_pair<T>(s, lane, ch, quiet == 2);
        else sweep_cols_pair<T, CONN>(s, lane, ch, quiet == 2);
        wave_lds_sync();
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            if (wave_any(ch[d])) {
                quiet = 0;
                changed = true;
            } else if (++quiet >= 3) {
                return true;
            }
        }
    }
    return false;
}

// The fetched marker values of the megapixel's border pixels, taken right
// after tile_fetch: lo/hi halves = top/bottom row at column `lane` (tb), and
// left/right column at row `lane` (lr).
struct Border {
    uint32_t tb, lr;
};
template <class T>
DEV Border tile_border(const lds8* s) {
    const int lane = lane_id();
    return Border{(tld<T>(s, 0, lane) & 0xFFFFu) | (tld<T>(s, kMP - 1, lane) << 16),
                  (tld<T>(s, lane, 0) & 0xFFFFu) | (tld<T>(s, lane, kMP - 1) << 16)};
}

template <class T>
DEV void st_px_coh(uint8_t* p, uint32_t v) {
    __hip_atomic_store(gp(reinterpret_cast<T*>(p)), (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The marker values of a store lane's PPW = 4 / sizeof(T) consecutive pixels
// of one tile row, packed as in memory (one 32-bit word); p points at the
// first pixel's tile entry (4-byte aligned).
template <class T>
DEV uint32_t tile_marker_word(const lds8* p) {
    const lds32* q = reinterpret_cast<const lds32*>(p);
    if constexpr (sizeof(T) == 1) return __builtin_amdgcn_perm(q[1], q[0], 0x06040200u);
    else return __builtin_amdgcn_perm(q[1], q[0], 0x05040100u);
}

// Write-back for a wave that is the megapixel's only writer (mutex queue, UQ):
// memory still holds what the wave fetched, so nothing is read back.  When the
// visit raised any pixel, every pixel of the megapixel is written (32-bit
// words; pixel stores for a word that runs past the image's right edge, whose
// other bytes are not the image's); otherwise nothing.  The rose flags (same
// layout as tile_store) come from the border values taken after the fetch.
// Returns without waiting for the stores.
template <class T>
DEV void tile_store_excl(const lds8* s, lds8* rose, Border ob, bool changed,
                         const Image& im, int mx, int my) {
    constexpr int PPW = 4 / (int)sizeof(T), WPR = 64 / PPW, RPI = 64 / WPR;
    constexpr int BITS = 8 * (int)sizeof(T);
    constexpr int RB = TileFmt<T>::RB;
    const int lane = lane_id();
    if (!changed) {
        reinterpret_cast<lds32*>(rose)[lane] = 0u;
        wave_lds_sync();
        return;
    }
    rose[lane] = (tld<T>(s, 0, lane) & 0xFFFFu) != (ob.tb & 0xFFFFu);
    rose[64 + lane] = (tld<T>(s, kMP - 1, lane) & 0xFFFFu) != (ob.tb >> 16);
    rose[128 + lane] = (tld<T>(s, lane, 0) & 0xFFFFu) != (ob.lr & 0xFFFFu);
    rose[192 + lane] = (tld<T>(s, lane, kMP - 1) & 0xFFFFu) != (ob.lr >> 16);
    const int x0 = mx * kMP, y0 = my * kMP;
    const int wl = lane % WPR, rl = lane / WPR;
    const int xw = x0 + wl * PPW;
    if (xw < im.W) {
        const bool whole = xw + PPW <= im.W;
        const lds8* src = s + toff<T>(rl, wl * PPW);
        uint8_t* base = im.J + (size_t)y0 * im.pj;  // scalar base, 32-bit lane offset
        const uint32_t lo = (uint32_t)rl * (uint32_t)im.pj + (uint32_t)(xw * (int)sizeof(T));
#pragma unroll
        for (int it = 0; it < kMP / RPI; ++it) {
            const int y = y0 + it * RPI + rl;
            if (y >= im.H) continue;
            const uint32_t f = tile_marker_word<T>(src + it * RPI * RB);
            uint8_t* wp = base + (size_t)(it * RPI) * im.pj + lo;
            if (whole) {
                st_coh(wp, f);
            } else {
                for (int k = 0; k < PPW && xw + k < im.W; ++k)
                    st_px_coh<T>(wp + k * sizeof(T), f >> (BITS * k));
            }
        }
    }
    wave_lds_sync();
}

// Neighbour bits of a megapixel.
enum : uint32_t {
    NB_SELF = 1u << 0, NB_T = 1u << 1, NB_B = 1u << 2, NB_L = 1u << 3, NB_R = 1u << 4,
    NB_TL = 1u << 5, NB_TR = 1u << 6, NB_BL = 1u << 7, NB_BR = 1u << 8
};

DEV int neighbour_index(int m, int nmx, uint32_t bit) {
    switch (bit) {
        case NB_T: return m - nmx;
        case NB_B: return m + nmx;
        case NB_L: return m - 1;
        case NB_R: return m + 1;
        case NB_TL: return m - nmx - 1;
        case NB_TR: return m - nmx + 1;
        case NB_BL: return m + nmx - 1;
        case NB_BR: return m + nmx + 1;
        default: return m;
    }
}

// Writes back the pixels that rose when a duplicate visitor of the same
// megapixel may write concurrently (DQ): reads the marker again and merges
// with a 32-bit compare-and-swap of the per-pixel max.  Records which border
// pixels rose (rose[0..63] top row, [64..127] bottom row, [128..191] left
// column, [192..255] right column).  Returns true when some pixel in memory
// rose.
template <class T>
DEV bool tile_store(const lds8* s, lds8* rose, const Image& im, int mx, int my) {
    constexpr int PPW = 4 / (int)sizeof(T), WPR = 64 / PPW, RPI = 64 / WPR;
    constexpr int BITS = 8 * (int)sizeof(T);
    constexpr uint32_t PM = (uint32_t)((1u << BITS) - 1u);
    const int lane = lane_id();
    const int x0 = mx * kMP, y0 = my * kMP;
    const int wl = lane % WPR, rl = lane / WPR;
    const int xw = x0 + wl * PPW;
    reinterpret_cast<lds32*>(rose)[lane] = 0u;
    wave_lds_sync();
    bool any = false;
    if (xw < im.W) {
        constexpr int NIT = kMP / RPI;
        uint32_t gv[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) {  // all re-reads in flight together
            const int yc = min(y0 + it * RPI + rl, im.H - 1);
            gv[it] = ld_coh(im.J + (size_t)yc * im.pj + (size_t)xw * sizeof(T));
        }
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int row = it * RPI + rl;
            const int y = y0 + row;
            if (y >= im.H) continue;
            uint32_t* wp = reinterpret_cast<uint32_t*>(im.J + (size_t)y * im.pj + (size_t)xw * sizeof(T));
            const uint32_t g = gv[it];
            const uint32_t tw = tile_marker_word<T>(s + toff<T>(row, wl * PPW));
            uint32_t f = g, up = 0;
#pragma unroll
            for (int k = 0; k < PPW; ++k) {
                const uint32_t v = (tw >> (BITS * k)) & PM;
                const uint32_t old = (g >> (BITS * k)) & PM;
                if (xw + k < im.W && v > old) {
                    f = (f & ~(PM << (BITS * k))) | (v << (BITS * k));
                    up |= 1u << k;
                }
            }
            if (up) {
                uint32_t cur = g;
                while (true) {
                    uint32_t want = cur;
#pragma unroll
                    for (int k = 0; k < PPW; ++k) {
                        const uint32_t a = (cur >> (BITS * k)) & PM, b = (f >> (BITS * k)) & PM;
                        if (b > a) want = (want & ~(PM << (BITS * k))) | (b << (BITS * k));
                    }
                    if (want == cur) break;
                    const uint32_t seen = a_cas(wp, cur, want);
                    if (seen == cur) break;
                    cur = seen;
                }
                any = true;
#pragma unroll
                for (int k = 0; k < PPW; ++k) {
                    if (!((up >> k) & 1u)) continue;
                    const int c = wl * PPW + k;
                    if (row == 0) rose[c] = 1;
                    if (row == kMP - 1) rose[64 + c] = 1;
                    if (c == 0) rose[128 + row] = 1;
                    if (c == kMP - 1) rose[192 + row] = 1;
                }
            }
        }
    }
    wave_lds_sync();
    return wave_any(any);
}

// Which neighbours hold a pixel q next to a pixel p of this megapixel that
// rose in this visit, with min(J(p), I(q)) > J(q) on the fetched halo.
template <class T, int CONN>
DEV uint32_t tile_notify_mask(const lds8* s, const lds8* rose, const Image& im, int mx,
                              int my) {
    const int lane = lane_id();
    const int x0 = mx * kMP, y0 = my * kMP;
    const bool hasT = y0 > 0, hasB = y0 + kMP < im.H, hasL = x0 > 0, hasR = x0 + kMP < im.W;
    uint32_t nb = 0;
    // top / bottom rows: lane = column
    bool t = false, b = false, l = false, r = false;
    if (hasT && rose[lane]) {
        const uint32_t p = tld<T>(s, 0, lane) & 0xFFFFu;
        t = raises(p, tld<T>(s, -1, lane));
        if (CONN == 8) {
            if (lane > 0) t |= raises(p, tld<T>(s, -1, lane - 1));
            if (lane < 63) t |= raises(p, tld<T>(s, -1, lane + 1));
        }
    }
    if (hasB && rose[64 + lane]) {
        const uint32_t p = tld<T>(s, 63, lane) & 0xFFFFu;
        b = raises(p, tld<T>(s, 64, lane));
        if (CONN == 8) {
            if (lane > 0) b |= raises(p, tld<T>(s, 64, lane - 1));
            if (lane < 63) b |= raises(p, tld<T>(s, 64, lane + 1));
        }
    }
    // left / right columns: lane = row
    if (hasL && rose[128 + lane]) {
        const uint32_t p = tld<T>(s, lane, 0) & 0xFFFFu;
        l = raises(p, tld<T>(s, lane, -1));
        if (CONN == 8) {
            if (lane > 0) l |= raises(p, tld<T>(s, lane - 1, -1));
            if (lane < 63) l |= raises(p, tld<T>(s, lane + 1, -1));
        }
    }
    if (hasR && rose[192 + lane]) {
        const uint32_t p = tld<T>(s, lane, 63) & 0xFFFFu;
        r = raises(p, tld<T>(s, lane, 64));
        if (CONN == 8) {
            if (lane > 0) r |= raises(p, tld<T>(s, lane - 1, 64));
            if (lane < 63) r |= raises(p, tld<T>(s, lane + 1, 64));
        }
    }
    if (wave_any(t)) nb |= NB_T;
    if (wave_any(b)) nb |= NB_B;
    if (wave_any(l)) nb |= NB_L;
    if (wave_any(r)) nb |= NB_R;
    if constexpr (CONN == 8) {
        if (hasT && hasL && rose[0] && raises(tld<T>(s, 0, 0) & 0xFFFFu, tld<T>(s, -1, -1))) nb |= NB_TL;
        if (hasT && hasR && rose[63] && raises(tld<T>(s, 0, 63) & 0xFFFFu, tld<T>(s, -1, 64))) nb |= NB_TR;
        if (hasB && hasL && rose[64] && raises(tld<T>(s, 63, 0) & 0xFFFFu, tld<T>(s, 64, -1))) nb |= NB_BL;
        if (hasB && hasR && rose[127] && raises(tld<T>(s, 63, 63) & 0xFFFFu, tld<T>(s, 64, 64)))
            nb |= NB_BR;
    }
    return nb;
}

template <class T>
DEV uint32_t ld_px_coh(const uint8_t* p) {
    return __hip_atomic_load(gp(reinterpret_cast<const T*>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// The neighbours that the fetched halo says to notify (nb, from
// tile_notify_mask), checked again against their pixels as memory holds them
// now: a neighbour visited since this visit's fetch may already be high
// enough, and notifying it then costs a visit that changes nothing (22% of
// the visits at 32768^2, 32% at 4096^2).  Leaving out such a neighbour is
// exact: only its own visits write its pixels and they only rise, so a pixel
// that is high enough in memory is high enough in the neighbour's final
// values.  Only the sides in nb are loaded (a row: one cache line; a column:
// one line per row; a corner: one pixel), with the mask from the fetched halo
// (masks do not change).  The loads complete with this visit's stores.
template <class T, int CONN>
DEV uint32_t tile_fresh_filter(const lds8* s, const lds8* rose, const Image& im, int mx, int my,
                               uint32_t nb) {
    constexpr uint32_t SIDES = NB_T | NB_B | NB_L | NB_R | NB_TL | NB_TR | NB_BL | NB_BR;
    if ((nb & SIDES) == 0u) return nb;
    const int lane = lane_id();
    const int x0 = mx * kMP, y0 = my * kMP;
    const int xc = min(x0 + lane, im.W - 1), yc = min(y0 + lane, im.H - 1);
    auto at = [&](int y, int x) { return im.J + (size_t)y * im.pj + (size_t)x * sizeof(T); };
    // (a side or corner in nb lies inside the image)
    uint32_t qt = 0, qb = 0, ql = 0, qr = 0, k0 = 0, k1 = 0, k2 = 0, k3 = 0;
    if (nb & NB_T) qt = ld_px_coh<T>(at(y0 - 1, xc));
    if (nb & NB_B) qb = ld_px_coh<T>(at(y0 + kMP, xc));
    if (nb & NB_L) ql = ld_px_coh<T>(at(yc, x0 - 1));
    if (nb & NB_R) qr = ld_px_coh<T>(at(yc, x0 + kMP));
    if constexpr (CONN == 8) {
        if (nb & NB_TL) k0 = ld_px_coh<T>(at(y0 - 1, x0 - 1));
        if (nb & NB_TR) k1 = ld_px_coh<T>(at(y0 - 1, x0 + kMP));
        if (nb & NB_BL) k2 = ld_px_coh<T>(at(y0 + kMP, x0 - 1));
        if (nb & NB_BR) k3 = ld_px_coh<T>(at(y0 + kMP, x0 + kMP));
    }
    // a neighbour pixel as a word: the fresh marker, the mask from the halo
    auto fw = [&](uint32_t j, int r, int c) { return j | (tld<T>(s, r, c) & 0xFFFF0000u); };
    // one side: lane = position along it; p the rose flag's pixel, q(r, c)
    // the halo pixel at tile coordinates (r, c); DPP moves outside branches
    auto side = [&](uint32_t bit, uint32_t q, int roff, int pr, int pc, int hr, int hc, bool along_row) {
        const uint32_t qa = from_left(q, 0u), qz = from_right(q, 0u);  // position - 1, + 1
        bool hit = false;
        if (rose[roff + lane]) {
            const uint32_t p = tld<T>(s, pr, pc) & 0xFFFFu;
            hit = raises(p, fw(q, hr, hc));
            if constexpr (CONN == 8) {
                if (lane > 0) hit |= raises(p, along_row ? fw(qa, hr, hc - 1) : fw(qa, hr - 1, hc));
                if (lane < 63) hit |= raises(p, along_row ? fw(qz, hr, hc + 1) : fw(qz, hr + 1, hc));
            }
        }
        return wave_any(hit) ? bit : 0u;
    };
    uint32_t out = nb & ~SIDES;
    if (nb & NB_T) out |= side(NB_T, qt, 0, 0, lane, -1, lane, true);
    if (nb & NB_B) out |= side(NB_B, qb, 64, kMP - 1, lane, kMP, lane, true);
    if (nb & NB_L) out |= side(NB_L, ql, 128, lane, 0, lane, -1, false);
    if (nb & NB_R) out |= side(NB_R, qr, 192, lane, kMP - 1, lane, kMP, false);
    if constexpr (CONN == 8) {
        if ((nb & NB_TL) && raises(tld<T>(s, 0, 0) & 0xFFFFu, fw(k0, -1, -1))) out |= NB_TL;
        if ((nb & NB_TR) && raises(tld<T>(s, 0, kMP - 1) & 0xFFFFu, fw(k1, -1, kMP))) out |= NB_TR;
        if ((nb & NB_BL) && raises(tld<T>(s, kMP - 1, 0) & 0xFFFFu, fw(k2, kMP, -1))) out |= NB_BL;
        if ((nb & NB_BR) && raises(tld<T>(s, kMP - 1, kMP - 1) & 0xFFFFu, fw(k3, kMP, kMP)))
            out |= NB_BR;
    }
    return out;
}

// ------------------------------------------------------------ seeding -------
// Streaming form of the seeding test: a wave sweeps one band of 64 image rows
// (one megapixel row) over a strip of 64 lanes x VB bytes, keeping the rows
// above / at / below in registers; left and right neighbours come from the
// adjacent lanes (DPP) and, at the strip edges, from memory.  Each lane ORs
// its pixels' instability; a ballot folds the lanes of one megapixel.
template <class T, int VB, int CONN>
__global__ __launch_bounds__(kWG) void k_unstable_scan(Descs ds, uint8_t* flags_all,
                                                        uint32_t* count) {
    using V = VecIO<T, VB>;
    using L = LanePx<T, VB>;
    constexpr int NW = V::NW, NPX = V::NPX, NP = L::NP;
    constexpr int LPM = kMP / NPX;          // lanes per megapixel
    constexpr int MPS = 64 / LPM;           // megapixels per strip
    const int lane = lane_id();
    // the wave's index, in a scalar register: every address derived from it
    // (block geometry, row bases) is then computed once per wave, not per lane
    const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (w >= ds.total) return;
    const ImgDesc dsc = scalar_desc(desc(ds, find_sw(ds, w)));
    const Image im = image_of(dsc);
    const int nmx = dsc.nmx;
    uint8_t* flags = flags_all + dsc.mp_base;
    const int local = w - dsc.sw_base;
    const int strip = local % dsc.strips, my = local / dsc.strips;
    const int rowbytes = im.W * (int)sizeof(T);
    const int xb0 = strip * 64 * VB;
    if (xb0 >= rowbytes) return;
    const int o = xb0 + lane * VB;
    const int xs0 = xb0 / (int)sizeof(T), xs1 = (xb0 + 64 * VB) / (int)sizeof(T);
    const int y0 = my * kMP, y1 = min(im.H, y0 + kMP);
    // Pixels as the lane's registers (16-bit pairs with 16-byte lanes, pixel k
    // and k + NP in one register).
    auto load_row = [&](const uint8_t* base, size_t pitch, int y, uint32_t (&px)[NP]) {
        uint32_t w[NW];
        if (y >= 0 && y < im.H) V::load(base + (size_t)y * pitch, o, rowbytes, w);
        else {
#pragma unroll
            for (int i = 0; i < NW; ++i) w[i] = 0;
        }
        L::split(w, px);
    };
    auto edge = [&](int y, int x) -> uint32_t {
        return (y >= 0 && y < im.H && x >= 0 && x < im.W) ? ldp<T>(im.J, im.pj, x, y) : 0u;
    };
    // Neighbour of the lane's first register to the left / last to the right:
    // the adjacent lane's pixel (outside the strip: `fill`) joined with this
    // lane's own pixel.
    auto left_of = [&](const uint32_t (&r)[NP], uint32_t fill) -> uint32_t {
        if constexpr (L::PAIRS) return from_left(r[NP - 1] >> 16, fill) | (r[NP - 1] << 16);
        else return from_left(r[NP - 1], fill);
    };
    auto right_of = [&](const uint32_t (&r)[NP], uint32_t fill) -> uint32_t {
        if constexpr (L::PAIRS) return (r[0] >> 16) | (from_right(r[0] & 0xFFFFu, fill) << 16);
        else return from_right(r[0], fill);
    };
    uint32_t up[NP], cur[NP];
    load_row(im.J, im.pj, y0 - 1, up);
    load_row(im.J, im.pj, y0, cur);
    uint32_t upL = lane == 0 ? edge(y0 - 1, xs0 - 1) : 0u, upR = lane == 63 ? edge(y0 - 1, xs1) : 0u;
    uint32_t cuL = lane == 0 ? edge(y0, xs0 - 1) : 0u, cuR = lane == 63 ? edge(y0, xs1) : 0u;
    bool un = false;
    // Groups of SU rows: the loads of the group (next marker rows, mask rows,
    // outside neighbours) are issued before any row of it is tested.
    constexpr int SU = VB == 16 ? 4 : 8;
    for (int g = y0; g < y1; g += SU) {
        uint32_t dnw[SU][NW], mkw[SU][NW], dL[SU], dR[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int yd = g + u + 1;                    // marker row below row g+u
            const int ym = min(g + u, y1 - 1);           // mask row g+u (clamped)
            const int ydc = min(yd, im.H - 1);           // in-bounds row to load
            V::load(im.J + (size_t)ydc * im.pj, o, rowbytes, dnw[u]);
            V::load(im.I + (size_t)ym * im.pi, o, rowbytes, mkw[u]);
            if (yd >= im.H) {
#pragma unroll
                for (int i = 0; i < NW; ++i) dnw[u][i] = 0;  // below the image: absent
            }
            dL[u] = lane == 0 ? edge(yd, xs0 - 1) : 0u;
            dR[u] = lane == 63 ? edge(yd, xs1) : 0u;
        }
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            if (g + u >= y1) break;  // the same in every lane
            uint32_t dn[NP], msk[NP], m[NP];
            L::split(dnw[u], dn);
            L::split(mkw[u], msk);
            const uint32_t dnL = dL[u], dnR = dR[u];
            // m = the largest neighbour: 8-connectivity, the 3 x 3 block's
            // column maxima (the pixel itself included: it cannot raise
            // itself) maxed over left, centre and right; 4-connectivity, up,
            // down, left and right.
            if constexpr (CONN == 8) {
                uint32_t vcol[NP];
#pragma unroll
                for (int k = 0; k < NP; ++k) vcol[k] = L::vmax(up[k], L::vmax(cur[k], dn[k]));
                const uint32_t vl = left_of(vcol, max(upL, max(cuL, dnL)));
                const uint32_t vr = right_of(vcol, max(upR, max(cuR, dnR)));
#pragma unroll
                for (int k = 0; k < NP; ++k)
                    m[k] = L::vmax(vcol[k], L::vmax(k > 0 ? vcol[k - 1] : vl, k < NP - 1 ? vcol[k + 1] : vr));
            } else {
                const uint32_t cl = left_of(cur, cuL), cr = right_of(cur, cuR);
#pragma unroll
                for (int k = 0; k < NP; ++k)
                    m[k] = L::vmax(L::vmax(up[k], dn[k]),
                                   L::vmax(k > 0 ? cur[k - 1] : cl, k < NP - 1 ? cur[k + 1] : cr));
            }
            // unstable where min(m, mask) > marker in some pixel
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const uint32_t t = L::vmin(m[k], msk[k]);
                un |= L::vmax(t, cur[k]) != cur[k];
            }
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                up[k] = cur[k];
                cur[k] = dn[k];
            }
            upL = cuL; upR = cuR; cuL = dnL; cuR = dnR;
        }
    }
    const uint64_t b = ballot(un);
    const int mx0 = xs0 / kMP;
    if (lane < MPS && mx0 + lane < nmx) {
        const bool f = ((b >> (lane * LPM)) & ((LPM == 64) ? ~0ull : ((1ull << LPM) - 1ull))) != 0;
        *gp(flags + (size_t)my * nmx + mx0 + lane) = f ? 1 : 0;
        if (f && count) a_add(count, 1u);
    }
}

// -------------------------------------------------------------- utilities ---
// The seeding test for the pixel pairs the seeding column pass leaves out:
// those across an edge between two strips (spx pixels wide) or a boundary
// between two blocks (R rows), on final values (after the pass).  A wave
// takes 64 rows of one strip edge (lane = row) or 64 columns of one block
// boundary (lane = column) and tests both sides.  The strip edges read the
// border columns the pass wrote to `edges` (one word per row and side,
// marker | mask << 16): in the image a pixel per row would cost a cache line.
// Waves of several images: ds.total in all, se_base the first of each image.
template <class T, int CONN>
__global__ __launch_bounds__(kWG) void k_seed_edges(Descs ds, const uint32_t* edges_all, int spx,
                                                     uint8_t* flags) {
    const int gw = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6)));
    const int lane = lane_id();
    if (gw >= ds.total) return;  // the whole wave
    const ImgDesc d = scalar_desc(desc(ds, search_se(ds, gw)));
    const Image im = image_of(d);
    const int w = gw - d.se_base, R = d.R, mp_base = d.mp_base, nmx = d.nmx;
    const int nedges = d.strips - 1, nbounds = (im.H + R - 1) / R - 1;
    const uint32_t* edges = edges_all + d.edge_base;
    const int nbands = (im.H + 63) / 64, nchunks = (im.W + 63) / 64;
    auto J = [&](int y, int x) -> uint32_t {
        return *gp(reinterpret_cast<const T*>(im.J + (size_t)y * im.pj + (size_t)x * sizeof(T)));
    };
    auto I = [&](int y, int x) -> uint32_t {
        return *gp(reinterpret_cast<const T*>(im.I + (size_t)y * im.pi + (size_t)x * sizeof(T)));
    };
    auto flag = [&](int y, int x) { flags[mp_base + (y >> 6) * nmx + (x >> 6)] = 1; };
    if (w < nedges * nbands) {  // strip edge between columns xe - 1 and xe
        const int se = w / nbands, xe = (se + 1) * spx, yb0 = (w % nbands) * 64, y = yb0 + lane;
        if (xe >= im.W) return;  // the whole wave
        const uint32_t* EL = edges + ((size_t)se * 2 + 1) * im.H;  // column xe - 1
        const uint32_t* ER = edges + ((size_t)se * 2 + 2) * im.H;  // column xe
        // each lane loads its row; the rows above and below come from the
        // neighbouring lanes (DPP), at the band's ends from one more load
        const bool in = y < im.H;
        const int yc = in ? y : im.H - 1;
        const uint32_t el = in ? *gp(EL + yc) : 0u, er = in ? *gp(ER + yc) : 0u;
        const uint32_t jl = el & 0xFFFFu, jr = er & 0xFFFFu;
        uint32_t qL = jl, qR = jr;
        if constexpr (CONN == 8) {
            uint32_t al = 0, ar = 0, bl = 0, br = 0;  // the rows just above / below the band
            if (lane == 0 && yb0 > 0) { al = *gp(EL + yb0 - 1) & 0xFFFFu; ar = *gp(ER + yb0 - 1) & 0xFFFFu; }
            if (lane == 63 && yb0 + 64 < im.H) { bl = *gp(EL + yb0 + 64) & 0xFFFFu; br = *gp(ER + yb0 + 64) & 0xFFFFu; }
            qL = max(qL, max(from_left(jl, al), from_right(jl, bl)));
            qR = max(qR, max(from_left(jr, ar), from_right(jr, br)));
        }
        if (!in) return;
        if (min(qL, er >> 16) > jr) flag(y, xe);
        if (min(qR, el >> 16) > jl) flag(y, xe - 1);
        return;
    }
    const int wb = w - nedges * nbands;
    if (wb >= nbounds * nchunks) return;
    const int yb = (wb / nchunks + 1) * R, x = (wb % nchunks) * 64 + lane;  // rows yb - 1 | yb
    if (yb >= im.H || x >= im.W) return;
    uint32_t qT = 0, qB = 0;
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
        if (CONN == 4 && dx != 0) continue;
        const int xx = x + dx;
        if (xx < 0 || xx >= im.W) continue;
        qT = max(qT, J(yb - 1, xx));
        qB = max(qB, J(yb, xx));
    }
    if (min(qT, I(yb, x)) > J(yb, x)) flag(yb, x);
    if (min(qB, I(yb - 1, x)) > J(yb - 1, x)) flag(yb - 1, x);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// Before a queue phase, in one launch: the control block and the megapixel
// state words (and the check words, when on) cleared, and every slot of the
// mutex ring set to expect its push ticket (ring == nullptr: DQ / UQ).  Short
// launches in a row are spaced by the host's launch rate, not by their work.
__global__ void k_queue_reset(uint32_t* ctl, size_t nctl, uint32_t* st, size_t nmp, uint32_t* chk,
                              unsigned long long* ring, size_t cap) {
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = i0; i < nctl; i += step) ctl[i] = 0u;
    for (size_t i = i0; i < nmp; i += step) st[i] = 0u;
    if (chk)
        for (size_t i = i0; i < nmp; i += step) chk[i] = 0u;
    if (ring)
        for (size_t i = i0; i < cap; i += step) ring[i] = (unsigned long long)i << 32;
}

// The queue's starting counters, once every seeded entry is reserved: the
// read-queue size / pending count / mutex tickets from the tail, clamped to
// the capacity (an overflow is flagged and the host re-seeds).
DEV void prepare_queue(const QueueArgs& q) {
    const uint32_t n = a_load(ctl_at(q.ctl, C_TAIL));
    const uint32_t k = n < q.cap ? n : q.cap;
    a_store(ctl_at(q.ctl, C_RSIZE), k);
    a_store(ctl_at(q.ctl, C_PENDING), k);
    if (n > q.cap) a_store(ctl_at(q.ctl, C_OVERFLOW), 1u);
    a_store(ctl_at(q.ctl, C_TAIL), k);
    a_store(ctl_at(q.ctl, C_SEEDED), k);  // mutex: seeded tickets are claimed apart,
    a_store(ctl_at(q.ctl, C_HEAD), k);    // single pops start after them
}

// Compacts the unstable flags into queue A and marks each queued megapixel
// in st:
//   mutex: st = PENDING, ring slot sequence = index + 1, pending += count
//   UQ:    st = 1 (already in the next read queue)
// A wave takes K consecutive 64-flag chunks (ballot + popcount per chunk) and
// reserves their queue entries with one atomicAdd on the tail: atomics on one
// address serialise, one per chunk cost 40 us at 262144 megapixels.  Entries
// keep the chunks' order inside a wave.  The last workgroup to finish sets the
// queue's starting counters (prepare_queue): each workgroup counts itself
// done (C_SEEDDONE, cleared by k_queue_reset) after its tail reservations have
// returned, so the last one's read of the tail follows all of them.
template <int K>
__global__ __launch_bounds__(kWG) void k_seed_queue(const uint8_t* flags, int nmp, int kind,
                                                     QueueArgs q) {
    const int lane = lane_id();
    for (int base = ((blockIdx.x * kWG + (int)threadIdx.x) >> 6) * 64 * K; base < nmp;
         base += gridDim.x * kWG * K) {
        int m[K];
        uint64_t b[K];
        uint32_t total = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int pos_in = base + k * 64 + lane;
            m[k] = pos_in < nmp ? pos_in : nmp;
            if (q.order_tiles && pos_in < nmp) {  // tile-ordered seeding
                int lo = 0, hi = q.order_nt - 1;  // last rank with start <= pos
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if ((int)q.order_start[mid] <= pos_in) lo = mid;
                    else hi = mid - 1;
                }
                const int t = (int)q.order_tiles[lo];
                const int local = pos_in - (int)q.order_start[lo];
                const int tx = t % q.order_ntx, ty = t / q.order_ntx;
                const int wmp = min(q.order_mpt, q.order_nmx - tx * q.order_mpt);
                m[k] = (ty * q.order_mpt + local / wmp) * q.order_nmx + tx * q.order_mpt +
                       local % wmp;
            }
            b[k] = ballot(m[k] < nmp && flags[m[k]] != 0);
            total += (uint32_t)__popcll(b[k]);
        }
        if (total == 0) continue;
        uint32_t start = 0;
        if (lane == 0) start = a_add(ctl_at(q.ctl, C_TAIL), total);
        start = first_lane(start);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t off = (uint32_t)__popcll(b[k] & ((1ull << lane) - 1ull));
            if ((b[k] >> lane) & 1ull) {
                const uint32_t pos = start + off;
                if (pos < q.cap) {
                    if (kind != Q_MUTEX) q.qa[pos] = (uint32_t)m[k];
                    if (kind == Q_MUTEX) {
                        q.st[m[k]] = ST_PENDING;
                        q.ring[pos] = ((unsigned long long)(pos + 1) << 32) | (uint32_t)m[k];
                    } else if (kind == Q_UQ) {
                        q.st[m[k]] = 1u;
                    }
                }
            }
            start += (uint32_t)__popcll(b[k]);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && a_add(ctl_at(q.ctl, C_SEEDDONE), 1u) == gridDim.x - 1u) prepare_queue(q);
}

// ------------------------------------------------------------ phase timers ---
// Per-wave accumulators of wall-clock ticks, added to the control block when
// the wave leaves (only when QueueArgs::trace is set).
struct PhaseTimer {
    unsigned long long acc[5] = {0, 0, 0, 0, 0};
    unsigned long long t = 0;
    bool on;
    __device__ explicit PhaseTimer(bool enabled) : on(enabled) {
        if (on) t = wall_clock64();
    }
    __device__ void lap(int k) {
        if (!on) return;
        const unsigned long long n = wall_clock64();
        acc[k] += n - t;
        t = n;
    }
    __device__ void flush(uint32_t* ctl) {
        if (!on || lane_id() != 0) return;
        for (int k = 0; k < 5; ++k)
            __hip_atomic_fetch_add(gp(reinterpret_cast<unsigned long long*>(ctl_at(ctl, C_PH_POP + k))),
                                   (unsigned long long)acc[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// ----------------------------------------------------------- timeouts -------
DEV unsigned long long now_ticks() { return wall_clock64(); }

// Bounded wait helper: returns false when the limit passed (and raises the flag).
struct Waiter {
    unsigned long long t0 = 0, limit;
    int spins = 0;
    __device__ explicit Waiter(unsigned long long lim) : limit(lim) {}
    __device__ bool step(uint32_t* ctl) {
        if (spins == 0) t0 = now_ticks();
        ++spins;
        __builtin_amdgcn_s_sleep(2);
        if ((spins & 63) == 0 && now_ticks() - t0 > limit) {
            a_store(ctl_at(ctl, C_TIMEOUT), 1u);
            return false;
        }
        return true;
    }
};

// ------------------------------------------- workgroup-local queue (LDS) ----
// A ring of LQ_CAP megapixel entries in LDS, shared by the waves of one
// workgroup of the mutex kernel.  A wave puts the neighbours its visit made
// pending here (while fewer than LQ_SOFT are queued) instead of in the global
// ring, and takes work from here before the global ring: the hand-off costs
// LDS atomics instead of two round trips to the device's shared counters and
// ring, and the next visit lands next to the last one.  The megapixel state
// protocol is unchanged (an entry here is a PENDING megapixel and counts in
// `pending`), so where an entry waits does not matter for correctness.
// Tickets with sequence numbers as in the global ring: slot t % LQ_CAP holds
// ((t + 1) << 32 | m) once ticket t is filled and (t + LQ_CAP) << 32 when it is
// free for ticket t + LQ_CAP.  `count` = entries reserved by pushes and not yet
// claimed by pops (at most LQ_CAP), so a push never waits on a full ring, only
// (briefly) on the pop that is freeing its slot.
constexpr int LQ_CAP = 16;
constexpr int LQ_SOFT = 8;
struct LocalQueue {
    uint32_t count, head, tail, pad;
    unsigned long long slot[LQ_CAP];
};
typedef __attribute__((address_space(3))) LocalQueue lds_lq;
typedef __attribute__((address_space(3))) unsigned long long lds64;

DEV uint32_t l_load(const lds32* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
DEV uint32_t l_add(lds32* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
DEV unsigned long long l_load64(const lds64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
DEV void l_store64(lds64* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Called by every wave of the workgroup before any uses the queue.
DEV void lq_init(lds_lq* L) {
    const int t = (int)threadIdx.x;
    if (t < LQ_CAP) l_store64(&L->slot[t], (unsigned long long)t << 32);
    if (t == 0) {
        __hip_atomic_store(&L->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_store(&L->head, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_store(&L->tail, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
}

// Takes one entry; false when none is reserved for taking.  The wait for the
// slot ends when its pusher, which has reserved it and runs meanwhile, writes
// it; it is bounded like every other wait.
DEV bool lq_pop(lds_lq* L, const QueueArgs& q, uint32_t& out) {
    uint32_t got = 0, ok = 0;
    if (lane_id() == 0 && l_load(&L->count) != 0u) {
        const uint32_t c = l_add(&L->count, 0u - 1u);  // value before
        if ((int)c <= 0) {
            l_add(&L->count, 1u);
        } else {
            const uint32_t h = l_add(&L->head, 1u);
            lds64* sl = &L->slot[h % LQ_CAP];
            Waiter w(q.timeout_ticks);
            unsigned long long e;
            while ((uint32_t)((e = l_load64(sl)) >> 32) != h + 1u) {
                if (!w.step(q.ctl)) break;
            }
            if ((uint32_t)(e >> 32) == h + 1u) {
                got = (uint32_t)e;
                ok = 1;
                l_store64(sl, (unsigned long long)(h + LQ_CAP) << 32);
            }
        }
    }
    ok = first_lane(ok);
    if (!ok) return false;
    out = first_lane(got);
    return true;
}

// Lanes with mine = true each put `value` into the local queue, as many as
// it takes (up to LQ_SOFT queued); returns which lanes it took.
DEV bool lq_push_some(lds_lq* L, const QueueArgs& q, bool mine, uint32_t value) {
    const int lane = lane_id();
    const uint64_t b = ballot(mine);
    const uint32_t n = (uint32_t)__popcll(b);
    if (n == 0) return false;
    uint32_t k = 0, t0 = 0;
    if (lane == 0) {
        const uint32_t c = l_load(&L->count);
        k = (int)c >= LQ_SOFT ? 0u : min(n, (uint32_t)LQ_SOFT - c);
        if (k) {
            const uint32_t r = l_add(&L->count, k);
            if (r + k > (uint32_t)LQ_CAP) {
                l_add(&L->count, 0u - k);
                k = 0;
            } else {
                t0 = l_add(&L->tail, k);
            }
        }
    }
    k = first_lane(k);
    t0 = first_lane(t0);
    const uint32_t rank = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
    const bool local = mine && rank < k;
    if (local) {
        const uint32_t t = t0 + rank;
        lds64* sl = &L->slot[t % LQ_CAP];
        Waiter w(q.timeout_ticks);
        bool free = true;
        while ((uint32_t)(l_load64(sl) >> 32) != t) {
            if (!w.step(q.ctl)) { free = false; break; }
        }
        if (free) l_store64(sl, ((unsigned long long)(t + 1u) << 32) | value);
    }
    return local;
}

// ------------------------------------------------- mutex variant (async) ----
// One global ring queue consumed by a persistent kernel.  Every megapixel is
// in a queue (the ring or one workgroup's local queue) at most once and is
// never visited by two waves at once:
//   notify(n):  old = st[n] |= PENDING; if old == 0, pending++ and push n
//   pop(m):     st[m] = RUNNING                        (PENDING cleared)
//   finish(m):  old = st[m] &= ~RUNNING; if old has PENDING, push m again,
//               else pending--
// A wave leaves when the ring is empty and pending == 0 (nothing queued, nothing
// running, so nothing can be pushed any more).
DEV bool ring_push_many(const QueueArgs& q, bool mine, uint32_t value) {
    // Wave-aggregated push: lanes with mine=true each append one value.
    const int lane = lane_id();
    const uint64_t b = ballot(mine);
    const uint32_t n = (uint32_t)__popcll(b);
    if (n == 0) return true;
    uint32_t start = 0;
    if (lane == 0) start = a_add(ctl_at(q.ctl, C_TAIL), n);
    start = first_lane(start);
    bool ok = true;
    if (mine) {
        const uint32_t t = start + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
        const uint32_t slot = t & (q.cap - 1u);
        Waiter w(q.timeout_ticks);
        while ((uint32_t)(a_load64(&q.ring[slot]) >> 32) != t) {
            if (!w.step(q.ctl)) { ok = false; break; }
        }
        if (ok) a_store64(&q.ring[slot], ((unsigned long long)(t + 1u) << 32) | value);
    }
    return !wave_any(!ok);
}

// Ticket value that tells a waiting wave the queue phase is over.
constexpr uint32_t RING_DONE = 0xFFFFFFFFu;

// Called by the wave whose decrement brought `pending` to zero: nothing is
// queued or running, so nothing will be pushed again and every ticket below
// the tail has been consumed.  A waiting wave holds one ticket at or after the
// tail and takes at most one more, so the tickets [tail, tail + nwaves) (all
// free: the ring holds more than nmp >= nwaves entries) are every ticket any
// wave can still wait on; each gets RING_DONE.  Waiting waves then poll only
// their own slot.
DEV void ring_close(const QueueArgs& q, uint32_t nwaves) {
    const int lane = lane_id();
    uint32_t tail = 0;
    if (lane == 0) tail = a_load(ctl_at(q.ctl, C_TAIL));
    tail = first_lane(tail);
    for (uint32_t i = (uint32_t)lane; i < nwaves; i += 64) {
        const uint32_t t = tail + i;
        a_store64(&q.ring[t & (q.cap - 1u)], ((unsigned long long)(t + 1u) << 32) | RING_DONE);
    }
}

// Takes `owed` finished megapixels off `pending`; the wave that reaches zero
// closes the ring.
DEV void pending_settle(const QueueArgs& q, uint32_t& owed, uint32_t nwaves) {
    if (!owed) return;
    uint32_t before = 0;
    if (lane_id() == 0) before = a_add(ctl_at(q.ctl, C_PENDING), 0u - owed);
    if (first_lane(before) == owed) ring_close(q, nwaves);
    owed = 0;
}

// Taking work, in this order: (1) the seeded tickets 0..S-1, filled before
// the kernel starts and claimed in batches of B through their own counter
// (every one is filled, nothing waits; slots freed at once) -- one atomic on a
// shared counter per B visits instead of one per visit; (2) the workgroup's
// local queue; (3) one ticket of the global ring from C_HEAD.  A wave holds at
// most one ring ticket: while it waits for that ticket's entry it also
// watches its local queue, takes local work when some arrives, and goes back
// to the ticket afterwards (the ticket's entry, once pushed, waits for it).
// A wave's decrements of `pending` for finished megapixels are held back
// while it still has work in hand and flushed before it waits or leaves: the
// count is never below the true number of queued and running megapixels, so
// no wave leaves early.
//
// A visit's memory round trips overlap where the protocol allows: the next
// megapixel m2 (claimed or local) is taken (st[m2] = RUNNING) while this
// visit's stores drain; after one wait, the notifications of this visit and
// the release of m go out as one round of atomics, and m2's tile loads are
// issued behind them.  m2 is not notified by this visit: its fetch comes after
// this visit's stores are complete.  Taking m2 before m is released only holds
// two megapixels at once; a notification that reaches m2 while it is taken
// queues it again, as for any running megapixel.
template <class T, int CONN, bool ONE>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(kQueueWaves<T>)))
void k_queue_mutex(Descs ds, QueueArgs q) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int lane = lane_id();
    lds_lq* L = (lds_lq*)(lds_raw + (kWG / 64) * WAVE_LDS<T>);
    lds8* s = (lds8*)lds_raw + (threadIdx.x >> 6) * WAVE_LDS<T>;
    lds8* rose = s + TILE_BYTES<T>;
    unsigned long long visits = 0;
    PhaseTimer pt(q.trace != 0);
    const uint32_t seeded = a_load(ctl_at(q.ctl, C_SEEDED));
    const uint32_t nwaves = gridDim.x * (kWG / 64);
    // Batch: the seeded region shared out evenly, one batch per wave, 1..16.
    // Consecutive tickets are neighbouring megapixels: one wave visiting them
    // in turn fetches halos its previous visit already raised, so larger
    // batches also cut re-visits.  Measured at 32768^2 after one raster round
    // (33K seeded megapixels, 4096 waves): a third of the share per batch
    // 0.77 ms (72K visits), the share 0.64-0.67 ms (65K), twice the share
    // 0.70-0.76 ms (too few waves get work); 16384^2 u16 0.275 / 0.281 /
    // 0.349 ms.
    const uint32_t B = max(1u, min(16u, seeded / nwaves));
    if (seeded == 0) return;  // `pending` starts at zero: nothing to do
    lq_init(L);
    uint32_t claimed = 0;  // lane k < nclaimed holds the k-th claimed megapixel
    int nclaimed = 0, next = 0;
    bool seeds_left = true;
    uint32_t owed = 0;     // finished megapixels not yet taken off `pending`
    // Next claimed megapixel, if any (claims a new batch when this one is used
    // up).  Wave w's first batch is tickets [wB, wB + B), taken without an
    // atomic (thousands of waves starting on one counter queue up); later
    // batches come from the counter, after the first nwaves * B tickets.
    bool first_claim = true;
    auto take_claimed = [&](uint32_t& out) -> bool {
        if (next >= nclaimed && seeds_left) {
            uint32_t c;
            if (first_claim) {
                c = (blockIdx.x * (kWG / 64) + first_lane(threadIdx.x >> 6)) * B;
                seeds_left = nwaves * B < seeded;  // tickets left for the counter
                first_claim = false;
            } else {
                uint32_t cc = 0;
                if (lane == 0) cc = a_add(ctl_at(q.ctl, C_SCLAIM), B);
                c = nwaves * B + first_lane(cc);
                seeds_left = c + B < seeded;
            }
            nclaimed = c >= seeded ? 0 : (int)min(B, seeded - c);
            next = 0;
            claimed = 0;
            if (lane < nclaimed) {
                const uint32_t t = c + (uint32_t)lane;
                const uint32_t slot = t & (q.cap - 1u);
                claimed = (uint32_t)a_load64(&q.ring[slot]);  // filled before the launch
                a_store64(&q.ring[slot], (unsigned long long)(t + q.cap) << 32);
            }
        }
        if (next >= nclaimed) return false;
        out = rdlane(claimed, next++);
        return true;
    };
    bool held = false;  // a ring ticket h taken and not yet served
    uint32_t h = 0;
    uint32_t m = 0;
    bool have = false;  // m is taken and its tile is in LDS
    while (true) {
        if (!have) {
            int state = 0;  // 1 = got an entry, 2 = done, 3 = timeout
            if (take_claimed(m) || lq_pop(L, q, m)) {
                state = 1;
            } else {
                pending_settle(q, owed, nwaves);
                while (true) {
                    if (!held) {
                        uint32_t hh = 0;
                        if (lane == 0) hh = a_add(ctl_at(q.ctl, C_HEAD), 1u);
                        h = first_lane(hh);
                        held = true;
                    }
                    const uint32_t slot = h & (q.cap - 1u);
                    uint32_t got = 0;
                    if (lane == 0) {
                        // The wave polls its own ring slot, which gets an
                        // entry or RING_DONE, and its workgroup's local queue
                        // (4: local work arrived).  Reads of the shared
                        // counters, which every working wave updates, queue up
                        // with those updates and slow the waves that still
                        // work: the timeout flag (and `pending`, a second way
                        // out) are read on every 256th poll only.
                        Waiter w(q.timeout_ticks);
                        for (uint32_t k = 0;; ++k) {
                            const unsigned long long e = a_load64(&q.ring[slot]);
                            if ((uint32_t)(e >> 32) == h + 1u) {
                                got = (uint32_t)e;
                                state = got == RING_DONE ? 2 : 1;
                                break;
                            }
                            if (l_load(&L->count) != 0u) { state = 4; break; }
                            if ((k & 255u) == 255u) {
                                if (a_load(ctl_at(q.ctl, C_TIMEOUT)) != 0u) { state = 3; break; }
                                if (a_load(ctl_at(q.ctl, C_PENDING)) == 0u) { state = 2; break; }
                            }
                            if (!w.step(q.ctl)) { state = 3; break; }
                        }
                        if (state == 1) a_store64(&q.ring[slot], (unsigned long long)(h + q.cap) << 32);
                    }
                    state = (int)first_lane((uint32_t)state);
                    if (state == 1) {
                        held = false;
                        m = first_lane(got);
                        break;
                    }
                    if (state == 4) {
                        if (lq_pop(L, q, m)) { state = 1; break; }
                        continue;  // taken by a sibling: back to the ticket
                    }
                    break;
                }
                if (state != 1) break;
            }
            if (lane == 0) {
                a_xchg(&q.st[m], ST_RUNNING);
                if (q.chk && a_add(&q.chk[m], 1u) != 0u) a_add(ctl_at(q.ctl, C_REPEAT), 1u);
            }
            mem_wait();
            pt.lap(0);
            const ImgDesc d0 = mp_desc<ONE>(ds, m);
            const int ml0 = (int)m - d0.mp_base;
            tile_fetch<T>(image_of(d0), ml0 % d0.nmx, ml0 / d0.nmx, s);
            pt.lap(1);
        }
        const ImgDesc dsc = mp_desc<ONE>(ds, m);
        const Image im = image_of(dsc);
        const int ml = (int)m - dsc.mp_base;
        const int mx = ml % dsc.nmx, my = ml / dsc.nmx;
        const Border ob = tile_border<T>(s);
        bool changed;
        const bool conv = tile_converge<T, CONN>(s, 4 * q.max_iters, changed);
        pt.lap(2);
        tile_store_excl<T>(s, rose, ob, changed, im, mx, my);
        uint32_t m2 = 0;
        const bool have2 = take_claimed(m2) || lq_pop(L, q, m2);
        if (have2 && lane == 0) {
            a_xchg(&q.st[m2], ST_RUNNING);
            if (q.chk && a_add(&q.chk[m2], 1u) != 0u) a_add(ctl_at(q.ctl, C_REPEAT), 1u);
        }
        if (q.chk && lane == 0) a_add(&q.chk[m], 0u - 1u);  // before m's release below
        uint32_t nb = tile_notify_mask<T, CONN>(s, rose, im, mx, my);
        nb = tile_fresh_filter<T, CONN>(s, rose, im, mx, my, nb);
        if (!conv) nb |= NB_SELF;
        mem_wait();  // this visit's stores are visible and m2 is taken
        pt.lap(3);
        ++visits;
        // notify neighbours (lane b handles bit b) and release m (lane 9)
        bool tried = false;
        uint32_t target = 0, old = 0;
        if (lane < 9 && ((nb >> lane) & 1u)) {
            target = (uint32_t)(dsc.mp_base + neighbour_index(ml, dsc.nmx, 1u << lane));
            tried = !(have2 && target == m2);
            if (tried) old = a_or(&q.st[target], ST_PENDING);
        }
        if (lane == 9) old = a_and(&q.st[m], ~ST_RUNNING);
        if (have2) {  // m2's loads go out behind the atomics
            const ImgDesc d2 = mp_desc<ONE>(ds, m2);
            const int ml2 = (int)m2 - d2.mp_base;
            tile_fetch<T>(image_of(d2), ml2 % d2.nmx, ml2 / d2.nmx, s);
        }
        const bool again = (rdlane(old, 9) & ST_PENDING) != 0u;
        // newly pending neighbours, and m itself when it was notified while
        // it ran: to the local queue first, the rest to the global ring
        const bool push = (tried && old == 0u) || (lane == 9 && again);
        if (lane == 9) target = m;
        const uint32_t npush = (uint32_t)__popcll(ballot(tried && old == 0u));
        if (npush && lane == 0) a_add(ctl_at(q.ctl, C_PENDING), npush);
        const bool local = lq_push_some(L, q, push, target);
        const bool ok = ring_push_many(q, push && !local, target);
        if (!again) ++owed;
        pt.lap(4);
        if (!ok) break;
        m = m2;
        have = have2;
    }
    pending_settle(q, owed, nwaves);
    pt.lap(0);
    pt.flush(q.ctl);
    if (lane == 0 && visits) {
        __hip_atomic_fetch_add(gp(reinterpret_cast<unsigned long long*>(ctl_at(q.ctl, C_VISITS_LO))),
                               (unsigned long long)visits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ------------------------------------------------ DQ / UQ (BSP rounds) ------
// Persistent cooperative kernel.  Round k consumes the read queue (claimed
// in batches), appends to the write queue (through a register buffer), and
// ends with one grid barrier that also flips the queues.
// Start gate of the round-based kernels: every block of the grid must be
// running at once (the rounds end in grid barriers).  A cooperative launch
// guarantees that; the ordinary launch used when the runtime refuses one
// does not.  Blocks that find the others within start_ticks go on; otherwise
// the gate raises C_NOTRES and every block leaves at once (late blocks see
// the flag on arrival), and the host relaunches with a smaller grid -- a few
// milliseconds lost instead of a barrier timing out.
DEV bool start_gate(const QueueArgs& q, uint32_t nblocks) {
    __shared__ int ok_s;
    if (threadIdx.x == 0) {
        ok_s = 0;
        if (a_load(ctl_at(q.ctl, C_NOTRES)) == 0u) {
            a_add(ctl_at(q.ctl, C_START), 1u);
            const unsigned long long t0 = now_ticks();
            for (uint32_t k = 0;; ++k) {
                if (a_load(ctl_at(q.ctl, C_START)) >= nblocks) { ok_s = 1; break; }
                if (a_load(ctl_at(q.ctl, C_NOTRES)) != 0u) break;
                __builtin_amdgcn_s_sleep(2);
                if ((k & 15u) == 15u && now_ticks() - t0 > q.start_ticks) {
                    a_store(ctl_at(q.ctl, C_NOTRES), 1u);
                    break;
                }
            }
        }
    }
    __syncthreads();
    return ok_s != 0;
}

// The end of a round: a grid barrier whose last arriving block also flips the
// queues (the write queue becomes the next read queue) before it releases the
// others, so a round costs one barrier.  Every append of the round is done
// before its block arrives.
DEV bool round_barrier(const QueueArgs& q, uint32_t nblocks, uint32_t round) {
    mem_wait();
    __syncthreads();
    __shared__ int ok_s;
    if (threadIdx.x == 0) {
        ok_s = 1;
        const uint32_t gen = a_load(ctl_at(q.ctl, C_BAR_GEN));
        const uint32_t arrived = a_add(ctl_at(q.ctl, C_BAR_COUNT), 1u);
        if (arrived == nblocks - 1u) {
            uint32_t ws = a_load(ctl_at(q.ctl, C_WSIZE));
            if (ws > q.cap) ws = q.cap;
            a_store(ctl_at(q.ctl, C_RSIZE), ws);
            a_store(ctl_at(q.ctl, C_WSIZE), 0u);
            a_store(ctl_at(q.ctl, C_CURSOR), 0u);
            a_store(ctl_at(q.ctl, C_ROUND), round + 1u);
            a_store(ctl_at(q.ctl, C_BAR_COUNT), 0u);
            mem_wait();
            a_store(ctl_at(q.ctl, C_BAR_GEN), gen + 1u);
        } else {
            Waiter w(q.timeout_ticks);
            while (a_load(ctl_at(q.ctl, C_BAR_GEN)) == gen) {
                if (!w.step(q.ctl)) { ok_s = 0; break; }
            }
        }
    }
    __syncthreads();
    mem_wait();
    return ok_s != 0;
}

template <class T, int CONN, int KIND, bool ONE>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(kQueueWaves<T>)))
void k_queue_bsp(Descs ds, QueueArgs q) {
    constexpr bool EXCL = KIND == Q_UQ;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const int lane = lane_id();
    lds8* s = (lds8*)lds_raw + (threadIdx.x >> 6) * WAVE_LDS<T>;
    lds8* rose = s + TILE_BYTES<T>;
    unsigned long long visits = 0;
    if (!start_gate(q, gridDim.x)) return;
    PhaseTimer pt(q.trace != 0);
    const uint32_t nwaves = gridDim.x * (kWG / 64);
    while (true) {
        const uint32_t round = a_load(ctl_at(q.ctl, C_ROUND));
        const uint32_t nr = a_load(ctl_at(q.ctl, C_RSIZE));
        uint32_t* rq = (round & 1u) ? q.qb : q.qa;
        uint32_t* wq = (round & 1u) ? q.qa : q.qb;
        // Read-queue entries are claimed B at a time (about three batches per
        // wave, 1..16): wave w's first batch is [wB, wB + B), later ones come
        // from C_CURSOR after the first nwaves * B -- one atomic on the shared
        // cursor per B visits.  Appends to the write queue collect in a
        // register buffer (lane i holds entry i) and go out 64 at a time with
        // one atomic on C_WSIZE and one coalesced store (the reference's block
        // write buffer, IWPP MP DQ/src/MorphologicalReconstruction.cu:555-589),
        // and before the round's barrier.
        const uint32_t B = max(1u, min(16u, nr / (3u * nwaves)));
        uint32_t claimed = 0, wbuf = 0;
        int nclaimed = 0, next = 0, wcount = 0;
        bool first_claim = true;
        auto flush = [&]() {
            if (wcount == 0) return;
            uint32_t start = 0;
            if (lane == 0) start = a_add(ctl_at(q.ctl, C_WSIZE), (uint32_t)wcount);
            start = first_lane(start);
            if (lane < wcount) {
                const uint32_t pos = start + (uint32_t)lane;
                if (pos < q.cap) st_coh(&wq[pos], wbuf);
                else a_store(ctl_at(q.ctl, C_OVERFLOW), 1u);
            }
            wcount = 0;
        };
        // Next claimed megapixel (claims a new batch when this one is used up).
        auto take = [&](uint32_t& out) -> bool {
            if (next >= nclaimed) {
                uint32_t c;
                if (first_claim) {
                    c = (blockIdx.x * (kWG / 64) + first_lane(threadIdx.x >> 6)) * B;
                    first_claim = false;
                } else {
                    uint32_t cc = 0;
                    if (lane == 0) cc = a_add(ctl_at(q.ctl, C_CURSOR), B);
                    c = nwaves * B + first_lane(cc);
                }
                if (c >= nr) {
                    nclaimed = next = 0;
                    return false;
                }
                nclaimed = (int)min(B, nr - c);
                next = 0;
                claimed = lane < nclaimed ? ld_coh(&rq[c + (uint32_t)lane]) : 0u;
            }
            out = rdlane(claimed, next++);
            if (lane == 0) {
                if (EXCL) a_xchg(&q.st[out], 0u);  // may be queued again for the next round
                if (q.chk && a_xchg(&q.chk[out], round + 1u) == round + 1u)
                    a_add(ctl_at(q.ctl, C_REPEAT), 1u);
            }
            return true;
        };
        bool more = true;
        uint32_t m = 0;
        bool have = false;  // m's tile is in LDS
        while (true) {
            if (!have) {
                if (!more || !take(m)) break;
                mem_wait();
                pt.lap(0);
                const ImgDesc d0 = mp_desc<ONE>(ds, m);
                const int ml0 = (int)m - d0.mp_base;
                tile_fetch<T>(image_of(d0), ml0 % d0.nmx, ml0 / d0.nmx, s);
                pt.lap(1);
            }
            const ImgDesc dsc = mp_desc<ONE>(ds, m);
            const Image im = image_of(dsc);
            const int ml = (int)m - dsc.mp_base;
            const int mx = ml % dsc.nmx, my = ml / dsc.nmx;
            const Border ob = tile_border<T>(s);
            bool changed;
            const bool conv = tile_converge<T, CONN>(s, 4 * q.max_iters, changed);
            pt.lap(2);
            bool rose_any = changed;
            if constexpr (EXCL) tile_store_excl<T>(s, rose, ob, changed, im, mx, my);
            else rose_any = tile_store<T>(s, rose, im, mx, my);
            mem_wait();
            pt.lap(3);
            uint32_t nb = tile_notify_mask<T, CONN>(s, rose, im, mx, my);
            wave_lds_sync();
            // DQ: a duplicate visitor of the same megapixel may have used stale
            // pixels, so any rise re-queues the megapixel (as the reference does).
            if (!conv || (!EXCL && rose_any)) nb |= NB_SELF;
            ++visits;
            uint32_t target = 0, old = 0;
            const bool cand = lane < 9 && ((nb >> lane) & 1u);
            if (cand) {
                target = (uint32_t)(dsc.mp_base + neighbour_index(ml, dsc.nmx, 1u << lane));
                if (EXCL) old = a_or(&q.st[target], 1u);
            }
            // the next megapixel's loads go out behind the atomics (its fetch
            // follows this visit's completed stores)
            uint32_t m2 = 0;
            const bool have2 = more && take(m2);
            more = have2;
            if (have2) {
                const ImgDesc d2 = mp_desc<ONE>(ds, m2);
                const int ml2 = (int)m2 - d2.mp_base;
                tile_fetch<T>(image_of(d2), ml2 % d2.nmx, ml2 / d2.nmx, s);
            }
            const bool push = cand && (!EXCL || old == 0u);
            const uint64_t b = ballot(push);
            const int n = __popcll(b);
            if (n) {
                if (wcount + n > 64) flush();
                // into the buffer through the (now free) rose area
                lds32* xs = reinterpret_cast<lds32*>(rose);
                if (push) xs[__popcll(b & ((1ull << lane) - 1ull))] = target;
                wave_lds_sync();
                if (lane >= wcount && lane < wcount + n) wbuf = xs[lane - wcount];
                wave_lds_sync();
                wcount += n;
            }
            pt.lap(4);
            m = m2;
            have = have2;
        }
        flush();
        if (!round_barrier(q, gridDim.x, round)) break;
        if (a_load(ctl_at(q.ctl, C_RSIZE)) == 0u) break;
        if (a_load(ctl_at(q.ctl, C_OVERFLOW)) != 0u) break;
        if (a_load(ctl_at(q.ctl, C_TIMEOUT)) != 0u) break;
        if ((int)round + 1 >= q.max_rounds) break;
    }
    pt.lap(0);
    pt.flush(q.ctl);
    if (lane == 0 && visits) {
        __hip_atomic_fetch_add(gp(reinterpret_cast<unsigned long long*>(ctl_at(q.ctl, C_VISITS_LO))),
                               (unsigned long long)visits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// After seeding: read-queue size / pending count from the seeding count;
// an over-full seeding raises the overflow flag.

// Queue checks: how many megapixel state words are not clear.
__global__ __launch_bounds__(kWG) void k_state_left(QueueArgs q) {
    uint32_t n = 0;
    for (int i = blockIdx.x * kWG + (int)threadIdx.x; i < q.nmp; i += gridDim.x * kWG)
        n += q.st[i] != 0u;
    if (n) a_add(ctl_at(q.ctl, C_STATE_LEFT), n);  // (none when the phase ended clean)
}

// Block (i, b) of a batch's work-item tables: image i's ranges, in strides
// over the gridDim.y blocks of the image.
__global__ __launch_bounds__(kWG) void k_image_tables(const ImgDesc* ds, uint32_t* img_of_mp,
                                                      uint32_t* img_of_row, uint32_t* img_of_cw,
                                                      uint32_t* img_of_sw) {
    const int i = (int)blockIdx.x;
    const ImgDesc d = ds[i];
    const size_t t0 = (size_t)blockIdx.y * kWG + threadIdx.x, step = (size_t)gridDim.y * kWG;
    const size_t nmp = (size_t)d.nmx * d.nmy, ncw = (size_t)d.strips * ((d.H + d.R - 1) / d.R);
    const size_t nsw = (size_t)d.strips * d.nmy;
    for (size_t t = t0; t < nmp; t += step) img_of_mp[d.mp_base + t] = (uint32_t)i;
    for (size_t t = t0; t < (size_t)d.H; t += step) img_of_row[d.row_base + t] = (uint32_t)i;
    for (size_t t = t0; t < ncw; t += step) img_of_cw[d.cw_base + t] = (uint32_t)i;
    for (size_t t = t0; t < nsw; t += step) img_of_sw[d.sw_base + t] = (uint32_t)i;
}

// High bit of every zero byte of w (exact: the low seven bits plus 0x7F carry
// into the high bit, never into the next byte).
DEV uint32_t zero_bytes(uint32_t w) { return ~(((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u; }

// Seed class of the markers (autotuner key): over every 64th row of every
// image, the pixels where min(marker, mask) > 0 and the pixels sampled
// (counts[0], counts[1]).  Block (i, b) samples image i's rows 64 (4 b + k),
// k < 4 (none past its last row), in 32-bit words, the four rows' loads
// issued together (rows start 4-byte aligned; a word past the row end lies
// inside the pitch and its bytes beyond the row are not counted):
// min(marker, mask) > 0 where both are non-zero, counted per byte (uint8) or
// half-word (uint16) from their zero masks.  One pair of atomics per block:
// atomics on one address serialise (one pair per wave and sampled row cost
// 0.36 ms over 256 images).
constexpr int kClassRows = 4;
template <class T>
__global__ __launch_bounds__(kWG) void k_marker_class(Descs ds, unsigned long long* counts) {
    const ImgDesc d = desc(ds, (int)blockIdx.x);
    const int rowbytes = d.W * (int)sizeof(T), nwords = (rowbytes + 3) / 4;
    const int r0 = (int)blockIdx.y * kClassRows;
    if (r0 * 64 >= d.H) return;  // the whole workgroup
    uint32_t nz = 0, n = 0;
    for (int w = (int)threadIdx.x; w < nwords; w += kWG) {
        const int vb = min(4, rowbytes - 4 * w);  // bytes of the word inside the row
        const uint32_t keep = vb >= 4 ? 0xFFFFFFFFu : (1u << (8 * vb)) - 1u;
        uint32_t jv[kClassRows], iv[kClassRows];
#pragma unroll
        for (int k = 0; k < kClassRows; ++k) {
            const int y = min((r0 + k) * 64, d.H - 1);  // (rows past the image not counted)
            jv[k] = *gp(reinterpret_cast<const uint32_t*>(d.J + (size_t)y * d.pj) + w);
            iv[k] = *gp(reinterpret_cast<const uint32_t*>(d.I + (size_t)y * d.pi) + w);
        }
#pragma unroll
        for (int k = 0; k < kClassRows; ++k) {
            if ((r0 + k) * 64 >= d.H) break;  // the same in every lane
            const uint32_t j = jv[k] & keep, i = iv[k] & keep;
            uint32_t both;
            if constexpr (sizeof(T) == 1) {
                // high bit of each byte that is non-zero in both
                both = ~(zero_bytes(j) | zero_bytes(i)) & keep & 0x80808080u;
            } else {
                both = ((j & 0xFFFFu) && (i & 0xFFFFu) ? 1u : 0u) |
                       ((j >> 16) && (i >> 16) ? 0x10000u : 0u);
            }
            nz += (uint32_t)__popc(both);
            n += (uint32_t)vb / (uint32_t)sizeof(T);
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {  // wave sums, then the workgroup's in LDS
        nz += __shfl_xor(nz, m, 64);
        n += __shfl_xor(n, m, 64);
    }
    __shared__ uint32_t red[2][kWG / 64];
    const int wv = (int)threadIdx.x >> 6;
    if (lane_id() == 0) {
        red[0][wv] = nz;
        red[1][wv] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int k = 0; k < kWG / 64; ++k) {
            a += red[0][k];
            b += red[1][k];
        }
        if (b) {
            __hip_atomic_fetch_add(gp(counts), a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(gp(counts + 1), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ============================================================ launchers ====
static int cu_count() {
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
    }
    return n;
}

static bool aligned16(const void* p, size_t pitch) {
    return ((uintptr_t)p % 16 == 0) && (pitch % 16 == 0);
}

hipError_t launch_clip(const Image& im, int dtype, hipStream_t s) {
    const bool v16 = aligned16(im.J, im.pj) && aligned16(im.I, im.pi);
    const int vb = v16 ? 16 : 4;
    const long nvec = ((long)im.W * dtype + vb - 1) / vb * im.H;
    const int grid = (int)std::min<long>((nvec + kWG - 1) / kWG, (long)cu_count() * 16);
    if (dtype == 1) {
        if (v16) hipLaunchKernelGGL((k_clip<uint8_t, 16>), grid, kWG, 0, s, im);
        else hipLaunchKernelGGL((k_clip<uint8_t, 4>), grid, kWG, 0, s, im);
    } else {
        if (v16) hipLaunchKernelGGL((k_clip<uint16_t, 16>), grid, kWG, 0, s, im);
        else hipLaunchKernelGGL((k_clip<uint16_t, 4>), grid, kWG, 0, s, im);
    }
    return hipGetLastError();
}

template <class T, int NCH>
static void row_pass_wave_t(const Descs& ds, bool clip, uint32_t* changed, hipStream_t s) {
    const int grid = (int)std::min<long>(((long)ds.total + 3) / 4, 65535);
    if (clip) hipLaunchKernelGGL((k_row_pass_wave<T, 16, NCH, true>), grid, kWG, 0, s, ds, changed);
    else hipLaunchKernelGGL((k_row_pass_wave<T, 16, NCH, false>), grid, kWG, 0, s, ds, changed);
}

// One wave per row of at most 8 KiB (16-byte lanes, 1 KiB per chunk).
static void row_pass_wave(const Descs& ds, int dtype, bool clip, int rowbytes, uint32_t* changed,
                          hipStream_t s) {
    if (dtype == 1) {
        if (rowbytes <= 4096) row_pass_wave_t<uint8_t, 4>(ds, clip, changed, s);
        else row_pass_wave_t<uint8_t, 8>(ds, clip, changed, s);
    } else {
        if (rowbytes <= 4096) row_pass_wave_t<uint16_t, 4>(ds, clip, changed, s);
        else row_pass_wave_t<uint16_t, 8>(ds, clip, changed, s);
    }
}
constexpr int kWaveRowMax = 8192;

constexpr int kHeldLanes = 256;

template <class T, int VB, bool CLIP>
static void row_pass_t(const Image& im, const TileMask& tm, uint32_t* changed, hipStream_t s) {
    // Rows of 4-32 KiB (16-byte lanes) are held in registers by one
    // 256-lane workgroup (k_row_pass_held): read once, written once.
    // Measured on MI355X, one isolated pass, algorithmic bytes (3 per u8
    // pixel, 6 per u16): 32768-wide u8 rows 0.580 ms = 5.6 TB/s held
    // (512 lanes 5.3, 1024 lanes 4.9) against 0.705 ms = 4.6 TB/s streamed
    // with the backward re-read; 16384-wide u16 rows 0.309 ms = 5.2 TB/s
    // against 0.425 ms for the earlier 1024-lane form with a barrier per
    // chunk; 6000-wide u8 rows 3.7 TB/s against 3.2 streamed.  Wider rows
    // are streamed chunk by chunk, forward and backward in one launch.
    const int rowbytes = im.W * (int)sizeof(T);
    const int grid = std::min(im.H, 65535);
    if (VB == 16 && rowbytes > 256 * VB && rowbytes <= 32768) {
        constexpr int NT = kHeldLanes, CB = NT * VB;  // lanes, bytes per chunk
        if (rowbytes <= CB)
            hipLaunchKernelGGL((k_row_pass_held<T, NT, 1, CLIP>), grid, NT, 0, s, im, tm, changed);
        else if (rowbytes <= 2 * CB)
            hipLaunchKernelGGL((k_row_pass_held<T, NT, 2, CLIP>), grid, NT, 0, s, im, tm, changed);
        else if (rowbytes <= 4 * CB)
            hipLaunchKernelGGL((k_row_pass_held<T, NT, 4, CLIP>), grid, NT, 0, s, im, tm, changed);
        else  // (the 8-chunk form exists for the lane counts that can need it)
            hipLaunchKernelGGL((k_row_pass_held<T, NT, NT <= 512 ? 8 : 4, CLIP>), grid, NT, 0, s, im,
                               tm, changed);
    } else if (rowbytes <= 256 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 1, 4, CLIP>), grid, 256, 0, s, im, tm, changed);
    else if (sizeof(T) == 2 && rowbytes <= 2 * 256 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 2, 4, CLIP>), grid, 256, 0, s, im, tm, changed);
    else if (sizeof(T) == 2 && rowbytes <= 1024 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 1, 16, CLIP>), grid, 1024, 0, s, im, tm, changed);
    else if (sizeof(T) == 2 && rowbytes <= 2 * 1024 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 2, 16, CLIP>), grid, 1024, 0, s, im, tm, changed);
    else
        hipLaunchKernelGGL((k_row_pass_long<T, VB, 0, CLIP>), grid, kWG, 0, s, im, tm, changed);
}

hipError_t launch_row_pass(const Image& im, int dtype, bool clip, const TileMask& tm,
                           uint32_t* changed, hipStream_t s) {
    const bool v16 = aligned16(im.J, im.pj) && aligned16(im.I, im.pi);
    const int rowbytes = im.W * dtype;
    // Rows of up to 2 KiB: one wave per row (measured on MI355X: 1024-wide
    // u8 rows 0.010 vs 0.012 ms per pass with a workgroup per row; at 4 KiB
    // the workgroup kernel is faster, 0.020 vs 0.022 ms).
    if (v16 && rowbytes <= 2048) {
        Descs ds{};
        ds.one.J = im.J;
        ds.one.I = im.I;
        ds.one.pj = im.pj;
        ds.one.pi = im.pi;
        ds.one.W = im.W;
        ds.one.H = im.H;
        ds.one.tm = tm;
        ds.n = 1;
        ds.total = im.H;
        row_pass_wave(ds, dtype, clip, rowbytes, changed, s);
        return hipGetLastError();
    }
    if (dtype == 1) {
        if (v16) clip ? row_pass_t<uint8_t, 16, true>(im, tm, changed, s)
                      : row_pass_t<uint8_t, 16, false>(im, tm, changed, s);
        else clip ? row_pass_t<uint8_t, 4, true>(im, tm, changed, s)
                  : row_pass_t<uint8_t, 4, false>(im, tm, changed, s);
    } else {
        if (v16) clip ? row_pass_t<uint16_t, 16, true>(im, tm, changed, s)
                      : row_pass_t<uint16_t, 16, false>(im, tm, changed, s);
        else clip ? row_pass_t<uint16_t, 4, true>(im, tm, changed, s)
                  : row_pass_t<uint16_t, 4, false>(im, tm, changed, s);
    }
    return hipGetLastError();
}

// 16-byte lanes need 16-byte aligned rows; narrow images use 4-byte lanes so
// that there are enough column strips to fill the GPU.
// 16-byte lanes need 16-byte aligned rows; narrow images use 4-byte lanes so
// that there are enough column strips to fill the GPU.
bool col_wide(const void* src, const void* dst, const void* mask, size_t ps, size_t pd,
              size_t pi, int W, int dtype) {
    return aligned16(src, ps) && aligned16(dst, pd) && aligned16(mask, pi) &&
           (size_t)W * dtype >= 16384;
}

int col_block_width(const void* src, const void* dst, const void* mask, size_t ps, size_t pd,
                    size_t pi, int W, int dtype) {
    return 64 * (col_wide(src, dst, mask, ps, pd, pi, W, dtype) ? 16 : 4) / dtype;
}

int col_strips(int W, int dtype, bool wide) {
    const int vb = wide ? 16 : 4;
    return (W * dtype + 64 * vb - 1) / (64 * vb);
}

int col_auto_rows(int W, int H, int dtype, bool wide) {
    const long strips = col_strips(W, dtype, wide);
    long r = strips * (long)H / 8192;  // aim at >= 8192 waves
    r = r / 32 * 32;
    return (int)std::min<long>(512, std::max<long>(32, r));
}


template <class T, int VB>
static void row_pass_batch_t(const Descs& ds, bool clip, uint32_t* changed, hipStream_t s) {
    const int grid = std::min(ds.total, 65535);
    if (clip) hipLaunchKernelGGL((k_row_pass_batch<T, VB, true>), grid, kWG, 0, s, ds, changed);
    else hipLaunchKernelGGL((k_row_pass_batch<T, VB, false>), grid, kWG, 0, s, ds, changed);
}

hipError_t launch_row_pass_batch(const Descs& ds, int dtype, bool clip, bool wide,
                                 int wave_rowbytes, uint32_t* changed, hipStream_t s) {
    if (ds.total <= 0) return hipSuccess;
    if (wave_rowbytes > 0 && wave_rowbytes <= kWaveRowMax) {
        row_pass_wave(ds, dtype, clip, wave_rowbytes, changed, s);
        return hipGetLastError();
    }
    if (dtype == 1) {
        if (wide) row_pass_batch_t<uint8_t, 16>(ds, clip, changed, s);
        else row_pass_batch_t<uint8_t, 4>(ds, clip, changed, s);
    } else {
        if (wide) row_pass_batch_t<uint16_t, 16>(ds, clip, changed, s);
        else row_pass_batch_t<uint16_t, 4>(ds, clip, changed, s);
    }
    return hipGetLastError();
}

template <class T, int VB, int CONN, int DIR>
static void col_pass_dir(const Descs& ds, int mode, uint32_t* changed, uint8_t* flags,
                         uint32_t* edges, hipStream_t s) {
    const int grid = (ds.total + 3) / 4;
    if (mode == 0) hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 0>), grid, kWG, 0, s, ds, changed, flags, edges);
    else if (mode == 1) hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 1>), grid, kWG, 0, s, ds, changed, flags, edges);
    else if (mode == 2) hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 2>), grid, kWG, 0, s, ds, changed, flags, edges);
    else if (mode == 3) hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 3>), grid, kWG, 0, s, ds, changed, flags, edges);
    else if (mode == 4) hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 4>), grid, kWG, 0, s, ds, changed, flags, edges);
    else if (mode == 5) hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 5>), grid, kWG, 0, s, ds, changed, flags, edges);
    else if (mode == 6) {  // the clipping pass is the first (downward) pass
        if constexpr (DIR > 0)
            hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 6>), grid, kWG, 0, s, ds, changed, flags, edges);
    } else {  // the seeding pass is the last (upward) pass
        if constexpr (DIR < 0)
            hipLaunchKernelGGL((k_col_pass<T, VB, CONN, DIR, 7>), grid, kWG, 0, s, ds, changed, flags, edges);
    }
}

template <class T, int VB>
static void col_pass_t(const Descs& ds, int conn, int dir, int mode, uint32_t* changed,
                       uint8_t* flags, uint32_t* edges, hipStream_t s) {
    if (conn == 8) {
        if (dir > 0) col_pass_dir<T, VB, 8, 1>(ds, mode, changed, flags, edges, s);
        else col_pass_dir<T, VB, 8, -1>(ds, mode, changed, flags, edges, s);
    } else {
        if (dir > 0) col_pass_dir<T, VB, 4, 1>(ds, mode, changed, flags, edges, s);
        else col_pass_dir<T, VB, 4, -1>(ds, mode, changed, flags, edges, s);
    }
}

hipError_t launch_col_pass_batch(const Descs& ds, int dtype, int conn, int dir, int mode,
                                 bool wide, uint32_t* changed, hipStream_t s, uint8_t* flags,
                                 uint32_t* edges) {
    if (ds.total <= 0) return hipSuccess;
    if ((mode == COL_IN_PLACE_SEED || mode == COL_RASTER_SEED) && (!flags || !edges))
        return hipErrorInvalidValue;
    if (mode < COL_J_TO_P || mode > COL_RASTER_SEED || (mode == COL_RASTER_CLIP && dir < 0) ||
        (mode == COL_RASTER_SEED && dir > 0))
        return hipErrorInvalidValue;
    if (dtype == 1) {
        if (wide) col_pass_t<uint8_t, 16>(ds, conn, dir, mode, changed, flags, edges, s);
        else col_pass_t<uint8_t, 4>(ds, conn, dir, mode, changed, flags, edges, s);
    } else {
        if (wide) col_pass_t<uint16_t, 16>(ds, conn, dir, mode, changed, flags, edges, s);
        else col_pass_t<uint16_t, 4>(ds, conn, dir, mode, changed, flags, edges, s);
    }
    return hipGetLastError();
}

long seed_edge_waves(const ImgDesc& d) {
    const long nedges = std::max(0, d.strips - 1), nbounds = (d.H + d.R - 1) / d.R - 1;
    return nedges * ((d.H + 63) / 64) + std::max(0L, nbounds) * ((d.W + 63) / 64);
}

hipError_t launch_seed_edges(const Descs& ds, int dtype, int conn, bool wide,
                             const uint32_t* edges, uint8_t* flags, hipStream_t s) {
    if (ds.total <= 0) return hipSuccess;
    const int spx = 64 * (wide ? 16 : 4) / dtype;
    const int grid = (ds.total + 3) / 4;
    if (dtype == 1) {
        if (conn == 8) hipLaunchKernelGGL((k_seed_edges<uint8_t, 8>), grid, kWG, 0, s, ds, edges, spx, flags);
        else hipLaunchKernelGGL((k_seed_edges<uint8_t, 4>), grid, kWG, 0, s, ds, edges, spx, flags);
    } else {
        if (conn == 8) hipLaunchKernelGGL((k_seed_edges<uint16_t, 8>), grid, kWG, 0, s, ds, edges, spx, flags);
        else hipLaunchKernelGGL((k_seed_edges<uint16_t, 4>), grid, kWG, 0, s, ds, edges, spx, flags);
    }
    return hipGetLastError();
}

hipError_t launch_unstable_batch(const Descs& ds, int dtype, int conn, bool wide,
                                 uint8_t* flags, uint32_t* count, hipStream_t s) {
    if (ds.total <= 0) return hipSuccess;
    const int grid = (ds.total + 3) / 4;
#define IWPP_UN(TT, VBB, C) \
    hipLaunchKernelGGL((k_unstable_scan<TT, VBB, C>), grid, kWG, 0, s, ds, flags, count)
    if (dtype == 1) {
        if (wide) { if (conn == 8) IWPP_UN(uint8_t, 16, 8); else IWPP_UN(uint8_t, 16, 4); }
        else { if (conn == 8) IWPP_UN(uint8_t, 4, 8); else IWPP_UN(uint8_t, 4, 4); }
    } else {
        if (wide) { if (conn == 8) IWPP_UN(uint16_t, 16, 8); else IWPP_UN(uint16_t, 16, 4); }
        else { if (conn == 8) IWPP_UN(uint16_t, 4, 8); else IWPP_UN(uint16_t, 4, 4); }
    }
#undef IWPP_UN
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int grid = (int)std::min<size_t>((n + kWG - 1) / kWG, (size_t)cu_count() * 8);
    hipLaunchKernelGGL(k_fill_u32, grid, kWG, 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_marker_class(const Descs& ds, int dtype, unsigned long long* counts, int max_h,
                               hipStream_t s) {
    // one workgroup per image and kClassRows sampled rows of the tallest image
    const int rows = (max_h + 63) / 64;
    const dim3 grid((unsigned)ds.n, (unsigned)std::max(1, (rows + kClassRows - 1) / kClassRows));
    if (dtype == 1) hipLaunchKernelGGL((k_marker_class<uint8_t>), grid, kWG, 0, s, ds, counts);
    else hipLaunchKernelGGL((k_marker_class<uint16_t>), grid, kWG, 0, s, ds, counts);
    return hipGetLastError();
}

hipError_t launch_state_left(const QueueArgs& q, hipStream_t s) {
    const int grid = std::max(1, std::min((q.nmp + kWG - 1) / kWG, cu_count() * 4));
    hipLaunchKernelGGL(k_state_left, grid, kWG, 0, s, q);
    return hipGetLastError();
}

hipError_t launch_image_tables(const ImgDesc* d, int n, uint32_t* img_of_mp, uint32_t* img_of_row,
                               uint32_t* img_of_cw, uint32_t* img_of_sw, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int per_image = std::max(1, std::min(16, (cu_count() * 8 + n - 1) / n));
    hipLaunchKernelGGL(k_image_tables, dim3((unsigned)n, (unsigned)per_image), kWG, 0, s, d,
                       img_of_mp, img_of_row, img_of_cw, img_of_sw);
    return hipGetLastError();
}

hipError_t launch_queue_reset(uint32_t* ctl, size_t nctl, uint32_t* st, size_t nmp, uint32_t* chk,
                              unsigned long long* ring, size_t cap, hipStream_t s) {
    const size_t n = std::max(std::max(nctl, nmp), ring ? cap : (size_t)0);
    const int grid = (int)std::max<size_t>(1, std::min<size_t>((n + kWG - 1) / kWG,
                                                                (size_t)cu_count() * 8));
    hipLaunchKernelGGL(k_queue_reset, grid, kWG, 0, s, ctl, nctl, st, nmp, chk, ring, cap);
    return hipGetLastError();
}

hipError_t launch_seed_queue(const uint8_t* flags, int nmp, int kind, const QueueArgs& q,
                             hipStream_t s) {
    // small images: a chunk per wave, for more waves
    const int k = nmp >= 16384 ? 8 : 1;
    const int grid = std::max(1, std::min((nmp + kWG * k - 1) / (kWG * k), cu_count() * 8));
    if (k == 8) hipLaunchKernelGGL((k_seed_queue<8>), grid, kWG, 0, s, flags, nmp, kind, q);
    else hipLaunchKernelGGL((k_seed_queue<1>), grid, kWG, 0, s, flags, nmp, kind, q);
    return hipGetLastError();
}

template <class T, int CONN, bool ONE>
static hipError_t queue_phase_t(const Descs& ds, int kind, const QueueArgs& q, int wpc,
                                hipStream_t s) {
    const void* fn;
    if (kind == Q_MUTEX) fn = (const void*)k_queue_mutex<T, CONN, ONE>;
    else if (kind == Q_DQ) fn = (const void*)k_queue_bsp<T, CONN, Q_DQ, ONE>;
    else fn = (const void*)k_queue_bsp<T, CONN, Q_UQ, ONE>;
    int per_cu = 0;
    // tiles, and the mutex kernel's local queue (u16: 75 KiB, beyond the 64 KiB
    // default: say so)
    const size_t lds_bytes = (kWG / 64) * WAVE_LDS<T> + (kind == Q_MUTEX ? sizeof(LocalQueue) : 0);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWG, lds_bytes);
    if (e != hipSuccess) return e;
    if (per_cu <= 0) return hipErrorInvalidConfiguration;
    if (wpc > 0) per_cu = std::min(per_cu, wpc);
    int grid = per_cu * cu_count() / std::max(1, q.grid_div);
    // Test hook (failure handling): IWPP_TEST_OVERSUBSCRIBE=k launches the
    // round-based kernels k times larger than fits, without the cooperative
    // launch, so that the start gate must send them back.
    int over = 1;
    if (kind != Q_MUTEX) {
        if (const char* env = std::getenv("IWPP_TEST_OVERSUBSCRIBE")) over = std::max(1, std::atoi(env));
    }
    grid = std::max(1, std::min(grid, (q.nmp + 3) / 4));
    if (over > 1) grid = per_cu * cu_count() * over / std::max(1, q.grid_div);
    Descs dc = ds;
    QueueArgs qc = q;
    if (kind == Q_MUTEX) {
        hipLaunchKernelGGL((k_queue_mutex<T, CONN, ONE>), grid, kWG, lds_bytes, s, dc, qc);
        return hipGetLastError();
    }
    void* args[] = {&dc, &qc};
    if (over == 1) {
        e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kWG), args, lds_bytes, s);
        if (e == hipSuccess) return e;
    }
    // The runtime refuses cooperative launches in some states (for instance
    // while another process holds the device).  The grid is sized to fit the
    // device at once; if it still does not run co-resident, the kernel's
    // start gate sends every block back within q.start_ticks and the host
    // relaunches a smaller grid (every device-side wait is bounded in time).
    (void)hipGetLastError();
    if (kind == Q_DQ)
        hipLaunchKernelGGL((k_queue_bsp<T, CONN, Q_DQ, ONE>), grid, kWG, lds_bytes, s, dc, qc);
    else
        hipLaunchKernelGGL((k_queue_bsp<T, CONN, Q_UQ, ONE>), grid, kWG, lds_bytes, s, dc, qc);
    return hipGetLastError();
}

template <bool ONE>
static hipError_t queue_phase_one(const Descs& ds, int dtype, int conn, int kind,
                                  const QueueArgs& q, int wpc, hipStream_t s) {
    if (dtype == 1) {
        return conn == 8 ? queue_phase_t<uint8_t, 8, ONE>(ds, kind, q, wpc, s)
                         : queue_phase_t<uint8_t, 4, ONE>(ds, kind, q, wpc, s);
    }
    return conn == 8 ? queue_phase_t<uint16_t, 8, ONE>(ds, kind, q, wpc, s)
                     : queue_phase_t<uint16_t, 4, ONE>(ds, kind, q, wpc, s);
}

hipError_t launch_queue_phase(const Descs& ds, int dtype, int conn, int kind,
                              const QueueArgs& q, int wpc, hipStream_t s) {
    // a single image's descriptor is the kernel argument itself
    return ds.d == nullptr ? queue_phase_one<true>(ds, dtype, conn, kind, q, wpc, s)
                           : queue_phase_one<false>(ds, dtype, conn, kind, q, wpc, s);
}

}  // namespace iwpp
