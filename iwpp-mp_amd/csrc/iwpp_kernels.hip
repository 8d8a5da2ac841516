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

namespace iwpp {

#define DEV static __device__ __forceinline__

// ---------------------------------------------------------------- atomics ---
DEV uint32_t a_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_load_rlx(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV void a_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
DEV void a_store_rlx(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_or(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_and(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_and(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_xchg(uint32_t* p, uint32_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
DEV uint32_t a_cas(uint32_t* p, uint32_t expected, uint32_t desired) {
    __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                         __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    return expected;  // the value seen before the exchange
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

// ------------------------------------------------------ vector row access ---
template <class T, int VB>
struct VecIO {
    static constexpr int NW = VB / 4;                 // 32-bit words per lane
    static constexpr int NPX = VB / (int)sizeof(T);   // pixels per lane
    static constexpr int PPW = 4 / (int)sizeof(T);    // pixels per word
    static constexpr int BITS = 8 * (int)sizeof(T);
    static constexpr uint32_t PM = (uint32_t)((1u << BITS) - 1u);

    DEV void load(const uint8_t* row, int o, int rowbytes, uint32_t (&w)[NW]) {
        if (o + VB <= rowbytes) {
            if constexpr (VB == 16) {
                const uint4 t = *reinterpret_cast<const uint4*>(row + o);
                w[0] = t.x; w[1] = t.y; w[2] = t.z; w[3] = t.w;
            } else {
                w[0] = *reinterpret_cast<const uint32_t*>(row + o);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NW; ++i) w[i] = 0;
            if (o < rowbytes) {
#pragma unroll
                for (int k = 0; k < NPX; ++k) {
                    const int ob = o + k * (int)sizeof(T);
                    if (ob < rowbytes)
                        w[k / PPW] |= (uint32_t)(*reinterpret_cast<const T*>(row + ob))
                                      << (BITS * (k % PPW));
                }
            }
        }
    }
    DEV void store(uint8_t* row, int o, int rowbytes, const uint32_t (&w)[NW]) {
        if (o + VB <= rowbytes) {
            if constexpr (VB == 16) {
                *reinterpret_cast<uint4*>(row + o) = make_uint4(w[0], w[1], w[2], w[3]);
            } else {
                *reinterpret_cast<uint32_t*>(row + o) = w[0];
            }
        } else if (o < rowbytes) {
#pragma unroll
            for (int k = 0; k < NPX; ++k) {
                const int ob = o + k * (int)sizeof(T);
                if (ob < rowbytes)
                    *reinterpret_cast<T*>(row + ob) = (T)((w[k / PPW] >> (BITS * (k % PPW))) & PM);
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

DEV bool tile_on(const TileMask& tm, int x, int y) {
    return tm.active == nullptr || tm.active[(size_t)(y / tm.th) * tm.ntx + x / tm.tw] != 0;
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
    constexpr int NW = V::NW, NPX = V::NPX, NT = NWV * 64;
    __shared__ uint32_t sP[2][NCH][NWV], sQ[2][NCH][NWV];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rowbytes = im.W * (int)sizeof(T);
    bool any_change = false;
    for (int y = blockIdx.x; y < im.H; y += gridDim.x) {
        uint8_t* jr = im.J + (size_t)y * im.pj;
        const uint8_t* ir = im.I + (size_t)y * im.pi;
        uint32_t jv[NCH][NW], iv[NCH][NW];
        bool dch[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int o = (ch * NT + tid) * VB;
            V::load(jr, o, rowbytes, jv[ch]);
            V::load(ir, o, rowbytes, iv[ch]);
            dch[ch] = false;
            if (CLIP) {
                uint32_t a[NPX], b[NPX];
                V::unpack(jv[ch], a);
                V::unpack(iv[ch], b);
#pragma unroll
                for (int k = 0; k < NPX; ++k) {
                    dch[ch] |= a[k] > b[k];
                    a[k] = min(a[k], b[k]);
                }
                V::pack(a, jv[ch]);
            }
            if (o < rowbytes && !tile_on(tm, o / (int)sizeof(T), y)) {
#pragma unroll
                for (int i = 0; i < NW; ++i) iv[ch][i] = 0;  // wall: no rise, no conduction
            }
        }
        // forward (left to right)
        uint32_t carry = 0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            uint32_t v[NPX], c[NPX];
            V::unpack(jv[ch], v);
            V::unpack(iv[ch], c);
            uint32_t P = 0, Q = 0xFFFFFFFFu;
#pragma unroll
            for (int k = 0; k < NPX; ++k) {
                P = max(v[k], min(P, c[k]));
                Q = min(Q, c[k]);
            }
            uint32_t Pi = P, Qi = Q;
            scan_lanes_fwd(Pi, Qi);
            const uint32_t Pe = from_left(Pi, 0u), Qe = from_left(Qi, 0xFFFFFFFFu);
            if (lane == 63) {
                sP[0][ch][wv] = Pi;
                sQ[0][ch][wv] = Qi;
            }
            __syncthreads();
            uint32_t xin = carry, xall = carry;
#pragma unroll
            for (int u = 0; u < NWV; ++u) {
                const uint32_t tp = sP[0][ch][u], tq = sQ[0][ch][u];
                if (u < wv) xin = max(tp, min(xin, tq));
                xall = max(tp, min(xall, tq));
            }
            carry = xall;
            uint32_t x = max(Pe, min(xin, Qe));  // value left of this lane's first pixel
            bool d = false;
#pragma unroll
            for (int k = 0; k < NPX; ++k) {
                const uint32_t out = max(v[k], min(x, c[k]));
                d |= out != v[k];
                v[k] = out;
                x = out;
            }
            dch[ch] |= d;
            V::pack(v, jv[ch]);
        }
        // backward (right to left)
        carry = 0;
#pragma unroll
        for (int ch = NCH - 1; ch >= 0; --ch) {
            uint32_t v[NPX], c[NPX];
            V::unpack(jv[ch], v);
            V::unpack(iv[ch], c);
            uint32_t P = 0, Q = 0xFFFFFFFFu;
#pragma unroll
            for (int k = NPX - 1; k >= 0; --k) {
                P = max(v[k], min(P, c[k]));
                Q = min(Q, c[k]);
            }
            uint32_t Pi = P, Qi = Q;
            scan_lanes_bwd(Pi, Qi, lane);
            const uint32_t Pe = from_right(Pi, 0u), Qe = from_right(Qi, 0xFFFFFFFFu);
            if (lane == 0) {
                sP[1][ch][wv] = Pi;
                sQ[1][ch][wv] = Qi;
            }
            __syncthreads();
            uint32_t xin = carry, xall = carry;
#pragma unroll
            for (int u = NWV - 1; u >= 0; --u) {
                const uint32_t tp = sP[1][ch][u], tq = sQ[1][ch][u];
                if (u > wv) xin = max(tp, min(xin, tq));
                xall = max(tp, min(xall, tq));
            }
            carry = xall;
            uint32_t x = max(Pe, min(xin, Qe));  // value right of this lane's last pixel
            bool d = false;
#pragma unroll
            for (int k = NPX - 1; k >= 0; --k) {
                const uint32_t out = max(v[k], min(x, c[k]));
                d |= out != v[k];
                v[k] = out;
                x = out;
            }
            dch[ch] |= d;
            V::pack(v, jv[ch]);
        }
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            if (dch[ch]) V::store(jr, (ch * NT + tid) * VB, rowbytes, jv[ch]);
            any_change |= dch[ch];
        }
    }
    if (any_change && changed) *changed = 1u;
}

// Rows wider than the register-resident kernel: one direction per launch,
// chunks streamed with a carry, every chunk stored.
template <class T, int VB, int DIR, bool CLIP>
__global__ __launch_bounds__(kWG) void k_row_pass_long(Image im, TileMask tm, uint32_t* changed) {
    using V = VecIO<T, VB>;
    constexpr int NW = V::NW, NPX = V::NPX;
    __shared__ uint32_t sP[2][4], sQ[2][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rowbytes = im.W * (int)sizeof(T);
    const int nch = (rowbytes + kWG * VB - 1) / (kWG * VB);
    bool any_change = false;
    for (int y = blockIdx.x; y < im.H; y += gridDim.x) {
        uint8_t* jr = im.J + (size_t)y * im.pj;
        const uint8_t* ir = im.I + (size_t)y * im.pi;
        uint32_t carry = 0;
        for (int n = 0; n < nch; ++n) {
            const int ch = DIR > 0 ? n : nch - 1 - n;
            const int o = (ch * kWG + tid) * VB;
            uint32_t jw[NW], iw[NW], v[NPX], c[NPX], ov[NPX];
            V::load(jr, o, rowbytes, jw);
            V::load(ir, o, rowbytes, iw);
            V::unpack(jw, v);
            V::unpack(iw, c);
#pragma unroll
            for (int k = 0; k < NPX; ++k) {
                ov[k] = v[k];
                if (CLIP) v[k] = min(v[k], c[k]);
            }
            if (o < rowbytes && !tile_on(tm, o / (int)sizeof(T), y)) {
#pragma unroll
                for (int k = 0; k < NPX; ++k) c[k] = 0;
            }
            uint32_t P = 0, Q = 0xFFFFFFFFu;
#pragma unroll
            for (int kk = 0; kk < NPX; ++kk) {
                const int k = DIR > 0 ? kk : NPX - 1 - kk;
                P = max(v[k], min(P, c[k]));
                Q = min(Q, c[k]);
                v[k] = P;
            }
            uint32_t Pi = P, Qi = Q, Pe, Qe;
            if (DIR > 0) {
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
                const int u = DIR > 0 ? uu : 3 - uu;
                const uint32_t tp = sP[n & 1][u], tq = sQ[n & 1][u];
                if (DIR > 0 ? (u < wv) : (u > wv)) xin = max(tp, min(xin, tq));
                xall = max(tp, min(xall, tq));
            }
            carry = xall;
            const uint32_t xl = max(Pe, min(xin, Qe));
            uint32_t q = 0xFFFFFFFFu;
            bool d = false;
#pragma unroll
            for (int kk = 0; kk < NPX; ++kk) {
                const int k = DIR > 0 ? kk : NPX - 1 - kk;
                q = min(q, c[k]);
                v[k] = max(v[k], min(xl, q));
                d |= v[k] != ov[k];
            }
            if (d) {
                V::pack(v, jw);
                V::store(jr, o, rowbytes, jw);
            }
            any_change |= d;
            __syncthreads();  // the LDS slot pair is reused two chunks later
        }
    }
    if (any_change && changed) *changed = 1u;
}

// ----------------------------------------------------------- column pass ----
// Each wave sweeps one block of (64 lanes x VB bytes) columns x R rows,
// reading src and writing dst (out of place).  Inside the block each row
// uses the block's previous output row; neighbours outside the block come
// from src, so the pass is deterministic.
template <class T, int VB, int CONN, int DIR>
__global__ __launch_bounds__(kWG) void k_col_pass(const uint8_t* __restrict__ S,
                                                   uint8_t* __restrict__ D,
                                                   const uint8_t* __restrict__ I, size_t ps,
                                                   size_t pd, size_t pi, int W, int H, int R,
                                                   TileMask tm, uint32_t* changed) {
    using V = VecIO<T, VB>;
    constexpr int NW = V::NW, NPX = V::NPX;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rowbytes = W * (int)sizeof(T);
    const int strip = blockIdx.x * 4 + wv;
    const int xb0 = strip * 64 * VB;  // first byte of the strip
    if (xb0 >= rowbytes) return;      // whole wave past the row end (no barriers below)
    const int o = xb0 + lane * VB;
    const int xs0 = xb0 / (int)sizeof(T);             // first pixel of the strip
    const int xs1 = (xb0 + 64 * VB) / (int)sizeof(T); // first pixel right of the strip
    const int y0 = blockIdx.y * R, y1 = min(H, y0 + R);
    const bool on = o < rowbytes ? tile_on(tm, o / (int)sizeof(T), y0) : false;
    uint32_t prev[NPX];
    {
        const int yh = DIR > 0 ? y0 - 1 : y1;
        uint32_t w[NW];
        if (yh >= 0 && yh < H) V::load(S + (size_t)yh * ps, o, rowbytes, w);
        else {
#pragma unroll
            for (int i = 0; i < NW; ++i) w[i] = 0;
        }
        V::unpack(w, prev);
    }
    bool any_change = false;
    for (int i = 0; i < y1 - y0; ++i) {
        const int y = DIR > 0 ? y0 + i : y1 - 1 - i;
        const int yp = y - DIR;
        uint32_t jw[NW], iw[NW], v[NPX], c[NPX];
        V::load(S + (size_t)y * ps, o, rowbytes, jw);
        V::load(I + (size_t)y * pi, o, rowbytes, iw);
        V::unpack(jw, v);
        V::unpack(iw, c);
        if (!on || (tm.active && !tile_on(tm, o / (int)sizeof(T), y))) {
#pragma unroll
            for (int k = 0; k < NPX; ++k) c[k] = 0;
        }
        uint32_t m[NPX];
        if constexpr (CONN == 8) {
            uint32_t l_edge = 0, r_edge = 0;
            const bool yp_in = yp >= 0 && yp < H;
            if (lane == 0 && yp_in && xs0 > 0)
                l_edge = *reinterpret_cast<const T*>(S + (size_t)yp * ps + (size_t)(xs0 - 1) * sizeof(T));
            if (lane == 63 && yp_in && xs1 < W)
                r_edge = *reinterpret_cast<const T*>(S + (size_t)yp * ps + (size_t)xs1 * sizeof(T));
            const uint32_t left = from_left(prev[NPX - 1], l_edge);
            const uint32_t right = from_right(prev[0], r_edge);
#pragma unroll
            for (int k = 0; k < NPX; ++k) {
                const uint32_t a = k > 0 ? prev[k - 1] : left;
                const uint32_t b = k < NPX - 1 ? prev[k + 1] : right;
                m[k] = max(prev[k], max(a, b));
            }
        } else {
#pragma unroll
            for (int k = 0; k < NPX; ++k) m[k] = prev[k];
        }
        bool d = false;
#pragma unroll
        for (int k = 0; k < NPX; ++k) {
            const uint32_t out = max(v[k], min(m[k], c[k]));
            d |= out != v[k];
            prev[k] = out;
        }
        any_change |= d;
        V::pack(prev, jw);
        V::store(D + (size_t)y * pd, o, rowbytes, jw);
    }
    if (any_change && changed) *changed = 1u;
}

// ------------------------------------------------- megapixel (64 x 64) tile --
// Lane c holds column x0+c: rows 0..63 in registers.  Halo columns sit in two
// registers each (lane k holds row y0-1+k, k < 64; lanes 0,1 of the second
// register hold k = 64, 65), read back with readlane at compile-time lanes.
template <class T>
DEV uint32_t ldp(const uint8_t* base, size_t pitch, int x, int y) {
    return *reinterpret_cast<const T*>(base + (size_t)y * pitch + (size_t)x * sizeof(T));
}

struct Tile {
    uint32_t j[kMP], m[kMP];                // marker, mask
    uint32_t roseLo, roseHi;                // bit r: row r of this lane's column rose
    uint32_t hT, iT, hB, iB;                // halo rows above/below (this lane's column)
    uint32_t hLa, hLb, iLa, iLb, hRa, hRb, iRa, iRb;  // halo columns
};

template <int K>
DEV uint32_t halo(uint32_t a, uint32_t b) {
    if constexpr (K < 64) return rdlane(a, K);
    else return rdlane(b, K - 64);
}

template <class T>
DEV void tile_fetch(const Image& im, int mx, int my, Tile& t) {
    const int lane = lane_id();
    const int x0 = mx * kMP, y0 = my * kMP, x = x0 + lane;
    const bool inx = x < im.W;
#pragma unroll
    for (int r = 0; r < kMP; ++r) {
        const int y = y0 + r;
        const bool in = inx && y < im.H;
        t.j[r] = in ? ldp<T>(im.J, im.pj, x, y) : 0u;
        t.m[r] = in ? ldp<T>(im.I, im.pi, x, y) : 0u;
    }
    t.roseLo = 0u;
    t.roseHi = 0u;
    const bool top = inx && y0 > 0, bot = inx && y0 + kMP < im.H;
    t.hT = top ? ldp<T>(im.J, im.pj, x, y0 - 1) : 0u;
    t.iT = top ? ldp<T>(im.I, im.pi, x, y0 - 1) : 0u;
    t.hB = bot ? ldp<T>(im.J, im.pj, x, y0 + kMP) : 0u;
    t.iB = bot ? ldp<T>(im.I, im.pi, x, y0 + kMP) : 0u;
    const bool hasL = x0 > 0, hasR = x0 + kMP < im.W;
    {
        const int y = y0 - 1 + lane;
        const bool iny = y >= 0 && y < im.H;
        t.hLa = (iny && hasL) ? ldp<T>(im.J, im.pj, x0 - 1, y) : 0u;
        t.iLa = (iny && hasL) ? ldp<T>(im.I, im.pi, x0 - 1, y) : 0u;
        t.hRa = (iny && hasR) ? ldp<T>(im.J, im.pj, x0 + kMP, y) : 0u;
        t.iRa = (iny && hasR) ? ldp<T>(im.I, im.pi, x0 + kMP, y) : 0u;
    }
    {
        const int y = y0 + 63 + lane;
        const bool iny = lane < 2 && y < im.H;
        t.hLb = (iny && hasL) ? ldp<T>(im.J, im.pj, x0 - 1, y) : 0u;
        t.iLb = (iny && hasL) ? ldp<T>(im.I, im.pi, x0 - 1, y) : 0u;
        t.hRb = (iny && hasR) ? ldp<T>(im.J, im.pj, x0 + kMP, y) : 0u;
        t.iRb = (iny && hasR) ? ldp<T>(im.I, im.pi, x0 + kMP, y) : 0u;
    }
}

template <int R>
DEV void tile_set(Tile& t, uint32_t out, bool& ch) {
    const bool c = out != t.j[R];
    ch |= c;
    if constexpr (R < 32) t.roseLo |= (uint32_t)c << R;
    else t.roseHi |= (uint32_t)c << (R - 32);
    t.j[R] = out;
}
template <int R>
DEV bool tile_rose(const Tile& t) {
    if constexpr (R < 32) return (t.roseLo >> R) & 1u;
    else return (t.roseHi >> (R - 32)) & 1u;
}
DEV bool rose_bit(uint32_t lo, uint32_t hi, int r) {
    return ((r < 32 ? lo >> r : hi >> (r - 32)) & 1u) != 0u;
}

template <int R>
DEV void tile_xf_row(Tile& t, bool& ch) {
    uint32_t P = t.j[R], Q = t.m[R];
    scan_lanes_fwd(P, Q);
    const uint32_t out = max(P, min(halo<R + 1>(t.hLa, t.hLb), Q));
    tile_set<R>(t, out, ch);
}
template <int R>
DEV void tile_xb_row(Tile& t, bool& ch, int lane) {
    uint32_t P = t.j[R], Q = t.m[R];
    scan_lanes_bwd(P, Q, lane);
    const uint32_t out = max(P, min(halo<R + 1>(t.hRa, t.hRb), Q));
    tile_set<R>(t, out, ch);
}
template <int... Rs>
DEV void tile_xf_all(Tile& t, bool& ch, std::integer_sequence<int, Rs...>) {
    (tile_xf_row<Rs>(t, ch), ...);
}
template <int... Rs>
DEV void tile_xb_all(Tile& t, bool& ch, int lane, std::integer_sequence<int, Rs...>) {
    (tile_xb_row<Rs>(t, ch, lane), ...);
}

// Downward column sweep over the tile; K is the halo index of row R-1.
template <int CONN, int R>
DEV void tile_yd_row(Tile& t, uint32_t& prev, bool& ch) {
    uint32_t mx = prev;
    if constexpr (CONN == 8) {
        mx = max(mx, from_left(prev, halo<R>(t.hLa, t.hLb)));
        mx = max(mx, from_right(prev, halo<R>(t.hRa, t.hRb)));
    }
    const uint32_t out = max(t.j[R], min(mx, t.m[R]));
    tile_set<R>(t, out, ch);
    prev = out;
}
template <int CONN, int R>
DEV void tile_yu_row(Tile& t, uint32_t& prev, bool& ch) {
    uint32_t mx = prev;
    if constexpr (CONN == 8) {
        mx = max(mx, from_left(prev, halo<R + 2>(t.hLa, t.hLb)));
        mx = max(mx, from_right(prev, halo<R + 2>(t.hRa, t.hRb)));
    }
    const uint32_t out = max(t.j[R], min(mx, t.m[R]));
    tile_set<R>(t, out, ch);
    prev = out;
}
template <int CONN, int... Rs>
DEV void tile_yd_all(Tile& t, bool& ch, std::integer_sequence<int, Rs...>) {
    uint32_t prev = t.hT;
    (tile_yd_row<CONN, Rs>(t, prev, ch), ...);
}
template <int CONN, int... Rs>
DEV void tile_yu_all(Tile& t, bool& ch, std::integer_sequence<int, Rs...>) {
    uint32_t prev = t.hB;
    (tile_yu_row<CONN, kMP - 1 - Rs>(t, prev, ch), ...);
}

using Rows = std::make_integer_sequence<int, kMP>;

// Local sweeps until nothing changes (or max_iters rounds).  Returns true when
// the tile converged.
template <int CONN>
DEV bool tile_converge(Tile& t, int max_iters) {
    const int lane = lane_id();
    for (int it = 0; it < max_iters; ++it) {
        bool ch = false;
        tile_xf_all(t, ch, Rows{});
        tile_yd_all<CONN>(t, ch, Rows{});
        tile_xb_all(t, ch, lane, Rows{});
        tile_yu_all<CONN>(t, ch, Rows{});
        if (!wave_any(ch)) return true;
    }
    return false;
}

// Neighbour bits of a megapixel.
enum : uint32_t {
    NB_SELF = 1u << 0, NB_T = 1u << 1, NB_B = 1u << 2, NB_L = 1u << 3, NB_R = 1u << 4,
    NB_TL = 1u << 5, NB_TR = 1u << 6, NB_BL = 1u << 7, NB_BR = 1u << 8
};

// Which neighbours hold a pixel q next to a pixel p of this tile that rose in
// this visit, with min(J(p), I(q)) > J(q) on the fetched halo values.
template <int CONN>
DEV uint32_t tile_notify_mask(const Tile& t, const Image& im, int mx, int my) {
    const int lane = lane_id();
    const int x0 = mx * kMP, y0 = my * kMP;
    const bool hasT = y0 > 0, hasB = y0 + kMP < im.H, hasL = x0 > 0, hasR = x0 + kMP < im.W;
    uint32_t nb = 0;
    if (hasT) {
        const uint32_t p = t.j[0];
        const bool up = tile_rose<0>(t);
        bool c = up && min(p, t.iT) > t.hT;
        if constexpr (CONN == 8) {
            const uint32_t hl = from_left(t.hT, 0u), il = from_left(t.iT, 0u);
            const uint32_t hr = from_right(t.hT, 0u), ir = from_right(t.iT, 0u);
            c |= up && lane > 0 && min(p, il) > hl;
            c |= up && lane < 63 && min(p, ir) > hr;
        }
        if (wave_any(c)) nb |= NB_T;
    }
    if (hasB) {
        const uint32_t p = t.j[kMP - 1];
        const bool up = tile_rose<kMP - 1>(t);
        bool c = up && min(p, t.iB) > t.hB;
        if constexpr (CONN == 8) {
            const uint32_t hl = from_left(t.hB, 0u), il = from_left(t.iB, 0u);
            const uint32_t hr = from_right(t.hB, 0u), ir = from_right(t.iB, 0u);
            c |= up && lane > 0 && min(p, il) > hl;
            c |= up && lane < 63 && min(p, ir) > hr;
        }
        if (wave_any(c)) nb |= NB_B;
    }
    // Left/right: gather column 0 / column 63 into lane r (row r) so every lane
    // tests one row against the halo column held in lane r.
    if (hasL || hasR) {
        uint32_t cl = 0, cr = 0;
#pragma unroll
        for (int r = 0; r < kMP; ++r) {
            const uint32_t a = rdlane(t.j[r], 0);
            const uint32_t b = rdlane(t.j[r], 63);
            if (lane == r) { cl = a; cr = b; }
        }
        const bool upL = rose_bit(rdlane(t.roseLo, 0), rdlane(t.roseHi, 0), lane);
        const bool upR = rose_bit(rdlane(t.roseLo, 63), rdlane(t.roseHi, 63), lane);
        if (hasL) {
            const bool up = upL;
            // lane r: halo row r is at index r+1 -> register a, lane r+1
            const uint32_t h = from_right(t.hLa, rdlane(t.hLb, 0));
            const uint32_t i = from_right(t.iLa, rdlane(t.iLb, 0));
            bool c = up && min(cl, i) > h;
            if constexpr (CONN == 8) {
                // rows r-1 (index r) and r+1 (index r+2), kept inside rows 0..63
                const uint32_t h0 = t.hLa, i0 = t.iLa;
                const uint32_t h2 = from_right(h, rdlane(t.hLb, 1));
                const uint32_t i2 = from_right(i, rdlane(t.iLb, 1));
                c |= up && lane > 0 && min(cl, i0) > h0;
                c |= up && lane < 63 && min(cl, i2) > h2;
            }
            if (wave_any(c)) nb |= NB_L;
        }
        if (hasR) {
            const bool up = upR;
            const uint32_t h = from_right(t.hRa, rdlane(t.hRb, 0));
            const uint32_t i = from_right(t.iRa, rdlane(t.iRb, 0));
            bool c = up && min(cr, i) > h;
            if constexpr (CONN == 8) {
                const uint32_t h0 = t.hRa, i0 = t.iRa;
                const uint32_t h2 = from_right(h, rdlane(t.hRb, 1));
                const uint32_t i2 = from_right(i, rdlane(t.iRb, 1));
                c |= up && lane > 0 && min(cr, i0) > h0;
                c |= up && lane < 63 && min(cr, i2) > h2;
            }
            if (wave_any(c)) nb |= NB_R;
        }
    }
    if constexpr (CONN == 8) {
        const uint32_t j00 = rdlane(t.j[0], 0), j063 = rdlane(t.j[0], 63);
        const uint32_t j630 = rdlane(t.j[kMP - 1], 0), j6363 = rdlane(t.j[kMP - 1], 63);
        const bool u00 = rdlane(t.roseLo, 0) & 1u, u063 = rdlane(t.roseLo, 63) & 1u;
        const bool u630 = rdlane(t.roseHi, 0) >> 31, u6363 = rdlane(t.roseHi, 63) >> 31;
        if (hasT && hasL && u00 && min(j00, halo<0>(t.iLa, t.iLb)) > halo<0>(t.hLa, t.hLb))
            nb |= NB_TL;
        if (hasT && hasR && u063 &&
            min(j063, halo<0>(t.iRa, t.iRb)) > halo<0>(t.hRa, t.hRb))
            nb |= NB_TR;
        if (hasB && hasL && u630 &&
            min(j630, halo<65>(t.iLa, t.iLb)) > halo<65>(t.hLa, t.hLb))
            nb |= NB_BL;
        if (hasB && hasR && u6363 &&
            min(j6363, halo<65>(t.iRa, t.iRb)) > halo<65>(t.hRa, t.hRb))
            nb |= NB_BR;
    }
    return nb;
}

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

// Write back the rows that rose.  EXCL: the visitor is the only writer of its
// pixels (plain stores).  Otherwise other visitors of the same megapixel may
// write concurrently: 32-bit compare-and-swap of the per-pixel max.
// Returns true when a pixel in memory rose.
template <class T, bool EXCL>
DEV bool tile_store(const Tile& t, const Image& im, int mx, int my) {
    const int lane = lane_id();
    const int x = mx * kMP + lane, y0 = my * kMP;
    bool rose = false;
    if constexpr (EXCL) {
#pragma unroll
        for (int r = 0; r < kMP; ++r) {
            if (rose_bit(t.roseLo, t.roseHi, r)) {
                *reinterpret_cast<T*>(im.J + (size_t)(y0 + r) * im.pj + (size_t)x * sizeof(T)) =
                    (T)t.j[r];
                rose = true;
            }
        }
    } else {
        constexpr int PPW = 4 / (int)sizeof(T), BITS = 8 * (int)sizeof(T);
        constexpr uint32_t PM = (uint32_t)((1u << BITS) - 1u);
        // Lanes are pixels; the word holding pixel x starts at pixel x & ~(PPW-1).
        // The lane owning the word's first pixel (or the first in-tile pixel of a
        // word that straddles the tile's left edge) gathers its PPW pixels.
        const int xw = x & ~(PPW - 1);
        const int sub = x - xw;
        const bool leader = (sub == 0) || lane == 0;
        uint32_t rlo[PPW], rhi[PPW];
#pragma unroll
        for (int k = 0; k < PPW; ++k) {
            const int src = lane - sub + k;
            const bool valid = src >= 0 && src < 64;
            rlo[k] = valid ? __shfl(t.roseLo, src & 63) : 0u;
            rhi[k] = valid ? __shfl(t.roseHi, src & 63) : 0u;
        }
#pragma unroll
        for (int r = 0; r < kMP; ++r) {
            // new value and rise flag of pixel xw+k, taken from lane (lane - sub + k)
            uint32_t nv[PPW];
            bool any_up = false;
#pragma unroll
            for (int k = 0; k < PPW; ++k) {
                const int src = lane - sub + k;
                const uint32_t v = __shfl(t.j[r], src & 63);
                const bool valid = src >= 0 && src < 64;
                nv[k] = valid ? v : 0u;
                any_up |= rose_bit(rlo[k], rhi[k], r);
            }
            if (leader && any_up && x < im.W && y0 + r < im.H) {
                uint32_t* wp = reinterpret_cast<uint32_t*>(im.J + (size_t)(y0 + r) * im.pj +
                                                           (size_t)xw * sizeof(T));
                uint32_t cur = *wp;
                while (true) {
                    uint32_t want = cur;
#pragma unroll
                    for (int k = 0; k < PPW; ++k) {
                        const uint32_t c = (cur >> (BITS * k)) & PM;
                        if (nv[k] > c) want = (want & ~(PM << (BITS * k))) | (nv[k] << (BITS * k));
                    }
                    if (want == cur) break;
                    const uint32_t seen = a_cas(wp, cur, want);
                    if (seen == cur) { rose = true; break; }
                    cur = seen;
                }
            }
        }
    }
    return wave_any(rose);
}

// ------------------------------------------------------------ seeding -------
// flags[m] = 1 when megapixel m holds a pixel p with a neighbour q such that
// min(J(q), I(p)) > J(p).  One wave per megapixel.
template <class T, int CONN>
__global__ __launch_bounds__(kWG) void k_unstable(Image im, int nmx, int nmy, uint8_t* flags,
                                                   uint32_t* count) {
    const int lane = lane_id();
    const int wid = (blockIdx.x * kWG + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * kWG) >> 6;
    for (int m = wid; m < nmx * nmy; m += nwaves) {
        const int mx = m % nmx, my = m / nmx;
        Tile t;
        tile_fetch<T>(im, mx, my, t);
        bool un = false;
#pragma unroll
        for (int r = 0; r < kMP; ++r) {
            const uint32_t up = r > 0 ? t.j[r - 1] : t.hT;
            const uint32_t dn = r < kMP - 1 ? t.j[r + 1] : t.hB;
            uint32_t mx4 = max(up, dn);
            uint32_t lft, rgt;
            if (r + 1 < 64) {
                lft = from_left(t.j[r], rdlane(t.hLa, r + 1));
                rgt = from_right(t.j[r], rdlane(t.hRa, r + 1));
            } else {
                lft = from_left(t.j[r], rdlane(t.hLb, r + 1 - 64));
                rgt = from_right(t.j[r], rdlane(t.hRb, r + 1 - 64));
            }
            mx4 = max(mx4, max(lft, rgt));
            if constexpr (CONN == 8) {
                uint32_t ul, ur, dl, drr;
                ul = from_left(up, r < 64 ? rdlane(t.hLa, r) : 0u);
                ur = from_right(up, r < 64 ? rdlane(t.hRa, r) : 0u);
                if (r + 2 < 64) {
                    dl = from_left(dn, rdlane(t.hLa, r + 2));
                    drr = from_right(dn, rdlane(t.hRa, r + 2));
                } else {
                    dl = from_left(dn, rdlane(t.hLb, r + 2 - 64));
                    drr = from_right(dn, rdlane(t.hRb, r + 2 - 64));
                }
                mx4 = max(mx4, max(max(ul, ur), max(dl, drr)));
            }
            un |= min(mx4, t.m[r]) > t.j[r];
        }
        const bool u = wave_any(un);
        if (lane == 0) {
            flags[m] = u ? 1 : 0;
            if (u && count) atomicAdd(count, 1u);
        }
    }
}

// -------------------------------------------------------------- utilities ---
__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

// Compacts the unstable flags into queue A (ballot + popcount inside a wave,
// one atomicAdd per wave) and marks each queued megapixel in st:
//   mutex: st = PENDING, ring slot sequence = index + 1, pending += count
//   UQ:    st = 1 (already in the next read queue)
__global__ __launch_bounds__(kWG) void k_seed_queue(const uint8_t* flags, int nmp, int kind,
                                                     QueueArgs q) {
    const int lane = lane_id();
    for (int base = (blockIdx.x * kWG + threadIdx.x) & ~63; base < nmp;
         base += gridDim.x * kWG) {
        const int m = base + lane;
        const bool f = m < nmp && flags[m] != 0;
        const uint64_t b = ballot(f);
        const uint32_t n = (uint32_t)__popcll(b);
        if (n == 0) continue;
        uint32_t start = 0;
        if (lane == 0) start = a_add(ctl_at(q.ctl, C_TAIL), n);
        start = first_lane(start);
        const uint32_t off = (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
        if (f) {
            const uint32_t pos = start + off;
            if (pos < q.cap) {
                q.qa[pos] = (uint32_t)m;
                if (kind == Q_MUTEX) {
                    q.st[m] = ST_PENDING;
                    q.seq[pos] = pos + 1;
                } else if (kind == Q_UQ) {
                    q.st[m] = 1u;
                }
            }
        }
    }
}

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

// ------------------------------------------------- mutex variant (async) ----
// One global ring queue consumed by a persistent kernel.  Every megapixel is
// in the ring at most once and is never visited by two waves at once:
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
        while (a_load(&q.seq[slot]) != t) {
            if (!w.step(q.ctl)) { ok = false; break; }
        }
        if (ok) {
            q.qa[slot] = value;
            a_store(&q.seq[slot], t + 1u);
        }
    }
    return !wave_any(!ok);
}

template <class T, int CONN>
__global__ __launch_bounds__(kWG) void k_queue_mutex(Image im, QueueArgs q) {
    const int lane = lane_id();
    unsigned long long visits = 0;
    while (true) {
        // pop
        uint32_t h = 0;
        if (lane == 0) h = a_add(ctl_at(q.ctl, C_HEAD), 1u);
        h = first_lane(h);
        const uint32_t slot = h & (q.cap - 1u);
        int state = 0;  // 1 = got an entry, 2 = done, 3 = timeout
        if (lane == 0) {
            Waiter w(q.timeout_ticks);
            while (true) {
                if (a_load(&q.seq[slot]) == h + 1u) { state = 1; break; }
                if (a_load(ctl_at(q.ctl, C_PENDING)) == 0u) { state = 2; break; }
                if (a_load_rlx(ctl_at(q.ctl, C_TIMEOUT)) != 0u || !w.step(q.ctl)) { state = 3; break; }
            }
        }
        state = (int)first_lane((uint32_t)state);
        if (state != 1) break;
        uint32_t m = 0;
        if (lane == 0) {
            m = q.qa[slot];
            a_store(&q.seq[slot], h + q.cap);
            a_xchg(&q.st[m], ST_RUNNING);
        }
        m = first_lane(m);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int mx = (int)m % q.nmx, my = (int)m / q.nmx;
        Tile t;
        tile_fetch<T>(im, mx, my, t);
        const bool conv = tile_converge<CONN>(t, q.max_iters);
        tile_store<T, true>(t, im, mx, my);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t nb = tile_notify_mask<CONN>(t, im, mx, my);
        if (!conv) nb |= NB_SELF;
        ++visits;
        // notify neighbours: lane b handles bit b
        bool push = false;
        uint32_t target = 0;
        if (lane < 9 && ((nb >> lane) & 1u)) {
            target = (uint32_t)neighbour_index((int)m, q.nmx, 1u << lane);
            const uint32_t old = a_or(&q.st[target], ST_PENDING);
            push = old == 0u;
        }
        const uint32_t npush = (uint32_t)__popcll(ballot(push));
        if (npush && lane == 0) a_add(ctl_at(q.ctl, C_PENDING), npush);
        bool ok = ring_push_many(q, push, target);
        // finish
        uint32_t old = 0;
        if (lane == 0) old = a_and(&q.st[m], ~ST_RUNNING);
        old = first_lane(old);
        if (old & ST_PENDING) {
            ok = ring_push_many(q, lane == 0, m) && ok;
        } else if (lane == 0) {
            a_add(ctl_at(q.ctl, C_PENDING), 0xFFFFFFFFu);
        }
        if (!ok) break;
    }
    if (lane == 0 && visits) {
        atomicAdd(reinterpret_cast<unsigned long long*>(ctl_at(q.ctl, C_VISITS_LO)), visits);
    }
}

// ------------------------------------------------ DQ / UQ (BSP rounds) ------
// Persistent cooperative kernel.  Round k consumes the read queue (dynamic
// claiming by atomicAdd), appends to the write queue, then a grid barrier,
// a queue flip by one thread, and a second barrier.
DEV bool grid_barrier(uint32_t* ctl, uint32_t nblocks, unsigned long long limit) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __shared__ int ok_s;
    if (threadIdx.x == 0) {
        ok_s = 1;
        const uint32_t gen = a_load(ctl_at(ctl, C_BAR_GEN));
        const uint32_t arrived = a_add(ctl_at(ctl, C_BAR_COUNT), 1u);
        if (arrived == nblocks - 1u) {
            a_store_rlx(ctl_at(ctl, C_BAR_COUNT), 0u);
            a_store(ctl_at(ctl, C_BAR_GEN), gen + 1u);
        } else {
            Waiter w(limit);
            while (a_load(ctl_at(ctl, C_BAR_GEN)) == gen) {
                if (!w.step(ctl)) { ok_s = 0; break; }
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok_s != 0;
}

template <class T, int CONN, int KIND>
__global__ __launch_bounds__(kWG) void k_queue_bsp(Image im, QueueArgs q) {
    constexpr bool EXCL = KIND == Q_UQ;
    const int lane = lane_id();
    unsigned long long visits = 0;
    while (true) {
        const uint32_t round = a_load(ctl_at(q.ctl, C_ROUND));
        const uint32_t nr = a_load(ctl_at(q.ctl, C_RSIZE));
        uint32_t* rq = (round & 1u) ? q.qb : q.qa;
        uint32_t* wq = (round & 1u) ? q.qa : q.qb;
        while (true) {
            uint32_t idx = 0;
            if (lane == 0) idx = a_add(ctl_at(q.ctl, C_CURSOR), 1u);
            idx = first_lane(idx);
            if (idx >= nr) break;
            uint32_t m = 0;
            if (lane == 0) {
                m = rq[idx];
                if (EXCL) a_xchg(&q.st[m], 0u);  // may be queued again for the next round
            }
            m = first_lane(m);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const int mx = (int)m % q.nmx, my = (int)m / q.nmx;
            Tile t;
            tile_fetch<T>(im, mx, my, t);
            const bool conv = tile_converge<CONN>(t, q.max_iters);
            const bool rose = tile_store<T, EXCL>(t, im, mx, my);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            uint32_t nb = tile_notify_mask<CONN>(t, im, mx, my);
            // DQ: a duplicate visitor of the same megapixel may have used stale
            // pixels, so any rise re-queues the megapixel (as the reference does).
            if (!conv || (!EXCL && rose)) nb |= NB_SELF;
            ++visits;
            bool push = false;
            uint32_t target = 0;
            if (lane < 9 && ((nb >> lane) & 1u)) {
                target = (uint32_t)neighbour_index((int)m, q.nmx, 1u << lane);
                push = EXCL ? (a_or(&q.st[target], 1u) == 0u) : true;
            }
            const uint64_t b = ballot(push);
            const uint32_t n = (uint32_t)__popcll(b);
            if (n) {
                uint32_t start = 0;
                if (lane == 0) start = a_add(ctl_at(q.ctl, C_WSIZE), n);
                start = first_lane(start);
                if (push) {
                    const uint32_t pos = start + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
                    if (pos < q.cap) wq[pos] = target;
                    else a_store_rlx(ctl_at(q.ctl, C_OVERFLOW), 1u);
                }
            }
        }
        if (!grid_barrier(q.ctl, gridDim.x, q.timeout_ticks)) break;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            uint32_t ws = a_load(ctl_at(q.ctl, C_WSIZE));
            if (ws > q.cap) ws = q.cap;
            a_store_rlx(ctl_at(q.ctl, C_RSIZE), ws);
            a_store_rlx(ctl_at(q.ctl, C_WSIZE), 0u);
            a_store_rlx(ctl_at(q.ctl, C_CURSOR), 0u);
            a_store(ctl_at(q.ctl, C_ROUND), round + 1u);
        }
        if (!grid_barrier(q.ctl, gridDim.x, q.timeout_ticks)) break;
        if (a_load(ctl_at(q.ctl, C_RSIZE)) == 0u) break;
        if (a_load(ctl_at(q.ctl, C_OVERFLOW)) != 0u) break;
        if (a_load(ctl_at(q.ctl, C_TIMEOUT)) != 0u) break;
        if ((int)round + 1 >= q.max_rounds) break;
    }
    if (lane == 0 && visits) {
        atomicAdd(reinterpret_cast<unsigned long long*>(ctl_at(q.ctl, C_VISITS_LO)), visits);
    }
}

// After seeding: read-queue size / pending count from the seeding count;
// an over-full seeding raises the overflow flag.
__global__ void k_prepare_queue(QueueArgs q) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t n = a_load(ctl_at(q.ctl, C_TAIL));
        const uint32_t k = n < q.cap ? n : q.cap;
        a_store_rlx(ctl_at(q.ctl, C_RSIZE), k);
        a_store_rlx(ctl_at(q.ctl, C_PENDING), k);
        if (n > q.cap) a_store_rlx(ctl_at(q.ctl, C_OVERFLOW), 1u);
        a_store(ctl_at(q.ctl, C_TAIL), k);
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

template <class T, int VB, bool CLIP>
static void row_pass_t(const Image& im, const TileMask& tm, uint32_t* changed, hipStream_t s) {
    const int rowbytes = im.W * (int)sizeof(T);
    const int grid = std::min(im.H, 65535);
    if (rowbytes <= 256 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 1, 4, CLIP>), grid, 256, 0, s, im, tm, changed);
    else if (rowbytes <= 2 * 256 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 2, 4, CLIP>), grid, 256, 0, s, im, tm, changed);
    else if (rowbytes <= 1024 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 1, 16, CLIP>), grid, 1024, 0, s, im, tm, changed);
    else if (rowbytes <= 2 * 1024 * VB)
        hipLaunchKernelGGL((k_row_pass<T, VB, 2, 16, CLIP>), grid, 1024, 0, s, im, tm, changed);
    else {
        hipLaunchKernelGGL((k_row_pass_long<T, VB, 1, CLIP>), grid, kWG, 0, s, im, tm, changed);
        hipLaunchKernelGGL((k_row_pass_long<T, VB, -1, false>), grid, kWG, 0, s, im, tm, changed);
    }
}

hipError_t launch_row_pass(const Image& im, int dtype, bool clip, const TileMask& tm,
                           uint32_t* changed, hipStream_t s) {
    const bool v16 = aligned16(im.J, im.pj) && aligned16(im.I, im.pi);
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

int col_block_width(const void* src, const void* dst, const void* mask, size_t ps, size_t pd,
                    size_t pi, int dtype) {
    const bool v16 = aligned16(src, ps) && aligned16(dst, pd) && aligned16(mask, pi);
    return 64 * (v16 ? 16 : 4) / dtype;
}

template <class T, int VB>
static void col_pass_t(const uint8_t* src, uint8_t* dst, const uint8_t* mask, size_t ps,
                       size_t pd, size_t pi, int W, int H, int conn, int dir, int R,
                       const TileMask& tm, uint32_t* changed, hipStream_t s) {
    const int rowbytes = W * (int)sizeof(T);
    const int strips = (rowbytes + 64 * VB - 1) / (64 * VB);
    const dim3 grid((strips + 3) / 4, (H + R - 1) / R);
#define IWPP_COL(C, D)                                                                      \
    hipLaunchKernelGGL((k_col_pass<T, VB, C, D>), grid, kWG, 0, s, src, dst, mask, ps, pd, pi, \
                       W, H, R, tm, changed)
    if (conn == 8) {
        if (dir > 0) IWPP_COL(8, 1);
        else IWPP_COL(8, -1);
    } else {
        if (dir > 0) IWPP_COL(4, 1);
        else IWPP_COL(4, -1);
    }
#undef IWPP_COL
}

hipError_t launch_col_pass(const uint8_t* src, uint8_t* dst, const uint8_t* mask, size_t ps,
                           size_t pd, size_t pi, int W, int H, int dtype, int conn, int dir,
                           int R, const TileMask& tm, uint32_t* changed, hipStream_t s) {
    const bool v16 = aligned16(src, ps) && aligned16(dst, pd) && aligned16(mask, pi);
    if (dtype == 1) {
        if (v16) col_pass_t<uint8_t, 16>(src, dst, mask, ps, pd, pi, W, H, conn, dir, R, tm, changed, s);
        else col_pass_t<uint8_t, 4>(src, dst, mask, ps, pd, pi, W, H, conn, dir, R, tm, changed, s);
    } else {
        if (v16) col_pass_t<uint16_t, 16>(src, dst, mask, ps, pd, pi, W, H, conn, dir, R, tm, changed, s);
        else col_pass_t<uint16_t, 4>(src, dst, mask, ps, pd, pi, W, H, conn, dir, R, tm, changed, s);
    }
    return hipGetLastError();
}

hipError_t launch_unstable(const Image& im, int dtype, int conn, uint8_t* flags,
                           uint32_t* count, hipStream_t s) {
    const int nmx = (im.W + kMP - 1) / kMP, nmy = (im.H + kMP - 1) / kMP;
    const int nmp = nmx * nmy;
    const int grid = std::min((nmp + 3) / 4, cu_count() * 8);
#define IWPP_UN(TT, C) \
    hipLaunchKernelGGL((k_unstable<TT, C>), grid, kWG, 0, s, im, nmx, nmy, flags, count)
    if (dtype == 1) {
        if (conn == 8) IWPP_UN(uint8_t, 8);
        else IWPP_UN(uint8_t, 4);
    } else {
        if (conn == 8) IWPP_UN(uint16_t, 8);
        else IWPP_UN(uint16_t, 4);
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

hipError_t launch_iota_u32(uint32_t* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const int grid = (int)std::min<size_t>((n + kWG - 1) / kWG, (size_t)cu_count() * 8);
    hipLaunchKernelGGL(k_iota_u32, grid, kWG, 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_prepare_queue(const QueueArgs& q, hipStream_t s) {
    hipLaunchKernelGGL(k_prepare_queue, 1, 64, 0, s, q);
    return hipGetLastError();
}

hipError_t launch_seed_queue(const uint8_t* flags, int nmp, int kind, const QueueArgs& q,
                             hipStream_t s) {
    const int grid = std::min((nmp + kWG - 1) / kWG, cu_count() * 8);
    hipLaunchKernelGGL(k_seed_queue, std::max(grid, 1), kWG, 0, s, flags, nmp, kind, q);
    return hipGetLastError();
}

template <class T, int CONN>
static hipError_t queue_phase_t(const Image& im, int kind, const QueueArgs& q, int wpc,
                                hipStream_t s) {
    const void* fn;
    if (kind == Q_MUTEX) fn = (const void*)k_queue_mutex<T, CONN>;
    else if (kind == Q_DQ) fn = (const void*)k_queue_bsp<T, CONN, Q_DQ>;
    else fn = (const void*)k_queue_bsp<T, CONN, Q_UQ>;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWG, 0);
    if (e != hipSuccess) return e;
    if (per_cu <= 0) return hipErrorInvalidConfiguration;
    if (wpc > 0) per_cu = std::min(per_cu, wpc);
    int grid = per_cu * cu_count();
    const int nmp = q.nmx * q.nmy;
    grid = std::max(1, std::min(grid, (nmp + 3) / 4));
    Image imc = im;
    QueueArgs qc = q;
    if (kind == Q_MUTEX) {
        hipLaunchKernelGGL((k_queue_mutex<T, CONN>), grid, kWG, 0, s, imc, qc);
        return hipGetLastError();
    }
    void* args[] = {&imc, &qc};
    return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kWG), args, 0, s);
}

hipError_t launch_queue_phase(const Image& im, int dtype, int conn, int kind, const QueueArgs& q,
                              int wpc, hipStream_t s) {
    if (dtype == 1) {
        return conn == 8 ? queue_phase_t<uint8_t, 8>(im, kind, q, wpc, s)
                         : queue_phase_t<uint8_t, 4>(im, kind, q, wpc, s);
    }
    return conn == 8 ? queue_phase_t<uint16_t, 8>(im, kind, q, wpc, s)
                     : queue_phase_t<uint16_t, 4>(im, kind, q, wpc, s);
}

}  // namespace iwpp
