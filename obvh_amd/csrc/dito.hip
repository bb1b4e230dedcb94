// dito.hip -- DiTO-14 oriented bounding boxes of point sets on gfx950.
//
// Method: Larsson & Kallberg DiTO-14 exactly as the reference CPU code
// computes it (src/common/algorithm/dito/dito.cpp:478-538), plus the 0.001
// extent floor of kr_points_obb (src/common/geometry.cpp:249).  The GPU
// driver of the reference (dito.cu:614-744: four synchronised launches,
// float CAS atomics per point, last-writer index selection) is replaced by:
//
//  setup     one workgroup scans the set sizes into a tile table (a set is
//            cut into tiles of `tile` points) and writes the result of empty
//            sets; a grid kernel fills tile -> set.
//  pass 1    workgroups (4 x wave64) take tiles in grid-stride order (one
//            tile per workgroup by default).  Lane l of a
//            wave reads points chunk + j*64 + l, j < 8 (coalesced SoA or AoS),
//            forms the 7 slab projections, and keeps per slab side the running
//            extreme with the index where it was first reached: one strict
//            compare and two selects per point and side (a group-extreme
//            scheme that searched the group on improvement ran nearly every
//            search, since some lane of a wave improves on most groups, and
//            issued 3.5x the VALU work of pass 2).  (value, index) pairs reduce
//            lexicographically (ties keep the lower index): order free, and
//            equal to the reference's sequential strict-compare scan.
//            Output: 14 pairs per tile.
//  cand      one wave per set reduces its tile partials, then runs the
//            candidate search (base triangle + ditetrahedron, up to 21 axis
//            triples) over a 14-entry point array held in registers: every
//            lane computes the common prefix, 21 lanes evaluate one axis
//            triple each, and a (quality, id) minimum picks the winner.
//  pass 2    the same tiles project every point on the 3 chosen axes (fused
//            min/max of 3 dot products), 6 floats per tile.
//  final     one wave per set reduces and finalises the box.
//
// No atomics or cross-workgroup fences: kernel boundaries order the phases.
// Each pass reads 12 B per point, 24 B per point per fit.
#include <float.h>

#include "common.hpp"
#include "dito.hpp"

namespace obvh {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kG = 8;               // points per lane per chunk
constexpr int kChunk = kWave * kG;  // points per wave per chunk
constexpr uint32_t kNoIdx = 0xFFFFFFFFu;

struct PtsSoA {
    const float* x;
    const float* y;
    const float* z;
    __device__ __forceinline__ f3 load(uint64_t i) const { return f3{x[i], y[i], z[i]}; }
};
struct PtsAoS {
    const float* xyz;
    __device__ __forceinline__ f3 load(uint64_t i) const {
        return f3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    }
};

__device__ __forceinline__ uint64_t set_off(const uint64_t* offs, uint64_t single, uint32_t s) {
    return offs ? offs[s] : (s ? single : 0);
}

// ---------------------------------------------------------------- setup --

// A set whose offsets decrease or run past `total` is invalid: it gets no
// tiles, a NaN box and degenerate = 4, so no later pass reads or writes
// outside the point arrays or the carved tile arrays (`cap_tiles` entries).
__device__ __forceinline__ bool set_valid(uint64_t b, uint64_t e, uint64_t total) { return b <= e && e <= total; }

__global__ __launch_bounds__(1024) void dito_setup(const uint64_t* offs, uint64_t single_count,
                                                   uint32_t nsets, uint32_t tile, uint64_t total,
                                                   uint32_t cap_tiles, DitoWs ws,
                                                   kr_obb3* out, obvh_dito_trace* trace) {
    __shared__ uint32_t carry;
    __shared__ uint32_t wsum[1024 / kWave];
    if (threadIdx.x == 0) carry = 0;
    if (threadIdx.x == 0 && offs == nullptr) {
        ws.offsets_single[0] = 0;
        ws.offsets_single[1] = single_count;
    }
    __syncthreads();
    for (uint32_t base = 0; base < nsets; base += blockDim.x) {
        const uint32_t s = base + threadIdx.x;
        uint32_t nt = 0;
        if (s < nsets) {
            const uint64_t b = set_off(offs, single_count, s), e = set_off(offs, single_count, s + 1);
            const uint64_t n = e - b;
            if (!set_valid(b, e, total)) {
                const float nan = __builtin_nanf("");
                kr_obb3 o;
                o.mid = o.v0 = o.v1 = o.v2 = o.ext = kr_vec3{nan, nan, nan};
                out[s] = o;
                obvh_dito_trace t = {};
                t.pair_k = t.p2_sel = t.q0_sel = t.q1_sel = -1;
                t.best = -1;
                t.degenerate = 4;
                trace[s] = t;
            } else if (n != 0) {
                nt = (uint32_t)((n + tile - 1) / tile);
            } else {
                // DiTO_14 of an empty set (dito.cpp:498-502) + extent floor.
                kr_obb3 o;
                o.mid = kr_vec3{0, 0, 0};
                o.v0 = kr_vec3{1, 0, 0};
                o.v1 = kr_vec3{0, 1, 0};
                o.v2 = kr_vec3{0, 0, 1};
                o.ext = kr_vec3{0.001f, 0.001f, 0.001f};
                out[s] = o;
                obvh_dito_trace t = {};
                t.pair_k = t.p2_sel = t.q0_sel = t.q1_sel = -1;
                t.best = -1;
                t.degenerate = 3;
                t.result_aabb = 1;
                trace[s] = t;
            }
        }
        uint32_t v = nt;
        const int lane = lane_id();
#pragma unroll
        for (int m = 1; m < kWave; m <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)v, m, kWave);
            if (lane >= m) v += o;
        }
        if (lane == kWave - 1) wsum[wave_id()] = v;
        __syncthreads();
        uint32_t wpre = 0;
        for (int w = 0; w < wave_id(); ++w) wpre += wsum[w];
        const uint32_t incl = carry + wpre + v;
        if (s < nsets) ws.first_tile[s] = incl - nt;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) carry = incl;
        __syncthreads();
    }
    // valid offsets need at most cap_tiles - 1 tiles; the clamp only guards
    // the carved arrays
    if (threadIdx.x == 0) ws.first_tile[nsets] = carry < cap_tiles ? carry : cap_tiles;
}

__global__ __launch_bounds__(256) void dito_tile_table(uint32_t nsets, DitoWs ws) {
    const uint32_t T = ws.first_tile[nsets];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nsets;  // first_tile[lo] <= t < first_tile[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (ws.first_tile[mid] <= t) lo = mid; else hi = mid;
        }
        ws.tile_set[t] = lo;  // last set starting at or before t: skips empty sets
    }
}

// --------------------------------------------------------------- pass 1 --

// Slab extremes of points [base, end) of the set starting at sbeg: 14
// (value, index within the set) pairs written to v_out / i_out by threads
// 0..13 of the workgroup.
template <class P>
__device__ __forceinline__ void tile_pass1(P pts, uint64_t sbeg, uint64_t base, uint64_t end,
                                           float (&sh_v)[kWaves][14], uint32_t (&sh_i)[kWaves][14],
                                           float* v_out, uint32_t* i_out) {
    const int lane = lane_id(), wave = wave_id();
    const float NaN = __builtin_nanf("");
    const uint32_t nchunks = (uint32_t)((end - base + kChunk - 1) / kChunk);
    // per lane and slab side: running extreme and the index where it was
    // first reached (points of one lane are visited in increasing index)
    float rmn[7], rmx[7];
    uint32_t imn[7], imx[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        rmn[k] = NaN;  // NaN never compares: the first real value wins
        rmx[k] = NaN;
        imn[k] = imx[k] = kNoIdx;
    }
    for (uint32_t c = wave; c < nchunks; c += kWaves) {
        const uint64_t cb = base + (uint64_t)c * kChunk + lane;
        const bool full = base + (uint64_t)(c + 1) * kChunk <= end;
        f3 p[kG];
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            const uint64_t idx = cb + (uint64_t)j * kWave;
            p[j] = (full || idx < end) ? pts.load(idx) : f3{NaN, NaN, NaN};
        }
        // Per point and slab side a strict compare against the running
        // extreme keeps the first point that reaches it, as the sequential
        // scan does; !(v >= cur) also takes the first value, while cur is
        // NaN.  NaN values (the short last chunk's missing points, or
        // non-finite input) are skipped here; the candidate step applies the
        // reference scan's rule for a NaN at point 0.
        const uint32_t i0 = (uint32_t)(cb - sbeg);
#pragma unroll
        for (int j = 0; j < kG; ++j) {
            float pr[7];
            slab7(p[j], pr);
            const uint32_t idx = i0 + (uint32_t)j * kWave;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                if (!(pr[k] >= rmn[k]) && pr[k] == pr[k]) { rmn[k] = pr[k]; imn[k] = idx; }
                if (!(pr[k] <= rmx[k]) && pr[k] == pr[k]) { rmx[k] = pr[k]; imx[k] = idx; }
            }
        }
    }
    // wave: (value, index) lexicographic, then the workgroup over waves
#pragma unroll
    for (int side = 0; side < 14; ++side) {
        const int k = side < 7 ? side : side - 7;
        float v = side < 7 ? rmn[k] : rmx[k];
        uint32_t i = side < 7 ? imn[k] : imx[k];
        if (i == kNoIdx) v = side < 7 ? FLT_MAX * 2.0f : -FLT_MAX * 2.0f;
        if (side < 7) wave_pair<true>(v, i); else wave_pair<false>(v, i);
        if (lane == 0) {
            sh_v[wave][side] = v;
            sh_i[wave][side] = i;
        }
    }
    __syncthreads();
    if (threadIdx.x < 14) {
        const int side = threadIdx.x;
        float v = sh_v[0][side];
        uint32_t i = sh_i[0][side];
        for (int w = 1; w < kWaves; ++w) {
            const float ov = sh_v[w][side];
            const uint32_t oi = sh_i[w][side];
            if (oi == kNoIdx) continue;
            const bool better = (i == kNoIdx) ||
                                (side < 7 ? pair_better<true>(v, i, ov, oi) : pair_better<false>(v, i, ov, oi));
            if (better) { v = ov; i = oi; }
        }
        v_out[side] = v;
        i_out[side] = i;
    }
    __syncthreads();
}

template <class P>
__global__ __launch_bounds__(kBlock) void dito_pass1(P pts, const uint64_t* offs, uint32_t nsets,
                                                     uint32_t tile, DitoWs ws) {
    __shared__ float sh_v[kWaves][14];
    __shared__ uint32_t sh_i[kWaves][14];
    const uint32_t T = ws.first_tile[nsets];
    for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
        const uint32_t s = ws.tile_set[t];
        const uint64_t sbeg = offs[s], send = offs[s + 1];
        const uint32_t ft = ws.first_tile[s];
        const uint64_t base = sbeg + (uint64_t)(t - ft) * tile;
        const uint64_t end = base + tile < send ? base + tile : send;
        tile_pass1(pts, sbeg, base, end, sh_v, sh_i, &ws.part1_v[(size_t)t * 14], &ws.part1_i[(size_t)t * 14]);
    }
}

// ----------------------------------------- candidate search (one wave) --
//
// dito.cpp:347-418 over a 14-entry point list `sel`.  Sets of n <= 14 points
// use their n points followed by copies of sel[0]: a copy never wins a strict
// comparison against the first occurrence, so every result equals the
// n-point search.  Candidate ids: 0..2 base triangle, 3..11 upper tetra,
// 12..20 lower tetra.  Every lane computes the common prefix (extreme pair,
// base triangle, normal, apexes); lane c < 21 evaluates candidate c with the
// same operations as the sequential scan.  The scan keeps the first candidate
// whose quality is strictly below the best so far, starting from the AABB's:
// that is the lexicographic minimum of (quality, id) over the candidates,
// kept when it is strictly below the AABB's quality; a NaN quality or an
// absent candidate (apex not valid) never wins.

__device__ __forceinline__ void ext_projs(const f3* sel, f3 d, float& mn, float& mx) {
    float p = dot3(sel[0], d);
    mn = p;
    mx = p;
#pragma unroll
    for (int i = 1; i < 14; ++i) {
        p = dot3(sel[i], d);
        if (p < mn) mn = p;  // dito.cpp:198-202
        if (p > mx) mx = p;
    }
}

// Quality of candidate c (dito.cpp:309-345 for one edge of one triangle), its
// axes in b0..b2; +inf when the candidate does not exist.
__device__ __forceinline__ float eval_candidate(const f3* sel, int c, f3 n, f3 e0, f3 e1, f3 e2, f3 p0, f3 p1,
                                                f3 p2, f3 q0, f3 q1, bool q0v, bool q1v, f3& b0, f3& b1, f3& b2) {
    const int face = c / 3, k = c - 3 * face;
    f3 nf = n, ea = e0, eb = e1, ec = e2;
    if (face > 0) {
        const bool up = face <= 3;
        if (up ? !q0v : !q1v) return FLT_MAX * 2.0f;
        const f3 q = up ? q0 : q1;
        const f3 f0 = dm::normalize(sub3(q, p0));
        const f3 f1 = dm::normalize(sub3(q, p1));
        const f3 f2 = dm::normalize(sub3(q, p2));
        const int fi = (face - 1) % 3;  // dito.cpp:394-417
        if (fi == 0) {
            nf = dm::normalize(dm::cross(f1, e0)); ea = e0; eb = f1; ec = f0;
        } else if (fi == 1) {
            nf = dm::normalize(dm::cross(f2, e1)); ea = e1; eb = f2; ec = f1;
        } else {
            nf = dm::normalize(dm::cross(f0, e2)); ea = e2; eb = f0; ec = f2;
        }
    }
    const f3 e = k == 0 ? ea : (k == 1 ? eb : ec);
    float nmin, nmax, emin, emax, mmin, mmax;
    ext_projs(sel, nf, nmin, nmax);
    const f3 m = dm::cross(e, nf);
    ext_projs(sel, e, emin, emax);
    ext_projs(sel, m, mmin, mmax);
    b0 = e;
    b1 = nf;
    b2 = m;
    return quality(f3{emax - emin, nmax - nmin, mmax - mmin});
}

// Run by every lane of a wave; the results are uniform.
__device__ void dito_candidate_search(const float mn[7], const float mx[7], const f3 minv[7], const f3 maxv[7],
                                      const f3 sel[14], SetState& st, obvh_dito_trace& tr) {
    const float eps = 0.000001f;
    const f3 almid = f3{(mn[0] + mx[0]) * 0.5f, (mn[1] + mx[1]) * 0.5f, (mn[2] + mx[2]) * 0.5f};
    const f3 allen = f3{mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    st.almid[0] = almid.x; st.almid[1] = almid.y; st.almid[2] = almid.z;
    st.allen[0] = allen.x; st.allen[1] = allen.y; st.allen[2] = allen.z;
    st.alval = quality(allen);

    int kstar = 0;
    float dbest = dm::sqdist(maxv[0], minv[0]);
#pragma unroll
    for (int k = 1; k < 7; ++k) {
        const float d = dm::sqdist(maxv[k], minv[k]);
        if (d > dbest) { dbest = d; kstar = k; }
    }
    f3 p0 = minv[0], p1 = maxv[0];
#pragma unroll
    for (int k = 1; k < 7; ++k)
        if (kstar == k) { p0 = minv[k]; p1 = maxv[k]; }
    tr.pair_k = kstar;
    if (dm::sqdist(p0, p1) < eps) {  // degenerate case 1 -> AABB
        st.mode = 1;
        tr.degenerate = 1;
        return;
    }
    const f3 e0 = dm::normalize(sub3(p0, p1));
    int ip2 = 0;
    f3 p2 = sel[0];
    float dmax;
    {
        f3 u0 = sub3(sel[0], p0);
        float t = dot3(e0, u0);
        dmax = len2(u0) - t * t / len2(e0);
#pragma unroll
        for (int i = 1; i < 14; ++i) {
            u0 = sub3(sel[i], p0);
            t = dot3(e0, u0);
            const float d = len2(u0) - t * t / len2(e0);
            if (d > dmax) { dmax = d; ip2 = i; p2 = sel[i]; }
        }
    }
    tr.p2_sel = ip2;
    if (dmax < eps) {  // degenerate case 2 -> line-aligned box (dito.cpp:451-476)
        const f3 u = e0;
        f3 r = u;
        if (fabsf(u.x) > fabsf(u.y) && fabsf(u.x) > fabsf(u.z)) r.x = 0;
        else if (fabsf(u.y) > fabsf(u.z)) r.y = 0;
        else r.z = 0;
        if (len2(r) < eps) r = f3{1, 1, 1};
        const f3 v = dm::normalize(dm::cross(u, r));
        const f3 w = dm::normalize(dm::cross(u, v));
        st.ax[0] = u.x; st.ax[1] = u.y; st.ax[2] = u.z;
        st.ax[3] = v.x; st.ax[4] = v.y; st.ax[5] = v.z;
        st.ax[6] = w.x; st.ax[7] = w.y; st.ax[8] = w.z;
        st.mode = 2;
        tr.degenerate = 2;
        return;
    }
    const f3 e1 = dm::normalize(sub3(p1, p2));
    const f3 e2 = dm::normalize(sub3(p2, p0));
    const f3 n = dm::normalize(dm::cross(e1, e0));

    float pr = dot3(sel[0], n);
    float lo = pr, hi = pr;
    int ilo = 0, ihi = 0;
    f3 q1 = sel[0], q0 = sel[0];
#pragma unroll
    for (int i = 1; i < 14; ++i) {
        pr = dot3(sel[i], n);
        if (pr < lo) { lo = pr; ilo = i; q1 = sel[i]; }
        if (pr > hi) { hi = pr; ihi = i; q0 = sel[i]; }
    }
    const float tri = dot3(p0, n);
    const int q0v = (hi - eps > tri) ? 1 : 0;
    const int q1v = (lo + eps < tri) ? 1 : 0;
    tr.q0_sel = ihi;
    tr.q1_sel = ilo;
    tr.q0_valid = q0v;
    tr.q1_valid = q1v;

    const int lane = lane_id();
    const float INF = FLT_MAX * 2.0f;
    f3 b0 = f3{1, 0, 0}, b1 = f3{0, 1, 0}, b2 = f3{0, 0, 1};
    float q = INF;
    if (lane < 21) q = eval_candidate(sel, lane, n, e0, e1, e2, p0, p1, p2, q0, q1, q0v != 0, q1v != 0, b0, b1, b2);
    if (!(q == q)) q = INF;  // NaN never compares below anything
    int id = lane;
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        const float oq = __shfl_xor(q, m, kWave);
        const int oid = __shfl_xor(id, m, kWave);
        if (oq < q || (oq == q && oid < id)) { q = oq; id = oid; }
    }
    int best_id = -1;
    if (q < st.alval) {  // strictly better than the AABB (and than every earlier candidate)
        best_id = id;
        b0 = f3{__shfl(b0.x, id), __shfl(b0.y, id), __shfl(b0.z, id)};
        b1 = f3{__shfl(b1.x, id), __shfl(b1.y, id), __shfl(b1.z, id)};
        b2 = f3{__shfl(b2.x, id), __shfl(b2.y, id), __shfl(b2.z, id)};
    } else {
        b0 = f3{1, 0, 0};
        b1 = f3{0, 1, 0};
        b2 = f3{0, 0, 1};
    }
    st.ax[0] = b0.x; st.ax[1] = b0.y; st.ax[2] = b0.z;
    st.ax[3] = b1.x; st.ax[4] = b1.y; st.ax[5] = b1.z;
    st.ax[6] = b2.x; st.ax[7] = b2.y; st.ax[8] = b2.z;
    st.mode = 0;
    tr.degenerate = 0;
    tr.best = best_id;
}

// One wave: reduce the nt tile partials of a set (points [sbeg, send),
// first tile ft), then search; state and trace written by lane 0.
template <class P>
__device__ __forceinline__ void set_cand(P pts, uint64_t sbeg, uint64_t send, uint32_t ft, uint32_t nt,
                                         const float* part1_v, const uint32_t* part1_i, SetState* state,
                                         obvh_dito_trace* trace) {
    const int lane = lane_id();
    // Lane l reduces side l % 16 (sides 0..13) over the tiles j = l / 16
    // (mod 4); two shuffle steps join the four groups.  (value, index) pairs
    // order lexicographically (no NaN values: pass 1 never takes one), so
    // the order of the reduction does not matter.
    float v[14];
    uint32_t ix[14];
    const f3 pt0 = pts.load(sbeg);  // for the NaN rule below, read beside the partials
    {
        const int side = lane & 15;
        const bool mn = side < 7;
        float a = mn ? FLT_MAX * 2.0f : -FLT_MAX * 2.0f;
        uint32_t ai = kNoIdx;
        if (side < 14) {
            // sixteen tiles' pairs are read before any is compared: one
            // memory round trip per 64 tiles instead of one per 4
            constexpr int kB = 16;
            for (uint32_t j0 = (uint32_t)(lane >> 4); j0 < nt; j0 += 4 * kB) {
                float ov[kB];
                uint32_t oi[kB];
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    const uint32_t j = j0 + 4u * u;
                    ov[u] = j < nt ? part1_v[(size_t)(ft + j) * 14 + side] : 0.0f;
                    oi[u] = j < nt ? part1_i[(size_t)(ft + j) * 14 + side] : kNoIdx;
                }
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    if (j0 + 4u * u >= nt) continue;
                    const bool better = (ai == kNoIdx) || (mn ? pair_better<true>(a, ai, ov[u], oi[u])
                                                              : pair_better<false>(a, ai, ov[u], oi[u]));
                    if (better) { a = ov[u]; ai = oi[u]; }
                }
            }
        }
#pragma unroll
        for (int m = 16; m < kWave; m <<= 1) {
            const float ov = __shfl_xor(a, m, kWave);
            const uint32_t oi = (uint32_t)__shfl_xor((int)ai, m, kWave);
            if (mn ? pair_better<true>(a, ai, ov, oi) : pair_better<false>(a, ai, ov, oi)) { a = ov; ai = oi; }
        }
#pragma unroll
        for (int k = 0; k < 14; ++k) {
            v[k] = __shfl(a, k, kWave);
            ix[k] = (uint32_t)__shfl((int)ai, k, kWave);
        }
    }
    const uint64_t n = send - sbeg;
    // The reference scan starts from point 0 and compares strictly
    // (dito.cpp:116-173): a slab whose value at point 0 is NaN stays (NaN, 0),
    // since no later compare against NaN succeeds; otherwise NaN values are
    // never taken and the result is the (value, first index) minimum over
    // the other points, which is what the reductions above computed.
    {
        float pr0[7];
        slab7(pt0, pr0);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if (pr0[k] != pr0[k]) {
                v[k] = v[7 + k] = pr0[k];
                ix[k] = ix[7 + k] = 0;
            }
        }
    }
#pragma unroll
    for (int side = 0; side < 14; ++side)
        if ((uint64_t)ix[side] >= n) ix[side] = 0;  // not reachable (a slab that is NaN everywhere is NaN at point 0)
    obvh_dito_trace tr = {};
    tr.pair_k = tr.p2_sel = tr.q0_sel = tr.q1_sel = -1;
    tr.best = -1;
    float mn[7], mx[7];
    f3 minv[7], maxv[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        mn[k] = v[k];
        mx[k] = v[7 + k];
        tr.slab_min_idx[k] = ix[k];
        tr.slab_max_idx[k] = ix[7 + k];
        minv[k] = pts.load(sbeg + ix[k]);
        maxv[k] = pts.load(sbeg + ix[7 + k]);
    }
    f3 sel[14];
    if (n > 14) {
#pragma unroll
        for (int k = 0; k < 7; ++k) { sel[k] = minv[k]; sel[7 + k] = maxv[k]; }
    } else {
        sel[0] = pts.load(sbeg);
#pragma unroll
        for (int i = 1; i < 14; ++i) sel[i] = (uint64_t)i < n ? pts.load(sbeg + i) : sel[0];
    }
    SetState st = {};
    dito_candidate_search(mn, mx, minv, maxv, sel, st, tr);
    if (lane == 0) {
        *state = st;
        *trace = tr;
    }
}

// One wave per set: reduce the tile partials, then the wave searches.
template <class P>
__global__ __launch_bounds__(kBlock) void dito_cand(P pts, const uint64_t* offs, uint32_t nsets, uint64_t total,
                                                    DitoWs ws, obvh_dito_trace* trace) {
    const uint32_t s = blockIdx.x * kWaves + wave_id();
    if (s >= nsets) return;
    const uint64_t sbeg = offs[s], send = offs[s + 1];
    if (send == sbeg || !set_valid(sbeg, send, total)) return;
    const uint32_t ft = ws.first_tile[s], nt = ws.first_tile[s + 1] - ft;
    set_cand(pts, sbeg, send, ft, nt, ws.part1_v, ws.part1_i, &ws.state[s], &trace[s]);
}

// --------------------------------------------------------------- pass 2 --

// Ranges of the points [base, end) on the set's three axes: 6 floats (min
// b0..b2, max b0..b2) written to out6 by threads 0..5.
template <class P>
__device__ __forceinline__ void tile_pass2(P pts, uint64_t base, uint64_t end, const SetState* st,
                                           float (&sh)[kWaves][6], float* out6) {
    const int lane = lane_id(), wave = wave_id();
    const float INF = FLT_MAX * 2.0f, NaN = __builtin_nanf("");
    float r[6] = {INF, INF, INF, -INF, -INF, -INF};
    if (st->mode != 1) {
        const f3 b0 = f3{st->ax[0], st->ax[1], st->ax[2]};
        const f3 b1 = f3{st->ax[3], st->ax[4], st->ax[5]};
        const f3 b2 = f3{st->ax[6], st->ax[7], st->ax[8]};
        const uint32_t nchunks = (uint32_t)((end - base + kChunk - 1) / kChunk);
        for (uint32_t c = wave; c < nchunks; c += kWaves) {
            const uint64_t cb = base + (uint64_t)c * kChunk + lane;
            const bool full = base + (uint64_t)(c + 1) * kChunk <= end;
            f3 p[kG];
#pragma unroll
            for (int j = 0; j < kG; ++j) {
                const uint64_t idx = cb + (uint64_t)j * kWave;
                p[j] = (full || idx < end) ? pts.load(idx) : f3{NaN, NaN, NaN};
            }
#pragma unroll
            for (int j = 0; j < kG; ++j) {
                const float d0 = dot3(p[j], b0), d1 = dot3(p[j], b1), d2 = dot3(p[j], b2);
                r[0] = fminf(r[0], d0); r[3] = fmaxf(r[3], d0);
                r[1] = fminf(r[1], d1); r[4] = fmaxf(r[4], d1);
                r[2] = fminf(r[2], d2); r[5] = fmaxf(r[5], d2);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { r[k] = wave_min(r[k]); r[3 + k] = wave_max(r[3 + k]); }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) sh[wave][k] = r[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float a = sh[0][k];
        for (int w = 1; w < kWaves; ++w) a = k < 3 ? fminf(a, sh[w][k]) : fmaxf(a, sh[w][k]);
        out6[k] = a;
    }
    __syncthreads();
}

template <class P>
__global__ __launch_bounds__(kBlock) void dito_pass2(P pts, const uint64_t* offs, uint32_t nsets,
                                                     uint32_t tile, DitoWs ws) {
    __shared__ float sh[kWaves][6];
    const uint32_t T = ws.first_tile[nsets];
    for (uint32_t t = blockIdx.x; t < T; t += gridDim.x) {
        const uint32_t s = ws.tile_set[t];
        const uint64_t sbeg = offs[s], send = offs[s + 1];
        const uint32_t ft = ws.first_tile[s];
        const uint64_t base = sbeg + (uint64_t)(t - ft) * tile;
        const uint64_t end = base + tile < send ? base + tile : send;
        tile_pass2(pts, base, end, &ws.state[s], sh, &ws.part2[(size_t)t * 6]);
    }
}

// ---------------------------------------------------------------- final --

// One wave: reduce the nt tile ranges of a set (first tile ft) and write
// its box (lane 0).
template <class P>
__device__ __forceinline__ void set_final(P pts, uint64_t sbeg, uint32_t ft, uint32_t nt, const float* part2,
                                          const SetState* state, kr_obb3* out, obvh_dito_trace* trace) {
    const float INF = FLT_MAX * 2.0f;
    float a[6] = {INF, INF, INF, -INF, -INF, -INF};
    for (uint32_t j = lane_id(); j < nt; j += kWave) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float o = part2[(size_t)(ft + j) * 6 + k];
            a[k] = k < 3 ? fminf(a[k], o) : fmaxf(a[k], o);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { a[k] = wave_min(a[k]); a[3 + k] = wave_max(a[3 + k]); }
    if (lane_id() != 0) return;
    const SetState st = *state;
    const f3 b0 = f3{st.ax[0], st.ax[1], st.ax[2]};
    const f3 b1 = f3{st.ax[3], st.ax[4], st.ax[5]};
    const f3 b2 = f3{st.ax[6], st.ax[7], st.ax[8]};
    // computeObbDimensions (dito.cpp:421-426) starts from point 0 with strict
    // compares: a NaN projection of point 0 stays, later NaNs are skipped
    // (fminf / fmaxf above skip them too).
    if (st.mode != 1) {
        const f3 p0 = pts.load(sbeg);
        const float d[3] = {dot3(p0, b0), dot3(p0, b1), dot3(p0, b2)};
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (d[k] != d[k]) a[k] = a[3 + k] = d[k];
    }
    const f3 bmn = f3{a[0], a[1], a[2]}, bmx = f3{a[3], a[4], a[5]};
    const f3 blen = sub3(bmx, bmn);
    kr_obb3 o;
    int result_aabb;
    if (st.mode == 2 || (st.mode == 0 && quality(blen) < st.alval)) {
        // finalizeOBB, dito.cpp:428-440
        const f3 ext = dm::scale(0.5f, blen);
        const f3 q = dm::scale(0.5f, add3(bmn, bmx));
        f3 mid = dm::scale(q.x, b0);
        mid = add3(mid, dm::scale(q.y, b1));
        mid = add3(mid, dm::scale(q.z, b2));
        o.mid = kr_vec3{mid.x, mid.y, mid.z};
        o.v0 = kr_vec3{b0.x, b0.y, b0.z};
        o.v1 = kr_vec3{b1.x, b1.y, b1.z};
        o.v2 = kr_vec3{b2.x, b2.y, b2.z};
        o.ext = kr_vec3{ext.x, ext.y, ext.z};
        result_aabb = 0;
    } else {  // finalizeAxisAlignedOBB, dito.cpp:442-449
        const f3 e = dm::scale(0.5f, f3{st.allen[0], st.allen[1], st.allen[2]});
        o.mid = kr_vec3{st.almid[0], st.almid[1], st.almid[2]};
        o.v0 = kr_vec3{1, 0, 0};
        o.v1 = kr_vec3{0, 1, 0};
        o.v2 = kr_vec3{0, 0, 1};
        o.ext = kr_vec3{e.x, e.y, e.z};
        result_aabb = 1;
    }
    // kr_points_obb extent floor (geometry.cpp:249)
    o.ext.x = kmax(o.ext.x, 0.001f);
    o.ext.y = kmax(o.ext.y, 0.001f);
    o.ext.z = kmax(o.ext.z, 0.001f);
    *out = o;
    trace->result_aabb = result_aabb;
}

template <class P>
__global__ __launch_bounds__(kBlock) void dito_final(P pts, const uint64_t* offs, uint32_t nsets, uint64_t total,
                                                     DitoWs ws, kr_obb3* out, obvh_dito_trace* trace) {
    const uint32_t s = blockIdx.x * kWaves + wave_id();
    if (s >= nsets) return;
    const uint64_t sbeg = offs[s], send = offs[s + 1];
    if (send == sbeg || !set_valid(sbeg, send, total)) return;
    const uint32_t ft = ws.first_tile[s], nt = ws.first_tile[s + 1] - ft;
    set_final(pts, sbeg, ft, nt, ws.part2, &ws.state[s], &out[s], &trace[s]);
}

// ------------------------------------------------- one set, two launches --
//
// A single set of n > 0 points in tiles of `tile` points, one workgroup per
// tile (grid = tile count).  Launch A: slab extremes per tile; the last
// workgroup to finish (completion counter done[0]) reduces them and runs the
// candidate search.  Launch B: axis ranges per tile; the last workgroup
// (done[1]) reduces them and writes the box.  A workgroup publishes its tile
// result with one agent-scope release (the atomic on the counter); the last
// one acquires before reading the others'.  The counters must be zero before
// the first call on a workspace (dito_one_clear, or a clear at allocation);
// each last workgroup resets its counter, so later calls need no clear.

__global__ void dito_one_clear(DitoWs ws) {
    if (threadIdx.x < 2) ws.done[threadIdx.x] = 0;
}

template <class P>
__global__ __launch_bounds__(kBlock) void dito_one_a(P pts, uint64_t n, uint32_t tile, DitoWs ws,
                                                     obvh_dito_trace* trace) {
    __shared__ float sh_v[kWaves][14];
    __shared__ uint32_t sh_i[kWaves][14];
    __shared__ uint32_t last;
    const uint32_t t = blockIdx.x, T = gridDim.x;
    const uint64_t base = (uint64_t)t * tile, end = base + tile < n ? base + tile : n;
    // threads 0..13 (wave 0) write the tile's pairs; thread 0's release
    // orders them before its count
    tile_pass1(pts, 0, base, end, sh_v, sh_i, &ws.part1_v[(size_t)t * 14], &ws.part1_i[(size_t)t * 14]);
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(&ws.done[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == T - 1;
    __syncthreads();
    if (!last || wave_id() != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    set_cand(pts, 0, n, 0, T, ws.part1_v, ws.part1_i, ws.state, trace);
    if (lane_id() == 0) __hip_atomic_store(&ws.done[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class P>
__global__ __launch_bounds__(kBlock) void dito_one_b(P pts, uint64_t n, uint32_t tile, DitoWs ws, kr_obb3* out,
                                                     obvh_dito_trace* trace) {
    __shared__ float sh[kWaves][6];
    __shared__ uint32_t last;
    const uint32_t t = blockIdx.x, T = gridDim.x;
    const uint64_t base = (uint64_t)t * tile, end = base + tile < n ? base + tile : n;
    tile_pass2(pts, base, end, ws.state, sh, &ws.part2[(size_t)t * 6]);  // threads 0..5 write
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(&ws.done[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == T - 1;
    __syncthreads();
    if (!last || wave_id() != 0) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    set_final(pts, 0, 0, T, ws.part2, ws.state, out, trace);
    if (lane_id() == 0) __hip_atomic_store(&ws.done[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ----------------------------------------------------------------- host --

size_t dito_max_tiles(uint32_t nsets, uint64_t total, uint32_t tile) {
    return (size_t)(total / tile) + nsets + 1;
}

size_t dito_carve(DitoWs* ws, void* base, size_t cap, uint32_t nsets, uint64_t total, uint32_t tile) {
    Carver c{(char*)base, 0, cap};
    const size_t T = dito_max_tiles(nsets, total, tile);
    ws->offsets_single = c.take<uint64_t>(2);
    ws->done = c.take<uint32_t>(2);
    ws->first_tile = c.take<uint32_t>(nsets + 1);
    ws->tile_set = c.take<uint32_t>(T);
    ws->part1_v = c.take<float>(T * 14);
    ws->part1_i = c.take<uint32_t>(T * 14);
    ws->part2 = c.take<float>(T * 6);
    ws->state = c.take<SetState>(nsets);
    ws->trace_tmp = c.take<obvh_dito_trace>(nsets);
    return c.used;
}

uint32_t dito_auto_tile(uint32_t nsets, uint64_t total) {
    // about 4096 tiles for the whole job, between 512 and 65536 points each
    (void)nsets;
    uint64_t want = total / 4096;
    uint64_t t = (want + kChunk - 1) / kChunk * kChunk;
    if (t < (uint64_t)kChunk) t = kChunk;
    if (t > 65536) t = 65536;
    return (uint32_t)t;
}

template <class P>
static void launch_one(P p, uint64_t n, uint32_t tile, uint32_t T, DitoWs ws, kr_obb3* out, obvh_dito_trace* tr,
                       bool clear, hipStream_t stream) {
    if (clear) hipLaunchKernelGGL(dito_one_clear, dim3(1), dim3(64), 0, stream, ws);
    hipLaunchKernelGGL(dito_one_a<P>, dim3(T), dim3(kBlock), 0, stream, p, n, tile, ws, tr);
    hipLaunchKernelGGL(dito_one_b<P>, dim3(T), dim3(kBlock), 0, stream, p, n, tile, ws, out, tr);
}

kr_error dito_one_launch(const float* x, const float* y, const float* z, const float* xyz, uint64_t n,
                         kr_obb3* out, obvh_dito_trace* trace, void* workspace, size_t ws_bytes, bool clear,
                         hipStream_t stream) {
    if (n == 0) return "obvh_dito: the two-launch fit needs at least one point";
    // Tiles of 2048 points (one chunk per wave) up to 1024 tiles, larger
    // beyond: each workgroup pays one agent-scope release and the last one
    // reduces every tile, so a small set wants few tiles.  At least the
    // batched path's tile, so the workspace size of obvh_dito_workspace_size
    // always suffices.
    uint64_t tl = (n + 1023) / 1024;
    tl = (tl + kChunk - 1) / kChunk * kChunk;
    const uint32_t tile = (uint32_t)max(max(tl, (uint64_t)4 * kChunk), (uint64_t)dito_auto_tile(1, n));
    DitoWs ws;
    const size_t need = dito_carve(&ws, workspace, ws_bytes, 1, n, tile);
    if (workspace == nullptr || ws_bytes < need) return "obvh_dito: workspace too small";
    const uint64_t T = (n + tile - 1) / tile;
    if (T > dito_max_tiles(1, n, tile)) return "obvh_dito: tile count";
    obvh_dito_trace* tr = trace ? trace : ws.trace_tmp;
    if (xyz) launch_one(PtsAoS{xyz}, n, tile, (uint32_t)T, ws, out, tr, clear, stream);
    else launch_one(PtsSoA{x, y, z}, n, tile, (uint32_t)T, ws, out, tr, clear, stream);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

template <class P>
static void launch_passes(P p, const uint64_t* offs, uint32_t nsets, uint64_t total, uint32_t tile, uint32_t workers,
                          DitoWs ws, kr_obb3* out, obvh_dito_trace* tr, hipStream_t stream) {
    const dim3 set_grid((nsets + kWaves - 1) / kWaves);
    hipLaunchKernelGGL(dito_pass1<P>, dim3(workers), dim3(kBlock), 0, stream, p, offs, nsets, tile, ws);
    hipLaunchKernelGGL(dito_cand<P>, set_grid, dim3(kBlock), 0, stream, p, offs, nsets, total, ws, tr);
    hipLaunchKernelGGL(dito_pass2<P>, dim3(workers), dim3(kBlock), 0, stream, p, offs, nsets, tile, ws);
    hipLaunchKernelGGL(dito_final<P>, set_grid, dim3(kBlock), 0, stream, p, offs, nsets, total, ws, out, tr);
}

kr_error dito_launch(const float* x, const float* y, const float* z, const float* xyz,
                     const uint64_t* offsets, uint64_t single_count, uint32_t nsets, uint64_t total,
                     kr_obb3* out, obvh_dito_trace* trace, uint32_t tile, uint32_t workers,
                     void* workspace, size_t ws_bytes, hipStream_t stream) {
    if (nsets == 0) return nullptr;
    if (tile == 0) tile = dito_auto_tile(nsets, total);
    if (tile % kChunk != 0) return "obvh_dito: tile_points must be a multiple of 512";
    DitoWs ws;
    const size_t need = dito_carve(&ws, workspace, ws_bytes, nsets, total, tile);
    if (workspace == nullptr || ws_bytes < need) return "obvh_dito: workspace too small";
    const uint64_t* offs = offsets ? offsets : ws.offsets_single;
    obvh_dito_trace* tr = trace ? trace : ws.trace_tmp;
    const size_t T = dito_max_tiles(nsets, total, tile);
    // default: one workgroup per tile (measured faster than persistent
    // workgroups looping over tiles)
    if (workers == 0) workers = (uint32_t)(T < 65536 ? T : 65536);
    if (workers > T) workers = (uint32_t)T;
    if (T > 0xFFFFFFF0ull) return "obvh_dito: too many tiles";
    hipLaunchKernelGGL(dito_setup, dim3(1), dim3(1024), 0, stream, offsets, single_count, nsets, tile, total,
                       (uint32_t)T, ws, out, tr);
    uint32_t tb = (uint32_t)((T + 255) / 256);
    if (tb > 4096) tb = 4096;
    hipLaunchKernelGGL(dito_tile_table, dim3(tb), dim3(256), 0, stream, nsets, ws);
    if (xyz) launch_passes(PtsAoS{xyz}, offs, nsets, total, tile, workers, ws, out, tr, stream);
    else launch_passes(PtsSoA{x, y, z}, offs, nsets, total, tile, workers, ws, out, tr, stream);
    OBVH_HIP_TRY(hipGetLastError());
    return nullptr;
}

}  // namespace obvh
