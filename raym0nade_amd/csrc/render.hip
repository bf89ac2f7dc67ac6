// Wavefront integrator behind rm_render: the reference's per-pixel recursion
// (renderPixel -> sampleIndirectLightFromFirstIntersection -> sampleRay,
// src/render.cpp:121-551) restated as kernels over path queues.
//
//   k_primary   one lane per pixel: camera ray, closest hit, G-buffer, red
//               nudge, sample budget (renderPixel, :448-501)
//   k_direct    one lane per pixel: spp_direct light samples at the primary
//               vertex with their shadow rays (:425-437,518-522)
//   per round j of indirect samples:
//     k_spawn   one lane per accumulator set: the depth-0 vertex of one
//               indirect sample (:314-397) -> compacted path queue
//     k_bounce  one lane per live path, once per depth: closest hit and the
//               whole sampleRay vertex (:121-312) -> next queue, or the path's
//               light samples straight into its accumulator set
//   k_finish    one lane per pixel: merge the pixel's sets, drop the outlier
//               sample, calcVar (:529-550)
//
// Sample streams: indirect sample s of pixel p draws from Philox stream
// (seed, TAG_INDIRECT, p, s), primary light sample T from (seed, TAG_DIRECT,
// p, T), in the reference's draw order.
//
// Accumulation: in round j pixel p runs M_p paths (K for opaque hits, 16K for
// transparent ones, whose budget is 16x: :498-501), sample s = j*M_p + m into
// accumulator set m.  A set receives at most one path per round, so sets are
// updated without atomics and the result is deterministic.  The reference
// drops a sample iff Clum/(sum-Clum+1e-4) > 16 (:534-547); only the sample with
// the largest Clum can satisfy that (Clum >= 0), so each set tracks its
// maximum and its increments, and k_finish subtracts the global maximum's
// increments when it is dropped.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>
#include "rm_frame.h"
#include "integrator.h"

namespace rm {

constexpr int RB = 256;      // block size of the render kernels
static_assert(RB == WALK_BLOCK, "LDS walk stacks assume blocks of WALK_BLOCK lanes");
constexpr int NCOUNT = 32;   // queue counters per round (>= max_depth + 1)

// ---------------------------------------------------------------- medium
// Medium (src/render.cpp:13-42): a multimap id -> (ior, absorb) holding air
// (-1, 1.0, white) plus the entered dielectrics, kept sorted by id with equal
// ids in insertion order.  Air contributes exact factors, so only the entered
// media are stored.  Every entry's ior is its material's ior and its absorb
// the material's transmitting colour -- the baseColor of a transmitting hit
// (src/material.cpp:300-335) -- nudged like the primary albedo
// (src/render.cpp:492-495) when the medium was entered at the primary vertex,
// whose baseColor is the nudged one.  So an entry is the material id, with
// MED_NUDGED for the nudge, and ior and absorb are read from the material.
// A path enters at most one medium per vertex that continues, and at most
// max_depth <= 16 vertices continue, so 16 entries hold the reference's
// stack exactly; an insert into a full stack (only possible at the last
// vertex, where nothing reads the stack afterwards) is dropped and counted.
// The loops below run over the live entries only, so a path outside every
// medium never touches the array, which lives in per-lane scratch.
constexpr int MED_CAP = 16;
constexpr uint32_t MED_NUDGED = 0x80000000u;
struct Med {
    int n;
    uint32_t e[MED_CAP];
};
__device__ __forceinline__ int med_key(uint32_t e) { return (int)(e & ~MED_NUDGED); }
__device__ __forceinline__ void med_clear(Med &m) { m.n = 0; }
__device__ __forceinline__ V3 med_entry_absorb(const SceneView &S, uint32_t e) {
    V3 a = ld3(S.mats[med_key(e)].trans);
    if (e & MED_NUDGED) nudge_base(a);
    return a;
}
__device__ __forceinline__ float med_ior(const SceneView &S, const Med &m) {
    float r = 1.0f;
#pragma nounroll
    for (int i = 0; i < m.n; ++i) r = gmax(r, S.mats[med_key(m.e[i])].ior);  // std::max(ior, medium.ior)
    return r;
}
__device__ __forceinline__ V3 med_absorb(const SceneView &S, const Med &m) {
    V3 a = v3(1.0f);
#pragma nounroll
    for (int i = 0; i < m.n; ++i) a *= med_entry_absorb(S, m.e[i]);
    return a;
}
// multimap::erase(id): every entry of `id` goes; the others keep their order.
__device__ __forceinline__ void med_erase(Med &m, int id) {
    int k = 0;
#pragma nounroll
    for (int i = 0; i < m.n; ++i)
        if (med_key(m.e[i]) != id) m.e[k++] = m.e[i];
    m.n = k;
}
// multimap::emplace(id, ...): after the entries with keys <= id.  Returns
// false when the stack is full and the entry is dropped (counted as
// RM_STAT_MEDIUM_OVERFLOW by the caller).
__device__ __forceinline__ bool med_insert(Med &m, int id, bool nudged) {
    if (m.n >= MED_CAP) return false;
    int pos = m.n;
#pragma nounroll
    while (pos > 0 && med_key(m.e[pos - 1]) > id) {
        m.e[pos] = m.e[pos - 1];
        --pos;
    }
    m.e[pos] = (uint32_t)id | (nudged ? MED_NUDGED : 0u);
    ++m.n;
    return true;
}

// getAbsorb (src/render.cpp:83-87)
__device__ __forceinline__ V3 get_absorb(V3 a, float dis) {
    float c = dot(a, W_RGB());
    if (!(c < 1.0f - T_EPS)) return v3(1.0f);
    float k = 32.0f * dis;
    V3 b{a.x < 0.0f ? 0.0f : a.x, a.y < 0.0f ? 0.0f : a.y, a.z < 0.0f ? 0.0f : a.z};
    return V3{powf(b.x, k), powf(b.y, k), powf(b.z, k)};
}

// calcEta (src/render.cpp:89-99) at a transmitting hit, whose eta and
// baseColor are its material's ior and transmitting colour; returns false
// when the re-inserted entry did not fit.
__device__ __forceinline__ bool calc_eta(const SceneView &S, Med &m, Surf &s) {
    float e1 = med_ior(S, m), e2;
    bool ok = true;
    if (s.entering) {
        e2 = gmax(e1, s.eta);
    } else {
        med_erase(m, s.id);
        e2 = med_ior(S, m);
        ok = med_insert(m, s.id, false);
    }
    s.eta = e1 / e2;
    return ok;
}

// ---------------------------------------------------------------- buffers
struct PathQ {
    float4 *f[12];
};
// f0 o.xyz | pixel      f1 d.xyz | sample     f2 T.xyz | W        f3 p0.xyz | rf
// f4 dPdx  | rng n      f5 dPdy  | set        f6 dDdx  | depth+exclude<<8
// f7 dDdy  | medium n + job<<8              f8-f11 medium entries 0..15
// (only the float4s holding live entries are written and read)

// The medium stack travels beside the path (load_path / store_path take it
// separately): its entries are indexed dynamically and so stay in scratch,
// and inside Path they would keep the whole record there with them.
struct Path {
    V3 o, d, T, p0;
    float W, rf;
    uint32_t pix, smp, n, set;
    int depth, job;
    bool exclude;
    Diff diff;
};

__device__ __forceinline__ float ub(uint32_t x) { return __uint_as_float(x); }
__device__ __forceinline__ uint32_t bu(float x) { return __float_as_uint(x); }

__device__ __forceinline__ void store_path(const PathQ &Q, uint32_t q, const Path &P, const Med &med) {
    Q.f[0][q] = make_float4(P.o.x, P.o.y, P.o.z, ub(P.pix));
    Q.f[1][q] = make_float4(P.d.x, P.d.y, P.d.z, ub(P.smp));
    Q.f[2][q] = make_float4(P.T.x, P.T.y, P.T.z, P.W);
    Q.f[3][q] = make_float4(P.p0.x, P.p0.y, P.p0.z, P.rf);
    Q.f[4][q] = make_float4(P.diff.dPdx.x, P.diff.dPdx.y, P.diff.dPdx.z, ub(P.n));
    Q.f[5][q] = make_float4(P.diff.dPdy.x, P.diff.dPdy.y, P.diff.dPdy.z, ub(P.set));
    Q.f[6][q] = make_float4(P.diff.dDdx.x, P.diff.dDdx.y, P.diff.dDdx.z, ub((uint32_t)P.depth | (P.exclude ? 256u : 0u)));
    Q.f[7][q] = make_float4(P.diff.dDdy.x, P.diff.dDdy.y, P.diff.dDdy.z, ub((uint32_t)med.n | ((uint32_t)P.job << 8)));
#pragma nounroll
    for (int k = 0; 4 * k < med.n; ++k)
        Q.f[8 + k][q] = make_float4(ub(med.e[4 * k]), ub(med.e[4 * k + 1]), ub(med.e[4 * k + 2]), ub(med.e[4 * k + 3]));
}

__device__ __forceinline__ void load_path(const PathQ &Q, uint32_t q, Path &P, Med &med) {
    float4 a = Q.f[0][q], b = Q.f[1][q], c = Q.f[2][q], d = Q.f[3][q];
    float4 e = Q.f[4][q], f = Q.f[5][q], g = Q.f[6][q], h = Q.f[7][q];
    P.o = V3{a.x, a.y, a.z}; P.pix = bu(a.w);
    P.d = V3{b.x, b.y, b.z}; P.smp = bu(b.w);
    P.T = V3{c.x, c.y, c.z}; P.W = c.w;
    P.p0 = V3{d.x, d.y, d.z}; P.rf = d.w;
    P.diff.dPdx = V3{e.x, e.y, e.z}; P.n = bu(e.w);
    P.diff.dPdy = V3{f.x, f.y, f.z}; P.set = bu(f.w);
    P.diff.dDdx = V3{g.x, g.y, g.z};
    uint32_t dx = bu(g.w);
    P.depth = (int)(dx & 255u);
    P.exclude = (dx & 256u) != 0;
    P.diff.dDdy = V3{h.x, h.y, h.z};
    med.n = (int)(bu(h.w) & 255u);
    P.job = (int)(bu(h.w) >> 8);
#pragma nounroll
    for (int k = 0; 4 * k < med.n; ++k) {
        const float4 m = Q.f[8 + k][q];
        med.e[4 * k] = bu(m.x);
        med.e[4 * k + 1] = bu(m.y);
        med.e[4 * k + 2] = bu(m.z);
        med.e[4 * k + 3] = bu(m.w);
    }
}

// Per-pixel primary-vertex state kept between kernels.
//   pi0 = nudged base colour, roughness
//   pi1 = metallic, opacity, eta (absolute), id | entering << 30
//   pi2 = saved base colour, n_indirect (int bits)
struct PixBufs {
    float4 *pi0, *pi1, *pi2;
    int32_t *M;        // sets per pixel (0: no indirect work)
    int32_t *set_off;  // exclusive prefix of M
    int32_t *set_pix;  // set -> pixel
};

// Accumulator set: 5 float4 = Id, Is, (S, maxC, count, -), maxD, maxS.
struct AccBufs {
    float4 *a[5];
};

// One frame of a (possibly batched) launch.  Per-pixel work buffers are
// indexed by the global pixel index (jobs concatenated); the frame buffers by
// the job-local pixel index.
struct Job {
    SceneView S;
    RArgs A;
    FrameView F;
    int64_t pix_off, np;
    // launches of several jobs: the job's counters in NSHARD shards, xstride
    // counters apart, added into F.stats when the launch ends (every wave of
    // a batch updating one job's counters would serialise on their addresses)
    unsigned long long *xs = nullptr;
    int xstride = 0;
};
constexpr int NSHARD = 16;

// Where a wave adds to counter k of job jb: the frame's own counter, or the
// wave's shard of it.
__device__ __forceinline__ unsigned long long *stat_at(const Job &jb, int k) {
    if (!jb.xs) return jb.F.stats + k;
    const uint32_t w = (blockIdx.x * (RB / 64u) + (threadIdx.x >> 6)) & (uint32_t)(NSHARD - 1);
    return jb.xs + (size_t)w * (size_t)jb.xstride + k;
}

// Runs body(j) for the lanes whose job is j, once per distinct job among the
// active lanes of the wave, with j wave-uniform (readfirstlane), so the job's
// scene and arguments are loaded as scalars.  A wave usually holds one job.
template <class F>
__device__ __forceinline__ void for_each_job(int myjob, bool active, F body) {
    uint64_t pend = __ballot(active);
    while (pend) {
        int lead = (int)__ffsll((unsigned long long)pend) - 1;
        int j = __builtin_amdgcn_readfirstlane(__shfl(myjob, lead, 64));
        bool mine = active && myjob == j;
        if (mine) body(j);
        pend &= ~__ballot(mine);
    }
}

__device__ __forceinline__ int job_of(const Job *J, int nj, int64_t p) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (J[mid].pix_off <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Adds per-lane counts v (v < 2^BITS) to counter k of the frame of each
// lane's job (RM_STAT_*): one atomic per distinct job among the counting lanes
// of the wave.  Every active lane of the wave calls it; a lane with v == 0
// may pass any job.
template <int BITS>
__device__ __forceinline__ void job_stat_add(const Job *J, int myjob, uint32_t v, int k) {
    uint64_t pend = __ballot(v != 0u);
    while (pend) {
        const int lead = (int)__ffsll((unsigned long long)pend) - 1;
        const int j = __builtin_amdgcn_readfirstlane(__shfl(myjob, lead, 64));
        const bool mine = v != 0u && myjob == j;
        unsigned long long sum = 0;
        for (int b = 0; b < BITS; ++b)
            sum += (unsigned long long)__popcll(__ballot(mine && ((v >> b) & 1u))) << b;
        if ((int)(threadIdx.x & 63u) == lead) atomicAdd(stat_at(J[j], k), sum);
        pend &= ~__ballot(mine);
    }
}

// The largest v (< 2^BITS) of each job in the wave, one atomicMax per job
// and wave.  Called by every lane of the wave together; lanes with v == 0
// take no part.
template <int BITS>
__device__ __forceinline__ void job_stat_max(const Job *J, int myjob, uint32_t v, int k) {
    uint64_t pend = __ballot(v != 0u);
    while (pend) {
        const int lead = (int)__ffsll((unsigned long long)pend) - 1;
        const int j = __builtin_amdgcn_readfirstlane(__shfl(myjob, lead, 64));
        const bool mine = v != 0u && myjob == j;
        uint32_t m = 0;
        for (int b = BITS - 1; b >= 0; --b) {
            const uint32_t c = m | (1u << b);
            if (__ballot(mine && v >= c)) m = c;
        }
        if ((int)(threadIdx.x & 63u) == lead) atomicMax(stat_at(J[j], k), (unsigned long long)m);
        pend &= ~__ballot(mine);
    }
}

// Per-lane render counters of a persistent kernel, kept in the block's LDS
// (rows [N + 1][RB]: the lane's job + 1, then N counters; max-type where
// bit k of MAXMASK is set, sums otherwise), so that they cost the kernel no
// registers.  A lane whose next item belongs to another job first adds its
// counts to the old job's RM_STAT_* keys[k] (note); the kernel ends with
// flush_wave, which every lane of the wave calls together and which adds
// each job's counts over the wave with one atomic per counter.  Counting per
// 64 items instead costs a contended atomic each time.  For launches of one
// job; in a batch, whose queues interleave the jobs' items, the kernels add
// per wave and item batch (job_stat_add), one atomic per job present.
template <int N, uint32_t MAXMASK>
struct LaneStats {
    uint32_t *s;  // this lane's column
    __device__ __forceinline__ void init(uint32_t *rows) {
        s = rows + threadIdx.x;
#pragma unroll
        for (int k = 0; k <= N; ++k) s[k * RB] = 0u;
    }
    __device__ __forceinline__ void add_lane(const Job *J, const int (&keys)[N], int job) const {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const uint32_t x = s[(k + 1) * RB];
            if (!x) continue;
            if ((MAXMASK >> k) & 1u)
                atomicMax(stat_at(J[job], keys[k]), (unsigned long long)x);
            else
                atomicAdd(stat_at(J[job], keys[k]), (unsigned long long)x);
        }
    }
    __device__ __forceinline__ void note(const Job *J, const int (&keys)[N], int job, const uint32_t (&x)[N]) {
        const uint32_t cur = s[0];
        if (cur != (uint32_t)job + 1u) {
            if (cur) add_lane(J, keys, (int)cur - 1);
            s[0] = (uint32_t)job + 1u;
#pragma unroll
            for (int k = 0; k < N; ++k) s[(k + 1) * RB] = 0u;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            uint32_t &c = s[(k + 1) * RB];
            c = ((MAXMASK >> k) & 1u) ? (x[k] > c ? x[k] : c) : c + x[k];
        }
    }
    __device__ __forceinline__ void flush_wave(const Job *J, const int (&keys)[N]) const {
        const uint32_t lane = threadIdx.x & 63u;
        const int job = (int)s[0] - 1;
        uint32_t v[N];
        bool has = false;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            v[k] = s[(k + 1) * RB];
            has = has || v[k] != 0u;
        }
        has = has && job >= 0;
        uint64_t pending = __ballot(has);
        while (pending) {
            const int lead = (int)__ffsll((unsigned long long)pending) - 1;
            const int lj = __builtin_amdgcn_readfirstlane(__shfl(job, lead, 64));
            const bool mine = has && job == lj;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                uint32_t x = mine ? v[k] : 0u;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const uint32_t y = __shfl_xor(x, off, 64);
                    x = ((MAXMASK >> k) & 1u) ? (x > y ? x : y) : x + y;
                }
                if (lane == (uint32_t)lead && x) {
                    if ((MAXMASK >> k) & 1u)
                        atomicMax(stat_at(J[lj], keys[k]), (unsigned long long)x);
                    else
                        atomicAdd(stat_at(J[lj], keys[k]), (unsigned long long)x);
                }
            }
            if (mine) has = false;
            pending &= ~__ballot(mine);
        }
    }
};

// p: global pixel (work buffers), lp: job-local pixel (frame buffers)
__device__ __forceinline__ Surf pixel_surf(const PixBufs &P, const FrameView &F, int64_t p, int64_t lp) {
    Surf s = surf_default();
    float4 a = P.pi0[p], b = P.pi1[p];
    s.pos = ld3(F.pos + lp * 3);
    s.shapeN = ld3(F.shapeN + lp * 3);
    s.surfN = ld3(F.surfN + lp * 3);
    s.emission = ld3(F.emis + lp * 3);
    s.base = V3{a.x, a.y, a.z};
    s.roughness = a.w;
    s.metallic = b.x;
    s.opacity = b.y;
    s.eta = b.z;
    uint32_t bits = bu(b.w);
    s.id = (int)(bits & 0x3fffffffu);
    s.entering = (bits >> 30) & 1u;
    return s;
}

// TPS (here and in k_direct, k_extend, k_nee): triangles per walk step.
// MAPS (here and in k_bounce): some job's scene has 8-bit texture maps.
template <int TPS, bool MAPS>
__global__ __launch_bounds__(RB) void k_primary(const Job *__restrict__ J, int nj, int64_t total_px, PixBufs P) {
    const int64_t gp = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (gp >= total_px) return;
    const Job &jb = J[job_of(J, nj, gp)];
    const SceneView &S = jb.S;
    const RArgs &A = jb.A;
    const FrameView &F = jb.F;
    const int64_t p = gp - jb.pix_off;  // frame-local pixel; work buffers use gp
    int x = (int)(p % A.w), y = (int)(p / A.w);
    V3 Dir;
    Diff base;
    camera_ray(A, x, y, Dir, base);
    RayPre r = ray_pre(A.pos, Dir);
    float t = INFINITY;
    int32_t slot = model_closest<true, TPS>(S, A.pos, Dir, r, t);
    const int ji = (int)(&jb - J);
    job_stat_add<1>(J, ji, 1u, RM_STAT_PRIMARY_RAYS);
    job_stat_add<1>(J, ji, slot >= 0 ? 1u : 0u, RM_STAT_HITS);
    Surf s = surf_default();
    float4 zero = make_float4(0, 0, 0, 0);
    if (slot < 0) {
        // miss: position NaN, sky emission; nothing else is computed (:479-485)
        st3(F.pos, p, vnan());
        st3(F.shapeN, p, vnan());
        st3(F.surfN, p, vnan());
        st3(F.emis, p, sky_get(S, Dir));
        st3(F.base, p, v3(0.0f));
        st3(F.mat, p, V3{0.0f, 0.8f, 1.0f});  // HitInfo's defaults (src/model.cpp:4-7)
        P.pi0[gp] = zero;
        P.pi1[gp] = zero;
        P.pi2[gp] = zero;
        P.M[gp] = 0;
        return;
    }
    s.pos = A.pos + t * Dir;
    V3 dPdx, dPdy;
    hit_info<MAPS>(S, slot, Dir, t, base, dPdx, dPdy, s);
    bool opaque = s.opacity > 1.0f - T_EPS;
    V3 sav = s.base;
    nudge_base(s.base);
    int spp_direct = (int)(float(A.spp) * A.p_direct);
    int n_ind = (A.spp - spp_direct) * (opaque ? 1 : 16);
    st3(F.pos, p, s.pos);
    st3(F.shapeN, p, s.shapeN);
    st3(F.surfN, p, s.surfN);
    st3(F.emis, p, s.emission);
    st3(F.base, p, s.base);  // the nudged colour stays when no indirect sample exists (:529-530)
    st3(F.mat, p, V3{s.metallic, s.roughness, s.opacity});
    P.pi0[gp] = make_float4(s.base.x, s.base.y, s.base.z, s.roughness);
    P.pi1[gp] = make_float4(s.metallic, s.opacity, s.eta, ub((uint32_t)s.id | (s.entering ? (1u << 30) : 0u)));
    P.pi2[gp] = make_float4(sav.x, sav.y, sav.z, ub((uint32_t)n_ind));
    bool indirect = !(length(s.emission) > 0.0f) && n_ind > 0;
    P.M[gp] = indirect ? (opaque ? A.K : 16 * A.K) : 0;
}

__device__ __forceinline__ void calc_var(Acc4 &a, float e) {
    a.r *= e; a.g *= e; a.b *= e;
    a.v *= e * e;
    a.v -= dot(V3{a.r, a.g, a.b}, V3{a.r, a.g, a.b});
    if (a.v < 0.0f) a.v = 0.0f;
}

// MINW as for k_bounce / k_nee (register cap chosen by tuning).
template <int MINW, int TPS>
__global__ __launch_bounds__(RB, MINW) void k_direct(const Job *__restrict__ J, int nj, int64_t total_px, PixBufs P) {
    const int64_t gp = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (gp >= total_px) return;
    const Job &jb = J[job_of(J, nj, gp)];
    const SceneView &S = jb.S;
    const RArgs &A = jb.A;
    const FrameView &F = jb.F;
    const int64_t p = gp - jb.pix_off;  // the RNG streams are keyed by the frame-local pixel
    Acc4 d{0, 0, 0, 0}, sp{0, 0, 0, 0};
    V3 pos = ld3(F.pos + p * 3);
    V3 em = ld3(F.emis + p * 3);
    uint32_t nrays = 0;
    if (any_finite(pos) && !(length(em) > 0.0f)) {
        Surf s = pixel_surf(P, F, gp, p);
        V3 in = normalize(s.pos - A.pos);
        V3 V = -in;
        int cnt = (int)(float(A.spp) * A.p_direct);
        float total = direct_total(S, s, V);
        if (total != 0.0f) {
            TraceCounts tc{};
            for (int T = 0; T < cnt; ++T) {
                Stream g{A.seed, TAG_DIRECT, (uint32_t)p, (uint32_t)T, 0u};
                LSample ls;
                if (direct_sample<false, true, TPS>(S, s, V, total, cnt, g, ls, tc, nullptr, &nrays))
                    accumulate(s.base, ls, d, sp);
            }
        }
    }
    job_stat_add<31>(J, (int)(&jb - J), nrays, RM_STAT_SHADOW_RAYS);
    calc_var(d, A.exposure);
    calc_var(sp, A.exposure);
    reinterpret_cast<float4 *>(F.Dd)[p] = make_float4(d.r, d.g, d.b, d.v);
    reinterpret_cast<float4 *>(F.Ds)[p] = make_float4(sp.r, sp.g, sp.b, sp.v);
}

// The NEE pass split in three, with the same result: k_nee_gen draws every
// light sample of each NEE record (sampleCount[depth] <= NEE_MAX of them) up
// to its shadow ray, k_shadow traces all shadow rays in a lean persistent
// any-hit walk, and k_nee_add adds the clear samples to the path's set in
// the reference order.  Drawing ahead is exact because a shadow ray draws
// no random numbers.  Entry e = record * NEE_MAX + i of the shadow queue:
// o = (origin, aim), d = (direction, job + 1 or 0 when there is no sample),
// x = (light * bsdf * factor, weight), vis = 1 when the ray is clear.

constexpr int NEE_MAX = 6;  // sample_count(16)
// The queue: rays appended by the generating kernels (one atomic per wave)
// at o = (origin, aim), d = (direction, job index), with vis written by
// k_shadow; and per light-sample entry e, ref[e] = index of its queued ray,
// or SH_NONE (no sample), SH_CLEAR / SH_BLOCKED (a job whose tree is
// shallower than the launch's queue depth: there a walk costs less than
// queueing it, so the generating kernel traced it itself), and its payload x
// (and x2).
constexpr uint32_t SH_NONE = 0xffffffffu, SH_CLEAR = 0xfffffffeu, SH_BLOCKED = 0xfffffffdu;
struct ShadowQ {
    float4 *o, *d;
    uint8_t *vis;
    uint32_t *ref;
    float4 *x;
    float4 *x2;  // the primary-vertex pass: (light, 0); x then holds (bsdf, weight)
};
__device__ __forceinline__ bool shadow_clear(const ShadowQ &Q, uint32_t r) {
    return r == SH_CLEAR || (r < SH_BLOCKED && Q.vis[r]);
}

// Wave-aggregated append: one atomic per wave.
__device__ __forceinline__ uint32_t append(uint32_t *counter, bool want) {
    uint64_t mask = __ballot(want);
    if (!want) return 0xffffffffu;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
    uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader, 64);
    return base + rank;
}

// Block-wide append: one atomic per block instead of one per wave (appends
// to one counter from every wave of the launch serialise on its address).
// Every thread of the block calls it together; s holds RB / 64 + 2 words of
// LDS, and keeps the block's base and total in its last two.  Returns the
// slot of a lane that wants one.
__device__ __forceinline__ uint32_t append_block(uint32_t *counter, bool want, uint32_t *s) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(want);
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) s[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < RB / 64; ++w) {
            const uint32_t c = s[w];
            s[w] = t;
            t += c;
        }
        s[RB / 64] = t ? atomicAdd(counter, t) : 0u;
        s[RB / 64 + 1] = t;
    }
    __syncthreads();
    return want ? s[RB / 64] + s[wave] + rank : 0xffffffffu;
}

// Work claims of the persistent item kernels: a wave takes 64 queue items
// per iteration, claimed `per` batches at a time with one atomic.  per is up
// to 4 when the queue holds 16 such claims for every wave of the launch, so
// that the last claims, which end the launch, stay a small part of it;
// shorter queues claim one batch at a time.  Wave-uniform.
struct Claims {
    uint32_t next = 0, end = 0, per;
    __device__ __forceinline__ explicit Claims(uint32_t n) {
        const uint32_t waves = gridDim.x * (RB / 64u);
        per = (uint32_t)max(1, min(4, (int)(n / (64u * 16u * waves))));
    }
    __device__ __forceinline__ uint32_t take(uint32_t *work) {
        if (next >= end) {
            uint32_t b = 0;
            if ((threadIdx.x & 63u) == 0) b = atomicAdd(work, 64u * per);
            next = __shfl(b, 0, 64);
            end = next + 64u * per;
        }
        const uint32_t r = next;
        next += 64u;
        return r;
    }
};

// Queues the shadow ray of light-sample entry e, or traces it here for a
// shallow tree; sets ref[e].
template <int TPS>
__device__ __forceinline__ void shadow_entry(const ShadowQ &Q, int64_t e, const SceneView &S, int job, V3 o,
                                             const ShadowRay &sr, int queue_depth, uint32_t *queued) {
    // k_shadow carries only the walk's part of the scene, so shadow rays of a
    // scene with alpha cut-outs are traced here, through Model::rayHit_test.
    const bool q = S.depth >= queue_depth && !S.cutout;
    const uint32_t slot = append(queued, q);
    if (q) {
        Q.o[slot] = make_float4(o.x, o.y, o.z, sr.aim);
        Q.d[slot] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, ub((uint32_t)job));
        Q.ref[e] = slot;
    } else {
        RayPre r = ray_pre(o, sr.dir);
        Q.ref[e] = model_clear<true, TPS>(S, o, sr.dir, r, sr.aim) ? SH_CLEAR : SH_BLOCKED;
    }
}

// The primary-vertex light samples (k_direct) split like the NEE pass, in
// chunks of C samples per pixel: k_direct_gen draws sample t0 + t of pixel gp
// in lane gp * C + t (each sample has its own stream, so they are
// independent) up to its shadow ray; k_shadow traces them; k_direct_add
// continues each pixel's sums in sample order in the frame's Dd / Ds; after
// the last chunk k_direct_final applies calcVar.  Same result as k_direct.
// A chunk covers the pixels [0, npx) of the jobs with more than t0 samples.
template <int MINW, int TPS>
__global__ __launch_bounds__(RB, MINW) void k_direct_gen(const Job *__restrict__ J, int nj, int64_t npx, PixBufs P,
                                                   int C, int t0, ShadowQ SQ, int queue_depth, uint32_t *queued) {
    __shared__ uint32_t s_app[RB / 64 + 2];
    const int64_t g = (int64_t)blockIdx.x * RB + threadIdx.x;
    const bool in = g < npx * C;
    const int64_t gp = in ? g / C : 0;
    const int jid = job_of(J, nj, gp);
    const Job &jb = J[jid];
    const SceneView &S = jb.S;
    bool ok = false;
    ShadowRay sr;
    V3 o;
    if (in) {
        const int T = t0 + (int)(g - gp * C);
        const RArgs &A = jb.A;
        const FrameView &F = jb.F;
        const int64_t p = gp - jb.pix_off;
        const int cnt = (int)(float(A.spp) * A.p_direct);
        if (T < cnt) {
            V3 pos = ld3(F.pos + p * 3);
            V3 em = ld3(F.emis + p * 3);
            if (any_finite(pos) && !(length(em) > 0.0f)) {
                Surf s = pixel_surf(P, F, gp, p);
                V3 V = -normalize(s.pos - A.pos);
                float total = direct_total(S, s, V);
                if (total != 0.0f) {
                    Stream rng{A.seed, TAG_DIRECT, (uint32_t)p, (uint32_t)T, 0u};
                    LSample ls;
                    if (direct_ray(S, s, V, total, cnt, rng, sr, ls)) {
                        SQ.x[g] = make_float4(ls.bsdf.x, ls.bsdf.y, ls.bsdf.z, ls.weight);
                        SQ.x2[g] = make_float4(ls.light.x, ls.light.y, ls.light.z, 0.0f);
                        o = s.pos;
                        ok = true;
                    }
                }
            }
        }
    }
    // k_shadow carries only the walk's part of the scene, so shadow rays of a
    // scene with alpha cut-outs are traced here, through Model::rayHit_test.
    const bool q = ok && S.depth >= queue_depth && !S.cutout;
    const uint32_t slot = append_block(queued, q, s_app);
    if (q) {
        SQ.o[slot] = make_float4(o.x, o.y, o.z, sr.aim);
        SQ.d[slot] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, ub((uint32_t)jid));
        SQ.ref[g] = slot;
    } else if (ok) {
        RayPre r = ray_pre(o, sr.dir);
        SQ.ref[g] = model_clear<true, TPS>(S, o, sr.dir, r, sr.aim) ? SH_CLEAR : SH_BLOCKED;
    } else if (in) {
        SQ.ref[g] = SH_NONE;
    }
    if (nj == 1) {
        // one job: its shadow-ray count summed over the block (queued and
        // traced here alike), one atomic per block
        __syncthreads();
        const uint64_t m = __ballot(ok);
        if ((threadIdx.x & 63u) == 0) s_app[threadIdx.x >> 6] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int w = 0; w < RB / 64; ++w) t += s_app[w];
            if (t) atomicAdd(J[0].F.stats + RM_STAT_SHADOW_RAYS, (unsigned long long)t);
        }
    } else {
        job_stat_add<1>(J, jid, ok ? 1u : 0u, RM_STAT_SHADOW_RAYS);
    }
}

__global__ __launch_bounds__(RB) void k_direct_add(const Job *__restrict__ J, int nj, int64_t npx, PixBufs P, int C,
                                                   int t0, ShadowQ SQ) {
    const int64_t gp = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (gp >= npx) return;
    const Job &jb = J[job_of(J, nj, gp)];
    const FrameView &F = jb.F;
    const int64_t p = gp - jb.pix_off;
    float4 *Dd = reinterpret_cast<float4 *>(F.Dd), *Ds = reinterpret_cast<float4 *>(F.Ds);
    Acc4 d{0, 0, 0, 0}, sp{0, 0, 0, 0};
    if (t0 > 0) {
        const float4 a = Dd[p], b = Ds[p];
        d = Acc4{a.x, a.y, a.z, a.w};
        sp = Acc4{b.x, b.y, b.z, b.w};
    }
    const float4 b0 = P.pi0[gp];
    const V3 base{b0.x, b0.y, b0.z};
    for (int t = 0; t < C; ++t) {
        const int64_t e = gp * C + t;
        if (!shadow_clear(SQ, SQ.ref[e])) continue;
        const float4 x = SQ.x[e], l = SQ.x2[e];
        accumulate(base, LSample{V3{x.x, x.y, x.z}, V3{l.x, l.y, l.z}, x.w}, d, sp);
    }
    Dd[p] = make_float4(d.r, d.g, d.b, d.v);
    Ds[p] = make_float4(sp.r, sp.g, sp.b, sp.v);
}

__global__ __launch_bounds__(RB) void k_direct_final(const Job *__restrict__ J, int nj, int64_t total_px) {
    const int64_t gp = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (gp >= total_px) return;
    const Job &jb = J[job_of(J, nj, gp)];
    const FrameView &F = jb.F;
    const int64_t p = gp - jb.pix_off;
    float4 *Dd = reinterpret_cast<float4 *>(F.Dd), *Ds = reinterpret_cast<float4 *>(F.Ds);
    Acc4 d{0, 0, 0, 0}, sp{0, 0, 0, 0};
    if ((int)(float(jb.A.spp) * jb.A.p_direct) > 0) {  // an add pass wrote this pixel's sums
        const float4 a = Dd[p], b = Ds[p];
        d = Acc4{a.x, a.y, a.z, a.w};
        sp = Acc4{b.x, b.y, b.z, b.w};
    }
    calc_var(d, jb.A.exposure);
    calc_var(sp, jb.A.exposure);
    Dd[p] = make_float4(d.r, d.g, d.b, d.v);
    Ds[p] = make_float4(sp.r, sp.g, sp.b, sp.v);
}


// The depth-0 vertex (sampleIndirectLightFromFirstIntersection, :314-397).
__global__ __launch_bounds__(RB) void k_spawn(const Job *__restrict__ J, int nj, PixBufs P, int round,
                                              uint32_t total_sets, PathQ Q, uint32_t *count) {
    __shared__ uint32_t s_app[RB / 64 + 2], s_ovf[RB / 64], s_max[RB / 64];
    uint32_t g = blockIdx.x * RB + threadIdx.x;
    bool want = false;
    Path path;
    Med med;
    med_clear(med);
    bool active = false;
    int64_t gp = 0;
    uint32_t s_idx = 0;
    int myjob = -1;
    bool ovf = false;
    if (g < total_sets) {
        gp = P.set_pix[g];
        int M = P.M[gp];
        int m = (int)(g - (uint32_t)P.set_off[gp]);
        uint32_t n_ind = bu(P.pi2[gp].w);
        s_idx = (uint32_t)round * (uint32_t)M + (uint32_t)m;
        active = s_idx < n_ind;
        if (active) myjob = job_of(J, nj, gp);
    }
    for_each_job(myjob, active, [&](int ji) {
            const Job &jb = J[ji];
            const RArgs &A = jb.A;
            const int64_t p = gp - jb.pix_off;
            Surf s = pixel_surf(P, jb.F, gp, p);
            int x = (int)(p % A.w), y = (int)(p / A.w);
            V3 Dir;
            Diff cam;
            camera_ray(A, x, y, Dir, cam);
            V3 in = normalize(s.pos - A.pos);
            float hit_t = length(s.pos - A.pos);
            V3 V = -in;
            float rf = s.roughness * 1.0f;
            s.eta = 1.0f / s.eta;
            float P_reflect = 1.0f, Fr = 0.0f;
            if (s.opacity < T_EPS) {
                V3 rd;
                precise_refraction(s, V, rd, Fr);
                P_reflect = 0.24f + (1.0f - 0.24f) * Fr;
            }
            Stream rng{A.seed, TAG_INDIRECT, (uint32_t)p, s_idx, 0u};
            V3 nd, bp;
            int fails = 0;
            Diff next;
            if (rng.next() < P_reflect) {
                calc_dPdxy(in, hit_t, s.shapeN, cam, next.dPdx, next.dPdy);
                calc_dDdxy(in, s.surfN, cam, next.dDdx, next.dDdy);
                sample_reflection(s, V, rng, nd, bp, fails);
                bp /= P_reflect;
            } else {
                sample_btdf(s, V, rng, nd, bp, fails);
                bp *= (1.0f - Fr);
                bp /= (1.0f - P_reflect);
                next = cam;
                // the primary vertex's baseColor is the nudged one (k_primary)
                if (s.entering)
                    ovf = !med_insert(med, s.id, true);
                else
                    med_erase(med, s.id);
            }
            if (any_finite(nd)) {
                want = true;
                path.o = s.pos;
                path.d = nd;
                path.T = v3(1.0f);
                path.p0 = bp;
                path.W = fails > 0 ? 1.0f / (float)(1 + fails) : 1.0f;
                path.rf = rf;
                path.pix = (uint32_t)gp;  // global; the RNG key is the frame-local pixel
                path.job = ji;
                path.smp = s_idx;
                path.n = rng.n;
                path.set = g;
                path.depth = 1;
                path.exclude = true;
                path.diff = next;
            }
    });
    // one append and, for a launch of one job, one atomic per counter per
    // block: counters that every wave of the launch updates serialise
    const uint32_t q = append_block(count, want, s_app);
    if (want) store_path(Q, q, path, med);
    if (nj == 1) {
        const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
        const uint32_t ov = (uint32_t)__popcll(__ballot(ovf));
        uint32_t mx = 0;  // med.n <= MED_CAP = 16: five bits
        for (int b = 4; b >= 0; --b) {
            const uint32_t c = mx | (1u << b);
            if (__ballot(want && (uint32_t)med.n >= c)) mx = c;
        }
        if (lane == 0) {
            s_ovf[wave] = ov;
            s_max[wave] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t o = 0, m = 0;
            for (int w = 0; w < RB / 64; ++w) {
                o += s_ovf[w];
                m = s_max[w] > m ? s_max[w] : m;
            }
            unsigned long long *stats = J[0].F.stats;
            if (s_app[RB / 64 + 1]) atomicAdd(stats + RM_STAT_PATHS, (unsigned long long)s_app[RB / 64 + 1]);
            if (o) atomicAdd(stats + RM_STAT_MEDIUM_OVERFLOW, (unsigned long long)o);
            if (m) atomicMax(stats + RM_STAT_MEDIUM_MAX, (unsigned long long)m);
        }
    } else {
        job_stat_add<1>(J, myjob, want ? 1u : 0u, RM_STAT_PATHS);
        job_stat_add<1>(J, myjob, ovf ? 1u : 0u, RM_STAT_MEDIUM_OVERFLOW);
        job_stat_max<5>(J, myjob, want ? (uint32_t)med.n : 0u, RM_STAT_MEDIUM_MAX);
    }
}

// Light sample of a finished path -> its accumulator set.
// Light sample k of indirect sample s is named s * 8 + k (k < 6): the name
// k_finish records for the dropped outlier, so the adjoint skips the same one.
struct SetAcc {
    Acc4 d, s;
    float S, maxC, cnt;
    uint32_t maxid;
    Acc4 md, ms;
};

__device__ __forceinline__ void add_sample(SetAcc &a, V3 base, V3 p0, V3 light, float w, uint32_t id, Acc4 &dd,
                                           Acc4 &ds) {
    float c = dot(p0 * light, W_RGB()) * w;
    dd = Acc4{0, 0, 0, 0};
    ds = Acc4{0, 0, 0, 0};
    accumulate(base, LSample{p0, light, w}, dd, ds);
    a.d.r += dd.r; a.d.g += dd.g; a.d.b += dd.b; a.d.v += dd.v;
    a.s.r += ds.r; a.s.g += ds.g; a.s.b += ds.b; a.s.v += ds.v;
    a.S += c;
    a.cnt += 1.0f;
    if (c > a.maxC) {
        a.maxC = c;
        a.maxid = id;
        a.md = dd;
        a.ms = ds;
    }
}

__device__ __forceinline__ SetAcc load_set(const AccBufs &B, uint32_t g) {
    SetAcc a;
    float4 x0 = B.a[0][g], x1 = B.a[1][g], x2 = B.a[2][g], x3 = B.a[3][g], x4 = B.a[4][g];
    a.d = Acc4{x0.x, x0.y, x0.z, x0.w};
    a.s = Acc4{x1.x, x1.y, x1.z, x1.w};
    a.S = x2.x; a.maxC = x2.y; a.cnt = x2.z; a.maxid = bu(x2.w);
    a.md = Acc4{x3.x, x3.y, x3.z, x3.w};
    a.ms = Acc4{x4.x, x4.y, x4.z, x4.w};
    return a;
}
__device__ __forceinline__ void store_set(const AccBufs &B, uint32_t g, const SetAcc &a) {
    B.a[0][g] = make_float4(a.d.r, a.d.g, a.d.b, a.d.v);
    B.a[1][g] = make_float4(a.s.r, a.s.g, a.s.b, a.s.v);
    B.a[2][g] = make_float4(a.S, a.maxC, a.cnt, ub(a.maxid));
    B.a[3][g] = make_float4(a.md.r, a.md.g, a.md.b, a.md.v);
    B.a[4][g] = make_float4(a.ms.r, a.ms.g, a.ms.b, a.ms.v);
}

// One sampleRay vertex (src/render.cpp:121-312) for every live path.
// Persistent: each wave claims 64 queue entries per atomic on `work` until the
// queue (count_in, written by the previous stage) is exhausted, so the grid is
// sized by the machine, not by the live count, and no host round trip is
// needed between depths.
// A path that ends in sampleRay's NEE-and-stop branch (:147-160,229-240)
// leaves the shading state of its last vertex here; k_nee draws that vertex's
// light samples (sampleDirectLight with the path's own stream) and adds them
// to the path's accumulator set.  9 float4 per record:
//   f0 pos | pixel   f1 shapeN | set   f2 surfN | sample   f3 V | rng counter
//   f4 base | roughness   f5 metallic, opacity, eta, flags   f6 T | W
//   f7 p0 | factor   f8 extra | specular
//   flags = entering | extra_on << 1 | light-sample count << 2 | job << 8
struct NeeQ {
    float4 *f[9];
};
struct NeeRec {
    Surf s;
    V3 V, extra;
    uint32_t n;
    int cnt;
    float factor;
    bool extra_on;
};
__device__ __forceinline__ void store_nee(const NeeQ &Q, uint32_t q, const Path &P, const NeeRec &R) {
    const Surf &s = R.s;
    Q.f[0][q] = make_float4(s.pos.x, s.pos.y, s.pos.z, ub(P.pix));
    Q.f[1][q] = make_float4(s.shapeN.x, s.shapeN.y, s.shapeN.z, ub(P.set));
    Q.f[2][q] = make_float4(s.surfN.x, s.surfN.y, s.surfN.z, ub(P.smp));
    Q.f[3][q] = make_float4(R.V.x, R.V.y, R.V.z, ub(R.n));
    Q.f[4][q] = make_float4(s.base.x, s.base.y, s.base.z, s.roughness);
    uint32_t flags = (s.entering ? 1u : 0u) | (R.extra_on ? 2u : 0u) | ((uint32_t)R.cnt << 2) | ((uint32_t)P.job << 8);
    Q.f[5][q] = make_float4(s.metallic, s.opacity, s.eta, ub(flags));
    Q.f[6][q] = make_float4(P.T.x, P.T.y, P.T.z, P.W);
    Q.f[7][q] = make_float4(P.p0.x, P.p0.y, P.p0.z, R.factor);
    Q.f[8][q] = make_float4(R.extra.x, R.extra.y, R.extra.z, s.specular);
}

// Adds light sample x of a finished path (sample name smp * 8 + k) to its
// accumulator set, passBsdf'd up to the first vertex: light *= terminal bsdf
// (and, on the refraction side, the medium's absorb) and every deeper bsdfPdf.
// A read-modify-write: the set belongs to this path for the round.
// With `rec` (the sample's record, rm_frame_set_recording) the sample's
// increments and w L are also summed into the record.
__device__ __forceinline__ void add_light_sample(const PixBufs &P, const AccBufs &AB, uint32_t pix, uint32_t set,
                                                 uint32_t smp, uint32_t k, V3 T, float W, V3 p0, const LSample &x,
                                                 bool extra_on, V3 extra, float4 *rec) {
    float4 b0 = P.pi0[pix], b1 = P.pi1[pix];
    V3 base = b1.y < T_EPS ? v3(0.0f) : V3{b0.x, b0.y, b0.z};
    float inv = 1.0f / (float)bu(P.pi2[pix].w);
    V3 L = x.light * x.bsdf;
    if (extra_on) L = L * extra;
    L = L * T;
    const float w = x.weight * W * inv;
    SetAcc acc = load_set(AB, set);
    Acc4 dd, ds;
    add_sample(acc, base, p0, L, w, smp * 8u + k, dd, ds);
    store_set(AB, set, acc);
    if (rec && !(length(L) < T_EPS)) {
        float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
        rec[0] = make_float4(r0.x + dd.r, r0.y + dd.g, r0.z + dd.b, 1.0f);
        rec[1] = make_float4(r1.x + ds.r, r1.y + ds.g, r1.z + ds.b, 0.0f);
        V3 wl = w * L;
        rec[2] = make_float4(r2.x + wl.x, r2.y + wl.y, r2.z + wl.z, 0.0f);
    }
}

// The record slot of indirect sample `smp` of global pixel `gp` in job jb.
__device__ __forceinline__ float4 *rec_slot(const Job &jb, uint32_t gp, uint32_t smp) {
    if (!jb.F.rec) return nullptr;
    return jb.F.rec + ((int64_t)(gp - jb.pix_off) * jb.F.rec_stride + smp) * 3;
}

// Closest hit of every queued path (sampleRay's model.rayHit, :127), ahead of
// the shading kernel: a lean persistent traversal.  A wave hands a new queue
// entry to each of its idle lanes once at least `fetch_min` lanes are idle
// (one atomic per refill), so lanes do not wait for the wave's longest walk.
// Writes (t, triangle slot or -1) per entry.  The traversal is the same walk
// as everywhere else, so the hits are bit-identical.
// LDS: the walk's first lds_entries pending entries in the block's dynamic
// LDS (lds_entries * RB * 8 bytes), deeper ones in per-lane scratch.
// COUNT: each walk's node visits and tests go to its frame's RM_STAT_WALK_*.
// A lane sums the walks of one job in registers and adds them when its next
// walk belongs to another job (add_walk_counts) and, summed over the wave per
// job, when the kernel ends (flush_walk_counts).
__device__ __forceinline__ void add_walk_counts(unsigned long long *stats, TraceCounts &c) {
    if (c.interior | c.leaves) {
        atomicAdd(stats + RM_STAT_WALK_INTERIOR, (unsigned long long)c.interior);
        atomicAdd(stats + RM_STAT_WALK_LEAVES, (unsigned long long)c.leaves);
        atomicAdd(stats + RM_STAT_WALK_BOXES, (unsigned long long)c.boxes);
        atomicAdd(stats + RM_STAT_WALK_TRIS, (unsigned long long)c.tris);
    }
    c = TraceCounts{};
}

// Called by every lane of the wave together.
__device__ __forceinline__ void flush_walk_counts(const Job *__restrict__ J, int job, TraceCounts &c) {
    const uint32_t lane = threadIdx.x & 63u;
    bool has = (c.interior | c.leaves) != 0u;
    uint64_t pending = __ballot(has);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lj = __shfl(job, leader, 64);
        const bool mine = has && job == lj;
        uint32_t v[4] = {mine ? c.interior : 0u, mine ? c.leaves : 0u, mine ? c.boxes : 0u, mine ? c.tris : 0u};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
        if (lane == (uint32_t)leader) {
            unsigned long long *stats = J[lj].F.stats;
            atomicAdd(stats + RM_STAT_WALK_INTERIOR, (unsigned long long)v[0]);
            atomicAdd(stats + RM_STAT_WALK_LEAVES, (unsigned long long)v[1]);
            atomicAdd(stats + RM_STAT_WALK_BOXES, (unsigned long long)v[2]);
            atomicAdd(stats + RM_STAT_WALK_TRIS, (unsigned long long)v[3]);
        }
        if (mine) has = false;
        pending &= ~__ballot(mine);
    }
    c = TraceCounts{};
}

template <bool LDS, int TPS, bool COUNT = false>
__global__ __launch_bounds__(RB) void k_extend(const Job *__restrict__ J, PathQ Q, const uint32_t *count_in,
                                               uint32_t *work, uint2 *__restrict__ hit, int fetch_min,
                                               int lds_entries, int phases) {
    const uint32_t n = *count_in;
    const uint32_t lane = threadIdx.x & 63u;
    ScratchEntries ent;
    Walker<false, COUNT, typename std::conditional<LDS, LdsStack, PrivateStack>::type, TPS> w;
    if constexpr (LDS)
        w.stk.bind(&ent, lds_entries);
    else
        w.stk.bind(&ent);
    RayPre r;
    SceneView S;  // per lane: the nodes, triangles and root of its path's scene
    int64_t q = -1;
    int job = 0;
    bool exhausted = false;
    TraceCounts c{};
    while (true) {
        bool idle = q < 0;
        uint64_t idle_mask = __ballot(idle);
        uint64_t busy_mask = __ballot(!idle);
        if (exhausted && busy_mask == 0) break;
        if (!exhausted && idle_mask && (__popcll(idle_mask) >= fetch_min || busy_mask == 0)) {
            uint32_t cnt = (uint32_t)__popcll(idle_mask);
            uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle_mask) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(work, cnt);
            base = __shfl(base, (int)leader, 64);
            if (base + cnt >= n) exhausted = true;
            if (idle) {
                uint32_t slot = base + (uint32_t)__popcll(idle_mask & ((1ull << lane) - 1ull));
                if (slot < n) {
                    q = slot;
                    float4 a = Q.f[0][slot], b = Q.f[1][slot];
                    const int nj = (int)(bu(Q.f[7][slot].w) >> 8);
                    if (COUNT && nj != job) add_walk_counts(J[job].F.stats, c);
                    job = nj;
                    const SceneView &SJ = J[job].S;
                    S.nodes = SJ.nodes;
                    S.tris = SJ.tris;
                    S.root_leaf = SJ.root_leaf;
                    r = ray_pre(V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z});
                    w.init(S, r, INFINITY);
                }
            }
        }
        const bool turn = walk_turn(w.cur, phases, q >= 0);
        if (q >= 0 && turn && w.step(S, r, T_EPS, -INFINITY, c)) {
            hit[q] = make_uint2(bu(w.tmax), (uint32_t)w.best);
            q = -1;
        }
    }
    if (COUNT) flush_walk_counts(J, job, c);
}

// MINW: minimum resident waves per SIMD the register allocation must allow
// (1: no cap; 3: at most 168 VGPRs; 4: at most 128, the rest spilled).
// Chosen by tuning.
// The 8-bit map code would take the uncapped variant one register past 256
// (one wave per SIMD); it is held to two waves.
template <int MINW, bool MAPS>
__global__ __launch_bounds__(RB, (MAPS && MINW < 2) ? 2 : MINW) void k_bounce(const Job *__restrict__ J, int nj, PixBufs P, AccBufs AB, PathQ Qin,
                                               const uint32_t *count_in, uint32_t *work, const uint2 *__restrict__ hit,
                                               PathQ Qout, uint32_t *count_out, NeeQ NQ, uint32_t *nee_count) {
  const uint32_t n_in = *count_in;
  const uint32_t lane = threadIdx.x & 63u;
  const int keys[4] = {RM_STAT_EXTENSION_RAYS, RM_STAT_HITS, RM_STAT_MEDIUM_OVERFLOW, RM_STAT_MEDIUM_MAX};
  __shared__ uint32_t stat_rows[5 * RB];
  LaneStats<4, 8u> st;
  st.init(stat_rows);
  Claims cl(n_in);
  while (true) {
    const uint32_t base = cl.take(work);
    if (base >= n_in) break;
    uint32_t q = base + lane;
    bool live = q < n_in;
    bool cont = false;
    Path path;
    Med med;
    med_clear(med);
    uint2 hq = make_uint2(0u, 0u);
    if (live) {
        load_path(Qin, q, path, med);
        hq = hit[q];
    }
    bool nee_out = false;
    bool hit_any = false;
    uint32_t ovf = 0;
    NeeRec nee_rec;
    for_each_job(live ? path.job : -1, live, [&](int jid) {
        const Job &jb = J[jid];
        const SceneView &S = jb.S;
        const RArgs &A = jb.A;
        Stream rng{A.seed, TAG_INDIRECT, (uint32_t)(path.pix - jb.pix_off), path.smp, path.n};
        const int depth = path.depth;
        const int maxD = A.max_depth;
        float t = ub(hq.x);
        TraceCounts tc{};
        int32_t slot = (int32_t)hq.y;
        // k_extend walks once; the rest of Model::rayHit for alpha cut-outs
        // (this launch has no walk LDS)
        if (S.cutout) slot = model_hit_rest<false, 0>(S, path.o, path.d, t, slot);
        V3 extra = v3(1.0f);  // refract-side NEE multiplies by absorb after the samples (passBsdf)
        bool extra_on = false;
        // an emitter or sky hit ends the path with one light sample
        auto emit = [&](const LSample &x) {
            add_light_sample(P, AB, path.pix, path.set, path.smp, 0u, path.T, path.W, path.p0, x, false, extra,
                             rec_slot(jb, path.pix, path.smp));
        };
        if (slot < 0) {
            if (!(path.exclude || !S.sky)) emit(LSample{v3(1.0f), sky_get(S, path.d), 1.0f});
        } else {
            hit_any = true;
            Surf s = surf_default();
            s.pos = path.o + path.d * t;
            V3 dPdx, dPdy;
            hit_info<MAPS>(S, slot, path.d, t, path.diff, dPdx, dPdy, s);
            V3 V = -path.d;
            path.rf = gmax(path.rf, 1.0f * s.roughness);
            s.roughness = gmax(s.roughness, path.rf);
            const float P_RR = 1.0f + (0.5f - 1.0f) * sqrtf(s.roughness);
            bool doNEE = P_RR < 0.9f && depth >= A.rr_min;
            V3 absorb = get_absorb(med_absorb(S, med), t);
            if (length(s.emission) > T_EPS) {
                if (!path.exclude) emit(LSample{absorb, s.emission, 1.0f});
            } else {
                float P_reflect = 1.0f, Fr = 0.0f;
                if (s.opacity < T_EPS) {
                    if (!calc_eta(S, med, s)) ++ovf;
                    V3 rd;
                    precise_refraction(s, V, rd, Fr);
                    P_reflect = med.n == 0 ? 0.24f + (1.0f - 0.24f) * Fr : Fr;
                    P_reflect = gmax(P_reflect - 1e-3f, 0.0f);
                }
                V3 nd = vnan(), bp = vnan();
                int fails = 0;
                Diff next;
                bool nee = false;
                float factor = 1.0f;
                if (rng.next() <= P_reflect) {
                    doNEE = doNEE && s.entering && med.n == 0;
                    if (doNEE && (depth >= maxD || rng.next() > P_RR)) {
                        nee = true;
                        factor = 1.0f / P_reflect;
                        if (depth < maxD) factor /= (1.0f - P_RR);
                    } else {
                        sample_reflection(s, V, rng, nd, bp, fails);
                        bp /= P_reflect;
                        if (doNEE) bp /= P_RR;
                        next.dPdx = dPdx;
                        next.dPdy = dPdy;
                        calc_dDdxy(path.d, s.surfN, path.diff, next.dDdx, next.dDdy);
                    }
                } else {
                    doNEE = doNEE && !s.entering && med.n == 1;
                    if (doNEE && (depth >= maxD || rng.next() > P_RR)) {
                        nee = true;
                        factor = 1.0f / (1.0f - P_reflect);
                        if (depth < maxD) factor /= (1.0f - P_RR);
                        extra = absorb;
                        extra_on = true;
                    } else {
                        if (fabsf(s.eta - 1.0f) < T_EPS) {
                            nd = path.d;
                            bp = v3(1.0f);
                        } else {
                            sample_btdf(s, V, rng, nd, bp, fails);
                            bp *= (1.0f - Fr);
                        }
                        next = path.diff;
                        bp /= (1.0f - P_reflect);
                        if (doNEE) bp /= P_RR;
                        if (s.entering) {
                            if (!med_insert(med, s.id, false)) ++ovf;
                        } else {
                            med_erase(med, s.id);
                        }
                    }
                }
                if (nee) {
                    // the light samples of this vertex are drawn by k_nee,
                    // from the same stream position
                    nee_out = true;
                    nee_rec.s = s;
                    nee_rec.V = V;
                    nee_rec.n = rng.n;
                    nee_rec.cnt = sample_count(depth);
                    nee_rec.factor = factor;
                    nee_rec.extra = extra;
                    nee_rec.extra_on = extra_on;
                } else if (any_finite(nd) && depth != maxD) {
                    bp *= absorb;
                    cont = true;
                    path.T = path.T * bp;
                    if (fails > 0) path.W *= 1.0f / (float)(1 + fails);
                    path.exclude = doNEE;
                    path.depth = depth + 1;
                    path.o = s.pos;
                    path.d = nd;
                    path.diff = next;
                    path.n = rng.n;
                }
            }
        }
    });
    uint32_t slot_out = append(count_out, cont);
    if (cont) store_path(Qout, slot_out, path, med);
    uint32_t slot_nee = append(nee_count, nee_out);
    if (nee_out) store_nee(NQ, slot_nee, path, nee_rec);
    if (nj > 1) {
        const int myjob = live ? path.job : -1;
        job_stat_add<1>(J, myjob, live ? 1u : 0u, RM_STAT_EXTENSION_RAYS);
        job_stat_add<1>(J, myjob, hit_any ? 1u : 0u, RM_STAT_HITS);
        job_stat_add<2>(J, myjob, ovf, RM_STAT_MEDIUM_OVERFLOW);
        job_stat_max<5>(J, cont ? path.job : -1, cont ? (uint32_t)med.n : 0u, RM_STAT_MEDIUM_MAX);
    } else if (live) {
        const uint32_t x[4] = {1u, hit_any ? 1u : 0u, ovf, cont ? (uint32_t)med.n : 0u};
        st.note(J, keys, path.job, x);
    }
  }
  st.flush_wave(J, keys);
}

// sampleDirectLight at the last vertex of every path that ended with
// NEE-and-stop this round (src/render.cpp:147-160,229-240 with
// src/sampling.cpp:467-527): up to sampleCount[depth] light samples, each
// with its shadow ray, scaled by the branch factor, into the path's set.
template <int MINW, int TPS>
__global__ __launch_bounds__(RB, MINW) void k_nee(const Job *__restrict__ J, int nj, PixBufs P, AccBufs AB, NeeQ NQ,
                                            const uint32_t *count, uint32_t *work) {
  const uint32_t n = *count;
  const uint32_t lane = threadIdx.x & 63u;
  const int keys[1] = {RM_STAT_SHADOW_RAYS};
  __shared__ uint32_t stat_rows[2 * RB];
  LaneStats<1, 0u> st;
  st.init(stat_rows);
  while (true) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(work, 64u);
    base = __shfl(base, 0, 64);
    if (base >= n) break;
    const uint32_t q = base + lane;
    const bool live = q < n;
    float4 f5 = make_float4(0, 0, 0, 0);
    if (live) f5 = NQ.f[5][q];
    const uint32_t flags = bu(f5.w);
    uint32_t nrays = 0;
    for_each_job(live ? (int)(flags >> 8) : -1, live, [&](int jid) {
        const Job &jb = J[jid];
        const SceneView &S = jb.S;
        const RArgs &A = jb.A;
        float4 f0 = NQ.f[0][q], f1 = NQ.f[1][q], f2 = NQ.f[2][q], f3 = NQ.f[3][q], f4 = NQ.f[4][q];
        float4 f6 = NQ.f[6][q], f7 = NQ.f[7][q], f8 = NQ.f[8][q];
        Surf s = surf_default();
        s.pos = V3{f0.x, f0.y, f0.z};
        s.shapeN = V3{f1.x, f1.y, f1.z};
        s.surfN = V3{f2.x, f2.y, f2.z};
        s.base = V3{f4.x, f4.y, f4.z};
        s.roughness = f4.w;
        s.metallic = f5.x;
        s.opacity = f5.y;
        s.eta = f5.z;
        s.entering = (flags & 1u) != 0;
        s.specular = f8.w;
        const V3 V{f3.x, f3.y, f3.z};
        const uint32_t pix = bu(f0.w), set = bu(f1.w), smp = bu(f2.w);
        const bool extra_on = (flags & 2u) != 0;
        const int cnt = (int)((flags >> 2) & 63u);
        const V3 T{f6.x, f6.y, f6.z}, p0{f7.x, f7.y, f7.z}, extra{f8.x, f8.y, f8.z};
        const float W = f6.w, factor = f7.w;
        Stream rng{A.seed, TAG_INDIRECT, (uint32_t)(pix - jb.pix_off), smp, bu(f3.w)};
        float total = direct_total(S, s, V);
        if (total != 0.0f) {
            uint32_t k = 0;
            TraceCounts tc{};
            for (int Ti = 0; Ti < cnt; ++Ti) {
                LSample x;
                if (direct_sample<false, true, TPS>(S, s, V, total, cnt, rng, x, tc, nullptr, &nrays)) {
                    x.bsdf *= factor;  // scaling(samples, factor)
                    add_light_sample(P, AB, pix, set, smp, k, T, W, p0, x, extra_on, extra, rec_slot(jb, pix, smp));
                    ++k;
                }
            }
        }
    });
    if (nj > 1) {
        job_stat_add<3>(J, live ? (int)(flags >> 8) : -1, nrays, RM_STAT_SHADOW_RAYS);
    } else if (live) {
        const uint32_t x[1] = {nrays};
        st.note(J, keys, (int)(flags >> 8), x);
    }
  }
  st.flush_wave(J, keys);
}

template <int MINW, int TPS>
__global__ __launch_bounds__(RB, MINW) void k_nee_gen(const Job *__restrict__ J, int nj, NeeQ NQ, const uint32_t *count,
                                                uint32_t *work, ShadowQ SQ, int queue_depth, uint32_t *queued) {
  const uint32_t n = *count;
  const uint32_t lane = threadIdx.x & 63u;
  const int keys[1] = {RM_STAT_SHADOW_RAYS};
  __shared__ uint32_t stat_rows[2 * RB];
  // the queued rays of this lane's record, (direction, aim) per sample: the
  // wave appends them all with one atomic after the record's samples are drawn
  // (appends to one counter from every wave serialise on its address)
  __shared__ float4 rbuf[NEE_MAX * RB];
  LaneStats<1, 0u> st;
  st.init(stat_rows);
  Claims cl(n);
  while (true) {
    const uint32_t base = cl.take(work);
    if (base >= n) break;
    const uint32_t q = base + lane;
    const bool live = q < n;
    float4 f5 = make_float4(0, 0, 0, 0);
    if (live) f5 = NQ.f[5][q];
    const uint32_t flags = bu(f5.w);
    uint32_t nrays = 0, qmask = 0;
    V3 org = v3(0.0f);
    for_each_job(live ? (int)(flags >> 8) : -1, live, [&](int jid) {
        const Job &jb = J[jid];
        const SceneView &S = jb.S;
        const RArgs &A = jb.A;
        float4 f0 = NQ.f[0][q], f1 = NQ.f[1][q], f2 = NQ.f[2][q], f3 = NQ.f[3][q], f4 = NQ.f[4][q];
        float4 f8 = NQ.f[8][q];
        Surf s = surf_default();
        s.pos = V3{f0.x, f0.y, f0.z};
        s.shapeN = V3{f1.x, f1.y, f1.z};
        s.surfN = V3{f2.x, f2.y, f2.z};
        s.base = V3{f4.x, f4.y, f4.z};
        s.roughness = f4.w;
        s.metallic = f5.x;
        s.opacity = f5.y;
        s.eta = f5.z;
        s.entering = (flags & 1u) != 0;
        s.specular = f8.w;
        const V3 V{f3.x, f3.y, f3.z};
        const uint32_t pix = bu(f0.w), smp = bu(f2.w);
        const int cnt = (int)((flags >> 2) & 63u);
        const float factor = NQ.f[7][q].w;
        // k_shadow carries only the walk's part of the scene, so shadow rays of
        // a scene with alpha cut-outs are traced here, through Model::rayHit_test
        const bool to_queue = S.depth >= queue_depth && !S.cutout;
        org = s.pos;
        Stream rng{A.seed, TAG_INDIRECT, (uint32_t)(pix - jb.pix_off), smp, bu(f3.w)};
        const float total = direct_total(S, s, V);
        for (int Ti = 0; Ti < NEE_MAX; ++Ti) {
            const int64_t e = (int64_t)q * NEE_MAX + Ti;
            ShadowRay sr;
            LSample x;
            if (total != 0.0f && Ti < cnt && direct_ray(S, s, V, total, cnt, rng, sr, x)) {
                x.bsdf *= factor;  // scaling(samples, factor)
                const V3 L = x.light * x.bsdf;
                SQ.x[e] = make_float4(L.x, L.y, L.z, x.weight);
                if (to_queue) {
                    rbuf[Ti * RB + threadIdx.x] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, sr.aim);
                    qmask |= 1u << Ti;
                } else {
                    RayPre r = ray_pre(s.pos, sr.dir);
                    SQ.ref[e] = model_clear<true, TPS>(S, s.pos, sr.dir, r, sr.aim) ? SH_CLEAR : SH_BLOCKED;
                }
                ++nrays;
            } else {
                SQ.ref[e] = SH_NONE;
            }
        }
    });
    // one append for the wave: lane l's rays take slots base + (rays of the
    // lanes below l), in sample order
    const uint32_t c = (uint32_t)__popc(qmask);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t rank = 0, wtotal = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const uint64_t m = __ballot((c >> b) & 1u);
        rank += (uint32_t)__popcll(m & below) << b;
        wtotal += (uint32_t)__popcll(m) << b;
    }
    if (wtotal) {
        uint32_t wb = 0;
        if (lane == 0) wb = atomicAdd(queued, wtotal);
        uint32_t slot = __shfl(wb, 0, 64) + rank;
        const uint32_t job = flags >> 8;
        for (int Ti = 0; Ti < NEE_MAX; ++Ti) {
            if (!((qmask >> Ti) & 1u)) continue;
            const float4 r = rbuf[Ti * RB + threadIdx.x];
            SQ.o[slot] = make_float4(org.x, org.y, org.z, r.w);
            SQ.d[slot] = make_float4(r.x, r.y, r.z, ub(job));
            SQ.ref[(int64_t)q * NEE_MAX + Ti] = slot;
            ++slot;
        }
    }
    if (nj > 1) {
        job_stat_add<3>(J, live ? (int)(flags >> 8) : -1, nrays, RM_STAT_SHADOW_RAYS);
    } else if (live) {
        const uint32_t x[1] = {nrays};
        st.note(J, keys, (int)(flags >> 8), x);
    }
  }
  st.flush_wave(J, keys);
}

// Any-hit walk of every shadow ray of the round (the k_extend scheme with
// ANY: it stops once the running closest t is below aim).
template <bool LDS, int TPS, bool COUNT = false>
__global__ __launch_bounds__(RB) void k_shadow(const Job *__restrict__ J, ShadowQ SQ, const uint32_t *count,
                                               uint32_t *work, int fetch_min, int lds_entries, int phases) {
    const uint32_t n = *count;
    const uint32_t lane = threadIdx.x & 63u;
    ScratchEntries ent;
    Walker<true, COUNT, typename std::conditional<LDS, LdsStack, PrivateStack>::type, TPS> w;
    if constexpr (LDS)
        w.stk.bind(&ent, lds_entries);
    else
        w.stk.bind(&ent);
    RayPre r;
    SceneView S;  // per lane: the nodes, triangles and root of its ray's scene
    float aim = 0.0f;
    int64_t q = -1;
    uint32_t job = 0;
    bool exhausted = false;
    TraceCounts c{};
    while (true) {
        bool idle = q < 0;
        uint64_t idle_mask = __ballot(idle);
        uint64_t busy_mask = __ballot(!idle);
        if (exhausted && busy_mask == 0) break;
        if (!exhausted && idle_mask && (__popcll(idle_mask) >= fetch_min || busy_mask == 0)) {
            uint32_t cnt = (uint32_t)__popcll(idle_mask);
            uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle_mask) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(work, cnt);
            base = __shfl(base, (int)leader, 64);
            if (base + cnt >= n) exhausted = true;
            if (idle) {
                uint32_t slot = base + (uint32_t)__popcll(idle_mask & ((1ull << lane) - 1ull));
                if (slot < n) {
                    q = slot;
                    const float4 oo = SQ.o[slot], dd = SQ.d[slot];
                    const uint32_t nj = bu(dd.w);
                    if (COUNT && nj != job) add_walk_counts(J[job].F.stats, c);
                    job = nj;
                    const SceneView &SJ = J[job].S;
                    S.nodes = SJ.nodes;
                    S.tris = SJ.tris;
                    S.root_leaf = SJ.root_leaf;
                    aim = oo.w;
                    r = ray_pre(V3{oo.x, oo.y, oo.z}, V3{dd.x, dd.y, dd.z});
                    w.init(S, r, aim + T_EPS);
                }
            }
        }
        const bool turn = walk_turn(w.cur, phases, q >= 0);
        if (q >= 0 && turn && w.step(S, r, T_EPS, aim, c)) {
            SQ.vis[q] = (w.tmax < aim) ? 0 : 1;
            q = -1;
        }
    }
    if (COUNT) flush_walk_counts(J, (int)job, c);
}

template <int MINW>
__global__ __launch_bounds__(RB, MINW) void k_nee_add(const Job *__restrict__ J, PixBufs P, AccBufs AB, NeeQ NQ,
                                                const uint32_t *count, uint32_t *work, ShadowQ SQ) {
  const uint32_t n = *count;
  const uint32_t lane = threadIdx.x & 63u;
  Claims cl(n);
  while (true) {
    const uint32_t base = cl.take(work);
    if (base >= n) break;
    const uint32_t q = base + lane;
    const bool live = q < n;
    float4 f5 = make_float4(0, 0, 0, 0);
    if (live) f5 = NQ.f[5][q];
    const uint32_t flags = bu(f5.w);
    for_each_job(live ? (int)(flags >> 8) : -1, live, [&](int jid) {
        const Job &jb = J[jid];
        const uint32_t pix = bu(NQ.f[0][q].w), set = bu(NQ.f[1][q].w), smp = bu(NQ.f[2][q].w);
        const bool extra_on = (flags & 2u) != 0;
        const int cnt = (int)((flags >> 2) & 63u);
        const float4 f6 = NQ.f[6][q], f7 = NQ.f[7][q], f8 = NQ.f[8][q];
        const V3 T{f6.x, f6.y, f6.z}, p0{f7.x, f7.y, f7.z}, extra{f8.x, f8.y, f8.z};
        const float W = f6.w;
        uint32_t k = 0;
        for (int Ti = 0; Ti < cnt; ++Ti) {
            const int64_t e = (int64_t)q * NEE_MAX + Ti;
            if (!shadow_clear(SQ, SQ.ref[e])) continue;
            const float4 xx = SQ.x[e];
            // light * bsdf was formed when the sample was drawn; times 1 is exact
            const LSample x{V3{xx.x, xx.y, xx.z}, v3(1.0f), xx.w};
            add_light_sample(P, AB, pix, set, smp, k, T, W, p0, x, extra_on, extra, rec_slot(jb, pix, smp));
            ++k;
        }
    });
  }
}

__global__ __launch_bounds__(RB) void k_clear_sets(AccBufs AB, uint32_t n) {
    uint32_t g = blockIdx.x * RB + threadIdx.x;
    if (g >= n) return;
    float4 z = make_float4(0, 0, 0, 0);
    AB.a[0][g] = z;
    AB.a[1][g] = z;
    AB.a[2][g] = make_float4(0.0f, -1.0f, 0.0f, 0.0f);
    AB.a[3][g] = z;
    AB.a[4][g] = z;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Adds job blockIdx.x's counter shards into its frame's counters (the
// maximum for RM_STAT_MEDIUM_MAX).
__global__ void k_stats_reduce(const Job *__restrict__ J) {
    const Job &jb = J[blockIdx.x];
    const int k = (int)threadIdx.x;
    if (!jb.xs || k >= RM_STAT_COUNT) return;
    unsigned long long v = 0;
    for (int w = 0; w < NSHARD; ++w) {
        const unsigned long long x = jb.xs[(size_t)w * (size_t)jb.xstride + k];
        v = k == RM_STAT_MEDIUM_MAX ? (x > v ? x : v) : v + x;
    }
    if (!v) return;
    if (k == RM_STAT_MEDIUM_MAX)
        atomicMax(jb.F.stats + k, v);
    else
        atomicAdd(jb.F.stats + k, v);
}

__global__ void k_zero_u32(uint32_t *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_zero_f4(float4 *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ __launch_bounds__(RB) void k_finish(const Job *__restrict__ J, int nj, int64_t total_px, PixBufs P,
                                               AccBufs AB) {
    const int64_t gp = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (gp >= total_px) return;
    const Job &jb = J[job_of(J, nj, gp)];
    const RArgs &A = jb.A;
    const FrameView &F = jb.F;
    const int64_t p = gp - jb.pix_off;
    int M = P.M[gp];
    Acc4 d{0, 0, 0, 0}, s{0, 0, 0, 0};
    int32_t dropped = -1;
    if (M > 0) {
        int off = P.set_off[gp];
        float S = 0.0f, cnt = 0.0f, maxC = -1.0f;
        uint32_t maxid = 0;
        Acc4 md{0, 0, 0, 0}, ms{0, 0, 0, 0};
        for (int k = 0; k < M; ++k) {
            SetAcc a = load_set(AB, (uint32_t)(off + k));
            d.r += a.d.r; d.g += a.d.g; d.b += a.d.b; d.v += a.d.v;
            s.r += a.s.r; s.g += a.s.g; s.b += a.s.b; s.v += a.s.v;
            S += a.S;
            cnt += a.cnt;
            if (a.maxC > maxC) { maxC = a.maxC; maxid = a.maxid; md = a.md; ms = a.ms; }
        }
        if (cnt > 0.0f) {
            if (maxC >= 0.0f && maxC / (S - maxC + T_EPS) > 16.0f) {
                d.r -= md.r; d.g -= md.g; d.b -= md.b; d.v -= md.v;
                s.r -= ms.r; s.g -= ms.g; s.b -= ms.b; s.v -= ms.v;
                dropped = (int32_t)maxid;
            }
            calc_var(d, A.exposure);
            calc_var(s, A.exposure);
            float4 sv = P.pi2[gp];
            st3(F.base, p, V3{sv.x, sv.y, sv.z});
        }
    }
    reinterpret_cast<float4 *>(F.Id)[p] = make_float4(d.r, d.g, d.b, d.v);
    reinterpret_cast<float4 *>(F.Is)[p] = make_float4(s.r, s.g, s.b, s.v);
    F.drop[p] = dropped;
}

// Exclusive prefix of M over pixels, two passes with the block totals scanned
// on the host, then the set -> pixel map.
constexpr int SB = 1024;
__global__ __launch_bounds__(SB) void k_block_sums(const int32_t *M, int64_t n, int64_t *sums) {
    __shared__ int64_t sh[SB / 64];
    int64_t i = (int64_t)blockIdx.x * SB + threadIdx.x;
    int64_t v = i < n ? M[i] : 0;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int k = 0; k < SB / 64; ++k) t += sh[k];
        sums[blockIdx.x] = t;
    }
}
__global__ __launch_bounds__(SB) void k_block_scan(const int32_t *M, int64_t n, const int64_t *block_off, int32_t *off,
                                                   int32_t *set_pix) {
    __shared__ int32_t sh[SB];
    int64_t i = (int64_t)blockIdx.x * SB + threadIdx.x;
    int32_t v = i < n ? M[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < SB; s <<= 1) {
        int32_t add = threadIdx.x >= (unsigned)s ? sh[threadIdx.x - s] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    if (i < n) {
        int32_t o = (int32_t)block_off[blockIdx.x] + sh[threadIdx.x] - v;
        off[i] = o;
        for (int k = 0; k < v; ++k) set_pix[o + k] = (int32_t)i;
    }
}

template <class T>
T *carve(char *&p, size_t count) {
    T *r = reinterpret_cast<T *>(p);
    p += (count * sizeof(T) + 255) & ~(size_t)255;
    return r;
}

// ---------------------------------------------------------------- host side
struct HostJob {
    const rm_scene *scene;
    const rm_render_args *args;
    rm_frame *frame;
    int post;  // rm_postprocess options, 0 = none
};

// all_px: pixels of the whole launch; the automatic paths-per-round keeps
// the live accumulator sets of a launch near 16M (a single 1080p frame: 8
// paths per round; more rounds of fewer paths cost a launch sequence each,
// and every path slot holds about 1 KB of work buffers).
static int32_t make_job(const HostJob &h, Job &J, int &rounds, int &maxd, int64_t all_px, const char *fn) {
    const rm_render_args *a = h.args;
    rm_frame *f = h.frame;
    if (!h.scene || !a || !f) return fail(RM_ERR_INVALID, "%s: NULL argument", fn);
    if (a->width != f->w || a->height != f->h)
        return fail(RM_ERR_INVALID, "%s: frame is %dx%d, args ask %dx%d", fn, f->w, f->h, a->width, a->height);
    if (a->spp < 0) return fail(RM_ERR_INVALID, "%s: negative spp", fn);
    int max_depth = a->max_depth > 0 ? a->max_depth : 16;
    if (max_depth > 16) return fail(RM_ERR_INVALID, "%s: max_depth %d exceeds 16 (sampleCount table)", fn, max_depth);
    if (!(a->p_direct >= 0.0f && a->p_direct <= 1.0f)) return fail(RM_ERR_INVALID, "%s: p_direct outside [0,1]", fn);
    if (h.scene->view.depth > MAX_STACK) return fail(RM_ERR_UNSUPPORTED, "%s: tree too deep", fn);
    const int64_t np = (int64_t)f->w * f->h;
    if (f->drop.bytes < (size_t)np * 4) {
        int32_t rc = f->drop.alloc((size_t)np * 4);
        if (rc) return rc;
    }
    RArgs &A = J.A;
    A = rargs_from(a);
    int K = a->paths_per_round;
    if (K <= 0) K = (int)std::min<int64_t>(16, std::max<int64_t>(1, (int64_t)(16 << 20) / std::max<int64_t>(1, all_px)));
    A.K = K;
    const int spp_direct = (int)(float(a->spp) * a->p_direct);
    const int n_opaque = a->spp - spp_direct;
    rounds = n_opaque > 0 ? (n_opaque + K - 1) / K : 0;
    maxd = max_depth;
    J.S = h.scene->view;
    FrameView &F = J.F;
    F.pos = f->buf[RM_BUF_POSITION].as<float>();
    F.shapeN = f->buf[RM_BUF_SHAPE_NORMAL].as<float>();
    F.surfN = f->buf[RM_BUF_SURFACE_NORMAL].as<float>();
    F.emis = f->buf[RM_BUF_EMISSION].as<float>();
    F.base = f->buf[RM_BUF_BASE_COLOR].as<float>();
    F.mat = f->buf[RM_BUF_MATERIAL].as<float>();
    F.Dd = f->buf[RM_BUF_DD].as<float>();
    F.Ds = f->buf[RM_BUF_DS].as<float>();
    F.Id = f->buf[RM_BUF_ID].as<float>();
    F.Is = f->buf[RM_BUF_IS].as<float>();
    F.drop = f->drop.as<int32_t>();
    F.stats = f->stats.as<unsigned long long>();
    F.rec = nullptr;
    F.rec_stride = 0;
    f->rec_valid = false;
    if (f->record) {
        // one record per (pixel, indirect sample); only for scenes without
        // transmission (16x sample budgets on transparent pixels)
        bool opaque = true;
        for (const MatRec &m : h.scene->host.mats) opaque = opaque && m.opacity > 1.0f - T_EPS;
        const int stride = a->spp - (int)(float(a->spp) * a->p_direct);
        if (opaque && stride > 0) {
            const size_t need = (size_t)np * stride * 48;
            if (f->rec.bytes < need) {
                int32_t rc = f->rec.alloc(need);
                if (rc) return rc;
            }
            F.rec = f->rec.as<float4>();
            F.rec_stride = stride;
            f->rec_stride = stride;
        }
    }
    J.np = np;
    return RM_OK;
}

// Renders every job in one set of launches over their concatenated pixels.
// Host round trip: one, after the primary pass (the accumulator-set total
// sizes the work buffer).  Everything after it is a fixed launch sequence
// with device-side queue counts; with `graph` it is captured once into a
// hipGraph and launched as one graph.
// Primary-vertex light samples per pixel of a job (spp * P_Direct).
static int direct_count(const HostJob &h) {
    return h.args ? std::max(0, (int)(float(h.args->spp) * h.args->p_direct)) : 0;
}

static int32_t render_jobs(const std::vector<HostJob> &hj_in, DevBuf &work, hipStream_t st, bool graph) {
    const int nj = (int)hj_in.size();
    if (nj <= 0) return RM_OK;
    if (nj > (1 << 23)) return fail(RM_ERR_INVALID, "too many jobs");
    // Jobs in order of falling light samples per pixel, so that each chunk of
    // the split primary-vertex pass covers a prefix of the pixels.  Every job's
    // result is independent of the order.
    std::vector<HostJob> hj(hj_in);
    std::stable_sort(hj.begin(), hj.end(),
                     [](const HostJob &a, const HostJob &b) { return direct_count(a) > direct_count(b); });
    std::vector<Job> J(nj);
    int R = 0, D = 1;
    int64_t all_px = 0;
    for (int i = 0; i < nj; ++i)
        if (hj[i].args) all_px += (int64_t)std::max(0, hj[i].args->width) * std::max(0, hj[i].args->height);
    int64_t total_px = 0;
    bool maps = false;  // the shading kernels with 8-bit texture lookups
    for (int i = 0; i < nj; ++i) {
        int r = 0, d = 0;
        int32_t rc = make_job(hj[i], J[i], r, d, all_px, nj > 1 ? "rm_render_batch" : "rm_render");
        if (rc) return rc;
        maps = maps || hj[i].scene->host.maps8 != 0;
        J[i].pix_off = total_px;
        total_px += J[i].np;
        R = std::max(R, r);
        D = std::max(D, d);
    }
    const int64_t nblk = (total_px + SB - 1) / SB;
    auto pad = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const bool shard = nj > 1;
    const size_t nshard = shard ? (size_t)NSHARD * nj * RM_STAT_COUNT : 0;
    const size_t pix_bytes = pad(nj * sizeof(Job)) + 3 * pad(total_px * 16) + 2 * pad(total_px * 4) + pad(nblk * 8) +
                             pad(4 * NCOUNT * 4) + pad(nshard * 8);
    if (work.bytes < pix_bytes) {
        int32_t rc = work.alloc(pix_bytes);
        if (rc) return rc;
    }
    char *cur = work.as<char>();
    Job *dJ = carve<Job>(cur, nj);
    PixBufs P;
    P.pi0 = carve<float4>(cur, total_px);
    P.pi1 = carve<float4>(cur, total_px);
    P.pi2 = carve<float4>(cur, total_px);
    P.M = carve<int32_t>(cur, total_px);
    P.set_off = carve<int32_t>(cur, total_px);
    int64_t *bsum = carve<int64_t>(cur, nblk);
    // counts[d]: paths entering depth d+1; counts[NCOUNT + d]: shading claims of
    // depth d; counts[2 NCOUNT + d]: traversal claims of depth d;
    // counts[3 NCOUNT]: NEE records of the round, counts[3 NCOUNT + 1..3]: claims of
    // the NEE kernels (one pass, or generate / shadow / add), counts[3 NCOUNT + 4]:
    // queued shadow rays; counts[3 NCOUNT + 8, 9]: the same for each chunk of
    // the primary-vertex pass
    uint32_t *counts = carve<uint32_t>(cur, 4 * NCOUNT);
    unsigned long long *shards = carve<unsigned long long>(cur, nshard);
    for (int i = 0; i < nj; ++i) {
        J[i].xs = shard ? shards + (size_t)i * RM_STAT_COUNT : nullptr;
        J[i].xstride = nj * RM_STAT_COUNT;
    }
    // Trace variant 3 keeps the first pending entries of every walk in LDS:
    // each kernel of the launch that traverses gets nl entries per lane (at
    // most the deepest tree's depth), and the job table's views say so.
    const bool lds_walk = tuning().trace_variant == 3;
    int max_depth = 1;
    for (int i = 0; i < nj; ++i) max_depth = std::max(max_depth, (int)J[i].S.depth);
    const int nl = lds_walk ? std::min(max_depth, tuning().lds_stack_entries) : 0;
    const size_t walk_lds = (size_t)nl * RB * 8;
    // LdsStack's launch contract: every kernel below that walks is launched
    // with blocks of RB lanes and walk_lds bytes of dynamic LDS
    if (RB != WALK_BLOCK || walk_lds < (size_t)nl * WALK_BLOCK * 8 || nl > MAX_STACK)
        return fail(RM_ERR_INVALID, "render: LDS walk launch (%d lanes, %zu bytes) does not hold %d entries", RB,
                    walk_lds, nl);
    for (int i = 0; i < nj; ++i) J[i].S.walk_lds = nl;
    RM_HIP(hipMemcpyAsync(dJ, J.data(), nj * sizeof(Job), hipMemcpyHostToDevice, st));
    for (int i = 0; i < nj; ++i)
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(RM_STAT_COUNT), 0, st, J[i].F.stats, (int)RM_STAT_COUNT);
    if (shard)
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((nshard + 255) / 256)), dim3(256), 0, st, shards, (int)nshard);
    for (int i = 0; i < nj; ++i)
        if (J[i].F.rec) {
            const int64_t n4 = J[i].np * J[i].F.rec_stride * 3;
            hipLaunchKernelGGL(k_zero_f4, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, st, J[i].F.rec, n4);
        }
    const unsigned gp = (unsigned)((total_px + RB - 1) / RB);
    // triangles per walk step: any of 0, 1, 2, 4 for the extension kernel;
    // the shading kernels' walks take a whole leaf (0) or two triangles
    const int tps = tuning().tris_per_step;
    hipLaunchKernelGGL(maps ? (tps ? k_primary<2, true> : k_primary<0, true>)
                            : (tps ? k_primary<2, false> : k_primary<0, false>),
                       dim3(gp), dim3(RB), walk_lds, st, (const Job *)dJ, nj, total_px, P);
    hipLaunchKernelGGL(k_block_sums, dim3((unsigned)nblk), dim3(SB), 0, st, P.M, total_px, bsum);
    RM_HIP(hipGetLastError());
    std::vector<int64_t> hs(nblk);
    RM_HIP(hipMemcpyAsync(hs.data(), bsum, nblk * 8, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int64_t b = 0; b < nblk; ++b) {
        int64_t v = hs[b];
        hs[b] = total;
        total += v;
    }
    if (total > 0x7fffffffLL) return fail(RM_ERR_UNSUPPORTED, "%lld path slots per round", (long long)total);
    // The split shadow passes' queue: NEE_MAX entries per path slot for the
    // NEE pass, C per pixel for each chunk of the primary-vertex pass (C up to
    // 32, such that a chunk fits the NEE pass's entries, or 16M entries when
    // those are fewer: fewer, larger chunks spare launches and walk tails).
    // k_shadow counts in 32 bits.
    int max_direct = 0;
    for (int i = 0; i < nj; ++i) max_direct = std::max(max_direct, (int)(float(J[i].A.spp) * J[i].A.p_direct));
    const int64_t dcap = std::max<int64_t>((int64_t)16 << 20, total * NEE_MAX);
    const int dC = (int)std::max<int64_t>(1, std::min<int64_t>(32, dcap / std::max<int64_t>(1, total_px)));
    // split: 1 queues every shadow ray; 2 queues those of jobs whose tree has
    // SPLIT_DEPTH levels or more and traces the others where they are drawn
    // (on shallow trees a walk costs less than queueing it), and keeps the
    // single-kernel passes when no job is that deep
    constexpr int SPLIT_DEPTH = 12;
    // counting renders send every shadow ray through the counting k_shadow
    const bool counting = tuning().render_count != 0;
    const int split_mode = counting ? 1 : tuning().nee_split;
    const bool split = (split_mode == 1 || (split_mode == 2 && max_depth >= SPLIT_DEPTH)) &&
                       total * NEE_MAX <= 0xffffffffLL && total_px * dC <= 0xffffffffLL;
    const int queue_depth = split_mode == 1 ? 0 : SPLIT_DEPTH;
    const int64_t ndirect = split ? total_px * dC : 0;
    const int64_t nshadow = split ? std::max<int64_t>(total * NEE_MAX, ndirect) : 0;
    const size_t set_bytes = pad(total * 4) + 5 * pad(total * 16) + 24 * pad(total * 16) + pad(total * 8) +
                             9 * pad(total * 16) + 3 * pad(nshadow * 16) + pad(nshadow * 4) + pad(nshadow) +
                             pad(ndirect * 16);
    const size_t need = pix_bytes + set_bytes;
    // Grow when short.  Give memory back only when the buffer exceeds twice
    // the largest need of the last 64 launches (a kept workspace would
    // otherwise hold its largest launch's size for good, while jobs of mixed
    // sizes must not reallocate back and forth).  The buffer holds the
    // primary pass's results, so a reallocation starts over.
    work.window_max = std::max(work.window_max, need);
    size_t target = 0;
    if (work.bytes < need) {
        target = need;
    } else if (++work.window_calls >= 64) {
        if (work.bytes > 2 * work.window_max + ((size_t)1 << 30)) target = work.window_max;
        work.window_max = 0;
        work.window_calls = 0;
    }
    if (target) {
        int32_t rc = work.alloc(target);
        if (rc) return rc;
        return render_jobs(hj, work, st, graph);
    }
    for (int i = 0; i < nj; ++i)
        if (hj[i].post) {
            int32_t rc = post_prepare(hj[i].frame, hj[i].post, true);
            if (rc) return rc;
        }
    RM_HIP(hipMemcpyAsync(bsum, hs.data(), nblk * 8, hipMemcpyHostToDevice, st));
    RM_HIP(hipStreamSynchronize(st));  // `hs` is pageable host memory owned by this call
    P.set_pix = carve<int32_t>(cur, total);
    AccBufs AB;
    for (int k = 0; k < 5; ++k) AB.a[k] = carve<float4>(cur, total);
    PathQ Q[2];
    for (int qi = 0; qi < 2; ++qi)
        for (int k = 0; k < 12; ++k) Q[qi].f[k] = carve<float4>(cur, total);
    uint2 *hits = carve<uint2>(cur, total);
    NeeQ NQ;
    for (int k = 0; k < 9; ++k) NQ.f[k] = carve<float4>(cur, total);
    ShadowQ SQ;
    SQ.o = carve<float4>(cur, nshadow);
    SQ.d = carve<float4>(cur, nshadow);
    SQ.x = carve<float4>(cur, nshadow);
    SQ.ref = carve<uint32_t>(cur, nshadow);
    SQ.vis = carve<uint8_t>(cur, nshadow);
    SQ.x2 = carve<float4>(cur, ndirect);
    hipLaunchKernelGGL(k_block_scan, dim3((unsigned)nblk), dim3(SB), 0, st, P.M, total_px, bsum, P.set_off, P.set_pix);
    RM_HIP(hipGetLastError());

    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    if (graph) RM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int mw = tuning().render_min_waves;  // 1, 3 or 4
    const unsigned gpers = (unsigned)(device_cu_count() * tuning().persistent_blocks_per_cu);
    const int fetch_min = tuning().fetch_min;
    const int phases = tuning().walk_phases;
    auto shadow = counting ? (lds_walk ? (tps == 1 ? k_shadow<true, 1, true> : tps == 2 ? k_shadow<true, 2, true>
                                          : tps == 4 ? k_shadow<true, 4, true> : k_shadow<true, 0, true>)
                                       : (tps == 1 ? k_shadow<false, 1, true> : tps == 2 ? k_shadow<false, 2, true>
                                          : tps == 4 ? k_shadow<false, 4, true> : k_shadow<false, 0, true>))
                  : (lds_walk ? (tps == 1 ? k_shadow<true, 1> : tps == 2 ? k_shadow<true, 2>
                                 : tps == 4 ? k_shadow<true, 4> : k_shadow<true, 0>)
                              : (tps == 1 ? k_shadow<false, 1> : tps == 2 ? k_shadow<false, 2>
                                 : tps == 4 ? k_shadow<false, 4> : k_shadow<false, 0>));
    if (split) {
        uint32_t *claims = counts + 3 * NCOUNT + 8;  // and [9]: queued rays; both zeroed for each chunk
        auto direct_gen = mw == 4 ? (tps ? k_direct_gen<4, 2> : k_direct_gen<4, 0>)
                          : mw == 3 ? (tps ? k_direct_gen<3, 2> : k_direct_gen<3, 0>)
                                    : (tps ? k_direct_gen<1, 2> : k_direct_gen<1, 0>);
        // samples [t0, t0 + C) of the pixels [0, npx) whose jobs have more than
        // t0 samples (jobs are in order of falling sample counts)
        for (int t0 = 0; t0 < max_direct;) {
            int64_t npx = total_px;
            for (int i = 0; i < nj; ++i)
                if (direct_count(hj[i]) <= t0) { npx = J[i].pix_off; break; }
            const int C = (int)std::max<int64_t>(
                1, std::min<int64_t>({(int64_t)max_direct - t0, ndirect / std::max<int64_t>(1, npx)}));
            const int64_t n = npx * C;
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(128), 0, st, claims, 2);
            hipLaunchKernelGGL(direct_gen, dim3((unsigned)((n + RB - 1) / RB)), dim3(RB), walk_lds, st,
                               (const Job *)dJ, nj, npx, P, C, t0, SQ, queue_depth, claims + 1);
            hipLaunchKernelGGL(shadow, dim3(gpers), dim3(RB), walk_lds, st, (const Job *)dJ, SQ,
                               (const uint32_t *)(claims + 1), claims, fetch_min, nl, phases);
            hipLaunchKernelGGL(k_direct_add, dim3((unsigned)((npx + RB - 1) / RB)), dim3(RB), 0, st, (const Job *)dJ, nj,
                               npx, P, C, t0, SQ);
            t0 += C;
        }
        hipLaunchKernelGGL(k_direct_final, dim3(gp), dim3(RB), 0, st, (const Job *)dJ, nj, total_px);
    } else {
        auto direct = mw == 4 ? (tps ? k_direct<4, 2> : k_direct<4, 0>)
                      : mw == 3 ? (tps ? k_direct<3, 2> : k_direct<3, 0>) : (tps ? k_direct<1, 2> : k_direct<1, 0>);
        hipLaunchKernelGGL(direct, dim3(gp), dim3(RB), walk_lds, st, (const Job *)dJ, nj, total_px, P);
    }
    const uint32_t nsets = (uint32_t)total;
    if (nsets > 0) {
        const unsigned gs = (nsets + RB - 1) / RB;
        const unsigned gb = (unsigned)(device_cu_count() * tuning().render_blocks_per_cu);
        hipLaunchKernelGGL(k_clear_sets, dim3(gs), dim3(RB), 0, st, AB, nsets);
        const unsigned ge = gpers;
        auto bounce = maps ? (mw == 4 ? k_bounce<4, true> : mw == 3 ? k_bounce<3, true> : k_bounce<1, true>)
                           : (mw == 4 ? k_bounce<4, false> : mw == 3 ? k_bounce<3, false> : k_bounce<1, false>);
        auto nee = mw == 4 ? (tps ? k_nee<4, 2> : k_nee<4, 0>)
                   : mw == 3 ? (tps ? k_nee<3, 2> : k_nee<3, 0>) : (tps ? k_nee<1, 2> : k_nee<1, 0>);
        auto nee_gen = mw == 4 ? (tps ? k_nee_gen<4, 2> : k_nee_gen<4, 0>)
                       : mw == 3 ? (tps ? k_nee_gen<3, 2> : k_nee_gen<3, 0>) : (tps ? k_nee_gen<1, 2> : k_nee_gen<1, 0>);
        auto nee_add = mw == 4 ? k_nee_add<4> : mw == 3 ? k_nee_add<3> : k_nee_add<1>;
        auto extend_lds = counting ? (tps == 1 ? k_extend<true, 1, true> : tps == 2 ? k_extend<true, 2, true>
                                      : tps == 4 ? k_extend<true, 4, true> : k_extend<true, 0, true>)
                                   : (tps == 1 ? k_extend<true, 1> : tps == 2 ? k_extend<true, 2>
                                      : tps == 4 ? k_extend<true, 4> : k_extend<true, 0>);
        auto extend_scratch = counting ? (tps == 1 ? k_extend<false, 1, true> : tps == 2 ? k_extend<false, 2, true>
                                          : tps == 4 ? k_extend<false, 4, true> : k_extend<false, 0, true>)
                                       : (tps == 1 ? k_extend<false, 1> : tps == 2 ? k_extend<false, 2>
                                          : tps == 4 ? k_extend<false, 4> : k_extend<false, 0>);
        for (int j = 0; j < R; ++j) {
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(128), 0, st, counts, 4 * NCOUNT);
            hipLaunchKernelGGL(k_spawn, dim3(gs), dim3(RB), 0, st, (const Job *)dJ, nj, P, j, nsets, Q[0], counts);
            for (int d = 1; d <= D; ++d) {
                const uint32_t *n_in = counts + d - 1;
                if (lds_walk)
                    hipLaunchKernelGGL(extend_lds, dim3(ge), dim3(RB), walk_lds, st, (const Job *)dJ,
                                       Q[(d - 1) & 1], n_in, counts + 2 * NCOUNT + d, hits, fetch_min, nl, phases);
                else
                    hipLaunchKernelGGL(extend_scratch, dim3(ge), dim3(RB), 0, st, (const Job *)dJ, Q[(d - 1) & 1],
                                       n_in, counts + 2 * NCOUNT + d, hits, fetch_min, 0, phases);
                hipLaunchKernelGGL(bounce, dim3(gb), dim3(RB), 0, st, (const Job *)dJ, nj, P, AB, Q[(d - 1) & 1], n_in,
                                   counts + NCOUNT + d, (const uint2 *)hits, Q[d & 1], counts + d, NQ,
                                   counts + 3 * NCOUNT);
            }
            const uint32_t *n_nee = counts + 3 * NCOUNT;
            if (split) {
                hipLaunchKernelGGL(nee_gen, dim3(gb), dim3(RB), walk_lds, st, (const Job *)dJ, nj, NQ, n_nee,
                                   counts + 3 * NCOUNT + 1, SQ, queue_depth, counts + 3 * NCOUNT + 4);
                hipLaunchKernelGGL(shadow, dim3(ge), dim3(RB), walk_lds, st, (const Job *)dJ, SQ,
                                   (const uint32_t *)(counts + 3 * NCOUNT + 4), counts + 3 * NCOUNT + 2, fetch_min, nl,
                                   phases);
                hipLaunchKernelGGL(nee_add, dim3(gb), dim3(RB), 0, st, (const Job *)dJ, P, AB, NQ, n_nee,
                                   counts + 3 * NCOUNT + 3, SQ);
            } else {
                hipLaunchKernelGGL(nee, dim3(gb), dim3(RB), walk_lds, st, (const Job *)dJ, nj, P, AB, NQ, n_nee,
                                   counts + 3 * NCOUNT + 1);
            }
        }
    }
    hipLaunchKernelGGL(k_finish, dim3(gp), dim3(RB), 0, st, (const Job *)dJ, nj, total_px, P, AB);
    if (shard) hipLaunchKernelGGL(k_stats_reduce, dim3(nj), dim3(RM_STAT_COUNT), 0, st, (const Job *)dJ);
    for (int i = 0; i < nj; ++i)
        if (hj[i].post) {
            int32_t rc = post_launch(hj[i].frame, hj[i].post, hj[i].args->exposure, st);
            if (rc) {
                if (graph) {
                    hipGraph_t tmp;
                    (void)hipStreamEndCapture(st, &tmp);
                    if (tmp) (void)hipGraphDestroy(tmp);
                }
                return rc;
            }
        }
    if (graph) {
        RM_HIP(hipStreamEndCapture(st, &g));
        hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return hip_fail(e, "hipGraphInstantiate");
        }
        e = hipGraphLaunch(ge, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
    }
    RM_HIP(hipGetLastError());
    // what each frame now holds, for rm_adjoint
    for (int i = 0; i < nj; ++i) {
        rm_frame *f = hj[i].frame;
        f->stamp_scene = hj[i].scene;
        f->stamp_version = hj[i].scene->version;
        f->stamp_args = *hj[i].args;
        f->stamped = true;
        f->dof_focus = hj[i].args->focus;
        f->dof_coc = hj[i].args->coc;
        for (int k = 0; k < 3; ++k) f->dof_cam[k] = hj[i].args->position[k];
        f->rec_valid = J[i].F.rec != nullptr;
    }
    return RM_OK;
}

}  // namespace rm

using namespace rm;

// Integrator work buffers (path queues, accumulator sets, shadow queues) of
// single renders: one per (device, stream), reused by every render queued on
// that stream, whose order makes the reuse safe; kept until the process ends
// (render_jobs gives memory back when it stays oversized).  Renders on
// different streams use different buffers and do not wait for each other.
struct StreamWork {
    std::mutex mu;
    DevBuf buf;
};
static int32_t stream_work(hipStream_t st, StreamWork **out) {
    int dev = 0;
    RM_HIP(hipGetDevice(&dev));
    static std::mutex mu;
    static auto *table = new std::map<std::pair<int, uintptr_t>, StreamWork *>();
    std::lock_guard<std::mutex> lock(mu);
    StreamWork *&w = (*table)[{dev, (uintptr_t)st}];
    if (!w) w = new StreamWork();
    *out = w;
    return RM_OK;
}

extern "C" int32_t rm_render(const rm_scene *scene, const rm_render_args *a, rm_frame *f, void *stream) {
    if (!f) return fail(RM_ERR_INVALID, "rm_render: frame is NULL");
    std::vector<HostJob> hj{HostJob{scene, a, f, 0}};
    StreamWork *w = nullptr;
    int32_t rc = stream_work((hipStream_t)stream, &w);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(w->mu);
    return render_jobs(hj, w->buf, (hipStream_t)stream, false);
}

extern "C" int32_t rm_render_batch(const rm_batch_job *jobs, int32_t n, int32_t use_graph, void *stream) {
    if (n < 0 || (n > 0 && !jobs)) return fail(RM_ERR_INVALID, "rm_render_batch: bad job list");
    if (n == 0) return RM_OK;
    std::vector<HostJob> hj(n);
    for (int i = 0; i < n; ++i) {
        if (!jobs[i].frame) return fail(RM_ERR_INVALID, "rm_render_batch: job %d has no frame", i);
        for (int k = 0; k < i; ++k)
            if (jobs[k].frame == jobs[i].frame) return fail(RM_ERR_INVALID, "rm_render_batch: frame of job %d reused", i);
        hj[i] = HostJob{jobs[i].scene, &jobs[i].args, jobs[i].frame, jobs[i].post_options};
    }
    // a graph cannot be captured on the legacy default stream
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    if (use_graph && !st) {
        RM_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
        st = own;
    }
    // one workspace per device, kept between batches (a serving loop would
    // otherwise allocate gigabytes per call); batches on one device take turns
    // (never destroyed: freeing at exit could run after the HIP runtime's teardown)
    static std::mutex mu;
    static DevBuf *work = new DevBuf[64];
    std::lock_guard<std::mutex> lock(mu);
    int dev = 0;
    RM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(RM_ERR_UNSUPPORTED, "rm_render_batch: device %d", dev);
    int32_t rc = render_jobs(hj, work[dev], st, use_graph != 0);
    hipError_t e = hipStreamSynchronize(st);
    if (own) (void)hipStreamDestroy(own);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "rm_render_batch");
    return RM_OK;
}

