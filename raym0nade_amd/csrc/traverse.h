// gfx950 BVH traversal that reproduces the reference's recursive walk
// (BVH::dfs_rayHit, src/bvh.cpp:56-88) result for result.
//
// Reference order, restated iteratively:
//   visit(u interior): tL_k,tR_k = rayInBox(child k) on [t_min, t_max(now)];
//     first = child0 if tL0 < tL1 else child1 (ties -> child1);
//     visit(first)  if tL_first  < tR_first;
//     visit(second) if tL_second < tR_second && tL_second < t_max(after first).
//   visit(leaf): Moller-Trumbore on faces in order, accept t_min < t < t_max.
// The second child is pushed with tL_second when tL_second < tR_second holds and
// the t_max test is made when it is popped, after the whole first subtree: the
// same sequence of box tests, triangle tests and t_max updates as the recursion.
//
// Exactness: every operation is the reference's fp32 operation in glm order
// (see rm_vec.h), built with -ffp-contract=off and IEEE division/sqrt.  Three
// values the reference recomputes per test are precomputed per ray or per
// triangle with the identical operations (1/d_i, v1-v0, v2-v0), and
// |a|/length(e1) < 1e-4 is replaced by the exactly equivalent |a| < thr.
#pragma once
#include <hip/hip_runtime.h>
#include "rm_layout.h"
#include "rm_vec.h"

namespace rm {

constexpr float T_EPS = 1e-4f;   // eps_zero

struct RayPre {
    V3 o, d;
    float inv[3];
    bool par[3];
};

__device__ __forceinline__ RayPre ray_pre(V3 o, V3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    r.par[0] = fabsf(d.x) < T_EPS;
    r.par[1] = fabsf(d.y) < T_EPS;
    r.par[2] = fabsf(d.z) < T_EPS;
    r.inv[0] = 1.0f / d.x;
    r.inv[1] = 1.0f / d.y;
    r.inv[2] = 1.0f / d.z;
    return r;
}

// rayInBox (src/geometry.cpp:40-61) on one slab.  Returns false when the
// reference would return early (tR=-1 or tL>tR), leaving tL/tR as it does.
__device__ __forceinline__ bool slab(float o, float inv, bool par, float lo, float hi, float &tL, float &tR) {
    if (par) {
        if (o < lo || o > hi) {
            tR = -1.0f;
            return false;
        }
        return true;
    }
    float a = (lo - o) * inv, b = (hi - o) * inv;
    float en = inv >= 0.0f ? a : b, ex = inv >= 0.0f ? b : a;
    // std::fmax / std::fmin; all operands are finite or +-inf here, where the
    // hardware max/min agree with them up to the sign of a zero, which no
    // later comparison can observe.
    tL = fmaxf(tL, en);
    tR = fminf(tR, ex);
    tR += T_EPS;
    return !(tL > tR);
}

__device__ __forceinline__ void box_test(const RayPre &r, float lx, float ly, float lz, float hx, float hy, float hz,
                                         float &tL, float &tR) {
    if (!slab(r.o.x, r.inv[0], r.par[0], lx, hx, tL, tR)) return;
    if (!slab(r.o.y, r.inv[1], r.par[1], ly, hy, tL, tR)) return;
    slab(r.o.z, r.inv[2], r.par[2], lz, hz, tL, tR);
}

// RayTriangleIntersection (src/geometry.cpp:63-87).
__device__ __forceinline__ float tri_test(const RayPre &r, const TriRec &T) {
    V3 v0{T.p[0], T.p[1], T.p[2]};
    V3 e1{T.e1[0], T.e1[1], T.e1[2]};
    V3 e2{T.e2[0], T.e2[1], T.e2[2]};
    V3 h = cross(r.d, e2);
    float a = dot(e1, h);
    if (fabsf(a) < T.p[3]) return INFINITY;
    float f = 1.0f / a;
    V3 s = r.o - v0;
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return INFINITY;
    V3 q = cross(s, e1);
    float v = f * dot(r.d, q);
    if (v < 0.0f || u + v > 1.0f) return INFINITY;
    return f * dot(e2, q);
}

__device__ __forceinline__ TriRec load_tri(const TriRec *tris, uint32_t i) {
    TriRec T;
    const float4 *p = reinterpret_cast<const float4 *>(tris + i);
    float4 x = p[0], y = p[1], z = p[2];
    T.p[0] = x.x; T.p[1] = x.y; T.p[2] = x.z; T.p[3] = x.w;
    T.e1[0] = y.x; T.e1[1] = y.y; T.e1[2] = y.z; T.e1[3] = y.w;
    T.e2[0] = z.x; T.e2[1] = z.y; T.e2[2] = z.z; T.e2[3] = z.w;
    return T;
}

struct TraceCounts {
    uint32_t interior, leaves, boxes, tris;
};

// Leaf test: faces [first, first+count) in order, strict acceptance
// (src/bvh.cpp:58-67).
template <bool COUNT>
__device__ __forceinline__ void leaf_test(const SceneView &S, const RayPre &r, uint32_t code, float tmin, float &tmax,
                                          int32_t &best, TraceCounts &c) {
    uint32_t first = code >> LEAF_SHIFT, cnt = code & LEAF_MASK;
    if (COUNT) { c.leaves++; c.tris += cnt; }
    for (uint32_t i = 0; i < cnt; ++i) {
        TriRec T = load_tri(S.tris, first + i);
        float t = tri_test(r, T);
        if (tmin < t && t < tmax) {
            tmax = t;
            best = (int32_t)(first + i);
        }
    }
}

// Closest hit on (tmin, tmax).  With ANY, returns as soon as tmax < aim
// (Model::rayHit_test's boolean is decided then).  Returns the slot of the
// closest accepted triangle (BVH order) or -1; tmax holds its t.
template <bool ANY, bool COUNT>
__device__ __forceinline__ int32_t trace(const SceneView &S, const RayPre &r, float tmin, float &tmax, float aim,
                                         TraceCounts &c) {
    int32_t best = -1;
    if (S.root_leaf) {
        leaf_test<COUNT>(S, r, (uint32_t)S.root_leaf, tmin, tmax, best, c);
        return best;
    }
    uint32_t stk_ref[MAX_STACK];
    float stk_tl[MAX_STACK];
    int sp = 0;
    uint32_t u = 1;
    while (true) {
        // ---- interior visit
        if (COUNT) { c.interior++; c.boxes += 2; }
        const float4 *P = reinterpret_cast<const float4 *>(S.nodes + u);
        float4 A = P[0], Bq = P[1], C = P[2];
        uint2 Lf = *reinterpret_cast<const uint2 *>(&P[3]);
        float tL0 = tmin, tR0 = tmax, tL1 = tmin, tR1 = tmax;
        box_test(r, A.x, A.y, A.z, A.w, Bq.x, Bq.y, tL0, tR0);
        box_test(r, Bq.z, Bq.w, C.x, C.y, C.z, C.w, tL1, tR1);
        uint32_t ref0 = Lf.x ? (Lf.x | 0x80000000u) : (u << 1);
        uint32_t ref1 = Lf.y ? (Lf.y | 0x80000000u) : (u << 1 | 1u);
        bool f0 = tL0 < tL1;
        float tLf = f0 ? tL0 : tL1, tRf = f0 ? tR0 : tR1;
        float tLs = f0 ? tL1 : tL0, tRs = f0 ? tR1 : tR0;
        uint32_t rf = f0 ? ref0 : ref1, rs = f0 ? ref1 : ref0;
        if (tLs < tRs) {
            stk_ref[sp] = rs;
            stk_tl[sp] = tLs;
            ++sp;
        }
        uint32_t next = (tLf < tRf) ? rf : 0u;
        // ---- resolve leaves and pops until the next interior node
        while (true) {
            if (next == 0u) {
                while (sp > 0) {
                    --sp;
                    if (stk_tl[sp] < tmax) {
                        next = stk_ref[sp];
                        break;
                    }
                }
                if (next == 0u) return best;
            }
            if (next & 0x80000000u) {
                if (COUNT) {}
                leaf_test<COUNT>(S, r, next & 0x7fffffffu, tmin, tmax, best, c);
                if (ANY && tmax < aim) return best;
                next = 0u;
                continue;
            }
            u = next;
            break;
        }
    }
}

// ---------------------------------------------------------------------------
// Step-wise form of the same walk: one call of step() processes exactly one
// node (an interior visit or a whole leaf), so lanes of a wave stay on the
// same code path, and a lane whose ray is finished can take a new ray.
//
// Branch-free slab test: for a non-empty box (lo <= hi) and a non-parallel
// axis, min((lo-o)*inv, (hi-o)*inv) is exactly the reference's entry distance
// and max(...) its exit distance (the one the `invD >= 0` test selects), so
// the three axes are evaluated without branches and the state the reference
// returns with (after the first axis where tL > tR, or after all three) is
// selected at the end.  Rays with a parallel axis (|d_i| < 1e-4) take the
// branching test.

__device__ __forceinline__ void box_test_fast(const RayPre &r, float lx, float ly, float lz, float hx, float hy,
                                              float hz, float &tL, float &tR) {
    float ax = (lx - r.o.x) * r.inv[0], bx = (hx - r.o.x) * r.inv[0];
    float ay = (ly - r.o.y) * r.inv[1], by = (hy - r.o.y) * r.inv[1];
    float az = (lz - r.o.z) * r.inv[2], bz = (hz - r.o.z) * r.inv[2];
    float L1 = fmaxf(tL, fminf(ax, bx)), R1 = fminf(tR, fmaxf(ax, bx)) + T_EPS;
    float L2 = fmaxf(L1, fminf(ay, by)), R2 = fminf(R1, fmaxf(ay, by)) + T_EPS;
    float L3 = fmaxf(L2, fminf(az, bz)), R3 = fminf(R2, fmaxf(az, bz)) + T_EPS;
    bool s1 = L1 > R1, s2 = L2 > R2;
    tL = s1 ? L1 : (s2 ? L2 : L3);
    tR = s1 ? R1 : (s2 ? R2 : R3);
}

// Storage of a walker's pending entries.  PrivateStack: per-lane arrays, which
// the compiler places in scratch (backed by HBM through the caches).
// LdsStack: entries [0, nl) in the block's dynamic LDS (blocks of WALK_BLOCK
// lanes), entry i of lane t holding its node at word [2 i WALK_BLOCK + t] and
// its entry distance at [(2 i + 1) WALK_BLOCK + t], so a wave's accesses to one
// entry fall in distinct banks; entries from nl on (deeper than the LDS share)
// in per-lane scratch.  nl * WALK_BLOCK * 8 bytes of LDS; nl >= the tree depth
// keeps every entry in LDS.
constexpr int WALK_BLOCK = 256;

// Per-lane scratch entries of a walk.  Kept apart from the walker's scalar
// state: a struct holding a dynamically indexed array stays in scratch as a
// whole, so the walker only points at this one.
struct ScratchEntries {
    uint32_t ref[MAX_STACK];
    float tl[MAX_STACK];
};

struct PrivateStack {
    ScratchEntries *e;
    __device__ __forceinline__ void bind(ScratchEntries *m) { e = m; }
    __device__ __forceinline__ void put(int i, uint32_t r, float t) { e->ref[i] = r; e->tl[i] = t; }
    __device__ __forceinline__ uint32_t ref_at(int i) const { return e->ref[i]; }
    __device__ __forceinline__ float tl_at(int i) const { return e->tl[i]; }
};
struct LdsStack {
    uint32_t *col;     // this lane's column
    int nl;            // entries held in LDS
    ScratchEntries *e; // the deeper ones
    // Binds the launching kernel's dynamic LDS as this lane's first n entries.
    __device__ __forceinline__ void bind(ScratchEntries *m, int n) {
        extern __shared__ uint32_t walk_lds[];
        col = walk_lds + threadIdx.x;
        nl = n;
        e = m;
    }
    __device__ __forceinline__ void put(int i, uint32_t r, float t) {
        if (i < nl) {
            col[i * (2 * WALK_BLOCK)] = r;
            col[i * (2 * WALK_BLOCK) + WALK_BLOCK] = __float_as_uint(t);
        } else {
            e->ref[i] = r;
            e->tl[i] = t;
        }
    }
    // The LDS reads are relaxed atomic loads only so that the compiler keeps
    // them apart from the scratch reads (merged, both become flat loads).
    __device__ __forceinline__ uint32_t lds_at(int w) const {
        return __hip_atomic_load(col + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ uint32_t ref_at(int i) const { return i < nl ? lds_at(i * (2 * WALK_BLOCK)) : e->ref[i]; }
    __device__ __forceinline__ float tl_at(int i) const {
        return i < nl ? __uint_as_float(lds_at(i * (2 * WALK_BLOCK) + WALK_BLOCK)) : e->tl[i];
    }
};

// TPS: triangles a step tests (0: a whole leaf).  Splitting a leaf's tests
// over steps keeps their order, so the result is the same; lanes of a wave
// that stand at interior nodes then wait one triangle, not a whole leaf.
template <bool ANY, bool COUNT, class Stack = PrivateStack, int TPS = 0>
struct Walker {
    Stack stk;
    int sp;
    uint32_t cur;  // interior node index, leaf ref (bit 31), or 0 when finished
    int32_t best;
    float tmax;
    bool anypar;
    uint32_t li;   // TPS > 0: next triangle of the current leaf

    __device__ __forceinline__ void init(const SceneView &S, const RayPre &r, float tmax0) {
        sp = 0;
        best = -1;
        tmax = tmax0;
        cur = S.root_leaf ? ((uint32_t)S.root_leaf | 0x80000000u) : 1u;
        anypar = r.par[0] || r.par[1] || r.par[2];
        li = 0;
    }

    // The record of the interior node about to be visited (nothing for a leaf).
    struct Node {
        float4 a, b, c;
        uint2 leaf;
    };
    __device__ __forceinline__ Node fetch(const SceneView &S) const {
        Node n;
        if (!(cur & 0x80000000u)) {
            const float4 *P = reinterpret_cast<const float4 *>(S.nodes + cur);
            n.a = P[0];
            n.b = P[1];
            n.c = P[2];
            n.leaf = *reinterpret_cast<const uint2 *>(&P[3]);
        }
        return n;
    }

    // Returns true when the walk has finished (result in best / tmax); see trace_walk.
    __device__ __forceinline__ bool step(const SceneView &S, const RayPre &r, float tmin, float aim, TraceCounts &c) {
        return step_with(S, r, tmin, aim, c, fetch(S));
    }

    // step() with the current interior node's record already loaded.
    __device__ __forceinline__ bool step_with(const SceneView &S, const RayPre &r, float tmin, float aim, TraceCounts &c,
                                              const Node &nd) {
        if (!(cur & 0x80000000u)) {
            if (COUNT) { c.interior++; c.boxes += 2; }
            const uint32_t u = cur;
            const float4 A = nd.a, Bq = nd.b, C = nd.c;
            const uint2 Lf = nd.leaf;
            float tL0 = tmin, tR0 = tmax, tL1 = tmin, tR1 = tmax;
            if (!anypar) {
                box_test_fast(r, A.x, A.y, A.z, A.w, Bq.x, Bq.y, tL0, tR0);
                box_test_fast(r, Bq.z, Bq.w, C.x, C.y, C.z, C.w, tL1, tR1);
            } else {
                box_test(r, A.x, A.y, A.z, A.w, Bq.x, Bq.y, tL0, tR0);
                box_test(r, Bq.z, Bq.w, C.x, C.y, C.z, C.w, tL1, tR1);
            }
            uint32_t ref0 = Lf.x ? (Lf.x | 0x80000000u) : (u << 1);
            uint32_t ref1 = Lf.y ? (Lf.y | 0x80000000u) : (u << 1 | 1u);
            bool f0 = tL0 < tL1;
            float tLf = f0 ? tL0 : tL1, tRf = f0 ? tR0 : tR1;
            float tLs = f0 ? tL1 : tL0, tRs = f0 ? tR1 : tR0;
            uint32_t rf = f0 ? ref0 : ref1, rs = f0 ? ref1 : ref0;
            if (tLs < tRs) {
                stk.put(sp, rs, tLs);
                ++sp;
            }
            cur = (tLf < tRf) ? rf : 0u;
        } else if (TPS == 0) {
            leaf_test<COUNT>(S, r, cur & 0x7fffffffu, tmin, tmax, best, c);
            cur = 0u;
            if (ANY && tmax < aim) return true;
        } else {
            const uint32_t code = cur & 0x7fffffffu;
            const uint32_t first = code >> LEAF_SHIFT, cnt = code & LEAF_MASK;
            if (COUNT && li == 0) { c.leaves++; c.tris += cnt; }
            const uint32_t end = li + TPS < cnt ? li + TPS : cnt;
            for (uint32_t i = li; i < end; ++i) {
                const TriRec T = load_tri(S.tris, first + i);
                const float t = tri_test(r, T);
                if (tmin < t && t < tmax) {
                    tmax = t;
                    best = (int32_t)(first + i);
                }
            }
            li = end;
            if (ANY && tmax < aim) return true;
            if (li < cnt) return false;
            li = 0;
            cur = 0u;
        }
        if (cur == 0u) {
            while (sp > 0) {
                --sp;
                if (stk.tl_at(sp) < tmax) {
                    cur = stk.ref_at(sp);
                    break;
                }
            }
            if (cur == 0u) return true;
        }
        return false;
    }
};

// Whether a walking lane, standing at `cur` (a walker's current node: bit 31
// for a leaf), takes its step in this iteration of a persistent walk.  With
// phases p in 1..64 a wave takes one kind of step at a time: triangle tests
// when at least p/64 of its walking lanes stand at a leaf, else interior
// nodes; each ray's own sequence of steps is unchanged.  Called by every lane
// of the wave together; `walking` is false for lanes without a ray.
__device__ __forceinline__ bool walk_turn(uint32_t cur, int phases, bool walking) {
    if (!phases) return true;
    const bool leaf = walking && (cur & 0x80000000u);
    const uint32_t nlf = (uint32_t)__popcll(__ballot(leaf)), nact = (uint32_t)__popcll(__ballot(walking));
    return leaf == (64u * nlf >= (uint32_t)phases * nact);
}

// The whole walk for one ray with the step-wise walker: the same result as
// trace<ANY, false>, with far fewer registers, for kernels that do more than
// traversal.  The pending entries live in scratch, or their first S.walk_lds
// in the kernel's dynamic LDS (the launch then allocates S.walk_lds *
// WALK_BLOCK * 8 bytes and runs blocks of WALK_BLOCK lanes).
template <bool ANY, class Stack, int TPS>
__device__ __forceinline__ int32_t walk_with(Walker<ANY, false, Stack, TPS> &w, const SceneView &S, const RayPre &r,
                                             float tmin, float &tmax, float aim) {
    w.init(S, r, tmax);
    TraceCounts c{};
    while (!w.step(S, r, tmin, aim, c)) {
    }
    tmax = w.tmax;
    return w.best;
}

// LDSW = false compiles the scratch walk only (kernels that never get LDS
// stacks); TPS as for Walker.
template <bool ANY, bool LDSW = true, int TPS = 0>
__device__ __forceinline__ int32_t trace_walk(const SceneView &S, const RayPre &r, float tmin, float &tmax, float aim) {
    if constexpr (LDSW) {
        if (S.walk_lds) {
            ScratchEntries deep;
            Walker<ANY, false, LdsStack, TPS> w;
            w.stk.bind(&deep, S.walk_lds);
            return walk_with(w, S, r, tmin, tmax, aim);
        }
    }
    ScratchEntries ent;
    Walker<ANY, false, PrivateStack, TPS> w;
    w.stk.bind(&ent);
    return walk_with(w, S, r, tmin, tmax, aim);
}

}  // namespace rm
