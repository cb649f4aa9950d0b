// mfg_trace.hip -- HIP kernels of the MI355X test-particle tracer (gfx950).
//
// k_trace: one persistent launch traces a whole batch of desorptions. Each
// lane owns one particle in registers; dead lanes refill from a global work
// queue with one atomic per wave (ballot + mbcnt), first from the particles
// the previous launch parked. Per event a lane walks an 8-wide BVH of 128 B
// nodes with 8-bit quantised FP32 boxes (the top of the tree copied to LDS
// per block, traversal stack in LDS laid out [entry][lane]); leaf children
// go on the stack and the wave tests them together once enough lanes hold
// one; facets are tested in FP64. It then applies the event of reference
// src/Simulation/Particle.cpp:207-396 (desorption, bounce, absorption,
// transparent pass, teleport, link, leak, volume events). Tallies are u64 /
// FP64 atomics: lanes of a wave that update the same facet counter (or
// desorb from the same facet) add through one set of atomics, the counters
// of a few hot facets are summed per block in LDS, and the global counters
// are kept per lane and flushed once per wave at exit.
//
// Random numbers: Philox4x32-10 keyed by the variant seed, counter
// (particle index, block, 0); a lane draws them in the same order as the
// CPU restatement, so seeded runs are comparable event by event.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mfg_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInv2p32 = 2.3283064365386962890625e-10;

// ------------------------------------------------------------------ RNG
__device__ __forceinline__ void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                         uint32_t k1, uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t& o3) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

__device__ __forceinline__ double u2d(uint32_t x) { return ((double)x + 0.5) * kInv2p32; }

struct Rng {
    uint32_t k0, k1, p0, p1, blk, b0, b1, b2, b3, avail;
    __device__ void init(uint64_t seed, uint64_t pid) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        p0 = (uint32_t)pid; p1 = (uint32_t)(pid >> 32);
        blk = 0; avail = 0;
    }
    __device__ __forceinline__ double rnd() {
        if (avail == 0) {
            philox10(p0, p1, blk, 0u, k0, k1, b0, b1, b2, b3);
            ++blk;
            avail = 4;
        }
        uint32_t r = b0;
        b0 = b1; b1 = b2; b2 = b3;
        --avail;
        return u2d(r);
    }
    __device__ double coin(uint32_t event, uint32_t facet) const {
        uint32_t o0, o1, o2, o3;
        philox10(p0, p1, event, 0x80000000u | facet, k0, k1, o0, o1, o2, o3);
        return u2d(o0);
    }
};

// ------------------------------------------------------------------ math helpers
struct d3 { double x, y, z; };
__device__ __forceinline__ d3 mk(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 add(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 scl(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double fdot(d3 a, d3 b) { return fma(a.x, b.x, fma(a.y, b.y, a.z * b.z)); }
__device__ __forceinline__ d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double nrm(d3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ d3 unitv(d3 a) { double n = nrm(a); return n > 0 ? scl(1.0 / n, a) : a; }
__device__ __forceinline__ d3 ld3(const double* p) { return {p[0], p[1], p[2]}; }

__device__ __forceinline__ void atomic_addd(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_addu(void* p, unsigned long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), v);
}

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ------------------------------------------------------------------ tables
struct XYRef { const double* x; const double* y; uint32_t n; };

__device__ double interp_y(double x, XYRef t, bool logx, bool logy, bool extrapolate) {
    uint32_t n = t.n;
    if (n == 0) return 0.0;
    if (n == 1) return t.y[0];
    uint32_t i;
    if (x <= t.x[0]) {
        if (!extrapolate) return t.y[0];
        i = 0;
    } else if (x >= t.x[n - 1]) {
        if (!extrapolate) return t.y[n - 1];
        i = n - 2;
    } else {
        uint32_t lo = 0, hi = n - 1;
        while (hi - lo > 1) { uint32_t mid = (lo + hi) / 2; if (t.x[mid] <= x) lo = mid; else hi = mid; }
        i = lo;
    }
    double x0 = t.x[i], x1 = t.x[i + 1], y0 = t.y[i], y1 = t.y[i + 1];
    if (logx) { x0 = log10(x0); x1 = log10(x1); x = log10(x); }
    if (logy) { y0 = log10(y0); y1 = log10(y1); }
    double y = (x1 == x0) ? y0 : y0 + (x - x0) * (y1 - y0) / (x1 - x0);
    return logy ? pow(10.0, y) : y;
}

// interp_x with a guide table: guide[k] is the last point whose y is at or
// below k/K of the table's last y, so a lookup starts one or two points
// before the answer; the interval found is the binary search's.
__device__ double interp_x_guided(double y, XYRef t, const uint32_t* guide, uint32_t K) {
    uint32_t n = t.n;
    uint32_t i;
    double ylast = t.y[n - 1];
    if (y >= ylast) {
        i = n - 2;
    } else {
        uint32_t k = (uint32_t)fmax(0.0, fmin((double)(K - 1), y / ylast * (double)K));
        i = guide[k];
        while (i > 0 && t.y[i] > y) --i;
        while (i + 1 < n - 1 && t.y[i + 1] <= y) ++i;
    }
    double y0 = t.y[i], y1 = t.y[i + 1], x0 = t.x[i], x1 = t.x[i + 1];
    if (y1 == y0) return x0;
    return x0 + (y - y0) * (x1 - x0) / (y1 - y0);
}

__device__ double interp_x(double y, XYRef t) {
    uint32_t n = t.n;
    if (n == 1) return t.x[0];
    uint32_t i;
    if (y < t.y[0]) i = 0;
    else if (y >= t.y[n - 1]) i = n - 2;
    else {
        uint32_t lo = 0, hi = n - 1;
        while (hi - lo > 1) { uint32_t mid = (lo + hi) / 2; if (t.y[mid] <= y) lo = mid; else hi = mid; }
        i = lo;
    }
    double y0 = t.y[i], y1 = t.y[i + 1], x0 = t.x[i], x1 = t.x[i + 1];
    if (y1 == y0) return x0;
    return x0 + (y - y0) * (x1 - x0) / (y1 - y0);
}

// ------------------------------------------------------------------ geometry
__device__ __forceinline__ bool in_poly(const double* __restrict__ p, uint32_t n, double u, double v) {
    bool inside = false;
    double uj = p[2 * (n - 1)], vj = p[2 * (n - 1) + 1];
    for (uint32_t i = 0; i < n; ++i) {
        double ui = p[2 * i], vi = p[2 * i + 1];
        if ((vi > v) != (vj > v)) {
            double x = (uj - ui) * (v - vi) / (vj - vi) + ui;
            if (u < x) inside = !inside;
        }
        uj = ui; vj = vi;
    }
    return inside;
}

// in_poly for more than 4 vertices through the slab index stored after the
// (u,v) list: the same crossing predicate on the edges that straddle v
// (exactly those the loop above toggles on), found by a binary search over
// the sorted vertex heights.
__device__ __forceinline__ uint32_t low_bits(double d) { return (uint32_t)__double_as_longlong(d); }
__device__ bool in_poly_slabs(const double* __restrict__ p, uint32_t n, double u, double v) {
    const double* S = p + 2 * (size_t)n;
    const uint32_t m = low_bits(S[0]);
    const double* w = S + 1;
    if (!(v >= w[0] && v < w[m - 1])) return false;   // no straddling edge (NaN included)
    uint32_t lo = 0, hi = m - 1;                       // w[lo] <= v < w[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (w[mid] <= v) lo = mid; else hi = mid;
    }
    const uint32_t e0 = low_bits(S[1 + m + lo]), e1 = low_bits(S[2 + m + lo]);
    const double2* E = reinterpret_cast<const double2*>(S + 2 + 2 * (size_t)m);
    bool inside = false;
    for (uint32_t e = e0; e < e1; ++e) {
        const double2 a = E[2 * e], b = E[2 * e + 1];   // (u_i, v_i), (u_j, v_j)
        const double x = (b.x - a.x) * (v - a.y) / (b.y - a.y) + a.x;
        if (u < x) inside = !inside;
    }
    return inside;
}
__device__ __forceinline__ bool in_polygon(const double* __restrict__ p, uint32_t n, double u, double v) {
    return n > 4 ? in_poly_slabs(p, n, u, v) : in_poly(p, n, u, v);
}

// in_poly for a triangle held in registers: the same crossing tests in the
// same order as the loop above with n = 3.
__device__ __forceinline__ bool crossing(double ui, double vi, double uj, double vj, double u, double v) {
    if ((vi > v) != (vj > v)) {
        double x = (uj - ui) * (v - vi) / (vj - vi) + ui;
        return u < x;
    }
    return false;
}
// in_poly for a quadrilateral held in registers (n = 4, same order)
__device__ __forceinline__ bool in_quad(const LeafUV* __restrict__ LU, double u, double v) {
    const double2* q = reinterpret_cast<const double2*>(LU->uv);
    const double2 p0 = q[0], p1 = q[1], p2 = q[2], p3 = q[3];
    bool inside = false;
    if (crossing(p0.x, p0.y, p3.x, p3.y, u, v)) inside = !inside;
    if (crossing(p1.x, p1.y, p0.x, p0.y, u, v)) inside = !inside;
    if (crossing(p2.x, p2.y, p1.x, p1.y, u, v)) inside = !inside;
    if (crossing(p3.x, p3.y, p2.x, p2.y, u, v)) inside = !inside;
    return inside;
}
__device__ __forceinline__ bool in_tri(const LeafUV* __restrict__ LU, double u, double v) {
    const double2* q = reinterpret_cast<const double2*>(LU->uv);
    const double2 p0 = q[0], p1 = q[1], p2 = q[2];
    bool inside = false;
    if (crossing(p0.x, p0.y, p2.x, p2.y, u, v)) inside = !inside;
    if (crossing(p1.x, p1.y, p0.x, p0.y, u, v)) inside = !inside;
    if (crossing(p2.x, p2.y, p1.x, p1.y, u, v)) inside = !inside;
    return inside;
}

// Ray/facet test; the same arithmetic, operation for operation, as the CPU
// restatement (explicit fma dot products, contraction off elsewhere).
__device__ __forceinline__ bool facet_test(const DFacet* __restrict__ F, const double* __restrict__ poly, d3 P, d3 D,
                                           double& t, double& u, double& v) {
    d3 N = ld3(F->N);
    double denom = fdot(N, D);
    uint32_t flags = F->flags;
    if (denom == 0.0) return false;
    if (!(flags & FF_2SIDED) && denom > 0.0) return false;
    d3 O = ld3(F->O);
    d3 w = sub(O, P);
    t = fdot(N, w) / denom;
    if (!(t > 0.0)) return false;
    d3 X = mk(fma(t, D.x, P.x), fma(t, D.y, P.y), fma(t, D.z, P.z));
    d3 r = sub(X, O);
    u = fdot(r, ld3(F->U)) * F->invU2;
    v = fdot(r, ld3(F->V)) * F->invV2;
    if (u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0) return false;
    return in_polygon(poly + 2 * (size_t)F->poly_off, F->poly_n, u, v);
}

// facet_test on a leaf record: the same operations in the same order. The
// record's 128 B are loaded together before the first decision (one round
// trip); the (u,v) polygon of a triangle or quadrilateral is fetched only
// when the ray meets the facet's rectangle.
__device__ __forceinline__ bool leaf_test(const LeafFacet* __restrict__ L, const LeafUV* __restrict__ LU,
                                          const double* __restrict__ poly, d3 P, d3 D, int32_t skip, double& t,
                                          double& u, double& v, int32_t& id, uint32_t& flags) {
    const uint4 hd = reinterpret_cast<const uint4*>(L)[0];
    const double2* q = reinterpret_cast<const double2*>(L);
    const double2 q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6], q7 = q[7];
    id = (int32_t)hd.x;
    flags = hd.y;
    const d3 O = mk(q1.x, q1.y, q2.x), N = mk(q2.y, q3.x, q3.y);
    const d3 U = mk(q4.x, q4.y, q5.x), V = mk(q5.y, q6.x, q6.y);
    const double iu = q7.x, iv = q7.y;
    double denom = fdot(N, D);
    d3 w = sub(O, P);
    t = fdot(N, w) / denom;
    d3 X = mk(fma(t, D.x, P.x), fma(t, D.y, P.y), fma(t, D.z, P.z));
    d3 r = sub(X, O);
    u = fdot(r, U) * iu;
    v = fdot(r, V) * iv;
    const uint32_t n = hd.w;
    // non-short-circuit: every operand (the whole record) is needed here
    bool ok = (denom != 0.0) & (((flags & FF_2SIDED) != 0u) | !(denom > 0.0)) & (t > 0.0) &
              !((u < 0.0) | (u > 1.0) | (v < 0.0) | (v > 1.0)) & (id != skip) & (n >= 3u) & (hd.z != ~0u);
    if (!ok) return false;
    if (n == 3) return in_tri(LU, u, v);
    if (n == 4) return in_quad(LU, u, v);
    return in_poly_slabs(poly + 2 * (size_t)hd.z, n, u, v);
}

struct Ray32 { float ox, oy, oz, ix, iy, iz; };

__device__ __forceinline__ float safe_inv(double d) {
    float f = (float)d;
    if (fabsf(f) < 1e-30f) f = (f < 0.0f) ? -1e-30f : 1e-30f;
    return 1.0f / f;
}

__device__ __forceinline__ Ray32 make_ray32(d3 P, d3 D) {
    return {(float)P.x, (float)P.y, (float)P.z, safe_inv(D.x), safe_inv(D.y), safe_inv(D.z)};
}

// A node's 128 B as eight 16-byte loads issued together (one round trip).
// Words: c0 = origin x,y,z + (ex, ey, ez, leaf mask) bytes; c1 = node_base,
// leaf_base, leaf counts, inner mask (byte 0); c2..c4 = low planes of x, y,
// z and c5..c7 = high planes, two FP16 steps per word (slot 2i low half,
// 2i + 1 high half).
struct NodeRegs {
    uint4 c0, c1, c2, c3, c4, c5, c6, c7;
};

__device__ __forceinline__ NodeRegs load_node(const BvhNode* __restrict__ nd) {
    const uint4* q = reinterpret_cast<const uint4*>(nd);
    NodeRegs n;
    n.c0 = q[0]; n.c1 = q[1]; n.c2 = q[2]; n.c3 = q[3]; n.c4 = q[4]; n.c5 = q[5]; n.c6 = q[6]; n.c7 = q[7];
    return n;
}

// A node from the block's LDS copy of the top of the tree, else from HBM.
__device__ __forceinline__ NodeRegs load_node_at(const BvhNode* __restrict__ nodes, const uint4* lnodes,
                                                 uint32_t n_lds, int32_t node) {
    if ((uint32_t)node < n_lds) {
        const uint4* q = lnodes + (size_t)node * 8;
        NodeRegs n;
        n.c0 = q[0]; n.c1 = q[1]; n.c2 = q[2]; n.c3 = q[3]; n.c4 = q[4]; n.c5 = q[5]; n.c6 = q[6]; n.c7 = q[7];
        return n;
    }
    return load_node(nodes + node);
}

__device__ __forceinline__ uint32_t node_leaves(const NodeRegs& n) { return n.c0.w >> 24; }
__device__ __forceinline__ uint32_t node_inner(const NodeRegs& n) { return n.c1.w & 0xffu; }
// inner child of slot c: the node_base-th node plus the inner slots before c
__device__ __forceinline__ int32_t node_inner_child(const NodeRegs& n, uint32_t c) {
    return (int32_t)n.c1.x + (int32_t)__popc(node_inner(n) & ((1u << c) - 1u));
}
// leaf child of slot c: (first record, facets); its records follow those of
// the leaf slots before c, each holding 1 + its 4-bit count facets
__device__ __forceinline__ uint32_t nibble_sum(uint32_t v) {
    const uint32_t t = (v & 0x0f0f0f0fu) + ((v >> 4) & 0x0f0f0f0fu);
    return (t * 0x01010101u) >> 24;
}
__device__ __forceinline__ void node_leaf_child(const NodeRegs& n, uint32_t c, uint32_t& first, uint32_t& cnt) {
    const uint32_t below = (1u << c) - 1u;
    const uint32_t cw = n.c1.z;
    first = n.c1.y + (uint32_t)__popc(node_leaves(n) & below) + nibble_sum(cw & ((1u << (4u * c)) - 1u));
    cnt = ((cw >> (4u * c)) & 15u) + 1u;
}
// 2^e as a float (|e| < 127)
__device__ __forceinline__ float exp2i(int e) { return __int_as_float((e + 127) << 23); }
// the FP16 step of slot 2i (lo) or 2i + 1 (hi) of a plane word, as a float
// (converted inside the v_fma_mix_f32 that uses it)
__device__ __forceinline__ float step_lo(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu)); }
__device__ __forceinline__ float step_hi(uint32_t w) { return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16)); }
__device__ __forceinline__ uint32_t word_of(const uint4& v, uint32_t i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// Slab test of the node's children against [0, tmax]: the hit mask.
// A bound q (an integer step) on an axis sits at o + q 2^e, so its ray
// parameter is (o + q 2^e - ray_o) / d = fma(q, 2^e/d, (o - ray_o)/d): two
// per-node constants per axis, then one fma per bound (q held as an FP16
// value, so conversion and fma are one v_fma_mix_f32 with the same single
// rounding). The quantised boxes enclose the FP64 boxes padded by 1e-5 of
// the model extent, far above the FP32 rounding of these steps, so the
// test never drops a box that holds an FP64 hit: a ray through a point of
// the unpadded box spends at least 1e-5 extent / |d| of its parameter
// inside the padded one (within [0, tmax] as well, since tmax is at or
// beyond that hit), against FP32 errors near 1e-7 extent / |d|, so no
// tolerance is added to a <= b.
__device__ __forceinline__ uint32_t node_hits(const NodeRegs& n, const Ray32& r, float tmax) {
    const float ox = __uint_as_float(n.c0.x), oy = __uint_as_float(n.c0.y), oz = __uint_as_float(n.c0.z);
    const uint32_t ew = n.c0.w;
    const float ax = exp2i((int)(int8_t)(ew & 0xffu)) * r.ix, ay = exp2i((int)(int8_t)((ew >> 8) & 0xffu)) * r.iy;
    const float az = exp2i((int)(int8_t)((ew >> 16) & 0xffu)) * r.iz;
    const float bx = (ox - r.ox) * r.ix, by = (oy - r.oy) * r.iy, bz = (oz - r.oz) * r.iz;
    // fma(q, a, b) is monotonic in q, rising when the axis' 1/d is >= 0: the
    // low bound is then the entry plane and the high one the exit plane, and
    // the reverse otherwise (an empty slot, low 255 > high 0, still misses)
    const bool px = r.ix >= 0.0f, py = r.iy >= 0.0f, pz = r.iz >= 0.0f;
    const uint4 nx = px ? n.c2 : n.c5, fx = px ? n.c5 : n.c2;
    const uint4 ny = py ? n.c3 : n.c6, fy = py ? n.c6 : n.c3;
    const uint4 nz = pz ? n.c4 : n.c7, fz = pz ? n.c7 : n.c4;
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) {
        const uint32_t i = c >> 1;
        const uint32_t wnx = word_of(nx, i), wfx = word_of(fx, i), wny = word_of(ny, i);
        const uint32_t wfy = word_of(fy, i), wnz = word_of(nz, i), wfz = word_of(fz, i);
        const bool hi = (c & 1u) != 0;
        const float tnx = fmaf(hi ? step_hi(wnx) : step_lo(wnx), ax, bx), tfx = fmaf(hi ? step_hi(wfx) : step_lo(wfx), ax, bx);
        const float tny = fmaf(hi ? step_hi(wny) : step_lo(wny), ay, by), tfy = fmaf(hi ? step_hi(wfy) : step_lo(wfy), ay, by);
        const float tnz = fmaf(hi ? step_hi(wnz) : step_lo(wnz), az, bz), tfz = fmaf(hi ? step_hi(wfz) : step_lo(wfz), az, bz);
        float a = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
        float b = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
        if (a <= b) mask |= 1u << c;
    }
    return mask & (node_leaves(n) | node_inner(n));
}

__device__ __forceinline__ float widen(double t) {
    if (t > 3.0e38) return 3.0e38f;
    return (float)t * 1.000004f + 1e-30f;
}

struct HitRec { int32_t facet; double t, u, v; };

// ---------------------------------------------------------------- DES_ANGLEMAP
// Sampling of a recorded incident-angle map (AnglemapGeneration.cpp:15-410),
// on the table block described in TraceParams::am_src. lower_index is the
// index of the last CDF value <= key (-1 below the first); weigh(a, b, w) =
// a + (b - a) * w.
struct AmSrc {
    int W, L, H;
    double limit, ratio, total;
    const double *line, *cdf, *pc, *pp;
};
__device__ __forceinline__ AmSrc am_view(const double* b) {
    AmSrc A;
    A.W = (int)b[0]; A.L = (int)b[1]; A.H = (int)b[2];
    A.limit = b[3]; A.ratio = b[4]; A.total = b[5];
    const int T = A.L + A.H;
    A.line = b + 6; A.cdf = A.line + T; A.pc = A.cdf + T; A.pp = A.pc + (size_t)T * A.W;
    return A;
}
__device__ __forceinline__ double am_weigh(double a, double b, double w) { return a + (b - a) * w; }
__device__ int am_lower_index(double key, const double* A, int n) {
    int lo = 0, hi = n;   // first element > key
    while (lo < hi) { int mid = (lo + hi) / 2; if (A[mid] <= key) lo = mid + 1; else hi = mid; }
    return lo - 1;
}
__device__ int am_weighed_lower_index(double key, double w, const double* A, const double* B, int n) {
    int lo = 0, hi = n;
    while (lo < hi) { int mid = (lo + hi) / 2; if (am_weigh(A[mid], B[mid], w) <= key) lo = mid + 1; else hi = mid; }
    return lo - 1;
}
__device__ __forceinline__ double am_theta(const AmSrc& A, double ti) {
    if ((size_t)ti < (size_t)A.L) return A.limit * ti / (double)A.L;
    return A.limit + (kPi / 2.0 - A.limit) * (ti - (double)A.L) / (double)A.H;
}
__device__ __forceinline__ double am_phi(const AmSrc& A, double pi_) {
    double w = (double)A.W;
    double c = (pi_ < w) ? pi_ : pi_ - w;
    return -kPi + 2.0 * kPi * c / w;
}
__device__ double am_phi_pdf(const AmSrc& A, double ti, int plo) {
    const int W = A.W, T = A.L + A.H;
    const size_t p = (size_t)(((plo % W) + W) % W);
    if (ti < 0.5) return A.pp[p];
    if (ti > (double)T - 0.5) return A.pp[(size_t)(T - 1) * W + p];
    size_t tl = (size_t)(ti - 0.5);
    double ov = ti - 0.5 - (double)tl;
    return am_weigh(A.pp[tl * W + p], A.pp[(tl + 1) * W + p], ov);
}
__device__ double am_phi_cdf(const AmSrc& A, double ti, int plo) {
    const int W = A.W, L = A.L, T = A.L + A.H;
    auto line = [&](size_t t) { return (plo < W) ? A.pc[t * W + plo] : 1.0 + A.pc[t * W]; };
    if (ti < (double)L) {
        if (ti < 0.5) return line(0);
        if (ti > (double)L - 0.5) return line((size_t)(L - 1));
    } else {
        if (ti < (double)L + 0.5) return line((size_t)L);
        if (ti > (double)T - 0.5) return line((size_t)(T - 1));
    }
    size_t tl = (size_t)(ti - 0.5);
    double ov = ti - 0.5 - (double)tl;
    return am_weigh(line(tl), line(tl + 1), ov);
}
__device__ void am_generate_theta(const AmSrc& A, double r, double& theta, int& tlo, double& ov) {
    const int L = A.L, H = A.H;
    const double* cdf = A.cdf;
    auto quad = [&](int t, double c) {
        double step = am_theta(A, (double)t + 1.5) - am_theta(A, (double)t + 0.5);
        double b = A.line[t] / A.total / step;
        double a = 0.5 * (A.line[t + 1] - A.line[t]) / A.total / (step * step);
        double dy = r - c;
        double dx = (-b + sqrt(b * b + 4 * a * dy)) / (2 * a);
        return dx / step;
    };
    if (r < A.ratio) {
        tlo = am_lower_index(r, cdf, L);
        if (tlo == -1) {
            ov = 0.5 * r / cdf[0];
            theta = am_theta(A, ov);
        } else if (tlo == L - 1) {
            ov = 0.5 * (r - cdf[tlo]) / (A.ratio - cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        } else {
            if (A.line[tlo] == A.line[tlo + 1]) ov = (r - cdf[tlo]) / (cdf[tlo + 1] - cdf[tlo]);
            else ov = quad(tlo, cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        }
    } else {
        tlo = L + am_lower_index(r, cdf + L, H);
        if (tlo == L - 1) {
            ov = 0.5 * (r - A.ratio) / (cdf[L] - A.ratio);
            theta = am_theta(A, (double)L + ov);
        } else if (tlo == L + H - 1) {
            ov = 0.5 * (r - cdf[tlo]) / (1.0 - cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        } else {
            if (A.line[tlo] == A.line[tlo + 1]) ov = (r - cdf[tlo]) / (cdf[tlo + 1] - cdf[tlo]);
            else ov = quad(tlo, cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        }
    }
}
__device__ double am_generate_phi(const AmSrc& A, int tlo, double ov, double r) {
    const int W = A.W, L = A.L, T = A.L + A.H;
    if (W == 1) return -kPi + 2.0 * kPi * r;
    int plo;
    double w;
    const double* C = A.pc;
    auto edge = [&](int t) {
        r += C[(size_t)t * W];
        plo = am_lower_index(r, C + (size_t)t * W, W);
        w = ov;
    };
    auto between = [&](int t) {
        double div = A.line[t] * (1.0 - ov) + A.line[t + 1] * ov;
        w = div > 0.0 ? (ov * A.line[t + 1]) / div : ov;
        r += am_weigh(C[(size_t)t * W], C[(size_t)(t + 1) * W], w);
        plo = am_weighed_lower_index(r, w, C + (size_t)t * W, C + (size_t)(t + 1) * W, W);
    };
    if (tlo < L) {
        if (tlo == -1) edge(0);
        else if (tlo == L - 1) edge(tlo);
        else between(tlo);
    } else {
        if (tlo == L - 1) edge(L);
        else if (tlo == T - 1) edge(tlo);
        else between(tlo);
    }
    const double ti = (double)tlo + 0.5 + w;
    double pov;
    if (plo == -1) {
        pov = 0.5 * r / am_phi_cdf(A, ti, 0);
        return am_phi(A, pov);
    }
    double p0 = am_phi_pdf(A, ti, plo), p1 = am_phi_pdf(A, ti, plo + 1);
    if (p0 == p1) {
        double c0 = am_phi_cdf(A, ti, plo);
        pov = (r - c0) / (am_phi_cdf(A, ti, plo + 1) - c0);
    } else {
        double step = 2.0 * kPi / (double)W;
        double c = am_phi_cdf(A, ti, plo);
        double b = p0 / step;
        double a = 0.5 * (p1 - p0) / (step * step);
        double dy = r - c;
        double D = b * b + 4 * a * dy;
        if (D < 0) pov = (r - c) / (am_phi_cdf(A, ti, ((plo + 1) % W + W) % W) - c);
        else pov = ((-b + sqrt(b * b + 4 * a * dy)) / (2 * a)) / step;
    }
    return am_phi(A, (double)plo + 0.5 + pov);
}

__device__ __forceinline__ double td_opacity(const TraceParams& P, const DVariant* var, int32_t f, double time) {
    const DPhys& ph = P.phys[f];
    const DTable tb = P.params[var->param_base + ph.opacity_param];
    return interp_y(time, XYRef{P.tab_x + tb.off, P.tab_y + tb.off, tb.n}, tb.log_x, tb.log_y, false);
}

// Traversal stack of one lane: the first `lds_cap` entries in LDS (laid out
// [entry][lane]), deeper entries (rare) in a per-lane global area.
struct LaneStack {
    int32_t* lds;
    uint32_t stride, lds_cap, cap;
    int32_t* spill;
    // entry sp of a lane sits sp * stride words after its first one (both
    // below 2^24: a full-rate 24-bit multiply)
    __device__ __forceinline__ void push(uint32_t& sp, int32_t v) const {
        if (sp < lds_cap) {
            lds[__umul24(sp, stride)] = v;
            ++sp;
        } else if (sp < cap) {
            spill[sp - lds_cap] = v;
            ++sp;
        }
    }
    __device__ __forceinline__ int32_t pop(uint32_t& sp) const {
        --sp;
        return sp < lds_cap ? lds[__umul24(sp, stride)] : spill[sp - lds_cap];
    }
};

// Nearest hard hit below tmax. With TRANSP, non-opaque facets are hard with
// probability `opacity` (coin per event and facet); opaque_only ignores
// opacity (fixed ray batches).
//
// Each lane holds one stack entry: a node (>= 0) or a leaf (< 0: ~(first
// record << 4 | facets - 1)). A node step tests the children's boxes and
// pushes the inner children (slot order), then the leaf children on top,
// so that leaves come off first and shrink tmax early. Each iteration the
// wave votes: lanes holding a leaf test its facets together once at least
// kLeafBatch of them hold one, or twice as many as hold a node, or when no
// lane holds a node; otherwise the node lanes step. Testing a node's leaves
// inside the node step instead ran the facet loop about 49 times per
// traversal of a beam-pipe wave with 7 of 64 lanes active; the vote cuts
// the wave's node plus facet iterations by about 29 % there (CPU emulation
// of both on the same tree) and leaves every hit unchanged: the nearest
// hit does not depend on the order facets are tested in (ties go to the
// lower facet id), and the opacity coin is a function of event and facet.
// A near-to-far sort of the children left the node visits per ray within
// 2 % on these trees but cost about a hundred instructions per node.
constexpr uint32_t kLeafBatch = 24;

__device__ __forceinline__ int32_t leaf_entry(uint32_t first, uint32_t cnt) {
    return ~(int32_t)((first << 4) | (cnt - 1u));
}

template <bool TRANSP>
__device__ HitRec traverse(const TraceParams& P, int32_t root, d3 Pp, d3 D, int32_t skip, double tmax,
                           const LaneStack& S, const Rng* rng,
                           uint32_t event, const DVariant* var, double time, bool opaque_only,
                           const uint4* lnodes) {
    HitRec h{-1, tmax, 0.0, 0.0};
    if (root < 0) return h;
    Ray32 r = make_ray32(Pp, D);
    float tbf = widen(tmax);
    uint32_t sp = 0;
    int32_t cur = root;
    bool live = true;
    for (;;) {
        const bool on_leaf = live && cur < 0, on_node = live && cur >= 0;
        const uint32_t nl = (uint32_t)__popcll(__ballot(on_node));
        const uint32_t ll = (uint32_t)__popcll(__ballot(on_leaf));
        if (nl + ll == 0) break;
        if (ll >= kLeafBatch || ll >= 2 * nl) {
            if (on_leaf) {
                const uint32_t e = ~(uint32_t)cur, first = e >> 4, cnt = (e & 15u) + 1u;
                for (uint32_t k = 0; k < cnt; ++k) {
                    double t, u, v;
                    int32_t f;
                    uint32_t fl;
                    bool hit = leaf_test(P.leaves + first + k, P.leaf_uv + first + k, P.poly, Pp, D, skip, t, u, v, f, fl);
                    if (!hit) continue;
                    if (!(t < h.t || (t == h.t && f < h.facet))) continue;
                    bool hard = true;
                    if constexpr (TRANSP) {
                        if (!opaque_only && !(fl & FF_OPAQUE)) {
                            double op = (fl & FF_TD_OPACITY) ? td_opacity(P, var, f, time) : (double)P.phys[f].opacity;
                            if (op < 1.0) hard = (op > 0.0) && rng->coin(event, (uint32_t)(f - (int32_t)var->facet_base)) < op;
                        }
                    }
                    if (hard) {
                        h.facet = f; h.t = t; h.u = u; h.v = v;
                        tbf = widen(t);
                    }
                }
                if (sp == 0) live = false;
                else cur = S.pop(sp);
            }
        } else if (on_node) {
            const NodeRegs nd = load_node_at(P.nodes, lnodes, P.lds_nodes, cur);
            const uint32_t m = node_hits(nd, r, tbf);   // children entering before tmax
            uint32_t im = m & node_inner(nd);
            uint32_t lm = m & node_leaves(nd);
            // the entry the stack would hand back first (the lowest leaf
            // slot, else the lowest inner slot) becomes the next one
            // directly, without its push and pop
            int32_t next = 0;
            const bool direct = (im | lm) != 0;
            if (lm) {
                const uint32_t c = (uint32_t)__builtin_ctz(lm);
                lm &= lm - 1u;
                uint32_t first, cnt;
                node_leaf_child(nd, c, first, cnt);
                next = leaf_entry(first, cnt);
            } else if (im) {
                const uint32_t c = (uint32_t)__builtin_ctz(im);
                im &= im - 1u;
                next = node_inner_child(nd, c);
            }
#pragma unroll
            for (int c = 7; c >= 0; --c)
                if (im & (1u << c)) S.push(sp, node_inner_child(nd, (uint32_t)c));
            while (lm) {
                const uint32_t c = 31u - (uint32_t)__builtin_clz(lm);
                lm &= ~(1u << c);
                uint32_t first, cnt;
                node_leaf_child(nd, c, first, cnt);
                S.push(sp, leaf_entry(first, cnt));
            }
            if (direct) cur = next;
            else if (sp == 0) live = false;
            else cur = S.pop(sp);
        }
    }
    return h;
}

// ------------------------------------------------------------------ tracer
template <uint32_t FEAT>
struct Tracer {
    const TraceParams& P;
    LaneStack S;
    // particle state (reference Particle.h:95-122)
    d3 pos, dir;
    double time, vel, ori, gen_time, dist, decay_t, scatter_path, hu, hv;
    int32_t structure, last_hit, tele_from;
    uint32_t nb_bounces, nb_events;
    uint64_t pid;
    Rng rng;
    const DVariant* var;
    // lane accumulators (flushed once per wave)
    uint32_t acc_hits, acc_des;
    double acc_hit_equiv, acc_abs_equiv, acc_dist_total, acc_dist_full;
    uint32_t acc_leak, acc_decay;
    uint32_t tr_n;
    uint32_t* adj_cnt;
    uint32_t seg_n;                   // adjoint: target scores since the last survived sticky collision
    int32_t* seg_t;
    double* seg_x;
    double* hot;                      // LDS table of the hot facets' slot-0 counters (8 words each)
    const uint4* lnodes;              // LDS copy of the first P.lds_nodes BVH nodes
    // pending slot-0 facet counter update of this lane
    int32_t pend_f;
    uint32_t pend_hits;
    double pend_he, pend_abs, pend_s1v, pend_svo, pend_s1pv;

    __device__ Tracer(const TraceParams& p, const LaneStack& s) : P(p), S(s) {
        acc_hits = acc_des = 0;
        acc_hit_equiv = acc_abs_equiv = acc_dist_total = acc_dist_full = 0.0;
        acc_leak = acc_decay = 0;
        tr_n = 0;
        adj_cnt = nullptr;
        seg_n = 0;
        seg_t = nullptr;
        seg_x = nullptr;
        var = nullptr;
        hot = nullptr;
        lnodes = nullptr;
        clear_pending();
    }

    __device__ int32_t local(int32_t f) const { return f - (int32_t)var->facet_base; }

    __device__ void trace(int32_t f, int32_t kind) {
        if constexpr ((FEAT & FEAT_TRACE) != 0) {
            if (pid < P.trace_n && tr_n < P.trace_k) {
                size_t at = (size_t)pid * P.trace_k + tr_n;
                P.trace_facet[at] = f < 0 ? f : local(f);
                P.trace_kind[at] = kind;
                ++tr_n;
            }
        }
        (void)f; (void)kind;
    }
    __device__ int moment(double t) const {
        if constexpr ((FEAT & FEAT_MOMENTS) != 0) {
            uint32_t n = var->n_moments;
            if (n == 0) return -1;
            const double* S = P.mom_start + var->mom_base;
            uint32_t lo = 0, hi = n;  // first start >= t
            while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (S[mid] < t) lo = mid + 1; else hi = mid; }
            uint32_t lb = lo;
            if (lb != 0) --lb;
            if (S[lb] <= t && t < P.mom_end[var->mom_base + lb]) return (int)lb + 1;
        }
        (void)t;
        return -1;
    }
    __device__ double param_at(int32_t id, double t, double c) const {
        if constexpr ((FEAT & FEAT_TDPARAM) != 0) {
            if (id >= 0) {
                const DTable tb = P.params[var->param_base + id];
                return interp_y(t, XYRef{P.tab_x + tb.off, P.tab_y + tb.off, tb.n}, tb.log_x, tb.log_y, false);
            }
        }
        (void)id; (void)t;
        return c;
    }
    __device__ double sticking(const DPhys& ph) const {
        if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
            if (ph.adj_group >= 0) return P.adj_sticking[ph.adj_group];
        }
        return param_at(ph.sticking_param, time, ph.sticking);
    }
    __device__ double temperature(const DPhys& ph) const { return param_at(ph.temperature_param, time, ph.temperature); }

    __device__ mfg_facet_counters* fc(const DPhys& ph, int slot) const {
        return P.fc + var->fc_base + (size_t)slot * var->n_facets + (uint32_t)ph.local_id;
    }

    // IncreaseFacetCounter (Particle.cpp:1193-1230). Slot 0 goes to the
    // lane's pending record (summed across the wave before the atomics);
    // moment slots and force terms go straight to memory.
    __device__ void counter(const DPhys& ph, int m, uint32_t hit, uint32_t des, uint32_t abs_, double s1v, double svo,
                            d3 imp, d3 imp2, d3 impm) {
        double he = (double)hit * ori;
        double s1pv = (hit || des) ? (he + (double)des) / vel : 0.0;
        bool force = false;
        if constexpr ((FEAT & FEAT_FORCE) != 0) force = var->enable_force != 0;
        int32_t gf = ph.local_id + (int32_t)var->facet_base;
        if (!force && des == 0) {
            if (pend_f >= 0 && pend_f != gf) flush_pending_direct();
            pend_f = gf;
            pend_hits += hit;
            pend_he += he;
            pend_abs += (double)abs_ * ori;
            pend_s1v += ori * s1v;
            pend_svo += ori * svo;
            pend_s1pv += s1pv;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            int slot = k == 0 ? 0 : m;
            if (k == 1 && m <= 0) break;
            if (k == 0 && !force && des == 0) continue;
            mfg_facet_counters* c = fc(ph, slot);
            if (hit) { atomic_addu(&c->nb_mc_hit, hit); atomic_addd(&c->nb_hit_equiv, he); }
            if (des) atomic_addu(&c->nb_desorbed, des);
            if (abs_) atomic_addd(&c->nb_abs_equiv, (double)abs_ * ori);
            if (s1v != 0.0) atomic_addd(&c->sum_1_per_ort_velocity, ori * s1v);
            if (svo != 0.0) atomic_addd(&c->sum_v_ort, ori * svo);
            if (hit || des) atomic_addd(&c->sum_1_per_velocity, s1pv);
            if constexpr ((FEAT & FEAT_FORCE) != 0) {
                if (force) {
                    double* fo = P.force + var->force_base + ((size_t)slot * var->n_facets + (uint32_t)ph.local_id) * 9;
                    atomic_addd(fo + 0, ori * imp.x); atomic_addd(fo + 1, ori * imp.y); atomic_addd(fo + 2, ori * imp.z);
                    atomic_addd(fo + 3, ori * imp2.x); atomic_addd(fo + 4, ori * imp2.y); atomic_addd(fo + 5, ori * imp2.z);
                    atomic_addd(fo + 6, ori * impm.x); atomic_addd(fo + 7, ori * impm.y); atomic_addd(fo + 8, ori * impm.z);
                }
            }
        }
        if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
            if (hit && ph.adj_target >= 0) adjoint_score(ph, he, false);
        }
        (void)imp; (void)imp2; (void)impm;
    }
    // Without low-flux weighting every particle weighs 1, so the slot-0
    // nb_hit_equiv equals nb_mc_hit: it is not added here (one atomic less
    // per hit) and the host fills it in when it reads the counters.
    __device__ bool weighted() const {
        if constexpr ((FEAT & FEAT_LOWFLUX) != 0) return var->low_flux != 0;
        return false;
    }
    // Atomics of this lane's pending slot-0 record, then clear it.
    __device__ void flush_pending_direct() {
        if (pend_f < 0) return;
        const DPhys& ph = P.phys[pend_f];
        const bool w = weighted();
        if (ph.hot_idx >= 0) {
            double* h = hot + 8 * ph.hot_idx;   // same field order as mfg_facet_counters
            if (pend_hits) {
                atomicAdd(reinterpret_cast<unsigned long long*>(h), (unsigned long long)pend_hits);
                if (w) atomic_addd(h + 1, pend_he);
            }
            if (pend_abs != 0.0) atomic_addd(h + 3, pend_abs);
            if (pend_s1v != 0.0) atomic_addd(h + 4, pend_s1v);
            if (pend_svo != 0.0) atomic_addd(h + 5, pend_svo);
            if (pend_s1pv != 0.0) atomic_addd(h + 6, pend_s1pv);
            clear_pending();
            return;
        }
        mfg_facet_counters* cc = P.fc + var->fc_base + (uint32_t)ph.local_id;
        if (pend_hits) {
            atomic_addu(&cc->nb_mc_hit, pend_hits);
            if (w) atomic_addd(&cc->nb_hit_equiv, pend_he);
        }
        if (pend_abs != 0.0) atomic_addd(&cc->nb_abs_equiv, pend_abs);
        if (pend_s1v != 0.0) atomic_addd(&cc->sum_1_per_ort_velocity, pend_s1v);
        if (pend_svo != 0.0) atomic_addd(&cc->sum_v_ort, pend_svo);
        if (pend_s1pv != 0.0) atomic_addd(&cc->sum_1_per_velocity, pend_s1pv);
        clear_pending();
    }
    __device__ void clear_pending() {
        pend_f = -1; pend_hits = 0;
        pend_he = pend_abs = pend_s1v = pend_svo = pend_s1pv = 0.0;
    }
    // Path score of a hit-equivalent x on a target:
    // d x / d s_g = x * (-n_g / (1 - s_g)) (+ x / s_t for the absorbed part).
    __device__ void adjoint_score(const DPhys& ph, double x, bool absorbed_part) {
        if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
            int t = ph.adj_target;
            atomic_addd(P.adj_j0 + t, x);
            uint32_t G = P.adj_groups;
            for (uint32_t g = 0; g < G; ++g) {
                double s = P.adj_sticking[g];
                double d = -x * (double)adj_cnt[g] / (1.0 - s);
                if (absorbed_part && ph.adj_group == (int)g) d += x / s;
                if (d != 0.0) atomic_addd(P.adj_jac + (size_t)t * G + g, d);
            }
            // kept for the cutoff term until the particle survives a sticky collision
            if (seg_n < MFG_ADJ_SEG) { seg_t[seg_n] = t; seg_x[seg_n] = x; }
            ++seg_n;
        }
        (void)ph; (void)x; (void)absorbed_part;
    }
    // The low-flux cutoff's term of dE[J]/ds_g (not a derivative of any
    // single path): the particle ends at collision K, W_{K-1} (1 - s_K) <=
    // c < W_{K-1}; a rise ds of s_g ends it at K-1 instead with probability
    // f(0) n_g ds / (1 - s_g) (n_g: its collisions with group g up to K-1,
    // f(0): density of ln(W_{K-1}/c) at 0), losing the scores recorded since
    // it survived K-1. In the renewal limit ln(W_{K-1}/c) is uniform on
    // (0, -ln(1 - s_K)] given K, so the term is
    //   dJ_t/ds_g += -sum x_t n_g / ((1 - s_g)(-ln(1 - s_K))).
    __device__ void cutoff_boundary(const DPhys& ph, double sK) {
        if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
            if (seg_n > MFG_ADJ_SEG) atomicOr(P.adj_seg_overflow, 1ull);
            const double dK = -log1p(-sK);
            const uint32_t n_seg = seg_n < MFG_ADJ_SEG ? seg_n : MFG_ADJ_SEG;
            if (!(dK > 0.0) || n_seg == 0) return;
            const uint32_t G = P.adj_groups;
            for (uint32_t i = 0; i < n_seg; ++i) {
                const int t = seg_t[i];
                const double x = seg_x[i];
                for (uint32_t g = 0; g < G; ++g) {
                    const double n = (double)adj_cnt[g] - (ph.adj_group == (int)g ? 1.0 : 0.0);
                    if (n > 0.0) atomic_addd(P.adj_jac + (size_t)t * G + g, -x * n / ((1.0 - P.adj_sticking[g]) * dK));
                }
            }
        }
        (void)ph; (void)sK;
    }
    __device__ void counter_force(const DPhys& ph, int m, uint32_t hit, uint32_t des, uint32_t abs_, double s1v,
                                  double svo, double sign) {
        d3 z = mk(0, 0, 0), v2 = z, mom = z, vv = z;
        if constexpr ((FEAT & FEAT_FORCE) != 0) {
            if (var->enable_force) {
                vv = scl(sign * vel, dir);
                v2 = mk(vv.x * vv.x, vv.y * vv.y, vv.z * vv.z);
                mom = cross(sub(pos, ld3(var->torque)), vv);
            }
        }
        (void)sign;
        counter(ph, m, hit, des, abs_, s1v, svo, vv, v2, mom);
    }
    __device__ bool cell(const DPhys& ph, uint32_t& add) const {
        int tu = (int)(hu * ph.tex_wp), tv = (int)(hv * ph.tex_hp);
        tu = tu > ph.tex_w - 1 ? ph.tex_w - 1 : tu;
        tv = tv > ph.tex_h - 1 ? ph.tex_h - 1 : tv;
        if (tu < 0 || tv < 0) return false;
        add = (uint32_t)tu + (uint32_t)tv * (uint32_t)ph.tex_w;
        return true;
    }
    __device__ double ort(const DFacet* F) const { return fabs(dot(dir, ld3(F->N))); }
    // RecordHitOnTexture (Particle.cpp:977-1004)
    __device__ void texture(const DFacet* F, const DPhys& ph, int m, bool count, double vf, double of) {
        uint32_t add;
        if (!cell(ph, add)) return;
        double ov = var->c_factor * vel * ort(F);
        double inc = P.tex_inc[var->tex_base + (size_t)ph.tex_off + add];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            int slot = k == 0 ? 0 : m;
            if (k == 1 && m <= 0) break;
            mfg_texture_cell* c = P.tex + var->tex_base + (size_t)slot * var->n_tex + (size_t)ph.tex_off + add;
            if (count) atomic_addd(&c->count_equiv, ori);
            atomic_addd(&c->sum_1_per_ort_velocity, ori * vf / ov);
            atomic_addd(&c->sum_v_ort_per_area, ori * of * ov * inc);
        }
    }
    // RecordDirectionVector (Particle.cpp:1006-1023)
    __device__ void direction(const DPhys& ph, int m) {
        if (ph.dir_off < 0) return;
        uint32_t add;
        if (!cell(ph, add)) return;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            int slot = k == 0 ? 0 : m;
            if (k == 1 && m <= 0) break;
            mfg_direction_cell* c = P.dir + var->dir_base + (size_t)slot * var->n_dir + (size_t)ph.dir_off + add;
            atomic_addd(&c->dir[0], ori * dir.x * vel);
            atomic_addd(&c->dir[1], ori * dir.y * vel);
            atomic_addd(&c->dir[2], ori * dir.z * vel);
            atomic_addu(&c->count, 1ull);
        }
    }
    // ProfileFacet (Particle.cpp:1025-1085)
    __device__ void profile(const DFacet* F, const DPhys& ph, int m, bool count, double vf, double of) {
        if (ph.prof_off < 0) return;
        int pt = ph.profile_type;
        double a = ort(F);
        mfg_profile_slice* base = P.prof + var->prof_base + (size_t)ph.prof_off;
        if (count && pt == MFG_PROFILE_ANGULAR) {
            double th = acos(fmin(1.0, a));
            long p = (long)(th / (kPi / 2) * (double)MFG_PROFILE_SIZE);
            p = p < 0 ? 0 : (p > MFG_PROFILE_SIZE - 1 ? MFG_PROFILE_SIZE - 1 : p);
            for (int k = 0; k < 2; ++k) {
                int slot = k == 0 ? 0 : m;
                if (k == 1 && m <= 0) break;
                atomic_addd(&base[(size_t)slot * var->n_prof + p].count_equiv, ori);
            }
        } else if (pt == MFG_PROFILE_U || pt == MFG_PROFILE_V) {
            double c = pt == MFG_PROFILE_U ? hu : hv;
            if (!(c >= 0.0)) return;
            double x = c * (double)MFG_PROFILE_SIZE;
            if (!(x < (double)MFG_PROFILE_SIZE)) return;
            size_t p = (size_t)x;
            double ov = vel * a;
            for (int k = 0; k < 2; ++k) {
                int slot = k == 0 ? 0 : m;
                if (k == 1 && m <= 0) break;
                mfg_profile_slice* s = &base[(size_t)slot * var->n_prof + p];
                if (count) atomic_addd(&s->count_equiv, ori);
                atomic_addd(&s->sum_1_per_ort_velocity, ori * vf / ov);
                atomic_addd(&s->sum_v_ort, ori * of * var->c_factor * ov);
            }
        } else if (count && (pt == MFG_PROFILE_VELOCITY || pt == MFG_PROFILE_ORT_VELOCITY || pt == MFG_PROFILE_TAN_VELOCITY)) {
            double d = pt == MFG_PROFILE_VELOCITY ? 1.0 : (pt == MFG_PROFILE_ORT_VELOCITY ? a : sqrt(fmax(0.0, 1 - a * a)));
            double x = d * vel / ph.max_speed * (double)MFG_PROFILE_SIZE;
            if (!(x >= 0.0) || !(x < (double)MFG_PROFILE_SIZE)) return;
            size_t p = (size_t)x;
            for (int k = 0; k < 2; ++k) {
                int slot = k == 0 ? 0 : m;
                if (k == 1 && m <= 0) break;
                atomic_addd(&base[(size_t)slot * var->n_prof + p].count_equiv, ori);
            }
        }
    }
    __device__ void polar(const DFacet* F, const DPhys& ph, double& th, double& phv) const {
        double lu = dot(dir, ld3(ph.nU)), lv = dot(dir, ld3(ph.nV)), ln = dot(dir, ld3(F->N));
        th = acos(fmax(-1.0, fmin(1.0, ln)));
        phv = atan2(lv, lu);
    }
    // RecordAngleMap (Particle.cpp:1104-1136). The angles only select a bin,
    // so they are evaluated in FP32.
    __device__ void angle_map(const DFacet* F, const DPhys& ph) {
        if (!(ph.flags & FF_ANGLEMAP) || ph.am_off < 0) return;
        double lu = dot(dir, ld3(ph.nU)), lv = dot(dir, ld3(ph.nV)), ln = dot(dir, ld3(F->N));
        double th = (double)acosf((float)fmax(-1.0, fmin(1.0, ln)));
        double phv = (double)atan2f((float)lv, (float)lu);
        if (th > kPi / 2.0) th = fabs(kPi - th);
        long ti;
        if (th < ph.am_theta_limit) {
            if (ph.am_lo == 0) return;
            ti = (long)(th / ph.am_theta_limit * (double)ph.am_lo);
        } else {
            if (ph.am_hi == 0) return;
            ti = (long)ph.am_lo + (long)((th - ph.am_theta_limit) / (kPi / 2.0 - ph.am_theta_limit) * (double)ph.am_hi);
        }
        long nt = (long)(ph.am_lo + ph.am_hi);
        ti = ti < 0 ? 0 : (ti > nt - 1 ? nt - 1 : ti);
        long pi_ = (long)((phv + 3.1415926) / (2.0 * kPi) * (double)ph.am_phi);
        pi_ = pi_ < 0 ? 0 : (pi_ > (long)ph.am_phi - 1 ? (long)ph.am_phi - 1 : pi_);
        atomic_addu(P.am + var->am_base + (size_t)ph.am_off + (size_t)ti * ph.am_phi + pi_, 1ull);
    }
    // LogHit (Particle.cpp:1087-1102)
    __device__ void log_hit(const DFacet* F, const DPhys& ph, int32_t f) {
        if constexpr ((FEAT & FEAT_LOG) != 0) {
            if (var->log_facet == f) {
                unsigned long long at = atomicAdd(P.log_count, 1ull);
                if (at < var->log_limit) {
                    mfg_log_entry e;
                    e.u = hu; e.v = hv;
                    polar(F, ph, e.hit_theta, e.hit_phi);
                    e.ori_ratio = ori; e.particle_decay_moment = decay_t; e.time = time; e.velocity = vel;
                    P.log[at] = e;
                }
            }
        }
        (void)F; (void)ph; (void)f;
    }
    __device__ void hist_add(double* g, uint32_t nb, uint32_t nd, uint32_t nt, uint32_t bbin, double dbin, double tbin) {
        uint32_t off = 0;
        if (nb) { uint32_t b = nb_bounces / bbin; atomic_addd(g + off + (b < nb - 1 ? b : nb - 1), ori); off += nb; }
        if (nd) {
            double x = dist / dbin;
            uint32_t b = x < (double)(nd - 1) ? (uint32_t)x : nd - 1;
            atomic_addd(g + off + b, ori); off += nd;
        }
        if (nt) {
            double x = (time - gen_time) / tbin;
            uint32_t b = x < (double)(nt - 1) ? (uint32_t)x : nt - 1;
            atomic_addd(g + off + b, ori);
        }
    }
    // RecordHistograms (Particle.cpp:926-975), called on absorption
    __device__ void histograms(const DPhys& ph, int m) {
        if constexpr ((FEAT & FEAT_HIST) != 0) {
            for (int k = 0; k < 2; ++k) {
                int slot = k == 0 ? 0 : m;
                if (k == 1 && m <= 0) break;
                if (var->n_gh)
                    hist_add(P.gh + var->gh_base + (size_t)slot * var->n_gh, var->gh_nb_bounce, var->gh_nb_dist,
                             var->gh_nb_time, var->gh_bounce_bin, var->gh_dist_bin, var->gh_time_bin);
                if (ph.fh_off >= 0)
                    hist_add(P.fh + var->fh_base + (size_t)slot * var->n_fh + ph.fh_off, ph.fh_nb_bounce, ph.fh_nb_dist,
                             ph.fh_nb_time, ph.fh_bounce_bin, ph.fh_dist_bin, ph.fh_time_bin);
            }
        }
        (void)ph; (void)m;
    }
    __device__ d3 polar_to_cart(const DFacet* F, const DPhys& ph, double cos_t, double sin_t, double phv, bool reverse) const {
        double s, c;
        sincos(phv, &s, &c);
        double n = reverse ? -cos_t : cos_t;
        return add(add(scl(sin_t * c, ld3(ph.nU)), scl(sin_t * s, ld3(ph.nV))), scl(n, ld3(F->N)));
    }
    // TreatMovingFacet (Physics.cpp:22-39)
    __device__ void moving(d3& d, double& v) const {
        if constexpr ((FEAT & FEAT_MOVING) != 0) {
            d3 a = mk(0, 0, 0);
            if (var->motion_type == 1) a = ld3(var->motion2);
            else if (var->motion_type == 2) a = cross(ld3(var->motion2), scl(0.01, sub(pos, ld3(var->motion1))));
            d3 nv = add(scl(v, d), a);
            v = nrm(nv);
            d = unitv(nv);
        }
        (void)d; (void)v;
    }
    __device__ double maxwell_speed(double sq) {
        XYRef t{P.tab_x + var->maxwell_base, P.tab_y + var->maxwell_base, 100};
        return interp_x_guided(rng.rnd(), t, P.guides + var->maxwell_guide, MFG_MAXWELL_GUIDE) * sq;
    }
    __device__ d3 emit(const DFacet* F, const DPhys& ph, int type, double n_exp, bool reverse) {
        if (type == MFG_DES_UNIFORM) {
            double c = rng.rnd(); double phv = rng.rnd() * 2.0 * kPi;
            return polar_to_cart(F, ph, c, sqrt(1 - c * c), phv, reverse);
        }
        if (type == MFG_DES_COSINE_N) {
            double c = pow(rng.rnd(), 1.0 / (n_exp + 1.0)); double phv = rng.rnd() * 2.0 * kPi;
            return polar_to_cart(F, ph, c, sqrt(1 - c * c), phv, reverse);
        }
        if constexpr ((FEAT & FEAT_AMSRC) != 0) {
            if (type == MFG_DES_ANGLEMAP) {   // Particle.cpp:599-621: the map holds incident angles
                const AmSrc A = am_view(P.am_src + ph.am_src_off);
                double th, ov;
                int tlo;
                am_generate_theta(A, rng.rnd(), th, tlo, ov);
                double phv = am_generate_phi(A, tlo, ov, rng.rnd());
                double t2 = kPi - th;
                return polar_to_cart(F, ph, cos(t2), sin(t2), phv, false);
            }
        }
        double a = rng.rnd(); double phv = rng.rnd() * 2.0 * kPi;
        return polar_to_cart(F, ph, sqrt(a), sqrt(1 - a), phv, reverse);
    }

    // StartFromSource (Particle.cpp:407-685) without the source-facet
    // counter, which the caller adds (summed across the wave).
    __device__ int32_t start(uint64_t local_pid) {
        pid = local_pid;
        rng.init(var->seed, pid);
        nb_events = 0;
        tr_n = 0;
        if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
            for (uint32_t g = 0; g < P.adj_groups; ++g) adj_cnt[g] = 0;
            seg_n = 0;
        }
        double r = rng.rnd() * var->total_des;
        const double* cum = P.src_cum + var->src_base;
        uint32_t n = var->src_n;
        uint32_t lo = 0, hi = n;  // first cum > r
        while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (cum[mid] <= r) lo = mid + 1; else hi = mid; }
        uint32_t k = lo < n ? lo : n - 1;
        int32_t src = P.src_facet[var->src_base + k];
        const DFacet* F = P.facets + src;
        const DPhys& ph = P.phys[src];
        bool reverse = false;
        if (F->flags & FF_2SIDED) reverse = rng.rnd() > 0.5;
        long map_w = -1, map_h = -1;
        if constexpr ((FEAT & FEAT_OGMAP) != 0) {
            if (ph.flags & FF_OGMAP) {
                double base = k ? cum[k - 1] : 0.0;
                double rem = (r - base) / var->latest * (1.38E-23 * ph.temperature);
                const double* oc = P.og_cdf + ph.og_off;
                uint32_t nc = ph.og_w * ph.og_h, a = 0, b = nc;
                while (a < b) { uint32_t mid = (a + b) / 2; if (oc[mid] <= rem) a = mid + 1; else b = mid; }
                uint32_t cellv = a < nc ? a : nc - 1;
                map_h = (long)(cellv / ph.og_w);
                map_w = (long)cellv - map_h * (long)ph.og_w;
            }
        }
        last_hit = src;
        double rt = rng.rnd();
        time = rt * var->latest;
        if constexpr ((FEAT & FEAT_TDPARAM) != 0) {
            if (ph.id_index >= 0) {
                const DTable tb = P.ids[var->id_base + ph.id_index];
                XYRef t{P.tab_x + tb.off, P.tab_y + tb.off, tb.n};
                time = interp_x(rt * t.y[t.n - 1], t);
            }
        }
        gen_time = time;
        if (var->use_maxwell) vel = maxwell_speed(ph.sqrt_temp);
        else vel = 145.469 * sqrt(ph.temperature / var->gas_mass);
        ori = 1.0;
        decay_t = 1e100;
        scatter_path = 1e100;
        if constexpr ((FEAT & FEAT_VOLUME) != 0) {
            if (var->enable_decay) decay_t = time + (1.44269 * var->half_life) * -log(rng.rnd());
            if (var->scatter_enabled) scatter_path = var->scatter_mfp * -log(rng.rnd());
        }
        nb_bounces = 0;
        dist = 0.0;
        bool found = false;
        const double* poly = P.poly + 2 * (size_t)F->poly_off;
        uint32_t pn = F->poly_n;
        d3 O = ld3(F->O), U = ld3(F->U), V = ld3(F->V);
        for (int tr = 0; tr < 1000 && !found; ++tr) {
            double u, v;
            if ((FEAT & FEAT_OGMAP) != 0 && map_w >= 0) {
                double a = rng.rnd(), b = rng.rnd();
                double wp = ph.og_wp, hp = ph.og_hp;
                u = (map_w < (long)ph.og_w - 1) ? ((double)map_w + a) / wp : ((double)map_w + a * (wp - ((double)ph.og_w - 1.0))) / wp;
                v = (map_h < (long)ph.og_h - 1) ? ((double)map_h + b) / hp : ((double)map_h + b * (hp - ((double)ph.og_h - 1.0))) / hp;
            } else {
                u = rng.rnd();
                v = rng.rnd();
            }
            if (in_polygon(poly, pn, u, v)) {
                pos = add(add(O, scl(u, U)), scl(v, V));
                hu = u; hv = v;
                found = true;
            }
        }
        if (!found) {
            if ((FEAT & FEAT_OGMAP) != 0 && map_w >= 0) {
                hu = ((double)map_w + 0.5) / ph.og_wp;
                hv = ((double)map_h + 0.5) / ph.og_hp;
                pos = add(add(O, scl(hu, U)), scl(hv, V));
            } else {
                hu = 0.5; hv = 0.5;
                pos = ld3(ph.centre);
            }
        }
        dir = emit(F, ph, ph.desorb_type, ph.des_exp, reverse);
        structure = ph.super_idx;
        tele_from = -1;
        if constexpr ((FEAT & FEAT_MOVING) != 0) {
            if ((ph.flags & FF_MOVING) && var->motion_type) moving(dir, vel);
        }
        return src;
    }

    // Records of a desorption other than the facet counter of slot 0.
    __device__ void desorption_records(int32_t src, int m, double ov) {
        const DFacet* F = P.facets + src;
        const DPhys& ph = P.phys[src];
        if (m > 0 || ((FEAT & FEAT_FORCE) != 0 && var->enable_force)) {
            // moment slot (and force terms) per lane
            d3 z = mk(0, 0, 0), v2 = z, mom = z, vv = z;
            if constexpr ((FEAT & FEAT_FORCE) != 0) {
                if (var->enable_force) {
                    vv = scl(vel, dir);
                    v2 = mk(vv.x * vv.x, vv.y * vv.y, vv.z * vv.z);
                    mom = cross(sub(pos, ld3(var->torque)), vv);
                    double* fo = P.force + var->force_base + (size_t)(uint32_t)ph.local_id * 9;
                    atomic_addd(fo + 0, vv.x); atomic_addd(fo + 1, vv.y); atomic_addd(fo + 2, vv.z);
                    atomic_addd(fo + 3, v2.x); atomic_addd(fo + 4, v2.y); atomic_addd(fo + 5, v2.z);
                    atomic_addd(fo + 6, mom.x); atomic_addd(fo + 7, mom.y); atomic_addd(fo + 8, mom.z);
                }
            }
            if (m > 0) {
                mfg_facet_counters* c = fc(ph, m);
                atomic_addu(&c->nb_desorbed, 1ull);
                atomic_addd(&c->sum_1_per_ort_velocity, 2.0 / ov);
                atomic_addd(&c->sum_v_ort, var->c_factor * ov);
                atomic_addd(&c->sum_1_per_velocity, 1.0 / vel);
                if constexpr ((FEAT & FEAT_FORCE) != 0) {
                    if (var->enable_force) {
                        double* fo = P.force + var->force_base + ((size_t)m * var->n_facets + (uint32_t)ph.local_id) * 9;
                        atomic_addd(fo + 0, vv.x); atomic_addd(fo + 1, vv.y); atomic_addd(fo + 2, vv.z);
                        atomic_addd(fo + 3, v2.x); atomic_addd(fo + 4, v2.y); atomic_addd(fo + 5, v2.z);
                        atomic_addd(fo + 6, mom.x); atomic_addd(fo + 7, mom.y); atomic_addd(fo + 8, mom.z);
                    }
                }
            }
        }
        profile(F, ph, m, false, 2.0, 1.0);
        log_hit(F, ph, src);
        if ((ph.flags & FF_COUNT_DES) && ph.tex_off >= 0) texture(F, ph, m, true, 2.0, 1.0);
        trace(src, 0);
    }

    // RegisterTransparentPass (Particle.cpp:1232-1257)
    __device__ void transparent_pass(int32_t f, double tp) {
        const DFacet* F = P.facets + f;
        const DPhys& ph = P.phys[f];
        double df = ort(F);
        int m = moment(time + tp / 100.0 / vel);
        d3 z = mk(0, 0, 0);
        counter(ph, m, 1, 0, 0, 2.0 / (vel * df), 2.0 * var->c_factor * vel * df, z, z, z);
        if ((ph.flags & FF_COUNT_TRANS) && ph.tex_off >= 0) texture(F, ph, m, true, 2.0, 2.0);
        if (ph.flags & FF_COUNT_DIR) direction(ph, m);
        log_hit(F, ph, f);
        profile(F, ph, m, true, 2.0, 2.0);
        angle_map(F, ph);
        trace(f, 3);
    }
    // RecordAbsorb (Particle.cpp:888-924)
    __device__ void record_absorb(int32_t f) {
        const DFacet* F = P.facets + f;
        const DPhys& ph = P.phys[f];
        ++acc_hits; acc_hit_equiv += ori; acc_abs_equiv += ori;
        int m = moment(time);
        histograms(ph, m);
        double ov = vel * ort(F);
        bool scored = false;
        if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
            if (ph.adj_target >= 0) {
                // the absorbed part carries its own d/ds_t term
                for (int k = 0; k < 2; ++k) {
                    int slot = k == 0 ? 0 : m;
                    if (k == 1 && m <= 0) break;
                    mfg_facet_counters* c = fc(ph, slot);
                    atomic_addu(&c->nb_mc_hit, 1ull);
                    atomic_addd(&c->nb_hit_equiv, ori);
                    atomic_addd(&c->nb_abs_equiv, ori);
                    atomic_addd(&c->sum_1_per_ort_velocity, ori * 2.0 / ov);
                    atomic_addd(&c->sum_v_ort, ori * var->c_factor * ov);
                    atomic_addd(&c->sum_1_per_velocity, ori / vel);
                }
                adjoint_score(ph, ori, true);
                scored = true;
            }
        }
        if (!scored) counter_force(ph, m, 1, 0, 1, 2.0 / ov, var->c_factor * ov, 1.0);
        log_hit(F, ph, f);
        profile(F, ph, m, true, 2.0, 1.0);
        angle_map(F, ph);
        if ((ph.flags & FF_COUNT_ABS) && ph.tex_off >= 0) texture(F, ph, m, true, 2.0, 1.0);
        if (ph.flags & FF_COUNT_DIR) direction(ph, m);
    }
    // UpdateVelocity (Particle.cpp:1138-1163)
    __device__ void update_velocity(const DPhys& ph) {
        double T = temperature(ph);
        double sq = ph.temperature_param == -1 ? ph.sqrt_temp : sqrt(T);
        if (ph.flags & FF_FULL_ACC) {
            vel = var->use_maxwell ? maxwell_speed(sq) : 145.469 * sq / var->gas_mass;
        } else {
            double o2 = vel * vel, n2;
            if (var->use_maxwell) { double vv = maxwell_speed(sq); n2 = vv * vv; }
            else n2 = 29369.939 * (T / var->gas_mass);
            vel = sqrt(o2 + (n2 - o2) * ph.accomodation);
        }
    }
    // PerformBounce (Particle.cpp:691-838)
    __device__ void bounce(int32_t f) {
        const DFacet* F = P.facets + f;
        const DPhys& ph = P.phys[f];
        ++acc_hits; acc_hit_equiv += ori;
        if constexpr ((FEAT & FEAT_TELEPORT) != 0) {
            if (ph.super_dest) {
                int m = moment(time);
                d3 z = mk(0, 0, 0);
                counter(ph, m, 1, 0, 0, 0.0, 0.0, z, z, z);
                structure = ph.super_dest - 1;
                if constexpr ((FEAT & FEAT_MOVING) != 0) { if (ph.flags & FF_MOVING) moving(dir, vel); }
                log_hit(F, ph, f);
                profile(F, ph, m, true, 2.0, 2.0);
                angle_map(F, ph);
                if ((ph.flags & FF_COUNT_TRANS) && ph.tex_off >= 0) texture(F, ph, m, true, 2.0, 2.0);
                if (ph.flags & FF_COUNT_DIR) direction(ph, m);
                trace(f, 5);
                return;
            }
        }
        d3 N = ld3(F->N);
        double dn = dot(dir, N);
        bool revert = (F->flags & FF_2SIDED) && dn > 0.0;
        double ov_in = vel * fabs(dn);
        int m = moment(time);
        counter_force(ph, m, 1, 0, 0, 1.0 / ov_in, var->c_factor * ov_in, 1.0);
        ++nb_bounces;
        if ((ph.flags & FF_COUNT_REFL) && ph.tex_off >= 0) texture(F, ph, m, true, 1.0, 1.0);
        if (ph.flags & FF_COUNT_DIR) direction(ph, m);
        log_hit(F, ph, f);
        profile(F, ph, m, true, 1.0, 1.0);
        angle_map(F, ph);
        update_velocity(ph);
        int m_out = m;
        if constexpr ((FEAT & FEAT_SOJOURN) != 0) {
            if (ph.flags & FF_SOJOURN) {
                double A = exp(-ph.sojourn_e / (8.31 * temperature(ph)));
                time += -log(rng.rnd()) / (A * ph.sojourn_freq);
                m_out = moment(time);
            }
        }
        if (ph.flags & FF_DIFFUSE) {
            double a = rng.rnd(); double phv = rng.rnd() * 2.0 * kPi;
            dir = polar_to_cart(F, ph, sqrt(a), sqrt(1 - a), phv, revert);
        } else {
            double rt = rng.rnd();
            if (rt < ph.diffuse) {
                double a = rng.rnd(); double phv = rng.rnd() * 2.0 * kPi;
                dir = polar_to_cart(F, ph, sqrt(a), sqrt(1 - a), phv, revert);
            } else if (rt < ph.diffuse + ph.specular) {
                dir = sub(dir, scl(2.0 * dot(dir, N), N));
            } else {
                double c = pow(rng.rnd(), 1.0 / (ph.cos_exp + 1.0)); double phv = rng.rnd() * 2.0 * kPi;
                dir = polar_to_cart(F, ph, c, sqrt(1 - c * c), phv, revert);
            }
        }
        if constexpr ((FEAT & FEAT_MOVING) != 0) {
            if ((ph.flags & FF_MOVING) && var->motion_type) moving(dir, vel);
        }
        double ov_out = vel * ort(F);
        counter_force(ph, m_out, 0, 0, 0, 1.0 / ov_out, var->c_factor * ov_out, -1.0);
        if ((ph.flags & FF_COUNT_REFL) && ph.tex_off >= 0) texture(F, ph, m_out, false, 1.0, 1.0);
        profile(F, ph, m_out, false, 1.0, 1.0);
        last_hit = f;
        trace(f, 1);
    }
    // PerformTeleport (Particle.cpp:76-183). "Back to origin" (teleport_dest
    // = -1) goes to the destination of the particle's previous teleport:
    // the reference records teleportedFrom = destIndex (:122). A destination
    // in every structure (super_idx -1) keeps the particle's structure; the
    // reference would index its structure list with -1 there (:121, :245).
    __device__ void teleport(int32_t f) {
        const DFacet* F = P.facets + f;
        const DPhys& ph = P.phys[f];
        trace(f, 4);
        int32_t dest = ph.teleport_dest == -1 ? tele_from : ph.teleport_dest - 1;
        if (dest < 0 || dest >= (int32_t)var->n_facets) { last_hit = f; return; }
        int32_t dg = dest + (int32_t)var->facet_base;
        const DFacet* D = P.facets + dg;
        const DPhys& dph = P.phys[dg];
        if (dph.super_idx >= 0) structure = dph.super_idx;
        tele_from = dest;
        int m = moment(time);
        if ((ph.flags & FF_COUNT_TRANS) && ph.tex_off >= 0) texture(F, ph, m, true, 2.0, 2.0);
        if (ph.flags & FF_COUNT_DIR) direction(ph, m);
        profile(F, ph, m, true, 2.0, 2.0);
        log_hit(F, ph, f);
        angle_map(F, ph);
        double lu = dot(dir, ld3(ph.nU)), lv = dot(dir, ld3(ph.nV)), ln = dot(dir, ld3(F->N));
        dir = add(add(scl(lu, ld3(dph.nU)), scl(lv, ld3(dph.nV))), scl(ln, ld3(D->N)));
        double u = hu, v = hv;
        d3 O = ld3(D->O), U = ld3(D->U), V = ld3(D->V);
        pos = add(add(O, scl(u, U)), scl(v, V));
        const double* poly = P.poly + 2 * (size_t)D->poly_off;
        if (!in_polygon(poly, D->poly_n, u, v)) {
            for (int tr = 0; tr < 1000; ++tr) {
                u = rng.rnd(); v = rng.rnd();
                if (in_polygon(poly, D->poly_n, u, v)) { pos = add(add(O, scl(u, U)), scl(v, V)); break; }
            }
        }
        last_hit = dg;
        double ov = vel * ort(F);
        d3 z = mk(0, 0, 0);
        counter(ph, m, 1, 0, 0, 2.0 / ov, 2.0 * var->c_factor * ov, z, z, z);
    }
    // PerformScatter (Particle.cpp:840-886, Physics.cpp:41-63)
    __device__ bool scatter() {
        ++acc_hits; acc_hit_equiv += ori;
        ++nb_bounces;
        double rho = var->scatter_rho, mr = var->scatter_mr;
        double b = rho * sqrt(rng.rnd());
        double a = sqrt(rho * rho - b * b);
        double ctg = 1.0 / (2.0 * a * b) * (rho * rho * mr + b * b - a * a);
        double th = atan(1.0 / ctg);
        double az = 2.0 * kPi * rng.rnd();
        d3 rv = dir.x != 0.0 ? mk(0, 0, 1) : (dir.y != 0.0 ? mk(1, 0, 1) : mk(1, 0, 0));
        d3 un = unitv(cross(dir, rv)), vn = unitv(cross(dir, un));
        double st = sin(th), ct = cos(th), sa, ca;
        sincos(az, &sa, &ca);
        dir = add(add(scl(st * ca, un), scl(st * sa, vn)), scl(ct, dir));
        double q = mr * rho / b * st;
        vel = sqrt(vel * vel / (1.0 + q * q));
        scatter_path = dist + var->scatter_mfp * -log(rng.rnd());
        trace(-1, 8);
        return !(var->scatter_cutoff && vel < var->scatter_cut);
    }

    // Add this lane's global counters to its variant and clear them.
    // ---- particles kept between launches
    __device__ void save(CarryRec& r, uint64_t events_left, uint32_t vi) const {
        r.pos[0] = pos.x; r.pos[1] = pos.y; r.pos[2] = pos.z;
        r.dir[0] = dir.x; r.dir[1] = dir.y; r.dir[2] = dir.z;
        r.time = time; r.vel = vel; r.ori = ori; r.gen_time = gen_time; r.dist = dist;
        r.decay_t = decay_t; r.scatter_path = scatter_path; r.hu = hu; r.hv = hv;
        r.pid = pid; r.events_left = events_left;
        r.k0 = rng.k0; r.k1 = rng.k1; r.blk = rng.blk;
        r.b0 = rng.b0; r.b1 = rng.b1; r.b2 = rng.b2; r.b3 = rng.b3; r.avail = rng.avail;
        r.structure = structure; r.last_hit = last_hit; r.tele_from = tele_from;
        r.nb_bounces = nb_bounces; r.nb_events = nb_events; r.variant = vi; r.tr_n = tr_n;
    }
    __device__ uint64_t restore(const CarryRec& r) {
        pos = mk(r.pos[0], r.pos[1], r.pos[2]);
        dir = mk(r.dir[0], r.dir[1], r.dir[2]);
        time = r.time; vel = r.vel; ori = r.ori; gen_time = r.gen_time; dist = r.dist;
        decay_t = r.decay_t; scatter_path = r.scatter_path; hu = r.hu; hv = r.hv;
        pid = r.pid;
        rng.k0 = r.k0; rng.k1 = r.k1; rng.p0 = (uint32_t)r.pid; rng.p1 = (uint32_t)(r.pid >> 32);
        rng.blk = r.blk; rng.b0 = r.b0; rng.b1 = r.b1; rng.b2 = r.b2; rng.b3 = r.b3; rng.avail = r.avail;
        structure = r.structure; last_hit = r.last_hit; tele_from = r.tele_from;
        nb_bounces = r.nb_bounces; nb_events = r.nb_events; tr_n = r.tr_n;
        return r.events_left;
    }

    __device__ void flush_lane() {
        if (var == nullptr) return;
        DGlobal* gl = P.glob + var->glob_index;
        if (acc_hits) atomic_addu(&gl->nb_mc_hit, acc_hits);
        if (acc_des) atomic_addu(&gl->nb_desorbed, acc_des);
        if (acc_hit_equiv != 0.0) atomic_addd(&gl->nb_hit_equiv, acc_hit_equiv);
        if (acc_abs_equiv != 0.0) atomic_addd(&gl->nb_abs_equiv, acc_abs_equiv);
        if (acc_leak) atomic_addu(&gl->nb_leak, acc_leak);
        if (acc_dist_total != 0.0) atomic_addd(&gl->dist_total, acc_dist_total);
        if (acc_dist_full != 0.0) atomic_addd(&gl->dist_full, acc_dist_full);
        if (acc_decay) atomic_addu(&gl->nb_decayed, acc_decay);
        acc_hits = acc_des = acc_leak = acc_decay = 0;
        acc_hit_equiv = acc_abs_equiv = acc_dist_total = acc_dist_full = 0.0;
    }

    // One event of an alive particle (one SimulationMCStep iteration).
    // Returns false when the particle ends.
    __device__ bool step() {
        constexpr bool TR = (FEAT & FEAT_TRANSPARENT) != 0;
        ++nb_events;
        int32_t root = P.roots[var->root_base + (uint32_t)structure];
        HitRec h = traverse<TR>(P, root, pos, dir, last_hit, 1e300, S, &rng, nb_events, var,
                                time, false, lnodes);
        if (h.facet < 0) {  // leak
            ++acc_leak;
            unsigned long long at = atomicAdd(P.leak_count, 1ull);
            if (at < P.leak_cap) {
                double* L = P.leaks + at * 6;
                L[0] = pos.x; L[1] = pos.y; L[2] = pos.z; L[3] = dir.x; L[4] = dir.y; L[5] = dir.z;
            }
            trace(-1, 6);
            return false;
        }
        double travel = h.t;
        int ev = 0;
        if constexpr ((FEAT & FEAT_VOLUME) != 0) {
            if (!var->calc_const_flow) { double dl = vel * 100.0 * (var->latest - time); if (dl < travel) { travel = dl; ev = 1; } }
            if (var->enable_decay) { double dd = vel * 100.0 * (decay_t - time); if (dd < travel) { travel = dd; ev = 2; } }
            if (var->scatter_enabled) { double ds = scatter_path - dist; if (ds < travel) { travel = ds; ev = 3; } }
        }
        if constexpr (TR) {
            if (var->flags & FEAT_TRANSPARENT) {
                // second walk: every transparent pass closer than `travel`
                Ray32 r = make_ray32(pos, dir);
                float tbf = widen(travel);
                uint32_t sp = 0;
                int32_t node = root;
                double save_u = hu, save_v = hv;
                for (;;) {
                    const NodeRegs nd = load_node_at(P.nodes, lnodes, P.lds_nodes, node);
                    uint32_t msk = node_hits(nd, r, tbf);
                    while (msk) {
                        uint32_t c = (uint32_t)__builtin_ctz(msk);
                        msk &= msk - 1u;
                        if (node_inner(nd) & (1u << c)) {
                            S.push(sp, node_inner_child(nd, c));
                            continue;
                        }
                        uint32_t first, cnt;
                        node_leaf_child(nd, c, first, cnt);
                        for (uint32_t kk = 0; kk < cnt; ++kk) {
                            double t, u, v;
                            int32_t f;
                            uint32_t fl;
                            bool hit = leaf_test(P.leaves + first + kk, P.leaf_uv + first + kk, P.poly, pos, dir, last_hit,
                                                 t, u, v, f, fl);
                            if (!hit || (fl & FF_OPAQUE)) continue;
                            if (!(t < travel)) continue;
                            double op = (fl & FF_TD_OPACITY) ? td_opacity(P, var, f, time) : (double)P.phys[f].opacity;
                            bool hard = op >= 1.0 || (op > 0.0 && rng.coin(nb_events, (uint32_t)local(f)) < op);
                            if (hard) continue;
                            hu = u; hv = v;
                            transparent_pass(f, t);
                        }
                    }
                    if (sp == 0) break;
                    node = S.pop(sp);
                }
                hu = save_u; hv = save_v;
            }
        }
        pos = add(pos, scl(travel, dir));
        time += travel / 100.0 / vel;
        if (ev == 0) {
            int32_t f = h.facet;
            const DPhys& ph = P.phys[f];
            hu = h.u; hv = h.v;
            acc_dist_total += travel; acc_dist_full += travel; dist += travel;
            if constexpr ((FEAT & FEAT_TELEPORT) != 0) {
                if (ph.teleport_dest != 0) { teleport(f); return true; }
            }
            double s = sticking(ph);
            bool lowflux = false;
            if constexpr ((FEAT & FEAT_LOWFLUX) != 0) lowflux = var->low_flux != 0;
            if (!lowflux) {
                if (s == 1.0 || (s > 0.0 && rng.rnd() < s)) { record_absorb(f); trace(f, 2); return false; }
                bounce(f);
                return true;
            }
            if (s > 0.0) {
                double w0 = ori;
                ori *= s;
                record_absorb(f);
                ori = w0 * (1.0 - s);
                if constexpr ((FEAT & FEAT_ADJOINT) != 0) { if (ph.adj_group >= 0) adj_cnt[ph.adj_group] += 1; }
            } else {
                ori *= (1.0 - s);
            }
            if (ori > var->low_flux_cutoff) {
                if constexpr ((FEAT & FEAT_ADJOINT) != 0) { if (s > 0.0) seg_n = 0; }
                bounce(f);
                return true;
            }
            if constexpr ((FEAT & FEAT_ADJOINT) != 0) { if (s > 0.0) cutoff_boundary(ph, s); }
            trace(f, 2);
            return false;
        }
        if (ev == 1 || ev == 2) {
            acc_dist_total += travel;
            ++acc_decay;
            trace(-1, 7);
            return false;
        }
        acc_dist_total += travel; acc_dist_full += travel; dist += travel;
        bool ok = scatter();
        last_hit = -1;
        if (!ok) { ++acc_decay; return false; }
        return true;
    }
};

__device__ __forceinline__ uint32_t find_variant(const uint64_t* prefix, uint32_t n, uint64_t g) {
    uint32_t lo = 0, hi = n;  // largest v with prefix[v] <= g
    while (hi - lo > 1) { uint32_t mid = (lo + hi) / 2; if (prefix[mid] <= g) lo = mid; else hi = mid; }
    return lo;
}

template <uint32_t FEAT>
__device__ __forceinline__ void trace_body(const TraceParams& P) {
    extern __shared__ int32_t lds_stack[];
    const uint32_t lane = __lane_id();
    const size_t gl = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const LaneStack lane_stack{lds_stack + threadIdx.x, blockDim.x, P.stack_size, P.stack_total,
                               P.stack_spill ? P.stack_spill + gl * (P.stack_total - P.stack_size)
                                             : lds_stack + threadIdx.x};
    Tracer<FEAT> T(P, lane_stack);
    // hot-facet counters after the stacks (stack bytes are a multiple of 8)
    double* hot = reinterpret_cast<double*>(lds_stack + (size_t)blockDim.x * P.stack_size);
    for (uint32_t i = threadIdx.x; i < 8u * P.n_hot; i += blockDim.x) hot[i] = 0.0;
    // then the block's copy of the top of the tree (128 B per node; the
    // offsets above are multiples of 64 B)
    uint4* lnodes = reinterpret_cast<uint4*>(hot + 8u * P.n_hot);
    {
        const uint4* g = reinterpret_cast<const uint4*>(P.nodes);
        for (uint32_t i = threadIdx.x; i < 8u * P.lds_nodes; i += blockDim.x) lnodes[i] = g[i];
    }
    __syncthreads();
    T.hot = hot;
    T.lnodes = lnodes;
    if constexpr ((FEAT & FEAT_ADJOINT) != 0) {
        T.adj_cnt = P.adj_counts + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * P.adj_groups;
        T.seg_t = P.adj_seg_t + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * MFG_ADJ_SEG;
        T.seg_x = P.adj_seg_x + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * MFG_ADJ_SEG;
    }
    bool alive = false, done = false;
    uint32_t vi = 0;
    uint64_t events_left = 0;
    // work items: first the particles the previous launch left in flight,
    // then this launch's desorptions
    uint64_t n_in = 0;
    const CarryRec* carry_in = nullptr;
    CarryRec* carry_out = nullptr;
    unsigned long long* n_out = nullptr;
    if (P.carry) {
        uint32_t half = (uint32_t)P.carry_ctl[2] & 1u;
        n_in = P.carry_ctl[half];
        carry_in = P.carry + (size_t)half * P.carry_cap;
        carry_out = P.carry + (size_t)(half ^ 1u) * P.carry_cap;
        n_out = P.carry_ctl + (half ^ 1u);
    }
    const uint64_t items = n_in + P.total;
    const uint64_t t_start = wall_clock64();
    uint32_t iter = 0;   // the loop is wave-uniform, and so is this count
    // the stop flag lives in host memory: a read crosses the bus, so it is
    // polled every 64 iterations, not at every refill
    int stop = __builtin_amdgcn_readfirstlane(*P.stop_flag);
    for (;;) {
        // the wall clock and the stop flag (read every 64 iterations) are
        // one value for the wave: expired and stop are uniform
        bool expired = false;
        if ((++iter & 63u) == 0u) {
            expired = __builtin_amdgcn_readfirstlane((int)((uint64_t)(wall_clock64() - t_start) > P.time_limit)) != 0;
            stop = __builtin_amdgcn_readfirstlane(*P.stop_flag);
        }
        if (expired && lane == 0) atomicMax(P.expired, 1ull);
        bool need = !alive && !done;
        uint64_t mask = expired ? 0ull : __ballot(need);
        bool dry = false;   // some lane of this wave found the queue exhausted
        if (mask) {
            uint32_t cnt = (uint32_t)__popcll(mask);
            uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(P.next_particle, (unsigned long long)cnt);
            base = __shfl(base, leader, 64);
            // after a stop request no new particle starts; resumed ones still
            // run (to the end, or back into the carry buffer)
            const uint64_t limit = stop ? n_in : items;
            bool started = false;
            int32_t src = -1;
            if (need) {
                uint64_t g = base + mbcnt(mask);
                if (g < n_in) {
                    const CarryRec& r = carry_in[g];
                    vi = r.variant;
                    if (T.var != P.variants + vi) T.flush_lane();
                    T.var = P.variants + vi;
                    events_left = T.restore(r);
                    alive = true;
                } else if (g < limit) {
                    uint64_t d = g - n_in;
                    vi = P.n_variants > 1 ? find_variant(P.var_prefix, P.n_variants, d) : 0;
                    if (T.var != P.variants + vi) T.flush_lane();
                    T.var = P.variants + vi;
                    ++T.acc_des;
                    uint64_t lp = d - P.var_prefix[vi] + T.var->pid_base;
                    src = T.start(lp);
                    started = true;
                    alive = true;
                    events_left = P.max_events ? P.max_events : ~0ull;
                } else {
                    done = true;
                }
            }
            dry = __ballot(done && need) != 0;
            // desorption counters: lanes sharing the leader's source facet
            // add through one set of atomics
            uint64_t smask = __ballot(started);
            if (smask) {
                const DPhys* ph = started ? &P.phys[src] : nullptr;
                double ov = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
                int m = -1;
                if (started) {
                    ov = T.vel * fabs(dot(T.dir, ld3(P.facets[src].N)));
                    w1 = 2.0 / ov; w2 = T.var->c_factor * ov; w3 = 1.0 / T.vel;
                    m = T.moment(T.time);
                }
                uint32_t sl = (uint32_t)__ffsll((unsigned long long)smask) - 1;
                int32_t key = __shfl(src, sl, 64);
                bool grp = started && src == key;
                uint64_t gmask = __ballot(grp);
                double s1 = wave_sum(grp ? w1 : 0.0), s2 = wave_sum(grp ? w2 : 0.0), s3 = wave_sum(grp ? w3 : 0.0);
                if (grp && lane == sl) {
                    if (ph->hot_idx >= 0) {
                        double* h = hot + 8 * ph->hot_idx;
                        atomicAdd(reinterpret_cast<unsigned long long*>(h + 2), (unsigned long long)__popcll(gmask));
                        atomic_addd(h + 4, s1);
                        atomic_addd(h + 5, s2);
                        atomic_addd(h + 6, s3);
                    } else {
                        mfg_facet_counters* c = T.fc(*ph, 0);
                        atomic_addu(&c->nb_desorbed, (unsigned long long)__popcll(gmask));
                        atomic_addd(&c->sum_1_per_ort_velocity, s1);
                        atomic_addd(&c->sum_v_ort, s2);
                        atomic_addd(&c->sum_1_per_velocity, s3);
                    }
                }
                if (started && !grp) {
                    mfg_facet_counters* c = T.fc(*ph, 0);
                    atomic_addu(&c->nb_desorbed, 1ull);
                    atomic_addd(&c->sum_1_per_ort_velocity, w1);
                    atomic_addd(&c->sum_v_ort, w2);
                    atomic_addd(&c->sum_1_per_velocity, w3);
                }
                if (started) T.desorption_records(src, m, ov);
            }
        }
        if ((P.carry_keep && dry) || expired) {
            // the queue is empty (or the time limit is reached): park this
            // wave's particles for the next launch instead of running their
            // tails here; without carry, expiry ends them where they are
            uint64_t am = P.carry_keep ? __ballot(alive) : 0ull;
            if (am) {
                uint32_t al = (uint32_t)__ffsll((unsigned long long)am) - 1;
                unsigned long long ob = 0;
                if (lane == al) ob = atomicAdd(n_out, (unsigned long long)__popcll(am));
                ob = __shfl(ob, al, 64);
                if (alive) T.save(carry_out[ob + mbcnt(am)], events_left, vi);
            }
            break;
        }
        if (!__ballot(!done)) break;
        if (alive) {
            alive = T.step();
            if (alive && --events_left == 0) alive = false;
        }
        // slot-0 facet counters of this iteration: lanes that updated the
        // same facet as the first pending lane add through one set of atomics
        uint64_t pm = __ballot(T.pend_f >= 0);
        if (pm) {
            uint32_t ld = (uint32_t)__ffsll((unsigned long long)pm) - 1;
            int32_t key = __shfl(T.pend_f, ld, 64);
            bool grp = T.pend_f == key;
            uint64_t gm = __ballot(grp);
            if (__popcll(gm) > 1) {
                double he = 0.0;
                if constexpr ((FEAT & FEAT_LOWFLUX) != 0) he = wave_sum(grp ? T.pend_he : 0.0);
                double ab = wave_sum(grp ? T.pend_abs : 0.0);
                double s1 = wave_sum(grp ? T.pend_s1v : 0.0), sv = wave_sum(grp ? T.pend_svo : 0.0);
                double sp = wave_sum(grp ? T.pend_s1pv : 0.0);
                unsigned long long hc = wave_sum_u(grp ? (unsigned long long)T.pend_hits : 0ull);
                if (lane == ld) {
                    T.pend_hits = (uint32_t)hc; T.pend_he = he; T.pend_abs = ab;
                    T.pend_s1v = s1; T.pend_svo = sv; T.pend_s1pv = sp;
                } else if (grp) {
                    T.clear_pending();
                }
            }
            if (T.pend_f >= 0) T.flush_pending_direct();
        }
    }
    // flush lane accumulators: one set of atomics per wave for a single
    // variant, per lane otherwise
    if (P.n_variants == 1) {
        double he = wave_sum(T.acc_hit_equiv), ae = wave_sum(T.acc_abs_equiv);
        double d1 = wave_sum(T.acc_dist_total), d2 = wave_sum(T.acc_dist_full);
        unsigned long long hits = wave_sum_u(T.acc_hits), des = wave_sum_u(T.acc_des);
        unsigned long long lk = wave_sum_u(T.acc_leak), dc = wave_sum_u(T.acc_decay);
        if (lane == 0) {
            DGlobal* gl = P.glob + P.variants[0].glob_index;
            if (hits) atomic_addu(&gl->nb_mc_hit, hits);
            if (des) atomic_addu(&gl->nb_desorbed, des);
            if (he != 0.0) atomic_addd(&gl->nb_hit_equiv, he);
            if (ae != 0.0) atomic_addd(&gl->nb_abs_equiv, ae);
            if (lk) atomic_addu(&gl->nb_leak, lk);
            if (d1 != 0.0) atomic_addd(&gl->dist_total, d1);
            if (d2 != 0.0) atomic_addd(&gl->dist_full, d2);
            if (dc) atomic_addu(&gl->nb_decayed, dc);
        }
    } else {
        T.flush_lane();
    }
    // every wave of the block has left the loop: add the block's hot-facet
    // sums to the global records
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 8u * P.n_hot; i += blockDim.x) {
        const uint32_t k = i >> 3, fld = i & 7u;
        if (fld == 7u) continue;
        double* g = reinterpret_cast<double*>(P.fc + P.hot_fc[k]) + fld;
        if (fld == 0u || fld == 2u) {
            unsigned long long v = *reinterpret_cast<unsigned long long*>(hot + i);
            if (v) atomic_addu(reinterpret_cast<unsigned long long*>(g), v);
        } else if (hot[i] != 0.0) {
            atomic_addd(g, hot[i]);
        }
    }
}

// One instantiation per occupancy bound: WAVES waves per SIMD caps the
// VGPR budget (cold per-particle state spills to scratch at the higher
// bounds); the host autotuner picks the fastest for a model.
template <uint32_t FEAT, int WAVES>
__global__ void __launch_bounds__(256, WAVES) k_trace(TraceParams P) {
    trace_body<FEAT>(P);
}

// One block per CU of 256 lanes per wave slot (768 lanes at 3 waves per
// SIMD, 1024 at 4): all waves of a CU share one LDS copy of the top of the
// tree, which can then hold several times the nodes of a 256-lane block's
// copy (same per-lane LDS budget).
template <uint32_t FEAT, int WAVES>
__global__ void __launch_bounds__(256 * WAVES, WAVES) k_trace_cu(TraceParams P) {
    trace_body<FEAT>(P);
}

// Parked particles that a launch did not claim: the queue hands out items
// [0, claimed) and the first n_in of them are the resumed particles, so
// [min(claimed, n_in), n_in) were never taken. That happens only when the
// launch time limit stops the claims (a dry queue means every item was
// claimed). These records are appended to the output half after the ones
// the waves parked themselves.
__device__ __forceinline__ void carry_counts(const unsigned long long* ctl, const unsigned long long* claimed,
                                             uint64_t& half, uint64_t& first, uint64_t& n, uint64_t& n_out) {
    half = ctl[2] & 1ull;
    const uint64_t n_in = ctl[half];
    first = *claimed < n_in ? *claimed : n_in;
    n = n_in - first;
    n_out = ctl[half ^ 1ull];
}
__global__ void k_carry_rescue(const unsigned long long* ctl, const unsigned long long* claimed, CarryRec* carry,
                               uint64_t cap) {
    uint64_t half, first, n, n_out;
    carry_counts(ctl, claimed, half, first, n, n_out);
    const CarryRec* in = carry + half * cap;
    CarryRec* out = carry + (half ^ 1ull) * cap;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (n_out + i < cap) out[n_out + i] = in[first + i];
}
// After a launch (and the rescue): the half it resumed from is consumed,
// and the half that received its leftovers becomes the next launch's
// input. overflow is set if the rescued records did not fit.
__global__ void k_carry_flip(unsigned long long* ctl, const unsigned long long* claimed, uint64_t cap,
                             unsigned long long* overflow) {
    if (threadIdx.x == 0) {
        uint64_t half, first, n, n_out;
        carry_counts(ctl, claimed, half, first, n, n_out);
        uint64_t total = n_out + n;
        if (total > cap) { *overflow = 1ull; total = cap; }
        ctl[half ^ 1ull] = total;
        ctl[half] = 0ull;
        ctl[2] = half ^ 1ull;
    }
}

__global__ void __launch_bounds__(256) k_intersect(IsectParams p) {
    extern __shared__ int32_t lds_stack[];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    TraceParams tp{};
    tp.facets = p.facets; tp.poly = p.poly; tp.nodes = p.nodes; tp.leaves = p.leaves; tp.leaf_uv = p.leaf_uv;
    d3 P = ld3(p.origins + 3 * i), D = ld3(p.dirs + 3 * i);
    int32_t skip = p.last_hit ? p.last_hit[i] : -1;
    if (skip >= 0) skip += (int32_t)p.facet_base;
    // the whole stack is in LDS here (cap == lds_cap): the spill pointer is never used
    const LaneStack S{lds_stack + threadIdx.x, blockDim.x, p.stack_size, p.stack_size, lds_stack + threadIdx.x};
    HitRec h = traverse<false>(tp, p.root, P, D, skip, 1e300, S,
                               nullptr, 0, nullptr, 0.0, true, nullptr);
    p.out_facet[i] = h.facet >= 0 ? h.facet - (int32_t)p.facet_base : -1;
    p.out_t[i] = h.facet >= 0 ? h.t : 0.0;
    p.out_u[i] = h.u;
    p.out_v[i] = h.v;
}

__global__ void k_zero(uint64_t* p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

__global__ void k_backward(const double* jac, const double* up, double* grad, uint32_t nt, uint32_t ng, double scale) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    double s = 0.0;
    for (uint32_t t = 0; t < nt; ++t) s += up[t] * jac[(size_t)t * ng + g];
    grad[g] = s * scale;
}

// Instantiated feature sets: a model runs on the smallest one covering its
// features (fewer live registers, higher occupancy).
constexpr uint32_t kLean = 0;
constexpr uint32_t kLowFlux = FEAT_LOWFLUX;
constexpr uint32_t kLowFluxAdj = FEAT_LOWFLUX | FEAT_ADJOINT;
constexpr uint32_t kFull = FEAT_ALL;
constexpr uint32_t kFullAdj = FEAT_ALL | FEAT_ADJOINT;

}  // namespace

uint32_t mfg_trace_kernel_features(uint32_t model_features) {
    const uint32_t rest = model_features & FEAT_ALL & ~FEAT_LOWFLUX;
    if (model_features & FEAT_ADJOINT) return rest ? kFullAdj : kLowFluxAdj;
    if (rest) return kFull;
    return (model_features & FEAT_LOWFLUX) ? kLowFlux : kLean;
}

namespace {
typedef void (*TraceKernel)(TraceParams);
// [feature set][waves per SIMD - 1]; null where not instantiated
// The one-block-per-CU kernels (lean and low-flux feature sets, 3 or 4
// waves per SIMD: 768 or 1024 lanes; low flux with the adjoint at 3).
TraceKernel kernel_cu(uint32_t features, int waves) {
    if (waves == 3) {
        if (features == kLean) return k_trace_cu<kLean, 3>;
        if (features == kLowFlux) return k_trace_cu<kLowFlux, 3>;
        if (features == kLowFluxAdj) return k_trace_cu<kLowFluxAdj, 3>;
    } else if (waves == 4) {
        if (features == kLean) return k_trace_cu<kLean, 4>;
        if (features == kLowFlux) return k_trace_cu<kLowFlux, 4>;
    }
    return nullptr;
}

TraceKernel kernel_for(uint32_t features, int waves) {
    constexpr int N = 5;   // occupancy bounds 1..5 waves per SIMD
    static TraceKernel lean[N] = {nullptr, k_trace<kLean, 2>, k_trace<kLean, 3>, k_trace<kLean, 4>, k_trace<kLean, 5>};
    static TraceKernel full[N] = {k_trace<kFull, 1>, k_trace<kFull, 2>, k_trace<kFull, 3>, nullptr, nullptr};
    static TraceKernel adj[N] = {k_trace<kFullAdj, 1>, k_trace<kFullAdj, 2>, nullptr, nullptr, nullptr};
    static TraceKernel lowflux[N] = {nullptr, k_trace<kLowFlux, 2>, k_trace<kLowFlux, 3>, k_trace<kLowFlux, 4>,
                                     k_trace<kLowFlux, 5>};
    static TraceKernel lowflux_adj[N] = {nullptr, k_trace<kLowFluxAdj, 2>, k_trace<kLowFluxAdj, 3>,
                                         k_trace<kLowFluxAdj, 4>, nullptr};
    TraceKernel* t = features == kLean ? lean
                   : features == kFull ? full
                   : features == kLowFlux ? lowflux
                   : features == kLowFluxAdj ? lowflux_adj : adj;
    if (waves < 1) waves = 1;
    if (waves > N) waves = N;
    for (int w = waves; w >= 1; --w) if (t[w - 1]) return t[w - 1];
    for (int w = waves; w <= N; ++w) if (t[w - 1]) return t[w - 1];
    return t[0];
}
}  // namespace

int mfg_trace_variants(uint32_t features, int* waves_out, int cap) {
    uint32_t f = mfg_trace_kernel_features(features);
    int lo = 1, hi = 2;
    if (f == kLean || f == kLowFlux) { lo = 2; hi = 5; }
    else if (f == kLowFluxAdj) { lo = 2; hi = 4; }
    else if (f == kFull) hi = 3;
    int n = 0;
    for (int w = lo; w <= hi && n < cap; ++w) waves_out[n++] = w;
    return n;
}

int mfg_trace_has_cu_block(uint32_t features, int block) {
    return block % 256 == 0 && kernel_cu(mfg_trace_kernel_features(features), block / 256) != nullptr;
}

int mfg_trace_occupancy(uint32_t features, int waves, int block, size_t lds) {
    int n = 0;
    const uint32_t f = mfg_trace_kernel_features(features);
    TraceKernel k = block > 256 ? kernel_cu(f, block / 256) : kernel_for(f, waves);
    if (!k) return 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, block, lds);
    return e == hipSuccess ? n : 1;
}

int mfg_launch_trace(const TraceParams& p, uint32_t features, int waves, int grid, int block, void* stream) {
    size_t lds = (size_t)block * p.stack_size * sizeof(int32_t) + (size_t)p.n_hot * 8 * sizeof(double) +
                 (size_t)p.lds_nodes * sizeof(BvhNode);
    const uint32_t f = mfg_trace_kernel_features(features);
    TraceKernel k = block > 256 ? kernel_cu(f, block / 256) : kernel_for(f, waves);
    if (!k) return -1;
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, (hipStream_t)stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mfg_launch_carry_flip(unsigned long long* ctl, const unsigned long long* claimed, CarryRec* carry, uint64_t cap,
                          unsigned long long* overflow, void* stream) {
    uint64_t blocks = (cap + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_carry_rescue, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, ctl, claimed, carry,
                       cap);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(k_carry_flip, dim3(1), dim3(64), 0, (hipStream_t)stream, ctl, claimed, cap, overflow);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mfg_launch_intersect(const IsectParams& p, void* stream) {
    const int block = 256;
    uint64_t grid = (p.n + block - 1) / block;
    if (grid == 0) return 0;
    size_t lds = (size_t)block * p.stack_size * sizeof(int32_t);
    hipLaunchKernelGGL(k_intersect, dim3((uint32_t)grid), dim3(block), lds, (hipStream_t)stream, p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mfg_launch_zero(void* ptr, uint64_t bytes, void* stream) {
    uint64_t n = bytes / 8;
    if (n == 0) return 0;
    uint64_t grid = (n + 255) / 256;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_zero, dim3((uint32_t)grid), dim3(256), 0, (hipStream_t)stream, (uint64_t*)ptr, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Device staging of the allocation-free fixed-ray-batch entry
// (mfg_intersect_batch_static): origins, directions, skip ids and the four
// outputs of up to MFG_STATIC_MAX_RAYS rays, at the fixed offsets of
// StaticRays. Every GPU the library runs on holds its own copy until the
// process ends; host code reaches it only through the address that
// mfg_scratch_address returns for the current device.
static __device__ __align__(256) unsigned char scratch[sizeof(StaticRays)];

namespace {
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MFG_MAX_DEVICES] = {};
}  // namespace

int mfg_scratch_address(int device, StaticRays** out) {
    if (!out || device < 0 || device >= MFG_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        if (hipGetDevice(&cur) != hipSuccess || cur != device) return -1;   // the device must be current
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<StaticRays*>(g_scratch_ptr[device]);
    return 0;
}

int mfg_launch_backward(const double* jac, const double* up, double* grad, uint32_t nt, uint32_t ng, double scale,
                        void* stream) {
    hipLaunchKernelGGL(k_backward, dim3((ng + 63) / 64), dim3(64), 0, (hipStream_t)stream, jac, up, grad, nt, ng, scale);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
