// mfg_host.cpp -- host runtime and C ABI of the MI355X tracer.
//
// Model preparation restates what the reference does before a run
// (MolflowSimulationModel::PrepareToRun, MolflowSimGeom.cpp:194-335;
// texture cell increments, Initializer.cpp:433-456; Maxwell CDF,
// CDFGeneration.cpp:55-83; integrated desorption, IDGeneration.cpp:45-206;
// moments, TimeMoments.cpp:58-83), builds the per-structure BVH
// (BuildAccelStructure, MolflowSimGeom.cpp:124-184) as 8-wide (or 4-wide)
// quantised nodes, and moves everything to HBM once. Runs launch the
// persistent trace kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "mfg_internal.h"

namespace {

constexpr double kBoltz = 1.38e-23;
constexpr double kMbarlsToPam3s = 0.1;
constexpr int kMaxwellPoints = 100;
constexpr uint32_t kLeakCap = 2048;

// Texture cells along one side for a precise size: the ceiling, except
// that a size above a whole number by rounding noise only (relative 1e-9)
// counts as that number. The rule of MolFlow's own (absent) facet code as
// its stored results show it: test case 02c stores texDimY
// 5.000000000000002 with 5 rows of cells.
int32_t texture_cells(double precise) {
    int32_t n = (int32_t)std::ceil(precise);
    if (n > 1 && precise - (n - 1) <= 1e-9 * precise) --n;
    return n;
}

struct Err : std::runtime_error {
    int code;
    Err(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw Err(MFG_ERR_DEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { double n = norm(a); return n > 0 ? (1.0 / n) * a : a; }

struct XY { double x, y; };

double interp_y(double x, const XY* t, size_t n, bool logx, bool logy, bool extrapolate) {
    if (n == 0) return 0.0;
    if (n == 1) return t[0].y;
    size_t i;
    if (x <= t[0].x) {
        if (!extrapolate) return t[0].y;
        i = 0;
    } else if (x >= t[n - 1].x) {
        if (!extrapolate) return t[n - 1].y;
        i = n - 2;
    } else {
        size_t lo = 0, hi = n - 1;
        while (hi - lo > 1) { size_t mid = (lo + hi) / 2; if (t[mid].x <= x) lo = mid; else hi = mid; }
        i = lo;
    }
    double x0 = t[i].x, x1 = t[i + 1].x, y0 = t[i].y, y1 = t[i + 1].y;
    if (logx) { x0 = std::log10(x0); x1 = std::log10(x1); x = std::log10(x); }
    if (logy) { y0 = std::log10(y0); y1 = std::log10(y1); }
    double y = (x1 == x0) ? y0 : y0 + (x - x0) * (y1 - y0) / (x1 - x0);
    return logy ? std::pow(10.0, y) : y;
}

std::vector<XY> maxwell_cdf(double gas_mass) {
    std::vector<XY> cdf;
    const double T = 1.0, Kb = 1.38E-23, R = 8.3144621;
    const double a = std::sqrt(Kb * T / (gas_mass * 1.67E-27));
    const double vp = std::sqrt(2.0 * R * T / (gas_mass / 1000.0));
    const double bin = 4.0 * vp / (double)kMaxwellPoints;
    for (int i = 0; i < kMaxwellPoints; ++i) {
        double x = (double)i * bin;
        double q = std::pow(x, 2.0) / (2.0 * std::pow(a, 2.0));
        cdf.push_back({x, 1.0 - std::exp(-q) * (q + 1.0)});
    }
    return cdf;
}

std::vector<double> parse_user_moment(const char* content, double window) {
    std::vector<double> out;
    double b, s, e;
    int nb = std::sscanf(content, "%lf,%lf,%lf", &b, &s, &e);
    if (nb == 1 && b >= 0.0) out.push_back(b);
    else if (nb == 3 && b >= 0.0 && e > b && s < (e - b) && !(s < window))
        for (double t = b; t <= e; t += s) out.push_back(t);
    return out;
}

bool is_equal(double a, double b) { return std::fabs(a - b) < 1e-10; }

std::vector<XY> integrated_desorption(const XY* par, size_t np, bool logx, bool logy, double latest) {
    std::vector<XY> ID, og;
    if (np == 0) return ID;
    size_t after;
    for (after = 0; after < np && par[after].x < latest; ++after) {}
    bool last_before = false;
    if (after >= np) { after = np - 1; last_before = true; }
    ID.push_back({0.0, 0.0});
    if (par[0].x > 0.0) og.push_back({0.0, par[0].y});
    if (last_before) {
        for (size_t i = 0; i < np; ++i) og.push_back(par[i]);
        og.push_back({latest, og.back().y});
    } else {
        for (size_t i = 0; i < after; ++i) og.push_back(par[i]);
        if (og.empty() || !is_equal(og.back().x, latest))
            og.push_back({latest, interp_y(latest, par, np, logx, logy, true)});
    }
    for (size_t i = 1; i < og.size(); ++i) {
        if (is_equal(og[i].y, og[i - 1].y)) {
            ID.push_back({og[i].x, ID.back().y + (og[i].x - og[i - 1].x) * og[i].y * kMbarlsToPam3s});
            continue;
        }
        const int steps = 10;
        double t0 = og[i - 1].x, t1 = og[i].x, dt = t1 - t0, delta = 1.0 / steps;
        double sub_dt = 0, sub_logdt = 0, lt0 = 0, lt1 = 0, ldt = 0;
        if (logx) {
            lt0 = (t0 == 0) ? -99 : std::log10(t0);
            lt1 = std::log10(t1);
            ldt = lt1 - lt0;
            sub_logdt = ldt * delta;
        } else sub_dt = dt * delta;
        double prev_v = og[i - 1].y, prev_t = t0;
        for (double frac = delta; frac < 1.0001; frac += delta) {
            double te = logx ? std::pow(10.0, lt0 + frac * ldt) : t0 + frac * dt;
            double ve = interp_y(te, og.data(), og.size(), logx, logy, true);
            double gas;
            if (!logx && !logy) gas = sub_dt * 0.5 * (prev_v + ve);
            else if (logx && !logy) gas = (ve - prev_v) * (prev_t - te) / std::log(te / prev_t) - prev_t * prev_v + te * ve;
            else if (!logx && logy) {
                double lve = ve > 0 ? std::log10(ve) : -99;
                double lpv = prev_t > 0 ? std::log10(prev_v) : -99;
                gas = sub_dt * (ve - prev_v) * std::log10(std::exp(1.0)) / (lve - lpv);
            } else {
                double lve = ve > 0 ? std::log10(ve) : -99;
                double lpv = prev_t > 0 ? std::log10(prev_v) : -99;
                double m = (lve - lpv) / sub_logdt;
                if (m != -1.0) gas = prev_v / std::pow(prev_t, m) / (m + 1.0) * (std::pow(te, m + 1.0) - std::pow(prev_t, m + 1.0));
                else gas = prev_t * prev_v * sub_logdt;
            }
            ID.push_back({te, ID.back().y + gas * kMbarlsToPam3s});
            prev_v = ve; prev_t = te;
        }
    }
    return ID;
}

double poly_area(const std::vector<double>& p) {
    size_t n = p.size() / 2;
    double a = 0;
    for (size_t i = 0, j = n - 1; i < n; j = i++) a += p[2 * j] * p[2 * i + 1] - p[2 * i] * p[2 * j + 1];
    return 0.5 * std::fabs(a);
}

double clip_area(const std::vector<double>& p, double u0, double u1, double v0, double v1) {
    std::vector<double> cur = p, nxt;
    for (int side = 0; side < 4; ++side) {
        nxt.clear();
        size_t n = cur.size() / 2;
        if (n == 0) break;
        auto inside = [&](double u, double v) {
            switch (side) { case 0: return u >= u0; case 1: return u <= u1; case 2: return v >= v0; default: return v <= v1; }
        };
        auto isect = [&](double au, double av, double bu, double bv, double& ou, double& ov) {
            double t;
            switch (side) {
                case 0: t = (u0 - au) / (bu - au); break;
                case 1: t = (u1 - au) / (bu - au); break;
                case 2: t = (v0 - av) / (bv - av); break;
                default: t = (v1 - av) / (bv - av); break;
            }
            ou = au + t * (bu - au); ov = av + t * (bv - av);
        };
        for (size_t i = 0; i < n; ++i) {
            double au = cur[2 * ((i + n - 1) % n)], av = cur[2 * ((i + n - 1) % n) + 1];
            double bu = cur[2 * i], bv = cur[2 * i + 1];
            bool ain = inside(au, av), bin = inside(bu, bv);
            double iu, iv;
            if (bin) {
                if (!ain) { isect(au, av, bu, bv, iu, iv); nxt.push_back(iu); nxt.push_back(iv); }
                nxt.push_back(bu); nxt.push_back(bv);
            } else if (ain) { isect(au, av, bu, bv, iu, iv); nxt.push_back(iu); nxt.push_back(iv); }
        }
        cur.swap(nxt);
    }
    return cur.size() >= 6 ? poly_area(cur) : 0.0;
}

int hist_bounce_bins(int rec, uint32_t bin, uint32_t max) { return rec && bin ? (int)(max / bin + 2) : 0; }
int hist_real_bins(int rec, double bin, double max) { return rec && bin > 0 ? (int)((size_t)(max / bin) + 1) : 0; }

// ------------------------------------------------------------------ BVH
struct Box { double lo[3], hi[3]; };
struct B2 { Box b; int left = -1, right = -1, first = 0, count = 0; };

inline void grow(Box& a, const Box& b) {
    for (int k = 0; k < 3; ++k) { a.lo[k] = std::min(a.lo[k], b.lo[k]); a.hi[k] = std::max(a.hi[k], b.hi[k]); }
}
inline Box empty_box() { Box b; for (int k = 0; k < 3; ++k) { b.lo[k] = 1e300; b.hi[k] = -1e300; } return b; }
inline double area(const Box& b) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = b.hi[k] - b.lo[k];
    if (d[0] < 0) return 0.0;
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

// Binned-SAH binary tree (the reference builds BVHAccel with SAH and at most
// `leaf` facets per leaf, MolflowSimulation.cpp:79).
void build_b2(std::vector<B2>& nodes, std::vector<int>& prims, const std::vector<Box>& boxes, int leaf) {
    nodes.clear();
    if (prims.empty()) return;
    struct Item { int lo, hi, node; };
    nodes.push_back(B2());
    std::vector<Item> st{{0, (int)prims.size(), 0}};
    const int NB = 16;
    while (!st.empty()) {
        Item it = st.back(); st.pop_back();
        B2 nd;
        nd.b = empty_box();
        Box cb = empty_box();
        for (int i = it.lo; i < it.hi; ++i) {
            const Box& b = boxes[prims[i]];
            grow(nd.b, b);
            for (int k = 0; k < 3; ++k) { double c = 0.5 * (b.lo[k] + b.hi[k]); cb.lo[k] = std::min(cb.lo[k], c); cb.hi[k] = std::max(cb.hi[k], c); }
        }
        int n = it.hi - it.lo, split = -1;
        if (n > leaf) {
            double best = 1e300; int bax = -1, bbin = -1;
            for (int ax = 0; ax < 3; ++ax) {
                double a0 = cb.lo[ax], a1 = cb.hi[ax];
                if (!(a1 - a0 > 0)) continue;
                int cnt[NB] = {0};
                Box bb[NB];
                for (int k = 0; k < NB; ++k) bb[k] = empty_box();
                for (int i = it.lo; i < it.hi; ++i) {
                    const Box& b = boxes[prims[i]];
                    double c = 0.5 * (b.lo[ax] + b.hi[ax]);
                    int k = std::min(NB - 1, (int)((c - a0) / (a1 - a0) * NB));
                    cnt[k]++; grow(bb[k], b);
                }
                Box lb[NB], rb[NB]; int lc[NB], rc[NB];
                Box acc = empty_box(); int ac = 0;
                for (int k = 0; k < NB; ++k) { grow(acc, bb[k]); ac += cnt[k]; lb[k] = acc; lc[k] = ac; }
                acc = empty_box(); ac = 0;
                for (int k = NB - 1; k >= 0; --k) { grow(acc, bb[k]); ac += cnt[k]; rb[k] = acc; rc[k] = ac; }
                for (int s = 1; s < NB; ++s) {
                    if (!lc[s - 1] || !rc[s]) continue;
                    double cost = lc[s - 1] * area(lb[s - 1]) + rc[s] * area(rb[s]);
                    if (cost < best) { best = cost; bax = ax; bbin = s; }
                }
            }
            if (bax >= 0) {
                double a0 = cb.lo[bax], a1 = cb.hi[bax];
                auto mid = std::partition(prims.begin() + it.lo, prims.begin() + it.hi, [&](int id) {
                    const Box& b = boxes[id];
                    double c = 0.5 * (b.lo[bax] + b.hi[bax]);
                    return std::min(NB - 1, (int)((c - a0) / (a1 - a0) * NB)) < bbin;
                });
                split = (int)(mid - prims.begin());
            }
            if (split <= it.lo || split >= it.hi) split = it.lo + n / 2;
        }
        if (split < 0) {
            nd.first = it.lo; nd.count = n;
            nodes[it.node] = nd;
        } else {
            nd.left = (int)nodes.size(); nodes.push_back(B2());
            nd.right = (int)nodes.size(); nodes.push_back(B2());
            nodes[it.node] = nd;
            st.push_back({split, it.hi, nd.right});
            st.push_back({it.lo, split, nd.left});
        }
    }
}

inline float round_down(double x) { float f = (float)x; if ((double)f > x) f = std::nextafter(f, -INFINITY); return f; }
inline float round_up(double x) { float f = (float)x; if ((double)f < x) f = std::nextafter(f, INFINITY); return f; }

// Quantised child bounds of one axis (boxes padded by `pad` and rounded
// outwards): the FP32 values the kernel computes, o + q * 2^e (one
// rounding), must enclose [lo, hi].
struct AxisQuant {
    float o;
    int e;
    static float deq(float o, int e, int q) { return o + (float)q * std::ldexp(1.0f, e); }
    // smallest step 2^e with which every child box fits in 0..255 steps
    static AxisQuant make(double node_lo, double node_hi) {
        AxisQuant a;
        a.o = round_down(node_lo);
        double ext = std::max(node_hi - (double)a.o, 1e-30);
        a.e = (int)std::ceil(std::log2(ext / 250.0));
        a.e = std::max(-126, std::min(a.e, 120));
        while (a.e < 120 && (double)deq(a.o, a.e, 255) < node_hi) ++a.e;
        return a;
    }
    uint8_t qlo(double lo) const {
        int q = (int)std::floor(((double)lo - (double)o) / std::ldexp(1.0, e));
        q = std::max(0, std::min(255, q));
        while (q > 0 && (double)deq(o, e, q) > lo) --q;
        return (uint8_t)q;
    }
    uint8_t qhi(double hi) const {
        int q = (int)std::ceil(((double)hi - (double)o) / std::ldexp(1.0, e));
        q = std::max(0, std::min(255, q));
        while (q < 255 && (double)deq(o, e, q) < hi) ++q;
        if ((double)deq(o, e, q) < hi) throw Err(MFG_ERR_INVALID, "BVH box quantisation failed");
        return (uint8_t)q;
    }
};

// A wide node while the tree is built: explicit child indices, which
// pack_node turns into the device node's implicit indexing.
struct WideNode {
    float o[3];
    int8_t e[3];
    uint8_t valid, leaf_mask;
    uint8_t qlo[3][8], qhi[3][8];   // steps 0..255 (an empty slot: lo 255, hi 0)
    int32_t child[8];               // inner: node index; leaf: first leaf record
    uint8_t count[8];               // facets of a leaf child (0 for inner/empty)
};

// Collapses the binary tree into nodes of up to `width` children (4 or 8):
// the child with the largest surface is opened until the node is full. A
// node's leaf records are appended before its inner children are built, so
// they are consecutive in slot order.
struct Collapse {
    const std::vector<B2>& b2;
    const std::vector<int>& prims;
    std::vector<WideNode>& out;
    std::vector<int32_t>& leaf_ids;
    int32_t node_base, leaf_base;
    double pad;
    int width = MFG_BVH_WIDTH;
    int max_depth = 0;
    int build(int n2, int depth) {
        max_depth = std::max(max_depth, depth);
        int ch[8]; int nc = 0;
        const B2& root = b2[n2];
        if (root.left < 0) { ch[nc++] = n2; }
        else { ch[nc++] = root.left; ch[nc++] = root.right; }
        while (nc < width) {
            int best = -1; double ba = -1;
            for (int i = 0; i < nc; ++i) if (b2[ch[i]].left >= 0 && area(b2[ch[i]].b) > ba) { ba = area(b2[ch[i]].b); best = i; }
            if (best < 0) break;
            int c = ch[best];
            ch[best] = b2[c].left;
            ch[nc++] = b2[c].right;
        }
        int idx = (int)out.size();
        out.push_back(WideNode());
        WideNode nd;
        std::memset(&nd, 0, sizeof(nd));
        Box all = empty_box();
        for (int c = 0; c < nc; ++c) grow(all, b2[ch[c]].b);
        AxisQuant q[3];
        for (int k = 0; k < 3; ++k) q[k] = AxisQuant::make(all.lo[k] - pad, all.hi[k] + pad);
        for (int k = 0; k < 3; ++k) { nd.o[k] = q[k].o; nd.e[k] = (int8_t)q[k].e; }
        for (int c = 0; c < 8; ++c) {
            if (c >= nc) {   // empty slot: an inverted box, never hit
                for (int k = 0; k < 3; ++k) { nd.qlo[k][c] = 255; nd.qhi[k][c] = 0; }
                continue;
            }
            const B2& cn = b2[ch[c]];
            nd.valid |= (uint8_t)(1u << c);
            for (int k = 0; k < 3; ++k) {
                nd.qlo[k][c] = q[k].qlo(cn.b.lo[k] - pad);
                nd.qhi[k][c] = q[k].qhi(cn.b.hi[k] + pad);
            }
            if (cn.left < 0) {   // leaf: its records now, in slot order
                if (cn.count < 1 || cn.count > 16) throw Err(MFG_ERR_INVALID, "BVH leaf outside 1..16 facets");
                nd.child[c] = (int32_t)leaf_ids.size() + leaf_base;
                for (int i = 0; i < cn.count; ++i) leaf_ids.push_back(prims[cn.first + i]);
                nd.count[c] = (uint8_t)cn.count;
                nd.leaf_mask |= (uint8_t)(1u << c);
            }
        }
        for (int c = 0; c < nc; ++c)
            if (b2[ch[c]].left >= 0) nd.child[c] = build(ch[c], depth + 1) + node_base;
        out[idx] = nd;
        return idx;
    }
};

// FP16 bit pattern of an integer 0..255 (exact)
inline uint16_t half_of_step(uint8_t q) {
    if (q == 0) return 0;
    int e = 31 - __builtin_clz((unsigned)q);          // 2^e <= q < 2^(e+1)
    return (uint16_t)(((e + 15) << 10) | (((unsigned)q << (10 - e)) & 0x3ffu));
}

// The device node of a wide node: inner children must be consecutive nodes
// and leaf records consecutive in slot order (breadth-first numbering and
// Collapse's record order give both; checked here).
BvhNode pack_node(const WideNode& w) {
    BvhNode p;
    std::memset(&p, 0, sizeof(p));
    p.ox = w.o[0]; p.oy = w.o[1]; p.oz = w.o[2];
    p.ex = w.e[0]; p.ey = w.e[1]; p.ez = w.e[2];
    p.leaf_mask = w.leaf_mask;
    p.inner_mask = (uint8_t)(w.valid & ~w.leaf_mask);
    int inner = 0, leafs = 0;
    for (int c = 0; c < 8; ++c) {
        for (int k = 0; k < 3; ++k) { p.qlo[k][c] = half_of_step(w.qlo[k][c]); p.qhi[k][c] = half_of_step(w.qhi[k][c]); }
        if (!(w.valid & (1u << c))) continue;
        if (w.leaf_mask & (1u << c)) {
            if (leafs == 0) p.leaf_base = w.child[c];
            if (w.child[c] != p.leaf_base + leafs) throw Err(MFG_ERR_INVALID, "BVH leaf records of a node not consecutive");
            if (w.count[c] < 1 || w.count[c] > 16) throw Err(MFG_ERR_INVALID, "BVH leaf outside 1..16 facets");
            p.counts |= (uint32_t)(w.count[c] - 1) << (4 * c);
            leafs += w.count[c];
        } else {
            if (inner == 0) p.node_base = w.child[c];
            if (w.child[c] != p.node_base + inner) throw Err(MFG_ERR_INVALID, "BVH inner children of a node not consecutive");
            ++inner;
        }
    }
    return p;
}

// ------------------------------------------------------------------ host model
struct Param { std::vector<XY> pts; bool logx, logy; };

struct HostModel {
    mfg_settings s{};
    uint32_t n_struct = 1;
    std::vector<DFacet> facets;
    std::vector<DPhys> phys;
    std::vector<double> poly;
    std::vector<Box> boxes;
    std::vector<double> tex_inc;
    std::vector<Param> params;
    std::vector<std::vector<XY>> IDs;
    std::vector<XY> maxwell;
    std::vector<double> mom_start, mom_end;
    std::vector<double> src_cum;
    std::vector<int32_t> src_facet;   // local
    std::vector<double> og_cdf;
    std::vector<double> am_src;       // DES_ANGLEMAP blocks (variant-local offsets in DPhys)
    double latest = 0, total_des = 0, final_rate = 0, final_rate_pam3s = 0, c_factor = 1.0;
    uint32_t n_slots = 1;
    uint64_t n_tex = 0, n_dir = 0, n_prof = 0, n_am = 0, n_fh = 0;
    uint32_t gh_b = 0, gh_d = 0, gh_t = 0;
    uint32_t features = 0;
    std::vector<std::vector<int>> struct_facets;
    double extent = 1.0;
};

// DES_ANGLEMAP source table of one facet (InitializeAngleMap,
// src/Simulation/MolflowSimFacet.cpp:92-222): phi line sums, theta CDF at
// line midpoints, per-line midpoint phi CDFs and normalised phi pdfs; the
// lower and higher theta parts are stored as one run of L + H lines.
void angle_source_block(std::vector<double>& out, const mfg_facet_desc& d, const uint64_t* pdf, uint64_t n,
                        uint32_t f) {
    const int W = (int)d.am_phi_width, L = (int)d.am_theta_lower_res, H = (int)d.am_theta_higher_res, T = L + H;
    if (W <= 0 || T <= 0 || n != (uint64_t)T * (uint64_t)W)
        throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + ": DES_ANGLEMAP map does not match its shape");
    std::vector<uint64_t> sum(T, 0);
    uint64_t lower = 0, total;
    for (int t = 0; t < L; ++t) {
        for (int p = 0; p < W; ++p) sum[t] += pdf[(size_t)t * W + p];
        lower += sum[t];
    }
    total = lower;
    for (int t = L; t < T; ++t) {
        for (int p = 0; p < W; ++p) sum[t] += pdf[(size_t)t * W + p];
        total += sum[t];
    }
    if (total == 0) throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + ": DES_ANGLEMAP map is all zero");
    const double ratio = (double)lower / (double)total;
    const size_t base = out.size();
    out.resize(base + 6 + 2 * (size_t)T + 2 * (size_t)T * W);
    double* b = out.data() + base;
    b[0] = W; b[1] = L; b[2] = H; b[3] = d.am_theta_limit; b[4] = ratio; b[5] = (double)total;
    double *line = b + 6, *cdf = line + T, *pc = cdf + T, *pp = pc + (size_t)T * W;
    const double tn = 1.0 / (double)total;
    for (int t = 0; t < T; ++t) {
        line[t] = (double)sum[t];
        if (t == 0) cdf[t] = 0.5 * (double)sum[0] * tn;
        else if (t == L) cdf[t] = ratio + 0.5 * (double)sum[L] * tn;
        else cdf[t] = cdf[t - 1] + (double)(sum[t - 1] + sum[t]) * 0.5 * tn;
        const double pn = 1.0 / (double)sum[t];
        for (int p = 0; p < W; ++p) {
            const size_t i = (size_t)t * W + p;
            if (sum[t] == 0) {
                pc[i] = (0.5 + (double)p) / (double)W;
                pp[i] = 1.0 / (double)W;
            } else {
                if (p == 0) pc[i] = 0.5 * (double)pdf[(size_t)t * W] * pn;
                else pc[i] = pc[i - 1] + (double)(pdf[i - 1] + pdf[i]) * 0.5 * pn;
                pp[i] = (double)pdf[i] * pn;
            }
        }
    }
}

// Slab index of a polygon of more than 4 vertices, stored right after its
// normalised (u,v) list in the polygon array. The crossing-number test
// toggles on every edge with exactly one end above v; sorted by v, the
// distinct vertex heights w_0 < ... < w_{m-1} cut [w_0, w_{m-1}) into slabs,
// and all v of one slab [w_k, w_{k+1}) see the same such edges: those with
// min(v_i, v_j) <= w_k and max(v_i, v_j) >= w_{k+1}. A lookup finds v's slab
// and evaluates the crossing predicate of the polygon loop, operand for
// operand, on those edges only (two for a convex polygon); the parity of the
// toggles does not depend on their order. Layout, in doubles: m (bits), the
// m heights, m edge offsets (bits; slab k is [off_k, off_{k+1})), one pad,
// then four doubles (u_i, v_i, u_j, v_j) per listed edge, j being i's
// predecessor as in the loop; padded to an even count (the array holds
// (u,v) pairs).
void append_slabs(std::vector<double>& out, const std::vector<double>& pl) {
    const size_t n = pl.size() / 2;
    std::vector<double> w(n);
    for (size_t i = 0; i < n; ++i) w[i] = pl[2 * i + 1];
    std::sort(w.begin(), w.end());
    w.erase(std::unique(w.begin(), w.end()), w.end());
    const size_t m = w.size();
    auto bits = [](uint64_t x) { double d; std::memcpy(&d, &x, sizeof d); return d; };
    std::vector<double> edges;
    std::vector<uint64_t> off(m, 0);
    for (size_t k = 0; k + 1 < m; ++k) {
        off[k] = edges.size() / 4;
        for (size_t i = 0; i < n; ++i) {
            const size_t j = i == 0 ? n - 1 : i - 1;
            const double vi = pl[2 * i + 1], vj = pl[2 * j + 1];
            if (std::min(vi, vj) <= w[k] && std::max(vi, vj) >= w[k + 1]) {
                edges.push_back(pl[2 * i]); edges.push_back(vi);
                edges.push_back(pl[2 * j]); edges.push_back(vj);
            }
        }
    }
    off[m - 1] = edges.size() / 4;
    const size_t start = out.size();
    out.push_back(bits(m));
    out.insert(out.end(), w.begin(), w.end());
    for (size_t k = 0; k < m; ++k) out.push_back(bits(off[k]));
    out.push_back(0.0);   // the edges start on a 16 B boundary
    out.insert(out.end(), edges.begin(), edges.end());
    if ((out.size() - start) % 2) out.push_back(0.0);
}

void prepare(HostModel& M, const mfg_model_view* v) {
    if (!v) throw Err(MFG_ERR_INVALID, "null model view");
    if (v->struct_size != sizeof(mfg_model_view)) throw Err(MFG_ERR_INVALID, "mfg_model_view size mismatch");
    if (v->abi_version != MFG_ABI_VERSION) throw Err(MFG_ERR_INVALID, "ABI version mismatch");
    if (!v->settings || !v->facets || !v->vertices || !v->facet_vertex_offset || !v->facet_vertex_index)
        throw Err(MFG_ERR_INVALID, "model view has null arrays");
    if (v->n_facets == 0) throw Err(MFG_ERR_INVALID, "model has no facets");
    M.s = *v->settings;
    if (!(M.s.gas_mass > 0)) throw Err(MFG_ERR_INVALID, "gas mass must be positive");
    M.n_struct = std::max<uint32_t>(1, v->n_structures);
    for (uint32_t p = 0; p < v->n_params; ++p) {
        Param P;
        for (uint32_t k = v->param_offset[p]; k < v->param_offset[p + 1]; ++k) P.pts.push_back({v->param_x[k], v->param_y[k]});
        P.logx = v->param_log_x ? v->param_log_x[p] : 0;
        P.logy = v->param_log_y ? v->param_log_y[p] : 0;
        M.params.push_back(P);
    }
    std::vector<std::pair<double, double>> mom;
    for (uint32_t i = 0; i < v->n_moments; ++i) mom.push_back({v->moment_time[i], v->moment_window[i]});
    std::sort(mom.begin(), mom.end(), [](auto a, auto b) { return a.first - .5 * a.second < b.first - .5 * b.second; });
    for (auto& mm : mom) { M.mom_start.push_back(mm.first - .5 * mm.second); M.mom_end.push_back(mm.first + .5 * mm.second); }
    M.latest = mom.empty() ? M.s.time_window * .5 : mom.back().first + .5 * mom.back().second;
    M.n_slots = 1 + (uint32_t)mom.size();
    M.maxwell = maxwell_cdf(M.s.gas_mass);
    M.c_factor = M.s.use_maxwell ? 1.0 : 1.1781;
    uint32_t feat = 0;
    if (!mom.empty()) feat |= FEAT_MOMENTS;
    if (!M.s.calc_constant_flow || M.s.enable_decay || M.s.scatter_enabled) feat |= FEAT_VOLUME;
    if (M.s.low_flux_mode) feat |= FEAT_LOWFLUX;
    if (M.s.enable_force) feat |= FEAT_FORCE;
    if (M.s.log_enabled) feat |= FEAT_LOG;

    uint32_t nF = v->n_facets;
    M.facets.resize(nF);
    M.phys.resize(nF);
    M.boxes.resize(nF);
    std::vector<int> id_of_param(M.params.size(), -1);
    double glo[3] = {1e300, 1e300, 1e300}, ghi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t f = 0; f < nF; ++f) {
        const mfg_facet_desc& d = v->facets[f];
        DFacet& F = M.facets[f];
        DPhys& ph = M.phys[f];
        std::memset(&F, 0, sizeof(F));
        std::memset(&ph, 0, sizeof(ph));
        uint32_t a = v->facet_vertex_offset[f], b = v->facet_vertex_offset[f + 1];
        if (b < a + 3) throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + " has fewer than 3 vertices");
        std::vector<V3> P;
        for (uint32_t k = a; k < b; ++k) {
            uint32_t vi = v->facet_vertex_index[k];
            if (vi >= v->n_vertices) throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + " refers to a missing vertex");
            P.push_back({v->vertices[3 * vi], v->vertices[3 * vi + 1], v->vertices[3 * vi + 2]});
        }
        size_t n = P.size();
        // frame: Newell normal, nU along the first edge, nV = N x nU
        V3 N = {0, 0, 0};
        for (size_t i = 0; i < n; ++i) {
            const V3& p = P[i]; const V3& q = P[(i + 1) % n];
            N.x += (p.y - q.y) * (p.z + q.z);
            N.y += (p.z - q.z) * (p.x + q.x);
            N.z += (p.x - q.x) * (p.y + q.y);
        }
        N = unit(N);
        V3 nU = unit(P[1] - P[0]);
        V3 nV = unit(cross(N, nU));
        double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
        std::vector<double> pu(n), pv(n);
        for (size_t i = 0; i < n; ++i) {
            V3 r = P[i] - P[0];
            pu[i] = dot(r, nU); pv[i] = dot(r, nV);
            umin = std::min(umin, pu[i]); umax = std::max(umax, pu[i]);
            vmin = std::min(vmin, pv[i]); vmax = std::max(vmax, pv[i]);
        }
        if (!(umax > umin) || !(vmax > vmin)) throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + " is degenerate");
        V3 O = P[0] + umin * nU + vmin * nV;
        V3 U = (umax - umin) * nU, V = (vmax - vmin) * nV;
        F.O[0] = O.x; F.O[1] = O.y; F.O[2] = O.z;
        F.N[0] = N.x; F.N[1] = N.y; F.N[2] = N.z;
        F.U[0] = U.x; F.U[1] = U.y; F.U[2] = U.z;
        F.V[0] = V.x; F.V[1] = V.y; F.V[2] = V.z;
        F.invU2 = 1.0 / dot(U, U);
        F.invV2 = 1.0 / dot(V, V);
        F.poly_off = (uint32_t)(M.poly.size() / 2);
        F.poly_n = (uint32_t)n;
        std::vector<double> pl(2 * n);
        for (size_t i = 0; i < n; ++i) {
            pl[2 * i] = (pu[i] - umin) / (umax - umin);
            pl[2 * i + 1] = (pv[i] - vmin) / (vmax - vmin);
        }
        M.poly.insert(M.poly.end(), pl.begin(), pl.end());
        if (n > 4) append_slabs(M.poly, pl);
        double lu = norm(U), lv = norm(V);
        Box bx = empty_box();
        V3 c = {0, 0, 0};
        for (auto& p : P) {
            c = c + p;
            double pp[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k) {
                bx.lo[k] = std::min(bx.lo[k], pp[k]); bx.hi[k] = std::max(bx.hi[k], pp[k]);
                glo[k] = std::min(glo[k], pp[k]); ghi[k] = std::max(ghi[k], pp[k]);
            }
        }
        c = (1.0 / n) * c;
        M.boxes[f] = bx;
        for (int k = 0; k < 3; ++k) { ph.nU[k] = k == 0 ? nU.x : k == 1 ? nU.y : nU.z; ph.nV[k] = k == 0 ? nV.x : k == 1 ? nV.y : nV.z; }
        ph.centre[0] = c.x; ph.centre[1] = c.y; ph.centre[2] = c.z;
        double opacity = std::min(1.0, std::max(0.0, d.opacity));
        int32_t sp = d.sticking_param, op = d.opacity_param, tp = d.temperature_param, gp = d.outgassing_param;
        if (d.super_dest) { op = -1; sp = -1; }
        for (int32_t pid : {sp, op, tp, gp})
            if (pid >= (int32_t)M.params.size()) throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + " refers to a missing parameter");
        ph.sticking = d.sticking; ph.opacity = opacity; ph.temperature = d.temperature;
        ph.sqrt_temp = std::sqrt(d.temperature); ph.accomodation = d.accomodation;
        ph.diffuse = d.diffuse_part; ph.specular = d.specular_part; ph.cos_exp = d.cosine_exponent;
        ph.des_exp = d.desorb_exponent; ph.sojourn_freq = d.sojourn_freq; ph.sojourn_e = d.sojourn_e;
        ph.max_speed = 4.0 * std::sqrt(2.0 * 8.31 * d.temperature / 0.001 / M.s.gas_mass);
        ph.sticking_param = sp; ph.opacity_param = op; ph.temperature_param = tp; ph.outgassing_param = gp;
        ph.super_idx = d.super_idx; ph.super_dest = d.super_dest; ph.teleport_dest = d.teleport_dest;
        ph.profile_type = d.profile_type; ph.desorb_type = d.desorb_type;
        ph.local_id = (int32_t)f;
        ph.tex_off = ph.dir_off = -1; ph.prof_off = ph.am_off = ph.fh_off = -1;
        ph.id_index = -1; ph.adj_group = -1; ph.adj_target = -1; ph.og_off = -1;
        uint32_t fl = 0;
        if (d.is_2sided) fl |= FF_2SIDED;
        if (op >= 0) { fl |= FF_TD_OPACITY; feat |= FEAT_TRANSPARENT | FEAT_TDPARAM; }
        else if (opacity >= 1.0) fl |= FF_OPAQUE;
        else { if (opacity <= 0.0) fl |= FF_TRANSP; feat |= FEAT_TRANSPARENT; }
        if (sp >= 0 || tp >= 0 || gp >= 0) feat |= FEAT_TDPARAM;
        if (d.is_moving) { fl |= FF_MOVING; if (M.s.motion_type) feat |= FEAT_MOVING; }
        if (d.enable_sojourn) { fl |= FF_SOJOURN; feat |= FEAT_SOJOURN; }
        if (d.teleport_dest != 0) { fl |= FF_TELEPORT; feat |= FEAT_TELEPORT; }
        if (d.super_dest != 0) {
            if (d.super_dest - 1 >= (int)M.n_struct || d.super_dest < 0) throw Err(MFG_ERR_INVALID, "invalid structure link on facet " + std::to_string(f));
            fl |= FF_LINK; feat |= FEAT_TELEPORT;
        }
        // structure -1 means "in every structure"; anything else must name one
        if (d.super_idx < -1 || d.super_idx >= (int)M.n_struct)
            throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + ": structure index " + std::to_string(d.super_idx) +
                                           " out of range [-1, " + std::to_string(M.n_struct) + ")");
        // teleport destination + 1, 0 none, -1 back to the previous destination
        if (d.teleport_dest < -1 || d.teleport_dest > (int32_t)nF)
            throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + ": teleport destination " +
                                           std::to_string(d.teleport_dest) + " out of range [-1, " + std::to_string(nF) + "]");
        if (d.diffuse_part > 0.999999) fl |= FF_DIFFUSE;
        if (d.accomodation > 0.9999) fl |= FF_FULL_ACC;
        if (d.count_des) fl |= FF_COUNT_DES;
        if (d.count_abs) fl |= FF_COUNT_ABS;
        if (d.count_refl) fl |= FF_COUNT_REFL;
        if (d.count_trans) fl |= FF_COUNT_TRANS;
        if (d.count_dir) fl |= FF_COUNT_DIR;
        ph.am_src_off = -1;
        ph.hot_idx = -1;
        if (d.desorb_type == MFG_DES_ANGLEMAP) {
            if (!v->am_pdf_offset || !v->am_pdf || v->am_pdf_offset[f + 1] == v->am_pdf_offset[f])
                throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + ": DES_ANGLEMAP without a recorded angle map");
            ph.am_src_off = (int64_t)M.am_src.size();
            angle_source_block(M.am_src, d, v->am_pdf + v->am_pdf_offset[f], v->am_pdf_offset[f + 1] - v->am_pdf_offset[f], f);
            feat |= FEAT_AMSRC;
        }
        // textures
        if (d.is_textured && d.tex_width_precise > 0 && d.tex_height_precise > 0) {
            fl |= FF_TEXTURE;
            ph.tex_w = texture_cells(d.tex_width_precise);
            ph.tex_h = texture_cells(d.tex_height_precise);
            ph.tex_wp = d.tex_width_precise; ph.tex_hp = d.tex_height_precise;
            ph.tex_off = (int64_t)M.n_tex;
            if (d.count_dir) { ph.dir_off = (int64_t)M.n_dir; M.n_dir += (uint64_t)ph.tex_w * ph.tex_h; }
            M.n_tex += (uint64_t)ph.tex_w * ph.tex_h;
            double rw = lu / d.tex_width_precise, rh = lv / d.tex_height_precise;
            double full = rw * rh, two = d.is_2sided ? 2.0 : 1.0;
            for (int j = 0; j < ph.tex_h; ++j)
                for (int i = 0; i < ph.tex_w; ++i) {
                    double u0 = i / d.tex_width_precise, u1 = (i + 1) / d.tex_width_precise;
                    double v0 = j / d.tex_height_precise, v1 = (j + 1) / d.tex_height_precise;
                    double a2 = clip_area(pl, u0, u1, v0, v1) * lu * lv;
                    double inc;
                    if (a2 >= full * (1.0 - 1e-9)) inc = 1.0 / (full * two);
                    else if (a2 > 0) inc = 1.0 / (a2 * two);
                    else inc = 0.0;
                    M.tex_inc.push_back(inc);
                }
        }
        if (d.profile_type != MFG_PROFILE_NONE) { fl |= FF_PROFILE; ph.prof_off = (int32_t)M.n_prof; M.n_prof += MFG_PROFILE_SIZE; }
        if (d.am_record) {
            fl |= FF_ANGLEMAP;
            ph.am_off = (int32_t)M.n_am; ph.am_phi = d.am_phi_width; ph.am_lo = d.am_theta_lower_res; ph.am_hi = d.am_theta_higher_res;
            ph.am_theta_limit = d.am_theta_limit;
            M.n_am += (uint64_t)d.am_phi_width * (d.am_theta_lower_res + d.am_theta_higher_res);
        }
        uint32_t fb = hist_bounce_bins(d.fh_record_bounce, d.fh_bounce_binsize, d.fh_bounce_max);
        uint32_t fd = hist_real_bins(d.fh_record_distance, d.fh_distance_binsize, d.fh_distance_max);
        uint32_t ft = hist_real_bins(d.fh_record_time, d.fh_time_binsize, d.fh_time_max);
        if (fb + fd + ft) {
            fl |= FF_FHIST; feat |= FEAT_HIST;
            ph.fh_off = (int32_t)M.n_fh; M.n_fh += fb + fd + ft;
            ph.fh_nb_bounce = fb; ph.fh_nb_dist = fd; ph.fh_nb_time = ft;
            ph.fh_bounce_bin = d.fh_bounce_binsize; ph.fh_dist_bin = d.fh_distance_binsize; ph.fh_time_bin = d.fh_time_binsize;
        }
        // outgassing map
        if (d.use_outgassing_map && v->og_map_offset && v->og_map_values) {
            uint64_t o0 = v->og_map_offset[f], o1 = v->og_map_offset[f + 1];
            if (o1 - o0 != (uint64_t)d.og_map_width * d.og_map_height || o1 == o0)
                throw Err(MFG_ERR_INVALID, "outgassing map size mismatch on facet " + std::to_string(f));
            fl |= FF_OGMAP; feat |= FEAT_OGMAP;
            ph.og_off = (int64_t)M.og_cdf.size();
            double acc = 0;
            for (uint64_t k = o0; k < o1; ++k) { acc += v->og_map_values[k]; M.og_cdf.push_back(acc); }
            ph.og_w = d.og_map_width; ph.og_h = d.og_map_height;
            ph.og_wp = lu * d.og_ratio_u; ph.og_hp = lv * d.og_ratio_v;
        }
        if (gp >= 0) {
            if (id_of_param[gp] < 0) {
                id_of_param[gp] = (int)M.IDs.size();
                const Param& P2 = M.params[gp];
                M.IDs.push_back(integrated_desorption(P2.pts.data(), P2.pts.size(), P2.logx, P2.logy, M.latest));
            }
            ph.id_index = id_of_param[gp];
        }
        ph.flags = fl;
        F.flags = fl;
        F.opacity = (float)opacity;
    }
    M.extent = std::max({ghi[0] - glo[0], ghi[1] - glo[1], ghi[2] - glo[2], 1e-6});
    // source table (CalcTotalOutgassing order, MolflowSimGeom.cpp:290-335)
    double acc = 0;
    for (uint32_t f = 0; f < nF; ++f) {
        const mfg_facet_desc& d = v->facets[f];
        const DPhys& ph = M.phys[f];
        if (d.desorb_type == MFG_DES_NONE) continue;
        double kT = kBoltz * d.temperature, w;
        if (ph.flags & FF_OGMAP) {
            double sum = M.og_cdf[ph.og_off + (uint64_t)ph.og_w * ph.og_h - 1];
            M.total_des += M.latest * sum / kT;
            M.final_rate += sum / kT;
            M.final_rate_pam3s += sum;
            if (!(d.total_outgassing > 0)) continue;
            w = M.latest * d.total_outgassing / kT;
        } else if (ph.outgassing_param < 0) {
            w = M.latest * d.outgassing / kT;
            M.total_des += w;
            M.final_rate += d.outgassing / kT;
            M.final_rate_pam3s += d.outgassing;
        } else {
            w = M.IDs[ph.id_index].back().y / kT;
            M.total_des += w;
            const double last = M.params[ph.outgassing_param].pts.back().y * kMbarlsToPam3s;
            M.final_rate += last / kT;
            M.final_rate_pam3s += last;
        }
        if (w <= 0) continue;
        if (d.super_idx == -1) throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + " is in all structures and can't desorb");
        acc += w;
        M.src_cum.push_back(acc);
        M.src_facet.push_back((int32_t)f);
    }
    if (M.src_cum.empty()) throw Err(MFG_ERR_DESORPTION, "no desorbing facet (total outgassing is zero)");
    M.gh_b = hist_bounce_bins(M.s.gh_record_bounce, M.s.gh_bounce_binsize, M.s.gh_bounce_max);
    M.gh_d = hist_real_bins(M.s.gh_record_distance, M.s.gh_distance_binsize, M.s.gh_distance_max);
    M.gh_t = hist_real_bins(M.s.gh_record_time, M.s.gh_time_binsize, M.s.gh_time_max);
    if (M.gh_b + M.gh_d + M.gh_t) feat |= FEAT_HIST;
    M.struct_facets.assign(M.n_struct, {});
    for (uint32_t f = 0; f < nF; ++f) {
        int si = v->facets[f].super_idx;
        if (si < 0) { for (auto& sf : M.struct_facets) sf.push_back((int)f); }
        else M.struct_facets[si].push_back((int)f);
    }
    M.features = feat;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    void alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) { p = nullptr; n = 0; throw Err(MFG_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"); }
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct VariantLayout {
    uint32_t n_facets, n_slots, n_struct;
    uint64_t fc, force, tex, n_tex, dir, n_dir, prof, n_prof, am, n_am, gh, n_gh, fh, n_fh;
    std::vector<int64_t> off_tex, off_dir, off_prof, off_am, off_fh;
    uint32_t log_limit;
    bool enable_force;
};

}  // namespace

struct mfg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int n_cu = 0;
    uint32_t block = 256, blocks_per_cu = 0;
    int waves = 0;                       // occupancy bound of the trace kernel, 0 = default
    bool carry_on = false;               // keep particles in flight between launches
    DevBuf<CarryRec> carry;              // two halves of carry_cap records
    uint64_t carry_cap = 0;
    uint64_t in_flight = 0;              // particles resident in the carry buffer
    std::vector<std::unique_ptr<HostModel>> models;
    std::vector<VariantLayout> layouts;
    std::vector<DVariant> dvars;
    std::vector<uint64_t> pid_base;
    uint64_t seed = 42;
    uint32_t features = 0;
    bool loaded = false, accel = false;
    uint32_t stack_size = 0;             // traversal stack entries a walk may need
    DevBuf<int32_t> stack_spill;         // per-lane entries beyond the LDS part
    int bvh_width = MFG_BVH_WIDTH, leaf_size = 2;   // children per BVH node, facets per leaf
    // BVH nodes each block keeps in LDS (top of the breadth-first array):
    // 16 KB, which leaves 256-thread blocks their occupancy up to 5 waves
    // per SIMD
    uint32_t node_cache = 128;
    size_t lds_per_block = 65536;        // the device's LDS limit for one block
    // device data
    DevBuf<DFacet> facets;
    DevBuf<DPhys> phys;
    DevBuf<double> poly, src_cum, og_cdf, tab_x, tab_y, mom_start, mom_end, tex_inc, am_src;
    DevBuf<int32_t> src_facet, roots;
    DevBuf<LeafFacet> leaves;
    DevBuf<LeafUV> leaf_uv;
    std::vector<DPhys> h_phys;           // host copy of the uploaded physics records
    std::vector<int32_t> hot;            // hot facets (global index), their counters summed in LDS
    DevBuf<uint64_t> hot_fc;
    std::vector<DFacet> h_facets;        // host copies of the concatenated facets and polygons
    std::vector<double> h_poly;
    DevBuf<uint32_t> guides;
    DevBuf<BvhNode> nodes;
    DevBuf<DVariant> variants;
    DevBuf<uint64_t> prefix;
    DevBuf<DTable> params, ids;
    // tallies
    DevBuf<mfg_facet_counters> fc;
    DevBuf<double> force, gh, fh;
    DevBuf<mfg_texture_cell> tex;
    DevBuf<mfg_direction_cell> dir;
    DevBuf<mfg_profile_slice> prof;
    DevBuf<unsigned long long> am, counters;   // counters: next_particle, log_count, leak_count, time-limit flag, carry control[3]
    DevBuf<DGlobal> glob;
    DevBuf<mfg_log_entry> log;
    DevBuf<double> leaks;
    uint32_t log_cap = 0;
    // host-mapped stop flag
    int32_t* stop_host = nullptr;
    double time_limit_s = 60.0;          // wall-clock limit of one trace launch
    uint64_t clock_khz = 0;              // device wall-clock rate
    int32_t* stop_dev = nullptr;
    // pinned run block (variants + prefix), read by graph replays
    DVariant* pin_vars = nullptr;
    uint64_t* pin_prefix = nullptr;
    size_t pin_cap = 0;
    // trace
    uint32_t trace_n = 0, trace_k = 0;
    DevBuf<int32_t> trace_f, trace_kind;
    // adjoint
    uint32_t adj_groups = 0, adj_targets = 0;
    DevBuf<double> adj_j0, adj_jac, adj_s, adj_up, adj_grad;
    DevBuf<uint32_t> adj_counts;
    DevBuf<int32_t> adj_seg_t;           // per lane: target scores of the current segment (cutoff term)
    DevBuf<double> adj_seg_x;
    std::vector<int32_t> adj_group_of, adj_target_list;
    uint64_t adj_des = 0;
    bool adj_s_ok = false;               // every group sticking inside (0,1)
    // graph
    bool graph_on = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<uint64_t> graph_key;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // staging array of mfg_intersect_batch_static: the device address of
    // the library's static `scratch` array on this context's device
    StaticRays* scratch = nullptr;
    std::vector<int32_t> h_roots;        // host copy of the BVH roots
};

namespace {

int fail(mfg_context* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <typename F>
int guarded(mfg_context* c, F&& fn) {
    if (!c) return MFG_ERR_INVALID;
    try {
        HIPCHK(hipSetDevice(c->device));
        fn();
        c->err.clear();
        return MFG_OK;
    } catch (const Err& e) {
        return fail(c, e.code, e.what());
    } catch (const std::exception& e) {
        return fail(c, MFG_ERR_INVALID, e.what());
    }
}

void reset_graph(mfg_context* c) {
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    c->graph_key.clear();
}

// Marks the hot facets in the physics records and stores the index of each
// one's slot-0 counter record.
void set_hot_facets(mfg_context* c, const std::vector<int32_t>& hot, bool upload) {
    for (int32_t f : c->hot) c->h_phys[(size_t)f].hot_idx = -1;
    c->hot = hot;
    std::vector<uint64_t> idx;
    for (size_t k = 0; k < hot.size(); ++k) {
        DPhys& ph = c->h_phys[(size_t)hot[k]];
        ph.hot_idx = (int32_t)k;
        size_t vi = 0;
        while (vi + 1 < c->dvars.size() && c->dvars[vi + 1].facet_base <= (uint32_t)hot[k]) ++vi;
        idx.push_back(c->dvars[vi].fc_base + (uint64_t)(uint32_t)ph.local_id);
    }
    c->hot_fc.upload(idx, c->stream);
    if (upload) c->phys.upload(c->h_phys, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    reset_graph(c);
}

void upload_models(mfg_context* c) {
    reset_graph(c);
    std::vector<DFacet> facets;
    std::vector<DPhys> phys;
    std::vector<double> poly, src_cum, og_cdf, tab_x, tab_y, mom_s, mom_e, tex_inc, am_src;
    std::vector<int32_t> src_facet;
    std::vector<uint32_t> guides;
    std::vector<DTable> params, ids;
    c->dvars.clear();
    c->layouts.clear();
    uint64_t fc = 0, force = 0, tex = 0, dir = 0, prof = 0, am = 0, gh = 0, fh = 0;
    uint32_t feat = 0, log_cap = 0;
    for (size_t vi = 0; vi < c->models.size(); ++vi) {
        HostModel& M = *c->models[vi];
        DVariant dv;
        std::memset(&dv, 0, sizeof(dv));
        uint32_t fbase = (uint32_t)facets.size(), pbase = (uint32_t)(poly.size() / 2);
        dv.facet_base = fbase;
        dv.n_facets = (uint32_t)M.facets.size();
        dv.n_struct = M.n_struct;
        for (size_t f = 0; f < M.facets.size(); ++f) {
            DFacet F = M.facets[f];
            F.poly_off += pbase;
            facets.push_back(F);
            DPhys ph = M.phys[f];
            if (ph.og_off >= 0) ph.og_off += (int64_t)og_cdf.size();
            if (ph.am_src_off >= 0) ph.am_src_off += (int64_t)am_src.size();
            phys.push_back(ph);
        }
        poly.insert(poly.end(), M.poly.begin(), M.poly.end());
        og_cdf.insert(og_cdf.end(), M.og_cdf.begin(), M.og_cdf.end());
        am_src.insert(am_src.end(), M.am_src.begin(), M.am_src.end());
        dv.src_base = (uint32_t)src_cum.size();
        dv.src_n = (uint32_t)M.src_cum.size();
        src_cum.insert(src_cum.end(), M.src_cum.begin(), M.src_cum.end());
        for (int32_t s : M.src_facet) src_facet.push_back(s + (int32_t)fbase);
        dv.total_des = M.src_cum.back();
        dv.latest = M.latest; dv.gas_mass = M.s.gas_mass; dv.half_life = M.s.half_life;
        dv.low_flux_cutoff = M.s.low_flux_cutoff; dv.c_factor = M.c_factor;
        dv.scatter_mr = M.s.scatter_mass_ratio; dv.scatter_mfp = M.s.scatter_mean_free_path_cm;
        dv.scatter_cut = M.s.scatter_cutoff_speed; dv.scatter_rho = M.s.scatter_rho;
        for (int k = 0; k < 3; ++k) { dv.motion1[k] = M.s.motion_vector1[k]; dv.motion2[k] = M.s.motion_vector2[k]; dv.torque[k] = M.s.torque_ref[k]; }
        dv.gh_dist_bin = M.s.gh_distance_binsize; dv.gh_time_bin = M.s.gh_time_binsize;
        dv.flags = M.features;
        dv.use_maxwell = M.s.use_maxwell; dv.calc_const_flow = M.s.calc_constant_flow; dv.enable_decay = M.s.enable_decay;
        dv.scatter_enabled = M.s.scatter_enabled; dv.scatter_cutoff = M.s.scatter_enable_cutoff;
        dv.low_flux = M.s.low_flux_mode; dv.motion_type = M.s.motion_type; dv.enable_force = M.s.enable_force;
        dv.gh_bounce_bin = M.s.gh_bounce_binsize; dv.gh_nb_bounce = M.gh_b; dv.gh_nb_dist = M.gh_d; dv.gh_nb_time = M.gh_t;
        dv.log_facet = M.s.log_enabled && M.s.log_facet_id >= 0 ? M.s.log_facet_id + (int32_t)fbase : -1;
        dv.log_limit = M.s.log_enabled ? M.s.log_limit : 0;
        log_cap = std::max(log_cap, dv.log_limit);
        dv.maxwell_base = (uint32_t)tab_x.size();
        for (auto& p : M.maxwell) { tab_x.push_back(p.x); tab_y.push_back(p.y); }
        dv.maxwell_guide = (uint32_t)guides.size();
        {
            const double ylast = M.maxwell.back().y;
            for (uint32_t k = 0; k < MFG_MAXWELL_GUIDE; ++k) {
                double start = ylast * (double)k / (double)MFG_MAXWELL_GUIDE;
                uint32_t i = 0;
                while (i + 2 < M.maxwell.size() && M.maxwell[i + 1].y <= start) ++i;
                guides.push_back(i);
            }
        }
        dv.param_base = (uint32_t)params.size(); dv.n_params = (uint32_t)M.params.size();
        for (auto& p : M.params) {
            params.push_back(DTable{(uint32_t)tab_x.size(), (uint32_t)p.pts.size(), p.logx, p.logy});
            for (auto& q : p.pts) { tab_x.push_back(q.x); tab_y.push_back(q.y); }
        }
        dv.id_base = (uint32_t)ids.size(); dv.n_ids = (uint32_t)M.IDs.size();
        for (auto& t : M.IDs) {
            ids.push_back(DTable{(uint32_t)tab_x.size(), (uint32_t)t.size(), 0, 0});
            for (auto& q : t) { tab_x.push_back(q.x); tab_y.push_back(q.y); }
        }
        dv.mom_base = (uint32_t)mom_s.size(); dv.n_moments = (uint32_t)M.mom_start.size();
        mom_s.insert(mom_s.end(), M.mom_start.begin(), M.mom_start.end());
        mom_e.insert(mom_e.end(), M.mom_end.begin(), M.mom_end.end());
        VariantLayout L;
        L.n_facets = dv.n_facets; L.n_slots = M.n_slots; L.n_struct = M.n_struct;
        L.fc = fc; fc += (uint64_t)M.n_slots * dv.n_facets;
        L.enable_force = M.s.enable_force != 0;
        L.force = force; if (L.enable_force) force += (uint64_t)M.n_slots * dv.n_facets * 9;
        L.tex = tex; L.n_tex = M.n_tex; tex += M.n_slots * M.n_tex;
        L.dir = dir; L.n_dir = M.n_dir; dir += M.n_slots * M.n_dir;
        L.prof = prof; L.n_prof = M.n_prof; prof += M.n_slots * M.n_prof;
        L.am = am; L.n_am = M.n_am; am += M.n_am;
        L.n_gh = M.gh_b + M.gh_d + M.gh_t; L.gh = gh; gh += M.n_slots * L.n_gh;
        L.fh = fh; L.n_fh = M.n_fh; fh += M.n_slots * M.n_fh;
        L.log_limit = dv.log_limit;
        for (auto& ph : M.phys) {
            L.off_tex.push_back(ph.tex_off); L.off_dir.push_back(ph.dir_off); L.off_prof.push_back(ph.prof_off);
            L.off_am.push_back(ph.am_off); L.off_fh.push_back(ph.fh_off);
        }
        // texture increments: one copy (slot-independent), indexed like slot 0
        tex_inc.insert(tex_inc.end(), M.tex_inc.begin(), M.tex_inc.end());
        for (uint64_t k = 0; k + 1 < M.n_slots; ++k) tex_inc.insert(tex_inc.end(), M.tex_inc.size(), 0.0);
        dv.fc_base = L.fc; dv.force_base = L.force; dv.tex_base = L.tex; dv.n_tex = L.n_tex;
        dv.dir_base = L.dir; dv.n_dir = L.n_dir; dv.prof_base = L.prof; dv.n_prof = L.n_prof;
        dv.am_base = L.am; dv.gh_base = L.gh; dv.n_gh = L.n_gh; dv.fh_base = L.fh; dv.n_fh = L.n_fh;
        dv.n_slots = M.n_slots; dv.glob_index = (uint32_t)vi;
        feat |= M.features;
        c->dvars.push_back(dv);
        c->layouts.push_back(std::move(L));
    }
    hipStream_t s = c->stream;
    c->h_phys = phys;
    // default hot facets of each variant: its sources and absorbers when
    // there are only a few of them (the autotuner refines the choice from
    // measured counts)
    std::vector<int32_t> hot;
    for (size_t vi = 0; vi < c->models.size(); ++vi) {
        const HostModel& M = *c->models[vi];
        const uint32_t fb = c->dvars[vi].facet_base;
        std::vector<int32_t> src(M.src_facet.begin(), M.src_facet.end()), ab;
        for (size_t f = 0; f < M.phys.size(); ++f)
            if (M.phys[f].sticking > 0.0 || M.phys[f].sticking_param >= 0) ab.push_back((int32_t)f);
        std::vector<int32_t> pick;
        if (src.size() <= 4) pick.insert(pick.end(), src.begin(), src.end());
        if (ab.size() <= 8) pick.insert(pick.end(), ab.begin(), ab.end());
        std::sort(pick.begin(), pick.end());
        pick.erase(std::unique(pick.begin(), pick.end()), pick.end());
        for (int32_t f : pick) if (hot.size() < MFG_MAX_HOT) hot.push_back(f + (int32_t)fb);
    }
    c->hot.clear();
    set_hot_facets(c, hot, false);
    c->facets.upload(facets, s); c->phys.upload(c->h_phys, s); c->poly.upload(poly, s);
    c->h_facets = facets;
    c->h_poly = poly;
    c->src_cum.upload(src_cum, s); c->src_facet.upload(src_facet, s); c->og_cdf.upload(og_cdf, s);
    c->am_src.upload(am_src, s);
    c->tab_x.upload(tab_x, s); c->tab_y.upload(tab_y, s); c->guides.upload(guides, s); c->params.upload(params, s); c->ids.upload(ids, s);
    c->mom_start.upload(mom_s, s); c->mom_end.upload(mom_e, s); c->tex_inc.upload(tex_inc, s);
    c->fc.alloc(std::max<uint64_t>(fc, 1)); c->force.alloc(std::max<uint64_t>(force, 1));
    c->tex.alloc(std::max<uint64_t>(tex, 1)); c->dir.alloc(std::max<uint64_t>(dir, 1));
    c->prof.alloc(std::max<uint64_t>(prof, 1)); c->am.alloc(std::max<uint64_t>(am, 1));
    c->gh.alloc(std::max<uint64_t>(gh, 1)); c->fh.alloc(std::max<uint64_t>(fh, 1));
    c->glob.alloc(c->models.size());
    c->counters.alloc(9);   // [8]: adjoint segment overflow flag
    c->log_cap = log_cap;
    c->log.alloc(std::max<uint32_t>(log_cap, 1));
    c->leaks.alloc((size_t)kLeakCap * 6);
    c->variants.alloc(c->dvars.size());
    c->prefix.alloc(c->dvars.size() + 1);
    if (c->pin_cap < c->dvars.size() + 1) {
        if (c->pin_vars) (void)hipHostFree(c->pin_vars);
        if (c->pin_prefix) (void)hipHostFree(c->pin_prefix);
        c->pin_vars = nullptr; c->pin_prefix = nullptr;
        HIPCHK(hipHostMalloc((void**)&c->pin_vars, (c->dvars.size() + 1) * sizeof(DVariant), hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void**)&c->pin_prefix, (c->dvars.size() + 1) * sizeof(uint64_t), hipHostMallocDefault));
        c->pin_cap = c->dvars.size() + 1;
    }
    c->features = feat;
    c->pid_base.assign(c->models.size(), 0);
    c->loaded = true;
    c->accel = false;
    HIPCHK(hipStreamSynchronize(s));
}

// Zero the device tallies. keep_particles = true is the tmpState->Reset()
// of a periodic report (Particle.cpp:1294-1312): the parked particles, the
// carry control words and the particle streams stay as they are. Otherwise
// (ResetSimulation) the parked particles are dropped too.
void zero_tallies(mfg_context* c, bool keep_particles = false) {
    hipStream_t s = c->stream;
    auto z = [&](void* p, size_t b) { if (mfg_launch_zero(p, b, s)) throw Err(MFG_ERR_DEVICE, "zero kernel launch failed"); };
    z(c->fc.p, c->fc.bytes()); z(c->force.p, c->force.bytes()); z(c->tex.p, c->tex.bytes());
    z(c->dir.p, c->dir.bytes()); z(c->prof.p, c->prof.bytes()); z(c->am.p, c->am.bytes());
    z(c->gh.p, c->gh.bytes()); z(c->fh.p, c->fh.bytes()); z(c->glob.p, c->glob.bytes());
    if (keep_particles) {
        // log and leak counts, time-limit flag; not the carry words (4..6)
        z(c->counters.p + 1, 3 * sizeof(unsigned long long));
        z(c->counters.p + 7, sizeof(unsigned long long));
    } else {
        z(c->counters.p, c->counters.bytes());
    }
    if (c->adj_j0.p) { z(c->adj_j0.p, c->adj_j0.bytes()); z(c->adj_jac.p, c->adj_jac.bytes()); }
    HIPCHK(hipStreamSynchronize(s));
    c->adj_des = 0;
    if (!keep_particles) c->in_flight = 0;   // the counters above include the carry control words
}

// Renumber a variant's nodes breadth-first from its roots, so the top of
// every tree is one prefix of the array (the part a block copies to LDS).
void breadth_first(std::vector<WideNode>& nodes, std::vector<int32_t>& roots) {
    std::vector<int32_t> order, idx(nodes.size(), -1);
    order.reserve(nodes.size());
    for (int32_t r : roots)
        if (r >= 0) { idx[(size_t)r] = (int32_t)order.size(); order.push_back(r); }
    for (size_t h = 0; h < order.size(); ++h) {
        const WideNode& nd = nodes[(size_t)order[h]];
        for (int k = 0; k < 8; ++k) {
            if (!(nd.valid & (1u << k)) || (nd.leaf_mask & (1u << k))) continue;
            int32_t ch = nd.child[k];
            idx[(size_t)ch] = (int32_t)order.size();
            order.push_back(ch);
        }
    }
    if (order.size() != nodes.size()) throw Err(MFG_ERR_INVALID, "BVH nodes unreachable from the roots");
    std::vector<WideNode> out(nodes.size());
    for (size_t i = 0; i < order.size(); ++i) {
        WideNode nd = nodes[(size_t)order[i]];
        for (int k = 0; k < 8; ++k)
            if ((nd.valid & (1u << k)) && !(nd.leaf_mask & (1u << k))) nd.child[k] = idx[(size_t)nd.child[k]];
        out[i] = nd;
    }
    nodes.swap(out);
    for (int32_t& r : roots)
        if (r >= 0) r = idx[(size_t)r];
}

void build_accel(mfg_context* c, int width, int leaf) {
    if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
    if (width == 0) width = MFG_BVH_WIDTH;
    if (width != 4 && width != 8) throw Err(MFG_ERR_INVALID, "BVH width must be 4 or 8 (0 = 8)");
    c->bvh_width = width;
    if (leaf <= 0) leaf = 2;
    // a traversal stack entry holds a leaf as 27 bits of first record and
    // 4 bits of facet count
    if (leaf > 16) throw Err(MFG_ERR_INVALID, "leaf size above 16");
    c->leaf_size = leaf;
    size_t nv = c->models.size();
    std::vector<std::vector<WideNode>> vnodes(nv);
    std::vector<std::vector<int32_t>> vleaf(nv), vroots(nv);
    std::vector<int> vdepth(nv, 0);
    std::atomic<size_t> next{0};
    std::vector<std::string> errors(nv);
    auto work = [&]() {
        for (;;) {
            size_t vi = next.fetch_add(1);
            if (vi >= nv) return;
            try {
                HostModel& M = *c->models[vi];
                double pad = 1e-5 * M.extent + 1e-9;
                for (uint32_t s = 0; s < M.n_struct; ++s) {
                    std::vector<int> prims = M.struct_facets[s];
                    if (prims.empty()) { vroots[vi].push_back(-1); continue; }
                    std::vector<B2> b2;
                    build_b2(b2, prims, M.boxes, leaf);
                    for (auto& p : prims) p += (int)c->dvars[vi].facet_base;
                    // indices are variant-local here, made global after all variants are built
                    Collapse col{b2, prims, vnodes[vi], vleaf[vi], 0, 0, pad, width};
                    int r = col.build(0, 1);
                    vroots[vi].push_back(r);
                    vdepth[vi] = std::max(vdepth[vi], col.max_depth);
                }
                breadth_first(vnodes[vi], vroots[vi]);
            } catch (const std::exception& e) {
                errors[vi] = e.what();
            }
        }
    };
    unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16));
    nt = (unsigned)std::min<size_t>(nt, nv);
    std::vector<std::thread> th;
    for (unsigned i = 1; i < nt; ++i) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    for (auto& e : errors) if (!e.empty()) throw Err(MFG_ERR_INVALID, e);
    // concatenate with global node / leaf offsets
    std::vector<BvhNode> nodes;
    std::vector<int32_t> leafs, roots;
    int depth = 0;
    for (size_t vi = 0; vi < nv; ++vi) {
        int32_t nb = (int32_t)nodes.size(), lb = (int32_t)leafs.size();
        for (WideNode nd : vnodes[vi]) {
            for (int k = 0; k < 8; ++k) {
                if (!(nd.valid & (1u << k))) continue;
                nd.child[k] += (nd.leaf_mask & (1u << k)) ? lb : nb;
            }
            nodes.push_back(pack_node(nd));
        }
        leafs.insert(leafs.end(), vleaf[vi].begin(), vleaf[vi].end());
        c->dvars[vi].root_base = (uint32_t)roots.size();
        for (int32_t r : vroots[vi]) roots.push_back(r < 0 ? -1 : r + nb);
        depth = std::max(depth, vdepth[vi]);
    }
    // stack entries: a walk pushes at most width-1 per level
    uint32_t need = (uint32_t)((width - 1) * depth + width);
    if (need > MFG_MAX_STACK) throw Err(MFG_ERR_LIMIT, "BVH too deep for the LDS traversal stack (depth " + std::to_string(depth) + ")");
    c->stack_size = need;
    if (leafs.size() >= ((size_t)1 << 27)) throw Err(MFG_ERR_LIMIT, "more than 2^27 BVH leaf records");
    hipStream_t s = c->stream;
    c->nodes.upload(nodes, s);
    // leaf-ordered copies of the test data (each facet sits in one leaf)
    std::vector<LeafFacet> lf(leafs.size());
    std::vector<LeafUV> luv(leafs.size());
    for (size_t i = 0; i < leafs.size(); ++i) {
        const DFacet& F = c->h_facets.at((size_t)leafs[i]);
        LeafFacet& L = lf[i];
        std::memset(&L, 0, sizeof(L));
        std::memset(&luv[i], 0, sizeof(LeafUV));
        for (int k = 0; k < 3; ++k) { L.O[k] = F.O[k]; L.N[k] = F.N[k]; L.U[k] = F.U[k]; L.V[k] = F.V[k]; }
        L.invU2 = F.invU2; L.invV2 = F.invV2;
        L.id = leafs[i]; L.flags = F.flags; L.poly_off = F.poly_off; L.poly_n = F.poly_n;
        if (F.poly_n == 3 || F.poly_n == 4)
            for (uint32_t k = 0; k < 2 * F.poly_n; ++k) luv[i].uv[k] = c->h_poly[2 * (size_t)F.poly_off + k];
    }
    c->leaves.upload(lf, s);
    c->leaf_uv.upload(luv, s);
    c->roots.upload(roots, s);
    c->h_roots = roots;
    HIPCHK(hipStreamSynchronize(s));
    c->accel = true;
    reset_graph(c);
}

// Nodes read from LDS: the top of the tree of a single-variant model.
uint32_t lds_node_count(const mfg_context* c, int block) {
    if (c->dvars.size() != 1) return 0;
    size_t n = c->node_cache;
    if (block > 256) {
        // a block of 768 or 1024 lanes serves a whole CU: the LDS its stacks and hot
        // facets leave (at least four times a 256-lane block's copy)
        const size_t used = (size_t)block * std::min<uint32_t>(c->stack_size, MFG_LDS_STACK) * sizeof(int32_t) +
                            c->hot.size() * 8 * sizeof(double);
        const size_t room = c->lds_per_block > used + 2048 ? (c->lds_per_block - used - 2048) / sizeof(BvhNode) : 0;
        n = std::max<size_t>(4 * (size_t)c->node_cache, room);
    }
    return (uint32_t)std::min<size_t>(n, c->nodes.n);
}

TraceParams make_params(mfg_context* c, uint64_t total, uint64_t max_events) {
    TraceParams p{};
    p.facets = c->facets.p; p.phys = c->phys.p; p.poly = c->poly.p; p.nodes = c->nodes.p;
    p.leaves = c->leaves.p; p.leaf_uv = c->leaf_uv.p; p.roots = c->roots.p; p.variants = c->variants.p;
    p.var_prefix = c->prefix.p; p.n_variants = (uint32_t)c->dvars.size();
    p.stack_size = std::min<uint32_t>(c->stack_size, MFG_LDS_STACK);
    p.stack_total = c->stack_size;
    p.stack_spill = p.stack_total > p.stack_size ? c->stack_spill.p : nullptr;
    p.src_cum = c->src_cum.p; p.src_facet = c->src_facet.p; p.og_cdf = c->og_cdf.p; p.am_src = c->am_src.p;
    p.hot_fc = c->hot_fc.p; p.n_hot = (uint32_t)c->hot.size();
    p.guides = c->guides.p; p.tab_x = c->tab_x.p; p.tab_y = c->tab_y.p; p.params = c->params.p; p.ids = c->ids.p;
    p.mom_start = c->mom_start.p; p.mom_end = c->mom_end.p;
    p.fc = c->fc.p; p.force = c->force.p; p.tex = c->tex.p; p.tex_inc = c->tex_inc.p; p.dir = c->dir.p;
    p.prof = c->prof.p; p.am = c->am.p; p.gh = c->gh.p; p.fh = c->fh.p; p.glob = c->glob.p;
    p.log = c->log.p; p.log_count = c->counters.p + 1; p.leaks = c->leaks.p; p.leak_count = c->counters.p + 2;
    p.leak_cap = kLeakCap;
    p.next_particle = c->counters.p;
    p.total = total;
    p.max_events = max_events;
    p.stop_flag = c->stop_dev;
    p.time_limit = (uint64_t)(c->time_limit_s * (double)c->clock_khz * 1000.0);
    p.expired = c->counters.p + 3;
    p.carry = c->carry.p; p.carry_ctl = c->counters.p + 4; p.carry_cap = c->carry_cap;
    p.carry_keep = c->carry_on ? 1u : 0u;
    p.trace_facet = c->trace_f.p; p.trace_kind = c->trace_kind.p; p.trace_n = c->trace_n; p.trace_k = c->trace_k;
    p.adj_j0 = c->adj_j0.p; p.adj_jac = c->adj_jac.p; p.adj_sticking = c->adj_s.p; p.adj_counts = c->adj_counts.p;
    p.adj_seg_t = c->adj_seg_t.p; p.adj_seg_x = c->adj_seg_x.p; p.adj_seg_overflow = c->counters.p + 8;
    p.adj_groups = c->adj_groups; p.adj_targets = c->adj_targets;
    return p;
}

int default_waves(uint32_t feats) {
    uint32_t k = mfg_trace_kernel_features(feats);
    if ((k & FEAT_ALL & ~FEAT_LOWFLUX) == 0) return 3;   // lean and low-flux sets
    return (k & FEAT_ADJOINT) ? 1 : 2;
}

void launch_shape(mfg_context* c, uint32_t feats, int& grid, int& block, int& waves) {
    block = (int)c->block;
    waves = c->waves ? c->waves : default_waves(feats);
    if (block > 256) {
        // the one-block-per-CU kernels exist for the lean and low-flux
        // feature sets (768 lanes: 3 waves per SIMD, 1024: 4); other runs
        // take 256-lane blocks
        if (mfg_trace_has_cu_block(feats, block)) waves = block / 256;
        else block = 256;
    }
    auto lds_bytes = [&](int b) {
        return (size_t)b * std::min<uint32_t>(c->stack_size, MFG_LDS_STACK) * sizeof(int32_t) +
               c->hot.size() * 8 * sizeof(double) + (size_t)lds_node_count(c, b) * sizeof(BvhNode);
    };
    if (block > 256 && mfg_trace_occupancy(feats, waves, block, lds_bytes(block)) < 1) {
        block = 256;   // the block's LDS does not fit
        waves = c->waves ? c->waves : default_waves(feats);
    }
    uint32_t per_cu = c->blocks_per_cu;
    if (block > 256) per_cu = 1;
    if (per_cu == 0) {
        // fill every CU to its occupancy limit (persistent launch)
        int occ = mfg_trace_occupancy(feats, waves, block, lds_bytes(block));
        per_cu = (uint32_t)std::max(1, occ);
    }
    grid = (int)(per_cu * (uint32_t)c->n_cu);
}

// The carry buffer holds at most one particle per lane of a launch. Growing
// it keeps the particles already parked in the half the next launch reads.
void ensure_carry(mfg_context* c, uint64_t lanes) {
    if (c->carry_cap >= lanes) return;
    uint64_t cap = std::max<uint64_t>(lanes, (uint64_t)c->n_cu * 2048);
    DevBuf<CarryRec> nb;
    nb.alloc(2 * cap);
    unsigned long long ctl[3] = {0, 0, 0};
    if (c->carry.p) {
        HIPCHK(hipMemcpy(ctl, c->counters.p + 4, sizeof(ctl), hipMemcpyDeviceToHost));
        uint64_t half = ctl[2] & 1ull;
        if (ctl[half]) HIPCHK(hipMemcpy(nb.p + half * cap, c->carry.p + half * c->carry_cap, ctl[half] * sizeof(CarryRec),
                                        hipMemcpyDeviceToDevice));
    }
    std::swap(c->carry.p, nb.p);
    std::swap(c->carry.n, nb.n);
    c->carry_cap = cap;
    reset_graph(c);
}

void run_variants(mfg_context* c, const uint64_t* n_des, const uint64_t* seeds, uint64_t max_events, mfg_step_result* res) {
    if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
    if (!c->accel) throw Err(MFG_ERR_STATE, "acceleration structure not built (call mfg_build_accel)");
    size_t nv = c->dvars.size();
    uint64_t total = 0;
    c->pin_prefix[0] = 0;
    for (size_t v = 0; v < nv; ++v) {
        total += n_des[v];
        c->pin_prefix[v + 1] = total;
        c->dvars[v].seed = seeds ? seeds[v] : c->seed;
        c->dvars[v].pid_base = c->pid_base[v];
        c->pin_vars[v] = c->dvars[v];
    }
    uint32_t feats = c->features;
    if (c->trace_n) feats |= FEAT_TRACE;
    if (c->adj_groups) feats |= FEAT_ADJOINT;
    int grid, block, waves;
    launch_shape(c, feats, grid, block, waves);
    if (c->adj_groups) {
        if (c->carry_on || c->in_flight) throw Err(MFG_ERR_STATE, "the adjoint traces every particle to its end: disable carry and drain first");
        if (!c->adj_s_ok)
            throw Err(MFG_ERR_STATE, "a sticking group's value is not inside (0,1): set it with mfg_adjoint_set_sticking");
        size_t need = (size_t)grid * block * c->adj_groups;
        if (c->adj_counts.n < need) c->adj_counts.alloc(need);
        const size_t seg = (size_t)grid * block * MFG_ADJ_SEG;
        if (c->adj_seg_t.n < seg) { c->adj_seg_t.alloc(seg); c->adj_seg_x.alloc(seg); reset_graph(c); }
    }
    if (c->carry_on) ensure_carry(c, (uint64_t)grid * (uint64_t)block);
    if (c->stack_size > MFG_LDS_STACK) {
        size_t need = (size_t)grid * block * (c->stack_size - MFG_LDS_STACK);
        if (c->stack_spill.n < need) { c->stack_spill.alloc(need); reset_graph(c); }
    }
    // read the global counters before the run to report this run's hits
    std::vector<DGlobal> before(nv);
    hipStream_t s = c->stream;
    HIPCHK(hipMemcpyAsync(before.data(), c->glob.p, nv * sizeof(DGlobal), hipMemcpyDeviceToHost, s));
    *c->stop_host = 0;
    TraceParams p = make_params(c, total, max_events);
    p.lds_nodes = lds_node_count(c, block);
    std::vector<uint64_t> key(n_des, n_des + nv);
    key.push_back(max_events); key.push_back(feats); key.push_back((uint64_t)grid); key.push_back(block);
    key.push_back((uint64_t)waves);
    key.push_back((uint64_t)(uintptr_t)p.carry); key.push_back(p.carry_keep);
    key.push_back((uint64_t)(uintptr_t)p.hot_fc); key.push_back(p.n_hot); key.push_back(p.time_limit);
    key.push_back(p.lds_nodes);
    // the events bracket the trace kernel alone (eager path) or the whole
    // graph (replay path)
    auto enqueue = [&](bool events) {
        HIPCHK(hipMemcpyAsync(c->variants.p, c->pin_vars, nv * sizeof(DVariant), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->prefix.p, c->pin_prefix, (nv + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        if (mfg_launch_zero(c->counters.p, sizeof(unsigned long long), s) ||
            mfg_launch_zero(c->counters.p + 3, sizeof(unsigned long long), s) ||
            (c->adj_groups && mfg_launch_zero(c->counters.p + 8, sizeof(unsigned long long), s)))
            throw Err(MFG_ERR_DEVICE, "zero kernel launch failed");
        if (events) HIPCHK(hipEventRecord(c->ev0, s));
        if (mfg_launch_trace(p, feats, waves, grid, block, s)) throw Err(MFG_ERR_DEVICE, std::string("trace launch failed: ") + hipGetErrorString(hipGetLastError()));
        if (p.carry && mfg_launch_carry_flip(p.carry_ctl, p.next_particle, p.carry, p.carry_cap, c->counters.p + 7, s))
            throw Err(MFG_ERR_DEVICE, "carry flip launch failed");
    };
    if (c->graph_on) {
        HIPCHK(hipEventRecord(c->ev0, s));
        if (!c->graph_exec || c->graph_key != key) {
            reset_graph(c);
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try { enqueue(false); } catch (...) { hipGraph_t g = nullptr; (void)hipStreamEndCapture(s, &g); if (g) (void)hipGraphDestroy(g); throw; }
            HIPCHK(hipStreamEndCapture(s, &c->graph));
            HIPCHK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
            c->graph_key = key;
        }
        HIPCHK(hipGraphLaunch(c->graph_exec, s));
    } else {
        enqueue(true);
    }
    HIPCHK(hipEventRecord(c->ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    std::vector<DGlobal> after(nv);
    HIPCHK(hipMemcpy(after.data(), c->glob.p, nv * sizeof(DGlobal), hipMemcpyDeviceToHost));
    unsigned long long expired = 0;
    HIPCHK(hipMemcpy(&expired, c->counters.p + 3, sizeof(expired), hipMemcpyDeviceToHost));
    if (p.carry) {
        unsigned long long ctl[4];   // carry control words and the overflow flag
        HIPCHK(hipMemcpy(ctl, c->counters.p + 4, sizeof(ctl), hipMemcpyDeviceToHost));
        c->in_flight = ctl[ctl[2] & 1ull];
        if (ctl[3]) throw Err(MFG_ERR_STATE, "carry buffer overflow: parked particles were lost");
    }
    uint64_t hits = 0, des = 0;
    for (size_t v = 0; v < nv; ++v) {
        hits += after[v].nb_mc_hit - before[v].nb_mc_hit;
        des += after[v].nb_desorbed - before[v].nb_desorbed;
        c->pid_base[v] += n_des[v];
    }
    if (c->adj_groups) {
        unsigned long long ov = 0;
        HIPCHK(hipMemcpy(&ov, c->counters.p + 8, sizeof(ov), hipMemcpyDeviceToHost));
        if (ov) throw Err(MFG_ERR_LIMIT, "adjoint: a particle scored on targets more than " + std::to_string(MFG_ADJ_SEG) +
                                             " times between two sticky collisions (cutoff term incomplete)");
        c->adj_des += des;
    }
    if (res) {
        res->status = 0;
        res->stopped = expired ? 2 : *c->stop_host ? 1 : 0;
        res->desorbed = des;
        res->hits = hits;
        res->kernel_ms = ms;
        if (des < total && !res->stopped) res->status = 2;
    }
}

void add_tallies(mfg_context* c, uint32_t vi, const mfg_tally_view* o) {
    if (!o || o->struct_size != sizeof(mfg_tally_view)) throw Err(MFG_ERR_INVALID, "mfg_tally_view size mismatch");
    if (vi >= c->layouts.size()) throw Err(MFG_ERR_INVALID, "variant index out of range");
    const VariantLayout& L = c->layouts[vi];
    auto fetch = [&](auto* dev, uint64_t off, uint64_t n) {
        using T = std::remove_pointer_t<decltype(dev)>;
        std::vector<T> h(n);
        if (n) HIPCHK(hipMemcpy(h.data(), dev + off, n * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    };
    if (o->global) {
        auto g = fetch(c->glob.p, vi, 1);
        mfg_global_counters& a = *o->global;
        a.nb_mc_hit += g[0].nb_mc_hit; a.nb_hit_equiv += g[0].nb_hit_equiv; a.nb_desorbed += g[0].nb_desorbed;
        a.nb_abs_equiv += g[0].nb_abs_equiv; a.nb_leak_total += g[0].nb_leak; a.dist_traveled_total += g[0].dist_total;
        a.dist_traveled_full_hits_only += g[0].dist_full; a.nb_decayed += g[0].nb_decayed;
    }
    uint64_t S = L.n_slots;
    if (o->facets) {
        auto h = fetch(c->fc.p, L.fc, S * L.n_facets);
        // without low-flux weighting the kernel leaves the slot-0 hit
        // equivalents to this read: every hit weighs 1
        if (!c->dvars[vi].low_flux)
            for (uint64_t i = 0; i < L.n_facets; ++i) h[i].nb_hit_equiv = (double)h[i].nb_mc_hit;
        for (size_t i = 0; i < h.size(); ++i) {
            auto& a = o->facets[i]; const auto& b = h[i];
            a.nb_mc_hit += b.nb_mc_hit; a.nb_hit_equiv += b.nb_hit_equiv; a.nb_desorbed += b.nb_desorbed;
            a.nb_abs_equiv += b.nb_abs_equiv; a.sum_1_per_ort_velocity += b.sum_1_per_ort_velocity;
            a.sum_v_ort += b.sum_v_ort; a.sum_1_per_velocity += b.sum_1_per_velocity;
        }
    }
    if (o->forces && L.enable_force) {
        auto h = fetch(c->force.p, L.force, S * L.n_facets * 9);
        for (size_t i = 0; i < h.size(); ++i) o->forces[i] += h[i];
    }
    if (o->textures && L.n_tex) {
        auto h = fetch(c->tex.p, L.tex, S * L.n_tex);
        for (size_t i = 0; i < h.size(); ++i) {
            o->textures[i].count_equiv += h[i].count_equiv;
            o->textures[i].sum_v_ort_per_area += h[i].sum_v_ort_per_area;
            o->textures[i].sum_1_per_ort_velocity += h[i].sum_1_per_ort_velocity;
        }
    }
    if (o->directions && L.n_dir) {
        auto h = fetch(c->dir.p, L.dir, S * L.n_dir);
        for (size_t i = 0; i < h.size(); ++i) {
            for (int k = 0; k < 3; ++k) o->directions[i].dir[k] += h[i].dir[k];
            o->directions[i].count += h[i].count;
        }
    }
    if (o->profiles && L.n_prof) {
        auto h = fetch(c->prof.p, L.prof, S * L.n_prof);
        for (size_t i = 0; i < h.size(); ++i) {
            o->profiles[i].count_equiv += h[i].count_equiv;
            o->profiles[i].sum_v_ort += h[i].sum_v_ort;
            o->profiles[i].sum_1_per_ort_velocity += h[i].sum_1_per_ort_velocity;
        }
    }
    if (o->angle_maps && L.n_am) {
        auto h = fetch(c->am.p, L.am, L.n_am);
        for (size_t i = 0; i < h.size(); ++i) o->angle_maps[i] += h[i];
    }
    if (o->global_hist && L.n_gh) {
        auto h = fetch(c->gh.p, L.gh, S * L.n_gh);
        for (size_t i = 0; i < h.size(); ++i) o->global_hist[i] += h[i];
    }
    if (o->facet_hist && L.n_fh) {
        auto h = fetch(c->fh.p, L.fh, S * L.n_fh);
        for (size_t i = 0; i < h.size(); ++i) o->facet_hist[i] += h[i];
    }
    if (o->log && o->log_count && c->log_cap) {
        auto n = fetch(c->counters.p, 1, 1);
        uint64_t cnt = std::min<uint64_t>(n[0], L.log_limit);
        auto h = fetch(c->log.p, 0, cnt);
        for (uint64_t i = 0; i < cnt; ++i) o->log[i] = h[i];
        *o->log_count = cnt;
    }
    if (o->leak_pos_dir && o->leak_count) {
        auto n = fetch(c->counters.p, 2, 1);
        uint64_t cnt = std::min<uint64_t>(n[0], kLeakCap);
        auto h = fetch(c->leaks.p, 0, cnt * 6);
        if (!h.empty()) std::memcpy(o->leak_pos_dir, h.data(), h.size() * sizeof(double));
        *o->leak_count = n[0];
    }
}

void fill_layout(mfg_context* c, uint32_t vi, mfg_tally_layout* L) {
    if (!L) throw Err(MFG_ERR_INVALID, "null layout");
    if (vi >= c->layouts.size()) throw Err(MFG_ERR_INVALID, "variant index out of range");
    const VariantLayout& V = c->layouts[vi];
    std::memset(L, 0, sizeof(*L));
    L->struct_size = sizeof(*L);
    L->n_facets = V.n_facets; L->n_slots = V.n_slots; L->n_structures = V.n_struct;
    L->n_tex_cells = V.n_tex; L->n_dir_cells = V.n_dir; L->n_profile_slices = V.n_prof;
    L->n_angle_cells = V.n_am; L->n_global_hist_bins = V.n_gh; L->n_facet_hist_bins = V.n_fh;
    L->log_capacity = V.log_limit; L->leak_capacity = kLeakCap;
}

void load_all(mfg_context* c, const mfg_model_view* const* views, uint32_t n) {
    if (n == 0) throw Err(MFG_ERR_INVALID, "no models");
    std::vector<std::unique_ptr<HostModel>> ms(n);
    std::vector<std::string> errs(n);
    std::vector<int> codes(n, 0);
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= n) return;
            try {
                ms[i] = std::make_unique<HostModel>();
                prepare(*ms[i], views[i]);
            } catch (const Err& e) { errs[i] = e.what(); codes[i] = e.code; }
            catch (const std::exception& e) { errs[i] = e.what(); codes[i] = MFG_ERR_INVALID; }
        }
    };
    unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16));
    nt = std::min<unsigned>(nt, n);
    std::vector<std::thread> th;
    for (unsigned i = 1; i < nt; ++i) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    for (uint32_t i = 0; i < n; ++i)
        if (codes[i]) throw Err(codes[i], (n > 1 ? "variant " + std::to_string(i) + ": " : std::string()) + errs[i]);
    c->models = std::move(ms);
    c->adj_groups = 0; c->adj_targets = 0;
    upload_models(c);
    zero_tallies(c);
}

}  // namespace

namespace {
// Checks shared by both fixed-ray-batch entry points.
void check_batch(mfg_context* c, uint64_t n, const double* o, const double* d, const int32_t* last, int32_t structure,
                 int32_t* of, double* ot, double* ou, double* ov) {
    if (!c->accel) throw Err(MFG_ERR_STATE, "acceleration structure not built");
    if (!o || !d || !of || !ot || !ou || !ov) throw Err(MFG_ERR_INVALID, "null buffer");
    if (structure < 0 || (uint32_t)structure >= c->dvars[0].n_struct) throw Err(MFG_ERR_INVALID, "structure out of range");
    if (last)
        for (uint64_t i = 0; i < n; ++i)
            if (last[i] >= (int32_t)c->dvars[0].n_facets) throw Err(MFG_ERR_INVALID, "last_hit out of range");
}

// First hits of n rays whose inputs and outputs sit in device buffers.
void launch_batch(mfg_context* c, uint64_t n, const double* dO, const double* dD, const int32_t* dL, int32_t structure,
                  int32_t* dF, double* dT, double* dU, double* dV) {
    IsectParams p{};
    p.facets = c->facets.p; p.poly = c->poly.p; p.nodes = c->nodes.p; p.leaves = c->leaves.p; p.leaf_uv = c->leaf_uv.p;
    p.root = c->h_roots.at(c->dvars[0].root_base + (uint32_t)structure);
    p.stack_size = c->stack_size; p.n = n; p.origins = dO; p.dirs = dD; p.last_hit = dL;
    p.out_facet = dF; p.out_t = dT; p.out_u = dU; p.out_v = dV; p.facet_base = c->dvars[0].facet_base;
    if (mfg_launch_intersect(p, c->stream)) throw Err(MFG_ERR_DEVICE, "intersect launch failed");
}
}  // namespace

extern "C" {

int mfg_abi_version(void) { return MFG_ABI_VERSION; }

int mfg_create(mfg_context** out, int device) {
    if (!out) return MFG_ERR_INVALID;
    *out = nullptr;
    auto c = std::make_unique<mfg_context>();
    c->device = device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MFG_ERR_DEVICE;
    if (device < 0 || device >= ndev) return MFG_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return MFG_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MFG_ERR_DEVICE;
    c->n_cu = prop.multiProcessorCount;
    c->lds_per_block = prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return MFG_ERR_DEVICE;
    if (hipHostMalloc((void**)&c->stop_host, sizeof(int32_t), hipHostMallocMapped) != hipSuccess) return MFG_ERR_DEVICE;
    *c->stop_host = 0;
    if (hipHostGetDevicePointer((void**)&c->stop_dev, c->stop_host, 0) != hipSuccess) return MFG_ERR_DEVICE;
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) return MFG_ERR_DEVICE;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
        return MFG_ERR_DEVICE;
    c->clock_khz = (uint64_t)khz;
    if (mfg_scratch_address(device, &c->scratch) != 0 || !c->scratch) return MFG_ERR_DEVICE;
    *out = c.release();
    return MFG_OK;
}

void mfg_destroy(mfg_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    reset_graph(c);
    if (c->stop_host) (void)hipHostFree(c->stop_host);
    if (c->pin_vars) (void)hipHostFree(c->pin_vars);
    if (c->pin_prefix) (void)hipHostFree(c->pin_prefix);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    hipStream_t s = c->stream;
    delete c;
    if (s) (void)hipStreamDestroy(s);
}

const char* mfg_last_error(const mfg_context* c) { return c ? c->err.c_str() : "null context"; }

int mfg_load_model(mfg_context* c, const mfg_model_view* m) {
    return guarded(c, [&] { const mfg_model_view* v[1] = {m}; load_all(c, v, 1); });
}

int mfg_load_variants(mfg_context* c, const mfg_model_view* const* models, uint32_t n) {
    return guarded(c, [&] { if (!models) throw Err(MFG_ERR_INVALID, "null model list"); load_all(c, models, n); });
}

int mfg_build_accel(mfg_context* c, int width, int leaf) {
    return guarded(c, [&] { build_accel(c, width, leaf); });
}

int mfg_reset(mfg_context* c, uint64_t seed, uint64_t offset) {
    return guarded(c, [&] {
        if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
        c->seed = seed;
        for (auto& b : c->pid_base) b = offset;
        zero_tallies(c);
    });
}

int mfg_run(mfg_context* c, uint64_t n_des, uint64_t max_events, mfg_step_result* r) {
    return guarded(c, [&] {
        if (c->loaded && c->dvars.size() != 1) throw Err(MFG_ERR_STATE, "several variants loaded: use mfg_run_variants");
        uint64_t s = c->seed;
        run_variants(c, &n_des, &s, max_events, r);
    });
}

int mfg_run_variants(mfg_context* c, const uint64_t* n_des, const uint64_t* seeds, uint64_t max_events,
                     mfg_step_result* r) {
    return guarded(c, [&] {
        if (!n_des) throw Err(MFG_ERR_INVALID, "null desorption counts");
        run_variants(c, n_des, seeds, max_events, r);
    });
}

int mfg_set_time_limit(mfg_context* c, double seconds) {
    return guarded(c, [&] {
        if (!(seconds > 0.0 && seconds <= 86400.0)) throw Err(MFG_ERR_INVALID, "time limit must be in (0, 86400] s");
        c->time_limit_s = seconds;
    });
}

int mfg_request_stop(mfg_context* c) {
    if (!c || !c->stop_host) return MFG_ERR_INVALID;
    __atomic_store_n(c->stop_host, 1, __ATOMIC_SEQ_CST);
    return MFG_OK;
}

int mfg_get_layout(const mfg_context* c, mfg_tally_layout* L) {
    return guarded(const_cast<mfg_context*>(c), [&] { if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded"); fill_layout(const_cast<mfg_context*>(c), 0, L); });
}

int mfg_get_variant_layout(const mfg_context* c, uint32_t vi, mfg_tally_layout* L) {
    return guarded(const_cast<mfg_context*>(c), [&] { fill_layout(const_cast<mfg_context*>(c), vi, L); });
}

int mfg_get_facet_offsets(const mfg_context* c, int kind, int64_t* out, size_t n) {
    return guarded(const_cast<mfg_context*>(c), [&] {
        if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
        const VariantLayout& L = c->layouts[0];
        if (!out || n < L.n_facets) throw Err(MFG_ERR_INVALID, "offset buffer too small");
        const std::vector<int64_t>* v = kind == 0 ? &L.off_tex : kind == 1 ? &L.off_dir : kind == 2 ? &L.off_prof
                                       : kind == 3 ? &L.off_am : kind == 4 ? &L.off_fh : nullptr;
        if (!v) throw Err(MFG_ERR_INVALID, "unknown offset kind");
        std::copy(v->begin(), v->end(), out);
    });
}

int mfg_read_tallies(mfg_context* c, const mfg_tally_view* o) {
    return guarded(c, [&] { if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded"); add_tallies(c, 0, o); });
}

int mfg_read_variant_tallies(mfg_context* c, uint32_t vi, const mfg_tally_view* o) {
    return guarded(c, [&] { if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded"); add_tallies(c, vi, o); });
}

int mfg_clear_tallies(mfg_context* c) {
    return guarded(c, [&] {
        if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
        zero_tallies(c, true);
    });
}

int mfg_take_tallies(mfg_context* c, const mfg_tally_view* o) {
    return guarded(c, [&] {
        if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
        if (c->dvars.size() != 1)
            throw Err(MFG_ERR_STATE, "several variants loaded: read each with mfg_read_variant_tallies, then mfg_clear_tallies");
        add_tallies(c, 0, o);
        zero_tallies(c, true);
    });
}

int mfg_reinit_particle_log(mfg_context* c) {
    return guarded(c, [&] {
        if (!c->loaded) throw Err(MFG_ERR_STATE, "no model loaded");
        if (mfg_launch_zero(c->counters.p + 1, sizeof(unsigned long long), c->stream))
            throw Err(MFG_ERR_DEVICE, "zero kernel launch failed");
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int mfg_sanity_check(const mfg_model_view* v, char* out, size_t cap, uint32_t* n_problems) {
    if (!n_problems) return MFG_ERR_INVALID;
    std::vector<std::string> log;
    try {
        HostModel M;
        prepare(M, v);
    } catch (const std::exception& e) {
        log.push_back(e.what());
    }
    // the reference's own checks (MolflowSimulationModel::SanityCheck,
    // MolflowSimGeom.cpp:337-404), on what the view holds
    if (v && v->struct_size == sizeof(mfg_model_view) && v->facets && v->settings) {
        bool any_source = false;
        for (uint32_t f = 0; f < v->n_facets; ++f) {
            const mfg_facet_desc& d = v->facets[f];
            const bool counters = d.count_des || d.count_abs || d.count_refl || d.count_trans || d.count_dir;
            if (!d.is_textured && d.tex_width_precise * d.tex_height_precise > 0)
                log.push_back("[Facet #" + std::to_string(f + 1) + "] Untextured facet with texture size");
            else if (!d.is_textured && counters)
                log.push_back("[Facet #" + std::to_string(f + 1) + "] Untextured facet with texture counters");
            if (d.desorb_type != MFG_DES_NONE && d.temperature_param >= 0)
                log.push_back("[Facet " + std::to_string(f + 1) +
                              "]: Time-dependent temperature not allowed on facets with outgassing");
            any_source = any_source || d.desorb_type != MFG_DES_NONE;
        }
        if (v->settings->enable_decay && v->settings->half_life <= 0.0)
            log.push_back("Particle decay is set, but half life was not set");
        if (!any_source) log.push_back("No desorption facet found");
    }
    std::string all;
    for (auto& s : log) all += s + "\n";
    if (out && cap) {
        size_t n = std::min(cap - 1, all.size());
        std::memcpy(out, all.data(), n);
        out[n] = '\0';
    }
    *n_problems = (uint32_t)log.size();
    return MFG_OK;
}

int mfg_model_totals(const mfg_model_view* v, double* out) {
    if (!out) return MFG_ERR_INVALID;
    try {
        HostModel M;
        prepare(M, v);
        out[0] = M.final_rate;
        out[1] = M.final_rate_pam3s;
        out[2] = M.total_des;
        out[3] = M.latest;
    } catch (const Err& e) {
        return e.code;
    } catch (const std::exception&) {
        return MFG_ERR_INVALID;
    }
    return MFG_OK;
}

int mfg_hits_size(const mfg_context* c, uint64_t* bytes) {
    return guarded(const_cast<mfg_context*>(c), [&] {
        if (!bytes) throw Err(MFG_ERR_INVALID, "null output");
        uint64_t b = 0;
        for (auto& L : c->layouts) {
            b += sizeof(mfg_global_counters) + (uint64_t)L.n_slots * L.n_facets * sizeof(mfg_facet_counters);
            if (L.enable_force) b += (uint64_t)L.n_slots * L.n_facets * 9 * sizeof(double);
            b += L.n_slots * (L.n_tex * sizeof(mfg_texture_cell) + L.n_dir * sizeof(mfg_direction_cell) +
                              L.n_prof * sizeof(mfg_profile_slice) + (L.n_gh + L.n_fh) * sizeof(double));
            b += L.n_am * sizeof(uint64_t);
        }
        *bytes = b;
    });
}

int mfg_intersect_batch(mfg_context* c, uint64_t n, const double* o, const double* d, const int32_t* last,
                        int32_t structure, int32_t* of, double* ot, double* ou, double* ov) {
    return guarded(c, [&] {
        check_batch(c, n, o, d, last, structure, of, ot, ou, ov);
        if (n == 0) return;
        DevBuf<double> dO, dD, dT, dU, dV;
        DevBuf<int32_t> dL, dF;
        dO.alloc(3 * n); dD.alloc(3 * n); dT.alloc(n); dU.alloc(n); dV.alloc(n); dF.alloc(n);
        hipStream_t s = c->stream;
        HIPCHK(hipMemcpyAsync(dO.p, o, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dD.p, d, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        if (last) { dL.alloc(n); HIPCHK(hipMemcpyAsync(dL.p, last, n * sizeof(int32_t), hipMemcpyHostToDevice, s)); }
        launch_batch(c, n, dO.p, dD.p, last ? dL.p : nullptr, structure, dF.p, dT.p, dU.p, dV.p);
        HIPCHK(hipMemcpyAsync(of, dF.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ot, dT.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ou, dU.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ov, dV.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int mfg_intersect_batch_static(mfg_context* c, uint64_t n, const double* o, const double* d, const int32_t* last,
                               int32_t structure, int32_t* of, double* ot, double* ou, double* ov) {
    return guarded(c, [&] {
        if (n > MFG_STATIC_MAX_RAYS)
            throw Err(MFG_ERR_LIMIT, "mfg_intersect_batch_static takes at most " + std::to_string(MFG_STATIC_MAX_RAYS) +
                                         " rays per call");
        check_batch(c, n, o, d, last, structure, of, ot, ou, ov);
        if (n == 0) return;
        // every buffer below is a field of the static staging array; nothing
        // is allocated on the device (the outputs are written in full, so
        // nothing needs clearing between calls)
        StaticRays* S = c->scratch;
        hipStream_t s = c->stream;
        HIPCHK(hipMemcpyAsync(S->origins, o, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(S->dirs, d, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        if (last) HIPCHK(hipMemcpyAsync(S->last_hit, last, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
        launch_batch(c, n, S->origins, S->dirs, last ? S->last_hit : nullptr, structure, S->facet, S->t, S->u, S->v);
        HIPCHK(hipMemcpyAsync(of, S->facet, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ot, S->t, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ou, S->u, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(ov, S->v, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int mfg_maxwell_cdf(double gas_mass, double* v, double* p, size_t n) {
    if (!v || !p || !(gas_mass > 0)) return MFG_ERR_INVALID;
    auto c = maxwell_cdf(gas_mass);
    for (size_t i = 0; i < n && i < c.size(); ++i) { v[i] = c[i].x; p[i] = c[i].y; }
    return MFG_OK;
}

int mfg_parse_user_moment(const char* content, double window, double* out, size_t cap, size_t* n) {
    if (!content || !n) return MFG_ERR_INVALID;
    auto v = parse_user_moment(content, window);
    for (size_t i = 0; i < v.size() && i < cap && out; ++i) out[i] = v[i];
    *n = v.size();
    return MFG_OK;
}

int mfg_integrated_desorption(const double* x, const double* y, size_t np, int lx, int ly, double latest,
                              double* ot, double* oc, size_t cap, size_t* n) {
    if (!x || !y || !n || np == 0) return MFG_ERR_INVALID;
    std::vector<XY> p;
    for (size_t i = 0; i < np; ++i) p.push_back({x[i], y[i]});
    auto id = integrated_desorption(p.data(), p.size(), lx != 0, ly != 0, latest);
    for (size_t i = 0; i < id.size() && i < cap; ++i) { ot[i] = id[i].x; oc[i] = id[i].y; }
    *n = id.size();
    return MFG_OK;
}

int mfg_set_trace(mfg_context* c, uint32_t n, uint32_t k) {
    return guarded(c, [&] {
        if ((uint64_t)n * k > (1ull << 28)) throw Err(MFG_ERR_LIMIT, "trace buffer too large");
        c->trace_n = (n && k) ? n : 0;
        c->trace_k = (n && k) ? k : 0;
        if (c->trace_n) {
            std::vector<int32_t> f((size_t)n * k, -2), kd((size_t)n * k, -1);
            c->trace_f.upload(f, c->stream);
            c->trace_kind.upload(kd, c->stream);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        reset_graph(c);
    });
}

int mfg_get_trace(mfg_context* c, int32_t* facet, int32_t* kind, size_t n) {
    return guarded(c, [&] {
        size_t need = (size_t)c->trace_n * c->trace_k;
        if (!facet || !kind || n < need) throw Err(MFG_ERR_INVALID, "trace buffers too small");
        if (need) {
            HIPCHK(hipMemcpy(facet, c->trace_f.p, need * sizeof(int32_t), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(kind, c->trace_kind.p, need * sizeof(int32_t), hipMemcpyDeviceToHost));
        }
    });
}

int mfg_set_launch(mfg_context* c, uint32_t block, uint32_t per_cu) {
    return guarded(c, [&] {
        if (block != 0 && block != 768 && block != 1024 && (block % 64 != 0 || block > 256))
            throw Err(MFG_ERR_INVALID, "block must be a multiple of 64 up to 256, 768 or 1024");
        if (per_cu > 64) throw Err(MFG_ERR_INVALID, "blocks per CU above 64");
        c->block = block ? block : 256;
        c->blocks_per_cu = per_cu;
        reset_graph(c);
    });
}

int mfg_set_carry(mfg_context* c, int enable) {
    return guarded(c, [&] {
        if (enable && c->adj_groups) throw Err(MFG_ERR_STATE, "the adjoint traces every particle to its end (no carry)");
        c->carry_on = enable != 0;
        reset_graph(c);
    });
}

int mfg_in_flight(const mfg_context* c, uint64_t* n) {
    return guarded(const_cast<mfg_context*>(c), [&] {
        if (!n) throw Err(MFG_ERR_INVALID, "null output");
        *n = c->in_flight;
    });
}

int mfg_set_kernel_variant(mfg_context* c, uint32_t waves) {
    return guarded(c, [&] {
        if (waves > 5) throw Err(MFG_ERR_INVALID, "waves per SIMD must be 0..5");
        c->waves = (int)waves;
        reset_graph(c);
    });
}

int mfg_autotune(mfg_context* c, uint64_t n_probe, mfg_tune_result* out) {
    return guarded(c, [&] {
        if (!c->accel) throw Err(MFG_ERR_STATE, "acceleration structure not built");
        if (n_probe == 0) throw Err(MFG_ERR_INVALID, "probe size must be positive");
        uint32_t feats = c->features | (c->trace_n ? FEAT_TRACE : 0u) | (c->adj_groups ? FEAT_ADJOINT : 0u);
        int wl[8];
        int nw = mfg_trace_variants(feats, wl, 8);
        const uint32_t blocks[] = {128, 256};
        size_t nv = c->dvars.size();
        std::vector<uint64_t> nd(nv, std::max<uint64_t>(1, n_probe / nv)), seeds(nv, c->seed);
        std::vector<uint64_t> save_pid = c->pid_base;
        const bool save_carry = c->carry_on;
        // probes keep particles in flight between launches (when the adjoint
        // allows it), so each timed probe runs at the steady-state rate
        // rather than waiting for its longest trajectories
        c->carry_on = c->adj_groups == 0;
        double best = -1;
        mfg_tune_result r{};
        // BVH shapes tried: 8-wide with 2 and 1 facets per leaf, 4-wide with 2
        // (a leaf of one facet needs no per-leaf loop; more nodes in exchange)
        const int shapes[][2] = {{8, 2}, {8, 1}, {4, 2}};
        for (const auto& sh : shapes) {
            const int width = sh[0], leaf = sh[1];
            if (width != c->bvh_width || leaf != c->leaf_size) build_accel(c, width, leaf);
            // one untimed probe per shape loads the code objects and fills the lanes
            c->block = 256; c->blocks_per_cu = 0; c->waves = wl[0];
            run_variants(c, nd.data(), seeds.data(), 0, nullptr);
            for (int i = 0; i < nw + 2; ++i)
                for (uint32_t b : blocks) {
                    // last: one block per CU of 768 and of 1024 lanes, where the kernel exists
                    const uint32_t cu_block = i == nw ? 768u : 1024u;
                    if (i >= nw && (b != blocks[0] || !mfg_trace_has_cu_block(feats, (int)cu_block))) continue;
                    c->waves = i < nw ? wl[i] : (int)cu_block / 256; c->block = i < nw ? b : cu_block; c->blocks_per_cu = 0;
                    mfg_step_result sr{};
                    run_variants(c, nd.data(), seeds.data(), 0, &sr);
                    double hs = sr.kernel_ms > 0 ? (double)sr.hits / (sr.kernel_ms * 1e-3) : 0.0;
                    if (hs > best) {
                        best = hs;
                        int grid, block, waves;
                        launch_shape(c, feats, grid, block, waves);
                        r.waves = (uint32_t)waves; r.block = (uint32_t)block; r.blocks_per_cu = (uint32_t)(grid / c->n_cu);
                        r.hits_per_s = hs; r.bvh_width = (uint32_t)width; r.leaf_size = (uint32_t)leaf;
                    }
                }
        }
        if ((int)r.bvh_width != c->bvh_width || (int)r.leaf_size != c->leaf_size)
            build_accel(c, (int)r.bvh_width, (int)r.leaf_size);
        c->carry_on = save_carry;
        c->waves = (int)r.waves; c->block = r.block; c->blocks_per_cu = 0;
        // hot facets: those that took at least 1/2 % of the probes' slot-0
        // counter updates (hits + desorptions), most frequent first
        {
            std::vector<std::pair<uint64_t, int32_t>> ev;
            uint64_t total = 0;
            for (size_t vi = 0; vi < c->dvars.size(); ++vi) {
                const DVariant& dv = c->dvars[vi];
                std::vector<mfg_facet_counters> h(dv.n_facets);
                HIPCHK(hipMemcpy(h.data(), c->fc.p + dv.fc_base, dv.n_facets * sizeof(mfg_facet_counters), hipMemcpyDeviceToHost));
                for (uint32_t f = 0; f < dv.n_facets; ++f) {
                    uint64_t e = h[f].nb_mc_hit + h[f].nb_desorbed;
                    total += e;
                    if (e) ev.push_back({e, (int32_t)(dv.facet_base + f)});
                }
            }
            std::sort(ev.begin(), ev.end(), [](const auto& a, const auto& b) { return a.first > b.first || (a.first == b.first && a.second < b.second); });
            std::vector<int32_t> hot;
            for (auto& e : ev) {
                if (hot.size() >= MFG_MAX_HOT || e.first * 200 < total) break;
                hot.push_back(e.second);
            }
            std::sort(hot.begin(), hot.end());
            set_hot_facets(c, hot, true);
            r.n_hot = (uint32_t)hot.size();
        }
        c->pid_base = save_pid;
        zero_tallies(c);
        reset_graph(c);
        if (out) *out = r;
    });
}

int mfg_enable_graph(mfg_context* c, int enable) {
    return guarded(c, [&] { c->graph_on = enable != 0; reset_graph(c); });
}

int mfg_adjoint_setup(mfg_context* c, const int32_t* group_of_facet, uint32_t G, const int32_t* targets, uint32_t T) {
    return guarded(c, [&] {
        if (!c->loaded || c->dvars.size() != 1) throw Err(MFG_ERR_STATE, "adjoint needs exactly one loaded model");
        HostModel& M = *c->models[0];
        if (!M.s.low_flux_mode) throw Err(MFG_ERR_INVALID, "adjoint requires low-flux mode");
        if (G == 0 || G > MFG_MAX_GROUPS) throw Err(MFG_ERR_LIMIT, "number of sticking groups must be 1..64");
        if (!group_of_facet || !targets || T == 0) throw Err(MFG_ERR_INVALID, "null adjoint arrays");
        uint32_t nF = (uint32_t)M.phys.size();
        std::vector<DPhys>& phys = c->h_phys;   // one variant: global = local facet index
        for (uint32_t f = 0; f < nF; ++f) {
            int32_t g = group_of_facet[f];
            if (g >= (int32_t)G) throw Err(MFG_ERR_INVALID, "group index out of range");
            phys[f].adj_group = g < 0 ? -1 : g;
            phys[f].adj_target = -1;
            phys[f].sticking_param = g >= 0 ? -1 : M.phys[f].sticking_param;
        }
        for (uint32_t t = 0; t < T; ++t) {
            if (targets[t] < 0 || (uint32_t)targets[t] >= nF) throw Err(MFG_ERR_INVALID, "target facet out of range");
            phys[targets[t]].adj_target = (int32_t)t;
        }
        c->phys.upload(phys, c->stream);
        c->adj_groups = G; c->adj_targets = T;
        c->adj_j0.alloc(T); c->adj_jac.alloc((size_t)T * G); c->adj_s.alloc(G);
        c->adj_up.alloc(T); c->adj_grad.alloc(G);
        // each group starts from its facets' common constant sticking; a
        // group whose facets disagree is refused, and one whose value is not
        // inside (0,1) (where d/ds = -n/(1-s) + 1/s is defined) must be set
        // with mfg_adjoint_set_sticking before the next run
        std::vector<double> s0(G, 0.0);
        std::vector<char> seen(G, 0);
        for (uint32_t f = 0; f < nF; ++f) {
            const int32_t g = group_of_facet[f];
            if (g < 0) continue;
            if (M.phys[f].sticking_param >= 0)
                throw Err(MFG_ERR_INVALID, "facet " + std::to_string(f) + " of sticking group " + std::to_string(g) +
                                               " has a time-dependent sticking");
            if (seen[g] && s0[g] != M.phys[f].sticking)
                throw Err(MFG_ERR_INVALID, "sticking group " + std::to_string(g) + " mixes facets of different sticking");
            s0[g] = M.phys[f].sticking;
            seen[g] = 1;
        }
        c->adj_s_ok = true;
        for (uint32_t g = 0; g < G; ++g)
            if (!(s0[g] > 0.0 && s0[g] < 1.0)) c->adj_s_ok = false;
        HIPCHK(hipMemcpyAsync(c->adj_s.p, s0.data(), G * sizeof(double), hipMemcpyHostToDevice, c->stream));
        c->features |= FEAT_ADJOINT;
        zero_tallies(c);
        reset_graph(c);
    });
}

int mfg_adjoint_set_sticking(mfg_context* c, const double* s, uint32_t G) {
    return guarded(c, [&] {
        if (!c->adj_groups || G != c->adj_groups || !s) throw Err(MFG_ERR_INVALID, "adjoint not set up or wrong group count");
        for (uint32_t g = 0; g < G; ++g)
            if (!(s[g] > 0.0 && s[g] < 1.0)) throw Err(MFG_ERR_INVALID, "group sticking must be in (0,1)");
        HIPCHK(hipMemcpy(c->adj_s.p, s, G * sizeof(double), hipMemcpyHostToDevice));
        c->adj_s_ok = true;
    });
}

int mfg_adjoint_read(mfg_context* c, double* j0, double* jac) {
    return guarded(c, [&] {
        if (!c->adj_groups) throw Err(MFG_ERR_STATE, "adjoint not set up");
        if (j0) HIPCHK(hipMemcpy(j0, c->adj_j0.p, c->adj_targets * sizeof(double), hipMemcpyDeviceToHost));
        if (jac) HIPCHK(hipMemcpy(jac, c->adj_jac.p, (size_t)c->adj_targets * c->adj_groups * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int mfg_adjoint_backward(mfg_context* c, const double* up, double* grad) {
    return guarded(c, [&] {
        if (!c->adj_groups) throw Err(MFG_ERR_STATE, "adjoint not set up");
        if (!up || !grad) throw Err(MFG_ERR_INVALID, "null buffer");
        HIPCHK(hipMemcpyAsync(c->adj_up.p, up, c->adj_targets * sizeof(double), hipMemcpyHostToDevice, c->stream));
        double scale = c->adj_des ? 1.0 / (double)c->adj_des : 0.0;
        if (mfg_launch_backward(c->adj_jac.p, c->adj_up.p, c->adj_grad.p, c->adj_targets, c->adj_groups, scale, c->stream))
            throw Err(MFG_ERR_DEVICE, "backward launch failed");
        HIPCHK(hipMemcpyAsync(grad, c->adj_grad.p, c->adj_groups * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
