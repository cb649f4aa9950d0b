/*
 * mfcpu.cpp -- CPU oracle: FP64 restatement of MolFlow+'s ParticleTracer.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and timed CPU baseline); the
 * product library never links or calls this file.
 *
 * What it restates, function by function (reference paths):
 *   SimulationMCStep ............ src/Simulation/Particle.cpp:207-396
 *   StartFromSource ............. Particle.cpp:407-685
 *   PerformBounce ............... Particle.cpp:691-838
 *   PerformScatter .............. Particle.cpp:840-886, Physics.cpp:41-63
 *   RecordAbsorb ................ Particle.cpp:888-924
 *   RecordHistograms ............ Particle.cpp:926-975
 *   RecordHitOnTexture .......... Particle.cpp:977-1004
 *   RecordDirectionVector ....... Particle.cpp:1006-1023
 *   ProfileFacet ................ Particle.cpp:1025-1085
 *   LogHit ...................... Particle.cpp:1087-1102
 *   RecordAngleMap .............. Particle.cpp:1104-1136
 *   UpdateVelocity .............. Particle.cpp:1138-1163
 *   IncreaseFacetCounter ........ Particle.cpp:1193-1230
 *   RegisterTransparentPass ..... Particle.cpp:1232-1257
 *   PerformTeleport ............. Particle.cpp:76-183
 *   LookupMomentIndex ........... Particle.cpp:1364-1379
 *   GetStickingAt/Opacity/Temp .. src/Simulation/SimulationMC.cpp:15-39
 *   Maxwell CDF ................. src/Simulation/CDFGeneration.cpp:55-83
 *   GenerateRandomVelocity,
 *   GenerateDesorptionTime,
 *   TreatMovingFacet ............ src/Simulation/Physics.cpp:4-39
 *   Generate_ID ................. src/Simulation/IDGeneration.cpp:45-206
 *   PrepareToRun, intervals,
 *   CalcTotalOutgassing ......... src/Simulation/MolflowSimGeom.cpp:194-335
 *   ParseUserMoment ............. src/TimeMoments.cpp:58-83
 *   texture cell increments ..... src/Initializer.cpp:433-456
 * The src_shared pieces that the reference calls but that are absent from
 * the checkout (BVH Intersect, IsInFacet, PolarToCartesian, InterpolateX/Y,
 * MersenneTwister) are restated from their documented roles:
 *   - facet frame: N = Newell normal, nU = unit(v1-v0), nV = N x nU; the
 *     (u,v) rectangle is the bounding box of the polygon in that frame;
 *   - ray/facet test: plane hit with explicit fma dot products, then
 *     even-odd crossing test in normalised (u,v); a one-sided facet is only
 *     seen from the side its normal points to;
 *   - RNG: Philox4x32-10 (Salmon et al., SC'11) keyed by the seed, counter
 *     (particle index, block); uniforms are (x + 0.5) * 2^-32.
 * The GPU library follows the same specification so that seeded runs draw
 * the same random numbers in the same order.
 */
#include "mfcpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kBoltz = 1.38e-23;        // Particle.cpp:427,446,454
constexpr double kMbarlsToPam3s = 0.1;     // MolflowTypes.h MBARLS_TO_PAM3S
constexpr int kProfileSize = MFG_PROFILE_SIZE;
constexpr int kMaxwellPoints = 100;        // CDFGeneration.cpp:4
constexpr size_t kLeakCap = 2048;

// ------------------------------------------------------------------ vectors
struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// The dot product used by the ray/facet test: fixed fma order.
inline double fdot(V3 a, V3 b) { return std::fma(a.x, b.x, std::fma(a.y, b.y, a.z * b.z)); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { double n = norm(a); return n > 0 ? (1.0 / n) * a : a; }

// ------------------------------------------------------------------ Philox4x32-10
void philox(const uint32_t c[4], const uint32_t k[2], uint32_t o[4]) {
    uint32_t x0 = c[0], x1 = c[1], x2 = c[2], x3 = c[3], k0 = k[0], k1 = k[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x0, p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0, y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1, y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o[0] = x0; o[1] = x1; o[2] = x2; o[3] = x3;
}

inline double u32_to_unit(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
// Texture cells along one side for a precise size: the ceiling, except
// that a size above a whole number by rounding noise only (relative 1e-9)
// counts as that number. The rule of MolFlow's own (absent) facet code,
// as its stored results show it: test case 02c stores texDimY
// 5.000000000000002 with 5 rows of cells.
inline int texture_cells(double precise) {
    int n = (int)std::ceil(precise);
    if (n > 1 && precise - (n - 1) <= 1e-9 * precise) --n;
    return n;
}

struct Rng {
    uint32_t key[2];
    uint32_t pid_lo, pid_hi;
    uint32_t block = 0;
    uint32_t buf[4];
    int avail = 0;
    void init(uint64_t seed, uint64_t pid) {
        key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32);
        pid_lo = (uint32_t)pid; pid_hi = (uint32_t)(pid >> 32);
        block = 0; avail = 0;
    }
    double rnd() {
        if (avail == 0) {
            uint32_t c[4] = {pid_lo, pid_hi, block++, 0u};
            philox(c, key, buf);
            avail = 4;
        }
        return u32_to_unit(buf[4 - avail--]);
    }
    // Opacity coin of one (event, facet) pair: independent of the stream.
    double coin(uint32_t event, uint32_t facet) const {
        uint32_t c[4] = {pid_lo, pid_hi, event, 0x80000000u | facet};
        uint32_t o[4];
        philox(c, key, o);
        return u32_to_unit(o[0]);
    }
};

// ------------------------------------------------------------------ tables
struct XY { double x, y; };

// Linear interpolation of y at x over sorted points, optionally in log10
// space; outside the table the end value is held unless extrapolation is
// allowed (the InterpolateY role, used at SimulationMC.cpp:20).
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
        size_t lo = 0, hi = n - 1;  // t[lo].x <= x < t[hi].x
        while (hi - lo > 1) { size_t mid = (lo + hi) / 2; if (t[mid].x <= x) lo = mid; else hi = mid; }
        i = lo;
    }
    double x0 = t[i].x, x1 = t[i + 1].x, y0 = t[i].y, y1 = t[i + 1].y;
    if (logx) { x0 = std::log10(x0); x1 = std::log10(x1); x = std::log10(x); }
    if (logy) { y0 = std::log10(y0); y1 = std::log10(y1); }
    double y = (x1 == x0) ? y0 : y0 + (x - x0) * (y1 - y0) / (x1 - x0);
    return logy ? std::pow(10.0, y) : y;
}

// Inverse lookup: x where the (monotone) y column reaches `y`, linear, with
// extrapolation past the last point (InterpolateX role, Physics.cpp:5,13).
double interp_x(double y, const XY* t, size_t n) {
    if (n == 1) return t[0].x;
    size_t i;
    if (y < t[0].y) i = 0;
    else if (y >= t[n - 1].y) i = n - 2;
    else {
        size_t lo = 0, hi = n - 1;  // t[lo].y <= y < t[hi].y
        while (hi - lo > 1) { size_t mid = (lo + hi) / 2; if (t[mid].y <= y) lo = mid; else hi = mid; }
        i = lo;
    }
    double y0 = t[i].y, y1 = t[i + 1].y, x0 = t[i].x, x1 = t[i + 1].x;
    if (y1 == y0) return x0;
    return x0 + (y - y0) * (x1 - x0) / (y1 - y0);
}

std::vector<XY> maxwell_cdf(double gas_mass) {  // CDFGeneration.cpp:55-83 at T = 1 K
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

std::vector<double> parse_user_moment(const char* content, double window) {  // TimeMoments.cpp:58-83
    std::vector<double> out;
    double b, s, e;
    int nb = std::sscanf(content, "%lf,%lf,%lf", &b, &s, &e);
    if (nb == 1 && b >= 0.0) out.push_back(b);
    else if (nb == 3 && b >= 0.0 && e > b && s < (e - b)) {
        if (!(s < window))
            for (double t = b; t <= e; t += s) out.push_back(t);
    }
    return out;
}

bool is_equal(double a, double b) { return std::fabs(a - b) < 1e-10; }

// Generate_ID (IDGeneration.cpp:45-206): cumulative desorbed gas (Pa*m3)
// versus time, from t=0 to latestMoment.
std::vector<XY> integrated_desorption(const XY* par, size_t np, bool logx, bool logy, double latest) {
    std::vector<XY> ID, og;
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
        if (!is_equal(og.back().x, latest))
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
                double lpv = prev_t > 0 ? std::log10(prev_v) : -99;  // reference tests the time here
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

// ------------------------------------------------------------------ geometry
// DES_ANGLEMAP source distribution of one facet, built from its recorded
// incident-angle map like InitializeAngleMap (src/Simulation/
// MolflowSimFacet.cpp:92-222). Lines of the lower and higher theta parts
// are concatenated: line t < L is lower line t, line t >= L is higher line
// t - L.
struct AngleSource {
    int W = 0, L = 0, H = 0;
    double limit = 0;
    std::vector<uint64_t> pdf;        // (L+H)*W recorded counts
    std::vector<uint64_t> line_sum;   // phi line sums (phi_pdfsums_*Theta)
    std::vector<double> theta_cdf;    // CDF at line midpoints (theta_CDF_lower/higher)
    std::vector<double> phi_cdf;      // (L+H)*W midpoint CDFs (phi_CDFs_*Theta)
    std::vector<double> phi_pdf;      // (L+H)*W normalised pdfs (phi_pdfs_*Theta)
    uint64_t sum_lower = 0, sum_higher = 0;   // theta_CDFsum_lower / _higher
    double lower_ratio = 0;                   // thetaLowerRatio
};

struct Facet {
    mfg_facet_desc d;
    V3 O, U, V, N, nU, nV, center;
    double invU2, invV2, area, sqrt_temp, max_speed;
    std::vector<double> poly;   // normalised (u,v) pairs
    V3 lo, hi;
    int tex_w = 0, tex_h = 0;
    std::vector<double> tex_inc;
    int64_t tex_off = -1, dir_off = -1, prof_off = -1, am_off = -1, fh_off = -1;
    int fh_bounce_bins = 0, fh_dist_bins = 0, fh_time_bins = 0;
    // source table
    std::vector<double> og_cdf;  // cumulative outgassing map (Pa*m3/s)
    int id_index = -1;
    AngleSource am_src;
};

bool in_poly(const std::vector<double>& p, double u, double v) {
    size_t n = p.size() / 2;
    bool inside = false;
    for (size_t i = 0, j = n - 1; i < n; j = i++) {
        double ui = p[2 * i], vi = p[2 * i + 1], uj = p[2 * j], vj = p[2 * j + 1];
        if ((vi > v) != (vj > v)) {
            double x = (uj - ui) * (v - vi) / (vj - vi) + ui;
            if (u < x) inside = !inside;
        }
    }
    return inside;
}

void build_angle_source(Facet& F, const std::vector<uint64_t>& pdf) {
    AngleSource& A = F.am_src;
    A.W = (int)F.d.am_phi_width; A.L = (int)F.d.am_theta_lower_res; A.H = (int)F.d.am_theta_higher_res;
    A.limit = F.d.am_theta_limit;
    const int W = A.W, L = A.L, H = A.H, T = L + H;
    if (W <= 0 || T <= 0) throw std::runtime_error("DES_ANGLEMAP facet with an empty angle map shape");
    if (pdf.size() != (size_t)T * W) throw std::runtime_error("DES_ANGLEMAP map size does not match its shape");
    A.pdf = pdf;
    A.line_sum.assign(T, 0);
    A.theta_cdf.assign(T, 0.0);
    A.phi_cdf.assign((size_t)T * W, 0.0);
    A.phi_pdf.assign((size_t)T * W, 0.0);
    // first pass: phi line sums, lower and total sums
    A.sum_lower = 0;
    for (int t = 0; t < L; ++t) {
        for (int p = 0; p < W; ++p) A.line_sum[t] += pdf[(size_t)t * W + p];
        A.sum_lower += A.line_sum[t];
    }
    A.sum_higher = A.sum_lower;   // the higher sum includes the lower part
    for (int t = L; t < T; ++t) {
        for (int p = 0; p < W; ++p) A.line_sum[t] += pdf[(size_t)t * W + p];
        A.sum_higher += A.line_sum[t];
    }
    if (A.sum_higher == 0) throw std::runtime_error("DES_ANGLEMAP facet with an all-zero angle map");
    A.lower_ratio = (double)A.sum_lower / (double)A.sum_higher;
    // second pass: midpoint CDFs and normalised phi pdfs
    const double tn = 1.0 / (double)A.sum_higher;
    for (int t = 0; t < T; ++t) {
        const bool first_of_part = (t == 0) || (t == L);
        if (t == 0) A.theta_cdf[t] = 0.5 * (double)A.line_sum[0] * tn;
        else if (t == L) A.theta_cdf[t] = A.lower_ratio + 0.5 * (double)A.line_sum[L] * tn;
        else A.theta_cdf[t] = A.theta_cdf[t - 1] + (double)(A.line_sum[t - 1] + A.line_sum[t]) * 0.5 * tn;
        (void)first_of_part;
        const double pn = 1.0 / (double)A.line_sum[t];
        for (int p = 0; p < W; ++p) {
            const size_t i = (size_t)t * W + p;
            if (A.line_sum[t] == 0) {   // empty line: uniform phi
                A.phi_cdf[i] = (0.5 + (double)p) / (double)W;
                A.phi_pdf[i] = 1.0 / (double)W;
            } else {
                if (p == 0) A.phi_cdf[i] = 0.5 * (double)pdf[(size_t)t * W] * pn;
                else A.phi_cdf[i] = A.phi_cdf[i - 1] + (double)(pdf[i - 1] + pdf[i]) * 0.5 * pn;
                A.phi_pdf[i] = (double)pdf[i] * pn;
            }
        }
    }
}

// Helpers whose reference versions live in the absent shared library,
// restated from their documented roles:
//   lower_index: index of the last CDF value <= key, -1 below the first;
//   weighed_lower_index: the same on the line (1-w)*A + w*B;
//   weigh(a, b, w) = a + (b - a) * w; IDX: circular index.
int lower_index(double key, const double* A, int n) {
    return (int)(std::upper_bound(A, A + n, key) - A) - 1;
}
double weigh(double a, double b, double w) { return a + (b - a) * w; }
int weighed_lower_index(double key, double w, const double* A, const double* B, int n) {
    int lo = 0, hi = n;   // first index whose weighed value is > key
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (weigh(A[mid], B[mid], w) <= key) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}
int circ(int i, int n) { return ((i % n) + n) % n; }

// GetTheta / GetPhi (AnglemapGeneration.cpp): bin index -> angle
double am_theta(const AngleSource& A, double ti) {
    if ((size_t)ti < (size_t)A.L) return A.limit * ti / (double)A.L;
    return A.limit + (kPi / 2.0 - A.limit) * (ti - (double)A.L) / (double)A.H;
}
double am_phi(const AngleSource& A, double pi_) {
    double w = (double)A.W;
    double c = (pi_ < w) ? pi_ : pi_ - w;
    return -kPi + 2.0 * kPi * c / w;
}
// GetPhiNormalizedPdfValue
double am_phi_pdf(const AngleSource& A, double ti, int plo) {
    const int W = A.W, T = A.L + A.H;
    const size_t p = (size_t)circ(plo, W);
    if (ti < 0.5) return A.phi_pdf[p];
    if (ti > (double)T - 0.5) return A.phi_pdf[(size_t)(T - 1) * W + p];
    size_t tl = (size_t)(ti - 0.5);
    double ov = ti - 0.5 - (double)tl;
    return weigh(A.phi_pdf[tl * W + p], A.phi_pdf[(tl + 1) * W + p], ov);
}
// GetPhiCDFValue
double am_phi_cdf(const AngleSource& A, double ti, int plo) {
    const int W = A.W, L = A.L, T = A.L + A.H;
    auto line = [&](size_t t) { return (plo < W) ? A.phi_cdf[t * W + plo] : 1.0 + A.phi_cdf[t * W]; };
    if (ti < (double)L) {
        if (ti < 0.5) return line(0);
        if (ti > (double)L - 0.5) return line((size_t)(L - 1));
    } else {
        if (ti < (double)L + 0.5) return line((size_t)L);
        if (ti > (double)T - 0.5) return line((size_t)(T - 1));
    }
    size_t tl = (size_t)(ti - 0.5);
    double ov = ti - 0.5 - (double)tl;
    return weigh(line(tl), line(tl + 1), ov);
}

// GenerateThetaFromAngleMap: theta, lower line index (-1 .. T-1), overshoot
void am_generate_theta(const AngleSource& A, double r, double& theta, int& tlo, double& ov) {
    const int L = A.L, H = A.H;
    const double* cdf = A.theta_cdf.data();
    auto quad = [&](int t, double c) {   // linear pdf between line midpoints t and t+1
        double step = am_theta(A, (double)t + 1.5) - am_theta(A, (double)t + 0.5);
        double b = (double)A.line_sum[t] / (double)A.sum_higher / step;
        double a = 0.5 * ((double)A.line_sum[t + 1] - (double)A.line_sum[t]) / (double)A.sum_higher / (step * step);
        double dy = r - c;
        double dx = (-b + std::sqrt(b * b + 4 * a * dy)) / (2 * a);
        return dx / step;
    };
    if (r < A.lower_ratio) {
        tlo = lower_index(r, cdf, L);
        if (tlo == -1) {
            ov = 0.5 * r / cdf[0];
            theta = am_theta(A, ov);
        } else if (tlo == L - 1) {
            ov = 0.5 * (r - cdf[tlo]) / (A.lower_ratio - cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        } else {
            if (A.line_sum[tlo] == A.line_sum[tlo + 1]) ov = (r - cdf[tlo]) / (cdf[tlo + 1] - cdf[tlo]);
            else ov = quad(tlo, cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        }
    } else {
        tlo = L + lower_index(r, cdf + L, H);
        if (tlo == L - 1) {
            ov = 0.5 * (r - A.lower_ratio) / (cdf[L] - A.lower_ratio);
            theta = am_theta(A, (double)L + ov);
        } else if (tlo == L + H - 1) {
            ov = 0.5 * (r - cdf[tlo]) / (1.0 - cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        } else {
            if (A.line_sum[tlo] == A.line_sum[tlo + 1]) ov = (r - cdf[tlo]) / (cdf[tlo + 1] - cdf[tlo]);
            else ov = quad(tlo, cdf[tlo]);
            theta = am_theta(A, (double)tlo + 0.5 + ov);
        }
    }
}

// GeneratePhiFromAngleMap
double am_generate_phi(const AngleSource& A, int tlo, double ov, double r) {
    const int W = A.W, L = A.L, T = A.L + A.H;
    if (W == 1) return -kPi + 2.0 * kPi * r;
    int plo;
    double w;
    const double* C = A.phi_cdf.data();
    auto edge = [&](int t) {   // a single line, shifted by its first midpoint value
        r += C[(size_t)t * W];
        plo = lower_index(r, C + (size_t)t * W, W);
        w = ov;
    };
    auto between = [&](int t) {   // weighed between lines t and t+1
        double div = (double)A.line_sum[t] * (1.0 - ov) + (double)A.line_sum[t + 1] * ov;
        w = div > 0.0 ? (ov * (double)A.line_sum[t + 1]) / div : ov;
        r += weigh(C[(size_t)t * W], C[(size_t)(t + 1) * W], w);
        plo = weighed_lower_index(r, w, C + (size_t)t * W, C + (size_t)(t + 1) * W, W);
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
    double phi, pov;
    if (plo == -1) {
        pov = 0.5 * r / am_phi_cdf(A, ti, 0);
        phi = am_phi(A, pov);
    } else {
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
            if (D < 0) pov = (r - c) / (am_phi_cdf(A, ti, circ(plo + 1, W)) - c);
            else pov = ((-b + std::sqrt(b * b + 4 * a * dy)) / (2 * a)) / step;
        }
        phi = am_phi(A, (double)plo + 0.5 + pov);
    }
    return phi;
}

double poly_area(const std::vector<double>& p) {
    size_t n = p.size() / 2;
    double a = 0;
    for (size_t i = 0, j = n - 1; i < n; j = i++) a += p[2 * j] * p[2 * i + 1] - p[2 * i] * p[2 * j + 1];
    return 0.5 * std::fabs(a);
}

// Sutherland-Hodgman clip of a polygon against an axis-aligned box.
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

// ------------------------------------------------------------------ BVH (binned SAH, 2 facets per leaf)
struct BNode { V3 lo, hi; int left = -1, right = -1, first = 0, count = 0; };
struct Bvh { std::vector<BNode> nodes; std::vector<int> prims; };

void build_bvh(Bvh& b, const std::vector<Facet>& F, std::vector<int> prims, int leaf) {
    b.nodes.clear(); b.prims.clear();
    if (prims.empty()) return;
    struct Item { int lo, hi, node; };
    std::vector<int>& P = prims;
    b.nodes.push_back(BNode());
    std::vector<Item> stack{{0, (int)P.size(), 0}};
    while (!stack.empty()) {
        Item it = stack.back(); stack.pop_back();
        BNode nd;
        nd.lo = {1e300, 1e300, 1e300}; nd.hi = {-1e300, -1e300, -1e300};
        V3 clo = nd.lo, chi = nd.hi;
        for (int i = it.lo; i < it.hi; ++i) {
            const Facet& f = F[P[i]];
            nd.lo = {std::min(nd.lo.x, f.lo.x), std::min(nd.lo.y, f.lo.y), std::min(nd.lo.z, f.lo.z)};
            nd.hi = {std::max(nd.hi.x, f.hi.x), std::max(nd.hi.y, f.hi.y), std::max(nd.hi.z, f.hi.z)};
            V3 c = 0.5 * (f.lo + f.hi);
            clo = {std::min(clo.x, c.x), std::min(clo.y, c.y), std::min(clo.z, c.z)};
            chi = {std::max(chi.x, c.x), std::max(chi.y, c.y), std::max(chi.z, c.z)};
        }
        int n = it.hi - it.lo;
        int split = -1;
        if (n > leaf) {
            double best = 1e300; int bestAxis = -1, bestBin = -1;
            const int NB = 16;
            for (int ax = 0; ax < 3; ++ax) {
                double a0 = ax == 0 ? clo.x : ax == 1 ? clo.y : clo.z;
                double a1 = ax == 0 ? chi.x : ax == 1 ? chi.y : chi.z;
                if (a1 - a0 <= 0) continue;
                int cnt[NB] = {0}; V3 blo[NB], bhi[NB];
                for (int k = 0; k < NB; ++k) { blo[k] = {1e300, 1e300, 1e300}; bhi[k] = {-1e300, -1e300, -1e300}; }
                for (int i = it.lo; i < it.hi; ++i) {
                    const Facet& f = F[P[i]];
                    V3 c = 0.5 * (f.lo + f.hi);
                    double cv = ax == 0 ? c.x : ax == 1 ? c.y : c.z;
                    int k = std::min(NB - 1, (int)((cv - a0) / (a1 - a0) * NB));
                    cnt[k]++;
                    blo[k] = {std::min(blo[k].x, f.lo.x), std::min(blo[k].y, f.lo.y), std::min(blo[k].z, f.lo.z)};
                    bhi[k] = {std::max(bhi[k].x, f.hi.x), std::max(bhi[k].y, f.hi.y), std::max(bhi[k].z, f.hi.z)};
                }
                auto area = [](V3 lo, V3 hi) { V3 d = hi - lo; if (d.x < 0) return 0.0; return d.x * d.y + d.y * d.z + d.z * d.x; };
                for (int s = 1; s < NB; ++s) {
                    V3 l0 = {1e300, 1e300, 1e300}, l1 = {-1e300, -1e300, -1e300}, r0 = l0, r1 = l1;
                    int nl = 0, nr = 0;
                    for (int k = 0; k < s; ++k) if (cnt[k]) { nl += cnt[k]; l0 = {std::min(l0.x, blo[k].x), std::min(l0.y, blo[k].y), std::min(l0.z, blo[k].z)}; l1 = {std::max(l1.x, bhi[k].x), std::max(l1.y, bhi[k].y), std::max(l1.z, bhi[k].z)}; }
                    for (int k = s; k < NB; ++k) if (cnt[k]) { nr += cnt[k]; r0 = {std::min(r0.x, blo[k].x), std::min(r0.y, blo[k].y), std::min(r0.z, blo[k].z)}; r1 = {std::max(r1.x, bhi[k].x), std::max(r1.y, bhi[k].y), std::max(r1.z, bhi[k].z)}; }
                    if (!nl || !nr) continue;
                    double cost = nl * area(l0, l1) + nr * area(r0, r1);
                    if (cost < best) { best = cost; bestAxis = ax; bestBin = s; }
                }
            }
            if (bestAxis >= 0) {
                double a0 = bestAxis == 0 ? clo.x : bestAxis == 1 ? clo.y : clo.z;
                double a1 = bestAxis == 0 ? chi.x : bestAxis == 1 ? chi.y : chi.z;
                auto mid = std::partition(P.begin() + it.lo, P.begin() + it.hi, [&](int id) {
                    V3 c = 0.5 * (F[id].lo + F[id].hi);
                    double cv = bestAxis == 0 ? c.x : bestAxis == 1 ? c.y : c.z;
                    return std::min(NB - 1, (int)((cv - a0) / (a1 - a0) * NB)) < bestBin;
                });
                split = (int)(mid - P.begin());
            } else {
                split = it.lo + n / 2;  // coincident centroids: median split
            }
            if (split <= it.lo || split >= it.hi) split = it.lo + n / 2;
        }
        if (split < 0) {
            nd.first = (int)b.prims.size(); nd.count = n;
            for (int i = it.lo; i < it.hi; ++i) b.prims.push_back(P[i]);
            b.nodes[it.node] = nd;
        } else {
            nd.left = (int)b.nodes.size(); b.nodes.push_back(BNode());
            nd.right = (int)b.nodes.size(); b.nodes.push_back(BNode());
            b.nodes[it.node] = nd;
            stack.push_back({split, it.hi, nd.right});
            stack.push_back({it.lo, split, nd.left});
        }
    }
}

// ------------------------------------------------------------------ model
struct Param { std::vector<XY> pts; bool logx, logy; };

struct Model {
    mfg_settings s;
    std::vector<Facet> F;
    std::vector<Param> params;
    std::vector<double> mom_start, mom_end;   // intervalCache
    std::vector<std::vector<XY>> IDs;
    std::vector<XY> maxwell;
    std::vector<std::vector<uint64_t>> am_pdf;  // DES_ANGLEMAP input maps (unused by sampling here)
    double latest = 0, total_des = 0, final_rate = 0, final_rate_pam3s = 0;
    int n_struct = 1;
    std::vector<Bvh> bvh;
    std::vector<std::vector<int>> struct_facets;
    // source table: cumulative weights over desorbing facets in facet order
    std::vector<double> src_cum;
    std::vector<int> src_facet;
    // layout
    uint64_t n_tex = 0, n_dir = 0, n_prof = 0, n_am = 0, n_fh = 0;
    int gh_bounce = 0, gh_dist = 0, gh_time = 0;
    uint32_t n_slots = 1;
    double c_factor = 1.0;  // 1.0 with Maxwell speeds, else 1.1781 (Particle.cpp:752)
};

int hist_bounce_bins(int rec, uint32_t bin, uint32_t max) { return rec && bin ? (int)(max / bin + 2) : 0; }
int hist_real_bins(int rec, double bin, double max) { return rec && bin > 0 ? (int)((size_t)(max / bin) + 1) : 0; }

double param_at(const Model& m, int id, double t, double constant) {
    if (id < 0) return constant;
    const Param& p = m.params[id];
    return interp_y(t, p.pts.data(), p.pts.size(), p.logx, p.logy, false);
}

int moment_index(const Model& m, double t) {  // LookupMomentIndex: 1-based slot, -1 outside
    const auto& S = m.mom_start;
    if (S.empty()) return -1;
    size_t lb = std::lower_bound(S.begin(), S.end(), t) - S.begin();  // first start >= t
    if (lb != 0) --lb;
    if (S[lb] <= t && t < m.mom_end[lb]) return (int)lb + 1;
    return -1;
}

void load_model(Model& M, const mfg_model_view* v) {
    if (!v || v->struct_size != sizeof(mfg_model_view)) throw std::runtime_error("model view size mismatch");
    if (!v->settings) throw std::runtime_error("missing settings");
    M.s = *v->settings;
    M.n_struct = std::max<uint32_t>(1, v->n_structures);
    for (uint32_t p = 0; p < v->n_params; ++p) {
        Param P;
        for (uint32_t k = v->param_offset[p]; k < v->param_offset[p + 1]; ++k) P.pts.push_back({v->param_x[k], v->param_y[k]});
        P.logx = v->param_log_x ? v->param_log_x[p] : 0;
        P.logy = v->param_log_y ? v->param_log_y[p] : 0;
        M.params.push_back(P);
    }
    // moments sorted by start time (ParseAndCheckUserMoments)
    std::vector<std::pair<double, double>> mom;
    for (uint32_t i = 0; i < v->n_moments; ++i) mom.push_back({v->moment_time[i], v->moment_window[i]});
    std::sort(mom.begin(), mom.end(), [](auto a, auto b) { return a.first - .5 * a.second < b.first - .5 * b.second; });
    for (auto& mm : mom) { M.mom_start.push_back(mm.first - .5 * mm.second); M.mom_end.push_back(mm.first + .5 * mm.second); }
    M.latest = mom.empty() ? M.s.time_window * .5 : mom.back().first + .5 * mom.back().second;
    M.n_slots = 1 + (uint32_t)mom.size();
    M.maxwell = maxwell_cdf(M.s.gas_mass);
    M.c_factor = M.s.use_maxwell ? 1.0 : 1.1781;

    M.F.resize(v->n_facets);
    std::vector<int> id_of_param(M.params.size(), -1);
    for (uint32_t f = 0; f < v->n_facets; ++f) {
        Facet& F = M.F[f];
        F.d = v->facets[f];
        uint32_t a = v->facet_vertex_offset[f], b = v->facet_vertex_offset[f + 1];
        if (b - a < 3) throw std::runtime_error("facet with fewer than 3 vertices");
        std::vector<V3> P;
        for (uint32_t k = a; k < b; ++k) {
            uint32_t vi = v->facet_vertex_index[k];
            if (vi >= v->n_vertices) throw std::runtime_error("vertex index out of range");
            P.push_back({v->vertices[3 * vi], v->vertices[3 * vi + 1], v->vertices[3 * vi + 2]});
        }
        size_t n = P.size();
        V3 N = {0, 0, 0};
        for (size_t i = 0; i < n; ++i) {
            const V3& p = P[i]; const V3& q = P[(i + 1) % n];
            N.x += (p.y - q.y) * (p.z + q.z);
            N.y += (p.z - q.z) * (p.x + q.x);
            N.z += (p.x - q.x) * (p.y + q.y);
        }
        F.N = unit(N);
        F.nU = unit(P[1] - P[0]);
        F.nV = unit(cross(F.N, F.nU));
        double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
        std::vector<double> pu(n), pv(n);
        for (size_t i = 0; i < n; ++i) {
            V3 r = P[i] - P[0];
            pu[i] = dot(r, F.nU); pv[i] = dot(r, F.nV);
            umin = std::min(umin, pu[i]); umax = std::max(umax, pu[i]);
            vmin = std::min(vmin, pv[i]); vmax = std::max(vmax, pv[i]);
        }
        F.O = P[0] + umin * F.nU + vmin * F.nV;
        F.U = (umax - umin) * F.nU;
        F.V = (vmax - vmin) * F.nV;
        F.invU2 = 1.0 / dot(F.U, F.U);
        F.invV2 = 1.0 / dot(F.V, F.V);
        F.poly.resize(2 * n);
        for (size_t i = 0; i < n; ++i) {
            F.poly[2 * i] = (pu[i] - umin) / (umax - umin);
            F.poly[2 * i + 1] = (pv[i] - vmin) / (vmax - vmin);
        }
        double lu = norm(F.U), lv = norm(F.V);
        F.area = lu * lv * poly_area(F.poly);
        V3 c = {0, 0, 0};
        F.lo = P[0]; F.hi = P[0];
        for (auto& p : P) {
            c = c + p;
            F.lo = {std::min(F.lo.x, p.x), std::min(F.lo.y, p.y), std::min(F.lo.z, p.z)};
            F.hi = {std::max(F.hi.x, p.x), std::max(F.hi.y, p.y), std::max(F.hi.z, p.z)};
        }
        F.center = (1.0 / n) * c;
        F.d.opacity = std::min(1.0, std::max(0.0, F.d.opacity));
        if (F.d.super_dest) { F.d.opacity_param = -1; F.d.sticking_param = -1; }  // InitializeLinkFacet
        F.sqrt_temp = std::sqrt(F.d.temperature);
        // Upper edge of velocity profiles: 4 x most probable speed at the facet temperature.
        F.max_speed = 4.0 * std::sqrt(2.0 * 8.31 * F.d.temperature / 0.001 / M.s.gas_mass);
        for (int pid : {F.d.sticking_param, F.d.opacity_param, F.d.outgassing_param, F.d.temperature_param})
            if (pid >= (int)M.params.size()) throw std::runtime_error("facet refers to a missing parameter");
        // textures (Initializer.cpp:433-456)
        if (F.d.is_textured && F.d.tex_width_precise > 0 && F.d.tex_height_precise > 0) {
            F.tex_w = texture_cells(F.d.tex_width_precise);
            F.tex_h = texture_cells(F.d.tex_height_precise);
            double rw = lu / F.d.tex_width_precise, rh = lv / F.d.tex_height_precise;
            double full = rw * rh, two = F.d.is_2sided ? 2.0 : 1.0;
            F.tex_inc.resize((size_t)F.tex_w * F.tex_h);
            for (int j = 0; j < F.tex_h; ++j)
                for (int i = 0; i < F.tex_w; ++i) {
                    double u0 = i / F.d.tex_width_precise, u1 = (i + 1) / F.d.tex_width_precise;
                    double v0 = j / F.d.tex_height_precise, v1 = (j + 1) / F.d.tex_height_precise;
                    double a2 = clip_area(F.poly, u0, u1, v0, v1) * lu * lv;
                    double inc;
                    if (a2 >= full * (1.0 - 1e-9)) inc = 1.0 / (full * two);
                    else if (a2 > 0) inc = 1.0 / (a2 * two);
                    else inc = 0.0;
                    F.tex_inc[(size_t)j * F.tex_w + i] = inc;
                }
        }
        // outgassing map CDF (MolflowSimFacet.cpp:224-239)
        if (F.d.use_outgassing_map && v->og_map_offset && v->og_map_values) {
            double acc = 0;
            for (uint64_t k = v->og_map_offset[f]; k < v->og_map_offset[f + 1]; ++k) { acc += v->og_map_values[k]; F.og_cdf.push_back(acc); }
            if (F.og_cdf.size() != (size_t)F.d.og_map_width * F.d.og_map_height) throw std::runtime_error("outgassing map size mismatch");
        } else F.d.use_outgassing_map = 0;
        if (F.d.desorb_type == MFG_DES_ANGLEMAP) {
            if (!v->am_pdf_offset || !v->am_pdf) throw std::runtime_error("DES_ANGLEMAP facet without a recorded angle map");
            std::vector<uint64_t> pdf(v->am_pdf + v->am_pdf_offset[f], v->am_pdf + v->am_pdf_offset[f + 1]);
            build_angle_source(F, pdf);
        }
        // integrated desorption per distinct outgassing parameter
        if (F.d.outgassing_param >= 0) {
            int p = F.d.outgassing_param;
            if (id_of_param[p] < 0) {
                id_of_param[p] = (int)M.IDs.size();
                const Param& P2 = M.params[p];
                M.IDs.push_back(integrated_desorption(P2.pts.data(), P2.pts.size(), P2.logx, P2.logy, M.latest));
            }
            F.id_index = id_of_param[p];
        }
    }
    // CalcTotalOutgassing (MolflowSimGeom.cpp:290-335) and the source table
    double acc = 0;
    for (uint32_t f = 0; f < v->n_facets; ++f) {
        Facet& F = M.F[f];
        if (F.d.desorb_type == MFG_DES_NONE) continue;
        double kT = kBoltz * F.d.temperature, w;
        if (F.d.use_outgassing_map) {
            double sum = F.og_cdf.empty() ? 0 : F.og_cdf.back();
            M.total_des += M.latest * sum / kT;
            M.final_rate += sum / kT; M.final_rate_pam3s += sum;
            if (!(F.d.total_outgassing > 0)) continue;
            w = M.latest * F.d.total_outgassing / kT;
        } else if (F.d.outgassing_param < 0) {
            w = M.latest * F.d.outgassing / kT;
            M.total_des += w; M.final_rate += F.d.outgassing / kT; M.final_rate_pam3s += F.d.outgassing;
        } else {
            w = M.IDs[F.id_index].back().y / kT;
            M.total_des += w;
            const Param& P2 = M.params[F.d.outgassing_param];
            double last = P2.pts.back().y * kMbarlsToPam3s;
            M.final_rate += last / kT; M.final_rate_pam3s += last;
        }
        if (w <= 0) continue;
        if (F.d.super_idx == -1) throw std::runtime_error("a facet in all structures can't desorb");
        acc += w;
        M.src_cum.push_back(acc);
        M.src_facet.push_back((int)f);
    }
    // layout
    M.gh_bounce = hist_bounce_bins(M.s.gh_record_bounce, M.s.gh_bounce_binsize, M.s.gh_bounce_max);
    M.gh_dist = hist_real_bins(M.s.gh_record_distance, M.s.gh_distance_binsize, M.s.gh_distance_max);
    M.gh_time = hist_real_bins(M.s.gh_record_time, M.s.gh_time_binsize, M.s.gh_time_max);
    for (auto& F : M.F) {
        if (F.tex_w > 0) { F.tex_off = (int64_t)M.n_tex; M.n_tex += (uint64_t)F.tex_w * F.tex_h; }
        if (F.tex_w > 0 && F.d.count_dir) { F.dir_off = (int64_t)M.n_dir; M.n_dir += (uint64_t)F.tex_w * F.tex_h; }
        if (F.d.profile_type != MFG_PROFILE_NONE) { F.prof_off = (int64_t)M.n_prof; M.n_prof += kProfileSize; }
        if (F.d.am_record) { F.am_off = (int64_t)M.n_am; M.n_am += (uint64_t)F.d.am_phi_width * (F.d.am_theta_lower_res + F.d.am_theta_higher_res); }
        F.fh_bounce_bins = hist_bounce_bins(F.d.fh_record_bounce, F.d.fh_bounce_binsize, F.d.fh_bounce_max);
        F.fh_dist_bins = hist_real_bins(F.d.fh_record_distance, F.d.fh_distance_binsize, F.d.fh_distance_max);
        F.fh_time_bins = hist_real_bins(F.d.fh_record_time, F.d.fh_time_binsize, F.d.fh_time_max);
        int nb = F.fh_bounce_bins + F.fh_dist_bins + F.fh_time_bins;
        if (nb > 0) { F.fh_off = (int64_t)M.n_fh; M.n_fh += nb; }
    }
    // structures and BVHs
    M.struct_facets.assign(M.n_struct, {});
    for (uint32_t f = 0; f < v->n_facets; ++f) {
        int si = M.F[f].d.super_idx;
        if (si < -1 || si >= M.n_struct) throw std::runtime_error("facet structure index out of range");
        int td = M.F[f].d.teleport_dest;
        if (td < -1 || td > (int)v->n_facets) throw std::runtime_error("teleport destination out of range");
        if (si < 0) { for (auto& sf : M.struct_facets) sf.push_back((int)f); }
        else M.struct_facets[si].push_back((int)f);
        int sd = M.F[f].d.super_dest;
        if (sd && (sd - 1 >= M.n_struct || sd < 0)) throw std::runtime_error("invalid structure link");
    }
    M.bvh.resize(M.n_struct);
    for (int s = 0; s < M.n_struct; ++s) build_bvh(M.bvh[s], M.F, M.struct_facets[s], 2);
}

// ------------------------------------------------------------------ tallies
struct Tally {
    mfg_global_counters g{};
    std::vector<mfg_facet_counters> fc;
    std::vector<double> force;
    std::vector<mfg_texture_cell> tex;
    std::vector<mfg_direction_cell> dir;
    std::vector<mfg_profile_slice> prof;
    std::vector<uint64_t> am;
    std::vector<double> gh, fh;
    std::vector<mfg_log_entry> log;
    std::vector<double> leaks;
    uint64_t n_leaks = 0;
    size_t nF = 0, slots = 1;
    uint64_t ntex = 0, ndir = 0, nprof = 0, ngh = 0, nfh = 0;
    void init(const Model& M) {
        nF = M.F.size(); slots = M.n_slots;
        ntex = M.n_tex; ndir = M.n_dir; nprof = M.n_prof;
        ngh = (uint64_t)M.gh_bounce + M.gh_dist + M.gh_time; nfh = M.n_fh;
        fc.assign(slots * nF, mfg_facet_counters{});
        if (M.s.enable_force) force.assign(slots * nF * 9, 0.0);
        tex.assign(slots * ntex, mfg_texture_cell{});
        dir.assign(slots * ndir, mfg_direction_cell{});
        prof.assign(slots * nprof, mfg_profile_slice{});
        am.assign(M.n_am, 0);
        gh.assign(slots * ngh, 0.0);
        fh.assign(slots * nfh, 0.0);
    }
};

// ------------------------------------------------------------------ intersection
struct Hit { int facet = -1; double t = 1e300, u = 0, v = 0; };

// Plane + polygon test. Returns true with t,u,v on a hit.
inline bool facet_test(const Facet& F, V3 P, V3 D, double& t, double& u, double& v) {
    double denom = fdot(F.N, D);
    if (denom == 0.0) return false;
    if (!F.d.is_2sided && denom > 0.0) return false;
    V3 w = F.O - P;
    t = fdot(F.N, w) / denom;
    if (!(t > 0.0)) return false;
    V3 X = {std::fma(t, D.x, P.x), std::fma(t, D.y, P.y), std::fma(t, D.z, P.z)};
    V3 r = X - F.O;
    u = fdot(r, F.U) * F.invU2;
    v = fdot(r, F.V) * F.invV2;
    if (u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0) return false;
    return in_poly(F.poly, u, v);
}

inline bool better(double t, int f, const Hit& h) { return t < h.t || (t == h.t && f < h.facet); }

inline bool box_hit(const BNode& n, V3 P, V3 inv, double tmax) {
    double t0 = 0.0, t1 = tmax;
    double a = (n.lo.x - P.x) * inv.x, b = (n.hi.x - P.x) * inv.x; if (a > b) std::swap(a, b);
    t0 = std::max(t0, a); t1 = std::min(t1, b * (1 + 1e-12));
    a = (n.lo.y - P.y) * inv.y; b = (n.hi.y - P.y) * inv.y; if (a > b) std::swap(a, b);
    t0 = std::max(t0, a); t1 = std::min(t1, b * (1 + 1e-12));
    a = (n.lo.z - P.z) * inv.z; b = (n.hi.z - P.z) * inv.z; if (a > b) std::swap(a, b);
    t0 = std::max(t0, a); t1 = std::min(t1, b * (1 + 1e-12));
    return t0 <= t1 * (1 + 1e-12) + 1e-12;
}

struct TPass { int facet; double t, u, v; };

// Nearest hard hit of a ray in structure s. Facets with opacity < 1 are hard
// with probability `opacity` (coin per event and facet); the others are
// collected as transparent passes. opaque_only treats every facet as hard.
Hit intersect(const Model& M, int s, V3 P, V3 D, int skip, double time, const Rng* rng, uint32_t event,
              std::vector<TPass>* passes, bool brute, bool opaque_only, uint64_t* nodes_visited = nullptr,
              uint64_t* prims_visited = nullptr) {
    Hit h;
    auto test = [&](int f) {
        if (f == skip) return;
        const Facet& F = M.F[f];
        double t, u, v;
        if (prims_visited) ++*prims_visited;
        if (!facet_test(F, P, D, t, u, v)) return;
        bool hard = true;
        if (!opaque_only) {
            double op = F.d.opacity_param >= 0 ? param_at(M, F.d.opacity_param, time, F.d.opacity) : F.d.opacity;
            if (op < 1.0) hard = (op > 0.0) && rng->coin(event, (uint32_t)f) < op;
        }
        if (hard) { if (better(t, f, h)) { h.facet = f; h.t = t; h.u = u; h.v = v; } }
        else if (passes) passes->push_back({f, t, u, v});
    };
    if (brute) {
        for (int f : M.struct_facets[s]) test(f);
        return h;
    }
    const Bvh& B = M.bvh[s];
    if (B.nodes.empty()) return h;
    V3 inv = {1.0 / D.x, 1.0 / D.y, 1.0 / D.z};
    int stack[128]; int sp = 0; stack[sp++] = 0;
    while (sp) {
        const BNode& n = B.nodes[stack[--sp]];
        if (nodes_visited) ++*nodes_visited;
        if (!box_hit(n, P, inv, h.t)) continue;
        if (n.left < 0) { for (int i = 0; i < n.count; ++i) test(B.prims[n.first + i]); continue; }
        stack[sp++] = n.right; stack[sp++] = n.left;
        if (sp > 126) throw std::runtime_error("BVH too deep");
    }
    return h;
}

// ------------------------------------------------------------------ tracer
struct Adjoint {
    const int32_t* group = nullptr;   // per facet
    uint32_t G = 0;
    const double* s = nullptr;        // per group
    std::vector<int> target_of;       // facet -> target index or -1
    uint32_t T = 0;
    std::vector<double> j0, jac;      // per thread
};

struct Tracer {
    const Model& M;
    Tally& T;
    Rng rng;
    Adjoint* adj = nullptr;
    std::vector<uint32_t> ncoll;  // adjoint: sticky collisions per group
    std::vector<std::pair<int, double>> seg;   // adjoint: (target, score) since the last survived collision
    // particle state (Particle.h:95-122)
    V3 pos{}, dir{};
    double time = 0, velocity = 0, ori = 1, gen_time = 0, dist = 0, decay_t = 1e100, scatter_path = 1e100;
    int structure = 0, last_hit = -1, teleported_from = -1;
    uint64_t nb_bounces = 0;
    uint32_t nb_events = 0;
    double hu = 0, hv = 0;   // (u,v) of the facet being recorded
    // trace
    int32_t* tr_f = nullptr; int32_t* tr_k = nullptr; uint32_t tr_cap = 0, tr_n = 0;

    Tracer(const Model& m, Tally& t) : M(m), T(t) {}

    void trace(int f, int kind) {
        if (tr_n < tr_cap) { tr_f[tr_n] = f; tr_k[tr_n] = kind; ++tr_n; }
    }
    double sticking(const Facet& F, int fid) const {
        if (adj && adj->group[fid] >= 0) return adj->s[adj->group[fid]];
        return param_at(M, F.d.sticking_param, time, F.d.sticking);
    }
    double temperature(const Facet& F) const { return param_at(M, F.d.temperature_param, time, F.d.temperature); }

    void facet_counter(int fid, int m, uint64_t hit, uint64_t des, uint64_t abs, double s1v, double svo, V3 imp, V3 imp2, V3 impm) {
        double he = (double)hit * ori;
        for (int slot : {0, m}) {
            if (slot < 0 || (slot == m && m == 0)) continue;
            mfg_facet_counters& c = T.fc[(size_t)slot * T.nF + fid];
            c.nb_mc_hit += hit; c.nb_hit_equiv += he; c.nb_desorbed += des;
            c.nb_abs_equiv += (double)abs * ori;
            c.sum_1_per_ort_velocity += ori * s1v;
            c.sum_v_ort += ori * svo;
            c.sum_1_per_velocity += (he + (double)des) / velocity;
            if (M.s.enable_force) {
                double* fo = &T.force[((size_t)slot * T.nF + fid) * 9];
                fo[0] += ori * imp.x; fo[1] += ori * imp.y; fo[2] += ori * imp.z;
                fo[3] += ori * imp2.x; fo[4] += ori * imp2.y; fo[5] += ori * imp2.z;
                fo[6] += ori * impm.x; fo[7] += ori * impm.y; fo[8] += ori * impm.z;
            }
        }
        if (adj && hit && adj->target_of[fid] >= 0) adjoint_score(fid, he, false);
    }
    // Path-score contribution of a hit-equivalent x recorded on target fid:
    // d x / d s_g = x * (-n_g / (1 - s_g)) (+ x / s_t for the absorbed part).
    // The score is also kept in `seg` (scores since the last collision the
    // particle survived) for the cutoff term of cutoff_boundary.
    void adjoint_score(int fid, double x, bool absorbed_part) {
        int t = adj->target_of[fid];
        adj->j0[t] += x;
        for (uint32_t g = 0; g < adj->G; ++g) {
            double d = -x * (double)ncoll[g] / (1.0 - adj->s[g]);
            if (absorbed_part && adj->group[fid] == (int)g) d += x / adj->s[g];
            adj->jac[(size_t)t * adj->G + g] += d;
        }
        seg.push_back({t, x});
    }
    // The low-flux cutoff's term of d E[J] / d s_g (not a derivative of any
    // single path). A particle ends at the first collision K whose weight
    // falls to the cutoff c: W_K = W_{K-1} (1 - s_K) <= c < W_{K-1}. A rise
    // ds of s_g lowers W_{K-1} by W_{K-1} n_g ds / (1 - s_g) (n_g: the
    // particle's collisions with group g up to K-1), so with probability
    // f(0) n_g ds / (1 - s_g) the particle ends at K-1 instead, losing the
    // scores recorded after it survived K-1 (`seg`); f(0) is the density of
    // ln(W_{K-1} / c) at 0. In the renewal limit (ln(1/c) spans many
    // collisions) ln(W_{K-1} / c) is uniform on (0, -ln(1 - s_K)] given the
    // collision K that ends the particle, so f(0) = 1 / -ln(1 - s_K):
    //   dJ_t / ds_g += -sum_seg x_t n_g / ((1 - s_g) (-ln(1 - s_K))).
    void cutoff_boundary(int fid, double sK) {
        const double dK = -std::log1p(-sK);
        if (!(dK > 0.0)) return;
        const int gK = adj->group[fid];
        for (const auto& e : seg)
            for (uint32_t g = 0; g < adj->G; ++g) {
                const double n = (double)ncoll[g] - (gK == (int)g ? 1.0 : 0.0);
                if (n > 0.0) adj->jac[(size_t)e.first * adj->G + g] -= e.second * n / ((1.0 - adj->s[g]) * dK);
            }
    }
    void force_terms(V3 vel, V3& v2, V3& mom) {
        v2 = {vel.x * vel.x, vel.y * vel.y, vel.z * vel.z};
        V3 tr = {M.s.torque_ref[0], M.s.torque_ref[1], M.s.torque_ref[2]};
        mom = cross(pos - tr, vel);
    }
    void counter_with_force(int fid, int m, uint64_t hit, uint64_t des, uint64_t abs, double s1v, double svo, double sign) {
        V3 z{0, 0, 0}, v2 = z, mom = z, vel = z;
        if (M.s.enable_force) { vel = (sign * velocity) * dir; force_terms(vel, v2, mom); }
        facet_counter(fid, m, hit, des, abs, s1v, svo, vel, v2, mom);
    }
    void texture(const Facet& F, int m, bool count, double vf, double of) {
        int tu = std::min(F.tex_w - 1, (int)(hu * F.d.tex_width_precise));
        int tv = std::min(F.tex_h - 1, (int)(hv * F.d.tex_height_precise));
        if (tu < 0 || tv < 0) return;
        size_t add = (size_t)tu + (size_t)tv * F.tex_w;
        double ov = M.c_factor * velocity * std::fabs(dot(dir, F.N));
        for (int slot : {0, m}) {
            if (slot < 0 || (slot == m && m == 0)) continue;
            mfg_texture_cell& c = T.tex[(size_t)slot * T.ntex + F.tex_off + add];
            if (count) c.count_equiv += ori;
            c.sum_1_per_ort_velocity += ori * vf / ov;
            c.sum_v_ort_per_area += ori * of * ov * F.tex_inc[add];
        }
    }
    void direction(const Facet& F, int m) {
        if (F.dir_off < 0) return;
        int tu = std::min(F.tex_w - 1, (int)(hu * F.d.tex_width_precise));
        int tv = std::min(F.tex_h - 1, (int)(hv * F.d.tex_height_precise));
        if (tu < 0 || tv < 0) return;
        size_t add = (size_t)tu + (size_t)tv * F.tex_w;
        for (int slot : {0, m}) {
            if (slot < 0 || (slot == m && m == 0)) continue;
            mfg_direction_cell& c = T.dir[(size_t)slot * T.ndir + F.dir_off + add];
            c.dir[0] += ori * dir.x * velocity; c.dir[1] += ori * dir.y * velocity; c.dir[2] += ori * dir.z * velocity;
            c.count++;
        }
    }
    void profile(const Facet& F, int m, bool count, double vf, double of) {
        if (F.prof_off < 0) return;
        int pt = F.d.profile_type;
        auto slots = [&](auto&& fn) { for (int slot : {0, m}) { if (slot < 0 || (slot == m && m == 0)) continue; fn(T.prof[(size_t)slot * T.nprof + F.prof_off]); } };
        (void)slots;
        if (count && pt == MFG_PROFILE_ANGULAR) {
            double th = std::acos(std::min(1.0, std::fabs(dot(F.N, dir))));
            long pos_ = (long)(th / (kPi / 2) * (double)kProfileSize);
            pos_ = std::max(0L, std::min((long)kProfileSize - 1, pos_));
            for (int slot : {0, m}) { if (slot < 0 || (slot == m && m == 0)) continue; T.prof[(size_t)slot * T.nprof + F.prof_off + pos_].count_equiv += ori; }
        } else if (pt == MFG_PROFILE_U || pt == MFG_PROFILE_V) {
            double c = pt == MFG_PROFILE_U ? hu : hv;
            if (!(c >= 0.0)) return;
            size_t pos_ = (size_t)(c * (double)kProfileSize);
            if (pos_ >= (size_t)kProfileSize) return;
            double ov = velocity * std::fabs(dot(F.N, dir));
            for (int slot : {0, m}) {
                if (slot < 0 || (slot == m && m == 0)) continue;
                mfg_profile_slice& p = T.prof[(size_t)slot * T.nprof + F.prof_off + pos_];
                if (count) p.count_equiv += ori;
                p.sum_1_per_ort_velocity += ori * vf / ov;
                p.sum_v_ort += ori * of * M.c_factor * ov;
            }
        } else if (count && (pt == MFG_PROFILE_VELOCITY || pt == MFG_PROFILE_ORT_VELOCITY || pt == MFG_PROFILE_TAN_VELOCITY)) {
            double d;
            if (pt == MFG_PROFILE_VELOCITY) d = 1.0;
            else if (pt == MFG_PROFILE_ORT_VELOCITY) d = std::fabs(dot(F.N, dir));
            else { double a = std::fabs(dot(F.N, dir)); d = std::sqrt(std::max(0.0, 1 - a * a)); }
            double x = d * velocity / F.max_speed * (double)kProfileSize;
            if (!(x >= 0.0)) return;
            size_t pos_ = (size_t)x;
            if (pos_ >= (size_t)kProfileSize) return;
            for (int slot : {0, m}) { if (slot < 0 || (slot == m && m == 0)) continue; T.prof[(size_t)slot * T.nprof + F.prof_off + pos_].count_equiv += ori; }
        }
    }
    void polar(const Facet& F, double& th, double& ph) const {
        double lu = dot(dir, F.nU), lv = dot(dir, F.nV), ln = dot(dir, F.N);
        th = std::acos(std::max(-1.0, std::min(1.0, ln)));
        ph = std::atan2(lv, lu);
    }
    void angle_map(const Facet& F) {
        if (!F.d.am_record || F.am_off < 0) return;
        double th, ph; polar(F, th, ph);
        if (th > kPi / 2.0) th = std::fabs(kPi - th);
        const auto& d = F.d;
        long ti;
        if (th < d.am_theta_limit) {
            if (d.am_theta_lower_res == 0) return;
            ti = (long)(th / d.am_theta_limit * (double)d.am_theta_lower_res);
        } else {
            if (d.am_theta_higher_res == 0) return;
            ti = (long)d.am_theta_lower_res + (long)((th - d.am_theta_limit) / (kPi / 2.0 - d.am_theta_limit) * (double)d.am_theta_higher_res);
        }
        long nt = (long)(d.am_theta_lower_res + d.am_theta_higher_res);
        ti = std::max(0L, std::min(nt - 1, ti));
        long pi_ = (long)((ph + 3.1415926) / (2.0 * kPi) * (double)d.am_phi_width);
        pi_ = std::max(0L, std::min((long)d.am_phi_width - 1, pi_));
        T.am[F.am_off + (size_t)ti * d.am_phi_width + pi_]++;
    }
    void log_hit(int fid) {
        if (!M.s.log_enabled || M.s.log_facet_id != fid || T.log.size() >= M.s.log_limit) return;
        mfg_log_entry e;
        e.u = hu; e.v = hv;
        polar(M.F[fid], e.hit_theta, e.hit_phi);
        e.ori_ratio = ori; e.particle_decay_moment = decay_t; e.time = time; e.velocity = velocity;
        T.log.push_back(e);
    }
    void histograms(const Facet& F, int m) {  // RecordHistograms: called on absorption
        for (int slot : {0, m}) {
            if (slot < 0) return;
            if (T.ngh) {
                double* g = &T.gh[(size_t)slot * T.ngh];
                int off = 0;
                if (M.gh_bounce) { g[off + std::min<uint64_t>(nb_bounces / M.s.gh_bounce_binsize, M.gh_bounce - 1)] += ori; off += M.gh_bounce; }
                if (M.gh_dist) { g[off + std::min<size_t>((size_t)(dist / M.s.gh_distance_binsize), M.gh_dist - 1)] += ori; off += M.gh_dist; }
                if (M.gh_time) { g[off + std::min<size_t>((size_t)((time - gen_time) / M.s.gh_time_binsize), M.gh_time - 1)] += ori; }
            }
            if (F.fh_off >= 0) {
                double* g = &T.fh[(size_t)slot * T.nfh + F.fh_off];
                int off = 0;
                if (F.fh_bounce_bins) { g[off + std::min<uint64_t>(nb_bounces / F.d.fh_bounce_binsize, F.fh_bounce_bins - 1)] += ori; off += F.fh_bounce_bins; }
                if (F.fh_dist_bins) { g[off + std::min<size_t>((size_t)(dist / F.d.fh_distance_binsize), F.fh_dist_bins - 1)] += ori; off += F.fh_dist_bins; }
                if (F.fh_time_bins) { g[off + std::min<size_t>((size_t)((time - gen_time) / F.d.fh_time_binsize), F.fh_time_bins - 1)] += ori; }
            }
            if (slot == m) return;  // m == 0 case already handled
        }
    }
    V3 polar_to_cart(const Facet& F, double cos_t, double sin_t, double phi, bool reverse) const {
        double c = std::cos(phi), s = std::sin(phi);
        double n = reverse ? -cos_t : cos_t;
        return (sin_t * c) * F.nU + (sin_t * s) * F.nV + n * F.N;
    }
    void moving(V3& d, double& vel) const {  // TreatMovingFacet, Physics.cpp:22-39
        V3 add{0, 0, 0};
        if (M.s.motion_type == 1) add = {M.s.motion_vector2[0], M.s.motion_vector2[1], M.s.motion_vector2[2]};
        else if (M.s.motion_type == 2) {
            V3 base = {M.s.motion_vector1[0], M.s.motion_vector1[1], M.s.motion_vector1[2]};
            V3 ax = {M.s.motion_vector2[0], M.s.motion_vector2[1], M.s.motion_vector2[2]};
            add = cross(ax, 0.01 * (pos - base));
        }
        V3 nv = vel * d + add;
        vel = norm(nv); d = unit(nv);
    }
    double maxwell_speed(double sqrtT) { return interp_x(rng.rnd(), M.maxwell.data(), M.maxwell.size()) * sqrtT; }

    bool start_from_source() {  // StartFromSource
        double r = rng.rnd() * M.total_des;
        size_t k = std::upper_bound(M.src_cum.begin(), M.src_cum.end(), r) - M.src_cum.begin();
        if (M.src_cum.empty()) return false;
        if (k >= M.src_cum.size()) k = M.src_cum.size() - 1;
        int src = M.src_facet[k];
        const Facet& F = M.F[src];
        bool reverse = false;
        if (F.d.is_2sided) reverse = rng.rnd() > 0.5;
        long map_w = -1, map_h = -1;
        if (F.d.use_outgassing_map) {
            double base = k ? M.src_cum[k - 1] : 0.0;
            double rem = (r - base) / M.latest * (kBoltz * F.d.temperature);
            size_t cell = std::upper_bound(F.og_cdf.begin(), F.og_cdf.end(), rem) - F.og_cdf.begin();
            if (cell >= F.og_cdf.size()) cell = F.og_cdf.size() - 1;
            map_h = (long)(cell / F.d.og_map_width);
            map_w = (long)cell - map_h * (long)F.d.og_map_width;
        }
        last_hit = src;
        double rt = rng.rnd();
        if (F.d.outgassing_param >= 0) {
            const auto& ID = M.IDs[F.id_index];
            time = interp_x(rt * ID.back().y, ID.data(), ID.size());
        } else time = rt * M.latest;
        gen_time = time;
        if (M.s.use_maxwell) velocity = maxwell_speed(F.sqrt_temp);
        else velocity = 145.469 * std::sqrt(F.d.temperature / M.s.gas_mass);
        ori = 1.0;
        decay_t = M.s.enable_decay ? time + (1.44269 * M.s.half_life) * -std::log(rng.rnd()) : 1e100;
        scatter_path = M.s.scatter_enabled ? M.s.scatter_mean_free_path_cm * -std::log(rng.rnd()) : 1e100;
        nb_bounces = 0; dist = 0;
        bool found = false;
        double wp = F.d.og_map_width ? norm(F.U) * F.d.og_ratio_u : 0, hp = F.d.og_map_height ? norm(F.V) * F.d.og_ratio_v : 0;
        for (int tr = 0; tr < 1000 && !found; ++tr) {
            double u, v;
            if (map_w >= 0) {
                double a = rng.rnd(), b = rng.rnd();
                u = (map_w < (long)F.d.og_map_width - 1) ? ((double)map_w + a) / wp : ((double)map_w + a * (wp - ((double)F.d.og_map_width - 1.0))) / wp;
                v = (map_h < (long)F.d.og_map_height - 1) ? ((double)map_h + b) / hp : ((double)map_h + b * (hp - ((double)F.d.og_map_height - 1.0))) / hp;
            } else { u = rng.rnd(); v = rng.rnd(); }
            if (in_poly(F.poly, u, v)) { pos = F.O + u * F.U + v * F.V; hu = u; hv = v; found = true; }
        }
        if (!found) {
            if (map_w >= 0) { hu = ((double)map_w + 0.5) / wp; hv = ((double)map_h + 0.5) / hp; pos = F.O + hu * F.U + hv * F.V; }
            else { hu = 0.5; hv = 0.5; pos = F.center; }
        }
        switch (F.d.desorb_type) {
            case MFG_DES_UNIFORM: { double c = rng.rnd(); double ph = rng.rnd() * 2.0 * kPi; dir = polar_to_cart(F, c, std::sqrt(1 - c * c), ph, reverse); break; }
            case MFG_DES_COSINE_N: { double c = std::pow(rng.rnd(), 1.0 / (F.d.desorb_exponent + 1.0)); double ph = rng.rnd() * 2.0 * kPi; dir = polar_to_cart(F, c, std::sqrt(1 - c * c), ph, reverse); break; }
            case MFG_DES_ANGLEMAP: {   // Particle.cpp:599-621
                double th, ov;
                int tlo;
                am_generate_theta(F.am_src, rng.rnd(), th, tlo, ov);
                double ph = am_generate_phi(F.am_src, tlo, ov, rng.rnd());
                double t2 = kPi - th;   // the map holds incident angles
                dir = polar_to_cart(F, std::cos(t2), std::sin(t2), ph, false);
                break;
            }
            default: { double a = rng.rnd(); double ph = rng.rnd() * 2.0 * kPi; dir = polar_to_cart(F, std::sqrt(a), std::sqrt(1 - a), ph, reverse); break; }
        }
        structure = F.d.super_idx;
        teleported_from = -1;
        T.g.nb_desorbed++;
        if (F.d.is_moving && M.s.motion_type) moving(dir, velocity);
        double ov = velocity * std::fabs(dot(dir, F.N));
        int m = moment_index(M, time);
        counter_with_force(src, m, 0, 1, 0, 2.0 / ov, M.c_factor * ov, 1.0);
        profile(F, m, false, 2.0, 1.0);
        log_hit(src);
        if (F.d.count_des && F.tex_off >= 0) texture(F, m, true, 2.0, 1.0);
        trace(src, 0);
        return true;
    }
    void transparent_pass(int fid, double tp) {  // RegisterTransparentPass
        const Facet& F = M.F[fid];
        double df = std::fabs(dot(dir, F.N));
        int m = moment_index(M, time + tp / 100.0 / velocity);
        facet_counter(fid, m, 1, 0, 0, 2.0 / (velocity * df), 2.0 * M.c_factor * velocity * df, {0, 0, 0}, {0, 0, 0}, {0, 0, 0});
        if (F.d.count_trans && F.tex_off >= 0) texture(F, m, true, 2.0, 2.0);
        if (F.d.count_dir) direction(F, m);
        log_hit(fid);
        profile(F, m, true, 2.0, 2.0);
        angle_map(F);
        trace(fid, 3);
    }
    void record_absorb(int fid) {  // RecordAbsorb
        const Facet& F = M.F[fid];
        T.g.nb_mc_hit++; T.g.nb_hit_equiv += ori; T.g.nb_abs_equiv += ori;
        int m = moment_index(M, time);
        histograms(F, m);
        double ov = velocity * std::fabs(dot(dir, F.N));
        bool adj_abs = adj && adj->target_of[fid] >= 0;
        Adjoint* saved = adj;
        if (adj_abs) adj = nullptr;  // scored below with the absorbed-part term
        counter_with_force(fid, m, 1, 0, 1, 2.0 / ov, M.c_factor * ov, 1.0);
        adj = saved;
        if (adj_abs) adjoint_score(fid, ori, true);
        log_hit(fid);
        profile(F, m, true, 2.0, 1.0);
        angle_map(F);
        if (F.d.count_abs && F.tex_off >= 0) texture(F, m, true, 2.0, 1.0);
        if (F.d.count_dir) direction(F, m);
    }
    void update_velocity(const Facet& F) {  // UpdateVelocity
        double T_ = temperature(F);
        double sq = F.d.temperature_param == -1 ? F.sqrt_temp : std::sqrt(T_);
        if (F.d.accomodation > 0.9999) {
            velocity = M.s.use_maxwell ? maxwell_speed(sq) : 145.469 * sq / M.s.gas_mass;
        } else {
            double o2 = velocity * velocity, n2;
            if (M.s.use_maxwell) { double vv = maxwell_speed(sq); n2 = vv * vv; }
            else n2 = 29369.939 * (T_ / M.s.gas_mass);
            velocity = std::sqrt(o2 + (n2 - o2) * F.d.accomodation);
        }
    }
    void perform_bounce(int fid) {  // PerformBounce
        const Facet& F = M.F[fid];
        T.g.nb_mc_hit++; T.g.nb_hit_equiv += ori;
        if (F.d.super_dest) {
            int m = moment_index(M, time);
            facet_counter(fid, m, 1, 0, 0, 0, 0, {0, 0, 0}, {0, 0, 0}, {0, 0, 0});
            structure = F.d.super_dest - 1;
            if (F.d.is_moving) moving(dir, velocity);
            log_hit(fid);
            profile(F, m, true, 2.0, 2.0);
            angle_map(F);
            if (F.d.count_trans && F.tex_off >= 0) texture(F, m, true, 2.0, 2.0);
            if (F.d.count_dir) direction(F, m);
            trace(fid, 5);
            return;
        }
        bool revert = F.d.is_2sided && dot(dir, F.N) > 0.0;
        double ov = velocity * std::fabs(dot(dir, F.N));
        int m = moment_index(M, time);
        counter_with_force(fid, m, 1, 0, 0, 1.0 / ov, M.c_factor * ov, 1.0);
        nb_bounces++;
        if (F.d.count_refl && F.tex_off >= 0) texture(F, m, true, 1.0, 1.0);
        if (F.d.count_dir) direction(F, m);
        log_hit(fid);
        profile(F, m, true, 1.0, 1.0);
        angle_map(F);
        update_velocity(F);
        if (F.d.enable_sojourn) {
            double A = std::exp(-F.d.sojourn_e / (8.31 * temperature(F)));
            time += -std::log(rng.rnd()) / (A * F.d.sojourn_freq);
            m = moment_index(M, time);
        }
        const auto& d = F.d;
        if (d.diffuse_part > 0.999999) {
            double a = rng.rnd(); double ph = rng.rnd() * 2.0 * kPi;
            dir = polar_to_cart(F, std::sqrt(a), std::sqrt(1 - a), ph, revert);
        } else {
            double rt = rng.rnd();
            if (rt < d.diffuse_part) {
                double a = rng.rnd(); double ph = rng.rnd() * 2.0 * kPi;
                dir = polar_to_cart(F, std::sqrt(a), std::sqrt(1 - a), ph, revert);
            } else if (rt < d.diffuse_part + d.specular_part) {
                dir = dir - (2.0 * dot(dir, F.N)) * F.N;
            } else {
                double c = std::pow(rng.rnd(), 1.0 / (d.cosine_exponent + 1.0)); double ph = rng.rnd() * 2.0 * kPi;
                dir = polar_to_cart(F, c, std::sqrt(1 - c * c), ph, revert);
            }
        }
        if (F.d.is_moving && M.s.motion_type) moving(dir, velocity);
        ov = velocity * std::fabs(dot(dir, F.N));
        counter_with_force(fid, m, 0, 0, 0, 1.0 / ov, M.c_factor * ov, -1.0);
        if (F.d.count_refl && F.tex_off >= 0) texture(F, m, false, 1.0, 1.0);
        profile(F, m, false, 1.0, 1.0);
        last_hit = fid;
        trace(fid, 1);
    }
    // PerformTeleport (Particle.cpp:76-183). teleported_from becomes the
    // destination, as in the reference's live code (:122); "back to origin"
    // (teleport_dest -1) therefore returns to the last teleport destination.
    // The reference's retry loop for a point outside the destination polygon
    // (:146-158) increments its try counter after the loop, so it only ends
    // on success; here it ends after 1000 tries. A destination in every
    // structure (super_idx -1) keeps the current structure; the reference
    // would set ray.structure = -1 (:121) and fail at its next Intersect.
    void perform_teleport(int fid) {
        const Facet& F = M.F[fid];
        int dest = F.d.teleport_dest == -1 ? teleported_from : F.d.teleport_dest - 1;
        trace(fid, 4);
        if (dest < 0 || dest >= (int)M.F.size()) { last_hit = fid; return; }
        const Facet& D = M.F[dest];
        if (D.d.super_idx >= 0) structure = D.d.super_idx;
        teleported_from = dest;
        int m = moment_index(M, time);
        if (F.d.count_trans && F.tex_off >= 0) texture(F, m, true, 2.0, 2.0);
        if (F.d.count_dir) direction(F, m);
        profile(F, m, true, 2.0, 2.0);
        log_hit(fid);
        angle_map(F);
        double lu = dot(dir, F.nU), lv = dot(dir, F.nV), ln = dot(dir, F.N);
        dir = lu * D.nU + lv * D.nV + ln * D.N;
        double u = hu, v = hv;
        pos = D.O + u * D.U + v * D.V;
        if (!in_poly(D.poly, u, v)) {
            for (int tr = 0; tr < 1000; ++tr) {
                u = rng.rnd(); v = rng.rnd();
                if (in_poly(D.poly, u, v)) { pos = D.O + u * D.U + v * D.V; break; }
            }
        }
        last_hit = dest;
        double ov = velocity * std::fabs(dot(dir, F.N));
        facet_counter(fid, m, 1, 0, 0, 2.0 / ov, 2.0 * M.c_factor * ov, {0, 0, 0}, {0, 0, 0}, {0, 0, 0});
    }
    bool perform_scatter() {  // PerformScatter
        T.g.nb_mc_hit++; T.g.nb_hit_equiv += ori;
        nb_bounces++;
        double rho = M.s.scatter_rho, mr = M.s.scatter_mass_ratio;
        double b = rho * std::sqrt(rng.rnd());
        double a = std::sqrt(rho * rho - b * b);
        double ctg = 1.0 / (2.0 * a * b) * (rho * rho * mr + b * b - a * a);
        double th = std::atan(1.0 / ctg);
        double az = 2.0 * kPi * rng.rnd();
        V3 rv = dir.x != 0.0 ? V3{0, 0, 1} : (dir.y != 0.0 ? V3{1, 0, 1} : V3{1, 0, 0});
        V3 un = unit(cross(dir, rv)), vn = unit(cross(dir, un));
        dir = (std::sin(th) * std::cos(az)) * un + (std::sin(th) * std::sin(az)) * vn + std::cos(th) * dir;
        velocity = std::sqrt(velocity * velocity / (1.0 + std::pow(mr * rho / b * std::sin(th), 2.0)));
        scatter_path = dist + M.s.scatter_mean_free_path_cm * -std::log(rng.rnd());
        trace(-1, 8);
        return !(M.s.scatter_enable_cutoff && velocity < M.s.scatter_cutoff_speed);
    }

    // One particle from desorption to its end (SimulationMCStep loop).
    void run_particle(uint64_t seed, uint64_t pid, uint64_t max_events, std::vector<TPass>& passes) {
        rng.init(seed, pid);
        nb_events = 0;
        if (adj) std::fill(ncoll.begin(), ncoll.end(), 0u);
        seg.clear();
        if (!start_from_source()) return;
        for (;;) {
            if (max_events && nb_events >= max_events) return;
            ++nb_events;
            passes.clear();
            Hit h = intersect(M, structure, pos, dir, last_hit, time, &rng, nb_events, &passes, false, false);
            if (h.facet < 0) {  // leak
                T.g.nb_leak_total++;
                if (T.leaks.size() < 6 * kLeakCap) { for (double x : {pos.x, pos.y, pos.z, dir.x, dir.y, dir.z}) T.leaks.push_back(x); }
                T.n_leaks++;
                trace(-1, 6);
                return;
            }
            double travel = h.t;
            int ev = 0;  // 0 facet, 1 overtime, 2 decay, 3 scatter
            if (!M.s.calc_constant_flow) { double dl = velocity * 100.0 * (M.latest - time); if (dl < travel) { travel = dl; ev = 1; } }
            if (M.s.enable_decay) { double dd = velocity * 100.0 * (decay_t - time); if (dd < travel) { travel = dd; ev = 2; } }
            if (M.s.scatter_enabled) { double ds = scatter_path - dist; if (ds < travel) { travel = ds; ev = 3; } }
            for (const TPass& p : passes)
                if (p.t < travel) { hu = p.u; hv = p.v; transparent_pass(p.facet, p.t); }
            pos = pos + travel * dir;
            time += travel / 100.0 / velocity;
            if (ev == 0) {
                const Facet& F = M.F[h.facet];
                hu = h.u; hv = h.v;
                T.g.dist_traveled_total += travel; T.g.dist_traveled_full_hits_only += travel; dist += travel;
                if (F.d.teleport_dest != 0) { perform_teleport(h.facet); continue; }
                double s = sticking(F, h.facet);
                if (!M.s.low_flux_mode) {
                    if (s == 1.0 || (s > 0.0 && rng.rnd() < s)) { record_absorb(h.facet); trace(h.facet, 2); return; }
                    perform_bounce(h.facet);
                } else {
                    if (s > 0.0) {
                        double w0 = ori;
                        ori *= s;
                        record_absorb(h.facet);
                        ori = w0 * (1.0 - s);
                        if (adj && adj->group[h.facet] >= 0) ncoll[adj->group[h.facet]]++;
                    } else ori *= (1.0 - s);
                    if (ori > M.s.low_flux_cutoff) {
                        if (s > 0.0) seg.clear();
                        perform_bounce(h.facet);
                    } else {
                        if (adj && s > 0.0) cutoff_boundary(h.facet, s);
                        trace(h.facet, 2);
                        return;
                    }
                }
            } else if (ev == 1 || ev == 2) {
                T.g.dist_traveled_total += travel;
                T.g.nb_decayed++;
                trace(-1, 7);
                return;
            } else {
                T.g.dist_traveled_total += travel; T.g.dist_traveled_full_hits_only += travel; dist += travel;
                bool ok = perform_scatter();
                last_hit = -1;
                if (!ok) { T.g.nb_decayed++; return; }
            }
        }
    }
};

void add_into(const Model& M, const Tally& T, const mfg_tally_view* o) {
    if (o->global) {
        auto& g = *o->global;
        g.nb_mc_hit += T.g.nb_mc_hit; g.nb_hit_equiv += T.g.nb_hit_equiv; g.nb_desorbed += T.g.nb_desorbed;
        g.nb_abs_equiv += T.g.nb_abs_equiv; g.nb_leak_total += T.g.nb_leak_total;
        g.dist_traveled_total += T.g.dist_traveled_total; g.dist_traveled_full_hits_only += T.g.dist_traveled_full_hits_only;
        g.nb_decayed += T.g.nb_decayed;
    }
    if (o->facets) for (size_t i = 0; i < T.fc.size(); ++i) {
        auto& a = o->facets[i]; const auto& b = T.fc[i];
        a.nb_mc_hit += b.nb_mc_hit; a.nb_hit_equiv += b.nb_hit_equiv; a.nb_desorbed += b.nb_desorbed;
        a.nb_abs_equiv += b.nb_abs_equiv; a.sum_1_per_ort_velocity += b.sum_1_per_ort_velocity;
        a.sum_v_ort += b.sum_v_ort; a.sum_1_per_velocity += b.sum_1_per_velocity;
    }
    if (o->forces && !T.force.empty()) for (size_t i = 0; i < T.force.size(); ++i) o->forces[i] += T.force[i];
    if (o->textures) for (size_t i = 0; i < T.tex.size(); ++i) { o->textures[i].count_equiv += T.tex[i].count_equiv; o->textures[i].sum_v_ort_per_area += T.tex[i].sum_v_ort_per_area; o->textures[i].sum_1_per_ort_velocity += T.tex[i].sum_1_per_ort_velocity; }
    if (o->directions) for (size_t i = 0; i < T.dir.size(); ++i) { for (int k = 0; k < 3; ++k) o->directions[i].dir[k] += T.dir[i].dir[k]; o->directions[i].count += T.dir[i].count; }
    if (o->profiles) for (size_t i = 0; i < T.prof.size(); ++i) { o->profiles[i].count_equiv += T.prof[i].count_equiv; o->profiles[i].sum_v_ort += T.prof[i].sum_v_ort; o->profiles[i].sum_1_per_ort_velocity += T.prof[i].sum_1_per_ort_velocity; }
    if (o->angle_maps) for (size_t i = 0; i < T.am.size(); ++i) o->angle_maps[i] += T.am[i];
    if (o->global_hist) for (size_t i = 0; i < T.gh.size(); ++i) o->global_hist[i] += T.gh[i];
    if (o->facet_hist) for (size_t i = 0; i < T.fh.size(); ++i) o->facet_hist[i] += T.fh[i];
    (void)M;
}

}  // namespace

struct mfcpu_model { Model M; };

extern "C" {

mfcpu_model* mfcpu_load(const mfg_model_view* view, char* err, size_t cap) {
    try {
        auto m = std::make_unique<mfcpu_model>();
        load_model(m->M, view);
        return m.release();
    } catch (const std::exception& e) {
        if (err && cap) { std::snprintf(err, cap, "%s", e.what()); }
        return nullptr;
    }
}

void mfcpu_free(mfcpu_model* m) { delete m; }

int mfcpu_layout(const mfcpu_model* mm, mfg_tally_layout* L) {
    if (!mm || !L) return MFG_ERR_INVALID;
    const Model& M = mm->M;
    std::memset(L, 0, sizeof(*L));
    L->struct_size = sizeof(*L);
    L->n_facets = (uint32_t)M.F.size(); L->n_slots = M.n_slots; L->n_structures = (uint32_t)M.n_struct;
    L->n_tex_cells = M.n_tex; L->n_dir_cells = M.n_dir; L->n_profile_slices = M.n_prof; L->n_angle_cells = M.n_am;
    L->n_global_hist_bins = (uint64_t)M.gh_bounce + M.gh_dist + M.gh_time; L->n_facet_hist_bins = M.n_fh;
    L->log_capacity = M.s.log_enabled ? M.s.log_limit : 0; L->leak_capacity = kLeakCap;
    return 0;
}

int mfcpu_facet_offsets(const mfcpu_model* mm, int kind, int64_t* out, size_t n) {
    if (!mm || !out || n < mm->M.F.size()) return MFG_ERR_INVALID;
    for (size_t f = 0; f < mm->M.F.size(); ++f) {
        const Facet& F = mm->M.F[f];
        out[f] = kind == 0 ? F.tex_off : kind == 1 ? F.dir_off : kind == 2 ? F.prof_off : kind == 3 ? F.am_off : F.fh_off;
    }
    return 0;
}

int mfcpu_run(mfcpu_model* mm, uint64_t seed, uint64_t first, uint64_t n_des, uint64_t max_events, int threads,
              const mfg_tally_view* out) {
    if (!mm || !out) return MFG_ERR_INVALID;
    const Model& M = mm->M;
    if (M.src_cum.empty()) return MFG_ERR_DESORPTION;
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
    std::vector<Tally> tallies(nt);
    for (auto& t : tallies) t.init(M);
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Tracer tr(M, tallies[tid]);
        std::vector<TPass> passes;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < (int64_t)n_des; ++i) tr.run_particle(seed, first + (uint64_t)i, max_events, passes);
    }
    for (auto& t : tallies) add_into(M, t, out);
    if (out->log && out->log_count) {
        uint64_t n = 0;
        for (auto& t : tallies) for (auto& e : t.log) if (n < M.s.log_limit) out->log[n++] = e;
        *out->log_count = n;
    }
    if (out->leak_pos_dir && out->leak_count) {
        uint64_t n = 0, total = 0;
        for (auto& t : tallies) { total += t.n_leaks; for (size_t i = 0; i + 6 <= t.leaks.size() && n < kLeakCap; i += 6, ++n) std::memcpy(out->leak_pos_dir + 6 * n, &t.leaks[i], 6 * sizeof(double)); }
        *out->leak_count = total;
    }
    return 0;
}

int mfcpu_trace(mfcpu_model* mm, uint64_t seed, uint64_t first, uint32_t n, uint32_t k, int32_t* facet, int32_t* kind) {
    if (!mm) return MFG_ERR_INVALID;
    Tally t; t.init(mm->M);
    Tracer tr(mm->M, t);
    std::vector<TPass> passes;
    for (uint32_t i = 0; i < n; ++i) {
        for (uint32_t j = 0; j < k; ++j) { facet[(size_t)i * k + j] = -2; kind[(size_t)i * k + j] = -1; }
        tr.tr_f = facet + (size_t)i * k; tr.tr_k = kind + (size_t)i * k; tr.tr_cap = k; tr.tr_n = 0;
        tr.run_particle(seed, first + i, k + 1, passes);
    }
    return 0;
}

int mfcpu_intersect_batch(mfcpu_model* mm, uint64_t n, const double* o, const double* d, const int32_t* last,
                          int32_t s, int brute, int32_t* of, double* ot, double* ou, double* ov) {
    if (!mm || s < 0 || s >= mm->M.n_struct) return MFG_ERR_INVALID;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        V3 P = {o[3 * i], o[3 * i + 1], o[3 * i + 2]}, D = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        Hit h = intersect(mm->M, s, P, D, last ? last[i] : -1, 0.0, nullptr, 0, nullptr, brute != 0, true);
        of[i] = h.facet; ot[i] = h.facet >= 0 ? h.t : 0.0; ou[i] = h.u; ov[i] = h.v;
    }
    return 0;
}

int mfcpu_bvh_stats(mfcpu_model* mm, uint64_t n, const double* o, const double* d, const int32_t* last, int32_t s,
                    double* mean_nodes, double* mean_prims) {
    if (!mm || n == 0) return MFG_ERR_INVALID;
    uint64_t nodes = 0, prims = 0;
    for (uint64_t i = 0; i < n; ++i) {
        V3 P = {o[3 * i], o[3 * i + 1], o[3 * i + 2]}, D = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        intersect(mm->M, s, P, D, last ? last[i] : -1, 0.0, nullptr, 0, nullptr, false, true, &nodes, &prims);
    }
    *mean_nodes = (double)nodes / n; *mean_prims = (double)prims / n;
    return 0;
}

void mfcpu_philox(const uint32_t c[4], const uint32_t k[2], uint32_t o[4]) { philox(c, k, o); }

int mfcpu_maxwell_cdf(double gas_mass, double* v, double* p, size_t n) {
    auto c = maxwell_cdf(gas_mass);
    for (size_t i = 0; i < n && i < c.size(); ++i) { v[i] = c[i].x; p[i] = c[i].y; }
    return 0;
}

int mfcpu_parse_user_moment(const char* content, double window, double* out, size_t cap, size_t* n) {
    auto v = parse_user_moment(content, window);
    for (size_t i = 0; i < v.size() && i < cap; ++i) out[i] = v[i];
    *n = v.size();
    return 0;
}

int mfcpu_integrated_desorption(const double* x, const double* y, size_t np, int lx, int ly, double latest,
                                double* ot, double* oc, size_t cap, size_t* n) {
    std::vector<XY> p;
    for (size_t i = 0; i < np; ++i) p.push_back({x[i], y[i]});
    auto id = integrated_desorption(p.data(), p.size(), lx != 0, ly != 0, latest);
    for (size_t i = 0; i < id.size() && i < cap; ++i) { ot[i] = id[i].x; oc[i] = id[i].y; }
    *n = id.size();
    return 0;
}

double mfcpu_facet_area(const mfcpu_model* m, uint32_t f) { return m->M.F.at(f).area; }

int mfcpu_facet_frame(const mfcpu_model* m, uint32_t f, double* o) {
    const Facet& F = m->M.F.at(f);
    V3 a[4] = {F.O, F.U, F.V, F.N};
    for (int i = 0; i < 4; ++i) { o[3 * i] = a[i].x; o[3 * i + 1] = a[i].y; o[3 * i + 2] = a[i].z; }
    return 0;
}

int mfcpu_texture_increments(const mfcpu_model* m, uint32_t f, double* out, size_t n) {
    const Facet& F = m->M.F.at(f);
    if (n < F.tex_inc.size()) return MFG_ERR_INVALID;
    std::copy(F.tex_inc.begin(), F.tex_inc.end(), out);
    return 0;
}

double mfcpu_total_desorbed_molecules(const mfcpu_model* m) { return m->M.total_des; }
double mfcpu_final_outgassing_rate(const mfcpu_model* m) { return m->M.final_rate; }
double mfcpu_latest_moment(const mfcpu_model* m) { return m->M.latest; }

// One PerformScatter (Particle.cpp:840-886, Physics.cpp:41-63) of n
// particles moving along +z at speed 1 with mass ratio and rho as given;
// returns the new direction's z component and the new speed of each. A
// test hook for the scattering laws; draws use particle i's stream.
int mfcpu_scatter_sample(double mass_ratio, double rho, uint64_t n, uint64_t seed, double* out_cos,
                         double* out_speed) {
    if (!out_cos || !out_speed || !(rho > 0)) return MFG_ERR_INVALID;
    Model M{};
    std::memset(&M.s, 0, sizeof(M.s));
    M.s.scatter_enabled = 1;
    M.s.scatter_mass_ratio = mass_ratio;
    M.s.scatter_rho = rho;
    M.s.scatter_mean_free_path_cm = 1.0;
    Tally T;
    Tracer tr(M, T);
    for (uint64_t i = 0; i < n; ++i) {
        tr.rng.init(seed, i);
        tr.dir = {0.0, 0.0, 1.0};
        tr.velocity = 1.0;
        tr.dist = 0.0;
        tr.perform_scatter();
        out_cos[i] = tr.dir.z;
        out_speed[i] = tr.velocity;
    }
    return 0;
}

int mfcpu_adjoint_run(mfcpu_model* mm, uint64_t seed, uint64_t first, uint64_t n_des, const int32_t* group,
                      uint32_t G, const double* gs, const int32_t* targets, uint32_t nT, int threads,
                      double* j0, double* jac, uint64_t max_events) {
    if (!mm || !mm->M.s.low_flux_mode) return MFG_ERR_INVALID;
    const Model& M = mm->M;
    // the path score d/ds = -n/(1-s) (+ 1/s on the absorbed part) needs every
    // group sticking strictly inside (0,1)
    if (!group || !gs || !targets || G == 0 || nT == 0) return MFG_ERR_INVALID;
    for (uint32_t g = 0; g < G; ++g)
        if (!(gs[g] > 0.0 && gs[g] < 1.0)) return MFG_ERR_INVALID;
    for (uint32_t t = 0; t < nT; ++t)
        if (targets[t] < 0 || (size_t)targets[t] >= M.F.size()) return MFG_ERR_INVALID;
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
    std::vector<Tally> tallies(nt);
    std::vector<Adjoint> adjs(nt);
    for (int i = 0; i < nt; ++i) {
        tallies[i].init(M);
        Adjoint& a = adjs[i];
        a.group = group; a.G = G; a.s = gs; a.T = nT;
        a.target_of.assign(M.F.size(), -1);
        for (uint32_t t = 0; t < nT; ++t) a.target_of[targets[t]] = (int)t;
        a.j0.assign(nT, 0.0); a.jac.assign((size_t)nT * G, 0.0);
    }
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Tracer tr(M, tallies[tid]);
        tr.adj = &adjs[tid];
        tr.ncoll.assign(G, 0);
        std::vector<TPass> passes;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < (int64_t)n_des; ++i) tr.run_particle(seed, first + (uint64_t)i, max_events, passes);
    }
    for (uint32_t t = 0; t < nT; ++t) { j0[t] = 0; for (uint32_t g = 0; g < G; ++g) jac[(size_t)t * G + g] = 0; }
    for (auto& a : adjs) {
        for (uint32_t t = 0; t < nT; ++t) j0[t] += a.j0[t];
        for (size_t i = 0; i < (size_t)nT * G; ++i) jac[i] += a.jac[i];
    }
    return 0;
}

}  // extern "C"
