// CPU oracle for the MPCD time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is a single-threaded C++ restatement of
// the reference CPU backend (azantop/cudaMPCD, src/libmpcd/cpu_backend/).  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker.  The product (cudampcd_amd) never links or calls it.
//
// What it restates (reference file:line):
//   distribute_particles      simulation_functions.cpp:7-33
//   translate_particles       simulation_functions.cpp:38-103  (standard geometry)
//   srd_collision             simulation_functions.cpp:108-202
//   extended_collision        simulation_functions.cpp:207-406
//   SymmetricMatrix::inverse  common/mechanic.hpp:52-79
//   VolumeContainer::getIndex / getPosition / getZIdx   cpu_backend/volume_container.hpp:69-95
//   MPCCell::unlockedAdd / average / averageReduceOnly  common/mpc_cell.hpp:67-86
//   CPUBackend step / stepAndAccumulateSample / getSampleMean  cpu_backend.cpp:23-122
//   Xoshiro128Plus (seed, next, uniform, int, gaussianf, unit vector)  common/random.hpp:24-258
//
// Randomness is supplied by one of two providers that are called in exactly
// the order the reference draws from its single Xoshiro128+ stream:
//   * mode 0 "reference": a restated Xoshiro128+ stream with the reference's
//     transforms (libm logf/sinf/cosf), so that, given the same seed, the
//     oracle reproduces the reference CPU trajectory bit for bit;
//   * mode 1 "counter": stateless Philox4x32-10 draws keyed by
//     (seed, step, cell, index, purpose) with float transforms built from
//     +,-,*,/,sqrt and integer bit operations only, so that a GPU using the
//     same key layout produces the same bits.
//
// Each function is pinned in tests/test_oracle_pinning.py against the compiled
// reference CPU backend, closed forms, invariants or brute force on tiny inputs.
//
// Every floating-point expression keeps the operation order of the reference
// (vector_3d.hpp operator definitions), and the file is compiled with
// -ffp-contract=off so no fused multiply-add changes a rounding.
//
// Per-cell sums come in two orders (orc_params::sum_order):
//   * 0 "serial": the reference's, members added one after another in member
//     order (table particles, then ghosts);
//   * 1 "lane groups": the GPU's, whose per-cell stages give each cell a group
//     of G = group_lanes(n) lanes: member i is added by lane i mod G, each lane
//     in member order, and the G partial sums are combined pairwise,
//     ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)), as the GPU's butterfly
//     of cross-lane moves does.  Every other operation is the serial order's.
//     The two orders differ by float rounding only; tests/test_oracle_pinning.py
//     checks order 1 against order 0 (per-particle velocities within 1e-5,
//     cell momentum, SRD energy).

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace oracle {

// ---------------------------------------------------------------- vectors ---
struct V3 {
    float x, y, z;
};
static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
// vector_3d.hpp:223-225: x*o.x + y*o.y + z*o.z (left to right)
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vector_3d.hpp:233-235
static inline V3 cross(V3 a, V3 o) {
    return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x};
}

// ------------------------------------------------------ Xoshiro128+ stream ---
// Restated from Blackman & Vigna's xoshiro128+ (random.hpp:80-92) with the
// reference's seeding (random.hpp:66-74) and float transforms (:126-164, :237-253).
struct Xoshiro {
    uint32_t s[4] = {0, 0, 0, 0};
    bool gen_f = false;
    float z1_f = 0.f;

    void seed(uint64_t a, uint64_t b) {
        s[0] = static_cast<uint32_t>(a >> 2);
        s[1] = static_cast<uint32_t>(b);
        s[2] = static_cast<uint32_t>(b >> 2);
        s[3] = static_cast<uint32_t>(a);
        z1_f = 0.f;
        gen_f = false;
    }
    static inline uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
    uint32_t next() {
        uint32_t const result = s[0] + s[3];
        uint32_t const t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 11);
        return result;
    }
    float uniform() { return static_cast<float>(next()) * (1.0f / static_cast<float>(UINT32_MAX)); }
    int uniform_int(int lo, int hi) { return static_cast<int>(next() / 2) % (hi - lo + 1) + lo; }
    float gaussian() {
        gen_f = !gen_f;
        if (!gen_f) return z1_f;
        float u1 = logf(uniform() + FLT_MIN);
        float u2 = static_cast<float>(next()) * (2.0f / static_cast<float>(UINT32_MAX));
        z1_f = sinf(u2 * float(M_PI));
        u2 = cosf(u2 * float(M_PI));
        z1_f *= sqrtf(-2 * u1);
        return u2 * sqrtf(-2 * u1);
    }
    V3 unit_vector() {
        float rsq, rd1, rd2;
        do {
            rd1 = 1.0f - 2.0f * uniform();
            rd2 = 1.0f - 2.0f * uniform();
            rsq = rd1 * rd1 + rd2 * rd2;
        } while (rsq > 1.0f);
        float rdh = 2.0f * sqrtf(1.0f - rsq);
        return {rd1 * rdh, rd2 * rdh, (1.0f - 2.0f * rsq)};
    }
};

// ------------------------------------------------------- Philox4x32-10 ---
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11.  Round: (c0,c1,c2,c3) -> (hi(B*c2)^c1^k0, lo(B*c2), hi(A*c0)^c3^k1, lo(A*c0))
// with A = 0xD2511F53, B = 0xCD9E8D57, Weyl key bumps 0x9E3779B9 / 0xBB67AE85.
struct U4 {
    uint32_t v[4];
};
static inline U4 philox(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    for (int r = 0; r < 10; ++r) {
        uint64_t const p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        uint64_t const p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        uint32_t const n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
        uint32_t const n1 = static_cast<uint32_t>(p1);
        uint32_t const n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
        uint32_t const n3 = static_cast<uint32_t>(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {{c0, c1, c2, c3}};
}

// ------------------------------------------- portable float transforms ---
// Built only from IEEE-exact operations (+,-,*,/,sqrt, floor, bit moves) so the
// same source order gives the same bits on any IEEE-754 binary32 machine.
static inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// uniform in [0,1]: same transform as the reference genUniformFloat (random.hpp:126-129)
static inline float uf(uint32_t u) { return static_cast<float>(u) * (1.0f / 4294967296.0f); }

// natural log for positive normal x: x = m*2^e, m in (sqrt(1/2), sqrt(2)],
// log(m) = 2 atanh((m-1)/(m+1)) by its odd series (fdlibm-style ln2 split).
static inline float p_logf(float x) {
    uint32_t b = f2u(x);
    int e = static_cast<int>((b >> 23) & 0xffu) - 127;
    float m = u2f((b & 0x7fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float s = f * f;
    float p = f * (2.0f + s * (0.666666667f + s * (0.4f + s * (0.285714286f + s * 0.222222222f))));
    float fe = static_cast<float>(e);
    return fe * 0.693145751953125f + (fe * 1.42860682e-06f + p);
}

// sin(pi t), cos(pi t) for t in [0, 2]: quadrant reduction then Taylor series.
static inline void p_sincospi(float t, float* s_out, float* c_out) {
    float q = floorf(2.0f * t + 0.5f);
    float r = t - 0.5f * q;                // exact, |r| <= 1/4
    float x = r * 3.14159265f;
    float x2 = x * x;
    float s = x + x * x2 * (-0.166666667f + x2 * (0.00833333333f + x2 * (-0.000198412698f + x2 * 2.75573192e-06f)));
    float c = 1.0f + x2 * (-0.5f + x2 * (0.0416666667f + x2 * (-0.00138888889f + x2 * (2.48015873e-05f + x2 * -2.75573192e-07f))));
    int qi = static_cast<int>(q) & 3;
    float so, co;
    if (qi == 0) { so = s; co = c; }
    else if (qi == 1) { so = c; co = -s; }
    else if (qi == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *s_out = so;
    *c_out = co;
}

// exp(x): x = k ln2 + r, |r| <= ln2/2, degree-7 Taylor polynomial, scale by 2^k.
static inline float p_expf(float x) {
    if (x != x) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 89.0f) return INFINITY;
    float k = floorf(x * 1.44269504f + 0.5f);
    float r = (x - k * 0.693145751953125f) - k * 1.42860682e-06f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (0.166666667f + r * (0.0416666667f + r * (0.00833333333f + r * (0.00138888889f + r * 0.000198412698f))))));
    int ki = static_cast<int>(k);
    if (ki >= -126) {
        if (ki > 127) return INFINITY;
        return p * u2f(static_cast<uint32_t>(ki + 127) << 23);
    }
    return (p * u2f(static_cast<uint32_t>(ki + 127 + 64) << 23)) * u2f(static_cast<uint32_t>(127 - 64) << 23);
}

// Box-Muller pair from two words: (cos part, sin part), same structure as the
// reference gaussianf (random.hpp:141-164).
static inline void box_muller(uint32_t a, uint32_t b, float* g0, float* g1) {
    float u1 = p_logf(uf(a) + FLT_MIN);
    float t = static_cast<float>(b) * (2.0f / 4294967296.0f);
    float s, c;
    p_sincospi(t, &s, &c);
    float rad = sqrtf(-2.0f * u1);
    *g0 = c * rad;
    *g1 = s * rad;
}

enum Purpose : uint32_t {
    P_SHIFT = 1,
    P_GHOST_COUNT = 2,
    P_GHOST_POS = 3,
    P_GHOST_VEL = 4,
    P_AXIS = 5,
    P_EXT = 6,
    P_THERM = 7,
    P_INIT_POS = 8,
    P_INIT_VEL = 9,
};

// ------------------------------------------------------ randomness source ---
struct Draws {
    int mode = 0;              // 0 reference stream, 1 counter
    Xoshiro xo;                // mode 0
    uint32_t k0 = 0, k1 = 0;   // mode 1 key
    uint32_t system = 0;       // mode 1: ensemble member id
    int ghost_xy_order = 0;    // mode 0: order of the two ghost x/y uniforms

    U4 ctr(uint32_t purpose, uint32_t c0, uint32_t cell, uint32_t step) const {
        return philox(k0, k1, c0, cell, step, (purpose << 24) | (system & 0xffffffu));
    }

    V3 shift(uint32_t step) {
        if (mode == 0) {
            float a = xo.uniform() - 0.5f;
            float b = xo.uniform() - 0.5f;
            float c = xo.uniform() - 0.5f;
            return {a, b, c};
        }
        U4 r = ctr(P_SHIFT, 0, 0, step);
        return {uf(r.v[0]) - 0.5f, uf(r.v[1]) - 0.5f, uf(r.v[2]) - 0.5f};
    }
    float ghost_count_u(uint32_t step, uint32_t cell, uint32_t i) {
        if (mode == 0) return xo.uniform();
        U4 r = ctr(P_GHOST_COUNT, i >> 2, cell, step);
        return uf(r.v[i & 3]);
    }
    // ghost position inside the cell, each coordinate in [0,1]; z is drawn
    // from the region the reference's rejection loop accepts
    // (simulation_functions.cpp:159-161).
    V3 ghost_pos(uint32_t step, uint32_t cell, uint32_t k, float shift, bool layer, bool sign) {
        if (mode == 0) {
            float z;
            do {
                z = xo.uniform();
            } while (((z < shift) ^ layer) ^ sign);
            float a = xo.uniform();
            float b = xo.uniform();
            if (ghost_xy_order == 0) return {b, a, z};  // g++ evaluates constructor arguments right to left
            return {a, b, z};
        }
        U4 r = ctr(P_GHOST_POS, k, cell, step);
        float uz = uf(r.v[2]);
        float z = (layer ^ sign) ? shift * uz : shift + (1.0f - shift) * uz;
        return {uf(r.v[0]), uf(r.v[1]), z};
    }
    V3 ghost_vel(uint32_t step, uint32_t cell, uint32_t k) {
        if (mode == 0) {
            float a = xo.gaussian();
            float b = xo.gaussian();
            float c = xo.gaussian();
            return {a, b, c};
        }
        U4 r = ctr(P_GHOST_VEL, k, cell, step);
        float g0, g1, g2, g3;
        box_muller(r.v[0], r.v[1], &g0, &g1);
        box_muller(r.v[2], r.v[3], &g2, &g3);
        return {g0, g1, g2};
    }
    // mode 0: Marsaglia's rejection (random.hpp:237-253); mode 1: the same
    // uniform distribution on the sphere from one counter draw, z uniform on
    // [-1, 1] and the azimuth uniform on [0, 2 pi) (Archimedes)
    V3 srd_axis(uint32_t step, uint32_t cell) {
        if (mode == 0) return xo.unit_vector();
        U4 r = ctr(P_AXIS, 0, cell, step);
        float const z = 1.0f - 2.0f * uf(r.v[0]);
        float s, c;
        p_sincospi(2.0f * uf(r.v[1]), &s, &c);
        float const rho = sqrtf((1.0f - z) * (1.0f + z));
        return {rho * c, rho * s, z};
    }
    // mode 1: the standard normal triple whose multiple is the sum of a SRD
    // wall cell's ghost velocities (Oracle::srd_ghost_sum)
    V3 ghost_vel_sum(uint32_t step, uint32_t cell) {
        U4 r = ctr(P_GHOST_VEL, 0, cell, step);
        float g0, g1, g2, g3;
        box_muller(r.v[0], r.v[1], &g0, &g1);
        box_muller(r.v[2], r.v[3], &g2, &g3);
        return {g0, g1, g2};
    }
    // extended: (discrete direction 0..12, sign bit) then gate uniform
    void ext_axis(uint32_t step, uint32_t cell, int* select, int* flip) {
        if (mode == 0) {
            *select = xo.uniform_int(0, 12);
            *flip = xo.uniform_int(0, 1);
            return;
        }
        U4 r = ctr(P_EXT, 0, cell, step);
        *select = static_cast<int>(r.v[0] / 2) % 13;
        *flip = static_cast<int>(r.v[1] / 2) % 2;
    }
    float ext_gate(uint32_t step, uint32_t cell) {
        if (mode == 0) return xo.uniform();
        U4 r = ctr(P_EXT, 0, cell, step);
        return uf(r.v[2]);
    }
    float ext_gauss(uint32_t step, uint32_t cell, uint32_t i) {
        if (mode == 0) return xo.gaussian();
        U4 r = ctr(P_THERM, i >> 2, cell, step);
        float g[4];
        box_muller(r.v[0], r.v[1], &g[0], &g[1]);
        box_muller(r.v[2], r.v[3], &g[2], &g[3]);
        return g[i & 3];
    }
    // attempt > 0: redraws of the channel geometry's rejection loop
    V3 init_pos_u(uint64_t i, uint32_t attempt = 0) {
        if (mode == 0) {
            float a = xo.uniform();
            float b = xo.uniform();
            float c = xo.uniform();
            return {a, b, c};
        }
        U4 r = ctr(P_INIT_POS, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), attempt);
        return {uf(r.v[0]), uf(r.v[1]), uf(r.v[2])};
    }
    V3 init_vel(uint64_t i) {
        if (mode == 0) {
            float a = xo.gaussian();
            float b = xo.gaussian();
            float c = xo.gaussian();
            return {a, b, c};
        }
        U4 r = ctr(P_INIT_VEL, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), 0);
        float g0, g1, g2, g3;
        box_muller(r.v[0], r.v[1], &g0, &g1);
        box_muller(r.v[2], r.v[3], &g2, &g3);
        return {g0, g1, g2};
    }
};

// ------------------------------------------------------------- params ---
}  // namespace oracle

extern "C" {
struct orc_params {
    float volume[3];        // cells per axis, wall layers included
    int periodic[3];
    unsigned n;             // particles per cell
    unsigned long long N;   // particles
    float delta_t;
    float drag;
    float thermal_velocity;
    float rotation_deg;     // SRD rotation angle in degrees (reference: 120)
    int algorithm;          // 0 srd, 1 extended
    int rng_mode;           // 0 reference Xoshiro128+, 1 counter (Philox)
    unsigned long long seed_a, seed_b;  // mode 0: Xoshiro seed(a,b); mode 1: Philox key = seed_a
    unsigned system;        // mode 1: ensemble member id
    int ghost_xy_order;     // mode 0 only
    int experiment;         // 0 standard, 1 channel (cylinder about the x axis)
    int sum_order;          // per-cell sums: 0 serial (reference order), 1 lane groups (GPU order)
};
}

namespace oracle {

struct Particle {
    V3 pos, vel;
    uint32_t cell;
};

// Lanes per cell of the GPU's per-cell stages: a function of the particles per
// cell only, so that results do not depend on how cells are tiled.
static inline uint32_t group_lanes(uint32_t n) { return n < 6 ? 1u : (n < 12 ? 2u : (n < 24 ? 4u : 8u)); }

// Partial sums of the G lanes of a cell group combined pairwise (G a power of
// two; G = 1 returns the one serial sum).
template <class T, class Add>
static T pairwise(std::vector<T> v, Add add) {
    for (size_t w = 1; w < v.size(); w *= 2)
        for (size_t j = 0; j + w < v.size(); j += 2 * w) v[j] = add(v[j], v[j + w]);
    return v[0];
}
static inline float addf(float a, float b) { return a + b; }

struct Sample {
    float density;
    V3 mean;
};

struct Oracle {
    orc_params p;
    V3 L;                    // volume edges
    uint32_t X, Y, Z, XY, C;
    V3 half;                 // L * 0.5
    std::vector<Particle> parts;
    std::vector<uint32_t> counter, list;
    uint32_t cap;            // 4 n table slots per cell
    Draws rng;
    uint64_t step_index = 0;  // counter-mode step key
    V3 shift = {0, 0, 0};
    bool fixed_shift = false;  // known-answer tests: use `forced` instead of a drawn shift
    V3 forced = {0, 0, 0};
    float sin_a, cos_a;
    // x body force per cell, v_x += g dt + f dt in the SRD rule (empty: none;
    // the reference has no force field)
    std::vector<float> force;
    // draws of the last collision, when logging (for injecting them into
    // another implementation): the shift; SRD: every colliding cell's axis
    // (zero for the others); extended: every colliding cell's direction
    // index, sign and gate uniform (bits) and the thermal Gaussians of its
    // members (table particles, then ghosts; 4n per cell, zero where unused);
    // both: per wall-layer cell w = (upper layer ? XY : 0) + x + X y its ghost
    // count and its n ghost slots, the first `count` drawn: velocity (SRD, 3
    // floats, thermal velocity applied) or position and velocity (extended, 6)
    bool log_draws = false;
    V3 log_shift = {0, 0, 0};
    std::vector<float> log_axes, log_gvel, log_gauss;
    std::vector<uint32_t> log_gcnt, log_ext;

    // sampling, reference semantics (cpu_backend.cpp:72-122) in FP32 Kahan
    bool sampling_done = true;
    size_t sample_counter = 0;
    std::vector<Sample> states, kahan;
    // sampling in FP64 (sum of per-step means, same S-1 normalisation)
    std::vector<double> acc_d, acc_v;

    explicit Oracle(orc_params const& q) : p(q) {
        L = {p.volume[0], p.volume[1], p.volume[2]};
        X = static_cast<uint32_t>(L.x);
        Y = static_cast<uint32_t>(L.y);
        Z = static_cast<uint32_t>(L.z);
        XY = X * Y;
        C = XY * Z;
        half = mul(L, 0.5f);
        cap = 4 * p.n;
        parts.resize(p.N);
        counter.assign(C, 0);
        list.assign(static_cast<size_t>(C) * cap, 0);
        states.assign(C, {0, {0, 0, 0}});
        kahan.assign(C, {0, {0, 0, 0}});
        acc_d.assign(C, 0.0);
        acc_v.assign(3 * static_cast<size_t>(C), 0.0);
        rng.mode = p.rng_mode;
        rng.ghost_xy_order = p.ghost_xy_order;
        if (p.rng_mode == 0) {
            rng.xo.seed(p.seed_a, p.seed_b);
        } else {
            rng.k0 = static_cast<uint32_t>(p.seed_a);
            rng.k1 = static_cast<uint32_t>(p.seed_a >> 32);
            rng.system = p.system;
        }
        // simulation_functions.cpp:118-119 (angle fixed to 120 degrees there)
        sin_a = sinf((M_PI * static_cast<double>(p.rotation_deg)) / 180);
        cos_a = cosf((M_PI * static_cast<double>(p.rotation_deg)) / 180);
    }

    // volume_container.hpp:69-75
    uint32_t index(V3 r) const {
        uint32_t x = static_cast<uint32_t>(floorf(r.x + half.x)) % X;
        uint32_t y = static_cast<uint32_t>(floorf(r.y + half.y)) % Y;
        uint32_t z = static_cast<uint32_t>(floorf(r.z + half.z)) % Z;
        return x + (y * X) + (z * XY);
    }
    // volume_container.hpp:86-91
    V3 cell_centre(uint32_t idx) const {
        V3 v = {0.5f + static_cast<float>(idx % X), 0.5f + static_cast<float>((idx % XY) / X),
                0.5f + static_cast<float>(idx / XY)};
        return sub(v, half);
    }
    float zidx(uint32_t idx) const { return static_cast<float>(idx / XY); }

    // simulation_functions.cpp:46-51
    V3 wrap(V3 r) const {
        r.x = fmodf(r.x + 1.5f * L.x, L.x) - L.x * 0.5f;
        r.y = fmodf(r.y + 1.5f * L.y, L.y) - L.y * 0.5f;
        r.z = fmodf(r.z + 1.5f * L.z, L.z) - L.z * 0.5f;
        return r;
    }

    // simulation_functions.cpp:7-33 (standard geometry)
    void distribute() {
        V3 scale = {L.x - static_cast<float>(2 * (1 - p.periodic[0])),
                    L.y - static_cast<float>(2 * (1 - p.periodic[1])),
                    L.z - static_cast<float>(2 * (1 - p.periodic[2]))};
        // channel: rejection outside the radius taken from the x extent (:11, :23-24)
        float const r2 = (p.volume[0] - 2) * (p.volume[0] - 2) * 0.25f;
        for (size_t i = 0; i < parts.size(); ++i) {
            Particle q = {};
            for (uint32_t attempt = 0;; ++attempt) {
                V3 u = rng.init_pos_u(i, attempt);
                V3 r = sub(u, V3{0.5f, 0.5f, 0.5f});
                q.pos = {r.x * scale.x, r.y * scale.y, r.z * scale.z};
                if (!(p.experiment == 1 && (q.pos.z * q.pos.z + q.pos.y * q.pos.y) > r2)) break;
            }
            q.vel = mul(rng.init_vel(i), p.thermal_velocity);
            q.cell = index(q.pos);
            parts[i] = q;
        }
    }

    // cpu_backend.cpp:23-33 + simulation_functions.cpp:38-103
    void translate() {
        shift = fixed_shift ? forced : rng.shift(static_cast<uint32_t>(step_index));
        std::fill(counter.begin(), counter.end(), 0u);
        float const dt = p.delta_t;
        // channel: radius from the z extent (:45), unlike distribute
        float const r2 = (L.z - 2) * (L.z - 2) * 0.25f;
        for (size_t idx = 0; idx < parts.size(); ++idx) {
            Particle q = parts[idx];
            if (p.experiment == 1) {
                // simulation_functions.cpp:75-82: undo a move that leaves the cylinder, reverse the velocity
                q.pos = add(q.pos, mul(q.vel, dt));
                if ((q.pos.z * q.pos.z + q.pos.y * q.pos.y) > r2) {
                    q.pos = sub(q.pos, mul(q.vel, dt));
                    q.vel = neg(q.vel);
                }
            } else if (!p.periodic[2]) {
                float z_wall = (0.5f * L.z - 1);
                float next_z = q.pos.z + q.vel.z * dt;
                if (fabsf(next_z) > z_wall) {
                    float t = (next_z > 0 ? z_wall - q.pos.z : -z_wall - q.pos.z) / q.vel.z;
                    q.pos = add(q.pos, mul(q.vel, t));
                    q.vel = neg(q.vel);
                    q.pos = add(q.pos, mul(q.vel, dt - t));
                } else {
                    q.pos = add(q.pos, mul(q.vel, dt));
                }
            } else {
                q.pos = add(q.pos, mul(q.vel, dt));
            }
            q.pos = wrap(add(q.pos, shift));
            q.cell = index(q.pos);
            uint32_t slot = counter[q.cell];
            counter[q.cell] += 1;
            if (slot < cap)
                list[q.cell + static_cast<size_t>(slot) * C] = static_cast<uint32_t>(idx);
            else
                q.pos = wrap(sub(q.pos, shift));
            parts[idx] = q;
        }
    }

    // wall-layer test shared by both collision rules (simulation_functions.cpp:137-138)
    void wall_flags(uint32_t c, bool sign, bool* layer, bool* ghosts) const {
        *layer = (zidx(c) == (L.z - (sign ? 1 : 2)));
        *ghosts = (!p.periodic[2]) && ((zidx(c) == static_cast<float>(sign)) || *layer);
    }

    // Counter stream, SRD rule: the sum of a wall cell's `added` ghost
    // velocities (each Maxwell-Boltzmann at the thermal velocity,
    // simulation_functions.cpp:165) drawn at once, as a sum of independent
    // Gaussians is one Gaussian of the summed variance: sqrt(added) times the
    // thermal velocity times a standard normal triple.  The SRD collision uses
    // ghosts only through this sum (the cell mean; getCorrection returns the
    // mean, mpc_cell.hpp:82-84).
    V3 srd_ghost_sum(uint32_t step, uint32_t c, uint32_t added) {
        return mul(rng.ghost_vel_sum(step, c), p.thermal_velocity * sqrtf(static_cast<float>(added)));
    }

    // simulation_functions.cpp:108-202
    void srd() {
        float const sh = fabsf(shift.z);
        bool const sign = shift.z > 0;
        float const drag = p.drag * p.delta_t;
        uint32_t const step = static_cast<uint32_t>(step_index);
        std::vector<V3> bp(cap + p.n + 1), bv(cap + p.n + 1);
        std::vector<int64_t> bidx(cap + p.n + 1);
        uint32_t const G = p.sum_order ? group_lanes(p.n) : 1u;
        bool const cstream = rng.mode == 1;   // counter stream
        if (log_draws) {
            log_shift = shift;
            log_axes.assign(3 * static_cast<size_t>(C), 0.0f);
            log_gcnt.assign(2 * static_cast<size_t>(XY), 0u);
            log_gvel.assign(2 * static_cast<size_t>(XY) * p.n * 3, 0.0f);
        }
        for (uint32_t c = 0; c < C; ++c) {
            uint32_t m = std::min(cap, counter[c]);
            bool layer, ghosts;
            wall_flags(c, sign, &layer, &ghosts);
            uint32_t added = 0;
            if (ghosts)
                for (uint32_t i = 0; i < p.n; ++i)
                    if ((rng.ghost_count_u(step, c, i) > sh) ^ layer ^ sign) ++added;
            uint32_t total = m + added;
            size_t const w = (layer ? XY : 0u) + c % XY;   // wall-layer cell (when ghosts)
            if (log_draws && ghosts) log_gcnt[w] = added;
            for (uint32_t i = 0; i < m; ++i) {
                bidx[i] = list[c + static_cast<size_t>(i) * C];
                bp[i] = parts[bidx[i]].pos;
                bv[i] = parts[bidx[i]].vel;
            }
            // the counter stream keeps only the ghosts' velocity sum (they
            // enter the collision through the cell mean alone); the reference
            // stream draws every ghost, in the reference's order
            V3 const gsum = cstream && added ? srd_ghost_sum(step, c, added) : V3{0, 0, 0};
            uint32_t const members = cstream ? m : total;   // held in the buffers
            if (ghosts && !cstream) {
                V3 centre = cell_centre(c);
                for (uint32_t i = m; i < total; ++i) {
                    V3 g = rng.ghost_pos(step, c, i - m, sh, layer, sign);
                    bidx[i] = -1;
                    bp[i] = add(V3{g.x - 0.5f, g.y - 0.5f, g.z - 0.5f}, centre);
                    bv[i] = mul(rng.ghost_vel(step, c, i - m), p.thermal_velocity);
                    if (log_draws) {
                        float* const gv = &log_gvel[(w * p.n + (i - m)) * 3];
                        gv[0] = bv[i].x; gv[1] = bv[i].y; gv[2] = bv[i].z;
                    }
                }
            }
            if (total > 1) {
                std::vector<V3> part(G, V3{0, 0, 0});
                for (uint32_t i = 0; i < members; ++i) part[i % G] = add(part[i % G], bv[i]);
                V3 sum = pairwise(part, add);
                if (cstream && added) sum = add(sum, gsum);
                float inv = 1.0f / static_cast<float>(total);
                V3 mean = mul(sum, inv);
                V3 axis = rng.srd_axis(step, c);
                if (log_draws) {
                    log_axes[3 * c] = axis.x; log_axes[3 * c + 1] = axis.y; log_axes[3 * c + 2] = axis.z;
                }
                for (uint32_t i = 0; i < members; ++i) {
                    V3 v = sub(bv[i], mean);
                    V3 vpar = mul(axis, dot(v, axis));
                    V3 vperp = sub(v, vpar);
                    bv[i] = add(add(vpar, mul(vperp, cos_a)), mul(cross(vperp, axis), sin_a));
                }
                float const fdt = force.empty() ? 0.0f : force[c] * p.delta_t;
                for (uint32_t i = 0; i < members; ++i) {
                    bv[i] = add(bv[i], mean);
                    bp[i] = wrap(sub(bp[i], shift));
                    if (force.empty()) bv[i].x += drag; else bv[i].x += drag + fdt;
                }
            }
            for (uint32_t i = 0; i < m; ++i) {
                Particle& q = parts[bidx[i]];
                q.pos = bp[i];
                q.vel = bv[i];
                q.cell = index(bp[i]);
            }
        }
    }

    // the 13 discrete directions of simulation_functions.cpp:272-293
    static V3 ext_direction(int select) {
        int const steps = 4;
        float theta, phi;
        int phi_i = select % steps + 1, theta_i = select / steps + 1;
        if ((theta_i % 2) && (phi_i % 2))
            theta = theta_i == 1 ? 0.95531661812450927816f : float(M_PI) - 0.95531661812450927816f;
        else
            theta = theta_i * (float(M_PI) / steps);
        phi = phi_i * (float(M_PI) / steps);
        return {sinf(theta) * cosf(phi), sinf(theta) * sinf(phi), cosf(theta)};
    }

    struct Sym {
        float xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
    };
    // mechanic.hpp:26 and :52-79
    static Sym inverse(Sym const& m, float tol) {
        float det = m.xx * (m.yy * m.zz - m.yz * m.yz) + m.xz * (2 * m.xy * m.yz - m.yy * m.xz) - (m.xy * m.xy * m.zz);
        if (det < 0) det = 0;
        Sym o;
        if (det > tol) {
            float id = 1 / det;
            o.xx = (m.yy * m.zz - m.yz * m.yz) * id;
            o.yy = (m.xx * m.zz - m.xz * m.xz) * id;
            o.zz = (m.xx * m.yy - m.xy * m.xy) * id;
            o.xy = (m.xz * m.yz - m.zz * m.xy) * id;
            o.xz = (m.xy * m.yz - m.yy * m.xz) * id;
            o.yz = (m.xy * m.xz - m.xx * m.yz) * id;
            return o;
        }
        float tr = m.xx + m.yy + m.zz;
        if (tr > tol) {
            float s = 4 / (tr * tr);
            o.xx = m.xx * s; o.yy = m.yy * s; o.zz = m.zz * s;
            o.xy = m.xy * s; o.xz = m.xz * s; o.yz = m.yz * s;
        }
        return o;
    }
    static V3 apply(Sym const& m, V3 v) {
        return {(m.xx * v.x) + (m.xy * v.y) + (m.xz * v.z), (m.xy * v.x) + (m.yy * v.y) + (m.yz * v.z),
                (m.xz * v.x) + (m.yz * v.y) + (m.zz * v.z)};
    }

    // simulation_functions.cpp:207-406
    void extended() {
        float const sh = fabsf(shift.z);
        bool const sign = shift.z > 0;
        float const scale = 0.01f;
        float const drag = p.drag * p.delta_t;
        float const tv = p.thermal_velocity;
        uint32_t const step = static_cast<uint32_t>(step_index);
        std::vector<V3> bp(cap + p.n + 1), bv(cap + p.n + 1);
        std::vector<int64_t> bidx(cap + p.n + 1);
        std::vector<uint8_t> side(cap + p.n + 1);
        uint32_t const G = p.sum_order ? group_lanes(p.n) : 1u;
        if (log_draws) {
            log_shift = shift;
            log_ext.assign(3 * static_cast<size_t>(C), 0u);
            log_gauss.assign(static_cast<size_t>(C) * cap, 0.0f);
            log_gcnt.assign(2 * static_cast<size_t>(XY), 0u);
            log_gvel.assign(2 * static_cast<size_t>(XY) * p.n * 6, 0.0f);
        }
        for (uint32_t c = 0; c < C; ++c) {
            uint32_t m = std::min(cap, counter[c]);
            bool layer, ghosts;
            wall_flags(c, sign, &layer, &ghosts);
            uint32_t added = 0;
            if (ghosts)
                for (uint32_t i = 0; i < p.n; ++i)
                    if ((rng.ghost_count_u(step, c, i) > sh) ^ layer ^ sign) ++added;
            size_t const w = (layer ? XY : 0u) + c % XY;   // wall-layer cell (when ghosts)
            if (log_draws && ghosts) log_gcnt[w] = added;
            uint32_t total = std::min(m + added, cap);
            uint32_t reals = total - added;   // may drop table entries when ghosts fill the cap
            for (uint32_t i = 0; i < reals; ++i) {
                bidx[i] = list[c + static_cast<size_t>(i) * C];
                bp[i] = parts[bidx[i]].pos;
                bv[i] = parts[bidx[i]].vel;
            }
            V3 offset = cell_centre(c);
            if (ghosts) {
                for (uint32_t i = reals; i < total; ++i) {
                    V3 g = rng.ghost_pos(step, c, i - reals, sh, layer, sign);
                    bidx[i] = -1;
                    bp[i] = add(V3{g.x - 0.5f, g.y - 0.5f, g.z - 0.5f}, offset);
                    bv[i] = mul(rng.ghost_vel(step, c, i - reals), tv);
                    if (log_draws) {
                        float* const gv = &log_gvel[(w * p.n + (i - reals)) * 6];
                        gv[0] = bp[i].x; gv[1] = bp[i].y; gv[2] = bp[i].z;
                        gv[3] = bv[i].x; gv[4] = bv[i].y; gv[5] = bv[i].z;
                    }
                }
            }
            if (total > 1) {
                int select, flip;
                rng.ext_axis(step, c, &select, &flip);
                if (log_draws) {
                    log_ext[3 * c] = static_cast<uint32_t>(select);
                    log_ext[3 * c + 1] = static_cast<uint32_t>(flip);
                }
                V3 axis = ext_direction(select);
                if (flip) axis = neg(axis);

                // per-lane partial sums (one lane: the serial order)
                std::vector<V3> pcom(G, V3{0, 0, 0}), pmean(G, V3{0, 0, 0});
                for (uint32_t i = 0; i < total; ++i) {
                    pcom[i % G] = add(pcom[i % G], sub(bp[i], offset));
                    pmean[i % G] = add(pmean[i % G], bv[i]);
                }
                V3 com = pairwise(pcom, add), mean = pairwise(pmean, add);
                com = add(mul(com, 1.0f / total), offset);
                mean = mul(mean, 1.0f / total);

                float proj0 = 0.0f, proj1 = 0.0f;
                int size0 = 0;
                std::vector<float> pproj(G, 0.0f);
                for (uint32_t i = 0; i < total; ++i) {
                    bp[i] = sub(bp[i], com);
                    bv[i] = sub(bv[i], mean);
                    side[i] = dot(bp[i], axis) < 0;
                    if (side[i]) {
                        pproj[i % G] += dot(bv[i], axis);
                        size0 += 1;
                    }
                }
                proj0 = pairwise(pproj, addf);
                int size1 = static_cast<int>(total) - size0;
                proj1 = -proj0 / size1;
                proj0 = proj0 / size0;
                float arg = scale * (proj1 - proj0) * size0 * size1;
                float probability = 1 - (p.rng_mode == 0 ? std::exp(arg) : p_expf(arg));

                V3 dL = {0, 0, 0};
                float const gate = rng.ext_gate(step, c);
                if (log_draws) log_ext[3 * c + 2] = f2u(gate);
                bool collide = gate < probability;
                if (collide) {
                    float t0 = proj1 * (float(size1) / size0), t1 = proj0 * (float(size0) / size1);
                    // per lane: the two side sums, the angular momentum (added
                    // to here, subtracted from in the next pass) and the inertia
                    std::vector<float> pm0(G, 0.0f), pm1(G, 0.0f);
                    std::vector<V3> pL(G, V3{0, 0, 0});
                    std::vector<Sym> pI(G);
                    for (uint32_t i = 0; i < total; ++i) {
                        bool s = side[i];
                        float const xi = rng.ext_gauss(step, c, i);
                        if (log_draws) log_gauss[static_cast<size_t>(c) * cap + i] = xi;
                        float vr = xi * tv;
                        (s ? pm0 : pm1)[i % G] += vr;
                        pL[i % G] = add(pL[i % G], cross(bp[i], bv[i]));
                        bv[i] = add(bv[i], mul(axis, vr - dot(bv[i], axis) + (s ? t0 : t1)));
                        V3 q = {bp[i].x * bp[i].x, bp[i].y * bp[i].y, bp[i].z * bp[i].z};
                        Sym& I = pI[i % G];
                        I.xx += q.y + q.z;
                        I.yy += q.x + q.z;
                        I.zz += q.x + q.y;
                        I.xy += -bp[i].x * bp[i].y;
                        I.xz += -bp[i].x * bp[i].z;
                        I.yz += -bp[i].y * bp[i].z;
                    }
                    float const mean0 = pairwise(pm0, addf), mean1 = pairwise(pm1, addf);
                    V3 vm0 = mul(axis, mean0 / size0), vm1 = mul(axis, mean1 / size1);
                    for (uint32_t i = 0; i < total; ++i) {
                        bv[i] = sub(bv[i], side[i] ? vm0 : vm1);
                        pL[i % G] = sub(pL[i % G], cross(bp[i], bv[i]));
                    }
                    dL = pairwise(pL, add);
                    Sym I = pairwise(pI, [](Sym a, Sym b) {
                        Sym o;
                        o.xx = a.xx + b.xx; o.yy = a.yy + b.yy; o.zz = a.zz + b.zz;
                        o.xy = a.xy + b.xy; o.xz = a.xz + b.xz; o.yz = a.yz + b.yz;
                        return o;
                    });
                    I = inverse(I, 1e-5f);
                    dL = apply(I, dL);
                }
                for (uint32_t i = 0; i < reals; ++i) {
                    bv[i] = add(bv[i], mean);
                    bv[i].x += drag;
                    bv[i] = sub(bv[i], cross(bp[i], dL));
                    bp[i] = wrap(sub(add(com, bp[i]), shift));
                }
            }
            for (uint32_t i = 0; i < reals; ++i) {
                Particle& q = parts[bidx[i]];
                q.pos = bp[i];
                q.vel = bv[i];
                q.cell = index(bp[i]);
            }
        }
    }

    void collide() {
        if (p.algorithm == 0) srd(); else extended();
    }

    void step(int n) {
        for (int i = 0; i < n; ++i) {
            translate();
            collide();
            ++step_index;
        }
    }

    // cpu_backend.cpp:72-109 plus an FP64 accumulator of the same per-step fields
    void step_and_sample(int n) {
        std::vector<uint32_t> dens(C);
        std::vector<V3> sum(C);
        std::vector<double> sd(3 * static_cast<size_t>(C));
        for (int s = 0; s < n; ++s) {
            step(1);
            std::fill(dens.begin(), dens.end(), 0u);
            std::fill(sum.begin(), sum.end(), V3{0, 0, 0});
            std::fill(sd.begin(), sd.end(), 0.0);
            for (auto const& q : parts) {
                uint32_t c = index(q.pos);
                dens[c] += 1;
                sum[c] = add(sum[c], q.vel);
                sd[3 * c + 0] += q.vel.x;
                sd[3 * c + 1] += q.vel.y;
                sd[3 * c + 2] += q.vel.z;
            }
            std::vector<Sample> cur(C);
            for (uint32_t c = 0; c < C; ++c) {
                V3 mv = sum[c];
                if (dens[c] > 0) mv = mul(mv, 1.0f / static_cast<float>(dens[c]));  // mpc_cell.hpp:84
                cur[c] = {static_cast<float>(dens[c]), mv};
            }
            bool first = sampling_done;
            if (sampling_done) {
                sampling_done = false;
                sample_counter = 0;
                std::fill(kahan.begin(), kahan.end(), Sample{0, {0, 0, 0}});
                states = cur;
            } else {
                sample_counter++;
                for (uint32_t c = 0; c < C; ++c) {
                    {
                        float y = cur[c].density - kahan[c].density;
                        float t = states[c].density + y;
                        kahan[c].density = (t - states[c].density) - y;
                        states[c].density = t;
                    }
                    V3 y = sub(cur[c].mean, kahan[c].mean);
                    V3 t = add(states[c].mean, y);
                    kahan[c].mean = sub(sub(t, states[c].mean), y);
                    states[c].mean = t;
                }
            }
            for (uint32_t c = 0; c < C; ++c) {
                double d = static_cast<double>(dens[c]);
                double inv = dens[c] > 0 ? 1.0 / d : 0.0;
                if (first) {
                    acc_d[c] = d;
                    for (int k = 0; k < 3; ++k) acc_v[3 * c + k] = sd[3 * c + k] * inv;
                } else {
                    acc_d[c] += d;
                    for (int k = 0; k < 3; ++k) acc_v[3 * c + k] += sd[3 * c + k] * inv;
                }
            }
        }
    }

    // Frozen-assignment model of an SRD run (counter mode), for checking a
    // pathwise gradient: the real trajectory is stepped `steps` times with the
    // oracle's own parameters, and beside it a double-precision "shadow"
    // velocity of every particle follows the same cells, tables, bounces, axes
    // and ghosts with the drag g, rotation angle and force field given here:
    //   v' = -v if the particle bounced, else v;
    //   table particles of cells with total > 1: v'' = u + R(angle)(v' - u) + (g + f_c) dt e_x,
    //   u = (sum of table v' + ghost velocities) / total;   others: v'' = v'.
    // Returns L = mean over the z layers [zlo, zhi) of (profile(z) - target(z))^2
    // with profile(z) the mean over steps of the layer's mean x velocity
    // (layer = z cell of the real position after the step), computed from the
    // shadow velocities, or from the real ones when use_real is set.  The
    // oracle's state is restored afterwards.
    double frozen_profile_loss(int steps, double g, double angle_deg, const double* f, const double* target,
                               uint32_t zlo, uint32_t zhi, int use_real, double* profile_out) {
        if (p.rng_mode != 1 || p.algorithm != 0) return NAN;  // the shadow redraws: stateless draws, SRD only
        std::vector<Particle> const saved = parts;
        uint64_t const saved_step = step_index;
        size_t const Np = parts.size();
        std::vector<double> sv(3 * Np);
        for (size_t i = 0; i < Np; ++i) {
            sv[3 * i] = parts[i].vel.x;
            sv[3 * i + 1] = parts[i].vel.y;
            sv[3 * i + 2] = parts[i].vel.z;
        }
        double const ca = std::cos(M_PI * angle_deg / 180.0), sa = std::sin(M_PI * angle_deg / 180.0);
        double const dt = p.delta_t;
        std::vector<double> prof(zhi - zlo, 0.0);
        std::vector<uint8_t> bounced(Np);
        for (int t = 0; t < steps; ++t) {
            if (!p.periodic[2]) {
                float const z_wall = (0.5f * L.z - 1);
                for (size_t i = 0; i < Np; ++i)
                    bounced[i] = fabsf(parts[i].pos.z + parts[i].vel.z * p.delta_t) > z_wall;
            }
            translate();
            for (size_t i = 0; i < Np; ++i)
                if (bounced[i])
                    for (int k = 0; k < 3; ++k) sv[3 * i + k] = -sv[3 * i + k];
            // shadow collision with the draws of this step
            float const sh = fabsf(shift.z);
            bool const sign = shift.z > 0;
            uint32_t const step = static_cast<uint32_t>(step_index);
            for (uint32_t c = 0; c < C; ++c) {
                uint32_t const m = std::min(cap, counter[c]);
                bool layer, ghosts;
                wall_flags(c, sign, &layer, &ghosts);
                uint32_t added = 0;
                if (ghosts)
                    for (uint32_t i = 0; i < p.n; ++i)
                        if ((rng.ghost_count_u(step, c, i) > sh) ^ layer ^ sign) ++added;
                uint32_t const total = m + added;
                if (total <= 1) continue;
                double u[3] = {0, 0, 0};
                for (uint32_t i = 0; i < m; ++i) {
                    size_t const q = list[c + static_cast<size_t>(i) * C];
                    for (int k = 0; k < 3; ++k) u[k] += sv[3 * q + k];
                }
                if (added) {
                    V3 const gv = srd_ghost_sum(step, c, added);
                    u[0] += gv.x;
                    u[1] += gv.y;
                    u[2] += gv.z;
                }
                for (double& x : u) x /= total;
                V3 const a3 = rng.srd_axis(step, c);
                double const a[3] = {a3.x, a3.y, a3.z};
                double const kick = (g + (f ? f[c] : 0.0)) * dt;
                for (uint32_t i = 0; i < m; ++i) {
                    size_t const q = list[c + static_cast<size_t>(i) * C];
                    double w[3], par[3], perp[3];
                    for (int k = 0; k < 3; ++k) w[k] = sv[3 * q + k] - u[k];
                    double const d = w[0] * a[0] + w[1] * a[1] + w[2] * a[2];
                    for (int k = 0; k < 3; ++k) {
                        par[k] = a[k] * d;
                        perp[k] = w[k] - par[k];
                    }
                    double const cr[3] = {perp[1] * a[2] - perp[2] * a[1], perp[2] * a[0] - perp[0] * a[2],
                                          perp[0] * a[1] - perp[1] * a[0]};
                    for (int k = 0; k < 3; ++k) sv[3 * q + k] = u[k] + par[k] + ca * perp[k] + sa * cr[k];
                    sv[3 * q] += kick;
                }
            }
            srd();
            ++step_index;
            std::vector<double> lsum(Z, 0.0);
            std::vector<uint32_t> lcnt(Z, 0u);
            for (size_t i = 0; i < Np; ++i) {
                uint32_t const iz = static_cast<uint32_t>(floorf(parts[i].pos.z + half.z)) % Z;
                lcnt[iz] += 1;
                lsum[iz] += use_real ? static_cast<double>(parts[i].vel.x) : sv[3 * i];
            }
            for (uint32_t z = zlo; z < zhi; ++z)
                if (lcnt[z]) prof[z - zlo] += lsum[z] / lcnt[z] / steps;
        }
        double loss = 0.0;
        for (uint32_t z = zlo; z < zhi; ++z) {
            double const d = prof[z - zlo] - target[z - zlo];
            loss += d * d / (zhi - zlo);
        }
        if (profile_out) std::copy(prof.begin(), prof.end(), profile_out);
        parts = saved;
        step_index = saved_step;
        return loss;
    }

    // cpu_backend.cpp:111-122 (sum over S samples times 1/(S-1))
    void sample_mean(float* dens, float* vel) {
        double inverse = 1.0 / sample_counter;
        for (uint32_t c = 0; c < C; ++c) {
            dens[c] = states[c].density * inverse;
            vel[3 * c + 0] = states[c].mean.x * inverse;
            vel[3 * c + 1] = states[c].mean.y * inverse;
            vel[3 * c + 2] = states[c].mean.z * inverse;
        }
        sampling_done = true;
    }
    void sample_mean_f64(double* dens, double* vel) const {
        double inverse = 1.0 / static_cast<double>(sample_counter);
        for (uint32_t c = 0; c < C; ++c) {
            dens[c] = acc_d[c] * inverse;
            for (int k = 0; k < 3; ++k) vel[3 * c + k] = acc_v[3 * c + k] * inverse;
        }
    }
};

}  // namespace oracle

using oracle::Oracle;

extern "C" {

void* orc_create(const orc_params* p) { return new Oracle(*p); }
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }
void orc_distribute(void* h) { static_cast<Oracle*>(h)->distribute(); }
unsigned long long orc_n_particles(void* h) { return static_cast<Oracle*>(h)->parts.size(); }
unsigned orc_n_cells(void* h) { return static_cast<Oracle*>(h)->C; }
void orc_set_step(void* h, unsigned long long s) { static_cast<Oracle*>(h)->step_index = s; }
unsigned long long orc_get_step(void* h) { return static_cast<Oracle*>(h)->step_index; }

void orc_set_particles(void* h, const float* pos, const float* vel, unsigned long long n) {
    auto* o = static_cast<Oracle*>(h);
    o->parts.resize(n);
    for (size_t i = 0; i < n; ++i) {
        o->parts[i].pos = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        o->parts[i].vel = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        o->parts[i].cell = o->index(o->parts[i].pos);
    }
}
void orc_get_particles(void* h, float* pos, float* vel) {
    auto* o = static_cast<Oracle*>(h);
    for (size_t i = 0; i < o->parts.size(); ++i) {
        auto const& q = o->parts[i];
        if (pos) { pos[3 * i] = q.pos.x; pos[3 * i + 1] = q.pos.y; pos[3 * i + 2] = q.pos.z; }
        if (vel) { vel[3 * i] = q.vel.x; vel[3 * i + 1] = q.vel.y; vel[3 * i + 2] = q.vel.z; }
    }
}

// streaming half-step: draws the shift, streams, fills counter/list (no collision)
void orc_translate(void* h, float* shift_out) {
    auto* o = static_cast<Oracle*>(h);
    o->translate();
    if (shift_out) { shift_out[0] = o->shift.x; shift_out[1] = o->shift.y; shift_out[2] = o->shift.z; }
}
void orc_force_shift(void* h, const float* s) {
    auto* o = static_cast<Oracle*>(h);
    o->fixed_shift = s != nullptr;
    if (s) o->forced = {s[0], s[1], s[2]};
}
void orc_collide(void* h) {
    auto* o = static_cast<Oracle*>(h);
    o->collide();
    ++o->step_index;
}
// per-particle cell index, per-cell counter, and the cell list flattened in
// cell order (first min(counter, 4n) entries of each cell); returns entries written
unsigned long long orc_get_cells(void* h, unsigned* cell_of_particle, unsigned* counter, unsigned* flat_list) {
    auto* o = static_cast<Oracle*>(h);
    for (size_t i = 0; i < o->parts.size(); ++i) cell_of_particle[i] = o->parts[i].cell;
    std::memcpy(counter, o->counter.data(), sizeof(unsigned) * o->C);
    size_t k = 0;
    for (uint32_t c = 0; c < o->C; ++c) {
        uint32_t m = std::min(o->cap, o->counter[c]);
        for (uint32_t i = 0; i < m; ++i) flat_list[k++] = o->list[c + static_cast<size_t>(i) * o->C];
    }
    return k;
}
void orc_step(void* h, int n) { static_cast<Oracle*>(h)->step(n); }
// x body force per cell (n_cells floats; NULL removes it)
// draws of the SRD collisions from now on are logged (on != 0); the last
// collision's: shift[3], axes[3 C] (reference cell order), ghost counts[2 XY]
// and ghost velocities[2 XY n 3] (wall-layer cell w = (upper ? XY : 0) + x + X y)
void orc_log_draws(void* h, int on) { static_cast<Oracle*>(h)->log_draws = on != 0; }
void orc_get_draws(void* h, float* shift, float* axes, unsigned* gcnt, float* gvel) {
    auto* o = static_cast<Oracle*>(h);
    shift[0] = o->log_shift.x; shift[1] = o->log_shift.y; shift[2] = o->log_shift.z;
    std::copy(o->log_axes.begin(), o->log_axes.end(), axes);
    std::copy(o->log_gcnt.begin(), o->log_gcnt.end(), gcnt);
    std::copy(o->log_gvel.begin(), o->log_gvel.end(), gvel);
}
// extended rule: shift[3], per cell (direction index, sign, gate bits)[3 C],
// thermal Gaussians[C 4n], ghost counts[2 XY], ghost positions and velocities[2 XY n 6]
void orc_get_ext_draws(void* h, float* shift, unsigned* ext, float* gauss, unsigned* gcnt, float* gvel) {
    auto* o = static_cast<Oracle*>(h);
    shift[0] = o->log_shift.x; shift[1] = o->log_shift.y; shift[2] = o->log_shift.z;
    std::copy(o->log_ext.begin(), o->log_ext.end(), ext);
    std::copy(o->log_gauss.begin(), o->log_gauss.end(), gauss);
    std::copy(o->log_gcnt.begin(), o->log_gcnt.end(), gcnt);
    std::copy(o->log_gvel.begin(), o->log_gvel.end(), gvel);
}
void orc_set_force(void* h, const float* f) {
    Oracle* o = static_cast<Oracle*>(h);
    if (f) o->force.assign(f, f + o->C); else o->force.clear();
}
double orc_frozen_profile_loss(void* h, int steps, double drag, double angle_deg, const double* force,
                               const double* target, unsigned zlo, unsigned zhi, int use_real, double* profile) {
    return static_cast<Oracle*>(h)->frozen_profile_loss(steps, drag, angle_deg, force, target, zlo, zhi, use_real, profile);
}
void orc_step_and_sample(void* h, int n) { static_cast<Oracle*>(h)->step_and_sample(n); }
void orc_get_sample_mean(void* h, float* dens, float* vel) { static_cast<Oracle*>(h)->sample_mean(dens, vel); }
void orc_get_sample_mean_f64(void* h, double* dens, double* vel) { static_cast<Oracle*>(h)->sample_mean_f64(dens, vel); }
unsigned long long orc_sample_count(void* h) { return static_cast<Oracle*>(h)->sample_counter; }

// building blocks, exported for the known-answer tests
void orc_philox(unsigned k0, unsigned k1, const unsigned* ctr, unsigned* out) {
    oracle::U4 r = oracle::philox(k0, k1, ctr[0], ctr[1], ctr[2], ctr[3]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}
float orc_logf(float x) { return oracle::p_logf(x); }
float orc_expf(float x) { return oracle::p_expf(x); }
void orc_sincospi(float t, float* s, float* c) { oracle::p_sincospi(t, s, c); }
void orc_box_muller(unsigned a, unsigned b, float* g0, float* g1) { oracle::box_muller(a, b, g0, g1); }
void orc_xoshiro_uniform(unsigned long long a, unsigned long long b, unsigned long long n, float* out) {
    oracle::Xoshiro x;
    x.seed(a, b);
    for (size_t i = 0; i < n; ++i) out[i] = x.uniform();
}
void orc_xoshiro_gaussian(unsigned long long a, unsigned long long b, unsigned long long n, float* out) {
    oracle::Xoshiro x;
    x.seed(a, b);
    for (size_t i = 0; i < n; ++i) out[i] = x.gaussian();
}
// counter-mode draws, for statistics tests
void orc_counter_draws(void* h, int purpose, unsigned step, unsigned cell, unsigned count, float* out) {
    auto* o = static_cast<Oracle*>(h);
    for (unsigned i = 0; i < count; ++i) {
        if (purpose == oracle::P_THERM) out[i] = o->rng.ext_gauss(step, cell, i);
        else if (purpose == oracle::P_GHOST_COUNT) out[i] = o->rng.ghost_count_u(step, cell, i);
        else if (purpose == oracle::P_AXIS) {
            oracle::V3 a = o->rng.srd_axis(step, cell + i);
            out[3 * i] = a.x; out[3 * i + 1] = a.y; out[3 * i + 2] = a.z;
        }
    }
}
// what the SRD rule adds to a wall cell's velocity sum for `added` ghosts, for
// `count` cells from `cell`: mode 1 the counter stream's one-draw sum, mode 0
// the sum of `added` thermal velocities drawn one by one from the reference
// stream (its Maxwell-Boltzmann triples times the thermal velocity, summed in
// draw order), the reference's construction
void orc_srd_ghost_sums(void* h, unsigned step, unsigned cell, unsigned count, unsigned added, float* out3) {
    auto* o = static_cast<Oracle*>(h);
    for (unsigned i = 0; i < count; ++i) {
        oracle::V3 g{0, 0, 0};
        if (o->rng.mode == 1) {
            g = o->srd_ghost_sum(step, cell + i, added);
        } else {
            for (unsigned k = 0; k < added; ++k) g = oracle::add(g, oracle::mul(o->rng.ghost_vel(step, cell + i, k),
                                                                          o->p.thermal_velocity));
        }
        out3[3 * i] = g.x;
        out3[3 * i + 1] = g.y;
        out3[3 * i + 2] = g.z;
    }
}
// `count` wall-ghost positions of one cell (x, y, z in [0, 1] each, before the
// cell centre is added): mode 0 from the stream with the reference's rejection
// loop, mode 1 ghost k = i of the counter stream
void orc_ghost_pos_draws(void* h, unsigned step, unsigned cell, unsigned count, float shift, int layer, int sign,
                         float* out3) {
    auto* o = static_cast<Oracle*>(h);
    for (unsigned i = 0; i < count; ++i) {
        oracle::V3 const g = o->rng.ghost_pos(step, cell, i, shift, layer != 0, sign != 0);
        out3[3 * i] = g.x;
        out3[3 * i + 1] = g.y;
        out3[3 * i + 2] = g.z;
    }
}
// extended-rule direction draws of cells [cell0, cell0 + count): the index of
// the discrete direction (0..12) and the sign bit, then the gate uniform
void orc_ext_axis_draws(void* h, unsigned step, unsigned cell0, unsigned count, int* select, int* flip, float* gate) {
    auto* o = static_cast<Oracle*>(h);
    for (unsigned i = 0; i < count; ++i) {
        o->rng.ext_axis(step, cell0 + i, select + i, flip + i);
        gate[i] = o->rng.ext_gate(step, cell0 + i);
    }
}
void orc_ext_direction(int select, float* out) {
    oracle::V3 a = Oracle::ext_direction(select);
    out[0] = a.x; out[1] = a.y; out[2] = a.z;
}
void orc_sym_inverse(const float* m6, float tol, float* out6) {
    Oracle::Sym m;
    m.xx = m6[0]; m.yy = m6[1]; m.zz = m6[2]; m.xy = m6[3]; m.xz = m6[4]; m.yz = m6[5];
    Oracle::Sym o = Oracle::inverse(m, tol);
    out6[0] = o.xx; out6[1] = o.yy; out6[2] = o.zz; out6[3] = o.xy; out6[4] = o.xz; out6[5] = o.yz;
}

}  // extern "C"
