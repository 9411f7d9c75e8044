// MPCD time step for MI355X (gfx950, wave64): device kernels, host
// orchestration and the C-ABI declared in include/mpcd_hip.h.
//
// One step (reference order: shift -> stream -> collide, cuda_backend_impl.cu:111-131):
//
//   scan      offsets[c] = exclusive prefix of the per-cell counts of this
//             step (counts were accumulated by the previous collide launch);
//             counts are cleared for the next step in the same pass.
//   perm      cell-ordered permutation of the particle array: a block of
//             particles counts its keys in LDS windows and reserves one range
//             per distinct key with a single device atomic; perm[slot] = j.
//   draws     SRD only: the rotation axis of every cell and the ghost counts
//             and velocities of the wall cells, from the counter-based stream.
//   collide   one workgroup per tile of consecutive cells: gathers its cells'
//             particles through perm into LDS, orders each cell by source
//             index (stable counting-sort order, identical to the reference's
//             cell list, simulation_functions.cpp:84-94), re-applies the
//             streaming of this step, runs the SRD or extended collision, writes
//             the particles out in cell order (contiguous records), and already
//             streams them with the next step's shift to count the next step's
//             cells (fused histogram) and, when sampling, accumulates the mean
//             fields in fixed point (LDS bins per batch, then per-step sums
//             indexed by cell key; the host reorders them to reference order
//             where they leave the device).
//
// Particle state is one FP32 record per particle (x, y, z, vx, vy, vz: 24 B),
// double-buffered; the collide launch reads one buffer and writes the other.
// The gather of a particle then touches one or two cache lines instead of
// six, and the cell-ordered write-out stays contiguous.  All floating-point code is compiled without
// contraction so that per-particle arithmetic matches the CPU restatement
// bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#pragma clang fp contract(off)

#include "mpcd_hip.h"
#include "rng.h"

namespace mpcd_hip {

constexpr int WG = 256;          // threads per workgroup (4 wave64)
constexpr double FIX_SCALE = 1099511627776.0;   // 2^40, fixed-point velocity sums

// Parameters of one system.  A handle holds a table of them in device memory
// (one entry per ensemble member; a plain simulation is a table of one), and
// every kernel reads the entry of the system its block works on.  All systems
// share flattened arrays: system s owns particle slots [pb, pb + N), cell keys
// [cb, cb + C) and wall cells [wb, wb + 2 X Y).
struct KParams {
    float L[3];          // volume edges (wall layers included)
    float half[3];       // L / 2
    float L15[3], L2[3], L25[3];  // 1.5 L, 2 L, 2.5 L (exact: L is a whole number up to 2^21)
    uint32_t L3bits[3];  // bit pattern of 3 L
    uint32_t dim[3];     // X, Y, Z
    uint32_t ord[3];     // axis at internal position 0 (fastest), 1, 2
    uint32_t D0, D1, D2; // extents of the internal axes
    uint32_t Bsh;        // rows of the slowest axis per block = 1 << Bsh
    uint32_t nblk, Blast, BLK;  // blocks, rows of the last block, cells per full block
    uint32_t mBLK[2], mD0[2], mBlast[2];  // division magics (multiplier, shift)
    uint32_t C;          // cells of this system
    int perio_z;
    int channel;         // channel geometry: cylinder about the x axis (simulation_functions.cpp:11-24, 75-82)
    float r2_init, r2_chan;  // its squared radius at initialisation (from X) and in streaming (from Z)
    float dt, drag_dt, tv, sin_a, cos_a;
    uint32_t n, cap;     // particles per cell, cell table capacity 4n
    uint32_t mn[2];      // division magic for n (n > 1)
    Stream rs;
    uint32_t tile_cells; // cells per collide tile
    uint32_t lds_cap;    // LDS particle capacity of this system's collide tiles (host-side choice)
    uint32_t tile_req;   // requested cells per tile (0: about one batch)
    uint32_t gid;        // launch set: systems of one capacity and set share a collide launch
    uint32_t N;          // particles of this system
    uint32_t cb, pb, wb; // first cell key, particle slot, wall cell in the shared arrays
    uint64_t gvb;        // first ghost float of this system (3 per slot SRD, 6 extended)
    // step state: written by the host at set-up, advanced on the device by k_advance
    uint32_t step;
    float shift[3], shift_next[3];
    // extended rule, a step with injected draws (mpcd_hip_step_injected_extended;
    // null otherwise): per reference cell the direction index, sign and gate
    // uniform (bits), and the members' thermal Gaussians, cap per cell
    uint32_t const* inj_ext;
    float const* inj_gauss;
};

// The system table and the block -> system maps of one handle.
struct Sys {
    KParams* P;                  // [ns]
    uint32_t const* blk_part;    // [ns + 1] first block of each system in per-particle kernels
    uint32_t const* blk_cell;    // [ns + 1] ... in per-cell kernels
    uint32_t const* blk_wall;    // [ns + 1] ... in per-wall-cell kernels
    uint16_t const* tile_sys;    // [tiles] system of each collide tile (ns > 1)
    uint32_t const* tile0;       // [ns + 1] first collide tile of each system
    uint32_t ns;
    uint32_t CT;                 // cells of all systems (stride of the axis planes)
    uint64_t NT;                 // particles of all systems (stride of the state planes)
};

// Cells larger than the LDS tile are listed after the status word (status[1]
// count, then (system, local key) pairs) for the big-cell launch, which runs
// BIG_BLOCKS workgroups over the list.  The list holds one entry per
// N / (smallest tile capacity + 1) + 1, so it cannot overflow.
constexpr uint32_t BIG_BLOCKS = 256;
constexpr uint64_t big_list_words(uint64_t n) { return 2 + 2 * (n / 257 + 2); }

// the 6 floats of particle i: x y z vx vy vz
__host__ __device__ __forceinline__ uint64_t rec(uint64_t i) { return 6 * i; }

__device__ __forceinline__ void load_rec(const float* __restrict__ st, uint64_t i, float* v) {
    float2 const* const q = reinterpret_cast<float2 const*>(st + rec(i));
    float2 const a = q[0], b = q[1], c = q[2];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
__device__ __forceinline__ void store_rec(float* __restrict__ st, uint64_t i, float px, float py, float pz, float vx,
                                          float vy, float vz) {
    float2* const q = reinterpret_cast<float2*>(st + rec(i));
    q[0] = make_float2(px, py);
    q[1] = make_float2(pz, vx);
    q[2] = make_float2(vy, vz);
}

// system whose block range holds block b (base is non-decreasing, base[0] = 0)
__device__ __forceinline__ uint32_t find_sys(uint32_t const* base, uint32_t ns, uint32_t b) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        uint32_t const mid = (lo + hi) >> 1;
        if (base[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// the 13 directions of the extended rule (simulation_functions.cpp:272-290),
// evaluated on the host with the reference's float expressions
__constant__ float c_axes[13][3];

struct SampleBufs {
    uint32_t* cnt;                  // per reference cell
    unsigned long long* sum;        // 3 per reference cell, fixed point 2^40, two's complement
};

// Adjoint runs (SRD, one system): the body-force field and what one step
// records for the reverse pass.  Null pointers record nothing / apply no force.
struct Tape {
    float* vp;             // [3][N] streamed velocity of every output slot, before the collision
    float* u;              // [3][C] mean velocity of every cell (particles + ghosts; 0 if not collided)
    uint32_t* tot;         // [C] particles in the table + ghosts of every cell
    float const* force;    // [C] x body force per reference cell, applied as v_x += g dt + f dt
};
constexpr uint32_t BOUNCED = 0x80000000u;  // flag in a recorded permutation entry: the particle bounced

// ------------------------------------------------------------ helpers ---

// fmodf for the arguments the fast path below does not cover (never on a
// normal trajectory): one out-of-line copy instead of one per call site
__device__ __noinline__ float wrap_slow(float x, float L) { return fmodf(x, L); }

// periodic wrap fmodf(r + 1.5 L, L) - L/2 (simulation_functions.cpp:46-51) on
// axis a.  For x = r + 1.5 L in [0, 3L), fmodf(x, L) = x - k L with k = 0, 1
// or 2, and that subtraction is exact (Sterbenz); k L + L/2 is exact as well,
// so one subtraction x - (k L + L/2) rounds exactly as fmodf(x, L) - L/2.
// The three candidates are formed and the right one selected.  x outside
// [0, 3L) (never on a normal trajectory) takes fmodf itself; as unsigned
// integers the bits of a non-negative float order like its value, and those
// of negative numbers and NaN lie above those of 3L.
__device__ __forceinline__ float wrap1(float r, KParams const& P, int a) {
    float const x = r + P.L15[a];
    float y0 = x - P.half[a], y1 = x - P.L15[a], y2 = x - P.L25[a];
    // keeps the three subtractions (the compiler would otherwise select the
    // constant first and need register copies of the scalar operands)
    asm("" : "+v"(y0), "+v"(y1), "+v"(y2));
    float out = x >= P.L2[a] ? y2 : (x >= P.L[a] ? y1 : y0);
    if (__builtin_expect(__float_as_uint(x) >= P.L3bits[a], 0)) out = wrap_slow(x, P.L[a]) - P.half[a];
    return out;
}

// u % D with the common cases as a compare and a subtraction
__device__ __forceinline__ uint32_t mod_dim(uint32_t u, uint32_t D) {
    if (u < D) return u;
    if (u - D < D) return u - D;
    return u % D;
}

// streaming with z-wall bounce-back (simulation_functions.cpp:56-74), then
// shift and periodic wrap (:84); returns whether the particle bounced
__device__ __forceinline__ bool stream_shift(KParams const& P, float const* sh, float& px, float& py, float& pz,
                                             float& vx, float& vy, float& vz) {
    float const dt = P.dt;
    bool bounced = false;
    if (P.channel) {
        // undo a move that leaves the cylinder and reverse the velocity (:75-82)
        px = px + vx * dt;
        py = py + vy * dt;
        pz = pz + vz * dt;
        if ((pz * pz + py * py) > P.r2_chan) {
            px = px - vx * dt;
            py = py - vy * dt;
            pz = pz - vz * dt;
            vx = -vx;
            vy = -vy;
            vz = -vz;
            bounced = true;
        }
    } else if (!P.perio_z) {
        float const z_wall = (0.5f * P.L[2] - 1.0f);
        float const next_z = pz + vz * dt;
        if (fabsf(next_z) > z_wall) {
            float const t = (next_z > 0 ? z_wall - pz : -z_wall - pz) / vz;
            px = px + vx * t;
            py = py + vy * t;
            pz = pz + vz * t;
            vx = -vx;
            vy = -vy;
            vz = -vz;
            bounced = true;
            float const rem = dt - t;
            px = px + vx * rem;
            py = py + vy * rem;
            pz = pz + vz * rem;
        } else {
            px = px + vx * dt;
            py = py + vy * dt;
            pz = pz + vz * dt;
        }
    } else {
        px = px + vx * dt;
        py = py + vy * dt;
        pz = pz + vz * dt;
    }
    px = wrap1(px + sh[0], P, 0);
    py = wrap1(py + sh[1], P, 1);
    pz = wrap1(pz + sh[2], P, 2);
    return bounced;
}

// cell coordinates (volume_container.hpp:69-75)
__device__ __forceinline__ void cell_of(KParams const& P, float px, float py, float pz, uint32_t& ix, uint32_t& iy,
                                        uint32_t& iz) {
    ix = mod_dim(static_cast<uint32_t>(floorf(px + P.half[0])), P.dim[0]);
    iy = mod_dim(static_cast<uint32_t>(floorf(py + P.half[1])), P.dim[1]);
    iz = mod_dim(static_cast<uint32_t>(floorf(pz + P.half[2])), P.dim[2]);
}

// Internal cell key: axes ordered by extent (i0 fastest); the slowest axis is
// cut into blocks of 2^Bsh rows, and inside a block the order is (i1, row, i0),
// so a cell's neighbours along all three axes lie within one block.
__host__ __device__ __forceinline__ uint32_t encode_key(KParams const& P, uint32_t i0, uint32_t i1, uint32_t i2) {
    uint32_t const b = i2 >> P.Bsh;
    uint32_t const r2 = i2 - (b << P.Bsh);
    uint32_t const Bb = b + 1 == P.nblk ? P.Blast : (1u << P.Bsh);
    return b * P.BLK + (i1 * Bb + r2) * P.D0 + i0;
}
// a / d for 32-bit a by a precomputed (multiplier, shift) pair
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994: m = floor(2^32 (2^l - d) / d) + 1, l = ceil(log2 d)).
__host__ __device__ __forceinline__ uint32_t udiv(uint32_t a, uint32_t const* mg) {
    uint32_t const t = static_cast<uint32_t>((static_cast<uint64_t>(a) * mg[0]) >> 32);
    return (t + ((a - t) >> 1)) >> mg[1];
}
inline void udiv_magic(uint32_t d, uint32_t* mg) {
    if (d <= 1) {  // not representable; callers divide by 1 without udiv
        mg[0] = 0;
        mg[1] = 0;
        return;
    }
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    mg[0] = static_cast<uint32_t>(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    mg[1] = l - 1;
}

__host__ __device__ __forceinline__ void decode_key3(KParams const& P, uint32_t key, uint32_t& i0, uint32_t& i1,
                                                     uint32_t& i2) {
    uint32_t const b = P.mBLK[0] || P.mBLK[1] ? udiv(key, P.mBLK) : key / P.BLK;
    uint32_t const rem = key - b * P.BLK;
    bool const last = b + 1 == P.nblk;
    uint32_t const Bb = last ? P.Blast : (1u << P.Bsh);
    uint32_t const t = P.D0 > 1 ? udiv(rem, P.mD0) : rem;
    i0 = rem - t * P.D0;
    i1 = last ? (P.Blast > 1 ? udiv(t, P.mBlast) : t) : (t >> P.Bsh);
    i2 = (b << P.Bsh) + (t - i1 * Bb);
}
__host__ __device__ __forceinline__ uint32_t key_of(KParams const& P, uint32_t ix, uint32_t iy, uint32_t iz) {
    uint32_t const i[3] = {ix, iy, iz};
    return encode_key(P, i[P.ord[0]], i[P.ord[1]], i[P.ord[2]]);
}
__host__ __device__ __forceinline__ uint32_t ref_of(KParams const& P, uint32_t ix, uint32_t iy, uint32_t iz) {
    return ix + P.dim[0] * (iy + P.dim[1] * iz);
}
__host__ __device__ __forceinline__ void decode_key(KParams const& P, uint32_t key, uint32_t& ix, uint32_t& iy,
                                                    uint32_t& iz) {
    uint32_t q0, q1, q2;
    decode_key3(P, key, q0, q1, q2);
    ix = P.ord[0] == 0 ? q0 : (P.ord[1] == 0 ? q1 : q2);
    iy = P.ord[0] == 1 ? q0 : (P.ord[1] == 1 ? q1 : q2);
    iz = P.ord[0] == 2 ? q0 : (P.ord[1] == 2 ? q1 : q2);
}

// stride of the middle internal axis at local key `key` (cells of one block
// of the slowest axis), the key distance of a neighbour along that axis
__device__ __forceinline__ uint32_t mid_stride(KParams const& P, uint32_t key) {
    uint32_t const b = P.mBLK[0] || P.mBLK[1] ? udiv(key, P.mBLK) : key / P.BLK;
    return (b + 1 == P.nblk ? P.Blast : (1u << P.Bsh)) * P.D0;
}

// Groups the lanes of a wave by key without touching memory: every lane
// learns the leader lane of its group and its position inside the group, the
// leader learns the group size.  Must be called by whole waves in uniform
// control flow; invalid lanes join no group.
struct WaveGroup {
    int leader;
    uint32_t pos;
    uint32_t size;  // valid on the leader
};
__device__ __forceinline__ WaveGroup wave_group(uint32_t key, bool valid) {
    uint32_t const lane = __lane_id();
    uint64_t const below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    WaveGroup g = {static_cast<int>(lane), 0u, 0u};
    bool active = valid;
    for (;;) {
        uint64_t const m = __ballot(active);
        if (m == 0) break;
        int const leader = __builtin_ctzll(m);
        uint32_t const lk = __builtin_amdgcn_readlane(key, leader);
        bool const mine = active && key == lk;
        uint64_t const grp = __ballot(mine);
        if (mine) {
            g.leader = leader;
            g.pos = static_cast<uint32_t>(__popcll(grp & below));
            active = false;
        }
        if (static_cast<int>(lane) == leader) g.size = static_cast<uint32_t>(__popcll(grp));
    }
    return g;
}

// histogram update: one non-returning atomic per distinct key of the wave
__device__ __forceinline__ void wave_count(uint32_t* count, uint32_t key, bool valid) {
    WaveGroup const g = wave_group(key, valid);
    if (valid && g.leader == static_cast<int>(__lane_id())) atomicAdd(count + key, g.size);
}

// Per-cell stages give each cell a group of G = 2^lg aligned lanes (G a
// function of the particles per cell n only, so results do not depend on the
// tiling): member i of the cell is added by lane i mod G, each lane in member
// order, and group_sum combines the G partial sums pairwise,
// ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)).  Each butterfly stage is
// one DPP move inside a quad (lanes i ^ 1, then i ^ 2) or a half row (lane
// 7 - i, which after the quad stages holds the other quad's sum); both lanes
// of a pair add the same two values, so every lane ends with the same bits.
// All G lanes of a group must be active.
__host__ __device__ constexpr uint32_t cell_lanes_log2(uint32_t n) {
    return n < 6 ? 0u : (n < 12 ? 1u : (n < 24 ? 2u : 3u));
}
constexpr int DPP_QUAD_SWAP1 = 0xB1;    // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_SWAP2 = 0x4E;    // quad_perm [2, 3, 0, 1]
constexpr int DPP_HALF_MIRROR = 0x141;  // row_half_mirror: lane 7 - i of each 8
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ float group_sum(float v, uint32_t lg) {
    if (lg >= 1) v = v + dpp_f<DPP_QUAD_SWAP1>(v);
    if (lg >= 2) v = v + dpp_f<DPP_QUAD_SWAP2>(v);
    if (lg >= 3) v = v + dpp_f<DPP_HALF_MIRROR>(v);
    return v;
}
__device__ __forceinline__ int group_sum(int v, uint32_t lg) {
    if (lg >= 1) v = v + dpp_i<DPP_QUAD_SWAP1>(v);
    if (lg >= 2) v = v + dpp_i<DPP_QUAD_SWAP2>(v);
    if (lg >= 3) v = v + dpp_i<DPP_HALF_MIRROR>(v);
    return v;
}

__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t G) {
    // workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous run of tiles so neighbouring tiles share its L2
    uint32_t const q = G / 8, r = G % 8;
    uint32_t const x = b % 8, i = b / 8;
    return x * q + (x < r ? x : r) + i;
}

// ------------------------------------------------------------- kernels ---

// seeded initial distribution (simulation_functions.cpp:7-33): positions
// uniform in the fluid region (edge = volume - 2 (1 - periodic), :8-10),
// Maxwell-Boltzmann velocities
constexpr uint32_t INIT_ATTEMPTS = 1u << 16;   // channel rejection loop bound

__global__ __launch_bounds__(WG) void k_init(Sys sys, float* __restrict__ st, uint32_t* __restrict__ status) {
    uint32_t const s = find_sys(sys.blk_part, sys.ns, blockIdx.x);
    KParams const P = sys.P[s];
    uint32_t const b0 = sys.blk_part[s], nb = sys.blk_part[s + 1] - b0;
    float const sx = P.L[0], sy = P.L[1], sz = P.perio_z ? P.L[2] : P.L[2] - 2.0f;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x - b0) * WG + threadIdx.x; i < P.N;
         i += static_cast<uint64_t>(nb) * WG) {
        uint32_t const lo = static_cast<uint32_t>(i);
        // channel: redraw (attempt a keys the draw) until inside the cylinder (:11-24)
        float px, py, pz;
        for (uint32_t att = 0;; ++att) {
            Words const a = P.rs.at(P_INIT_POS, lo, 0u, att);
            px = (uniform(a.w[0]) - 0.5f) * sx;
            py = (uniform(a.w[1]) - 0.5f) * sy;
            pz = (uniform(a.w[2]) - 0.5f) * sz;
            if (!(P.channel && (pz * pz + py * py) > P.r2_init)) break;
            if (att + 1 == INIT_ATTEMPTS) {
                atomicOr(status, 2u);
                break;
            }
        }
        Words const b = P.rs.at(P_INIT_VEL, lo, 0u, 0);
        float g0, g1, g2, g3;
        box_muller(b.w[0], b.w[1], &g0, &g1);
        box_muller(b.w[2], b.w[3], &g2, &g3);
        uint64_t const j = P.pb + i;
        store_rec(st, j, px, py, pz, g0 * P.tv, g1 * P.tv, g2 * P.tv);
    }
}

// streaming half-step of an unsorted state: cell key per particle and the histogram
__global__ __launch_bounds__(WG) void k_keys(Sys sys, const float* __restrict__ st, uint32_t* __restrict__ count,
                                             uint32_t* __restrict__ keys) {
    uint32_t const s = find_sys(sys.blk_part, sys.ns, blockIdx.x);
    KParams const P = sys.P[s];
    uint32_t const b0 = sys.blk_part[s], nb = sys.blk_part[s + 1] - b0;
    // the loop bound is uniform: wave_count needs whole waves
    for (uint64_t i0 = static_cast<uint64_t>(blockIdx.x - b0) * WG; i0 < P.N; i0 += static_cast<uint64_t>(nb) * WG) {
        uint64_t const i = i0 + threadIdx.x;
        bool const valid = i < P.N;
        uint64_t const j = P.pb + i;
        uint32_t key = 0;
        if (valid) {
            float v[6];
            load_rec(st, j, v);
            float px = v[0], py = v[1], pz = v[2], vx = v[3], vy = v[4], vz = v[5];
            stream_shift(P, P.shift, px, py, pz, vx, vy, vz);
            uint32_t ix, iy, iz;
            cell_of(P, px, py, pz, ix, iy, iz);
            key = P.cb + key_of(P, ix, iy, iz);
        }
        wave_count(count, key, valid);
        if (valid) keys[j] = key;
    }
}

// ---- exclusive scan of the per-cell counts (3 passes) ----
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = WG * SCAN_ITEMS;

template <int NT>
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_w, uint32_t* total) {
    uint32_t const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t const y = __shfl_up(x, d, 64);
        if (lane >= static_cast<uint32_t>(d)) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) {
        uint32_t const t = s_w[k];
        if (static_cast<uint32_t>(k) < w) pre += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

__global__ __launch_bounds__(WG) void k_scan_reduce(const uint32_t* __restrict__ count, uint32_t C,
                                                    uint32_t* __restrict__ partial) {
    __shared__ uint32_t s_w[WG / 64];
    uint64_t const base = static_cast<uint64_t>(blockIdx.x) * SCAN_CHUNK;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint64_t const idx = base + j * WG + threadIdx.x;
        if (idx < C) s += count[idx];
    }
    uint32_t tot;
    block_exclusive<WG>(s, s_w, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_partials(uint32_t* __restrict__ partial, uint32_t nb) {
    __shared__ uint32_t s_w[1024 / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 1024) {
        uint32_t const i = base + threadIdx.x;
        uint32_t const v = i < nb ? partial[i] : 0u;
        uint32_t tot;
        uint32_t const ex = block_exclusive<1024>(v, s_w, &tot);
        if (i < nb) partial[i] = carry + ex;
        carry += tot;
    }
}

// fold: `partial` holds the blocks' totals and every block adds up those
// before it (grids of at most SCAN_FOLD_MAX blocks, which then skip
// k_scan_partials and its launch); else it holds their exclusive prefix
constexpr uint32_t SCAN_FOLD_MAX = 1024;
__global__ __launch_bounds__(WG) void k_scan_down(uint32_t* __restrict__ count, uint32_t C,
                                                  const uint32_t* __restrict__ partial, uint32_t* __restrict__ offset,
                                                  uint32_t* __restrict__ cursor, uint64_t N, int fold) {
    __shared__ uint32_t s_v[SCAN_CHUNK + SCAN_CHUNK / 32];
    __shared__ uint32_t s_w[WG / 64];
    __shared__ uint32_t s_f[WG / 64];
    uint64_t const base = static_cast<uint64_t>(blockIdx.x) * SCAN_CHUNK;
    auto pad = [](uint32_t i) { return i + (i >> 5); };
    uint32_t before = 0;
    if (fold) {
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += WG) before += partial[i];
        for (int dd = 32; dd >= 1; dd >>= 1) before += static_cast<uint32_t>(__shfl_xor(before, dd, 64));
        if ((threadIdx.x & 63u) == 0) s_f[threadIdx.x >> 6] = before;
    }
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint32_t const li = j * WG + threadIdx.x;
        uint64_t const idx = base + li;
        s_v[pad(li)] = idx < C ? count[idx] : 0u;
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) run += s_v[pad(threadIdx.x * SCAN_ITEMS + j)];
    uint32_t tot;
    uint32_t pre = block_exclusive<WG>(run, s_w, &tot);
    if (fold) {
        // s_f was written before the block_exclusive barriers
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < WG / 64; ++k) t += s_f[k];
        pre += t;
    } else {
        pre += partial[blockIdx.x];
    }
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint32_t const li = pad(threadIdx.x * SCAN_ITEMS + j);
        uint32_t const x = s_v[li];
        s_v[li] = pre;
        pre += x;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint32_t const li = j * WG + threadIdx.x;
        uint64_t const idx = base + li;
        if (idx < C) {
            uint32_t const o = s_v[pad(li)];
            offset[idx] = o;
            cursor[idx] = o;
            count[idx] = 0u;
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) offset[C] = static_cast<uint32_t>(N);
}

// cell-ordered permutation: each particle takes a free slot of its cell.
// A block takes PERM_ITEMS * WG consecutive particles.  They sat in a short run
// of cells and move at most about one cell, so their keys fall into three
// windows: around the run's first key and one middle-axis stride either side.
// The block counts them there in LDS (returning LDS atomics give each particle
// its place inside the block's share of the key), reserves one range per
// distinct key with a single device atomic (non-empty bins listed first, so the
// atomics run with full waves) and writes the slots.  Keys outside the windows
// take a device atomic of their own.  The order inside a cell is not the final
// one: the collide kernel orders every cell by source slot.
constexpr int PERM_ITEMS = 8;
constexpr uint32_t PERM_PW = 512;             // bins per window
constexpr uint32_t PERM_WIN = 3 * PERM_PW;

// one block of the permutation: particles [blk ITEMS, (blk + 1) ITEMS)
__device__ __forceinline__ void perm_block(Sys const& sys, const uint32_t* __restrict__ keys,
                                           uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm, uint64_t N,
                                           uint32_t blk) {
    constexpr uint32_t ITEMS = WG * PERM_ITEMS;
    // 17.4 KB in all, so eight blocks (the 32 waves a CU holds) fit the LDS
    static_assert(ITEMS <= 65536 && PERM_WIN < 65536, "prefixes, bins and entries in 16 bits");
    __shared__ uint32_t cnt[PERM_WIN];   // count per bin, then the bin's reserved device base
    __shared__ uint16_t pre[PERM_WIN];   // block-local exclusive prefix per bin
    __shared__ uint32_t sj[ITEMS];       // first the non-empty bin list, then the entries in bin
                                         // order: bin << 16 | particle - base
    __shared__ uint32_t s_w[WG / 64];
    uint64_t const base = static_cast<uint64_t>(blk) * ITEMS;
    for (uint32_t i = threadIdx.x; i < PERM_WIN; i += WG) cnt[i] = 0u;
    uint32_t key[PERM_ITEMS], local[PERM_ITEMS], bin[PERM_ITEMS];
#pragma unroll
    // a lane takes PERM_ITEMS consecutive particles: neighbouring lanes then
    // mostly hold particles of different cells, so the histogram's LDS
    // atomics below meet fewer lanes on one address (lane-interleaved items
    // measured 0.15 ms slower at 1000x1000x22 cells, n = 20)
    for (int q = 0; q < PERM_ITEMS; ++q) {
        uint64_t const j = base + threadIdx.x * PERM_ITEMS + q;
        key[q] = j < N ? keys[j] : ~0u;
    }
    // window bases from the block's first key, computed by every lane (the
    // values are uniform, so the key and the system's parameters are read
    // alongside the keys above, not after them)
    uint32_t const k0 = keys[base];   // base < N for every launched block
    uint32_t lo = 0, hi = sys.ns;      // system holding key k0
    while (hi - lo > 1) {
        uint32_t const mid = (lo + hi) >> 1;
        if (sys.P[mid].cb <= k0) lo = mid; else hi = mid;
    }
    KParams const& P = sys.P[lo];
    uint32_t const smid = mid_stride(P, k0 - P.cb);
    // the window starts one fastest-axis row (plus one) before the first key;
    // bases may wrap below zero (distances are taken modulo 2^32)
    uint32_t const w0 = k0 - min(P.D0 + 2u, PERM_PW / 2), wm = w0 - smid, wp = w0 + smid;
    __syncthreads();   // the bins are cleared
#pragma unroll
    for (int q = 0; q < PERM_ITEMS; ++q) {
        uint32_t w = key[q] - w0;
        uint32_t b = ~0u;
        if (w < PERM_PW) b = w;
        else if ((w = key[q] - wm) < PERM_PW) b = PERM_PW + w;
        else if ((w = key[q] - wp) < PERM_PW) b = 2 * PERM_PW + w;
        if (key[q] == ~0u) b = ~0u;
        bin[q] = b;
        local[q] = b != ~0u ? atomicAdd(&cnt[b], 1u) : ~0u;
    }
    __syncthreads();
    // thread t owns bins [t B, t B + B): one scan gives both the block-local
    // prefix of the counts (high half) and the places of the non-empty bins in
    // the bin list (low half); both totals stay below 2^16
    constexpr int PB = PERM_WIN / WG;
    uint32_t const b0 = threadIdx.x * PB;
    uint32_t packed = 0;
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        uint32_t const c = cnt[b0 + k];
        packed += (c << 16) + (c != 0u ? 1u : 0u);
    }
    uint32_t tot_packed;
    uint32_t const ex = block_exclusive<WG>(packed, s_w, &tot_packed);
    uint32_t p = ex >> 16, at = ex & 0xffffu;
    uint32_t const nz_tot = tot_packed & 0xffffu, in_win = tot_packed >> 16;
    uint32_t* const list = sj;
#pragma unroll
    for (int k = 0; k < PB; ++k) {
        uint32_t const c = cnt[b0 + k];
        pre[b0 + k] = static_cast<uint16_t>(p);
        p += c;
        if (c) list[at++] = b0 + k;
    }
    __syncthreads();
    // one device range per non-empty bin, with full waves
    for (uint32_t t = threadIdx.x; t < nz_tot; t += WG) {
        uint32_t const i = list[t];
        uint32_t const wdw = i / PERM_PW;
        uint32_t const k = (wdw == 0 ? w0 : (wdw == 1 ? wm : wp)) + (i - wdw * PERM_PW);
        cnt[i] = atomicAdd(cursor + k, cnt[i]);
    }
    __syncthreads();
    // entries in bin order; keys outside the windows take their own slot
#pragma unroll
    for (int q = 0; q < PERM_ITEMS; ++q) {
        uint64_t const j = base + threadIdx.x * PERM_ITEMS + q;
        if (key[q] == ~0u) continue;
        if (bin[q] != ~0u) {
            uint32_t const o = pre[bin[q]] + local[q];
            sj[o] = (bin[q] << 16) | static_cast<uint32_t>(threadIdx.x * PERM_ITEMS + q);
        } else {
            perm[atomicAdd(cursor + key[q], 1u)] = static_cast<uint32_t>(j);
        }
    }
    __syncthreads();
    // neighbouring lanes write neighbouring slots of a cell
    for (uint32_t o = threadIdx.x; o < in_win; o += WG) {
        uint32_t const e = sj[o], b = e >> 16;
        perm[cnt[b] + (o - pre[b])] = static_cast<uint32_t>(base) + (e & 0xffffu);
    }
}



// per-cell ghost count in a wall layer (simulation_functions.cpp:137-146)
__device__ __forceinline__ uint32_t ghost_count(KParams const& P, uint32_t ref, float sh, bool layer, bool sign) {
    uint32_t added = 0;
    for (uint32_t i = 0; i < P.n; i += 4) {
        Words const r = P.rs.at(P_GHOST_COUNT, i >> 2, ref, P.step);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (i + k < P.n && (((uniform(r.w[k]) > sh) ^ layer) ^ sign)) ++added;
    }
    return added;
}

// Random unit vector, uniform on the sphere, from one counter draw: z uniform
// on [-1, 1] and the azimuth uniform on [0, 2 pi) (Archimedes: the projection
// of the uniform sphere measure on an axis is uniform).  The reference draws
// the same distribution by Marsaglia's rejection (random.hpp:237-253); this
// form needs no loop, so the lanes of a wave never wait on a redraw.
__device__ __forceinline__ void srd_axis(KParams const& P, uint32_t ref, float& ax, float& ay, float& az) {
    Words const r = P.rs.at(P_AXIS, 0u, ref, P.step);
    float const z = 1.0f - 2.0f * uniform(r.w[0]);
    float sn, cs;
    sincospi(2.0f * uniform(r.w[1]), &sn, &cs);
    float const rho = sqrtf((1.0f - z) * (1.0f + z));
    ax = rho * cs;
    ay = rho * sn;
    az = z;
}

struct Sym {
    float xx, yy, zz, xy, xz, yz;
};

// mechanic.hpp:26, 52-79
__device__ __forceinline__ Sym sym_inverse(Sym const& m, float tol) {
    float det = m.xx * (m.yy * m.zz - m.yz * m.yz) + m.xz * (2.0f * m.xy * m.yz - m.yy * m.xz) - (m.xy * m.xy * m.zz);
    if (det < 0) det = 0;
    Sym o = {0, 0, 0, 0, 0, 0};
    if (det > tol) {
        float const id = 1.0f / det;
        o.xx = (m.yy * m.zz - m.yz * m.yz) * id;
        o.yy = (m.xx * m.zz - m.xz * m.xz) * id;
        o.zz = (m.xx * m.yy - m.xy * m.xy) * id;
        o.xy = (m.xz * m.yz - m.zz * m.xy) * id;
        o.xz = (m.xy * m.yz - m.yy * m.xz) * id;
        o.yz = (m.xy * m.xz - m.xx * m.yz) * id;
        return o;
    }
    float const tr = m.xx + m.yy + m.zz;
    if (tr > tol) {
        float const s = 4.0f / (tr * tr);
        o.xx = m.xx * s;
        o.yy = m.yy * s;
        o.zz = m.zz * s;
        o.xy = m.xy * s;
        o.xz = m.xz * s;
        o.yz = m.yz * s;
    }
    return o;
}

// Per-cell draws computed by fully populated waves instead of inside the
// particle tiles:
//   gcnt[w]        ghost count of wall cell w (simulation_functions.cpp:141-146)
//   gvel           SRD: gvel[w][a], the sum of the cell's ghost velocities
//                  (one draw, below); extended rule: gvel[w][g][6], each
//                  ghost's position and velocity (:147-167)
//   axis[a][key]   SRD rotation axis of every cell (random.hpp:237-253), only
//                  for steps that keep their axes in memory (taped adjoint
//                  steps); other steps draw the axis in the collide tile
// Wall cells: the two z layers next to the walls for this step's shift,
// w = (upper ? X*Y : 0) + x + X*y.
__device__ __forceinline__ void axes_block(Sys const& sys, float* __restrict__ axis, uint32_t blk) {
    uint32_t const s = find_sys(sys.blk_cell, sys.ns, blk);
    KParams const P = sys.P[s];
    uint32_t const b0 = sys.blk_cell[s], nb = sys.blk_cell[s + 1] - b0;
    uint64_t const CT = sys.CT;
    for (uint32_t key = (blk - b0) * WG + threadIdx.x; key < P.C; key += nb * WG) {
        uint32_t ix, iy, iz;
        decode_key(P, key, ix, iy, iz);
        float ax, ay, az;
        srd_axis(P, ref_of(P, ix, iy, iz), ax, ay, az);
        uint64_t const g = P.cb + key;
        axis[g] = ax;
        axis[CT + g] = ay;
        axis[2 * CT + g] = az;
    }
}

// A block takes WG wall cells at a time: one lane per cell draws the ghost
// count.  SRD: the ghosts enter the collision only through the sum of their
// velocities (the cell mean, simulation_functions.cpp:168-171; getCorrection
// returns the mean, mpc_cell.hpp:82-84), and the sum of `added` independent
// thermal velocities is one Gaussian of `added` times their variance, so the
// same lane draws that sum at once: sqrt(added) sqrt(T) (g0, g1, g2), 3 floats
// per wall cell.  Extended rule (ext): every ghost's position and velocity
// enter the collision; the cells' n ghost slots are spread over all lanes
// (slot g of cell k on lane k n + g), so the draws run on full waves and
// neighbouring lanes store neighbouring ghosts, 6 floats per slot (position
// in the cell, then velocity), the positions' z in the part of the cell the
// reference's rejection loop accepts (simulation_functions.cpp:147-167).
__device__ __forceinline__ void ghosts_block(Sys const& sys, uint32_t* __restrict__ gcnt, float* __restrict__ gvel,
                                             uint32_t blk, bool ext) {
    __shared__ uint32_t s_cnt[WG];
    uint32_t const s = find_sys(sys.blk_wall, sys.ns, blk);
    KParams const P = sys.P[s];
    uint32_t const b0 = sys.blk_wall[s], nb = sys.blk_wall[s + 1] - b0;
    uint32_t const XY = P.dim[0] * P.dim[1];
    float const sh = fabsf(P.shift[2]);
    bool const sign = P.shift[2] > 0;
    uint32_t const iz_layer = static_cast<uint32_t>(P.L[2] - (sign ? 1.0f : 2.0f));
    // reference index of wall cell w: w < XY the layer at z = sign, else the upper one
    auto ref_of_wall = [&](uint32_t w) {
        bool const layer = w >= XY;
        return (layer ? w - XY : w) + XY * (layer ? iz_layer : static_cast<uint32_t>(sign));
    };
    // the chunk start is the same on every lane, so the barriers are uniform
    for (uint32_t c0 = (blk - b0) * WG; c0 < 2 * XY; c0 += nb * WG) {
        uint32_t const w = c0 + threadIdx.x;
        uint32_t added = 0;
        if (w < 2 * XY) {
            added = ghost_count(P, ref_of_wall(w), sh, w >= XY, sign);
            gcnt[P.wb + w] = added;
            if (!ext && added) {
                Words const v = P.rs.at(P_GHOST_VEL, 0u, ref_of_wall(w), P.step);
                float g0, g1, g2, g3;
                box_muller(v.w[0], v.w[1], &g0, &g1);
                box_muller(v.w[2], v.w[3], &g2, &g3);
                float const sc = P.tv * sqrtf(static_cast<float>(added));
                float* const out = gvel + P.gvb + 3ull * w;
                out[0] = g0 * sc;
                out[1] = g1 * sc;
                out[2] = g2 * sc;
            }
        }
        if (!ext) continue;   // uniform over the block
        s_cnt[threadIdx.x] = added;
        __syncthreads();
        uint32_t const ncell = min(static_cast<uint32_t>(WG), 2 * XY - c0);
        constexpr uint32_t per = 6u;
        for (uint32_t t = threadIdx.x; t < ncell * P.n; t += WG) {
            uint32_t const k = t / P.n, g = t - k * P.n;
            if (g < s_cnt[k]) {
                uint32_t const w = c0 + k, ref = ref_of_wall(w);
                Words const v = P.rs.at(P_GHOST_VEL, g, ref, P.step);
                float g0, g1, g2, g3;
                box_muller(v.w[0], v.w[1], &g0, &g1);
                box_muller(v.w[2], v.w[3], &g2, &g3);
                float* out = gvel + P.gvb + static_cast<uint64_t>(per) * P.n * c0 + static_cast<uint64_t>(per) * t;
                {
                    // ghost position: uniform in x and y, z in the accepted
                    // part of the cell, about the cell centre (volume_container.hpp:86-91)
                    bool const layer = w >= XY;
                    uint32_t const wl = layer ? w - XY : w;
                    uint32_t const iy = wl / P.dim[0], ix = wl - iy * P.dim[0];
                    uint32_t const iz = layer ? iz_layer : static_cast<uint32_t>(sign);
                    float const ox = (0.5f + static_cast<float>(ix)) - P.half[0];
                    float const oy = (0.5f + static_cast<float>(iy)) - P.half[1];
                    float const oz = (0.5f + static_cast<float>(iz)) - P.half[2];
                    Words const r = P.rs.at(P_GHOST_POS, g, ref, P.step);
                    float const uz = uniform(r.w[2]);
                    float const z = (layer ^ sign) ? sh * uz : sh + (1.0f - sh) * uz;
                    out[0] = (uniform(r.w[0]) - 0.5f) + ox;
                    out[1] = (uniform(r.w[1]) - 0.5f) + oy;
                    out[2] = (z - 0.5f) + oz;
                    out += 3;
                }
                out[0] = g0 * P.tv;
                out[1] = g1 * P.tv;
                out[2] = g2 * P.tv;
            }
        }
        __syncthreads();
    }
}

// The per-cell draws of a step in one launch: blocks below nbw draw the wall
// cells' ghosts (fewer, compute-heavy blocks, dispatched first), the rest (SRD
// only) the cells' axes (memory-heavy), so the two kinds overlap.
__global__ __launch_bounds__(WG) void k_cell_draws(Sys sys, float* __restrict__ axis, uint32_t* __restrict__ gcnt,
                                                   float* __restrict__ gvel, uint32_t nbw, int ext) {
    if (blockIdx.x < nbw) ghosts_block(sys, gcnt, gvel, blockIdx.x, ext != 0);
    else axes_block(sys, axis, blockIdx.x - nbw);
}

__global__ __launch_bounds__(WG) void k_perm(Sys sys, const uint32_t* __restrict__ keys, uint32_t* __restrict__ cursor,
                                             uint32_t* __restrict__ perm, uint64_t N) {
    // neighbouring blocks fill the same cells: keep them on one XCD so their
    // partial cache lines meet in one L2
    perm_block(sys, keys, cursor, perm, N, xcd_tile(blockIdx.x, gridDim.x));
}

// the step's permutation (blocks [0, np)) and, in the same launch, its
// per-cell draws (the blocks after them: the work of k_cell_draws), so a step
// pays for one launch less; the two parts touch disjoint memory.  A kernel of
// its own: the draws' registers would lower k_perm's occupancy
__global__ __launch_bounds__(WG) void k_perm_draws(Sys sys, const uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm,
                                                   uint64_t N, uint32_t np, float* __restrict__ axis,
                                                   uint32_t* __restrict__ gcnt, float* __restrict__ gvel,
                                                   uint32_t nbw, int ext) {
    if (blockIdx.x < np) {
        perm_block(sys, keys, cursor, perm, N, xcd_tile(blockIdx.x, np));
    } else {
        uint32_t const b = blockIdx.x - np;
        if (b < nbw) ghosts_block(sys, gcnt, gvel, b, ext != 0);
        else axes_block(sys, axis, b - nbw);
    }
}

// one step forward of every system's step state: the shift of step t + 1 and
// the next shift (the grid shift is a pure function of (seed, system, step))
__device__ __forceinline__ void advance_system(Sys const& sys, uint32_t* __restrict__ big, uint32_t s) {
    if (s == 0) big[0] = 0u;   // the step's big cells are done
    if (s >= sys.ns) return;
    KParams* const P = sys.P + s;
    uint32_t const step = P->step + 1;
    float nx[3];
    grid_shift(P->rs, step + 1, nx);
    P->step = step;
    for (int a = 0; a < 3; ++a) {
        P->shift[a] = P->shift_next[a];
        P->shift_next[a] = nx[a];
    }
}

// a single system's step state at `step` (its shift and the next one), set
// on the device so that a stream can restart a run without the host table
__global__ void k_set_step(Sys sys, uint32_t step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        KParams* const P = sys.P;
        float s[3], nx[3];
        grid_shift(P->rs, step, s);
        grid_shift(P->rs, step + 1, nx);
        P->step = step;
        for (int a = 0; a < 3; ++a) {
            P->shift[a] = s[a];
            P->shift_next[a] = nx[a];
        }
    }
}

__global__ void k_advance(Sys sys, uint32_t* __restrict__ big) {
    advance_system(sys, big, blockIdx.x * blockDim.x + threadIdx.x);
}

// LDS tile of the collide kernel: up to CAP particles of whole cells, up to
// tile_cell_max(CAP, ALGO) cells.  Sizes are compile-time so every LDS access
// is an immediate offset.
constexpr int TCM = 64;   // largest cells per tile of any tile
// Small tiles keep fewer cells so that more workgroups fit the 160 KB of LDS
// of a CU (allocated in 512-byte granules): SRD 384-particle tiles 24 cells
// (thirteen two-wave workgroups, six or seven waves per SIMD), SRD 448 tiles
// 32 cells (ten workgroups).  Extended 512-particle tiles take 48 cells (eight
// workgroups with the 100-slot ghost region): fewer, fuller tiles win, each
// tile's fixed costs (offsets, permutation and record round trips, barriers)
// being the larger part of its life (500x500x40 cells, n = 10, lane groups of
// two: k_collide 2.82 ms at 48 cells, 2.97 at 40, 3.31 at 32, 3.79 at 24)
constexpr int tile_cell_max(int cap, int algo) {
    return algo == 0 ? (cap <= 384 ? 24 : (cap <= 448 ? 32 : TCM)) : (cap <= 512 ? 48 : TCM);
}
// ghost slots of an extended tile's ghost region (the SRD rule keeps no ghosts
// in the tile: a wall cell's ghosts enter it as one velocity sum).
// Extended 512-particle tiles keep 100 ghost slots: a 48-cell batch rarely
// holds more than a few wall cells, a batch whose ghosts would overflow ends
// early, and the smaller region lets eight workgroups share a CU instead of
// six (500x500x40 cells, n = 10: 256 slots 4.63 ms per step, 192 slots and
// seven workgroups 4.16, 112 slots and eight 3.79; 100 leave room for the
// ghosts' thermal Gaussians at eight).  A region must hold one wall cell's n
// ghosts (choose_tiling).
constexpr int ghost_slots(int cap) { return cap == 512 ? 100 : cap / 2; }

// threads of the collide workgroup for an LDS capacity: more, smaller
// workgroups keep more independent tiles in flight per CU
constexpr int tile_threads(int cap) { return cap <= 256 ? 64 : (cap <= 640 ? 128 : 256); }

// Second launch bound of the collide kernel.  HIP reads it as waves per SIMD,
// so these targets (12 for the 384-particle tile) are more than the LDS lets
// run; the compiler then schedules for the smallest register count it can
// reach.  The bound that matches the real occupancy (6, 5 or 4 waves) was
// measured slower for the extended rule (500x500x40 cells, n = 10: 5.87 ->
// 6.25 ms per step) and equal for SRD, so the larger target stays.
constexpr int tile_blocks(int cap, int algo) {
    return algo == 0 ? (cap <= 384 ? 1536 : (cap <= 448 ? 1280 : 1024)) / tile_threads(cap)
                     : (cap <= 512 ? 1024 : 768) / tile_threads(cap);
}

template <int CAP, int ALGO>
struct Tile {
    static constexpr int TC = tile_cell_max(CAP, ALGO);
    float p[6][CAP];                 // x y z vx vy vz, cell order
    uint32_t perm[CAP];              // source slot, arrival order; later scratch
    float cm[6][TC];                 // per cell: mean velocity xyz, axis xyz
    uint32_t off[TC + 1];            // cell offsets of the tile
    uint8_t go[TC];                  // collision happened
    uint32_t cref[ALGO == 1 ? TC : 1];   // extended: reference cell index
    float cforce[ALGO == 0 ? TC : 1];  // SRD adjoint runs: force term f dt of each cell
    uint32_t gcnt[TC];               // ghost count
    uint32_t goff[ALGO == 1 ? TC : 1];   // extended: table particles kept
    uint32_t wcell[ALGO == 1 ? TC : 1];  // extended: first ghost slot
    uint32_t wpre[ALGO == 1 ? TC + 1 : 1];  // extended: wall cells of the tile before each cell
    // extended: thermal draws (quads) of the batch before each cell (at most
    // (CAP + GS) / 4 + TC: a byte up to the 512 tile)
    std::conditional_t<(CAP <= 512), uint8_t, uint16_t> qb[ALGO == 1 ? TC + 1 : 1];
    uint8_t wk[ALGO == 1 ? TC : 1];         // extended: cell of each wall cell of the batch, in order
    uint8_t cell[CAP];               // tile-local cell of each slot
    static constexpr int GS = ALGO == 1 ? ghost_slots(CAP) : 1;
    float g[ALGO == 1 ? 6 : 1][GS];  // extended: ghosts, n slots per wall cell of the batch
    float gg[GS];                    // extended: the ghosts' thermal Gaussians
};
// eight extended 512-particle workgroups share the 160 KiB of a CU (beside
// the tile: the unsampled bins, 32 bytes, and the workgroup reduction of
// __syncthreads_count, about 230 bytes)
static_assert(sizeof(Tile<512, 1>) + 32 + 240 <= 20480, "extended 512 tile within a CU's eighth of the LDS");
static_assert((512 + ghost_slots(512)) / 4 + tile_cell_max(512, 1) < 256, "quad prefix of the 512 tile in a byte");
// thirteen SRD 384-particle workgroups share a CU (24 LDS granules of 512
// bytes each; 68 VGPRs would allow fourteen)
static_assert(sizeof(Tile<384, 0>) + 64 <= 12288, "SRD 384 tile within 24 LDS granules");

// sampled steps: per-batch LDS bins of the mean-field accumulators (count and
// the three fixed-point velocity sums of each cell, indexed by cell key)
// Stable LSD radix sort of one cell's m source slots by one workgroup, for a
// cell larger than the LDS tile (rare): 8-bit digits, `passes` of them (2
// when every slot is below 2^16, else 4), from src through tmp into dst
// (src -> tmp -> dst [-> tmp -> dst]).  Each chunk of NT slots is placed with
// the digit's running base, the counts of that digit in lower waves, and the
// lane's rank among the wave's lanes holding the same digit (eight ballots).
// sm: 256 * (1 + NT / 64) words of LDS.
template <uint32_t NT>
__device__ void sort_slots(const uint32_t* src, uint32_t* tmp, uint32_t* dst, uint32_t m, int passes,
                           uint32_t* sm) {
    constexpr uint32_t NW = NT / 64;
    static_assert(NT % 64 == 0, "whole waves");
    uint32_t* const base = sm;
    uint32_t* const wc = sm + 256;
    uint32_t const tid = threadIdx.x, w = tid / 64;
    for (int pass = 0; pass < passes; ++pass) {
        const uint32_t* const a = pass == 0 ? src : ((pass & 1) ? tmp : dst);
        uint32_t* const b = (pass & 1) ? dst : tmp;
        uint32_t const sh = 8u * static_cast<uint32_t>(pass);
        for (uint32_t i = tid; i < 256 * (NW + 1); i += NT) sm[i] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < m; i += NT) atomicAdd(&base[(a[i] >> sh) & 255u], 1u);
        __syncthreads();
        if (tid == 0) {
            uint32_t acc = 0;
            for (uint32_t d = 0; d < 256; ++d) {
                uint32_t const c = base[d];
                base[d] = acc;
                acc += c;
            }
        }
        __syncthreads();
        for (uint32_t c0 = 0; c0 < m; c0 += NT) {   // uniform bound: whole waves take the ballots
            uint32_t const i = c0 + tid;
            bool const valid = i < m;
            uint32_t const v = valid ? a[i] : 0u;
            uint32_t const d = (v >> sh) & 255u;
            uint64_t mask = __ballot(valid);
            for (uint32_t bit = 0; bit < 8; ++bit) {
                bool const on = ((d >> bit) & 1u) != 0;
                uint64_t const bb = __ballot(valid && on);
                mask &= on ? bb : ~bb;
            }
            uint32_t const rank = static_cast<uint32_t>(__popcll(mask & __lanemask_lt()));
            if (valid && rank == 0) wc[w * 256 + d] = static_cast<uint32_t>(__popcll(mask));
            __syncthreads();
            if (valid) {
                uint32_t pos = base[d] + rank;
                for (uint32_t q = 0; q < w; ++q) pos += wc[q * 256 + d];
                b[pos] = v;
            }
            __syncthreads();
            for (uint32_t dd = tid; dd < 256; dd += NT) {
                uint32_t t = 0;
                for (uint32_t q = 0; q < NW; ++q) {
                    t += wc[q * 256 + dd];
                    wc[q * 256 + dd] = 0u;
                }
                base[dd] += t;
            }
            __syncthreads();
        }
    }
}

// Sampling windows of a batch (cells outside them go to memory).  A particle's
// un-shifted cell is its collision cell or the neighbour against the shift
// along each axis, so its key lies in one of four ranges of span + 1 keys: the
// batch's keys widened by one cell along the fastest axis on the shift's side,
// that range one row (D0 keys) along the slowest axis, and both one
// middle-axis step away; four windows of tile_cell_max + 1 keys cover them
// for every grid.  The SRD 512-particle tile (up to 64 cells a batch) keeps two
// windows of capacity / 5 keys each (the first two ranges as one contiguous
// window, and the same keys one middle-axis step away): four windows of 65
// keys would cost it a workgroup per CU.
__host__ __device__ constexpr bool sample_four(int cap, int algo) { return !(algo == 0 && cap == 512); }
__host__ __device__ constexpr uint32_t sample_window(int cap, int algo) {
    return sample_four(cap, algo) ? static_cast<uint32_t>(tile_cell_max(cap, algo)) + 1u : static_cast<uint32_t>(cap) / 5;
}
template <uint32_t NB>
struct SampleBins {
    unsigned long long sum[3][NB];
    uint32_t cnt[NB];
};

// SRD per-cell stage of one batch (simulation_functions.cpp:133-182): the mean
// velocity of every cell over its table particles (slot order), summed by the
// cell's lane group (cell_lanes_log2: lane j adds members j, j + G, ...,
// group_sum combines the lanes), plus, for a wall cell with s.gcnt[k] > 0
// ghosts, the sum of their velocities, which the per-cell draws left in the
// cell's mean slots s.cm[0..2][k]; divided by table particles plus ghosts.
template <int CAP, bool ADJ>
__device__ void srd_cells(KParams const& P, Tile<CAP, 0>& s, uint32_t ncell, uint32_t const* off, uint32_t gb,
                          uint32_t key0, Tape const& T, uint64_t CT) {
    constexpr uint32_t NT = tile_threads(CAP);
    uint32_t const lg = cell_lanes_log2(P.n), G = 1u << lg;
    // groups are aligned and NT is a multiple of 64, so a group's lanes are
    // all active or all idle in every round
    for (uint32_t t = threadIdx.x; t < (ncell << lg); t += NT) {
        uint32_t const k = t >> lg, j = t & (G - 1u);
        uint32_t const a = off[k] - gb, mc = off[k + 1] - gb - a;
        uint32_t const reals = mc < P.cap ? mc : P.cap;
        uint32_t const added = s.gcnt[k];
        uint32_t const total = reals + added;
        float mx = 0.0f, my = 0.0f, mz = 0.0f;
        if (total > 1) {   // uniform over the group
            float const* const v = &s.p[3][a];
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            uint32_t i = j;
            // this lane's table particles, loads issued two members at a time
            for (; i + G < reals; i += 2 * G) {
                float const x0 = v[i], y0 = v[CAP + i], z0 = v[2 * CAP + i];
                float const x1 = v[i + G], y1 = v[CAP + i + G], z1 = v[2 * CAP + i + G];
                sx = sx + x0; sy = sy + y0; sz = sz + z0;
                sx = sx + x1; sy = sy + y1; sz = sz + z1;
            }
            for (; i < reals; i += G) {
                sx = sx + v[i]; sy = sy + v[CAP + i]; sz = sz + v[2 * CAP + i];
            }
            sx = group_sum(sx, lg);
            sy = group_sum(sy, lg);
            sz = group_sum(sz, lg);
            // the ghosts' sum (read by every lane of the group before lane 0
            // overwrites the slots with the mean below, in this wave's order)
            if (added) {
                sx = sx + s.cm[0][k];
                sy = sy + s.cm[1][k];
                sz = sz + s.cm[2][k];
            }
            float const inv = 1.0f / static_cast<float>(total);
            mx = sx * inv; my = sy * inv; mz = sz * inv;
            if (j == 0) {
                s.cm[0][k] = mx; s.cm[1][k] = my; s.cm[2][k] = mz;
            }
        }
        if (j == 0) s.go[k] = total > 1;
        if constexpr (ADJ) {
            if (j == 0) {
                uint64_t const gk = P.cb + key0 + k;
                if (T.tot) {
                    T.tot[gk] = total;
                    T.u[gk] = mx; T.u[CT + gk] = my; T.u[2 * CT + gk] = mz;
                }
                float f = 0.0f;
                if (T.force) {
                    uint32_t ix, iy, iz;
                    decode_key(P, key0 + k, ix, iy, iz);
                    f = T.force[P.cb + ref_of(P, ix, iy, iz)] * P.dt;
                }
                s.cforce[k] = f;
            }
        }
    }
}

// Extended rule (simulation_functions.cpp:207-406) in stages so that the
// draws run on densely packed lanes and the per-cell sums run on lane groups:
//   ext_cell_setup   one lane per cell: wall flags, ghost count (:137-146),
//                    table size; stores ref, ghost count (bit 31: upper wall
//                    layer) and the number of table particles kept
//   ext_ghost        one lane per ghost slot of the batch's wall cells: its
//                    position and velocity (:147-167), drawn for the step by
//                    k_cell_draws like the ghost count
//   thermal quads    one lane per draw of four Gaussians (draw q of a cell
//                    covers its members 4q .. 4q + 3, table particles then
//                    ghosts, as in the serial rule), for every member of each
//                    cell that collides
//   extended_cell    one lane group per cell (cell_lanes_log2): the collision itself
constexpr uint32_t UPPER_LAYER = 0x80000000u;

// whether cell layer iz receives ghosts this step: the z layers next to the
// walls for the current shift (simulation_functions.cpp:137-146); `layer`
// tells the upper one
__device__ __forceinline__ bool ghost_layer(KParams const& P, uint32_t iz, bool& layer) {
    bool const sign = P.shift[2] > 0;
    float const zf = static_cast<float>(iz);
    layer = (zf == (P.L[2] - (sign ? 1.0f : 2.0f)));
    return (!P.perio_z) && ((zf == static_cast<float>(sign)) || layer);
}

// returns the cell's thermal quads: (members + 3) / 4 when it collides, else 0
template <int CAP>
__device__ uint32_t ext_cell_setup(KParams const& P, Tile<CAP, 1>& s, uint32_t k, uint32_t mc, uint32_t key,
                                   uint32_t const* __restrict__ gcnt) {
    uint32_t ix, iy, iz;
    decode_key(P, key, ix, iy, iz);
    uint32_t const ref = ref_of(P, ix, iy, iz);
    bool layer;
    bool const ghosts = ghost_layer(P, iz, layer);
    uint32_t const reals_in = mc < P.cap ? mc : P.cap;
    // ghost count drawn for the step by k_cell_draws (wall cell index w)
    uint32_t const added = ghosts ? gcnt[P.wb + (layer ? P.dim[0] * P.dim[1] : 0u) + ix + P.dim[0] * iy] : 0u;
    uint32_t total = reals_in + added;
    if (total > P.cap) total = P.cap;
    s.cref[k] = ref;
    s.gcnt[k] = added | (layer ? UPPER_LAYER : 0u);
    s.goff[k] = total - added;   // table particles kept
    s.go[k] = total > 1;
    return total > 1 ? (total + 3u) >> 2 : 0u;
}

// ghost g of wall cell k of the batch, drawn for the step by k_cell_draws
// (position and velocity, 6 floats), into its ghost slot
template <int CAP>
__device__ __forceinline__ void ext_ghost(KParams const& P, Tile<CAP, 1>& s, uint32_t k, uint32_t g, uint32_t key,
                                          float const* __restrict__ gvel) {
    uint32_t ix, iy, iz;
    decode_key(P, key, ix, iy, iz);
    bool const layer = (s.gcnt[k] & UPPER_LAYER) != 0u;
    uint32_t const w = (layer ? P.dim[0] * P.dim[1] : 0u) + ix + P.dim[0] * iy;
    float2 const* const q = reinterpret_cast<float2 const*>(gvel + P.gvb + 6ull * (static_cast<uint64_t>(P.n) * w + g));
    float2 const a = q[0], b = q[1], c = q[2];
    uint32_t const o = s.wcell[k] + g;   // the cell's first ghost slot + g
    s.g[0][o] = a.x;
    s.g[1][o] = a.y;
    s.g[2][o] = b.x;
    s.g[3][o] = b.y;
    s.g[4][o] = c.x;
    s.g[5][o] = c.y;
}

// the four thermal Gaussians of draw `q` of cell `ref`
__device__ __forceinline__ void thermal_quad(KParams const& P, uint32_t ref, uint32_t q, float* gq) {
    Words const r = P.rs.at(P_THERM, q, ref, P.step);
    box_muller(r.w[0], r.w[1], &gq[0], &gq[1]);
    box_muller(r.w[2], r.w[3], &gq[2], &gq[3]);
}

// extended collision of one cell by its lane group (lane j of G = 2^lg: the
// cell's members j, j + G, ..., table particles first, then ghosts; sums over
// the members combined by group_sum, so every lane of the group holds the
// cell's sums and takes the same decisions); ghosts and the table particles'
// thermal Gaussians (in `gauss`, by LDS slot) were drawn by the stages above.
// Lane 0 of the group leaves the cell's results for ext_finish.
template <int CAP>
__device__ void extended_cell(KParams const& P, Tile<CAP, 1>& s, uint32_t k, uint32_t j, uint32_t lg, uint32_t a,
                              uint32_t key, float const* gauss) {
    uint32_t const G = 1u << lg;
    uint32_t ix, iy, iz;
    decode_key(P, key, ix, iy, iz);
    uint32_t const ref = s.cref[k];
    uint32_t const added = s.gcnt[k] & ~UPPER_LAYER;
    uint32_t const rr = s.goff[k];
    uint32_t const total = rr + added;
    float const ox = (0.5f + static_cast<float>(ix)) - P.half[0];
    float const oy = (0.5f + static_cast<float>(iy)) - P.half[1];
    float const oz = (0.5f + static_cast<float>(iz)) - P.half[2];
    uint32_t const gb = s.wcell[k];
    if (total <= 1) return;   // uniform over the group

    int select, flip;
    float gate;
    if (P.inj_ext) {
        select = static_cast<int>(P.inj_ext[3 * ref]);
        flip = static_cast<int>(P.inj_ext[3 * ref + 1]);
        gate = __uint_as_float(P.inj_ext[3 * ref + 2]);
    } else {
        Words const e = P.rs.at(P_EXT, 0, ref, P.step);
        select = static_cast<int>(e.w[0] / 2) % 13;
        flip = static_cast<int>(e.w[1] / 2) % 2;
        gate = uniform(e.w[2]);
    }
    float ax = c_axes[select][0], ay = c_axes[select][1], az = c_axes[select][2];
    if (flip) { ax = -ax; ay = -ay; az = -az; }

    // this lane's members in order: table particles (LDS slots a + i, i < rr),
    // then ghosts (ghost slots gb + i - rr); f(base, stride, local index,
    // member index, the member's six coordinates).  Coordinate c of a member
    // sits at base[c * S + local]: S = CAP for particles, GS for ghosts,
    // compile-time so the six rows are immediate LDS offsets.
    auto members = [&](auto&& f) {
        uint32_t i = j;
        {
            float* const base = &s.p[0][0] + a;
            constexpr uint32_t S = static_cast<uint32_t>(CAP);
            for (; i < rr; i += G) {
                float x[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) x[c] = base[c * S + i];
                f(base, S, i, i, x);
            }
        }
        {
            float* const base = &s.g[0][0] + gb;
            constexpr uint32_t S = static_cast<uint32_t>(Tile<CAP, 1>::GS);
            for (; i < total; i += G) {
                uint32_t const li = i - rr;
                float x[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) x[c] = base[c * S + li];
                f(base, S, li, i, x);
            }
        }
    };

    float cx = 0, cy = 0, cz = 0, mx = 0, my = 0, mz = 0;
    members([&](float*, uint32_t, uint32_t, uint32_t, float const* x) {
        cx = cx + (x[0] - ox);
        cy = cy + (x[1] - oy);
        cz = cz + (x[2] - oz);
        mx = mx + x[3];
        my = my + x[4];
        mz = mz + x[5];
    });
    cx = group_sum(cx, lg); cy = group_sum(cy, lg); cz = group_sum(cz, lg);
    mx = group_sum(mx, lg); my = group_sum(my, lg); mz = group_sum(mz, lg);
    float const inv = 1.0f / static_cast<float>(total);
    cx = cx * inv + ox;
    cy = cy * inv + oy;
    cz = cz * inv + oz;
    mx = mx * inv;
    my = my * inv;
    mz = mz * inv;

    float proj0 = 0.0f;
    int size0 = 0;
    members([&](float* base, uint32_t S, uint32_t i, uint32_t, float const* x) {
        float const rx = x[0] - cx, ry = x[1] - cy, rz = x[2] - cz;
        float const ux = x[3] - mx, uy = x[4] - my, uz = x[5] - mz;
        base[i] = rx; base[S + i] = ry; base[2 * S + i] = rz;
        base[3 * S + i] = ux; base[4 * S + i] = uy; base[5 * S + i] = uz;
        if (rx * ax + ry * ay + rz * az < 0) {
            proj0 += ux * ax + uy * ay + uz * az;
            size0 += 1;
        }
    });
    proj0 = group_sum(proj0, lg);
    size0 = group_sum(size0, lg);
    int const size1 = static_cast<int>(total) - size0;
    float const proj1 = -proj0 / static_cast<float>(size1);
    proj0 = proj0 / static_cast<float>(size0);
    float const arg = 0.01f * (proj1 - proj0) * static_cast<float>(size0) * static_cast<float>(size1);
    float const probability = 1.0f - exp_f(arg);
    bool const collide = gate < probability;   // uniform over the group

    float Lx = 0, Ly = 0, Lz = 0;
    if (collide) {
        float const t0 = proj1 * (static_cast<float>(size1) / static_cast<float>(size0));
        float const t1 = proj0 * (static_cast<float>(size0) / static_cast<float>(size1));
        float mean0 = 0.0f, mean1 = 0.0f;
        Sym I = {0, 0, 0, 0, 0, 0};
        members([&](float* base, uint32_t S, uint32_t li, uint32_t i, float const* x) {
            // member i's Gaussian of quad i / 4 (drawn for the batch beforehand)
            float const gi = i < rr ? gauss[a + li] : s.gg[gb + li];
            float const rx = x[0], ry = x[1], rz = x[2];
            float ux = x[3], uy = x[4], uz = x[5];
            bool const side = rx * ax + ry * ay + rz * az < 0;
            float const vr = gi * P.tv;
            if (side) mean0 += vr; else mean1 += vr;
            Lx = Lx + (ry * uz - rz * uy);
            Ly = Ly + (rz * ux - rx * uz);
            Lz = Lz + (rx * uy - ry * ux);
            float const sc = vr - (ux * ax + uy * ay + uz * az) + (side ? t0 : t1);
            ux = ux + ax * sc;
            uy = uy + ay * sc;
            uz = uz + az * sc;
            base[3 * S + li] = ux; base[4 * S + li] = uy; base[5 * S + li] = uz;
            float const qx = rx * rx, qy = ry * ry, qz = rz * rz;
            I.xx += qy + qz;
            I.yy += qx + qz;
            I.zz += qx + qy;
            I.xy += -rx * ry;
            I.xz += -rx * rz;
            I.yz += -ry * rz;
        });
        mean0 = group_sum(mean0, lg);
        mean1 = group_sum(mean1, lg);
        float const a0 = mean0 / static_cast<float>(size0), a1 = mean1 / static_cast<float>(size1);
        float const v0x = ax * a0, v0y = ay * a0, v0z = az * a0;
        float const v1x = ax * a1, v1y = ay * a1, v1z = az * a1;
        members([&](float* base, uint32_t S, uint32_t li, uint32_t, float const* x) {
            float const rx = x[0], ry = x[1], rz = x[2];
            bool const side = rx * ax + ry * ay + rz * az < 0;
            float const ux = x[3] - (side ? v0x : v1x);
            float const uy = x[4] - (side ? v0y : v1y);
            float const uz = x[5] - (side ? v0z : v1z);
            base[3 * S + li] = ux; base[4 * S + li] = uy; base[5 * S + li] = uz;
            Lx = Lx - (ry * uz - rz * uy);
            Ly = Ly - (rz * ux - rx * uz);
            Lz = Lz - (rx * uy - ry * ux);
        });
        Lx = group_sum(Lx, lg); Ly = group_sum(Ly, lg); Lz = group_sum(Lz, lg);
        I.xx = group_sum(I.xx, lg); I.yy = group_sum(I.yy, lg); I.zz = group_sum(I.zz, lg);
        I.xy = group_sum(I.xy, lg); I.xz = group_sum(I.xz, lg); I.yz = group_sum(I.yz, lg);
        Sym const J = sym_inverse(I, 1e-5f);
        float const nx = (J.xx * Lx) + (J.xy * Ly) + (J.xz * Lz);
        float const ny = (J.xy * Lx) + (J.yy * Ly) + (J.yz * Lz);
        float const nz = (J.xz * Lx) + (J.yz * Ly) + (J.zz * Lz);
        Lx = nx; Ly = ny; Lz = nz;
    }
    // the way back to the lab frame is per particle: ext_finish runs it on all
    // lanes from these (this cell's ghost count, reference index and ghost
    // slot are not read again, so they carry the correction)
    if (j == 0) {
        s.cm[0][k] = cx; s.cm[1][k] = cy; s.cm[2][k] = cz;
        s.cm[3][k] = mx; s.cm[4][k] = my; s.cm[5][k] = mz;
        s.gcnt[k] = __float_as_uint(Lx);
        s.cref[k] = __float_as_uint(Ly);
        s.wcell[k] = __float_as_uint(Lz);
    }
}

// table particle of cell k after extended_cell, its position and velocity
// relative to the cell (p, v, as extended_cell left them in the tile) in
// registers: velocity back to the lab frame with drag and the
// angular-momentum correction, position back from the centre of mass and
// un-shifted (simulation_functions.cpp:384-391)
template <int CAP>
__device__ __forceinline__ void ext_finish(KParams const& P, Tile<CAP, 1> const& s, uint32_t k, float& px,
                                           float& py, float& pz, float& vx, float& vy, float& vz) {
    float const cx = s.cm[0][k], cy = s.cm[1][k], cz = s.cm[2][k];
    float const mx = s.cm[3][k], my = s.cm[4][k], mz = s.cm[5][k];
    float const Lx = __uint_as_float(s.gcnt[k]), Ly = __uint_as_float(s.cref[k]), Lz = __uint_as_float(s.wcell[k]);
    float const rx = px, ry = py, rz = pz;
    float ux = vx + mx, uy = vy + my, uz = vz + mz;
    ux += P.drag_dt;
    ux = ux - (ry * Lz - rz * Ly);
    uy = uy - (rz * Lx - rx * Lz);
    uz = uz - (rx * Ly - ry * Lx);
    vx = ux; vy = uy; vz = uz;
    px = wrap1((cx + rx) - P.shift[0], P, 0);
    py = wrap1((cy + ry) - P.shift[1], P, 1);
    pz = wrap1((cz + rz) - P.shift[2], P, 2);
}

// particles gathered per thread per round (one for the 384-particle tile,
// which keeps it within 80 VGPRs)
template <int CAP>
constexpr int gather_quota() { return CAP <= 384 ? 1 : 2; }

template <int ALGO, bool SAMPLE, int CAP, bool ADJ, bool BIG>
__device__ __forceinline__ void collide_tile(Sys const& sys, const float* __restrict__ in, float* __restrict__ out,
                                             const uint32_t* __restrict__ perm, const uint32_t* __restrict__ offset,
                                             uint32_t* __restrict__ count_next, uint32_t* __restrict__ keys,
                                             uint32_t* __restrict__ perm_out, SampleBufs const& S,
                                             uint32_t* __restrict__ status, float const* __restrict__ axis,
                                             uint32_t const* __restrict__ gcnt, float const* __restrict__ gvel,
                                             Tape const& T, uint32_t io, uint32_t tile) {
    // io bit 0: `in` holds this step's streamed state (the previous launch
    // wrote it), so the gather skips the streaming; bit 1: write the next
    // step's streamed state instead of the post-collision state.  Both forms
    // are bit-identical to streaming the post-collision state here.
    bool const in_streamed = (io & 1u) != 0, out_streamed = (io & 2u) != 0;
    constexpr uint32_t NT = tile_threads(CAP);
    __shared__ Tile<CAP, ALGO> s;
    constexpr uint32_t WS = sample_window(CAP, ALGO);
    constexpr uint32_t NWIN = sample_four(CAP, ALGO) ? 4u : 2u;
    __shared__ SampleBins<SAMPLE ? NWIN * WS : 1u> sb;
    // BIG: `tile` is an entry of the big-cell list, the cell as a one-cell tile
    uint32_t* const big = status + 1;
    uint32_t const sy = BIG ? big[1 + 2 * tile] : (sys.ns > 1 ? sys.tile_sys[tile] : 0u);
    KParams const P = sys.P[sy];
    uint32_t const cap = CAP, TC = BIG ? 1u : P.tile_cells;
    uint32_t const tid = threadIdx.x;
    uint64_t const N = sys.NT;     // stride of the state planes
    uint64_t const CT = sys.CT;    // stride of the axis planes
    // tile -> cells [c0, c0 + nc) of system sy (local keys); g0 = global key of c0
    uint32_t const c0 = BIG ? big[2 + 2 * tile] : (tile - (sys.ns > 1 ? sys.tile0[sy] : 0u)) * TC;
    if (c0 >= P.C) return;
    uint32_t const nc = min(TC, P.C - c0);
    uint32_t const g0 = P.cb + c0;
    for (uint32_t k = tid; k <= nc; k += NT) s.off[k] = offset[g0 + k];
    if constexpr (ALGO == 1) {
        // wall cells of the tile (those that receive ghosts) counted before
        // each cell: a batch holds at most GS ghosts, n per wall cell
        if (tid < 64) {
            bool wall = false;
            if (tid < nc) {
                uint32_t ix, iy, iz;
                bool layer;
                decode_key(P, c0 + tid, ix, iy, iz);
                wall = ghost_layer(P, iz, layer);
            }
            uint64_t const mask = __ballot(wall);
            if (tid <= nc) s.wpre[tid] = static_cast<uint32_t>(__popcll(mask & ((1ull << tid) - 1ull)));
            if constexpr (Tile<CAP, ALGO>::TC >= 64) {
                if (tid == 0 && nc == 64) s.wpre[64] = static_cast<uint32_t>(__popcll(mask));
            }
        }
    }
    // (s.off and s.wpre are first read after the barrier that follows the
    // batch's permutation load)

    uint32_t kb = 0;
    while (kb < nc) {
        // batch = longest run of whole cells from kb that fits the LDS tile
        // (offsets are non-decreasing, so the fitting cells form a prefix).
        // Every wave finds it on its own from the offsets in memory (the same
        // few words for all; at most 64 cells per tile, one per lane), so no
        // barrier is needed: lane l tests cells [kb, kb + l]
        uint32_t const lane = tid & 63u;
        uint32_t const ob = offset[g0 + kb];
        uint32_t const k1 = kb + 1u + lane;
        uint32_t const ok = k1 <= nc ? offset[g0 + k1] : 0u;
        bool fits = k1 <= nc && ok - ob <= cap;
        if constexpr (ALGO == 1) {
            // and the ghosts of its wall cells fit the ghost region
            bool wall = false;
            if (kb + lane < nc) {
                uint32_t ix, iy, iz;
                bool layer;
                decode_key(P, c0 + kb + lane, ix, iy, iz);
                wall = ghost_layer(P, iz, layer);
            }
            uint64_t const wm = __ballot(wall);
            uint64_t const upto = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);
            fits = fits && static_cast<uint32_t>(__popcll(wm & upto)) * P.n <= static_cast<uint32_t>(s.GS);
        }
        uint32_t const nfit = static_cast<uint32_t>(__popcll(__ballot(fits)));
        uint32_t const ke = kb + (nfit > 0 ? nfit : 1u);
        uint32_t const gb = ob;
        uint32_t m = __builtin_amdgcn_readlane(ok, ke - kb - 1u) - gb;
        // particles [i_begin, i_end) of the batch's range, source slots at
        // srcs[gb + i], moved through without collision: streamed, un-shifted
        // (simulation_functions.cpp:90-93), written, sampled, streamed with the
        // next shift and counted for the next step
        auto pass_through = [&](const uint32_t* srcs, uint32_t i_begin, uint32_t i_end) {
            for (uint32_t i0 = i_begin; i0 < i_end; i0 += NT) {
                uint32_t const i = i0 + tid;
                bool const valid = i < i_end;
                uint32_t knext = 0;
                if (valid) {
                    uint64_t const o = static_cast<uint64_t>(gb) + i;
                    uint32_t const src = srcs[o];
                    float v[6];
                    load_rec(in, src, v);
                    float px = v[0], py = v[1], pz = v[2], vx = v[3], vy = v[4], vz = v[5];
                    if (!in_streamed) stream_shift(P, P.shift, px, py, pz, vx, vy, vz);
                    px = wrap1(px - P.shift[0], P, 0);
                    py = wrap1(py - P.shift[1], P, 1);
                    pz = wrap1(pz - P.shift[2], P, 2);
                    if (!out_streamed) store_rec(out, o, px, py, pz, vx, vy, vz);
                    if (perm_out) perm_out[o] = src;
                    uint32_t ix, iy, iz;
                    if constexpr (SAMPLE) {
                        cell_of(P, px, py, pz, ix, iy, iz);
                        uint64_t const ks = P.cb + key_of(P, ix, iy, iz);
                        atomicAdd(S.cnt + ks, 1u);
                        atomicAdd(S.sum + 3ull * ks + 0, static_cast<unsigned long long>(static_cast<long long>(rint(static_cast<double>(vx) * FIX_SCALE))));
                        atomicAdd(S.sum + 3ull * ks + 1, static_cast<unsigned long long>(static_cast<long long>(rint(static_cast<double>(vy) * FIX_SCALE))));
                        atomicAdd(S.sum + 3ull * ks + 2, static_cast<unsigned long long>(static_cast<long long>(rint(static_cast<double>(vz) * FIX_SCALE))));
                    }
                    stream_shift(P, P.shift_next, px, py, pz, vx, vy, vz);
                    if (out_streamed) store_rec(out, o, px, py, pz, vx, vy, vz);
                    cell_of(P, px, py, pz, ix, iy, iz);
                    knext = P.cb + key_of(P, ix, iy, iz);
                }
                wave_count(count_next, knext, valid);
                if (valid) keys[static_cast<uint64_t>(gb) + i] = knext;
            }
        };
        if (m > cap && (ADJ || (BIG && P.cap > cap))) {
            // A single cell larger than the LDS tile in a taped run, or whose
            // collision table (4n) would not fit the tile either: its particles
            // are moved through without collision so the state and the next
            // step's cell counts stay consistent, and the step reports the
            // condition as an error.
            if (tid == 0) atomicOr(status, 1u);
            pass_through(perm, 0u, m);
            __syncthreads();
            kb = ke;
            continue;
        }
        if (!BIG && m > cap) {
            // A single cell larger than the LDS tile: listed for the big-cell
            // launch that follows
            if (tid == 0) {
                uint32_t const e = atomicAdd(big, 1u);
                big[1 + 2 * e] = sy;
                big[2 + 2 * e] = c0 + kb;
            }
            kb = ke;
            continue;
        }
        // BIG: its source slots are sorted (into its share of the key array,
        // through its share of the output state, both free until written
        // below), its first P.cap particles collide as a one-cell batch, and
        // the rest (the reference's table overflow) are moved through after it.
        const uint32_t* psrc = perm;
        uint32_t tail = 0;
        if (BIG && m > cap) {
            static_assert(6 * CAP >= 256 * (1 + NT / 64), "sort scratch fits the particle arrays");
            sort_slots<NT>(perm + gb, reinterpret_cast<uint32_t*>(out) + rec(gb), keys + gb, m, N > 65536u ? 4 : 2,
                           reinterpret_cast<uint32_t*>(&s.p[0][0]));
            psrc = keys;
            tail = m - P.cap;
            m = P.cap;
            if (tid == 0) s.off[ke] = gb + m;
            __syncthreads();
        }
        uint32_t const ncell = ke - kb;
        for (uint32_t i = tid; i < m; i += NT) s.perm[i] = psrc[gb + i];
        if constexpr (ALGO == 0) {
            // per-cell draws while the permutation loads travel: the axis
            // (drawn here, or from `axis` for steps that keep their axes in
            // memory), and for wall cells the ghost count and the sum of the
            // ghosts' velocities (simulation_functions.cpp:137-171), which
            // waits in the cell's mean slots until the per-cell stage adds it
            bool const sign = P.shift[2] > 0;
            uint32_t const XY = P.dim[0] * P.dim[1];
            for (uint32_t k = tid; k < ncell; k += NT) {
                uint32_t const key = c0 + kb + k;
                uint32_t ix, iy, iz;
                decode_key(P, key, ix, iy, iz);
                if (axis) {
                    uint64_t const gk = P.cb + key;
                    s.cm[3][k] = axis[gk];
                    s.cm[4][k] = axis[CT + gk];
                    s.cm[5][k] = axis[2 * CT + gk];
                } else {
                    float ax, ay, az;
                    srd_axis(P, ref_of(P, ix, iy, iz), ax, ay, az);
                    s.cm[3][k] = ax;
                    s.cm[4][k] = ay;
                    s.cm[5][k] = az;
                }
                float const zf = static_cast<float>(iz);
                bool const layer = (zf == (P.L[2] - (sign ? 1.0f : 2.0f)));
                bool const ghosts = (!P.perio_z) && ((zf == static_cast<float>(sign)) || layer);
                uint32_t gc = 0u;
                float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                if (ghosts) {
                    uint32_t const w = (layer ? XY : 0u) + ix + P.dim[0] * iy;
                    gc = gcnt[P.wb + w];
                    if (gc) {
                        float const* const g = gvel + P.gvb + 3ull * w;
                        gx = g[0];
                        gy = g[1];
                        gz = g[2];
                    }
                }
                s.gcnt[k] = gc;
                s.cm[0][k] = gx;
                s.cm[1][k] = gy;
                s.cm[2][k] = gz;
            }
        }
        __syncthreads();
        // tile-local cell of every slot, four lanes per cell
        for (uint32_t t = tid; t < 4 * ncell; t += NT) {
            uint32_t const k = t / 4;
            uint32_t const a = s.off[kb + k] - gb, e = s.off[kb + k + 1] - gb;
            for (uint32_t i = a + (t & 3u); i < e; i += 4) s.cell[i] = static_cast<uint8_t>(k);
        }
        __syncthreads();
        // gather this thread's particles, GQ at a time with all their loads in
        // flight together, then order each cell by source slot and re-apply
        // this step's streaming
        constexpr int GQ = gather_quota<CAP>();
        for (uint32_t i0 = 0; i0 < m; i0 += GQ * NT) {
            float g[GQ][6];
            uint32_t srcs[GQ];
#pragma unroll
            for (int q = 0; q < GQ; ++q) {
                uint32_t const i = i0 + tid + q * NT;
                srcs[q] = 0;
                if (i < m) {
                    uint32_t const src = s.perm[i];
                    srcs[q] = src;
                    load_rec(in, src, g[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < GQ; ++q) {
                uint32_t const i = i0 + tid + q * NT;
                if (i < m) {
                    uint32_t const k = s.cell[i];
                    uint32_t const a = s.off[kb + k] - gb, e = s.off[kb + k + 1] - gb;
                    uint32_t const src = srcs[q];
                    uint32_t r = 0, j = a;
                    for (; j + 4 <= e; j += 4) {
                        uint32_t const p0 = s.perm[j], p1 = s.perm[j + 1], p2 = s.perm[j + 2], p3 = s.perm[j + 3];
                        r += (p0 < src) + (p1 < src) + (p2 < src) + (p3 < src);
                    }
                    for (; j < e; ++j) r += s.perm[j] < src;
                    uint32_t const d = a + r;
                    float px = g[q][0], py = g[q][1], pz = g[q][2];
                    float vx = g[q][3], vy = g[q][4], vz = g[q][5];
                    bool const bounced = in_streamed ? false : stream_shift(P, P.shift, px, py, pz, vx, vy, vz);
                    s.p[0][d] = px; s.p[1][d] = py; s.p[2][d] = pz;
                    s.p[3][d] = vx; s.p[4][d] = vy; s.p[5][d] = vz;
                    uint64_t const o = static_cast<uint64_t>(gb) + d;
                    if constexpr (ADJ) {
                        if (perm_out) perm_out[o] = src | (bounced ? BOUNCED : 0u);
                        if (T.vp) {
                            T.vp[o] = vx;
                            T.vp[N + o] = vy;
                            T.vp[2 * N + o] = vz;
                        }
                    } else {
                        (void)bounced;
                        if (perm_out) perm_out[o] = src;
                    }
                }
            }
        }
        __syncthreads();
        // next-step histogram bins (in the permutation scratch, free from here
        // on for SRD; the extended rule keeps its draws there until after its
        // per-cell stage) and sampling bins cleared
        constexpr uint32_t NW = 3 * (static_cast<uint32_t>(CAP) / 3);   // histogram words
        auto clear_bins = [&]() {
            for (uint32_t i = tid; i < NW; i += NT) s.perm[i] = 0u;
            if constexpr (SAMPLE) {
                for (uint32_t i = tid; i < NWIN * WS; i += NT) {
                    sb.cnt[i] = 0u;
                    sb.sum[0][i] = 0ull;
                    sb.sum[1][i] = 0ull;
                    sb.sum[2][i] = 0ull;
                }
            }
        };
        // per-cell stage
        if constexpr (ALGO == 0) {
            clear_bins();
            srd_cells<CAP, ADJ>(P, s, ncell, s.off + kb, gb, c0 + kb, T, CT);
        } else {
            // per-cell set-up, one lane of the first wave per cell, and the
            // thermal draws of the batch: quads of four Gaussians covering the
            // members (table particles, then ghosts) of every cell that
            // collides, counted before each cell by a scan of that wave (each
            // lane holds its own cell's count, so no barrier between the two)
            static_assert(Tile<CAP, ALGO>::TC <= 64, "one wave sets up the cells and scans their quads");
            if (tid < 64) {
                uint32_t q = 0;
                if (tid < ncell) {
                    uint32_t const k = tid;
                    q = ext_cell_setup<CAP>(P, s, k, s.off[kb + k + 1] - s.off[kb + k], c0 + kb + k, gcnt);
                    uint32_t const w = s.wpre[kb + k] - s.wpre[kb];
                    s.wcell[k] = w * P.n;
                    if (s.wpre[kb + k + 1] != s.wpre[kb + k]) s.wk[w] = static_cast<uint8_t>(k);
                }
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    uint32_t const y = __shfl_up(q, d, 64);
                    if (tid >= static_cast<uint32_t>(d)) q += y;
                }
                if (tid < ncell) s.qb[tid + 1] = static_cast<std::remove_reference_t<decltype(s.qb[0])>>(q);
                if (tid == 0) s.qb[0] = 0;
            }
            __syncthreads();
            // ghosts, one lane per ghost slot (n slots per wall cell of the batch)
            uint32_t const nslot = (s.wpre[kb + ncell] - s.wpre[kb]) * P.n;
            for (uint32_t t = tid; t < nslot; t += NT) {
                uint32_t const w = P.n > 1 ? udiv(t, P.mn) : t;
                uint32_t const g = t - w * P.n;
                uint32_t const k = s.wk[w];
                if (g < (s.gcnt[k] & ~UPPER_LAYER)) ext_ghost<CAP>(P, s, k, g, c0 + kb + k, gvel);
            }
            // thermal Gaussians, one lane per quad: the table particles' into the
            // permutation scratch by slot (free until the histogram below), the
            // ghosts' by ghost slot
            float* const gauss = reinterpret_cast<float*>(s.perm);
            uint32_t const nq = s.qb[ncell];
            for (uint32_t q = tid; q < nq; q += NT) {
                uint32_t lo = 0, hi = ncell;   // the cell of quad q: the last k with qb[k] <= q
                while (hi - lo > 1) {
                    uint32_t const mid = (lo + hi) >> 1;
                    if (s.qb[mid] <= q) lo = mid; else hi = mid;
                }
                uint32_t const k = lo, j = q - s.qb[k];
                float gq[4];
                if (P.inj_gauss) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) gq[u] = P.inj_gauss[static_cast<uint64_t>(s.cref[k]) * P.cap + 4 * j + u];
                } else {
                    thermal_quad(P, s.cref[k], j, gq);
                }
                uint32_t const a = s.off[kb + k] - gb, rr = s.goff[k];
                uint32_t const tot = rr + (s.gcnt[k] & ~UPPER_LAYER);
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    uint32_t const i = 4 * j + u;
                    if (i < rr) gauss[a + i] = gq[u];
                    else if (i < tot) s.gg[s.wcell[k] + (i - rr)] = gq[u];
                }
            }
            __syncthreads();
            {
                // lane groups of cell_lanes_log2(n) lanes per cell; groups are
                // aligned and NT is a multiple of 64, so a group's lanes are
                // all active or all idle in every round
                uint32_t const lg = cell_lanes_log2(P.n);
                for (uint32_t t = tid; t < (ncell << lg); t += NT) {
                    uint32_t const k = t >> lg;
                    extended_cell<CAP>(P, s, k, t & ((1u << lg) - 1u), lg, s.off[kb + k] - gb, c0 + kb + k, gauss);
                }
            }
            __syncthreads();
            // the permutation scratch held the thermal draws until here
            clear_bins();
        }
        __syncthreads();
        // next step's cell counts: keys near this batch's cells go to three LDS
        // windows in the free permutation scratch (the batch's own rows and the
        // rows one step away along the middle axis, smid keys away), the rest
        // straight to memory.  Window bases may wrap below zero: distances are
        // taken modulo 2^32, and only bins of real keys are ever non-zero.
        // Bins are 16 bits, two to a word (bin g < NW in the low half of word
        // g, bin NW + g in its high half, so neighbouring keys fall on
        // different words; a bin counts at most one batch): NB = 2 NW bins in
        // all, the own window's first, then the two middle-axis windows'.  A
        // window centred on the batch covers moves of up to (width - span) / 2
        // keys; a move along the second-fastest axis is D0 keys.  Equal thirds
        // cover every neighbour while span + 2 (D0 + 1) fits a third (D0 up to
        // about 110 with 384-particle batches); beyond that the middle-axis
        // windows shrink to the batch's keys and the own window takes the rest
        // (D0 up to about 300), so that only moves along both slow axes at
        // once go to memory.
        uint32_t* const hist = s.perm;
        constexpr uint32_t NB = 2 * NW;                                  // bins
        uint32_t const span = ke - kb;
        uint32_t W1 = NB / 3;
        if (span + 2 * (P.D0 + 1) > W1) W1 = min(NB / 3, span + 4);
        uint32_t const W0 = NB - 2 * W1;
        uint32_t const wb0 = g0 + kb - (span < W0 ? (W0 - span) / 2 : 0u);
        uint32_t const smid = mid_stride(P, c0 + kb);
        uint32_t const wb1 = g0 + kb - (span < W1 ? (W1 - span) / 2 : 0u);
        uint32_t const wbm = wb1 - smid, wbp = wb1 + smid;
        // sampling: un-shifted cells lie in the batch's cells or one step back
        // along each axis the shift is positive on, one step ahead where it is
        // negative; the windows (sample_window) take their counts and
        // fixed-point sums in LDS.  wq: first key of each window (modulo 2^32)
        float const sh0 = P.ord[0] == 0 ? P.shift[0] : (P.ord[0] == 1 ? P.shift[1] : P.shift[2]);
        float const shm = P.ord[1] == 0 ? P.shift[0] : (P.ord[1] == 1 ? P.shift[1] : P.shift[2]);
        float const sh2 = P.ord[2] == 0 ? P.shift[0] : (P.ord[2] == 1 ? P.shift[1] : P.shift[2]);
        uint32_t const dslow = sh2 > 0.0f ? 0u - P.D0 : P.D0;   // one row along the slowest axis
        uint32_t const dmid = shm > 0.0f ? 0u - smid : smid;    // one middle-axis step
        uint32_t wq[NWIN];
        if constexpr (NWIN == 4) {
            uint32_t const r0 = g0 + kb - (sh0 > 0.0f ? 1u : 0u);
            wq[0] = r0;
            wq[1] = r0 + dslow;
            wq[2] = r0 + dmid;
            wq[3] = r0 + dslow + dmid;
        } else {
            uint32_t const lo_off = (sh0 > 0.0f ? 1u : 0u) + (sh2 > 0.0f ? P.D0 : 0u);   // keys below the batch
            uint32_t const need = span + 1u + P.D0;
            wq[0] = g0 + kb - lo_off - (need < WS ? (WS - need) / 2 : 0u);
            wq[1] = wq[0] + dmid;
        }
        // per-particle finish, write-out, next-step cell count, sampling
        for (uint32_t i0 = 0; i0 < m; i0 += NT) {
            uint32_t const i = i0 + tid;
            bool const valid = i < m;
            uint32_t knext = 0;
            if (valid) {
                uint32_t const k = s.cell[i];
                uint32_t const slot = i - (s.off[kb + k] - gb);
                float px = s.p[0][i], py = s.p[1][i], pz = s.p[2][i];
                float vx = s.p[3][i], vy = s.p[4][i], vz = s.p[5][i];
                if (slot >= P.cap) {
                    // table overflow: un-shifted, no collision (simulation_functions.cpp:90-93)
                    px = wrap1(px - P.shift[0], P, 0);
                    py = wrap1(py - P.shift[1], P, 1);
                    pz = wrap1(pz - P.shift[2], P, 2);
                } else if constexpr (ALGO == 1) {
                    // table particles of cells with more than one member
                    if (s.go[k] && slot < s.goff[k]) ext_finish<CAP>(P, s, k, px, py, pz, vx, vy, vz);
                } else if (s.go[k]) {
                    // SRD rotation about the cell axis (simulation_functions.cpp:176-188)
                    float const mx = s.cm[0][k], my = s.cm[1][k], mz = s.cm[2][k];
                    float const ax = s.cm[3][k], ay = s.cm[4][k], az = s.cm[5][k];
                    float const ux = vx - mx, uy = vy - my, uz = vz - mz;
                    float const d = ux * ax + uy * ay + uz * az;
                    float const wx = ax * d, wy = ay * d, wz = az * d;
                    float const qx = ux - wx, qy = uy - wy, qz = uz - wz;
                    float const cxp = qy * az - qz * ay, cyp = qz * ax - qx * az, czp = qx * ay - qy * ax;
                    vx = ((wx + qx * P.cos_a) + cxp * P.sin_a) + mx;
                    vy = ((wy + qy * P.cos_a) + cyp * P.sin_a) + my;
                    vz = ((wz + qz * P.cos_a) + czp * P.sin_a) + mz;
                    px = wrap1(px - P.shift[0], P, 0);
                    py = wrap1(py - P.shift[1], P, 1);
                    pz = wrap1(pz - P.shift[2], P, 2);
                    if constexpr (ADJ) {
                        vx += P.drag_dt + s.cforce[k];   // f = 0 adds exactly drag_dt
                    } else {
                        vx += P.drag_dt;
                    }
                }
                uint64_t const o = static_cast<uint64_t>(gb) + i;
                if (!out_streamed) store_rec(out, o, px, py, pz, vx, vy, vz);
                if constexpr (SAMPLE) {
                    uint32_t ix, iy, iz;
                    cell_of(P, px, py, pz, ix, iy, iz);
                    uint32_t const ks = P.cb + key_of(P, ix, iy, iz);
                    unsigned long long const fx = static_cast<unsigned long long>(static_cast<long long>(rint(static_cast<double>(vx) * FIX_SCALE)));
                    unsigned long long const fy = static_cast<unsigned long long>(static_cast<long long>(rint(static_cast<double>(vy) * FIX_SCALE)));
                    unsigned long long const fz = static_cast<unsigned long long>(static_cast<long long>(rint(static_cast<double>(vz) * FIX_SCALE)));
                    uint32_t b = ~0u;
#pragma unroll
                    for (int w = NWIN - 1; w >= 0; --w)
                        if (ks - wq[w] < WS) b = static_cast<uint32_t>(w) * WS + (ks - wq[w]);
                    if (b != ~0u) {
                        atomicAdd(&sb.cnt[b], 1u);
                        atomicAdd(&sb.sum[0][b], fx);
                        atomicAdd(&sb.sum[1][b], fy);
                        atomicAdd(&sb.sum[2][b], fz);
                    } else {
                        atomicAdd(S.cnt + ks, 1u);
                        atomicAdd(S.sum + 3ull * ks + 0, fx);
                        atomicAdd(S.sum + 3ull * ks + 1, fy);
                        atomicAdd(S.sum + 3ull * ks + 2, fz);
                    }
                }
                stream_shift(P, P.shift_next, px, py, pz, vx, vy, vz);
                if (out_streamed) store_rec(out, o, px, py, pz, vx, vy, vz);
                uint32_t ix, iy, iz;
                cell_of(P, px, py, pz, ix, iy, iz);
                knext = P.cb + key_of(P, ix, iy, iz);
            }
            if (valid) {
                keys[static_cast<uint64_t>(gb) + i] = knext;
                uint32_t bin = knext - wb0;
                if (bin >= W0) {
                    uint32_t w = knext - wbm;
                    if (w < W1) {
                        bin = W0 + w;
                    } else {
                        w = knext - wbp;
                        bin = w < W1 ? W0 + W1 + w : ~0u;
                    }
                }
                if (bin < NB) atomicAdd(hist + (bin < NW ? bin : bin - NW), bin < NW ? 1u : 0x10000u);
                else atomicAdd(count_next + knext, 1u);
            }
        }
        __syncthreads();
        // key of bin g, modulo 2^32 (a window base may lie below zero)
        auto bin_key = [&](uint32_t g) -> uint32_t {
            return g < W0 ? wb0 + g : (g < W0 + W1 ? wbm + (g - W0) : wbp + (g - W0 - W1));
        };
        for (uint32_t i = tid; i < NW; i += NT) {
            uint32_t const c = hist[i];
            if (c & 0xffffu) atomicAdd(count_next + bin_key(i), c & 0xffffu);
            if (c >> 16) atomicAdd(count_next + bin_key(NW + i), c >> 16);
        }
        if constexpr (SAMPLE) {
            for (uint32_t i = tid; i < NWIN * WS; i += NT) {
                uint32_t const c = sb.cnt[i];
                if (c) {
                    // consecutive bins are consecutive keys: coalesced atomics
                    uint32_t const w = i / WS;
                    uint32_t base = wq[0];   // wq[w] by selects (no indexed register array)
#pragma unroll
                    for (uint32_t v = 1; v < NWIN; ++v)
                        if (w == v) base = wq[v];
                    uint64_t const rc = base + (i - w * WS);
                    atomicAdd(S.cnt + rc, c);
                    atomicAdd(S.sum + 3ull * rc + 0, sb.sum[0][i]);
                    atomicAdd(S.sum + 3ull * rc + 1, sb.sum[1][i]);
                    atomicAdd(S.sum + 3ull * rc + 2, sb.sum[2][i]);
                }
            }
        }
        // LDS is reused by a next batch or the moved-through tail (after the
        // tile, the big-cell launch's loop has its own barrier)
        if (ke < nc || (BIG && tail)) __syncthreads();
        if (BIG && tail) {
            pass_through(psrc, m, m + tail);
            __syncthreads();
            if (tid == 0) s.off[ke] = gb + m + tail;
            __syncthreads();
        }
        kb = ke;
    }
}

// ------------------------------------------------------ adjoint kernels ---
// Loss of an adjoint run: the x velocity averaged over each z layer of the
// cell grid (the layer of a particle is the z cell index of its stored
// position, as in sampling), then over the steps of the run.

constexpr int LAYERS_MAX = 1024;
constexpr uint64_t LAYER_BLOCK_MAX = 65536;   // particles per k_layers block
constexpr uint32_t LAYER_SLOTS = 16;          // copies of each step's layer totals (by block)
constexpr int LAYER_UNROLL = 8;               // particles per lane with their loads in flight together
constexpr int LAYER_REP_WORDS = 2048;         // LDS histogram entries (layers times lane replicas)

// Per-layer statistics of one step's state, and with `lz` the layer of every
// particle (taped steps).  Statistics: count and x-velocity sum of every z
// layer in fixed point, so the sums do not depend on the order of the
// atomics: each velocity is rounded to a multiple of 2^-40, q = rint(v 2^40).
// A block sums at most LAYER_BLOCK_MAX particles per layer in 64-bit LDS words
// (exact while |v| < 2^23 / LAYER_BLOCK_MAX = 128: a velocity outside, or
// not a number, sets bit 2 of the status word and the run fails); lanes spread over R
// replicas of the histogram (lane mod R), since neighbouring particles share
// a cell and would otherwise queue on one LDS word.  The block adds its sum Q
// to copy blockIdx mod LAYER_SLOTS of the step's totals, split as Q >> 20
// (units of 2^-20) and Q & (2^20 - 1) (units of 2^-40), which cannot
// overflow for fewer than 2^32 particles; the host adds the copies.
// cnt: [LAYER_SLOTS][Z]; sum: [2][LAYER_SLOTS][Z] (hi, then lo).
// With `big`, block 0 also moves the system to its next step (the work of
// k_advance): the pass reads only the system's z extent and half-width, which
// that does not change.
constexpr uint32_t STATUS_LAYER_VELOCITY = 4u;
__global__ __launch_bounds__(WG) void k_layers(Sys sys, const float* __restrict__ st, uint32_t* __restrict__ cnt,
                                               unsigned long long* __restrict__ sum, uint16_t* __restrict__ lz,
                                               uint32_t* __restrict__ status, uint32_t* __restrict__ big) {
    __shared__ uint32_t hc[LAYER_REP_WORDS];
    __shared__ unsigned long long hs[LAYER_REP_WORDS];
    uint64_t const N = sys.NT;
    uint32_t const Z = sys.P[0].dim[2];
    float const half_z = sys.P[0].half[2];
    if (big && blockIdx.x == 0 && threadIdx.x < max(sys.ns, 1u)) advance_system(sys, big, threadIdx.x);
    if (!cnt && !lz) return;
    // replicas: the largest power of two up to 16 that fits (2 or more, as Z <= LAYERS_MAX)
    uint32_t const R = 1u << (31 - __clz(min(16u, static_cast<uint32_t>(LAYER_REP_WORDS) / Z)));
    uint32_t const rep = (threadIdx.x & (R - 1u)) * Z;
    bool const stats = cnt != nullptr;
    if (stats) {
        for (uint32_t i = threadIdx.x; i < R * Z; i += WG) {
            hc[i] = 0u;
            hs[i] = 0ull;
        }
        __syncthreads();
    }
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * WG * LAYER_UNROLL;
    for (uint64_t j0 = blockIdx.x * static_cast<uint64_t>(WG) * LAYER_UNROLL + threadIdx.x; j0 < N; j0 += stride) {
        float2 r[LAYER_UNROLL];
#pragma unroll
        for (int q = 0; q < LAYER_UNROLL; ++q) {
            uint64_t const j = j0 + static_cast<uint64_t>(q) * WG;
            r[q] = j < N ? *reinterpret_cast<float2 const*>(st + rec(j) + 2) : make_float2(0.0f, 0.0f);
        }
#pragma unroll
        for (int q = 0; q < LAYER_UNROLL; ++q) {
            uint64_t const j = j0 + static_cast<uint64_t>(q) * WG;
            if (j < N) {
                uint32_t const iz = mod_dim(static_cast<uint32_t>(floorf(r[q].x + half_z)), Z);
                if (lz) lz[j] = static_cast<uint16_t>(iz);
                if (stats) {
                    if (!(fabsf(r[q].y) < 128.0f)) atomicOr(status, STATUS_LAYER_VELOCITY);
                    atomicAdd(&hc[rep + iz], 1u);
                    atomicAdd(&hs[rep + iz], static_cast<unsigned long long>(static_cast<long long>(
                                                 rint(static_cast<double>(r[q].y) * FIX_SCALE))));
                }
            }
        }
    }
    if (stats) {
        __syncthreads();
        uint32_t const slot = (blockIdx.x % LAYER_SLOTS) * Z;
        for (uint32_t i = threadIdx.x; i < Z; i += WG) {
            uint32_t c = 0;
            long long q = 0;
            for (uint32_t k = 0; k < R; ++k) {
                c += hc[k * Z + i];
                q += static_cast<long long>(hs[k * Z + i]);
            }
            if (c) {
                atomicAdd(cnt + slot + i, c);
                atomicAdd(sum + slot + i, static_cast<unsigned long long>(q >> 20));
                atomicAdd(sum + LAYER_SLOTS * Z + slot + i, static_cast<unsigned long long>(q & 0xfffffll));
            }
        }
    }
}

// Reverse of one recorded SRD step (pathwise: cells, tables, bounces, axes and
// ghosts of the step held fixed).  With v' the streamed velocity, u the cell
// mean over the table and the ghosts, R the rotation, the step maps
//   v'' = u + R (v' - u) + (g + f_c) dt e_x   (table particles of cells with more than one member)
//   v'' = v'                                   (overflow particles, lone cells)
// and v' = s v with s = -1 for a bounced particle.  The adjoint of a table
// particle is mu_i = R^T lam_i + (I - R^T) sum_table(lam) / total, and
//   dL/dg += dt sum lam_x,  dL/df_c += dt sum_c lam_x,  dL/dalpha += sum lam . dR/dalpha (v' - u).
// lam_in: dL/dv'' of the step's output slots (x, y, z planes), the loss seed
// of this step added to its x plane; lam_out: dL/dv of the input slots.
// part[b], part[G + b]: this block's running sums of dL/dg and dL/dalpha (radians);
// gf: dL/df per reference cell.
// ADJ_CELLS cells per block.  The kernel is bound by the latency of its
// dependent memory rounds (a launch is about one wave of blocks), so it has
// two: first the cell offsets, the cells' tape (axis, mean, total), their
// force-gradient entries and the step's loss-seed table; then every slot's
// adjoint, layer, source slot and streamed velocity, held in registers.  One
// lane per cell forms the cell's table sums (in slot order, FP64) from LDS and
// its rotation of the mean; then every lane finishes its own slots and
// scatters them.  A block whose cells hold more particles than the staging
// buffer runs the same arithmetic one lane per cell, reading memory directly.
constexpr uint32_t ADJ_CELLS = 64;     // cells per block
constexpr uint32_t ADJ_STAGE = 768;     // particles staged per block (12 n cells' worth at n = 10)

struct AdjIn {
    const uint32_t* off;
    const float* u;
    const float* ax;
    const uint32_t* tot;
    const uint32_t* perm;
    const float* vp;
    const uint16_t* lz;     // z layer of every output slot
    const float* stab;      // loss seed dL/dv_x of a particle in each layer after the step
    const float* lam_in;
    float* lam_out;
};

// R^T x: x rotated by -alpha about the axis (axx, axy, axz)
__device__ __forceinline__ void adj_rot_t(KParams const& P, float axx, float axy, float axz, float x, float y, float z,
                                          float* o3) {
    float const d = x * axx + y * axy + z * axz;
    float const px = x - axx * d, py = y - axy * d, pz = z - axz * d;
    float const cx = py * axz - pz * axy, cy = pz * axx - px * axz, cz = px * axy - py * axx;
    o3[0] = axx * d + P.cos_a * px - P.sin_a * cx;
    o3[1] = axy * d + P.cos_a * py - P.sin_a * cy;
    o3[2] = axz * d + P.cos_a * pz - P.sin_a * cz;
}

// scatter of one particle's adjoint back to its source slot (sign of a bounce)
__device__ __forceinline__ void adj_put(AdjIn const& A, uint64_t N, uint32_t o, float const* m) {
    uint32_t const pe = A.perm[o];
    uint32_t const src = pe & ~BOUNCED;
    float const sg = (pe & BOUNCED) ? -1.0f : 1.0f;
    A.lam_out[src] = sg * m[0];
    A.lam_out[N + src] = sg * m[1];
    A.lam_out[2 * N + src] = sg * m[2];
}

// dL/dalpha contribution of table particle o: l . dR/dalpha (v' - u), with
// dR/dalpha (v' - u) = -sin q + cos (q x a), q the part of v' - u normal to a
__device__ __forceinline__ double adj_dalpha_w(KParams const& P, float const* v, float const* l, float axx, float axy,
                                               float axz, float ux, float uy, float uz) {
    float const wx = v[0] - ux, wy = v[1] - uy, wz = v[2] - uz;
    float const d = wx * axx + wy * axy + wz * axz;
    float const qx = wx - axx * d, qy = wy - axy * d, qz = wz - axz * d;
    float const cx = qy * axz - qz * axy, cy = qz * axx - qx * axz, cz = qx * axy - qy * axx;
    return static_cast<double>(l[0]) * (-P.sin_a * qx + P.cos_a * cx) +
           static_cast<double>(l[1]) * (-P.sin_a * qy + P.cos_a * cy) +
           static_cast<double>(l[2]) * (-P.sin_a * qz + P.cos_a * cz);
}
__device__ __forceinline__ double adj_dalpha(KParams const& P, AdjIn const& A, uint64_t N, uint32_t o, float const* l,
                                             float axx, float axy, float axz, float ux, float uy, float uz) {
    float const v[3] = {A.vp[o], A.vp[N + o], A.vp[2 * N + o]};
    return adj_dalpha_w(P, v, l, axx, axy, axz, ux, uy, uz);
}

// the whole reverse step of cell `key` on one lane (blocks too large to stage)
__device__ void adjoint_cell_serial(KParams const& P, AdjIn const& A, uint64_t N, uint64_t C, uint32_t key,
                                    double* __restrict__ gf, double& dg, double& da) {
    uint32_t const a = A.off[key], e = A.off[key + 1];
    uint32_t const total = A.tot[key];
    uint32_t const reals = min(e - a, P.cap);
    auto lam = [&](uint32_t o, float* l) {
        l[0] = A.lam_in[o] + A.stab[A.lz[o]];
        l[1] = A.lam_in[N + o];
        l[2] = A.lam_in[2 * N + o];
    };
    uint32_t first_plain = a;
    if (total > 1) {
        float const axx = A.ax[key], axy = A.ax[C + key], axz = A.ax[2 * C + key];
        float const ux = A.u[key], uy = A.u[C + key], uz = A.u[2 * C + key];
        double sx = 0, sy = 0, sz = 0;
        for (uint32_t o = a; o < a + reals; ++o) {
            float l[3];
            lam(o, l);
            sx += l[0];
            sy += l[1];
            sz += l[2];
        }
        float const inv = static_cast<float>(1.0 / static_cast<double>(total));
        float const mx = static_cast<float>(sx) * inv, my = static_cast<float>(sy) * inv, mz = static_cast<float>(sz) * inv;
        float rm[3];
        adj_rot_t(P, axx, axy, axz, mx, my, mz, rm);
        float const cmx = mx - rm[0], cmy = my - rm[1], cmz = mz - rm[2];
        for (uint32_t o = a; o < a + reals; ++o) {
            float l[3];
            lam(o, l);
            float rl[3];
            adj_rot_t(P, axx, axy, axz, l[0], l[1], l[2], rl);
            float const m[3] = {rl[0] + cmx, rl[1] + cmy, rl[2] + cmz};
            da += adj_dalpha(P, A, N, o, l, axx, axy, axz, ux, uy, uz);
            adj_put(A, N, o, m);
        }
        // dL/dg: dt times the x adjoint summed over the table (the same sum as sx)
        double const gx = sx * static_cast<double>(P.dt);
        dg += gx;
        uint32_t ix, iy, iz;
        decode_key(P, key, ix, iy, iz);
        gf[P.cb + ref_of(P, ix, iy, iz)] += gx;
        first_plain = a + reals;
    }
    for (uint32_t o = first_plain; o < e; ++o) {
        float l[3];
        lam(o, l);
        adj_put(A, N, o, l);
    }
}

__global__ __launch_bounds__(WG) void k_adjoint_step(Sys sys, AdjIn A, double* __restrict__ part,
                                                     double* __restrict__ gf) {
    constexpr uint32_t PPT = ADJ_STAGE / WG;   // staged slots per lane
    static_assert(PPT * WG == ADJ_STAGE, "whole slots per lane");
    __shared__ float sl[3][ADJ_STAGE];   // adjoint (plus loss seed) of the block's particles
    __shared__ uint8_t scell[ADJ_STAGE]; // block-local cell of each staged particle
    __shared__ uint32_t soff[ADJ_CELLS + 1];
    __shared__ uint32_t sreal[ADJ_CELLS];      // table particles of a colliding cell, else 0
    __shared__ float scell_v[9][ADJ_CELLS];    // axis, cell mean u, (I - R^T) applied to the table mean
    __shared__ float sstab[LAYERS_MAX];        // the step's loss seed per layer
    __shared__ double red[2][WG / 64];
    KParams const P = sys.P[0];
    uint64_t const N = sys.NT, C = sys.CT;
    uint32_t const tid = threadIdx.x;
    static_assert(ADJ_CELLS <= WG && ADJ_CELLS <= 256, "one lane per cell, cell index in a byte");
    // neighbouring cell ranges scatter into the same source windows: keep them
    // on one XCD so their partial cache lines meet in one L2
    uint32_t const bt = xcd_tile(blockIdx.x, gridDim.x);
    uint32_t const k0 = bt * ADJ_CELLS;
    uint32_t const ncell = min(ADJ_CELLS, P.C - k0);   // k0 < P.C for every launched block
    // round 1
    for (uint32_t k = tid; k <= ncell; k += WG) soff[k] = A.off[k0 + k];
    for (uint32_t z = tid; z < P.dim[2]; z += WG) sstab[z] = A.stab[z];
    uint32_t ctot = 0, cref = 0;
    float cax = 0, cay = 0, caz = 0, cux = 0, cuy = 0, cuz = 0;
    double gf_old = 0.0;
    if (tid < ncell) {
        uint32_t const key = k0 + tid;
        ctot = A.tot[key];
        cax = A.ax[key]; cay = A.ax[C + key]; caz = A.ax[2 * C + key];
        cux = A.u[key]; cuy = A.u[C + key]; cuz = A.u[2 * C + key];
        uint32_t ix, iy, iz;
        decode_key(P, key, ix, iy, iz);
        cref = P.cb + ref_of(P, ix, iy, iz);
        if (ctot > 1) gf_old = gf[cref];
    }
    __syncthreads();
    uint32_t const pa = soff[0], m = soff[ncell] - pa;
    double dg = 0.0, da = 0.0;
    if (m <= ADJ_STAGE) {
        // round 2
        float l[PPT][3], w[PPT][3];
        uint32_t pe[PPT];
        uint16_t lzv[PPT];
#pragma unroll
        for (uint32_t q = 0; q < PPT; ++q) {
            uint32_t const i = tid + q * WG;
            if (i < m) {
                uint64_t const o = pa + i;
                l[q][0] = A.lam_in[o];
                l[q][1] = A.lam_in[N + o];
                l[q][2] = A.lam_in[2 * N + o];
                lzv[q] = A.lz[o];
                pe[q] = A.perm[o];
                w[q][0] = A.vp[o];
                w[q][1] = A.vp[N + o];
                w[q][2] = A.vp[2 * N + o];
            }
        }
        if (tid < ncell)
            for (uint32_t i = soff[tid] - pa; i < soff[tid + 1] - pa; ++i) scell[i] = static_cast<uint8_t>(tid);
#pragma unroll
        for (uint32_t q = 0; q < PPT; ++q) {
            uint32_t const i = tid + q * WG;
            if (i < m) {
                l[q][0] += sstab[lzv[q]];
                sl[0][i] = l[q][0];
                sl[1][i] = l[q][1];
                sl[2][i] = l[q][2];
            }
        }
        __syncthreads();
        if (tid < ncell) {
            uint32_t const a = soff[tid] - pa, e = soff[tid + 1] - pa;
            uint32_t const reals = min(e - a, P.cap);
            sreal[tid] = 0u;
            if (ctot > 1) {
                double sx = 0, sy = 0, sz = 0;
                for (uint32_t i = a; i < a + reals; ++i) {
                    sx += sl[0][i];
                    sy += sl[1][i];
                    sz += sl[2][i];
                }
                float const inv = static_cast<float>(1.0 / static_cast<double>(ctot));
                float const mx = static_cast<float>(sx) * inv, my = static_cast<float>(sy) * inv,
                            mz = static_cast<float>(sz) * inv;
                float rm[3];
                adj_rot_t(P, cax, cay, caz, mx, my, mz, rm);
                scell_v[0][tid] = cax; scell_v[1][tid] = cay; scell_v[2][tid] = caz;
                scell_v[3][tid] = cux; scell_v[4][tid] = cuy; scell_v[5][tid] = cuz;
                scell_v[6][tid] = mx - rm[0]; scell_v[7][tid] = my - rm[1]; scell_v[8][tid] = mz - rm[2];
                sreal[tid] = reals;
                double const gx = sx * static_cast<double>(P.dt);
                dg = gx;
                gf[cref] = gf_old + gx;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < PPT; ++q) {
            uint32_t const i = tid + q * WG;
            if (i < m) {
                uint32_t const k = scell[i];
                float mm[3] = {l[q][0], l[q][1], l[q][2]};
                if (i - (soff[k] - pa) < sreal[k]) {
                    float const axx = scell_v[0][k], axy = scell_v[1][k], axz = scell_v[2][k];
                    float rl[3];
                    adj_rot_t(P, axx, axy, axz, l[q][0], l[q][1], l[q][2], rl);
                    mm[0] = rl[0] + scell_v[6][k];
                    mm[1] = rl[1] + scell_v[7][k];
                    mm[2] = rl[2] + scell_v[8][k];
                    da += adj_dalpha_w(P, w[q], l[q], axx, axy, axz, scell_v[3][k], scell_v[4][k], scell_v[5][k]);
                }
                uint32_t const src = pe[q] & ~BOUNCED;
                float const sg = (pe[q] & BOUNCED) ? -1.0f : 1.0f;
                A.lam_out[src] = sg * mm[0];
                A.lam_out[N + src] = sg * mm[1];
                A.lam_out[2 * N + src] = sg * mm[2];
            }
        }
    } else if (tid < ncell) {
        adjoint_cell_serial(P, A, N, C, k0 + tid, gf, dg, da);
    }
    // block sums, kept per block so the final sum has a fixed order
    for (int d = 32; d >= 1; d >>= 1) {
        dg += __shfl_down(dg, d, 64);
        da += __shfl_down(da, d, 64);
    }
    uint32_t const lane = tid & 63u, wv = tid >> 6;
    if (lane == 0) {
        red[0][wv] = dg;
        red[1][wv] = da;
    }
    __syncthreads();
    if (tid == 0) {
        double tg = 0, ta = 0;
        for (int k = 0; k < WG / 64; ++k) {
            tg += red[0][k];
            ta += red[1][k];
        }
        part[bt] += tg;
        part[gridDim.x + bt] += ta;
    }
}

// largest per-cell count (the capacity a batch needs for the fullest cell)
__global__ __launch_bounds__(WG) void k_count_max(const uint32_t* __restrict__ count, uint32_t C,
                                                  uint32_t* __restrict__ out) {
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(WG) + threadIdx.x; i < C;
         i += static_cast<uint64_t>(gridDim.x) * WG)
        m = max(m, count[i]);
    for (int d = 32; d >= 1; d >>= 1) m = max(m, static_cast<uint32_t>(__shfl_down(m, d, 64)));
    if ((threadIdx.x & 63u) == 0) atomicMax(out, m);
}

// fold one sampled step into the FP64 accumulators; the per-step sums and the
// accumulators are both indexed by global cell key (reference order is
// restored where they leave the device)
__global__ __launch_bounds__(WG) void k_sample_fold(uint32_t C, uint32_t* __restrict__ cnt,
                                                    unsigned long long* __restrict__ sum, double* __restrict__ acc,
                                                    int first) {
    for (uint64_t c = blockIdx.x * static_cast<uint64_t>(WG) + threadIdx.x; c < C;
         c += static_cast<uint64_t>(gridDim.x) * WG) {
        uint32_t const n = cnt[c];
        double const d = static_cast<double>(n);
        double m[3];
        for (int k = 0; k < 3; ++k) {
            long long const v = static_cast<long long>(sum[3 * c + k]);
            m[k] = n > 0 ? (static_cast<double>(v) / FIX_SCALE) / d : 0.0;
            sum[3 * c + k] = 0ull;
        }
        cnt[c] = 0u;
        if (first) {
            acc[4 * c] = d;
            for (int k = 0; k < 3; ++k) acc[4 * c + 1 + k] = m[k];
        } else {
            acc[4 * c] += d;
            for (int k = 0; k < 3; ++k) acc[4 * c + 1 + k] += m[k];
        }
    }
}

// plain device copy, the practical HBM reference point for the roofline
// Copy of n float4 for the bandwidth ceiling: a block moves WG U float4 in
// one pass, each lane with its U loads in flight together (coalesced, lane
// stride 16 bytes), stores plain or nontemporal
using Vec4 = float __attribute__((ext_vector_type(4)));
template <int U, bool NTS>
__global__ __launch_bounds__(WG) void k_copy(const Vec4* __restrict__ src, Vec4* __restrict__ dst, uint64_t n) {
    uint64_t const base = blockIdx.x * static_cast<uint64_t>(WG) * U + threadIdx.x;
    Vec4 v[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
        uint64_t const i = base + static_cast<uint64_t>(q) * WG;
        if (i < n) v[q] = src[i];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
        uint64_t const i = base + static_cast<uint64_t>(q) * WG;
        if (i < n) {
            if constexpr (NTS) __builtin_nontemporal_store(v[q], dst + i);
            else dst[i] = v[q];
        }
    }
}

__global__ void k_selftest_philox(uint32_t k0, uint32_t k1, const uint32_t* ctr, uint32_t* out, size_t n) {
    size_t const i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        Words const r = philox(k0, k1, ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]);
        for (int k = 0; k < 4; ++k) out[4 * i + k] = r.w[k];
    }
}

__global__ void k_selftest_math(const float* in, float* out, size_t n, int which) {
    size_t const i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    float const x = in[i];
    float r0 = 0, r1 = 0;
    switch (which) {
        case 0: r0 = log_pos(x); break;
        case 1: sincospi(x, &r0, &r1); break;
        case 2: r0 = exp_f(x); break;
        case 3: r0 = sqrtf(x); break;
        case 4: r0 = fmodf(x, 7.0f); r1 = fmodf(x + 1.5f * 7.0f, 7.0f) - 7.0f * 0.5f; break;
        case 5: r0 = 1.0f / x; r1 = x / 3.0f; break;
        default: break;
    }
    out[2 * i] = r0;
    out[2 * i + 1] = r1;
}

}  // namespace mpcd_hip

// Scratch of the allocation-free step (mpcd_hip_step_noalloc): the step's
// cell-sort and per-cell arrays for a system of up to NOALLOC_MAX_PARTICLES
// particles and NOALLOC_MAX_CELLS cells, 14 MiB per device, held from the
// library's load until the process ends.  In host code `scratch` names a
// host-side stand-in; its device address comes from hipGetSymbolAddress, once
// per device (noalloc_scratch below).
constexpr uint64_t NOALLOC_MAX_PARTICLES = uint64_t(1) << 20;
constexpr uint64_t NOALLOC_MAX_CELLS = uint64_t(1) << 17;
constexpr uint64_t NOALLOC_MAX_GHOST_FLOATS = uint64_t(1) << 20;   // wall ghosts: (3 SRD, 6 n extended) 2 X Y
constexpr uint64_t noalloc_align(uint64_t b) { return (b + 255) / 256 * 256; }
// byte offsets of the regions
constexpr uint64_t NA_KP = 0;
constexpr uint64_t NA_BLK = NA_KP + noalloc_align(sizeof(mpcd_hip::KParams));
constexpr uint64_t NA_PERM = NA_BLK + noalloc_align(4 * 8);
constexpr uint64_t NA_KEYS = NA_PERM + noalloc_align(4 * NOALLOC_MAX_PARTICLES);
constexpr uint64_t NA_COUNT = NA_KEYS + noalloc_align(4 * NOALLOC_MAX_PARTICLES);
constexpr uint64_t NA_OFFSET = NA_COUNT + noalloc_align(4 * NOALLOC_MAX_CELLS);
constexpr uint64_t NA_CURSOR = NA_OFFSET + noalloc_align(4 * (NOALLOC_MAX_CELLS + 1));
constexpr uint64_t NA_PARTIAL = NA_CURSOR + noalloc_align(4 * NOALLOC_MAX_CELLS);
constexpr uint64_t NA_STATUS = NA_PARTIAL + noalloc_align(4 * (NOALLOC_MAX_CELLS / mpcd_hip::SCAN_CHUNK + 1));
constexpr uint64_t NA_GCNT = NA_STATUS + noalloc_align(4 * (2 + 2 * (NOALLOC_MAX_PARTICLES / 257 + 2)));
constexpr uint64_t NA_GVEL = NA_GCNT + noalloc_align(4 * NOALLOC_MAX_CELLS);
constexpr uint64_t NOALLOC_SCRATCH_BYTES = NA_GVEL + noalloc_align(4 * NOALLOC_MAX_GHOST_FLOATS);
static_assert(NA_STATUS + 4 * mpcd_hip::big_list_words(NOALLOC_MAX_PARTICLES) <= NA_GCNT, "big-cell list fits");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// ============================================================== host side ===

using namespace mpcd_hip;

namespace {

std::mutex g_err_mu;
std::string g_create_err;

struct HipError {
    std::string msg;
};

inline void ck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError{std::string(what) + ": " + hipGetErrorString(e)};
}

}  // namespace

struct mpcd_hip_s {
    std::vector<mpcd_params_t> prm;  // one per system
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<KParams> kp;         // host copy of the system table
    uint32_t ns = 0;
    int algorithm = 0;
    uint64_t N = 0, C = 0, W = 0;    // particles, cells, wall cells of all systems
    uint64_t GV = 0;                 // ghost-velocity floats of all systems
    float* st[2] = {nullptr, nullptr};
    int cur = 0;
    uint32_t* perm = nullptr;
    uint32_t* keys = nullptr;
    uint32_t* cursor = nullptr;
    uint32_t* count = nullptr;
    uint32_t* offset = nullptr;
    uint32_t* partial = nullptr;
    uint32_t nb = 0;
    uint32_t* status = nullptr;
    uint32_t* cmax = nullptr;        // largest per-cell count (k_count_max)
    uint32_t* perm_rec = nullptr;
    bool record = false, recorded = false;
    float* axis = nullptr;           // SRD per-cell axes of the current step, when kept in memory
    // the step's SRD axes go through `axis` (drawn by k_cell_draws or given by
    // the caller) instead of being drawn in the collide tiles: taped adjoint
    // steps (their reverse pass reads them) and injected steps
    bool axes_mem = false;
    uint32_t* wgcnt = nullptr;       // wall-cell ghost counts of the step
    float* wgvel = nullptr;          // wall-cell ghosts of the step: velocity sums (SRD), positions and velocities (extended)
    // draws of the next step given by the caller (mpcd_hip_step_injected):
    // axes by internal cell key (3 planes), wall-cell ghost counts and
    // ghosts in the layout of wgcnt / wgvel; null: the step draws its own
    struct Injected {
        float shift[3];
        std::vector<float> axes;
        std::vector<uint32_t> gcnt;
        std::vector<float> gvel;
    };
    Injected const* inject = nullptr;
    // system table and block -> system maps on the device
    KParams* d_kp = nullptr;
    uint32_t* d_blk = nullptr;       // blk_part | blk_cell | blk_wall | tile0, ns + 1 entries each
    uint16_t* d_tile_sys = nullptr;
    uint64_t tile_sys_cap = 0;
    uint32_t blocks_part = 0, blocks_cell = 0, blocks_wall = 0, ntiles = 0;
    Sys sys{};
    // sampling
    uint32_t* scnt = nullptr;
    unsigned long long* ssum = nullptr;
    double* acc = nullptr;
    bool sampling_done = true;
    uint64_t sample_counter = 0;
    uint64_t step_index = 0;
    bool keys_valid = false, sorted_valid = false;
    bool streamed_pending = false;   // the state buffer holds an intermediate (streamed) state
    uint32_t lds_cap = 0;            // largest LDS capacity of the systems' tiles
    // collide launches: systems sharing a compiled LDS capacity form a group
    // whose tiles are contiguous, [t0, t1), and take one launch
    struct Group {
        uint32_t cap, t0, t1, gid;
    };
    static constexpr int MAX_GROUPS = 16;   // 7 capacities x 2 launch sets
    Group grp[MAX_GROUPS] = {};
    uint32_t ngrp = 0;
    // autotuning: HIP events around each group's collide launch of eager steps,
    // and the milliseconds of each group summed over the steps
    bool group_timing = false;
    hipEvent_t gev[2 * MAX_GROUPS] = {};
    double group_ms[MAX_GROUPS] = {};
    // captured steps, by input buffer and sampling
    bool use_graph = true;
    hipGraphExec_t graph[2][2][4] = {};   // [input buffer][sampling][streamed in/out]
    int graph_launches[2] = {0, 0};
    // adjoint runs: tape of one checkpoint segment (two when there are more
    // segments: the forward pass tapes the last two, and later one is
    // recomputed while the other is run backwards), checkpoints, adjoint state
    struct Adj {
        uint32_t K = 0, nck = 0, steps = 0, halves = 0;
        hipStream_t recompute = nullptr;   // second stream of the segment recomputes
        hipEvent_t written[2] = {nullptr, nullptr}, read[2] = {nullptr, nullptr};   // per tape half
        uint32_t* perm = nullptr;    // [K][N] recorded permutation with the bounce flag
        float* vp = nullptr;         // [K][3N] streamed velocities
        uint16_t* lz = nullptr;      // [K][N] z layer of every output slot
        uint32_t* off = nullptr;     // [K][C+1] cell offsets (the step's sort writes them here)
        float* u = nullptr;          // [K][3C] cell means
        float* ax = nullptr;         // [K][3C] cell axes (the step's draws write them here)
        uint32_t* tot = nullptr;     // [K][C] cell totals
        float* ckpt = nullptr;       // [nck][6N] states at the checkpoints (first: the start state)
        float* lam[2] = {nullptr, nullptr};
        uint32_t* lcnt = nullptr;    // [steps][LAYER_SLOTS][Z] layer counts
        unsigned long long* lsum = nullptr;  // [steps][2][LAYER_SLOTS][Z] layer x-velocity sums, fixed point
        float* stab = nullptr;       // [steps][Z] loss seed of a particle in each layer after each step
        double* part = nullptr;      // [2][blocks] per-block gradient sums
        double* gf = nullptr;        // [C] dL/df
        float* force = nullptr;      // [C] body force of the run
    } adj;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    double t_collide = 0, t_total = 0;
    int launches = 0;
};

namespace {

void fill_geometry(mpcd_params_t const& p, KParams& k) {
    k = KParams{};
    for (int a = 0; a < 3; ++a) {
        k.L[a] = p.volume_size[a];
        k.half[a] = p.volume_size[a] * 0.5f;
        k.L15[a] = p.volume_size[a] * 1.5f;
        k.L2[a] = p.volume_size[a] * 2.0f;
        k.L25[a] = p.volume_size[a] * 2.5f;
        float const l3 = p.volume_size[a] * 3.0f;
        std::memcpy(&k.L3bits[a], &l3, 4);
        k.dim[a] = static_cast<uint32_t>(p.volume_size[a]);
    }
    k.C = k.dim[0] * k.dim[1] * k.dim[2];
    // internal cell order: smallest extent fastest, the slowest axis in
    // blocks of up to 128 rows (one block when it is shorter).  Neighbours
    // along the middle axis then lie one block height of keys apart (the
    // permutation and the next-step histogram give them windows of their
    // own), and only moves across a block edge land far away, where each
    // takes a device atomic of its own.  At 1000x1000x22 cells, n = 20:
    // 16-row blocks 9.4 ms per step, 128 rows 8.55 ms; 256 and 1024 rows were
    // as fast, but put the middle-axis neighbours beyond what the L2 keeps
    // (k_collide fetched 18.6 GB per launch instead of 13.9).
    uint32_t ord[3] = {0, 1, 2};
    std::stable_sort(ord, ord + 3, [&](uint32_t x, uint32_t y) { return k.dim[x] < k.dim[y]; });
    for (int a = 0; a < 3; ++a) k.ord[a] = ord[a];
    k.D0 = k.dim[ord[0]];
    k.D1 = k.dim[ord[1]];
    k.D2 = k.dim[ord[2]];
    k.Bsh = 0;
    while ((1u << k.Bsh) < k.D2 && k.Bsh < 7) ++k.Bsh;
    k.nblk = (k.D2 + (1u << k.Bsh) - 1) >> k.Bsh;
    k.Blast = k.D2 - ((k.nblk - 1) << k.Bsh);
    // one block: every key is below C, so key / C = 0 selects it (2^Bsh D1 D0
    // could pass 2^32 there); several blocks: D2 > 128, so 128 D1 D0 < C
    k.BLK = k.nblk == 1 ? k.C : (1u << k.Bsh) * k.D1 * k.D0;
    udiv_magic(k.BLK, k.mBLK);
    udiv_magic(k.D0, k.mD0);
    udiv_magic(k.Blast, k.mBlast);
    k.perio_z = p.periodicity[2];
    k.channel = p.experiment == 1;
    k.r2_init = (p.volume_size[0] - 2) * (p.volume_size[0] - 2) * 0.25f;
    k.r2_chan = (p.volume_size[2] - 2) * (p.volume_size[2] - 2) * 0.25f;
    k.dt = p.delta_t;
    k.drag_dt = p.drag * p.delta_t;
    k.tv = p.thermal_velocity;
    k.sin_a = sinf((M_PI * static_cast<double>(p.rotation_angle)) / 180);
    k.cos_a = cosf((M_PI * static_cast<double>(p.rotation_angle)) / 180);
    k.n = p.n;
    k.cap = 4 * p.n;
    udiv_magic(p.n, k.mn);
    k.rs.k0 = static_cast<uint32_t>(p.seed);
    k.rs.k1 = static_cast<uint32_t>(p.seed >> 32);
    k.rs.sys = p.system_id;
    k.N = static_cast<uint32_t>(p.N);
}

// fills c_axes on the current device (once per device)
void upload_axes(int device) {
    static std::mutex mu;
    static std::vector<int> done;
    std::lock_guard<std::mutex> g(mu);
    if (std::find(done.begin(), done.end(), device) != done.end()) return;
    float axes[13][3];
    for (int sel = 0; sel < 13; ++sel) {
        int const steps = 4;
        int const phi_i = sel % steps + 1, theta_i = sel / steps + 1;
        float theta;
        if ((theta_i % 2) && (phi_i % 2))
            theta = theta_i == 1 ? 0.95531661812450927816f : float(M_PI) - 0.95531661812450927816f;
        else
            theta = theta_i * (float(M_PI) / steps);
        float const phi = phi_i * (float(M_PI) / steps);
        axes[sel][0] = sinf(theta) * cosf(phi);
        axes[sel][1] = sinf(theta) * sinf(phi);
        axes[sel][2] = cosf(theta);
    }
    ck(hipMemcpyToSymbol(HIP_SYMBOL(c_axes), axes, sizeof(axes)), "upload direction table");
    done.push_back(device);
}

// LDS tile capacities compiled into k_collide (particles per tile)
constexpr uint32_t TILE_CAPS[7] = {256, 384, 448, 512, 640, 1152, 2048};
constexpr uint32_t TILE_CAP_MAX = TILE_CAPS[6];

// Cells per tile of every system for the LDS capacity `cap` (0: keep) and the
// request `tc` (0: the largest that fits).  SRD: on average about 80% of the
// tile's particle capacity; extended: the ghosts of a batch (at most n per
// cell) must fit the cap/2 ghost region.
// The compiled capacity a request of `cap` particles runs with (0: the
// default): the smallest at least as large that the rule can use (the
// extended rule needs 512 or more and room for one wall cell's n ghosts in its
// ghost region).
uint32_t pick_capacity(int algorithm, uint32_t nmax, uint32_t cap, uint64_t cells = 0) {
    // default: SRD 384-particle batches on two-wave workgroups (thirteen per
    // CU); below 16 particles per cell in large grids 512 (more cells per
    // batch: 1000x500x20 cells, n = 5: 64 cells of 512 1.20 ms per collide
    // launch, 24 of 384 1.43; n = 10: 51 of 512 1.94, 24 of 384 2.02; small
    // grids keep 384, whose more tiles fill the GPU better: 64x64x50 cells,
    // n = 10, 21.2 ms per 200-step loss against 24.1 with 512); the extended
    // rule 512; larger where the density makes single cells of several
    // hundred particles likely
    bool const sparse_large = nmax < 16 && cells >= (uint64_t(1) << 21);
    if (cap == 0) cap = std::max(algorithm == 0 && !sparse_large ? 384u : 512u, 6 * nmax);
    for (uint32_t c : TILE_CAPS)
        if (c >= cap && (algorithm == 0 || (c >= 512 && static_cast<uint32_t>(ghost_slots(static_cast<int>(c))) >= nmax)))
            return c;
    return TILE_CAP_MAX;
}

// Capacities the autotuner tries: every compiled one the rule can run whose
// batch holds a whole collision table (4n of the densest system).  With a
// smaller one, a cell fuller than the batch could not collide in the big-cell
// launch either, and the step would fail.  When no capacity holds the table
// (n above 512), only `current`.
std::vector<uint32_t> tune_capacities(int algorithm, uint32_t nmax, uint32_t current) {
    std::vector<uint32_t> out;
    for (uint32_t c : TILE_CAPS) {
        uint32_t const chosen = pick_capacity(algorithm, nmax, c);
        if (chosen >= 4 * nmax && std::find(out.begin(), out.end(), chosen) == out.end()) out.push_back(chosen);
    }
    if (out.empty()) out.push_back(current);
    return out;
}

// Tile shape of system s for the LDS capacity request `cap` (0: keep the
// system's, or its density's default, pick_capacity) and cells-per-tile
// request `tc` (0: about one batch).
void choose_tiling_sys(mpcd_hip_s* h, uint32_t s, uint32_t tc, uint32_t cap) {
    KParams& k = h->kp[s];
    uint32_t const n = std::max<uint32_t>(1, k.n);
    if (h->algorithm == 1 && TILE_CAP_MAX / 2 < n) throw HipError{"the extended rule supports n up to 1024"};
    if (cap == 0) cap = k.lds_cap;
    uint32_t const chosen = pick_capacity(h->algorithm, n, cap, k.C);
    k.lds_cap = chosen;
    k.tile_req = tc;
    // cells per tile: about one batch of particles by default; a tile whose
    // cells hold more (also on an explicit request) is processed in several batches
    uint32_t const t = tc ? tc : std::max<uint32_t>(1, (chosen + n / 2) / n);
    k.tile_cells = std::min<uint32_t>(t, static_cast<uint32_t>(tile_cell_max(static_cast<int>(chosen), h->algorithm)));
}

// the same request for every system
void choose_tiling(mpcd_hip_s* h, uint32_t tc, uint32_t cap) {
    for (uint32_t s = 0; s < h->ns; ++s) choose_tiling_sys(h, s, tc, cap);
}

void drop_graphs(mpcd_hip_s* h) {
    for (auto& plane : h->graph)
        for (auto& row : plane)
            for (auto& g : row)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
}

// the system table with the step state of h->step_index
void upload_table(mpcd_hip_s* h) {
    for (KParams& k : h->kp) {
        k.step = static_cast<uint32_t>(h->step_index);
        grid_shift(k.rs, k.step, k.shift);
        grid_shift(k.rs, k.step + 1, k.shift_next);
        // a step with injected draws (single system) uses the caller's shift
        if (h->inject)
            for (int a = 0; a < 3; ++a) k.shift[a] = h->inject->shift[a];
    }
    ck(hipStreamSynchronize(h->stream), "sync");
    ck(hipMemcpy(h->d_kp, h->kp.data(), sizeof(KParams) * h->ns, hipMemcpyHostToDevice), "upload system table");
}

// places every system in the shared arrays and builds the block -> system maps
void layout(mpcd_hip_s* h) {
    uint32_t const ns = h->ns;
    std::vector<uint32_t> blk(4 * (ns + 1), 0u);
    uint32_t* bp = blk.data();
    uint32_t* bc = bp + ns + 1;
    uint32_t* bw = bc + ns + 1;
    uint32_t* bt = bw + ns + 1;
    auto blocks = [](uint64_t items) { return static_cast<uint32_t>(std::min<uint64_t>((items + WG - 1) / WG, 65536)); };
    uint64_t cb = 0, pb = 0, wb = 0, gvb = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        KParams& k = h->kp[s];
        k.cb = static_cast<uint32_t>(cb);
        k.pb = static_cast<uint32_t>(pb);
        k.wb = static_cast<uint32_t>(wb);
        k.gvb = gvb;
        uint64_t const walls = k.perio_z ? 0 : 2ull * k.dim[0] * k.dim[1];
        bp[s + 1] = bp[s] + blocks(k.N);
        bc[s + 1] = bc[s] + blocks(k.C);
        bw[s + 1] = bw[s] + blocks(walls);
        cb += k.C;
        pb += k.N;
        wb += walls;
        gvb += (h->algorithm == 1 ? 6ull * k.n : 3ull) * walls;   // floats per wall cell
    }
    // collide tiles: the systems of each compiled capacity together (one
    // launch per capacity), bt[s] = first tile of system s, bt[ns] = all tiles
    h->ngrp = 0;
    h->lds_cap = 0;
    uint64_t tiles = 0;
    for (uint32_t gid = 0; gid < 2; ++gid)
        for (uint32_t cap : TILE_CAPS) {
            uint64_t const t0 = tiles;
            for (uint32_t s = 0; s < ns; ++s) {
                KParams const& k = h->kp[s];
                if (k.lds_cap != cap || k.gid != gid) continue;
                bt[s] = static_cast<uint32_t>(tiles);
                tiles += (k.C + k.tile_cells - 1) / k.tile_cells;
                if (tiles >= 0xffffffffull) throw HipError{"too many collide tiles"};
            }
            if (tiles > t0) {
                h->grp[h->ngrp++] = mpcd_hip_s::Group{cap, static_cast<uint32_t>(t0), static_cast<uint32_t>(tiles), gid};
                h->lds_cap = std::max(h->lds_cap, cap);
            }
        }
    bt[ns] = static_cast<uint32_t>(tiles);
    h->blocks_part = bp[ns];
    h->blocks_cell = bc[ns];
    h->blocks_wall = bw[ns];
    h->ntiles = bt[ns];
    h->W = wb;
    h->GV = gvb;
    ck(hipStreamSynchronize(h->stream), "sync");
    if (!h->d_blk) ck(hipMalloc(&h->d_blk, 4 * blk.size()), "alloc block maps");
    ck(hipMemcpy(h->d_blk, blk.data(), 4 * blk.size(), hipMemcpyHostToDevice), "upload block maps");
    if (ns > 1) {
        std::vector<uint16_t> ts(h->ntiles);
        for (uint32_t s = 0; s < ns; ++s) {
            KParams const& k = h->kp[s];
            uint32_t const nt = (k.C + k.tile_cells - 1) / k.tile_cells;
            for (uint32_t t = bt[s]; t < bt[s] + nt; ++t) ts[t] = static_cast<uint16_t>(s);
        }
        if (h->tile_sys_cap < ts.size()) {
            if (h->d_tile_sys) ck(hipFree(h->d_tile_sys), "free tile map");
            h->d_tile_sys = nullptr;
            ck(hipMalloc(&h->d_tile_sys, 2 * ts.size()), "alloc tile map");
            h->tile_sys_cap = ts.size();
        }
        ck(hipMemcpy(h->d_tile_sys, ts.data(), 2 * ts.size(), hipMemcpyHostToDevice), "upload tile map");
    }
    h->sys = Sys{h->d_kp, h->d_blk, h->d_blk + (ns + 1), h->d_blk + 2 * (ns + 1), h->d_tile_sys,
                 h->d_blk + 3 * (ns + 1), ns, static_cast<uint32_t>(h->C), h->N};
    drop_graphs(h);
    upload_table(h);
}

void set_err(mpcd_hip_s* h, std::string const& m) {
    if (h) h->err = m;
}

template <typename F>
int guarded(mpcd_hip_s* h, F&& f) {
    if (!h) return 1;
    if (h->streamed_pending) {
        // a step() call failed after writing an intermediate state: the
        // post-collision state the API describes is gone
        set_err(h, "the particle state was lost when an earlier step failed; destroy this handle");
        return 3;
    }
    try {
        f();
        return 0;
    } catch (HipError const& e) {
        set_err(h, e.msg);
    } catch (std::exception const& e) {
        set_err(h, e.what());
    }
    return 2;
}

// blocks of the step's per-cell draws: wall ghosts, and the SRD axes of a
// step that keeps them in memory (other steps, and the extended rule, draw
// their axes in the collision)
uint32_t draw_blocks(mpcd_hip_s const* h) {
    return h->blocks_wall + (h->algorithm == 0 && h->axes_mem ? h->blocks_cell : 0u);
}

// the per-cell draws of the step in the system table alone (k_cell_draws),
// for a step that needs no sort
int launch_draws(mpcd_hip_s* h) {
    uint32_t const nb = draw_blocks(h);
    if (!nb) return 0;
    hipLaunchKernelGGL(k_cell_draws, dim3(nb), dim3(WG), 0, h->stream, h->sys, h->axis, h->wgcnt, h->wgvel,
                       h->blocks_wall, h->algorithm);
    ck(hipGetLastError(), "cell draws");
    return 1;
}

void launch_keys(mpcd_hip_s* h) {
    ck(hipMemsetAsync(h->count, 0, 4 * h->C, h->stream), "clear counts");
    if (h->blocks_part)
        hipLaunchKernelGGL(k_keys, dim3(h->blocks_part), dim3(WG), 0, h->stream, h->sys, h->st[h->cur], h->count, h->keys);
    ck(hipGetLastError(), "k_keys");
    h->keys_valid = true;
    h->sorted_valid = false;
}

// After the cells were counted from scratch (initial state, set_particles,
// set_step_index): a cell fuller than one LDS batch would be streamed without
// collision and fail the step, so the batch capacity is raised to fit it when
// a compiled one does (up to 2048 particles; never lowered again here).
void fit_capacity(mpcd_hip_s* h) {
    ck(hipMemsetAsync(h->cmax, 0, 4, h->stream), "clear");
    uint32_t const blocks = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((h->C + WG - 1) / WG, 1024)));
    hipLaunchKernelGGL(k_count_max, dim3(blocks), dim3(WG), 0, h->stream, h->count, static_cast<uint32_t>(h->C), h->cmax);
    ck(hipGetLastError(), "k_count_max");
    uint32_t m = 0;
    ck(hipMemcpyAsync(&m, h->cmax, 4, hipMemcpyDeviceToHost, h->stream), "count max");
    ck(hipStreamSynchronize(h->stream), "count max");
    bool raised = false;
    for (uint32_t s = 0; s < h->ns; ++s)
        if (m > h->kp[s].lds_cap && m <= TILE_CAP_MAX) {
            choose_tiling_sys(h, s, h->kp[s].tile_req, m);
            raised = true;
        }
    if (raised) layout(h);
}

void launch_scan(mpcd_hip_s* h) {
    uint32_t const C = static_cast<uint32_t>(h->C);
    int const fold = h->nb <= SCAN_FOLD_MAX ? 1 : 0;
    hipLaunchKernelGGL(k_scan_reduce, dim3(h->nb), dim3(WG), 0, h->stream, h->count, C, h->partial);
    if (!fold) hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, h->stream, h->partial, h->nb);
    hipLaunchKernelGGL(k_scan_down, dim3(h->nb), dim3(WG), 0, h->stream, h->count, C, h->partial, h->offset, h->cursor,
                       h->N, fold);
}

uint32_t perm_blocks(mpcd_hip_s const* h) {
    return static_cast<uint32_t>((h->N + WG * PERM_ITEMS - 1) / (WG * PERM_ITEMS));
}

// The per-cell draws ride in the permutation's launch when they are few (a
// launch less per step: 64x64x52 cells, n = 10, 16.8 us against 12.7 + 11.5
// with the axes drawn there); many of them run better on their own, at the
// occupancy of their small LDS footprint rather than the permutation's
// (1000x1000x22 cells, SRD wall ghosts alone, 7813 blocks: 1.18 ms against
// 1.06 + 0.03; the 64-channel ensemble 12.97 against 12.58 ms per step)
constexpr uint32_t DRAWS_FOLD_MAX = 1024;   // draw blocks
bool fold_draws(mpcd_hip_s const* h) { return draw_blocks(h) <= DRAWS_FOLD_MAX; }

// the sort of the step (scan and permutation); with `draws` the step's
// per-cell draws ride in the permutation's launch
int launch_sort(mpcd_hip_s* h, bool draws = false) {
    launch_scan(h);
    uint32_t const np = perm_blocks(h), nd = draws ? draw_blocks(h) : 0u;
    if (nd)
        hipLaunchKernelGGL(k_perm_draws, dim3(np + nd), dim3(WG), 0, h->stream, h->sys, h->keys, h->cursor, h->perm,
                           h->N, np, h->axis, h->wgcnt, h->wgvel, h->blocks_wall, h->algorithm);
    else if (np)
        hipLaunchKernelGGL(k_perm, dim3(np), dim3(WG), 0, h->stream, h->sys, h->keys, h->cursor, h->perm, h->N);
    ck(hipGetLastError(), "sort");
    return 4;
}


// The collide launch: one workgroup per tile; BIG: the step's second launch,
// BIG_BLOCKS workgroups over the cells the first listed as larger than their
// tile (usually none: every workgroup returns at once)
template <int ALGO, bool SAMPLE, int CAP, bool ADJ = false, bool BIG = false>
__global__ __launch_bounds__(tile_threads(CAP), tile_blocks(CAP, ALGO)) void k_collide(
    Sys sys, const float* __restrict__ in, float* __restrict__ out, const uint32_t* __restrict__ perm,
    const uint32_t* __restrict__ offset, uint32_t* __restrict__ count_next, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ perm_out, SampleBufs S, uint32_t* __restrict__ status, float const* __restrict__ axis,
    uint32_t const* __restrict__ gcnt, float const* __restrict__ gvel, Tape T, uint32_t io, uint32_t toff) {
    if constexpr (BIG) {
        uint32_t const n = status[1];
        for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
            collide_tile<ALGO, SAMPLE, CAP, false, true>(sys, in, out, perm, offset, count_next, keys, perm_out, S,
                                                         status, axis, gcnt, gvel, T, io, e);
            __syncthreads();
        }
    } else {
        // tiles [toff, toff + gridDim.x) of the handle's tile space
        collide_tile<ALGO, SAMPLE, CAP, ADJ, false>(sys, in, out, perm, offset, count_next, keys, perm_out, S, status,
                                                    axis, gcnt, gvel, T, io, toff + xcd_tile(blockIdx.x, gridDim.x));
    }
}

// one collide launch: tiles [t0, t1) (BIG: the big-cell launch, BIG_BLOCKS
// workgroups over the cells listed as larger than their tile)
template <int ALGO, bool SAMPLE, int CAP, bool ADJ = false, bool BIG = false>
void launch_collide_t(mpcd_hip_s* h, uint32_t* rec, Tape const& T, uint32_t io, uint32_t t0, uint32_t t1) {
    SampleBufs S{h->scnt, h->ssum};
    uint32_t const grid = BIG ? BIG_BLOCKS : t1 - t0;
    if (!grid) return;
    hipLaunchKernelGGL((k_collide<ALGO, SAMPLE, CAP, ADJ, BIG>), dim3(grid), dim3(tile_threads(CAP)), 0, h->stream,
                       h->sys, h->st[h->cur], h->st[1 - h->cur], h->perm, h->offset, h->count, h->keys, rec, S,
                       h->status, h->axes_mem ? h->axis : nullptr, h->wgcnt, h->wgvel, T, io, t0);
}

template <int ALGO, bool SAMPLE, bool BIG>
void launch_collide_at(mpcd_hip_s* h, uint32_t cap, uint32_t* rec, uint32_t io, uint32_t t0, uint32_t t1) {
    switch (cap) {
        case 256:
            if constexpr (ALGO == 0) launch_collide_t<ALGO, SAMPLE, 256, false, BIG>(h, rec, Tape{}, io, t0, t1);
            else throw HipError{"the extended rule needs an LDS capacity of at least 512"};
            break;
        case 384:
            if constexpr (ALGO == 0) launch_collide_t<ALGO, SAMPLE, 384, false, BIG>(h, rec, Tape{}, io, t0, t1);
            else throw HipError{"the extended rule needs an LDS capacity of at least 512"};
            break;
        case 448:
            if constexpr (ALGO == 0) launch_collide_t<ALGO, SAMPLE, 448, false, BIG>(h, rec, Tape{}, io, t0, t1);
            else throw HipError{"the extended rule needs an LDS capacity of at least 512"};
            break;
        case 512: launch_collide_t<ALGO, SAMPLE, 512, false, BIG>(h, rec, Tape{}, io, t0, t1); break;
        case 640:
            if constexpr (ALGO == 0) launch_collide_t<ALGO, SAMPLE, 640, false, BIG>(h, rec, Tape{}, io, t0, t1);
            else launch_collide_t<ALGO, SAMPLE, 1152, false, BIG>(h, rec, Tape{}, io, t0, t1);
            break;
        case 1152: launch_collide_t<ALGO, SAMPLE, 1152, false, BIG>(h, rec, Tape{}, io, t0, t1); break;
        default: launch_collide_t<ALGO, SAMPLE, 2048, false, BIG>(h, rec, Tape{}, io, t0, t1); break;
    }
}

// The collide launches of a step: one per group of systems that share a
// compiled LDS capacity, then the big-cell launch at the largest capacity
// (cells the first launches listed as larger than their tile; usually none,
// every workgroup returns at once)
template <int ALGO, bool SAMPLE>
void launch_collide_cap(mpcd_hip_s* h, uint32_t* rec, uint32_t io) {
    for (uint32_t g = 0; g < h->ngrp; ++g) {
        if (h->group_timing) ck(hipEventRecord(h->gev[2 * g], h->stream), "event");
        launch_collide_at<ALGO, SAMPLE, false>(h, h->grp[g].cap, rec, io, h->grp[g].t0, h->grp[g].t1);
        if (h->group_timing) ck(hipEventRecord(h->gev[2 * g + 1], h->stream), "event");
    }
    launch_collide_at<ALGO, SAMPLE, true>(h, h->lds_cap, rec, io, 0, 0);
}

void ensure_sampling(mpcd_hip_s* h) {
    if (h->acc) return;
    ck(hipMalloc(&h->scnt, 4 * h->C), "alloc sample counts");
    ck(hipMalloc(&h->ssum, 8 * 3 * h->C), "alloc sample sums");
    ck(hipMalloc(&h->acc, 8 * 4 * h->C), "alloc accumulators");
    ck(hipMemsetAsync(h->scnt, 0, 4 * h->C, h->stream), "clear");
    ck(hipMemsetAsync(h->ssum, 0, 8 * 3 * h->C, h->stream), "clear");
    ck(hipMemsetAsync(h->acc, 0, 8 * 4 * h->C, h->stream), "clear");
}

// The launches of one step on h->stream (also what a captured step graph
// holds): sort, per-cell draws, collide, sampling fold, step-state advance.
// Uses no memset, so the sequence can be captured and replayed.
// io: bit 0 the input buffer holds this step's streamed state, bit 1 write the
// next step's streamed state (see k_collide)
int enqueue_step(mpcd_hip_s* h, bool sample, bool first, bool sort, uint32_t* rec, hipEvent_t e0, hipEvent_t e1,
                 uint32_t io) {
    int launches = 0;
    bool const fold = sort && !h->inject && fold_draws(h);
    if (sort) launches += launch_sort(h, fold);
    if (h->inject) {
        // the caller's draws in place of this step's (pageable host memory:
        // the copies complete before the collide launch reads them)
        auto const& d = *h->inject;
        if (!d.axes.empty())
            ck(hipMemcpyAsync(h->axis, d.axes.data(), 4 * d.axes.size(), hipMemcpyHostToDevice, h->stream),
               "injected axes");
        if (!d.gcnt.empty()) {
            ck(hipMemcpyAsync(h->wgcnt, d.gcnt.data(), 4 * d.gcnt.size(), hipMemcpyHostToDevice, h->stream),
               "injected ghost counts");
            ck(hipMemcpyAsync(h->wgvel, d.gvel.data(), 4 * d.gvel.size(), hipMemcpyHostToDevice, h->stream),
               "injected ghosts");
        }
    } else if (!fold) {
        launches += launch_draws(h);
    }
    if (e0) ck(hipEventRecord(e0, h->stream), "event");
    if (h->algorithm == 0) {
        if (sample) launch_collide_cap<0, true>(h, rec, io); else launch_collide_cap<0, false>(h, rec, io);
    } else {
        if (sample) launch_collide_cap<1, true>(h, rec, io); else launch_collide_cap<1, false>(h, rec, io);
    }
    ck(hipGetLastError(), "k_collide");
    if (e1) ck(hipEventRecord(e1, h->stream), "event");
    launches += 1;
    if (sample) {
        uint64_t const blocks = std::max<uint64_t>(1, std::min<uint64_t>((h->C + WG - 1) / WG, 1u << 16));
        hipLaunchKernelGGL(k_sample_fold, dim3(static_cast<uint32_t>(blocks)), dim3(WG), 0, h->stream,
                           static_cast<uint32_t>(h->C), h->scnt, h->ssum, h->acc, first ? 1 : 0);
        ck(hipGetLastError(), "k_sample_fold");
        launches += 1;
    }
    hipLaunchKernelGGL(k_advance, dim3((h->ns + 63) / 64), dim3(64), 0, h->stream, h->sys, h->status + 1);
    ck(hipGetLastError(), "k_advance");
    return launches + 1;
}

// the captured step reading h->st[h->cur] (captured on first use)
hipGraphExec_t step_graph(mpcd_hip_s* h, bool sample, uint32_t io) {
    hipGraphExec_t& ge = h->graph[h->cur][sample ? 1 : 0][io & 3u];
    if (ge) return ge;
    hipGraph_t g = nullptr;
    ck(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    int n = 0;
    try {
        n = enqueue_step(h, sample, false, true, nullptr, nullptr, nullptr, io);
    } catch (...) {
        (void)hipStreamEndCapture(h->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    ck(hipStreamEndCapture(h->stream, &g), "end capture");
    hipError_t const e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    ck(e, "instantiate step graph");
    h->graph_launches[sample ? 1 : 0] = n;
    return ge;
}

void run_steps(mpcd_hip_s* h, int n_steps, bool sample) {
    if (n_steps <= 0) return;
    if (sample) ensure_sampling(h);
    if (h->record && !h->perm_rec) ck(hipMalloc(&h->perm_rec, 4 * std::max<uint64_t>(h->N, 1)), "alloc perm record");
    if (h->timing) {
        if (h->ev.size() < 2 * static_cast<size_t>(n_steps)) {
            size_t const old = h->ev.size();
            h->ev.resize(2 * static_cast<size_t>(n_steps));
            for (size_t i = old; i < h->ev.size(); ++i) ck(hipEventCreate(&h->ev[i]), "event");
        }
        if (!h->ev_begin) {
            ck(hipEventCreate(&h->ev_begin), "event");
            ck(hipEventCreate(&h->ev_end), "event");
        }
        ck(hipEventRecord(h->ev_begin, h->stream), "event");
    }
    h->launches = 0;
    for (int s = 0; s < n_steps; ++s) {
        if (!h->keys_valid) {
            launch_keys(h);
            h->launches += 2;
            fit_capacity(h);
        }
        bool const first = sample && h->sampling_done;
        bool const graphed = h->use_graph && !h->timing && !h->record && !first && !h->sorted_valid && !h->group_timing &&
                             !h->inject;
        // inside one call the intermediate states stay streamed; the last step
        // writes the post-collision state the API reads
        uint32_t const io = (s > 0 ? 1u : 0u) | (s + 1 < n_steps ? 2u : 0u);
        if (io & 2u) h->streamed_pending = true;
        if (graphed) {
            ck(hipGraphLaunch(step_graph(h, sample, io), h->stream), "launch step graph");
            h->launches += h->graph_launches[sample ? 1 : 0];
        } else {
            hipEvent_t const e0 = h->timing ? h->ev[2 * s] : nullptr;
            hipEvent_t const e1 = h->timing ? h->ev[2 * s + 1] : nullptr;
            h->launches += enqueue_step(h, sample, first, !h->sorted_valid, h->record ? h->perm_rec : nullptr, e0, e1, io);
            if (h->group_timing) {
                ck(hipStreamSynchronize(h->stream), "step");
                for (uint32_t g = 0; g < h->ngrp; ++g) {
                    float ms = 0;
                    ck(hipEventElapsedTime(&ms, h->gev[2 * g], h->gev[2 * g + 1]), "elapsed");
                    h->group_ms[g] += ms;
                }
            }
        }
        if (!(io & 2u)) h->streamed_pending = false;   // the post-collision state is written again
        if (sample) {
            if (first) {
                h->sampling_done = false;
                h->sample_counter = 0;
            } else {
                h->sample_counter++;
            }
        }
        h->cur = 1 - h->cur;
        h->step_index++;
        h->keys_valid = true;
        h->sorted_valid = false;
        h->recorded = h->record;
    }
    if (h->timing) ck(hipEventRecord(h->ev_end, h->stream), "event");
    ck(hipStreamSynchronize(h->stream), "step");
    uint32_t st = 0;
    ck(hipMemcpy(&st, h->status, 4, hipMemcpyDeviceToHost), "status");
    if (h->timing) {
        float ms = 0;
        h->t_collide = 0;
        for (int s = 0; s < n_steps; ++s) {
            ck(hipEventElapsedTime(&ms, h->ev[2 * s], h->ev[2 * s + 1]), "elapsed");
            h->t_collide += ms;
        }
        ck(hipEventElapsedTime(&ms, h->ev_begin, h->ev_end), "elapsed");
        h->t_total = ms;
    }
    if (st & 1u) {
        ck(hipMemsetAsync(h->status, 0, 4, h->stream), "status");
        ck(hipStreamSynchronize(h->stream), "status");
        throw HipError{"a cell holds more particles than one LDS tile (" + std::to_string(h->lds_cap) +
                       ") and its collision table (4n) does not fit the tile either, or a taped run met such a "
                       "cell; raise the tile capacity with mpcd_hip_set_tiling"};
    }
}

void check_particles(mpcd_hip_s* h, size_t n) {
    if (n != h->N)
        throw HipError{"particle array holds " + std::to_string(n) + " particles, the system has " + std::to_string(h->N)};
}

void check_sys(mpcd_hip_s const* h, uint32_t s) {
    if (s >= h->ns) throw HipError{"system index " + std::to_string(s) + " out of range (" + std::to_string(h->ns) + " systems)"};
}

// xyz-interleaved host array <-> three fields of the particle records, slots
// [first, first + n); moved in chunks to bound the host buffer
constexpr uint64_t XFER_CHUNK = uint64_t(1) << 22;

void put_component(mpcd_hip_s* h, const float* src, int first_field) {
    std::vector<float> tmp(6 * std::min<uint64_t>(h->N, XFER_CHUNK));
    ck(hipStreamSynchronize(h->stream), "sync");
    for (uint64_t c0 = 0; c0 < h->N; c0 += XFER_CHUNK) {
        uint64_t const n = std::min<uint64_t>(XFER_CHUNK, h->N - c0);
        float* const d = h->st[h->cur] + rec(c0);
        ck(hipMemcpy(tmp.data(), d, 24 * n, hipMemcpyDeviceToHost), "download");
        for (uint64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) tmp[rec(i) + first_field + a] = src[3 * (c0 + i) + a];
        ck(hipMemcpy(d, tmp.data(), 24 * n, hipMemcpyHostToDevice), "upload");
    }
    h->keys_valid = false;
    h->sorted_valid = false;
}

void get_component(mpcd_hip_s* h, float* dst, int first_field, uint64_t first, uint64_t n_all) {
    std::vector<float> tmp(6 * std::min<uint64_t>(std::max<uint64_t>(n_all, 1), XFER_CHUNK));
    ck(hipStreamSynchronize(h->stream), "sync");
    for (uint64_t c0 = 0; c0 < n_all; c0 += XFER_CHUNK) {
        uint64_t const n = std::min<uint64_t>(XFER_CHUNK, n_all - c0);
        ck(hipMemcpy(tmp.data(), h->st[h->cur] + rec(first + c0), 24 * n, hipMemcpyDeviceToHost), "download");
        for (uint64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) dst[3 * (c0 + i) + a] = tmp[rec(i) + first_field + a];
    }
}

void adj_free(mpcd_hip_s* h);

void destroy(mpcd_hip_s* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    adj_free(h);
    for (auto* p : {h->st[0], h->st[1]}) if (p) (void)hipFree(p);
    for (void* p : {static_cast<void*>(h->perm), static_cast<void*>(h->keys), static_cast<void*>(h->cursor), static_cast<void*>(h->count),
                    static_cast<void*>(h->offset), static_cast<void*>(h->partial), static_cast<void*>(h->status),
                    static_cast<void*>(h->cmax),
                    static_cast<void*>(h->perm_rec), static_cast<void*>(h->scnt), static_cast<void*>(h->ssum),
                    static_cast<void*>(h->axis), static_cast<void*>(h->wgcnt), static_cast<void*>(h->wgvel),
                    static_cast<void*>(h->acc), static_cast<void*>(h->d_kp), static_cast<void*>(h->d_blk),
                    static_cast<void*>(h->d_tile_sys)})
        if (p) (void)hipFree(p);
    for (auto e : h->ev) (void)hipEventDestroy(e);
    if (h->ev_begin) (void)hipEventDestroy(h->ev_begin);
    if (h->ev_end) (void)hipEventDestroy(h->ev_end);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// ------------------------------------------------------------ adjoint runs

void adj_free(mpcd_hip_s* h) {
    auto& a = h->adj;
    if (a.recompute) {
        (void)hipStreamSynchronize(a.recompute);
        (void)hipStreamDestroy(a.recompute);
    }
    for (hipEvent_t e : {a.written[0], a.written[1], a.read[0], a.read[1]})
        if (e) (void)hipEventDestroy(e);
    for (void* p : {static_cast<void*>(a.perm), static_cast<void*>(a.vp), static_cast<void*>(a.lz),
                    static_cast<void*>(a.off), static_cast<void*>(a.u), static_cast<void*>(a.ax),
                    static_cast<void*>(a.tot), static_cast<void*>(a.ckpt), static_cast<void*>(a.lam[0]),
                    static_cast<void*>(a.lam[1]), static_cast<void*>(a.lcnt), static_cast<void*>(a.lsum),
                    static_cast<void*>(a.stab), static_cast<void*>(a.part), static_cast<void*>(a.gf),
                    static_cast<void*>(a.force)})
        if (p) (void)hipFree(p);
    a = mpcd_hip_s::Adj{};
}

template <typename T>
void dalloc(T*& p, uint64_t count, const char* what) {
    ck(hipMalloc(&p, sizeof(T) * std::max<uint64_t>(count, 1)), what);
}

// buffers for `steps` steps with checkpoints every K steps (K = 0: forward
// only): K taped steps (2 K with more than one segment), and the start state
// of every segment (those of the last two, which the forward pass tapes, go
// unused)
void adj_reserve(mpcd_hip_s* h, uint32_t steps, uint32_t K) {
    auto& a = h->adj;
    uint32_t const nck = K ? (steps + K - 1) / K : 1;
    uint32_t const halves = nck > 1 ? 2u : 1u;
    if (a.ckpt && a.K >= K && a.nck >= nck && a.steps >= steps && a.halves >= halves) return;
    ck(hipStreamSynchronize(h->stream), "sync");
    adj_free(h);
    uint64_t const N = h->N, C = h->C, Z = h->kp[0].dim[2];
    dalloc(a.ckpt, 6 * N * nck, "alloc checkpoints");
    dalloc(a.lcnt, uint64_t(steps) * LAYER_SLOTS * Z, "alloc layer counts");
    dalloc(a.lsum, uint64_t(steps) * 2 * LAYER_SLOTS * Z, "alloc layer sums");
    dalloc(a.force, C, "alloc force field");
    if (K) {
        uint64_t const S = uint64_t(K) * halves;   // tape slots
        dalloc(a.stab, uint64_t(steps) * Z, "alloc loss seeds");
        dalloc(a.perm, S * N, "alloc tape");
        dalloc(a.vp, S * 3 * N, "alloc tape");
        dalloc(a.lz, S * N, "alloc tape");
        dalloc(a.off, S * (C + 1), "alloc tape");
        dalloc(a.u, S * 3 * C, "alloc tape");
        dalloc(a.ax, S * 3 * C, "alloc tape");
        dalloc(a.tot, S * C, "alloc tape");
        dalloc(a.lam[0], 3 * N, "alloc adjoint");
        dalloc(a.lam[1], 3 * N, "alloc adjoint");
        dalloc(a.part, 2 * ((C + ADJ_CELLS - 1) / ADJ_CELLS), "alloc gradient partials");
        dalloc(a.gf, C, "alloc force gradient");
    }
    if (halves > 1) {
        ck(hipStreamCreateWithFlags(&a.recompute, hipStreamNonBlocking), "recompute stream");
        for (int i = 0; i < 2; ++i) {
            ck(hipEventCreateWithFlags(&a.written[i], hipEventDisableTiming), "event");
            ck(hipEventCreateWithFlags(&a.read[i], hipEventDisableTiming), "event");
        }
    }
    a.K = K;
    a.nck = nck;
    a.steps = steps;
    a.halves = halves;
}

// blocks of k_layers: one pass of WG * LAYER_UNROLL particles each (the
// kernel waits on its loads, so all of them go out at once) up to 1024 blocks,
// each flushing its histogram once with three atomics per layer (spread over
// LAYER_SLOTS copies: 64x64x54 cells, n = 10: 256 blocks of four passes
// 14.4 us per launch); no block sums more than LAYER_BLOCK_MAX particles (the
// exactness bound of its fixed-point words)
uint32_t layer_blocks(mpcd_hip_s* h) {
    uint64_t const b = std::max<uint64_t>(std::min<uint64_t>((h->N + WG * LAYER_UNROLL - 1) / (WG * LAYER_UNROLL), 1024),
                                          (h->N + LAYER_BLOCK_MAX - 1) / LAYER_BLOCK_MAX);
    return static_cast<uint32_t>(std::max<uint64_t>(1, b));
}

// one SRD step with the body force, then the layer statistics of its output
// into (cnt, sum) when given; records tape slot `j` when j >= 0: the step's
// sort writes its offsets and the draws their axes straight into the slot
void adj_step(mpcd_hip_s* h, float const* force, int j, uint32_t* cnt, unsigned long long* sum) {
    auto& a = h->adj;
    uint64_t const N = h->N, C = h->C;
    Tape T{};
    T.force = force;
    uint32_t* rec = nullptr;
    uint16_t* lz = nullptr;
    uint32_t* const own_offset = h->offset;
    float* const own_axis = h->axis;
    h->axes_mem = true;
    if (j >= 0) {
        T.vp = a.vp + uint64_t(j) * 3 * N;
        T.u = a.u + uint64_t(j) * 3 * C;
        T.tot = a.tot + uint64_t(j) * C;
        rec = a.perm + uint64_t(j) * N;
        lz = a.lz + uint64_t(j) * N;
        h->offset = a.off + uint64_t(j) * (C + 1);
        h->axis = a.ax + uint64_t(j) * 3 * C;
    }
    try {
        if (!h->keys_valid) launch_keys(h);
        bool const fold = !h->sorted_valid && fold_draws(h);
        if (!h->sorted_valid) launch_sort(h, fold);
        if (!fold) launch_draws(h);
        switch (h->lds_cap) {
            case 256: launch_collide_t<0, false, 256, true>(h, rec, T, 0, 0, h->ntiles); break;
            case 384: launch_collide_t<0, false, 384, true>(h, rec, T, 0, 0, h->ntiles); break;
            case 448: launch_collide_t<0, false, 448, true>(h, rec, T, 0, 0, h->ntiles); break;
            case 512: launch_collide_t<0, false, 512, true>(h, rec, T, 0, 0, h->ntiles); break;
            case 640: launch_collide_t<0, false, 640, true>(h, rec, T, 0, 0, h->ntiles); break;
            case 1152: launch_collide_t<0, false, 1152, true>(h, rec, T, 0, 0, h->ntiles); break;
            default: launch_collide_t<0, false, 2048, true>(h, rec, T, 0, 0, h->ntiles); break;
        }
        ck(hipGetLastError(), "adjoint-run step");
    } catch (...) {
        h->offset = own_offset;
        h->axis = own_axis;
        h->axes_mem = false;
        throw;
    }
    h->offset = own_offset;
    h->axis = own_axis;
    h->axes_mem = false;
    h->cur = 1 - h->cur;
    h->step_index++;
    h->keys_valid = true;
    h->sorted_valid = false;
    // the layer pass also moves the system to the next step
    hipLaunchKernelGGL(k_layers, dim3(layer_blocks(h)), dim3(WG), 0, h->stream, h->sys, h->st[h->cur], cnt, sum, lz,
                       h->status, h->status + 1);
}

struct ProfileResult {
    double loss = 0, grad_drag = 0, grad_angle = 0;
    std::vector<double> profile;
};

// Loss (and with K > 0 its pathwise gradient) of `steps` SRD steps from the
// current state with drag g, rotation angle and body force f:
//   L = mean over fluid z layers of (<vbar_x>(z) - target(z))^2,
//   <vbar_x>(z) = mean over the steps of the layer's mean x velocity after each step.
// The handle's state, step index and parameters are restored afterwards.
ProfileResult profile_run(mpcd_hip_s* h, int steps, int K, float drag, float angle, float const* force_host,
                          double const* target, size_t n_layers, float* grad_force) {
    if (h->ns != 1) throw HipError{"adjoint runs take a single system"};
    if (h->algorithm != 0) throw HipError{"adjoint runs implement the SRD rule"};
    if (h->kp[0].channel) throw HipError{"adjoint runs implement the standard geometry (walls in z)"};
    if (steps < 1) throw HipError{"steps must be positive"};
    if (K < 0) throw HipError{"checkpoint interval must be positive"};
    if (h->N >= BOUNCED) throw HipError{"adjoint runs take fewer than 2^31 particles"};
    KParams& P = h->kp[0];
    uint32_t const Z = P.dim[2];
    if (Z > static_cast<uint32_t>(LAYERS_MAX)) throw HipError{"adjoint runs take at most 1024 z layers"};
    uint32_t const zlo = P.perio_z ? 0u : 1u, zhi = P.perio_z ? Z : Z - 1;
    if (n_layers != zhi - zlo)
        throw HipError{"target holds " + std::to_string(n_layers) + " layers, the fluid has " + std::to_string(zhi - zlo)};
    uint32_t const T = static_cast<uint32_t>(steps);
    uint32_t const Kc = K ? std::min<uint32_t>(static_cast<uint32_t>(K), T) : 0u;
    adj_reserve(h, T, Kc);
    auto& a = h->adj;
    uint64_t const N = h->N, C = h->C;
    float const* force = nullptr;
    if (force_host) {
        ck(hipStreamSynchronize(h->stream), "sync");
        ck(hipMemcpy(a.force, force_host, 4 * C, hipMemcpyHostToDevice), "upload force field");
        force = a.force;
    }

    // save the start state and parameters, set the run's parameters
    uint64_t const s0 = h->step_index;
    KParams const saved = P;
    ck(hipMemcpyAsync(a.ckpt, h->st[h->cur], 24 * N, hipMemcpyDeviceToDevice, h->stream), "checkpoint");
    P.drag_dt = drag * P.dt;
    P.sin_a = sinf((M_PI * static_cast<double>(angle)) / 180);
    P.cos_a = cosf((M_PI * static_cast<double>(angle)) / 180);
    auto restore = [&]() {
        P = saved;
        ck(hipMemcpyAsync(h->st[h->cur], a.ckpt, 24 * N, hipMemcpyDeviceToDevice, h->stream), "restore");
        h->step_index = s0;
        h->keys_valid = false;
        h->sorted_valid = false;
        upload_table(h);
    };
    ProfileResult res;
    try {
        h->keys_valid = false;
        h->sorted_valid = false;
        upload_table(h);
        // forward pass: layer statistics of every step, the start state of
        // every segment but the last two, and the tape of the last segment
        // (half 0) and of the one before it (half 1), which the reverse pass
        // then need not recompute (the whole run when K >= steps or two
        // segments cover it: nothing is recomputed then)
        uint64_t const LZ = uint64_t(LAYER_SLOTS) * Z;
        ck(hipMemsetAsync(a.lcnt, 0, 4ull * T * LZ, h->stream), "clear");
        ck(hipMemsetAsync(a.lsum, 0, 16ull * T * LZ, h->stream), "clear");
        uint32_t const nseg = Kc ? (T + Kc - 1) / Kc : 1;
        uint32_t const tlast = (nseg - 1) * Kc;                   // steps before the last segment
        uint32_t const ttape = nseg > 1 ? (nseg - 2) * Kc : tlast;   // steps before the forward pass's tape
        for (uint32_t t = 1; t <= T; ++t) {
            int j = -1;
            if (Kc && t > tlast) j = static_cast<int>(t - tlast - 1);
            else if (Kc && t > ttape) j = static_cast<int>(Kc + (t - ttape - 1));
            adj_step(h, force, j, a.lcnt + uint64_t(t - 1) * LZ, a.lsum + uint64_t(t - 1) * 2 * LZ);
            if (Kc && t % Kc == 0 && t < ttape)
                ck(hipMemcpyAsync(a.ckpt + uint64_t(t / Kc) * 6 * N, h->st[h->cur], 24 * N, hipMemcpyDeviceToDevice,
                                  h->stream), "checkpoint");
        }
        ck(hipGetLastError(), "forward pass");
        std::vector<uint32_t> cnt(uint64_t(T) * LZ);
        std::vector<unsigned long long> sum(uint64_t(T) * 2 * LZ);
        ck(hipStreamSynchronize(h->stream), "forward pass");
        uint32_t st = 0;
        ck(hipMemcpy(&st, h->status, 4, hipMemcpyDeviceToHost), "status");
        if (st & 1u) {
            ck(hipMemset(h->status, 0, 4), "status");
            throw HipError{"a cell holds more particles than one LDS tile"};
        }
        if (st & STATUS_LAYER_VELOCITY) {
            ck(hipMemset(h->status, 0, 4), "status");
            throw HipError{"a particle's x velocity reached 128 (or is not a number): the profile loss sums the "
                           "layers' velocities in fixed point, exact only below that"};
        }
        ck(hipMemcpy(cnt.data(), a.lcnt, 4ull * T * LZ, hipMemcpyDeviceToHost), "download layers");
        ck(hipMemcpy(sum.data(), a.lsum, 16ull * T * LZ, hipMemcpyDeviceToHost), "download layers");
        // per step and layer: particle count, and the sum of rint(v 2^40) as
        // hi 2^20 + lo (both exact integers; the copies add up exactly)
        std::vector<uint32_t> ct(uint64_t(T) * Z, 0u);
        res.profile.assign(zhi - zlo, 0.0);
        for (uint32_t t = 0; t < T; ++t)
            for (uint32_t z = zlo; z < zhi; ++z) {
                uint32_t c = 0;
                long long hi = 0;
                unsigned long long lo = 0;
                for (uint32_t k = 0; k < LAYER_SLOTS; ++k) {
                    uint64_t const i = uint64_t(t) * LZ + uint64_t(k) * Z + z;
                    c += cnt[i];
                    hi += static_cast<long long>(sum[2 * uint64_t(t) * LZ + uint64_t(k) * Z + z]);
                    lo += sum[2 * uint64_t(t) * LZ + LZ + uint64_t(k) * Z + z];
                }
                ct[uint64_t(t) * Z + z] = c;
                double const m = c ? ((static_cast<double>(hi) * 1048576.0 + static_cast<double>(lo)) / FIX_SCALE) / c : 0.0;
                res.profile[z - zlo] += m / T;
            }
        std::vector<double> w(Z, 0.0);
        double const nl = static_cast<double>(zhi - zlo);
        for (uint32_t z = zlo; z < zhi; ++z) {
            double const d = res.profile[z - zlo] - target[z - zlo];
            res.loss += d * d / nl;
            w[z] = 2.0 * d / nl / T;
        }
        if (Kc) {
            // loss seed of a particle in layer z after step t: w[z] / count
            std::vector<float> stab(uint64_t(T) * Z, 0.0f);
            for (uint32_t t = 0; t < T; ++t)
                for (uint32_t z = 0; z < Z; ++z) {
                    uint32_t const c = ct[uint64_t(t) * Z + z];
                    stab[uint64_t(t) * Z + z] = c ? static_cast<float>(w[z] / static_cast<double>(c)) : 0.0f;
                }
            ck(hipMemcpy(a.stab, stab.data(), 4ull * T * Z, hipMemcpyHostToDevice), "upload loss seeds");
            // reverse pass, one segment at a time from the last; every
            // segment but the last two is first recomputed from its start
            // state with the tape
            uint32_t const gb = static_cast<uint32_t>((C + ADJ_CELLS - 1) / ADJ_CELLS);
            ck(hipMemsetAsync(a.lam[0], 0, 12 * N, h->stream), "clear");
            ck(hipMemsetAsync(a.part, 0, 16ull * gb, h->stream), "clear");
            ck(hipMemsetAsync(a.gf, 0, 8 * C, h->stream), "clear");
            int li = 0;
            // Segments are run backwards from the last; the forward pass
            // taped the last two (halves 0 and 1), and every earlier one is
            // recomputed from its checkpoint into the half of the segment two
            // after it on a second stream, once that segment has been run
            // backwards, while the segment after it runs backwards (the
            // reverse step reads only its tape, the adjoint state and the
            // system's fixed geometry; the recompute writes the state, the
            // sort arrays and its own half).
            hipStream_t const R = h->stream;
            bool const two = nseg > 1;
            auto half_of = [&](int seg) { return static_cast<uint32_t>((static_cast<int>(nseg) - 1 - seg) & 1); };
            if (two) {
                ck(hipEventRecord(a.written[0], R), "event");   // the forward pass and the clears above
                ck(hipStreamWaitEvent(a.recompute, a.written[0], 0), "wait");
            }
            for (int seg = static_cast<int>(nseg) - 1; seg >= 0; --seg) {
                uint32_t const t0 = static_cast<uint32_t>(seg) * Kc, t1 = std::min(T, t0 + Kc);
                if (seg >= 1 && seg + 1 < static_cast<int>(nseg)) {
                    // the segment before this one, into the other half, once the
                    // reverse pass that last read that half (segment seg + 1) is done
                    int const ps = seg - 1;
                    uint32_t const hp = half_of(ps);
                    uint32_t const p0 = static_cast<uint32_t>(ps) * Kc, p1 = std::min(T, p0 + Kc);
                    ck(hipStreamWaitEvent(a.recompute, a.read[hp], 0), "wait");
                    h->stream = a.recompute;
                    try {
                        ck(hipMemcpyAsync(h->st[h->cur], a.ckpt + uint64_t(ps) * 6 * N, 24 * N, hipMemcpyDeviceToDevice,
                                          h->stream), "restore checkpoint");
                        h->step_index = s0 + p0;
                        h->keys_valid = false;
                        h->sorted_valid = false;
                        hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, h->stream, h->sys, static_cast<uint32_t>(s0 + p0));
                        for (uint32_t t = p0 + 1; t <= p1; ++t)
                            adj_step(h, force, static_cast<int>(hp * Kc + (t - p0 - 1)), nullptr, nullptr);
                        ck(hipEventRecord(a.written[hp], h->stream), "event");
                    } catch (...) {
                        h->stream = R;
                        throw;
                    }
                    h->stream = R;
                }
                uint32_t const hs = half_of(seg);
                if (two && seg + 2 < static_cast<int>(nseg)) ck(hipStreamWaitEvent(R, a.written[hs], 0), "wait");
                for (uint32_t t = t1; t > t0; --t) {
                    uint64_t const j = hs * Kc + (t - t0 - 1);
                    AdjIn const in{a.off + j * (C + 1), a.u + j * 3 * C, a.ax + j * 3 * C, a.tot + j * C, a.perm + j * N,
                                   a.vp + j * 3 * N, a.lz + j * N, a.stab + uint64_t(t - 1) * Z, a.lam[li], a.lam[1 - li]};
                    hipLaunchKernelGGL(k_adjoint_step, dim3(gb), dim3(WG), 0, R, h->sys, in, a.part, a.gf);
                    li = 1 - li;
                }
                ck(hipGetLastError(), "reverse pass");
                if (two) ck(hipEventRecord(a.read[hs], R), "event");
            }
            std::vector<double> part(2ull * gb), gf(C);
            ck(hipStreamSynchronize(h->stream), "reverse pass");
            ck(hipMemcpy(part.data(), a.part, 16ull * gb, hipMemcpyDeviceToHost), "download gradients");
            for (uint32_t b = 0; b < gb; ++b) {
                res.grad_drag += part[b];
                res.grad_angle += part[gb + b];
            }
            res.grad_angle *= M_PI / 180.0;  // per degree
            if (grad_force) {
                ck(hipMemcpy(gf.data(), a.gf, 8 * C, hipMemcpyDeviceToHost), "download force gradient");
                for (uint64_t c = 0; c < C; ++c) grad_force[c] = static_cast<float>(gf[c]);
            }
        }
        restore();
        ck(hipStreamSynchronize(h->stream), "restore");
    } catch (...) {
        // a segment recompute may still be writing the state
        if (a.recompute) (void)hipStreamSynchronize(a.recompute);
        try {
            restore();
        } catch (...) {
        }
        throw;
    }
    return res;
}

std::string validate(mpcd_params_t const& p) {
    if (p.experiment != 0 && p.experiment != 1) return "unknown experiment (standard or channel)";
    if (p.experiment == 1 && !(p.volume_size[0] > 2.0f))
        return "the channel needs more than 2 cells along x (its radius is (X - 2) / 2 at initialisation)";
    if (p.algorithm != 0 && p.algorithm != 1) return "Unsupported MPCD algorithm";
    for (int a = 0; a < 3; ++a) {
        if (!(p.volume_size[a] >= 1.0f) || p.volume_size[a] != floorf(p.volume_size[a]))
            return "volume_size must hold positive whole numbers of cells";
        if (p.volume_size[a] > 2097152.0f) return "volume_size must stay at or below 2^21 cells per axis";
    }
    if (p.periodicity[0] == 0 || p.periodicity[1] == 0)
        return "the hip backend supports walls in z only (periodicity (1, 1, 0) or (1, 1, 1))";
    if (p.periodicity[2] == 0 && p.volume_size[2] < 3.0f) return "walls in z need at least 3 cell layers";
    if (p.n == 0) return "n must be positive";
    if (p.N == 0) return "particle count N must be positive";
    if (p.N >= 0xffffffffull) return "particle count must be below 2^32";
    double const cells = static_cast<double>(p.volume_size[0]) * p.volume_size[1] * p.volume_size[2];
    if (cells >= 4294967295.0) return "cell count must be below 2^32";
    return "";
}

int create_impl(const mpcd_params_t* params, uint32_t n_systems, mpcd_hip_t** out) {
    auto fail = [](std::string const& m) {
        std::lock_guard<std::mutex> g(g_err_mu);
        g_create_err = m;
        return 1;
    };
    if (!params || !out) return fail("null argument");
    *out = nullptr;
    if (n_systems == 0 || n_systems > 65535) return fail("an ensemble holds 1 to 65535 systems");
    uint64_t N = 0, C = 0;
    for (uint32_t s = 0; s < n_systems; ++s) {
        std::string const e = validate(params[s]);
        if (!e.empty()) return fail(n_systems > 1 ? "system " + std::to_string(s) + ": " + e : e);
        if (params[s].algorithm != params[0].algorithm) return fail("all systems of an ensemble use the same algorithm");
        if (params[s].device_id != params[0].device_id) return fail("all systems of an ensemble use the same device");
        N += params[s].N;
        C += static_cast<uint64_t>(params[s].volume_size[0]) * static_cast<uint64_t>(params[s].volume_size[1]) *
             static_cast<uint64_t>(params[s].volume_size[2]);
    }
    if (N >= 0xffffffffull) return fail("particle count of all systems must be below 2^32");
    if (C >= 0xffffffffull) return fail("cell count of all systems must be below 2^32");
    auto* h = new mpcd_hip_s();
    h->prm.assign(params, params + n_systems);
    h->ns = n_systems;
    h->algorithm = params[0].algorithm;
    h->N = N;
    h->C = C;
    try {
        int ndev = 0;
        ck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (ndev <= 0) throw HipError{"No HIP device found"};
        if (static_cast<int>(params[0].device_id) >= ndev) throw HipError{"device_id out of range"};
        h->device = static_cast<int>(params[0].device_id);
        ck(hipSetDevice(h->device), "hipSetDevice");
        ck(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "stream");
        h->kp.resize(n_systems);
        for (uint32_t s = 0; s < n_systems; ++s) fill_geometry(params[s], h->kp[s]);
        upload_axes(h->device);
        ck(hipMalloc(&h->d_kp, sizeof(KParams) * n_systems), "alloc system table");
        choose_tiling(h, 0, 0);
        layout(h);
        uint64_t const Na = std::max<uint64_t>(h->N, 1);
        ck(hipMalloc(&h->st[0], 24 * Na), "alloc state");
        ck(hipMalloc(&h->st[1], 24 * Na), "alloc state");
        ck(hipMalloc(&h->perm, 4 * Na), "alloc perm");
        ck(hipMalloc(&h->keys, 4 * Na), "alloc keys");
        ck(hipMalloc(&h->count, 4 * h->C), "alloc counts");
        ck(hipMalloc(&h->offset, 4 * (h->C + 1)), "alloc offsets");
        ck(hipMalloc(&h->cursor, 4 * h->C), "alloc cursors");
        if (h->algorithm == 0) ck(hipMalloc(&h->axis, 12 * h->C), "alloc axes");
        ck(hipMalloc(&h->wgcnt, 4 * std::max<uint64_t>(h->W, 1)), "alloc ghost counts");
        ck(hipMalloc(&h->wgvel, 4 * std::max<uint64_t>(h->GV, 1)), "alloc ghost draws");
        h->nb = static_cast<uint32_t>((h->C + SCAN_CHUNK - 1) / SCAN_CHUNK);
        ck(hipMalloc(&h->partial, 4 * std::max<uint32_t>(h->nb, 1)), "alloc scan partials");
        ck(hipMalloc(&h->status, 4 * big_list_words(h->N)), "alloc status and big-cell list");
        ck(hipMalloc(&h->cmax, 4), "alloc count max");
        ck(hipMemsetAsync(h->status, 0, 4 * big_list_words(h->N), h->stream), "clear status");
        ck(hipMemsetAsync(h->count, 0, 4 * h->C, h->stream), "clear counts");
        if (h->blocks_part)
            hipLaunchKernelGGL(k_init, dim3(h->blocks_part), dim3(WG), 0, h->stream, h->sys, h->st[0], h->status);
        ck(hipGetLastError(), "k_init");
        ck(hipStreamSynchronize(h->stream), "init");
        uint32_t st = 0;
        ck(hipMemcpy(&st, h->status, 4, hipMemcpyDeviceToHost), "status");
        if (st & 2u) throw HipError{"channel initialisation: the cylinder holds too little of the box"};
    } catch (HipError const& e) {
        destroy(h);
        return fail(e.msg);
    }
    *out = h;
    return 0;
}

// global cell key -> global reference cell index (system's first cell +
// x + X (y + Y z)); both orders cover the same range [cb, cb + C) per system
std::vector<uint32_t> ref_of_keys(mpcd_hip_t* h) {
    std::vector<uint32_t> r(h->C);
    for (KParams const& k : h->kp)
        for (uint32_t key = 0; key < k.C; ++key) {
            uint32_t ix, iy, iz;
            decode_key(k, key, ix, iy, iz);
            r[k.cb + key] = k.cb + ref_of(k, ix, iy, iz);
        }
    return r;
}

// FP64 accumulators of cells [first, first + n) in reference order (one whole
// system or the whole grid) scaled by 1/sample_counter (the reference divides
// the sum of S samples by S-1, cpu_backend.cpp:85-114)
void sample_mean_impl(mpcd_hip_t* h, uint64_t first, uint64_t n, float* d32, float* v32, double* d64, double* v64) {
    std::vector<double> acc(4 * n, 0.0);
    if (h->acc) {
        std::vector<double> by_key(4 * n);
        ck(hipStreamSynchronize(h->stream), "sync");
        ck(hipMemcpy(by_key.data(), h->acc + 4 * first, 8 * 4 * n, hipMemcpyDeviceToHost), "download accumulators");
        std::vector<uint32_t> const ref = ref_of_keys(h);
        for (uint64_t g = 0; g < n; ++g) {
            uint64_t const c = ref[first + g] - first;
            for (int k = 0; k < 4; ++k) acc[4 * c + k] = by_key[4 * g + k];
        }
    }
    double const inverse = 1.0 / static_cast<double>(h->sample_counter);
    for (uint64_t c = 0; c < n; ++c) {
        if (d32) {
            d32[c] = static_cast<float>(acc[4 * c] * inverse);
            for (int k = 0; k < 3; ++k) v32[3 * c + k] = static_cast<float>(acc[4 * c + 1 + k] * inverse);
        }
        if (d64) {
            d64[c] = acc[4 * c] * inverse;
            for (int k = 0; k < 3; ++k) v64[3 * c + k] = acc[4 * c + 1 + k] * inverse;
        }
    }
}

// ---------------------------------------------------- allocation-free step ---
// mpcd_hip_step_noalloc: n_steps of one system whose state the caller holds in
// device memory, with every array of the step in `scratch` instead of buffers
// of a handle.  The launches are the handle's own (enqueue_step) on a
// handle-shaped object on the stack whose pointers address the scratch
// regions; nothing is allocated on the device.

constexpr int NOALLOC_DEVICES = 64;
std::mutex g_noalloc_mu;                       // one call at a time in this process
unsigned char* g_noalloc_ptr[NOALLOC_DEVICES] = {};

// device address of `scratch` on the current device (looked up on the first
// call on that device; each device holds its own copy of the array)
unsigned char* noalloc_scratch(int device) {
    if (device < 0 || device >= NOALLOC_DEVICES) throw HipError{"device id above the allocation-free step's table"};
    if (!g_noalloc_ptr[device]) {
        void* ptr = nullptr;
        ck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) throw HipError{"hipGetSymbolAddress(scratch) returned no address"};
        g_noalloc_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    return g_noalloc_ptr[device];
}

void step_noalloc_impl(mpcd_params_t const& prm, uint64_t step_index, int n_steps, float* state_in,
                       float* state_out, size_t n_particles) {
    std::string const e = validate(prm);
    if (!e.empty()) throw HipError{e};
    if (n_steps < 1) throw HipError{"n_steps must be positive"};
    if (!state_in || !state_out || state_in == state_out) throw HipError{"state_in and state_out must be two device buffers"};
    if (n_particles != prm.N) throw HipError{"n_particles must equal params->N"};
    uint64_t const C = static_cast<uint64_t>(prm.volume_size[0]) * static_cast<uint64_t>(prm.volume_size[1]) *
                       static_cast<uint64_t>(prm.volume_size[2]);
    uint64_t const walls = prm.periodicity[2] ? 0 : 2ull * static_cast<uint64_t>(prm.volume_size[0]) *
                                                       static_cast<uint64_t>(prm.volume_size[1]);
    if (prm.N > NOALLOC_MAX_PARTICLES || C > NOALLOC_MAX_CELLS)
        throw HipError{"the allocation-free step takes at most " + std::to_string(NOALLOC_MAX_PARTICLES) +
                       " particles and " + std::to_string(NOALLOC_MAX_CELLS) + " cells (got " + std::to_string(prm.N) +
                       " and " + std::to_string(C) + ")"};
    // floats per wall cell: the SRD ghosts' velocity sum, or n extended ghosts
    uint64_t const per_wall = prm.algorithm == 1 ? 6ull * prm.n : 3ull;
    if (per_wall * walls > NOALLOC_MAX_GHOST_FLOATS)
        throw HipError{"the allocation-free step holds the ghosts of at most " +
                       std::to_string(NOALLOC_MAX_GHOST_FLOATS / per_wall) + " wall-layer cells (2 X Y) at this n"};
    int ndev = 0;
    ck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (static_cast<int>(prm.device_id) >= ndev) throw HipError{"device_id out of range"};
    ck(hipSetDevice(static_cast<int>(prm.device_id)), "hipSetDevice");
    unsigned char* const base = noalloc_scratch(static_cast<int>(prm.device_id));
    if (prm.algorithm == 1) upload_axes(static_cast<int>(prm.device_id));

    // the system table entry and the launch shape (choose_tiling / layout for one system)
    KParams kp;
    fill_geometry(prm, kp);
    uint32_t const cap = pick_capacity(prm.algorithm, kp.n, 0, C);
    kp.tile_cells = std::min<uint32_t>(std::max<uint32_t>(1, (cap + kp.n / 2) / kp.n),
                                       static_cast<uint32_t>(tile_cell_max(static_cast<int>(cap), prm.algorithm)));
    kp.step = static_cast<uint32_t>(step_index);
    grid_shift(kp.rs, kp.step, kp.shift);
    grid_shift(kp.rs, kp.step + 1, kp.shift_next);
    auto blocks = [](uint64_t items) { return static_cast<uint32_t>(std::min<uint64_t>((items + WG - 1) / WG, 65536)); };
    uint32_t const tiles = static_cast<uint32_t>((C + kp.tile_cells - 1) / kp.tile_cells);
    uint32_t const blk[8] = {0u, blocks(prm.N), 0u, blocks(C), 0u, blocks(walls), 0u, tiles};

    mpcd_hip_s g;   // pointers into `scratch`; owns nothing
    g.ns = 1;
    g.algorithm = prm.algorithm;
    g.N = prm.N;
    g.C = C;
    g.W = walls;
    g.GV = per_wall * walls;
    g.stream = nullptr;
    g.lds_cap = cap;
    g.grp[0] = mpcd_hip_s::Group{cap, 0u, tiles, 0u};
    g.ngrp = 1;
    g.blocks_part = blk[1];
    g.blocks_cell = blk[3];
    g.blocks_wall = blk[5];
    g.ntiles = tiles;
    g.nb = static_cast<uint32_t>((C + SCAN_CHUNK - 1) / SCAN_CHUNK);
    g.perm = reinterpret_cast<uint32_t*>(base + NA_PERM);
    g.keys = reinterpret_cast<uint32_t*>(base + NA_KEYS);
    g.count = reinterpret_cast<uint32_t*>(base + NA_COUNT);
    g.offset = reinterpret_cast<uint32_t*>(base + NA_OFFSET);
    g.cursor = reinterpret_cast<uint32_t*>(base + NA_CURSOR);
    g.partial = reinterpret_cast<uint32_t*>(base + NA_PARTIAL);
    g.status = reinterpret_cast<uint32_t*>(base + NA_STATUS);
    g.axis = nullptr;   // the collide tiles draw the axes
    g.wgcnt = reinterpret_cast<uint32_t*>(base + NA_GCNT);
    g.wgvel = reinterpret_cast<float*>(base + NA_GVEL);
    g.d_kp = reinterpret_cast<KParams*>(base + NA_KP);
    g.d_blk = reinterpret_cast<uint32_t*>(base + NA_BLK);
    g.sys = Sys{g.d_kp, g.d_blk, g.d_blk + 2, g.d_blk + 4, nullptr, g.d_blk + 6, 1u, static_cast<uint32_t>(C), prm.N};
    // the result lands in state_out: for an even number of steps the input is
    // first copied there and state_in is the other buffer
    bool const even = (n_steps % 2) == 0;
    g.st[0] = even ? state_out : state_in;
    g.st[1] = even ? state_in : state_out;
    g.cur = 0;

    // set-up on the null stream; the host arrays live until the final synchronisation
    ck(hipMemcpyAsync(g.d_kp, &kp, sizeof(KParams), hipMemcpyHostToDevice, nullptr), "upload system table");
    ck(hipMemcpyAsync(g.d_blk, blk, sizeof(blk), hipMemcpyHostToDevice, nullptr), "upload block maps");
    ck(hipMemsetAsync(g.status, 0, 8, nullptr), "clear status");   // error bits and the big-cell count
    if (even) ck(hipMemcpyAsync(state_out, state_in, 24 * prm.N, hipMemcpyDeviceToDevice, nullptr), "copy state");
    launch_keys(&g);
    for (int s = 0; s < n_steps; ++s) {
        uint32_t const io = (s > 0 ? 1u : 0u) | (s + 1 < n_steps ? 2u : 0u);
        enqueue_step(&g, false, false, true, nullptr, nullptr, nullptr, io);
        g.cur = 1 - g.cur;
    }
    ck(hipStreamSynchronize(nullptr), "allocation-free step");
    uint32_t st = 0;
    ck(hipMemcpy(&st, g.status, 4, hipMemcpyDeviceToHost), "status");
    if (st & 1u)
        throw HipError{"a cell holds more particles than one LDS tile (" + std::to_string(cap) +
                       ") and its collision table (4n) does not fit the tile either"};
}

}  // namespace

// ================================================================== C-ABI ===
extern "C" {

const char* mpcd_hip_last_error(const mpcd_hip_t* h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> g(g_err_mu);
    return g_create_err.c_str();
}

int mpcd_hip_create(const mpcd_params_t* params, mpcd_hip_t** out) { return create_impl(params, 1, out); }

int mpcd_hip_create_ensemble(const mpcd_params_t* params, uint32_t n_systems, mpcd_hip_t** out) {
    return create_impl(params, n_systems, out);
}

int mpcd_hip_destroy(mpcd_hip_t* h) {
    destroy(h);
    return 0;
}

int mpcd_hip_step(mpcd_hip_t* h, int n_steps) {
    return guarded(h, [&] {
        ck(hipSetDevice(h->device), "hipSetDevice");
        run_steps(h, n_steps, false);
    });
}

int mpcd_hip_step_and_sample(mpcd_hip_t* h, int n_steps) {
    return guarded(h, [&] {
        ck(hipSetDevice(h->device), "hipSetDevice");
        run_steps(h, n_steps, true);
    });
}

int mpcd_hip_step_injected(mpcd_hip_t* h, const float* shift, const float* axes, size_t n_cells,
                           const uint32_t* ghost_counts, const float* ghost_velocities, size_t n_wall_cells,
                           size_t n_ghost_values) {
    return guarded(h, [&] {
        if (!shift || !axes) throw HipError{"injected step: shift and axes are required"};
        if (h->ns != 1 || h->algorithm != 0) throw HipError{"injected step: single-system SRD handles only"};
        KParams& k = h->kp[0];
        uint32_t const X = k.dim[0], Y = k.dim[1], Z = k.dim[2];
        uint64_t const XY = uint64_t(X) * Y;
        bool const walls = !k.perio_z;
        if (n_cells != h->C) throw HipError{"injected step: axes for a different number of cells"};
        if (n_wall_cells != (walls ? 2 * XY : 0))
            throw HipError{"injected step: ghost arrays for a different number of wall-layer cells"};
        if (walls && (!ghost_counts || !ghost_velocities))
            throw HipError{"injected step: a system with walls needs ghost counts and velocities"};
        if (n_ghost_values != (walls ? 2 * XY * 3 * k.n : 0))
            throw HipError{"injected step: ghost velocities must hold n slots of 3 values per wall-layer cell"};
        ck(hipSetDevice(h->device), "hipSetDevice");
        mpcd_hip_s::Injected d;
        d.axes.assign(3 * h->C, 0.0f);
        for (uint32_t iz = 0; iz < Z; ++iz)
            for (uint32_t iy = 0; iy < Y; ++iy)
                for (uint32_t ix = 0; ix < X; ++ix) {
                    uint64_t const c = ix + X * (iy + uint64_t(Y) * iz);
                    uint64_t const key = k.cb + key_of(k, ix, iy, iz);
                    for (int a = 0; a < 3; ++a) d.axes[a * h->C + key] = axes[3 * c + a];
                }
        if (walls) {
            d.gcnt.assign(ghost_counts, ghost_counts + 2 * XY);
            for (uint32_t g : d.gcnt)
                if (g > k.n) throw HipError{"injected step: a ghost count above n"};
            // the SRD collision takes a wall cell's ghosts as the sum of their
            // velocities: summed here in ghost order
            d.gvel.assign(3 * 2 * XY, 0.0f);
            for (uint64_t w = 0; w < 2 * XY; ++w)
                for (uint32_t g = 0; g < d.gcnt[w]; ++g)
                    for (int a = 0; a < 3; ++a) d.gvel[3 * w + a] += ghost_velocities[(w * k.n + g) * 3 + a];
        }
        // the step's state with the injected shift, its cells counted from scratch
        for (int a = 0; a < 3; ++a) d.shift[a] = shift[a];
        h->inject = &d;
        h->axes_mem = true;
        h->keys_valid = false;
        h->sorted_valid = false;
        try {
            upload_table(h);
            run_steps(h, 1, false);
        } catch (...) {
            (void)hipStreamSynchronize(h->stream);
            h->inject = nullptr;
            h->axes_mem = false;
            upload_table(h);
            throw;
        }
        h->inject = nullptr;
        h->axes_mem = false;
        upload_table(h);   // the host table follows the device's (the next step's shifts)
    });
}

int mpcd_hip_step_injected_extended(mpcd_hip_t* h, const float* shift, const uint32_t* cell_draws,
                                    const float* thermal, size_t n_cells, size_t n_thermal,
                                    const uint32_t* ghost_counts, const float* ghosts, size_t n_wall_cells,
                                    size_t n_ghost_values) {
    return guarded(h, [&] {
        if (!shift || !cell_draws || !thermal)
            throw HipError{"injected step: shift, cell draws and thermal draws are required"};
        if (h->ns != 1 || h->algorithm != 1) throw HipError{"injected step: single-system extended-rule handles only"};
        KParams& k = h->kp[0];
        uint64_t const XY = uint64_t(k.dim[0]) * k.dim[1];
        bool const walls = !k.perio_z;
        if (n_cells != h->C) throw HipError{"injected step: cell draws for a different number of cells"};
        if (n_wall_cells != (walls ? 2 * XY : 0))
            throw HipError{"injected step: ghost arrays for a different number of wall-layer cells"};
        if (walls && (!ghost_counts || !ghosts))
            throw HipError{"injected step: a system with walls needs ghost counts, positions and velocities"};
        if (n_thermal != h->C * k.cap) throw HipError{"injected step: thermal draws must hold 4 n values per cell"};
        if (n_ghost_values != (walls ? 2 * XY * 6 * k.n : 0))
            throw HipError{"injected step: ghosts must hold n slots of 6 values per wall-layer cell"};
        for (uint64_t c = 0; c < h->C; ++c)
            if (cell_draws[3 * c] > 12 || cell_draws[3 * c + 1] > 1)
                throw HipError{"injected step: a direction index above 12 or a sign above 1"};
        ck(hipSetDevice(h->device), "hipSetDevice");
        mpcd_hip_s::Injected d;
        for (int a = 0; a < 3; ++a) d.shift[a] = shift[a];
        if (walls) {
            d.gcnt.assign(ghost_counts, ghost_counts + 2 * XY);
            for (uint32_t g : d.gcnt)
                if (g > k.n) throw HipError{"injected step: a ghost count above n"};
            d.gvel.assign(ghosts, ghosts + 2 * XY * 6 * k.n);
        }
        uint32_t* dev_ext = nullptr;
        float* dev_gauss = nullptr;
        auto release = [&] {
            k.inj_ext = nullptr;
            k.inj_gauss = nullptr;
            h->inject = nullptr;
            (void)hipFree(dev_ext);
            (void)hipFree(dev_gauss);
        };
        try {
            ck(hipMalloc(&dev_ext, 12 * h->C), "alloc injected cell draws");
            ck(hipMalloc(&dev_gauss, 4 * h->C * k.cap), "alloc injected thermal draws");
            ck(hipMemcpy(dev_ext, cell_draws, 12 * h->C, hipMemcpyHostToDevice), "injected cell draws");
            ck(hipMemcpy(dev_gauss, thermal, 4 * h->C * k.cap, hipMemcpyHostToDevice), "injected thermal draws");
            k.inj_ext = dev_ext;
            k.inj_gauss = dev_gauss;
            h->inject = &d;
            h->keys_valid = false;
            h->sorted_valid = false;
            upload_table(h);
            run_steps(h, 1, false);
        } catch (...) {
            ck(hipStreamSynchronize(h->stream), "sync");
            release();
            upload_table(h);
            throw;
        }
        release();
        upload_table(h);   // the host table follows the device's (the next step's shifts)
    });
}

int mpcd_hip_get_sample_mean(mpcd_hip_t* h, float* density, float* velocity, size_t n_cells) {
    return guarded(h, [&] {
        if (n_cells != h->C) throw HipError{"cell array size does not match the grid"};
        sample_mean_impl(h, 0, h->C, density, velocity, nullptr, nullptr);
        h->sampling_done = true;
    });
}

int mpcd_hip_get_sample_mean_f64(mpcd_hip_t* h, double* density, double* velocity, size_t n_cells) {
    return guarded(h, [&] {
        if (n_cells != h->C) throw HipError{"cell array size does not match the grid"};
        sample_mean_impl(h, 0, h->C, nullptr, nullptr, density, velocity);
    });
}

int mpcd_hip_n_particles(const mpcd_hip_t* h, uint64_t* out) {
    if (!h || !out) return 1;
    *out = h->N;
    return 0;
}

int mpcd_hip_n_cells(const mpcd_hip_t* h, uint64_t* out) {
    if (!h || !out) return 1;
    *out = h->C;
    return 0;
}

int mpcd_hip_get_positions(mpcd_hip_t* h, float* out, size_t n) {
    return guarded(h, [&] { check_particles(h, n); get_component(h, out, 0, 0, h->N); });
}
int mpcd_hip_get_velocities(mpcd_hip_t* h, float* out, size_t n) {
    return guarded(h, [&] { check_particles(h, n); get_component(h, out, 3, 0, h->N); });
}
int mpcd_hip_set_positions(mpcd_hip_t* h, const float* pos, size_t n) {
    return guarded(h, [&] { check_particles(h, n); put_component(h, pos, 0); });
}
int mpcd_hip_set_velocities(mpcd_hip_t* h, const float* vel, size_t n) {
    return guarded(h, [&] { check_particles(h, n); put_component(h, vel, 3); });
}
int mpcd_hip_set_particles(mpcd_hip_t* h, const float* pos, const float* vel, size_t n) {
    return guarded(h, [&] {
        check_particles(h, n);
        put_component(h, pos, 0);
        put_component(h, vel, 3);
    });
}

int mpcd_hip_get_step_index(const mpcd_hip_t* h, uint64_t* out) {
    if (!h || !out) return 1;
    *out = h->step_index;
    return 0;
}
int mpcd_hip_set_step_index(mpcd_hip_t* h, uint64_t step) {
    return guarded(h, [&] {
        h->step_index = step;
        h->keys_valid = false;
        h->sorted_valid = false;
        upload_table(h);
    });
}

int mpcd_hip_grid_shift(const mpcd_hip_t* h, uint64_t step, float* out3) {
    if (!h || !out3) return 1;
    grid_shift(h->kp[0].rs, static_cast<uint32_t>(step), out3);
    return 0;
}

int mpcd_hip_peek_sort(mpcd_hip_t* h, uint32_t* cell_of_particle, uint32_t* counts, uint32_t* permutation) {
    return guarded(h, [&] {
        ck(hipSetDevice(h->device), "hipSetDevice");
        if (!h->keys_valid) launch_keys(h);
        if (!h->sorted_valid) {
            launch_sort(h);
            h->sorted_valid = true;
        }
        ck(hipStreamSynchronize(h->stream), "sync");
        std::vector<uint32_t> kr(h->N);
        std::vector<uint32_t> off(h->C + 1), pm(h->N);
        if (h->N) {
            ck(hipMemcpy(kr.data(), h->keys, 4 * h->N, hipMemcpyDeviceToHost), "download keys");
            ck(hipMemcpy(pm.data(), h->perm, 4 * h->N, hipMemcpyDeviceToHost), "download perm");
        }
        ck(hipMemcpy(off.data(), h->offset, 4 * (h->C + 1), hipMemcpyDeviceToHost), "download offsets");
        std::vector<uint32_t> const ref_of_key = ref_of_keys(h);
        std::vector<uint32_t> key_of_ref(h->C);
        for (uint32_t key = 0; key < h->C; ++key) key_of_ref[ref_of_key[key]] = key;
        if (cell_of_particle)
            for (uint64_t j = 0; j < h->N; ++j) cell_of_particle[j] = ref_of_key[kr[j]];
        size_t w = 0;
        for (uint32_t c = 0; c < h->C; ++c) {
            uint32_t const key = key_of_ref[c];
            uint32_t const a = off[key], e = off[key + 1];
            if (counts) counts[c] = e - a;
            if (permutation) {
                std::vector<uint32_t> seg(pm.begin() + a, pm.begin() + e);
                std::sort(seg.begin(), seg.end());
                for (uint32_t v : seg) permutation[w++] = v;
            }
        }
    });
}

int mpcd_hip_record_permutation(mpcd_hip_t* h, int enable) {
    return guarded(h, [&] { h->record = enable != 0; });
}

int mpcd_hip_last_permutation(mpcd_hip_t* h, uint32_t* out, size_t n) {
    return guarded(h, [&] {
        check_particles(h, n);
        if (!h->recorded || !h->perm_rec) throw HipError{"no recorded permutation (enable recording before stepping)"};
        ck(hipStreamSynchronize(h->stream), "sync");
        ck(hipMemcpy(out, h->perm_rec, 4 * h->N, hipMemcpyDeviceToHost), "download permutation");
    });
}

int mpcd_hip_enable_timing(mpcd_hip_t* h, int enable) {
    return guarded(h, [&] { h->timing = enable != 0; });
}

int mpcd_hip_last_timing(const mpcd_hip_t* h, double* collide_ms, double* total_ms, int* launches) {
    if (!h) return 1;
    if (collide_ms) *collide_ms = h->t_collide;
    if (total_ms) *total_ms = h->t_total;
    if (launches) *launches = h->launches;
    return 0;
}

int mpcd_hip_set_tiling(mpcd_hip_t* h, uint32_t cells_per_tile, uint32_t lds_capacity) {
    return guarded(h, [&] {
        if (lds_capacity > TILE_CAP_MAX)
            throw HipError{"tile capacity above " + std::to_string(TILE_CAP_MAX) + " particles"};
        choose_tiling(h, cells_per_tile, lds_capacity);
        layout(h);
    });
}

int mpcd_hip_system_tiling(const mpcd_hip_t* h, uint32_t system, uint32_t* cells_per_tile, uint32_t* lds_capacity) {
    if (!h || system >= h->ns) return 1;
    if (cells_per_tile) *cells_per_tile = h->kp[system].tile_cells;
    if (lds_capacity) *lds_capacity = h->kp[system].lds_cap;
    return 0;
}

int mpcd_hip_get_tiling(const mpcd_hip_t* h, uint32_t* cells_per_tile, uint32_t* lds_capacity) {
    if (!h) return 1;
    if (cells_per_tile) *cells_per_tile = h->kp[0].tile_cells;
    if (lds_capacity) *lds_capacity = h->kp[0].lds_cap;
    return 0;
}

/* ---- adjoint runs ---- */

int mpcd_hip_profile_loss(mpcd_hip_t* h, int steps, float drag, float rotation_angle, const float* force,
                          const double* target, size_t n_layers, double* loss, double* profile) {
    return guarded(h, [&] {
        ck(hipSetDevice(h->device), "hipSetDevice");
        if (!target) throw HipError{"target is required"};
        ProfileResult const r = profile_run(h, steps, 0, drag, rotation_angle, force, target, n_layers, nullptr);
        if (loss) *loss = r.loss;
        if (profile) std::copy(r.profile.begin(), r.profile.end(), profile);
    });
}

int mpcd_hip_profile_loss_grad(mpcd_hip_t* h, int steps, int checkpoint_every, float drag, float rotation_angle,
                               const float* force, const double* target, size_t n_layers, double* loss,
                               double* profile, double* grad_drag, double* grad_angle, float* grad_force) {
    return guarded(h, [&] {
        ck(hipSetDevice(h->device), "hipSetDevice");
        if (!target) throw HipError{"target is required"};
        if (checkpoint_every < 1) throw HipError{"checkpoint_every must be positive"};
        ProfileResult const r =
            profile_run(h, steps, checkpoint_every, drag, rotation_angle, force, target, n_layers, grad_force);
        if (loss) *loss = r.loss;
        if (profile) std::copy(r.profile.begin(), r.profile.end(), profile);
        if (grad_drag) *grad_drag = r.grad_drag;
        if (grad_angle) *grad_angle = r.grad_angle;
    });
}

/* ---- ensembles ---- */

int mpcd_hip_n_systems(const mpcd_hip_t* h, uint32_t* out) {
    if (!h || !out) return 1;
    *out = h->ns;
    return 0;
}

int mpcd_hip_system_info(const mpcd_hip_t* h, uint32_t s, uint64_t* first_particle, uint64_t* n_particles,
                         uint64_t* first_cell, uint64_t* n_cells) {
    if (!h || s >= h->ns) return 1;
    KParams const& k = h->kp[s];
    if (first_particle) *first_particle = k.pb;
    if (n_particles) *n_particles = k.N;
    if (first_cell) *first_cell = k.cb;
    if (n_cells) *n_cells = k.C;
    return 0;
}

int mpcd_hip_system_get_particles(mpcd_hip_t* h, uint32_t s, float* pos, float* vel, size_t n) {
    return guarded(h, [&] {
        check_sys(h, s);
        KParams const& k = h->kp[s];
        if (n != k.N) throw HipError{"particle array size does not match the system"};
        if (pos) get_component(h, pos, 0, k.pb, k.N);
        if (vel) get_component(h, vel, 3, k.pb, k.N);
    });
}

int mpcd_hip_system_sample_mean(mpcd_hip_t* h, uint32_t s, float* density, float* velocity, size_t n_cells) {
    return guarded(h, [&] {
        check_sys(h, s);
        if (n_cells != h->kp[s].C) throw HipError{"cell array size does not match the system"};
        sample_mean_impl(h, h->kp[s].cb, h->kp[s].C, density, velocity, nullptr, nullptr);
    });
}

int mpcd_hip_system_sample_mean_f64(mpcd_hip_t* h, uint32_t s, double* density, double* velocity, size_t n_cells) {
    return guarded(h, [&] {
        check_sys(h, s);
        if (n_cells != h->kp[s].C) throw HipError{"cell array size does not match the system"};
        sample_mean_impl(h, h->kp[s].cb, h->kp[s].C, nullptr, nullptr, density, velocity);
    });
}

int mpcd_hip_reset_sampling(mpcd_hip_t* h) {
    return guarded(h, [&] { h->sampling_done = true; });
}

int mpcd_hip_get_sampling_state(mpcd_hip_t* h, double* acc, size_t n_cells, uint64_t* sample_counter, int* active) {
    return guarded(h, [&] {
        if (n_cells != h->C) throw HipError{"cell array size does not match the grid"};
        if (h->acc) {
            std::vector<double> by_key(4 * h->C);
            ck(hipStreamSynchronize(h->stream), "sync");
            ck(hipMemcpy(by_key.data(), h->acc, 32 * h->C, hipMemcpyDeviceToHost), "download accumulators");
            std::vector<uint32_t> const ref = ref_of_keys(h);
            for (uint64_t g = 0; g < h->C; ++g)
                for (int k = 0; k < 4; ++k) acc[4 * uint64_t(ref[g]) + k] = by_key[4 * g + k];
        } else {
            std::fill(acc, acc + 4 * h->C, 0.0);
        }
        if (sample_counter) *sample_counter = h->sample_counter;
        if (active) *active = h->sampling_done ? 0 : 1;
    });
}

int mpcd_hip_set_sampling_state(mpcd_hip_t* h, const double* acc, size_t n_cells, uint64_t sample_counter, int active) {
    return guarded(h, [&] {
        if (n_cells != h->C) throw HipError{"cell array size does not match the grid"};
        ensure_sampling(h);
        std::vector<double> by_key(4 * h->C);
        std::vector<uint32_t> const ref = ref_of_keys(h);
        for (uint64_t g = 0; g < h->C; ++g)
            for (int k = 0; k < 4; ++k) by_key[4 * g + k] = acc[4 * uint64_t(ref[g]) + k];
        ck(hipStreamSynchronize(h->stream), "sync");
        ck(hipMemcpy(h->acc, by_key.data(), 32 * h->C, hipMemcpyHostToDevice), "upload accumulators");
        h->sample_counter = sample_counter;
        h->sampling_done = active == 0;
    });
}

int mpcd_hip_use_graph(mpcd_hip_t* h, int enable) {
    return guarded(h, [&] {
        h->use_graph = enable != 0;
        if (!h->use_graph) {
            ck(hipStreamSynchronize(h->stream), "sync");
            drop_graphs(h);
        }
    });
}

int mpcd_hip_autotune(mpcd_hip_t* h, int steps_per_trial, uint32_t* cells_per_tile, uint32_t* lds_capacity,
                      double* ms_per_step) {
    return guarded(h, [&] {
        ck(hipSetDevice(h->device), "hipSetDevice");
        if (steps_per_trial < 1) throw HipError{"steps_per_trial must be positive"};
        struct Cand {
            uint32_t tc, cap;
        };
        // Density classes (the systems of one n), the most particles first.  A
        // class's tile shape changes only its own collide work, so each class
        // is tuned on the time of its own collide launches: while it is tuned
        // its systems form launches of their own (launch set 1), timed with
        // events over eager steps, the other classes at their current shapes.
        std::vector<uint32_t> dens;
        std::vector<uint64_t> mass;
        for (KParams const& k : h->kp) {
            auto const it = std::find(dens.begin(), dens.end(), k.n);
            if (it == dens.end()) {
                dens.push_back(k.n);
                mass.push_back(k.N);
            } else {
                mass[it - dens.begin()] += k.N;
            }
        }
        std::vector<size_t> order(dens.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return mass[a] > mass[b]; });
        std::vector<Cand> before(h->ns);
        for (uint32_t s = 0; s < h->ns; ++s) before[s] = {h->kp[s].tile_req, h->kp[s].lds_cap};
        bool const timing = h->timing;
        h->timing = false;
        for (hipEvent_t& e : h->gev) ck(hipEventCreate(&e), "event");
        auto cleanup = [&]() {
            h->group_timing = false;
            for (hipEvent_t& e : h->gev) {
                if (e) (void)hipEventDestroy(e);
                e = nullptr;
            }
            for (KParams& k : h->kp) k.gid = 0;
            h->timing = timing;
        };
        // milliseconds per step of the collide launches of launch set 1: the
        // fastest of steps_per_trial single steps (after one untimed step), so
        // that a step slowed by something else does not decide the ranking
        auto class_trial = [&]() {
            layout(h);
            h->group_timing = true;
            run_steps(h, 1, false);
            double best_ms = 1e300;
            for (int t = 0; t < steps_per_trial; ++t) {
                for (double& v : h->group_ms) v = 0.0;
                run_steps(h, 1, false);
                double ms = 0;
                for (uint32_t g = 0; g < h->ngrp; ++g)
                    if (h->grp[g].gid == 1) ms += h->group_ms[g];
                best_ms = std::min(best_ms, ms);
            }
            h->group_timing = false;
            return best_ms;
        };
        double step_ms = 0;
        try {
            for (size_t ci : order) {
                uint32_t const n = dens[ci];
                std::vector<uint32_t> members;
                for (uint32_t s = 0; s < h->ns; ++s) h->kp[s].gid = h->kp[s].n == n ? 1u : 0u;
                for (uint32_t s = 0; s < h->ns; ++s)
                    if (h->kp[s].n == n) members.push_back(s);
                uint32_t const m0 = members[0];
                auto apply = [&](Cand const& c) {
                    for (uint32_t s : members) choose_tiling_sys(h, s, c.tc, c.cap);
                };
                std::vector<std::pair<double, Cand>> tried;
                std::vector<std::pair<uint32_t, uint32_t>> seen;
                for (uint32_t cap : tune_capacities(h->algorithm, n, h->kp[m0].lds_cap))
                    for (uint32_t tc : {0u, 48u, 32u, 16u, 8u}) {
                        Cand const c = {tc, cap};
                        apply(c);
                        std::pair<uint32_t, uint32_t> const key{h->kp[m0].tile_cells, h->kp[m0].lds_cap};
                        if (std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
                        seen.push_back(key);
                        tried.push_back({class_trial(), c});
                    }
                // the three fastest once more, in reverse order; each keeps its better time
                std::stable_sort(tried.begin(), tried.end(),
                                 [](auto const& a, auto const& b) { return a.first < b.first; });
                size_t const again = std::min<size_t>(3, tried.size());
                for (size_t i = again; i-- > 0;) {
                    apply(tried[i].second);
                    tried[i].first = std::min(tried[i].first, class_trial());
                }
                Cand pick = {h->kp[m0].tile_req, h->kp[m0].lds_cap};
                double best = 1e300;
                for (auto const& t : tried)
                    if (t.first < best) {
                        best = t.first;
                        pick = t.second;
                    }
                apply(pick);
            }
            // the whole step with the chosen shapes, replayed from its graph
            for (KParams& k : h->kp) k.gid = 0;
            layout(h);
            hipEvent_t e0 = h->gev[0], e1 = h->gev[1];
            run_steps(h, 1, false);   // captures the step graph
            ck(hipEventRecord(e0, h->stream), "event");
            run_steps(h, steps_per_trial, false);
            ck(hipEventRecord(e1, h->stream), "event");
            ck(hipEventSynchronize(e1), "event");
            float ms = 0;
            ck(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            step_ms = ms / steps_per_trial;
        } catch (...) {
            cleanup();
            // the handle keeps the tile shapes it had before tuning
            try {
                for (uint32_t s = 0; s < h->ns; ++s) choose_tiling_sys(h, s, before[s].tc, before[s].cap);
                layout(h);
            } catch (...) {
            }
            throw;
        }
        cleanup();
        layout(h);
        if (cells_per_tile) *cells_per_tile = h->kp[0].tile_cells;
        if (lds_capacity) *lds_capacity = h->kp[0].lds_cap;
        if (ms_per_step) *ms_per_step = step_ms;
    });
}

int mpcd_hip_step_noalloc(const mpcd_params_t* params, uint64_t step_index, int n_steps, float* state_in,
                          float* state_out, size_t n_particles) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    try {
        if (!params) throw HipError{"null params"};
        step_noalloc_impl(*params, step_index, n_steps, state_in, state_out, n_particles);
        return 0;
    } catch (HipError const& e) {
        std::lock_guard<std::mutex> g(g_err_mu);
        g_create_err = e.msg;
    } catch (std::exception const& e) {
        std::lock_guard<std::mutex> g(g_err_mu);
        g_create_err = e.what();
    }
    return 1;
}

int mpcd_hip_tune_capacities(int algorithm, uint32_t nmax, uint32_t* caps, uint32_t max_caps, uint32_t* n_caps) {
    if (!caps || !n_caps || (algorithm != 0 && algorithm != 1)) return 1;
    std::vector<uint32_t> const c = tune_capacities(algorithm, std::max<uint32_t>(nmax, 1), 0);
    *n_caps = static_cast<uint32_t>(c.size());
    for (uint32_t i = 0; i < c.size() && i < max_caps; ++i) caps[i] = c[i];
    return 0;
}

int mpcd_hip_bench_copy(size_t bytes, int reps, double* ms_per_copy) {
    if (!ms_per_copy || reps <= 0 || bytes < 16) return 1;
    uint64_t const n = bytes / 16;
    Vec4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, 16 * n) != hipSuccess) return 1;
    if (hipMalloc(&b, 16 * n) != hipSuccess) { (void)hipFree(a); return 1; }
    rc |= hipMemset(a, 0, 16 * n) != hipSuccess;
    rc |= hipEventCreate(&e0) != hipSuccess;
    rc |= hipEventCreate(&e1) != hipSuccess;
    // the fastest of four forms: 4 or 8 float4 per lane, plain or nontemporal stores
    auto run = [&](auto kernel, int U) {
        uint64_t const blocks = (n + static_cast<uint64_t>(WG) * U - 1) / (static_cast<uint64_t>(WG) * U);
        hipLaunchKernelGGL(kernel, dim3(static_cast<uint32_t>(blocks)), dim3(WG), 0, 0, a, b, n);
        rc |= hipEventRecord(e0, 0) != hipSuccess;
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(kernel, dim3(static_cast<uint32_t>(blocks)), dim3(WG), 0, 0, a, b, n);
        rc |= hipEventRecord(e1, 0) != hipSuccess;
        rc |= hipEventSynchronize(e1) != hipSuccess;
        float ms = 0;
        rc |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
        return static_cast<double>(ms) / reps;
    };
    if (!rc && n / (WG * 4ull) < (1ull << 31)) {
        double best = run(k_copy<4, false>, 4);
        best = std::min(best, run(k_copy<4, true>, 4));
        best = std::min(best, run(k_copy<8, false>, 8));
        best = std::min(best, run(k_copy<8, true>, 8));
        *ms_per_copy = best;
    } else {
        rc = 1;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int mpcd_hip_selftest_philox(uint32_t k0, uint32_t k1, const uint32_t* ctr, uint32_t* out, size_t n) {
    uint32_t *dc = nullptr, *dout = nullptr;
    int rc = 0;
    if (hipMalloc(&dc, 16 * std::max<size_t>(n, 1)) != hipSuccess) return 1;
    if (hipMalloc(&dout, 16 * std::max<size_t>(n, 1)) != hipSuccess) { (void)hipFree(dc); return 1; }
    if (hipMemcpy(dc, ctr, 16 * n, hipMemcpyHostToDevice) != hipSuccess) rc = 1;
    if (!rc && n) hipLaunchKernelGGL(k_selftest_philox, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, 0, k0, k1, dc, dout, n);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = 1;
    if (!rc && hipMemcpy(out, dout, 16 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = 1;
    (void)hipFree(dc);
    (void)hipFree(dout);
    return rc;
}

int mpcd_hip_selftest_math(const float* in, float* out, size_t n, int which) {
    float *di = nullptr, *dout = nullptr;
    int rc = 0;
    if (hipMalloc(&di, 4 * std::max<size_t>(n, 1)) != hipSuccess) return 1;
    if (hipMalloc(&dout, 8 * std::max<size_t>(n, 1)) != hipSuccess) { (void)hipFree(di); return 1; }
    if (hipMemcpy(di, in, 4 * n, hipMemcpyHostToDevice) != hipSuccess) rc = 1;
    if (!rc && n) hipLaunchKernelGGL(k_selftest_math, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, 0, di, dout, n, which);
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = 1;
    if (!rc && hipMemcpy(out, dout, 8 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = 1;
    (void)hipFree(di);
    (void)hipFree(dout);
    return rc;
}

}  // extern "C"
