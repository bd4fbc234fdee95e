// Internal declarations of libstarfish_hip.so (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/starfish_hip.h"

namespace sf {

// Physical constants (core/common/Constants.java:20-33).
constexpr double EPS0 = 8.85418782e-12;
constexpr double K_B = 1.3806503e-23;
constexpr double QE = 1.60217646e-19;
constexpr double AMU = 1.66053886e-27;
constexpr double FLT_EPS = 1e-7;
constexpr double PI = 3.14159265358979323846;

// The top four bits of the Philox counter's tag word name the module that draws, so that MCC, DSMC
// and the device injection never share a stream, whatever tags their callers choose; the library
// also refuses two live MCC (or DSMC) objects with the same seed and tag.
constexpr uint32_t RNG_TAG_MASK = 0x0FFFFFFFu;
constexpr uint32_t RNG_MODULE_MCC = 1u << 28;
constexpr uint32_t RNG_MODULE_DSMC = 2u << 28;
constexpr uint32_t RNG_MODULE_INJECT = 3u << 28;

// Tile of cells owned by one workgroup of the push/deposit pass.
constexpr int TILE_W = 32;                      // cells along i
constexpr int TILE_H = 4;                       // cells along j
constexpr int TILE_HALO = 3;                    // cells of halo on every side (drift between sorts)
constexpr int TILE_NW = TILE_W + 1 + 2 * TILE_HALO;   // nodes along i
constexpr int TILE_NH = TILE_H + 1 + 2 * TILE_HALO;
constexpr int TILE_NODES = TILE_NW * TILE_NH;
constexpr int TILE_CW = TILE_W + 2 * TILE_HALO; // cells incl. halo
constexpr int TILE_CH = TILE_H + 2 * TILE_HALO;
constexpr int TILE_CELLS = TILE_CW * TILE_CH;
constexpr int NCH = 7;                          // deposit channels: n, nu, nv, nw, nuu, nvv, nww
constexpr int NGAS = 10;                        // computeFields outputs kept on the device
// bits of Species::flags, set by the push kernels
constexpr int PUSH_FLAG_FAST = 1;               // a velocity beyond the fixed-point fast range: exact slow quantisation
constexpr int PUSH_FLAG_OVERFLOW = 2;           // a moment contribution beyond the int64 range: moments invalid
constexpr int PUSH_THREADS = 256;
constexpr int CHUNK = 256;                      // particles staged in LDS per round
constexpr int CHUNK_STRIDE = 151;               // odd stride of the lane->particle permutation
// Mover sums are added into STAT_SLOTS copies (by workgroup index) of STAT_STRIDE int64 each, one
// 128-byte line apart, so that the atomics of thousands of waves do not queue on five addresses.
constexpr int STAT_SLOTS = 64;
constexpr int STAT_STRIDE = 16;

// Particle state, structure of arrays, one buffer.
struct PState {
    int32_t* cell;   // packed (j << 16) | i of the cell, -1 never stored
    float* fx;       // offset inside the cell along i (lc0 = i + fx)
    float* fy;       // offset along j
    float* vx;
    float* vy;
    float* vz;
    int32_t* id;     // optional
};

// Mesh parameters passed by value to kernels.
struct MeshP {
    int ni, nj, nci, ncj;
    int domain;
    int ntx, nty;     // tiles
    double x0[2], dh[2], inv_dh[2];
};

// Segment record for boundary processing (LinearSegment.java).
struct DevSeg {
    double x1[2], x2[2];
    double nrm[2];
    int btype;
    int absorb;
    int bnd;       // boundary index for the impact tallies (-1: not tallied)
};

// Device pointers of the mesh.
struct MeshD {
    const int8_t* face;        // [2nj+2ni] face types (RIGHT, TOP, LEFT, BOTTOM)
    const int32_t* nseg_ptr;   // [ni*nj+1] node segment lists (DIRICHLET/SINK only)
    const int32_t* nseg_idx;
    const DevSeg* seg;
    const uint8_t* node_hasseg;// [ni*nj] 1 if the node lists a DIRICHLET/SINK segment
    const float2* ef;          // [ni*nj] E field (i, j components)
    const float2* bf;          // [ni*nj] B field (i, j components); null when no field is loaded
    unsigned long long* bnd_hits;   // [nbnd] particles of the pushed species absorbed per boundary (or null)
};

// Per-species push constants.
struct PushC {
    double qm_dt;          // q/m * dt
    double qm;             // q/m
    double dt;
    // fixed-point scales 2^k of this step: density channel, the three velocity channels, the
    // three second moments (one value per group, so the push holds three scalars, not seven)
    double kn, kv, kvv;
    double spwt;
    // the velocity channels carry v - c and (v - c)^2 for the step's reference velocity c (the
    // species' mean velocity), so their scales follow the spread of the velocities about it
    double c0, c1, c2;
    // scale of the mover sums (momentum and speed, physical velocities)
    double kst;
    // deviations max_k |v_k - c_k| beyond vfast need the exact (slow) quantisation, beyond vover
    // a contribution does not fit an int64 (the step's moments are then flagged as invalid)
    double vfast, vover;
};

struct Species {
    sf_species_desc desc;
    double charge, mass, qm;
    int64_t capacity = 0;
    int64_t np = 0, nsorted = 0;   // slots in use (incl. dead holes between sorts), sorted prefix
    int64_t nalive = 0;            // particles alive after the last push or add
    int sort_interval = 4;         // steps between counting sorts; other steps push in place
    int since_sort = 0;
    int cur = 0;
    PState buf[2]{};
    int32_t* key = nullptr;        // new cell key per particle (-1 dead); the push's slow lists
    int* qcount = nullptr;         // [2] particles the last push left to the reference loop, work items
    int32_t* tile_nslow = nullptr; // [tiles] length of each push tile's slow list (kept in key)
    int2* slow_items = nullptr;    // (tile, chunk) work items of the slow lists
    unsigned long long* acc = nullptr;   // [NCH][nn] int64 deposit accumulators
    double* sums = nullptr;        // [NCH][nn] running moment sums, physical units (FP64)
    long long* mpc_sum = nullptr;  // [ncells]
    double* gas = nullptr;         // [NGAS][nn] computeFields results (t, t1-t3, nd-ave, u/v/w-ave, p, mpc)
    int* flags = nullptr;          // device status word of the push (SF_PUSH_FLAG_*)
    double* den = nullptr;         // [4][nn] nd, u, v, w
    int32_t* counts = nullptr;     // [ncells] histogram of new cells
    int32_t* offsets = nullptr;    // [ncells+1] exclusive scan (cell ranges of the sorted array)
    int32_t* cursor = nullptr;     // [ncells]
    int32_t* blocksum = nullptr;   // scan scratch
    long long* stats = nullptr;    // [STAT_SLOTS][STAT_STRIDE] int64 fixed-point mover sums
    int id_counter = 0;
    int num_samples = 0;
    bool first_time = true;        // KineticMaterial.first_time: computeFields on the first update
    int64_t pushes = 0;            // sf_push_deposit calls (the step counter when none is set)
    double kscale[NCH];            // fixed-point scales of the last push (set per push)
    double kstat = 0;              // fixed-point scale of the mover sums (from the species' vmax)
    // reference velocity and a bound of max_k |v_k - c_k| over the particles, from the last push
    // (its mover sums and deviations); invalid until the first push or after particles are loaded
    double vref[3] = {0, 0, 0};
    double vbound = 0;
    bool vrange_valid = false;
    double stats_host[5] = {0, 0, 0, 0, 0};
    unsigned long long* bnd_hits = nullptr;   // [ctx nbnd] particles absorbed per boundary since the last read
};

// Host-side java.util.Random (JDK-specified 48-bit LCG), the shared Starfish generator
// (core/common/Starfish.java:238-295).
struct JavaRandom {
    uint64_t seed = 0;
    void set_seed(int64_t s) { seed = ((uint64_t)s ^ 0x5DEECE66DULL) & ((1ULL << 48) - 1); }
    int32_t next(int bits) {
        seed = (seed * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
        return (int32_t)(uint32_t)(seed >> (48 - bits));
    }
    double next_double() { return (double)(((int64_t)next(26) << 27) + next(27)) * 0x1.0p-53; }
    double rnd2() { return -1.0 + 2 * next_double(); }
};

struct HostSource {
    int species;
    std::vector<double> seg;       // x1,y1,x2,y2 per linear segment
    std::vector<double> cum;
    double area = 0, mdot = 0, v_drift = 0, v_th = 0, rem = 0;
    double v_az = 0;               // MaxwellianSource v_az, added to the third velocity component
    int start_it = 0, stop_it = -1;   // Source start_it / stop_it (end_it); -1: no end
    int64_t added = 0;             // particles this source has placed in the domain so far
    double* d_geom = nullptr;      // device copy of seg then cum, for sampling on the device
};

struct PoissonState;
void poisson_free(PoissonState* p);

}  // namespace sf

struct sf_ctx {
    int device = 0;
    int ncu = 0;                    // compute units of the device
    hipStream_t stream = nullptr;
    std::string err;
    double dt = 1e-7;
    sf::JavaRandom rng;
    bool mesh_ready = false;
    sf_mesh_desc mdesc{};
    sf::MeshP mp{};
    // host copies
    std::vector<int> h_face_type;
    std::vector<double> h_face_value;
    std::vector<int> h_node_type;
    std::vector<double> h_node_bc, h_node_vol;
    // device mesh
    int8_t* d_face = nullptr;
    int32_t* d_nseg_ptr = nullptr;
    int32_t* d_nseg_idx = nullptr;
    sf::DevSeg* d_seg = nullptr;
    uint8_t* d_hasseg = nullptr;
    uint8_t* d_cellslow = nullptr; // [ncells] 1: a particle of this cell may meet a segment or a mesh face
    double* d_node_vol = nullptr;
    // fields
    double* d_phi = nullptr;
    double* d_rho = nullptr;
    double* d_efi = nullptr;
    double* d_efj = nullptr;
    float2* d_ef32 = nullptr;
    float2* d_bf32 = nullptr;   // magnetic field, FP32 like E; allocated only while nonzero
    std::vector<sf::Species*> species;
    std::vector<sf::HostSource> sources;
    uint32_t inject_step = 0;      // batches sampled on the device (counter of their Philox streams)
    int it = 0;                    // TimeModule step (Starfish.getIt()); 0 = not set by the host
    int source_calls = 0;          // sf_sources_sample* calls: the sources' step while `it` is not set
    // (seed, module-tagged stream tag) of every live MCC and DSMC object: no two share a stream
    std::vector<std::pair<uint64_t, uint32_t>> rng_streams;
    sf::PoissonState* poisson = nullptr;
    bool timing = false;
    hipEvent_t ev[8]{};
    double last_ms[4] = {0, 0, 0, 0};
    double last_push_ms[3] = {0, 0, 0};   // tile kernel ms, tail kernel ms, tail slots
    double last_queue[2] = {0, 0};        // queue kernel ms, particles the fast tile kernel queued
    bool no_fast_push = false;            // SF_NO_FAST_PUSH set: every step runs the full tile kernel
    int nbnd = 0;                         // boundaries of the impact tallies (sf_mesh_segment_boundaries)
    int nseg = 0;                         // segments uploaded by sf_mesh_setup
    sf::MeshD meshd() const {
        sf::MeshD m;
        m.face = d_face;
        m.nseg_ptr = d_nseg_ptr;
        m.nseg_idx = d_nseg_idx;
        m.seg = d_seg;
        m.node_hasseg = d_hasseg;
        m.ef = d_ef32;
        m.bf = d_bf32;
        m.bnd_hits = nullptr;
        return m;
    }
};

// error helpers
int sf_fail(sf_ctx* ctx, int code, const char* fmt, ...);
#define SF_HIP(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return sf_fail((ctx), SF_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                                    \
    } while (0)
#define SF_CHECK_CTX(ctx)                          \
    do {                                           \
        if (!(ctx)) return SF_ERR_ARG;             \
    } while (0)

namespace sf {
// Claims the counter-based stream (seed, module | tag) for a new MCC or DSMC object; tag
// SF_TAG_AUTO picks the smallest free tag.  Fails if another live object holds the stream.
int claim_stream(sf_ctx* ctx, uint64_t seed, uint32_t module, uint32_t tag, uint32_t* stream_tag);
void release_stream(sf_ctx* ctx, uint64_t seed, uint32_t stream_tag);
int species_alloc_buffers(sf_ctx* ctx, Species* sp);
int push_deposit(sf_ctx* ctx, Species* sp);
int sort_species(sf_ctx* ctx, Species* sp);
int add_particles_device(sf_ctx* ctx, Species* sp, int64_t n, const double* x, const double* y, const double* u,
                         const double* v, const double* w, const int* ids, int mode);
int noalloc_scratch_init(sf_ctx* ctx);
int add_particles_noalloc(sf_ctx* ctx, Species* sp, int64_t n, const double* x, const double* y, const double* u,
                          const double* v, const double* w, const int* ids, int mode);
int species_field(sf_ctx* ctx, Species* sp, int which, double* out);
int efield_update(sf_ctx* ctx);
int ef32_refresh(sf_ctx* ctx);
int rho_update(sf_ctx* ctx);
int poisson_setup(sf_ctx* ctx, const sf_poisson_desc* d);
int poisson_solve(sf_ctx* ctx, sf_solve_info* info);
int operator_apply(sf_ctx* ctx, const double* x, double* y);
int preload(sf_ctx* ctx, Species* sp, int64_t n, const double* drift, double temp, uint64_t seed);
int inject_maxwellian_device(sf_ctx* ctx, HostSource& s, int num, uint64_t seed, uint32_t step, uint32_t tag,
                             int64_t* added);
// SourceModule.sampleSources' window (SourceModule.java:240-241): the step of this sampling call
// (the host's step counter, else the number of sampling calls so far)
inline int source_step(sf_ctx* ctx);
inline bool source_active(const HostSource& s, int step) {
    return !(step < s.start_it || (s.stop_it >= 0 && step > s.stop_it));
}
}  // namespace sf

inline int sf::source_step(sf_ctx* ctx) {
    ctx->source_calls++;
    return ctx->it > 0 ? ctx->it : ctx->source_calls;
}
