// Push + deposit + per-cell counting sort of one kinetic species on gfx950.
//
// One step of KineticMaterial.updateFields() (core/materials/KineticMaterial.java:117-197,
// 1553-1595) runs as:
//   push    (push_deposit_tile / push_deposit_tail): gather E, push, boundary processing and
//           deposit of the 7 moment channels into LDS int64 accumulators of a 32x4-cell tile,
//           LDS histogram of the new cells; the pushed state is written back to the
//           particle's own slot (a dead particle becomes a hole, cell = -1).
//   scan    (scan_*, sort steps): exclusive scan of the cell histogram -> cell ranges, cursors,
//           mpc-sum.  On other steps the histogram only feeds mpc-sum (mpc_accumulate).
//   scatter (copy_scatter, sort steps): each pushed particle to its slot of its new cell
//           (counting sort + compaction of the holes).
//   normalise (normalize_nodes): int64 channels -> nd, u, v, w; running sums += channels.
// Particles are read coalesced in cell-sorted order.  The deposit and the sums are 64-bit
// fixed point, so every density and moment is independent of the order of the additions.
#include <cstring>
#include <mutex>

#include "sf_particle_math.h"
#include "sf_rng.h"

namespace sf {

// ----------------------------------------------------------------------------------------
// deposit helpers
// ----------------------------------------------------------------------------------------

// The 7 channels of one particle at unit weight: 1, v - c and (v - c)^2 per component, each at
// its fixed-point scale.  normalize_nodes turns the step's integer totals back into the
// reference's sums of mpw, mpw v and mpw v^2.
__device__ __forceinline__ void channel_values(const PushC& pc, double v0, double v1, double v2, double* val) {
    const double d0 = v0 - pc.c0, d1 = v1 - pc.c1, d2 = v2 - pc.c2;
    val[0] = pc.kn;
    val[1] = d0 * pc.kv;
    val[2] = d1 * pc.kv;
    val[3] = d2 * pc.kv;
    val[4] = d0 * d0 * pc.kvv;
    val[5] = d1 * d1 * pc.kvv;
    val[6] = d2 * d2 * pc.kvv;
}

// Global-memory deposit (used for the unsorted tail and for particles that left the tile).
__device__ __forceinline__ void deposit_global(const MeshP& m, unsigned long long* acc, size_t nn, int i, int j,
                                               double di, double dj, const double* val) {
    double w[4];
    scatter_weights(m, i, j, di, dj, w);
    const int ci[4] = {0, 1, 1, 0}, cj[4] = {0, 0, 1, 1};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        size_t u = (size_t)(j + cj[c]) * m.ni + (i + ci[c]);
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            long long q = quant(w[c], val[ch]);
            if (q != 0) atomicAdd(acc + ch * nn + u, (unsigned long long)q);
        }
    }
}

// Deposit of a particle whose speed is beyond the range of the one-FMA quantisation (a
// contribution of 2^51 or more): rounded with a full conversion instead, into global memory.
// Deterministic like the fast path (a particle takes one path or the other by its own speed).
__device__ __forceinline__ void deposit_global_exact(const MeshP& m, unsigned long long* acc, size_t nn, int i, int j,
                                                     double di, double dj, const double* val) {
    double w[4];
    scatter_weights(m, i, j, di, dj, w);
    const int ci[4] = {0, 1, 1, 0}, cj[4] = {0, 0, 1, 1};
    for (int c = 0; c < 4; c++) {
        size_t u = (size_t)(j + cj[c]) * m.ni + (i + ci[c]);
        for (int ch = 0; ch < NCH; ch++) {
            long long q = __double2ll_rn(w[c] * val[ch]);
            if (q != 0) atomicAdd(acc + ch * nn + u, (unsigned long long)q);
        }
    }
}

// Largest deviation of a velocity component from the reference velocity: the one-FMA
// quantisation needs it below pc.vfast.
__device__ __forceinline__ double vdev(const PushC& pc, double v0, double v1, double v2) {
    return fmax(fabs(v0 - pc.c0), fmax(fabs(v1 - pc.c1), fabs(v2 - pc.c2)));
}

// Load of the FP32 particle state, conversion to logical coordinates.
struct PIn {
    int ci, cj;
    double lc0, lc1, v0, v1, v2;
};
__device__ __forceinline__ PIn load_particle(const PState& s, int64_t p) {
    PIn q;
    int32_t c = s.cell[p];
    q.ci = cell_i(c);
    q.cj = cell_j(c);
    q.lc0 = q.ci + (double)s.fx[p];
    q.lc1 = q.cj + (double)s.fy[p];
    q.v0 = s.vx[p];
    q.v1 = s.vy[p];
    q.v2 = s.vz[p];
    return q;
}

// New cell of an alive particle (alive particles satisfy 0 <= lc <= n-1).
__device__ __forceinline__ void new_cell(const MeshP& m, double lc0, double lc1, int& ci, int& cj) {
    ci = (int)lc0;
    cj = (int)lc1;
    if (ci > m.nci - 1) ci = m.nci - 1;
    if (cj > m.ncj - 1) cj = m.ncj - 1;
    if (ci < 0) ci = 0;
    if (cj < 0) cj = 0;
}

// Cell and FP32 offsets of an alive particle.  An offset that rounds up to 1 in FP32 stands for
// the same position as offset 0 in the next cell and is stored that way, so the cell of every
// stored particle is floor(lc) (an offset of exactly 1 remains only in the last cell row/column).
__device__ __forceinline__ void cell_offsets(const MeshP& m, double lc0, double lc1, int& ci, int& cj, float& fx,
                                             float& fy) {
    new_cell(m, lc0, lc1, ci, cj);
    fx = (float)(lc0 - ci);
    fy = (float)(lc1 - cj);
    if (fx >= 1.0f && ci < m.nci - 1) {
        ci++;
        fx = 0.0f;
    }
    if (fy >= 1.0f && cj < m.ncj - 1) {
        cj++;
        fy = 0.0f;
    }
}

// Mover sums of the surviving particles (count, momentum and speed at the scale kst) and the
// largest deviation max_k |v_k - c_k|, which bounds the next step's channel scales.
struct StatAcc {
    long long n, p0, p1, p2, e;
    double dmax;
};

__device__ __forceinline__ void stat_add(StatAcc& s, const PushC& pc, double v0, double v1, double v2) {
    s.n += 1;
    s.p0 += quant(1.0, v0 * pc.kst);
    s.p1 += quant(1.0, v1 * pc.kst);
    s.p2 += quant(1.0, v2 * pc.kst);
    s.e += quant(1.0, sqrt(v0 * v0 + v1 * v1 + v2 * v2) * pc.kst);
    s.dmax = fmax(s.dmax, vdev(pc, v0, v1, v2));
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Wave sums of the mover sums, added into the workgroup's slot of the stats array (the host adds
// the slots).  A single copy would serialise the atomics of every wave of the launch at its end.
__device__ void stat_flush(StatAcc s, long long* stats) {
    s.n = wave_sum_ll(s.n);
    s.p0 = wave_sum_ll(s.p0);
    s.p1 = wave_sum_ll(s.p1);
    s.p2 = wave_sum_ll(s.p2);
    s.e = wave_sum_ll(s.e);
    for (int o = 32; o > 0; o >>= 1) s.dmax = fmax(s.dmax, __shfl_xor(s.dmax, o));
    if ((threadIdx.x & 63) == 0) {
        unsigned long long* slot = (unsigned long long*)stats + (size_t)(blockIdx.x % STAT_SLOTS) * STAT_STRIDE;
        atomicAdd(&slot[0], (unsigned long long)s.n);
        atomicAdd(&slot[1], (unsigned long long)s.p0);
        atomicAdd(&slot[2], (unsigned long long)s.p1);
        atomicAdd(&slot[3], (unsigned long long)s.p2);
        atomicAdd(&slot[4], (unsigned long long)s.e);
        // a non-negative double orders like its bit pattern
        if (s.dmax > 0) atomicMax(&slot[5], (unsigned long long)__double_as_longlong(s.dmax));
    }
}

// Workgroup barrier that waits for the wave's LDS operations only.  __syncthreads() also waits
// for all outstanding global loads and stores (vmcnt(0)); the window loop exchanges data through
// LDS alone, so its prefetched loads and its stores may stay in flight across the barrier.
__device__ __forceinline__ void lds_barrier() {
    __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ----------------------------------------------------------------------------------------
// push of the cell-sorted particles: one workgroup per tile of TILE_W x TILE_H cells
// ----------------------------------------------------------------------------------------
struct TileLds {
    unsigned long long acc[NCH][TILE_NODES];
    int hist[TILE_CELLS];
    float2 ef[TILE_NODES];
    uint8_t hs[TILE_NODES];
    int seg_start[TILE_H], seg_pre[TILE_H + 1];
    // results of one window, by slot: written in the permuted order, stored coalesced
    int32_t o_cell[PUSH_THREADS];
    float o_fx[PUSH_THREADS], o_fy[PUSH_THREADS], o_vx[PUSH_THREADS], o_vy[PUSH_THREADS], o_vz[PUSH_THREADS];
};

// Particle state of one window slot in registers.
struct PReg {
    int64_t g;
    int32_t c;
    float fx, fy, vx, vy, vz;
};

// The pushed state is written back to the particle's own slot; dead particles become holes
// (cell = -1) until the next sort.
// The tile's particles (its cell rows, contiguous after the last sort) are visited in windows of
// PUSH_THREADS; thread t takes the window's particle (t * CHUNK_STRIDE) mod PUSH_THREADS, so the
// lanes of a wave fall about one cell apart and their LDS atomics hit different nodes.  Each
// thread loads its particle straight into registers.  Results go back through
// LDS by slot and are stored coalesced: stores in the permuted order write partial cache lines
// and run at about half the bandwidth of coalesced ones on this part.
template <bool HAS_B>
__device__ __forceinline__ void push_deposit_tile_body(const MeshP& m, const MeshD& md, const PushC& pc,
                                                       const PState& s, const int32_t* __restrict__ offsets,
                                                       unsigned long long* __restrict__ acc,
                                                       int32_t* __restrict__ counts, long long* stats,
                                                       int* flags) {
    __shared__ TileLds L;
    const int t = threadIdx.x;
    StatAcc st{0, 0, 0, 0, 0, 0.0};
    // One workgroup per tile.  The dispatcher starts a waiting tile as soon as a workgroup slot
    // frees up, so tiles that take longer (bounces near objects, particles that left the LDS
    // window) do not leave the rest of the device idle at the end of the launch; this balances
    // better than persistent workgroups taking tiles from a counter.
    const int tile = blockIdx.x;
    const int tx = tile % m.ntx, ty = tile / m.ntx;
    const int ti0 = tx * TILE_W, tj0 = ty * TILE_H;            // first cell of the tile
    const int n_i0 = ti0 - TILE_HALO, n_j0 = tj0 - TILE_HALO;  // first node of the LDS tile
    const int c_i0 = ti0 - TILE_HALO, c_j0 = tj0 - TILE_HALO;  // first cell of the LDS histogram
    const size_t nn = (size_t)m.ni * m.nj;
    const int pt = (t * CHUNK_STRIDE) & (PUSH_THREADS - 1);   // this thread's slot of a window

    for (int k = t; k < NCH * TILE_NODES; k += PUSH_THREADS) (&L.acc[0][0])[k] = 0ull;
    for (int k = t; k < TILE_CELLS; k += PUSH_THREADS) L.hist[k] = 0;
    for (int k = t; k < TILE_NODES; k += PUSH_THREADS) {
        int i = n_i0 + k % TILE_NW, j = n_j0 + k / TILE_NW;
        i = min(max(i, 0), m.ni - 1);
        j = min(max(j, 0), m.nj - 1);
        L.ef[k] = md.ef[(size_t)j * m.ni + i];
        L.hs[k] = md.node_hasseg[(size_t)j * m.ni + i];
    }
    // The row segments (uniform, from the cell offsets of the last sort).  Every thread knows them,
    // so the first window's particle loads are issued before the barrier and overlap the staging
    // of the tile; thread 0 also leaves them in LDS for the window loop.
    int first_row = TILE_H, first_start = 0, first_len = 0;
    {
        int pre = 0;
#pragma unroll
        for (int r = 0; r < TILE_H; r++) {
            const int row = tj0 + r;
            int st = 0, len = 0;
            if (row < m.ncj) {
                const int k0 = row * m.nci + ti0;
                const int k1 = row * m.nci + min(ti0 + TILE_W, m.nci);
                st = offsets[k0];
                len = offsets[k1] - st;
            }
            if (first_row == TILE_H && len > 0) {
                first_row = r;
                first_start = st;
                first_len = len;
            }
            if (t == 0) {
                L.seg_start[r] = st;
                L.seg_pre[r] = pre;
            }
            pre += len;
        }
        if (t == 0) L.seg_pre[TILE_H] = pre;
    }
    PReg first;
    if (pt < first_len) {
        first.g = first_start + pt;
        first.c = s.cell[first.g];
        first.fx = s.fx[first.g];
        first.fy = s.fy[first.g];
        first.vx = s.vx[first.g];
        first.vy = s.vy[first.g];
        first.vz = s.vz[first.g];
    } else {
        first.g = 0;
        first.c = -1;
        first.fx = first.fy = first.vx = first.vy = first.vz = 0.0f;
    }
    __syncthreads();

    // clamped loads reproduce the global array for in-mesh nodes (the window is clamped above)
    auto load_e = [&](int i, int j) -> float2 {
        int li = i - n_i0, lj = j - n_j0;
        if (li >= 0 && li < TILE_NW && lj >= 0 && lj < TILE_NH) return L.ef[lj * TILE_NW + li];
        return md.ef[(size_t)j * m.ni + i];
    };
    auto hasseg = [&](int i, int j) -> int {
        int li = i - n_i0, lj = j - n_j0;
        if (li >= 0 && li < TILE_NW && lj >= 0 && lj < TILE_NH) return L.hs[lj * TILE_NW + li];
        return md.node_hasseg[(size_t)j * m.ni + i];
    };
    // The tile's particles: TILE_H contiguous row segments, each visited in windows of
    // PUSH_THREADS slots (the last window of a row may be partial).  The barriers between
    // windows wait for LDS only, so the previous window's stores stay in flight across them.
    // A window's loads are issued after the stores of the window before, ahead of the barrier
    // between them.  (Loading it before the previous window's push measured slower: the extra
    // registers cost more than the latency it hides at four workgroups per CU.)
    auto row_len = [&](int r) { return L.seg_pre[r + 1] - L.seg_pre[r]; };
    auto fetch = [&](int r, int b, PReg& q) {
        if (b + pt < row_len(r)) {
            q.g = L.seg_start[r] + b + pt;
            q.c = s.cell[q.g];
            q.fx = s.fx[q.g];
            q.fy = s.fy[q.g];
            q.vx = s.vx[q.g];
            q.vy = s.vy[q.g];
            q.vz = s.vz[q.g];
        } else {
            q.g = 0;
            q.c = -1;
            q.fx = q.fy = q.vx = q.vy = q.vz = 0.0f;
        }
    };
    int cr = first_row, cb = 0;
    PReg cur = first;   // the loads of each window are issued at the end of the window before it
    while (cr < TILE_H) {
        const int rlen = row_len(cr);
        const int64_t rst = L.seg_start[cr];
        const bool ok = cb + pt < rlen;
        int nr = cr, nb = cb + PUSH_THREADS;
        if (nb >= rlen) {
            nr = cr + 1;
            nb = 0;
            while (nr < TILE_H && row_len(nr) == 0) nr++;
        }
        int32_t key = -1;
        float fxn = cur.fx, fyn = cur.fy, vxn = cur.vx, vyn = cur.vy, vzn = cur.vz;
        if (ok && cur.c >= 0) {
            double lc0 = cell_i(cur.c) + (double)cur.fx;
            double lc1 = cell_j(cur.c) + (double)cur.fy;
            double v0 = cur.vx, v1 = cur.vy, v2 = cur.vz;
            bool alive = push_particle<HAS_B>(m, md, pc, load_e, hasseg, lc0, lc1, v0, v1, v2);
            if (alive) {
                int ci, cj;
                cell_offsets(m, lc0, lc1, ci, cj, fxn, fyn);
                key = pack_cell(ci, cj);
                vxn = (float)v0;
                vyn = (float)v1;
                vzn = (float)v2;
                stat_add(st, pc, vxn, vyn, vzn);
                // histogram of the new cell
                int hi = ci - c_i0, hj = cj - c_j0;
                if (hi >= 0 && hi < TILE_CW && hj >= 0 && hj < TILE_CH) atomicAdd(&L.hist[hj * TILE_CW + hi], 1);
                else atomicAdd(&counts[cj * m.nci + ci], 1);
                // deposit from the stored (FP32) state; Field2D.scatter skips lc >= n-1
                double di = (double)fxn, dj = (double)fyn;
                if (!(ci + di >= m.ni - 1 || cj + dj >= m.nj - 1)) {
                    double val[NCH];
                    channel_values(pc, vxn, vyn, vzn, val);
                    int li = ci - n_i0, lj = cj - n_j0;
                    const double vm = vdev(pc, vxn, vyn, vzn);
                    if (vm >= pc.vfast) {
                        deposit_global_exact(m, acc, nn, ci, cj, di, dj, val);
                        atomicOr(flags, vm >= pc.vover ? (PUSH_FLAG_FAST | PUSH_FLAG_OVERFLOW) : PUSH_FLAG_FAST);
                    } else if (li >= 0 && li + 1 < TILE_NW && lj >= 0 && lj + 1 < TILE_NH) {
                        double w4[4];
                        scatter_weights(m, ci, cj, di, dj, w4);
                        int l0 = lj * TILE_NW + li;
                        // Lanes of one cell would all add to the same node in one instruction;
                        // each lane starts at corner (lane mod 4), so those lanes spread over the
                        // four corners and the same-address conflicts drop about fourfold.
                        const int rot = t & 3;
                        double wr[4];
                        int offr[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int cq = (q + rot) & 3;
                            wr[q] = cq == 0 ? w4[0] : (cq == 1 ? w4[1] : (cq == 2 ? w4[2] : w4[3]));
                            offr[q] = cq == 0 ? 0 : (cq == 1 ? 1 : (cq == 2 ? TILE_NW + 1 : TILE_NW));
                        }
#pragma unroll
                        for (int q = 0; q < 4; q++) {
#pragma unroll
                            for (int ch = 0; ch < NCH; ch++) {
                                long long v = quant(wr[q], val[ch]);
                                atomicAdd(&L.acc[ch][l0 + offr[q]], (unsigned long long)v);
                            }
                        }
                    } else {
                        deposit_global(m, acc, nn, ci, cj, di, dj, val);
                    }
                }
            }
        }
        if (ok) {
            L.o_cell[pt] = key;   // -1: dead, a hole until the next sort
            L.o_fx[pt] = fxn;
            L.o_fy[pt] = fyn;
            L.o_vx[pt] = vxn;
            L.o_vy[pt] = vyn;
            L.o_vz[pt] = vzn;
        }
        lds_barrier();
        const int wlen = min(PUSH_THREADS, rlen - cb);
        if (t < wlen) {
            const int64_t g = rst + cb + t;
            const int32_t kc = L.o_cell[t];
            s.cell[g] = kc;
            if (kc >= 0) {
                s.fx[g] = L.o_fx[t];
                s.fy[g] = L.o_fy[t];
                s.vx[g] = L.o_vx[t];
                s.vy[g] = L.o_vy[t];
                s.vz[g] = L.o_vz[t];
            }
        }
        if (nr < TILE_H) fetch(nr, nb, cur);
        lds_barrier();
        cr = nr;
        cb = nb;
    }
    __syncthreads();
    // flush LDS accumulators (nonzero entries) with 64-bit integer atomics
    for (int k = t; k < TILE_NODES; k += PUSH_THREADS) {
        int i = n_i0 + k % TILE_NW, j = n_j0 + k / TILE_NW;
        if (i < 0 || j < 0 || i >= m.ni || j >= m.nj) continue;
        size_t u = (size_t)j * m.ni + i;
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            unsigned long long v = L.acc[ch][k];
            if (v != 0ull) atomicAdd(acc + ch * nn + u, v);
        }
    }
    for (int k = t; k < TILE_CELLS; k += PUSH_THREADS) {
        int h = L.hist[k];
        if (h == 0) continue;
        int ci = c_i0 + k % TILE_CW, cj = c_j0 + k / TILE_CW;
        atomicAdd(&counts[cj * m.nci + ci], h);
    }
    stat_flush(st, stats);
}

// Four workgroups per CU (35 KB of LDS each, at most 128 VGPRs).
__global__ void __launch_bounds__(PUSH_THREADS, 4) push_deposit_tile_inplace(MeshP m, MeshD md, PushC pc, PState s,
                                                                           const int32_t* __restrict__ offsets,
                                                                           unsigned long long* __restrict__ acc,
                                                                           int32_t* __restrict__ counts,
                                                                           long long* stats, int* flags) {
    push_deposit_tile_body<false>(m, md, pc, s, offsets, acc, counts, stats, flags);
}
// The same pass with a magnetic field loaded (Boris where the gathered B is nonzero).
__global__ void __launch_bounds__(PUSH_THREADS) push_deposit_tile_inplace_b(MeshP m, MeshD md, PushC pc, PState s,
                                                                          const int32_t* __restrict__ offsets,
                                                                          unsigned long long* __restrict__ acc,
                                                                          int32_t* __restrict__ counts,
                                                                          long long* stats, int* flags) {
    push_deposit_tile_body<true>(m, md, pc, s, offsets, acc, counts, stats, flags);
}

// ----------------------------------------------------------------------------------------
// push of one particle with global memory only: the reference loop in FP64 (every boundary
// case), deposit and histogram with global atomics.  Used for the unsorted tail (particles added
// since the last sort) and for the particles the fast tile kernel leaves to it.
// ----------------------------------------------------------------------------------------
template <bool HAS_B>
__device__ __forceinline__ void push_one_global(const MeshP& m, const MeshD& md, const PushC& pc, const PState& s,
                                                int64_t p, unsigned long long* __restrict__ acc,
                                                int32_t* __restrict__ counts, int* flags, StatAcc& st) {
    const size_t nn = (size_t)m.ni * m.nj;
    auto load_e = [&](int i, int j) -> float2 { return md.ef[(size_t)j * m.ni + i]; };
    auto hasseg = [&](int i, int j) -> int { return md.node_hasseg[(size_t)j * m.ni + i]; };
    PIn q = load_particle(s, p);
    bool alive = push_particle<HAS_B>(m, md, pc, load_e, hasseg, q.lc0, q.lc1, q.v0, q.v1, q.v2);
    int32_t key = -1;
    if (alive) {
        int ci, cj;
        float fxn, fyn;
        cell_offsets(m, q.lc0, q.lc1, ci, cj, fxn, fyn);
        key = pack_cell(ci, cj);
        float vxn = (float)q.v0, vyn = (float)q.v1, vzn = (float)q.v2;
        stat_add(st, pc, vxn, vyn, vzn);
        atomicAdd(&counts[cj * m.nci + ci], 1);
        double di = fxn, dj = fyn;
        if (!(ci + di >= m.ni - 1 || cj + dj >= m.nj - 1)) {
            double val[NCH];
            channel_values(pc, vxn, vyn, vzn, val);
            const double vm = vdev(pc, vxn, vyn, vzn);
            if (vm >= pc.vfast) {
                deposit_global_exact(m, acc, nn, ci, cj, di, dj, val);
                atomicOr(flags, vm >= pc.vover ? (PUSH_FLAG_FAST | PUSH_FLAG_OVERFLOW) : PUSH_FLAG_FAST);
            } else {
                deposit_global(m, acc, nn, ci, cj, di, dj, val);
            }
        }
        s.fx[p] = fxn;
        s.fy[p] = fyn;
        s.vx[p] = vxn;
        s.vy[p] = vyn;
        s.vz[p] = vzn;
    }
    s.cell[p] = key;
}

template <bool HAS_B>
__global__ void __launch_bounds__(256) push_deposit_tail(MeshP m, MeshD md, PushC pc, PState s, int64_t p0,
                                                         int64_t p1, unsigned long long* __restrict__ acc,
                                                         int32_t* __restrict__ counts, long long* stats,
                                                         int* flags) {
    StatAcc st{0, 0, 0, 0, 0, 0.0};
    for (int64_t p = p0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < p1;
         p += (int64_t)gridDim.x * blockDim.x) {
        if (s.cell[p] < 0) continue;   // hole left by a dead particle
        push_one_global<HAS_B>(m, md, pc, s, p, acc, counts, flags, st);
    }
    stat_flush(st, stats);
}

// ----------------------------------------------------------------------------------------
// fast tile kernel (XY, no magnetic field): the push of every particle whose cell is away from
// segments and mesh faces, in FP32 relative to its cell; everything else goes to the queue.
//
// A particle of a cell without the slow flag that moves less than one cell in each direction
// (and is not faster than the fixed-point fast range) is exactly the reference loop's first and
// only sub-move: ProcessBoundary finds no segment in its bounding box and no face crossing, so
// the particle lives and its new position is old + v dt.  In cell units that is
// f' = f + v dt/dh with the new cell i + floor(f'), evaluated in FP32 (the stored precision).
// The deposit adds the raw bit patterns fma(w, value, 1.5 2^52) of the fixed-point rounding into
// LDS; every particle adds exactly one such pattern per channel to each corner of its new cell,
// so the flush subtracts (particles of the four adjacent cells) x bits(1.5 2^52), counted by the
// tile's cell histogram.  Integer sums modulo 2^64 keep this exact and order independent.
// ----------------------------------------------------------------------------------------
struct FastC {
    float qmdt, cdx, cdy;          // q/m dt, dt/dx, dt/dy
    double vfast;                  // fast range of the quantisation (deviation from c)
    double kn, kv, kvv, kst;       // fixed-point scales of the channels and of the mover sums
    double c0, c1, c2;             // reference velocity of the channels
    int ni, nj, nci, ncj, ntx;
};

struct FastLds {
    unsigned long long acc[NCH][TILE_NODES];
    int hist[TILE_CELLS];
    float2 ef[TILE_NODES];
    uint8_t slow[TILE_CELLS];
    int seg_start[TILE_H], seg_pre[TILE_H + 1];
    int nslow;
};

constexpr int32_t SLOT_KEEP = -2;   // marker: the slot is not written by the fast kernel
constexpr int SLOW_CHUNK = 256;     // slow-list particles per work item of push_tile_slow

__device__ __forceinline__ unsigned long long qbits(double w, double v) {
    return (unsigned long long)__double_as_longlong(fma(w, v, QMAGIC));
}

// Row segments of a tile's particles: row r of the tile (cells ti0 .. ti0+TILE_W-1 of row
// tj0 + r) holds slots [seg_start[r], seg_start[r] + len_r) of the cell-sorted array, and
// seg_pre[r] = len_0 + ... + len_{r-1}.  The concatenated segments number the tile's slots
// 0 .. seg_pre[TILE_H)-1; tile_slot maps such an index back to its slot.
__device__ __forceinline__ void tile_segments(const int32_t* __restrict__ offsets, int nci, int ncj, int ti0, int tj0,
                                              int* seg_start, int* seg_pre) {
    int pre = 0;
#pragma unroll
    for (int r = 0; r < TILE_H; r++) {
        const int row = tj0 + r;
        int st = 0, len = 0;
        if (row < ncj) {
            st = offsets[row * nci + ti0];
            len = offsets[row * nci + min(ti0 + TILE_W, nci)] - st;
        }
        seg_start[r] = st;
        seg_pre[r] = pre;
        pre += len;
    }
    seg_pre[TILE_H] = pre;
}
__device__ __forceinline__ int64_t tile_slot(const int* seg_start, const int* seg_pre, int k) {
    int r = 0;
#pragma unroll
    for (int q = 1; q < TILE_H; q++) r += k >= seg_pre[q] ? 1 : 0;
    return (int64_t)seg_start[r] + (k - seg_pre[r]);
}

// Flush of a tile's LDS accumulators (raw patterns fma(w, value, 1.5 2^52), one per channel
// and corner from every particle counted in hist): node (a, b) of the window received one
// pattern per channel from every particle of window cells (a, b), (a-1, b), (a, b-1),
// (a-1, b-1), so it subtracts that many bit patterns of 1.5 2^52 (integer sums modulo 2^64
// keep this exact and order independent).  The cell histogram goes to the global counts.
__device__ __forceinline__ void flush_tile(const unsigned long long (*lacc)[TILE_NODES], const int* hist, int w_i0,
                                           int w_j0, int ni, int nj, int nci, unsigned long long* __restrict__ acc,
                                           int32_t* __restrict__ counts, int nthreads) {
    const int t = threadIdx.x;
    const size_t nn = (size_t)ni * nj;
    const unsigned long long MAGIC_BITS = (unsigned long long)__double_as_longlong(QMAGIC);
    for (int k = t; k < TILE_NODES; k += nthreads) {
        const int a = k % TILE_NW, b = k / TILE_NW;
        const int i = w_i0 + a, j = w_j0 + b;
        if (i < 0 || j < 0 || i >= ni || j >= nj) continue;
        int np = 0;
        if (a < TILE_CW && b < TILE_CH) np += hist[b * TILE_CW + a];
        if (a > 0 && b < TILE_CH) np += hist[b * TILE_CW + a - 1];
        if (a < TILE_CW && b > 0) np += hist[(b - 1) * TILE_CW + a];
        if (a > 0 && b > 0) np += hist[(b - 1) * TILE_CW + a - 1];
        if (np == 0) continue;
        const unsigned long long bias = (unsigned long long)np * MAGIC_BITS;
        const size_t u = (size_t)j * ni + i;
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            const unsigned long long v = lacc[ch][k] - bias;
            if (v != 0ull) atomicAdd(acc + ch * nn + u, v);
        }
    }
    for (int k = t; k < TILE_CELLS; k += nthreads) {
        const int h = hist[k];
        if (h == 0) continue;
        const int ci = w_i0 + k % TILE_CW, cj = w_j0 + k / TILE_CW;
        atomicAdd(&counts[cj * nci + ci], h);
    }
}

__global__ void __launch_bounds__(PUSH_THREADS, 4) push_tile_fast_xy(FastC fc, const float2* __restrict__ ef,
                                                                    const uint8_t* __restrict__ cellslow, PState s,
                                                                    const int32_t* __restrict__ offsets,
                                                                    unsigned long long* __restrict__ acc,
                                                                    int32_t* __restrict__ counts,
                                                                    long long* __restrict__ stats,
                                                                    int32_t* __restrict__ slow_list,
                                                                    int32_t* __restrict__ tile_nslow,
                                                                    int2* __restrict__ slow_items,
                                                                    int* __restrict__ qcount) {
    __shared__ FastLds L;
    const int t = threadIdx.x;
    const int tile = blockIdx.x;
    const int tx = tile % fc.ntx, ty = tile / fc.ntx;
    const int ti0 = tx * TILE_W, tj0 = ty * TILE_H;
    // the LDS window: nodes and cells from (ti0 - HALO, tj0 - HALO); node (a, b) of the window is
    // corner (0, 0) of window cell (a, b)
    const int w_i0 = ti0 - TILE_HALO, w_j0 = tj0 - TILE_HALO;
    const size_t nn = (size_t)fc.ni * fc.nj;
    const int pt = (t * CHUNK_STRIDE) & (PUSH_THREADS - 1);
    const unsigned long long MAGIC_BITS = (unsigned long long)__double_as_longlong(QMAGIC);

    for (int k = t; k < NCH * TILE_NODES; k += PUSH_THREADS) (&L.acc[0][0])[k] = 0ull;
    if (t == 0) L.nslow = 0;
    for (int k = t; k < TILE_CELLS; k += PUSH_THREADS) {
        L.hist[k] = 0;
        const int i = w_i0 + k % TILE_CW, j = w_j0 + k / TILE_CW;
        L.slow[k] = (i >= 0 && j >= 0 && i < fc.nci && j < fc.ncj) ? cellslow[(size_t)j * fc.nci + i] : 1;
    }
    for (int k = t; k < TILE_NODES; k += PUSH_THREADS) {
        int i = w_i0 + k % TILE_NW, j = w_j0 + k / TILE_NW;
        i = min(max(i, 0), fc.ni - 1);
        j = min(max(j, 0), fc.nj - 1);
        L.ef[k] = ef[(size_t)j * fc.ni + i];
    }
    if (t == 0) tile_segments(offsets, fc.nci, fc.ncj, ti0, tj0, L.seg_start, L.seg_pre);
    __syncthreads();
    // The tile's particles, numbered 0 .. total-1 across its row segments, are taken in windows of
    // PUSH_THREADS; within a window thread t takes particle pt.  The waves run through the windows
    // independently (no barrier): each loads its next window's particles before pushing the
    // current ones and stores each pushed particle straight back to its slot.
    const int total = L.seg_pre[TILE_H];
    const int sg0 = L.seg_start[0], sg1 = L.seg_start[1], sg2 = L.seg_start[2], sg3 = L.seg_start[3];
    const int pr1 = L.seg_pre[1], pr2 = L.seg_pre[2], pr3 = L.seg_pre[3];
    // The slot and the loaded state are computed as values and assigned whole: selects or
    // conditional stores over the addresses of locals or struct fields would put them in
    // scratch memory, and its loads wait for every outstanding global store.
    const int d1 = sg1 - pr1 - sg0, d2 = (sg2 - pr2) - (sg1 - pr1), d3 = (sg3 - pr3) - (sg2 - pr2);
    auto slot_of = [sg0, d1, d2, d3, pr1, pr2, pr3](int k) -> int64_t {
        // segment starts minus prefixes telescope: pr1 <= pr2 <= pr3
        const int b = sg0 + ((k >= pr1 ? -1 : 0) & d1) + ((k >= pr2 ? -1 : 0) & d2) + ((k >= pr3 ? -1 : 0) & d3);
        return (int64_t)b + k;
    };
    auto load = [&](PReg& r, int k) {
        const bool in = k < total;
        const int64_t g = in ? slot_of(k) : 0;
        int32_t c = -1;
        float fx = 0.f, fy = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
        if (in) {
            c = s.cell[g];
            fx = s.fx[g];
            fy = s.fy[g];
            vx = s.vx[g];
            vy = s.vy[g];
            vz = s.vz[g];
        }
        r = PReg{g, c, fx, fy, vx, vy, vz};
    };
    // per-lane mover sums: raw FMA bit patterns, the bias removed with the lane's count at the end
    long long sn = 0;
    unsigned long long sp0 = 0, sp1 = 0, sp2 = 0, se = 0;
    double dmax = 0.0;
    PReg nxt;
    load(nxt, pt);
    for (int w0 = 0; w0 < total; w0 += PUSH_THREADS) {
        const PReg cur = nxt;
        const bool ok = w0 + pt < total;
        load(nxt, w0 + PUSH_THREADS + pt);
        int32_t key = SLOT_KEEP;
        float fxn = 0.f, fyn = 0.f, vxn = 0.f, vyn = 0.f, vzn = 0.f;
        bool slow = false;
        if (ok && cur.c >= 0) {
            const int ci = cur.c & 0xFFFF, cj = (cur.c >> 16) & 0xFFFF;
            const int hi = ci - w_i0, hj = cj - w_j0;
            bool fast = (unsigned)hi < (unsigned)TILE_CW && (unsigned)hj < (unsigned)TILE_CH;
            if (fast) fast = L.slow[hj * TILE_CW + hi] == 0;
            if (fast) {
                // Field2D.gather: bilinear in the cell (the slow flag keeps the cell off the last node line)
                const int l = hj * TILE_NW + hi;
                const float2 e00 = L.ef[l], e10 = L.ef[l + 1], e01 = L.ef[l + TILE_NW], e11 = L.ef[l + TILE_NW + 1];
                const float gx = 1.0f - cur.fx, gy = 1.0f - cur.fy;
                const float ex = gy * (gx * e00.x + cur.fx * e10.x) + cur.fy * (gx * e01.x + cur.fx * e11.x);
                const float ey = gy * (gx * e00.y + cur.fx * e10.y) + cur.fy * (gx * e01.y + cur.fx * e11.y);
                vxn = cur.vx + fc.qmdt * ex;
                vyn = cur.vy + fc.qmdt * ey;
                vzn = cur.vz;
                const float dx = vxn * fc.cdx, dy = vyn * fc.cdy;
                const double d0 = (double)vxn - fc.c0, d1 = (double)vyn - fc.c1, d2 = (double)vzn - fc.c2;
                const double dev = fmax(fabs(d0), fmax(fabs(d1), fabs(d2)));
                // |d| < 0.99 keeps f + d strictly inside (-1, 2) after FP32 rounding
                fast = fabsf(dx) < 0.99f && fabsf(dy) < 0.99f && dev < fc.vfast;
                if (fast) {
                    const float xn = cur.fx + dx, yn = cur.fy + dy;
                    const float flx = floorf(xn), fly = floorf(yn);
                    fxn = xn - flx;
                    fyn = yn - fly;
                    int nci = ci + (int)flx, ncj = cj + (int)fly;
                    // x' - floor(x') can round up to 1 for x' just below an integer: that position
                    // is offset 0 of the next cell (the stored cell is floor(lc))
                    if (fxn >= 1.0f) { fxn = 0.0f; nci++; }
                    if (fyn >= 1.0f) { fyn = 0.0f; ncj++; }
                    key = pack_cell(nci, ncj);
                    // mover sums (KineticMaterial.java:247-259, 406-413)
                    sn++;
                    sp0 += qbits(1.0, (double)vxn * fc.kst);
                    sp1 += qbits(1.0, (double)vyn * fc.kst);
                    sp2 += qbits(1.0, (double)vzn * fc.kst);
                    se += qbits(1.0, (double)sqrtf(vxn * vxn + vyn * vyn + vzn * vzn) * fc.kst);
                    dmax = fmax(dmax, dev);
                    // deposit of the 7 channels from the stored FP32 state (Field2D.scatter)
                    const double dfx = fxn, dfy = fyn;
                    const double w00 = (1.0 - dfx) * (1.0 - dfy), w10 = dfx * (1.0 - dfy), w11 = dfx * dfy,
                                 w01 = (1.0 - dfx) * dfy;
                    const double val[NCH] = {fc.kn, d0 * fc.kv, d1 * fc.kv, d2 * fc.kv, d0 * d0 * fc.kvv,
                                             d1 * d1 * fc.kvv, d2 * d2 * fc.kvv};
                    const int hn = nci - w_i0, hm = ncj - w_j0;
                    if ((unsigned)hn < (unsigned)TILE_CW && (unsigned)hm < (unsigned)TILE_CH) {
                        atomicAdd(&L.hist[hm * TILE_CW + hn], 1);
                        const int l0 = hm * TILE_NW + hn;
                        // lanes of one cell start at different corners (lane mod 4), so the
                        // same-address conflicts of an instruction drop about fourfold
                        const int rot = t & 3;
                        double wr[4];
                        int offr[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int cq = (q + rot) & 3;
                            wr[q] = cq == 0 ? w00 : (cq == 1 ? w10 : (cq == 2 ? w11 : w01));
                            offr[q] = cq == 0 ? 0 : (cq == 1 ? 1 : (cq == 2 ? TILE_NW + 1 : TILE_NW));
                        }
#pragma unroll
                        for (int q = 0; q < 4; q++)
#pragma unroll
                            for (int ch = 0; ch < NCH; ch++) atomicAdd(&L.acc[ch][l0 + offr[q]], qbits(wr[q], val[ch]));
                    } else {
                        atomicAdd(&counts[ncj * fc.nci + nci], 1);
                        const double w4[4] = {w00, w10, w11, w01};
                        const int ci4[4] = {0, 1, 1, 0}, cj4[4] = {0, 0, 1, 1};
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            const size_t u = (size_t)(ncj + cj4[c]) * fc.ni + (nci + ci4[c]);
#pragma unroll
                            for (int ch = 0; ch < NCH; ch++) {
                                const long long qv = quant(w4[c], val[ch]);
                                if (qv != 0) atomicAdd(acc + ch * nn + u, (unsigned long long)qv);
                            }
                        }
                    }
                }
            }
            slow = !fast;
        }
        // slow particles: the slot goes to the tile's slow list (stored in the tile's own slots
        // of slow_list); push_tile_slow pushes and writes it
        {
            const unsigned long long mk = __ballot(slow);
            if (mk) {
                const int lane = t & 63;
                const int leader = __ffsll((long long)mk) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&L.nslow, __popcll(mk));
                base = __shfl(base, leader);
                if (slow)
                    slow_list[tile_slot(L.seg_start, L.seg_pre, base + __popcll(mk & ((1ull << lane) - 1ull)))] =
                        (int32_t)cur.g;
            }
        }
        if (key != SLOT_KEEP) {
            s.cell[cur.g] = key;
            s.fx[cur.g] = fxn;
            s.fy[cur.g] = fyn;
            s.vx[cur.g] = vxn;
            s.vy[cur.g] = vyn;
            s.vz[cur.g] = vzn;
        }
    }
    __syncthreads();
    if (t == 0) {
        // the slow list goes to push_tile_slow as work items of SLOW_CHUNK particles
        const int n = L.nslow;
        tile_nslow[tile] = n;
        if (n) {
            atomicAdd(&qcount[0], n);
            const int nch = (n + SLOW_CHUNK - 1) / SLOW_CHUNK;
            const int base = atomicAdd(&qcount[1], nch);
            for (int c = 0; c < nch; c++) slow_items[base + c] = make_int2(tile, c);
        }
    }
    flush_tile(L.acc, L.hist, w_i0, w_j0, fc.ni, fc.nj, fc.nci, acc, counts, PUSH_THREADS);
    const unsigned long long sb = (unsigned long long)sn * MAGIC_BITS;
    StatAcc st{sn, (long long)(sp0 - sb), (long long)(sp1 - sb), (long long)(sp2 - sb), (long long)(se - sb), dmax};
    stat_flush(st, stats);
}

// The particles push_tile_fast_xy left to the reference loop (cells next to segments or mesh
// faces, moves of 0.99 cell or more, deviations beyond the fast quantisation range): the full
// FP64 push with ProcessBoundary and the deposit into the tile's LDS window (raw patterns and
// histogram as in the fast kernel), into global memory for particles that land outside it, or
// with the exact quantisation past the fast range.  The work is a list of (tile, chunk) items of
// SLOW_CHUNK particles each, taken by the workgroups in turn, so a tile with a long slow list
// (a row of cells along a mesh face, or the cells around an object) is spread over many of them.
struct SlowLds {
    unsigned long long acc[NCH][TILE_NODES];
    int hist[TILE_CELLS];
    int seg_start[TILE_H], seg_pre[TILE_H + 1];
};

__global__ void __launch_bounds__(SLOW_CHUNK) push_tile_slow(MeshP m, MeshD md, PushC pc, PState s,
                                                             const int32_t* __restrict__ offsets,
                                                             const int32_t* __restrict__ slow_list,
                                                             const int32_t* __restrict__ tile_nslow,
                                                             const int2* __restrict__ slow_items,
                                                             const int* __restrict__ qcount,
                                                             unsigned long long* __restrict__ acc,
                                                             int32_t* __restrict__ counts, long long* stats,
                                                             int* flags) {
    __shared__ SlowLds L;
    const int t = threadIdx.x;
    const int nitems = qcount[1];
    const size_t nn = (size_t)m.ni * m.nj;
    auto load_e = [&](int i, int j) -> float2 { return md.ef[(size_t)j * m.ni + i]; };
    auto hasseg = [&](int i, int j) -> int { return md.node_hasseg[(size_t)j * m.ni + i]; };
    StatAcc st{0, 0, 0, 0, 0, 0.0};
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int2 item = slow_items[it];
        const int tile = item.x;
        const int tx = tile % m.ntx, ty = tile / m.ntx;
        const int ti0 = tx * TILE_W, tj0 = ty * TILE_H;
        const int w_i0 = ti0 - TILE_HALO, w_j0 = tj0 - TILE_HALO;
        for (int k = t; k < NCH * TILE_NODES; k += SLOW_CHUNK) (&L.acc[0][0])[k] = 0ull;
        for (int k = t; k < TILE_CELLS; k += SLOW_CHUNK) L.hist[k] = 0;
        if (t == 0) tile_segments(offsets, m.nci, m.ncj, ti0, tj0, L.seg_start, L.seg_pre);
        __syncthreads();
        const int k = item.y * SLOW_CHUNK + t;
        if (k < tile_nslow[tile]) {
            const int64_t g = slow_list[tile_slot(L.seg_start, L.seg_pre, k)];
            PIn q = load_particle(s, g);
            const bool alive = push_particle<false>(m, md, pc, load_e, hasseg, q.lc0, q.lc1, q.v0, q.v1, q.v2);
            int32_t key = -1;
            if (alive) {
                int ci, cj;
                float fxn, fyn;
                cell_offsets(m, q.lc0, q.lc1, ci, cj, fxn, fyn);
                key = pack_cell(ci, cj);
                const float vxn = (float)q.v0, vyn = (float)q.v1, vzn = (float)q.v2;
                stat_add(st, pc, vxn, vyn, vzn);
                const double di = fxn, dj = fyn;
                const int hn = ci - w_i0, hm = cj - w_j0;
                bool counted = false;
                if (!(ci + di >= m.ni - 1 || cj + dj >= m.nj - 1)) {
                    double val[NCH];
                    channel_values(pc, vxn, vyn, vzn, val);
                    const double vm = vdev(pc, vxn, vyn, vzn);
                    if (vm >= pc.vfast) {
                        deposit_global_exact(m, acc, nn, ci, cj, di, dj, val);
                        atomicOr(flags, vm >= pc.vover ? (PUSH_FLAG_FAST | PUSH_FLAG_OVERFLOW) : PUSH_FLAG_FAST);
                    } else if ((unsigned)hn < (unsigned)TILE_CW && (unsigned)hm < (unsigned)TILE_CH) {
                        double w4[4];
                        scatter_weights(m, ci, cj, di, dj, w4);
                        const int l0 = hm * TILE_NW + hn;
                        const int off4[4] = {0, 1, TILE_NW + 1, TILE_NW};
#pragma unroll
                        for (int c = 0; c < 4; c++)
#pragma unroll
                            for (int ch = 0; ch < NCH; ch++)
                                atomicAdd(&L.acc[ch][l0 + off4[c]], qbits(w4[c], val[ch]));
                        atomicAdd(&L.hist[hm * TILE_CW + hn], 1);
                        counted = true;
                    } else {
                        deposit_global(m, acc, nn, ci, cj, di, dj, val);
                    }
                }
                if (!counted) atomicAdd(&counts[cj * m.nci + ci], 1);
                s.fx[g] = fxn;
                s.fy[g] = fyn;
                s.vx[g] = vxn;
                s.vy[g] = vyn;
                s.vz[g] = vzn;
            }
            s.cell[g] = key;
        }
        __syncthreads();
        flush_tile(L.acc, L.hist, w_i0, w_j0, m.ni, m.nj, m.nci, acc, counts, SLOW_CHUNK);
        __syncthreads();   // the next item reuses the LDS window
    }
    stat_flush(st, stats);
}

// cell counts of an in-place step only feed the macroparticle-per-cell samples
__global__ void mpc_accumulate(int n, int32_t* __restrict__ counts, long long* __restrict__ mpc_sum) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        mpc_sum[c] += counts[c];
        counts[c] = 0;
    }
}

// ----------------------------------------------------------------------------------------
// counting-sort scan over cells: counts -> offsets (exclusive, ncells+1), cursor = offsets,
// mpc-sum += counts, counts = 0
// ----------------------------------------------------------------------------------------
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& total) {
    // wave inclusive scan
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int nw = blockDim.x >> 6;
        int w = lane < nw ? sh[lane] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(w, o);
            if (lane >= o) w += y;
        }
        if (lane < nw) sh[lane] = w;
    }
    __syncthreads();
    int wave_prefix = wid > 0 ? sh[wid - 1] : 0;
    total = sh[(blockDim.x >> 6) - 1];
    __syncthreads();
    return wave_prefix + x - v;
}

__global__ void __launch_bounds__(SCAN_THREADS) scan_reduce(const int32_t* __restrict__ counts, int n,
                                                            int32_t* __restrict__ blocksum) {
    __shared__ int sh[32];
    int base = blockIdx.x * SCAN_TILE;
    int s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int idx = base + threadIdx.x * SCAN_ITEMS + k;
        if (idx < n) s += counts[idx];
    }
    int total;
    block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_THREADS) scan_blocks(int32_t* __restrict__ blocksum, int nb) {
    __shared__ int sh[32];
    int carry = 0;
    for (int base = 0; base < nb; base += SCAN_THREADS) {
        int idx = base + threadIdx.x;
        int v = idx < nb ? blocksum[idx] : 0;
        int total;
        int ex = block_exclusive_scan(v, sh, total);
        if (idx < nb) blocksum[idx] = ex + carry;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) blocksum[nb] = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS) scan_write(int32_t* __restrict__ counts, int n,
                                                           const int32_t* __restrict__ blocksum,
                                                           int32_t* __restrict__ offsets, int32_t* __restrict__ cursor,
                                                           long long* __restrict__ mpc_sum, int nb) {
    __shared__ int sh[32];
    int base = blockIdx.x * SCAN_TILE;
    int v[SCAN_ITEMS];
    int s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int idx = base + threadIdx.x * SCAN_ITEMS + k;
        v[k] = idx < n ? counts[idx] : 0;
        s += v[k];
    }
    int total;
    int ex = block_exclusive_scan(s, sh, total) + blocksum[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int idx = base + threadIdx.x * SCAN_ITEMS + k;
        if (idx < n) {
            offsets[idx] = ex;
            cursor[idx] = ex;
            if (mpc_sum) mpc_sum[idx] += v[k];
            counts[idx] = 0;
        }
        ex += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) offsets[n] = blocksum[nb];
}

// ----------------------------------------------------------------------------------------
// sort steps: each pushed particle to its slot in the new cell order
// ----------------------------------------------------------------------------------------
// Slot of each valid lane in its cell: the lanes of one key are grouped by ballots, the
// lowest lane of each group reserves the group's slots with one atomic on the cell cursor, and
// the other lanes take consecutive slots after it.  The atomics of all groups of the wave are
// issued together, so the wave waits for one round trip however many cells it holds.
__device__ __forceinline__ int64_t wave_rank(bool valid, int key, int32_t* cursor) {
    const int lane = threadIdx.x & 63;
    unsigned long long active = __ballot(valid);
    int my_leader = lane;
    unsigned long long my_group = 0ull;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const int lk = __shfl(key, leader);
        const unsigned long long mm = __ballot(valid && key == lk);
        if (valid && key == lk) {
            my_leader = leader;
            my_group = mm;
        }
        active &= ~mm;
    }
    int b = 0;
    if (valid && lane == my_leader) b = atomicAdd(&cursor[key], __popcll(my_group));
    b = __shfl(b, my_leader);
    return valid ? b + __popcll(my_group & ((1ull << lane) - 1ull)) : -1;
}

// sort without a push (after particles were loaded, before DSMC): key = current cell.  In a
// cell-ordered array the lanes of a wave hold few cells: up to 4 groups of lanes with one cell
// each add their count with one atomic by their lowest lane; the lanes left (an unordered array)
// add their own.
__global__ void __launch_bounds__(256) count_current(MeshP m, PState s, int64_t np, int32_t* __restrict__ keys,
                                                     int32_t* __restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nloop = (np + stride - 1) / stride * stride;   // the lanes of a wave iterate together
    const int lane = threadIdx.x & 63;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nloop; p += stride) {
        const bool valid = p < np;
        const int32_t c = valid ? s.cell[p] : -1;
        if (valid) keys[p] = c;
        const int lk = c >= 0 ? cell_j(c) * m.nci + cell_i(c) : -1;
        unsigned long long rem = __ballot(lk >= 0);
        for (int g = 0; g < 4 && rem; g++) {
            const int leader = __ffsll((long long)rem) - 1;
            const int key = __shfl(lk, leader);
            const unsigned long long grp = __ballot(lk == key);
            if (lane == leader) atomicAdd(&counts[key], __popcll(grp));
            rem &= ~grp;
        }
        if ((rem >> lane) & 1ull) atomicAdd(&counts[lk], 1);
    }
}

template <bool IDS>
__global__ void __launch_bounds__(256) copy_scatter(MeshP m, PState s, PState d, int64_t np,
                                                    const int32_t* __restrict__ keys, int32_t* __restrict__ cursor) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nloop = (np + stride - 1) / stride * stride;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nloop; p += stride) {
        bool valid = p < np;
        int32_t key = valid ? keys[p] : -1;
        valid = valid && key >= 0;
        int lkey = valid ? cell_j(key) * m.nci + cell_i(key) : 0;
        // the state is loaded before the slot is known, so the loads overlap the cursor atomics
        float fx = 0, fy = 0, vx = 0, vy = 0, vz = 0;
        int32_t idv = 0;
        if (valid) {
            fx = s.fx[p];
            fy = s.fy[p];
            vx = s.vx[p];
            vy = s.vy[p];
            vz = s.vz[p];
            if (IDS) idv = s.id[p];
        }
        int64_t dst = wave_rank(valid, lkey, cursor);
        if (valid) {
            d.cell[dst] = key;
            d.fx[dst] = fx;
            d.fy[dst] = fy;
            d.vx[dst] = vx;
            d.vy[dst] = vy;
            d.vz[dst] = vz;
            if (IDS) d.id[dst] = idv;
        }
    }
}

// ----------------------------------------------------------------------------------------
// node-wise normalisation (KineticMaterial.updateFields tail, Field2D.divideByField and
// scaleByVol, Material.update nd_back) and sample accumulation (updateSamples)
// ----------------------------------------------------------------------------------------
// The running sums are FP64 in physical units (mpw-weighted): each step adds its exact integer
// totals, converted once, so they are deterministic and cannot wrap however long sampling runs.
// fs_n, fs_v, fs_vv are spwt / 2^k of the density, velocity and second-moment channels.
// The channels hold W = sum w, D_k = sum w (v_k - c_k) and Q_k = sum w (v_k - c_k)^2 (unit
// weight); the reference's sums are W, D_k + c_k W and Q_k + 2 c_k D_k + c_k^2 W, times mpw.
__global__ void __launch_bounds__(256) normalize_nodes(int64_t nn, unsigned long long* __restrict__ acc,
                                                       double* __restrict__ sums, const double* __restrict__ vol,
                                                       double* __restrict__ den, double inv_kw, double inv_kv,
                                                       double inv_kvv, double c0, double c1, double c2, double spwt,
                                                       double nd_back, int samples) {
    const double c[3] = {c0, c1, c2};
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        long long a[NCH];
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            a[ch] = (long long)acc[ch * nn + u];
            acc[ch * nn + u] = 0ull;
        }
        const double W = (double)a[0] * inv_kw;
        double D[3], Q[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            D[k] = (double)a[1 + k] * inv_kv;
            Q[k] = (double)a[4 + k] * inv_kvv;
        }
        if (samples) {
            sums[u] += W * spwt;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                sums[(1 + k) * nn + u] += (D[k] + c[k] * W) * spwt;
                sums[(4 + k) * nn + u] += (Q[k] + 2.0 * c[k] * D[k] + c[k] * c[k] * W) * spwt;
            }
        }
        const double nd = W * spwt / vol[u];
        double vel[3] = {0, 0, 0};
        if (a[0] != 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) vel[k] = D[k] / W + c[k];
        }
        den[u] = nd + nd_back;
        den[nn + u] = vel[0];
        den[2 * nn + u] = vel[1];
        den[3 * nn + u] = vel[2];
    }
}

// KineticMaterial.computeFields(MeshData) (core/materials/KineticMaterial.java:1606-1671) over
// the running sums, in the reference's order of operations: temperatures where count-sum > 0
// (init_vals.T elsewhere), nd-ave = count-sum * (1/num_samples) / nodeVol, u/v/w-ave = sum /
// count-sum (0 where the count is 0, Field2D.divideByField), p = nd-ave K T, and mpc =
// mpc-sum * (1/num_samples) for the cells (row j*ni + i of cell (i, j), 0 on the last node row
// and column).  gas: [NGAS][nn] t, t1, t2, t3, nd-ave, u-ave, v-ave, w-ave, p, mpc.
__global__ void __launch_bounds__(256) compute_fields_kernel(MeshP m, const double* __restrict__ sums,
                                                             const long long* __restrict__ mpc_sum,
                                                             const double* __restrict__ vol,
                                                             double* __restrict__ gas, double f, double init_T,
                                                             double inv_ns) {
    const int64_t nn = (int64_t)m.ni * m.nj;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        const double count = sums[u];
        double T = init_T, T1 = init_T, T2 = init_T, T3 = init_T;
        if (count > 0) {
            const double uu = sums[nn + u] / count, vv = sums[2 * nn + u] / count, ww = sums[3 * nn + u] / count;
            const double uu2 = sums[4 * nn + u] / count - uu * uu;
            const double vv2 = sums[5 * nn + u] / count - vv * vv;
            const double ww2 = sums[6 * nn + u] / count - ww * ww;
            T = (uu2 + vv2 + ww2) * f / 3.0;
            T1 = uu2 * f;
            T2 = vv2 * f;
            T3 = ww2 * f;
        }
        const double nd_ave = count * inv_ns / vol[u];
        gas[u] = T;
        gas[nn + u] = T1;
        gas[2 * nn + u] = T2;
        gas[3 * nn + u] = T3;
        gas[4 * nn + u] = nd_ave;
        gas[5 * nn + u] = count != 0 ? sums[nn + u] / count : 0.0;
        gas[6 * nn + u] = count != 0 ? sums[2 * nn + u] / count : 0.0;
        gas[7 * nn + u] = count != 0 ? sums[3 * nn + u] / count : 0.0;
        gas[8 * nn + u] = nd_ave * K_B * T;
        const int i = (int)(u % m.ni), j = (int)(u / m.ni);
        gas[9 * nn + u] = (i < m.nci && j < m.ncj) ? (double)mpc_sum[(size_t)j * m.nci + i] * inv_ns : 0.0;
    }
}

// ----------------------------------------------------------------------------------------
// particle addition (KineticMaterial.addParticle) and device preload
// ----------------------------------------------------------------------------------------
// KineticMaterial.addParticle of one accepted particle into slot p of the state (XtoL, clamp,
// half-step rewind with the current E when `rewind`, FP32 storage).
__device__ __forceinline__ void add_one(const MeshP& m, const MeshD& md, const PState& d, int64_t p, double x,
                                        double y, double u, double v, double w, double qm, double qm_half_dt,
                                        double half_dt, int rewind) {
    double lc0 = (x - m.x0[0]) * m.inv_dh[0];
    double lc1 = (y - m.x0[1]) * m.inv_dh[1];
    if (rewind) {
        if (lc0 >= m.ni) lc0 = m.ni - 1;
        if (lc1 >= m.nj) lc1 = m.nj - 1;
        // part.dt = -dt/2 (KineticMaterial.java:777-794)
        auto load_e = [&](int i, int j) -> float2 { return md.ef[(size_t)j * m.ni + i]; };
        if (md.bf) kick<true>(m, md, qm, qm_half_dt, half_dt, load_e, lc0, lc1, u, v, w);
        else kick<false>(m, md, qm, qm_half_dt, half_dt, load_e, lc0, lc1, u, v, w);
    }
    int ci = (int)lc0, cj = (int)lc1;
    ci = min(max(ci, 0), m.nci - 1);
    cj = min(max(cj, 0), m.ncj - 1);
    d.cell[p] = pack_cell(ci, cj);
    d.fx[p] = (float)(lc0 - ci);
    d.fy[p] = (float)(lc1 - cj);
    d.vx[p] = (float)u;
    d.vy[p] = (float)v;
    d.vz[p] = (float)w;
}

// `in` holds the columns x, y, u, v, w, `ld` doubles apart (ld = n for a packed [5][n] block).
__global__ void __launch_bounds__(256) add_particles_kernel_strided(MeshP m, MeshD md, PState d, int64_t base,
                                                                    int64_t n, const double* __restrict__ in,
                                                                    int64_t ld, double qm, double qm_half_dt,
                                                                    double half_dt, int rewind,
                                                                    const int32_t* __restrict__ ids, int id0,
                                                                    int with_ids) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = base + k;
        add_one(m, md, d, p, in[k], in[ld + k], in[2 * ld + k], in[3 * ld + k], in[4 * ld + k], qm, qm_half_dt,
                half_dt, rewind);
        if (with_ids) d.id[p] = ids ? ids[k] : id0 + (int)k;
    }
}

// MeshData.addParticle's acceptance (KineticMaterial.java:1356-1361: finite velocity) and
// DomainModule.getMesh's (DomainModule.java:106-117: inside the mesh, strictly or within
// FLT_EPS in logical coordinates), the same tests as the host filter of add_particles_device.
__device__ __forceinline__ bool accept_particle(const MeshP& m, double xd0, double xd1, double x, double y, double u,
                                                double v, double w) {
    if (!(isfinite(u) && isfinite(v) && isfinite(w))) return false;
    const double lc0 = (x - m.x0[0]) / m.dh[0], lc1 = (y - m.x0[1]) / m.dh[1];
    const bool strict = x >= m.x0[0] && x < xd0 && y >= m.x0[1] && y < xd1;
    const bool loose = !(lc0 < -FLT_EPS || lc1 < -FLT_EPS || lc0 > (m.ni - 1 + FLT_EPS) || lc1 > (m.nj - 1 + FLT_EPS));
    return strict || loose;
}

constexpr int FILTER_THREADS = 1024;

// Acceptance and addParticle of n <= FILTER_THREADS * 64 particles in one workgroup.  Thread t
// takes the contiguous chunk [t c, (t + 1) c) of the input; the workgroup's exclusive scan of the
// accepted counts numbers every accepted particle by its rank among the accepted ones in input
// order, as the host filter does, so slots and new ids are those add_particles_device gives.
// `in` holds the columns x, y, u, v, w, `ld` doubles apart; the accepted count goes to *kept.
__global__ void __launch_bounds__(FILTER_THREADS) add_particles_filtered(MeshP m, MeshD md, PState d, int64_t base,
                                                                         int n, const double* __restrict__ in,
                                                                         int64_t ld, double xd0, double xd1, double qm,
                                                                         double qm_half_dt, double half_dt, int rewind,
                                                                         const int32_t* __restrict__ ids, int id0,
                                                                         int with_ids, int* __restrict__ kept) {
    __shared__ int wave_tot[FILTER_THREADS / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int c = (n + FILTER_THREADS - 1) / FILTER_THREADS;
    const int k0 = min(n, t * c), k1 = min(n, k0 + c);
    auto take = [&](int k) {
        return accept_particle(m, xd0, xd1, in[k], in[ld + k], in[2 * ld + k], in[3 * ld + k], in[4 * ld + k]);
    };
    int cnt = 0;
    for (int k = k0; k < k1; k++) cnt += take(k) ? 1 : 0;
    // inclusive scan over the wave, then over the wave totals
    int inc = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(inc, off);
        if (lane >= off) inc += o;
    }
    if (lane == 63) wave_tot[wv] = inc;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < wv; q++) before += wave_tot[q];
    int r = before + inc - cnt;   // accepted particles ahead of this chunk
    for (int k = k0; k < k1; k++) {
        if (!take(k)) continue;
        const int64_t p = base + r;
        add_one(m, md, d, p, in[k], in[ld + k], in[2 * ld + k], in[3 * ld + k], in[4 * ld + k], qm, qm_half_dt,
                half_dt, rewind);
        if (with_ids) d.id[p] = ids ? ids[k] : id0 + r;
        r++;
    }
    if (t == FILTER_THREADS - 1) *kept = r;
}

// MaxwellianSource.sampleParticle + Source.sampleKinetic + KineticMaterial.addParticle for one
// batch of a boundary source, sampled on the device (sources/MaxwellianSource.java:93-125,
// core/source/Source.java:123-198, Spline.randomT for a polyline, Utils.SampleMaxw3D with its
// 1.2 factor).  Particle k of the batch draws from its own Philox stream (id k, the injection
// step, the source tag): draw 0 places it along the boundary; attempt a of the velocity uses
// draws 2 + 38 a .. 37 + 38 a (12 per component), repeated while the velocity does not point
// into the domain.  Particles outside the mesh, or still without an inward velocity after 64
// attempts (a source whose drift points out of the domain), become holes (cell -1).
__global__ void __launch_bounds__(256) inject_maxwellian_kernel(MeshP m, MeshD md, PState d, int64_t base, int n,
                                                                const double* __restrict__ geom, int nseg, double area,
                                                                double v_drift, double v_th, double v_az, double qm,
                                                                double qm_half_dt, double half_dt, double dt,
                                                                uint32_t k0, uint32_t k1, uint32_t step, uint32_t tag,
                                                                int with_ids, int id0,
                                                                unsigned long long* __restrict__ accepted) {
    const double* seg = geom;
    const double* cum = geom + 4 * nseg;
    unsigned long long n_ok = 0;   // added to accepted once per wave
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        ParticleRng rng{k0, k1, (uint32_t)k, step, tag};
        // Spline.randomT (XY): uniform in length, binary search over the cumulative lengths
        const double A1 = rng.draw(0) * area;
        int i1 = 0, i2 = nseg + 1, si;
        if (A1 < cum[0]) si = -1;
        else if (A1 > cum[nseg]) si = nseg + 1;
        else {
            while (true) {
                const int mid = (int)(0.5 * (i1 + i2));
                if (A1 < cum[mid]) i2 = mid;
                else if (A1 > cum[mid]) i1 = mid;
                else { i1 = mid; break; }
                if ((i2 - i1) <= 1) break;
            }
            si = i1;
        }
        si = min(max(si, 0), nseg - 1);
        const double frac = (A1 - cum[si]) / (cum[si + 1] - cum[si]);
        double t = si + frac;
        int ks = (int)t;
        double st = t - ks;
        if (ks > nseg - 1) { ks = nseg - 1; st = 1.0; }
        const double* g = seg + 4 * ks;
        const double px = g[0] + st * (g[2] - g[0]), py = g[1] + st * (g[3] - g[1]);
        const double ddx = g[2] - g[0], ddy = g[3] - g[1], len = sqrt(ddx * ddx + ddy * ddy);
        const double n0 = -ddy / len, n1 = ddx / len;
        double vel[3] = {0, 0, 0};
        const double sh = sqrt(0.5);
        bool inward = false;
        for (int a = 0; a < 64 && !inward; a++) {   // the reference repeats without a limit
            const uint32_t q = 2 + 38 * a;
            double vm[3];
            for (int c = 0; c < 3; c++) vm[c] = (v_th * 1.2) * (sh * 1.0 * ((rng.sum12(q + 12 * c) - 0.5 * 12) * 1.0));
            vel[0] = vm[0] + n0 * v_drift;
            vel[1] = vm[1] + n1 * v_drift;
            vel[2] = vm[2] + v_az;
            inward = n0 * vel[0] + n1 * vel[1] > 0;
        }
        // Source.sampleKinetic nudge, then KineticMaterial.addParticle
        const double x = px + vel[0] * 1e-6 * dt, y = py + vel[1] * 1e-6 * dt;
        const int64_t p = base + k;
        double lc0 = (x - m.x0[0]) * m.inv_dh[0];
        double lc1 = (y - m.x0[1]) * m.inv_dh[1];
        const bool inside = !(lc0 < -FLT_EPS || lc1 < -FLT_EPS || lc0 > (m.ni - 1 + FLT_EPS) || lc1 > (m.nj - 1 + FLT_EPS));
        if (!inward || !inside || !(isfinite(vel[0]) && isfinite(vel[1]) && isfinite(vel[2]))) {
            d.cell[p] = -1;
            continue;
        }
        if (lc0 >= m.ni) lc0 = m.ni - 1;
        if (lc1 >= m.nj) lc1 = m.nj - 1;
        auto load_e = [&](int i, int j) -> float2 { return md.ef[(size_t)j * m.ni + i]; };
        if (md.bf) kick<true>(m, md, qm, qm_half_dt, half_dt, load_e, lc0, lc1, vel[0], vel[1], vel[2]);
        else kick<false>(m, md, qm, qm_half_dt, half_dt, load_e, lc0, lc1, vel[0], vel[1], vel[2]);
        int ci = (int)lc0, cj = (int)lc1;
        ci = min(max(ci, 0), m.nci - 1);
        cj = min(max(cj, 0), m.ncj - 1);
        d.cell[p] = pack_cell(ci, cj);
        d.fx[p] = (float)(lc0 - ci);
        d.fy[p] = (float)(lc1 - cj);
        d.vx[p] = (float)vel[0];
        d.vy[p] = (float)vel[1];
        d.vz[p] = (float)vel[2];
        if (with_ids) d.id[p] = id0 + k;
        n_ok++;
    }
    for (int o = 32; o > 0; o >>= 1) n_ok += __shfl_xor(n_ok, o);
    if ((threadIdx.x & 63) == 0 && n_ok) atomicAdd(accepted, n_ok);
}

int inject_maxwellian_device(sf_ctx* ctx, HostSource& s, int num, uint64_t seed, uint32_t step, uint32_t tag,
                             int64_t* added) {
    *added = 0;
    if (num <= 0) return SF_OK;
    Species* sp = ctx->species[s.species];
    if (sp->np + num > sp->capacity)
        return sf_fail(ctx, SF_ERR_OOM, "species capacity %lld exceeded (have %lld, injecting %d)",
                       (long long)sp->capacity, (long long)sp->np, num);
    const int nseg = (int)s.seg.size() / 4;
    if (!s.d_geom) {
        std::vector<double> g(s.seg);
        g.insert(g.end(), s.cum.begin(), s.cum.end());
        SF_HIP(ctx, hipMalloc((void**)&s.d_geom, g.size() * sizeof(double)));
        SF_HIP(ctx, hipMemcpy(s.d_geom, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    unsigned long long* d_acc = nullptr;
    SF_HIP(ctx, hipMallocAsync((void**)&d_acc, sizeof(unsigned long long), ctx->stream));
    SF_HIP(ctx, hipMemsetAsync(d_acc, 0, sizeof(unsigned long long), ctx->stream));
    hipLaunchKernelGGL(inject_maxwellian_kernel, dim3(std::min(4096, (num + 255) / 256)), dim3(256), 0, ctx->stream, ctx->mp,
                       ctx->meshd(), sp->buf[sp->cur], sp->np, num, (const double*)s.d_geom, nseg, s.area, s.v_drift,
                       s.v_th, s.v_az, sp->qm, -0.5 * ctx->dt * sp->qm, -0.5 * ctx->dt, ctx->dt, (uint32_t)seed,
                       (uint32_t)(seed >> 32), step, tag, sp->desc.track_ids, sp->id_counter, d_acc);
    SF_HIP(ctx, hipGetLastError());
    unsigned long long acc = 0;
    SF_HIP(ctx, hipMemcpyAsync(&acc, d_acc, sizeof(acc), hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipFreeAsync(d_acc, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sp->id_counter += num;
    sp->np += num;             // rejected slots are holes until the next sort
    sp->nalive += (int64_t)acc;
    *added = (int64_t)acc;
    return SF_OK;
}

// counter-based generator for the preload: splitmix64 finaliser over (seed, particle, draw)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t p, uint32_t k) {
    uint64_t h = mix64(seed ^ mix64(p * 0x100000001B3ULL + k));
    return (double)(h >> 11) * 0x1.0p-53;
}

__global__ void __launch_bounds__(256) preload_kernel(MeshP m, PState d, int64_t base, int64_t n,
                                                      const int32_t* __restrict__ open_cells,
                                                      const double* __restrict__ cdf, int nopen, double dx_,
                                                      double dy_, double dz_, double vth, uint64_t seed,
                                                      int with_ids, int id0) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        double u = u01(seed, k, 0) * cdf[nopen];
        int lo = 0, hi = nopen - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (cdf[mid + 1] <= u) lo = mid + 1;
            else hi = mid;
        }
        int32_t c = open_cells[lo];
        int ci = c & 0xFFFF, cj = c >> 16;
        double fx = u01(seed, k, 1), fy = u01(seed, k, 2);
        if (m.domain == SF_DOMAIN_RZ) {
            // uniform in volume: r^2 uniform between the cell faces
            double r1 = m.x0[0] + ci * m.dh[0], r2 = r1 + m.dh[0];
            double r = sqrt(r1 * r1 + fx * (r2 * r2 - r1 * r1));
            fx = (r - r1) * m.inv_dh[0];
        }
        double vel[3];
        for (int a = 0; a < 3; a++) {     // Utils.SampleMaxw1D with M = 12, widened by 1.2
            double s = 0;
            for (int q = 0; q < 12; q++) s += u01(seed, k, 3 + a * 12 + q);
            vel[a] = (vth * 1.2) * sqrt(0.5) * (s - 6.0);
        }
        int64_t p = base + k;
        d.cell[p] = pack_cell(ci, cj);
        d.fx[p] = (float)fx;
        d.fy[p] = (float)fy;
        d.vx[p] = (float)(vel[0] + dx_);
        d.vy[p] = (float)(vel[1] + dy_);
        d.vz[p] = (float)(vel[2] + dz_);
        if (with_ids) d.id[p] = id0 + (int)k;
    }
}

__global__ void ef32_kernel(int64_t nn, const double* __restrict__ efi, const double* __restrict__ efj,
                            float2* __restrict__ ef) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x)
        ef[u] = make_float2((float)efi[u], (float)efj[u]);
}

// ----------------------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------------------
static int grid_for(int64_t n, int threads, int max_blocks = 65536) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (int)b;
}

// Channel scales of a step: a particle whose velocity components deviate from c by at most B
// contributes at most 2^46 to a velocity channel (a corner weight is at most 1), so a node holds
// 2^17 such contributions per step before its int64 total wraps (a contribution's typical size
// is a few percent of that).  A temperature is <(v-c)^2> - <v-c>^2 over a node's weight W, and
// the absolute quantum 2^-k of a contribution enters it as 2^-k / W: the scales are as fine as
// that margin allows, since W can be a small fraction of one particle's weight.
static void set_channel_scales(Species* sp, const double c[3], double B) {
    B = std::max(B, 1.0);
    const int kv = (int)std::floor(std::log2(std::ldexp(1.0, 46) / B));
    const int kvv = (int)std::floor(std::log2(std::ldexp(1.0, 46) / (B * B)));
    sp->kscale[0] = std::ldexp(1.0, 40);
    for (int k = 1; k <= 3; k++) sp->kscale[k] = std::ldexp(1.0, kv);
    for (int k = 4; k <= 6; k++) sp->kscale[k] = std::ldexp(1.0, kvv);
    for (int k = 0; k < 3; k++) sp->vref[k] = c[k];
}

static PushC make_pushc(sf_ctx* ctx, Species* sp) {
    PushC pc;
    pc.qm_dt = sp->qm * ctx->dt;
    pc.qm = sp->qm;
    pc.dt = ctx->dt;
    pc.kn = sp->kscale[0];
    pc.kv = sp->kscale[1];    // = kscale[2] = kscale[3]
    pc.kvv = sp->kscale[4];   // = kscale[5] = kscale[6]
    pc.spwt = sp->desc.spwt;
    pc.c0 = sp->vref[0];
    pc.c1 = sp->vref[1];
    pc.c2 = sp->vref[2];
    pc.kst = sp->kstat;
    // A contribution is at most d kv or d^2 kvv for a deviation d: the one-FMA rounding holds
    // below 2^51 and the fast path stops at 2^50 (a deviation of 16 B in the velocity channels,
    // 4 B in the second moments); beyond 2^62 a contribution does not fit an int64 with margin.
    // The mover sums round |v| kst <= sqrt(3) max|v_c| kst with the same FMA.
    pc.vfast = std::min(std::ldexp(1.0, 50) / pc.kv, std::sqrt(std::ldexp(1.0, 50) / pc.kvv));
    pc.vover = std::min(std::sqrt(std::ldexp(1.0, 62) / pc.kvv), std::ldexp(1.0, 62) / pc.kv);
    return pc;
}

// Mover sums and largest deviation from c of the particles in the array (holes skipped): the
// first push of a species (or the first after a load) takes its reference velocity and scales
// from them.
__global__ void __launch_bounds__(256) velocity_range_kernel(PState s, int64_t n, PushC pc, long long* stats) {
    StatAcc st{0, 0, 0, 0, 0, 0.0};
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (s.cell[p] < 0) continue;
        stat_add(st, pc, s.vx[p], s.vy[p], s.vz[p]);
    }
    stat_flush(st, stats);
}

// Adds the STAT_SLOTS copies of the mover sums into out[0..4] and returns the largest deviation.
static double combine_stat_slots(const long long* slots, long long* out) {
    for (int q = 0; q < 5; q++) out[q] = 0;
    double dmax = 0;
    for (int k = 0; k < STAT_SLOTS; k++) {
        for (int q = 0; q < 5; q++) out[q] += slots[k * STAT_STRIDE + q];
        double d;
        std::memcpy(&d, &slots[k * STAT_STRIDE + 5], sizeof(double));
        dmax = std::max(dmax, d);
    }
    return dmax;
}

// Reference velocity and channel scales of the coming push.  After a push they follow from its
// mover sums (the mean velocity) and its largest deviation; otherwise two passes over the
// particles measure them (mean, then deviation about the mean).
static int prepare_channel_scales(sf_ctx* ctx, Species* sp) {
    if (sp->vrange_valid) return SF_OK;
    const double zero[3] = {0, 0, 0};
    const double vmax = sp->desc.vmax > 0 ? sp->desc.vmax : 1e5;
    set_channel_scales(sp, zero, vmax);
    if (sp->np == 0) return SF_OK;
    long long slots[STAT_SLOTS * STAT_STRIDE], st[5];
    double c[3] = {0, 0, 0};
    for (int pass = 0; pass < 2; pass++) {
        PushC pc = make_pushc(ctx, sp);
        SF_HIP(ctx, hipMemsetAsync(sp->stats, 0, sizeof(slots), ctx->stream));
        hipLaunchKernelGGL(velocity_range_kernel, dim3(grid_for(sp->np, 256, 2048)), dim3(256), 0, ctx->stream,
                           sp->buf[sp->cur], sp->np, pc, sp->stats);
        SF_HIP(ctx, hipGetLastError());
        SF_HIP(ctx, hipMemcpyAsync(slots, sp->stats, sizeof(slots), hipMemcpyDeviceToHost, ctx->stream));
        SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        const double dmax = combine_stat_slots(slots, st);
        if (st[0] == 0) return SF_OK;
        if (pass == 0) {
            for (int k = 0; k < 3; k++) c[k] = (double)st[1 + k] / sp->kstat / (double)st[0];
            set_channel_scales(sp, c, vmax);
        } else {
            sp->vbound = dmax;
            set_channel_scales(sp, c, 1.25 * dmax);
            sp->vrange_valid = true;
        }
    }
    return SF_OK;
}

static int run_scan(sf_ctx* ctx, Species* sp, bool samples) {
    const int ncells = ctx->mp.nci * ctx->mp.ncj;
    const int nb = (ncells + SCAN_TILE - 1) / SCAN_TILE;
    hipLaunchKernelGGL(scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, ctx->stream, sp->counts, ncells, sp->blocksum);
    hipLaunchKernelGGL(scan_blocks, dim3(1), dim3(SCAN_THREADS), 0, ctx->stream, sp->blocksum, nb);
    hipLaunchKernelGGL(scan_write, dim3(nb), dim3(SCAN_THREADS), 0, ctx->stream, sp->counts, ncells, sp->blocksum,
                       sp->offsets, sp->cursor, samples ? sp->mpc_sum : (long long*)nullptr, nb);
    SF_HIP(ctx, hipGetLastError());
    return SF_OK;
}

static int read_total(sf_ctx* ctx, Species* sp, int64_t* total) {
    const int ncells = ctx->mp.nci * ctx->mp.ncj;
    int32_t t = 0;
    SF_HIP(ctx, hipMemcpyAsync(&t, sp->offsets + ncells, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *total = t;
    return SF_OK;
}

int sort_species(sf_ctx* ctx, Species* sp) {
    if (sp->np == 0) { sp->nsorted = 0; return SF_OK; }
    PState s = sp->buf[sp->cur], d = sp->buf[1 - sp->cur];
    hipLaunchKernelGGL(count_current, dim3(grid_for(sp->np, 256)), dim3(256), 0, ctx->stream, ctx->mp, s, sp->np,
                       sp->key, sp->counts);
    int rc = run_scan(ctx, sp, false);
    if (rc) return rc;
    if (sp->desc.track_ids)
        hipLaunchKernelGGL(copy_scatter<true>, dim3(grid_for(sp->np, 256)), dim3(256), 0, ctx->stream, ctx->mp, s,
                           d, sp->np, sp->key, sp->cursor);
    else
        hipLaunchKernelGGL(copy_scatter<false>, dim3(grid_for(sp->np, 256)), dim3(256), 0, ctx->stream, ctx->mp, s,
                           d, sp->np, sp->key, sp->cursor);
    SF_HIP(ctx, hipGetLastError());
    int64_t total;
    rc = read_total(ctx, sp, &total);
    if (rc) return rc;
    if (total != sp->nalive)
        return sf_fail(ctx, SF_ERR_STATE, "sort kept %lld particles of %lld", (long long)total, (long long)sp->nalive);
    sp->cur = 1 - sp->cur;
    sp->np = total;
    sp->nsorted = total;
    sp->since_sort = 0;
    return SF_OK;
}

// one workgroup per tile
static int tile_grid(const sf_ctx* ctx) { return std::max(1, ctx->mp.ntx * ctx->mp.nty); }

int push_deposit(sf_ctx* ctx, Species* sp) {
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    const MeshP& m = ctx->mp;
    MeshD md = ctx->meshd();
    md.bnd_hits = sp->bnd_hits;
    const int rc_scales = prepare_channel_scales(ctx, sp);
    if (rc_scales) return rc_scales;
    PushC pc = make_pushc(ctx, sp);
    PState s = sp->buf[sp->cur], d = sp->buf[1 - sp->cur];
    const int64_t nn = (int64_t)m.ni * m.nj;
    const int ncells = m.nci * m.ncj;
    // a sort step compacts and re-sorts; other steps push in place (particles drift at most
    // sort_interval-1 cells from their tile, inside the LDS halo for sub-cell steps)
    const bool sort_step = (sp->since_sort + 1 >= sp->sort_interval) || sp->nsorted == 0 ||
                           (sp->np - sp->nsorted) > sp->np / 8;
    const int64_t np_before = sp->np, nsorted_before = sp->nsorted;
    SF_HIP(ctx, hipMemsetAsync(sp->stats, 0, (size_t)STAT_SLOTS * STAT_STRIDE * sizeof(long long), ctx->stream));
    SF_HIP(ctx, hipMemsetAsync(sp->flags, 0, sizeof(int), ctx->stream));
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    const bool hb = md.bf != nullptr;
    // XY without a magnetic field: the fast tile kernel, with the particles it leaves queued for
    // the full reference loop; RZ / ZR / with B: the tile kernel that runs the loop itself
    const bool fastk = m.domain == SF_DOMAIN_XY && !hb && !ctx->no_fast_push;
    const dim3 tg(tile_grid(ctx)), tb(PUSH_THREADS);
    const dim3 g_tail(grid_for(sp->np - sp->nsorted, 256, 4096)), b256(256);
    // Every step pushes in place (a dead particle becomes a hole, cell = -1).  A sort step then
    // moves the pushed particles into cell order by their new cell, compacting the holes.
    SF_HIP(ctx, hipMemsetAsync(sp->qcount, 0, 2 * sizeof(int), ctx->stream));
    if (sp->nsorted > 0) {
        if (fastk) {
            FastC fc;
            fc.qmdt = (float)pc.qm_dt;
            fc.cdx = (float)(pc.dt * m.inv_dh[0]);
            fc.cdy = (float)(pc.dt * m.inv_dh[1]);
            fc.vfast = pc.vfast;
            fc.kn = pc.kn;
            fc.kv = pc.kv;
            fc.kvv = pc.kvv;
            fc.kst = pc.kst;
            fc.c0 = pc.c0;
            fc.c1 = pc.c1;
            fc.c2 = pc.c2;
            fc.ni = m.ni;
            fc.nj = m.nj;
            fc.nci = m.nci;
            fc.ncj = m.ncj;
            fc.ntx = m.ntx;
            hipLaunchKernelGGL(push_tile_fast_xy, tg, tb, 0, ctx->stream, fc, (const float2*)ctx->d_ef32,
                               (const uint8_t*)ctx->d_cellslow, s, (const int32_t*)sp->offsets, sp->acc, sp->counts,
                               sp->stats, sp->key, sp->tile_nslow, sp->slow_items, sp->qcount);
        } else if (hb) {
            hipLaunchKernelGGL(push_deposit_tile_inplace_b, tg, tb, 0, ctx->stream, m, md, pc, s, sp->offsets,
                               sp->acc, sp->counts, sp->stats, sp->flags);
        } else {
            hipLaunchKernelGGL(push_deposit_tile_inplace, tg, tb, 0, ctx->stream, m, md, pc, s, sp->offsets,
                               sp->acc, sp->counts, sp->stats, sp->flags);
        }
    }
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    if (sp->nsorted > 0 && fastk)   // the item count is on the device; idle workgroups exit
        hipLaunchKernelGGL(push_tile_slow, dim3(2048), dim3(SLOW_CHUNK), 0, ctx->stream, m, md, pc, s,
                           (const int32_t*)sp->offsets, (const int32_t*)sp->key, (const int32_t*)sp->tile_nslow,
                           (const int2*)sp->slow_items, (const int*)sp->qcount, sp->acc, sp->counts, sp->stats,
                           sp->flags);
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
    if (sp->np > sp->nsorted) {
        if (hb)
            hipLaunchKernelGGL(push_deposit_tail<true>, g_tail, b256, 0, ctx->stream, m, md, pc, s, sp->nsorted,
                               sp->np, sp->acc, sp->counts, sp->stats, sp->flags);
        else
            hipLaunchKernelGGL(push_deposit_tail<false>, g_tail, b256, 0, ctx->stream, m, md, pc, s, sp->nsorted,
                               sp->np, sp->acc, sp->counts, sp->stats, sp->flags);
    }
    SF_HIP(ctx, hipGetLastError());
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    if (sort_step) {
        int rc = run_scan(ctx, sp, true);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(mpc_accumulate, dim3(grid_for(ncells, 256, 4096)), dim3(256), 0, ctx->stream, ncells,
                           sp->counts, sp->mpc_sum);
    }
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    if (sort_step && sp->np > 0) {
        const dim3 gs(grid_for(sp->np, 256, 1 << 20));
        if (sp->desc.track_ids)
            hipLaunchKernelGGL(copy_scatter<true>, gs, b256, 0, ctx->stream, m, s, d, sp->np, (const int32_t*)s.cell,
                               sp->cursor);
        else
            hipLaunchKernelGGL(copy_scatter<false>, gs, b256, 0, ctx->stream, m, s, d, sp->np, (const int32_t*)s.cell,
                               sp->cursor);
        SF_HIP(ctx, hipGetLastError());
    }
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    const double spwt = sp->desc.spwt;
    hipLaunchKernelGGL(normalize_nodes, dim3(grid_for(nn, 256, 8192)), dim3(256), 0, ctx->stream, nn, sp->acc,
                       sp->sums, (const double*)ctx->d_node_vol, sp->den, 1.0 / sp->kscale[0], 1.0 / sp->kscale[1],
                       1.0 / sp->kscale[4], pc.c0, pc.c1, pc.c2, spwt, sp->desc.nd_back, 1);
    SF_HIP(ctx, hipGetLastError());
    if (ctx->timing) SF_HIP(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    // KineticMaterial.updateGasProperties (KineticMaterial.java:1553-1563): one more sample, and
    // computeFields on the first update and on steps that are multiples of 10
    sp->num_samples++;
    sp->pushes++;
    const int64_t step = ctx->it > 0 ? (int64_t)ctx->it : sp->pushes;
    if (sp->first_time || step % 10 == 0) {
        hipLaunchKernelGGL(compute_fields_kernel, dim3(grid_for(nn, 256, 8192)), dim3(256), 0, ctx->stream, m,
                           (const double*)sp->sums, (const long long*)sp->mpc_sum, (const double*)ctx->d_node_vol,
                           sp->gas, sp->mass / K_B, sp->desc.init_T, 1.0 / sp->num_samples);
        SF_HIP(ctx, hipGetLastError());
    }
    sp->first_time = false;
    // mover sums (KineticMaterial.java:247-259); the count is the number of survivors
    long long slots[STAT_SLOTS * STAT_STRIDE], st[5] = {0, 0, 0, 0, 0};
    SF_HIP(ctx, hipMemcpyAsync(slots, sp->stats, sizeof(slots), hipMemcpyDeviceToHost, ctx->stream));
    int64_t total = 0;
    if (sort_step) SF_HIP(ctx, hipMemcpyAsync(&total, sp->offsets + ncells, sizeof(int32_t), hipMemcpyDeviceToHost,
                                              ctx->stream));
    int flags = 0, queued = 0;
    SF_HIP(ctx, hipMemcpyAsync(&flags, sp->flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipMemcpyAsync(&queued, sp->qcount, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double dmax = combine_stat_slots(slots, st);
    const double inv_kst = 1.0 / sp->kstat;
    sp->stats_host[0] = (double)st[0] * sp->desc.spwt * sp->mass;
    sp->stats_host[1] = (double)st[1] * inv_kst * sp->desc.spwt * sp->mass;
    sp->stats_host[2] = (double)st[2] * inv_kst * sp->desc.spwt * sp->mass;
    sp->stats_host[3] = (double)st[3] * inv_kst * sp->desc.spwt * sp->mass;
    sp->stats_host[4] = (double)st[4] * inv_kst * sp->desc.spwt * sp->mass;
    // the next push's reference velocity (the survivors' mean) and scales: the survivors deviate
    // from the new c by at most dmax + |c_new - c|
    if (st[0] > 0) {
        double c[3], shift = 0;
        for (int k = 0; k < 3; k++) {
            c[k] = (double)st[1 + k] * inv_kst / (double)st[0];
            shift = std::max(shift, std::fabs(c[k] - sp->vref[k]));
        }
        sp->vbound = dmax + shift;
        set_channel_scales(sp, c, 1.25 * sp->vbound);
        sp->vrange_valid = true;
    }
    ctx->last_queue[1] = queued;
    if (ctx->timing) {
        float a, b, c, e, t0 = 0.f, t1 = 0.f;
        hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]);
        hipEventElapsedTime(&t0, ctx->ev[0], ctx->ev[5]);
        hipEventElapsedTime(&t1, ctx->ev[0], ctx->ev[6]);
        ctx->last_push_ms[0] = t0;
        ctx->last_push_ms[1] = a - t1;
        ctx->last_push_ms[2] = (double)(np_before - nsorted_before);
        ctx->last_queue[0] = t1 - t0;
        hipEventElapsedTime(&b, ctx->ev[1], ctx->ev[2]);
        hipEventElapsedTime(&c, ctx->ev[2], ctx->ev[3]);
        hipEventElapsedTime(&e, ctx->ev[3], ctx->ev[4]);
        ctx->last_ms[0] = a;
        ctx->last_ms[1] = b;
        ctx->last_ms[2] = sort_step ? c : 0.0;   // no scatter on a step without a sort
        ctx->last_ms[3] = e;
    }
    sp->nalive = st[0];
    if (sort_step) {
        if (total != st[0])
            return sf_fail(ctx, SF_ERR_STATE, "counting sort kept %lld particles, the mover %lld", (long long)total,
                           (long long)st[0]);
        sp->np = total;
        sp->nsorted = total;
        sp->cur = 1 - sp->cur;
        sp->since_sort = 0;
    } else {
        sp->since_sort++;
    }
    if (flags & PUSH_FLAG_OVERFLOW)
        return sf_fail(ctx, SF_WARN_MOMENT_RANGE,
                       "a particle velocity deviating by more than %.3g m/s from the species' mean overflowed "
                       "the 64-bit moment channels of this step", pc.vover);
    return SF_OK;
}

// ----------------------------------------------------------------------------------------
// sf_particles_add_noalloc: the staging of add_particles_device in a fixed device array
// ----------------------------------------------------------------------------------------
// Layout: the five FP64 input columns (x, y, u, v, w) SF_NOALLOC_MAX_ADD apart, the int32 ids,
// then the accepted count.  Every GPU holds its own copy until the process ends; calls that
// overlap on one GPU share it and race.
constexpr size_t NOALLOC_COLS = (size_t)5 * SF_NOALLOC_MAX_ADD * sizeof(double);
constexpr size_t NOALLOC_IDS = (size_t)SF_NOALLOC_MAX_ADD * sizeof(int32_t);
constexpr size_t NOALLOC_BYTES = NOALLOC_COLS + NOALLOC_IDS + 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
static_assert(SF_NOALLOC_MAX_ADD <= FILTER_THREADS * 64, "the filter kernel takes 64 particles per thread at most");

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device with the device
// current (the array's name is a host-side stand-in and must not reach a kernel).
int noalloc_scratch_init(sf_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return sf_fail(ctx, SF_ERR_ARG, "device %d: the fixed staging array supports devices 0-%d", ctx->device,
                       NOALLOC_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[ctx->device]) return SF_OK;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    void* p = nullptr;
    SF_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return sf_fail(ctx, SF_ERR_HIP, "hipGetSymbolAddress returned no address for the staging array");
    g_scratch_ptr[ctx->device] = p;
    return SF_OK;
}

int add_particles_noalloc(sf_ctx* ctx, Species* sp, int64_t n, const double* x, const double* y, const double* u,
                          const double* v, const double* w, const int* ids, int mode) {
    if (n <= 0) return SF_OK;
    if (n > SF_NOALLOC_MAX_ADD)
        return sf_fail(ctx, SF_ERR_ARG,
                       "sf_particles_add_noalloc: %lld particles exceed the %d of its fixed staging array "
                       "(split the call, or use sf_particles_add)", (long long)n, SF_NOALLOC_MAX_ADD);
    if (sp->np + n > sp->capacity)
        return sf_fail(ctx, SF_ERR_OOM, "species capacity %lld exceeded (have %lld, adding %lld)",
                       (long long)sp->capacity, (long long)sp->np, (long long)n);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    if (!base) return sf_fail(ctx, SF_ERR_STATE, "sf_particles_add_noalloc: staging array not set up on device %d",
                              ctx->device);
    double* d_cols = (double*)base;
    int32_t* d_ids = (int32_t*)((char*)base + NOALLOC_COLS);
    int* d_kept = (int*)((char*)base + NOALLOC_COLS + NOALLOC_IDS);
    const double* src[5] = {x, y, u, v, w};
    for (int c = 0; c < 5; c++)
        SF_HIP(ctx, hipMemcpyAsync(d_cols + (size_t)c * SF_NOALLOC_MAX_ADD, src[c], n * sizeof(double),
                                   hipMemcpyHostToDevice, ctx->stream));
    const bool with_ids = ids && sp->desc.track_ids;
    if (with_ids)
        SF_HIP(ctx, hipMemcpyAsync(d_ids, ids, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    const MeshP& m = ctx->mp;
    const double xd0 = m.x0[0] + (m.ni - 1) * m.dh[0], xd1 = m.x0[1] + (m.nj - 1) * m.dh[1];
    hipLaunchKernelGGL(add_particles_filtered, dim3(1), dim3(FILTER_THREADS), 0, ctx->stream, ctx->mp, ctx->meshd(),
                       sp->buf[sp->cur], sp->np, (int)n, (const double*)d_cols, (int64_t)SF_NOALLOC_MAX_ADD, xd0, xd1,
                       sp->qm, -0.5 * ctx->dt * sp->qm, -0.5 * ctx->dt, mode,
                       with_ids ? (const int32_t*)d_ids : (const int32_t*)nullptr, sp->id_counter,
                       sp->desc.track_ids, d_kept);
    SF_HIP(ctx, hipGetLastError());
    int kept = 0;
    SF_HIP(ctx, hipMemcpyAsync(&kept, d_kept, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!with_ids) sp->id_counter += kept;
    if ((int64_t)kept * 8 >= sp->nalive) sp->vrange_valid = false;   // a bulk load: re-measure the velocities
    sp->np += kept;
    sp->nalive += kept;
    return SF_OK;
}

int add_particles_device(sf_ctx* ctx, Species* sp, int64_t n, const double* x, const double* y, const double* u,
                         const double* v, const double* w, const int* ids, int mode) {
    if (n <= 0) return SF_OK;
    if (sp->np + n > sp->capacity)
        return sf_fail(ctx, SF_ERR_OOM, "species capacity %lld exceeded (have %lld, adding %lld)",
                       (long long)sp->capacity, (long long)sp->np, (long long)n);
    // MeshData.addParticle drops non-finite velocities (KineticMaterial.java:1356-1361); also
    // only particles inside the mesh are accepted (DomainModule.getMesh, DomainModule.java:106-117)
    std::vector<double> h;
    h.reserve(5 * n);
    std::vector<int> hid;
    std::vector<double> cols[5];
    const MeshP& m = ctx->mp;
    const double xd0 = m.x0[0] + (m.ni - 1) * m.dh[0], xd1 = m.x0[1] + (m.nj - 1) * m.dh[1];
    int64_t kept = 0;
    for (int c = 0; c < 5; c++) cols[c].reserve(n);
    for (int64_t k = 0; k < n; k++) {
        double vx = u[k], vy = v[k], vz = w[k];
        if (!(std::isfinite(vx) && std::isfinite(vy) && std::isfinite(vz))) continue;
        double lc0 = (x[k] - m.x0[0]) / m.dh[0], lc1 = (y[k] - m.x0[1]) / m.dh[1];
        bool strict = x[k] >= m.x0[0] && x[k] < xd0 && y[k] >= m.x0[1] && y[k] < xd1;
        bool loose = !(lc0 < -FLT_EPS || lc1 < -FLT_EPS || lc0 > (m.ni - 1 + FLT_EPS) || lc1 > (m.nj - 1 + FLT_EPS));
        if (!(strict || loose)) continue;
        cols[0].push_back(x[k]);
        cols[1].push_back(y[k]);
        cols[2].push_back(vx);
        cols[3].push_back(vy);
        cols[4].push_back(vz);
        if (ids) hid.push_back(ids[k]);
        kept++;
    }
    if (kept == 0) return SF_OK;
    for (int c = 0; c < 5; c++) h.insert(h.end(), cols[c].begin(), cols[c].end());
    double* d_in = nullptr;
    int32_t* d_ids = nullptr;
    SF_HIP(ctx, hipMallocAsync((void**)&d_in, 5 * kept * sizeof(double), ctx->stream));
    SF_HIP(ctx, hipMemcpyAsync(d_in, h.data(), 5 * kept * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (ids && sp->desc.track_ids) {
        SF_HIP(ctx, hipMallocAsync((void**)&d_ids, kept * sizeof(int32_t), ctx->stream));
        SF_HIP(ctx, hipMemcpyAsync(d_ids, hid.data(), kept * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    }
    hipLaunchKernelGGL(add_particles_kernel_strided, dim3(grid_for(kept, 256, 4096)), dim3(256), 0, ctx->stream,
                       ctx->mp, ctx->meshd(), sp->buf[sp->cur], sp->np, kept, (const double*)d_in, kept, sp->qm,
                       -0.5 * ctx->dt * sp->qm, -0.5 * ctx->dt, mode, (const int32_t*)d_ids, sp->id_counter, sp->desc.track_ids);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipFreeAsync(d_in, ctx->stream));
    if (d_ids) SF_HIP(ctx, hipFreeAsync(d_ids, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!(ids && sp->desc.track_ids)) sp->id_counter += (int)kept;
    if (kept * 8 >= sp->nalive) sp->vrange_valid = false;   // a bulk load: re-measure the velocities
    sp->np += kept;
    sp->nalive += kept;
    return SF_OK;
}

int preload(sf_ctx* ctx, Species* sp, int64_t n, const double* drift, double temp, uint64_t seed) {
    if (n <= 0) return SF_OK;
    if (sp->np + n > sp->capacity) return sf_fail(ctx, SF_ERR_OOM, "species capacity exceeded by preload");
    const MeshP& m = ctx->mp;
    std::vector<int32_t> open;
    std::vector<double> cdf(1, 0.0);
    for (int j = 0; j < m.ncj; j++)
        for (int i = 0; i < m.nci; i++) {
            auto nt = [&](int a, int b) { return ctx->h_node_type[(size_t)b * m.ni + a]; };
            if (nt(i, j) == SF_NODE_OPEN && nt(i + 1, j) == SF_NODE_OPEN && nt(i + 1, j + 1) == SF_NODE_OPEN &&
                nt(i, j + 1) == SF_NODE_OPEN) {
                open.push_back((j << 16) | i);
                double wgt = 1.0;
                if (m.domain == SF_DOMAIN_RZ) {
                    double r1 = m.x0[0] + i * m.dh[0], r2 = r1 + m.dh[0];
                    wgt = r2 * r2 - r1 * r1;
                }
                cdf.push_back(cdf.back() + wgt);
            }
        }
    if (open.empty()) return sf_fail(ctx, SF_ERR_STATE, "preload: no open cells");
    int32_t* d_open = nullptr;
    double* d_cdf = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&d_open, open.size() * sizeof(int32_t)));
    SF_HIP(ctx, hipMalloc((void**)&d_cdf, cdf.size() * sizeof(double)));
    SF_HIP(ctx, hipMemcpy(d_open, open.data(), open.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(d_cdf, cdf.data(), cdf.size() * sizeof(double), hipMemcpyHostToDevice));
    double vth = sqrt(2 * K_B * temp / sp->mass);   // Utils.computeVth
    hipLaunchKernelGGL(preload_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, ctx->stream, ctx->mp,
                       sp->buf[sp->cur], sp->np, n, (const int32_t*)d_open, (const double*)d_cdf, (int)open.size(),
                       drift[0], drift[1], drift[2], vth, seed, sp->desc.track_ids, sp->id_counter);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(d_open);
    hipFree(d_cdf);
    sp->id_counter += (int)n;
    sp->np += n;
    sp->nalive += n;
    sp->vrange_valid = false;   // re-measure the velocities at the next push
    return SF_OK;
}

int ef32_refresh(sf_ctx* ctx) {
    int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    hipLaunchKernelGGL(ef32_kernel, dim3(grid_for(nn, 256, 8192)), dim3(256), 0, ctx->stream, nn,
                       (const double*)ctx->d_efi, (const double*)ctx->d_efj, ctx->d_ef32);
    SF_HIP(ctx, hipGetLastError());
    return SF_OK;
}

// B field (DomainModule "bfi"/"bfj", DomainModule.java:174) stored as FP32 pairs like E.  A field
// that is zero on every node frees the buffer: the mover then takes the electric kick everywhere,
// which is what the reference does where the gathered B is zero.
int bfield_set(sf_ctx* ctx, const double* bfi, const double* bfj) {
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    bool any = false;
    std::vector<float2> h;
    if (bfi && bfj) {
        h.resize(nn);
        for (int64_t u = 0; u < nn; u++) {
            if (!std::isfinite(bfi[u]) || !std::isfinite(bfj[u]))
                return sf_fail(ctx, SF_ERR_ARG, "sf_bfield_set: non-finite value at node %lld", (long long)u);
            h[u] = make_float2((float)bfi[u], (float)bfj[u]);
            any = any || h[u].x != 0.0f || h[u].y != 0.0f;
        }
    }
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!any) {
        if (ctx->d_bf32) SF_HIP(ctx, hipFree(ctx->d_bf32));
        ctx->d_bf32 = nullptr;
        return SF_OK;
    }
    if (!ctx->d_bf32) SF_HIP(ctx, hipMalloc((void**)&ctx->d_bf32, nn * sizeof(float2)));
    SF_HIP(ctx, hipMemcpy(ctx->d_bf32, h.data(), nn * sizeof(float2), hipMemcpyHostToDevice));
    return SF_OK;
}

}  // namespace sf

extern "C" int sf_bfield_set(sf_ctx* ctx, const double* bfi, const double* bfj) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_bfield_set: mesh not set up");
    if ((bfi == nullptr) != (bfj == nullptr)) return sf_fail(ctx, SF_ERR_ARG, "sf_bfield_set: give both components or neither");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return sf::bfield_set(ctx, bfi, bfj);
}
