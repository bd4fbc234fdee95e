// Boltzmann-electron nonlinear Poisson solve (pic/PoissonSolver.java, core/solver/Solver.java)
// on gfx950: matrix-free operator, FP64 Newton, FP32 Jacobi-PCG with FP64 residual refinement.
//
// Reference operator.  Solver.initCoefficients assembles, on a uniform mesh:
//   * Dirichlet / internal nodes: identity rows                         (Solver.java:463-468)
//   * nodes on a non-MESH mesh face: the gradient row Gi (left/right faces, which take
//     precedence at corners) or Gj (bottom/top), i.e. (phi_in - phi_face)/dh = 0
//                                                                        (Solver.java:486-499)
//   * every other node: the finite-volume Laplacian sum_f R_f (grad phi . n) l_f / V with
//     V = nodeVol (2*pi*A*r in RZ), which on a uniform mesh is the 5-point stencil
//     XY: (phi_E - 2 phi + phi_W)/dx^2 + (phi_N - 2 phi + phi_S)/dy^2
//     RZ: [r_{i+1/2}(phi_E - phi) - r_{i-1/2}(phi - phi_W)]/(2 pi r_i dr^2)
//         + (phi_N - 2 phi + phi_S)/(2 pi dz^2)                          (Solver.java:503-552)
// Eliminating the identity rows (values into the right-hand side) and the one-sided rows
// (phi_face = phi_inside) and multiplying each remaining row by its control volume W_i gives
// a symmetric positive-definite system L x + W b(x) = f with the same discrete solution.
// Node classes are derived on the host; a free node stores, per neighbour, whether that
// neighbour is free, fixed (value known) or a face node tied back to the node itself.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "sf_internal.h"

namespace sf {

enum { NB_SELF = 0, NB_FREE = 1, NB_FIXED = 2 };
constexpr uint8_t CLS_NOT_FREE = 0xFF;

constexpr int NK_RING = 8;
// Forcing terms of the inexact Newton steps: a linear solve stops at eta times its starting
// residual.  ETA_COLD for a solve that starts far from its solution (first residual at least half
// the reference residual, as from phi = 0), ETA_WARM for one started from the previous step's phi.
// Measured on the 1024x512 cylinder deck: warm solves take 0.82, 0.76, 0.74, 0.70, 0.67 and
// 0.68 ms at eta = 1e-3, 3e-3, 1e-2, 3e-2, 5e-2 and 1e-1 (fewer PCG iterations, a few more Newton
// updates); a cold solve with 1e-2 takes 117 instead of 70 ms.  On meshes too large for one
// wide PCG tile per CU (narrow tiles, choose_tile_shape) a PCG iteration costs more and the
// linear solves converge more slowly, and looser inner solves pay: on the 2048 x 1024 RZ deck warm
// solves take 18.1, 17.6, 16.2 and 19.0 ms at eta = 2e-2, 5e-2, 1e-1 and 2e-1 (the 1024 x 512
// deck: 0.63 ms at 5e-2, 0.66 at 1e-1, 0.70 at 2e-1).
constexpr double ETA_COLD = 1e-3, ETA_WARM = 5e-2, ETA_WARM_LARGE = 1e-1;
struct NkState {
    double tol2;   // PCG tolerance (squared) of the current linear solve
    double r2;     // ||r||^2 of the Newton residual at the current linear solve's start (at stop: final)
    int k;         // PCG iteration within the current linear solve
    int newton;    // Newton updates so far
    int pcg;       // PCG iterations so far
    int done;      // 0 running, 1 converged, 2 Newton iteration cap, 3 residual not finite
};
struct PoissonState {
    sf_poisson_desc desc{};
    double eps = EPS0, C = 0;     // C = e/eps
    int nfree = 0;
    // device
    uint8_t* cls = nullptr;       // [nn] neighbour classes (2 bits each: E, W, N, S) or CLS_NOT_FREE
    double* W = nullptr;          // [nn] row weight (control volume)
    double* bfix = nullptr;       // [nn] sum of c' * phi_fixed over fixed neighbours
    int32_t* src = nullptr;       // [nn] node whose value a node takes (itself if free, -1 if fixed)
    double* fixval = nullptr;     // [nn] value of fixed nodes
    double* x = nullptr;          // [nn] FP64 potential on free nodes (solution vector)
    double* r = nullptr;          // [nn] FP64 Newton residual
    double* y = nullptr;          // [nn] FP64 accumulated correction
    double* rr = nullptr;         // [nn] FP64 refinement residual
    float* D = nullptr;           // [nn] FP32 Jacobian diagonal
    double* P = nullptr;          // [nn] FP64 W * b'(x)
    float *d = nullptr, *res = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    double* partial = nullptr;    // [nblk_elem] per-block partials of the Newton-level kernels
    double* pz = nullptr;         // [2][nblk_elem][2] (r.z, r.r) partials of PCG, by iteration parity
    double* ppq = nullptr;        // [nblk_tile] p.q partials of PCG
    double* scal = nullptr;       // device scalars
    int* istate = nullptr;        // [0] done (1 converged, 2 not finite), [1] iterations at done
    int nblk_tile = 0, nblk_elem = 0;
    bool wide_tiles = true;       // TileWide (else TileNarrow) for the tile kernels
    int last_its = 7;             // PCG iterations of the last linear solve (sizes the next batch)
    int last_pairs = 24;          // step pairs of the last device-driven solve (sizes its first batch)
    bool host_newton = false;     // SF_HOST_NEWTON: keep every solve on the host-driven loop
    bool warm_hint = false;       // the last solve converged (the next one likely starts warm)
    NkState* nk = nullptr;        // [NK_RING] device-driven Newton-PCG state ring
    double* nkc = nullptr;        // [NK_RING][2] (r.z, r.r) combined by pair i's nk_step_a
    NkState* host_nk = nullptr;   // pinned copy of the last state read
    double* host_scal = nullptr;  // pinned
    int* host_istate = nullptr;
    hipEvent_t e0{}, e1{};
};

// device scalars layout (FP64): PCG tolerance, squared norms from the finalize kernel
enum { S_TOL2 = 0, S_RZ0, S_NORM2, S_R0, S_YMAX, S_YNORM2, S_NSCAL = 8 };

// Reductions without grid-wide synchronisation.  A kernel leaves one partial per block
// (block_partials); the next kernel on the stream combines them (combine_partials, every
// block redundantly, in a fixed order) or a one-block finalize_partials kernel stores the total
// in the scalar array.  Device-scope fences and ticket atomics are avoided on purpose: on a
// multi-XCD part a release fence per block costs tens of microseconds per launch.
template <bool MAX>
__device__ __forceinline__ double red_op(double a, double b) { return MAX ? fmax(a, b) : a + b; }

template <int NV, bool MAX = false>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partial) {
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        for (int o = 32; o > 0; o >>= 1) v[k] = red_op<MAX>(v[k], __shfl_xor(v[k], o));
        if (lane == 0) red[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) {
            double t = red[k][0];
            for (int w = 1; w < nw; w++) t = red_op<MAX>(t, red[k][w]);
            partial[NV * blockIdx.x + k] = t;
        }
}

// Totals of nb partial NV-vectors written by an earlier kernel; every thread receives them.
template <int NV, bool MAX = false>
__device__ __forceinline__ void combine_partials(const double* __restrict__ partial, int nb, double (&tot)[NV]) {
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    double a[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) a[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) a[k] = red_op<MAX>(a[k], partial[NV * b + k]);
#pragma unroll
    for (int k = 0; k < NV; k++) {
        for (int o = 32; o > 0; o >>= 1) a[k] = red_op<MAX>(a[k], __shfl_xor(a[k], o));
        if (lane == 0) red[k][wave] = a[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double t = red[k][0];
        for (int w = 1; w < nw; w++) t = red_op<MAX>(t, red[k][w]);
        tot[k] = t;
    }
    __syncthreads();   // red[] may be reused by the next call
}

template <int NV, bool MAX = false>
__global__ void __launch_bounds__(256) finalize_partials(const double* __restrict__ partial, int nb,
                                                         double* __restrict__ scal, int slot) {
    double tot[NV];
    combine_partials<NV, MAX>(partial, nb, tot);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) scal[slot + k] = tot[k];
}

void poisson_free(PoissonState* p) {
    if (!p) return;
    void* ptrs[] = {p->cls, p->W,  p->bfix, p->src, p->fixval, p->x,  p->r,       p->y,    p->rr,
                    p->D,   p->P,  p->d,    p->res, p->z,      p->p0, p->p1,      p->q,    p->partial,
                    p->pz,  p->ppq, p->scal, p->istate};
    for (void* q : ptrs)
        if (q) hipFree(q);
    if (p->host_scal) hipHostFree(p->host_scal);
    if (p->host_istate) hipHostFree(p->host_istate);
    if (p->host_nk) hipHostFree(p->host_nk);
    if (p->nk) hipFree(p->nk);
    if (p->nkc) hipFree(p->nkc);
    if (p->e0) hipEventDestroy(p->e0);
    if (p->e1) hipEventDestroy(p->e1);
    delete p;
}

// symmetric stencil coefficients c' of a free node (row scaled by its control volume)
struct Stencil {
    double cE, cW, cN, cS;
};
__host__ __device__ __forceinline__ Stencil stencil(const MeshP& m, int i, int j) {
    Stencil s;
    if (m.domain == SF_DOMAIN_RZ) {
        double dr = m.dh[0], dz = m.dh[1];
        double ri = m.x0[0] + i * dr;
        s.cE = (ri + 0.5 * dr) * dz / dr;
        s.cW = (ri - 0.5 * dr) * dz / dr;
        s.cN = ri * dr / dz;
        s.cS = s.cN;
    } else if (m.domain == SF_DOMAIN_ZR) {
        double dz = m.dh[0], dr = m.dh[1];
        double rj = m.x0[1] + j * dr;
        s.cN = (rj + 0.5 * dr) * dz / dr;
        s.cS = (rj - 0.5 * dr) * dz / dr;
        s.cE = rj * dr / dz;
        s.cW = s.cE;
    } else {
        s.cE = s.cW = m.dh[1] / m.dh[0];
        s.cN = s.cS = m.dh[0] / m.dh[1];
    }
    return s;
}
// control volume that scales a free row (Mesh.nodeVol for interior nodes, 2*pi factor kept)
__host__ __device__ __forceinline__ double row_weight(const MeshP& m, int i, int j) {
    if (m.domain == SF_DOMAIN_RZ) return 2 * PI * m.dh[0] * m.dh[1] * (m.x0[0] + i * m.dh[0]);
    if (m.domain == SF_DOMAIN_ZR) return 2 * PI * m.dh[0] * m.dh[1] * (m.x0[1] + j * m.dh[1]);
    return m.dh[0] * m.dh[1];
}
// reference row scale: the assembled RZ row carries R_f but is divided by 2*pi*A*r, so the
// symmetric coefficients above (without 2*pi) equal W * A_ref with W = 2*pi*A*r.

__device__ __forceinline__ int nbc(uint8_t c, int dir) { return (c >> (2 * dir)) & 3; }

// L x restricted to free neighbours (FP64): diag*x_i - sum_free c' x_nb
__device__ __forceinline__ double lred_apply64(const MeshP& m, uint8_t c, int i, int j, size_t u, const double* x,
                                               double& diagL) {
    Stencil s = stencil(m, i, j);
    double xi = x[u];
    double acc = 0;
    diagL = 0;
    int e = nbc(c, 0), w = nbc(c, 1), n = nbc(c, 2), so = nbc(c, 3);
    if (e != NB_SELF) { diagL += s.cE; if (e == NB_FREE) acc -= s.cE * x[u + 1]; }
    if (w != NB_SELF) { diagL += s.cW; if (w == NB_FREE) acc -= s.cW * x[u - 1]; }
    if (n != NB_SELF) { diagL += s.cN; if (n == NB_FREE) acc -= s.cN * x[u + m.ni]; }
    if (so != NB_SELF) { diagL += s.cS; if (so == NB_FREE) acc -= s.cS * x[u - m.ni]; }
    return acc + diagL * xi;
}

// Newton residual r = f + bfix - L x - W b(x), Jacobian diagonal D = diag(L) + W b'(x);
// also ||r||^2 partials.  f = W * rho / eps.  The accumulated correction y is zero at this point,
// so r is also the first refinement residual: the FP32 PCG starts from d = 0, res = r.
__global__ void newton_residual(MeshP m, const uint8_t* __restrict__ cls, const double* __restrict__ W,
                                const double* __restrict__ bfix, const double* __restrict__ rho,
                                const double* __restrict__ x, double* __restrict__ r, float* __restrict__ D,
                                double* __restrict__ P, float* __restrict__ d, float* __restrict__ res,
                                double inv_eps, double Cn0, double phi0, double inv_kTe, int linear,
                                double* __restrict__ partial, const int* __restrict__ go) {
    if (go != nullptr && go[0] != 1) return;   // queued behind a linear solve that has not converged
    const int64_t nn = (int64_t)m.ni * m.nj;
    double acc = 0;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        uint8_t c = cls[u];
        d[u] = 0.0f;
        if (c == CLS_NOT_FREE) { r[u] = 0; D[u] = 1.0f; P[u] = 0; res[u] = 0.0f; continue; }
        int i = (int)(u % m.ni), j = (int)(u / m.ni);
        double diagL;
        double Lx = lred_apply64(m, c, i, j, u, x, diagL);
        double Wu = W[u];
        double bx = 0, bp = 0;
        if (!linear) {
            bx = Cn0 * exp((x[u] - phi0) * inv_kTe);    // PoissonSolver.java:229-244
            bp = bx * inv_kTe;
        }
        double ru = Wu * rho[u] * inv_eps + bfix[u] - Lx - Wu * bx;
        r[u] = ru;
        res[u] = (float)ru;
        P[u] = Wu * bp;
        D[u] = (float)(diagL + Wu * bp);
        acc += ru * ru;
    }
    double v[1] = {acc};
    block_partials<1>(v, partial);
}

// ||f + bfix - W b(0)||^2 for the fixed reference norm
// ||r_ref||^2 partials (the residual at x = 0 on the free nodes), and x gathered from phi (0 on
// the nodes that are not free) in the same pass
__global__ void ref_norm_kernel(MeshP m, const uint8_t* __restrict__ cls, const double* __restrict__ W,
                                const double* __restrict__ bfix, const double* __restrict__ rho, double inv_eps,
                                double Cn0, double phi0, double inv_kTe, int linear,
                                double* __restrict__ partial, const double* __restrict__ phi,
                                double* __restrict__ x) {
    const int64_t nn = (int64_t)m.ni * m.nj;
    double acc = 0;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        const bool free_node = cls[u] != CLS_NOT_FREE;
        x[u] = free_node ? phi[u] : 0.0;
        if (!free_node) continue;
        double b0 = linear ? 0.0 : Cn0 * exp((0.0 - phi0) * inv_kTe);
        double v = W[u] * rho[u] * inv_eps + bfix[u] - W[u] * b0;
        acc += v * v;
    }
    double v[1] = {acc};
    block_partials<1>(v, partial);
}

// refinement residual rr = r - (L + P) y (FP64); also initialises the FP32 PCG (d = 0,
// res = rr) and leaves the partials of ||rr||^2 (its total is the refinement norm)
__global__ void refine_residual(MeshP m, const uint8_t* __restrict__ cls, const double* __restrict__ r,
                                const double* __restrict__ y, const double* __restrict__ P,
                                double* __restrict__ rr, float* __restrict__ d, float* __restrict__ res,
                                double* __restrict__ partial, const int* __restrict__ go) {
    if (go != nullptr && go[0] != 1) return;
    const int64_t nn = (int64_t)m.ni * m.nj;
    double a_rr = 0;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        uint8_t c = cls[u];
        float zero = 0.0f;
        if (c == CLS_NOT_FREE) {
            rr[u] = 0; d[u] = zero; res[u] = zero;
            continue;
        }
        int i = (int)(u % m.ni), j = (int)(u / m.ni);
        double diagL;
        double Jy = lred_apply64(m, c, i, j, u, y, diagL) + P[u] * y[u];
        double v = r[u] - Jy;
        rr[u] = v;
        d[u] = zero;
        res[u] = (float)v;
        a_rr += v * v;
    }
    double v[1] = {a_rr};
    block_partials<1>(v, partial);
}

// Preconditioned CG in FP32, two kernels per iteration k, no grid-wide synchronisation: each
// kernel combines the partial sums the previous one left.  (r.z, r.r) partials alternate between
// pz[k & 1] (read at iteration k) and pz[(k + 1) & 1] (written by its update), so r.z of
// iteration k - 1 is still available for beta.  Convergence (r.r <= tol^2) is detected by
// every block of the step-1 kernel from the same sums; block 0 records it in istate, which
// turns every later queued kernel into a no-op.
// The tile kernels are templates over the tile shape: W x H nodes per block of NT threads; thread
// t owns column t % W and rows t / W + RSTEP s (s < RPT) of the tile.  MINW: the waves per SIMD
// nk_step_b is compiled for (its register budget).
template <int W_, int H_, int NT_, int MINW_>
struct TileShape {
    static constexpr int W = W_, H = H_, NT = NT_, MINW = MINW_;
    static constexpr int RSTEP = NT_ / W_;
    static constexpr int RPT = H_ / RSTEP;
    static_assert(NT_ % W_ == 0 && H_ % RSTEP == 0, "tile shape");
};
// 128 x 16 tiles of 1024 threads: the stronger block preconditioner, for meshes whose tiles fit in
// one workgroup per CU.  64 x 16 tiles of 512 threads on larger meshes, where several workgroups
// per CU hide each other's barriers (choose_tile_shape); there nk_step_b is held to 64 registers so
// that four of them fit a CU (69 registers allow three).
using TileWide = TileShape<128, 16, 1024, 4>;
using TileNarrow = TileShape<64, 16, 512, 8>;

// step 1: beta = rz_k / rz_{k-1} (0 at k = 0); p = z + beta p_old; q = J p; p.q partials
template <class TS>
__global__ void __launch_bounds__(TS::NT) pcg_spmv(MeshP m, const uint8_t* __restrict__ cls,
                                                const float* __restrict__ D, const float* __restrict__ z,
                                                const float* __restrict__ pold, float* __restrict__ pnew,
                                                float* __restrict__ q, const double* __restrict__ pz_cur,
                                                const double* __restrict__ pz_prev, int nbz, int k,
                                                double* __restrict__ ppq, double tol2, int* istate) {
    if (istate[0]) return;
    double cur[2];
    combine_partials<2>(pz_cur, nbz, cur);   // (rz_k, rr_k)
    if (!(cur[1] > tol2)) {                  // converged, or not finite
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            istate[0] = (cur[1] == cur[1]) ? 1 : 2;
            istate[1] = k;
        }
        return;
    }
    float beta = 0.0f;
    if (k > 0) {
        double prev[2];
        combine_partials<2>(pz_prev, nbz, prev);
        beta = (float)(cur[0] / prev[0]);
    }
    __shared__ float ps[(TS::H + 2) * (TS::W + 2)];
    const int nbx = (m.ni + TS::W - 1) / TS::W;
    const int bx = blockIdx.x % nbx, by = blockIdx.x / nbx;
    const int i0 = bx * TS::W - 1, j0 = by * TS::H - 1;
    for (int t = threadIdx.x; t < (TS::H + 2) * (TS::W + 2); t += blockDim.x) {
        int i = i0 + t % (TS::W + 2), j = j0 + t / (TS::W + 2);
        float v = 0.0f;
        if (i >= 0 && j >= 0 && i < m.ni && j < m.nj) {
            size_t u = (size_t)j * m.ni + i;
            v = k == 0 ? z[u] : z[u] + beta * pold[u];
        }
        ps[t] = v;
    }
    __syncthreads();
    const int li = threadIdx.x % TS::W;
    double acc = 0;
    for (int lj = threadIdx.x / TS::W; lj < TS::H; lj += TS::RSTEP) {
        const int i = i0 + 1 + li, j = j0 + 1 + lj;
        if (i >= m.ni || j >= m.nj) continue;
        size_t u = (size_t)j * m.ni + i;
        uint8_t c = cls[u];
        int l = (lj + 1) * (TS::W + 2) + (li + 1);
        float pv = ps[l];
        float qv = 0.0f;
        if (c != CLS_NOT_FREE) {
            Stencil s = stencil(m, i, j);
            qv = D[u] * pv;
            if (nbc(c, 0) == NB_FREE) qv -= (float)s.cE * ps[l + 1];
            if (nbc(c, 1) == NB_FREE) qv -= (float)s.cW * ps[l - 1];
            if (nbc(c, 2) == NB_FREE) qv -= (float)s.cN * ps[l + TS::W + 2];
            if (nbc(c, 3) == NB_FREE) qv -= (float)s.cS * ps[l - TS::W - 2];
        }
        pnew[u] = pv;
        q[u] = qv;
        acc += (double)pv * (double)qv;
    }
    double v[1] = {acc};
    block_partials<1>(v, ppq);
}

// Preconditioner M: SGS_SWEEPS symmetric Gauss-Seidel sweeps of J in red-black order (red,
// black, red, ..., red) on each TS::W x TS::H tile, with the values outside the tile taken as zero.
// That is block Jacobi over the tiles with a symmetric inner solve, so M is symmetric positive
// definite like the Jacobi preconditioner it replaces (the reference's LinearSolverPCG uses
// M = diag(A), LinearSolverPCG.java:48-49).  On the 1024x512 cylinder deck a cold solve takes
// 7242 PCG iterations with Jacobi, 4048 with one sweep and 2623 with three, and the sweeps stay
// inside LDS.
constexpr int SGS_SWEEPS = 3;
// Thread t owns column t % TS::W and rows t / TS::W + TS::RSTEP s of the tile.
template <class TS>
struct TileRows {
    float Dv[TS::RPT], cE[TS::RPT], cW[TS::RPT], cN[TS::RPT], cS[TS::RPT];
    uint8_t c[TS::RPT];
    bool in[TS::RPT];
};
template <class TS>
__device__ __forceinline__ void tile_rows(const MeshP& m, const uint8_t* __restrict__ cls, const float* __restrict__ D,
                                          int i0, int j0, TileRows<TS>& T) {
    const int li = threadIdx.x % TS::W;
#pragma unroll
    for (int s = 0; s < TS::RPT; s++) {
        const int lj = threadIdx.x / TS::W + TS::RSTEP * s;
        const int i = i0 + li, j = j0 + lj;
        T.in[s] = i < m.ni && j < m.nj;
        T.c[s] = CLS_NOT_FREE;
        T.Dv[s] = 1.0f;
        T.cE[s] = T.cW[s] = T.cN[s] = T.cS[s] = 0.0f;
        if (T.in[s]) {
            const size_t u = (size_t)j * m.ni + i;
            const uint8_t c = cls[u];
            T.c[s] = c;
            if (c != CLS_NOT_FREE) {
                const Stencil st = stencil(m, i, j);
                T.Dv[s] = D[u];
                // only free neighbours inside the tile couple
                T.cE[s] = (nbc(c, 0) == NB_FREE && li + 1 < TS::W) ? (float)st.cE : 0.0f;
                T.cW[s] = (nbc(c, 1) == NB_FREE && li > 0) ? (float)st.cW : 0.0f;
                T.cN[s] = (nbc(c, 2) == NB_FREE && lj + 1 < TS::H) ? (float)st.cN : 0.0f;
                T.cS[s] = (nbc(c, 3) == NB_FREE && lj > 0) ? (float)st.cS : 0.0f;
            }
        }
    }
}
// z = M^-1 r on the tile: rs holds r (0 at nodes that are not free or outside the mesh),
// zs receives z.
template <class TS>
__device__ __forceinline__ void tile_sgs(const TileRows<TS>& T, int i0, int j0, const float* rs, float* zs) {
    const int li = threadIdx.x % TS::W;
#pragma unroll
    for (int s = 0; s < TS::RPT; s++) zs[(threadIdx.x / TS::W + TS::RSTEP * s) * TS::W + li] = 0.0f;
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2 * SGS_SWEEPS + 1; h++) {
        const int color = h & 1;   // red (i + j even), black, red, ...
#pragma unroll
        for (int s = 0; s < TS::RPT; s++) {
            const int lj = threadIdx.x / TS::W + TS::RSTEP * s;
            if (((i0 + li + j0 + lj) & 1) != color || T.c[s] == CLS_NOT_FREE || !T.in[s]) continue;
            const int l = lj * TS::W + li;
            float acc = rs[l];
            if (T.cE[s] != 0.0f) acc += T.cE[s] * zs[l + 1];
            if (T.cW[s] != 0.0f) acc += T.cW[s] * zs[l - 1];
            if (T.cN[s] != 0.0f) acc += T.cN[s] * zs[l + TS::W];
            if (T.cS[s] != 0.0f) acc += T.cS[s] * zs[l - TS::W];
            zs[l] = acc / T.Dv[s];
        }
        __syncthreads();
    }
}

// z = M^-1 res of the refinement residual and the (r.z, r.r) partials of iteration 0; clears the
// PCG state word for the iterations queued behind it.
template <class TS>
__global__ void __launch_bounds__(TS::NT) pcg_precond(MeshP m, const uint8_t* __restrict__ cls,
                                                   const float* __restrict__ D, const float* __restrict__ res,
                                                   float* __restrict__ z, double* __restrict__ pz0,
                                                   int* __restrict__ istate) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        istate[0] = 0;
        istate[1] = 0;
    }
    __shared__ float rs[TS::H * TS::W], zs[TS::H * TS::W];
    const int nbx = (m.ni + TS::W - 1) / TS::W;
    const int i0 = (blockIdx.x % nbx) * TS::W, j0 = (blockIdx.x / nbx) * TS::H;
    const int li = threadIdx.x % TS::W;
    TileRows<TS> T;
    tile_rows<TS>(m, cls, D, i0, j0, T);
    float rv[TS::RPT];
#pragma unroll
    for (int s = 0; s < TS::RPT; s++) {
        const int lj = threadIdx.x / TS::W + TS::RSTEP * s;
        rv[s] = (T.in[s] && T.c[s] != CLS_NOT_FREE) ? res[(size_t)(j0 + lj) * m.ni + i0 + li] : 0.0f;
        rs[lj * TS::W + li] = rv[s];
    }
    tile_sgs<TS>(T, i0, j0, rs, zs);
    double a_rz = 0, a_rr = 0;
#pragma unroll
    for (int s = 0; s < TS::RPT; s++) {
        const int lj = threadIdx.x / TS::W + TS::RSTEP * s;
        if (!T.in[s]) continue;
        const float zv = zs[lj * TS::W + li];
        z[(size_t)(j0 + lj) * m.ni + i0 + li] = zv;
        a_rz += (double)rv[s] * (double)zv;
        a_rr += (double)rv[s] * (double)rv[s];
    }
    double v[2] = {a_rz, a_rr};
    block_partials<2>(v, pz0);
}

// step 2: alpha = rz_k / p.q; d += alpha p; res -= alpha q; z = M^-1 res; (r.z, r.r) partials
// of iteration k + 1 into pz_next.
template <class TS>
__global__ void __launch_bounds__(TS::NT) pcg_update(MeshP m, const uint8_t* __restrict__ cls,
                                                  const float* __restrict__ D, const float* __restrict__ p,
                                                  const float* __restrict__ q, float* __restrict__ d,
                                                  float* __restrict__ res, float* __restrict__ z,
                                                  const double* __restrict__ ppq, int nbq,
                                                  const double* __restrict__ pz_cur, int nbz,
                                                  double* __restrict__ pz_next, const int* istate) {
    if (istate[0]) return;
    double pq[1], cur[2];
    combine_partials<1>(ppq, nbq, pq);
    combine_partials<2>(pz_cur, nbz, cur);
    const float alpha = (float)(cur[0] / pq[0]);
    __shared__ float rs[TS::H * TS::W], zs[TS::H * TS::W];
    const int nbx = (m.ni + TS::W - 1) / TS::W;
    const int i0 = (blockIdx.x % nbx) * TS::W, j0 = (blockIdx.x / nbx) * TS::H;
    const int li = threadIdx.x % TS::W;
    TileRows<TS> T;
    tile_rows<TS>(m, cls, D, i0, j0, T);
    float rv[TS::RPT];
#pragma unroll
    for (int s = 0; s < TS::RPT; s++) {
        const int lj = threadIdx.x / TS::W + TS::RSTEP * s;
        rv[s] = 0.0f;
        if (T.in[s]) {
            const size_t u = (size_t)(j0 + lj) * m.ni + i0 + li;
            const float pv = p[u], qv = q[u];
            d[u] = d[u] + alpha * pv;
            rv[s] = res[u] - alpha * qv;
            res[u] = rv[s];
            if (T.c[s] == CLS_NOT_FREE) rv[s] = 0.0f;
        }
        rs[lj * TS::W + li] = rv[s];
    }
    tile_sgs<TS>(T, i0, j0, rs, zs);
    double a_rz = 0, a_rr = 0;
#pragma unroll
    for (int s = 0; s < TS::RPT; s++) {
        const int lj = threadIdx.x / TS::W + TS::RSTEP * s;
        if (!T.in[s]) continue;
        const float zv = zs[lj * TS::W + li];
        z[(size_t)(j0 + lj) * m.ni + i0 + li] = zv;
        a_rz += (double)rv[s] * (double)zv;
        a_rr += (double)rv[s] * (double)rv[s];
    }
    double v[2] = {a_rz, a_rr};
    block_partials<2>(v, pz_next);
}

// ------------------------------------------------------------------------------------------
// Device-driven Newton-PCG (warm solves).  The host queues pairs of kernels (nk_step_a,
// nk_step_b) without reading anything back in between; each pair is one PCG iteration or one
// Newton update, decided by every block from the same partial sums and the state the previous
// pair left in a small ring (state[i % NK_RING] is written by block 0 of pair i's second
// kernel and read by pair i + 1).  Pair i keeps its (r.z, r.r) partials in pz[i & 1] and its
// search direction in p[i & 1], so no kernel reads a buffer a block of the same kernel writes.
//   A (tile kernel): at the start of a linear solve (k = 0) the partials hold ||r||^2 of the
//     Newton residual: below (rel_tol ||r_ref||)^2 the solve is converged; otherwise the PCG
//     tolerance is set and the first direction computed.  Later: if the PCG residual is below
//     the tolerance (or k reached max_it), x += limited d (the Newton update, Solver.java:366-368
//     with the step limit of newton_update); else p = z + beta p_old, q = J p, p.q partials.
//   B (tile kernel): after a Newton update, the FP64 residual r = f + bfix - L x - W b(x), the
//     Jacobian diagonal, d = 0, res = r, z = M^-1 r and the partials of iteration 0 (r.r in
//     FP64); else the PCG update d += alpha p, res -= alpha q, z = M^-1 res and its partials.
// ------------------------------------------------------------------------------------------
struct NkParams {
    const double* ref2;   // ||r_ref||^2 on the device (the finalize of ref_norm_kernel)
    double rel2;          // rel_tol^2
    double eta2;          // squared forcing term (max(eta, 1e-5))^2
    double lim;           // Newton step limit (kTe)
    int max_it, nl_max;
};

enum { NK_PCG = 0, NK_UPDATE = 1, NK_STOP = 2 };
enum { NK_CONVERGED = 1, NK_CAP = 2, NK_NOT_FINITE = 3, NK_COLD = 4 };
// ||r_ref||^2 as the host path uses it (1 when the reference residual vanishes)
__device__ __forceinline__ double nk_ref2(const NkParams& np) {
    const double r = *np.ref2;
    return r > 0 ? r : 1.0;
}
// The decision of pair i from the previous state and the partials of the previous pair.  A solve
// whose first residual is above half the reference residual is not warm: it stops with NK_COLD
// before any update, and the host runs it on its own loop instead.
__device__ __forceinline__ int nk_decide(const NkState& S, const NkParams& np, const double cur[2], int& code) {
    code = 0;
    if (S.done) return NK_STOP;
    if (S.k == 0) {
        const double R = nk_ref2(np);
        if (!(cur[1] == cur[1]) || cur[1] > 1e300) { code = NK_NOT_FINITE; return NK_STOP; }
        if (cur[1] <= np.rel2 * R) { code = NK_CONVERGED; return NK_STOP; }
        if (S.newton == 0 && cur[1] > 0.25 * R) { code = NK_COLD; return NK_STOP; }
        if (S.newton >= np.nl_max) { code = NK_CAP; return NK_STOP; }
        return NK_PCG;
    }
    if (!(cur[1] > S.tol2) || S.k >= np.max_it) return NK_UPDATE;
    return NK_PCG;
}

template <class TS>
__global__ void __launch_bounds__(TS::NT) nk_step_a(MeshP m, const uint8_t* __restrict__ cls,
                                                 const float* __restrict__ D, const float* __restrict__ z,
                                                 float* __restrict__ p_even, float* __restrict__ p_odd,
                                                 float* __restrict__ q, double* __restrict__ x,
                                                 const float* __restrict__ d, const double* __restrict__ pz_even,
                                                 const double* __restrict__ pz_odd, int nbz,
                                                 double* __restrict__ ppq, const NkState* __restrict__ state,
                                                 double* __restrict__ nkc, NkParams np, int i) {
    const NkState S = state[(i - 1) % NK_RING];
    if (S.done) return;
    double cur[2];
    combine_partials<2>((i - 1) & 1 ? pz_odd : pz_even, nbz, cur);
    // every block combines the same partials in the same order; block 0 keeps the totals for
    // nk_step_b of this pair and nk_step_a of the next one
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        nkc[2 * (i % NK_RING)] = cur[0];
        nkc[2 * (i % NK_RING) + 1] = cur[1];
    }
    int code;
    const int act = nk_decide(S, np, cur, code);
    if (act == NK_STOP) return;
    const int nbx = (m.ni + TS::W - 1) / TS::W;
    const int bx = blockIdx.x % nbx, by = blockIdx.x / nbx;
    if (act == NK_UPDATE) {
        // x += limited d on the tile's free nodes
        const int li = threadIdx.x % TS::W;
        for (int lj = threadIdx.x / TS::W; lj < TS::H; lj += TS::RSTEP) {
            const int ii = bx * TS::W + li, jj = by * TS::H + lj;
            if (ii >= m.ni || jj >= m.nj) continue;
            const size_t u = (size_t)jj * m.ni + ii;
            if (cls[u] == CLS_NOT_FREE) continue;
            double v = (double)d[u];
            if (np.lim > 0) {
                if (v > np.lim) v = np.lim * (1.0 + log(v / np.lim));
                v = fmax(v, -10.0 * np.lim);
            }
            x[u] += v;
        }
        return;
    }
    float beta = 0.0f;
    if (S.k > 0) beta = (float)(cur[0] / nkc[2 * ((i - 1) % NK_RING)]);   // r.z of pair i-2's partials
    const float* pold = (i - 1) & 1 ? p_odd : p_even;
    float* pnew = i & 1 ? p_odd : p_even;
    __shared__ float ps[(TS::H + 2) * (TS::W + 2)];
    const int i0 = bx * TS::W - 1, j0 = by * TS::H - 1;
    for (int t = threadIdx.x; t < (TS::H + 2) * (TS::W + 2); t += blockDim.x) {
        const int ii = i0 + t % (TS::W + 2), jj = j0 + t / (TS::W + 2);
        float v = 0.0f;
        if (ii >= 0 && jj >= 0 && ii < m.ni && jj < m.nj) {
            const size_t u = (size_t)jj * m.ni + ii;
            v = S.k == 0 ? z[u] : z[u] + beta * pold[u];
        }
        ps[t] = v;
    }
    __syncthreads();
    const int li = threadIdx.x % TS::W;
    double acc = 0;
    for (int lj = threadIdx.x / TS::W; lj < TS::H; lj += TS::RSTEP) {
        const int ii = i0 + 1 + li, jj = j0 + 1 + lj;
        if (ii >= m.ni || jj >= m.nj) continue;
        const size_t u = (size_t)jj * m.ni + ii;
        const uint8_t c = cls[u];
        const int l = (lj + 1) * (TS::W + 2) + (li + 1);
        const float pv = ps[l];
        float qv = 0.0f;
        if (c != CLS_NOT_FREE) {
            const Stencil st = stencil(m, ii, jj);
            qv = D[u] * pv;
            if (nbc(c, 0) == NB_FREE) qv -= (float)st.cE * ps[l + 1];
            if (nbc(c, 1) == NB_FREE) qv -= (float)st.cW * ps[l - 1];
            if (nbc(c, 2) == NB_FREE) qv -= (float)st.cN * ps[l + TS::W + 2];
            if (nbc(c, 3) == NB_FREE) qv -= (float)st.cS * ps[l - TS::W - 2];
        }
        pnew[u] = pv;
        q[u] = qv;
        acc += (double)pv * (double)qv;
    }
    double v[1] = {acc};
    block_partials<1>(v, ppq);
}

// init = 1: the residual branch unconditionally (the state of pair 0, written to state[0]).
template <class TS>
__global__ void __launch_bounds__(TS::NT, TS::MINW) nk_step_b(MeshP m, const uint8_t* __restrict__ cls,
                                                 const double* __restrict__ W, const double* __restrict__ bfix,
                                                 const double* __restrict__ rho, const double* __restrict__ x,
                                                 double* __restrict__ r, float* __restrict__ D,
                                                 double* __restrict__ P, float* __restrict__ d,
                                                 float* __restrict__ res, float* __restrict__ z,
                                                 const float* __restrict__ p_even, const float* __restrict__ p_odd,
                                                 const float* __restrict__ q, const double* __restrict__ ppq, int nbq,
                                                 double* __restrict__ pz_even, double* __restrict__ pz_odd, int nbz,
                                                 NkState* __restrict__ state, const double* __restrict__ nkc,
                                                 NkParams np, double inv_eps, double Cn0, double phi0, double inv_kTe,
                                                 int i, int init) {
    NkState S{0.0, 0.0, 0, 0, 0, 0};
    int act = NK_UPDATE;   // init: compute the residual
    double cur[2] = {0.0, 0.0};
    if (!init) {
        S = state[(i - 1) % NK_RING];
        if (!S.done) {   // the totals nk_step_a of this pair combined
            cur[0] = nkc[2 * (i % NK_RING)];
            cur[1] = nkc[2 * (i % NK_RING) + 1];
        }
        int code;
        act = nk_decide(S, np, cur, code);
        if (act == NK_STOP) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                NkState o = S;
                if (!S.done) {
                    o.done = code;
                    o.r2 = cur[1];
                }
                state[i % NK_RING] = o;
            }
            return;
        }
    }
    float alpha = 0.0f;
    if (act == NK_PCG) {
        double pq[1];
        combine_partials<1>(ppq, nbq, pq);
        alpha = (float)(cur[0] / pq[0]);
    }
    const float* p = i & 1 ? p_odd : p_even;
    __shared__ float rs[TS::H * TS::W], zs[TS::H * TS::W];
    const int nbx = (m.ni + TS::W - 1) / TS::W;
    const int i0 = (blockIdx.x % nbx) * TS::W, j0 = (blockIdx.x / nbx) * TS::H;
    const int li = threadIdx.x % TS::W;
    double rr64[TS::RPT] = {};
    float rv[TS::RPT];
    if (act == NK_UPDATE) {
        // FP64 Newton residual and Jacobian diagonal of the tile (the newton_residual kernel's rows)
#pragma unroll
        for (int s4 = 0; s4 < TS::RPT; s4++) {
            const int lj = threadIdx.x / TS::W + TS::RSTEP * s4;
            const int ii = i0 + li, jj = j0 + lj;
            rv[s4] = 0.0f;
            if (ii < m.ni && jj < m.nj) {
                const size_t u = (size_t)jj * m.ni + ii;
                const uint8_t c = cls[u];
                d[u] = 0.0f;
                if (c == CLS_NOT_FREE) {
                    r[u] = 0;
                    D[u] = 1.0f;
                    P[u] = 0;
                    res[u] = 0.0f;
                } else {
                    double diagL;
                    const double Lx = lred_apply64(m, c, ii, jj, u, x, diagL);
                    const double Wu = W[u];
                    const double bx = Cn0 * exp((x[u] - phi0) * inv_kTe);
                    const double bp = bx * inv_kTe;
                    const double ru = Wu * rho[u] * inv_eps + bfix[u] - Lx - Wu * bx;
                    r[u] = ru;
                    P[u] = Wu * bp;
                    D[u] = (float)(diagL + Wu * bp);
                    rv[s4] = (float)ru;
                    res[u] = rv[s4];
                    rr64[s4] = ru;
                }
            }
            rs[lj * TS::W + li] = rv[s4];
        }
        __syncthreads();   // the tile's D values, read by tile_rows below
    }
    TileRows<TS> T;
    tile_rows<TS>(m, cls, D, i0, j0, T);
    if (act == NK_PCG) {
#pragma unroll
        for (int s4 = 0; s4 < TS::RPT; s4++) {
            const int lj = threadIdx.x / TS::W + TS::RSTEP * s4;
            rv[s4] = 0.0f;
            if (T.in[s4]) {
                const size_t u = (size_t)(j0 + lj) * m.ni + i0 + li;
                const float pv = p[u], qv = q[u];
                d[u] = d[u] + alpha * pv;
                rv[s4] = res[u] - alpha * qv;
                res[u] = rv[s4];
                if (T.c[s4] == CLS_NOT_FREE) rv[s4] = 0.0f;
            }
            rs[lj * TS::W + li] = rv[s4];
        }
    }
    tile_sgs<TS>(T, i0, j0, rs, zs);
    double a_rz = 0, a_rr = 0;
#pragma unroll
    for (int s4 = 0; s4 < TS::RPT; s4++) {
        const int lj = threadIdx.x / TS::W + TS::RSTEP * s4;
        if (!T.in[s4]) continue;
        const float zv = zs[lj * TS::W + li];
        z[(size_t)(j0 + lj) * m.ni + i0 + li] = zv;
        if (act == NK_UPDATE) {
            a_rz += rr64[s4] * (double)zv;
            a_rr += rr64[s4] * rr64[s4];
        } else {
            a_rz += (double)rv[s4] * (double)zv;
            a_rr += (double)rv[s4] * (double)rv[s4];
        }
    }
    double v[2] = {a_rz, a_rr};
    block_partials<2>(v, i & 1 ? pz_odd : pz_even);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        NkState o = S;
        if (act == NK_UPDATE) {
            if (!init) o.newton = S.newton + 1;
            o.k = 0;
        } else {
            if (S.k == 0) {   // the linear solve's tolerance from its first residual
                o.r2 = cur[1];
                o.tol2 = fmax(np.eta2 * cur[1], 0.25 * np.rel2 * nk_ref2(np));
            }
            o.k = S.k + 1;
            o.pcg = S.pcg + 1;
        }
        state[i % NK_RING] = o;
    }
}

template <class TS>
static int tile_count(int ni, int nj) { return ((ni + TS::W - 1) / TS::W) * ((nj + TS::H - 1) / TS::H); }
// Wide tiles while the mesh needs at most one of their workgroups per CU.  nk_step_b's 69 registers
// leave room for one 16-wave workgroup per CU, so more wide tiles than CUs run in rounds with no
// second workgroup to overlap a round's barriers and tail.  Measured warm solves: 1024 x 512 (256
// wide tiles) 0.62 ms with wide tiles, 0.65 with narrow ones; 2048 x 1024 (1024 wide tiles) 21.9 ms
// with wide tiles, 17.6 with narrow ones held to 64 registers (17.9 with wide ones held to 64).
static bool choose_tile_shape(int ni, int nj) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
        ncu = 256;
    return tile_count<TileWide>(ni, nj) <= ncu;
}
template <class F>
static void with_tile_shape(bool wide, F&& f) {
    if (wide) f(TileWide{});
    else f(TileNarrow{});
}

// y += d (FP64 accumulation of the FP32 correction)
__global__ void accumulate_y(int64_t nn, const float* __restrict__ d, double* __restrict__ y,
                             const int* __restrict__ go) {
    if (go != nullptr && go[0] != 1) return;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x)
        y[u] += (double)d[u];
}
// max |y| over the grid (for the Newton step limit)
__global__ void absmax_kernel(int64_t nn, const double* __restrict__ y, double* __restrict__ partial) {
    double acc = 0;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x)
        acc = fmax(acc, fabs(y[u]));
    double v[1] = {acc};
    block_partials<1, true>(v, partial);
}

// x += limit(y) (Newton update, Solver.java:366-368 with the sign of the reduced system); y = 0.
// kte > 0 limits each node's update: a rise above kTe is damped to kTe (1 + ln(y / kTe)), a fall
// is clipped at -10 kTe; kte <= 0 takes the full step.
// x += limited y on the free nodes; y is cleared.  ypart receives per-block partial sums of y^2
// (the Newton step before the limit), the quantity of the reference's convergence test
// ||y||_2 / N < nl_tol (Solver.java:371-381, Vec.norm).
// dlast (may be null): the last FP32 PCG correction, added to y first (y += d fused in).
__global__ void __launch_bounds__(256) newton_update(int64_t nn, const uint8_t* __restrict__ cls,
                                                     double* __restrict__ x, double* __restrict__ y,
                                                     const float* __restrict__ dlast, double kte,
                                                     const int* __restrict__ go, double* __restrict__ ypart) {
    if (go != nullptr && go[0] != 1) return;
    double v2[1] = {0.0};
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        double v = y[u];
        if (dlast != nullptr) v += (double)dlast[u];
        if (cls[u] != CLS_NOT_FREE) v2[0] += v * v;
        if (kte > 0) {
            if (v > kte) v = kte * (1.0 + log(v / kte));
            v = fmax(v, -10.0 * kte);
        }
        if (cls[u] != CLS_NOT_FREE) x[u] += v;
        y[u] = 0;
    }
    block_partials<1>(v2, ypart);
}

// full potential from the free-node solution: fixed nodes take their value, face nodes the
// value of the node their one-sided row ties them to
__global__ void expand_phi(int64_t nn, const int32_t* __restrict__ src, const double* __restrict__ fixval,
                           const double* __restrict__ x, double* __restrict__ phi) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        int32_t s = src[u];
        phi[u] = s < 0 ? fixval[u] : x[s] + fixval[u];
    }
}

// E = -grad(phi): Solver.evaluateGradient with the assembled Gi/Gj rows, which on a uniform
// mesh are central differences inside and one-sided differences on the mesh faces.
__global__ void efield_kernel(MeshP m, const double* __restrict__ phi, double* __restrict__ efi,
                              double* __restrict__ efj, float2* __restrict__ ef32) {
    const int64_t nn = (int64_t)m.ni * m.nj;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(u % m.ni), j = (int)(u / m.ni);
        double gi, gj;
        if (i == 0) gi = (phi[u + 1] - phi[u]) * m.inv_dh[0];
        else if (i == m.ni - 1) gi = (phi[u] - phi[u - 1]) * m.inv_dh[0];
        else gi = (phi[u + 1] - phi[u - 1]) * 0.5 * m.inv_dh[0];
        if (j == 0) gj = (phi[u + m.ni] - phi[u]) * m.inv_dh[1];
        else if (j == m.nj - 1) gj = (phi[u] - phi[u - m.ni]) * m.inv_dh[1];
        else gj = (phi[u + m.ni] - phi[u - m.ni]) * 0.5 * m.inv_dh[1];
        efi[u] = -gi;
        efj[u] = -gj;
        ef32[u] = make_float2((float)(-gi), (float)(-gj));
    }
}

// y = A_ref x on the full grid (reference rows), for the operator parity check
__global__ void apply_reference_operator(MeshP m, const uint8_t* __restrict__ rowtype, const double* __restrict__ x,
                                         double* __restrict__ y) {
    const int64_t nn = (int64_t)m.ni * m.nj;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(u % m.ni), j = (int)(u / m.ni);
        uint8_t t = rowtype[u];
        double v;
        switch (t) {
            case 0: v = x[u]; break;                                           // identity
            case 1: v = (x[u + 1] - x[u]) * m.inv_dh[0]; break;                  // Gi, left face
            case 2: v = (x[u] - x[u - 1]) * m.inv_dh[0]; break;                  // Gi, right face
            case 3: v = (x[u + m.ni] - x[u]) * m.inv_dh[1]; break;               // Gj, bottom face
            case 4: v = (x[u] - x[u - m.ni]) * m.inv_dh[1]; break;               // Gj, top face
            default: {
                Stencil s = stencil(m, i, j);
                double Wu = row_weight(m, i, j);
                v = (s.cE * (x[u + 1] - x[u]) + s.cW * (x[u - 1] - x[u]) + s.cN * (x[u + m.ni] - x[u]) +
                     s.cS * (x[u - m.ni] - x[u])) / Wu;
            }
        }
        y[u] = v;
    }
}

// ----------------------------------------------------------------------------------------
// host
// ----------------------------------------------------------------------------------------
static int grid_elem(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 512) b = 512;
    if (b < 1) b = 1;
    return (int)b;
}

// row types of the reference assembly: 0 identity, 1..4 one-sided gradient rows, 5 Laplacian
static int reference_row_types(sf_ctx* ctx, std::vector<uint8_t>& rt) {
    const MeshP& m = ctx->mp;
    const int ni = m.ni, nj = m.nj;
    rt.assign((size_t)ni * nj, 5);
    auto face = [&](int f, int idx) {
        int off = f == 0 ? 0 : (f == 1 ? nj : (f == 2 ? nj + ni : 2 * nj + ni));
        return ctx->h_face_type[off + idx];
    };
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
            size_t u = (size_t)j * ni + i;
            int nt = ctx->h_node_type[u];
            if (nt == SF_NODE_DIRICHLET || nt == SF_NODE_INTERNAL) { rt[u] = 0; continue; }
            if (nt == SF_NODE_INSULATOR)
                return sf_fail(ctx, SF_ERR_UNSUPPORTED, "insulator nodes are not supported by the GPU solver");
            if (i == 0 && face(2, j) != SF_FACE_MESH) { rt[u] = 1; continue; }
            if (i == ni - 1 && face(0, j) != SF_FACE_MESH) { rt[u] = 2; continue; }
            if (j == 0 && face(3, i) != SF_FACE_MESH) { rt[u] = 3; continue; }
            if (j == nj - 1 && face(1, i) != SF_FACE_MESH) { rt[u] = 4; continue; }
            if (i == 0 || j == 0 || i == ni - 1 || j == nj - 1)
                return sf_fail(ctx, SF_ERR_UNSUPPORTED, "MESH faces (multi-domain) are not supported");
        }
    return SF_OK;
}

int poisson_setup(sf_ctx* ctx, const sf_poisson_desc* dsc) {
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    const MeshP& m = ctx->mp;
    const int ni = m.ni, nj = m.nj;
    const size_t nn = (size_t)ni * nj;
    std::vector<uint8_t> rt;
    int rc = reference_row_types(ctx, rt);
    if (rc) return rc;
    // resolve the value source of every node: free nodes are their own source, fixed nodes
    // have none (-1), one-sided rows follow their tie (chains at corners)
    std::vector<int32_t> src(nn);
    for (size_t u = 0; u < nn; u++) src[u] = rt[u] == 0 ? -1 : (int32_t)u;
    auto tie = [&](size_t u) -> int64_t {
        switch (rt[u]) {
            case 1: return (int64_t)u + 1;
            case 2: return (int64_t)u - 1;
            case 3: return (int64_t)u + ni;
            case 4: return (int64_t)u - ni;
        }
        return -2;
    };
    // A one-sided face row reads (phi_in - phi_f)/dh = b_f (left, bottom) or (phi_f - phi_in)/dh = b_f
    // (right, top), so phi_f = phi_in + off_f.  In linear mode b_f is the node's bc_value
    // (PoissonSolver.update, PoissonSolver.java:140-153); in nonlinear mode Solver.solveNonLinearNR
    // zeroes b on these nodes (Solver.java:325-342) and every tie is exact (off = 0).  fixval
    // holds the value of fixed nodes and the offset of tied ones.
    const bool lin = dsc->linear != 0;
    auto step_off = [&](size_t v) -> double {
        if (!lin) return 0.0;
        const double b = ctx->h_node_bc[v];
        switch (rt[v]) {
            case 1: return -m.dh[0] * b;
            case 2: return m.dh[0] * b;
            case 3: return -m.dh[1] * b;
            case 4: return m.dh[1] * b;
        }
        return 0.0;
    };
    std::vector<double> fixval(nn, 0.0);
    for (size_t u = 0; u < nn; u++)
        if (rt[u] == 0) fixval[u] = ctx->h_node_bc[u];
    for (size_t u = 0; u < nn; u++) {
        if (rt[u] == 0 || rt[u] == 5) continue;
        size_t v = u;
        int guard = 0;
        double off = 0.0;
        while (rt[v] >= 1 && rt[v] <= 4 && guard++ < 8) {
            off += step_off(v);
            v = (size_t)tie(v);
        }
        if (rt[v] >= 1 && rt[v] <= 4) return sf_fail(ctx, SF_ERR_UNSUPPORTED, "cyclic face rows at node %zu", u);
        if (rt[v] == 0) { src[u] = -1; fixval[u] = ctx->h_node_bc[v] + off; }
        else { src[u] = (int32_t)v; fixval[u] = off; }
    }
    std::vector<uint8_t> cls(nn, CLS_NOT_FREE);
    std::vector<double> W(nn, 0.0), bfix(nn, 0.0);
    int nfree = 0;
    for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
            size_t u = (size_t)j * ni + i;
            if (rt[u] != 5) continue;
            nfree++;
            Stencil s = stencil(m, i, j);
            const size_t nb[4] = {u + 1, u - 1, u + ni, u - ni};
            const double cc[4] = {s.cE, s.cW, s.cN, s.cS};
            uint8_t c = 0;
            for (int d = 0; d < 4; d++) {
                size_t v = nb[d];
                int kind;
                if (rt[v] == 5) kind = NB_FREE;
                else if (src[v] < 0) { kind = NB_FIXED; bfix[u] += cc[d] * fixval[v]; }
                else if ((size_t)src[v] == u) { kind = NB_SELF; bfix[u] += cc[d] * fixval[v]; }
                else return sf_fail(ctx, SF_ERR_UNSUPPORTED, "node %zu couples to a face row of another node", u);
                c |= (uint8_t)(kind << (2 * d));
            }
            cls[u] = c;
            W[u] = row_weight(m, i, j);
        }
    PoissonState* P = ctx->poisson;
    if (!P) {
        P = new PoissonState();
        ctx->poisson = P;
        auto A = [&](void** p, size_t bytes) { return hipMalloc(p, bytes ? bytes : 8); };
        hipError_t e = hipSuccess;
        P->wide_tiles = choose_tile_shape(ni, nj);
        const int nb_tile = P->wide_tiles ? tile_count<TileWide>(ni, nj) : tile_count<TileNarrow>(ni, nj);
        P->nblk_tile = nb_tile;
        P->nblk_elem = grid_elem((int64_t)nn);
        size_t npart = (size_t)std::max(nb_tile, P->nblk_elem) * 4 + 16;
        e = A((void**)&P->cls, nn);
        if (!e) e = A((void**)&P->W, nn * 8);
        if (!e) e = A((void**)&P->bfix, nn * 8);
        if (!e) e = A((void**)&P->src, nn * 4);
        if (!e) e = A((void**)&P->fixval, nn * 8);
        if (!e) e = A((void**)&P->x, nn * 8);
        if (!e) e = A((void**)&P->r, nn * 8);
        if (!e) e = A((void**)&P->y, nn * 8);
        if (!e) e = A((void**)&P->rr, nn * 8);
        if (!e) e = A((void**)&P->D, nn * 4);
        if (!e) e = A((void**)&P->P, nn * 8);
        if (!e) e = A((void**)&P->d, nn * 4);
        if (!e) e = A((void**)&P->res, nn * 4);
        if (!e) e = A((void**)&P->z, nn * 4);
        if (!e) e = A((void**)&P->p0, nn * 4);
        if (!e) e = A((void**)&P->p1, nn * 4);
        if (!e) e = A((void**)&P->q, nn * 4);
        if (!e) e = A((void**)&P->partial, npart * 8);
        if (!e) e = A((void**)&P->pz, (size_t)4 * std::max(P->nblk_elem, nb_tile) * 8);
        if (!e) e = A((void**)&P->ppq, (size_t)P->nblk_tile * 8);
        if (!e) e = A((void**)&P->scal, S_NSCAL * 8 * 2);
        if (!e) e = A((void**)&P->istate, 16);
        if (!e) e = hipHostMalloc((void**)&P->host_scal, S_NSCAL * 8 * 2);
        if (!e) e = hipHostMalloc((void**)&P->host_istate, 16);
        if (!e) e = A((void**)&P->nk, NK_RING * sizeof(NkState));
        if (!e) e = A((void**)&P->nkc, NK_RING * 2 * sizeof(double));
        if (!e) e = hipHostMalloc((void**)&P->host_nk, sizeof(NkState));
        if (!e) e = hipEventCreate(&P->e0);
        if (!e) e = hipEventCreate(&P->e1);
        if (e) return sf_fail(ctx, SF_ERR_OOM, "poisson allocation failed: %s", hipGetErrorString(e));
        hipMemset(P->istate, 0, 16);
        hipMemset(P->nk, 0, NK_RING * sizeof(NkState));
        hipMemset(P->nkc, 0, NK_RING * 2 * sizeof(double));
        P->host_newton = getenv("SF_HOST_NEWTON") != nullptr;
        hipMemset(P->pz, 0, (size_t)4 * std::max(P->nblk_elem, nb_tile) * 8);
        hipMemset(P->ppq, 0, (size_t)P->nblk_tile * 8);
        hipMemset(P->y, 0, nn * 8);
        hipMemset(P->x, 0, nn * 8);
        hipMemset(P->d, 0, nn * 4);
        hipMemset(P->res, 0, nn * 4);
        hipMemset(P->z, 0, nn * 4);
        hipMemset(P->p0, 0, nn * 4);
        hipMemset(P->p1, 0, nn * 4);
        hipMemset(P->q, 0, nn * 4);
    }
    P->desc = *dsc;
    P->eps = EPS0 * (dsc->eps_r > 0 ? dsc->eps_r : 1.0);
    P->C = QE / P->eps;
    P->nfree = nfree;
    SF_HIP(ctx, hipMemcpy(P->cls, cls.data(), nn, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(P->W, W.data(), nn * 8, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(P->bfix, bfix.data(), nn * 8, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(P->src, src.data(), nn * 4, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(P->fixval, fixval.data(), nn * 8, hipMemcpyHostToDevice));
    // phi starts with the Dirichlet values on fixed nodes (Node.bc_value)
    std::vector<double> phi(nn, 0.0);
    for (size_t u = 0; u < nn; u++) phi[u] = src[u] < 0 ? fixval[u] : 0.0;
    SF_HIP(ctx, hipMemcpy(ctx->d_phi, phi.data(), nn * 8, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipDeviceSynchronize());
    return SF_OK;
}

// total of the per-block partials left by the last kernel, into scal[slot..slot+NV)
template <int NV, bool MAX = false>
static void finalize(PoissonState* P, hipStream_t st, const double* partial, int nb, int slot) {
    hipLaunchKernelGGL((finalize_partials<NV, MAX>), dim3(1), dim3(256), 0, st, partial, nb, P->scal, slot);
}

static int read_scalars(sf_ctx* ctx, PoissonState* P) {
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipMemcpyAsync(P->host_scal, P->scal, S_NSCAL * 8, hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SF_OK;
}

int poisson_solve(sf_ctx* ctx, sf_solve_info* info) {
    PoissonState* P = ctx->poisson;
    if (!P) return sf_fail(ctx, SF_ERR_STATE, "poisson solver not set up");
    const MeshP& m = ctx->mp;
    const int64_t nn = (int64_t)m.ni * m.nj;
    const sf_poisson_desc& D = P->desc;
    hipStream_t st = ctx->stream;
    const int ge = P->nblk_elem, gt = P->nblk_tile;
    const double inv_eps = 1.0 / P->eps;
    const double Cn0 = D.linear ? 0.0 : P->C * D.n0;
    const double inv_kTe = D.linear ? 0.0 : 1.0 / D.Te0;
    const double rel_tol = D.rel_tol > 0 ? D.rel_tol : 1e-8;
    const int max_it = D.max_it > 0 ? D.max_it : 5000;
    const int nl_max = D.linear ? 1 : (D.nl_max_it > 0 ? D.nl_max_it : 50);
    double* pz0 = P->pz;
    double* pz1 = P->pz + 2 * std::max(ge, gt);   // partials of both parities, by block
    double* ypart = P->partial + 2 * std::max(ge, gt);   // Newton step norm partials (not overwritten by residuals)
    // the reference's Newton stop ||y||_2 / N < nl_tol (N = all nodes), used when nl_tol > 0
    const double nl_tol = D.linear ? 0.0 : D.nl_tol;
    bool nl_stop = false;
    SF_HIP(ctx, hipEventRecord(P->e0, st));
    hipLaunchKernelGGL(ref_norm_kernel, dim3(ge), dim3(256), 0, st, m, (const uint8_t*)P->cls, (const double*)P->W,
                       (const double*)P->bfix, (const double*)ctx->d_rho, inv_eps, Cn0, D.phi0, inv_kTe, D.linear,
                       P->partial, (const double*)ctx->d_phi, P->x);
    finalize<1>(P, st, P->partial, ge, S_R0);
    int newton = 0, pcg_total = 0;
    double rel = 0, rnorm = 0, ref = 1.0;
    bool converged = false;
    const bool debug = getenv("SF_DEBUG_SOLVE") != nullptr;
    const double lim0 = D.linear ? 0.0 : D.Te0;
    // A warm nonlinear solve (from near its solution) runs as device-driven Newton-PCG: the pairs
    // of step kernels are queued in batches and the host reads one state word per batch.  Returns
    // the stop code (NK_COLD: the solve was not warm and nothing was changed) or an error.
    const bool device_ok = !D.linear && !(nl_tol > 0) && !debug && !P->host_newton;
    auto device_newton = [&](double eta_d) -> int {
        const double ee = std::max(eta_d, 1e-5);
        NkParams np{(const double*)(P->scal + S_R0), rel_tol * rel_tol, ee * ee, lim0, max_it, nl_max};
        auto step_b = [&](int i, int init) {
            with_tile_shape(P->wide_tiles, [&](auto ts) {
                using TS = decltype(ts);
                hipLaunchKernelGGL(nk_step_b<TS>, dim3(gt), dim3(TS::NT), 0, st, m, (const uint8_t*)P->cls,
                                   (const double*)P->W, (const double*)P->bfix, (const double*)ctx->d_rho,
                                   (const double*)P->x, P->r, P->D, P->P, P->d, P->res, P->z, (const float*)P->p0,
                                   (const float*)P->p1, (const float*)P->q, (const double*)P->ppq, gt, pz0, pz1, gt,
                                   P->nk, (const double*)P->nkc, np, inv_eps, Cn0, D.phi0, inv_kTe, i, init);
            });
        };
        step_b(0, 1);
        int i = 1, batch = std::min(64, std::max(8, P->last_pairs + 2));
        while (true) {
            for (int b = 0; b < batch; b++, i++) {
                with_tile_shape(P->wide_tiles, [&](auto ts) {
                    using TS = decltype(ts);
                    hipLaunchKernelGGL(nk_step_a<TS>, dim3(gt), dim3(TS::NT), 0, st, m, (const uint8_t*)P->cls,
                                       (const float*)P->D, (const float*)P->z, P->p0, P->p1, P->q, P->x,
                                       (const float*)P->d, (const double*)pz0, (const double*)pz1, gt, P->ppq,
                                       (const NkState*)P->nk, P->nkc, np, i);
                });
                step_b(i, 0);
            }
            SF_HIP(ctx, hipGetLastError());
            SF_HIP(ctx, hipMemcpyAsync(P->host_nk, P->nk + (i - 1) % NK_RING, sizeof(NkState),
                                       hipMemcpyDeviceToHost, st));
            SF_HIP(ctx, hipMemcpyAsync(P->host_scal, P->scal, S_NSCAL * 8, hipMemcpyDeviceToHost, st));
            // phi from x, queued before the host reads the state: when the batch ended the solve the
            // potential is already written (otherwise the next batch's expansion overwrites it)
            hipLaunchKernelGGL(expand_phi, dim3(ge), dim3(256), 0, st, nn, (const int32_t*)P->src,
                               (const double*)P->fixval, (const double*)P->x, ctx->d_phi);
            SF_HIP(ctx, hipStreamSynchronize(st));
            if (P->host_nk->done) break;
            if (i > 4 * (max_it + 1) * (nl_max + 1)) return sf_fail(ctx, SF_ERR_STATE, "device Newton-PCG did not stop");
            batch = std::min(2 * batch, 256);
        }
        const NkState fin = *P->host_nk;
        if (fin.done == NK_COLD) return NK_COLD;
        P->last_pairs = fin.newton + fin.pcg + 1;
        ref = std::sqrt(P->host_scal[S_R0]);
        if (!(ref > 0)) ref = 1.0;
        newton = fin.newton;
        pcg_total = fin.pcg;
        rnorm = std::sqrt(fin.r2);
        rel = rnorm / ref;
        if (fin.done == NK_NOT_FINITE || !std::isfinite(rel))
            return sf_fail(ctx, SF_ERR_STATE, "Newton residual is not finite (diverged)");
        converged = fin.done == NK_CONVERGED;
        return fin.done;
    };
    const double eta_warm = P->wide_tiles ? ETA_WARM : ETA_WARM_LARGE;
    int rc = 0;
    // the last solve converged and this one starts from its potential: go straight to the device
    // loop (it hands a cold start back untouched)
    bool device_done = false;
    if (device_ok && P->warm_hint) {
        const int code = device_newton(eta_warm);
        if (code < 0) return code;
        device_done = code != NK_COLD;
    }
    if (!device_done) {
    rc = read_scalars(ctx, P);
    if (rc) return rc;
    double ref2 = P->host_scal[S_R0];
    ref = std::sqrt(ref2);
    if (!(ref > 0)) ref = 1.0;
    double eta = ETA_COLD;
    bool eta_chosen = false;
    // y is zero at the start of every Newton iteration (newton_update clears it), so the first
    // refinement residual r - J y is the Newton residual itself: one read serves both tests.
    SF_HIP(ctx, hipMemsetAsync(P->y, 0, nn * sizeof(double), st));
    // Newton step limit, node by node.  The linearised exponential underestimates how fast the
    // electron density grows with phi, so a large rise overshoots (and overflows exp when the
    // solve starts far below a dense plasma): rises beyond kTe are damped logarithmically,
    // falls are clipped at 10 kTe.  Nodes already close stay on the quadratic path, and the
    // limit changes the path only, not the discrete solution.  (Against a plain 10 kTe clip
    // this cuts the steady-state Newton iterations of the cylinder deck from 8.4 to 5.4.)
    const double lim = D.linear ? 0.0 : D.Te0;
    auto launch_residual = [&](const int* go) {
        hipLaunchKernelGGL(newton_residual, dim3(ge), dim3(256), 0, st, m, (const uint8_t*)P->cls,
                           (const double*)P->W, (const double*)P->bfix, (const double*)ctx->d_rho,
                           (const double*)P->x, P->r, P->D, P->P, P->d, P->res, inv_eps, Cn0, D.phi0, inv_kTe,
                           D.linear, P->partial, go);
        finalize<1>(P, st, P->partial, ge, S_NORM2);
    };
    bool have_residual = false;   // the residual of the current x is already in host_scal
    for (int it = 0; it <= nl_max; it++) {
        if (!have_residual) {
            launch_residual(nullptr);
            rc = read_scalars(ctx, P);
            if (rc) return rc;
        }
        have_residual = false;
        rnorm = std::sqrt(P->host_scal[S_NORM2]);
        rel = rnorm / ref;
        if (!std::isfinite(rel)) return sf_fail(ctx, SF_ERR_STATE, "Newton residual is not finite (diverged)");
        if (rel <= rel_tol) { converged = true; break; }
        if (nl_stop) { converged = true; break; }    // the reference's criterion met on the last step
        if (it == nl_max) break;
        if (!eta_chosen) {
            eta = rel < 0.5 ? eta_warm : ETA_COLD;
            eta_chosen = true;
        }
        if (it == 0 && rel < 0.5 && device_ok) {
            const int code = device_newton(eta);
            if (code < 0) return code;
            if (code != NK_COLD) break;
        }
        newton++;
        // inexact Newton: solve J y = r to max(eta*||r||, 0.5*tol*ref) with FP32 PCG + FP64 refinement
        double lin_target = std::max(D.linear ? 0.5 * rel_tol * ref : eta * rnorm, 0.5 * rel_tol * ref);
        int refine = 0;
        bool stepped = false;   // the update and the next residual ran behind the linear solve
        while (true) {
            double rr = rnorm;   // the first refinement residual was computed with the Newton residual
            if (refine > 0) {
                hipLaunchKernelGGL(refine_residual, dim3(ge), dim3(256), 0, st, m, (const uint8_t*)P->cls,
                                   (const double*)P->r, (const double*)P->y, (const double*)P->P, P->rr, P->d,
                                   P->res, P->partial, nullptr);
                finalize<1>(P, st, P->partial, ge, S_NORM2);
                rc = read_scalars(ctx, P);
                if (rc) return rc;
                rr = std::sqrt(P->host_scal[S_NORM2]);
            }
            if (rr <= lin_target || refine >= 8 || pcg_total >= max_it * nl_max) break;
            // inner FP32 tolerance: a relative reduction FP32 can deliver
            double inner_tol = std::max(lin_target, 1e-5 * rr);
            const double tol2 = inner_tol * inner_tol;
            with_tile_shape(P->wide_tiles, [&](auto ts) {
                using TS = decltype(ts);
                hipLaunchKernelGGL(pcg_precond<TS>, dim3(gt), dim3(TS::NT), 0, st, m, (const uint8_t*)P->cls,
                                   (const float*)P->D, (const float*)P->res, P->z, pz0, P->istate);
            });
            // When this linear solve is the Newton step's last (it aims at the step's own target),
            // the update and the next Newton residual are queued behind the first batch of PCG
            // iterations.  They run only if that batch converged (the kernels read istate), and
            // then one host synchronisation serves both the PCG check and the next residual.
            const bool ahead = !debug && inner_tol == lin_target;
            float* pa = P->p0;
            float* pb = P->p1;
            // iterations queued between host checks: one more than the last linear solve needed,
            // doubling while unconverged (a converged iteration turns the rest into no-ops)
            int k = 0, its = 0, batch = std::min(8, std::max(3, P->last_its + 1));
            bool first_batch = true, done_in_first = false;
            while (true) {
                const int nb = std::min(batch, max_it - k);
                for (int b = 0; b < nb; b++, k++) {
                    double* cur = (k & 1) ? pz1 : pz0;
                    double* nxt = (k & 1) ? pz0 : pz1;
                    with_tile_shape(P->wide_tiles, [&](auto ts) {
                        using TS = decltype(ts);
                        hipLaunchKernelGGL(pcg_spmv<TS>, dim3(gt), dim3(TS::NT), 0, st, m, (const uint8_t*)P->cls,
                                           (const float*)P->D, (const float*)P->z, (const float*)pa, pb, P->q,
                                           (const double*)cur, (const double*)nxt, gt, k, P->ppq, tol2, P->istate);
                        hipLaunchKernelGGL(pcg_update<TS>, dim3(gt), dim3(TS::NT), 0, st, m, (const uint8_t*)P->cls,
                                           (const float*)P->D, (const float*)pb, (const float*)P->q, P->d, P->res,
                                           P->z, (const double*)P->ppq, gt, (const double*)cur, gt, nxt,
                                           (const int*)P->istate);
                    });
                    std::swap(pa, pb);
                }
                const bool queued = ahead && first_batch;
                if (queued) {
                    const int* go = (const int*)P->istate;
                    hipLaunchKernelGGL(newton_update, dim3(ge), dim3(256), 0, st, nn, (const uint8_t*)P->cls, P->x,
                                       P->y, (const float*)P->d, lim, go, ypart);
                    if (nl_tol > 0) finalize<1>(P, st, ypart, ge, S_YNORM2);
                    launch_residual(go);
                }
                SF_HIP(ctx, hipGetLastError());
                SF_HIP(ctx, hipMemcpyAsync(P->host_istate, P->istate, 8, hipMemcpyDeviceToHost, st));
                if (queued) SF_HIP(ctx, hipMemcpyAsync(P->host_scal, P->scal, S_NSCAL * 8, hipMemcpyDeviceToHost, st));
                SF_HIP(ctx, hipStreamSynchronize(st));
                if (P->host_istate[0]) {
                    its = P->host_istate[1];
                    done_in_first = first_batch;
                    break;
                }
                if (k >= max_it) { its = k; break; }
                batch = std::min(2 * batch, 64);
                first_batch = false;
            }
            pcg_total += its;
            P->last_its = its;
            if (ahead && done_in_first && P->host_istate[0] == 1) {
                stepped = true;
                break;
            }
            hipLaunchKernelGGL(accumulate_y, dim3(ge), dim3(256), 0, st, nn, (const float*)P->d, P->y, nullptr);
            refine++;
            // PCG aimed at the Newton step's own target (far above FP32's drift) and reached it:
            // no FP64 check is needed; the next Newton residual verifies the step anyway.
            if (P->host_istate[0] == 1 && inner_tol == lin_target) break;
        }
        if (stepped) {
            have_residual = true;
            SF_HIP(ctx, hipGetLastError());
            if (nl_tol > 0 && std::sqrt(P->host_scal[S_YNORM2]) / (double)nn < nl_tol) nl_stop = true;
            continue;
        }
        if (debug && !D.linear) {
            hipLaunchKernelGGL(absmax_kernel, dim3(ge), dim3(256), 0, st, nn, (const double*)P->y, P->partial);
            finalize<1, true>(P, st, P->partial, ge, S_YMAX);
        }
        hipLaunchKernelGGL(newton_update, dim3(ge), dim3(256), 0, st, nn, (const uint8_t*)P->cls, P->x, P->y,
                           (const float*)nullptr, lim, nullptr, ypart);
        if (nl_tol > 0) finalize<1>(P, st, ypart, ge, S_YNORM2);
        if (nl_tol > 0) {
            rc = read_scalars(ctx, P);
            if (rc) return rc;
            if (std::sqrt(P->host_scal[S_YNORM2]) / (double)nn < nl_tol) nl_stop = true;
        }
        if (debug) {
            rc = read_scalars(ctx, P);
            if (rc) return rc;
            fprintf(stderr, "[newton %d] rel %.3e lin_target/ref %.3e refine %d pcg %d ymax %.3e\n", it, rel,
                    lin_target / ref, refine, pcg_total, D.linear ? 0.0 : P->host_scal[S_YMAX]);
        }
        SF_HIP(ctx, hipGetLastError());
    }
    }   // !device_done
    P->warm_hint = converged;
    if (!device_done)   // (the device loop expanded phi with its last batch)
        hipLaunchKernelGGL(expand_phi, dim3(ge), dim3(256), 0, st, nn, (const int32_t*)P->src,
                           (const double*)P->fixval, (const double*)P->x, ctx->d_phi);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipEventRecord(P->e1, st));
    SF_HIP(ctx, hipEventSynchronize(P->e1));
    float ms = 0;
    hipEventElapsedTime(&ms, P->e0, P->e1);
    if (info) {
        info->newton_iterations = newton;
        info->pcg_iterations = pcg_total;
        info->rel_residual = rel;
        info->ref_norm = rnorm / (double)nn;
        info->ms = ms;
    }
    return converged ? SF_OK : SF_WARN_NOT_CONVERGED;
}

int efield_update(sf_ctx* ctx) {
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    hipLaunchKernelGGL(efield_kernel, dim3(grid_elem(nn)), dim3(256), 0, ctx->stream, ctx->mp,
                       (const double*)ctx->d_phi, ctx->d_efi, ctx->d_efj, ctx->d_ef32);
    SF_HIP(ctx, hipGetLastError());
    return SF_OK;
}

// rho = sum_s q_s nd_s (SolverModule.updateRho)
__global__ void rho_accumulate(int64_t nn, const double* __restrict__ den, double q, double* __restrict__ rho,
                               int first) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x)
        rho[u] = (first ? 0.0 : rho[u]) + den[u] * q;
}

int rho_update(sf_ctx* ctx) {
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    bool first = true;
    for (Species* sp : ctx->species) {
        if (sp->charge == 0) continue;
        hipLaunchKernelGGL(rho_accumulate, dim3(grid_elem(nn)), dim3(256), 0, ctx->stream, nn, (const double*)sp->den,
                           sp->charge, ctx->d_rho, first ? 1 : 0);
        first = false;
    }
    if (first) SF_HIP(ctx, hipMemsetAsync(ctx->d_rho, 0, nn * 8, ctx->stream));
    SF_HIP(ctx, hipGetLastError());
    return SF_OK;
}

int operator_apply(sf_ctx* ctx, const double* x, double* y) {
    std::vector<uint8_t> rt;
    int rc = reference_row_types(ctx, rt);
    if (rc) return rc;
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    uint8_t* d_rt = nullptr;
    double *d_x = nullptr, *d_y = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&d_rt, nn));
    SF_HIP(ctx, hipMalloc((void**)&d_x, nn * 8));
    SF_HIP(ctx, hipMalloc((void**)&d_y, nn * 8));
    SF_HIP(ctx, hipMemcpy(d_rt, rt.data(), nn, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(d_x, x, nn * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(apply_reference_operator, dim3(grid_elem(nn)), dim3(256), 0, ctx->stream, ctx->mp,
                       (const uint8_t*)d_rt, (const double*)d_x, d_y);
    SF_HIP(ctx, hipGetLastError());
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(y, d_y, nn * 8, hipMemcpyDeviceToHost));
    hipFree(d_rt);
    hipFree(d_x);
    hipFree(d_y);
    return SF_OK;
}

}  // namespace sf

using namespace sf;

extern "C" {

int sf_poisson_setup(sf_ctx* ctx, const sf_poisson_desc* desc) {
    SF_CHECK_CTX(ctx);
    if (!desc) return sf_fail(ctx, SF_ERR_ARG, "sf_poisson_setup: null desc");
    if (!desc->linear && !(desc->Te0 > 0)) return sf_fail(ctx, SF_ERR_ARG, "Te0 must be > 0 in nonlinear mode");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return poisson_setup(ctx, desc);
}

int sf_rho_update(sf_ctx* ctx) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return rho_update(ctx);
}

int sf_rho_set(sf_ctx* ctx, const double* rho) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !rho) return sf_fail(ctx, SF_ERR_ARG, "sf_rho_set: bad state or argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(ctx->d_rho, rho, (size_t)ctx->mp.ni * ctx->mp.nj * 8, hipMemcpyHostToDevice));
    return SF_OK;
}

int sf_rho_get(sf_ctx* ctx, double* rho) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !rho) return sf_fail(ctx, SF_ERR_ARG, "sf_rho_get: bad state or argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(rho, ctx->d_rho, (size_t)ctx->mp.ni * ctx->mp.nj * 8, hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_poisson_solve(sf_ctx* ctx, double* phi_io, sf_solve_info* info) {
    SF_CHECK_CTX(ctx);
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    if (phi_io) {
        SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        SF_HIP(ctx, hipMemcpy(ctx->d_phi, phi_io, nn * 8, hipMemcpyHostToDevice));
    }
    int rc = poisson_solve(ctx, info);
    if (rc < 0) return rc;
    if (phi_io) SF_HIP(ctx, hipMemcpy(phi_io, ctx->d_phi, nn * 8, hipMemcpyDeviceToHost));
    return rc;
}

int sf_phi_set(sf_ctx* ctx, const double* phi) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !phi) return sf_fail(ctx, SF_ERR_ARG, "sf_phi_set: bad state or argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(ctx->d_phi, phi, (size_t)ctx->mp.ni * ctx->mp.nj * 8, hipMemcpyHostToDevice));
    return SF_OK;
}

int sf_phi_get(sf_ctx* ctx, double* phi) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !phi) return sf_fail(ctx, SF_ERR_ARG, "sf_phi_get: bad state or argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(phi, ctx->d_phi, (size_t)ctx->mp.ni * ctx->mp.nj * 8, hipMemcpyDeviceToHost));
    return SF_OK;
}

// BoltzmannElectronsMaterial.electronModelBoltzmann (BoltzmannElectronsMaterial.java:179-215):
// n_e = den0 exp((phi - phi0) / kTe0) on every node, floored at 1 m^-3
__global__ void boltzmann_density_kernel(int64_t nn, const double* __restrict__ phi, double den0, double inv_kte,
                                         double phi0, double* __restrict__ den) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        double v = den0 * exp((phi[u] - phi0) * inv_kte);
        den[u] = v < 1.0 ? 1.0 : v;
    }
}

int sf_boltzmann_density(sf_ctx* ctx, double den0, double kTe0, double phi0, double* den) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !den) return sf_fail(ctx, SF_ERR_ARG, "sf_boltzmann_density: bad state or argument");
    // values not given (<= 0) come from the potential solver, as the reference does
    if (den0 <= 0 || kTe0 <= 0) {
        if (!ctx->poisson)
            return sf_fail(ctx, SF_ERR_STATE, "sf_boltzmann_density: den0/kTe0 not given and no solver set up");
        den0 = ctx->poisson->desc.n0;
        kTe0 = ctx->poisson->desc.Te0;
        phi0 = ctx->poisson->desc.phi0;
    }
    if (!(kTe0 > 0)) return sf_fail(ctx, SF_ERR_ARG, "sf_boltzmann_density: non zero temperature kTe0 is needed");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    double* d_den = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&d_den, nn * 8));
    hipLaunchKernelGGL(boltzmann_density_kernel, dim3(grid_elem(nn)), dim3(256), 0, ctx->stream, nn,
                       (const double*)ctx->d_phi, den0, 1.0 / kTe0, phi0, d_den);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(den, d_den, nn * 8, hipMemcpyDeviceToHost);
    hipFree(d_den);
    if (e != hipSuccess) return sf_fail(ctx, SF_ERR_HIP, "sf_boltzmann_density: %s", hipGetErrorString(e));
    return SF_OK;
}

int sf_efield_update(sf_ctx* ctx) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return efield_update(ctx);
}

int sf_efield_set(sf_ctx* ctx, const double* efi, const double* efj) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !efi || !efj) return sf_fail(ctx, SF_ERR_ARG, "sf_efield_set: bad state or argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    // a non-finite E would turn particle states into NaN, which no fixed-point deposit can hold
    for (size_t u = 0; u < nn; u++)
        if (!std::isfinite(efi[u]) || !std::isfinite(efj[u]))
            return sf_fail(ctx, SF_ERR_ARG, "sf_efield_set: non-finite value at node %zu", u);
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    SF_HIP(ctx, hipMemcpy(ctx->d_efi, efi, nn * 8, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(ctx->d_efj, efj, nn * 8, hipMemcpyHostToDevice));
    return ef32_refresh(ctx);
}

int sf_efield_get(sf_ctx* ctx, double* efi, double* efj) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (efi) SF_HIP(ctx, hipMemcpy(efi, ctx->d_efi, nn * 8, hipMemcpyDeviceToHost));
    if (efj) SF_HIP(ctx, hipMemcpy(efj, ctx->d_efj, nn * 8, hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_operator_apply(sf_ctx* ctx, const double* x, double* y) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready || !x || !y) return sf_fail(ctx, SF_ERR_ARG, "sf_operator_apply: bad state or argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return operator_apply(ctx, x, y);
}

}  // extern "C"

namespace sf {
// node classification of the reduced system, read by the learned-guess network (sf_cnn.hip)
const int32_t* poisson_src(sf_ctx* ctx) { return ctx->poisson ? ctx->poisson->src : nullptr; }
const uint8_t* poisson_cls(sf_ctx* ctx) { return ctx->poisson ? ctx->poisson->cls : nullptr; }
const double* poisson_fixval(sf_ctx* ctx) { return ctx->poisson ? ctx->poisson->fixval : nullptr; }
}  // namespace sf
namespace sf {
const sf_poisson_desc* poisson_desc(sf_ctx* ctx) { return ctx->poisson ? &ctx->poisson->desc : nullptr; }
}  // namespace sf

namespace sf {
// Relative nonlinear residual ||F(phi)|| / ||F(0)|| of the context's current potential, with the
// norms of the solve's stopping test.  Overwrites the solver's residual work vectors only.
int poisson_rel_residual(sf_ctx* ctx, double* rel) {
    PoissonState* P = ctx->poisson;
    if (!P) return sf_fail(ctx, SF_ERR_STATE, "poisson solver not set up");
    const MeshP& m = ctx->mp;
    const int64_t nn = (int64_t)m.ni * m.nj;
    const sf_poisson_desc& D = P->desc;
    hipStream_t st = ctx->stream;
    const int ge = P->nblk_elem;
    const double inv_eps = 1.0 / P->eps;
    const double Cn0 = D.linear ? 0.0 : P->C * D.n0;
    const double inv_kTe = D.linear ? 0.0 : 1.0 / D.Te0;
    hipLaunchKernelGGL(ref_norm_kernel, dim3(ge), dim3(256), 0, st, m, (const uint8_t*)P->cls, (const double*)P->W,
                       (const double*)P->bfix, (const double*)ctx->d_rho, inv_eps, Cn0, D.phi0, inv_kTe, D.linear,
                       P->partial, (const double*)ctx->d_phi, P->x);
    finalize<1>(P, st, P->partial, ge, S_R0);
    hipLaunchKernelGGL(newton_residual, dim3(ge), dim3(256), 0, st, m, (const uint8_t*)P->cls, (const double*)P->W,
                       (const double*)P->bfix, (const double*)ctx->d_rho, (const double*)P->x, P->r, P->D, P->P,
                       P->d, P->res, inv_eps, Cn0, D.phi0, inv_kTe, D.linear, P->partial, nullptr);
    finalize<1>(P, st, P->partial, ge, S_NORM2);
    int rc = read_scalars(ctx, P);
    if (rc) return rc;
    double ref = std::sqrt(P->host_scal[S_R0]);
    if (!(ref > 0)) ref = 1.0;
    *rel = std::sqrt(P->host_scal[S_NORM2]) / ref;
    return SF_OK;
}
}  // namespace sf
