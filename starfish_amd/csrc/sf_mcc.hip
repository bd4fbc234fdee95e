// Monte Carlo collisions of a kinetic species with a background material on the device
// (collisions/MCC.java:167-297): constant cross-section (Sigma.SigmaConst), charge exchange
// (ModelCEX) or elastic (ModelMEX) scattering.  One pass visits every particle slot once; the
// random draws come from Philox4x32-10 keyed by the seed and indexed by (particle id, step), so a
// particle's fate does not depend on the order or the thread it is processed in.
#include <cmath>
#include <cstring>
#include <vector>

#include "sf_particle_math.h"
#include "sf_rng.h"

namespace sf {

// Field2D.gather with the gather_safe fallback (Field2D.java:300-390) on an FP64 node field.
__device__ __forceinline__ double gather_node(const MeshP& m, const double* __restrict__ f, double fi, double fj) {
    int i = (int)fi, j = (int)fj;
    double di = fi - i, dj = fj - j;
    auto at = [&](int a, int b) { return f[(size_t)b * m.ni + a]; };
    if (i < 0 || j < 0 || i + 1 >= m.ni || j + 1 >= m.nj) {
        if (i < 0) { i = 0; di = 0; }
        if (j < 0) { j = 0; dj = 0; }
        if (i >= m.ni - 1) { i = m.ni - 1; di = 0; }
        if (j >= m.nj - 1) { j = m.nj - 1; dj = 0; }
        double v = (1 - di) * (1 - dj) * at(i, j);
        if (di > 0) v += di * (1 - dj) * at(i + 1, j);
        if (di > 0 && dj > 0) v += di * dj * at(i + 1, j + 1);
        if (dj > 0) v += (1 - di) * dj * at(i, j + 1);
        return v;
    }
    double v = (1 - di) * (1 - dj) * at(i, j);
    v += di * (1 - dj) * at(i + 1, j);
    v += di * dj * at(i + 1, j + 1);
    v += (1 - di) * dj * at(i, j + 1);
    return v;
}

struct MccTarget {
    const double *den, *u, *v, *w, *T;
};
struct MccParams {
    double sigma, dt, source_mass, target_mass, max_T, mpw;
    int model;   // 0 CEX, 1 MEX
    uint32_t k0, k1, step, tag;
};

__global__ void __launch_bounds__(256) mcc_kernel(MeshP m, PState s, int64_t np, MccTarget tg, MccParams pp,
                                                  unsigned long long* __restrict__ count, double* __restrict__ S,
                                                  unsigned long long* __restrict__ ncol) {
    unsigned long long ncol_local = 0;   // added to ncol once per wave
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = s.cell[p];
        if (c < 0) continue;   // hole left by a dead particle
        const double lc0 = cell_i(c) + (double)s.fx[p], lc1 = cell_j(c) + (double)s.fy[p];
        const double den_a = gather_node(m, tg.den, lc0, lc1);
        if (den_a <= 0) continue;
        // Material.sampleVelocity: the target's stream velocity at the particle
        const double t0 = gather_node(m, tg.u, lc0, lc1), t1 = gather_node(m, tg.v, lc0, lc1),
                     t2 = gather_node(m, tg.w, lc0, lc1);
        double v0 = s.vx[p], v1 = s.vy[p], v2 = s.vz[p];
        const double g0 = t0 - v0, g1 = t1 - v1, g2 = t2 - v2;
        const double g = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        const double P = 1 - exp(-pp.sigma * g * pp.dt * den_a);
        ParticleRng rng{pp.k0, pp.k1, (uint32_t)s.id[p], pp.step, pp.tag};
        if (P < rng.draw(0)) continue;   // no collision
        // Material.sampleMaxwellianVelocity(mesh, lc, 0, max_T) with Utils.SampleMaxw3D
        double T = gather_node(m, tg.T, lc0, lc1);
        if (T < 0) T = 0;
        if (pp.max_T > 0 && T > pp.max_T) T = pp.max_T;
        const double vth = sqrt(2 * K_B * T / pp.target_mass);
        const double sh = sqrt(0.5);
        // draws 2..13, 14..25 and 26..37: one component each
        const double x0 = t0 + (vth * 1.2) * (sh * 1 * ((rng.sum12(2) - 0.5 * 12) * 1.0));
        const double x1 = t1 + (vth * 1.2) * (sh * 1 * ((rng.sum12(14) - 0.5 * 12) * 1.0));
        const double x2 = t2 + (vth * 1.2) * (sh * 1 * ((rng.sum12(26) - 0.5 * 12) * 1.0));
        const double E1 = 0.5 * pp.source_mass * sqrt(v0 * v0 + v1 * v1 + v2 * v2);
        if (pp.model == 0) {   // ModelCEX: the ion takes the neutral's velocity
            v0 = x0;
            v1 = x1;
            v2 = x2;
        } else {               // ModelMEX: isotropic scattering in the centre-of-mass frame
            const double rm1 = pp.source_mass / (pp.source_mass + pp.target_mass);
            const double rm2 = pp.target_mass / (pp.source_mass + pp.target_mass);
            const double h0 = v0 - x0, h1 = v1 - x1, h2 = v2 - x2;
            const double gm = sqrt(h0 * h0 + h1 * h1 + h2 * h2);
            const double c0 = rm1 * v0 + rm2 * x0, c1 = rm1 * v1 + rm2 * x1, c2 = rm1 * v2 + rm2 * x2;
            const double B = 2. * rng.draw(38) - 1.0;
            const double A = sqrt(1. - B * B);
            const double C = 2. * M_PI * rng.draw(39);
            v0 = c0 + (B * gm) * rm2;
            v1 = c1 + (A * cos(C) * gm) * rm2;
            v2 = c2 + (A * sin(C) * gm) * rm2;
        }
        const float f0 = (float)v0, f1 = (float)v1, f2 = (float)v2;
        s.vx[p] = f0;
        s.vy[p] = f1;
        s.vz[p] = f2;
        const double E2 = 0.5 * pp.source_mass * sqrt(v0 * v0 + v1 * v1 + v2 * v2);
        const double dS = pp.mpw * (E1 - E2) / (pp.dt * node_vol_at(m, lc0, lc1));
        // target.getS(mesh).scatter(lc, dS) (Field2D.scatter, Field2D.java:244-295)
        const int si = (int)lc0, sj = (int)lc1;
        if (!(si < 0 || sj < 0 || si >= m.ni - 1 || sj >= m.nj - 1)) {
            double w[4];
            scatter_weights(m, si, sj, lc0 - si, lc1 - sj, w);
            atomicAdd(S + (size_t)sj * m.ni + si, w[0] * dS);
            atomicAdd(S + (size_t)sj * m.ni + si + 1, w[1] * dS);
            atomicAdd(S + (size_t)(sj + 1) * m.ni + si + 1, w[2] * dS);
            atomicAdd(S + (size_t)(sj + 1) * m.ni + si, w[3] * dS);
        }
        atomicAdd(count + (size_t)sj * m.ni + si, 1ull);   // col-count-sum at node ((int)lc0, (int)lc1)
        ncol_local++;
    }
    for (int o = 32; o > 0; o >>= 1) ncol_local += __shfl_xor(ncol_local, o);
    if ((threadIdx.x & 63) == 0 && ncol_local) atomicAdd(ncol, ncol_local);
}

__global__ void mcc_fields_kernel(int64_t nn, int which, const unsigned long long* __restrict__ count,
                                  const double* __restrict__ S, const double* __restrict__ node_vol, double mpw,
                                  double inv_samples, double inv_time, double* __restrict__ out) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        const double n = (double)count[u];
        double v;
        switch (which) {
            case 0: v = n; break;                                          // col-count-sum
            case 1: v = n * mpw; break;                                    // col-real-sum
            case 2: v = n * inv_samples; break;                            // col-count
            case 3: v = n * mpw * inv_time / node_vol[u]; break;           // nu
            default: v = S[u]; break;                                      // target S
        }
        out[u] = v;
    }
}

}  // namespace sf

using namespace sf;

struct sf_mcc {
    sf_ctx* ctx = nullptr;
    sf_mcc_desc desc{};
    double target_mass = 0;
    double* d_target = nullptr;          // [5][nn] den, u, v, w, T
    unsigned long long* d_count = nullptr;
    double* d_S = nullptr;
    unsigned long long* d_ncol = nullptr;
    bool have_target = false;
    int num_samples = 0;
    double time_start = 0, time_last = 0;
    uint32_t stream_tag = 0;             // module-tagged counter word of the Philox streams
};

static int grid_n(int64_t n, int max_blocks) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    return (int)(b < max_blocks ? b : max_blocks);
}

extern "C" {

int sf_mcc_create(sf_ctx* ctx, const sf_mcc_desc* desc, sf_mcc** out) {
    SF_CHECK_CTX(ctx);
    if (!desc || !out) return sf_fail(ctx, SF_ERR_ARG, "sf_mcc_create: null argument");
    *out = nullptr;
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_mcc_create: mesh not set up");
    if (desc->source_species < 0 || desc->source_species >= (int)ctx->species.size())
        return sf_fail(ctx, SF_ERR_ARG, "sf_mcc_create: no species %d", desc->source_species);
    if (!ctx->species[desc->source_species]->desc.track_ids)
        return sf_fail(ctx, SF_ERR_STATE, "sf_mcc_create: the source species must track particle ids");
    if (desc->model != SF_MCC_CEX && desc->model != SF_MCC_MEX)
        return sf_fail(ctx, SF_ERR_UNSUPPORTED, "sf_mcc_create: model %d is not supported", desc->model);
    if (!(desc->sigma >= 0) || !(desc->target_molwt > 0))
        return sf_fail(ctx, SF_ERR_ARG, "sf_mcc_create: sigma must be >= 0 and target_molwt > 0");
    uint32_t stream_tag = 0;
    {
        int rc = claim_stream(ctx, desc->seed, RNG_MODULE_MCC, desc->tag, &stream_tag);
        if (rc) return rc;
    }
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    sf_mcc* c = new sf_mcc();
    c->ctx = ctx;
    c->desc = *desc;
    c->stream_tag = stream_tag;
    if (c->desc.frequency < 1) c->desc.frequency = 1;
    c->target_mass = desc->target_molwt * AMU;
    hipError_t e = hipMalloc((void**)&c->d_target, 5 * nn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_count, nn * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_S, nn * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&c->d_ncol, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->d_target, 0, 5 * nn * sizeof(double));
    if (e == hipSuccess) e = hipMemset(c->d_count, 0, nn * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->d_S, 0, nn * sizeof(double));
    if (e != hipSuccess) {
        hipFree(c->d_target);
        hipFree(c->d_count);
        hipFree(c->d_S);
        hipFree(c->d_ncol);
        delete c;
        release_stream(ctx, desc->seed, stream_tag);
        return sf_fail(ctx, SF_ERR_HIP, "sf_mcc_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return SF_OK;
}

int sf_mcc_destroy(sf_mcc* c) {
    if (!c) return SF_ERR_ARG;
    hipSetDevice(c->ctx->device);
    hipStreamSynchronize(c->ctx->stream);
    hipFree(c->d_target);
    hipFree(c->d_count);
    hipFree(c->d_S);
    hipFree(c->d_ncol);
    release_stream(c->ctx, c->desc.seed, c->stream_tag);
    delete c;
    return SF_OK;
}

int sf_mcc_set_target(sf_mcc* c, const double* den, const double* u, const double* v, const double* w,
                      const double* temp) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    if (!den) return sf_fail(ctx, SF_ERR_ARG, "sf_mcc_set_target: the target density is required");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double* src[5] = {den, u, v, w, temp};
    for (int k = 0; k < 5; k++) {
        if (src[k]) SF_HIP(ctx, hipMemcpy(c->d_target + k * nn, src[k], nn * sizeof(double), hipMemcpyHostToDevice));
        else SF_HIP(ctx, hipMemset(c->d_target + k * nn, 0, nn * sizeof(double)));
    }
    c->have_target = true;
    return SF_OK;
}

int sf_mcc_perform(sf_mcc* c, int it, double time, int64_t* collisions) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    if (collisions) *collisions = 0;
    if (!c->have_target) return sf_fail(ctx, SF_ERR_STATE, "sf_mcc_perform: no target fields set");
    if (c->desc.stop_it > 0 && it > c->desc.stop_it) return SF_OK;
    if (it % c->desc.frequency != 0) return SF_OK;
    if (c->num_samples == 0) c->time_start = time;
    c->num_samples++;
    c->time_last = time;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    Species* sp = ctx->species[c->desc.source_species];
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    MccTarget tg{c->d_target, c->d_target + nn, c->d_target + 2 * nn, c->d_target + 3 * nn, c->d_target + 4 * nn};
    MccParams pp;
    pp.sigma = c->desc.sigma;
    pp.dt = c->desc.frequency * ctx->dt;
    pp.source_mass = sp->mass;
    pp.target_mass = c->target_mass;
    pp.max_T = c->desc.max_target_temp;
    pp.mpw = sp->desc.spwt;
    pp.model = c->desc.model;
    pp.k0 = (uint32_t)c->desc.seed;
    pp.k1 = (uint32_t)(c->desc.seed >> 32);
    pp.step = (uint32_t)it;
    pp.tag = c->stream_tag;
    SF_HIP(ctx, hipMemsetAsync(c->d_ncol, 0, sizeof(unsigned long long), ctx->stream));
    if (sp->np > 0) {
        hipLaunchKernelGGL(mcc_kernel, dim3(grid_n(sp->np, 1 << 16)), dim3(256), 0, ctx->stream, ctx->mp,
                           sp->buf[sp->cur], sp->np, tg, pp, c->d_count, c->d_S, c->d_ncol);
        SF_HIP(ctx, hipGetLastError());
    }
    unsigned long long n = 0;
    SF_HIP(ctx, hipMemcpyAsync(&n, c->d_ncol, sizeof(n), hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (collisions) *collisions = (int64_t)n;
    return SF_OK;
}

int sf_mcc_field(sf_mcc* c, int which, double* out) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    if (!out || which < 0 || which > 4) return sf_fail(ctx, SF_ERR_ARG, "sf_mcc_field: bad argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    Species* sp = ctx->species[c->desc.source_species];
    double* d_out = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&d_out, nn * sizeof(double)));
    // the reference divides by the sampling time without a guard (MCC.java:152-153)
    const double inv_samples = c->num_samples > 0 ? 1.0 / c->num_samples : 0.0;
    const double inv_time = 1.0 / (c->time_last - c->time_start);
    hipLaunchKernelGGL(mcc_fields_kernel, dim3(grid_n(nn, 8192)), dim3(256), 0, ctx->stream, nn, which,
                       (const unsigned long long*)c->d_count, (const double*)c->d_S, (const double*)ctx->d_node_vol,
                       sp->desc.spwt, inv_samples, inv_time, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, nn * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d_out);
    if (e != hipSuccess) return sf_fail(ctx, SF_ERR_HIP, "sf_mcc_field: %s", hipGetErrorString(e));
    return SF_OK;
}

int sf_mcc_clear_samples(sf_mcc* c, int clear_source) {
    if (!c) return SF_ERR_ARG;
    sf_ctx* ctx = c->ctx;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    SF_HIP(ctx, hipMemsetAsync(c->d_count, 0, nn * sizeof(unsigned long long), ctx->stream));
    if (clear_source) SF_HIP(ctx, hipMemsetAsync(c->d_S, 0, nn * sizeof(double), ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    c->num_samples = 0;
    return SF_OK;
}

}  // extern "C"
