// DSMC collisions on the device (collisions/DSMC.java:167-430): Bird's no-time-counter scheme
// per cell after the counting sort, with Boyd's weighting for two species, a constant or VHS
// (SigmaPlus.SigmaBird463) cross-section and the elastic VHS/VSS scattering of ModelElastic.
// One wave per cell.  The cell's particle list is the cell range of the sorted species, ordered
// by particle id in LDS, so that the pair choice (an index into that list) and the cell's
// Philox stream give the same collisions on every run.  The pairs of a cell are processed in
// sequence, as in the reference, by the wave's first lane.
#include <cmath>
#include <vector>

#include "sf_particle_math.h"
#include "sf_rng.h"

namespace sf {

constexpr int DSMC_MAXP = 2048;   // particles of one species in one cell held in LDS
// Cells whose particle ranges are at most DSMC_SMALL long go through a copy of the kernel with
// DSMC_SMALL-entry LDS lists (8 KB per one-wave workgroup instead of 32 KB), so that a CU holds
// four times as many of the latency-bound cell loops; the rest through the DSMC_MAXP copy.
constexpr int DSMC_SMALL = 512;

struct DsmcParams {
    int same, sigma_type;
    double c0, sigma0, ref_temp, vti, mr, gamma, vss_inv;
    double mass_a, mass_b, spwt_a, spwt_b, delta_t;
    uint32_t k0, k1, step, tag;
};

__device__ __forceinline__ double dsmc_sigma(const DsmcParams& pp, double g) {
    if (pp.sigma_type == 0) return pp.c0;
    return pp.sigma0 * pow(2. * K_B * pp.ref_temp / (pp.mr * g * g), pp.vti - 0.5) / pp.gamma;   // Bird eq. 4.63
}

// Particles of the cell that the reference bins into it: (int)lc inside the cells (a particle on
// the last node line has offset 1 in the last cell and is skipped, DSMC.java:206-214).
__device__ __forceinline__ bool dsmc_in_cell(const MeshP& m, const PState& s, int64_t p) {
    const int32_t c = s.cell[p];
    if (c < 0) return false;
    return cell_i(c) + (double)s.fx[p] < m.ni - 1 && cell_j(c) + (double)s.fy[p] < m.nj - 1;
}

// Number of binned particles in the cell range [a, b): the range length, except in the last
// cell row and column where particles may sit on the last node line.
__device__ int dsmc_count(const MeshP& m, const PState& s, int a, int b, bool edge) {
    if (!edge) return b - a;
    int n = 0;
    for (int p = a + (int)(threadIdx.x & 63); p < b; p += 64) n += dsmc_in_cell(m, s, p) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
    return n;
}

// Loads (id << 32 | slot) of the binned particles of [a, b) into key[0..n) and sorts key[0..pow2)
// by id (bitonic, padded with all-ones).  One wave.
__device__ void dsmc_load_sorted(const MeshP& m, const PState& s, int a, int b, unsigned long long* key, int& n) {
    const int lane = threadIdx.x & 63;
    int base = 0;
    for (int p0 = a; p0 < b; p0 += 64) {
        const int p = p0 + lane;
        const bool in = p < b && dsmc_in_cell(m, s, p);
        const unsigned long long mask = __ballot(in);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (in) key[base + before] = ((unsigned long long)(uint32_t)s.id[p] << 32) | (uint32_t)p;
        base += __popcll(mask);
    }
    n = base;
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int k = n + lane; k < np2; k += 64) key[k] = ~0ull;
    __syncthreads();
    for (int size = 2; size <= np2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = lane; k < np2; k += 64) {
                const int partner = k ^ stride;
                if (partner > k) {
                    const bool up = (k & size) == 0;
                    const unsigned long long x = key[k], y = key[partner];
                    if ((x > y) == up) {
                        key[k] = y;
                        key[partner] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// (held to 128 registers, four waves per SIMD: the small copy's cell loops then fill a CU four
// deep at the cost of 172 bytes of spills, 6.4 -> 5.5 ms per pass on 5e7 particles)
template <int CAP>
__global__ void __launch_bounds__(64, 4) dsmc_kernel(MeshP m, PState sa, const int32_t* __restrict__ off_a, PState sb,
                                                  const int32_t* __restrict__ off_b, DsmcParams pp,
                                                  double* __restrict__ sig_cr_max, double* __restrict__ rem,
                                                  double* __restrict__ count_sum, double* __restrict__ real_sum,
                                                  unsigned long long* __restrict__ ncol, int* __restrict__ overflow) {
    __shared__ unsigned long long key_a[CAP], key_b[CAP];
    const int lane = threadIdx.x;
    const int ncells = m.nci * m.ncj;
    for (int c = blockIdx.x; c < ncells; c += gridDim.x) {
        const int ci = c % m.nci, cj = c / m.nci;
        const bool edge = ci == m.nci - 1 || cj == m.ncj - 1;
        const int a0 = off_a[c], a1 = off_a[c + 1];
        int b0 = a0, b1 = a1;
        if (!pp.same) {
            b0 = off_b[c];
            b1 = off_b[c + 1];
        }
        // each cell goes through exactly one of the two copies, by its range lengths
        const bool small = a1 - a0 <= DSMC_SMALL && b1 - b0 <= DSMC_SMALL;
        if (small != (CAP == DSMC_SMALL)) continue;
        const double np1 = dsmc_count(m, sa, a0, a1, edge);
        double np2 = np1;
        if (!pp.same) np2 = dsmc_count(m, sb, b0, b1, edge);
        // collideCell (DSMC.java:238-350): NTC pair count with Boyd's weights (eq. 5); evaluated
        // without fused multiply-adds, so that its integer part is the reference's
        const double scm0 = sig_cr_max[c];
        double nsel_f;
        {
#pragma clang fp contract(off)
            const double cell_volume = node_vol_at(m, ci + 0.5, cj + 0.5);
            double Pab = pp.spwt_b / pp.spwt_a, Pba = pp.spwt_a / pp.spwt_b;
            if (Pab > 1) Pab = 1;
            if (Pba > 1) Pba = 1;
            nsel_f = ((np1 * pp.spwt_a / cell_volume) * np2 * pp.delta_t * scm0) / (Pab + (pp.spwt_b / pp.spwt_a) * Pba);
            if (!pp.same) nsel_f *= 2.0;
            nsel_f += rem[c];
        }
        int nsel = (int)nsel_f;
        if ((pp.same && (np1 < 2 || np2 < 2)) || np1 < 1 || np2 < 1) continue;
        double new_rem = nsel_f - nsel;   // an exact subtraction
        if (nsel > np1 * np2) {
            nsel = (int)(np1 * np2);
            new_rem = 0;
        }
        if (nsel == 0) {
            if (lane == 0) rem[c] = new_rem;
            continue;
        }
        if (np1 > CAP || np2 > CAP) {
            if (lane == 0) atomicMax(overflow, (int)fmax(np1, np2));
            continue;
        }
        int n1 = 0, n2 = 0;
        dsmc_load_sorted(m, sa, a0, a1, key_a, n1);
        unsigned long long* kb = key_a;
        if (!pp.same) {
            dsmc_load_sorted(m, sb, b0, b1, key_b, n2);
            kb = key_b;
        }
        if (lane == 0) {
            ParticleRng rng{pp.k0, pp.k1, (uint32_t)c, pp.step, pp.tag};
            uint32_t k = 0;
            double scm = 0, sums0 = 0, sums1 = 0;
            const double rm1 = pp.mass_a / (pp.mass_a + pp.mass_b), rm2 = pp.mass_b / (pp.mass_a + pp.mass_b);
            for (int q = 0; q < nsel; q++) {
                const int p1 = (int)(rng.draw(k++) * np1);
                int p2;
                do {
                    p2 = (int)(rng.draw(k++) * np2);
                } while (pp.same && p1 == p2);
                const int64_t ga = (int64_t)(uint32_t)key_a[p1];
                const int64_t gb = (int64_t)(uint32_t)kb[p2];
                const PState& Sb = pp.same ? sa : sb;
                double va[3] = {sa.vx[ga], sa.vy[ga], sa.vz[ga]};
                double vb[3] = {Sb.vx[gb], Sb.vy[gb], Sb.vz[gb]};
                const double g0 = va[0] - vb[0], g1 = va[1] - vb[1], g2 = va[2] - vb[2];
                const double cr = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
                const double sigma_cr = cr * dsmc_sigma(pp, cr);
                if (sigma_cr > scm) scm = sigma_cr;
                const double P = sigma_cr / scm0;
                if (rng.draw(k++) < P) {
                    // ModelElastic (DSMC.java:360-430), Bird's DSMC0 scattering
                    double vc[3], vr[3];
                    for (int d = 0; d < 3; d++) vc[d] = rm1 * va[d] + rm2 * vb[d];
                    if (fabs(pp.vss_inv - 1.) < 1.e-4) {
                        const double B = 2. * rng.draw(k++) - 1.0;
                        const double A = sqrt(1. - B * B);
                        vr[0] = B * cr;
                        const double C = 2. * M_PI * rng.draw(k++);
                        vr[1] = A * cos(C) * cr;
                        vr[2] = A * sin(C) * cr;
                    } else {
                        const double B = 2. * (pow(rng.draw(k++), pp.vss_inv)) - 1.;
                        const double A = sqrt(1. - B * B);
                        const double C = 2. * M_PI * rng.draw(k++);
                        const double OC = cos(C), SC = sin(C);
                        const double D = sqrt(g1 * g1 + g2 * g2);
                        if (D > 1.e-6) {
                            vr[0] = B * g0 + A * SC * D;
                            vr[1] = B * g1 + A * (cr * g2 * OC - g0 * g1 * SC) / D;
                            vr[2] = B * g2 - A * (cr * g1 * OC + g0 * g2 * SC) / D;
                        } else {
                            vr[0] = B * g1;
                            vr[1] = A * OC * g1;
                            vr[2] = A * SC * g1;
                        }
                    }
                    bool wa = true, wb = true;
                    if (pp.spwt_a != pp.spwt_b) {   // Boyd's probabilities for unequal weights
                        wa = rng.draw(k++) < pp.spwt_b / pp.spwt_a;
                        wb = rng.draw(k++) < pp.spwt_a / pp.spwt_b;
                    }
                    if (wa) {
                        sa.vx[ga] = (float)(vc[0] + vr[0] * rm2);
                        sa.vy[ga] = (float)(vc[1] + vr[1] * rm2);
                        sa.vz[ga] = (float)(vc[2] + vr[2] * rm2);
                    }
                    if (wb) {
                        Sb.vx[gb] = (float)(vc[0] - vr[0] * rm1);
                        Sb.vy[gb] = (float)(vc[1] - vr[1] * rm1);
                        Sb.vz[gb] = (float)(vc[2] - vr[2] * rm1);
                    }
                    sums0 += 1.0;
                    sums1 += 0.5 * (pp.spwt_a + pp.spwt_b);
                }
            }
            rem[c] = new_rem;
            if (scm > 0) sig_cr_max[c] = scm;
            const size_t u = (size_t)cj * m.ni + ci;
            count_sum[u] += sums0;
            real_sum[u] += sums1;
            if (sums0 > 0) atomicAdd(ncol, (unsigned long long)sums0);
        }
        __syncthreads();   // the LDS lists are reused by the next cell
    }
}

__global__ void dsmc_fields_kernel(int64_t nn, int which, const double* __restrict__ count_sum,
                                   const double* __restrict__ real_sum, double inv_samples, double inv_time,
                                   double* __restrict__ out) {
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < nn; u += (int64_t)gridDim.x * blockDim.x) {
        double v;
        switch (which) {
            case 0: v = count_sum[u]; break;                // col-count-sum
            case 1: v = real_sum[u]; break;                 // col-real-sum
            case 2: v = count_sum[u] * inv_samples; break;  // col-count
            default: v = real_sum[u] * inv_time; break;     // nu
        }
        out[u] = v;
    }
}

// Utils.gamma (core/common/Utils.java:218-236): Bird's approximation of the Gamma function
static double bird_gamma(double X) {
    double A = 1., Y = X;
    if (Y < 1.0) {
        A = A / Y;
    } else {
        do {
            Y = Y - 1;
            if (Y >= 1.) A = A * Y;
        } while (Y >= 1.);
    }
    return A * (1. - 0.5748646 * Y + 0.9512363 * Y * Y - 0.6998588 * Y * Y * Y + 0.4245549 * Y * Y * Y * Y -
                0.1010678 * Y * Y * Y * Y * Y);
}

}  // namespace sf

using namespace sf;

struct sf_dsmc {
    sf_ctx* ctx = nullptr;
    sf_dsmc_desc desc{};
    DsmcParams pp{};
    double* d_scm = nullptr;      // [ncells] sig_cr_max
    double* d_rem = nullptr;      // [ncells]
    double* d_count = nullptr;    // [nn] col-count-sum
    double* d_real = nullptr;     // [nn] col-real-sum
    unsigned long long* d_ncol = nullptr;
    int* d_overflow = nullptr;
    int num_samples = 0;
};

extern "C" {

int sf_dsmc_create(sf_ctx* ctx, const sf_dsmc_desc* desc, sf_dsmc** out) {
    SF_CHECK_CTX(ctx);
    if (!desc || !out) return sf_fail(ctx, SF_ERR_ARG, "sf_dsmc_create: null argument");
    *out = nullptr;
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_dsmc_create: mesh not set up");
    const int ns = (int)ctx->species.size();
    if (desc->species_a < 0 || desc->species_a >= ns || desc->species_b < 0 || desc->species_b >= ns)
        return sf_fail(ctx, SF_ERR_ARG, "sf_dsmc_create: bad species pair %d, %d", desc->species_a, desc->species_b);
    Species* A = ctx->species[desc->species_a];
    Species* B = ctx->species[desc->species_b];
    if (!A->desc.track_ids || !B->desc.track_ids)
        return sf_fail(ctx, SF_ERR_STATE, "sf_dsmc_create: both species must track particle ids");
    if (desc->sigma_type != SF_SIGMA_CONST && desc->sigma_type != SF_SIGMA_BIRD463)
        return sf_fail(ctx, SF_ERR_UNSUPPORTED, "sf_dsmc_create: sigma type %d is not supported", desc->sigma_type);
    if (!(desc->vss_alpha_a > 0) || !(desc->vss_alpha_b > 0))
        return sf_fail(ctx, SF_ERR_ARG, "sf_dsmc_create: vss_alpha must be > 0");
    DsmcParams pp{};
    pp.same = desc->species_a == desc->species_b;
    pp.sigma_type = desc->sigma_type;
    pp.c0 = desc->sigma_c0;
    const double d_ref = desc->diam_a + desc->diam_b;
    pp.sigma0 = 0.25 * M_PI * d_ref * d_ref;
    if (pp.sigma_type == SF_SIGMA_BIRD463 && !(pp.sigma0 > 0))
        return sf_fail(ctx, SF_ERR_ARG, "sf_dsmc_create: material diameter <diam> not defined");
    pp.ref_temp = 0.5 * (desc->ref_temp_a + desc->ref_temp_b);
    pp.vti = 0.5 * (desc->visc_temp_index_a + desc->visc_temp_index_b);
    pp.mass_a = A->mass;
    pp.mass_b = B->mass;
    pp.mr = A->mass * B->mass / (A->mass + B->mass);
    pp.gamma = bird_gamma(2.5 - pp.vti);
    pp.vss_inv = 0.5 * (1.0 / desc->vss_alpha_a + 1.0 / desc->vss_alpha_b);
    pp.spwt_a = A->desc.spwt;
    pp.spwt_b = B->desc.spwt;
    pp.k0 = (uint32_t)desc->seed;
    pp.k1 = (uint32_t)(desc->seed >> 32);
    {
        int rc = claim_stream(ctx, desc->seed, RNG_MODULE_DSMC, desc->tag, &pp.tag);
        if (rc) return rc;
    }
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj, nc = (size_t)ctx->mp.nci * ctx->mp.ncj;
    sf_dsmc* d = new sf_dsmc();
    d->ctx = ctx;
    d->desc = *desc;
    if (d->desc.frequency < 1) d->desc.frequency = 1;
    d->pp = pp;
    std::vector<double> scm(nc, desc->sig_cr_max0 > 0 ? desc->sig_cr_max0 : 1e-16);
    hipError_t e = hipMalloc((void**)&d->d_scm, nc * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d->d_rem, nc * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d->d_count, nn * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d->d_real, nn * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d->d_ncol, 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d->d_overflow, 4);
    if (e == hipSuccess) e = hipMemcpy(d->d_scm, scm.data(), nc * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d->d_rem, 0, nc * 8);
    if (e == hipSuccess) e = hipMemset(d->d_count, 0, nn * 8);
    if (e == hipSuccess) e = hipMemset(d->d_real, 0, nn * 8);
    if (e != hipSuccess) {
        hipFree(d->d_scm);
        hipFree(d->d_rem);
        hipFree(d->d_count);
        hipFree(d->d_real);
        hipFree(d->d_ncol);
        hipFree(d->d_overflow);
        delete d;
        release_stream(ctx, desc->seed, pp.tag);
        return sf_fail(ctx, SF_ERR_HIP, "sf_dsmc_create: %s", hipGetErrorString(e));
    }
    *out = d;
    return SF_OK;
}

int sf_dsmc_destroy(sf_dsmc* d) {
    if (!d) return SF_ERR_ARG;
    hipSetDevice(d->ctx->device);
    hipStreamSynchronize(d->ctx->stream);
    hipFree(d->d_scm);
    hipFree(d->d_rem);
    hipFree(d->d_count);
    hipFree(d->d_real);
    hipFree(d->d_ncol);
    hipFree(d->d_overflow);
    release_stream(d->ctx, d->desc.seed, d->pp.tag);
    delete d;
    return SF_OK;
}

int sf_dsmc_perform(sf_dsmc* d, int it, int64_t* collisions) {
    if (!d) return SF_ERR_ARG;
    sf_ctx* ctx = d->ctx;
    if (collisions) *collisions = 0;
    if (it % d->desc.frequency != 0) return SF_OK;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    d->num_samples++;
    Species* A = ctx->species[d->desc.species_a];
    Species* B = ctx->species[d->desc.species_b];
    // the cell lists are the cell ranges of freshly sorted, compacted species
    for (Species* sp : {A, B}) {
        if (sp->since_sort != 0 || sp->nsorted != sp->np) {
            int rc = sort_species(ctx, sp);
            if (rc) return rc;
        }
    }
    DsmcParams pp = d->pp;
    pp.delta_t = d->desc.frequency * ctx->dt;
    pp.step = (uint32_t)it;
    SF_HIP(ctx, hipMemsetAsync(d->d_ncol, 0, 8, ctx->stream));
    SF_HIP(ctx, hipMemsetAsync(d->d_overflow, 0, 4, ctx->stream));
    const int ncells = ctx->mp.nci * ctx->mp.ncj;
    const int grid = ncells < 65536 ? ncells : 65536;
    hipLaunchKernelGGL(dsmc_kernel<DSMC_SMALL>, dim3(grid), dim3(64), 0, ctx->stream, ctx->mp, A->buf[A->cur],
                       (const int32_t*)A->offsets, B->buf[B->cur], (const int32_t*)B->offsets, pp, d->d_scm, d->d_rem,
                       d->d_count, d->d_real, d->d_ncol, d->d_overflow);
    hipLaunchKernelGGL(dsmc_kernel<DSMC_MAXP>, dim3(grid), dim3(64), 0, ctx->stream, ctx->mp, A->buf[A->cur],
                       (const int32_t*)A->offsets, B->buf[B->cur], (const int32_t*)B->offsets, pp, d->d_scm, d->d_rem,
                       d->d_count, d->d_real, d->d_ncol, d->d_overflow);
    SF_HIP(ctx, hipGetLastError());
    unsigned long long n = 0;
    int over = 0;
    SF_HIP(ctx, hipMemcpyAsync(&n, d->d_ncol, 8, hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipMemcpyAsync(&over, d->d_overflow, 4, hipMemcpyDeviceToHost, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (collisions) *collisions = (int64_t)n;
    if (over > 0)
        return sf_fail(ctx, SF_ERR_UNSUPPORTED, "sf_dsmc_perform: a cell holds %d particles of one species (limit %d);"
                       " its collisions were skipped", over, DSMC_MAXP);
    return SF_OK;
}

int sf_dsmc_field(sf_dsmc* d, int which, double* out) {
    if (!d) return SF_ERR_ARG;
    sf_ctx* ctx = d->ctx;
    if (!out || which < 0 || which > 4) return sf_fail(ctx, SF_ERR_ARG, "sf_dsmc_field: bad argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (which == 4) {   // sig_cr_max per cell, index j*(ni-1)+i
        const size_t nc = (size_t)ctx->mp.nci * ctx->mp.ncj;
        SF_HIP(ctx, hipMemcpy(out, d->d_scm, nc * 8, hipMemcpyDeviceToHost));
        return SF_OK;
    }
    const int64_t nn = (int64_t)ctx->mp.ni * ctx->mp.nj;
    double* d_out = nullptr;
    SF_HIP(ctx, hipMalloc((void**)&d_out, nn * 8));
    const double inv_samples = d->num_samples > 0 ? 1.0 / d->num_samples : 0.0;
    // fc_nu = real_sum / (num_samples * frequency * dt) (DSMC.java:186-188)
    const double inv_time = 1.0 / (d->num_samples * d->desc.frequency * ctx->dt);
    hipLaunchKernelGGL(dsmc_fields_kernel, dim3((unsigned)((nn + 255) / 256 < 8192 ? (nn + 255) / 256 : 8192)),
                       dim3(256), 0, ctx->stream, nn, which, (const double*)d->d_count, (const double*)d->d_real,
                       inv_samples, inv_time, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, nn * 8, hipMemcpyDeviceToHost);
    hipFree(d_out);
    if (e != hipSuccess) return sf_fail(ctx, SF_ERR_HIP, "sf_dsmc_field: %s", hipGetErrorString(e));
    return SF_OK;
}

int sf_dsmc_clear_samples(sf_dsmc* d) {
    if (!d) return SF_ERR_ARG;
    sf_ctx* ctx = d->ctx;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    SF_HIP(ctx, hipMemsetAsync(d->d_count, 0, nn * 8, ctx->stream));
    SF_HIP(ctx, hipMemsetAsync(d->d_real, 0, nn * 8, ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    d->num_samples = 0;
    return SF_OK;
}

}  // extern "C"
