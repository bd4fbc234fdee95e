// Context, mesh, species and particle I/O entry points of the C-ABI.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "sf_internal.h"

using namespace sf;

int sf_fail(sf_ctx* ctx, int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

namespace {
template <class T>
int dalloc(sf_ctx* ctx, T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return sf_fail(ctx, SF_ERR_OOM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                                        hipGetErrorString(e));
    return SF_OK;
}
template <class T>
void dfree(T*& p) {
    if (p) hipFree((void*)p);
    p = nullptr;
}
#define TRY(x)                 \
    do {                       \
        int rc_ = (x);         \
        if (rc_ < 0) return rc_; \
    } while (0)

Species* get_species(sf_ctx* ctx, int s) {
    if (!ctx || s < 0 || s >= (int)ctx->species.size()) return nullptr;
    return ctx->species[s];
}
}  // namespace

namespace sf {

int claim_stream(sf_ctx* ctx, uint64_t seed, uint32_t module, uint32_t tag, uint32_t* stream_tag) {
    auto taken = [&](uint32_t t) {
        for (auto& s : ctx->rng_streams)
            if (s.first == seed && s.second == (module | t)) return true;
        return false;
    };
    if (tag == SF_TAG_AUTO) {
        tag = 0;
        while (taken(tag)) tag++;
    } else if (tag > RNG_TAG_MASK) {
        return sf_fail(ctx, SF_ERR_ARG, "stream tag %u is not below 2^28", tag);
    } else if (taken(tag)) {
        return sf_fail(ctx, SF_ERR_ARG,
                       "another live object of the same kind draws from seed %llu with tag %u; give it another tag",
                       (unsigned long long)seed, tag);
    }
    ctx->rng_streams.emplace_back(seed, module | tag);
    *stream_tag = module | tag;
    return SF_OK;
}

void release_stream(sf_ctx* ctx, uint64_t seed, uint32_t stream_tag) {
    auto& v = ctx->rng_streams;
    for (size_t k = 0; k < v.size(); k++)
        if (v[k].first == seed && v[k].second == stream_tag) {
            v.erase(v.begin() + k);
            return;
        }
}

}  // namespace sf

extern "C" {

const char* sf_version(void) { return "starfish_hip 0.1 (gfx950)"; }

int sf_ctx_create(int device, sf_ctx** out) {
    if (!out) return SF_ERR_ARG;
    *out = nullptr;
    if (device == -1) {
        // host-only context: geometry, generator and sources; every device entry point fails
        sf_ctx* c = new sf_ctx();
        c->device = -1;
        c->rng.set_seed(0);
        *out = c;
        return SF_OK;
    }
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return SF_ERR_HIP;
    if (device < 0 || device >= ndev) return SF_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return SF_ERR_HIP;
    sf_ctx* c = new sf_ctx();
    c->device = device;
    c->rng.set_seed(0);    // Starfish's shared generator is new Random(0)
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->ncu <= 0)
        c->ncu = 256;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SF_ERR_HIP;
    }
    for (auto& ev : c->ev) hipEventCreate(&ev);
    c->no_fast_push = getenv("SF_NO_FAST_PUSH") != nullptr;
    // the fixed staging array of sf_particles_add_noalloc: its address on this device
    const int rc = noalloc_scratch_init(c);
    if (rc) {
        for (auto& ev : c->ev) hipEventDestroy(ev);
        hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return SF_OK;
}

static void free_species(Species* sp) {
    for (int b = 0; b < 2; b++) {
        dfree(sp->buf[b].cell);
        dfree(sp->buf[b].fx);
        dfree(sp->buf[b].fy);
        dfree(sp->buf[b].vx);
        dfree(sp->buf[b].vy);
        dfree(sp->buf[b].vz);
        dfree(sp->buf[b].id);
    }
    dfree(sp->key);
    dfree(sp->acc);
    dfree(sp->sums);
    dfree(sp->mpc_sum);
    dfree(sp->gas);
    dfree(sp->flags);
    dfree(sp->qcount);
    dfree(sp->tile_nslow);
    dfree(sp->slow_items);
    dfree(sp->bnd_hits);
    dfree(sp->den);
    dfree(sp->counts);
    dfree(sp->offsets);
    dfree(sp->cursor);
    dfree(sp->blocksum);
    dfree(sp->stats);
    delete sp;
}

static void free_mesh(sf_ctx* c) {
    for (auto& src : c->sources) dfree(src.d_geom);
    dfree(c->d_face);
    dfree(c->d_nseg_ptr);
    dfree(c->d_nseg_idx);
    dfree(c->d_seg);
    dfree(c->d_hasseg);
    dfree(c->d_cellslow);
    dfree(c->d_node_vol);
    dfree(c->d_phi);
    dfree(c->d_rho);
    dfree(c->d_efi);
    dfree(c->d_efj);
    dfree(c->d_ef32);
    dfree(c->d_bf32);
}

int sf_ctx_destroy(sf_ctx* ctx) {
    if (!ctx) return SF_ERR_ARG;
    if (ctx->device < 0) {
        delete ctx;
        return SF_OK;
    }
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (Species* sp : ctx->species) free_species(sp);
    if (ctx->poisson) poisson_free(ctx->poisson);
    free_mesh(ctx);
    for (auto& ev : ctx->ev) hipEventDestroy(ev);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return SF_OK;
}

const char* sf_last_error(const sf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sf_ctx_set_dt(sf_ctx* ctx, double dt) {
    SF_CHECK_CTX(ctx);
    if (!(dt > 0) || !std::isfinite(dt)) return sf_fail(ctx, SF_ERR_ARG, "dt must be positive");
    ctx->dt = dt;
    return SF_OK;
}

int sf_ctx_set_rng(sf_ctx* ctx, int64_t seed, int raw_state) {
    SF_CHECK_CTX(ctx);
    if (raw_state) ctx->rng.seed = (uint64_t)seed & ((1ULL << 48) - 1);
    else ctx->rng.set_seed(seed);
    return SF_OK;
}

int sf_ctx_get_rng_state(const sf_ctx* ctx, int64_t* state) {
    if (!ctx || !state) return SF_ERR_ARG;
    *state = (int64_t)ctx->rng.seed;
    return SF_OK;
}

int sf_ctx_synchronize(sf_ctx* ctx) {
    SF_CHECK_CTX(ctx);
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SF_OK;
}

int sf_ctx_set_it(sf_ctx* ctx, int it) {
    SF_CHECK_CTX(ctx);
    if (it < 1) return sf_fail(ctx, SF_ERR_ARG, "sf_ctx_set_it: the step counter starts at 1 (got %d)", it);
    ctx->it = it;
    return SF_OK;
}

int sf_mesh_setup(sf_ctx* ctx, const sf_mesh_desc* mesh, const int* face_type, const double* face_value,
                  const int* node_type, const double* node_bc, const double* node_vol, int nseg,
                  const double* seg_xy, const int* seg_btype, const int* seg_absorb, const int* node_seg_ptr,
                  const int* node_seg_idx) {
    SF_CHECK_CTX(ctx);
    if (!mesh || !face_type || !node_type || !node_vol || !node_seg_ptr)
        return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: null argument");
    if (mesh->ni < 2 || mesh->nj < 2 || mesh->ni > 65536 || mesh->nj > 65536)
        return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: mesh size %d x %d outside [2, 65536]", mesh->ni, mesh->nj);
    if (!(mesh->dh[0] > 0 && mesh->dh[1] > 0)) return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: spacing must be > 0");
    if (mesh->domain != SF_DOMAIN_XY && mesh->domain != SF_DOMAIN_RZ && mesh->domain != SF_DOMAIN_ZR)
        return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: unknown domain type %d", mesh->domain);
    if (nseg > 0 && (!seg_xy || !seg_btype || !seg_absorb || !node_seg_idx))
        return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: segment arrays missing");
    if (ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_mesh_setup: mesh already set");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    const int ni = mesh->ni, nj = mesh->nj;
    const size_t nn = (size_t)ni * nj;
    const int nface = 2 * ni + 2 * nj;
    ctx->mdesc = *mesh;
    MeshP& m = ctx->mp;
    m.ni = ni;
    m.nj = nj;
    m.nci = ni - 1;
    m.ncj = nj - 1;
    m.domain = mesh->domain;
    m.ntx = (m.nci + TILE_W - 1) / TILE_W;
    m.nty = (m.ncj + TILE_H - 1) / TILE_H;
    for (int d = 0; d < 2; d++) {
        m.x0[d] = mesh->x0[d];
        m.dh[d] = mesh->dh[d];
        m.inv_dh[d] = 1.0 / mesh->dh[d];
    }
    ctx->h_face_type.assign(face_type, face_type + nface);
    ctx->h_face_value.assign(nface, 0.0);
    if (face_value) ctx->h_face_value.assign(face_value, face_value + nface);
    ctx->h_node_type.assign(node_type, node_type + nn);
    ctx->h_node_bc.assign(nn, 0.0);
    if (node_bc) ctx->h_node_bc.assign(node_bc, node_bc + nn);
    ctx->h_node_vol.assign(node_vol, node_vol + nn);
    for (size_t u = 0; u < nn; u++)
        if (!(node_vol[u] > 0) || !std::isfinite(node_vol[u]))
            return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: node volume %zu is not positive", u);
    // device copies
    std::vector<int8_t> face8(nface);
    for (int k = 0; k < nface; k++) face8[k] = (int8_t)face_type[k];
    // node segment lists restricted to segments the mover tests (DIRICHLET / SINK)
    std::vector<int32_t> ptr(nn + 1, 0), idx;
    std::vector<uint8_t> has(nn, 0);
    for (size_t u = 0; u < nn; u++) {
        for (int q = node_seg_ptr[u]; q < node_seg_ptr[u + 1]; q++) {
            int s = node_seg_idx[q];
            if (s < 0 || s >= nseg) return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: segment index out of range");
            if (seg_btype[s] == SF_BND_DIRICHLET || seg_btype[s] == SF_BND_SINK) idx.push_back(s);
        }
        ptr[u + 1] = (int32_t)idx.size();
        has[u] = ptr[u + 1] > ptr[u];
    }
    std::vector<DevSeg> segs(nseg);
    for (int s = 0; s < nseg; s++) {
        DevSeg& g = segs[s];
        g.x1[0] = seg_xy[4 * s];
        g.x1[1] = seg_xy[4 * s + 1];
        g.x2[0] = seg_xy[4 * s + 2];
        g.x2[1] = seg_xy[4 * s + 3];
        double dx = g.x2[0] - g.x1[0], dy = g.x2[1] - g.x1[1];
        double len = std::sqrt(dx * dx + dy * dy);
        if (!(len > 0)) return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_setup: zero-length segment %d", s);
        dx /= len;
        dy /= len;
        g.nrm[0] = -dy;            // LinearSegment normal (LinearSegment.java:41-44)
        g.nrm[1] = dx;
        g.btype = seg_btype[s];
        g.absorb = seg_absorb[s];
        g.bnd = -1;
    }
    TRY(dalloc(ctx, &ctx->d_face, nface));
    TRY(dalloc(ctx, &ctx->d_nseg_ptr, nn + 1));
    TRY(dalloc(ctx, &ctx->d_nseg_idx, idx.size()));
    TRY(dalloc(ctx, &ctx->d_seg, (size_t)nseg));
    ctx->nseg = nseg;
    ctx->nbnd = 0;
    TRY(dalloc(ctx, &ctx->d_hasseg, nn));
    // Cells whose particles can take the push's fast path.  The fast path moves a particle by
    // less than 0.99 cell per axis, so from cell (i, j) (lc in [i, i+1)) the new lc lies in
    // (i - 0.99, i + 1.99): ProcessBoundary's bounding box (KineticMaterial.java:482-512) spans
    // nodes [i-1, i+1] x [j-1, j+1], and the particle stays strictly inside the mesh when
    // 1 <= i <= nci-2 and 1 <= j <= ncj-2 (no face is crossed).  Any other cell, or one whose
    // box lists a DIRICHLET/SINK segment, is "slow": its particles run the full reference loop.
    const int nci = ni - 1, ncj = nj - 1;
    std::vector<uint8_t> cslow((size_t)nci * ncj, 1);
    for (int j = 1; j <= ncj - 2; j++)
        for (int i = 1; i <= nci - 2; i++) {
            uint8_t any = 0;
            for (int b = j - 1; b <= j + 1 && !any; b++)
                for (int a = i - 1; a <= i + 1; a++)
                    if (has[(size_t)b * ni + a]) { any = 1; break; }
            cslow[(size_t)j * nci + i] = any;
        }
    TRY(dalloc(ctx, &ctx->d_cellslow, cslow.size()));
    SF_HIP(ctx, hipMemcpy(ctx->d_cellslow, cslow.data(), cslow.size(), hipMemcpyHostToDevice));
    TRY(dalloc(ctx, &ctx->d_node_vol, nn));
    TRY(dalloc(ctx, &ctx->d_phi, nn));
    TRY(dalloc(ctx, &ctx->d_rho, nn));
    TRY(dalloc(ctx, &ctx->d_efi, nn));
    TRY(dalloc(ctx, &ctx->d_efj, nn));
    TRY(dalloc(ctx, &ctx->d_ef32, nn));
    SF_HIP(ctx, hipMemcpy(ctx->d_face, face8.data(), nface, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(ctx->d_nseg_ptr, ptr.data(), (nn + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!idx.empty())
        SF_HIP(ctx, hipMemcpy(ctx->d_nseg_idx, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (nseg > 0) SF_HIP(ctx, hipMemcpy(ctx->d_seg, segs.data(), nseg * sizeof(DevSeg), hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(ctx->d_hasseg, has.data(), nn, hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemcpy(ctx->d_node_vol, node_vol, nn * sizeof(double), hipMemcpyHostToDevice));
    SF_HIP(ctx, hipMemset(ctx->d_phi, 0, nn * sizeof(double)));
    SF_HIP(ctx, hipMemset(ctx->d_rho, 0, nn * sizeof(double)));
    SF_HIP(ctx, hipMemset(ctx->d_efi, 0, nn * sizeof(double)));
    SF_HIP(ctx, hipMemset(ctx->d_efj, 0, nn * sizeof(double)));
    SF_HIP(ctx, hipMemset(ctx->d_ef32, 0, nn * sizeof(float2)));
    SF_HIP(ctx, hipDeviceSynchronize());
    ctx->mesh_ready = true;
    return SF_OK;
}

int sf_species_create(sf_ctx* ctx, const sf_species_desc* desc, int* species) {
    SF_CHECK_CTX(ctx);
    if (!desc || !species) return sf_fail(ctx, SF_ERR_ARG, "sf_species_create: null argument");
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_species_create: mesh not set up");
    if (!(desc->molwt > 0)) return sf_fail(ctx, SF_ERR_ARG, "sf_species_create: molwt must be > 0");
    if (!(desc->spwt > 0)) return sf_fail(ctx, SF_ERR_ARG, "sf_species_create: spwt must be > 0");
    if (desc->capacity < 1 || desc->capacity > (int64_t)2000000000)
        return sf_fail(ctx, SF_ERR_ARG, "sf_species_create: capacity must be in [1, 2e9]");
    if (!(desc->vmax >= 0) || !std::isfinite(desc->vmax))
        return sf_fail(ctx, SF_ERR_ARG, "sf_species_create: vmax must be finite and >= 0");
    if (!std::isfinite(desc->init_T)) return sf_fail(ctx, SF_ERR_ARG, "sf_species_create: init_T must be finite");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    Species* sp = new Species();
    sp->desc = *desc;
    sp->charge = desc->charge * QE;       // Material.java:64-65
    sp->mass = desc->molwt * AMU;
    sp->qm = sp->charge / sp->mass;       // Material.init (Material.java:711)
    sp->capacity = desc->capacity;
    double vmax = desc->vmax > 0 ? desc->vmax : 1e5;
    // fixed-point scale of the mover sums: a particle at vmax contributes ~2^40; the deposit
    // channels take their scales from the particles at each push
    sp->kstat = std::ldexp(1.0, (int)std::floor(std::log2(std::ldexp(1.0, 40) / vmax)));
    sp->kscale[0] = std::ldexp(1.0, 40);
    for (int c = 1; c < NCH; c++) sp->kscale[c] = 1.0;
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj;
    const size_t ncells = (size_t)ctx->mp.nci * ctx->mp.ncj;
    const size_t cap = (size_t)desc->capacity;
    int rc = 0;
    for (int b = 0; b < 2 && !rc; b++) {
        rc = dalloc(ctx, &sp->buf[b].cell, cap);
        if (!rc) rc = dalloc(ctx, &sp->buf[b].fx, cap);
        if (!rc) rc = dalloc(ctx, &sp->buf[b].fy, cap);
        if (!rc) rc = dalloc(ctx, &sp->buf[b].vx, cap);
        if (!rc) rc = dalloc(ctx, &sp->buf[b].vy, cap);
        if (!rc) rc = dalloc(ctx, &sp->buf[b].vz, cap);
        if (!rc && desc->track_ids) rc = dalloc(ctx, &sp->buf[b].id, cap);
    }
    if (!rc) rc = dalloc(ctx, &sp->key, cap);
    if (!rc) rc = dalloc(ctx, &sp->acc, NCH * nn);
    if (!rc) rc = dalloc(ctx, &sp->sums, NCH * nn);
    if (!rc) rc = dalloc(ctx, &sp->mpc_sum, ncells);
    if (!rc) rc = dalloc(ctx, &sp->gas, NGAS * nn);
    if (!rc) rc = dalloc(ctx, &sp->flags, 1);
    const size_t ntiles = (size_t)std::max(1, ctx->mp.ntx * ctx->mp.nty);
    if (!rc) rc = dalloc(ctx, &sp->qcount, 2);
    if (!rc) rc = dalloc(ctx, &sp->tile_nslow, ntiles);
    // a tile of n slow particles makes ceil(n / 256) items: at most cap / 256 + tiles in all
    if (!rc) rc = dalloc(ctx, &sp->slow_items, cap / 256 + ntiles + 1);
    if (!rc && ctx->nbnd > 0) {
        rc = dalloc(ctx, &sp->bnd_hits, (size_t)ctx->nbnd);
        if (!rc) SF_HIP(ctx, hipMemset(sp->bnd_hits, 0, ctx->nbnd * sizeof(unsigned long long)));
    }
    if (!rc) rc = dalloc(ctx, &sp->den, 4 * nn);
    if (!rc) rc = dalloc(ctx, &sp->counts, ncells);
    if (!rc) rc = dalloc(ctx, &sp->offsets, ncells + 1);
    if (!rc) rc = dalloc(ctx, &sp->cursor, ncells);
    if (!rc) rc = dalloc(ctx, &sp->blocksum, ncells / 4096 + 2);
    if (!rc) rc = dalloc(ctx, &sp->stats, (size_t)STAT_SLOTS * STAT_STRIDE);
    if (rc) {
        free_species(sp);
        return rc;
    }
    hipMemset(sp->acc, 0, NCH * nn * sizeof(unsigned long long));
    hipMemset(sp->sums, 0, NCH * nn * sizeof(double));
    hipMemset(sp->mpc_sum, 0, ncells * sizeof(long long));
    hipMemset(sp->flags, 0, sizeof(int));
    // Material.init: the "t" field starts at init_vals.T, t1-t3 and the averages at zero
    // (Material.java:718-736, KineticMaterial.java:103-110)
    {
        std::vector<double> g(NGAS * nn, 0.0);
        for (size_t u = 0; u < nn; u++) g[u] = desc->init_T;
        hipMemcpy(sp->gas, g.data(), NGAS * nn * sizeof(double), hipMemcpyHostToDevice);
    }
    hipMemset(sp->counts, 0, ncells * sizeof(int32_t));
    hipMemset(sp->offsets, 0, (ncells + 1) * sizeof(int32_t));
    // Material.init: nd starts at nd_back, u = v = w = 0
    std::vector<double> init(4 * nn, 0.0);
    for (size_t u = 0; u < nn; u++) init[u] = desc->nd_back;
    hipMemcpy(sp->den, init.data(), 4 * nn * sizeof(double), hipMemcpyHostToDevice);
    SF_HIP(ctx, hipDeviceSynchronize());
    ctx->species.push_back(sp);
    *species = (int)ctx->species.size() - 1;
    return SF_OK;
}

int sf_particles_add(sf_ctx* ctx, int species, int64_t n, const double* x, const double* y, const double* u,
                     const double* v, const double* w, const int* ids, int mode) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp) return sf_fail(ctx, SF_ERR_ARG, "unknown species %d", species);
    if (n < 0 || (n > 0 && (!x || !y || !u || !v || !w))) return sf_fail(ctx, SF_ERR_ARG, "sf_particles_add: bad arrays");
    if (mode != 0 && mode != 1) return sf_fail(ctx, SF_ERR_ARG, "sf_particles_add: mode must be 0 or 1");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return add_particles_device(ctx, sp, n, x, y, u, v, w, ids, mode);
}

int sf_particles_add_noalloc(sf_ctx* ctx, int species, int64_t n, const double* x, const double* y,
                             const double* u, const double* v, const double* w, const int* ids, int mode) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp) return sf_fail(ctx, SF_ERR_ARG, "unknown species %d", species);
    if (n < 0 || (n > 0 && (!x || !y || !u || !v || !w)))
        return sf_fail(ctx, SF_ERR_ARG, "sf_particles_add_noalloc: bad arrays");
    if (mode != 0 && mode != 1) return sf_fail(ctx, SF_ERR_ARG, "sf_particles_add_noalloc: mode must be 0 or 1");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return add_particles_noalloc(ctx, sp, n, x, y, u, v, w, ids, mode);
}

int sf_particles_count(sf_ctx* ctx, int species, int64_t* n) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || !n) return sf_fail(ctx, SF_ERR_ARG, "sf_particles_count: bad argument");
    *n = sp->nalive;
    return SF_OK;
}

int sf_particles_download(sf_ctx* ctx, int species, double* x, double* y, double* lc0, double* lc1, double* u,
                          double* v, double* w, int* ids) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp) return sf_fail(ctx, SF_ERR_ARG, "unknown species %d", species);
    if (ids && !sp->desc.track_ids) return sf_fail(ctx, SF_ERR_STATE, "species does not track ids");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t n = sp->np;
    if (n == 0) return SF_OK;
    const PState& s = sp->buf[sp->cur];
    std::vector<int32_t> cell(n), id;
    std::vector<float> fx(n), fy(n), vx(n), vy(n), vz(n);
    SF_HIP(ctx, hipMemcpy(cell.data(), s.cell, n * 4, hipMemcpyDeviceToHost));
    SF_HIP(ctx, hipMemcpy(fx.data(), s.fx, n * 4, hipMemcpyDeviceToHost));
    SF_HIP(ctx, hipMemcpy(fy.data(), s.fy, n * 4, hipMemcpyDeviceToHost));
    SF_HIP(ctx, hipMemcpy(vx.data(), s.vx, n * 4, hipMemcpyDeviceToHost));
    SF_HIP(ctx, hipMemcpy(vy.data(), s.vy, n * 4, hipMemcpyDeviceToHost));
    SF_HIP(ctx, hipMemcpy(vz.data(), s.vz, n * 4, hipMemcpyDeviceToHost));
    if (ids) {
        id.resize(n);
        SF_HIP(ctx, hipMemcpy(id.data(), s.id, n * 4, hipMemcpyDeviceToHost));
    }
    const MeshP& m = ctx->mp;
    int64_t o = 0;   // dead particles leave holes (cell < 0) until the next sort
    for (int64_t k = 0; k < n; k++) {
        if (cell[k] < 0) continue;
        double l0 = (cell[k] & 0xFFFF) + (double)fx[k];
        double l1 = ((cell[k] >> 16) & 0xFFFF) + (double)fy[k];
        if (x) x[o] = m.x0[0] + l0 * m.dh[0];
        if (y) y[o] = m.x0[1] + l1 * m.dh[1];
        if (lc0) lc0[o] = l0;
        if (lc1) lc1[o] = l1;
        if (u) u[o] = vx[k];
        if (v) v[o] = vy[k];
        if (w) w[o] = vz[k];
        if (ids) ids[o] = id[k];
        o++;
    }
    if (o != sp->nalive) return sf_fail(ctx, SF_ERR_STATE, "download found %lld particles, expected %lld",
                                        (long long)o, (long long)sp->nalive);
    return SF_OK;
}

// ---- restart block of a kinetic species (KineticMaterial.saveRestartData / loadRestartData,
// core/materials/KineticMaterial.java:904-1000), in java.io.DataOutputStream's big-endian encoding
static void be_put(uint8_t*& p, uint64_t bits, int nbytes) {
    for (int k = nbytes - 1; k >= 0; k--) *p++ = (uint8_t)(bits >> (8 * k));
}
static void be_double(uint8_t*& p, double v) {
    uint64_t b;
    std::memcpy(&b, &v, 8);
    be_put(p, b, 8);
}
static uint64_t be_get(const uint8_t*& p, int nbytes) {
    uint64_t b = 0;
    for (int k = 0; k < nbytes; k++) b = (b << 8) | *p++;
    return b;
}
static double be_get_double(const uint8_t*& p) {
    uint64_t b = be_get(p, 8);
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}
static const int RESTART_FIELDS[9] = {0, 16, 12, 1, 2, 3, 17, 18, 19};   // den, den-ave, T, U, V, W, u/v/w-ave
static const int64_t RESTART_PART_BYTES = 11 * 8 + 4 + 4;

int sf_species_restart_size(sf_ctx* ctx, int species, int64_t* bytes) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || !bytes) return sf_fail(ctx, SF_ERR_ARG, "sf_species_restart_size: bad argument");
    *bytes = 4 + 8 + sp->nalive * RESTART_PART_BYTES + 9 * (int64_t)ctx->mp.ni * ctx->mp.nj * 8;
    return SF_OK;
}

int sf_species_restart_write(sf_ctx* ctx, int species, uint8_t* buf, int64_t bytes) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    int64_t need = 0;
    if (!sp || !buf || sf_species_restart_size(ctx, species, &need) != SF_OK || bytes < need)
        return sf_fail(ctx, SF_ERR_ARG, "sf_species_restart_write: bad argument or buffer shorter than %lld bytes",
                       (long long)need);
    const int64_t n = sp->nalive;
    std::vector<double> x(n), y(n), l0(n), l1(n), u(n), v(n), w(n);
    std::vector<int> id(sp->desc.track_ids ? n : 0);
    int rc = sf_particles_download(ctx, species, x.data(), y.data(), l0.data(), l1.data(), u.data(), v.data(), w.data(),
                                   sp->desc.track_ids ? id.data() : nullptr);
    if (rc) return rc;
    uint8_t* p = buf;
    be_put(p, (uint32_t)sp->num_samples, 4);
    be_put(p, (uint64_t)n, 8);
    for (int64_t k = 0; k < n; k++) {
        // pos[i], vel[i] interleaved; pos[2] (z in XY, theta in RZ) is not kept on the device
        be_double(p, x[k]);
        be_double(p, u[k]);
        be_double(p, y[k]);
        be_double(p, v[k]);
        be_double(p, 0.0);
        be_double(p, w[k]);
        be_double(p, l0[k]);
        be_double(p, l1[k]);
        be_double(p, 0.0);                 // dt: no time left after a completed push
        be_double(p, sp->desc.spwt);       // mpw
        be_double(p, sp->mass);
        be_put(p, 0u, 4);                  // born_it is not tracked
        be_put(p, (uint32_t)(sp->desc.track_ids ? id[k] : (int)k), 4);
    }
    const MeshP& m = ctx->mp;
    const size_t nn = (size_t)m.ni * m.nj;
    std::vector<double> f(nn);
    for (int q = 0; q < 9; q++) {
        rc = species_field(ctx, sp, RESTART_FIELDS[q], f.data());
        if (rc) return rc;
        for (int i = 0; i < m.ni; i++)          // Field2D.binaryWrite: data[i][j], j fastest
            for (int j = 0; j < m.nj; j++) be_double(p, f[(size_t)j * m.ni + i]);
    }
    return SF_OK;
}

int sf_species_restart_read(sf_ctx* ctx, int species, const uint8_t* buf, int64_t bytes, int64_t* consumed) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || !buf || bytes < 12) return sf_fail(ctx, SF_ERR_ARG, "sf_species_restart_read: bad argument");
    const uint8_t* p = buf;
    const int num_samples = (int)(int32_t)be_get(p, 4);
    const int64_t n = (int64_t)be_get(p, 8);
    const MeshP& m = ctx->mp;
    const int64_t total = 12 + n * RESTART_PART_BYTES + 9 * (int64_t)m.ni * m.nj * 8;
    if (n < 0 || bytes < total)
        return sf_fail(ctx, SF_ERR_ARG, "sf_species_restart_read: block of %lld particles needs %lld bytes, got %lld",
                       (long long)n, (long long)total, (long long)bytes);
    std::vector<double> x(n), y(n), u(n), v(n), w(n);
    for (int64_t k = 0; k < n; k++) {
        x[k] = be_get_double(p);
        u[k] = be_get_double(p);
        y[k] = be_get_double(p);
        v[k] = be_get_double(p);
        be_get_double(p);                  // pos[2]
        w[k] = be_get_double(p);
        p += 8 * 5 + 4 + 4;                // lc (recomputed from pos), dt, mpw, mass, born_it, id
    }
    // the fields are overwritten by the next push and sampling; they are skipped
    p += 9 * (size_t)m.ni * m.nj * 8;
    // KineticMaterial.loadRestartData adds every particle with addParticle: velocity rewind with
    // the current E and a new id (KineticMaterial.java:759-802)
    if (n > 0) {
        int rc = sf_particles_add(ctx, species, n, x.data(), y.data(), u.data(), v.data(), w.data(), nullptr, 1);
        if (rc < 0) return rc;
    }
    sp->num_samples = num_samples;
    if (consumed) *consumed = (int64_t)(p - buf);
    return SF_OK;
}

int sf_particles_preload(sf_ctx* ctx, int species, int64_t n, const double* drift3, double temp, uint64_t seed) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || !drift3 || n < 0 || !(temp >= 0)) return sf_fail(ctx, SF_ERR_ARG, "sf_particles_preload: bad argument");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    TRY(preload(ctx, sp, n, drift3, temp, seed));
    return sort_species(ctx, sp);
}

int sf_push_deposit(sf_ctx* ctx, int species) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp) return sf_fail(ctx, SF_ERR_ARG, "unknown species %d", species);
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return push_deposit(ctx, sp);
}

int sf_species_set_sort_interval(sf_ctx* ctx, int species, int interval) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp) return sf_fail(ctx, SF_ERR_ARG, "unknown species %d", species);
    if (interval < 1 || interval > 64) return sf_fail(ctx, SF_ERR_ARG, "sort interval must be in [1, 64]");
    sp->sort_interval = interval;
    return SF_OK;
}

int sf_species_clear_samples(sf_ctx* ctx, int species) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp) return sf_fail(ctx, SF_ERR_ARG, "unknown species %d", species);
    const size_t nn = (size_t)ctx->mp.ni * ctx->mp.nj, ncells = (size_t)ctx->mp.nci * ctx->mp.ncj;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipMemsetAsync(sp->sums, 0, NCH * nn * sizeof(double), ctx->stream));
    SF_HIP(ctx, hipMemsetAsync(sp->mpc_sum, 0, ncells * sizeof(long long), ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sp->num_samples = 0;
    return SF_OK;
}

int sf_mesh_segment_boundaries(sf_ctx* ctx, int nseg, const int* seg_bnd, int nbnd) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "sf_mesh_segment_boundaries: mesh not set up");
    if (nseg != ctx->nseg || nbnd < 0 || (nseg > 0 && !seg_bnd))
        return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_segment_boundaries: %d segments given, the mesh has %d", nseg,
                       ctx->nseg);
    for (int s = 0; s < nseg; s++)
        if (seg_bnd[s] < -1 || seg_bnd[s] >= nbnd)
            return sf_fail(ctx, SF_ERR_ARG, "sf_mesh_segment_boundaries: segment %d names boundary %d of %d", s,
                           seg_bnd[s], nbnd);
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<DevSeg> segs(nseg);
    if (nseg > 0) {
        SF_HIP(ctx, hipMemcpy(segs.data(), ctx->d_seg, nseg * sizeof(DevSeg), hipMemcpyDeviceToHost));
        for (int s = 0; s < nseg; s++) segs[s].bnd = seg_bnd[s];
        SF_HIP(ctx, hipMemcpy(ctx->d_seg, segs.data(), nseg * sizeof(DevSeg), hipMemcpyHostToDevice));
    }
    ctx->nbnd = nbnd;
    for (Species* sp : ctx->species) {
        dfree(sp->bnd_hits);
        if (nbnd > 0) {
            TRY(dalloc(ctx, &sp->bnd_hits, (size_t)nbnd));
            SF_HIP(ctx, hipMemset(sp->bnd_hits, 0, nbnd * sizeof(unsigned long long)));
        }
    }
    return SF_OK;
}

int sf_species_boundary_hits(sf_ctx* ctx, int species, int nbnd, int64_t* counts, int clear) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || nbnd != ctx->nbnd || (nbnd > 0 && !counts))
        return sf_fail(ctx, SF_ERR_ARG, "sf_species_boundary_hits: bad argument (the mesh tallies %d boundaries)",
                       ctx->nbnd);
    if (nbnd == 0) return SF_OK;
    SF_HIP(ctx, hipSetDevice(ctx->device));
    SF_HIP(ctx, hipMemcpyAsync(counts, sp->bnd_hits, nbnd * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    if (clear) SF_HIP(ctx, hipMemsetAsync(sp->bnd_hits, 0, nbnd * sizeof(unsigned long long), ctx->stream));
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SF_OK;
}

int sf_species_stats(sf_ctx* ctx, int species, double* out5) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || !out5) return sf_fail(ctx, SF_ERR_ARG, "sf_species_stats: bad argument");
    for (int k = 0; k < 5; k++) out5[k] = sp->stats_host[k];
    return SF_OK;
}

int sf_species_field(sf_ctx* ctx, int species, int which, double* out) {
    SF_CHECK_CTX(ctx);
    Species* sp = get_species(ctx, species);
    if (!sp || !out) return sf_fail(ctx, SF_ERR_ARG, "sf_species_field: bad argument");
    if (which < 0 || which > 21) return sf_fail(ctx, SF_ERR_ARG, "sf_species_field: unknown field %d", which);
    SF_HIP(ctx, hipSetDevice(ctx->device));
    return species_field(ctx, sp, which, out);
}

int sf_last_timings(sf_ctx* ctx, double* out4) {
    SF_CHECK_CTX(ctx);
    if (!out4) return SF_ERR_ARG;
    for (int k = 0; k < 4; k++) out4[k] = ctx->last_ms[k];
    return SF_OK;
}

int sf_last_push_timings(sf_ctx* ctx, double* out3) {
    SF_CHECK_CTX(ctx);
    if (!out3) return SF_ERR_ARG;
    for (int k = 0; k < 3; k++) out3[k] = ctx->last_push_ms[k];
    return SF_OK;
}

int sf_last_push_queue(sf_ctx* ctx, double* out2) {
    SF_CHECK_CTX(ctx);
    if (!out2) return SF_ERR_ARG;
    out2[0] = ctx->last_queue[0];
    out2[1] = ctx->last_queue[1];
    return SF_OK;
}

int sf_set_timing(sf_ctx* ctx, int enabled) {
    SF_CHECK_CTX(ctx);
    ctx->timing = enabled != 0;
    return SF_OK;
}

}  // extern "C"

namespace sf {

// Downloads the instantaneous fields (nd, u, v, w), the running moment sums (FP64 on the
// device), mpc-sum, and the computeFields outputs the push last refreshed.
int species_field(sf_ctx* ctx, Species* sp, int which, double* out) {
    const MeshP& m = ctx->mp;
    const size_t nn = (size_t)m.ni * m.nj;
    SF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (which <= 3) {
        SF_HIP(ctx, hipMemcpy(out, sp->den + which * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
        return SF_OK;
    }
    if (which >= 4 && which <= 10) {
        // count-sum, u-sum, v-sum, w-sum, uu-sum, vv-sum, ww-sum
        SF_HIP(ctx, hipMemcpy(out, sp->sums + (which - 4) * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
        return SF_OK;
    }
    if (which == 11) {
        // mpc-sum is cell data, stored on node (i,j) of the cell
        const size_t ncells = (size_t)m.nci * m.ncj;
        std::vector<long long> mpc(ncells);
        SF_HIP(ctx, hipMemcpy(mpc.data(), sp->mpc_sum, ncells * sizeof(long long), hipMemcpyDeviceToHost));
        for (size_t u = 0; u < nn; u++) out[u] = 0;
        for (int j = 0; j < m.ncj; j++)
            for (int i = 0; i < m.nci; i++) out[(size_t)j * m.ni + i] = (double)mpc[(size_t)j * m.nci + i];
        return SF_OK;
    }
    // 12..21: t, t1, t2, t3, nd-ave, u-ave, v-ave, w-ave, p, mpc
    SF_HIP(ctx, hipMemcpy(out, sp->gas + (size_t)(which - 12) * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
    return SF_OK;
}

}  // namespace sf
