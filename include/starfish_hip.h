/*
 * starfish_hip.h -- C-ABI of libstarfish_hip.so, the MI355X (gfx950) implementation of
 * the Starfish ES-PIC hot path: KineticMaterial particle move, Field2D charge scatter and
 * moment sampling, and the Boltzmann-electron nonlinear Poisson solve, on uniform XY and
 * RZ meshes.
 *
 * Conventions
 *   - Every entry point returns int: 0 (SF_OK) on success, < 0 on error, > 0 for
 *     "completed with a warning" (e.g. SF_WARN_NOT_CONVERGED).  No exceptions or abort()
 *     cross the ABI; sf_last_error() returns the message of the last failure on a context.
 *   - Node arrays are contiguous FP64 in the reference's flattened order u = j*ni + i
 *     (Mesh.IJtoN, core/domain/Mesh.java:893-897; Vec.deflate, core/common/Vec.java:165-183),
 *     which is what Solver.MeshData and PoissonSolver hold (pic/PoissonSolver.java:123-125).
 *   - Face arrays are [2*nj + 2*ni] in the reference's Face order: RIGHT[nj], TOP[ni],
 *     LEFT[nj], BOTTOM[ni] (core/domain/Mesh.java:109-121).
 *   - Host buffers are caller-owned and only read/written during the call; device buffers
 *     are owned by the context.  One host thread per context; calls are synchronous.
 *   - Paths in citations are relative to the reference's src/starfish/ directory.
 */
#ifndef STARFISH_HIP_H
#define STARFISH_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_OK 0
#define SF_WARN_NOT_CONVERGED 1
/* sf_push_deposit: a particle's moment contribution exceeded the int64 fixed-point range (a
   velocity deviating from the species' mean by about 2^8 times the previous step's largest
   deviation); the moment channels of that step are not valid */
#define SF_WARN_MOMENT_RANGE 2
#define SF_ERR_ARG (-1)
#define SF_ERR_HIP (-2)
#define SF_ERR_STATE (-3)
#define SF_ERR_OOM (-4)
#define SF_ERR_UNSUPPORTED (-5)
#define SF_ERR_TIMEOUT (-6)

/* domain types (core/domain/DomainModule.java DomainType) */
#define SF_DOMAIN_XY 0
#define SF_DOMAIN_RZ 1
#define SF_DOMAIN_ZR 2

/* mesh face boundary types (Mesh.DomainBoundaryType, core/domain/Mesh.java:146-160) */
#define SF_FACE_OPEN (-1)
#define SF_FACE_DIRICHLET 0
#define SF_FACE_NEUMANN 1
#define SF_FACE_PERIODIC 2
#define SF_FACE_SYMMETRY 3
#define SF_FACE_MESH 4
#define SF_FACE_SINK 5
#define SF_FACE_CIRCUIT 6

/* node types (Mesh.NodeType, core/domain/Mesh.java:176-190) */
#define SF_NODE_UNKNOWN (-2)
#define SF_NODE_OPEN (-1)
#define SF_NODE_DIRICHLET 0
#define SF_NODE_INSULATOR 1
#define SF_NODE_INTERNAL 2

/* boundary (spline) types (Boundary.BoundaryType, core/boundaries/Boundary.java:120-131) */
#define SF_BND_DIRICHLET 0
#define SF_BND_OPEN 1
#define SF_BND_SINK 2
#define SF_BND_VIRTUAL 3
#define SF_BND_INSULATOR 4

typedef struct sf_ctx sf_ctx;
typedef struct sf_geometry sf_geometry;

/* UniformMesh metrics (core/domain/UniformMesh.java:114-161) */
typedef struct {
    int domain;          /* SF_DOMAIN_* */
    int ni, nj;          /* number of nodes */
    double x0[2];        /* origin */
    double dh[2];        /* spacing */
} sf_mesh_desc;

/* ------------------------------------------------------------------------------------------
 * Library / context
 * ------------------------------------------------------------------------------------------ */
const char* sf_version(void);
/* Creates a context on HIP device `device` (device = -1: host-only context that supports the
   geometry builder, the shared generator and nothing that touches a GPU); replaces the per-run state Starfish keeps in
   KineticMaterial.MeshData and Solver.MeshData (core/materials/KineticMaterial.java:1314-1440,
   core/solver/Solver.java:119-135). */
int sf_ctx_create(int device, sf_ctx** out);
int sf_ctx_destroy(sf_ctx* ctx);
const char* sf_last_error(const sf_ctx* ctx);
/* Time step used by push, injection rewind and sources (TimeModule dt, core/common/TimeModule.java). */
int sf_ctx_set_dt(sf_ctx* ctx, double dt);
/* Seeds the context's java.util.Random-compatible shared generator (Starfish.random,
   core/common/Starfish.java:238); `raw_state` != 0 sets the 48-bit LCG state directly. */
int sf_ctx_set_rng(sf_ctx* ctx, int64_t seed, int raw_state);
int sf_ctx_get_rng_state(const sf_ctx* ctx, int64_t* state);
/* Blocks until all queued device work of the context is done. */
int sf_ctx_synchronize(sf_ctx* ctx);
/* Current time step of TimeModule (Starfish.getIt(), core/common/TimeModule.java:131-142; the
   first step is 1).  The push refreshes the computeFields outputs when the step is a multiple of
   10 and on a species' first update (KineticMaterial.updateGasProperties,
   core/materials/KineticMaterial.java:1553-1563).  While the host never sets it, each species
   counts its own sf_push_deposit calls as the step. */
int sf_ctx_set_it(sf_ctx* ctx, int it);

/* ------------------------------------------------------------------------------------------
 * Host-side mesh geometry (what Starfish's Mesh.setBoundaries + Mesh.initNodes produce,
 * core/domain/Mesh.java:364-402, 1204-1468, called from DomainModule.start,
 * core/domain/DomainModule.java:267-279).  A Java host already has these arrays and passes
 * them to sf_mesh_setup directly; non-Java hosts build them here.
 * Boundaries are linear polylines (Spline.setPath "M ... L ...", core/boundaries/Spline.java:126-285).
 * The Monte-Carlo interface-node volumes draw from the context's shared generator, as the
 * reference draws from Starfish.rnd2() (Mesh.computeInterfaceNodeVol, Mesh.java:1053-1084).
 * ------------------------------------------------------------------------------------------ */
int sf_geometry_build(sf_ctx* ctx, const sf_mesh_desc* mesh, const int* face_type, const double* face_value,
                      int nbnd, const int* bnd_type, const double* bnd_value, const int* bnd_has_material,
                      const int* bnd_npts, const double* bnd_pts, const int* bnd_reverse, sf_geometry** out);
/* sizes: nseg (linear segments), nlist (total length of the node segment lists) */
int sf_geometry_sizes(const sf_geometry* g, int* nseg, int* nlist);
int sf_geometry_export(const sf_geometry* g, int* node_type, double* node_bc, double* node_vol, double* seg_xy,
                       int* seg_btype, int* seg_absorb, int* node_seg_ptr, int* node_seg_idx);
/* seg_bnd: [nseg] index of the boundary each segment belongs to (Segment.getBoundary) */
int sf_geometry_segment_boundaries(const sf_geometry* g, int* seg_bnd);
int sf_geometry_free(sf_geometry* g);

/* ------------------------------------------------------------------------------------------
 * Mesh upload (device-resident for the run).  node_type/node_bc/node_vol: [ni*nj];
 * face_type/face_value: face layout; seg_*: [nseg] linear boundary segments with parent
 * boundary type and "has a material" (default surface interaction absorbs,
 * Material.performSurfaceInteraction, core/materials/Material.java:279-336); node segment lists
 * in CSR form (Node.segments, Mesh.java:127-135).
 * ------------------------------------------------------------------------------------------ */
int sf_mesh_setup(sf_ctx* ctx, const sf_mesh_desc* mesh, const int* face_type, const double* face_value,
                  const int* node_type, const double* node_bc, const double* node_vol, int nseg,
                  const double* seg_xy, const int* seg_btype, const int* seg_absorb, const int* node_seg_ptr,
                  const int* node_seg_idx);
/* Boundary index of each uploaded segment (seg_bnd [nseg], -1 = not tallied; the order of
   Starfish.getBoundaryList()).  From then on every species counts, per boundary, the particles
   a surface absorbs (Material.addSurfaceMassDeposit, core/materials/Material.java:784-795),
   which StatsModule divides by boundary area and time for its flux columns
   (Material.getFluxInst, Material.java:806-813; core/diagnostics/StatsModule.java:141-147). */
int sf_mesh_segment_boundaries(sf_ctx* ctx, int nseg, const int* seg_bnd, int nbnd);

/* ------------------------------------------------------------------------------------------
 * Kinetic species (KineticMaterial, core/materials/KineticMaterial.java; Material keys
 * charge/molwt/spwt/init nd_back/init T, Material.java:62-80, 580-608).
 * vmax: expected maximum particle speed (m/s, 0 selects 1e5), sets the fixed-point scale of the
 * mover sums (momentum, speed).  The moment channels carry v - c and (v - c)^2 about the species'
 * mean velocity c, at scales set at every push from the particles' largest deviation from c
 * (measured on the first push after a bulk load, then carried by each push); particles beyond
 * the scales' fast range are deposited with an exact slower quantisation.
 * capacity: maximum particles held on the device.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    double charge;       /* in units of e (Material "charge") */
    double molwt;        /* in amu (Material "molwt") */
    double spwt;         /* specific weight (KineticMaterial "spwt") */
    double nd_back;      /* background density added after deposit (init nd_back) */
    double vmax;         /* fixed-point velocity scale hint, m/s */
    int64_t capacity;    /* particle capacity */
    int track_ids;       /* 1: carry an int32 particle id through sort/compaction */
    double init_T;       /* init T (Material init_vals.T, default 0): temperature of empty nodes */
} sf_species_desc;
int sf_species_create(sf_ctx* ctx, const sf_species_desc* desc, int* species);

/* Adds particles from FP64 host arrays (positions x,y and velocities u,v,w; z / theta is not
   kept).  mode 1 = KineticMaterial.addParticle(MeshData, Particle) semantics
   (KineticMaterial.java:759-802: XtoL, clamp to ni-1/nj-1, half-step velocity rewind with the
   current E, new id); mode 0 = load the state as given (restart-style, no rewind).
   ids may be NULL (then ids continue the species counter). */
int sf_particles_add(sf_ctx* ctx, int species, int64_t n, const double* x, const double* y, const double* u,
                     const double* v, const double* w, const int* ids, int mode);
/* No counterpart in the reference.  sf_particles_add for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
   the same arguments, acceptance rules, slots and ids, but the inputs are staged in a fixed
   device array of the library (one per GPU, held until the process ends) instead of per-call
   allocations, so at most SF_NOALLOC_MAX_ADD particles per call (SF_ERR_ARG beyond that).  Calls
   that overlap on one GPU share that array and race. */
#define SF_NOALLOC_MAX_ADD 65536
int sf_particles_add_noalloc(sf_ctx* ctx, int species, int64_t n, const double* x, const double* y,
                             const double* u, const double* v, const double* w, const int* ids, int mode);
int sf_particles_count(sf_ctx* ctx, int species, int64_t* n);
/* Downloads the sf_particles_count() live particles: physical position (x,y), logical coordinates
   (lc0, lc1), velocities, ids.  Any pointer may be NULL.  Order is the device order (cell order
   right after a counting sort). */
int sf_particles_download(sf_ctx* ctx, int species, double* x, double* y, double* lc0, double* lc1, double* u,
                          double* v, double* w, int* ids);
/* Restart block of a kinetic species, the bytes KineticMaterial.saveRestartData writes to
   restart.bin (core/materials/KineticMaterial.java:904-951, big-endian as java.io.DataOutputStream;
   RestartModule writes the step and the material count before the blocks,
   core/common/RestartModule.java:120-140): num_samples, np, per particle pos/vel interleaved,
   lc, dt, mpw, mass, born_it, id, then den, den-ave, T, U, V, W, u-ave, v-ave, w-ave as
   Field2D.binaryWrite (i-major).  pos[2], dt and born_it are written as 0 (not kept on the
   device).  Reading follows loadRestartData (KineticMaterial.java:953-1000): every particle goes
   through addParticle (velocity rewind with the current E, new id), num_samples is restored and
   the fields are skipped (the next step recomputes them). */
int sf_species_restart_size(sf_ctx* ctx, int species, int64_t* bytes);
int sf_species_restart_write(sf_ctx* ctx, int species, uint8_t* buf, int64_t bytes);
int sf_species_restart_read(sf_ctx* ctx, int species, const uint8_t* buf, int64_t bytes, int64_t* consumed);
/* Device-side preload: n particles uniformly over the mesh cells whose four nodes are OPEN,
   velocities drifting Maxwellian (drift vector + temperature in K, Utils.SampleMaxw3D's 1.2
   widening kept, core/common/Utils.java:49-57), counter-based generator keyed by seed.
   No rewind (a preloaded population starts at the half step). */
int sf_particles_preload(sf_ctx* ctx, int species, int64_t n, const double* drift3, double temp, uint64_t seed);

/* ------------------------------------------------------------------------------------------
 * Hot path: one KineticMaterial.updateFields() for a species
 * (KineticMaterial.java:117-197 + updateGasProperties 1553-1595 + Material.update nd_back
 * 618-632): gather E, leapfrog push, XY drift / RZ rotation, boundary processing (segments,
 * faces), deposit of n, nu, nv, nw and the moment sums, counting sort + compaction of dead
 * particles.  Deposits are 64-bit fixed point, so densities are bit-identical run to run.
 * The species' E field is the context's (sf_poisson_solve / sf_efield_set).
 * ------------------------------------------------------------------------------------------ */
int sf_push_deposit(sf_ctx* ctx, int species);
/* Steps between per-cell counting sorts (KineticMaterial.sortParticlesToCells,
   KineticMaterial.java:1150-1179, orders the particles only for merging; here the order serves
   the tiled deposit).  Steps in between push in place and leave dead particles as holes that the
   next sort compacts.  Results do not depend on the interval.  Default 4, minimum 1. */
int sf_species_set_sort_interval(sf_ctx* ctx, int species, int interval);
/* which: 0 nd, 1 u, 2 v, 3 w (instantaneous, KineticMaterial.updateFields);
   4 count-sum, 5 u-sum, 6 v-sum, 7 w-sum, 8 uu-sum, 9 vv-sum, 10 ww-sum, 11 mpc-sum
   (KineticMaterial.updateSamples, running FP64 sums on the device); 12 t, 13 t1, 14 t2, 15 t3,
   16 nd-ave, 17 u-ave, 18 v-ave, 19 w-ave, 20 p, 21 mpc (KineticMaterial.computeFields,
   KineticMaterial.java:1606-1671, as last refreshed by the push: on the species' first update
   and on steps that are multiples of 10; empty nodes hold init_T).  mpc is cell data, written
   at node (i, j) of cell (i, j). */
int sf_species_field(sf_ctx* ctx, int species, int which, double* out);
/* KineticMaterial.clearSamples (KineticMaterial.java:1509-1529) */
int sf_species_clear_samples(sf_ctx* ctx, int species);
/* mover sums of the last push, multiplied by mass: mass, momentum[3], "energy" = mass*sum(mpw*|v|)
   (KineticMaterial.java:247-259, 406-413) */
int sf_species_stats(sf_ctx* ctx, int species, double* out5);
/* counts[nbnd]: particles of the species absorbed on each boundary since the last call with
   clear != 0 (deprate_collection_inst; multiply by spwt for Material.getFluxInst's sum) */
int sf_species_boundary_hits(sf_ctx* ctx, int species, int nbnd, int64_t* counts, int clear);

/* ------------------------------------------------------------------------------------------
 * Field solve: PoissonSolver (pic/PoissonSolver.java) with SolverModule's keys
 * (core/solver/SolverModule.java:71-86).  The operator is the reference's finite-volume
 * assembly (Solver.initCoefficients, core/solver/Solver.java:204-256, 451-726) applied
 * matrix-free; Dirichlet and one-sided Neumann rows are eliminated into an equivalent
 * symmetric system solved with Jacobi-PCG in FP32 and FP64 residual refinement inside the
 * FP64 Newton iteration on the Boltzmann term (Solver.solveNonLinearNR, Solver.java:293-387).
 * ------------------------------------------------------------------------------------------ */
typedef struct {
    int linear;          /* PoissonSolver "linear" */
    double n0, Te0, phi0;/* Boltzmann reference density (m^-3), temperature (eV), potential (V) */
    double eps_r;        /* relative permittivity */
    int max_it;          /* max PCG iterations per linear solve (SolverModule "max_it") */
    double rel_tol;      /* stop when ||F||_2 / ||F_ref||_2 <= rel_tol (FP64 residual) */
    int nl_max_it;       /* max Newton iterations (SolverModule "nl_max_it") */
    double nl_tol;       /* > 0: Newton also stops once a step y has ||y||_2 / N < nl_tol, N the node
                            count (the reference's criterion, Solver.java:371-381); <= 0: only rel_tol */
} sf_poisson_desc;
typedef struct {
    int newton_iterations;
    int pcg_iterations;      /* summed over Newton steps and refinements */
    double rel_residual;     /* final ||F||_2 / ||F_ref||_2 */
    double ref_norm;         /* reference-style norm ||F||_2 / N */
    double ms;               /* device time of the solve */
} sf_solve_info;
int sf_poisson_setup(sf_ctx* ctx, const sf_poisson_desc* desc);
/* SolverModule.updateRho (SolverModule.java:144-171): rho = sum over created species of q*nd */
int sf_rho_update(sf_ctx* ctx);
int sf_rho_set(sf_ctx* ctx, const double* rho);
int sf_rho_get(sf_ctx* ctx, double* rho);
/* PoissonSolver.update: solves for phi (device-resident, starting from the current phi).
   phi_io may be NULL; if not NULL its content is uploaded as the initial guess and the
   result is downloaded into it. */
int sf_poisson_solve(sf_ctx* ctx, double* phi_io, sf_solve_info* info);
int sf_phi_set(sf_ctx* ctx, const double* phi);
int sf_phi_get(sf_ctx* ctx, double* phi);
/* BoltzmannElectronsMaterial.electronModelBoltzmann (core/materials/BoltzmannElectronsMaterial.java:
   179-215): den = den0 exp((phi - phi0)/kTe0) floored at 1 m^-3, from the device-resident phi, in
   node order u = j*ni + i.  den0 <= 0 or kTe0 <= 0 takes den0, kTe0 and phi0 from the solver (as
   the reference takes them from the PotentialSolver).  The electron temperature field is the
   constant kTe0 * EVtoK and is left to the caller. */
int sf_boltzmann_density(sf_ctx* ctx, double den0, double kTe0, double phi0, double* den);
/* PotentialSolver.updateGradientField (pic/PotentialSolver.java:55-70): E = -grad(phi) */
int sf_efield_update(sf_ctx* ctx);
int sf_efield_set(sf_ctx* ctx, const double* efi, const double* efj);
int sf_efield_get(sf_ctx* ctx, double* efi, double* efj);
/* Magnetic field of the domain (DomainModule "bfi"/"bfj", core/domain/DomainModule.java:174,
   filled by load_field in the reference), node order u = j*ni + i, tesla, kept in FP32 like E.
   Where the gathered (Bi, Bj) is nonzero the mover and the addParticle rewind use
   UpdateVelocityBoris (core/materials/KineticMaterial.java:336-353, 779-794, 847-893); elsewhere
   the electric kick.  Both NULL, or a field that is zero everywhere, removes it. */
int sf_bfield_set(sf_ctx* ctx, const double* bfi, const double* bfj);
/* y = A x with the reference operator (diagnostic; x, y host [ni*nj] FP64) */
int sf_operator_apply(sf_ctx* ctx, const double* x, double* y);

/* ------------------------------------------------------------------------------------------
 * Sources (sources/MaxwellianSource.java:44-125, core/source/Source.java:123-198): sampled on
 * the host from the shared generator in the reference's order, then added with
 * sf_particles_add mode 1.
 * ------------------------------------------------------------------------------------------ */
int sf_source_maxwellian(sf_ctx* ctx, int species, int npts, const double* pts, int reverse, double mdot,
                         double v_drift, double temp, int* source);
/* MaxwellianSource <v_az> (sources/MaxwellianSource.java:62, 117: added to the third velocity
   component) and Source <start_it> / <stop_it> (or <end_it>) (core/source/Source.java:68-71): the
   source samples only on steps start_it <= it <= stop_it (stop_it < 0: no end), and outside that
   window it does not regenerate either (SourceModule.java:240-241).  The step is the one set with
   sf_ctx_set_it, else the count of sampling calls. */
int sf_source_options(sf_ctx* ctx, int source, double v_az, int start_it, int stop_it);
/* SourceModule.sampleSources (core/source/SourceModule.java:231-283): every active source injects */
int sf_sources_sample(sf_ctx* ctx, int64_t* injected);
/* The same injection with the particles sampled on the device (no host sampling or upload):
   identical particle counts per step, positions along the boundary and drifting-Maxwellian
   velocities (with Utils.SampleMaxw3D's 1.2 factor and the inward-velocity rejection of
   MaxwellianSource.sampleParticle) drawn from Philox4x32-10 streams keyed by seed and counted by
   (particle of the batch, batch, source).  The shared java.util.Random stream is not consumed, so
   the draws match the reference's in distribution, not value by value. */
int sf_sources_sample_device(sf_ctx* ctx, uint64_t seed, int64_t* injected);
/* Replaces Source.getMassGeneratedInst / clearMassGeneratedInst bookkeeping (core/source/Source.java:34-47,
   mass_generated_inst += mpw * mass per added particle at :193), which StatsModule.saveStats writes:
   the number of particles source `source` (0-based, in creation order) has added to the domain since
   the context was created, by either sampling path.  The mass is count * spwt * species mass. */
int sf_source_added(sf_ctx* ctx, int source, int64_t* count);

/* ------------------------------------------------------------------------------------------
 * Timing of the last sf_push_deposit (device ms, with sf_set_timing(ctx, 1)): out4[0] the push
 * (push + deposit, in place), out4[1] cell scan (sort step) or macro-particle count (other steps),
 * out4[2] the scatter into cell order (exactly 0 on a step without a sort), out4[3] normalisation.
 * ------------------------------------------------------------------------------------------ */
int sf_last_timings(sf_ctx* ctx, double* out4);
/* Split of out4[0]: out3[0] the tiled kernel over the cell-sorted particles, out3[1] the kernel
 * over the particles appended since the last sort (injected ones), out3[2] how many slots that
 * second kernel covered. */
int sf_last_push_timings(sf_ctx* ctx, double* out3);
/* The XY tile kernel pushes in FP32 every particle whose cell is away from segments and mesh faces
 * and leaves the others to per-tile slow lists that a second kernel pushes with the full reference
 * loop: out2[0] that kernel's ms (with sf_set_timing), out2[1] how many particles the last push
 * left to it.  The environment variable SF_NO_FAST_PUSH (at context creation) makes every step
 * run the full loop in the general tile kernel. */
int sf_last_push_queue(sf_ctx* ctx, double* out2);
int sf_set_timing(sf_ctx* ctx, int enabled);

/* ------------------------------------------------------------------------------------------
 * Monte Carlo collisions with a background material: MCC (collisions/MCC.java:45-221) with a
 * constant cross-section (interactions/Sigma.java:108-113, <sigma>const</sigma>) and the CEX
 * (MCC.java:287-297) or MEX/elastic (MCC.java:242-285) model; the IONIZATION model creates
 * particles and is not provided.  The source species must track ids: Starfish.rnd() is replaced
 * by Philox4x32-10 keyed by seed and counted by (particle id, step, tag), so results do not depend
 * on the order particles are visited in.  The target material enters through its node fields
 * (density, u-ave, v-ave, w-ave, temperature; Material.sampleMaxwellianVelocity,
 * core/materials/Material.java:859-896), node order u = j*ni + i.
 * ------------------------------------------------------------------------------------------ */
#define SF_MCC_CEX 0
#define SF_MCC_MEX 1
/* tag value that lets the library choose the smallest tag not used by another live object of
   the same kind and seed; explicit tags must be below 2^28 */
#define SF_TAG_AUTO 0xFFFFFFFFu
typedef struct {
    int source_species;      /* <source> (kinetic) */
    int model;               /* SF_MCC_CEX or SF_MCC_MEX (<model>) */
    double sigma;            /* constant cross-section, m^2 (<sigma_coeffs> of <sigma>const</sigma>) */
    double target_molwt;     /* target material's molecular weight, amu */
    int frequency;           /* <frequency>: steps between passes (>= 1) */
    int stop_it;             /* <stop_it>: no passes after this step (<= 0: none) */
    double max_target_temp;  /* <max_target_temp> (<= 0: no limit) */
    uint64_t seed;           /* Philox key */
    uint32_t tag;            /* distinguishes the streams of several interactions (< 2^28, or
                                SF_TAG_AUTO); MCC, DSMC and device injection streams never overlap,
                                and two live MCC objects with the same seed and tag are refused */
} sf_mcc_desc;
typedef struct sf_mcc sf_mcc;
int sf_mcc_create(sf_ctx* ctx, const sf_mcc_desc* desc, sf_mcc** out);
int sf_mcc_destroy(sf_mcc* mcc);
/* target node fields; u, v, w, temp may be NULL (zero) */
int sf_mcc_set_target(sf_mcc* mcc, const double* den, const double* u, const double* v, const double* w,
                      const double* temp);
/* MCC.perform (MCC.java:130-221) at step `it` and simulation time `time`: skipped after stop_it
   and unless it % frequency == 0; velocities of collided particles are replaced in place */
int sf_mcc_perform(sf_mcc* mcc, int it, double time, int64_t* collisions);
/* which: 0 col-count-sum, 1 col-real-sum, 2 col-count (per pass), 3 nu (real sum per second per
   node volume, MCC.java:147-160), 4 the target's energy source S (W/m^3, summed since the last
   clear; E = 0.5 m |v| as MCC.java:205-214 writes it) */
int sf_mcc_field(sf_mcc* mcc, int which, double* out);
/* MCC.clearSamples (MCC.java:120-124); clear_source also zeroes S */
int sf_mcc_clear_samples(sf_mcc* mcc, int clear_source);

/* ------------------------------------------------------------------------------------------
 * DSMC: Bird's no-time-counter collisions of a pair of kinetic species (collisions/DSMC.java:
 * 45-430) with Boyd's weights for unequal specific weights, the elastic model (VHS scattering,
 * or VSS when the pair's vss_alpha differs from 1) and a constant or Bird eq. 4.63 VHS
 * cross-section (collisions/SigmaPlus.java:38-78, from the materials' <diam>, <ref_temp>,
 * <visc_temp_index>).  Both species must track ids: a cell's list is ordered by particle id and
 * its draws come from Philox4x32-10 counted by (cell, step, tag).  A species that was pushed in
 * place since its last sort is sorted first.  At most 2048 particles of one species per cell
 * (SF_ERR_UNSUPPORTED beyond).  The reference's steady-state clearSamples is left to the caller.
 * ------------------------------------------------------------------------------------------ */
#define SF_SIGMA_CONST 0
#define SF_SIGMA_BIRD463 1
typedef struct {
    int species_a, species_b;        /* <pair>; the same index for a single species */
    int sigma_type;                  /* SF_SIGMA_CONST or SF_SIGMA_BIRD463 */
    double sigma_c0;                 /* constant cross-section, m^2 */
    double diam_a, diam_b;           /* materials' <diam> */
    double ref_temp_a, ref_temp_b;   /* <ref_temp> (reference default 275) */
    double visc_temp_index_a, visc_temp_index_b;   /* <visc_temp_index> (default 0.85) */
    double vss_alpha_a, vss_alpha_b; /* <vss_alpha> (default 1: VHS) */
    double sig_cr_max0;              /* <sig_cr_max> initial value (reference default 1e-16) */
    int frequency;                   /* <frequency> */
    uint64_t seed;
    uint32_t tag;                    /* < 2^28 or SF_TAG_AUTO, unique per seed among live DSMC objects */
} sf_dsmc_desc;
typedef struct sf_dsmc sf_dsmc;
int sf_dsmc_create(sf_ctx* ctx, const sf_dsmc_desc* desc, sf_dsmc** out);
int sf_dsmc_destroy(sf_dsmc* dsmc);
/* DSMC.perform (DSMC.java:163-190) at step `it`; skipped unless it % frequency == 0 */
int sf_dsmc_perform(sf_dsmc* dsmc, int it, int64_t* collisions);
/* which: 0 col-count-sum, 1 col-real-sum, 2 col-count, 3 nu (node fields at the cell's first
   node), 4 sig_cr_max per cell (index j*(ni-1)+i) */
int sf_dsmc_field(sf_dsmc* dsmc, int which, double* out);
int sf_dsmc_clear_samples(sf_dsmc* dsmc);

/* ------------------------------------------------------------------------------------------
 * Learned initial guess for the potential solve.  No counterpart in the reference: the
 * reference solvers start from the previous potential (LinearSolverPCG.java:33-37,
 * LinearSolverGS.java:41-53).  A 3-level encoder-decoder CNN (32/64/128 channels, 3x3
 * convolutions, BF16 weights and activations, FP32 accumulation on the matrix cores) maps the
 * solver's input (rho, Dirichlet mask, boundary phi, Boltzmann inversion of rho) to phi.  It is
 * trained on the device by the library's own forward/backward kernels and Adam.  The guess
 * only replaces the starting point: sf_poisson_solve still runs to its residual tolerance.
 * ------------------------------------------------------------------------------------------ */
typedef struct sf_cnn sf_cnn;
/* network for an ni x nj node grid (ni, nj <= 0: the context's mesh); max_batch <= 64 */
int sf_cnn_create(sf_ctx* ctx, int ni, int nj, int max_batch, uint64_t seed, sf_cnn** out);
int sf_cnn_destroy(sf_cnn* cnn);
/* padded grid (multiples of 4) and number of FP32 parameters */
int sf_cnn_dims(const sf_cnn* cnn, int* H, int* W, int64_t* nparams);
/* training set: snapshot of a solved context (its rho, node classes and phi) or host arrays
   x4 [H][W][4], target [H][W] (phi / 100 V), mask [H][W] (1 on free nodes) */
int sf_cnn_snapshot(sf_cnn* cnn, sf_ctx* solved);
int sf_cnn_add_sample(sf_cnn* cnn, const float* x4, const float* target, const uint8_t* mask);
int sf_cnn_get_sample(sf_cnn* cnn, int idx, float* x4, float* target, uint8_t* mask);
int sf_cnn_num_samples(const sf_cnn* cnn);
/* Adam on random mini-batches; *loss_out = mean masked MSE over the last tenth of the steps */
int sf_cnn_train(sf_cnn* cnn, int steps, int batch, double lr, double* loss_out);
/* loss, network output y [H][W] and all parameter gradients (order of sf_cnn_params) of one
   sample; any output pointer may be NULL */
int sf_cnn_loss_and_grads(sf_cnn* cnn, int sample, double* loss_out, float* y_out, float* grads_out,
                          int64_t* n_out);
/* FP32 parameters: per convolution W[9][N][K] then b[N] (10 convolutions), then the head
   w[32], b[1].  get = 1 downloads into buf, get = 0 uploads from buf. */
int sf_cnn_params(sf_cnn* cnn, int get, float* buf, int64_t* n_out);
/* writes the network's phi for the current rho / boundary values of `target` into its phi,
   the starting point of the next sf_poisson_solve */
int sf_cnn_predict(sf_cnn* cnn, sf_ctx* target);
/* sf_poisson_solve started from the network's guess.  policy 0: from the guess; policy 1: from
   whichever of the current phi and the guess has the smaller nonlinear residual.  Under either
   policy a guess with a non-finite residual is replaced by the current phi, and a solve that
   diverges from the guess is repeated from the current phi, so the plain solve stays the
   fallback.  *used_guess (may be NULL) reports whether the solve started from the guess.  Same
   tolerance and status as sf_poisson_solve. */
int sf_poisson_solve_learned(sf_ctx* ctx, sf_cnn* cnn, int policy, sf_solve_info* info, int* used_guess);
/* diagnostics: activation `which` of batch slot 0 after the last forward pass, as FP32
   (0 input [H][W][32], 1 e1a, 2 u1 = up(d2)++e1, 3 p1, 4 e2a, 5 u2 = up(e3)++e2, 6 p2, 7 e3a,
   8 e3, 9 d2a, 10 d2, 11 d1a, 12 d1, 13 output y); *count = number of values (out may be NULL) */
int sf_cnn_activation(sf_cnn* cnn, int which, float* out, int64_t* count);
/* d[16][16] = a[16][32] * b[32][16] through one BF16 matrix-core instruction with the operand
   layout the convolution kernels use (a, b rounded to BF16) */
int sf_cnn_mfma_probe(sf_ctx* ctx, const float* a, const float* b, float* d);

#ifdef __cplusplus
}
#endif
#endif
