/*
 * starfish_oracle.h -- C API of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++17 FP64 restatement of
 * the Java ES-PIC hot path of particleincell/Starfish (particle move, Field2D
 * scatter, moment sampling, finite-volume operator assembly, Newton-Raphson
 * with the reference's PCG and GS linear solvers, gradient field).  It exists
 * to check the HIP implementation and to time the reference algorithm on host
 * cores.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product library never links or calls it.
 *
 * All node arrays use the reference's flattened IJtoN order u = j*ni + i
 * (core/domain/Mesh.java:893-897, core/common/Vec.java:165-212).
 */
#ifndef STARFISH_ORACLE_H
#define STARFISH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* mesh + geometry inputs (what Starfish's Mesh/DomainModule hold after init) */
typedef struct {
    int domain;              /* 0 XY, 1 RZ, 2 ZR (DomainModule.DomainType) */
    int ni, nj;              /* nodes */
    double x0[2], dh[2];     /* UniformMesh origin and spacing */
    const int* face_type;    /* [2*nj+2*ni]: RIGHT[nj], TOP[ni], LEFT[nj], BOTTOM[ni]; Mesh.DomainBoundaryType values */
    const double* face_value;/* same layout, mesh-bc values */
    const int* node_type;    /* [ni*nj] Mesh.NodeType values */
    const double* node_bc;   /* [ni*nj] Node.bc_value */
    const double* node_vol;  /* [ni*nj] NodeVol field (incl. Monte-Carlo interface correction) */
    int nseg;                /* linear boundary segments */
    const double* seg_xy;    /* [nseg*4] x1,y1,x2,y2 */
    const int* seg_btype;    /* Boundary.BoundaryType of the parent boundary */
    const int* seg_absorb;   /* 1: parent boundary has a material (default surface interaction absorbs) */
    const int* node_seg_ptr; /* [ni*nj+1] CSR of Node.segments */
    const int* node_seg_idx;
    int64_t seed;            /* java.util.Random seed of the shared RNG */
    int nthreads;            /* mover threads (max_cores) */
    double dt;               /* time step */
    const int* seg_bnd;      /* [nseg] boundary index of each segment (Segment.getBoundary), may be NULL */
    int nbnd;                /* number of boundaries */
} orc_mesh_desc;

orc_sim* orc_create(const orc_mesh_desc* desc);
void orc_destroy(orc_sim* s);
/* particles of species sp absorbed on each boundary since the last call with clear != 0
   (Material.addSurfaceMassDeposit / getFluxInst, core/materials/Material.java:784-813) */
int orc_boundary_hits(orc_sim* s, int sp, int64_t* out, int nbnd, int clear);

int orc_add_species(orc_sim* s, double charge_e, double molwt, double spwt, double nd_back);
void orc_set_solver(orc_sim* s, int method /*0 GS, 1 PCG*/, int linear, double n0, double Te0, double phi0,
                    double eps_r, int max_it, double tol, int nl_max_it, double nl_tol);
/* Maxwellian boundary source on a polyline (MaxwellianSource.java); returns source index */
int orc_add_source(orc_sim* s, int sp, int npts, const double* pts, int reverse, double mdot, double v_drift,
                   double temp);

/* particles: pos3/vel3 are [n*3]; rewind=1 follows KineticMaterial.addParticle (half-step rewind) */
void orc_load_particles(orc_sim* s, int sp, int64_t n, const double* pos3, const double* vel3, const double* mpw,
                        const int* id, int rewind);
int64_t orc_num_particles(const orc_sim* s, int sp);
void orc_get_particles(const orc_sim* s, int sp, double* pos3, double* vel3, double* lc2, double* mpw, int* id);

/* fields: name in {phi, rho, efi, efj} (sp ignored) or a species field
   {nd, u, v, w, count-sum, u-sum, v-sum, w-sum, uu-sum, vv-sum, ww-sum, mpc-sum, t, t1, t2, t3, nd-ave,
    u-ave, v-ave, w-ave, p, mpc} */
int orc_set_field(orc_sim* s, int sp, const char* name, const double* data);
int orc_get_field(const orc_sim* s, int sp, const char* name, double* out);

void orc_move(orc_sim* s);                 /* KineticMaterial.moveParticles(false) for every species */
void orc_update_species_fields(orc_sim* s);/* updateFields(md) + updateGasProperties + nd_back (Material.update) */
void orc_compute_fields(orc_sim* s);       /* KineticMaterial.computeFields for every species */
void orc_update_rho(orc_sim* s);           /* SolverModule.updateRho */
int orc_solve_potential(orc_sim* s);       /* PoissonSolver.update; returns NR iterations */
void orc_update_efield(orc_sim* s);        /* PotentialSolver.updateGradientField */
void orc_initial_solve(orc_sim* s);        /* SolverModule.updateFields before the main loop */
void orc_step(orc_sim* s, int nsteps);     /* Starfish.MainLoop body */
/* MaxwellianSource v_az and Source start_it / stop_it (end_it) */
void orc_source_options(orc_sim* s, int src, double v_az, int start_it, int stop_it);
/* Material init_vals.T (the temperature computeFields writes on empty nodes) */
void orc_species_init_T(orc_sim* s, int sp, double T);
/* TimeModule step counter (Starfish.getIt(); step() advances it) */
void orc_set_it(orc_sim* s, int it);
void orc_sample_sources(orc_sim* s);       /* SourceModule.sampleSources at the current step */
int orc_get_it(const orc_sim* s);
void orc_get_stats(const orc_sim* s, int sp, double* out5); /* mass, momentum[3], energy sums */
void orc_get_timers(const orc_sim* s, double* out4);         /* seconds: source, move, deposit+samples, solve+E */
int orc_last_lin_iterations(const orc_sim* s);

/* assembled operators: which = 0 A, 1 Gi, 2 Gj (Solver.initCoefficients) */
int64_t orc_operator_nnz(orc_sim* s, int which);
void orc_operator_csr(orc_sim* s, int which, int64_t* rowptr, int* col, double* val);
int orc_fixed_nodes(orc_sim* s, int* out);

/* java.util.Random known-answer hooks */
void orc_set_rng_state(orc_sim* s, int64_t state); /* raw 48-bit LCG state */
void orc_javarandom_doubles(int64_t seed, int n, double* out);
void orc_javarandom_ints(int64_t seed, int n, int* out);
void orc_javarandom_ints_bounded(int64_t seed, int bound, int n, int* out);

#ifdef __cplusplus
}
#endif
#endif
