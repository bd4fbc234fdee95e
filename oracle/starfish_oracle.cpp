/*
 * starfish_oracle.cpp -- CPU FP64 restatement of the Starfish ES-PIC hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked into the product
 * library.  Every function cites the Java it restates; paths are relative to
 * the reference's src/starfish/ directory.  Numerics are plain IEEE double, as
 * in the JVM; the only deviations are stated next to the code:
 *   - java.lang.Math.exp/sqrt/asin are replaced by glibc's (exp may differ by
 *     1 ulp from the JDK's; parity for that boundary is unpinned),
 *   - the sparse matrix is a row-ordered map instead of HashMap rows (affects
 *     only the summation order of a row, i.e. the last bits),
 *   - the threaded GS sweep is run with its chunks in order (the reference's
 *     unsynchronised chunks make its iterates timing dependent; the discrete
 *     solution is the parity target),
 *   - MCC and DSMC draw from Philox4x32-10 streams (per particle id, per cell)
 *     instead of Starfish.rnd(), and DSMC cell lists are ordered by particle
 *     id; the generator is pinned by its published known-answer vectors and
 *     the collision passes by closed forms (tests/test_oracle_mcc.py,
 *     tests/test_oracle_dsmc.py).
 * Pins of every restated function: tests/test_oracle_*.py.  Parity unpinned:
 * the glibc/JDK math functions above and the Monte-Carlo interface node
 * volumes (reproduced draw for draw from the replica generator, never run
 * against Java).
 */
#include "starfish_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

/* core/common/Constants.java:20-33 */
const double EPS0 = 8.85418782e-12;
const double K_B = 1.3806503e-23;
const double QE = 1.60217646e-19;
const double AMU = 1.66053886e-27;
const double FLT_EPS = 1e-7;
const double PI = 3.14159265358979323846;
const double KtoEV = K_B / QE;
const double EVtoK = 1.0 / KtoEV;

/* ---------------------------------------------------------------------------
 * java.util.Random: 48-bit LCG as specified in the JDK documentation of
 * java.util.Random (setSeed, next, nextDouble, nextInt(bound)).  Starfish's
 * shared generator is `new Random(0)` (core/common/Starfish.java:238).
 * ------------------------------------------------------------------------- */
struct JavaRandom {
    uint64_t seed;
    explicit JavaRandom(int64_t s) { set_seed(s); }
    void set_seed(int64_t s) { seed = ((uint64_t)s ^ 0x5DEECE66DULL) & ((1ULL << 48) - 1); }
    int32_t next(int bits) {
        seed = (seed * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
        return (int32_t)(uint32_t)(seed >> (48 - bits));
    }
    double next_double() { return (double)(((int64_t)next(26) << 27) + next(27)) * 0x1.0p-53; }
    int32_t next_int() { return next(32); }
    int32_t next_int(int32_t bound) {
        if ((bound & -bound) == bound) return (int32_t)(((int64_t)bound * (int64_t)next(31)) >> 31);
        int32_t bits, val;
        do {
            bits = next(31);
            val = bits % bound;
        } while ((int32_t)((uint32_t)bits - (uint32_t)val + (uint32_t)(bound - 1)) < 0);
        return val;
    }
};

enum DomainType { XY = 0, RZ = 1, ZR = 2 };
/* core/domain/Mesh.java:109-121 Face; 146-160 DomainBoundaryType; 176-190 NodeType */
enum Face { RIGHT = 0, TOP = 1, LEFT = 2, BOTTOM = 3 };
enum DBT { DB_OPEN = -1, DB_DIRICHLET = 0, DB_NEUMANN = 1, DB_PERIODIC = 2, DB_SYMMETRY = 3, DB_MESH = 4,
           DB_SINK = 5, DB_CIRCUIT = 6 };
enum NodeType { NT_BAD = -99, NT_UNKNOWN = -2, NT_OPEN = -1, NT_DIRICHLET = 0, NT_INSULATOR = 1, NT_INTERNAL = 2 };
/* core/boundaries/Boundary.java:120-131 */
enum BType { BT_DIRICHLET = 0, BT_OPEN = 1, BT_SINK = 2, BT_VIRTUAL = 3, BT_INSULATOR = 4 };

/* core/boundaries/LinearSegment.java */
struct Segment {
    double x1[2], x2[2];
    double normal[2], tangent[2];
    double length;
    int btype;
    int absorb;
    int bnd = -1;        /* index of the parent boundary */
    void init(const double* a, const double* b) {
        x1[0] = a[0]; x1[1] = a[1]; x2[0] = b[0]; x2[1] = b[1];
        double dx = x2[0] - x1[0], dy = x2[1] - x1[1];
        length = std::sqrt(dx * dx + dy * dy);
        dx /= length; dy /= length;
        tangent[0] = dx; tangent[1] = dy;
        normal[0] = -dy; normal[1] = dx;           /* LinearSegment.java:41-44 */
    }
    void pos(double t, double* x) const {
        x[0] = x1[0] + t * (x2[0] - x1[0]);
        x[1] = x1[1] + t * (x2[1] - x1[1]);
    }
    /* LinearSegment.InfiniteLineIntersect (LinearSegment.java:164-179) */
    static bool line_intersect(const double* p1, const double* p2, const double* p3, const double* p4, double* xp) {
        double X1 = p1[0], X2 = p2[0], X3 = p3[0], X4 = p4[0];
        double Y1 = p1[1], Y2 = p2[1], Y3 = p3[1], Y4 = p4[1];
        double den = (X1 - X2) * (Y3 - Y4) - (Y1 - Y2) * (X3 - X4);
        if (den == 0) return false;
        xp[0] = ((X1 * Y2 - Y1 * X2) * (X3 - X4) - (X1 - X2) * (X3 * Y4 - Y3 * X4)) / den;
        xp[1] = ((X1 * Y2 - Y1 * X2) * (Y3 - Y4) - (Y1 - Y2) * (X3 * Y4 - Y3 * X4)) / den;
        return true;
    }
    /* LinearSegment.intersect(p3,p4) (LinearSegment.java:113-160): t[0] along this, t[1] along p3->p4 */
    void intersect(const double* p3, const double* p4, double* t) const {
        double xp[2];
        if (!line_intersect(x1, x2, p3, p4, xp)) { t[0] = -1; t[1] = -1; return; }
        if (std::fabs(x2[0] - x1[0]) > 1e-6) t[0] = (xp[0] - x1[0]) / (x2[0] - x1[0]);
        else t[0] = (xp[1] - x1[1]) / (x2[1] - x1[1]);
        if (t[0] < -FLT_EPS || t[0] > (1 + FLT_EPS)) { t[0] = -1; t[1] = -1; return; }
        if (std::fabs(p4[0] - p3[0]) > 1e-6) t[1] = (xp[0] - p3[0]) / (p4[0] - p3[0]);
        else t[1] = (xp[1] - p3[1]) / (p4[1] - p3[1]);
        if (t[1] < -FLT_EPS || t[1] > (1 + FLT_EPS)) { t[0] = -1; t[1] = -1; return; }
        if (t[0] < 0) t[0] = 0;
        if (t[1] < 0) t[1] = 0;
        if (t[0] > 1) t[0] = 1;
        if (t[1] > 1) t[1] = 1;
    }
};

struct Particle {
    double pos[3], vel[3], lc[2];
    double dt, mpw;
    int id;
};

/* node field in IJtoN order */
struct Field {
    int ni = 0, nj = 0;
    std::vector<double> d;
    void init(int ni_, int nj_, double v = 0) { ni = ni_; nj = nj_; d.assign((size_t)ni * nj, v); }
    double& at(int i, int j) { return d[(size_t)j * ni + i]; }
    double at(int i, int j) const { return d[(size_t)j * ni + i]; }
    void clear() { std::fill(d.begin(), d.end(), 0.0); }
};

/* ---------------------------------------------------------------------------
 * UniformMesh (core/domain/UniformMesh.java, core/domain/Mesh.java)
 * ------------------------------------------------------------------------- */
struct Mesh {
    int domain = XY;
    int ni = 0, nj = 0;
    double x0[2] = {0, 0}, dh[2] = {1, 1}, xd[2] = {0, 0};
    std::vector<int> face_type[4];
    std::vector<double> face_value[4];
    std::vector<int> node_type;
    std::vector<double> node_bc;
    Field node_vol;
    std::vector<Segment> seg;
    std::vector<int> nseg_ptr, nseg_idx;

    int ntype(int i, int j) const { return node_type[(size_t)j * ni + i]; }
    double& bc(int i, int j) { return node_bc[(size_t)j * ni + i]; }
    int btype(int face, int idx) const { return face_type[face][idx]; }
    bool is_dirichlet(int i, int j) const {           /* Mesh.java:332-334 */
        int t = ntype(i, j);
        return t == NT_DIRICHLET || t == NT_INTERNAL;
    }
    void pos(double i, double j, double* x) const {   /* UniformMesh.java:135-142 */
        x[0] = x0[0] + i * dh[0];
        x[1] = x0[1] + j * dh[1];
    }
    void XtoL(const double* x, double* lc) const {    /* UniformMesh.java:154-161 */
        lc[0] = (x[0] - x0[0]) / dh[0];
        lc[1] = (x[1] - x0[1]) / dh[1];
    }
    double R(double i, double j) const {              /* Mesh.java:824-831 */
        if (domain == RZ) return x0[0] + i * dh[0];
        if (domain == ZR) return x0[1] + j * dh[1];
        return 1;
    }
    int IJtoN(int i, int j) const {                   /* Mesh.java:893-897 */
        if (i >= 0 && i < ni && j >= 0 && j < nj) return j * ni + i;
        return -1;
    }
    bool contains_strict(const double* x) const {     /* UniformMesh.java:163-170 */
        return x[0] >= x0[0] && x[0] < xd[0] && x[1] >= x0[1] && x[1] < xd[1];
    }
    bool contains(const double* x) const {            /* Mesh.java:1473-1481 */
        double lc[2];
        XtoL(x, lc);
        if (lc[0] < -FLT_EPS || lc[1] < -FLT_EPS || lc[0] > (ni - 1 + FLT_EPS) || lc[1] > (nj - 1 + FLT_EPS))
            return false;
        return true;
    }
    static int minmax(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

    /* Mesh.controlVolumeLCs (Mesh.java:656-747), single-mesh domain */
    void control_volume_lcs(double i0, double j0, double delta, double lcs[4][2]) const {
        double i = -1, j = -1;
        for (int k = 0; k < 4; k++) {
            switch (k) {
                case 0: i = i0 - delta; j = j0 - delta; break;
                case 1: i = i0 + delta; j = j0 - delta; break;
                case 2: i = i0 + delta; j = j0 + delta; break;
                case 3: i = i0 - delta; j = j0 + delta; break;
            }
            int ii = (int)i0, jj = (int)j0;
            if (k > 1 && jj > 1 && (ntype(ii, jj) == NT_INSULATOR && ntype(ii, jj - 1) == NT_OPEN))
                j = j0;
            else if ((k == 0 || k == 3) && ii > 1 && ii + 1 < ni &&
                     (ntype(ii, jj) == NT_INSULATOR && ntype(ii + 1, jj) == NT_OPEN))
                i = i0;
            else if ((k == 1 || k == 2) && ii < ni - 1 && ii >= 1 &&
                     (ntype(ii, jj) == NT_INSULATOR && ntype(ii - 1, jj) == NT_OPEN))
                i = i0;
            if (i < 0 && btype(LEFT, minmax((int)(j0 + 0.5), 0, nj - 1)) != DB_MESH) i = 0;
            if (i > ni - 1 && btype(RIGHT, minmax((int)(j0 + 0.5), 0, nj - 1)) != DB_MESH) i = ni - 1;
            if (j < 0 && btype(BOTTOM, minmax((int)(i0 + 0.5), 0, ni - 1)) != DB_MESH) j = 0;
            if (j > nj - 1 && btype(TOP, minmax((int)(i0 + 0.5), 0, ni - 1)) != DB_MESH) j = nj - 1;
            /* no neighbour mesh exists in a single-mesh domain: out-of-range points are clamped */
            if (i < 0 || j < 0 || i > ni - 1 || j > nj - 1) {
                if (i < 0) i = 0;
                if (j < 0) j = 0;
                if (i >= ni - 1) i = ni - 1;
                if (j >= nj - 1) j = nj - 1;
            }
            lcs[k][0] = i;
            lcs[k][1] = j;
        }
        lcs[0][1] = std::max(lcs[0][1], lcs[1][1]);
        lcs[1][1] = lcs[0][1];
        lcs[1][0] = std::min(lcs[1][0], lcs[2][0]);
        lcs[2][0] = lcs[1][0];
        lcs[1][0] = std::min(lcs[1][0], lcs[2][0]);
        lcs[2][0] = lcs[1][0];
        lcs[2][1] = std::max(lcs[2][1], lcs[3][1]);
        lcs[3][1] = lcs[2][1];
        lcs[3][0] = std::max(lcs[3][0], lcs[0][0]);
        lcs[0][0] = lcs[3][0];
    }
    /* Mesh.area (Mesh.java:965-983), two triangles of the CV quadrilateral */
    double area(double i0, double j0, double delta = 0.5) const {
        double lcs[4][2], V[4][2];
        control_volume_lcs(i0, j0, delta, lcs);
        for (int k = 0; k < 4; k++) pos(lcs[k][0], lcs[k][1], V[k]);
        double A = 0.5 * std::fabs(V[0][0] * V[1][1] - V[0][0] * V[2][1] + V[1][0] * V[2][1] - V[1][0] * V[0][1] +
                                   V[2][0] * V[0][1] - V[2][0] * V[1][1]);
        A += 0.5 * std::fabs(V[0][0] * V[2][1] - V[0][0] * V[3][1] + V[2][0] * V[3][1] - V[2][0] * V[0][1] +
                             V[3][0] * V[0][1] - V[3][0] * V[2][1]);
        return A;
    }
    /* Mesh.nodeVol (Mesh.java:994-1020) */
    double node_volume(double i0, double j0) const {
        if (domain == XY) return 1.0 * area(i0, j0);
        double r = R(i0, j0);
        if (domain == RZ) {
            if (i0 == 0) r = R(i0 + 0.25, j0);
            else if (i0 == ni - 1) r = R(i0 - 0.25, j0);
        } else if (domain == ZR) {
            if (j0 == 0) r = R(i0, j0 + 0.25);
            else if (j0 == nj - 1) r = R(i0, j0 - 0.25);
        }
        return 2 * PI * area(i0, j0) * r;
    }
    /* Mesh.NodeIndexOffset (Mesh.java:1164-1198) */
    void node_index_offset(double im, double jm, double di, double dj, int* ii) const {
        double i = im + di, j = jm + dj;
        double fi = i - (int)i, fj = j - (int)j;
        ii[0] = (fi != 0) ? (int)im : (int)i;
        ii[1] = (fj != 0) ? (int)jm : (int)j;
        if (ii[0] < 0 && btype(LEFT, minmax(ii[1], 0, nj - 1)) != DB_MESH) ii[0] = 0;
        if (ii[1] < 0 && btype(BOTTOM, minmax(ii[0], 0, ni - 1)) != DB_MESH) ii[1] = 0;
        if (ii[0] >= ni && btype(RIGHT, minmax(ii[1], 0, nj - 1)) != DB_MESH) ii[0] = ni - 1;
        if (ii[1] >= nj && btype(TOP, minmax(ii[0], 0, ni - 1)) != DB_MESH) ii[1] = nj - 1;
    }
};

/* ---------------------------------------------------------------------------
 * Field2D scatter / gather (core/domain/Field2D.java:244-390)
 * ------------------------------------------------------------------------- */
void scatter(const Mesh& m, Field& f, double fi, double fj, double val) {
    int i = (int)fi, j = (int)fj;
    double di = fi - i, dj = fj - j;
    if (i < 0 || j < 0 || i >= f.ni - 1 || j >= f.nj - 1) return;
    if (m.domain == RZ) {                               /* Ruyten eq. 4.2, Field2D.java:272-276 */
        double rp = m.R(i + 1, fj), rm = m.R(i, fj), r = m.R(fi, fj);
        di = 1 - (0.5 * (rp - r) * (2 * rp + 3 * rm - r) / (rp * rp - rm * rm));
    } else if (m.domain == ZR) {
        double rp = m.R(fi, j + 1), rm = m.R(fi, j), r = m.R(fi, fj);
        di = fi - i;
        dj = 1 - (0.5 * (rp - r) * (2 * rp + 3 * rm - r) / (rp * rp - rm * rm));
    }
    f.at(i, j) += (1 - di) * (1 - dj) * val;
    f.at(i + 1, j) += di * (1 - dj) * val;
    f.at(i + 1, j + 1) += di * dj * val;
    f.at(i, j + 1) += (1 - di) * dj * val;
}
double gather_safe(const Field& f, double fi, double fj) {      /* Field2D.java:368-390 */
    int i = (int)fi, j = (int)fj;
    double di = fi - i, dj = fj - j;
    if (i < 0) { i = 0; di = 0; }
    if (j < 0) { j = 0; dj = 0; }
    if (i >= f.ni - 1) { i = f.ni - 1; di = 0; }
    if (j >= f.nj - 1) { j = f.nj - 1; dj = 0; }
    double v = (1 - di) * (1 - dj) * f.at(i, j);
    if (di > 0) v += di * (1 - dj) * f.at(i + 1, j);
    if (di > 0 && dj > 0) v += di * dj * f.at(i + 1, j + 1);
    if (dj > 0) v += (1 - di) * dj * f.at(i, j + 1);
    return v;
}
/* Field2D.gather(double[]) (Field2D.java:300-347): the bilinear form; the Java array
   access throws IndexOutOfBoundsException when i+1/j+1 (or i, j) leave the array, and the
   exception handler falls back to gather_safe. */
double gather(const Field& f, double fi, double fj) {
    int i = (int)fi, j = (int)fj;
    if (i < 0 || j < 0 || i + 1 >= f.ni || j + 1 >= f.nj) return gather_safe(f, fi, fj);
    double di = fi - i, dj = fj - j;
    double v = (1 - di) * (1 - dj) * f.at(i, j);
    v += di * (1 - dj) * f.at(i + 1, j);
    v += di * dj * f.at(i + 1, j + 1);
    v += (1 - di) * dj * f.at(i, j + 1);
    return v;
}

/* row-ordered sparse matrix standing in for core/solver/Matrix.java (HashMap rows) */
struct Matrix {
    std::vector<std::map<int, double>> rows;
    explicit Matrix(int n = 0) : rows(n) {}
    int nr() const { return (int)rows.size(); }
    double get(int i, int j) const {
        auto it = rows[i].find(j);
        return it == rows[i].end() ? 0.0 : it->second;
    }
    void set(int i, int j, double v) { rows[i][j] = v; }
    void add(int i, int j, double v) { set(i, j, get(i, j) + v); }
    void clear_row(int i) { rows[i].clear(); }
    void mult_row(int i, double s) { for (auto& kv : rows[i]) kv.second *= s; }
    void mult(const std::vector<double>& x, std::vector<double>& r) const {   /* Matrix.java:210-223 */
        r.assign(rows.size(), 0.0);
        for (size_t i = 0; i < rows.size(); i++) {
            double p = 0;
            for (auto& kv : rows[i]) p += kv.second * x[kv.first];
            r[i] = p;
        }
    }
    double mult_row_nondiag(const std::vector<double>& x, int i) const {    /* Matrix.java:252-263 */
        double p = 0;
        for (auto& kv : rows[i])
            if (kv.first != i) p += kv.second * x[kv.first];
        return p;
    }
};

double vnorm(const std::vector<double>& v) {   /* Vec.norm (Vec.java:139-145): sqrt(sum v^2)/N */
    double s = 0;
    for (double a : v) s += a * a;
    return std::sqrt(s) / (double)v.size();
}
double vdot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
    return s;
}

struct Species {
    double charge, mass, spwt, qm, nd_back;
    std::vector<Particle> parts;
    std::unique_ptr<std::atomic<long long>[]> bnd_hits;   /* absorbed particles per boundary (deprate_inst) */
    int nbnd = 0;
    int id_counter = 0;
    Field den, U, V, W;
    Field count_sum, u_sum, v_sum, w_sum, uu_sum, vv_sum, ww_sum, mpc_sum;
    Field nd_ave, u_ave, v_ave, w_ave, T, T1, T2, T3, p, mpc;
    int num_samples = 0;
    bool first_time = true;
    double init_T = 0;                    /* Material init_vals.T default */
    double mass_sum = 0, mom_sum[3] = {0, 0, 0}, energy_sum = 0;
};

/* Maxwellian boundary source (sources/MaxwellianSource.java, core/source/Source.java) */
struct Source {
    int sp;
    std::vector<Segment> segs;
    std::vector<double> cum_area;
    double spline_area = 0;
    double mdot, v_drift, v_th, v_az = 0;
    double mp_rem = 0;
    int num_mp = 0;
    int start_it = 0, stop_it = -1;       /* Source.java:68-71 */
};

struct Gradient {                         /* Solver.java:62-74 */
    double Gi[6] = {0, 0, 0, 0, 0, 0}, Gj[6] = {0, 0, 0, 0, 0, 0};
    double R = 0;
    int u[6] = {-1, -1, -1, -1, -1, -1};
};
struct EdgeData {                         /* Solver.java:611-622 */
    double im, jm, ndl_i, ndl_j;
    int ij[4][2];
};

}  // namespace

struct orc_sim {
    Mesh mesh;
    JavaRandom rnd{0};
    int nthreads = 1;
    int nbnd = 0;            /* boundaries of the impact tallies */
    double dt = 1e-7;
    int it = 1;
    double time = 0;
    std::vector<std::unique_ptr<Species>> species;
    std::vector<Source> sources;
    Field phi, rho, efi, efj;
    Field bfi, bfj;               /* DomainModule "bfi"/"bfj" (DomainModule.java:174); zero unless loaded */
    std::atomic<double> boundary_charge{0.0};
    void add_boundary_charge(double q) {
        double cur = boundary_charge.load();
        while (!boundary_charge.compare_exchange_weak(cur, cur + q)) {
        }
    }
    /* solver */
    int method = 0, linear = 0;
    double n0 = 1e15, kTe0 = 1, phi0 = 0, eps = EPS0;
    int lin_max_it = 5000, nl_max_it = 50;
    double lin_tol = 1e-6, nl_tol = 1e-4;
    bool solver_ready = false, solver_first = true, pcg_first_time = true;
    Matrix A, Gi, Gj;
    std::vector<char> fixed;
    int last_lin_it = 0;
    double timers[4] = {0, 0, 0, 0};

    /* ---------------- solver assembly (core/solver/Solver.java:204-726) ---------------- */
    void compute_edge_data(double i0, double j0, EdgeData e[4]) const {      /* Solver.java:634-704 */
        double lcs[4][2];
        mesh.control_volume_lcs(i0, j0, 0.5, lcs);
        const int faces[4] = {RIGHT, TOP, LEFT, BOTTOM};
        for (int fidx = 0; fidx < 4; fidx++) {
            int face = faces[fidx];
            int v1 = 0;
            double im = 0, jm = 0;
            switch (face) {
                case BOTTOM: v1 = 0; im = i0; jm = j0 - 0.5; break;
                case RIGHT: v1 = 1; im = i0 + 0.5; jm = j0; break;
                case TOP: v1 = 2; im = i0; jm = j0 + 0.5; break;
                case LEFT: v1 = 3; im = i0 - 0.5; jm = j0; break;
            }
            int v2 = v1 + 1;
            if (v2 > 3) v2 = 0;
            if (im < lcs[0][0]) im = lcs[0][0];
            if (im < lcs[3][0]) im = lcs[3][0];
            if (im > lcs[1][0]) im = lcs[1][0];
            if (im > lcs[2][0]) im = lcs[2][0];
            if (jm < lcs[0][1]) jm = lcs[0][1];
            if (jm < lcs[1][1]) jm = lcs[1][1];
            if (jm > lcs[2][1]) jm = lcs[2][1];
            if (jm > lcs[3][1]) jm = lcs[3][1];
            EdgeData& d = e[face];
            double x1[2], x2[2];
            mesh.pos(lcs[v1][0], lcs[v1][1], x1);
            mesh.pos(lcs[v2][0], lcs[v2][1], x2);
            d.ndl_i = x2[1] - x1[1];
            d.ndl_j = -(x2[0] - x1[0]);
            mesh.node_index_offset(im, jm, -0.5, -0.5, d.ij[0]);
            mesh.node_index_offset(im, jm, 0.5, -0.5, d.ij[1]);
            mesh.node_index_offset(im, jm, 0.5, 0.5, d.ij[2]);
            mesh.node_index_offset(im, jm, -0.5, 0.5, d.ij[3]);
            d.im = im;
            d.jm = jm;
        }
    }
    Gradient node_gradient(double i, double j) const {                      /* Solver.java:560-609 */
        Gradient G;
        EdgeData ed[4];
        compute_edge_data(i, j, ed);
        for (int f = 0; f < 4; f++) {
            const EdgeData& e = ed[f];
            for (int n = 0; n < 4; n++) {
                for (int k = 0; k < 6; k++) {
                    int u = mesh.IJtoN(e.ij[n][0], e.ij[n][1]);
                    if (u == G.u[k] || G.u[k] < 0) {
                        G.u[k] = u;
                        G.Gi[k] += 0.25 * e.ndl_i;
                        G.Gj[k] += 0.25 * e.ndl_j;
                        break;
                    }
                }
            }
        }
        double Ar = mesh.area(i, j);
        for (int k = 0; k < 6; k++) { G.Gj[k] *= 1 / Ar; G.Gi[k] *= 1 / Ar; }
        G.R = mesh.R(i, j);
        return G;
    }
    void set_node_fvm(int i, int j) {                                         /* Solver.java:451-553 */
        const int ni = mesh.ni, nj = mesh.nj;
        int u = mesh.IJtoN(i, j);
        if (mesh.is_dirichlet(i, j)) {
            A.clear_row(u);
            A.set(u, u, 1);
            fixed[u] = 1;
            return;
        }
        if (mesh.ntype(i, j) == NT_INSULATOR) {
            /* insulator nodes need Node.surf_normal, which this oracle does not carry */
            A.rows[u] = Gi.rows[u];
            return;
        }
        if ((i == 0 && mesh.btype(LEFT, j) != DB_MESH) || (i == ni - 1 && mesh.btype(RIGHT, j) != DB_MESH)) {
            A.rows[u] = Gi.rows[u];                   /* copyRow(Gi) */
            return;
        }
        if ((j == 0 && mesh.btype(BOTTOM, i) != DB_MESH) || (j == nj - 1 && mesh.btype(TOP, i) != DB_MESH)) {
            A.rows[u] = Gj.rows[u];                   /* copyRow(Gj) */
            return;
        }
        EdgeData e[4];
        compute_edge_data(i, j, e);
        for (int f = 0; f < 4; f++) {
            Gradient G = node_gradient(e[f].im, e[f].jm);
            double R = G.R;
            for (int k = 0; k < 6; k++) { G.Gi[k] *= e[f].ndl_i * R; G.Gj[k] *= e[f].ndl_j * R; }
            for (int v = 0; v < 6 && G.u[v] >= 0; v++) {
                double val = G.Gj[v] + G.Gi[v];
                if (val != 0) A.add(u, G.u[v], val);
            }
        }
        double Vn = mesh.node_volume(i, j);
        A.mult_row(u, 1.0 / Vn);
        mesh.bc(i, j) /= Vn;
    }
    void init_solver() {                                                      /* Solver.init, initCoefficients */
        const int ni = mesh.ni, nj = mesh.nj, nu = ni * nj;
        A = Matrix(nu);
        Gi = Matrix(nu);
        Gj = Matrix(nu);
        fixed.assign(nu, 0);
        for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) {
                Gradient G = node_gradient(i, j);
                int u = mesh.IJtoN(i, j);
                for (int v = 0; v < 6 && G.u[v] >= 0; v++) {
                    if (G.Gi[v] != 0) Gi.add(u, G.u[v], G.Gi[v]);
                    if (G.Gj[v] != 0) Gj.add(u, G.u[v], G.Gj[v]);
                }
            }
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) set_node_fvm(i, j);
        solver_ready = true;
    }

    /* ---------------- linear solvers ---------------- */
    /* LinearSolverPCG.solve (core/solver/LinearSolverPCG.java:21-122), single mesh */
    int pcg(const Matrix& Am, std::vector<double>& x, const std::vector<double>& b, int max_it, double tol) {
        double norm = tol;
        if (pcg_first_time) x = b;
        int n = Am.nr();
        std::vector<double> Mi(n), r(n), z(n), p(n), Ax(n), Ap(n);
        for (int i = 0; i < n; i++) Mi[i] = 1.0 / Am.get(i, i);
        Am.mult(x, Ax);
        for (int i = 0; i < n; i++) r[i] = b[i] - Ax[i];
        for (int i = 0; i < n; i++) z[i] = Mi[i] * r[i];
        p = z;
        pcg_first_time = false;
        int it = 1;
        while (it <= max_it) {
            norm = 0;
            double norm_m = vnorm(r);
            if (norm_m > norm) norm = norm_m;
            if (!(norm < tol)) {
                Am.mult(p, Ap);
                double alpha = vdot(r, z) / vdot(p, Ap);
                for (int i = 0; i < n; i++) x[i] += p[i] * alpha;
                double zr_dot = vdot(z, r);
                Am.mult(p, Ap);                        /* A*p is evaluated a second time, as in the reference */
                for (int i = 0; i < n; i++) r[i] -= Ap[i] * alpha;
                for (int i = 0; i < n; i++) z[i] = Mi[i] * r[i];
                double beta = vdot(z, r) / zr_dot;
                for (int i = 0; i < n; i++) p[i] = z[i] + p[i] * beta;
            }
            if (norm < tol) break;
            it++;
        }
        it--;
        return it;
    }
    /* LinearSolverGS.solve (core/solver/LinearSolverGS.java:47-152), SOR w=1.4, chunks run in order */
    int gs(const Matrix& Am, std::vector<double>& x, const std::vector<double>& b, int max_it, double tol) {
        int n = Am.nr();
        std::vector<double> diag(n), Ax(n);
        for (int i = 0; i < n; i++) diag[i] = Am.get(i, i);
        int it = 1;
        double norm = 1e66;
        while (it <= max_it) {
            if (it % 25 == 0) {
                Am.mult(x, Ax);                        /* Solver.calculateResidue */
                std::vector<double> res(n);
                for (int i = 0; i < n; i++) res[i] = Ax[i] - b[i];
                norm = vnorm(res);
                norm /= n;
                if (norm < tol) break;
            }
            for (int u = 0; u < n; u++) {
                double tau = Am.mult_row_nondiag(x, u);
                if (std::fabs(diag[u]) < 1e-8) continue;
                double g = (b[u] - 0.0 - tau) / diag[u];
                x[u] = x[u] + 1.4 * (g - x[u]);
            }
            it++;
        }
        it--;
        return it;
    }
    int lin_solve(const Matrix& Am, std::vector<double>& x, const std::vector<double>& b, int max_it, double tol) {
        return method == 1 ? pcg(Am, x, b, max_it, tol) : gs(Am, x, b, max_it, tol);
    }

    /* Solver.solveNonLinearNR (Solver.java:293-387) with PoissonSolver's Boltzmann term (PoissonSolver.java:221-252) */
    int solve_nonlinear(std::vector<double>& x, const std::vector<double>& b0) {
        const int ni = mesh.ni, nj = mesh.nj, n = ni * nj;
        const double C = QE / eps;
        std::vector<double> y(n, 0.0), bx(n), P(n), Ax(n), F(n);
        int it;
        double norm = -1;
        for (it = 0; it < nl_max_it; it++) {
            for (int k = 0; k < n; k++) {
                bx[k] = fixed[k] ? 0 : C * n0 * std::exp((x[k] - phi0) / kTe0);
                P[k] = fixed[k] ? 0 : C * n0 * std::exp((x[k] - phi0) / kTe0) / kTe0;
            }
            std::vector<double> b(n);
            for (int k = 0; k < n; k++) b[k] = b0[k] + bx[k];
            for (int i = 0; i < ni; i++)
                for (int j = 0; j < nj; j++) {
                    int u = mesh.IJtoN(i, j);
                    if (mesh.is_dirichlet(i, j)) continue;
                    if ((i == 0 && mesh.btype(LEFT, j) != DB_MESH && mesh.btype(LEFT, j) != DB_DIRICHLET) ||
                        (i == ni - 1 && mesh.btype(RIGHT, j) != DB_MESH && mesh.btype(RIGHT, j) != DB_DIRICHLET) ||
                        (j == 0 && mesh.btype(BOTTOM, i) != DB_MESH && mesh.btype(BOTTOM, i) != DB_DIRICHLET) ||
                        (j == nj - 1 && mesh.btype(TOP, i) != DB_MESH && mesh.btype(TOP, i) != DB_DIRICHLET)) {
                        b[u] = 0;
                        P[u] = 0;
                    }
                }
            A.mult(x, Ax);
            for (int k = 0; k < n; k++) F[k] = Ax[k] - b[k];
            Matrix J = A;                               /* subtractDiag copies the matrix */
            for (int k = 0; k < n; k++) J.set(k, k, A.get(k, k) - P[k]);
            last_lin_it = lin_solve(J, y, F, lin_max_it, lin_tol);
            for (int k = 0; k < n; k++) x[k] -= y[k];
            norm = vnorm(y);
            if (norm < nl_tol) break;
        }
        return it;
    }

    /* PoissonSolver.update (pic/PoissonSolver.java:110-214), qn_switch off, den0_pos unused */
    int solve_potential() {
        if (!solver_ready) init_solver();
        const int ni = mesh.ni, nj = mesh.nj, n = ni * nj;
        std::vector<double> x(phi.d), b(n);
        for (int k = 0; k < n; k++) b[k] = rho.d[k] * (-1 / eps);
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) {
                if (mesh.is_dirichlet(i, j)) { b[mesh.IJtoN(i, j)] = mesh.bc(i, j); continue; }
                if (mesh.ntype(i, j) == NT_INSULATOR) { b[mesh.IJtoN(i, j)] = 0; continue; }
                if (i != 0 && i != ni - 1 && j != 0 && j != nj - 1) continue;
                if (i == 0 && mesh.btype(LEFT, j) != DB_MESH) b[mesh.IJtoN(0, j)] = mesh.bc(0, j);
                if (i == ni - 1 && mesh.btype(RIGHT, j) != DB_MESH) b[mesh.IJtoN(ni - 1, j)] = mesh.bc(ni - 1, j);
                if (j == 0 && mesh.btype(BOTTOM, i) != DB_MESH) b[mesh.IJtoN(i, 0)] = mesh.bc(i, 0);
                /* the reference compares j with ni-1 here (PoissonSolver.java:152) */
                if (j == ni - 1 && mesh.btype(TOP, i) != DB_MESH) b[mesh.IJtoN(i, nj - 1)] = mesh.bc(i, nj - 1);
            }
        int itn;
        if (linear) itn = last_lin_it = lin_solve(A, x, b, lin_max_it, lin_tol);
        else itn = solve_nonlinear(x, b);
        phi.d = x;
        return itn;
    }
    /* PotentialSolver.updateGradientField + Solver.evaluateGradient (scale -1) */
    void update_efield() {
        if (!solver_ready) init_solver();
        std::vector<double> gi, gj;
        Gi.mult(phi.d, gi);
        Gj.mult(phi.d, gj);
        for (size_t k = 0; k < gi.size(); k++) { efi.d[k] = -gi[k]; efj.d[k] = -gj[k]; }
    }
    /* SolverModule.updateRho (core/solver/SolverModule.java:144-171) */
    void update_rho() {
        rho.clear();
        for (auto& sp : species) {
            if (sp->charge == 0) continue;
            for (size_t k = 0; k < rho.d.size(); k++) rho.d[k] += sp->den.d[k] * sp->charge;
        }
    }

    /* ---------------- particles ---------------- */
    /* UpdateVelocityBoris (KineticMaterial.java:847-893) with E = (Ei, Ej, 0) and B = (Bi, Bj, 0):
       half electric kick, rotation by t = (q/m) B dt/2 and s = 2t/(1+|t|^2), half kick */
    static void boris(double qm, double dtp, const double E[3], const double B[3], double vel[3]) {
        double t[3], s[3], vm[3], vp[3], c[3];
        for (int d = 0; d < 3; d++) t[d] = qm * B[d] * 0.5 * dtp;
        double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        for (int d = 0; d < 3; d++) s[d] = 2 * t[d] / (1 + t2);
        for (int d = 0; d < 3; d++) vm[d] = vel[d] + qm * E[d] * 0.5 * dtp;
        c[0] = vm[1] * t[2] - vm[2] * t[1];
        c[1] = -vm[0] * t[2] + vm[2] * t[0];
        c[2] = vm[0] * t[1] - vm[1] * t[0];
        for (int d = 0; d < 3; d++) vp[d] = vm[d] + c[d];
        c[0] = vp[1] * s[2] - vp[2] * s[1];
        c[1] = -vp[0] * s[2] + vp[2] * s[0];
        c[2] = vp[0] * s[1] - vp[1] * s[0];
        for (int d = 0; d < 3; d++) vel[d] = vm[d] + c[d] + qm * E[d] * 0.5 * dtp;
    }
    /* velocity update of ParticleMover.run and addParticle (KineticMaterial.java:336-353, 779-794):
       the electric kick over part.dt when the gathered B is zero, Boris otherwise */
    void update_velocity(const Species& sp, Particle& part) {
        double E[3] = {gather(efi, part.lc[0], part.lc[1]), gather(efj, part.lc[0], part.lc[1]), 0};
        double B[3] = {gather(bfi, part.lc[0], part.lc[1]), gather(bfj, part.lc[0], part.lc[1]), 0};
        if (B[0] == 0 && B[1] == 0) {
            part.vel[0] += sp.qm * E[0] * part.dt;
            part.vel[1] += sp.qm * E[1] * part.dt;
        } else {
            boris(sp.qm, part.dt, E, B, part.vel);
        }
    }

    /* KineticMaterial.addParticle(MeshData, Particle) (KineticMaterial.java:759-802) */
    void add_particle(Species& sp, Particle part, bool have_lc) {
        if (!have_lc) {
            mesh.XtoL(part.pos, part.lc);
            if (part.lc[0] >= mesh.ni) part.lc[0] = mesh.ni - 1;
            if (part.lc[1] >= mesh.nj) part.lc[1] = mesh.nj - 1;
        }
        part.dt = -0.5 * dt;
        update_velocity(sp, part);
        part.dt = 0;
        part.id = sp.id_counter++;
        /* MeshData.addParticle drops non-finite velocities (KineticMaterial.java:1356-1361) */
        if (!(std::isfinite(part.vel[0]) && std::isfinite(part.vel[1]) && std::isfinite(part.vel[2]))) return;
        sp.parts.push_back(part);
    }

    /* KineticMaterial.ProcessBoundary (KineticMaterial.java:471-750) */
    bool process_boundary(Species& sp, Particle& part, const double* old, const double* lc_old) {
        const int ni = mesh.ni, nj = mesh.nj;
        bool alive = true;
        double dt0 = part.dt;
        part.dt = 0;
        double lcmin0 = std::min(part.lc[0], lc_old[0]), lcmin1 = std::min(part.lc[1], lc_old[1]);
        double lcmax0 = std::max(part.lc[0], lc_old[0]), lcmax1 = std::max(part.lc[1], lc_old[1]);
        int i_min = (int)lcmin0, i_max = (int)lcmax0, j_min = (int)lcmin1, j_max = (int)lcmax1;
        if (i_min < 0) i_min = 0;
        if (j_min < 0) j_min = 0;
        if (i_max >= ni) i_max = ni - 1;
        if (j_max >= nj) j_max = nj - 1;
        /* the HashSet of candidate segments: duplicates give identical t and never win a strict '<' */
        double tp_min = 2.0, tsurf_min = 0;
        int seg_min = -1;
        for (int i = i_min; i <= i_max; i++)
            for (int j = j_min; j <= j_max; j++) {
                int nidx = j * ni + i;
                for (int q = mesh.nseg_ptr[nidx]; q < mesh.nseg_ptr[nidx + 1]; q++) {
                    int s = mesh.nseg_idx[q];
                    const Segment& seg = mesh.seg[s];
                    if (!(seg.btype == BT_DIRICHLET || seg.btype == BT_SINK)) continue;
                    double t[2];
                    seg.intersect(old, part.pos, t);
                    double t_part = t[1];
                    if (t_part > 0) {
                        double vm = std::sqrt(part.vel[0] * part.vel[0] + part.vel[1] * part.vel[1]);
                        double acos = (seg.normal[0] * part.vel[0] + seg.normal[1] * part.vel[1]) / vm;
                        if (t_part < FLT_EPS && (acos > 0 || seg.btype == BT_VIRTUAL)) continue;
                        if (t_part < tp_min) { tp_min = t_part; tsurf_min = t[0]; seg_min = s; }
                    }
                }
            }
        (void)tsurf_min;
        if (seg_min >= 0) {
            tp_min *= 0.9999;
            part.pos[0] = old[0] + tp_min * (part.pos[0] - old[0]);
            part.pos[1] = old[1] + tp_min * (part.pos[1] - old[1]);
            mesh.XtoL(part.pos, part.lc);
            part.dt = dt0 * (1 - tp_min);
            if (part.lc[0] < 0 && part.lc[0] > -FLT_EPS) part.lc[0] = 0;
            if (part.lc[1] < 0 && part.lc[1] > -FLT_EPS) part.lc[1] = 0;
            const Segment& sg = mesh.seg[seg_min];
            if (sg.absorb) alive = false;              /* Material.performSurfaceInteraction default: absorb */
            if (!alive) add_boundary_charge(part.mpw * sp.charge);
            if (sg.btype == BT_SINK) alive = false;
            /* addSurfaceMassDeposit on the boundary hit (KineticMaterial.java:596-600); counted in
               particles, the caller multiplies by spwt */
            if (!alive && sg.bnd >= 0 && sg.bnd < sp.nbnd) sp.bnd_hits[sg.bnd].fetch_add(1);
            if (!alive) return alive;
        }
        if (part.lc[0] < 0 || part.lc[1] < 0 || part.lc[0] >= ni - 1 || part.lc[1] >= nj - 1) {
            double t_right = 99, t_top = 99, t_left = 99, t_bottom = 99;
            if (part.lc[0] >= ni - 1) t_right = (ni - 1.0 - lc_old[0]) / (part.lc[0] - lc_old[0]);
            if (part.lc[1] >= nj - 1) t_top = (nj - 1.0 - lc_old[1]) / (part.lc[1] - lc_old[1]);
            if (part.lc[0] < 0) t_left = lc_old[0] / (lc_old[0] - part.lc[0]);
            if (part.lc[1] < 0) t_bottom = lc_old[1] / (lc_old[1] - part.lc[1]);
            int exit_face = RIGHT;
            double t = t_right;
            if (t_top < t) { exit_face = TOP; t = t_top; }
            if (t_left < t) { exit_face = LEFT; t = t_left; }
            if (t_bottom < t) { exit_face = BOTTOM; t = t_bottom; }
            part.lc[0] = lc_old[0] + t * (part.lc[0] - lc_old[0]);
            part.lc[1] = lc_old[1] + t * (part.lc[1] - lc_old[1]);
            if (part.lc[0] < 0) part.lc[0] = 0;
            else if (part.lc[0] > ni - 1) part.lc[0] = ni - 1;
            if (part.lc[1] < 0) part.lc[1] = 0;
            else if (part.lc[1] > nj - 1) part.lc[1] = nj - 1;
            double x[2];
            mesh.pos(part.lc[0], part.lc[1], x);
            part.pos[0] = x[0];
            part.pos[1] = x[1];
            part.dt = dt0 * (1 - t);
            int i = (int)part.lc[0], j = (int)part.lc[1];
            if (exit_face == TOP) j++;
            if (exit_face == RIGHT) i++;
            if (i < 0) i = 0;
            if (j < 0) j = 0;
            if (i >= ni - 1) i = ni - 1;
            if (j >= nj - 1) j = nj - 1;
            int type = (exit_face == LEFT || exit_face == RIGHT) ? mesh.btype(exit_face, j) : mesh.btype(exit_face, i);
            switch (type) {
                case DB_OPEN: return false;
                case DB_SYMMETRY: {
                    double nrm[3] = {0, 0, 0};                   /* UniformMesh.faceNormal */
                    if (exit_face == LEFT) nrm[0] = 1;
                    else if (exit_face == RIGHT) nrm[0] = -1;
                    else if (exit_face == BOTTOM) nrm[1] = 1;
                    else nrm[1] = -1;
                    double tm = part.vel[0] * nrm[0] + part.vel[1] * nrm[1] + part.vel[2] * nrm[2];
                    for (int d = 0; d < 3; d++) {              /* Vec.mirror (Vec.java:406-419) */
                        double tc = nrm[d] * tm;
                        double nc = part.vel[d] - tc;
                        part.vel[d] = -tc + nc;
                    }
                    return true;
                }
                case DB_PERIODIC:
                    if (exit_face == LEFT) part.pos[0] += (mesh.xd[0] - mesh.x0[0]);
                    else if (exit_face == RIGHT) part.pos[0] -= (mesh.xd[0] - mesh.x0[0]);
                    else if (exit_face == BOTTOM) part.pos[1] += (mesh.xd[1] - mesh.x0[1]);
                    else part.pos[1] -= (mesh.xd[1] - mesh.x0[1]);
                    return true;
                case DB_CIRCUIT: {
                    if (sp.charge >= 0) return false;
                    if (boundary_charge.load() / (-sp.charge) > part.mpw) {
                        add_boundary_charge(part.mpw * sp.charge);
                        return false;
                    }
                    for (int d = 0; d < 3; d++) part.vel[d] = -part.vel[d];
                    return true;
                }
                default: return false;                   /* DIRICHLET, MESH (no neighbour), SINK, ... */
            }
        }
        return true;
    }

    /* ParticleMover.run (KineticMaterial.java:299-422) over particles [p0,p1); returns alive flags */
    void move_range(Species& sp, size_t p0, size_t p1, std::vector<char>& alive_flag, double* sums) {
        const int max_bounces = 10;
        double N_sum = 0, P_sum[3] = {0, 0, 0}, E_sum = 0;
        for (size_t p = p0; p < p1; p++) {
            Particle& part = sp.parts[p];
            alive_flag[p] = 1;
            if (part.mpw <= 0) { alive_flag[p] = 0; continue; }
            part.dt += dt;
            update_velocity(sp, part);
            int bounces = 0;
            bool alive = true;
            while (part.dt > 0 && bounces++ < max_bounces) {
                double old[2] = {part.pos[0], part.pos[1]};
                double old_lc[2] = {part.lc[0], part.lc[1]};
                part.pos[0] += part.vel[0] * part.dt;
                part.pos[1] += part.vel[1] * part.dt;
                if (mesh.domain == RZ) {                 /* rotateToRZ (KineticMaterial.java:424-442) */
                    double Aa = part.vel[2] * part.dt, B = part.pos[0];
                    double Rr = std::sqrt(Aa * Aa + B * B);
                    double cs = B / Rr, sn = Aa / Rr;
                    part.pos[2] -= std::asin(sn);
                    double v1 = part.vel[0], v2 = part.vel[2];
                    part.pos[0] = Rr;
                    part.vel[0] = cs * v1 + sn * v2;
                    part.vel[2] = -sn * v1 + cs * v2;
                } else if (mesh.domain == ZR) {          /* rotateToZR (KineticMaterial.java:444-462) */
                    double Aa = part.vel[2] * part.dt, B = part.pos[1];
                    double Rr = std::sqrt(Aa * Aa + B * B);
                    double cs = B / Rr, sn = Aa / Rr;
                    part.pos[2] += std::acos(cs);
                    double v1 = part.vel[1], v2 = part.vel[2];
                    part.pos[1] = Rr;
                    part.vel[1] = cs * v1 + sn * v2;
                    part.vel[2] = -sn * v1 + cs * v2;
                } else {
                    part.pos[2] += part.vel[2] * part.dt;
                }
                mesh.XtoL(part.pos, part.lc);
                alive = process_boundary(sp, part, old, old_lc);
                if (!alive) { alive_flag[p] = 0; break; }
            }
            if (alive) {
                N_sum += part.mpw;
                P_sum[0] += part.mpw * part.vel[0];
                P_sum[1] += part.mpw * part.vel[1];
                P_sum[2] += part.mpw * part.vel[2];
                E_sum += part.mpw * std::sqrt(part.vel[0] * part.vel[0] + part.vel[1] * part.vel[1] +
                                              part.vel[2] * part.vel[2]);   /* Vec.mag3 */
            }
        }
        sums[0] = N_sum; sums[1] = P_sum[0]; sums[2] = P_sum[1]; sums[3] = P_sum[2]; sums[4] = E_sum;
    }
    /* KineticMaterial.moveParticles(false) (KineticMaterial.java:200-261): one mover per block */
    void move_species(Species& sp) {
        size_t np = sp.parts.size();
        std::vector<char> alive(np, 1);
        int nb = std::max(1, nthreads);
        std::vector<double> sums((size_t)nb * 5, 0.0);
        if (nb == 1 || np < 1000) {
            move_range(sp, 0, np, alive, &sums[0]);
            for (int b = 1; b < nb; b++) for (int q = 0; q < 5; q++) sums[b * 5 + q] = 0;
        } else {
            std::vector<std::thread> th;
            for (int b = 0; b < nb; b++) {
                size_t p0 = np * b / nb, p1 = np * (b + 1) / nb;
                th.emplace_back([this, &sp, p0, p1, &alive, &sums, b]() { move_range(sp, p0, p1, alive, &sums[b * 5]); });
            }
            for (auto& t : th) t.join();
        }
        sp.mass_sum = 0; sp.mom_sum[0] = sp.mom_sum[1] = sp.mom_sum[2] = 0; sp.energy_sum = 0;
        for (int b = 0; b < nb; b++) {
            sp.mass_sum += sums[b * 5 + 0] * sp.mass;
            sp.mom_sum[0] += sums[b * 5 + 1] * sp.mass;
            sp.mom_sum[1] += sums[b * 5 + 2] * sp.mass;
            sp.mom_sum[2] += sums[b * 5 + 3] * sp.mass;
            sp.energy_sum += sums[b * 5 + 4] * sp.mass;
        }
        size_t w = 0;                                       /* iterator.remove() of dead particles */
        for (size_t p = 0; p < np; p++)
            if (alive[p]) sp.parts[w++] = sp.parts[p];
        sp.parts.resize(w);
    }

    /* KineticMaterial.updateFields(MeshData) (KineticMaterial.java:168-197) */
    void deposit(Species& sp) {
        sp.den.clear(); sp.U.clear(); sp.V.clear(); sp.W.clear();
        for (const Particle& p : sp.parts) {
            scatter(mesh, sp.den, p.lc[0], p.lc[1], p.mpw);
            scatter(mesh, sp.U, p.lc[0], p.lc[1], p.vel[0] * p.mpw);
            scatter(mesh, sp.V, p.lc[0], p.lc[1], p.vel[1] * p.mpw);
            scatter(mesh, sp.W, p.lc[0], p.lc[1], p.vel[2] * p.mpw);
        }
        divide_by(sp.U, sp.den);
        divide_by(sp.V, sp.den);
        divide_by(sp.W, sp.den);
        scale_by_vol(sp.den);
    }
    static void divide_by(Field& f, const Field& g) {     /* Field2D.divideByField (Field2D.java:403-414) */
        for (size_t k = 0; k < f.d.size(); k++) f.d[k] = (g.d[k] != 0) ? f.d[k] / g.d[k] : 0;
    }
    void scale_by_vol(Field& f) {                         /* Field2D.scaleByVol (Field2D.java:418-431) */
        for (size_t k = 0; k < f.d.size(); k++) f.d[k] /= mesh.node_vol.d[k];
    }
    /* KineticMaterial.updateSamples (KineticMaterial.java:1570-1595) */
    void update_samples(Species& sp) {
        for (const Particle& p : sp.parts) {
            scatter(mesh, sp.u_sum, p.lc[0], p.lc[1], p.mpw * p.vel[0]);
            scatter(mesh, sp.v_sum, p.lc[0], p.lc[1], p.mpw * p.vel[1]);
            scatter(mesh, sp.w_sum, p.lc[0], p.lc[1], p.mpw * p.vel[2]);
            scatter(mesh, sp.uu_sum, p.lc[0], p.lc[1], p.mpw * p.vel[0] * p.vel[0]);
            scatter(mesh, sp.vv_sum, p.lc[0], p.lc[1], p.mpw * p.vel[1] * p.vel[1]);
            scatter(mesh, sp.ww_sum, p.lc[0], p.lc[1], p.mpw * p.vel[2] * p.vel[2]);
            scatter(mesh, sp.count_sum, p.lc[0], p.lc[1], p.mpw);
            sp.mpc_sum.at((int)p.lc[0], (int)p.lc[1]) += 1;
        }
    }
    /* KineticMaterial.computeFields(MeshData) (KineticMaterial.java:1606-1671) */
    void compute_fields(Species& sp) {
        for (size_t k = 0; k < sp.count_sum.d.size(); k++) {
            double count = sp.count_sum.d[k];
            if (count > 0) {
                double u = sp.u_sum.d[k] / count, v = sp.v_sum.d[k] / count, w = sp.w_sum.d[k] / count;
                double uu = sp.uu_sum.d[k] / count - u * u;
                double vv = sp.vv_sum.d[k] / count - v * v;
                double ww = sp.ww_sum.d[k] / count - w * w;
                double f = sp.mass / K_B;
                sp.T.d[k] = (uu + vv + ww) * f / 3.0;
                sp.T1.d[k] = uu * f;
                sp.T2.d[k] = vv * f;
                sp.T3.d[k] = ww * f;
            } else {
                sp.T.d[k] = sp.T1.d[k] = sp.T2.d[k] = sp.T3.d[k] = sp.init_T;
            }
        }
        sp.nd_ave.d = sp.count_sum.d;
        for (double& a : sp.nd_ave.d) a *= 1.0 / sp.num_samples;
        scale_by_vol(sp.nd_ave);
        sp.u_ave.d = sp.u_sum.d; divide_by(sp.u_ave, sp.count_sum);
        sp.v_ave.d = sp.v_sum.d; divide_by(sp.v_ave, sp.count_sum);
        sp.w_ave.d = sp.w_sum.d; divide_by(sp.w_ave, sp.count_sum);
        for (size_t k = 0; k < sp.p.d.size(); k++) sp.p.d[k] = sp.nd_ave.d[k] * K_B * sp.T.d[k];
        sp.mpc.d = sp.mpc_sum.d;
        for (double& a : sp.mpc.d) a *= 1. / sp.num_samples;
    }
    /* Material.update -> KineticMaterial.updateFields (KineticMaterial.java:117-163, Material.java:618-632),
       after the move */
    void update_species_fields(Species& sp) {
        deposit(sp);
        update_samples(sp);
        sp.num_samples++;
        if (sp.first_time || it % 10 == 0) compute_fields(sp);
        sp.first_time = false;
        if (sp.nd_back != 0)
            for (double& a : sp.den.d) a += sp.nd_back;
    }

    /* ---------------- sources ---------------- */
    double randomT(const Source& s) {                     /* Spline.randomT (Spline.java:597-650), XY */
        double A1 = rnd.next_double() * s.spline_area;
        int i;
        const std::vector<double>& v = s.cum_area;        /* Vec.binarySearch (Vec.java:529-546) */
        if (A1 < v[0]) i = -1;
        else if (A1 > v.back()) i = (int)v.size();
        else {
            int i1 = 0, i2 = (int)v.size();
            while (true) {
                int mid = (int)(0.5 * (i1 + i2));
                if (A1 < v[mid]) i2 = mid;
                else if (A1 > v[mid]) i1 = mid;
                else { i1 = mid; break; }
                if ((i2 - i1) <= 1) break;
            }
            i = i1;
        }
        double seg_area = s.segs[i].length;
        double frac = (A1 - s.cum_area[i]) / seg_area;
        return i + frac;
    }
    double sample_maxw1d() {                              /* Utils.SampleMaxw1D (Utils.java:27-37) */
        double sum = 0;
        for (int i = 0; i < 12; i++) sum += rnd.next_double();
        double fm = (sum - 0.5 * 12) * std::sqrt(12.0 / 12);
        return std::sqrt(0.5) * 1.0 * fm;
    }
    void sample_source(Source& s) {
        Species& sp = *species[s.sp];
        double mp = (s.mdot * dt) / (sp.mass * sp.spwt) + s.mp_rem;  /* Source.regenerate (Source.java:123-134) */
        s.num_mp = (int)mp;
        s.mp_rem = mp - s.num_mp;
        while (s.num_mp > 0) {                            /* Source.sampleKinetic (Source.java:167-198) */
            Particle part{};
            part.mpw = sp.spwt;
            double t = randomT(s);
            int si = (int)t;
            double seg_t = t - si;
            if (si > (int)s.segs.size() - 1) { si = (int)s.segs.size() - 1; seg_t = 1.0; }
            double x[2];
            s.segs[si].pos(seg_t, x);
            const double* n = s.segs[std::min(si, (int)s.segs.size() - 1)].normal;
            part.pos[0] = x[0];
            part.pos[1] = x[1];
            part.pos[2] = 0;
            if (mesh.domain != XY) part.pos[2] = rnd.next_double() * 2 * PI;
            do {                                          /* MaxwellianSource.sampleParticle (93-125) */
                double vm[3];
                vm[0] = (s.v_th * 1.2) * sample_maxw1d();
                vm[1] = (s.v_th * 1.2) * sample_maxw1d();
                vm[2] = (s.v_th * 1.2) * sample_maxw1d();
                part.vel[0] = vm[0] + n[0] * s.v_drift;
                part.vel[1] = vm[1] + n[1] * s.v_drift;
                part.vel[2] = vm[2] + s.v_az;
                part.dt = rnd.next_double() * dt;
            } while (n[0] * part.vel[0] + n[1] * part.vel[1] <= 0);
            s.num_mp--;
            for (int d = 0; d < 3; d++) part.pos[d] += part.vel[d] * 1e-6 * dt;
            /* KineticMaterial.addParticle(Particle): DomainModule.getMesh (DomainModule.java:106-117) */
            if (!(mesh.contains_strict(part.pos) || mesh.contains(part.pos))) continue;
            add_particle(sp, part, false);
        }
    }

    /* SourceModule.sampleSources (SourceModule.java:231-283): sources outside their window skip */
    void sample_sources() {
        for (auto& s : sources) {
            if (it < s.start_it || (s.stop_it >= 0 && it > s.stop_it)) continue;
            sample_source(s);
        }
    }
    /* Starfish.MainLoop body (core/common/Starfish.java:77-121) */
    void step() {
        using clk = std::chrono::steady_clock;
        auto t0 = clk::now();
        sample_sources();
        auto t1 = clk::now();
        for (auto& sp : species) move_species(*sp);
        auto t2 = clk::now();
        for (auto& sp : species) update_species_fields(*sp);
        auto t3 = clk::now();
        update_rho();
        solve_potential();
        update_efield();
        auto t4 = clk::now();
        timers[0] += std::chrono::duration<double>(t1 - t0).count();
        timers[1] += std::chrono::duration<double>(t2 - t1).count();
        timers[2] += std::chrono::duration<double>(t3 - t2).count();
        timers[3] += std::chrono::duration<double>(t4 - t3).count();
        it++;
        time += dt;
    }

    Field* species_field(Species& sp, const std::string& n) {
        if (n == "nd") return &sp.den;
        if (n == "u") return &sp.U;
        if (n == "v") return &sp.V;
        if (n == "w") return &sp.W;
        if (n == "count-sum") return &sp.count_sum;
        if (n == "u-sum") return &sp.u_sum;
        if (n == "v-sum") return &sp.v_sum;
        if (n == "w-sum") return &sp.w_sum;
        if (n == "uu-sum") return &sp.uu_sum;
        if (n == "vv-sum") return &sp.vv_sum;
        if (n == "ww-sum") return &sp.ww_sum;
        if (n == "mpc-sum") return &sp.mpc_sum;
        if (n == "t") return &sp.T;
        if (n == "t1") return &sp.T1;
        if (n == "t2") return &sp.T2;
        if (n == "t3") return &sp.T3;
        if (n == "nd-ave") return &sp.nd_ave;
        if (n == "u-ave") return &sp.u_ave;
        if (n == "v-ave") return &sp.v_ave;
        if (n == "w-ave") return &sp.w_ave;
        if (n == "p") return &sp.p;
        if (n == "mpc") return &sp.mpc;
        return nullptr;
    }
    Field* field(int spi, const std::string& n) {
        if (n == "phi") return &phi;
        if (n == "rho") return &rho;
        if (n == "efi") return &efi;
        if (n == "efj") return &efj;
        if (n == "bfi") return &bfi;
        if (n == "bfj") return &bfj;
        if (spi < 0 || spi >= (int)species.size()) return nullptr;
        return species_field(*species[spi], n);
    }
};

extern "C" {

orc_sim* orc_create(const orc_mesh_desc* d) {
    orc_sim* s = new orc_sim();
    Mesh& m = s->mesh;
    m.domain = d->domain;
    m.ni = d->ni;
    m.nj = d->nj;
    m.x0[0] = d->x0[0]; m.x0[1] = d->x0[1];
    m.dh[0] = d->dh[0]; m.dh[1] = d->dh[1];
    m.xd[0] = m.x0[0] + (m.ni - 1) * m.dh[0];      /* UniformMesh.setXd */
    m.xd[1] = m.x0[1] + (m.nj - 1) * m.dh[1];
    const int len[4] = {m.nj, m.ni, m.nj, m.ni};
    int off = 0;
    for (int f = 0; f < 4; f++) {
        m.face_type[f].assign(d->face_type + off, d->face_type + off + len[f]);
        m.face_value[f].assign(d->face_value + off, d->face_value + off + len[f]);
        off += len[f];
    }
    size_t nn = (size_t)m.ni * m.nj;
    m.node_type.assign(d->node_type, d->node_type + nn);
    m.node_bc.assign(d->node_bc, d->node_bc + nn);
    m.node_vol.init(m.ni, m.nj);
    std::copy(d->node_vol, d->node_vol + nn, m.node_vol.d.begin());
    m.seg.resize(d->nseg);
    for (int q = 0; q < d->nseg; q++) {
        m.seg[q].init(d->seg_xy + 4 * q, d->seg_xy + 4 * q + 2);
        m.seg[q].btype = d->seg_btype[q];
        m.seg[q].absorb = d->seg_absorb[q];
        m.seg[q].bnd = d->seg_bnd ? d->seg_bnd[q] : -1;
    }
    m.nseg_ptr.assign(d->node_seg_ptr, d->node_seg_ptr + nn + 1);
    m.nseg_idx.assign(d->node_seg_idx, d->node_seg_idx + m.nseg_ptr[nn]);
    s->rnd.set_seed(d->seed);
    s->nthreads = d->nthreads > 0 ? d->nthreads : 1;
    s->nbnd = d->seg_bnd ? std::max(0, d->nbnd) : 0;
    s->dt = d->dt;
    s->phi.init(m.ni, m.nj);
    s->rho.init(m.ni, m.nj);
    s->efi.init(m.ni, m.nj);
    s->bfi.init(m.ni, m.nj);
    s->bfj.init(m.ni, m.nj);
    s->efj.init(m.ni, m.nj);
    return s;
}

void orc_destroy(orc_sim* s) { delete s; }

int orc_boundary_hits(orc_sim* s, int sp, int64_t* out, int nbnd, int clear) {
    if (sp < 0 || sp >= (int)s->species.size() || nbnd != s->nbnd) return -1;
    Species& S = *s->species[sp];
    for (int b = 0; b < nbnd; b++) {
        out[b] = S.bnd_hits[b].load();
        if (clear) S.bnd_hits[b] = 0;
    }
    return 0;
}

int orc_add_species(orc_sim* s, double charge_e, double molwt, double spwt, double nd_back) {
    auto sp = std::make_unique<Species>();
    sp->charge = charge_e * QE;                  /* Material.java:64-65 */
    sp->mass = molwt * AMU;
    sp->spwt = spwt;
    sp->qm = sp->charge / sp->mass;              /* Material.init (Material.java:711) */
    sp->nd_back = nd_back;
    int ni = s->mesh.ni, nj = s->mesh.nj;
    for (Field* f : {&sp->den, &sp->U, &sp->V, &sp->W, &sp->count_sum, &sp->u_sum, &sp->v_sum, &sp->w_sum,
                     &sp->uu_sum, &sp->vv_sum, &sp->ww_sum, &sp->mpc_sum, &sp->nd_ave, &sp->u_ave, &sp->v_ave,
                     &sp->w_ave, &sp->T, &sp->T1, &sp->T2, &sp->T3, &sp->p, &sp->mpc})
        f->init(ni, nj);
    for (double& a : sp->den.d) a += nd_back;    /* Material.init adds nd_back */
    sp->nbnd = s->nbnd;
    sp->bnd_hits.reset(new std::atomic<long long>[std::max(1, s->nbnd)]);
    for (int b = 0; b < std::max(1, s->nbnd); b++) sp->bnd_hits[b] = 0;
    s->species.push_back(std::move(sp));
    return (int)s->species.size() - 1;
}

void orc_set_solver(orc_sim* s, int method, int linear, double n0, double Te0, double phi0, double eps_r,
                    int max_it, double tol, int nl_max_it, double nl_tol) {
    s->method = method;
    s->linear = linear;
    s->n0 = n0;
    s->kTe0 = Te0;
    s->phi0 = phi0;
    s->eps = EPS0 * eps_r;
    s->lin_max_it = max_it;
    s->lin_tol = tol;
    s->nl_max_it = nl_max_it;
    s->nl_tol = nl_tol;
}

int orc_add_source(orc_sim* s, int sp, int npts, const double* pts, int reverse, double mdot, double v_drift,
                   double temp) {
    Source src;
    src.sp = sp;
    for (int k = 0; k + 1 < npts; k++) {         /* Spline.addLinearSegment (Spline.java:307-313) */
        Segment g{};
        if (!reverse) g.init(pts + 2 * k, pts + 2 * k + 2);
        else g.init(pts + 2 * k + 2, pts + 2 * k);
        src.segs.push_back(g);
    }
    if (reverse) std::reverse(src.segs.begin(), src.segs.end());
    src.cum_area.assign(src.segs.size() + 1, 0.0);     /* Spline.computeGeometry, XY area = length */
    for (size_t k = 0; k < src.segs.size(); k++) {
        src.spline_area += src.segs[k].length;
        src.cum_area[k + 1] = src.cum_area[k] + src.segs[k].length;
    }
    src.mdot = mdot;
    src.v_drift = v_drift;
    src.v_th = std::sqrt(2 * K_B * temp / s->species[sp]->mass);   /* Utils.computeVth */
    s->sources.push_back(src);
    return (int)s->sources.size() - 1;
}

void orc_species_init_T(orc_sim* s, int sp, double T) { s->species[sp]->init_T = T; }
void orc_set_it(orc_sim* s, int it) { s->it = it; }
void orc_sample_sources(orc_sim* s) { s->sample_sources(); }
int orc_get_it(const orc_sim* s) { return s->it; }

void orc_source_options(orc_sim* s, int src, double v_az, int start_it, int stop_it) {
    Source& so = s->sources[src];
    so.v_az = v_az;                                    /* MaxwellianSource.java:62, 117 */
    so.start_it = start_it;
    so.stop_it = stop_it;
}

void orc_load_particles(orc_sim* s, int spi, int64_t n, const double* pos3, const double* vel3, const double* mpw,
                        const int* id, int rewind) {
    Species& sp = *s->species[spi];
    for (int64_t p = 0; p < n; p++) {
        Particle part{};
        for (int d = 0; d < 3; d++) { part.pos[d] = pos3[3 * p + d]; part.vel[d] = vel3[3 * p + d]; }
        part.mpw = mpw ? mpw[p] : sp.spwt;
        if (rewind) {
            s->add_particle(sp, part, false);
        } else {
            s->mesh.XtoL(part.pos, part.lc);
            part.dt = 0;
            part.id = id ? id[p] : sp.id_counter;
            sp.id_counter++;
            sp.parts.push_back(part);
        }
    }
}

int64_t orc_num_particles(const orc_sim* s, int sp) { return (int64_t)s->species[sp]->parts.size(); }

void orc_get_particles(const orc_sim* s, int spi, double* pos3, double* vel3, double* lc2, double* mpw, int* id) {
    const Species& sp = *s->species[spi];
    for (size_t p = 0; p < sp.parts.size(); p++) {
        const Particle& q = sp.parts[p];
        for (int d = 0; d < 3; d++) {
            if (pos3) pos3[3 * p + d] = q.pos[d];
            if (vel3) vel3[3 * p + d] = q.vel[d];
        }
        if (lc2) { lc2[2 * p] = q.lc[0]; lc2[2 * p + 1] = q.lc[1]; }
        if (mpw) mpw[p] = q.mpw;
        if (id) id[p] = q.id;
    }
}

int orc_set_field(orc_sim* s, int sp, const char* name, const double* data) {
    Field* f = s->field(sp, name);
    if (!f) return -1;
    std::copy(data, data + f->d.size(), f->d.begin());
    return 0;
}
int orc_get_field(const orc_sim* s, int sp, const char* name, double* out) {
    Field* f = const_cast<orc_sim*>(s)->field(sp, name);
    if (!f) return -1;
    std::copy(f->d.begin(), f->d.end(), out);
    return 0;
}

void orc_move(orc_sim* s) { for (auto& sp : s->species) s->move_species(*sp); }
void orc_update_species_fields(orc_sim* s) { for (auto& sp : s->species) s->update_species_fields(*sp); }
void orc_compute_fields(orc_sim* s) { for (auto& sp : s->species) s->compute_fields(*sp); }
void orc_update_rho(orc_sim* s) { s->update_rho(); }
int orc_solve_potential(orc_sim* s) { return s->solve_potential(); }
void orc_update_efield(orc_sim* s) { s->update_efield(); }
void orc_initial_solve(orc_sim* s) {
    s->update_rho();
    s->solve_potential();
    s->update_efield();
}
void orc_step(orc_sim* s, int nsteps) { for (int k = 0; k < nsteps; k++) s->step(); }
void orc_get_stats(const orc_sim* s, int spi, double* out5) {
    const Species& sp = *s->species[spi];
    out5[0] = sp.mass_sum; out5[1] = sp.mom_sum[0]; out5[2] = sp.mom_sum[1]; out5[3] = sp.mom_sum[2];
    out5[4] = sp.energy_sum;
}
void orc_get_timers(const orc_sim* s, double* out4) { for (int k = 0; k < 4; k++) out4[k] = s->timers[k]; }
int orc_last_lin_iterations(const orc_sim* s) { return s->last_lin_it; }

int64_t orc_operator_nnz(orc_sim* s, int which) {
    if (!s->solver_ready) s->init_solver();
    const Matrix& M = which == 0 ? s->A : (which == 1 ? s->Gi : s->Gj);
    int64_t nnz = 0;
    for (auto& r : M.rows) nnz += (int64_t)r.size();
    return nnz;
}
void orc_operator_csr(orc_sim* s, int which, int64_t* rowptr, int* col, double* val) {
    if (!s->solver_ready) s->init_solver();
    const Matrix& M = which == 0 ? s->A : (which == 1 ? s->Gi : s->Gj);
    int64_t k = 0;
    rowptr[0] = 0;
    for (size_t i = 0; i < M.rows.size(); i++) {
        for (auto& kv : M.rows[i]) { col[k] = kv.first; val[k] = kv.second; k++; }
        rowptr[i + 1] = k;
    }
}
int orc_fixed_nodes(orc_sim* s, int* out) {
    if (!s->solver_ready) s->init_solver();
    int c = 0;
    for (size_t k = 0; k < s->fixed.size(); k++) { out[k] = s->fixed[k]; c += s->fixed[k]; }
    return c;
}

void orc_javarandom_doubles(int64_t seed, int n, double* out) {
    JavaRandom r(seed);
    for (int k = 0; k < n; k++) out[k] = r.next_double();
}
void orc_javarandom_ints(int64_t seed, int n, int* out) {
    JavaRandom r(seed);
    for (int k = 0; k < n; k++) out[k] = r.next_int();
}
void orc_javarandom_ints_bounded(int64_t seed, int bound, int n, int* out) {
    JavaRandom r(seed);
    for (int k = 0; k < n; k++) out[k] = r.next_int(bound);
}

}  // extern "C"

/* ---------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): the counter-based generator that replaces Starfish.rnd() in the MCC restatement, so
 * that the draws of a particle depend only on (seed, particle id, step) and not on the order in
 * which particles are visited.
 * ------------------------------------------------------------------------- */
static void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; round++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    for (int k = 0; k < 4; k++) out[k] = c[k];
}
extern "C" void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox4x32_10(ctr, key, out);
}

/* draw k of a particle's stream: block k/2 of the counter (block, id, step, tag), 53 bits from
   two 32-bit words, in [0, 1) like java.util.Random.nextDouble */
struct CounterRng {
    uint32_t key[2], id, step, tag;
    double draw(uint32_t k) const {
        uint32_t ctr[4] = {k >> 1, id, step, tag}, o[4];
        philox4x32_10(ctr, key, o);
        uint64_t bits = (k & 1) ? (((uint64_t)o[3] << 32) | o[2]) : (((uint64_t)o[1] << 32) | o[0]);
        return (double)(bits >> 11) * 0x1.0p-53;
    }
};

/* MCC.perform(Mesh) (collisions/MCC.java:167-221) over one kinetic source species with a
   constant cross-section (Sigma.SigmaConst, interactions/Sigma.java:108-113) and the CEX
   (MCC.java:287-297) or MEX/elastic (MCC.java:242-285) model.  The target material enters as its
   node fields: density, stream velocity (u-ave, v-ave, w-ave) and temperature
   (Material.sampleVelocity / sampleMaxwellianVelocity, Material.java:859-896; Utils.SampleMaxw3D
   with its 1.2 factor, Utils.java:27-57).  Starfish.rnd() is replaced by the counter-based
   stream: draw 0 is the collision test, draws 2..37 the target's Maxwellian (12 per component,
   each component starting a new Philox block), draws 38 and 39 the elastic scattering angles.  Accumulates col-count-sum and col-real-sum at
   node ((int)lc0, (int)lc1) and the target's energy source S (Field2D.scatter of
   mpw (E1 - E2) / (dt vol), with E = 0.5 m |v| as the reference writes it).  Returns the number
   of collisions. */
extern "C" int64_t orc_mcc_perform(orc_sim* s, int spi, const double* den, const double* uave, const double* vave,
                                   const double* wave, const double* temp, double target_mass, double sigma,
                                   int model, double max_T, uint64_t seed, uint32_t tag, int it, int frequency,
                                   double* count_sum, double* real_sum, double* S) {
    Species& sp = *s->species[spi];
    const Mesh& m = s->mesh;
    Field fden, fu, fv, fw, fT, fS;
    fden.init(m.ni, m.nj); fu.init(m.ni, m.nj); fv.init(m.ni, m.nj); fw.init(m.ni, m.nj); fT.init(m.ni, m.nj);
    fS.init(m.ni, m.nj);
    size_t nn = (size_t)m.ni * m.nj;
    std::copy(den, den + nn, fden.d.begin());
    std::copy(uave, uave + nn, fu.d.begin());
    std::copy(vave, vave + nn, fv.d.begin());
    std::copy(wave, wave + nn, fw.d.begin());
    std::copy(temp, temp + nn, fT.d.begin());
    std::copy(S, S + nn, fS.d.begin());
    const double dt = frequency * s->dt;
    CounterRng rng;
    rng.key[0] = (uint32_t)seed;
    rng.key[1] = (uint32_t)(seed >> 32);
    rng.step = (uint32_t)it;
    rng.tag = (1u << 28) | tag;          /* MCC streams carry module 1 in the tag's top bits */
    int64_t ncol = 0;
    for (Particle& part : sp.parts) {
        if (part.mpw <= 0) continue;
        double den_a = gather(fden, part.lc[0], part.lc[1]);
        if (den_a <= 0) continue;
        double tv[3] = {gather(fu, part.lc[0], part.lc[1]), gather(fv, part.lc[0], part.lc[1]),
                        gather(fw, part.lc[0], part.lc[1])};
        double gv[3];
        for (int i = 0; i < 3; i++) gv[i] = tv[i] - part.vel[i];
        double g = std::sqrt(gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2]);
        double P = 1 - std::exp(-sigma * g * dt * den_a);
        rng.id = (uint32_t)part.id;
        if (P < rng.draw(0)) continue;
        /* target.sampleMaxwellianVelocity(mesh, lc, 0, max_T) */
        double vt[3] = {tv[0], tv[1], tv[2]};
        double T = gather(fT, part.lc[0], part.lc[1]);
        if (T < 0) T = 0;
        if (max_T > 0 && T > max_T) T = max_T;
        double vth = std::sqrt(2 * K_B * T / target_mass);   /* Utils.computeVth */
        for (int c = 0; c < 3; c++) {
            double sum = 0;
            for (int k = 0; k < 12; k++) sum += rng.draw(2 + 12 * c + k);
            double fm = (sum - 0.5 * 12) * std::sqrt(12.0 / 12);
            vt[c] += (vth * 1.2) * (std::sqrt(0.5) * 1 * fm);
        }
        double E1 = 0.5 * sp.mass * std::sqrt(part.vel[0] * part.vel[0] + part.vel[1] * part.vel[1] +
                                              part.vel[2] * part.vel[2]);
        if (model == 0) {
            for (int c = 0; c < 3; c++) part.vel[c] = vt[c];
        } else {
            double rm1 = sp.mass / (sp.mass + target_mass), rm2 = target_mass / (sp.mass + target_mass);
            double gg[3], vcm[3], vr[3];
            for (int c = 0; c < 3; c++) gg[c] = part.vel[c] - vt[c];
            double gmag = std::sqrt(gg[0] * gg[0] + gg[1] * gg[1] + gg[2] * gg[2]);
            for (int c = 0; c < 3; c++) vcm[c] = rm1 * part.vel[c] + rm2 * vt[c];
            double B = 2. * rng.draw(38) - 1.0;
            double A = std::sqrt(1. - B * B);
            vr[0] = B * gmag;
            double C = 2. * PI * rng.draw(39);
            vr[1] = A * std::cos(C) * gmag;
            vr[2] = A * std::sin(C) * gmag;
            for (int c = 0; c < 3; c++) part.vel[c] = vcm[c] + vr[c] * rm2;
        }
        double E2 = 0.5 * sp.mass * std::sqrt(part.vel[0] * part.vel[0] + part.vel[1] * part.vel[1] +
                                              part.vel[2] * part.vel[2]);
        double vol = m.node_volume(part.lc[0], part.lc[1]);
        double dS = part.mpw * (E1 - E2) / (dt * vol);
        scatter(m, fS, part.lc[0], part.lc[1], dS);
        int i = (int)part.lc[0], j = (int)part.lc[1];
        count_sum[(size_t)j * m.ni + i] += 1;
        real_sum[(size_t)j * m.ni + i] += part.mpw;
        ncol++;
    }
    std::copy(fS.d.begin(), fS.d.end(), S);
    return ncol;
}

/* Utils.gamma (core/common/Utils.java:218-236): Bird's approximation of the Gamma function */
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

extern "C" double orc_bird_gamma(double x) { return bird_gamma(x); }

/* DSMC pair parameters: the cross-section (SigmaConst, or SigmaBird463 of
   collisions/SigmaPlus.java:38-78 from the two materials' diam, ref_temp, visc_temp_index) and the
   VSS exponent vss_inv = (1/alpha1 + 1/alpha2)/2 (DSMC.java:83) */
struct DsmcPair {
    int sigma_type;   /* 0 const, 1 Bird463 */
    double c0, sigma0, ref_temp, visc_temp_index, reduced_mass, gamma, vss_inv;
    double eval(double g) const {
        if (sigma_type == 0) return c0;
        return sigma0 * std::pow(2. * K_B * ref_temp / (reduced_mass * g * g), visc_temp_index - 0.5) / gamma;
    }
};

/* DSMC.perform(Mesh) + collideCell + ModelElastic (collisions/DSMC.java:167-430) for species
   spa, spb (the same index for a single-species pair), one pass at step `it`.  Cell lists hold the
   particles with (int)lc inside the cells, ordered by particle id; Starfish.rnd() is replaced by
   the counter-based stream of the cell (draw k of counter (k/2, cell, step, tag)), consumed in
   the reference's order.  sig_cr_max and rem are the per-cell state (cell index j*(ni-1)+i),
   count_sum / real_sum the node fields at node (i, j).  Returns the number of collisions. */
extern "C" int64_t orc_dsmc_perform(orc_sim* s, int spa, int spb, int sigma_type, double c0, double diam_a,
                                    double diam_b, double ref_temp_a, double ref_temp_b, double vti_a, double vti_b,
                                    double vss_alpha_a, double vss_alpha_b, uint64_t seed, uint32_t tag, int it,
                                    int frequency, double* sig_cr_max, double* rem, double* count_sum,
                                    double* real_sum) {
    const Mesh& m = s->mesh;
    Species& A = *s->species[spa];
    Species& B = *s->species[spb];
    const bool same = spa == spb;
    DsmcPair pr;
    pr.sigma_type = sigma_type;
    pr.c0 = c0;
    double d_ref = diam_a + diam_b;
    pr.sigma0 = 0.25 * PI * d_ref * d_ref;
    pr.ref_temp = 0.5 * (ref_temp_a + ref_temp_b);
    pr.visc_temp_index = 0.5 * (vti_a + vti_b);
    pr.reduced_mass = A.mass * B.mass / (A.mass + B.mass);
    pr.gamma = bird_gamma(2.5 - pr.visc_temp_index);
    pr.vss_inv = 0.5 * (1.0 / vss_alpha_a + 1.0 / vss_alpha_b);
    const int nci = m.ni - 1, ncj = m.nj - 1;
    std::vector<std::vector<Particle*>> la((size_t)nci * ncj), lb(same ? 0 : (size_t)nci * ncj);
    auto bin = [&](Species& sp, std::vector<std::vector<Particle*>>& L) {
        for (Particle& p : sp.parts) {
            if (p.mpw <= 0) continue;
            int i = (int)p.lc[0], j = (int)p.lc[1];
            if (i >= m.ni - 1 || j >= m.nj - 1) continue;
            L[(size_t)j * nci + i].push_back(&p);
        }
        for (auto& c : L) std::sort(c.begin(), c.end(), [](const Particle* x, const Particle* y) { return x->id < y->id; });
    };
    bin(A, la);
    if (!same) bin(B, lb);
    const double delta_t = frequency * s->dt;
    const double spwt1 = A.spwt, spwt2 = B.spwt;
    CounterRng rng;
    rng.key[0] = (uint32_t)seed;
    rng.key[1] = (uint32_t)(seed >> 32);
    rng.step = (uint32_t)it;
    rng.tag = (2u << 28) | tag;          /* DSMC streams carry module 2 in the tag's top bits */
    int64_t ntot = 0;
    for (int j = 0; j < ncj; j++)
        for (int i = 0; i < nci; i++) {
            const size_t c = (size_t)j * nci + i;
            std::vector<Particle*>& l1 = la[c];
            std::vector<Particle*>& l2 = same ? la[c] : lb[c];
            double np1 = (double)l1.size(), np2 = (double)l2.size();
            double cell_volume = m.node_volume(i + 0.5, j + 0.5);   /* Mesh.cellVol */
            double Pab = spwt2 / spwt1, Pba = spwt1 / spwt2;
            if (Pab > 1) Pab = 1;
            if (Pba > 1) Pba = 1;
            double nsel_f = ((np1 * spwt1 / cell_volume) * np2 * delta_t * sig_cr_max[c]) / (Pab + (spwt2 / spwt1) * Pba);
            if (!same) nsel_f *= 2.0;
            nsel_f += rem[c];
            int nsel = (int)nsel_f;
            if ((same && (np1 < 2 || np2 < 2)) || np1 < 1 || np2 < 1) continue;
            rem[c] = nsel_f - nsel;
            if (nsel > np1 * np2) {
                nsel = (int)(np1 * np2);
                rem[c] = 0;
            }
            rng.id = (uint32_t)c;
            uint32_t k = 0;
            double scm = 0, sums0 = 0, sums1 = 0;
            for (int q = 0; q < nsel; q++) {
                int p1 = (int)(rng.draw(k++) * np1), p2;
                do {
                    p2 = (int)(rng.draw(k++) * np2);
                } while (same && p1 == p2);
                Particle& a = *l1[p1];
                Particle& b = *l2[p2];
                double cr[3] = {a.vel[0] - b.vel[0], a.vel[1] - b.vel[1], a.vel[2] - b.vel[2]};
                double cr_mag = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
                double sigma_cr = cr_mag * pr.eval(cr_mag);
                if (sigma_cr > scm) scm = sigma_cr;
                double P = sigma_cr / sig_cr_max[c];
                if (rng.draw(k++) < P) {
                    /* ModelElastic (DSMC.java:360-430) */
                    double ma = A.mass, mb = B.mass;
                    double rm1 = ma / (ma + mb), rm2 = mb / (ma + mb);
                    double g[3] = {a.vel[0] - b.vel[0], a.vel[1] - b.vel[1], a.vel[2] - b.vel[2]};
                    double g_mag = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
                    double vc[3], vr[3];
                    for (int d = 0; d < 3; d++) vc[d] = rm1 * a.vel[d] + rm2 * b.vel[d];
                    if (std::fabs(pr.vss_inv - 1.) < 1.e-4) {
                        double Bc = 2. * rng.draw(k++) - 1.0;
                        double Ac = std::sqrt(1. - Bc * Bc);
                        vr[0] = Bc * g_mag;
                        double Cc = 2. * PI * rng.draw(k++);
                        vr[1] = Ac * std::cos(Cc) * g_mag;
                        vr[2] = Ac * std::sin(Cc) * g_mag;
                    } else {
                        double Bc = 2. * (std::pow(rng.draw(k++), pr.vss_inv)) - 1.;
                        double Ac = std::sqrt(1. - Bc * Bc);
                        double Cc = 2. * PI * rng.draw(k++);
                        double OC = std::cos(Cc), SC = std::sin(Cc);
                        double D = std::sqrt(g[1] * g[1] + g[2] * g[2]);
                        if (D > 1.e-6) {
                            vr[0] = Bc * g[0] + Ac * SC * D;
                            vr[1] = Bc * g[1] + Ac * (g_mag * g[2] * OC - g[0] * g[1] * SC) / D;
                            vr[2] = Bc * g[2] - Ac * (g_mag * g[1] * OC + g[0] * g[2] * SC) / D;
                        } else {
                            vr[0] = Bc * g[1];
                            vr[1] = Ac * OC * g[1];
                            vr[2] = Ac * SC * g[1];
                        }
                    }
                    if (a.mpw == b.mpw) {
                        for (int d = 0; d < 3; d++) {
                            a.vel[d] = vc[d] + vr[d] * rm2;
                            b.vel[d] = vc[d] - vr[d] * rm1;
                        }
                    } else {
                        double pab = b.mpw / a.mpw, pba = a.mpw / b.mpw;
                        if (rng.draw(k++) < pab)
                            for (int d = 0; d < 3; d++) a.vel[d] = vc[d] + rm2 * vr[d];
                        if (rng.draw(k++) < pba)
                            for (int d = 0; d < 3; d++) b.vel[d] = vc[d] - rm1 * vr[d];
                    }
                    sums0 += 1.0;
                    sums1 += 0.5 * (a.mpw + b.mpw);
                }
            }
            if (scm > 0) sig_cr_max[c] = scm;
            count_sum[(size_t)j * m.ni + i] += sums0;
            real_sum[(size_t)j * m.ni + i] += sums1;
            ntot += (int64_t)sums0;
        }
    return ntot;
}

// BoltzmannElectronsMaterial.electronModelBoltzmann (core/materials/BoltzmannElectronsMaterial.java:
// 200-210): den = den0 * Math.exp((phi - phi0) / kTe0), then den < 1 -> 1
extern "C" void orc_boltzmann_density(int64_t n, const double* phi, double den0, double kTe0, double phi0,
                                      double* out) {
    for (int64_t u = 0; u < n; u++) {
        double d = den0 * std::exp((phi[u] - phi0) / kTe0);
        out[u] = d < 1 ? 1 : d;
    }
}

extern "C" void orc_set_rng_state(orc_sim* s, int64_t state) { s->rnd.seed = (uint64_t)state & ((1ULL << 48) - 1); }

// Test hook: rounds every particle's in-cell offsets and velocities to FP32, as the device stores
// them (int32 cell + FP32 offset, FP32 velocity); positions follow from the rounded offsets.
extern "C" void orc_round_state_fp32(orc_sim* s, int spi) {
    Species& sp = *s->species[spi];
    for (auto& q : sp.parts) {
        for (int d = 0; d < 2; d++) {
            double c = std::floor(q.lc[d]);
            q.lc[d] = c + (double)(float)(q.lc[d] - c);
        }
        for (int d = 0; d < 3; d++) q.vel[d] = (double)(float)q.vel[d];
        double x[2];
        s->mesh.pos(q.lc[0], q.lc[1], x);
        q.pos[0] = x[0];
        q.pos[1] = x[1];
    }
}
