// Host-side setup that feeds the device hot path: mesh geometry (node types, node segment
// lists, node volumes) and Maxwellian boundary sources.  A Java host gets the geometry from
// Starfish's own Mesh; these entry points serve hosts without it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sf_internal.h"

using namespace sf;

namespace {

struct LSeg {
    double x1[2], x2[2], nrm[2], length;
    int btype, bnd, id, absorb;
    void init(const double* a, const double* b) {
        x1[0] = a[0]; x1[1] = a[1]; x2[0] = b[0]; x2[1] = b[1];
        double dx = x2[0] - x1[0], dy = x2[1] - x1[1];
        length = std::sqrt(dx * dx + dy * dy);
        nrm[0] = -dy / length;
        nrm[1] = dx / length;
    }
    void pos(double t, double* x) const {
        x[0] = x1[0] + t * (x2[0] - x1[0]);
        x[1] = x1[1] + t * (x2[1] - x1[1]);
    }
    // LinearSegment.intersect (core/boundaries/LinearSegment.java:113-160), t[0] along this
    double intersect0(const double* p3, const double* p4) const {
        double X1 = x1[0], X2 = x2[0], X3 = p3[0], X4 = p4[0];
        double Y1 = x1[1], Y2 = x2[1], Y3 = p3[1], Y4 = p4[1];
        double den = (X1 - X2) * (Y3 - Y4) - (Y1 - Y2) * (X3 - X4);
        if (den == 0) return -1;
        double xp0 = ((X1 * Y2 - Y1 * X2) * (X3 - X4) - (X1 - X2) * (X3 * Y4 - Y3 * X4)) / den;
        double xp1 = ((X1 * Y2 - Y1 * X2) * (Y3 - Y4) - (Y1 - Y2) * (X3 * Y4 - Y3 * X4)) / den;
        double t0, t1;
        if (std::fabs(x2[0] - x1[0]) > 1e-6) t0 = (xp0 - x1[0]) / (x2[0] - x1[0]);
        else t0 = (xp1 - x1[1]) / (x2[1] - x1[1]);
        if (t0 < -FLT_EPS || t0 > 1 + FLT_EPS) return -1;
        if (std::fabs(p4[0] - p3[0]) > 1e-6) t1 = (xp0 - p3[0]) / (p4[0] - p3[0]);
        else t1 = (xp1 - p3[1]) / (p4[1] - p3[1]);
        if (t1 < -FLT_EPS || t1 > 1 + FLT_EPS) return -1;
        return t0 < 0 ? 0 : (t0 > 1 ? 1 : t0);
    }
};

bool in_box(const double* p, const double* b1, const double* b2) {
    return p[0] >= b1[0] && p[0] <= b2[0] && p[1] >= b1[1] && p[1] <= b2[1];
}

// Segment.segmentInBox (core/boundaries/Segment.java:236-283)
bool segment_in_box(const LSeg& s, const double* b1, const double* b2) {
    if (in_box(s.x1, b1, b2) || in_box(s.x2, b1, b2)) return true;
    double f[4][4] = {{b1[0], b1[1], b2[0], b1[1]},    // BOTTOM
                      {b2[0], b1[1], b2[0], b2[1]},    // RIGHT
                      {b2[0], b2[1], b1[0], b2[1]},    // TOP
                      {b1[0], b2[1], b1[0], b1[1]}};   // LEFT
    for (auto& e : f)
        if (s.intersect0(e, e + 2) >= 0) return true;
    return false;
}

// Spline.isOnLine (core/boundaries/Spline.java:835-862)
bool is_on_line(const double* xp, const LSeg& s) {
    double t;
    if (s.x2[0] != s.x1[0]) t = (xp[0] - s.x1[0]) / (s.x2[0] - s.x1[0]);
    else t = (xp[1] - s.x1[1]) / (s.x2[1] - s.x1[1]);
    if (t < -FLT_EPS || t > 1.0 + FLT_EPS) return false;
    double p0 = s.x1[0] + t * (s.x2[0] - s.x1[0]), p1 = s.x1[1] + t * (s.x2[1] - s.x1[1]);
    double d1 = p0 - xp[0], d2 = p1 - xp[1];
    double d = (d1 * d1 + d2 * d2) / (s.length * s.length);
    return !(d > FLT_EPS);
}

// Spline.isInternal(xp, segment) (Spline.java:806-833); closestPos of a linear segment is 0.5
bool is_internal(const double* xp, const LSeg& s) {
    if (is_on_line(xp, s)) return true;
    double m[2];
    s.pos(0.5, m);
    double r0 = xp[0] - m[0], r1 = xp[1] - m[1];
    double rm = std::sqrt(r0 * r0 + r1 * r1);
    double c = (r0 * s.nrm[0] + r1 * s.nrm[1]) / rm;
    return c <= 0;
}

struct Geo {
    sf_mesh_desc m;
    std::vector<LSeg> segs;
    std::vector<std::vector<int>> nseg;   // node segment lists
    std::vector<int> type;
    std::vector<double> bc, vol;
    std::vector<int> face;                // face types (RIGHT, TOP, LEFT, BOTTOM)
    std::vector<double> facev;
    std::vector<int> bnd_type;
    std::vector<double> bnd_value;

    int ni() const { return m.ni; }
    int nj() const { return m.nj; }
    size_t u(int i, int j) const { return (size_t)j * m.ni + i; }
    void pos(double i, double j, double* x) const {
        x[0] = m.x0[0] + i * m.dh[0];
        x[1] = m.x0[1] + j * m.dh[1];
    }
    int face_type(int f, int idx) const {
        int off = f == 0 ? 0 : (f == 1 ? m.nj : (f == 2 ? m.nj + m.ni : 2 * m.nj + m.ni));
        return face[off + idx];
    }
    // Spline.visibleSegment (Spline.java:877-926), linear segments only
    int visible_segment(const double* xp, const std::vector<int>& list) const {
        for (int a : list) {
            const LSeg& s = segs[a];
            if (s.btype == SF_BND_VIRTUAL || s.btype == SF_BND_OPEN) continue;
            if (is_on_line(xp, s)) return a;
            double xc[2];
            s.pos(0.5, xc);
            bool hit = false;
            for (int b : list) {
                if (b == a) continue;
                const LSeg& s2 = segs[b];
                if (s2.btype == SF_BND_VIRTUAL || s2.btype == SF_BND_OPEN) continue;
                if (s2.intersect0(xp, xc) >= 0) { hit = true; break; }
            }
            if (!hit) return a;
        }
        return -1;
    }
    // Mesh.isInternalPoint (Mesh.java:1307-1334)
    bool internal_point(double l0, double l1) const {
        int i1 = (int)l0, j1 = (int)l1;
        int i2 = i1 + 1, j2 = j1 + 1;
        if (i2 >= ni()) i2 = ni() - 1;
        if (j2 >= nj()) j2 = nj() - 1;
        double x[2];
        pos(l0, l1, x);
        for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++) {
                const std::vector<int>& bl = nseg[u(i, j)];
                if (bl.empty()) continue;
                int s = visible_segment(x, bl);
                if (s < 0) continue;
                return is_internal(x, segs[s]);
            }
        return false;
    }
    static int minmax(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
    // Mesh.controlVolumeLCs (Mesh.java:656-747) for a single mesh
    void cv_lcs(double i0, double j0, double delta, double l[4][2]) const {
        double i = -1, j = -1;
        for (int k = 0; k < 4; k++) {
            switch (k) {
                case 0: i = i0 - delta; j = j0 - delta; break;
                case 1: i = i0 + delta; j = j0 - delta; break;
                case 2: i = i0 + delta; j = j0 + delta; break;
                case 3: i = i0 - delta; j = j0 + delta; break;
            }
            int ii = (int)i0, jj = (int)j0;
            auto t = [&](int a, int b) { return type[u(a, b)]; };
            if (k > 1 && jj > 1 && (t(ii, jj) == SF_NODE_INSULATOR && t(ii, jj - 1) == SF_NODE_OPEN)) j = j0;
            else if ((k == 0 || k == 3) && ii > 1 && ii + 1 < ni() &&
                     (t(ii, jj) == SF_NODE_INSULATOR && t(ii + 1, jj) == SF_NODE_OPEN))
                i = i0;
            else if ((k == 1 || k == 2) && ii < ni() - 1 && ii >= 1 &&
                     (t(ii, jj) == SF_NODE_INSULATOR && t(ii - 1, jj) == SF_NODE_OPEN))
                i = i0;
            if (i < 0 && face_type(2, minmax((int)(j0 + 0.5), 0, nj() - 1)) != SF_FACE_MESH) i = 0;
            if (i > ni() - 1 && face_type(0, minmax((int)(j0 + 0.5), 0, nj() - 1)) != SF_FACE_MESH) i = ni() - 1;
            if (j < 0 && face_type(3, minmax((int)(i0 + 0.5), 0, ni() - 1)) != SF_FACE_MESH) j = 0;
            if (j > nj() - 1 && face_type(1, minmax((int)(i0 + 0.5), 0, ni() - 1)) != SF_FACE_MESH) j = nj() - 1;
            if (i < 0 || j < 0 || i > ni() - 1 || j > nj() - 1) {
                if (i < 0) i = 0;
                if (j < 0) j = 0;
                if (i >= ni() - 1) i = ni() - 1;
                if (j >= nj() - 1) j = nj() - 1;
            }
            l[k][0] = i;
            l[k][1] = j;
        }
        l[0][1] = std::max(l[0][1], l[1][1]);
        l[1][1] = l[0][1];
        l[1][0] = std::min(l[1][0], l[2][0]);
        l[2][0] = l[1][0];
        l[1][0] = std::min(l[1][0], l[2][0]);
        l[2][0] = l[1][0];
        l[2][1] = std::max(l[2][1], l[3][1]);
        l[3][1] = l[2][1];
        l[3][0] = std::max(l[3][0], l[0][0]);
        l[0][0] = l[3][0];
    }
    double area(double i0, double j0) const {   // Mesh.area (Mesh.java:965-983)
        double l[4][2], V[4][2];
        cv_lcs(i0, j0, 0.5, l);
        for (int k = 0; k < 4; k++) pos(l[k][0], l[k][1], V[k]);
        double A = 0.5 * std::fabs(V[0][0] * V[1][1] - V[0][0] * V[2][1] + V[1][0] * V[2][1] - V[1][0] * V[0][1] +
                                   V[2][0] * V[0][1] - V[2][0] * V[1][1]);
        A += 0.5 * std::fabs(V[0][0] * V[2][1] - V[0][0] * V[3][1] + V[2][0] * V[3][1] - V[2][0] * V[0][1] +
                             V[3][0] * V[0][1] - V[3][0] * V[2][1]);
        return A;
    }
    double node_vol(int i0, int j0) const {     // Mesh.nodeVol (Mesh.java:994-1020)
        if (m.domain == SF_DOMAIN_XY) return area(i0, j0);
        double r;
        if (m.domain == SF_DOMAIN_RZ) {
            r = m.x0[0] + i0 * m.dh[0];
            if (i0 == 0) r = m.x0[0] + (i0 + 0.25) * m.dh[0];
            else if (i0 == ni() - 1) r = m.x0[0] + (i0 - 0.25) * m.dh[0];
        } else {
            r = m.x0[1] + j0 * m.dh[1];
            if (j0 == 0) r = m.x0[1] + (j0 + 0.25) * m.dh[1];
            else if (j0 == nj() - 1) r = m.x0[1] + (j0 - 0.25) * m.dh[1];
        }
        return 2 * PI * area(i0, j0) * r;
    }
};

}  // namespace

struct sf_geometry {
    Geo g;
};

extern "C" {

int sf_geometry_build(sf_ctx* ctx, const sf_mesh_desc* mesh, const int* face_type, const double* face_value,
                      int nbnd, const int* bnd_type, const double* bnd_value, const int* bnd_has_material,
                      const int* bnd_npts, const double* bnd_pts, const int* bnd_reverse, sf_geometry** out) {
    SF_CHECK_CTX(ctx);
    if (!mesh || !face_type || !out || (nbnd > 0 && (!bnd_type || !bnd_npts || !bnd_pts)))
        return sf_fail(ctx, SF_ERR_ARG, "sf_geometry_build: null argument");
    if (mesh->ni < 2 || mesh->nj < 2 || !(mesh->dh[0] > 0) || !(mesh->dh[1] > 0))
        return sf_fail(ctx, SF_ERR_ARG, "sf_geometry_build: bad mesh");
    sf_geometry* G = new sf_geometry();
    Geo& g = G->g;
    g.m = *mesh;
    const int ni = mesh->ni, nj = mesh->nj;
    const size_t nn = (size_t)ni * nj;
    const int nface = 2 * ni + 2 * nj;
    g.face.assign(face_type, face_type + nface);
    g.facev.assign(nface, 0.0);
    if (face_value) g.facev.assign(face_value, face_value + nface);
    g.type.assign(nn, SF_NODE_UNKNOWN);
    g.bc.assign(nn, 0.0);
    g.nseg.assign(nn, {});
    // linear segments of every boundary (Spline.setPath / addLinearSegment / reverseSegmentOrdering)
    const double* pts = bnd_pts;
    for (int b = 0; b < nbnd; b++) {
        int np = bnd_npts[b];
        std::vector<LSeg> ss;
        bool rev = bnd_reverse && bnd_reverse[b];
        for (int k = 0; k + 1 < np; k++) {
            LSeg s{};
            if (!rev) s.init(pts + 2 * k, pts + 2 * k + 2);
            else s.init(pts + 2 * k + 2, pts + 2 * k);
            if (!(s.length > 0)) {
                delete G;
                return sf_fail(ctx, SF_ERR_ARG, "boundary %d has a zero-length segment", b);
            }
            s.btype = bnd_type[b];
            s.bnd = b;
            s.absorb = bnd_has_material ? bnd_has_material[b] : 1;
            ss.push_back(s);
        }
        if (rev) std::reverse(ss.begin(), ss.end());
        for (size_t k = 0; k < ss.size(); k++) {
            ss[k].id = (int)k;
            g.segs.push_back(ss[k]);
        }
        g.bnd_type.push_back(bnd_type[b]);
        g.bnd_value.push_back(bnd_value ? bnd_value[b] : 0.0);
        pts += 2 * np;
    }
    // Mesh.setBoundaries -> setNodeControlVolumes (Mesh.java:1219-1295)
    if (!g.segs.empty()) {
        for (size_t a = 0; a < g.segs.size(); a++) {
            const LSeg& s = g.segs[a];
            double bmin[2] = {std::min(s.x1[0], s.x2[0]), std::min(s.x1[1], s.x2[1])};
            double bmax[2] = {std::max(s.x1[0], s.x2[0]), std::max(s.x1[1], s.x2[1])};
            int lcm[2] = {(int)((bmin[0] - mesh->x0[0]) / mesh->dh[0]), (int)((bmin[1] - mesh->x0[1]) / mesh->dh[1])};
            int lcp[2] = {(int)((bmax[0] - mesh->x0[0]) / mesh->dh[0]), (int)((bmax[1] - mesh->x0[1]) / mesh->dh[1])};
            lcm[0]--; lcm[1]--;
            lcp[0] += 2; lcp[1] += 2;
            if (lcm[0] < 0) lcm[0] = 0;
            if (lcm[1] < 0) lcm[1] = 0;
            if (lcp[0] > ni - 1) lcp[0] = ni - 1;
            if (lcp[1] > nj - 1) lcp[1] = nj - 1;
            for (int j = lcm[1]; j <= lcp[1]; j++)
                for (int i = lcm[0]; i <= lcp[0]; i++) {
                    const double bsize = 1.01;
                    double m1[2], p1[2];
                    g.pos(i - bsize, j - bsize, m1);
                    g.pos(i + bsize, j + bsize, p1);
                    if (!segment_in_box(s, m1, p1)) continue;
                    std::vector<int>& lst = g.nseg[g.u(i, j)];
                    bool found = false;
                    for (int q : lst)
                        if (g.segs[q].bnd == s.bnd && g.segs[q].id == s.id) { found = true; break; }
                    if (!found) lst.push_back((int)a);
                }
        }
        // setInterfaceNodeLocation (Mesh.java:1337-1378)
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) {
                const std::vector<int>& bl = g.nseg[g.u(i, j)];
                if (bl.empty()) continue;
                double x[2];
                g.pos(i, j, x);
                int s = g.visible_segment(x, bl);
                if (s < 0) continue;
                size_t uu = g.u(i, j);
                if (is_internal(x, g.segs[s])) {
                    int bt = g.segs[s].btype;
                    if (bt == SF_BND_INSULATOR) {
                        g.type[uu] = SF_NODE_INSULATOR;
                        g.bc[uu] = g.bnd_value[g.segs[s].bnd];
                    } else if (bt == SF_BND_DIRICHLET) {
                        g.type[uu] = SF_NODE_DIRICHLET;
                        g.bc[uu] = g.bnd_value[g.segs[s].bnd];
                    } else {
                        delete G;
                        return sf_fail(ctx, SF_ERR_ARG, "unknown internal boundary type %d", bt);
                    }
                } else if (g.type[uu] == SF_NODE_UNKNOWN) {
                    g.type[uu] = SF_NODE_OPEN;
                    g.bc[uu] = 0;
                }
            }
        // performFloodFill (Mesh.java:1384-1460)
        bool has = false;
        for (size_t k = 0; k < nn && !has; k++) has = g.type[k] != SF_NODE_UNKNOWN;
        if (!has) {
            for (size_t k = 0; k < nn; k++) g.type[k] = SF_NODE_OPEN;
        } else {
            auto ok = [&](int i, int j) {
                int t = g.type[g.u(i, j)];
                return t == SF_NODE_OPEN || t == SF_NODE_DIRICHLET || t == SF_NODE_INTERNAL;
            };
            for (long pass = 0; pass < 20L * ni * nj; pass++) {
                int count = 0;
                for (int i = 0; i < ni; i++)
                    for (int j = 0; j < nj; j++) {
                        size_t uu = g.u(i, j);
                        if (g.type[uu] != SF_NODE_UNKNOWN) continue;
                        long from = -1;
                        if (i > 0 && ok(i - 1, j)) from = (long)g.u(i - 1, j);
                        else if (i < ni - 1 && ok(i + 1, j)) from = (long)g.u(i + 1, j);
                        else if (j > 0 && ok(i, j - 1)) from = (long)g.u(i, j - 1);
                        else if (j < nj - 1 && ok(i, j + 1)) from = (long)g.u(i, j + 1);
                        if (from >= 0) {
                            g.type[uu] = g.type[from];
                            g.bc[uu] = g.bc[from];
                            count++;
                            if (g.type[uu] == SF_NODE_DIRICHLET) g.type[uu] = SF_NODE_INTERNAL;
                        }
                    }
                if (count == 0) break;
            }
        }
    } else {
        for (size_t k = 0; k < nn; k++) g.type[k] = SF_NODE_UNKNOWN;
    }
    // Mesh.initNodes (Mesh.java:364-402): node volumes, Monte-Carlo interface correction
    g.vol.assign(nn, 0.0);
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) g.vol[g.u(i, j)] = g.node_vol(i, j);
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            size_t uu = g.u(i, j);
            if (g.nseg[uu].empty()) continue;
            if (g.internal_point(i, j)) continue;
            int inside = 0, good = 0;
            for (int d = 0; d < 1000; d++) {
                double l0 = i + 0.5 * ctx->rng.rnd2();
                double l1 = j + 0.5 * ctx->rng.rnd2();
                if (l0 < 0 || l1 < 0 || l0 >= ni - 1 || l1 >= nj - 1) continue;
                inside++;
                if (!g.internal_point(l0, l1)) good++;
            }
            if (good > 0) g.vol[uu] *= good / (double)inside;
        }
    for (int i = 0; i < ni; i++) {
        if (g.face_type(3, i) == SF_FACE_DIRICHLET) {
            g.type[g.u(i, 0)] = SF_NODE_DIRICHLET;
            g.bc[g.u(i, 0)] = g.facev[2 * nj + ni + i];
        }
        if (g.face_type(1, i) == SF_FACE_DIRICHLET) {
            g.type[g.u(i, nj - 1)] = SF_NODE_DIRICHLET;
            g.bc[g.u(i, nj - 1)] = g.facev[nj + i];
        }
    }
    for (int j = 0; j < nj; j++) {
        if (g.face_type(2, j) == SF_FACE_DIRICHLET) {
            g.type[g.u(0, j)] = SF_NODE_DIRICHLET;
            g.bc[g.u(0, j)] = g.facev[nj + ni + j];
        }
        if (g.face_type(0, j) == SF_FACE_DIRICHLET) {
            g.type[g.u(ni - 1, j)] = SF_NODE_DIRICHLET;
            g.bc[g.u(ni - 1, j)] = g.facev[j];
        }
    }
    // a mesh without boundaries leaves nodes UNKNOWN in the reference until used; treat as open
    for (size_t k = 0; k < nn; k++)
        if (g.type[k] == SF_NODE_UNKNOWN) g.type[k] = SF_NODE_OPEN;
    *out = G;
    return SF_OK;
}

int sf_geometry_sizes(const sf_geometry* G, int* nseg, int* nlist) {
    if (!G || !nseg || !nlist) return SF_ERR_ARG;
    *nseg = (int)G->g.segs.size();
    size_t tot = 0;
    for (auto& l : G->g.nseg) tot += l.size();
    *nlist = (int)tot;
    return SF_OK;
}

int sf_geometry_segment_boundaries(const sf_geometry* G, int* seg_bnd) {
    if (!G || !seg_bnd) return SF_ERR_ARG;
    for (size_t s = 0; s < G->g.segs.size(); s++) seg_bnd[s] = G->g.segs[s].bnd;
    return SF_OK;
}

int sf_geometry_export(const sf_geometry* G, int* node_type, double* node_bc, double* node_vol, double* seg_xy,
                       int* seg_btype, int* seg_absorb, int* node_seg_ptr, int* node_seg_idx) {
    if (!G) return SF_ERR_ARG;
    const Geo& g = G->g;
    const size_t nn = (size_t)g.m.ni * g.m.nj;
    for (size_t k = 0; k < nn; k++) {
        if (node_type) node_type[k] = g.type[k];
        if (node_bc) node_bc[k] = g.bc[k];
        if (node_vol) node_vol[k] = g.vol[k];
    }
    for (size_t s = 0; s < g.segs.size(); s++) {
        if (seg_xy) {
            seg_xy[4 * s] = g.segs[s].x1[0];
            seg_xy[4 * s + 1] = g.segs[s].x1[1];
            seg_xy[4 * s + 2] = g.segs[s].x2[0];
            seg_xy[4 * s + 3] = g.segs[s].x2[1];
        }
        if (seg_btype) seg_btype[s] = g.segs[s].btype;
        if (seg_absorb) seg_absorb[s] = g.segs[s].absorb;
    }
    if (node_seg_ptr) {
        int k = 0;
        node_seg_ptr[0] = 0;
        for (size_t u = 0; u < nn; u++) {
            for (int q : g.nseg[u]) {
                if (node_seg_idx) node_seg_idx[k] = q;
                k++;
            }
            node_seg_ptr[u + 1] = k;
        }
    }
    return SF_OK;
}

int sf_geometry_free(sf_geometry* G) {
    delete G;
    return SF_OK;
}

// MaxwellianSource (sources/MaxwellianSource.java:44-125) on a linear polyline boundary
int sf_source_maxwellian(sf_ctx* ctx, int species, int npts, const double* pts, int reverse, double mdot,
                         double v_drift, double temp, int* source) {
    SF_CHECK_CTX(ctx);
    if (species < 0 || species >= (int)ctx->species.size() || npts < 2 || !pts || !(mdot >= 0) || !(temp >= 0))
        return sf_fail(ctx, SF_ERR_ARG, "sf_source_maxwellian: bad argument");
    if (ctx->mdesc.domain != SF_DOMAIN_XY)
        return sf_fail(ctx, SF_ERR_UNSUPPORTED, "sf_source_maxwellian: only XY boundary sources are supported");
    HostSource s;
    s.species = species;
    std::vector<LSeg> ss;
    for (int k = 0; k + 1 < npts; k++) {
        LSeg g{};
        if (!reverse) g.init(pts + 2 * k, pts + 2 * k + 2);
        else g.init(pts + 2 * k + 2, pts + 2 * k);
        ss.push_back(g);
    }
    if (reverse) std::reverse(ss.begin(), ss.end());
    s.cum.push_back(0.0);
    for (auto& g : ss) {
        s.seg.insert(s.seg.end(), {g.x1[0], g.x1[1], g.x2[0], g.x2[1]});
        s.area += g.length;                       // XY segment area = length (depth 1)
        s.cum.push_back(s.cum.back() + g.length);
    }
    s.mdot = mdot;
    s.v_drift = v_drift;
    s.v_th = std::sqrt(2 * K_B * temp / ctx->species[species]->mass);   // Utils.computeVth
    ctx->sources.push_back(s);
    if (source) *source = (int)ctx->sources.size() - 1;
    return SF_OK;
}

int sf_source_options(sf_ctx* ctx, int source, double v_az, int start_it, int stop_it) {
    SF_CHECK_CTX(ctx);
    if (source < 0 || source >= (int)ctx->sources.size())
        return sf_fail(ctx, SF_ERR_ARG, "sf_source_options: bad source index %d", source);
    if (!std::isfinite(v_az)) return sf_fail(ctx, SF_ERR_ARG, "sf_source_options: v_az must be finite");
    HostSource& s = ctx->sources[source];
    s.v_az = v_az;
    s.start_it = start_it;
    s.stop_it = stop_it < 0 ? -1 : stop_it;
    return SF_OK;
}

// SourceModule.sampleSources with the particles sampled on the device: the same particle counts
// (mdot dt / (m spwt) plus the carried remainder), positions and velocities drawn from
// per-particle Philox streams keyed by seed instead of the shared generator
int sf_sources_sample_device(sf_ctx* ctx, uint64_t seed, int64_t* injected) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    SF_HIP(ctx, hipSetDevice(ctx->device));
    int64_t total = 0;
    const uint32_t step = ctx->inject_step++;
    const int it = source_step(ctx);
    for (size_t q = 0; q < ctx->sources.size(); q++) {
        HostSource& s = ctx->sources[q];
        if (!source_active(s, it)) continue;
        Species* sp = ctx->species[s.species];
        double mp = (s.mdot * ctx->dt) / (sp->mass * sp->desc.spwt) + s.rem;
        int num = (int)mp;
        s.rem = mp - num;
        int64_t added = 0;
        int rc = inject_maxwellian_device(ctx, s, num, seed, step, RNG_MODULE_INJECT | (uint32_t)q, &added);
        if (rc) return rc;
        s.added += added;
        total += added;
    }
    if (injected) *injected = total;
    return SF_OK;
}

// SourceModule.sampleSources -> Source.regenerate / sampleKinetic / MaxwellianSource.sampleParticle
int sf_sources_sample(sf_ctx* ctx, int64_t* injected) {
    SF_CHECK_CTX(ctx);
    if (!ctx->mesh_ready) return sf_fail(ctx, SF_ERR_STATE, "mesh not set up");
    int64_t total = 0;
    const double dt = ctx->dt;
    const int it = source_step(ctx);
    for (HostSource& s : ctx->sources) {
        if (!source_active(s, it)) continue;
        Species* sp = ctx->species[s.species];
        double mp = (s.mdot * dt) / (sp->mass * sp->desc.spwt) + s.rem;
        int num = (int)mp;
        s.rem = mp - num;
        std::vector<double> x, y, u, v, w;
        x.reserve(num); y.reserve(num); u.reserve(num); v.reserve(num); w.reserve(num);
        const int nsg = (int)s.seg.size() / 4;
        while (num > 0) {
            // Spline.randomT (XY): uniform in length, Vec.binarySearch over cumulative lengths
            double A1 = ctx->rng.next_double() * s.area;
            int i1 = 0, i2 = (int)s.cum.size();
            int si;
            if (A1 < s.cum[0]) si = -1;
            else if (A1 > s.cum.back()) si = (int)s.cum.size();
            else {
                while (true) {
                    int mid = (int)(0.5 * (i1 + i2));
                    if (A1 < s.cum[mid]) i2 = mid;
                    else if (A1 > s.cum[mid]) i1 = mid;
                    else { i1 = mid; break; }
                    if ((i2 - i1) <= 1) break;
                }
                si = i1;
            }
            double seg_len = s.cum[si + 1] - s.cum[si];
            double frac = (A1 - s.cum[si]) / seg_len;
            double t = si + frac;
            int k = (int)t;
            double st = t - k;
            if (k > nsg - 1) { k = nsg - 1; st = 1.0; }
            const double* g = &s.seg[4 * k];
            double px = g[0] + st * (g[2] - g[0]), py = g[1] + st * (g[3] - g[1]);
            double dx = g[2] - g[0], dy = g[3] - g[1], len = std::sqrt(dx * dx + dy * dy);
            double n0 = -dy / len, n1 = dx / len;
            double vel[3];
            do {
                double vm[3];
                for (int a = 0; a < 3; a++) {     // Utils.SampleMaxw3D: (v_th*1.2) * SampleMaxw1D(1)
                    double sum = 0;
                    for (int q = 0; q < 12; q++) sum += ctx->rng.next_double();
                    double fm = (sum - 0.5 * 12) * std::sqrt(12.0 / 12);
                    vm[a] = (s.v_th * 1.2) * (std::sqrt(0.5) * 1.0 * fm);
                }
                vel[0] = vm[0] + n0 * s.v_drift;
                vel[1] = vm[1] + n1 * s.v_drift;
                vel[2] = vm[2] + s.v_az;
                (void)ctx->rng.next_double();    // part.dt = rnd*dt, overwritten by addParticle
            } while (n0 * vel[0] + n1 * vel[1] <= 0);
            num--;
            // Source.sampleKinetic nudge: pos += vel * 1e-6 * dt
            x.push_back(px + vel[0] * 1e-6 * dt);
            y.push_back(py + vel[1] * 1e-6 * dt);
            u.push_back(vel[0]);
            v.push_back(vel[1]);
            w.push_back(vel[2]);
        }
        if (!x.empty()) {
            int64_t before = sp->np;
            SF_HIP(ctx, hipSetDevice(ctx->device));
            int rc = add_particles_device(ctx, sp, (int64_t)x.size(), x.data(), y.data(), u.data(), v.data(), w.data(),
                                      nullptr, 1);
            if (rc) return rc;
            s.added += sp->np - before;
            total += sp->np - before;
        }
    }
    if (injected) *injected = total;
    return SF_OK;
}

// Source.mass_generated bookkeeping (Source.java:34-47, 193): the count of particles the source
// has added to the domain; mass = count * spwt * species mass.
int sf_source_added(sf_ctx* ctx, int source, int64_t* count) {
    SF_CHECK_CTX(ctx);
    if (source < 0 || source >= (int)ctx->sources.size() || !count)
        return sf_fail(ctx, SF_ERR_ARG, "sf_source_added: bad source index %d", source);
    *count = ctx->sources[source].added;
    return SF_OK;
}

}  // extern "C"
