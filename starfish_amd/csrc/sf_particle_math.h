// Device-side particle mechanics shared by the push kernels: field gather, leapfrog push,
// XY drift / RZ rotation and boundary processing.  Computation is FP64; particle state is
// stored as a packed cell index plus FP32 offsets and FP32 velocities.
#pragma once
#include "sf_internal.h"

namespace sf {

// Field2D.gather / gather_safe (core/domain/Field2D.java:300-390) on the FP32 E field.
// The reference evaluates the bilinear form and, when the Java array access goes out of
// bounds, falls back to gather_safe (clamped corner, zero weight beyond the edge).
template <class Load>
__device__ __forceinline__ void gather_e(const MeshP& m, double fi, double fj, Load load, double& ex, double& ey) {
    int i = (int)fi, j = (int)fj;
    double di = fi - i, dj = fj - j;
    if (i < 0 || j < 0 || i + 1 >= m.ni || j + 1 >= m.nj) {
        if (i < 0) { i = 0; di = 0; }
        if (j < 0) { j = 0; dj = 0; }
        if (i >= m.ni - 1) { i = m.ni - 1; di = 0; }
        if (j >= m.nj - 1) { j = m.nj - 1; dj = 0; }
        float2 e00 = load(i, j);
        double w00 = (1 - di) * (1 - dj);
        ex = w00 * e00.x;
        ey = w00 * e00.y;
        if (di > 0) {
            float2 e = load(i + 1, j);
            ex += di * (1 - dj) * e.x;
            ey += di * (1 - dj) * e.y;
        }
        if (di > 0 && dj > 0) {
            float2 e = load(i + 1, j + 1);
            ex += di * dj * e.x;
            ey += di * dj * e.y;
        }
        if (dj > 0) {
            float2 e = load(i, j + 1);
            ex += (1 - di) * dj * e.x;
            ey += (1 - di) * dj * e.y;
        }
        return;
    }
    float2 e00 = load(i, j), e10 = load(i + 1, j), e11 = load(i + 1, j + 1), e01 = load(i, j + 1);
    ex = (1 - di) * (1 - dj) * e00.x;
    ex += di * (1 - dj) * e10.x;
    ex += di * dj * e11.x;
    ex += (1 - di) * dj * e01.x;
    ey = (1 - di) * (1 - dj) * e00.y;
    ey += di * (1 - dj) * e10.y;
    ey += di * dj * e11.y;
    ey += (1 - di) * dj * e01.y;
}

__device__ __forceinline__ int face_type(const MeshP& m, const MeshD& md, int face, int idx) {
    // face layout RIGHT[nj], TOP[ni], LEFT[nj], BOTTOM[ni]
    int off = 0;
    if (face == 1) off = m.nj;
    else if (face == 2) off = m.nj + m.ni;
    else if (face == 3) off = 2 * m.nj + m.ni;
    return (int)md.face[off + idx];
}

// LinearSegment.intersect(p3,p4) (core/boundaries/LinearSegment.java:113-160); returns t along
// the particle path in t1 and along the segment in t0, or -1.
__device__ __forceinline__ void seg_intersect(const DevSeg& s, const double* p3, const double* p4, double& t0,
                                              double& t1) {
    const double* p1 = s.x1;
    const double* p2 = s.x2;
    double X1 = p1[0], X2 = p2[0], X3 = p3[0], X4 = p4[0];
    double Y1 = p1[1], Y2 = p2[1], Y3 = p3[1], Y4 = p4[1];
    double den = (X1 - X2) * (Y3 - Y4) - (Y1 - Y2) * (X3 - X4);
    if (den == 0) { t0 = -1; t1 = -1; return; }
    double xp0 = ((X1 * Y2 - Y1 * X2) * (X3 - X4) - (X1 - X2) * (X3 * Y4 - Y3 * X4)) / den;
    double xp1 = ((X1 * Y2 - Y1 * X2) * (Y3 - Y4) - (Y1 - Y2) * (X3 * Y4 - Y3 * X4)) / den;
    if (fabs(p2[0] - p1[0]) > 1e-6) t0 = (xp0 - p1[0]) / (p2[0] - p1[0]);
    else t0 = (xp1 - p1[1]) / (p2[1] - p1[1]);
    if (t0 < -FLT_EPS || t0 > (1 + FLT_EPS)) { t0 = -1; t1 = -1; return; }
    if (fabs(p4[0] - p3[0]) > 1e-6) t1 = (xp0 - p3[0]) / (p4[0] - p3[0]);
    else t1 = (xp1 - p3[1]) / (p4[1] - p3[1]);
    if (t1 < -FLT_EPS || t1 > (1 + FLT_EPS)) { t0 = -1; t1 = -1; return; }
    if (t0 < 0) t0 = 0;
    if (t1 < 0) t1 = 0;
    if (t0 > 1) t0 = 1;
    if (t1 > 1) t1 = 1;
}

// KineticMaterial.ProcessBoundary (core/materials/KineticMaterial.java:471-750) in logical
// coordinates.  lc/v are updated in place; dtr receives the remaining time.  Returns alive.
__device__ __forceinline__ bool process_boundary(const MeshP& m, const MeshD& md, double& lc0, double& lc1, double o0,
                                              double o1, double& v0, double& v1, double& v2, double& dtr) {
    const int ni = m.ni, nj = m.nj;
    double dt0 = dtr;
    dtr = 0;
    int i_min = (int)fmin(lc0, o0), i_max = (int)fmax(lc0, o0);
    int j_min = (int)fmin(lc1, o1), j_max = (int)fmax(lc1, o1);
    if (i_min < 0) i_min = 0;
    if (j_min < 0) j_min = 0;
    if (i_max >= ni) i_max = ni - 1;
    if (j_max >= nj) j_max = nj - 1;
    bool any = false;
    for (int i = i_min; i <= i_max && !any; i++)
        for (int j = j_min; j <= j_max; j++)
            if (md.node_hasseg[j * ni + i]) { any = true; break; }
    if (any) {
        double oldp[2] = {m.x0[0] + o0 * m.dh[0], m.x0[1] + o1 * m.dh[1]};
        double newp[2] = {m.x0[0] + lc0 * m.dh[0], m.x0[1] + lc1 * m.dh[1]};
        double tp_min = 2.0;
        int seg_min = -1;
        double vmag = sqrt(v0 * v0 + v1 * v1);
        for (int i = i_min; i <= i_max; i++)
            for (int j = j_min; j <= j_max; j++) {
                int n = j * ni + i;
                for (int q = md.nseg_ptr[n]; q < md.nseg_ptr[n + 1]; q++) {
                    int s = md.nseg_idx[q];
                    const DevSeg sg = md.seg[s];
                    double t0, t1;
                    seg_intersect(sg, oldp, newp, t0, t1);
                    if (t1 > 0) {
                        double acos_ = (sg.nrm[0] * v0 + sg.nrm[1] * v1) / vmag;
                        if (t1 < FLT_EPS && (acos_ > 0 || sg.btype == SF_BND_VIRTUAL)) continue;
                        if (t1 < tp_min) { tp_min = t1; seg_min = s; }
                    }
                }
            }
        if (seg_min >= 0) {
            tp_min *= 0.9999;
            double px = oldp[0] + tp_min * (newp[0] - oldp[0]);
            double py = oldp[1] + tp_min * (newp[1] - oldp[1]);
            lc0 = (px - m.x0[0]) * m.inv_dh[0];
            lc1 = (py - m.x0[1]) * m.inv_dh[1];
            dtr = dt0 * (1 - tp_min);
            if (lc0 < 0 && lc0 > -FLT_EPS) lc0 = 0;
            if (lc1 < 0 && lc1 > -FLT_EPS) lc1 = 0;
            const DevSeg sg = md.seg[seg_min];
            // default surface interaction of a boundary with a material absorbs; SINK always does.
            // An absorbed particle is tallied on its boundary (Material.addSurfaceMassDeposit,
            // Material.java:784-795, feeding StatsModule's flux columns through getFluxInst).
            if (sg.absorb || sg.btype == SF_BND_SINK) {
                if (md.bnd_hits != nullptr && sg.bnd >= 0) atomicAdd(md.bnd_hits + sg.bnd, 1ull);
                return false;
            }
        }
    }
    if (lc0 < 0 || lc1 < 0 || lc0 >= ni - 1 || lc1 >= nj - 1) {
        double t_right = 99, t_top = 99, t_left = 99, t_bottom = 99;
        if (lc0 >= ni - 1) t_right = (ni - 1.0 - o0) / (lc0 - o0);
        if (lc1 >= nj - 1) t_top = (nj - 1.0 - o1) / (lc1 - o1);
        if (lc0 < 0) t_left = o0 / (o0 - lc0);
        if (lc1 < 0) t_bottom = o1 / (o1 - lc1);
        int face = 0;  // RIGHT
        double t = t_right;
        if (t_top < t) { face = 1; t = t_top; }
        if (t_left < t) { face = 2; t = t_left; }
        if (t_bottom < t) { face = 3; t = t_bottom; }
        lc0 = o0 + t * (lc0 - o0);
        lc1 = o1 + t * (lc1 - o1);
        if (lc0 < 0) lc0 = 0;
        else if (lc0 > ni - 1) lc0 = ni - 1;
        if (lc1 < 0) lc1 = 0;
        else if (lc1 > nj - 1) lc1 = nj - 1;
        dtr = dt0 * (1 - t);
        int i = (int)lc0, j = (int)lc1;
        if (face == 1) j++;
        if (face == 0) i++;
        if (i < 0) i = 0;
        if (j < 0) j = 0;
        if (i >= ni - 1) i = ni - 1;
        if (j >= nj - 1) j = nj - 1;
        int type = (face == 2 || face == 0) ? face_type(m, md, face, j) : face_type(m, md, face, i);
        switch (type) {
            case SF_FACE_SYMMETRY:
                // Vec.mirror about the inward face normal flips the normal velocity component
                if (face == 0 || face == 2) v0 = -v0;
                else v1 = -v1;
                return true;
            case SF_FACE_PERIODIC:
                // the reference shifts the physical position; here the logical coordinate
                if (face == 2) lc0 += (ni - 1);
                else if (face == 0) lc0 -= (ni - 1);
                else if (face == 3) lc1 += (nj - 1);
                else lc1 -= (nj - 1);
                return true;
            default:
                return false;  // OPEN, DIRICHLET, NEUMANN, MESH without neighbour, SINK, CIRCUIT (ions)
        }
    }
    return true;
}

// One particle through ParticleMover.run (KineticMaterial.java:299-422), B = 0 branch.
// Inputs: logical coordinate and velocity; outputs updated in place.  Returns alive.
// HasSeg: (i, j) -> nonzero if node (i, j) lists a DIRICHLET/SINK segment (md.node_hasseg, or a
// staged copy of it).
// UpdateVelocityBoris (KineticMaterial.java:847-893) for E = (ex, ey, 0) and B = (bx, by, 0):
// half electric kick, rotation by t = (q/m) B dtp/2 and s = 2t / (1 + |t|^2), half kick.
// The z components of t and s are zero, so their cross-product terms drop out.
__device__ __forceinline__ void boris_velocity(double qm, double dtp, double ex, double ey, double bx, double by,
                                               double& v0, double& v1, double& v2) {
    const double t0 = qm * bx * 0.5 * dtp, t1 = qm * by * 0.5 * dtp;
    const double f = 2 / (1 + (t0 * t0 + t1 * t1));
    const double s0 = f * t0, s1 = f * t1;
    const double hx = qm * ex * 0.5 * dtp, hy = qm * ey * 0.5 * dtp;
    const double m0 = v0 + hx, m1 = v1 + hy, m2 = v2;
    const double p0 = m0 - m2 * t1, p1 = m1 + m2 * t0, p2 = m2 + (m0 * t1 - m1 * t0);
    v0 = m0 - p2 * s1 + hx;
    v1 = m1 + p2 * s0 + hy;
    v2 = m2 + (p0 * s1 - p1 * s0);
}

// Velocity update of ParticleMover.run / addParticle (KineticMaterial.java:336-353, 779-794):
// the electric kick over dtp where the gathered B is zero, Boris elsewhere.  HAS_B = false is
// the path with no magnetic field loaded anywhere.
template <bool HAS_B, class Load>
__device__ __forceinline__ void kick(const MeshP& m, const MeshD& md, double qm, double qm_dtp, double dtp, Load load,
                                     double lc0, double lc1, double& v0, double& v1, double& v2) {
    double ex, ey;
    gather_e(m, lc0, lc1, load, ex, ey);
    if (HAS_B) {
        double bx, by;
        auto load_b = [&](int i, int j) -> float2 { return md.bf[(size_t)j * m.ni + i]; };
        gather_e(m, lc0, lc1, load_b, bx, by);
        if (bx != 0 || by != 0) {
            boris_velocity(qm, dtp, ex, ey, bx, by, v0, v1, v2);
            return;
        }
    }
    v0 += qm_dtp * ex;
    v1 += qm_dtp * ey;
}

template <bool HAS_B = false, class Load, class HasSeg>
__device__ __forceinline__ bool push_particle(const MeshP& m, const MeshD& md, const PushC& pc, Load load,
                                              HasSeg hasseg, double& lc0, double& lc1, double& v0, double& v1,
                                              double& v2) {
    kick<HAS_B>(m, md, pc.qm, pc.qm_dt, pc.dt, load, lc0, lc1, v0, v1, v2);
    // Fast path: a single sub-move that stays inside the mesh and whose bounding box of nodes
    // lists no DIRICHLET/SINK segment.  ProcessBoundary then returns alive with no time left,
    // so this is the reference loop's first and only iteration.  Anything else takes the loop.
    {
        double n0, n1, nv0 = v0, nv1 = v1, nv2 = v2;
        if (m.domain == SF_DOMAIN_RZ) {
            double xr = m.x0[0] + lc0 * m.dh[0] + v0 * pc.dt;
            n1 = lc1 + v1 * pc.dt * m.inv_dh[1];
            double A = v2 * pc.dt;
            double R = sqrt(A * A + xr * xr);
            double c = xr / R, s = A / R;
            nv0 = c * v0 + s * v2;
            nv2 = -s * v0 + c * v2;
            n0 = (R - m.x0[0]) * m.inv_dh[0];
        } else if (m.domain == SF_DOMAIN_ZR) {
            double yr = m.x0[1] + lc1 * m.dh[1] + v1 * pc.dt;
            n0 = lc0 + v0 * pc.dt * m.inv_dh[0];
            double A = v2 * pc.dt;
            double R = sqrt(A * A + yr * yr);
            double c = yr / R, s = A / R;
            nv1 = c * v1 + s * v2;
            nv2 = -s * v1 + c * v2;
            n1 = (R - m.x0[1]) * m.inv_dh[1];
        } else {
            n0 = lc0 + v0 * pc.dt * m.inv_dh[0];
            n1 = lc1 + v1 * pc.dt * m.inv_dh[1];
        }
        if (n0 >= 0 && n1 >= 0 && n0 < m.ni - 1 && n1 < m.nj - 1) {
            int imin = (int)fmin(n0, lc0), imax = (int)fmax(n0, lc0);
            int jmin = (int)fmin(n1, lc1), jmax = (int)fmax(n1, lc1);
            if (imin >= 0 && jmin >= 0 && imax - imin <= 1 && jmax - jmin <= 1) {
                int seg = hasseg(imin, jmin) | hasseg(imax, jmin) | hasseg(imin, jmax) | hasseg(imax, jmax);
                if (!seg) {
                    lc0 = n0;
                    lc1 = n1;
                    v0 = nv0;
                    v1 = nv1;
                    v2 = nv2;
                    return true;
                }
            }
        }
    }
    double dtr = pc.dt;
    int bounces = 0;
    bool alive = true;
    while (dtr > 0 && bounces++ < 10) {
        double o0 = lc0, o1 = lc1;
        if (m.domain == SF_DOMAIN_RZ) {
            // rotateToRZ (KineticMaterial.java:424-442): r after the planar move, rotated by w*dt
            double xr = m.x0[0] + lc0 * m.dh[0] + v0 * dtr;
            lc1 += v1 * dtr * m.inv_dh[1];
            double A = v2 * dtr;
            double R = sqrt(A * A + xr * xr);
            double c = xr / R, s = A / R;
            double nv0 = c * v0 + s * v2;
            double nv2 = -s * v0 + c * v2;
            v0 = nv0;
            v2 = nv2;
            lc0 = (R - m.x0[0]) * m.inv_dh[0];
        } else if (m.domain == SF_DOMAIN_ZR) {
            // rotateToZR (KineticMaterial.java:444-462)
            double yr = m.x0[1] + lc1 * m.dh[1] + v1 * dtr;
            lc0 += v0 * dtr * m.inv_dh[0];
            double A = v2 * dtr;
            double R = sqrt(A * A + yr * yr);
            double c = yr / R, s = A / R;
            double nv1 = c * v1 + s * v2;
            double nv2 = -s * v1 + c * v2;
            v1 = nv1;
            v2 = nv2;
            lc1 = (R - m.x0[1]) * m.inv_dh[1];
        } else {
            lc0 += v0 * dtr * m.inv_dh[0];
            lc1 += v1 * dtr * m.inv_dh[1];
        }
        alive = process_boundary(m, md, lc0, lc1, o0, o1, v0, v1, v2, dtr);
        if (!alive) break;
    }
    return alive;
}

// Corner weights of Field2D.scatter (core/domain/Field2D.java:244-295) for a particle at
// cell (i,j) + offset (di,dj); RZ uses Ruyten's eq. 4.2 for the radial weight.
__device__ __forceinline__ void scatter_weights(const MeshP& m, int i, int j, double di, double dj, double* w) {
    if (m.domain == SF_DOMAIN_RZ) {
        double fi = i + di;
        double rp = m.x0[0] + (i + 1) * m.dh[0], rm = m.x0[0] + i * m.dh[0], r = m.x0[0] + fi * m.dh[0];
        di = 1 - (0.5 * (rp - r) * (2 * rp + 3 * rm - r) / (rp * rp - rm * rm));
    } else if (m.domain == SF_DOMAIN_ZR) {
        double fj = j + dj;
        double rp = m.x0[1] + (j + 1) * m.dh[1], rm = m.x0[1] + j * m.dh[1], r = m.x0[1] + fj * m.dh[1];
        dj = 1 - (0.5 * (rp - r) * (2 * rp + 3 * rm - r) / (rp * rp - rm * rm));
    }
    w[0] = (1 - di) * (1 - dj);   // (i, j)
    w[1] = di * (1 - dj);         // (i+1, j)
    w[2] = di * dj;               // (i+1, j+1)
    w[3] = (1 - di) * dj;         // (i, j+1)
}

// Mesh.nodeVol(i0, j0) (Mesh.java:998-1031) at a fractional position on a single uniform mesh:
// the corners (i0 -/+ 1/2, j0 -/+ 1/2) clamped to the mesh (controlVolumeLCs, Mesh.java:656-716),
// their positions, and the area of the quadrilateral as the sum of two triangles (Mesh.area,
// Mesh.java:965-985), evaluated in the reference's order without fused multiply-adds so that it
// rounds as the reference does (a DSMC pair count is the integer part of a product with it);
// depth 1 in XY, 2 pi A r in RZ/ZR with r moved a quarter cell inwards on the first/last line.
__device__ __forceinline__ double node_vol_at(const MeshP& m, double i0, double j0) {
#pragma clang fp contract(off)
    double V[4][2];
    const double di[4] = {-0.5, 0.5, 0.5, -0.5}, dj[4] = {-0.5, -0.5, 0.5, 0.5};
    for (int k = 0; k < 4; k++) {
        double i = i0 + di[k], j = j0 + dj[k];
        if (i < 0) i = 0;
        if (i > m.ni - 1) i = m.ni - 1;
        if (j < 0) j = 0;
        if (j > m.nj - 1) j = m.nj - 1;
        V[k][0] = m.x0[0] + i * m.dh[0];
        V[k][1] = m.x0[1] + j * m.dh[1];
    }
    double A = 0.5 * fabs(V[0][0] * V[1][1] - V[0][0] * V[2][1] + V[1][0] * V[2][1] - V[1][0] * V[0][1] +
                          V[2][0] * V[0][1] - V[2][0] * V[1][1]);
    A += 0.5 * fabs(V[0][0] * V[2][1] - V[0][0] * V[3][1] + V[2][0] * V[3][1] - V[2][0] * V[0][1] +
                    V[3][0] * V[0][1] - V[3][0] * V[2][1]);
    if (m.domain == SF_DOMAIN_XY) return 1.0 * A;
    double r;
    if (m.domain == SF_DOMAIN_RZ) {
        const double ir = i0 == 0 ? i0 + 0.25 : (i0 == m.ni - 1 ? i0 - 0.25 : i0);
        r = m.x0[0] + ir * m.dh[0];
    } else {
        const double jr = j0 == 0 ? j0 + 0.25 : (j0 == m.nj - 1 ? j0 - 0.25 : j0);
        r = m.x0[1] + jr * m.dh[1];
    }
    return 2 * M_PI * A * r;
}

__device__ __forceinline__ int cell_i(int32_t c) { return c & 0xFFFF; }
__device__ __forceinline__ int cell_j(int32_t c) { return (c >> 16) & 0xFFFF; }
__device__ __forceinline__ int32_t pack_cell(int i, int j) { return (j << 16) | i; }

// Exact round-to-nearest of a double below 2^51 in magnitude to int64 (1.5*2^52 trick):
// one FMA and one integer subtraction, order independent.
constexpr double QMAGIC = 6755399441055744.0;
__device__ __forceinline__ long long quant(double w, double val) {
    return (long long)(__double_as_longlong(fma(w, val, QMAGIC)) - __double_as_longlong(QMAGIC));
}

}  // namespace sf
